// extern "C" layer (include/ausaxs_hip.h). Exceptions never cross the ABI:
// each entry catches, stores the message in a thread-local last-error string
// and returns a non-zero status (the reference's execute_with_catch,
// source/api/api_helper.h:25-45).
#include "../../include/ausaxs_hip.h"

#include "fitter.hpp"
#include "grid.hpp"
#include "hydrate.hpp"
#include "hist.hpp"
#include "rigid.hpp"
#include "serve.hpp"
#include "symmetry.hpp"
#include "synth.hpp"

#include <cstring>
#include <string>

using namespace ausaxs::gpu;

struct ausaxs_structure {
    std::unique_ptr<ResidentStructure> s;
    bool timed = false;
    EngineTiming last{};
};

struct ausaxs_histogram {
    std::unique_ptr<CompositeDistanceHistogramHIP> h;
};

struct ausaxs_rigid {
    std::unique_ptr<RigidScan> r;
};

struct ausaxs_symmetry {
    std::unique_ptr<SymmetryStructure> s;
};

struct ausaxs_scalable {
    std::unique_ptr<ScalableGridStructure> s;
};

struct ausaxs_fitter {
    std::unique_ptr<HistogramFitter> f;
};

struct ausaxs_server {
    std::unique_ptr<Server> s;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return 1;
}

void require(bool ok, const char* msg) {
    if (!ok) throw std::runtime_error(msg);
}

HistSettings to_settings(const ausaxs_settings* s) {
    HistSettings h;
    if (s) {
        h.bin_width = s->bin_width;
        require(s->bin_count >= 10, "bin_count must be >= 10");
        h.bin_count = (uint32_t)s->bin_count;
        h.weighted_bins = s->weighted_bins != 0;
        h.debye_terms = s->debye_terms;
        h.tile_j = s->tile_j;
        h.lattice = s->lattice;
        h.surface_thickness = s->surface_thickness;
    }
    require(h.surface_thickness >= 0 && h.surface_thickness <= 64, "surface_thickness must lie in [0, 64] A");
    require(h.lattice >= 0 && h.lattice <= 2, "lattice must be 0, 1 or 2");
    require(h.bin_width > 0, "bin_width must be positive");
    require(h.debye_terms >= 0 && h.debye_terms <= 3, "debye_terms must be 0, 1, 2 or 3");
    return h;
}

Mode to_mode(int m) {
    require(m == AUSAXS_MODE_FFAVG || m == AUSAXS_MODE_FFEXPLICIT || m == AUSAXS_MODE_SIMPLE,
            "unknown mode (FFGrid structures are made by ausaxs_structure_create_grid)");
    return (Mode)m;
}
}  // namespace

extern "C" {

void ausaxs_default_settings(ausaxs_settings* s) {
    if (!s) return;
    HistSettings h;
    s->bin_width = h.bin_width;
    s->bin_count = (int)h.bin_count;
    s->weighted_bins = h.weighted_bins ? 1 : 0;
    s->debye_terms = h.debye_terms;
    s->tile_j = h.tile_j;
    s->lattice = h.lattice;
    s->surface_thickness = h.surface_thickness;
}

int ausaxs_get_last_error(char* buf, int len) {
    if (!buf || len <= 0) return 1;
    std::strncpy(buf, g_last_error.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return 0;
}

int ausaxs_device_info(char* name, int len, int* cus, size_t* lds) {
    return guarded([&] {
        Context& c = Context::get();
        if (name && len > 0) {
            std::strncpy(name, c.prop.gcnArchName, (size_t)len - 1);
            name[len - 1] = 0;
        }
        if (cus) *cus = c.prop.multiProcessorCount;
        if (lds) *lds = c.prop.sharedMemPerBlock;
    });
}

int ausaxs_structure_create(int mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                            const ausaxs_settings* settings, double z_exv, ausaxs_structure** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || atoms, "atoms is NULL");
        require(n_waters == 0 || waters, "waters is NULL");
        auto s = std::make_unique<ausaxs_structure>();
        s->s = std::make_unique<ResidentStructure>(to_mode(mode), atoms, n_atoms, waters, n_waters,
                                                   to_settings(settings), z_exv);
        *out = s.release();
    });
}

int ausaxs_structure_create_grid(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                 const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                 ausaxs_structure** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_waters == 0 || waters_xyz, "waters is NULL");
        auto s = std::make_unique<ausaxs_structure>();
        s->s = make_grid_structure(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters, to_settings(settings));
        *out = s.release();
    });
}

int ausaxs_structure_create_ffavg(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                  const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                  double* z_exv, ausaxs_structure** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_waters == 0 || waters_xyz, "waters is NULL");
        auto s = std::make_unique<ausaxs_structure>();
        s->s = make_ffavg_structure(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters, to_settings(settings));
        if (z_exv) *z_exv = s->s->z_exv();
        *out = s.release();
    });
}

int ausaxs_structure_create_grid_surface(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                         const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                         ausaxs_structure** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_waters == 0 || waters_xyz, "waters is NULL");
        auto s = std::make_unique<ausaxs_structure>();
        s->s = make_grid_surface_structure(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters, to_settings(settings));
        *out = s.release();
    });
}

int ausaxs_grid_exv_surface(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                            int n_waters, double* interior_xyz, int capacity_interior, int* n_interior,
                            double* surface_xyz, int capacity_surface, int* n_surface, double surface_thickness) {
    return guarded([&] {
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        GridExv g = make_grid_exv(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters, true, surface_thickness);
        const int ni = (int)(g.xyz.size() / 3), ns = (int)(g.surface.size() / 3);
        if (n_interior) *n_interior = ni;
        if (n_surface) *n_surface = ns;
        require((!interior_xyz || capacity_interior >= ni) && (!surface_xyz || capacity_surface >= ns),
                "capacity too small");
        if (interior_xyz) std::copy(g.xyz.begin(), g.xyz.end(), interior_xyz);
        if (surface_xyz) std::copy(g.surface.begin(), g.surface.end(), surface_xyz);
    });
}

void ausaxs_default_hydrate_settings(ausaxs_hydrate_settings* s) {
    if (!s) return;
    const HydrateSettings d;
    s->atomic_radius = d.atomic_radius;
    s->hydration_radius = d.hydration_radius;
    s->shell_correction = d.shell_correction;
    s->explicit_grid = 0;
    for (int k = 0; k < 3; ++k) s->grid_min[k] = s->grid_max[k] = 0;
}

int ausaxs_hydrate_radial(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const int* body_sizes,
                          int n_bodies, const double* noise, double* waters_xyz, int capacity, int* n_waters,
                          int* per_body, const ausaxs_hydrate_settings* settings) {
    return guarded([&] {
        require(n_atoms >= 0 && n_bodies >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_bodies == 0 || body_sizes, "body_sizes is NULL");
        std::vector<int> bodies(body_sizes, body_sizes + n_bodies);
        HydrateSettings hs;
        if (settings) {
            hs.atomic_radius = settings->atomic_radius;
            hs.hydration_radius = settings->hydration_radius;
            hs.shell_correction = settings->shell_correction;
            hs.explicit_grid = settings->explicit_grid != 0;
            for (int k = 0; k < 3; ++k) hs.grid_min[k] = settings->grid_min[k], hs.grid_max[k] = settings->grid_max[k];
        }
        Hydration h = hydrate_radial(atoms_xyz, atom_ff, n_atoms, bodies, noise, hs);
        const int m = (int)(h.waters.size() / 3);
        if (n_waters) *n_waters = m;
        if (per_body) std::copy(h.body_count.begin(), h.body_count.end(), per_body);
        require(!waters_xyz || capacity >= m, "waters capacity too small");
        if (waters_xyz) std::copy(h.waters.begin(), h.waters.end(), waters_xyz);
    });
}

int ausaxs_grid_exv(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                    int n_waters, double* out_xyz, int capacity, int* n_out, double* volume) {
    return guarded([&] {
        require(n_out != nullptr, "n_out is NULL");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_waters == 0 || waters_xyz, "waters is NULL");
        GridExv g = make_grid_exv(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters);
        const int n = (int)(g.xyz.size() / 3);
        *n_out = n;
        if (volume) *volume = g.volume;
        require(capacity >= n && (n == 0 || out_xyz), "out_xyz capacity too small");
        std::memcpy(out_xyz, g.xyz.data(), g.xyz.size() * sizeof(double));
    });
}

int ausaxs_histogram_grid_axes(const ausaxs_histogram* h, double* d_ax, double* d_xx) {
    return guarded([&] {
        require(h && d_ax && d_xx, "NULL argument");
        require(h->h->mode == Mode::FFGrid || h->h->mode == Mode::FFGridSurface, "not an FFGrid histogram");
        std::memcpy(d_ax, h->h->d_axis_ax.data(), h->h->d_axis_ax.size() * sizeof(double));
        std::memcpy(d_xx, h->h->d_axis_xx.data(), h->h->d_axis_xx.size() * sizeof(double));
    });
}

int ausaxs_structure_calculate(ausaxs_structure* s, ausaxs_histogram** out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        auto h = std::make_unique<ausaxs_histogram>();
        h->h = s->s->calculate(s->timed);
        s->last = s->s->engine().timing();
        *out = h.release();
    });
}

int ausaxs_structure_free(ausaxs_structure* s) {
    delete s;
    return 0;
}

int ausaxs_fitter_create(ausaxs_histogram* h, const double* q, const double* I, const double* err, int n,
                         unsigned enabled, const double* q_model, int nq_model, ausaxs_fitter** out) {
    return guarded([&] {
        require(h && out, "NULL argument");
        require(n >= 3 && q && I, "a fit needs at least three data points");
        require(nq_model >= 0 && (nq_model == 0 || q_model), "q_model is NULL");
        auto f = std::make_unique<ausaxs_fitter>();
        f->f = std::make_unique<HistogramFitter>(
            *h->h, std::vector<double>(q, q + n), std::vector<double>(I, I + n),
            err ? std::vector<double>(err, err + n) : std::vector<double>{}, enabled,
            std::vector<double>(q_model, q_model + nq_model));
        *out = f.release();
    });
}

int ausaxs_scalable_create(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                          int n_waters, const ausaxs_settings* settings, ausaxs_scalable** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_atoms >= 0 && n_waters >= 0, "negative sizes");
        require(n_atoms == 0 || (atoms_xyz && atom_ff), "atoms is NULL");
        require(n_waters == 0 || waters_xyz, "waters is NULL");
        auto s = std::make_unique<ausaxs_scalable>();
        s->s = std::make_unique<ScalableGridStructure>(atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters,
                                                       to_settings(settings));
        *out = s.release();
    });
}

int ausaxs_scalable_calculate(ausaxs_scalable* s, double k, ausaxs_histogram** out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        auto h = std::make_unique<ausaxs_histogram>();
        h->h = s->s->calculate(k);
        *out = h.release();
    });
}

int ausaxs_scalable_free(ausaxs_scalable* s) {
    delete s;
    return 0;
}

int ausaxs_fitter_create_scalable(ausaxs_scalable* s, const double* q, const double* I, const double* err, int n,
                                  unsigned enabled, const double* q_model, int nq_model, ausaxs_fitter** out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        require(n >= 3 && q && I, "a fit needs at least three data points");
        require(nq_model >= 0 && (nq_model == 0 || q_model), "q_model is NULL");
        auto f = std::make_unique<ausaxs_fitter>();
        f->f = std::make_unique<HistogramFitter>(
            *s->s, std::vector<double>(q, q + n), std::vector<double>(I, I + n),
            err ? std::vector<double>(err, err + n) : std::vector<double>{}, enabled,
            std::vector<double>(q_model, q_model + nq_model));
        *out = f.release();
    });
}

int ausaxs_fitter_params(const ausaxs_fitter* f, int* n_params, unsigned* enabled) {
    return guarded([&] {
        require(f != nullptr, "NULL fitter");
        if (n_params) *n_params = f->f->n_params();
        if (enabled) *enabled = f->f->enabled();
    });
}

int ausaxs_fitter_evaluate(ausaxs_fitter* f, const double* params, int n_sets, double* out) {
    return guarded([&] {
        require(f && out && n_sets >= 0, "bad argument");
        require(n_sets == 0 || f->f->n_params() == 0 || params, "params is NULL");
        f->f->evaluate(params, n_sets, out);
    });
}

int ausaxs_fitter_fit(ausaxs_fitter* f, int algorithm, int scan_evals, double* params, double* param_err,
                      double* lin, int* dof, int* fevals, double* I_fit, double* residuals) {
    return guarded([&] {
        require(f != nullptr, "NULL fitter");
        require(algorithm == 0 || algorithm == 2, "unknown minimizer (0 = golden, 2 = scan)");
        const FitOutcome r = f->f->fit((FitAlgorithm)algorithm, scan_evals);
        if (params) std::copy(r.params.begin(), r.params.end(), params);
        if (param_err) std::copy(r.param_err.begin(), r.param_err.end(), param_err);
        if (lin) {
            lin[0] = r.a, lin[1] = r.b, lin[2] = r.a_err, lin[3] = r.b_err, lin[4] = r.chi2;
        }
        if (dof) *dof = r.dof;
        if (fevals) *fevals = r.fevals;
        if (I_fit) std::copy(r.I_fit.begin(), r.I_fit.end(), I_fit);
        if (residuals) std::copy(r.residuals.begin(), r.residuals.end(), residuals);
    });
}

int ausaxs_fitter_free(ausaxs_fitter* f) {
    delete f;
    return 0;
}

namespace {
int symmetry_create(const float* atoms, const int* body_sizes, int n_bodies, const int* sym_per_body,
                    const int* sym_reps, const double* transforms, const int* sched_len, const int* schedule,
                    const float* waters, int n_waters, const ausaxs_settings* settings, bool partial,
                    ausaxs_symmetry** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_bodies >= 0 && n_waters >= 0, "negative sizes");
        require(n_bodies == 0 || (body_sizes && sym_per_body), "body arrays are NULL");
        require(n_waters == 0 || waters, "waters is NULL");
        std::vector<int> sizes(body_sizes, body_sizes + n_bodies);
        size_t n_atoms = 0;
        for (int v : sizes) {
            require(v >= 0, "negative body size");
            n_atoms += (size_t)v;
        }
        require(n_atoms == 0 || atoms, "atoms is NULL");
        std::vector<std::vector<SymmetrySpec>> syms(n_bodies);
        size_t k = 0, t = 0, e = 0;
        for (int b = 0; b < n_bodies; ++b) {
            require(sym_per_body[b] >= 0, "negative symmetry count");
            for (int i = 0; i < sym_per_body[b]; ++i, ++k) {
                require(sym_reps && sched_len, "symmetry arrays are NULL");
                SymmetrySpec sp;
                sp.reps = sym_reps[k];
                require(sp.reps >= 1, "a symmetry needs at least one repetition");
                require(transforms != nullptr, "transforms is NULL");
                sp.transforms.assign(transforms + 12 * t, transforms + 12 * (t + sp.reps));
                t += (size_t)sp.reps;
                require(sched_len[k] >= 0, "negative schedule length");
                require(sched_len[k] == 0 || schedule, "schedule is NULL");
                for (int j = 0; j < sched_len[k]; ++j, ++e)
                    sp.schedule.push_back(CopyPair{schedule[3 * e], schedule[3 * e + 1], schedule[3 * e + 2]});
                syms[b].push_back(std::move(sp));
            }
        }
        auto s = std::make_unique<ausaxs_symmetry>();
        s->s = std::make_unique<SymmetryStructure>(atoms, sizes, std::move(syms), waters, n_waters,
                                                   to_settings(settings), partial);
        *out = s.release();
    });
}
}  // namespace

int ausaxs_symmetry_create(const float* atoms, const int* body_sizes, int n_bodies, const int* sym_per_body,
                           const int* sym_reps, const double* transforms, const int* sched_len, const int* schedule,
                           const float* waters, int n_waters, const ausaxs_settings* settings,
                           ausaxs_symmetry** out) {
    return symmetry_create(atoms, body_sizes, n_bodies, sym_per_body, sym_reps, transforms, sched_len, schedule,
                           waters, n_waters, settings, false, out);
}

int ausaxs_symmetry_create_partial(const float* atoms, const int* body_sizes, int n_bodies, const int* sym_per_body,
                                   const int* sym_reps, const double* transforms, const int* sched_len,
                                   const int* schedule, const float* waters, int n_waters,
                                   const ausaxs_settings* settings, ausaxs_symmetry** out) {
    return symmetry_create(atoms, body_sizes, n_bodies, sym_per_body, sym_reps, transforms, sched_len, schedule,
                           waters, n_waters, settings, true, out);
}

int ausaxs_symmetry_set_body(ausaxs_symmetry* s, int body, const float* atoms, int internal) {
    return guarded([&] {
        require(s && atoms, "NULL argument");
        s->s->set_body(body, atoms, internal != 0);
    });
}

int ausaxs_symmetry_set_symmetry(ausaxs_symmetry* s, int body, int symmetry, int reps, const double* transforms,
                                 int sched_len, const int* schedule) {
    return guarded([&] {
        require(s != nullptr, "NULL structure");
        require(reps >= 1 && transforms, "a symmetry needs at least one repetition");
        require(sched_len >= 0 && (sched_len == 0 || schedule), "schedule is NULL");
        SymmetrySpec sp;
        sp.reps = reps;
        sp.transforms.assign(transforms, transforms + 12 * (size_t)reps);
        for (int j = 0; j < sched_len; ++j)
            sp.schedule.push_back(CopyPair{schedule[3 * j], schedule[3 * j + 1], schedule[3 * j + 2]});
        s->s->set_symmetry(body, symmetry, std::move(sp));
    });
}

int ausaxs_symmetry_set_waters(ausaxs_symmetry* s, const float* waters, int n_waters) {
    return guarded([&] {
        require(s != nullptr && n_waters >= 0 && (n_waters == 0 || waters), "bad argument");
        s->s->set_waters(waters, n_waters);
    });
}

int ausaxs_symmetry_calculate(ausaxs_symmetry* s, ausaxs_histogram** out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        auto h = std::make_unique<ausaxs_histogram>();
        h->h = s->s->calculate();
        *out = h.release();
    });
}

int ausaxs_symmetry_free(ausaxs_symmetry* s) {
    delete s;
    return 0;
}

int ausaxs_symmetry_info(const ausaxs_symmetry* s, long long* pairs, int* n_items) {
    return guarded([&] {
        require(s != nullptr, "NULL structure");
        if (pairs) *pairs = s->s->engine().pairs();
        if (n_items) *n_items = s->s->engine().n_items();
    });
}

namespace {
int put_schedule(const std::vector<CopyPair>& v, int* out, int capacity, int* n_out) {
    if (n_out) *n_out = (int)v.size();
    require(capacity >= (int)v.size() || !out, "schedule capacity too small");
    if (out)
        for (size_t i = 0; i < v.size(); ++i) {
            out[3 * i] = v[i].a;
            out[3 * i + 1] = v[i].b;
            out[3 * i + 2] = v[i].scale;
        }
    return 0;
}
}  // namespace

int ausaxs_symmetry_chain_schedule(int reps, int closed, int* out, int capacity, int* n_out) {
    return guarded([&] { put_schedule(chain_pair_schedule(reps, closed != 0), out, capacity, n_out); });
}

int ausaxs_symmetry_placement_schedule(const double* placements, int n, int* out, int capacity, int* n_out) {
    return guarded([&] {
        require(n >= 0 && (n == 0 || placements), "placements is NULL");
        put_schedule(placement_pair_schedule(std::vector<double>(placements, placements + 12 * (size_t)n)), out,
                     capacity, n_out);
    });
}

int ausaxs_symmetry_cyclic(const double* cm, const double* offset, const double* step, const double* axis,
                           double angle, int reps, double* out, int* closed) {
    return guarded([&] {
        require(cm && offset && step && axis && out, "NULL argument");
        bool c = false;
        const std::vector<double> t = cyclic_transforms(cm, offset, step, axis, angle, reps, &c);
        std::copy(t.begin(), t.end(), out);
        if (closed) *closed = c ? 1 : 0;
    });
}

int ausaxs_structure_set_timing(ausaxs_structure* s, int on) {
    return guarded([&] {
        require(s != nullptr, "NULL structure");
        s->timed = on != 0;
    });
}

int ausaxs_last_timing(const ausaxs_structure* s, double* pair_ms, double* total_ms, long long* pairs,
                       int* n_items, long long* kernel_pairs, long long* lattice_pairs) {
    return guarded([&] {
        require(s != nullptr, "NULL structure");
        if (pair_ms) *pair_ms = s->last.pair_ms;
        if (total_ms) *total_ms = s->last.total_ms;
        if (pairs) *pairs = s->last.pairs;
        if (n_items) *n_items = s->last.n_items;
        if (kernel_pairs) *kernel_pairs = s->last.kernel_pairs;
        if (lattice_pairs) *lattice_pairs = s->last.lattice_pairs;
    });
}

int ausaxs_last_merge_timing(const ausaxs_structure* s, double* merge_ms, long long* merge_bytes) {
    return guarded([&] {
        require(s != nullptr, "NULL structure");
        if (merge_ms) *merge_ms = s->last.merge_ms;
        if (merge_bytes) *merge_bytes = s->last.merge_bytes;
    });
}

int ausaxs_hist_calculate(int mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                          const ausaxs_settings* settings, double z_exv, ausaxs_histogram** out) {
    ausaxs_structure* s = nullptr;
    int rc = ausaxs_structure_create(mode, atoms, n_atoms, waters, n_waters, settings, z_exv, &s);
    if (rc) return rc;
    rc = ausaxs_structure_calculate(s, out);
    ausaxs_structure_free(s);
    return rc;
}

int ausaxs_histogram_info(const ausaxs_histogram* h, int* mode, int* max_bin) {
    return guarded([&] {
        require(h != nullptr, "NULL histogram");
        if (mode) *mode = (int)h->h->mode;
        if (max_bin) *max_bin = (int)h->h->max_bin;
    });
}

int ausaxs_histogram_get(const ausaxs_histogram* h, int which, double* value, uint32_t* count, double* center) {
    return guarded([&] {
        require(h != nullptr, "NULL histogram");
        const CompositeDistanceHistogramHIP& c = *h->h;
        const std::vector<WeightedEntry>* v = nullptr;
        std::vector<WeightedEntry> tmp;
        switch (which) {
            case AUSAXS_AA:
            case AUSAXS_AW: {
                const EntryRows& e = which == AUSAXS_AA ? c.aa : c.aw;
                tmp.resize(e.size());
                e.dense(tmp.data());
                v = &tmp;
                break;
            }
            case AUSAXS_WW: v = &c.ww; break;
            case AUSAXS_PTOT: v = &c.p_tot; break;
            case AUSAXS_PTOT_AX: v = &c.p_tot_ax; break;
            case AUSAXS_XX: v = &c.p_xx; break;
            case AUSAXS_XX_INTERIOR: v = &c.xx_interior; break;
            case AUSAXS_XX_SURFACE: v = &c.xx_surface; break;
            case AUSAXS_XX_CROSS: v = &c.xx_cross; break;
            case AUSAXS_WX_INTERIOR: v = &c.wx_interior; break;
            case AUSAXS_WX_SURFACE: v = &c.wx_surface; break;
            case AUSAXS_AX_INTERIOR:
            case AUSAXS_AX_SURFACE: {
                const EntryRows& e = which == AUSAXS_AX_INTERIOR ? c.ax_interior : c.ax_surface;
                require(c.mode == Mode::FFGridSurface, "not an FFGridSurface histogram");
                tmp.resize(e.size());
                e.dense(tmp.data());
                v = &tmp;
                break;
            }
            default: throw std::runtime_error("unknown distribution id");
        }
        for (size_t i = 0; i < v->size(); ++i) {
            if (value) value[i] = (*v)[i].value;
            if (count) count[i] = (*v)[i].count;
            if (center) center[i] = (*v)[i].bin_center;
        }
    });
}

int ausaxs_histogram_d_axis(const ausaxs_histogram* h, double* d) {
    return guarded([&] {
        require(h && d, "NULL argument");
        std::memcpy(d, h->h->d_axis.data(), h->h->d_axis.size() * sizeof(double));
    });
}

int ausaxs_histogram_set_params(ausaxs_histogram* h, double cw, double cx, double crho, double dw_atomic,
                                double dw_exv) {
    return guarded([&] {
        require(h != nullptr, "NULL histogram");
        h->h->params = FreeParams{cw, cx, crho, dw_atomic, dw_exv};
    });
}

int ausaxs_histogram_debye(ausaxs_histogram* h, const double* q, int nq, double* I, double* profiles) {
    return guarded([&] {
        require(h && q && I && nq >= 0, "bad argument");
        std::vector<double> qv(q, q + nq), prof;
        auto res = h->h->debye_transform(qv, profiles ? &prof : nullptr);
        std::memcpy(I, res.data(), (size_t)nq * sizeof(double));
        if (profiles) std::memcpy(profiles, prof.data(), prof.size() * sizeof(double));
    });
}

int ausaxs_histogram_free(ausaxs_histogram* h) {
    delete h;
    return 0;
}

int ausaxs_default_q_axis(double qmin, double qmax, double* q, int* nq) {
    return guarded([&] {
        require(q && nq, "NULL argument");
        auto v = default_q_axis(qmin, qmax);
        std::memcpy(q, v.data(), v.size() * sizeof(double));
        *nq = (int)v.size();
    });
}

int ausaxs_pair_histogram(const float* a, int n, const float* b, int m, double bin_width, int bin_count,
                          uint64_t* count, double* dsum) {
    return guarded([&] {
        require(a && count && dsum && n >= 0 && m >= 0 && bin_count > 0, "bad argument");
        HistSettings s;
        s.bin_width = bin_width;
        s.bin_count = (uint32_t)bin_count;
        raw_pair_histogram(a, n, b, m, s, count, dsum);
    });
}

int ausaxs_debye_contract(const double* H, int rows, int K, const double* d, const double* q, int nq, int terms,
                          double* out) {
    return guarded([&] {
        require(H && d && q && out && rows >= 0 && K >= 0 && nq >= 0, "bad argument");
        require(terms >= 0 && terms <= 3, "terms must be 0..3");
        Context& ctx = Context::get();
        DevBuf<double> dH, dd, dq, dout;
        DevBuf<char> scratch;
        dH.upload(H, (size_t)rows * K, ctx.stream());
        dd.upload(d, K, ctx.stream());
        dq.upload(q, nq, ctx.stream());
        dout.ensure((size_t)rows * nq);
        scratch.ensure(debye_scratch_bytes(rows, K, nq));
        DebyeLaunch dl{dH.p, dd.p, dq.p, rows, K, nq, terms, dout.p, scratch.p};
        check(terms > 0 ? launch_debye(dl, ctx.stream()) : launch_debye_f64(dl, ctx.stream()), "debye");
        check(hipMemcpyAsync(out, dout.p, (size_t)rows * nq * sizeof(double), hipMemcpyDeviceToHost, ctx.stream()),
              "D2H");
        ctx.sync();
    });
}

int ausaxs_debye_contract_noalloc(const double* H, int rows, int K, const double* d, const double* q, int nq,
                                  int terms, double* out, void* stream) {
    return guarded([&] {
        require(H && d && q && out && rows >= 0 && K >= 0 && nq >= 0, "bad argument");
        require(rows <= (1 << 20) && K <= (1 << 24) && nq <= (1 << 24), "rows, K or nq out of range");
        require(terms >= 0 && terms <= 3, "terms must be 0..3");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        DebyeLaunch dl{H, d, q, rows, K, nq, terms, out, nullptr};
        if (terms == 0) {
            check(launch_debye_f64(dl, s), "debye_f64");
            return;
        }
        if (debye_scratch_bytes(rows, K, nq) > kDebyeNoAllocBytes)
            throw std::runtime_error("ausaxs_hip: this contraction needs " +
                                     std::to_string(debye_scratch_bytes(rows, K, nq)) +
                                     " B of work space; the allocation-free entry point holds " +
                                     std::to_string(kDebyeNoAllocBytes) + " B (use ausaxs_debye_contract)");
        check(debye_noalloc_scratch(&dl.scratch), "hipGetSymbolAddress");
        check(launch_debye(dl, s), "debye");
    });
}

int ausaxs_bin_check(uint32_t lo, uint32_t hi, double bin_width, unsigned long long* mism) {
    return guarded([&] {
        require(mism && lo <= hi, "bad argument");
        Context& ctx = Context::get();
        DevBuf<unsigned long long> d;
        d.ensure(6);
        check(hipMemsetAsync(d.p, 0, 6 * sizeof(unsigned long long), ctx.stream()), "memset");
        check(launch_bin_check(lo, hi, (float)(1.0 / bin_width), d.p, ctx.stream()), "bin_check_kernel");
        check(hipMemcpyAsync(mism, d.p, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx.stream()), "D2H");
        ctx.sync();
    });
}

int ausaxs_copy_probe(size_t bytes, int reps, double* ms) {
    return guarded([&] {
        require(bytes % 16 == 0 && reps > 0, "bytes must be a multiple of 16, reps > 0");
        Context& ctx = Context::get();
        DevBuf<char> a, b;
        a.ensure(bytes);
        b.ensure(bytes);
        check(hipMemsetAsync(a.p, 1, bytes, ctx.stream()), "memset");
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        check(hipEventRecord(e0, ctx.stream()), "hipEventRecord");
        for (int r = 0; r < reps; ++r) check(launch_copy(a.p, b.p, bytes, ctx.stream()), "copy_kernel");
        check(hipEventRecord(e1, ctx.stream()), "hipEventRecord");
        ctx.sync();
        float t = 0;
        check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        if (ms) *ms = t / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int ausaxs_issue_probe(int nbins, int reps, double* rates) {
    return guarded([&] {
        require(rates != nullptr && reps > 0, "rates is NULL or reps < 1");
        require(nbins >= 1 && (size_t)nbins * sizeof(uint64_t) <= 64 * 1024, "nbins out of range");
        Context& ctx = Context::get();
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        DevBuf<float> out;
        out.ensure(probe_out_floats(cus));
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        for (int which = 0; which < 4; ++which) {
            double wi = 0;
            check(launch_issue_probe(which, cus, nbins, out.p, ctx.stream(), &wi), "issue probe");  // warm-up
            check(hipEventRecord(e0, ctx.stream()), "hipEventRecord");
            for (int r = 0; r < reps; ++r)
                check(launch_issue_probe(which, cus, nbins, out.p, ctx.stream(), &wi), "issue probe");
            check(hipEventRecord(e1, ctx.stream()), "hipEventRecord");
            ctx.sync();
            float t = 0;
            check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
            rates[which] = wi * reps / ((double)t * 1e6) / cus;  // per CU per ns
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int ausaxs_synth_globule(uint64_t seed, int n, double ax, double ay, double az, double r_excl, float* out) {
    return guarded([&] {
        require(out != nullptr || n == 0, "out is NULL");
        synth_globule(seed, n, ax, ay, az, r_excl, out);
    });
}

int ausaxs_synth_shell(uint64_t seed, int m, double ax, double ay, double az, double offset, float* out) {
    return guarded([&] {
        require(out != nullptr || m == 0, "out is NULL");
        synth_shell(seed, m, ax, ay, az, offset, out);
    });
}


int ausaxs_rigid_create(const double* xyz, const float* w, const int32_t* body_sizes, int n_bodies,
                        const double* waters_xyz, const float* waters_w, int n_waters, const ausaxs_settings* settings,
                        ausaxs_rigid** out) {
    return guarded([&] {
        require(out && body_sizes && n_bodies >= 1 && n_waters >= 0, "bad argument");
        int64_t n = 0;
        for (int b = 0; b < n_bodies; ++b) n += body_sizes[b];
        require(n == 0 || (xyz && w), "bad argument: body coordinates");
        require(n_waters == 0 || (waters_xyz && waters_w), "bad argument: waters");
        HistSettings s = to_settings(settings);
        if (settings) s.tile_j = settings->tile_j;  // 0: autotune
        auto r = std::make_unique<ausaxs_rigid>();
        r->r = std::make_unique<RigidScan>(xyz, w, body_sizes, n_bodies, waters_xyz, waters_w, n_waters, s);
        *out = r.release();
    });
}

int ausaxs_rigid_scan(ausaxs_rigid* r, const double* transforms, int n_conf, const double* q, int nq, double cw,
                      double* I) {
    return guarded([&] {
        require(r && r->r && n_conf >= 0 && nq >= 0, "bad argument");
        require((n_conf == 0 || transforms) && (nq == 0 || q) && (n_conf == 0 || nq == 0 || I), "bad argument");
        r->r->scan(transforms, n_conf, q, nq, cw, I);
    });
}

int ausaxs_rigid_scan_hydrated(ausaxs_rigid* r, const double* transforms, int n_conf, const double* waters_xyz,
                               const float* waters_w, const int32_t* water_counts, const double* q, int nq,
                               double cw, double* I) {
    return guarded([&] {
        require(r && r->r && n_conf >= 0 && nq >= 0, "bad argument");
        require((n_conf == 0 || (transforms && water_counts)) && (nq == 0 || q) && (n_conf == 0 || nq == 0 || I),
                "bad argument");
        r->r->scan_hydrated(transforms, n_conf, waters_xyz, waters_w, water_counts, q, nq, cw, I);
    });
}

int ausaxs_rigid_histogram(ausaxs_rigid* r, int conformation, ausaxs_histogram** out) {
    return guarded([&] {
        require(r && r->r && out, "bad argument");
        auto h = std::make_unique<ausaxs_histogram>();
        h->h = r->r->histogram(conformation);
        *out = h.release();
    });
}

int ausaxs_rigid_info(const ausaxs_rigid* r, int* tile_j, int* grid, int* copies, int* batch, int* n_items,
                      long long* pairs, double* scan_ms) {
    return guarded([&] {
        require(r && r->r, "bad argument");
        const RigidInfo& i = r->r->info();
        if (tile_j) *tile_j = i.tile_j;
        if (copies) *copies = i.copies;
        if (grid) *grid = i.grid;
        if (batch) *batch = i.batch;
        if (n_items) *n_items = i.n_items;
        if (pairs) *pairs = (long long)i.pairs;
        if (scan_ms) *scan_ms = i.total_ms;
    });
}

int ausaxs_rigid_free(ausaxs_rigid* r) {
    return guarded([&] { delete r; });
}

int ausaxs_server_create(int slots, double budget_gb, ausaxs_server** out) {
    return guarded([&] {
        require(out && budget_gb > 0, "bad argument");
        auto s = std::make_unique<ausaxs_server>();
        s->s = std::make_unique<Server>(slots, (size_t)(budget_gb * (1ull << 30)));
        *out = s.release();
    });
}

int ausaxs_server_submit(ausaxs_server* s, int mode, const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                         const double* waters_xyz, int n_waters, const double* q, int nq,
                         const ausaxs_settings* settings, double z_exv, long long* id) {
    return guarded([&] {
        require(s && s->s && id, "bad argument");
        require(n_atoms >= 0 && n_waters >= 0 && nq >= 0, "negative sizes");
        require((n_atoms == 0 || (atoms_xyz && atom_ff)) && (n_waters == 0 || waters_xyz) && (nq == 0 || q),
                "bad argument");
        require(mode == AUSAXS_MODE_FFAVG || mode == AUSAXS_MODE_FFEXPLICIT || mode == AUSAXS_MODE_FFGRID,
                "the server scores FFAVG, FFEXPLICIT and FFGRID requests");
        *id = s->s->submit((Mode)mode, atoms_xyz, atom_ff, n_atoms, waters_xyz, n_waters, q, nq,
                           to_settings(settings), z_exv);
    });
}

int ausaxs_server_wait(ausaxs_server* s, long long id, double timeout_s, double* I, double* latency_ms,
                       double* queue_ms, int* ready) {
    return guarded([&] {
        require(s && s->s && ready, "bad argument");
        std::vector<double> v;
        double lat = 0, qu = 0;
        *ready = s->s->wait(id, timeout_s, v, lat, qu) ? 1 : 0;
        if (!*ready) return;
        if (I) std::copy(v.begin(), v.end(), I);
        if (latency_ms) *latency_ms = lat;
        if (queue_ms) *queue_ms = qu;
    });
}

int ausaxs_server_stats(ausaxs_server* s, long long* submitted, long long* completed, long long* failed,
                        int* in_flight, int* peak_in_flight, double* peak_gb, long long* pairs,
                        long long* kernel_pairs) {
    return guarded([&] {
        require(s && s->s, "bad argument");
        const ServeStats st = s->s->stats();
        if (submitted) *submitted = st.submitted;
        if (completed) *completed = st.completed;
        if (failed) *failed = st.failed;
        if (in_flight) *in_flight = st.in_flight;
        if (peak_in_flight) *peak_in_flight = st.peak_in_flight;
        if (peak_gb) *peak_gb = (double)st.peak_in_use / (1ull << 30);
        if (pairs) *pairs = st.pairs;
        if (kernel_pairs) *kernel_pairs = st.kernel_pairs;
    });
}

int ausaxs_server_batch_stats(ausaxs_server* s, long long* cache_hits, long long* cache_misses, long long* batches,
                              long long* batched, long long* evictions, double* cached_gb, long long* express) {
    return guarded([&] {
        require(s && s->s, "bad argument");
        const ServeStats st = s->s->stats();
        if (cache_hits) *cache_hits = st.cache_hits;
        if (cache_misses) *cache_misses = st.cache_misses;
        if (batches) *batches = st.batches;
        if (batched) *batched = st.batched;
        if (evictions) *evictions = st.evictions;
        if (cached_gb) *cached_gb = (double)st.cached_bytes / (1ull << 30);
        if (express) *express = st.express;
    });
}

int ausaxs_pool_stats(size_t* device_held, size_t* device_in_use, long long* device_mallocs,
                      long long* host_mallocs, long long* reuses) {
    return guarded([&] {
        const PoolStats p = pool_stats();
        if (device_held) *device_held = p.device_held;
        if (device_in_use) *device_in_use = p.device_in_use;
        if (device_mallocs) *device_mallocs = p.device_mallocs;
        if (host_mallocs) *host_mallocs = p.host_mallocs;
        if (reuses) *reuses = p.reuses;
    });
}

int ausaxs_pool_trim(void) {
    return guarded([&] {
        Context::get();  // a device must be present
        pool_trim();
    });
}

int ausaxs_server_free(ausaxs_server* s) {
    return guarded([&] { delete s; });
}

int ausaxs_cubic_spline(const double* x, const double* y, int n, const double* z, int m, double* out) {
    return guarded([&] {
        require(n >= 0 && m >= 0, "negative sizes");
        require((x && y) || n == 0, "x or y is NULL");
        require((z && out) || m == 0, "z or out is NULL");
        const CubicSpline sp(std::vector<double>(x, x + n), std::vector<double>(y, y + n));
        for (int i = 0; i < m; ++i) out[i] = sp(z[i]);
    });
}

int ausaxs_lsq(const double* x, const double* y, const double* err, int n, int n_models, double* out) {
    return guarded([&] {
        require(n >= 3 && n_models >= 0 && (n_models == 0 || (x && y && out)), "bad argument (n >= 3 points)");
        Context& ctx = Context::get();
        DevBuf<double> dx, dy, de, dout;
        dx.upload(x, (size_t)n_models * n, ctx.stream());
        dy.upload(y, n, ctx.stream());
        if (err) de.upload(err, n, ctx.stream());
        dout.ensure((size_t)n_models * 5);
        check(launch_lsq(dx.p, dy.p, err ? de.p : nullptr, n, n_models, dout.p, ctx.stream()), "lsq_kernel");
        check(hipMemcpyAsync(out, dout.p, (size_t)n_models * 5 * sizeof(double), hipMemcpyDeviceToHost,
                             ctx.stream()),
              "D2H");
        ctx.sync();
    });
}

int ausaxs_exact_debye(const float* xyzw, int n, const double* q, int nq, double* I) {
    return guarded([&] {
        require(n >= 0 && nq >= 0 && (n == 0 || xyzw) && (nq == 0 || (q && I)), "bad argument");
        if (nq == 0) return;
        Context& ctx = Context::get();
        DevBuf<float4> dp;
        DevBuf<double> dq, dpart, dout;
        std::vector<float4> pts(n);
        for (int i = 0; i < n; ++i) pts[i] = make_float4(xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], xyzw[4 * i + 3]);
        dp.upload(pts.data(), pts.size(), ctx.stream());
        dq.upload(q, nq, ctx.stream());
        dpart.ensure((size_t)nq * std::max(1, exact_debye_blocks(n)));
        dout.ensure(nq);
        check(launch_exact_debye(dp.p, n, dq.p, nq, dpart.p, dout.p, ctx.stream()), "exact_debye");
        check(hipMemcpyAsync(I, dout.p, (size_t)nq * sizeof(double), hipMemcpyDeviceToHost, ctx.stream()), "D2H");
        ctx.sync();
    });
}

}  // extern "C"
