/*
 * ausaxs_oracle.c — CPU restatement of AUSAXS's Debye hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker. Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Build: see oracle/Makefile (IEEE FP32: -ffp-contract=off, no fast-math, so
 * the SIMD-body rule keeps its separately rounded products and sums and the
 * scalar rule gets its explicit fmaf).
 *
 * Parity pins (tests/test_oracle.py):
 *   - binning rules: lattice inputs with closed-form sqrt(k) distances,
 *     brute force over float d2 ranges, hand-derived ties.
 *   - manager conventions: 8-atom cube closed form (histogram_manager.cpp:30-146),
 *     partial-sum invariants, count totals n(n-1)/2.
 *   - sinc/Debye: closed forms of the cube I(q) (composite_distance_histogram_ff_avg.cpp:23-190),
 *     exact Debye vs binned Debye, carbon_sphere_waxs golden curve.
 *   - form factors: I0 = sum(a)+c closed form, q-monotonicity, ExvFormFactor closed form.
 */
#include "ausaxs_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* binning rules                                                            */
/* ------------------------------------------------------------------------ */

/* SIMD body: CompactCoordinatesXYZFF.h:514-545 (AVX2), :397-420 (SSE),
 * CompactCoordinatesXYZW.h:491-535. d = this - other, squares, (x+y)+z,
 * sqrt, mul by inv_width, cvtps2dq (round-half-even). */
static inline void rule_simd(const float* a, const float* b, float inv_w, float* d_out, int32_t* bin_out) {
    float dx = a[0] - b[0];
    float dy = a[1] - b[1];
    float dz = a[2] - b[2];
    float sx = dx * dx;
    float sy = dy * dy;
    float sz = dz * dz;
    float d2 = (sx + sy) + sz;
    float d = sqrtf(d2);
    float x = d * inv_w;
    *d_out = d;
    *bin_out = (int32_t)nearbyintf(x);
}

/* Scalar tail: evaluate_scalar (CompactCoordinatesXYZFF.h:313-329,
 * CompactCoordinatesXYZW.h:303-315) as compiled with the reference's release
 * flags: FMA-contracted squared distance and std::round lowered to
 * trunc(x + copysign(0.49999997f, x)). */
static inline void rule_scalar(const float* a, const float* b, float inv_w, float* d_out, int32_t* bin_out) {
    float dx = a[0] - b[0];
    float dy = a[1] - b[1];
    float dz = a[2] - b[2];
    float d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
    float d = sqrtf(d2);
    float x = d * inv_w;
    *d_out = d;
    *bin_out = (int32_t)truncf(x + copysignf(0.49999997f, x));
}

void orc_bin_pair(const float* a, const float* b, float inv_w, int rule, float* d_out, int32_t* bin_out) {
    if (rule) rule_scalar(a, b, inv_w, d_out, bin_out);
    else rule_simd(a, b, inv_w, d_out, bin_out);
}

/* Loop geometry of every self loop in the reference (e.g.
 * HistogramManagerMTFFAvg.cpp:214-234): j = i+1, then blocks of 16, 8, 4 while
 * they fit, then evaluate1 for the rest: the last (n-i-1) mod 4 columns. */
int orc_rule_self(int n, int i, int j) {
    int tail_start = n - ((n - i - 1) % 4);
    return j >= tail_start;
}

/* Cross loops (e.g. HistogramManagerMTFFAvg.cpp:237-257): j from 0 over m
 * columns; the last m mod 4 use evaluate1. */
int orc_rule_cross(int m, int j) {
    return j >= m - (m % 4);
}

/* ------------------------------------------------------------------------ */
/* weighted entries (WeightedEntry.h:64-81)                                 */
/* ------------------------------------------------------------------------ */

static inline void inc_n(orc_entry* e, int n, float d) {
    e->count += (uint32_t)n;
    e->bin_center += (double)((float)n * d);
    e->value += n;
}

static inline void add_n(orc_entry* e, int n, float d, float w) {
    e->count += (uint32_t)n;
    e->bin_center += (double)((float)n * d);
    e->value += n * (double)w;
}

static inline void entry_add(orc_entry* a, const orc_entry* b) {
    a->value += b->value;
    a->count += b->count;
    a->bin_center += b->bin_center;
}

/* WeightedEntry operator*(entry, double) (WeightedEntry.h:113-115): the count
 * goes through double and is truncated back to unsigned. */
static inline void entry_scale(orc_entry* e, double f) {
    e->value *= f;
    e->count = (uint32_t)(f * e->count);
    e->bin_center *= f;
}

static inline int ff_of(const float* p) {
    int32_t v;
    memcpy(&v, &p[3], 4);
    return v;
}

/* ------------------------------------------------------------------------ */
/* thread pool over 800-row tasks (GeneralSettings.cpp:50)                  */
/* ------------------------------------------------------------------------ */

#define ORC_JOB 800

typedef struct job_ctx job_ctx;
typedef int (*row_fn)(job_ctx* c, void* priv, int i);

struct job_ctx {
    int n_rows;
    atomic_int next;
    atomic_int err;
    row_fn fn;
    void* args;
    void** privs;
    int tid_counter;
    pthread_mutex_t mu;
};

typedef struct {
    job_ctx* c;
    int tid;
} worker_arg;

static void* worker(void* p) {
    worker_arg* wa = (worker_arg*)p;
    job_ctx* c = wa->c;
    void* priv = c->privs[wa->tid];
    for (;;) {
        int start = atomic_fetch_add(&c->next, ORC_JOB);
        if (start >= c->n_rows) break;
        int end = start + ORC_JOB < c->n_rows ? start + ORC_JOB : c->n_rows;
        for (int i = start; i < end; ++i) {
            int e = c->fn(c, priv, i);
            if (e) atomic_store(&c->err, e);
        }
    }
    return NULL;
}

static int run_rows(int n_rows, int nthreads, row_fn fn, void* args, void** privs) {
    job_ctx c;
    c.n_rows = n_rows;
    atomic_init(&c.next, 0);
    atomic_init(&c.err, 0);
    c.fn = fn;
    c.args = args;
    c.privs = privs;
    if (nthreads <= 1) {
        worker_arg wa = {&c, 0};
        worker(&wa);
        return atomic_load(&c.err);
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    worker_arg* wa = (worker_arg*)malloc(sizeof(worker_arg) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        wa[t].c = &c;
        wa[t].tid = t;
        pthread_create(&th[t], NULL, worker, &wa[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(wa);
    return atomic_load(&c.err);
}

/* ------------------------------------------------------------------------ */
/* FF managers                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    int mode;
    const float* a;
    int na;
    const float* w;
    int nw;
    float inv_w;
    int nb;
} ff_args;

typedef struct {
    orc_entry* aa; /* [15][15][nb] */
    orc_entry* aw; /* [15][nb] */
    orc_entry* ww; /* [nb] */
} ff_priv;

#define IDX3(f1, f2, b, nb) (((size_t)(f1) * ORC_MAX_FF + (f2)) * (size_t)(nb) + (b))
#define IDX2(f, b, nb) ((size_t)(f) * (size_t)(nb) + (b))

/* evaluate a block of k columns with one rule, then scatter — the shape of
 * add16/add8/add4 + increment_* in the reference (TemplateHelperBase.h). */
static int aa_row(job_ctx* c, void* pv, int i) {
    ff_args* A = (ff_args*)c->args;
    ff_priv* P = (ff_priv*)pv;
    const float* a = A->a;
    int n = A->na, nb = A->nb;
    const float* pi = a + 4 * (size_t)i;
    int fi = ff_of(pi);
    int tail = n - ((n - i - 1) % 4);
    float d[16];
    int32_t bin[16];
    int j = i + 1;
    while (j < n) {
        int k;
        if (j < tail) {
            k = (j + 15 < n) ? 16 : (j + 7 < n) ? 8 : 4;
            for (int t = 0; t < k; ++t) rule_simd(pi, a + 4 * (size_t)(j + t), A->inv_w, &d[t], &bin[t]);
        } else {
            k = 1;
            rule_scalar(pi, a + 4 * (size_t)j, A->inv_w, &d[0], &bin[0]);
        }
        for (int t = 0; t < k; ++t) {
            if (bin[t] < 0 || bin[t] >= nb) return -2;
            int fj = ff_of(a + 4 * (size_t)(j + t));
            inc_n(&P->aa[IDX3(fi, fj, bin[t], nb)], 2, d[t]);
            if (A->mode == ORC_FFAVG) {
                /* HistogramManagerMTFFAvg.cpp:27-34: [ff_i][exv] and [exv][exv] */
                inc_n(&P->aa[IDX3(fi, ORC_EXV, bin[t], nb)], 2, d[t]);
                inc_n(&P->aa[IDX3(ORC_EXV, ORC_EXV, bin[t], nb)], 2, d[t]);
            }
        }
        j += k;
    }
    return 0;
}

/* atoms (rows) x waters (columns): FFAvg :237-257, FFExplicit :260-280 */
static int aw_row(job_ctx* c, void* pv, int i) {
    ff_args* A = (ff_args*)c->args;
    ff_priv* P = (ff_priv*)pv;
    int m = A->nw, nb = A->nb;
    const float* pi = A->a + 4 * (size_t)i;
    int fi = ff_of(pi);
    int tail = m - (m % 4);
    for (int j = 0; j < m; ++j) {
        float d;
        int32_t bin;
        if (j < tail) rule_simd(pi, A->w + 4 * (size_t)j, A->inv_w, &d, &bin);
        else rule_scalar(pi, A->w + 4 * (size_t)j, A->inv_w, &d, &bin);
        if (bin < 0 || bin >= nb) return -2;
        inc_n(&P->aw[IDX2(fi, bin, nb)], 1, d);
        if (A->mode == ORC_FFAVG) inc_n(&P->aw[IDX2(ORC_EXV, bin, nb)], 1, d); /* FFAvg.cpp:111-118 */
    }
    return 0;
}

static int ww_row(job_ctx* c, void* pv, int i) {
    ff_args* A = (ff_args*)c->args;
    ff_priv* P = (ff_priv*)pv;
    int n = A->nw, nb = A->nb;
    const float* pi = A->w + 4 * (size_t)i;
    int tail = n - ((n - i - 1) % 4);
    for (int j = i + 1; j < n; ++j) {
        float d;
        int32_t bin;
        if (j < tail) rule_simd(pi, A->w + 4 * (size_t)j, A->inv_w, &d, &bin);
        else rule_scalar(pi, A->w + 4 * (size_t)j, A->inv_w, &d, &bin);
        if (bin < 0 || bin >= nb) return -2;
        inc_n(&P->ww[bin], 2, d);
    }
    return 0;
}

static int max_bin_of(const orc_entry* p, int nb) {
    /* HistogramManagerMTFFAvg.cpp:340-347 — compares `value` (operator==) */
    for (int i = nb - 1; i >= 10; --i)
        if (p[i].value != 0) return i + 1;
    return 10;
}

int orc_hist_ff(int mode, const float* atoms, int n_a, const float* waters, int n_w,
                float inv_w, int nb, double z_exv, int nthreads,
                orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot) {
    for (int i = 0; i < n_a; ++i) {
        int f = ff_of(atoms + 4 * (size_t)i);
        if (f < 1 || f >= ORC_MAX_FF) return -3; /* CompactCoordinatesFF.h:21-32 */
    }
    if (nthreads < 1) nthreads = 1;
    size_t n3 = (size_t)ORC_MAX_FF * ORC_MAX_FF * nb, n2 = (size_t)ORC_MAX_FF * nb;
    ff_priv* privs = (ff_priv*)calloc(nthreads, sizeof(ff_priv));
    void** pp = (void**)malloc(sizeof(void*) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        privs[t].aa = (orc_entry*)calloc(n3, sizeof(orc_entry));
        privs[t].aw = (orc_entry*)calloc(n2, sizeof(orc_entry));
        privs[t].ww = (orc_entry*)calloc(nb, sizeof(orc_entry));
        pp[t] = &privs[t];
    }
    ff_args A = {mode, atoms, n_a, waters, n_w, inv_w, nb};
    int err = 0;
    err |= run_rows(n_a, nthreads, aa_row, &A, pp);
    err |= run_rows(n_a, nthreads, aw_row, &A, pp);
    err |= run_rows(n_w, nthreads, ww_row, &A, pp);

    /* ThreadLocalWrapper::merge (ThreadLocalWrapper.h:89-105) */
    memset(aa, 0, n3 * sizeof(orc_entry));
    memset(aw, 0, n2 * sizeof(orc_entry));
    memset(ww, 0, nb * sizeof(orc_entry));
    for (int t = 0; t < nthreads; ++t) {
        for (size_t k = 0; k < n3; ++k) entry_add(&aa[k], &privs[t].aa[k]);
        for (size_t k = 0; k < n2; ++k) entry_add(&aw[k], &privs[t].aw[k]);
        for (int k = 0; k < nb; ++k) entry_add(&ww[k], &privs[t].ww[k]);
        free(privs[t].aa);
        free(privs[t].aw);
        free(privs[t].ww);
    }
    free(privs);
    free(pp);
    if (err) return -2;

    /* self correlations: FFAvg.cpp:308-323, FFExplicit.cpp:331-346 */
    for (int i = 0; i < n_a; ++i) {
        int f = ff_of(atoms + 4 * (size_t)i);
        inc_n(&aa[IDX3(f, f, 0, nb)], 1, 0.0f);
    }
    if (mode == ORC_FFAVG) {
        orc_entry e = {(double)n_a, (uint32_t)n_a, 0};
        entry_add(&aa[IDX3(ORC_EXV, ORC_EXV, 0, nb)], &e);
    }
    {
        orc_entry e = {(double)n_w, (uint32_t)n_w, 0};
        entry_add(&ww[0], &e);
    }

    /* p_tot: FFAvg.cpp:325-337 (aa ff>=1 x ff>=1, aw ff>=1, ww) */
    memset(ptot, 0, nb * sizeof(orc_entry));
    for (int f1 = 1; f1 < ORC_MAX_FF; ++f1) {
        for (int f2 = 1; f2 < ORC_MAX_FF; ++f2)
            for (int b = 0; b < nb; ++b) entry_add(&ptot[b], &aa[IDX3(f1, f2, b, nb)]);
        if (mode == ORC_FFEXPLICIT)
            for (int b = 0; b < nb; ++b) entry_add(&ptot[b], &aw[IDX2(f1, b, nb)]);
    }
    if (mode == ORC_FFAVG)
        for (int f1 = 1; f1 < ORC_MAX_FF; ++f1)
            for (int b = 0; b < nb; ++b) entry_add(&ptot[b], &aw[IDX2(f1, b, nb)]);
    for (int b = 0; b < nb; ++b) entry_add(&ptot[b], &ww[b]);
    int max_bin = max_bin_of(ptot, nb);

    /* Z_exv scaling of the exv rows: FFAvg.cpp:355-364 */
    if (mode == ORC_FFAVG) {
        for (int f = 1; f < ORC_MAX_FF; ++f)
            for (int b = 0; b < max_bin; ++b) entry_scale(&aa[IDX3(f, ORC_EXV, b, nb)], z_exv);
        for (int b = 0; b < max_bin; ++b) entry_scale(&aa[IDX3(ORC_EXV, ORC_EXV, b, nb)], z_exv * z_exv);
        for (int b = 0; b < max_bin; ++b) entry_scale(&aw[IDX2(ORC_EXV, b, nb)], z_exv);
    }
    return max_bin;
}

/* ------------------------------------------------------------------------ */
/* grid excluded volume                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    orc_entry* h;
} simple_priv;

enum { G_EMPTY = 0, G_ACENTER = 1, G_AAREA = 2, G_WCENTER = 4, G_WAREA = 8, G_VOLUME = 16 };

/* constants::radius::get_vdw_radius(form_factor_t) (FormFactorType.h:174-193), vdwTable.h */
static double orc_vdw(int t) {
    switch (t) {
        case 2: return 1.1;
        case 3: case 4: case 5: case 6: return 1.96;
        case 7: case 8: case 9: case 10: return 1.79;
        case 1: case 11: return 1.71;
        case 12: case 13: return 2.06;
        case 14: return 1.88;
        default: return -1;
    }
}

/* One octant of a sphere in 8 angular divisions, reflected to the full sphere
 * in the order +++, -++, +-+, --+, ++-, -+-, +--, ---, components below 1e-5
 * set to 0 and points within 1e-5 of an earlier one dropped
 * (RadialLineGenerator::generate, RadialLineGenerator.cpp:15-63, and
 * RadialHydration::prepare_rotations, RadialHydration.cpp:82-139). Returns
 * the count (26). */
static int sphere_dirs(double dirs[][3], int cap) {
    const double pi = 3.14159265358979323846, ang = 2 * pi / 8;
    int nd = 0;
    for (double theta = 0; theta <= pi * 0.5; theta += ang)
        for (double phi = 0; phi <= pi * 0.5; phi += ang) {
            const double base[3] = {cos(phi) * sin(theta), sin(phi) * sin(theta), cos(theta)};
            for (int r = 0; r < 8; ++r) {
                double p[3] = {(r & 1) ? -base[0] : base[0], (r & 2) ? -base[1] : base[1],
                               (r & 4) ? -base[2] : base[2]};
                for (int c = 0; c < 3; ++c)
                    if (fabs(p[c]) < 1e-5) p[c] = 0;
                int dup = 0;
                for (int e = 0; e < nd && !dup; ++e) {
                    double dx = dirs[e][0] - p[0], dy = dirs[e][1] - p[1], dz = dirs[e][2] - p[2];
                    dup = sqrt(dx * dx + dy * dy + dz * dz) < 1e-5;
                }
                if (!dup) {
                    if (nd >= cap) return -1;
                    memcpy(dirs[nd++], p, sizeof p);
                }
            }
        }
    return nd;
}

/* The bin offsets of the collision radii of collision_check
 * (RawGridWithSurfaceExv.cpp:15-16) at width 1: off[d][r] for direction d at
 * radius r. Returns the count. */
static int radial_offsets(int off[][3][3], int cap) {
    const double radius[3] = {sqrt(1.0) + 1e-3, 2.0, 3.0};
    double dirs[64][3];
    int nd = sphere_dirs(dirs, 64);
    if (nd < 0 || nd > cap) return -1;
    for (int d = 0; d < nd; ++d)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) off[d][r][c] = (int)(round(radius[r] * dirs[d][c]) * 1.0);
    return nd;
}

/* Grid(bodies) (Grid.cpp:34-110, to_bins :414-419) with Grid::add (:300-352)
 * marking atom and water centres, then SphericalExpander<true,false>::
 * expand_volume for the atoms (SphericalExpander.h:29-90); rvol and ra as for
 * grid_exv_core. Returns the state grid (calloc'd) with its minimum, bin
 * counts, the atoms' bins (malloc'd) and Grid::get_volume's count. */
/* lim: NULL for Grid(bodies) (Grid.cpp:34-104), else {xmin, ymin, zmin, xmax,
 * ymax, zmax} for Grid(Limit3D) (Grid.cpp:30, Axis3D.cpp:17: span / width bins) */
static uint8_t* grid_build(const double* atoms, const int32_t* ff, int n, const double* waters, int m, double rvol,
                           double ra, const double* lim, double gmin[3], int bins[3], int** loc_out, long* vol_out) {
    const double width = 1.0;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (atoms[3 * i + k] < mn[k]) mn[k] = atoms[3 * i + k];
            if (atoms[3 * i + k] > mx[k]) mx[k] = atoms[3 * i + k];
        }
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) {
            if (waters[3 * i + k] < mn[k]) mn[k] = waters[3 * i + k];
            if (waters[3 * i + k] > mx[k]) mx[k] = waters[3 * i + k];
        }
    double scaling = 0.25;
    if (mx[0] - mn[0] < 50 || mx[1] - mn[1] < 50 || mx[2] - mn[2] < 50) scaling = 1;
    const double margin = 2.14 + 1.71 + (-0.35) + 3 * width; /* vdw P + vdw O + shell_correction + 3 cells */
    for (int k = 0; k < 3; ++k) {
        double diff = mx[k] - mn[k];
        double expand = 0.5 * diff * scaling > margin ? 0.5 * diff * scaling : margin;
        gmin[k] = floor(mn[k] - expand - width);
        double gmax = ceil(mx[k] + expand + width);
        bins[k] = (int)((gmax - gmin[k]) / width);
        if (lim) {
            gmin[k] = lim[k];
            bins[k] = (int)((lim[3 + k] - lim[k]) / width);
        }
    }
    size_t total = (size_t)bins[0] * bins[1] * bins[2];
    uint8_t* g = (uint8_t*)calloc(total, 1);
#define G(i, j, k) g[((size_t)(i) * bins[1] + (j)) * bins[2] + (k)]
    long vol = 0;
    int* loc = (int*)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    /* Grid::add: atoms then waters (one body) */
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) loc[3 * i + k] = (int)round((atoms[3 * i + k] - gmin[k]) / width);
        uint8_t* s = &G(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]);
        vol += (*s == G_EMPTY || (*s & (G_WAREA | G_WCENTER))); /* is_empty_or_water */
        *s |= G_ACENTER;
    }
    for (int i = 0; i < m; ++i) {
        int l[3];
        for (int k = 0; k < 3; ++k) l[k] = (int)round((waters[3 * i + k] - gmin[k]) / width);
        G(l[0], l[1], l[2]) |= G_WCENTER;
    }
    /* expand_volume: atoms first (waters afterwards only add W_AREA marks) */
    for (int a = 0; a < n; ++a) {
        double rvdw = ra < 0 ? orc_vdw(ff[a]) : ra;
        double rvdw2 = pow(rvdw, 2), rvol2 = pow(rvol, 2);
        double br = ceil((rvdw > rvol ? rvdw : rvol) / width);
        int bx = loc[3 * a], by = loc[3 * a + 1], bz = loc[3 * a + 2];
        int xm = bx - br > 0 ? (int)(bx - br) : 0, xp = bx + br + 1 < bins[0] ? (int)(bx + br + 1) : bins[0];
        int ym = by - br > 0 ? (int)(by - br) : 0, yp = by + br + 1 < bins[1] ? (int)(by + br + 1) : bins[1];
        int zm = bz - br > 0 ? (int)(bz - br) : 0, zp = bz + br + 1 < bins[2] ? (int)(bz + br + 1) : bins[2];
        const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
        for (int i = xm; i < xp; ++i) {
            double x2 = pow(gmin[0] + i * width - ax, 2);
            for (int j = ym; j < yp; ++j) {
                double x2y2 = x2 + pow(gmin[1] + j * width - ay, 2);
                for (int k = zm; k < zp; ++k) {
                    double dist = x2y2 + pow(gmin[2] + k * width - az, 2);
                    uint8_t* s = &G(i, j, k);
                    int fill = dist <= rvdw2;
                    if (fill) {
                        if (!(*s == G_EMPTY || (*s & G_VOLUME))) continue; /* is_empty_or_volume */
                        vol += !(*s & G_VOLUME);
                        *s |= G_AAREA;
                    }
                    if (!fill && dist <= rvol2) {
                        if (*s != G_EMPTY) continue;
                        *s |= G_VOLUME;
                        vol++;
                    }
                }
            }
        }
    }
#undef G
    *loc_out = loc;
    *vol_out = vol;
    return g;
}

static int grid_exv_core(const double* atoms, const int32_t* ff, int n, const double* waters, int m,
                         double* out_xyz, int cap, double* surf_xyz, int* n_surf, double* volume, double rvol,
                         double ra, double thickness) {
    const double width = 1.0;
    double gmin[3];
    int bins[3], *loc = NULL;
    long vol = 0;
    uint8_t* g = grid_build(atoms, ff, n, waters, m, rvol, ra, NULL, gmin, bins, &loc, &vol);
#define G(i, j, k) g[((size_t)(i) * bins[1] + (j)) * bins[2] + (k)]
    /* RawGridExv::create over the atoms' bounding_box_index +- buffer */
    int imin[3] = {1 << 30, 1 << 30, 1 << 30}, imax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int a = 0; a < n; ++a)
        for (int k = 0; k < 3; ++k) {
            if (imin[k] > loc[3 * a + k]) imin[k] = loc[3 * a + k];
            if (imax[k] < loc[3 * a + k]) imax[k] = loc[3 * a + k] + 1;
        }
    if (n == 0)
        for (int k = 0; k < 3; ++k) imin[k] = imax[k] = 0;
    /* max(1, round(max(min_exv_radius, 2) / cell_width)) (RawGridWithSurfaceExv.cpp:107) */
    const int buffer = (int)fmax(1.0, round(fmax(rvol, 2.0) / width));
    int off[64][3][3];
    int n_dirs = surf_xyz ? radial_offsets(off, 64) : 0;
    if (n_dirs < 0) {
        free(g), free(loc);
        return -1;
    }
#define OOB(a, b, c) ((a) < 0 || (a) >= bins[0] || (b) < 0 || (b) >= bins[1] || (c) < 0 || (c) >= bins[2])
#define EMPTY_OR_WATER(a, b, c) (G(a, b, c) == G_EMPTY || (G(a, b, c) & (G_WAREA | G_WCENTER)))
    /* RawGridWithSurfaceExv::create (:196-204): a surface thickness other than
     * one cell takes the helper's non-unity branch, whose first pass only marks
     * the surface dummies (RESERVED_1, here mark bit 1) */
    const int thick = (int)round(thickness / width);
    uint8_t* mark = NULL;
    if (surf_xyz && thick != 1) {
        mark = calloc((size_t)bins[0] * bins[1] * bins[2], 1);
        if (!mark) {
            free(g), free(loc);
            return -1;
        }
    }
#define MARK(i, j, k) mark[((size_t)(i) * bins[1] + (j)) * bins[2] + (k)]
    int count = 0, ns = 0, ok = 1;
    for (int i = imin[0] - buffer > 0 ? imin[0] - buffer : 0; i < (imax[0] + buffer + 1 < bins[0] ? imax[0] + buffer + 1 : bins[0]); ++i)
        for (int j = imin[1] - buffer > 0 ? imin[1] - buffer : 0; j < (imax[1] + buffer + 1 < bins[1] ? imax[1] + buffer + 1 : bins[1]); ++j)
            for (int k = imin[2] - buffer > 0 ? imin[2] - buffer : 0; k < (imax[2] + buffer + 1 < bins[2] ? imax[2] + buffer + 1 : bins[2]); ++k) {
                if (!(G(i, j, k) & (G_VOLUME | G_ACENTER | G_AAREA))) continue;
                int surface = 0;
                if (surf_xyz) {
                    /* collision_check (RawGridWithSurfaceExv.cpp:14-80): per
                     * direction, the first of r1, r2, r3 that is out of
                     * bounds (+7) or occupied (+0, +3, +5) ends it; +7 if all
                     * three are empty or water. Interior iff score >= 42. */
                    int score = 0;
                    for (int d = 0; d < n_dirs; ++d) {
                        static const int gain[3] = {0, 3, 5};
                        int r = 0;
                        for (; r < 3; ++r) {
                            int a = i + off[d][r][0], b = j + off[d][r][1], c = k + off[d][r][2];
                            if (OOB(a, b, c)) {
                                score += 7;
                                break;
                            }
                            if (!EMPTY_OR_WATER(a, b, c)) {
                                score += gain[r];
                                break;
                            }
                        }
                        if (r == 3) score += 7;
                    }
                    surface = score >= 42;
                    if (mark) {
                        if (surface) MARK(i, j, k) |= 1;
                        continue;
                    }
                }
                double* dst = surface ? surf_xyz : out_xyz;
                int* cnt = surface ? &ns : &count;
                if (*cnt < cap) {
                    dst[3 * *cnt] = gmin[0] + i * width;
                    dst[3 * *cnt + 1] = gmin[1] + j * width;
                    dst[3 * *cnt + 2] = gmin[2] + k * width;
                } else {
                    ok = 0;
                }
                ++*cnt;
            }
    if (mark) {
        /* expand = round(thickness / cell_width) * 0.5 as an int (:140-155):
         * every exv-state voxel within squared integer distance expand^2 of a
         * marked voxel gets RESERVED_2 (mark bit 2). The reference indexes the
         * neighbours without a bounds check; here voxels off the grid are
         * skipped (they hold no dummies). */
        const int e = (int)(thick * 0.5), e2 = e * e;
        for (int i = imin[0] - buffer > 0 ? imin[0] - buffer : 0; i < (imax[0] + buffer + 1 < bins[0] ? imax[0] + buffer + 1 : bins[0]); ++i)
            for (int j = imin[1] - buffer > 0 ? imin[1] - buffer : 0; j < (imax[1] + buffer + 1 < bins[1] ? imax[1] + buffer + 1 : bins[1]); ++j)
                for (int k = imin[2] - buffer > 0 ? imin[2] - buffer : 0; k < (imax[2] + buffer + 1 < bins[2] ? imax[2] + buffer + 1 : bins[2]); ++k) {
                    if (!(MARK(i, j, k) & 1)) continue;
                    for (int l = -e; l <= e; ++l)
                        for (int mm = -e; mm <= e; ++mm)
                            for (int nn = -e; nn <= e; ++nn) {
                                const int a = i + l, b = j + mm, c = k + nn;
                                if (OOB(a, b, c)) continue;
                                if ((G(a, b, c) & (G_VOLUME | G_ACENTER | G_AAREA)) && l * l + mm * mm + nn * nn <= e2)
                                    MARK(a, b, c) |= 2;
                            }
                }
        /* collection (:158-180): marked voxels are surface, other exv voxels interior */
        for (int i = imin[0] - buffer > 0 ? imin[0] - buffer : 0; i < (imax[0] + buffer + 1 < bins[0] ? imax[0] + buffer + 1 : bins[0]); ++i)
            for (int j = imin[1] - buffer > 0 ? imin[1] - buffer : 0; j < (imax[1] + buffer + 1 < bins[1] ? imax[1] + buffer + 1 : bins[1]); ++j)
                for (int k = imin[2] - buffer > 0 ? imin[2] - buffer : 0; k < (imax[2] + buffer + 1 < bins[2] ? imax[2] + buffer + 1 : bins[2]); ++k) {
                    int surface;
                    if (MARK(i, j, k)) surface = 1;
                    else if (G(i, j, k) & (G_VOLUME | G_ACENTER | G_AAREA)) surface = 0;
                    else continue;
                    double* dst = surface ? surf_xyz : out_xyz;
                    int* cnt = surface ? &ns : &count;
                    if (*cnt < cap) {
                        dst[3 * *cnt] = gmin[0] + i * width;
                        dst[3 * *cnt + 1] = gmin[1] + j * width;
                        dst[3 * *cnt + 2] = gmin[2] + k * width;
                    } else {
                        ok = 0;
                    }
                    ++*cnt;
                }
        free(mark);
    }
#undef MARK
#undef OOB
#undef EMPTY_OR_WATER
#undef G
    free(g);
    free(loc);
    if (volume) *volume = vol * width * width * width;
    if (n_surf) *n_surf = ns;
    return ok ? count : -1;
}

/* RadialHydration::generate_explicit_hydration (RadialHydration.cpp:40-80)
 * per body on the molecule's grid (GridBasedHydration::hydrate,
 * GridBasedHydration.cpp:36-79, NoCulling), with collision_check (:141-206). */
int orc_hydrate_radial(const double* atoms, const int32_t* ff, int n, const int* body_sizes, int n_bodies,
                       const double* noise, double* out, int cap, int* per_body) {
    orc_hydrate_opts o = {-1, 1.71, -0.35, 0, {0, 0, 0, 0, 0, 0}};
    return orc_hydrate_radial_opts(atoms, ff, n, body_sizes, n_bodies, noise, &o, out, cap, per_body);
}

int orc_hydrate_radial_opts(const double* atoms, const int32_t* ff, int n, const int* body_sizes, int n_bodies,
                            const double* noise, const orc_hydrate_opts* o, double* out, int cap, int* per_body) {
    if (n <= 0) return 0;
    if (o->explicit_grid)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                const int b = (int)round((atoms[3 * i + k] - o->grid_lim[k]) / 1.0);
                if (b < 0 || b >= (int)((o->grid_lim[3 + k] - o->grid_lim[k]) / 1.0)) return -2; /* outside the grid */
            }
    double gmin[3];
    int bins[3], *loc = NULL;
    long vol = 0;
    uint8_t* g = grid_build(atoms, ff, n, NULL, 0, 2.15, o->atomic_radius, o->explicit_grid ? o->grid_lim : NULL, gmin, bins, &loc, &vol);
#define G(i, j, k) g[((size_t)(i) * bins[1] + (j)) * bins[2] + (k)]
#define OOB(a, b, c) ((a) < 0 || (a) >= bins[0] || (b) < 0 || (b) >= bins[1] || (c) < 0 || (c) >= bins[2])
#define ONLY_EMPTY_OR_VOLUME(v) ((v) == G_EMPTY || (v) == G_VOLUME)
    double dirs[64][3];
    const int nd = sphere_dirs(dirs, 64);
    /* Grid::get_hydration_radius (vdw O by default) + settings::hydrate::shell_correction */
    const double rh = o->hydration_radius + o->shell_correction;
    int p1[64][3], p3[64][3], p5[64][3];
    for (int d = 0; d < nd; ++d)
        for (int c = 0; c < 3; ++c) {
            p1[d][c] = (int)(round(rh * dirs[d][c]) / 1.0);
            p3[d][c] = (int)(round(3 * rh * dirs[d][c]) / 1.0);
            p5[d][c] = (int)(round(5 * rh * dirs[d][c]) / 1.0);
        }
    int count = 0, ok = 1, one_body = n;
    if (n_bodies <= 0) body_sizes = &one_body, n_bodies = 1;
    int a = 0;
    for (int b = 0; b < n_bodies; ++b) {
        int placed = 0;
        for (int e = 0; e < body_sizes[b]; ++e, ++a) {
            const double reff = (o->atomic_radius < 0 ? orc_vdw(ff[a]) : o->atomic_radius) + rh;
            for (int d = 0; d < nd; ++d) {
                double w[3];
                int wb[3];
                for (int c = 0; c < 3; ++c) {
                    double nz = noise ? noise[((size_t)a * nd + d) * 3 + c] : 0.0;
                    w[c] = atoms[3 * a + c] + dirs[d][c] * reff + nz;
                    wb[c] = (int)round((w[c] - gmin[c]) / 1.0);
                }
                if (OOB(wb[0], wb[1], wb[2])) continue;
                if (!ONLY_EMPTY_OR_VOLUME(G(wb[0], wb[1], wb[2]))) continue;
                /* collision_check */
                int score = 0, inside = 0, reject = 0;
                for (int d2 = 0; d2 < nd && !reject; ++d2) {
                    int x = wb[0] + p1[d2][0], y = wb[1] + p1[d2][1], z = wb[2] + p1[d2][2];
                    x = x < 0 ? 0 : x >= bins[0] ? bins[0] - 1 : x;
                    y = y < 0 ? 0 : y >= bins[1] ? bins[1] - 1 : y;
                    z = z < 0 ? 0 : z >= bins[2] ? bins[2] - 1 : z;
                    if (!ONLY_EMPTY_OR_VOLUME(G(x, y, z))) {
                        if (++inside > 2) reject = 1;
                        continue;
                    }
                    x = wb[0] + p3[d2][0], y = wb[1] + p3[d2][1], z = wb[2] + p3[d2][2];
                    if (OOB(x, y, z)) {
                        score += 2;
                        continue;
                    }
                    if (!ONLY_EMPTY_OR_VOLUME(G(x, y, z))) {
                        score -= 2;
                        continue;
                    }
                    score += 2;
                    x = wb[0] + p5[d2][0], y = wb[1] + p5[d2][1], z = wb[2] + p5[d2][2];
                    if (OOB(x, y, z)) {
                        score += 1;
                        continue;
                    }
                    if (!ONLY_EMPTY_OR_VOLUME(G(x, y, z))) score -= 1;
                }
                if (reject || score <= 0.25 * (3.0 * nd)) continue;
                /* Grid::add(water, true): centre, then W_AREA on empty voxels
                 * within the hydration radius of the water (SphericalExpander.h:92-140) */
                G(wb[0], wb[1], wb[2]) |= G_WCENTER;
                const double rw2 = pow(o->hydration_radius, 2);
                const int br = (int)ceil(o->hydration_radius / 1.0);
                for (int i = wb[0] - br > 0 ? wb[0] - br : 0; i < (wb[0] + br + 1 < bins[0] ? wb[0] + br + 1 : bins[0]); ++i) {
                    double x2 = pow(gmin[0] + i - w[0], 2);
                    for (int j = wb[1] - br > 0 ? wb[1] - br : 0; j < (wb[1] + br + 1 < bins[1] ? wb[1] + br + 1 : bins[1]); ++j) {
                        double x2y2 = x2 + pow(gmin[1] + j - w[1], 2);
                        for (int k = wb[2] - br > 0 ? wb[2] - br : 0; k < (wb[2] + br + 1 < bins[2] ? wb[2] + br + 1 : bins[2]); ++k) {
                            double dist = x2y2 + pow(gmin[2] + k - w[2], 2);
                            if (dist <= rw2 && G(i, j, k) == G_EMPTY) G(i, j, k) |= G_WAREA;
                        }
                    }
                }
                if (count < cap) memcpy(out + 3 * (size_t)count, w, sizeof w);
                else ok = 0;
                ++count, ++placed;
            }
        }
        if (per_body) per_body[b] = placed;
    }
#undef G
#undef OOB
#undef ONLY_EMPTY_OR_VOLUME
    free(g), free(loc);
    return ok ? count : -1;
}

int orc_grid_exv(const double* atoms, const int32_t* ff, int n, const double* waters, int m, double* out_xyz,
                 int cap, double* volume) {
    return grid_exv_core(atoms, ff, n, waters, m, out_xyz, cap, NULL, NULL, volume, 2.15, -1, 1.0);
}

int orc_grid_exv_surface(const double* atoms, const int32_t* ff, int n, const double* waters, int m,
                         double* interior_xyz, double* surface_xyz, int cap, int* n_surface, double* volume,
                         double rvol, double ra, double surface_thickness) {
    return grid_exv_core(atoms, ff, n, waters, m, interior_xyz, cap, surface_xyz, n_surface, volume, rvol, ra,
                         surface_thickness);
}

/* grid pair loops: xx self (factor 2, 1D), ax atoms x dummies (factor 1,
 * [ff_atom]), wx waters x dummies (factor 1, 1D) —
 * HistogramManagerMTFFGrid.cpp:68-137, TemplateHelperGrid.h */
typedef struct {
    const float* rows;
    int nr;
    const float* cols;
    int nc;
    float inv_w;
    int nb;
    int self;   /* xx */
    int typed;  /* ax: 2D by row ff */
    int factor; /* template N: 2 for dummy-dummy, 1 for atom/water-dummy */
} grid_args;

static int grid_row(job_ctx* c, void* pv, int i) {
    grid_args* A = (grid_args*)c->args;
    orc_entry* h = ((simple_priv*)pv)->h;
    const float* pi = A->rows + 4 * (size_t)i;
    const int m = A->self ? A->nr : A->nc;
    const float* cols = A->self ? A->rows : A->cols;
    const int j0 = A->self ? i + 1 : 0;
    const int tail = A->self ? m - ((m - i - 1) % 4) : m - (m % 4);
    const int factor = A->factor;
    const size_t off = A->typed ? (size_t)ff_of(pi) * A->nb : 0;
    for (int j = j0; j < m; ++j) {
        float d;
        int32_t bin;
        if (j < tail) rule_simd(pi, cols + 4 * (size_t)j, A->inv_w, &d, &bin);
        else rule_scalar(pi, cols + 4 * (size_t)j, A->inv_w, &d, &bin);
        if (bin < 0 || bin >= A->nb) return -2;
        inc_n(&h[off + bin], factor, d);
    }
    return 0;
}

static int run_grid_f(const float* rows, int nr, const float* cols, int nc, int self, int typed, int factor,
                      float inv_w, int nb, int nthreads, orc_entry* out) {
    const size_t n = (size_t)(typed ? ORC_MAX_FF : 1) * nb;
    simple_priv* privs = (simple_priv*)calloc(nthreads, sizeof(simple_priv));
    void** pp = (void**)malloc(sizeof(void*) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        privs[t].h = (orc_entry*)calloc(n, sizeof(orc_entry));
        pp[t] = &privs[t];
    }
    grid_args A = {rows, nr, cols, nc, inv_w, nb, self, typed, factor};
    int err = run_rows(nr, nthreads, grid_row, &A, pp);
    memset(out, 0, n * sizeof(orc_entry));
    for (int t = 0; t < nthreads; ++t) {
        for (size_t b = 0; b < n; ++b) entry_add(&out[b], &privs[t].h[b]);
        free(privs[t].h);
    }
    free(privs);
    free(pp);
    return err;
}

static int run_grid(const float* rows, int nr, const float* cols, int nc, int self, int typed, float inv_w, int nb,
                    int nthreads, orc_entry* out) {
    return run_grid_f(rows, nr, cols, nc, self, typed, self ? 2 : 1, inv_w, nb, nthreads, out);
}

static void weighted_axis_n(const orc_entry* p, int n_entries, int n_axis, double bin_width, double* d_out) {
    double width = (n_axis * bin_width) / n_axis;
    for (int i = 0; i < n_entries; ++i) {
        double dv = i * width;
        double bc = p[i].bin_center;
        d_out[i] = ((double)(!bc) * dv + bc) / (double)((!p[i].count) + p[i].count);
    }
}

int orc_hist_ffgrid(const float* atoms, int n_a, const float* waters, int n_w, const float* dummies, int n_x,
                    double volume, float inv_w, int nb, int nthreads, orc_entry* aa, orc_entry* aw, orc_entry* ww,
                    orc_entry* ptot, orc_entry* ptot_ax, orc_entry* pxx) {
    if (nthreads < 1) nthreads = 1;
    /* HistogramManagerMTFFAvg<true, vbw>::calculate_all with Z from the grid volume */
    double z = n_a == 0 ? 0 : volume * 0.334 / n_a;
    int base = orc_hist_ff(ORC_FFAVG, atoms, n_a, waters, n_w, inv_w, nb, z, nthreads, aa, aw, ww, ptot);
    if (base < 0) return base;
    orc_entry* xx = (orc_entry*)calloc(nb, sizeof(orc_entry));
    orc_entry* ax = (orc_entry*)calloc((size_t)ORC_MAX_FF * nb, sizeof(orc_entry));
    orc_entry* wx = (orc_entry*)calloc(nb, sizeof(orc_entry));
    int err = 0;
    err |= run_grid(dummies, n_x, NULL, 0, 1, 0, inv_w, nb, nthreads, xx);
    err |= run_grid(atoms, n_a, dummies, n_x, 0, 1, inv_w, nb, nthreads, ax);
    err |= run_grid(waters, n_w, dummies, n_x, 0, 0, inv_w, nb, nthreads, wx);
    if (err) {
        free(xx), free(ax), free(wx);
        return -2;
    }
    {
        orc_entry e = {(double)n_x, (uint32_t)n_x, 0};
        entry_add(&xx[0], &e);
    }
    int mbx = 10;
    for (int i = nb - 1; i >= 10; --i)
        if (xx[i].value != 0 || wx[i].value != 0) {
            mbx = i + 1;
            break;
        }
    /* weighted counts of the FFAvg result at the first default q (FFAvgBase.cpp:77-109) */
    double* dax = (double*)calloc(base, sizeof(double));
    orc_weighted_axis(ptot, base, 1.0 / inv_w, dax);
    double* wc = (double*)calloc(base, sizeof(double));
    const double q0 = 1e-4;
    for (int f1 = 1; f1 < ORC_MAX_FF; ++f1)
        for (int f2 = 1; f2 < ORC_MAX_FF; ++f2) {
            double w = orc_ff_atomic(f1, q0) * orc_ff_atomic(f2, q0);
            for (int b = 0; b < base; ++b) wc[b] += aa[IDX3(f1, f2, b, nb)].value * w;
        }
    for (int f = 1; f < ORC_MAX_FF; ++f) {
        double w = 2 * 1.0 * (orc_ff_atomic(f, q0) * orc_ff_atomic(ORC_WATER, q0));
        for (int b = 0; b < base; ++b) wc[b] += aw[IDX2(f, b, nb)].value * w;
    }
    {
        double w = 1.0 * 1.0 * (orc_ff_atomic(ORC_WATER, q0) * orc_ff_atomic(ORC_WATER, q0));
        for (int b = 0; b < base; ++b) wc[b] += ww[b].value * w;
    }
    int max_bin = mbx > base ? mbx : base;
    memset(ptot, 0, (size_t)nb * sizeof(orc_entry));
    for (int b = 0; b < base; ++b) {
        ptot[b].value = wc[b];
        ptot[b].count = 0;
        ptot[b].bin_center = dax[b];
    }
    /* FFAvg rows beyond its own size are zero */
    for (int f1 = 0; f1 < ORC_MAX_FF; ++f1) {
        for (int f2 = 0; f2 < ORC_MAX_FF; ++f2)
            for (int b = base; b < nb; ++b) memset(&aa[IDX3(f1, f2, b, nb)], 0, sizeof(orc_entry));
        for (int b = base; b < nb; ++b) memset(&aw[IDX2(f1, b, nb)], 0, sizeof(orc_entry));
    }
    for (int b = base; b < nb; ++b) memset(&ww[b], 0, sizeof(orc_entry));
    memset(ptot_ax, 0, (size_t)nb * sizeof(orc_entry));
    for (int b = 0; b < max_bin; ++b) entry_add(&ptot_ax[b], &wx[b]);
    for (int f = 0; f < ORC_MAX_FF; ++f)
        for (int b = 0; b < max_bin; ++b) entry_add(&ptot_ax[b], &ax[IDX2(f, b, nb)]);
    for (int f = 0; f < ORC_MAX_FF; ++f)
        for (int b = 0; b < max_bin; ++b) aa[IDX3(f, ORC_EXV, b, nb)] = ax[IDX2(f, b, nb)];
    for (int b = 0; b < max_bin; ++b) aw[IDX2(ORC_EXV, b, nb)] = wx[b];
    for (int b = 0; b < max_bin; ++b) aa[IDX3(ORC_EXV, ORC_EXV, b, nb)] = xx[b];
    memcpy(pxx, xx, (size_t)nb * sizeof(orc_entry));
    free(xx), free(ax), free(wx), free(dax), free(wc);
    return max_bin;
}

int orc_hist_ffgrid_surface(const float* atoms, int n_a, const float* waters, int n_w, const float* xi, int n_i,
                            const float* xs, int n_s, double volume, float inv_w, int nb, int nthreads,
                            orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot, orc_entry* ptot_ax,
                            orc_entry* pxx, orc_entry* xxi, orc_entry* xxs, orc_entry* xxc, orc_entry* wxi,
                            orc_entry* wxs, orc_entry* axi, orc_entry* axs) {
    if (nthreads < 1) nthreads = 1;
    /* the FFAvg base with Z from the grid volume (HistogramManagerMTFFGridSurface.cpp:28) */
    double z = n_a == 0 ? 0 : volume * 0.334 / n_a;
    int base = orc_hist_ff(ORC_FFAVG, atoms, n_a, waters, n_w, inv_w, nb, z, nthreads, aa, aw, ww, ptot);
    if (base < 0) return base;
    int err = 0;
    /* :52-196: xx interior/surface self, interior x surface (factor 2), ax and wx against each */
    err |= run_grid_f(xi, n_i, NULL, 0, 1, 0, 2, inv_w, nb, nthreads, xxi);
    err |= run_grid_f(xs, n_s, NULL, 0, 1, 0, 2, inv_w, nb, nthreads, xxs);
    err |= run_grid_f(xi, n_i, xs, n_s, 0, 0, 2, inv_w, nb, nthreads, xxc);
    err |= run_grid_f(atoms, n_a, xi, n_i, 0, 1, 1, inv_w, nb, nthreads, axi);
    err |= run_grid_f(atoms, n_a, xs, n_s, 0, 1, 1, inv_w, nb, nthreads, axs);
    err |= run_grid_f(waters, n_w, xi, n_i, 0, 0, 1, inv_w, nb, nthreads, wxi);
    err |= run_grid_f(waters, n_w, xs, n_s, 0, 0, 1, inv_w, nb, nthreads, wxs);
    if (err) return -2;
    {
        orc_entry e = {(double)n_i, (uint32_t)n_i, 0};
        entry_add(&xxi[0], &e);
        orc_entry f = {(double)n_s, (uint32_t)n_s, 0};
        entry_add(&xxs[0], &f);
    }
    int mbx = 10;
    for (int i = nb - 1; i >= 10; --i)
        if (xxs[i].value != 0 || xxi[i].value != 0 || wxs[i].value != 0 || wxi[i].value != 0) {
            mbx = i + 1;
            break;
        }
    /* p_tot: the FFAvg weighted counts at the first default q, centres = FFAvg d axis */
    double* dax = (double*)calloc(base, sizeof(double));
    orc_weighted_axis(ptot, base, 1.0 / inv_w, dax);
    double* wc = (double*)calloc(base, sizeof(double));
    const double q0 = 1e-4;
    for (int f1 = 1; f1 < ORC_MAX_FF; ++f1)
        for (int f2 = 1; f2 < ORC_MAX_FF; ++f2) {
            double w = orc_ff_atomic(f1, q0) * orc_ff_atomic(f2, q0);
            for (int b = 0; b < base; ++b) wc[b] += aa[IDX3(f1, f2, b, nb)].value * w;
        }
    for (int f = 1; f < ORC_MAX_FF; ++f) {
        double w = 2 * 1.0 * (orc_ff_atomic(f, q0) * orc_ff_atomic(ORC_WATER, q0));
        for (int b = 0; b < base; ++b) wc[b] += aw[IDX2(f, b, nb)].value * w;
    }
    {
        double w = 1.0 * 1.0 * (orc_ff_atomic(ORC_WATER, q0) * orc_ff_atomic(ORC_WATER, q0));
        for (int b = 0; b < base; ++b) wc[b] += ww[b].value * w;
    }
    int max_bin = mbx > base ? mbx : base;
    memset(ptot, 0, (size_t)nb * sizeof(orc_entry));
    for (int b = 0; b < base; ++b) {
        ptot[b].value = wc[b];
        ptot[b].count = 0;
        ptot[b].bin_center = dax[b];
    }
    for (int f1 = 0; f1 < ORC_MAX_FF; ++f1) {
        for (int f2 = 0; f2 < ORC_MAX_FF; ++f2)
            for (int b = base; b < nb; ++b) memset(&aa[IDX3(f1, f2, b, nb)], 0, sizeof(orc_entry));
        for (int b = base; b < nb; ++b) memset(&aw[IDX2(f1, b, nb)], 0, sizeof(orc_entry));
    }
    for (int b = base; b < nb; ++b) memset(&ww[b], 0, sizeof(orc_entry));
    /* :233-247: the weighted totals that give the ax and xx distance axes */
    memset(ptot_ax, 0, (size_t)nb * sizeof(orc_entry));
    memset(pxx, 0, (size_t)nb * sizeof(orc_entry));
    for (int b = 0; b < max_bin; ++b) {
        entry_add(&ptot_ax[b], &wxi[b]);
        entry_add(&ptot_ax[b], &wxs[b]);
    }
    for (int f = 0; f < ORC_MAX_FF; ++f)
        for (int b = 0; b < max_bin; ++b) {
            entry_add(&ptot_ax[b], &axi[IDX2(f, b, nb)]);
            entry_add(&ptot_ax[b], &axs[IDX2(f, b, nb)]);
        }
    for (int b = 0; b < max_bin; ++b) {
        entry_add(&pxx[b], &xxi[b]);
        entry_add(&pxx[b], &xxs[b]);
        entry_add(&pxx[b], &xxc[b]);
    }
    /* :248-256 and CompositeDistanceHistogramFFGridSurface::initialize (:88-102):
     * the exv rows are cleared, then hold the dummy profiles at unit scaling */
    for (int f = 0; f < ORC_MAX_FF; ++f)
        for (int b = 0; b < nb; ++b) memset(&aa[IDX3(f, ORC_EXV, b, nb)], 0, sizeof(orc_entry));
    for (int b = 0; b < nb; ++b) memset(&aw[IDX2(ORC_EXV, b, nb)], 0, sizeof(orc_entry));
    for (int f = 1; f < ORC_MAX_FF; ++f)
        for (int b = 0; b < max_bin; ++b) {
            entry_add(&aa[IDX3(f, ORC_EXV, b, nb)], &axi[IDX2(f, b, nb)]);
            entry_add(&aa[IDX3(f, ORC_EXV, b, nb)], &axs[IDX2(f, b, nb)]);
        }
    for (int b = 0; b < max_bin; ++b) {
        aa[IDX3(ORC_EXV, ORC_EXV, b, nb)] = pxx[b];
        entry_add(&aw[IDX2(ORC_EXV, b, nb)], &wxi[b]);
        entry_add(&aw[IDX2(ORC_EXV, b, nb)], &wxs[b]);
    }
    free(dax), free(wc);
    return max_bin;
}

/* weighted axis of a histogram that is n_axis bins long (first n_entries) */
void orc_weighted_axis_n(const orc_entry* p, int n_entries, int n_axis, double bin_width, double* d_out) {
    weighted_axis_n(p, n_entries, n_axis, bin_width, d_out);
}

/* ------------------------------------------------------------------------ */
/* raw pair histogram                                                       */
/* ------------------------------------------------------------------------ */

int orc_pair_hist(const float* a, int n, const float* b, int m, float inv_w, int nb, int factor, orc_entry* out) {
    memset(out, 0, (size_t)nb * sizeof(orc_entry));
    if (!b) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                float d;
                int32_t bin;
                orc_bin_pair(a + 4 * (size_t)i, a + 4 * (size_t)j, inv_w, orc_rule_self(n, i, j), &d, &bin);
                if (bin < 0 || bin >= nb) return -2;
                inc_n(&out[bin], factor, d);
            }
    } else {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                float d;
                int32_t bin;
                orc_bin_pair(a + 4 * (size_t)i, b + 4 * (size_t)j, inv_w, orc_rule_cross(m, j), &d, &bin);
                if (bin < 0 || bin >= nb) return -2;
                inc_n(&out[bin], factor, d);
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* simple (XYZW) manager                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float* rows;
    int nr;
    const float* cols;
    int nc;
    float inv_w;
    int nb;
    int self;
    int factor; /* template N of add<N>: 2 * scaling */
} simple_args;


static int simple_row(job_ctx* c, void* pv, int i) {
    simple_args* A = (simple_args*)c->args;
    orc_entry* h = ((simple_priv*)pv)->h;
    const float* pi = A->rows + 4 * (size_t)i;
    int m = A->self ? A->nr : A->nc;
    const float* cols = A->self ? A->rows : A->cols;
    int j0 = A->self ? i + 1 : 0;
    int tail = A->self ? m - ((m - i - 1) % 4) : m - (m % 4);
    for (int j = j0; j < m; ++j) {
        const float* pj = cols + 4 * (size_t)j;
        float d;
        int32_t bin;
        if (j < tail) rule_simd(pi, pj, A->inv_w, &d, &bin);
        else rule_scalar(pi, pj, A->inv_w, &d, &bin);
        if (bin < 0 || bin >= A->nb) return -2;
        /* TemplateHelperSimple.h:15-61: add<2 * scaling>(bin, d, w_i*w_j) (float product) */
        add_n(&h[bin], A->factor, d, pi[3] * pj[3]);
    }
    return 0;
}

/* SimpleCalculator::enqueue_calculate_self<scaling> / _cross<scaling>
 * (SimpleCalculator.h:94-211): rows over data_2, columns over data_1 */
static int run_simple_scaled(const float* rows, int nr, const float* cols, int nc, int self, int scaling,
                             float inv_w, int nb, int nthreads, orc_entry* out) {
    simple_priv* privs = (simple_priv*)calloc(nthreads, sizeof(simple_priv));
    void** pp = (void**)malloc(sizeof(void*) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        privs[t].h = (orc_entry*)calloc(nb, sizeof(orc_entry));
        pp[t] = &privs[t];
    }
    simple_args A = {rows, nr, cols, nc, inv_w, nb, self, 2 * scaling};
    int err = run_rows(nr, nthreads, simple_row, &A, pp);
    memset(out, 0, (size_t)nb * sizeof(orc_entry));
    for (int t = 0; t < nthreads; ++t) {
        for (int b = 0; b < nb; ++b) entry_add(&out[b], &privs[t].h[b]);
        free(privs[t].h);
    }
    free(privs);
    free(pp);
    if (self) {
        /* SimpleCalculator.h:144-160: bin 0 += WeightedEntry(t, (unsigned) t, 0),
         * t = scaling * sum (w*w) (float products summed in double) */
        double tw = 0;
        for (int i = 0; i < nr; ++i) {
            const float w = rows[4 * (size_t)i + 3];
            tw += (double)(w * w);
        }
        tw = scaling * tw;
        orc_entry e = {tw, (uint32_t)tw, 0};
        entry_add(&out[0], &e);
    }
    return err;
}

static int run_simple(const float* rows, int nr, const float* cols, int nc, int self, float inv_w, int nb,
                      int nthreads, orc_entry* out) {
    return run_simple_scaled(rows, nr, cols, nc, self, 1, inv_w, nb, nthreads, out);
}

int orc_hist_simple(const float* atoms, int n_a, const float* waters, int n_w, float inv_w, int nb, int nthreads,
                    orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot) {
    if (nthreads < 1) nthreads = 1;
    int err = 0;
    err |= run_simple(atoms, n_a, NULL, 0, 1, inv_w, nb, nthreads, aa);
    err |= run_simple(waters, n_w, NULL, 0, 1, inv_w, nb, nthreads, ww);
    /* enqueue_calculate_cross(data_a, data_w): rows over data_2 = waters,
     * columns over data_1 = atoms (SimpleCalculator.h:188-211) */
    err |= run_simple(waters, n_w, atoms, n_a, 0, inv_w, nb, nthreads, aw);
    if (err) return -2;
    for (int b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(&ptot[b], &ww[b]);
        entry_add(&ptot[b], &aw[b]);
    }
    return max_bin_of(ptot, nb);
}

void orc_rigid_transform(const double* xyz, const float* w, int n, const double* T, float* out) {
    for (int i = 0; i < n; ++i) {
        const double x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
        for (int k = 0; k < 3; ++k) {
            const double* r = T + 3 * k;
            const double v = ((r[0] * x + r[1] * y) + r[2] * z) + T[9 + k];
            out[4 * (size_t)i + k] = (float)v;
        }
        out[4 * (size_t)i + 3] = w[i];
    }
}

int orc_hist_partial(const float* self_bodies, const float* bodies, const int* sizes, int n_bodies,
                     const float* waters, int n_w, float inv_w, int nb, int nthreads,
                     orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot) {
    if (nthreads < 1) nthreads = 1;
    size_t* off = (size_t*)malloc(sizeof(size_t) * (n_bodies + 1));
    off[0] = 0;
    for (int b = 0; b < n_bodies; ++b) off[b + 1] = off[b] + (size_t)sizes[b];
    orc_entry* tmp = (orc_entry*)malloc(sizeof(orc_entry) * (size_t)nb);
    memset(aa, 0, sizeof(orc_entry) * (size_t)nb);
    memset(aw, 0, sizeof(orc_entry) * (size_t)nb);
    int err = 0;
    for (int i = 0; i < n_bodies; ++i) {
        /* self-correlation cached from the coordinates it was computed on */
        err |= run_simple(self_bodies + 4 * off[i], sizes[i], NULL, 0, 1, inv_w, nb, nthreads, tmp);
        for (int k = 0; k < nb; ++k) entry_add(&aa[k], &tmp[k]);
        for (int j = 0; j < i; ++j) {
            err |= run_simple(bodies + 4 * off[j], sizes[j], bodies + 4 * off[i], sizes[i], 0, inv_w, nb, nthreads,
                              tmp);
            for (int k = 0; k < nb; ++k) entry_add(&aa[k], &tmp[k]);
        }
        err |= run_simple(waters, n_w, bodies + 4 * off[i], sizes[i], 0, inv_w, nb, nthreads, tmp);
        for (int k = 0; k < nb; ++k) entry_add(&aw[k], &tmp[k]);
    }
    err |= run_simple(waters, n_w, NULL, 0, 1, inv_w, nb, nthreads, ww);
    free(tmp);
    free(off);
    if (err) return -2;
    for (int b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(&ptot[b], &aw[b]);
        entry_add(&ptot[b], &ww[b]);
    }
    return max_bin_of(ptot, nb);
}

void orc_affine_float(const float* pts, int n, const double* T, float* out) {
    for (int i = 0; i < n; ++i) {
        const double x = pts[4 * (size_t)i], y = pts[4 * (size_t)i + 1], z = pts[4 * (size_t)i + 2];
        for (int k = 0; k < 3; ++k) {
            const double* r = T + 3 * k;
            out[4 * (size_t)i + k] = (float)(((r[0] * x + r[1] * y) + r[2] * z) + T[9 + k]);
        }
        out[4 * (size_t)i + 3] = pts[4 * (size_t)i + 3];
    }
}

int orc_hist_symmetry(const float* atoms, const int* sizes, int n_bodies, const int* sym_per_body,
                      const int* sym_reps, const double* transforms, const int* sched_len, const int* schedule,
                      const float* waters, int n_w, float inv_w, int nb, int nthreads, orc_entry* aa,
                      orc_entry* aw, orc_entry* ww, orc_entry* ptot) {
    if (nthreads < 1) nthreads = 1;
    int B = n_bodies, n_sym = 0;
    for (int b = 0; b < B; ++b) n_sym += sym_per_body[b];
    /* body offsets, first symmetry of each body, copy offsets per symmetry */
    size_t* off = (size_t*)calloc(B + 1, sizeof(size_t));
    int* sym0 = (int*)calloc(B + 1, sizeof(int));
    size_t* t0 = (size_t*)calloc(n_sym + 1, sizeof(size_t));
    size_t* e0 = (size_t*)calloc(n_sym + 1, sizeof(size_t));
    for (int b = 0; b < B; ++b) {
        off[b + 1] = off[b] + (size_t)sizes[b];
        sym0[b + 1] = sym0[b] + sym_per_body[b];
    }
    for (int k = 0; k < n_sym; ++k) {
        t0[k + 1] = t0[k] + (size_t)sym_reps[k];
        e0[k + 1] = e0[k] + (size_t)sched_len[k];
    }
    /* copies: generate_transformed_data (SymmetryHelpers.cpp:16-43) */
    float** copy = (float**)calloc(t0[n_sym] + 1, sizeof(float*));
    for (int b = 0; b < B; ++b)
        for (int k = sym0[b]; k < sym0[b + 1]; ++k)
            for (int r = 0; r < sym_reps[k]; ++r) {
                size_t c = t0[k] + (size_t)r;
                copy[c] = (float*)malloc(sizeof(float) * 4 * ((size_t)sizes[b] + 1));
                orc_affine_float(atoms + 4 * off[b], sizes[b], transforms + 12 * c, copy[c]);
            }
#define BODY(b) (atoms + 4 * off[b])
#define AT(b, k, rep) ((rep) == 0 ? BODY(b) : copy[t0[k] + (size_t)(rep)-1])
    orc_entry* tmp = (orc_entry*)malloc(sizeof(orc_entry) * (size_t)nb);
    memset(aa, 0, sizeof(orc_entry) * (size_t)nb);
    memset(aw, 0, sizeof(orc_entry) * (size_t)nb);
    memset(ww, 0, sizeof(orc_entry) * (size_t)nb);
    int err = 0;
#define ACC(dst, expr)                                                  \
    do {                                                                \
        err |= (expr);                                                  \
        for (int q_ = 0; q_ < nb; ++q_) entry_add(&(dst)[q_], &tmp[q_]); \
    } while (0)
    /* SymmetryManagerMT.cpp:57-132; cross(d1, d2) = rows d2, columns d1 */
    for (int b1 = 0; b1 < B; ++b1) {
        int total = 0;
        for (int k = sym0[b1]; k < sym0[b1 + 1]; ++k) total += sym_reps[k];
        ACC(aa, run_simple_scaled(BODY(b1), sizes[b1], NULL, 0, 1, 1 + total, inv_w, nb, nthreads, tmp));
        if (n_w > 0) ACC(aw, run_simple_scaled(BODY(b1), sizes[b1], waters, n_w, 0, 1, inv_w, nb, nthreads, tmp));
        for (int k1 = sym0[b1]; k1 < sym0[b1 + 1]; ++k1) {
            for (size_t e = e0[k1]; e < e0[k1 + 1]; ++e) {
                const int ra = schedule[3 * e], rb = schedule[3 * e + 1], sc = schedule[3 * e + 2];
                ACC(aa, run_simple_scaled(AT(b1, k1, rb), sizes[b1], AT(b1, k1, ra), sizes[b1], 0, sc, inv_w, nb,
                                          nthreads, tmp));
            }
            for (int r1 = 1; r1 <= sym_reps[k1]; ++r1) {
                const float* c1 = AT(b1, k1, r1);
                if (n_w > 0) ACC(aw, run_simple_scaled(c1, sizes[b1], waters, n_w, 0, 1, inv_w, nb, nthreads, tmp));
                for (int b2 = b1 + 1; b2 < B; ++b2) {
                    ACC(aa, run_simple_scaled(c1, sizes[b1], BODY(b2), sizes[b2], 0, 1, inv_w, nb, nthreads, tmp));
                    for (int k2 = sym0[b2]; k2 < sym0[b2 + 1]; ++k2)
                        for (int r2 = 1; r2 <= sym_reps[k2]; ++r2)
                            ACC(aa, run_simple_scaled(AT(b2, k2, r2), sizes[b2], c1, sizes[b1], 0, 1, inv_w, nb,
                                                      nthreads, tmp));
                }
                for (int k2 = k1 + 1; k2 < sym0[b1 + 1]; ++k2)
                    for (int r2 = 1; r2 <= sym_reps[k2]; ++r2)
                        ACC(aa, run_simple_scaled(AT(b1, k2, r2), sizes[b1], c1, sizes[b1], 0, 1, inv_w, nb,
                                                  nthreads, tmp));
            }
        }
        for (int b2 = b1 + 1; b2 < B; ++b2) {
            ACC(aa, run_simple_scaled(BODY(b2), sizes[b2], BODY(b1), sizes[b1], 0, 1, inv_w, nb, nthreads, tmp));
            for (int k2 = sym0[b2]; k2 < sym0[b2 + 1]; ++k2)
                for (int r2 = 1; r2 <= sym_reps[k2]; ++r2)
                    ACC(aa, run_simple_scaled(AT(b2, k2, r2), sizes[b2], BODY(b1), sizes[b1], 0, 1, inv_w, nb,
                                              nthreads, tmp));
        }
    }
    if (n_w > 0) ACC(ww, run_simple_scaled(waters, n_w, NULL, 0, 1, 1, inv_w, nb, nthreads, tmp));
#undef ACC
#undef AT
#undef BODY
    for (size_t c = 0; c < t0[n_sym]; ++c) free(copy[c]);
    free(copy);
    free(tmp);
    free(off);
    free(sym0);
    free(t0);
    free(e0);
    if (err) return -2;
    for (int b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(&ptot[b], &ww[b]);
        entry_add(&ptot[b], &aw[b]);
    }
    return max_bin_of(ptot, nb);
}

int orc_spline(const double* x, const double* y, int n, const double* z, int m, double* out) {
    if (n < 4) return -1;
    double* h = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* D = (double*)malloc(sizeof(double) * n);
    double* Q = (double*)malloc(sizeof(double) * n);
    double* B = (double*)malloc(sizeof(double) * n);
    double* b = (double*)malloc(sizeof(double) * n);
    double* c = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n - 1; ++i) {
        h[i] = x[i + 1] - x[i];
        p[i] = (y[i + 1] - y[i]) / h[i];
    }
    D[0] = 2;
    Q[0] = 1;
    B[0] = 3 * p[0];
    D[n - 1] = 2;
    B[n - 1] = 3 * p[n - 2];
    for (int i = 0; i < n - 2; ++i) {
        D[i + 1] = 2 * h[i] / h[i + 1] + 2;
        Q[i + 1] = h[i] / h[i + 1];
        B[i + 1] = 3 * (p[i] + p[i + 1] * h[i] / h[i + 1]);
    }
    for (int i = 1; i < n; ++i) {
        D[i] -= Q[i - 1] / D[i - 1];
        B[i] -= B[i - 1] / D[i - 1];
    }
    b[n - 1] = B[n - 1] / D[n - 1];
    for (int i = n - 2; i >= 0; --i) b[i] = (B[i] - Q[i] * b[i + 1]) / D[i];
    for (int i = 0; i < n - 1; ++i) {
        c[i] = (-2 * b[i] - b[i + 1] + 3 * p[i]) / h[i];
        d[i] = (b[i] + b[i + 1] - 2 * p[i]) / pow(h[i], 2);
    }
    for (int k = 0; k < m; ++k) {
        /* CubicSpline::search: binary search for the last knot <= z (0 below
         * the range), the last interval at or past the end */
        int lo = 0, hi = n;
        while (lo < hi) {
            int mid = lo + (hi - lo) / 2;
            if (z[k] < x[mid]) hi = mid;
            else lo = mid + 1;
        }
        int i = lo - 1 < 0 ? 0 : lo - 1;
        if (i == n - 1) --i;
        double t = z[k] - x[i];
        out[k] = y[i] + b[i] * t + c[i] * pow(t, 2) + d[i] * pow(t, 3);
    }
    free(h), free(p), free(D), free(Q), free(B), free(b), free(c), free(d);
    return 0;
}

int orc_hist_symmetry_partial(const float* atoms, const float* self_atoms, const int* self_scale, const int* sizes,
                              int n_bodies, const int* sym_per_body, const int* sym_reps, const double* transforms,
                              const int* sched_len, const int* schedule, const float* waters, int n_w, float inv_w,
                              int nb, int nthreads, orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot) {
    if (nthreads < 1) nthreads = 1;
    int B = n_bodies, n_sym = 0;
    for (int b = 0; b < B; ++b) n_sym += sym_per_body[b];
    size_t* off = (size_t*)calloc(B + 1, sizeof(size_t));
    int* sym0 = (int*)calloc(B + 1, sizeof(int));
    size_t* t0 = (size_t*)calloc(n_sym + 1, sizeof(size_t));
    size_t* e0 = (size_t*)calloc(n_sym + 1, sizeof(size_t));
    for (int b = 0; b < B; ++b) {
        off[b + 1] = off[b] + (size_t)sizes[b];
        sym0[b + 1] = sym0[b] + sym_per_body[b];
    }
    for (int k = 0; k < n_sym; ++k) {
        t0[k + 1] = t0[k] + (size_t)sym_reps[k];
        e0[k + 1] = e0[k] + (size_t)sched_len[k];
    }
    float** copy = (float**)calloc(t0[n_sym] + 1, sizeof(float*));
    for (int b = 0; b < B; ++b)
        for (int k = sym0[b]; k < sym0[b + 1]; ++k)
            for (int r = 0; r < sym_reps[k]; ++r) {
                size_t c = t0[k] + (size_t)r;
                copy[c] = (float*)malloc(sizeof(float) * 4 * ((size_t)sizes[b] + 1));
                orc_affine_float(atoms + 4 * off[b], sizes[b], transforms + 12 * c, copy[c]);
            }
#define BODY(b) (atoms + 4 * off[b])
/* symmetry slot s of body b (0 = the body, s >= 1 = symmetry s-1), repetition r */
#define COORD(b, s, r) ((s) == 0 ? BODY(b) : copy[t0[sym0[b] + (s)-1] + (size_t)(r)])
#define REPS(b, s) ((s) == 0 ? 1 : sym_reps[sym0[b] + (s)-1])
    orc_entry* tmp = (orc_entry*)malloc(sizeof(orc_entry) * (size_t)nb);
    memset(aa, 0, sizeof(orc_entry) * (size_t)nb);
    memset(aw, 0, sizeof(orc_entry) * (size_t)nb);
    memset(ww, 0, sizeof(orc_entry) * (size_t)nb);
    int err = 0;
#define ACC(dst, expr)                                                  \
    do {                                                                \
        err |= (expr);                                                  \
        for (int q_ = 0; q_ < nb; ++q_) entry_add(&(dst)[q_], &tmp[q_]); \
    } while (0)
    for (int b1 = 0; b1 < B; ++b1) {
        const int ns1 = sym_per_body[b1];
        /* calc_aa_self: the self-term coordinates and scale */
        ACC(aa, run_simple_scaled(self_atoms + 4 * off[b1], sizes[b1], NULL, 0, 1, self_scale[b1], inv_w, nb,
                                  nthreads, tmp));
        /* calc_aa(b1, s1, b2, s2), b2 < b1: rows over b2, columns over b1 (:620-678) */
        for (int b2 = 0; b2 < b1; ++b2)
            for (int s1 = 0; s1 <= ns1; ++s1)
                for (int s2 = 0; s2 <= sym_per_body[b2]; ++s2)
                    for (int r1 = 0; r1 < REPS(b1, s1); ++r1)
                        for (int r2 = 0; r2 < REPS(b2, s2); ++r2)
                            ACC(aa, run_simple_scaled(COORD(b2, s2, r2), sizes[b2], COORD(b1, s1, r1), sizes[b1], 0,
                                                      1, inv_w, nb, nthreads, tmp));
        for (int s1 = 1; s1 <= ns1; ++s1) {
            /* calc_aa(b1, s1, b1, 0): the pair schedule (:604-618) */
            const int k1 = sym0[b1] + s1 - 1;
            for (size_t e = e0[k1]; e < e0[k1 + 1]; ++e) {
                const int ra = schedule[3 * e], rb = schedule[3 * e + 1], sc = schedule[3 * e + 2];
                const float* A = ra == 0 ? BODY(b1) : copy[t0[k1] + (size_t)ra - 1];
                const float* Bp = rb == 0 ? BODY(b1) : copy[t0[k1] + (size_t)rb - 1];
                ACC(aa, run_simple_scaled(Bp, sizes[b1], A, sizes[b1], 0, sc, inv_w, nb, nthreads, tmp));
            }
            /* calc_aa(b1, s1, b1, s2), s2 < s1: rows over symmetry s2 */
            for (int s2 = 1; s2 < s1; ++s2)
                for (int r1 = 0; r1 < REPS(b1, s1); ++r1)
                    for (int r2 = 0; r2 < REPS(b1, s2); ++r2)
                        ACC(aa, run_simple_scaled(COORD(b1, s2, r2), sizes[b1], COORD(b1, s1, r1), sizes[b1], 0, 1,
                                                  inv_w, nb, nthreads, tmp));
        }
        /* calc_aw(b1, s1): rows over the waters (:685-706) */
        if (n_w > 0)
            for (int s1 = 0; s1 <= ns1; ++s1)
                for (int r1 = 0; r1 < REPS(b1, s1); ++r1)
                    ACC(aw, run_simple_scaled(waters, n_w, COORD(b1, s1, r1), sizes[b1], 0, 1, inv_w, nb, nthreads,
                                              tmp));
    }
    if (n_w > 0) ACC(ww, run_simple_scaled(waters, n_w, NULL, 0, 1, 1, inv_w, nb, nthreads, tmp));
#undef ACC
#undef REPS
#undef COORD
#undef BODY
    for (size_t c = 0; c < t0[n_sym]; ++c) free(copy[c]);
    free(copy);
    free(tmp);
    free(off);
    free(sym0);
    free(t0);
    free(e0);
    if (err) return -2;
    for (int b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(&ptot[b], &ww[b]);
        entry_add(&ptot[b], &aw[b]);
    }
    return max_bin_of(ptot, nb);
}

/* ------------------------------------------------------------------------ */
/* weighted axis, sinc table, contraction                                   */
/* ------------------------------------------------------------------------ */

void orc_weighted_axis(const orc_entry* p, int n, double bin_width, double* d_out) {
    /* Axis(0, n*bin_width, n).as_vector(): i*width with width = (n*bw)/n */
    double width = (n * bin_width) / n;
    for (int i = 0; i < n; ++i) {
        double dv = 0 + i * width;
        double bc = p[i].bin_center;
        d_out[i] = ((double)(!bc) * dv + bc) / (double)((!p[i].count) + p[i].count);
    }
    if (n > 0) d_out[0] = 0;
}

void orc_sinc_table(const double* q, int nq, const double* d, int nd, double* out) {
    const double tol = 1e-3, inv6 = 1. / 6, inv120 = 1. / 120;
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nd; ++j) {
            double qd = q[i] * d[j];
            double v;
            if (qd < tol) {
                double qd2 = qd * qd;
                v = 1 - qd2 * inv6 + qd2 * qd2 * inv120;
            } else {
                v = sin(qd) / qd;
            }
            out[(size_t)i * nd + j] = v;
        }
}

void orc_contract(const double* H, int rows, int nd, const double* S, int nq, double* out) {
    for (int r = 0; r < rows; ++r)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            const double* h = H + (size_t)r * nd;
            const double* t = S + (size_t)i * nd;
            for (int b = 0; b < nd; ++b) s += h[b] * t[b];
            out[(size_t)r * nq + i] = s;
        }
}

void orc_contract_ld(const double* H, int rows, int nd, const double* S, int nq, double* out) {
    for (int r = 0; r < rows; ++r)
        for (int i = 0; i < nq; ++i) {
            long double s = 0.0L;
            const double* h = H + (size_t)r * nd;
            const double* t = S + (size_t)i * nd;
            for (int b = 0; b < nd; ++b) s += (long double)h[b] * (long double)t[b];
            out[(size_t)r * nq + i] = (double)s;
        }
}

/* ------------------------------------------------------------------------ */
/* form factors                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    double a[5], b[5], c;
} ff5;

#define PI_ 3.14159265358979323846
/* FormFactorTable.h:15-23: b given in s units, converted by 1/(16 pi^2) */
#define S2Q (1. / (4 * 4 * PI_ * PI_))

/* coefficient sets: FormFactorTable.h:28-120 */
static const ff5 FF_H = {{0.489918, 0.262003, 0.196767, 0.049879, 0}, {20.6593, 7.74039, 49.5519, 2.20159, 0}, 0.001305};
static const ff5 FF_O = {{2.960427, 2.508818, 0.637853, 0.722838, 1.142756}, {14.182259, 5.936858, 0.112726, 34.958481, 0.390240}, 0.027014};
static const ff5 FF_N = {{11.893780, 3.277479, 1.858092, 0.858927, 0.912985}, {0.000158, 10.232723, 30.344690, 0.656065, 0.217287}, -11.804902};
static const ff5 FF_S = {{6.362157, 5.154568, 1.473732, 1.635073, 1.209372}, {1.514347, 22.092528, 0.061373, 55.445176, 0.646925}, 0.154722};
static const ff5 FF_C = {{2.657506, 1.078079, 1.490909, -4.241070, 0.713791}, {14.780758, 0.776775, 42.086843, -0.000294, 0.239535}, 4.297983};
static const ff5 FF_AR = {{7.188004, 6.638454, 0.454180, 1.929593, 1.523654}, {0.956221, 15.339877, 15.339862, 39.043824, 0.062409}, 0.265954};
static const ff5 FF_CH = {{2.909530, 0.485267, 1.516151, 0.206905, 1.541626}, {13.933084, 23.221524, 41.990403, 4.974183, 0.679266}, 0.337670};
static const ff5 FF_CH2 = {{3.275723, 0.870037, 1.534606, 0.395078, 1.544562}, {13.408502, 23.785175, 41.922444, 5.019072, 0.724439}, 0.377096};
static const ff5 FF_CH3 = {{3.681341, 1.228691, 1.549320, 0.574033, 1.554377}, {13.026207, 24.131974, 41.869426, 4.984373, 0.765769}, 0.409294};
static const ff5 FF_OH = {{0.456221, 3.219608, 0.812773, 2.666928, 1.380927}, {21.503498, 13.397134, 34.547137, 5.826620, 0.412902}, 0.463202};
static const ff5 FF_NH = {{1.650531, 0.429639, 2.144736, 1.851894, 1.408921}, {10.603730, 6.987283, 29.939901, 10.573859, 0.611678}, 0.510589};
static const ff5 FF_NH2 = {{1.904157, 1.942536, 2.435585, 0.730512, 1.379728}, {10.803702, 10.792421, 29.610479, 6.847755, 0.709687}, 0.603738};
static const ff5 FF_NH3 = {{1.882162, 1.933200, 2.465843, 0.927311, 1.190889}, {10.975157, 10.956008, 29.208572, 6.663555, 0.843650}, 0.597322};
static const ff5 FF_SH = {{0.570042, 6.337416, 1.641643, 5.398549, 1.527982}, {11.447986, 1.197657, 55.401032, 22.420955, 2.356552}, 1.523944};

static double eval5(const ff5* f, double q, int convert) {
    double s = 0;
    for (int i = 0; i < 5; ++i) {
        double b = convert ? f->b[i] * S2Q : f->b[i];
        s += f->a[i] * exp(-b * q * q);
    }
    return s + f->c;
}

double orc_ff_atomic(int type, double q) {
    switch (type) {
        case 0: { /* excluded_volume: a={1}, b={r_avg^2/2} in q units (FormFactorTable.h) */
            double r = 1.62;
            return exp(-(r * r / 2) * q * q);
        }
        case 1: return eval5(&FF_OH, q, 1);  /* WATER == OH -> OH_alc */
        case 2: return eval5(&FF_H, q, 1);
        case 3: return eval5(&FF_C, q, 1);
        case 4: return eval5(&FF_CH, q, 1);  /* CH_sp3 */
        case 5: return eval5(&FF_CH2, q, 1); /* CH2_sp3 */
        case 6: return eval5(&FF_CH3, q, 1); /* CH3_sp3 */
        case 7: return eval5(&FF_N, q, 1);
        case 8: return eval5(&FF_NH, q, 1);
        case 9: return eval5(&FF_NH2, q, 1);
        case 10: return eval5(&FF_NH3, q, 1); /* NH3_plus */
        case 11: return eval5(&FF_O, q, 1);
        case 12: return eval5(&FF_S, q, 1);
        case 13: return eval5(&FF_SH, q, 1);
        case 14: return eval5(&FF_AR, q, 1); /* OTHER: argon */
        default: return NAN;
    }
}

/* MinimumFluctuation_implicit_H (ExvTable.h), OTHER -> Ar = 4pi/3 1.88^3 */
static double exv_volume(int type) {
    switch (type) {
        case 1: return 20.911;  /* WATER == OH */
        case 2: return 0;       /* H */
        case 3: return 12.352;  /* C */
        case 4: return 11.640;  /* CH */
        case 5: return 34.583;  /* CH2 */
        case 6: return 41.851;  /* CH3 */
        case 7: return 0.027;   /* N */
        case 8: return 2.181;   /* NH */
        case 9: return 20.562;  /* NH2 */
        case 10: return 20.722; /* NH3 */
        case 11: return 14.238; /* O */
        case 12: return 15.413; /* S */
        case 13: return 28.529; /* SH */
        case 14: return 4 * PI_ / 3 * 1.88 * 1.88 * 1.88;
        default: return NAN;
    }
}

double orc_ff_exv(int type, double q) {
    double v = exv_volume(type);
    double exponent = pow(v, 2. / 3) / (4 * PI_);
    double q0 = v * 0.334;
    return q0 * exp(-exponent * q * q);
}

/* ------------------------------------------------------------------------ */
/* intensity profiles                                                       */
/* ------------------------------------------------------------------------ */

#define S3(f1, f2, k) sinqd_aa[((size_t)(f1) * ORC_MAX_FF + (f2)) * nq + (k)]
#define S2(f, k) sinqd_aw[(size_t)(f) * nq + (k)]

void orc_profiles(int mode, const double* q, int nq, const double* sinqd_aa, const double* sinqd_aw,
                  const double* sinqd_ww, orc_params p, double* out) {
    double* aa = out;
    double* ax = out + nq;
    double* aw = out + 2 * nq;
    double* xx = out + 3 * nq;
    double* wx = out + 4 * nq;
    double* ww = out + 5 * nq;
    for (int k = 0; k < 6 * nq; ++k) out[k] = 0;
    const double rm2 = 1.62 * 1.62 / 4;
    for (int k = 0; k < nq; ++k) {
        double qk = q[k];
        double cx = mode != ORC_FFEXPLICIT ? p.cx : pow(p.cx, 3) * exp(-rm2 * (pow(p.cx, 2) - 1) * qk * qk);
        double fw = orc_ff_atomic(ORC_WATER, qk);
        /* exv form factor: FFAvg's raw EXCLUDED_VOLUME entry, or FFGrid's
         * ExvFormFactor(V) (CompositeDistanceHistogramFFGrid.cpp:36-75,
         * ExvFormFactor.h:23-34): q0 = 0.334 V, exponent V^(2/3) / (4 pi) */
        double gv = p.grid_v > 0 ? p.grid_v : 1.0;
        double fx = mode == ORC_FFGRID ? 0.334 * gv * exp(-(pow(gv, 2. / 3) / (4 * PI_)) * qk * qk)
                                       : orc_ff_atomic(ORC_EXV, qk);
        for (int f1 = 1; f1 < ORC_MAX_FF; ++f1) {
            double a1 = orc_ff_atomic(f1, qk);
            for (int f2 = 1; f2 < ORC_MAX_FF; ++f2) {
                double a2 = orc_ff_atomic(f2, qk);
                double s = S3(f1, f2, k);
                aa[k] += s * a1 * a2;
                if (mode == ORC_FFEXPLICIT) {
                    /* raw_cross_table(i,j) = atomic(i)*exv(j); raw_exv_table = exv*exv */
                    double cross = a1 * orc_ff_exv(f2, qk) + a2 * orc_ff_exv(f1, qk);
                    ax[k] += p.crho * cx * s * cross;
                    xx[k] += pow(cx * p.crho, 2) * s * orc_ff_exv(f1, qk) * orc_ff_exv(f2, qk);
                }
            }
            if (mode != ORC_FFEXPLICIT) {
                ax[k] += 2 * p.crho * cx * S3(f1, ORC_EXV, k) * a1 * fx;
                aw[k] += 2 * p.cw * S2(f1, k) * a1 * fw;
            } else {
                aw[k] += 2 * p.cw * S2(f1, k) * a1 * fw;
                wx[k] += p.crho * cx * p.cw * S2(f1, k) * (fw * orc_ff_exv(f1, qk) + a1 * orc_ff_exv(ORC_WATER, qk));
            }
        }
        if (mode != ORC_FFEXPLICIT) {
            xx[k] += pow(cx * p.crho, 2) * S3(ORC_EXV, ORC_EXV, k) * fx * fx;
            wx[k] += 2 * p.crho * cx * p.cw * S2(ORC_EXV, k) * fx * fw;
        }
        ww[k] += p.cw * p.cw * sinqd_ww[k] * fw * fw;

        /* Debye-Waller: FFAvgBase.cpp:310-374 */
        if (p.dw_atomic != 0 || p.dw_exv != 0) {
            double ba = exp(-qk * qk * p.dw_atomic * p.dw_atomic * 0.5);
            double bx = exp(-qk * qk * p.dw_exv * p.dw_exv * 0.5);
            aa[k] *= ba * ba;
            xx[k] *= bx * bx;
            ax[k] *= ba * bx;
            aw[k] *= ba;
            wx[k] *= bx;
        }
    }
}

void orc_total(const double* prof, int nq, double* out) {
    for (int k = 0; k < nq; ++k)
        out[k] = prof[k] - prof[nq + k] + prof[2 * nq + k] + prof[3 * nq + k] - prof[4 * nq + k] + prof[5 * nq + k];
}

/* ------------------------------------------------------------------------ */
/* exact Debye and LSQ                                                      */
/* ------------------------------------------------------------------------ */

void orc_exact_debye(const float* p, int n, const double* q, int nq, double* out) {
    for (int k = 0; k < nq; ++k) {
        double sum = 0;
        for (int i = 0; i < n; ++i) {
            const float* pi = p + 4 * (size_t)i;
            for (int j = i + 1; j < n; ++j) {
                float d;
                int32_t bin;
                orc_bin_pair(pi, p + 4 * (size_t)j, 1.0f, orc_rule_self(n, i, j), &d, &bin);
                double qr = q[k] * d;
                float w2 = 2 * (pi[3] * p[4 * (size_t)j + 3]);
                sum += qr < 1e-9 ? w2 : w2 * sin(qr) / qr;
            }
            sum += pow(pi[3], 2);
        }
        out[k] = sum * exp(-q[k] * q[k]);
    }
}

void orc_lsq(const double* x, const double* y, const double* err, int n, double* out) {
    double S = 0, Sx = 0, Sy = 0, Sxx = 0, Sxy = 0;
    for (int i = 0; i < n; ++i) {
        double is = err ? 1. / err[i] : 1.;
        double w = is * is;
        S += w;
        Sx += x[i] * w;
        Sy += y[i] * w;
        Sxx += x[i] * x[i] * w;
        Sxy += x[i] * y[i] * w;
    }
    double delta = S * Sxx - Sx * Sx;
    out[0] = (S * Sxy - Sx * Sy) / delta;
    out[1] = (Sxx * Sy - Sx * Sxy) / delta;
    out[2] = S / delta;
    out[3] = Sxx / delta;
    /* Fitter::chi2 (Fitter.cpp:10-13) over LinearLeastSquares::get_residuals
     * (LinearLeastSquares.cpp:71-78): sum of ((y - (a x + b)) / err)^2 */
    double c = 0;
    for (int i = 0; i < n; ++i) {
        double is = err ? 1. / err[i] : 1.;
        double r = (y[i] - (out[0] * x[i] + out[1])) * is;
        c += r * r;
    }
    out[4] = c;
}
