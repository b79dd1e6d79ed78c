// A caller on the reference's side of the C ABI, compiled against
// include/ausaxs_hip.h alone and linked with libausaxs_hip.so.
//
// It does what the histogram-manager binding does (a HistogramManager
// subclass in place of HistogramManagerMTFFExplicit::calculate_all,
// HistogramManagerMTFFExplicit.cpp:216-436, and of HistogramManagerMTFFAvg,
// HistogramManagerMTFFAvg.cpp:194-371): pack the atoms as CompactCoordinatesFF
// (16 bytes per atom), run the manager, and copy every distribution into the
// reference's layouts — Container3D [ff1][ff2][bin] for p_aa, Container2D
// [ff][bin] for p_aw, Container1D for p_ww, and WeightedEntry {value, count,
// bin_center = sum of distances} for p_tot — with plain vectors standing in
// for the Distribution classes. Then I(q) through ausaxs_histogram_debye.
//
// Output: one JSON object on stdout (numbers printed with 17 significant
// digits, so a reader can compare them exactly). Exit status 0 on success,
// 3 when a library call fails (the message from ausaxs_get_last_error is in
// the JSON), 2 on a usage error.
#include "ausaxs_hip.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

constexpr int kFF = 15;  // form_factor::get_count_without_excluded_volume() + excluded volume

int fail(const char* what) {
    char msg[512] = {0};
    ausaxs_get_last_error(msg, sizeof msg);
    std::string esc;
    for (const char* p = msg; *p; ++p) esc += (*p == '"' || *p == '\\') ? '\'' : *p;
    std::printf("{\"error\": \"%s: %s\"}\n", what, esc.c_str());
    return 3;
}

// WeightedDistribution1D's entries
struct Entry {
    double value = 0;
    uint32_t count = 0;
    double bin_center = 0;
};

struct Distributions {
    int max_bin = 0;
    std::vector<double> aa, aw, ww;  // Distribution3D / 2D / 1D contents
    std::vector<Entry> p_tot;
    std::vector<double> d;  // WeightedDistribution1D::get_weighted_axis with d[0] = 0
};

bool copy_out(const ausaxs_histogram* h, Distributions& out) {
    int mode = 0;
    if (ausaxs_histogram_info(h, &mode, &out.max_bin)) return false;
    const size_t mb = (size_t)out.max_bin;
    out.aa.assign(kFF * kFF * mb, 0.0);
    out.aw.assign(kFF * mb, 0.0);
    out.ww.assign(mb, 0.0);
    out.d.assign(mb, 0.0);
    std::vector<double> v(mb), c(mb);
    std::vector<uint32_t> n(mb);
    if (ausaxs_histogram_get(h, AUSAXS_AA, out.aa.data(), nullptr, nullptr)) return false;
    if (ausaxs_histogram_get(h, AUSAXS_AW, out.aw.data(), nullptr, nullptr)) return false;
    if (ausaxs_histogram_get(h, AUSAXS_WW, out.ww.data(), nullptr, nullptr)) return false;
    if (ausaxs_histogram_get(h, AUSAXS_PTOT, v.data(), n.data(), c.data())) return false;
    if (ausaxs_histogram_d_axis(h, out.d.data())) return false;
    out.p_tot.resize(mb);
    for (size_t k = 0; k < mb; ++k) out.p_tot[k] = Entry{v[k], n[k], c[k]};
    return true;
}

double sum(const std::vector<double>& x) {
    double s = 0;
    for (double v : x) s += v;
    return s;
}

void print_array(const char* key, const std::vector<double>& x) {
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < x.size(); ++i) std::printf("%s%.17g", i ? ", " : "", x[i]);
    std::printf("]");
}

void print_distributions(const char* key, const Distributions& d, const std::vector<double>& I) {
    double pv = 0, pc = 0;
    unsigned long long pn = 0;
    for (const Entry& e : d.p_tot) pv += e.value, pn += e.count, pc += e.bin_center;
    std::printf("\"%s\": {\"max_bin\": %d, \"aa_sum\": %.17g, \"aw_sum\": %.17g, \"ww_sum\": %.17g, "
                "\"ptot_value_sum\": %.17g, \"ptot_count_sum\": %llu, \"ptot_center_sum\": %.17g, "
                "\"d_sum\": %.17g, ",
                key, d.max_bin, sum(d.aa), sum(d.aw), sum(d.ww), pv, pn, pc, sum(d.d));
    print_array("I", I);
    std::printf("}");
}

}  // namespace

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 1500;
    if (n < 1 || n > 1000000) {
        std::fprintf(stderr, "usage: %s [atoms in 1..1000000]\n", argv[0]);
        return 2;
    }
    const int m = n / 4 > 0 ? n / 4 : 1;
    const double ax = 18, ay = 16, az = 14;

    // CompactCoordinatesFF {x, y, z, form factor as int32 bits}
    std::vector<float> atoms(4 * (size_t)n), waters(4 * (size_t)m);
    if (ausaxs_synth_globule(7, n, ax, ay, az, 2.0, atoms.data())) return fail("synth_globule");
    if (ausaxs_synth_shell(8, m, ax, ay, az, 3.0, waters.data())) return fail("synth_shell");

    ausaxs_settings s;
    ausaxs_default_settings(&s);

    int nq = 0;
    std::vector<double> q(200);  // the default axis has at most 200 points
    if (ausaxs_default_q_axis(1e-4, 0.5, q.data(), &nq)) return fail("default_q_axis");
    q.resize(nq);
    std::vector<double> I(nq);

    // FF-explicit manager: one-shot create + calculate + free
    ausaxs_histogram* h = nullptr;
    if (ausaxs_hist_calculate(AUSAXS_MODE_FFEXPLICIT, atoms.data(), n, waters.data(), m, &s, 0.0, &h))
        return fail("hist_calculate");
    Distributions explicit_ff;
    if (!copy_out(h, explicit_ff)) return fail("copy FF-explicit");
    if (ausaxs_histogram_debye(h, q.data(), nq, I.data(), nullptr)) return fail("debye FF-explicit");
    ausaxs_histogram_free(h);
    const std::vector<double> I_explicit = I;

    // FF-average manager from FP64 coordinates, with its own Z_exv = V_grid * 0.334 / N
    std::vector<double> axyz(3 * (size_t)n), wxyz(3 * (size_t)m);
    std::vector<int32_t> ff(n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) axyz[3 * i + k] = atoms[4 * i + k];
        std::int32_t t;
        static_assert(sizeof t == sizeof atoms[0], "int32 form factor in a float slot");
        __builtin_memcpy(&t, &atoms[4 * i + 3], sizeof t);
        ff[i] = t;
    }
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) wxyz[3 * i + k] = waters[4 * i + k];
    ausaxs_structure* st = nullptr;
    double z_exv = 0;
    if (ausaxs_structure_create_ffavg(axyz.data(), ff.data(), n, wxyz.data(), m, &s, &z_exv, &st))
        return fail("structure_create_ffavg");
    h = nullptr;
    if (ausaxs_structure_calculate(st, &h)) return fail("structure_calculate FF-avg");
    Distributions avg_ff;
    if (!copy_out(h, avg_ff)) return fail("copy FF-avg");
    if (ausaxs_histogram_debye(h, q.data(), nq, I.data(), nullptr)) return fail("debye FF-avg");
    ausaxs_histogram_free(h);
    ausaxs_structure_free(st);

    std::printf("{\"n_atoms\": %d, \"n_waters\": %d, \"nq\": %d, ", n, m, nq);
    print_distributions("ffexplicit", explicit_ff, I_explicit);
    std::printf(", ");
    print_distributions("ffavg", avg_ff, I);
    std::printf(", \"z_exv\": %.17g}\n", z_exv);
    return 0;
}
