/*
 * ausaxs_oracle — CPU restatement of AUSAXS's Debye hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it. The product
 * path (ausaxs_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the AUSAXS 1.2.9 checkout). Parity pins are listed per function; a function
 * without one is marked "parity unpinned".
 */
#ifndef AUSAXS_ORACLE_H
#define AUSAXS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_FF 15          /* settings/FormFactorSettings.h:12 max_ff_types   */
#define ORC_EXV 0              /* form_factor/FormFactorType.h:14-33 EXCLUDED_VOLUME */
#define ORC_WATER 1            /* ... WATER == OH                                  */

/* WeightedEntry{double value; unsigned count; double bin_center}
 * distribution/detail/WeightedEntry.h:16-50 */
typedef struct {
    double value;
    uint32_t count;
    double bin_center;
} orc_entry;

enum { ORC_FFAVG = 0, ORC_FFEXPLICIT = 1, ORC_FFGRID = 2 };

/* ---- binning rules (CompactCoordinatesXYZFF.h) -------------------------- */
/* rule 0: SIMD body (AVX2 :514-545, SSE :397-420): d2=(dx²+dy²)+dz², sqrt_rn,
 *         x=d*inv_w, round-half-even.
 * rule 1: scalar tail (evaluate_scalar :313-329 built with -ffast-math):
 *         d2=fma(dz,dz,fma(dx,dx,dy*dy)), sqrt_rn, trunc(x+0.49999997f). */
void orc_bin_pair(const float* a, const float* b, float inv_w, int rule, float* d_out, int32_t* bin_out);

/* Which rule the reference loop uses for a pair (Appendix of the loop
 * geometry in HistogramManagerMTFFAvg.cpp:214-280): self loops j>i use the
 * scalar rule for the last (n-i-1) mod 4 columns; cross loops for the last
 * m mod 4 columns. */
int orc_rule_self(int n, int i, int j);
int orc_rule_cross(int m, int j);

/* ---- FF managers ---------------------------------------------------------
 * HistogramManagerMTFFAvg::calculate_all (HistogramManagerMTFFAvg.cpp:194-371)
 * HistogramManagerMTFFExplicit::calculate_all (...FFExplicit.cpp:216-436)
 * atoms/waters: n x 4 floats {x,y,z, ff as int32 bits}; ff already remapped.
 * aa: [15][15][bin_count], aw: [15][bin_count], ww/ptot: [bin_count].
 * z_exv: FFAvg only (V_grid*0.334/N); ignored for FFExplicit.
 * nthreads: worker threads (800-row tasks, thread-private histograms, merge,
 * GeneralSettings.cpp:50, ThreadLocalWrapper.h:89-105).
 * Returns max_bin (>= 10) or a negative status: -2 bin out of range, -3 bad ff. */
int orc_hist_ff(int mode, const float* atoms, int n_a, const float* waters, int n_w,
                float inv_w, int bin_count, double z_exv, int nthreads,
                orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot);

/* Grid excluded-volume dummies, restating the reference's voxel-state
 * procedure literally: Grid(bodies) (Grid.cpp:34-110, to_bins :414-419),
 * Grid::add (:300-352), SphericalExpander<true,false>::expand_volume for the
 * atoms (SphericalExpander.h:29-90), GridObj state predicates
 * (GridObj.cpp:16-43) and RawGridExv::create (RawGridExv.cpp:10-59).
 * atoms/waters: n x 3 FP64; out_xyz: cap x 3. Returns the dummy count
 * (or -1 if cap is too small); *volume = Grid::get_volume(). */
int orc_grid_exv(const double* atoms, const int32_t* ff, int n, const double* waters, int m, double* out_xyz,
                 int cap, double* volume);

/* RadialHydration with NoCulling (RadialHydration.cpp:40-206,
 * GridBasedHydration.cpp:36-79): the molecule's grid from the atoms, expanded,
 * then per body per atom per direction a candidate at vdw + 1.36 A (+ noise
 * [n][26][3], NULL for none), accepted when its voxel is empty or volume only
 * and collision_check passes, each acceptance marking the grid. out: cap x 3.
 * Returns the water count (-1 if cap is too small); per_body (optional). */
int orc_hydrate_radial(const double* atoms, const int32_t* ff, int n, const int* body_sizes, int n_bodies,
                       const double* noise, double* out, int cap, int* per_body);
/* The same with the knobs of the reference's own hydration tests
 * (tests/feature/grid/grid.cpp:707-748): GridDebug's atomic and hydration
 * radius overrides, settings::hydrate::shell_correction and a grid built from
 * explicit limits (Grid(Limit3D), Grid.cpp:30). Defaults: -1 (each atom's
 * vdW radius), 1.71, -0.35, 0 (Grid(bodies)). */
typedef struct {
    double atomic_radius;
    double hydration_radius;
    double shell_correction;
    int explicit_grid;
    double grid_lim[6]; /* xmin, ymin, zmin, xmax, ymax, zmax */
} orc_hydrate_opts;
int orc_hydrate_radial_opts(const double* atoms, const int32_t* ff, int n, const int* body_sizes, int n_bodies,
                            const double* noise, const orc_hydrate_opts* o, double* out, int cap, int* per_body);

/* RawGridWithSurfaceExv::create with surface detection at unit cell width
 * (RawGridWithSurfaceExv.cpp:82-127): the dummies of orc_grid_exv split by
 * collision_check (:14-80) over the radial lines of RadialLineGenerator
 * (RadialLineGenerator.cpp:15-63). interior_xyz / surface_xyz: cap x 3 each.
 * rvol: settings::grid::min_exv_radius (2.15 by default); ra < 0: each
 * atom's van der Waals radius, else ra for every atom (GridDebug of the
 * reference's grid tests). surface_thickness: settings::grid::exv::
 * surface_thickness (A, cell width 1); other than one cell, the non-unity
 * branch (RawGridWithSurfaceExv.cpp:137-190) grows the surface by the
 * dummies within int(round(thickness) * 0.5) cells of it. Returns the
 * interior count (or -1 if cap is too small); *n_surface. */
int orc_grid_exv_surface(const double* atoms, const int32_t* ff, int n, const double* waters, int m,
                         double* interior_xyz, double* surface_xyz, int cap, int* n_surface, double* volume,
                         double rvol, double ra, double surface_thickness);

/* HistogramManagerMTFFGrid::calculate_all (HistogramManagerMTFFGrid.cpp:43-234)
 * on top of the FFAvg manager. atoms/waters: n x 4 floats {x,y,z,ff} as for
 * orc_hist_ff; dummies: nx x 4 floats (integer lattice points); volume: grid
 * volume. Outputs (each sized as in orc_hist_ff, bin_count wide): aa, aw, ww,
 * ptot (FFGrid p_tot: value = weighted counts, bin_center = FFAvg d axis),
 * ptot_ax, pxx. Returns max_bin or a negative status. */
int orc_hist_ffgrid(const float* atoms, int n_a, const float* waters, int n_w, const float* dummies, int n_x,
                    double volume, float inv_w, int bin_count, int nthreads, orc_entry* aa, orc_entry* aw,
                    orc_entry* ww, orc_entry* ptot, orc_entry* ptot_ax, orc_entry* pxx);

/* weighted axis of a histogram n_axis bins long, first n_entries values, no
 * forced zero at bin 0 (the FFGrid ax/xx axes, CompositeDistanceHistogramFFGrid.cpp:25-27) */
void orc_weighted_axis_n(const orc_entry* p, int n_entries, int n_axis, double bin_width, double* d_out);

/* Raw pair histogram of one pair set without manager conventions:
 * self (b == NULL) or cross (rows a, cols b); factor = reference template N.
 * out[bin_count] entries. Returns 0 or -2 on a bin >= bin_count. */
int orc_pair_hist(const float* a, int n, const float* b, int m, float inv_w, int bin_count,
                  int factor, orc_entry* out);

/* HistogramManagerMT::calculate_all with SimpleCalculator (XYZW weights):
 * HistogramManagerMT.cpp:23-75, SimpleCalculator.h:94-284, TemplateHelperSimple.h:15-61.
 * atoms/waters: n x 4 floats {x,y,z,w}. aa/aw/ww/ptot: [bin_count].
 * Returns max_bin or negative status. */
int orc_hist_simple(const float* atoms, int n_a, const float* waters, int n_w,
                    float inv_w, int bin_count, int nthreads,
                    orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot);

/* Rigid-body transform of FP64 body coordinates (n x 3) to the float points
 * a CompactCoordinates rebuild holds after the move (Body::rotate/translate
 * act on FP64 atoms; CompactCoordinatesXYZW.h:106-165 casts each coordinate):
 *   out_x = (float)(((R00*x + R01*y) + R02*z) + t0), likewise y, z;
 * T = {R00,R01,R02, R10,R11,R12, R20,R21,R22, t0,t1,t2}; out[4*i+3] = w[i]. */
void orc_rigid_transform(const double* xyz, const float* w, int n, const double* T, float* out);

/* PartialHistogramManagerMT::calculate_all (PartialHistogramManagerMT.cpp:42-257)
 * in the state reached after every body moved once the self terms were cached:
 *   aa = sum_b self(self_bodies[b]) + sum_{i>j} cross(rows body j, cols body i)
 *        (calc_aa(i, j) -> enqueue_calculate_cross(a[i], a[j]): rows over data_2)
 *   aw = sum_i cross(rows waters, cols body i)      (calc_aw(i))
 *   ww = self(waters)                               (calc_ww)
 *   ptot = aa + aw + ww (the master histogram), max_bin as calculate() (:133-141).
 * self_bodies/bodies: XYZW points concatenated per body, sizes[n_bodies].
 * Returns max_bin or a negative status. */
int orc_hist_partial(const float* self_bodies, const float* bodies, const int* sizes, int n_bodies,
                     const float* waters, int n_w, float inv_w, int nb, int nthreads,
                     orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot);

/* FP64 affine map of float XYZW points as generate_transformed_data applies
 * a symmetry copy (SymmetryHelpers.cpp:26-36, ISymmetry.h:24, Vector3.h:169-181):
 *   out_x = (float)(((R00*x + R01*y) + R02*z) + t0) with x = (double) pts_x;
 * T = {R00..R22, t0, t1, t2}; the weight is copied. */
void orc_affine_float(const float* pts, int n, const double* T, float* out);

/* SymmetryManagerMT::calculate (SymmetryManagerMT.cpp:39-170) with XYZW
 * weights: per body a self histogram scaled by 1 + its copies, its pair
 * schedules (repA, repB, scale), copy x other-body / other-symmetry crosses,
 * body x water crosses for the original and every copy, water self.
 * Layout as ausaxs_symmetry_create. Returns max_bin or a negative status. */
int orc_hist_symmetry(const float* atoms, const int* sizes, int n_bodies, const int* sym_per_body,
                      const int* sym_reps, const double* transforms, const int* sched_len, const int* schedule,
                      const float* waters, int n_w, float inv_w, int nb, int nthreads, orc_entry* aa,
                      orc_entry* aw, orc_entry* ww, orc_entry* ptot);

/* math::CubicSpline (CubicSpline.cpp): knots (x, y), n >= 4, evaluated at
 * m points z. Returns 0, or -1 for fewer than four knots. */
int orc_spline(const double* x, const double* y, int n, const double* z, int m, double* out);

/* PartialSymmetryManagerMT::calculate_all (PartialSymmetryManagerMT.cpp:78-450)
 * in the state after the given changes: self terms from self_atoms with
 * factors self_scale[b] (calc_aa_self at the last internal change), every
 * other block from atoms with the manager's roles (calc_aa: rows over the
 * lower body / the lower symmetry; calc_aw: rows over the waters). */
int orc_hist_symmetry_partial(const float* atoms, const float* self_atoms, const int* self_scale, const int* sizes,
                              int n_bodies, const int* sym_per_body, const int* sym_reps, const double* transforms,
                              const int* sched_len, const int* schedule, const float* waters, int n_w, float inv_w,
                              int nb, int nthreads, orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot);

/* HistogramManagerMTFFGridSurface::calculate_all (HistogramManagerMTFFGridSurface.cpp:16-274)
 * with interior dummies xi and surface dummies xs (n x 4 floats, type EXV):
 * the FFAvg base with Z from the grid volume, then xx interior / surface /
 * interior x surface (factor 2), ax and wx against each (factor 1); exv rows
 * of aa / aw set to the profiles at unit surface scaling. Profiles: [nb]
 * (xx*, wx*), [15][nb] (ax*). Returns max_bin or a negative status. */
int orc_hist_ffgrid_surface(const float* atoms, int n_a, const float* waters, int n_w, const float* xi, int n_i,
                            const float* xs, int n_s, double volume, float inv_w, int nb, int nthreads,
                            orc_entry* aa, orc_entry* aw, orc_entry* ww, orc_entry* ptot, orc_entry* ptot_ax,
                            orc_entry* pxx, orc_entry* xxi, orc_entry* xxs, orc_entry* xxc, orc_entry* wxi,
                            orc_entry* wxs, orc_entry* axi, orc_entry* axs);

/* ---- weighted axis, sinc, Debye ------------------------------------------ */
/* WeightedDistribution1D::get_weighted_axis (WeightedDistribution1D.cpp:51-58)
 * followed by DistanceHistogram::initialize (DistanceHistogram.cpp:44-47, d[0]=0). */
void orc_weighted_axis(const orc_entry* p, int n, double bin_width, double* d_out);

/* VectorDebyeTable::initialize (VectorDebyeTable.cpp:24-38): out[iq*nd + id]. */
void orc_sinc_table(const double* q, int nq, const double* d, int nd, double* out);

/* std::inner_product per (row, q): out[r*nq + iq] = sum_b H[r*nd+b]*S[iq*nd+b]
 * (CompositeDistanceHistogramFFAvgBase.cpp:427-467). */
void orc_contract(const double* H, int rows, int nd, const double* S, int nq, double* out);

/* Kahan-free long-double contraction used as the accuracy reference for the
 * device GEMM (same operands as orc_contract, 80-bit accumulation). */
void orc_contract_ld(const double* H, int rows, int nd, const double* S, int nq, double* out);

/* ---- form factors -------------------------------------------------------- */
/* FormFactor::evaluate (FormFactor.h:33-39) with the raw 5-Gaussian tables of
 * FormFactorTable.h:28-120 mapped per lookup::atomic::raw::get (FormFactor.h:93-117). */
double orc_ff_atomic(int type, double q);
/* ExvFormFactor(volume).evaluate (ExvFormFactor.h:16-42) with the default
 * MinimumFluctuation_implicit_H set (ExvTable.h, ExvSettings.h:57-69). */
double orc_ff_exv(int type, double q);

/* Free parameters of ICompositeDistanceHistogramExv (FFAvgBase.h). */
typedef struct {
    double cw, cx, crho, dw_atomic, dw_exv;
    double grid_v; /* FFGrid: ExvFormFactor volume of the dummies (0 reads as 1 A^3) */
} orc_params;

/* Six partial profiles {aa, ax, aw, xx, wx, ww} on an arbitrary q grid:
 * FFAvg: cache_refresh_intensity_profiles (FFAvgBase.cpp:470-558) + DW
 * (:310-374); FFExplicit: FFExplicitBase.cpp:105-212 with exv_factor :54-57.
 * sinqd_aa: [15][15][nq], sinqd_aw: [15][nq], sinqd_ww: [nq] (FFAvg also
 * reads the EXV rows/columns of aa/aw). out: [6][nq]. */
void orc_profiles(int mode, const double* q, int nq, const double* sinqd_aa, const double* sinqd_aw,
                  const double* sinqd_ww, orc_params p, double* out);

/* I = aa - ax + aw + xx - wx + ww (FFAvgBase.cpp:57-76). */
void orc_total(const double* profiles, int nq, double* out);

/* Exact O(N^2) Debye with f = exp(-q^2/2) per atom, FP64 over FP32-rounded
 * distances (ExactDebyeCalculator.cpp:13-57). xyzw: n x 4 floats. */
void orc_exact_debye(const float* xyzw, int n, const double* q, int nq, double* out);

/* LinearLeastSquares::fit_params_only (LinearLeastSquares.cpp:26-44):
 * weighted fit y ~ a*x + b with weights 1/err^2. out[0]=a, out[1]=b,
 * out[2]=S/delta, out[3]=Sxx/delta, out[4]=chi2 (Fitter.cpp:10-13). */
void orc_lsq(const double* x, const double* y, const double* err, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif
