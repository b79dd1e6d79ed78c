/*
 * ausaxs_hip.h — C ABI of the MI355X-native AUSAXS Debye hot path.
 *
 * Plain C: pointers and sizes only. Every function returns 0 on success and a
 * non-zero status on failure; the message of the last failure on the calling
 * thread is available from ausaxs_get_last_error. This mirrors the reference's
 * C API conventions (source/api/api_helper.h:15-45: status out-parameter plus
 * get_last_error_msg), with the status as the return value.
 *
 * Point arrays are n x 4 floats in the reference's CompactCoordinates layout
 * (hist/detail/data/CompactCoordinatesXYZFF.h:107-169, ...XYZW.h:106-165):
 *   XYZFF: {x, y, z, form-factor type as int32 bits}   (FF managers)
 *   XYZW : {x, y, z, weight}                           (plain manager)
 * Coordinates are the FP64 atom coordinates cast to float by the caller, as
 * the reference's CompactCoordinates constructors do.
 */
#ifndef AUSAXS_HIP_H
#define AUSAXS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* manager choices (settings/HistogramSettings.h:23-42) */
#define AUSAXS_MODE_FFAVG 0      /* HistogramManagerMTFFAvg       */
#define AUSAXS_MODE_FFEXPLICIT 1 /* HistogramManagerMTFFExplicit  */
#define AUSAXS_MODE_SIMPLE 2     /* HistogramManagerMT (XYZW)     */
#define AUSAXS_MODE_FFGRID 3     /* HistogramManagerMTFFGrid      */
#define AUSAXS_MODE_FFGRID_SURFACE 4 /* HistogramManagerMTFFGridSurface */

/* which distribution ausaxs_histogram_get returns */
#define AUSAXS_AA 0
#define AUSAXS_AW 1
#define AUSAXS_WW 2
#define AUSAXS_PTOT 3
#define AUSAXS_PTOT_AX 4 /* FFGrid: weighted atom/water-dummy totals [max_bin] */
#define AUSAXS_XX 5      /* FFGrid: weighted dummy-dummy histogram [max_bin]   */
/* FFGridSurface: interior / surface dummy profiles (CompositeDistanceHistogramFFGridSurface.cpp:33-46) */
#define AUSAXS_XX_INTERIOR 6 /* [max_bin] */
#define AUSAXS_XX_SURFACE 7  /* [max_bin] */
#define AUSAXS_XX_CROSS 8    /* [max_bin], interior x surface */
#define AUSAXS_WX_INTERIOR 9 /* [max_bin] */
#define AUSAXS_WX_SURFACE 10 /* [max_bin] */
#define AUSAXS_AX_INTERIOR 11 /* [15][max_bin] */
#define AUSAXS_AX_SURFACE 12  /* [15][max_bin] */

typedef struct {
    double bin_width;  /* settings::axes::bin_width (default 0.25)  */
    int bin_count;     /* settings::axes::bin_count (default 8000)  */
    int weighted_bins; /* settings::hist::weighted_bins (default 1) */
    int debye_terms;   /* bf16 terms per GEMM operand: 2 = hi+lo, 3 = hi+mid+lo (default); 0 = FP64 */
    int tile_j;        /* columns per pair work item (default 2000, at most 8064) */
    int lattice;       /* grid dummy-dummy pairs by lattice correlation instead of pair by pair
                          (identical histograms): 0 off, 1 when cheaper (default), 2 always */
    double surface_thickness; /* grid-surface excluded volume: settings::grid::exv::surface_thickness
                                 (GridSettings.cpp:17; A, default 1, in [0, 64]) */
} ausaxs_settings;

typedef struct ausaxs_structure ausaxs_structure; /* resident coordinates + work lists */
typedef struct ausaxs_histogram ausaxs_histogram; /* ICompositeDistanceHistogram(Exv)  */

void ausaxs_default_settings(ausaxs_settings* s);
int ausaxs_get_last_error(char* buf, int len);
int ausaxs_device_info(char* name, int len, int* compute_units, size_t* lds_per_block);

/* Upload coordinates and build the pair work lists once.
 * Replaces the construction half of HistogramManagerMTFFAvg/FFExplicit/MT
 * (HistogramManagerMTFFAvg.cpp:203-208: CompactCoordinatesFF of bodies and waters).
 * z_exv: FFAvg only, V_grid * 0.334 / N_atoms (HistogramManagerMTFFAvg.cpp:355-359). */
int ausaxs_structure_create(int mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                            const ausaxs_settings* settings, double z_exv, ausaxs_structure** out);

/* Grid excluded-volume structure (HistogramManagerMTFFGrid, HistogramManagerMTFFGrid.cpp:43-234).
 * Coordinates are FP64 (n x 3), as the grid is built from the FP64 atoms
 * (Grid.cpp:34-110); they are cast to float for the pair kernels like
 * CompactCoordinatesFF does. Dummies come from ausaxs_grid_exv. */
int ausaxs_structure_create_grid(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                 const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                 ausaxs_structure** out);

/* Grid excluded volume with a surface layer (HistogramManagerMTFFGridSurface,
 * HistogramManagerMTFFGridSurface.cpp:16-274; CompositeDistanceHistogramFFGridSurface.cpp).
 * The dummies split into interior and surface ones (RawGridWithSurfaceExv::create,
 * RawGridWithSurfaceExv.cpp:7-155); the excluded-volume parameter (cx of
 * ausaxs_histogram_set_params) scales the surface dummies only. */
int ausaxs_structure_create_grid_surface(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                         const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                         ausaxs_structure** out);
/* HistogramManagerMTFFAvg from FP64 atoms (n x 3, form-factor types 1..14)
 * and waters (m x 3) with the manager's own excluded-volume charge:
 * Z_exv = V_grid * 0.334 / N (HistogramManagerMTFFAvg.cpp:355-359), V_grid the
 * volume of the molecule's grid (Molecule::get_volume_grid, Molecule.cpp:
 * 147-150; the same grid as ausaxs_grid_exv, whose volume it is). *z_exv
 * (optional) receives the value used. The atoms and waters are cast to float
 * as CompactCoordinatesFF holds them. */
int ausaxs_structure_create_ffavg(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                                  const double* waters_xyz, int n_waters, const ausaxs_settings* settings,
                                  double* z_exv, ausaxs_structure** out);
/* Radial hydration shell: RadialHydration::generate_explicit_hydration
 * (RadialHydration.cpp:40-80) per body on the molecule's grid, as
 * GridBasedHydration::hydrate runs it (GridBasedHydration.cpp:36-79) with the
 * default NoCulling; candidates at vdw + 1.36 A along 26 directions from each
 * atom, accepted by RadialHydration::collision_check (:141-206), each accepted
 * water marking the grid before the next candidate. atoms: n x 3 FP64 with
 * form-factor types; body_sizes: n_bodies sizes summing to n (NULL / 0: one
 * body); noise: NULL (none; the reference's tests) or n x 26 x 3 displacements
 * (the reference draws N(0, 0.75) per component). waters: capacity x 3, in
 * placement order (capacity 26 n always suffices; NULL: count only);
 * per_body: n_bodies counts or NULL. settings: NULL for the reference's
 * defaults, else the radii, shell correction and grid extent below. Host
 * only (no device work). */
typedef struct ausaxs_hydrate_settings {
    double atomic_radius;    /* Grid::get_atomic_radius for every atom; < 0: each type's vdW radius (default -1) */
    double hydration_radius; /* Grid::get_hydration_radius (default 1.71, the vdW radius of O) */
    double shell_correction; /* settings::hydrate::shell_correction (default -0.35) */
    int explicit_grid;       /* 1: the grid of Grid(Limit3D(grid_min, grid_max)) with 1 A cells; 0: Grid(bodies) */
    double grid_min[3], grid_max[3];
} ausaxs_hydrate_settings;
void ausaxs_default_hydrate_settings(ausaxs_hydrate_settings* s);
int ausaxs_hydrate_radial(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const int* body_sizes,
                          int n_bodies, const double* noise, double* waters_xyz, int capacity, int* n_waters,
                          int* per_body, const ausaxs_hydrate_settings* settings);
/* The interior and surface dummies of RawGridWithSurfaceExv::create, each in the
 * reference's voxel order; counts are returned also when a capacity is too small.
 * surface_thickness: settings::grid::exv::surface_thickness (A; 1 is the default). */
int ausaxs_grid_exv_surface(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                            int n_waters, double* interior_xyz, int capacity_interior, int* n_interior,
                            double* surface_xyz, int capacity_surface, int* n_surface, double surface_thickness);

/* Excluded-volume dummy atoms of the default grid (RawGridExv::create,
 * RawGridExv.cpp:10-59; Grid.cpp:34-110; SphericalExpander.h:29-90) in the
 * reference's order, and the grid volume (Grid::get_volume, Grid.cpp:429-432).
 * out_xyz: capacity x 3 doubles; if capacity is too small, *n_out still
 * receives the required count and status 1 is returned. */
int ausaxs_grid_exv(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                    int n_waters, double* out_xyz, int capacity, int* n_out, double* volume);

/* FFGrid distance axes of the ax and xx rows (CompositeDistanceHistogramFFGrid.cpp:130-136): [max_bin] each. */
int ausaxs_histogram_grid_axes(const ausaxs_histogram* h, double* d_ax, double* d_xx);

/* IHistogramManager::calculate_all (IHistogramManager.h:16-29) on resident data:
 * HistogramManagerMTFFAvg.cpp:194-371, HistogramManagerMTFFExplicit.cpp:216-436,
 * HistogramManagerMT.cpp:23-75. */
int ausaxs_structure_calculate(ausaxs_structure* s, ausaxs_histogram** out);
int ausaxs_structure_free(ausaxs_structure* s);

/* One-shot manager: create + calculate + free the structure. */
int ausaxs_hist_calculate(int mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                          const ausaxs_settings* settings, double z_exv, ausaxs_histogram** out);

/* Shape of a composite histogram: max_bin after truncation
 * (HistogramManagerMTFFAvg.cpp:340-352). */
int ausaxs_histogram_info(const ausaxs_histogram* h, int* mode, int* max_bin);

/* Reference-layout distributions (WeightedEntry fields split into arrays).
 * FF modes: AA is [15][15][max_bin] (Container3D.h:48), AW is [15][max_bin],
 * WW and PTOT are [max_bin]. SIMPLE mode: all are [max_bin].
 * Any of value/count/bin_center may be NULL. */
int ausaxs_histogram_get(const ausaxs_histogram* h, int which, double* value, uint32_t* count, double* bin_center);

/* Weighted distance axis (WeightedDistribution1D::get_weighted_axis,
 * WeightedDistribution1D.cpp:51-58, with d[0] = 0 per DistanceHistogram.cpp:44-47). */
int ausaxs_histogram_d_axis(const ausaxs_histogram* h, double* d);

/* apply_water_scaling_factor, apply_excluded_volume_scaling_factor,
 * apply_solvent_density_scaling_factor, apply_{atomic,exv}_debye_waller_factor
 * (ICompositeDistanceHistogramExv.h:12-107). */
int ausaxs_histogram_set_params(ausaxs_histogram* h, double cw, double cx, double crho, double dw_atomic,
                                double dw_exv);

/* debye_transform on an explicit q grid (DistanceHistogram.cpp:70-87 for the
 * plain histogram; CompositeDistanceHistogramFFAvgBase.cpp:57-76 for the
 * composite ones). I: [nq]. profiles (optional): [6][nq] = aa, ax, aw, xx, wx, ww. */
int ausaxs_histogram_debye(ausaxs_histogram* h, const double* q, int nq, double* I, double* profiles);

int ausaxs_histogram_free(ausaxs_histogram* h);

/* ---- batched rigid-body scan (PartialHistogramManagerMT) ------------------- */

typedef struct ausaxs_rigid ausaxs_rigid;

/* Bodies for PartialHistogramManagerMT (PartialHistogramManagerMT.cpp:42-341):
 * xyz: FP64 coordinates (n x 3, bodies concatenated in order), w: XYZW weight
 * per atom (CompactCoordinatesXYZW.h:106-165), body_sizes[n_bodies]. Waters
 * (optional, FP64 m x 3 + weights) stay fixed. The self-correlation of every
 * body and of the waters is computed here, once, from these coordinates
 * (initialize / calc_self_correlation / calc_ww, :259-294), as the reference
 * caches it while bodies only move rigidly. settings->tile_j = 0 selects the
 * autotuner (tile width, LDS histogram copies and persistent grid, timed on
 * the first batch of each shape). */
int ausaxs_rigid_create(const double* xyz, const float* w, const int32_t* body_sizes, int n_bodies,
                        const double* waters_xyz, const float* waters_w, int n_waters, const ausaxs_settings* settings,
                        ausaxs_rigid** out);

/* Score n_conf conformations. transforms: [n_conf][n_bodies][12] = {R00..R22
 * row-major, t0, t1, t2}; each body's atoms become float(R x + t) (FP64,
 * no contraction), as a CompactCoordinates rebuild of moved bodies holds
 * (update_compact_representation_body, :165-169). Per conformation this is
 * PartialHistogramManagerMT::calculate_all after every moved body's inter-body
 * and water blocks were recomputed (calc_aa, calc_aw; :96-108), followed by
 * CompositeDistanceHistogram::debye_transform on q with water scaling cw
 * (CompositeDistanceHistogram.cpp:55-70, DistanceHistogram.cpp:54-87).
 * Bodies whose transform is the same in every conformation keep their
 * mutual blocks from one evaluation. I: [n_conf][nq]. */
int ausaxs_rigid_scan(ausaxs_rigid* r, const double* transforms, int n_conf, const double* q, int nq, double cw,
                      double* I);

/* The hydration-modified step of PartialHistogramManagerMT::calculate_all
 * (PartialHistogramManagerMT.cpp:83-105: calc_ww, and calc_aw for every body),
 * as after SimpleController::step re-hydrates the moved molecule
 * (source/rigidbody/controller/SimpleController.cpp:78). Conformation c has its
 * own waters: water_counts[c] consecutive rows of waters_xyz (FP64 x, y, z in
 * the frame of the moved bodies; cast to float as CompactCoordinates does) and
 * waters_w, in conformation order. They replace the waters given at creation;
 * each weight must be one that a body point or creation water has, unless the
 * scan has more than 64 distinct weights. The bodies' self terms stay cached.
 * ausaxs_rigid_histogram then returns this scan's conformations. */
int ausaxs_rigid_scan_hydrated(ausaxs_rigid* r, const double* transforms, int n_conf, const double* waters_xyz,
                               const float* waters_w, const int32_t* water_counts, const double* q, int nq,
                               double cw, double* I);

/* The composite histogram (SIMPLE layout) of one conformation of the last
 * batch of the last scan (batches hold up to 1024 conformations). */
int ausaxs_rigid_histogram(ausaxs_rigid* r, int conformation, ausaxs_histogram** out);

/* Launch shape and timing of the last scan: tile width, persistent grid
 * (0 = one workgroup per item), LDS histogram copies, conformations per graph
 * launch, work items per launch, pair evaluations, HIP-event time of the
 * graph launches. Any pointer may be NULL. */
int ausaxs_rigid_info(const ausaxs_rigid* r, int* tile_j, int* grid, int* copies, int* batch, int* n_items,
                      long long* pairs, double* scan_ms);
int ausaxs_rigid_free(ausaxs_rigid* r);

/* ---- serving: profile requests with continuous batching ------------------- */

typedef struct ausaxs_server ausaxs_server;

/* A request loop with `slots` device slots (worker threads with their own
 * streams, 1..16) and a device-memory budget in GiB. Requests are admitted
 * in arrival order as soon as a slot is idle and their estimated footprint
 * fits the budget, and run concurrently with those already on the GPU. With
 * two or more slots the last one is an express slot that admits only small
 * requests (below 2e9 estimated pair evaluations), the first one queued, so
 * they do not wait while large requests hold the other slots.
 * Structures stay resident across requests (keyed by their exact content,
 * mode and settings; idle ones are evicted least recently used first when
 * the budget needs the memory), and small requests queued together run as
 * one merged pair launch. There
 * is no reference counterpart; each request is one manager evaluation
 * (HistogramManagerMTFF{Avg,Explicit,Grid}::calculate_all) plus
 * debye_transform on its q grid, identical to ausaxs_structure_create[_grid] +
 * ausaxs_structure_calculate + ausaxs_histogram_debye on the same input. */
int ausaxs_server_create(int slots, double budget_gb, ausaxs_server** out);

/* Queue one request (inputs are copied): FP64 atoms (n x 3) with form-factor
 * types (1..14), FP64 waters (m x 3), q grid. mode: AUSAXS_MODE_FFAVG,
 * _FFEXPLICIT or _FFGRID. A request whose footprint exceeds the whole budget
 * is refused (status 1). */
int ausaxs_server_submit(ausaxs_server* s, int mode, const double* atoms_xyz, const int32_t* atom_ff, int n_atoms,
                         const double* waters_xyz, int n_waters, const double* q, int nq,
                         const ausaxs_settings* settings, double z_exv, long long* id);

/* Wait up to timeout_s for a request. *ready = 1: I[nq] holds the profile,
 * latency_ms = completion - submission, queue_ms = admission - submission,
 * and the request is released. *ready = 0: still running (status 0). A
 * failed request returns status 1 with its error message. */
int ausaxs_server_wait(ausaxs_server* s, long long id, double timeout_s, double* I, double* latency_ms,
                       double* queue_ms, int* ready);

/* Counters since creation; peak_gb = largest sum of admitted footprints;
 * pairs: pairs in the histograms of completed requests, kernel_pairs: of
 * which the pair kernel evaluated (grid dummy pairs go by lattice correlation). */
int ausaxs_server_stats(ausaxs_server* s, long long* submitted, long long* completed, long long* failed,
                        int* in_flight, int* peak_in_flight, double* peak_gb, long long* pairs,
                        long long* kernel_pairs);

/* Batching and residency counters: requests served from a resident
 * structure (cache_hits) or built (cache_misses), merged launches (batches)
 * and the requests they ran (batched), evicted structures, the device
 * memory idle resident structures hold (GiB), and the requests the express
 * slot admitted. No reference counterpart. Any pointer may be NULL. */
int ausaxs_server_batch_stats(ausaxs_server* s, long long* cache_hits, long long* cache_misses, long long* batches,
                              long long* batched, long long* evictions, double* cached_gb, long long* express);

/* Stops the loop: running requests finish, queued ones fail. */
int ausaxs_server_free(ausaxs_server* s);

/* The library's caching allocator (every device and pinned host buffer):
 * bytes of device blocks held and handed out, hipMalloc / hipHostMalloc calls
 * so far, allocations served from freed blocks. A freed buffer returns to the
 * pool, not to HIP (hipFree waits for the whole device); ausaxs_pool_trim
 * returns the free blocks to HIP. No reference counterpart. */
int ausaxs_pool_stats(size_t* device_held, size_t* device_in_use, long long* device_mallocs,
                      long long* host_mallocs, long long* reuses);
int ausaxs_pool_trim(void);

/* The reference's default q sub-axis (ConstantsAxes.h:14, Axis::sub_axis,
 * Axis.cpp:40-59). q must hold 200 doubles; *nq receives the length. */
int ausaxs_default_q_axis(double qmin, double qmax, double* q, int* nq);

/* ---- fitting and validation --------------------------------------------- */

/* LinearLeastSquares::fit_params_only (fitter/detail/LinearLeastSquares.cpp:26-44)
 * and Fitter::chi2 (fitter/Fitter.cpp:10-13) in FP64 on the device, for
 * n_models model curves x[n_models][n] against one data set y[n] with errors
 * err[n] (NULL: unit errors): y ~ a x + b. out[n_models][5] = a, b, S/delta,
 * Sxx/delta, chi2. Scores a whole rigid-body scan against data in one call. */
int ausaxs_lsq(const double* x, const double* y, const double* err, int n, int n_models, double* out);

/* math::CubicSpline (math/CubicSpline.cpp): the natural cubic spline through
 * (x[k], y[k]), k < n (n >= 4, x strictly increasing), at z[i], i < m, the end
 * intervals' cubics continuing outside the knots. The fitter splines model
 * curves onto the data's q values with it. Host only. */
int ausaxs_cubic_spline(const double* x, const double* y, int n, const double* z, int m, double* out);

/* hist::exact_debye_transform (intensity_calculator/ExactDebyeCalculator.cpp:13-57):
 * I(q) = exp(-q^2) [sum_i w_i^2 + sum_{i<j} 2 w_i w_j sinc(q d_ij)] over XYZW
 * points (n x 4 floats), distances with the reference's evaluate_8/4/1 rules,
 * FP64 sums. */
int ausaxs_exact_debye(const float* xyzw, int n, const double* q, int nq, double* I);

/* ---- kernel-level entries (tests, tools) -------------------------------- */

/* One pair set with the reference's loop geometry (self if b == NULL, else
 * rows a x columns b): per-bin pair counts and summed distances. Replaces the
 * evaluate/increment loop of SimpleCalculator.h:94-215 for one set. */
int ausaxs_pair_histogram(const float* a, int n, const float* b, int m, double bin_width, int bin_count,
                          uint64_t* count, double* dsum);

/* Debye contraction out[r][k] = sum_b H[r][b] sinc(q_k d_b) on the device
 * (cache_refresh_sinqd, CompositeDistanceHistogramFFAvgBase.cpp:427-467).
 * terms: 2/3 = bf16-split MFMA, 0 = FP64. */
int ausaxs_debye_contract(const double* H, int rows, int K, const double* d, const double* q, int nq, int terms,
                          double* out);

/* No counterpart in the reference: ausaxs_debye_contract for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). H (rows x K), d (K), q (nq) and out
 * (rows x nq) are device pointers; the work is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting.
 * It allocates nothing: its work space is a static device array of
 * AUSAXS_DEBYE_NOALLOC_SCRATCH_BYTES on every GPU the library runs on, held
 * until the process ends. Contractions that need more (bf16-split work space
 * 6 (nq + rows) K bytes plus the partial tiles) fail with an error. Calls
 * that overlap on one GPU share that array and race; serialise them. */
#define AUSAXS_DEBYE_NOALLOC_SCRATCH_BYTES (32u << 20)
int ausaxs_debye_contract_noalloc(const double* H, int rows, int K, const double* d, const double* q, int nq,
                                  int terms, double* out, void* stream);

/* Exhaustive device check of the binning arithmetic over float d2 bit
 * patterns [lo_bits, hi_bits]: out[0] = pair-kernel bins that differ from
 * the correctly rounded rule, out[1] = device sqrt results that are not
 * correctly rounded, out[2] = results of the tie fallback's residual
 * correction (normal d2) that differ from the correctly rounded sqrt,
 * out[3] = pair-kernel offset codes more than 1 + 2^-24 q units (q = units
 * per bin) from the exact offset of the real root, out[4] = the sum of the code errors in units of 2^-16
 * (two's complement), out[5] = the same sum for the correctly rounded root.
 * out: 6 entries. No counterpart in the reference: the pin of the pair
 * kernel's arithmetic behind CompactCoordinatesXYZFF.h:514. */
int ausaxs_bin_check(uint32_t lo_bits, uint32_t hi_bits, double bin_width, unsigned long long* out);

/* Device-to-device copy of `bytes` (multiple of 16) by a plain streaming
 * kernel, `reps` times; *ms = mean kernel time. A traffic-counter reference of
 * known size (no counterpart in the reference). */
int ausaxs_copy_probe(size_t bytes, int reps, double* ms);

/* The device's own issue ceilings for the pair kernel's instruction classes
 * (no counterpart in the reference), each the mean over `reps` launches after
 * one warm-up, in wave instructions per CU per nanosecond: rates[0] FP32 adds
 * with vector-register operands, rates[1] the same with a scalar-register
 * operand, rates[2] v_rsq_f32, rates[3] ds_add_u64 at hashed cells of an
 * nbins-cell LDS histogram (three 512-lane workgroups per CU; nbins * 8 <=
 * 65536). rates: 4 entries. */
int ausaxs_issue_probe(int nbins, int reps, double* rates);

/* HIP-event timings of the last ausaxs_structure_calculate with timing on.
 * pairs: all pairs in the histograms; kernel_pairs: of which pair_hist_kernel
 * evaluated; lattice_pairs: of which lattice correlation counted. */
int ausaxs_structure_set_timing(ausaxs_structure* s, int on);
int ausaxs_last_timing(const ausaxs_structure* s, double* pair_ms, double* total_ms, long long* pairs,
                       int* n_items, long long* kernel_pairs, long long* lattice_pairs);
/* The merge of that calculation (merge_chunks_kernel + merge_rows_kernel,
 * HIP events): its time and its HBM bytes (partial rows read, chunk sums
 * written and read back, result rows written), for its achieved GB/s (no
 * counterpart in the reference). */
int ausaxs_last_merge_timing(const ausaxs_structure* s, double* merge_ms, long long* merge_bytes);

/* ---- bodies with symmetry copies ------------------------------------------ */

/* Replaces SymmetryManagerMT<weighted_bins, vbw>::calculate_all
 * (SymmetryManagerMT.cpp:29-170) for XYZW bodies. atoms: sum(body_sizes) x 4
 * floats {x, y, z, w}, body after body. Body b has sym_per_body[b]
 * symmetries; symmetry k (counted over all bodies) has sym_reps[k] copies,
 * each an affine map of 12 doubles in `transforms` (rotation row-major, then
 * translation; ISymmetry.h:17-25), applied in FP64 to the float coordinates
 * as generate_transformed_data does (SymmetryHelpers.cpp:16-43), and a pair
 * schedule of sched_len[k] entries (repA, repB, scale) in `schedule`
 * (ISymmetry::internal_pair_schedule, ISymmetry.h:26-39; repetition 0 = the
 * original). Any positive scale is accepted (the reference's switch stops at
 * 30 and 60, SimpleCalculator.h:286-377). Result: a SIMPLE-mode histogram
 * (aa, aw, ww, p_tot, max_bin) as the reference's CompositeDistanceHistogram. */
typedef struct ausaxs_symmetry ausaxs_symmetry;
int ausaxs_symmetry_create(const float* atoms, const int* body_sizes, int n_bodies, const int* sym_per_body,
                           const int* sym_reps, const double* transforms, const int* sched_len, const int* schedule,
                           const float* waters, int n_waters, const ausaxs_settings* settings,
                           ausaxs_symmetry** out);
/* Replaces PartialSymmetryManagerMT<weighted_bins, vbw> (PartialSymmetryManagerMT.cpp:78-450):
 * same arguments; the pairs are taken with that manager's row/column roles
 * (calc_aa, calc_aw, :596-706) and a body's self term keeps the coordinates
 * and copy count of its last internal modification (:265-276). */
int ausaxs_symmetry_create_partial(const float* atoms, const int* body_sizes, int n_bodies, const int* sym_per_body,
                                   const int* sym_reps, const double* transforms, const int* sched_len,
                                   const int* schedule, const float* waters, int n_waters,
                                   const ausaxs_settings* settings, ausaxs_symmetry** out);
/* State changes between calculations (the state manager's signals,
 * StateManager.h:16-128): new float coordinates of one body (an external,
 * rigid move; internal = 1 for a change of its own distances or weights), one
 * symmetry's copies and schedule, or the waters. */
int ausaxs_symmetry_set_body(ausaxs_symmetry* s, int body, const float* atoms, int internal);
int ausaxs_symmetry_set_symmetry(ausaxs_symmetry* s, int body, int symmetry, int reps, const double* transforms,
                                 int sched_len, const int* schedule);
int ausaxs_symmetry_set_waters(ausaxs_symmetry* s, const float* waters, int n_waters);
int ausaxs_symmetry_calculate(ausaxs_symmetry* s, ausaxs_histogram** out);
int ausaxs_symmetry_free(ausaxs_symmetry* s);
/* Pair evaluations of one calculate. */
int ausaxs_symmetry_info(const ausaxs_symmetry* s, long long* pairs, int* n_items);

/* ISymmetry::internal_pair_schedule (ISymmetry.cpp:43-66): the cyclic-chain
 * schedule of a symmetry with reps copies. out: 3 ints per entry; *n_out =
 * entries (also when capacity is too small: status 1). */
int ausaxs_symmetry_chain_schedule(int reps, int closed, int* out, int capacity, int* n_out);
/* compute_pair_schedule (PairScheduleUtil.cpp:33-60): n placements of 12
 * doubles (placement 0 = the original body). */
int ausaxs_symmetry_placement_schedule(const double* placements, int n, int* out, int capacity, int* n_out);
/* CyclicSymmetry::_make_transform / is_closed (CyclicSymmetry.cpp:85-118):
 * the maps of copies 1..reps (12 doubles each) for a body with centre cm. */
int ausaxs_symmetry_cyclic(const double* cm, const double* offset, const double* step, const double* axis,
                           double angle, int reps, double* out, int* closed);

/* ---- grid excluded volume with stretchable dummies -------------------------- */

/* Replaces HistogramManagerMTFFGridScalableExv (HistogramManagerMTFFGridScalableExv.cpp:16-244)
 * and CompositeDistanceHistogramFFGridScalableExv (:9-19). The grid and its
 * dummies are built once (as ausaxs_structure_create_grid); calculate(k)
 * returns the FFGrid histogram with the dummies' float coordinates multiplied
 * by k and the dummy form factor of volume k^3 (exv width 1 A). */
typedef struct ausaxs_scalable ausaxs_scalable;
int ausaxs_scalable_create(const double* atoms_xyz, const int32_t* atom_ff, int n_atoms, const double* waters_xyz,
                          int n_waters, const ausaxs_settings* settings, ausaxs_scalable** out);
int ausaxs_scalable_calculate(ausaxs_scalable* s, double k, ausaxs_histogram** out);
int ausaxs_scalable_free(ausaxs_scalable* s);

/* ---- fitting a histogram to measured data ---------------------------------- */

/* Replaces SmartFitter(data, histogram) (SmartFitter.cpp:50-58). The
 * histogram must outlive the fitter. Data: n points (q, I, err; err NULL =
 * unit errors). enabled: AUSAXS_FIT_* bits (EnabledFitParameters,
 * SmartFitter.cpp:11-30); SIMPLE histograms keep only the hydration bit
 * (validate_model, :32-49). q_model NULL: the default q axis on [1e-4, 0.5]
 * (DistanceHistogram::get_q_axis). */
#define AUSAXS_FIT_HYDRATION 1
#define AUSAXS_FIT_EXCLUDED_VOLUME 2
#define AUSAXS_FIT_SOLVENT_DENSITY 4
#define AUSAXS_FIT_ATOMIC_DEBYE_WALLER 8
#define AUSAXS_FIT_EXV_DEBYE_WALLER 16
#define AUSAXS_MINI_GOLDEN 0 /* mini::algorithm::GOLDEN */
#define AUSAXS_MINI_SCAN 2   /* mini::algorithm::SCAN */
typedef struct ausaxs_fitter ausaxs_fitter;
int ausaxs_fitter_create(ausaxs_histogram* h, const double* q, const double* I, const double* err, int n,
                         unsigned enabled, const double* q_model, int nq_model, ausaxs_fitter** out);
/* The same fitter on a scalable grid structure: the excluded-volume parameter
 * stretches the dummies (apply_excluded_volume_scaling_factor of
 * CompositeDistanceHistogramFFGridScalableExv) and recomputes the histograms. */
int ausaxs_fitter_create_scalable(ausaxs_scalable* s, const double* q, const double* I, const double* err, int n,
                                  unsigned enabled, const double* q_model, int nq_model, ausaxs_fitter** out);
/* Number of free parameters and the enabled mask after validation. */
int ausaxs_fitter_params(const ausaxs_fitter* f, int* n_params, unsigned* enabled);
/* SmartFitter::chi2 (Fitter.cpp:10-13) for n_sets parameter vectors
 * (n_sets x n_params, in bit order) in one batch: out[5*k..] = a, b, var a,
 * var b, chi2 of the inner linear fit (LinearLeastSquares.cpp:26-44). */
int ausaxs_fitter_evaluate(ausaxs_fitter* f, const double* params, int n_sets, double* out);
/* SmartFitter::fit (SmartFitter.cpp:148-170) with mini::Golden or mini::Scan
 * (Golden.cpp:22-70, Scan.cpp:26-59; one free parameter at most, as the
 * reference's minimizers without dlib). params / param_err: n_params values;
 * lin: a, b, a_err, b_err, chi2; I_fit and residuals: n values (optional). */
int ausaxs_fitter_fit(ausaxs_fitter* f, int algorithm, int scan_evals, double* params, double* param_err,
                      double* lin, int* dof, int* fevals, double* I_fit, double* residuals);
int ausaxs_fitter_free(ausaxs_fitter* f);

/* ---- synthetic structures (deterministic, xoshiro256** seeded by SplitMix64) */

/* Self-avoiding 3.8 A chain packed in an ellipsoid with semi-axes (ax, ay, az),
 * minimum non-bonded distance r_excl; elements C 62% / N 17% / O 20% / S 1%
 * written as form-factor types (C=3, N=7, O=11, S=12) in the fourth slot. */
int ausaxs_synth_globule(uint64_t seed, int n, double ax, double ay, double az, double r_excl, float* out);
/* m waters on the ellipsoid surface pushed out by offset along the normal;
 * fourth slot = WATER type (1). */
int ausaxs_synth_shell(uint64_t seed, int m, double ax, double ay, double az, double offset, float* out);

#ifdef __cplusplus
}
#endif
#endif
