/*
 * dida_letkf.h -- C ABI of the MI355X LETKF local-analysis engine.
 *
 * Drop-in boundary for the per-grid-point local analysis of the DIDA DA
 * cycle.  All entry points are extern "C", take plain pointers and sizes,
 * and return an int status (DIDA_OK = 0) unless they replace a Fortran
 * subroutine, in which case they keep its void signature.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   letkf_core_        code/x86/TradDA/DA/LETKF/src/letkf_core.F90:14-15
 *   letkf_core_batch_  code/sw/TradDA/DA/LETKF/src/letkf_core.F90:16-17
 *   letkf_core_r4_, letkf_core_batch_r4_  the same with yb_type = real(4)
 *                      (code/x86/TradDA/DA/utils/da_config.inc:7-10)
 *   dida_letkf_sweep   the per-cube loop of letkf_compute_nooverlap,
 *                      code/x86/TradDA/DA/LETKF/src/letkf_compute.F90:12-126
 *                      (gen_yb_dis :72, letkf_core :97, xb_2_xa :111)
 *   dida_letkf_post    the post-analysis blend and spreads of run_da_parallel,
 *                      code/x86/TradDA/DA/EnDNN/src/module_da_parallel.F90:216-259
 *   dida_struct_*      struct_atm2endnn / struct_endnn2atm (and _a),
 *                      code/x86/TradDA/DA/EnDNN/src/data_trans.F90:2-57
 */
#ifndef DIDA_LETKF_H
#define DIDA_LETKF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DIDA_OK 0
#define DIDA_WARN_NOCONV 1      /* some eigensolves hit the sweep cap */
#define DIDA_ERR_ARG (-1)       /* invalid argument */
#define DIDA_ERR_HIP (-2)       /* HIP runtime error */
#define DIDA_ERR_NOMEM (-3)     /* device allocation failed */
#define DIDA_ERR_CAPACITY (-4)  /* selection lists exceeded the workspace */
#define DIDA_ERR_UNSUPPORTED (-5)
#define DIDA_ERR_NONFINITE (-6) /* some analyses met a non-finite Yb, d or R^-1: their
                                   transforms (and so their columns of Xa', xbar_a)
                                   are NaN; the others are unaffected */

/* Precision of the Yb^T R^-1 Yb contraction. */
#define DIDA_GRAM_FP64 0        /* FP64 MFMA (the plain drop-in) */
#define DIDA_GRAM_BF16X3 1      /* BF16 MFMA, FP32 accumulate, split hi/lo x3 */
#define DIDA_GRAM_FP16X3 2      /* FP16 MFMA, FP32 accumulate, split hi/lo x3 */
#define DIDA_GRAM_BF16 3        /* BF16 MFMA, FP32 accumulate, one pass */

/* Execution form of a sweep. */
#define DIDA_FORM_PLAIN 0       /* separate transform / update kernels, grid order */
#define DIDA_FORM_BATCHED 1     /* chains of adjacent analyses from a work queue, each
                                   eigensolve warm-started from the previous one: one
                                   fused wavefront per analysis (k <= 64), chained
                                   workgroups (64 < k <= 128) */

typedef struct dida_letkf_params {
    int32_t nanals;             /* k, ensemble size (letkf.F90:24 ens) */
    int32_t nrow;               /* n, state rows per column (letkf.F90:26) */
    int32_t nlat, nlon;         /* grid: column c = ilat + nlat * ilon (nx, ny) */
    int32_t max_obs;            /* MAX_OBS (letkf.F90:155; 10000) */
    int32_t gc_filter;          /* 1: Gaspari-Cohn schr2, 0: boxcar schr */
    double local_dist;          /* L in km (da_loc_distance) */
    double obs_polar;           /* obs with |lat| > obs_polar are discarded */
    double inflation_factor;    /* rho, multiplicative (da_inflation_factor) */
    int32_t rtps;               /* 1: relax-to-prior-spread on Xa' with rtps_alpha */
    int32_t form;               /* DIDA_FORM_* */
    double rtps_alpha;
    int32_t gram;               /* DIDA_GRAM_* */
    int32_t jacobi_max_sweeps;  /* eigensolver sweep cap */
    double jacobi_tol;          /* rotate while |a_p.a_q| > tol |a_p||a_q|
                                   (0: 1e-13 with an FP64 Gram, 1e-8 otherwise) */
    double bin_deg;             /* obs bin size in degrees (0: automatic) */
    int32_t chain_len;          /* batched form: adjacent columns per work item,
                                   each warm-starting the next one's eigensolve
                                   (0: automatic) */
    int32_t warm_start;         /* batched form: 1 to warm-start along chains */
    int32_t jacobi_predict;     /* 1 ends the eigensolve after a sweep whose largest
                                   |a_p.a_q|^2 / (|a_p|^2 |a_q|^2) is <= tol (the
                                   next sweep would rotate nothing if convergence
                                   were quadratic); 0 runs until a sweep rotates
                                   nothing; -1 (default): 1 with a half-precision
                                   Gram, 0 with the FP64 Gram, whose 1e-10 bar
                                   the prediction can miss on cold starts with
                                   clustered eigenvalues */
    /* Vertical localisation (letkf_compute_nooverlap_vertical,
     * letkf_compute.F90:358-526; letkf.F90:228 picks it when vert_dist > 0):
     * one analysis per (column, cube of cube_nz levels), rloc scaled by
     * schr2 (gc_filter_vertical 1) or schr (0) of |ln p_cube - ln p_obs| /
     * vert_dist.  Batched form only; needs dida_letkf_set_levels and obs ph. */
    double vert_dist;           /* da_vert_distance, in ln p; <= 0: off (default) */
    int32_t cube_nz;            /* da_vertical_interval, levels per cube (default 1,
                                   as namelists.input_100km_32mem:69) */
    int32_t gc_filter_vertical; /* da_gc_enable_vertical (default 1, as :67) */
    /* Horizontal cubes (letkf_compute_nooverlap with cube_n > 1,
     * letkf_compute.F90:12-126; letkf_compute for overlap > 0, :128-228 with
     * xb_add2_xa :304-354): cubes of cube_n x cube_n columns whose starts
     * advance by cube_n - overlap (cube_n without overlap) from the first
     * column, clipped at the last; one analysis per cube, selected around the
     * cube's lower-middle column (gen_yb_dis, letkf_yb.F90:19-20); every
     * column takes the mean of the analyses of the cubes covering it.  Both
     * forms; with vertical localisation overlap = 0 (as the reference's
     * vertical sweep, letkf_compute.F90:408-425: one analysis per horizontal
     * and vertical cube, levels over the centre column's surface pressure). */
    int32_t cube_n;             /* da_block_interval (default 1, as
                                   namelists.input_100km_32mem:68) */
    int32_t overlap;            /* da_block_overlap, 0 <= overlap < cube_n (default 0) */
    int32_t xcd_tiles;          /* batched form: 1 drains the chains by longitude stripes,
                                   one stripe per accelerator die (XCD) first, so that
                                   concurrent analyses of a die share its L2; 0 (default)
                                   one queue, most expensive chains first */
    int32_t bucket_chain[4];    /* batched form, k <= 64: > 0 gives analyses of local
                                   count p >= 0, 64, 512, 2048 (the four buckets) their
                                   own chain length (1: cold starts; 0: chain_len);
                                   all 0 (default): chain_len for every analysis */
    int32_t pad_;
} dida_letkf_params;

/* Observations, device pointers, one entry per scalar observation (the
 * reference's obs_list flattened in its (obs, level, variable) order,
 * read_obs_mod.F90:15-27). */
typedef struct dida_letkf_obs {
    int32_t nobs;
    int32_t pad_;
    const double *lat, *lon;    /* degrees */
    const double *yb;           /* [nobs * k] H(Xb) - mean, ensemble fastest (pro_yb) */
    const double *dep;          /* [nobs] yo - mean H(Xb) (dis_yo) */
    const float *rinv;          /* [nobs] R^-1, REAL(4) as in the reference */
    const double *ph;           /* [nobs] obs pressure (Pa), read only with vertical
                                   localisation (obs_list ph, read_obs_mod.F90:15-27) */
} dida_letkf_obs;

/* Diagnostics of the last sweep (dida_letkf_stats synchronises). */
typedef struct dida_letkf_stats {
    int64_t nanal;              /* analyses in the sweep */
    int64_t nobsl_sum;          /* sum of local obs counts */
    int32_t nobsl_max;
    int32_t nobsl_capped;       /* analyses truncated at MAX_OBS */
    int32_t nonconverged;       /* eigensolves that hit the sweep cap */
    int32_t max_sweeps_used;
    int32_t capacity_overflow;  /* 1 if the selection lists did not fit */
    int32_t nonfinite;          /* analyses with a non-finite M or b (info -2) */
    int64_t sweeps_sum;         /* Jacobi sweeps summed over converged analyses */
} dida_letkf_stats;

typedef struct dida_letkf_ctx dida_letkf_ctx;

/* Defaults: max_obs 10000, gc_filter 1, obs_polar 90, rho 1, plain FP64,
 * automatic Jacobi tolerance, 30 sweeps, automatic bins, automatic chains
 * (8 columns, or 32 analyses for k > 64) with warm start (batched form). */
void dida_letkf_params_default(dida_letkf_params *prm);

/* ------------------------------------------------------------------------ */
/* letkf_core_ : per-analysis drop-in for the external Fortran subroutine
 * letkf_core (code/x86/TradDA/DA/LETKF/src/letkf_core.F90:14-15), same
 * by-reference arguments and meaning.  Host pointers.  hxens is
 * (nanals, nobsl) ensemble fastest; trans(i, j) = W_ij + wbar_i at
 * trans[i + j * nanals].  Runs on the current HIP device; errors are printed
 * to stderr (the reference has no status argument). */
void letkf_core_(const int32_t *nanals, const int32_t *nobsl, const double *hxens,
                 const float *rdiaginv, const double *dep, const double *rloc, double *trans,
                 double *overline_w, const double *inflation_factor, const int32_t *pflag);

/* letkf_core_batch_ : drop-in for the Sunway batched core
 * (code/sw/TradDA/DA/LETKF/src/letkf_core.F90:16-17).  Host pointers:
 * hxens (nanals, max_obs, batch), rdiaginv/dep/rloc (max_obs, batch),
 * trans (nanals, nanals, batch), overline_w (nanals, batch).  Unlike the
 * reference, hxens is left unmodified. */
void letkf_core_batch_(const int32_t *batch, const int32_t *batch_large_matrix,
                       const int32_t *max_obs, const int32_t *nanals, const int32_t *nobsl,
                       double *hxens, const float *rdiaginv, const double *dep,
                       const double *rloc, double *trans, double *overline_w,
                       const double *inflation_factor, const int32_t *pflag);

/* REAL(4) forms of the two drop-ins: the reference's letkf_core and
 * letkf_core_batch compiled with yb_type = ayb_type = real(4)
 * (code/x86/TradDA/DA/utils/da_config.inc:7-10; letkf_core.F90:14-29 and
 * code/sw/TradDA/DA/LETKF/src/letkf_core.F90:16-35 with those types), where
 * hxens, dep, rloc, trans and overline_w are REAL(4) (rdiaginv is REAL(4) in
 * both builds).  Same arguments and layouts otherwise.  The reference then
 * computes in single precision (sgemm, ssyevd); here the values are analysed
 * in FP64 and the results rounded once to REAL(4), so they agree with the
 * FP64 analysis of the same values to one REAL(4) rounding and with the
 * reference's single-precision arithmetic to its own rounding error. */
void letkf_core_r4_(const int32_t *nanals, const int32_t *nobsl, const float *hxens,
                    const float *rdiaginv, const float *dep, const float *rloc, float *trans,
                    float *overline_w, const double *inflation_factor, const int32_t *pflag);
void letkf_core_batch_r4_(const int32_t *batch, const int32_t *batch_large_matrix,
                          const int32_t *max_obs, const int32_t *nanals, const int32_t *nobsl,
                          float *hxens, const float *rdiaginv, const float *dep,
                          const float *rloc, float *trans, float *overline_w,
                          const double *inflation_factor, const int32_t *pflag);

/* Status-returning batched core on device pointers (same layout as
 * letkf_core_batch_), enqueued on `stream` (a hipStream_t, NULL = default).
 * info[batch] (optional) receives the Jacobi sweep count, -1 when the
 * eigensolve hit the sweep cap, or -2 when M or b was not finite (trans and
 * overline_w are then NaN). */
int dida_letkf_core_batch_dev(int32_t batch, int32_t max_obs, int32_t nanals,
                              const int32_t *nobsl, const double *hxens, const float *rdiaginv,
                              const double *dep, const double *rloc, double *trans,
                              double *overline_w, double inflation_factor, int32_t gram,
                              int32_t *info, void *stream);

/* The same batched core without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory -- code recorded
 * in a HIP graph, tight serving loops, harnesses that forbid hipMalloc.  The
 * per-analysis obs offsets that dida_letkf_core_batch_dev stages in a buffer
 * it allocates live here in a fixed array of the library on each device (8
 * MiB, held until the process ends), written by a kernel on `stream`; no host
 * copy, no synchronisation.  batch is at most 1048576 (larger batches return
 * DIDA_ERR_ARG before anything runs).  Every call on a device shares that
 * device's array: calls that overlap on one device (unordered streams) race
 * on it; calls ordered on one stream do not.  Same arguments, layouts and
 * results as dida_letkf_core_batch_dev. */
int dida_letkf_core_batch_dev_noalloc(int32_t batch, int32_t max_obs, int32_t nanals,
                                      const int32_t *nobsl, const double *hxens,
                                      const float *rdiaginv, const double *dep,
                                      const double *rloc, double *trans, double *overline_w,
                                      double inflation_factor, int32_t gram, int32_t *info,
                                      void *stream);

/* ------------------------------------------------------------------------ */
/* Whole-sweep engine: selection -> contractions -> eigensolve -> transform ->
 * update for every column of the grid, the per-cube loop of
 * letkf_compute_nooverlap (letkf_compute.F90:12-126) with cube_n = 1.
 *
 * nobs_max bounds the observation count of later sweeps (below 2^29; for
 * k <= 64 the binned rows, nobs_max x 32 or 64 values of 4 or 8 bytes, must
 * also stay below 4 GiB; else DIDA_ERR_UNSUPPORTED); sel_capacity is the
 * number of (column, obs) selection entries to reserve (0: sized by the
 * sweeps).  Every eager sweep reads its number of selection entries back once
 * (one stream synchronisation, after the count pass).  Sized by the sweeps,
 * the workspace is allocated by the first and grows (by 1/8 more) when a
 * later one needs more; growing replaces its buffers, so a graph captured on
 * the context before is dropped (capture again) and a group graph holding the
 * context refuses to launch.  With a reserved workspace the plain form and
 * the batched form with k > 64 run each sweep in runs of consecutive
 * analyses whose lists fit it, so workloads whose selection lists exceed the
 * device memory left next to the state still run; the fused batched form
 * (k <= 64) returns DIDA_ERR_CAPACITY before any analysis runs.  A captured
 * graph whose lists outgrow the workspace on replay analyses the columns past
 * it without obs and dida_letkf_stats_get reports DIDA_ERR_CAPACITY. */
int dida_letkf_create(dida_letkf_ctx **ctx, const dida_letkf_params *prm, int32_t nobs_max,
                      int64_t sel_capacity);
int dida_letkf_destroy(dida_letkf_ctx *ctx);

/* Grid coordinates (REAL(4), degrees, device pointers): grid_lat[nlat],
 * grid_lon[nlon] (pro_info.F90:11-12). */
int dida_letkf_set_grid(dida_letkf_ctx *ctx, const float *grid_lat, const float *grid_lon);

/* One analysis sweep.  Device pointers: xb = Xb' (k, n, nlat, nlon) ensemble
 * fastest, axb = xbar_b (n, nlat, nlon); outputs xa = Xa' (re-centred) and
 * axa = xbar_a in the same layouts (letkf_run, letkf.F90:157).  Enqueued on
 * `stream`; no host synchronisation once the selection workspace is sized. */
int dida_letkf_sweep(dida_letkf_ctx *ctx, const dida_letkf_obs *obs, const double *xb,
                     const double *axb, double *xa, double *axa, void *stream);

/* The same sweep with the ensemble perturbations stored in FP32 (the
 * state type of the reference is the xb_type macro of
 * code/x86/TradDA/DA/utils/da_config.inc:4-5): xb = Xb'
 * and xa = Xa' as float in the layout above, means axb, axa FP64; the update
 * reads Xb' once, computes in FP64 and rounds Xa' once.  Plain form (any k)
 * or batched form with 64 < k <= 128, horizontal cubes included; the fused
 * batched form (k <= 64) returns DIDA_ERR_UNSUPPORTED.  Runs eagerly (no
 * graph capture form). */
int dida_letkf_sweep_f32(dida_letkf_ctx *ctx, const dida_letkf_obs *obs, const float *xb,
                         const double *axb, float *xa, double *axa, void *stream);

/* Capture one sweep into a HIP graph owned by the context (inputs and
 * outputs fixed at capture), then replay it with dida_letkf_graph_launch. */
int dida_letkf_graph_capture(dida_letkf_ctx *ctx, const dida_letkf_obs *obs, const double *xb,
                             const double *axb, double *xa, double *axa, void *stream);
int dida_letkf_graph_launch(dida_letkf_ctx *ctx, void *stream);

/* Autotuner.  Times sweeps of the caller's workload (inputs as for
 * dida_letkf_sweep) over candidate layouts and keeps the fastest for the
 * sweeps that follow: the length of the warm-start chains of the batched form
 * (2, 4, 8, 16, 32, 64 adjacent columns per work item: longer chains warm-start
 * more eigensolves, shorter ones balance the cost-sorted queue better) and
 * the obs bin size (theta/2, theta/4, theta/8, theta = support radius:
 * coarser bins scan more candidates, finer ones visit more bins; a column
 * capped at MAX_OBS keeps the obs of smallest caller index whatever the bin
 * size, as the reference's in-order scan does); then, in the fused batched
 * form (k <= 64), a chain length per local-obs-count bucket (p < 64, < 512,
 * < 2048, >= 2048; 1, 2, 4, ..., 64, where 1 starts every analysis of the
 * bucket cold), one bucket at a time: chains are then runs of adjacent
 * columns of one bucket, built on the device every sweep from its counts,
 * drained from one cost-sorted queue.  First of all it picks the tile of
 * columns (4 x 1, 4 x 2 or 4 x 4) that one wavefront of the selection passes
 * tests each candidate obs against.  Each candidate is
 * timed twice (HIP events, minimum); xa/axa hold the result of a sweep with
 * the kept layout on return.  Synchronises. */
typedef struct dida_letkf_tuning {
    int32_t chain_len;          /* kept (batched form; 0 for the plain form) */
    int32_t ncand;              /* layouts timed */
    double bin_deg;             /* kept */
    double ms_initial;          /* sweep time with the layout on entry and the kept
                                   selection tile (chosen first, by the binning and
                                   selection stages alone) */
    double ms_best;             /* sweep time with the kept layout */
    int32_t xcd_tiles;          /* kept queue order (batched form) */
    int32_t bucketed;           /* 1: chain length per obs-count bucket (below) */
    int32_t bucket_lo[4];       /* buckets: analyses with local count p >= bucket_lo[b]
                                   (and below the next bound) */
    int32_t bucket_chain[4];    /* kept chain length per bucket (1: cold starts) */
    int64_t bucket_count[4];    /* analyses per bucket in the tuned sweep */
    int32_t select_tile;        /* kept tile of columns of the selection passes:
                                   0: 4 x 2, 1: 4 x 1, 2: 4 x 4 (latitudes x longitudes) */
    int32_t pad_;
} dida_letkf_tuning;
int dida_letkf_autotune(dida_letkf_ctx *ctx, const dida_letkf_obs *obs, const double *xb,
                        const double *axb, double *xa, double *axa, void *stream,
                        dida_letkf_tuning *out);

/* Several analysis domains in one HIP graph (the batching form's regional
 * case: independent domains of different sizes and ensemble sizes, each with
 * its own context).  The graph forks from `stream` into one branch per
 * context, runs one sweep of each with the given inputs and outputs, and
 * joins back, so the domains run concurrently; dida_letkf_group_launch
 * replays it.  Every context must have run one eager sweep first (it sizes
 * the selection lists).  Contexts must not be used elsewhere while their
 * group graph runs.  No counterpart in the reference, which runs domains as
 * separate jobs. */
typedef struct dida_letkf_group dida_letkf_group;
int dida_letkf_group_capture(dida_letkf_group **grp, int32_t n, dida_letkf_ctx *const *ctx,
                             const dida_letkf_obs *obs, const double *const *xb,
                             const double *const *axb, double *const *xa,
                             double *const *axa, void *stream);
int dida_letkf_group_launch(dida_letkf_group *grp, void *stream);
int dida_letkf_group_destroy(dida_letkf_group *grp);

/* Timing of eager sweeps.  `on` bit 0 (DIDA_TIMING_STAGES): HIP events on
 * the sweep's stream around each stage (the reference's GPTL regions
 * "Generate Yb matrix", "Lektf computation core", "Convert xb to xa",
 * letkf_compute.F90:69-120), read by dida_letkf_stage_ms for the last sweep
 * (ms[DIDA_STAGE_*]; synchronises).  Bit 1 (DIDA_TIMING_PHASES): device
 * clocks per phase inside the batched kernel (dida_letkf_phase_cycles);
 * these add atomics to the kernel, so leave them off where it is timed. */
#define DIDA_TIMING_STAGES 1
#define DIDA_TIMING_PHASES 2
#define DIDA_STAGE_BIN 0        /* obs binning and gather */
#define DIDA_STAGE_SELECT 1     /* local selection: count, scan, fill */
#define DIDA_STAGE_TRANSFORM 2  /* contractions, eigensolve, W and wbar */
#define DIDA_STAGE_UPDATE 3     /* ensemble update */
#define DIDA_NSTAGES 4
int dida_letkf_set_timing(dida_letkf_ctx *ctx, int32_t on);
int dida_letkf_stage_ms(dida_letkf_ctx *ctx, double ms[DIDA_NSTAGES]);

/* Batched form with DIDA_TIMING_PHASES: device clock cycles summed over workgroups
 * per phase of the fused kernel (Gram, warm start, Jacobi, transform,
 * update; the reference's "eig decomposition" and core timers,
 * letkf_core.F90:94-117).  Synchronises. */
#define DIDA_NPHASES 5
int dida_letkf_phase_cycles(dida_letkf_ctx *ctx, double cyc[DIDA_NPHASES]);

/* Diagnostics of the last sweep; synchronises the context's last stream.
 * Returns DIDA_ERR_CAPACITY, else DIDA_ERR_NONFINITE, else DIDA_WARN_NOCONV
 * when the sweep had such analyses (the reference prints a failed eigensolve,
 * letkf_core.F90:119-122, and its Sunway caller stops on a NaN transform,
 * letkf_compute.F90:168-187), DIDA_OK otherwise. */
int dida_letkf_stats_get(dida_letkf_ctx *ctx, dida_letkf_stats *st);

/* Device arrays of the last sweep (valid until the next sweep):
 * nobsl[ncol], offset[ncol + 1] (int64), csr_idx / csr_w (binned obs index
 * and Gaspari-Cohn weight per selection entry), perm[nobs] (binned position
 * -> caller's obs index), info[ncol] (Jacobi sweeps, -1 not converged,
 * -2 not finite). */
int dida_letkf_debug_arrays(dida_letkf_ctx *ctx, const int32_t **nobsl, const int64_t **offset,
                            const int32_t **csr_idx, const double **csr_w, const int32_t **perm,
                            const int32_t **info);

/* ------------------------------------------------------------------------ */
/* Observation operator and centring: the step before the analysis, which
 * produces its obs arrays (letkf_obs.F90:120-178 with find_obs_grid :187-243;
 * obs_interpolation_3d_2 and ps2ph_3d_interp, obs_make.F90:233-381).
 *
 * Model levels: half-level pressure of level j = 0..nz from the member's
 * surface pressure ps (vert_coord_mod.F90:547-563):
 *   DIDA_VCOORD_SIGMA:  c1[j] (ps - ptop) + ptop
 *   DIDA_VCOORD_HYBRID: c1[j] (p0 - ptop) + c2[j] (ps - ptop) + ptop
 * Full levels are the means of adjacent half levels. */
#define DIDA_VCOORD_SIGMA 0
#define DIDA_VCOORD_HYBRID 1
typedef struct dida_vert_coord {
    int32_t type;
    int32_t nz;                 /* full levels */
    const double *c1, *c2;      /* device [nz + 1] (c2: hybrid only) */
    double ptop, p0;            /* Pa */
} dida_vert_coord;

/* Observation records (read_obs_mod.F90:15-27 obs_list): nrec records at
 * (lat, lon) with onz levels of pressure ph[r * onz + l] (Pa, ascending
 * within a record, as the reference's level search assumes) and onv
 * variables vid[v] (indices into the state variables); values
 * yo[(r * onz + l) * onv + v].  Device pointers except vid (host). */
typedef struct dida_obs_records {
    int32_t nrec, onz, onv, pad_;
    const int32_t *vid;
    const double *lat, *lon, *ph, *yo;
} dida_obs_records;

/* H(Xb) for every member and the centring: outputs the analysis's scalar obs
 * s = (r * onz + l) * onv + v: lat[s], lon[s], yb[s * k + e] = H(x_e) -
 * mean_e H(x_e) (pro_yb), dep[s] = yo[s] - mean_e H(x_e) (dis_yo), and, if ph is not NULL,
 * ph[s] = the record's pressure at level l (for vertical localisation); R^-1 of
 * the same scalars is the caller's.  State as for dida_letkf_sweep (Xb' and
 * xbar_b, device); state variable v occupies rows var_start[v] ..
 * var_start[v] + var_lev[v] - 1 (var_lev 1: a 2-D variable; host arrays,
 * nvar <= 32, onv <= 32); surface pressure at row ps_row.  Grid brackets as
 * find_obs_grid (first grid point >= the obs); outside the first / last grid
 * latitude the nearest row; periodic_lon wraps the longitude.  k <= 128.
 * Enqueued on `stream`. */
int dida_letkf_obs_operator(int32_t k, int32_t nrow, int32_t nlat, int32_t nlon,
                            const float *grid_lat, const float *grid_lon, int32_t periodic_lon,
                            const double *xb, const double *axb, int32_t ps_row,
                            const dida_vert_coord *vc, int32_t nvar, const int32_t *var_start,
                            const int32_t *var_lev, const dida_obs_records *rec, double *lat,
                            double *lon, double *yb, double *dep, double *ph, void *stream);

/* Model levels for vertical localisation (params vert_dist > 0): the
 * coordinate table (device arrays, kept by reference), the surface-pressure
 * row and the variable layout (host arrays, copied): variable v occupies rows
 * var_start[v] .. var_start[v] + var_lev[v] - 1, var_lev[v] = 1 (2-D) or
 * vc->nz (3-D), every row of the column exactly once (var_info.F90:22).
 * Cube q = levels sk..ek (sk = q cube_nz) updates rows var_start + sk..ek of
 * every 3-D variable; the last cube also updates the 2-D variables
 * (letkf_compute.F90:438-520).  The cube's ln p is that of full level
 * (sk + ek) / 2 (1-based) over the column's mean surface pressure
 * (ps2alnp, obs_make.F90:418-437, whose unset bottom half level ph_lev(nz+1)
 * is here taken from the table).  Call before the first sweep; calling it
 * again replaces the tables and drops a captured graph of this context
 * (capture again). */
int dida_letkf_set_levels(dida_letkf_ctx *ctx, const dida_vert_coord *vc, int32_t ps_row,
                          int32_t nvar, const int32_t *var_start, const int32_t *var_lev);

/* ------------------------------------------------------------------------ */
/* Post-analysis (run_da_parallel, module_da_parallel.F90:216-259): inside
 * |grid_lat| <= polar_lat, xa_out = gamma1 (xbar_b + Xb') + gamma2 (xbar_a +
 * Xa'); outside, the background; spread_b/a = sqrt(sum_e (.)^2 / k).
 * Device pointers; xa_out is (k, n, nlat, nlon), spreads (n, nlat, nlon). */
int dida_letkf_post(int32_t k, int32_t nrow, int32_t nlat, int32_t nlon, const float *grid_lat,
                    double polar_lat, double gamma1, double gamma2, const double *xb,
                    const double *axb, const double *xa, const double *axa, double *xa_out,
                    double *spread_b, double *spread_a, void *stream);

/* ------------------------------------------------------------------------ */
/* Coupler layout conversions (data_trans.F90:2-57, called per member group
 * and variable at module_da_coupler.F90:378 and :405).  The coupler's
 * member-fastest block xb_atm(ens, ny, nx, n) -- ny longitudes, nx latitudes,
 * n levels of one variable, as unpacked from the MCT attribute vector -- is
 * copied into, or out of, the analysis state xb(ens_max, n_max, nx, ny) at
 * members ens_start .. ens_start + ens - 1 and rows n_start .. n_start + n - 1
 * (1-based, as the reference passes them); state entries outside the block are
 * left as they are.  The _a forms move the means axb_atm(ny, nx, n) <->
 * axb(n_max, nx, ny).  Device pointers, enqueued on `stream`. */
int dida_struct_atm2endnn(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                          int32_t n_max, int32_t ens_start, int32_t n_start,
                          const double *xb_atm, double *xb, void *stream);
int dida_struct_endnn2atm(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                          int32_t n_max, int32_t ens_start, int32_t n_start, const double *xb,
                          double *xb_atm, void *stream);
int dida_struct_atm2endnn_a(int32_t n, int32_t nx, int32_t ny, int32_t n_max, int32_t n_start,
                            const double *axb_atm, double *axb, void *stream);
int dida_struct_endnn2atm_a(int32_t n, int32_t nx, int32_t ny, int32_t n_max, int32_t n_start,
                            const double *axb, double *axb_atm, void *stream);

/* Last error message of this thread (static storage). */
const char *dida_letkf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
