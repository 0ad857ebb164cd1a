/*
 * Thin C ABI between the C host layer (letkf_host.c) and the HIP kernels.
 * Every launcher enqueues on `stream` (a hipStream_t) and returns a
 * hipError_t value as int; nothing here synchronises or allocates, except
 * where a function says it only queries a size.
 */
#ifndef DIDA_LETKF_KERNELS_H
#define DIDA_LETKF_KERNELS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Observation bins: nband bands of dlat = 180/nband degrees and nlonbin
 * longitude bins of dlon = 360/nlonbin degrees (an exact tiling, so bin
 * indices wrap at the date line); band = floor((lat + 90) / dlat), lon bin =
 * floor(wrap360(lon) / dlon); |lat| > obs_polar goes to bin nbins. */
typedef struct dk_bin_geom {
    double dlat, dlon;
    double obs_polar;
    int32_t nband, nlonbin, nbins;
    int32_t pad_;
} dk_bin_geom;

/* Observations in binned order, with the per-obs factors of the distance
 * formula precomputed (cos/sin of the latitude in the reference's radians
 * and the longitude in the reference's radians) and the unit vector
 * (cx cos yr, cx sin yr, sx) rounded to FP32, used to decide the candidates
 * away from the support radius, and in FP64, for the weights of the obs
 * surely inside it. */
typedef struct dk_obs_binned {
    const double *cx, *sx, *yr;   /* [nobs] */
    const float *u4;              /* [nobs * 4]: unit vector x, y, z, 0 */
    const double *ux, *uy, *uz;   /* [nobs]: unit vector in FP64 */
    const double *yb;             /* [nobs * k] ensemble fastest */
    const double *dep;            /* [nobs] */
    const float *rinv;            /* [nobs] */
    const int32_t *bin_start;     /* [nbins + 2] */
} dk_obs_binned;

typedef struct dk_select_args {
    int32_t nlat, nlon;
    const float *grid_lat, *grid_lon; /* REAL(4) grid coordinates, degrees */
    dk_bin_geom geom;
    dk_obs_binned obs;
    double local_dist;   /* L, km */
    double theta_deg;    /* angular radius of the support plus margin, deg */
    double cos_far;      /* central-angle cosine below which gc = 0 surely */
    double cos_near;     /* central-angle cosine above which gc > 0 surely */
    double cos_keep;     /* cos(2 L / a): kept iff the central-angle cosine >= it */
    int32_t gc_filter;
    int32_t max_obs;
    int32_t *nobsl;      /* [ncol] count pass output / fill pass input */
    const int64_t *offset; /* [ncol + 1] fill pass */
    int32_t *csr_idx;    /* fill pass: binned obs index */
    double *csr_w;       /* fill pass: rloc */
    int64_t capacity;    /* fill pass: entries available in csr_idx / csr_w */
    int32_t *overflow;   /* fill pass: set to 1 when a column does not fit */
    /* fill pass over analyses [col0, col0 + ncols) only (ncols 0: all),
     * writing entry offset[col] - base */
    int64_t col0, ncols, base;
    /* MAX_OBS truncation in the caller's obs order: perm[l] = caller index of
     * binned obs l (< 2^perm_bits); thr[col] (cap pass output, fill pass
     * input) = one past the largest caller index a column capped at max_obs
     * keeps, INT32_MAX for a column that keeps every obs with gc > 0 */
    const int32_t *perm;
    const int32_t *iperm; /* [nobs]: binned position of caller index i (cap pass) */
    int32_t *thr;
    int32_t perm_bits;
    int32_t nobs;
    int32_t tile;        /* count and fill: DK_SELECT_TILE_* columns per wavefront */
    /* count and fill over all columns: wavefront w of the launch takes tile
     * tile_order[w] (largest candidate count first); NULL: grid order */
    const int32_t *tile_order;
} dk_select_args;

/* Tile shapes (latitudes x longitudes of the analysis grid) of the count and
 * fill passes. */
#define DK_SELECT_TILE_4X2 0
#define DK_SELECT_TILE_4X1 1
#define DK_SELECT_TILE_4X4 2

/* Fill n int32 with v (a kernel, so that captured graphs hold no memset). */
int dk_fill_i32(int32_t *p, int32_t v, int64_t n, void *stream);

/* Stage 1: keys and identity values for the sort. */
int dk_bin_keys(int32_t nobs, const double *lat, const double *lon, dk_bin_geom g,
                uint32_t *key, int32_t *iota, void *stream);
/* Stable radix sort of (key, iota) by key.  tmp == NULL: *tmp_bytes is set. */
int dk_sort_pairs_u32(void *tmp, size_t *tmp_bytes, const uint32_t *kin, uint32_t *kout,
                      const int32_t *vin, int32_t *vout, int32_t n, int end_bit, void *stream);
/* bin_start[b] = number of obs with key < b, b = 0 .. nbins + 1, from the
 * nobs sorted keys (unread when nobs is 0). */
int dk_bin_starts(int32_t nbins, int32_t nobs, const uint32_t *key_sorted, int32_t *bin_start,
                  void *stream);
/* Gather obs into binned order and precompute the distance factors.  yb
 * rows are copied into rows of ldy >= k values, zero past k: in FP64 into
 * yb_b and/or rounded to FP32 into yb32_b (either may be NULL). */
int dk_obs_gather(int32_t nobs, int32_t k, int32_t ldy, const int32_t *perm, const double *lat,
                  const double *lon, const double *yb, const double *dep, const float *rinv,
                  double *cx, double *sx, double *yr, float *u4, double *ux, double *uy,
                  double *uz, int32_t *iperm, double *yb_b, float *yb32_b, double *dep_b,
                  float *rinv_b, void *stream);
/* Selection passes: count (capped at max_obs) and fill (entries with gc > 0
 * and perm < thr, in bin order) by tiles of columns (DK_SELECT_TILE_*), one
 * wavefront a tile reading each candidate once for its columns; cap (thr of every
 * column; a column whose count reached max_obs scans all its candidates for
 * the caller-order threshold by a radix select over perm), one wavefront a
 * column. */
/* Tiles of the count and fill passes (shape a->tile) over the analysis grid. */
int64_t dk_select_ntiles(int32_t nlat, int32_t nlon, int32_t tile);
/* order[0 .. ntiles) = the tiles by the obs count of their candidate bins,
 * largest first (so that the longest tiles do not start last); key, id and
 * key_sorted hold ntiles entries; tmp as dk_sort_pairs (query with tmp NULL:
 * *tmp_bytes for ntiles). */
int dk_select_tile_order(const dk_select_args *a, int32_t ntiles, uint32_t *key, int32_t *id,
                         uint32_t *key_sorted, int32_t *order, void *tmp, size_t *tmp_bytes,
                         void *stream);
int dk_select_count(const dk_select_args *a, void *stream);
int dk_select_cap(const dk_select_args *a, void *stream);
int dk_select_fill(const dk_select_args *a, void *stream);
/* offset[0] = 0, offset[i+1] = offset[i] + nobsl[i]; tmp as dk_sort_pairs. */
int dk_scan_offsets(void *tmp, size_t *tmp_bytes, const int32_t *nobsl, int64_t *offset,
                    int32_t n, void *stream);

/* Transform (letkf_core) for a set of analyses. */
typedef struct dk_core_args {
    int32_t k;              /* nanals */
    int32_t nanal;          /* analyses in this launch */
    const int32_t *list;    /* analysis ids (NULL: aid0 + block index) */
    int32_t aid0;           /* first analysis id; trans/wbar hold id - aid0 */
    int32_t gram;           /* DK_GRAM_*: FP64 (every kernel); half-precision ones (chained) */
    const int64_t *offset;  /* obs offset of each analysis */
    const int32_t *nobsl;   /* obs count of each analysis */
    int64_t capacity;       /* entries readable at idx/rloc; an analysis whose
                               range does not fit is treated as having no obs */
    int32_t max_p;          /* obs counts are clamped to [0, max_p] */
    int32_t pad1_;
    const int32_t *idx;     /* CSR obs index (NULL: dense, index = offset + o) */
    const double *rloc;     /* [offset + o] */
    const double *yb;       /* obs ensemble perturbations, k per obs */
    const float *rinv;      /* per obs index */
    const double *dep;      /* per obs index */
    double rho;             /* multiplicative inflation */
    double tol;             /* Jacobi: rotate while |g| > tol sqrt(a b) */
    int32_t max_sweeps;
    int32_t predict;        /* Jacobi: stop when the next sweep would rotate nothing */
    int64_t base;           /* idx / rloc hold entries offset[id] - base */
    double *trans;          /* [(id - aid0) * k * k] trans(i, j) at i + j k */
    double *wbar;           /* [(id - aid0) * k] (optional) */
    int32_t *info;          /* [id]: sweeps, or -1 when not converged (optional) */
    const float *yb32;      /* half-precision Grams (chained): yb rounded to FP32,
                               read instead of yb, rows ldy32 apart (a multiple of 4,
                               zero past k) */
    int32_t ldy32;
    int32_t pad2_;
    const double *ymax;     /* FP16X3 (chained): max_e |yb| sqrt(R^-1) per obs index */
} dk_core_args;

int dk_letkf_core_f64(const dk_core_args *a, void *stream);

/* Sets the thread's last error message (dida_letkf_last_error) and returns
 * code; for entry points implemented in HIP sources. */
int dk_set_error(int code, const char *msg);

/* Chained analyses (64 < k <= 128): chains of chain_len consecutive
 * analyses of [aid0, aid0 + nanal), drained by nblocks persistent
 * workgroups from *head (zero before the launch); each workgroup keeps the
 * last eigenvectors in vscr + blockIdx * 128 * 128 and warm-starts the next
 * analysis of its chain from them. */
typedef struct dk_chain_args {
    int32_t chain_len, nchain;
    int32_t *head;
    double *vscr;
    unsigned long long *phase_clk; /* optional [DK_NPHASE] device clock sums
                                      (gram, warm, jacobi, transform) */
    const int32_t *order;          /* optional: chain taken at queue position q */
} dk_chain_args;
/* Sort keys of the chains of one chained launch: 2^32-1 - (sum of the local
 * obs counts of the chain's analyses + 10 k each), so that an ascending
 * sort drains the most expensive chains first. */
int dk_chain_costs(const dk_core_args *a, int32_t chain_len, int32_t nchain, uint32_t *key,
                   int32_t *id, void *stream);
int dk_letkf_chain_f64(const dk_core_args *a, const dk_chain_args *ch, int nblocks,
                       void *stream);

/* Update (xb_2_xa) for a set of columns, with optional RTPS. */
typedef struct dk_update_args {
    int32_t k, nrow;
    int32_t ncol;           /* columns in this launch */
    int32_t col0;           /* first column id (identity list) */
    const double *trans;    /* [(col - col0) * k * k] */
    const double *xb;       /* [col * nrow * k] */
    const double *axb;      /* [col * nrow] */
    double *xa;
    double *axa;
    int32_t rtps;           /* 1: relax to prior spread with alpha */
    int32_t pad_;
    double rtps_alpha;
    /* horizontal cubes (cube_n > 1): column col = ilat + nlat ilon takes the
     * mean of the transforms trans[a * k * k] of the cubes covering it, a =
     * aj + nalat ai over cube starts aj * step (latitude), ai * step
     * (longitude), each cube cube_n columns wide (clipped at the last) */
    int32_t nlat, nlon, cube_n, step, nalat, nalon;
    /* FP32 perturbations (dk_update_f64 only): when xa32 is set, Xb' is read
     * from xb32 and Xa' written to xa32 instead of xb / xa */
    const float *xb32;
    float *xa32;
} dk_update_args;

int dk_update_f64(const dk_update_args *a, void *stream);
/* The same update with FP64 MFMA, one wavefront per column (k <= 64). */
int dk_update_cols(const dk_update_args *a, void *stream);

/* Centres of horizontal cubes: ca_lat[aj] = grid_lat[c(aj)], c(aj) the
 * 0-based lower middle of columns aj * step .. min(aj * step + cube_n, nlat) - 1
 * (gen_yb_dis, letkf_yb.F90:19-20); likewise ca_lon. */
int dk_cube_centres(int32_t nlat, int32_t nlon, int32_t cube_n, int32_t step,
                    const float *grid_lat, const float *grid_lon, float *ca_lat, float *ca_lon,
                    void *stream);

/* Precision of the Gram in the batched form (values of DIDA_GRAM_*). */
#define DK_GRAM_FP64 0
#define DK_GRAM_BF16X3 1
#define DK_GRAM_FP16X3 2
#define DK_GRAM_BF16 3

/* Batched form: chains of chain_len adjacent columns along a meridian,
 * drained from a queue by a persistent kernel. */
typedef struct dk_batched_args {
    int32_t k, nrow, nlat, nlon;
    int32_t nchain, chains_per_lon, chain_len, gram;
    int32_t warm_start, rtps, max_sweeps, max_p;
    int32_t predict, ldy;   /* ldy: row stride of yb, the kernel's padded k (32 or 64) */
    double rtps_alpha, rho, tol;
    int64_t capacity;
    const int64_t *offset;
    const int32_t *nobsl;
    const int32_t *idx;     /* CSR binned obs index */
    const double *rloc;
    const double *yb;       /* binned obs rows of ldy values, zero past k (FP64 Gram) */
    const float *yb32;      /* the rows rounded to FP32 (split half-precision Grams,
                               whose hi + lo parts hold 16 to 22 bits) */
    const float *rinv;
    const double *dep;
    const double *ymax;     /* per binned obs: max |yb| sqrt(R^-1) (FP16 scaling) */
    const double *xb, *axb;
    double *xa, *axa;
    int32_t *info;          /* [col] Jacobi sweeps or -1 */
    double *wbar_out;       /* [col * k] (optional) */
    const int32_t *queue;   /* chain ids in processing order (NULL: identity) */
    int32_t *queue_head;    /* work counter, zero before the launch */
    unsigned long long *phase_clk; /* optional [DK_NPHASE] device clock sums */
    /* vertical localisation (ncube > 0): one analysis per (column, cube),
     * analysis id col * ncube + cube; rloc scaled by gc_weight(|cube_lnp -
     * obs_lnp| / vert_dist, gc_vert); cube q updates rows
     * cube_rows[cube_row_off[q] .. cube_row_off[q + 1]) */
    int32_t ncube, gc_vert;
    double vert_dist;
    const double *obs_lnp;      /* [binned obs] ln p */
    const float *sel_key;       /* per selection entry: (float) ln p, ascending within
                                   each column (idx / rloc reordered alike) */
    const double *cube_lnp;     /* [col * ncube + q] ln p of the cube's centre level */
    const int32_t *cube_row_off, *cube_rows;
    /* optional: write trans(i, e) = W + wbar 1^T of analysis col at
     * trans_out[col * k * k + i + e k] instead of updating the state */
    double *trans_out;
    /* vertical localisation with horizontal cubes (cube_n > 1): analysis
     * aj + nlat ai updates the columns of cube aj, ai of the state grid
     * grid_nlat x grid_nlon (cubes start every cube_n columns) */
    int32_t cube_n, grid_nlat, grid_nlon, pad4_;
    /* ngroup > 0: the queue holds ngroup groups of chains (longitude stripe
     * g = columns [g nlon / ngroup, (g + 1) nlon / ngroup), chains in
     * latitude-band-major order), group g starting at stripe_lo(g) *
     * chains_per_lon; workgroup b drains group b % ngroup first (one group per
     * XCD, so neighbouring analyses share an L2), then the others;
     * queue_head holds ngroup counters */
    int32_t ngroup, pad3_;
    /* per-bucket chains (bucketed != 0, one queue): chain c covers the
     * chain_ncol[c] analyses of one longitude from column chain_col0[c];
     * *nchain_dev chains (queue entries past it are never taken) */
    int32_t bucketed, pad5_;
    const int32_t *chain_col0, *chain_ncol;
    const int64_t *nchain_dev;
    /* selection entries of analysis a at offset[a] - base (runs of a sweep
     * over a fixed workspace); trans_out holds analysis a at (a - trans_aid0) k^2 */
    int64_t base, trans_aid0;
} dk_batched_args;

#define DK_XCD_GROUPS 8

/* Chains per obs-count bucket: analysis of local count p is in bucket b, the
 * last with p >= lo[b] (lo[0] = 0); a chain is a run of adjacent latitudes
 * of one longitude in the same bucket, at most len[b] long (1: every analysis
 * starts cold).  Built per sweep: count pass (line_cnt[ilon]), exclusive scan
 * of the counts into line_off[nlon + 1], write pass (descriptors). */
#define DK_NBUCKET 4
typedef struct dk_chain_build_args {
    int32_t nlat, nlon;
    int32_t nbucket, pad_;
    int32_t lo[DK_NBUCKET], len[DK_NBUCKET];
    const int32_t *nobsl;
    int32_t *line_cnt;
    const int64_t *line_off;
    int32_t *chain_col0, *chain_ncol;
} dk_chain_build_args;
int dk_chain_build(const dk_chain_build_args *a, int write, void *stream);

/* Phases of the batched kernel timed by phase_clk (wave 0 of each block). */
#define DK_PHASE_GRAM 0
#define DK_PHASE_WARM 1
#define DK_PHASE_JACOBI 2
#define DK_PHASE_TRANSFORM 3
#define DK_PHASE_UPDATE 4
#define DK_NPHASE 5

/* Sort keys of the chains: 2^32-1 - cost (ascending sort = largest first). */
int dk_chain_keys(const dk_batched_args *a, uint32_t *key, int32_t *id, void *stream);
/* Grouped queue (a->ngroup > 0): order[position] = chain, no sort needed. */
int dk_chain_groups(const dk_batched_args *a, int32_t *order, void *stream);
/* Batched kernel (one 64-thread wavefront per analysis): resident
 * workgroups per CU for (k, gram), and the persistent launch. */
int dk_wave_blocks_per_cu(int k, int gram);
int dk_letkf_wave(const dk_batched_args *a, int nblocks, void *stream);
/* Plain form, k <= 64: the transforms of analyses [aid0, aid0 + nanal) in
 * grid order, one wavefront each, FP64 Gram, cold-started register Jacobi;
 * trans (and wbar, info when set) as dk_letkf_core_f64.  The binned Yb rows
 * are KP = 32 or 64 values apart (zero past k), as for the batched form. */
int dk_letkf_plain_wave(const dk_core_args *a, void *stream);
/* ymax[i] = max_e |yb[i k + e]| sqrt(rinv[i]); the second form reads FP32 rows. */
int dk_obs_ymax(int32_t nobs, int32_t k, int32_t ldy, const double *yb, const float *rinv,
                double *ymax, void *stream);
int dk_obs_ymax32(int32_t nobs, int32_t k, int32_t ldy, const float *yb, const float *rinv,
                  double *ymax, void *stream);

/* Vertical localisation inputs (letkf_compute.F90:424-447, obs_make.F90:418-437):
 * cube_lnp[col * ncube + q] = ln of the full-level pressure of level
 * (sk + ek) / 2 (1-based) of cube q = levels sk..ek, from the column's mean
 * surface pressure axb[col * nrow + ps_row] and the half levels c1/c2. */
typedef struct dk_cube_lnp_args {
    int32_t nrow, ps_row, nz, vtype;
    int32_t ncube, cube_nz;
    int64_t ncol;               /* horizontal analyses (columns, or cubes of columns) */
    /* horizontal cubes of cube_n x cube_n columns (analysis aj + nalat ai):
     * the surface pressure is the 1-based centre column's (cube_n 1: the
     * column's own) */
    int32_t cube_n, nalat, grid_nlat, grid_nlon;
    const double *c1, *c2;
    double ptop, p0;
    const double *axb;
    double *cube_lnp;
} dk_cube_lnp_args;
int dk_cube_lnp(const dk_cube_lnp_args *a, void *stream);
/* Vertical localisation: every column's selection list [offset[c],
 * offset[c+1]) reordered by ascending obs ln p (stable): key_out = (float)
 * lnp of the entries, idx2 / w2 = the reordered idx / w.  tmp == NULL queries
 * the rocPRIM work space for `capacity` entries. */
int dk_seg_sort_lnp(void *tmp, size_t *tmp_bytes, int64_t nseg, int64_t capacity,
                    const int64_t *offset, const int32_t *idx, const double *w,
                    const double *lnp, float *key_in, float *key_out, int32_t *pos_in,
                    int32_t *pos_out, int32_t *idx2, double *w2, void *stream);
/* lnp[i] = log(ph[perm[i]]) (obs pressures in bin order). */
int dk_obs_lnp(int32_t nobs, const int32_t *perm, const double *ph, double *lnp, void *stream);

/* Observation operator H(Xb) and centring, one wavefront per record. */
#define DK_OBSOP_MAXVAR 32
typedef struct dk_obsop_args {
    int32_t k, nrow, nlat, nlon;
    int32_t periodic, ps_row, nz, vtype;   /* vtype 0 sigma, 1 hybrid */
    int32_t onz, onv;
    int64_t nrec;
    const float *grid_lat, *grid_lon;
    const double *xb, *axb;
    const double *c1, *c2;                 /* nz + 1 half levels */
    double ptop, p0;
    int32_t vstart[DK_OBSOP_MAXVAR], vlev[DK_OBSOP_MAXVAR], vid[DK_OBSOP_MAXVAR];
    const double *rec_lat, *rec_lon, *rec_ph, *yo;
    double *lat, *lon, *yb, *dep;
    double *ph;                            /* optional: scalar obs pressure */
} dk_obsop_args;

int dk_obs_operator(const dk_obsop_args *a, void *stream);

/* Post-analysis blend of run_da_parallel (RTPP-style gamma blend inside
 * |lat| <= polar_lat, background outside) and spreads. */
typedef struct dk_post_args {
    int32_t k, nrow, nlat, nlon;
    const float *grid_lat;
    double polar_lat, gamma1, gamma2;
    const double *xb, *axb, *xa, *axa;
    double *xa_out, *spread_b, *spread_a;
} dk_post_args;

int dk_post(const dk_post_args *a, void *stream);

/* Coupler layout conversions (0-based starts): to_state 1 copies the
 * coupler block into the state, 0 the state block out. */
int dk_struct_ens(int to_state, int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                  int32_t n_max, int32_t e0, int32_t n0, const double *src, double *dst,
                  void *stream);
int dk_struct_mean(int to_state, int32_t n, int32_t nx, int32_t ny, int32_t n_max, int32_t n0,
                   const double *src, double *dst, void *stream);

#ifdef __cplusplus
}
#endif
#endif
