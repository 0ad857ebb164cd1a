/*
 * Host layer of the LETKF engine, in C: the public C ABI of dida_letkf.h on
 * top of the thin kernel ABI of letkf_kernels.h.  It owns device workspaces,
 * orders the kernel stages of a sweep on the caller's stream, and implements
 * the reference's Fortran-callable letkf_core_ / letkf_core_batch_.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dida_letkf.h"
#include "letkf_kernels.h"

/* ------------------------------------------------------------------------ */
static _Thread_local char g_err[512];

static int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *dida_letkf_last_error(void) { return g_err; }

/* The error message of an entry point implemented in a HIP source. */
int dk_set_error(int code, const char *msg)
{
    return set_err(code, "%s", msg);
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        int e_ = (int)(expr);                                                            \
        if (e_ != 0)                                                                     \
            return set_err(DIDA_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                 \
                           hipGetErrorString((hipError_t)e_), __FILE__, __LINE__);       \
    } while (0)

static const double kEarthA = (double)6371.393f;

void dida_letkf_params_default(dida_letkf_params *p)
{
    memset(p, 0, sizeof *p);
    p->nanals = 0;
    p->max_obs = 10000;
    p->gc_filter = 1;
    p->local_dist = 1000.0;
    p->obs_polar = 90.0;
    p->inflation_factor = 1.0;
    p->rtps = 0;
    p->rtps_alpha = 0.0;
    p->form = DIDA_FORM_PLAIN;
    p->gram = DIDA_GRAM_FP64;
    p->jacobi_max_sweeps = 30;
    p->jacobi_tol = 0.0;
    p->bin_deg = 0.0;
    p->chain_len = 0;
    p->warm_start = 1;
    p->jacobi_predict = -1;
    p->vert_dist = 0.0;
    p->cube_nz = 1;
    p->gc_filter_vertical = 1;
    p->cube_n = 1;
    p->overlap = 0;
}

/* Starts of horizontal cubes advance by cube_n (no overlap) or cube_n -
 * overlap (letkf_compute.F90:55,61 and :180,188). */
static int cube_step(const dida_letkf_params *p)
{
    return p->overlap > 0 ? p->cube_n - p->overlap : p->cube_n;
}

static double jacobi_tol_of(const dida_letkf_params *p)
{
    if (p->jacobi_tol > 0.0) return p->jacobi_tol;
    return p->gram == DIDA_GRAM_FP64 ? 1e-13 : 1e-8;
}

static int predict_of(const dida_letkf_params *p)
{
    if (p->jacobi_predict < 0) return p->gram != DIDA_GRAM_FP64;
    return p->jacobi_predict != 0;
}

static int check_params(const dida_letkf_params *p)
{
    if (p->nanals < 2 || p->nanals > 128)
        return set_err(DIDA_ERR_UNSUPPORTED, "nanals = %d: supported range is 2..128", p->nanals);
    if (p->form == DIDA_FORM_BATCHED && p->nanals > 64 && p->vert_dist > 0.0)
        return set_err(DIDA_ERR_UNSUPPORTED,
                       "nanals = %d: the batched form takes k > 64 without vertical "
                       "localisation only (vertical: 2..64)", p->nanals);
    if (p->nrow < 1 || p->nlat < 1 || p->nlon < 1)
        return set_err(DIDA_ERR_ARG, "grid %d x %d x %d rows is empty", p->nlat, p->nlon, p->nrow);
    if (p->max_obs < 0) return set_err(DIDA_ERR_ARG, "max_obs < 0");
    if (!(p->local_dist > 0.0)) return set_err(DIDA_ERR_ARG, "local_dist must be > 0");
    if (!(p->inflation_factor > 0.0)) return set_err(DIDA_ERR_ARG, "inflation_factor must be > 0");
    if (p->jacobi_max_sweeps < 1) return set_err(DIDA_ERR_ARG, "jacobi_max_sweeps < 1");
    if (p->jacobi_tol < 0.0) return set_err(DIDA_ERR_ARG, "jacobi_tol must be >= 0");
    if (p->chain_len < 0) return set_err(DIDA_ERR_ARG, "chain_len must be >= 0");
    if (p->form != DIDA_FORM_PLAIN && p->form != DIDA_FORM_BATCHED)
        return set_err(DIDA_ERR_UNSUPPORTED, "form %d is not available", p->form);
    if (p->gram < DIDA_GRAM_FP64 || p->gram > DIDA_GRAM_BF16)
        return set_err(DIDA_ERR_UNSUPPORTED, "gram precision %d is not available", p->gram);
    if (p->form == DIDA_FORM_PLAIN && p->gram != DIDA_GRAM_FP64)
        return set_err(DIDA_ERR_UNSUPPORTED, "the plain form computes the Gram in FP64 only");
    if (p->cube_n < 1 || p->overlap < 0 || cube_step(p) < 1)
        return set_err(DIDA_ERR_ARG, "cube_n = %d, overlap = %d: need cube_n >= 1 and "
                       "0 <= overlap < cube_n", p->cube_n, p->overlap);
    int nbucket_set = 0;
    for (int b = 0; b < 4; ++b) {
        if (p->bucket_chain[b] < 0) return set_err(DIDA_ERR_ARG, "bucket_chain must be >= 0");
        nbucket_set += p->bucket_chain[b] > 0;
    }
    if (nbucket_set && !(p->form == DIDA_FORM_BATCHED && p->nanals <= 64))
        return set_err(DIDA_ERR_UNSUPPORTED,
                       "bucket_chain: chains per obs-count bucket are a batched-form option "
                       "for k <= 64");
    if (p->vert_dist > 0.0) {
        if (p->overlap != 0)
            return set_err(DIDA_ERR_UNSUPPORTED,
                           "vertical localisation runs with overlap = 0 only (the reference's "
                           "overlapping-cube sweep has none)");
        if (p->form != DIDA_FORM_BATCHED)
            return set_err(DIDA_ERR_UNSUPPORTED,
                           "vertical localisation runs in the batched form only");
        if (p->cube_nz < 1) return set_err(DIDA_ERR_ARG, "cube_nz must be >= 1");
    }
    return DIDA_OK;
}

/* ------------------------------------------------------------------------ */
#define DIDA_MAX_MARKS 160

struct dida_letkf_ctx {
    dida_letkf_params prm;
    int32_t nobs_max;
    int64_t ncol;
    /* analysis grid: one selection list per column (cube_n = 1) or per
     * horizontal cube, nsel = nalat * nalon; cube centres and transforms */
    int32_t nalat, nalon, step;
    int64_t nsel;
    float *ca_lat, *ca_lon;
    double *trans_all;
    const float *grid_lat, *grid_lon;
    dk_bin_geom geom;
    int end_bit;
    int32_t nbins_max;   /* bin arrays are sized for the finest bins tried */
    double bin_deg;
    double theta_deg, cos_far, cos_near, cos_keep;
    /* binning */
    uint32_t *key, *key_sorted;
    int32_t *iota, *perm, *bin_start;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    double *cx, *sx, *yr, *yb_b, *dep_b;
    float *u4;           /* FP32 unit vector of every binned obs (x, y, z, 0) */
    double *u3;          /* FP64 unit vector of every binned obs: x, y, z blocks of nobs_max */
    int32_t *iperm;      /* binned position of every caller obs index */
    float *yb32_b;       /* batched half-precision Grams: binned Yb in FP32 (no yb_b) */
    int32_t ldy;         /* row stride of yb_b: k (plain), the padded k (batched) */
    float *rinv_b;
    /* selection */
    int32_t *nobsl;
    int32_t *thr;        /* caller-index threshold of columns capped at MAX_OBS */
    int64_t *offset;
    void *scan_tmp;
    size_t scan_tmp_bytes;
    /* count / fill tiles, largest candidate count first (sized for the
     * smallest tile shape) */
    int32_t ntile_max;
    uint32_t *tile_key, *tile_key_sorted;
    int32_t *tile_id, *tile_order;
    void *tile_tmp;
    size_t tile_tmp_bytes;
    int32_t *csr_idx;
    double *csr_w;
    int64_t capacity;
    int32_t user_capacity; /* capacity given at create: plain sweeps run in chunks */
    int32_t *flags; /* [0] capacity overflow */
    /* transform (plain form, chunked) */
    int32_t chunk;
    double *trans;
    int32_t *info;
    /* batched form: chain queue */
    int32_t chain_len, chains_per_lon, nchain, nblocks;
    int32_t ngroup;      /* 0: one cost-sorted queue; DK_XCD_GROUPS: stripes per XCD */
    int32_t chained;     /* batched form, k > 64: chained LDS analyses (vscr per CU) */
    int32_t ncu;
    double *vscr;
    int32_t nchain_max;  /* chain arrays are sized for chains of one column */
    uint32_t *chain_key, *chain_key_sorted;
    int32_t *chain_id, *chain_order, *queue_head;
    void *chain_tmp;
    size_t chain_tmp_bytes;
    double *ymax;
    /* per-bucket chains (fused batched form): on when bucketed, chain length
     * bucket_len[b] for analyses of local count in bucket b (thresholds
     * kBucketLo); the chain descriptors are built on the device every sweep */
    int32_t sel_tile;    /* DK_SELECT_TILE_* of the count and fill passes */
    int32_t bucketed;
    int32_t bucket_len[DK_NBUCKET];
    int32_t *line_cnt, *chain_col0, *chain_ncol;
    int64_t *line_off;
    /* vertical localisation: analyses per column, binned obs ln p, cube
     * centre ln p [ncol * ncube], rows per cube (CSR over cubes) */
    int32_t ncube, ps_row;
    dida_vert_coord vc;
    double *obs_lnp, *cube_lnp;
    int32_t *cube_row_off, *cube_rows;
    /* selection lists ordered by obs ln p (capacity entries each) */
    float *vkey_in, *vkey_out;
    int32_t *vpos_in, *vpos_out, *vidx;
    double *vw;
    void *vsort_tmp;
    size_t vsort_tmp_bytes;
    unsigned long long *phase_clk; /* [DK_NPHASE], allocated when timing is on */
    /* FP32 perturbations of the dida_letkf_sweep_f32 call in progress */
    const float *xb32;
    float *xa32;
    /* per-stage HIP event timing (eager sweeps only) */
    int timing;
    int nmark;
    hipEvent_t ev[DIDA_MAX_MARKS];
    int ev_stage[DIDA_MAX_MARKS];
    /* graph */
    hipGraph_t graph;
    hipGraphExec_t gexec;
    hipStream_t last_stream;
    int32_t last_nobs;
    int32_t last_nobs_pending; /* obs count of the sweep being enqueued */
    int64_t generation;        /* bumped when the selection workspace is replaced */
};

static void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

int dida_letkf_destroy(dida_letkf_ctx *c)
{
    if (!c) return DIDA_OK;
    for (int i = 0; i < DIDA_MAX_MARKS; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    void *ptrs[] = {c->key, c->key_sorted, c->iota, c->perm, c->bin_start,
                    c->sort_tmp, c->cx, c->sx, c->yr, c->u4, c->u3, c->iperm, c->yb_b,
                    c->yb32_b, c->dep_b, c->rinv_b,
                    c->nobsl, c->thr, c->offset, c->scan_tmp, c->csr_idx, c->csr_w, c->flags,
                    c->trans, c->info, c->chain_key, c->chain_key_sorted, c->chain_id,
                    c->chain_order, c->queue_head, c->chain_tmp, c->ymax, c->phase_clk,
                    c->obs_lnp, c->cube_lnp, c->cube_row_off, c->cube_rows, c->ca_lat,
                    c->ca_lon, c->trans_all, c->vkey_in, c->vkey_out, c->vpos_in,
                    c->vpos_out, c->vidx, c->vw, c->vsort_tmp, c->vscr, c->line_cnt,
                    c->chain_col0, c->chain_ncol, c->line_off, c->tile_key,
                    c->tile_key_sorted, c->tile_id, c->tile_order, c->tile_tmp};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) dfree(ptrs[i]);
    free(c);
    return DIDA_OK;
}

#define DMALLOC(ptr, bytes)                                                              \
    do {                                                                                 \
        size_t b_ = (size_t)(bytes);                                                     \
        if (hipMalloc((void **)&(ptr), b_ ? b_ : 16) != hipSuccess) {                    \
            dida_letkf_destroy(c);                                                       \
            return set_err(DIDA_ERR_NOMEM, "hipMalloc(%zu) failed for %s", b_, #ptr);    \
        }                                                                                \
    } while (0)

/* Obs bins of bd degrees (clamped to [0.25, 15]): nband bands of 180 / nband
 * and nlonbin longitude bins of 360 / nlonbin tile the sphere exactly. */
static dk_bin_geom bin_geom(double bd, double obs_polar)
{
    if (bd < 0.25) bd = 0.25;
    if (bd > 15.0) bd = 15.0;
    dk_bin_geom g;
    memset(&g, 0, sizeof g);
    g.obs_polar = obs_polar;
    g.nband = (int32_t)ceil(180.0 / bd);
    g.nlonbin = (int32_t)ceil(360.0 / bd);
    g.dlat = 180.0 / g.nband;
    g.dlon = 360.0 / g.nlonbin;
    g.nbins = g.nband * g.nlonbin;
    return g;
}

/* Lower bounds of the local-obs-count buckets of the per-bucket chains. */
static const int32_t kBucketLo[DK_NBUCKET] = {0, 64, 512, 2048};

/* Bin size, chain length and queue order of the sweeps that follow (within
 * the sizes the context was created for); blen (fused batched form): one
 * chain length per obs-count bucket instead of chain_len, one queue. */
static int set_layout(dida_letkf_ctx *c, double bin_deg, int32_t chain_len, int32_t tiles,
                      const int32_t *blen)
{
    dk_bin_geom g = bin_geom(bin_deg, c->prm.obs_polar);
    if (g.nbins > c->nbins_max) return set_err(DIDA_ERR_ARG, "bins finer than allocated");
    c->geom = g;
    c->bin_deg = bin_deg;
    c->end_bit = 1;
    while ((1u << c->end_bit) <= (unsigned)g.nbins) ++c->end_bit;
    if (c->prm.form == DIDA_FORM_BATCHED) {
        if (chain_len < 1) chain_len = 1;
        if (chain_len > c->nalat) chain_len = c->nalat;
        c->chain_len = chain_len;
        c->chains_per_lon = (c->nalat + chain_len - 1) / chain_len;
        c->nchain = c->chains_per_lon * c->nalon;
        c->ngroup = tiles ? DK_XCD_GROUPS : 0;
        c->bucketed = blen != NULL && !c->chained;
        if (c->bucketed) {
            c->ngroup = 0;
            for (int b = 0; b < DK_NBUCKET; ++b)
                c->bucket_len[b] = blen[b] < 1 ? 1 : (blen[b] > c->nalat ? c->nalat : blen[b]);
        }
    }
    return DIDA_OK;
}

int dida_letkf_create(dida_letkf_ctx **out, const dida_letkf_params *prm, int32_t nobs_max,
                      int64_t sel_capacity)
{
    if (!out || !prm) return set_err(DIDA_ERR_ARG, "null argument");
    *out = NULL;
    int rc = check_params(prm);
    if (rc) return rc;
    if (nobs_max < 0 || sel_capacity < 0) return set_err(DIDA_ERR_ARG, "negative size");
    /* the selection's fill addresses the binned obs with 32-bit byte offsets */
    if (nobs_max >= (1 << 29))
        return set_err(DIDA_ERR_UNSUPPORTED, "nobs_max = %d: at most %d obs", (int)nobs_max,
                       (1 << 29) - 1);
    dida_letkf_ctx *c = (dida_letkf_ctx *)calloc(1, sizeof *c);
    if (!c) return set_err(DIDA_ERR_NOMEM, "calloc");
    c->prm = *prm;
    c->nobs_max = nobs_max;
    c->ncol = (int64_t)prm->nlat * prm->nlon;
    c->step = cube_step(prm);
    c->nalat = (prm->nlat + c->step - 1) / c->step;
    c->nalon = (prm->nlon + c->step - 1) / c->step;
    c->nsel = (int64_t)c->nalat * c->nalon;
    const int k = prm->nanals;

    /* support radius in degrees, widened so that no kept obs falls outside
     * the visited bins (the distance formula uses a float-rounded pi) */
    const double pi = 3.14159265358979323846;
    double theta_rad = 2.0 * prm->local_dist / kEarthA;
    c->theta_deg = theta_rad * (180.0 / pi) * (1.0 + 1e-6) + 1e-6;
    c->cos_far = (theta_rad * (1.0 + 1e-7) < pi) ? cos(theta_rad * (1.0 + 1e-7)) : -2.0;
    /* candidates closer than theta (1 - 1e-7) have gc > 0 (r < 2) for sure */
    c->cos_near = (theta_rad * (1.0 - 1e-7) < pi) ? cos(theta_rad * (1.0 - 1e-7)) : -1.0;
    /* gc > 0 exactly when r = d / L <= 2 (gc(2) is about 1.4e-6 with the
     * reference's REAL(4) constants), that is central angle <= 2 L / a */
    c->cos_keep = theta_rad < pi ? cos(theta_rad) : -2.0;
    /* bins: the given size, or theta / 4; arrays sized for the finest size
     * the autotuner tries (theta / 8) */
    const double bd = prm->bin_deg > 0.0 ? prm->bin_deg : c->theta_deg / 4.0;
    {
        const dk_bin_geom g0 = bin_geom(bd, prm->obs_polar);
        const dk_bin_geom g1 = bin_geom(c->theta_deg / 8.0, prm->obs_polar);
        c->nbins_max = g0.nbins > g1.nbins ? g0.nbins : g1.nbins;
    }
    c->nchain_max = (int32_t)c->nsel;
    /* automatic chain length: 8 columns for the fused kernel (its autotuner
     * refines it), 32 analyses for the chained workgroups (measured best) */
    const int32_t cl0 = prm->form == DIDA_FORM_BATCHED && k > 64 ? 32 : 8;
    const int32_t cl1 = prm->chain_len > 0 ? prm->chain_len : cl0;
    int32_t blen[DK_NBUCKET];
    int bset = 0;
    for (int b = 0; b < DK_NBUCKET; ++b) {
        blen[b] = prm->bucket_chain[b] > 0 ? prm->bucket_chain[b] : cl1;
        bset |= prm->bucket_chain[b] > 0;
    }
    rc = set_layout(c, bd, cl1, prm->xcd_tiles, bset ? blen : NULL);
    if (rc) {
        free(c);
        return rc;
    }
    int end_bit_max = 1;
    while ((1u << end_bit_max) <= (unsigned)c->nbins_max) ++end_bit_max;

    const size_t no = (size_t)(nobs_max > 0 ? nobs_max : 1);
    DMALLOC(c->key, no * sizeof(uint32_t));
    DMALLOC(c->key_sorted, no * sizeof(uint32_t));
    DMALLOC(c->iota, no * sizeof(int32_t));
    DMALLOC(c->perm, no * sizeof(int32_t));
    DMALLOC(c->bin_start, (size_t)(c->nbins_max + 2) * sizeof(int32_t));
    c->sort_tmp_bytes = 0;
    if (dk_sort_pairs_u32(NULL, &c->sort_tmp_bytes, c->key, c->key_sorted, c->iota, c->perm,
                          (int32_t)no, end_bit_max, NULL)) {
        dida_letkf_destroy(c);
        return set_err(DIDA_ERR_HIP, "sort size query failed");
    }
    DMALLOC(c->sort_tmp, c->sort_tmp_bytes);
    DMALLOC(c->cx, no * sizeof(double));
    DMALLOC(c->sx, no * sizeof(double));
    DMALLOC(c->yr, no * sizeof(double));
    DMALLOC(c->u4, no * 4 * sizeof(float));
    DMALLOC(c->u3, no * 3 * sizeof(double));
    DMALLOC(c->iperm, no * sizeof(int32_t));
    c->chained = prm->form == DIDA_FORM_BATCHED && k > 64;
    /* the one-wavefront kernels (batched form, and the plain form with k <= 64)
     * read binned Yb rows of 32 or 64 values, zero past k */
    c->ldy = k <= 64 && !c->chained ? (k <= 32 ? 32 : 64) : k;
    if (c->chained && prm->gram != DIDA_GRAM_FP64) c->ldy = (k + 3) & ~3; /* float4 rows */
    /* the split half-precision Grams read FP32 rows (their hi + lo parts
     * hold at most 22 bits) */
    const size_t row_bytes = prm->form == DIDA_FORM_BATCHED && prm->gram != DIDA_GRAM_FP64
                                 ? sizeof(float) : sizeof(double);
    /* the one-wavefront kernels address the rows with 32-bit byte offsets */
    if (k <= 64 && !c->chained && no * (size_t)c->ldy * row_bytes > 0xFFFFFFFFull) {
        dida_letkf_destroy(c);
        return set_err(DIDA_ERR_UNSUPPORTED,
                       "nobs_max = %lld: the binned rows exceed 4 GiB (at most %lld obs at k = %d)",
                       (long long)no, (long long)(0xFFFFFFFFull / ((size_t)c->ldy * row_bytes)),
                       k);
    }
    if (row_bytes == sizeof(float))
        DMALLOC(c->yb32_b, no * (size_t)c->ldy * sizeof(float));
    else
        DMALLOC(c->yb_b, no * (size_t)c->ldy * sizeof(double));
    DMALLOC(c->dep_b, no * sizeof(double));
    DMALLOC(c->rinv_b, no * sizeof(float));

    DMALLOC(c->nobsl, (size_t)c->nsel * sizeof(int32_t));
    DMALLOC(c->thr, (size_t)c->nsel * sizeof(int32_t));
    DMALLOC(c->offset, (size_t)(c->nsel + 1) * sizeof(int64_t));
    dk_scan_offsets(NULL, &c->scan_tmp_bytes, NULL, NULL, (int32_t)c->nsel, NULL);
    DMALLOC(c->scan_tmp, c->scan_tmp_bytes);
    c->ntile_max = (int32_t)dk_select_ntiles(c->nalat, c->nalon, DK_SELECT_TILE_4X1);
    DMALLOC(c->tile_key, (size_t)c->ntile_max * sizeof(uint32_t));
    DMALLOC(c->tile_key_sorted, (size_t)c->ntile_max * sizeof(uint32_t));
    DMALLOC(c->tile_id, (size_t)c->ntile_max * sizeof(int32_t));
    DMALLOC(c->tile_order, (size_t)c->ntile_max * sizeof(int32_t));
    if (dk_select_tile_order(NULL, c->ntile_max, NULL, NULL, NULL, NULL, NULL, &c->tile_tmp_bytes,
                             NULL)) {
        dida_letkf_destroy(c);
        return set_err(DIDA_ERR_HIP, "tile sort size query failed");
    }
    DMALLOC(c->tile_tmp, c->tile_tmp_bytes);
    DMALLOC(c->flags, 4 * sizeof(int32_t));
    c->capacity = sel_capacity;
    c->user_capacity = sel_capacity > 0;
    if (sel_capacity > 0) {
        DMALLOC(c->csr_idx, (size_t)sel_capacity * sizeof(int32_t));
        DMALLOC(c->csr_w, (size_t)sel_capacity * sizeof(double));
    }
    DMALLOC(c->info, (size_t)c->nsel * sizeof(int32_t));
    if (prm->cube_n > 1) {
        DMALLOC(c->ca_lat, (size_t)c->nalat * sizeof(float));
        DMALLOC(c->ca_lon, (size_t)c->nalon * sizeof(float));
        if (!(prm->vert_dist > 0.0))  /* vertical cubes update in the fused kernel */
            DMALLOC(c->trans_all, (size_t)c->nsel * k * k * sizeof(double));
    }
    if (prm->form == DIDA_FORM_PLAIN || c->chained) {
        /* analyses per transform launch; the chained form takes more, since
         * each launch ends with a tail of whole chains on few workgroups */
        const int64_t cmax = c->chained ? 131072 : 32768;
        c->chunk = (int32_t)(c->nsel < cmax ? c->nsel : cmax);
        if (prm->cube_n == 1) DMALLOC(c->trans, (size_t)c->chunk * k * k * sizeof(double));
    }
    if (prm->form == DIDA_FORM_BATCHED) {
        DMALLOC(c->chain_key, (size_t)c->nchain_max * sizeof(uint32_t));
        DMALLOC(c->chain_key_sorted, (size_t)c->nchain_max * sizeof(uint32_t));
        DMALLOC(c->chain_id, (size_t)c->nchain_max * sizeof(int32_t));
        DMALLOC(c->chain_order, (size_t)c->nchain_max * sizeof(int32_t));
        DMALLOC(c->queue_head, 16 * sizeof(int32_t));
        DMALLOC(c->ymax, no * sizeof(double));
        if (dk_sort_pairs_u32(NULL, &c->chain_tmp_bytes, c->chain_key, c->chain_key_sorted,
                              c->chain_id, c->chain_order, c->nchain_max, 32, NULL)) {
            dida_letkf_destroy(c);
            return set_err(DIDA_ERR_HIP, "chain sort size query failed");
        }
        DMALLOC(c->chain_tmp, c->chain_tmp_bytes);
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu <= 0)
            ncu = 256;
        c->nblocks = ncu * dk_wave_blocks_per_cu(k, prm->gram);
        c->ncu = ncu;
        if (!c->chained) {
            DMALLOC(c->line_cnt, (size_t)c->nalon * sizeof(int32_t));
            DMALLOC(c->line_off, (size_t)(c->nalon + 1) * sizeof(int64_t));
            DMALLOC(c->chain_col0, (size_t)c->nchain_max * sizeof(int32_t));
            DMALLOC(c->chain_ncol, (size_t)c->nchain_max * sizeof(int32_t));
        }
        if (prm->vert_dist > 0.0) DMALLOC(c->obs_lnp, no * sizeof(double));
        /* chained analyses: one workgroup per CU, 128 x 128 eigenvectors each */
        if (c->chained) DMALLOC(c->vscr, (size_t)ncu * 128 * 128 * sizeof(double));
    }
    *out = c;
    return DIDA_OK;
}

int dida_letkf_set_grid(dida_letkf_ctx *c, const float *grid_lat, const float *grid_lon)
{
    if (!c || !grid_lat || !grid_lon) return set_err(DIDA_ERR_ARG, "null argument");
    c->grid_lat = grid_lat;
    c->grid_lon = grid_lon;
    return DIDA_OK;
}

/* Buffers that order the selection lists by obs ln p (vertical
 * localisation), once the selection capacity is known. */
static int alloc_vertical_sort(dida_letkf_ctx *c)
{
    if (c->prm.vert_dist <= 0.0 || c->capacity <= 0 || c->vkey_in) return DIDA_OK;
    /* rocPRIM's segmented sort counts entries and segments in 32 bits */
    if (c->capacity > (int64_t)UINT32_MAX || c->nsel > (int64_t)UINT32_MAX)
        return set_err(DIDA_ERR_UNSUPPORTED,
                       "vertical localisation: a selection workspace of %lld entries is above "
                       "2^32 - 1 (pass a smaller capacity to run the sweep in pieces)",
                       (long long)c->capacity);
    const size_t cap = (size_t)c->capacity;
    size_t tb = 0;
    if (dk_seg_sort_lnp(NULL, &tb, c->nsel, c->capacity, c->offset, NULL, NULL, NULL, NULL, NULL,
                        NULL, NULL, NULL, NULL, NULL))
        return set_err(DIDA_ERR_HIP, "segmented sort size query failed");
    if (hipMalloc((void **)&c->vkey_in, cap * sizeof(float)) ||
        hipMalloc((void **)&c->vkey_out, cap * sizeof(float)) ||
        hipMalloc((void **)&c->vpos_in, cap * sizeof(int32_t)) ||
        hipMalloc((void **)&c->vpos_out, cap * sizeof(int32_t)) ||
        hipMalloc((void **)&c->vidx, cap * sizeof(int32_t)) ||
        hipMalloc((void **)&c->vw, cap * sizeof(double)) ||
        hipMalloc(&c->vsort_tmp, tb ? tb : 16))
        return set_err(DIDA_ERR_NOMEM, "vertical ordering buffers of %zu entries", cap);
    c->vsort_tmp_bytes = tb;
    return DIDA_OK;
}

int dida_letkf_set_levels(dida_letkf_ctx *c, const dida_vert_coord *vc, int32_t ps_row,
                          int32_t nvar, const int32_t *var_start, const int32_t *var_lev)
{
    if (!c || !vc || !var_start || !var_lev) return set_err(DIDA_ERR_ARG, "null argument");
    if (!(c->prm.vert_dist > 0.0))
        return set_err(DIDA_ERR_ARG, "set_levels: vertical localisation is off (vert_dist <= 0)");
    const int32_t nrow = c->prm.nrow, nz = vc->nz;
    if (vc->type != DIDA_VCOORD_SIGMA && vc->type != DIDA_VCOORD_HYBRID)
        return set_err(DIDA_ERR_ARG, "set_levels: vertical coordinate type");
    if (nz < 1 || !vc->c1 || (vc->type == DIDA_VCOORD_HYBRID && !vc->c2))
        return set_err(DIDA_ERR_ARG, "set_levels: vertical coordinate");
    if (ps_row < 0 || ps_row >= nrow) return set_err(DIDA_ERR_ARG, "set_levels: ps_row");
    if (nvar < 1) return set_err(DIDA_ERR_ARG, "set_levels: no variables");
    const int32_t cnz = c->prm.cube_nz;
    const int32_t ncube = (nz + cnz - 1) / cnz;
    int32_t *seen = (int32_t *)calloc((size_t)nrow, sizeof(int32_t));
    int32_t *off = (int32_t *)malloc((size_t)(ncube + 1) * sizeof(int32_t));
    int32_t *rows = (int32_t *)malloc((size_t)nrow * sizeof(int32_t));
    int rc = DIDA_OK;
    if (!seen || !off || !rows) rc = set_err(DIDA_ERR_NOMEM, "set_levels: host tables");
    for (int v = 0; v < nvar && !rc; ++v) {
        if (var_start[v] < 0 || (var_lev[v] != 1 && var_lev[v] != nz) ||
            var_start[v] + var_lev[v] > nrow) {
            rc = set_err(DIDA_ERR_ARG, "set_levels: variable %d rows (levels must be 1 or nz)", v);
            break;
        }
        for (int l = 0; l < var_lev[v]; ++l) seen[var_start[v] + l]++;
    }
    for (int r = 0; r < nrow && !rc; ++r)
        if (seen[r] != 1) rc = set_err(DIDA_ERR_ARG, "set_levels: row %d is in %d variables", r, seen[r]);
    /* cube q: levels sk..ek of every 3-D variable, then the 2-D variables
     * with the last cube (letkf_compute.F90:438-520) */
    int32_t nr = 0;
    for (int q = 0; q < ncube && !rc; ++q) {
        const int sk = q * cnz, ek = sk + cnz - 1 < nz - 1 ? sk + cnz - 1 : nz - 1;
        off[q] = nr;
        for (int v = 0; v < nvar; ++v) {
            if (var_lev[v] == 1) {
                if (q == ncube - 1) rows[nr++] = var_start[v];
            } else {
                for (int l = sk; l <= ek; ++l) rows[nr++] = var_start[v] + l;
            }
        }
    }
    if (!rc) off[ncube] = nr;
    if (!rc && c->gexec) {
        /* a captured sweep refers to the tables replaced below */
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = NULL;
        if (c->graph) (void)hipGraphDestroy(c->graph);
        c->graph = NULL;
    }
    if (!rc) {
        /* hipFree waits for the device work that may still read them */
        dfree(c->cube_row_off);
        dfree(c->cube_rows);
        dfree(c->cube_lnp);
        c->cube_row_off = NULL;
        c->cube_rows = NULL;
        c->cube_lnp = NULL;
        c->ncube = 0;
        if (hipMalloc((void **)&c->cube_row_off, (size_t)(ncube + 1) * sizeof(int32_t)) ||
            hipMalloc((void **)&c->cube_rows, (size_t)nrow * sizeof(int32_t)) ||
            hipMalloc((void **)&c->cube_lnp, (size_t)c->nsel * ncube * sizeof(double)))
            rc = set_err(DIDA_ERR_NOMEM, "set_levels: device tables");
        else if (hipMemcpy(c->cube_row_off, off, (size_t)(ncube + 1) * sizeof(int32_t),
                           hipMemcpyHostToDevice) ||
                 hipMemcpy(c->cube_rows, rows, (size_t)nrow * sizeof(int32_t),
                           hipMemcpyHostToDevice))
            rc = set_err(DIDA_ERR_HIP, "set_levels: copy");
    }
    if (!rc) {
        /* info holds one entry per analysis */
        int32_t *info = NULL;
        if (hipMalloc((void **)&info, (size_t)c->nsel * ncube * sizeof(int32_t)))
            rc = set_err(DIDA_ERR_NOMEM, "set_levels: info");
        else {
            dfree(c->info);
            c->info = info;
        }
    }
    if (!rc) {
        c->vc = *vc;
        c->ps_row = ps_row;
        c->ncube = ncube;
        rc = alloc_vertical_sort(c);
    }
    free(seen);
    free(off);
    free(rows);
    return rc;
}

/* Record a timing mark closing `stage` (-1: the start mark). */
static void tmark(dida_letkf_ctx *c, int stage, hipStream_t s)
{
    if (!c->timing || c->nmark >= DIDA_MAX_MARKS) return;
    if (!c->ev[c->nmark] && hipEventCreate(&c->ev[c->nmark]) != hipSuccess) return;
    c->ev_stage[c->nmark] = stage;
    (void)hipEventRecord(c->ev[c->nmark], s);
    c->nmark++;
}

/* Stages up to and including the count pass and the offset scan. */
static int sweep_select_count(dida_letkf_ctx *c, const dida_letkf_obs *o, hipStream_t s,
                              dk_select_args *sa)
{
    const int k = c->prm.nanals;
    c->last_nobs_pending = o->nobs;
    c->nmark = 0;
    tmark(c, -1, s);
    HIP_TRY(dk_fill_i32(c->flags, 0, 4, s));
    HIP_TRY(dk_bin_keys(o->nobs, o->lat, o->lon, c->geom, c->key, c->iota, s));
    if (o->nobs > 0) {
        size_t tb = c->sort_tmp_bytes;
        HIP_TRY(dk_sort_pairs_u32(c->sort_tmp, &tb, c->key, c->key_sorted, c->iota, c->perm,
                                  o->nobs, c->end_bit, s));
    }
    HIP_TRY(dk_bin_starts(c->geom.nbins, o->nobs, c->key_sorted, c->bin_start, s));
    HIP_TRY(dk_obs_gather(o->nobs, k, c->ldy, c->perm, o->lat, o->lon, o->yb, o->dep, o->rinv,
                          c->cx, c->sx, c->yr, c->u4, c->u3, c->u3 + c->nobs_max,
                          c->u3 + 2 * c->nobs_max, c->iperm, c->yb_b, c->yb32_b, c->dep_b,
                          c->rinv_b, s));
    if (c->ncube > 0) HIP_TRY(dk_obs_lnp(o->nobs, c->perm, o->ph, c->obs_lnp, s));
    if (c->chained && c->prm.gram == DIDA_GRAM_FP16X3)
        HIP_TRY(dk_obs_ymax32(o->nobs, k, c->ldy, c->yb32_b, c->rinv_b, c->ymax, s));
    tmark(c, DIDA_STAGE_BIN, s);
    memset(sa, 0, sizeof *sa);
    sa->nlat = c->nalat;
    sa->nlon = c->nalon;
    sa->grid_lat = c->grid_lat;
    sa->grid_lon = c->grid_lon;
    if (c->prm.cube_n > 1) {
        /* one selection per horizontal cube, at its centre column */
        HIP_TRY(dk_cube_centres(c->prm.nlat, c->prm.nlon, c->prm.cube_n, c->step, c->grid_lat,
                                c->grid_lon, c->ca_lat, c->ca_lon, s));
        sa->grid_lat = c->ca_lat;
        sa->grid_lon = c->ca_lon;
    }
    sa->geom = c->geom;
    sa->obs.cx = c->cx;
    sa->obs.sx = c->sx;
    sa->obs.yr = c->yr;
    sa->obs.u4 = c->u4;
    sa->obs.ux = c->u3;
    sa->obs.uy = c->u3 + c->nobs_max;
    sa->obs.uz = c->u3 + 2 * c->nobs_max;
    sa->obs.yb = c->yb_b;
    sa->obs.dep = c->dep_b;
    sa->obs.rinv = c->rinv_b;
    sa->obs.bin_start = c->bin_start;
    sa->local_dist = c->prm.local_dist;
    sa->theta_deg = c->theta_deg;
    sa->cos_far = c->cos_far;
    sa->cos_near = c->cos_near;
    sa->cos_keep = c->cos_keep;
    sa->gc_filter = c->prm.gc_filter;
    sa->max_obs = c->prm.max_obs;
    sa->nobsl = c->nobsl;
    sa->offset = c->offset;
    sa->overflow = c->flags;
    sa->perm = c->perm;
    sa->iperm = c->iperm;
    sa->nobs = o->nobs;
    sa->thr = c->thr;
    sa->tile = c->sel_tile;
    sa->perm_bits = 1;
    while (sa->perm_bits < 31 && ((int64_t)1 << sa->perm_bits) < (int64_t)o->nobs) ++sa->perm_bits;
    /* the tiles of the count and fill passes, largest first */
    {
        const int64_t nt = dk_select_ntiles(sa->nlat, sa->nlon, sa->tile);
        size_t tb = c->tile_tmp_bytes;
        sa->tile_order = NULL;
        if (nt <= c->ntile_max) {
            HIP_TRY(dk_select_tile_order(sa, (int32_t)nt, c->tile_key, c->tile_id,
                                         c->tile_key_sorted, c->tile_order, c->tile_tmp, &tb, s));
            sa->tile_order = c->tile_order;
        }
    }
    HIP_TRY(dk_select_count(sa, s));
    /* columns capped at MAX_OBS: the caller-order threshold of what they keep */
    HIP_TRY(dk_select_cap(sa, s));
    size_t tb = c->scan_tmp_bytes;
    HIP_TRY(dk_scan_offsets(c->scan_tmp, &tb, c->nobsl, c->offset, (int32_t)c->nsel, s));
    return DIDA_OK;
}

/* Drop the context's captured graph (it refers to buffers being replaced). */
static void drop_graph(dida_letkf_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = NULL;
    c->graph = NULL;
}

/* Eager sweeps: the selection entries of this sweep (offset[nsel], read
 * back once, a stream synchronisation) against the workspace.  Sized by the
 * sweep (no capacity given at create): the workspace grows when a sweep's
 * lists exceed it, which replaces its buffers, so graphs captured before
 * are dropped (group graphs refuse to launch).  Fixed at create: the plain
 * form and the chained form run in runs that fit; the fused batched form
 * reports DIDA_ERR_CAPACITY before any analysis runs. */
static int ensure_capacity(dida_letkf_ctx *c, hipStream_t s)
{
    int64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, c->offset + c->nsel, sizeof total, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (c->user_capacity) {
        if (total > c->capacity && c->prm.form == DIDA_FORM_BATCHED && !c->chained)
            return set_err(DIDA_ERR_CAPACITY,
                           "the sweep's selection lists hold %lld entries, the workspace %lld "
                           "(the fused batched form runs whole sweeps only)",
                           (long long)total, (long long)c->capacity);
        return DIDA_OK;
    }
    if (total <= c->capacity && c->capacity > 0) return DIDA_OK;
    drop_graph(c);
    c->generation++;
    dfree(c->csr_idx);
    dfree(c->csr_w);
    c->csr_idx = NULL;
    c->csr_w = NULL;
    void *vb[] = {c->vkey_in, c->vkey_out, c->vpos_in, c->vpos_out, c->vidx, c->vw, c->vsort_tmp};
    for (size_t i = 0; i < sizeof vb / sizeof vb[0]; ++i) dfree(vb[i]);
    c->vkey_in = c->vkey_out = NULL;
    c->vpos_in = c->vpos_out = c->vidx = NULL;
    c->vw = NULL;
    c->vsort_tmp = NULL;
    c->capacity = 0;
    int64_t cap = total + total / 8 + 1024;
    if (hipMalloc((void **)&c->csr_idx, (size_t)cap * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&c->csr_w, (size_t)cap * sizeof(double)) != hipSuccess)
        return set_err(DIDA_ERR_NOMEM, "selection workspace of %lld entries", (long long)cap);
    c->capacity = cap;
    return alloc_vertical_sort(c);
}

/* Horizontal cubes: every column from the mean transform of its covering
 * cubes (trans_all). */
static int cube_update(dida_letkf_ctx *c, const double *xb, const double *axb, double *xa,
                       double *axa, hipStream_t s)
{
    dk_update_args ua;
    memset(&ua, 0, sizeof ua);
    ua.k = c->prm.nanals;
    ua.nrow = c->prm.nrow;
    ua.ncol = (int32_t)c->ncol;
    ua.col0 = 0;
    ua.trans = c->trans_all;
    ua.xb = xb;
    ua.axb = axb;
    ua.xa = xa;
    ua.axa = axa;
    ua.rtps = c->prm.rtps;
    ua.rtps_alpha = c->prm.rtps_alpha;
    ua.nlat = c->prm.nlat;
    ua.nlon = c->prm.nlon;
    ua.cube_n = c->prm.cube_n;
    ua.step = c->step;
    ua.nalat = c->nalat;
    ua.nalon = c->nalon;
    ua.xb32 = c->xb32;
    ua.xa32 = c->xa32;
    if (ua.k <= 64 && !ua.xa32) HIP_TRY(dk_update_cols(&ua, s));
    else HIP_TRY(dk_update_f64(&ua, s));
    return DIDA_OK;
}

/* Batched form: chain queue by cost, then the persistent fused kernel. */
static int sweep_batched(dida_letkf_ctx *c, const double *xb, const double *axb, double *xa,
                         double *axa, hipStream_t s)
{
    dk_batched_args ba;
    memset(&ba, 0, sizeof ba);
    ba.k = c->prm.nanals;
    ba.nrow = c->prm.nrow;
    ba.nlat = c->nalat;
    ba.nlon = c->nalon;
    ba.trans_out = c->prm.cube_n > 1 ? c->trans_all : NULL;
    ba.nchain = c->nchain;
    ba.chains_per_lon = c->chains_per_lon;
    ba.chain_len = c->chain_len;
    ba.gram = c->prm.gram;
    ba.warm_start = c->prm.warm_start ? 1 : 0;
    ba.predict = predict_of(&c->prm);
    ba.rtps = c->prm.rtps;
    ba.max_sweeps = c->prm.jacobi_max_sweeps;
    ba.max_p = c->prm.max_obs;
    ba.rtps_alpha = c->prm.rtps_alpha;
    ba.rho = c->prm.inflation_factor;
    ba.tol = jacobi_tol_of(&c->prm);
    ba.capacity = c->capacity;
    ba.offset = c->offset;
    ba.nobsl = c->nobsl;
    ba.idx = c->csr_idx;
    ba.rloc = c->csr_w;
    ba.yb = c->yb_b;
    ba.yb32 = c->yb32_b;
    ba.ldy = c->ldy;
    ba.rinv = c->rinv_b;
    ba.dep = c->dep_b;
    ba.ymax = c->ymax;
    ba.xb = xb;
    ba.axb = axb;
    ba.xa = xa;
    ba.axa = axa;
    ba.info = c->info;
    ba.queue = c->chain_order;
    ba.queue_head = c->queue_head;
    if ((c->timing & 2) && c->phase_clk) {
        HIP_TRY(hipMemsetAsync(c->phase_clk, 0, DK_NPHASE * sizeof(unsigned long long), s));
        ba.phase_clk = c->phase_clk;
    }
    if (c->ncube > 0) {
        dk_cube_lnp_args la;
        memset(&la, 0, sizeof la);
        la.nrow = c->prm.nrow;
        la.ps_row = c->ps_row;
        la.nz = c->vc.nz;
        la.vtype = c->vc.type;
        la.ncube = c->ncube;
        la.cube_nz = c->prm.cube_nz;
        la.ncol = c->nsel;
        la.cube_n = c->prm.cube_n;
        la.nalat = c->nalat;
        la.grid_nlat = c->prm.nlat;
        la.grid_nlon = c->prm.nlon;
        la.c1 = c->vc.c1;
        la.c2 = c->vc.c2 ? c->vc.c2 : c->vc.c1;
        la.ptop = c->vc.ptop;
        la.p0 = c->vc.p0;
        la.axb = axb;
        la.cube_lnp = c->cube_lnp;
        HIP_TRY(dk_cube_lnp(&la, s));
        ba.ncube = c->ncube;
        ba.gc_vert = c->prm.gc_filter_vertical ? 1 : 0;
        ba.vert_dist = c->prm.vert_dist;
        ba.obs_lnp = c->obs_lnp;
        ba.cube_lnp = c->cube_lnp;
        ba.cube_row_off = c->cube_row_off;
        ba.cube_rows = c->cube_rows;
        ba.cube_n = c->prm.cube_n;
        ba.grid_nlat = c->prm.nlat;
        ba.grid_nlon = c->prm.nlon;
        ba.trans_out = NULL;
        if (c->vkey_in) {
            size_t tb = c->vsort_tmp_bytes;
            HIP_TRY(dk_seg_sort_lnp(c->vsort_tmp, &tb, c->nsel, c->capacity, c->offset,
                                    c->csr_idx, c->csr_w, c->obs_lnp, c->vkey_in, c->vkey_out,
                                    c->vpos_in, c->vpos_out, c->vidx, c->vw, s));
            ba.idx = c->vidx;
            ba.rloc = c->vw;
            ba.sel_key = c->vkey_out;
        }
    }
    if (c->prm.gram == DIDA_GRAM_FP16X3)
        HIP_TRY(dk_obs_ymax32(c->last_nobs_pending, ba.k, c->ldy, c->yb32_b, c->rinv_b, c->ymax,
                              s));
    ba.ngroup = c->ngroup;
    if (c->bucketed) {
        /* chains of one obs-count bucket, built from this sweep's counts */
        dk_chain_build_args cb;
        memset(&cb, 0, sizeof cb);
        cb.nlat = c->nalat;
        cb.nlon = c->nalon;
        cb.nbucket = DK_NBUCKET;
        for (int b = 0; b < DK_NBUCKET; ++b) {
            cb.lo[b] = kBucketLo[b];
            cb.len[b] = c->bucket_len[b];
        }
        cb.nobsl = c->nobsl;
        cb.line_cnt = c->line_cnt;
        cb.line_off = c->line_off;
        cb.chain_col0 = c->chain_col0;
        cb.chain_ncol = c->chain_ncol;
        HIP_TRY(dk_chain_build(&cb, 0, s));
        size_t tb = c->scan_tmp_bytes;
        HIP_TRY(dk_scan_offsets(c->scan_tmp, &tb, c->line_cnt, c->line_off, c->nalon, s));
        HIP_TRY(dk_chain_build(&cb, 1, s));
        ba.bucketed = 1;
        ba.chain_col0 = c->chain_col0;
        ba.chain_ncol = c->chain_ncol;
        ba.nchain_dev = c->line_off + c->nalon;
        ba.nchain = c->nchain_max;
        ba.ngroup = 0;
        HIP_TRY(dk_chain_keys(&ba, c->chain_key, c->chain_id, s));
        tb = c->chain_tmp_bytes;
        HIP_TRY(dk_sort_pairs_u32(c->chain_tmp, &tb, c->chain_key, c->chain_key_sorted,
                                  c->chain_id, c->chain_order, c->nchain_max, 32, s));
    } else if (c->ngroup > 0) {
        HIP_TRY(dk_chain_groups(&ba, c->chain_order, s));
    } else {
        HIP_TRY(dk_chain_keys(&ba, c->chain_key, c->chain_id, s));
        size_t tb = c->chain_tmp_bytes;
        HIP_TRY(dk_sort_pairs_u32(c->chain_tmp, &tb, c->chain_key, c->chain_key_sorted,
                                  c->chain_id, c->chain_order, c->nchain, 32, s));
    }
    HIP_TRY(dk_fill_i32(c->queue_head, 0, 16, s));
    tmark(c, DIDA_STAGE_SELECT, s);
    HIP_TRY(dk_letkf_wave(&ba, c->nblocks, s));
    tmark(c, DIDA_STAGE_TRANSFORM, s);
    if (c->prm.cube_n > 1 && c->ncube == 0) {
        HIP_TRY(cube_update(c, xb, axb, xa, axa, s));
        tmark(c, DIDA_STAGE_UPDATE, s);
    }
    return DIDA_OK;
}

/* Transforms of analyses [aid0, aid0 + nanal): one workgroup each, or, for
 * the batched form with k > 64, chains of warm-started analyses. */
static int core_launch(dida_letkf_ctx *c, dk_core_args *ca, hipStream_t s)
{
    if (!c->chained)
        return c->prm.nanals <= 64 ? dk_letkf_plain_wave(ca, s) : dk_letkf_core_f64(ca, s);
    dk_chain_args ch;
    ch.chain_len = c->chain_len;
    /* chains on multiples of chain_len from analysis 0 (a first partial one
     * when aid0 is inside a chain) */
    ch.nchain = (ca->aid0 % c->chain_len + ca->nanal + c->chain_len - 1) / c->chain_len;
    ch.head = c->queue_head;
    ch.vscr = c->vscr;
    ch.phase_clk = NULL;
    /* chains drained most expensive first, so the launch ends on cheap ones */
    int e = dk_chain_costs(ca, ch.chain_len, ch.nchain, c->chain_key, c->chain_id, s);
    if (e) return e;
    size_t tb = c->chain_tmp_bytes;
    e = dk_sort_pairs_u32(c->chain_tmp, &tb, c->chain_key, c->chain_key_sorted, c->chain_id,
                          c->chain_order, ch.nchain, 32, s);
    if (e) return e;
    ch.order = c->chain_order;
    e = dk_fill_i32(c->queue_head, 0, 1, s);
    if (e) return e;
    if ((c->timing & 2) && c->phase_clk) {
        e = (int)hipMemsetAsync(c->phase_clk, 0, DK_NPHASE * sizeof(unsigned long long), s);
        if (e) return e;
        ch.phase_clk = c->phase_clk;
    }
    return dk_letkf_chain_f64(ca, &ch, c->ncu, s);
}

/* Plain form: transforms (and, for columns, updates) of analyses [a0, a1)
 * whose selection entries sit at offset[] - base; `last` closes the sweep
 * (horizontal cubes update every column once all transforms exist). */
static int plain_transform(dida_letkf_ctx *c, int64_t a0, int64_t a1, int64_t base, int last,
                           const double *xb, const double *axb, double *xa, double *axa,
                           hipStream_t s)
{
    const int k = c->prm.nanals;
    dk_core_args ca;
    memset(&ca, 0, sizeof ca);
    ca.k = k;
    ca.offset = c->offset;
    ca.base = base;
    ca.nobsl = c->nobsl;
    ca.idx = c->csr_idx;
    ca.rloc = c->csr_w;
    ca.capacity = c->capacity;
    ca.max_p = c->prm.max_obs;
    ca.yb = c->yb_b;
    ca.yb32 = c->yb32_b;
    ca.ldy32 = c->ldy;
    ca.ymax = c->ymax;
    ca.rinv = c->rinv_b;
    ca.dep = c->dep_b;
    ca.gram = c->chained ? c->prm.gram : DIDA_GRAM_FP64;
    ca.rho = c->prm.inflation_factor;
    ca.tol = jacobi_tol_of(&c->prm);
    ca.max_sweeps = c->prm.jacobi_max_sweeps;
    ca.predict = predict_of(&c->prm);
    ca.trans = c->trans;
    ca.wbar = NULL;
    ca.info = c->info;
    dk_update_args ua;
    memset(&ua, 0, sizeof ua);
    ua.k = k;
    ua.nrow = c->prm.nrow;
    ua.trans = c->trans;
    ua.xb = xb;
    ua.axb = axb;
    ua.xa = xa;
    ua.axa = axa;
    ua.xb32 = c->xb32;
    ua.xa32 = c->xa32;
    ua.rtps = c->prm.rtps;
    ua.rtps_alpha = c->prm.rtps_alpha;
    /* chained analyses: launches end at multiples of the chain length, so
     * every chain is the same whichever launches a run is cut in */
    const int64_t L = c->chained ? c->chain_len : 1;
    const int64_t chunk = c->chunk >= L ? c->chunk / L * L : c->chunk;
#define LAUNCH_END(c0)                                                               \
    (chunk >= L ? ((c0) / L * L + chunk < a1 ? (c0) / L * L + chunk : a1)              \
                : ((c0) + chunk < a1 ? (c0) + chunk : a1))
    if (c->prm.cube_n > 1) {
        for (int64_t c0 = a0, c1; c0 < a1; c0 = c1) {
            c1 = LAUNCH_END(c0);
            int32_t n = (int32_t)(c1 - c0);
            ca.aid0 = (int32_t)c0;
            ca.nanal = n;
            ca.trans = c->trans_all + c0 * k * k;
            HIP_TRY(core_launch(c, &ca, s));
        }
        tmark(c, DIDA_STAGE_TRANSFORM, s);
        if (last) {
            HIP_TRY(cube_update(c, xb, axb, xa, axa, s));
            tmark(c, DIDA_STAGE_UPDATE, s);
        }
        return DIDA_OK;
    }
    for (int64_t c0 = a0, c1; c0 < a1; c0 = c1) {
        c1 = LAUNCH_END(c0);
        int32_t n = (int32_t)(c1 - c0);
        ca.aid0 = (int32_t)c0;
        ca.nanal = n;
        HIP_TRY(core_launch(c, &ca, s));
        tmark(c, DIDA_STAGE_TRANSFORM, s);
        ua.col0 = (int32_t)c0;
        ua.ncol = n;
        if (k <= 64 && !ua.xa32) HIP_TRY(dk_update_cols(&ua, s));
        else HIP_TRY(dk_update_f64(&ua, s));
        tmark(c, DIDA_STAGE_UPDATE, s);
    }
#undef LAUNCH_END
    return DIDA_OK;
}

/* Plain form with a workspace fixed at create: the analyses run in runs
 * whose selection lists fit it (the host reads the offsets once per sweep;
 * one run when everything fits). */
static int sweep_plain_chunked(dida_letkf_ctx *c, dk_select_args *sa, const double *xb,
                               const double *axb, double *xa, double *axa, hipStream_t s)
{
    const int64_t n = c->nsel, cap = c->capacity;
    int64_t *off = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    if (!off) return set_err(DIDA_ERR_NOMEM, "host offsets");
    int rc = DIDA_OK;
    if (hipMemcpyAsync(off, c->offset, (size_t)(n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = set_err(DIDA_ERR_HIP, "offsets copy failed");
    sa->csr_idx = c->csr_idx;
    sa->csr_w = c->csr_w;
    sa->capacity = cap;
    int64_t a0 = 0;
    while (!rc && a0 < n) {
        if (off[a0 + 1] - off[a0] > cap) {
            rc = set_err(DIDA_ERR_CAPACITY, "analysis %lld alone needs %lld selection entries > %lld",
                         (long long)a0, (long long)(off[a0 + 1] - off[a0]), (long long)cap);
            break;
        }
        int64_t lo = a0 + 1, hi = n; /* largest a1 with off[a1] - off[a0] <= cap */
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo + 1) / 2;
            if (off[mid] - off[a0] <= cap) lo = mid;
            else hi = mid - 1;
        }
        int64_t a1 = lo;
        if (c->chained && a1 < n) {
            /* end the run at a chain boundary when one is in reach (chains
             * are counted from analysis 0, so a run that ends inside a chain
             * is continued by the next one all the same) */
            const int64_t al = a1 / c->chain_len * c->chain_len;
            if (al > a0) a1 = al;
        }
        sa->col0 = a0;
        sa->ncols = a1 - a0;
        sa->base = off[a0];
        if (dk_select_fill(sa, s)) {
            rc = set_err(DIDA_ERR_HIP, "selection fill failed");
            break;
        }
        tmark(c, DIDA_STAGE_SELECT, s);
        rc = plain_transform(c, a0, a1, off[a0], a1 == n, xb, axb, xa, axa, s);
        a0 = a1;
    }
    sa->col0 = sa->ncols = sa->base = 0;
    free(off);
    return rc;
}

static int sweep_rest(dida_letkf_ctx *c, dk_select_args *sa, const double *xb,
                      const double *axb, double *xa, double *axa, hipStream_t s)
{
    sa->csr_idx = c->csr_idx;
    sa->csr_w = c->csr_w;
    sa->capacity = c->capacity;
    HIP_TRY(dk_select_fill(sa, s));
    tmark(c, DIDA_STAGE_SELECT, s);

    if (c->prm.form == DIDA_FORM_BATCHED && !c->chained)
        return sweep_batched(c, xb, axb, xa, axa, s);

    return plain_transform(c, 0, c->nsel, 0, 1, xb, axb, xa, axa, s);
}

static int check_sweep_args(dida_letkf_ctx *c, const dida_letkf_obs *o, const void *xb,
                            const double *axb, void *xa, double *axa)
{
    if (!c || !o || !xb || !axb || !xa || !axa) return set_err(DIDA_ERR_ARG, "null argument");
    if (!c->grid_lat) return set_err(DIDA_ERR_ARG, "grid not set (dida_letkf_set_grid)");
    if (o->nobs < 0 || o->nobs > c->nobs_max)
        return set_err(DIDA_ERR_ARG, "nobs = %d outside [0, %d]", o->nobs, c->nobs_max);
    if (o->nobs > 0 && (!o->lat || !o->lon || !o->yb || !o->dep || !o->rinv))
        return set_err(DIDA_ERR_ARG, "null obs array");
    if (c->prm.vert_dist > 0.0) {
        if (c->ncube <= 0)
            return set_err(DIDA_ERR_ARG, "vertical localisation: levels not set (dida_letkf_set_levels)");
        if (o->nobs > 0 && !o->ph)
            return set_err(DIDA_ERR_ARG, "vertical localisation: obs ph is null");
    }
    return DIDA_OK;
}

/* Selection, transforms and update of one sweep (arguments checked); with
 * c->xa32 set the perturbations are the FP32 ones and xb, xa are unused. */
static int sweep_impl(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                      const double *axb, double *xa, double *axa, void *stream)
{
    int rc;
    hipStream_t s = (hipStream_t)stream;
    dk_select_args sa;
    rc = sweep_select_count(c, o, s, &sa);
    if (rc) return rc;
    rc = ensure_capacity(c, s);
    if (rc) return rc;
    if ((c->prm.form == DIDA_FORM_PLAIN || c->chained) && c->user_capacity)
        rc = sweep_plain_chunked(c, &sa, xb, axb, xa, axa, s);
    else
        rc = sweep_rest(c, &sa, xb, axb, xa, axa, s);
    c->last_stream = s;
    c->last_nobs = o->nobs;
    return rc;
}

int dida_letkf_sweep(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                     const double *axb, double *xa, double *axa, void *stream)
{
    int rc = check_sweep_args(c, o, xb, axb, xa, axa);
    if (rc) return rc;
    return sweep_impl(c, o, xb, axb, xa, axa, stream);
}

int dida_letkf_sweep_f32(dida_letkf_ctx *c, const dida_letkf_obs *o, const float *xb,
                         const double *axb, float *xa, double *axa, void *stream)
{
    int rc = check_sweep_args(c, o, xb, axb, xa, axa);
    if (rc) return rc;
    if (c->prm.form == DIDA_FORM_BATCHED && !c->chained)
        return set_err(DIDA_ERR_UNSUPPORTED,
                       "FP32 state: plain form, or batched form with 64 < k <= 128 (the fused "
                       "k <= 64 kernel streams FP64 state)");
    c->xb32 = xb;
    c->xa32 = xa;
    rc = sweep_impl(c, o, NULL, axb, NULL, axa, stream);
    c->xb32 = NULL;
    c->xa32 = NULL;
    return rc;
}

int dida_letkf_graph_capture(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                             const double *axb, double *xa, double *axa, void *stream)
{
    int rc = check_sweep_args(c, o, xb, axb, xa, axa);
    if (rc) return rc;
    if (c->capacity <= 0)
        return set_err(DIDA_ERR_ARG, "run one sweep (or give sel_capacity) before capture");
    hipStream_t s = (hipStream_t)stream;
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = NULL;
    }
    if (c->graph) {
        (void)hipGraphDestroy(c->graph);
        c->graph = NULL;
    }
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int timing = c->timing;
    c->timing = 0;
    dk_select_args sa;
    rc = sweep_select_count(c, o, s, &sa);
    if (!rc) rc = sweep_rest(c, &sa, xb, axb, xa, axa, s);
    c->timing = timing;
    hipGraph_t g = NULL;
    int e = (int)hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e) return set_err(DIDA_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString((hipError_t)e));
    c->graph = g;
    HIP_TRY(hipGraphInstantiate(&c->gexec, g, NULL, NULL, 0));
    c->last_nobs = o->nobs;
    return DIDA_OK;
}

static int timed_sweep(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                       const double *axb, double *xa, double *axa, hipStream_t s,
                       hipEvent_t e0, hipEvent_t e1, double *ms)
{
    *ms = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
        HIP_TRY(hipEventRecord(e0, s));
        int rc = dida_letkf_sweep(c, o, xb, axb, xa, axa, s);
        if (rc < 0) return rc;
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        if (t < *ms) *ms = t;
    }
    return DIDA_OK;
}

/* Best of two sweeps' binning and selection stages alone, from the stage
 * marks (ms). */
static int timed_select(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                        const double *axb, double *xa, double *axa, hipStream_t s, double *ms)
{
    *ms = 1e300;
    const int timing = c->timing;
    c->timing = 1;
    int rc = DIDA_OK;
    for (int rep = 0; rep < 2 && !rc; ++rep) {
        rc = dida_letkf_sweep(c, o, xb, axb, xa, axa, s);
        if (rc < 0) break;
        rc = DIDA_OK;
        double st[DIDA_NSTAGES];
        rc = dida_letkf_stage_ms(c, st);
        if (!rc && st[DIDA_STAGE_BIN] + st[DIDA_STAGE_SELECT] < *ms)
            *ms = st[DIDA_STAGE_BIN] + st[DIDA_STAGE_SELECT];
    }
    c->timing = timing;
    return rc;
}

int dida_letkf_autotune(dida_letkf_ctx *c, const dida_letkf_obs *o, const double *xb,
                        const double *axb, double *xa, double *axa, void *stream,
                        dida_letkf_tuning *out)
{
    if (!out) return set_err(DIDA_ERR_ARG, "null tuning output");
    memset(out, 0, sizeof *out);
    int rc = check_sweep_args(c, o, xb, axb, xa, axa);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = dida_letkf_sweep(c, o, xb, axb, xa, axa, stream); /* sizes the lists */
    if (rc < 0) return rc;
    const int timing = c->timing;
    c->timing = 0;
    hipEvent_t e0 = NULL, e1 = NULL;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        c->timing = timing;
        if (e0) (void)hipEventDestroy(e0);
        return set_err(DIDA_ERR_HIP, "autotune: events");
    }
    const double bd0 = c->bin_deg;
    const int32_t cl0 = c->chain_len, tl0 = c->ngroup > 0;
    const int bucketed0 = c->bucketed;
    int32_t bl0[DK_NBUCKET];
    memcpy(bl0, c->bucket_len, sizeof bl0);
    double best_bd = bd0, best_ms = 0.0;
    int32_t best_cl = cl0, best_tl = tl0;
    int best_bk = bucketed0;
    int32_t best_bl[DK_NBUCKET];
    memcpy(best_bl, bl0, sizeof best_bl);
    /* 0. the tile of columns of the selection passes, judged by the binning
     * and selection stages alone (the whole sweep varies between runs by
     * more than the tiles differ); the initial time is taken with it */
    double best_sel = 1e300;
    int32_t best_tile = c->sel_tile;
    for (int32_t tt = DK_SELECT_TILE_4X2; tt <= DK_SELECT_TILE_4X4 && !rc; ++tt) {
        c->sel_tile = tt;
        double ms = 0.0;
        rc = timed_select(c, o, xb, axb, xa, axa, s, &ms);
        if (!rc && ms < best_sel) {
            best_sel = ms;
            best_tile = tt;
        }
    }
    c->sel_tile = best_tile;
    if (!rc) rc = timed_sweep(c, o, xb, axb, xa, axa, s, e0, e1, &best_ms);
    out->ms_initial = best_ms;
    out->ncand = 4;
    const int batched = c->prm.form == DIDA_FORM_BATCHED;
    /* one candidate: time it and keep it if faster */
#define TRY_LAYOUT(BD, CL, TL, BL)                                                       \
    do {                                                                                 \
        rc = set_layout(c, (BD), (CL), (TL), (BL));                                      \
        if (rc) break;                                                                   \
        double ms_ = 0.0;                                                                \
        rc = timed_sweep(c, o, xb, axb, xa, axa, s, e0, e1, &ms_);                       \
        out->ncand++;                                                                    \
        if (!rc && ms_ < best_ms) {                                                      \
            best_ms = ms_;                                                               \
            best_bd = (BD);                                                              \
            best_cl = (CL);                                                              \
            best_tl = (TL);                                                              \
            best_bk = (BL) != NULL;                                                      \
            if (BL) memcpy(best_bl, (BL), sizeof best_bl);                               \
        }                                                                                \
    } while (0)
    /* 1. the bin size (a capped column keeps the same obs at every size) */
    const double bins[3] = {c->theta_deg / 2.0, c->theta_deg / 4.0, c->theta_deg / 8.0};
    for (int ib = 0; ib < 3 && !rc; ++ib)
        if (bins[ib] != bd0 || bucketed0) TRY_LAYOUT(bins[ib], cl0, tl0, NULL);
    /* 2. one chain length for all analyses, and the queue order */
    const int32_t chains[6] = {2, 4, 8, 16, 32, 64};
    const double bd1 = best_bd;
    for (int ic = 0; ic < 6 && batched && !rc; ++ic)
        for (int it = 0; it < 2 && !rc; ++it)
            if (!(bd1 == best_bd && chains[ic] == best_cl && it == best_tl && !best_bk))
                TRY_LAYOUT(bd1, chains[ic], it, NULL);
    /* 3. fused form: a chain length per obs-count bucket (1: cold starts),
     * one bucket at a time from the best single length */
    int64_t bcount[DK_NBUCKET] = {0, 0, 0, 0};
    if (batched && !c->chained && !rc) {
        int32_t *nb = (int32_t *)malloc((size_t)c->nsel * sizeof(int32_t));
        if (!nb) rc = set_err(DIDA_ERR_NOMEM, "autotune: host counts");
        else if (hipMemcpy(nb, c->nobsl, (size_t)c->nsel * sizeof(int32_t),
                           hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_err(DIDA_ERR_HIP, "autotune: counts copy");
        for (int64_t i = 0; i < c->nsel && !rc; ++i) {
            int b = 0;
            while (b + 1 < DK_NBUCKET && nb[i] >= kBucketLo[b + 1]) ++b;
            bcount[b]++;
        }
        free(nb);
        int32_t bl[DK_NBUCKET];
        for (int b = 0; b < DK_NBUCKET; ++b) bl[b] = best_bk ? best_bl[b] : best_cl;
        const int32_t blens[7] = {1, 2, 4, 8, 16, 32, 64};
        const double bd2 = best_bd;
        const int32_t cl2 = best_cl;
        if (!best_bk && !rc) TRY_LAYOUT(bd2, cl2, 0, bl);
        for (int b = 0; b < DK_NBUCKET && !rc; ++b) {
            if (!bcount[b]) continue;
            const int32_t keep = bl[b];
            int32_t bestb = keep;
            for (int j = 0; j < 7 && !rc; ++j) {
                if (blens[j] == keep) continue;
                bl[b] = blens[j];
                const double before = best_ms;
                TRY_LAYOUT(bd2, cl2, 0, bl);
                if (best_ms < before) bestb = blens[j];
            }
            bl[b] = bestb;
        }
    }
#undef TRY_LAYOUT
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->timing = timing;
    int rc2 = set_layout(c, best_bd, best_cl, best_tl, best_bk ? best_bl : NULL);
    if (!rc) rc = rc2;
    if (rc) return rc;
    rc = dida_letkf_sweep(c, o, xb, axb, xa, axa, stream);
    if (rc < 0) return rc;
    out->chain_len = batched ? c->chain_len : 0;
    out->bin_deg = c->bin_deg;
    out->ms_best = best_ms;
    out->xcd_tiles = batched ? c->ngroup > 0 : 0;
    out->bucketed = c->bucketed;
    out->select_tile = c->sel_tile;
    for (int b = 0; b < DK_NBUCKET; ++b) {
        out->bucket_lo[b] = kBucketLo[b];
        out->bucket_chain[b] = c->bucketed ? c->bucket_len[b] : 0;
        out->bucket_count[b] = bcount[b];
    }
    return DIDA_OK;
}

struct dida_letkf_group {
    int32_t n;
    hipStream_t *branch;
    hipEvent_t *ev;      /* [n + 1]: fork, then one join event per branch */
    hipGraph_t graph;
    hipGraphExec_t exec;
    dida_letkf_ctx **ctx;  /* the contexts and their workspace generation at capture */
    int64_t *gen;
};

int dida_letkf_group_destroy(dida_letkf_group *g)
{
    if (!g) return DIDA_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (int i = 0; g->ev && i <= g->n; ++i)
        if (g->ev[i]) (void)hipEventDestroy(g->ev[i]);
    for (int i = 0; g->branch && i < g->n; ++i)
        if (g->branch[i]) (void)hipStreamDestroy(g->branch[i]);
    free(g->ev);
    free(g->branch);
    free(g->ctx);
    free(g->gen);
    free(g);
    return DIDA_OK;
}

int dida_letkf_group_capture(dida_letkf_group **out, int32_t n, dida_letkf_ctx *const *ctx,
                             const dida_letkf_obs *obs, const double *const *xb,
                             const double *const *axb, double *const *xa, double *const *axa,
                             void *stream)
{
    if (!out || n < 1 || !ctx || !obs || !xb || !axb || !xa || !axa)
        return set_err(DIDA_ERR_ARG, "group capture: null argument or n < 1");
    *out = NULL;
    for (int i = 0; i < n; ++i) {
        int rc = check_sweep_args(ctx[i], &obs[i], xb[i], axb[i], xa[i], axa[i]);
        if (rc) return rc;
        if (ctx[i]->capacity <= 0)
            return set_err(DIDA_ERR_ARG, "group capture: context %d has not run a sweep", i);
        for (int j = 0; j < i; ++j)
            if (ctx[j] == ctx[i]) return set_err(DIDA_ERR_ARG, "group capture: context repeated");
    }
    dida_letkf_group *g = (dida_letkf_group *)calloc(1, sizeof *g);
    if (!g) return set_err(DIDA_ERR_NOMEM, "group");
    g->n = n;
    g->branch = (hipStream_t *)calloc((size_t)n, sizeof(hipStream_t));
    g->ev = (hipEvent_t *)calloc((size_t)n + 1, sizeof(hipEvent_t));
    g->ctx = (dida_letkf_ctx **)calloc((size_t)n, sizeof(dida_letkf_ctx *));
    g->gen = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    if (!g->branch || !g->ev || !g->ctx || !g->gen) {
        dida_letkf_group_destroy(g);
        return set_err(DIDA_ERR_NOMEM, "group");
    }
    int rc = DIDA_OK;
    for (int i = 0; i < n && !rc; ++i)
        if (hipStreamCreateWithFlags(&g->branch[i], hipStreamNonBlocking) != hipSuccess)
            rc = set_err(DIDA_ERR_HIP, "group: stream");
    for (int i = 0; i <= n && !rc; ++i)
        if (hipEventCreateWithFlags(&g->ev[i], hipEventDisableTiming) != hipSuccess)
            rc = set_err(DIDA_ERR_HIP, "group: event");
    if (rc) {
        dida_letkf_group_destroy(g);
        return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        dida_letkf_group_destroy(g);
        return set_err(DIDA_ERR_HIP, "group: begin capture");
    }
    /* fork: every branch waits for the origin stream's point of capture */
    if (hipEventRecord(g->ev[0], s) != hipSuccess) rc = set_err(DIDA_ERR_HIP, "group: fork");
    for (int i = 0; i < n && !rc; ++i) {
        dida_letkf_ctx *c = ctx[i];
        hipStream_t b = g->branch[i];
        g->ctx[i] = c;
        g->gen[i] = c->generation;
        if (hipStreamWaitEvent(b, g->ev[0], 0) != hipSuccess) {
            rc = set_err(DIDA_ERR_HIP, "group: fork wait");
            break;
        }
        const int timing = c->timing;
        c->timing = 0;
        dk_select_args sa;
        rc = sweep_select_count(c, &obs[i], b, &sa);
        if (!rc) rc = sweep_rest(c, &sa, xb[i], axb[i], xa[i], axa[i], b);
        c->timing = timing;
        c->last_nobs = obs[i].nobs;
        c->last_stream = s;
        if (!rc && hipEventRecord(g->ev[i + 1], b) != hipSuccess)
            rc = set_err(DIDA_ERR_HIP, "group: join record");
        if (!rc && hipStreamWaitEvent(s, g->ev[i + 1], 0) != hipSuccess)
            rc = set_err(DIDA_ERR_HIP, "group: join wait");
    }
    hipGraph_t graph = NULL;
    int e = (int)hipStreamEndCapture(s, &graph);
    if (!rc && e) rc = set_err(DIDA_ERR_HIP, "group: end capture: %s",
                               hipGetErrorString((hipError_t)e));
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        dida_letkf_group_destroy(g);
        return rc;
    }
    g->graph = graph;
    if (hipGraphInstantiate(&g->exec, graph, NULL, NULL, 0) != hipSuccess) {
        dida_letkf_group_destroy(g);
        return set_err(DIDA_ERR_HIP, "group: instantiate");
    }
    *out = g;
    return DIDA_OK;
}

int dida_letkf_group_launch(dida_letkf_group *g, void *stream)
{
    if (!g || !g->exec) return set_err(DIDA_ERR_ARG, "no group graph");
    for (int i = 0; i < g->n; ++i)
        if (g->ctx[i]->generation != g->gen[i])
            return set_err(DIDA_ERR_ARG, "group graph: context %d replaced its selection "
                           "workspace after the capture (capture again)", i);
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return DIDA_OK;
}

int dida_letkf_graph_launch(dida_letkf_ctx *c, void *stream)
{
    if (!c || !c->gexec) return set_err(DIDA_ERR_ARG, "no captured graph");
    HIP_TRY(hipGraphLaunch(c->gexec, (hipStream_t)stream));
    c->last_stream = (hipStream_t)stream;
    return DIDA_OK;
}

int dida_letkf_set_timing(dida_letkf_ctx *c, int32_t on)
{
    if (!c) return set_err(DIDA_ERR_ARG, "null context");
    c->timing = on & 3;
    c->nmark = 0;
    if ((c->timing & 2) && !c->phase_clk && c->prm.form == DIDA_FORM_BATCHED &&
        hipMalloc((void **)&c->phase_clk, DK_NPHASE * sizeof(unsigned long long)) != hipSuccess)
        return set_err(DIDA_ERR_NOMEM, "phase clocks");
    return DIDA_OK;
}

int dida_letkf_phase_cycles(dida_letkf_ctx *c, double cyc[DIDA_NPHASES])
{
    if (!c || !cyc) return set_err(DIDA_ERR_ARG, "null argument");
    for (int i = 0; i < DIDA_NPHASES; ++i) cyc[i] = 0.0;
    if (!c->phase_clk) return DIDA_OK;
    unsigned long long h[DK_NPHASE];
    HIP_TRY(hipStreamSynchronize(c->last_stream));
    HIP_TRY(hipMemcpy(h, c->phase_clk, sizeof h, hipMemcpyDeviceToHost));
    for (int i = 0; i < DK_NPHASE && i < DIDA_NPHASES; ++i) cyc[i] = (double)h[i];
    return DIDA_OK;
}

int dida_letkf_stage_ms(dida_letkf_ctx *c, double ms[DIDA_NSTAGES])
{
    if (!c || !ms) return set_err(DIDA_ERR_ARG, "null argument");
    for (int i = 0; i < DIDA_NSTAGES; ++i) ms[i] = 0.0;
    if (c->nmark < 2) return DIDA_OK;
    HIP_TRY(hipEventSynchronize(c->ev[c->nmark - 1]));
    for (int i = 1; i < c->nmark; ++i) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, c->ev[i - 1], c->ev[i]));
        int st = c->ev_stage[i];
        if (st >= 0 && st < DIDA_NSTAGES) ms[st] += t;
    }
    return DIDA_OK;
}

int dida_letkf_stats_get(dida_letkf_ctx *c, dida_letkf_stats *st)
{
    if (!c || !st) return set_err(DIDA_ERR_ARG, "null argument");
    memset(st, 0, sizeof *st);
    HIP_TRY(hipStreamSynchronize(c->last_stream));
    const size_t n = (size_t)c->nsel;
    const size_t nq = c->ncube > 0 ? (size_t)c->ncube : 1; /* analyses per selection */
    int32_t *nb = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t *inf = (int32_t *)malloc(n * nq * sizeof(int32_t));
    int32_t flags[4];
    if (!nb || !inf) {
        free(nb);
        free(inf);
        return set_err(DIDA_ERR_NOMEM, "host buffers");
    }
    int rc = DIDA_OK;
    if (hipMemcpy(nb, c->nobsl, n * sizeof(int32_t), hipMemcpyDeviceToHost) ||
        hipMemcpy(inf, c->info, n * nq * sizeof(int32_t), hipMemcpyDeviceToHost) ||
        hipMemcpy(flags, c->flags, sizeof flags, hipMemcpyDeviceToHost)) {
        rc = set_err(DIDA_ERR_HIP, "stats copy failed");
    } else {
        st->nanal = (int64_t)(n * nq);
        for (size_t i = 0; i < n * nq; ++i) {
            const int32_t p = nb[i / nq];
            st->nobsl_sum += p;
            if (p > st->nobsl_max) st->nobsl_max = p;
            if (p >= c->prm.max_obs) st->nobsl_capped++;
            if (inf[i] == -2) st->nonfinite++;
            else if (inf[i] < 0) st->nonconverged++;
            else {
                st->sweeps_sum += inf[i];
                if (inf[i] > st->max_sweeps_used) st->max_sweeps_used = inf[i];
            }
        }
        st->capacity_overflow = flags[0];
        if (flags[0]) rc = set_err(DIDA_ERR_CAPACITY, "selection lists exceeded the workspace");
        else if (st->nonfinite)
            rc = set_err(DIDA_ERR_NONFINITE, "%d analyses met non-finite inputs (NaN transforms)",
                         st->nonfinite);
        else if (st->nonconverged) rc = DIDA_WARN_NOCONV;
    }
    free(nb);
    free(inf);
    return rc;
}

int dida_letkf_debug_arrays(dida_letkf_ctx *c, const int32_t **nobsl, const int64_t **offset,
                            const int32_t **csr_idx, const double **csr_w, const int32_t **perm,
                            const int32_t **info)
{
    if (!c) return set_err(DIDA_ERR_ARG, "null context");
    if (nobsl) *nobsl = c->nobsl;
    if (offset) *offset = c->offset;
    if (csr_idx) *csr_idx = c->csr_idx;
    if (csr_w) *csr_w = c->csr_w;
    if (perm) *perm = c->perm;
    if (info) *info = c->info;
    return DIDA_OK;
}

int dida_letkf_post(int32_t k, int32_t nrow, int32_t nlat, int32_t nlon, const float *grid_lat,
                    double polar_lat, double gamma1, double gamma2, const double *xb,
                    const double *axb, const double *xa, const double *axa, double *xa_out,
                    double *spread_b, double *spread_a, void *stream)
{
    if (k < 1 || nrow < 1 || nlat < 1 || nlon < 1) return set_err(DIDA_ERR_ARG, "empty shape");
    if (!grid_lat || !xb || !axb || !xa || !axa || !xa_out || !spread_b || !spread_a)
        return set_err(DIDA_ERR_ARG, "null argument");
    dk_post_args a = {k, nrow, nlat, nlon, grid_lat, polar_lat, gamma1, gamma2, xb, axb,
                      xa, axa, xa_out, spread_b, spread_a};
    HIP_TRY(dk_post(&a, stream));
    return DIDA_OK;
}

/* ------------------------------------------------------------------------ */
/* Dense batched core on device pointers. */
static _Thread_local int64_t *tl_off = NULL;
static _Thread_local int32_t tl_off_n = 0;
static _Thread_local int32_t tl_off_stride = 0;

int dida_letkf_core_batch_dev(int32_t batch, int32_t max_obs, int32_t nanals,
                              const int32_t *nobsl, const double *hxens, const float *rdiaginv,
                              const double *dep, const double *rloc, double *trans,
                              double *overline_w, double inflation_factor, int32_t gram,
                              int32_t *info, void *stream)
{
    if (batch < 0 || max_obs < 0) return set_err(DIDA_ERR_ARG, "negative size");
    if (nanals < 2 || nanals > 128)
        return set_err(DIDA_ERR_UNSUPPORTED, "nanals = %d: supported range is 2..128", nanals);
    if (gram != DIDA_GRAM_FP64) return set_err(DIDA_ERR_UNSUPPORTED, "gram %d", gram);
    if (!(inflation_factor > 0.0)) return set_err(DIDA_ERR_ARG, "inflation_factor must be > 0");
    if (batch == 0) return DIDA_OK;
    if (!nobsl || !trans || (max_obs > 0 && (!hxens || !rdiaginv || !dep || !rloc)))
        return set_err(DIDA_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    /* per-analysis offsets a * max_obs, kept per thread */
    if (tl_off_n < batch || tl_off_stride != max_obs) {
        if (tl_off) (void)hipFree(tl_off);
        tl_off = NULL;
        tl_off_n = 0;
        int64_t *h = (int64_t *)malloc((size_t)batch * sizeof(int64_t));
        if (!h) return set_err(DIDA_ERR_NOMEM, "host offsets");
        for (int32_t i = 0; i < batch; ++i) h[i] = (int64_t)i * max_obs;
        if (hipMalloc((void **)&tl_off, (size_t)batch * sizeof(int64_t)) != hipSuccess) {
            free(h);
            return set_err(DIDA_ERR_NOMEM, "offsets");
        }
        int e = (int)hipMemcpy(tl_off, h, (size_t)batch * sizeof(int64_t), hipMemcpyHostToDevice);
        free(h);
        if (e) return set_err(DIDA_ERR_HIP, "offset copy: %s", hipGetErrorString((hipError_t)e));
        tl_off_n = batch;
        tl_off_stride = max_obs;
    }
    dk_core_args ca;
    memset(&ca, 0, sizeof ca);
    ca.k = nanals;
    ca.nanal = batch;
    ca.offset = tl_off;
    ca.nobsl = nobsl;
    ca.capacity = (int64_t)batch * max_obs;
    ca.max_p = max_obs;
    ca.idx = NULL;
    ca.rloc = rloc;
    ca.yb = hxens;
    ca.rinv = rdiaginv;
    ca.dep = dep;
    ca.rho = inflation_factor;
    ca.tol = 1e-13;
    ca.max_sweeps = 30;
    ca.predict = 0;
    ca.trans = trans;
    ca.wbar = overline_w;
    ca.info = info;
    HIP_TRY(dk_letkf_core_f64(&ca, s));
    return DIDA_OK;
}

/* Host-pointer batch: stages through device buffers kept per thread. */
static _Thread_local struct {
    size_t c_hx, c_dep, c_rloc, c_ri, c_tr, c_wb, c_nb, c_info;
    double *hx, *dep, *rloc, *tr, *wb;
    float *ri;
    int32_t *nb, *info;
} tl;

static int grow(void **p, size_t *cap, size_t need)
{
    if (need == 0) need = 16;
    if (*cap >= need && *p) return 0;
    if (*p) (void)hipFree(*p);
    *p = NULL;
    *cap = 0;
    if (hipMalloc(p, need) != hipSuccess) return -1;
    *cap = need;
    return 0;
}

static int core_batch_host(int32_t batch, int32_t max_obs, int32_t k, const int32_t *nobsl,
                           const double *hxens, const float *rdiaginv, const double *dep,
                           const double *rloc, double *trans, double *wbar, double rho,
                           int32_t *info_h)
{
    if (batch < 0 || max_obs < 0 || k < 1) return set_err(DIDA_ERR_ARG, "bad sizes");
    const size_t nob = (size_t)batch * (size_t)max_obs;
    if (grow((void **)&tl.hx, &tl.c_hx, nob * (size_t)k * sizeof(double)) ||
        grow((void **)&tl.dep, &tl.c_dep, nob * sizeof(double)) ||
        grow((void **)&tl.rloc, &tl.c_rloc, nob * sizeof(double)) ||
        grow((void **)&tl.ri, &tl.c_ri, nob * sizeof(float)) ||
        grow((void **)&tl.tr, &tl.c_tr, (size_t)batch * k * k * sizeof(double)) ||
        grow((void **)&tl.wb, &tl.c_wb, (size_t)batch * k * sizeof(double)) ||
        grow((void **)&tl.nb, &tl.c_nb, (size_t)batch * sizeof(int32_t)) ||
        grow((void **)&tl.info, &tl.c_info, (size_t)batch * sizeof(int32_t)))
        return set_err(DIDA_ERR_NOMEM, "device staging buffers");
    if (nob) {
        HIP_TRY(hipMemcpy(tl.hx, hxens, nob * (size_t)k * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tl.dep, dep, nob * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tl.rloc, rloc, nob * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tl.ri, rdiaginv, nob * sizeof(float), hipMemcpyHostToDevice));
    }
    if (batch == 0) return DIDA_OK;
    HIP_TRY(hipMemcpy(tl.nb, nobsl, (size_t)batch * sizeof(int32_t), hipMemcpyHostToDevice));
    int rc = dida_letkf_core_batch_dev(batch, max_obs, k, tl.nb, tl.hx, tl.ri, tl.dep, tl.rloc,
                                       tl.tr, tl.wb, rho, DIDA_GRAM_FP64, tl.info, NULL);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(trans, tl.tr, (size_t)batch * k * k * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(wbar, tl.wb, (size_t)batch * k * sizeof(double), hipMemcpyDeviceToHost));
    if (info_h)
        HIP_TRY(hipMemcpy(info_h, tl.info, (size_t)batch * sizeof(int32_t), hipMemcpyDeviceToHost));
    return DIDA_OK;
}

void letkf_core_(const int32_t *nanals, const int32_t *nobsl, const double *hxens,
                 const float *rdiaginv, const double *dep, const double *rloc, double *trans,
                 double *overline_w, const double *inflation_factor, const int32_t *pflag)
{
    int32_t info = 0;
    int32_t p = *nobsl;
    int rc = core_batch_host(1, p, *nanals, &p, hxens, rdiaginv, dep, rloc, trans, overline_w,
                             *inflation_factor, &info);
    if (rc) fprintf(stderr, "letkf_core_: %s\n", dida_letkf_last_error());
    else if (info == -2) fprintf(stderr, "letkf_core_: non-finite input, trans is NaN\n");
    else if (info < 0) fprintf(stderr, "letkf_core_: eigensolver did not converge\n");
    if (pflag && *pflag) {
        printf("nobsl %d  jacobi sweeps %d\n", p, info);
        printf("overline_w");
        for (int i = 0; i < *nanals; ++i) printf(" %.6g", overline_w[i]);
        printf("\n");
    }
}

void letkf_core_batch_(const int32_t *batch, const int32_t *batch_large_matrix,
                       const int32_t *max_obs, const int32_t *nanals, const int32_t *nobsl,
                       double *hxens, const float *rdiaginv, const double *dep,
                       const double *rloc, double *trans, double *overline_w,
                       const double *inflation_factor, const int32_t *pflag)
{
    (void)batch_large_matrix;
    int rc = core_batch_host(*batch, *max_obs, *nanals, nobsl, hxens, rdiaginv, dep, rloc, trans,
                             overline_w, *inflation_factor, NULL);
    if (rc) fprintf(stderr, "letkf_core_batch_: %s\n", dida_letkf_last_error());
    if (pflag && *pflag) printf("letkf_core_batch_: batch %d done\n", *batch);
}

/* ------------------------------------------------------------------------ */
/* REAL(4) forms: the reference's letkf_core / letkf_core_batch built with
 * yb_type = ayb_type = real(4) (da_config.inc:7-10), where hxens, dep, rloc,
 * trans and overline_w are REAL(4).  The values are widened to FP64 on the
 * host, analysed by the FP64 kernels and rounded once to REAL(4). */
static double *widen(const float *x, size_t n)
{
    double *y = (double *)malloc((n ? n : 1) * sizeof(double));
    if (y)
        for (size_t i = 0; i < n; ++i) y[i] = (double)x[i];
    return y;
}

static int core_batch_host_r4(int32_t batch, int32_t max_obs, int32_t k, const int32_t *nobsl,
                              const float *hxens, const float *rdiaginv, const float *dep,
                              const float *rloc, float *trans, float *wbar, double rho,
                              int32_t *info)
{
    if (batch < 0 || max_obs < 0 || k < 1) return set_err(DIDA_ERR_ARG, "bad sizes");
    const size_t nob = (size_t)batch * (size_t)max_obs;
    const size_t ntr = (size_t)batch * k * k, nwb = (size_t)batch * k;
    double *hx = widen(hxens, nob * (size_t)k), *dp = widen(dep, nob), *rl = widen(rloc, nob);
    double *tr = (double *)malloc((ntr ? ntr : 1) * sizeof(double));
    double *wb = (double *)malloc((nwb ? nwb : 1) * sizeof(double));
    int rc = DIDA_OK;
    if (!hx || !dp || !rl || !tr || !wb) rc = set_err(DIDA_ERR_NOMEM, "host buffers");
    if (!rc)
        rc = core_batch_host(batch, max_obs, k, nobsl, hx, rdiaginv, dp, rl, tr, wb, rho, info);
    if (!rc) {
        for (size_t i = 0; i < ntr; ++i) trans[i] = (float)tr[i];
        for (size_t i = 0; i < nwb; ++i) wbar[i] = (float)wb[i];
    }
    free(hx);
    free(dp);
    free(rl);
    free(tr);
    free(wb);
    return rc;
}

void letkf_core_r4_(const int32_t *nanals, const int32_t *nobsl, const float *hxens,
                    const float *rdiaginv, const float *dep, const float *rloc, float *trans,
                    float *overline_w, const double *inflation_factor, const int32_t *pflag)
{
    int32_t info = 0;
    int32_t p = *nobsl;
    int rc = core_batch_host_r4(1, p, *nanals, &p, hxens, rdiaginv, dep, rloc, trans,
                                overline_w, *inflation_factor, &info);
    if (rc) fprintf(stderr, "letkf_core_r4_: %s\n", dida_letkf_last_error());
    else if (info == -2) fprintf(stderr, "letkf_core_r4_: non-finite input, trans is NaN\n");
    else if (info < 0) fprintf(stderr, "letkf_core_r4_: eigensolver did not converge\n");
    if (pflag && *pflag) {
        printf("nobsl %d  jacobi sweeps %d\n", p, info);
        printf("overline_w");
        for (int i = 0; i < *nanals; ++i) printf(" %.6g", (double)overline_w[i]);
        printf("\n");
    }
}

void letkf_core_batch_r4_(const int32_t *batch, const int32_t *batch_large_matrix,
                          const int32_t *max_obs, const int32_t *nanals, const int32_t *nobsl,
                          float *hxens, const float *rdiaginv, const float *dep,
                          const float *rloc, float *trans, float *overline_w,
                          const double *inflation_factor, const int32_t *pflag)
{
    (void)batch_large_matrix;
    int rc = core_batch_host_r4(*batch, *max_obs, *nanals, nobsl, hxens, rdiaginv, dep, rloc,
                                trans, overline_w, *inflation_factor, NULL);
    if (rc) fprintf(stderr, "letkf_core_batch_r4_: %s\n", dida_letkf_last_error());
    if (pflag && *pflag) printf("letkf_core_batch_r4_: batch %d done\n", *batch);
}

/* ------------------------------------------------------------------------ */
int dida_letkf_obs_operator(int32_t k, int32_t nrow, int32_t nlat, int32_t nlon,
                            const float *grid_lat, const float *grid_lon, int32_t periodic_lon,
                            const double *xb, const double *axb, int32_t ps_row,
                            const dida_vert_coord *vc, int32_t nvar, const int32_t *var_start,
                            const int32_t *var_lev, const dida_obs_records *rec, double *lat,
                            double *lon, double *yb, double *dep, double *ph, void *stream)
{
    if (!vc || !rec || !var_start || !var_lev || !grid_lat || !grid_lon || !xb || !axb)
        return set_err(DIDA_ERR_ARG, "obs operator: null argument");
    if (k < 1 || k > 128) return set_err(DIDA_ERR_UNSUPPORTED, "obs operator: k = %d (1..128)", k);
    if (nrow < 1 || nlat < 1 || nlon < 1) return set_err(DIDA_ERR_ARG, "obs operator: empty grid");
    if (nvar < 1 || nvar > DK_OBSOP_MAXVAR || rec->onv < 0 || rec->onv > DK_OBSOP_MAXVAR)
        return set_err(DIDA_ERR_UNSUPPORTED, "obs operator: at most %d variables",
                       DK_OBSOP_MAXVAR);
    if (rec->nrec < 0 || rec->onz < 0) return set_err(DIDA_ERR_ARG, "obs operator: negative size");
    if (ps_row < 0 || ps_row >= nrow) return set_err(DIDA_ERR_ARG, "obs operator: ps_row");
    if (vc->type != DIDA_VCOORD_SIGMA && vc->type != DIDA_VCOORD_HYBRID)
        return set_err(DIDA_ERR_ARG, "obs operator: vertical coordinate type");
    if (vc->nz < 0 || !vc->c1 || (vc->type == DIDA_VCOORD_HYBRID && !vc->c2))
        return set_err(DIDA_ERR_ARG, "obs operator: vertical coordinate");
    dk_obsop_args a;
    memset(&a, 0, sizeof a);
    for (int v = 0; v < nvar; ++v) {
        if (var_start[v] < 0 || var_lev[v] < 1 || var_start[v] + var_lev[v] > nrow ||
            (var_lev[v] > 1 && var_lev[v] != vc->nz))
            return set_err(DIDA_ERR_ARG, "obs operator: variable %d rows", v);
        a.vstart[v] = var_start[v];
        a.vlev[v] = var_lev[v];
    }
    for (int v = 0; v < rec->onv; ++v) {
        if (!rec->vid || rec->vid[v] < 0 || rec->vid[v] >= nvar)
            return set_err(DIDA_ERR_ARG, "obs operator: obs variable %d", v);
        a.vid[v] = rec->vid[v];
    }
    if (rec->nrec == 0 || rec->onz == 0 || rec->onv == 0) return DIDA_OK;
    if (!rec->lat || !rec->lon || !rec->ph || !rec->yo || !lat || !lon || !yb || !dep)
        return set_err(DIDA_ERR_ARG, "obs operator: null record or output array");
    a.k = k;
    a.nrow = nrow;
    a.nlat = nlat;
    a.nlon = nlon;
    a.periodic = periodic_lon ? 1 : 0;
    a.ps_row = ps_row;
    a.nz = vc->nz;
    a.vtype = vc->type;
    a.onz = rec->onz;
    a.onv = rec->onv;
    a.nrec = rec->nrec;
    a.grid_lat = grid_lat;
    a.grid_lon = grid_lon;
    a.xb = xb;
    a.axb = axb;
    a.c1 = vc->c1;
    a.c2 = vc->c2 ? vc->c2 : vc->c1;
    a.ptop = vc->ptop;
    a.p0 = vc->p0;
    a.rec_lat = rec->lat;
    a.rec_lon = rec->lon;
    a.rec_ph = rec->ph;
    a.yo = rec->yo;
    a.lat = lat;
    a.lon = lon;
    a.yb = yb;
    a.dep = dep;
    a.ph = ph;
    HIP_TRY(dk_obs_operator(&a, stream));
    return DIDA_OK;
}

static int check_struct(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                        int32_t n_max, int32_t ens_start, int32_t n_start, const void *a,
                        const void *b)
{
    if (!a || !b) return set_err(DIDA_ERR_ARG, "struct: null array");
    if (ens < 0 || n < 0 || nx < 0 || ny < 0) return set_err(DIDA_ERR_ARG, "struct: negative size");
    if (ens_start < 1 || ens_start + ens - 1 > ens_max || n_start < 1 || n_start + n - 1 > n_max)
        return set_err(DIDA_ERR_ARG,
                       "struct: block (members %d..%d, rows %d..%d) outside (%d, %d)", ens_start,
                       ens_start + ens - 1, n_start, n_start + n - 1, ens_max, n_max);
    return DIDA_OK;
}

int dida_struct_atm2endnn(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                          int32_t n_max, int32_t ens_start, int32_t n_start,
                          const double *xb_atm, double *xb, void *stream)
{
    int rc = check_struct(ens, n, nx, ny, ens_max, n_max, ens_start, n_start, xb_atm, xb);
    if (rc) return rc;
    HIP_TRY(dk_struct_ens(1, ens, n, nx, ny, ens_max, n_max, ens_start - 1, n_start - 1, xb_atm,
                          xb, stream));
    return DIDA_OK;
}

int dida_struct_endnn2atm(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                          int32_t n_max, int32_t ens_start, int32_t n_start, const double *xb,
                          double *xb_atm, void *stream)
{
    int rc = check_struct(ens, n, nx, ny, ens_max, n_max, ens_start, n_start, xb, xb_atm);
    if (rc) return rc;
    HIP_TRY(dk_struct_ens(0, ens, n, nx, ny, ens_max, n_max, ens_start - 1, n_start - 1, xb,
                          xb_atm, stream));
    return DIDA_OK;
}

int dida_struct_atm2endnn_a(int32_t n, int32_t nx, int32_t ny, int32_t n_max, int32_t n_start,
                            const double *axb_atm, double *axb, void *stream)
{
    int rc = check_struct(1, n, nx, ny, 1, n_max, 1, n_start, axb_atm, axb);
    if (rc) return rc;
    if (nx > 65535) return set_err(DIDA_ERR_UNSUPPORTED, "struct: nx > 65535");
    HIP_TRY(dk_struct_mean(1, n, nx, ny, n_max, n_start - 1, axb_atm, axb, stream));
    return DIDA_OK;
}

int dida_struct_endnn2atm_a(int32_t n, int32_t nx, int32_t ny, int32_t n_max, int32_t n_start,
                            const double *axb, double *axb_atm, void *stream)
{
    int rc = check_struct(1, n, nx, ny, 1, n_max, 1, n_start, axb, axb_atm);
    if (rc) return rc;
    if (nx > 65535) return set_err(DIDA_ERR_UNSUPPORTED, "struct: nx > 65535");
    HIP_TRY(dk_struct_mean(0, n, nx, ny, n_max, n_start - 1, axb, axb_atm, stream));
    return DIDA_OK;
}
