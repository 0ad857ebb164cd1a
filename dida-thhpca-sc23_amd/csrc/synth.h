/* Synthetic LETKF workload generator (benchmark and test inputs). */
#ifndef DIDA_SYNTH_H
#define DIDA_SYNTH_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct dida_synth_cfg {
    int32_t nlat, nlon, nz, k;
    int32_t nobs, nclusters;
    int32_t lmax_pert, pad_;
    double cluster_frac, cluster_radius_km;
    double sig_T, sig_uv;            /* obs error std before scaling */
    double rscale_min, rscale_max;   /* R scale, log-uniform */
    double pert_T, pert_uv, pert_ps; /* background perturbation std */
    double pert_len_km;              /* smoothing length of the perturbations */
    double vert_corr;                /* AR(1) coefficient between levels */
    uint64_t seed;
} dida_synth_cfg;
/* Number of spectral perturbation coefficients per (member, row). */
int dida_synth_pert_ncoef(const dida_synth_cfg *c);
/* Device scratch needed by dida_synth_generate. */
size_t dida_synth_scratch_bytes(const dida_synth_cfg *c);
/* Grid, state (xb FP64 or, if xb == NULL, xb_f32) and obs, all device
 * pointers, enqueued on stream.  Obs need the FP64 state. */
int dida_synth_generate(const dida_synth_cfg *c, void *scratch, float *glat, float *glon,
                        double *xb, float *xb_f32, double *axb, double *lat, double *lon,
                        double *yb, double *dep, float *rinv, void *stream);
#ifdef __cplusplus
}
#endif
#endif
