/*
 * CPU oracle for the LETKF local-analysis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke entry point and the CPU-baseline leg of the benchmark link it; the
 * product library never does.
 *
 * It is a plain-C FP64 restatement of the reference x86 path:
 *   orc_latlon_dist   code/x86/TradDA/DA/utils/pro_info.F90:148-161
 *   orc_schr2/schr    code/x86/TradDA/DA/LETKF/src/letkf_yb.F90:245-285
 *   orc_select_*      code/x86/TradDA/DA/LETKF/src/letkf_yb.F90:3-133 (gen_yb_dis)
 *   orc_letkf_core    code/x86/TradDA/DA/LETKF/src/letkf_core.F90:14-233
 *   orc_xb_2_xa       code/x86/TradDA/DA/LETKF/src/letkf_compute.F90:231-301
 *   orc_sweep         code/x86/TradDA/DA/LETKF/src/letkf_compute.F90:12-126
 *   orc_centre_obs    code/x86/TradDA/DA/LETKF/src/letkf_obs.F90:156-178
 *   orc_post          code/x86/TradDA/DA/EnDNN/src/module_da_parallel.F90:216-259
 *   orc_sweep_vertical code/x86/TradDA/DA/LETKF/src/letkf_compute.F90:358-526
 *                     (letkf_compute_nooverlap_vertical), letkf_yb.F90:137-243
 *                     (gen_yb_dis_vertical, vertical_localization),
 *                     obs_make.F90:418-437 (ps2alnp)
 *   orc_sweep_cubes   code/x86/TradDA/DA/LETKF/src/letkf_compute.F90:12-126 with
 *                     cube_n > 1, and :128-228 (letkf_compute, overlapping cubes)
 *                     with xb_add2_xa :304-354 and the cube_count average
 *   orc_struct_*      code/x86/TradDA/DA/EnDNN/src/data_trans.F90:2-57
 *   orc_obs_operator  code/x86/TradDA/DA/LETKF/src/letkf_obs.F90:120-178, 187-243
 *                     (find_obs_grid), obs_make.F90:233-381 (obs_interpolation_3d_2,
 *                     ps2ph_3d_interp), DA/utils/vert_coord_mod.F90:547-563
 * and keeps the reference's numeric quirks: single-precision literals in the
 * distance and Gaspari-Cohn code, REAL(4) R^-1 and grid coordinates, strict
 * ">" thresholds and MAX_OBS truncation in scan order.
 *
 * Pinning (tests/test_oracle_*.py): Gaspari-Cohn against its exact rational
 * closed form and the values the single-precision literals fix; distance
 * against the haversine formula; letkf_core against p = 0 (W = sqrt(rho) I),
 * p = 1 (Sherman-Morrison closed form), the invariants W = W^T,
 * W^2 = (k-1) Pa, Pa^-1 = (k-1)/rho I + C, W 1 = sqrt(rho) 1, 1^T wbar = 0,
 * and scipy eigh / long-double brute force; selection against a haversine
 * brute force; the update against a dense product.
 */
#ifndef DIDA_LETKF_ORACLE_H
#define DIDA_LETKF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Great-circle distance in km, FP64 arithmetic with the reference's REAL(4)
 * constants (a = 6371.393f, pi = 4*atan(1.0f)).  x = latitude, y = longitude,
 * degrees. */
double orc_latlon_dist(double x1, double y1, double x2, double y2);

/* Gaspari-Cohn fifth-order compact support (schr2) and the boxcar (schr). */
double orc_schr2(double r);
double orc_schr(double r);

/* Symmetric eigensolver (cyclic two-sided Jacobi, FP64).  a: n*n column-major,
 * only read as symmetric; on return holds eigenvectors (columns), w ascending
 * eigenvalues.  Returns the number of sweeps, or -1 if not converged. */
int orc_syev_jacobi(int n, double *a, double *w, int max_sweeps);

/* letkf_core: trans[i + j*k] = W_ij + wbar_i, wbar[k].  hxens: k*p, ensemble
 * fastest.  Returns eigensolver sweeps (or -1). */
int orc_letkf_core(int k, int p, const double *hxens, const float *rdiaginv,
                   const double *dep, const double *rloc, double *trans,
                   double *wbar, double inflation_factor);

/* xb_2_xa for ncells rows of one cube: xa[e + c*k] = sum_i trans[i + e*k] *
 * xb[i + c*k] + axb[c], then re-centred; axa[c] = ensemble mean. */
void orc_xb_2_xa(int k, int ncells, const double *xb, const double *axb,
                 const double *trans, double *xa, double *axa);

/* Observation arrays, scalar records in scan order (one entry per
 * (obs, level, variable) of the reference's obs_list, flattened in its
 * (l, m, n) loop order). */
typedef struct orc_obs {
    int32_t nobs;
    const double *lat;      /* degrees */
    const double *lon;      /* degrees */
    const double *yb;       /* k * nobs, ensemble fastest (pro_yb) */
    const double *dep;      /* nobs (dis_yo, already centred) */
    const float *rinv;      /* nobs (obs_rinv, REAL(4)) */
} orc_obs;

typedef struct orc_params {
    int32_t k;              /* ensemble size (nanals) */
    int32_t nrow;           /* state rows per column (n) */
    int32_t nlat, nlon;     /* grid; column c = ilat + nlat*ilon */
    int32_t max_obs;        /* MAX_OBS */
    int32_t gc_filter;      /* 1: schr2, 0: schr */
    double local_dist;      /* L, km */
    double obs_polar;       /* |lat| <= obs_polar */
    double inflation;       /* rho */
} orc_params;

/* One column's local selection by a scan of every obs in the given order
 * (gen_yb_dis with a single neighbour block).  Writes up to max_obs indices
 * and weights; returns nobsl. */
int orc_select_column(const orc_params *prm, const orc_obs *obs,
                      double center_lat, double center_lon,
                      int32_t *idx, double *rloc);

/* The observation bins used to prune the scan (shared geometry with the
 * device path): nband = ceil(180 / bin_deg) bands of 180 / nband degrees,
 * nlonbin = ceil(360 / bin_deg) longitude bins of 360 / nlonbin degrees, so
 * the bins tile the sphere exactly; band = floor((lat + 90) / dlat), lon bin
 * = floor(lon_wrapped / dlon); obs with |lat| > obs_polar go to the sentinel
 * bin nbins.  Fills key[nobs]; returns nbins. */
int orc_bin_keys(const orc_obs *obs, double bin_deg, double obs_polar,
                 int32_t *key, int32_t *nband, int32_t *nlonbin);

/* Stable counting sort of obs indices by key: perm[i] = original index of
 * the i-th obs in binned order; bin_start[nbins + 2]. */
void orc_bin_sort(int nobs, const int32_t *key, int nbins, int32_t *perm,
                  int32_t *bin_start);

/* Candidate bin ranges of a column (pairs of inclusive bin ids, ascending);
 * ranges must hold 4 * nband ints.  Returns the number of ranges. */
int orc_candidate_ranges(double clat, double clon, double local_dist,
                         double bin_deg, int nband, int nlonbin, int32_t *ranges);

/* Selection over the binned list (obs already permuted into binned order,
 * perm[l] = caller index of binned obs l, NULL for the identity), visiting
 * only candidate bins.  The kept set is that of orc_select_column on the
 * caller's list: every obs with gc > 0 or, when more than max_obs are kept,
 * the max_obs of smallest caller index (gen_yb_dis stops at MAX_OBS in its
 * scan order); listed in bin order.  Returns the count (-1: no memory). */
int orc_select_column_binned(const orc_params *prm, const orc_obs *obs,
                             const int32_t *bin_start, int nband, int nlonbin,
                             double bin_deg, const int32_t *perm, double center_lat,
                             double center_lon, int32_t *idx, double *rloc);

typedef struct orc_bins {
    const int32_t *bin_start; /* nbins + 2 */
    int32_t nband, nlonbin;
    double bin_deg;
    const int32_t *perm;      /* binned position -> caller index (NULL: identity) */
} orc_bins;

/* Whole sweep (letkf_compute_nooverlap with cube_n = 1): one analysis per
 * column, OpenMP over columns.  If cols != NULL only the ncols listed
 * columns are analysed (others untouched).  Outputs may be NULL except xa,
 * axa.  nobsl_out[col], trans_out[col*k*k], wbar_out[col*k] optional.
 * bins != NULL: obs are in binned order and only candidate bins are scanned.
 * Returns the number of analyses whose eigensolve did not converge. */
int orc_sweep(const orc_params *prm, const float *grid_lat,
              const float *grid_lon, const orc_obs *obs, const double *xb,
              const double *axb, double *xa, double *axa, const int32_t *cols,
              int32_t ncols, int32_t *nobsl_out, double *trans_out,
              double *wbar_out, const orc_bins *bins, int nthreads);

/* Horizontal cubes: the grid is tiled by cubes of cube_n x cube_n columns
 * whose starts advance by step = cube_n - overlap (cube_n when overlap = 0)
 * from the first column in each direction, clipped at the last one.  One
 * analysis per cube, with the selection centred on column
 * ((sj + ej) / 2, (si + ei) / 2) (1-based integer division, gen_yb_dis,
 * letkf_yb.F90:19-20).  overlap = 0: the cube's transform updates each of its
 * columns (xb_2_xa).  overlap > 0: each column accumulates, over the cubes
 * covering it in the reference's loop order (longitude starts outer,
 * latitude starts inner), dist_e = xbar_b + sum_m Xb'_m trans(m, e), its
 * member mean and dist_e - mean (xb_add2_xa), and is divided by the number of
 * covering cubes.  trans_out[a * k * k] (optional) holds the transform of
 * analysis a = aj + nalat ai.  Returns non-converged eigensolves. */
int orc_sweep_cubes(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                    const orc_obs *obs, const double *xb, const double *axb, double *xa,
                    double *axa, const orc_bins *bins, int cube_n, int overlap,
                    double *trans_out, int nthreads);

/* The sweep restricted to the listed columns with compact state arrays:
 * xb/xa [ncols * nrow * k] and axb/axa [ncols * nrow] hold column cols[w] at
 * position w.  Used for sampled CPU baselines. */
int orc_sweep_compact(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                      const orc_obs *obs, const orc_bins *bins, const int32_t *cols,
                      int32_t ncols, const double *xb, const double *axb, double *xa,
                      double *axa, int32_t *nobsl_out, int nthreads);

/* Vertical coordinate of the model levels: half-level pressure of level
 * j = 0..nz (vert_coord_mod.F90:547-563):
 *   sigma  (type 0): c1[j] (ps - ptop) + ptop
 *   hybrid (type 1): c1[j] (p0 - ptop) + c2[j] (ps - ptop) + ptop */
typedef struct orc_vcoord {
    int32_t type, nz;
    const double *c1, *c2;
    double ptop, p0;
} orc_vcoord;

/* Vertical localisation: for every horizontal cube of cube_n x cube_n
 * columns (cube_n = 1: every column; starts every cube_n columns, clipped,
 * selection and surface pressure at the 1-based centre (s + e) / 2), the
 * horizontal selection of orc_sweep, then for each vertical cube of cube_nz levels
 * (sk..ek, the last one shorter) an analysis with rloc multiplied by
 * schr2 (gc_filter_vertical) or schr of |ln p_c - ln p_obs| / vert_dist,
 * where p_c is the full-level pressure of level (sk + ek) / 2 (1-based) from
 * the centre column's mean surface pressure (axb at ps_row; levels from vc),
 * ln p_obs = log(obs_ph).  Each cube's transform updates rows sk..ek of every
 * 3-D variable and, with the last cube, the 2-D variables, in every column of
 * the horizontal cube; the rows are then
 * averaged and re-centred (xa, axa as orc_sweep).  The reference leaves the
 * bottom half level of ps2alnp unset (obs_make.F90:429-434); it is computed
 * here from vc.  bins as orc_sweep.  Returns non-converged eigensolves. */
int orc_sweep_vertical(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                       const orc_obs *obs, const double *obs_ph, const double *xb,
                       const double *axb, double *xa, double *axa, const orc_bins *bins,
                       const orc_vcoord *vc, int ps_row, int nvar, const int32_t *var_start,
                       const int32_t *var_lev, double vert_dist, int cube_nz,
                       int gc_filter_vertical, int cube_n, int nthreads);

/* Observation centring: ybar = mean_e hxb; dep = yo - ybar; yb = hxb - ybar. */
void orc_centre_obs(int k, int nobs, const double *hxb, const double *yo,
                    double *yb, double *dep);


/* Observation operator H(Xb) and centring.  State xb[(col * nrow + row) * k
 * + e] (perturbations) and axb[col * nrow + row] (mean), col = ilat + nlat
 * ilon, grid_lat/grid_lon ascending (REAL(4)); ps at row ps_row; variable
 * v of the state starts at row var_start[v] with var_lev[v] levels (1: a
 * 2-D variable).  Records r = 0..nrec-1 at (rec_lat, rec_lon) with onz
 * levels of pressure rec_ph[r * onz + l] (Pa) and onv variables vid[]; the
 * outputs are scalar obs s = (r * onz + l) * onv + v, in the order the
 * reference's selection appends them: lat/lon, yb[s * k + e] = H(x_e) -
 * mean_e, dep[s] = yo[s] - mean_e.  Per member: bilinear surface pressure,
 * full-level pressures (mean of half levels, level nz repeated), the level
 * search of obs_interpolation_3d_2 (carried over the record's levels in
 * order), log-pressure interpolation between the bracketing levels (the
 * nearest level outside them), bilinear in the horizontal.  Grid brackets:
 * the first grid latitude (longitude) >= the obs; an obs at or before the
 * first one takes that point alone (weight 0 on the missing neighbour, as
 * the reference's halo point gets); past the last latitude the last row
 * alone (the reference has no point there on a global grid); with
 * periodic_lon the longitude after the last wraps to the first + 360. */
void orc_obs_operator(int k, int nrow, int nlat, int nlon, const float *grid_lat,
                      const float *grid_lon, int periodic_lon, const double *xb,
                      const double *axb, int ps_row, const orc_vcoord *vc,
                      const int32_t *var_start, const int32_t *var_lev, int nrec, int onz,
                      int onv, const int32_t *vid, const double *rec_lat,
                      const double *rec_lon, const double *rec_ph, const double *yo,
                      double *lat, double *lon, double *yb, double *dep);

/* Post-analysis blend and spreads (run_da_parallel):
 *   mode 0: xa_out = gamma1 (xb + axb) + gamma2 (xa + axa) inside
 *           |lat| <= polar_lat, background outside; spreads as reference.
 *   mode 1: RTPS (not in the reference): xa' <- xa' (alpha sb/sa + 1 - alpha),
 *           xa_out = axa + xa'.
 * xb, xa: k*nrow*ncol; axb, axa: nrow*ncol; spread_b/a: nrow*ncol. */
void orc_post(int mode, int k, int nrow, int nlat, int nlon,
              const float *grid_lat, double polar_lat, double gamma1,
              double gamma2, double alpha, const double *xb,
              const double *axb, const double *xa, const double *axa,
              double *xa_out, double *spread_b, double *spread_a);

#ifdef __cplusplus
}
#endif
/* Coupler layout conversions (data_trans.F90:2-57), Fortran loops restated
 * with 1-based starts: xb_atm(ens, ny, nx, n) <-> xb(ens_max, n_max, nx, ny)
 * at members ens_start.., rows n_start..; means axb_atm(ny, nx, n) <->
 * axb(n_max, nx, ny). */
void orc_struct_atm2endnn(int ens, int n, int nx, int ny, int ens_max, int n_max, int ens_start,
                          int n_start, const double *xb_atm, double *xb);
void orc_struct_endnn2atm(int ens, int n, int nx, int ny, int ens_max, int n_max, int ens_start,
                          int n_start, const double *xb, double *xb_atm);
void orc_struct_atm2endnn_a(int n, int nx, int ny, int n_max, int n_start, const double *axb_atm,
                            double *axb);
void orc_struct_endnn2atm_a(int n, int nx, int ny, int n_max, int n_start, const double *axb,
                            double *axb_atm);

#endif
