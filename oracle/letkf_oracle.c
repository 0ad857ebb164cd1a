/*
 * CPU oracle for the LETKF local-analysis path -- TEST INFRASTRUCTURE ONLY.
 * Only tests, the smoke entry point and the benchmark's CPU-baseline leg use
 * it, as the checker.  See letkf_oracle.h for the reference functions each
 * routine restates and for how each one is pinned.
 *
 * Built with -ffp-contract=off so that every expression rounds in the order
 * it is written, as the reference's Fortran does.
 */
#include "letkf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* REAL(4) literals of the reference, widened exactly to double.
 * pro_info.F90:148 `#define a 6371.393` and :156 `pi = atan(1.0) * 4`
 * (single-precision atan of 1.0 times 4 is the float nearest pi). */
#define ORC_EARTH_A ((double)6371.393f)
#define ORC_PI_F32 ((double)3.14159274101257324f)
/* letkf_yb.F90:256,257,261 single-precision literals */
#define ORC_GC_5_3 ((double)1.666667f)
#define ORC_GC_2_3 ((double)0.6666667f)

/* pro_info.F90:148-161 latlon_dist_r8: x = latitude, y = longitude (deg). */
double orc_latlon_dist(double x1, double y1, double x2, double y2)
{
    const double pi = ORC_PI_F32;
    double x1r = x1 / 180.0 * pi;
    double x2r = x2 / 180.0 * pi;
    double y1r = y1 / 180.0 * pi;
    double y2r = y2 / 180.0 * pi;
    double v = cos(x1r) * cos(x2r) * cos(y1r - y2r) + sin(x1r) * sin(x2r);
    if (!(v < 1.0)) v = 1.0; /* Fortran min(v, 1.0) */
    return ORC_EARTH_A * acos(v);
}

/* letkf_yb.F90:245-269 schr2: Gaspari-Cohn (1999) eq. 4.10, r = d / L. */
double orc_schr2(double ri)
{
    double rci, rj, rk;
    if (ri > 2.0) {
        rci = 0.0;
    } else if (ri > 1.0) {
        rj = ri / 12.0 - 0.5;
        rj = rj * ri + 0.625;
        rj = rj * ri + ORC_GC_5_3;
        rci = (rj * ri - 5.0) * ri + 4.0 - ORC_GC_2_3 / ri;
    } else if (ri > 0.0) {
        rk = ri * ri;
        rj = (-0.25 * ri + 0.5) * ri + 0.625;
        rci = (rj * ri - ORC_GC_5_3) * rk + 1.0;
    } else {
        rci = 1.0;
    }
    return rci;
}

/* letkf_yb.F90:270-285 schr: boxcar of half-width 2L. */
double orc_schr(double ri) { return ri > 2.0 ? 0.0 : 1.0; }

/* ------------------------------------------------------------------------ */
/* Symmetric eigensolver: cyclic two-sided Jacobi with eigenvector
 * accumulation (Jacobi 1846; Rutishauser's formulation of the rotation).
 * Stands in for MKL dsyevd('V','L') at letkf_core.F90:98: same outputs
 * (ascending eigenvalues, orthonormal eigenvectors in columns). */
int orc_syev_jacobi(int n, double *a, double *w, int max_sweeps)
{
    double *v = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *s = (double *)malloc(sizeof(double) * (size_t)n * n);
    int i, j, r, sweep, conv = -1;
#define A_(i, j) s[(size_t)(i) + (size_t)(j) * n]
#define V_(i, j) v[(size_t)(i) + (size_t)(j) * n]
    /* symmetrise from the lower triangle ('L') */
    for (j = 0; j < n; ++j)
        for (i = j; i < n; ++i) {
            A_(i, j) = a[(size_t)i + (size_t)j * n];
            A_(j, i) = A_(i, j);
        }
    for (j = 0; j < n; ++j)
        for (i = 0; i < n; ++i) V_(i, j) = (i == j) ? 1.0 : 0.0;

    for (sweep = 1; sweep <= max_sweeps; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i) {
                double x = A_(i, j) * A_(i, j);
                tot += x;
                if (i != j) off += x;
            }
        if (off == 0.0 || off <= tot * (DBL_EPSILON * DBL_EPSILON) * 1e-4) {
            conv = sweep - 1;
            break;
        }
        int rotated = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A_(p, q);
                if (apq == 0.0) continue;
                double app = A_(p, p), aqq = A_(q, q);
                /* negligible against both diagonal entries: drop it */
                if (fabs(apq) <= 0.25 * DBL_EPSILON * sqrt(fabs(app * aqq)) &&
                    sweep > 3) {
                    A_(p, q) = A_(q, p) = 0.0;
                    continue;
                }
                double theta = (aqq - app) / (2.0 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
                double c = 1.0 / sqrt(t * t + 1.0);
                double sn = t * c;
                for (r = 0; r < n; ++r) {
                    if (r == p || r == q) continue;
                    double arp = A_(r, p), arq = A_(r, q);
                    A_(r, p) = c * arp - sn * arq;
                    A_(r, q) = sn * arp + c * arq;
                    A_(p, r) = A_(r, p);
                    A_(q, r) = A_(r, q);
                }
                A_(p, p) = app - t * apq;
                A_(q, q) = aqq + t * apq;
                A_(p, q) = A_(q, p) = 0.0;
                for (r = 0; r < n; ++r) {
                    double vp = V_(r, p), vq = V_(r, q);
                    V_(r, p) = c * vp - sn * vq;
                    V_(r, q) = sn * vp + c * vq;
                }
                rotated = 1;
            }
        if (!rotated) {
            conv = sweep;
            break;
        }
    }
    /* ascending order (LAPACK convention); selection sort keeps it simple */
    int *ord = (int *)malloc(sizeof(int) * (size_t)n);
    for (i = 0; i < n; ++i) ord[i] = i;
    for (i = 0; i < n; ++i) {
        int m = i;
        for (j = i + 1; j < n; ++j)
            if (A_(ord[j], ord[j]) < A_(ord[m], ord[m])) m = j;
        int tmp = ord[i];
        ord[i] = ord[m];
        ord[m] = tmp;
    }
    for (j = 0; j < n; ++j) {
        w[j] = A_(ord[j], ord[j]);
        for (i = 0; i < n; ++i) a[(size_t)i + (size_t)j * n] = V_(i, ord[j]);
    }
#undef A_
#undef V_
    free(ord);
    free(s);
    free(v);
    return conv;
}

/* ------------------------------------------------------------------------ */
/* letkf_core.F90:14-233, same operation order:
 *   rrloc = sqrt(R^-1 rho_loc)                         :64-66
 *   Yhat = Yb rrloc ; work1 = Yhat Yhat^T               :67, :74
 *   (V, lambda) = eig(work1) ; lambda += (k-1)/rho      :98, :110-113
 *   Pa = (V Lambda^-1) V^T                              :129-139
 *   Yhat *= rrloc ; work2 = Pa Yhat ; wbar = work2 dep  :146-148, :164, :190-195
 *   W = (V sqrt((k-1)/Lambda)) V^T ; trans = W + wbar   :198-206, :221-225 */
int orc_letkf_core(int k, int p, const double *hxens, const float *rdiaginv,
                   const double *dep, const double *rloc, double *trans,
                   double *wbar, double inflation_factor)
{
    const size_t kk = (size_t)k * k;
    double *hx = (double *)malloc(sizeof(double) * ((size_t)k * (p > 0 ? p : 1)));
    double *rr = (double *)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
    double *work1 = (double *)malloc(sizeof(double) * kk);
    double *eivec = (double *)malloc(sizeof(double) * kk);
    double *pa = (double *)malloc(sizeof(double) * kk);
    double *eival = (double *)malloc(sizeof(double) * (size_t)k);
    double *work2 = (double *)malloc(sizeof(double) * ((size_t)k * (p > 0 ? p : 1)));
    int i, j, m, o;

    for (o = 0; o < p; ++o) {
        rr[o] = (double)rdiaginv[o] * rloc[o];
        rr[o] = sqrt(rr[o]);
        for (i = 0; i < k; ++i)
            hx[(size_t)i + (size_t)o * k] = hxens[(size_t)i + (size_t)o * k] * rr[o];
    }
    /* dgemm('n','t',k,k,p): work1(i,j) = sum_o hx(i,o) hx(j,o) */
    for (j = 0; j < k; ++j)
        for (i = 0; i < k; ++i) {
            double s = 0.0;
            for (o = 0; o < p; ++o)
                s += hx[(size_t)i + (size_t)o * k] * hx[(size_t)j + (size_t)o * k];
            work1[(size_t)i + (size_t)j * k] = s;
        }

    int sweeps = orc_syev_jacobi(k, work1, eival, 100);

    for (m = 0; m < k; ++m) eival[m] = eival[m] * 1.0 + (double)(k - 1) / inflation_factor;

    for (j = 0; j < k; ++j)
        for (i = 0; i < k; ++i) {
            eivec[(size_t)i + (size_t)j * k] = work1[(size_t)i + (size_t)j * k];
            work1[(size_t)i + (size_t)j * k] = work1[(size_t)i + (size_t)j * k] / eival[j];
        }
    /* pa = work1 eivec^T */
    for (j = 0; j < k; ++j)
        for (i = 0; i < k; ++i) {
            double s = 0.0;
            for (m = 0; m < k; ++m)
                s += work1[(size_t)i + (size_t)m * k] * eivec[(size_t)j + (size_t)m * k];
            pa[(size_t)i + (size_t)j * k] = s;
        }
    for (o = 0; o < p; ++o)
        for (i = 0; i < k; ++i) hx[(size_t)i + (size_t)o * k] *= rr[o];
    /* work2 = pa hx (k x p) */
    for (o = 0; o < p; ++o)
        for (i = 0; i < k; ++i) {
            double s = 0.0;
            for (m = 0; m < k; ++m)
                s += pa[(size_t)i + (size_t)m * k] * hx[(size_t)m + (size_t)o * k];
            work2[(size_t)i + (size_t)o * k] = s;
        }
    for (i = 0; i < k; ++i) wbar[i] = 0.0;
    for (o = 0; o < p; ++o)
        for (i = 0; i < k; ++i) wbar[i] = wbar[i] + work2[(size_t)i + (size_t)o * k] * dep[o];
    /* W = (eivec sqrt((k-1)/eival)) eivec^T */
    for (j = 0; j < k; ++j) {
        double rho = sqrt((double)(k - 1) / eival[j]);
        for (i = 0; i < k; ++i)
            work1[(size_t)i + (size_t)j * k] = eivec[(size_t)i + (size_t)j * k] * rho;
    }
    for (j = 0; j < k; ++j)
        for (i = 0; i < k; ++i) {
            double s = 0.0;
            for (m = 0; m < k; ++m)
                s += work1[(size_t)i + (size_t)m * k] * eivec[(size_t)j + (size_t)m * k];
            trans[(size_t)i + (size_t)j * k] = s;
        }
    for (j = 0; j < k; ++j)
        for (i = 0; i < k; ++i) trans[(size_t)i + (size_t)j * k] += wbar[i];

    free(work2);
    free(eival);
    free(pa);
    free(eivec);
    free(work1);
    free(rr);
    free(hx);
    return sweeps;
}

/* letkf_compute.F90:231-301 xb_2_xa for a set of ncells rows. */
void orc_xb_2_xa(int k, int ncells, const double *xb, const double *axb,
                 const double *trans, double *xa, double *axa)
{
    for (int c = 0; c < ncells; ++c) {
        const double *x = xb + (size_t)c * k;
        double *y = xa + (size_t)c * k;
        for (int e = 0; e < k; ++e) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += trans[(size_t)i + (size_t)e * k] * x[i];
            y[e] = s;
        }
        for (int e = 0; e < k; ++e) y[e] = y[e] + axb[c];
        double m = 0.0;
        for (int e = 0; e < k; ++e) m = m + y[e];
        m = m / k;
        axa[c] = m;
        for (int e = 0; e < k; ++e) y[e] = y[e] - m;
    }
}

/* ------------------------------------------------------------------------ */
/* gen_yb_dis (letkf_yb.F90:71-131) with a single neighbour block: scan the
 * obs in order, keep |lat| <= obs_polar and gc > 0, stop at MAX_OBS. */
static inline int select_one(const orc_params *prm, const orc_obs *obs, int l,
                             double clat, double clon, double *gc_out)
{
    if (!(fabs(obs->lat[l]) <= prm->obs_polar)) return 0;
    double d = orc_latlon_dist(obs->lat[l], obs->lon[l], clat, clon);
    double r = d * 1.0 / prm->local_dist;
    double gc = prm->gc_filter ? orc_schr2(r) : orc_schr(r);
    *gc_out = gc;
    return gc > 0.0;
}

int orc_select_column(const orc_params *prm, const orc_obs *obs,
                      double center_lat, double center_lon, int32_t *idx,
                      double *rloc)
{
    int p = 0;
    for (int l = 0; l < obs->nobs; ++l) {
        double gc;
        if (select_one(prm, obs, l, center_lat, center_lon, &gc)) {
            idx[p] = l;
            rloc[p] = gc;
            ++p;
            if (p >= prm->max_obs) break;
        }
    }
    return p;
}

/* ------------------------------------------------------------------------ */
/* Observation bins (this build's pruning structure; not in the reference,
 * which prunes by MPI neighbour blocks).  The visiting order of candidate
 * bins is ascending bin id, so the selected sequence is the same as a full
 * scan of the binned list. */
static inline double wrap360(double lon)
{
    double w = fmod(lon, 360.0);
    if (w < 0.0) w += 360.0;
    if (w >= 360.0) w -= 360.0;
    return w;
}

int orc_bin_keys(const orc_obs *obs, double bin_deg, double obs_polar,
                 int32_t *key, int32_t *nband_out, int32_t *nlonbin_out)
{
    int nband = (int)ceil(180.0 / bin_deg);
    int nlonbin = (int)ceil(360.0 / bin_deg);
    int nbins = nband * nlonbin;
    const double dlat = 180.0 / nband, dlon = 360.0 / nlonbin;
    for (int l = 0; l < obs->nobs; ++l) {
        double lat = obs->lat[l];
        if (!(fabs(lat) <= obs_polar)) {
            key[l] = nbins;
            continue;
        }
        int b = (int)floor((lat + 90.0) / dlat);
        if (b < 0) b = 0;
        if (b > nband - 1) b = nband - 1;
        int c = (int)floor(wrap360(obs->lon[l]) / dlon);
        if (c < 0) c = 0;
        if (c > nlonbin - 1) c = nlonbin - 1;
        key[l] = b * nlonbin + c;
    }
    *nband_out = nband;
    *nlonbin_out = nlonbin;
    return nbins;
}

void orc_bin_sort(int nobs, const int32_t *key, int nbins, int32_t *perm,
                  int32_t *bin_start)
{
    for (int b = 0; b <= nbins + 1; ++b) bin_start[b] = 0;
    for (int l = 0; l < nobs; ++l) bin_start[key[l] + 1]++;
    for (int b = 0; b <= nbins; ++b) bin_start[b + 1] += bin_start[b];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nbins + 1));
    memcpy(fill, bin_start, sizeof(int32_t) * (size_t)(nbins + 1));
    for (int l = 0; l < nobs; ++l) perm[fill[key[l]]++] = l;
    free(fill);
}

/* Candidate bins of one column: bands b0..b1 and, per band, up to two
 * ascending lon-bin ranges.  The angular radius of the support (2L on a
 * sphere of radius a) is widened by 1e-6 relative plus 1e-6 deg so that the
 * float-rounded pi of the distance formula can never put a kept obs outside
 * the visited bins. */
int orc_candidate_ranges(double clat, double clon, double local_dist,
                         double bin_deg, int nband, int nlonbin, int32_t *ranges)
{
    const double pi = 3.14159265358979323846;
    const double dlat = 180.0 / nband, dlonb = 360.0 / nlonbin;
    (void)bin_deg;
    double theta = 2.0 * local_dist / ORC_EARTH_A * (180.0 / pi);
    theta = theta * (1.0 + 1e-6) + 1e-6;
    int b0 = (int)floor((clat - theta + 90.0) / dlat);
    int b1 = (int)floor((clat + theta + 90.0) / dlat);
    if (b0 < 0) b0 = 0;
    if (b1 > nband - 1) b1 = nband - 1;
    int full = 0;
    double dlon = 0.0;
    if (fabs(clat) + theta >= 90.0) {
        full = 1;
    } else {
        double sv = sin(theta * pi / 180.0) / cos(clat * pi / 180.0);
        if (sv >= 1.0) full = 1;
        else dlon = asin(sv) * 180.0 / pi * (1.0 + 1e-6) + 1e-6;
    }
    int lo = 0, hi = nlonbin - 1;
    if (!full) {
        double cw = wrap360(clon);
        lo = (int)floor((cw - dlon) / dlonb);
        hi = (int)floor((cw + dlon) / dlonb);
        if (hi - lo + 1 >= nlonbin) full = 1;
    }
    int nr = 0;
    for (int b = b0; b <= b1; ++b) {
        int base = b * nlonbin;
        if (full) {
            ranges[2 * nr] = base;
            ranges[2 * nr + 1] = base + nlonbin - 1;
            ++nr;
        } else if (lo < 0) {
            ranges[2 * nr] = base;
            ranges[2 * nr + 1] = base + hi;
            ++nr;
            ranges[2 * nr] = base + lo + nlonbin;
            ranges[2 * nr + 1] = base + nlonbin - 1;
            ++nr;
        } else if (hi > nlonbin - 1) {
            ranges[2 * nr] = base;
            ranges[2 * nr + 1] = base + hi - nlonbin;
            ++nr;
            ranges[2 * nr] = base + lo;
            ranges[2 * nr + 1] = base + nlonbin - 1;
            ++nr;
        } else {
            ranges[2 * nr] = base + lo;
            ranges[2 * nr + 1] = base + hi;
            ++nr;
        }
    }
    return nr;
}

static int cmp_i32(const void *a, const void *b)
{
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* The binned list is the caller's list permuted; perm[l] (NULL: l) is the
 * caller index of binned obs l.  gen_yb_dis appends kept obs in its scan
 * order and stops at MAX_OBS (letkf_yb.F90:98-115, letkf.F90:155); with one
 * global block that order is the caller's, so when more than max_obs obs
 * are kept the list holds the max_obs of smallest caller index (listed in
 * bin order).  Uncapped columns keep every obs with gc > 0. */
int orc_select_column_binned(const orc_params *prm, const orc_obs *obs,
                             const int32_t *bin_start, int nband, int nlonbin,
                             double bin_deg, const int32_t *perm, double center_lat,
                             double center_lon, int32_t *idx, double *rloc)
{
    int32_t ranges[4 * 400];
    int nr = orc_candidate_ranges(center_lat, center_lon, prm->local_dist, bin_deg,
                                  nband, nlonbin, ranges);
    const int max_obs = prm->max_obs;
    if (max_obs <= 0) return 0;
    int32_t *keys = NULL;
    size_t nkeys = 0, cap = 0;
    int p = 0;
    for (int q = 0; q < nr; ++q) {
        int l0 = bin_start[ranges[2 * q]], l1 = bin_start[ranges[2 * q + 1] + 1];
        for (int l = l0; l < l1; ++l) {
            double gc;
            if (!select_one(prm, obs, l, center_lat, center_lon, &gc)) continue;
            if (p < max_obs) {
                idx[p] = l;
                rloc[p] = gc;
                ++p;
            }
            if (nkeys == cap) {
                cap = cap ? 2 * cap : 4096;
                int32_t *nk = (int32_t *)realloc(keys, cap * sizeof(int32_t));
                if (!nk) {
                    free(keys);
                    return -1;
                }
                keys = nk;
            }
            keys[nkeys++] = perm ? perm[l] : l;
        }
    }
    if (nkeys <= (size_t)max_obs) {
        free(keys);
        return p;
    }
    /* MAX_OBS binds: keep the caller indices up to the max_obs-th smallest */
    qsort(keys, nkeys, sizeof(int32_t), cmp_i32);
    const int32_t last = keys[max_obs - 1];
    free(keys);
    p = 0;
    for (int q = 0; q < nr; ++q) {
        int l0 = bin_start[ranges[2 * q]], l1 = bin_start[ranges[2 * q + 1] + 1];
        for (int l = l0; l < l1; ++l) {
            double gc;
            if (!select_one(prm, obs, l, center_lat, center_lon, &gc)) continue;
            if ((perm ? perm[l] : l) <= last) {
                idx[p] = l;
                rloc[p] = gc;
                ++p;
            }
        }
    }
    return p;
}

/* ------------------------------------------------------------------------ */
int orc_sweep(const orc_params *prm, const float *grid_lat,
              const float *grid_lon, const orc_obs *obs, const double *xb,
              const double *axb, double *xa, double *axa, const int32_t *cols,
              int32_t ncols, int32_t *nobsl_out, double *trans_out,
              double *wbar_out, const orc_bins *bins, int nthreads)
{
    const int k = prm->k, n = prm->nrow;
    const int64_t ncol_all = (int64_t)prm->nlat * prm->nlon;
    const int64_t nwork = cols ? ncols : ncol_all;
    int nfail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : nfail)
    {
        int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)prm->max_obs);
        double *rl = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *hx = (double *)malloc(sizeof(double) * (size_t)prm->max_obs * k);
        double *dp = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        float *ri = (float *)malloc(sizeof(float) * (size_t)prm->max_obs);
        double *tr = (double *)malloc(sizeof(double) * (size_t)k * k);
        double *wb = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(dynamic, 4)
        for (int64_t w = 0; w < nwork; ++w) {
            int64_t col = cols ? cols[w] : w;
            int ilat = (int)(col % prm->nlat), ilon = (int)(col / prm->nlat);
            double clat = (double)grid_lat[ilat], clon = (double)grid_lon[ilon];
            int p = bins ? orc_select_column_binned(prm, obs, bins->bin_start, bins->nband,
                                                    bins->nlonbin, bins->bin_deg, bins->perm, clat,
                                                    clon, idx, rl)
                         : orc_select_column(prm, obs, clat, clon, idx, rl);
            for (int o = 0; o < p; ++o) {
                memcpy(hx + (size_t)o * k, obs->yb + (size_t)idx[o] * k, sizeof(double) * k);
                dp[o] = obs->dep[idx[o]];
                ri[o] = obs->rinv[idx[o]];
            }
            if (orc_letkf_core(k, p, hx, ri, dp, rl, tr, wb, prm->inflation) < 0) ++nfail;
            orc_xb_2_xa(k, n, xb + (size_t)col * n * k, axb + (size_t)col * n, tr,
                        xa + (size_t)col * n * k, axa + (size_t)col * n);
            if (nobsl_out) nobsl_out[col] = p;
            if (trans_out) memcpy(trans_out + (size_t)col * k * k, tr, sizeof(double) * k * k);
            if (wbar_out) memcpy(wbar_out + (size_t)col * k, wb, sizeof(double) * k);
        }
        free(wb);
        free(tr);
        free(ri);
        free(dp);
        free(hx);
        free(rl);
        free(idx);
    }
    return nfail;
}

int orc_sweep_cubes(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                    const orc_obs *obs, const double *xb, const double *axb, double *xa,
                    double *axa, const orc_bins *bins, int cube_n, int overlap,
                    double *trans_out, int nthreads)
{
    const int k = prm->k, n = prm->nrow, nlat = prm->nlat, nlon = prm->nlon;
    const int step = overlap > 0 ? cube_n - overlap : cube_n;
    if (cube_n < 1 || overlap < 0 || step < 1) return -1;
    const int nalat = (nlat + step - 1) / step, nalon = (nlon + step - 1) / step;
    const int64_t nanal = (int64_t)nalat * nalon;
    double *trans = trans_out ? trans_out
                              : (double *)malloc(sizeof(double) * (size_t)nanal * k * k);
    if (!trans) return -1;
    int nfail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    /* one analysis per cube (gen_yb_dis + letkf_core) */
#pragma omp parallel reduction(+ : nfail)
    {
        int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)prm->max_obs);
        double *rl = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *hx = (double *)malloc(sizeof(double) * (size_t)prm->max_obs * k);
        double *dp = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        float *ri = (float *)malloc(sizeof(float) * (size_t)prm->max_obs);
        double *wb = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(dynamic, 4)
        for (int64_t a = 0; a < nanal; ++a) {
            const int aj = (int)(a % nalat), ai = (int)(a / nalat);
            const int sj = aj * step, si = ai * step;
            const int ej = sj + cube_n - 1 < nlat - 1 ? sj + cube_n - 1 : nlat - 1;
            const int ei = si + cube_n - 1 < nlon - 1 ? si + cube_n - 1 : nlon - 1;
            /* 1-based (s + e) / 2, back to 0-based */
            const int cj = (sj + 1 + ej + 1) / 2 - 1, ci = (si + 1 + ei + 1) / 2 - 1;
            const double clat = (double)grid_lat[cj], clon = (double)grid_lon[ci];
            int p = bins ? orc_select_column_binned(prm, obs, bins->bin_start, bins->nband,
                                                    bins->nlonbin, bins->bin_deg, bins->perm, clat,
                                                    clon, idx, rl)
                         : orc_select_column(prm, obs, clat, clon, idx, rl);
            for (int o = 0; o < p; ++o) {
                memcpy(hx + (size_t)o * k, obs->yb + (size_t)idx[o] * k, sizeof(double) * k);
                dp[o] = obs->dep[idx[o]];
                ri[o] = obs->rinv[idx[o]];
            }
            if (orc_letkf_core(k, p, hx, ri, dp, rl, trans + (size_t)a * k * k, wb,
                               prm->inflation) < 0)
                ++nfail;
        }
        free(wb);
        free(ri);
        free(dp);
        free(hx);
        free(rl);
        free(idx);
    }
    /* the update of every column from the cubes covering it */
#pragma omp parallel
    {
        double *dist = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(static)
        for (int64_t col = 0; col < (int64_t)nlat * nlon; ++col) {
            const int ilat = (int)(col % nlat), ilon = (int)(col / nlat);
            const double *xbc = xb + (size_t)col * n * k;
            const double *axbc = axb + (size_t)col * n;
            double *xac = xa + (size_t)col * n * k;
            double *axac = axa + (size_t)col * n;
            if (overlap == 0) {
                const int64_t a = ilat / step + (int64_t)nalat * (ilon / step);
                orc_xb_2_xa(k, n, xbc, axbc, trans + (size_t)a * k * k, xac, axac);
                continue;
            }
            memset(xac, 0, sizeof(double) * (size_t)n * k);
            memset(axac, 0, sizeof(double) * (size_t)n);
            int count = 0;
            for (int ai = 0; ai < nalon; ++ai) {
                const int si = ai * step;
                const int ei = si + cube_n - 1 < nlon - 1 ? si + cube_n - 1 : nlon - 1;
                if (ilon < si || ilon > ei) continue;
                for (int aj = 0; aj < nalat; ++aj) {
                    const int sj = aj * step;
                    const int ej = sj + cube_n - 1 < nlat - 1 ? sj + cube_n - 1 : nlat - 1;
                    if (ilat < sj || ilat > ej) continue;
                    const double *tr = trans + ((size_t)aj + (size_t)nalat * ai) * k * k;
                    for (int v = 0; v < n; ++v) {
                        double mean = 0.0;
                        for (int e = 0; e < k; ++e) {
                            double d = 0.0;
                            for (int m = 0; m < k; ++m) d += xbc[(size_t)v * k + m] * tr[m + (size_t)e * k];
                            dist[e] = axbc[v] + d;
                            mean += dist[e];
                        }
                        mean /= k;
                        axac[v] += mean;
                        for (int e = 0; e < k; ++e) xac[(size_t)v * k + e] += dist[e] - mean;
                    }
                    ++count;
                }
            }
            for (int v = 0; v < n; ++v) {
                axac[v] /= count;
                for (int e = 0; e < k; ++e) xac[(size_t)v * k + e] /= count;
            }
        }
        free(dist);
    }
    if (!trans_out) free(trans);
    return nfail;
}

int orc_sweep_compact(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                      const orc_obs *obs, const orc_bins *bins, const int32_t *cols,
                      int32_t ncols, const double *xb, const double *axb, double *xa,
                      double *axa, int32_t *nobsl_out, int nthreads)
{
    const int k = prm->k, n = prm->nrow;
    int nfail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : nfail)
    {
        int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)prm->max_obs);
        double *rl = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *hx = (double *)malloc(sizeof(double) * (size_t)prm->max_obs * k);
        double *dp = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        float *ri = (float *)malloc(sizeof(float) * (size_t)prm->max_obs);
        double *tr = (double *)malloc(sizeof(double) * (size_t)k * k);
        double *wb = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(dynamic, 1)
        for (int32_t w = 0; w < ncols; ++w) {
            int64_t col = cols[w];
            int ilat = (int)(col % prm->nlat), ilon = (int)(col / prm->nlat);
            double clat = (double)grid_lat[ilat], clon = (double)grid_lon[ilon];
            int p = bins ? orc_select_column_binned(prm, obs, bins->bin_start, bins->nband,
                                                    bins->nlonbin, bins->bin_deg, bins->perm, clat,
                                                    clon, idx, rl)
                         : orc_select_column(prm, obs, clat, clon, idx, rl);
            for (int o = 0; o < p; ++o) {
                memcpy(hx + (size_t)o * k, obs->yb + (size_t)idx[o] * k, sizeof(double) * k);
                dp[o] = obs->dep[idx[o]];
                ri[o] = obs->rinv[idx[o]];
            }
            if (orc_letkf_core(k, p, hx, ri, dp, rl, tr, wb, prm->inflation) < 0) ++nfail;
            orc_xb_2_xa(k, n, xb + (size_t)w * n * k, axb + (size_t)w * n, tr,
                        xa + (size_t)w * n * k, axa + (size_t)w * n);
            if (nobsl_out) nobsl_out[w] = p;
        }
        free(wb);
        free(tr);
        free(ri);
        free(dp);
        free(hx);
        free(rl);
        free(idx);
    }
    return nfail;
}

/* letkf_obs.F90:156-178 */
void orc_centre_obs(int k, int nobs, const double *hxb, const double *yo,
                    double *yb, double *dep)
{
    for (int l = 0; l < nobs; ++l) {
        double s = 0.0;
        for (int e = 0; e < k; ++e) s = s + hxb[(size_t)l * k + e];
        s = s / k;
        dep[l] = yo[l] - s;
        for (int e = 0; e < k; ++e) yb[(size_t)l * k + e] = hxb[(size_t)l * k + e] - s;
    }
}

/* module_da_parallel.F90:216-259 (mode 0), plus RTPS (mode 1). */
void orc_post(int mode, int k, int nrow, int nlat, int nlon,
              const float *grid_lat, double polar_lat, double gamma1,
              double gamma2, double alpha, const double *xb,
              const double *axb, const double *xa, const double *axa,
              double *xa_out, double *spread_b, double *spread_a)
{
    for (int ilon = 0; ilon < nlon; ++ilon)
        for (int ilat = 0; ilat < nlat; ++ilat) {
            size_t col = (size_t)ilat + (size_t)nlat * ilon;
            int inside = fabs((double)grid_lat[ilat]) <= polar_lat;
            for (int r = 0; r < nrow; ++r) {
                size_t cr = col * nrow + r;
                const double *b = xb + cr * k;
                const double *a = xa + cr * k;
                double *o = xa_out + cr * k;
                double sb = 0.0, sa = 0.0;
                for (int e = 0; e < k; ++e) sb = sb + b[e] * b[e];
                if (mode == 0) {
                    if (inside) {
                        for (int e = 0; e < k; ++e)
                            o[e] = gamma1 * (axb[cr] + b[e]) + gamma2 * (axa[cr] + a[e]);
                        for (int e = 0; e < k; ++e)
                            sa = sa + (o[e] - axa[cr]) * (o[e] - axa[cr]);
                    } else {
                        for (int e = 0; e < k; ++e) o[e] = b[e] + axb[cr];
                        sa = sb;
                    }
                } else {
                    double s2 = 0.0;
                    for (int e = 0; e < k; ++e) s2 = s2 + a[e] * a[e];
                    double sig_b = sqrt(sb / k), sig_a = sqrt(s2 / k);
                    double fac = (sig_a > 0.0) ? alpha * sig_b / sig_a + 1.0 - alpha : 1.0;
                    for (int e = 0; e < k; ++e) o[e] = axa[cr] + a[e] * fac;
                    for (int e = 0; e < k; ++e)
                        sa = sa + (o[e] - axa[cr]) * (o[e] - axa[cr]);
                }
                spread_b[cr] = sqrt(sb / k);
                spread_a[cr] = sqrt(sa / k);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Observation operator (letkf_obs.F90:120-178, 187-243; obs_make.F90:233-381;
 * vert_coord_mod.F90:547-563).  Operation order follows the reference. */

/* find_obs_grid for one axis: i1 = first index with g[i1] >= x; returns the
 * weight of i0 (the point before).  Differences of REAL(4) grid values are
 * formed in single precision, as the reference's expression is. */
static double bracket(const float *g, int n, int periodic, double x, int *i0, int *i1)
{
    int j = 0;
    while (j < n && (double)g[j] < x) ++j;
    if (j == 0) { /* at or before the first point: that point alone */
        *i0 = *i1 = 0;
        return 0.0;
    }
    if (j == n) {
        if (!periodic) { /* past the last point: that point alone */
            *i0 = *i1 = n - 1;
            return 1.0;
        }
        const float gw = g[0] + 360.0f; /* the wrapped first point */
        if (!((double)gw >= x)) {
            *i0 = *i1 = 0;
            return 0.0;
        }
        const double d = (double)(float)(gw - g[n - 1]);
        *i0 = n - 1;
        *i1 = 0;
        return (d - x + (double)g[n - 1]) / d;
    }
    const double d = (double)(float)(g[j] - g[j - 1]);
    *i0 = j - 1;
    *i1 = j;
    return (d - x + (double)g[j - 1]) / d;
}

static double ph_half(const orc_vcoord *vc, int j, double ps)
{
    if (vc->type == 0) return vc->c1[j] * (ps - vc->ptop) + vc->ptop;
    return vc->c1[j] * (vc->p0 - vc->ptop) + vc->c2[j] * (ps - vc->ptop) + vc->ptop;
}

void orc_obs_operator(int k, int nrow, int nlat, int nlon, const float *grid_lat,
                      const float *grid_lon, int periodic_lon, const double *xb,
                      const double *axb, int ps_row, const orc_vcoord *vc,
                      const int32_t *var_start, const int32_t *var_lev, int nrec, int onz,
                      int onv, const int32_t *vid, const double *rec_lat,
                      const double *rec_lon, const double *rec_ph, const double *yo,
                      double *lat, double *lon, double *yb, double *dep)
{
    const int nz = vc->nz;
    double *ph = (double *)malloc(sizeof(double) * (size_t)(nz + 2));
    int *pos = (int *)malloc(sizeof(int) * (size_t)(onz > 0 ? onz : 1));
    double *h = (double *)malloc(sizeof(double) * (size_t)k);
    for (int r = 0; r < nrec; ++r) {
        int a0, a1, b0, b1;
        const double t1 = bracket(grid_lat, nlat, 0, rec_lat[r], &a0, &a1);
        const double t2 = bracket(grid_lon, nlon, periodic_lon, rec_lon[r], &b0, &b1);
        const size_t c00 = (size_t)a0 + (size_t)nlat * b0, c10 = (size_t)a1 + (size_t)nlat * b0;
        const size_t c01 = (size_t)a0 + (size_t)nlat * b1, c11 = (size_t)a1 + (size_t)nlat * b1;
#define X(col, row, e) (xb[((col) * (size_t)nrow + (size_t)(row)) * k + (e)] + \
                        axb[(col) * (size_t)nrow + (size_t)(row)])
#define HORIZ(row, e) (((X(c00, row, e)) * t1 + (X(c10, row, e)) * (1 - t1)) * t2 + \
                       ((X(c01, row, e)) * t1 + (X(c11, row, e)) * (1 - t1)) * (1 - t2))
        for (int l = 0; l < onz; ++l)
            for (int v = 0; v < onv; ++v) {
                const size_t sidx = ((size_t)r * onz + l) * onv + v;
                lat[sidx] = rec_lat[r];
                lon[sidx] = rec_lon[r];
            }
        for (int e = 0; e < k; ++e) {
            /* ps2ph_3d_interp */
            const double r1 = X(c00, ps_row, e) * t1 + X(c10, ps_row, e) * (1 - t1);
            const double r2 = X(c01, ps_row, e) * t1 + X(c11, ps_row, e) * (1 - t1);
            const double ps = r1 * t2 + r2 * (1 - t2);
            for (int j = 0; j < nz; ++j)
                ph[j] = 0.5 * (ph_half(vc, j, ps) + ph_half(vc, j + 1, ps));
            if (nz > 0) ph[nz] = ph[nz - 1];
            /* level search, carried over the record's levels */
            int j = 0;
            for (int l = 0; l < onz; ++l) {
                while (j < nz && rec_ph[(size_t)r * onz + l] > ph[j]) ++j;
                pos[l] = j; /* the reference's ph_pos (1-based) */
            }
            for (int l = 0; l < onz; ++l)
                for (int v = 0; v < onv; ++v) {
                    const int var = vid[v];
                    const size_t sidx = ((size_t)r * onz + l) * onv + v;
                    double val;
                    if (var_lev[var] > 1) {
                        const int p = pos[l];
                        if (p >= 1 && p < nz) {
                            const int tl = var_start[var] + p - 1;
                            const double po = rec_ph[(size_t)r * onz + l];
                            const double t3 = (log(ph[p]) - log(po)) / (log(ph[p]) - log(ph[p - 1]));
                            val = HORIZ(tl, e) * t3 + HORIZ(tl + 1, e) * (1 - t3);
                        } else {
                            const int tl = p == 0 ? var_start[var] : var_start[var] + nz - 1;
                            val = HORIZ(tl, e);
                        }
                    } else {
                        val = HORIZ(var_start[var], e);
                    }
                    yb[sidx * k + e] = val;
                }
        }
#undef HORIZ
#undef X
    }
    /* centring (letkf_obs.F90:156-178) */
    const int ns = nrec * onz * onv;
    for (int s2 = 0; s2 < ns; ++s2) {
        double s = 0.0;
        for (int e = 0; e < k; ++e) s = s + yb[(size_t)s2 * k + e];
        s = s / k;
        dep[s2] = yo[s2] - s;
        for (int e = 0; e < k; ++e) yb[(size_t)s2 * k + e] = yb[(size_t)s2 * k + e] - s;
    }
    free(h);
    free(pos);
    free(ph);
}

/* letkf_compute_nooverlap_vertical (letkf_compute.F90:358-526). */
int orc_sweep_vertical(const orc_params *prm, const float *grid_lat, const float *grid_lon,
                       const orc_obs *obs, const double *obs_ph, const double *xb,
                       const double *axb, double *xa, double *axa, const orc_bins *bins,
                       const orc_vcoord *vc, int ps_row, int nvar, const int32_t *var_start,
                       const int32_t *var_lev, double vert_dist, int cube_nz,
                       int gc_filter_vertical, int cube_n, int nthreads)
{
    const int k = prm->k, n = prm->nrow, nz = vc->nz, nlat = prm->nlat, nlon = prm->nlon;
    if (cube_n < 1) cube_n = 1;
    const int nalat = (nlat + cube_n - 1) / cube_n, nalon = (nlon + cube_n - 1) / cube_n;
    const int64_t nanal = (int64_t)nalat * nalon;
    int nfail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : nfail)
    {
        int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)prm->max_obs);
        double *rl = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *rlv = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *olnp = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        double *hx = (double *)malloc(sizeof(double) * (size_t)prm->max_obs * k);
        double *dp = (double *)malloc(sizeof(double) * (size_t)prm->max_obs);
        float *ri = (float *)malloc(sizeof(float) * (size_t)prm->max_obs);
        double *tr = (double *)malloc(sizeof(double) * (size_t)k * k);
        double *wb = (double *)malloc(sizeof(double) * (size_t)k);
        double *alnp = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
#pragma omp for schedule(dynamic, 4)
        for (int64_t an = 0; an < nanal; ++an) {
            /* horizontal cube: columns sj..ej x si..ei, centred on the
             * 1-based (s + e) / 2 (letkf_compute.F90:409-425) */
            const int aj = (int)(an % nalat), ai = (int)(an / nalat);
            const int sj = aj * cube_n, si = ai * cube_n;
            const int ej = sj + cube_n - 1 < nlat - 1 ? sj + cube_n - 1 : nlat - 1;
            const int ei = si + cube_n - 1 < nlon - 1 ? si + cube_n - 1 : nlon - 1;
            const int cj = (sj + 1 + ej + 1) / 2 - 1, ci = (si + 1 + ei + 1) / 2 - 1;
            const int64_t ccol = cj + (int64_t)nlat * ci;
            double clat = (double)grid_lat[cj], clon = (double)grid_lon[ci];
            int p = bins ? orc_select_column_binned(prm, obs, bins->bin_start, bins->nband,
                                                    bins->nlonbin, bins->bin_deg, bins->perm, clat,
                                                    clon, idx, rl)
                         : orc_select_column(prm, obs, clat, clon, idx, rl);
            for (int o = 0; o < p; ++o) {
                memcpy(hx + (size_t)o * k, obs->yb + (size_t)idx[o] * k, sizeof(double) * k);
                dp[o] = obs->dep[idx[o]];
                ri[o] = obs->rinv[idx[o]];
                olnp[o] = log(obs_ph[idx[o]]);
            }
            /* ps2alnp from the centre column's mean surface pressure */
            const double aps = axb[(size_t)ccol * n + ps_row];
            for (int i = 0; i < nz; ++i) {
                alnp[i] = 0.5 * (ph_half(vc, i, aps) + ph_half(vc, i + 1, aps));
                alnp[i] = log(alnp[i]);
            }
            for (int sk = 0; sk < nz; sk += cube_nz) {
                int ek = sk + cube_nz - 1;
                if (ek > nz - 1) ek = nz - 1;
                const int ck = (sk + 1 + ek + 1) / 2 - 1;
                for (int o = 0; o < p; ++o) {
                    const double r = fabs(alnp[ck] - olnp[o]) / vert_dist;
                    rlv[o] = rl[o] * (gc_filter_vertical ? orc_schr2(r) : orc_schr(r));
                }
                if (orc_letkf_core(k, p, hx, ri, dp, rlv, tr, wb, prm->inflation) < 0) ++nfail;
                for (int ilon = si; ilon <= ei; ++ilon)
                    for (int ilat = sj; ilat <= ej; ++ilat) {
                        const int64_t col = ilat + (int64_t)nlat * ilon;
                        for (int v = 0; v < nvar; ++v) {
                            int r0, r1;
                            if (var_lev[v] == 1) {
                                if (ek != nz - 1) continue;
                                r0 = r1 = var_start[v];
                            } else {
                                r0 = var_start[v] + sk;
                                r1 = var_start[v] + ek;
                            }
                            for (int row = r0; row <= r1; ++row)
                                orc_xb_2_xa(k, 1, xb + ((size_t)col * n + row) * k,
                                            axb + (size_t)col * n + row, tr,
                                            xa + ((size_t)col * n + row) * k,
                                            axa + (size_t)col * n + row);
                        }
                    }
            }
        }
        free(alnp);
        free(wb);
        free(tr);
        free(ri);
        free(dp);
        free(hx);
        free(olnp);
        free(rlv);
        free(rl);
        free(idx);
    }
    return nfail;
}

/* data_trans.F90:2-20 (struct_atm2endnn): xb2(ens_start:, nn, i, j) =
 * xb(1:ens, j, i, nn - n_start + 1), Fortran column-major. */
void orc_struct_atm2endnn(int ens, int n, int nx, int ny, int ens_max, int n_max, int ens_start,
                          int n_start, const double *xb_atm, double *xb)
{
    for (int nn = n_start; nn <= n_start + n - 1; ++nn)
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j)
                for (int e = 1; e <= ens; ++e) {
                    const size_t src = (size_t)(e - 1) +
                                       (size_t)ens * ((j - 1) + (size_t)ny * ((i - 1) +
                                       (size_t)nx * (nn - n_start)));
                    const size_t dst = (size_t)(ens_start + e - 2) +
                                       (size_t)ens_max * ((nn - 1) + (size_t)n_max * ((i - 1) +
                                       (size_t)nx * (j - 1)));
                    xb[dst] = xb_atm[src];
                }
}

/* data_trans.F90:22-38 (struct_endnn2atm). */
void orc_struct_endnn2atm(int ens, int n, int nx, int ny, int ens_max, int n_max, int ens_start,
                          int n_start, const double *xb, double *xb_atm)
{
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int nn = 1; nn <= n; ++nn)
                for (int e = 1; e <= ens; ++e) {
                    const size_t dst = (size_t)(e - 1) +
                                       (size_t)ens * ((j - 1) + (size_t)ny * ((i - 1) +
                                       (size_t)nx * (nn - 1)));
                    const size_t src = (size_t)(ens_start + e - 2) +
                                       (size_t)ens_max * ((nn + n_start - 2) + (size_t)n_max *
                                       ((i - 1) + (size_t)nx * (j - 1)));
                    xb_atm[dst] = xb[src];
                }
}

/* data_trans.F90:40-50 (struct_atm2endnn_a). */
void orc_struct_atm2endnn_a(int n, int nx, int ny, int n_max, int n_start, const double *axb_atm,
                            double *axb)
{
    for (int nn = n_start; nn <= n_start + n - 1; ++nn)
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j)
                axb[(size_t)(nn - 1) + (size_t)n_max * ((i - 1) + (size_t)nx * (j - 1))] =
                    axb_atm[(size_t)(j - 1) + (size_t)ny * ((i - 1) + (size_t)nx * (nn - n_start))];
}

/* data_trans.F90:52-57 (struct_endnn2atm_a). */
void orc_struct_endnn2atm_a(int n, int nx, int ny, int n_max, int n_start, const double *axb,
                            double *axb_atm)
{
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int nn = 1; nn <= n; ++nn)
                axb_atm[(size_t)(j - 1) + (size_t)ny * ((i - 1) + (size_t)nx * (nn - 1))] =
                    axb[(size_t)(nn + n_start - 2) + (size_t)n_max * ((i - 1) + (size_t)nx * (j - 1))];
}
