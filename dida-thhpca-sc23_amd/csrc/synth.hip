// Synthetic LETKF workloads on the GPU (benchmark and test inputs; not part
// of the analysis library).  Everything is a pure function of the seed and the
// element index (counter-based RNG), so a workload is reproducible bit for
// bit on any MI355X.
//
//   grid      lat_j = -90 + (j + 1/2) 180/nlat, lon_i = i 360/nlon (REAL(4))
//   truth     real spherical harmonics of degree <= lmax_truth per variable,
//             linear in level between two random fields; T 250 +- 30 K,
//             u/v +- 30 m/s, ps 1000 +- 20 hPa
//   members   truth + smoothed noise: spectral coefficients N(0,1) damped by
//             exp(-l(l+1) (s/a)^2 / 2) (a Gaussian kernel of length s on the
//             sphere), AR(1) in level; sigma 1 K, 1.5 m/s, 1 hPa
//   state     xbar_b = member mean, Xb' = members - xbar_b, rows ps, u(nz),
//             v(nz), T(nz), layout (k, n, nlat, nlon)
//   obs       uniform on the sphere or in Gaussian clusters; one of u/v/T at a
//             random continuous level; H bilinear in lat/lon and linear in
//             level; yo = H(truth) + N(0, R); R = sigma^2 * scale with scale
//             log-uniform; R^-1 in REAL(4); centring as letkf_obs.F90:156-178.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "synth.h"

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline double unif(uint64_t seed, uint32_t stream, uint64_t i) {
  uint64_t r = mix64(seed ^ mix64(((uint64_t)stream << 44) ^ i));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double gauss(uint64_t seed, uint32_t stream, uint64_t i) {
  double u1 = unif(seed, stream, 2 * i), u2 = unif(seed, stream, 2 * i + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

enum Stream : uint32_t {
  kTruth = 1, kPert = 2, kObsPos = 3, kObsVar = 4, kObsLev = 5, kObsNoise = 6,
  kObsScale = 7, kCluster = 8
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kA = 6371.0;

// Fully normalised associated Legendre functions P[l(l+1)/2 + m], l <= L.
__host__ __device__ inline void legendre(int L, double x, double* P) {
  const double c = sqrt(fmax(0.0, 1.0 - x * x));
  P[0] = 1.0;
  for (int m = 0; m <= L; ++m) {
    double pmm;
    if (m == 0) pmm = 1.0;
    else pmm = sqrt((2.0 * m + 1.0) / (2.0 * m)) * c * P[(m - 1) * m / 2 + (m - 1)];
    if (m > 0) P[m * (m + 1) / 2 + m] = pmm;
    if (m + 1 <= L) P[(m + 1) * (m + 2) / 2 + m] = sqrt(2.0 * m + 3.0) * x * pmm;
    for (int l = m + 2; l <= L; ++l) {
      double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
      double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
      P[l * (l + 1) / 2 + m] = a * (x * P[(l - 1) * l / 2 + m] - b * P[(l - 2) * (l - 1) / 2 + m]);
    }
  }
}

__global__ void k_grid(int nlat, int nlon, float* glat, float* glon) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nlat) glat[t] = (float)(-90.0 + (t + 0.5) * 180.0 / nlat);
  if (t < nlon) glon[t] = (float)(t * 360.0 / nlon);
}

// Truth: 4 variables x 2 level bases of degree <= LT harmonics.
constexpr int kLT = 10;
constexpr int kNT = (kLT + 1) * (kLT + 2) / 2;

__global__ void k_truth(dida_synth_cfg c, const float* glat, const float* glon, double* truth) {
  const int64_t ncol = (int64_t)c.nlat * c.nlon;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int j = (int)(col % c.nlat), i = (int)(col / c.nlat);
  const double phi = glat[j] * kPi / 180.0, lam = glon[i] * kPi / 180.0;
  double P[kNT];
  legendre(kLT, sin(phi), P);
  double Y[4][2];
  for (int v = 0; v < 4; ++v)
    for (int b = 0; b < 2; ++b) {
      double s = 0.0;
      int q = 0;
      for (int l = 0; l <= kLT; ++l)
        for (int m = 0; m <= l; ++m, ++q) {
          const uint64_t id = ((uint64_t)(v * 2 + b) << 16) + (uint64_t)q * 2;
          const double fm = m ? 1.4142135623730951 : 1.0;
          s += P[l * (l + 1) / 2 + m] * fm *
               (gauss(c.seed, kTruth, id) * cos(m * lam) + (m ? gauss(c.seed, kTruth, id + 1) * sin(m * lam) : 0.0));
        }
      Y[v][b] = s / sqrt((double)((kLT + 1) * (kLT + 1)));
    }
  const int n = 1 + 3 * c.nz;
  double* t = truth + col * n;
  t[0] = 1000.0 + 20.0 * Y[0][0];
  for (int v = 0; v < 3; ++v) {
    const double amp = v < 2 ? 30.0 : 30.0, base = v < 2 ? 0.0 : 250.0;
    for (int z = 0; z < c.nz; ++z) {
      const double w = c.nz > 1 ? (double)z / (c.nz - 1) : 0.0;
      t[1 + v * c.nz + z] = base + amp * (Y[v + 1][0] + w * (Y[v + 1][1] - Y[v + 1][0]));
    }
  }
}

// Perturbation spectral coefficients coef[(e * n + row) * NC + q], AR(1) in
// level within each 3-D variable; q runs over (l, m, cos/sin).
__global__ void k_pert_coef(dida_synth_cfg c, int NC, float* coef) {
  const int LP = c.lmax_pert;
  const int n = 1 + 3 * c.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)c.k * 4 * NC;
  if (t >= total) return;
  const int q = (int)(t % NC);
  const int v = (int)((t / NC) % 4);
  const int e = (int)(t / ((int64_t)NC * 4));
  // q -> degree l
  int l = 0, rem = q;
  while (rem >= 2 * l + 1) { rem -= 2 * l + 1; ++l; }
  const double s = c.pert_len_km / kA;
  double norm = 0.0;
  for (int ll = 0; ll <= LP; ++ll) norm += (2.0 * ll + 1.0) * exp(-ll * (ll + 1.0) * s * s);
  const double amp = exp(-0.5 * l * (l + 1.0) * s * s) / sqrt(norm);
  const uint64_t base = (((uint64_t)e * 4 + v) << 24) + (uint64_t)q * 64;
  if (v == 0) {
    coef[((int64_t)e * n + 0) * NC + q] = (float)(amp * gauss(c.seed, kPert, base));
    return;
  }
  const double rv = c.vert_corr, sv = sqrt(1.0 - rv * rv);
  double x = gauss(c.seed, kPert, base);
  for (int z = 0; z < c.nz; ++z) {
    if (z > 0) x = rv * x + sv * gauss(c.seed, kPert, base + z);
    coef[((int64_t)e * n + 1 + (v - 1) * c.nz + z) * NC + q] = (float)(amp * x);
  }
}

// Legendre table Ptab[j * NL + l(l+1)/2 + m] per grid latitude.
__global__ void k_legendre_tab(int nlat, int LP, const float* glat, double* Ptab) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlat) return;
  legendre(LP, sin(glat[j] * kPi / 180.0), Ptab + (int64_t)j * ((LP + 1) * (LP + 2) / 2));
}

// Stage 1: per (e, row, lat j) the Fourier coefficients A_m, B_m.
__global__ void k_pert_lat(dida_synth_cfg c, int NC, const float* coef, const double* Ptab,
                           float* AB) {
  const int LP = c.lmax_pert;
  const int NL = (LP + 1) * (LP + 2) / 2;
  const int n = 1 + 3 * c.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)c.k * n * c.nlat;
  if (t >= total) return;
  const int j = (int)(t % c.nlat);
  const int64_t er = t / c.nlat;
  const double* P = Ptab + (int64_t)j * NL;
  const float* cf = coef + er * NC;
  float* out = AB + t * (2 * (LP + 1));
  for (int m = 0; m <= LP; ++m) {
    double a = 0.0, b = 0.0;
    for (int l = m; l <= LP; ++l) {
      // q index of (l, m, cos) = l^2 + 2m - 1 (m > 0) or l^2 (m = 0); sin = cos + 1
      const int qc = l * l + (m ? 2 * m - 1 : 0);
      a += P[l * (l + 1) / 2 + m] * cf[qc];
      if (m) b += P[l * (l + 1) / 2 + m] * cf[qc + 1];
    }
    const double fm = m ? 1.4142135623730951 : 1.0;
    out[2 * m] = (float)(a * fm);
    out[2 * m + 1] = (float)(b * fm);
  }
}

__device__ inline double row_sigma(const dida_synth_cfg& c, int row) {
  if (row == 0) return c.pert_ps;
  const int v = (row - 1) / c.nz;
  return v < 2 ? c.pert_uv : c.pert_T;
}

// Stage 2: per (row, col) all members; writes xbar_b and Xb'.
template <typename T>
__global__ void k_pert_state(dida_synth_cfg c, const float* AB, const float* glon,
                             const double* truth, T* xb, double* axb) {
  const int LP = c.lmax_pert;
  const int n = 1 + 3 * c.nz;
  const int64_t ncol = (int64_t)c.nlat * c.nlon;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * n) return;
  const int row = (int)(t % n);
  const int64_t col = t / n;
  const int j = (int)(col % c.nlat), i = (int)(col / c.nlat);
  const double lam = glon[i] * kPi / 180.0;
  float cs[2 * 65];
  for (int m = 0; m <= LP; ++m) {
    cs[2 * m] = (float)cos(m * lam);
    cs[2 * m + 1] = (float)sin(m * lam);
  }
  const double sig = row_sigma(c, row);
  const double tr = truth[col * n + row];
  const float* ab0 = AB + ((int64_t)row * c.nlat + j) * (2 * (LP + 1));
  const int64_t estride = (int64_t)n * c.nlat * (2 * (LP + 1));
  // two passes over the members (mean, then perturbations) keep no per-member array
  double mean = 0.0;
  for (int e = 0; e < c.k; ++e) {
    const float* ab = ab0 + e * estride;
    float s = 0.0f;
    for (int m = 0; m <= LP; ++m) s += ab[2 * m] * cs[2 * m] + ab[2 * m + 1] * cs[2 * m + 1];
    mean += tr + sig * (double)s;
  }
  mean /= c.k;
  axb[col * n + row] = mean;
  T* x = xb + (col * n + row) * c.k;
  for (int e = 0; e < c.k; ++e) {
    const float* ab = ab0 + e * estride;
    float s = 0.0f;
    for (int m = 0; m <= LP; ++m) s += ab[2 * m] * cs[2 * m] + ab[2 * m + 1] * cs[2 * m + 1];
    x[e] = (T)(tr + sig * (double)s - mean);
  }
}

__device__ inline double wrap360(double x) {
  double w = fmod(x, 360.0);
  if (w < 0.0) w += 360.0;
  if (w >= 360.0) w -= 360.0;
  return w;
}

__global__ void k_obs(dida_synth_cfg c, const double* cl_lat, const double* cl_lon,
                      const double* truth, const double* xb, const double* axb, double* lat,
                      double* lon, double* yb, double* dep, float* rinv) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= c.nobs) return;
  const int n = 1 + 3 * c.nz;
  const int nclu = (int)(c.cluster_frac * c.nobs);
  double la, lo;
  if (o >= c.nobs - nclu && c.nclusters > 0) {
    const int cl = o % c.nclusters;
    const double dx = gauss(c.seed, kObsPos, 2 * (uint64_t)o) * c.cluster_radius_km;
    const double dy = gauss(c.seed, kObsPos, 2 * (uint64_t)o + 1) * c.cluster_radius_km;
    la = cl_lat[cl] + dy / (kA * kPi / 180.0);
    la = fmax(-89.9, fmin(89.9, la));
    lo = wrap360(cl_lon[cl] + dx / (kA * kPi / 180.0 * fmax(0.05, cos(cl_lat[cl] * kPi / 180.0))));
  } else {
    la = asin(2.0 * unif(c.seed, kObsPos, 2 * (uint64_t)o) - 1.0) * 180.0 / kPi;
    lo = 360.0 * unif(c.seed, kObsPos, 2 * (uint64_t)o + 1);
  }
  lat[o] = la;
  lon[o] = lo;
  const int var = (int)(3.0 * unif(c.seed, kObsVar, o)) % 3;  // 0 u, 1 v, 2 T
  const double z = c.nz > 1 ? (c.nz - 1) * unif(c.seed, kObsLev, o) : 0.0;
  int z0 = (int)floor(z);
  if (z0 > c.nz - 2) z0 = c.nz > 1 ? c.nz - 2 : 0;
  const double tz = c.nz > 1 ? z - z0 : 0.0;
  const double dphi = 180.0 / c.nlat, dlam = 360.0 / c.nlon;
  double fj = (la + 90.0) / dphi - 0.5;
  int j0 = (int)floor(fj);
  if (j0 < 0) j0 = 0;
  if (j0 > c.nlat - 2) j0 = c.nlat - 2;
  double tj = fmin(1.0, fmax(0.0, fj - j0));
  double fi = wrap360(lo) / dlam;
  int i0 = (int)floor(fi);
  double ti = fi - i0;
  i0 %= c.nlon;
  const int i1 = (i0 + 1) % c.nlon;
  const int64_t cols[4] = {(int64_t)i0 * c.nlat + j0, (int64_t)i0 * c.nlat + j0 + 1,
                           (int64_t)i1 * c.nlat + j0, (int64_t)i1 * c.nlat + j0 + 1};
  const double wh[4] = {(1 - ti) * (1 - tj), (1 - ti) * tj, ti * (1 - tj), ti * tj};
  const int r0 = 1 + var * c.nz + z0;
  const int r1 = c.nz > 1 ? r0 + 1 : r0;
  double htruth = 0.0;
  for (int q = 0; q < 4; ++q)
    htruth += wh[q] * ((1 - tz) * truth[cols[q] * n + r0] + tz * truth[cols[q] * n + r1]);
  const double sig = var < 2 ? c.sig_uv : c.sig_T;
  const double lsc = log(c.rscale_min) +
                     (log(c.rscale_max) - log(c.rscale_min)) * unif(c.seed, kObsScale, o);
  const double scale = exp(lsc);
  const double sd = sig * sqrt(scale);
  const double yo = htruth + sd * gauss(c.seed, kObsNoise, o);
  const float sdf = (float)sd;
  rinv[o] = 1.0f / (sdf * sdf);
  // H(Xb) per member, then centring (letkf_obs.F90:156-178)
  double* y = yb + (int64_t)o * c.k;
  double sum = 0.0;
  for (int e = 0; e < c.k; ++e) {
    double h = 0.0;
    for (int q = 0; q < 4; ++q) {
      const int64_t b0 = cols[q] * n + r0, b1 = cols[q] * n + r1;
      const double x0 = axb[b0] + xb[b0 * c.k + e];
      const double x1 = axb[b1] + xb[b1 * c.k + e];
      h += wh[q] * ((1 - tz) * x0 + tz * x1);
    }
    y[e] = h;
    sum += h;
  }
  const double ybar = sum / c.k;
  dep[o] = yo - ybar;
  for (int e = 0; e < c.k; ++e) y[e] = y[e] - ybar;
}

}  // namespace

extern "C" int dida_synth_pert_ncoef(const dida_synth_cfg* c) {
  return (c->lmax_pert + 1) * (c->lmax_pert + 1);
}

extern "C" size_t dida_synth_scratch_bytes(const dida_synth_cfg* c) {
  const int64_t n = 1 + 3 * (int64_t)c->nz;
  const int64_t NC = dida_synth_pert_ncoef(c);
  const int64_t ncol = (int64_t)c->nlat * c->nlon;
  size_t b = 0;
  b += (size_t)c->k * n * NC * sizeof(float);                               // coef
  b += (size_t)c->k * n * c->nlat * 2 * (c->lmax_pert + 1) * sizeof(float); // AB
  b += (size_t)ncol * n * sizeof(double);                                    // truth
  b += 2 * 64 * sizeof(double);                                              // clusters
  b += (size_t)c->nlat * (c->lmax_pert + 1) * (c->lmax_pert + 2) / 2 * sizeof(double); // Ptab
  return b + 2048;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" int dida_synth_generate(const dida_synth_cfg* cp, void* scratch, float* glat,
                                   float* glon, double* xb, float* xb_f32, double* axb,
                                   double* lat, double* lon, double* yb, double* dep,
                                   float* rinv, void* stream) {
  const dida_synth_cfg c = *cp;
  hipStream_t s = (hipStream_t)stream;
  if (c.k < 2 || c.k > 128 || c.lmax_pert > 64 || c.nlat < 2 || c.nlon < 2 || c.nz < 1 ||
      c.nclusters > 64)
    return (int)hipErrorInvalidValue;
  const int n = 1 + 3 * c.nz;
  const int NC = dida_synth_pert_ncoef(&c);
  const int64_t ncol = (int64_t)c.nlat * c.nlon;
  char* p = (char*)scratch;
  float* coef = (float*)p;
  p += align_up((size_t)c.k * n * NC * sizeof(float));
  float* AB = (float*)p;
  p += align_up((size_t)c.k * n * c.nlat * 2 * (c.lmax_pert + 1) * sizeof(float));
  double* truth = (double*)p;
  p += align_up((size_t)ncol * n * sizeof(double));
  double* clus = (double*)p;
  p += align_up(128 * sizeof(double));
  double* Ptab = (double*)p;

  // cluster centres on the host (a few numbers), copied with the stream
  double hc[128];
  for (int q = 0; q < 64; ++q) {
    hc[q] = asin(2.0 * unif(c.seed, kCluster, 2 * (uint64_t)q) - 1.0) * 180.0 / kPi * 0.8;
    hc[64 + q] = 360.0 * unif(c.seed, kCluster, 2 * (uint64_t)q + 1);
  }
  hipError_t e = hipMemcpyAsync(clus, hc, sizeof hc, hipMemcpyHostToDevice, s);
  if (e) return (int)e;
  const int nmax = c.nlat > c.nlon ? c.nlat : c.nlon;
  k_grid<<<(nmax + 255) / 256, 256, 0, s>>>(c.nlat, c.nlon, glat, glon);
  k_truth<<<(unsigned)((ncol + 127) / 128), 128, 0, s>>>(c, glat, glon, truth);
  const int64_t ncoefthreads = (int64_t)c.k * 4 * NC;
  k_pert_coef<<<(unsigned)((ncoefthreads + 255) / 256), 256, 0, s>>>(c, NC, coef);
  const int64_t nlt = (int64_t)c.k * n * c.nlat;
  k_legendre_tab<<<(c.nlat + 63) / 64, 64, 0, s>>>(c.nlat, c.lmax_pert, glat, Ptab);
  k_pert_lat<<<(unsigned)((nlt + 127) / 128), 128, 0, s>>>(c, NC, coef, Ptab, AB);
  const int64_t nst = ncol * n;
  if (xb) k_pert_state<double><<<(unsigned)((nst + 127) / 128), 128, 0, s>>>(c, AB, glon, truth, xb, axb);
  if (!xb && xb_f32)
    k_pert_state<float><<<(unsigned)((nst + 127) / 128), 128, 0, s>>>(c, AB, glon, truth, xb_f32, axb);
  if (c.nobs > 0 && xb)
    k_obs<<<(c.nobs + 127) / 128, 128, 0, s>>>(c, clus, clus + 64, truth, xb, axb, lat, lon, yb,
                                               dep, rinv);
  return (int)hipGetLastError();
}
