// Device-side local observation selection (gen_yb_dis, letkf_yb.F90:3-133).
//
// Observations are binned on a regular lat/lon lattice (stable sort by bin
// id), then one wavefront per tile of grid columns visits the candidate bins
// in ascending bin id, decides per candidate whether the reference's
// Gaspari-Cohn weight is above zero (an FP32 unit-vector test, the thin shell
// it leaves open by the FP64 unit-vector cosine), takes the weight of kept
// ones, and compacts them with a ballot.  Three passes: count (capped at
// MAX_OBS), cap and fill into a CSR list.
//
// MAX_OBS truncation.  gen_yb_dis appends the kept obs in its scan order and
// stops at MAX_OBS (letkf_yb.F90:98-115, letkf.F90:155); over one global
// block that order is the caller's obs order, so a column with more than
// max_obs kept obs keeps the max_obs of smallest caller index.  The cap pass
// finds, for every column whose count reached max_obs, the caller index of
// its max_obs-th kept obs by a radix select (11-bit digits of perm, one
// histogram in LDS per wave, one scan of the candidates per digit); the fill
// pass then keeps the obs below that threshold, so what a capped column
// keeps does not depend on the bin size.
#include "letkf_device.h"
#include "letkf_kernels.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

using namespace dida;

namespace {

__device__ __forceinline__ double wrap360(double lon) {
  double w = fmod(lon, 360.0);
  if (w < 0.0) w += 360.0;
  if (w >= 360.0) w -= 360.0;
  return w;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_bin_keys(int32_t nobs, const double* __restrict__ lat,
                           const double* __restrict__ lon, dk_bin_geom g,
                           uint32_t* __restrict__ key, int32_t* __restrict__ iota) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nobs) return;
  double la = lat[i];
  uint32_t kk;
  if (!(fabs(la) <= g.obs_polar)) {
    kk = (uint32_t)g.nbins;
  } else {
    int b = (int)floor((la + 90.0) / g.dlat);
    b = b < 0 ? 0 : (b > g.nband - 1 ? g.nband - 1 : b);
    int c = (int)floor(wrap360(lon[i]) / g.dlon);
    c = c < 0 ? 0 : (c > g.nlonbin - 1 ? g.nlonbin - 1 : c);
    kk = (uint32_t)(b * g.nlonbin + c);
  }
  key[i] = kk;
  iota[i] = i;
}

// bin_start[b] = the number of obs whose key is below b, b = 0 .. nbins + 1
// (bin nbins holds the obs beyond obs_polar): a lower-bound search of the
// sorted keys, one thread a bin.
__global__ void k_bin_starts(int32_t nbins, int32_t nobs, const uint32_t* __restrict__ key_sorted,
                             int32_t* __restrict__ bin_start) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins + 1) return;
  int lo = 0, hi = nobs;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (key_sorted[mid] < (uint32_t)b) lo = mid + 1;
    else hi = mid;
  }
  bin_start[b] = lo;
}

__global__ void k_obs_gather(int32_t nobs, const int32_t* __restrict__ perm,
                             const double* __restrict__ lat, const double* __restrict__ lon,
                             const double* __restrict__ dep, const float* __restrict__ rinv,
                             double* __restrict__ cx, double* __restrict__ sx,
                             double* __restrict__ yr, float4* __restrict__ u4,
                             double* __restrict__ ux, double* __restrict__ uy,
                             double* __restrict__ uz, int32_t* __restrict__ iperm,
                             double* __restrict__ dep_b, float* __restrict__ rinv_b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nobs) return;
  int s = perm[i];
  iperm[s] = i;
  double x1r = deg2rad_ref(lat[s]);
  const double c1 = cos(x1r), s1 = sin(x1r), y1 = deg2rad_ref(lon[s]);
  cx[i] = c1;
  sx[i] = s1;
  yr[i] = y1;
  const double vx = c1 * cos(y1), vy = c1 * sin(y1);
  u4[i] = make_float4((float)vx, (float)vy, (float)s1, 0.0f);
  ux[i] = vx;
  uy[i] = vy;
  uz[i] = s1;
  dep_b[i] = dep[s];
  rinv_b[i] = rinv[s];
}

// Yb rows in bin order (row stride ldy, zero past k), in FP64 or rounded to
// FP32 (the half-precision Gram of the chained form splits FP32 values): a
// wavefront a row, four rows a workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_obs_gather_yb(int32_t nobs, int32_t k, int32_t ldy,
                                                       const int32_t* __restrict__ perm,
                                                       const double* __restrict__ yb,
                                                       T* __restrict__ yb_b) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nobs) return;
  const int lane = threadIdx.x & 63;
  const double* __restrict__ src = yb + (int64_t)perm[i] * k;
  T* __restrict__ dst = yb_b + i * ldy;
  for (int e = lane; e < ldy; e += kWave) dst[e] = e < k ? (T)src[e] : (T)0;
}

template <typename T>
__global__ void k_obs_ymax(int32_t nobs, int32_t k, int32_t ldy, const T* __restrict__ yb,
                           const float* __restrict__ rinv, double* __restrict__ ymax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nobs) return;
  double m = 0.0;
  for (int e = 0; e < k; ++e) m = fmax(m, fabs((double)yb[(int64_t)i * ldy + e]));
  ymax[i] = m * sqrt((double)rinv[i]);
}

// Candidate region of analysis columns: bands b0..b1 and, unless `full`,
// the longitudes [lo_deg, hi_deg] (degrees, unwrapped: lo_deg may be below 0
// and hi_deg above 360).  For one column: the spherical cap of angular radius
// theta_deg around it, widened so that no kept obs falls outside the visited
// bins (the reference's distance uses a float-rounded pi).
struct CandRegion {
  int b0, b1;
  bool full;
  double lo_deg, hi_deg;
};

__device__ __forceinline__ CandRegion column_region(const dk_select_args& a, double clat,
                                                    double cw) {
  const double kPi = 3.14159265358979323846;
  const double th = a.theta_deg;
  const dk_bin_geom& g = a.geom;
  CandRegion r;
  r.b0 = (int)floor((clat - th + 90.0) / g.dlat);
  r.b1 = (int)floor((clat + th + 90.0) / g.dlat);
  r.b0 = r.b0 < 0 ? 0 : r.b0;
  r.b1 = r.b1 > g.nband - 1 ? g.nband - 1 : r.b1;
  r.full = false;
  double dlon = 0.0;
  if (fabs(clat) + th >= 90.0) {
    r.full = true;
  } else {
    const double sv = sin(th * kPi / 180.0) / cos(clat * kPi / 180.0);
    if (sv >= 1.0) r.full = true;
    else dlon = asin(sv) * 180.0 / kPi * (1.0 + 1e-6) + 1e-6;
  }
  r.lo_deg = cw - dlon;
  r.hi_deg = cw + dlon;
  return r;
}

// Visit the candidate bins of a region; call f(l0, l1) per contiguous obs
// range in ascending bin order (so every column's kept obs come in the order
// of a full scan of the binned list); f returns false to stop.
template <class F>
__device__ __forceinline__ void for_region_ranges(const dk_select_args& a, const CandRegion& r,
                                                  F&& f) {
  const dk_bin_geom& g = a.geom;
  bool full = r.full;
  int lo = 0, hi = g.nlonbin - 1;
  if (!full) {
    lo = (int)floor(r.lo_deg / g.dlon);
    hi = (int)floor(r.hi_deg / g.dlon);
    if (hi - lo + 1 >= g.nlonbin) full = true;
  }
  const int32_t* bs = a.obs.bin_start;
  for (int b = r.b0; b <= r.b1; ++b) {
    int base = b * g.nlonbin;
    int r0a, r0b, r1a = -1, r1b = -2;
    if (full) {
      r0a = base; r0b = base + g.nlonbin - 1;
    } else if (lo < 0) {
      r0a = base; r0b = base + hi;
      r1a = base + lo + g.nlonbin; r1b = base + g.nlonbin - 1;
    } else if (hi > g.nlonbin - 1) {
      r0a = base; r0b = base + hi - g.nlonbin;
      r1a = base + lo; r1b = base + g.nlonbin - 1;
    } else {
      r0a = base + lo; r0b = base + hi;
    }
    if (!f(bs[r0a], bs[r0b + 1])) return;
    if (r1a <= r1b && !f(bs[r1a], bs[r1b + 1])) return;
  }
}

// The column's geometry for the candidate test.
struct ColGeo {
  double cx2, sx2, y2r;  // cos / sin of the latitude, longitude (reference radians)
  double ux, uy, uz;     // unit vector
  float vx, vy, vz;      // unit vector, FP32
};

__device__ __forceinline__ ColGeo col_geo(double clat, double clon) {
  ColGeo q;
  const double x2r = deg2rad_ref(clat);
  q.cx2 = cos(x2r);
  q.sx2 = sin(x2r);
  q.y2r = deg2rad_ref(clon);
  q.ux = q.cx2 * cos(q.y2r);
  q.uy = q.cx2 * sin(q.y2r);
  q.uz = q.sx2;
  q.vx = (float)q.ux;
  q.vy = (float)q.uy;
  q.vz = (float)q.uz;
  return q;
}

// Margins of the FP32 unit-vector test: its dot product differs from the
// exact cosine by less than 4e-7 (FP32 rounding of both vectors, the
// products and the sums), so outside [cos_far - 1e-6, cos_near + 1e-6] it
// decides a candidate as the reference formula would (candidates surely
// beyond the support are rejected, those surely inside it have gc > 0); the
// thin shell between is decided by shell_keep.
__device__ __forceinline__ float far_lo32(const dk_select_args& a) {
  return (float)(a.cos_far - 1e-6);
}
__device__ __forceinline__ float near_hi32(const dk_select_args& a) {
  return (float)(a.cos_near + 1e-6);
}
__device__ __forceinline__ float dot32(const float4& u, float vx, float vy, float vz) {
  return fmaf(u.z, vz, fmaf(u.y, vy, u.x * vx));
}

// The reference's weight of binned obs l (letkf_yb.F90:84-96): gc of the
// distance in the reference's formula (0 beyond the support).
__device__ __forceinline__ void ref_weight(const dk_select_args& a, const ColGeo& q, int l,
                                           double& gc) {
#pragma clang fp contract(off)
  gc = 0.0;
  const double v = central_cos(a.obs.cx[l], a.obs.sx[l], a.obs.yr[l], q.cx2, q.sx2, q.y2r);
  if (!(v >= a.cos_far)) return;
  const double d = dist_from_cos(v);
  const double r = d * 1.0 / a.local_dist;
  gc = gc_weight(r, a.gc_filter);
}

// Central angle (rad) between unit vectors p and q from their difference d
// = p - q: |d| = 2 sin(theta / 2), so theta = 2 asin(|d| / 2), summed as the
// Maclaurin series of asin, x sum_k (2k)! / (4^k (k!)^2 (2k + 1)) x^(2k).
// kTerms = 13 for theta <= 0.4 (x^2 < 0.04), 8 for theta <= 0.197 (x^2 <
// 0.0097): the first term left out is below 1e-18 of the sum either way.
// Unlike acos(p . q) it keeps its digits at short distances.  |d| is
// c2 / sqrt(c2) from the Newton-refined reciprocal square root (a few ulp).
constexpr double kChordMaxDeg = 22.9;       // 0.4 rad
constexpr double kChordShortMaxDeg = 11.3;  // 0.197 rad
// a * b + c as the three-operand FP64 FMA with c in scalar registers: left to
// itself the compiler takes the two-operand form, whose accumulator is the
// addend, and copies each loop-invariant series coefficient into it first.
__device__ __forceinline__ double fma3(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}
template <int kTerms>
__device__ __forceinline__ double chord_angle(double dx, double dy, double dz) {
  static_assert(kTerms == 13 || kTerms == 8, "13 or 8 terms");
  const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double x2 = 0.25 * c2;
  double s;
  if constexpr (kTerms == 13) {
    s = fma(0.006447210311889649,          // 676039 / 104857600
            x2, 0.0073125258735988454);    // 88179 / 12058624
    s = fma3(s, x2, 0.008390335809616815);   // 46189 / 5505024
    s = fma3(s, x2, 0.009761609529194078);   // 12155 / 1245184
    s = fma3(s, x2, 0.011551800896139705);   // 6435 / 557056
    s = fma3(s, x2, 0.01396484375);          // 143 / 10240
    s = fma3(s, x2, 0.017352764423076924);   // 231 / 13312
  } else {
    s = fma(0.01396484375,                 // 143 / 10240
            x2, 0.017352764423076924);     // 231 / 13312
  }
  s = fma3(s, x2, 0.022372159090909092);     // 63 / 2816
  s = fma3(s, x2, 0.030381944444444444);     // 35 / 1152
  s = fma3(s, x2, 0.044642857142857144);     // 5 / 112
  s = fma3(s, x2, 0.075);                    // 3 / 40
  s = fma3(s, x2, 0.16666666666666666);      // 1 / 6
  s = fma(s, x2, 1.0);
  const double c = c2 > 0.0 ? c2 * rsq_nr(c2) : 0.0;  // |d|
  return c * s;                              // 2 x s
}

// gc_weight for a candidate surely inside the support (r <= 2): the same
// polynomials, with r / 12 as a product and 0.6666667 / r from the
// Newton-refined reciprocal (agrees with gc_weight to a few ulp).
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(fma(-x, y, 1.0), y, y);
  y = fma(fma(-x, y, 1.0), y, y);
  return y;
}
__device__ __forceinline__ double gc_weight_inner(double ri, int gc_filter) {
  if (!gc_filter || ri > 2.0) return ri > 2.0 ? 0.0 : 1.0;
  if (ri > 1.0) {
    double rj = fma(ri, 1.0 / 12.0, -0.5);
    rj = fma(rj, ri, 0.625);
    rj = fma(rj, ri, kGc53);
    return fma(fma(rj, ri, -5.0), ri, 4.0) - kGc23 * rcp_nr(ri);
  }
  if (ri > 0.0) {
    const double rk = ri * ri;
    double rj = fma(fma(-0.25, ri, 0.5), ri, 0.625);
    return fma(fma(rj, ri, -kGc53), rk, 1.0);
  }
  return 1.0;
}

// The test of a candidate in the shell that the FP32 test leaves open: the
// central-angle cosine from the FP64 unit vectors against cos(2 L / a).  In
// the reference's radians the dot product of the unit vectors is the
// reference's cos(x1) cos(x2) cos(y1 - y2) + sin(x1) sin(x2) up to rounding
// (a few 1e-16), and its gc > 0 holds exactly when r = d / L <= 2 (gc(2) is
// about 1.4e-6 with its REAL(4) constants), so the two tests can differ only
// for an obs within about 1e-15 of the support edge in cosine.
__device__ __forceinline__ bool shell_keep(const dk_select_args& a, double qx, double qy,
                                           double qz, int l) {
  const double v = fma(a.obs.ux[l], qx, fma(a.obs.uy[l], qy, a.obs.uz[l] * qz));
  return v >= a.cos_keep;
}

// Whether binned obs l is kept (gc > 0), without its weight.
__device__ __forceinline__ bool keep_candidate(const dk_select_args& a, const ColGeo& q, int l,
                                               const float4& u) {
  const float d = dot32(u, q.vx, q.vy, q.vz);
  if (d < far_lo32(a)) return false;
  if (d > near_hi32(a)) return true;
  return shell_keep(a, q.ux, q.uy, q.uz, l);
}

// Tiles of TJ x TI analysis columns (latitude-fastest within the tile), one
// wavefront a tile, four tiles a workgroup; TJ x TI <= kNT.
constexpr int kNT = 16;

// Geometry and output offset of a tile's columns, read by the weight and
// exact-test code (once per buffered wavefront of candidates).
struct TileCols {
  double cx2[kNT], sx2[kNT], y2r[kNT], ux[kNT], uy[kNT], uz[kNT];
  int64_t out[kNT];
  float4 v[kNT];  // FP32 unit vector (read by the scan as one broadcast)
};

// Fill pass: per column, the candidates that passed the FP32 test, in scan
// order, waiting for a full wavefront of weights and shell tests (up to 63
// waiting plus the 64 of one chunk): l, or ~l for a candidate in the shell
// that shell_keep decides.
template <int NTT>
struct TileBuf {
  int32_t l[NTT][128];
};

__device__ __forceinline__ double wave_min(double v) {
  for (int s = 32; s >= 1; s >>= 1) v = fmin(v, __shfl_xor(v, s, kWave));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int s = 32; s >= 1; s >>= 1) v = fmax(v, __shfl_xor(v, s, kWave));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int s = 32; s >= 1; s >>= 1) v = min(v, __shfl_xor(v, s, kWave));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int s = 32; s >= 1; s >>= 1) v = max(v, __shfl_xor(v, s, kWave));
  return v;
}

// Per-column state held lane-wise (lane j holds column j's): the value of
// lane j (j wave-uniform) as a scalar, and v with lane j set to x.
__device__ __forceinline__ int lane_get(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float lane_get(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ int lane_set(int v, int j, int x) {
  return (int)(threadIdx.x & 63) == j ? x : v;
}
// Ballot and any-lane of a condition as a lane mask, without widening it to
// an int first (HIP's __any and __ballot take an int).
__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ __forceinline__ bool any_lane(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }
// A wave-uniform mask in scalar registers: the compiler cannot always prove
// a ballot-derived value uniform after divergent stores, and would run the
// loops over it with exec masks and vector bit scans.
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
using u32x2 = unsigned int __attribute__((__vector_size__(2 * sizeof(unsigned int))));
// Raw buffer over p (32-bit byte offsets, no range clamp below 4 GiB).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0xFFFFFFFF, 0x00020000);
}
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __longlong_as_double((long long)__builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// Count (kFill false: nobsl[col] = min(kept, max_obs)) or fill (kept obs with
// caller index below thr[col], in bin order, at offset[col] - base) of the
// columns of a tile: the union of the columns' candidate regions is scanned
// once, 64 candidates at a time, each chunk tested against every column of
// the tile.  Fill: columns outside [col0, col0 + ncols) are skipped.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kTerms (fill): weights from the chord angle, its series of 13 or 8 terms
// by the support (kChordMaxDeg, kChordShortMaxDeg), agreeing with the
// reference formula to about 1e-13 (the rounding of its acos near 1); 0:
// from the reference's formula, evaluated for every kept obs.
template <bool kFill, int TJ, int TI, int kTerms>
__global__ __launch_bounds__(256) void k_select_tiles(dk_select_args a) {
  constexpr int kTJ = TJ, kTI = TI, kNTT = TJ * TI;
  static_assert(kNTT <= kNT, "tile larger than its column state");
  __shared__ TileCols tc_all[4];
  __shared__ TileBuf<kFill ? kNTT : 1> tb_all[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  TileCols& tc = tc_all[wave];
  auto& tb = tb_all[wave];
  const int ntj = (a.nlat + kTJ - 1) / kTJ;
  const bool run = kFill && a.ncols > 0;
  const int ti0 = run ? (int)(a.col0 / a.nlat) / kTI : 0;
  int64_t tile = (int64_t)ti0 * ntj + (int64_t)blockIdx.x * 4 + wave;
  if (a.tile_order && !run) {
    const int64_t ntiles = (int64_t)ntj * ((a.nlon + kTI - 1) / kTI);
    if (tile >= ntiles) return;
    tile = a.tile_order[tile];
  }
  const int tj = (int)(tile % ntj), ti = (int)(tile / ntj);
  const int64_t c_lo = kFill && a.ncols > 0 ? a.col0 : 0;
  const int64_t c_hi = kFill && a.ncols > 0 ? a.col0 + a.ncols : (int64_t)a.nlat * a.nlon;
  // lane j < kNTT: column j of the tile
  const int jj = lane % kTJ, ii = lane / kTJ;
  const int ilat = tj * kTJ + jj, ilon = ti * kTI + ii;
  const int64_t col = (int64_t)ilat + (int64_t)a.nlat * ilon;
  const bool mine = lane < kNTT && ilat < a.nlat && ilon < a.nlon && col >= c_lo && col < c_hi;
  if (!any_lane(mine)) return;
  double clat = 0.0, clon = 0.0;
  if (mine) {
    clat = (double)a.grid_lat[ilat];
    clon = (double)a.grid_lon[ilon];
  }
  // the union of the columns' regions, longitudes unwrapped around the first
  // column's
  const int src = __ffsll((unsigned long long)ballot(mine)) - 1;
  const double ref = wrap360(__shfl(clon, src, kWave));
  double cw = wrap360(clon);
  if (cw - ref > 180.0) cw -= 360.0;
  if (cw - ref < -180.0) cw += 360.0;
  const CandRegion rc = column_region(a, clat, cw);
  CandRegion reg;
  reg.b0 = wave_min_i(mine ? rc.b0 : 1 << 30);
  reg.b1 = wave_max_i(mine ? rc.b1 : -1);
  reg.full = any_lane(mine && rc.full);
  reg.lo_deg = wave_min(mine ? rc.lo_deg : 1e300);
  reg.hi_deg = wave_max(mine ? rc.hi_deg : -1e300);
  const ColGeo q = col_geo(clat, clon);
  if (lane < kNTT) {
    tc.cx2[lane] = q.cx2;
    tc.sx2[lane] = q.sx2;
    tc.y2r[lane] = q.y2r;
    tc.ux[lane] = q.ux;
    tc.uy[lane] = q.uy;
    tc.uz[lane] = q.uz;
    tc.out[lane] = kFill && mine ? a.offset[col] - a.base : 0;
    tc.v[lane] = make_float4(q.vx, q.vy, q.vz, 0.0f);
    if (kFill && mine && a.offset[col] - a.base + a.nobsl[col] > a.capacity) *a.overflow = 1;
  }
  // lane-wise column state: how many obs the column takes (0 for lanes that
  // are not columns of this run), caller-index threshold, kept so far,
  // candidates waiting in the buffer
  const int climit = mine ? (kFill ? a.nobsl[col] : a.max_obs) : 0;
  const int cthr = kFill && mine ? a.thr[col] : INT32_MAX;
  int ccnt = 0, cnb = 0;
  wave_sync_lds();
  const bool capped = any_lane(cthr != INT32_MAX);
  // columns still counting / filling (bit j)
  uint32_t open = uniform_u32((uint32_t)ballot(climit > 0));
  const float flo = far_lo32(a), nhi = near_hi32(a);
  const double r_per_rad = kEarthA / a.local_dist;
  const __amdgpu_buffer_rsrc_t rux = buf_rsrc(a.obs.ux), ruy = buf_rsrc(a.obs.uy),
                               ruz = buf_rsrc(a.obs.uz);
  // The test of the lanes' shell candidates l against column j.
  auto shell_j = [&](int j, int l, bool shell) -> bool {
    return shell && shell_keep(a, tc.ux[j], tc.uy[j], tc.uz[j], l);
  };
  // Fill: the weight (and, in the shell, the FP64 test) of the first n
  // buffered candidates of column j, one per lane; kept ones are written in
  // buffer (scan) order.
  auto flush = [&](int j, int n) {
    const bool live = lane < n;
    const int lb = live ? tb.l[j][lane] : 0;
    const int l = lb < 0 ? ~lb : lb;  // lanes past n: obs 0, never stored
    const uint64_t mlive = n >= kWave ? ~0ull : (1ull << n) - 1;
    const uint64_t mshell = ballot(lb < 0) & mlive;
    // kept: the candidates surely inside, and those of the shell that pass
    uint64_t bm = mlive & ~mshell;
    bool keep = live & (lb >= 0);
    double gc;
    if constexpr (kTerms != 0) {
      // each candidate's FP64 unit vector serves the shell test and the
      // weight (binned index < nobs < 2^29: 32-bit byte offsets)
      const uint32_t off = (uint32_t)l << 3;
      const double px = buf_f64(rux, off), py = buf_f64(ruy, off), pz = buf_f64(ruz, off);
      const double qx = tc.ux[j], qy = tc.uy[j], qz = tc.uz[j];
      if (mshell) {
        // shell_keep's cosine
        const bool sk = live & (lb < 0) & (fma(px, qx, fma(py, qy, pz * qz)) >= a.cos_keep);
        bm |= ballot(sk);
        keep = keep | sk;
      }
      gc = gc_weight_inner(chord_angle<kTerms>(px - qx, py - qy, pz - qz) * r_per_rad,
                           a.gc_filter);
    } else {
      if (mshell) {
        const bool sk = live & (lb < 0) && shell_keep(a, tc.ux[j], tc.uy[j], tc.uz[j], l);
        bm |= ballot(sk);
        keep = keep | sk;
      }
      ColGeo g;
      g.cx2 = tc.cx2[j];
      g.sx2 = tc.sx2[j];
      g.y2r = tc.y2r[j];
      ref_weight(a, g, l, gc);
    }
    const int c0 = lane_get(ccnt, j), lim = lane_get(climit, j);
    // the column's run of the lists, entries past the workspace dropped
    // (the overflow flag is set); 32-bit offsets from a scalar base
    const int64_t oj = uniform_i64(tc.out[j]);
    const int64_t room = a.capacity - oj;
    const int lim_w = room < (int64_t)lim ? (room > 0 ? (int)room : 0) : lim;
    if (keep) {
      const int pos = c0 + (int)__builtin_amdgcn_mbcnt_hi(
                               (uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
      if (pos < lim_w) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)l, buf_rsrc(a.csr_idx + oj),
                                              (uint32_t)pos * 4u, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, gc),
                                              buf_rsrc(a.csr_w + oj), (uint32_t)pos * 8u, 0, 0);
      }
    }
    const int c1 = c0 + __popcll(bm);
    ccnt = lane_set(ccnt, j, c1);
    if (c1 >= lim) open &= ~(1u << j);
  };
  if (open) {
    const float4* __restrict__ u4 = reinterpret_cast<const float4*>(a.obs.u4);
    for_region_ranges(a, reg, [&](int l0, int l1) -> bool {
      // (ranges of empty bins come too: nothing to load)
      if (l0 >= l1) return true;
      // each chunk's unit vectors (and caller indices) are loaded while the
      // one before is tested; indices clamped into the range need no branch
      float4 u_next = u4[l0 + lane < l1 ? l0 + lane : l1 - 1];
      int32_t pv_next = kFill && capped ? a.perm[l0 + lane < l1 ? l0 + lane : l1 - 1] : 0;
      for (int base = l0; base < l1; base += kWave) {
        const int l = base + lane;
        const bool valid = l < l1;
        const float4 u = u_next;
        const int32_t pv = pv_next;
        if (base + kWave < l1) {
          const int ln = l + kWave < l1 ? l + kWave : l1 - 1;
          u_next = u4[ln];
          if (kFill && capped) pv_next = a.perm[ln];
        }
        // the FP32 test's bounds, out of reach for lanes past the range, so
        // that each test is one compare
        const float flo_l = valid ? flo : 3.0f, nhi_l = valid ? nhi : 3.0f;
        for (uint32_t m = uniform_u32(open); m; m &= m - 1) {
          const int j = __builtin_ctz(m);
          const float4 cv = tc.v[j];
          const float d = dot32(u, cv.x, cv.y, cv.z);
          if constexpr (kFill) {
            // a capped column's caller-index threshold (INT32_MAX, above
            // every caller index, for the others), then the FP32 test;
            // survivors wait in the column's buffer
            uint64_t bm = ballot(d >= flo_l);
            if (capped) bm &= ballot(pv < lane_get(cthr, j));
            if (!bm) continue;
            const bool cand = (d >= flo_l) & (!capped | (pv < lane_get(cthr, j)));
            const int nb0 = lane_get(cnb, j);
            if (cand)
              tb.l[j][nb0 + (int)__builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(bm >> 32),
                                __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] = d > nhi ? l : ~l;
            int nb = nb0 + __popcll(bm);
            if (nb >= kWave) {
              wave_sync_lds();
              flush(j, kWave);
              // the rest moves to the front of the buffer
              const int rest = nb - kWave;
              const int v = lane < rest ? tb.l[j][kWave + lane] : 0;
              wave_sync_lds();
              if (lane < rest) tb.l[j][lane] = v;
              nb = rest;
            }
            cnb = lane_set(cnb, j, nb);
          } else {
            const bool cand = d >= flo_l;
            if (!any_lane(cand)) continue;
            bool keep = d > nhi_l;
            const bool shell = cand & !keep;
            if (any_lane(shell)) keep = keep || shell_j(j, l, shell);
            const int c1 = lane_get(ccnt, j) + __popcll(ballot(keep));
            ccnt = lane_set(ccnt, j, c1);
            if (c1 >= lane_get(climit, j)) open &= ~(1u << j);
          }
        }
        if (!open) return false;
      }
      return true;
    });
    if constexpr (kFill) {
      // the candidates still waiting
      wave_sync_lds();
      for (uint32_t m = uniform_u32(open); m; m &= m - 1) {
        const int j = __builtin_ctz(m);
        const int nb = lane_get(cnb, j);
        if (nb > 0) flush(j, nb);
      }
    }
  }
  if (!kFill && mine) a.nobsl[col] = ccnt < a.max_obs ? ccnt : a.max_obs;
}

// Sort key of every tile of TJ x TI columns: the obs count of the bins its
// wavefront scans (the union of its columns' candidate regions, formed as
// the tile kernel forms it), largest first.
template <int TJ, int TI>
__global__ void k_tile_cost(dk_select_args a, int32_t ntiles, uint32_t* __restrict__ key,
                            int32_t* __restrict__ id) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int ntj = (a.nlat + TJ - 1) / TJ;
  const int tj = t % ntj, ti = t / ntj;
  CandRegion reg;
  reg.b0 = 1 << 30;
  reg.b1 = -1;
  reg.full = false;
  reg.lo_deg = 1e300;
  reg.hi_deg = -1e300;
  double ref = 0.0;
  bool first = true;
  for (int ii = 0; ii < TI; ++ii)
    for (int jj = 0; jj < TJ; ++jj) {
      const int ilat = tj * TJ + jj, ilon = ti * TI + ii;
      if (ilat >= a.nlat || ilon >= a.nlon) continue;
      const double clat = (double)a.grid_lat[ilat];
      double cw = wrap360((double)a.grid_lon[ilon]);
      if (first) ref = cw;
      first = false;
      if (cw - ref > 180.0) cw -= 360.0;
      if (cw - ref < -180.0) cw += 360.0;
      const CandRegion rc = column_region(a, clat, cw);
      reg.b0 = min(reg.b0, rc.b0);
      reg.b1 = max(reg.b1, rc.b1);
      reg.full = reg.full || rc.full;
      reg.lo_deg = fmin(reg.lo_deg, rc.lo_deg);
      reg.hi_deg = fmax(reg.hi_deg, rc.hi_deg);
    }
  int64_t n = 0;
  if (!first)
    for_region_ranges(a, reg, [&](int l0, int l1) -> bool {
      n += l1 - l0;
      return true;
    });
  key[t] = 0xFFFFFFFFu - (uint32_t)(n < 0xFFFFFFFE ? n : 0xFFFFFFFE);
  id[t] = t;
}

// LDS writes of this wave visible to its later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kCapDigit = 11;               // bits of caller index per radix level
constexpr int kCapBins = 1 << kCapDigit;
constexpr int kCapEnumBits = 12;            // digit ranges of up to 2^12 caller indices
                                            // are enumerated instead of rescanned
constexpr int kCapUnroll = 4;               // candidate chunks in flight

// thr[col] of every column: INT32_MAX unless the count pass stopped at
// max_obs; then, if more than max_obs obs are kept, one past the caller index
// of the max_obs-th smallest kept one.  Level by level, the kept obs whose
// caller index lies in the open range [lo, lo + 2^width) are counted per
// 11-bit digit, and the digit holding rank `rank` narrows the range; once
// the range holds at most 2^12 caller indices they are tested one by one in
// caller order (the inverse permutation gives their binned positions; obs in
// the polar bin are never candidates), which replaces a rescan of the
// column's candidates.
__global__ __launch_bounds__(256) void k_select_cap(dk_select_args a) {
  // one workgroup a column: the four wavefronts share the candidate scan
  // (chunks dealt round-robin) and one histogram; wavefront 0 picks the
  // digit and enumerates
  __shared__ uint32_t hist[kCapBins];
  __shared__ int sel[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t col = blockIdx.x;
  if (a.max_obs <= 0 || a.nobsl[col] < a.max_obs) {
    if (threadIdx.x == 0) a.thr[col] = INT32_MAX;
    return;
  }
  const int ilat = (int)(col % a.nlat), ilon = (int)(col / a.nlat);
  const double clat = (double)a.grid_lat[ilat];
  const double clon = (double)a.grid_lon[ilon];
  const ColGeo q = col_geo(clat, clon);
  const CandRegion reg = column_region(a, clat, wrap360(clon));
  const float flo = far_lo32(a), nhi = near_hi32(a);
  int width = a.perm_bits;  // caller indices < 2^perm_bits
  uint32_t lo = 0;
  int rank = a.max_obs;
  int32_t thr = INT32_MAX;
  for (int level = 0;; ++level) {
    const int dbits = width < kCapDigit ? width : kCapDigit;
    const int shift = width - dbits;
    const int nb = 1 << dbits;
    for (int i = (int)threadIdx.x; i < nb; i += 256) hist[i] = 0u;
    __syncthreads();
    const uint32_t span = width >= 31 ? 0x80000000u : (1u << width);
    for_region_ranges(a, reg, [&](int l0, int l1) -> bool {
      // kCapUnroll chunks a step, their loads issued together
      for (int base = l0 + wave * kCapUnroll * kWave; base < l1;
           base += 4 * kCapUnroll * kWave) {
        // indices clamped into the range, so that the loads need no branch
        // and all issue before the first use
        int32_t pv[kCapUnroll];
        float4 u[kCapUnroll];
#pragma unroll
        for (int k = 0; k < kCapUnroll; ++k) {
          const int l = base + k * kWave + lane;
          const int lc = l < l1 ? l : l1 - 1;
          pv[k] = a.perm[lc];
          u[k] = reinterpret_cast<const float4*>(a.obs.u4)[lc];
        }
#pragma unroll
        for (int k = 0; k < kCapUnroll; ++k) {
          const int l = base + k * kWave + lane;
          // keep_candidate's tests as single compares; lanes past the range
          // or outside the caller-index range have rel = span
          const uint32_t rel = l < l1 ? (uint32_t)pv[k] - lo : span;
          const float d = dot32(u[k], q.vx, q.vy, q.vz);
          const bool in = rel < span;
          bool keep = in & (d > nhi);
          if (ballot(in) & ballot(d >= flo) & ~ballot(d > nhi))
            keep = keep || ((in & (d >= flo) & !(d > nhi)) && shell_keep(a, q.ux, q.uy, q.uz, l));
          if (keep) atomicAdd(&hist[rel >> shift], 1u);
        }
      }
      return true;
    });
    __syncthreads();
    // the digit whose cumulative count reaches rank: bins split over lanes,
    // a wave scan of the lane sums, then the owning lane walks its bins
    if (wave == 0) {
      const int per = (nb + kWave - 1) / kWave;
      const int b0 = lane * per;
      int sum = 0;
      for (int i = 0; i < per; ++i)
        if (b0 + i < nb) sum += (int)hist[b0 + i];
      int inc = sum;
      for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += t;
      }
      const int exc = inc - sum;
      // the level's kept obs; at the first level all of them, and a column
      // whose kept obs all fit needs no threshold
      if (lane == kWave - 1) sel[3] = inc;
      const uint64_t m = ballot(exc < rank && rank <= inc);
      const int src = m ? __ffsll((unsigned long long)m) - 1 : 0;
      if (lane == src) {
        int cum = exc, dsel = 0, before = 0;
        for (int i = 0; i < per; ++i) {
          const int h = b0 + i < nb ? (int)hist[b0 + i] : 0;
          if (cum + h >= rank) {
            dsel = b0 + i;
            before = cum;
            break;
          }
          cum += h;
        }
        sel[0] = dsel;
        sel[1] = before;
        sel[2] = m != 0;  // cannot fail with consistent passes
      }
    }
    __syncthreads();
    if (level == 0 && sel[3] <= a.max_obs) break;  // every kept obs fits
    if (!sel[2]) break;
    lo += (uint32_t)sel[0] << shift;
    rank -= sel[1];
    width = shift;
    if (width == 0) {
      thr = (int32_t)(lo + 1u);
      break;
    }
    if (a.iperm && width <= kCapEnumBits) {
      if (wave == 0) {
        const uint32_t n = 1u << width;
        const int32_t polar0 = a.obs.bin_start[a.geom.nbins];
        int cum = 0;  // kept obs of the range before this chunk
        for (uint32_t base = 0; base < n; base += kWave) {
          const uint32_t ci = lo + base + (uint32_t)lane;
          bool keep = false;
          if (base + (uint32_t)lane < n && ci < (uint32_t)a.nobs) {
            const int l = a.iperm[ci];
            keep = l < polar0 &&
                   keep_candidate(a, q, l, reinterpret_cast<const float4*>(a.obs.u4)[l]);
          }
          const uint64_t km = ballot(keep);
          const int cnt = __popcll(km);
          if (cum + cnt >= rank) {
            // the lane holding the (rank - cum)-th kept obs of the chunk
            const int below = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0u));
            const uint64_t hit = ballot(keep && below == rank - cum - 1);
            thr = (int32_t)(lo + base + (uint32_t)(__ffsll((unsigned long long)hit) - 1) + 1u);
            break;
          }
          cum += cnt;
        }
      }
      break;
    }
    __syncthreads();  // sel and the bins are read before the next level clears them
  }
  if (threadIdx.x == 0) a.thr[col] = thr;
}

}  // namespace

extern "C" int dk_fill_i32(int32_t* p, int32_t v, int64_t n, void* stream) {
  if (n <= 0) return 0;
  k_fill_i32<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(p, v, n);
  return (int)hipGetLastError();
}

extern "C" int dk_bin_keys(int32_t nobs, const double* lat, const double* lon, dk_bin_geom g,
                           uint32_t* key, int32_t* iota, void* stream) {
  if (nobs <= 0) return 0;
  k_bin_keys<<<(nobs + 255) / 256, 256, 0, (hipStream_t)stream>>>(nobs, lat, lon, g, key, iota);
  return (int)hipGetLastError();
}

extern "C" int dk_sort_pairs_u32(void* tmp, size_t* tmp_bytes, const uint32_t* kin,
                                 uint32_t* kout, const int32_t* vin, int32_t* vout, int32_t n,
                                 int end_bit, void* stream) {
  return (int)rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0,
                                        end_bit, (hipStream_t)stream, false);
}

extern "C" int dk_bin_starts(int32_t nbins, int32_t nobs, const uint32_t* key_sorted,
                             int32_t* bin_start, void* stream) {
  if (nbins < 0 || nobs < 0) return (int)hipErrorInvalidValue;
  k_bin_starts<<<(unsigned)((nbins + 2 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      nbins, nobs, key_sorted, bin_start);
  return (int)hipGetLastError();
}

extern "C" int dk_obs_gather(int32_t nobs, int32_t k, int32_t ldy, const int32_t* perm,
                             const double* lat,
                             const double* lon, const double* yb, const double* dep,
                             const float* rinv, double* cx, double* sx, double* yr, float* u4,
                             double* ux, double* uy, double* uz, int32_t* iperm, double* yb_b,
                             float* yb32_b, double* dep_b, float* rinv_b, void* stream) {
  if (nobs <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  k_obs_gather<<<(nobs + 255) / 256, 256, 0, s>>>(nobs, perm, lat, lon, dep, rinv, cx, sx, yr,
                                                  (float4*)u4, ux, uy, uz, iperm, dep_b, rinv_b);
  const unsigned blocks = (unsigned)((nobs + 3) / 4);
  if (yb_b) k_obs_gather_yb<double><<<blocks, 256, 0, s>>>(nobs, k, ldy, perm, yb, yb_b);
  if (yb32_b) k_obs_gather_yb<float><<<blocks, 256, 0, s>>>(nobs, k, ldy, perm, yb, yb32_b);
  return (int)hipGetLastError();
}

extern "C" int dk_obs_ymax(int32_t nobs, int32_t k, int32_t ldy, const double* yb,
                           const float* rinv, double* ymax, void* stream) {
  if (nobs <= 0) return 0;
  k_obs_ymax<double><<<(nobs + 255) / 256, 256, 0, (hipStream_t)stream>>>(nobs, k, ldy, yb,
                                                                         rinv, ymax);
  return (int)hipGetLastError();
}

extern "C" int dk_obs_ymax32(int32_t nobs, int32_t k, int32_t ldy, const float* yb,
                             const float* rinv, double* ymax, void* stream) {
  if (nobs <= 0) return 0;
  k_obs_ymax<float><<<(nobs + 255) / 256, 256, 0, (hipStream_t)stream>>>(nobs, k, ldy, yb,
                                                                        rinv, ymax);
  return (int)hipGetLastError();
}

// Workgroups of four tiles over the tiles of the analysis grid (fill over a
// run of columns: the tiles of its longitudes).
static unsigned tile_blocks(const dk_select_args* a, bool fill, int tj, int ti) {
  const int64_t ntj = (a->nlat + tj - 1) / tj;
  int64_t ti0 = 0, ti1 = (a->nlon + ti - 1) / ti;
  if (fill && a->ncols > 0) {
    ti0 = (a->col0 / a->nlat) / ti;
    ti1 = ((a->col0 + a->ncols - 1) / a->nlat) / ti + 1;
  }
  return (unsigned)(((ti1 - ti0) * ntj + 3) / 4);
}

template <bool kFill, int kTerms>
static void select_tiles_launch(const dk_select_args* a, hipStream_t s) {
  switch (a->tile) {
    case DK_SELECT_TILE_4X1:
      k_select_tiles<kFill, 4, 1, kTerms><<<tile_blocks(a, kFill, 4, 1), 256, 0, s>>>(*a);
      break;
    case DK_SELECT_TILE_4X4:
      k_select_tiles<kFill, 4, 4, kTerms><<<tile_blocks(a, kFill, 4, 4), 256, 0, s>>>(*a);
      break;
    default:
      k_select_tiles<kFill, 4, 2, kTerms><<<tile_blocks(a, kFill, 4, 2), 256, 0, s>>>(*a);
      break;
  }
}

// The count pass needs no weights; the fill takes them from the chord angle
// when the support is within its range, with the shorter series when that
// suffices.
template <bool kFill>
static int select_tiles(const dk_select_args* a, hipStream_t s) {
  if constexpr (!kFill) {
    select_tiles_launch<false, 13>(a, s);
  } else {
    if (a->theta_deg <= kChordShortMaxDeg) select_tiles_launch<true, 8>(a, s);
    else if (a->theta_deg <= kChordMaxDeg) select_tiles_launch<true, 13>(a, s);
    else select_tiles_launch<true, 0>(a, s);
  }
  return (int)hipGetLastError();
}

static void tile_shape(int32_t tile, int* tj, int* ti) {
  *tj = 4;
  *ti = tile == DK_SELECT_TILE_4X1 ? 1 : tile == DK_SELECT_TILE_4X4 ? 4 : 2;
}

extern "C" int64_t dk_select_ntiles(int32_t nlat, int32_t nlon, int32_t tile) {
  int tj, ti;
  tile_shape(tile, &tj, &ti);
  return (int64_t)((nlat + tj - 1) / tj) * ((nlon + ti - 1) / ti);
}

extern "C" int dk_select_tile_order(const dk_select_args* a, int32_t ntiles, uint32_t* key,
                                    int32_t* id, uint32_t* key_sorted, int32_t* order, void* tmp,
                                    size_t* tmp_bytes, void* stream) {
  if (tmp == nullptr)
    return dk_sort_pairs_u32(nullptr, tmp_bytes, key, key_sorted, id, order, ntiles, 32, stream);
  if (ntiles <= 0) return 0;
  if (dk_select_ntiles(a->nlat, a->nlon, a->tile) != ntiles) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)((ntiles + 255) / 256);
  switch (a->tile) {
    case DK_SELECT_TILE_4X1: k_tile_cost<4, 1><<<blocks, 256, 0, s>>>(*a, ntiles, key, id); break;
    case DK_SELECT_TILE_4X4: k_tile_cost<4, 4><<<blocks, 256, 0, s>>>(*a, ntiles, key, id); break;
    default: k_tile_cost<4, 2><<<blocks, 256, 0, s>>>(*a, ntiles, key, id); break;
  }
  const int e = (int)hipGetLastError();
  if (e) return e;
  return dk_sort_pairs_u32(tmp, tmp_bytes, key, key_sorted, id, order, ntiles, 32, stream);
}

extern "C" int dk_select_count(const dk_select_args* a, void* stream) {
  const int64_t ncol = (int64_t)a->nlat * a->nlon;
  if (ncol <= 0) return 0;
  return select_tiles<false>(a, (hipStream_t)stream);
}

extern "C" int dk_select_cap(const dk_select_args* a, void* stream) {
  const int64_t ncol = (int64_t)a->nlat * a->nlon;
  if (ncol <= 0) return 0;
  if (a->perm_bits < 1 || a->perm_bits > 31) return (int)hipErrorInvalidValue;
  k_select_cap<<<(unsigned)ncol, 256, 0, (hipStream_t)stream>>>(*a);
  return (int)hipGetLastError();
}

extern "C" int dk_select_fill(const dk_select_args* a, void* stream) {
  const int64_t ncol = a->ncols > 0 ? a->ncols : (int64_t)a->nlat * a->nlon;
  if (ncol <= 0) return 0;
  // the fill's 32-bit byte offsets (binned obs, a column's entries)
  if (a->nobs >= (1 << 29)) return (int)hipErrorInvalidValue;
  return select_tiles<true>(a, (hipStream_t)stream);
}

namespace {

// Three-phase exclusive scan of per-column obs counts into int64 offsets:
// per-block totals, a one-block scan of the totals, then per-block scans.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void k_scan_block_sums(const int32_t* __restrict__ v,
                                                                int32_t n,
                                                                int64_t* __restrict__ bsum) {
  __shared__ int64_t s[kScanBlock / kWave];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  int64_t x = i < n ? (int64_t)v[i] : 0;
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) t += s[w];
    bsum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(int64_t* __restrict__ bsum, int32_t nb) {
  __shared__ int64_t s[kScanBlock];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += kScanBlock) {
    int i = base + threadIdx.x;
    int64_t v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kScanBlock; d <<= 1) {
      int64_t t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += s[kScanBlock - 1];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int32_t* __restrict__ v, int32_t n,
                                                           const int64_t* __restrict__ bsum,
                                                           int64_t* __restrict__ out) {
  __shared__ int64_t s[kScanBlock];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  int64_t x = i < n ? (int64_t)v[i] : 0;
  s[threadIdx.x] = x;
  __syncthreads();
  for (int d = 1; d < kScanBlock; d <<= 1) {
    int64_t t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t incl = bsum[blockIdx.x] + s[threadIdx.x];
  if (i < n) out[i] = incl - x;
  if (i == n - 1) out[n] = incl;
}

}  // namespace

// tmp must hold one int64 per block of 1024 counts.
extern "C" int dk_scan_offsets(void* tmp, size_t* tmp_bytes, const int32_t* nobsl,
                               int64_t* offset, int32_t n, void* stream) {
  const int32_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (tmp == nullptr) {
    *tmp_bytes = sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1);
    return 0;
  }
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0) return 0;
  int64_t* bsum = (int64_t*)tmp;
  k_scan_block_sums<<<nb, kScanBlock, 0, s>>>(nobsl, n, bsum);
  k_scan_sums<<<1, kScanBlock, 0, s>>>(bsum, nb);
  k_scan_apply<<<nb, kScanBlock, 0, s>>>(nobsl, n, bsum, offset);
  return (int)hipGetLastError();
}

// Vertical localisation: each column's selection list ordered by obs ln p.
// Keys (float ln p) and in-segment positions, one block per column.
__global__ void k_seg_keys(int64_t nseg, const int64_t* __restrict__ offset,
                           const int32_t* __restrict__ idx, const double* __restrict__ lnp,
                           float* __restrict__ key, int32_t* __restrict__ pos) {
  const int64_t col = blockIdx.x;
  if (col >= nseg) return;
  const int64_t o0 = offset[col], o1 = offset[col + 1];
  for (int64_t j = o0 + threadIdx.x; j < o1; j += blockDim.x) {
    key[j] = (float)lnp[idx[j]];
    pos[j] = (int32_t)(j - o0);
  }
}

__global__ void k_seg_gather(int64_t nseg, const int64_t* __restrict__ offset,
                             const int32_t* __restrict__ pos, const int32_t* __restrict__ idx,
                             const double* __restrict__ w, int32_t* __restrict__ idx2,
                             double* __restrict__ w2) {
  const int64_t col = blockIdx.x;
  if (col >= nseg) return;
  const int64_t o0 = offset[col], o1 = offset[col + 1];
  for (int64_t j = o0 + threadIdx.x; j < o1; j += blockDim.x) {
    const int64_t src = o0 + pos[j];
    idx2[j] = idx[src];
    w2[j] = w[src];
  }
}

// Segment bounds clamped to the workspace (lists past it are not sorted;
// their analyses see no obs and the sweep reports the overflow).
__global__ void k_clamp_offsets(int64_t n, const int64_t* __restrict__ offset, int64_t capacity,
                                int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) out[i] = offset[i] < capacity ? offset[i] : capacity;
}

extern "C" int dk_seg_sort_lnp(void* tmp, size_t* tmp_bytes, int64_t nseg, int64_t capacity,
                               const int64_t* offset, const int32_t* idx, const double* w,
                               const double* lnp, float* key_in, float* key_out,
                               int32_t* pos_in, int32_t* pos_out, int32_t* idx2, double* w2,
                               void* stream) {
  // work space: rocPRIM's, then the clamped offsets (nseg + 1, 256-aligned)
  const size_t off_bytes = ((size_t)(nseg + 1) * sizeof(int64_t) + 255) & ~(size_t)255;
  hipStream_t s = (hipStream_t)stream;
  if (tmp == nullptr) {
    size_t rb = 0;
    const int64_t* null_off = nullptr;
    int e = (int)rocprim::segmented_radix_sort_pairs(
        nullptr, rb, key_in, key_out, pos_in, pos_out, (unsigned int)capacity,
        (unsigned int)nseg, null_off, null_off, 0, 32, s, false);
    *tmp_bytes = off_bytes + ((rb + 255) & ~(size_t)255);
    return e;
  }
  if (nseg <= 0) return 0;
  int64_t* coff = (int64_t*)tmp;
  void* rtmp = (char*)tmp + off_bytes;
  size_t rb = *tmp_bytes - off_bytes;
  k_clamp_offsets<<<(unsigned)((nseg + 1 + 255) / 256), 256, 0, s>>>(nseg, offset, capacity, coff);
  k_seg_keys<<<(unsigned)nseg, 256, 0, s>>>(nseg, coff, idx, lnp, key_in, pos_in);
  int e = (int)rocprim::segmented_radix_sort_pairs(rtmp, rb, key_in, key_out, pos_in, pos_out,
                                                   (unsigned int)capacity, (unsigned int)nseg,
                                                   (const int64_t*)coff, (const int64_t*)coff + 1,
                                                   0, 32, s, false);
  if (e) return e;
  k_seg_gather<<<(unsigned)nseg, 256, 0, s>>>(nseg, coff, pos_out, idx, w, idx2, w2);
  return (int)hipGetLastError();
}
