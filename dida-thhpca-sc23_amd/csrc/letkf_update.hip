// Ensemble update (xb_2_xa, letkf_compute.F90:231-301) and the post-analysis
// blend of run_da_parallel (module_da_parallel.F90:216-259).
//
// Update, one workgroup per column: trans is staged in LDS as T[i][e] (member
// e contiguous, conflict-free), the column's Xb' rows stream through LDS a few
// rows at a time (ensemble-fastest, so every row is one coalesced load), and
// each lane forms Xa(e) = sum_i trans(i, e) Xb'(i) + xbar_b, the row mean, and
// the re-centred perturbation.  Optional RTPS (Whitaker & Hamill 2012; not in
// the reference): Xa' <- Xa' (alpha sigma_b / sigma_a + 1 - alpha).
#include "letkf_device.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

constexpr int kThreads = 256;

// Sum of v over the KP lanes of this thread's row slot.
template <int KP>
__device__ __forceinline__ double slot_sum(double v, double* sred) {
  if constexpr (KP <= 64) {
#pragma unroll
    for (int s = 1; s < KP; s <<= 1) v += shfl_xor_d(v, s);
    return v;
  } else {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) v += shfl_xor_d(v, s);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sred[w] = v;
    __syncthreads();
    const int slot0 = (w / (KP / 64)) * (KP / 64);
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < KP / 64; ++q) t += sred[slot0 + q];
    return t;
  }
}

// S: storage type of Xb' and Xa' (double, or float read from xb32 and
// written to xa32); the arithmetic is FP64 either way.
template <int KP, typename S>
__global__ __launch_bounds__(kThreads) void k_update_f64(dk_update_args u) {
  constexpr int RP = kThreads / KP;
  constexpr int LDT = KP + 1;
  __shared__ double sT[KP * LDT];
  __shared__ double sX[RP * KP];
  __shared__ double sred[kThreads / 64];
  const int c = blockIdx.x;
  const int64_t col = (int64_t)u.col0 + c;
  const int k = u.k, nrow = u.nrow;
  if (u.cube_n > 1) {
    // mean transform of the cubes covering the column, summed in the
    // reference's cube order (longitude starts outer, latitude inner);
    // averaging the transforms equals averaging the cubes' re-centred
    // analyses (xb_add2_xa / cube_count), which are linear in trans
    const int ilat = (int)(col % u.nlat), ilon = (int)(col / u.nlat);
    const int st = u.step, cn = u.cube_n;
    const int aj0 = ilat - cn + 1 > 0 ? (ilat - cn + 1 + st - 1) / st : 0;
    const int aj1 = min(ilat / st, u.nalat - 1);
    const int ai0 = ilon - cn + 1 > 0 ? (ilon - cn + 1 + st - 1) / st : 0;
    const int ai1 = min(ilon / st, u.nalon - 1);
    const double cnt = (double)((aj1 - aj0 + 1) * (ai1 - ai0 + 1));
    for (int t = threadIdx.x; t < k * k; t += kThreads) {
      double sum = 0.0;
      for (int ai = ai0; ai <= ai1; ++ai)
        for (int aj = aj0; aj <= aj1; ++aj)
          sum += u.trans[((int64_t)aj + (int64_t)u.nalat * ai) * k * k + t];
      const int i = t % k, e = t / k;
      sT[i * LDT + e] = sum / cnt;
    }
  } else {
    const double* T = u.trans + (int64_t)c * k * k;
    for (int t = threadIdx.x; t < k * k; t += kThreads) {
      const int i = t % k, e = t / k;
      sT[i * LDT + e] = T[t];
    }
  }
  const S* xbs;
  S* xas;
  if constexpr (sizeof(S) == sizeof(float)) {
    xbs = u.xb32;
    xas = u.xa32;
  } else {
    xbs = u.xb;
    xas = u.xa;
  }
  const int e = threadIdx.x % KP, slot = threadIdx.x / KP;
  const double rk = 1.0 / (double)k;
  for (int row0 = 0; row0 < nrow; row0 += RP) {
    const int row = row0 + slot;
    const bool rowok = row < nrow;
    const bool ok = rowok && e < k;
    const int64_t cr = col * nrow + row;
    const double xv = ok ? (double)xbs[cr * k + e] : 0.0;
    __syncthreads();
    sX[slot * KP + e] = xv;
    __syncthreads();
    double acc = 0.0;
    if (rowok)
      for (int i = 0; i < k; ++i) acc += sT[i * LDT + e] * sX[slot * KP + i];
    const double xbar = rowok ? u.axb[cr] : 0.0;
    const double v = ok ? acc + xbar : 0.0;
    const double mean = slot_sum<KP>(v, sred) / (double)k;
    double xa = ok ? v - mean : 0.0;
    if (u.rtps) {
      const double sb2 = slot_sum<KP>(xv * xv, sred) * rk;
      const double sa2 = slot_sum<KP>(xa * xa, sred) * rk;
      if (sa2 > 0.0) xa *= u.rtps_alpha * sqrt(sb2) / sqrt(sa2) + 1.0 - u.rtps_alpha;
    }
    if (ok) xas[cr * k + e] = (S)xa;
    if (rowok && e == 0) u.axa[cr] = mean;
  }
}

// 64 < k <= 128: the same update on FP64 MFMA 16x16x4, one 256-thread
// workgroup per column.  Xa(row, e) = sum_i Xb'(i, row) trans(i, e) + xbar_b
// with rows in tiles of 16 (M), members in tiles of 16 (N; wave w owns
// member tiles 2w and 2w + 1 and holds their trans columns in registers)
// and i in steps of 4 (K).  Each tile of 16 rows of Xb' is staged once in
// LDS as FP64 and read by all four waves; the row mean (and the RTPS
// spreads) meet across the waves in LDS.  Horizontal cubes take the mean
// transform of the covering cubes, as k_update_f64.
template <typename S>
__global__ __launch_bounds__(kThreads) void k_update_mfma128(dk_update_args u) {
  constexpr int KP = 128, NS = KP / 4, LDX = KP + 4;
  __shared__ double sX[16 * LDX];
  __shared__ double sv[4][16], sb[4][16], sa[4][16];
  const int c = blockIdx.x;
  const int64_t col = (int64_t)u.col0 + c;
  const int k = u.k, nrow = u.nrow;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, r16 = lane & 15, q4 = lane >> 4;
  const S* xbs;
  S* xas;
  if constexpr (sizeof(S) == sizeof(float)) {
    xbs = u.xb32;
    xas = u.xa32;
  } else {
    xbs = u.xb;
    xas = u.xa;
  }
  // B operands: trans(i = 4s + q4, e = (2w + j) * 16 + r16), zero past k
  double bv[2][NS];
  if (u.cube_n > 1) {
    const int ilat = (int)(col % u.nlat), ilon = (int)(col / u.nlat);
    const int st = u.step, cn = u.cube_n;
    const int aj0 = ilat - cn + 1 > 0 ? (ilat - cn + 1 + st - 1) / st : 0;
    const int aj1 = min(ilat / st, u.nalat - 1);
    const int ai0 = ilon - cn + 1 > 0 ? (ilon - cn + 1 + st - 1) / st : 0;
    const int ai1 = min(ilon / st, u.nalon - 1);
    const double cnt = (double)((aj1 - aj0 + 1) * (ai1 - ai0 + 1));
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int i = 4 * s + q4, e = (2 * w + j) * 16 + r16;
        double sum = 0.0;
        if (i < k && e < k)
          for (int ai = ai0; ai <= ai1; ++ai)
            for (int aj = aj0; aj <= aj1; ++aj)
              sum += u.trans[((int64_t)aj + (int64_t)u.nalat * ai) * k * k + i + (int64_t)e * k];
        bv[j][s] = sum / cnt;
      }
  } else {
    const double* T = u.trans + (int64_t)c * k * k;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int i = 4 * s + q4, e = (2 * w + j) * 16 + r16;
        bv[j][s] = i < k && e < k ? T[i + (int64_t)e * k] : 0.0;
      }
  }
  const double rk = 1.0 / (double)k;
  for (int r0 = 0; r0 < nrow; r0 += 16) {
    __syncthreads();  // the previous tile's sX, sv, sb, sa are read
    for (int t = tid; t < 16 * KP; t += kThreads) {
      const int rr = t / KP, e = t % KP, row = r0 + rr;
      sX[rr * LDX + e] = row < nrow && e < k ? (double)xbs[(col * nrow + row) * k + e] : 0.0;
    }
    __syncthreads();
    double4_t acc[2] = {double4_t{0.0, 0.0, 0.0, 0.0}, double4_t{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double a = sX[r16 * LDX + 4 * s + q4];
      acc[0] = mfma_f64_16x16x4(a, bv[0][s], acc[0]);
      acc[1] = mfma_f64_16x16x4(a, bv[1][s], acc[1]);
    }
    // acc[j][r]: row r0 + 4r + q4, member (2w + j) * 16 + r16
    double v[2][4], pv[4], pb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rl = 4 * r + q4, row = r0 + rl;
      const double xbar = row < nrow ? u.axb[col * nrow + row] : 0.0;
      pv[r] = 0.0;
      pb[r] = 0.0;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int e = (2 * w + j) * 16 + r16;
        const bool ok = row < nrow && e < k;
        v[j][r] = ok ? acc[j][r] + xbar : 0.0;
        const double xv = ok ? sX[rl * LDX + e] : 0.0;
        pv[r] += v[j][r];
        pb[r] += xv * xv;
      }
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        pv[r] += shfl_xor_d(pv[r], m);
        pb[r] += shfl_xor_d(pb[r], m);
      }
      if (r16 == 0) {
        sv[w][rl] = pv[r];
        sb[w][rl] = pb[r];
      }
    }
    __syncthreads();
    double mean[4], xa[2][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rl = 4 * r + q4;
      mean[r] = ((sv[0][rl] + sv[1][rl]) + (sv[2][rl] + sv[3][rl])) * rk;
      double pa = 0.0;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int e = (2 * w + j) * 16 + r16;
        const bool ok = r0 + rl < nrow && e < k;
        xa[j][r] = ok ? v[j][r] - mean[r] : 0.0;
        pa += xa[j][r] * xa[j][r];
      }
      if (u.rtps) {
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) pa += shfl_xor_d(pa, m);
        if (r16 == 0) sa[w][rl] = pa;
      }
    }
    if (u.rtps) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rl = 4 * r + q4;
        const double sb2 = ((sb[0][rl] + sb[1][rl]) + (sb[2][rl] + sb[3][rl])) * rk;
        const double sa2 = ((sa[0][rl] + sa[1][rl]) + (sa[2][rl] + sa[3][rl])) * rk;
        if (sa2 > 0.0) {
          const double f = u.rtps_alpha * sqrt(sb2) / sqrt(sa2) + 1.0 - u.rtps_alpha;
          xa[0][r] *= f;
          xa[1][r] *= f;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rl = 4 * r + q4, row = r0 + rl;
      if (row < nrow) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int e = (2 * w + j) * 16 + r16;
          if (e < k) xas[(col * nrow + row) * k + e] = (S)xa[j][r];
        }
        if (w == 0 && r16 == 0) u.axa[col * nrow + row] = mean[r];
      }
    }
  }
}

__global__ void k_cube_centres(int32_t nlat, int32_t nlon, int32_t cube_n, int32_t step,
                               const float* grid_lat, const float* grid_lon, float* ca_lat,
                               float* ca_lon) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nalat = (nlat + step - 1) / step, nalon = (nlon + step - 1) / step;
  if (t < nalat) {
    const int s = t * step, e = min(s + cube_n - 1, nlat - 1);
    ca_lat[t] = grid_lat[(s + e) / 2];
  } else if (t < nalat + nalon) {
    const int s = (t - nalat) * step, e = min(s + cube_n - 1, nlon - 1);
    ca_lon[t - nalat] = grid_lon[(s + e) / 2];
  }
}

__global__ void k_post(dk_post_args a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncr = (int64_t)a.nlat * a.nlon * a.nrow;
  if (t >= ncr) return;
  const int64_t col = t / a.nrow;
  const int ilat = (int)(col % a.nlat);
  const int k = a.k;
  const double* b = a.xb + t * k;
  const double* x = a.xa + t * k;
  double* o = a.xa_out + t * k;
  const double axb = a.axb[t], axa = a.axa[t];
  double sb = 0.0, sa = 0.0;
  for (int e = 0; e < k; ++e) sb += b[e] * b[e];
  if (fabs((double)a.grid_lat[ilat]) <= a.polar_lat) {
    for (int e = 0; e < k; ++e) {
      const double v = a.gamma1 * (axb + b[e]) + a.gamma2 * (axa + x[e]);
      o[e] = v;
      sa += (v - axa) * (v - axa);
    }
  } else {
    for (int e = 0; e < k; ++e) o[e] = b[e] + axb;
    sa = sb;
  }
  a.spread_b[t] = sqrt(sb / k);
  a.spread_a[t] = sqrt(sa / k);
}

// Coupler block xb_atm[((nn nx + i) ny + j) ens + e] <-> state
// xb[((j nx + i) n_max + n0 + nn) ens_max + e0 + e], one thread per element in
// coupler order: coupler-side accesses are fully coalesced, state-side ones
// contiguous over the ens members of a (row, column).
template <bool TO_STATE>
__global__ void k_struct_ens(int32_t ens, int32_t n, int32_t nx, int32_t ny, int32_t ens_max,
                             int32_t n_max, int32_t e0, int32_t n0, const double* __restrict__ src,
                             double* __restrict__ dst) {
  const int64_t total = (int64_t)ens * n * nx * ny;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int e = (int)(t % ens);
    int64_t r = t / ens;
    const int j = (int)(r % ny);
    r /= ny;
    const int i = (int)(r % nx);
    const int nn = (int)(r / nx);
    const int64_t s = (((int64_t)j * nx + i) * n_max + n0 + nn) * ens_max + e0 + e;
    if (TO_STATE) dst[s] = src[t];
    else dst[t] = src[s];
  }
}

// Means: axb_atm[(nn nx + i) ny + j] <-> axb[(j nx + i) n_max + n0 + nn], a
// transpose of (nn, j) per latitude i through a 32 x 33 LDS tile so that
// both sides are read and written along their contiguous index.
template <bool TO_STATE>
__global__ __launch_bounds__(256) void k_struct_mean(int32_t n, int32_t nx, int32_t ny,
                                                     int32_t n_max, int32_t n0,
                                                     const double* __restrict__ src,
                                                     double* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int j0 = blockIdx.x * 32, nn0 = blockIdx.y * 32, i = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  auto atm = [&](int nn, int j) { return ((int64_t)nn * nx + i) * ny + j; };
  auto st = [&](int nn, int j) { return ((int64_t)j * nx + i) * n_max + n0 + nn; };
  if (TO_STATE) {
    for (int y = ty; y < 32; y += 8) {
      const int nn = nn0 + y, j = j0 + tx;
      if (nn < n && j < ny) tile[y][tx] = src[atm(nn, j)];
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      const int j = j0 + y, nn = nn0 + tx;
      if (nn < n && j < ny) dst[st(nn, j)] = tile[tx][y];
    }
  } else {
    for (int y = ty; y < 32; y += 8) {
      const int j = j0 + y, nn = nn0 + tx;
      if (nn < n && j < ny) tile[tx][y] = src[st(nn, j)];
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      const int nn = nn0 + y, j = j0 + tx;
      if (nn < n && j < ny) dst[atm(nn, j)] = tile[y][tx];
    }
  }
}

}  // namespace

extern "C" int dk_struct_ens(int to_state, int32_t ens, int32_t n, int32_t nx, int32_t ny,
                             int32_t ens_max, int32_t n_max, int32_t e0, int32_t n0,
                             const double* src, double* dst, void* stream) {
  const int64_t total = (int64_t)ens * n * nx * ny;
  if (total <= 0) return 0;
  const int64_t nb = (total + 255) / 256;
  const unsigned blocks = (unsigned)(nb < 65536 * 8 ? nb : 65536 * 8);
  hipStream_t s = (hipStream_t)stream;
  if (to_state)
    k_struct_ens<true><<<blocks, 256, 0, s>>>(ens, n, nx, ny, ens_max, n_max, e0, n0, src, dst);
  else
    k_struct_ens<false><<<blocks, 256, 0, s>>>(ens, n, nx, ny, ens_max, n_max, e0, n0, src, dst);
  return (int)hipGetLastError();
}

extern "C" int dk_struct_mean(int to_state, int32_t n, int32_t nx, int32_t ny, int32_t n_max,
                              int32_t n0, const double* src, double* dst, void* stream) {
  if ((int64_t)n * nx * ny <= 0) return 0;
  if (nx > 65535) return (int)hipErrorInvalidValue;
  const dim3 grid((unsigned)((ny + 31) / 32), (unsigned)((n + 31) / 32), (unsigned)nx);
  hipStream_t s = (hipStream_t)stream;
  if (to_state) k_struct_mean<true><<<grid, 256, 0, s>>>(n, nx, ny, n_max, n0, src, dst);
  else k_struct_mean<false><<<grid, 256, 0, s>>>(n, nx, ny, n_max, n0, src, dst);
  return (int)hipGetLastError();
}

extern "C" int dk_update_f64(const dk_update_args* u, void* stream) {
  if (u->ncol <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int k = u->k;
  if (k < 2 || k > 128) return (int)hipErrorInvalidValue;
  if (u->xa32) {
    if (k <= 16) k_update_f64<16, float><<<u->ncol, kThreads, 0, s>>>(*u);
    else if (k <= 32) k_update_f64<32, float><<<u->ncol, kThreads, 0, s>>>(*u);
    else if (k <= 64) k_update_f64<64, float><<<u->ncol, kThreads, 0, s>>>(*u);
    else k_update_mfma128<float><<<u->ncol, kThreads, 0, s>>>(*u);
  } else {
    if (k <= 16) k_update_f64<16, double><<<u->ncol, kThreads, 0, s>>>(*u);
    else if (k <= 32) k_update_f64<32, double><<<u->ncol, kThreads, 0, s>>>(*u);
    else if (k <= 64) k_update_f64<64, double><<<u->ncol, kThreads, 0, s>>>(*u);
    else k_update_mfma128<double><<<u->ncol, kThreads, 0, s>>>(*u);
  }
  return (int)hipGetLastError();
}

extern "C" int dk_cube_centres(int32_t nlat, int32_t nlon, int32_t cube_n, int32_t step,
                               const float* grid_lat, const float* grid_lon, float* ca_lat,
                               float* ca_lon, void* stream) {
  if (nlat < 1 || nlon < 1 || cube_n < 1 || step < 1) return (int)hipErrorInvalidValue;
  const int n = (nlat + step - 1) / step + (nlon + step - 1) / step;
  k_cube_centres<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(nlat, nlon, cube_n, step,
                                                                   grid_lat, grid_lon, ca_lat,
                                                                   ca_lon);
  return (int)hipGetLastError();
}

extern "C" int dk_post(const dk_post_args* a, void* stream) {
  const int64_t ncr = (int64_t)a->nlat * a->nlon * a->nrow;
  if (ncr <= 0) return 0;
  k_post<<<(unsigned)((ncr + 255) / 256), 256, 0, (hipStream_t)stream>>>(*a);
  return (int)hipGetLastError();
}
