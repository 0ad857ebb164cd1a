// Batched LETKF form: one persistent kernel drains a work queue of column
// chains sorted by local obs count; each workgroup (4 wavefronts) runs the
// whole local analysis of every column of its chain and the ensemble update,
// with nothing but the analysis ensemble going back to HBM.
//
// Per column:
//   1. Gram C = Yhat Yhat^T on BF16 or FP16 MFMA (32x32x16) with FP32
//      accumulation.  Yhat = Yb sqrt(R^-1 rloc) is split into two half-width
//      pieces hi + lo and C ~ hi hi^T + hi lo^T + lo hi^T (split x3).  FP16
//      pieces are scaled per analysis by a power of two so the largest entry
//      is ~2^11.  b = Yhat^T (sqrt(R^-1 rloc) dep) accumulates in FP64 from
//      the FP64 obs values (the mean increment stays FP64).  GRAM_FP64 runs the
//      Gram on FP64 MFMA (16x16x4) instead.
//   2. M = C + (k-1)/rho I.  For all but the first column of a chain the
//      eigenvector basis V of the previous (adjacent) column warm-starts the
//      one-sided Jacobi: A0 = M V (FP64 MFMA).  M is positive definite, so the
//      converged columns lambda_j v_j keep every direction and V stays a full
//      orthonormal basis.
//   3. W = V diag(sqrt((k-1)/lambda)) V^T (FP64 MFMA), wbar = V diag(1/lambda)
//      V^T b, trans = W + wbar 1^T (letkf_core.F90:97-225 with the shift of
//      :110-113 applied before the eigensolve).
//   4. Xa = trans^T Xb' + xbar_b, mean, re-centre (xb_2_xa,
//      letkf_compute.F90:261-299) on FP64 MFMA, rows streamed from HBM, with
//      optional RTPS.
#include "letkf_device.h"
#include "letkf_jacobi.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

constexpr int kThreads = 256;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

template <int KP>
struct BSmem {
  static constexpr int LDA = KP + 1;
  double A[KP * LDA];  // M -> Jacobi array -> trans (stored T[e + i LDA] = trans(i, e))
  double V[KP * LDA];  // eigenvectors of the current / previous column
  double b[KP];
  double nrm[KP];
  double lam[KP];
  double u[KP];
  double wb[KP];
  double red[4];
  int flag[2];
  int item;
  // per-obs scalars of the current block of kObsBlock local obs
  int32_t ogi[256];
  double orr[256];
  double osd[256];
};

constexpr int kObsBlock = 256;

// Stage obs [base, base + kObsBlock) of the analysis: binned index,
// sqrt(R^-1 rloc) and sqrt(R^-1 rloc) dep.  Ends with a barrier.
template <int KP>
__device__ __forceinline__ void stage_obs(const dk_batched_args& a, int64_t off, int p, int base,
                                          BSmem<KP>& sm) {
  const int o = base + threadIdx.x;
  int gi = 0;
  double rr = 0.0, sd = 0.0;
  if (o < p) {
    gi = a.idx[off + o];
    rr = sqrt((double)a.rinv[gi] * a.rloc[off + o]);
    sd = rr * a.dep[gi];
  }
  __syncthreads();
  sm.ogi[threadIdx.x] = gi;
  sm.orr[threadIdx.x] = rr;
  sm.osd[threadIdx.x] = sd;
  __syncthreads();
}

template <int GRAM>
struct Half;
template <>
struct Half<DK_GRAM_BF16X3> {
  typedef bf16x8_t vec;
  typedef __bf16 elt;
  static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Half<DK_GRAM_FP16X3> {
  typedef f16x8_t vec;
  typedef _Float16 elt;
  static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Half<DK_GRAM_BF16> : Half<DK_GRAM_BF16X3> {};

// Largest |Yhat| entry bound of one analysis (FP16 scaling).
__device__ double analysis_ymax(const dk_batched_args& a, int64_t off, int p, double* red) {
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int o = tid; o < p; o += kThreads) {
    const int gi = a.idx[off + o];
    mx = fmax(mx, a.ymax[gi] * sqrt(a.rloc[off + o]));
  }
  for (int s = 32; s >= 1; s >>= 1) mx = fmax(mx, __shfl_xor(mx, s, kWave));
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Mixed-precision Gram (+ FP64 b) into sm.A / sm.b; KP = 32 * NB.
template <int KP, int GRAM>
__device__ void gram_half(const dk_batched_args& a, int64_t off, int p, BSmem<KP>& sm) {
  constexpr int NB = KP / 32;
  constexpr int NTL = NB * (NB + 1) / 2;
  constexpr int LDA = BSmem<KP>::LDA;
  constexpr bool kSplit = GRAM != DK_GRAM_BF16;
  typedef typename Half<GRAM>::vec hvec;
  typedef typename Half<GRAM>::elt helt;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r32 = lane & 31, g2 = lane >> 5;
  const int k = a.k;
  double scale = 1.0, inv2 = 1.0;
  if (GRAM == DK_GRAM_FP16X3) {
    const double mx = analysis_ymax(a, off, p, sm.red);
    if (mx > 0.0) {
      int e;
      frexp(mx, &e);                 // mx < 2^e
      scale = ldexp(1.0, 11 - e);    // max |Yhat| scale < 2^11
      inv2 = ldexp(1.0, 2 * (e - 11));
    }
  }
  f32x16_t acc[NTL];
#pragma unroll
  for (int t = 0; t < NTL; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
  double bacc[NB];
#pragma unroll
  for (int I = 0; I < NB; ++I) bacc[I] = 0.0;

  for (int base = 0; base < p; base += kObsBlock) {
    stage_obs<KP>(a, off, p, base, sm);
    const int nb = min(kObsBlock, p - base);
    // chunks of 16 obs; wave w takes chunks w, w + 4, ...
    for (int c = wave; c * 16 < nb; c += 4) {
      hvec hi[NB], lo[NB];
      double x[8][NB];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int ol = c * 16 + 8 * g2 + e;
        const double rr = sm.orr[ol];
        const double* y = a.yb + (int64_t)sm.ogi[ol] * k;
#pragma unroll
        for (int I = 0; I < NB; ++I) {
          const int row = I * 32 + r32;
          x[e][I] = (row < k && ol < nb) ? y[row] * rr : 0.0;
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int ol = c * 16 + 8 * g2 + e;
        const double sd = ol < nb ? sm.osd[ol] : 0.0;
#pragma unroll
        for (int I = 0; I < NB; ++I) {
          bacc[I] += x[e][I] * sd;
          const float xs = (float)(x[e][I] * scale);
          const helt h = (helt)xs;
          hi[I][e] = h;
          if (kSplit) lo[I][e] = (helt)(xs - (float)h);
        }
      }
#pragma unroll
      for (int I = 0; I < NB; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          const int t = I * (I + 1) / 2 + J;
          acc[t] = Half<GRAM>::mfma(hi[I], hi[J], acc[t]);
          if (kSplit) {
            acc[t] = Half<GRAM>::mfma(hi[I], lo[J], acc[t]);
            acc[t] = Half<GRAM>::mfma(lo[I], hi[J], acc[t]);
          }
        }
    }
  }
#pragma unroll
  for (int I = 0; I < NB; ++I) bacc[I] += shfl_xor_d(bacc[I], 32);
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int I = 0; I < NB; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          const int t = I * (I + 1) / 2 + J;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            // 32x32 result layout: row 8 (r/4) + 4 (l/32) + r%4, column l%32
            const int i = I * 32 + 8 * (r >> 2) + 4 * g2 + (r & 3), j = J * 32 + r32;
            const double v = (double)acc[t][r] * inv2;
            if (w == 0) {
              sm.A[i + j * LDA] = v;
              if (I != J) sm.A[j + i * LDA] = v;
            } else {
              sm.A[i + j * LDA] += v;
              if (I != J) sm.A[j + i * LDA] += v;
            }
          }
        }
      if (g2 == 0) {
#pragma unroll
        for (int I = 0; I < NB; ++I) {
          if (w == 0) sm.b[I * 32 + r32] = bacc[I];
          else sm.b[I * 32 + r32] += bacc[I];
        }
      }
    }
    __syncthreads();
  }
}

// FP64 Gram (+ b) on 16x16x4 MFMA into sm.A / sm.b (KP <= 64).
template <int KP>
__device__ void gram_f64(const dk_batched_args& a, int64_t off, int p, BSmem<KP>& sm) {
  constexpr int NT = KP / 16;
  constexpr int NTILE = NT * (NT + 1) / 2;
  constexpr int LDA = BSmem<KP>::LDA;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = lane >> 4;
  const int k = a.k;
  double4_t acc[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  double bacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) bacc[t] = 0.0;
  for (int base = 0; base < p; base += kObsBlock) {
    stage_obs<KP>(a, off, p, base, sm);
    const int nb = min(kObsBlock, p - base);
    // chunks of 4 obs, four chunks per step; wave w takes steps w, w + 4, ...
    for (int c0 = wave * 4; c0 * 4 < nb; c0 += 16) {
      double f[4][NT], sd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ol = (c0 + u) * 4 + q4;
        const bool ok = ol < nb;
        const double rr = ok ? sm.orr[ol] : 0.0;
        sd[u] = ok ? sm.osd[ol] : 0.0;
        const double* y = a.yb + (int64_t)(ok ? sm.ogi[ol] : 0) * k;
#pragma unroll
        for (int T = 0; T < NT; ++T) {
          const int row = T * 16 + r16;
          f[u][T] = (ok && row < k) ? y[row] * rr : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int I = 0; I < NT; ++I)
#pragma unroll
          for (int J = 0; J <= I; ++J)
            acc[I * (I + 1) / 2 + J] = mfma_f64_16x16x4(f[u][I], f[u][J], acc[I * (I + 1) / 2 + J]);
#pragma unroll
        for (int T = 0; T < NT; ++T) bacc[T] += f[u][T] * sd[u];
      }
    }
  }
#pragma unroll
  for (int T = 0; T < NT; ++T) {
    bacc[T] += shfl_xor_d(bacc[T], 16);
    bacc[T] += shfl_xor_d(bacc[T], 32);
  }
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          const double4_t v = acc[I * (I + 1) / 2 + J];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = I * 16 + 4 * r + q4, j = J * 16 + r16;
            if (w == 0) {
              sm.A[i + j * LDA] = v[r];
              if (I != J) sm.A[j + i * LDA] = v[r];
            } else {
              sm.A[i + j * LDA] += v[r];
              if (I != J) sm.A[j + i * LDA] += v[r];
            }
          }
        }
      if (q4 == 0) {
#pragma unroll
        for (int T = 0; T < NT; ++T) {
          if (w == 0) sm.b[T * 16 + r16] = bacc[T];
          else sm.b[T * 16 + r16] += bacc[T];
        }
      }
    }
    __syncthreads();
  }
}

// A <- A V on FP64 MFMA (A, V in LDS, KP x KP, column-major).
template <int KP>
__device__ void lds_gemm_av(BSmem<KP>& sm) {
  constexpr int NT = KP / 16;
  constexpr int LDA = BSmem<KP>::LDA;
  constexpr int TPW = (NT * NT + 3) / 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = lane >> 4;
  double4_t d[TPW];
#pragma unroll
  for (int s = 0; s < TPW; ++s) {
    d[s] = double4_t{0.0, 0.0, 0.0, 0.0};
    const int t = wave + 4 * s;
    if (t < NT * NT) {
      const int I = t / NT, J = t % NT;
      for (int j0 = 0; j0 < KP; j0 += 4) {
        const double fa = sm.A[I * 16 + r16 + (j0 + q4) * LDA];
        const double fb = sm.V[(j0 + q4) + (J * 16 + r16) * LDA];
        d[s] = mfma_f64_16x16x4(fa, fb, d[s]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < TPW; ++s) {
    const int t = wave + 4 * s;
    if (t < NT * NT) {
      const int I = t / NT, J = t % NT;
#pragma unroll
      for (int r = 0; r < 4; ++r) sm.A[(I * 16 + 4 * r + q4) + (J * 16 + r16) * LDA] = d[s][r];
    }
  }
  __syncthreads();
}

template <int KP, int GRAM>
__device__ void analyse_column(const dk_batched_args& a, int64_t col, bool warm, BSmem<KP>& sm) {
  PhaseClock clk(a.phase_clk);
  constexpr int NT = KP / 16;
  constexpr int LDA = BSmem<KP>::LDA;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, r16 = lane & 15, q4 = lane >> 4;
  const int k = a.k;
  const int m = (k + 1) & ~1;
  const int64_t off = a.offset[col];
  int p = a.nobsl[col];
  p = p < 0 ? 0 : (p > a.max_p ? a.max_p : p);
  if (off < 0 || off + p > a.capacity) p = 0;

  if constexpr (GRAM == DK_GRAM_FP64) gram_f64<KP>(a, off, p, sm);
  else gram_half<KP, GRAM>(a, off, p, sm);
  clk.mark(DK_PHASE_GRAM);

  // M = C + (k-1)/rho I, then the warm start A0 = M V_prev
  const double km1 = (double)(k - 1);
  const double a0 = km1 / a.rho;
  for (int i = tid; i < k; i += kThreads) sm.A[i + i * LDA] += a0;
  __syncthreads();
  if (warm) lds_gemm_av<KP>(sm);
  clk.mark(DK_PHASE_WARM);

  bool conv;
  const int sweeps = jacobi_lds<KP, LDA, kThreads>(sm.A, sm.nrm, sm.flag, a.tol, a.max_sweeps,
                                                   &conv);
  clk.mark(DK_PHASE_JACOBI);
  // lambda_j = |a_j|, v_j = a_j / lambda_j
  for (int j = tid; j < KP; j += kThreads) {
    double s = 0.0;
    if (j < m)
      for (int i = 0; i < k; ++i) s += sm.A[i + j * LDA] * sm.A[i + j * LDA];
    sm.lam[j] = sqrt(s);
  }
  __syncthreads();
  for (int t = tid; t < KP * KP; t += kThreads) {
    const int i = t % KP, j = t / KP;
    const double l = sm.lam[j];
    sm.V[i + j * LDA] = (i < k && j < m && l > 0.0) ? sm.A[i + j * LDA] / l : 0.0;
  }
  __syncthreads();
  // u = diag(1/lambda) V^T b ; wbar = V u
  for (int j = tid; j < KP; j += kThreads) {
    double s = 0.0;
    const double l = sm.lam[j];
    if (j < m && l > 0.0) {
      for (int i = 0; i < k; ++i) s += sm.V[i + j * LDA] * sm.b[i];
      s /= l;
    }
    sm.u[j] = s;
  }
  __syncthreads();
  for (int i = tid; i < KP; i += kThreads) {
    double s = 0.0;
    if (i < k)
      for (int j = 0; j < m; ++j) s += sm.V[i + j * LDA] * sm.u[j];
    sm.wb[i] = s;
  }
  __syncthreads();
  // trans(i, e) = W(i, e) + wbar(i), stored at A[e + i LDA]
  for (int t = wave; t < NT * NT; t += 4) {
    const int I = t / NT, I2 = t % NT;
    double4_t d = double4_t{0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < m; j0 += 4) {
      const int jj = j0 + q4;
      double fa = 0.0, fb = 0.0;
      const double l = sm.lam[jj < KP ? jj : 0];
      if (jj < m && l > 0.0) {
        fa = sm.V[I * 16 + r16 + jj * LDA] * sqrt(km1 / l);
        fb = sm.V[I2 * 16 + r16 + jj * LDA];
      }
      d = mfma_f64_16x16x4(fa, fb, d);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I * 16 + 4 * r + q4, e = I2 * 16 + r16;
      sm.A[e + i * LDA] = (i < k && e < k) ? d[r] + sm.wb[i] : 0.0;
    }
  }
  __syncthreads();
  if (a.info && tid == 0) a.info[col] = conv ? sweeps : -1;
  if (a.wbar_out)
    for (int i = tid; i < k; i += kThreads) a.wbar_out[col * k + i] = sm.wb[i];

  clk.mark(DK_PHASE_TRANSFORM);
  // Update: per 16-row block, D(e, row) = sum_i trans(i, e) Xb'(i, row)
  const int nrow = a.nrow;
  const int nrb = (nrow + 15) / 16;
  const double* xbc = a.xb + col * (int64_t)nrow * k;
  double* xac = a.xa + col * (int64_t)nrow * k;
  for (int rb = wave; rb < nrb; rb += 4) {
    const int row = rb * 16 + r16;
    const bool rok = row < nrow;
    double4_t d[NT];
#pragma unroll
    for (int E = 0; E < NT; ++E) d[E] = double4_t{0.0, 0.0, 0.0, 0.0};
    double ss = 0.0;
    for (int i0 = 0; i0 < k; i0 += 4) {
      const int i = i0 + q4;
      const double xv = (rok && i < k) ? xbc[(int64_t)row * k + i] : 0.0;
      ss += xv * xv;
#pragma unroll
      for (int E = 0; E < NT; ++E) {
        const double fa = sm.A[E * 16 + r16 + i * LDA];
        d[E] = mfma_f64_16x16x4(fa, xv, d[E]);
      }
    }
    // D[e_local = 4r + q4][row = r16]
    const double xbar = rok ? a.axb[col * nrow + row] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int E = 0; E < NT; ++E)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = E * 16 + 4 * r + q4;
        d[E][r] = e < k ? d[E][r] + xbar : 0.0;
        s += d[E][r];
      }
    s += shfl_xor_d(s, 16);
    s += shfl_xor_d(s, 32);
    const double mean = s / (double)k;
    double sa = 0.0;
#pragma unroll
    for (int E = 0; E < NT; ++E)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = E * 16 + 4 * r + q4;
        d[E][r] = e < k ? d[E][r] - mean : 0.0;
        sa += d[E][r] * d[E][r];
      }
    double fac = 1.0;
    if (a.rtps) {
      ss += shfl_xor_d(ss, 16);
      ss += shfl_xor_d(ss, 32);
      sa += shfl_xor_d(sa, 16);
      sa += shfl_xor_d(sa, 32);
      if (sa > 0.0) fac = a.rtps_alpha * sqrt(ss) / sqrt(sa) + 1.0 - a.rtps_alpha;
    }
    if (rok) {
#pragma unroll
      for (int E = 0; E < NT; ++E)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = E * 16 + 4 * r + q4;
          if (e < k) xac[(int64_t)row * k + e] = d[E][r] * fac;
        }
      if (q4 == 0) a.axa[col * nrow + row] = mean;
    }
  }
  __syncthreads();
  clk.mark(DK_PHASE_UPDATE);
}

template <int KP, int GRAM>
__global__ __launch_bounds__(kThreads, 2) void k_letkf_batched(dk_batched_args a) {
  __shared__ BSmem<KP> sm;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) sm.item = atomicAdd(a.queue_head, 1);
    __syncthreads();
    const int item = sm.item;
    __syncthreads();
    if (item >= a.nchain) break;
    const int chain = a.queue ? a.queue[item] : item;
    const int ilon = chain / a.chains_per_lon;
    const int lat0 = (chain % a.chains_per_lon) * a.chain_len;
    const int lat1 = min(lat0 + a.chain_len, a.nlat);
    for (int ilat = lat0; ilat < lat1; ++ilat) {
      const int64_t col = (int64_t)ilon * a.nlat + ilat;
      analyse_column<KP, GRAM>(a, col, a.warm_start && ilat > lat0, sm);
    }
  }
}

__global__ void k_chain_keys(dk_batched_args a, uint32_t* key, int32_t* id) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nchain) return;
  const int ilon = c / a.chains_per_lon;
  const int lat0 = (c % a.chains_per_lon) * a.chain_len;
  const int lat1 = min(lat0 + a.chain_len, a.nlat);
  // work ~ 4 k^2 p (Gram) + ~ 40 k^3 (eigensolve, transform, update) per column
  double cost = 0.0;
  for (int ilat = lat0; ilat < lat1; ++ilat)
    cost += (double)a.nobsl[(int64_t)ilon * a.nlat + ilat] + 10.0 * a.k;
  const double c32 = fmin(cost, 4.0e9);
  key[c] = 0xFFFFFFFFu - (uint32_t)c32;
  id[c] = c;
}

template <int KP, int GRAM>
int launch(const dk_batched_args* a, int nblocks, hipStream_t s) {
  k_letkf_batched<KP, GRAM><<<nblocks, kThreads, 0, s>>>(*a);
  return (int)hipGetLastError();
}

template <int KP, int GRAM>
int occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_letkf_batched<KP, GRAM>, kThreads, 0))
    return 1;
  return nb > 0 ? nb : 1;
}

template <int KP>
int dispatch(const dk_batched_args* a, int nblocks, hipStream_t s) {
  switch (a->gram) {
    case DK_GRAM_FP64: return launch<KP, DK_GRAM_FP64>(a, nblocks, s);
    case DK_GRAM_BF16X3: return launch<KP, DK_GRAM_BF16X3>(a, nblocks, s);
    case DK_GRAM_FP16X3: return launch<KP, DK_GRAM_FP16X3>(a, nblocks, s);
    case DK_GRAM_BF16: return launch<KP, DK_GRAM_BF16>(a, nblocks, s);
  }
  return (int)hipErrorInvalidValue;
}

template <int KP>
int occ_dispatch(int gram) {
  switch (gram) {
    case DK_GRAM_FP64: return occupancy<KP, DK_GRAM_FP64>();
    case DK_GRAM_BF16X3: return occupancy<KP, DK_GRAM_BF16X3>();
    case DK_GRAM_FP16X3: return occupancy<KP, DK_GRAM_FP16X3>();
    case DK_GRAM_BF16: return occupancy<KP, DK_GRAM_BF16>();
  }
  return 1;
}

}  // namespace

extern "C" int dk_batched_blocks_per_cu(int k, int gram) {
  if (k <= 32) return occ_dispatch<32>(gram);
  return occ_dispatch<64>(gram);
}

extern "C" int dk_chain_keys(const dk_batched_args* a, uint32_t* key, int32_t* id, void* stream) {
  if (a->nchain <= 0) return 0;
  k_chain_keys<<<(a->nchain + 255) / 256, 256, 0, (hipStream_t)stream>>>(*a, key, id);
  return (int)hipGetLastError();
}

extern "C" int dk_letkf_batched(const dk_batched_args* a, int nblocks, void* stream) {
  if (a->nchain <= 0) return 0;
  if (a->k < 2 || a->k > 64) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (a->k <= 32) return dispatch<32>(a, nblocks, s);
  return dispatch<64>(a, nblocks, s);
}
