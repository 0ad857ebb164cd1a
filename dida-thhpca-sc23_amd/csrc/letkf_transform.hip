// LETKF transform (letkf_core, letkf_core.F90:14-233), plain FP64 form.
//
// One workgroup (4 wavefronts) per local analysis:
//   1. Gram C = Yhat Yhat^T on FP64 MFMA (16x16x4), Yhat = Yb sqrt(R^-1 rloc),
//      and b = Yhat (sqrt(R^-1 rloc) dep), both accumulated straight from the
//      gathered obs columns; the four waves split the obs and reduce in LDS.
//   2. One-sided (Hestenes) Jacobi, resident in LDS, on M = C + (k-1)/rho I
//      (the reference shifts the eigenvalues of C by (k-1)/rho,
//      letkf_core.F90:110-113): column pairs of a round-robin ordering are
//      orthogonalised in parallel until a sweep makes no rotation.  M is
//      symmetric positive definite, so the converged columns are lambda_j v_j
//      with every lambda_j > 0 and no separate eigenvector array is needed.
//   3. W = V diag(sqrt((k-1)/lambda)) V^T and wbar = V diag(1/lambda) V^T b
//      (letkf_core.F90:129-210), trans = W + wbar 1^T.
#include "letkf_device.h"
#include "letkf_jacobi.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

constexpr int kThreads = 256;

template <int KP>
struct CoreSmem {
  static constexpr int LDA = KP + 1;
  double A[KP * LDA];
  double b[KP];
  double nrm[KP];
  double sg[KP];
  double sh[KP];
  double u[KP];
  double wb[KP];
  int flag[4];
};

// KP = 128 Gram: wave W owns the lower tiles of tile rows W and NT-1-W
// (9 tiles each, fixed at compile time); every wave reads every obs.  Obs go
// in stages of U steps of 4 (lane group q4 takes obs 4u + q4 of a step): the
// rows and weights of stage s + 1 are in flight while stage s computes, the
// indices of stage s + 2 behind them.  Wave 0 also forms b.
template <int KP, int W, int U>
__device__ __forceinline__ void gram_f64_rows(const dk_core_args& a, int64_t off, int p,
                                              CoreSmem<KP>& sm) {
  constexpr int NT = KP / 16;
  constexpr int IA = W, IB = NT - 1 - W;  // IA < IB
  constexpr int NF = IB + 1;              // tile rows this wave reads
  constexpr int LDA = CoreSmem<KP>::LDA;
  const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
  const int k = a.k;
  double4_t accA[IA + 1], accB[IB + 1];
#pragma unroll
  for (int J = 0; J <= IA; ++J) accA[J] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int J = 0; J <= IB; ++J) accB[J] = double4_t{0.0, 0.0, 0.0, 0.0};
  double bacc[NT];
#pragma unroll
  for (int T = 0; T < NT; ++T) bacc[T] = 0.0;
  const int nst = (p + 4 * U - 1) / (4 * U);
  int gn[U];
  auto load_idx = [&](int st) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int o = st * 4 * U + 4 * u + q4;
      const int oc = o < p ? o : 0;
      gn[u] = a.idx ? a.idx[off + oc] : (int)(off + oc);
    }
  };
  // rows T*16 + r16 (clamped to k - 1, masked later) and sqrt(R^-1 rloc),
  // its product with d, of stage st
  auto load_rows = [&](int st, double (&f)[U][NF], double (&rr)[U], double (&sd)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int o = st * 4 * U + 4 * u + q4;
      const bool ok = o < p;
      const int64_t gi = gn[u];
      const double* y = a.yb + gi * k;
#pragma unroll
      for (int T = 0; T < NF; ++T) {
        const int row = T * 16 + r16;
        f[u][T] = y[row < k ? row : k - 1];
      }
      const double w = ok ? (double)a.rinv[gi] * a.rloc[off + (ok ? o : 0)] : 0.0;
      rr[u] = w;
      sd[u] = a.dep[gi];
    }
  };
  auto compute = [&](double (&f)[U][NF], const double (&rr)[U], const double (&sd)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double r = sqrt(rr[u]), d = r * sd[u];
#pragma unroll
      for (int T = 0; T < NF; ++T) f[u][T] = (T * 16 + r16 < k) ? f[u][T] * r : 0.0;
#pragma unroll
      for (int J = 0; J <= IA; ++J) accA[J] = mfma_f64_16x16x4(f[u][IA], f[u][J], accA[J]);
#pragma unroll
      for (int J = 0; J <= IB; ++J) accB[J] = mfma_f64_16x16x4(f[u][IB], f[u][J], accB[J]);
      if constexpr (W == 0)
#pragma unroll
        for (int T = 0; T < NT; ++T) bacc[T] += f[u][T] * d;
    }
  };
  if (nst > 0) {
    double fA[U][NF], fB[U][NF], rA[U], rB[U], sA[U], sB[U];
    load_idx(0);
    load_rows(0, fA, rA, sA);
    if (nst > 1) load_idx(1);
    for (int st = 0; st < nst; st += 2) {
      if (st + 1 < nst) load_rows(st + 1, fB, rB, sB);
      if (st + 2 < nst) load_idx(st + 2);
      __builtin_amdgcn_sched_barrier(0);
      compute(fA, rA, sA);
      if (st + 1 >= nst) break;
      if (st + 2 < nst) load_rows(st + 2, fA, rA, sA);
      if (st + 3 < nst) load_idx(st + 3);
      __builtin_amdgcn_sched_barrier(0);
      compute(fB, rB, sB);
    }
  }
  // D[4r + q4][r16] of tile (I, J) = C(I*16 + 4r + q4, J*16 + r16)
#pragma unroll
  for (int J = 0; J <= IA; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = IA * 16 + 4 * r + q4, j = J * 16 + r16;
      sm.A[i + j * LDA] = accA[J][r];
      if (IA != J) sm.A[j + i * LDA] = accA[J][r];
    }
#pragma unroll
  for (int J = 0; J <= IB; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = IB * 16 + 4 * r + q4, j = J * 16 + r16;
      sm.A[i + j * LDA] = accB[J][r];
      if (IB != J) sm.A[j + i * LDA] = accB[J][r];
    }
  if constexpr (W == 0) {
#pragma unroll
    for (int T = 0; T < NT; ++T) {
      bacc[T] += shfl_xor_d(bacc[T], 16);
      bacc[T] += shfl_xor_d(bacc[T], 32);
    }
    if (q4 == 0)
#pragma unroll
      for (int T = 0; T < NT; ++T) sm.b[T * 16 + r16] = bacc[T];
  }
}

template <int KP, int U>
__device__ __forceinline__ void gram_f64_tiles(const dk_core_args& a, int64_t off, int p,
                                               CoreSmem<KP>& sm) {
  static_assert(KP / 16 == 8, "four waves own two tile rows each");
  switch (threadIdx.x >> 6) {
    case 0: gram_f64_rows<KP, 0, U>(a, off, p, sm); break;
    case 1: gram_f64_rows<KP, 1, U>(a, off, p, sm); break;
    case 2: gram_f64_rows<KP, 2, U>(a, off, p, sm); break;
    default: gram_f64_rows<KP, 3, U>(a, off, p, sm); break;
  }
  __syncthreads();
}

// Gram + b.  KP <= 64: the obs are split across the four waves (all lower
// tiles of C fit in one wave's accumulators) and the partial sums meet in
// LDS.  KP = 128: the 36 lower tiles are split across the waves (tile t to
// wave t % 4), every wave reads every obs, and wave 0 also forms b.
template <int KP>
__device__ __forceinline__ void gram_f64_obs(const dk_core_args& a, int64_t off, int p,
                                             CoreSmem<KP>& sm) {
  constexpr int NT = KP / 16;
  constexpr int NTILE = NT * (NT + 1) / 2;
  constexpr int LDA = CoreSmem<KP>::LDA;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = lane >> 4;
  const int k = a.k;
  double4_t acc[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  double bacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) bacc[t] = 0.0;

  // four chunks of 4 obs per step; obs past p read entry 0 of the analysis
  // (in bounds: p > 0 inside the loop) with weight 0, so every load is
  // unconditional: indices first, then the obs rows and scalars
  constexpr int U = 4;
  for (int c0 = wave * U; c0 * 4 < p; c0 += 4 * U) {
    int64_t gi[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int o = (c0 + u) * 4 + q4;
      ok[u] = o < p;
      const int oc = ok[u] ? o : 0;
      gi[u] = a.idx ? (int64_t)a.idx[off + oc] : off + oc;
    }
    __builtin_amdgcn_sched_barrier(0);
    double f[U][NT], ri[U], rl[U], dp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int oc = ok[u] ? (c0 + u) * 4 + q4 : 0;
      ri[u] = (double)a.rinv[gi[u]];
      rl[u] = a.rloc[off + oc];
      dp[u] = a.dep[gi[u]];
      const double* y = a.yb + gi[u] * k;
#pragma unroll
      for (int T = 0; T < NT; ++T) {
        const int row = T * 16 + r16;
        f[u][T] = y[row < k ? row : k - 1];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double rr = ok[u] ? sqrt(ri[u] * rl[u]) : 0.0;
      const double sd = rr * dp[u];
#pragma unroll
      for (int T = 0; T < NT; ++T) f[u][T] = (T * 16 + r16 < k) ? f[u][T] * rr : 0.0;
#pragma unroll
      for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J)
          acc[I * (I + 1) / 2 + J] = mfma_f64_16x16x4(f[u][I], f[u][J], acc[I * (I + 1) / 2 + J]);
#pragma unroll
      for (int T = 0; T < NT; ++T) bacc[T] += f[u][T] * sd;
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

// KP = 128 Gram in half precision (batched form): Yhat = Yb sqrt(R^-1 rloc)
// in FP32 (Yb read from its FP32 copy, half the bytes of the gather), split
// as hi = h(Yhat), lo = h(Yhat - hi) (h: BF16 rounding), and C = hi hi^T +
// hi lo^T + lo hi^T (hi hi^T alone for GRAM = BF16; h: FP16 rounding of a
// power-of-two scaled Yhat for GRAM = FP16X3) on 16x16x32 MFMA with
// FP32 accumulation; b = Yb^T (R^-1 rloc d) accumulated in FP64.
// Blocks of 64 obs are converted once by the whole workgroup into LDS, in
// the MFMA operand layout: while block i is multiplied, block i + 1 is split
// and the rows of block i + 2 are requested (the obs metadata runs two blocks
// further ahead).  Thread (rq = tid % 32, og = tid / 32) loads rows
// 4 rq .. 4 rq + 3 of the 8 obs of octet og (one 16-byte load each) and
// writes their four 16-byte hi and lo vectors.  The staging area aliases A,
// which is free until C is written.  Wave W owns the lower tiles of tile
// rows W and NT-1-W.
template <int KP, typename HV>
struct H16Stage {
  HV hi[2][8][KP];        // [block & 1][obs octet][row]: obs 8 o .. 8 o + 7
  HV lo[2][8][KP];
  double rr[4][64];       // [block & 3]: sqrt(R^-1 rloc), 0 past p
  double w[4][64];        // R^-1 rloc dep
  int32_t gi[4][64];      // obs index
};

template <int KP, int GRAM, int W>
__device__ __forceinline__ void gram_h16_rows(const dk_core_args& a, int64_t off, int p,
                                              CoreSmem<KP>& sm) {
  typedef typename H16<GRAM>::vec hvec;
  typedef typename H16<GRAM>::elt helt;
  constexpr bool kSplit = GRAM != DK_GRAM_BF16;
  constexpr int NT = KP / 16;
  constexpr int IA = W, IB = NT - 1 - W;
  constexpr int NF = IB + 1;
  constexpr int LDA = CoreSmem<KP>::LDA;
  static_assert(sizeof(H16Stage<KP, hvec>) <= sizeof(sm.A), "staging fits in A");
  H16Stage<KP, hvec>& st = *reinterpret_cast<H16Stage<KP, hvec>*>(sm.A);
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, c = lane & 15;
  const int rq = tid & 31, og = tid >> 5;  // staging rows 4 rq.., obs octet og
  const int k = a.k;
  const bool rows_ok = 4 * rq < k;
  const int nblk = (p + 63) / 64;
  f32x4_t accA[IA + 1], accB[IB + 1];
#pragma unroll
  for (int J = 0; J <= IA; ++J) accA[J] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int J = 0; J <= IB; ++J) accB[J] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  double bacc[4] = {0.0, 0.0, 0.0, 0.0};
  f32x4_t rawA[8], rawB[8];  // rows of the even and the odd blocks in flight
  // FP16: a power-of-two scale per analysis puts the largest |Yhat| near
  // 2^11 (FP16's range); C is scaled back by its square
  double scale = 1.0, inv2 = 1.0;
  if constexpr (GRAM == DK_GRAM_FP16X3) {
    double mx = 0.0;
    for (int o = tid; o < p; o += 256) {
      const int gi = a.idx ? a.idx[off + o] : (int)(off + o);
      mx = fmax(mx, a.ymax[gi] * sqrt(a.rloc[off + o]));
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) mx = fmax(mx, __shfl_xor(mx, s, kWave));
    if (lane == 0) sm.nrm[tid >> 6] = mx;
    __syncthreads();
    mx = fmax(fmax(sm.nrm[0], sm.nrm[1]), fmax(sm.nrm[2], sm.nrm[3]));
    if (mx > 0.0) {
      int e;
      frexp(mx, &e);
      scale = ldexp(1.0, 11 - e);
      inv2 = ldexp(1.0, 2 * (e - 11));
    }
  }

  // obs metadata of block blk, threads 0-63 (obs past p take the index of
  // obs 0 of the analysis and zero weights), in two steps so that their
  // latencies hide behind the block being multiplied and split: the index
  // (meta_idx), then R^-1, rloc and d (meta_vals), then the LDS stores
  struct Meta {
    int gi, oc;
    bool ok;
    float ri;
    double rl, dp;
  };
  auto meta_idx = [&](int blk, Meta& m) {
    const int o = blk * 64 + tid;
    m.ok = o < p;
    m.oc = m.ok ? o : 0;
    m.gi = a.idx ? a.idx[off + m.oc] : (int)(off + m.oc);
  };
  auto meta_vals = [&](Meta& m) {
    m.ri = a.rinv[m.gi];
    m.rl = a.rloc[off + m.oc];
    m.dp = a.dep[m.gi];
  };
  auto meta_store = [&](const Meta& m, int buf) {
    const double r2 = m.ok ? (double)m.ri * m.rl : 0.0;
    st.gi[buf][tid] = m.gi;
    st.rr[buf][tid] = sqrt(r2);
    st.w[buf][tid] = r2 * m.dp;
  };
  auto meta = [&](int blk) {
    if (tid < 64) {
      Meta m;
      meta_idx(blk, m);
      meta_vals(m);
      meta_store(m, blk & 3);
    }
  };
  auto load = [&](int blk, f32x4_t (&raw)[8]) {
    const int ms = blk & 3;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      raw[e] = rows_ok ? *reinterpret_cast<const f32x4_t*>(
                             a.yb32 + (int64_t)st.gi[ms][og * 8 + e] * a.ldy32 + 4 * rq)
                       : f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  };
  auto split = [&](int blk, const f32x4_t (&raw)[8]) {
    const int ms = blk & 3, buf = blk & 1;
    float rs[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const double w = st.w[ms][og * 8 + e];
      rs[e] = (float)(st.rr[ms][og * 8 + e] * scale);
#pragma unroll
      for (int r = 0; r < 4; ++r) bacc[r] = fma((double)raw[e][r], w, bacc[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      hvec hv, lv;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float xs = raw[e][r] * rs[e];
        const helt h = (helt)xs;
        hv[e] = h;
        lv[e] = (helt)(xs - (float)h);
      }
      st.hi[buf][og][4 * rq + r] = hv;
      if (kSplit) st.lo[buf][og][4 * rq + r] = lv;
    }
  };
  auto mma = [&](int blk) {
    const int buf = blk & 1;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      hvec fh[NF], fl[NF];
#pragma unroll
      for (int T = 0; T < NF; ++T) {
        fh[T] = st.hi[buf][s * 4 + g][T * 16 + c];
        if (kSplit) fl[T] = st.lo[buf][s * 4 + g][T * 16 + c];
      }
#pragma unroll
      for (int J = 0; J <= IA; ++J) {
        accA[J] = H16<GRAM>::mfma(fh[IA], fh[J], accA[J]);
        if (kSplit) {
          accA[J] = H16<GRAM>::mfma(fh[IA], fl[J], accA[J]);
          accA[J] = H16<GRAM>::mfma(fl[IA], fh[J], accA[J]);
        }
      }
#pragma unroll
      for (int J = 0; J <= IB; ++J) {
        accB[J] = H16<GRAM>::mfma(fh[IB], fh[J], accB[J]);
        if (kSplit) {
          accB[J] = H16<GRAM>::mfma(fh[IB], fl[J], accB[J]);
          accB[J] = H16<GRAM>::mfma(fl[IB], fh[J], accB[J]);
        }
      }
    }
  };
  // one step: block blk is multiplied, blk + 1 split (its rows arrived
  // during the previous step, in rs), the rows of blk + 2 requested (into
  // rl) and the metadata of blk + 4 staged in the slot of block blk, which
  // no step reads any more; it is first read two steps later, so one
  // barrier a step orders everything
  auto pipe_step = [&](int blk, f32x4_t (&rl)[8], const f32x4_t (&rs)[8]) {
    const bool more = tid < 64 && blk + 4 < nblk;
    Meta m;
    if (more) meta_idx(blk + 4, m);
    if (blk + 2 < nblk) load(blk + 2, rl);
    __builtin_amdgcn_sched_barrier(0);
    mma(blk);
    if (more) meta_vals(m);
    if (blk + 1 < nblk) split(blk + 1, rs);
    if (more) meta_store(m, (blk + 4) & 3);
    __syncthreads();
  };
  if (nblk > 0) {
    for (int b = 0; b < 4 && b < nblk; ++b) meta(b);
    __syncthreads();
    load(0, rawA);
    if (nblk > 1) load(1, rawB);
    split(0, rawA);
    __syncthreads();
    for (int blk = 0; blk < nblk; blk += 2) {
      pipe_step(blk, rawA, rawB);
      if (blk + 1 >= nblk) break;
      pipe_step(blk + 1, rawB, rawA);
    }
  }
  // D[4g + r][c] of tile (I, J) = C(I*16 + 4g + r, J*16 + c)
#pragma unroll
  for (int J = 0; J <= IA; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = IA * 16 + 4 * g + r, j = J * 16 + c;
      sm.A[i + j * LDA] = (double)accA[J][r] * inv2;
      if (IA != J) sm.A[j + i * LDA] = (double)accA[J][r] * inv2;
    }
#pragma unroll
  for (int J = 0; J <= IB; ++J)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = IB * 16 + 4 * g + r, j = J * 16 + c;
      sm.A[i + j * LDA] = (double)accB[J][r] * inv2;
      if (IB != J) sm.A[j + i * LDA] = (double)accB[J][r] * inv2;
    }
  // b: the partials of the eight obs groups of each row meet in LDS (sm.u
  // and sm.sg, which the transform fills later)
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) bacc[r] += shfl_xor_d(bacc[r], 32);  // og pairs of a wave
  const int wv = tid >> 6;
  if (lane < 32) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (wv == 0) sm.u[4 * rq + r] = bacc[r];
      if (wv == 1) sm.sg[4 * rq + r] = bacc[r];
    }
  }
  __syncthreads();
  if (lane < 32 && wv >= 2) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (wv == 2) sm.u[4 * rq + r] += bacc[r];
      if (wv == 3) bacc[r] += sm.sg[4 * rq + r];
    }
  }
  __syncthreads();
  if (lane < 32 && wv == 3) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 4 * rq + r;
      sm.b[i] = i < k ? bacc[r] + sm.u[i] : 0.0;
    }
  }
}

template <int KP, int GRAM>
__device__ __forceinline__ void gram_h16_tiles(const dk_core_args& a, int64_t off, int p,
                                               CoreSmem<KP>& sm) {
  static_assert(KP == 128, "two k-steps of 128 rows per 64-obs block");
  switch (threadIdx.x >> 6) {
    case 0: gram_h16_rows<KP, GRAM, 0>(a, off, p, sm); break;
    case 1: gram_h16_rows<KP, GRAM, 1>(a, off, p, sm); break;
    case 2: gram_h16_rows<KP, GRAM, 2>(a, off, p, sm); break;
    default: gram_h16_rows<KP, GRAM, 3>(a, off, p, sm); break;
  }
  __syncthreads();
}

// U: k = 128 Gram stage length (steps of 4 obs), traded against registers
template <int KP, int U, int GRAM>
__device__ __forceinline__ void gram_any(const dk_core_args& a, int64_t off, int p,
                                         CoreSmem<KP>& sm) {
  if constexpr (GRAM != DK_GRAM_FP64) gram_h16_tiles<KP, GRAM>(a, off, p, sm);
  else if constexpr (KP > 64) gram_f64_tiles<KP, U>(a, off, p, sm);
  else gram_f64_obs<KP>(a, off, p, sm);
}

// One analysis by the workgroup.  vprev: eigenvectors of the previous
// analysis of a chain (column-major, KP x KP, in HBM), from which the
// eigensolve starts (A = M V_prev); vnext: where this analysis leaves its own.
// GRAM: precision of C (b, the eigensolve and the transform are FP64).
// A non-finite Yb, d or R^-1 reaches M (then an eigenvalue) or b; the
// transform is then NaN, as the reference's NaN would propagate (the Sunway
// caller checks trans and overline_w for it, letkf_compute.F90:168-187), and
// info is -2.  Returns whether the eigenvectors may warm-start the next
// analysis: not after a non-finite one, nor after one whose eigensolve hit
// the sweep cap (its basis is not orthonormal).
template <int KP, int U, int GRAM>
__device__ bool analysis_f64(const dk_core_args& a, int aid, CoreSmem<KP>& sm,
                             const double* vprev, double* vnext,
                             unsigned long long* phase_clk) {
  constexpr int NT = KP / 16;
  constexpr int LDA = CoreSmem<KP>::LDA;
  const int64_t slot = (int64_t)aid - a.aid0;
  const int k = a.k;
  const int64_t off = a.offset[aid] - a.base;
  int p = a.nobsl[aid];
  p = p < 0 ? 0 : (p > a.max_p ? a.max_p : p);
  if (off < 0 || off + p > a.capacity) p = 0;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, r16 = lane & 15, q4 = lane >> 4;

  // the previous analysis of a chain may still be reading A (its W tiles)
  __syncthreads();
  PhaseClock clk(phase_clk);
  gram_any<KP, U, GRAM>(a, off, p, sm);
  const double km1 = (double)(k - 1);
  const double a0 = km1 / a.rho;
  // M = C + (k-1)/rho I on the first k diagonal entries (letkf_core.F90:110-113
  // applies the shift to the eigenvalues; M is positive definite, so every
  // converged column is a non-null lambda_j v_j)
  for (int i = tid; i < k; i += kThreads) sm.A[i + i * LDA] += a0;
  __syncthreads();
  clk.mark(DK_PHASE_GRAM);
  if constexpr (KP == 128) {
    if (vprev) {
      // warm start A = M V_prev on MFMA: wave w owns the tiles of tile
      // columns w and w + 4 (all 8 tile rows); V_prev is read coherently
      // with L2 (another analysis of this workgroup wrote it)
      double4_t d[NT][2];
#pragma unroll
      for (int I = 0; I < NT; ++I) d[I][0] = d[I][1] = double4_t{0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < KP; k0 += 4) {
        const int kk = k0 + q4;
        const double b0 = __hip_atomic_load(vprev + kk + (wave * 16 + r16) * KP,
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double b1 = __hip_atomic_load(vprev + kk + ((wave + 4) * 16 + r16) * KP,
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int I = 0; I < NT; ++I) {
          const double am = sm.A[(I * 16 + r16) + kk * LDA];
          d[I][0] = mfma_f64_16x16x4(am, b0, d[I][0]);
          d[I][1] = mfma_f64_16x16x4(am, b1, d[I][1]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int jx = 0; jx < 2; ++jx)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            sm.A[(I * 16 + 4 * r + q4) + ((wave + 4 * jx) * 16 + r16) * LDA] = d[I][jx][r];
      __syncthreads();
    }
  }
  clk.mark(DK_PHASE_WARM);

  bool conv;
  int sweeps;
  if constexpr (KP == 128)
    sweeps = jacobi_regs128<LDA, kThreads>(sm.A, sm.flag, a.tol, a.max_sweeps, a.predict != 0,
                                           &conv);
  else
    sweeps = jacobi_lds<KP, LDA, kThreads>(sm.A, sm.nrm, sm.flag, a.tol, a.max_sweeps,
                                           a.predict != 0, &conv);
  clk.mark(DK_PHASE_JACOBI);
  const int m = (k + 1) & ~1;

  // exact column norms: alpha_j = lambda_j^2, lambda_j eigenvalues of M
  int nonfinite = 0;
  for (int j = tid; j < m; j += kThreads) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += sm.A[i + j * LDA] * sm.A[i + j * LDA];
    sm.nrm[j] = s;
    nonfinite |= !__builtin_isfinite(s);
  }
  for (int i = tid; i < k; i += kThreads) nonfinite |= !__builtin_isfinite(sm.b[i]);
  const bool bad = __syncthreads_or(nonfinite) != 0;
  const double nan = __builtin_nan("");
  if (vnext) {
    // v_j = a_j / lambda_j for the next analysis of the chain (1/lambda_j
    // staged in u, which is filled again below).  Only this workgroup reads
    // vnext back, on the same CU and L2, with loads that bypass the L1: the
    // stores need to have reached the L2, which the workgroup-scope release
    // and the barrier at the start of the next analysis ensure (an agent-
    // scope fence would also write back the L2 of the whole die).
    for (int j = tid; j < KP; j += kThreads) {
      const double al = j < m ? sm.nrm[j] : 0.0;
      sm.u[j] = al > 0.0 ? 1.0 / sqrt(al) : 0.0;
    }
    __syncthreads();
    for (int t = tid; t < KP * KP; t += kThreads) {
      const int i = t % KP, j = t / KP;
      vnext[t] = i < k ? sm.A[i + j * LDA] * sm.u[j] : 0.0;
    }
    __threadfence_block();
  }
  // with v_j = a_j / lambda_j: W = sum_j sqrt((k-1)/lambda_j) v_j v_j^T
  // = A diag(g) A^T, g_j = sqrt((k-1)/lambda_j) / alpha_j, and
  // wbar = sum_j v_j (v_j^T b) / lambda_j = A diag(h) A^T b, h_j = 1/(lambda_j alpha_j)
  for (int j = tid; j < KP; j += kThreads) {
    const double al = j < m ? sm.nrm[j] : 0.0;
    double g = 0.0, h = 0.0;
    if (al > 0.0) {
      const double lam = sqrt(al);
      g = sqrt(km1 / lam) / al;
      h = 1.0 / (lam * al);
    }
    sm.sg[j] = g;
    sm.sh[j] = h;
  }
  __syncthreads();
  for (int j = tid; j < KP; j += kThreads) {
    double s = 0.0;
    if (j < m)
      for (int i = 0; i < k; ++i) s += sm.A[i + j * LDA] * sm.b[i];
    sm.u[j] = s * sm.sh[j];
  }
  __syncthreads();
  for (int i = tid; i < KP; i += kThreads) {
    double s = 0.0;
    if (i < k)
      for (int j = 0; j < m; ++j) s += sm.A[i + j * LDA] * sm.u[j];
    sm.wb[i] = s;
  }
  __syncthreads();

  // W = (A diag(sg)) A^T on MFMA, written with wbar as trans.  The tile of
  // trans rows I, columns I2 is formed as W(I2 block, I block), whose result
  // layout puts 16 consecutive rows of a trans column in 16 lanes (W is
  // symmetric), so the stores are whole 128-byte lines.
  double* trans = a.trans + slot * k * k;
  // (columns past m are zero and sg is zero there, so the contraction runs
  // over all KP columns at compile-time trip counts, in two chains)
  for (int t = wave; t < NT * NT; t += 4) {
    const int I = t / NT, I2 = t % NT;
    double4_t d = double4_t{0.0, 0.0, 0.0, 0.0}, d1 = d;
#pragma unroll
    for (int j0 = 0; j0 < KP; j0 += 8) {
      const int ja = j0 + q4, jb = j0 + 4 + q4;
      d = mfma_f64_16x16x4(sm.A[I2 * 16 + r16 + ja * LDA] * sm.sg[ja],
                           sm.A[I * 16 + r16 + ja * LDA], d);
      d1 = mfma_f64_16x16x4(sm.A[I2 * 16 + r16 + jb * LDA] * sm.sg[jb],
                            sm.A[I * 16 + r16 + jb * LDA], d1);
    }
    d += d1;
    // d[r] = W(I2*16 + 4r + q4, I*16 + r16) = W(i, i2)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I * 16 + r16, i2 = I2 * 16 + 4 * r + q4;
      if (i < k && i2 < k) trans[i + (int64_t)i2 * k] = bad ? nan : d[r] + sm.wb[i];
    }
  }
  if (a.wbar)
    for (int i = tid; i < k; i += kThreads) a.wbar[slot * k + i] = bad ? nan : sm.wb[i];
  if (a.info && tid == 0) a.info[aid] = bad ? -2 : (conv ? sweeps : -1);
  clk.mark(DK_PHASE_TRANSFORM);
  return conv && !bad;
}


template <int KP>
__global__ __launch_bounds__(kThreads) void k_letkf_core_f64(dk_core_args a) {
  __shared__ CoreSmem<KP> sm;
  const int aid = a.list ? a.list[blockIdx.x] : a.aid0 + (int)blockIdx.x;
  analysis_f64<KP, 6, DK_GRAM_FP64>(a, aid, sm, nullptr, nullptr, nullptr);
}

// Analyses [c0, c1) of chain q of a launch over [aid0, aid0 + nanal): chains
// start at multiples of chain_len counted from analysis 0, so a launch that
// starts inside a chain runs the rest of it as its first chain, and how a
// sweep is cut into launches does not change its chains.
__device__ __forceinline__ void chain_range(const dk_core_args& a, int chain_len, int q, int* c0,
                                            int* c1) {
  const int s0 = a.aid0 - a.aid0 % chain_len + q * chain_len;
  *c0 = max(s0, a.aid0);
  *c1 = min(s0 + chain_len, a.aid0 + a.nanal);
}

// Chains of chain_len consecutive analyses drained from a work counter by
// persistent workgroups; every analysis after the first of a chain starts its
// eigensolve from the previous one's eigenvectors (a per-workgroup KP x KP
// buffer in HBM).
template <int KP, int GRAM>
__global__ __launch_bounds__(kThreads) void k_letkf_chain_f64(dk_core_args a, dk_chain_args ch) {
  __shared__ CoreSmem<KP> sm;
  __shared__ int item_s;
  double* vs = ch.vscr + (int64_t)blockIdx.x * KP * KP;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(ch.head, 1);
    __syncthreads();
    const int item = item_s;
    __syncthreads();
    if (item >= ch.nchain) break;
    int c0, c1;
    chain_range(a, ch.chain_len, ch.order ? ch.order[item] : item, &c0, &c1);
    bool prev_ok = false;  // whether the previous analysis may warm-start this one
    for (int aid = c0; aid < c1; ++aid)
      prev_ok = analysis_f64<KP, 4, GRAM>(a, aid, sm, prev_ok ? vs : nullptr,
                                          aid + 1 < c1 ? vs : nullptr, ch.phase_clk);
  }
}

__global__ void k_chain_costs(dk_core_args a, int32_t chain_len, int32_t nchain, uint32_t* key,
                              int32_t* id) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nchain) return;
  int c0, c1;
  chain_range(a, chain_len, q, &c0, &c1);
  double cost = 0.0;
  for (int aid = c0; aid < c1; ++aid) {
    const int p = a.nobsl[aid];
    cost += (double)(p < a.max_p ? p : a.max_p) + 10.0 * a.k;
  }
  key[q] = 0xFFFFFFFFu - (uint32_t)fmin(cost, 4.0e9);
  id[q] = q;
}

}  // namespace

extern "C" int dk_chain_costs(const dk_core_args* a, int32_t chain_len, int32_t nchain,
                              uint32_t* key, int32_t* id, void* stream) {
  if (nchain <= 0) return 0;
  if (chain_len < 1) return (int)hipErrorInvalidValue;
  k_chain_costs<<<(nchain + 255) / 256, 256, 0, (hipStream_t)stream>>>(*a, chain_len, nchain,
                                                                       key, id);
  return (int)hipGetLastError();
}

extern "C" int dk_letkf_chain_f64(const dk_core_args* a, const dk_chain_args* ch, int nblocks,
                                  void* stream) {
  if (a->nanal <= 0 || ch->nchain <= 0) return 0;
  if (a->k <= 64 || a->k > 128 || ch->chain_len < 1) return (int)hipErrorInvalidValue;
  if (a->gram != DK_GRAM_FP64 && !a->yb32) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  switch (a->gram) {
    case DK_GRAM_FP64: k_letkf_chain_f64<128, DK_GRAM_FP64><<<nblocks, kThreads, 0, s>>>(*a, *ch); break;
    case DK_GRAM_BF16X3: k_letkf_chain_f64<128, DK_GRAM_BF16X3><<<nblocks, kThreads, 0, s>>>(*a, *ch); break;
    case DK_GRAM_BF16: k_letkf_chain_f64<128, DK_GRAM_BF16><<<nblocks, kThreads, 0, s>>>(*a, *ch); break;
    case DK_GRAM_FP16X3:
      if (!a->ymax) return (int)hipErrorInvalidValue;
      k_letkf_chain_f64<128, DK_GRAM_FP16X3><<<nblocks, kThreads, 0, s>>>(*a, *ch);
      break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

extern "C" int dk_letkf_core_f64(const dk_core_args* a, void* stream) {
  if (a->nanal <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int k = a->k;
  if (k < 2 || k > 128 || a->gram != DK_GRAM_FP64) return (int)hipErrorInvalidValue;
  if (k <= 16) k_letkf_core_f64<16><<<a->nanal, kThreads, 0, s>>>(*a);
  else if (k <= 32) k_letkf_core_f64<32><<<a->nanal, kThreads, 0, s>>>(*a);
  else if (k <= 64) k_letkf_core_f64<64><<<a->nanal, kThreads, 0, s>>>(*a);
  else k_letkf_core_f64<128><<<a->nanal, kThreads, 0, s>>>(*a);
  return (int)hipGetLastError();
}
