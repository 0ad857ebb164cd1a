// Batched LETKF form, one wavefront per local analysis.
//
// A persistent grid of one-wave workgroups drains a queue of column chains
// (sorted by local obs count).  Each wave runs the whole analysis of every
// column of its chain with the KP x KP working matrix held in registers in
// the MFMA result layout: lane (g = l/16, c = l%16) owns, for each column
// slot j = J*16 + c (J < KP/16), the rows row(I, r, g) (I < KP/16, r < 4),
// where row() is I*16 + 4g + r after a 16x16x32 half-precision Gram and
// I*16 + 4r + g after an FP64 16x16x4 product.  Because M = C + (k-1)/rho I
// is symmetric, these registers are also the A operand of M V, and the tiles
// of W come out in the A-operand layout of the update, so the only LDS state
// is the eigenvector basis V (the warm start of the next column).
//
//   Gram      C = Yhat Yhat^T, Yhat = Yb sqrt(R^-1 rloc): FP16/BF16 MFMA with
//             FP32 accumulation and the split hi + lo (x3), FP16 scaled per
//             analysis by a power of two; or FP64 MFMA.  b = Yhat^T
//             (sqrt(R^-1 rloc) dep) in FP64 (letkf_core.F90:58-79, 146-195).
//   warm      A0 = M V_prev on FP64 MFMA for all but the first column of a
//             chain; M is positive definite, so V stays a full basis.
//   Jacobi    one-sided, FP64, pairs (columns J = 2P, 2P+1 of a lane group)
//             rotated in parallel; the round ordering pairs the two halves of
//             every block of positions, recursively (KP - 1 rounds a sweep),
//             with only the bottom columns moving between rounds.
//   W, wbar   W = V diag(sqrt((k-1)/lambda)) V^T, wbar = V diag(1/lambda) V^T b
//             on FP64 MFMA; trans = W + wbar 1^T (letkf_core.F90:129-225).
//   update    Xa = trans^T Xb' + xbar_b, mean, re-centre, RTPS (xb_2_xa,
//             letkf_compute.F90:261-299) on FP64 MFMA.
#include "letkf_device.h"
#include "letkf_jacobi.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int GRAM>
struct H16;
template <>
struct H16<DK_GRAM_BF16X3> {
  typedef bf16x8_t vec;
  typedef __bf16 elt;
  static __device__ __forceinline__ f32x4_t mfma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct H16<DK_GRAM_FP16X3> {
  typedef f16x8_t vec;
  typedef _Float16 elt;
  static __device__ __forceinline__ f32x4_t mfma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct H16<DK_GRAM_BF16> : H16<DK_GRAM_BF16X3> {};

template <int KP>
struct WSmem {
  static constexpr int LDV = KP + 1;
  double V[KP * LDV];  // eigenvector basis: column slot j at V[i + j LDV]
  double b[KP];
  double f[KP];
  double u[KP];
  int32_t ogi[64];
  double orr[64];
  double osd[64];
};

__device__ __forceinline__ int rowA(int I, int r, int g) { return I * 16 + 4 * g + r; }
__device__ __forceinline__ int rowB(int I, int r, int g) { return I * 16 + 4 * r + g; }
__device__ __forceinline__ int row_of(bool patB, int I, int r, int g) {
  return patB ? rowB(I, r, g) : rowA(I, r, g);
}

__device__ __forceinline__ double sum_g(double v) {
  v += shfl_xor_d(v, 16);
  v += shfl_xor_d(v, 32);
  return v;
}

// Stage obs [base, base + 64) of the analysis in LDS (binned index,
// sqrt(R^-1 rloc), sqrt(R^-1 rloc) dep).
template <int KP>
__device__ __forceinline__ void stage64(const dk_batched_args& a, int64_t off, int p, int base,
                                        WSmem<KP>& sm) {
  const int l = threadIdx.x;
  const int o = base + l;
  int gi = 0;
  double rr = 0.0, sd = 0.0;
  if (o < p) {
    gi = a.idx[off + o];
    rr = sqrt((double)a.rinv[gi] * a.rloc[off + o]);
    sd = rr * a.dep[gi];
  }
  __syncthreads();
  sm.ogi[l] = gi;
  sm.orr[l] = rr;
  sm.osd[l] = sd;
  __syncthreads();
}

template <int KP, int GRAM>
struct Wave {
  static constexpr int NJ = KP / 16;   // column sets per lane (and row blocks)
  static constexpr int NQ = NJ * 4;    // rows per column per lane
  static constexpr int NP = NJ / 2;    // column pairs per lane group
  static constexpr int LDV = WSmem<KP>::LDV;

  double X[NJ][NQ];  // working matrix (see header)
  bool patB;         // current row pattern

  // ---- Gram ---------------------------------------------------------------
  __device__ void gram(const dk_batched_args& a, int64_t off, int p, WSmem<KP>& sm) {
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const int k = a.k;
    double bacc[NJ];
#pragma unroll
    for (int I = 0; I < NJ; ++I) bacc[I] = 0.0;
    if constexpr (GRAM == DK_GRAM_FP64) {
      double4_t acc[NJ][NJ];
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J) acc[I][J] = double4_t{0.0, 0.0, 0.0, 0.0};
      for (int base = 0; base < p; base += 64) {
        stage64<KP>(a, off, p, base, sm);
        const int nb = min(64, p - base);
        for (int c0 = 0; c0 * 4 < nb; c0 += 4) {
          double fr[4][NJ], sd[4];
          // loads first (unconditional, in bounds), then the arithmetic
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double* y = a.yb + (int64_t)sm.ogi[(c0 + u) * 4 + g] * k;
#pragma unroll
            for (int I = 0; I < NJ; ++I) {
              const int row = I * 16 + c;
              fr[u][I] = y[row < k ? row : k - 1];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int ol = (c0 + u) * 4 + g;
            const double rr = sm.orr[ol];
            sd[u] = sm.osd[ol];
#pragma unroll
            for (int I = 0; I < NJ; ++I) fr[u][I] = (I * 16 + c < k) ? fr[u][I] * rr : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int I = 0; I < NJ; ++I) {
              bacc[I] += fr[u][I] * sd[u];
#pragma unroll
              for (int J = 0; J < NJ; ++J) acc[I][J] = mfma_f64_16x16x4(fr[u][I], fr[u][J], acc[I][J]);
            }
          }
        }
      }
      // D[4r + g][c] of tile (I, J) = C(I*16 + 4r + g, J*16 + c)
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) X[J][I * 4 + r] = acc[I][J][r];
      patB = true;
    } else {
      typedef typename H16<GRAM>::vec hvec;
      typedef typename H16<GRAM>::elt helt;
      constexpr bool kSplit = GRAM != DK_GRAM_BF16;
      double scale = 1.0, inv2 = 1.0;
      if (GRAM == DK_GRAM_FP16X3) {
        double mx = 0.0;
        for (int o = lane; o < p; o += 64) {
          const int gi = a.idx[off + o];
          mx = fmax(mx, a.ymax[gi] * sqrt(a.rloc[off + o]));
        }
        for (int s = 32; s >= 1; s >>= 1) mx = fmax(mx, __shfl_xor(mx, s, kWave));
        if (mx > 0.0) {
          int e;
          frexp(mx, &e);
          scale = ldexp(1.0, 11 - e);
          inv2 = ldexp(1.0, 2 * (e - 11));
        }
      }
      f32x4_t acc[NJ][NJ];
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J) acc[I][J] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
      for (int base = 0; base < p; base += 64) {
        stage64<KP>(a, off, p, base, sm);
        const int nb = min(64, p - base);
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          if (ch * 32 >= nb) break;  // uniform across the wave
          // all 8 x NJ loads of the chunk first (staged entries past p carry
          // index 0 and weight 0, so they are unconditional and in bounds),
          // then the arithmetic
          double raw[8][NJ];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const double* y = a.yb + (int64_t)sm.ogi[ch * 32 + 8 * g + e] * k;
#pragma unroll
            for (int I = 0; I < NJ; ++I) {
              const int row = I * 16 + c;
              raw[e][I] = y[row < k ? row : k - 1];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          hvec hi[NJ], lo[NJ];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int ol = ch * 32 + 8 * g + e;
            const double rr = sm.orr[ol];
            const double sd = sm.osd[ol];
            const float rs = (float)(rr * scale);
#pragma unroll
            for (int I = 0; I < NJ; ++I) {
              const bool rowok = I * 16 + c < k;
              const double xd = rowok ? raw[e][I] * rr : 0.0;
              bacc[I] += xd * sd;
              const float xs = rowok ? (float)raw[e][I] * rs : 0.0f;
              const helt h = (helt)xs;
              hi[I][e] = h;
              if (kSplit) lo[I][e] = (helt)(xs - (float)h);
            }
          }
#pragma unroll
          for (int I = 0; I < NJ; ++I)
#pragma unroll
            for (int J = 0; J < NJ; ++J) {
              acc[I][J] = H16<GRAM>::mfma(hi[I], hi[J], acc[I][J]);
              if (kSplit) {
                acc[I][J] = H16<GRAM>::mfma(hi[I], lo[J], acc[I][J]);
                acc[I][J] = H16<GRAM>::mfma(lo[I], hi[J], acc[I][J]);
              }
            }
        }
      }
      // D[4g + r][c] of tile (I, J) = C(I*16 + 4g + r, J*16 + c)
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) X[J][I * 4 + r] = (double)acc[I][J][r] * inv2;
      patB = false;
    }
    // b: lanes hold partials of rows I*16 + c over their obs; sum over g
#pragma unroll
    for (int I = 0; I < NJ; ++I) bacc[I] = sum_g(bacc[I]);
    __syncthreads();
    if (g == 0)
#pragma unroll
      for (int I = 0; I < NJ; ++I) sm.b[I * 16 + c] = bacc[I];
  }

  // M = C + a0 I on the first k diagonal entries
  __device__ void shift(int k, double a0) {
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int J = 0; J < NJ; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool diag = patB ? (4 * r + g == c) : (4 * g + r == c);
        if (diag && J * 16 + c < k) X[J][J * 4 + r] += a0;
      }
  }

  // X <- M V with V from LDS (the basis of the previous column).
  __device__ void warm(const WSmem<KP>& sm) {
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    double4_t d[NJ][NJ];
#pragma unroll
    for (int J = 0; J < NJ; ++J)
#pragma unroll
      for (int J2 = 0; J2 < NJ; ++J2) d[J][J2] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int kk = row_of(patB, I, r, g);  // contraction index of this chunk
#pragma unroll
        for (int J2 = 0; J2 < NJ; ++J2) {
          const double vb = sm.V[kk + (J2 * 16 + c) * LDV];
#pragma unroll
          for (int J = 0; J < NJ; ++J) d[J][J2] = mfma_f64_16x16x4(X[J][I * 4 + r], vb, d[J][J2]);
        }
      }
    // D[4r' + g][c] of (J, J2) = A0(J*16 + 4r' + g, J2*16 + c)
#pragma unroll
    for (int J = 0; J < NJ; ++J)
#pragma unroll
      for (int J2 = 0; J2 < NJ; ++J2)
#pragma unroll
        for (int r = 0; r < 4; ++r) X[J2][J * 4 + r] = d[J][J2][r];
    patB = true;
  }

  __device__ double colnorm(int J) const {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) s += X[J][q] * X[J][q];
    return sum_g(s);
  }

  // Exchange the bottom column of pair P with lane-group distance d (xor).
  __device__ void xswap(int P, int d, bool first, double* nb, double* nt) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double v = first ? X[2 * P + 1][q] : X[2 * P][q];
      const double w = shfl_xor_d(v, d);
      if (first) X[2 * P + 1][q] = w;
      else X[2 * P][q] = w;
    }
    const double v = first ? nb[P] : nt[P];
    const double w = shfl_xor_d(v, d);
    if (first) nb[P] = w;
    else nt[P] = w;
  }

  // One-sided Jacobi on X.  Returns sweeps; *conv when a sweep rotated
  // nothing, or (predict) when the largest relative off-diagonal of a sweep
  // was below sqrt(tol), so that the next sweep would rotate nothing.
  __device__ int jacobi(double tol, int max_sweeps, bool predict, bool* conv) {
    const int lane = threadIdx.x, c = lane & 15;
    constexpr int NPS = KP / 2;  // pair slots
    double nt[NP], nb[NP];
    int sweep = 0;
    *conv = false;
    for (; sweep < max_sweeps; ++sweep) {
#pragma unroll
      for (int P = 0; P < NP; ++P) {
        nt[P] = colnorm(2 * P);
        nb[P] = colnorm(2 * P + 1);
      }
      bool rotated = false;
      double dmax = 0.0;
      for (int G = NPS; G >= 1; G >>= 1) {
        // G rounds pairing the two halves of each block of 2G positions
        for (int rr = 0; rr < G; ++rr) {
          // all pairs of the lane are handled branch-free and side by side:
          // dot products in four partial sums, then the rotation, applied
          // as the identity (c = 1, s = t = 0) where none is needed
          double gmv[NP];
#pragma unroll
          for (int P = 0; P < NP; ++P) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q += 4) {
              s0 += X[2 * P][q] * X[2 * P + 1][q];
              s1 += X[2 * P][q + 1] * X[2 * P + 1][q + 1];
              s2 += X[2 * P][q + 2] * X[2 * P + 1][q + 2];
              s3 += X[2 * P][q + 3] * X[2 * P + 1][q + 3];
            }
            gmv[P] = (s0 + s1) + (s2 + s3);
          }
#pragma unroll
          for (int P = 0; P < NP; ++P) gmv[P] = sum_g(gmv[P]);
#pragma unroll
          for (int P = 0; P < NP; ++P) {
            const double gm = gmv[P];
            const double al = nt[P], be = nb[P];
            const double den = sqrt(al * be);
            dmax = fmax(dmax, den > 0.0 ? fabs(gm) / den : 0.0);
            const bool need = gm != 0.0 && fabs(gm) > tol * den;
            rotated |= need;
            double cs, sn, t;
            jacobi_rotation(need ? al : 0.0, need ? be : 1.0, need ? gm : 0.0, &cs, &sn, &t);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              const double x = X[2 * P][q], y = X[2 * P + 1][q];
              X[2 * P][q] = cs * x - sn * y;
              X[2 * P + 1][q] = sn * x + cs * y;
            }
            nt[P] = al - t * gm;
            nb[P] = be + t * gm;
          }
          if (rr + 1 < G) {
            // bottoms move one position down within the block of G positions:
            // position (c, P) takes the bottom of (c, P + 1), and the last pair
            // of a lane group the bottom of pair 0 of the next lane group
            // (cyclic within G / NP lane groups)
            if (G >= NP) {
              const int nlg = G / NP;  // lane groups per block
              const int cn = (c & ~(nlg - 1)) | ((c + 1) & (nlg - 1));
              const int src = (lane & ~15) | cn;
              double first[NQ];
              double n0 = nb[0];
#pragma unroll
              for (int q = 0; q < NQ; ++q) first[q] = X[1][q];
#pragma unroll
              for (int P = 0; P + 1 < NP; ++P) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) X[2 * P + 1][q] = X[2 * P + 3][q];
                nb[P] = nb[P + 1];
              }
              if (nlg > 1) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) first[q] = __shfl(first[q], src, kWave);
                n0 = __shfl(n0, src, kWave);
              }
#pragma unroll
              for (int q = 0; q < NQ; ++q) X[2 * NP - 1][q] = first[q];
              nb[NP - 1] = n0;
            }
          }
        }
        if (G > 1) {
          // next level: position p (first half of its block of 2G) swaps its
          // bottom with the top of p + G/2
          const int half = G / 2;  // in pair slots
          if (half >= NP) {
            const int d = half / NP;  // lane groups
            const bool first = ((c / d) & 1) == 0;
#pragma unroll
            for (int P = 0; P < NP; ++P) xswap(P, d, first, nb, nt);
          } else {
            // half < NP: partners are pairs P and P + half of the same lane
#pragma unroll
            for (int P = 0; P < NP; ++P) {
              if ((P / half) % 2 == 0) {
                const int P2 = P + half;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                  const double tmp = X[2 * P + 1][q];
                  X[2 * P + 1][q] = X[2 * P2][q];
                  X[2 * P2][q] = tmp;
                }
                const double tn = nb[P];
                nb[P] = nt[P2];
                nt[P2] = tn;
              }
            }
          }
        }
      }
      const bool any_rot = __any(rotated);
      for (int s = 32; s >= 1; s >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, s, kWave));
      if (!any_rot || (predict && dmax * dmax <= tol)) {
        *conv = true;
        ++sweep;
        break;
      }
    }
    return sweep;
  }
};

template <int KP, int GRAM>
__device__ void wave_column(const dk_batched_args& a, int64_t col, bool warm, WSmem<KP>& sm,
                            Wave<KP, GRAM>& w) {
  constexpr int NJ = Wave<KP, GRAM>::NJ;
  constexpr int NQ = Wave<KP, GRAM>::NQ;
  constexpr int LDV = WSmem<KP>::LDV;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int k = a.k;
  PhaseClock clk(a.phase_clk);
  const int64_t off = a.offset[col];
  int p = a.nobsl[col];
  p = p < 0 ? 0 : (p > a.max_p ? a.max_p : p);
  if (off < 0 || off + p > a.capacity) p = 0;

  w.gram(a, off, p, sm);
  const double km1 = (double)(k - 1);
  w.shift(k, km1 / a.rho);
  clk.mark(DK_PHASE_GRAM);
  if (warm) w.warm(sm);
  clk.mark(DK_PHASE_WARM);
  bool conv;
  const int sweeps = w.jacobi(a.tol, a.max_sweeps, GRAM != DK_GRAM_FP64, &conv);
  clk.mark(DK_PHASE_JACOBI);

  // lambda_j = |a_j|; V = A / lambda (to LDS); f = sqrt((k-1)/lambda);
  // u = V^T b / lambda
  double lam[NJ];
#pragma unroll
  for (int J = 0; J < NJ; ++J) lam[J] = sqrt(w.colnorm(J));
  __syncthreads();
#pragma unroll
  for (int J = 0; J < NJ; ++J) {
    const double il = lam[J] > 0.0 ? 1.0 / lam[J] : 0.0;
    double ub = 0.0;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row_of(w.patB, I, r, g);
        const double v = w.X[J][I * 4 + r] * il;
        sm.V[i + (J * 16 + c) * LDV] = v;
        ub += v * sm.b[i];
      }
    ub = sum_g(ub);
    if (g == 0) {
      sm.f[J * 16 + c] = lam[J] > 0.0 ? sqrt(km1 / lam[J]) : 0.0;
      sm.u[J * 16 + c] = ub * il;
    }
  }
  __syncthreads();

  // trans(i, e) = W(i, e) + wbar(i): tiles D[4r + g][c] = (I*16 + 4r + g, I2*16 + c)
  double4_t T[NJ][NJ], wb[NJ];
#pragma unroll
  for (int I = 0; I < NJ; ++I) {
    wb[I] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int I2 = 0; I2 < NJ; ++I2) T[I][I2] = double4_t{0.0, 0.0, 0.0, 0.0};
  }
  for (int j0 = 0; j0 < KP; j0 += 4) {
    const int jj = j0 + g;
    const double fj = sm.f[jj], uj = sm.u[jj];
    double v[NJ];
#pragma unroll
    for (int I = 0; I < NJ; ++I) v[I] = sm.V[I * 16 + c + jj * LDV];
#pragma unroll
    for (int I = 0; I < NJ; ++I) {
      wb[I] = mfma_f64_16x16x4(v[I], uj, wb[I]);
#pragma unroll
      for (int I2 = 0; I2 < NJ; ++I2) T[I][I2] = mfma_f64_16x16x4(v[I] * fj, v[I2], T[I][I2]);
    }
  }
#pragma unroll
  for (int I = 0; I < NJ; ++I)
#pragma unroll
    for (int I2 = 0; I2 < NJ; ++I2)
#pragma unroll
      for (int r = 0; r < 4; ++r) T[I][I2][r] += wb[I][r];
  if (a.info && lane == 0) a.info[col] = conv ? sweeps : -1;
  if (a.wbar_out && c == 0)
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rowB(I, r, g);
        if (i < k) a.wbar_out[col * k + i] = wb[I][r];
      }
  clk.mark(DK_PHASE_TRANSFORM);

  // update, 16 state rows at a time: D_E[4r' + g][c] = Xa(E*16 + 4r' + g, R0 + c)
  const int nrow = a.nrow;
  const double* xbc = a.xb + col * (int64_t)nrow * k;
  double* xac = a.xa + col * (int64_t)nrow * k;
  for (int R0 = 0; R0 < nrow; R0 += 16) {
    const int xr = R0 + c;
    const bool rok = xr < nrow;
    double4_t D[NJ];
#pragma unroll
    for (int E = 0; E < NJ; ++E) D[E] = double4_t{0.0, 0.0, 0.0, 0.0};
    double xl[NJ][4];
    const double* xrow = xbc + (int64_t)(rok ? xr : nrow - 1) * k;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rowB(I, r, g);
        xl[I][r] = xrow[i < k ? i : k - 1];
      }
    __builtin_amdgcn_sched_barrier(0);
    double ssb = 0.0;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rowB(I, r, g);
        const double xv = (rok && i < k) ? xl[I][r] : 0.0;
        ssb += xv * xv;
#pragma unroll
        for (int E = 0; E < NJ; ++E) D[E] = mfma_f64_16x16x4(T[I][E][r], xv, D[E]);
      }
    const double xbar = a.axb[col * nrow + (rok ? xr : nrow - 1)];
    double s = 0.0;
#pragma unroll
    for (int E = 0; E < NJ; ++E)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = rowB(E, r, g);
        D[E][r] = e < k ? D[E][r] + xbar : 0.0;
        s += D[E][r];
      }
    const double mean = sum_g(s) / (double)k;
    double ssa = 0.0;
#pragma unroll
    for (int E = 0; E < NJ; ++E)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = rowB(E, r, g);
        D[E][r] = e < k ? D[E][r] - mean : 0.0;
        ssa += D[E][r] * D[E][r];
      }
    double fac = 1.0;
    if (a.rtps) {
      ssb = sum_g(ssb);
      ssa = sum_g(ssa);
      if (ssa > 0.0) fac = a.rtps_alpha * sqrt(ssb) / sqrt(ssa) + 1.0 - a.rtps_alpha;
    }
    if (rok) {
#pragma unroll
      for (int E = 0; E < NJ; ++E)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = rowB(E, r, g);
          if (e < k) xac[(int64_t)xr * k + e] = D[E][r] * fac;
        }
      if (g == 0) a.axa[col * nrow + xr] = mean;
    }
  }
  clk.mark(DK_PHASE_UPDATE);
  (void)NQ;
}

template <int KP, int GRAM>
__global__ __launch_bounds__(64, 1) void k_letkf_wave(dk_batched_args a) {
  __shared__ WSmem<KP> sm;
  Wave<KP, GRAM> w;
  const int lane = threadIdx.x;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(a.queue_head, 1);
    item = __shfl(item, 0, kWave);
    if (item >= a.nchain) break;
    const int chain = a.queue ? a.queue[item] : item;
    const int ilon = chain / a.chains_per_lon;
    const int lat0 = (chain % a.chains_per_lon) * a.chain_len;
    const int lat1 = min(lat0 + a.chain_len, a.nlat);
    for (int ilat = lat0; ilat < lat1; ++ilat) {
      const int64_t col = (int64_t)ilon * a.nlat + ilat;
      wave_column<KP, GRAM>(a, col, a.warm_start && ilat > lat0, sm, w);
    }
  }
}

template <int KP, int GRAM>
int wlaunch(const dk_batched_args* a, int nblocks, hipStream_t s) {
  k_letkf_wave<KP, GRAM><<<nblocks, 64, 0, s>>>(*a);
  return (int)hipGetLastError();
}

template <int KP, int GRAM>
int wocc() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_letkf_wave<KP, GRAM>, 64, 0)) return 1;
  return nb > 0 ? nb : 1;
}

template <int KP>
int wdispatch(const dk_batched_args* a, int nblocks, hipStream_t s) {
  switch (a->gram) {
    case DK_GRAM_FP64: return wlaunch<KP, DK_GRAM_FP64>(a, nblocks, s);
    case DK_GRAM_BF16X3: return wlaunch<KP, DK_GRAM_BF16X3>(a, nblocks, s);
    case DK_GRAM_FP16X3: return wlaunch<KP, DK_GRAM_FP16X3>(a, nblocks, s);
    case DK_GRAM_BF16: return wlaunch<KP, DK_GRAM_BF16>(a, nblocks, s);
  }
  return (int)hipErrorInvalidValue;
}

template <int KP>
int wocc_dispatch(int gram) {
  switch (gram) {
    case DK_GRAM_FP64: return wocc<KP, DK_GRAM_FP64>();
    case DK_GRAM_BF16X3: return wocc<KP, DK_GRAM_BF16X3>();
    case DK_GRAM_FP16X3: return wocc<KP, DK_GRAM_FP16X3>();
    case DK_GRAM_BF16: return wocc<KP, DK_GRAM_BF16>();
  }
  return 1;
}

}  // namespace

extern "C" int dk_wave_blocks_per_cu(int k, int gram) {
  if (k <= 32) return wocc_dispatch<32>(gram);
  return wocc_dispatch<64>(gram);
}

extern "C" int dk_letkf_wave(const dk_batched_args* a, int nblocks, void* stream) {
  if (a->nchain <= 0) return 0;
  if (a->k < 2 || a->k > 64) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (a->k <= 32) return wdispatch<32>(a, nblocks, s);
  return wdispatch<64>(a, nblocks, s);
}
