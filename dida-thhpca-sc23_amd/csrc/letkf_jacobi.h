// One-sided (Hestenes) Jacobi on a KP x KP block resident in LDS.
//
// The KP columns of A (column-major, leading dimension LDA; rows >= k and
// columns >= k are zero padding) are orthogonalised pairwise; in each of the
// KP - 1 rounds of a round-robin (circle) ordering the KP/2 disjoint pairs are
// rotated in parallel, TPP = 2 * NTHREADS / KP threads per pair, each owning
// KP / TPP rows of both columns (compile-time trip counts, so all LDS loads of
// a round are issued together).  A pair (p, q) is rotated while
// |a_p . a_q| > tol |a_p| |a_q|; column norms are recomputed at the start of
// each sweep and updated exactly in between (alpha' = alpha - t g,
// beta' = beta + t g).  A sweep without any rotation ends the iteration, or,
// with predict, a sweep in which no pair had g^2 > tol alpha beta (Jacobi
// converges quadratically, so the next sweep would rotate nothing).
// flag[4] is LDS scratch.
// For a symmetric positive (semi)definite input the converged columns are
// lambda_j v_j (Hestenes 1958; Demmel & Veselic 1992).
#pragma once

#include "letkf_device.h"

namespace dida {

// Rotation (c, s, t) that makes a_p' = c a_p - s a_q and a_q' = s a_p + c a_q
// orthogonal, for alpha = |a_p|^2, beta = |a_q|^2, g = a_p . a_q != 0:
// t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (beta - alpha) / (2 g),
// written as t = sign(d e) |e| / (|d| + sqrt(d^2 + e^2)), d = beta - alpha,
// e = 2 g, which needs one division.
__device__ __forceinline__ void jacobi_rotation(double al, double be, double g, double* c,
                                                double* s, double* t) {
  const double d = be - al, e = 2.0 * g;
  const double den = fabs(d) + sqrt(d * d + e * e);
  double tt = fabs(e) / den;
  if ((d < 0.0) != (e < 0.0) && d != 0.0) tt = -tt;
  const double cc = 1.0 / sqrt(1.0 + tt * tt);
  *c = cc;
  *s = cc * tt;
  *t = tt;
}

template <int KP, int LDA, int NTHREADS>
__device__ int jacobi_lds(double* A, double* nrm, int* flag, double tol, int max_sweeps,
                          bool predict, bool* converged) {
  constexpr int NPAIR = KP / 2;
  constexpr int TPP0 = NTHREADS / NPAIR;
  constexpr int TPP = TPP0 > 32 ? 32 : TPP0;   // threads per pair (power of two)
  constexpr int RPT = (KP + TPP - 1) / TPP;    // rows per thread
  const int tid = threadIdx.x;
  const int pair = tid / TPP, sub = tid % TPP;
  const bool active = pair < NPAIR;
  if (tid == 0) {
    flag[0] = 0;
    flag[2] = 0;
  }
  int sweep = 0;
  *converged = false;
  for (; sweep < max_sweeps; ++sweep) {
    for (int j = tid; j < KP; j += NTHREADS) {
      double s = 0.0;
#pragma unroll 16
      for (int i = 0; i < KP; ++i) s += A[i + j * LDA] * A[i + j * LDA];
      nrm[j] = s;
    }
    __syncthreads();
    bool rotated = false, big = false;
    for (int r = 0; r < KP - 1; ++r) {
      if (active) {
        int p, q;
        if (pair == 0) {
          p = r;
          q = KP - 1;
        } else {
          p = (r - pair + (KP - 1)) % (KP - 1);
          q = (r + pair) % (KP - 1);
          if (p > q) {
            const int t = p;
            p = q;
            q = t;
          }
        }
        double* ap = A + p * LDA;
        double* aq = A + q * LDA;
        double x[RPT], y[RPT];
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int i = sub + u * TPP;
          x[u] = i < KP ? ap[i] : 0.0;
          y[u] = i < KP ? aq[i] : 0.0;
        }
        double g = 0.0;
#pragma unroll
        for (int u = 0; u < RPT; ++u) g += x[u] * y[u];
#pragma unroll
        for (int s = 1; s < TPP; s <<= 1) g += shfl_xor_d(g, s);
        const double al = nrm[p], be = nrm[q];
        big |= g * g > tol * (al * be);
        if (g != 0.0 && fabs(g) > tol * sqrt(al * be)) {
          double c, s, t;
          jacobi_rotation(al, be, g, &c, &s, &t);
#pragma unroll
          for (int u = 0; u < RPT; ++u) {
            const int i = sub + u * TPP;
            if (i < KP) {
              ap[i] = c * x[u] - s * y[u];
              aq[i] = s * x[u] + c * y[u];
            }
          }
          if (sub == 0) {
            nrm[p] = al - t * g;
            nrm[q] = be + t * g;
          }
          rotated = true;
        }
      }
      __syncthreads();
    }
    if (rotated) flag[sweep & 1] = 1;
    if (big) flag[2 + (sweep & 1)] = 1;
    __syncthreads();
    if (flag[sweep & 1] == 0 || (predict && flag[2 + (sweep & 1)] == 0)) {
      *converged = true;
      ++sweep;
      break;
    }
    if (tid == 0) {
      flag[(sweep + 1) & 1] = 0;
      flag[2 + ((sweep + 1) & 1)] = 0;
    }
  }
  return sweep;
}

// ---------------------------------------------------------------------------
// One-sided Jacobi on a 128 x 128 block by 256 threads with the columns in
// registers (same result contract and stopping rule as jacobi_lds).
//
// Groups of eight threads (gp = tid / 8: 32 groups, 8 per wavefront) hold
// four column slots X[0..3] each (initially columns J * 32 + gp); thread sub
// of a group holds rows 8u + sub (u < 16) of its four columns.  Columns are
// kept as sc_j x_j with the scaled rotations of rotation_angle (two FMAs a
// row).  Each round rotates two independent pairs per group, (X0, X1) and
// (X2, X3) or (X0, X3) and (X2, X1), so renaming the slots changes partners
// for free.  Ordering (127 rounds a sweep; the fused kernel's order for two
// pairs per lane group, over 32 groups): for G = 64, 32, ..., 4, blocks of
// G rounds in steps of four, slots 1 and 3 moving once inside the step (from
// group gp ^ 1) and once between steps (from gp ^ lowbit(r + 4) / 2); then
// slots 1 and 3 of the first half of every block of G/4 groups swap with
// slots 0 and 2 of the partner group; three final rounds pair the four
// slots of each group.  Moves along group bits 0-2 stay in the wavefront
// (DPP and lane-swap VALU ops), the six a sweep along bits 3-4 go through
// LDS, which aliases A (the matrix is in registers meanwhile).  Every column returns to
// A at the index it came from, so zero padding columns stay at >= k.
struct JacobiXbuf {
  static constexpr int LD = 132;  // column stride (bank spread)
  double col[64 * LD];            // [2 gp + q][row]: two columns a group
  double nrm[64], sc[64], isc[64];
  int32_t id[64];
};

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// sum over the eight lanes of a group (quad_perm [1,0,3,2], [2,3,0,1], then
// row_half_mirror); every lane of the group gets the same bits
__device__ __forceinline__ double oct_sum(double v) {
  v = v + dpp_d<0xB1>(v);
  v = v + dpp_d<0x4E>(v);
  return v + dpp_d<0x141>(v);
}

// v of lane l ^ V (V = 8: DPP row_ror:8 inside each 16-lane row; V = 16, 32:
// v_permlane16/32_swap of v with itself, whose first result holds the value
// of the lower row (lanes l & ~V) and the second that of the upper one)
template <int V>
__device__ __forceinline__ unsigned lane_xor_u(unsigned v) {
  if constexpr (V == 8) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
  } else if constexpr (V == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (threadIdx.x & 16) ? r[0] : r[1];
  } else {
    static_assert(V == 32, "lane xor of 8, 16 or 32");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
  }
}
template <int V>
__device__ __forceinline__ double lane_xor(double v) {
  return __hiloint2double((int)lane_xor_u<V>((unsigned)__double2hiint(v)),
                          (int)lane_xor_u<V>((unsigned)__double2loint(v)));
}

struct Jacobi128 {
  static constexpr int NR = 16;  // rows per thread
  double X[4][NR];               // column slots of the group
  double nrm[4], sc[4], isc[4];  // squared norms, scales, inverse scales (between sweeps)
  int id[4];                     // column index in A
  // Within a sweep the state lives with the pair that is rotated: lanes 0-3
  // of a group hold the top tp and bottom bt of the round's first pair,
  // lanes 4-7 those of the second (row_half_mirror, lane i <-> 7 - i, trades
  // them between the halves).
  struct ColSt {
    double nrm, sc, isc;
  };
  ColSt tp, bt;

  template <int P, int Q>
  __device__ __forceinline__ double dot() const {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int u = 0; u < NR; u += 4) {
      a0 = fma(X[P][u], X[Q][u], a0);
      a1 = fma(X[P][u + 1], X[Q][u + 1], a1);
      a2 = fma(X[P][u + 2], X[Q][u + 2], a2);
      a3 = fma(X[P][u + 3], X[Q][u + 3], a3);
    }
    return (a0 + a1) + (a2 + a3);
  }
  template <int P, int Q>
  __device__ __forceinline__ void apply(double t, double gm, double cs, double ic) {
    const double ta = t * (sc[Q] * isc[P]), tb = t * (sc[P] * isc[Q]);
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double x = X[P][u], y = X[Q][u];
      X[P][u] = fma(-ta, y, x);
      X[Q][u] = fma(tb, x, y);
    }
    nrm[P] = fma(-t, gm, nrm[P]);
    nrm[Q] = fma(t, gm, nrm[Q]);
    sc[P] *= cs;
    sc[Q] *= cs;
    isc[P] *= ic;
    isc[Q] *= ic;
  }
  template <int P, int Q>
  __device__ __forceinline__ void rotate_x(double ta, double tb) {
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double x = X[P][u], y = X[Q][u];
      X[P][u] = fma(-ta, y, x);
      X[Q][u] = fma(tb, x, y);
    }
  }
  // Lanes 0-3 of a group take the angle of the first pair, lanes 4-7 that
  // of the second: the quad sums of both dot products are exchanged across
  // the halves (row_half_mirror pairs lane i with 7 - i), each half finishes
  // its own pair's sum and angle from its state, updates that state, and the
  // two FMA factors of each pair come back the same way.
  template <int T0, int B0, int T1, int B1>
  __device__ __forceinline__ void round(double tol2, double tolp, bool& rotated, bool& big) {
    double x = dot<T0, B0>(), y = dot<T1, B1>();
    x = x + dpp_d<0xB1>(x);
    y = y + dpp_d<0xB1>(y);
    x = x + dpp_d<0x4E>(x);
    y = y + dpp_d<0x4E>(y);
    const bool second = (threadIdx.x & 4) != 0;
    const double gs = (second ? y : x) + dpp_d<0x141>(second ? x : y);
    double t, gm, cs, ic;
    rotation_angle(gs, tp.nrm, bt.nrm, tp.sc, bt.sc, tol2, tolp, rotated, big, t, gm, cs, ic);
    const double ta = t * (bt.sc * tp.isc), tb = t * (tp.sc * bt.isc);
    tp.nrm = fma(-t, gm, tp.nrm);
    bt.nrm = fma(t, gm, bt.nrm);
    tp.sc *= cs;
    bt.sc *= cs;
    tp.isc *= ic;
    bt.isc *= ic;
    const double tao = dpp_d<0x141>(ta), tbo = dpp_d<0x141>(tb);
    rotate_x<T0, B0>(second ? tao : ta, second ? tbo : tb);
    rotate_x<T1, B1>(second ? ta : tao, second ? tb : tbo);
  }
  // the bottoms' state trades halves (pairing (0, 1), (2, 3) <-> (0, 3), (2, 1))
  __device__ __forceinline__ void swap_bt() {
    bt.nrm = dpp_d<0x141>(bt.nrm);
    bt.sc = dpp_d<0x141>(bt.sc);
    bt.isc = dpp_d<0x141>(bt.isc);
  }
  // (0, 3), (2, 1) -> (0, 2), (3, 1): lanes 0-3 keep their top (column 0) and
  // take the top of lanes 4-7 (column 2) as bottom; lanes 4-7 take the bottom
  // of lanes 0-3 (column 3) as top and keep their bottom (column 1)
  __device__ __forceinline__ void cross_tb() {
    const bool second = (threadIdx.x & 4) != 0;
    const ColSt mt{dpp_d<0x141>(tp.nrm), dpp_d<0x141>(tp.sc), dpp_d<0x141>(tp.isc)};
    const ColSt mb{dpp_d<0x141>(bt.nrm), dpp_d<0x141>(bt.sc), dpp_d<0x141>(bt.isc)};
    if (second) tp = mb;
    else bt = mt;
  }
  // the pair state at a sweep's start (lanes 0-3: slots 0, 1; 4-7: 2, 3)
  __device__ __forceinline__ void state_start() {
    const bool second = (threadIdx.x & 4) != 0;
    tp = ColSt{second ? nrm[2] : nrm[0], 1.0, 1.0};
    bt = ColSt{second ? nrm[3] : nrm[1], 1.0, 1.0};
  }
  // the scales of all four slots at a sweep's end (lanes 0-3 hold columns
  // 0 and 2, lanes 4-7 columns 3 and 1)
  __device__ __forceinline__ void state_end() {
    const bool second = (threadIdx.x & 4) != 0;
    const double mt = dpp_d<0x141>(tp.sc), mb = dpp_d<0x141>(bt.sc);
    sc[0] = second ? mt : tp.sc;
    sc[1] = second ? bt.sc : mb;
    sc[2] = second ? mb : bt.sc;
    sc[3] = second ? tp.sc : mt;
  }
  // slot J takes slot J of the group V / 8 groups away (lanes l ^ V)
  template <int J, int V>
  __device__ __forceinline__ void move_lane() {
#pragma unroll
    for (int u = 0; u < NR; ++u) X[J][u] = lane_xor<V>(X[J][u]);
    id[J] = (int)lane_xor_u<V>((unsigned)id[J]);
  }
  template <int V>
  __device__ __forceinline__ void move_bt_lane() {
    bt.nrm = lane_xor<V>(bt.nrm);
    bt.sc = lane_xor<V>(bt.sc);
    bt.isc = lane_xor<V>(bt.isc);
  }
  // the column of slot J (q: 0 or 1 of the group's two exchange places) to
  // the exchange area, and back from group src
  template <int J>
  __device__ __forceinline__ void put(JacobiXbuf& xb, int gp, int sub, int q) const {
#pragma unroll
    for (int u = 0; u < NR; ++u) xb.col[(2 * gp + q) * JacobiXbuf::LD + 8 * u + sub] = X[J][u];
    if (sub == 0) xb.id[2 * gp + q] = id[J];
  }
  template <int J>
  __device__ __forceinline__ void get(const JacobiXbuf& xb, int src, int sub, int q) {
#pragma unroll
    for (int u = 0; u < NR; ++u) X[J][u] = xb.col[(2 * src + q) * JacobiXbuf::LD + 8 * u + sub];
    id[J] = xb.id[2 * src + q];
  }
  // a half's state c to exchange place q of the group (lanes 0 and 4 write)
  __device__ __forceinline__ static void put_st(JacobiXbuf& xb, int gp, int sub, int q,
                                                const ColSt& c) {
    if ((sub & 3) == 0) {
      xb.nrm[2 * gp + q] = c.nrm;
      xb.sc[2 * gp + q] = c.sc;
      xb.isc[2 * gp + q] = c.isc;
    }
  }
  __device__ __forceinline__ static void get_st(const JacobiXbuf& xb, int src, int q, ColSt& c) {
    c.nrm = xb.nrm[2 * src + q];
    c.sc = xb.sc[2 * src + q];
    c.isc = xb.isc[2 * src + q];
  }
  // slots 1 and 3 from group gp ^ v with their state (v >= 8, through LDS:
  // in the pairing (0, 1), (2, 3), where the bottom state of lanes 0-3 is
  // slot 1's and of lanes 4-7 slot 3's)
  __device__ __forceinline__ void move_bottoms(JacobiXbuf& xb, int gp, int sub, int v) {
    if (v == 1) {
      move_lane<1, 8>();
      move_lane<3, 8>();
      move_bt_lane<8>();
    } else if (v == 2) {
      move_lane<1, 16>();
      move_lane<3, 16>();
      move_bt_lane<16>();
    } else if (v == 4) {
      move_lane<1, 32>();
      move_lane<3, 32>();
      move_bt_lane<32>();
    } else {
      const int qh = sub >> 2;  // the half's exchange place: slot 1 or 3
      put<1>(xb, gp, sub, 0);
      put<3>(xb, gp, sub, 1);
      put_st(xb, gp, sub, qh, bt);
      __syncthreads();
      get<1>(xb, gp ^ v, sub, 0);
      get<3>(xb, gp ^ v, sub, 1);
      get_st(xb, gp ^ v, qh, bt);
      __syncthreads();
    }
  }
  // first: slot B takes the partner's slot T; otherwise slot T takes the
  // partner's slot B (partner: lanes l ^ V); the pair's state alike
  template <int T, int B, int V>
  __device__ __forceinline__ void xswap_lane(bool first) {
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double w = lane_xor<V>(first ? X[B][u] : X[T][u]);
      X[B][u] = first ? w : X[B][u];
      X[T][u] = first ? X[T][u] : w;
    }
    const int wi = (int)lane_xor_u<V>((unsigned)(first ? id[B] : id[T]));
    id[B] = first ? wi : id[B];
    id[T] = first ? id[T] : wi;
  }
  template <int V>
  __device__ __forceinline__ void xswap_st_lane(bool first) {
    double w = lane_xor<V>(first ? bt.nrm : tp.nrm);
    bt.nrm = first ? w : bt.nrm;
    tp.nrm = first ? tp.nrm : w;
    w = lane_xor<V>(first ? bt.sc : tp.sc);
    bt.sc = first ? w : bt.sc;
    tp.sc = first ? tp.sc : w;
    w = lane_xor<V>(first ? bt.isc : tp.isc);
    bt.isc = first ? w : bt.isc;
    tp.isc = first ? tp.isc : w;
  }
  // (0, 1) and (2, 3) swapped with group gp ^ d, in the pairing (0, 1), (2, 3)
  __device__ __forceinline__ void xswap(JacobiXbuf& xb, int gp, int sub, int d) {
    const bool first = (gp & d) == 0;
    if (d == 1) {
      xswap_lane<0, 1, 8>(first);
      xswap_lane<2, 3, 8>(first);
      xswap_st_lane<8>(first);
    } else if (d == 2) {
      xswap_lane<0, 1, 16>(first);
      xswap_lane<2, 3, 16>(first);
      xswap_st_lane<16>(first);
    } else if (d == 4) {
      xswap_lane<0, 1, 32>(first);
      xswap_lane<2, 3, 32>(first);
      xswap_st_lane<32>(first);
    } else {
      const int qh = sub >> 2;  // the half's pair: (0, 1) or (2, 3)
      if (first) {
        put<1>(xb, gp, sub, 0);
        put<3>(xb, gp, sub, 1);
        put_st(xb, gp, sub, qh, bt);
      } else {
        put<0>(xb, gp, sub, 0);
        put<2>(xb, gp, sub, 1);
        put_st(xb, gp, sub, qh, tp);
      }
      __syncthreads();
      if (first) {
        get<1>(xb, gp ^ d, sub, 0);
        get<3>(xb, gp ^ d, sub, 1);
        get_st(xb, gp ^ d, qh, bt);
      } else {
        get<0>(xb, gp ^ d, sub, 0);
        get<2>(xb, gp ^ d, sub, 1);
        get_st(xb, gp ^ d, qh, tp);
      }
      __syncthreads();
    }
  }
  __device__ __forceinline__ void fold() {
#pragma unroll
    for (int J = 0; J < 4; ++J) {
#pragma unroll
      for (int u = 0; u < NR; ++u) X[J][u] *= sc[J];
      sc[J] = 1.0;
      isc[J] = 1.0;
    }
  }
};

template <int LDA, int NTHREADS>
__device__ int jacobi_regs128(double* A, int* flag, double tol, int max_sweeps, bool predict,
                              bool* converged) {
  static_assert(NTHREADS == 256, "32 groups of eight threads");
  static_assert(sizeof(JacobiXbuf) <= sizeof(double) * 128 * LDA, "exchange area fits in A");
  constexpr int NR = Jacobi128::NR;
  const int tid = threadIdx.x, lane = tid & 63, sub = tid & 7, gp = tid >> 3;
  JacobiXbuf& xb = *reinterpret_cast<JacobiXbuf*>(A);
  Jacobi128 j;
#pragma unroll
  for (int J = 0; J < 4; ++J) {
    j.id[J] = J * 32 + gp;
#pragma unroll
    for (int u = 0; u < NR; ++u) j.X[J][u] = A[8 * u + sub + j.id[J] * LDA];
    j.sc[J] = 1.0;
    j.isc[J] = 1.0;
  }
  if (tid == 0) {
    flag[0] = 0;
    flag[2] = 0;
  }
  __syncthreads();  // A is read before the exchange area is written
  const double tol2 = tol * tol;
  const double tolp = predict ? tol : __builtin_inf();
  int sweep = 0;
  *converged = false;
  for (; sweep < max_sweeps; ++sweep) {
    j.fold();
#pragma unroll
    for (int J = 0; J < 4; ++J) {
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int u = 0; u < NR; u += 2) {
        a0 = fma(j.X[J][u], j.X[J][u], a0);
        a1 = fma(j.X[J][u + 1], j.X[J][u + 1], a1);
      }
      j.nrm[J] = oct_sum(a0 + a1);
    }
    j.state_start();
    bool rotated = false, big = false;
    for (int G = 64; G >= 4; G >>= 1) {
      for (int r = 0; r < G; r += 4) {
        j.round<0, 1, 2, 3>(tol2, tolp, rotated, big);
        j.swap_bt();
        j.round<0, 3, 2, 1>(tol2, tolp, rotated, big);
        j.move_bottoms(xb, gp, sub, 1);  // lanes l ^ 8: the bottom state moves along
        j.round<0, 3, 2, 1>(tol2, tolp, rotated, big);
        j.swap_bt();
        j.round<0, 1, 2, 3>(tol2, tolp, rotated, big);
        if (r + 4 < G) j.move_bottoms(xb, gp, sub, ((r + 4) & -(r + 4)) >> 1);
      }
      j.xswap(xb, gp, sub, G >> 2);
    }
    j.round<0, 1, 2, 3>(tol2, tolp, rotated, big);
    j.swap_bt();
    j.round<0, 3, 2, 1>(tol2, tolp, rotated, big);
    j.cross_tb();
    j.round<0, 2, 3, 1>(tol2, tolp, rotated, big);
    j.state_end();
    const bool any_rot = __any(rotated), any_big = __any(big);
    if (lane == 0) {
      if (any_rot) flag[sweep & 1] = 1;
      if (any_big) flag[2 + (sweep & 1)] = 1;
    }
    __syncthreads();
    if (flag[sweep & 1] == 0 || (predict && flag[2 + (sweep & 1)] == 0)) {
      *converged = true;
      ++sweep;
      break;
    }
    if (tid == 0) {
      flag[(sweep + 1) & 1] = 0;
      flag[2 + ((sweep + 1) & 1)] = 0;
    }
  }
  j.fold();
  __syncthreads();  // the exchange area is no longer read
#pragma unroll
  for (int J = 0; J < 4; ++J)
#pragma unroll
    for (int u = 0; u < NR; ++u) A[8 * u + sub + j.id[J] * LDA] = j.X[J][u];
  __syncthreads();
  return sweep;
}

}  // namespace dida
