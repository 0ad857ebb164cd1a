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
// Pair slot s = tid / 4 (64 slots, 16 per wavefront) holds a top and a
// bottom column; the four threads of a slot hold rows 4u + sub (u < 32) of
// both.  Column j is kept as sc_j x_j with the scaled rotations of
// rotation_angle (two FMAs a row).  Ordering (KP - 1 = 127 rounds a sweep):
// for G = 64, 32, ..., 1, each block of G slots runs G rounds in which the
// top of slot s meets the bottom that started the level in slot
// s ^ gray(r); between rounds the bottoms move by the one bit in which
// gray(r) and gray(r + 1) differ; then the bottom of every slot in the
// first half of its block swaps with the top of slot s + G/2.  Moves along
// slot bits 0-3 stay inside a wavefront (ds_bpermute); the six a sweep
// along bits 4 and 5 go through LDS, which aliases A (the matrix is in
// registers meanwhile).  Every column returns to A at the index it came
// from, so zero padding columns stay at >= k.
struct JacobiXbuf {
  static constexpr int LD = 132;  // slot stride: 8 slots of 4 rows cover the banks
  double col[64 * LD];
  double nrm[64], sc[64], isc[64];
  int32_t id[64];
};

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// sum over the four lanes of a quad (quad_perm [1,0,3,2], then [2,3,0,1]);
// every lane of the quad gets the same bits
__device__ __forceinline__ double quad_sum(double v) {
  v = v + dpp_d<0xB1>(v);
  return v + dpp_d<0x4E>(v);
}

struct Jacobi128 {
  static constexpr int NR = 32;  // rows per thread
  double x[NR], y[NR];           // top and bottom column of the slot
  double nt, nb, st, sb, it, ib; // squared norms, scales, inverse scales
  int idt, idb;                  // their column index in A

  __device__ __forceinline__ void round(double tol2, double tolp, bool& rotated, bool& big) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int u = 0; u < NR; u += 4) {
      a0 = fma(x[u], y[u], a0);
      a1 = fma(x[u + 1], y[u + 1], a1);
      a2 = fma(x[u + 2], y[u + 2], a2);
      a3 = fma(x[u + 3], y[u + 3], a3);
    }
    const double gs = quad_sum((a0 + a1) + (a2 + a3));
    double t, gm, cs, ic;
    rotation_angle(gs, nt, nb, st, sb, tol2, tolp, rotated, big, t, gm, cs, ic);
    const double ta = t * (sb * it), tb = t * (st * ib);
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double xx = x[u], yy = y[u];
      x[u] = fma(-ta, yy, xx);
      y[u] = fma(tb, xx, yy);
    }
    nt = fma(-t, gm, nt);
    nb = fma(t, gm, nb);
    st *= cs;
    sb *= cs;
    it *= ic;
    ib *= ic;
  }
  // the bottom of the slot whose lanes sit at byte address addr
  __device__ __forceinline__ void move_lane(int addr) {
#pragma unroll
    for (int u = 0; u < NR; ++u) y[u] = bperm_d(addr, y[u]);
    nb = bperm_d(addr, nb);
    sb = bperm_d(addr, sb);
    ib = bperm_d(addr, ib);
    idb = __builtin_amdgcn_ds_bpermute(addr, idb);
  }
  // the bottom of slot src, through LDS
  __device__ __forceinline__ void move_lds(JacobiXbuf& xb, int s, int sub, int src) {
#pragma unroll
    for (int u = 0; u < NR; ++u) xb.col[s * JacobiXbuf::LD + 4 * u + sub] = y[u];
    if (sub == 0) {
      xb.nrm[s] = nb;
      xb.sc[s] = sb;
      xb.isc[s] = ib;
      xb.id[s] = idb;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NR; ++u) y[u] = xb.col[src * JacobiXbuf::LD + 4 * u + sub];
    nb = xb.nrm[src];
    sb = xb.sc[src];
    ib = xb.isc[src];
    idb = xb.id[src];
    __syncthreads();
  }
  // first: this slot's bottom takes the partner's top; otherwise this
  // slot's top takes the partner's bottom
  __device__ __forceinline__ void xswap_lane(int addr, bool first) {
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double w = bperm_d(addr, first ? y[u] : x[u]);
      y[u] = first ? w : y[u];
      x[u] = first ? x[u] : w;
    }
    double w = bperm_d(addr, first ? nb : nt);
    nb = first ? w : nb;
    nt = first ? nt : w;
    w = bperm_d(addr, first ? sb : st);
    sb = first ? w : sb;
    st = first ? st : w;
    w = bperm_d(addr, first ? ib : it);
    ib = first ? w : ib;
    it = first ? it : w;
    const int wi = __builtin_amdgcn_ds_bpermute(addr, first ? idb : idt);
    idb = first ? wi : idb;
    idt = first ? idt : wi;
  }
  __device__ __forceinline__ void xswap_lds(JacobiXbuf& xb, int s, int sub, int src,
                                            bool first) {
#pragma unroll
    for (int u = 0; u < NR; ++u) xb.col[s * JacobiXbuf::LD + 4 * u + sub] = first ? y[u] : x[u];
    if (sub == 0) {
      xb.nrm[s] = first ? nb : nt;
      xb.sc[s] = first ? sb : st;
      xb.isc[s] = first ? ib : it;
      xb.id[s] = first ? idb : idt;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const double w = xb.col[src * JacobiXbuf::LD + 4 * u + sub];
      y[u] = first ? w : y[u];
      x[u] = first ? x[u] : w;
    }
    const double wn = xb.nrm[src], ws = xb.sc[src], wi = xb.isc[src];
    const int wd = xb.id[src];
    if (first) {
      nb = wn;
      sb = ws;
      ib = wi;
      idb = wd;
    } else {
      nt = wn;
      st = ws;
      it = wi;
      idt = wd;
    }
    __syncthreads();
  }
  __device__ __forceinline__ void fold() {
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      x[u] *= st;
      y[u] *= sb;
    }
    st = sb = it = ib = 1.0;
  }
};

template <int LDA, int NTHREADS>
__device__ int jacobi_regs128(double* A, int* flag, double tol, int max_sweeps, bool predict,
                              bool* converged) {
  static_assert(NTHREADS == 256, "64 pair slots of four threads");
  static_assert(sizeof(JacobiXbuf) <= sizeof(double) * 128 * LDA, "exchange area fits in A");
  constexpr int NR = Jacobi128::NR;
  const int tid = threadIdx.x, lane = tid & 63, sub = tid & 3, s = tid >> 2;
  JacobiXbuf& xb = *reinterpret_cast<JacobiXbuf*>(A);
  Jacobi128 j;
  j.idt = 2 * s;
  j.idb = 2 * s + 1;
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    j.x[u] = A[4 * u + sub + j.idt * LDA];
    j.y[u] = A[4 * u + sub + j.idb * LDA];
  }
  j.st = j.sb = j.it = j.ib = 1.0;
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
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      a0 = fma(j.x[u], j.x[u], a0);
      a1 = fma(j.y[u], j.y[u], a1);
    }
    j.nt = quad_sum(a0);
    j.nb = quad_sum(a1);
    bool rotated = false, big = false;
    for (int G = 64; G >= 1; G >>= 1) {
      for (int r = 0; r < G; ++r) {
        j.round(tol2, tolp, rotated, big);
        if (r + 1 < G) {
          const int m = (r + 1) & -(r + 1);
          if (m < 16) j.move_lane((lane ^ (4 * m)) << 2);
          else j.move_lds(xb, s, sub, s ^ m);
        }
      }
      if (G > 1) {
        const int d = G >> 1;
        const bool first = (s & d) == 0;
        if (d < 16) j.xswap_lane((lane ^ (4 * d)) << 2, first);
        else j.xswap_lds(xb, s, sub, s ^ d, first);
      }
    }
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
  for (int u = 0; u < NR; ++u) {
    A[4 * u + sub + j.idt * LDA] = j.x[u];
    A[4 * u + sub + j.idb * LDA] = j.y[u];
  }
  __syncthreads();
  return sweep;
}

}  // namespace dida
