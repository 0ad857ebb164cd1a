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

}  // namespace dida
