// One-sided (Hestenes) Jacobi on a k x m block resident in LDS.
//
// The m = k rounded up to even columns of A (column-major, leading dimension
// LDA, rows 0..k-1 used) are orthogonalised pairwise; in each of the m - 1
// rounds of a round-robin (circle) ordering the m/2 disjoint pairs are rotated
// in parallel, tpp threads per pair.  A pair (p, q) is rotated while
// |a_p . a_q| > tol |a_p| |a_q|; column norms are recomputed at the start of
// each sweep and updated exactly in between (alpha' = alpha - t g,
// beta' = beta + t g).  A sweep without any rotation ends the iteration.
// For a symmetric positive (semi)definite input the converged columns are
// lambda_j v_j (Hestenes 1958; Demmel & Veselic 1992).
#pragma once

#include "letkf_device.h"

namespace dida {

template <int LDA, int NTHREADS>
__device__ int jacobi_lds(double* A, double* nrm, int* flag, int k, double tol, int max_sweeps,
                          bool* converged) {
  const int m = (k + 1) & ~1;
  const int npair = m >> 1;
  int tpp = 1;
  while (tpp * 2 * npair <= NTHREADS && tpp < 32) tpp <<= 1;
  const int tid = threadIdx.x;
  const int pair = tid / tpp, sub = tid % tpp;
  const bool active = pair < npair;
  if (tid == 0) flag[0] = 0;
  int sweep = 0;
  *converged = false;
  for (; sweep < max_sweeps; ++sweep) {
    for (int j = tid; j < m; j += NTHREADS) {
      double s = 0.0;
      for (int i = 0; i < k; ++i) s += A[i + j * LDA] * A[i + j * LDA];
      nrm[j] = s;
    }
    __syncthreads();
    bool rotated = false;
    for (int r = 0; r < m - 1; ++r) {
      if (active) {
        int p, q;
        if (pair == 0) {
          p = m - 1;
          q = r;
        } else {
          p = (r - pair + (m - 1)) % (m - 1);
          q = (r + pair) % (m - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double* ap = A + p * LDA;
        double* aq = A + q * LDA;
        double g = 0.0;
        for (int i = sub; i < k; i += tpp) g += ap[i] * aq[i];
        for (int s = 1; s < tpp; s <<= 1) g += shfl_xor_d(g, s);
        const double al = nrm[p], be = nrm[q];
        if (g != 0.0 && fabs(g) > tol * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * g);
          double t;
          if (fabs(zeta) > 1e100) t = 0.5 / zeta;
          else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t);
          const double s = c * t;
          for (int i = sub; i < k; i += tpp) {
            const double x = ap[i], y = aq[i];
            ap[i] = c * x - s * y;
            aq[i] = s * x + c * y;
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
    __syncthreads();
    if (flag[sweep & 1] == 0) {
      *converged = true;
      ++sweep;
      break;
    }
    if (tid == 0) flag[(sweep + 1) & 1] = 0;
  }
  return sweep;
}

}  // namespace dida
