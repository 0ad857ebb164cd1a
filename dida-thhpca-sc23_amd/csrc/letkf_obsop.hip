// Observation operator H(Xb) and centring on the GPU (the reference's
// letkf_obs.F90:120-178 with find_obs_grid :187-243, obs_interpolation_3d_2
// and ps2ph_3d_interp, obs_make.F90:233-381, and the half-level pressures of
// vert_coord_mod.F90:547-563).
//
// One wavefront per observation record, lanes over ensemble members (k <=
// 128, two members a lane past 64).  Per member: bilinear surface pressure,
// the level search over full-level pressures computed on the fly (carried
// over the record's levels in order, as the reference's loop does), log-p
// weights between the bracketing levels (the nearest level outside them),
// bilinear horizontal interpolation of member = mean + perturbation.  State
// loads are member-contiguous (coalesced across lanes); the ensemble mean of
// each scalar obs is a wave reduction.  Outputs are the scalar obs of the
// analysis in (record, level, variable) order.
#include "letkf_device.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

// First index i of ascending g[0..n) with g[i] >= x (n if none).
__device__ __forceinline__ int first_ge(const float* g, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((double)g[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// find_obs_grid for one axis: the bracketing indices and the weight of the
// first (REAL(4) grid differences formed in single precision).
__device__ __forceinline__ double bracket(const float* g, int n, bool periodic, double x,
                                          int* i0, int* i1) {
#pragma clang fp contract(off)
  const int j = first_ge(g, n, x);
  if (j == 0) {
    *i0 = *i1 = 0;
    return 0.0;
  }
  if (j == n) {
    if (!periodic) {
      *i0 = *i1 = n - 1;
      return 1.0;
    }
    const float gw = g[0] + 360.0f;
    if (!((double)gw >= x)) {
      *i0 = *i1 = 0;
      return 0.0;
    }
    const double d = (double)(float)(gw - g[n - 1]);
    *i0 = n - 1;
    *i1 = 0;
    return (d - x + (double)g[n - 1]) / d;
  }
  const double d = (double)(float)(g[j] - g[j - 1]);
  *i0 = j - 1;
  *i1 = j;
  return (d - x + (double)g[j - 1]) / d;
}

// half-level pressure j (0-based, top first) over surface pressure ps
// (vert_coord_mod.F90:547-563); A holds vtype, c1, c2, ptop, p0
template <class A>
__device__ __forceinline__ double ph_half(const A& a, int j, double ps) {
#pragma clang fp contract(off)
  if (a.vtype == 0) return a.c1[j] * (ps - a.ptop) + a.ptop;
  return a.c1[j] * (a.p0 - a.ptop) + a.c2[j] * (ps - a.ptop) + a.ptop;
}

// full-level pressure j (0-based) of a member with surface pressure ps
template <class A>
__device__ __forceinline__ double ph_full(const A& a, int j, double ps) {
#pragma clang fp contract(off)
  return 0.5 * (ph_half(a, j, ps) + ph_half(a, j + 1, ps));
}

__global__ __launch_bounds__(256) void k_obs_operator(dk_obsop_args a) {
#pragma clang fp contract(off)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + wave;
  if (r >= a.nrec) return;
  const int k = a.k, nrow = a.nrow, nlat = a.nlat, nz = a.nz, onz = a.onz, onv = a.onv;
  const double rlat = a.rec_lat[r], rlon = a.rec_lon[r];
  int a0, a1, b0, b1;
  const double t1 = bracket(a.grid_lat, nlat, false, rlat, &a0, &a1);
  const double t2 = bracket(a.grid_lon, a.nlon, a.periodic != 0, rlon, &b0, &b1);
  const int64_t c00 = (int64_t)a0 + (int64_t)nlat * b0, c10 = (int64_t)a1 + (int64_t)nlat * b0;
  const int64_t c01 = (int64_t)a0 + (int64_t)nlat * b1, c11 = (int64_t)a1 + (int64_t)nlat * b1;
  const int64_t s0 = r * onz * onv;
  for (int q = lane; q < onz * onv; q += 64) {
    a.lat[s0 + q] = rlat;
    a.lon[s0 + q] = rlon;
    if (a.ph) a.ph[s0 + q] = a.rec_ph[r * onz + q / onv];
  }
  // member e of row `row` at a corner column, and the bilinear value
  auto X = [&](int64_t col, int row, int e) {
    return a.xb[(col * nrow + row) * k + e] + a.axb[col * nrow + row];
  };
  auto horiz = [&](int row, int e) {
    const double r1 = X(c00, row, e) * t1 + X(c10, row, e) * (1 - t1);
    const double r2 = X(c01, row, e) * t1 + X(c11, row, e) * (1 - t1);
    return r1 * t2 + r2 * (1 - t2);
  };
  constexpr int kMaxM = 2;  // members per lane (k <= 128)
  double ps[kMaxM];
  int jpos[kMaxM];
#pragma unroll
  for (int m = 0; m < kMaxM; ++m) {
    const int e = lane + 64 * m;
    ps[m] = e < k ? horiz(a.ps_row, e) : 0.0;
    jpos[m] = 0;
  }
  for (int l = 0; l < onz; ++l) {
    const double po = a.rec_ph[r * onz + l];
    int pos[kMaxM];
#pragma unroll
    for (int m = 0; m < kMaxM; ++m) {
      const int e = lane + 64 * m;
      if (e < k)
        while (jpos[m] < nz && po > ph_full(a, jpos[m], ps[m])) ++jpos[m];
      pos[m] = jpos[m];
    }
    for (int v = 0; v < onv; ++v) {
      const int var = a.vid[v];
      const int vs = a.vstart[var];
      double h[kMaxM];
      double sum = 0.0;
#pragma unroll
      for (int m = 0; m < kMaxM; ++m) {
        const int e = lane + 64 * m;
        h[m] = 0.0;
        if (e >= k) continue;
        if (a.vlev[var] > 1) {
          const int p = pos[m];
          if (p >= 1 && p < nz) {
            const int tl = vs + p - 1;
            const double pu = ph_full(a, p - 1, ps[m]), pl = ph_full(a, p, ps[m]);
            const double t3 = (log(pl) - log(po)) / (log(pl) - log(pu));
            h[m] = horiz(tl, e) * t3 + horiz(tl + 1, e) * (1 - t3);
          } else {
            h[m] = horiz(p == 0 ? vs : vs + nz - 1, e);
          }
        } else {
          h[m] = horiz(vs, e);
        }
        sum += h[m];
      }
      for (int sh = 32; sh >= 1; sh >>= 1) sum += shfl_xor_d(sum, sh);
      const double mean = sum / k;
      const int64_t s = s0 + (int64_t)l * onv + v;
#pragma unroll
      for (int m = 0; m < kMaxM; ++m) {
        const int e = lane + 64 * m;
        if (e < k) a.yb[s * k + e] = h[m] - mean;
      }
      if (lane == 0) a.dep[s] = a.yo[s] - mean;
    }
  }
}

// ps2alnp (obs_make.F90:418-437) at the centre level of every cube: one
// thread per (column, cube).  The reference reads the bottom half level
// ph_lev(nz + 1) without setting it; here it is the surface pressure of the
// coordinate table, as for every other half level.
__global__ void k_cube_lnp(dk_cube_lnp_args a) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ncol * a.ncube) return;
  const int64_t an = t / a.ncube;
  const int q = (int)(t - an * a.ncube);
  const int sk = q * a.cube_nz;
  const int ek = min(sk + a.cube_nz - 1, a.nz - 1);
  const int ck = (sk + 1 + ek + 1) / 2 - 1;
  int64_t col = an;
  if (a.cube_n > 1) {
    const int aj = (int)(an % a.nalat), ai = (int)(an / a.nalat);
    const int sj = aj * a.cube_n, si = ai * a.cube_n;
    const int ej = min(sj + a.cube_n - 1, a.grid_nlat - 1);
    const int ei = min(si + a.cube_n - 1, a.grid_nlon - 1);
    col = (int64_t)((sj + ej) / 2) + (int64_t)a.grid_nlat * ((si + ei) / 2);
  }
  const double ps = a.axb[col * a.nrow + a.ps_row];
  a.cube_lnp[t] = log(ph_full(a, ck, ps));
}

__global__ void k_obs_lnp(int32_t nobs, const int32_t* __restrict__ perm,
                          const double* __restrict__ ph, double* __restrict__ lnp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nobs) lnp[i] = log(ph[perm[i]]);
}

}  // namespace

extern "C" int dk_cube_lnp(const dk_cube_lnp_args* a, void* stream) {
  const int64_t n = a->ncol * a->ncube;
  if (n <= 0) return 0;
  k_cube_lnp<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(*a);
  return (int)hipGetLastError();
}

extern "C" int dk_obs_lnp(int32_t nobs, const int32_t* perm, const double* ph, double* lnp,
                          void* stream) {
  if (nobs <= 0) return 0;
  k_obs_lnp<<<(nobs + 255) / 256, 256, 0, (hipStream_t)stream>>>(nobs, perm, ph, lnp);
  return (int)hipGetLastError();
}

extern "C" int dk_obs_operator(const dk_obsop_args* a, void* stream) {
  if (a->nrec <= 0) return 0;
  if (a->k < 1 || a->k > 128) return (int)hipErrorInvalidValue;
  k_obs_operator<<<(unsigned)((a->nrec + 3) / 4), 256, 0, (hipStream_t)stream>>>(*a);
  return (int)hipGetLastError();
}
