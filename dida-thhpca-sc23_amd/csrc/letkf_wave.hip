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
#include <cstring>
#include <type_traits>

#include "letkf_device.h"
#include "letkf_kernels.h"

using namespace dida;

namespace {

// FULLV: the whole eigenvector basis V (the warm start of the next analysis
// of a chain, column slot j at V[i + j LDV]); otherwise 16 eigenvectors at a
// time (column slot j at V[(j % 16) LDV + i]), for cold starts only, so that
// two wavefronts fit a SIMD.
template <int KP, bool FULLV>
struct WSmem {
  static constexpr int LDV = KP + 1;
  double V[(FULLV ? KP : 16) * LDV];
  double b[KP];
  double f[KP];
  double u[KP];
  int32_t ogi[2][64];  // staged obs blocks (two buffers): binned index,
  double orr[2][64];   // sqrt(R^-1 rloc),
  double ow[2][64];    // R^-1 rloc dep
  double ct[16 * 17];  // a 16 x 16 tile being transposed (Gram, W), or wbar
};

__device__ __forceinline__ int rowA(int I, int r, int g) { return I * 16 + 4 * g + r; }
__device__ __forceinline__ int rowB(int I, int r, int g) { return I * 16 + 4 * r + g; }
__device__ __forceinline__ int row_of(bool patB, int I, int r, int g) {
  return patB ? rowB(I, r, g) : rowA(I, r, g);
}

// v + v(lane ^ 32) and v + v(lane ^ 16) on the VALU (gfx950 row swaps:
// v_permlane32_swap exchanges lanes 32-63 of one operand with lanes 0-31 of
// the other, v_permlane16_swap odd 16-lane rows with even ones; with both
// operands equal to v the two results hold v and its partner, in either
// order, so their sum is the butterfly).
__device__ __forceinline__ double add_xor32(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}
__device__ __forceinline__ double add_xor16(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}
// sum over the four lane groups g (lanes c, c + 16, c + 32, c + 48)
__device__ __forceinline__ double sum_g(double v) { return add_xor16(add_xor32(v)); }

// Sums of x and y over the four lane groups, transposed: lanes of groups 0
// and 1 get sum_g(x), lanes of groups 2 and 3 get sum_g(y).  The row swap
// of x with y leaves x's groups 0, 1 and y's groups 0, 1 in one register and
// x's groups 2, 3 and y's groups 2, 3 in the other (in either order), so one
// add forms the partials; a row-pair butterfly finishes them.
__device__ __forceinline__ double sum_g_pair(double x, double y) {
  const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false,
                                                  false);
  const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false,
                                                  false);
  const double s = __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  return add_xor16(s);
}

// v0 = v of lane groups 0 and 1, v1 = v of lane groups 2 and 3, in every lane
// (for v equal within each of those pairs of groups): v_permlane32_swap of v
// with itself returns lanes 0-31 of v twice in its first result and lanes
// 32-63 twice in its second.
__device__ __forceinline__ void split_g(double v, double& v0, double& v1) {
  const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false,
                                                  false);
  const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false,
                                                  false);
  v0 = __hiloint2double(b[0], a[0]);
  v1 = __hiloint2double(b[1], a[1]);
}

// v of lane (l ^ 1), same 16-lane row (ds_swizzle, bit mode: and 0x1f, xor 1)
__device__ __forceinline__ double swz_xor1(double v) {
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), 0x1F | (1 << 10));
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), 0x1F | (1 << 10));
  return __hiloint2double(hi, lo);
}
// Metadata of one 64-obs block of an analysis, lane l holding obs base + l:
// binned index, sqrt(R^-1 rloc) and R^-1 rloc dep.  Entries past p take the
// index of obs 0 and zero weights, so every load is unconditional and in
// bounds (callers only use blocks with base < p, p > 0).  The loads are split
// in two so the dependent ones can be issued once the indices have arrived.
constexpr int kModeColumn = 0;
constexpr int kModeVertical = 1;
constexpr int kModeTransOut = 2;

struct ObsMeta {
  int gi;
  bool ok;
  double rl, ri, dp, lp;
  __device__ __forceinline__ void load_index(const dk_batched_args& a, int64_t off, int p,
                                             int base) {
    const int o = base + (int)threadIdx.x;
    ok = o < p;
    const int oc = ok ? o : 0;
    gi = a.idx[off + oc];
    rl = a.rloc[off + oc];
  }
  template <bool VERT>
  __device__ __forceinline__ void load_obs(const dk_batched_args& a) {
    ri = (double)a.rinv[gi];
    dp = a.dep[gi];
    if constexpr (VERT) lp = a.obs_lnp[gi];
  }
  template <class SM>
  __device__ __forceinline__ void store_index(SM& sm, int buf) const {
    sm.ogi[buf][threadIdx.x] = gi;
  }
  // sqrt(R^-1 rloc) and R^-1 rloc d; with vertical localisation rloc is
  // scaled by the weight of |ln p_cube - ln p_obs| / vert_dist
  // (letkf_yb.F90:223-243)
  template <bool VERT, class SM>
  __device__ __forceinline__ void store_weights(SM& sm, int buf, const dk_batched_args& a,
                                                double clnp) const {
    const int l = threadIdx.x;
    double rv = rl;
    if constexpr (VERT) rv = rl * gc_weight(fabs(clnp - lp) / a.vert_dist, a.gc_vert);
    const double r2 = ok ? ri * rv : 0.0;
    const double rr = sqrt(r2);
    sm.orr[buf][l] = rr;
    sm.ow[buf][l] = rr * (rr * dp);
  }
  // The same for the split half-precision Grams, in FP32 and with the
  // analysis's power-of-two scale applied to sqrt(R^-1 rloc), in the first
  // half of the same LDS arrays (read as FP32 only by those Grams).
  template <bool VERT, class SM>
  __device__ __forceinline__ void store_weights32(SM& sm, int buf, const dk_batched_args& a,
                                                  double clnp, double scale) const {
    const int l = threadIdx.x;
    double rv = rl;
    if constexpr (VERT) rv = rl * gc_weight(fabs(clnp - lp) / a.vert_dist, a.gc_vert);
    const double r2 = ok ? ri * rv : 0.0;
    const double rr = sqrt(r2);
    reinterpret_cast<float*>(sm.orr[buf])[l] = (float)(rr * scale);
    reinterpret_cast<float*>(sm.ow[buf])[l] = (float)(rr * (rr * dp));
  }
};

template <int KP, int GRAM>
struct Wave {
  static constexpr int NJ = KP / 16;   // column sets per lane (and row blocks)
  static constexpr int NQ = NJ * 4;    // rows per column per lane
  static constexpr int NP = NJ / 2;    // column pairs per lane group
  // Newton steps of the rotation cosine: one for the split half-precision
  // Grams (about 2^-44, far below their own error), two for the FP64 Gram
  static constexpr int kNewton = GRAM == DK_GRAM_FP64 ? 2 : 1;
  double X[NJ][NQ];  // working matrix (see header)
  bool patB;         // current row pattern
  // Yb rows as read: FP64 for the FP64 Gram, FP32 for the half-precision ones
  typedef typename std::conditional<GRAM == DK_GRAM_FP64, double, float>::type RawT;

  // obs e (< 8) of half h of a 64-obs block in the Gram MFMA operand layout:
  // FP64 16x16x4 k-steps of 4 obs (obs 4e + g), half 16x16x32 (obs 8g + e)
  static __device__ __forceinline__ int half_obs(int h, int e, int g) {
    return h * 32 + (GRAM == DK_GRAM_FP64 ? 4 * e + g : 8 * g + e);
  }

  // Yb rows I*16 + c of the 8 obs of half h in buffer buf (binned rows are
  // KP values apart and zero past k, so every row is in bounds).  Buffer
  // loads with a 32-bit byte offset a lane and obs, the I offsets in the
  // instruction (the host keeps the rows below 4 GiB): one add per obs
  // instead of 64-bit address arithmetic.
  static constexpr int kRowShift = (KP == 64 ? 6 : 5) + (sizeof(RawT) == 8 ? 3 : 2);
  template <class SM>
  __device__ __forceinline__ void load_rows(const dk_batched_args& a, const SM& sm,
                                            int buf, int h, RawT (&raw)[8][NJ]) const {
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const void* yb;
    if constexpr (GRAM == DK_GRAM_FP64) yb = a.yb;
    else yb = a.yb32;
    const __amdgpu_buffer_rsrc_t rows =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(yb), (short)0, 0xFFFFFFFF, 0x00020000);
    const uint32_t lane_off = (uint32_t)c * (uint32_t)sizeof(RawT);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint32_t o = ((uint32_t)sm.ogi[buf][half_obs(h, e, g)] << kRowShift) + lane_off;
#pragma unroll
      for (int I = 0; I < NJ; ++I) {
        if constexpr (sizeof(RawT) == 8) {
          raw[e][I] = __longlong_as_double((long long)__builtin_amdgcn_raw_buffer_load_b64(
              rows, o + (uint32_t)(I * 16 * 8), 0, 0));
        } else {
          raw[e][I] = __int_as_float((int)__builtin_amdgcn_raw_buffer_load_b32(
              rows, o + (uint32_t)(I * 16 * 4), 0, 0));
        }
      }
    }
  }

  // Accumulate half h of the staged block: C on MFMA, b in FP64
  template <typename AccD, typename AccH, class SM>
  __device__ __forceinline__ void gram_half(const SM& sm, int buf, int h,
                                            const RawT (&raw)[8][NJ], double scale,
                                            double (&bacc)[NJ], AccD& accd, AccH& acch) const {
    const int g = threadIdx.x >> 4;
    if constexpr (GRAM == DK_GRAM_FP64) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int ol = half_obs(h, e, g);
        const double rr = sm.orr[buf][ol], w = sm.ow[buf][ol];
        double f[NJ];
#pragma unroll
        for (int I = 0; I < NJ; ++I) {
          f[I] = raw[e][I] * rr;
          bacc[I] = fma(raw[e][I], w, bacc[I]);
        }
#pragma unroll
        for (int I = 0; I < NJ; ++I)
#pragma unroll
          for (int J = 0; J < NJ; ++J) accd[I][J] = mfma_f64_16x16x4(f[I], f[J], accd[I][J]);
      }
    } else {
      typedef typename H16<GRAM>::vec hvec;
      typedef typename H16<GRAM>::vec2 hvec2;
      constexpr bool kSplit = GRAM != DK_GRAM_BF16;
      hvec hi[NJ], lo[NJ];
      // b: the half's 8 products a row summed in FP32, then added in FP64
      // (relative error ~1e-7, against the split Gram's ~1e-6)
      float bh[NJ];
#pragma unroll
      for (int I = 0; I < NJ; ++I) bh[I] = 0.0f;
      // two obs at a time, so one instruction converts both elements of a pair
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const int o0 = half_obs(h, e, g), o1 = half_obs(h, e + 1, g);
        const float* orr = reinterpret_cast<const float*>(sm.orr[buf]);
        const float* ow = reinterpret_cast<const float*>(sm.ow[buf]);
        const f32x2_t rs = {orr[o0], orr[o1]};
        const float w0 = ow[o0], w1 = ow[o1];
#pragma unroll
        for (int I = 0; I < NJ; ++I) {
          bh[I] = fmaf(raw[e][I], w0, bh[I]);
          bh[I] = fmaf(raw[e + 1][I], w1, bh[I]);
          const f32x2_t xs = f32x2_t{raw[e][I], raw[e + 1][I]} * rs;
          const hvec2 hv = __builtin_convertvector(xs, hvec2);
          hi[I][e] = hv[0];
          hi[I][e + 1] = hv[1];
          if (kSplit) {
            const hvec2 lv = __builtin_convertvector(xs - __builtin_convertvector(hv, f32x2_t),
                                                     hvec2);
            lo[I][e] = lv[0];
            lo[I][e + 1] = lv[1];
          }
        }
      }
#pragma unroll
      for (int I = 0; I < NJ; ++I) bacc[I] += (double)bh[I];
      // C is symmetric: the tiles I >= J only
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          acch[I][J] = H16<GRAM>::mfma(hi[I], hi[J], acch[I][J]);
          if (kSplit) {
            acch[I][J] = H16<GRAM>::mfma(hi[I], lo[J], acch[I][J]);
            acch[I][J] = H16<GRAM>::mfma(lo[I], hi[J], acch[I][J]);
          }
        }
    }
  }

  // ---- Gram ---------------------------------------------------------------
  // C = Yhat Yhat^T and b = Yhat^T (sqrt(R^-1 rloc) d) over 64-obs blocks of
  // two 32-obs halves.  Rows past k read row k - 1 (in bounds) and are masked
  // once at the end (a row of Yhat only reaches the row and column of C it
  // indexes).
  template <bool VERT, class SM>
  __device__ void gram(const dk_batched_args& a, int64_t off, int p, double clnp, SM& sm) {
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const int k = a.k;
    constexpr bool kF64 = GRAM == DK_GRAM_FP64;
    double bacc[NJ];
#pragma unroll
    for (int I = 0; I < NJ; ++I) bacc[I] = 0.0;
    double scale = 1.0, inv2 = 1.0;
    if (GRAM == DK_GRAM_FP16X3) {
      // power-of-two scale so the largest |Yhat| entry is ~2^11 in FP16
      double mx = 0.0;
      for (int o = lane; o < p; o += 64) {
        const int gi = a.idx[off + o];
        double rv = a.rloc[off + o];
        if constexpr (VERT)
          rv *= gc_weight(fabs(clnp - a.obs_lnp[gi]) / a.vert_dist, a.gc_vert);
        mx = fmax(mx, a.ymax[gi] * sqrt(rv));
      }
      for (int s = 32; s >= 1; s >>= 1) mx = fmax(mx, __shfl_xor(mx, s, kWave));
      if (mx > 0.0) {
        int e;
        frexp(mx, &e);
        scale = ldexp(1.0, 11 - e);
        inv2 = ldexp(1.0, 2 * (e - 11));
      }
    }
    double4_t accd[kF64 ? NJ : 1][kF64 ? NJ : 1];
    f32x4_t acch[kF64 ? 1 : NJ][kF64 ? 1 : NJ];
    if constexpr (kF64) {
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J) accd[I][J] = double4_t{0.0, 0.0, 0.0, 0.0};
    } else {
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J) acch[I][J] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    }
    if (p > 0 && !kF64) {
      // half-precision Grams (FP32 rows, half the registers of FP64 ones):
      // the rows of block b + 1 are issued while block b computes; the
      // metadata of block b + 3 is loaded during block b and staged in the
      // LDS buffer of its parity at the end of block b + 1, so its gathers
      // have a whole block to arrive (loads past p read obs 0, in bounds)
      ObsMeta m0, m1, MA, MB;
      m0.load_index(a, off, p, 0);
      m1.load_index(a, off, p, 64);
      MA.load_index(a, off, p, 128);
      m0.load_obs<VERT>(a);
      m1.load_obs<VERT>(a);
      MA.load_obs<VERT>(a);
      m0.store_index(sm, 0);
      m1.store_index(sm, 1);
      m0.store_weights32<VERT>(sm, 0, a, clnp, scale);
      m1.store_weights32<VERT>(sm, 1, a, clnp, scale);
      __syncthreads();
      RawT A0[8][NJ], A1[8][NJ], B0[8][NJ], B1[8][NJ];
      load_rows(a, sm, 0, 0, A0);
      load_rows(a, sm, 0, 1, A1);
      // block b (at base): its rows in c0/c1, block b + 1 staged in LDS,
      // block b + 2's metadata in `cur` (loaded), block b + 3's into `nxt`
      auto step = [&](int base, RawT (&c0)[8][NJ], RawT (&c1)[8][NJ], RawT (&n0)[8][NJ],
                      RawT (&n1)[8][NJ], ObsMeta& cur, ObsMeta& nxt) {
        const int buf = (base >> 6) & 1;
        nxt.load_index(a, off, p, base + 192);
        load_rows(a, sm, buf ^ 1, 0, n0);
        load_rows(a, sm, buf ^ 1, 1, n1);
        __builtin_amdgcn_sched_barrier(0);
        gram_half(sm, buf, 0, c0, scale, bacc, accd, acch);
        nxt.load_obs<VERT>(a);
        if (base + 32 < p) gram_half(sm, buf, 1, c1, scale, bacc, accd, acch);
        __syncthreads();  // block b's buffer has been read
        cur.store_index(sm, buf);
        cur.store_weights32<VERT>(sm, buf, a, clnp, scale);
        __syncthreads();
      };
      for (int base = 0; base < p; base += 128) {
        step(base, A0, A1, B0, B1, MA, MB);
        if (base + 64 >= p) break;
        step(base + 64, B0, B1, A0, A1, MB, MA);
      }
    } else if (p > 0) {
      // software pipeline over 32-obs halves: the rows of a half are issued
      // while the previous half computes; the next block's indices are
      // staged in the other LDS buffer during the current block
      ObsMeta m;
      m.load_index(a, off, p, 0);
      m.load_obs<VERT>(a);
      m.store_index(sm, 0);
      m.store_weights<VERT>(sm, 0, a, clnp);
      __syncthreads();
      RawT rawA[8][NJ], rawB[8][NJ];
      load_rows(a, sm, 0, 0, rawA);
      for (int base = 0; base < p; base += 64) {
        const int buf = (base >> 6) & 1;
        ObsMeta n;
        n.load_index(a, off, p, base + 64);
        load_rows(a, sm, buf, 1, rawB);
        __builtin_amdgcn_sched_barrier(0);
        gram_half(sm, buf, 0, rawA, scale, bacc, accd, acch);
        n.load_obs<VERT>(a);
        n.store_index(sm, buf ^ 1);
        __syncthreads();
        load_rows(a, sm, buf ^ 1, 0, rawA);
        __builtin_amdgcn_sched_barrier(0);
        if (base + 32 < p) gram_half(sm, buf, 1, rawB, scale, bacc, accd, acch);
        n.store_weights<VERT>(sm, buf ^ 1, a, clnp);
        __syncthreads();
      }
    }
    if constexpr (kF64) {
      // D[4r + g][c] of tile (I, J) = C(I*16 + 4r + g, J*16 + c)
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J < NJ; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool in = I * 16 + 4 * r + g < k && J * 16 + c < k;
            X[J][I * 4 + r] = in ? accd[I][J][r] : 0.0;
          }
      patB = true;
    } else {
      // D[4g + r][c] of tile (I, J) = C(I*16 + 4g + r, J*16 + c), formed for
      // I >= J; tile (I, J) of I < J is the transpose of tile (J, I), passed
      // through LDS (written as element (4g + r, c), read as (c, 4g + r))
#pragma unroll
      for (int I = 0; I < NJ; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool in = I * 16 + 4 * g + r < k && J * 16 + c < k;
            X[J][I * 4 + r] = in ? (double)acch[I][J][r] * inv2 : 0.0;
          }
#pragma unroll
      for (int J = 1; J < NJ; ++J)
#pragma unroll
        for (int I = 0; I < J; ++I) {
#pragma unroll
          for (int r = 0; r < 4; ++r) sm.ct[(4 * g + r) * 17 + c] = acch[J][I][r];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool in = I * 16 + 4 * g + r < k && J * 16 + c < k;
            X[J][I * 4 + r] = in ? (double)sm.ct[c * 17 + 4 * g + r] * inv2 : 0.0;
          }
          __syncthreads();  // read before the next tile overwrites it
        }
      patB = false;
    }
    // b: lanes hold partials of rows I*16 + c over their obs; sum over g
#pragma unroll
    for (int I = 0; I < NJ; ++I) bacc[I] = sum_g(bacc[I]);
    __syncthreads();
    if (g == 0)
#pragma unroll
      for (int I = 0; I < NJ; ++I) sm.b[I * 16 + c] = I * 16 + c < k ? bacc[I] : 0.0;
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

  // X <- M V with V from LDS (the basis of the previous analysis).
  template <class SM>
  __device__ void warm(const SM& sm) {
    constexpr int LDV = SM::LDV;
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

  // ---- one-sided Jacobi with scaled rotations ------------------------------
  // Column j of the working matrix is sc[j] X[j]; nrm[j] is its squared norm
  // (kept by alpha' = alpha - t gamma, beta' = beta + t gamma and recomputed
  // at the start of each sweep, when the scales are folded back into X).  The
  // rotation p' = c p - s q, q' = s p + c q (t = s / c) is applied as
  // x_p' = x_p - t (sc_q / sc_p) x_q, x_q' = x_q + t (sc_p / sc_q) x_p with
  // sc_p' = c sc_p, sc_q' = c sc_q: two FMAs a row instead of four.  |t| <= 1,
  // so c >= 1/sqrt(2) and a sweep (KP - 1 rotations per column) changes a
  // scale by at most 2^((KP - 1) / 2).
  double nrm[NJ], sc[NJ], isc[NJ];
  // With two pairs a lane (NP = 2) the state lives with the pair that is
  // rotated: lane groups 0-1 hold the top tp and bottom bt of the first pair
  // of the round, groups 2-3 those of the second.  A round then broadcasts
  // only the two FMA factors of each pair, and between rounds the state
  // follows the pairing (nrm, sc, isc above hold it between sweeps).
  struct ColSt {
    double nrm, sc, isc;
  };
  ColSt tp, bt;

  template <int P, int Q>
  __device__ __forceinline__ double dot() const {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q += 2) {
      s0 = fma(X[P][q], X[Q][q], s0);
      s1 = fma(X[P][q + 1], X[Q][q + 1], s1);
    }
    return s0 + s1;
  }

  // The dot products of the two pairs of a round, their FMAs interleaved.
  template <int P0, int Q0, int P1, int Q1>
  __device__ __forceinline__ void dot2(double& r0, double& r1) const {
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q += 2) {
      a0 = fma(X[P0][q], X[Q0][q], a0);
      b0 = fma(X[P1][q], X[Q1][q], b0);
      a1 = fma(X[P0][q + 1], X[Q0][q + 1], a1);
      b1 = fma(X[P1][q + 1], X[Q1][q + 1], b1);
    }
    r0 = a0 + a1;
    r1 = b0 + b1;
  }

  // Rotate columns P, Q given gs = X[P] . X[Q] (summed over the wave).  A
  // pair is rotated while gamma^2 > tol^2 alpha beta; *big records a pair with
  // gamma^2 > tol alpha beta (the next sweep would still rotate, the test of
  // predicted convergence).  Pairs below the threshold take t = 0, c = 1.
  template <int P, int Q>
  __device__ __forceinline__ void rot(double gs, double tol2, double tolp, bool& rotated,
                                      bool& big) {
    const double al = nrm[P], be = nrm[Q];
    const double gm = gs * (sc[P] * sc[Q]);
    const double gg = gm * gm, ab = al * be;
    const bool need = gg > tol2 * ab;
    big |= gg > tolp * ab;
    rotated |= need;
    // t = sign(d e) |e| / (|d| + sqrt(d^2 + e^2)), d = beta - alpha, e = 2 gamma
    const double d = be - al, e = 2.0 * gm;
    const double r2 = fma(d, d, e * e);
    const double den = fma(r2, __builtin_amdgcn_rsq(r2), fabs(d));
    double t = copysign(fabs(e) * __builtin_amdgcn_rcp(den), d * e);
    t = need ? t : 0.0;
    const double u = fma(t, t, 1.0);
    double cs = rsq_nr<kNewton>(u);  // cos = 1 / sqrt(1 + t^2)
    double ic = u * cs;     // 1 / cos
    cs = need ? cs : 1.0;
    ic = need ? ic : 1.0;
    const double ta = t * (sc[Q] * isc[P]);
    const double tb = t * (sc[P] * isc[Q]);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double x = X[P][q], y = X[Q][q];
      X[P][q] = fma(-ta, y, x);
      X[Q][q] = fma(tb, x, y);
    }
    nrm[P] = fma(-t, gm, al);
    nrm[Q] = fma(t, gm, be);
    sc[P] *= cs;
    sc[Q] *= cs;
    isc[P] *= ic;
    isc[Q] *= ic;
  }

  // Apply the rotation (t, gm, cs, ic) to columns P, Q and their state.
  template <int P, int Q>
  __device__ __forceinline__ void apply(double t, double gm, double cs, double ic) {
    const double ta = t * (sc[Q] * isc[P]);
    const double tb = t * (sc[P] * isc[Q]);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double x = X[P][q], y = X[Q][q];
      X[P][q] = fma(-ta, y, x);
      X[Q][q] = fma(tb, x, y);
    }
    nrm[P] = fma(-t, gm, nrm[P]);
    nrm[Q] = fma(t, gm, nrm[Q]);
    sc[P] *= cs;
    sc[Q] *= cs;
    isc[P] *= ic;
    isc[Q] *= ic;
  }

  // x_P' = x_P - ta x_Q, x_Q' = x_Q + tb x_P over the lane's rows.
  template <int P, int Q>
  __device__ __forceinline__ void rotate_x(double ta, double tb) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double x = X[P][q], y = X[Q][q];
      X[P][q] = fma(-ta, y, x);
      X[Q][q] = fma(tb, x, y);
    }
  }

  // One round: pairs (T0, B0) and, with two pairs a lane, (T1, B1).  With two
  // pairs, lane groups 0-1 compute the angle of the first from their state
  // (tp, bt) and groups 2-3 that of the second; each half updates its state
  // and the FMA factors of both pairs are then broadcast.
  template <int T0, int B0, int T1, int B1>
  __device__ __forceinline__ void round(double tol2, double tolp, bool& rotated, bool& big) {
    if constexpr (NP == 2) {
      double d0, d1;
      dot2<T0, B0, T1, B1>(d0, d1);
      const double gs = sum_g_pair(d0, d1);
      double t, gm, cs, ic;
      rotation_angle<kNewton>(gs, tp.nrm, bt.nrm, tp.sc, bt.sc, tol2, tolp, rotated, big, t, gm,
                              cs, ic);
      const double ta = t * (bt.sc * tp.isc);
      const double tb = t * (tp.sc * bt.isc);
      tp.nrm = fma(-t, gm, tp.nrm);
      bt.nrm = fma(t, gm, bt.nrm);
      tp.sc *= cs;
      bt.sc *= cs;
      tp.isc *= ic;
      bt.isc *= ic;
      double ta0, ta1, tb0, tb1;
      split_g(ta, ta0, ta1);
      split_g(tb, tb0, tb1);
      rotate_x<T0, B0>(ta0, tb0);
      rotate_x<T1, B1>(ta1, tb1);
    } else {
      rot<T0, B0>(sum_g(dot<T0, B0>()), tol2, tolp, rotated, big);
    }
  }

  // NP = 2 state moves.  The bottoms' state trades lane halves (the pairing
  // (0, 1), (2, 3) to (0, 3), (2, 1) and back).
  __device__ __forceinline__ void swap_bt() {
    const bool second = threadIdx.x >= 32;
    double a, b;
    split_g(bt.nrm, a, b);
    bt.nrm = second ? a : b;
    split_g(bt.sc, a, b);
    bt.sc = second ? a : b;
    split_g(bt.isc, a, b);
    bt.isc = second ? a : b;
  }
  // x and y trade halves crosswise: x keeps groups 0-1 and takes y's groups
  // 0-1 into groups 2-3; y takes x's groups 2-3 into groups 0-1 and keeps
  // groups 2-3 (one v_permlane32_swap of the two registers).
  static __device__ __forceinline__ void cross(double& x, double& y) {
    const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false,
                                                    false);
    const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false,
                                                    false);
    x = __hiloint2double(b[0], a[0]);
    y = __hiloint2double(b[1], a[1]);
  }
  // From the pairing (0, 3), (2, 1) to (0, 2), (3, 1): groups 0-1 keep their
  // top (column 0) and take the top of groups 2-3 (column 2) as bottom;
  // groups 2-3 take the bottom of groups 0-1 (column 3) as top and keep their
  // bottom (column 1).
  __device__ __forceinline__ void cross_tb() {
    cross(tp.nrm, bt.nrm);
    cross(tp.sc, bt.sc);
    cross(tp.isc, bt.isc);
  }
  __device__ __forceinline__ void bt_xor1() {
    bt.nrm = swz_xor1(bt.nrm);
    bt.sc = swz_xor1(bt.sc);
    bt.isc = swz_xor1(bt.isc);
  }
  __device__ __forceinline__ void bt_perm(int addr) {
    bt.nrm = bperm_d(addr, bt.nrm);
    bt.sc = bperm_d(addr, bt.sc);
    bt.isc = bperm_d(addr, bt.isc);
  }
  __device__ __forceinline__ void tb_xswap(int addr, bool first) {
    double w = bperm_d(addr, first ? bt.nrm : tp.nrm);
    bt.nrm = first ? w : bt.nrm;
    tp.nrm = first ? tp.nrm : w;
    w = bperm_d(addr, first ? bt.sc : tp.sc);
    bt.sc = first ? w : bt.sc;
    tp.sc = first ? tp.sc : w;
    w = bperm_d(addr, first ? bt.isc : tp.isc);
    bt.isc = first ? w : bt.isc;
    tp.isc = first ? tp.isc : w;
  }
  template <int J>
  __device__ __forceinline__ void x_xor1() {
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[J][q] = swz_xor1(X[J][q]);
  }
  template <int J>
  __device__ __forceinline__ void x_perm(int addr) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[J][q] = bperm_d(addr, X[J][q]);
  }
  template <int T, int B>
  __device__ __forceinline__ void x_xswap(int addr, bool first) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double w = bperm_d(addr, first ? X[B][q] : X[T][q]);
      X[B][q] = first ? w : X[B][q];
      X[T][q] = first ? X[T][q] : w;
    }
  }

  // Column J takes the column J of lane l ^ 1 (with its state).
  template <int J>
  __device__ __forceinline__ void move_xor1() {
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[J][q] = swz_xor1(X[J][q]);
    nrm[J] = swz_xor1(nrm[J]);
    sc[J] = swz_xor1(sc[J]);
    isc[J] = swz_xor1(isc[J]);
  }
  // Column J takes the column J of the lane at byte address addr.
  template <int J>
  __device__ __forceinline__ void move_perm(int addr) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[J][q] = bperm_d(addr, X[J][q]);
    nrm[J] = bperm_d(addr, nrm[J]);
    sc[J] = bperm_d(addr, sc[J]);
    isc[J] = bperm_d(addr, isc[J]);
  }
  // Lanes of the first half (first) take the top T of their partner (byte
  // address addr) into their bottom B; the partner takes their bottom into
  // its top.  Each lane therefore sends what its partner takes.
  template <int T, int B>
  __device__ __forceinline__ void xswap(int addr, bool first) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double w = bperm_d(addr, first ? X[B][q] : X[T][q]);
      X[B][q] = first ? w : X[B][q];
      X[T][q] = first ? X[T][q] : w;
    }
    double w = bperm_d(addr, first ? nrm[B] : nrm[T]);
    nrm[B] = first ? w : nrm[B];
    nrm[T] = first ? nrm[T] : w;
    w = bperm_d(addr, first ? sc[B] : sc[T]);
    sc[B] = first ? w : sc[B];
    sc[T] = first ? sc[T] : w;
    w = bperm_d(addr, first ? isc[B] : isc[T]);
    isc[B] = first ? w : isc[B];
    isc[T] = first ? isc[T] : w;
  }

  __device__ __forceinline__ void fold_scales() {
#pragma unroll
    for (int J = 0; J < NJ; ++J) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) X[J][q] *= sc[J];
      sc[J] = 1.0;
      isc[J] = 1.0;
    }
  }

  // One-sided Jacobi on X.  Returns sweeps; *conv when a sweep rotated
  // nothing, or (predict) when no pair of a sweep had gamma^2 > tol alpha
  // beta, so that the next sweep would rotate nothing.
  //
  // Order of a sweep (KP - 1 rounds): pair slots s = c NP + P hold a top
  // (column 2P of lane group c) and a bottom (2P + 1).  For G = KP/2, KP/4,
  // ..., 1, each block of G slots runs G rounds in which the top of slot s
  // meets the bottom that started the level in slot s ^ gray(r); between
  // rounds the bottoms move by the one bit in which gray(r) and gray(r + 1)
  // differ (bit 0 within a lane when NP = 2, handled by renaming; higher bits
  // across lane groups).  Then the bottom of every slot in the first half of
  // its block swaps with the top of slot s + G/2, and the two halves recurse.
  __device__ int jacobi(double tol, int max_sweeps, bool predict, bool* conv) {
    const int lane = threadIdx.x, c = lane & 15;
    const double tol2 = tol * tol;
    const double tolp = predict ? tol : __builtin_inf();
#pragma unroll
    for (int J = 0; J < NJ; ++J) {
      sc[J] = 1.0;
      isc[J] = 1.0;
    }
    int sweep = 0;
    *conv = false;
    for (; sweep < max_sweeps; ++sweep) {
      fold_scales();
#pragma unroll
      for (int J = 0; J < NJ; ++J) nrm[J] = colnorm(J);
      bool rotated = false, big = false;
      if constexpr (NP == 2) {
        // groups 0-1: pair (0, 1), groups 2-3: pair (2, 3)
        const bool second = lane >= 32;
        tp.nrm = second ? nrm[2] : nrm[0];
        bt.nrm = second ? nrm[3] : nrm[1];
        tp.sc = bt.sc = tp.isc = bt.isc = 1.0;
        for (int G = KP / 2; G >= 4; G >>= 1) {
          for (int r = 0; r < G; r += 4) {
            round<0, 1, 2, 3>(tol2, tolp, rotated, big);
            swap_bt();
            round<0, 3, 2, 1>(tol2, tolp, rotated, big);
            x_xor1<1>();
            x_xor1<3>();
            bt_xor1();
            round<0, 3, 2, 1>(tol2, tolp, rotated, big);
            swap_bt();
            round<0, 1, 2, 3>(tol2, tolp, rotated, big);
            if (r + 4 < G) {
              const int addr = (lane ^ (((r + 4) & -(r + 4)) >> 1)) << 2;
              x_perm<1>(addr);
              x_perm<3>(addr);
              bt_perm(addr);
            }
          }
          const int d = G >> 2;
          const int addr = (lane ^ d) << 2;
          const bool first = (c & d) == 0;
          x_xswap<0, 1>(addr, first);
          x_xswap<2, 3>(addr, first);
          tb_xswap(addr, first);
        }
        round<0, 1, 2, 3>(tol2, tolp, rotated, big);
        swap_bt();
        round<0, 3, 2, 1>(tol2, tolp, rotated, big);
        cross_tb();
        round<0, 2, 3, 1>(tol2, tolp, rotated, big);
        // groups 0-1 hold columns 0 and 2, groups 2-3 columns 3 and 1
        split_g(tp.sc, sc[0], sc[3]);
        split_g(bt.sc, sc[2], sc[1]);
      } else {
        for (int G = KP / 2; G >= 1; G >>= 1) {
          for (int r = 0; r < G; ++r) {
            round<0, 1, 0, 1>(tol2, tolp, rotated, big);
            if (r + 1 < G) move_perm<1>((lane ^ ((r + 1) & -(r + 1))) << 2);
          }
          if (G > 1) {
            const int d = G >> 1;
            xswap<0, 1>((lane ^ d) << 2, (c & d) == 0);
          }
        }
      }
      const bool any_rot = __any(rotated);
      const bool any_big = __any(big);
      if (!any_rot || (predict && !any_big)) {
        *conv = true;
        ++sweep;
        break;
      }
    }
    fold_scales();
    return sweep;
  }
};

// Xa' (re-centred, optionally RTPS-relaxed) and xbar_a of nr rows of column
// col from the transform tiles T (T[I][E][r] at lane (g, c) = trans(I*16 + 4r
// + g, E*16 + c), zero past k), 16 rows at a time: D_E[4r' + g][c] =
// Xa(E*16 + 4r' + g, row R0 + c); rows rmap[0 .. nr) with MAP, else 0 .. nr.
template <int KP, bool MAP>
__device__ __forceinline__ void update_rows(int k, int nrow, const double* __restrict__ xb,
                                            const double* __restrict__ axb, double* xa,
                                            double* axa, bool rtps, double rtps_alpha,
                                            int64_t col,
                                            const double4_t (&T)[KP / 16][KP / 16],
                                            const int32_t* rmap, int nr) {
  constexpr int NJ = KP / 16;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const double* xbc = xb + col * (int64_t)nrow * k;
  double* xac = xa + col * (int64_t)nrow * k;
  auto phys = [&](int r) { return MAP ? rmap[r] : r; };
  // Xb' of rows R0 + c (clamped to the last row, in bounds), members of the
  // B-operand pattern; the next block's loads are issued before a block
  // computes
  auto load_xb = [&](int R0, double (&xl)[NJ][4]) {
    const int xr = R0 + c;
    const double* xrow = xbc + (int64_t)phys(xr < nr ? xr : nr - 1) * k;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rowB(I, r, g);
        xl[I][r] = xrow[i < k ? i : k - 1];
      }
  };
  double xl[NJ][4];
  if (nr > 0) load_xb(0, xl);
  for (int R0 = 0; R0 < nr; R0 += 16) {
    const int xr = R0 + c;
    const bool rok = xr < nr;
    const int pr = phys(rok ? xr : nr - 1);
    double4_t D[NJ];
#pragma unroll
    for (int E = 0; E < NJ; ++E) D[E] = double4_t{0.0, 0.0, 0.0, 0.0};
    double xn[NJ][4];
    load_xb(R0 + 16, xn);
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
    const double xbar = axb[col * nrow + pr];
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
    if (rtps) {
      ssb = sum_g(ssb);
      ssa = sum_g(ssa);
      if (ssa > 0.0) fac = rtps_alpha * sqrt(ssb) / sqrt(ssa) + 1.0 - rtps_alpha;
    }
    if (rok) {
#pragma unroll
      for (int E = 0; E < NJ; ++E)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = rowB(E, r, g);
          if (e < k) xac[(int64_t)pr * k + e] = D[E][r] * fac;
        }
      if (g == 0) axa[col * nrow + pr] = mean;
    }
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) xl[I][r] = xn[I][r];
  }
}

// MODE: kModeColumn (one analysis per column, fused update), kModeVertical
// (one per column and level cube, fused update of the cube's rows) or
// kModeTransOut (one per horizontal cube, transform written to HBM)
// Returns whether the analysis's eigenvectors may warm-start the next one:
// not when it was not finite (its transform is NaN) or its eigensolve hit the
// sweep cap (its basis is not orthonormal, and A = M V_prev would then not
// carry the eigenvectors of M).
template <int KP, int GRAM, int MODE, bool FULLV>
__device__ bool wave_column(const dk_batched_args& a, int64_t col, int cube, bool warm,
                            WSmem<KP, FULLV>& sm, Wave<KP, GRAM>& w) {
  constexpr int NJ = Wave<KP, GRAM>::NJ;
  constexpr int NQ = Wave<KP, GRAM>::NQ;
  constexpr int LDV = WSmem<KP, FULLV>::LDV;
  constexpr bool VERT = MODE == kModeVertical;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int k = a.k;
  PhaseClock clk(a.phase_clk);
  const int64_t off = a.offset[col] - a.base;
  int p = a.nobsl[col];
  p = p < 0 ? 0 : (p > a.max_p ? a.max_p : p);
  if (off < 0 || off + p > a.capacity) p = 0;
  // analysis id; the cube's centre ln p with vertical localisation
  const int64_t aid = VERT ? col * a.ncube + cube : col;
  const double clnp = VERT ? a.cube_lnp[aid] : 0.0;
  int lo = 0, hi = p;
  if (VERT && a.sel_key) {
    // the column's list is ordered by ln p: only [lo, hi) can lie within
    // 2 vert_dist of the cube's centre (float keys, so a margin of 1e-5 in
    // ln p on each side); the rest carry a vertical weight of exactly 0
    const double half = 2.0 * a.vert_dist * (1.0 + 1e-12) + 1e-5;
    const double klo = clnp - half, khi = clnp + half;
    const float* key = a.sel_key + off;
    int L = 0, R = p;
    while (L < R) {
      const int m = (L + R) >> 1;
      if ((double)key[m] < klo) L = m + 1;
      else R = m;
    }
    lo = L;
    R = p;
    while (L < R) {
      const int m = (L + R) >> 1;
      if ((double)key[m] <= khi) L = m + 1;
      else R = m;
    }
    hi = L;
  }

  w.template gram<VERT>(a, off + lo, hi - lo, clnp, sm);
  const double km1 = (double)(k - 1);
  w.shift(k, km1 / a.rho);
  clk.mark(DK_PHASE_GRAM);
  if (FULLV && warm) w.warm(sm);
  clk.mark(DK_PHASE_WARM);
  bool conv;
  const int sweeps = w.jacobi(a.tol, a.max_sweeps, a.predict != 0, &conv);
  clk.mark(DK_PHASE_JACOBI);

  // lambda_j = |a_j|; V = A / lambda (to LDS); f = sqrt((k-1)/lambda);
  // u = V^T b / lambda
  double lam[NJ];
  // a non-finite Yb, d or R^-1 reaches M (then lambda) or b: the analysis is
  // flagged (info -2) and its transform set to NaN, as the reference's NaN
  // would propagate (letkf_compute.F90:168-187 of the Sunway path checks for it)
  bool bad = false;
#pragma unroll
  for (int J = 0; J < NJ; ++J) {
    lam[J] = sqrt(w.colnorm(J));
    bad |= !__builtin_isfinite(lam[J]) || !__builtin_isfinite(sm.b[J * 16 + c]);
  }
  bad = __any(bad);
  __syncthreads();

  // trans(i, e) = W(i, e) + wbar(i): tiles D[4r + g][c] = (I*16 + 4r + g, I2*16 + c).
  // W = sum_j f_j v_j v_j^T on FP64 MFMA (the tiles on and below the
  // diagonal), wbar = sum_j u_j v_j on the VALU, over the eigenvectors
  // v_j = a_j / lambda_j (through LDS: rows to the lanes of the MFMA
  // operands), f_j = sqrt((k-1)/lambda_j), u_j = v_j^T b / lambda_j; FULLV:
  // all KP at once (V stays for the next warm start), otherwise 16 at a time
  double4_t T[NJ][NJ];
  double wp[NJ];  // partials of wbar(I*16 + c) over this lane's eigenvectors
#pragma unroll
  for (int I = 0; I < NJ; ++I) {
    wp[I] = 0.0;
#pragma unroll
    for (int I2 = 0; I2 < NJ; ++I2) T[I][I2] = double4_t{0.0, 0.0, 0.0, 0.0};
  }
  // eigenvector slot J to LDS, with f and u
  auto put_v = [&](int J) {
    const double il = lam[J] > 0.0 ? 1.0 / lam[J] : 0.0;
    double ub = 0.0;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row_of(w.patB, I, r, g);
        const double v = w.X[J][I * 4 + r] * il;
        sm.V[i + ((FULLV ? J * 16 : 0) + c) * LDV] = v;
        ub += v * sm.b[i];
      }
    ub = sum_g(ub);
    if (g == 0) {
      sm.f[J * 16 + c] = lam[J] > 0.0 ? sqrt(km1 / lam[J]) : 0.0;
      sm.u[J * 16 + c] = ub * il;
    }
  };
  // the contraction over eigenvectors [j0, j1) (LDS column slots from 0)
  auto mul_v = [&](int j0, int j1) {
    for (int jb = j0; jb < j1; jb += 4) {
      const int jj = jb + g;
      const double fj = sm.f[jj], uj = sm.u[jj];
      double v[NJ];
#pragma unroll
      for (int I = 0; I < NJ; ++I) v[I] = sm.V[I * 16 + c + (jj - j0) * LDV];
#pragma unroll
      for (int I = 0; I < NJ; ++I) {
        wp[I] = fma(v[I], uj, wp[I]);
        // W is symmetric: the tiles I2 <= I only
#pragma unroll
        for (int I2 = 0; I2 <= I; ++I2)
          T[I][I2] = mfma_f64_16x16x4(v[I] * fj, v[I2], T[I][I2]);
      }
    }
  };
  if constexpr (FULLV) {
#pragma unroll
    for (int J = 0; J < NJ; ++J) put_v(J);
    __syncthreads();
    mul_v(0, KP);
  } else {
#pragma unroll
    for (int J = 0; J < NJ; ++J) {
      put_v(J);
      __syncthreads();
      mul_v(J * 16, J * 16 + 16);
      __syncthreads();  // the slot is read before the next one overwrites it
    }
  }
  // wbar(i) = sum_j v_j(i) u_j: the partials summed over the lane groups,
  // then through LDS to the rows of the tile layout (wb[I][r] = wbar(I*16 +
  // 4r + g))
  double4_t wb[NJ];
#pragma unroll
  for (int I = 0; I < NJ; ++I) wp[I] = sum_g(wp[I]);
  if (g == 0)
#pragma unroll
    for (int I = 0; I < NJ; ++I) sm.ct[I * 16 + c] = wp[I];
  __syncthreads();
#pragma unroll
  for (int I = 0; I < NJ; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r) wb[I][r] = sm.ct[I * 16 + 4 * r + g];
  __syncthreads();
  // the tiles above the diagonal: transposes of the ones below, through LDS
  // (tile (I2, I) written as element (4r + g, c), read as (c, 4r + g))
#pragma unroll
  for (int I2 = 1; I2 < NJ; ++I2)
#pragma unroll
    for (int I = 0; I < I2; ++I) {
#pragma unroll
      for (int r = 0; r < 4; ++r) sm.ct[(4 * r + g) * 17 + c] = T[I2][I][r];
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; ++r) T[I][I2][r] = sm.ct[c * 17 + 4 * r + g];
      __syncthreads();  // read before the next tile overwrites it
    }
  const double nan = __builtin_nan("");
#pragma unroll
  for (int I = 0; I < NJ; ++I)
#pragma unroll
    for (int I2 = 0; I2 < NJ; ++I2)
#pragma unroll
      for (int r = 0; r < 4; ++r) T[I][I2][r] = bad ? nan : T[I][I2][r] + wb[I][r];
  if (a.info && lane == 0) a.info[aid] = bad ? -2 : (conv ? sweeps : -1);
  if (a.wbar_out && c == 0)
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rowB(I, r, g);
        if (i < k) a.wbar_out[(aid - a.trans_aid0) * k + i] = wb[I][r];
      }
  clk.mark(DK_PHASE_TRANSFORM);
  if constexpr (MODE == kModeTransOut) {
    // horizontal cubes: the transform goes to HBM; a per-column kernel
    // applies the mean of the covering cubes' transforms
    double* tr = a.trans_out + (aid - a.trans_aid0) * k * k;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int I2 = 0; I2 < NJ; ++I2)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = rowB(I, r, g), e = I2 * 16 + c;
          if (i < k && e < k) tr[i + (int64_t)e * k] = T[I][I2][r];
        }
    return conv && !bad;
  }

  // update: all nrow rows of the column, or with vertical localisation the
  // cube's list of rows
  const int32_t* rmap = VERT ? a.cube_rows + a.cube_row_off[cube] : nullptr;
  const int nr = VERT ? a.cube_row_off[cube + 1] - a.cube_row_off[cube] : a.nrow;
  if (VERT && a.cube_n > 1) {
    // horizontal cube of columns: every column takes the cube's transform
    const int aj = (int)(col % a.nlat), ai = (int)(col / a.nlat);
    const int sj = aj * a.cube_n, si = ai * a.cube_n;
    const int ej = min(sj + a.cube_n - 1, a.grid_nlat - 1);
    const int ei = min(si + a.cube_n - 1, a.grid_nlon - 1);
    for (int ilon = si; ilon <= ei; ++ilon)
      for (int ilat = sj; ilat <= ej; ++ilat)
        update_rows<KP, VERT>(k, a.nrow, a.xb, a.axb, a.xa, a.axa, a.rtps != 0, a.rtps_alpha,
                              (int64_t)ilat + (int64_t)a.grid_nlat * ilon, T, rmap, nr);
  } else {
    update_rows<KP, VERT>(k, a.nrow, a.xb, a.axb, a.xa, a.axa, a.rtps != 0, a.rtps_alpha, col,
                          T, rmap, nr);
  }
  clk.mark(DK_PHASE_UPDATE);
  (void)NQ;
  return conv && !bad;
}

// The plain form: one wavefront per analysis of [aid0, aid0 + gridDim.x) in
// grid order, every eigensolve cold (no chains), transform to HBM.  Two
// wavefronts a SIMD (256 registers each: the few spills cost less than the
// latency a second wavefront hides; config 2: 210 -> 198 ms a sweep).
template <int KP>
__global__ __launch_bounds__(64, 2) void k_letkf_plain(dk_batched_args a, int32_t aid0) {
  __shared__ WSmem<KP, false> sm;
  Wave<KP, DK_GRAM_FP64> w;
  wave_column<KP, DK_GRAM_FP64, kModeTransOut, false>(a, (int64_t)aid0 + blockIdx.x, 0, false,
                                                       sm, w);
}

// First longitude of stripe g of n groups (groups of whole longitudes).
__device__ __forceinline__ int stripe_lo(int g, int n, int nlon) {
  return (int)(((int64_t)g * nlon) / n);
}

// One wavefront per SIMD: the fused analysis needs about 490 registers
// (at 256, two waves a SIMD, it spills 300-700 of them and runs slower).
template <int KP, int GRAM, int MODE>
__global__ __launch_bounds__(64, 1) void k_letkf_wave(dk_batched_args a) {
  __shared__ WSmem<KP, true> sm;
  Wave<KP, GRAM> w;
  const int lane = threadIdx.x;
  // one queue (ngroup 0) or ngroup group queues, own group first
  const int ng = a.ngroup > 0 ? a.ngroup : 1;
  const int g0 = a.ngroup > 0 ? (int)(blockIdx.x % (unsigned)ng) : 0;
  int s = 0, gq = g0, beg = 0;
  int cnt = a.bucketed ? (int)*a.nchain_dev : a.nchain;
  if (a.ngroup > 0) {
    beg = stripe_lo(gq, ng, a.nlon) * a.chains_per_lon;
    cnt = stripe_lo(gq + 1, ng, a.nlon) * a.chains_per_lon - beg;
  }
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(a.queue_head + gq, 1);
    item = __shfl(item, 0, kWave);
    if (item >= cnt) {
      if (++s >= ng) break;
      gq = (g0 + s) % ng;
      beg = stripe_lo(gq, ng, a.nlon) * a.chains_per_lon;
      cnt = stripe_lo(gq + 1, ng, a.nlon) * a.chains_per_lon - beg;
      continue;
    }
    const int chain = a.queue ? a.queue[beg + item] : beg + item;
    int ilon, lat0, lat1;
    if (a.bucketed) {
      const int c0 = a.chain_col0[chain];
      ilon = c0 / a.nlat;
      lat0 = c0 - ilon * a.nlat;
      lat1 = lat0 + a.chain_ncol[chain];
    } else {
      ilon = chain / a.chains_per_lon;
      lat0 = (chain % a.chains_per_lon) * a.chain_len;
      lat1 = min(lat0 + a.chain_len, a.nlat);
    }
    // an analysis warm-starts from the previous one of its chain unless that
    // one was not finite or did not converge
    bool prev_ok = false;
    for (int ilat = lat0; ilat < lat1; ++ilat) {
      const int64_t col = (int64_t)ilon * a.nlat + ilat;
      if constexpr (MODE == kModeVertical) {
        // the cubes of a column in alternating order, so that every analysis
        // warm-starts from the adjacent cube of its column or the same cube
        // of the previous column
        const bool down = ((ilat - lat0) & 1) != 0;
        for (int t = 0; t < a.ncube; ++t)
          prev_ok = wave_column<KP, GRAM, MODE, true>(a, col, down ? a.ncube - 1 - t : t,
                                                      a.warm_start && prev_ok, sm, w);
      } else {
        prev_ok = wave_column<KP, GRAM, MODE, true>(a, col, 0, a.warm_start && prev_ok, sm, w);
      }
    }
  }
}

// Sort key of chain c: 2^32 - 1 - cost, so that an ascending radix sort puts
// the most expensive chains first.  Cost of a column ~ its local obs count
// (Gram) + 10 k (eigensolve, transform and update), times the number of
// vertical cubes.
__global__ void k_chain_keys(dk_batched_args a, uint32_t* key, int32_t* id) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nchain) return;
  int ilon, lat0, lat1;
  if (a.bucketed) {
    // a.nchain entries are sorted; those past the built chains go last
    if (c >= (int)*a.nchain_dev) {
      key[c] = 0xFFFFFFFFu;
      id[c] = c;
      return;
    }
    const int c0 = a.chain_col0[c];
    ilon = c0 / a.nlat;
    lat0 = c0 - ilon * a.nlat;
    lat1 = lat0 + a.chain_ncol[c];
  } else {
    ilon = c / a.chains_per_lon;
    lat0 = (c % a.chains_per_lon) * a.chain_len;
    lat1 = min(lat0 + a.chain_len, a.nlat);
  }
  double cost = 0.0;
  for (int ilat = lat0; ilat < lat1; ++ilat)
    cost += (double)a.nobsl[(int64_t)ilon * a.nlat + ilat] + 10.0 * a.k;
  if (a.ncube > 0) cost *= a.ncube;
  key[c] = 0xFFFFFFFFu - (uint32_t)fmin(cost, 4.0e9);
  id[c] = c;
}

// Update of every column from a stored transform: the column's own (trans +
// (col - col0) k^2) or, with horizontal cubes, the mean of the covering
// cubes' transforms (summed in the reference's cube order: longitude starts
// outer, latitude inner; averaging transforms equals averaging the cubes'
// re-centred analyses, which are linear in trans).  One wave per column,
// FP64 MFMA as the fused kernel's update.
template <int KP>
__global__ __launch_bounds__(64) void k_update_cols(dk_update_args u) {
  constexpr int NJ = KP / 16;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int64_t col = (int64_t)u.col0 + blockIdx.x;
  const int k = u.k;
  double4_t T[NJ][NJ];
  if (u.cube_n > 1) {
    const int ilat = (int)(col % u.nlat), ilon = (int)(col / u.nlat);
    const int st = u.step, cn = u.cube_n;
    const int aj0 = ilat - cn + 1 > 0 ? (ilat - cn + 1 + st - 1) / st : 0;
    const int aj1 = min(ilat / st, u.nalat - 1);
    const int ai0 = ilon - cn + 1 > 0 ? (ilon - cn + 1 + st - 1) / st : 0;
    const int ai1 = min(ilon / st, u.nalon - 1);
    const double cnt = (double)((aj1 - aj0 + 1) * (ai1 - ai0 + 1));
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int E = 0; E < NJ; ++E) T[I][E] = double4_t{0.0, 0.0, 0.0, 0.0};
    for (int ai = ai0; ai <= ai1; ++ai)
      for (int aj = aj0; aj <= aj1; ++aj) {
        const double* tr = u.trans + ((int64_t)aj + (int64_t)u.nalat * ai) * k * k;
#pragma unroll
        for (int I = 0; I < NJ; ++I)
#pragma unroll
          for (int E = 0; E < NJ; ++E)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = rowB(I, r, g), e = E * 16 + c;
              T[I][E][r] += (i < k && e < k) ? tr[i + (int64_t)e * k] : 0.0;
            }
      }
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int E = 0; E < NJ; ++E)
#pragma unroll
        for (int r = 0; r < 4; ++r) T[I][E][r] /= cnt;
  } else {
    const double* tr = u.trans + (int64_t)blockIdx.x * k * k;
#pragma unroll
    for (int I = 0; I < NJ; ++I)
#pragma unroll
      for (int E = 0; E < NJ; ++E)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = rowB(I, r, g), e = E * 16 + c;
          T[I][E][r] = (i < k && e < k) ? tr[i + (int64_t)e * k] : 0.0;
        }
  }
  update_rows<KP, false>(k, u.nrow, u.xb, u.axb, u.xa, u.axa, u.rtps != 0, u.rtps_alpha, col, T,
                         nullptr, u.nrow);
}

__device__ __forceinline__ int bucket_of(int p, const dk_chain_build_args& a) {
  int b = 0;
  while (b + 1 < a.nbucket && p >= a.lo[b + 1]) ++b;
  return b;
}

// One wavefront per longitude: the chains of its latitudes, counted (kWrite
// false) or written.  A chain is a piece of a run of columns of one bucket:
// the run is cut every len[bucket] columns from its start.  Lane t holds
// latitudes [t per, (t + 1) per): the run start entering each lane comes
// from a max-scan of the lanes' last run starts, a chain starts where the
// offset from its run start is a multiple of the length, and a chain ends at
// the next chain start (a suffix min-scan of the lanes' first ones).
template <bool kWrite>
__global__ __launch_bounds__(256) void k_chain_build(dk_chain_build_args a) {
  const int lane = threadIdx.x & 63;
  const int ilon = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (ilon >= a.nlon) return;
  const int n = a.nlat;
  const int32_t* p = a.nobsl + (int64_t)ilon * n;
  const int per = (n + kWave - 1) / kWave;
  const int i0 = min(lane * per, n), i1 = min(i0 + per, n);
  const int prevb = i0 > 0 && i0 < n ? bucket_of(p[i0 - 1], a) : -1;
  // the lane's last run start (or -1), then the one entering the lane
  int last = -1;
  int b = prevb;
  for (int i = i0; i < i1; ++i) {
    const int bi = bucket_of(p[i], a);
    if (bi != b) last = i;
    b = bi;
  }
  int inc = last;
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_up(inc, o, kWave);
    if (lane >= o) inc = max(inc, t);
  }
  int rs = __shfl_up(inc, 1, kWave);
  if (lane == 0) rs = -1;
  // chain starts of the lane: count and first
  int cnt = 0, first = n;
  {
    int r = rs;
    b = prevb;
    for (int i = i0; i < i1; ++i) {
      const int bi = bucket_of(p[i], a);
      if (bi != b) r = i;
      b = bi;
      const int len = a.len[bi] > 0 ? a.len[bi] : 1;
      if ((i - r) % len == 0) {
        first = min(first, i);
        ++cnt;
      }
    }
  }
  int before = cnt;  // inclusive scan of the counts
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_up(before, o, kWave);
    if (lane >= o) before += t;
  }
  const int total = __shfl(before, kWave - 1, kWave);
  before -= cnt;
  if (!kWrite) {
    if (lane == 0) a.line_cnt[ilon] = total;
    return;
  }
  // the first chain start above the lane (n if none)
  int nx = first;
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_down(nx, o, kWave);
    if (lane + o < kWave) nx = min(nx, t);
  }
  nx = __shfl_down(nx, 1, kWave);
  if (lane == kWave - 1) nx = n;
  const int64_t out = a.line_off[ilon] + before;
  int r = rs, prev = -1, k = 0;
  b = prevb;
  for (int i = i0; i < i1; ++i) {
    const int bi = bucket_of(p[i], a);
    if (bi != b) r = i;
    b = bi;
    const int len = a.len[bi] > 0 ? a.len[bi] : 1;
    if ((i - r) % len == 0) {
      if (prev >= 0) a.chain_ncol[out + k - 1] = i - prev;
      a.chain_col0[out + k] = ilon * n + i;
      prev = i;
      ++k;
    }
  }
  if (prev >= 0) a.chain_ncol[out + k - 1] = nx - prev;
}

// Grouped queue: chain (ilon, latchain) of stripe g goes to position
// stripe_lo(g) cpl + latchain * width(g) + (ilon - stripe_lo(g)).
__global__ void k_chain_groups(dk_batched_args a, int32_t* order) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nchain) return;
  const int cpl = a.chains_per_lon, ng = a.ngroup;
  const int ilon = c / cpl, lc = c % cpl;
  int g = 0;
  while (g + 1 < ng && stripe_lo(g + 1, ng, a.nlon) <= ilon) ++g;
  const int l0 = stripe_lo(g, ng, a.nlon), wdt = stripe_lo(g + 1, ng, a.nlon) - l0;
  order[l0 * cpl + lc * wdt + (ilon - l0)] = c;
}

template <int KP, int GRAM>
int wlaunch(const dk_batched_args* a, int nblocks, hipStream_t s) {
  if (a->ncube > 0) k_letkf_wave<KP, GRAM, kModeVertical><<<nblocks, 64, 0, s>>>(*a);
  else if (a->trans_out) k_letkf_wave<KP, GRAM, kModeTransOut><<<nblocks, 64, 0, s>>>(*a);
  else k_letkf_wave<KP, GRAM, kModeColumn><<<nblocks, 64, 0, s>>>(*a);
  return (int)hipGetLastError();
}

template <int KP, int GRAM>
int wocc() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_letkf_wave<KP, GRAM, kModeColumn>, 64,
                                                   0))
    return 1;
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

extern "C" int dk_letkf_plain_wave(const dk_core_args* c, void* stream) {
  if (c->nanal <= 0) return 0;
  if (c->k < 2 || c->k > 64 || c->gram != DK_GRAM_FP64 || !c->idx) return (int)hipErrorInvalidValue;
  dk_batched_args a;
  memset(&a, 0, sizeof a);
  a.k = c->k;
  a.ldy = c->k <= 32 ? 32 : 64;
  a.max_sweeps = c->max_sweeps;
  a.max_p = c->max_p;
  a.predict = c->predict;
  a.rho = c->rho;
  a.tol = c->tol;
  a.capacity = c->capacity;
  a.offset = c->offset;
  a.nobsl = c->nobsl;
  a.idx = c->idx;
  a.rloc = c->rloc;
  a.yb = c->yb;
  a.rinv = c->rinv;
  a.dep = c->dep;
  a.info = c->info;
  a.base = c->base;
  a.trans_out = c->trans;
  a.trans_aid0 = c->aid0;
  a.wbar_out = c->wbar;
  hipStream_t s = (hipStream_t)stream;
  if (c->k <= 32) k_letkf_plain<32><<<c->nanal, 64, 0, s>>>(a, c->aid0);
  else k_letkf_plain<64><<<c->nanal, 64, 0, s>>>(a, c->aid0);
  return (int)hipGetLastError();
}

extern "C" int dk_update_cols(const dk_update_args* u, void* stream) {
  if (u->ncol <= 0) return 0;
  if (u->k < 2 || u->k > 64) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (u->k <= 32) k_update_cols<32><<<u->ncol, 64, 0, s>>>(*u);
  else k_update_cols<64><<<u->ncol, 64, 0, s>>>(*u);
  return (int)hipGetLastError();
}

extern "C" int dk_chain_groups(const dk_batched_args* a, int32_t* order, void* stream) {
  if (a->nchain <= 0 || a->ngroup <= 0) return 0;
  k_chain_groups<<<(a->nchain + 255) / 256, 256, 0, (hipStream_t)stream>>>(*a, order);
  return (int)hipGetLastError();
}

extern "C" int dk_chain_build(const dk_chain_build_args* a, int write, void* stream) {
  if (a->nlon <= 0) return 0;
  if (a->nbucket < 1 || a->nbucket > DK_NBUCKET) return (int)hipErrorInvalidValue;
  const unsigned nb = (unsigned)((a->nlon + 3) / 4);
  if (write) k_chain_build<true><<<nb, 256, 0, (hipStream_t)stream>>>(*a);
  else k_chain_build<false><<<nb, 256, 0, (hipStream_t)stream>>>(*a);
  return (int)hipGetLastError();
}

extern "C" int dk_chain_keys(const dk_batched_args* a, uint32_t* key, int32_t* id, void* stream) {
  if (a->nchain <= 0) return 0;
  k_chain_keys<<<(a->nchain + 255) / 256, 256, 0, (hipStream_t)stream>>>(*a, key, id);
  return (int)hipGetLastError();
}

extern "C" int dk_letkf_wave(const dk_batched_args* a, int nblocks, void* stream) {
  if (a->nchain <= 0) return 0;
  if (a->k < 2 || a->k > 64) return (int)hipErrorInvalidValue;
  if (a->ldy != (a->k <= 32 ? 32 : 64)) return (int)hipErrorInvalidValue;
  if (a->gram == DK_GRAM_FP64 ? !a->yb : !a->yb32) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (a->k <= 32) return wdispatch<32>(a, nblocks, s);
  return wdispatch<64>(a, nblocks, s);
}
