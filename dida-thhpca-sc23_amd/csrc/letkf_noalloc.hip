// The batched core without any device allocation.
//
// dida_letkf_core_batch_dev stages the per-analysis obs offsets (analysis b
// at b * max_obs) in a device buffer it allocates and fills from the host.
// dida_letkf_core_batch_dev_noalloc computes the same offsets with a kernel
// into `scratch`, a fixed array of this library on every device, and then
// runs the same transform kernel, so it calls no allocator and no host copy:
// it can be recorded in a HIP graph or called from a serving loop.  Calls on
// one device share its array and the last call's offsets stay in it; calls
// that overlap on one device (different streams, no ordering) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/dida_letkf.h"
#include "letkf_kernels.h"

// Analyses per call: the offsets of at most this many (8 MiB of int64 per
// device, held until the process ends).
constexpr int32_t kNoAllocMaxBatch = 1 << 20;

static __device__ __align__(256) int64_t scratch[kNoAllocMaxBatch];

namespace {

// offset[b] = b * stride: dense obs blocks of max_obs entries per analysis
__global__ void k_dense_offsets(int64_t* __restrict__ offset, int32_t batch, int32_t stride) {
  const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < batch) offset[b] = (int64_t)b * stride;
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int64_t* g_scratch[kMaxDevices];  // device address of `scratch` on each device

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
int scratch_on_device(int64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return dk_set_error(DIDA_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return dk_set_error(DIDA_ERR_UNSUPPORTED, "core_batch_dev_noalloc: device index >= 64");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dk_set_error(DIDA_ERR_HIP, "core_batch_dev_noalloc: hipGetSymbolAddress failed");
    g_scratch[dev] = (int64_t*)p;
  }
  *out = g_scratch[dev];
  return DIDA_OK;
}

}  // namespace

extern "C" int dida_letkf_core_batch_dev_noalloc(int32_t batch, int32_t max_obs, int32_t nanals,
                                                 const int32_t* nobsl, const double* hxens,
                                                 const float* rdiaginv, const double* dep,
                                                 const double* rloc, double* trans,
                                                 double* overline_w, double inflation_factor,
                                                 int32_t gram, int32_t* info, void* stream) {
  if (batch < 0 || max_obs < 0) return dk_set_error(DIDA_ERR_ARG, "negative size");
  if (batch > kNoAllocMaxBatch)
    return dk_set_error(DIDA_ERR_ARG,
                        "core_batch_dev_noalloc: batch above 1048576 analyses (its fixed "
                        "offset array); split the batch or call dida_letkf_core_batch_dev");
  if (nanals < 2 || nanals > 128)
    return dk_set_error(DIDA_ERR_UNSUPPORTED, "nanals: supported range is 2..128");
  if (gram != DIDA_GRAM_FP64) return dk_set_error(DIDA_ERR_UNSUPPORTED, "gram: FP64 only");
  if (!(inflation_factor > 0.0))
    return dk_set_error(DIDA_ERR_ARG, "inflation_factor must be > 0");
  if (batch == 0) return DIDA_OK;
  if (!nobsl || !trans || (max_obs > 0 && (!hxens || !rdiaginv || !dep || !rloc)))
    return dk_set_error(DIDA_ERR_ARG, "null argument");
  int64_t* offset = nullptr;
  int rc = scratch_on_device(&offset);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  k_dense_offsets<<<(unsigned)((batch + 255) / 256), 256, 0, s>>>(offset, batch, max_obs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return dk_set_error(DIDA_ERR_HIP, hipGetErrorString(e));
  dk_core_args ca = {};
  ca.k = nanals;
  ca.nanal = batch;
  ca.offset = offset;
  ca.nobsl = nobsl;
  ca.capacity = (int64_t)batch * max_obs;
  ca.max_p = max_obs;
  ca.idx = nullptr;
  ca.rloc = rloc;
  ca.yb = hxens;
  ca.rinv = rdiaginv;
  ca.dep = dep;
  ca.rho = inflation_factor;
  ca.tol = 1e-13;
  ca.max_sweeps = 30;
  ca.predict = 0;
  ca.trans = trans;
  ca.wbar = overline_w;
  ca.info = info;
  rc = dk_letkf_core_f64(&ca, s);
  if (rc) return dk_set_error(DIDA_ERR_HIP, hipGetErrorString((hipError_t)rc));
  return DIDA_OK;
}
