// lh_noalloc.hip -- lh_divergence_noalloc: ComputeDivergence (LiquidModel.cpp:1065-1083) on
// caller-owned device arrays with no allocation, for callers that cannot call hipMalloc (code
// under HIP graph capture, serving loops).  The per-block partial sums that lh_divergence keeps
// in its context live here in one file-scope device array per GPU.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <mutex>
#include <string>

#include "lh_ctx.h"

namespace lh {

// the dense divergence's block count (the context's reduction uses the same, so the sums come in
// the same order and the results are bit-identical)
constexpr int kNoAllocBlocks = 1024;

// per-block partials: sums at [b], maxima at [nb + b], counts at [2 nb + b] (24 KiB per GPU, held
// until the process ends).  In host code this name is a host-side stand-in: only the address
// hipGetSymbolAddress returns for the current device is passed to kernels.
static __device__ __align__(256) double scratch[3 * kNoAllocBlocks];

namespace {

constexpr int kMaxDim = 4096;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of scratch, per device
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// the device address of scratch on the current device, looked up once per device
int scratch_ptr(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return fail(LH_ERR_DEVICE, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return fail(LH_ERR_DEVICE, "device index beyond the supported 64");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    double* p = nullptr;
    hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LH_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return LH_OK;
}

}  // namespace

const char* noalloc_last_error() { return g_err.empty() ? "null context" : g_err.c_str(); }

}  // namespace lh

extern "C" int lh_divergence_noalloc(const float* vx, const float* vy, const float* vz, const float* lwx,
                                     const float* lwy, const float* lwz, const float* phi, const int dims[3],
                                     float dx, double* out, void* stream) {
  using namespace lh;
  if (!vx || !vy || !vz || !lwx || !lwy || !lwz || !phi || !dims || !out)
    return fail(LH_ERR_ARG, "lh_divergence_noalloc: null pointer");
  long long cells = 1;
  for (int a = 0; a < 3; ++a) {
    if (dims[a] < 1 || dims[a] > kMaxDim)
      return fail(LH_ERR_ARG, "lh_divergence_noalloc: each grid dimension must be in [1, 4096]");
    cells *= dims[a];
  }
  if (cells > (long long)INT_MAX)
    return fail(LH_ERR_ARG, "lh_divergence_noalloc: more than 2^31 - 1 cells (the cell count is an int)");
  if (!(dx > 0.f) || !std::isfinite(dx)) return fail(LH_ERR_ARG, "lh_divergence_noalloc: dx must be positive");
  double* part = nullptr;
  int rc = scratch_ptr(&part);
  if (rc) return rc;
  GridP g;
  for (int a = 0; a < 3; ++a) {
    g.N[a] = dims[a];
    g.mn[a] = 0.f;
    g.dx[a] = dx;
  }
  const float* v[3] = {vx, vy, vz};
  const float* lw[3] = {lwx, lwy, lwz};
  launch_divergence_arrays(static_cast<hipStream_t>(stream), v, lw, phi, g, part, kNoAllocBlocks, out, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LH_ERR_DEVICE, std::string("lh_divergence_noalloc: ") + hipGetErrorString(e));
  return LH_OK;
}
