// mlp_static.hip -- surrogate inference without device allocation (no
// reference counterpart).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the query
// rows and the two outputs of one call are staged in `scratch`, a file-scope
// device array of 4 MiB, one copy per GPU, held until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns on
// each device (scratch_address, cached per device); the kernels (the fused
// ensemble pass of mlp.hip) take that address as an argument.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU race.  The
// array keeps the last call's contents; nothing in it must start at zero
// (each call writes the rows it reads back).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mlp.h"

constexpr size_t kStaticScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

size_t up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

}  // namespace

int mlp_static_scratch(int device, void** out, std::string& err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device index out of range for the static staging array";
    return CLR_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_addr[device]) {
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return CLR_E_HIP;
    }
    g_addr[device] = p;
  }
  *out = g_addr[device];
  return 0;
}

size_t mlp_static_max_rows(const MlpState* st) {
  const clr_mlp_shape* sh = mlp_get_shape(st);
  const size_t per = sizeof(float) * ((size_t)sh->in_dim + 2 * (size_t)sh->out_dim);
  return (kStaticScratchBytes - 3 * kAlign) / per;
}

int mlp_infer_static(MlpState* st, void* scratch_ptr, const float* x, int64_t n, float* mean, float* stdev,
                     hipStream_t s, std::string& err) {
  if (!mlp_fused_ok(st)) {
    err = "allocation-free inference needs the fused shape: hidden <= 256, in_dim <= 32, members * out_dim <= 80";
    return CLR_E_UNSUPPORTED;
  }
  const int64_t cap = (int64_t)mlp_static_max_rows(st);
  if (n < 1 || n > cap) {
    err = "allocation-free inference takes 1.." + std::to_string(cap) + " queries per call (4 MiB of staging)";
    return CLR_E_RANGE;
  }
  const clr_mlp_shape* sh = mlp_get_shape(st);
  const size_t bx = up(sizeof(float) * (size_t)n * sh->in_dim), bm = up(sizeof(float) * (size_t)n * sh->out_dim);
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr);
  float* xd = reinterpret_cast<float*>(base);
  float* md = reinterpret_cast<float*>(base + bx);
  float* sd = reinterpret_cast<float*>(base + bx + bm);
  hipError_t e = hipMemcpyAsync(xd, x, sizeof(float) * (size_t)n * sh->in_dim, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) { err = std::string("staging queries: ") + hipGetErrorString(e); return CLR_E_HIP; }
  if (int rc = mlp_infer_device(st, xd, n, md, stdev ? sd : nullptr, s, err)) return rc;
  e = hipMemcpyAsync(mean, md, sizeof(float) * (size_t)n * sh->out_dim, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && stdev)
    e = hipMemcpyAsync(stdev, sd, sizeof(float) * (size_t)n * sh->out_dim, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { err = std::string("allocation-free inference: ") + hipGetErrorString(e); return CLR_E_HIP; }
  return 0;
}
