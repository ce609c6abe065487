// Sigma-vectors that allocate nothing: adc2g_sigma_device_noalloc.
//
// adc2g_sigma_device grows the handle's workspaces (pair tensors, packed 1h1p
// blocks, GEMM split-K partials) with hipMalloc on demand.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops) use this
// variant instead: its workspaces are regions of one file-scope device array,
// `scratch`, of kScratchBytes per GPU, which each GPU the library runs on holds
// until the process ends.  The array's device address is resolved once per
// device with hipGetSymbolAddress when the first handle is created there; host
// code uses only that address.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race, and the array keeps the last call's
// contents.  Nothing in it needs to start at zero: each pass writes every
// element it later reads (the pair tensors, the packed 1h1p blocks and the
// split-K partials are written before they are read).
#include <mutex>

#include "handle.hpp"

namespace adc2g {

constexpr size_t kScratchBytes = (size_t)256 << 20;  // 256 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Resolve the device address of `scratch` on `device` (the current device),
// once per device.  Called by Handle::create.
void resolve_scratch(int device) {
  ARG_CHECK(device >= 0 && device < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = (char*)p;
}

size_t scratch_capacity() { return kScratchBytes; }

// W = H X for nvec device columns with the workspaces in the device's scratch.
void sigma_noalloc(Handle& H, int nvec, const double* dX, double* dW, Prec p, hipStream_t s) {
  ARG_CHECK(!H.parent, "the no-allocation sigma-vector works on the full operator only (a CVS view stages "
                       "through buffers of its own)");
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[H.device];
  }
  if (!base) throw Error(ADC2G_ESTATE, "the scratch of this device has not been resolved");
  H.scratch = base;
  H.scratch_bytes = kScratchBytes;
  try {
    H.sigma(dX, H.d.ndim, dW, H.d.ndim, nvec, p, s);
  } catch (...) {
    H.scratch = nullptr;
    H.scratch_bytes = 0;
    H.gws.ext = nullptr;
    H.gws.ext_size = 0;
    throw;
  }
  H.scratch = nullptr;
  H.scratch_bytes = 0;
  H.gws.ext = nullptr;
  H.gws.ext_size = 0;
}

}  // namespace adc2g
