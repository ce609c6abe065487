// Shared declarations of the adc2g library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/adc2g.h"

namespace adc2g {

// Error carrying an ADC2G_E* code; the C ABI layer turns it into a return code
// and the thread's last-error message.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::adc2g::Error(ADC2G_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_) + \
                                           " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define ARG_CHECK(cond, msg)                                  \
  do {                                                        \
    if (!(cond)) throw ::adc2g::Error(ADC2G_EARG, (msg));     \
  } while (0)

// ADC2G_POISON=1 fills every new floating-point device buffer with 0xff
// bytes (NaN): a debugging aid that turns a read of memory no kernel has
// written into NaN results.  Integer (index) buffers are left alone.
inline bool poison_new_buffers() {
  static const bool on = [] {
    const char* e = std::getenv("ADC2G_POISON");
    return e != nullptr && e[0] == '1';
  }();
  return on;
}

// Device buffer owned by the handle.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipDeviceSynchronize();  // work in flight may still use the old buffer
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(ADC2G_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                    hipGetErrorString(e));
    }
    n = count;
    if constexpr (!std::is_integral<T>::value || std::is_same<T, char>::value) {
      if (poison_new_buffers()) {
        (void)hipMemset(p, 0xff, count * sizeof(T));
        (void)hipDeviceSynchronize();
      }
    }
  }
  T* get() const { return p; }
};

// Sizes of the ADC(2)-s configuration space (source/adclib/select_fano.f90:539-718).
struct Dims {
  int o = 0, v = 0, naux = 0;
  int64_t n1 = 0, npo = 0, npv = 0;
  int64_t cnt[5] = {0, 0, 0, 0, 0};   // t1..t5
  int64_t off[6] = {0, 0, 0, 0, 0, 0}; // starts of t1..t5 in the full vector, off[5] = ndim
  int64_t n2 = 0, ndim = 0;
  void set(int o_, int v_, int naux_) {
    o = o_; v = v_; naux = naux_;
    n1 = (int64_t)o * v;
    npo = (int64_t)o * (o - 1) / 2;
    npv = (int64_t)v * (v - 1) / 2;
    cnt[0] = (int64_t)o * v;
    cnt[1] = npv * o;
    cnt[2] = npo * v;
    cnt[3] = npo * npv;
    cnt[4] = npo * npv;
    off[0] = n1;
    for (int t = 0; t < 5; ++t) off[t + 1] = off[t] + cnt[t];
    ndim = off[5];
    n2 = ndim - n1;
  }
};

inline int grid_for(int64_t n, int block = 256, int cap = 65536) {
  int64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace adc2g
