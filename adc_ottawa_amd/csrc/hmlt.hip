// The reference's on-disk matrix files written from the device operator:
//   hmlt.dia*: record (maxbl, nbuf) int32, record H_jj (ndim doubles)
//              (wrtdg, source/adclib/filetools.f90:24-34);
//   hmlt.off*: nbuf records of (buff(maxbl) double, oi(maxbl) int32,
//              oj(maxbl) int32, nlim int32) with the first nlim entries valid
//              (wrtoffdg, :36-44), as get_offdiag_adc2_save_omp
//              (source/adclib/get_matrix.f90:705-1186) fills them: the upper
//              triangle i < j, 1-based indices, elements with |h_ij| > minc,
//              full records of maxbl and a last record with the rest (possibly
//              empty), nbuf = floor(noffd / maxbl) + 1.
// Readers: rdham_on / rdham_off (source/eigen/block_davidson.f90:600-686).
// Fortran sequential unformatted layout: each record is framed by its byte
// length as a 4-byte little-endian integer before and after.
//
// The columns of H come from the sigma-vector of blocks of unit vectors;
// one workgroup per column keeps the rows i < j with |h_ij| > minc in row
// order (a workgroup-wide prefix sum per chunk of rows) and only those are
// copied to the host.  Elements are written column by column; the reference
// writes them in its evaluation order, which its readers do not depend on.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int HM_THREADS = 256;
constexpr int HM_COLS = 128;  // columns per sigma pass

// cnt[c] = number of rows i < j0 + c with |W[i, c]| > minc
__global__ void __launch_bounds__(HM_THREADS) hm_count_kernel(int64_t n, int64_t j0, const double* __restrict__ W,
                                                              double minc, int64_t* __restrict__ cnt) {
  __shared__ int red[HM_THREADS / 64];
  const int c = blockIdx.x;
  const int64_t j = j0 + c;
  const double* w = W + n * c;
  int s = 0;
  for (int64_t i = threadIdx.x; i < j; i += HM_THREADS) s += fabs(w[i]) > minc;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int k = 0; k < HM_THREADS / 64; ++k) t += red[k];
    cnt[c] = t;
  }
}

// rows and values of the kept elements of column c at out + off[c], in row order
__global__ void __launch_bounds__(HM_THREADS) hm_compact_kernel(int64_t n, int64_t j0, const double* __restrict__ W,
                                                                double minc, const int64_t* __restrict__ off,
                                                                int32_t* __restrict__ rows,
                                                                double* __restrict__ vals) {
  __shared__ int wsum[HM_THREADS / 64];
  __shared__ int64_t base;
  const int c = blockIdx.x;
  const int64_t j = j0 + c;
  const double* w = W + n * c;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) base = off[c];
  __syncthreads();
  for (int64_t i0 = 0; i0 < j; i0 += HM_THREADS) {
    const int64_t i = i0 + threadIdx.x;
    const double x = (i < j) ? w[i] : 0.0;
    const bool keep = (i < j) && fabs(x) > minc;
    const uint64_t m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(m);
    __syncthreads();
    int64_t pos = base + before;
    for (int k = 0; k < wv; ++k) pos += wsum[k];
    if (keep) {
      rows[pos] = (int32_t)(i + 1);
      vals[pos] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int k = 0; k < HM_THREADS / 64; ++k) t += wsum[k];
      base += t;
    }
    __syncthreads();
  }
}

struct FFile {
  FILE* f = nullptr;
  std::string path;
  explicit FFile(const char* p) : path(p) {
    f = std::fopen(p, "wb");
    if (!f) throw Error(ADC2G_EARG, std::string("cannot open ") + p + " for writing");
  }
  ~FFile() {
    if (f) std::fclose(f);
  }
  void put(const void* p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw Error(ADC2G_EARG, "write to " + path + " failed");
  }
  void marker(size_t bytes) {
    const int32_t m = (int32_t)bytes;
    put(&m, 4);
  }
  void close() {
    if (f && std::fclose(f) != 0) {
      f = nullptr;
      throw Error(ADC2G_EARG, "closing " + path + " failed");
    }
    f = nullptr;
  }
};

}  // namespace

void hmlt_write(Handle& H, const char* dia_path, const char* off_path, double minc, int maxbl, int64_t* noffd_out,
                int32_t* nbuf_out) {
  const int64_t n = H.d.ndim;
  ARG_CHECK(dia_path && off_path, "NULL path");
  ARG_CHECK(std::isfinite(minc) && minc >= 0.0, "minc must be finite and >= 0");
  ARG_CHECK(maxbl >= 1 && (int64_t)maxbl * 16 + 4 <= 2147483647LL, "maxbl out of range");
  ARG_CHECK(n * 8 <= 2147483647LL, "ndim too large for one Fortran record of the diagonal");
  ARG_CHECK(n <= 2147483647LL, "ndim exceeds the int32 indices of the record format");
  hipStream_t st = H.st;

  // the off-diagonal file: records filled column by column
  FFile off(off_path);
  std::vector<double> buff(maxbl, 0.0);
  std::vector<int32_t> oi(maxbl, 0), oj(maxbl, 0);
  int fill = 0;
  int64_t noffd = 0;
  int32_t nbuf = 0;
  const size_t rec_bytes = (size_t)maxbl * 16 + 4;
  auto flush = [&](int nlim) {
    off.marker(rec_bytes);
    off.put(buff.data(), sizeof(double) * maxbl);
    off.put(oi.data(), sizeof(int32_t) * maxbl);
    off.put(oj.data(), sizeof(int32_t) * maxbl);
    const int32_t nl = nlim;
    off.put(&nl, 4);
    off.marker(rec_bytes);
    ++nbuf;
  };

  const int cb = (int)std::min<int64_t>(HM_COLS, n);
  DevBuf<double> X, W;
  DevBuf<int64_t> cnt, offs;
  X.alloc((size_t)n * cb);
  W.alloc((size_t)n * cb);
  cnt.alloc(cb);
  offs.alloc(cb);
  DevBuf<int32_t> rows;
  DevBuf<double> vals;
  std::vector<int64_t> hc(cb), ho(cb);
  std::vector<int32_t> hr;
  std::vector<double> hv;
  DevBuf<int64_t> idx;
  idx.alloc(cb);
  std::vector<int64_t> hidx(cb);
  for (int64_t j0 = 0; j0 < n; j0 += cb) {
    const int nc = (int)std::min<int64_t>(cb, n - j0);
    for (int c = 0; c < nc; ++c) hidx[c] = j0 + c;
    HIP_CHECK(hipMemcpyAsync(idx.get(), hidx.data(), sizeof(int64_t) * nc, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(X.get(), 0, sizeof(double) * n * nc, st));
    launch_unit_vectors(nc, idx.get(), X.get(), n, st);
    H.sigma(X.get(), n, W.get(), n, nc, Prec::F64, st);
    hipLaunchKernelGGL(hm_count_kernel, dim3(nc), dim3(HM_THREADS), 0, st, n, j0, W.get(), minc, cnt.get());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hc.data(), cnt.get(), sizeof(int64_t) * nc, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    int64_t tot = 0;
    for (int c = 0; c < nc; ++c) {
      ho[c] = tot;
      tot += hc[c];
    }
    if (tot == 0) continue;
    rows.alloc((size_t)tot);
    vals.alloc((size_t)tot);
    HIP_CHECK(hipMemcpyAsync(offs.get(), ho.data(), sizeof(int64_t) * nc, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(hm_compact_kernel, dim3(nc), dim3(HM_THREADS), 0, st, n, j0, W.get(), minc, offs.get(),
                       rows.get(), vals.get());
    HIP_CHECK(hipGetLastError());
    hr.resize(tot);
    hv.resize(tot);
    HIP_CHECK(hipMemcpyAsync(hr.data(), rows.get(), sizeof(int32_t) * tot, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hv.data(), vals.get(), sizeof(double) * tot, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int c = 0; c < nc; ++c) {
      const int32_t jj = (int32_t)(j0 + c + 1);
      for (int64_t k = ho[c]; k < ho[c] + hc[c]; ++k) {
        buff[fill] = hv[k];
        oi[fill] = hr[k];
        oj[fill] = jj;
        if (++fill == maxbl) {
          flush(maxbl);
          fill = 0;
        }
      }
    }
    noffd += tot;
  }
  flush(fill);  // the last record, possibly empty
  off.close();

  // the diagonal file
  std::vector<double> hii(n);
  HIP_CHECK(hipMemcpyAsync(hii.data(), H.hii.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  FFile dia(dia_path);
  const int32_t head[2] = {maxbl, nbuf};
  dia.marker(8);
  dia.put(head, 8);
  dia.marker(8);
  dia.marker((size_t)n * 8);
  dia.put(hii.data(), (size_t)n * 8);
  dia.marker((size_t)n * 8);
  dia.close();
  if (noffd_out) *noffd_out = noffd;
  if (nbuf_out) *nbuf_out = nbuf;
}

}  // namespace adc2g
