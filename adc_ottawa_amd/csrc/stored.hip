// A handle on a stored matrix: the reference's on-disk files hmlt.dia* and
// hmlt.off* (wrtdg / wrtoffdg, source/adclib/filetools.f90:24-44), loaded as
// its in-core Davidson loads them (rdham_on / rdham_off,
// source/eigen/block_davidson.f90:588-686), with the product of
// hxvec_incore (:844-883):
//   W(i) = H_ii X(i) + sum_k [i == oi_k] h_k X(oj_k) + [i == oj_k] h_k X(oi_k)
// over the stored elements k.  On the device the elements sit in CSR form
// with both (oi, oj) and (oj, oi) entered per element, 32-bit columns, so
// that one group of lanes per row sums its row with no atomics; the solvers
// then run on the matrix as on an operator built from integrals.  Files
// whose records are split into gfortran sub-records (records over 2 GiB) are
// rejected.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

// one Fortran sequential unformatted record: 4-byte length, payload, 4-byte length
bool read_record(FILE* f, std::vector<char>& buf, const std::string& path) {
  int32_t head = 0, tail = 0;
  const size_t got = std::fread(&head, 4, 1, f);
  if (got != 1) return false;
  if (head < 0) throw Error(ADC2G_EARG, path + ": records split into sub-records are not supported");
  buf.resize((size_t)head);
  if (head > 0 && std::fread(buf.data(), 1, (size_t)head, f) != (size_t)head)
    throw Error(ADC2G_EARG, path + ": truncated record");
  if (std::fread(&tail, 4, 1, f) != 1 || tail != head) throw Error(ADC2G_EARG, path + ": corrupt record marker");
  return true;
}

constexpr int SP_THREADS = 256;
constexpr int SP_NV = 4;  // columns per pass

// W[:, s0..s0+3] = diag X + A X for the rows of this group (LN lanes per row)
template <int LN>
__global__ void __launch_bounds__(SP_THREADS) stored_spmm_kernel(int64_t n, int nvec, const int64_t* __restrict__ rowptr,
                                                                  const int32_t* __restrict__ col,
                                                                  const double* __restrict__ val,
                                                                  const double* __restrict__ diag,
                                                                  const double* __restrict__ X, int64_t ldx,
                                                                  double* __restrict__ W, int64_t ldw) {
  const int lane = threadIdx.x % LN;
  const int64_t rows_per_grid = (int64_t)gridDim.x * (SP_THREADS / LN);
  const int s0 = blockIdx.y * SP_NV;
  const int nv = min(SP_NV, nvec - s0);
  for (int64_t r = (int64_t)blockIdx.x * (SP_THREADS / LN) + threadIdx.x / LN; r < n; r += rows_per_grid) {
    double acc[SP_NV] = {0.0, 0.0, 0.0, 0.0};
    const int64_t e = rowptr[r + 1];
    for (int64_t k = rowptr[r] + lane; k < e; k += LN) {
      const int64_t c = col[k];
      const double h = val[k];
#pragma unroll
      for (int q = 0; q < SP_NV; ++q)
        if (q < nv) acc[q] += h * X[c + (int64_t)(s0 + q) * ldx];
    }
#pragma unroll
    for (int q = 0; q < SP_NV; ++q)
#pragma unroll
      for (int off = LN / 2; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, LN);
    if (lane == 0) {
      const double d = diag[r];
#pragma unroll
      for (int q = 0; q < SP_NV; ++q)
        if (q < nv) W[r + (int64_t)(s0 + q) * ldw] = d * X[r + (int64_t)(s0 + q) * ldx] + acc[q];
    }
  }
}

}  // namespace

void Handle::create_stored(const char* dia_path, const char* off_path, int dev) {
  ARG_CHECK(dia_path && off_path, "file names must not be NULL");
  device = dev;
  HIP_CHECK(hipSetDevice(device));
  HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  stored = true;
  std::vector<char> rec;
  // hmlt.dia: (maxbl, nbuf), then H_jj
  FILE* f = std::fopen(dia_path, "rb");
  if (!f) throw Error(ADC2G_EARG, std::string("cannot open ") + dia_path);
  int32_t maxbl = 0, nbuf = 0;
  std::vector<double> hdiag;
  try {
    if (!read_record(f, rec, dia_path) || rec.size() != 8)
      throw Error(ADC2G_EARG, std::string(dia_path) + ": first record is not (maxbl, nbuf)");
    std::memcpy(&maxbl, rec.data(), 4);
    std::memcpy(&nbuf, rec.data() + 4, 4);
    if (!read_record(f, rec, dia_path) || rec.empty() || rec.size() % 8 != 0)
      throw Error(ADC2G_EARG, std::string(dia_path) + ": second record is not the diagonal");
    hdiag.resize(rec.size() / 8);
    std::memcpy(hdiag.data(), rec.data(), rec.size());
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  ARG_CHECK(maxbl >= 1 && nbuf >= 0, "hmlt.dia: maxbl must be >= 1 and nbuf >= 0");
  const int64_t n = (int64_t)hdiag.size();
  ARG_CHECK(n < ((int64_t)1 << 31), "stored matrices are limited to 2^31 - 1 rows");
  d = Dims{};
  d.ndim = n;
  d.off[5] = n;
  // hmlt.off: nbuf records of (buff(maxbl), oi(maxbl), oj(maxbl), nlim)
  std::vector<double> hv;
  std::vector<int32_t> oi, oj;
  f = std::fopen(off_path, "rb");
  if (!f) throw Error(ADC2G_EARG, std::string("cannot open ") + off_path);
  try {
    const size_t want = (size_t)maxbl * 16 + 4;
    for (int32_t b = 0; b < nbuf; ++b) {
      if (!read_record(f, rec, off_path)) throw Error(ADC2G_EARG, std::string(off_path) + ": fewer records than nbuf");
      if (rec.size() != want) throw Error(ADC2G_EARG, std::string(off_path) + ": record size does not match maxbl");
      int32_t nlim = 0;
      std::memcpy(&nlim, rec.data() + (size_t)maxbl * 16, 4);
      if (nlim < 0 || nlim > maxbl) throw Error(ADC2G_EARG, std::string(off_path) + ": nlim out of range");
      const size_t b0 = hv.size();
      hv.resize(b0 + nlim);
      oi.resize(b0 + nlim);
      oj.resize(b0 + nlim);
      std::memcpy(hv.data() + b0, rec.data(), (size_t)nlim * 8);
      std::memcpy(oi.data() + b0, rec.data() + (size_t)maxbl * 8, (size_t)nlim * 4);
      std::memcpy(oj.data() + b0, rec.data() + (size_t)maxbl * 12, (size_t)nlim * 4);
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  // CSR with both (i, j) and (j, i) per element (0-based)
  const size_t ne = hv.size();
  std::vector<int64_t> rp((size_t)n + 1, 0);
  for (size_t k = 0; k < ne; ++k) {
    if (oi[k] < 1 || oi[k] > n || oj[k] < 1 || oj[k] > n)
      throw Error(ADC2G_EARG, std::string(off_path) + ": element index out of range");
    ++rp[(size_t)oi[k]];
    ++rp[(size_t)oj[k]];
  }
  for (int64_t r = 0; r < n; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
  std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<int32_t> cc(2 * ne);
  std::vector<double> vv(2 * ne);
  for (size_t k = 0; k < ne; ++k) {
    const int64_t i = oi[k] - 1, j = oj[k] - 1;
    int64_t p = fill[(size_t)i]++;
    cc[(size_t)p] = (int32_t)j;
    vv[(size_t)p] = hv[k];
    p = fill[(size_t)j]++;
    cc[(size_t)p] = (int32_t)i;
    vv[(size_t)p] = hv[k];
  }
  stored_noffd = (int64_t)ne;
  stored_per_row = n > 0 ? 2.0 * (double)ne / (double)n : 0.0;
  hii.alloc((size_t)n);
  st_rowptr.alloc((size_t)n + 1);
  st_col.alloc(std::max<size_t>(1, 2 * ne));
  st_val.alloc(std::max<size_t>(1, 2 * ne));
  HIP_CHECK(hipMemcpyAsync(hii.get(), hdiag.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(st_rowptr.get(), rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  if (ne > 0) {
    HIP_CHECK(hipMemcpyAsync(st_col.get(), cc.data(), sizeof(int32_t) * 2 * ne, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(st_val.get(), vv.data(), sizeof(double) * 2 * ne, hipMemcpyHostToDevice, st));
  }
  HIP_CHECK(hipStreamSynchronize(st));
}

void Handle::sigma_stored(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, hipStream_t s) {
  const int64_t n = d.ndim;
  const int yb = (nvec + SP_NV - 1) / SP_NV;
  ARG_CHECK(yb <= 65535, "too many columns for one stored-matrix product");
  if (stored_per_row > 24.0) {
    constexpr int LN = 32;
    const int gx = grid_for(n, SP_THREADS / LN, 8192);
    hipLaunchKernelGGL(stored_spmm_kernel<LN>, dim3(gx, yb), dim3(SP_THREADS), 0, s, n, nvec, st_rowptr.get(),
                       st_col.get(), st_val.get(), hii.get(), X, ldx, W, ldw);
  } else {
    constexpr int LN = 8;
    const int gx = grid_for(n, SP_THREADS / LN, 8192);
    hipLaunchKernelGGL(stored_spmm_kernel<LN>, dim3(gx, yb), dim3(SP_THREADS), 0, s, n, nvec, st_rowptr.get(),
                       st_col.get(), st_val.get(), hii.get(), X, ldx, W, ldw);
  }
  HIP_CHECK(hipGetLastError());
}

}  // namespace adc2g
