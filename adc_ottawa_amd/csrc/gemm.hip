// Instantiations and launcher of the MFMA GEMM (see gemm.hpp).
#include "gemm.hpp"

namespace adc2g {

namespace {

template <class T, int BM, int BN, int BK>
void launch(const GemmArgs& g, hipStream_t st) {
  const int tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  dim3 grid(tiles, g.batch, g.splitk);
  if (g.ev_begin) HIP_CHECK(hipEventRecord(g.ev_begin, st));
  if (g.tag == 1)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, double, 1>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, double, 0>), grid, dim3(256), 0, st, g);
  HIP_CHECK(hipGetLastError());
  if (g.ev_end) HIP_CHECK(hipEventRecord(g.ev_end, st));
}

template <class T>
void dispatch(GemmArgs g, GemmWorkspace& ws, hipStream_t st, int splitk_hint) {
  typedef typename Mma<T>::Acc Acc;
  constexpr int BK = (sizeof(T) == 8) ? 16 : (sizeof(T) == 4 ? 32 : 64);
  const bool small = (g.M <= 64 || g.N <= 64);
  const int BMs = small ? 64 : 128;
  const int tiles = ((g.M + BMs - 1) / BMs) * ((g.N + BMs - 1) / BMs) * g.batch;
  const int ktiles = (g.K + BK - 1) / BK;
  int split = splitk_hint;
  if (split <= 0) {
    split = 1;
    const int target = 1024;  // workgroups to fill 256 CUs with some slack
    if (tiles < target && ktiles >= 8) {
      split = (target + tiles - 1) / tiles;
      split = std::min(split, ktiles / 4);
      split = std::max(split, 1);
    }
  }
  // normalise so that no split is empty
  if (split > 1) {
    const int per = (ktiles + split - 1) / split;
    split = (ktiles + per - 1) / per;
  }
  g.splitk = split;
  const bool need_partial = (split > 1 || g.map.kind != MAP_STD || g.out_prec != 0);
  if (need_partial) {
    const size_t need = (size_t)g.batch * split * (size_t)g.M * g.N * sizeof(Acc);
    ws.buf.alloc(need);
    g.partial = ws.buf.get();
  }
  if (small)
    launch<T, 64, 64, BK>(g, st);
  else
    launch<T, 128, 128, BK>(g, st);
  if (need_partial) {
    const int64_t total = (int64_t)g.M * g.N * g.batch;
    if (split >= 8 && total <= (int64_t)1 << 22) {
      const int wb = (int)((total + 63) / 64);
      if (g.out_prec == 1)
        hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, float>), dim3(wb), dim3(256), 0, st, g);
      else if (g.out_prec == 2)
        hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, __bf16>), dim3(wb), dim3(256), 0, st, g);
      else
        hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, double>), dim3(wb), dim3(256), 0, st, g);
      HIP_CHECK(hipGetLastError());
      return;
    }
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    if (g.out_prec == 1)
      hipLaunchKernelGGL((splitk_reduce_kernel<Acc, float>), dim3(blocks), dim3(256), 0, st, g);
    else if (g.out_prec == 2)
      hipLaunchKernelGGL((splitk_reduce_kernel<Acc, __bf16>), dim3(blocks), dim3(256), 0, st, g);
    else
      hipLaunchKernelGGL((splitk_reduce_kernel<Acc, double>), dim3(blocks), dim3(256), 0, st, g);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace

void gemm(GemmArgs g, Prec prec, GemmWorkspace& ws, hipStream_t st, int splitk_hint) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  if (g.K <= 0) {
    throw Error(ADC2G_EARG, "gemm with K <= 0");
  }
  switch (prec) {
    case Prec::F64: dispatch<double>(g, ws, st, splitk_hint); break;
    case Prec::F32: dispatch<float>(g, ws, st, splitk_hint); break;
    case Prec::BF16: dispatch<__bf16>(g, ws, st, splitk_hint); break;
  }
}

}  // namespace adc2g
