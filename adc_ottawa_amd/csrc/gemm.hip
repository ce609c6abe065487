// Instantiations and launcher of the MFMA GEMM (see gemm.hpp).
#include "gemm.hpp"

namespace adc2g {

namespace {

// Output element type of the direct map epilogues: the down contraction writes
// Y in the GEMM precision, the W build writes FP64 (FP64 GEMMs) or FP32.
template <class T> struct WOut { typedef float type; };
template <> struct WOut<double> { typedef double type; };

template <class T> constexpr int prec_code() { return sizeof(T) == 8 ? 0 : (sizeof(T) == 4 ? 1 : 2); }

// Maps with a direct (no split-K) epilogue instantiation, given the output type.
template <class T>
bool direct_map(const GemmArgs& g) {
  switch (g.map.kind) {
    case MAP_STD: return true;
    case MAP_DOWNY: return g.out_prec == prec_code<T>();
    case MAP_WUPV:
    case MAP_WUPO: return g.out_prec == (sizeof(T) == 8 ? 0 : 1);
    default: return false;
  }
}

// Threads per workgroup: FP64 128x128 tiles run 8 waves (each 64x32) so that
// two workgroups per CU keep four waves per SIMD; the rest run 4 waves.
template <class T, int BM, int BN> constexpr int threads_for() {
  return (sizeof(T) == 8 && ((BM == 128 && BN == 128) || BM == 32)) ? 512 : 256;
}

template <class T, int BM, int BN, int BK, class OutT, int TAG, int MK>
void launch_k(const GemmArgs& g, dim3 grid, hipStream_t st) {
  constexpr int NT = threads_for<T, BM, BN>();
  if (g.vec)
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, OutT, TAG, true, MK, NT>), grid, dim3(NT), 0, st, g);
  else
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, OutT, TAG, false, MK, NT>), grid, dim3(NT), 0, st, g);
}

template <class T, int BM, int BN, int BK>
void launch(const GemmArgs& g, hipStream_t st) {
  const int tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  dim3 grid(tiles, g.batch, g.splitk);
  const int mk = g.direct ? g.map.kind : MAP_STD;
  if (g.ev_begin) HIP_CHECK(hipEventRecord(g.ev_begin, st));
  switch (mk) {
    case MAP_DOWNY: launch_k<T, BM, BN, BK, T, 1, MAP_DOWNY>(g, grid, st); break;
    case MAP_WUPV: launch_k<T, BM, BN, BK, typename WOut<T>::type, 0, MAP_WUPV>(g, grid, st); break;
    case MAP_WUPO: launch_k<T, BM, BN, BK, typename WOut<T>::type, 0, MAP_WUPO>(g, grid, st); break;
    default:
      if (g.tag == 1) launch_k<T, BM, BN, BK, double, 1, MAP_STD>(g, grid, st);
      else launch_k<T, BM, BN, BK, double, 0, MAP_STD>(g, grid, st);
  }
  HIP_CHECK(hipGetLastError());
  if (g.ev_end) HIP_CHECK(hipEventRecord(g.ev_end, st));
}

// Split-K / batch partials: workspace for batch * splitk * M * N accumulators
template <class Acc>
void* partial_buffer(const GemmArgs& g, GemmWorkspace& ws) {
  const size_t need = (size_t)g.batch * g.splitk * (size_t)g.M * g.N * sizeof(Acc);
  if (ws.ext) {
    if (need > ws.ext_size)
      throw Error(ADC2G_ENOMEM, "GEMM partials need " + std::to_string(need) + " bytes; the scratch leaves " +
                                    std::to_string(ws.ext_size));
    return ws.ext;
  }
  ws.buf.alloc(need);
  return ws.buf.get();
}

// Fold the partials (over splits, and over batches when they are terms of one
// sum) through the map.
template <class Acc>
void reduce_partials(GemmArgs g, hipStream_t st) {
  if (g.reduce_batch) {  // the batch partials are further splits of one sum
    g.splitk *= g.batch;
    g.batch = 1;
  }
  const int64_t total = (int64_t)g.M * g.N * g.batch;
  if (g.splitk >= 8 && total <= (int64_t)1 << 22) {  // many splits, few outputs
    const int wb = (int)((total + 63) / 64);
    if (g.out_prec == 1)
      hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, float>), dim3(wb), dim3(1024), 0, st, g);
    else if (g.out_prec == 2)
      hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, __bf16>), dim3(wb), dim3(1024), 0, st, g);
    else
      hipLaunchKernelGGL((splitk_reduce_wide_kernel<Acc, double>), dim3(wb), dim3(1024), 0, st, g);
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

template <class T>
void dispatch(GemmArgs g, GemmWorkspace& ws, hipStream_t st, int splitk_hint) {
  typedef typename Mma<T>::Acc Acc;
  constexpr int BK = (sizeof(T) == 8) ? 16 : (sizeof(T) == 4 ? 32 : 64);
  // tile shape: 128x32 for N <= 32, 64x64 when M or N is small; otherwise
  // the shape with the least padded output (within 2%), larger tiles first
  static const int SBM[9] = {0, 128, 64, 128, 128, 64, 64, 32, 32};
  static const int SBN[9] = {0, 128, 64, 32, 64, 128, 208, 208, 128};
  int shape = g.tile;
  // the 208-column tiles load their operands along K only
  if ((shape == 6 || shape == 7) && !(g.sak == 1 && g.sbk == 1)) shape = 0;
  if (shape <= 0) {
    auto padded = [&](int sh) {
      const int bm = SBM[sh], bn = SBN[sh];
      return (double)((g.M + bm - 1) / bm * bm) * (double)((g.N + bn - 1) / bn * bn);
    };
    if (g.N <= 32 && g.M > 64) {
      shape = 3;
    } else if (g.M <= 64 || g.N <= 64) {
      shape = 2;
    } else {
      const int order[4] = {1, 4, 5, 2};
      double best = padded(1);
      for (int sh : order) best = std::min(best, padded(sh));
      for (int sh : order)
        if (padded(sh) <= 1.02 * best) {
          shape = sh;
          break;
        }
    }
  }
  // thin products (M <= 32 rows, e.g. the coupling contractions of one or two
  // vectors): 32-row tiles, so that no MFMA works on padding rows
  if (g.M <= 32) {
    if (shape == 6) shape = 7;
    else if (shape == 2 && g.N >= 128) shape = 8;
  }
  const int BMs = SBM[shape], BNs = SBN[shape];
  const int tiles = ((g.M + BMs - 1) / BMs) * ((g.N + BNs - 1) / BNs) * g.batch;
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
  auto aligned = [](const void* p, int64_t stride, int64_t bstride) {
    return ((uintptr_t)p % 16 == 0) && ((stride * (int64_t)sizeof(T)) % 16 == 0) &&
           ((bstride * (int64_t)sizeof(T)) % 16 == 0);
  };
  constexpr int EV = 16 / sizeof(T);
  g.vec = (g.sak == 1 && g.sbk == 1 && g.K % EV == 0 && aligned(g.A, g.sam, g.sab) && aligned(g.B, g.sbn, g.sbb))
              ? 1 : 0;
  // without split-K the map is applied in-kernel when an instantiation has it,
  // otherwise by the reduction kernel from the partial
  g.direct = (split == 1 && !g.reduce_batch && direct_map<T>(g)) ? 1 : 0;
  const bool need_partial = !g.direct;
  if (need_partial) g.partial = partial_buffer<Acc>(g, ws);
  switch (shape) {
    case 2: launch<T, 64, 64, BK>(g, st); break;
    case 3: launch<T, 128, 32, BK>(g, st); break;
    case 4: launch<T, 128, 64, BK>(g, st); break;
    case 5: launch<T, 64, 128, BK>(g, st); break;
    case 6: launch<T, 64, 208, BK>(g, st); break;
    case 7: launch<T, 32, 208, BK>(g, st); break;
    case 8: launch<T, 32, 128, BK>(g, st); break;
    default: launch<T, 128, 128, BK>(g, st);
  }
  if (need_partial) reduce_partials<Acc>(g, st);
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
