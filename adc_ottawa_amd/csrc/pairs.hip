// Tiled CSF <-> pair-tensor kernels (replace the one-element-per-thread
// versions in kernels.hip; same maps, see the header there).
//
// One workgroup per (virtual index a, vector).  The t2/t4/t5 configurations
// of the virtual pairs (a, b), b > a, occupy consecutive pair numbers
// pv(a, b), so the workgroup walks b in chunks of BB:
//   pack   (CSF -> Z): stage x4/x5 at [q][pv(a, b)] and x2 at [pv(a,b)][k]
//          (contiguous runs over b) in LDS, then write the o x o tiles
//          Z(a,b,.,.) (rows a*o+k: contiguous over (b, l)) and Z(b,a,.,.)
//          (rows b*o+k: contiguous over l);
//   unpack (G -> CSF): stage the tiles G[(a,.),(b,.)] and G[(b,.),(a,.)] in
//          LDS, then write t4/t5 (contiguous over b for each occupied pair q)
//          and t2 (contiguous over (b, k)), adding the 2h2p diagonal times x.
// The diagonal tile b = a (t1, t3) is handled first by the same workgroup.
#include "kernels.hpp"

namespace adc2g {

namespace {

constexpr int PT = 256;
constexpr double SQRT2 = 1.4142135623730951;
constexpr double INV_SQRT2 = 0.7071067811865476;
constexpr double SQRT3_2 = 1.2247448713915890;

__device__ __forceinline__ int64_t pidx(int64_t lo, int64_t hi, int64_t n) {
  return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
}

template <class T>
__device__ __forceinline__ T cvt(double x) { return (T)x; }
template <>
__device__ __forceinline__ __bf16 cvt<__bf16>(double x) { return (__bf16)(float)x; }
template <class T>
__device__ __forceinline__ double dbl(T x) { return (double)x; }
template <>
__device__ __forceinline__ double dbl<__bf16>(__bf16 x) { return (double)(float)x; }

template <class T>
__global__ void __launch_bounds__(PT) pack_tiled_kernel(const double* __restrict__ X, int64_t ldx, DimsK d,
                                                         T* __restrict__ Z, PairLayout lay, int BB) {
  extern __shared__ double sm[];
  const int o = d.o, v = d.v, tid = threadIdx.x;
  const int a = blockIdx.x;
  const int chunk = blockIdx.y;
  const int64_t vec = blockIdx.z;
  const int b0 = a + 1 + chunk * BB;
  if (b0 >= v && chunk > 0) return;  // whole workgroup, before any barrier
  const int64_t n1 = d.n1, npo = d.npo, npv = d.npv;
  const double* x = X + ldx * vec;
  T* z = Z + vec * lay.sv;
  const int64_t sc = lay.sc;
  const int BP = BB + 1;            // padded run length (LDS banks differ across q)
  double* xs4 = sm;                 // [npo][BP]
  double* xs5 = xs4 + npo * BP;     // [npo][BP]
  double* xs2 = xs5 + npo * BP;     // [BB][o]
  // diagonal tile b = a (first chunk)
  if (chunk == 0) {
    for (int t = tid; t < o * o; t += PT) {
      const int k = t / o, l = t - k * o;
      double val;
      if (k == l) val = -SQRT2 * x[d.off1 + (int64_t)k * v + a];
      else val = x[d.off3 + pidx(min(k, l), max(k, l), o) * v + a];
      z[(int64_t)a * sc + (int64_t)k * n1 + (int64_t)a * o + l] = cvt<T>(val);
    }
  }
  if (b0 < v) {
    const int nb = min(BB, v - b0);
    const int64_t p0 = pidx(a, b0, v);  // pv(a, b0); pv(a, b0 + bb) = p0 + bb
    for (int t = tid; t < npo * nb; t += PT) {
      const int q = t / nb, bb = t - q * nb;
      const int64_t J = (int64_t)q * npv + p0 + bb;
      xs4[q * BP + bb] = INV_SQRT2 * x[d.off4 + J];
      xs5[q * BP + bb] = SQRT3_2 * x[d.off5 + J];
    }
    for (int t = tid; t < nb * o; t += PT) {
      const int bb = t / o, k = t - bb * o;
      xs2[bb * o + k] = -x[d.off2 + (p0 + bb) * o + k];
    }
    __syncthreads();
    // Z(a, b, k, l) (rows a*o + k, columns b*o + l) and Z(b, a, k, l) (rows
    // b*o + k, columns a*o + l): each thread keeps one (k, l) and walks the b
    // of the chunk, so the index decoding is done once per thread
    for (int kl = tid; kl < o * o; kl += PT) {
      const int k = kl / o, l = kl - k * o;
      const bool diag = (k == l);
      const int q = diag ? 0 : (int)pidx(min(k, l), max(k, l), o);
      const double s_ab = (k < l) ? 1.0 : -1.0;  // sign of x5 in Z(a,b,k,l); Z(b,a,k,l) takes the other
      T* zab = z + (int64_t)a * sc + (int64_t)k * n1 + (int64_t)b0 * o + l;
      T* zba = z + (int64_t)b0 * sc + (int64_t)k * n1 + (int64_t)a * o + l;
      for (int bb = 0; bb < nb; ++bb) {
        double vab, vba;
        if (diag) {
          vab = vba = xs2[bb * o + k];
        } else {
          const double p4 = xs4[q * BP + bb], p5 = xs5[q * BP + bb];
          vab = p4 + s_ab * p5;
          vba = p4 - s_ab * p5;
        }
        zab[(int64_t)bb * o] = cvt<T>(vab);
        zba[(int64_t)bb * sc] = cvt<T>(vba);
      }
    }
  }
}

template <class T>
__global__ void __launch_bounds__(PT) unpack_tiled_kernel(const T* __restrict__ G, const double* __restrict__ X,
                                                           int64_t ldx, double* __restrict__ W, int64_t ldw, DimsK d,
                                                           const double* __restrict__ eps, PairLayout lay, int BB,
                                                           int op) {
  extern __shared__ double sm[];
  const int o = d.o, v = d.v, tid = threadIdx.x;
  const int a = blockIdx.x;
  const int chunk = blockIdx.y;
  const int64_t vec = blockIdx.z;
  const int b0 = a + 1 + chunk * BB;
  if (b0 >= v && chunk > 0) return;  // whole workgroup, before any barrier
  const int64_t n1 = d.n1, npo = d.npo, npv = d.npv;
  const T* g = G + vec * lay.sv;
  const int64_t sc = lay.sc;
  const double* x = X + ldx * vec;
  double* w = W + ldw * vec;
  const int oo = o * o;
  // op: tile stride, o*o + 1 when that fits the LDS budget (LDS banks then
  // differ across b)
  double* tab = sm;            // [BB][o*o + 1]  G[(a,k),(b,l)]
  double* tba = sm + BB * op;  // [BB][o*o + 1]  G[(b,k),(a,l)]
  const double ea = eps[o + a];
  // diagonal tile b = a: t1 and t3 (first chunk)
  if (chunk == 0) {
    for (int t = tid; t < oo; t += PT) tab[t] = dbl(g[(int64_t)a * sc + (int64_t)(t / o) * n1 + (int64_t)a * o + t % o]);
    __syncthreads();
    for (int k = tid; k < o; k += PT) {
      const int64_t J = d.off1 + (int64_t)k * v + a;
      w[J] = -SQRT2 * tab[k * o + k] + (2.0 * ea - 2.0 * eps[k]) * x[J];
    }
    for (int q = tid; q < npo; q += PT) {
      // decode q -> (k, l) by walking the rows (npo is small)
      int k = 0;
      int64_t base = 0;
      while (base + (o - 1 - k) <= q) {
        base += o - 1 - k;
        ++k;
      }
      const int l = (int)(q - base) + k + 1;
      const int64_t J = d.off3 + (int64_t)q * v + a;
      w[J] = tab[k * o + l] + tab[l * o + k] + (2.0 * ea - eps[k] - eps[l]) * x[J];
    }
  }
  if (b0 < v) {
    const int nb = min(BB, v - b0);
    const int64_t p0 = pidx(a, b0, v);
    __syncthreads();
    for (int kl = tid; kl < oo; kl += PT) {
      const int k = kl / o, l = kl - k * o;
      const T* gab = g + (int64_t)a * sc + (int64_t)k * n1 + (int64_t)b0 * o + l;
      const T* gba = g + (int64_t)b0 * sc + (int64_t)k * n1 + (int64_t)a * o + l;
      for (int bb = 0; bb < nb; ++bb) {
        tab[bb * op + kl] = dbl(gab[(int64_t)bb * o]);
        tba[bb * op + kl] = dbl(gba[(int64_t)bb * sc]);
      }
    }
    __syncthreads();
    // t2 (a<b, k=l): contiguous over (b, k)
    for (int t = tid; t < nb * o; t += PT) {
      const int bb = t / o, k = t - bb * o;
      const int64_t J = d.off2 + (p0 + bb) * o + k;
      const double D = ea + eps[o + b0 + bb] - 2.0 * eps[k];
      w[J] = -(tab[bb * op + k * o + k] + tba[bb * op + k * o + k]) + D * x[J];
    }
    // t4, t5 (a<b, k<l): contiguous over b for each occupied pair q; the
    // pair (k, l) of q is decoded once per thread and step of q
    for (int t = tid; t < npo * nb; t += PT) {
      const int q = t / nb, bb = t - q * nb;
      int k = 0;
      int64_t base = 0;
      while (base + (o - 1 - k) <= q) {
        base += o - 1 - k;
        ++k;
      }
      const int l = (int)(q - base) + k + 1;
      const double* A = tab + bb * op;
      const double* Bt = tba + bb * op;
      const double g1 = A[k * o + l], g3 = A[l * o + k], g2 = Bt[k * o + l], g4 = Bt[l * o + k];
      const double D = ea + eps[o + b0 + bb] - eps[k] - eps[l];
      const int64_t J = (int64_t)q * npv + p0 + bb;
      w[d.off4 + J] = INV_SQRT2 * (g1 + g2 + g3 + g4) + D * x[d.off4 + J];
      w[d.off5 + J] = SQRT3_2 * (g1 - g2 - g3 + g4) + D * x[d.off5 + J];
    }
  }
}

}  // namespace

int pack_chunk(const DimsK& d) {
  const int64_t per = 2 * d.npo + d.o;
  int bb = (int)(4096 / std::max<int64_t>(per, 1));
  return std::max(1, std::min(bb, 16));
}

int unpack_chunk(const DimsK& d) {
  int bb = 2048 / std::max(1, d.o * d.o);
  return std::max(1, std::min(bb, 16));
}

template <class T>
void launch_pack_tiled(const double* X, int64_t ldx, int nvec, const DimsK& d, T* Z, const PairLayout& lay,
                       hipStream_t st) {
  const int BB = pack_chunk(d);
  const size_t lds = sizeof(double) * ((size_t)(BB + 1) * 2 * d.npo + (size_t)BB * d.o);
  const int chunks = (d.v - 1 + BB - 1) / BB;
  hipLaunchKernelGGL((pack_tiled_kernel<T>), dim3(d.v, std::max(chunks, 1), nvec), dim3(PT), lds, st, X, ldx, d, Z, lay, BB);
  HIP_CHECK(hipGetLastError());
}

template <class T>
void launch_unpack_tiled(const T* G, const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, const DimsK& d,
                         const double* eps, const PairLayout& lay, hipStream_t st) {
  const int BB = unpack_chunk(d);
  int op = d.o * d.o + 1;
  if ((size_t)2 * BB * op * sizeof(double) > 65536) op = d.o * d.o;
  const size_t lds = sizeof(double) * (size_t)2 * BB * op;
  const int chunks = (d.v - 1 + BB - 1) / BB;
  hipLaunchKernelGGL((unpack_tiled_kernel<T>), dim3(d.v, std::max(chunks, 1), nvec), dim3(PT), lds, st, G, X, ldx, W,
                     ldw, d, eps, lay, BB, op);
  HIP_CHECK(hipGetLastError());
}

#define INST(T)                                                                                           \
  template void launch_pack_tiled<T>(const double*, int64_t, int, const DimsK&, T*, const PairLayout&,      \
                                     hipStream_t);                                                          \
  template void launch_unpack_tiled<T>(const T*, const double*, int64_t, double*, int64_t, int, const DimsK&, \
                                       const double*, const PairLayout&, hipStream_t);
INST(double)
INST(float)
INST(__bf16)
#undef INST

}  // namespace adc2g
