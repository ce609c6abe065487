// Streaming kernels: factor extraction, CSF <-> pair-tensor maps with the
// fused 2h2p diagonal, the full diagonal, and the ph-ph build operands.
//
// Configuration order (source/adclib/select_fano.f90:539-718), 0-based:
//   t1 (a=b, k=l):  index k*v + a
//   t2 (a<b, k=l):  index pv(a,b)*o + k
//   t3 (a=b, k<l):  index po(k,l)*v + a
//   t4, t5 (a<b, k<l): index po(k,l)*npv + pv(a,b)
// with pv/po the lexicographic pair numbers of select_fano's nested loops.
//
// Pair tensor Z(c,b,k,l) (stored as Z[(c*o+k)*n1 + b*o+l] per vector) of the
// closed-shell amplitudes, chosen so that the five coupling functions of
// source/adclib/adc_ph.f90:974-1073 collapse into
//   sigma1(a,k) = sum_{bcl} (ac|lb) Z(c,b,k,l) - sum_{bml} (mk|lb) Z(a,b,m,l):
//   t1 -> Z(a,a,k,k) = -sqrt2 x;  t2 -> Z(a,b,k,k) = Z(b,a,k,k) = -x;
//   t3 -> Z(a,a,k,l) = Z(a,a,l,k) = x;
//   t4/t5 -> Z = x4/sqrt2 +- sqrt(3/2) x5, '+' when (c<b) == (k<l).
// pair_to_csf applies the transpose of this map to G(c,b,k,l).
#include "kernels.hpp"

namespace adc2g {

namespace {

constexpr double SQRT2 = 1.4142135623730951;
constexpr double INV_SQRT2 = 0.7071067811865476;
constexpr double SQRT3_2 = 1.2247448713915890;  // sqrt(3/2)

__device__ __forceinline__ int64_t pair_index(int64_t lo, int64_t hi, int64_t n) {
  return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
}

// inverse of pair_index: the largest lo with pair_index(lo, lo+1, n) <= p
__device__ __forceinline__ void pair_decode(int64_t p, int64_t n, int& lo, int& hi) {
  // start from the closed form, then correct rounding
  const double nn = (double)(2 * n - 1);
  int64_t a = (int64_t)floor((nn - sqrt(nn * nn - 8.0 * (double)p)) * 0.5);
  if (a < 0) a = 0;
  while (a > 0 && pair_index(a, a + 1, n) > p) --a;
  while (a + 1 < n - 1 && pair_index(a + 1, a + 2, n) <= p) ++a;
  lo = (int)a;
  hi = (int)(p - pair_index(a, a + 1, n) + a + 1);
}

__global__ void extract_factors_kernel(const double* __restrict__ B, int nbas, DimsK d, double* Bov,
                                       double* Bvv, double* Boo) {
  const int o = d.o, v = d.v;
  const int64_t per = (int64_t)nbas * nbas;
  const int64_t total = per * d.naux;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t P = idx / per;
    const int64_t pq = idx - P * per;
    const int p = (int)(pq / nbas), q = (int)(pq - (int64_t)p * nbas);
    const double x = B[idx];
    if (p >= o && q < o) Bov[P * d.n1 + (int64_t)(p - o) * o + q] = x;
    if (p >= o && q >= o) Bvv[P * (int64_t)v * v + (int64_t)(p - o) * v + (q - o)] = x;
    if (p < o && q < o) Boo[P * (int64_t)o * o + (int64_t)p * o + q] = x;
  }
}

__global__ void convert_f32_kernel(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}

__global__ void convert_bf16_kernel(const double* __restrict__ s, __bf16* __restrict__ d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (__bf16)(float)s[i];
}

// dst[c * R + r] = src[r * C + c] through 32 x 32 LDS tiles, converting the
// element type on the way.
template <class S, class D>
__global__ void __launch_bounds__(256) transpose_kernel(const S* __restrict__ src, int64_t R, int64_t C,
                                                        D* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int64_t tiles_c = (C + 31) / 32;
  const int64_t r0 = (blockIdx.x / tiles_c) * 32, c0 = (blockIdx.x % tiles_c) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int64_t r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < R && c < C) ? (double)src[r * C + c] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t c = c0 + j, r = r0 + tx;
    if (r < R && c < C) dst[c * R + r] = (D)tile[tx][j];
  }
}

template <>
__global__ void __launch_bounds__(256) transpose_kernel<double, __bf16>(const double* __restrict__ src, int64_t R,
                                                                       int64_t C, __bf16* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int64_t tiles_c = (C + 31) / 32;
  const int64_t r0 = (blockIdx.x / tiles_c) * 32, c0 = (blockIdx.x % tiles_c) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int64_t r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < R && c < C) ? src[r * C + c] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t c = c0 + j, r = r0 + tx;
    if (r < R && c < C) dst[c * R + r] = (__bf16)(float)tile[tx][j];
  }
}

template <>
__global__ void __launch_bounds__(256) transpose_kernel<float, __bf16>(const float* __restrict__ src, int64_t R,
                                                                      int64_t C, __bf16* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int64_t tiles_c = (C + 31) / 32;
  const int64_t r0 = (blockIdx.x / tiles_c) * 32, c0 = (blockIdx.x % tiles_c) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int64_t r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < R && c < C) ? (double)src[r * C + c] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t c = c0 + j, r = r0 + tx;
    if (r < R && c < C) dst[c * R + r] = (__bf16)(float)tile[tx][j];
  }
}

template <class S, class D>
__global__ void swap01_kernel(const S* __restrict__ src, int64_t A, int64_t B, int64_t C, D* __restrict__ dst) {
  const int64_t total = A * B * C;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx % C, ji = idx / C;
    const int64_t i = ji % A, j = ji / A;
    dst[idx] = (D)src[(i * B + j) * C + c];
  }
}

template <class T>
__global__ void pack_x1_kernel(const double* __restrict__ X, int64_t ldx, int nvec, DimsK d, T* __restrict__ X1c) {
  const int64_t total = d.n1 * nvec;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    // output order: a outer, vec, k inner
    const int64_t k = idx % d.o;
    const int64_t rest = idx / d.o;
    const int64_t vec = rest % nvec, a = rest / nvec;
    X1c[idx] = (T)X[a * d.o + k + ldx * vec];
  }
}

template <class T>
__device__ __forceinline__ double dbl_of(T x) { return (double)x; }
template <>
__device__ __forceinline__ double dbl_of<__bf16>(__bf16 x) { return (double)(float)x; }

// G[(c*nc + vec)][n] (+)= -sum_m x1(c,m;vec) Ot[m][n], n < o*n1: one column
// per thread with its o values of Ot in registers, RB rows per workgroup with
// their x1 values staged in LDS (read as broadcasts).
constexpr int OCC_RB = 32;
template <class T, int OMAX>
__global__ void __launch_bounds__(256) occ_up_kernel(const double* __restrict__ X, int64_t ldx, int nc, DimsK d,
                                                     const T* __restrict__ Ot, T* __restrict__ G, int accumulate) {
  __shared__ double xs[OCC_RB * OMAX];
  const int o = d.o;
  const int64_t ncol = (int64_t)o * d.n1;
  const int64_t rows = (int64_t)d.v * nc;
  const int64_t r0 = (int64_t)blockIdx.y * OCC_RB;
  const int nr = (int)min((int64_t)OCC_RB, rows - r0);
  for (int t = threadIdx.x; t < nr * o; t += blockDim.x) {
    const int r = t / o, m = t - r * o;
    const int64_t mu = r0 + r, c = mu / nc, vec = mu - c * nc;
    xs[r * OMAX + m] = X[c * o + m + ldx * vec];
  }
  __syncthreads();
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ncol) return;
  double ot[OMAX];
#pragma unroll
  for (int m = 0; m < OMAX; ++m) ot[m] = (m < o) ? dbl_of(Ot[(int64_t)m * ncol + n]) : 0.0;
  for (int r = 0; r < nr; ++r) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < OMAX; ++m)
      if (m < o) acc += xs[r * OMAX + m] * ot[m];
    T* g = G + (r0 + r) * ncol + n;
    *g = accumulate ? (T)(dbl_of(*g) - acc) : (T)(-acc);
  }
}

template <class T>
__global__ void pack_x1t_kernel(const double* __restrict__ X, int64_t ldx, int nvec, DimsK d, T* __restrict__ X1t) {
  const int64_t total = d.n1 * nvec;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    // output order: vec outer, k, a inner
    const int64_t a = idx % d.v;
    const int64_t rest = idx / d.v;
    const int64_t k = rest % d.o, vec = rest / d.o;
    X1t[idx] = (T)X[a * d.o + k + ldx * vec];
  }
}

template <class T>
__global__ void csf_to_pair_kernel(const double* __restrict__ X, int64_t ldx, int nvec, DimsK d,
                                   T* __restrict__ Z, PairLayout lay) {
  const int64_t n1 = d.n1, nn = n1 * n1;
  const int o = d.o, v = d.v;
  const int64_t total = nn * nvec;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vec = idx / nn;
    const int64_t rem = idx - vec * nn;
    const int64_t ck = rem / n1, bl = rem - ck * n1;
    const int c = (int)(ck / o), k = (int)(ck - (int64_t)c * o);
    const int b = (int)(bl / o), l = (int)(bl - (int64_t)b * o);
    const double* x = X + ldx * vec;
    double val;
    if (c == b) {
      if (k == l) {
        val = -SQRT2 * x[d.off1 + (int64_t)k * v + c];
      } else {
        const int lo = min(k, l), hi = max(k, l);
        val = x[d.off3 + pair_index(lo, hi, o) * v + c];
      }
    } else {
      const int lo = min(c, b), hi = max(c, b);
      const int64_t pv = pair_index(lo, hi, v);
      if (k == l) {
        val = -x[d.off2 + pv * o + k];
      } else {
        const int64_t J = pair_index(min(k, l), max(k, l), o) * d.npv + pv;
        const double s4 = INV_SQRT2 * x[d.off4 + J];
        const double s5 = SQRT3_2 * x[d.off5 + J];
        val = ((c < b) == (k < l)) ? (s4 + s5) : (s4 - s5);
      }
    }
    Z[(int64_t)c * lay.sc + vec * lay.sv + (int64_t)k * n1 + bl] = (T)val;
  }
}

template <class T>
__global__ void pair_to_csf_kernel(const T* __restrict__ G, const double* __restrict__ X, int64_t ldx,
                                   double* __restrict__ W, int64_t ldw, int nvec, DimsK d,
                                   const double* __restrict__ eps, PairLayout lay) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1;
  const int64_t units = d.off5 - d.off1;  // t1..t4 entries; a t4 unit also writes its t5 twin
  const int64_t total = units * nvec;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vec = idx / units;
    const int64_t u = idx - vec * units;
    const T* g = G + vec * lay.sv;
    const double* x = X + ldx * vec;
    double* w = W + ldw * vec;
    const int64_t J = d.off1 + u;
    auto gm = [&](int c, int k, int b, int l) -> double {
      return (double)g[(int64_t)c * lay.sc + (int64_t)k * n1 + (int64_t)b * o + l];
    };
    if (J < d.off2) {  // t1
      const int64_t t = J - d.off1;
      const int k = (int)(t / v), a = (int)(t - (int64_t)k * v);
      const double D = 2.0 * eps[o + a] - 2.0 * eps[k];
      w[J] = -SQRT2 * gm(a, k, a, k) + D * x[J];
    } else if (J < d.off3) {  // t2
      const int64_t t = J - d.off2;
      const int64_t p = t / o;
      const int k = (int)(t - p * o);
      int a, b;
      pair_decode(p, v, a, b);
      const double D = eps[o + a] + eps[o + b] - 2.0 * eps[k];
      w[J] = -(gm(a, k, b, k) + gm(b, k, a, k)) + D * x[J];
    } else if (J < d.off4) {  // t3
      const int64_t t = J - d.off3;
      const int64_t q = t / v;
      const int a = (int)(t - q * v);
      int k, l;
      pair_decode(q, o, k, l);
      const double D = 2.0 * eps[o + a] - eps[k] - eps[l];
      w[J] = gm(a, k, a, l) + gm(a, l, a, k) + D * x[J];
    } else {  // t4 and its t5 twin
      const int64_t t = J - d.off4;
      const int64_t q = t / d.npv, p = t - q * d.npv;
      int a, b, k, l;
      pair_decode(p, v, a, b);
      pair_decode(q, o, k, l);
      const double D = eps[o + a] + eps[o + b] - eps[k] - eps[l];
      const double g1 = gm(a, k, b, l), g2 = gm(b, k, a, l), g3 = gm(a, l, b, k), g4 = gm(b, l, a, k);
      w[J] = INV_SQRT2 * (g1 + g2 + g3 + g4) + D * x[J];
      const int64_t J5 = d.off5 + t;
      w[J5] = SQRT3_2 * (g1 - g2 - g3 + g4) + D * x[J5];
    }
  }
}

__global__ void diagonal_kernel(const double* __restrict__ H11, const double* __restrict__ eps, DimsK d,
                                double* __restrict__ hii) {
  const int o = d.o, v = d.v;
  for (int64_t J = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; J < d.ndim;
       J += (int64_t)gridDim.x * blockDim.x) {
    double D;
    if (J < d.n1) {
      D = H11[J * d.n1 + J];
    } else if (J < d.off2) {
      const int64_t t = J - d.off1;
      const int k = (int)(t / v), a = (int)(t - (int64_t)k * v);
      D = 2.0 * eps[o + a] - 2.0 * eps[k];
    } else if (J < d.off3) {
      const int64_t t = J - d.off2;
      const int64_t p = t / o;
      const int k = (int)(t - p * o);
      int a, b;
      pair_decode(p, v, a, b);
      D = eps[o + a] + eps[o + b] - 2.0 * eps[k];
    } else if (J < d.off4) {
      const int64_t t = J - d.off3;
      const int64_t q = t / v;
      const int a = (int)(t - q * v);
      int k, l;
      pair_decode(q, o, k, l);
      D = 2.0 * eps[o + a] - eps[k] - eps[l];
    } else {
      const int64_t t = (J < d.off5) ? J - d.off4 : J - d.off5;
      const int64_t q = t / d.npv, p = t - q * d.npv;
      int a, b, k, l;
      pair_decode(p, v, a, b);
      pair_decode(q, o, k, l);
      D = eps[o + a] + eps[o + b] - eps[k] - eps[l];
    }
    hii[J] = D;
  }
}

// CA operands; row a, K = (u*o + w)*v + r (u, w occupied, r virtual):
//   Kt(a;u,w,r) = (au|rw) = M[(a,u),(r,w)],  Lt = 2 Kt(a;u,w,r) - Kt(a;w,u,r),
//   Dr = e_r + e_a - e_u - e_w.   CA = 1/2 [Kt/Dr | Kt] [Lt | Lt/Dr]^T.
__global__ void ca_operands_kernel(const double* __restrict__ Mm, const double* __restrict__ eps, DimsK d,
                                   double* __restrict__ X1, double* __restrict__ X2) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1, K0 = (int64_t)o * o * v;
  const int64_t total = (int64_t)v * K0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = idx / K0, K = idx - a * K0;
    const int r = (int)(K % v);
    const int64_t uw = K / v;
    const int w = (int)(uw % o), u = (int)(uw / o);
    const double kt = Mm[(a * o + u) * n1 + (int64_t)r * o + w];
    const double kt2 = Mm[(a * o + w) * n1 + (int64_t)r * o + u];
    const double lt = 2.0 * kt - kt2;
    const double Dr = eps[o + r] + eps[o + a] - eps[u] - eps[w];
    const int64_t row = a * 2 * K0;
    X1[row + K] = kt / Dr;
    X1[row + K0 + K] = kt;
    X2[row + K] = lt;
    X2[row + K0 + K] = lt / Dr;
  }
}

// CB operands; row k, K = (c*v + dd)*o + i:
//   P(k;c,d,i) = M[(c,k),(d,i)],  Q = 2 P(k;c,d,i) - P(k;d,c,i),
//   E = e_i - e_c - e_d + e_k.   CB = -1/2 [P/E | P] [Q | Q/E]^T.
__global__ void cb_operands_kernel(const double* __restrict__ Mm, const double* __restrict__ eps, DimsK d,
                                   double* __restrict__ X1, double* __restrict__ X2) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1, K0 = (int64_t)v * v * o;
  const int64_t total = (int64_t)o * K0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = idx / K0, K = idx - k * K0;
    const int i = (int)(K % o);
    const int64_t cd = K / o;
    const int dd = (int)(cd % v), c = (int)(cd / v);
    const double p = Mm[((int64_t)c * o + k) * n1 + (int64_t)dd * o + i];
    const double p2 = Mm[((int64_t)dd * o + k) * n1 + (int64_t)c * o + i];
    const double q = 2.0 * p - p2;
    const double E = eps[i] - eps[o + c] - eps[o + dd] + eps[k];
    const int64_t row = k * 2 * K0;
    X1[row + K] = p / E;
    X1[row + K0 + K] = p;
    X2[row + K] = q;
    X2[row + K0 + K] = q / E;
  }
}

// CC operands; row (a,k), K = c*o + i:
//   T = 2 M[(a,k),(c,i)] - M[(a,i),(c,k)],  D = e_a - e_k + e_c - e_i.
//   CC = -1/2 [T/D | T] [T | T/D]^T.
__global__ void cc_operands_kernel(const double* __restrict__ Mm, const double* __restrict__ eps, DimsK d,
                                   double* __restrict__ X1, double* __restrict__ X2) {
  const int o = d.o;
  const int64_t n1 = d.n1;
  const int64_t total = n1 * n1;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ak = idx / n1, ci = idx - ak * n1;
    const int64_t a = ak / o, k = ak - a * o;
    const int64_t c = ci / o, i = ci - c * o;
    const double t = 2.0 * Mm[ak * n1 + ci] - Mm[(a * o + i) * n1 + c * o + k];
    const double D = eps[o + a] - eps[k] + eps[o + c] - eps[i];
    const int64_t row = ak * 2 * n1;
    X1[row + ci] = t / D;
    X1[row + n1 + ci] = t;
    X2[row + ci] = t;
    X2[row + n1 + ci] = t / D;
  }
}

__global__ void phph_assemble_kernel(double* __restrict__ H11, const double* __restrict__ Mm,
                                     const double* __restrict__ Jm, const double* __restrict__ CA,
                                     const double* __restrict__ CB, const double* __restrict__ eps, DimsK d,
                                     int second_order) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1, total = n1 * n1;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / n1, j = idx - i * n1;
    const int64_t a = i / o, k = i - a * o;
    const int64_t a2 = j / o, k2 = j - a2 * o;
    double h = 2.0 * Mm[idx] - Jm[(a * v + a2) * (int64_t)o * o + k * o + k2];
    if (second_order) {
      h += H11[idx];  // CC on entry
      if (k == k2) h += CA[a * v + a2];
      if (a == a2) h += CB[k * o + k2];
    }
    if (i == j) h += eps[o + a] - eps[k];
    H11[idx] = h;
  }
}

}  // namespace

void launch_extract_factors(const double* B, int nbas, const DimsK& d, double* Bov, double* Bvv, double* Boo,
                            hipStream_t st) {
  const int64_t total = (int64_t)nbas * nbas * d.naux;
  hipLaunchKernelGGL(extract_factors_kernel, dim3(grid_for(total)), dim3(256), 0, st, B, nbas, d, Bov, Bvv, Boo);
  HIP_CHECK(hipGetLastError());
}

void launch_convert_f32(const double* src, float* dst, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(convert_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, src, dst, n);
  HIP_CHECK(hipGetLastError());
}

void launch_convert_bf16(const double* src, __bf16* dst, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(convert_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, st, src, dst, n);
  HIP_CHECK(hipGetLastError());
}

template <class S, class D>
void launch_transpose(const S* src, int64_t R, int64_t C, D* dst, hipStream_t st) {
  const int64_t tiles = ((R + 31) / 32) * ((C + 31) / 32);
  if (tiles <= 0) return;
  if (tiles > 0x7fffffff) throw Error(ADC2G_EARG, "transpose too large");
  hipLaunchKernelGGL((transpose_kernel<S, D>), dim3((unsigned)tiles), dim3(256), 0, st, src, R, C, dst);
  HIP_CHECK(hipGetLastError());
}
template void launch_transpose<double, double>(const double*, int64_t, int64_t, double*, hipStream_t);
template void launch_transpose<double, float>(const double*, int64_t, int64_t, float*, hipStream_t);
template void launch_transpose<double, __bf16>(const double*, int64_t, int64_t, __bf16*, hipStream_t);
template void launch_transpose<float, float>(const float*, int64_t, int64_t, float*, hipStream_t);
template void launch_transpose<float, __bf16>(const float*, int64_t, int64_t, __bf16*, hipStream_t);

template <class S, class D>
void launch_swap01(const S* src, int64_t A, int64_t B, int64_t C, D* dst, hipStream_t st) {
  hipLaunchKernelGGL((swap01_kernel<S, D>), dim3(grid_for(A * B * C)), dim3(256), 0, st, src, A, B, C, dst);
  HIP_CHECK(hipGetLastError());
}
template void launch_swap01<double, double>(const double*, int64_t, int64_t, int64_t, double*, hipStream_t);
template void launch_swap01<double, float>(const double*, int64_t, int64_t, int64_t, float*, hipStream_t);
template void launch_swap01<double, __bf16>(const double*, int64_t, int64_t, int64_t, __bf16*, hipStream_t);

template <class T>
void launch_pack_x1(const double* X, int64_t ldx, int nvec, const DimsK& d, T* X1c, hipStream_t st) {
  hipLaunchKernelGGL((pack_x1_kernel<T>), dim3(grid_for(d.n1 * nvec)), dim3(256), 0, st, X, ldx, nvec, d, X1c);
  HIP_CHECK(hipGetLastError());
}

template <class T>
bool launch_occ_up(const double* X, int64_t ldx, int nc, const DimsK& d, const T* Ot, T* G, bool accumulate,
                   hipStream_t st) {
  if (d.o > 32) return false;
  const int64_t ncol = (int64_t)d.o * d.n1, rows = (int64_t)d.v * nc;
  dim3 grid((unsigned)((ncol + 255) / 256), (unsigned)((rows + OCC_RB - 1) / OCC_RB));
  if (d.o <= 16)
    hipLaunchKernelGGL((occ_up_kernel<T, 16>), grid, dim3(256), 0, st, X, ldx, nc, d, Ot, G, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL((occ_up_kernel<T, 32>), grid, dim3(256), 0, st, X, ldx, nc, d, Ot, G, accumulate ? 1 : 0);
  HIP_CHECK(hipGetLastError());
  return true;
}

template <class T>
void launch_pack_x1t(const double* X, int64_t ldx, int nvec, const DimsK& d, T* X1t, hipStream_t st) {
  hipLaunchKernelGGL((pack_x1t_kernel<T>), dim3(grid_for(d.n1 * nvec)), dim3(256), 0, st, X, ldx, nvec, d, X1t);
  HIP_CHECK(hipGetLastError());
}

template <class T>
void launch_csf_to_pair(const double* X, int64_t ldx, int nvec, const DimsK& d, T* Z, const PairLayout& lay,
                        hipStream_t st) {
  if (d.o <= 64) {
    launch_pack_tiled<T>(X, ldx, nvec, d, Z, lay, st);
    return;
  }
  hipLaunchKernelGGL((csf_to_pair_kernel<T>), dim3(grid_for(d.n1 * d.n1 * nvec)), dim3(256), 0, st, X, ldx,
                     nvec, d, Z, lay);
  HIP_CHECK(hipGetLastError());
}

template <class T>
void launch_pair_to_csf(const T* G, const double* X, int64_t ldx, double* W, int64_t ldw, int nvec,
                        const DimsK& d, const double* eps, const PairLayout& lay, hipStream_t st) {
  if (d.o <= 64) {
    launch_unpack_tiled<T>(G, X, ldx, W, ldw, nvec, d, eps, lay, st);
    return;
  }
  const int64_t units = (d.off5 - d.off1) * nvec;
  hipLaunchKernelGGL((pair_to_csf_kernel<T>), dim3(grid_for(units)), dim3(256), 0, st, G, X, ldx, W, ldw, nvec,
                     d, eps, lay);
  HIP_CHECK(hipGetLastError());
}

void launch_diagonal(const double* H11, const double* eps, const DimsK& d, double* hii, hipStream_t st) {
  hipLaunchKernelGGL(diagonal_kernel, dim3(grid_for(d.ndim)), dim3(256), 0, st, H11, eps, d, hii);
  HIP_CHECK(hipGetLastError());
}

void launch_ca_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st) {
  const int64_t total = (int64_t)d.v * d.o * d.o * d.v;
  hipLaunchKernelGGL(ca_operands_kernel, dim3(grid_for(total)), dim3(256), 0, st, Mm, eps, d, X1, X2);
  HIP_CHECK(hipGetLastError());
}

void launch_cb_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st) {
  const int64_t total = (int64_t)d.o * d.v * d.v * d.o;
  hipLaunchKernelGGL(cb_operands_kernel, dim3(grid_for(total)), dim3(256), 0, st, Mm, eps, d, X1, X2);
  HIP_CHECK(hipGetLastError());
}

void launch_cc_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st) {
  hipLaunchKernelGGL(cc_operands_kernel, dim3(grid_for(d.n1 * d.n1)), dim3(256), 0, st, Mm, eps, d, X1, X2);
  HIP_CHECK(hipGetLastError());
}

void launch_phph_assemble(double* H11, const double* Mm, const double* Jm, const double* CA, const double* CB,
                          const double* eps, const DimsK& d, int second_order, hipStream_t st) {
  hipLaunchKernelGGL(phph_assemble_kernel, dim3(grid_for(d.n1 * d.n1)), dim3(256), 0, st, H11, Mm, Jm, CA, CB,
                     eps, d, second_order);
  HIP_CHECK(hipGetLastError());
}

#define INST(T)                                                                                             \
  template void launch_pack_x1<T>(const double*, int64_t, int, const DimsK&, T*, hipStream_t);               \
  template void launch_pack_x1t<T>(const double*, int64_t, int, const DimsK&, T*, hipStream_t);              \
  template bool launch_occ_up<T>(const double*, int64_t, int, const DimsK&, const T*, T*, bool, hipStream_t);    \
  template void launch_csf_to_pair<T>(const double*, int64_t, int, const DimsK&, T*, const PairLayout&,       \
                                      hipStream_t);                                                           \
  template void launch_pair_to_csf<T>(const T*, const double*, int64_t, double*, int64_t, int, const DimsK&, \
                                      const double*, const PairLayout&, hipStream_t);
INST(double)
INST(float)
INST(__bf16)
#undef INST

}  // namespace adc2g
