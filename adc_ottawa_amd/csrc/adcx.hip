// ADC(2)-x: the first-order 2h2p-2h2p block (get_offdiag_adc2ext_save,
// source/adclib/get_matrix.f90:2435-2893, element functions C_1_1 ..
// C_4ii_4ii of source/adclib/adc_ph.f90:1102-1645; diagonal of
// get_diag_adc2ext_direct) added to the ADC(2)-s sigma-vector.
//
// In pair space, with Z the closed-shell pair tensor of a CSF 2h2p vector
// (the t5 amplitudes weighted by 1/3, the norm ratio of the CSF states) and
// T = 2Z - XZ, (XZ)(c,b,k,l) = Z(b,c,k,l):
//   G(c,b,k,l) = 1/2 sum_de (cd|be) T(d,e,k,l) + 1/2 sum_mn (km|ln) T(c,b,m,n)
//              + sum_dm [2 (ck|dm) - (cd|km)] T(d,b,m,l)
//              - sum_dm [(cl|dm) T(d,b,m,k) + (cd|lm) T(d,b,k,m)],
// and the block is the CSF projection of G (again with 1/3 on t5).  As
// matrix products over n1 = o v and v^2 / o^2 pair indices:
//   G1[(c,k)][(b,l)] = A1 T1,          A1 = 2 (ck|dm) - (cd|km),  T1[(d,m)][(b,l)] = T(d,b,m,l)
//   G2[(c,l)][(b,k)] = OV T1 + VO T3,  T3[(d,m)][(b,k)] = T(d,b,k,m)
//   G3[(c,b)][(k,l)] = (VVVV T4 + T4 OOOO) / 2,  T4[(d,e)][(k,l)] = T(d,e,k,l)
//   G(c,b,k,l) = G1[(c,k),(b,l)] - G2[(c,l),(b,k)] + G3[(c,b),(k,l)].
// (T1 also serves the (cl|dm) term: T(d,b,m,k) = T(b,d,k,m), the pair
// symmetry of Z.)  One vector at a time, FP64; the vvvv matrix is v^4 doubles.
#include "handle.hpp"

#include <algorithm>
#include <cmath>

namespace adc2g {

namespace {

constexpr double AX_SQ2 = 1.4142135623730951;
constexpr double AX_ISQ2 = 0.7071067811865476;
constexpr double AX_SQ32 = 1.2247448713915890;

__device__ __forceinline__ int64_t ax_pair(int64_t lo, int64_t hi, int64_t n) {
  return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
}

// p -> (lo, hi), lo < hi < n, the inverse of ax_pair
__device__ __forceinline__ void ax_unpair(int64_t p, int64_t n, int& lo, int& hi) {
  const double b = 2.0 * (double)n - 1.0;
  int a = (int)floor(0.5 * (b - sqrt(fmax(b * b - 8.0 * (double)p, 0.0))));
  a = max(a, 0);
  while (a > 0 && (int64_t)a * (2 * n - a - 1) / 2 > p) --a;
  while ((int64_t)(a + 1) * (2 * n - a - 2) / 2 <= p) ++a;
  lo = a;
  hi = (int)(p - (int64_t)a * (2 * n - a - 1) / 2) + a + 1;
}

// Z(c,b,k,l) of the full vector x (t5 weighted by 1/3)
__device__ double ax_z(const double* __restrict__ x, const DimsK& d, int c, int b, int k, int l) {
  const int o = d.o, v = d.v;
  if (c == b) {
    if (k == l) return -AX_SQ2 * x[d.off1 + (int64_t)k * v + c];
    return x[d.off3 + ax_pair(min(k, l), max(k, l), o) * v + c];
  }
  const int64_t pab = ax_pair(min(c, b), max(c, b), v);
  if (k == l) return -x[d.off2 + pab * o + k];
  const int64_t q = ax_pair(min(k, l), max(k, l), o) * d.npv + pab;
  const double s4 = AX_ISQ2 * x[d.off4 + q], s5 = (AX_SQ32 / 3.0) * x[d.off5 + q];
  return ((c < b) == (k < l)) ? s4 + s5 : s4 - s5;
}

__device__ __forceinline__ double ax_t(const double* x, const DimsK& d, int c, int b, int k, int l) {
  return 2.0 * ax_z(x, d, c, b, k, l) - ax_z(x, d, b, c, k, l);
}

// the three layouts of T for the nb vectors of a chunk (vector j =
// blockIdx.y), stored transposed -- row (j, column index), contiguous over
// the contraction index -- so that one matrix product with both operands
// K-contiguous serves the whole chunk:
//   T1t[(j,b,l)][(d,m)] = T(d,b,m,l), T3t[(j,b,k)][(d,m)] = T(d,b,k,m),
//   T4t[(j,k,l)][(d,e)] = T(d,e,k,l), and T4n[(d,e)][(j,k,l)] = T(d,e,k,l)
//   for the product with the (oo|oo) matrix
__global__ void ax_layouts_kernel(const double* __restrict__ X, int64_t ldx, DimsK d, int nb,
                                  double* __restrict__ T1, double* __restrict__ T3, double* __restrict__ T4,
                                  double* __restrict__ T4n) {
  const int o = d.o, v = d.v;
  const int j = blockIdx.y;
  const double* x = X + ldx * j;
  const int64_t n1 = d.n1, nn = n1 * n1, oo = (int64_t)o * o;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nn;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx / n1, q = idx - r * n1;
    const int dd = (int)(r / o), m = (int)(r - (int64_t)dd * o);
    const int b = (int)(q / o), l = (int)(q - (int64_t)b * o);
    const int64_t o13 = ((int64_t)j * n1 + q) * n1 + r;
    T1[o13] = ax_t(x, d, dd, b, m, l);
    T3[o13] = ax_t(x, d, dd, b, l, m);
    const int64_t de = idx / oo, kl = idx - de * oo;
    const int d4 = (int)(de / v), e4 = (int)(de - (int64_t)d4 * v);
    const int k4 = (int)(kl / o), l4 = (int)(kl - (int64_t)k4 * o);
    const double t4 = ax_t(x, d, d4, e4, k4, l4);
    T4[((int64_t)j * oo + kl) * (int64_t)v * v + de] = t4;
    T4n[(de * nb + j) * oo + kl] = t4;
  }
}

// G(c,b,k,l) of vector j of a chunk of nb (the column-block layouts above)
__device__ __forceinline__ double ax_g(const double* G1, const double* G2, const double* G3, const DimsK& d, int nb,
                                       int j, int c, int b, int k, int l) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1, oo = (int64_t)o * o;
  return G1[(((int64_t)c * o + k) * nb + j) * n1 + (int64_t)b * o + l] -
         G2[(((int64_t)c * o + l) * nb + j) * n1 + (int64_t)b * o + k] +
         G3[(((int64_t)c * v + b) * nb + j) * oo + (int64_t)k * o + l];
}

// 2h2p configuration J (0-based within the 2h2p part) -> type 1..5 and (a, b, k, l)
__device__ __forceinline__ int ax_decode(int64_t J, const DimsK& d, int& a, int& b, int& k, int& l) {
  const int o = d.o, v = d.v;
  const int64_t c1 = (int64_t)o * v, c2 = d.npv * o, c3 = d.npo * v, c4 = d.npo * d.npv;
  if (J < c1) {
    k = l = (int)(J / v);
    a = b = (int)(J - (int64_t)k * v);
    return 1;
  }
  J -= c1;
  if (J < c2) {
    k = l = (int)(J % o);
    ax_unpair(J / o, v, a, b);
    return 2;
  }
  J -= c2;
  if (J < c3) {
    a = b = (int)(J % v);
    ax_unpair(J / v, o, k, l);
    return 3;
  }
  J -= c3;
  const int t = (J < c4) ? 4 : 5;
  if (t == 5) J -= c4;
  ax_unpair(J % d.npv, v, a, b);
  ax_unpair(J / d.npv, o, k, l);
  return t;
}

// w[off1 + J] += (M22 x)_J from G (the CSF projection, 1/3 on t5), vector
// j = blockIdx.y of the chunk
__global__ void ax_project_kernel(const double* __restrict__ G1, const double* __restrict__ G2,
                                  const double* __restrict__ G3, DimsK d, int nb, double* __restrict__ W,
                                  int64_t ldw) {
  const int64_t n2 = d.ndim - d.n1;
  const int j = blockIdx.y;
  double* w = W + ldw * j;
  for (int64_t J = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; J < n2; J += (int64_t)gridDim.x * blockDim.x) {
    int a, b, k, l;
    const int t = ax_decode(J, d, a, b, k, l);
    double val;
    if (t == 1) {
      val = -AX_SQ2 * ax_g(G1, G2, G3, d, nb, j, a, a, k, k);
    } else if (t == 2) {
      val = -(ax_g(G1, G2, G3, d, nb, j, a, b, k, k) + ax_g(G1, G2, G3, d, nb, j, b, a, k, k));
    } else if (t == 3) {
      val = ax_g(G1, G2, G3, d, nb, j, a, a, k, l) + ax_g(G1, G2, G3, d, nb, j, a, a, l, k);
    } else {
      const double g1 = ax_g(G1, G2, G3, d, nb, j, a, b, k, l), g2 = ax_g(G1, G2, G3, d, nb, j, b, a, k, l);
      const double g3 = ax_g(G1, G2, G3, d, nb, j, a, b, l, k), g4 = ax_g(G1, G2, G3, d, nb, j, b, a, l, k);
      val = (t == 4) ? AX_ISQ2 * (g1 + g2 + g3 + g4) : (AX_SQ32 / 3.0) * (g1 - g2 - g3 + g4);
    }
    w[d.off1 + J] += val;
  }
}

struct AxMats {
  const double *vvvv, *oooo, *ov, *vo, *a1;
};

// pair-space matrix element of the operator acting on T (the G formula)
__device__ double ax_kpp(const AxMats& M, const DimsK& d, int c, int b, int k, int l, int dd, int e, int m, int n) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1;
  double r = 0.0;
  if (k == m && l == n) r += 0.5 * M.vvvv[(((int64_t)c * v + b) * v + dd) * v + e];
  if (c == dd && b == e) r += 0.5 * M.oooo[(((int64_t)m * o + n) * o + k) * o + l];
  if (b == e) {
    if (n == l) r += M.a1[((int64_t)c * o + k) * n1 + (int64_t)dd * o + m];
    if (n == k) r -= M.ov[((int64_t)c * o + l) * n1 + (int64_t)dd * o + m];
    if (m == k) r -= M.vo[((int64_t)c * o + l) * n1 + (int64_t)dd * o + n];
  }
  return r;
}

// diagonal of the block: sum over the (at most 4) pair positions x, y of the
// configuration's pair tensor P of P[x] P[y] (2 K''[x,y] - K''[x,Xy])
__global__ void ax_diag_kernel(AxMats M, DimsK d, double* __restrict__ diag) {
  const int64_t n2 = d.ndim - d.n1;
  for (int64_t J = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; J < n2; J += (int64_t)gridDim.x * blockDim.x) {
    int a, b, k, l;
    const int t = ax_decode(J, d, a, b, k, l);
    int pc[4], pb[4], pk[4], pl[4];
    double pw[4];
    int np;
    double nt = 1.0;
    if (t == 1) {
      np = 1;
      pc[0] = a; pb[0] = a; pk[0] = k; pl[0] = k; pw[0] = -AX_SQ2;
    } else if (t == 2) {
      np = 2;
      pc[0] = a; pb[0] = b; pk[0] = k; pl[0] = k; pw[0] = -1.0;
      pc[1] = b; pb[1] = a; pk[1] = k; pl[1] = k; pw[1] = -1.0;
    } else if (t == 3) {
      np = 2;
      pc[0] = a; pb[0] = a; pk[0] = k; pl[0] = l; pw[0] = 1.0;
      pc[1] = a; pb[1] = a; pk[1] = l; pl[1] = k; pw[1] = 1.0;
    } else {
      np = 4;
      const double s = (t == 4) ? AX_ISQ2 : AX_SQ32;
      const double x = (t == 4) ? AX_ISQ2 : -AX_SQ32;
      if (t == 5) nt = 1.0 / 3.0;
      pc[0] = a; pb[0] = b; pk[0] = k; pl[0] = l; pw[0] = s;
      pc[1] = b; pb[1] = a; pk[1] = l; pl[1] = k; pw[1] = s;
      pc[2] = b; pb[2] = a; pk[2] = k; pl[2] = l; pw[2] = x;
      pc[3] = a; pb[3] = b; pk[3] = l; pl[3] = k; pw[3] = x;
    }
    double sum = 0.0;
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) {
        const double kk = 2.0 * ax_kpp(M, d, pc[i], pb[i], pk[i], pl[i], pc[j], pb[j], pk[j], pl[j]) -
                          ax_kpp(M, d, pc[i], pb[i], pk[i], pl[i], pb[j], pc[j], pk[j], pl[j]);
        sum += pw[i] * pw[j] * kk;
      }
    diag[J] = nt * nt * sum;
  }
}

// out[i0][k][j][i3] = in[i0][j][k][i3]: in of shape (n0, nb, na, n3), out of
// shape (n0, na, nb, n3) -- the two middle indices swapped
__global__ void ax_swap_mid_kernel(const double* __restrict__ in, int64_t n0, int64_t na, int64_t nb, int64_t n3,
                                   double* __restrict__ out) {
  const int64_t total = n0 * na * nb * n3;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = idx;
    const int64_t i3 = r % n3; r /= n3;
    const int64_t j = r % nb; r /= nb;   // output index order: [i][j'][k'][i3] with j' over nb
    const int64_t k = r % na; r /= na;
    const int64_t i0 = r;
    // idx = ((i0*na + k)*nb + j)*n3 + i3 is out[i0][k][j][i3] = in[i0][j][k][i3]
    out[idx] = in[((i0 * nb + j) * na + k) * n3 + i3];
  }
}

__global__ void ax_axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                                const double* __restrict__ y, double* __restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    z[i] = a * x[i] + b * y[i];
}

__global__ void ax_add_kernel(int64_t n, double s, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += s * x[i];
}

// C (M x N, ldc) = alpha A (M x K, lda) Bt^T + beta C with Bt N x K (ldbt): both
// operands contiguous along K (the vector loads and the wide tiles)
void ax_gemm_bt(Handle& H, int M, int N, int K, const double* A, int64_t lda, const double* Bt, int64_t ldbt,
                double* C, int64_t ldc, double alpha, double beta, hipStream_t s) {
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.sam = lda; g.sak = 1;
  g.B = Bt; g.sbk = 1; g.sbn = ldbt;
  g.map.C = C; g.map.scm = ldc; g.map.scn = 1; g.map.alpha = alpha; g.map.beta = beta;
  gemm(g, Prec::F64, H.gws, s);
}


}  // namespace

void Handle::enable_adcx(bool on) {
  ARG_CHECK(!parent, "ADC(2)-x is switched on the full operator, before its views are made");
  ARG_CHECK(nviews == 0, "switch ADC(2)-x before making views (their diagonal is taken at creation)");
  if (on == adcx) return;
  HIP_CHECK(hipSetDevice(device));
  const int o = d.o, v = d.v, naux = d.naux;
  const int64_t n1 = d.n1, n2 = d.ndim - d.n1, vv = (int64_t)v * v, oo = (int64_t)o * o;
  if (on && !ax_d2.get()) {
    // (ck|dm) and (cd|km) as n1 x n1 matrices, A1 = 2 OV - VO
    ax_ov.alloc((size_t)n1 * n1);
    ax_vo.alloc((size_t)n1 * n1);
    ax_a1.alloc((size_t)n1 * n1);
    {
      GemmArgs g;
      g.M = (int)n1; g.N = (int)n1; g.K = naux;
      g.A = Bov.get(); g.sam = 1; g.sak = n1;
      g.B = Bov.get(); g.sbk = n1; g.sbn = 1;
      g.map.C = ax_ov.get(); g.map.scm = n1; g.map.scn = 1;
      gemm(g, Prec::F64, gws, st);
    }
    {
      DevBuf<double> J;  // [(c,d)][(k,m)] = (cd|km)
      J.alloc((size_t)vv * oo);
      GemmArgs g;
      g.M = (int)vv; g.N = (int)oo; g.K = naux;
      g.A = Bvv.get(); g.sam = 1; g.sak = vv;
      g.B = Boo.get(); g.sbk = oo; g.sbn = 1;
      g.map.C = J.get(); g.map.scm = oo; g.map.scn = 1;
      gemm(g, Prec::F64, gws, st);
      // [c][d][k][m] -> [c][k][d][m]
      hipLaunchKernelGGL(ax_swap_mid_kernel, dim3(grid_for(vv * oo)), dim3(256), 0, st, J.get(), (int64_t)v,
                         (int64_t)o, (int64_t)v, (int64_t)o, ax_vo.get());
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(ax_axpby_kernel, dim3(grid_for(n1 * n1)), dim3(256), 0, st, n1 * n1, 2.0, ax_ov.get(), -1.0,
                       ax_vo.get(), ax_a1.get());
    HIP_CHECK(hipGetLastError());
    // (cd|be) as [(c,b)][(d,e)] and (km|ln) as [(m,n)][(k,l)]
    ax_vvvv.alloc((size_t)vv * vv);
    {
      DevBuf<double> W;  // [(c,d)][(b,e)]
      W.alloc((size_t)vv * vv);
      GemmArgs g;
      g.M = (int)vv; g.N = (int)vv; g.K = naux;
      g.A = Bvv.get(); g.sam = 1; g.sak = vv;
      g.B = Bvv.get(); g.sbk = vv; g.sbn = 1;
      g.map.C = W.get(); g.map.scm = vv; g.map.scn = 1;
      gemm(g, Prec::F64, gws, st);
      hipLaunchKernelGGL(ax_swap_mid_kernel, dim3(grid_for(vv * vv)), dim3(256), 0, st, W.get(), (int64_t)v,
                         (int64_t)v, (int64_t)v, (int64_t)v, ax_vvvv.get());
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(st));
    }
    ax_oooo.alloc((size_t)oo * oo);
    {
      DevBuf<double> Q;  // [(k,m)][(l,n)] = (km|ln)
      Q.alloc((size_t)oo * oo);
      GemmArgs g;
      g.M = (int)oo; g.N = (int)oo; g.K = naux;
      g.A = Boo.get(); g.sam = 1; g.sak = oo;
      g.B = Boo.get(); g.sbk = oo; g.sbn = 1;
      g.map.C = Q.get(); g.map.scm = oo; g.map.scn = 1;
      gemm(g, Prec::F64, gws, st);
      // [k][m][l][n] -> [m][n][k][l]: two steps, swap(m,l) then transpose pairs
      DevBuf<double> R;
      R.alloc((size_t)oo * oo);
      // R[k][l][m][n] = Q[k][m][l][n]
      hipLaunchKernelGGL(ax_swap_mid_kernel, dim3(grid_for(oo * oo)), dim3(256), 0, st, Q.get(), (int64_t)o,
                         (int64_t)o, (int64_t)o, (int64_t)o, R.get());
      HIP_CHECK(hipGetLastError());
      // ax_oooo[(m,n)][(k,l)] = R[(k,l)][(m,n)]: a transpose of the oo x oo matrix
      launch_transpose<double, double>(R.get(), oo, oo, ax_oooo.get(), st);
      HIP_CHECK(hipStreamSynchronize(st));
    }
    ax_d2.alloc((size_t)n2);
    AxMats M{ax_vvvv.get(), ax_oooo.get(), ax_ov.get(), ax_vo.get(), ax_a1.get()};
    hipLaunchKernelGGL(ax_diag_kernel, dim3(grid_for(n2)), dim3(256), 0, st, M, dk, ax_d2.get());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
  }
  // the ADC(2)-x diagonal: hii of the 2h2p configurations plus the block's
  hipLaunchKernelGGL(ax_add_kernel, dim3(grid_for(n2)), dim3(256), 0, st, n2, on ? 1.0 : -1.0, ax_d2.get(),
                     hii.get() + n1);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
  adcx = on;
}

void Handle::adcx_apply(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, hipStream_t s) {
  ARG_CHECK(!scratch, "the no-allocation sigma-vector does not include the ADC(2)-x block");
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1, nn = n1 * n1, vv = (int64_t)v * v, oo = (int64_t)o * o;
  // vectors per chunk: seven n1 x n1 arrays per vector within about 4 GB
  const int64_t per_vec = 7 * nn * (int64_t)sizeof(double);
  const int cap = (int)std::max<int64_t>(1, std::min<int64_t>(16, ((int64_t)4 << 30) / per_vec));
  const int nbmax = std::min(nvec, cap);
  ax_t1.alloc((size_t)nbmax * nn);
  ax_t3.alloc((size_t)nbmax * nn);
  ax_t4.alloc((size_t)nbmax * nn);
  ax_t4n.alloc((size_t)nbmax * nn);
  ax_g1.alloc((size_t)nbmax * nn);
  ax_g2.alloc((size_t)nbmax * nn);
  ax_g3.alloc((size_t)nbmax * nn);
  for (int j0 = 0; j0 < nvec; j0 += nbmax) {
    const int nb = std::min(nbmax, nvec - j0);
    const int64_t w1 = (int64_t)nb * n1, w4 = (int64_t)nb * oo;  // row lengths of the chunk layouts
    hipLaunchKernelGGL(ax_layouts_kernel, dim3(grid_for(nn), nb), dim3(256), 0, s, X + ldx * j0, ldx, dk, nb,
                       ax_t1.get(), ax_t3.get(), ax_t4.get(), ax_t4n.get());
    HIP_CHECK(hipGetLastError());
    ax_gemm_bt(*this, (int)n1, (int)w1, (int)n1, ax_a1.get(), n1, ax_t1.get(), n1, ax_g1.get(), w1, 1.0, 0.0, s);
    ax_gemm_bt(*this, (int)n1, (int)w1, (int)n1, ax_ov.get(), n1, ax_t1.get(), n1, ax_g2.get(), w1, 1.0, 0.0, s);
    ax_gemm_bt(*this, (int)n1, (int)w1, (int)n1, ax_vo.get(), n1, ax_t3.get(), n1, ax_g2.get(), w1, 1.0, 1.0, s);
    ax_gemm_bt(*this, (int)vv, (int)w4, (int)vv, ax_vvvv.get(), vv, ax_t4.get(), vv, ax_g3.get(), w4, 0.5, 0.0, s);
    {  // G3 += T4 OOOO / 2 for each vector of the chunk (batch = vector); OOOO is symmetric
       // ((km|ln) = (mk|nl)), so B(mn, kl) = OOOO[(k,l)][(m,n)] is contiguous along K too
      GemmArgs g;
      g.M = (int)vv; g.N = (int)oo; g.K = (int)oo; g.batch = nb;
      g.A = ax_t4n.get(); g.sam = w4; g.sak = 1; g.sab = oo;
      g.B = ax_oooo.get(); g.sbk = 1; g.sbn = oo; g.sbb = 0;
      g.map.C = ax_g3.get(); g.map.scm = w4; g.map.scn = 1; g.map.scb = oo;
      g.map.alpha = 0.5; g.map.beta = 1.0;
      gemm(g, Prec::F64, gws, s);
    }
    hipLaunchKernelGGL(ax_project_kernel, dim3(grid_for(d.ndim - n1), nb), dim3(256), 0, s, ax_g1.get(),
                       ax_g2.get(), ax_g3.get(), dk, nb, W + ldw * j0, ldw);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace adc2g
