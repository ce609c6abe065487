// The ADC(2) F-vector f_J = <Psi_J|D|Psi_0> of a one-electron operator D
// (get_modifiedtm_adc2, source/adclib/get_moment.f90:153-442, with the term
// functions of source/adclib/dipole_ph.f90), closed shell, C1 symmetry.
//
// Every term is evaluated from the MP2 amplitudes t(i,j,a,b) = (ia|jb) / D_ij^ab
// with the restricted sums of the reference (b < c, l < m, each with its
// diagonal term) folded into full sums: each summand is symmetric under the
// exchange of the two restricted indices and its diagonal term is half the
// folded summand.  With the n1 x n1 pair matrices (rows and columns (a,i),
// index a*o + i)
//   G = (ai|bj),  X = t(i,j,a,b),  Y = 2 t(i,j,a,b) - t(i,j,b,a),
//   J = (ab|ij),  K = 2G - J,  L = 2J - G,
//   Y'[(c,m),(b,k)] = Y[(c,k),(b,m)],  X'[(a,m),(c,l)] = X[(a,l),(c,m)],
// the second-order 1h1p terms with dpl_ov become
//   F25 + F26 = sum_{m,c} dpl(m,c) / D_km^ac  (K Y + Y K)[(a,k),(c,m)]
//   F27 = -sum_{l,b} dpl(l,b) / D_kl^ab  (G Y + J Y')[(a,l),(b,k)]
//   F28 = -sum_{m,b} dpl(m,b) / D_km^ab  (X K + X' L)[(a,m),(b,k)]
// (six MFMA GEMMs of n1^3), the ladder term
//   F29 = sum_{b,l} dpl(l,b) / D_kl^ab  sum_{c,d} (ac|bd) Y[(c,k),(d,l)]
// from (ac|bd) built per a by a batched GEMM over the density-fitting
// factors and contracted by a second GEMM, and the remaining terms (F0, FA,
// FB, FC, F21..F24, F210 and their tau intermediates) by reductions with one
// workgroup per output element.  The 2h2p block is
//   R(a,b,k,l) = sum_m dpl(m,k) t(m,l,a,b) + sum_m dpl(m,l) t(k,m,a,b)
//              - sum_c dpl(a,c) t(k,l,c,b) - sum_c dpl(b,c) t(k,l,a,c)
// with FI = -R(a,a,k,k), FII = -sqrt(2) R(a,b,k,k), FIII = sqrt(2) R(a,a,k,l),
// FIV1 = R(a,b,k,l) + R(b,a,k,l) and FIV2 = sqrt(3) (R(a,b,k,l) - R(b,a,k,l)),
// one thread per configuration.
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int FV_THREADS = 256;

__device__ double fv_block_sum(double s) {
  __shared__ double red[FV_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = s;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < FV_THREADS / 64; ++k) t += red[k];
  return t;
}

__device__ __forceinline__ int64_t fv_pair_index(int64_t lo, int64_t hi, int64_t n) {
  return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
}

// (lo, hi) of pair p of the strictly ordered pairs lo < hi < n in row order
__device__ __forceinline__ void pair_decode_dev(int64_t p, int64_t n, int& lo, int& hi) {
  const double nn = (double)(2 * n - 1);
  int64_t a = (int64_t)floor((nn - sqrt(nn * nn - 8.0 * (double)p)) * 0.5);
  if (a < 0) a = 0;
  while (a > 0 && fv_pair_index(a, a + 1, n) > p) --a;
  while (a + 1 < n - 1 && fv_pair_index(a + 1, a + 2, n) <= p) ++a;
  lo = (int)a;
  hi = (int)(p - fv_pair_index(a, a + 1, n) + a + 1);
}

// out[r] (+)= sum_{i < n_in} f(r, i), one workgroup per output element r
template <class F>
__global__ void __launch_bounds__(FV_THREADS) fv_sum_kernel(int64_t n_in, F f, int accumulate,
                                                            double* __restrict__ out) {
  const int64_t r = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n_in; i += FV_THREADS) s += f(r, i);
  s = fv_block_sum(s);
  if (threadIdx.x == 0) out[r] = accumulate ? out[r] + s : s;
}

template <class F>
void fv_sum(int64_t n_out, int64_t n_in, F f, bool accumulate, double* out, hipStream_t st) {
  hipLaunchKernelGGL((fv_sum_kernel<F>), dim3((unsigned)n_out), dim3(FV_THREADS), 0, st, n_in, f,
                     accumulate ? 1 : 0, out);
  HIP_CHECK(hipGetLastError());
}

// X, Y, K, L, X', Y' and J from G and (ab|kl) = VVOO[(a*v+b)*o*o + k*o + l]
__global__ void __launch_bounds__(FV_THREADS) fv_pairs_kernel(int o, int v, const double* __restrict__ eps,
                                                              const double* __restrict__ G,
                                                              const double* __restrict__ VVOO, double* __restrict__ X,
                                                              double* __restrict__ Y, double* __restrict__ J,
                                                              double* __restrict__ Xs) {
  const int64_t n1 = (int64_t)o * v, total = n1 * n1;
  for (int64_t t = blockIdx.x * (int64_t)FV_THREADS + threadIdx.x; t < total; t += (int64_t)gridDim.x * FV_THREADS) {
    const int64_t m = t / n1, n = t - m * n1;
    const int a = (int)(m / o), i = (int)(m % o), b = (int)(n / o), j = (int)(n % o);
    const double D = eps[i] + eps[j] - eps[o + a] - eps[o + b];
    const double g = G[t];
    const double gs = G[(int64_t)(b * o + i) * n1 + (a * o + j)];  // (bi|aj)
    X[t] = g / D;
    Y[t] = (2.0 * g - gs) / D;
    J[t] = VVOO[((int64_t)a * v + b) * o * o + (int64_t)i * o + j];  // (ab|ij)
    // X'[(a,i),(b,j)] = X[(a,j),(b,i)] = (aj|bi) / D (D is symmetric in i, j)
    Xs[t] = G[(int64_t)(a * o + j) * n1 + (b * o + i)] / D;
  }
}

// K = 2G - J, L = 2J - G, Y'[(c,m),(b,k)] = Y[(c,k),(b,m)]
__global__ void __launch_bounds__(FV_THREADS) fv_pairs2_kernel(int o, int v, const double* __restrict__ G,
                                                               const double* __restrict__ J,
                                                               const double* __restrict__ Y, double* __restrict__ K,
                                                               double* __restrict__ L, double* __restrict__ Ys) {
  const int64_t n1 = (int64_t)o * v, total = n1 * n1;
  for (int64_t t = blockIdx.x * (int64_t)FV_THREADS + threadIdx.x; t < total; t += (int64_t)gridDim.x * FV_THREADS) {
    const int64_t m = t / n1, n = t - m * n1;
    const int c = (int)(m / o), mm = (int)(m % o), b = (int)(n / o), k = (int)(n % o);
    K[t] = 2.0 * G[t] - J[t];
    L[t] = 2.0 * J[t] - G[t];
    Ys[t] = Y[(int64_t)(c * o + k) * n1 + (b * o + mm)];
  }
}

// Y2T[(k,l)][(c,d)] = Y[(c,k),(d,l)]
__global__ void __launch_bounds__(FV_THREADS) fv_y2t_kernel(int o, int v, const double* __restrict__ Y,
                                                            double* __restrict__ Y2T) {
  const int64_t n1 = (int64_t)o * v, vv = (int64_t)v * v, total = vv * o * o;
  for (int64_t t = blockIdx.x * (int64_t)FV_THREADS + threadIdx.x; t < total; t += (int64_t)gridDim.x * FV_THREADS) {
    const int64_t kl = t / vv, cd = t - kl * vv;
    const int k = (int)(kl / o), l = (int)(kl % o), c = (int)(cd / v), d = (int)(cd % v);
    Y2T[t] = Y[(int64_t)(c * o + k) * n1 + (d * o + l)];
  }
}

// BT[(x)][P] = B[P][x] for x < nx
__global__ void __launch_bounds__(FV_THREADS) fv_transpose_kernel(int64_t nx, int naux, const double* __restrict__ B,
                                                                  double* __restrict__ BT) {
  const int64_t total = nx * naux;
  for (int64_t t = blockIdx.x * (int64_t)FV_THREADS + threadIdx.x; t < total; t += (int64_t)gridDim.x * FV_THREADS) {
    const int64_t x = t / naux, P = t - x * naux;
    BT[t] = B[P * nx + x];
  }
}

// the 2h2p block, one thread per configuration
__global__ void __launch_bounds__(FV_THREADS) fv_2h2p_kernel(DimsK d, const double* __restrict__ dp, int nbas,
                                                             const double* __restrict__ X, double* __restrict__ f) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1;
  auto t = [&](int i, int j, int a, int b) -> double { return X[(int64_t)(a * o + i) * n1 + (b * o + j)]; };
  auto dpl = [&](int p, int q) -> double { return dp[p + (int64_t)nbas * q]; };
  auto R = [&](int a, int b, int k, int l) -> double {
    double s = 0.0;
    for (int m = 0; m < o; ++m) s += dpl(m, k) * t(m, l, a, b) + dpl(m, l) * t(k, m, a, b);
    for (int c = 0; c < v; ++c) s -= dpl(o + a, o + c) * t(k, l, c, b) + dpl(o + b, o + c) * t(k, l, a, c);
    return s;
  };
  const int64_t n2 = d.ndim - n1;
  for (int64_t q = blockIdx.x * (int64_t)FV_THREADS + threadIdx.x; q < n2; q += (int64_t)gridDim.x * FV_THREADS) {
    const int64_t J = n1 + q;
    double val;
    if (J < d.off2) {  // t1: i outer, a inner
      const int64_t u = J - d.off1;
      const int k = (int)(u / v), a = (int)(u - (int64_t)k * v);
      val = -R(a, a, k, k);
    } else if (J < d.off3) {  // t2: (a<b) outer, i inner
      const int64_t u = J - d.off2;
      const int64_t p = u / o;
      const int k = (int)(u - p * o);
      int a, b;
      pair_decode_dev(p, v, a, b);
      val = -1.4142135623730951 * R(a, b, k, k);
    } else if (J < d.off4) {  // t3: (i<j) outer, a inner
      const int64_t u = J - d.off3;
      const int64_t qq = u / v;
      const int a = (int)(u - qq * v);
      int k, l;
      pair_decode_dev(qq, o, k, l);
      val = 1.4142135623730951 * R(a, a, k, l);
    } else {  // t4 / t5: (i<j) outer, (a<b) inner
      const bool t5 = J >= d.off5;
      const int64_t u = t5 ? J - d.off5 : J - d.off4;
      const int64_t qq = u / d.npv, p = u - qq * d.npv;
      int a, b, k, l;
      pair_decode_dev(p, v, a, b);
      pair_decode_dev(qq, o, k, l);
      const double r1 = R(a, b, k, l), r2 = R(b, a, k, l);
      val = t5 ? 1.7320508075688772 * (r1 - r2) : r1 + r2;
    }
    f[q] = val;
  }
}

// C[m*N + n] = sum_P A[P*sak + m] B[P*sbk + n]   (m < M, n < N, P < naux)
void df2(Handle& H, int M, const double* A, int64_t sak, int N, const double* B, int64_t sbk, double* C,
         hipStream_t st) {
  GemmArgs g;
  g.M = M; g.N = N; g.K = H.d.naux;
  g.A = A; g.sam = 1; g.sak = sak;
  g.B = B; g.sbk = sbk; g.sbn = 1;
  g.map.C = C; g.map.scm = N; g.map.scn = 1;
  gemm(g, Prec::F64, H.gws, st);
}

// C (+)= A . B for n1 x n1 row-major A and symmetric B
void pair_gemm(Handle& H, const double* A, const double* B, double* C, double beta, hipStream_t st) {
  const int n1 = (int)H.d.n1;
  GemmArgs g;
  g.M = n1; g.N = n1; g.K = n1;
  g.A = A; g.sam = n1; g.sak = 1;
  g.B = B; g.sbk = 1; g.sbn = n1;  // B(k, n) = B[n*n1 + k] = B[k*n1 + n]
  g.map.C = C; g.map.scm = n1; g.map.scn = 1; g.map.beta = beta;
  gemm(g, Prec::F64, H.gws, st);
}

}  // namespace

void fvector_run(Handle& H, const double* dpl_host, double* f_host) {
  ARG_CHECK(dpl_host && f_host, "NULL argument");
  if (H.parent) {
    // CVS view: get_modifiedtm_adc2 on the CVS configurations (lcvs skips the
    // I and II blocks, which the CVS space does not have; every other element
    // is the same function of its configuration), i.e. the full-space vector
    // at the view's configurations
    std::vector<double> full(H.parent->d.ndim);
    fvector_run(*H.parent, dpl_host, full.data());
    std::vector<int64_t> vm(H.d.ndim);
    HIP_CHECK(hipMemcpy(vm.data(), H.vmap.get(), sizeof(int64_t) * vm.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < H.d.ndim; ++i) f_host[i] = full[vm[i]];
    return;
  }
  const int o = H.d.o, v = H.d.v, naux = H.d.naux, nbas = o + v;
  const int64_t n1 = H.d.n1, nn = n1 * n1, vv = (int64_t)v * v, oo = (int64_t)o * o;
  hipStream_t st = H.st;
  const DimsK dk = dims_k(H.d);
  const double* eps = H.eps.get();

  DevBuf<double> dp, G, X, Y, J, Xs, K, L, Ys, Q, VVOO, OOOO, Vb, Ob, acc, tmp, tmp2;
  dp.alloc((size_t)nbas * nbas);
  HIP_CHECK(hipMemcpyAsync(dp.get(), dpl_host, sizeof(double) * nbas * nbas, hipMemcpyHostToDevice, st));
  G.alloc(nn); X.alloc(nn); Y.alloc(nn); J.alloc(nn); Xs.alloc(nn);
  VVOO.alloc(vv * oo);
  df2(H, (int)n1, H.Bov.get(), n1, (int)n1, H.Bov.get(), n1, G.get(), st);        // (ai|bj)
  df2(H, (int)vv, H.Bvv.get(), vv, (int)oo, H.Boo.get(), oo, VVOO.get(), st);     // (ab|kl)
  const int gsz = grid_for(nn, FV_THREADS, 16384);
  hipLaunchKernelGGL(fv_pairs_kernel, dim3(gsz), dim3(FV_THREADS), 0, st, o, v, eps, G.get(), VVOO.get(), X.get(),
                     Y.get(), J.get(), Xs.get());
  HIP_CHECK(hipGetLastError());
  K.alloc(nn); L.alloc(nn); Ys.alloc(nn); Q.alloc(nn);
  hipLaunchKernelGGL(fv_pairs2_kernel, dim3(gsz), dim3(FV_THREADS), 0, st, o, v, G.get(), J.get(), Y.get(), K.get(),
                     L.get(), Ys.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
  VVOO.release();

  // (ac|bl) as Vb[(a*v+c)*n1 + b*o+l], (km|bl) as Ob[(k*o+m)*n1 + b*o+l], (mk|nl) as OOOO[(m*o+k)*o*o + n*o+l]
  Vb.alloc(vv * n1);
  Ob.alloc(oo * n1);
  OOOO.alloc(oo * oo);
  df2(H, (int)vv, H.Bvv.get(), vv, (int)n1, H.Bov.get(), n1, Vb.get(), st);
  df2(H, (int)oo, H.Boo.get(), oo, (int)n1, H.Bov.get(), n1, Ob.get(), st);
  df2(H, (int)oo, H.Boo.get(), oo, (int)oo, H.Boo.get(), oo, OOOO.get(), st);

  const double* Xp = X.get();
  const double* Yp = Y.get();
  const double* Vp = Vb.get();
  const double* Op = Ob.get();
  const double* Dp = dp.get();
  const double* OOp = OOOO.get();
  // accessors (all captured by value in the device lambdas below)
#define T_(i, j, a, b) Xp[(int64_t)((a) * o + (i)) * n1 + ((b) * o + (j))]
#define DPL(p, q) Dp[(p) + (int64_t)nbas * (q)]
#define VB(a, c, b, l) Vp[((int64_t)(a) * v + (c)) * n1 + ((b) * o + (l))]
#define OB(k, m, b, l) Op[((int64_t)(k) * o + (m)) * n1 + ((b) * o + (l))]
#define E_(p) eps[p]

  acc.alloc(n1);  // the 1h1p sum, rows a*o + k
  tmp.alloc(std::max<int64_t>(vv, n1));
  tmp2.alloc(n1);
  double* A = acc.get();
  double* T1 = tmp.get();
  double* T2 = tmp2.get();

  // F0 (dipole_ph.f90:24-46), with dpl(a,k) as the first element of the sum
  fv_sum(n1, (int64_t)v * o, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), b = (int)(i / o), l = (int)(i % o);
    const double s = DPL(o + b, l) * (2.0 * T_(k, l, a, b) - T_(l, k, a, b));
    return i == 0 ? s + DPL(o + a, k) : s;
  }, false, A, st);

  // tauA(c,a) (:896-935) at T1[c*v + a]; FA = sum_c tauA(c,a) dpl(c,k)
  fv_sum(vv, (int64_t)o * o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int c = (int)(r / v), a = (int)(r % v);
    const int l = (int)(i % o), m = (int)((i / o) % o), b = (int)(i / (o * o));
    return -0.5 * T_(l, m, a, b) * (2.0 * T_(l, m, c, b) - T_(l, m, b, c));
  }, false, T1, st);
  fv_sum(n1, v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), c = (int)i;
    return T1[(int64_t)c * v + a] * DPL(o + c, k);
  }, true, A, st);

  // tauB(m,k) (:939-981) at T1[m*o + k]; FB = sum_l dpl(a,l) tauB(l,k)
  fv_sum(oo, (int64_t)v * o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int m = (int)(r / o), k = (int)(r % o);
    const int b = (int)(i % v), l = (int)((i / v) % o), c = (int)(i / (v * o));
    return -0.5 * T_(k, l, b, c) * (2.0 * T_(m, l, b, c) - T_(m, l, c, b));
  }, false, T1, st);
  fv_sum(n1, o, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), l = (int)i;
    return DPL(o + a, l) * T1[(int64_t)l * o + k];
  }, true, A, st);

  // tau21(a,m) (:985-1029) and tau22(a,m) (:1033-1076) at T2[a*o + m];
  // F21 + F22 = sum_l (tau21 + tau22)(a,l) dpl(l,k)
  fv_sum(n1, (int64_t)o * v * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), m = (int)(r % o);
    const int l = (int)(i % o), b = (int)((i / o) % v), c = (int)(i / (o * v));
    return -T_(l, m, b, c) * (2.0 * VB(a, c, b, l) - VB(a, b, c, l)) / (E_(m) - E_(o + a));
  }, false, T2, st);
  fv_sum(n1, (int64_t)o * v * o, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), n = (int)(r % o);
    const int l = (int)(i % o), b = (int)((i / o) % v), m = (int)(i / (o * v));
    return OB(m, n, b, l) * (2.0 * T_(l, m, b, a) - T_(m, l, b, a)) / (E_(n) - E_(o + a));
  }, true, T2, st);
  fv_sum(n1, o, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), l = (int)i;
    return T2[(int64_t)a * o + l] * DPL(l, k);
  }, true, A, st);

  // tau23(b,k) (:1080-1125) and tau24(b,k) (:1129-1177) at T2[b*o + k];
  // F23 + F24 = sum_c dpl(a,c) (tau23 + tau24)(c,k)
  fv_sum(n1, (int64_t)o * v * o, [=] __device__(int64_t r, int64_t i) -> double {
    const int b = (int)(r / o), k = (int)(r % o);
    const int m = (int)(i % o), c = (int)((i / o) % v), l = (int)(i / (o * v));
    return -T_(l, m, b, c) * (2.0 * OB(l, k, c, m) - OB(m, k, c, l)) / (E_(k) - E_(o + b));
  }, false, T2, st);
  fv_sum(n1, (int64_t)o * v * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int dd = (int)(r / o), k = (int)(r % o);
    const int l = (int)(i % o), c = (int)((i / o) % v), b = (int)(i / (o * v));
    return T_(k, l, b, c) * (2.0 * VB(dd, b, c, l) - VB(dd, c, b, l)) / (E_(k) - E_(o + dd));
  }, true, T2, st);
  fv_sum(n1, v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), c = (int)i;
    return DPL(o + a, o + c) * T2[(int64_t)c * o + k];
  }, true, A, st);

  // tauC(b,l) (:1181-1209) at T2[b*o + l]; FC (FC_ph_new, :1213-1242)
  fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int b = (int)(r / o), l = (int)(r % o), m = (int)(i % o), c = (int)(i / o);
    return 0.5 * DPL(o + c, m) * (2.0 * T_(m, l, c, b) - T_(m, l, b, c));
  }, false, T2, st);
  fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), l = (int)(i % o), b = (int)(i / o);
    return T2[(int64_t)b * o + l] * (2.0 * T_(k, l, a, b) - T_(l, k, a, b));
  }, true, A, st);

  // F210 (:610-656): sum_{b,l} dpl(l,b) / D_kl^ab sum_{m,n} t(m,n,a,b) (2 (mk|nl) - (ml|nk))
  fv_sum(n1, (int64_t)o * o * o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o);
    const int m = (int)(i % o), n = (int)((i / o) % o), l = (int)((i / (o * o)) % o), b = (int)(i / (o * o * o));
    const double x = 2.0 * OOp[((int64_t)m * o + k) * o * o + n * o + l] - OOp[((int64_t)m * o + l) * o * o + n * o + k];
    return DPL(l, o + b) / (E_(k) + E_(l) - E_(o + a) - E_(o + b)) * T_(m, n, a, b) * x;
  }, true, A, st);
  HIP_CHECK(hipStreamSynchronize(st));
  Vb.release();
  Ob.release();

  // F25 + F26 (:390-469): Q = K Y + Y K
  pair_gemm(H, K.get(), Y.get(), Q.get(), 0.0, st);
  pair_gemm(H, Y.get(), K.get(), Q.get(), 1.0, st);
  const double* Qp = Q.get();
  fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), m = (int)(i % o), c = (int)(i / o);
    return DPL(m, o + c) / (E_(k) + E_(m) - E_(o + a) - E_(o + c)) * Qp[r * n1 + (c * o + m)];
  }, true, A, st);
  // F27 (:473-513): Q = G Y + J Y'
  pair_gemm(H, G.get(), Y.get(), Q.get(), 0.0, st);
  pair_gemm(H, J.get(), Ys.get(), Q.get(), 1.0, st);
  fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), l = (int)(i % o), b = (int)(i / o);
    return -DPL(l, o + b) / (E_(k) + E_(l) - E_(o + a) - E_(o + b)) *
           Qp[((int64_t)a * o + l) * n1 + (b * o + k)];
  }, true, A, st);
  // F28 (:517-557): Q = X K + X' L
  pair_gemm(H, X.get(), K.get(), Q.get(), 0.0, st);
  pair_gemm(H, Xs.get(), L.get(), Q.get(), 1.0, st);
  fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
    const int a = (int)(r / o), k = (int)(r % o), m = (int)(i % o), b = (int)(i / o);
    return -DPL(m, o + b) / (E_(k) + E_(m) - E_(o + a) - E_(o + b)) *
           Qp[((int64_t)a * o + m) * n1 + (b * o + k)];
  }, true, A, st);
  HIP_CHECK(hipStreamSynchronize(st));
  K.release(); L.release(); Ys.release(); Xs.release(); J.release(); Q.release(); G.release();

  // F29 (:561-606): Z[(a,b),(k,l)] = sum_{c,d} (ac|bd) Y[(c,k),(d,l)], (ac|bd)
  // per a as W[b][(c,d)] from the transposed factors
  {
    DevBuf<double> BT, W, Y2T, Z;
    BT.alloc(vv * naux);
    hipLaunchKernelGGL(fv_transpose_kernel, dim3(grid_for(vv * naux, FV_THREADS, 16384)), dim3(FV_THREADS), 0, st,
                       vv, naux, H.Bvv.get(), BT.get());
    HIP_CHECK(hipGetLastError());
    Y2T.alloc(vv * oo);
    hipLaunchKernelGGL(fv_y2t_kernel, dim3(grid_for(vv * oo, FV_THREADS, 16384)), dim3(FV_THREADS), 0, st, o, v,
                       Y.get(), Y2T.get());
    HIP_CHECK(hipGetLastError());
    W.alloc(vv * v);
    Z.alloc(vv * oo);
    for (int a = 0; a < v; ++a) {
      GemmArgs g;  // W[b][c][d] = sum_P BT[(b,d)][P] BT[(a,c)][P]
      g.M = v; g.N = v; g.K = naux; g.batch = v;
      g.A = BT.get(); g.sam = naux; g.sak = 1; g.sab = (int64_t)v * naux;
      g.B = BT.get() + (int64_t)a * v * naux; g.sbk = 1; g.sbn = naux; g.sbb = 0;
      g.map.C = W.get(); g.map.scm = 1; g.map.scn = v; g.map.scb = vv;
      gemm(g, Prec::F64, H.gws, st);
      GemmArgs z;  // Z[(a,b)][(k,l)] = sum_{(c,d)} W[b][(c,d)] Y2T[(k,l)][(c,d)]
      z.M = v; z.N = (int)oo; z.K = (int)vv;
      z.A = W.get(); z.sam = vv; z.sak = 1;
      z.B = Y2T.get(); z.sbk = 1; z.sbn = vv;
      z.map.C = Z.get() + (int64_t)a * v * oo; z.map.scm = oo; z.map.scn = 1;
      gemm(z, Prec::F64, H.gws, st);
    }
    const double* Zp = Z.get();
    fv_sum(n1, (int64_t)o * v, [=] __device__(int64_t r, int64_t i) -> double {
      const int a = (int)(r / o), k = (int)(r % o), l = (int)(i % o), b = (int)(i / o);
      return DPL(l, o + b) / (E_(k) + E_(l) - E_(o + a) - E_(o + b)) * Zp[((int64_t)a * v + b) * oo + k * o + l];
    }, true, A, st);
    HIP_CHECK(hipStreamSynchronize(st));
  }

  // 1h1p: f = -sqrt(2) * (the sum) (get_moment.f90:367)
  std::vector<double> h1(n1);
  HIP_CHECK(hipMemcpyAsync(h1.data(), A, sizeof(double) * n1, hipMemcpyDeviceToHost, st));
  // 2h2p (get_moment.f90:370-435)
  DevBuf<double> f2;
  const int64_t n2 = H.d.ndim - n1;
  f2.alloc(n2);
  hipLaunchKernelGGL(fv_2h2p_kernel, dim3(grid_for(n2, FV_THREADS, 65536)), dim3(FV_THREADS), 0, st, dk, Dp, nbas,
                     Xp, f2.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(f_host + n1, f2.get(), sizeof(double) * n2, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  for (int64_t r = 0; r < n1; ++r) f_host[r] = -std::sqrt(2.0) * h1[r];
#undef T_
#undef DPL
#undef VB
#undef OB
#undef E_
}

}  // namespace adc2g
