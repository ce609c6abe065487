// MP2 ground-state quantities on the device: the energy (mp2_ener,
// source/adclib/mp2.f90:68-154), the D2 diagnostic (mp2_d2, :161-276) and the
// ground-state density matrix (rho_mp2, :301-413, with rhogs2_uu, rhogs2_oo,
// rhogs2_ou, :421-596), closed shell; on a frozen-core view with the frozen
// occupied orbitals out of every occupied sum, as lifrzcore sets occ1.
//
// With G[(a,i),(b,j)] = (ai|bj) = sum_P B_P,ai B_P,bj (one GEMM over the
// density-fitting factors), D_ijab = e_i + e_j - e_a - e_b,
// X = G / D (the amplitudes t_ij^ab) and Y[(a,i),(b,j)] = 2 X[(a,i),(b,j)] -
// X[(b,i),(a,j)]:
//   E      = sum G (2 G - G[(b,i),(a,j)]) / D
//   (T^o T^o)_ij = sum_{a,(b,k)} X[(a,i),(b,k)] X[(a,j),(b,k)],
//   (T^v T^v)_ab = sum_{i,(c,j)} X[(a,i),(c,j)] X[(b,i),(c,j)],
//   oo(i,j) = -2 sum_{a,(b,k)} X[(a,i),(b,k)] Y[(a,j),(b,k)],
//   uu(a,b) =  2 sum_{i,(c,j)} X[(a,i),(c,j)] Y[(b,i),(c,j)],
//   ou(i,a) = 2 [ sum_{b,(c,j)} X[(b,i),(c,j)] ((ac|bj) - 2 (ab|cj))
//               + sum_{j,(b,k)} (ij|bk) Y[(a,j),(b,k)] ] / (e_a - e_i),
// which are the reference's element sums with the two terms of each summand
// folded by the symmetry of the dummy indices (X and Y are symmetric
// matrices).  The (vv|ov) and (oo|ov) blocks for ou come from two more
// GEMMs.  The density is assembled as the reference's loops leave it: the
// off-diagonal oo elements receive oo(i,j) + oo(j,i) (the loop adds and
// mirrors, then adds the transposed element to the mirrored value).
#include <cmath>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int MP_THREADS = 256;

__device__ double mp_block_sum(double s) {
  __shared__ double red[MP_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = s;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < MP_THREADS / 64; ++k) t += red[k];
  return t;
}

// X = G / D, Y = (2 G - G_swap) / D; part[blk] = partial energy
// frz (o flags, or NULL): amplitudes with a frozen occupied index are zero,
// which restricts every occupied sum to the active orbitals (occ1 = ncore + 1)
__global__ void __launch_bounds__(MP_THREADS) mp2_amp_kernel(int n1, int o, const double* __restrict__ eps,
                                                             const double* __restrict__ G, const char* __restrict__ frz,
                                                             double* __restrict__ X, double* __restrict__ Y,
                                                             double* __restrict__ part) {
  const int64_t total = (int64_t)n1 * n1;
  double s = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)MP_THREADS + threadIdx.x; t < total; t += (int64_t)gridDim.x * MP_THREADS) {
    const int m = (int)(t % n1), n = (int)(t / n1);
    const int a = m / o, i = m % o, b = n / o, j = n % o;
    if (frz && (frz[i] || frz[j])) {
      X[t] = 0.0;
      Y[t] = 0.0;
      continue;
    }
    const double D = eps[i] + eps[j] - eps[o + a] - eps[o + b];
    const double g = G[t];
    const double gs = G[(int64_t)(b * o + i) + (int64_t)n1 * (a * o + j)];
    X[t] = g / D;
    Y[t] = (2.0 * g - gs) / D;
    s += g * (2.0 * g - gs) / D;
  }
  s = mp_block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[p + P q] = scale sum_{r < R} sum_{c < n1} A[c + n1 row(p,r)] B[c + n1 row(q,r)],
// row(p, r) = p sp + r sr; one workgroup per (p, q)
__global__ void __launch_bounds__(MP_THREADS) mp2_pair_kernel(const double* __restrict__ A,
                                                              const double* __restrict__ B, int n1, int P, int sp,
                                                              int sr, int R, double scale, double* __restrict__ out) {
  const int p = blockIdx.x % P, q = blockIdx.x / P;
  double s = 0.0;
  for (int r = 0; r < R; ++r) {
    const double* a = A + (int64_t)n1 * (p * sp + r * sr);
    const double* b = B + (int64_t)n1 * (q * sp + r * sr);
    for (int c = threadIdx.x; c < n1; c += MP_THREADS) s += a[c] * b[c];
  }
  s = mp_block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = scale * s;
}

// s1[i + o a] = sum_{b, (c,j)} X[(c,j) + n1 (b o + i)] (V[(a v + c) n1 + b o + j] - 2 V[(a v + b) n1 + c o + j])
__global__ void __launch_bounds__(MP_THREADS) mp2_ou1_kernel(int n1, int o, int v, const double* __restrict__ X,
                                                             const double* __restrict__ V, double* __restrict__ s1) {
  const int i = blockIdx.x % o, a = blockIdx.x / o;
  double s = 0.0;
  for (int b = 0; b < v; ++b) {
    const double* x = X + (int64_t)n1 * (b * o + i);
    const double* vab = V + ((int64_t)a * v + b) * n1;
    for (int t = threadIdx.x; t < n1; t += MP_THREADS) {
      const int c = t / o, j = t % o;
      const double vacbj = V[((int64_t)a * v + c) * n1 + b * o + j];
      s += x[t] * (vacbj - 2.0 * vab[t]);
    }
  }
  s = mp_block_sum(s);
  if (threadIdx.x == 0) s1[blockIdx.x] = s;
}

// s2[i + o a] = sum_{j, (b,k)} O[(i o + j) n1 + (b,k)] Y[(b,k) + n1 (a o + j)]
__global__ void __launch_bounds__(MP_THREADS) mp2_ou2_kernel(int n1, int o, const double* __restrict__ Ob,
                                                             const double* __restrict__ Y, double* __restrict__ s2) {
  const int i = blockIdx.x % o, a = blockIdx.x / o;
  double s = 0.0;
  for (int j = 0; j < o; ++j) {
    const double* ob = Ob + (int64_t)(i * o + j) * n1;
    const double* y = Y + (int64_t)n1 * (a * o + j);
    for (int t = threadIdx.x; t < n1; t += MP_THREADS) s += ob[t] * y[t];
  }
  s = mp_block_sum(s);
  if (threadIdx.x == 0) s2[blockIdx.x] = s;
}

// C[m scm + n] = sum_P A[P sak + m] B[P n1 + n] (m < M, n < n1): an integral block
// from the factors, rows m of the first pair, columns (b,l) of the ov pair
void df_block(Handle& H, int M, const double* A, int64_t sak, double* C, hipStream_t st) {
  GemmArgs g;
  g.M = M;
  g.N = (int)H.d.n1;
  g.K = H.d.naux;
  g.A = A;
  g.sam = 1;
  g.sak = sak;
  g.B = H.Bov.get();
  g.sbk = H.d.n1;
  g.sbn = 1;
  g.map.C = C;
  g.map.scm = H.d.n1;
  g.map.scn = 1;
  g.map.alpha = 1.0;
  g.map.beta = 0.0;
  gemm(g, Prec::F64, H.gws, st);
}

}  // namespace

void mp2_run(Handle& Hv, double* e_mp2, double* d2, double* rho) {
  ARG_CHECK(!Hv.parent || !Hv.frozen_occ.empty(), "MP2 quantities belong to the full operator or a frozen-core "
                                                  "view, not to a CVS view");
  // a frozen-core view: the parent's integrals with the frozen occupied
  // orbitals left out of every occupied sum (lifrzcore, mp2.f90:84-90, 333-337)
  Handle& H = Hv.parent ? *Hv.parent : Hv;
  const std::vector<char>* frz = Hv.parent ? &Hv.frozen_occ : nullptr;
  const int o = H.d.o, v = H.d.v, n1 = (int)H.d.n1, nbas = o + v;
  DevBuf<char> dfrz;
  if (frz) {
    dfrz.alloc((size_t)o);
    HIP_CHECK(hipMemcpyAsync(dfrz.get(), frz->data(), (size_t)o, hipMemcpyHostToDevice, H.st));
  }
  auto frozen = [&](int i) { return frz && (*frz)[i]; };
  hipStream_t st = H.st;
  const int64_t nn = (int64_t)n1 * n1;
  DevBuf<double> G, X, Y, part, sums, toto, tvtv;
  G.alloc(nn);
  X.alloc(nn);
  Y.alloc(nn);
  // G[(a,i),(b,j)] = sum_P Bov[P][(a,i)] Bov[P][(b,j)] (column-major; symmetric)
  df_block(H, n1, H.Bov.get(), n1, G.get(), st);
  const int nblk = (int)std::min<int64_t>(4096, (nn + MP_THREADS - 1) / MP_THREADS);
  part.alloc(nblk);
  sums.alloc(1);
  hipLaunchKernelGGL(mp2_amp_kernel, dim3(nblk), dim3(MP_THREADS), 0, st, n1, o, H.eps.get(), G.get(),
                     frz ? dfrz.get() : nullptr, X.get(), Y.get(), part.get());
  HIP_CHECK(hipGetLastError());
  launch_colsum(nblk, 1, part.get(), sums.get(), st);
  double e = 0.0;
  HIP_CHECK(hipMemcpyAsync(&e, sums.get(), sizeof(double), hipMemcpyDeviceToHost, st));
  // T^o T^o (rows (a,i): p = i, r = a) and T^v T^v (p = a, r = i)
  toto.alloc((size_t)o * o);
  tvtv.alloc((size_t)v * v);
  hipLaunchKernelGGL(mp2_pair_kernel, dim3(o * o), dim3(MP_THREADS), 0, st, X.get(), X.get(), n1, o, 1, o, v, 1.0,
                     toto.get());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mp2_pair_kernel, dim3(v * v), dim3(MP_THREADS), 0, st, X.get(), X.get(), n1, v, o, 1, o, 1.0,
                     tvtv.get());
  HIP_CHECK(hipGetLastError());
  std::vector<double> hto((size_t)o * o), htv((size_t)v * v);
  HIP_CHECK(hipMemcpyAsync(hto.data(), toto.get(), sizeof(double) * hto.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(htv.data(), tvtv.get(), sizeof(double) * htv.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (e_mp2) *e_mp2 = e;
  if (d2) {
    std::vector<double> w(std::max(o, v)), Z((size_t)v * v);
    if (!host_syev(o, hto.data(), w.data(), Z.data()) ) throw Error(ADC2G_EEIG, "mp2_d2: T^o T^o eigensolve failed");
    double dd = std::sqrt(w[o - 1]);
    if (!host_syev(v, htv.data(), w.data(), Z.data())) throw Error(ADC2G_EEIG, "mp2_d2: T^v T^v eigensolve failed");
    if (std::sqrt(w[v - 1]) > dd) dd = std::sqrt(w[v - 1]);
    *d2 = dd;
  }
  if (!rho) return;
  DevBuf<double> oo, uu, s1, s2, Vb, Ob;
  oo.alloc((size_t)o * o);
  uu.alloc((size_t)v * v);
  s1.alloc((size_t)o * v);
  s2.alloc((size_t)o * v);
  hipLaunchKernelGGL(mp2_pair_kernel, dim3(o * o), dim3(MP_THREADS), 0, st, X.get(), Y.get(), n1, o, 1, o, v, -2.0,
                     oo.get());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mp2_pair_kernel, dim3(v * v), dim3(MP_THREADS), 0, st, X.get(), Y.get(), n1, v, o, 1, o, 2.0,
                     uu.get());
  HIP_CHECK(hipGetLastError());
  // (vv|ov) as V[(a,c)][(b,l)] and (oo|ov) as O[(k,m)][(b,l)]
  Vb.alloc((size_t)v * v * n1);
  Ob.alloc((size_t)o * o * n1);
  df_block(H, v * v, H.Bvv.get(), (int64_t)v * v, Vb.get(), st);
  df_block(H, o * o, H.Boo.get(), (int64_t)o * o, Ob.get(), st);
  hipLaunchKernelGGL(mp2_ou1_kernel, dim3(o * v), dim3(MP_THREADS), 0, st, n1, o, v, X.get(), Vb.get(), s1.get());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mp2_ou2_kernel, dim3(o * v), dim3(MP_THREADS), 0, st, n1, o, Ob.get(), Y.get(), s2.get());
  HIP_CHECK(hipGetLastError());
  std::vector<double> hoo((size_t)o * o), huu((size_t)v * v), h1((size_t)o * v), h2((size_t)o * v), he(nbas);
  HIP_CHECK(hipMemcpyAsync(hoo.data(), oo.get(), sizeof(double) * hoo.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(huu.data(), uu.get(), sizeof(double) * huu.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(h1.data(), s1.get(), sizeof(double) * h1.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(h2.data(), s2.get(), sizeof(double) * h2.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(he.data(), H.eps.get(), sizeof(double) * nbas, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  // assembly in the order of rho_mp2's loops (column-major nbas x nbas)
  auto R = [&](int p, int q) -> double& { return rho[p + (size_t)nbas * q]; };
  for (size_t k = 0; k < (size_t)nbas * nbas; ++k) rho[k] = 0.0;
  for (int i = 0; i < o; ++i) R(i, i) = 2.0;
  for (int i = 0; i < o; ++i)
    for (int j = 0; j < o; ++j) {
      if (frozen(i) || frozen(j)) continue;  // do i = occ1, nocc; do j = occ1, nocc
      R(i, j) = R(i, j) + hoo[i + (size_t)o * j];
      R(j, i) = R(i, j);
    }
  for (int a = 0; a < v; ++a)
    for (int b = a; b < v; ++b) {
      R(o + a, o + b) = huu[a + (size_t)v * b];
      R(o + b, o + a) = R(o + a, o + b);
    }
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) {
      if (frozen(i)) continue;
      const double val = 2.0 * (h1[i + (size_t)o * a] + h2[i + (size_t)o * a]) / (he[o + a] - he[i]);
      R(i, o + a) = val;
      R(o + a, i) = val;
    }
}

}  // namespace adc2g
