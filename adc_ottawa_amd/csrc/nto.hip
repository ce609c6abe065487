// The occupied-virtual block of the ADC(2) ground-to-excited-state transition
// density of adc2_trdens_gs_virt_occ (source/adclib/density_matrix.f90:497-636),
// the input of the natural transition orbitals of adc2_nto_gs and of the
// time-dependent NTOs of tdadc2_nto (source/adclib/nto.f90:76-260, 382-547).
//
// With X[a,i] the 1h1p amplitudes of a vector (configuration a*o + i; the
// 2h2p amplitudes do not enter), rho the MP2 density as rho_mp2 assembles it
// (2 + oo(i,i) on the occupied diagonal) and the symmetric n1 x n1 matrix
//   M[(a,i),(b,j)] = (2 (ai|bj) - (aj|bi)) / (e_a + e_b - e_i - e_j),
// the reference's five contributions are
//   T = sqrt(2) (-X + M X + rho_vv X / 4 - X rho_oo / 4 - M M X / 2):
// zeroth order (:525-535), first order (:540-556), the two MP2-density terms
// with rhomp2(a,b) and rhomp2(i,j) (:561-589), and the third second-order
// term (:594-627), whose intermediate lambda is -M X.  M comes from the
// density-fitting factors once per operator and stays in device memory with
// rho; each call is then two n1 x n1 x nvec FP64 GEMMs and one kernel.  On a
// frozen-core view rho is the frozen-core MP2 density (the frozen occupied
// orbitals out of its occupied sums), on a CVS view the full one; the view's
// 1h1p amplitudes enter at their configurations, and M keeps every orbital,
// as the reference's loops over the view's kpqf do.
#include <cmath>
#include <vector>

#include "gemm.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

// M from G = (ai|bj) (row-major, rows and columns a*o + i)
__global__ void nto_m_kernel(int o, int v, const double* __restrict__ eps, const double* __restrict__ G,
                             double* __restrict__ M) {
  const int64_t n1 = (int64_t)o * v, nn = n1 * n1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / n1, c = t % n1;
    const int a = (int)(r / o), i = (int)(r % o), b = (int)(c / o), j = (int)(c % o);
    const double ex = G[((int64_t)a * o + j) * n1 + ((int64_t)b * o + i)];  // (aj|bi)
    M[t] = (2.0 * G[t] - ex) / (eps[o + a] + eps[o + b] - eps[i] - eps[j]);
  }
}

// out[s][a + v*i] = sqrt(2) (-X - P2/2 + P1 + (rho_vv X)/4 - (X rho_oo)/4)
// with P1 = M X and P2 = M P1 (n1 x nvec, column s at s*n1)
__global__ void nto_final_kernel(int o, int v, int nvec, const double* __restrict__ x, int64_t ldx,
                                 const double* __restrict__ P1, const double* __restrict__ P2,
                                 const double* __restrict__ rho, double* __restrict__ out) {
  const int64_t n1 = (int64_t)o * v, nbas = o + v;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n1 * nvec;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / n1, r = t % n1;
    const int a = (int)(r / o), i = (int)(r % o);
    const double* xs = x + s * ldx;
    double rv = 0.0, ro = 0.0;
    for (int b = 0; b < v; ++b) rv += rho[(o + a) + nbas * (o + b)] * xs[(int64_t)b * o + i];
    for (int j = 0; j < o; ++j) ro += xs[(int64_t)a * o + j] * rho[j + nbas * i];
    const double val = -xs[r] + P1[s * n1 + r] + 0.25 * rv - 0.25 * ro - 0.5 * P2[s * n1 + r];
    out[s * n1 + a + (int64_t)v * i] = std::sqrt(2.0) * val;
  }
}

// x (the view's 1h1p rows, ld ldx) scattered into the full 1h1p rows (ld n1)
__global__ void nto_scatter_kernel(int64_t n1v, int nvec, const int64_t* __restrict__ vmap,
                                   const double* __restrict__ x, int64_t ldx, double* __restrict__ xf, int64_t n1) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n1v * nvec;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / n1v, r = t % n1v;
    xf[vmap[r] + s * n1] = x[r + s * ldx];
  }
}

// a frozen-core view works on its parent's orbitals with the frozen-core MP2
// density; its 1h1p configurations (the view's first rows) map into the
// parent's (the frozen ones stay zero), as the reference's loops over kpqf
// of the frozen-core space leave them
void nto_prepare(Handle& H) {
  if (H.nto_ready) return;
  ARG_CHECK(!H.stored, "the transition density needs the operator built from integrals, not a stored matrix");
  Handle& F = H.parent ? *H.parent : H;
  const int o = F.d.o, v = F.d.v, nbas = o + v;
  const int64_t n1 = F.d.n1, nn = n1 * n1;
  hipStream_t st = H.st;
  std::vector<double> rho((size_t)nbas * nbas);
  // a frozen-core view: the frozen-core density; a CVS view: the full one
  // (rho_mp2 changes only with lifrzcore)
  const bool cvs = H.parent && H.frozen_occ.empty();
  mp2_run(cvs ? F : H, nullptr, nullptr, rho.data());
  H.nto_rho.alloc((size_t)nbas * nbas);
  HIP_CHECK(hipMemcpyAsync(H.nto_rho.get(), rho.data(), sizeof(double) * rho.size(), hipMemcpyHostToDevice, st));
  DevBuf<double> G;
  G.alloc(nn);
  GemmArgs g;  // G[(a,i)*n1 + (b,j)] = sum_P Bov[P][(a,i)] Bov[P][(b,j)]
  g.M = (int)n1; g.N = (int)n1; g.K = F.d.naux;
  g.A = F.Bov.get(); g.sam = 1; g.sak = n1;
  g.B = F.Bov.get(); g.sbk = n1; g.sbn = 1;
  g.map.C = G.get(); g.map.scm = n1; g.map.scn = 1;
  gemm(g, Prec::F64, H.gws, st);
  H.nto_M.alloc(nn);
  hipLaunchKernelGGL(nto_m_kernel, dim3(grid_for(nn, 256, 65536)), dim3(256), 0, st, o, v, F.eps.get(), G.get(),
                     H.nto_M.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
  H.nto_ready = true;
}

}  // namespace

void trdens_vo_device(Handle& H, int nvec, const double* x, int64_t ldx, double* out) {
  ARG_CHECK(nvec >= 1 && ldx >= H.d.n1, "nvec must be >= 1 and ldx >= the 1h1p count");
  nto_prepare(H);
  const Handle& F = H.parent ? *H.parent : H;
  const int64_t n1 = F.d.n1;
  hipStream_t st = H.st;
  if (H.parent) {
    H.nto_x.alloc((size_t)n1 * nvec);
    HIP_CHECK(hipMemsetAsync(H.nto_x.get(), 0, sizeof(double) * n1 * nvec, st));
    hipLaunchKernelGGL(nto_scatter_kernel, dim3(grid_for(H.d.n1 * nvec, 256, 16384)), dim3(256), 0, st, H.d.n1, nvec,
                       H.vmap.get(), x, ldx, H.nto_x.get(), n1);
    HIP_CHECK(hipGetLastError());
    x = H.nto_x.get();
    ldx = n1;
  }
  H.nto_p.alloc(2 * (size_t)n1 * nvec);
  double* P1 = H.nto_p.get();
  double* P2 = P1 + n1 * nvec;
  for (int pass = 0; pass < 2; ++pass) {
    GemmArgs g;  // P = M . (X or P1); M symmetric, row-major
    g.M = (int)n1; g.N = nvec; g.K = (int)n1;
    g.A = H.nto_M.get(); g.sam = n1; g.sak = 1;
    g.B = pass == 0 ? x : P1; g.sbk = 1; g.sbn = pass == 0 ? ldx : n1;
    g.map.C = pass == 0 ? P1 : P2; g.map.scm = 1; g.map.scn = n1;
    gemm(g, Prec::F64, H.gws, st);
  }
  hipLaunchKernelGGL(nto_final_kernel, dim3(grid_for(n1 * nvec, 256, 16384)), dim3(256), 0, st, F.d.o, F.d.v, nvec, x,
                     ldx, P1, P2, H.nto_rho.get(), out);
  HIP_CHECK(hipGetLastError());
}

void trdens_vo_run(Handle& H, int nvec, const double* x_host, int64_t ldx, double* out_host) {
  ARG_CHECK(x_host && out_host, "NULL argument");
  ARG_CHECK(nvec >= 1 && ldx >= H.d.ndim, "nvec must be >= 1 and ldx >= ndim");
  const int64_t n1v = H.d.n1, n1 = H.parent ? H.parent->d.n1 : H.d.n1;
  DevBuf<double> dx, dout;
  dx.alloc((size_t)n1v * nvec);
  dout.alloc((size_t)n1 * nvec);
  nto_prepare(H);
  // only the 1h1p rows enter
  HIP_CHECK(hipMemcpy2DAsync(dx.get(), sizeof(double) * n1v, x_host, sizeof(double) * ldx, sizeof(double) * n1v, nvec,
                             hipMemcpyHostToDevice, H.st));
  trdens_vo_device(H, nvec, dx.get(), n1v, dout.get());
  HIP_CHECK(hipMemcpyAsync(out_host, dout.get(), sizeof(double) * n1 * nvec, hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
}

}  // namespace adc2g
