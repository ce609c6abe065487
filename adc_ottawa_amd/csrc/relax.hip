// Imaginary-time relaxation to the lowest eigenstates on the GPU
// sigma-vector: relaxation (source/propagation/relaxation.f90:36-133) with
// its default integrator, XSIL (xsil_rlx, :139-309).
//
// States are relaxed one after another.  One timestep of a state:
//  * collapse (:1163-1203): when the variational subspace holds maxdim/krydim
//    timesteps, the oldest block of krydim + 1 vectors and its rows and
//    columns of the subspace matrices are dropped;
//  * Lanczos block (get_lancvecs_current, :1207-1326): the first vector is
//    psi, orthogonalised against the converged states one after another and
//    normalised; krydim Lanczos steps r = H q - beta v, alpha = q.r,
//    r -= alpha q, beta = |r|, and the (krydim+1)-th vector r / beta.  The
//    steps run back to back on the stream: alpha and beta^2 stay in device
//    memory and the kernels that need them read them there;
//  * matrix elements with the earlier blocks (subspace_matrices,
//    :1330-1381): S = L_prev^T L_cur in one tall-skinny product, and H_ij from
//    the Lanczos relation, H l_j = beta_{k-1} l_{j-1} + alpha_k l_j + beta_k l_{j+1};
//  * propagation (solve_subspace_tdse, :1385-1491, canonical_ortho,
//    :1495-1606): on the host, the overlap is diagonalised, eigenvalues
//    |sigma| <= 1e-6 are dropped, and H' = X^T H X is diagonalised.  Then
//    c = X exp(-H' dtau) X^+ e_first, normalised, and psi = L c on the
//    device.  The propagated state is normalised; the reference's
//    orthogonalisation before that subtracts the projection on
//    vec_conv(:,s), which is zero until state s converges, so only the
//    normalisation is kept;
//  * E = psi.H psi and |H psi - E psi| (residual_1vec, :1610-1651).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int RL_THREADS = 256;

__device__ __forceinline__ void rl_block_partial(double s, double* part) {
  __shared__ double red[RL_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < RL_THREADS / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

#define RL_LOOP(i) \
  for (int64_t i = blockIdx.x * (int64_t)RL_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RL_THREADS)

// part[blk] = partial a.b
__global__ void __launch_bounds__(RL_THREADS) rl_dot_kernel(int64_t n, const double* __restrict__ a,
                                                            const double* __restrict__ b, double* part) {
  double s = 0.0;
  RL_LOOP(i) s += a[i] * b[i];
  rl_block_partial(s, part);
}

// y -= coef[0] x
__global__ void __launch_bounds__(RL_THREADS) rl_axmy_kernel(int64_t n, const double* __restrict__ coef,
                                                             const double* __restrict__ x, double* __restrict__ y) {
  const double c = coef[0];
  RL_LOOP(i) y[i] -= c * x[i];
}

// out = r / sqrt(b2[0]) (out may be r)
__global__ void __launch_bounds__(RL_THREADS) rl_scale_kernel(int64_t n, const double* r, const double* b2,
                                                              double* out) {
  const double inv = 1.0 / sqrt(b2[0]);
  RL_LOOP(i) out[i] = r[i] * inv;
}

// Lanczos step, first half: r -= sqrt(b2[0]) v (when b2 is given); part = partial q.r
__global__ void __launch_bounds__(RL_THREADS) rl_lanc_a_kernel(int64_t n, double* __restrict__ r,
                                                               const double* __restrict__ v, const double* b2,
                                                               const double* __restrict__ q, double* part) {
  const double beta = b2 ? sqrt(b2[0]) : 0.0;
  double s = 0.0;
  RL_LOOP(i) {
    double x = r[i];
    if (b2) x -= beta * v[i];
    r[i] = x;
    s += q[i] * x;
  }
  rl_block_partial(s, part);
}

// second half: r -= alpha[0] q; part = partial r.r
__global__ void __launch_bounds__(RL_THREADS) rl_lanc_b_kernel(int64_t n, double* __restrict__ r,
                                                               const double* __restrict__ q, const double* alpha,
                                                               double* part) {
  const double a = alpha[0];
  double s = 0.0;
  RL_LOOP(i) {
    const double x = r[i] - a * q[i];
    r[i] = x;
    s += x * x;
  }
  rl_block_partial(s, part);
}

// part = partial (hpsi - E[0] psi)^2
__global__ void __launch_bounds__(RL_THREADS) rl_res_kernel(int64_t n, const double* __restrict__ hpsi,
                                                            const double* __restrict__ psi, const double* E,
                                                            double* part) {
  const double e = E[0];
  double s = 0.0;
  RL_LOOP(i) {
    const double x = hpsi[i] - e * psi[i];
    s += x * x;
  }
  rl_block_partial(s, part);
}

#undef RL_LOOP

// sigma-vector time from event pairs read after the stream has been synchronised
struct SigmaClock {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  double seconds = 0.0;
  int64_t count = 0;
  ~SigmaClock() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  hipEvent_t next() {
    if (used == ev.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      ev.push_back(e);
    }
    return ev[used++];
  }
  // after a stream synchronisation
  void collect() {
    for (size_t i = 0; i + 1 < used; i += 2) {
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      seconds += ms * 1e-3;
    }
    used = 0;
  }
};

}  // namespace

#define RL_LOOP(i) \
  for (int64_t i = blockIdx.x * (int64_t)RL_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RL_THREADS)

// ---- Bulirsch-Stoer (bslib.f90: BSStep, ModMidpoint, PolyExtrapol, AbsBSError)
// for the relaxation right-hand side f(psi) = -H psi.  The reference works in
// complex arithmetic, but with a real start vector and f real every
// imaginary part stays exactly zero, so the vectors here are real.

// psi1 = saved, psi2 = saved + h dt
__global__ void __launch_bounds__(RL_THREADS) bs_mm_start_kernel(int64_t n, const double* __restrict__ saved,
                                                                 const double* __restrict__ dt, double h,
                                                                 double* __restrict__ psi1, double* __restrict__ psi2) {
  RL_LOOP(i) {
    psi1[i] = saved[i];
    psi2[i] = saved[i] + h * dt[i];
  }
}

// est = -w;  swap = psi1 + 2h est;  psi1 = psi2;  psi2 = swap
__global__ void __launch_bounds__(RL_THREADS) bs_mm_step_kernel(int64_t n, const double* __restrict__ w, double h2,
                                                                double* __restrict__ psi1, double* __restrict__ psi2) {
  RL_LOOP(i) {
    const double swap = psi1[i] + h2 * (-w[i]);
    psi1[i] = psi2[i];
    psi2[i] = swap;
  }
}

// out = 0.5 (psi1 + psi2 + h est), est = -w
__global__ void __launch_bounds__(RL_THREADS) bs_mm_final_kernel(int64_t n, const double* __restrict__ w, double h,
                                                                 const double* __restrict__ psi1,
                                                                 const double* __restrict__ psi2,
                                                                 double* __restrict__ out) {
  RL_LOOP(i) out[i] = 0.5 * (psi1[i] + psi2[i] + h * (-w[i]));
}

constexpr int BS_MAXORDER = 16;
struct BSFactors {
  double f[BS_MAXORDER + 1], g[BS_MAXORDER + 1];
};

// PolyExtrapol for one column (tableau columns 1..IntOrder-1 at T + (c-1)*n)
// fused with AbsBSError: part[blk] = max |DeltaPsi| over the block
__global__ void __launch_bounds__(RL_THREADS) bs_extrapol_kernel(int64_t n, int column, int int_order, BSFactors fac,
                                                                 double* __restrict__ psi, double* __restrict__ T,
                                                                 double* __restrict__ part) {
  double mx = 0.0;
  RL_LOOP(i) {
    double x = psi[i];
    double dpsi = x;
    if (column == 1) {
      T[i] = x;
    } else {
      double diff = x;
      for (int q = 1; q < column; ++q) {
        double* Tq = T + (size_t)n * (q - 1);
        const double interim = Tq[i];
        Tq[i] = dpsi;
        const double delta = diff - interim;
        dpsi = fac.f[q] * delta;
        diff = fac.g[q] * delta;
        x = x + dpsi;
      }
      if (column < int_order) T[(size_t)n * (column - 1) + i] = dpsi;
      psi[i] = x;
    }
    mx = fmax(mx, fabs(dpsi));
  }
  __shared__ double red[RL_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < RL_THREADS / 64; ++k) t = fmax(t, red[k]);
    part[blockIdx.x] = t;
  }
}

// ---- RKF45 (rkf45rlxlib.f90: r8_fehl, r8_rkf45) for f(y) = -H y; the
// stage combinations keep the reference's grouping (no contraction)

// f = -w
__global__ void __launch_bounds__(RL_THREADS) rk_neg_kernel(int64_t n, const double* __restrict__ w,
                                                            double* __restrict__ f) {
  RL_LOOP(i) f[i] = -w[i];
}

// x = y + ch * (stage combination of yp, f1..f5)
__global__ void __launch_bounds__(RL_THREADS) rk_stage_kernel(int64_t n, int stage, double ch,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ yp,
                                                              const double* __restrict__ f1,
                                                              const double* __restrict__ f2,
                                                              const double* __restrict__ f3,
                                                              const double* __restrict__ f4,
                                                              const double* __restrict__ f5, double* __restrict__ x) {
#pragma clang fp contract(off)
  RL_LOOP(i) {
    double c;
    switch (stage) {
      case 1: c = yp[i]; break;
      case 2: c = yp[i] + 3.0 * f1[i]; break;
      case 3: c = 1932.0 * yp[i] + (7296.0 * f2[i] - 7200.0 * f1[i]); break;
      case 4: c = (8341.0 * yp[i] - 845.0 * f3[i]) + (29440.0 * f2[i] - 32832.0 * f1[i]); break;
      case 5: c = (-6080.0 * yp[i] + (9295.0 * f3[i] - 5643.0 * f4[i])) + (41040.0 * f1[i] - 28352.0 * f2[i]); break;
      default:
        c = (902880.0 * yp[i] + (3855735.0 * f3[i] - 1371249.0 * f4[i])) + (3953664.0 * f2[i] + 277020.0 * f5[i]);
        break;
    }
    x[i] = y[i] + ch * c;
  }
}

// part[blk] = max_k ee_k / et_k, et = |y| + |s| + ae
__global__ void __launch_bounds__(RL_THREADS) rk_err_kernel(int64_t n, double ae, const double* __restrict__ y,
                                                            const double* __restrict__ s,
                                                            const double* __restrict__ yp,
                                                            const double* __restrict__ f2,
                                                            const double* __restrict__ f3,
                                                            const double* __restrict__ f4,
                                                            const double* __restrict__ f5, double* __restrict__ part) {
#pragma clang fp contract(off)
  double mx = 0.0;
  RL_LOOP(i) {
    const double et = fabs(y[i]) + fabs(s[i]) + ae;
    const double ee = fabs((-2090.0 * yp[i] + (21970.0 * f3[i] - 15048.0 * f4[i])) +
                           (22528.0 * f2[i] - 27360.0 * f5[i]));
    mx = fmax(mx, ee / et);
  }
  __shared__ double red[RL_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < RL_THREADS / 64; ++k) t = fmax(t, red[k]);
    part[blockIdx.x] = t;
  }
}

// r8_rkf45's save variables
struct RKState {
  double abserr_save = -1.0, h = -1.0, relerr_save = -1.0;
  int flag_save = -1000, init = -1000, kflag = -1000, kop = -1, nfe = -1;
};

#undef RL_LOOP

// BSStep with its Fortran save variables as members (they persist from call
// to call, as in the reference).
struct BSEngine {
  int optimal_column = 0;
  int nsteps[BS_MAXORDER + 2] = {};
  double int_accuracy = 0.0, new_time = 0.0;
  double expense[BS_MAXORDER + 2] = {};
  double correction[BS_MAXORDER + 1][BS_MAXORDER + 1] = {};
  double old_accuracy = -1.0;
  int old_int_order = -1;
  bool first_call = true;
  double old_sq[BS_MAXORDER + 1] = {};
};

void relax_run(Handle& H, const adc2g_relax_params& p, const double* guess_h, double* eigval, double* eigvec,
               double* resid_out, int32_t* niter_out, adc2g_relax_stats* stats) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  const int64_t n = H.d.ndim;
  const int ns = p.nstates, kd = p.krydim;
  ARG_CHECK(ns >= 1 && kd >= 1 && p.maxiter >= 1, "nstates, krydim and maxiter must be >= 1");
  ARG_CHECK(std::isfinite(p.step) && p.step > 0.0, "step must be finite and positive");
  ARG_CHECK(p.tol >= 0.0, "tol must be >= 0");
  ARG_CHECK((int64_t)kd + 1 <= n && (int64_t)ns <= n, "krydim + 1 and nstates must not exceed ndim");
  const int maxdim = p.maxdim <= 0 ? 2 * kd : p.maxdim;
  ARG_CHECK(maxdim % kd == 0, "maxdim must be an integer multiple of krydim");
  const int maxvec = maxdim + maxdim / kd;  // stored vectors, krydim + 1 per timestep
  const int kb = kd + 1;
  const Prec pr = prec_from_int(p.precision);
  hipStream_t st = H.st;
  const int nblk = reduce_blocks(n);
  const bool bs_mode = p.integrator == 2;
  const bool rk_mode = p.integrator == 3;
  ARG_CHECK(p.integrator >= 0 && p.integrator <= 3,
            "integrator must be 0 or 1 (XSIL), 2 (Bulirsch-Stoer) or 3 (RKF45)");
  ARG_CHECK(!(bs_mode || rk_mode) || (std::isfinite(p.toler) && p.toler > 0.0),
            "toler must be positive for Bulirsch-Stoer and RKF45");

  DevBuf<double> L, r, psi, hpsi, guess, conv, part, sc, coef, Sx;
  L.alloc((size_t)n * maxvec);
  r.alloc(n);
  psi.alloc(n);
  hpsi.alloc(n);
  guess.alloc((size_t)n * ns);
  conv.alloc((size_t)n * ns);
  part.alloc(nblk);
  // device scalars: alpha[kd], beta^2[kd], then 8 work slots
  sc.alloc((size_t)2 * kd + 8);
  coef.alloc(maxvec);
  Sx.alloc((size_t)maxvec * kb);
  double* alpha_d = sc.get();
  double* b2_d = sc.get() + kd;
  double* w_d = sc.get() + 2 * kd;

  if (guess_h) {
    HIP_CHECK(hipMemcpyAsync(guess.get(), guess_h, sizeof(double) * n * ns, hipMemcpyHostToDevice, st));
  } else {
    // initvec_ondiag: unit vectors on the ns smallest |H_jj| (ties in index order)
    std::vector<double> hii(n);
    HIP_CHECK(hipMemcpyAsync(hii.data(), H.hii.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + ns, order.end(), [&](int64_t a, int64_t b) {
      const double fa = std::fabs(hii[a]), fb = std::fabs(hii[b]);
      return fa < fb || (fa == fb && a < b);
    });
    DevBuf<int64_t> didx;
    didx.alloc(ns);
    HIP_CHECK(hipMemcpyAsync(didx.get(), order.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(guess.get(), 0, sizeof(double) * n * ns, st));
    launch_unit_vectors(ns, didx.get(), guess.get(), n, st);
    HIP_CHECK(hipStreamSynchronize(st));
  }

  SigmaClock clock;
  auto sigma1 = [&](const double* x, double* w) {
    HIP_CHECK(hipEventRecord(clock.next(), st));
    H.sigma(x, n, w, n, 1, pr, st);
    HIP_CHECK(hipEventRecord(clock.next(), st));
    ++clock.count;
  };
  auto dot = [&](const double* a, const double* b, double* out_d) {
    hipLaunchKernelGGL(rl_dot_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, a, b, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 1, part.get(), out_d, st);
  };
  auto normalise = [&](double* x, double* slot_d) {
    dot(x, x, slot_d);
    hipLaunchKernelGGL(rl_scale_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, x, slot_d, x);
    HIP_CHECK(hipGetLastError());
  };
  // E = x.Hx and |Hx - E x| (one sigma-vector, one download)
  auto residual = [&](const double* x, double& E, double& res) {
    sigma1(x, hpsi.get());
    dot(x, hpsi.get(), w_d + 3);
    hipLaunchKernelGGL(rl_res_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, hpsi.get(), x, w_d + 3, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 1, part.get(), w_d + 4, st);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, w_d + 3, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    clock.collect();
    E = h[0];
    res = std::sqrt(h[1]);
  };

  // host copies of the subspace matrices (column-major, ld = maxvec)
  std::vector<double> subh((size_t)maxvec * maxvec), subs((size_t)maxvec * maxvec), tmp((size_t)maxvec * maxvec);
  std::vector<double> alpha(kd), beta(kd), b2h(kd), sx((size_t)maxvec * kb);
  auto SH = [&](std::vector<double>& m, int i, int j) -> double& { return m[i + (size_t)maxvec * j]; };

  // Bulirsch-Stoer state (bs_rlx: intorder 16, nextstep 0, time 0) and
  // vectors: saved psi, dtpsi (as H psi), the midpoint pair, the tableau
  BSEngine bse;
  int bs_order = BS_MAXORDER;
  double bs_next = 0.0, bs_time = 0.0;
  DevBuf<double> bs_saved, bs_w, bs_p1, bs_p2, bs_tab, bs_dt;
  std::vector<double> bs_part(nblk);
  if (bs_mode) {
    bs_saved.alloc(n);
    bs_w.alloc(n);
    bs_dt.alloc(n);
    bs_p1.alloc(n);
    bs_p2.alloc(n);
    bs_tab.alloc((size_t)n * BS_MAXORDER);
  }
  // one BSStep call; psi holds Psi, bs_w holds H psi (DtPsi = -bs_w)
  auto bs_step = [&](double int_period, double abs_time, double large_step, double tol_error,
                     double& actual_out) -> int {
    const double RelativeMinStep = 1.0e-10, MaxScale = 0.1, ErrorSafety = 0.25, ReductionSafety = 0.7,
                 MinReduction = 0.5, MaxReduction = 1.0e-5, Tiny = 1.0e-25;
    int& io = bs_order;
    if (io < 2 || io > BS_MAXORDER) return 1;
    const double min_large = RelativeMinStep * int_period;
    double next_large = -Tiny;
    bse.new_time = -Tiny;
    if (tol_error != bse.old_accuracy || io != bse.old_int_order) {
      bse.old_accuracy = tol_error;
      bse.int_accuracy = ErrorSafety * tol_error;
      bse.old_int_order = io;
      for (int q = 1; q <= io + 1; ++q) bse.nsteps[q] = 2 * q;
      bse.expense[1] = bse.nsteps[1] + 1;
      for (int q = 2; q <= io + 1; ++q) bse.expense[q] = bse.expense[q - 1] + bse.nsteps[q];
      for (int q = 2; q <= io; ++q)
        for (int q1 = 1; q1 < q; ++q1)
          bse.correction[q1][q] =
              std::pow(bse.int_accuracy, (bse.expense[q1 + 1] - bse.expense[q + 1]) /
                                             ((2 * q1 + 1) * (bse.expense[q + 1] - bse.expense[1] + 1.0)));
      int oc = 2;
      while (oc <= io - 1) {
        if (bse.expense[oc + 1] > bse.expense[oc] * bse.correction[oc - 1][oc]) break;
        ++oc;
      }
      bse.optimal_column = oc;
      io = oc;
    }
    double actual = large_step;
    bool reduced = false;
    double min_work = 1.0e35;
    HIP_CHECK(hipMemcpyAsync(bs_saved.get(), psi.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    // DtPsi = -H psi, kept apart from bs_w, which the midpoint steps reuse
    HIP_CHECK(hipMemcpyAsync(bs_dt.get(), bs_w.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (actual != next_large || abs_time != bse.new_time) {
      bse.first_call = true;
      bse.optimal_column = io;
    }
    double err[BS_MAXORDER] = {};
    int q = 0, q1 = 0;
    for (;;) {  // 300
      double reduction = 0.0;
      bool conv = false;
      q1 = 0;
      for (q = 1; q <= io; ++q) {
        bse.new_time = abs_time + actual;
        const double sq = (actual / bse.nsteps[q]) * (actual / bse.nsteps[q]);
        if (actual < min_large && bse.new_time < int_period) return 2;
        // ModMidpoint: dt = -H saved is -bs_dt
        const int ns_q = bse.nsteps[q];
        const double h = actual / ns_q, h2 = 2.0 * h;
        hipLaunchKernelGGL(bs_mm_start_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_saved.get(), bs_dt.get(),
                           -h, bs_p1.get(), bs_p2.get());
        HIP_CHECK(hipGetLastError());
        sigma1(bs_p2.get(), bs_w.get());
        for (int k = 2; k <= ns_q; ++k) {
          hipLaunchKernelGGL(bs_mm_step_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_w.get(), h2, bs_p1.get(),
                             bs_p2.get());
          HIP_CHECK(hipGetLastError());
          sigma1(bs_p2.get(), bs_w.get());
        }
        hipLaunchKernelGGL(bs_mm_final_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_w.get(), h, bs_p1.get(),
                           bs_p2.get(), psi.get());
        HIP_CHECK(hipGetLastError());
        // PolyExtrapol + AbsBSError
        bse.old_sq[q] = sq;
        BSFactors fac{};
        for (int pp = 1; pp < q; ++pp) {
          const double hl = 1.0 / (bse.old_sq[q - pp] - sq);
          fac.f[pp] = sq * hl;
          fac.g[pp] = bse.old_sq[q - pp] * hl;
        }
        hipLaunchKernelGGL(bs_extrapol_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, q, io, fac, psi.get(),
                           bs_tab.get(), part.get());
        HIP_CHECK(hipGetLastError());
        double max_error = 0.0;
        q1 = 0;
        if (q >= 2) {
          HIP_CHECK(hipMemcpyAsync(bs_part.data(), part.get(), sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
          HIP_CHECK(hipStreamSynchronize(st));
          clock.collect();
          for (double x : bs_part) max_error = std::max(max_error, x);
          q1 = q - 1;
          err[q1] = std::pow(max_error / bse.int_accuracy, 1.0 / (2 * q1 + 1.0));
        }
        if (q >= 2 && (q >= bse.optimal_column - 1 || bse.first_call)) {
          if (max_error < tol_error) {
            conv = true;
            break;
          }
          const int oc = bse.optimal_column;
          if (q == io || q == oc + 1) {
            reduction = ReductionSafety / err[q1];
            break;
          } else if (q == oc) {
            if (bse.correction[oc - 1][oc] < err[q1]) {
              reduction = 1.0 / err[q1];
              break;
            }
          } else if (oc == io) {
            if (bse.correction[q1][io - 1] < err[q1]) {
              reduction = ReductionSafety * bse.correction[q1][io - 1] / err[q1];
              break;
            }
          } else if (bse.correction[q1][oc] < err[q1]) {
            reduction = bse.correction[q1][oc - 1] / err[q1];
            break;
          }
        }
      }
      if (conv) break;
      reduction = std::max(std::min(reduction, MinReduction), MaxReduction);  // 400
      actual *= reduction;
      reduced = true;
    }
    bse.first_call = false;  // 500
    double scale = 0.0;
    for (int q2 = 1; q2 <= q1; ++q2) {
      const double wf = std::max(err[q2], MaxScale);
      const double work = wf * bse.expense[q2 + 1];
      if (work < min_work) {
        scale = wf;
        min_work = work;
        bse.optimal_column = q2 + 1;
      }
    }
    next_large = actual / scale;
    const int oc = bse.optimal_column;
    if (oc >= q && oc < io && !reduced) {
      const double wf = std::max(scale / bse.correction[oc - 1][oc], MaxScale);
      if (bse.expense[oc + 1] * wf <= min_work) {
        next_large = actual / wf;
        bse.optimal_column = oc + 1;
      }
    }
    bs_next = next_large;
    actual_out = actual;
    return 0;
  };

  // RKF45 state and vectors (y is psi; yp persists, as the reference's dtpsi)
  RKState rks;
  DevBuf<double> rk_yp, rk_f[5], rk_x, rk_w;
  std::vector<double> rk_part(nblk);
  if (rk_mode) {
    rk_yp.alloc(n);
    for (auto& b : rk_f) b.alloc(n);
    rk_x.alloc(n);
    rk_w.alloc(n);
    HIP_CHECK(hipMemsetAsync(rk_yp.get(), 0, sizeof(double) * n, st));
  }
  auto rk_f_of = [&](const double* x, double* out) {  // out = -H x
    sigma1(x, rk_w.get());
    hipLaunchKernelGGL(rk_neg_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, rk_w.get(), out);
    HIP_CHECK(hipGetLastError());
  };
  auto rk_stage = [&](int stage, double ch, double* out) {
    hipLaunchKernelGGL(rk_stage_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, stage, ch, psi.get(), rk_yp.get(),
                       rk_f[0].get(), rk_f[1].get(), rk_f[2].get(), rk_f[3].get(), rk_f[4].get(), out);
    HIP_CHECK(hipGetLastError());
  };
  // r8_fehl: the solution s ends in rk_f[0] (the reference passes f1 as s)
  auto rk_fehl = [&](double h) {
    rk_stage(1, h / 4.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[0].get());
    rk_stage(2, 3.0 * h / 32.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[1].get());
    rk_stage(3, h / 2197.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[2].get());
    rk_stage(4, h / 4104.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[3].get());
    rk_stage(5, h / 20520.0, rk_x.get());  // the reference stores this point in f1, used once
    rk_f_of(rk_x.get(), rk_f[4].get());
    rk_stage(6, h / 7618050.0, rk_f[0].get());
  };
  // one r8_rkf45 call (flag 1 or 2 as rkf45_rlx uses it); returns the flag
  auto rk_step = [&](double& t, double tout, double relerr, double abserr, int flag) -> int {
    const double eps = DBL_EPSILON;
    if (relerr < 0.0 || abserr < 0.0 || flag == 0 || flag > 8 || flag < -2) return 8;
    int mflag = std::abs(flag);
    if (mflag != 1) {
      if (t == tout && rks.kflag != 3) return 8;
      if (mflag == 2) {
        if (rks.kflag == 3) {
          flag = rks.flag_save;
          mflag = std::abs(flag);
        } else if (rks.init == 0) {
          flag = rks.flag_save;
        } else if (rks.kflag == 4) {
          rks.nfe = 0;
        } else if (rks.kflag == 5 && abserr == 0.0) {
          return 5;
        } else if (rks.kflag == 6 && relerr <= rks.relerr_save && abserr <= rks.abserr_save) {
          return 6;
        }
      }
    }
    rks.flag_save = flag;
    rks.kflag = 0;
    rks.relerr_save = relerr;
    rks.abserr_save = abserr;
    if (relerr < 2.0 * eps + 1.0e-12) {
      rks.kflag = 3;
      return 3;
    }
    double dt = tout - t;
    if (mflag == 1) {
      rks.init = 0;
      rks.kop = 0;
      rk_f_of(psi.get(), rk_yp.get());
      rks.nfe = 1;
      if (t == tout) return 2;
    }
    if (rks.init == 0) {
      // the initial step from the components, in order (host copies of y, yp)
      rks.init = 1;
      std::vector<double> hy(n), hyp(n);
      HIP_CHECK(hipMemcpyAsync(hy.data(), psi.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(hyp.data(), rk_yp.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      clock.collect();
      double h = std::fabs(dt), toln = 0.0;
      for (int64_t k = 0; k < n; ++k) {
        const double tol = relerr * std::fabs(hy[k]) + abserr;
        if (tol > 0.0) {
          toln = tol;
          const double ypk = std::fabs(hyp[k]);
          if (tol < ypk * std::pow(h, 5)) h = std::pow(tol / ypk, 0.2);
        }
      }
      if (toln <= 0.0) h = 0.0;
      rks.h = std::max(h, 26.0 * eps * std::max(std::fabs(t), std::fabs(dt)));
      rks.flag_save = flag >= 0 ? 2 : -2;
    }
    rks.h = dt >= 0.0 ? std::fabs(rks.h) : -std::fabs(rks.h);
    if (2.0 * std::fabs(dt) <= std::fabs(rks.h)) ++rks.kop;
    if (rks.kop == 10000) {
      rks.kop = 0;
      return 7;
    }
    if (std::fabs(dt) <= 26.0 * eps * std::fabs(t)) {
      t = tout;
      rk_stage(1, dt, rk_x.get());  // y + dt yp
      HIP_CHECK(hipMemcpyAsync(psi.get(), rk_x.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      rk_f_of(psi.get(), rk_yp.get());
      ++rks.nfe;
      return 2;
    }
    bool output = false;
    const double scale = 2.0 / relerr, ae = scale * abserr;
    for (;;) {
      bool hfaild = false;
      const double hmin = 26.0 * eps * std::fabs(t);
      dt = tout - t;
      if (!(2.0 * std::fabs(rks.h) <= std::fabs(dt))) {
        if (std::fabs(dt) <= std::fabs(rks.h)) {
          output = true;
          rks.h = dt;
        } else {
          rks.h = 0.5 * dt;
        }
      }
      double esttol = 0.0;
      for (;;) {
        if (rks.nfe > 30000) {
          rks.kflag = 4;
          return 4;
        }
        rk_fehl(rks.h);
        rks.nfe += 5;
        hipLaunchKernelGGL(rk_err_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, ae, psi.get(), rk_f[0].get(),
                           rk_yp.get(), rk_f[1].get(), rk_f[2].get(), rk_f[3].get(), rk_f[4].get(), part.get());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(rk_part.data(), part.get(), sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        clock.collect();
        double eeoet = 0.0;
        for (double x : rk_part) eeoet = std::max(eeoet, x);
        esttol = std::fabs(rks.h) * eeoet * scale / 752400.0;
        if (esttol <= 1.0) break;
        hfaild = true;
        output = false;
        const double sc = esttol < 59049.0 ? 0.9 / std::pow(esttol, 0.2) : 0.1;
        rks.h = sc * rks.h;
        if (std::fabs(rks.h) < hmin) {
          rks.kflag = 6;
          return 6;
        }
      }
      t = t + rks.h;
      HIP_CHECK(hipMemcpyAsync(psi.get(), rk_f[0].get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      rk_f_of(psi.get(), rk_yp.get());
      ++rks.nfe;
      double sc = 0.0001889568 < esttol ? 0.9 / std::pow(esttol, 0.2) : 5.0;
      if (hfaild) sc = std::min(sc, 1.0);
      const double hh = std::max(sc * std::fabs(rks.h), hmin);
      rks.h = rks.h >= 0.0 ? hh : -hh;
      if (output) {
        t = tout;
        return 2;
      }
      if (flag <= 0) return -2;
    }
  };

  int nconv = 0;
  int status = ADC2G_OK;
  std::string fail_msg;
  for (int s = 0; s < ns; ++s) {
    std::fill(subh.begin(), subh.end(), 0.0);
    std::fill(subs.begin(), subs.end(), 0.0);
    int ncurr = 0;
    HIP_CHECK(hipMemcpyAsync(psi.get(), guess.get() + (size_t)n * s, sizeof(double) * n, hipMemcpyDeviceToDevice,
                             st));
    double E = 0.0, res = 0.0;
    residual(psi.get(), E, res);
    if (p.verbose) std::fprintf(stderr, "relax state %d step %5d  E = %.12f  residual = %.4e\n", s + 1, 0, E, res);
    int it_done = p.maxiter;
    bool converged = false;
    if (bs_mode) {
      // bs_rlx (relaxation.f90:315-495): int(tfinal/intperiod) timesteps with
      // tfinal = maxiter*step, each covered by BSStep calls
      const double intperiod = p.step, tfinal = (double)p.maxiter * p.step, tiny = 1e-9;
      const int nt = (int)(tfinal / intperiod);
      it_done = nt;
      for (int it = 1; it <= nt; ++it) {
        double inttime = 0.0;
        for (;;) {
          double stepsize;
          if (inttime == 0.0) stepsize = intperiod;
          else if (inttime + bs_next > intperiod) stepsize = intperiod - inttime;
          else stepsize = bs_next;
          sigma1(psi.get(), bs_w.get());  // dtpsi = -H psi
          double true_step = 0.0;
          const int code = bs_step(intperiod, bs_time, stepsize, p.toler, true_step);
          if (code != 0)
            throw Error(ADC2G_EBREAK, std::string("Failure in the Bulirsch-Stoer integrator: ") +
                                          (code == 1 ? "Illegal integration order" : "Stepsize underflow"));
          bs_time += true_step;
          inttime += true_step;
          if (!(std::fabs(intperiod - inttime) > std::fabs(tiny * intperiod))) break;
        }
        normalise(psi.get(), w_d + 2);
        residual(psi.get(), E, res);
        if (!std::isfinite(E) || !std::isfinite(res))
          throw Error(ADC2G_EBREAK, "relaxation: non-finite energy at state " + std::to_string(s + 1));
        if (p.verbose)
          std::fprintf(stderr, "relax state %d step %5d  E = %.12f  residual = %.4e\n", s + 1, it, E, res);
        if (res <= p.tol) {
          converged = true;
          it_done = it;
          break;
        }
      }
    }
    if (rk_mode) {
      // rkf45_rlx (relaxation.f90:501-621): tout = i*step, relerr = abserr =
      // siltol, flag 1 at the start of the state
      double t = 0.0;
      int flag = 1;
      for (int it = 1; it <= p.maxiter; ++it) {
        flag = rk_step(t, it * p.step, p.toler, p.toler, flag);
        if (flag != 2)
          throw Error(ADC2G_EBREAK, "Error in the RKF45 integrator. Error code: " + std::to_string(flag));
        normalise(psi.get(), w_d + 2);
        residual(psi.get(), E, res);
        if (!std::isfinite(E) || !std::isfinite(res))
          throw Error(ADC2G_EBREAK, "relaxation: non-finite energy at state " + std::to_string(s + 1));
        if (p.verbose)
          std::fprintf(stderr, "relax state %d step %5d  E = %.12f  residual = %.4e\n", s + 1, it, E, res);
        if (res <= p.tol) {
          converged = true;
          it_done = it;
          break;
        }
      }
    }
    for (int it = 1; !bs_mode && !rk_mode && it <= p.maxiter; ++it) {
      if (ncurr * kb == maxvec) {  // collapse_subspace
        for (int blk = 1; blk < maxvec / kb; ++blk)
          HIP_CHECK(hipMemcpyAsync(L.get() + (size_t)n * kb * (blk - 1), L.get() + (size_t)n * kb * blk,
                                   sizeof(double) * n * kb, hipMemcpyDeviceToDevice, st));
        const int keep = maxvec - kd;
        for (auto* m : {&subs, &subh}) {
          for (int j = 0; j < keep; ++j)
            for (int i = 0; i < keep; ++i) tmp[i + (size_t)keep * j] = SH(*m, kd + i, kd + j);
          for (int j = 0; j < keep; ++j)
            for (int i = 0; i < keep; ++i) SH(*m, i, j) = tmp[i + (size_t)keep * j];
        }
        --ncurr;
      }
      ++ncurr;
      const int k1 = (ncurr - 1) * kb;  // first vector of this timestep in L
      const int b0 = (ncurr - 1) * kd;  // its row/column in the subspace matrices
      double* Lk = L.get() + (size_t)n * k1;
      // first vector: psi orthogonalised against the converged states, normalised
      HIP_CHECK(hipMemcpyAsync(Lk, psi.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      for (int c = 0; c < nconv; ++c) {
        const double* vc = conv.get() + (size_t)n * c;
        dot(vc, Lk, w_d);
        hipLaunchKernelGGL(rl_axmy_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, w_d, vc, Lk);
        HIP_CHECK(hipGetLastError());
      }
      normalise(Lk, w_d + 1);
      // krydim Lanczos steps and the extra vector
      for (int j = 0; j < kd; ++j) {
        const double* q = Lk + (size_t)n * j;
        sigma1(q, r.get());
        hipLaunchKernelGGL(rl_lanc_a_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, r.get(),
                           j ? q - n : nullptr, j ? b2_d + j - 1 : nullptr, q, part.get());
        HIP_CHECK(hipGetLastError());
        launch_colsum(nblk, 1, part.get(), alpha_d + j, st);
        hipLaunchKernelGGL(rl_lanc_b_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, r.get(), q, alpha_d + j,
                           part.get());
        HIP_CHECK(hipGetLastError());
        launch_colsum(nblk, 1, part.get(), b2_d + j, st);
        hipLaunchKernelGGL(rl_scale_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, r.get(), b2_d + j,
                           Lk + (size_t)n * (j + 1));
        HIP_CHECK(hipGetLastError());
      }
      if (ncurr > 1) gemm_tn(H, k1, kb, n, L.get(), n, Lk, n, Sx.get(), k1, 1.0, 0.0);
      HIP_CHECK(hipMemcpyAsync(alpha.data(), alpha_d, sizeof(double) * kd, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(b2h.data(), b2_d, sizeof(double) * kd, hipMemcpyDeviceToHost, st));
      if (ncurr > 1)
        HIP_CHECK(hipMemcpyAsync(sx.data(), Sx.get(), sizeof(double) * k1 * kb, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      clock.collect();
      for (int j = 0; j < kd; ++j) {
        if (!(b2h[j] > 0.0) || !std::isfinite(b2h[j]) || !std::isfinite(alpha[j]))
          throw Error(ADC2G_EBREAK, "relaxation: Lanczos breakdown at state " + std::to_string(s + 1) +
                                        ", step " + std::to_string(it));
        beta[j] = std::sqrt(b2h[j]);
      }
      for (int i = 0; i < kd; ++i) {
        SH(subh, b0 + i, b0 + i) = alpha[i];
        SH(subs, b0 + i, b0 + i) = 1.0;
      }
      for (int i = 0; i + 1 < kd; ++i) {
        SH(subh, b0 + i, b0 + i + 1) = beta[i];
        SH(subh, b0 + i + 1, b0 + i) = beta[i];
      }
      // matrix elements with the vectors of the earlier timesteps
      for (int np = 0; np + 1 < ncurr; ++np)
        for (int ii = 0; ii < kd; ++ii) {
          const int i = np * kb + ii, i1 = np * kd + ii;
          auto ov = [&](int col) { return sx[i + (size_t)k1 * col]; };  // l_i . l_{k1+col}
          for (int k = 0; k < kd; ++k) {
            const int j1 = b0 + k;
            const double hij = (k == 0) ? alpha[k] * ov(k) + beta[k] * ov(k + 1)
                                        : beta[k - 1] * ov(k - 1) + alpha[k] * ov(k) + beta[k] * ov(k + 1);
            SH(subh, i1, j1) = hij;
            SH(subh, j1, i1) = hij;
            SH(subs, i1, j1) = ov(k);
            SH(subs, j1, i1) = ov(k);
          }
        }
      // canonical orthogonalisation and propagation in the subspace (host)
      const int nv = ncurr * kd;
      std::vector<double> S((size_t)nv * nv), Hs((size_t)nv * nv), sig(nv), U((size_t)nv * nv);
      for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nv; ++i) {
          S[i + (size_t)nv * j] = SH(subs, i, j);
          Hs[i + (size_t)nv * j] = SH(subh, i, j);
        }
      if (!host_syev(nv, S.data(), sig.data(), U.data()))
        throw Error(ADC2G_EEIG, "relaxation: diagonalisation of the subspace overlap matrix failed");
      int nlin = 0;
      for (int i = 0; i < nv; ++i)
        if (std::fabs(sig[i]) > 1e-6) ++nlin;
      const int nnull = nv - nlin;
      std::vector<double> X((size_t)nv * nlin), c0(nlin), T1((size_t)nv * nlin), H1((size_t)nlin * nlin);
      for (int c = 0; c < nlin; ++c) {
        const double sv = sig[nnull + c], rs = std::sqrt(sv);
        for (int i = 0; i < nv; ++i) X[i + (size_t)nv * c] = U[i + (size_t)nv * (nnull + c)] / rs;
        c0[c] = U[b0 + (size_t)nv * (nnull + c)] * rs;
      }
      // H1 = X^T (Hs X)
      for (int c = 0; c < nlin; ++c)
        for (int i = 0; i < nv; ++i) {
          double t = 0.0;
          for (int k = 0; k < nv; ++k) t += Hs[i + (size_t)nv * k] * X[k + (size_t)nv * c];
          T1[i + (size_t)nv * c] = t;
        }
      for (int c = 0; c < nlin; ++c)
        for (int rr = 0; rr < nlin; ++rr) {
          double t = 0.0;
          for (int k = 0; k < nv; ++k) t += X[k + (size_t)nv * rr] * T1[k + (size_t)nv * c];
          H1[rr + (size_t)nlin * c] = t;
        }
      std::vector<double> lam(nlin), V((size_t)nlin * nlin);
      if (!host_syev(nlin, H1.data(), lam.data(), V.data()))
        throw Error(ADC2G_EEIG, "relaxation: diagonalisation of the subspace Hamiltonian failed");
      // c1 = V exp(-lam dtau) V^T c0, normalised; coefficients X c1
      std::vector<double> y(nlin), c1(nlin, 0.0);
      for (int k = 0; k < nlin; ++k) {
        double t = 0.0;
        for (int i = 0; i < nlin; ++i) t += V[i + (size_t)nlin * k] * c0[i];
        y[k] = std::exp(-lam[k] * p.step) * t;
      }
      for (int k = 0; k < nlin; ++k)
        for (int i = 0; i < nlin; ++i) c1[i] += V[i + (size_t)nlin * k] * y[k];
      double nrm = 0.0;
      for (int i = 0; i < nlin; ++i) nrm += c1[i] * c1[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < nlin; ++i) c1[i] /= nrm;
      // coefficients over the stored vectors (zero on the extra vector of each block)
      const int ncols = ncurr * kb;
      std::vector<double> cf(ncols, 0.0);
      for (int blk = 0; blk < ncurr; ++blk)
        for (int i = 0; i < kd; ++i) {
          double t = 0.0;
          for (int c = 0; c < nlin; ++c) t += X[(blk * kd + i) + (size_t)nv * c] * c1[c];
          cf[blk * kb + i] = t;
        }
      HIP_CHECK(hipMemcpyAsync(coef.get(), cf.data(), sizeof(double) * ncols, hipMemcpyHostToDevice, st));
      gemm_nn(H, n, 1, ncols, L.get(), n, coef.get(), ncols, psi.get(), n, 1.0, 0.0);
      normalise(psi.get(), w_d + 2);
      residual(psi.get(), E, res);
      if (!std::isfinite(E) || !std::isfinite(res))
        throw Error(ADC2G_EBREAK, "relaxation: non-finite energy at state " + std::to_string(s + 1));
      if (p.verbose)
        std::fprintf(stderr, "relax state %d step %5d  E = %.12f  residual = %.4e\n", s + 1, it, E, res);
      if (res <= p.tol) {
        converged = true;
        it_done = it;
        break;
      }
    }
    eigval[s] = E;
    if (resid_out) resid_out[s] = res;
    if (niter_out) niter_out[s] = it_done;
    if (eigvec)
      HIP_CHECK(hipMemcpyAsync(eigvec + (size_t)n * s, psi.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(conv.get() + (size_t)n * s, psi.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (!converged) {
      status = ADC2G_ENOCONV;
      fail_msg = "Convergence not reached for state: " + std::to_string(s + 1);
      break;
    }
    ++nconv;
  }
  if (stats) {
    stats->nconv = nconv;
    stats->reserved = 0;
    stats->nsigma = clock.count;
    stats->time_sigma = clock.seconds;
    stats->time_total = std::chrono::duration<double>(clk::now() - t_start).count();
  }
  if (status != ADC2G_OK) throw Error(status, fail_msg);
}

}  // namespace adc2g
