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

// ---- Extrapolation integrator, device side: the modified midpoint rule and
// the polynomial extrapolation tableau (integrator 2 of the reference,
// source/propagation/bslib.f90) for the relaxation right-hand side
// f(psi) = -H psi.  The reference works in complex arithmetic, but with a
// real start vector and f real every imaginary part stays exactly zero, so
// the vectors here are real.

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

// Extrapolation of column `column` against the earlier ones (Neville-type
// recursion in h^2 with the factors f[q], g[q]; tableau columns 1..K-1 at
// T + (c-1)*n), fused with the error norm: part[blk] = max |dpsi| over the
// block, dpsi the last correction
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

// ---- RKF45, device side (integrator 3 of the reference,
// source/propagation/rkf45rlxlib.f90) for f(y) = -H y: the Fehlberg stage
// combinations with the published integer coefficients, grouped as the
// reference groups them and with contraction off, so the rounding (and with
// it every step decision) is the reference's

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

#undef RL_LOOP

// ---- Extrapolation integrator: step-size and order control
//
// Gragg-Bulirsch-Stoer extrapolation with Deuflhard's order and step-size
// control (P. Deuflhard, Numer. Math. 41 (1983) 399-422; E. Hairer,
// S. P. Norsett, G. Wanner, Solving Ordinary Differential Equations I,
// 2nd ed., Sec. II.9), for the step-number sequence n_j = 2j, polynomial
// extrapolation in h^2 and the absolute max-norm error of the last
// extrapolation correction.  The tolerances, safety factors and decision
// rules are the ones BSStep (source/propagation/bslib.f90:118-366) uses, so
// the steps taken are the reference's:
//  * work of column j: A_1 = n_1 + 1, A_j = A_{j-1} + n_j; convergence
//    factors alpha(i, j) = e^((A_{i+1} - A_{j+1}) / ((2i + 1)(A_{j+1} - A_1 + 1)))
//    with e = tol / 4;
//  * whenever tol or the highest column K changed since the last step, K is
//    lowered to the first j >= 2 with A_{j+1} > A_j alpha(j-1, j) (one more
//    column no longer pays for itself); the table is rebuilt with the new K
//    at the next step, so K settles over the first steps;
//  * every step opens the whole order window (BSStep compares the requested
//    step and time with values it resets at entry, so its restart test always
//    fires); with err_j = (|dy_j|_max / e)^(1/(2j+1)) after column j + 1,
//    the step is accepted once |dy_j|_max < tol; otherwise it is shrunk at
//    column K by 0.7 / err_{K-1}, and at a column j + 1 < K when
//    alpha(j, K-1) < err_j by 0.7 alpha(j, K-1) / err_j, each factor
//    clamped to [1e-5, 0.5], and the step is redone from column 1;
//  * after a step accepted at column k, the next step is h / s, where
//    s = max(err_q, 0.1) for the q < k of least work A_{q+1} s (first one on
//    ties); if that q is k - 1, k < K and the step was not shrunk, one column
//    more is tried: s' = max(s / alpha(k-1, k), 0.1), taken when
//    A_{k+1} s' does not exceed that least work.
class ExtrapolationControl {
 public:
  static constexpr int kMaxColumns = 16;
  static constexpr double kMinStepRel = 1e-10;   // smallest step, relative to the interval
  static constexpr double kTolFraction = 0.25;   // e = tol / 4
  static constexpr double kScaleFloor = 0.1;     // lower bound of err_j in the work estimate
  static constexpr double kOrderSafety = 0.7;
  static constexpr double kShrinkMax = 0.5, kShrinkMin = 1e-5;

  enum class Verdict { kAccept, kShrink, kNextColumn };

  // Rebuild the tables when tol or K changed; returns the (lowered) K.
  int configure(double tol, int K) {
    if (tol == tol_ && K == K_) return K;
    tol_ = tol;
    K_ = K;
    e_ = kTolFraction * tol;
    for (int j = 1; j <= K + 1; ++j) n_[j] = 2 * j;
    A_[1] = n_[1] + 1;
    for (int j = 2; j <= K + 1; ++j) A_[j] = A_[j - 1] + n_[j];
    for (int j = 2; j <= K; ++j)
      for (int i = 1; i < j; ++i)
        alpha_[i][j] = std::pow(e_, (A_[i + 1] - A_[j + 1]) / ((2 * i + 1) * (A_[j + 1] - A_[1] + 1.0)));
    int k = 2;
    for (; k <= K - 1; ++k)
      if (A_[k + 1] > A_[k] * alpha_[k - 1][k]) break;
    return k;
  }

  int substeps(int j) const { return n_[j]; }
  double scaled_error(int j, double max_dy) const { return std::pow(max_dy / e_, 1.0 / (2 * j + 1.0)); }

  // Convergence monitor after column k >= 2 (err_{k-1} in err[k-1]).
  Verdict judge(int k, int K, double max_dy, double tol, const double* err, double* shrink) const {
    if (max_dy < tol) return Verdict::kAccept;
    const int j = k - 1;
    if (k == K) {
      *shrink = kOrderSafety / err[j];
      return Verdict::kShrink;
    }
    if (alpha_[j][K - 1] < err[j]) {
      *shrink = kOrderSafety * alpha_[j][K - 1] / err[j];
      return Verdict::kShrink;
    }
    return Verdict::kNextColumn;
  }

  static double clamp_shrink(double s) { return std::max(std::min(s, kShrinkMax), kShrinkMin); }

  // Step proposed for the next call after acceptance at column k.
  double next_step(double h, int k, int K, bool shrunk, const double* err) const {
    double s = 0.0, least = 1.0e35;
    int best = 0;
    for (int q = 1; q <= k - 1; ++q) {
      const double sq = std::max(err[q], kScaleFloor);
      const double w = sq * A_[q + 1];
      if (w < least) {
        s = sq;
        least = w;
        best = q;
      }
    }
    double next = h / s;
    if (best == k - 1 && k < K && !shrunk) {
      const double s1 = std::max(s / alpha_[k - 1][k], kScaleFloor);
      if (A_[k + 1] * s1 <= least) next = h / s1;
    }
    return next;
  }

 private:
  double tol_ = -1.0, e_ = 0.0;
  int K_ = -1;
  int n_[kMaxColumns + 2] = {};
  double A_[kMaxColumns + 2] = {};
  double alpha_[kMaxColumns + 1][kMaxColumns + 1] = {};
};

// ---- RKF45: Fehlberg's 4(5) pair with the Shampine-Watts step control
//
// L. F. Shampine and H. A. Watts, "Practical solution of ordinary
// differential equations by Runge-Kutta methods", Sandia report SAND76-0585
// (1976); G. E. Forsythe, M. A. Malcolm and C. B. Moler, Computer Methods for
// Mathematical Computations (1977), Sec. 6.5.  The quantities are the
// published ones, as rkf45_rlx (source/propagation/relaxation.f90:501-621)
// drives the driver (source/propagation/rkf45rlxlib.f90): one call per
// output time tout = i*step, relerr = abserr = tol.
//  * Error test: with s the fifth-order solution and
//    ee_k = |-2090 y'_k + 21970 f3_k - 15048 f4_k + 22528 f2_k - 27360 f5_k|,
//    the step passes when |h| max_k ee_k / (|y_k| + |s_k| + abserr/relerr)
//    * (2 / relerr) / 752400 <= 1.
//  * Step factors (safety 0.9, order 1/5): after a failure
//    0.9 / est^(1/5), at least 0.1 (est >= 9^5); after a success
//    0.9 / est^(1/5), at most 5 (est <= (0.9/5)^5), at most 1 right after a
//    failure; never below hmin = 26 eps |t|.
//  * The first step of a state comes from the tolerance and the derivative,
//    component by component: h = |tout - t|, reduced to (tol_k / |y'_k|)^(1/5)
//    wherever tol_k < |y'_k| h^5, and at least 26 eps max(|t|, |tout - t|).
//  * The step is cut to reach tout exactly (half the remaining interval
//    when the step exceeds half of it); an interval below 26 eps |t| is
//    covered by one Euler step.
//  * Limits: 30000 derivative evaluations per state (code 4); a step
//    reduced below hmin (code 6); relerr below 2 eps + 1e-12 (code 3).  The
//    published routine also stops after 100 calls whose step exceeds twice
//    the output interval (code 7); rkf45rlxlib.f90:593 raises that count to
//    10000, which is kept so that the same runs finish.
//  * The derivative y' carries over from one call to the next, also when the
//    caller normalises y in between (as rkf45_rlx does).
class Rkf45Control {
 public:
  static constexpr int kMaxEvals = 30000;
  static constexpr int kMaxOversizeCalls = 10000;
  static constexpr double kErrDen = 752400.0;
  static constexpr double kFailEstCap = 59049.0;        // 9^5: failure factor floor 0.1
  static constexpr double kPassEstFloor = 0.0001889568;  // (0.9/5)^5: growth cap 5

  // Status codes of the published driver (2: reached tout).
  enum { kReached = 2, kRelTolTooSmall = 3, kTooManyEvals = 4, kStepTooSmall = 6, kTooManyOversize = 7,
         kBadInput = 8 };

  void begin_state() {
    have_step_ = false;
    oversize_calls_ = 0;
    nfev_ = 0;
  }
  void count_evals(int k) { nfev_ += k; }
  bool evals_exhausted() const { return nfev_ > kMaxEvals; }
  bool have_step() const { return have_step_; }
  double step() const { return h_; }
  void set_step(double h) { h_ = h; }

  // First step from the host copies of y and y' (component order as published).
  void initial_step(const std::vector<double>& y, const std::vector<double>& yp, double t, double dt,
                    double relerr, double abserr) {
    double h = std::fabs(dt), last_tol = 0.0;
    for (size_t k = 0; k < y.size(); ++k) {
      const double tol = relerr * std::fabs(y[k]) + abserr;
      if (tol > 0.0) {
        last_tol = tol;
        const double ypk = std::fabs(yp[k]);
        if (tol < ypk * std::pow(h, 5)) h = std::pow(tol / ypk, 0.2);
      }
    }
    if (last_tol <= 0.0) h = 0.0;
    h_ = std::max(h, 26.0 * DBL_EPSILON * std::max(std::fabs(t), std::fabs(dt)));
    have_step_ = true;
  }

  // Counts a call whose step exceeds twice the output interval.
  bool oversize_limit(double dt) {
    if (2.0 * std::fabs(dt) <= std::fabs(h_)) ++oversize_calls_;
    if (oversize_calls_ == kMaxOversizeCalls) {
      oversize_calls_ = 0;
      return true;
    }
    return false;
  }

  static double fail_factor(double est) { return est < kFailEstCap ? 0.9 / std::pow(est, 0.2) : 0.1; }
  static double pass_factor(double est, bool failed) {
    const double s = kPassEstFloor < est ? 0.9 / std::pow(est, 0.2) : 5.0;
    return failed ? std::min(s, 1.0) : s;
  }

 private:
  double h_ = 0.0;
  bool have_step_ = false;
  int oversize_calls_ = 0;
  int nfev_ = 0;
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

  // Extrapolation integrator state (bs_rlx: highest column 16, no proposed
  // step yet, time 0) and vectors: the saved state, its derivative -H psi,
  // the midpoint pair and the extrapolation tableau
  ExtrapolationControl bsc;
  int bs_K = ExtrapolationControl::kMaxColumns;
  double bs_next = 0.0, bs_time = 0.0;
  DevBuf<double> bs_saved, bs_w, bs_p1, bs_p2, bs_tab, bs_dt;
  std::vector<double> bs_part(nblk);
  if (bs_mode) {
    bs_saved.alloc(n);
    bs_w.alloc(n);
    bs_dt.alloc(n);
    bs_p1.alloc(n);
    bs_p2.alloc(n);
    bs_tab.alloc((size_t)n * ExtrapolationControl::kMaxColumns);
  }
  // Modified midpoint rule over h with m substeps from the saved state and
  // its derivative -bs_dt; the smoothed end value goes to psi.
  auto bs_midpoint = [&](double h, int m) {
    const double hs = h / m;
    hipLaunchKernelGGL(bs_mm_start_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_saved.get(), bs_dt.get(), -hs,
                       bs_p1.get(), bs_p2.get());
    HIP_CHECK(hipGetLastError());
    sigma1(bs_p2.get(), bs_w.get());
    for (int k = 2; k <= m; ++k) {
      hipLaunchKernelGGL(bs_mm_step_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_w.get(), 2.0 * hs,
                         bs_p1.get(), bs_p2.get());
      HIP_CHECK(hipGetLastError());
      sigma1(bs_p2.get(), bs_w.get());
    }
    hipLaunchKernelGGL(bs_mm_final_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, bs_w.get(), hs, bs_p1.get(),
                       bs_p2.get(), psi.get());
    HIP_CHECK(hipGetLastError());
  };
  // One extrapolation step of at most h inside [0, period] from time t: psi
  // holds the state, bs_w holds H psi.  Returns 0, 1 (highest column out of
  // range) or 2 (step below 1e-10 of the period); the step taken goes to
  // taken and the proposal for the next one to bs_next.
  auto bs_step = [&](double period, double t, double h, double tol, double& taken) -> int {
    using EC = ExtrapolationControl;
    if (bs_K < 2 || bs_K > EC::kMaxColumns) return 1;
    bs_K = bsc.configure(tol, bs_K);
    const int K = bs_K;
    const double hmin = EC::kMinStepRel * period;
    HIP_CHECK(hipMemcpyAsync(bs_saved.get(), psi.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipMemcpyAsync(bs_dt.get(), bs_w.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double err[EC::kMaxColumns + 1] = {}, hsq[EC::kMaxColumns + 1] = {};
    bool shrunk = false;
    int k_acc = 0;
    for (;;) {
      double shrink = 0.0;
      for (int k = 1; k <= K && !k_acc; ++k) {
        if (h < hmin && t + h < period) return 2;
        const int m = bsc.substeps(k);
        hsq[k] = (h / m) * (h / m);
        bs_midpoint(h, m);
        // polynomial extrapolation in h^2: column k against the earlier ones
        BSFactors fac{};
        for (int q = 1; q < k; ++q) {
          const double r = 1.0 / (hsq[k - q] - hsq[k]);
          fac.f[q] = hsq[k] * r;
          fac.g[q] = hsq[k - q] * r;
        }
        hipLaunchKernelGGL(bs_extrapol_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, k, K, fac, psi.get(),
                           bs_tab.get(), part.get());
        HIP_CHECK(hipGetLastError());
        if (k < 2) continue;
        HIP_CHECK(hipMemcpyAsync(bs_part.data(), part.get(), sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        clock.collect();
        double max_dy = 0.0;
        for (double x : bs_part) max_dy = std::max(max_dy, x);
        err[k - 1] = bsc.scaled_error(k - 1, max_dy);
        const auto v = bsc.judge(k, K, max_dy, tol, err, &shrink);
        if (v == EC::Verdict::kAccept) k_acc = k;
        else if (v == EC::Verdict::kShrink) break;
      }
      if (k_acc) break;
      h *= EC::clamp_shrink(shrink);
      shrunk = true;
    }
    bs_next = bsc.next_step(h, k_acc, K, shrunk, err);
    taken = h;
    return 0;
  };

  // RKF45 state and vectors (y is psi; y' persists from call to call)
  Rkf45Control rkc;
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
  // One Fehlberg step of size h from (psi, y'): stages 2..6 at
  // h/4, 3h/8, 12h/13, h, h/2 into rk_f[0..4]; the fifth-order solution then
  // replaces rk_f[0] (the first stage is no longer needed).
  auto rk_fehlberg = [&](double h) {
    rk_stage(1, h / 4.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[0].get());
    rk_stage(2, 3.0 * h / 32.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[1].get());
    rk_stage(3, h / 2197.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[2].get());
    rk_stage(4, h / 4104.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[3].get());
    rk_stage(5, h / 20520.0, rk_x.get());
    rk_f_of(rk_x.get(), rk_f[4].get());
    rk_stage(6, h / 7618050.0, rk_f[0].get());
  };
  // max_k ee_k / (|y_k| + |s_k| + ae) of the step just taken
  auto rk_error_ratio = [&](double ae) -> double {
    hipLaunchKernelGGL(rk_err_kernel, dim3(nblk), dim3(RL_THREADS), 0, st, n, ae, psi.get(), rk_f[0].get(),
                       rk_yp.get(), rk_f[1].get(), rk_f[2].get(), rk_f[3].get(), rk_f[4].get(), part.get());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(rk_part.data(), part.get(), sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    clock.collect();
    double r = 0.0;
    for (double x : rk_part) r = std::max(r, x);
    return r;
  };
  // Advance psi from t to tout with relerr = abserr = tol (first: the first
  // call of a state, which evaluates y' and restarts the counters).  Returns
  // Rkf45Control::kReached or an error code of the published driver.
  auto rk_advance = [&](double& t, double tout, double tol, bool first) -> int {
    using RC = Rkf45Control;
    const double relerr = tol, abserr = tol, tiny = 26.0 * DBL_EPSILON;
    if (relerr < 0.0 || abserr < 0.0) return RC::kBadInput;
    if (!first && t == tout) return RC::kBadInput;
    if (relerr < 2.0 * DBL_EPSILON + 1.0e-12) return RC::kRelTolTooSmall;
    double dt = tout - t;
    if (first) {
      rkc.begin_state();
      rk_f_of(psi.get(), rk_yp.get());
      rkc.count_evals(1);
      if (t == tout) return RC::kReached;
    }
    if (!rkc.have_step()) {
      std::vector<double> hy(n), hyp(n);
      HIP_CHECK(hipMemcpyAsync(hy.data(), psi.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(hyp.data(), rk_yp.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      clock.collect();
      rkc.initial_step(hy, hyp, t, dt, relerr, abserr);
    }
    rkc.set_step(dt >= 0.0 ? std::fabs(rkc.step()) : -std::fabs(rkc.step()));
    if (rkc.oversize_limit(dt)) return RC::kTooManyOversize;
    if (std::fabs(dt) <= tiny * std::fabs(t)) {  // one Euler step covers it
      t = tout;
      rk_stage(1, dt, rk_x.get());
      HIP_CHECK(hipMemcpyAsync(psi.get(), rk_x.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      rk_f_of(psi.get(), rk_yp.get());
      rkc.count_evals(1);
      return RC::kReached;
    }
    const double scale = 2.0 / relerr, ae = scale * abserr;
    for (;;) {
      bool failed = false, last = false;
      const double hmin = tiny * std::fabs(t);
      dt = tout - t;
      double h = rkc.step();
      if (!(2.0 * std::fabs(h) <= std::fabs(dt))) {
        if (std::fabs(dt) <= std::fabs(h)) {
          last = true;
          h = dt;
        } else {
          h = 0.5 * dt;
        }
      }
      double est = 0.0;
      for (;;) {
        if (rkc.evals_exhausted()) return RC::kTooManyEvals;
        rk_fehlberg(h);
        rkc.count_evals(5);
        est = std::fabs(h) * rk_error_ratio(ae) * scale / RC::kErrDen;
        if (est <= 1.0) break;
        failed = true;
        last = false;
        h = RC::fail_factor(est) * h;
        if (std::fabs(h) < hmin) return RC::kStepTooSmall;
      }
      t = t + h;
      HIP_CHECK(hipMemcpyAsync(psi.get(), rk_f[0].get(), sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      rk_f_of(psi.get(), rk_yp.get());
      rkc.count_evals(1);
      const double hh = std::max(RC::pass_factor(est, failed) * std::fabs(h), hmin);
      rkc.set_step(h >= 0.0 ? hh : -hh);
      if (last) {
        t = tout;
        return RC::kReached;
      }
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
      // tfinal = maxiter*step, each covered by extrapolation steps
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
      // siltol, the driver restarted at the start of the state
      double t = 0.0;
      for (int it = 1; it <= p.maxiter; ++it) {
        const int code = rk_advance(t, it * p.step, p.toler, it == 1);
        if (code != Rkf45Control::kReached)
          throw Error(ADC2G_EBREAK, "Error in the RKF45 integrator. Error code: " + std::to_string(code));
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
