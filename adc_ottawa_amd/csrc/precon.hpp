// Learned preconditioner: a small MLP that rescales the DPR correction
// element by element (see mlp.hip for the network and its kernels).
#pragma once

#include "handle.hpp"

namespace adc2g {

// Harvest hook: records (features, DPR step, exact correction) samples of a
// replayed Davidson trajectory.
struct DavTrace;

bool precon_ready(const Handle& H);
// T[:, i] = d_i * (1 + m(features)) for roots with use_dpr[i] == 0 and
// T[:, i] = d_i otherwise, where d is the DPR step (base == NULL) or the
// given base step (base may be T itself: the step is scaled in place).
void precon_learned_step(Handle& H, int64_t n, int nc, const double* theta_d, const double* rnorm_h,
                         const int* use_dpr_h, const double* Res, double* T, int64_t ld,
                         const double* base = nullptr);
// records of one iteration; base: the base step of each root (NULL: DPR)
void davtrace_iteration(Handle& H, DavTrace& tr, int it, int bs_curr, int ns_curr, const double* theta_h,
                        const double* norm_h, const double* X, const double* Res, int64_t n,
                        const double* base = nullptr);

}  // namespace adc2g
