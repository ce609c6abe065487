// Learned preconditioner: a small MLP that rescales the DPR correction
// element by element (see mlp.hip for the network and its kernels).
#pragma once

#include "handle.hpp"

namespace adc2g {

// Harvest hook: records (features, DPR step, exact correction) samples of a
// replayed Davidson trajectory.
struct DavTrace;

bool precon_ready(const Handle& H);
// T[:, i] = DPR_i * (1 + m(features)) for roots with use_dpr[i] == 0, plain DPR otherwise.
void precon_learned_step(Handle& H, int64_t n, int nc, const double* theta_d, const double* rnorm_h,
                         const int* use_dpr_h, const double* Res, double* T, int64_t ld);
void davtrace_iteration(Handle& H, DavTrace& tr, int it, int bs_curr, int ns_curr, const double* theta_h,
                        const double* norm_h, const double* X, const double* Res, int64_t n);

}  // namespace adc2g
