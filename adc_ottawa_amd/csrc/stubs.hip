// Temporary: solver entry points not yet built.
#include "solvers.hpp"

namespace adc2g {

int precon_nparams() { return 0; }
void precon_harvest(Handle&, const adc2g_dav_params&, int64_t, int64_t*) { throw Error(ADC2G_ESTATE, "not built yet"); }
void precon_train(Handle&, const Handle&, int, int, double, uint64_t, double*) { throw Error(ADC2G_ESTATE, "not built yet"); }
void precon_get(const Handle&, float*) { throw Error(ADC2G_ESTATE, "not built yet"); }
void precon_set(Handle&, const float*) { throw Error(ADC2G_ESTATE, "not built yet"); }
void precon_eval(Handle&, int64_t, const float*, const int32_t*, const float*, const float*, float*, float*, double*) {
  throw Error(ADC2G_ESTATE, "not built yet");
}

}  // namespace adc2g

#include "precon.hpp"
namespace adc2g {
bool precon_ready(const Handle&) { return false; }
void precon_learned_step(Handle&, int64_t, int, const double*, const double*, const int*, const double*, double*,
                         int64_t) {
  throw Error(ADC2G_ESTATE, "not built yet");
}
void davtrace_iteration(Handle&, DavTrace&, int, int, int, const double*, const double*, const double*,
                        const double*, int64_t) {}
}  // namespace adc2g
