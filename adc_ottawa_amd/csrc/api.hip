// C ABI (include/adc2g.h): argument checks, exception -> error code mapping,
// host <-> device staging for the host-pointer entry points.
#include <cmath>
#include <cstring>
#include <memory>

#include "coupled.hpp"
#include "handle.hpp"
#include "solvers.hpp"

struct adc2g_handle {
  adc2g::Handle h;
};

namespace adc2g {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

template <class F>
static int32_t guarded(F&& f) {
  try {
    f();
    return ADC2G_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return ADC2G_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ADC2G_EHIP;
  }
}

}  // namespace adc2g

using namespace adc2g;

extern "C" {

// the entry points that need the orbital structure of the operator
static void require_orbitals(const adc2g::Handle& h, const char* what) {
  if (h.stored)
    throw adc2g::Error(ADC2G_EARG, std::string(what) + " needs the orbital structure: not available on a stored-matrix "
                                   "handle (adc2g_create_from_hmlt)");
}

int32_t adc2g_create_from_hmlt(const char* dia_path, const char* off_path, int32_t device, adc2g_handle** out) {
  return guarded([&] {
    ARG_CHECK(out != nullptr, "out must not be NULL");
    *out = nullptr;
    auto hp = std::make_unique<adc2g_handle>();
    hp->h.create_stored(dia_path, off_path, device);
    *out = hp.release();
  });
}

int32_t adc2g_create(int32_t nocc, int32_t nvirt, int32_t naux, const double* eps, const double* B,
                     const adc2g_opts* opts, adc2g_handle** out) {
  return guarded([&] {
    ARG_CHECK(out != nullptr, "out must not be NULL");
    *out = nullptr;
    adc2g_opts o{};
    o.device = 0;
    o.build_phph = 1;
    o.keep_lowprec = 0;
    if (opts) o = *opts;
    auto hp = std::make_unique<adc2g_handle>();
    hp->h.create(nocc, nvirt, naux, eps, B, o);
    *out = hp.release();
  });
}

int32_t adc2g_create_cvs(adc2g_handle* full, int32_t ncore, const int32_t* core, adc2g_handle** out) {
  return guarded([&] {
    ARG_CHECK(full && out, "NULL argument");
    require_orbitals(full->h, "a CVS view");
    *out = nullptr;
    HIP_CHECK(hipSetDevice(full->h.device));
    auto hp = std::make_unique<adc2g_handle>();
    hp->h.create_cvs(full->h, ncore, core);
    *out = hp.release();
  });
}

int32_t adc2g_create_frozen_core(adc2g_handle* full, int32_t nfrozen, const int32_t* frozen, adc2g_handle** out) {
  return guarded([&] {
    ARG_CHECK(full && out, "NULL argument");
    require_orbitals(full->h, "a frozen-core view");
    *out = nullptr;
    HIP_CHECK(hipSetDevice(full->h.device));
    auto hp = std::make_unique<adc2g_handle>();
    hp->h.create_frozen_core(full->h, nfrozen, frozen);
    *out = hp.release();
  });
}

int32_t adc2g_set_adc2x(adc2g_handle* h, int32_t enable) {
  return guarded([&] {
    ARG_CHECK(h, "NULL handle");
    require_orbitals(h->h, "ADC(2)-x");
    HIP_CHECK(hipSetDevice(h->h.device));
    h->h.enable_adcx(enable != 0);
  });
}

int32_t adc2g_destroy(adc2g_handle* h) {
  return guarded([&] {
    if (!h) return;
    if (h->h.nviews > 0)
      throw Error(ADC2G_ESTATE, "the operator still has " + std::to_string(h->h.nviews) +
                                    " CVS view(s); destroy them first");
    (void)hipSetDevice(h->h.device);
    if (h->h.st) (void)hipStreamSynchronize(h->h.st);
    delete h;
  });
}

int32_t adc2g_dims(const adc2g_handle* h, int64_t* ndim, int64_t* n1, int64_t counts[5]) {
  return guarded([&] {
    ARG_CHECK(h != nullptr, "NULL handle");
    if (ndim) *ndim = h->h.d.ndim;
    if (n1) *n1 = h->h.d.n1;
    if (counts)
      for (int t = 0; t < 5; ++t) counts[t] = h->h.d.cnt[t];
  });
}

int32_t adc2g_coupling_form(const adc2g_handle* h, int32_t* form) {
  return guarded([&] {
    ARG_CHECK(h && form, "NULL argument");
    *form = h->h.stored ? ADC2G_COUPLING_STORED
                        : (h->h.form == Handle::FORM_INT ? ADC2G_COUPLING_INT : ADC2G_COUPLING_DF);
  });
}

int32_t adc2g_diagonal(adc2g_handle* h, double* hii) {
  return guarded([&] {
    ARG_CHECK(h && hii, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    HIP_CHECK(hipMemcpyAsync(hii, h->h.hii.get(), h->h.d.ndim * sizeof(double), hipMemcpyDeviceToHost, h->h.st));
    HIP_CHECK(hipStreamSynchronize(h->h.st));
  });
}

int32_t adc2g_phph(adc2g_handle* h, double* h11) {
  return guarded([&] {
    ARG_CHECK(h && h11, "NULL argument");
    require_orbitals(h->h, "the ph-ph block");
    ARG_CHECK(!h->h.parent, "the ph-ph block belongs to the full operator, not to a CVS view");
    HIP_CHECK(hipSetDevice(h->h.device));
    const size_t n = (size_t)h->h.d.n1 * h->h.d.n1;
    HIP_CHECK(hipMemcpyAsync(h11, h->h.H11.get(), n * sizeof(double), hipMemcpyDeviceToHost, h->h.st));
    HIP_CHECK(hipStreamSynchronize(h->h.st));
  });
}

int32_t adc2g_coupled_schur(adc2g_handle* h, double sigma, int32_t precision, double* a0) {
  return guarded([&] {
    ARG_CHECK(h && a0, "NULL argument");
    require_orbitals(h->h, "the Schur complement");
    ARG_CHECK(!h->h.parent, "the Schur complement belongs to the full operator, not to a CVS view");
    ARG_CHECK(std::isfinite(sigma), "sigma must be finite");
    HIP_CHECK(hipSetDevice(h->h.device));
    CoupledPrecon cp;
    cp.init(h->h, prec_from_int(precision));
    ARG_CHECK(sigma < cp.dmin, "sigma must lie below every 2h2p diagonal element");
    cp.form_schur(sigma, cp.prec);
    const size_t n = (size_t)h->h.d.n1 * h->h.d.n1;
    HIP_CHECK(hipMemcpyAsync(a0, cp.Qm.get(), n * sizeof(double), hipMemcpyDeviceToHost, h->h.st));
    HIP_CHECK(hipStreamSynchronize(h->h.st));
  });
}

int32_t adc2g_sigma(adc2g_handle* h, int32_t nvec, const double* X, double* W, int32_t precision) {
  return guarded([&] {
    ARG_CHECK(h && X && W, "NULL argument");
    ARG_CHECK(nvec >= 0, "nvec must be >= 0");
    Handle& H = h->h;
    HIP_CHECK(hipSetDevice(H.device));
    const Prec p = prec_from_int(precision);
    if (nvec == 0) return;
    const size_t n = (size_t)H.d.ndim * nvec;
    DevBuf<double> dX, dW;
    dX.alloc(n);
    dW.alloc(n);
    HIP_CHECK(hipMemcpyAsync(dX.get(), X, n * sizeof(double), hipMemcpyHostToDevice, H.st));
    H.sigma(dX.get(), H.d.ndim, dW.get(), H.d.ndim, nvec, p, H.st);
    HIP_CHECK(hipMemcpyAsync(W, dW.get(), n * sizeof(double), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
  });
}

int32_t adc2g_sigma_device(adc2g_handle* h, int32_t nvec, const double* dX, double* dW, int32_t precision,
                           void* stream) {
  return guarded([&] {
    ARG_CHECK(h && dX && dW, "NULL argument");
    ARG_CHECK(nvec >= 0, "nvec must be >= 0");
    Handle& H = h->h;
    HIP_CHECK(hipSetDevice(H.device));
    H.sigma(dX, H.d.ndim, dW, H.d.ndim, nvec, prec_from_int(precision), (hipStream_t)stream);
  });
}

int32_t adc2g_sigma_device_noalloc(adc2g_handle* h, int32_t nvec, const double* dX, double* dW, int32_t precision,
                                   void* stream) {
  return guarded([&] {
    ARG_CHECK(h && dX && dW, "NULL argument");
    require_orbitals(h->h, "the sigma-vector without allocation");
    ARG_CHECK(nvec >= 0, "nvec must be >= 0");
    Handle& H = h->h;
    HIP_CHECK(hipSetDevice(H.device));
    sigma_noalloc(H, nvec, dX, dW, prec_from_int(precision), (hipStream_t)stream);
  });
}

int32_t adc2g_matxvec_treal(adc2g_handle* h, const double* v1, double* v2, int32_t precision) {
  return guarded([&] {
    ARG_CHECK(h && v1 && v2, "NULL argument");
    Handle& H = h->h;
    HIP_CHECK(hipSetDevice(H.device));
    matxvec_treal_host(H, v1, v2, prec_from_int(precision));
  });
}

int32_t adc2g_davidson(adc2g_handle* h, const adc2g_dav_params* p, double* eigval, double* eigvec,
                       adc2g_dav_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && eigval, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    ARG_CHECK(p->guess != ADC2G_GUESS_VECTORS, "guess = ADC2G_GUESS_VECTORS needs adc2g_davidson_vectors");
    davidson_run(h->h, *p, eigval, eigvec, stats, nullptr);
  });
}

int32_t adc2g_davidson_vectors(adc2g_handle* h, const adc2g_dav_params* p, const double* guess, double* eigval,
                               double* eigvec, adc2g_dav_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && guess && eigval, "NULL argument");
    ARG_CHECK(p->blocksize > 0, "blocksize must be given with an initial block");
    HIP_CHECK(hipSetDevice(h->h.device));
    adc2g_dav_params q = *p;
    q.guess = ADC2G_GUESS_VECTORS;
    davidson_run(h->h, q, eigval, eigvec, stats, nullptr, guess);
  });
}

int32_t adc2g_sil_autocorr(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec, double* auto0,
                           int64_t nrec_cap, int64_t* nrec_out, adc2g_sil_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && fvec && auto0, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    sil_run(h->h, *p, fvec, auto0, nrec_cap, nrec_out, stats);
  });
}

int32_t adc2g_sil_fdstates(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec, int32_t nfbas,
                           const double* ebound, int32_t iwfunc, int32_t nsel, const double* fbas2eig,
                           double* vecs, double* energies, adc2g_sil_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && fvec && ebound && fbas2eig && energies, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    sil_fdstates_run(h->h, *p, fvec, nfbas, ebound, iwfunc, nsel, fbas2eig, vecs, energies, stats);
  });
}

int32_t adc2g_set_dipole_is(adc2g_handle* h, int32_t component, const double* diag, int64_t nnz,
                            const int64_t* row, const int64_t* col, const double* val, double d00,
                            const double* d0j) {
  return guarded([&] {
    ARG_CHECK(h, "NULL handle");
    HIP_CHECK(hipSetDevice(h->h.device));
    h->h.set_dipole_is(component, diag, nnz, row, col, val, d00, d0j);
  });
}

int32_t adc2g_set_cap_is(adc2g_handle* h, const double* diag, int64_t nnz, const int64_t* row, const int64_t* col,
                         const double* val, double w00, const double* w0j) {
  return guarded([&] {
    ARG_CHECK(h, "NULL handle");
    HIP_CHECK(hipSetDevice(h->h.device));
    h->h.set_cap_is(diag, nnz, row, col, val, w00, w0j);
  });
}

int32_t adc2g_set_cap_projector(adc2g_handle* h, int32_t nvec, const double* vecs, int32_t ground) {
  return guarded([&] {
    ARG_CHECK(h, "NULL handle");
    h->h.set_cap_projector(nvec, vecs, ground != 0);
  });
}

int32_t adc2g_matxvec_treal_laser(adc2g_handle* h, double time, int32_t npulse, const adc2g_pulse* pulses,
                                  const double* v1, double* v2) {
  return guarded([&] {
    ARG_CHECK(h && v1 && v2, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    matxvec_treal_laser_host(h->h, time, npulse, pulses, v1, v2);
  });
}

int32_t adc2g_sil_laser(adc2g_handle* h, const adc2g_laser_params* p, int32_t npulse, const adc2g_pulse* pulses,
                        double* psi, double* norms, double* flux, double* trdens, int64_t nrec_cap,
                        int64_t* nrec_out, adc2g_sil_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && psi && norms, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    sil_laser_run(h->h, *p, npulse, pulses, psi, norms, flux, trdens, nrec_cap, nrec_out, stats);
  });
}

int32_t adc2g_trdens_vo(adc2g_handle* h, int32_t nvec, const double* x, int64_t ldx, double* trdens) {
  return guarded([&] {
    ARG_CHECK(h && x && trdens, "NULL argument");
    require_orbitals(h->h, "the transition density");
    HIP_CHECK(hipSetDevice(h->h.device));
    trdens_vo_run(h->h, nvec, x, ldx, trdens);
  });
}

int32_t adc2g_set_flux_theta(adc2g_handle* h, const double* diag, int64_t nnz, const int64_t* row,
                             const int64_t* col, const double* val, double theta00, const double* theta0j) {
  return guarded([&] {
    ARG_CHECK(h, "NULL handle");
    HIP_CHECK(hipSetDevice(h->h.device));
    h->h.set_flux_theta(diag, nnz, row, col, val, theta00, theta0j);
  });
}

int32_t adc2g_sil_autocorr_orders(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec, int32_t autoord,
                                  double* auto0, double* auto1, double* auto2, int64_t nrec_cap, int64_t* nrec_out,
                                  adc2g_sil_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && fvec && auto0, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    sil_run(h->h, *p, fvec, auto0, nrec_cap, nrec_out, stats, autoord, auto1, auto2);
  });
}

int32_t adc2g_precon_harvest(adc2g_handle* h, const adc2g_dav_params* p, int64_t max_records, int64_t* added) {
  return guarded([&] {
    ARG_CHECK(h && p, "NULL argument");
    require_orbitals(h->h, "the learned preconditioner");
    ARG_CHECK(!h->h.parent, "the learned preconditioner works on the full operator only");
    HIP_CHECK(hipSetDevice(h->h.device));
    precon_harvest(h->h, *p, max_records, added);
  });
}

int32_t adc2g_precon_train(adc2g_handle* h, const adc2g_handle* data, int32_t epochs, int32_t batch, double lr,
                           uint64_t seed, double* loss_out) {
  return guarded([&] {
    ARG_CHECK(h && data, "NULL argument");
    ARG_CHECK(!h->h.parent && !data->h.parent, "the learned preconditioner works on the full operator only");
    HIP_CHECK(hipSetDevice(h->h.device));
    precon_train(h->h, data->h, epochs, batch, lr, seed, loss_out);
  });
}

int32_t adc2g_precon_absorb(adc2g_handle* dst, const adc2g_handle* src, int64_t* nrec) {
  return guarded([&] {
    ARG_CHECK(dst && src, "NULL argument");
    HIP_CHECK(hipSetDevice(dst->h.device));
    precon_absorb(dst->h, src->h, nrec);
  });
}

int32_t adc2g_precon_nparams(void) { return precon_nparams(); }

int32_t adc2g_precon_get(const adc2g_handle* h, float* params) {
  return guarded([&] {
    ARG_CHECK(h && params, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    precon_get(h->h, params);
  });
}

int32_t adc2g_precon_set(adc2g_handle* h, const float* params) {
  return guarded([&] {
    ARG_CHECK(h && params, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    precon_set(h->h, params);
  });
}

int32_t adc2g_precon_eval(adc2g_handle* h, int64_t n, const float* feat, const int32_t* cls, const float* dn,
                          const float* en, float* m_out, float* grad_out, double* loss_out) {
  return guarded([&] {
    ARG_CHECK(h && feat && cls && dn && en, "NULL argument");
    ARG_CHECK(n > 0, "n must be > 0");
    HIP_CHECK(hipSetDevice(h->h.device));
    precon_eval(h->h, n, feat, cls, dn, en, m_out, grad_out, loss_out);
  });
}

int32_t adc2g_last_error(char* buf, int32_t len) {
  if (!buf || len <= 0) return ADC2G_EARG;
  std::strncpy(buf, g_last_error.c_str(), (size_t)len - 1);
  buf[len - 1] = '\0';
  return ADC2G_OK;
}

int32_t adc2g_cheby_auto(adc2g_handle* h, const double* bounds, const double* q0, int32_t order,
                         int32_t precision, double* auto_out) {
  return guarded([&] {
    ARG_CHECK(h && bounds && q0 && auto_out, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    cheby_run(h->h, bounds, q0, order, prec_from_int(precision), auto_out);
  });
}

int32_t adc2g_cheby_fdstates(adc2g_handle* h, const double* bounds, const double* q0, int32_t order, int32_t nsel,
                             const double* coef, int32_t precision, double* vecs, double* energies) {
  return guarded([&] {
    ARG_CHECK(h && bounds && q0 && coef && energies, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    cheby_fdstates_run(h->h, bounds, q0, order, nsel, coef, prec_from_int(precision), vecs, energies);
  });
}

int32_t adc2g_block_lanczos(adc2g_handle* h, const adc2g_lanczos_params* p, const double* q0,
                            double* eigval, double* tmat, const double* travec, double* tranmom,
                            double* ritz, adc2g_lanczos_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && q0 && eigval, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    lanczos_run(h->h, *p, q0, eigval, tmat, travec, tranmom, ritz, stats);
  });
}

int32_t adc2g_relax(adc2g_handle* h, const adc2g_relax_params* p, const double* guess, double* eigval,
                    double* eigvec, double* residual, int32_t* niter, adc2g_relax_stats* stats) {
  return guarded([&] {
    ARG_CHECK(h && p && eigval, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    relax_run(h->h, *p, guess, eigval, eigvec, residual, niter, stats);
  });
}

int32_t adc2g_write_hmlt(adc2g_handle* h, const char* dia_path, const char* off_path, double minc,
                         int32_t maxbl, int64_t* noffd, int32_t* nbuf) {
  return guarded([&] {
    ARG_CHECK(h != nullptr, "NULL handle");
    HIP_CHECK(hipSetDevice(h->h.device));
    hmlt_write(h->h, dia_path, off_path, minc, maxbl > 0 ? maxbl : 655360, noffd, nbuf);
  });
}

int32_t adc2g_mp2(adc2g_handle* h, double* e_mp2, double* d2, double* rho) {
  return guarded([&] {
    ARG_CHECK(h != nullptr, "NULL handle");
    require_orbitals(h->h, "MP2");
    HIP_CHECK(hipSetDevice(h->h.device));
    mp2_run(h->h, e_mp2, d2, rho);
  });
}

int32_t adc2g_fvector(adc2g_handle* h, const double* dpl, double* f) {
  return guarded([&] {
    ARG_CHECK(h && dpl && f, "NULL argument");
    require_orbitals(h->h, "the F-vector");
    HIP_CHECK(hipSetDevice(h->h.device));
    fvector_run(h->h, dpl, f);
  });
}

int32_t adc2g_lanczos_guess(adc2g_handle* h, int32_t kind, int32_t subdim, int32_t blocksize, double* q0) {
  return guarded([&] {
    ARG_CHECK(h && q0, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    const int64_t n = h->h.d.ndim;
    adc2g::DevBuf<double> g;
    g.alloc((size_t)n * (blocksize > 0 ? blocksize : 1));
    lanczos_guess(h->h, kind, subdim > 0 ? subdim : 800, blocksize, g.get());
    HIP_CHECK(hipMemcpy(q0, g.get(), sizeof(double) * n * blocksize, hipMemcpyDeviceToHost));
  });
}

int32_t adc2g_spectral_bounds(adc2g_handle* h, int32_t method, int32_t ncycles, double* bounds) {
  return guarded([&] {
    ARG_CHECK(h && bounds, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    spectral_bounds_run(h->h, method, ncycles, bounds);
  });
}

int32_t adc2g_power(adc2g_handle* h, int32_t subdim, int32_t maxit, double thrsh, double* eval, double* vec,
                    int32_t* niter) {
  return guarded([&] {
    ARG_CHECK(h && eval, "NULL argument");
    HIP_CHECK(hipSetDevice(h->h.device));
    int it = 0;
    try {
      power_run(h->h, subdim, maxit > 0 ? maxit : 350, thrsh > 0.0 ? thrsh : 1e-8, eval, vec, &it);
    } catch (...) {
      if (niter) *niter = it;
      throw;
    }
    if (niter) *niter = it;
  });
}

int32_t adc2g_sigma_profile(adc2g_handle* h, int32_t enable, double* out) {
  return guarded([&] {
    ARG_CHECK(h != nullptr, "NULL handle");
    Handle& H = h->h;
    HIP_CHECK(hipSetDevice(H.device));
    if (out) H.prof_collect(out);
    H.prof_on = enable != 0;
    if (!H.prof_on) {
      H.prof.clear();
      H.ev_used = 0;
    }
  });
}

}  // extern "C"
