// The opaque handle: device-resident ADC(2)-s operator and solver state.
#pragma once

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

struct adc2g_handle;

namespace adc2g {

struct SigmaTimer;
struct PreconNet;

struct Handle {
  Dims d;
  DimsK dk;
  int device = 0;
  hipStream_t st = nullptr;
  int second_order = 1;

  // operator (FP64, always present)
  DevBuf<double> eps, Bov, Bvv, Boo, H11, hii;
  // low-precision copies of the coupling factor (mixed-precision sigma)
  DevBuf<float> Bov32;
  DevBuf<__bf16> Bov16;

  // sigma workspaces (grown on demand, sized per vector chunk)
  DevBuf<double> X1c, Yvv, Yoo, Wup;
  DevBuf<char> ZG;          // Z then G, in the chunk's GEMM precision
  DevBuf<char> Wup_lp;      // Wup in low precision for the up GEMM
  GemmWorkspace gws;
  DevBuf<double> skinny_ws;  // per-block partials of the tall-skinny products
  int chunk = 32;           // vectors per sigma pass

  // optional per-kernel device timing of sigma-vectors (benchmark): event
  // pairs around the coupling GEMMs (class 0), the CSF->pair packing (1) and
  // the pair->CSF projection with the 2h2p diagonal (2), with the algorithmic
  // flops (class 0) or bytes (1, 2) each launch is credited with.
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double work;
  };
  bool prof_on = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<ProfRec> prof;
  hipEvent_t prof_event();
  void prof_collect(double out[9]);

  // learned preconditioner (network + harvested training records)
  PreconNet* net = nullptr;

  Handle() = default;
  ~Handle();

  void create(int o, int v, int naux, const double* eps_h, const double* B_h, const adc2g_opts& opts);
  void build_phph(const double* dB, int nbas);
  void ensure_lowprec(Prec p);
  void sigma(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s);
  size_t sigma_workspace_bytes(int nc, Prec p) const;
};

Prec prec_from_int(int p);
void precon_net_free(Handle& H);

}  // namespace adc2g
