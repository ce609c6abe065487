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
  // coupling factors in each GEMM precision (0 FP64, 1 FP32, 2 BF16): Bov
  // [P][(a,k)], its transpose BovT [(a,k)][P] (K-contiguous operand of the up
  // contraction), Bvv [P][a][b], Boo [P][k][l].  The FP64 set holds only BovT
  // and otherwise uses the buffers above.
  struct FactorSet {
    DevBuf<char> Bov, BovT, Bvv, Boo;
    bool ready = false;
  };
  FactorSet fs[3];
  // coupling form: FORM_DF contracts the amplitudes with the factors (sums
  // over naux), FORM_INT with the integral blocks assembled once from them
  // (sums over v and o):
  //   V [a][c][(b,l)] = (ac|bl),  Vt [c][(b,l)][a] = (ca|bl),
  //   O [k][m][(b,l)] = (km|bl),  Ot [m][k][(b,l)] = (km|bl).
  enum { FORM_DF = 0, FORM_INT = 1 };
  int form = FORM_DF;
  struct IntSet {
    DevBuf<char> V, Vt, O, Ot;
    bool ready = false;
  };
  IntSet is[3];
  void build_integrals();
  size_t integral_bytes(Prec p) const;
  const void* fBov(Prec p) const;
  const void* fBovT(Prec p) const;
  const void* fBvv(Prec p) const;
  const void* fBoo(Prec p) const;

  // sigma workspaces (grown on demand, sized per vector chunk), typed by the
  // chunk's GEMM precision
  DevBuf<char> X1c, X1t, Yvv, Yoo;
  DevBuf<char> Wup;         // W[P][vec][(c,k)] (accumulator type: FP64, or FP32 for FP32/BF16)
  DevBuf<char> WupT;        // W transposed, [(vec,c,k)][P], GEMM precision
  DevBuf<char> ZG;          // Z then G, in the chunk's GEMM precision
  // workspaces of the current sigma pass: the buffers above, or regions of
  // a caller's scratch (sigma with `scratch` set allocates nothing)
  void *ws_x1c = nullptr, *ws_x1t = nullptr, *ws_zg = nullptr, *ws_yvv = nullptr, *ws_yoo = nullptr,
       *ws_wup = nullptr, *ws_wupt = nullptr;
  char* scratch = nullptr;
  size_t scratch_bytes = 0;
  bool lowprec_ready(Prec p) const;
  int scratch_chunk(Prec p) const;
  GemmWorkspace gws;
  DevBuf<double> skinny_ws;  // per-block partials of the tall-skinny products
  int chunk = 32;           // vectors per sigma pass

  // optional per-kernel device timing of sigma-vectors (benchmark): event
  // pairs around the main coupling GEMMs (class 0: the virtual-index
  // contractions, or the naux contractions of the DF form), the CSF->pair
  // packing (1), the pair->CSF projection with the 2h2p diagonal (2) and the
  // remaining coupling GEMMs (3), with the algorithmic flops (classes 0, 3)
  // or bytes (1, 2) each launch is credited with.
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
  void prof_collect(double out[12]);

  // learned preconditioner (network + harvested training records)
  PreconNet* net = nullptr;

  // sub-space view (CVS): this handle's vectors are the configurations
  // vmap[0..d.ndim) of the parent's space, in that order; sigma scatters into
  // the parent's space, applies the parent's sigma-vector and gathers.  d of
  // a view holds only ndim, n1 and the type counts (vcnt).
  Handle* parent = nullptr;
  // frozen-core view: the frozen occupied orbitals (1 = frozen), empty otherwise
  std::vector<char> frozen_occ;
  // stored matrix (stored.hip): the reference's hmlt files in CSR form with
  // both triangles; d.ndim only (no orbital structure), H_jj in hii
  bool stored = false;
  int64_t stored_noffd = 0;
  double stored_per_row = 0.0;
  DevBuf<int64_t> st_rowptr;
  DevBuf<int32_t> st_col;
  DevBuf<double> st_val;
  void create_stored(const char* dia_path, const char* off_path, int dev);
  void sigma_stored(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, hipStream_t s);
  int nviews = 0;                 // views alive on this handle
  DevBuf<int64_t> vmap, vinv;     // view -> parent index, parent -> view index or -1
  DevBuf<double> vX, vW;          // parent-size staging (columns of a chunk)
  int64_t vcnt[5] = {0, 0, 0, 0, 0};
  void create_cvs(Handle& full, int ncore, const int32_t* core);
  void create_frozen_core(Handle& full, int nfrozen, const int32_t* frozen);
  void finish_view(Handle& full, const std::vector<int64_t>& map, int64_t n1v, const int64_t cnt[5]);
  void sigma_view(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s);

  // ADC(2)-x: the first-order 2h2p-2h2p block added to the sigma-vector
  // (FP64, one vector at a time) and its diagonal to hii.  Integral matrices
  // (built on enable): ax_vvvv [(c,b)][(d,e)] = (cd|be), ax_oooo [(m,n)][(k,l)]
  // = (km|ln), ax_ov [(c,k)][(d,m)] = (ck|dm), ax_vo [(c,k)][(d,m)] = (cd|km),
  // ax_a1 = 2 ax_ov - ax_vo; per-vector layouts of T = 2Z - XZ and of the
  // three partial results (adcx.hip).
  bool adcx = false;
  DevBuf<double> ax_vvvv, ax_oooo, ax_ov, ax_vo, ax_a1, ax_d2;
  DevBuf<double> ax_t1, ax_t3, ax_t4, ax_t4n, ax_g1, ax_g2, ax_g3;
  void enable_adcx(bool on);
  void adcx_apply(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, hipStream_t s);

  // Intermediate-state representations of the dipole components for laser
  // propagation (laser.hip): diagonal, off-diagonal part in CSR with both
  // triangles, ground-state element and column.
  struct DipoleIS {
    bool set = false;
    double d00 = 0.0;
    double per_row = 0.0;  // stored off-diagonal elements per row (both triangles)
    DevBuf<double> diag, val, d0j;
    DevBuf<int64_t> rowptr;
    DevBuf<int32_t> col;
  };
  DipoleIS dip[3];
  // the complex absorbing potential W (same layout; d00 = w00, d0j = w0j)
  // and the projector of the CAP term: orthonormal IS vectors (ndim x
  // capproj_n) and whether the ground state is projected out
  DipoleIS cap;
  DipoleIS flux_theta;  // projector onto the absorbing region for the flux (theta00 = d00)
  DevBuf<double> capproj;
  int capproj_n = 0;
  bool capproj_ground = false;
  DevBuf<double> laser_part, laser_u, laser_c, laser_coef, laser_vz;  // work arrays of the laser operator
  // transition density (nto.hip): the n1 x n1 matrix M, the MP2 density and work space
  bool nto_ready = false;
  DevBuf<double> nto_M, nto_rho, nto_p, nto_x;
  void set_dipole_is(int c, const double* diag, int64_t nnz, const int64_t* row, const int64_t* col,
                     const double* val, double d00, const double* d0j);
  void set_cap_is(const double* diag, int64_t nnz, const int64_t* row, const int64_t* col, const double* val,
                  double w00, const double* w0j);
  void set_cap_projector(int nvec, const double* vecs, bool ground);
  void set_flux_theta(const double* diag, int64_t nnz, const int64_t* row, const int64_t* col, const double* val,
                      double theta00, const double* theta0j);

  Handle() = default;
  ~Handle();

  void create(int o, int v, int naux, const double* eps_h, const double* B_h, const adc2g_opts& opts);
  void build_phph(const double* dB, int nbas);
  void ensure_lowprec(Prec p);
  void sigma(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s);
  // Half sigma-vectors for the coupled preconditioner: SIGMA_DOWN writes only
  // the 1h1p rows W1 = H11 x1 + C x2, SIGMA_UP only the 2h2p rows
  // W2 = C^T x1 + D x2 (the other rows of W are left undefined).  The INT form
  // skips the contractions of the other half; the DF form and CVS views
  // compute the whole sigma-vector.
  enum { SIGMA_ALL = 0, SIGMA_DOWN = 1, SIGMA_UP = 2 };
  int part = SIGMA_ALL;
  void sigma_part(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, int which, hipStream_t s);
  void sigma_df(const double* X, int64_t ldx, double* W, int64_t ldw, int nc, Prec p, hipStream_t s);
  void sigma_int(const double* X, int64_t ldx, double* W, int64_t ldw, int nc, Prec p, hipStream_t s);
  // INT form: W1 += C x2 for the nc pair tensors Z (virtual-major layout)
  void int_down(const void* Z, int nc, Prec p, double* W1, int64_t ldw, hipStream_t s);
  size_t sigma_workspace_bytes(int nc, Prec p) const;
};

Prec prec_from_int(int p);
void precon_net_free(Handle& H);
// The no-allocation sigma-vector (noalloc.hip): the device's scratch address
// is resolved once per device (at handle creation); sigma_noalloc runs
// H.sigma with every workspace inside that scratch.
void resolve_scratch(int device);
size_t scratch_capacity();
void sigma_noalloc(Handle& H, int nvec, const double* dX, double* dW, Prec p, hipStream_t s);

}  // namespace adc2g
