// Operator construction (ph-ph block, diagonal, factors) and the sigma-vector.
//
// The reference builds the whole matrix once as a COO list on disk
// (write_fspace_adc2_1, source/adclib/fspacetrial.f90:533-562 ->
// get_offdiag_adc2_save_omp, source/adclib/get_matrix.f90:705-1186) and then
// multiplies with it (hxvec_incore, source/eigen/block_davidson.f90:844-883).
// Here only the dense ph-ph block (n1 x n1) is built; the 1h1p<->2h2p
// coupling is applied on the fly as MFMA contractions of the pair tensor of
// the 2h2p amplitudes, either with the density-fitting factors (DF form, sums
// over naux) or with the (vv|ov) and (oo|ov) blocks assembled once from them
// (INT form, sums over v and o), and the 2h2p-2h2p block (diagonal,
// adc_ph.f90:1090-1096) is fused into the pair-to-CSF projection.
#include "handle.hpp"

#include <algorithm>

namespace adc2g {

Prec prec_from_int(int p) {
  switch (p) {
    case ADC2G_PREC_FP64: return Prec::F64;
    case ADC2G_PREC_FP32: return Prec::F32;
    case ADC2G_PREC_BF16: return Prec::BF16;
    default: throw Error(ADC2G_EARG, "unknown precision " + std::to_string(p));
  }
}

static size_t prec_size(Prec p) { return p == Prec::F64 ? 8 : (p == Prec::F32 ? 4 : 2); }

Handle::~Handle() {
  if (parent) --parent->nviews;
  precon_net_free(*this);
  for (auto& e : ev_pool)
    if (e) (void)hipEventDestroy(e);
  if (st) (void)hipStreamDestroy(st);
}

hipEvent_t Handle::prof_event() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

void Handle::prof_collect(double out[12]) {
  for (int i = 0; i < 12; ++i) out[i] = 0.0;
  HIP_CHECK(hipDeviceSynchronize());  // records may sit on a caller's stream
  for (const auto& r : prof) {
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    out[3 * r.cls] += ms;
    out[3 * r.cls + 1] += 1.0;
    out[3 * r.cls + 2] += r.work;
  }
  prof.clear();
  ev_used = 0;
}

void Handle::create(int o, int v, int naux, const double* eps_h, const double* B_h, const adc2g_opts& opts) {
  ARG_CHECK(o >= 2 && v >= 2 && naux >= 1, "need nocc >= 2, nvirt >= 2, naux >= 1");
  ARG_CHECK(eps_h && B_h, "eps and B must not be NULL");
  ARG_CHECK(opts.coupling >= ADC2G_COUPLING_AUTO && opts.coupling <= ADC2G_COUPLING_INT, "unknown coupling form");
  device = opts.device;
  HIP_CHECK(hipSetDevice(device));
  HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  d.set(o, v, naux);
  dk = dims_k(d);
  second_order = opts.build_phph ? 1 : 0;
  const int nbas = o + v;
  eps.alloc(nbas);
  HIP_CHECK(hipMemcpyAsync(eps.get(), eps_h, nbas * sizeof(double), hipMemcpyHostToDevice, st));
  DevBuf<double> dB;
  dB.alloc((size_t)naux * nbas * nbas);
  HIP_CHECK(hipMemcpyAsync(dB.get(), B_h, dB.n * sizeof(double), hipMemcpyHostToDevice, st));
  Bov.alloc((size_t)naux * d.n1);
  Bvv.alloc((size_t)naux * v * v);
  Boo.alloc((size_t)naux * o * o);
  launch_extract_factors(dB.get(), nbas, dk, Bov.get(), Bvv.get(), Boo.get(), st);
  build_phph(dB.get(), nbas);
  HIP_CHECK(hipStreamSynchronize(st));
  // coupling form: the integral blocks need (v + o)/naux of the flops of the
  // factor contractions, for v*v*n1 + o*o*n1 stored elements (twice: both
  // operand layouts) per precision
  size_t need = integral_bytes(Prec::F64);
  if (opts.keep_lowprec) need += integral_bytes(Prec::F32) + integral_bytes(Prec::BF16);
  const size_t int_budget = (size_t)96 << 30;
  if (opts.coupling == ADC2G_COUPLING_INT)
    form = FORM_INT;
  else if (opts.coupling == ADC2G_COUPLING_DF)
    form = FORM_DF;
  else
    form = ((double)(o + v) < 0.6 * naux && need <= int_budget) ? FORM_INT : FORM_DF;
  ensure_lowprec(Prec::F64);
  if (opts.keep_lowprec) {
    ensure_lowprec(Prec::F32);
    ensure_lowprec(Prec::BF16);
  }
  // default chunk: keep the sigma workspace within ~12 GB
  const size_t budget = (size_t)12 << 30;
  chunk = 64;
  while (chunk > 1 && sigma_workspace_bytes(chunk, Prec::F64) > budget) chunk /= 2;
  resolve_scratch(device);
}

static int prec_index(Prec p) { return p == Prec::F64 ? 0 : (p == Prec::F32 ? 1 : 2); }

size_t Handle::integral_bytes(Prec p) const {
  return prec_size(p) * 2 * ((size_t)d.v * d.v + (size_t)d.o * d.o) * (size_t)d.n1;
}

// FP64 integral blocks of the INT form, assembled from the factors by GEMMs:
// V[(a,c)][bl] = sum_P Bvv[P][a][c] Bov[P][bl], O[(k,m)][bl] = sum_P
// Boo[P][k][m] Bov[P][bl]; Vt and Ot are the same numbers in the layouts the
// up contractions read.
void Handle::build_integrals() {
  IntSet& s = is[0];
  if (s.ready) return;
  const int o = d.o, v = d.v, naux = d.naux;
  const int64_t n1 = d.n1;
  const int64_t nV = (int64_t)v * v * n1, nO = (int64_t)o * o * n1;
  s.V.alloc(sizeof(double) * nV);
  s.Vt.alloc(sizeof(double) * nV);
  s.O.alloc(sizeof(double) * nO);
  s.Ot.alloc(sizeof(double) * nO);
  for (int which = 0; which < 2; ++which) {
    GemmArgs g;
    const int r = which == 0 ? v : o;
    g.M = r * r; g.N = (int)n1; g.K = naux;
    g.A = which == 0 ? (const void*)Bvv.get() : (const void*)Boo.get();
    g.sam = 1; g.sak = (int64_t)r * r;
    g.B = Bov.get(); g.sbk = n1; g.sbn = 1;
    g.map.C = which == 0 ? s.V.get() : s.O.get();
    g.map.scm = n1; g.map.scn = 1;
    gemm(g, Prec::F64, gws, st);
  }
  const double* V = (const double*)s.V.get();
  for (int c = 0; c < v; ++c)
    launch_transpose<double, double>(V + (int64_t)c * v * n1, v, n1, (double*)s.Vt.get() + (int64_t)c * n1 * v, st);
  launch_swap01<double, double>((const double*)s.O.get(), o, o, n1, (double*)s.Ot.get(), st);
  HIP_CHECK(hipStreamSynchronize(st));
  s.ready = true;
}

// Build the coupling operands of precision p (once): the factor set (DF form)
// or the integral blocks (INT form, rounded from the FP64 ones).
void Handle::ensure_lowprec(Prec p) {
  if (form == FORM_INT) {
    build_integrals();
    IntSet& s = is[prec_index(p)];
    if (s.ready) return;
    const IntSet& f = is[0];
    const int o = d.o, v = d.v;
    const int64_t n1 = d.n1;
    const int64_t nV = (int64_t)v * v * n1, nO = (int64_t)o * o * n1;
    const size_t es = prec_size(p);
    s.V.alloc(es * nV);
    s.Vt.alloc(es * nV);
    s.O.alloc(es * nO);
    s.Ot.alloc(es * nO);
    const double* V = (const double*)f.V.get();
    const double* O = (const double*)f.O.get();
    if (p == Prec::F32) {
      launch_convert_f32(V, (float*)s.V.get(), nV, st);
      launch_convert_f32(O, (float*)s.O.get(), nO, st);
      for (int c = 0; c < v; ++c)
        launch_transpose<double, float>(V + (int64_t)c * v * n1, v, n1, (float*)s.Vt.get() + (int64_t)c * n1 * v, st);
      launch_swap01<double, float>(O, o, o, n1, (float*)s.Ot.get(), st);
    } else {
      launch_convert_bf16(V, (__bf16*)s.V.get(), nV, st);
      launch_convert_bf16(O, (__bf16*)s.O.get(), nO, st);
      for (int c = 0; c < v; ++c)
        launch_transpose<double, __bf16>(V + (int64_t)c * v * n1, v, n1, (__bf16*)s.Vt.get() + (int64_t)c * n1 * v,
                                         st);
      launch_swap01<double, __bf16>(O, o, o, n1, (__bf16*)s.Ot.get(), st);
    }
    HIP_CHECK(hipStreamSynchronize(st));
    s.ready = true;
    return;
  }
  FactorSet& f = fs[prec_index(p)];
  if (f.ready) return;
  const int64_t n1 = d.n1, naux = d.naux;
  const int64_t nov = naux * n1, nvv = (int64_t)naux * d.v * d.v, noo = (int64_t)naux * d.o * d.o;
  if (p == Prec::F64) {
    f.BovT.alloc(sizeof(double) * nov);
    launch_transpose<double, double>(Bov.get(), naux, n1, (double*)f.BovT.get(), st);
  } else if (p == Prec::F32) {
    f.Bov.alloc(sizeof(float) * nov);
    f.BovT.alloc(sizeof(float) * nov);
    f.Bvv.alloc(sizeof(float) * nvv);
    f.Boo.alloc(sizeof(float) * noo);
    launch_convert_f32(Bov.get(), (float*)f.Bov.get(), nov, st);
    launch_transpose<double, float>(Bov.get(), naux, n1, (float*)f.BovT.get(), st);
    launch_convert_f32(Bvv.get(), (float*)f.Bvv.get(), nvv, st);
    launch_convert_f32(Boo.get(), (float*)f.Boo.get(), noo, st);
  } else {
    f.Bov.alloc(sizeof(__bf16) * nov);
    f.BovT.alloc(sizeof(__bf16) * nov);
    f.Bvv.alloc(sizeof(__bf16) * nvv);
    f.Boo.alloc(sizeof(__bf16) * noo);
    launch_convert_bf16(Bov.get(), (__bf16*)f.Bov.get(), nov, st);
    launch_transpose<double, __bf16>(Bov.get(), naux, n1, (__bf16*)f.BovT.get(), st);
    launch_convert_bf16(Bvv.get(), (__bf16*)f.Bvv.get(), nvv, st);
    launch_convert_bf16(Boo.get(), (__bf16*)f.Boo.get(), noo, st);
  }
  HIP_CHECK(hipStreamSynchronize(st));
  f.ready = true;
}

const void* Handle::fBov(Prec p) const { return p == Prec::F64 ? (const void*)Bov.get() : fs[prec_index(p)].Bov.get(); }
const void* Handle::fBovT(Prec p) const { return fs[prec_index(p)].BovT.get(); }
const void* Handle::fBvv(Prec p) const { return p == Prec::F64 ? (const void*)Bvv.get() : fs[prec_index(p)].Bvv.get(); }
const void* Handle::fBoo(Prec p) const { return p == Prec::F64 ? (const void*)Boo.get() : fs[prec_index(p)].Boo.get(); }

// ph-ph block from the factors (first order + CA, CB, CC as products with
// separable denominators; the algebra is stated in kernels.hip).
void Handle::build_phph(const double* dB, int nbas) {
  (void)dB;
  (void)nbas;
  const int o = d.o, v = d.v, naux = d.naux;
  const int64_t n1 = d.n1;
  DevBuf<double> Mm, Jm;
  Mm.alloc(n1 * n1);
  Jm.alloc((size_t)v * v * o * o);
  H11.alloc(n1 * n1);
  // M = Bov^T Bov, (ak|ci)
  {
    GemmArgs g;
    g.M = (int)n1; g.N = (int)n1; g.K = naux;
    g.A = Bov.get(); g.sam = 1; g.sak = n1;
    g.B = Bov.get(); g.sbk = n1; g.sbn = 1;
    g.map.C = Mm.get(); g.map.scm = n1; g.map.scn = 1;
    gemm(g, Prec::F64, gws, st);
  }
  // J[(a,a'),(k,k')] = (aa'|kk')
  {
    GemmArgs g;
    g.M = v * v; g.N = o * o; g.K = naux;
    g.A = Bvv.get(); g.sam = 1; g.sak = (int64_t)v * v;
    g.B = Boo.get(); g.sbk = (int64_t)o * o; g.sbn = 1;
    g.map.C = Jm.get(); g.map.scm = (int64_t)o * o; g.map.scn = 1;
    gemm(g, Prec::F64, gws, st);
  }
  DevBuf<double> CA, CB;
  CA.alloc((size_t)v * v);
  CB.alloc((size_t)o * o);
  if (second_order) {
    {
      const int64_t K0 = (int64_t)o * o * v;
      DevBuf<double> X1, X2;
      X1.alloc((size_t)v * 2 * K0);
      X2.alloc((size_t)v * 2 * K0);
      launch_ca_operands(Mm.get(), eps.get(), dk, X1.get(), X2.get(), st);
      GemmArgs g;
      g.M = v; g.N = v; g.K = (int)(2 * K0);
      g.A = X1.get(); g.sam = 2 * K0; g.sak = 1;
      g.B = X2.get(); g.sbk = 1; g.sbn = 2 * K0;
      g.map.C = CA.get(); g.map.scm = v; g.map.scn = 1; g.map.alpha = 0.5;
      gemm(g, Prec::F64, gws, st);
      HIP_CHECK(hipStreamSynchronize(st));
    }
    {
      const int64_t K0 = (int64_t)v * v * o;
      DevBuf<double> X1, X2;
      X1.alloc((size_t)o * 2 * K0);
      X2.alloc((size_t)o * 2 * K0);
      launch_cb_operands(Mm.get(), eps.get(), dk, X1.get(), X2.get(), st);
      GemmArgs g;
      g.M = o; g.N = o; g.K = (int)(2 * K0);
      g.A = X1.get(); g.sam = 2 * K0; g.sak = 1;
      g.B = X2.get(); g.sbk = 1; g.sbn = 2 * K0;
      g.map.C = CB.get(); g.map.scm = o; g.map.scn = 1; g.map.alpha = -0.5;
      gemm(g, Prec::F64, gws, st);
      HIP_CHECK(hipStreamSynchronize(st));
    }
    {
      DevBuf<double> X1, X2;
      X1.alloc((size_t)n1 * 2 * n1);
      X2.alloc((size_t)n1 * 2 * n1);
      launch_cc_operands(Mm.get(), eps.get(), dk, X1.get(), X2.get(), st);
      GemmArgs g;
      g.M = (int)n1; g.N = (int)n1; g.K = (int)(2 * n1);
      g.A = X1.get(); g.sam = 2 * n1; g.sak = 1;
      g.B = X2.get(); g.sbk = 1; g.sbn = 2 * n1;
      g.map.C = H11.get(); g.map.scm = n1; g.map.scn = 1; g.map.alpha = -0.5;
      gemm(g, Prec::F64, gws, st);
      HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  launch_phph_assemble(H11.get(), Mm.get(), Jm.get(), CA.get(), CB.get(), eps.get(), dk, second_order, st);
  hii.alloc(d.ndim);
  launch_diagonal(H11.get(), eps.get(), dk, hii.get(), st);
  HIP_CHECK(hipStreamSynchronize(st));
}

size_t Handle::sigma_workspace_bytes(int nc, Prec p) const {
  const size_t n1 = d.n1, naux = d.naux, es = prec_size(p);
  size_t b = n1 * n1 * nc * es;  // Z / G
  if (form == FORM_INT) {
    b += 2 * n1 * nc * es;                         // x1 in two layouts
    b += (size_t)d.v * d.v * d.o * nc * sizeof(double);  // per-c partials of the down contraction
    return b;
  }
  b += 2 * naux * n1 * nc * es;                  // Yvv, Yoo
  b += naux * n1 * nc * (p == Prec::F64 ? 8 : 4);  // W
  b += naux * n1 * nc * es;                      // W transposed
  b += naux * n1 * nc * sizeof(double);          // GEMM partials
  return b;
}

bool Handle::lowprec_ready(Prec p) const {
  return form == FORM_INT ? is[prec_index(p)].ready : (p == Prec::F64 ? fs[0].ready : fs[prec_index(p)].ready);
}

// Vectors per pass that fit the caller's scratch: the workspace of the pass
// (256-byte aligned regions) plus at least 32 MiB, or the per-vector
// workspace if larger, for the GEMM split-K partials.
int Handle::scratch_chunk(Prec p) const {
  const size_t per = sigma_workspace_bytes(1, p) + 6 * 256;
  const size_t reserve = std::max<size_t>((size_t)32 << 20, per);
  if (scratch_bytes < per + reserve) return 0;
  return (int)std::min<size_t>((size_t)chunk, (scratch_bytes - reserve) / per);
}

// W = H X for nvec columns (device pointers, ld = ldx / ldw), in chunks of at
// most `chunk` vectors.  The 1h1p-1h1p block H11 x1 is always FP64; every
// coupling contraction runs in the GEMM precision p (FP32 accumulation for
// FP32/BF16), the 2h2p diagonal in FP64 inside the pair->CSF projection.
void Handle::sigma(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s) {
  ARG_CHECK(ldx >= d.ndim && ldw >= d.ndim, "leading dimensions must be >= ndim");
  if (nvec <= 0) return;
  if (parent) {
    sigma_view(X, ldx, W, ldw, nvec, p, s);
    return;
  }
  if (!s) s = st;
  if (stored) {  // FP64 at every requested precision
    ARG_CHECK(!scratch, "a stored-matrix handle has no scratch-buffer sigma-vector");
    sigma_stored(X, ldx, W, ldw, nvec, s);
    return;
  }
  const size_t es = prec_size(p);
  const size_t n1 = d.n1, naux = d.naux;
  const size_t wes = (p == Prec::F64) ? 8 : 4;  // DF W accumulator type
  int nc_max = std::min(chunk, nvec);
  if (scratch) {
    // caller-provided scratch: no allocation; the operands must be resident
    ARG_CHECK(lowprec_ready(p), "the operands of this precision are not resident (create the handle with "
                                "keep_lowprec = 1 or run one sigma-vector of this precision first)");
    nc_max = std::min(nvec, scratch_chunk(p));
    ARG_CHECK(nc_max >= 1, "one vector needs " + std::to_string(sigma_workspace_bytes(1, p)) +
                               " bytes of scratch; the scratch holds " + std::to_string(scratch_bytes));
    char* q = scratch;
    auto carve = [&](size_t bytes) {
      char* r = q;
      q += (bytes + 255) / 256 * 256;
      return (void*)r;
    };
    ws_x1c = carve(n1 * nc_max * es);
    ws_zg = carve(n1 * n1 * nc_max * es);
    if (form == FORM_INT) {
      ws_x1t = carve(n1 * nc_max * es);
    } else {
      ws_yvv = carve(naux * n1 * nc_max * es);
      ws_yoo = carve(naux * n1 * nc_max * es);
      ws_wup = carve(naux * n1 * nc_max * wes);
      ws_wupt = carve(naux * n1 * nc_max * es);
    }
    ARG_CHECK(q <= scratch + scratch_bytes, "scratch layout exceeds the scratch");
    gws.ext = q;
    gws.ext_size = scratch_bytes - (size_t)(q - scratch);
  } else {
    ensure_lowprec(p);
    gws.ext = nullptr;
    gws.ext_size = 0;
    X1c.alloc(n1 * nc_max * es);
    ZG.alloc(n1 * n1 * nc_max * es);
    ws_x1c = X1c.get();
    ws_zg = ZG.get();
    if (form == FORM_INT) {
      X1t.alloc(n1 * nc_max * es);
      ws_x1t = X1t.get();
    } else {
      Yvv.alloc(naux * n1 * nc_max * es);
      Yoo.alloc(naux * n1 * nc_max * es);
      Wup.alloc(naux * n1 * nc_max * wes);
      WupT.alloc(naux * n1 * nc_max * es);
      ws_yvv = Yvv.get();
      ws_yoo = Yoo.get();
      ws_wup = Wup.get();
      ws_wupt = WupT.get();
    }
  }
  for (int c0 = 0; c0 < nvec; c0 += nc_max) {
    const int nc = std::min(nc_max, nvec - c0);
    if (form == FORM_INT)
      sigma_int(X + ldx * c0, ldx, W + ldw * c0, ldw, nc, p, s);
    else
      sigma_df(X + ldx * c0, ldx, W + ldw * c0, ldw, nc, p, s);
  }
  if (adcx && part == SIGMA_ALL) adcx_apply(X, ldx, W, ldw, nvec, s);
}

void Handle::sigma_part(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, int which,
                        hipStream_t s) {
  part = parent ? SIGMA_ALL : which;
  try {
    sigma(X, ldx, W, ldw, nvec, p, s);
  } catch (...) {
    part = SIGMA_ALL;
    throw;
  }
  part = SIGMA_ALL;
}

namespace {

template <class T>
void pack_chunk_t(const double* Xc, int64_t ldx, int nc, const DimsK& dk, void* Z, void* X1c, void* X1t,
                  const PairLayout& lay, hipStream_t s) {
  launch_csf_to_pair<T>(Xc, ldx, nc, dk, (T*)Z, lay, s);
  launch_pack_x1<T>(Xc, ldx, nc, dk, (T*)X1c, s);
  if (X1t) launch_pack_x1t<T>(Xc, ldx, nc, dk, (T*)X1t, s);
}

void pack_chunk(Prec p, const double* Xc, int64_t ldx, int nc, const DimsK& dk, void* Z, void* X1c, void* X1t,
                const PairLayout& lay, hipStream_t s) {
  switch (p) {
    case Prec::F64: pack_chunk_t<double>(Xc, ldx, nc, dk, Z, X1c, X1t, lay, s); break;
    case Prec::F32: pack_chunk_t<float>(Xc, ldx, nc, dk, Z, X1c, X1t, lay, s); break;
    case Prec::BF16: pack_chunk_t<__bf16>(Xc, ldx, nc, dk, Z, X1c, X1t, lay, s); break;
  }
}

void pack_x1_only(Prec p, const double* Xc, int64_t ldx, int nc, const DimsK& dk, void* X1c, void* X1t,
                  hipStream_t s) {
  switch (p) {
    case Prec::F64:
      launch_pack_x1<double>(Xc, ldx, nc, dk, (double*)X1c, s);
      launch_pack_x1t<double>(Xc, ldx, nc, dk, (double*)X1t, s);
      break;
    case Prec::F32:
      launch_pack_x1<float>(Xc, ldx, nc, dk, (float*)X1c, s);
      launch_pack_x1t<float>(Xc, ldx, nc, dk, (float*)X1t, s);
      break;
    case Prec::BF16:
      launch_pack_x1<__bf16>(Xc, ldx, nc, dk, (__bf16*)X1c, s);
      launch_pack_x1t<__bf16>(Xc, ldx, nc, dk, (__bf16*)X1t, s);
      break;
  }
}

void unpack_chunk(Prec p, const void* G, const double* Xc, int64_t ldx, double* Wc, int64_t ldw, int nc,
                  const DimsK& dk, const double* eps, const PairLayout& lay, hipStream_t s) {
  switch (p) {
    case Prec::F64: launch_pair_to_csf<double>((const double*)G, Xc, ldx, Wc, ldw, nc, dk, eps, lay, s); break;
    case Prec::F32: launch_pair_to_csf<float>((const float*)G, Xc, ldx, Wc, ldw, nc, dk, eps, lay, s); break;
    case Prec::BF16: launch_pair_to_csf<__bf16>((const __bf16*)G, Xc, ldx, Wc, ldw, nc, dk, eps, lay, s); break;
  }
}

}  // namespace

// INT form, down half of one chunk: W1 += C x2 for the nc pair tensors in ZG
// (virtual-major layout, GEMM precision p), W1 the 1h1p rows (ld ldw):
//   s1(a,k)  += sum_c sum_bl V[a][c][bl] Z[c][vec][k][bl]   (batched over c, summed)
//   s1(a,k)  -= sum_{m,bl} Z[a][vec][m][bl] O[k][m][bl]
void Handle::int_down(const void* Z, int nc, Prec p, double* Wc, int64_t ldw, hipStream_t s) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1;
  const IntSet& I = is[prec_index(p)];
  auto span = [&](int cls, double work) -> std::pair<hipEvent_t, hipEvent_t> {
    if (!prof_on) return {nullptr, nullptr};
    hipEvent_t a = prof_event(), b = prof_event();
    prof.push_back({cls, a, b, work});
    return {a, b};
  };
  {  // sigma1(a, k + o*vec) += sum_{c,bl} Z[c][vec][k][bl] V[a][c][bl], computed as its
     // transpose (rows (vec,k), columns a) so that v = 200 fills a 208-column tile
    GemmArgs g;
    g.M = o * nc; g.N = v; g.K = (int)n1; g.batch = v; g.reduce_batch = 1;
    g.A = Z; g.sam = n1; g.sak = 1; g.sab = (int64_t)nc * o * n1;
    g.B = I.V.get(); g.sbn = (int64_t)v * n1; g.sbk = 1; g.sbb = n1;
    g.map.kind = MAP_SIG1T; g.map.C = Wc; g.map.beta = 1.0;
    g.map.q[0] = o; g.map.q[1] = ldw;
    if (v > 192 && v <= 208) g.tile = 6;
    g.tag = 1;
    auto e = span(0, 2.0 * v * (double)n1 * n1 * nc);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
  {  // sigma1(a*nc + vec, k) -= sum_{m,bl} Z[a][vec][m][bl] O[k][m][bl]
    GemmArgs g;
    g.M = v * nc; g.N = o; g.K = (int)(o * n1);
    g.A = Z; g.sam = (int64_t)o * n1; g.sak = 1;
    g.B = I.O.get(); g.sbn = (int64_t)o * n1; g.sbk = 1;
    g.map.kind = MAP_SIG1O; g.map.C = Wc; g.map.alpha = -1.0; g.map.beta = 1.0;
    g.map.q[0] = o; g.map.q[1] = ldw; g.map.q[2] = nc;
    auto e = span(3, 2.0 * o * (double)n1 * n1 * nc);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
}

// INT form, one chunk of nc vectors.  With the pair tensor stored
// virtual-major, Z[c][vec][k][(b,l)] = Z(c,b,k,l) of vector vec:
//   s1(a,k)   = H11 x1                                              (FP64)
//   s1(a,k)  += sum_c sum_bl V[a][c][bl] Z[c][vec][k][bl]           (down, batched over c, summed)
//   s1(a,k)  -= sum_{m,bl} Z[a][vec][m][bl] O[k][m][bl]             (down, occupied)
//   G         = -sum_m x1(c,m) Ot[m][k][bl]   [c][vec][k][bl]        (up, occupied)
//   G        += sum_a x1(a,k) Vt[c][bl][a]                           (up, batched over c)
//              (BF16: the two up terms in the opposite order)
//   s2        = projection of G onto the CSFs + D x2                 (unpack)
// which is sigma1(a,k) = sum_{P,c} B[P,a,c] Y[P,c,k] - sum_{P,m} B[P,k,m]
// Y[P,a,m] and G = sum_P W[P,c,k] B[P,l,b] of the DF form with the sum over P
// taken first.
void Handle::sigma_int(const double* Xc, int64_t ldx, double* Wc, int64_t ldw, int nc, Prec p, hipStream_t s) {
  const int o = d.o, v = d.v;
  const int64_t n1 = d.n1;
  const int out_lp = (p == Prec::F64) ? 0 : (p == Prec::F32 ? 1 : 2);
  const IntSet& I = is[prec_index(p)];
  const PairLayout lay = pair_layout_virt_major(dk, nc);
  const double esd = (double)prec_size(p);
  auto span = [&](int cls, double work) -> std::pair<hipEvent_t, hipEvent_t> {
    if (!prof_on) return {nullptr, nullptr};
    hipEvent_t a = prof_event(), b = prof_event();
    prof.push_back({cls, a, b, work});
    return {a, b};
  };
  const double flops_v = 2.0 * v * (double)n1 * n1 * nc;  // each virtual contraction
  const double flops_o = 2.0 * o * (double)n1 * n1 * nc;  // each occupied contraction
  const bool down = part != SIGMA_UP, up = part != SIGMA_DOWN;
  if (down) {
    auto ep = span(1, (double)nc * (d.n2 * 8.0 + (double)n1 * n1 * esd));
    if (ep.first) HIP_CHECK(hipEventRecord(ep.first, s));
    pack_chunk(p, Xc, ldx, nc, dk, ws_zg, ws_x1c, ws_x1t, lay, s);
    if (ep.second) HIP_CHECK(hipEventRecord(ep.second, s));
  } else {
    pack_x1_only(p, Xc, ldx, nc, dk, ws_x1c, ws_x1t, s);
  }
  if (down) {  // sigma1 = H11 x1 (FP64)
    GemmArgs g;
    g.M = (int)n1; g.N = nc; g.K = (int)n1;
    g.A = H11.get(); g.sam = n1; g.sak = 1;
    g.B = Xc; g.sbk = 1; g.sbn = ldx;
    g.map.C = Wc; g.map.scm = 1; g.map.scn = ldw;
    gemm(g, Prec::F64, gws, s);
  }
  if (down) int_down(ws_zg, nc, p, Wc, ldw, s);
  // occupied up term: streaming kernel (nocc <= 32) or GEMM.  BF16 writes G
  // with the virtual contraction first and adds this term afterwards (no G
  // read in the GEMM epilogue); FP64/FP32 write this term first and let the
  // GEMM epilogue add to it (measured faster each way at o=16, v=200).
  auto occ_kernel = [&](bool accumulate) -> bool {
    auto e = span(3, flops_o);
    if (e.first) HIP_CHECK(hipEventRecord(e.first, s));
    bool ok = false;
    switch (p) {
      case Prec::F64:
        ok = launch_occ_up<double>(Xc, ldx, nc, dk, (const double*)I.Ot.get(), (double*)ws_zg, accumulate, s);
        break;
      case Prec::F32:
        ok = launch_occ_up<float>(Xc, ldx, nc, dk, (const float*)I.Ot.get(), (float*)ws_zg, accumulate, s);
        break;
      case Prec::BF16:
        ok = launch_occ_up<__bf16>(Xc, ldx, nc, dk, (const __bf16*)I.Ot.get(), (__bf16*)ws_zg, accumulate, s);
        break;
    }
    if (e.second) HIP_CHECK(hipEventRecord(e.second, s));
    if (!ok && prof_on) prof.pop_back();  // nothing was launched between the two events
    return ok;
  };
  if (!up) return;
  const bool occ_last = (p == Prec::BF16) && d.o <= 32;
  if (!occ_last && !occ_kernel(false)) {  // G[(c*nc + vec)][(k,bl)] = -sum_m x1(c,m) Ot[m][k][bl]   (overwrites Z)
    GemmArgs g;
    g.M = v * nc; g.N = (int)(o * n1); g.K = o;
    g.A = ws_x1c; g.sam = o; g.sak = 1;
    g.B = I.Ot.get(); g.sbk = (int64_t)o * n1; g.sbn = 1;
    g.map.C = ws_zg; g.map.scm = (int64_t)o * n1; g.map.scn = 1; g.map.alpha = -1.0;
    g.out_prec = out_lp;
    auto e = span(3, flops_o);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
  {  // G[c][(vec,k)][bl] += sum_a x1t[(vec,k)][a] Vt[c][bl][a]
    GemmArgs g;
    g.tile = 2;  // 64x64: K = v is short and the epilogue reads and writes G; measured
                 // ahead of 64x128 and 128x128 at o=16, v=200, 20 vectors
    g.M = o * nc; g.N = (int)n1; g.K = v; g.batch = v;
    g.A = ws_x1t; g.sam = v; g.sak = 1; g.sab = 0;
    g.B = I.Vt.get(); g.sbn = v; g.sbk = 1; g.sbb = n1 * v;
    g.map.C = ws_zg; g.map.scm = n1; g.map.scn = 1; g.map.scb = (int64_t)nc * o * n1;
    g.map.beta = occ_last ? 0.0 : 1.0;
    g.out_prec = out_lp;
    g.tag = 1;
    auto e = span(0, flops_v);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
  if (occ_last) occ_kernel(true);
  auto eu = span(2, (double)nc * (3.0 * d.n2 * 8.0 + (double)n1 * n1 * esd));
  if (eu.first) HIP_CHECK(hipEventRecord(eu.first, s));
  unpack_chunk(p, ws_zg, Xc, ldx, Wc, ldw, nc, dk, eps.get(), lay, s);
  if (eu.second) HIP_CHECK(hipEventRecord(eu.second, s));
}

// DF form, one chunk of nc vectors (T the GEMM precision p):
//   Z    = pairs(x2)                              [vec][(c,k)][(b,l)]   (pack)
//   s1   = H11 x1                                                      (FP64)
//   Y    = Bov . Z   -> Yvv [(c)][(P)]..., Yoo                         (down)
//   s1  += Bvv . Yvv - Yoo . Boo
//   W    = Bvv . x1 - x1 . Boo                    [P][vec][(c,k)]
//   Wt   = W^T                                    [(vec,c,k)][P]
//   G    = Wt . BovT^T                            [vec][(c,k)][(b,l)]   (up)
//   s2   = projection of G onto the CSFs + D x2                        (unpack)
void Handle::sigma_df(const double* Xc, int64_t ldx, double* Wc, int64_t ldw, int nc, Prec p, hipStream_t s) {
  const int o = d.o, v = d.v, naux = d.naux;
  const int64_t n1 = d.n1;
  const int out_lp = (p == Prec::F64) ? 0 : (p == Prec::F32 ? 1 : 2);
  const int out_w = (p == Prec::F64) ? 0 : 1;
  const PairLayout lay = pair_layout_vec_major(dk);
  const void* BovP = fBov(p);
  const void* BovTP = fBovT(p);
  const void* BvvP = fBvv(p);
  const void* BooP = fBoo(p);
  const double esd = (double)prec_size(p);
  auto span = [&](int cls, double work) -> std::pair<hipEvent_t, hipEvent_t> {
    if (!prof_on) return {nullptr, nullptr};
    hipEvent_t a = prof_event(), b = prof_event();
    prof.push_back({cls, a, b, work});
    return {a, b};
  };
  const double flops_w = 2.0 * (v + o) * (double)naux * n1 * nc;  // each of the sigma1 and W contraction pairs
  auto ep = span(1, (double)nc * (d.n2 * 8.0 + (double)n1 * n1 * esd));
  if (ep.first) HIP_CHECK(hipEventRecord(ep.first, s));
  pack_chunk(p, Xc, ldx, nc, dk, ws_zg, ws_x1c, nullptr, lay, s);
  if (ep.second) HIP_CHECK(hipEventRecord(ep.second, s));
  // sigma1 = H11 x1 (FP64)
  {
    GemmArgs g;
    g.M = (int)n1; g.N = nc; g.K = (int)n1;
    g.A = H11.get(); g.sam = n1; g.sak = 1;
    g.B = Xc; g.sbk = 1; g.sbn = ldx;
    g.map.C = Wc; g.map.scm = 1; g.map.scn = ldw;
    gemm(g, Prec::F64, gws, s);
  }
  // down: Y[P,(c,k)] = sum_{bl} B[P,bl] Z[(c,k),(b,l)]  -> Yvv and Yoo layouts
  {
    GemmArgs g;
    g.M = naux; g.N = (int)(n1 * nc); g.K = (int)n1;
    g.A = BovP; g.sam = n1; g.sak = 1;
    g.B = ws_zg; g.sbk = 1; g.sbn = n1;
    g.map.kind = MAP_DOWNY; g.map.C = ws_yvv; g.map.C2 = ws_yoo;
    g.map.q[0] = o; g.map.q[1] = nc; g.map.q[2] = n1; g.map.q[3] = naux;
    g.out_prec = out_lp;
    g.tag = 1;
    auto e = span(0, 2.0 * naux * (double)n1 * n1 * nc);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
  auto e1 = span(3, 2.0 * flops_w);
  if (e1.first) HIP_CHECK(hipEventRecord(e1.first, s));
  // sigma1(a,k) += sum_{P,c} B[P,a,c] Y[P,c,k]
  {
    GemmArgs g;
    g.M = v; g.N = o * nc; g.K = v * naux;
    g.A = BvvP; g.sam = 1; g.sak = v;
    g.B = ws_yvv; g.sbk = (int64_t)o * nc; g.sbn = 1;
    g.map.kind = MAP_SIG1V; g.map.C = Wc; g.map.beta = 1.0;
    g.map.q[0] = o; g.map.q[1] = ldw;
    gemm(g, p, gws, s);
  }
  // sigma1(a,k) -= sum_{P,m} B[P,k,m] Y[P,a,m]
  {
    GemmArgs g;
    g.M = v * nc; g.N = o; g.K = o * naux;
    g.A = ws_yoo; g.sam = (int64_t)o * naux; g.sak = 1;
    g.B = BooP; g.sbk = o; g.sbn = 1;
    g.map.kind = MAP_SIG1O; g.map.C = Wc; g.map.alpha = -1.0; g.map.beta = 1.0;
    g.map.q[0] = o; g.map.q[1] = ldw; g.map.q[2] = nc;
    gemm(g, p, gws, s);
  }
  // W[P,vec,c,k] = sum_a B[P,c,a] x1(a,k) - sum_m B[P,k,m] x1(c,m)
  {
    GemmArgs g;
    g.M = v * naux; g.N = o * nc; g.K = v;
    g.A = BvvP; g.sam = v; g.sak = 1;
    g.B = ws_x1c; g.sbk = (int64_t)o * nc; g.sbn = 1;
    g.map.kind = MAP_WUPV; g.map.C = ws_wup;
    g.map.q[0] = o; g.map.q[1] = v; g.map.q[2] = n1; g.map.q[3] = nc;
    g.out_prec = out_w;
    gemm(g, p, gws, s);
  }
  {
    GemmArgs g;
    g.M = v * nc; g.N = o * naux; g.K = o;
    g.A = ws_x1c; g.sam = o; g.sak = 1;
    g.B = BooP; g.sbk = 1; g.sbn = o;
    g.map.kind = MAP_WUPO; g.map.C = ws_wup; g.map.alpha = -1.0; g.map.beta = 1.0;
    g.map.q[0] = o; g.map.q[1] = v; g.map.q[2] = n1; g.map.q[3] = nc;
    g.out_prec = out_w;
    gemm(g, p, gws, s);
  }
  // Wt[(vec,c,k)][P] = W[P][(vec,c,k)] in the GEMM precision
  switch (p) {
    case Prec::F64:
      launch_transpose<double, double>((const double*)ws_wup, naux, n1 * nc, (double*)ws_wupt, s);
      break;
    case Prec::F32:
      launch_transpose<float, float>((const float*)ws_wup, naux, n1 * nc, (float*)ws_wupt, s);
      break;
    case Prec::BF16:
      launch_transpose<float, __bf16>((const float*)ws_wup, naux, n1 * nc, (__bf16*)ws_wupt, s);
      break;
  }
  if (e1.second) HIP_CHECK(hipEventRecord(e1.second, s));
  // up: G[(vec,c,k),(b,l)] = sum_P Wt[(vec,c,k),P] BovT[(b,l),P]  (both K-contiguous)
  {
    GemmArgs g;
    g.M = (int)(n1 * nc); g.N = (int)n1; g.K = naux;
    g.A = ws_wupt; g.sam = naux; g.sak = 1;
    g.B = BovTP; g.sbk = 1; g.sbn = naux;
    g.map.C = ws_zg; g.map.scm = n1; g.map.scn = 1;
    g.out_prec = out_lp;
    g.tag = 1;
    auto e = span(0, 2.0 * naux * (double)n1 * n1 * nc);
    g.ev_begin = e.first; g.ev_end = e.second;
    gemm(g, p, gws, s);
  }
  auto eu = span(2, (double)nc * (3.0 * d.n2 * 8.0 + (double)n1 * n1 * esd));
  if (eu.first) HIP_CHECK(hipEventRecord(eu.first, s));
  unpack_chunk(p, ws_zg, Xc, ldx, Wc, ldw, nc, dk, eps.get(), lay, s);
  if (eu.second) HIP_CHECK(hipEventRecord(eu.second, s));
}

}  // namespace adc2g

// ---- sub-space views (CVS)

namespace adc2g {

namespace {

// Xf[f + nf c] = X[inv[f] + ldx c] for the configurations of the view, 0 elsewhere
__global__ void view_scatter_kernel(int64_t nf, int nc, const int64_t* __restrict__ inv, const double* __restrict__ X,
                                    int64_t ldx, double* __restrict__ Xf) {
  const int64_t total = nf * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = t % nf, c = t / nf;
    const int64_t r = inv[f];
    Xf[t] = r >= 0 ? X[r + ldx * c] : 0.0;
  }
}

// W[r + ldw c] = Wf[map[r] + nf c]
__global__ void view_gather_kernel(int64_t nv, int nc, const int64_t* __restrict__ map, const double* __restrict__ Wf,
                                   int64_t nf, double* __restrict__ W, int64_t ldw) {
  const int64_t total = nv * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t % nv, c = t / nv;
    W[r + ldw * c] = Wf[map[r] + nf * c];
  }
}

}  // namespace

// CVS configuration space (select_atom_is_cvs, source/adclib/select_fano.f90:112-142;
// select_atom_d_cvs, :903-1027): the 1h1p (a, i) with i a core orbital (a outer,
// i inner), no t1 or t2, and the t3, t4, t5 configurations with exactly one of
// i < j a core orbital, each in the loop order of the full space.  The CVS
// matrix elements are those of the full matrix (get_offdiag_adc2_save_cvs,
// source/adclib/get_matrix.f90:1323-1726, calls the same element functions), so
// the view is the principal sub-matrix on these configurations, in the order of
// the full space.
// Orbital flags of a view: 1-based occupied indices, each at most once.
static std::vector<char> view_orbitals(int o, int n, const int32_t* orb, const char* what) {
  ARG_CHECK(n >= 1 && n < o && orb, std::string("need 1 <= n < nocc ") + what + " orbitals");
  std::vector<char> is(o, 0);
  for (int c = 0; c < n; ++c) {
    ARG_CHECK(orb[c] >= 1 && orb[c] <= o, std::string(what) + " orbitals are occupied orbitals 1..nocc");
    ARG_CHECK(!is[orb[c] - 1], std::string(what) + " orbital listed twice");
    is[orb[c] - 1] = 1;
  }
  return is;
}

void Handle::create_cvs(Handle& full, int ncore, const int32_t* core) {
  ARG_CHECK(!full.parent, "a view cannot be made from another view");
  const int o = full.d.o, v = full.d.v;
  const std::vector<char> isc = view_orbitals(o, ncore, core, "core");
  const Dims& F = full.d;
  std::vector<int64_t> map;
  map.reserve((size_t)F.ndim / 4);
  // 1h1p: idx = a o + i
  for (int a = 0; a < v; ++a)
    for (int i = 0; i < o; ++i)
      if (isc[i]) map.push_back((int64_t)a * o + i);
  int64_t cnt[5] = {0, 0, 0, 0, 0};
  // t3 (a = b, i < j): (i, j) outer, a inner
  int64_t pij = 0;
  for (int i = 0; i < o; ++i)
    for (int j = i + 1; j < o; ++j, ++pij)
      if (isc[i] + isc[j] == 1)
        for (int a = 0; a < v; ++a, ++cnt[2]) map.push_back(F.off[2] + pij * v + a);
  // t4 then t5 (a < b, i < j): (i, j) outer, (a, b) inner
  for (int t = 3; t <= 4; ++t) {
    pij = 0;
    for (int i = 0; i < o; ++i)
      for (int j = i + 1; j < o; ++j, ++pij)
        if (isc[i] + isc[j] == 1)
          for (int64_t pab = 0; pab < F.npv; ++pab, ++cnt[t]) map.push_back(F.off[t] + pij * F.npv + pab);
  }
  finish_view(full, map, (int64_t)ncore * v, cnt);
}

// Frozen-core configuration space (lifrzcore in select_atom_is, select_atom_d:
// source/adclib/select_fano.f90:27-40, 575, 610, 635-642, 665-672): every
// configuration none of whose holes is a frozen orbital, in the loop order of
// the full space.  Only the configurations are restricted: the matrix
// elements, and the sums over occupied orbitals inside them, are those of the
// full matrix, so the view is the principal sub-matrix on these
// configurations.  (The reference's MP2 ground state with frozen core sums
// over the active occupied orbitals only, source/adclib/mp2.f90:84-90: that is
// the MP2 of the system without the frozen orbitals.)
void Handle::create_frozen_core(Handle& full, int nfrozen, const int32_t* frozen) {
  ARG_CHECK(!full.parent, "a view cannot be made from another view");
  const int o = full.d.o, v = full.d.v;
  const std::vector<char> isf = view_orbitals(o, nfrozen, frozen, "frozen");
  frozen_occ = isf;
  const Dims& F = full.d;
  std::vector<int64_t> map;
  map.reserve((size_t)F.ndim);
  for (int a = 0; a < v; ++a)
    for (int i = 0; i < o; ++i)
      if (!isf[i]) map.push_back((int64_t)a * o + i);
  int64_t cnt[5] = {0, 0, 0, 0, 0};
  // t1 (a = b, i = j): i outer, a inner
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a)
      if (!isf[i]) {
        map.push_back(F.off[0] + (int64_t)i * v + a);
        ++cnt[0];
      }
  // t2 (a < b, i = j): (a, b) outer, i inner
  for (int64_t pab = 0; pab < F.npv; ++pab)
    for (int i = 0; i < o; ++i)
      if (!isf[i]) {
        map.push_back(F.off[1] + pab * o + i);
        ++cnt[1];
      }
  // t3 (a = b, i < j): (i, j) outer, a inner; t4, t5 (a < b, i < j): (i, j) outer, (a, b) inner
  int64_t pij = 0;
  for (int i = 0; i < o; ++i)
    for (int j = i + 1; j < o; ++j, ++pij)
      if (!isf[i] && !isf[j])
        for (int a = 0; a < v; ++a, ++cnt[2]) map.push_back(F.off[2] + pij * v + a);
  for (int t = 3; t <= 4; ++t) {
    pij = 0;
    for (int i = 0; i < o; ++i)
      for (int j = i + 1; j < o; ++j, ++pij)
        if (!isf[i] && !isf[j])
          for (int64_t pab = 0; pab < F.npv; ++pab, ++cnt[t]) map.push_back(F.off[t] + pij * F.npv + pab);
  }
  finish_view(full, map, (int64_t)(o - nfrozen) * v, cnt);
}

// A view on the configurations map[] of the full space (n1 of them 1h1p, then
// cnt[0..4] of t1..t5, each group in the order of the full space).
void Handle::finish_view(Handle& full, const std::vector<int64_t>& map, int64_t n1v, const int64_t cnt[5]) {
  const int o = full.d.o, v = full.d.v;
  const Dims& F = full.d;
  for (int t = 0; t < 5; ++t) vcnt[t] = cnt[t];
  parent = &full;
  ++full.nviews;
  device = full.device;
  form = full.form;
  HIP_CHECK(hipSetDevice(device));
  HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const int64_t nv = (int64_t)map.size();
  d = Dims();
  d.o = o;
  d.v = v;
  d.naux = F.naux;
  d.n1 = n1v;
  d.ndim = nv;
  d.n2 = nv - d.n1;
  for (int t = 0; t < 5; ++t) d.cnt[t] = vcnt[t];
  d.off[0] = d.n1;
  for (int t = 0; t < 5; ++t) d.off[t + 1] = d.off[t] + d.cnt[t];
  std::vector<int64_t> inv(F.ndim, -1);
  for (int64_t r = 0; r < nv; ++r) inv[map[r]] = r;
  vmap.alloc(nv);
  vinv.alloc(F.ndim);
  HIP_CHECK(hipMemcpyAsync(vmap.get(), map.data(), sizeof(int64_t) * nv, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(vinv.get(), inv.data(), sizeof(int64_t) * F.ndim, hipMemcpyHostToDevice, st));
  // the diagonal of the view
  hii.alloc(nv);
  HIP_CHECK(hipStreamSynchronize(full.st));
  hipLaunchKernelGGL(view_gather_kernel, dim3(grid_for(nv)), dim3(256), 0, st, nv, 1, vmap.get(), full.hii.get(),
                     F.ndim, hii.get(), nv);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
}

void Handle::sigma_view(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s) {
  if (!s) s = st;
  const int64_t nf = parent->d.ndim, nv = d.ndim;
  const int cw = std::min(nvec, parent->chunk);
  vX.alloc((size_t)nf * cw);
  vW.alloc((size_t)nf * cw);
  for (int c0 = 0; c0 < nvec; c0 += cw) {
    const int nc = std::min(cw, nvec - c0);
    hipLaunchKernelGGL(view_scatter_kernel, dim3(grid_for(nf * nc)), dim3(256), 0, s, nf, nc, vinv.get(),
                       X + ldx * c0, ldx, vX.get());
    HIP_CHECK(hipGetLastError());
    parent->sigma(vX.get(), nf, vW.get(), nf, nc, p, s);
    hipLaunchKernelGGL(view_gather_kernel, dim3(grid_for(nv * nc)), dim3(256), 0, s, nv, nc, vmap.get(), vW.get(), nf,
                       W + ldw * c0, ldw);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace adc2g
