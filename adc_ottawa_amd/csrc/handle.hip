// Operator construction (ph-ph block, diagonal, factors) and the sigma-vector.
//
// The reference builds the whole matrix once as a COO list on disk
// (write_fspace_adc2_1, source/adclib/fspacetrial.f90:533-562 ->
// get_offdiag_adc2_save_omp, source/adclib/get_matrix.f90:705-1186) and then
// multiplies with it (hxvec_incore, source/eigen/block_davidson.f90:844-883).
// Here only the dense ph-ph block (n1 x n1) is built; the 1h1p<->2h2p
// coupling is applied on the fly as density-fitted contractions, and the
// 2h2p-2h2p block (diagonal, adc_ph.f90:1090-1096) is fused into the
// pair-to-CSF projection.
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

void Handle::prof_collect(double out[9]) {
  for (int i = 0; i < 9; ++i) out[i] = 0.0;
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
  ensure_lowprec(Prec::F64);
  if (opts.keep_lowprec) {
    ensure_lowprec(Prec::F32);
    ensure_lowprec(Prec::BF16);
  }
  // default chunk: keep the sigma workspace within ~12 GB
  const size_t budget = (size_t)12 << 30;
  chunk = 64;
  while (chunk > 1 && sigma_workspace_bytes(chunk, Prec::F64) > budget) chunk /= 2;
}

static int prec_index(Prec p) { return p == Prec::F64 ? 0 : (p == Prec::F32 ? 1 : 2); }

// Build the factor set of precision p (once).
void Handle::ensure_lowprec(Prec p) {
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
  size_t b = n1 * n1 * nc * es;                  // Z / G
  b += 2 * naux * n1 * nc * es;                  // Yvv, Yoo
  b += naux * n1 * nc * (p == Prec::F64 ? 8 : 4);  // W
  b += naux * n1 * nc * es;                      // W transposed
  b += naux * n1 * nc * sizeof(double);          // GEMM partials
  return b;
}

// W = H X for nvec columns (device pointers, ld = ldx / ldw).
//
// Per chunk of nc vectors, with T the GEMM precision p (the 1h1p-1h1p block
// H11 x1 always FP64; every coupling contraction in T with FP32 accumulation
// for FP32/BF16):
//   Z    = pairs(x2)                              [vec][(c,k)][(b,l)]   (pack)
//   s1   = H11 x1                                                      (FP64)
//   Y    = Bov . Z   -> Yvv [(c)][(P)]..., Yoo                         (down)
//   s1  += Bvv . Yvv - Yoo . Boo
//   W    = Bvv . x1 - x1 . Boo                    [P][vec][(c,k)]
//   Wt   = W^T                                    [(vec,c,k)][P]
//   G    = Wt . BovT^T                            [vec][(c,k)][(b,l)]   (up)
//   s2   = projection of G onto the CSFs + D x2                        (unpack)
void Handle::sigma(const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, Prec p, hipStream_t s) {
  ARG_CHECK(ldx >= d.ndim && ldw >= d.ndim, "leading dimensions must be >= ndim");
  if (nvec <= 0) return;
  if (!s) s = st;
  ensure_lowprec(p);
  const int o = d.o, v = d.v, naux = d.naux;
  const int64_t n1 = d.n1;
  const int nc_max = std::min(chunk, nvec);
  const size_t es = prec_size(p);
  const size_t wes = (p == Prec::F64) ? 8 : 4;  // W accumulator type
  const int out_lp = (p == Prec::F64) ? 0 : (p == Prec::F32 ? 1 : 2);
  const int out_w = (p == Prec::F64) ? 0 : 1;
  X1c.alloc(n1 * nc_max * es);
  Yvv.alloc((size_t)naux * n1 * nc_max * es);
  Yoo.alloc((size_t)naux * n1 * nc_max * es);
  Wup.alloc((size_t)naux * n1 * nc_max * wes);
  WupT.alloc((size_t)naux * n1 * nc_max * es);
  ZG.alloc((size_t)n1 * n1 * nc_max * es);
  const void* BovP = fBov(p);
  const void* BovTP = fBovT(p);
  const void* BvvP = fBvv(p);
  const void* BooP = fBoo(p);
  const double esd = (double)es;
  auto span = [&](int cls, double work) -> std::pair<hipEvent_t, hipEvent_t> {
    if (!prof_on) return {nullptr, nullptr};
    hipEvent_t a = prof_event(), b = prof_event();
    prof.push_back({cls, a, b, work});
    return {a, b};
  };
  for (int c0 = 0; c0 < nvec; c0 += nc_max) {
    const int nc = std::min(nc_max, nvec - c0);
    const double* Xc = X + ldx * c0;
    double* Wc = W + ldw * c0;
    auto ep = span(1, (double)nc * (d.n2 * 8.0 + (double)n1 * n1 * esd));
    if (ep.first) HIP_CHECK(hipEventRecord(ep.first, s));
    switch (p) {
      case Prec::F64:
        launch_csf_to_pair<double>(Xc, ldx, nc, dk, (double*)ZG.get(), s);
        launch_pack_x1<double>(Xc, ldx, nc, dk, (double*)X1c.get(), s);
        break;
      case Prec::F32:
        launch_csf_to_pair<float>(Xc, ldx, nc, dk, (float*)ZG.get(), s);
        launch_pack_x1<float>(Xc, ldx, nc, dk, (float*)X1c.get(), s);
        break;
      case Prec::BF16:
        launch_csf_to_pair<__bf16>(Xc, ldx, nc, dk, (__bf16*)ZG.get(), s);
        launch_pack_x1<__bf16>(Xc, ldx, nc, dk, (__bf16*)X1c.get(), s);
        break;
    }
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
      g.B = ZG.get(); g.sbk = 1; g.sbn = n1;
      g.map.kind = MAP_DOWNY; g.map.C = Yvv.get(); g.map.C2 = Yoo.get();
      g.map.q[0] = o; g.map.q[1] = nc; g.map.q[2] = n1; g.map.q[3] = naux;
      g.out_prec = out_lp;
      g.tag = 1;
      auto e = span(0, 2.0 * naux * (double)n1 * n1 * nc);
      g.ev_begin = e.first; g.ev_end = e.second;
      gemm(g, p, gws, s);
    }
    // sigma1(a,k) += sum_{P,c} B[P,a,c] Y[P,c,k]
    {
      GemmArgs g;
      g.M = v; g.N = o * nc; g.K = v * naux;
      g.A = BvvP; g.sam = 1; g.sak = v;
      g.B = Yvv.get(); g.sbk = (int64_t)o * nc; g.sbn = 1;
      g.map.kind = MAP_SIG1V; g.map.C = Wc; g.map.beta = 1.0;
      g.map.q[0] = o; g.map.q[1] = ldw;
      gemm(g, p, gws, s);
    }
    // sigma1(a,k) -= sum_{P,m} B[P,k,m] Y[P,a,m]
    {
      GemmArgs g;
      g.M = v * nc; g.N = o; g.K = o * naux;
      g.A = Yoo.get(); g.sam = (int64_t)o * naux; g.sak = 1;
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
      g.B = X1c.get(); g.sbk = (int64_t)o * nc; g.sbn = 1;
      g.map.kind = MAP_WUPV; g.map.C = Wup.get();
      g.map.q[0] = o; g.map.q[1] = v; g.map.q[2] = n1; g.map.q[3] = nc;
      g.out_prec = out_w;
      gemm(g, p, gws, s);
    }
    {
      GemmArgs g;
      g.M = v * nc; g.N = o * naux; g.K = o;
      g.A = X1c.get(); g.sam = o; g.sak = 1;
      g.B = BooP; g.sbk = 1; g.sbn = o;
      g.map.kind = MAP_WUPO; g.map.C = Wup.get(); g.map.alpha = -1.0; g.map.beta = 1.0;
      g.map.q[0] = o; g.map.q[1] = v; g.map.q[2] = n1; g.map.q[3] = nc;
      g.out_prec = out_w;
      gemm(g, p, gws, s);
    }
    // Wt[(vec,c,k)][P] = W[P][(vec,c,k)] in the GEMM precision
    switch (p) {
      case Prec::F64:
        launch_transpose<double, double>((const double*)Wup.get(), naux, n1 * nc, (double*)WupT.get(), s);
        break;
      case Prec::F32:
        launch_transpose<float, float>((const float*)Wup.get(), naux, n1 * nc, (float*)WupT.get(), s);
        break;
      case Prec::BF16:
        launch_transpose<float, __bf16>((const float*)Wup.get(), naux, n1 * nc, (__bf16*)WupT.get(), s);
        break;
    }
    // up: G[(vec,c,k),(b,l)] = sum_P Wt[(vec,c,k),P] BovT[(b,l),P]  (both K-contiguous)
    {
      GemmArgs g;
      g.M = (int)(n1 * nc); g.N = (int)n1; g.K = naux;
      g.A = WupT.get(); g.sam = naux; g.sak = 1;
      g.B = BovTP; g.sbk = 1; g.sbn = naux;
      g.map.C = ZG.get(); g.map.scm = n1; g.map.scn = 1;
      g.out_prec = out_lp;
      g.tag = 1;
      auto e = span(0, 2.0 * naux * (double)n1 * n1 * nc);
      g.ev_begin = e.first; g.ev_end = e.second;
      gemm(g, p, gws, s);
    }
    auto eu = span(2, (double)nc * (3.0 * d.n2 * 8.0 + (double)n1 * n1 * esd));
    if (eu.first) HIP_CHECK(hipEventRecord(eu.first, s));
    switch (p) {
      case Prec::F64:
        launch_pair_to_csf<double>((const double*)ZG.get(), Xc, ldx, Wc, ldw, nc, dk, eps.get(), s);
        break;
      case Prec::F32:
        launch_pair_to_csf<float>((const float*)ZG.get(), Xc, ldx, Wc, ldw, nc, dk, eps.get(), s);
        break;
      case Prec::BF16:
        launch_pair_to_csf<__bf16>((const __bf16*)ZG.get(), Xc, ldx, Wc, ldw, nc, dk, eps.get(), s);
        break;
    }
    if (eu.second) HIP_CHECK(hipEventRecord(eu.second, s));
  }
}

}  // namespace adc2g
