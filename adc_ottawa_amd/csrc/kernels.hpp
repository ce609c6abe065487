// Streaming (HBM-bound) kernels of the sigma-vector and the ph-ph build.
#pragma once

#include "common.hpp"

namespace adc2g {

// Device copy of the configuration-space sizes passed by value to kernels.
struct DimsK {
  int o, v, naux;
  int64_t n1, npo, npv;
  int64_t off1, off2, off3, off4, off5, ndim;  // starts of t1..t5 and the end
};

inline DimsK dims_k(const Dims& d) {
  DimsK k;
  k.o = d.o; k.v = d.v; k.naux = d.naux;
  k.n1 = d.n1; k.npo = d.npo; k.npv = d.npv;
  k.off1 = d.off[0]; k.off2 = d.off[1]; k.off3 = d.off[2]; k.off4 = d.off[3]; k.off5 = d.off[4];
  k.ndim = d.off[5];
  return k;
}

// Storage of the pair tensors Z / G of a chunk of nc vectors: row (c, k) of
// vector vec starts at c*sc + vec*sv + k*n1 and column (b, l) is b*o + l.
//   vector-major  [vec][c][k][(b,l)]: sc = o*n1,    sv = n1*n1
//   virtual-major [c][vec][k][(b,l)]: sc = nc*o*n1, sv = o*n1
struct PairLayout {
  int64_t sc, sv;
};
inline PairLayout pair_layout_vec_major(const DimsK& d) { return {(int64_t)d.o * d.n1, d.n1 * d.n1}; }
inline PairLayout pair_layout_virt_major(const DimsK& d, int nc) {
  return {(int64_t)nc * d.o * d.n1, (int64_t)d.o * d.n1};
}

// ---- factor extraction: B[P][p][q] (nbas^2 per P) -> Bov[P][a*o+k], Bvv[P][a][b], Boo[P][k][l]
void launch_extract_factors(const double* B, int nbas, const DimsK& d, double* Bov, double* Bvv,
                            double* Boo, hipStream_t st);
void launch_convert_f32(const double* src, float* dst, int64_t n, hipStream_t st);
void launch_convert_bf16(const double* src, __bf16* dst, int64_t n, hipStream_t st);
// dst[c * R + r] = (D) src[r * C + c]  (R x C row-major -> C x R row-major)
template <class S, class D>
void launch_transpose(const S* src, int64_t R, int64_t C, D* dst, hipStream_t st);
// dst[(j*A + i)*C + c] = (D) src[(i*B + j)*C + c]  ([A][B][C] -> [B][A][C])
template <class S, class D>
void launch_swap01(const S* src, int64_t A, int64_t B, int64_t C, D* dst, hipStream_t st);

// ---- sigma-vector streaming kernels
// X1c[a][vec][k] <- X[a*o+k + ldx*vec]
template <class T>
void launch_pack_x1(const double* X, int64_t ldx, int nvec, const DimsK& d, T* X1c, hipStream_t st);
// X1t[(vec,k)][a] <- X[a*o+k + ldx*vec]
template <class T>
void launch_pack_x1t(const double* X, int64_t ldx, int nvec, const DimsK& d, T* X1t, hipStream_t st);
// Occupied term of the INT up contraction, virtual-major pair layout:
// G[(c*nc + vec)][(k,bl)] = -sum_m x1(c,m;vec) Ot[m][(k,bl)], or added to G
// when accumulate.  Returns false (nothing launched) for nocc > 32.
template <class T>
bool launch_occ_up(const double* X, int64_t ldx, int nc, const DimsK& d, const T* Ot, T* G, bool accumulate,
                   hipStream_t st);
// Z[(c,k),(b,l)] (n1 x n1 per vector, stored as `lay`) from the 2h2p amplitudes of X
template <class T>
void launch_csf_to_pair(const double* X, int64_t ldx, int nvec, const DimsK& d, T* Z, const PairLayout& lay,
                        hipStream_t st);
// W[n1 + J] = (projection of G onto CSF J) + D_J * X[n1 + J], for every 2h2p J
template <class T>
void launch_pair_to_csf(const T* G, const double* X, int64_t ldx, double* W, int64_t ldw, int nvec,
                        const DimsK& d, const double* eps, const PairLayout& lay, hipStream_t st);
// Tiled versions of the two maps above (pairs.hip), used for nocc <= 64.
template <class T>
void launch_pack_tiled(const double* X, int64_t ldx, int nvec, const DimsK& d, T* Z, const PairLayout& lay,
                       hipStream_t st);
template <class T>
void launch_unpack_tiled(const T* G, const double* X, int64_t ldx, double* W, int64_t ldw, int nvec, const DimsK& d,
                         const double* eps, const PairLayout& lay, hipStream_t st);
// hii[j] for the whole vector: diag(H11) and the 2h2p orbital-energy sums
void launch_diagonal(const double* H11, const double* eps, const DimsK& d, double* hii, hipStream_t st);

// ---- ph-ph build (second order as separable-denominator products)
// CA operands: X1[a][2K0] = [K/D | K], X2[a][2K0] = [L | L/D], K0 = o*o*v
void launch_ca_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st);
// CB operands: X1[k][2K0] = [P/E | P], X2[k][2K0] = [Q | Q/E], K0 = v*v*o
void launch_cb_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st);
// CC operands: X1[(a,k)][2 n1] = [T/D | T], X2 = [T | T/D]
void launch_cc_operands(const double* Mm, const double* eps, const DimsK& d, double* X1, double* X2,
                        hipStream_t st);
// H11 (holding CC on entry) += 2M - J + delta_kk' CA + delta_aa' CB + delta (e_a - e_k)
void launch_phph_assemble(double* H11, const double* Mm, const double* Jm, const double* CA,
                          const double* CB, const double* eps, const DimsK& d, int second_order,
                          hipStream_t st);


}  // namespace adc2g
