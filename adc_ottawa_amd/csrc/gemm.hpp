// MFMA GEMM for gfx950: C(map) = alpha * sum_k A(m,k) B(k,n) + beta * C(map).
//
// One kernel template serves every contraction of the sigma-vector, the
// ph-ph build and the Davidson subspace algebra.  Operands are addressed by
// element strides, so row-/column-major and the index-permuted views of the
// density-fitting factors need no copies:
//     A(m,k) = A[b*sab + m*sam + k*sak],   B(k,n) = B[b*sbb + k*sbk + n*sbn].
// The output goes through a small set of index maps (MapKind) that write the
// permuted layouts the next contraction reads.  Split-K writes fp partials
// that a reduction kernel folds through the same map.
//
// Tiling: NT threads = NT/64 waves in a 2 x (NT/128) grid, block tile BM x BN, K tile BK,
// each wave an (BM/2) x (BN/2) grid of 16x16 MFMA blocks.  Both operand tiles
// live in LDS K-contiguous (row = m or n, BK elements + pad), double
// buffered, so one barrier per K tile.  Within a K tile the K index is
// permuted (lane group g of an MFMA step s takes k = g*KV + s-th slice) so
// that every lane reads its MFMA operands as one contiguous LDS vector.
#pragma once

#include "common.hpp"

namespace adc2g {

enum MapKind : int {
  MAP_STD = 0,    // off = m*scm + n*scn + b*scb
  MAP_DOWNY = 1,  // DF down-contraction output -> Y_vv and Y_oo layouts
  MAP_SIG1V = 2,  // (a, k + o*vec) -> sigma1
  MAP_SIG1O = 3,  // (a*nvec + vec, k) -> sigma1
  MAP_WUPV = 4,   // (c + v*P, k + o*vec) -> W[P][vec][c][k]
  MAP_WUPO = 5,   // (c*nvec + vec, k + o*P) -> W[P][vec][c][k]
  MAP_SIG1T = 6,  // (k + o*vec, a) -> sigma1 (the transpose of MAP_SIG1V)
};

struct MapParams {
  int kind = MAP_STD;
  double alpha = 1.0, beta = 0.0;
  void* C = nullptr;
  void* C2 = nullptr;
  int64_t scm = 0, scn = 0, scb = 0;
  int64_t q[6] = {0, 0, 0, 0, 0, 0};
};

struct GemmArgs {
  int M = 0, N = 0, K = 0, batch = 1, splitk = 1;
  const void* A = nullptr;
  int64_t sam = 0, sak = 0, sab = 0;
  const void* B = nullptr;
  int64_t sbk = 0, sbn = 0, sbb = 0;
  MapParams map;
  int out_prec = 0;         // output element type: 0 double, 1 float, 2 bfloat16
  // set by the launcher: both operands' rows are K-contiguous, 16-byte aligned
  // and K is a whole number of 16-byte chunks, so tiles load with 16-byte
  // vectors (the VEC instantiation)
  int vec = 0;
  int direct = 0;           // set by the launcher: apply the map in the kernel epilogue
  // 1: the batches are terms of one sum (per-batch partials, reduced over
  // batches and splits into a single M x N output through the map)
  int reduce_batch = 0;
  // tile shape: 0 chosen by the launcher, 1 = 128x128, 2 = 64x64, 3 = 128x32,
  // 4 = 128x64, 5 = 64x128, 6 = 64x208 (N = 200-ish; both operands K-contiguous)
  int tile = 0;
  void* partial = nullptr;  // split-K workspace (accumulator type), batch*splitk*M*N
  // host-side only: kernel tag (1 = the DF coupling contractions, compiled as a
  // separately named kernel so profilers report them apart) and optional
  // events recorded around the main kernel launch
  int tag = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
};

// Output offset(s) of element (m, n) of batch b. Returns the number of outputs (1 or 2).
__device__ __forceinline__ int map_offsets(const MapParams& mp, int64_t m, int64_t n, int64_t b,
                                           int64_t& o1, int64_t& o2) {
  switch (mp.kind) {
    case MAP_DOWNY: {
      // m = P, n = ck + n1*vec; q = {o, nvec, n1, naux}
      const int64_t o = mp.q[0], nvec = mp.q[1], n1 = mp.q[2], naux = mp.q[3];
      const int64_t vec = n / n1, ck = n - vec * n1;
      const int64_t c = ck / o, k = ck - c * o;
      o1 = k + o * vec + o * nvec * c + n1 * nvec * m;
      o2 = k + o * m + o * naux * (c * nvec + vec);
      return 2;
    }
    case MAP_SIG1V: {
      const int64_t o = mp.q[0], ldw = mp.q[1];
      const int64_t vec = n / o, k = n - vec * o;
      o1 = m * o + k + ldw * vec;
      return 1;
    }
    case MAP_SIG1T: {
      const int64_t o = mp.q[0], ldw = mp.q[1];
      const int64_t vec = m / o, k = m - vec * o;
      o1 = n * o + k + ldw * vec;
      return 1;
    }
    case MAP_SIG1O: {
      const int64_t o = mp.q[0], ldw = mp.q[1], nvec = mp.q[2];
      const int64_t a = m / nvec, vec = m - a * nvec;
      o1 = a * o + n + ldw * vec;
      return 1;
    }
    case MAP_WUPV: {
      // m = c + v*P, n = k + o*vec; q = {o, v, n1, nvec}
      const int64_t o = mp.q[0], v = mp.q[1], n1 = mp.q[2], nvec = mp.q[3];
      const int64_t P = m / v, c = m - P * v;
      const int64_t vec = n / o, k = n - vec * o;
      o1 = k + o * c + n1 * vec + n1 * nvec * P;
      return 1;
    }
    case MAP_WUPO: {
      // m = c*nvec + vec, n = k + o*P
      const int64_t o = mp.q[0], n1 = mp.q[2], nvec = mp.q[3];
      const int64_t c = m / nvec, vec = m - c * nvec;
      const int64_t P = n / o, k = n - P * o;
      o1 = k + o * c + n1 * vec + n1 * nvec * P;
      return 1;
    }
    default:
      o1 = m * mp.scm + n * mp.scn + b * mp.scb;
      return 1;
  }
}

template <class T>
__device__ __forceinline__ T from_double(double x) { return (T)x; }
template <>
__device__ __forceinline__ __bf16 from_double<__bf16>(double x) { return (__bf16)(float)x; }
template <class T>
__device__ __forceinline__ double to_double(T x) { return (double)x; }
template <>
__device__ __forceinline__ double to_double<__bf16>(__bf16 x) { return (double)(float)x; }

template <class OutT>
__device__ __forceinline__ void map_store(const MapParams& mp, int64_t m, int64_t n, int64_t b, double val) {
  int64_t o1, o2 = 0;
  const int cnt = map_offsets(mp, m, n, b, o1, o2);
  OutT* C = (OutT*)mp.C;
  double r = mp.alpha * val;
  if (mp.beta != 0.0) r += mp.beta * to_double(C[o1]);
  C[o1] = from_double<OutT>(r);
  if (cnt == 2) {
    OutT* C2 = (OutT*)mp.C2;
    double r2 = mp.alpha * val;
    if (mp.beta != 0.0) r2 += mp.beta * to_double(C2[o2]);
    C2[o2] = from_double<OutT>(r2);
  }
}

// The permuting maps are separable: offset = mpart(m) + npart(n) (two outputs
// for MAP_DOWNY), so the direct epilogue decodes each row and each column of
// a tile once, with 32-bit divisions, and stores with additions only.
template <int MK>
__device__ __forceinline__ void map_npart(const MapParams& mp, uint32_t n, int64_t& g1, int64_t& g2) {
  if constexpr (MK == MAP_DOWNY) {
    // n = ck + n1*vec, ck = c*o + k:  o1 = k + o*vec + o*nvec*c (+ n1*nvec*m),
    //                                 o2 = k + o*naux*(c*nvec + vec) (+ o*m)
    const uint32_t o = (uint32_t)mp.q[0], nvec = (uint32_t)mp.q[1], n1 = (uint32_t)mp.q[2];
    const int64_t naux = mp.q[3];
    const uint32_t vec = n / n1, ck = n - vec * n1, c = ck / o, k = ck - c * o;
    g1 = (int64_t)k + (int64_t)o * vec + (int64_t)o * nvec * c;
    g2 = (int64_t)k + (int64_t)o * naux * ((int64_t)c * nvec + vec);
  } else if constexpr (MK == MAP_WUPV) {
    // n = k + o*vec -> k + n1*vec
    const uint32_t o = (uint32_t)mp.q[0];
    const int64_t n1 = mp.q[2];
    const uint32_t vec = n / o, k = n - vec * o;
    g1 = (int64_t)k + n1 * vec;
    g2 = 0;
  } else {
    static_assert(MK == MAP_WUPO, "map kind without a direct epilogue");
    // n = k + o*P -> k + n1*nvec*P
    const uint32_t o = (uint32_t)mp.q[0];
    const int64_t n1 = mp.q[2], nvec = mp.q[3];
    const uint32_t P = n / o, k = n - P * o;
    g1 = (int64_t)k + n1 * nvec * P;
    g2 = 0;
  }
}

template <int MK>
__device__ __forceinline__ void map_mpart(const MapParams& mp, uint32_t m, int64_t& f1, int64_t& f2) {
  if constexpr (MK == MAP_DOWNY) {
    f1 = mp.q[2] * mp.q[1] * (int64_t)m;  // n1*nvec*m
    f2 = mp.q[0] * (int64_t)m;            // o*m
  } else if constexpr (MK == MAP_WUPV) {
    // m = c + v*P -> o*c + n1*nvec*P
    const uint32_t v = (uint32_t)mp.q[1];
    const uint32_t P = m / v, c = m - P * v;
    f1 = mp.q[0] * (int64_t)c + mp.q[2] * mp.q[3] * (int64_t)P;
    f2 = 0;
  } else {
    // m = c*nvec + vec -> o*c + n1*vec
    const uint32_t nvec = (uint32_t)mp.q[3];
    const uint32_t c = m / nvec, vec = m - c * nvec;
    f1 = mp.q[0] * (int64_t)c + mp.q[2] * (int64_t)vec;
    f2 = 0;
  }
}

template <class OutT>
__device__ __forceinline__ void store_scaled(OutT* C, int64_t off, double alpha, double beta, double val) {
  double r = alpha * val;
  if (beta != 0.0) r += beta * to_double(C[off]);
  C[off] = from_double<OutT>(r);
}

// ---- per-type MFMA traits ---------------------------------------------------
// KV: K elements per lane per MFMA; KM: K per MFMA (= 4*KV).
template <class T>
struct Mma;

template <>
struct Mma<double> {
  typedef double Acc;
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double frag;
  static constexpr int KV = 1;
  static __device__ __forceinline__ frag load(const double* p) { return *p; }
  static __device__ __forceinline__ v4 mma(frag a, frag b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // register r of lane l holds D[row][l%16]
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <>
struct Mma<float> {
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float frag;
  static constexpr int KV = 1;
  static __device__ __forceinline__ frag load(const float* p) { return *p; }
  static __device__ __forceinline__ v4 mma(frag a, frag b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <>
struct Mma<__bf16> {
  typedef float Acc;
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  typedef v8 frag;
  static constexpr int KV = 8;
  static __device__ __forceinline__ frag load(const __bf16* p) { return *(const v8*)p; }
  static __device__ __forceinline__ v4 mma(frag a, frag b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// XCD-aware work order: the w-th workgroup dispatched (flattened over the
// grid) runs on XCD w % 8; give each XCD a contiguous range of the work
// sequence (tiles m-fastest, then batches, then K splits) so that tiles
// sharing an operand panel, and all tiles of one batch, share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int w, int T) {
  const int per = T >> 3, rem = T & 7;
  const int x = w & 7, slot = w >> 3;
  return x * per + (x < rem ? x : rem) + slot;
}

template <class T, int BM, int BN, int BK, class OutT, int TAG, bool VEC, int MK, int NT>
__global__ void __launch_bounds__(NT, NT / 128)
gemm_kernel(GemmArgs g) {
  typedef Mma<T> MT;
  typedef typename MT::Acc Acc;
  typedef typename MT::v4 v4;
  constexpr int KV = MT::KV;
  constexpr int KM = 4 * KV;           // K per MFMA
  constexpr int KS = BK / KM;          // MFMA steps per K tile
  constexpr int PAD = 16 / sizeof(T);  // 16 B of padding per LDS row
  constexpr int LDR = BK + PAD;
  constexpr int WGN = NT / 128;        // waves along N (2 x WGN wave grid)
  // 16-column blocks per wave; when they do not divide evenly (BN = 208:
  // 7 + 6) the last wave takes fewer, tn_eff below
  constexpr int NB16 = BN / 16;
  constexpr int TN = (NB16 + WGN - 1) / WGN;
  constexpr bool EVEN = (NB16 % WGN == 0);
  constexpr int WM = BM / 2, WN = TN * 16;
  constexpr int TM = WM / 16;
  static_assert(BN % 16 == 0 && BM % 32 == 0, "tile shape");
  constexpr int EA = BM * BK / NT;     // elements per thread per tile
  // (B: rounded up; when BN * BK is not a multiple of NT, as for 208-column
  // tiles on 512 threads, the K-contiguous load and store skip rows >= BN)
  constexpr int EB = (BN * BK + NT - 1) / NT;
  static_assert((BM * BK) % NT == 0, "A tile must divide over the threads");
  static_assert(TN >= 1 && EA >= 1 && EB >= 1, "tile too small for the thread count");
  static_assert(BK % KM == 0, "BK must be a multiple of the MFMA K");

  __shared__ __attribute__((aligned(16))) T sA[2][BM * LDR];
  __shared__ __attribute__((aligned(16))) T sB[2][BN * LDR];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int tn_eff = EVEN ? TN : min(TN, NB16 - wn * TN);  // wave-uniform

  const int tiles_m = (g.M + BM - 1) / BM;
  const int tiles_n = (g.N + BN - 1) / BN;
  const int T_all = tiles_m * tiles_n;
  const int w = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                         gridDim.x * gridDim.y * gridDim.z);
  const int t = w % T_all;
  const int b = (w / T_all) % gridDim.y;  // batch
  const int ks = w / (T_all * gridDim.y);  // split index
  const int tm = t % tiles_m, tn = t / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;

  // K range of this split, in whole K tiles
  const int ktiles_all = (g.K + BK - 1) / BK;
  const int kt_per = (ktiles_all + g.splitk - 1) / g.splitk;
  const int kt_begin = ks * kt_per;
  const int kt_end = min(ktiles_all, kt_begin + kt_per);

  const T* A = (const T*)g.A + (int64_t)b * g.sab;
  const T* B = (const T*)g.B + (int64_t)b * g.sbb;

  // load mapping: contiguous-K operands read rows along k, contiguous-M/N
  // operands read along m/n; both store K-contiguous rows in LDS.
  const bool a_kc = (g.sak == 1);
  const bool b_kc = (g.sbk == 1);

  T ra[VEC ? 1 : EA], rb[VEC ? 1 : EB];

  // Per-thread fixed coordinates of the load pattern.  K-contiguous: the
  // thread keeps column c and walks rows r0 + e*(256/BK).  MN-contiguous: it
  // keeps row r and walks columns c0 + e*(256/BM).
  constexpr int RSA = NT / BK;    // row step, K-contiguous
  constexpr int CSA = NT / BM;    // column step, M-contiguous
  constexpr int CSB = NT / BN;
  const int kc_c = tid % BK, kc_r = tid / BK;
  const int ma_r = tid % BM, ma_c = tid / BM;
  const int nb_r = tid % BN, nb_c = tid / BN;

  // 16-byte vector loads of K-contiguous operands: each row of a K tile is
  // BK*sizeof(T) = 128 bytes = CPR chunks; thread t takes chunk t % CPR of
  // rows t / CPR + e * RPP.
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  constexpr int EV = 16 / sizeof(T);
  constexpr int CPR = BK / EV;
  constexpr int RPP = NT / CPR;
  constexpr int VA = (BM + RPP - 1) / RPP, VB = (BN + RPP - 1) / RPP;
  static_assert(BK % EV == 0 && NT % CPR == 0, "vector load tiling");
  const int vc = tid % CPR, vr = tid / CPR;
  u32x4 va[VEC ? VA : 1], vb[VEC ? VB : 1];
  auto vload_a = [&](int kt) {
    const int k = kt * BK + vc * EV;
#pragma unroll
    for (int e = 0; e < VA; ++e) {
      const int m = m0 + vr + e * RPP;
      va[e] = (vr + e * RPP < BM && m < g.M && k < g.K) ? *(const u32x4*)(A + (int64_t)m * g.sam + k)
                                                         : (u32x4){0u, 0u, 0u, 0u};
    }
  };
  auto vload_b = [&](int kt) {
    const int k = kt * BK + vc * EV;
#pragma unroll
    for (int e = 0; e < VB; ++e) {
      const int n = n0 + vr + e * RPP;
      vb[e] = (vr + e * RPP < BN && n < g.N && k < g.K) ? *(const u32x4*)(B + (int64_t)n * g.sbn + k)
                                                         : (u32x4){0u, 0u, 0u, 0u};
    }
  };
  auto load_a = [&](int kt) {
    const int k0 = kt * BK;
    if constexpr (VEC) {
      vload_a(kt);
    } else if (a_kc) {
      const int k = k0 + kc_c;
      const T* p = A + (int64_t)(m0 + kc_r) * g.sam + k;
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int m = m0 + kc_r + e * RSA;
        ra[e] = (m < g.M && k < g.K) ? p[(int64_t)(e * RSA) * g.sam] : (T)0.0f;
      }
    } else {
      const int m = m0 + ma_r;
      const T* p = A + (int64_t)m * g.sam + (int64_t)(k0 + ma_c) * g.sak;
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int k = k0 + ma_c + e * CSA;
        ra[e] = (m < g.M && k < g.K) ? p[(int64_t)(e * CSA) * g.sak] : (T)0.0f;
      }
    }
  };
  auto load_b = [&](int kt) {
    const int k0 = kt * BK;
    if constexpr (VEC) {
      vload_b(kt);
    } else if (b_kc) {
      const int k = k0 + kc_c;
      const T* p = B + (int64_t)(n0 + kc_r) * g.sbn + k;
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int n = n0 + kc_r + e * RSA;
        rb[e] = (kc_r + e * RSA < BN && n < g.N && k < g.K) ? p[(int64_t)(e * RSA) * g.sbn] : (T)0.0f;
      }
    } else {
      const int n = n0 + nb_r;
      const T* p = B + (int64_t)n * g.sbn + (int64_t)(k0 + nb_c) * g.sbk;
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int k = k0 + nb_c + e * CSB;
        rb[e] = (n < g.N && k < g.K) ? p[(int64_t)(e * CSB) * g.sbk] : (T)0.0f;
      }
    }
  };
  // LDS column for logical k within the tile: step s, lane group gq, element e
  // of the lane's vector: k = s*KM + gq*KV + e is stored at gq*(KS*KV) + s*KV + e,
  // so a lane's operand vector for each step is contiguous.
  auto kpos = [&](int c) {
    const int s = c / KM, rem = c % KM, gq = rem / KV, e = rem % KV;
    return gq * (KS * KV) + s * KV + e;
  };
  // chunk -> LDS: element j of the chunk is logical k = vc*EV + j
  auto vstore = [&](T* base, const u32x4& v) {
    if constexpr (EV == KV) {
      *(u32x4*)(base + kpos(vc * EV)) = v;  // the chunk is one lane operand vector
    } else {
      const T* el = (const T*)&v;
#pragma unroll
      for (int j = 0; j < EV; ++j) base[kpos(vc * EV + j)] = el[j];
    }
  };
  auto store_a = [&](int buf) {
    if constexpr (VEC) {
#pragma unroll
      for (int e = 0; e < VA; ++e)
        if (vr + e * RPP < BM) vstore(&sA[buf][(vr + e * RPP) * LDR], va[e]);
    } else if (a_kc) {
      const int kp = kpos(kc_c);
#pragma unroll
      for (int e = 0; e < EA; ++e) sA[buf][(kc_r + e * RSA) * LDR + kp] = ra[e];
    } else {
#pragma unroll
      for (int e = 0; e < EA; ++e) sA[buf][ma_r * LDR + kpos(ma_c + e * CSA)] = ra[e];
    }
  };
  auto store_b = [&](int buf) {
    if constexpr (VEC) {
#pragma unroll
      for (int e = 0; e < VB; ++e)
        if (vr + e * RPP < BN) vstore(&sB[buf][(vr + e * RPP) * LDR], vb[e]);
    } else if (b_kc) {
      const int kp = kpos(kc_c);
#pragma unroll
      for (int e = 0; e < EB; ++e)
        if (kc_r + e * RSA < BN) sB[buf][(kc_r + e * RSA) * LDR + kp] = rb[e];
    } else {
#pragma unroll
      for (int e = 0; e < EB; ++e) sB[buf][nb_r * LDR + kpos(nb_c + e * CSB)] = rb[e];
    }
  };

  v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4){0, 0, 0, 0};

  const int gq = lane >> 4, li = lane & 15;

  if (kt_begin < kt_end) {
    load_a(kt_begin);
    load_b(kt_begin);
    store_a(0);
    store_b(0);
    __syncthreads();
    for (int kt = kt_begin; kt < kt_end; ++kt) {
      const int buf = (kt - kt_begin) & 1;
      const bool more = (kt + 1 < kt_end);
      if (more) {
        load_a(kt + 1);
        load_b(kt + 1);
      }
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        typename MT::frag fa[TM], fb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
          fa[i] = MT::load(&sA[buf][(wm * WM + i * 16 + li) * LDR + gq * (KS * KV) + s * KV]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          if (EVEN || j < tn_eff)
            fb[j] = MT::load(&sB[buf][(wn * WN + j * 16 + li) * LDR + gq * (KS * KV) + s * KV]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            if (EVEN || j < tn_eff) acc[i][j] = MT::mma(fa[i], fb[j], acc[i][j]);
      }
      if (more) {
        store_a(buf ^ 1);
        store_b(buf ^ 1);
      }
      __syncthreads();
    }
  }

  // epilogue.  Without split-K the tile goes straight through the map this
  // instantiation was compiled for (MK); split-K writes a row-major partial
  // that the reduction kernel folds through the map.  The loops stay fully
  // unrolled, so every index into acc is a compile-time constant.
  const bool direct = (g.direct != 0);
  Acc* part = (Acc*)g.partial + ((int64_t)b * g.splitk + ks) * (int64_t)g.M * g.N;
  OutT* C = (OutT*)g.map.C;
  const double alpha = g.map.alpha, beta = g.map.beta;
  if constexpr (MK != MAP_STD) {
    if (direct) {
      OutT* C2 = (OutT*)g.map.C2;
      int64_t gn1[TN], gn2[TN];
      bool nv[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * WN + j * 16 + li;
        nv[j] = n < g.N && (EVEN || j < tn_eff);
        map_npart<MK>(g.map, nv[j] ? (uint32_t)n : 0u, gn1[j], gn2[j]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm * WM + i * 16 + MT::row(lane, r);
          if (m >= g.M) continue;
          int64_t f1, f2;
          map_mpart<MK>(g.map, (uint32_t)m, f1, f2);
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if (!nv[j]) continue;
            const double val = (double)acc[i][j][r];
            store_scaled<OutT>(C, f1 + gn1[j], alpha, beta, val);
            if constexpr (MK == MAP_DOWNY) store_scaled<OutT>(C2, f2 + gn2[j], alpha, beta, val);
          }
        }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WM + i * 16 + MT::row(lane, r);
        const int n = n0 + wn * WN + j * 16 + li;
        if (m < g.M && n < g.N && (EVEN || j < tn_eff)) {
          if (MK == MAP_STD && direct) {
            const int64_t off = (int64_t)m * g.map.scm + (int64_t)n * g.map.scn + (int64_t)b * g.map.scb;
            if (g.out_prec == 0) {
              double val = alpha * (double)acc[i][j][r];
              if (beta != 0.0) val += beta * (double)C[off];
              C[off] = (OutT)val;
            } else if (g.out_prec == 1) {
              float* Cf = (float*)g.map.C;
              double val = alpha * (double)acc[i][j][r];
              if (beta != 0.0) val += beta * (double)Cf[off];
              Cf[off] = (float)val;
            } else {
              __bf16* Ch = (__bf16*)g.map.C;
              double val = alpha * (double)acc[i][j][r];
              if (beta != 0.0) val += beta * (double)(float)Ch[off];
              Ch[off] = (__bf16)(float)val;
            }
          } else {
            part[(int64_t)m * g.N + n] = acc[i][j][r];
          }
        }
      }
}

// Sum the split-K partials and write through the map.
template <class Acc, class OutT>
__global__ void splitk_reduce_kernel(GemmArgs g) {
  const int64_t MN = (int64_t)g.M * g.N;
  const int64_t total = MN * g.batch;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = idx / MN, mn = idx - b * MN;
    const Acc* p = (const Acc*)g.partial + b * g.splitk * MN + mn;
    double s = 0.0;
    for (int k = 0; k < g.splitk; ++k) s += (double)p[k * MN];
    map_store<OutT>(g.map, mn / g.N, mn % g.N, b, s);
  }
}

// Split-K reduction for many splits and few outputs: 64 consecutive outputs
// per workgroup (one per lane), the splits dealt round-robin to the 16 waves,
// each wave keeping 4 independent sums (4 loads in flight per lane), and the
// 64 sums added in a fixed order.
template <class Acc, class OutT>
__global__ void __launch_bounds__(1024) splitk_reduce_wide_kernel(GemmArgs g) {
  constexpr int NW = 16, U = 4;
  __shared__ double red[NW][64];
  const int64_t MN = (int64_t)g.M * g.N;
  const int64_t total = MN * g.batch;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t idx = (int64_t)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (idx < total) {
    const int64_t b = idx / MN, mn = idx - b * MN;
    const Acc* p = (const Acc*)g.partial + b * g.splitk * MN + mn;
    double acc[U] = {0.0, 0.0, 0.0, 0.0};
    int k = wv;
    for (; k + (U - 1) * NW < g.splitk; k += U * NW) {
#pragma unroll
      for (int u = 0; u < U; ++u) acc[u] += (double)p[(int64_t)(k + u * NW) * MN];
    }
    // fewer than U splits of this wave remain
    if (k < g.splitk) acc[0] += (double)p[(int64_t)k * MN];
    if (k + NW < g.splitk) acc[1] += (double)p[(int64_t)(k + NW) * MN];
    if (k + 2 * NW < g.splitk) acc[2] += (double)p[(int64_t)(k + 2 * NW) * MN];
    s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
  red[wv][lane] = s;
  __syncthreads();
  if (wv == 0 && idx < total) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += red[w][lane];
    const int64_t b = idx / MN, mn = idx - b * MN;
    map_store<OutT>(g.map, mn / g.N, mn % g.N, b, t);
  }
}

// Host launcher. Chooses tile sizes and split-K; `ws` is a workspace for the
// split-K partials that is grown on demand.
enum class Prec { F64, F32, BF16 };

struct GemmWorkspace {
  DevBuf<char> buf;
  // when set, split-K partials go here (no allocation; too small -> ADC2G_ENOMEM)
  char* ext = nullptr;
  size_t ext_size = 0;
};

void gemm(GemmArgs g, Prec prec, GemmWorkspace& ws, hipStream_t st, int splitk_hint = 0);

}  // namespace adc2g
