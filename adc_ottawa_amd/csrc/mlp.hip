// Learned preconditioner: network, training (forward/backward/Adam),
// inference inside the Davidson step, and harvesting of training records.
//
// No counterpart in the reference, whose correction steps are dpr/olsen
// (source/eigen/block_davidson.f90:1288-1419); the network rescales the DPR
// step element by element:
//   features (element j, root i): g = asinh(H_jj - theta_i), theta_i,
//            rh = asinh(r_j / |r_i|), class c_j (0 = 1h1p, 1..5 = 2h2p type)
//   h1 = silu(W1 [g, theta, rh] + E[c] + b1)   (64)
//   h2 = silu(W2 h1 + b2)                      (64)
//   m  = w3 . h2 + b3
//   t_j = d_j (1 + m),  d_j = r_j / (theta_i - H_jj)  (the DPR step)
// Training minimises mean (asinh(dn (1 + m)) - asinh(en))^2 over harvested
// records, dn and en being the DPR step and the exact correction x - x*/(x.x*)
// towards the converged eigenvector x*, both divided by the RMS element of
// x - x*/(x.x*); the asinh keeps the few elements with a near-zero DPR
// denominator from dominating the mean.
//
// Weights are kept as fp32 master copies (Adam) and used rounded to bf16.
// Training runs one 64-lane wave per record stream in fp32 with the bf16
// weights; inference in the Davidson step evaluates the 64x64 layer with
// bf16 MFMA (fp32 accumulate) and the first layer's activations rounded to
// bf16 as its operand.
#include <cmath>
#include <vector>

#include "linalg.hpp"
#include "precon.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int HID = 64;
constexpr int NCLS = 6;
constexpr int OFF_W1 = 0;
constexpr int OFF_E = OFF_W1 + HID * 3;
constexpr int OFF_B1 = OFF_E + NCLS * HID;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID;
constexpr int NP = OFF_B3 + 1;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

__device__ __forceinline__ float wave_sum(float x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// ---- training: one wave (64 lanes = hidden units) per block -----------------
__global__ void __launch_bounds__(64) fwdbwd_kernel(const float* __restrict__ P, int64_t n0, int64_t n,
                                                     const float* __restrict__ feat, const int* __restrict__ cls,
                                                     const float* __restrict__ dn, const float* __restrict__ en,
                                                     float inv_n, float* __restrict__ gpart,
                                                     double* __restrict__ lpart, float* __restrict__ mout) {
  __shared__ float sh1[HID], sgz[HID], sgE[NCLS * HID];
  const int j = threadIdx.x;
  float w1[3], w2r[HID], w2c[HID];
  for (int f = 0; f < 3; ++f) w1[f] = P[OFF_W1 + j * 3 + f];
  for (int k = 0; k < HID; ++k) {
    w2r[k] = P[OFF_W2 + j * HID + k];
    w2c[k] = P[OFF_W2 + k * HID + j];
  }
  const float b1 = P[OFF_B1 + j], b2 = P[OFF_B2 + j], w3 = P[OFF_W3 + j], b3 = P[OFF_B3];
  float gw1[3] = {0, 0, 0}, gw2[HID];
  for (int k = 0; k < HID; ++k) gw2[k] = 0.0f;
  for (int c = 0; c < NCLS; ++c) sgE[c * HID + j] = 0.0f;
  float gb1 = 0, gb2 = 0, gw3 = 0, gb3 = 0;
  double loss = 0.0;
  for (int64_t rec = n0 + blockIdx.x; rec < n; rec += gridDim.x) {
    const float x0 = feat[3 * rec], x1 = feat[3 * rec + 1], x2 = feat[3 * rec + 2];
    const int c = cls[rec];
    const float d = dn[rec], e = en[rec];
    const float z1 = w1[0] * x0 + w1[1] * x1 + w1[2] * x2 + P[OFF_E + c * HID + j] + b1;
    const float h1 = silu(z1);
    sh1[j] = h1;
    __syncthreads();
    float z2 = b2;
    for (int k = 0; k < HID; ++k) z2 += w2r[k] * sh1[k];
    const float h2 = silu(z2);
    const float m = wave_sum(w3 * h2) + b3;
    if (mout && j == 0) mout[rec] = m;
    const float t = d * (1.0f + m);
    const float r = asinhf(t) - asinhf(e);
    if (j == 0) loss += (double)r * r;
    const float gm = 2.0f * r * d * inv_n / sqrtf(1.0f + t * t);
    gw3 += gm * h2;
    if (j == 0) gb3 += gm;
    const float gz2 = gm * w3 * dsilu(z2);
    gb2 += gz2;
    for (int k = 0; k < HID; ++k) gw2[k] += gz2 * sh1[k];
    sgz[j] = gz2;
    __syncthreads();
    float gh1 = 0.0f;
    for (int i = 0; i < HID; ++i) gh1 += w2c[i] * sgz[i];
    const float gz1 = gh1 * dsilu(z1);
    gb1 += gz1;
    gw1[0] += gz1 * x0;
    gw1[1] += gz1 * x1;
    gw1[2] += gz1 * x2;
    sgE[c * HID + j] += gz1;
    __syncthreads();
  }
  float* g = gpart + (int64_t)blockIdx.x * NP;
  for (int f = 0; f < 3; ++f) g[OFF_W1 + j * 3 + f] = gw1[f];
  for (int c = 0; c < NCLS; ++c) g[OFF_E + c * HID + j] = sgE[c * HID + j];
  g[OFF_B1 + j] = gb1;
  for (int k = 0; k < HID; ++k) g[OFF_W2 + j * HID + k] = gw2[k];
  g[OFF_B2 + j] = gb2;
  g[OFF_W3 + j] = gw3;
  if (j == 0) {
    g[OFF_B3] = gb3;
    lpart[blockIdx.x] = loss;
  }
}

__global__ void grad_reduce_kernel(int nblk, const float* __restrict__ gpart, float* __restrict__ grad,
                                   const double* __restrict__ lpart, double* __restrict__ loss) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NP; p += gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += gpart[(int64_t)b * NP + p];
    grad[p] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += lpart[b];
    loss[0] = s;
  }
}

__global__ void adam_kernel(float* __restrict__ W, float* __restrict__ Mo, float* __restrict__ Vo,
                            const float* __restrict__ g, float* __restrict__ Pb, __bf16* __restrict__ P16, float lr,
                            float c1, float c2) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NP; p += gridDim.x * blockDim.x) {
    const float gg = g[p];
    const float m = b1 * Mo[p] + (1.0f - b1) * gg;
    const float v = b2 * Vo[p] + (1.0f - b2) * gg * gg;
    Mo[p] = m;
    Vo[p] = v;
    const float w = W[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    W[p] = w;
    Pb[p] = bf16r(w);
    P16[p] = (__bf16)w;
  }
}

__global__ void round_params_kernel(const float* __restrict__ W, float* __restrict__ Pb, __bf16* __restrict__ P16) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NP; p += gridDim.x * blockDim.x) {
    Pb[p] = bf16r(W[p]);
    P16[p] = (__bf16)W[p];
  }
}

// ---- inference inside the Davidson step (bf16 MFMA for the 64x64 layer) -----
struct InferArgs {
  int64_t n;
  int nc;
  int64_t ld;
  int64_t off[6];  // class boundaries: 1h1p | t1 | t2 | t3 | t4 | t5 | end
  const double* theta;
  const double* rnorm;
  const int* use_dpr;
  const double* hii;
  const double* Res;
  double* T;
  const double* base;   // base step (ld as Res), or NULL for the DPR step
  const float* P;       // bf16-rounded weights as fp32
  const __bf16* P16;    // bf16 weights
  int* nonfinite;       // count of elements that fell back to DPR
};

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int class_of(const InferArgs& a, int64_t j) {
  if (j < a.off[0]) return 0;
  if (j < a.off[1]) return 1;
  if (j < a.off[2]) return 2;
  if (j < a.off[3]) return 3;
  if (j < a.off[4]) return 4;
  return 5;
}

__device__ __forceinline__ double guard_den2(double den) {
  const double DEN_MIN = 1e-10;
  if (fabs(den) < DEN_MIN) return den < 0.0 ? -DEN_MIN : DEN_MIN;
  return den;
}

__global__ void __launch_bounds__(256) infer_kernel(InferArgs a) {
  __shared__ float sw[OFF_W2];  // W1, E, b1
  __shared__ float sb2[HID], sw3[HID];
  __shared__ float sb3;
  for (int i = threadIdx.x; i < OFF_W2; i += blockDim.x) sw[i] = a.P[i];
  for (int i = threadIdx.x; i < HID; i += blockDim.x) {
    sb2[i] = a.P[OFF_B2 + i];
    sw3[i] = a.P[OFF_W3 + i];
  }
  if (threadIdx.x == 0) sb3 = a.P[OFF_B3];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, li = lane & 15;
  // B operand: B[k][j] = W2[j][k]; tile t covers j = 16t + li, step s covers k = 32s + 8g + e
  bf8v bw[4][2];
  for (int t = 0; t < 4; ++t)
    for (int s = 0; s < 2; ++s) bw[t][s] = *(const bf8v*)(a.P16 + OFF_W2 + (16 * t + li) * HID + 32 * s + 8 * g);
  const int64_t chunks_per_col = (a.n + 15) / 16;
  const int64_t total = chunks_per_col * a.nc;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t ch = wave; ch < total; ch += nwave) {
    const int col = (int)(ch / chunks_per_col);
    const int64_t e0 = (ch - (int64_t)col * chunks_per_col) * 16;
    const double th = a.theta[col];
    const bool dpr_only = a.use_dpr[col] != 0;
    const double* res = a.Res + a.ld * col;
    double* T = a.T + a.ld * col;
    const double* bstep = a.base ? a.base + a.ld * col : nullptr;
    if (dpr_only) {
      if (li < 4) {
        for (int q = 0; q < 4; ++q) {
          const int64_t j = e0 + 4 * g + q;
          if (li == q && j < a.n) T[j] = bstep ? bstep[j] : res[j] / guard_den2(th - a.hii[j]);
        }
      }
      continue;
    }
    // layer 1 for element e0 + li (all four lane groups compute 16 of its 64 units)
    const int64_t jel = e0 + li;
    float x0 = 0.f, x2 = 0.f;
    int c = 0;
    if (jel < a.n) {
      x0 = asinhf((float)(a.hii[jel] - th));
      x2 = asinhf((float)(res[jel] / a.rnorm[col]));
      c = class_of(a, jel);
    }
    const float x1 = (float)th;
    bf8v ha[2];
    for (int s = 0; s < 2; ++s)
      for (int e = 0; e < 8; ++e) {
        const int kk = 32 * s + 8 * g + e;
        const float z = sw[OFF_W1 + kk * 3] * x0 + sw[OFF_W1 + kk * 3 + 1] * x1 + sw[OFF_W1 + kk * 3 + 2] * x2 +
                        sw[OFF_E + c * HID + kk] + sw[OFF_B1 + kk];
        ha[s][e] = (__bf16)silu(z);
      }
    // layer 2 on MFMA: rows = 16 elements, cols = 64 hidden units
    f4v acc[4];
    for (int t = 0; t < 4; ++t) {
      acc[t] = (f4v){0.f, 0.f, 0.f, 0.f};
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[0], bw[t][0], acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[1], bw[t][1], acc[t], 0, 0, 0);
    }
    // layer 3: register r holds row (element) 4g + r, column (unit) 16t + li
    float part[4];
    for (int r = 0; r < 4; ++r) {
      float s = 0.f;
      for (int t = 0; t < 4; ++t) {
        const int jj = 16 * t + li;
        s += sw3[jj] * silu(acc[t][r] + sb2[jj]);
      }
      part[r] = s;
    }
    for (int off = 1; off < 16; off <<= 1)
      for (int r = 0; r < 4; ++r) part[r] += __shfl_xor(part[r], off, 64);
    // lane (g, li < 4) writes element 4g + li
    if (li < 4) {
      const int64_t j = e0 + 4 * g + li;
      float m = part[0];
      if (li == 1) m = part[1];
      if (li == 2) m = part[2];
      if (li == 3) m = part[3];
      m += sb3;
      if (j < a.n) {
        const double dstep = bstep ? bstep[j] : res[j] / guard_den2(th - a.hii[j]);
        if (isfinite(m)) {
          T[j] = dstep * (1.0 + (double)m);
        } else {
          T[j] = dstep;
          atomicAdd(a.nonfinite, 1);
        }
      }
    }
  }
}

// ---- harvesting --------------------------------------------------------------
__global__ void coldot_kernel(int64_t n, int nc, const double* __restrict__ A, int64_t lda,
                              const double* __restrict__ B, int64_t ldb, double* __restrict__ part) {
  __shared__ double red[16];
  for (int c = 0; c < nc; ++c) {
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
      s += A[i + lda * c] * B[i + ldb * c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (ln == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
      part[(int64_t)blockIdx.x * nc + c] = t;
    }
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct SampleArgs {
  int64_t n, n1;
  int nroots, per_root;
  int64_t ld;
  int64_t off[6];
  uint64_t seed;
  const double* theta;   // per root (device)
  const double* rnorm;   // per root
  const double* s;       // x . x*
  const double* enorm;   // RMS element of x - x*/s
  const int* active;     // root sampled?
  const double* hii;
  const double* X;
  const double* Xs;      // converged eigenvectors (ld)
  const int* match;      // column of Xs for each root
  const double* Res;
  const double* bstep;   // base step of each root (ld), or NULL for the DPR step
  float* feat;
  int* cls;
  float* dn;
  float* en;
  int64_t base;          // record offset
};

__global__ void sample_kernel(SampleArgs a) {
  const int64_t total = (int64_t)a.nroots * a.per_root;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / a.per_root), q = (int)(idx % a.per_root);
    const int64_t out = a.base + idx;
    const uint64_t h = mix64(a.seed ^ ((uint64_t)idx * 0x100000001B3ull));
    const int64_t j = (q & 1) ? (a.n1 + (int64_t)(h % (uint64_t)(a.n - a.n1))) : (int64_t)(h % (uint64_t)a.n1);
    if (!a.active[r]) {
      a.feat[3 * out] = 0.f; a.feat[3 * out + 1] = 0.f; a.feat[3 * out + 2] = 0.f;
      a.cls[out] = 0; a.dn[out] = 0.f; a.en[out] = 0.f;
      continue;
    }
    const double th = a.theta[r];
    const double rj = a.Res[j + a.ld * r];
    const double dj = a.bstep ? a.bstep[j + a.ld * r] : rj / guard_den2(th - a.hii[j]);
    const double ej = (a.X[j + a.ld * r] - a.Xs[j + a.ld * a.match[r]] / a.s[r]) / a.enorm[r];
    a.feat[3 * out] = asinhf((float)(a.hii[j] - th));
    a.feat[3 * out + 1] = (float)th;
    a.feat[3 * out + 2] = asinhf((float)(rj / a.rnorm[r]));
    int c = 0;
    if (j >= a.off[0]) c = 1;
    if (j >= a.off[1]) c = 2;
    if (j >= a.off[2]) c = 3;
    if (j >= a.off[3]) c = 4;
    if (j >= a.off[4]) c = 5;
    a.cls[out] = c;
    a.dn[out] = (float)(dj / a.enorm[r]);
    a.en[out] = (float)ej;
  }
}

}  // namespace

// Network state and training records owned by a handle.
struct PreconNet {
  DevBuf<float> W, Mo, Vo, Pb, grad, gpart;
  DevBuf<__bf16> P16;
  DevBuf<double> lpart, loss;
  DevBuf<int> nonfinite;
  bool have = false;
  int64_t adam_t = 0;
  // records
  DevBuf<float> feat, dn, en;
  DevBuf<int> cls;
  int64_t nrec = 0, cap = 0;
  void init_buffers() {
    W.alloc(NP); Mo.alloc(NP); Vo.alloc(NP); Pb.alloc(NP); grad.alloc(NP); P16.alloc(NP);
    loss.alloc(1); nonfinite.alloc(1);
  }
  void reserve(int64_t want) {
    if (want <= cap) return;
    int64_t nc = std::max<int64_t>(want, cap * 2);
    DevBuf<float> f2, d2, e2;
    DevBuf<int> c2;
    f2.alloc(3 * nc); d2.alloc(nc); e2.alloc(nc); c2.alloc(nc);
    if (nrec > 0) {
      HIP_CHECK(hipMemcpy(f2.get(), feat.get(), sizeof(float) * 3 * nrec, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(d2.get(), dn.get(), sizeof(float) * nrec, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(e2.get(), en.get(), sizeof(float) * nrec, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(c2.get(), cls.get(), sizeof(int) * nrec, hipMemcpyDeviceToDevice));
    }
    std::swap(feat.p, f2.p); std::swap(feat.n, f2.n);
    std::swap(dn.p, d2.p); std::swap(dn.n, d2.n);
    std::swap(en.p, e2.p); std::swap(en.n, e2.n);
    std::swap(cls.p, c2.p); std::swap(cls.n, c2.n);
    cap = nc;
  }
};

struct DavTrace {
  DevBuf<double> Xs;         // converged eigenvectors, ndim x nstates
  int nstates = 0;
  int per_root = 16;
  int64_t max_records = 0;
  int64_t added = 0;
  uint64_t seed = 1;
  double tol = 1e-7;
  DevBuf<double> dbuf;       // theta, rnorm, s, enorm per root
  DevBuf<int> active, match;
  DevBuf<double> part, ovl;
};

static PreconNet& net_of(Handle& H) {
  if (!H.net) {
    H.net = new PreconNet();
    H.net->init_buffers();
  }
  return *H.net;
}

void precon_net_free(Handle& H) {
  delete H.net;
  H.net = nullptr;
}

int precon_nparams() { return NP; }

bool precon_ready(const Handle& H) { return H.net && H.net->have; }

void precon_set(Handle& H, const float* params) {
  PreconNet& N = net_of(H);
  HIP_CHECK(hipMemcpy(N.W.get(), params, sizeof(float) * NP, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemset(N.Mo.get(), 0, sizeof(float) * NP));
  HIP_CHECK(hipMemset(N.Vo.get(), 0, sizeof(float) * NP));
  hipLaunchKernelGGL(round_params_kernel, dim3((NP + 255) / 256), dim3(256), 0, H.st, N.W.get(), N.Pb.get(),
                     N.P16.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(H.st));
  N.adam_t = 0;
  N.have = true;
}

void precon_absorb(Handle& dst, const Handle& src, int64_t* nrec) {
  PreconNet& D = net_of(dst);
  const PreconNet* S = src.net;
  const int64_t m = S ? S->nrec : 0;
  if (m > 0) {
    D.reserve(D.nrec + m);
    HIP_CHECK(hipMemcpy(D.feat.get() + 3 * D.nrec, S->feat.get(), sizeof(float) * 3 * m, hipMemcpyDeviceToDevice));
    HIP_CHECK(hipMemcpy(D.dn.get() + D.nrec, S->dn.get(), sizeof(float) * m, hipMemcpyDeviceToDevice));
    HIP_CHECK(hipMemcpy(D.en.get() + D.nrec, S->en.get(), sizeof(float) * m, hipMemcpyDeviceToDevice));
    HIP_CHECK(hipMemcpy(D.cls.get() + D.nrec, S->cls.get(), sizeof(int) * m, hipMemcpyDeviceToDevice));
    D.nrec += m;
  }
  if (nrec) *nrec = D.nrec;
}

void precon_get(const Handle& H, float* params) {
  if (!precon_ready(H)) throw Error(ADC2G_ESTATE, "no network on this handle");
  HIP_CHECK(hipMemcpy(params, H.net->W.get(), sizeof(float) * NP, hipMemcpyDeviceToHost));
}

// Fixed initialisation (when training starts without set): small random first
// layers, zero output layer (the network starts as plain DPR).
static void init_default(Handle& H, uint64_t seed) {
  std::vector<float> p(NP, 0.0f);
  uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto nrm = [&]() {
    // Box-Muller from a 64-bit LCG
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const double u1 = ((s >> 11) + 0.5) / 9007199254740992.0;
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const double u2 = ((s >> 11) + 0.5) / 9007199254740992.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < HID * 3; ++i) p[OFF_W1 + i] = (float)(nrm() / std::sqrt(3.0));
  for (int i = 0; i < NCLS * HID; ++i) p[OFF_E + i] = (float)(0.1 * nrm());
  for (int i = 0; i < HID * HID; ++i) p[OFF_W2 + i] = (float)(nrm() / 8.0);
  precon_set(H, p.data());
}

// Mean loss and gradient over records [0, n) of the given arrays (device).
static double fwdbwd(Handle& H, PreconNet& N, const float* feat, const int* cls, const float* dn, const float* en,
                     int64_t n0, int64_t n, float* mout) {
  const int64_t cnt = n - n0;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(cnt, 2048));
  N.gpart.alloc((size_t)nblk * NP);
  N.lpart.alloc(nblk);
  hipLaunchKernelGGL(fwdbwd_kernel, dim3(nblk), dim3(64), 0, H.st, N.Pb.get(), n0, n, feat, cls, dn, en,
                     1.0f / (float)cnt, N.gpart.get(), N.lpart.get(), mout);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((NP + 255) / 256), dim3(256), 0, H.st, nblk, N.gpart.get(),
                     N.grad.get(), N.lpart.get(), N.loss.get());
  HIP_CHECK(hipGetLastError());
  double loss = 0;
  HIP_CHECK(hipMemcpyAsync(&loss, N.loss.get(), sizeof(double), hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  return loss / (double)cnt;
}

void precon_eval(Handle& H, int64_t n, const float* feat, const int32_t* cls, const float* dn, const float* en,
                 float* m_out, float* grad_out, double* loss_out) {
  if (!precon_ready(H)) throw Error(ADC2G_ESTATE, "no network on this handle");
  PreconNet& N = *H.net;
  DevBuf<float> f, d, e, m;
  DevBuf<int> c;
  f.alloc(3 * n); d.alloc(n); e.alloc(n); m.alloc(n); c.alloc(n);
  HIP_CHECK(hipMemcpy(f.get(), feat, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d.get(), dn, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(e.get(), en, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c.get(), cls, sizeof(int) * n, hipMemcpyHostToDevice));
  for (int64_t i = 0; i < n; ++i)
    ARG_CHECK(cls[i] >= 0 && cls[i] < NCLS, "class id out of range");
  const double loss = fwdbwd(H, N, f.get(), c.get(), d.get(), e.get(), 0, n, m.get());
  if (loss_out) *loss_out = loss;
  if (m_out) HIP_CHECK(hipMemcpy(m_out, m.get(), sizeof(float) * n, hipMemcpyDeviceToHost));
  if (grad_out) HIP_CHECK(hipMemcpy(grad_out, N.grad.get(), sizeof(float) * NP, hipMemcpyDeviceToHost));
}

void precon_train(Handle& H, const Handle& data, int epochs, int batch, double lr, uint64_t seed,
                  double* loss_out) {
  ARG_CHECK(epochs >= 0 && batch >= 1, "epochs must be >= 0 and batch >= 1");
  const PreconNet* D = data.net;
  if (!D || D->nrec == 0) throw Error(ADC2G_ESTATE, "no harvested records on the data handle");
  PreconNet& N = net_of(H);
  if (!N.have) init_default(H, seed);
  const int64_t nrec = D->nrec;
  // records are harvested in random element order already; mini-batches are
  // contiguous slices starting at a seed-dependent rotation each epoch
  uint64_t s = seed * 2862933555777941757ull + 3037000493ull;
  for (int ep = 0; ep < epochs; ++ep) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const int64_t start = (int64_t)((s >> 17) % (uint64_t)nrec);
    double acc = 0.0;
    int64_t nb = 0;
    for (int64_t b0 = 0; b0 < nrec; b0 += batch) {
      int64_t lo = (start + b0) % nrec;
      int64_t hi = std::min<int64_t>(lo + batch, nrec);
      const double l = fwdbwd(H, N, D->feat.get(), D->cls.get(), D->dn.get(), D->en.get(), lo, hi, nullptr);
      ++N.adam_t;
      const float c1 = 1.0f - std::pow(0.9f, (float)N.adam_t), c2 = 1.0f - std::pow(0.999f, (float)N.adam_t);
      hipLaunchKernelGGL(adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, H.st, N.W.get(), N.Mo.get(), N.Vo.get(),
                         N.grad.get(), N.Pb.get(), N.P16.get(), (float)lr, c1, c2);
      HIP_CHECK(hipGetLastError());
      acc += l;
      ++nb;
    }
    if (loss_out) loss_out[ep] = acc / (double)std::max<int64_t>(nb, 1);
  }
  HIP_CHECK(hipStreamSynchronize(H.st));
  N.have = true;
}

void precon_learned_step(Handle& H, int64_t n, int nc, const double* theta_d, const double* rnorm_h,
                         const int* use_dpr_h, const double* Res, double* T, int64_t ld, const double* base) {
  PreconNet& N = *H.net;
  DevBuf<double> rn;
  DevBuf<int> ud;
  rn.alloc(nc);
  ud.alloc(nc);
  HIP_CHECK(hipMemcpyAsync(rn.get(), rnorm_h, sizeof(double) * nc, hipMemcpyHostToDevice, H.st));
  HIP_CHECK(hipMemcpyAsync(ud.get(), use_dpr_h, sizeof(int) * nc, hipMemcpyHostToDevice, H.st));
  InferArgs a;
  a.n = n; a.nc = nc; a.ld = ld;
  for (int t = 0; t < 6; ++t) a.off[t] = H.d.off[t];
  a.theta = theta_d; a.rnorm = rn.get(); a.use_dpr = ud.get();
  a.hii = H.hii.get(); a.Res = Res; a.T = T; a.base = base;
  a.P = N.Pb.get(); a.P16 = N.P16.get(); a.nonfinite = N.nonfinite.get();
  const int64_t waves = std::min<int64_t>(((n + 15) / 16) * nc, (int64_t)2048 * 4);
  hipLaunchKernelGGL(infer_kernel, dim3((int)((waves + 3) / 4)), dim3(256), 0, H.st, a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(H.st));  // rn/ud are freed on return
}

void davtrace_iteration(Handle& H, DavTrace& tr, int it, int bs_curr, int ns_curr, const double* theta_h,
                        const double* norm_h, const double* X, const double* Res, int64_t n, const double* base) {
  (void)bs_curr;
  if (tr.added >= tr.max_records) return;
  const int nr = ns_curr;  // active (not yet deflated) roots
  if (nr <= 0) return;
  PreconNet& N = net_of(H);
  const int ns = tr.nstates;
  tr.dbuf.alloc(4 * (size_t)nr);
  tr.active.alloc(nr);
  tr.match.alloc(nr);
  tr.ovl.alloc((size_t)nr * ns);
  // overlaps O = X^T X* (nr x ns): each active root is matched to the
  // converged eigenvector it overlaps most (deflation reorders the roots)
  {
    GemmArgs g;
    g.M = nr; g.N = ns; g.K = (int)n;
    g.A = X; g.sam = n; g.sak = 1;
    g.B = tr.Xs.get(); g.sbk = 1; g.sbn = n;
    g.map.C = tr.ovl.get(); g.map.scm = 1; g.map.scn = nr;
    gemm(g, Prec::F64, H.gws, H.st);
  }
  std::vector<double> O((size_t)nr * ns);
  HIP_CHECK(hipMemcpyAsync(O.data(), tr.ovl.get(), sizeof(double) * O.size(), hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  std::vector<double> h(4 * (size_t)nr);
  std::vector<int> act(nr), mt(nr);
  int nact = 0;
  for (int i = 0; i < nr; ++i) {
    int best = 0;
    for (int j = 1; j < ns; ++j)
      if (std::fabs(O[i + (size_t)j * nr]) > std::fabs(O[i + (size_t)best * nr])) best = j;
    const double s = O[i + (size_t)best * nr];
    mt[i] = best;
    // |x - x*/s|^2 = 1/s^2 - 1 for unit x, x*; targets are divided by its RMS
    const double e2 = (s != 0.0) ? (1.0 / (s * s) - 1.0) : 0.0;
    act[i] = (norm_h[i] >= tr.tol && std::fabs(s) > 0.5 && e2 > 1e-24) ? 1 : 0;
    nact += act[i];
    h[i] = theta_h[i];
    h[nr + i] = norm_h[i];
    h[2 * nr + i] = s;
    h[3 * nr + i] = std::sqrt(std::max(e2, 1e-300) / (double)n);
  }
  if (nact == 0) return;
  HIP_CHECK(hipMemcpyAsync(tr.dbuf.get(), h.data(), sizeof(double) * 4 * nr, hipMemcpyHostToDevice, H.st));
  HIP_CHECK(hipMemcpyAsync(tr.active.get(), act.data(), sizeof(int) * nr, hipMemcpyHostToDevice, H.st));
  HIP_CHECK(hipMemcpyAsync(tr.match.get(), mt.data(), sizeof(int) * nr, hipMemcpyHostToDevice, H.st));
  const int64_t cnt = (int64_t)nr * tr.per_root;
  N.reserve(N.nrec + cnt);
  SampleArgs a;
  a.n = n; a.n1 = H.d.n1; a.nroots = nr; a.per_root = tr.per_root; a.ld = n;
  for (int t = 0; t < 6; ++t) a.off[t] = H.d.off[t];
  a.seed = tr.seed + (uint64_t)it * 0x9E3779B97F4A7C15ull;
  a.theta = tr.dbuf.get(); a.rnorm = tr.dbuf.get() + nr; a.s = tr.dbuf.get() + 2 * nr;
  a.enorm = tr.dbuf.get() + 3 * nr; a.active = tr.active.get();
  a.hii = H.hii.get(); a.X = X; a.Xs = tr.Xs.get(); a.match = tr.match.get(); a.Res = Res; a.bstep = base;
  a.feat = N.feat.get(); a.cls = N.cls.get(); a.dn = N.dn.get(); a.en = N.en.get();
  a.base = N.nrec;
  hipLaunchKernelGGL(sample_kernel, dim3(grid_for(cnt)), dim3(256), 0, H.st, a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(H.st));
  // keep only the active roots' records (inactive ones were written as zeros and are dropped by compaction)
  if (nact < nr) {
    std::vector<float> f(3 * cnt), d(cnt), e(cnt);
    std::vector<int> c(cnt);
    HIP_CHECK(hipMemcpy(f.data(), N.feat.get() + 3 * N.nrec, sizeof(float) * 3 * cnt, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(d.data(), N.dn.get() + N.nrec, sizeof(float) * cnt, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(e.data(), N.en.get() + N.nrec, sizeof(float) * cnt, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(c.data(), N.cls.get() + N.nrec, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    int64_t w = 0;
    for (int i = 0; i < nr; ++i) {
      if (!act[i]) continue;
      for (int q = 0; q < tr.per_root; ++q, ++w) {
        const int64_t src = (int64_t)i * tr.per_root + q;
        f[3 * w] = f[3 * src]; f[3 * w + 1] = f[3 * src + 1]; f[3 * w + 2] = f[3 * src + 2];
        d[w] = d[src]; e[w] = e[src]; c[w] = c[src];
      }
    }
    HIP_CHECK(hipMemcpy(N.feat.get() + 3 * N.nrec, f.data(), sizeof(float) * 3 * w, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(N.dn.get() + N.nrec, d.data(), sizeof(float) * w, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(N.en.get() + N.nrec, e.data(), sizeof(float) * w, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(N.cls.get() + N.nrec, c.data(), sizeof(int) * w, hipMemcpyHostToDevice));
    N.nrec += w;
    tr.added += w;
  } else {
    N.nrec += cnt;
    tr.added += cnt;
  }
}

void precon_harvest(Handle& H, const adc2g_dav_params& p, int64_t max_records, int64_t* added) {
  // pass 1: FP64 Davidson with the base step (DPR, or the coupled step for
  // the coupled preconditioners), with deflation, to convergence -> x*
  adc2g_dav_params q = p;
  q.precon = (p.precon == ADC2G_PRE_COUPLED || p.precon == ADC2G_PRE_LEARNED_COUPLED) ? ADC2G_PRE_COUPLED
                                                                                       : ADC2G_PRE_DPR;
  q.deflate = 1;
  q.precision = ADC2G_PREC_FP64;
  q.verbose = 0;
  const int64_t n = H.d.ndim;
  DavTrace tr;
  tr.nstates = q.nstates;
  tr.Xs.alloc((size_t)n * q.nstates);
  std::vector<double> ev(q.nstates), vec((size_t)n * q.nstates);
  davidson_run(H, q, ev.data(), vec.data(), nullptr, nullptr);
  HIP_CHECK(hipMemcpy(tr.Xs.get(), vec.data(), sizeof(double) * vec.size(), hipMemcpyHostToDevice));
  // pass 2: the same (deterministic) trajectory, sampling records each iteration
  tr.max_records = max_records > 0 ? max_records : ((int64_t)1 << 40);
  tr.tol = q.tol;
  tr.seed = 0x5EEDull ^ ((uint64_t)H.d.o << 32) ^ ((uint64_t)H.d.v << 8) ^ (uint64_t)H.d.naux;
  davidson_run(H, q, ev.data(), nullptr, nullptr, &tr);
  if (added) *added = tr.added;
}

}  // namespace adc2g
