// Laser-driven propagation: the time-dependent operator of propagate_laser_adc2
// (source/propagation/proplib_adc2.f90:37-630).
//
// The wavefunction has ndim + 1 components: the intermediate states, then the
// MP2 ground state Psi0.  With E(t) the sum of the pulses (efield,
// source/propagation/tdselib.f90:850-1029) and D_c the intermediate-state
// representation of dipole component c, matxvec_treal_laser
// (tdselib.f90:569-720) applies
//   IS rows:      -i H v + i sum_c E_c [(D_c + d00_c) v + d0j_c v0]
//   ground row:            i sum_c E_c (d00_c v0 + d0j_c . v)
// (<Psi0|H|Psi_J> = 0 and H00 = 0 for the ground-state-shifted H).  D_c comes
// from the caller as the reference stores it (SCRATCH/dipole_<c>, read by
// opxvec_ext, tdselib.f90:1031-1153): the diagonal, then each off-diagonal
// element once, applied to (i, j) and (j, i).  Here it is a CSR matrix with
// both triangles, one thread per row; vectors are planar complex (re column,
// im column) with leading dimension ndim + 1, so -i H v is one two-column
// sigma-vector of the first ndim rows.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

struct LaserComp {
  double E, d00;
  const double* diag;
  const double* val;
  const double* d0j;
  const int64_t* rowptr;
  const int32_t* col;
};

struct LaserDev {
  LaserComp c[3];
  int nc;
};

__device__ double laser_block_sum(double x, double* sh) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = x;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}

typedef double cplx2 __attribute__((ext_vector_type(2)));

// (re, im) interleaved copy of the first nd rows of a planar complex vector:
// the CSR products gather one 16-byte element per stored entry
__global__ void interleave_kernel(int64_t nd, int64_t n, const double* __restrict__ v, cplx2* __restrict__ vz) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd; i += (int64_t)gridDim.x * blockDim.x)
    vz[i] = (cplx2){v[i], v[n + i]};
}

// sum over row i of the off-diagonal part: the LN lanes of a group take every
// LN-th stored element (coalesced reads of col / val) and combine their sums
// with shuffles inside the group, so every lane of the group ends with the
// row's sum
template <int LN>
__device__ __forceinline__ void csr_row(const LaserComp& c, int64_t i, int s, const cplx2* __restrict__ vz,
                                        double& re, double& im) {
  re = 0.0;
  im = 0.0;
  const int64_t p1 = c.rowptr[i + 1];
  for (int64_t p = c.rowptr[i] + s; p < p1; p += LN) {
    const double w = c.val[p];
    const cplx2 z = vz[c.col[p]];
    re += w * z[0];
    im += w * z[1];
  }
#pragma unroll
  for (int off = LN / 2; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, 64);
    im += __shfl_xor(im, off, 64);
  }
}

// intermediate-state rows, LN lanes per row:
// out = -i hv + i sum_c E_c [(D_c + d00_c) v + d0j_c v0]
template <int LN>
__global__ void laser_is_kernel(int64_t nd, int64_t n, const double* __restrict__ hv, const double* __restrict__ v,
                                const cplx2* __restrict__ vz, double* __restrict__ out, LaserDev L) {
  const double v0re = v[nd], v0im = v[n + nd];
  const int s = threadIdx.x % LN;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x / LN);
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LN; i < nd; i += stride) {
    double sre = 0.0, sim = 0.0;
    for (int q = 0; q < L.nc; ++q) {
      const LaserComp& c = L.c[q];
      double dre, dim;
      csr_row<LN>(c, i, s, vz, dre, dim);
      const double dd = c.diag[i] + c.d00;
      dre += dd * v[i] + c.d0j[i] * v0re;
      dim += dd * v[n + i] + c.d0j[i] * v0im;
      sre += c.E * dre;
      sim += c.E * dim;
    }
    // -i (hre + i him) = him - i hre;  i (sre + i sim) = -sim + i sre
    if (s == 0) {
      out[i] = hv[n + i] - sim;
      out[n + i] = -hv[i] + sre;
    }
  }
}

// partial sums of sum_c E_c d0j_c . v over the intermediate states
__global__ void laser_ground_partial_kernel(int64_t nd, int64_t n, const double* __restrict__ v, LaserDev L,
                                            double* __restrict__ part) {
  __shared__ double sh[16];
  double sre = 0.0, sim = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd; i += (int64_t)gridDim.x * blockDim.x) {
    double w = 0.0;
    for (int q = 0; q < L.nc; ++q) w += L.c[q].E * L.c[q].d0j[i];
    sre += w * v[i];
    sim += w * v[n + i];
  }
  sre = laser_block_sum(sre, sh);
  sim = laser_block_sum(sim, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sre;
    part[2 * blockIdx.x + 1] = sim;
  }
}

// ground row: out0 = i (sum_c E_c d00_c v0 + the partial sums)
__global__ void laser_ground_final_kernel(int nblk, int64_t nd, int64_t n, const double* __restrict__ v, LaserDev L,
                                          const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[16];
  double sre = 0.0, sim = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    sre += part[2 * b];
    sim += part[2 * b + 1];
  }
  sre = laser_block_sum(sre, sh);
  sim = laser_block_sum(sim, sh);
  if (threadIdx.x == 0) {
    double e00 = 0.0;
    for (int q = 0; q < L.nc; ++q) e00 += L.c[q].E * L.c[q].d00;
    const double gre = sre + e00 * v[nd], gim = sim + e00 * v[n + nd];
    out[nd] = -gim;
    out[n + nd] = gre;
  }
}

// out = scale [op (diag u + off u) + add u + d0j u0] on the IS rows, LN
// lanes per row (no -i: the CAP term -W u of the reference takes op = 1,
// add = w00; the flux's operators op = 1 + their ground element, add = 0)
template <int LN>
__global__ void isop_is_kernel(int64_t nd, int64_t n, const double* __restrict__ u, const cplx2* __restrict__ uz,
                               double* __restrict__ out, LaserComp c, double scale, double op, double add) {
  const double u0re = u[nd], u0im = u[n + nd];
  const int s = threadIdx.x % LN;
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x / LN);
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LN; i < nd; i += stride) {
    double re, im;
    csr_row<LN>(c, i, s, uz, re, im);
    const double dg = c.diag[i];
    if (s == 0) {
      out[i] = scale * (op * (re + dg * u[i]) + add * u[i] + c.d0j[i] * u0re);
      out[n + i] = scale * (op * (im + dg * u[n + i]) + add * u[n + i] + c.d0j[i] * u0im);
    }
  }
}

__global__ void isop_ground_partial_kernel(int64_t nd, int64_t n, const double* __restrict__ u, LaserComp c,
                                           double* __restrict__ part) {
  __shared__ double sh[16];
  double sre = 0.0, sim = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd; i += (int64_t)gridDim.x * blockDim.x) {
    sre += c.d0j[i] * u[i];
    sim += c.d0j[i] * u[n + i];
  }
  sre = laser_block_sum(sre, sh);
  sim = laser_block_sum(sim, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sre;
    part[2 * blockIdx.x + 1] = sim;
  }
}

__global__ void isop_ground_final_kernel(int nblk, int64_t nd, int64_t n, const double* __restrict__ u, LaserComp c,
                                         double scale, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[16];
  double sre = 0.0, sim = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    sre += part[2 * b];
    sim += part[2 * b + 1];
  }
  sre = laser_block_sum(sre, sh);
  sim = laser_block_sum(sim, sh);
  if (threadIdx.x == 0) {
    out[nd] = scale * (sre + c.d00 * u[nd]);
    out[n + nd] = scale * (sim + c.d00 * u[n + nd]);
  }
}

__global__ void zero_ground_kernel(int64_t nd, int64_t n, double* __restrict__ x) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    x[nd] = 0.0;
    x[n + nd] = 0.0;
  }
}

__global__ void add_planar_kernel(int64_t m, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

// one launch of a row kernel with LN = 1, 4, 8, 16 or 32 lanes per row, from
// the stored off-diagonal elements per row (grid-stride over the rows)
template <class F>
void launch_rows(double per_row, int64_t nd, F&& f) {
  auto go = [&](auto ln) { f(ln, dim3(grid_for(nd * decltype(ln)::value))); };
  if (per_row <= 1.0) go(std::integral_constant<int, 1>());
  else if (per_row <= 6.0) go(std::integral_constant<int, 4>());
  else if (per_row <= 12.0) go(std::integral_constant<int, 8>());
  else if (per_row <= 24.0) go(std::integral_constant<int, 16>());
  else go(std::integral_constant<int, 32>());
}

double envelope_value(const adc2g_pulse& p, double t) {
  const double pi = 3.14159265358979323846;
  switch (p.envelope) {
    case 1: {  // cos^2 about tzero, half-width fwhm
      const double tz = p.envpar[0], fw = p.envpar[1];
      if (t > tz - fw && t < tz + fw) {
        const double c = std::cos(pi * (t - tz) / (2.0 * fw));
        return c * c;
      }
      return 0.0;
    }
    case 2: {  // sin^2 ramp of length tau, then 1
      const double tau = p.envpar[0];
      if (t < 0.0) return 0.0;
      if (t <= tau) {
        const double s = std::sin(pi * t / (2.0 * tau));
        return s * s;
      }
      return 1.0;
    }
    case 3:  // box on [ti, tf]
      return (t < p.envpar[0] || t > p.envpar[1]) ? 0.0 : 1.0;
    case 4: {  // sin^2 from tzero over 2 fwhm
      const double tz = p.envpar[0], fw = p.envpar[1];
      if (t > tz && t < tz + 2.0 * fw) {
        const double s = std::sin(pi * (t - tz) / (2.0 * fw));
        return s * s;
      }
      return 0.0;
    }
  }
  return 0.0;
}

}  // namespace

void laser_check_pulses(int npulse, const adc2g_pulse* pulses) {
  ARG_CHECK(npulse >= 0, "npulse must be >= 0");
  ARG_CHECK(npulse == 0 || pulses, "NULL pulses");
  for (int i = 0; i < npulse; ++i) {
    const adc2g_pulse& p = pulses[i];
    ARG_CHECK(p.carrier == 1 || p.carrier == 2, "pulse carrier must be 1 (cosine) or 2 (sine)");
    ARG_CHECK(p.envelope >= 1 && p.envelope <= 4, "pulse envelope must be 1 (cos^2), 2 (sin^2 ramp), 3 (box) or 4 (sin^2)");
    ARG_CHECK((p.envelope != 1 && p.envelope != 4) || p.envpar[1] > 0.0, "the cos^2 / sin^2 envelope needs fwhm > 0");
    ARG_CHECK(p.envelope != 2 || p.envpar[0] > 0.0, "the sin^2 ramp needs tau > 0");
  }
}

void laser_efield(int npulse, const adc2g_pulse* pulses, double t, double E[3]) {
  E[0] = E[1] = E[2] = 0.0;
  for (int i = 0; i < npulse; ++i) {
    const adc2g_pulse& p = pulses[i];
    const double arg = p.freq * (t - p.t0) - p.phase;
    const double carrier = (p.carrier == 1) ? std::cos(arg) : std::sin(arg);
    const double f = p.strength * envelope_value(p, t) * carrier;
    for (int c = 0; c < 3; ++c) E[c] += p.pulse_vec[c] * f;
  }
}

namespace {

// an IS operator in the reference's file layout -> D (CSR with both triangles)
void fill_is_operator(Handle& Hd, Handle::DipoleIS& D, const double* diag_h, int64_t nnz, const int64_t* row,
                      const int64_t* colh, const double* valh, double d00, const double* d0j_h) {
  ARG_CHECK(nnz >= 0, "nnz must be >= 0");
  const Dims& d = Hd.d;
  const hipStream_t st = Hd.st;
  if (nnz == 0 && !diag_h && !d0j_h && d00 == 0.0) {
    D.set = false;
    D.diag.release();
    D.val.release();
    D.d0j.release();
    D.rowptr.release();
    D.col.release();
    D.d00 = 0.0;
    D.per_row = 0.0;
    return;
  }
  ARG_CHECK(diag_h && d0j_h, "the diagonal and the ground-state column are required");
  ARG_CHECK(nnz == 0 || (row && colh && valh), "NULL off-diagonal arrays");
  const int64_t nd = d.ndim;
  ARG_CHECK(nd < ((int64_t)1 << 31), "the configuration space is too large for 32-bit column indices");
  std::vector<int64_t> rp(nd + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) {
    ARG_CHECK(row[k] >= 0 && row[k] < nd && colh[k] >= 0 && colh[k] < nd,
              "off-diagonal element " + std::to_string(k) + " has an index outside 0 .. ndim-1");
    ARG_CHECK(row[k] != colh[k], "off-diagonal element " + std::to_string(k) + " lies on the diagonal");
    ++rp[row[k] + 1];
    ++rp[colh[k] + 1];
  }
  for (int64_t i = 0; i < nd; ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<int32_t> cl(std::max<int64_t>(2 * nnz, 1));
  std::vector<double> vl(std::max<int64_t>(2 * nnz, 1));
  for (int64_t k = 0; k < nnz; ++k) {
    int64_t p = fill[row[k]]++;
    cl[p] = (int32_t)colh[k];
    vl[p] = valh[k];
    p = fill[colh[k]]++;
    cl[p] = (int32_t)row[k];
    vl[p] = valh[k];
  }
  HIP_CHECK(hipSetDevice(Hd.device));
  D.diag.alloc(nd);
  D.d0j.alloc(nd);
  D.rowptr.alloc(nd + 1);
  D.col.alloc(cl.size());
  D.val.alloc(vl.size());
  HIP_CHECK(hipMemcpyAsync(D.diag.get(), diag_h, sizeof(double) * nd, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(D.d0j.get(), d0j_h, sizeof(double) * nd, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(D.rowptr.get(), rp.data(), sizeof(int64_t) * (nd + 1), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(D.col.get(), cl.data(), sizeof(int32_t) * cl.size(), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(D.val.get(), vl.data(), sizeof(double) * vl.size(), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipStreamSynchronize(st));
  D.d00 = d00;
  D.per_row = (double)(2 * nnz) / (double)nd;
  D.set = true;
}

}  // namespace

void Handle::set_dipole_is(int c, const double* diag_h, int64_t nnz, const int64_t* row, const int64_t* colh,
                           const double* valh, double d00, const double* d0j_h) {
  ARG_CHECK(c >= 0 && c <= 2, "dipole component must be 0 (x), 1 (y) or 2 (z)");
  fill_is_operator(*this, dip[c], diag_h, nnz, row, colh, valh, d00, d0j_h);
}

void Handle::set_cap_is(const double* diag_h, int64_t nnz, const int64_t* row, const int64_t* colh,
                        const double* valh, double w00, const double* w0j_h) {
  fill_is_operator(*this, cap, diag_h, nnz, row, colh, valh, w00, w0j_h);
}

void Handle::set_flux_theta(const double* diag_h, int64_t nnz, const int64_t* row, const int64_t* colh,
                            const double* valh, double theta00, const double* theta0j_h) {
  fill_is_operator(*this, flux_theta, diag_h, nnz, row, colh, valh, theta00, theta0j_h);
}

void Handle::set_cap_projector(int nvec, const double* vecs, bool ground) {
  ARG_CHECK(nvec >= 0, "nvec must be >= 0");
  ARG_CHECK(nvec == 0 || vecs, "NULL projector vectors");
  HIP_CHECK(hipSetDevice(device));
  if (nvec > 0) {
    // the projector 1 - R R^T needs orthonormal columns (the reference's
    // eigenvectors are)
    double dev = 0.0;
    for (int a = 0; a < nvec; ++a)
      for (int b = 0; b <= a; ++b) {
        double g = 0.0;
        const double* x = vecs + (size_t)d.ndim * a;
        const double* y = vecs + (size_t)d.ndim * b;
        for (int64_t i = 0; i < d.ndim; ++i) g += x[i] * y[i];
        dev = std::max(dev, std::fabs(g - (a == b ? 1.0 : 0.0)));
      }
    ARG_CHECK(dev <= 1e-8, "the projector vectors must be orthonormal (max |R^T R - I| = " + std::to_string(dev) + ")");
    capproj.alloc((size_t)d.ndim * nvec);
    HIP_CHECK(hipMemcpyAsync(capproj.get(), vecs, sizeof(double) * d.ndim * nvec, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
  } else {
    capproj.release();
  }
  capproj_n = nvec;
  capproj_ground = ground;
}

// x <- P x: the selected bound states and, if chosen, the ground state
// projected out (planar complex, ndim + 1 rows)
void cap_project(Handle& H, double* x) {
  const int64_t nd = H.d.ndim, n = nd + 1;
  if (H.capproj_n > 0) {
    H.laser_coef.alloc(2 * (size_t)H.capproj_n);
    gemm_tn(H, H.capproj_n, 2, nd, H.capproj.get(), nd, x, n, H.laser_coef.get(), H.capproj_n, 1.0, 0.0);
    gemm_nn(H, nd, 2, H.capproj_n, H.capproj.get(), nd, H.laser_coef.get(), H.capproj_n, x, n, -1.0, 1.0);
  }
  if (H.capproj_ground) {
    hipLaunchKernelGGL(zero_ground_kernel, dim3(1), dim3(64), 0, H.st, nd, n, x);
    HIP_CHECK(hipGetLastError());
  }
}

// out = scale [op (D_IS u) + add u + d0j u0 ; d00 u0 + d0j . u] for an IS
// operator D (diagonal + CSR, ground element d00, column d0j)
void apply_is_operator(Handle& H, const Handle::DipoleIS& D, const double* u, double* out, double scale, double op,
                       double add) {
  const int64_t nd = H.d.ndim, n = nd + 1;
  const hipStream_t s = H.st;
  LaserComp w;
  w.E = 1.0;
  w.d00 = D.d00;
  w.diag = D.diag.get();
  w.val = D.val.get();
  w.d0j = D.d0j.get();
  w.rowptr = D.rowptr.get();
  w.col = D.col.get();
  H.laser_vz.alloc(2 * (size_t)nd);
  cplx2* uz = (cplx2*)H.laser_vz.get();
  if (D.per_row > 0.0) {
    hipLaunchKernelGGL(interleave_kernel, dim3(grid_for(nd)), dim3(256), 0, s, nd, n, u, uz);
    HIP_CHECK(hipGetLastError());
  }
  launch_rows(D.per_row, nd, [&](auto ln, dim3 grid) {
    hipLaunchKernelGGL(isop_is_kernel<decltype(ln)::value>, grid, dim3(256), 0, s, nd, n, u, uz, out, w, scale, op,
                       add);
  });
  HIP_CHECK(hipGetLastError());
  const int nbc = reduce_blocks(nd);
  H.laser_part.alloc(2 * (size_t)std::max(nbc, reduce_blocks(n)));
  hipLaunchKernelGGL(isop_ground_partial_kernel, dim3(nbc), dim3(256), 0, s, nd, n, u, w, H.laser_part.get());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(isop_ground_final_kernel, dim3(1), dim3(256), 0, s, nbc, nd, n, u, w, scale, H.laser_part.get(),
                     out);
  HIP_CHECK(hipGetLastError());
}

void laser_apply(Handle& H, const double E[3], const double* v, double* hv, double* out, Prec p, hipStream_t s) {
  const int64_t nd = H.d.ndim, n = nd + 1;
  H.sigma(v, n, hv, n, 2, p, s);
  LaserDev L{};
  L.nc = 0;
  double per_row = 0.0;  // lanes per row from the stored elements per row of the active components
  for (int c = 0; c < 3; ++c) {
    if (E[c] == 0.0) continue;
    const Handle::DipoleIS& D = H.dip[c];
    ARG_CHECK(D.set, std::string("the field has a ") + "xyz"[c] +
                         " component but that dipole component is not set (adc2g_set_dipole_is)");
    per_row = std::max(per_row, D.per_row);
    LaserComp& q = L.c[L.nc++];
    q.E = E[c];
    q.d00 = D.d00;
    q.diag = D.diag.get();
    q.val = D.val.get();
    q.d0j = D.d0j.get();
    q.rowptr = D.rowptr.get();
    q.col = D.col.get();
  }
  H.laser_vz.alloc(2 * (size_t)nd);
  cplx2* vz = (cplx2*)H.laser_vz.get();
  if (per_row > 0.0) {
    hipLaunchKernelGGL(interleave_kernel, dim3(grid_for(nd)), dim3(256), 0, s, nd, n, v, vz);
    HIP_CHECK(hipGetLastError());
  }
  launch_rows(per_row, nd, [&](auto ln, dim3 grid) {
    hipLaunchKernelGGL(laser_is_kernel<decltype(ln)::value>, grid, dim3(256), 0, s, nd, n, hv, v, vz, out, L);
  });
  HIP_CHECK(hipGetLastError());
  const int nb = L.nc > 0 ? reduce_blocks(nd) : 0;
  H.laser_part.alloc(2 * (size_t)std::max(nb, 1));
  if (nb > 0) {
    hipLaunchKernelGGL(laser_ground_partial_kernel, dim3(nb), dim3(256), 0, s, nd, n, v, L, H.laser_part.get());
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(laser_ground_final_kernel, dim3(1), dim3(256), 0, s, nb, nd, n, v, L, H.laser_part.get(), out);
  HIP_CHECK(hipGetLastError());
  if (!H.cap.set) return;
  // CAP (tdselib.f90:722-830): out += -P [(W + w00) u + w0j u0] (IS rows) and
  // -P (w00 u0 + w0j . u) (ground row), u = P v
  ARG_CHECK(s == H.st, "the CAP term runs on the handle's stream");
  H.laser_u.alloc(2 * (size_t)n);
  H.laser_c.alloc(2 * (size_t)n);
  HIP_CHECK(hipMemcpyAsync(H.laser_u.get(), v, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
  cap_project(H, H.laser_u.get());
  apply_is_operator(H, H.cap, H.laser_u.get(), H.laser_c.get(), -1.0, 1.0, H.cap.d00);
  cap_project(H, H.laser_c.get());
  hipLaunchKernelGGL(add_planar_kernel, dim3(grid_for(2 * n)), dim3(256), 0, s, 2 * n, H.laser_c.get(), out);
  HIP_CHECK(hipGetLastError());
}

// adc2_flux_cap (source/propagation/flux.f90:179-411):
//   flux = 2 Re <dpsi/dt | P Theta' P psi> + 2 Re <psi | P W' P psi>,
// Theta the IS representation of the projector onto the absorbing region
// (SCRATCH/theta, theta00, theta0j) and W that of the CAP.  The reference
// adds theta00 (w00) times the operator's IS output rather than times the
// vector on the IS rows (oppsi = oppsi + theta00 * oppsi); Theta' and W' are
// those operators as written there.
double laser_flux(Handle& H, const double* psi, const double* dtpsi) {
  ARG_CHECK(H.cap.set && H.flux_theta.set, "the flux needs a CAP and the Theta operator (adc2g_set_flux_theta)");
  const int64_t nd = H.d.ndim, n = nd + 1;
  const hipStream_t s = H.st;
  H.laser_u.alloc(2 * (size_t)n);
  H.laser_c.alloc(2 * (size_t)n);
  double* u = H.laser_u.get();
  double* c = H.laser_c.get();
  auto term = [&](const Handle::DipoleIS& D, const double* bra) {
    HIP_CHECK(hipMemcpyAsync(u, psi, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, s));
    cap_project(H, u);
    apply_is_operator(H, D, u, c, 1.0, 1.0 + D.d00, 0.0);
    cap_project(H, c);
    const int nbk = launch_cdot_partial(n, bra, c, n, H.laser_part.get(), s);
    H.laser_coef.alloc(2);
    launch_colsum(nbk, 2, H.laser_part.get(), H.laser_coef.get(), s);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, H.laser_coef.get(), sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return 2.0 * h[0];
  };
  H.laser_part.alloc(2 * (size_t)reduce_blocks(n));
  return term(H.flux_theta, dtpsi) + term(H.cap, psi);
}

void matxvec_treal_laser_host(Handle& H, double t, int npulse, const adc2g_pulse* pulses, const double* v1,
                              double* v2) {
  laser_check_pulses(npulse, pulses);
  const int64_t n = H.d.ndim + 1;
  double E[3];
  laser_efield(npulse, pulses, t, E);
  std::vector<double> planar(2 * n);
  for (int64_t i = 0; i < n; ++i) {
    planar[i] = v1[2 * i];
    planar[n + i] = v1[2 * i + 1];
  }
  DevBuf<double> dv, dh, dout;
  dv.alloc(2 * n);
  dh.alloc(2 * n);
  dout.alloc(2 * n);
  HIP_CHECK(hipMemcpyAsync(dv.get(), planar.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, H.st));
  laser_apply(H, E, dv.get(), dh.get(), dout.get(), Prec::F64, H.st);
  HIP_CHECK(hipMemcpyAsync(planar.data(), dout.get(), sizeof(double) * 2 * n, hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  for (int64_t i = 0; i < n; ++i) {
    v2[2 * i] = planar[i];
    v2[2 * i + 1] = planar[n + i];
  }
}

}  // namespace adc2g
