/*
 * Stored-matrix (COO) ADC(2)-s sigma-vector on the CPU.  TEST ORACLE AND CPU
 * BASELINE ONLY -- never linked into the product library.
 *
 * Restates the reference's default CPU path:
 *  - the off-diagonal build get_offdiag_adc2_save_omp
 *    (source/adclib/get_matrix.f90:705-1186): upper triangle only, ph-ph
 *    pairs i<j (:848-884), then the five 1h1p-2h2p blocks in the order
 *    t1 (C5), t2 (C4), t3 (C3), t4 (C1), t5 (C2) (:895-1054), each as
 *    "do i over 1h1p rows, do j over the configurations of the block",
 *    keeping |h| > minc = 1e-12 (source/include/parameters.f90:134);
 *  - the element functions C1..C5_ph_2p2h (source/adclib/adc_ph.f90:974-1073),
 *    evaluated literally from integral slices (pq|rs);
 *  - the sigma-vector hxvec_incore (source/eigen/block_davidson.f90:844-883):
 *    W = hii * V, then for every column (OpenMP over columns)
 *    W(i_k) += h_k V(j_k), W(j_k) += h_k V(i_k).
 *
 * Row sampling: the build covers the 1h1p rows of the virtual orbitals
 * [a0, a1) (all occupied k), so a bounded sample of the matrix can be timed;
 * the cost of hxvec_incore is linear in the number of stored elements.
 *
 * Integrals come from the density-fitting factors B[P][p][q]:
 *   Iv[a-a0][c][l][b] = (a c | l b)   (a in the sample, c, b virtual, l occupied)
 *   Io[k][x][y][b]    = (x k | y b)   (k, x, y occupied, b virtual)
 * which are all the integrals the coupling functions read for these rows.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MINC 1e-12

typedef struct {
  int o, v, naux, nbas;
  int64_t n1, npo, npv, off[6];
} dims_t;

static void set_dims(dims_t* d, int o, int v, int naux) {
  d->o = o; d->v = v; d->naux = naux; d->nbas = o + v;
  d->n1 = (int64_t)o * v;
  d->npo = (int64_t)o * (o - 1) / 2;
  d->npv = (int64_t)v * (v - 1) / 2;
  d->off[0] = d->n1;
  d->off[1] = d->off[0] + (int64_t)o * v;
  d->off[2] = d->off[1] + d->npv * o;
  d->off[3] = d->off[2] + d->npo * v;
  d->off[4] = d->off[3] + d->npo * d->npv;
  d->off[5] = d->off[4] + d->npo * d->npv;
}

int64_t coo_ndim(int o, int v) {
  dims_t d;
  set_dims(&d, o, v, 1);
  return d.off[5];
}

/* (pq|rs) from B for MO indices */
static double eri(const double* B, const dims_t* d, int p, int q, int r, int s) {
  const int64_t nb2 = (int64_t)d->nbas * d->nbas;
  double acc = 0.0;
  for (int P = 0; P < d->naux; ++P) acc += B[P * nb2 + (int64_t)p * d->nbas + q] * B[P * nb2 + (int64_t)r * d->nbas + s];
  return acc;
}

/* Iv and Io slices (sizes (a1-a0)*v*o*v and o*o*o*v) */
static void build_slices(const double* B, const dims_t* d, int a0, int a1, double* Iv, double* Io) {
  const int o = d->o, v = d->v, naux = d->naux, nb = d->nbas;
  const int64_t nb2 = (int64_t)nb * nb;
  const int na = a1 - a0;
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int ia = 0; ia < na; ++ia)
    for (int c = 0; c < v; ++c) {
      double* out = Iv + (((int64_t)ia * v + c) * o) * v;
      memset(out, 0, sizeof(double) * o * v);
      const int A = o + a0 + ia, Cc = o + c;
      for (int P = 0; P < naux; ++P) {
        const double bac = B[P * nb2 + (int64_t)A * nb + Cc];
        const double* row = B + P * nb2;
        for (int l = 0; l < o; ++l) {
          const double* bl = row + (int64_t)l * nb + o;
          double* ol = out + (int64_t)l * v;
          for (int b = 0; b < v; ++b) ol[b] += bac * bl[b];
        }
      }
    }
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int k = 0; k < o; ++k)
    for (int x = 0; x < o; ++x)
      for (int y = 0; y < o; ++y) {
        double* out = Io + (((int64_t)k * o + x) * o + y) * v;
        for (int b = 0; b < v; ++b) out[b] = eri(B, d, x, k, y, o + b);
      }
}

static inline int64_t pair_index(int64_t lo, int64_t hi, int64_t n) { return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1); }

/* coupling element of 1h1p row (a,k) [a virtual index, k occupied] with
 * 2h2p config (type t, a', b', k', l') -- adc_ph.f90:974-1073, indices are the
 * 0-based virtual/occupied numbers; ia = a - a0 indexes Iv. */
static double coupling(const dims_t* d, const double* Iv, const double* Io, int ia, int a, int k, int t, int ap,
                       int bp, int kp, int lp) {
  const int o = d->o, v = d->v;
#define IV(c, l, b) Iv[(((int64_t)ia * v + (c)) * o + (l)) * v + (b)]   /* (a c|l b) */
#define IO(x, y, b) Io[(((int64_t)k * o + (x)) * o + (y)) * v + (b)]    /* (x k|y b) */
  double f = 0.0;
  switch (t) {
    case 4: /* C1_ph_2p2h */
      if (a == ap) f -= IO(kp, lp, bp) + IO(lp, kp, bp);
      if (a == bp) f -= IO(kp, lp, ap) + IO(lp, kp, ap);
      if (k == kp) f += IV(ap, lp, bp) + IV(bp, lp, ap);
      if (k == lp) f += IV(ap, kp, bp) + IV(bp, kp, ap);
      return f / sqrt(2.0);
    case 5: /* C2_ph_2p2h */
      if (a == ap) f -= IO(kp, lp, bp) - IO(lp, kp, bp);
      if (a == bp) f += IO(kp, lp, ap) - IO(lp, kp, ap);
      if (k == kp) f += IV(ap, lp, bp) - IV(bp, lp, ap);
      if (k == lp) f -= IV(ap, kp, bp) - IV(bp, kp, ap);
      return sqrt(3.0) * f / sqrt(2.0);
    case 3: /* C3_ph_2p2h(j,k,ipr,kpr,lpr), a' = b' */
      if (a == ap) f -= IO(kp, lp, ap) + IO(lp, kp, ap);
      if (k == kp) f += IV(ap, lp, ap);
      if (k == lp) f += IV(ap, kp, ap);
      return f;
    case 2: /* C4_ph_2p2h(a,k,apr,bpr,kpr), k' = l' */
      if (a == ap) f += IO(kp, kp, bp);
      if (a == bp) f += IO(kp, kp, ap);
      if (k == kp) f -= IV(ap, kp, bp) + IV(bp, kp, ap);
      return f;
    default: /* 1: C5_ph_2p2h(a,k,apr,kpr) */
      if (a == ap) f += IO(kp, kp, ap);
      if (k == kp) f -= IV(ap, kp, ap);
      return sqrt(2.0) * f;
  }
#undef IV
#undef IO
}

/* pair tables: pair number -> (lo, hi), the inverse of pair_index */
typedef struct { int *pa, *pb, *pk, *pl; } tables_t;

static void make_tables(const dims_t* d, tables_t* T) {
  T->pa = (int*)malloc(sizeof(int) * (d->npv + 1));
  T->pb = (int*)malloc(sizeof(int) * (d->npv + 1));
  T->pk = (int*)malloc(sizeof(int) * (d->npo + 1));
  T->pl = (int*)malloc(sizeof(int) * (d->npo + 1));
  int64_t c = 0;
  for (int a = 0; a < d->v; ++a)
    for (int b = a + 1; b < d->v; ++b) { T->pa[c] = a; T->pb[c] = b; ++c; }
  c = 0;
  for (int k = 0; k < d->o; ++k)
    for (int l = k + 1; l < d->o; ++l) { T->pk[c] = k; T->pl[c] = l; ++c; }
}

static void free_tables(tables_t* T) { free(T->pa); free(T->pb); free(T->pk); free(T->pl); }

/* decode config j of block t (0-based within the block) */
static inline void decode(const dims_t* d, const tables_t* T, int t, int64_t j, int* ap, int* bp, int* kp, int* lp) {
  const int o = d->o, v = d->v;
  if (t == 1) { *kp = *lp = (int)(j / v); *ap = *bp = (int)(j % v); return; }
  if (t == 2) {
    const int64_t p = j / o; *kp = *lp = (int)(j % o);
    *ap = T->pa[p]; *bp = T->pb[p]; return;
  }
  if (t == 3) {
    const int64_t q = j / v; *ap = *bp = (int)(j % v);
    *kp = T->pk[q]; *lp = T->pl[q]; return;
  }
  {
    const int64_t q = j / d->npv, p = j % d->npv;
    *ap = T->pa[p]; *bp = T->pb[p]; *kp = T->pk[q]; *lp = T->pl[q];
  }
}

int64_t coo_row_coupling_count(int o, int v);

/* Upper bound of stored elements for rows of [a0,a1): ph-ph pairs + couplings */
int64_t coo_capacity(int o, int v, int a0, int a1) {
  const int64_t rows = (int64_t)(a1 - a0) * o;
  return rows * ((int64_t)o * v + coo_row_coupling_count(o, v));
}

/*
 * Build the COO records of the 1h1p rows with virtual index in [a0,a1).
 * h11 (n1 x n1, row-major, may be NULL) supplies the ph-ph values (the
 * second-order ph-ph terms are not restated here; the oracle's Python
 * restatement pins them).  Returns the number of stored elements.
 */
int64_t coo_build(int o, int v, int naux, const double* B, const double* h11, int a0, int a1, double* val,
                  int32_t* oi, int32_t* oj) {
  dims_t d;
  set_dims(&d, o, v, naux);
  const int na = a1 - a0;
  double* Iv = (double*)malloc(sizeof(double) * (size_t)na * v * o * v);
  double* Io = (double*)malloc(sizeof(double) * (size_t)o * o * o * v);
  build_slices(B, &d, a0, a1, Iv, Io);
  tables_t tb;
  make_tables(&d, &tb);
  const int64_t r0 = (int64_t)a0 * o, r1 = (int64_t)a1 * o;
  int64_t count = 0;
  /* ph-ph block, i < j */
  if (h11) {
    for (int64_t i = r0; i < r1; ++i)
      for (int64_t j = i + 1; j < d.n1; ++j) {
        const double h = h11[i * d.n1 + j];
        if (fabs(h) > MINC) { val[count] = h; oi[count] = (int32_t)(i + 1); oj[count] = (int32_t)(j + 1); ++count; }
      }
  }
  /* coupling blocks, type by type (t1..t5), rows outer, configs inner */
  for (int t = 1; t <= 5; ++t) {
    const int64_t c0 = d.off[t - 1], c1 = d.off[t];
    const int64_t nrow = r1 - r0;
    int64_t* rc = (int64_t*)calloc((size_t)nrow + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic)
    for (int64_t ii = 0; ii < nrow; ++ii) {
      const int64_t i = r0 + ii;
      const int a = (int)(i / o), k = (int)(i % o);
      int64_t c = 0;
      for (int64_t j = c0; j < c1; ++j) {
        int ap, bp, kp, lp;
        decode(&d, &tb, t, j - c0, &ap, &bp, &kp, &lp);
        if (t == 1 && !(a == ap || k == kp)) continue;
        if ((t == 2 || t == 4 || t == 5) && !(a == ap || a == bp || k == kp || k == lp)) continue;
        if (t == 3 && !(a == ap || k == kp || k == lp)) continue;
        const double h = coupling(&d, Iv, Io, a - a0, a, k, t, ap, bp, kp, lp);
        if (fabs(h) > MINC) ++c;
      }
      rc[ii + 1] = c;
    }
    for (int64_t ii = 0; ii < nrow; ++ii) rc[ii + 1] += rc[ii];
#pragma omp parallel for schedule(dynamic)
    for (int64_t ii = 0; ii < nrow; ++ii) {
      const int64_t i = r0 + ii;
      const int a = (int)(i / o), k = (int)(i % o);
      int64_t c = count + rc[ii];
      for (int64_t j = c0; j < c1; ++j) {
        int ap, bp, kp, lp;
        decode(&d, &tb, t, j - c0, &ap, &bp, &kp, &lp);
        if (t == 1 && !(a == ap || k == kp)) continue;
        if ((t == 2 || t == 4 || t == 5) && !(a == ap || a == bp || k == kp || k == lp)) continue;
        if (t == 3 && !(a == ap || k == kp || k == lp)) continue;
        const double h = coupling(&d, Iv, Io, a - a0, a, k, t, ap, bp, kp, lp);
        if (fabs(h) > MINC) { val[c] = h; oi[c] = (int32_t)(i + 1); oj[c] = (int32_t)(j + 1); ++c; }
      }
    }
    count += rc[nrow];
    free(rc);
  }
  free(Iv);
  free(Io);
  free_tables(&tb);
  return count;
}

/* hxvec_incore: W = hii*V over all ndim rows, then the COO contributions.
 * V, W column-major ndim x nvec; indices 1-based.  Returns seconds spent in
 * the off-diagonal part (the diagonal part time goes to *t_diag). */
double coo_hxvec(int64_t ndim, int nvec, const double* hii, int64_t nnz, const double* val, const int32_t* oi,
                 const int32_t* oj, const double* V, double* W, double* t_diag) {
  double t0 = omp_get_wtime();
#pragma omp parallel for
  for (int n = 0; n < nvec; ++n)
    for (int64_t m = 0; m < ndim; ++m) W[(int64_t)n * ndim + m] = hii[m] * V[(int64_t)n * ndim + m];
  double t1 = omp_get_wtime();
#pragma omp parallel for
  for (int n = 0; n < nvec; ++n) {
    const double* v = V + (int64_t)n * ndim;
    double* w = W + (int64_t)n * ndim;
    for (int64_t k = 0; k < nnz; ++k) {
      const int64_t i = oi[k] - 1, j = oj[k] - 1;
      w[i] += val[k] * v[j];
      w[j] += val[k] * v[i];
    }
  }
  double t2 = omp_get_wtime();
  if (t_diag) *t_diag = t1 - t0;
  return t2 - t1;
}

/* Structurally non-zero coupling elements per 1h1p row (every row has the
 * same count): used to scale a sampled timing to the full matrix. */
int64_t coo_row_coupling_count(int o, int v) {
  const int64_t npo = (int64_t)o * (o - 1) / 2, npv = (int64_t)v * (v - 1) / 2;
  const int64_t t1 = v + o - 1;
  const int64_t t2 = (int64_t)o * (v - 1) + npv - (v - 1);
  const int64_t t3 = npo + (int64_t)v * (o - 1) - (o - 1);
  const int64_t t4 = (int64_t)(v - 1) * npo + (int64_t)(o - 1) * npv - (int64_t)(v - 1) * (o - 1);
  return t1 + t2 + t3 + 2 * t4;
}
