/* Integrals over contracted Cartesian Gaussians (s, p, d) by the
 * McMurchie-Davidson scheme: overlap, kinetic energy, nuclear attraction,
 * dipole and two-electron repulsion integrals.
 *
 * Host code of the molecular front end (the reference computes its own AO
 * integrals from the GAMESS basis, source/qclib/load_electronic_structure.f90:153-342).
 * Every primitive carries the normalisation of x^l exp(-a r^2) times the
 * contraction coefficient; the caller renormalises each contracted Cartesian
 * component from the overlap diagonal.  Cartesian components of a shell in
 * GAMESS order: s; x y z; xx yy zz xy xz yz.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LMAX 2
#define LE (LMAX + 2)               /* i, j range of the E tables (kinetic needs j + 2) */
#define TE (2 * LMAX + 4)           /* t range of the E tables (one past the largest i + j, read as zero) */
#define LR (4 * LMAX)               /* highest total order of the ERI R tensor */
#define NR (LR + 1)

static const int CART[LMAX + 1][6][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
};
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* Boys function F_n(T), n = 0..nmax: downward recursion from the series for
 * T < 35, upward from F_0 = sqrt(pi/T)/2 beyond (erf(sqrt 35) = 1 - 1e-17). */
static void boys(int nmax, double T, double* F) {
  if (T < 1e-14) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1);
    return;
  }
  const double et = exp(-T);
  if (T < 35.0) {
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
  } else {
    F[0] = 0.5 * sqrt(M_PI / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
  }
}

/* Hermite expansion coefficients E[i][j][t] of the product of the 1D
 * Gaussians x_A^i exp(-a x_A^2) x_B^j exp(-b x_B^2), AB = A - B. */
static void etab(int imax, int jmax, double a, double b, double AB, double E[LE + 1][LE + 1][TE + 1]) {
  const double p = a + b, mu = a * b / p;
  const double XPA = -b * AB / p, XPB = a * AB / p, h = 0.5 / p;
  memset(E, 0, sizeof(double) * (LE + 1) * (LE + 1) * (TE + 1));
  E[0][0][0] = exp(-mu * AB * AB);
  for (int i = 0; i <= imax; ++i) {
    if (i > 0)
      for (int t = 0; t <= i; ++t)
        E[i][0][t] = (t > 0 ? h * E[i - 1][0][t - 1] : 0.0) + XPA * E[i - 1][0][t] + (t + 1) * E[i - 1][0][t + 1];
    for (int j = 1; j <= jmax; ++j)
      for (int t = 0; t <= i + j; ++t)
        E[i][j][t] = (t > 0 ? h * E[i][j - 1][t - 1] : 0.0) + XPB * E[i][j - 1][t] + (t + 1) * E[i][j - 1][t + 1];
  }
}

/* Hermite Coulomb integrals R[t][u][v] = R^0_{tuv}(alpha, PC), t+u+v <= L. */
static void rtab(int L, double alpha, double X, double Y, double Z, double R[NR][NR][NR]) {
  static __thread double Rn[NR + 1][NR][NR][NR];
  double F[NR + 1];
  boys(L, alpha * (X * X + Y * Y + Z * Z), F);
  double f = 1.0;
  for (int n = 0; n <= L; ++n) {
    Rn[n][0][0][0] = f * F[n];
    f *= -2.0 * alpha;
  }
  for (int n = L - 1; n >= 0; --n) {
    const int top = L - n;
    for (int s = 1; s <= top; ++s)
      for (int t = 0; t <= s; ++t)
        for (int u = 0; u <= s - t; ++u) {
          const int v = s - t - u;
          double r;
          if (t > 0)
            r = (t > 1 ? (t - 1) * Rn[n + 1][t - 2][u][v] : 0.0) + X * Rn[n + 1][t - 1][u][v];
          else if (u > 0)
            r = (u > 1 ? (u - 1) * Rn[n + 1][t][u - 2][v] : 0.0) + Y * Rn[n + 1][t][u - 1][v];
          else
            r = (v > 1 ? (v - 1) * Rn[n + 1][t][u][v - 2] : 0.0) + Z * Rn[n + 1][t][u][v - 1];
          Rn[n][t][u][v] = r;
        }
  }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

static double dfact(int n) { /* (n)!! for odd n >= -1 */
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

/* normalisation of x^l exp(-a r^2) */
static double pnorm(double a, int l) {
  return pow(2.0 * a / M_PI, 0.75) * pow(4.0 * a, 0.5 * l) / sqrt(dfact(2 * l - 1));
}

typedef struct {
  int nshell;
  const int *atom, *l, *nprim, *off, *ao;  /* ao: first AO index of the shell */
  const double *exps, *coefs, *xyz;        /* xyz: atom coordinates (bohr), 3 per atom */
} Basis;

static int check_basis(const Basis* B) {
  for (int s = 0; s < B->nshell; ++s)
    if (B->l[s] < 0 || B->l[s] > LMAX) return -1;
  return 0;
}

/* S, T, V (nuclear attraction with charges Zc at natom centres) and the three
 * dipole components (origin at the coordinate origin), nao x nao row-major. */
int ints_one_electron(int nshell, const int* atom, const int* l, const int* nprim, const int* off, const int* ao,
                      const double* exps, const double* coefs, const double* xyz, int natom, const double* Zc,
                      int nao, double* S, double* T, double* V, double* Dx, double* Dy, double* Dz) {
  Basis B = {nshell, atom, l, nprim, off, ao, exps, coefs, xyz};
  if (check_basis(&B)) return -1;
#pragma omp parallel for schedule(dynamic)
  for (int si = 0; si < nshell; ++si)
    for (int sj = 0; sj < nshell; ++sj) {
      const int la = l[si], lb = l[sj];
      const double* A = xyz + 3 * atom[si];
      const double* Bc = xyz + 3 * atom[sj];
      const int na = ncart(la), nb = ncart(lb);
      double s[6][6] = {{0}}, t[6][6] = {{0}}, v[6][6] = {{0}}, dx[6][6] = {{0}}, dy[6][6] = {{0}}, dz[6][6] = {{0}};
      for (int pa = 0; pa < nprim[si]; ++pa)
        for (int pb = 0; pb < nprim[sj]; ++pb) {
          const double a = exps[off[si] + pa], b = exps[off[sj] + pb];
          const double ca = coefs[off[si] + pa] * pnorm(a, la), cb = coefs[off[sj] + pb] * pnorm(b, lb);
          const double p = a + b;
          double Ex[LE + 1][LE + 1][TE + 1], Ey[LE + 1][LE + 1][TE + 1], Ez[LE + 1][LE + 1][TE + 1];
          etab(la, lb + 2, a, b, A[0] - Bc[0], Ex);
          etab(la, lb + 2, a, b, A[1] - Bc[1], Ey);
          etab(la, lb + 2, a, b, A[2] - Bc[2], Ez);
          const double P[3] = {(a * A[0] + b * Bc[0]) / p, (a * A[1] + b * Bc[1]) / p, (a * A[2] + b * Bc[2]) / p};
          const double sp = pow(M_PI / p, 1.5);
          /* nuclear attraction: R tensors per centre */
          const int L = la + lb;
          for (int c = 0; c < natom; ++c) {
            double R[NR][NR][NR];
            rtab(L, p, P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1], P[2] - xyz[3 * c + 2], R);
            const double pref = -Zc[c] * 2.0 * M_PI / p * ca * cb;
            for (int ia = 0; ia < na; ++ia)
              for (int ib = 0; ib < nb; ++ib) {
                const int *ea = CART[la][ia], *eb = CART[lb][ib];
                double acc = 0.0;
                for (int tt = 0; tt <= ea[0] + eb[0]; ++tt)
                  for (int uu = 0; uu <= ea[1] + eb[1]; ++uu)
                    for (int vv = 0; vv <= ea[2] + eb[2]; ++vv)
                      acc += Ex[ea[0]][eb[0]][tt] * Ey[ea[1]][eb[1]][uu] * Ez[ea[2]][eb[2]][vv] * R[tt][uu][vv];
                v[ia][ib] += pref * acc;
              }
          }
          for (int ia = 0; ia < na; ++ia)
            for (int ib = 0; ib < nb; ++ib) {
              const int *ea = CART[la][ia], *eb = CART[lb][ib];
              const double ov[3] = {Ex[ea[0]][eb[0]][0], Ey[ea[1]][eb[1]][0], Ez[ea[2]][eb[2]][0]};
              double kin[3], mom[3];
              const double (*Et[3])[LE + 1][TE + 1] = {Ex, Ey, Ez};
              for (int d = 0; d < 3; ++d) {
                const int i = ea[d], j = eb[d];
                const double (*E)[LE + 1][TE + 1] = Et[d];
                kin[d] = b * (2 * j + 1) * E[i][j][0] - 2.0 * b * b * E[i][j + 2][0] -
                         (j >= 2 ? 0.5 * j * (j - 1) * E[i][j - 2][0] : 0.0);
                mom[d] = E[i][j][1] + P[d] * E[i][j][0];
              }
              const double w = ca * cb * sp;
              s[ia][ib] += w * ov[0] * ov[1] * ov[2];
              t[ia][ib] += w * (kin[0] * ov[1] * ov[2] + ov[0] * kin[1] * ov[2] + ov[0] * ov[1] * kin[2]);
              dx[ia][ib] += w * mom[0] * ov[1] * ov[2];
              dy[ia][ib] += w * ov[0] * mom[1] * ov[2];
              dz[ia][ib] += w * ov[0] * ov[1] * mom[2];
            }
        }
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
          const int64_t k = (int64_t)(ao[si] + ia) * nao + ao[sj] + ib;
          S[k] = s[ia][ib];
          T[k] = t[ia][ib];
          V[k] = v[ia][ib];
          Dx[k] = dx[ia][ib];
          Dy[k] = dy[ia][ib];
          Dz[k] = dz[ia][ib];
        }
    }
  return 0;
}

/* (ab|cd) for every AO quadruple, nao^4 row-major; each unique shell quartet
 * is computed once and written to its eight permutations. */
int ints_eri(int nshell, const int* atom, const int* l, const int* nprim, const int* off, const int* ao,
             const double* exps, const double* coefs, const double* xyz, int nao, double* eri) {
  Basis B = {nshell, atom, l, nprim, off, ao, exps, coefs, xyz};
  if (check_basis(&B)) return -1;
  const int npair = nshell * (nshell + 1) / 2;
  const int64_t n = nao;
#pragma omp parallel for schedule(dynamic)
  for (int ij = 0; ij < npair; ++ij) {
    int si = 0;
    while ((si + 1) * (si + 2) / 2 <= ij) ++si;
    const int sj = ij - si * (si + 1) / 2;
    for (int kl = 0; kl <= ij; ++kl) {
      int sk = 0;
      while ((sk + 1) * (sk + 2) / 2 <= kl) ++sk;
      const int sl = kl - sk * (sk + 1) / 2;
      const int la = l[si], lb = l[sj], lc = l[sk], ld = l[sl];
      const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
      const double *A = xyz + 3 * atom[si], *Bc = xyz + 3 * atom[sj], *Cc = xyz + 3 * atom[sk],
                   *D = xyz + 3 * atom[sl];
      double out[6][6][6][6];
      memset(out, 0, sizeof(out));
      for (int pa = 0; pa < nprim[si]; ++pa)
        for (int pb = 0; pb < nprim[sj]; ++pb) {
          const double a = exps[off[si] + pa], b = exps[off[sj] + pb];
          const double cab = coefs[off[si] + pa] * pnorm(a, la) * coefs[off[sj] + pb] * pnorm(b, lb);
          const double p = a + b;
          const double P[3] = {(a * A[0] + b * Bc[0]) / p, (a * A[1] + b * Bc[1]) / p, (a * A[2] + b * Bc[2]) / p};
          double Eab[3][LE + 1][LE + 1][TE + 1];
          for (int d = 0; d < 3; ++d) etab(la, lb, a, b, A[d] - Bc[d], Eab[d]);
          for (int pc = 0; pc < nprim[sk]; ++pc)
            for (int pd = 0; pd < nprim[sl]; ++pd) {
              const double c = exps[off[sk] + pc], dd = exps[off[sl] + pd];
              const double ccd = coefs[off[sk] + pc] * pnorm(c, lc) * coefs[off[sl] + pd] * pnorm(dd, ld);
              const double q = c + dd;
              const double Q[3] = {(c * Cc[0] + dd * D[0]) / q, (c * Cc[1] + dd * D[1]) / q,
                                   (c * Cc[2] + dd * D[2]) / q};
              double Ecd[3][LE + 1][LE + 1][TE + 1];
              for (int d = 0; d < 3; ++d) etab(lc, ld, c, dd, Cc[d] - D[d], Ecd[d]);
              const double alpha = p * q / (p + q);
              double R[NR][NR][NR];
              const int L = la + lb + lc + ld;
              rtab(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
              const double pref = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * cab * ccd;
              for (int ic = 0; ic < nc; ++ic)
                for (int id = 0; id < nd; ++id) {
                  const int *ec = CART[lc][ic], *ed = CART[ld][id];
                  /* G[t][u][v] = sum_{tau,nu,phi} (-1)^{tau+nu+phi} Ecd R[t+tau][u+nu][v+phi] */
                  double G[2 * LMAX + 1][2 * LMAX + 1][2 * LMAX + 1];
                  const int Lab = la + lb;
                  for (int t = 0; t <= Lab; ++t)
                    for (int u = 0; u <= Lab - t; ++u)
                      for (int v = 0; v <= Lab - t - u; ++v) {
                        double g = 0.0;
                        for (int tau = 0; tau <= ec[0] + ed[0]; ++tau)
                          for (int nu = 0; nu <= ec[1] + ed[1]; ++nu)
                            for (int ph = 0; ph <= ec[2] + ed[2]; ++ph) {
                              const double sgn = ((tau + nu + ph) & 1) ? -1.0 : 1.0;
                              g += sgn * Ecd[0][ec[0]][ed[0]][tau] * Ecd[1][ec[1]][ed[1]][nu] *
                                   Ecd[2][ec[2]][ed[2]][ph] * R[t + tau][u + nu][v + ph];
                            }
                        G[t][u][v] = g;
                      }
                  for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib) {
                      const int *ea = CART[la][ia], *eb = CART[lb][ib];
                      double acc = 0.0;
                      for (int t = 0; t <= ea[0] + eb[0]; ++t)
                        for (int u = 0; u <= ea[1] + eb[1]; ++u)
                          for (int v = 0; v <= ea[2] + eb[2]; ++v)
                            acc += Eab[0][ea[0]][eb[0]][t] * Eab[1][ea[1]][eb[1]][u] * Eab[2][ea[2]][eb[2]][v] *
                                   G[t][u][v];
                      out[ia][ib][ic][id] += pref * acc;
                    }
                }
            }
        }
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
          for (int ic = 0; ic < nc; ++ic)
            for (int id = 0; id < nd; ++id) {
              const int64_t i = ao[si] + ia, j = ao[sj] + ib, k = ao[sk] + ic, m = ao[sl] + id;
              const double x = out[ia][ib][ic][id];
              eri[((i * n + j) * n + k) * n + m] = x;
              eri[((j * n + i) * n + k) * n + m] = x;
              eri[((i * n + j) * n + m) * n + k] = x;
              eri[((j * n + i) * n + m) * n + k] = x;
              eri[((k * n + m) * n + i) * n + j] = x;
              eri[((m * n + k) * n + i) * n + j] = x;
              eri[((k * n + m) * n + j) * n + i] = x;
              eri[((m * n + k) * n + j) * n + i] = x;
            }
    }
  }
  return 0;
}
