/*
 * evr_oracle.c -- CPU restatement of the reference's H|psi> path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * evr_oracle.h for the parity status of each routine; each function cites
 * the reference file:line it restates.  Abbreviations (files relative to
 * the reference root):
 *   HERM  Source_ElVibRot/sub_Basis/sub_quadra_herm.f90
 *   POLY  Source_Lib/sub_communf90/sub_math/sub_polyortho.f90
 *   BAS   Source_ElVibRot/sub_Basis/sub_module_basis.f90
 *   SPB   Source_ElVibRot/sub_Basis/sub_quadra_SparseBasis.f90
 *   SGT2  Source_ElVibRot/sub_Basis/sub_module_param_SGType2.f90
 *   NDI   Source_Lib/sub_nDindex/sub_module_nDindex.f90
 *   LTON  Source_ElVibRot/sub_Basis/sub_module_Basis_LTO_n.f90
 *   RD    Source_ElVibRot/sub_Basis/sub_read_data.f90
 *   BGSG4 Source_ElVibRot/sub_Basis/sub_Basis_SG4/sub_module_basis_BtoG_GtoB_SG4.f90
 *   BG    Source_ElVibRot/sub_Basis/sub_module_basis_BtoG_GtoB.f90
 *   OPSG4 Source_ElVibRot/sub_Operator/sub_OpPsi_SG4.f90
 *   OPPSI Source_ElVibRot/sub_Operator/sub_OpPsi.f90
 */
#include "evr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_L 3.141592653589793238462643383279502884L

/* ======================================================================
 * 1D Gauss-Hermite rule.  The reference calls hercom and then overwrites
 * nodes and w*exp(x^2) from 50-digit tables (HERM:318-347).  Here the
 * nodes are the roots of the orthonormal Hermite polynomial p_nq found by
 * Newton's method in long double (starting guesses: the classical
 * asymptotic formulas of Stroud & Secrest for Gauss-Hermite nodes), and
 * w_i*exp(x_i^2) = 1/(nq * phi_{nq-1}(x_i)^2) (Christoffel-Darboux) where
 * phi_k = p_k*exp(-x^2/2) is the orthonormal Hermite function.
 * ====================================================================== */
static void herm_orthofun(int n, long double x, long double *pn, long double *pnm1)
{
  /* orthonormal Hermite functions phi_k(x) by the three-term recurrence */
  long double p0 = powl(PI_L, -0.25L) * expl(-0.5L * x * x), p1 = 0.0L;
  long double pm = 0.0L;
  if (n == 0) { *pn = p0; *pnm1 = 0.0L; return; }
  p1 = sqrtl(2.0L) * x * p0;
  pm = p0;
  for (int k = 1; k < n; ++k) {
    long double pk1 = sqrtl(2.0L / (k + 1)) * x * p1 - sqrtl((long double)k / (k + 1)) * pm;
    pm = p1;
    p1 = pk1;
  }
  *pn = p1;
  *pnm1 = pm;
}

int orc_herm_quadrature(int nq, double *x, double *wex2)
{
  if (nq < 1) return -1;
  if (nq == 1) { x[0] = 0.0; wex2[0] = sqrt((double)PI_L); return 0; }
  long double *r = (long double *)malloc(sizeof(long double) * nq);
  int m = (nq + 1) / 2;
  long double z = 0.0L;
  for (int i = 0; i < m; ++i) {
    /* starting guesses for the i-th largest root */
    if (i == 0) z = sqrtl(2.0L * nq + 1) - 1.85575L * powl(2.0L * nq + 1, -1.0L / 6);
    else if (i == 1) z = z - 1.14L * powl((long double)nq, 0.426L) / z;
    else if (i == 2) z = 1.86L * z - 0.86L * r[0];
    else if (i == 3) z = 1.91L * z - 0.91L * r[1];
    else z = 2.0L * z - r[i - 2];
    for (int it = 0; it < 200; ++it) {
      /* Newton on p_nq using p_nq' = sqrt(2 nq) p_{nq-1}; the Hermite
       * function ratio equals the polynomial ratio. */
      long double pn, pnm1;
      herm_orthofun(nq, z, &pn, &pnm1);
      long double dz = pn / (sqrtl(2.0L * nq) * pnm1);
      z -= dz;
      if (fabsl(dz) <= 1e-19L * (1.0L + fabsl(z))) break;
    }
    r[i] = z;
  }
  for (int i = 0; i < m; ++i) {
    long double xi = r[i], pn, pnm1;
    herm_orthofun(nq, xi, &pn, &pnm1);
    long double w = 1.0L / ((long double)nq * pnm1 * pnm1);
    /* reference order: ascending nodes (tables are ascending) */
    x[nq - 1 - i] = (double)xi;
    x[i] = (double)(-xi);
    wex2[nq - 1 - i] = (double)w;
    wex2[i] = (double)w;
  }
  if (nq % 2 == 1) x[nq / 2] = 0.0;
  free(r);
  return 0;
}

/* ======================================================================
 * Hermite functions with derivatives (POLY:830-877 poly_Hermite,
 * POLY:1060-1122 d0d1d2poly_Hermite_exp, POLY:1133-1164 the grid loop).
 * ====================================================================== */
static double poly_hermite(double x, int l)
{
  /* POLY:830-877: unnormalised recurrence, then / sqrt(sqrt(pi) 2^l l!) */
  double norme = sqrt(3.14159265358979323846264338327950288);
  if (l == 0) return 1.0 / sqrt(norme);
  if (l == 1) { norme *= 2.0; return 2.0 * x / sqrt(norme); }
  double pl2 = 1.0, pl1 = 2.0 * x, pl0 = 0.0;
  norme *= 2.0;
  for (int i = 2; i <= l; ++i) {
    norme = norme * 2.0 * (double)i;
    pl0 = 2.0 * (x * pl1 - (double)(i - 1) * pl2);
    pl2 = pl1;
    pl1 = pl0;
  }
  return pl0 / sqrt(norme);
}

int orc_herm_basis(int nb, int nq, const double *x, double *B, double *B1, double *B2)
{
  for (int k = 0; k < nq; ++k) {
    for (int l = 0; l < nb; ++l) {
      double xx = x[k];
      double d0 = poly_hermite(xx, l), d1, d2;
      if (l == 0) { d1 = 0.0; d2 = 0.0; }
      else if (l == 1) { d1 = sqrt(2.0) * poly_hermite(xx, 0); d2 = 0.0; }
      else {
        d1 = sqrt(2.0 * (double)l) * poly_hermite(xx, l - 1);
        d2 = 2.0 * (xx * d1 - d0 * (double)l);
      }
      double pexp = exp(-0.5 * xx * xx);
      d2 = (d2 - 2.0 * xx * d1 + (xx * xx - 1.0) * d0) * pexp;
      d1 = (d1 - xx * d0) * pexp;
      d0 = d0 * pexp;
      B[k + (size_t)nq * l] = d0;
      B1[k + (size_t)nq * l] = d1;
      B2[k + (size_t)nq * l] = d2;
    }
  }
  return 0;
}

/* cyclic Jacobi eigensolver for a small symmetric matrix (column-major);
 * eigenvalues in ev, eigenvectors in the columns of V. */
static void jacobi_eig(int n, double *A, double *ev, double *V)
{
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i + n * j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p + n * q] * A[p + n * q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + n * q];
        if (fabs(apq) < 1e-300) continue;
        double theta = (A[q + n * q] - A[p + n * p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k + n * p], akq = A[k + n * q];
          A[k + n * p] = c * akp - s * akq;
          A[k + n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p + n * k], aqk = A[q + n * k];
          A[p + n * k] = c * apk - s * aqk;
          A[q + n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + n * p], vkq = V[k + n * q];
          V[k + n * p] = c * vkp - s * vkq;
          V[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i + n * i];
}

static void matmul_cm(int m, int k, int n, const double *A, const double *B, double *C)
{
  /* C(m x n) = A(m x k) * B(k x n), column-major */
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int l = 0; l < k; ++l) s += A[i + (size_t)m * l] * B[l + (size_t)k * j];
      C[i + (size_t)m * j] = s;
    }
}

int orc_herm_dngg(int nb, int nq, const double *B, const double *B1,
                  const double *B2, double *D0, double *D1, double *D2)
{
  /* BAS:2195-2460, PseudoInverse=.TRUE., SVD=.FALSE.:
   *   d0bxd0bT = B * B^T (nq x nq); eigenpairs sorted by decreasing
   *   eigenvalue; inverse on the nb largest; pinv = B^T * inv (nb x nq);
   *   dnRGG%dk = dnRGB%dk * pinv. */
  double *BBt = (double *)calloc((size_t)nq * nq, sizeof(double));
  double *ev = (double *)malloc(sizeof(double) * nq);
  double *V = (double *)malloc(sizeof(double) * nq * nq);
  double *inv = (double *)calloc((size_t)nq * nq, sizeof(double));
  double *pinv = (double *)malloc(sizeof(double) * nb * nq);
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) {
      double s = 0.0;
      for (int l = 0; l < nb; ++l) s += B[i + (size_t)nq * l] * B[j + (size_t)nq * l];
      BBt[i + nq * j] = s;
    }
  jacobi_eig(nq, BBt, ev, V);
  int *ord = (int *)malloc(sizeof(int) * nq);
  for (int i = 0; i < nq; ++i) ord[i] = i;
  for (int i = 0; i < nq; ++i)
    for (int j = i + 1; j < nq; ++j)
      if (ev[ord[j]] > ev[ord[i]]) { int t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
  for (int a = 0; a < nb; ++a) {
    int c = ord[a];
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) inv[i + nq * j] += V[i + nq * c] * V[j + nq * c] / ev[c];
  }
  /* pinv = B^T * inv */
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < nq; ++j) {
      double s = 0.0;
      for (int i = 0; i < nq; ++i) s += B[i + (size_t)nq * b] * inv[i + nq * j];
      pinv[b + nb * j] = s;
    }
  if (D0) matmul_cm(nq, nb, nq, B, pinv, D0);
  if (D1) matmul_cm(nq, nb, nq, B1, pinv, D1);
  if (D2) matmul_cm(nq, nb, nq, B2, pinv, D2);
  free(BBt); free(ev); free(V); free(inv); free(pinv); free(ord);
  return 0;
}

/* ======================================================================
 * SG4 basis.
 * ====================================================================== */
typedef struct {
  int nb, nq;
  double *x, *B, *W, *D1, *D2;  /* B: nq x nb, W: nb x nq, D1/D2: nq x nq */
} orc_1d;

struct orc_sg4 {
  int D, LG, LB, Lmin;
  long long nterms, nb_packed, sum_nb, sum_nq, max_nb, max_nq;
  int *levels;          /* nterms x D */
  double *weights;      /* nterms */
  long long *off_nb, *off_nq;
  int *packed;          /* nb_packed x D */
  int *map;             /* sum_nb */
  orc_1d *bas;          /* (LG+1) x D, index l + (LG+1)*d */
  int *nmax;            /* D: number of 1D functions of mode d in the packed basis */
  int nmaxall;          /* max_d nmax[d] */
  int *lev_of_fn;       /* D x nmaxall: level of 1D function i (0-based) of mode d
                           at lev_of_fn[d * nmaxall + i] */
};

static int law(int A, int B, int l) { return A + B * l; }

/* Smolyak term enumeration: type -5 nDindex (NDI:2463-2516): the first
 * index runs fastest, values 0..Lmax, tuples kept when
 * Lmin <= sum(l) <= Lmax (Analysis_nDval_nDindex_type5, NDI:2566-2587). */
static long long enumerate_terms(int D, int Lmin, int Lmax, int *out)
{
  int *l = (int *)calloc(D, sizeof(int));
  long long n = 0;
  for (;;) {
    int s = 0;
    for (int d = 0; d < D; ++d) s += l[d];
    if (s >= Lmin && s <= Lmax) {
      if (out) memcpy(out + (size_t)n * D, l, sizeof(int) * D);
      ++n;
    }
    /* colexicographic increment over the box [0,Lmax]^D; carries when the
     * first index leaves the range or the level sum exceeds Lmax (every
     * later value of that index would exceed it too) */
    int d = 0;
    for (;;) {
      l[d]++;
      int s2 = 0;
      for (int e = 0; e < D; ++e) s2 += l[e];
      if (l[d] <= Lmax && s2 <= Lmax) break;
      l[d] = 0;
      ++d;
      if (d == D) { free(l); return n; }
    }
  }
}

/* Packed basis: type 5 nDindex (NDI:2396-2461): last index fastest,
 * 1D function indices 0..nmax[d]-1 of mode d (level lev[d*stride + i]),
 * kept when the sum of levels is <= LB. */
static long long enumerate_packed(int D, const int *nmax, const int *lev, int stride, int LB,
                                  int *out)
{
  int *n = (int *)calloc(D, sizeof(int));
  long long cnt = 0;
  for (;;) {
    int s = 0;
    for (int d = 0; d < D; ++d) s += lev[d * stride + n[d]];
    if (s <= LB) {
      if (out) memcpy(out + (size_t)cnt * D, n, sizeof(int) * D);
      ++cnt;
    }
    int d = D - 1;
    for (;;) {
      n[d]++;
      int s2 = 0;
      for (int e = 0; e < D; ++e) s2 += lev[e * stride + (n[e] < nmax[e] ? n[e] : 0)];
      if (n[d] < nmax[d] && s2 <= LB) break;
      n[d] = 0;
      --d;
      if (d < 0) { free(n); return cnt; }
    }
  }
}

static double binomial(int n, int k)
{
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
  return r;
}

static void build_1d(orc_1d *b, int nb, int nq)
{
  b->nb = nb; b->nq = nq;
  b->x = (double *)malloc(sizeof(double) * nq);
  double *w = (double *)malloc(sizeof(double) * nq);
  orc_herm_quadrature(nq, b->x, w);
  b->B = (double *)malloc(sizeof(double) * nq * nb);
  double *B1 = (double *)malloc(sizeof(double) * nq * nb);
  double *B2 = (double *)malloc(sizeof(double) * nq * nb);
  orc_herm_basis(nb, nq, b->x, b->B, B1, B2);
  /* dnRBGwrho%d0(ib,:) = dnRGB%d0(:,ib) * wrho (BAS:2121-2124) */
  b->W = (double *)malloc(sizeof(double) * nb * nq);
  for (int i = 0; i < nb; ++i)
    for (int q = 0; q < nq; ++q) b->W[i + nb * q] = b->B[q + nq * i] * w[q];
  b->D1 = (double *)malloc(sizeof(double) * nq * nq);
  b->D2 = (double *)malloc(sizeof(double) * nq * nq);
  orc_herm_dngg(nb, nq, b->B, B1, B2, NULL, b->D1, b->D2);
  free(w); free(B1); free(B2);
}

static void free_1d(orc_1d *b)
{
  free(b->x); free(b->B); free(b->W); free(b->D1); free(b->D2);
}

orc_sg4 *orc_sg4_create(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB)
{
  if (D < 1 || D > 64) return NULL;
  int a[4][64];
  for (int d = 0; d < D; ++d) { a[0][d] = nbA; a[1][d] = nbB; a[2][d] = nqA; a[3][d] = nqB; }
  return orc_sg4_create_modes(D, LG, LB, a[0], a[1], a[2], a[3]);
}

orc_sg4 *orc_sg4_create_modes(int D, int LG, int LB, const int *nbA, const int *nbB,
                              const int *nqA, const int *nqB)
{
  if (D < 1 || D > 64 || LG < 0 || LB < 0 || LB > LG) return NULL;
  for (int d = 0; d < D; ++d)
    if (nbA[d] < 1 || nbB[d] < 0 || nqA[d] < 1 || nqB[d] < 0 ||
        nbA[d] > nqA[d] || nbB[d] > nqB[d]) return NULL;
  orc_sg4 *s = (orc_sg4 *)calloc(1, sizeof(orc_sg4));
  s->D = D; s->LG = LG; s->LB = LB;
  s->Lmin = LG - D + 1 > 0 ? LG - D + 1 : 0;          /* SPB:994 */
  /* primitive bases per level: nq(L), nb(min(L,LB)) (SPB:1059-1064), each
   * mode with its own law (the basis_nD items of the input, RD:511-512) */
  s->bas = (orc_1d *)calloc((size_t)(LG + 1) * D, sizeof(orc_1d));
  for (int d = 0; d < D; ++d)
    for (int l = 0; l <= LG; ++l) {
      int lb = l < LB ? l : LB;
      build_1d(&s->bas[l + (LG + 1) * d], law(nbA[d], nbB[d], lb), law(nqA[d], nqB[d], l));
    }
  /* level of a 1D function: smallest L with nb(L) >= n, n 1-based
   * (Get_L_FROM_Basis_L_TO_n, LTON:508-564; used at SPB:1162) */
  s->nmax = (int *)malloc(sizeof(int) * D);
  s->nmaxall = 0;
  for (int d = 0; d < D; ++d) {
    s->nmax[d] = law(nbA[d], nbB[d], LB);
    if (s->nmax[d] > s->nmaxall) s->nmaxall = s->nmax[d];
  }
  s->lev_of_fn = (int *)calloc((size_t)D * s->nmaxall, sizeof(int));
  for (int d = 0; d < D; ++d)
    for (int i = 0; i < s->nmax[d]; ++i) {
      int L = 0;
      while (law(nbA[d], nbB[d], L) < i + 1) ++L;
      s->lev_of_fn[d * s->nmaxall + i] = L;
    }
  s->nterms = enumerate_terms(D, s->Lmin, LG, NULL);
  s->levels = (int *)malloc(sizeof(int) * s->nterms * D);
  enumerate_terms(D, s->Lmin, LG, s->levels);
  s->weights = (double *)malloc(sizeof(double) * s->nterms);
  s->off_nb = (long long *)malloc(sizeof(long long) * (s->nterms + 1));
  s->off_nq = (long long *)malloc(sizeof(long long) * (s->nterms + 1));
  s->off_nb[0] = s->off_nq[0] = 0;
  for (long long t = 0; t < s->nterms; ++t) {
    const int *l = s->levels + t * D;
    int sum = 0;
    long long nb = 1, nq = 1;
    for (int d = 0; d < D; ++d) {
      sum += l[d];
      nb *= s->bas[l[d] + (LG + 1) * d].nb;
      nq *= s->bas[l[d] + (LG + 1) * d].nq;
    }
    /* calc_Weight_OF_SRep (SGT2:778-793) */
    int dl = LG - sum;
    if (dl < 0 || dl > D - 1) s->weights[t] = 0.0;
    else s->weights[t] = (dl % 2 == 0 ? 1.0 : -1.0) * binomial(D - 1, dl);
    s->off_nb[t + 1] = s->off_nb[t] + nb;
    s->off_nq[t + 1] = s->off_nq[t] + nq;
    if (nb > s->max_nb) s->max_nb = nb;
    if (nq > s->max_nq) s->max_nq = nq;
  }
  s->sum_nb = s->off_nb[s->nterms];
  s->sum_nq = s->off_nq[s->nterms];
  s->nb_packed = enumerate_packed(D, s->nmax, s->lev_of_fn, s->nmaxall, LB, NULL);
  s->packed = (int *)malloc(sizeof(int) * s->nb_packed * D);
  enumerate_packed(D, s->nmax, s->lev_of_fn, s->nmaxall, LB, s->packed);

  /* Map term-local basis index -> packed index (BGSG4:811-892, calc_nDI
   * NDI:2648).  The rank of a tuple in the type-5 order is computed with
   * suffix counts cnt[d][b] = #tuples of dims d..D-1 with level sum <= b. */
  int LBp = LB + 1;
  long long *cnt = (long long *)calloc((size_t)(D + 1) * LBp, sizeof(long long));
  for (int b = 0; b < LBp; ++b) cnt[D * LBp + b] = 1;
  for (int d = D - 1; d >= 0; --d)
    for (int b = 0; b < LBp; ++b) {
      long long c = 0;
      const int *lv = s->lev_of_fn + (size_t)d * s->nmaxall;
      for (int i = 0; i < s->nmax[d]; ++i)
        if (lv[i] <= b) c += cnt[(d + 1) * LBp + b - lv[i]];
      cnt[d * LBp + b] = c;
    }
  s->map = (int *)malloc(sizeof(int) * s->sum_nb);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long t = 0; t < s->nterms; ++t) {
    const int *l = s->levels + t * D;
    int nbd[64], n[64];
    for (int d = 0; d < D; ++d) { nbd[d] = s->bas[l[d] + (LG + 1) * d].nb; n[d] = 0; }
    long long nbt = s->off_nb[t + 1] - s->off_nb[t];
    for (long long e = 0; e < nbt; ++e) {
      /* tuple of term-local index e, first dim fastest (NDI:2613-2628) */
      long long r = e;
      for (int d = 0; d < D; ++d) { n[d] = (int)(r % nbd[d]); r /= nbd[d]; }
      int bud = LB, ok = 1;
      long long rank = 0;
      for (int d = 0; d < D && ok; ++d) {
        const int *lv = s->lev_of_fn + (size_t)d * s->nmaxall;
        for (int m = 0; m < n[d]; ++m)
          if (lv[m] <= bud) rank += cnt[(d + 1) * LBp + bud - lv[m]];
        bud -= lv[n[d]];
        if (bud < 0) ok = 0;
      }
      s->map[s->off_nb[t] + e] = ok ? (int)rank : -1;
    }
  }
  free(cnt);
  return s;
}

void orc_sg4_destroy(orc_sg4 *s)
{
  if (!s) return;
  for (int i = 0; i < (s->LG + 1) * s->D; ++i) free_1d(&s->bas[i]);
  free(s->bas); free(s->levels); free(s->weights); free(s->off_nb); free(s->off_nq);
  free(s->packed); free(s->map); free(s->lev_of_fn); free(s->nmax); free(s);
}

void orc_sg4_sizes(const orc_sg4 *s, long long *z)
{
  z[0] = s->nterms; z[1] = s->nb_packed; z[2] = s->sum_nb; z[3] = s->sum_nq;
  z[4] = s->max_nb; z[5] = s->max_nq; z[6] = s->D; z[7] = s->LG;
}

void orc_sg4_terms(const orc_sg4 *s, int *levels, double *weights, long long *off_nb, long long *off_nq)
{
  if (levels) memcpy(levels, s->levels, sizeof(int) * s->nterms * s->D);
  if (weights) memcpy(weights, s->weights, sizeof(double) * s->nterms);
  if (off_nb) memcpy(off_nb, s->off_nb, sizeof(long long) * (s->nterms + 1));
  if (off_nq) memcpy(off_nq, s->off_nq, sizeof(long long) * (s->nterms + 1));
}

void orc_sg4_packed(const orc_sg4 *s, int *tuples)
{
  memcpy(tuples, s->packed, sizeof(int) * s->nb_packed * s->D);
}

void orc_sg4_map(const orc_sg4 *s, int *map) { memcpy(map, s->map, sizeof(int) * s->sum_nb); }

void orc_sg4_grid(const orc_sg4 *s, double *q)
{
  const int D = s->D;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long t = 0; t < s->nterms; ++t) {
    const int *l = s->levels + t * D;
    long long nqt = s->off_nq[t + 1] - s->off_nq[t];
    for (long long p = 0; p < nqt; ++p) {
      long long r = p;
      for (int d = 0; d < D; ++d) {
        const orc_1d *b = &s->bas[l[d] + (s->LG + 1) * d];
        q[(s->off_nq[t] + p) * D + d] = b->x[r % b->nq];
        r /= b->nq;
      }
    }
  }
}

int orc_sg4_1d(const orc_sg4 *s, int d, int l, int *nb, int *nq, double *x,
               double *B, double *W, double *D1, double *D2)
{
  if (d < 0 || d >= s->D || l < 0 || l > s->LG) return -1;
  const orc_1d *b = &s->bas[l + (s->LG + 1) * d];
  *nb = b->nb; *nq = b->nq;
  if (x) memcpy(x, b->x, sizeof(double) * b->nq);
  if (B) memcpy(B, b->B, sizeof(double) * b->nq * b->nb);
  if (W) memcpy(W, b->W, sizeof(double) * b->nq * b->nb);
  if (D1) memcpy(D1, b->D1, sizeof(double) * b->nq * b->nq);
  if (D2) memcpy(D2, b->D2, sizeof(double) * b->nq * b->nq);
  return 0;
}

/* ======================================================================
 * Term-local transforms on first-dimension-fastest tensors.
 * ====================================================================== */

/* BDP_TO_GDP_OF_SmolyakRep (BGSG4:2385-2496): dims 1..D, at step i the
 * tensor is viewed as (nnq, nb_i, nnb) and RTempG(iq,:,ib) =
 * B_i * RTempB(iq,:,ib) (BGSG4:2457-2471). */
static void term_b2g(int D, const orc_1d *const *b, const double *in, double *out, double *tmp)
{
  long long nnb = 1, nnq = 1;
  for (int d = 0; d < D; ++d) nnb *= b[d]->nb;
  const double *src = in;
  double *bufs[2] = {out, tmp};
  /* pick buffers so that the last step lands in out */
  int cur = (D % 2 == 1) ? 0 : 1;
  for (int d = 0; d < D; ++d) {
    int nb2 = b[d]->nb, nq2 = b[d]->nq;
    nnb /= nb2;
    double *dst = bufs[cur];
    const double *M = b[d]->B; /* nq2 x nb2 */
    /* the same sums (k ascending for every element), the pencil index iq
     * innermost so the loads are contiguous */
    for (long long ib = 0; ib < nnb; ++ib)
      for (int q = 0; q < nq2; ++q) {
        double *o = dst + nnq * (q + (long long)nq2 * ib);
        for (long long iq = 0; iq < nnq; ++iq) o[iq] = 0.0;
        for (int k = 0; k < nb2; ++k) {
          const double m = M[q + (size_t)nq2 * k];
          const double *r = src + nnq * (k + (long long)nb2 * ib);
          for (long long iq = 0; iq < nnq; ++iq) o[iq] += m * r[iq];
        }
      }
    nnq *= nq2;
    src = dst;
    cur ^= 1;
  }
  if (D == 0) out[0] = in[0];
}

/* GDP_TO_BDP_OF_SmolyakRep (BGSG4:2497-2581): dims 1..D, tensor viewed as
 * (nnb, nq_i, nnq), RTempB(ib,:,iq) = W_i * RTempG(ib,:,iq)
 * (BGSG4:2549-2561). */
static void term_g2b(int D, const orc_1d *const *b, const double *in, double *out,
                     double *t1, double *t2)
{
  long long nnq = 1, nnb = 1;
  for (int d = 0; d < D; ++d) nnq *= b[d]->nq;
  const double *src = in;
  double *bufs[2] = {t1, t2};
  int cur = 0;
  for (int d = 0; d < D; ++d) {
    int nb2 = b[d]->nb, nq2 = b[d]->nq;
    nnq /= nq2;
    double *dst = bufs[cur];
    const double *M = b[d]->W; /* nb2 x nq2 */
    for (long long iq = 0; iq < nnq; ++iq)
      for (int k = 0; k < nb2; ++k) {
        double *o = dst + nnb * (k + (long long)nb2 * iq);
        for (long long ib = 0; ib < nnb; ++ib) o[ib] = 0.0;
        for (int q = 0; q < nq2; ++q) {
          const double m = M[k + (size_t)nb2 * q];
          const double *r = src + nnb * (q + (long long)nq2 * iq);
          for (long long ib = 0; ib < nnb; ++ib) o[ib] += m * r[ib];
        }
      }
    nnb *= nb2;
    src = dst;
    cur ^= 1;
  }
  memcpy(out, src, sizeof(double) * nnb);
}

/* DerivOp_TO_RDP_OF_SmolaykRep (BGSG4:2690-2795) for one dimension:
 * RG2(iq1,:,iq3) = M * RG2(iq1,:,iq3) (BGSG4:2764-2768), in place via tmp. */
static void term_deriv(int D, const orc_1d *const *b, int dim, const double *M,
                       double *R, double *tmp)
{
  long long nnq1 = 1, nnq3 = 1;
  for (int d = 0; d < dim; ++d) nnq1 *= b[d]->nq;
  for (int d = dim + 1; d < D; ++d) nnq3 *= b[d]->nq;
  int nq2 = b[dim]->nq;
  for (long long i3 = 0; i3 < nnq3; ++i3)
    for (int q = 0; q < nq2; ++q) {
      double *o = tmp + nnq1 * (q + (long long)nq2 * i3);
      for (long long i1 = 0; i1 < nnq1; ++i1) o[i1] = 0.0;
      for (int k = 0; k < nq2; ++k) {
        const double m = M[q + (size_t)nq2 * k];
        const double *r = R + nnq1 * (k + (long long)nq2 * i3);
        for (long long i1 = 0; i1 < nnq1; ++i1) o[i1] += m * r[i1];
      }
    }
  long long n = nnq1 * nq2 * nnq3;
  memcpy(R, tmp, sizeof(double) * n);
}

/* grid combine of one term, restating OPSG4:1448-1650 (nb0 = 1).
 * psiG is overwritten by (H psi)_G.  qpt: grid coordinates (nq x D). */
static void term_grid_op(int D, const orc_1d *const *b, const orc_op *op, long long nq,
                         const double *V, const double *J, const double *s, const double *Gq,
                         long long gstride, const double *qpt, const double *qmm, double *psiG,
                         double *w1, double *w2, double *PsiRj, double *PsiRi, double *OpPsiR,
                         double *PsiRall, const double *F2q, const double *F1q)
{
  if (op->type == 0) {
    /* OPSG4:1448-1488 */
    for (long long p = 0; p < nq; ++p) psiG[p] = (V ? V[p] : 0.0) * psiG[p];
    return;
  }
  if (op->type == 1) {
    /* OPSG4:1490-1546: OpPsi = sum over operator terms of
     * GridOp(term) * DerivOp(term) psi.  Terms: (0,0) -> V; (i,i) -> F2_ii
     * with dnRGG%d2; (i,j) i/=j -> F2_ij with d1 on dim i and d1 on dim j;
     * (i,0) -> F1_i with d1. */
    /* F2q/F1q: GridOp(:, iterm) per point of the term grid (F2q, F1q point
     * at the term's first grid point, column stride gstride) */
    memset(OpPsiR, 0, sizeof(double) * nq);
    if (V) for (long long p = 0; p < nq; ++p) OpPsiR[p] += V[p] * psiG[p];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        const double *fq = op->F2q ? F2q + gstride * (i + (long long)D * j) : NULL;
        double f = fq ? 1.0 : op->F2[i + D * j];
        if (f == 0.0) continue;
        memcpy(w1, psiG, sizeof(double) * nq);
        if (i == j) term_deriv(D, b, i, b[i]->D2, w1, w2);
        else { term_deriv(D, b, i, b[i]->D1, w1, w2); term_deriv(D, b, j, b[j]->D1, w1, w2); }
        if (fq) for (long long p = 0; p < nq; ++p) OpPsiR[p] += fq[p] * w1[p];
        else for (long long p = 0; p < nq; ++p) OpPsiR[p] += f * w1[p];
      }
    if (op->F1 || F1q)
      for (int i = 0; i < D; ++i) {
        const double *fq = F1q ? F1q + gstride * i : NULL;
        double f = fq ? 1.0 : op->F1[i];
        if (f == 0.0) continue;
        memcpy(w1, psiG, sizeof(double) * nq);
        term_deriv(D, b, i, b[i]->D1, w1, w2);
        if (fq) for (long long p = 0; p < nq; ++p) OpPsiR[p] += fq[p] * w1[p];
        else for (long long p = 0; p < nq; ++p) OpPsiR[p] += f * w1[p];
      }
    memcpy(psiG, OpPsiR, sizeof(double) * nq);
    return;
  }
  /* type 10, OPSG4:1566-1632 */
  double *VPsi = w1;
  for (long long p = 0; p < nq; ++p) VPsi[p] = (V ? V[p] : 0.0) * psiG[p];
  if (s) for (long long p = 0; p < nq; ++p) psiG[p] *= s[p];
  for (int j = 0; j < D; ++j) {
    memcpy(PsiRj + (size_t)nq * j, psiG, sizeof(double) * nq);
    term_deriv(D, b, j, b[j]->D1, PsiRj + (size_t)nq * j, w2);
  }
  memset(OpPsiR, 0, sizeof(double) * nq);
  if (!Gq && op->Ga) {
    /* G_ji(q) = G0_ji + sum_k Ga_jik q_k, evaluated per point as the
     * reference's get_d0GG does (OPSG4:2724), for a block of points at a time
     * (a small table GGiq(block, j, i)), then PsiRi = J sum_j GGiq(:,j,i)
     * PsiRj(:,j) for every i */
    enum { BLK = 64 };
    double *Gb = (double *)malloc(sizeof(double) * (size_t)D * D * BLK);
    for (long long p0 = 0; p0 < nq; p0 += BLK) {
      const int np = (int)(nq - p0 < BLK ? nq - p0 : BLK);
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          double *g = Gb + (size_t)BLK * (j + D * i);
          const double g0 = op->G0[j + D * i];
          for (int pp = 0; pp < np; ++pp) g[pp] = g0;
          for (int k = 0; k < D; ++k) {
            const double a = op->Ga[j + D * (i + D * k)];
            const double *qk = qmm + (size_t)nq * k + p0;
            for (int pp = 0; pp < np; ++pp) g[pp] += a * qk[pp];
          }
        }
      for (int i = 0; i < D; ++i) {
        double *o = PsiRall + (size_t)nq * i + p0;
        for (int pp = 0; pp < np; ++pp) o[pp] = 0.0;
        for (int j = 0; j < D; ++j) {
          const double *g = Gb + (size_t)BLK * (j + D * i);
          const double *pj = PsiRj + (size_t)nq * j + p0;
          for (int pp = 0; pp < np; ++pp) o[pp] += g[pp] * pj[pp];
        }
        if (J) for (int pp = 0; pp < np; ++pp) o[pp] *= J[p0 + pp];
      }
    }
    free(Gb);
    for (int i = 0; i < D; ++i) {
      term_deriv(D, b, i, b[i]->D1, PsiRall + (size_t)nq * i, w2);
      for (long long p = 0; p < nq; ++p) OpPsiR[p] += PsiRall[(size_t)nq * i + p];
    }
    for (long long p = 0; p < nq; ++p)
      psiG[p] = -0.5 * OpPsiR[p] / ((J ? J[p] : 1.0) * (s ? s[p] : 1.0)) + VPsi[p];
    return;
  }
  for (int i = 0; i < D; ++i) {
    /* PsiRi = sum_j GGiq(:,j,i) * PsiRj(:,j); GGiq(:,j,i) = G_ji(q), as
     * whole-grid vector operations (OPSG4:1590-1610) */
    if (Gq) {
      for (long long p = 0; p < nq; ++p) PsiRi[p] = 0.0;
      for (int j = 0; j < D; ++j) {
        const double *g = Gq + gstride * (j + (long long)D * i);   /* GGiq(:, j, i) */
        const double *pj = PsiRj + (size_t)nq * j;
        for (long long p = 0; p < nq; ++p) PsiRi[p] += g[p] * pj[p];
      }
      if (J) for (long long p = 0; p < nq; ++p) PsiRi[p] *= J[p];
    } else {
      for (long long p = 0; p < nq; ++p) {
        double acc = 0.0;
        for (int j = 0; j < D; ++j) {
          double g = op->G0[j + D * i];
          if (op->Ga)
            for (int k = 0; k < D; ++k) g += op->Ga[j + D * (i + D * k)] * qpt[p * D + k];
          acc += g * PsiRj[(size_t)nq * j + p];
        }
        PsiRi[p] = acc * (J ? J[p] : 1.0);
      }
    }
    term_deriv(D, b, i, b[i]->D1, PsiRi, w2);
    for (long long p = 0; p < nq; ++p) OpPsiR[p] += PsiRi[p];
  }
  for (long long p = 0; p < nq; ++p)
    psiG[p] = -0.5 * OpPsiR[p] / ((J ? J[p] : 1.0) * (s ? s[p] : 1.0)) + VPsi[p];
}

static int sg4_oppsi_list(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                          double *hpsi, int nthreads, const long long *list, long long nlist);

/* The potential at every point of one term grid, as the reference fills its
 * grid cache point by point from the potential routine (OPSG4:2754-2767):
 * coefficients v2[D], c[(i*D+j)*D+k] (i<=j<=k), d[i*D+j] (i<j), e[D]. */
static void poly_potential(int D, const double *cf, long long nq, const double *qpt, double *V)
{
  const double *v2 = cf, *c3 = cf + D, *d22 = c3 + (size_t)D * D * D, *e4 = d22 + (size_t)D * D;
  for (long long p = 0; p < nq; ++p) {
    const double *q = qpt + p * D;
    double val = 0.0;
    for (int i = 0; i < D; ++i) {
      double q2 = q[i] * q[i];
      val += (v2[i] + e4[i] * q2) * q2;
      for (int j = i + 1; j < D; ++j) val += d22[i * D + j] * q2 * q[j] * q[j];
      for (int j = i; j < D; ++j) {
        double qij = q[i] * q[j], acc = 0.0;
        for (int k = j; k < D; ++k) acc += c3[((size_t)i * D + j) * D + k] * q[k];
        val += qij * acc;
      }
    }
    V[p] = val;
  }
}

int orc_sg4_oppsi_range(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                        double *hpsi, int nthreads, long long t0, long long t1)
{
  if (t0 < 0) t0 = 0;
  if (t1 > s->nterms) t1 = s->nterms;
  if (t1 < t0) t1 = t0;
  long long *list = (long long *)malloc(sizeof(long long) * (t1 - t0 + 1));
  for (long long t = t0; t < t1; ++t) list[t - t0] = t;
  int rc = sg4_oppsi_list(s, op, nvec, psi, hpsi, nthreads, list, t1 - t0);
  free(list);
  return rc;
}

int orc_sg4_oppsi_terms(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                        double *hpsi, int nthreads, const long long *terms, long long nlist)
{
  for (long long i = 0; i < nlist; ++i)
    if (terms[i] < 0 || terms[i] >= s->nterms) return -3;
  return sg4_oppsi_list(s, op, nvec, psi, hpsi, nthreads, terms, nlist);
}

static int sg4_oppsi_list(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                          double *hpsi, int nthreads, const long long *list, long long nlist)
{
  const int D = s->D;
  if (nvec < 1) return -1;                                  /* OPSG4:737-750 */
  if (op->type != 0 && op->type != 1 && op->type != 10) return -2;
  /* OpPsi = 0 (OPSG4:761-767) */
  memset(hpsi, 0, sizeof(double) * s->nb_packed * nvec);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    long long mq = s->max_nq, mb = s->max_nb;
    double *G = (double *)malloc(sizeof(double) * mq);
    double *T = (double *)malloc(sizeof(double) * mq);
    double *w1 = (double *)malloc(sizeof(double) * mq);
    double *w2 = (double *)malloc(sizeof(double) * mq);
    double *PsiRj = (double *)malloc(sizeof(double) * mq * D);
    double *PsiRi = (double *)malloc(sizeof(double) * mq);
    double *OpPsiR = (double *)malloc(sizeof(double) * mq);
    double *qpt = (double *)malloc(sizeof(double) * mq * D);
    double *qmm = (double *)malloc(sizeof(double) * mq * D);   /* mode-major copy */
    double *Vt = (double *)malloc(sizeof(double) * mq);
    /* PsiRi of every i at once (D x nq) for the G0 + Ga q path */
    double *PsiRall = (op->type == 10 && op->Ga && !op->Gq)
                          ? (double *)malloc(sizeof(double) * mq * D) : NULL;
    double *x = (double *)malloc(sizeof(double) * mb);
    double *y = (double *)malloc(sizeof(double) * mb);
    const orc_1d *b[64];
    /* PSG4_omp=1: static schedule over terms (OPSG4:822-865) */
#pragma omp for schedule(static)
    for (long long it = 0; it < nlist; ++it) {
      const long long t = list[it];
      const int *l = s->levels + t * D;
      for (int d = 0; d < D; ++d) b[d] = &s->bas[l[d] + (s->LG + 1) * d];
      long long nbt = s->off_nb[t + 1] - s->off_nb[t];
      long long nqt = s->off_nq[t + 1] - s->off_nq[t];
      const int *map = s->map + s->off_nb[t];
      if ((op->type == 10 && op->Ga && !op->Gq) || (!op->V && op->Vpoly)) {
        for (long long p = 0; p < nqt; ++p) {
          long long r = p;
          for (int d = 0; d < D; ++d) {
            qpt[p * D + d] = qmm[d * nqt + p] = b[d]->x[r % b[d]->nq];
            r /= b[d]->nq;
          }
        }
      }
      const double *Gterm = op->Gq ? op->Gq + s->off_nq[t] : NULL;
      const double *V = op->V ? op->V + s->off_nq[t] : NULL;
      if (!V && op->Vpoly) {
        poly_potential(D, op->Vpoly, nqt, qpt, Vt);
        V = Vt;
      }
      const double *J = op->Jac ? op->Jac + s->off_nq[t] : NULL;
      const double *sr = op->sqRhoOVERJac ? op->sqRhoOVERJac + s->off_nq[t] : NULL;
      for (int v = 0; v < nvec; ++v) {
        const double *pv = psi + (size_t)s->nb_packed * v;
        /* gather (BGSG4:1220-1238) */
        for (long long e = 0; e < nbt; ++e) x[e] = map[e] >= 0 ? pv[map[e]] : 0.0;
        term_b2g(D, b, x, G, T);
        term_grid_op(D, b, op, nqt, V, J, sr, Gterm, s->sum_nq, qpt, qmm, G, w1, w2, PsiRj, PsiRi,
                     OpPsiR, PsiRall, op->F2q ? op->F2q + s->off_nq[t] : NULL,
                     op->F1q ? op->F1q + s->off_nq[t] : NULL);
        term_g2b(D, b, G, y, T, w2);
        /* weighted scatter (BGSG4:1273-1287) with OMP ATOMIC */
        double wt = s->weights[t];
        double *hv = hpsi + (size_t)s->nb_packed * v;
        for (long long e = 0; e < nbt; ++e)
          if (map[e] >= 0) {
#pragma omp atomic
            hv[map[e]] += wt * y[e];
          }
      }
    }
    free(G); free(T); free(w1); free(w2); free(PsiRj); free(PsiRi); free(OpPsiR);
    free(qpt); free(qmm); free(Vt); free(PsiRall); free(x); free(y);
  }
  return 0;
}

int orc_sg4_oppsi(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                  double *hpsi, int nthreads)
{
  return orc_sg4_oppsi_range(s, op, nvec, psi, hpsi, nthreads, 0, s->nterms);
}

int orc_sg4_b2g(const orc_sg4 *s, const double *psiB, double *psiG)
{
  const int D = s->D;
#pragma omp parallel
  {
    double *x = (double *)malloc(sizeof(double) * s->max_nb);
    double *T = (double *)malloc(sizeof(double) * s->max_nq);
    const orc_1d *b[64];
#pragma omp for schedule(static)
    for (long long t = 0; t < s->nterms; ++t) {
      const int *l = s->levels + t * D;
      for (int d = 0; d < D; ++d) b[d] = &s->bas[l[d] + (s->LG + 1) * d];
      long long nbt = s->off_nb[t + 1] - s->off_nb[t];
      const int *map = s->map + s->off_nb[t];
      for (long long e = 0; e < nbt; ++e) x[e] = map[e] >= 0 ? psiB[map[e]] : 0.0;
      term_b2g(D, b, x, psiG + s->off_nq[t], T);
    }
    free(x); free(T);
  }
  return 0;
}

int orc_sg4_g2b(const orc_sg4 *s, const double *psiG, double *psiB)
{
  const int D = s->D;
  memset(psiB, 0, sizeof(double) * s->nb_packed);
  double *y = (double *)malloc(sizeof(double) * s->max_nb);
  double *T = (double *)malloc(sizeof(double) * s->max_nq);
  double *T2 = (double *)malloc(sizeof(double) * s->max_nq);
  const orc_1d *b[64];
  for (long long t = 0; t < s->nterms; ++t) {
    if (fabs(s->weights[t]) < 1e-6) continue;               /* BGSG4:1004 */
    const int *l = s->levels + t * D;
    for (int d = 0; d < D; ++d) b[d] = &s->bas[l[d] + (s->LG + 1) * d];
    long long nbt = s->off_nb[t + 1] - s->off_nb[t];
    const int *map = s->map + s->off_nb[t];
    term_g2b(D, b, psiG + s->off_nq[t], y, T, T2);
    for (long long e = 0; e < nbt; ++e)
      if (map[e] >= 0) psiB[map[e]] += s->weights[t] * y[e];
  }
  free(y); free(T); free(T2);
  return 0;
}

/* DerivOp_TO_RDP_OF_SmolaykRep (BGSG4:2690-2795) on every term grid of the
 * concatenated SG4 grid: modes i, j 0-based, < 0 for none; i == j: dnRGG%d2,
 * i /= j: d1 along i then along j (BGSG4:2762-2768 per derivative). */
int orc_sg4_deriv(const orc_sg4 *s, int i, int j, const double *psiG, double *out)
{
  const int D = s->D;
  if (i >= D || j >= D) return -1;
  memcpy(out, psiG, sizeof(double) * s->sum_nq);
  double *T = (double *)malloc(sizeof(double) * s->max_nq);
  const orc_1d *b[64];
  for (long long t = 0; t < s->nterms; ++t) {
    const int *l = s->levels + t * D;
    for (int d = 0; d < D; ++d) b[d] = &s->bas[l[d] + (s->LG + 1) * d];
    double *R = out + s->off_nq[t];
    if (i >= 0 && i == j) {
      term_deriv(D, b, i, b[i]->D2, R, T);
    } else {
      if (i >= 0) term_deriv(D, b, i, b[i]->D1, R, T);
      if (j >= 0) term_deriv(D, b, j, b[j]->D1, R, T);
    }
  }
  free(T);
  return 0;
}

/* ======================================================================
 * Direct-product basis.  Basis and grid vectors are last-dimension-fastest
 * (BG:730-798, BG:134-221).
 * ====================================================================== */
struct orc_dp {
  int D;
  long long nb_tot, nq_tot;
  orc_1d *bas;
};

orc_dp *orc_dp_create(int D, const int *nb, const int *nq)
{
  orc_dp *p = (orc_dp *)calloc(1, sizeof(orc_dp));
  p->D = D;
  p->bas = (orc_1d *)calloc(D, sizeof(orc_1d));
  p->nb_tot = p->nq_tot = 1;
  for (int d = 0; d < D; ++d) {
    build_1d(&p->bas[d], nb[d], nq[d]);
    p->nb_tot *= nb[d];
    p->nq_tot *= nq[d];
  }
  return p;
}

void orc_dp_destroy(orc_dp *p)
{
  if (!p) return;
  for (int d = 0; d < p->D; ++d) free_1d(&p->bas[d]);
  free(p->bas); free(p);
}

void orc_dp_sizes(const orc_dp *p, long long *z) { z[0] = p->nb_tot; z[1] = p->nq_tot; }

void orc_dp_grid(const orc_dp *p, double *q)
{
  for (long long i = 0; i < p->nq_tot; ++i) {
    long long r = i;
    for (int d = p->D - 1; d >= 0; --d) {
      q[i * p->D + d] = p->bas[d].x[r % p->bas[d].nq];
      r /= p->bas[d].nq;
    }
  }
}

/* RecRvecB_TO_RVecG, direct product (BG:730-798): ibasis = D..1, tensor
 * viewed as (nnq, nb_i, nnb) with nnq the already-transformed (faster)
 * grid dims. */
static void dp_b2g(const orc_dp *p, const double *in, double *out)
{
  const int D = p->D;
  long long nnq = 1, nnb = p->nb_tot;
  double *cur = (double *)malloc(sizeof(double) * (p->nq_tot > p->nb_tot ? p->nq_tot : p->nb_tot) * 2);
  double *nxt = cur + (p->nq_tot > p->nb_tot ? p->nq_tot : p->nb_tot);
  memcpy(cur, in, sizeof(double) * p->nb_tot);
  for (int d = D - 1; d >= 0; --d) {
    int nb2 = p->bas[d].nb, nq2 = p->bas[d].nq;
    nnb /= nb2;
    /* RTempB(nnq, nbb) with nbb = nb2 x nnb (nb2 fastest within a block) */
    for (long long ib = 0; ib < nnb; ++ib)
      for (long long iq = 0; iq < nnq; ++iq)
        for (int q = 0; q < nq2; ++q) {
          double s = 0.0;
          for (int k = 0; k < nb2; ++k)
            s += p->bas[d].B[q + (size_t)nq2 * k] * cur[iq + nnq * (k + (long long)nb2 * ib)];
          nxt[iq + nnq * (q + (long long)nq2 * ib)] = s;
        }
    nnq *= nq2;
    double *t = cur; cur = nxt; nxt = t;
  }
  memcpy(out, cur, sizeof(double) * p->nq_tot);
  free(cur < nxt ? cur : nxt);
}

/* RecRVecG_TO_RvecB, direct product (BG:134-221): ibasis = 1..D, the
 * slowest remaining grid dim is contracted; the basis index of the
 * contracted dims accumulates with the later dim fastest. */
static void dp_g2b(const orc_dp *p, const double *in, double *out)
{
  const int D = p->D;
  long long nnq = p->nq_tot, nnb = 1;
  long long mx = p->nq_tot > p->nb_tot ? p->nq_tot : p->nb_tot;
  double *buf = (double *)malloc(sizeof(double) * mx * 2);
  double *cur = buf, *nxt = buf + mx;
  memcpy(cur, in, sizeof(double) * p->nq_tot);
  for (int d = 0; d < D; ++d) {
    int nb2 = p->bas[d].nb, nq2 = p->bas[d].nq;
    nnq /= nq2;
    /* RTempG(nnq, nq2, nnb) -> RTempB(nnq, nb2 x nnb) */
    for (long long ib = 0; ib < nnb; ++ib)
      for (long long iq = 0; iq < nnq; ++iq)
        for (int k = 0; k < nb2; ++k) {
          double s = 0.0;
          for (int q = 0; q < nq2; ++q)
            s += p->bas[d].W[k + (size_t)nb2 * q] * cur[iq + nnq * (q + (long long)nq2 * ib)];
          nxt[iq + nnq * (k + (long long)nb2 * ib)] = s;
        }
    nnb *= nb2;
    double *t = cur; cur = nxt; nxt = t;
  }
  memcpy(out, cur, sizeof(double) * p->nb_tot);
  free(buf);
}

/* DerivOp_TO_RVecG, direct product (BG:1291-1352): ibasis = D..1 with
 * the grid viewed as (nnq1, nq_i, nnq3), nnq1 = product of faster dims. */
static void dp_deriv(const orc_dp *p, int dim, const double *M, double *R, double *tmp)
{
  long long nnq1 = 1, nnq3 = 1;
  for (int d = dim + 1; d < p->D; ++d) nnq1 *= p->bas[d].nq;
  for (int d = 0; d < dim; ++d) nnq3 *= p->bas[d].nq;
  int nq2 = p->bas[dim].nq;
  for (long long i3 = 0; i3 < nnq3; ++i3)
    for (int q = 0; q < nq2; ++q) {
      double *o = tmp + nnq1 * (q + (long long)nq2 * i3);
      for (long long i1 = 0; i1 < nnq1; ++i1) o[i1] = 0.0;
      for (int k = 0; k < nq2; ++k) {
        const double m = M[q + (size_t)nq2 * k];
        const double *r = R + nnq1 * (k + (long long)nq2 * i3);
        for (long long i1 = 0; i1 < nnq1; ++i1) o[i1] += m * r[i1];
      }
    }
  memcpy(R, tmp, sizeof(double) * p->nq_tot);
}

int orc_dp_b2g(const orc_dp *p, const double *psiB, double *psiG) { dp_b2g(p, psiB, psiG); return 0; }
int orc_dp_g2b(const orc_dp *p, const double *psiG, double *psiB) { dp_g2b(p, psiG, psiB); return 0; }

/* DerivOp_TO_RVecG on the DP grid (BG:1291-1352), modes as orc_sg4_deriv. */
int orc_dp_deriv(const orc_dp *p, int i, int j, const double *psiG, double *out)
{
  if (i >= p->D || j >= p->D) return -1;
  memcpy(out, psiG, sizeof(double) * p->nq_tot);
  double *tmp = (double *)malloc(sizeof(double) * p->nq_tot);
  if (i >= 0 && i == j) {
    dp_deriv(p, i, p->bas[i].D2, out, tmp);
  } else {
    if (i >= 0) dp_deriv(p, i, p->bas[i].D1, out, tmp);
    if (j >= 0) dp_deriv(p, j, p->bas[j].D1, out, tmp);
  }
  free(tmp);
  return 0;
}

int orc_dp_oppsi(const orc_dp *p, const orc_op *op, int nvec, const double *psi, double *hpsi)
{
  const int D = p->D;
  const long long nq = p->nq_tot;
  double *G = (double *)malloc(sizeof(double) * nq);
  double *Op = (double *)malloc(sizeof(double) * nq);
  double *tmp = (double *)malloc(sizeof(double) * nq);
  double *w1 = (double *)malloc(sizeof(double) * nq);
  double *derRGi = (double *)malloc(sizeof(double) * nq * D);
  double *qpt = (double *)malloc(sizeof(double) * nq * D);
  orc_dp_grid(p, qpt);
  for (int v = 0; v < nvec; ++v) {
    dp_b2g(p, psi + (size_t)p->nb_tot * v, G);
    if (op->type == 0) {
      for (long long i = 0; i < nq; ++i) Op[i] = (op->V ? op->V[i] : 0.0) * G[i];
    } else if (op->type == 1) {
      memset(Op, 0, sizeof(double) * nq);
      if (op->V) for (long long i = 0; i < nq; ++i) Op[i] += op->V[i] * G[i];
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          double f = op->F2[i + D * j];
          if (f == 0.0) continue;
          memcpy(w1, G, sizeof(double) * nq);
          if (i == j) dp_deriv(p, i, p->bas[i].D2, w1, tmp);
          else { dp_deriv(p, i, p->bas[i].D1, w1, tmp); dp_deriv(p, j, p->bas[j].D1, w1, tmp); }
          for (long long k = 0; k < nq; ++k) Op[k] += f * w1[k];
        }
      if (op->F1)
        for (int i = 0; i < D; ++i) {
          if (op->F1[i] == 0.0) continue;
          memcpy(w1, G, sizeof(double) * nq);
          dp_deriv(p, i, p->bas[i].D1, w1, tmp);
          for (long long k = 0; k < nq; ++k) Op[k] += op->F1[i] * w1[k];
        }
    } else {
      /* OPPSI:1797-2070 */
      const double *J = op->Jac, *s = op->sqRhoOVERJac;
      if (s) for (long long i = 0; i < nq; ++i) G[i] *= s[i];          /* inv=.FALSE. */
      for (long long i = 0; i < nq; ++i)                                 /* OPPSI:1921-1936 */
        Op[i] = -2.0 * (J ? J[i] : 1.0) * (op->V ? op->V[i] : 0.0) * G[i];
      for (int i = 0; i < D; ++i) {                                      /* OPPSI:1954-1967 */
        memcpy(derRGi + (size_t)nq * i, G, sizeof(double) * nq);
        dp_deriv(p, i, p->bas[i].D1, derRGi + (size_t)nq * i, tmp);
      }
      for (int j = 0; j < D; ++j) {                                      /* OPPSI:1985-2046 */
        for (long long k = 0; k < nq; ++k) {
          double acc = 0.0;
          for (int i = 0; i < D; ++i) {
            double g = op->G0[i + D * j];
            if (op->Ga)
              for (int m = 0; m < D; ++m) g += op->Ga[i + D * (j + D * m)] * qpt[k * D + m];
            acc += derRGi[(size_t)nq * i + k] * g;
          }
          w1[k] = acc * (J ? J[k] : 1.0);
        }
        dp_deriv(p, j, p->bas[j].D1, w1, tmp);
        for (long long k = 0; k < nq; ++k) Op[k] += w1[k];
      }
      for (long long k = 0; k < nq; ++k) {                               /* OPPSI:2047-2049 */
        Op[k] = -0.5 * Op[k] / (J ? J[k] : 1.0);
        if (s) Op[k] /= s[k];
      }
    }
    dp_g2b(p, Op, hpsi + (size_t)p->nb_tot * v);
  }
  free(G); free(Op); free(tmp); free(w1); free(derRGi); free(qpt);
  return 0;
}
