// Host-side basis construction (see evr_setup.hpp).
//
// The 1D Hermite bases follow the definitions of the reference
// (sub_quadra_hermite, Source_ElVibRot/sub_Basis/sub_quadra_herm.f90:208-411;
// Hermite functions Source_Lib/sub_communf90/sub_math/sub_polyortho.f90:830-1164;
// grid-to-grid matrices Source_ElVibRot/sub_Basis/sub_module_basis.f90:2195-2460),
// computed here in extended precision by independent methods:
//  * nodes: eigenvalues of the Hermite Jacobi matrix by Sturm-sequence
//    bisection, weights from the Christoffel function;
//  * Hermite functions by the orthonormal three-term recurrence, derivatives
//    by the ladder relations;
//  * pinv(B) = (B^T B)^{-1} B^T by Cholesky (equal to the reference's
//    eigen-based pseudo-inverse for a full-column-rank B).
#include "evr_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace evr {

using ld = long double;
static const ld kPi = 3.141592653589793238462643383279502884L;

// number of eigenvalues of the n x n Hermite Jacobi matrix below lam
static int sturm_count(int n, ld lam)
{
  int cnt = 0;
  ld d = -lam;  // diagonal entries are 0
  if (d < 0) ++cnt;
  for (int k = 1; k < n; ++k) {
    ld b2 = (ld)k / 2;  // beta_k^2
    if (d == 0) d = 1e-4000L;
    d = -lam - b2 / d;
    if (d < 0) ++cnt;
  }
  return cnt;
}

static void herm_fun_all(int n, ld x, ld *phi)
{
  // orthonormal Hermite functions phi_0..phi_{n-1}
  phi[0] = powl(kPi, -0.25L) * expl(-0.5L * x * x);
  if (n > 1) phi[1] = sqrtl(2.0L) * x * phi[0];
  for (int k = 1; k + 1 < n; ++k)
    phi[k + 1] = sqrtl(2.0L / (k + 1)) * x * phi[k] - sqrtl((ld)k / (k + 1)) * phi[k - 1];
}

void hermite_rule(int nq, std::vector<double> &x, std::vector<double> &wex2)
{
  x.assign(nq, 0.0);
  wex2.assign(nq, 0.0);
  std::vector<ld> phi(nq + 1);
  ld bound = sqrtl(2.0L * nq + 2.0L);
  for (int i = 0; i < nq; ++i) {
    // i-th smallest eigenvalue: smallest lam with count(lam) > i
    ld lo = -bound, hi = bound;
    for (int it = 0; it < 200; ++it) {
      ld mid = 0.5L * (lo + hi);
      if (mid == lo || mid == hi) break;
      if (sturm_count(nq, mid) > i) hi = mid; else lo = mid;
    }
    ld xi = 0.5L * (lo + hi);
    if (nq % 2 == 1 && i == nq / 2) xi = 0.0L;
    herm_fun_all(nq, xi, phi.data());
    ld s = 0;
    for (int k = 0; k < nq; ++k) s += phi[k] * phi[k];
    x[i] = (double)xi;
    wex2[i] = (double)(1.0L / s);
  }
  // exact symmetry of the rule
  for (int i = 0; i < nq / 2; ++i) {
    double a = 0.5 * (x[nq - 1 - i] - x[i]);
    x[i] = -a;
    x[nq - 1 - i] = a;
    double w = 0.5 * (wex2[i] + wex2[nq - 1 - i]);
    wex2[i] = wex2[nq - 1 - i] = w;
  }
}

void hermite_slot(int nb, int nq, Slot1D &s)
{
  s.nb = nb;
  s.nq = nq;
  hermite_rule(nq, s.x, s.wex2);
  s.B.assign((size_t)nq * nb, 0.0);
  s.B1.assign((size_t)nq * nb, 0.0);
  s.B2.assign((size_t)nq * nb, 0.0);
  std::vector<ld> phi(nb + 2);
  for (int q = 0; q < nq; ++q) {
    ld xq = s.x[q];
    herm_fun_all(nb + 1, xq, phi.data());
    for (int n = 0; n < nb; ++n) {
      ld d1 = -sqrtl((ld)(n + 1) / 2) * phi[n + 1] + (n > 0 ? sqrtl((ld)n / 2) * phi[n - 1] : 0.0L);
      ld d2 = (xq * xq - (2.0L * n + 1.0L)) * phi[n];
      s.B[q + (size_t)nq * n] = (double)phi[n];
      s.B1[q + (size_t)nq * n] = (double)d1;
      s.B2[q + (size_t)nq * n] = (double)d2;
    }
  }
  s.W.assign((size_t)nb * nq, 0.0);
  for (int n = 0; n < nb; ++n)
    for (int q = 0; q < nq; ++q) s.W[n + (size_t)nb * q] = s.B[q + (size_t)nq * n] * s.wex2[q];
  // pinv(B) = (B^T B)^{-1} B^T via Cholesky in long double
  std::vector<ld> G((size_t)nb * nb, 0.0L), L((size_t)nb * nb, 0.0L);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      ld a = 0;
      for (int q = 0; q < nq; ++q) a += (ld)s.B[q + (size_t)nq * i] * s.B[q + (size_t)nq * j];
      G[i + (size_t)nb * j] = a;
    }
  for (int j = 0; j < nb; ++j) {
    ld d = G[j + (size_t)nb * j];
    for (int k = 0; k < j; ++k) d -= L[j + (size_t)nb * k] * L[j + (size_t)nb * k];
    ld ljj = sqrtl(d);
    L[j + (size_t)nb * j] = ljj;
    for (int i = j + 1; i < nb; ++i) {
      ld v = G[i + (size_t)nb * j];
      for (int k = 0; k < j; ++k) v -= L[i + (size_t)nb * k] * L[j + (size_t)nb * k];
      L[i + (size_t)nb * j] = v / ljj;
    }
  }
  std::vector<ld> pinv((size_t)nb * nq);
  for (int q = 0; q < nq; ++q) {
    // solve G y = B(q,:)^T
    std::vector<ld> y(nb);
    for (int i = 0; i < nb; ++i) {
      ld v = s.B[q + (size_t)nq * i];
      for (int k = 0; k < i; ++k) v -= L[i + (size_t)nb * k] * y[k];
      y[i] = v / L[i + (size_t)nb * i];
    }
    for (int i = nb - 1; i >= 0; --i) {
      ld v = y[i];
      for (int k = i + 1; k < nb; ++k) v -= L[k + (size_t)nb * i] * y[k];
      y[i] = v / L[i + (size_t)nb * i];
    }
    for (int i = 0; i < nb; ++i) pinv[i + (size_t)nb * q] = y[i];
  }
  s.D1.assign((size_t)nq * nq, 0.0);
  s.D2.assign((size_t)nq * nq, 0.0);
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) {
      ld a1 = 0, a2 = 0;
      for (int k = 0; k < nb; ++k) {
        a1 += (ld)s.B1[i + (size_t)nq * k] * pinv[k + (size_t)nb * j];
        a2 += (ld)s.B2[i + (size_t)nq * k] * pinv[k + (size_t)nb * j];
      }
      s.D1[i + (size_t)nq * j] = (double)a1;
      s.D2[i + (size_t)nq * j] = (double)a2;
    }
}

void keo_factors(const Slot1D &s, std::vector<double> &out)
{
  const int nb = s.nb, nq = s.nq;
  out.assign((size_t)kNumFactors * nb * nb, 0.0);
  // grid-side operator images of the basis: columns of nq x nb matrices
  std::vector<ld> DB((size_t)nq * nb), xB((size_t)nq * nb), DDB((size_t)nq * nb),
      DxB((size_t)nq * nb), xDB((size_t)nq * nb), DxDB((size_t)nq * nb), D2B((size_t)nq * nb);
  auto Dmul = [&](const double *M, const std::vector<ld> &in, std::vector<ld> &o) {
    for (int c = 0; c < nb; ++c)
      for (int i = 0; i < nq; ++i) {
        ld a = 0;
        for (int k = 0; k < nq; ++k) a += (ld)M[i + (size_t)nq * k] * in[k + (size_t)nq * c];
        o[i + (size_t)nq * c] = a;
      }
  };
  std::vector<ld> Bl((size_t)nq * nb);
  for (size_t i = 0; i < Bl.size(); ++i) Bl[i] = s.B[i];
  Dmul(s.D1.data(), Bl, DB);
  for (int c = 0; c < nb; ++c)
    for (int i = 0; i < nq; ++i) xB[i + (size_t)nq * c] = (ld)s.x[i] * Bl[i + (size_t)nq * c];
  Dmul(s.D1.data(), DB, DDB);
  Dmul(s.D1.data(), xB, DxB);
  for (int c = 0; c < nb; ++c)
    for (int i = 0; i < nq; ++i) xDB[i + (size_t)nq * c] = (ld)s.x[i] * DB[i + (size_t)nq * c];
  Dmul(s.D1.data(), xDB, DxDB);
  Dmul(s.D2.data(), Bl, D2B);
  const std::vector<ld> *img[kNumFactors] = {&DB, &xB, &DDB, &DxB, &xDB, &DxDB, &D2B};
  for (int f = 0; f < kNumFactors; ++f)
    for (int r = 0; r < nb; ++r)
      for (int c = 0; c < nb; ++c) {
        ld a = 0;
        for (int q = 0; q < nq; ++q) a += (ld)s.W[r + (size_t)nb * q] * (*img[f])[q + (size_t)nq * c];
        out[(size_t)f * nb * nb + r + (size_t)nb * c] = (double)a;
      }
}

void symmetrize_factors(int nb, std::vector<double> &f)
{
  const size_t nn = (size_t)nb * nb;
  auto M = [&](int fac, int r, int c) -> double & { return f[fac * nn + r + (size_t)nb * c]; };
  std::vector<double> RS(nn);
  for (int r = 0; r < nb; ++r)
    for (int c = 0; c < nb; ++c) RS[r + (size_t)nb * c] = M(kR, r, c) + M(kS, r, c);
  for (int r = 0; r < nb; ++r)
    for (int c = r; c < nb; ++c) {
      auto sym = [&](int fac) {
        double v = 0.5 * (M(fac, r, c) + M(fac, c, r));
        M(fac, r, c) = M(fac, c, r) = v;
      };
      auto anti = [&](int fac) {
        double v = 0.5 * (M(fac, r, c) - M(fac, c, r));
        M(fac, r, c) = v;
        M(fac, c, r) = -v;
      };
      anti(kP);
      sym(kQ);
      sym(kT);
      sym(kU);
      sym(kT2);
      double a = 0.25 * (RS[r + (size_t)nb * c] - RS[c + (size_t)nb * r]);
      M(kR, r, c) = M(kS, r, c) = a;
      M(kR, c, r) = M(kS, c, r) = -a;
    }
}

static void append_slot(BasisTables &t, const Slot1D &s)
{
  t.x.insert(t.x.end(), s.x.begin(), s.x.end());
  t.B.insert(t.B.end(), s.B.begin(), s.B.end());
  t.W.insert(t.W.end(), s.W.begin(), s.W.end());
  t.D1.insert(t.D1.end(), s.D1.begin(), s.D1.end());
  t.D2.insert(t.D2.end(), s.D2.begin(), s.D2.end());
}

std::string build_sg4_hermite(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB,
                              BasisTables &t)
{
  if (D < 1 || D > kMaxD) return "D out of range (1..64)";
  const std::vector<int> a(D, nbA), b(D, nbB), qa(D, nqA), qb(D, nqB);
  return build_sg4_hermite_modes(D, LG, LB, a.data(), b.data(), qa.data(), qb.data(), t);
}

std::string build_sg4_hermite_modes(int D, int LG, int LB, const int *nbA, const int *nbB,
                                    const int *nqA, const int *nqB, BasisTables &t)
{
  if (D < 1 || D > kMaxD) return "D out of range (1..64)";
  if (LG < 0 || LB < 0 || LB > LG) return "need 0 <= LB <= LG";
  for (int d = 0; d < D; ++d)
    if (nbA[d] < 1 || nqA[d] < 1 || nbB[d] < 0 || nqB[d] < 0) return "bad L_TO_n law";
  // per mode d: nb(l) = nbA + nbB l, nq(l) = nqA + nqB l (the basis_nD
  // L_TO_nq_A/B laws of each mode)
  auto nbL = [&](int d, int l) { return nbA[d] + nbB[d] * l; };
  auto nqL = [&](int d, int l) { return nqA[d] + nqB[d] * l; };
  for (int d = 0; d < D; ++d)
    for (int l = 0; l <= LG; ++l)
      if (nqL(d, l) < nbL(d, std::min(l, LB))) return "nq(l) < nb(l): grid too small for the basis";
  t = BasisTables();
  t.D = D;
  t.nlev = LG + 1;
  t.nb.assign((size_t)D * t.nlev, 0);
  t.nq.assign((size_t)D * t.nlev, 0);
  // primitive bases per mode and level: nq(L), nb(min(L,LB))
  for (int d = 0; d < D; ++d)
    for (int l = 0; l <= LG; ++l) {
      Slot1D s;
      hermite_slot(nbL(d, std::min(l, LB)), nqL(d, l), s);
      t.nb[d * t.nlev + l] = s.nb;
      t.nq[d * t.nlev + l] = s.nq;
      append_slot(t, s);
    }
  // Smolyak terms: levels with Lmin <= |l| <= LG, first mode fastest
  const int Lmin = std::max(0, LG - D + 1);
  std::vector<int> l(D, 0);
  auto binom = [](int n, int k) {
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
  };
  while (true) {
    int s = 0;
    for (int d = 0; d < D; ++d) s += l[d];
    if (s >= Lmin && s <= LG) {
      t.term_levels.insert(t.term_levels.end(), l.begin(), l.end());
      int dl = LG - s;
      t.weights.push_back((dl % 2 ? -1.0 : 1.0) * binom(D - 1, dl));
    }
    int d = 0;
    for (; d < D; ++d) {
      ++l[d];
      int s2 = 0;
      for (int e = 0; e < D; ++e) s2 += l[e];
      if (s2 <= LG) break;
      l[d] = 0;
    }
    if (d == D) break;
  }
  t.nterms = (int64_t)t.weights.size();
  // level of 1D function i of mode d: smallest L with nb_d(L) > i
  std::vector<std::vector<int>> lev_of(D);
  for (int d = 0; d < D; ++d) {
    const int nmax = nbL(d, LB);
    lev_of[d].resize(nmax);
    for (int i = 0; i < nmax; ++i) {
      int L = 0;
      while (nbL(d, L) <= i) ++L;
      lev_of[d][i] = L;
    }
  }
  // cnt[d][b]: tuples of modes d..D-1 whose level sum is <= b
  const int Bp = LB + 1;
  std::vector<int64_t> cnt((size_t)(D + 1) * Bp, 0);
  for (int b = 0; b < Bp; ++b) cnt[(size_t)D * Bp + b] = 1;
  for (int d = D - 1; d >= 0; --d)
    for (int b = 0; b < Bp; ++b) {
      int64_t c = 0;
      for (int lv : lev_of[d])
        if (lv <= b) c += cnt[(size_t)(d + 1) * Bp + b - lv];
      cnt[(size_t)d * Bp + b] = c;
    }
  t.nb_packed = cnt[LB];
  // map of every term-local function (first mode fastest) to its packed rank
  // in the last-mode-fastest ordering of {n : sum level(n_d) <= LB}
  std::vector<int64_t> off(t.nterms + 1, 0);
  for (int64_t k = 0; k < t.nterms; ++k) {
    int64_t nbt = 1;
    for (int d = 0; d < D; ++d) nbt *= t.nb[d * t.nlev + t.term_levels[k * D + d]];
    off[k + 1] = off[k] + nbt;
  }
  if (off[t.nterms] > INT32_MAX) return "sum of term basis sizes exceeds int32";
  t.map.assign(off[t.nterms], -1);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t k = 0; k < t.nterms; ++k) {
    int nbd[kMaxD], n[kMaxD];
    for (int d = 0; d < D; ++d) { nbd[d] = t.nb[d * t.nlev + t.term_levels[k * D + d]]; n[d] = 0; }
    const int64_t nbt = off[k + 1] - off[k];
    for (int64_t e = 0; e < nbt; ++e) {
      int bud = LB;
      int64_t rank = 0;
      bool ok = true;
      for (int d = 0; d < D && ok; ++d) {
        const std::vector<int> &lv = lev_of[d];
        for (int m = 0; m < n[d]; ++m)
          if (lv[m] <= bud) rank += cnt[(size_t)(d + 1) * Bp + bud - lv[m]];
        bud -= lv[n[d]];
        if (bud < 0) ok = false;
      }
      t.map[off[k] + e] = ok ? (int)rank : -1;
      for (int d = 0; d < D; ++d) {  // next tuple, first mode fastest
        if (++n[d] < nbd[d]) break;
        n[d] = 0;
      }
    }
  }
  return "";
}

std::string build_dp_hermite(int D, const int *nb, const int *nq, BasisTables &t)
{
  if (D < 1 || D > kMaxD) return "D out of range (1..64)";
  t = BasisTables();
  t.D = D;
  t.nlev = 1;
  t.dp_order = true;
  t.nb.assign(nb, nb + D);
  t.nq.assign(nq, nq + D);
  int64_t nbt = 1;
  for (int d = 0; d < D; ++d) {
    if (nb[d] < 1 || nq[d] < nb[d]) return "need 1 <= nb <= nq per mode";
    Slot1D s;
    hermite_slot(nb[d], nq[d], s);
    append_slot(t, s);
    nbt *= nb[d];
  }
  if (nbt > INT32_MAX) return "basis too large";
  t.nterms = 1;
  t.term_levels.assign(D, 0);
  t.weights.assign(1, 1.0);
  t.nb_packed = nbt;
  t.map.resize(nbt);
  // term-local index e (first mode fastest) -> DP index (last mode fastest)
  std::vector<int> n(D, 0);
  for (int64_t e = 0; e < nbt; ++e) {
    int64_t p = 0;
    for (int d = 0; d < D; ++d) p = p * nb[d] + n[d];
    t.map[e] = (int)p;
    for (int d = 0; d < D; ++d) {
      if (++n[d] < nb[d]) break;
      n[d] = 0;
    }
  }
  return "";
}

}  // namespace evr
