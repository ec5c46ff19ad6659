// Host-side construction of the vibrational basis: 1D Hermite primitive
// bases, Smolyak term tables, packed basis and the term -> packed map, plus
// the per-mode basis-space factors of the kinetic operator.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace evr {

constexpr int kMaxD = 64;

// One 1D primitive basis (a "slot": one mode at one Smolyak level).
// Matrices are column-major.
struct Slot1D {
  int nb = 0, nq = 0;
  std::vector<double> x;        // nq grid points
  std::vector<double> wex2;     // nq quadrature weights times exp(x^2)
  std::vector<double> B;        // nq x nb  basis on grid
  std::vector<double> B1, B2;   // nq x nb  first / second derivatives
  std::vector<double> W;        // nb x nq  B^T diag(wex2)
  std::vector<double> D1, D2;   // nq x nq  grid -> grid derivative matrices
};

// Gauss-Hermite nodes (ascending) and w*exp(x^2) for nq points.
void hermite_rule(int nq, std::vector<double> &x, std::vector<double> &wex2);
// Fill every matrix of a Hermite slot.
void hermite_slot(int nb, int nq, Slot1D &s);

// Basis-space factors of the grid kinetic operator for one slot, all nb x nb
// column-major, computed in extended precision from the double matrices:
//   P = W D1 B, Q = W diag(x) B, T = W D1 D1 B, R = W D1 diag(x) B,
//   S = W diag(x) D1 B, U = W D1 diag(x) D1 B, T2 = W D2 B,
//   Qg = diag-free grid coordinate factor W diag(x) B (same as Q).
enum KeoFactor { kP = 0, kQ, kT, kR, kS, kU, kT2, kNumFactors };
void keo_factors(const Slot1D &s, std::vector<double> &out /* kNumFactors*nb*nb */);
// Factors of the symmetric part (K + K^T)/2 of the kinetic operator:
// P -> antisym(P), Q -> sym(Q), T -> sym(T), U -> sym(U), T2 -> sym(T2),
// R = S -> antisym(R + S)/2 (the k=i and k=j terms of a symmetric a_ijk pair
// up as (R + S)_i x P_j).
void symmetrize_factors(int nb, std::vector<double> &f);

// Fully built basis (owned arrays behind evr_basis_desc).
struct BasisTables {
  int D = 0, nlev = 0;
  int64_t nterms = 0, nb_packed = 0;
  bool dp_order = false;                 // grid order of a direct-product basis
  std::vector<int> term_levels;          // nterms x D
  std::vector<double> weights;           // nterms
  std::vector<int> nb, nq;               // D x nlev
  std::vector<double> x, B, W, D1, D2;   // concatenated per slot
  std::vector<int> map;                  // sum nb_t
};

std::string build_sg4_hermite(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB,
                              BasisTables &out);
// The same with per-mode laws nb_d(l) = nbA[d] + nbB[d] l, nq_d(l) = nqA[d] + nqB[d] l.
std::string build_sg4_hermite_modes(int D, int LG, int LB, const int *nbA, const int *nbB,
                                    const int *nqA, const int *nqB, BasisTables &t);
std::string build_dp_hermite(int D, const int *nb, const int *nq, BasisTables &out);

}  // namespace evr
