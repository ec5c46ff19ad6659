/*
 * evr_oracle.h -- CPU restatement of ElVibRot-TnumTana's H|psi> path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (libevrhip) never links or calls it.
 *
 * Every routine restates one piece of the reference's algorithm in plain C and
 * cites the reference file:line it follows (paths relative to the reference
 * root).  Layouts follow the reference: Fortran column-major matrices,
 * term-local Smolyak tensors first-dimension-fastest, packed basis and
 * direct-product tensors last-dimension-fastest.
 *
 * Parity status of each routine is stated next to its declaration.
 */
#ifndef EVR_ORACLE_H
#define EVR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Hermite rule with nq points: nodes x and the weights the reference
 * stores in base%w, i.e. w_i*exp(x_i^2) (HERM:326-347 reads column 4 of
 * Internal_data/HermQuadra/herm{nq}.txt).  Pinned against those tables. */
int orc_herm_quadrature(int nq, double *x, double *wex2);

/* Hermite functions and first/second derivatives on the grid,
 * d0d1d2poly_Hermite_exp_grille (POLY:1133-1164, POLY:1060-1122, POLY:830-877).
 * B, B1, B2 are nq x nb column-major (dnRGB%d0, d1, d2).  Pinned by
 * orthonormality under the quadrature and by B1 = d/dx B (closed form). */
int orc_herm_basis(int nb, int nq, const double *x,
                   double *B, double *B1, double *B2);

/* Grid-to-grid matrices dnRGG%d0/d1/d2 = {B,B1,B2} * pinv(B) with pinv built
 * from the nb largest eigenpairs of B*B^T (BAS:2195-2460).  nq x nq each.
 * Pinned by D0*B = B, D1*B = B1, D2*B = B2 (the reference's own check,
 * BAS:2399-2440). */
int orc_herm_dngg(int nb, int nq, const double *B, const double *B1,
                  const double *B2, double *D0, double *D1, double *D2);

/* ------------------------------------------------------------------ */
/* Smolyak SG4 basis (SparseGrid_type=4, direct=4).                   */
/* ------------------------------------------------------------------ */
typedef struct orc_sg4 orc_sg4;

/* Build the SG4 basis for D modes, grid level LG, basis level LB, with
 * nb(l) = nbA + nbB*l and nq(l) = nqA + nqB*l (L_TO_n type 0,
 * LTON:308-311), 1D Hermite primitive bases with Q0=0, scaleQ=1.
 * Terms: SPB:924-1435 (Lmin = max(0, LG-D+1), SPB:969), ordered by the
 * type -5 increment (NDI:2463-2516).  Packed basis: type 5 increment
 * (NDI:2396-2461), level of a 1D function from L_TO_nb (SPB:1060-1075).
 * Map: BGSG4:625-949. Weights: SGT2:754-854. */
orc_sg4 *orc_sg4_create(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB);
/* The same with a law per mode: nb_d(l) = nbA[d] + nbB[d]*l, nq_d(l) =
 * nqA[d] + nqB[d]*l (one basis_nD item per mode, each with its own
 * L_TO_nb/L_TO_nq, Source_ElVibRot/sub_Basis/sub_read_data.f90:511-512 and
 * :738-740, nb = nq when L_TO_nb is not given); the level of a 1D function
 * of mode d follows that mode's law. */
orc_sg4 *orc_sg4_create_modes(int D, int LG, int LB, const int *nbA, const int *nbB,
                              const int *nqA, const int *nqB);
void orc_sg4_destroy(orc_sg4 *s);

/* sizes[0]=nterms, [1]=nb_packed, [2]=sum nb_t, [3]=sum nq_t,
 * [4]=max nb_t, [5]=max nq_t, [6]=D, [7]=LG */
void orc_sg4_sizes(const orc_sg4 *s, long long *sizes);
/* levels: nterms x D (term-major), weights: nterms,
 * off_nb/off_nq: nterms+1 prefix sums (tab_Sum_nb/nq_OF_SRep). */
void orc_sg4_terms(const orc_sg4 *s, int *levels, double *weights,
                   long long *off_nb, long long *off_nq);
/* packed basis tuples: nb_packed x D, 0-based 1D function indices. */
void orc_sg4_packed(const orc_sg4 *s, int *tuples);
/* map: sum nb_t entries, packed index (0-based) or -1 (tab_iB_OF_SRep_TO_iB). */
void orc_sg4_map(const orc_sg4 *s, int *map);
/* grid coordinates of every term grid point: sum nq_t x D (point-major). */
void orc_sg4_grid(const orc_sg4 *s, double *q);
/* 1D data of (dim d, level l): nb, nq and the matrices (column-major). */
int orc_sg4_1d(const orc_sg4 *s, int d, int l, int *nb, int *nq, double *x,
               double *B, double *W, double *D1, double *D2);

/* ------------------------------------------------------------------ */
/* Operators.                                                          */
/* ------------------------------------------------------------------ */
typedef struct {
  int type;            /* 0: V only; 1: sum F2_ij d_i d_j + sum F1_i d_i + V;
                          10: -1/2 sum_ij d_i J G_ij d_j (s psi)/(J s) + V   */
  const double *V;     /* grid potential, one value per grid point (term
                          concatenated for SG4), may be NULL (= 0)          */
  const double *G0;    /* D x D constant part of G (type 10)                 */
  const double *Ga;    /* D x D x D, G_ij(q) = G0_ij + sum_k Ga[i+D*(j+D*k)] q_k,
                          may be NULL                                        */
  const double *F2;    /* D x D (type 1)                                     */
  const double *F1;    /* D (type 1), may be NULL                            */
  const double *Jac;   /* per grid point Jacobian J, NULL = 1 (type 10)      */
  const double *sqRhoOVERJac; /* per grid point s = sqrt(rho/J), NULL = 1    */
  const double *Gq;    /* SG4 type 10: G per grid point in place of G0/Ga,
                          Fortran (sum_nq, D, D): Gq(g, j, i) = GGiq(:,j,i) of
                          get_OpGrid_type10_OF_ONEDP_FOR_SG4 (OPSG4:2538-2816),
                          the term-concatenated grid; NULL: G0 + Ga q        */
  const double *Vpoly; /* SG4 with V == NULL: the potential is evaluated at
                          every term-grid point as the reference's grid cache
                          is filled, by calling the potential routine per
                          point (OPSG4:2754-2767, calcN_op), here the
                          polynomial sum_i v2_i q_i^2 + sum_{i<=j<=k} c_ijk
                          q_i q_j q_k + sum_{i<j} d_ij q_i^2 q_j^2 + sum_i e_i
                          q_i^4 with coefficients v2[D], c[(i*D+j)*D+k],
                          d[i*D+j], e[D] in one array; NULL: no potential.
                          Pinned against an explicit V of the same formula. */
  const double *F2q;   /* SG4 type 1 with the operator grids per point, as
                          get_OpGrid_type1_OF_ONEDP_FOR_SG4 tabulates them
                          (OPSG4:2820-3084): Fortran (sum_nq, D, D), F2q(g,i,j)
                          the coefficient of d_i d_j at g (d_i d_i through
                          dnRGG%d2, i /= j through d1 on i and on j); NULL:
                          the constant F2                                   */
  const double *F1q;   /* (sum_nq, D): coefficient of d_i per point, or NULL */
} orc_op;

/* H|psi> on the SG4 basis: sub_TabOpPsi_FOR_SGtype4 (OPSG4:678-979) with
 * per-term sub_TabOpPsi_OF_ONEDP_FOR_SGtype4 (OPSG4:1354-1668): gather
 * (BGSG4:1176-1246), BDP_TO_GDP (BGSG4:2385-2496), DerivOp
 * (BGSG4:2690-2795), grid combine (type 0 OPSG4:1448-1488, type 1
 * OPSG4:1490-1546, type 10 OPSG4:1548-1650), GDP_TO_BDP (BGSG4:2497-2581),
 * weighted scatter (BGSG4:1250-1289).  OpenMP static over terms with an
 * atomic scatter, like PSG4_omp=1 (OPSG4:822-865).  psi/hpsi: nb_packed x
 * nvec column-major.  nthreads<=0 uses the OpenMP default. */
int orc_sg4_oppsi(const orc_sg4 *s, const orc_op *op, int nvec,
                  const double *psi, double *hpsi, int nthreads);
/* Same, restricted to terms [t0, t1) (bounded CPU-baseline samples). */
int orc_sg4_oppsi_range(const orc_sg4 *s, const orc_op *op, int nvec,
                        const double *psi, double *hpsi, int nthreads,
                        long long t0, long long t1);

/* Same, for an explicit list of terms. */
int orc_sg4_oppsi_terms(const orc_sg4 *s, const orc_op *op, int nvec, const double *psi,
                        double *hpsi, int nthreads, const long long *terms, long long nlist);

/* BasisTOGrid / GridTOBasis on SG4 (CS4): psi_G is the concatenation of the
 * term grids (tab_Sum_nq order).  b2g: tabPackedBasis_TO_SmolyakRepBasis +
 * BSmolyakRep_TO_GSmolyakRep (BGSG4:1032, BGSG4:2216).  g2b:
 * GSmolyakRep_TO_BSmolyakRep + SmolyakRepBasis_TO_tabPackedBasis
 * (BGSG4:2101, BGSG4:951-1028; terms with |w|<1e-6 skipped, BGSG4:1004). */
int orc_sg4_b2g(const orc_sg4 *s, const double *psiB, double *psiG);
int orc_sg4_g2b(const orc_sg4 *s, const double *psiG, double *psiB);
/* DerivOp_TO_RDP_OF_SmolaykRep (BGSG4:2690-2795) on every term grid: modes
 * i, j 0-based (< 0: none); i == j dnRGG%d2, else d1 along i and along j.
 * Pinned by the derivative of the grid image of a basis vector against the
 * closed-form derivatives of the Hermite functions (D1 B = B', D2 B = B''). */
int orc_sg4_deriv(const orc_sg4 *s, int i, int j, const double *psiG, double *out);

/* ------------------------------------------------------------------ */
/* Direct-product basis (SparseGrid_type=0).                           */
/* ------------------------------------------------------------------ */
typedef struct orc_dp orc_dp;
orc_dp *orc_dp_create(int D, const int *nb, const int *nq);
void orc_dp_destroy(orc_dp *p);
/* sizes[0] = nb_tot, [1] = nq_tot */
void orc_dp_sizes(const orc_dp *p, long long *sizes);
/* grid coordinates, nq_tot x D, last dimension fastest (BG:730-798). */
void orc_dp_grid(const orc_dp *p, double *q);
/* type-10 H|psi>: sub_TabOpPsi_WITH_MemGrid_BGG_Hamil10 (OPPSI:1797-2070)
 * with DP RecRvecB_TO_RVecG (BG:730-798), DerivOp_TO_RVecG (BG:1291-1352),
 * RecRVecG_TO_RvecB (BG:134-221), sub_sqRhoOVERJac_Psi (OPPSI:3067-3185).
 * Type 0 and type 1 are also accepted (same grid combine as SG4). */
int orc_dp_oppsi(const orc_dp *p, const orc_op *op, int nvec,
                 const double *psi, double *hpsi);
int orc_dp_b2g(const orc_dp *p, const double *psiB, double *psiG);
int orc_dp_g2b(const orc_dp *p, const double *psiG, double *psiB);
/* DerivOp_TO_RVecG on the DP grid (BG:1291-1352), modes as orc_sg4_deriv. */
int orc_dp_deriv(const orc_dp *p, int i, int j, const double *psiG, double *out);

#ifdef __cplusplus
}
#endif
#endif
