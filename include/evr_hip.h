/*
 * evr_hip.h -- C ABI of the MI355X H|psi> engine (libevrhip.so).
 *
 * Drop-in boundary for ElVibRot-TnumTana's operator action on Smolyak (SG4)
 * and direct-product (DP) vibrational bases.  Every entry point returns an
 * int status: 0 = OK, < 0 = error (message via evr_last_error()); it never
 * exits the process.  The reference STOPs on the same conditions; its Fortran
 * host maps a nonzero status to STOP (see evr_hip_f.f90).
 *
 * Conventions
 *  - Matrices are Fortran column-major, as in the reference's basis type
 *    (dnRGB%d0 is nq x nb, dnRBGwrho%d0 is nb x nq, dnRGG%d1/d2 are nq x nq).
 *  - Vectors psi are packed-basis, column-major nb_packed x nvec (one column
 *    per psi, the reference's RvecB layout, PSIAL:88).  Complex psi is passed
 *    as two real columns, as the reference does (OPPSI:392-408).
 *  - Term-local tensors are first-dimension-fastest (BGSG4:2443-2476); the
 *    concatenated Smolyak grid follows tab_Sum_nq_OF_SRep order.
 *  - The caller owns host arrays; a plan owns its device buffers and stream.
 *    One caller thread per plan; different plans are independent.
 *  - Reductions are deterministic: the weighted Smolyak scatter is a
 *    segmented sum in a fixed order, not atomics.
 *
 * Reference file abbreviations (relative to the reference root):
 *   OPPSI Source_ElVibRot/sub_Operator/sub_OpPsi.f90
 *   OPSG4 Source_ElVibRot/sub_Operator/sub_OpPsi_SG4.f90
 *   BGSG4 Source_ElVibRot/sub_Basis/sub_Basis_SG4/sub_module_basis_BtoG_GtoB_SG4.f90
 *   BG    Source_ElVibRot/sub_Basis/sub_module_basis_BtoG_GtoB.f90
 *   PSIBG Source_ElVibRot/sub_WP/sub_module_psi_B_TO_G.f90
 *   BSET  Source_ElVibRot/sub_Basis/sub_module_basis_set_alloc.f90
 *   SGT2  Source_ElVibRot/sub_Basis/sub_module_param_SGType2.f90
 *   DAV   Source_ElVibRot/sub_propagation/sub_module_Davidson.f90
 *   MARCH Source_ElVibRot/sub_propagation/sub_module_propa_march.f90
 *   CTRL  Source_ElVibRot/sub_propagation/sub_control.f90
 *   HMAX  Source_ElVibRot/sub_propagation/sub_Hmax.f90
 *   POLY  Source_Lib/sub_communf90/sub_math/sub_polyortho.f90
 */
#ifndef EVR_HIP_H
#define EVR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EVR_OK 0
#define EVR_ERR_ARG -1
#define EVR_ERR_HIP -2
#define EVR_ERR_SIZE -3
#define EVR_ERR_STATE -4
#define EVR_ERR_IO -5

/* evr_oppsi flags */
#define EVR_FP64 0          /* all contractions in FP64                         */
#define EVR_FP32 1          /* contractions in FP32, FP32 accumulation          */
#define EVR_DEVICE_PTRS 2   /* psi/hpsi are device pointers on the plan stream  */
#define EVR_MATRIX_FREE 4   /* ignore an assembled operator (sum-factorised path) */
#define EVR_AUTO_PATH 8     /* with an assembled operator, take whichever of the
                               assembled and sum-factorised paths a cost model
                               says is faster for this nvec (the assembled pool
                               is streamed once per call whatever nvec); the
                               propagation drivers always pass it              */

/* evr_plan_assemble precisions */
#define EVR_ASM_FP64 1
#define EVR_ASM_FP32 2

/* ------------------------------------------------------------------------
 * Basis description: what the reference's SG4 basis type holds
 * (tab_basisPrimSG(0:L,1:D) with dnRGB/dnRBGwrho/dnRGG, BSET:85-102;
 * param_SGType2 term table, weights and packed map, SGT2:54-102).  A DP basis
 * is one term with weight 1 whose map is the DP ordering.
 * ---------------------------------------------------------------------- */
typedef struct evr_basis_desc {
  int D;                     /* number of modes                                  */
  int nlev;                  /* 1D basis slots per mode (L+1 for SG4, 1 for DP)  */
  int64_t nterms;
  const int *term_levels;    /* nterms x D, term-major: slot of mode d in term t */
  const double *weights;     /* nterms: WeightSG (SGT2:754-854)                  */
  const int *nb;             /* D x nlev (index d*nlev + l)                      */
  const int *nq;             /* D x nlev                                         */
  const double *x;           /* grid points, per slot nq values, slots in order  */
  const double *B;           /* per slot nq x nb: dnRGB%d0                       */
  const double *W;           /* per slot nb x nq: dnRBGwrho%d0                   */
  const double *D1;          /* per slot nq x nq: dnRGG%d1                       */
  const double *D2;          /* per slot nq x nq: dnRGG%d2                       */
  int64_t nb_packed;
  const int *map;            /* sum_t nb_t: packed index (0-based) of each
                                term-local basis function, -1 if pruned
                                (tab_iB_OF_SRep_TO_iB, BGSG4:625-949)           */
  int grid_last_fastest;     /* 1 for a DP basis: V and grid vectors are in the
                                reference's DP grid order, last mode fastest
                                (BG:730-798); 0: concatenated term grids       */
} evr_basis_desc;

/* ------------------------------------------------------------------------
 * Operator description (nb0 = 1, one electronic surface).
 *  type 0 : V psi                                   (OPSG4:1448-1488)
 *  type 1 : sum_ij F2_ij d_i d_j + sum_i F1_i d_i + V, constant F2/F1,
 *           d_i d_i through dnRGG%d2                (OPSG4:1490-1546)
 *  type 10: -1/2 sum_ij d_i G_ij(q) d_j + V with G_ij(q) = G0_ij +
 *           sum_k Ga_ijk q_k, J = sqRhoOVERJac = 1   (OPSG4:1548-1650)
 * V is one value per grid point of the concatenated term grids (the
 * reference's OpGrid%Grid with Save_MemGrid, OPSG4:2656-2666), or NULL and
 * set later with evr_plan_set_potential / evr_plan_gen_potential_poly.
 * ---------------------------------------------------------------------- */
typedef struct evr_op_desc {
  int type;
  const double *V;           /* sum_t nq_t or NULL                               */
  const double *G0;          /* D x D                                            */
  const double *Ga;          /* D x D x D (i + D*(j + D*k)) or NULL              */
  const double *F2;          /* D x D (type 1)                                   */
  const double *F1;          /* D (type 1) or NULL                               */
  int hermitian;             /* 0: the reference operator H as is (OpPsi).
                                1: its symmetric part (H + H^T)/2, evaluated
                                exactly in the same pass: the V part is
                                symmetric (W = B^T diag(w)), P and Q are exactly
                                (anti)symmetric under the quadrature, and the
                                mode factors T, U and R+S are replaced by their
                                (anti)symmetric parts.  Needed by the symmetric
                                eigensolvers: with dnRGG%d1 = B' pinv(B) the
                                type-10 operator is not symmetric.             */
} evr_op_desc;

typedef struct evr_plan evr_plan;

/* Last error message of this thread (never NULL). */
const char *evr_last_error(void);

/* ---- basis construction helpers (host) -------------------------------- */
/* Hermite (HO) Smolyak SG4 basis with nb(l) = nbA + nbB*l, nq(l) = nqA +
 * nqB*l, grid level LG, basis level LB (RecSparseGrid_ForDP_type4,
 * Source_ElVibRot/sub_Basis/sub_quadra_SparseBasis.f90:924-1435).  The
 * returned descriptor owns its arrays; free with evr_basis_desc_free. */
int evr_sg4_hermite_desc(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB,
                         evr_basis_desc **out);
/* The same with a law per mode (the L_TO_nq_A / L_TO_nq_B of each basis_nD
 * block): nb_d(l) = nbA[d] + nbB[d] l, nq_d(l) = nqA[d] + nqB[d] l. */
int evr_sg4_hermite_desc_modes(int D, int LG, int LB, const int *nbA, const int *nbB, const int *nqA,
                               const int *nqB, evr_basis_desc **out);
/* Hermite direct-product basis, nb[d] functions on nq[d] points per mode,
 * packed order last-dimension-fastest (BG:730-798). */
int evr_dp_hermite_desc(int D, const int *nb, const int *nq, evr_basis_desc **out);
void evr_basis_desc_free(evr_basis_desc *d);
/* Sizes: out[0]=sum nb_t, out[1]=sum nq_t, out[2]=max nb_t, out[3]=max nq_t. */
int evr_basis_desc_sizes(const evr_basis_desc *d, int64_t *out);
/* Coordinates of every grid point (sum nq_t x D, point-major, host). */
int evr_basis_desc_grid(const evr_basis_desc *d, double *q);

/* ---- wavepacket files (host) ------------------------------------------- */
/* Write nwp wavepackets of nb0 x nb_packed rows (state-major: row
 * (i_e - 1) nb_packed + packed index; column-major; complex as (Re, Im)
 * column pairs as evr_propagate) to `path` in the reference's save-file
 * format, sub_save_psi (Source_ElVibRot/sub_WP/sub_module_psi_io.f90:365-424):
 *  version 0, formatted (the reference's default, FilePsiVersion = 0):
 *    "nb_psi nb_tot", then per wavepacket one line "n_1 .. n_D i_h i_e Re [Im]"
 *    per function and "end wp k" (Write_Psi_nDBasis, :1660-1686);
 *  version 1, formatted: namelist headerFile, the index list, then each
 *    wavepacket list-directed (:1441-1604, :1750-1762);
 *  version 1, unformatted: Fortran sequential records with 4-byte markers
 *    and 4-byte integers (:1480-1487).
 * n_d is the 1-based 1D function index of mode d (Rec_ndim_index), i_h = 1. */
int evr_psi_save(const evr_basis_desc *basis, int nb0, int nwp, int cplx, const double *psi, int version,
                 int formatted, const char *path);
/* Read a file in any of those forms (detected from the header, as
 * Read_header_saveFile_psi :1294-1439) into psi (nb0 x nb_packed x max_wp
 * columns, x2 for complex; zeroed first): each value goes to the function
 * with the same index tuple (n_1..n_D, i_h, i_e), so a file written on
 * another basis is projected onto this one (Read_list_nDindBasis1_TO_
 * nDindBasis2 :1143-1292).  *nwp = wavepackets read (at most max_wp). */
int evr_psi_read(const evr_basis_desc *basis, int nb0, int cplx, int max_wp, double *psi, int *nwp,
                 const char *path);

/* ---- plans -------------------------------------------------------------- */
/* Replaces the per-call setup of sub_TabOpPsi_FOR_SGtype4 (OPSG4:678-979):
 * uploads tables, 1D matrices, the basis-space KEO factors and V.
 * max_nvec bounds the vectors per evr_oppsi call.  D <= 12: every mode is a
 * position of the term kernel.  12 < D <= 64 (SG4 at low levels, e.g. the
 * reference's 21D and 24D working tests): the modes of a term with one
 * function on one grid point merge into one scalar position, so each term
 * may have at most 11 other modes; not for a DP basis, a Ga, or
 * evr_plan_gen_potential_poly.  1D slots: 1 <= nb <= nq <= 64.  A term runs
 * in one workgroup with its tensors in LDS when it has at most 8192
 * functions, slots of at most 31 functions and slabs that fit; any other
 * term (D <= 12) runs streamed: the reference's basis -> grid, grid
 * operator and grid -> basis sequence as whole-device passes over global
 * memory, one mode at a time, with a work block of (7 + D) x nq_t x k
 * values made here (k vectors per pass, within 4 GiB).  A plan with
 * streamed terms is not assembled and takes no operator grids per point;
 * its diagonal is read off H on blocks of unit vectors.  A one-term plan
 * (direct-product basis) of at least 256 functions streams its term for
 * H|psi> and the transforms even when it fits a workgroup, and keeps the
 * resident kernel for the assembly and the diagonal. */
int evr_plan_create(const evr_basis_desc *basis, const evr_op_desc *op, int max_nvec,
                    evr_plan **out);
int evr_plan_destroy(evr_plan *p);
/* Replace V (host array, sum nq_t values). */
int evr_plan_set_potential(evr_plan *p, const double *V);
/* Type 10 with the kinetic tables per grid point, as the reference's
 * get_OpGrid_type10_OF_ONEDP_FOR_SG4 (OPSG4:2538-2816) gets them from Tnum
 * (get_d0GG): G is Fortran (sum_nq, D, D) over the term-concatenated grid,
 * G(g, j, i) = GGiq(g, j, i); Jac and sqRhoOVERJac are sum_nq values or NULL
 * (= 1).  From then on the kinetic part is the reference's grid form
 * (OPSG4:1566-1632): -1/2 (1/(J s)) sum_i d_i [J sum_j G(:, j, i) d_j (s psi)]
 * with d = dnRGG%d1, on every term grid; G0/Ga are no longer used.  SG4 bases
 * with D <= 12, matrix-free only (EVR_ERR_ARG for an assembly); EVR_FP32
 * runs the transforms and keeps the grid columns in FP32 (tables and grid
 * arithmetic FP64).  A hermitian plan applies (H + H^T) / 2 of this grid
 * form, the transpose as W diag(1/w) Op^T diag(w) B with W = B^T diag(w) on
 * every term grid.  The plan diagonal then comes from H on blocks of unit
 * vectors.  Device
 * memory: (D^2 + 2) sum_nq tables and (D + 2) sum_nq work doubles per vector
 * processed at once. */
int evr_plan_set_keo_grid(evr_plan *p, const double *G, const double *Jac, const double *sqRhoOVERJac);
/* Type 1 with the operator grids per grid point, as the reference's
 * get_OpGrid_type1_OF_ONEDP_FOR_SG4 (OPSG4:2820-3084) tabulates them from
 * Tnum (TnumKEO_TO_tab_d0H, Source_PrimOperator/sub_PrimOp.f90:1540-1700) and
 * combines them (OPSG4:1490-1546): H psi_G = V psi_G + sum_ij F2(:, i, j)
 * d_i d_j psi_G + sum_i F1(:, i) d_i psi_G on every term grid, d_i d_i through
 * dnRGG%d2 and d_i d_j (i /= j) through dnRGG%d1 on both modes.  F2 is
 * Fortran (sum_nq, D, D) over the term-concatenated grid, F2(g, i, j) the
 * coefficient of d_i d_j (the reference's one term per pair i <= j: put its
 * value in one triangle and zero in the other); F1 is (sum_nq, D) or NULL.
 * Replaces the plan's constant F2/F1.  SG4 bases with D <= 12, a type-1 plan
 * matrix-free only (EVR_ERR_ARG for an assembly); EVR_FP32 and the
 * hermitian form (H + H^T) / 2 as for evr_plan_set_keo_grid (the transposed
 * operator with the tables w F2, w F1 made here); the plan diagonal then
 * comes from H on blocks of unit vectors. */
int evr_plan_set_type1_grid(evr_plan *p, const double *F2, const double *F1);
/* Generate V on the device: V(q) = sum_i v2_i q_i^2 + sum_{i<=j<=k} c_ijk
 * q_i q_j q_k + sum_{i<j} d_ij q_i^2 q_j^2 + sum_i e_i q_i^4, with c given as a
 * dense D^3 array (only i<=j<=k read), d as D x D (only i<j read). */
int evr_plan_gen_potential_poly(evr_plan *p, const double *v2, const double *c3,
                                const double *d22, const double *e4);
/* Sizes: out[0]=nb_packed, [1]=nterms, [2]=sum nb_t, [3]=sum nq_t,
 * [4]=max nb_t, [5]=max nq_t, [6]=D. */
int evr_plan_sizes(const evr_plan *p, int64_t *out);
/* hipStream_t of the plan (as void*), for callers that pass device pointers. */
void *evr_plan_stream(evr_plan *p);
/* Largest nvec one evr_oppsi call accepts (the max_nvec given at creation). */
int evr_plan_max_nvec(const evr_plan *p);

/* Assembled operator (no counterpart in the reference, which applies every
 * term matrix-free on each call): every term's nb_t x nb_t matrix H_t (its
 * kinetic and potential parts, Smolyak weight included) is built once by the
 * FP64 term kernel applied to the term's unit vectors and kept on the device
 * in MFMA fragment order, FP64 and/or FP32 (`which`: EVR_ASM_FP64 |
 * EVR_ASM_FP32; about sum_t nb_t^2 values each, padded to 16 x 16 tiles,
 * e.g. 144 GB FP64 + 72 GB FP32 at D=9 L=8, so this trades HBM capacity for
 * dense MFMA work; EVR_ERR_SIZE when the device memory is short).  evr_oppsi then computes
 * R_t = H_t X_t with MFMA and the same scatter, in the precision asked,
 * unless EVR_MATRIX_FREE is set.  Setting a new potential drops the assembled
 * operator.  evr_plan_assembled returns the mask held and the bytes used. */
int evr_plan_assemble(evr_plan *p, int which);
int evr_plan_assembled(const evr_plan *p, double *bytes);
int evr_plan_release_assembled(evr_plan *p);

/* Several systems as one block-diagonal operator (no counterpart in the
 * reference, which treats one system per run): the FP64 assembled operators
 * of n plans (each assembled with EVR_ASM_FP64, any sizes and dimensions)
 * are copied into one pool with concatenated tables, so one H|psi> over the
 * concatenated packed vectors of all systems is one launch sequence (gather,
 * MFMA apply over every system's terms, scatter).  The plans may be released
 * or destroyed afterwards.  nb_offsets (n + 1 values) gives where system k's
 * packed functions start in the concatenation. */
typedef struct evr_multi evr_multi;
int evr_multi_create(evr_plan *const *plans, int n, int max_nvec, evr_multi **out);
int evr_multi_destroy(evr_multi *m);
int evr_multi_offsets(const evr_multi *m, int64_t *nb_offsets);
/* Number of systems n. */
int evr_multi_count(const evr_multi *m);
void *evr_multi_stream(evr_multi *m);
/* hpsi = H psi for nvec columns of sum_k nb_k rows (column-major); flags:
 * EVR_DEVICE_PTRS for device pointers on evr_multi_stream. */
int evr_multi_oppsi(evr_multi *m, int nvec, const double *psi, double *hpsi, int flags);

/* nb0 coupled electronic states over one packed basis (the reference's
 * nb_be > 1 operators, V(:,ib0,jb0) per grid point, OPSG4:1566-1576): a
 * wavepacket holds nb0 blocks of nb_packed rows, state-major within each
 * column (vibrational index fastest), and (H psi)_i = sum_j H_ij psi_j.  The
 * blocks are plans on the same basis, row-major in blocks[i*nb0 + j] (NULL
 * for a zero block): a diagonal block carries the kinetic part and V_ii, an
 * off-diagonal block is a type-0 plan with V_ij.  The plans must outlive the
 * coupled operator; each block runs through its own fastest path (assembled
 * when present). */
typedef struct evr_coupled evr_coupled;
int evr_coupled_create(evr_plan *const *blocks, int nb0, evr_coupled **out);
int evr_coupled_destroy(evr_coupled *c);
/* nb_packed of the basis, nb0, and the largest nvec of one call. */
int evr_coupled_sizes(const evr_coupled *c, int64_t *nb, int *nb0, int *max_nvec);
void *evr_coupled_stream(evr_coupled *c);
/* hpsi = H psi for nvec columns of nb0 x nb_packed rows; flags as evr_oppsi
 * (EVR_DEVICE_PTRS: device pointers on evr_coupled_stream). */
int evr_coupled_oppsi(evr_coupled *c, int nvec, const double *psi, double *hpsi, int flags);
/* Per-term tile autotuner, the counterpart of the reference's run-time tuning
 * of its term loop (Tune_SG4threads_HPsi, Source_ElVibRot/sub_main/vib.f90:
 * 1598-1705, which times H|psi> for each OpenMP thread count and keeps the
 * fastest).  The sum-factorised term kernels come in EVR_NCLASSES work
 * classes (threads x elements per thread x vectors per workgroup, tile
 * shapes of growing size); the static rule gives each term the smallest
 * class that holds its nb_t functions.  The tuner takes the terms of each
 * static class (a bucket) in turn, times the whole sum-factorised H|psi> of
 * nvec vectors (flags EVR_FP64 or EVR_FP32; HIP events on the plan stream,
 * median of reps after a warm-up) with the bucket moved one and two classes
 * up, keeps a larger class only when it is more than 3% faster, and stops
 * at the first slower one.  The layout chosen stays with the plan
 * (evr_oppsi, the assembly).  chosen (EVR_NCLASSES ints, or NULL): the class
 * of each bucket; ms_static / ms_tuned: the H|psi> times of the static rule
 * and of the tuned layout.  Results are unchanged by the layout (the same
 * arithmetic per term). */
#define EVR_NCLASSES 8
int evr_plan_autotune(evr_plan *p, int nvec, int flags, int reps, int *chosen, double *ms_static,
                      double *ms_tuned);
/* Profiling (no counterpart in the reference, whose analogue is the
 * nb_mult_BTOG/GTOB counters, SYS:120): when on, every term-kernel launch is
 * bracketed by HIP events on the plan stream; evr_plan_kernel_stats returns,
 * per kernel class, the accumulated milliseconds, algorithmic flops and
 * launch count (names: 64 chars each).  Turning it on resets the counters. */
int evr_plan_set_profiling(evr_plan *p, int on);
int evr_plan_kernel_stats(evr_plan *p, int max, char *names, double *ms, double *flops,
                          int64_t *launches, int *n);

/* ---- operator action ---------------------------------------------------- */
/* H|psi> for nvec vectors: sub_TabOpPsi(Psi(:), OpPsi(:), para_Op)
 * (OPPSI:701-793) dispatched to sub_TabOpPsi_FOR_SGtype4 (OPSG4:678) for
 * SG4 and to sub_TabOpPsi_WITH_MemGrid_BGG_Hamil10 (OPPSI:1797) for DP.
 * psi, hpsi: nb_packed x nvec column-major.  flags: EVR_FP64 | EVR_FP32,
 * optionally | EVR_DEVICE_PTRS. */
int evr_oppsi(evr_plan *p, int nvec, const double *psi, double *hpsi, int flags);

/* H|psi> with no device allocation (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels and results as
 * evr_oppsi with EVR_DEVICE_PTRS (required in flags), but the packed staging
 * block, the term rows and the segment sums live in one static device array
 * of EVR_NOALLOC_SCRATCH_BYTES per GPU instead of the plan's buffers:
 * 8 x (nb_packed + term rows + segments) x nvec bytes, rounded to 256-byte
 * blocks, must fit (EVR_ERR_SIZE otherwise; e.g. D=6 SG4 L=5 takes 32
 * vectors, D=9 SG4 L=8 none).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (on different streams or host threads) race.
 * No part of the array needs clearing (each call writes what it reads).  The
 * plan's device must be current; FP32 needs a prior FP32 evr_oppsi call (the
 * FP32 potential copy) or an FP32 assembled operator; plans with G, J, s per
 * grid point are refused. */
#define EVR_NOALLOC_SCRATCH_BYTES (64u << 20)
int evr_oppsi_noalloc(evr_plan *p, int nvec, const double *psi, double *hpsi, int flags);

/* Diagonal of H in the packed basis (nb_packed values): V part exact
 * (sum_t w_t W^(x) diag(V) B^(x) diagonal, one grid->basis pass of V with
 * W o B^T), kinetic part from the diagonals of the per-mode factors.  Used as
 * the Davidson preconditioner (NewVec_type 2, DAV:1392-1413).  A plan with
 * G, J, s per grid point (evr_plan_set_keo_grid) reads its exact diagonal off
 * H applied to blocks of max_nvec unit vectors (nb_packed / max_nvec
 * applications).  flags: EVR_DEVICE_PTRS if diag is a device pointer. */
int evr_plan_diagonal(evr_plan *p, double *diag, int flags);

/* BasisTOGrid: sub_PsiBasisRep_TO_GridRep (PSIBG:60) -> RecRvecB_TO_RVecG
 * (BG:605), SG4 branch BSmolyakRep_TO_GSmolyakRep (BGSG4:2216).  psiG is the
 * concatenated term grid (sum nq_t).  One vector, host pointers. */
int evr_basis_to_grid(evr_plan *p, const double *psiB, double *psiG);
/* GridTOBasis: sub_PsiGridRep_TO_BasisRep (PSIBG:178) -> RecRVecG_TO_RvecB
 * (BG:59), SG4 branch with the weighted Smolyak sum (BGSG4:951-1028). */
int evr_grid_to_basis(evr_plan *p, const double *psiG, double *psiB);

/* Multi-vector forms of BasisTOGrid / GridTOBasis (sub_PsiBasisRep_TO_GridRep
 * PSIBG:60, sub_PsiGridRep_TO_BasisRep PSIBG:178, for a table of psi):
 * ncol columns, psiB nb_packed x ncol and psiG sum_nq x ncol column-major
 * (the grid in the caller's order: concatenated term grids for SG4, the
 * reference's DP order for a DP basis); flags EVR_DEVICE_PTRS for device
 * pointers on the plan stream (then no host copies are made). */
int evr_basis_to_grid_n(evr_plan *p, int ncol, const double *psiB, double *psiG, int flags);
int evr_grid_to_basis_n(evr_plan *p, int ncol, const double *psiG, double *psiB, int flags);
/* Grid derivative DerivOp_TO_RVecG(RVecG, nq, basis, tab_der) (BG:1210-1434;
 * SG4: DerivOp_TO_RDP_OF_SmolaykRep on every term grid, BGSG4:2690-2795) of
 * ncol grid columns: tab_der = (i, j), 1-based modes, 0 (or < 0) for none:
 * (i, i) dnRGG%d2 along i; (i, 0) or (0, i) dnRGG%d1 along i; (i, j) d1 along
 * both; (0, 0) a copy.  Layout and flags as evr_grid_to_basis_n; D <= 12. */
int evr_grid_deriv(evr_plan *p, int ncol, const int *tab_der, const double *psiG, double *out, int flags);

/* ---- eigen-solver --------------------------------------------------------- */
/* Block Davidson for the nev lowest levels: sub_propagation_Davidson
 * (DAV:54-748) with sub_MakeHPsi (DAV:880-1024), Rayleigh-Ritz in FP64
 * (DAV:1028-1143, here a device symmetric eigensolver), residuals
 * (DAV:1214-1288), new vectors (DAV:1290-1588: NewVec_type 2 when precond=1,
 * type 4 otherwise), collapse (DAV:516-606; max_sub >= nev + block).
 * Symmetric H expected (type 1, or a plan created with hermitian=1).
 * precision EVR_FP32 (mixed): the H|psi> of every new vector in FP32, the
 * subspace algebra in FP64, and FP64 defect corrections: the subspace is
 * collapsed on the Ritz vectors and H applied to them in FP64 when the
 * residuals first drop below refine_resi, stop falling (6 iterations without
 * a 10% decrease: the FP32 floor) or all meet the criteria; couplings of the
 * FP64 columns with FP32 ones come from the FP64 side; convergence is
 * declared on an all-FP64 H V only (FP64 residuals). */
typedef struct evr_davidson_opts {
  int nev;                 /* levels wanted                                     */
  int block;               /* new vectors per iteration (<= plan max_nvec)      */
  int max_sub;             /* subspace size before a collapse (0: the larger of
                              4 nev and nev + 4 block)                          */
  int max_it;              /* iterations (0: 200)                               */
  double conv_resi;        /* ||H x - E x|| threshold (energy units)            */
  double conv_ene;         /* |E(it) - E(it-1)| threshold                       */
  int precond;             /* 1: Davidson correction r/(E - diag)               */
  const double *diag;      /* host diagonal (nb_packed) or NULL: evr_plan_diagonal */
  const double *guess;     /* host start vectors nb_packed x nguess, or NULL:
                              unit vectors on the lowest diagonal entries       */
  int nguess;
  int precision;           /* EVR_FP64 or EVR_FP32 (mixed, see above)           */
  double refine_resi;      /* residual of the first FP64 defect correction      */
  int max_nvec_hint;       /* vectors per evr_oppsi call (0: block)             */
  int verbose;
  /* Wavepacket files (read_WP of the reference, DAV:215-216 with
   * ReadWP0_Davidson DAV:752 through sub_read_psi0; the levels' vectors saved
   * with sub_save_psi as DAV:218): basis is the descriptor the plan was made
   * from (file rows are matched by index tuple, evr_psi_read).  guess_file:
   * real start vectors read from it (up to max_sub, in place of guess);
   * save_file: the nev Ritz vectors written at the end (version 0,
   * formatted).  NULL: not used. */
  const evr_basis_desc *basis;
  const char *guess_file;
  const char *save_file;
  int graph;               /* 1: the H|psi> of each iteration's new vectors is
                              one captured HIP graph per (block size,
                              precision), replayed on the operator's stream
                              (results identical to graph = 0)               */
} evr_davidson_opts;

typedef struct evr_davidson_info {
  int iterations;
  int converged;           /* 1 when all nev levels met both criteria in FP64   */
  int nconv;
  int64_t hpsi_fp64;       /* vector H|psi> applications in FP64                */
  int64_t hpsi_fp32;       /* ... in FP32                                       */
  double max_resi;
  double seconds;
} evr_davidson_info;

/* evals: nev values (ascending); evecs: host nb_packed x nev or NULL. */
int evr_davidson(evr_plan *p, const evr_davidson_opts *opts, double *evals, double *evecs,
                 evr_davidson_info *info);

/* ---- time propagation ------------------------------------------------------ */
/* psi(t + dt) for a batch of nwp complex wavepackets: march_cheby
 * (MARCH:4142-4433) and march_SIL (MARCH:2899-3100).  psi is nb_packed x
 * (2 nwp) column-major, columns (Re psi_w, Im psi_w) for w = 0..nwp-1, and is
 * overwritten with psi(nsteps dt).
 *
 * Chebyshev: H_s = (H - E0)/Esc with E0 = Hmin + dE/2, Esc = dE/2, dE =
 * Hmax - Hmin (MARCH:4211-4231); coefficients c_0 = J_0(r), c_k = 2 J_k(r),
 * r = Esc dt (cof, MARCH:5734-5755); orders 0..npoly-1 at most, stopping at
 * the first order k >= 2 with |c_k| <w_k|w_k> < poly_tol for every wavepacket
 * (MARCH:4290-4335); on failure Hmax += dE/10 and the step is redone, at most
 * 10 times (MARCH:4348-4373); finally psi *= exp(-i E0 dt).  Hmin >= Hmax in
 * the options asks for the interval from a 60-step Lanczos estimate of the
 * extreme eigenvalues, widened by 2% of the range (the reference obtains it
 * from sub_Hmax, HMAX:43-328).
 * SIL: Krylov space of at most npoly vectors from psi_w, E0_w = <psi|H|psi>,
 * stopping when the last Krylov component of exp(-i T dt) e_1 is below
 * poly_tol (MARCH:2957-2990, UPsi_spec MARCH:3101-3172).  The Krylov vectors
 * are re-orthogonalised against all previous ones (classical Gram-Schmidt,
 * two passes); the reference keeps the three-term recurrence only.
 * Field: with field != NULL the step n (0-based) uses H - field[n*nwp + w] mu
 * for wavepacket w, mu being a plan on the same basis (a dipole surface as a
 * type-0 operator); the field is constant over a step (for Chebyshev the
 * spectral interval is widened by max|field| * |mu|_max).
 * Taylor (EVR_PROPA_TAYLOR, march_noD_field): psi(t_n + dt) = e^{-i E0 dt}
 * sum_i w_i, w_0 = psi, w_{i+1} = (-i dt/(i+1)) [(H - E0) w_i - mu sum_{k<=i}
 * (dt^k/k!) E^(k)(t_n) w_{i-k}] (the Taylor series of the time-dependent
 * equation, field derivatives up to the largest k with max|E^(k)| dt^(k+1)/k!
 * > poly_tol), orders up to npoly, stopping when every |w_{i+1}|^2 < poly_tol;
 * E0 = (Hmin + Hmax)/2 (Hmin >= Hmax: from the Lanczos interval).
 * RK4 (EVR_PROPA_RK4, march_RK4_field): classical Runge-Kutta on
 * dpsi/dt = -i [(H - E0) - E(t) mu] psi (fcn_field) with the field at t_n,
 * t_n + dt/2, t_n + dt, then psi *= e^{-i E0 dt}.
 * The operator must be symmetric (type 1, or a plan created hermitian=1). */
#define EVR_PROPA_CHEBY 1
#define EVR_PROPA_SIL 2
#define EVR_PROPA_TAYLOR 3   /* march_noD_field (MARCH:1635-1836)               */
#define EVR_PROPA_RK4 4      /* march_RK4_field with fcn_field (MARCH:392-543,
                                1434-1515)                                      */
typedef struct evr_propa_opts {
  int method;              /* EVR_PROPA_CHEBY or EVR_PROPA_SIL                  */
  double dt;               /* WPdeltaT                                          */
  int nsteps;
  int npoly;               /* max Chebyshev terms / Krylov dimension            */
  double poly_tol;
  double Hmin, Hmax;       /* spectral interval (Hmin >= Hmax: estimated)       */
  const double *field;     /* nsteps x nwp values (host) or NULL                */
  evr_plan *mu;            /* coupling operator when field != NULL              */
  const double *psi0;      /* autocorrelation reference, same layout as psi;
                              NULL: the initial psi                             */
  int precision;           /* EVR_FP64 (EVR_FP32: H|psi> in FP32)               */
  int device_ptrs;         /* psi / psi0 are device pointers                    */
  int graph;               /* Chebyshev: 1 = replay each step after the first as
                              one captured HIP graph of fixed order (the first
                              step's), the exit test |c_K| <w_K|w_K> < poly_tol
                              checked after every replay and the step redone
                              adaptively when it fails; 0 = adaptive order per
                              step with a host test per order (march_cheby)     */
  const double *field_sub; /* TAYLOR / RK4 (host): the time-dependent field of
                              step n at sub-points, field_sub[(n*nsub + j)*nwp
                              + w]: TAYLOR the derivatives E^(j)(t_n), j < nsub
                              (sub_dnE, sub_module_field.f90:541); RK4 E(t_n),
                              E(t_n + dt/2), E(t_n + dt) (nsub = 3); mu the
                              coupling plan; NULL: no field                   */
  int nsub;
} evr_propa_opts;

typedef struct evr_propa_info {
  int64_t hpsi;            /* real-column H|psi> applications                   */
  int max_order;           /* largest Chebyshev order / Krylov dimension used   */
  int retries;             /* Hmax enlargements                                 */
  double Hmin, Hmax;       /* interval used (Chebyshev)                         */
  double seconds;
} evr_propa_info;

/* autocorr: host nsteps x nwp complex values <psi0_w|psi_w(t_n)> (Re, Im
 * pairs, Calc_AutoCorr) or NULL. */
int evr_propagate(evr_plan *H, const evr_propa_opts *opts, int nwp, double *psi,
                  double *autocorr, evr_propa_info *info);
/* The same on coupled electronic states (evr_coupled_create): psi and psi0
 * columns hold nb0 x nb_packed rows; no field coupling (field and mu NULL). */
int evr_coupled_propagate(evr_coupled *H, const evr_propa_opts *opts, int nwp, double *psi,
                          double *autocorr, evr_propa_info *info);
/* Coupled states driven by a field (the reference's field term with the
 * dipole operator of the electronic states, MARCH:1434-1836): step n uses
 * H - field[n*nwp + w] mu, mu a coupled operator on the same states (its
 * blocks mu_ij typically type-0 plans of the transition and permanent
 * dipoles); opts->field as for evr_propagate, opts->mu NULL. */
int evr_coupled_propagate_field(evr_coupled *H, evr_coupled *mu, const evr_propa_opts *opts, int nwp,
                                double *psi, double *autocorr, evr_propa_info *info);

/* Diagonal of a coupled operator (nb0 x nb_packed values, state-major): the
 * diagonals of the blocks H_ii (evr_plan_diagonal), zero for a missing block.
 * flags: EVR_DEVICE_PTRS if diag is a device pointer. */
int evr_coupled_diagonal(evr_coupled *C, double *diag, int flags);

/* Lowest levels of coupled electronic states by the block Davidson of
 * evr_davidson (sub_propagation_Davidson on an operator with nb_bie > 1,
 * DAV:516-606); vectors of nb0 x nb_packed rows, the preconditioner diagonal
 * from evr_coupled_diagonal unless opts->diag is given; FP64 only
 * (opts->precision EVR_FP32 is refused). */
int evr_coupled_davidson(evr_coupled *C, const evr_davidson_opts *opts, double *evals, double *evecs,
                         evr_davidson_info *info);

/* march_cheby on a multi-system operator (evr_multi_create), one complex
 * wavepacket per system (psi: sum_k nb_k x 2, Re and Im columns), every
 * system with its own interval [Hmin[k], Hmax[k]] (its E0, Esc and Bessel
 * coefficients) and its own adaptive order per step (the exit test
 * |c_K| <w_K|w_K> < poly_tol per system: a system whose series has ended takes
 * no further terms while the others continue); all systems advance by one
 * H|psi> launch sequence per order.  No counterpart in the reference (one
 * system per run).  info->max_order is the largest order used; the intervals
 * must contain the spectra (EVR_ERR_STATE when a series diverges). */
int evr_multi_propagate(evr_multi *m, double dt, int nsteps, int npoly, double poly_tol,
                        const double *Hmin, const double *Hmax, double *psi, int device_ptrs,
                        evr_propa_info *info);

/* ---- optimal control: objective and adjoint gradient ----------------------- */
/* nf fields, each E_f(t) = sin^2(pi t/T) sum_{k=1..nk} a_kf sin(k pi t/T),
 * drive psi0 with H - E_f(t) mu by march_cheby steps of dt = T/nsteps (the
 * field of step n is its value at (n + 1/2) dt).  J_f = |<phi1|psi_f(T)>|^2,
 * the fidelity of calc_fidelity (CTRL:1000-1057).  The gradient of J_f with
 * respect to a_kf, to the mid-step field values and to the grid values of V
 * (the array of evr_plan_set_potential, so per term-grid point for SG4) is the
 * exact derivative of the computed propagation: a backward (costate) sweep
 * through the Chebyshev recurrences with the same H|psi>, from wavepackets
 * stored every `checkpoint` steps and recomputed.  The reference optimises
 * its field by Krotov / Zhu-Rabitz iterations instead (sub_Opt_control,
 * CTRL:290-520); the expansion and the gradient have no counterpart there. */
typedef struct evr_control_opts {
  double T;                /* final time                                        */
  int nsteps;              /* dt = T / nsteps                                   */
  int nk;                  /* number of field coefficients a_k                  */
  int checkpoint;          /* steps between stored wavepackets (0: 32)          */
  int npoly;               /* max Chebyshev terms per step                      */
  double poly_tol;
  double Hmin, Hmax;       /* interval of H (Hmin >= Hmax: estimated); it is
                              widened by max|E| * |mu|_max                      */
  int precision;           /* EVR_FP64; or EVR_FP32: the forward and backward
                              sweeps apply H and mu in FP32 (wavepackets,
                              recurrences and accumulations stay FP64), and J
                              is refined by one FP64 forward sweep; the
                              gradients are those of the FP32 sweeps           */
  int graph;               /* 1: forward steps after the first replayed as one
                              captured Chebyshev step (as evr_propa_opts.graph);
                              results identical to graph = 0                  */
} evr_control_opts;

typedef struct evr_control_info {
  int64_t hpsi;            /* real-column H|psi> applications (forward + backward) */
  int max_order;
  int retries;
  double Hmin, Hmax;
  double seconds_forward, seconds_backward;
} evr_control_info;

/* The same fields sampled for the time-dependent integrators of evr_propagate
 * (EVR_PROPA_TAYLOR: the derivatives E^(j)(t_n), j < nsub, as sub_dnE gives
 * them, Source_ElVibRot/sub_propagation/sub_module_field.f90:541; EVR_PROPA_RK4:
 * E(t_n), E(t_n + dt/2), E(t_n + dt), nsub = 3), t_n = n T/nsteps, in the
 * layout of evr_propa_opts.field_sub: field_sub[(n*nsub + j)*nf + f]. */
int evr_control_field_sub(const evr_control_opts *opts, int nf, const double *amp, int method, int nsub,
                          double *field_sub);
/* The reference's field update, build_field (CTRL:902-971) with the
 * Zhu-Rabitz / Krotov choice of sub_Opt_control (CTRL:290-520):
 * E_new(t_n) = [E_old(t_n) if krotov] - (1/alpha) s(t_n) Im(<psi|chi>
 * <chi|mu|psi>)(t_n), chi the target propagated back, s = envelopp(t, Tenv) =
 * sin^2(pi t/Tenv) (Tenv <= 0: s = 1), at the mid-steps t_n = (n + 1/2) dt of
 * the field evr_control_gradient holds; the bracket equals (dJ/dE_n)/(2 dt)
 * with grad_E from evr_control_gradient (nsteps x nf, field[n*nf + f]).
 * field_old / field_new: nsteps x nf. */
int evr_control_build_field(const evr_control_opts *opts, int nf, const double *grad_E, double alpha,
                            double Tenv, int krotov, const double *field_old, double *field_new);
/* Field values at the mid-steps: field[n*nf + f] (amp: nk x nf, column per field). */
int evr_control_field(const evr_control_opts *opts, int nf, const double *amp, double *field);
/* H: symmetric plan (hermitian=1 or type 1); mu: plan of the coupling on the
 * same basis (a type-0 plan whose V is mu(q)).  psi0, phi1: host nb_packed x 2
 * (Re, Im columns), shared by all fields.  Outputs (host): J (nf); grad_a
 * (nk x nf), grad_E (nsteps x nf: dJ/dE_n), grad_V (sum_nq x nf), each NULL
 * when not wanted (with all three NULL only the forward sweep runs). */
int evr_control_gradient(evr_plan *H, evr_plan *mu, const evr_control_opts *opts, int nf,
                         const double *amp, const double *psi0, const double *phi1, double *J,
                         double *grad_a, double *grad_E, double *grad_V, evr_control_info *info);

/* Propagation restart files, SaveWP_restart / ReadWP_restart
 * (Source_ElVibRot/sub_propagation/sub_module_propagation.f90:376-488): a
 * formatted list-directed file "T nb_WP size_WP" followed by each
 * wavepacket's complex coefficients (size_WP = nb0 x nb_packed values, here
 * one "(re,im)" per line).  psi: size_WP x (2 nwp) columns (Re, Im) as
 * evr_propagate.  Reading: *found = 0 and *T = 0 when the header cannot be
 * read (the reference's "No restart, T0=0"), EVR_ERR_IO when the sizes in the
 * file differ from (nwp, size_WP) (the reference STOPs). */
int evr_propa_save_restart(double T, int nwp, int64_t size_wp, const double *psi, const char *path);
int evr_propa_read_restart(const char *path, int nwp, int64_t size_wp, double *psi, double *T, int *found);

/* Extreme eigenvalues of a symmetric plan from `steps` Lanczos steps with full
 * re-orthogonalisation (min(steps, nb_packed) used), the estimate behind the
 * Chebyshev interval; the reference's sub_Hmax (HMAX:43-328) takes Hmin from
 * the minimum of the grid potential (HMAX:111-125) and Hmax from the last
 * basis function's <e_N|H|e_N> (HMAX:206-217), widened by DHmax. */
int evr_plan_bounds(evr_plan *p, int steps, double *Hmin, double *Hmax);

/* J_0(r) .. J_n(r) (out: n + 1 values), the Bessel functions behind the
 * Chebyshev coefficients (mmbsjn, POLY:1781), by Miller's backward recurrence. */
int evr_bessel_j(int n, double r, double *out);

#ifdef __cplusplus
}
#endif
#endif
