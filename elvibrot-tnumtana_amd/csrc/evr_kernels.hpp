// Device-side data layout shared by the plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace evr {

constexpr int kMaxDev = 12;   // modes supported by the term kernel (5-bit digits in a u64)
constexpr int kMaxNb = 31;    // largest nb of a 1D slot (5-bit basis digits)
constexpr int kMaxNq = 64;    // largest nq of a 1D slot
constexpr int kGenCache = 20; // slab contractions of up to this many inputs keep them in registers
constexpr int kKeoFactors = 7;
constexpr int kDp = 16;       // padded stride of the KEO coefficient strips

// Per Smolyak term, as the kernel needs it.
struct TermDev {
  int64_t off_nb;             // start of the term in the concatenated basis (map, R)
  int64_t off_nq;             // start of the term in the concatenated grid (V)
  int nbt, nqt;               // basis / grid size of the term
  int lds_mats;               // doubles of 1D matrices staged in LDS
  double w;                   // Smolyak weight (SGT2:754-854)
  int slot[kMaxDev];          // 1D slot of each mode
  // Reduced grid of the slab: the Dv modes vmode[0..Dv-1] with more than one
  // function.  H|psi> has the one-function modes summed into V_eff once per
  // potential (k_fold_v); the transforms expand / contract them on the fly.
  // Reduced modes 0..vm-1 are contracted inside an LDS slab (vnbA functions,
  // vQA points), the others (vnbB, vQB) through one Kronecker row per outer
  // point; vstep0: first of the 2 vm slab contraction steps (StepDev table).
  int Dv;
  int vmode[kMaxDev];
  int vm, vnbA, vQA, vnbB, vQB, vstep0;
  int64_t off_vq;             // start of the term in the concatenated V_eff
  // Kinetic part: the na modes amode[] carry non-vanishing P, Q, R factors;
  // the other modes (one function each) enter through scalar folds only.
  int na;
  int amode[kMaxDev];
  // Assembled operator: start of the term's tiled nb_t x nb_t matrix (MFMA
  // A-fragment order, see asm_index2 / asm_index4) in the assembled pools.
  int64_t off_asm;
  // Coefficient set of the term (index into the per-set KEO coefficients:
  // keoc + cix * keoc_stride, coefC + cix * D * D, coefF1 + cix * D).  All
  // terms share set 0 unless the plan compacts modes (more modes than the
  // kernel's positions): then each set holds the coefficients of one
  // position -> mode assignment.
  int cix;
};

// Position of element (row, col) of a term's assembled matrix (both pools):
// fragments of 16 rows x 4 columns (one K-step), each 64 values in the MFMA
// A-operand lane order (lane = 16 * (col % 4) + row % 16), with two
// consecutive K-steps interleaved per lane, K-step pairs ordered
// [row / 16][col / 8]: a lane reads the A operands of two MFMAs with one
// 16-byte (FP64) or 8-byte (FP32) load.
__host__ __device__ __forceinline__ int64_t asm_index2(int row, int col, int ks)
{
  const int kstep = col >> 2;
  return (((int64_t)(row >> 4) * (ks >> 1) + (kstep >> 1)) * 64 + (col & 3) * 16 + (row & 15)) * 2 +
         (kstep & 1);
}

// The same for the FP32 pool with four consecutive K-steps per lane: a lane
// reads the A operands of a group of four MFMAs with one 16-byte load.
__host__ __device__ __forceinline__ int64_t asm_index4(int row, int col, int ks)
{
  const int kstep = col >> 2;
  return (((int64_t)(row >> 4) * (ks >> 2) + (kstep >> 2)) * 64 + (col & 3) * 16 + (row & 15)) * 4 +
         (kstep & 3);
}

// K-steps (4 columns each) stored per 16-row tile of an nb_t x nb_t matrix:
// rounded up to a multiple of 4 (zero columns), so the apply kernel reads
// whole groups of four fragments without bounds tests.
__host__ __device__ __forceinline__ int asm_ksteps(int nbt) { return (((nbt + 3) >> 2) + 3) & ~3; }

// One contraction step inside a slab: out[i1 + n1*(q + nout*i3)] =
// sum_k M[q + nout*k] in[i1 + n1*(k + nin*i3)], M at pool offset `mat`.
struct StepDev {
  int n1, n3, nin, nout;
  float rn1;                  // 1/(n1*NVW) is formed on the device
  int pad;
  int64_t mat;
};

// Per 1D slot (mode, level): offsets into the matrix pools.
struct SlotDev {
  int nb, nq;
  int64_t offB;    // B: nq x nb  (pool "bw")
  int64_t offW;    // W: nb x nq  (pool "bw")
  int64_t offK;    // 7 factors of nb x nb (pool "keo")
  int64_t offX;    // nq grid points (pool "x")
  int64_t offD1;   // dnRGG%d1: nq x nq (pool "d1", grid form of type 10 with G per point)
};

// kModeAssemble: H|psi> of the term's own unit vectors e_c, c = c0 .. c0+NVW-1,
// written as columns of the term's assembled matrix.
enum TermMode { kModeHpsi = 0, kModeB2G = 1, kModeG2B = 2, kModeAssemble = 3 };

template <typename T>
struct TermParams {
  int D;
  int mode;
  int hasV, hasKeo, hasA, hasF1, diagT2;
  int diagKeo;                // G2B mode: add the kinetic diagonal (plan diagonal)
  int nvec;                   // vectors in this call (blockIdx decode)
  int64_t nunits;             // padded count of terms in this launch (multiple of 8)
  const int *order;           // term ids for this launch (-1 = padding)
  const TermDev *terms;
  const SlotDev *slots;
  const StepDev *steps;
  const T *bw;                // pool of B and W matrices
  const T *keo;               // pool of basis-space KEO factors
  const int *map;             // term-local -> packed
  const int *rowpos;          // term row -> its row of R (scatter order), -1 if pruned
  const T *V;                 // grid potential (term-concatenated)
  const double *coefC;        // D x D: c[i*D+j]
  const double *coefA;        // D^3: a[(i*D+j)*D+k]
  const double *coefF1;       // D
  const double *keoc;         // pre-zeroed coefficient strips (see plan), one block per set
  int64_t keoc_stride;        // doubles per coefficient set in keoc
  const T *X;                 // packed input, [nb_packed][nvec]
  T *R;                       // term results in scatter order, [rows][nvec]
  T *G;                       // grid output (B2G), column v at G + v*gld
  int64_t gld;                // column stride of G (sum nq_t)
  const int2 *items;          // assemble: (term, first column) per workgroup
  double *A64;                // assemble: FP64 pool (or NULL)
  float *A32;                 // assemble: FP32 pool (or NULL)
  const T *Gin;               // grid input (G2B), [sum nq_t]
};

}  // namespace evr
