// Eigen-solver driver over the H|psi> engine: block Davidson with FP64
// Rayleigh-Ritz, two-pass block orthogonalisation and optional FP32 H|psi>
// of the new vectors with FP64 defect corrections on the Ritz vectors.
//
// Follows the reference's sub_propagation_Davidson
// (Source_ElVibRot/sub_propagation/sub_module_Davidson.f90:54-748): one new
// vector per unconverged wanted state and iteration (residual_max_nb,
// DAV:1290-1588), convergence on |dE| < conv_ene and ||r|| < conv_resi
// (DAV:407-446), collapse of the subspace on the current Ritz vectors when it
// is full (DAV:516-606), NewVec_type 2 (Davidson correction r/(E - Hdiag),
// DAV:1392-1413) when a diagonal is given, plain residuals (type 4) otherwise.
#include <hip/hip_runtime.h>

#include <map>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/evr_hip.h"
#include "evr_internal.hpp"
#include "evr_propa.hpp"

using OpRef = evr::propa::OpRef;

namespace {

using evr::set_error;

#define HTRY(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return set_error(EVR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define BTRY(expr)                                                                           \
  do {                                                                                       \
    rocblas_status s_ = (expr);                                                              \
    if (s_ != rocblas_status_success)                                                        \
      return set_error(EVR_ERR_HIP, std::string(#expr) + ": rocblas status " + std::to_string((int)s_)); \
  } while (0)
#define RTRY(expr)                                                                           \
  do {                                                                                       \
    int r_ = (expr);                                                                         \
    if (r_ != EVR_OK) return r_;                                                             \
  } while (0)

struct DBuf {
  double *p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t k)
  {
    const hipError_t e = k ? hipMalloc((void **)&p, sizeof(double) * k) : hipSuccess;
    n = e == hipSuccess ? k : 0;
    if (e != hipSuccess) p = nullptr;
    return e;
  }
  ~DBuf()
  {
    if (p) (void)hipFree(p);
  }
};

struct Blas {
  rocblas_handle h = nullptr;
  ~Blas()
  {
    if (h) rocblas_destroy_handle(h);
  }
};

unsigned nblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

// R[:, c] = XH[:, sel[c]] - theta[sel[c]] XV[:, sel[c]]
__global__ void k_residual(int64_t nb, int nc, const int *sel, const double *theta, const double *XH,
                           const double *XV, double *R)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nc;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(g / nb);
    const int64_t i = g - (int64_t)c * nb;
    const int j = sel[c];
    R[g] = XH[i + nb * j] - theta[j] * XV[i + nb * j];
  }
}

// squared column norms in two passes of fixed order: blocks (segment s,
// column c) write partial sums part[c * kNormSeg + s], then one thread per
// column adds its segments in order (a long column spread over many blocks)
constexpr int kNormSeg = 64;
__global__ __launch_bounds__(256) void k_colnorm2_part(int64_t nb, const double *A, double *part)
{
  __shared__ double red[256];
  const int s = blockIdx.x, c = blockIdx.y;
  const int64_t len = (nb + kNormSeg - 1) / kNormSeg;
  const int64_t i0 = s * len, i1 = i0 + len < nb ? i0 + len : nb;
  const double *a = A + nb * c;
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc = fma(a[i], a[i], acc);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)c * kNormSeg + s] = red[0];
}

__global__ void k_colnorm2_fin(int nc, const double *part, double *out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double acc = 0.0;
  for (int s = 0; s < kNormSeg; ++s) acc += part[(int64_t)c * kNormSeg + s];
  out[c] = acc;
}

// C[i + ldc j] = sum over b (in order) of P[b m n + i + m j]: the partial
// products of a split inner-product gemm
__global__ void k_sum_parts(int m, int n, int nbat, const double *P, double *C, int ldc)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < (int64_t)m * n;
       g += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int b = 0; b < nbat; ++b) acc += P[(int64_t)b * m * n + g];
    const int i = (int)(g % m), j = (int)(g / m);
    C[i + (int64_t)ldc * j] = acc;
  }
}

// Rayleigh-Ritz matrix, columns m0..m-1 (and the matching rows) from
// C1 = V^T HV[:, m0:m] (m x nn) and C2 = V[:, m0:m]^T HV (nn x m):
// S(i, j) = S(j, i) = (C1(i, j - m0) + C2(j - m0, i)) / 2, the symmetric part.
// Columns below nex hold FP64 products H v (mixed precision: the others are
// FP32 products): a coupling of such a column with another one is taken from
// the FP64 product alone, v_j^T (H v_i) with i < nex <= j, so the FP32 error
// of H v_j enters the Ritz vectors of the wanted states only through their
// small components on v_j.
__global__ void k_place_sym(int m, int m0, const double *C1, const double *C2, double *S, int lds, int nex)
{
  const int nn = m - m0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < (int64_t)m * nn;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(g % m), jj = (int)(g / m), j = m0 + jj;
    const double a = C1[i + (int64_t)m * jj], b = C2[jj + (int64_t)nn * i];
    const double v = (i < nex && j >= nex) ? b : (j < nex && i >= nex) ? a : 0.5 * (a + b);
    S[i + (int64_t)lds * j] = v;
    S[j + (int64_t)lds * i] = v;
  }
}

// Davidson correction t = r / (theta - diag), guarded near zero
__global__ void k_precond(int64_t nb, int nc, const double *shift, const double *diag, double *T)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nc;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(g / nb);
    const int64_t i = g - (int64_t)c * nb;
    double den = shift[c] - diag[i];
    const double floor_ = 1e-8 * (fabs(shift[c]) + 1.0);
    if (fabs(den) < floor_) den = den < 0 ? -floor_ : floor_;
    T[g] /= den;
  }
}

__global__ void k_unit_cols(int64_t nb, int nc, const int64_t *idx, double *V)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nc;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(g / nb);
    const int64_t i = g - (int64_t)c * nb;
    V[g] = (i == idx[c]) ? 1.0 : 0.0;
  }
}

// cyclic Jacobi on a small symmetric matrix (host), eigenvalues ascending
void jacobi_sym(int n, std::vector<double> A, std::vector<double> &ev, std::vector<double> &V)
{
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += A[p + (size_t)n * p] * A[p + (size_t)n * p];
      for (int q = p + 1; q < n; ++q) off += A[p + (size_t)n * q] * A[p + (size_t)n * q];
    }
    if (off <= 1e-30 * diag) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p + (size_t)n * q];
        if (apq == 0.0) continue;
        const double th = (A[q + (size_t)n * q] - A[p + (size_t)n * p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
          A[k + (size_t)n * p] = c * akp - s * akq;
          A[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
          A[p + (size_t)n * k] = c * apk - s * aqk;
          A[q + (size_t)n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
          V[k + (size_t)n * p] = c * vkp - s * vkq;
          V[k + (size_t)n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return A[a + (size_t)n * a] < A[b + (size_t)n * b]; });
  ev.resize(n);
  std::vector<double> W((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    ev[k] = A[ord[k] + (size_t)n * ord[k]];
    for (int i = 0; i < n; ++i) W[i + (size_t)n * k] = V[i + (size_t)n * ord[k]];
  }
  V.swap(W);
}

struct Dav {
  OpRef op;  // one plan, or coupled electronic states
  hipStream_t st = nullptr;
  rocblas_handle bh = nullptr;
  int64_t nb = 0;
  int max_nvec = 0;
  int64_t hpsi64 = 0, hpsi32 = 0;

  // HX = H X for nc columns (chunks of the plan's max_nvec)
  int apply(const double *X, double *HX, int nc, int flags, bool count = true)
  {
    for (int c0 = 0; c0 < nc; c0 += max_nvec) {
      const int k = std::min(max_nvec, nc - c0);
      RTRY(op.plan ? evr_oppsi(op.plan, k, X + nb * c0, HX + nb * c0, flags | EVR_DEVICE_PTRS)
                   : evr_coupled_oppsi(op.coupled, k, X + nb * c0, HX + nb * c0, flags | EVR_DEVICE_PTRS));
      if (count) {
        if (flags & EVR_FP32) hpsi32 += k; else hpsi64 += k;
      }
    }
    return EVR_OK;
  }
  // The same for the iteration's new-vector block (X, HX the fixed buffers of
  // that block) as a captured HIP graph per (columns, precision): the first
  // call runs eagerly (it also makes lazy state such as the FP32 potential
  // copy) and records the launches; later calls replay them on the stream.
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  int apply_captured(const double *X, double *HX, int nc, int flags)
  {
    const auto key = std::make_pair(nc, flags);
    auto it = graphs.find(key);
    if (it != graphs.end()) {
      HTRY(hipGraphLaunch(it->second, st));
      if (flags & EVR_FP32) hpsi32 += nc; else hpsi64 += nc;
      return EVR_OK;
    }
    RTRY(apply(X, HX, nc, flags));
    hipGraph_t g = nullptr;
    HTRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = apply(X, HX, nc, flags, false);
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != EVR_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc != EVR_OK) return rc;
      HTRY(e);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HTRY(ei);
    graphs[key] = ex;
    return EVR_OK;
  }
  ~Dav()
  {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
  }

  // C (m x n) = A^T B with A nb x m, B nb x n.  The result is small and the
  // reduction long (nb rows), so the rows are split into nbat chunks
  // (strided-batched gemm, enough workgroups for the whole GPU) and the
  // partial products summed in a fixed order.
  DBuf part;
  int gemm_tn(int m, int n, const double *A, const double *B, double *C, int ldc)
  {
    const double one = 1.0, zero = 0.0;
    const int64_t tiles = (int64_t)((m + 127) / 128) * ((n + 127) / 128);
    int nbat = (int)std::min<int64_t>(64, std::max<int64_t>(1, 1024 / tiles));
    while (nbat > 1 && nb / nbat < 4096) nbat /= 2;
    if (nbat == 1) {
      BTRY(rocblas_dgemm(bh, rocblas_operation_transpose, rocblas_operation_none, m, n, (int)nb, &one, A,
                         (int)nb, B, (int)nb, &zero, C, ldc));
      return EVR_OK;
    }
    const size_t need = (size_t)nbat * m * n;
    if (part.n < need) {
      if (part.p) HTRY(hipFree(part.p));
      part.p = nullptr;
      HTRY(part.alloc(need));
    }
    const int64_t chunk = nb / nbat, rem = nb - chunk * nbat;
    BTRY(rocblas_dgemm_strided_batched(bh, rocblas_operation_transpose, rocblas_operation_none, m, n,
                                       (int)chunk, &one, A, (int)nb, chunk, B, (int)nb, chunk, &zero,
                                       part.p, m, (int64_t)m * n, nbat));
    if (rem)  // the last rows, into the first partial product
      BTRY(rocblas_dgemm(bh, rocblas_operation_transpose, rocblas_operation_none, m, n, (int)rem, &one,
                         A + chunk * nbat, (int)nb, B + chunk * nbat, (int)nb, &one, part.p, m));
    hipLaunchKernelGGL(k_sum_parts, dim3(nblocks((int64_t)m * n)), dim3(256), 0, st, m, n, nbat, part.p, C,
                       ldc);
    HTRY(hipGetLastError());
    return EVR_OK;
  }
  // out[c] = |A[:, c]|^2 for nc columns of nb rows
  DBuf npart;
  int colnorm2(const double *A, int nc, double *out)
  {
    if (npart.n < (size_t)nc * kNormSeg) {
      if (npart.p) HTRY(hipFree(npart.p));
      npart.p = nullptr;
      HTRY(npart.alloc((size_t)nc * kNormSeg));
    }
    hipLaunchKernelGGL(k_colnorm2_part, dim3(kNormSeg, nc), dim3(256), 0, st, nb, A, npart.p);
    hipLaunchKernelGGL(k_colnorm2_fin, dim3((nc + 63) / 64), dim3(64), 0, st, nc, npart.p, out);
    HTRY(hipGetLastError());
    return EVR_OK;
  }
  // C (nb x n) = alpha A (nb x m) B (m x n) + beta C
  int gemm_nn(int m, int n, double alpha, const double *A, const double *B, int ldb, double beta,
              double *C)
  {
    BTRY(rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_none, (int)nb, n, m, &alpha, A,
                       (int)nb, B, ldb, &beta, C, (int)nb));
    return EVR_OK;
  }

  // Orthonormalise the nc columns of T against V[:, :m] and among themselves
  // (classical Gram-Schmidt twice, then SVQB twice dropping dependent
  // directions).  Returns the number of columns kept (compacted to the front).
  int orthonormalise(const double *V, int m, double *T, int nc, double *tmpC, double *tmpT,
                     int &kept)
  {
    for (int pass = 0; pass < 2 && m > 0; ++pass) {
      RTRY(gemm_tn(m, nc, V, T, tmpC, m));
      RTRY(gemm_nn(m, nc, -1.0, V, tmpC, m, 1.0, T));
    }
    kept = nc;
    for (int pass = 0; pass < 2 && kept > 0; ++pass) {
      RTRY(gemm_tn(kept, kept, T, T, tmpC, kept));
      std::vector<double> G((size_t)kept * kept), ev, U;
      HTRY(hipMemcpyAsync(G.data(), tmpC, sizeof(double) * kept * kept, hipMemcpyDeviceToHost, st));
      HTRY(hipStreamSynchronize(st));
      // SVQB: scale to unit diagonal, keep the well-conditioned directions
      std::vector<double> sc(kept);
      for (int i = 0; i < kept; ++i) {
        const double g = G[i + (size_t)kept * i];
        sc[i] = g > 1e-300 ? 1.0 / std::sqrt(g) : 0.0;
      }
      for (int i = 0; i < kept; ++i)
        for (int j = 0; j < kept; ++j) G[i + (size_t)kept * j] *= sc[i] * sc[j];
      jacobi_sym(kept, G, ev, U);
      std::vector<double> M;
      int nk = 0;
      for (int k = kept - 1; k >= 0; --k) {
        if (ev[k] < 1e-10) continue;
        for (int i = 0; i < kept; ++i) M.push_back(sc[i] * U[i + (size_t)kept * k] / std::sqrt(ev[k]));
        ++nk;
      }
      if (nk == 0) { kept = 0; break; }
      HTRY(evr::upload_now(tmpC, M.data(), sizeof(double) * M.size(), st));
      RTRY(gemm_nn(kept, nk, 1.0, T, tmpC, kept, 0.0, tmpT));
      HTRY(hipMemcpyAsync(T, tmpT, sizeof(double) * nb * nk, hipMemcpyDeviceToDevice, st));
      kept = nk;
    }
    return EVR_OK;
  }
};

}  // namespace

static int davidson_impl(OpRef op, const evr_davidson_opts *o, double *evals, double *evecs,
                         evr_davidson_info *info);

extern "C" int evr_davidson(evr_plan *plan, const evr_davidson_opts *o, double *evals, double *evecs,
                            evr_davidson_info *info)
{
  if (!plan) return set_error(EVR_ERR_ARG, "NULL argument");
  return davidson_impl(OpRef(plan), o, evals, evecs, info);
}

extern "C" int evr_coupled_davidson(evr_coupled *C, const evr_davidson_opts *o, double *evals, double *evecs,
                                    evr_davidson_info *info)
{
  if (!C) return set_error(EVR_ERR_ARG, "NULL argument");
  if (o && o->precision == EVR_FP32)
    return set_error(EVR_ERR_ARG, "coupled-state Davidson is FP64 only (precision EVR_FP64)");
  return davidson_impl(OpRef(C), o, evals, evecs, info);
}

static int davidson_impl(OpRef op, const evr_davidson_opts *o, double *evals, double *evecs,
                         evr_davidson_info *info)
{
  if (!o || !evals) return set_error(EVR_ERR_ARG, "NULL argument");
  const auto t_start = std::chrono::steady_clock::now();
  Dav d;
  d.op = op;
  RTRY(op.rows(&d.nb));
  d.st = op.stream();
  const int nev = o->nev;
  const int blk = o->block > 0 ? o->block : std::min(nev, 16);
  if (nev < 1 || nev > d.nb) return set_error(EVR_ERR_ARG, "nev out of range");
  // default subspace: 4 nev (at 200 levels of D=9 L=8, 800 columns take 68
  // iterations where 600 take 78 and 400 take 132; the larger Rayleigh-Ritz
  // problem costs less than the extra iterations)
  int mmax = o->max_sub > 0 ? o->max_sub : std::max(4 * nev, nev + 4 * blk);
  mmax = (int)std::min<int64_t>(mmax, d.nb);
  // a collapse keeps the Ritz vectors of every wanted state and makes room
  // for one block of new vectors (DAV:516-606 keeps the wanted states)
  if (mmax < nev + blk && mmax < d.nb)
    return set_error(EVR_ERR_ARG, "max_sub must be at least nev + block (" + std::to_string(nev + blk) + ")");
  const int max_it = o->max_it > 0 ? o->max_it : 200;
  d.max_nvec = (int)o->max_nvec_hint;
  if (d.max_nvec <= 0) d.max_nvec = blk;
  Blas blas;
  BTRY(rocblas_create_handle(&blas.h));
  BTRY(rocblas_set_stream(blas.h, d.st));
  BTRY(rocblas_set_pointer_mode(blas.h, rocblas_pointer_mode_host));
  // no atomics in the library's gemms (and rocSOLVER's, on this handle): the
  // same inputs give the same iterates run after run, which the mixed
  // driver's stall test (a 10% drop within 6 iterations) is sensitive to
  BTRY(rocblas_set_atomics_mode(blas.h, rocblas_atomics_not_allowed));
  d.bh = blas.h;

  const int64_t nb = d.nb;
  const int kmax = std::min(mmax, nev + blk);
  DBuf V, HV, XV, XH, T, HT, S, Sk, C2, theta, E, tmpC, tmpT, nrm, dsel, dshift, diag, didx;
  int *dinfo = nullptr;
  HTRY(V.alloc(nb * mmax));
  HTRY(HV.alloc(nb * mmax));
  HTRY(XV.alloc(nb * kmax));
  HTRY(XH.alloc(nb * kmax));
  HTRY(T.alloc(nb * blk));
  HTRY(HT.alloc(nb * blk));
  HTRY(tmpT.alloc(nb * std::max(blk, 1)));
  HTRY(S.alloc((size_t)mmax * mmax));
  HTRY(Sk.alloc((size_t)mmax * mmax));
  HTRY(C2.alloc((size_t)mmax * blk));
  HTRY(theta.alloc(mmax));
  HTRY(E.alloc(mmax));
  HTRY(tmpC.alloc((size_t)mmax * std::max(blk, kmax)));
  HTRY(nrm.alloc(blk));
  HTRY(dsel.alloc(blk));
  HTRY(dshift.alloc(blk));
  HTRY(hipMalloc((void **)&dinfo, sizeof(int)));
  struct IntFree { int *p; ~IntFree() { if (p) (void)hipFree(p); } } dinfo_free{dinfo};

  // preconditioner diagonal
  const bool use_diag = o->precond != 0;
  std::vector<double> hdiag;
  if (use_diag || !(o->guess || o->guess_file)) {
    HTRY(diag.alloc(nb));
    if (o->diag) {
      HTRY(hipMemcpyAsync(diag.p, o->diag, sizeof(double) * nb, hipMemcpyHostToDevice, d.st));
      hdiag.assign(o->diag, o->diag + nb);
    } else {
      RTRY(op.plan ? evr_plan_diagonal(op.plan, diag.p, EVR_DEVICE_PTRS)
                   : evr_coupled_diagonal(op.coupled, diag.p, EVR_DEVICE_PTRS));
      hdiag.resize(nb);
      HTRY(hipMemcpyAsync(hdiag.data(), diag.p, sizeof(double) * nb, hipMemcpyDeviceToHost, d.st));
      HTRY(hipStreamSynchronize(d.st));
    }
  }

  // wavepacket files: the basis descriptor maps file rows to packed rows
  int nb0 = 1;
  if (o->guess_file || o->save_file) {
    if (!o->basis) return set_error(EVR_ERR_ARG, "guess_file / save_file need opts->basis");
    if (o->basis->nb_packed < 1 || nb % o->basis->nb_packed)
      return set_error(EVR_ERR_ARG, "opts->basis does not match the operator's vector length");
    nb0 = (int)(nb / o->basis->nb_packed);
  }
  std::vector<double> file_guess;
  int nfile = 0;
  if (o->guess_file) {
    file_guess.resize((size_t)nb * mmax);
    RTRY(evr_psi_read(o->basis, nb0, 0, mmax, file_guess.data(), &nfile, o->guess_file));
    if (nfile < 1) return set_error(EVR_ERR_IO, std::string("no wavepacket in ") + o->guess_file);
  }
  const double *guess = o->guess_file ? file_guess.data() : o->guess;
  const int nguess = o->guess_file ? nfile : o->nguess;

  // initial subspace: given vectors, else unit vectors on the lowest diagonal entries
  int m = 0;
  int k0 = guess ? std::min<int>(nguess, mmax) : std::min(mmax, std::max(nev, blk));
  if (guess) {
    HTRY(hipMemcpyAsync(V.p, guess, sizeof(double) * nb * k0, hipMemcpyHostToDevice, d.st));
  } else {
    std::vector<int64_t> idx(nb);
    for (int64_t i = 0; i < nb; ++i) idx[i] = i;
    std::partial_sort(idx.begin(), idx.begin() + k0, idx.end(),
                      [&](int64_t a, int64_t b) { return hdiag[a] < hdiag[b] || (hdiag[a] == hdiag[b] && a < b); });
    DBuf didx2;
    HTRY(didx2.alloc(k0));
    HTRY(evr::upload_now(didx2.p, idx.data(), sizeof(int64_t) * k0, d.st));
    hipLaunchKernelGGL(k_unit_cols, dim3(nblocks(nb * k0)), dim3(256), 0, d.st, nb, k0,
                       (const int64_t *)didx2.p, V.p);
    HTRY(hipGetLastError());
    HTRY(hipStreamSynchronize(d.st));
  }
  {
    // orthonormalise the start vectors in chunks of blk
    int done = 0;
    for (int c0 = 0; c0 < k0; c0 += blk) {
      const int nc = std::min(blk, k0 - c0);
      HTRY(hipMemcpyAsync(T.p, V.p + nb * c0, sizeof(double) * nb * nc, hipMemcpyDeviceToDevice, d.st));
      int kept = 0;
      RTRY(d.orthonormalise(V.p, done, T.p, nc, tmpC.p, tmpT.p, kept));
      HTRY(hipMemcpyAsync(V.p + nb * done, T.p, sizeof(double) * nb * kept, hipMemcpyDeviceToDevice, d.st));
      done += kept;
    }
    m = done;
  }
  // Precision.  EVR_FP64: every H|psi> in FP64.  EVR_FP32 (mixed): every
  // new vector's H|psi> in FP32; the subspace algebra (Rayleigh-Ritz,
  // orthonormalisation, residuals) stays FP64, and FP64 enters as a defect
  // correction: the subspace is collapsed on the Ritz vectors and H applied
  // to them in FP64 when the residuals computed from the mixed H V have
  // converged, stopped falling (the FP32 floor) or first dropped below
  // refine_resi.  The leading nex columns of HV are then FP64 products; their
  // couplings with later (FP32) columns are taken from the FP64 side
  // (k_place_sym), so the Ritz vectors of the wanted states keep converging
  // below the FP32 floor between corrections, and a collapse keeps them as
  // nearly exact columns.  Convergence is declared only on an all-FP64 H V
  // (right after a correction), so the reported residuals are FP64 residuals.
  const bool mixed = o->precision == EVR_FP32;
  const int flags = mixed ? EVR_FP32 : EVR_FP64;
  RTRY(d.apply(V.p, HV.p, m, flags));
  int nex = mixed ? 0 : m;  // leading columns of HV that are FP64 products
  int ncorr = 0;            // FP64 defect corrections made

  int m_s = 0;  // columns of V whose rows and columns are in Sk
  std::vector<double> th(mmax), th_old(nev, 1e300), rn(nev, 1e300);
  std::vector<int> conv(nev, 0);
  int it = 0, nconv = 0;
  double maxres = 0.0;
  bool done = false;
  bool fresh = false;  // this Rayleigh-Ritz step follows a defect correction
  double best_res = 1e300;
  int stall = 0;
  for (it = 1; it <= max_it; ++it) {
    // Rayleigh-Ritz in FP64.  Sk keeps the symmetric part of V^T H V over
    // the steps: only the columns appended since the last step (and their
    // rows) are computed, all of them after a collapse or the FP64 rebuild
    // of H V; the eigensolver works on a copy.
    if (m_s == 0) {
      RTRY(d.gemm_tn(m, m, V.p, HV.p, S.p, m));
      hipLaunchKernelGGL(k_place_sym, dim3(nblocks((int64_t)m * m)), dim3(256), 0, d.st, m, 0, S.p, S.p,
                         Sk.p, mmax, nex);
    } else if (m > m_s) {
      const int nn = m - m_s;
      RTRY(d.gemm_tn(m, nn, V.p, HV.p + nb * m_s, tmpC.p, m));
      RTRY(d.gemm_tn(nn, m, V.p + nb * m_s, HV.p, C2.p, nn));
      hipLaunchKernelGGL(k_place_sym, dim3(nblocks((int64_t)m * nn)), dim3(256), 0, d.st, m, m_s, tmpC.p,
                         C2.p, Sk.p, mmax, nex);
    }
    HTRY(hipGetLastError());
    m_s = m;
    HTRY(hipMemcpy2DAsync(S.p, sizeof(double) * m, Sk.p, sizeof(double) * mmax, sizeof(double) * m, m,
                          hipMemcpyDeviceToDevice, d.st));
    BTRY(rocsolver_dsyevd(d.bh, rocblas_evect_original, rocblas_fill_upper, m, S.p, m, theta.p, E.p,
                          dinfo));
    HTRY(hipMemcpyAsync(th.data(), theta.p, sizeof(double) * m, hipMemcpyDeviceToHost, d.st));
    int hinfo = 0;
    HTRY(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, d.st));
    HTRY(hipStreamSynchronize(d.st));
    if (hinfo != 0) return set_error(EVR_ERR_STATE, "subspace eigensolver failed");
    const int kr = std::min(m, kmax);
    const int ne = std::min(nev, m);
    // Ritz vectors and their H images: the ne wanted ones every step, the
    // kr - ne extra ones only when a collapse keeps them
    RTRY(d.gemm_nn(m, ne, 1.0, V.p, S.p, m, 0.0, XV.p));
    RTRY(d.gemm_nn(m, ne, 1.0, HV.p, S.p, m, 0.0, XH.p));
    auto ritz_extra = [&]() -> int {
      if (kr > ne) {
        RTRY(d.gemm_nn(m, kr - ne, 1.0, V.p, S.p + (size_t)m * ne, m, 0.0, XV.p + nb * ne));
        RTRY(d.gemm_nn(m, kr - ne, 1.0, HV.p, S.p + (size_t)m * ne, m, 0.0, XH.p + nb * ne));
      }
      return EVR_OK;
    };
    // residuals of the wanted states, in chunks of blk
    maxres = 0.0;
    for (int j0 = 0; j0 < ne; j0 += blk) {
      const int nc = std::min(blk, ne - j0);
      std::vector<int> sel(nc);
      for (int c = 0; c < nc; ++c) sel[c] = j0 + c;
      HTRY(evr::upload_now(dsel.p, sel.data(), sizeof(int) * nc, d.st));
      hipLaunchKernelGGL(k_residual, dim3(nblocks(nb * nc)), dim3(256), 0, d.st, nb, nc,
                         (const int *)dsel.p, theta.p, XH.p, XV.p, T.p);
      HTRY(hipGetLastError());
      RTRY(d.colnorm2(T.p, nc, nrm.p));
      std::vector<double> r2(nc);
      HTRY(hipMemcpyAsync(r2.data(), nrm.p, sizeof(double) * nc, hipMemcpyDeviceToHost, d.st));
      HTRY(hipStreamSynchronize(d.st));
      for (int c = 0; c < nc; ++c) {
        rn[j0 + c] = std::sqrt(r2[c]);
        maxres = std::max(maxres, rn[j0 + c]);
      }
    }
    // convergence: |r| < conv_resi and |E(it) - E(it-1)| < conv_ene
    // (DAV:407-446).  Right after a defect correction the energy change
    // measures the FP32 error of the previous step, not the iteration, so
    // the FP64 residual alone decides there (|dE| <= |r|^2 / gap).
    nconv = 0;
    for (int j = 0; j < ne; ++j) {
      conv[j] = (rn[j] < o->conv_resi && (fresh || std::fabs(th[j] - th_old[j]) < o->conv_ene)) ? 1 : 0;
      nconv += conv[j];
      th_old[j] = th[j];
    }
    fresh = false;
    if (o->verbose) {
      fprintf(stderr, "davidson it %d m %d (fp64 cols %d) conv %d/%d max|r| %.3e E0 %.10f\n", it, m, nex,
              nconv, nev, maxres, th[0]);
      if (o->verbose > 1)
        for (int j = 0; j < ne; ++j) fprintf(stderr, "   %3d %.10f %.3e\n", j, th[j], rn[j]);
    }
    const bool all_conv = nconv == ne && ne == nev;
    if (all_conv && nex == m) { done = true; break; }
    if (mixed) {
      if (maxres < 0.9 * best_res) {
        best_res = maxres;
        stall = 0;
      } else {
        ++stall;
      }
      if (all_conv || stall >= 6 || (ncorr == 0 && maxres < o->refine_resi)) {
        // defect correction: collapse on the Ritz vectors and apply H in FP64
        const int keep = kr;
        RTRY(ritz_extra());
        HTRY(hipMemcpyAsync(V.p, XV.p, sizeof(double) * nb * keep, hipMemcpyDeviceToDevice, d.st));
        RTRY(d.apply(V.p, HV.p, keep, EVR_FP64));
        m = keep;
        nex = keep;
        m_s = 0;
        ++ncorr;
        fresh = true;
        best_res = 1e300;
        stall = 0;
        continue;
      }
    }
    // new vectors: one per unconverged wanted state (DAV:1290-1588); when there
    // are more than the block holds, the states with the largest residuals
    std::vector<int> sel;
    for (int j = 0; j < ne; ++j)
      if (!conv[j]) sel.push_back(j);
    if ((int)sel.size() > blk) {
      std::stable_sort(sel.begin(), sel.end(), [&](int a, int b) { return rn[a] > rn[b]; });
      sel.resize(blk);
      std::sort(sel.begin(), sel.end());
    }
    if (sel.empty()) break;  // all converged on a mixed H V: handled above
    int nc = (int)sel.size();
    if (m + nc > mmax) {
      // collapse on the current Ritz vectors (XV, XH hold kr of them; at
      // least nev, since mmax >= nev + block)
      const int keep = std::min(kr, mmax - nc);
      RTRY(ritz_extra());
      HTRY(hipMemcpyAsync(V.p, XV.p, sizeof(double) * nb * keep, hipMemcpyDeviceToDevice, d.st));
      HTRY(hipMemcpyAsync(HV.p, XH.p, sizeof(double) * nb * keep, hipMemcpyDeviceToDevice, d.st));
      // FP64: all kept columns exact.  Mixed: the wanted Ritz vectors stay
      // nearly exact when the subspace had FP64 columns (their FP32 parts
      // carry small coefficients); the other kept vectors are FP32 products.
      nex = !mixed ? keep : (nex > 0 ? std::min(keep, nev) : 0);
      m = keep;
      m_s = 0;
    }
    HTRY(evr::upload_now(dsel.p, sel.data(), sizeof(int) * nc, d.st));
    hipLaunchKernelGGL(k_residual, dim3(nblocks(nb * nc)), dim3(256), 0, d.st, nb, nc,
                       (const int *)dsel.p, theta.p, XH.p, XV.p, T.p);
    HTRY(hipGetLastError());
    if (use_diag) {
      std::vector<double> sh(nc);
      for (int c = 0; c < nc; ++c) sh[c] = th[sel[c]];
      HTRY(evr::upload_now(dshift.p, sh.data(), sizeof(double) * nc, d.st));
      hipLaunchKernelGGL(k_precond, dim3(nblocks(nb * nc)), dim3(256), 0, d.st, nb, nc, dshift.p,
                         diag.p, T.p);
      HTRY(hipGetLastError());
    }
    int kept = 0;
    RTRY(d.orthonormalise(V.p, m, T.p, nc, tmpC.p, tmpT.p, kept));
    if (o->verbose) fprintf(stderr, "  new vectors %d kept %d (m %d)\n", nc, kept, m);
    if (kept == 0) break;  // the subspace cannot grow (DAV:1534 drops tiny vectors)
    if (o->graph) RTRY(d.apply_captured(T.p, HT.p, kept, flags));
    else RTRY(d.apply(T.p, HT.p, kept, flags));
    HTRY(hipMemcpyAsync(V.p + nb * m, T.p, sizeof(double) * nb * kept, hipMemcpyDeviceToDevice, d.st));
    HTRY(hipMemcpyAsync(HV.p + nb * m, HT.p, sizeof(double) * nb * kept, hipMemcpyDeviceToDevice, d.st));
    if (!mixed) nex = m + kept;
    m += kept;
  }
  // results: Ritz values and vectors of the last Rayleigh-Ritz step
  for (int j = 0; j < nev; ++j) evals[j] = j < m ? th[j] : NAN;
  if (evecs)
    HTRY(hipMemcpyAsync(evecs, XV.p, sizeof(double) * nb * std::min(nev, m), hipMemcpyDeviceToHost, d.st));
  HTRY(hipStreamSynchronize(d.st));
  if (o->save_file && m > 0) {
    const int ns = std::min(nev, m);
    std::vector<double> vs((size_t)nb * ns);
    HTRY(hipMemcpyAsync(vs.data(), XV.p, sizeof(double) * vs.size(), hipMemcpyDeviceToHost, d.st));
    HTRY(hipStreamSynchronize(d.st));
    RTRY(evr_psi_save(o->basis, nb0, ns, 0, vs.data(), 0, 1, o->save_file));
  }
  if (info) {
    info->iterations = std::min(it, max_it);
    info->converged = done ? 1 : 0;
    info->nconv = nconv;
    info->hpsi_fp64 = d.hpsi64;
    info->hpsi_fp32 = d.hpsi32;
    info->max_resi = maxres;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return EVR_OK;
}
