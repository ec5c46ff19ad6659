// Time propagation drivers over the H|psi> engine: Chebyshev (march_cheby)
// and short iterative Lanczos (march_SIL) for a batch of complex wavepackets,
// with an optional field coupling H_n = H - E_n mu held constant over a step.
//
// Reference: Source_ElVibRot/sub_propagation/sub_module_propa_march.f90
// (MARCH): march_cheby MARCH:4142-4433 (coefficients cof MARCH:5734-5755,
// sub_scaledOpPsi OPPSI:2823-2866, Hmax retry MARCH:4348-4373), march_SIL
// MARCH:2899-3100 with UPsi_spec MARCH:3101-3172.  A complex psi is carried
// as two real columns (Re, Im), as the reference does for its real operator
// kernels (OPPSI:392-408); the operator is real symmetric, so H acts on both
// columns independently and one evr_oppsi call covers a whole batch.
//
// Device layout: every wavepacket batch is nb x (2 nwp) column-major; the
// Krylov space of the SIL driver is one allocation of (npoly + 1) such blocks.
// The per-order vector work is one grid-stride kernel (recurrence, field term,
// series accumulation); reductions run one workgroup per (vector, wavepacket)
// in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "evr_propa.hpp"

namespace evr {
namespace propa {

#define HTRY(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(EVR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)
#define RTRY(expr)                                                                           \
  do {                                                                                       \
    int r_ = (expr);                                                                         \
    if (r_ != EVR_OK) return r_;                                                             \
  } while (0)

static unsigned nblocks(int64_t n)
{
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

// ---- batched complex kernels on (Re, Im) column pairs ------------------------

// One Chebyshev order for every wavepacket w (field E_w, shift e0, scale esc):
//   T = (HW - E_w MW - e0 W2) / esc
//   first order (W1 == nullptr):  W3 = -i T
//   later orders:                 W3 = W1 - 2i T
//   PSI += c W3
__global__ void k_cheb_order(int64_t nb, int nwp, const double *E, double e0, double esc,
                             const double *W1, const double *W2, const double *HW,
                             const double *MW, double *W3, double *PSI, double c)
{
  const double inv = 1.0 / esc;
  const double f = W1 ? 2.0 : 1.0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nwp;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(g / nb);
    const int64_t r = g + nb * w, m = r + nb;  // row i of column 2w, then of column 2w+1
    double tr = HW[r] - e0 * W2[r], ti = HW[m] - e0 * W2[m];
    if (MW) {
      tr -= E[w] * MW[r];
      ti -= E[w] * MW[m];
    }
    tr *= inv;
    ti *= inv;
    double xr = f * ti, xi = -f * tr;
    if (W1) {
      xr += W1[r];
      xi += W1[m];
    }
    W3[r] = xr;
    W3[m] = xi;
    PSI[r] += c * xr;
    PSI[m] += c * xi;
  }
}

// Y_w = beta_w Y_w + sum_j coef[j][w] Q_j,w  (complex coefficients; Q_j = Q + j*qs)
__global__ void k_zcomb(int64_t nb, int nwp, int nk, const double *Q, int64_t qs,
                        const double *coef, const double *beta, double *Y)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nwp;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(g / nb);
    const int64_t r = g + nb * w, m = r + nb;
    double yr = 0.0, yi = 0.0;
    if (beta) {
      const double br = beta[2 * w], bi = beta[2 * w + 1];
      const double ur = Y[r], ui = Y[m];
      yr = br * ur - bi * ui;
      yi = br * ui + bi * ur;
    }
    for (int j = 0; j < nk; ++j) {
      const double cr = coef[2 * (j * nwp + w)], ci = coef[2 * (j * nwp + w) + 1];
      const double qr = Q[j * qs + r], qi = Q[j * qs + m];
      yr += cr * qr - ci * qi;
      yi += cr * qi + ci * qr;
    }
    Y[r] = yr;
    Y[m] = yi;
  }
}

// HW_w -= E_w MW_w
__global__ void k_field(int64_t nb, int nwp, const double *E, const double *MW, double *HW)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < 2 * nb * nwp;
       g += (int64_t)gridDim.x * blockDim.x)
    HW[g] -= E[g / (2 * nb)] * MW[g];
}

// Kernels of the captured Chebyshev step: coefficients are kernel arguments
// (no uploads inside the graph) and the step's field values are picked on
// the device from all steps' values by a step counter the graph advances.
// Y = a X (real a) over n values
__global__ void k_rscale(int64_t n, double a, const double *X, double *Y)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    Y[g] = a * X[g];
}

// Y_w = (c + i s) Y_w for every wavepacket
__global__ void k_zphase(int64_t nb, int nwp, double c, double s, double *Y)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nb * nwp;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(g / nb);
    const int64_t r = g + nb * w, m = r + nb;
    const double yr = Y[r], yi = Y[m];
    Y[r] = c * yr - s * yi;
    Y[m] = s * yr + c * yi;
  }
}

// E[w] = field[step * nwp + w], step = ctr[0]
__global__ void k_field_pick(int nwp, const double *field, const double *ctr, double *E)
{
  const int64_t step = (int64_t)ctr[0];
  for (int w = threadIdx.x; w < nwp; w += blockDim.x) E[w] = field[step * nwp + w];
}

__global__ void k_step_next(double *ctr)
{
  if (threadIdx.x == 0) ctr[0] += 1.0;
}

// ---- several systems, one wavepacket each (Re column, then Im column, of
// nbt = sum_k nb_k rows; row i belongs to system sys[i]) -----------------------

// k_cheb_order with the shift, scale and series coefficient of each row's
// system; a system whose series has ended (act = 0) takes no further term
__global__ void k_mcheb_order(int64_t nbt, const int *sys, const double *e0, const double *esc,
                              const double *c, const int *act, const double *W1, const double *W2,
                              const double *HW, double *W3, double *PSI)
{
  const double f = W1 ? 2.0 : 1.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nbt;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int k = sys[r];
    const int64_t m = r + nbt;
    const double inv = 1.0 / esc[k];
    double tr = HW[r] - e0[k] * W2[r], ti = HW[m] - e0[k] * W2[m];
    tr *= inv;
    ti *= inv;
    double xr = f * ti, xi = -f * tr;
    if (W1) {
      xr += W1[r];
      xi += W1[m];
    }
    W3[r] = xr;
    W3[m] = xi;
    if (act[k]) {
      PSI[r] += c[k] * xr;
      PSI[m] += c[k] * xi;
    }
  }
}

// Y = a_k X per row (both columns)
__global__ void k_mscale(int64_t nbt, const int *sys, const double *a, const double *X, double *Y)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nbt;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double ak = a[sys[r]];
    Y[r] = ak * X[r];
    Y[r + nbt] = ak * X[r + nbt];
  }
}

// Y = (cs_k + i sn_k) Y per row
__global__ void k_mphase(int64_t nbt, const int *sys, const double *cs, const double *sn, double *Y)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nbt;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int k = sys[r];
    const double yr = Y[r], yi = Y[r + nbt];
    Y[r] = cs[k] * yr - sn[k] * yi;
    Y[r + nbt] = sn[k] * yr + cs[k] * yi;
  }
}

// out[k] = <w|w> over the rows of system k (one workgroup per system)
__global__ __launch_bounds__(256) void k_msys_norm(int64_t nbt, const int64_t *off, const double *W,
                                                   double *out)
{
  __shared__ double red[256];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int64_t r = off[k] + threadIdx.x; r < off[k + 1]; r += 256)
    s = fma(W[r], W[r], fma(W[r + nbt], W[r + nbt], s));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// out[j][w] = <X_j,w | Y_w> (complex), X_j = X + j*xs; grid (nwp, nk)
__global__ __launch_bounds__(256) void k_zdots(int64_t nb, int nwp, const double *X, int64_t xs,
                                               const double *Y, double *out)
{
  __shared__ double rr[256], ri[256];
  const int w = blockIdx.x, j = blockIdx.y;
  const double *xr = X + j * xs + nb * (2 * w), *xi = xr + nb;
  const double *yr = Y + nb * (2 * w), *yi = yr + nb;
  double sr = 0.0, si = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += 256) {
    sr = fma(xr[i], yr[i], fma(xi[i], yi[i], sr));
    si = fma(xr[i], yi[i], fma(-xi[i], yr[i], si));
  }
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      rr[threadIdx.x] += rr[threadIdx.x + s];
      ri[threadIdx.x] += ri[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * (j * nwp + w)] = rr[0];
    out[2 * (j * nwp + w) + 1] = ri[0];
  }
}

// real versions for one column: out[j] = <X_j|y>, X_j = X + j*xs; grid (nk)
__global__ __launch_bounds__(256) void k_rdots(int64_t nb, const double *X, int64_t xs,
                                               const double *y, double *out)
{
  __shared__ double rr[256];
  const double *x = X + blockIdx.x * xs;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += 256) s = fma(x[i], y[i], s);
  rr[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) rr[threadIdx.x] += rr[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = rr[0];
}

// y = b y + sum_j c_j X_j
__global__ void k_rcomb(int64_t nb, int nk, const double *X, int64_t xs, const double *c, double b,
                        double *y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = b == 0.0 ? 0.0 : b * y[i];
    for (int j = 0; j < nk; ++j) v = fma(c[j], X[j * xs + i], v);
    y[i] = v;
  }
}

// ---- host numerics ---------------------------------------------------------

std::vector<double> bessel_j(int n, double r)
{
  std::vector<double> out(n + 1, 0.0);
  if (r == 0.0) {
    out[0] = 1.0;
    return out;
  }
  // J_{k-1} = (2k/r) J_k - J_{k+1}, started far above max(n, r) and normalised
  // with J_0 + 2 sum_k J_2k = 1 (Abramowitz & Stegun 9.1.27, 9.1.46).
  const double ar = std::fabs(r);
  const int top = std::max(n, (int)ar);
  const int start = 2 * ((top + 30 + (int)std::sqrt(60.0 * (top + 1))) / 2);
  std::vector<long double> j(start + 1, 0.0L);
  long double jp = 0.0L, jc = 1e-30L;
  for (int k = start; k >= 0; --k) {
    j[k] = jc;
    if (k == 0) break;
    const long double jm = (2.0L * k / (long double)ar) * jc - jp;
    jp = jc;
    jc = jm;
    if (std::fabs((double)jc) > 1e200) {
      for (int q = k; q <= start; ++q) j[q] *= 1e-200L;
      jp *= 1e-200L;
      jc *= 1e-200L;
    }
  }
  long double norm = j[0];
  for (int k = 2; k <= start; k += 2) norm += 2.0L * j[k];
  for (int k = 0; k <= n; ++k) {
    long double v = j[k] / norm;
    if (r < 0.0 && (k & 1)) v = -v;  // J_k(-r) = (-1)^k J_k(r)
    out[k] = (double)v;
  }
  return out;
}

void sym_eig(int n, std::vector<double> A, std::vector<double> &ev, std::vector<double> &V)
{
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int p = 0; p < n; ++p) {
      dg += A[p + (size_t)n * p] * A[p + (size_t)n * p];
      for (int q = p + 1; q < n; ++q) off += A[p + (size_t)n * q] * A[p + (size_t)n * q];
    }
    if (off <= 1e-34 * (dg + 1e-300)) break;
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
  ev.resize(n);
  for (int i = 0; i < n; ++i) ev[i] = A[i + (size_t)n * i];
}

std::vector<cplx> expm_tridiag_e1(int k, const double *alpha, const double *beta, double dt)
{
  std::vector<double> T((size_t)k * k, 0.0), ev, V;
  for (int i = 0; i < k; ++i) {
    T[i + (size_t)k * i] = alpha[i];
    if (i + 1 < k) T[i + (size_t)k * (i + 1)] = T[(i + 1) + (size_t)k * i] = beta[i + 1];
  }
  sym_eig(k, T, ev, V);
  std::vector<cplx> u(k, 0.0);
  for (int e = 0; e < k; ++e) {
    const cplx f = std::exp(cplx(0.0, -ev[e] * dt)) * V[(size_t)k * e];
    for (int i = 0; i < k; ++i) u[i] += f * V[i + (size_t)k * e];
  }
  return u;
}

// ---- operator handle ---------------------------------------------------------

int OpRef::rows(int64_t *nb) const
{
  if (plan) {
    int64_t sz[7];
    RTRY(evr_plan_sizes(plan, sz));
    *nb = sz[0];
    return EVR_OK;
  }
  int64_t n = 0;
  int nb0 = 0;
  RTRY(evr_coupled_sizes(coupled, &n, &nb0, nullptr));
  *nb = n * nb0;
  return EVR_OK;
}

hipStream_t OpRef::stream() const
{
  return (hipStream_t)(plan ? evr_plan_stream(plan) : evr_coupled_stream(coupled));
}

int OpRef::max_nvec() const
{
  if (plan) return evr_plan_max_nvec(plan);
  int mv = 0;
  if (evr_coupled_sizes(coupled, nullptr, nullptr, &mv) != EVR_OK) return 0;
  return mv;
}

// The propagation drivers let each H|psi> take the cheaper of the assembled
// and sum-factorised paths for its number of columns (EVR_AUTO_PATH).
int OpRef::apply(int nvec, const double *X, double *Y, int flags) const
{
  const int f = flags | EVR_DEVICE_PTRS | ((flags & EVR_MATRIX_FREE) ? 0 : EVR_AUTO_PATH);
  return plan ? evr_oppsi(plan, nvec, X, Y, f) : evr_coupled_oppsi(coupled, nvec, X, Y, f);
}

int lanczos_bounds(OpRef p, int steps, double *lo, double *hi)
{
  int64_t nb = 0;
  RTRY(p.rows(&nb));
  hipStream_t st = p.stream();
  DBuf Q, w, red, cf;
  HTRY(Q.alloc(nb * (steps + 1)));
  HTRY(w.alloc(nb));
  HTRY(red.alloc(steps + 1));
  HTRY(cf.alloc(steps + 1));
  std::vector<double> h(nb);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  double nrm = 0.0;
  for (int64_t i = 0; i < nb; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    h[i] = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    nrm += h[i] * h[i];
  }
  for (auto &v : h) v /= std::sqrt(nrm);
  HTRY(upload_now(Q.p, h.data(), sizeof(double) * nb, st));
  std::vector<double> al, be(1, 0.0);
  for (int k = 0; k < steps; ++k) {
    RTRY(p.apply(1, Q.p + nb * k, w.p, EVR_FP64));
    // c_j = <q_j|H q_k>, j <= k; w -= sum_j c_j q_j (Lanczos with full
    // re-orthogonalisation: c_k is alpha_k, c_{k-1} is beta_k; a single pass
    // loses orthogonality within a few tens of steps)
    hipLaunchKernelGGL(k_rdots, dim3(k + 1), dim3(256), 0, st, nb, Q.p, nb, w.p, red.p);
    HTRY(hipGetLastError());
    std::vector<double> c(k + 1);
    HTRY(hipMemcpyAsync(c.data(), red.p, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, st));
    HTRY(hipStreamSynchronize(st));
    al.push_back(c[k]);
    for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
      if (pass == 1) {
        hipLaunchKernelGGL(k_rdots, dim3(k + 1), dim3(256), 0, st, nb, Q.p, nb, w.p, red.p);
        HTRY(hipGetLastError());
        HTRY(hipMemcpyAsync(c.data(), red.p, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, st));
        HTRY(hipStreamSynchronize(st));
      }
      for (auto &v : c) v = -v;
      HTRY(upload_now(cf.p, c.data(), sizeof(double) * (k + 1), st));
      hipLaunchKernelGGL(k_rcomb, dim3(nblocks(nb)), dim3(256), 0, st, nb, k + 1, Q.p, nb, cf.p, 1.0,
                         w.p);
      HTRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rdots, dim3(1), dim3(256), 0, st, nb, w.p, 0, w.p, red.p);
    HTRY(hipGetLastError());
    double b2 = 0.0;
    HTRY(hipMemcpyAsync(&b2, red.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HTRY(hipStreamSynchronize(st));
    const double b = std::sqrt(std::max(b2, 0.0));
    if (k + 1 == steps || b <= 1e-13 * (std::fabs(al.back()) + 1e-300)) break;
    be.push_back(b);
    const double ib = 1.0 / b;
    HTRY(upload_now(cf.p, &ib, sizeof(double), st));
    hipLaunchKernelGGL(k_rcomb, dim3(nblocks(nb)), dim3(256), 0, st, nb, 1, w.p, 0, cf.p, 0.0,
                       Q.p + nb * (k + 1));
    HTRY(hipGetLastError());
  }
  HTRY(hipStreamSynchronize(st));
  const int k = (int)al.size();
  std::vector<double> T((size_t)k * k, 0.0), ev, V;
  for (int i = 0; i < k; ++i) {
    T[i + (size_t)k * i] = al[i];
    if (i + 1 < k) T[i + (size_t)k * (i + 1)] = T[(i + 1) + (size_t)k * i] = be[i + 1];
  }
  sym_eig(k, T, ev, V);
  *lo = *std::min_element(ev.begin(), ev.end());
  *hi = *std::max_element(ev.begin(), ev.end());
  return EVR_OK;
}

// ---- batch context ---------------------------------------------------------

int Batch::init(OpRef H_, OpRef mu_, int nwp_, int flags_)
{
  H = H_;
  mu = mu_;
  nwp = nwp_;
  flags = flags_;
  RTRY(H.rows(&nb));
  st = H.stream();
  if (mu) {
    int64_t n2 = 0;
    RTRY(mu.rows(&n2));
    if (n2 != nb) return set_error(EVR_ERR_ARG, "coupling plan has another basis size");
  }
  HTRY(dred.alloc(2 * (int64_t)nwp * 64));
  HTRY(dcoef.alloc(2 * (int64_t)nwp * 65));
  HTRY(dE.alloc(nwp));
  if (!fork) HTRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  if (!join) HTRY(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  return EVR_OK;
}

Batch::~Batch()
{
  if (fork) (void)hipEventDestroy(fork);
  if (join) (void)hipEventDestroy(join);
}

int Batch::apply(OpRef p, const double *X, double *Y, int fl)
{
  const int ncol = 2 * nwp;
  const int mv = std::max(1, p.max_nvec());
  // the operator runs on its own stream: ordered after this batch's work and
  // before what follows by events (no host wait, so a captured step can take
  // the coupling plan's stream along)
  hipStream_t ps = p.stream();
  if (ps != st) {
    HTRY(hipEventRecord(fork, st));
    HTRY(hipStreamWaitEvent(ps, fork, 0));
  }
  for (int c0 = 0; c0 < ncol; c0 += mv) {
    const int k = std::min(mv, ncol - c0);
    RTRY(p.apply(k, X + nb * c0, Y + nb * c0, fl));
  }
  if (ps != st) {
    HTRY(hipEventRecord(join, ps));
    HTRY(hipStreamWaitEvent(st, join, 0));
  }
  if (p == H) hpsi += ncol;
  return EVR_OK;
}

int Batch::apply_field(const double *X, double *HX, double *MX, const double *dEp)
{
  RTRY(apply(H, X, HX, flags));
  if (mu && dEp) {
    if (!MX) return set_error(EVR_ERR_STATE, "no buffer for the coupling term");
    RTRY(apply(mu, X, MX, flags));
  }
  return EVR_OK;
}

int Batch::dots(const double *X, int64_t xs, int nk, const double *Y, std::vector<cplx> &out)
{
  if (nk < 1 || nk > 64) return set_error(EVR_ERR_ARG, "dot products: 1..64 vectors");
  hipLaunchKernelGGL(k_zdots, dim3(nwp, nk), dim3(256), 0, st, nb, nwp, X, xs, Y, dred.p);
  HTRY(hipGetLastError());
  std::vector<double> h(2 * (size_t)nwp * nk);
  HTRY(hipMemcpyAsync(h.data(), dred.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  out.resize((size_t)nwp * nk);
  for (size_t i = 0; i < out.size(); ++i) out[i] = cplx(h[2 * i], h[2 * i + 1]);
  return EVR_OK;
}

int Batch::comb(const double *Q, int64_t qs, int nk, const std::vector<cplx> &coef,
                const std::vector<cplx> *beta, double *Y)
{
  if (nk < 0 || nk > 64) return set_error(EVR_ERR_ARG, "linear combination: 0..64 vectors");
  std::vector<double> h(2 * (size_t)nwp * (nk + 1));
  for (size_t i = 0; i < (size_t)nwp * nk; ++i) {
    h[2 * i] = coef[i].real();
    h[2 * i + 1] = coef[i].imag();
  }
  const size_t bo = 2 * (size_t)nwp * nk;
  if (beta)
    for (int w = 0; w < nwp; ++w) {
      h[bo + 2 * w] = (*beta)[w].real();
      h[bo + 2 * w + 1] = (*beta)[w].imag();
    }
  HTRY(upload_now(dcoef.p, h.data(), sizeof(double) * h.size(), st));
  hipLaunchKernelGGL(k_zcomb, dim3(nblocks(nb * nwp)), dim3(256), 0, st, nb, nwp, nk, Q, qs,
                     dcoef.p, beta ? dcoef.p + bo : nullptr, Y);
  HTRY(hipGetLastError());
  return EVR_OK;
}

int Batch::scale(const std::vector<cplx> &a, double *Y)
{
  return comb(Y, 0, 0, std::vector<cplx>(), &a, Y);
}

int Batch::set_field(const double *E)
{
  HTRY(upload_now(dE.p, E, sizeof(double) * nwp, st));
  return EVR_OK;
}

int launch_cheb_order(const Batch &P, const double *dE, double e0, double esc, const double *W1,
                      const double *W2, const double *HW, const double *MW, double *W3, double *PSI,
                      double c)
{
  hipLaunchKernelGGL(k_cheb_order, dim3(nblocks(P.nb * P.nwp)), dim3(256), 0, P.st, P.nb, P.nwp, dE,
                     e0, esc, W1, W2, HW, MW, W3, PSI, c);
  HTRY(hipGetLastError());
  return EVR_OK;
}

int launch_field(const Batch &P, const double *dE, const double *MW, double *HW)
{
  hipLaunchKernelGGL(k_field, dim3(nblocks(2 * P.nb * P.nwp)), dim3(256), 0, P.st, P.nb, P.nwp, dE,
                     MW, HW);
  HTRY(hipGetLastError());
  return EVR_OK;
}

int ChebBufs::alloc(int64_t nblk)
{
  for (DBuf *b : {&W1, &W2, &W3, &SAVE, &HW}) {
    const hipError_t e = b->alloc(nblk);
    if (e != hipSuccess) return set_error(EVR_ERR_HIP, std::string("Chebyshev buffers: ") + hipGetErrorString(e));
  }
  return EVR_OK;
}

int cheb_step(Batch &P, ChebBufs &cw, double *PSI, const double *dEp, double *MWp, double dt,
              int npoly, double poly_tol, double &hmin, double &hmax, ChebRecord &rec, int &retries,
              double *store, bool fixed)
{
  const int64_t nblk = P.nb * 2 * P.nwp;
  const size_t bytes = sizeof(double) * nblk;
  HTRY(hipMemcpyAsync(cw.SAVE.p, PSI, bytes, hipMemcpyDeviceToDevice, P.st));
  for (int attempt = 0; attempt < 10; ++attempt) {
    double E0, esc;
    if (fixed) {
      E0 = rec.e0;
      esc = rec.esc;
    } else {
      const double dE_ = hmax - hmin;
      E0 = hmin + 0.5 * dE_;  // MARCH:4211-4231
      esc = 0.5 * dE_;
    }
    const int kmax = fixed ? rec.order : npoly - 1;
    const std::vector<double> J = bessel_j(kmax, esc * dt);
    // w_k lives at store + k*nblk, or in three rotating buffers
    double *w0 = store ? store : cw.W1.p, *w1 = store ? store + nblk : cw.W2.p,
           *w2 = store ? store + 2 * nblk : cw.W3.p;
    // order 0: w0 = psi, psi = J_0 psi
    HTRY(hipMemcpyAsync(w0, cw.SAVE.p, bytes, hipMemcpyDeviceToDevice, P.st));
    HTRY(hipMemcpyAsync(PSI, cw.SAVE.p, bytes, hipMemcpyDeviceToDevice, P.st));
    RTRY(P.scale(std::vector<cplx>(P.nwp, cplx(J[0], 0.0)), PSI));
    // order 1: w1 = -i Hs w0, psi += 2 J_1 w1
    RTRY(P.apply_field(w0, cw.HW.p, MWp, dEp));
    RTRY(launch_cheb_order(P, dEp, E0, esc, nullptr, w0, cw.HW.p, MWp, w1, PSI, 2.0 * J[1]));
    bool blow = false, conv = fixed;
    int order = 1;
    for (order = 2; order <= kmax; ++order) {
      if (store) w2 = store + order * nblk;
      RTRY(P.apply_field(w1, cw.HW.p, MWp, dEp));
      RTRY(launch_cheb_order(P, dEp, E0, esc, w0, w1, cw.HW.p, MWp, w2, PSI, 2.0 * J[order]));
      double *t = w0;
      w0 = w1;
      w1 = w2;
      w2 = t;
      if (fixed) continue;
      // exit test |c_k| <w_k|w_k> (MARCH:4306-4335), over every wavepacket
      std::vector<cplx> nn;
      RTRY(P.dots(w1, 0, 1, w1, nn));
      double crit = 0.0;
      for (int w = 0; w < P.nwp; ++w) crit = std::max(crit, std::fabs(nn[w].real() * 2.0 * J[order]));
      if (!(crit <= 1e15)) {
        blow = true;
        break;
      }
      if (crit < poly_tol) {
        conv = true;
        break;
      }
    }
    if (fixed) order = kmax;
    if (blow || !conv) {  // MARCH:4348-4373
      hmax += 0.1 * (hmax - hmin);
      ++retries;
      continue;
    }
    // phase exp(-i E0 dt) (MARCH:4386-4388)
    RTRY(P.scale(std::vector<cplx>(P.nwp, std::exp(cplx(0.0, -E0 * dt))), PSI));
    rec.e0 = E0;
    rec.esc = esc;
    rec.order = order;
    return EVR_OK;
  }
  return set_error(EVR_ERR_STATE, "Chebyshev expansion did not converge (Hmin " + std::to_string(hmin) +
                                      ", Hmax " + std::to_string(hmax) + "): reduce dt or raise npoly");
}

// One Chebyshev step of fixed order captured as a HIP graph: the launches of
// every order (H|psi> kernels on the plan streams, recurrence, series), the
// exit-test norm <w_K|w_K> into P.dred and the phase, with nothing read back
// inside.  Replaying it advances PSI by one step; with a field the graph
// first picks the step's values from all steps' values by a device counter
// and advances the counter last.

static int capture_cheb_step(Batch &P, ChebBufs &cw, double *PSI, double *MWp, double dt,
                             const double *field_all, double *ctr, const std::vector<double> &J,
                             const ChebRecord &rec)
{
  const int64_t nblk = P.nb * 2 * P.nwp;
  const size_t bytes = sizeof(double) * nblk;
  const double *dEp = field_all ? P.dE.p : nullptr;
  if (field_all) {
    hipLaunchKernelGGL(k_field_pick, dim3(1), dim3(256), 0, P.st, P.nwp, field_all, ctr, P.dE.p);
    HTRY(hipGetLastError());
  }
  double *w0 = cw.W1.p, *w1 = cw.W2.p, *w2 = cw.W3.p;
  HTRY(hipMemcpyAsync(cw.SAVE.p, PSI, bytes, hipMemcpyDeviceToDevice, P.st));
  HTRY(hipMemcpyAsync(w0, PSI, bytes, hipMemcpyDeviceToDevice, P.st));
  hipLaunchKernelGGL(k_rscale, dim3(nblocks(nblk)), dim3(256), 0, P.st, nblk, J[0], cw.SAVE.p, PSI);
  HTRY(hipGetLastError());
  RTRY(P.apply_field(w0, cw.HW.p, MWp, dEp));
  RTRY(launch_cheb_order(P, dEp, rec.e0, rec.esc, nullptr, w0, cw.HW.p, MWp, w1, PSI, 2.0 * J[1]));
  for (int order = 2; order <= rec.order; ++order) {
    RTRY(P.apply_field(w1, cw.HW.p, MWp, dEp));
    RTRY(launch_cheb_order(P, dEp, rec.e0, rec.esc, w0, w1, cw.HW.p, MWp, w2, PSI, 2.0 * J[order]));
    double *t = w0;
    w0 = w1;
    w1 = w2;
    w2 = t;
  }
  hipLaunchKernelGGL(k_zdots, dim3(P.nwp, 1), dim3(256), 0, P.st, P.nb, P.nwp, w1, (int64_t)0, w1,
                     P.dred.p);
  HTRY(hipGetLastError());
  hipLaunchKernelGGL(k_zphase, dim3(nblocks(P.nb * P.nwp)), dim3(256), 0, P.st, P.nb, P.nwp,
                     std::cos(rec.e0 * dt), -std::sin(rec.e0 * dt), PSI);
  HTRY(hipGetLastError());
  if (field_all) {
    hipLaunchKernelGGL(k_step_next, dim3(1), dim3(64), 0, P.st, ctr);
    HTRY(hipGetLastError());
  }
  return EVR_OK;
}

int build_cheb_graph(Batch &P, ChebBufs &cw, double *PSI, double *MWp, double dt,
                            const double *field_all, double *ctr, const ChebRecord &rec, ChebGraph &G)
{
  G.reset();
  const std::vector<double> J = bessel_j(rec.order, rec.esc * dt);
  const int64_t h0 = P.hpsi;
  HTRY(hipStreamBeginCapture(P.st, hipStreamCaptureModeThreadLocal));
  const int rc = capture_cheb_step(P, cw, PSI, MWp, dt, field_all, ctr, J, rec);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(P.st, &g);
  G.hpsi_per_step = P.hpsi - h0;
  P.hpsi = h0;  // nothing ran yet
  if (rc != EVR_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != EVR_OK) return rc;
    HTRY(e);
  }
  G.g = g;
  HTRY(hipGraphInstantiate(&G.x, g, nullptr, nullptr, 0));
  G.rec = rec;
  G.cK = 2.0 * J[rec.order];
  return EVR_OK;
}

// Replay one step; ok = the exit test |c_K| <w_K|w_K> < poly_tol held for
// every wavepacket (else PSI is restored from the saved step start).
int replay_cheb_graph(Batch &P, ChebBufs &cw, double *PSI, double poly_tol, ChebGraph &G,
                             bool &ok)
{
  HTRY(hipGraphLaunch(G.x, P.st));
  P.hpsi += G.hpsi_per_step;
  std::vector<double> h(2 * (size_t)P.nwp);
  HTRY(hipMemcpyAsync(h.data(), P.dred.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, P.st));
  HTRY(hipStreamSynchronize(P.st));
  double crit = 0.0;
  for (int w = 0; w < P.nwp; ++w) crit = std::max(crit, std::fabs(h[2 * w] * G.cK));
  ok = crit < poly_tol;  // NaN fails too
  if (!ok)
    HTRY(hipMemcpyAsync(PSI, cw.SAVE.p, sizeof(double) * P.nb * 2 * P.nwp, hipMemcpyDeviceToDevice,
                        P.st));
  return EVR_OK;
}

}  // namespace propa
}  // namespace evr

using namespace evr::propa;
using evr::set_error;

static int propagate_impl(OpRef H, OpRef mu, const evr_propa_opts *o, int nwp, double *psi, double *autocorr,
                          evr_propa_info *info);

extern "C" int evr_propagate(evr_plan *H, const evr_propa_opts *o, int nwp, double *psi,
                             double *autocorr, evr_propa_info *info)
{
  if (!H) return set_error(EVR_ERR_ARG, "NULL plan");
  return propagate_impl(OpRef(H), OpRef(o ? o->mu : nullptr), o, nwp, psi, autocorr, info);
}

extern "C" int evr_coupled_propagate(evr_coupled *C, const evr_propa_opts *o, int nwp, double *psi,
                                     double *autocorr, evr_propa_info *info)
{
  if (!C) return set_error(EVR_ERR_ARG, "NULL coupled operator");
  if (o && (o->field || o->mu))
    return set_error(EVR_ERR_ARG, "field coupling on coupled states: evr_coupled_propagate_field");
  return propagate_impl(OpRef(C), OpRef(), o, nwp, psi, autocorr, info);
}

extern "C" int evr_coupled_propagate_field(evr_coupled *C, evr_coupled *mu, const evr_propa_opts *o, int nwp,
                                           double *psi, double *autocorr, evr_propa_info *info)
{
  if (!C || !mu) return set_error(EVR_ERR_ARG, "NULL coupled operator");
  if (o && o->mu) return set_error(EVR_ERR_ARG, "the coupling is the coupled operator mu (opts->mu NULL)");
  return propagate_impl(OpRef(C), OpRef(mu), o, nwp, psi, autocorr, info);
}

// Time-dependent field integrators of the reference: march_noD_field (Taylor
// series of the time-dependent equation with the field derivatives) and
// march_RK4_field (classical Runge-Kutta on fcn_field), on the batch.
static int propagate_field_impl(OpRef H, OpRef mu, const evr_propa_opts *o, int nwp, double *psi,
                                double *autocorr, evr_propa_info *info)
{
  const bool taylor = o->method == EVR_PROPA_TAYLOR;
  if (o->nsteps < 0 || !(o->dt > 0.0)) return set_error(EVR_ERR_ARG, "bad dt or nsteps");
  if (o->field) return set_error(EVR_ERR_ARG, "TAYLOR and RK4 take the field in field_sub (field must be NULL)");
  if ((o->field_sub != nullptr) != (bool)mu)
    return set_error(EVR_ERR_ARG, "field_sub and the coupling plan go together");
  if (o->field_sub && (taylor ? o->nsub < 1 : o->nsub != 3))
    return set_error(EVR_ERR_ARG, taylor ? "Taylor needs nsub >= 1 field derivatives"
                                         : "RK4 needs nsub = 3 field values per step");
  if (taylor && (o->npoly < 2 || o->npoly > 64)) return set_error(EVR_ERR_ARG, "Taylor needs 2 <= npoly <= 64");
  const auto t_start = std::chrono::steady_clock::now();
  Batch P;
  RTRY(P.init(H, mu, nwp, o->precision == EVR_FP32 ? EVR_FP32 : EVR_FP64));
  const int64_t nb = P.nb, nblk = nb * 2 * nwp;
  const double dt = o->dt;
  const hipMemcpyKind h2d = o->device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  // E0: the centre of the spectral interval (the reference's para_H%E0 shift)
  double hmin = o->Hmin, hmax = o->Hmax;
  if (!(hmin < hmax)) RTRY(lanczos_bounds(H, 60, &hmin, &hmax));
  const double e0 = 0.5 * (hmin + hmax);
  const int nw = taylor ? o->npoly : 5;   // work vectors
  DBuf PSI, PSI0, W, HM;
  HTRY(PSI.alloc(nblk));
  HTRY(W.alloc(nblk * nw));
  HTRY(HM.alloc(2 * nblk));               // H x and mu x, contiguous
  double *HW = HM.p, *MW = HM.p + nblk;
  HTRY(hipMemcpyAsync(PSI.p, psi, sizeof(double) * nblk, h2d, P.st));
  if (autocorr) {
    HTRY(PSI0.alloc(nblk));
    HTRY(hipMemcpyAsync(PSI0.p, o->psi0 ? o->psi0 : psi, sizeof(double) * nblk, h2d, P.st));
  }
  HTRY(hipStreamSynchronize(P.st));
  const std::vector<cplx> one(nwp, 1.0), zero(nwp, 0.0), me0(nwp, -e0);
  const std::vector<cplx> phase(nwp, std::exp(cplx(0.0, -e0 * dt)));
  std::vector<double> E(nwp, 0.0);
  int max_order = 0;
  // y = -i [(H - E0) x - E mu x]  (fcn_field), the field values E already set
  auto fcn = [&](const double *x, double *y, bool field) -> int {
    RTRY(P.apply(H, x, HW, P.flags));
    if (field) {
      RTRY(P.apply(mu, x, MW, P.flags));
      RTRY(launch_field(P, P.dE.p, MW, HW));
    }
    RTRY(P.comb(x, 0, 1, me0, &one, HW));
    return P.comb(HW, 0, 1, std::vector<cplx>(nwp, cplx(0.0, -1.0)), nullptr, y);
  };
  for (int step = 0; step < o->nsteps; ++step) {
    const double *fs = o->field_sub ? o->field_sub + (int64_t)step * o->nsub * nwp : nullptr;
    if (taylor) {
      // ---- march_noD_field (MARCH:1635-1836): w_0 = psi, psi accumulates w_i
      int max_der = 0;
      std::vector<double> fk;   // fk[k*nwp + w] = -E^(k)(t_n)
      if (fs) {
        const int nb_der = std::min(o->nsub - 1, o->npoly - 1);
        fk.assign((size_t)(nb_der + 1) * nwp, 0.0);
        double fac = dt;
        for (int k = 0; k <= nb_der; ++k) {
          if (k > 0) fac *= dt / k;
          double mx = 0.0;
          for (int w = 0; w < nwp; ++w) {
            fk[(size_t)k * nwp + w] = -fs[(size_t)k * nwp + w];
            mx = std::max(mx, std::fabs(fs[(size_t)k * nwp + w]));
          }
          if (k > 0 && mx * fac > o->poly_tol) max_der = k;
        }
      }
      // the field term is skipped when every derivative is below poly_tol
      // (MARCH:1740-1742)
      bool field_on = false;
      for (double v : fk) field_on |= std::fabs(v) >= o->poly_tol;
      HTRY(hipMemcpyAsync(W.p, PSI.p, sizeof(double) * nblk, hipMemcpyDeviceToDevice, P.st));
      int order = 0;
      for (int i = 0; i + 1 < o->npoly; ++i) {
        const double *wi = W.p + nblk * i;
        double *wn = W.p + nblk * (i + 1);
        RTRY(P.apply(H, wi, HW, P.flags));
        RTRY(P.comb(wi, 0, 1, me0, &one, HW));                 // (H - E0) w_i
        if (field_on) {
          // s = sum_{k <= min(i, max_der)} (dt^k/k!) (-E^(k)) w_{i-k}; HW += mu s
          const int K = std::min(i, max_der);
          std::vector<cplx> c((size_t)(K + 1) * nwp);
          double f = 1.0;
          for (int k = 0; k <= K; ++k) {
            if (k > 0) f *= dt / k;
            for (int w = 0; w < nwp; ++w) c[(size_t)(K - k) * nwp + w] = f * fk[(size_t)k * nwp + w];
          }
          RTRY(P.comb(W.p + nblk * (i - K), nblk, K + 1, c, nullptr, wn));
          RTRY(P.apply(mu, wn, MW, P.flags));
          RTRY(P.comb(MW, 0, 1, one, &one, HW));
        }
        // w_{i+1} = (-i dt/(i+1)) HW;  psi += w_{i+1}
        RTRY(P.comb(HW, 0, 1, std::vector<cplx>(nwp, cplx(0.0, -dt / (i + 1))), nullptr, wn));
        RTRY(P.comb(wn, 0, 1, one, &one, PSI.p));
        order = i + 1;
        std::vector<cplx> nn;
        RTRY(P.dots(wn, 0, 1, wn, nn));
        bool done = true;
        for (int w = 0; w < nwp; ++w) {
          if (!(nn[w].real() < 1e15)) return set_error(EVR_ERR_STATE, "Taylor series diverges (dt too large)");
          if (nn[w].real() >= o->poly_tol) done = false;
        }
        if (done) break;
      }
      max_order = std::max(max_order, order);
      RTRY(P.scale(phase, PSI.p));
    } else {
      // ---- march_RK4_field (MARCH:392-543)
      double *w1 = W.p, *w2 = W.p + nblk, *w3 = W.p + 2 * nblk, *w4 = W.p + 3 * nblk, *w5 = W.p + 4 * nblk;
      auto set_E = [&](int j) -> int {
        if (!fs) return EVR_OK;
        for (int w = 0; w < nwp; ++w) E[w] = fs[(size_t)j * nwp + w];
        return P.set_field(E.data());
      };
      const std::vector<cplx> h2(nwp, 0.5 * dt), h1(nwp, dt), two(nwp, 2.0), h6(nwp, dt / 6.0);
      RTRY(set_E(0));
      RTRY(fcn(PSI.p, w1, fs != nullptr));                                   // w1 = f(t, psi)
      RTRY(P.comb(w1, 0, 1, h2, nullptr, w2));
      RTRY(P.comb(PSI.p, 0, 1, one, &one, w2));                               // w2 = psi + w1 dt/2
      RTRY(set_E(1));
      RTRY(fcn(w2, w3, fs != nullptr));                                       // w3 = f(t + dt/2, w2)
      RTRY(P.comb(w3, 0, 1, h2, nullptr, w2));
      RTRY(P.comb(PSI.p, 0, 1, one, &one, w2));                               // w2 = psi + w3 dt/2
      RTRY(fcn(w2, w4, fs != nullptr));                                       // w4 = f(t + dt/2, w2)
      RTRY(P.comb(w4, 0, 1, h1, nullptr, w2));
      RTRY(P.comb(PSI.p, 0, 1, one, &one, w2));                               // w2 = psi + w4 dt
      RTRY(P.comb(w3, 0, 1, one, &one, w4));                                  // w4 = w3 + w4
      RTRY(set_E(2));
      RTRY(fcn(w2, w3, fs != nullptr));                                       // w3 = f(t + dt, w2)
      RTRY(P.comb(w1, 0, 1, one, nullptr, w5));
      RTRY(P.comb(w3, 0, 1, one, &one, w5));
      RTRY(P.comb(w4, 0, 1, two, &one, w5));                                  // w5 = w1 + w3 + 2 w4
      RTRY(P.comb(w5, 0, 1, h6, &one, PSI.p));                                // psi += w5 dt/6
      RTRY(P.scale(phase, PSI.p));
      max_order = 4;
    }
    if (autocorr) {
      std::vector<cplx> ac;
      RTRY(P.dots(PSI0.p, 0, 1, PSI.p, ac));
      for (int w = 0; w < nwp; ++w) {
        autocorr[2 * ((int64_t)step * nwp + w)] = ac[w].real();
        autocorr[2 * ((int64_t)step * nwp + w) + 1] = ac[w].imag();
      }
    }
  }
  HTRY(hipMemcpyAsync(psi, PSI.p, sizeof(double) * nblk,
                      o->device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, P.st));
  HTRY(hipStreamSynchronize(P.st));
  if (info) {
    info->hpsi = P.hpsi;
    info->max_order = max_order;
    info->retries = 0;
    info->Hmin = hmin;
    info->Hmax = hmax;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return EVR_OK;
}

static int propagate_impl(OpRef H, OpRef mu, const evr_propa_opts *o, int nwp, double *psi, double *autocorr,
                          evr_propa_info *info)
{
  if (!H || !o || !psi || nwp < 1) return set_error(EVR_ERR_ARG, "bad argument");
  if (o->method == EVR_PROPA_TAYLOR || o->method == EVR_PROPA_RK4)
    return propagate_field_impl(H, mu, o, nwp, psi, autocorr, info);
  if (o->method != EVR_PROPA_CHEBY && o->method != EVR_PROPA_SIL)
    return set_error(EVR_ERR_ARG, "method must be EVR_PROPA_CHEBY, _SIL, _TAYLOR or _RK4");
  if (o->method == EVR_PROPA_CHEBY && o->npoly < 3)
    return set_error(EVR_ERR_ARG, "Chebyshev needs npoly >= 3");
  if (o->method == EVR_PROPA_SIL && (o->npoly < 2 || o->npoly > 63))
    return set_error(EVR_ERR_ARG, "SIL needs 2 <= npoly <= 63");
  if (o->nsteps < 0 || !(o->dt > 0.0)) return set_error(EVR_ERR_ARG, "bad dt or nsteps");
  if ((o->field != nullptr) != (bool)mu)
    return set_error(EVR_ERR_ARG, "field values and the coupling plan go together");
  const auto t_start = std::chrono::steady_clock::now();
  Batch P;
  RTRY(P.init(H, mu, nwp, o->precision == EVR_FP32 ? EVR_FP32 : EVR_FP64));
  const int64_t nb = P.nb, nblk = nb * 2 * nwp;
  const hipMemcpyKind h2d = o->device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  DBuf PSI, HW, MW, PSI0;
  HTRY(PSI.alloc(nblk));
  HTRY(HW.alloc(nblk));
  if (mu) HTRY(MW.alloc(nblk));
  HTRY(hipMemcpyAsync(PSI.p, psi, sizeof(double) * nblk, h2d, P.st));
  if (autocorr) {
    HTRY(PSI0.alloc(nblk));
    HTRY(hipMemcpyAsync(PSI0.p, o->psi0 ? o->psi0 : psi, sizeof(double) * nblk, h2d, P.st));
  }
  HTRY(hipStreamSynchronize(P.st));
  double fmax = 0.0;
  if (o->field)
    for (int64_t k = 0; k < (int64_t)o->nsteps * nwp; ++k) fmax = std::max(fmax, std::fabs(o->field[k]));
  double hmin = o->Hmin, hmax = o->Hmax;
  int retries = 0, max_order = 0;
  ChebBufs cw;
  DBuf Q;
  if (o->method == EVR_PROPA_CHEBY) {
    if (!(hmin < hmax)) {
      RTRY(lanczos_bounds(H, 60, &hmin, &hmax));
      const double pad = 0.02 * (hmax - hmin);
      hmin -= pad;
      hmax += pad;
    }
    if (mu && fmax > 0.0) {
      double mlo, mhi;
      RTRY(lanczos_bounds(mu, 60, &mlo, &mhi));
      const double mb = 1.05 * std::max(std::fabs(mlo), std::fabs(mhi));
      hmin -= fmax * mb;
      hmax += fmax * mb;
    }
    RTRY(cw.alloc(nblk));
  } else {
    HTRY(Q.alloc(nblk * (o->npoly + 1)));
  }
  std::vector<double> E(nwp, 0.0);
  // captured steps: all steps' field values and the step counter on the device
  const bool use_graph = o->graph && o->method == EVR_PROPA_CHEBY && o->nsteps > 1;
  ChebGraph G;
  DBuf field_all, ctr;
  if (use_graph && o->field) {
    HTRY(field_all.alloc((int64_t)o->nsteps * nwp));
    HTRY(ctr.alloc(1));
    HTRY(evr::upload_now(field_all.p, o->field, sizeof(double) * o->nsteps * nwp, P.st));
  }
  for (int step = 0; step < o->nsteps; ++step) {
    const double *dEp = o->field ? P.dE.p : nullptr;
    double *MWp = o->field ? MW.p : nullptr;
    if (o->method == EVR_PROPA_CHEBY && G.x) {
      // ---- a captured step of the last order; redone adaptively (and the
      // graph rebuilt) when its exit test fails
      bool ok = false;
      RTRY(replay_cheb_graph(P, cw, PSI.p, o->poly_tol, G, ok));
      if (ok) {
        max_order = std::max(max_order, G.rec.order);
      } else {
        if (o->field) {
          for (int w = 0; w < nwp; ++w) E[w] = o->field[(int64_t)step * nwp + w];
          RTRY(P.set_field(E.data()));
        }
        ChebRecord rec;
        RTRY(cheb_step(P, cw, PSI.p, dEp, MWp, o->dt, o->npoly, o->poly_tol, hmin, hmax, rec, retries,
                       nullptr, false));
        max_order = std::max(max_order, rec.order);
        if (rec.order > G.rec.order || rec.e0 != G.rec.e0 || rec.esc != G.rec.esc) {
          RTRY(build_cheb_graph(P, cw, PSI.p, MWp, o->dt, field_all.p, ctr.p, rec, G));
        }
      }
    } else if (o->method == EVR_PROPA_CHEBY) {
      // ---- march_cheby (MARCH:4142-4433)
      if (o->field) {
        for (int w = 0; w < nwp; ++w) E[w] = o->field[(int64_t)step * nwp + w];
        RTRY(P.set_field(E.data()));
      }
      ChebRecord rec;
      RTRY(cheb_step(P, cw, PSI.p, dEp, MWp, o->dt, o->npoly, o->poly_tol, hmin, hmax, rec, retries,
                     nullptr, false));
      max_order = std::max(max_order, rec.order);
      if (use_graph && step + 1 < o->nsteps) {
        if (o->field) {
          const double next = (double)(step + 1);
          HTRY(evr::upload_now(ctr.p, &next, sizeof(double), P.st));
        }
        RTRY(build_cheb_graph(P, cw, PSI.p, MWp, o->dt, field_all.p, ctr.p, rec, G));
      }
    } else {
      if (o->field) {
        for (int w = 0; w < nwp; ++w) E[w] = o->field[(int64_t)step * nwp + w];
        RTRY(P.set_field(E.data()));
      }
      // ---- march_SIL (MARCH:2899-3100)
      const int kmax = o->npoly;
      std::vector<cplx> nn;
      RTRY(P.dots(PSI.p, 0, 1, PSI.p, nn));
      std::vector<double> nu(nwp);
      std::vector<cplx> a(nwp), zero(nwp, 0.0), one(nwp, 1.0);
      for (int w = 0; w < nwp; ++w) {
        nu[w] = std::sqrt(std::max(nn[w].real(), 0.0));
        a[w] = nu[w] > 0.0 ? 1.0 / nu[w] : 0.0;
      }
      RTRY(P.comb(PSI.p, 0, 1, a, &zero, Q.p));
      std::vector<std::vector<double>> al(nwp), be(nwp, std::vector<double>(1, 0.0));
      std::vector<double> e0(nwp, 0.0);
      std::vector<std::vector<cplx>> U(nwp);
      int n = 0;
      for (int k = 0; k < kmax; ++k) {
        double *qk = Q.p + nblk * k;
        RTRY(P.apply_field(qk, HW.p, MWp, dEp));
        if (MWp) RTRY(launch_field(P, dEp, MWp, HW.p));
        // w = H q_k - beta_k q_{k-1} (the E0 q_k shift is folded into the projection)
        if (k > 0) {
          std::vector<cplx> c(nwp);
          for (int w = 0; w < nwp; ++w) c[w] = -be[w][k];
          RTRY(P.comb(Q.p + nblk * (k - 1), 0, 1, c, &one, HW.p));
        }
        // c_j = <q_j|w>, j <= k: c_k is alpha_k + E0, the others re-orthogonalise
        std::vector<cplx> c;
        RTRY(P.dots(Q.p, nblk, k + 1, HW.p, c));
        if (k == 0)
          for (int w = 0; w < nwp; ++w) e0[w] = c[w].real();  // E0 = <psi|H|psi>
        for (int w = 0; w < nwp; ++w) al[w].push_back(c[(size_t)k * nwp + w].real() - e0[w]);
        for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
          if (pass == 1) RTRY(P.dots(Q.p, nblk, k + 1, HW.p, c));
          for (auto &v : c) v = -v;
          RTRY(P.comb(Q.p, nblk, k + 1, c, &one, HW.p));
        }
        bool done = true;
        for (int w = 0; w < nwp; ++w) {
          U[w] = expm_tridiag_e1(k + 1, al[w].data(), be[w].data(), o->dt);
          if (std::abs(U[w][k]) >= o->poly_tol) done = false;
        }
        n = k + 1;
        if (done || k + 1 == kmax) break;
        std::vector<cplx> bb;
        RTRY(P.dots(HW.p, 0, 1, HW.p, bb));
        std::vector<cplx> ib(nwp);
        for (int w = 0; w < nwp; ++w) {
          const double b = std::sqrt(std::max(bb[w].real(), 0.0));
          be[w].push_back(b);
          ib[w] = b > 0.0 ? 1.0 / b : 0.0;
        }
        RTRY(P.comb(HW.p, 0, 1, ib, &zero, Q.p + nblk * (k + 1)));
      }
      max_order = std::max(max_order, n);
      // psi = nu exp(-i E0 dt) sum_k u_k q_k
      std::vector<cplx> coef((size_t)n * nwp);
      for (int k = 0; k < n; ++k)
        for (int w = 0; w < nwp; ++w)
          coef[(size_t)k * nwp + w] = nu[w] * U[w][k] * std::exp(cplx(0.0, -e0[w] * o->dt));
      RTRY(P.comb(Q.p, nblk, n, coef, nullptr, PSI.p));
    }
    if (autocorr) {
      std::vector<cplx> ac;
      RTRY(P.dots(PSI0.p, 0, 1, PSI.p, ac));
      for (int w = 0; w < nwp; ++w) {
        autocorr[2 * ((int64_t)step * nwp + w)] = ac[w].real();
        autocorr[2 * ((int64_t)step * nwp + w) + 1] = ac[w].imag();
      }
    }
  }
  HTRY(hipMemcpyAsync(psi, PSI.p, sizeof(double) * nblk,
                      o->device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, P.st));
  HTRY(hipStreamSynchronize(P.st));
  if (info) {
    info->hpsi = P.hpsi;
    info->max_order = max_order;
    info->retries = retries;
    info->Hmin = hmin;
    info->Hmax = hmax;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return EVR_OK;
}

extern "C" int evr_plan_bounds(evr_plan *p, int steps, double *Hmin, double *Hmax)
{
  if (!p || !Hmin || !Hmax || steps < 2) return set_error(EVR_ERR_ARG, "bad argument");
  int64_t sz[7];
  RTRY(evr_plan_sizes(p, sz));
  return lanczos_bounds(p, (int)std::min<int64_t>(steps, sz[0]), Hmin, Hmax);
}

extern "C" int evr_bessel_j(int n, double r, double *out)
{
  if (n < 0 || !out) return set_error(EVR_ERR_ARG, "bad argument");
  const std::vector<double> j = bessel_j(n, r);
  std::copy(j.begin(), j.end(), out);
  return EVR_OK;
}

// ---- march_cheby over a multi-system operator ---------------------------------
// Each system k keeps its own E0_k, Esc_k, Bessel coefficients and adaptive
// order (the exit test of march_cheby, MARCH:4306-4335, per system); one
// H|psi> launch sequence per order advances all of them.
extern "C" int evr_multi_propagate(evr_multi *m, double dt, int nsteps, int npoly, double poly_tol,
                                   const double *Hmin, const double *Hmax, double *psi, int device_ptrs,
                                   evr_propa_info *info)
{
  if (!m || !Hmin || !Hmax || !psi || npoly < 3 || nsteps < 0 || !(dt > 0.0))
    return set_error(EVR_ERR_ARG, "bad argument");
  const auto t_start = std::chrono::steady_clock::now();
  const int n = evr_multi_count(m);
  if (n < 1) return n;
  std::vector<int64_t> off(n + 1);
  RTRY(evr_multi_offsets(m, off.data()));
  const int64_t nbt = off[n], nblk = 2 * nbt;
  hipStream_t st = (hipStream_t)evr_multi_stream(m);
  std::vector<double> e0(n), esc(n), cs(n), sn(n);
  std::vector<std::vector<double>> J(n);
  for (int k = 0; k < n; ++k) {
    if (!(Hmin[k] < Hmax[k])) return set_error(EVR_ERR_ARG, "Hmin[k] < Hmax[k] needed for every system");
    const double dE = Hmax[k] - Hmin[k];
    e0[k] = Hmin[k] + 0.5 * dE;  // MARCH:4211-4231
    esc[k] = 0.5 * dE;
    J[k] = bessel_j(npoly - 1, esc[k] * dt);
    const std::complex<double> ph = std::exp(std::complex<double>(0.0, -e0[k] * dt));
    cs[k] = ph.real();
    sn[k] = ph.imag();
  }
  // coefficient table: row 0 = J_0, row j >= 1 = 2 J_j, one value per system
  std::vector<double> ctab((size_t)npoly * n);
  for (int j = 0; j < npoly; ++j)
    for (int k = 0; k < n; ++k) ctab[(size_t)j * n + k] = (j == 0 ? 1.0 : 2.0) * J[k][j];
  std::vector<int> sys(nbt);
  for (int k = 0; k < n; ++k)
    for (int64_t r = off[k]; r < off[k + 1]; ++r) sys[r] = k;
  DBuf PSI, W1, W2, W3, HW, dE0, dEsc, dC, dCs, dSn, dNorm;
  for (DBuf *b : {&PSI, &W1, &W2, &W3, &HW}) HTRY(b->alloc(nblk));
  HTRY(dE0.alloc(n)); HTRY(dEsc.alloc(n)); HTRY(dCs.alloc(n)); HTRY(dSn.alloc(n)); HTRY(dNorm.alloc(n));
  HTRY(dC.alloc((int64_t)npoly * n));
  int *dSys = nullptr, *dAct = nullptr;
  int64_t *dOff = nullptr;
  struct Free {
    void *p[3];
    ~Free() { for (void *q : p) if (q) (void)hipFree(q); }
  } guard{{nullptr, nullptr, nullptr}};
  HTRY(hipMalloc((void **)&dSys, sizeof(int) * std::max<int64_t>(nbt, 1)));
  guard.p[0] = dSys;
  HTRY(hipMalloc((void **)&dAct, sizeof(int) * n));
  guard.p[1] = dAct;
  HTRY(hipMalloc((void **)&dOff, sizeof(int64_t) * (n + 1)));
  guard.p[2] = dOff;
  HTRY(evr::upload_now(dSys, sys.data(), sizeof(int) * nbt, st));
  HTRY(evr::upload_now(dOff, off.data(), sizeof(int64_t) * (n + 1), st));
  HTRY(evr::upload_now(dE0.p, e0.data(), sizeof(double) * n, st));
  HTRY(evr::upload_now(dEsc.p, esc.data(), sizeof(double) * n, st));
  HTRY(evr::upload_now(dCs.p, cs.data(), sizeof(double) * n, st));
  HTRY(evr::upload_now(dSn.p, sn.data(), sizeof(double) * n, st));
  HTRY(evr::upload_now(dC.p, ctab.data(), sizeof(double) * ctab.size(), st));
  HTRY(hipMemcpyAsync(PSI.p, psi, sizeof(double) * nblk,
                      device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  const unsigned nbk = nblocks(nbt);
  int64_t hpsi = 0;
  int max_order = 0;
  std::vector<int> act(n);
  std::vector<double> nrm(n);
  for (int step = 0; step < nsteps; ++step) {
    std::fill(act.begin(), act.end(), 1);
    HTRY(evr::upload_now(dAct, act.data(), sizeof(int) * n, st));
    double *w0 = W1.p, *w1 = W2.p, *w2 = W3.p;
    // order 0: w0 = psi, psi = J_0 psi; order 1: w1 = -i Hs w0, psi += 2 J_1 w1
    HTRY(hipMemcpyAsync(w0, PSI.p, sizeof(double) * nblk, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_mscale, dim3(nbk), dim3(256), 0, st, nbt, dSys, dC.p, w0, PSI.p);
    HTRY(hipGetLastError());
    RTRY(evr_multi_oppsi(m, 2, w0, HW.p, EVR_DEVICE_PTRS));
    hipLaunchKernelGGL(k_mcheb_order, dim3(nbk), dim3(256), 0, st, nbt, dSys, dE0.p, dEsc.p, dC.p + n, dAct,
                       nullptr, w0, HW.p, w1, PSI.p);
    HTRY(hipGetLastError());
    hpsi += 2;
    int left = n, order = 1;
    for (order = 2; order < npoly && left > 0; ++order) {
      RTRY(evr_multi_oppsi(m, 2, w1, HW.p, EVR_DEVICE_PTRS));
      hipLaunchKernelGGL(k_mcheb_order, dim3(nbk), dim3(256), 0, st, nbt, dSys, dE0.p, dEsc.p,
                         dC.p + (size_t)order * n, dAct, w0, w1, HW.p, w2, PSI.p);
      HTRY(hipGetLastError());
      hpsi += 2;
      double *t = w0;
      w0 = w1;
      w1 = w2;
      w2 = t;
      // exit test per system (MARCH:4306-4335)
      hipLaunchKernelGGL(k_msys_norm, dim3(n), dim3(256), 0, st, nbt, dOff, w1, dNorm.p);
      HTRY(hipGetLastError());
      HTRY(hipMemcpyAsync(nrm.data(), dNorm.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
      HTRY(hipStreamSynchronize(st));
      bool changed = false;
      for (int k = 0; k < n; ++k) {
        if (!act[k]) continue;
        const double crit = std::fabs(nrm[k] * ctab[(size_t)order * n + k]);
        if (!(crit <= 1e15))
          return set_error(EVR_ERR_STATE, "Chebyshev series of system " + std::to_string(k) +
                                              " diverges: widen its interval");
        if (crit < poly_tol) {
          act[k] = 0;
          --left;
          changed = true;
          max_order = std::max(max_order, order);
        }
      }
      if (changed && left > 0) HTRY(evr::upload_now(dAct, act.data(), sizeof(int) * n, st));
    }
    if (left > 0)
      return set_error(EVR_ERR_STATE, "Chebyshev expansion did not converge within npoly terms for " +
                                          std::to_string(left) + " system(s): reduce dt or raise npoly");
    // phase exp(-i E0_k dt) (MARCH:4386-4388)
    hipLaunchKernelGGL(k_mphase, dim3(nbk), dim3(256), 0, st, nbt, dSys, dCs.p, dSn.p, PSI.p);
    HTRY(hipGetLastError());
  }
  HTRY(hipMemcpyAsync(psi, PSI.p, sizeof(double) * nblk,
                      device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  if (info) {
    info->hpsi = hpsi;
    info->max_order = max_order;
    info->retries = 0;
    info->Hmin = *std::min_element(Hmin, Hmin + n);
    info->Hmax = *std::max_element(Hmax, Hmax + n);
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return EVR_OK;
}
