// Shared pieces of the propagation drivers and the control gradient: the
// batch context (a plan, an optional coupling plan, nwp complex wavepackets
// stored as (Re, Im) column pairs) and the host numerics they share.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/evr_hip.h"
#include "evr_internal.hpp"

namespace evr {
namespace propa {

using cplx = std::complex<double>;

// The operator a driver applies: one plan, or coupled states over one basis.
struct OpRef {
  evr_plan *plan = nullptr;
  evr_coupled *coupled = nullptr;
  OpRef() = default;
  OpRef(evr_plan *p) : plan(p) {}
  OpRef(evr_coupled *c) : coupled(c) {}
  explicit operator bool() const { return plan || coupled; }
  bool operator==(const OpRef &o) const { return plan == o.plan && coupled == o.coupled; }
  bool operator!=(const OpRef &o) const { return !(*this == o); }
  int rows(int64_t *nb) const;   // length of one column
  hipStream_t stream() const;
  int max_nvec() const;
  // Y = H X for nvec columns (device pointers, on stream())
  int apply(int nvec, const double *X, double *Y, int flags) const;
};

struct DBuf {
  double *p = nullptr;
  hipError_t alloc(int64_t k)
  {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    return k > 0 ? hipMalloc((void **)&p, sizeof(double) * (size_t)k) : hipSuccess;
  }
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf()
  {
    if (p) (void)hipFree(p);
  }
};

// J_0(r) .. J_n(r) (Miller backward recurrence in long double).
std::vector<double> bessel_j(int n, double r);
// Eigenpairs of a small dense symmetric matrix (cyclic Jacobi; V column-major).
void sym_eig(int n, std::vector<double> A, std::vector<double> &ev, std::vector<double> &V);
// exp(-i T dt) e_1 for the k x k tridiagonal T (diagonal alpha[0..k-1],
// off-diagonal beta[1..k-1]).
std::vector<cplx> expm_tridiag_e1(int k, const double *alpha, const double *beta, double dt);
// Extreme eigenvalues of a symmetric plan from `steps` Lanczos steps.
int lanczos_bounds(OpRef p, int steps, double *lo, double *hi);

// nwp complex wavepackets on one plan (and an optional coupling plan).
struct Batch {
  OpRef H, mu;
  hipStream_t st = nullptr;
  int64_t nb = 0;
  int nwp = 0, flags = 0;
  int64_t hpsi = 0;
  DBuf dred, dcoef, dE;
  hipEvent_t fork = nullptr, join = nullptr;  // cross-stream order with a coupling plan

  Batch() = default;
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;
  ~Batch();

  int init(OpRef H, OpRef mu, int nwp, int flags);
  // Y = p X on all 2 nwp columns (chunks of the plan's max_nvec)
  int apply(OpRef p, const double *X, double *Y, int flags);
  // HX = H X and, when a field is set, MX = mu X
  int apply_field(const double *X, double *HX, double *MX, const double *dE);
  // out[j*nwp + w] = <X_j,w | Y_w>, X_j = X + j*xs, j < nk <= 64
  int dots(const double *X, int64_t xs, int nk, const double *Y, std::vector<cplx> &out);
  // Y_w = beta_w Y_w + sum_j coef[j*nwp + w] Q_j,w (beta NULL: Y not read)
  int comb(const double *Q, int64_t qs, int nk, const std::vector<cplx> &coef,
           const std::vector<cplx> *beta, double *Y);
  int scale(const std::vector<cplx> &a, double *Y);
  int set_field(const double *E);
};

// One Chebyshev order on the batch (see k_cheb_order) and HW -= E mu W.
int launch_cheb_order(const Batch &P, const double *dE, double e0, double esc, const double *W1,
                      const double *W2, const double *HW, const double *MW, double *W3, double *PSI,
                      double c);
int launch_field(const Batch &P, const double *dE, const double *MW, double *HW);

// Work buffers of one Chebyshev step (nblk = nb * 2 nwp each).
struct ChebBufs {
  DBuf W1, W2, W3, SAVE, HW;
  int alloc(int64_t nblk);
};
// What a step used: shift, scale and highest order.
struct ChebRecord {
  double e0 = 0.0, esc = 1.0;
  int order = 0;
};
// One march_cheby step on the batch, PSI (device, nb x 2nwp) in place, field
// values dE (device, per wavepacket) with MW a buffer for mu w (both NULL
// without a field).  Adaptive (fixed == false): orders until the exit test,
// Hmax widened and the step redone on failure; rec receives E0, Esc, order.
// Fixed: exactly rec.order orders with rec's E0/Esc (a recomputation).  With
// store != NULL, w_k is kept at store + k*nblk for k = 0..order (the store
// holds npoly blocks in adaptive mode).
int cheb_step(Batch &P, ChebBufs &cw, double *PSI, const double *dE, double *MW, double dt,
              int npoly, double poly_tol, double &hmin, double &hmax, ChebRecord &rec, int &retries,
              double *store, bool fixed);

// One Chebyshev step of fixed order captured as a HIP graph (the launches of
// cheb_step with rec's order, E0 and Esc): with field_all (all steps' field
// values on the device) the graph picks its step's values by the device
// counter ctr and advances it.  replay_cheb_graph runs it and checks the exit
// test |c_K| <w_K|w_K> < poly_tol (ok = false: PSI restored to the step start).
struct ChebGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  ChebRecord rec;
  double cK = 0.0;             // 2 J_K: the exit test's coefficient
  int64_t hpsi_per_step = 0;   // H|psi> columns one replay applies
  ChebGraph() = default;
  ChebGraph(const ChebGraph &) = delete;
  ChebGraph &operator=(const ChebGraph &) = delete;
  void reset()
  {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
  }
  ~ChebGraph() { reset(); }
};
int build_cheb_graph(Batch &P, ChebBufs &cw, double *PSI, double *MWp, double dt, const double *field_all,
                     double *ctr, const ChebRecord &rec, ChebGraph &G);
int replay_cheb_graph(Batch &P, ChebBufs &cw, double *PSI, double poly_tol, ChebGraph &G, bool &ok);

}  // namespace propa
}  // namespace evr
