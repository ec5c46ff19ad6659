// Internal helpers shared by the plan and driver translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/evr_hip.h"

namespace evr {
// Record the message returned by evr_last_error() and return `code`.
int set_error(int code, const std::string &msg);

// Host -> device copy of a transient host array (a local vector or scalar):
// enqueued on `s` after the work already there, and complete on return, so
// the source may be reused or freed at once.
inline hipError_t upload_now(void *dst, const void *src, size_t bytes, hipStream_t s)
{
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}

// Basis -> grid for ncol device columns on the plan stream: X is nb_packed x
// ncol, G is sum_nq x ncol (column-major, term-concatenated grid order).
int plan_b2g_device(evr_plan *p, int ncol, const double *X, double *G);
// Device array (sum_nq) of w_t * (quadrature weight) for every term-grid point:
// the V part of the operator is sum_g gw_g b_g b_g^T with b_g the basis
// functions at g, so d<x|H|y>/dV_g = gw_g (B x)_g (B y)_g.
int plan_grid_weights(evr_plan *p, const double **gw);
// Term-concatenated grid order -> the caller's grid order (DP: reference order).
int plan_grid_to_user(const evr_plan *p, const double *term_grid, double *user);
}  // namespace evr
