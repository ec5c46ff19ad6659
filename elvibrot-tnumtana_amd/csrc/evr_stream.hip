// Streamed terms: a term beyond what one workgroup holds in LDS (more than
// 8192 functions, e.g. a direct-product basis of 3 modes x 21 functions or an
// SG4 term of 6 modes at level 2; a 1D slot of 32 to 64 functions; slabs
// that do not fit), and the one term of a direct-product plan, runs the
// reference's per-term sequence as whole-device passes over global memory,
// one mode at a time:
//   basis -> grid   BDP_TO_GDP_OF_SmolyakRep (BGSG4:2385-2496)
//   grid operator   OpPsi_OF_ONEDP_FOR_SG4 (OPSG4:1448-1650): V, the type-1
//                   F2/F1 derivative sums or the type-10 -1/2 d_i G_ji(q) d_j
//   grid -> basis   GDP_TO_BDP_OF_SmolyakRep (BGSG4:2497-2581)
// with the grid derivative of DerivOp_TO_RDP_OF_SmolaykRep (BGSG4:2690-2795).
// Tensors are [point or function][vector]: the vectors of the chunk fastest,
// then mode 0, mode 1, ... (the term's own basis and grid orders).
#pragma once

#include "evr_kernels.hpp"

namespace evr {

// one mode pass: out[l + L (o + nout r)] = sum_i M(o, i) in[l + L (i + nin r)]
// with M(o, i) = M[o + ld i], or M[i + ld o] when trans; the matrix (at most
// 64 x 64, FP64 in memory) is staged in LDS in T as sM[o nin + i].  A thread
// keeps the nin inputs of its pencil (l, r) in registers and writes a chunk
// of osz consecutive outputs (nchunk chunks per pencil, the chunk index
// slowest), so every input is read nchunk times and consecutive threads take
// consecutive l (coalesced whenever L >= 64).  MAXIN >= nin.  With accum
// the pass adds scale times its result to out instead of storing it.
template <typename T, int MAXIN>
__global__ __launch_bounds__(256) void k_mode_pass(const T *__restrict__ in, T *__restrict__ out,
                                                   const double *__restrict__ M, int ld, int trans,
                                                   int64_t L, int nin, int nout, int64_t R, int osz,
                                                   int accum, double scale)
{
  __shared__ T sM[kMaxNq * kMaxNq];
  for (int e = threadIdx.x; e < nin * nout; e += blockDim.x) {
    const int o = e / nin, i = e - o * nin;
    sM[e] = T(trans ? M[i + (int64_t)ld * o] : M[o + (int64_t)ld * i]);
  }
  __syncthreads();
  const int64_t LR = L * R;
  const int nchunk = (nout + osz - 1) / osz;
  const int64_t total = LR * nchunk;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ch = idx / LR;
    const int64_t lr = idx - ch * LR;
    const int64_t r = lr / L;
    const int64_t l = lr - r * L;
    const T *src = in + l + L * nin * r;
    T *dst = out + l + L * nout * r;
    T x[MAXIN];
#pragma unroll
    for (int i = 0; i < MAXIN; ++i) x[i] = i < nin ? src[L * i] : T(0);
    const int o1 = min(nout, (int)(ch + 1) * osz);
    for (int o = (int)ch * osz; o < o1; ++o) {
      const T *m = sM + o * nin;
      T acc = T(0);
#pragma unroll
      for (int i = 0; i < MAXIN; ++i)
        if (i < nin) acc = fma(m[i], x[i], acc);
      if (accum) dst[L * o] = fma(T(scale), acc, dst[L * o]);
      else dst[L * o] = acc;
    }
  }
}

// the same mode pass on MFMA when L >= 16: a wave takes a tile of 16
// consecutive l of one r and all outputs, out(16 o x 16 l) = M(16 o x 4 i)
// in(4 i x 16 l) accumulated over K-steps of 4 (v_mfma_*_16x16x4): the
// matrix from LDS, zero-padded to 16 rows and 4 columns, the inputs loaded
// once from global into registers (zero past nin or L), each output written
// once.  A lane holds A(row lane % 16, k lane / 16), B(k lane / 16, column
// lane % 16) and D(row lane / 16 + 4 j for FP64, 4 (lane / 16) + j for FP32,
// column lane % 16).
template <typename T>
__global__ __launch_bounds__(256) void k_mode_mfma(const T *__restrict__ in, T *__restrict__ out,
                                                   const double *__restrict__ M, int ld, int trans,
                                                   int64_t L, int nin, int nout, int64_t R, int accum,
                                                   double scale)
{
  using acc_t = T __attribute__((ext_vector_type(4)));
  constexpr int KMAX = kMaxNq / 4;
  __shared__ T sA[kMaxNq * kMaxNq];
  const int nin_p = (nin + 3) & ~3, nout_p = (nout + 15) & ~15;
  for (int e = threadIdx.x; e < nout_p * nin_p; e += blockDim.x) {
    const int o = e / nin_p, i = e - o * nin_p;
    sA[e] = (o < nout && i < nin) ? T(trans ? M[i + (int64_t)ld * o] : M[o + (int64_t)ld * i]) : T(0);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ks = nin_p / 4, ot = nout_p / 16;
  const int64_t ltiles = (L + 15) / 16, ntiles = ltiles * R;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t r = tile / ltiles;
    const int64_t l = (tile - r * ltiles) * 16 + (lane & 15);
    const bool lok = l < L;
    const T *src = in + l + L * nin * r;
    T bf[KMAX];
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      const int i = 4 * s + (lane >> 4);
      bf[s] = (s < ks && lok && i < nin) ? src[L * i] : T(0);
    }
    T *dst = out + l + L * nout * r;
    for (int t = 0; t < ot; ++t) {
      acc_t acc = {T(0), T(0), T(0), T(0)};
      const T *arow = sA + (16 * t + (lane & 15)) * nin_p + (lane >> 4);
#pragma unroll
      for (int s = 0; s < KMAX; ++s)
        if (s < ks) {
          if constexpr (sizeof(T) == 8)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[4 * s], bf[s], acc, 0, 0, 0);
          else
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[4 * s], bf[s], acc, 0, 0, 0);
        }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 16 * t + (sizeof(T) == 8 ? (lane >> 4) + 4 * j : 4 * (lane >> 4) + j);
        if (lok && o < nout) {
          if (accum) dst[L * o] = fma(T(scale), acc[j], dst[L * o]);
          else dst[L * o] = acc[j];
        }
      }
    }
  }
}

// The elementwise kernels below index with 32 bits: the host keeps every
// tensor of a pass below 2^31 values (the work block is at most 4 GiB).

// out[r k + v] = X[map[r] nvec + c0 + v] (0 for a pruned function, map < 0)
template <typename T>
__global__ __launch_bounds__(256) void k_stream_gather(const T *__restrict__ X, int nvec, int c0, int k,
                                                       const int *__restrict__ map, unsigned n,
                                                       T *__restrict__ out)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned r = i / (unsigned)k;
    const int v = (int)(i - r * (unsigned)k);
    const int pk = map[r];
    out[i] = pk >= 0 ? X[(int64_t)pk * nvec + c0 + v] : T(0);
  }
}

// R[rowpos[r] nvec + c0 + v] = w Y[r k + v], or w (Y + Yt) / 2 with the
// transposed pass Yt (rows of pruned functions, rowpos < 0, are skipped)
template <typename T>
__global__ __launch_bounds__(256) void k_stream_scatter(const T *__restrict__ Y, const T *__restrict__ Yt,
                                                        double w, int nvec, int c0, int k,
                                                        const int *__restrict__ rowpos, unsigned n,
                                                        T *__restrict__ R)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned r = i / (unsigned)k;
    const int v = (int)(i - r * (unsigned)k);
    const int rp = rowpos[r];
    if (rp < 0) continue;
    const T y = Yt ? T(0.5) * (Y[i] + Yt[i]) : Y[i];
    R[(int64_t)rp * nvec + c0 + v] = T(w) * y;
  }
}

// grid columns <-> [point][vector]: G[(c0 + v) gld + q] with G at the term's
// first point
template <typename T>
__global__ __launch_bounds__(256) void k_stream_grid_out(const T *__restrict__ Gt, unsigned n, int k, int c0,
                                                         T *__restrict__ G, int64_t gld)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned q = i / (unsigned)k;
    const int v = (int)(i - q * (unsigned)k);
    G[(int64_t)(c0 + v) * gld + q] = Gt[i];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_stream_grid_in(const T *__restrict__ G, int64_t gld, unsigned n, int k,
                                                        int c0, T *__restrict__ Gt)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned q = i / (unsigned)k;
    const int v = (int)(i - q * (unsigned)k);
    Gt[i] = G[(int64_t)(c0 + v) * gld + q];
  }
}

// out[q k + v] = V[q] g[q k + v] (V NULL: 0)
template <typename T>
__global__ __launch_bounds__(256) void k_stream_vmul(const double *__restrict__ V, const T *__restrict__ g,
                                                     unsigned n, int k, T *__restrict__ out)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = V ? T(V[i / (unsigned)k]) * g[i] : T(0);
}

// grid of one streamed term: extents and coordinates of its modes
struct StreamGrid {
  int D;
  int nq[kMaxDev];
  int64_t offX[kMaxDev];  // first grid coordinate of each mode in the x pool
};

// type 10, one i: Q[q k + v] = sum_j c_ij(q) P_j[q k + v] with
// c_ij(q) = C[i D + j] + sum_m A[(i D + j) D + m] x_m(q) (= -1/2 G_ji(q),
// OPSG4:1590-1610); trans takes c_ji (the transposed operator).  P holds D
// tensors of n values.  One workgroup per 256 consecutive values: the c_ij of
// its grid points (at most 256 / k + 2) are formed once in LDS and shared by
// the k vectors of each point.
template <typename T>
__global__ __launch_bounds__(256) void k_stream_gcomb(StreamGrid sg, int i, int trans,
                                                      const double *__restrict__ C,
                                                      const double *__restrict__ A,
                                                      const double *__restrict__ xpool,
                                                      const T *__restrict__ P, unsigned n, int k,
                                                      T *__restrict__ Q)
{
  __shared__ T sc[258 * kMaxDev];
  const int D = sg.D;
  const unsigned e0 = blockIdx.x * 256u;
  const unsigned q0 = e0 / (unsigned)k;
  const unsigned elast = min(n, e0 + 256u) - 1u;
  const int np = (int)(elast / (unsigned)k - q0) + 1;
  for (int t = threadIdx.x; t < np * D; t += blockDim.x) {
    const int pt = t / D, j = t - pt * D;
    const int ij = trans ? j * D + i : i * D + j;
    double c = C[ij];
    if (A) {
      unsigned rem = q0 + (unsigned)pt;
      for (int d = 0; d < D; ++d) {
        const unsigned nx = rem / (unsigned)sg.nq[d];
        c += A[ij * D + d] * xpool[sg.offX[d] + (rem - nx * (unsigned)sg.nq[d])];
        rem = nx;
      }
    }
    sc[t] = T(c);
  }
  __syncthreads();
  const unsigned e = e0 + threadIdx.x;
  if (e >= n) return;
  const T *c = sc + (int)(e / (unsigned)k - q0) * D;
  T acc = T(0);
  for (int j = 0; j < D; ++j) acc = fma(c[j], P[(size_t)j * n + e], acc);
  Q[e] = acc;
}

}  // namespace evr
