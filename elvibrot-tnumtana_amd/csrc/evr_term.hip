// Smolyak-term kernels of the H|psi> engine.
//
// One workgroup applies one Smolyak term to one vector:
//   gather   x = psi[map]                          (BGSG4:1176-1246)
//   KEO      y = sum of Kronecker products of per-mode nb x nb factors
//            (the reference's grid form -1/2 sum_ij d_i G_ij(q) d_j of
//             OPSG4:1548-1650 sandwiched between GDP_TO_BDP and BDP_TO_GDP,
//             factored per mode because W_d B_d = I under the quadrature)
//   V part   y += W^(x) diag(V) B^(x) x on the term grid, streamed through
//            LDS slabs: modes 0..m-1 are contracted inside a slab, modes
//            m..D-1 through one Kronecker row per outer grid point
//            (BDP_TO_GDP BGSG4:2385-2496, V psi OPSG4:1566-1576,
//             GDP_TO_BDP BGSG4:2497-2581)
//   write    R[term rows] = WeightSG * y           (BGSG4:1250-1289, the
//            weighted scatter itself is the deterministic CSR reduction)
#include "evr_kernels.hpp"

namespace evr {

template <typename T>
__device__ __forceinline__ T apply_mode(const T *__restrict__ M, const T *__restrict__ buf,
                                        int e, int bd, int nbd, int st)
{
  // (M along one mode)(buf) at element e: sum_b' M[bd, b'] buf[e + (b'-bd) st]
  const T *row = M + bd;
  const T *col = buf + (e - bd * st);
  T s = T(0);
  for (int b = 0; b < nbd; ++b) s += row[nbd * b] * col[b * st];
  return s;
}

// out[i1 + n1*(q + nout*i3)] = sum_k M[q + nout*k] in[i1 + n1*(k + nin*i3)]
template <typename T, int NT>
__device__ __forceinline__ void contract_lds(const T *__restrict__ in, T *__restrict__ out,
                                             const T *__restrict__ M, int n1, int nin, int nout,
                                             int n3)
{
  const int npen = n1 * n3;
  for (int pp = threadIdx.x; pp < npen; pp += NT) {
    const int i3 = pp / n1;
    const int i1 = pp - i3 * n1;
    T vin[kMaxN];
#pragma unroll
    for (int k = 0; k < kMaxN; ++k)
      if (k < nin) vin[k] = in[i1 + n1 * (k + nin * i3)];
    for (int q = 0; q < nout; ++q) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < kMaxN; ++k)
        if (k < nin) s += M[q + nout * k] * vin[k];
      out[i1 + n1 * (q + nout * i3)] = s;
    }
  }
}

template <int EPT>
struct IGroup {
  static constexpr int value = EPT >= 16 ? 1 : (EPT >= 8 ? 2 : (EPT >= 4 ? 4 : (EPT >= 2 ? 6 : 12)));
};

template <typename T, int NT, int EPT>
__global__ __launch_bounds__(NT) void k_term(TermParams<T> p)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  __shared__ int s_nb[kMaxDev], s_nq[kMaxDev], s_offB[kMaxDev], s_offW[kMaxDev],
      s_offK[kMaxDev], s_st[kMaxDev];

  // XCD-aware decode: consecutive blocks go round-robin over the 8 XCDs;
  // all vectors of one term land on the same XCD (shared V slab in L2) and
  // the cost-sorted terms are dealt to the XCDs in turn.
  const int64_t b = blockIdx.x;
  const int xcd = (int)(b & 7);
  const int64_t slot = b >> 3;
  const int v = (int)(slot % p.nvec);
  const int64_t tr = (slot / p.nvec) * 8 + xcd;
  if (tr >= p.nunits) return;
  const int t = p.order[tr];
  if (t < 0) return;
  const TermDev &td = p.terms[t];
  const int D = p.D;
  const int tid = threadIdx.x;
  const int nbt = td.nbt;

  if (tid == 0) {
    int off = 0, st = 1;
    for (int d = 0; d < D; ++d) {
      const SlotDev &s = p.slots[td.slot[d]];
      s_nb[d] = s.nb;
      s_nq[d] = s.nq;
      s_st[d] = st;
      st *= s.nb;
      s_offB[d] = off;
      off += s.nq * s.nb;
      s_offW[d] = off;
      off += s.nq * s.nb;
      s_offK[d] = off;
      if (p.hasKeo) off += kKeoFactors * s.nb * s.nb;
    }
  }
  __syncthreads();
  // stage the 1D matrices of this term in LDS
  for (int d = 0; d < D; ++d) {
    const SlotDev &s = p.slots[td.slot[d]];
    const int nbq = s.nb * s.nq;
    for (int i = tid; i < nbq; i += NT) {
      smem[s_offB[d] + i] = p.bw[s.offB + i];
      smem[s_offW[d] + i] = p.bw[s.offW + i];
    }
    if (p.hasKeo) {
      const int nk = kKeoFactors * s.nb * s.nb;
      for (int i = tid; i < nk; i += NT) smem[s_offK[d] + i] = p.keo[s.offK + i];
    }
  }
  T *sx = smem + td.lds_mats;
  T *work = sx + nbt;
  const int64_t nvec = p.nvec;
  if (p.mode != kModeG2B) {
    const int *map = p.map + td.off_nb;
    for (int e = tid; e < nbt; e += NT) {
      const int mp = map[e];
      sx[e] = mp >= 0 ? p.X[(int64_t)mp * nvec + v] : T(0);
    }
  }
  __syncthreads();

  T yown[EPT];
  uint64_t dig[EPT];
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    yown[r] = T(0);
    dig[r] = 0;
    int e = tid + r * NT;
    if (e < nbt) {
      int rem = e;
      for (int d = 0; d < D; ++d) {
        const int nbd = s_nb[d];
        const int q = rem / nbd;
        dig[r] |= (uint64_t)(rem - q * nbd) << (5 * d);
        rem = q;
      }
    }
  }

  // ---------------------------------------------------------------- KEO --
  if (p.mode == kModeHpsi && p.hasKeo) {
    constexpr int IG = IGroup<EPT>::value;
    T *zb = work;
    const double *cC = p.coefC;
    const double *cA = p.coefA;
    for (int i0 = 0; i0 < D; i0 += IG) {
      T accM[IG][EPT], accN[IG][EPT], accO[IG][EPT];
#pragma unroll
      for (int ii = 0; ii < IG; ++ii)
#pragma unroll
        for (int r = 0; r < EPT; ++r) accM[ii][r] = accN[ii][r] = accO[ii][r] = T(0);
      for (int j = 0; j < D; ++j) {
        const int nbj = s_nb[j], stj = s_st[j];
        const T *Kj = smem + s_offK[j];
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt) {
            const int bj = (int)((dig[r] >> (5 * j)) & 31);
            zb[e] = apply_mode(Kj + 0 * nbj * nbj, sx, e, bj, nbj, stj);  // P_j x
          }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt) {
            const int bj = (int)((dig[r] >> (5 * j)) & 31);
            const T zj = zb[e];
            T sxv = T(0), qxv = T(0);
            if (p.hasA) {
              sxv = apply_mode(Kj + 4 * nbj * nbj, sx, e, bj, nbj, stj);  // S_j x
              qxv = apply_mode(Kj + 1 * nbj * nbj, sx, e, bj, nbj, stj);  // Q_j x
            }
#pragma unroll
            for (int ii = 0; ii < IG; ++ii) {
              const int i = i0 + ii;
              if (i < D && i != j) {
                accM[ii][r] += T(cC[i * D + j]) * zj;
                if (p.hasA) {
                  accM[ii][r] += T(cA[(i * D + j) * D + j]) * sxv;
                  accN[ii][r] += T(cA[(i * D + j) * D + i]) * zj;
                  accO[ii][r] += T(cA[(i * D + i) * D + j]) * qxv;
                }
              }
            }
            if (p.hasA) {
              for (int k = 0; k < D; ++k) {
                if (k == j) continue;
                const int bk = (int)((dig[r] >> (5 * k)) & 31);
                const int nbk = s_nb[k];
                const T qz = apply_mode(smem + s_offK[k] + 1 * nbk * nbk, zb, e, bk, nbk, s_st[k]);
#pragma unroll
                for (int ii = 0; ii < IG; ++ii) {
                  const int i = i0 + ii;
                  if (i < D && i != j && i != k) accM[ii][r] += T(cA[(i * D + j) * D + k]) * qz;
                }
              }
            }
          }
        }
        __syncthreads();
      }
      // y += P_i M_i + R_i N_i + Tdiag_i O_i + a_iii U_i x for the group
#pragma unroll
      for (int ii = 0; ii < IG; ++ii) {
        const int i = i0 + ii;
        if (i >= D) break;
        const int nbi = s_nb[i], sti = s_st[i];
        const T *Ki = smem + s_offK[i];
        const int nn = nbi * nbi;
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt) zb[e] = accM[ii][r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt)
            yown[r] += apply_mode(Ki + 0 * nn, zb, e, (int)((dig[r] >> (5 * i)) & 31), nbi, sti);
        }
        __syncthreads();
        if (p.hasA) {
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int e = tid + r * NT;
            if (e < nbt) zb[e] = accN[ii][r];
          }
          __syncthreads();
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int e = tid + r * NT;
            if (e < nbt)
              yown[r] += apply_mode(Ki + 3 * nn, zb, e, (int)((dig[r] >> (5 * i)) & 31), nbi, sti);
          }
          __syncthreads();
        }
        const T cii = T(cC[i * D + i]);
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt) zb[e] = accO[ii][r] + cii * sx[e];
        }
        __syncthreads();
        const int fdiag = p.diagT2 ? 6 : 2;
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int e = tid + r * NT;
          if (e < nbt) {
            const int bi = (int)((dig[r] >> (5 * i)) & 31);
            yown[r] += apply_mode(Ki + fdiag * nn, zb, e, bi, nbi, sti);
            if (p.hasA)
              yown[r] += T(cA[(i * D + i) * D + i]) * apply_mode(Ki + 5 * nn, sx, e, bi, nbi, sti);
          }
        }
        __syncthreads();
      }
    }
    if (p.hasF1) {
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int e = tid + r * NT;
        if (e < nbt)
          for (int i = 0; i < D; ++i) {
            const int nbi = s_nb[i];
            yown[r] += T(p.coefF1[i]) *
                       apply_mode(smem + s_offK[i], sx, e, (int)((dig[r] >> (5 * i)) & 31), nbi,
                                  s_st[i]);
          }
      }
    }
  }

  // ------------------------------------------------------------- V part --
  const bool vpart = (p.mode == kModeHpsi && p.hasV) || p.mode == kModeB2G || p.mode == kModeG2B;
  if (vpart) {
    const int m = td.m, QA = td.QA, nbA = td.nbA, nbB = td.nbB;
    T *slab0 = work;
    T *slab1 = work + QA;
    T *crow = work + 2 * QA;
    T *crowW = crow + nbB;
    for (int qB = 0; qB < td.QB; ++qB) {
      for (int bB = tid; bB < nbB; bB += NT) {
        T c = T(1), cw = T(1);
        int rq = qB, rb = bB;
        for (int d = m; d < D; ++d) {
          const int nqd = s_nq[d], nbd = s_nb[d];
          const int qd = rq % nqd;
          rq /= nqd;
          const int bd = rb % nbd;
          rb /= nbd;
          c *= smem[s_offB[d] + qd + nqd * bd];
          cw *= smem[s_offW[d] + bd + nbd * qd];
        }
        crow[bB] = c;
        crowW[bB] = cw;
      }
      __syncthreads();
      T *cur = slab0, *nxt = slab1;
      const int64_t gbase = td.off_nq + (int64_t)QA * qB;
      if (p.mode != kModeG2B) {
        for (int a = tid; a < nbA; a += NT) {
          T s = T(0);
          for (int bB = 0; bB < nbB; ++bB) s += crow[bB] * sx[a + nbA * bB];
          cur[a] = s;
        }
        __syncthreads();
        int n1 = 1, n3 = nbA;
        for (int d = 0; d < m; ++d) {
          const int nin = s_nb[d], nout = s_nq[d];
          n3 /= nin;
          contract_lds<T, NT>(cur, nxt, smem + s_offB[d], n1, nin, nout, n3);
          n1 *= nout;
          T *tmp = cur; cur = nxt; nxt = tmp;
          __syncthreads();
        }
        if (p.mode == kModeB2G) {
          for (int i = tid; i < QA; i += NT) p.G[gbase + i] = cur[i];
          __syncthreads();
          continue;
        }
        for (int i = tid; i < QA; i += NT) cur[i] *= p.V[gbase + i];
        __syncthreads();
      } else {
        for (int i = tid; i < QA; i += NT) cur[i] = p.Gin[gbase + i];
        __syncthreads();
      }
      int n1 = 1, n3 = QA;
      for (int d = 0; d < m; ++d) {
        const int nin = s_nq[d], nout = s_nb[d];
        n3 /= nin;
        contract_lds<T, NT>(cur, nxt, smem + s_offW[d], n1, nin, nout, n3);
        n1 *= nout;
        T *tmp = cur; cur = nxt; nxt = tmp;
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int e = tid + r * NT;
        if (e < nbt) {
          const int bB = e / nbA;
          yown[r] += crowW[bB] * cur[e - bB * nbA];
        }
      }
      __syncthreads();
    }
  }

  if (p.mode == kModeB2G) return;
  T w = T(td.w);
  if (p.mode == kModeG2B && fabs(td.w) < 1e-6) w = T(0);  // BGSG4:1004
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    const int e = tid + r * NT;
    if (e < nbt) p.R[(td.off_nb + e) * nvec + v] = w * yown[r];
  }
}

// Deterministic weighted Smolyak scatter: hpsi[pk, v] = sum over the term
// rows that map to pk (CSR, fixed order) of R.  Output column-major.
template <typename T>
__global__ __launch_bounds__(256) void k_reduce(const T *__restrict__ R, const int *__restrict__ ptr,
                                                const int *__restrict__ idx, int64_t nb_packed,
                                                int nvec, double *__restrict__ out)
{
  const int64_t n = nb_packed * nvec;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int v = (int)(g % nvec);
    const int64_t pk = g / nvec;
    T s = T(0);
    for (int r = ptr[pk]; r < ptr[pk + 1]; ++r) s += R[(int64_t)idx[r] * nvec + v];
    out[(int64_t)v * nb_packed + pk] = (double)s;
  }
}

// psi (column-major, nb_packed x nvec) -> X[pk][v]
template <typename T>
__global__ __launch_bounds__(256) void k_pack_in(const double *__restrict__ psi, int64_t nb_packed,
                                                 int nvec, T *__restrict__ X)
{
  const int64_t n = nb_packed * nvec;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int v = (int)(g % nvec);
    const int64_t pk = g / nvec;
    X[g] = T(psi[(int64_t)v * nb_packed + pk]);
  }
}

// V(q) of the polynomial generator, one workgroup per term.
__global__ __launch_bounds__(256) void k_gen_poly(int D, const TermDev *__restrict__ terms,
                                                  const SlotDev *__restrict__ slots,
                                                  const double *__restrict__ xpool,
                                                  const double *__restrict__ coef, double *V)
{
  // coef layout: v2[D], c3[D^3], d22[D^2], e4[D]
  const TermDev &td = terms[blockIdx.x];
  const double *v2 = coef, *c3 = coef + D, *d22 = c3 + D * D * D, *e4 = d22 + D * D;
  for (int pnt = threadIdx.x; pnt < td.nqt; pnt += blockDim.x) {
    double q[kMaxDev];
    int rem = pnt;
    for (int d = 0; d < D; ++d) {
      const SlotDev &s = slots[td.slot[d]];
      const int id = rem % s.nq;
      rem /= s.nq;
      q[d] = xpool[s.offX + id];
    }
    double val = 0.0;
    for (int i = 0; i < D; ++i) {
      const double qi2 = q[i] * q[i];
      val += v2[i] * qi2 + e4[i] * qi2 * qi2;
      for (int j = i; j < D; ++j) {
        if (j > i) val += d22[i * D + j] * qi2 * q[j] * q[j];
        for (int k = j; k < D; ++k) val += c3[(i * D + j) * D + k] * q[i] * q[j] * q[k];
      }
    }
    V[td.off_nq + pnt] = val;
  }
}

__global__ __launch_bounds__(256) void k_to_float(const double *__restrict__ a, float *__restrict__ b,
                                                  int64_t n)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    b[g] = (float)a[g];
}

// explicit instantiations are produced by the launch table in evr_plan.hip
}  // namespace evr
