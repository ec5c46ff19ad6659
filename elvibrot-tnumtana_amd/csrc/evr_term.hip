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

// Term classes whose kinetic part runs as pencil passes: those with two
// units per thread.  Measured on the D=9 L=8 operator (32 vectors, FP64):
// 512x2x2 21.9 -> 16.2 ms (and 1024x2x2 45.7 -> 42.7 ms, 1024x2x1 52.5 ->
// 45.6 ms, classes since replaced by four-unit ones); the one-unit classes
// (small terms, few pencils per barrier) are faster per unit (0.98 against
// 1.68 ms at 256x1x4), and four units per thread need three LDS buffers
// where the passes need six.
constexpr bool keo_pencil(int nt, int ept, int nvw)
{
  (void)nt;
  (void)nvw;
  return ept == 2;
}

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

// sum_b row[nbd*b] * col[b*st]: one row of a mode factor against a pencil.
// nbd is wave-uniform; the common sizes are unrolled.
template <typename T>
__device__ __forceinline__ T apply_k(const T *__restrict__ row, const T *__restrict__ col, int nbd,
                                     int st)
{
  if (nbd == 1) return row[0] * col[0];
  if (nbd == 3) return fma(row[0], col[0], fma(row[3], col[st], row[6] * col[2 * st]));
  if (nbd == 5)
    return fma(row[0], col[0],
               fma(row[5], col[st], fma(row[10], col[2 * st], fma(row[15], col[3 * st], row[20] * col[4 * st]))));
  T s = T(0);
  for (int b = 0; b < nbd; ++b) s = fma(row[nbd * b], col[b * st], s);
  return s;
}

// A contraction step read from LDS, made wave-uniform (scalar registers:
// the step's matrix is then read with scalar loads)
__device__ __forceinline__ StepDev uniform_step(const StepDev &s)
{
  StepDev r;
  r.n1 = __builtin_amdgcn_readfirstlane(s.n1);
  r.n3 = __builtin_amdgcn_readfirstlane(s.n3);
  r.nin = __builtin_amdgcn_readfirstlane(s.nin);
  r.nout = __builtin_amdgcn_readfirstlane(s.nout);
  r.rn1 = s.rn1;
  r.pad = 0;
  const uint64_t mat = (uint64_t)s.mat;
  r.mat = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(mat >> 32)) << 32) |
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)mat));
  return r;
}

// Integer division of x < 2^24 by a wave-uniform d via the float reciprocal
// (exact after one correction step).
__device__ __forceinline__ int fdiv(int x, int d, float rd)
{
  int q = (int)((float)x * rd);
  int r = x - q * d;
  if (r < 0) --q;
  else if (r >= d) ++q;
  return q;
}

// out[(i1 + n1*(q + nout*i3))*NVW + v] = sum_k M[q + nout*k] in[(i1 + n1*(k + nin*i3))*NVW + v]
// M is read from global memory at wave-uniform addresses (scalar loads, the
// coefficients live in SGPRs); sizes known at compile time are fully unrolled
// and each thread carries one pencil for the NVW vectors of the tile.
template <typename T, int NT, int NVW, int NIN, int NOUT>
__device__ __forceinline__ void contract_fix(const T *__restrict__ in, T *__restrict__ out,
                                             const T *__restrict__ M, int n1, int n3)
{
  const int npen = n1 * n3;
  const float rn1 = 1.0f / (float)n1;
  T m[NIN * NOUT];
#pragma unroll
  for (int i = 0; i < NIN * NOUT; ++i) m[i] = M[i];
  for (int pp = threadIdx.x; pp < npen; pp += NT) {
    const int i3 = fdiv(pp, n1, rn1);
    const int i1 = pp - i3 * n1;
    const T *src = in + (i1 + n1 * NIN * i3) * NVW;
    T *dst = out + (i1 + n1 * NOUT * i3) * NVW;
    T v[NIN][NVW];
#pragma unroll
    for (int k = 0; k < NIN; ++k)
#pragma unroll
      for (int c = 0; c < NVW; ++c) v[k][c] = src[k * n1 * NVW + c];
#pragma unroll
    for (int q = 0; q < NOUT; ++q)
#pragma unroll
      for (int c = 0; c < NVW; ++c) {
        T s = m[q] * v[0][c];
#pragma unroll
        for (int k = 1; k < NIN; ++k) s = fma(m[q + NOUT * k], v[k][c], s);
        dst[q * n1 * NVW + c] = s;
      }
  }
}

template <typename T, int NT, int NVW>
__device__ __forceinline__ void contract_gen(const T *__restrict__ in, T *__restrict__ out,
                                             const T *__restrict__ M, int n1, int nin, int nout,
                                             int n3)
{
  const int npen = n1 * n3 * NVW;
  const int n1v = n1 * NVW;
  const float rn1 = 1.0f / (float)n1v;
  if (nin > kGenCache) {
    // long inputs: each result re-reads its pencil from LDS
    for (int pp = threadIdx.x; pp < npen; pp += NT) {
      const int i3 = fdiv(pp, n1v, rn1);
      const int i1 = pp - i3 * n1v;
      const T *src = in + i1 + n1v * nin * i3;
      for (int q = 0; q < nout; ++q) {
        T s = T(0);
        for (int k = 0; k < nin; ++k) s = fma(M[q + nout * k], src[n1v * k], s);
        out[i1 + n1v * (q + nout * i3)] = s;
      }
    }
    return;
  }
  for (int pp = threadIdx.x; pp < npen; pp += NT) {
    const int i3 = fdiv(pp, n1v, rn1);
    const int i1 = pp - i3 * n1v;  // (i1, v) combined
    T vin[kGenCache];
#pragma unroll
    for (int k = 0; k < kGenCache; ++k)
      if (k < nin) vin[k] = in[i1 + n1v * (k + nin * i3)];
    for (int q = 0; q < nout; ++q) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < kGenCache; ++k)
        if (k < nin) s = fma(M[q + nout * k], vin[k], s);
      out[i1 + n1v * (q + nout * i3)] = s;
    }
  }
}

template <typename T, int NT, int NVW>
__device__ __forceinline__ void contract(const T *__restrict__ in, T *__restrict__ out,
                                         const T *__restrict__ M, int n1, int nin, int nout, int n3)
{
#define EVR_CONTRACT_CASE(A, B) \
  if (nin == A && nout == B) { contract_fix<T, NT, NVW, A, B>(in, out, M, n1, n3); return; }
  EVR_CONTRACT_CASE(1, 2)
  EVR_CONTRACT_CASE(2, 1)
  EVR_CONTRACT_CASE(3, 4)
  EVR_CONTRACT_CASE(4, 3)
  EVR_CONTRACT_CASE(5, 6)
  EVR_CONTRACT_CASE(6, 5)
  EVR_CONTRACT_CASE(1, 1)
  EVR_CONTRACT_CASE(3, 3)
  EVR_CONTRACT_CASE(5, 5)
#undef EVR_CONTRACT_CASE
  contract_gen<T, NT, NVW>(in, out, M, n1, nin, nout, n3);
}

// out (=, or += scale *) the factor K (nb x nb, K[a + nb b] = K(a, b))
// applied along one mode of the unit tensor `in` (mode stride st in units):
// one thread per pencil; the pencil is read from LDS once and the factor, at
// a wave-uniform global address, sits in scalar registers.
template <typename T, int NT, int NB>
__device__ __forceinline__ void pencil_fix(const T *__restrict__ K, const T *__restrict__ in,
                                           T *__restrict__ out, int st, int npen, bool add, T scale)
{
  T k[NB * NB];
#pragma unroll
  for (int i = 0; i < NB * NB; ++i) k[i] = K[i];
  const float rst = 1.0f / (float)st;
  for (int pp = threadIdx.x; pp < npen; pp += NT) {
    const int base = pp + fdiv(pp, st, rst) * st * (NB - 1);
    T v[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) v[b] = in[base + b * st];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      T s = k[a] * v[0];
#pragma unroll
      for (int b = 1; b < NB; ++b) s = fma(k[a + NB * b], v[b], s);
      T *o = out + base + a * st;
      *o = add ? fma(scale, s, *o) : s;
    }
  }
}

// pencil_fix for any nb: the factor read from its LDS copy and the pencil
// re-read per result (few registers; sizes without a fixed variant are rare)
template <typename T, int NT>
__device__ __forceinline__ void pencil_gen(const T *__restrict__ K, const T *__restrict__ in,
                                           T *__restrict__ out, int nb, int st, int npen, bool add,
                                           T scale)
{
  const float rst = 1.0f / (float)st;
  for (int pp = threadIdx.x; pp < npen; pp += NT) {
    const int base = pp + fdiv(pp, st, rst) * st * (nb - 1);
    for (int a = 0; a < nb; ++a) {
      T s = T(0);
      for (int b = 0; b < nb; ++b) s = fma(K[a + nb * b], in[base + b * st], s);
      T *o = out + base + a * st;
      *o = add ? fma(scale, s, *o) : s;
    }
  }
}

// One factor along one mode over nunit units; Kg: the factor in global
// memory (wave-uniform address), Kl: its LDS copy.  Fixed variants for the
// odd sizes of the Hermite levels that fill the large terms (3 and 5
// functions); more fixed sizes cost registers in every term class.
template <typename T, int NT>
__device__ __forceinline__ void pencil_pass(const T *Kg, const T *Kl, const T *in, T *out, int nb, int st,
                                            int nunit, bool add, T scale)
{
  const int npen = nunit / nb;
  switch (nb) {
    case 3: pencil_fix<T, NT, 3>(Kg, in, out, st, npen, add, scale); return;
    case 5: pencil_fix<T, NT, 5>(Kg, in, out, st, npen, add, scale); return;
    default: pencil_gen<T, NT>(Kl, in, out, nb, st, npen, add, scale); return;
  }
}

// Unit u = e*NVW + v handled by thread tid in its r-th slot.  When each thread
// holds as many units as the tile has vectors, it takes the NVW vectors of one
// element (element-major: shared digits, factor rows and coefficient loads);
// otherwise units are dealt round-robin over the threads.
template <int NT, int EPT, int NVW>
__device__ __forceinline__ int unit_of(int tid, int r)
{
  return (NVW > 1 && EPT == NVW) ? tid * EPT + r : tid + r * NT;
}

template <int EPT>
struct IGroup {
  static constexpr int value = EPT >= 8 ? 1 : (EPT >= 4 ? 3 : (EPT >= 2 ? 6 : 12));
};

// One workgroup applies one Smolyak term to a tile of NVW vectors.  LDS
// arrays are [element][vector] (vector fastest); a "unit" u = e*NVW + v.
// Below 1024 threads the register budget is set for four waves per SIMD
// (128 VGPRs; a few bytes of spill at one unit per thread): 512x1x4 2.79 ->
// 1.55 ms, 256x1x4 0.72 -> 0.60 ms at D=9 L=8.
template <typename T, int NT, int EPT, int NVW>
__global__ __launch_bounds__(NT, NT < 1024 ? 4 : 1) void k_term(TermParams<T> p)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  __shared__ int s_nb[kMaxDev], s_nq[kMaxDev], s_offB[kMaxDev], s_offW[kMaxDev],
      s_offK[kMaxDev], s_st[kMaxDev];

  // XCD-aware decode: consecutive blocks go round-robin over the 8 XCDs; all
  // vector tiles of one term land on the same XCD (shared V slab and
  // matrices in its L2) and the cost-sorted terms are dealt to the XCDs in turn.
  // Assembly: one workgroup per (term, first column) work item instead.
  const bool asmb = p.mode == kModeAssemble;
  const bool hmode = p.mode == kModeHpsi || asmb;  // the full operator is applied
  int t, v0;
  if (asmb) {
    const int2 it = p.items[blockIdx.x];
    t = it.x;
    v0 = it.y;
  } else {
    const int nvt = (p.nvec + NVW - 1) / NVW;
    const int64_t b = blockIdx.x;
    const int xcd = (int)(b & 7);
    const int64_t slot = b >> 3;
    const int vt = (int)(slot % nvt);
    const int64_t tr = (slot / nvt) * 8 + xcd;
    if (tr >= p.nunits) return;
    t = p.order[tr];
    v0 = vt * NVW;
  }
  if (t < 0) return;
  const TermDev &td = p.terms[t];
  const int D = p.D;
  const int tid = threadIdx.x;
  const int nbt = td.nbt;
  const int nunit = nbt * NVW;
  const int64_t nvec = asmb ? nbt : p.nvec;  // assembly: columns of the term

  // slot table of the term: one load per mode in parallel, then the LDS
  // layout (prefix sums) by one thread from LDS
  __shared__ int64_t s_gB[kMaxDev], s_gW[kMaxDev], s_gK[kMaxDev];
  __shared__ int s_mats;
  if (tid < D) {
    const SlotDev &s = p.slots[td.slot[tid]];
    s_nb[tid] = s.nb;
    s_nq[tid] = s.nq;
    s_gB[tid] = s.offB;
    s_gW[tid] = s.offW;
    s_gK[tid] = s.offK;
  }
  __syncthreads();
  if (tid == 0) {
    int off = 0, st = NVW;
    for (int d = 0; d < D; ++d) {
      const int nb = s_nb[d], nq = s_nq[d];
      s_st[d] = st;  // unit stride of mode d
      st *= nb;
      s_offB[d] = off;
      off += nq * nb;
      s_offW[d] = off;
      off += nq * nb;
      s_offK[d] = off;
      if (p.hasKeo) off += kKeoFactors * nb * nb;
    }
    s_mats = off;
  }
  __syncthreads();
  // stage the 1D matrices of this term in LDS: one flat pass over all modes
  // (every load in flight at once, not one mode after another)
  for (int i = tid; i < s_mats; i += NT) {
    int d = 0;
    while (d + 1 < D && s_offB[d + 1] <= i) ++d;
    const int nbq = s_nb[d] * s_nq[d];
    const int l = i - s_offB[d];
    smem[i] = l < nbq ? p.bw[s_gB[d] + l]
                      : (l < 2 * nbq ? p.bw[s_gW[d] + (l - nbq)] : p.keo[s_gK[d] + (l - 2 * nbq)]);
  }
  T *sx = smem + td.lds_mats;
  T *work = sx + nunit;
  if (asmb) {
    // unit vector e_{v0+c} of the term; a function pruned from the packed
    // basis has no input value, so its column of the assembled matrix is zero
    const int *map = p.map + td.off_nb;
    for (int u = tid; u < nunit; u += NT) {
      const int e = u / NVW, c = u - e * NVW;
      sx[u] = (e == v0 + c && map[e] >= 0) ? T(1) : T(0);
    }
  } else if (p.mode != kModeG2B) {
    const int *map = p.map + td.off_nb;
    for (int u = tid; u < nunit; u += NT) {
      const int e = u / NVW, c = u - e * NVW;
      const int mp = map[e];
      sx[u] = (mp >= 0 && v0 + c < nvec) ? p.X[(int64_t)mp * nvec + v0 + c] : T(0);
    }
  }
  __syncthreads();

  T yown[EPT];
  uint64_t dig[EPT];
  constexpr bool kElementMajor = NVW > 1 && EPT == NVW;
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    yown[r] = T(0);
    dig[r] = 0;
    const int u = unit_of<NT, EPT, NVW>(tid, r);
    if (kElementMajor && r > 0) {
      dig[r] = dig[0];  // same element, another vector
    } else if (u < nunit) {
      int rem = u / NVW;
      for (int d = 0; d < D; ++d) {
        const int nbd = s_nb[d];
        const int q = rem / nbd;
        dig[r] |= (uint64_t)(rem - q * nbd) << (5 * d);
        rem = q;
      }
    }
  }

  // ---------------------------------------------------------------- KEO --
  // Two forms of the kinetic part over the active modes for up to four
  // units per thread, chosen per term class (keo_pencil): one-mode factors
  // applied by pencil passes through LDS buffers (a thread per pencil, the
  // factor in scalar registers; fewer LDS reads, more barriers), or per unit
  // (each result reads its pencil and factor row; three LDS buffers).
  // Accumulators M_i, N_i, O_i of a group of IG active modes live in
  // registers per unit (IG 2 at four units: 127 VGPRs, no spill; 1024x4x1
  // 19.6 -> 11.9 ms against the all-mode form below at D=9 L=8); the
  // coefficient strips are pre-zeroed on the host (no branches, no index
  // tests) and read at wave-uniform addresses.  Eight units per thread keep
  // the all-mode form (three buffers of 8,192 units exceed the LDS).
  if constexpr (EPT <= 4) {
  if constexpr (keo_pencil(NT, EPT, NVW)) {
    if (hmode && p.hasKeo) {
      T *Zb = work;              // P_j x
      T *Sb = work + nunit;      // S_j x; M_i stage
      T *Qb = work + 2 * nunit;  // Q_j x; N_i stage
      T *Wb = work + 3 * nunit;  // Q_k Z_j (two buffers, alternating); O_i stage
      T *Yb = work + 5 * nunit;  // sum of the final factor passes
      // this term's coefficient set (strips in its mode positions)
      const double *kS = p.keoc + (int64_t)td.cix * p.keoc_stride;  // [j*Dp + i]
      const double *kN = kS + kDp * kDp;            // [j*Dp + i]
      const double *kO = kN + kDp * kDp;            // [k*Dp + i]
      const double *kC = kO + kDp * kDp;            // [j*Dp + i]
      const double *kCd = kC + kDp * kDp;           // [i]
      const double *kU = kCd + kDp;                 // [i]
      const double *kF = kU + kDp;                  // [i]
      const double *kA = kF + kDp;                  // [(j*Dp + k)*Dp + i], with Ga only
      const int fdiag = p.diagT2 ? 6 : 2;
      // Modes outside amode[] have one function and vanishing P, Q, R (odd
      // integrands on the symmetric quadrature): every product holding one of
      // those factors drops, the others fold into scalars of this term:
      //   c_ii Td_i + a_iii U_i           -> s_f0 x
      //   a_iik Td_i Q_k  (k active)      -> s_fq[k] Q_k x
      //   a_ijj P_i S_j   (i active)      -> s_fs[i] added to M_i
      __shared__ T s_fq[kMaxDev], s_fs[kMaxDev], s_f0;
      __shared__ int s_koff[kMaxDev];  // offset of mode d's factors in p.keo
      const int na = td.na;
      if (tid == 0) {
        bool act[kMaxDev];
        for (int d = 0; d < kMaxDev; ++d) {
          act[d] = false;
          s_fq[d] = T(0);
          s_fs[d] = T(0);
          s_koff[d] = d < D ? p.slots[td.slot[d]].offK : 0;
        }
        for (int a = 0; a < na; ++a) act[td.amode[a]] = true;
        T f0 = T(0);
        for (int i = 0; i < D; ++i) {
          if (act[i]) continue;
          const T *Ki = smem + s_offK[i];  // 1 x 1 factors
          f0 += T(kCd[i]) * Ki[fdiag];
          if (p.hasA) {
            f0 += T(kU[i]) * Ki[5];
            for (int k = 0; k < D; ++k)
              if (act[k]) {
                s_fq[k] += T(kO[k * kDp + i]) * Ki[fdiag];
                s_fs[k] += T(kS[i * kDp + k]) * Ki[4];
              }
          }
        }
        s_f0 = f0;
      }
      for (int u = tid; u < nunit; u += NT) Yb[u] = T(0);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int u = unit_of<NT, EPT, NVW>(tid, r);
        if (u < nunit) yown[r] = s_f0 * sx[u];
      }
      // factor f of mode d applied along d: out = K in, or out += scale K in
      auto pen = [&](int d, int f, const T *in, T *out, bool add, T scale) {
        const int nbd = s_nb[d];
        const int off = __builtin_amdgcn_readfirstlane(s_koff[d] + f * nbd * nbd);
        pencil_pass<T, NT>(p.keo + off, smem + s_offK[d] + f * nbd * nbd, in, out, nbd, s_st[d], nunit,
                           add, scale);
      };
      auto keo_groups = [&](auto igc) {
      constexpr int IG = decltype(igc)::value;
      for (int i0 = 0; i0 < na; i0 += IG) {
        T accM[IG][EPT], accN[IG][EPT], accO[IG][EPT];
#pragma unroll
        for (int ii = 0; ii < IG; ++ii)
#pragma unroll
          for (int r = 0; r < EPT; ++r) accM[ii][r] = accN[ii][r] = accO[ii][r] = T(0);
        // strip column of each accumulator: the active mode, or the all-zero
        // padding column kDp - 1 past the end of the list
        int icol[IG];
#pragma unroll
        for (int ii = 0; ii < IG; ++ii) icol[ii] = i0 + ii < na ? td.amode[i0 + ii] : kDp - 1;
        for (int ja = 0; ja < na; ++ja) {
          const int j = td.amode[ja];
          pen(j, 0, sx, Zb, false, T(1));              // Z_j = P_j x
          if (p.hasA) {
            pen(j, 4, sx, Sb, false, T(1));            // S_j x
            pen(j, 1, sx, Qb, false, T(1));            // Q_j x
          }
          __syncthreads();
          {
            T cz[IG], cs[IG], cn[IG], co[IG];
            const T fq = s_fq[j];
#pragma unroll
            for (int ii = 0; ii < IG; ++ii) {
              cz[ii] = T(kC[j * kDp + icol[ii]]);
              cs[ii] = T(kS[j * kDp + icol[ii]]);
              cn[ii] = T(kN[j * kDp + icol[ii]]);
              co[ii] = T(kO[j * kDp + icol[ii]]);
            }
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
              const int u = unit_of<NT, EPT, NVW>(tid, r);
              if (u < nunit) {
                const T zj = Zb[u];
#pragma unroll
                for (int ii = 0; ii < IG; ++ii) accM[ii][r] = fma(cz[ii], zj, accM[ii][r]);
                if (p.hasA) {
                  const T sxv = Sb[u], qxv = Qb[u];
#pragma unroll
                  for (int ii = 0; ii < IG; ++ii) {
                    accM[ii][r] = fma(cs[ii], sxv, accM[ii][r]);
                    accN[ii][r] = fma(cn[ii], zj, accN[ii][r]);
                    accO[ii][r] = fma(co[ii], qxv, accO[ii][r]);
                  }
                  if (i0 == 0) yown[r] = fma(fq, qxv, yown[r]);
                }
              }
            }
          }
          if (p.hasA) {
            // Q_k Z_j for the other active modes, into two alternating
            // buffers: a buffer is rewritten only after the barrier that
            // follows the pass after its last read
            int nk = 0;
            for (int ka = 0; ka < na; ++ka) {
              const int k = td.amode[ka];
              if (k == j) continue;
              T *Wk = Wb + (nk & 1) * nunit;
              ++nk;
              pen(k, 1, Zb, Wk, false, T(1));
              __syncthreads();
              T ca[IG];
#pragma unroll
              for (int ii = 0; ii < IG; ++ii) ca[ii] = T(kA[(j * kDp + k) * kDp + icol[ii]]);
#pragma unroll
              for (int r = 0; r < EPT; ++r) {
                const int u = unit_of<NT, EPT, NVW>(tid, r);
                if (u < nunit) {
                  const T qz = Wk[u];
#pragma unroll
                  for (int ii = 0; ii < IG; ++ii) accM[ii][r] = fma(ca[ii], qz, accM[ii][r]);
                }
              }
            }
          }
          __syncthreads();
        }
        // Y += P_i (M_i + (s_fs_i + F1_i) x) + R_i N_i + Tdiag_i (O_i + c_ii x) + a_iii U_i x
#pragma unroll
        for (int ii = 0; ii < IG; ++ii) {
          if (i0 + ii >= na) break;
          const int i = icol[ii];
          const T cd = T(kCd[i]), cu = T(kU[i]), fs = s_fs[i] + (p.hasF1 ? T(kF[i]) : T(0));
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit) {
              Sb[u] = fma(fs, sx[u], accM[ii][r]);
              Qb[u] = accN[ii][r];
              Wb[u] = fma(cd, sx[u], accO[ii][r]);
            }
          }
          __syncthreads();
          pen(i, 0, Sb, Yb, true, T(1));              // P_i M_i
          pen(i, fdiag, Wb, Yb, true, T(1));          // T_i O_i
          if (p.hasA) {
            pen(i, 3, Qb, Yb, true, T(1));            // R_i N_i
            pen(i, 5, sx, Yb, true, cu);              // a_iii U_i x
          }
          __syncthreads();
        }
      }
      };
      // group size fitted to the active modes (no accumulators for absent ones)
      if constexpr (EPT == 1) {
        if (na <= 4) keo_groups(std::integral_constant<int, 4>{});
        else keo_groups(std::integral_constant<int, 8>{});
      } else {
        if (na <= 3) keo_groups(std::integral_constant<int, 3>{});
        else keo_groups(std::integral_constant<int, 6>{});
      }
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int u = unit_of<NT, EPT, NVW>(tid, r);
        if (u < nunit) yown[r] += Yb[u];
      }
    }
  } else {
    if (hmode && p.hasKeo) {
      T *zb = work;
      T *zb2 = work + nunit;
      T *zb3 = work + 2 * nunit;
      // this term's coefficient set (strips in its mode positions)
      const double *kS = p.keoc + (int64_t)td.cix * p.keoc_stride;  // [j*Dp + i]
      const double *kN = kS + kDp * kDp;            // [j*Dp + i]
      const double *kO = kN + kDp * kDp;            // [k*Dp + i]
      const double *kC = kO + kDp * kDp;            // [j*Dp + i]
      const double *kCd = kC + kDp * kDp;           // [i]
      const double *kU = kCd + kDp;                 // [i]
      const double *kF = kU + kDp;                  // [i]
      const double *kA = kF + kDp;                  // [(j*Dp + k)*Dp + i], with Ga only
      // first unit of the mode-d pencil through unit u (slot r)
      auto ubase = [&](int r, int u, int d) { return u - (int)((dig[r] >> (5 * d)) & 31) * s_st[d]; };
      const int fdiag = p.diagT2 ? 6 : 2;
      // Modes outside amode[] have one function and vanishing P, Q, R (odd
      // integrands on the symmetric quadrature): every product holding one of
      // those factors drops, the others fold into scalars of this term:
      //   c_ii Td_i + a_iii U_i           -> s_f0 x
      //   a_iik Td_i Q_k  (k active)      -> s_fq[k] Q_k x
      //   a_ijj P_i S_j   (i active)      -> s_fs[i] added to M_i
      __shared__ T s_fq[kMaxDev], s_fs[kMaxDev], s_f0;
      const int na = td.na;
      if (tid == 0) {
        bool act[kMaxDev];
        for (int d = 0; d < kMaxDev; ++d) {
          act[d] = false;
          s_fq[d] = T(0);
          s_fs[d] = T(0);
        }
        for (int a = 0; a < na; ++a) act[td.amode[a]] = true;
        T f0 = T(0);
        for (int i = 0; i < D; ++i) {
          if (act[i]) continue;
          const T *Ki = smem + s_offK[i];  // 1 x 1 factors
          f0 += T(kCd[i]) * Ki[fdiag];
          if (p.hasA) {
            f0 += T(kU[i]) * Ki[5];
            for (int k = 0; k < D; ++k)
              if (act[k]) {
                s_fq[k] += T(kO[k * kDp + i]) * Ki[fdiag];
                s_fs[k] += T(kS[i * kDp + k]) * Ki[4];
              }
          }
        }
        s_f0 = f0;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int u = unit_of<NT, EPT, NVW>(tid, r);
        if (u < nunit) yown[r] = s_f0 * sx[u];
      }
      auto keo_groups = [&](auto igc) {
      constexpr int IG = decltype(igc)::value;
      for (int i0 = 0; i0 < na; i0 += IG) {
        T accM[IG][EPT], accN[IG][EPT], accO[IG][EPT];
#pragma unroll
        for (int ii = 0; ii < IG; ++ii)
#pragma unroll
          for (int r = 0; r < EPT; ++r) accM[ii][r] = accN[ii][r] = accO[ii][r] = T(0);
        // strip column of each accumulator: the active mode, or the all-zero
        // padding column kDp - 1 past the end of the list
        int icol[IG];
#pragma unroll
        for (int ii = 0; ii < IG; ++ii) icol[ii] = i0 + ii < na ? td.amode[i0 + ii] : kDp - 1;
        for (int ja = 0; ja < na; ++ja) {
          const int j = td.amode[ja];
          const int nbj = s_nb[j], stj = s_st[j], nnj = nbj * nbj;
          const T *Kj = smem + s_offK[j];
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit) {
              const int bj = (int)((dig[r] >> (5 * j)) & 31);
              zb[u] = apply_k(Kj + bj, sx + ubase(r, u, j), nbj, stj);  // Z_j = P_j x
            }
          }
          __syncthreads();
          {
            T cz[IG], cs[IG], cn[IG], co[IG];
            const T fq = s_fq[j];
#pragma unroll
            for (int ii = 0; ii < IG; ++ii) {
              cz[ii] = T(kC[j * kDp + icol[ii]]);
              cs[ii] = T(kS[j * kDp + icol[ii]]);
              cn[ii] = T(kN[j * kDp + icol[ii]]);
              co[ii] = T(kO[j * kDp + icol[ii]]);
            }
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
              const int u = unit_of<NT, EPT, NVW>(tid, r);
              if (u < nunit) {
                const T zj = zb[u];
#pragma unroll
                for (int ii = 0; ii < IG; ++ii) accM[ii][r] = fma(cz[ii], zj, accM[ii][r]);
                if (p.hasA) {
                  const int bj = (int)((dig[r] >> (5 * j)) & 31);
                  const T sxv = apply_k(Kj + 4 * nnj + bj, sx + ubase(r, u, j), nbj, stj);  // S_j x
                  const T qxv = apply_k(Kj + nnj + bj, sx + ubase(r, u, j), nbj, stj);      // Q_j x
#pragma unroll
                  for (int ii = 0; ii < IG; ++ii) {
                    accM[ii][r] = fma(cs[ii], sxv, accM[ii][r]);
                    accN[ii][r] = fma(cn[ii], zj, accN[ii][r]);
                    accO[ii][r] = fma(co[ii], qxv, accO[ii][r]);
                  }
                  if (i0 == 0) yown[r] = fma(fq, qxv, yown[r]);
                }
              }
            }
          }
          if (p.hasA) {
            for (int ka = 0; ka < na; ++ka) {
              const int k = td.amode[ka];
              if (k == j) continue;
              const int nbk = s_nb[k], stk = s_st[k];
              const T *Qk = smem + s_offK[k] + nbk * nbk;
              T ca[IG];
#pragma unroll
              for (int ii = 0; ii < IG; ++ii) ca[ii] = T(kA[(j * kDp + k) * kDp + icol[ii]]);
#pragma unroll
              for (int r = 0; r < EPT; ++r) {
                const int u = unit_of<NT, EPT, NVW>(tid, r);
                if (u < nunit) {
                  const int bk = (int)((dig[r] >> (5 * k)) & 31);
                  const T qz = apply_k(Qk + bk, zb + ubase(r, u, k), nbk, stk);  // Q_k Z_j
#pragma unroll
                  for (int ii = 0; ii < IG; ++ii) accM[ii][r] = fma(ca[ii], qz, accM[ii][r]);
                }
              }
            }
          }
          __syncthreads();
        }
        // y += P_i M_i + R_i N_i + Tdiag_i (O_i + c_ii x) + a_iii U_i x + F1_i P_i x
#pragma unroll
        for (int ii = 0; ii < IG; ++ii) {
          if (i0 + ii >= na) break;
          const int i = icol[ii];
          const T cd = T(kCd[i]), cu = T(kU[i]), cf = T(kF[i]), fs = s_fs[i];
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit) {
              zb[u] = fma(fs, sx[u], accM[ii][r]);
              zb2[u] = accN[ii][r];
              zb3[u] = fma(cd, sx[u], accO[ii][r]);
            }
          }
          __syncthreads();
          const int nbi = s_nb[i], sti = s_st[i], nn = nbi * nbi;
          const T *Ki = smem + s_offK[i];
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit) {
              const int bi = (int)((dig[r] >> (5 * i)) & 31);
              const int bo = ubase(r, u, i);
              T yv = apply_k(Ki + bi, zb + bo, nbi, sti);                        // P_i M_i
              yv += apply_k(Ki + fdiag * nn + bi, zb3 + bo, nbi, sti);            // T_i O_i
              if (p.hasA) {
                yv += apply_k(Ki + 3 * nn + bi, zb2 + bo, nbi, sti);              // R_i N_i
                yv = fma(cu, apply_k(Ki + 5 * nn + bi, sx + bo, nbi, sti), yv);   // U_i x
              }
              if (p.hasF1) yv = fma(cf, apply_k(Ki + bi, sx + bo, nbi, sti), yv); // F1_i P_i x
              yown[r] += yv;
            }
          }
          __syncthreads();
        }
      }
      };
      // group size fitted to the active modes (no accumulators for absent ones)
      if constexpr (EPT == 1) {
        if (na <= 4) keo_groups(std::integral_constant<int, 4>{});
        else keo_groups(std::integral_constant<int, 8>{});
      } else if constexpr (EPT == 2) {
        if (na <= 3) keo_groups(std::integral_constant<int, 3>{});
        else keo_groups(std::integral_constant<int, 5>{});
      } else {
        if (na <= 2) keo_groups(std::integral_constant<int, 2>{});
        else keo_groups(std::integral_constant<int, 2>{});
      }
    }
  }
  } else {
  if (hmode && p.hasKeo) {
    constexpr int IG = IGroup<EPT>::value;
    T *zb = work;
    const double *cC = p.coefC + (int64_t)td.cix * D * D;  // this term's coefficient set
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
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) {
            const int bj = (int)((dig[r] >> (5 * j)) & 31);
            zb[u] = apply_mode(Kj + 0 * nbj * nbj, sx, u, bj, nbj, stj);  // P_j x
          }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) {
            const int bj = (int)((dig[r] >> (5 * j)) & 31);
            const T zj = zb[u];
            T sxv = T(0), qxv = T(0);
            if (p.hasA) {
              sxv = apply_mode(Kj + 4 * nbj * nbj, sx, u, bj, nbj, stj);  // S_j x
              qxv = apply_mode(Kj + 1 * nbj * nbj, sx, u, bj, nbj, stj);  // Q_j x
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
                const T qz = apply_mode(smem + s_offK[k] + 1 * nbk * nbk, zb, u, bk, nbk, s_st[k]);
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
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) zb[u] = accM[ii][r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit)
            yown[r] += apply_mode(Ki + 0 * nn, zb, u, (int)((dig[r] >> (5 * i)) & 31), nbi, sti);
        }
        __syncthreads();
        if (p.hasA) {
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit) zb[u] = accN[ii][r];
          }
          __syncthreads();
#pragma unroll
          for (int r = 0; r < EPT; ++r) {
            const int u = unit_of<NT, EPT, NVW>(tid, r);
            if (u < nunit)
              yown[r] += apply_mode(Ki + 3 * nn, zb, u, (int)((dig[r] >> (5 * i)) & 31), nbi, sti);
          }
          __syncthreads();
        }
        const T cii = T(cC[i * D + i]);
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) zb[u] = accO[ii][r] + cii * sx[u];
        }
        __syncthreads();
        const int fdiag = p.diagT2 ? 6 : 2;
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) {
            const int bi = (int)((dig[r] >> (5 * i)) & 31);
            yown[r] += apply_mode(Ki + fdiag * nn, zb, u, bi, nbi, sti);
            if (p.hasA)
              yown[r] += T(cA[(i * D + i) * D + i]) * apply_mode(Ki + 5 * nn, sx, u, bi, nbi, sti);
          }
        }
        __syncthreads();
      }
    }
    if (p.hasF1) {
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int u = unit_of<NT, EPT, NVW>(tid, r);
        if (u < nunit)
          for (int i = 0; i < D; ++i) {
            const int nbi = s_nb[i];
            yown[r] += T(p.coefF1[(int64_t)td.cix * D + i]) *
                       apply_mode(smem + s_offK[i], sx, u, (int)((dig[r] >> (5 * i)) & 31), nbi,
                                  s_st[i]);
          }
      }
    }
  }

  }

  // ------------------------------------------------------------- V part --
  const bool vpart = (hmode && p.hasV) || p.mode == kModeB2G || p.mode == kModeG2B;
  if (vpart) {
    // The slab walks the reduced grid: the modes with more than one function.
    // H|psi> has the one-function modes folded into V_eff (p.V); the basis
    // <-> grid transforms expand the reduced values over those modes on the
    // way out (B->G: value x prod_d B_d[q_d, 0]) and contract them on the way
    // in (G->B: sum over their points of value x prod_d W_d[0, q_d]), so the
    // contractions run on the reduced grid only.
    const int Dg = td.Dv;
    const int m = td.vm, QA = td.vQA, nbA = td.vnbA, nbB = td.vnbB, QBg = td.vQB;
    const int64_t goff = hmode ? td.off_vq : td.off_nq;
    T *slab0 = work;
    T *slab1 = work + QA * NVW;
    T *crow = work + 2 * QA * NVW;
    T *crowW = crow + nbB;
    __shared__ int s_qd[kMaxDev];  // digits of the current outer grid point
    __shared__ int sv_nb[kMaxDev], sv_nq[kMaxDev], sv_offB[kMaxDev], sv_offW[kMaxDev];
    // transforms: full-grid stride of each reduced position, the one-function
    // modes (points, full-grid stride, B and W offsets in LDS), the count of
    // their point combinations and the full-grid offset of the outer point
    __shared__ int s_fst[kMaxDev], s_inq[kMaxDev], s_ifst[kMaxDev], s_ioB[kMaxDev], s_ioW[kMaxDev];
    __shared__ int s_ni, s_ncomb, s_bofs;
    if (tid < kMaxDev) s_qd[tid] = 0;
    if (tid < Dg) {
      const int d = td.vmode[tid];
      sv_nb[tid] = s_nb[d];
      sv_nq[tid] = s_nq[d];
      sv_offB[tid] = s_offB[d];
      sv_offW[tid] = s_offW[d];
    }
    if (tid == 0 && !hmode) {
      int fst = 1, ni = 0, nc = 1, k = 0;
      for (int d = 0; d < D; ++d) {
        if (k < Dg && td.vmode[k] == d) {
          s_fst[k++] = fst;
        } else {
          s_inq[ni] = s_nq[d];
          s_ifst[ni] = fst;
          s_ioB[ni] = s_offB[d];
          s_ioW[ni] = s_offW[d];
          nc *= s_nq[d];
          ++ni;
        }
        fst *= s_nq[d];
      }
      s_ni = ni;
      s_ncomb = nc;
      s_bofs = 0;
    }
    __syncthreads();
    // per-unit offsets of the outer accumulation (no division inside leaves)
    int aoff[EPT], boff[EPT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int u = unit_of<NT, EPT, NVW>(tid, r);
      const int e = u / NVW;
      const int bB = e / nbA;
      boff[r] = bB;
      aoff[r] = u - bB * nbA * NVW;
    }
    // outer basis digits of the crow entry this thread computes (bB = tid)
    uint64_t bdig = 0;
    {
      int rb = tid;
      for (int d = m; d < Dg; ++d) {
        const int nbd = sv_nb[d];
        const int q = rb / nbd;
        bdig |= (uint64_t)(rb - q * nbd) << (5 * d);
        rb = q;
      }
    }
    // the 2m contraction steps of the slab, staged in LDS by one parallel
    // load (read back as wave-uniform values)
    __shared__ StepDev s_stp[2 * kMaxDev];
    {
      const StepDev *stp = p.steps + td.vstep0;
      if (tid < 2 * m) s_stp[tid] = stp[tid];
    }
    __syncthreads();
    for (int qB = 0; qB < QBg; ++qB) {
      for (int bB = tid; bB < nbB; bB += NT) {
        uint64_t bd64 = bdig;
        if (bB != tid) {  // more outer functions than threads: digits on the fly
          bd64 = 0;
          int rb = bB;
          for (int d = m; d < Dg; ++d) {
            const int nbd = sv_nb[d];
            const int q = rb / nbd;
            bd64 |= (uint64_t)(rb - q * nbd) << (5 * d);
            rb = q;
          }
        }
        T c = T(1), cw = T(1);
        for (int d = m; d < Dg; ++d) {
          const int nqd = sv_nq[d], nbd = sv_nb[d];
          const int qd = s_qd[d];
          const int bd = (int)((bd64 >> (5 * d)) & 31);
          c *= smem[sv_offB[d] + qd + nqd * bd];
          cw *= smem[sv_offW[d] + bd + nbd * qd];
        }
        crow[bB] = c;
        crowW[bB] = cw;
      }
      __syncthreads();
      T *cur = slab0, *nxt = slab1;
      const int64_t gbase = goff + (int64_t)QA * qB;
      if (p.mode != kModeG2B) {
        const int na = nbA * NVW;
        for (int a2 = tid; a2 < na; a2 += NT) {
          const T *xs = sx + a2;
          T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
          int bB = 0;
          for (; bB + 4 <= nbB; bB += 4) {
            s0 = fma(crow[bB], xs[na * bB], s0);
            s1 = fma(crow[bB + 1], xs[na * (bB + 1)], s1);
            s2 = fma(crow[bB + 2], xs[na * (bB + 2)], s2);
            s3 = fma(crow[bB + 3], xs[na * (bB + 3)], s3);
          }
          for (; bB < nbB; ++bB) s0 = fma(crow[bB], xs[na * bB], s0);
          cur[a2] = (s0 + s1) + (s2 + s3);
        }
        __syncthreads();
        for (int d = 0; d < m; ++d) {
          const StepDev st = uniform_step(s_stp[d]);
          contract<T, NT, NVW>(cur, nxt, p.bw + st.mat, st.n1, st.nin, st.nout, st.n3);
          T *tmp = cur; cur = nxt; nxt = tmp;
          __syncthreads();
        }
        if (p.mode == kModeB2G) {
          // reduced point i of this slab -> every full-grid point over the
          // one-function modes (combination outermost: neighbouring threads
          // write neighbouring reduced points)
          const int ni = s_ni, nc = s_ncomb;
          for (int cb = 0; cb < nc; ++cb) {
            int r2 = cb, coff = s_bofs;
            T sc = T(1);
            for (int k = 0; k < ni; ++k) {
              const int q = r2 % s_inq[k];
              r2 /= s_inq[k];
              coff += q * s_ifst[k];
              sc *= smem[s_ioB[k] + q];
            }
            for (int i2 = tid; i2 < QA * NVW; i2 += NT) {
              const int i = i2 / NVW, c = i2 - i * NVW;
              if (v0 + c >= nvec) continue;
              int rem = i, off = coff;
              for (int k = 0; k < m; ++k) {
                const int q = rem % sv_nq[k];
                rem /= sv_nq[k];
                off += q * s_fst[k];
              }
              p.G[(int64_t)(v0 + c) * p.gld + goff + off] = sc * cur[i2];
            }
          }
        } else {
          for (int i2 = tid; i2 < QA * NVW; i2 += NT) cur[i2] *= p.V[gbase + i2 / NVW];
        }
        __syncthreads();
      } else {
        // grid input: column v0 + c at Gin + (v0 + c) * gld (lanes past nvec
        // read 0), contracted over the one-function modes with their W rows
        const int ni = s_ni, nc = s_ncomb;
        for (int i2 = tid; i2 < QA * NVW; i2 += NT) {
          const int i = i2 / NVW, c = i2 - i * NVW;
          T acc = T(0);
          if (v0 + c < nvec) {
            int rem = i, off = s_bofs;
            for (int k = 0; k < m; ++k) {
              const int q = rem % sv_nq[k];
              rem /= sv_nq[k];
              off += q * s_fst[k];
            }
            const T *Gc = p.Gin + (int64_t)(v0 + c) * p.gld + goff + off;
            for (int cb = 0; cb < nc; ++cb) {
              int r2 = cb, coff = 0;
              T sc = T(1);
              for (int k = 0; k < ni; ++k) {
                const int q = r2 % s_inq[k];
                r2 /= s_inq[k];
                coff += q * s_ifst[k];
                sc *= smem[s_ioW[k] + q];
              }
              acc = fma(sc, Gc[coff], acc);
            }
          }
          cur[i2] = acc;
        }
        __syncthreads();
      }
      if (p.mode != kModeB2G) {
        for (int d = 0; d < m; ++d) {
          const StepDev st = uniform_step(s_stp[m + d]);
          contract<T, NT, NVW>(cur, nxt, p.bw + st.mat, st.n1, st.nin, st.nout, st.n3);
          T *tmp = cur; cur = nxt; nxt = tmp;
          __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const int u = unit_of<NT, EPT, NVW>(tid, r);
          if (u < nunit) yown[r] = fma(crowW[boff[r]], cur[aoff[r]], yown[r]);
        }
      }
      // next outer grid point (digits of modes m..D-1, mode m fastest)
      if (tid == 0) {
        for (int d = m; d < Dg; ++d) {
          if (++s_qd[d] < sv_nq[d]) break;
          s_qd[d] = 0;
        }
        if (!hmode) {
          int bo = 0;
          for (int d = m; d < Dg; ++d) bo += s_qd[d] * s_fst[d];
          s_bofs = bo;
        }
      }
      __syncthreads();
    }
  }

  if (p.mode == kModeB2G) return;
  if (p.mode == kModeG2B && p.diagKeo && p.hasKeo) {
    // diagonal of the kinetic part: sum_i c_ii Tdiag_i[b_i,b_i] + a_iii U_i[b_i,b_i]
    // (the other factor products vanish on the diagonal: P and Q have a zero diagonal)
    const double *kCd = p.keoc + (int64_t)td.cix * p.keoc_stride + 4 * kDp * kDp, *kU = kCd + kDp;
    const int fdiag = p.diagT2 ? 6 : 2;
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int u = unit_of<NT, EPT, NVW>(tid, r);
      if (u < nunit)
        for (int i = 0; i < D; ++i) {
          const int nbi = s_nb[i], bi = (int)((dig[r] >> (5 * i)) & 31), nn = nbi * nbi;
          const T *Ki = smem + s_offK[i];
          yown[r] += T(kCd[i]) * Ki[fdiag * nn + bi + nbi * bi] + T(kU[i]) * Ki[5 * nn + bi + nbi * bi];
        }
    }
  }
  T w = T(td.w);
  if (p.mode == kModeG2B && !p.diagKeo && fabs(td.w) < 1e-6) w = T(0);  // BGSG4:1004
  if (asmb) {
    // column v0 + c of the weighted term matrix, in the tiled pools
    const int ks = asm_ksteps(nbt);
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int u = unit_of<NT, EPT, NVW>(tid, r);
      if (u < nunit) {
        const int e = u / NVW, c = u - e * NVW;
        if (v0 + c < nbt) {
          const double val = (double)(w * yown[r]);
          if (p.A64) p.A64[td.off_asm + asm_index2(e, v0 + c, ks)] = val;
          if (p.A32) p.A32[td.off_asm + asm_index4(e, v0 + c, ks)] = (float)val;
        }
      }
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    const int u = unit_of<NT, EPT, NVW>(tid, r);
    if (u < nunit) {
      const int e = u / NVW, c = u - e * NVW;
      const int rp = p.rowpos[td.off_nb + e];
      if (rp >= 0 && v0 + c < nvec) p.R[(int64_t)rp * nvec + v0 + c] = w * yown[r];
    }
  }
}

// Assembled H|psi>: R_t = H_t X_t for every term, X_t = the term's rows of the
// packed block gathered through the map.  One workgroup of four waves per
// (term, 256-row block); each wave owns up to four 16-row tiles (NM, a
// compile-time count, so the inner loop has no bounds tests) and NTILE
// 16-vector tiles of MFMA accumulators.  The term's X rows stream through LDS
// in chunks of kAsmKc rows; the matrix fragments are read once, coalesced,
// from HBM, one group of four K-steps ahead of the MFMAs that use them.
constexpr int kAsmKc = 128;
template <typename T, int NTILE, int NM>
__device__ __forceinline__ void asm_tiles(const T *__restrict__ At, const int *__restrict__ mp,
                                          const T *__restrict__ X, T *__restrict__ xs, int nbt, int ks,
                                          int mt0, int nvec, int vbase, const int *__restrict__ rpt,
                                          T *__restrict__ R)
{
  constexpr int NV = 16 * NTILE, LD = NV + 1, MA = NM > 0 ? NM : 1;
  constexpr int spc = kAsmKc / 4;          // K-steps per staged chunk of X rows
  constexpr int RPT = kAsmKc * NV / 256;   // X values staged per thread
  constexpr int KST = 256 / NV;            // row stride between them
  const int lane = threadIdx.x & 63;
  using acc_t = T __attribute__((ext_vector_type(4)));
  acc_t acc[MA][NTILE];
#pragma unroll
  for (int m = 0; m < MA; ++m)
#pragma unroll
    for (int n = 0; n < NTILE; ++n) acc[m][n] = acc_t{T(0), T(0), T(0), T(0)};
  // X rows [kc0, kc0 + kAsmKc) into LDS.  Every read goes to a clamped valid
  // address with no branch, so all RPT reads of a thread are in flight
  // together; rows past nb_t and pruned rows are then staged as zeros (they
  // meet zero columns of the assembled matrix, and a zero keeps a non-finite
  // value elsewhere in the block out of this term), and vectors past nvec
  // are not stored.
  const int sv = threadIdx.x % NV, sk = threadIdx.x / NV;
  const int vcol = vbase + sv < nvec ? vbase + sv : nvec - 1;
  auto stage = [&](int kc0) {
    __syncthreads();  // the previous chunk is consumed
    int pk[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int kg = kc0 + sk + KST * j;
      const int mk = mp[kg < nbt ? kg : nbt - 1];
      pk[j] = kg < nbt ? mk : -1;
    }
    T val[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const T x = X[(int64_t)(pk[j] >= 0 ? pk[j] : 0) * nvec + vcol];
      val[j] = pk[j] >= 0 ? x : T(0);
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) xs[(sk + KST * j) * LD + sv] = val[j];
    __syncthreads();
  };
  // the 4 x NM matrix fragments of K-steps s..s+3 (ks is a multiple of 4):
  // FP64, two 16-byte loads of K-step pairs (asm_index2) per tile; FP32, one
  // 16-byte load of the four K-steps (asm_index4; 13.2 against 14.5 ms per
  // FP32 apply at D=9 L=8, 32 vectors, in a same-box A/B with two 8-byte
  // loads).  The pool is read once per call: non-temporal loads keep it from
  // evicting the term's X rows from L2, which the other work items of the
  // term read again (9% faster FP64 apply in a same-box A/B).
  auto load = [&](T(&af)[4][MA], int s) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if constexpr (sizeof(T) == 8) {
        using T2 = T __attribute__((ext_vector_type(2)));
        const T2 *A2 = reinterpret_cast<const T2 *>(At);
        const int64_t b = ((int64_t)(mt0 + m) * (ks >> 1) + (s >> 1)) * 64 + lane;
        const T2 x0 = __builtin_nontemporal_load(A2 + b);
        const T2 x1 = __builtin_nontemporal_load(A2 + b + 64);
        af[0][m] = x0.x;
        af[1][m] = x0.y;
        af[2][m] = x1.x;
        af[3][m] = x1.y;
      } else {
        using T4 = T __attribute__((ext_vector_type(4)));
        const T4 *A4 = reinterpret_cast<const T4 *>(At);
        const int64_t b = ((int64_t)(mt0 + m) * (ks >> 2) + (s >> 2)) * 64 + lane;
        const T4 x = __builtin_nontemporal_load(A4 + b);
        af[0][m] = x.x;
        af[1][m] = x.y;
        af[2][m] = x.z;
        af[3][m] = x.w;
      }
    }
  };
  auto compute = [&](const T(&af)[4][MA], int s) {
    const int kl0 = 4 * (s % spc) + (lane >> 4);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T bf[NTILE];
#pragma unroll
      for (int n = 0; n < NTILE; ++n) bf[n] = xs[(kl0 + 4 * q) * LD + 16 * n + (lane & 15)];
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int n = 0; n < NTILE; ++n) {
          if constexpr (sizeof(T) == 8)
            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q][m], bf[n], acc[m][n], 0, 0, 0);
          else
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[q][m], bf[n], acc[m][n], 0, 0, 0);
        }
    }
  };
  // two groups in flight: the fragments of the next group are requested
  // before the MFMAs of the current one (scheduling barriers keep that order);
  // past the end the last group is requested again (cache hits) instead of a
  // test, so the loads never sit behind a branch.  Chunks of X start at even
  // groups (spc is a multiple of 8).
  static_assert(spc % 8 == 0, "X chunks must start at even groups");
  T a0[4][MA], a1[4][MA];
  load(a0, 0);
  const int ng = ks >> 2;
  int s = 0;
  for (int g = 0; g + 1 < ng; g += 2, s += 8) {
    if (s % spc == 0) stage(4 * s);
    load(a1, s + 4);
    __builtin_amdgcn_sched_barrier(0);
    compute(a0, s);
    __builtin_amdgcn_sched_barrier(0);
    load(a0, s + 8 < ks ? s + 8 : ks - 4);
    __builtin_amdgcn_sched_barrier(0);
    compute(a1, s + 4);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (ng & 1) {
    if (s % spc == 0) stage(4 * s);
    compute(a0, s);
  }
  // D layout: f64 row (lane / 16) + 4 r, f32 row 4 (lane / 16) + r; column lane % 16
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int n = 0; n < NTILE; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * (mt0 + m) + (sizeof(T) == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        const int v = vbase + 16 * n + (lane & 15);
        const int rp = row < nbt ? rpt[row] : -1;  // row of R in scatter order
        // R is read once, by the scatter: streamed past L2 like the pool (FP64)
        if (rp >= 0 && v < nvec) {
          if constexpr (sizeof(T) == 8) __builtin_nontemporal_store(acc[m][n][r], R + (int64_t)rp * nvec + v);
          else R[(int64_t)rp * nvec + v] = acc[m][n][r];
        }
      }
}

template <typename T, int NTILE>
// Three waves per SIMD (FP64: 168 VGPRs with a 2-dword spill; measured 3%
// faster than two waves with no spill, and FP32 3% faster than four waves).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_apply_asm(const TermDev *__restrict__ terms,
                                                   const int2 *__restrict__ items,
                                                   const T *__restrict__ A, const int *__restrict__ map,
                                                   const int *__restrict__ rowpos,
                                                   const T *__restrict__ X, int nvec, int vbase,
                                                   T *__restrict__ R)
{
  __shared__ T xs[kAsmKc * (16 * NTILE + 1)];
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int nbt = td.nbt, mtn = (nbt + 15) >> 4, ks = asm_ksteps(nbt);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int mt0 = it.y + 4 * wave;
  const int nm = mtn - mt0 < 0 ? 0 : (mtn - mt0 > 4 ? 4 : mtn - mt0);
  const T *At = A + td.off_asm;
  const int *mp = map + td.off_nb;
  const int *rpt = rowpos + td.off_nb;
  // every wave runs the same chunk loop (same barriers), with its own tile count
  switch (nm) {
    case 4: asm_tiles<T, NTILE, 4>(At, mp, X, xs, nbt, ks, mt0, nvec, vbase, rpt, R); break;
    case 3: asm_tiles<T, NTILE, 3>(At, mp, X, xs, nbt, ks, mt0, nvec, vbase, rpt, R); break;
    case 2: asm_tiles<T, NTILE, 2>(At, mp, X, xs, nbt, ks, mt0, nvec, vbase, rpt, R); break;
    case 1: asm_tiles<T, NTILE, 1>(At, mp, X, xs, nbt, ks, mt0, nvec, vbase, rpt, R); break;
    default: asm_tiles<T, NTILE, 0>(At, mp, X, xs, nbt, ks, mt0, nvec, vbase, rpt, R); break;
  }
}

// Deterministic weighted Smolyak scatter: hpsi[pk, v] = sum over the term
// rows that map to pk (fixed order) of R.  Output column-major.
//
// The term kernels write the rows of packed function pk, in increasing term
// order, to one contiguous stretch of R.  The stretches are very uneven (at
// D=9 L=8 the median function has 1 row, the ground function 24,310: it sits
// in every term), so each is cut into segments of at most kReduceSeg rows and
// the sum runs in two passes with a fixed order: segment sums (four
// interleaved partial sums), then the segments of each function in order.
constexpr int kReduceTile = 16;   // segments / functions per workgroup
constexpr int kReduceSeg = 128;   // rows per segment at most

// pass 1: P[s][v] = sum of rows seg[s] .. seg[s+1]-1 of R (coalesced rows in,
// coalesced rows out)
template <typename T>
__global__ __launch_bounds__(256) void k_reduce_seg(const T *__restrict__ R, const int *__restrict__ seg,
                                                    int64_t nseg, int nvec, double *__restrict__ P)
{
  const int64_t s0 = (int64_t)blockIdx.x * kReduceTile;
  const int ns = nseg - s0 < kReduceTile ? (int)(nseg - s0) : kReduceTile;
  for (int g = threadIdx.x; g < ns * nvec; g += blockDim.x) {
    const int sl = g / nvec, v = g - sl * nvec;
    const int r0 = seg[s0 + sl], n = seg[s0 + sl + 1] - r0;
    const T *Rv = R + (int64_t)r0 * nvec + v;
    // FP64 sums also for FP32 term rows: the signed Smolyak weights cancel
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = 0;
#pragma unroll 4
    for (; i + 4 <= n; i += 4) {
      a0 += (double)Rv[(int64_t)i * nvec];
      a1 += (double)Rv[(int64_t)(i + 1) * nvec];
      a2 += (double)Rv[(int64_t)(i + 2) * nvec];
      a3 += (double)Rv[(int64_t)(i + 3) * nvec];
    }
    for (; i < n; ++i) a0 += (double)Rv[(int64_t)i * nvec];
    P[(s0 + sl) * nvec + v] = (a0 + a1) + (a2 + a3);
  }
}

// pass 2: out[v][pk] = sum over the segments fseg[pk] .. fseg[pk+1]-1 of P,
// in order, through an LDS tile (the packed index fastest on the way out)
__global__ __launch_bounds__(256) void k_reduce_fn(const double *__restrict__ P, const int *__restrict__ fseg,
                                                   int64_t nb_packed, int nvec, double *__restrict__ out)
{
  extern __shared__ double tile[];  // kReduceTile x (nvec + 1)
  const int64_t pk0 = (int64_t)blockIdx.x * kReduceTile;
  const int np = nb_packed - pk0 < kReduceTile ? (int)(nb_packed - pk0) : kReduceTile;
  const int ld = nvec + 1;
  for (int g = threadIdx.x; g < np * nvec; g += blockDim.x) {
    const int pl = g / nvec, v = g - pl * nvec;
    const int a = fseg[pk0 + pl], b = fseg[pk0 + pl + 1];
    double s = 0.0;
#pragma unroll 8
    for (int k = a; k < b; ++k) s += P[(int64_t)k * nvec + v];
    tile[pl * ld + v] = s;
  }
  __syncthreads();
  for (int g = threadIdx.x; g < np * nvec; g += blockDim.x) {
    const int v = g / np, pl = g - v * np;
    out[(int64_t)v * nb_packed + pk0 + pl] = tile[pl * ld + v];
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

// V_eff of one term, one workgroup per term.  For a mode d with one basis
// function the term's V part factors as W_d diag(.) B_d = sum_z W_d[0,z]
// B_d[z,0] (.)_z, so
//   V_eff(r) = sum_z prod_{d one-function} W_d[0,z_d] B_d[z_d,0] V(r, z)
// over the grid points z of those modes, r running over the grid of the
// other modes (first mode fastest in both).  The same discrete operator, with
// the one-function sums taken once per potential instead of per H|psi>.
__global__ __launch_bounds__(256) void k_fold_v(int D, const TermDev *__restrict__ terms,
                                                const SlotDev *__restrict__ slots,
                                                const double *__restrict__ bw,
                                                const double *__restrict__ V, double *Veff)
{
  const TermDev &td = terms[blockIdx.x];
  int nq[kMaxDev], zm[kMaxDev];
  int64_t stride[kMaxDev];
  int64_t st = 1;
  for (int d = 0; d < D; ++d) {
    nq[d] = slots[td.slot[d]].nq;
    stride[d] = st;
    st *= nq[d];
  }
  // one-function modes: the complement of vmode[] (both ascending)
  int nz = 0, nzc = 1;
  for (int d = 0, j = 0; d < D; ++d) {
    if (j < td.Dv && td.vmode[j] == d) {
      ++j;
      continue;
    }
    zm[nz++] = d;
    nzc *= nq[d];
  }
  int64_t qv = 1;
  for (int j = 0; j < td.Dv; ++j) qv *= nq[td.vmode[j]];
  for (int64_t r = threadIdx.x; r < qv; r += blockDim.x) {
    int64_t rem = r, base = 0;
    for (int j = 0; j < td.Dv; ++j) {
      const int d = td.vmode[j];
      const int64_t q = rem / nq[d];
      base += (rem - q * nq[d]) * stride[d];
      rem = q;
    }
    double acc = 0.0;
    for (int z = 0; z < nzc; ++z) {
      int rz = z;
      int64_t idx = base;
      double c = 1.0;
      for (int k = 0; k < nz; ++k) {
        const int d = zm[k];
        const int iz = rz % nq[d];
        rz /= nq[d];
        idx += iz * stride[d];
        const SlotDev &s = slots[td.slot[d]];
        c *= bw[s.offW + iz] * bw[s.offB + iz];  // W[0, iz] B[iz, 0] (nb = 1)
      }
      acc = fma(c, V[td.off_nq + idx], acc);
    }
    Veff[td.off_vq + r] = acc;
  }
}

__global__ __launch_bounds__(256) void k_to_float(const double *__restrict__ a, float *__restrict__ b,
                                                  int64_t n)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    b[g] = (float)a[g];
}

// y += a x over n values
__global__ __launch_bounds__(256) void k_axpy(int64_t n, double a, const double *__restrict__ x,
                                              double *__restrict__ y)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    y[g] += a * x[g];
}

// dst[i] = src[i] + add, entries below zero (pruned) kept at -1
__global__ __launch_bounds__(256) void k_offset_ints(const int *__restrict__ src, int64_t n, int64_t add,
                                                     int *__restrict__ dst)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int v = src[g];
    dst[g] = v >= 0 ? (int)(v + add) : -1;
  }
}

// Zero fill with 64-bit indices (the assembled pools exceed 2^34 words; a
// fill kernel is also safe inside captured graphs, unlike memset nodes).
template <typename T>
__global__ __launch_bounds__(256) void k_zero(T *__restrict__ a, int64_t n)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    a[g] = T(0);
}

// ------------------------------------------- type 10 with G, J, s per point --
// The reference's grid form of one term (OPSG4:1566-1632) over the
// term-concatenated grids, k vectors at a time, column v of a grid array at
// v * sum_nq:  P_j = d_j (s psi_G);  h_i = J sum_j G(:, j, i) P_j;
// out = -1/2 (sum_i d_i h_i) / (J s).  A derivative along mode j is the slot's
// dnRGG%d1 (nq x nq, column-major) applied along that mode of the term grid
// (first mode fastest).  One thread per grid point; items = (term, first point).

// mode-j derivative of column `in` (optionally scaled by sc) at point p of a term
template <typename T>
__device__ __forceinline__ double grid_d1(const double *__restrict__ Dm, int nq, int qj, int64_t st,
                                          const T *__restrict__ in, const double *__restrict__ sc,
                                          int64_t p0)
{
  double acc = 0.0;
  for (int k = 0; k < nq; ++k) {
    const int64_t g = p0 + k * st;
    acc += Dm[qj + nq * k] * (sc ? sc[g] * (double)in[g] : (double)in[g]);
  }
  return acc;
}

// the same with the transposed 1D matrix (row qj of Dm^T)
template <typename T>
__device__ __forceinline__ double grid_d1t(const double *__restrict__ Dm, int nq, int qj, int64_t st,
                                           const T *__restrict__ in, const double *__restrict__ sc,
                                           int64_t p0)
{
  double acc = 0.0;
  for (int k = 0; k < nq; ++k) {
    const int64_t g = p0 + k * st;
    acc += Dm[k + nq * qj] * (sc ? sc[g] * (double)in[g] : (double)in[g]);
  }
  return acc;
}

// The grid-form kernels keep psi_G, the derivative columns and the output in
// T (FP64, or FP32 for the FP32 H|psi>) and do their arithmetic in FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_gk_deriv(int D, const TermDev *__restrict__ terms,
                                                  const SlotDev *__restrict__ slots,
                                                  const double *__restrict__ D1, const int2 *__restrict__ items,
                                                  const double *__restrict__ s, const T *__restrict__ G,
                                                  int64_t sum_nq, int k, T *__restrict__ P, int trans)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  int64_t st = 1;
  for (int j = 0; j < D; ++j) {
    const SlotDev &sl = slots[td.slot[j]];
    const int nq = sl.nq;
    const int qj = (int)((p / st) % nq);
    const int64_t p0 = td.off_nq + p - qj * st;
    for (int v = 0; v < k; ++v)
      P[((int64_t)v * D + j) * sum_nq + td.off_nq + p] =
          T(trans ? grid_d1t(D1 + sl.offD1, nq, qj, st, G + v * sum_nq, s, p0)
                  : grid_d1(D1 + sl.offD1, nq, qj, st, G + v * sum_nq, s, p0));
    st *= nq;
  }
}

// h_i = J sum_j G[(i D + j) sum_nq + g] P_j, in place (all P_j of the point
// are read before any h_i is written); trans takes G[(j D + i) sum_nq + g]
template <typename T>
__global__ __launch_bounds__(256) void k_gk_combine(int D, int64_t sum_nq, int k, const double *__restrict__ GG,
                                                    const double *__restrict__ J, T *__restrict__ P, int trans)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < sum_nq;
       g += (int64_t)gridDim.x * blockDim.x) {
    const double jg = J[g];
    for (int v = 0; v < k; ++v) {
      double pj[kMaxDev];
      for (int j = 0; j < D; ++j) pj[j] = (double)P[((int64_t)v * D + j) * sum_nq + g];
      for (int i = 0; i < D; ++i) {
        double h = 0.0;
        for (int j = 0; j < D; ++j) h += GG[((int64_t)(trans ? j * D + i : i * D + j)) * sum_nq + g] * pj[j];
        P[((int64_t)v * D + i) * sum_nq + g] = T(jg * h);
      }
    }
  }
}

// out = -1/2 (sum_i d_i h_i) / (J s) + V psi_G (V on the unscaled psi_G,
// OPSG4:1566-1576).  With wq (the quadrature weight of every point) the
// transposed form of the weighted operator: out = -1/2 (s / wq) sum_i d_i^T
// h_i + V psi_G.
template <typename T>
__global__ __launch_bounds__(256) void k_gk_out(int D, const TermDev *__restrict__ terms,
                                                const SlotDev *__restrict__ slots, const double *__restrict__ D1,
                                                const int2 *__restrict__ items, const double *__restrict__ J,
                                                const double *__restrict__ s, const double *__restrict__ V,
                                                const T *__restrict__ G, const T *__restrict__ P,
                                                int64_t sum_nq, int k, T *__restrict__ F,
                                                const double *__restrict__ wq)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  const int64_t g = td.off_nq + p;
  const double scale = wq ? -0.5 * s[g] / wq[g] : -0.5 / (J[g] * s[g]);
  for (int v = 0; v < k; ++v) {
    double o = 0.0;
    int64_t st = 1;
    for (int i = 0; i < D; ++i) {
      const SlotDev &sl = slots[td.slot[i]];
      const int nq = sl.nq;
      const int qi = (int)((p / st) % nq);
      const T *Pi = P + ((int64_t)v * D + i) * sum_nq;
      o += wq ? grid_d1t(D1 + sl.offD1, nq, qi, st, Pi, (const double *)nullptr, td.off_nq + p - qi * st)
              : grid_d1(D1 + sl.offD1, nq, qi, st, Pi, (const double *)nullptr, td.off_nq + p - qi * st);
      st *= nq;
    }
    F[v * sum_nq + g] = T(scale * o + V[g] * (double)G[v * sum_nq + g]);
  }
}

// a[i] = (a[i] + b[i]) / 2: the symmetric part from the rows of H and H^T
template <typename T>
__global__ __launch_bounds__(256) void k_half_sum(T *__restrict__ a, const T *__restrict__ b, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = T(0.5) * (a[i] + b[i]);
}

// ------------------------------------------- type 1 with F2, F1 per point --
// The reference's type-1 grid combine of one term (OPSG4:1490-1546) with its
// operator grids per point (get_OpGrid_type1_OF_ONEDP_FOR_SG4,
// OPSG4:2820-3084):  out = V psi_G + sum_i F1(:, i) d_i psi_G
//   + sum_i F2(:, i, i) d2_i psi_G + sum_{i != j} F2(:, i, j) d_i (d_j psi_G),
// d = dnRGG%d1, d2 = dnRGG%d2 along one mode of the term grid; P holds
// d_j psi_G (k_gk_deriv without a scale).  F2 is (sum_nq, D, D) Fortran,
// F1 (sum_nq, D) or NULL.  One thread per grid point.
template <typename T>
__global__ __launch_bounds__(256) void k_t1_out(int D, const TermDev *__restrict__ terms,
                                                const SlotDev *__restrict__ slots, const double *__restrict__ D1,
                                                const double *__restrict__ D2, const int2 *__restrict__ items,
                                                const double *__restrict__ F2, const double *__restrict__ F1,
                                                const double *__restrict__ V, const T *__restrict__ G,
                                                const T *__restrict__ P, int64_t sum_nq, int k,
                                                T *__restrict__ F)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  const int64_t g = td.off_nq + p;
  int qd[kMaxDev];
  int64_t st[kMaxDev];
  int64_t s = 1;
  for (int d = 0; d < D; ++d) {
    const int nq = slots[td.slot[d]].nq;
    qd[d] = (int)((p / s) % nq);
    st[d] = s;
    s *= nq;
  }
  for (int v = 0; v < k; ++v) {
    const T *Gv = G + (int64_t)v * sum_nq;
    double o = V[g] * (double)Gv[g];
    for (int i = 0; i < D; ++i) {
      const SlotDev &sl = slots[td.slot[i]];
      const int64_t p0 = td.off_nq + p - qd[i] * st[i];
      if (F1) o += F1[g + sum_nq * i] * (double)P[((int64_t)v * D + i) * sum_nq + g];
      const double fii = F2[g + sum_nq * (i + (int64_t)D * i)];
      if (fii != 0.0) o += fii * grid_d1(D2 + sl.offD1, sl.nq, qd[i], st[i], Gv, nullptr, p0);
      for (int j = 0; j < D; ++j) {
        if (j == i) continue;
        const double fij = F2[g + sum_nq * (i + (int64_t)D * j)];
        if (fij != 0.0)
          o += fij * grid_d1(D1 + sl.offD1, sl.nq, qd[i], st[i], P + ((int64_t)v * D + j) * sum_nq, nullptr, p0);
      }
    }
    F[(int64_t)v * sum_nq + g] = T(o);
  }
}

// The transposed type-1 grid form for the hermitian plan: with y = w psi_G
// (w the quadrature weight of the point) the weighted transpose is
//   out = V psi_G + (1/w) [sum_i d_i^T (F1_i y) + d2_i^T (F2_ii y)
//                          + sum_j d_j^T U_j],  U_j = sum_{i != j} d_i^T (F2_ij y),
// with F1w = w F1 and F2w = w F2 tabulated at set-up.  k_t1t_u writes U_j
// into P, k_t1t_out the output.  One thread per grid point.
template <typename T>
__global__ __launch_bounds__(256) void k_t1t_u(int D, const TermDev *__restrict__ terms,
                                               const SlotDev *__restrict__ slots, const double *__restrict__ D1,
                                               const int2 *__restrict__ items, const double *__restrict__ F2w,
                                               const T *__restrict__ G, int64_t sum_nq, int k, T *__restrict__ P)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  const int64_t g = td.off_nq + p;
  int qd[kMaxDev];
  int64_t st[kMaxDev];
  int64_t s = 1;
  for (int d = 0; d < D; ++d) {
    const int nq = slots[td.slot[d]].nq;
    qd[d] = (int)((p / s) % nq);
    st[d] = s;
    s *= nq;
  }
  for (int v = 0; v < k; ++v) {
    const T *Gv = G + (int64_t)v * sum_nq;
    for (int j = 0; j < D; ++j) {
      double u = 0.0;
      for (int i = 0; i < D; ++i) {
        if (i == j) continue;
        const SlotDev &sl = slots[td.slot[i]];
        u += grid_d1t(D1 + sl.offD1, sl.nq, qd[i], st[i], Gv, F2w + sum_nq * (i + (int64_t)D * j),
                      td.off_nq + p - qd[i] * st[i]);
      }
      P[((int64_t)v * D + j) * sum_nq + g] = T(u);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_t1t_out(int D, const TermDev *__restrict__ terms,
                                                 const SlotDev *__restrict__ slots, const double *__restrict__ D1,
                                                 const double *__restrict__ D2, const int2 *__restrict__ items,
                                                 const double *__restrict__ F2w, const double *__restrict__ F1w,
                                                 const double *__restrict__ wq, const double *__restrict__ V,
                                                 const T *__restrict__ G, const T *__restrict__ P, int64_t sum_nq,
                                                 int k, T *__restrict__ F)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  const int64_t g = td.off_nq + p;
  int qd[kMaxDev];
  int64_t st[kMaxDev];
  int64_t s = 1;
  for (int d = 0; d < D; ++d) {
    const int nq = slots[td.slot[d]].nq;
    qd[d] = (int)((p / s) % nq);
    st[d] = s;
    s *= nq;
  }
  for (int v = 0; v < k; ++v) {
    const T *Gv = G + (int64_t)v * sum_nq;
    double o = 0.0;
    for (int i = 0; i < D; ++i) {
      const SlotDev &sl = slots[td.slot[i]];
      const int64_t p0 = td.off_nq + p - qd[i] * st[i];
      if (F1w) o += grid_d1t(D1 + sl.offD1, sl.nq, qd[i], st[i], Gv, F1w + sum_nq * i, p0);
      o += grid_d1t(D2 + sl.offD1, sl.nq, qd[i], st[i], Gv, F2w + sum_nq * (i + (int64_t)D * i), p0);
      o += grid_d1t(D1 + sl.offD1, sl.nq, qd[i], st[i], P + ((int64_t)v * D + i) * sum_nq,
                    (const double *)nullptr, p0);
    }
    F[(int64_t)v * sum_nq + g] = T(V[g] * (double)Gv[g] + o / wq[g]);
  }
}

// ----------------------------------------------------- grid derivative --
// DerivOp_TO_RDP_OF_SmolaykRep (BGSG4:2690-2795) on every term grid, ncol
// columns of the concatenated term grids: i == j >= 0: dnRGG%d2 along mode
// i; one of them < 0: dnRGG%d1 along the other; i != j both >= 0: d1 along
// mode j, then along mode i; both < 0: a copy.  One thread per grid point.
__global__ __launch_bounds__(256) void k_grid_deriv(const TermDev *__restrict__ terms,
                                                    const SlotDev *__restrict__ slots,
                                                    const double *__restrict__ D1, const double *__restrict__ D2,
                                                    const int2 *__restrict__ items, int di, int dj,
                                                    const double *__restrict__ in, int64_t sum_nq, int ncol,
                                                    double *__restrict__ out)
{
  const int2 it = items[blockIdx.x];
  const TermDev &td = terms[it.x];
  const int p = it.y + threadIdx.x;
  if (p >= td.nqt) return;
  const int64_t g = td.off_nq + p;
  // stride and coordinate of modes di, dj at point p (first mode fastest)
  int64_t st = 1, sti = 1, stj = 1;
  int qi = 0, qj = 0, nqi = 1, nqj = 1;
  int64_t oi = 0, oj = 0;
  for (int d = 0; d <= (di > dj ? di : dj); ++d) {
    const SlotDev &sl = slots[td.slot[d]];
    const int q = (int)((p / st) % sl.nq);
    if (d == di) { sti = st; qi = q; nqi = sl.nq; oi = sl.offD1; }
    if (d == dj) { stj = st; qj = q; nqj = sl.nq; oj = sl.offD1; }
    st *= sl.nq;
  }
  for (int v = 0; v < ncol; ++v) {
    const double *x = in + (int64_t)v * sum_nq;
    double r;
    if (di < 0 && dj < 0) {
      r = x[g];
    } else if (di == dj || di < 0 || dj < 0) {
      const bool two = di == dj;
      const int m = di >= 0 ? di : dj;
      const int64_t s = m == di ? sti : stj;
      const int q = m == di ? qi : qj, n = m == di ? nqi : nqj;
      const double *M = (two ? D2 : D1) + (m == di ? oi : oj);
      r = grid_d1(M, n, q, s, x, nullptr, g - q * s);
    } else {
      r = 0.0;
      const int64_t p00 = g - qi * sti - qj * stj;
      for (int a = 0; a < nqi; ++a) {
        double inner = 0.0;
        for (int b = 0; b < nqj; ++b) inner += D1[oj + qj + nqj * b] * x[p00 + a * sti + b * stj];
        r += D1[oi + qi + nqi * a] * inner;
      }
    }
    out[(int64_t)v * sum_nq + g] = r;
  }
}

// Reference direct-product grid order (last mode fastest) <-> term-grid order
// (first mode fastest) for ncol columns: to_user: out[perm[i]] = in[i], else
// out[i] = in[perm[i]].
__global__ __launch_bounds__(256) void k_perm_cols(int64_t n, int ncol, const int *__restrict__ perm,
                                                   const double *__restrict__ in, double *__restrict__ out,
                                                   int to_user)
{
  const int64_t tot = n * ncol;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < tot;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k % n, c = k / n;
    if (to_user) out[perm[i] + c * n] = in[k];
    else out[k] = in[perm[i] + c * n];
  }
}

// explicit instantiations are produced by the launch table in evr_plan.hip
}  // namespace evr
