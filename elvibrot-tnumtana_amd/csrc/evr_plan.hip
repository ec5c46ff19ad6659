// Plans and the C ABI of the H|psi> engine (see include/evr_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/evr_hip.h"
#include "evr_internal.hpp"
#include "evr_kernels.hpp"
#include "evr_setup.hpp"
#include "evr_term.hip"
#include "evr_stream.hip"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(EVR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

template <typename X>
struct DevBuf {
  X *ptr = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count)
  {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)&ptr, sizeof(X) * count);
  }
  hipError_t upload(const X *h, size_t count, hipStream_t s)
  {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    e = hipMemcpyAsync(ptr, h, sizeof(X) * count, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  void release()
  {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
};

// work class of the term kernel: threads per workgroup x elements per thread
struct KClass {
  int nt, ept, nvw;
  int max_nbt() const { return nt * ept / nvw; }
};
// (nt, ept, nvw): ept == nvw > 1 is element-major (each thread the nvw vectors
// of one basis function).  The term kernels are bound by latency and barrier
// waits, so the classes below 1024 threads are built for four waves per SIMD
// (two or more workgroups per CU hide each other's barriers): nb_t 513-1024
// and 1025-2048 run as 512 x 4 rather than 1024 x 2 (38.4 against 42.7 ms
// and 40.7 against 45.5 ms at D=9 L=8, 32 vectors), 129-256 as 512 x 2 at 4
// vectors rather than 1024 x 1 (6.9 against 8.0 ms; 256 x 4 8.0 ms), 257-512
// stays 512 x 2 (16.9 ms; 256 x 4 measured 18.4 ms).
constexpr KClass kClasses[] = {{256, 1, 4},  {512, 1, 4},  {512, 2, 4}, {512, 2, 2},
                               {512, 4, 2}, {512, 4, 1}, {1024, 4, 1}, {1024, 8, 1}};
constexpr int kNumClasses = sizeof(kClasses) / sizeof(kClasses[0]);
// dynamic LDS per workgroup: 160 KiB less a reserve for the kernels' static
// arrays (about 2 KiB for the term kernels)
constexpr size_t kLdsLimit = 160 * 1024 - 4096;
// a one-term plan of at least this many functions runs its term streamed
constexpr int kStreamOneMin = 256;

}  // namespace

int evr::set_error(int code, const std::string &msg) { return fail(code, msg); }

struct evr_desc_owner {
  evr_basis_desc desc;
  evr::BasisTables tab;
};

struct evr_plan {
  int D = 0, type = 0, max_nvec = 0;  // D: the term kernel's mode positions
  int device = 0;                     // HIP device current at creation
  int ncu = 256;                      // compute units of that device
  int D_full = 0;                     // modes of the basis (> D when compacted)
  int64_t keoc_stride = 0;            // doubles per KEO coefficient set
  bool dp_order = false;
  int64_t nterms = 0, nb_packed = 0, sum_nb = 0, sum_nq = 0, max_nb = 0, max_nq = 0;
  int64_t sum_vq = 0;  // reduced grid points (V_eff) of all terms
  hipStream_t stream = nullptr;
  bool hasV = false, hasKeo = false, hasA = false, hasF1 = false, diagT2 = false;
  // host copies
  std::vector<evr::TermDev> terms;
  std::vector<evr::SlotDev> slots;
  std::vector<int> dp_perm;  // DP: reference grid index of each term-grid point
  DevBuf<int> d_dp_perm;     // the same on the device (multi-vector transforms)
  // device tables
  DevBuf<evr::TermDev> d_terms;
  DevBuf<evr::SlotDev> d_slots;
  std::vector<evr::StepDev> steps;
  DevBuf<evr::StepDev> d_steps;
  DevBuf<double> d_bw64, d_keo64, d_x, d_bwdiag, d_D1, d_D2;
  DevBuf<float> d_bw32, d_keo32;
  // type 10 with G, J, s per grid point (evr_plan_set_keo_grid): the tables,
  // the (term, first point) items of the grid kernels, and work grids for
  // gvec vectors at a time (psi_G, the D derivative columns, the KEO output)
  bool hermitian = false, keo_grid = false;
  int grid_type = 0;  // keo_grid: 10 (G, J, s per point) or 1 (F2, F1 per point)
  DevBuf<double> d_F2g, d_F1g;  // type 1 operator grids (sum_nq x D x D, sum_nq x D)
  int gvec = 0;
  int64_t n_gitems = 0;
  DevBuf<double> d_GG, d_Jg, d_Sg, d_Gg, d_Pg, d_Fg;
  DevBuf<int2> d_gitems;
  // hermitian grid form: quadrature weight per point, w / (J s), and the term
  // rows of the transposed sequence
  DevBuf<double> d_Wq, d_As, d_R2, d_F2w, d_F1w;
  // streamed terms (a basis beyond one workgroup's LDS, evr_stream.hip): their
  // ids, the largest term grid among them, the vectors per pass and the work
  // block of (7 + D) tensors of big_nq x big_k words
  std::vector<int64_t> big;
  // a one-term plan (direct-product basis) whose term fits a workgroup still
  // runs it streamed for H|psi> and the transforms: the resident kernel would
  // give the whole device one workgroup per vector group
  std::vector<int64_t> one;
  int64_t big_nq = 0;
  int big_k = 0;
  DevBuf<double> d_big;
  std::vector<double> hC, hF1;  // host copies of the KEO coefficients (kernel convention)
  DevBuf<int> d_map, d_rowpos;
  // reduction segments: rows seg[s] .. seg[s+1]-1 of R; the segments of packed
  // function pk are fseg[pk] .. fseg[pk+1]-1
  DevBuf<int> d_seg, d_fseg;
  int64_t nseg = 0;
  int64_t nrows = 0;   // rows of R: term rows that map to a packed function
  DevBuf<double> d_P;  // segment sums, nseg x max_nvec
  DevBuf<double> d_V;
  DevBuf<double> d_Veff;  // V summed over the one-function modes of each term
  DevBuf<float> d_Veff32;
  DevBuf<double> d_coefC, d_coefA, d_coefF1, d_keoc;
  DevBuf<int> d_order[kNumClasses];
  int64_t nunits[kNumClasses] = {};   // work list lengths (terms padded to a multiple of 8)
  int64_t nterms_class[kNumClasses] = {};
  std::vector<int> tclass;  // kernel class of every term (static rule or autotuned)
  bool keo_layout = false;  // the term layouts hold the KEO factors in LDS
  size_t lds_bytes64[kNumClasses] = {}, lds_bytes32[kNumClasses] = {};
  // work buffers
  DevBuf<double> d_X, d_R, d_psi, d_hpsi, d_G;
  DevBuf<double> d_gw;  // w_t * quadrature weight of every term-grid point (lazy)
  // assembled per-term operators (evr_plan_assemble)
  int64_t asm_elems = 0;           // sum over terms of the tiled nb_t x nb_t sizes
  DevBuf<double> d_A64;
  DevBuf<float> d_A32;
  DevBuf<int2> d_asm_items;        // (term, first 16-row tile) per apply workgroup
  int64_t n_asm_items = 0;
  double asm_flops_per_vec = 0.0;  // 2 sum nb_t^2
  double asm_bytes64 = 0.0;        // matrix bytes read per pass (FP64)
  bool V32_valid = false;
  // per-launch accounting: algorithmic flops per vector of each kernel class
  double flops_class[kNumClasses] = {};
  double flops_reduce = 0;
  bool profiling = false;
  hipEvent_t ev[2 * (kNumClasses + 1)] = {};
  double ms_acc[kNumClasses + 1] = {};
  double flops_acc[kNumClasses + 1] = {};
  int64_t launches[kNumClasses + 1] = {};

  evr_plan() = default;
  evr_plan(const evr_plan &) = delete;
  evr_plan &operator=(const evr_plan &) = delete;
  ~evr_plan()  // also on a failed evr_plan_create
  {
    if (stream) (void)hipStreamSynchronize(stream);
    d_terms.release(); d_slots.release(); d_steps.release(); d_bw64.release(); d_keo64.release();
    d_x.release(); d_bwdiag.release(); d_bw32.release(); d_keo32.release();
    d_map.release(); d_seg.release(); d_fseg.release(); d_P.release(); d_rowpos.release(); d_V.release();
    d_Veff.release(); d_Veff32.release(); d_gw.release();
    d_coefC.release(); d_coefA.release(); d_coefF1.release(); d_keoc.release();
    for (int c = 0; c < kNumClasses; ++c) d_order[c].release();
    d_X.release(); d_R.release(); d_psi.release(); d_hpsi.release(); d_G.release();
    d_A64.release(); d_A32.release(); d_asm_items.release(); d_D1.release(); d_D2.release();
    d_GG.release(); d_Jg.release(); d_Sg.release(); d_Gg.release(); d_Pg.release(); d_Fg.release();
    d_F2g.release(); d_F1g.release(); d_gitems.release(); d_dp_perm.release(); d_big.release();
    d_Wq.release(); d_As.release(); d_R2.release(); d_F2w.release(); d_F1w.release();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <typename T, int NT, int EPT, int NVW>
hipError_t launch_one(const evr::TermParams<T> &prm, int64_t units, int nvec, size_t lds,
                      hipStream_t s)
{
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&evr::k_term<T, NT, EPT, NVW>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int64_t nblk = units * ((nvec + NVW - 1) / NVW);
  if (nblk <= 0) return hipSuccess;
  // the dispatch grid is counted in work-items in 32 bits
  if (nblk > (int64_t)0xFFFFFFFFu / NT) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((evr::k_term<T, NT, EPT, NVW>), dim3((unsigned)nblk), dim3(NT), lds, s, prm);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_class(int c, const evr::TermParams<T> &prm, int64_t units, int nvec,
                        size_t lds, hipStream_t s)
{
  switch (c) {
    case 0: return launch_one<T, 256, 1, 4>(prm, units, nvec, lds, s);
    case 1: return launch_one<T, 512, 1, 4>(prm, units, nvec, lds, s);
    case 2: return launch_one<T, 512, 2, 4>(prm, units, nvec, lds, s);
    case 3: return launch_one<T, 512, 2, 2>(prm, units, nvec, lds, s);
    case 4: return launch_one<T, 512, 4, 2>(prm, units, nvec, lds, s);
    case 5: return launch_one<T, 512, 4, 1>(prm, units, nvec, lds, s);
    case 6: return launch_one<T, 1024, 4, 1>(prm, units, nvec, lds, s);
    case 7: return launch_one<T, 1024, 8, 1>(prm, units, nvec, lds, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
evr::TermParams<T> base_params(evr_plan *p, int mode, int nvec)
{
  evr::TermParams<T> prm{};
  prm.D = p->D;
  prm.mode = mode;
  prm.hasV = p->hasV;
  prm.hasKeo = p->hasKeo;
  prm.hasA = p->hasA;
  prm.hasF1 = p->hasF1;
  prm.diagT2 = p->diagT2;
  prm.nvec = nvec;
  prm.terms = p->d_terms.ptr;
  prm.slots = p->d_slots.ptr;
  prm.steps = p->d_steps.ptr;
  prm.map = p->d_map.ptr;
  prm.rowpos = p->d_rowpos.ptr;
  prm.coefC = p->d_coefC.ptr;
  prm.coefA = p->d_coefA.ptr;
  prm.coefF1 = p->d_coefF1.ptr;
  prm.keoc = p->d_keoc.ptr;
  prm.keoc_stride = p->keoc_stride;
  prm.gld = p->sum_nq;
  return prm;
}

// Algorithmic flops of one term for one vector (the single-pass algorithm:
// i-group re-computation is not counted).
double term_flops(const evr_plan *p, const evr::TermDev &td)
{
  const int D = p->D;
  int nb[evr::kMaxDev], nq[evr::kMaxDev];
  for (int d = 0; d < D; ++d) { nb[d] = p->slots[td.slot[d]].nb; nq[d] = p->slots[td.slot[d]].nq; }
  double f = 0.0;
  if (p->hasV) {
    // reduced grid: the modes with more than one function
    int vb[evr::kMaxDev], vq[evr::kMaxDev];
    for (int j = 0; j < td.Dv; ++j) { vb[j] = nb[td.vmode[j]]; vq[j] = nq[td.vmode[j]]; }
    double leaf = 2.0 * td.vnbA * td.vnbB + 2.0 * td.nbt + (double)td.vnbB * (td.Dv - td.vm) * 2.0 +
                  td.vQA;
    for (int d = 0; d < td.vm; ++d) {
      double n1 = 1, n3 = 1, m1 = 1, m3 = 1;
      for (int e = 0; e < d; ++e) { n1 *= vq[e]; m1 *= vb[e]; }
      for (int e = d + 1; e < td.vm; ++e) { n3 *= vb[e]; m3 *= vq[e]; }
      leaf += 2.0 * n1 * vb[d] * vq[d] * n3 + 2.0 * m1 * vq[d] * vb[d] * m3;
    }
    f += leaf * td.vQB;
  }
  if (p->hasKeo) {
    // modes with non-vanishing P, Q, R; the others are scalar folds
    const int na = td.na;
    double e = 2.0;
    for (int ja = 0; ja < na; ++ja) {
      const int j = td.amode[ja];
      e += 2.0 * nb[j] + 2.0 * (na - 1);
      if (p->hasA) {
        e += 4.0 * nb[j] + 6.0 * (na - 1) + 2.0;
        for (int ka = 0; ka < na; ++ka)
          if (ka != ja) e += 2.0 * nb[td.amode[ka]] + 2.0 * (na - 2);
      }
    }
    for (int ia = 0; ia < na; ++ia) {
      const int i = td.amode[ia];
      e += 4.0 * nb[i] + 2.0;
      if (p->hasA) e += 4.0 * nb[i] + 4.0;
      if (p->hasF1) e += 2.0 * nb[i] + 2.0;
    }
    f += e * td.nbt;
  }
  (void)D;
  return f;
}

// Slab split of term td for kernel class c over its reduced grid (the modes
// with more than one function; the largest first-m-modes slab that fits the
// class's LDS budget) and its 2m contraction steps, B->G then G->B, appended
// to steps; words: LDS doubles needed.
bool layout_term(const evr_plan *p, evr::TermDev &td, int c, std::vector<evr::StepDev> &steps, size_t &words)
{
  const int D = p->D;
  int nbd[evr::kMaxDev], nqd[evr::kMaxDev];
  for (int d = 0; d < D; ++d) {
    nbd[d] = p->slots[td.slot[d]].nb;
    nqd[d] = p->slots[td.slot[d]].nq;
  }
  const int64_t nbt = td.nbt;
  const int nvw = kClasses[c].nvw;
  const size_t keo_bufs = kClasses[c].ept <= 4 ? (evr::keo_pencil(kClasses[c].nt, kClasses[c].ept, nvw) ? 6 : 3) : 1;
  const size_t base = td.lds_mats + (size_t)nbt * nvw;
  const size_t keo_words = p->keo_layout ? keo_bufs * nbt * nvw : 0;
  // half the LDS when the term fits in it, so two workgroups share a CU
  const size_t half = (kLdsLimit / 2 - 1024) / 8;
  const size_t budget =
      base + std::max<size_t>(keo_words, 2 * (size_t)nvw + 2 * nbt) <= half ? half : kLdsLimit / 8;
  words = base + keo_words;
  auto split = [&](const int *ml, int nm, int &m_, int &nbA_, int &QA_, int &nbB_, int &QB_,
                   int &step0) -> bool {
    int best = -1;
    int64_t nbA = 1, QA = 1;
    for (int m = 0; m <= nm; ++m) {
      if (m > 0) {
        nbA *= nbd[ml[m - 1]];
        QA *= nqd[ml[m - 1]];
      }
      const size_t work = std::max<size_t>(keo_words, 2 * (size_t)QA * nvw + 2 * (nbt / nbA));
      if (base + work <= budget) best = m;
    }
    if (best < 0) return false;
    m_ = best;
    nbA_ = 1;
    QA_ = 1;
    int64_t qall = 1;
    for (int k = 0; k < nm; ++k) qall *= nqd[ml[k]];
    for (int k = 0; k < best; ++k) { nbA_ *= nbd[ml[k]]; QA_ *= nqd[ml[k]]; }
    nbB_ = (int)(nbt / nbA_);
    QB_ = (int)(qall / QA_);
    step0 = (int)steps.size();
    for (int k = 0; k < best; ++k) {
      evr::StepDev sd{};
      int n1 = 1, n3 = 1;
      for (int e2 = 0; e2 < k; ++e2) n1 *= nqd[ml[e2]];
      for (int e2 = k + 1; e2 < best; ++e2) n3 *= nbd[ml[e2]];
      sd.n1 = n1; sd.n3 = n3; sd.nin = nbd[ml[k]]; sd.nout = nqd[ml[k]];
      sd.mat = p->slots[td.slot[ml[k]]].offB;
      steps.push_back(sd);
    }
    for (int k = 0; k < best; ++k) {
      evr::StepDev sd{};
      int n1 = 1, n3 = 1;
      for (int e2 = 0; e2 < k; ++e2) n1 *= nbd[ml[e2]];
      for (int e2 = k + 1; e2 < best; ++e2) n3 *= nqd[ml[e2]];
      sd.n1 = n1; sd.n3 = n3; sd.nin = nqd[ml[k]]; sd.nout = nbd[ml[k]];
      sd.mat = p->slots[td.slot[ml[k]]].offW;
      steps.push_back(sd);
    }
    words = std::max(words, base + std::max<size_t>(keo_words, 2 * (size_t)QA_ * nvw + 2 * nbB_));
    return true;
  };
  return split(td.vmode, td.Dv, td.vm, td.vnbA, td.vQA, td.vnbB, td.vQB, td.vstep0);
}

// Lays out every term for its class p->tclass[t]: slab splits and steps, LDS
// sizes and flops per class, the per-class work lists (most expensive terms
// first); uploads the term table, the steps and the lists.
int apply_classes(evr_plan *p)
{
  std::vector<evr::StepDev> steps;
  for (int c = 0; c < kNumClasses; ++c) {
    p->lds_bytes64[c] = p->lds_bytes32[c] = 0;
    p->flops_class[c] = 0.0;
    p->nunits[c] = p->nterms_class[c] = 0;
    p->d_order[c].release();
  }
  std::vector<double> cost(p->nterms);
  for (int64_t t = 0; t < p->nterms; ++t) {
    evr::TermDev &td = p->terms[t];
    const int c = p->tclass[t];
    if (c < 0) continue;  // streamed term
    if (kClasses[c].max_nbt() < td.nbt) return fail(EVR_ERR_ARG, "term class too small for its basis");
    size_t words = 0;
    if (!layout_term(p, td, c, steps, words)) return fail(EVR_ERR_SIZE, "term does not fit in LDS");
    p->lds_bytes64[c] = std::max(p->lds_bytes64[c], words * 8);
    p->lds_bytes32[c] = std::max(p->lds_bytes32[c], words * 4);
    p->flops_class[c] += term_flops(p, td);
    double qv = 1.0;
    for (int j = 0; j < td.Dv; ++j) qv *= p->slots[td.slot[td.vmode[j]]].nq;
    cost[t] = qv * 50.0 + (double)td.nbt * (td.na + 1) * (td.na + 1) * 6.0;
  }
  p->steps.swap(steps);
  for (int c = 0; c < kNumClasses; ++c) {
    std::vector<int> ids;
    for (int64_t t = 0; t < p->nterms; ++t)
      if (p->tclass[t] == c) ids.push_back((int)t);
    if (ids.empty()) continue;
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    p->nterms_class[c] = (int64_t)ids.size();
    while (ids.size() % 8) ids.push_back(-1);
    p->nunits[c] = (int64_t)ids.size();
    HIP_TRY(p->d_order[c].upload(ids.data(), ids.size(), p->stream));
  }
  HIP_TRY(p->d_terms.upload(p->terms.data(), p->terms.size(), p->stream));
  HIP_TRY(p->d_steps.upload(p->steps.data(), p->steps.size(), p->stream));
  return EVR_OK;
}

// The streamed terms of a run_terms call (evr_stream.hip), one term and up to
// big_k vectors at a time: H|psi> reads X and writes the term rows of R, B2G
// reads X and writes the grid columns of G, G2B reads the grid columns of Gin
// and writes R.  Whole-device launches on the plan stream, nothing allocated.
template <typename T>
int run_streamed(evr_plan *p, const std::vector<int64_t> &list, int mode, int nvec, const T *X, T *R, T *G,
                 const T *Gin)
{
  const int D = p->D;
  const int64_t S = p->big_nq * p->big_k;
  T *base = reinterpret_cast<T *>(p->d_big.ptr);
  T *b_in = base, *b_g = base + S, *b_a = base + 2 * S, *b_b = base + 3 * S, *b_acc = base + 4 * S,
    *b_y0 = base + 5 * S, *b_y1 = base + 6 * S, *b_P = base + 7 * S;
  auto blocks = [](int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); };
  for (int64_t t : list) {
    const evr::TermDev &td = p->terms[t];
    int nb[evr::kMaxDev], nq[evr::kMaxDev];
    const evr::SlotDev *sl[evr::kMaxDev];
    for (int d = 0; d < D; ++d) {
      sl[d] = &p->slots[td.slot[d]];
      nb[d] = sl[d]->nb;
      nq[d] = sl[d]->nq;
    }
    const double *bw = p->d_bw64.ptr;
    for (int c0 = 0; c0 < nvec; c0 += p->big_k) {
      const int k = std::min(p->big_k, nvec - c0);
      const int64_t nbk = (int64_t)td.nbt * k, nqk = (int64_t)td.nqt * k;
      if (nqk >= INT32_MAX) return fail(EVR_ERR_SIZE, "streamed term grid too large for one pass");
      hipError_t e = hipSuccess;
      auto pass = [&](const T *in, T *out, const double *M, int ld, int trans, int64_t L, int nin, int nout,
                      int64_t Rr, int accum = 0, double scale = 1.0) {
        if (L >= 16) {  // tiles of 16 pencils on MFMA
          const int64_t tiles = (L + 15) / 16 * Rr;
          hipLaunchKernelGGL(evr::k_mode_mfma<T>, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 65536))),
                             dim3(256), 0, p->stream, in, out, M, ld, trans, L, nin, nout, Rr, accum, scale);
          if (e == hipSuccess) e = hipGetLastError();
          return;
        }
        // output chunks per pencil: enough threads for about 1024 per CU
        const int64_t want = (int64_t)p->ncu * 1024;
        const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(nout, want / std::max<int64_t>(1, L * Rr)));
        const int osz = (nout + nchunk - 1) / nchunk;
        const int64_t thr = L * Rr * ((nout + osz - 1) / osz);
        const dim3 g(blocks(thr)), b(256);
        if (nin <= 16)
          hipLaunchKernelGGL((evr::k_mode_pass<T, 16>), g, b, 0, p->stream, in, out, M, ld, trans, L, nin, nout, Rr, osz,
                             accum, scale);
        else if (nin <= 32)
          hipLaunchKernelGGL((evr::k_mode_pass<T, 32>), g, b, 0, p->stream, in, out, M, ld, trans, L, nin, nout, Rr, osz,
                             accum, scale);
        else
          hipLaunchKernelGGL((evr::k_mode_pass<T, 64>), g, b, 0, p->stream, in, out, M, ld, trans, L, nin, nout, Rr, osz,
                             accum, scale);
        if (e == hipSuccess) e = hipGetLastError();
      };
      // basis -> grid over all modes (B, or W^T for the transposed operator)
      auto b2g = [&](const T *src, T *dst, int trans) {
        int64_t L = k, Rr = td.nbt;
        const T *cur = src;
        for (int d = 0; d < D; ++d) {
          Rr /= nb[d];
          T *out = d == D - 1 ? dst : ((d & 1) ? b_b : b_a);
          if (trans) pass(cur, out, bw + sl[d]->offW, nb[d], 1, L, nb[d], nq[d], Rr);
          else pass(cur, out, bw + sl[d]->offB, nq[d], 0, L, nb[d], nq[d], Rr);
          L *= nq[d];
          cur = out;
        }
      };
      // grid -> basis over all modes (W, or B^T for the transposed operator)
      auto g2b = [&](const T *src, T *dst, int trans) {
        int64_t L = k, Rr = td.nqt;
        const T *cur = src;
        for (int d = 0; d < D; ++d) {
          Rr /= nq[d];
          T *out = d == D - 1 ? dst : ((d & 1) ? b_b : b_a);
          if (trans) pass(cur, out, bw + sl[d]->offB, nq[d], 1, L, nq[d], nb[d], Rr);
          else pass(cur, out, bw + sl[d]->offW, nb[d], 0, L, nq[d], nb[d], Rr);
          L *= nb[d];
          cur = out;
        }
      };
      // grid derivative on mode d (D1 or D2 of its slot, nq x nq); with
      // accum, dst += scale d(src)
      auto deriv = [&](const T *src, T *dst, int d, const double *Dm, int trans, int accum = 0,
                       double scale = 1.0) {
        int64_t L = k, Rr = 1;
        for (int m = 0; m < d; ++m) L *= nq[m];
        for (int m = d + 1; m < D; ++m) Rr *= nq[m];
        pass(src, dst, Dm + sl[d]->offD1, nq[d], trans, L, nq[d], nq[d], Rr, accum, scale);
      };
      // the grid operator of the term on b_g into b_acc (OPSG4:1448-1650)
      auto grid_op = [&](int trans) {
        const double *V = p->hasV ? p->d_V.ptr + td.off_nq : nullptr;
        hipLaunchKernelGGL(evr::k_stream_vmul<T>, dim3(blocks(nqk)), dim3(256), 0, p->stream, V, b_g,
                           (unsigned)nqk, k, b_acc);
        if (e == hipSuccess) e = hipGetLastError();
        if (p->type == 1) {
          for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) {
              const double f = p->hC[i * D + j];
              if (f == 0.0) continue;
              if (i == j) {
                deriv(b_g, b_acc, i, p->d_D2.ptr, trans, 1, f);
              } else {
                deriv(b_g, b_a, i, p->d_D1.ptr, trans);
                deriv(b_a, b_acc, j, p->d_D1.ptr, trans, 1, f);
              }
            }
          if (p->hasF1)
            for (int i = 0; i < D; ++i)
              if (p->hF1[i] != 0.0) deriv(b_g, b_acc, i, p->d_D1.ptr, trans, 1, p->hF1[i]);
        } else if (p->type == 10) {
          evr::StreamGrid sg{};
          sg.D = D;
          for (int d = 0; d < D; ++d) {
            sg.nq[d] = nq[d];
            sg.offX[d] = sl[d]->offX;
          }
          for (int j = 0; j < D; ++j) deriv(b_g, b_P + (size_t)j * nqk, j, p->d_D1.ptr, trans);
          for (int i = 0; i < D; ++i) {
            hipLaunchKernelGGL(evr::k_stream_gcomb<T>, dim3((unsigned)((nqk + 255) / 256)), dim3(256), 0, p->stream,
                               sg, i, trans,
                               p->d_coefC.ptr, p->hasA ? p->d_coefA.ptr : nullptr, p->d_x.ptr, b_P,
                               (unsigned)nqk, k, b_a);
            if (e == hipSuccess) e = hipGetLastError();
            deriv(b_a, b_acc, i, p->d_D1.ptr, trans, 1, 1.0);
          }
        }
      };
      if (mode == evr::kModeHpsi || mode == evr::kModeB2G) {
        hipLaunchKernelGGL(evr::k_stream_gather<T>, dim3(blocks(nbk)), dim3(256), 0, p->stream, X, nvec, c0, k,
                           p->d_map.ptr + td.off_nb, (unsigned)nbk, b_in);
        if (e == hipSuccess) e = hipGetLastError();
      }
      if (mode == evr::kModeB2G) {
        b2g(b_in, b_g, 0);
        hipLaunchKernelGGL(evr::k_stream_grid_out<T>, dim3(blocks(nqk)), dim3(256), 0, p->stream, b_g,
                           (unsigned)nqk, k, c0, G + td.off_nq, p->sum_nq);
        if (e == hipSuccess) e = hipGetLastError();
      } else if (mode == evr::kModeG2B) {
        hipLaunchKernelGGL(evr::k_stream_grid_in<T>, dim3(blocks(nqk)), dim3(256), 0, p->stream,
                           Gin + td.off_nq, p->sum_nq, (unsigned)nqk, k, c0, b_g);
        if (e == hipSuccess) e = hipGetLastError();
        g2b(b_g, b_y0, 0);
        const double w = std::fabs(td.w) < 1e-6 ? 0.0 : td.w;  // BGSG4:1004
        hipLaunchKernelGGL(evr::k_stream_scatter<T>, dim3(blocks(nbk)), dim3(256), 0, p->stream, b_y0,
                           (const T *)nullptr, w, nvec, c0, k, p->d_rowpos.ptr + td.off_nb, (unsigned)nbk, R);
        if (e == hipSuccess) e = hipGetLastError();
      } else {
        b2g(b_in, b_g, 0);
        grid_op(0);
        g2b(b_acc, b_y0, 0);
        // the hermitian form: (H_t + H_t^T) / 2, the transpose through W^T,
        // the transposed grid operator and B^T (the V part is symmetric)
        const bool sym = p->hermitian && p->hasKeo;
        if (sym) {
          b2g(b_in, b_g, 1);
          grid_op(1);
          g2b(b_acc, b_y1, 1);
        }
        hipLaunchKernelGGL(evr::k_stream_scatter<T>, dim3(blocks(nbk)), dim3(256), 0, p->stream, b_y0,
                           sym ? (const T *)b_y1 : (const T *)nullptr, td.w, nvec, c0, k,
                           p->d_rowpos.ptr + td.off_nb, (unsigned)nbk, R);
        if (e == hipSuccess) e = hipGetLastError();
      }
      if (e != hipSuccess) return fail(EVR_ERR_HIP, std::string("streamed term: ") + hipGetErrorString(e));
    }
  }
  return EVR_OK;
}

// Apply every term class for `nvec` vectors held in X ([nb_packed][nvec]).
template <typename T>
int run_terms(evr_plan *p, int mode, int nvec, const T *X, T *R, T *G, const T *Gin, const T *bw,
              const T *keo, const T *V, int diagKeo = 0)
{
  if (!p->one.empty() && !diagKeo && (mode == evr::kModeHpsi || mode == evr::kModeB2G || mode == evr::kModeG2B))
    return run_streamed<T>(p, p->one, mode, nvec, X, R, G, Gin);
  for (int c = 0; c < kNumClasses; ++c) {
    if (p->nunits[c] == 0) continue;
    evr::TermParams<T> prm = base_params<T>(p, mode, nvec);
    prm.diagKeo = diagKeo;
    prm.nunits = p->nunits[c];
    prm.order = p->d_order[c].ptr;
    prm.bw = bw;
    prm.keo = keo;
    prm.V = V;
    prm.X = X;
    prm.R = R;
    prm.G = G;
    prm.Gin = Gin;
    size_t lds = sizeof(T) == 8 ? p->lds_bytes64[c] : p->lds_bytes32[c];
    if (p->profiling) (void)hipEventRecord(p->ev[2 * c], p->stream);
    hipError_t e = launch_class<T>(c, prm, p->nunits[c], nvec, lds, p->stream);
    if (e != hipSuccess) return fail(EVR_ERR_HIP, std::string("term kernel: ") + hipGetErrorString(e));
    if (p->profiling) {
      (void)hipEventRecord(p->ev[2 * c + 1], p->stream);
      (void)hipEventSynchronize(p->ev[2 * c + 1]);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, p->ev[2 * c], p->ev[2 * c + 1]);
      p->ms_acc[c] += ms;
      p->flops_acc[c] += p->flops_class[c] * nvec;
      p->launches[c] += 1;
    }
  }
  if (!p->big.empty()) {
    if (diagKeo) return fail(EVR_ERR_ARG, "the kinetic diagonal of a streamed term is not formed here");
    return run_streamed<T>(p, p->big, mode, nvec, X, R, G, Gin);
  }
  return EVR_OK;
}

unsigned blocks_for(int64_t n)
{
  int64_t b = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

// a[0 .. n) = 0 with a 64-bit fill kernel (the assembled pools are well past
// 2^34 words)
template <typename T>
hipError_t zero_fill(T *a, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(evr::k_zero<T>, dim3(blocks_for(n)), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

// Columns c < nc of X (nb x nc, column-major) = unit vectors e_{b0 + c}.
__global__ void k_unit_cols(int64_t nb, int nc, int64_t b0, double *X)
{
  const int64_t n = nb * nc;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k % nb, c = k / nb;
    X[k] = i == b0 + c ? 1.0 : 0.0;
  }
}

// diag[b0 + c] = Y[b0 + c, c] for c < nc
__global__ void k_take_diag(int64_t nb, int nc, int64_t b0, const double *Y, double *diag)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) diag[b0 + c] = Y[b0 + c + (int64_t)c * nb];
}

// What the assembled apply and the scatter read: one plan's tables, or the
// concatenated tables of several plans (evr_multi).
struct AsmOp {
  const evr::TermDev *terms;
  const int2 *items;
  int64_t n_items;
  const int *map, *rowpos, *seg, *fseg;
  int64_t nseg, nb_packed;
  double *P;               // segment sums
  hipStream_t stream;
};

AsmOp plan_op(const evr_plan *p)
{
  return AsmOp{p->d_terms.ptr, p->d_asm_items.ptr, p->n_asm_items, p->d_map.ptr, p->d_rowpos.ptr,
               p->d_seg.ptr, p->d_fseg.ptr, p->nseg, p->nb_packed, p->d_P.ptr, p->stream};
}

// Deterministic Smolyak scatter R -> out (nb_packed x nvec, column-major).
template <typename T>
hipError_t reduce_op(const AsmOp &o, const T *R, int nvec, double *out)
{
  const int64_t sblocks = (o.nseg + evr::kReduceTile - 1) / evr::kReduceTile;
  hipLaunchKernelGGL(evr::k_reduce_seg<T>, dim3((unsigned)std::max<int64_t>(sblocks, 1)), dim3(256), 0,
                     o.stream, R, o.seg, o.nseg, nvec, o.P);
  const size_t lds = sizeof(double) * evr::kReduceTile * (size_t)(nvec + 1);
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&evr::k_reduce_fn),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int64_t blocks = (o.nb_packed + evr::kReduceTile - 1) / evr::kReduceTile;
  hipLaunchKernelGGL(evr::k_reduce_fn, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), lds,
                     o.stream, o.P, o.fseg, o.nb_packed, nvec, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_reduce(evr_plan *p, const T *R, int nvec, double *out, double *P = nullptr)
{
  AsmOp o = plan_op(p);
  if (P) o.P = P;
  return reduce_op<T>(o, R, nvec, out);
}

// R = H_t X_t for every term from the assembled pool A (32 vectors per pass).
template <typename T>
hipError_t apply_op(const AsmOp &o, const T *A, const T *X, int nvec, T *R)
{
  for (int vb = 0; vb < nvec; vb += 32) {
    const dim3 grid((unsigned)o.n_items);
    if (nvec - vb > 16)
      hipLaunchKernelGGL((evr::k_apply_asm<T, 2>), grid, dim3(256), 0, o.stream, o.terms, o.items, A,
                         o.map, o.rowpos, X, nvec, vb, R);
    else
      hipLaunchKernelGGL((evr::k_apply_asm<T, 1>), grid, dim3(256), 0, o.stream, o.terms, o.items, A,
                         o.map, o.rowpos, X, nvec, vb, R);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Work items of the assembled apply: (term, first of 16 row tiles).
// Workgroups are handed to the 8 XCDs round-robin (workgroup i on XCD i % 8),
// so all items of a term go to one XCD, in consecutive turns: the term's X
// rows are then read from HBM once and from that XCD's L2 afterwards.  Terms
// are spread over the XCDs by matrix size, largest first.
std::vector<int2> asm_items(const std::vector<int> &nbt)
{
  constexpr int kXcd = 8;
  const int64_t nt = (int64_t)nbt.size();
  std::vector<int64_t> order(nt);
  for (int64_t t = 0; t < nt; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return nbt[a] > nbt[b]; });
  std::vector<std::vector<int2>> queue(kXcd);
  std::vector<double> bytes(kXcd, 0.0);
  size_t total = 0;
  for (int64_t t : order) {
    const int mtn = (nbt[t] + 15) / 16;
    const int x = (int)(std::min_element(bytes.begin(), bytes.end()) - bytes.begin());
    for (int m0 = 0; m0 < mtn; m0 += 16) queue[x].push_back(int2{(int)t, m0});
    bytes[x] += (double)mtn * evr::asm_ksteps(nbt[t]);
    total += (mtn + 15) / 16;
  }
  std::vector<int2> items;
  items.reserve(total);
  std::vector<size_t> head(kXcd, 0);
  while (items.size() < total) {
    for (int x = 0; x < kXcd && items.size() < total; ++x) {
      if (head[x] < queue[x].size()) {
        items.push_back(queue[x][head[x]++]);
      } else {
        // this XCD's queue is empty: take the last item of the longest one
        int y = 0;
        for (int z = 1; z < kXcd; ++z)
          if (queue[z].size() - head[z] > queue[y].size() - head[y]) y = z;
        items.push_back(queue[y].back());
        queue[y].pop_back();
      }
    }
  }
  return items;
}

template <typename T>
hipError_t apply_assembled(evr_plan *p, const T *A, const T *X, int nvec, T *R)
{
  const int slot = kNumClasses;  // profiling slot of the assembled apply
  if (p->profiling) (void)hipEventRecord(p->ev[2 * slot], p->stream);
  const hipError_t e = apply_op<T>(plan_op(p), A, X, nvec, R);
  if (e != hipSuccess) return e;
  if (p->profiling) {
    (void)hipEventRecord(p->ev[2 * slot + 1], p->stream);
    (void)hipEventSynchronize(p->ev[2 * slot + 1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p->ev[2 * slot], p->ev[2 * slot + 1]);
    p->ms_acc[slot] += ms;
    p->flops_acc[slot] += p->asm_flops_per_vec * nvec;
    p->launches[slot] += 1;
  }
  return hipSuccess;
}

template <typename T, int NT, int EPT, int NVW>
hipError_t launch_items_one(const evr::TermParams<T> &prm, int64_t nitems, size_t lds, hipStream_t s)
{
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&evr::k_term<T, NT, EPT, NVW>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  // the dispatch grid is counted in work-items in 32 bits: at most
  // (2^32 - 1) / NT workgroups per launch, so long item lists go in pieces
  constexpr int64_t kMaxBlocks = (int64_t)0xFFFFFFFFu / NT;
  for (int64_t i0 = 0; i0 < nitems; i0 += kMaxBlocks) {
    evr::TermParams<T> q = prm;
    q.items = prm.items + i0;
    const int64_t nb = std::min<int64_t>(kMaxBlocks, nitems - i0);
    hipLaunchKernelGGL((evr::k_term<T, NT, EPT, NVW>), dim3((unsigned)nb), dim3(NT), lds, s, q);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_items(int c, const evr::TermParams<double> &prm, int64_t nitems, size_t lds,
                        hipStream_t s)
{
  switch (c) {
    case 0: return launch_items_one<double, 256, 1, 4>(prm, nitems, lds, s);
    case 1: return launch_items_one<double, 512, 1, 4>(prm, nitems, lds, s);
    case 2: return launch_items_one<double, 512, 2, 4>(prm, nitems, lds, s);
    case 3: return launch_items_one<double, 512, 2, 2>(prm, nitems, lds, s);
    case 4: return launch_items_one<double, 512, 4, 2>(prm, nitems, lds, s);
    case 5: return launch_items_one<double, 512, 4, 1>(prm, nitems, lds, s);
    case 6: return launch_items_one<double, 1024, 4, 1>(prm, nitems, lds, s);
    case 7: return launch_items_one<double, 1024, 8, 1>(prm, nitems, lds, s);
    default: return hipErrorInvalidValue;
  }
}

// H|psi> of a plan with G, J, s (or F2, F1) per grid point into dh (device,
// nb_packed x nvec): B->G, the reference's grid form of every term
// (k_gk_deriv, k_gk_combine, k_gk_out or k_t1_out, V included), then the
// weighted G->B and the deterministic scatter, gvec vectors at a time.  T is
// the precision of the transforms and of the grid columns (FP32 keeps the
// tables and the grid arithmetic in FP64; the work grids hold floats).
template <typename T>
int keo_grid_hpsi(evr_plan *p, int nvec, const double *dpsi, double *dh)
{
  const int D = p->D;
  const int64_t nq = p->sum_nq, nb = p->nb_packed;
  constexpr bool f32 = sizeof(T) == 4;
  T *Gg = reinterpret_cast<T *>(p->d_Gg.ptr), *Pg = reinterpret_cast<T *>(p->d_Pg.ptr),
    *Fg = reinterpret_cast<T *>(p->d_Fg.ptr), *R = reinterpret_cast<T *>(p->d_R.ptr);
  const T *bw = f32 ? reinterpret_cast<const T *>(p->d_bw32.ptr) : reinterpret_cast<const T *>(p->d_bw64.ptr);
  const T *keo = f32 ? reinterpret_cast<const T *>(p->d_keo32.ptr) : reinterpret_cast<const T *>(p->d_keo64.ptr);
  for (int c0 = 0; c0 < nvec; c0 += p->gvec) {
    const int k = std::min(p->gvec, nvec - c0);
    int rc;
    if constexpr (f32) {
      T *X = reinterpret_cast<T *>(p->d_X.ptr);
      hipLaunchKernelGGL(evr::k_pack_in<T>, dim3(blocks_for(nb * k)), dim3(256), 0, p->stream, dpsi + nb * c0,
                         nb, k, X);
      HIP_TRY(hipGetLastError());
      rc = run_terms<T>(p, evr::kModeB2G, k, X, nullptr, Gg, nullptr, bw, keo, nullptr);
    } else {
      rc = evr::plan_b2g_device(p, k, dpsi + nb * c0, p->d_Gg.ptr);
    }
    if (rc) return rc;
    const bool t1 = p->grid_type == 1;
    hipLaunchKernelGGL(evr::k_gk_deriv<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                       p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_gitems.ptr,
                       t1 ? nullptr : p->d_Sg.ptr, Gg, nq, k, Pg, 0);
    HIP_TRY(hipGetLastError());
    if (t1) {
      hipLaunchKernelGGL(evr::k_t1_out<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_D2.ptr, p->d_gitems.ptr,
                         p->d_F2g.ptr, p->d_F1g.ptr, p->d_V.ptr, Gg, Pg, nq, k, Fg);
      HIP_TRY(hipGetLastError());
    } else {
      hipLaunchKernelGGL(evr::k_gk_combine<T>, dim3(blocks_for(nq)), dim3(256), 0, p->stream, D, nq, k,
                         p->d_GG.ptr, p->d_Jg.ptr, Pg, 0);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(evr::k_gk_out<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_gitems.ptr, p->d_Jg.ptr,
                         p->d_Sg.ptr, p->d_V.ptr, Gg, Pg, nq, k, Fg, (const double *)nullptr);
      HIP_TRY(hipGetLastError());
    }
    rc = run_terms<T>(p, evr::kModeG2B, k, nullptr, R, nullptr, Fg, bw, keo, nullptr);
    if (rc) return rc;
    if (p->hermitian && !t1) {
      // H^T psi = W diag(1/w) Op^T diag(w) B psi (W = B^T diag(w) on every
      // term grid), so the transposed sequence reuses the B->G columns in Gg
      // and the weighted G->B: d_i^T on (w / (J s)) psi_G, J sum_i G_ji,
      // -1/2 (s / w) sum_j d_j^T; the rows of H and H^T are then averaged
      T *R2 = reinterpret_cast<T *>(p->d_R2.ptr);
      hipLaunchKernelGGL(evr::k_gk_deriv<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_gitems.ptr, p->d_As.ptr, Gg, nq, k,
                         Pg, 1);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(evr::k_gk_combine<T>, dim3(blocks_for(nq)), dim3(256), 0, p->stream, D, nq, k,
                         p->d_GG.ptr, p->d_Jg.ptr, Pg, 1);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(evr::k_gk_out<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_gitems.ptr, p->d_Jg.ptr,
                         p->d_Sg.ptr, p->d_V.ptr, Gg, Pg, nq, k, Fg, (const double *)p->d_Wq.ptr);
      HIP_TRY(hipGetLastError());
      rc = run_terms<T>(p, evr::kModeG2B, k, nullptr, R2, nullptr, Fg, bw, keo, nullptr);
      if (rc) return rc;
      const int64_t nr = p->nrows * k;
      hipLaunchKernelGGL(evr::k_half_sum<T>, dim3(blocks_for(nr)), dim3(256), 0, p->stream, R, (const T *)R2, nr);
      HIP_TRY(hipGetLastError());
    } else if (p->hermitian) {
      // type 1: the weighted transpose W diag(1/w) Op^T diag(w) B of the grid
      // operator (k_t1t_u, k_t1t_out) on the same B->G columns
      T *R2 = reinterpret_cast<T *>(p->d_R2.ptr);
      hipLaunchKernelGGL(evr::k_t1t_u<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_gitems.ptr, p->d_F2w.ptr, Gg, nq, k,
                         Pg);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(evr::k_t1t_out<T>, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream, D,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_D2.ptr, p->d_gitems.ptr, p->d_F2w.ptr,
                         p->d_F1w.ptr, p->d_Wq.ptr, p->d_V.ptr, Gg, Pg, nq, k, Fg);
      HIP_TRY(hipGetLastError());
      rc = run_terms<T>(p, evr::kModeG2B, k, nullptr, R2, nullptr, Fg, bw, keo, nullptr);
      if (rc) return rc;
      const int64_t nr = p->nrows * k;
      hipLaunchKernelGGL(evr::k_half_sum<T>, dim3(blocks_for(nr)), dim3(256), 0, p->stream, R, (const T *)R2, nr);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(launch_reduce<T>(p, R, k, dh + nb * c0));
  }
  return EVR_OK;
}

// V_eff = V summed over the one-function modes of every term (k_fold_v),
// queued on the plan stream after the V it reads.
int fold_v(evr_plan *p)
{
  if (p->nterms == 0) return EVR_OK;
  hipLaunchKernelGGL(evr::k_fold_v, dim3((unsigned)p->nterms), dim3(256), 0, p->stream, p->D,
                     p->d_terms.ptr, p->d_slots.ptr, p->d_bw64.ptr, p->d_V.ptr, p->d_Veff.ptr);
  HIP_TRY(hipGetLastError());
  return EVR_OK;
}

// Work items (term, first point) of the grid-form kernels and the work grids
// of gvec vectors at a time (psi_G, its D derivatives, the operator output):
// made once per plan.
int grid_items(evr_plan *p)
{
  if (p->d_gitems.ptr) return EVR_OK;
  std::vector<int2> items;
  for (int64_t t = 0; t < p->nterms; ++t)
    for (int q0 = 0; q0 < p->terms[t].nqt; q0 += 256) items.push_back(make_int2((int)t, q0));
  if ((int64_t)items.size() > (int64_t)0xFFFFFFFFu / 256)
    return fail(EVR_ERR_SIZE, "too many grid work items for one launch");
  p->n_gitems = (int64_t)items.size();
  HIP_TRY(p->d_gitems.upload(items.data(), items.size(), p->stream));
  return EVR_OK;
}

int grid_form_buffers(evr_plan *p)
{
  if (p->d_Gg.ptr) return EVR_OK;
  const int D = p->D;
  const int64_t nq = p->sum_nq;
  int rc = grid_items(p);
  if (rc) return rc;
  // (D + 2) x sum_nq doubles per vector
  p->gvec = (int)std::max<int64_t>(1, std::min<int64_t>(p->max_nvec, (int64_t)(8e9 / (8.0 * (D + 2) * nq))));
  HIP_TRY(p->d_Gg.alloc((size_t)nq * p->gvec));
  HIP_TRY(p->d_Pg.alloc((size_t)nq * D * p->gvec));
  HIP_TRY(p->d_Fg.alloc((size_t)nq * p->gvec));
  return EVR_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
static void recount_flops(evr_plan *p);

extern "C" {

const char *evr_last_error(void) { return g_err.c_str(); }

static int wrap_desc(std::unique_ptr<evr_desc_owner> &o, evr_basis_desc **out)
{
  evr::BasisTables &t = o->tab;
  evr_basis_desc &d = o->desc;
  d.D = t.D;
  d.nlev = t.nlev;
  d.nterms = t.nterms;
  d.term_levels = t.term_levels.data();
  d.weights = t.weights.data();
  d.nb = t.nb.data();
  d.nq = t.nq.data();
  d.x = t.x.data();
  d.B = t.B.data();
  d.W = t.W.data();
  d.D1 = t.D1.data();
  d.D2 = t.D2.data();
  d.nb_packed = t.nb_packed;
  d.map = t.map.data();
  d.grid_last_fastest = t.dp_order ? 1 : 0;
  *out = &o.release()->desc;  // desc is the first member of evr_desc_owner
  return EVR_OK;
}

int evr_sg4_hermite_desc(int D, int LG, int LB, int nbA, int nbB, int nqA, int nqB,
                         evr_basis_desc **out)
{
  if (!out) return fail(EVR_ERR_ARG, "out is NULL");
  std::unique_ptr<evr_desc_owner> o(new evr_desc_owner());
  std::string err = evr::build_sg4_hermite(D, LG, LB, nbA, nbB, nqA, nqB, o->tab);
  if (!err.empty()) return fail(EVR_ERR_ARG, err);
  return wrap_desc(o, out);
}

int evr_sg4_hermite_desc_modes(int D, int LG, int LB, const int *nbA, const int *nbB, const int *nqA,
                               const int *nqB, evr_basis_desc **out)
{
  if (!out || !nbA || !nbB || !nqA || !nqB) return fail(EVR_ERR_ARG, "NULL argument");
  std::unique_ptr<evr_desc_owner> o(new evr_desc_owner());
  std::string err = evr::build_sg4_hermite_modes(D, LG, LB, nbA, nbB, nqA, nqB, o->tab);
  if (!err.empty()) return fail(EVR_ERR_ARG, err);
  return wrap_desc(o, out);
}

int evr_dp_hermite_desc(int D, const int *nb, const int *nq, evr_basis_desc **out)
{
  if (!out || !nb || !nq) return fail(EVR_ERR_ARG, "NULL argument");
  std::unique_ptr<evr_desc_owner> o(new evr_desc_owner());
  std::string err = evr::build_dp_hermite(D, nb, nq, o->tab);
  if (!err.empty()) return fail(EVR_ERR_ARG, err);
  return wrap_desc(o, out);
}

void evr_basis_desc_free(evr_basis_desc *d)
{
  if (d) delete reinterpret_cast<evr_desc_owner *>(d);
}

static void term_sizes(const evr_basis_desc *d, int64_t t, int64_t &nb, int64_t &nq)
{
  nb = nq = 1;
  for (int k = 0; k < d->D; ++k) {
    const int s = k * d->nlev + d->term_levels[t * d->D + k];
    nb *= d->nb[s];
    nq *= d->nq[s];
  }
}

int evr_basis_desc_sizes(const evr_basis_desc *d, int64_t *out)
{
  if (!d || !out) return fail(EVR_ERR_ARG, "NULL argument");
  int64_t snb = 0, snq = 0, mnb = 0, mnq = 0;
  for (int64_t t = 0; t < d->nterms; ++t) {
    int64_t nb, nq;
    term_sizes(d, t, nb, nq);
    snb += nb;
    snq += nq;
    mnb = std::max(mnb, nb);
    mnq = std::max(mnq, nq);
  }
  out[0] = snb; out[1] = snq; out[2] = mnb; out[3] = mnq;
  return EVR_OK;
}

// offsets of the per-slot arrays inside the descriptor pools
static void slot_offsets(const evr_basis_desc *d, std::vector<int64_t> &ox, std::vector<int64_t> &oB,
                         std::vector<int64_t> &oD)
{
  const int ns = d->D * d->nlev;
  ox.assign(ns + 1, 0);
  oB.assign(ns + 1, 0);
  oD.assign(ns + 1, 0);
  for (int s = 0; s < ns; ++s) {
    ox[s + 1] = ox[s] + d->nq[s];
    oB[s + 1] = oB[s] + (int64_t)d->nq[s] * d->nb[s];
    oD[s + 1] = oD[s] + (int64_t)d->nq[s] * d->nq[s];
  }
}

int evr_basis_desc_grid(const evr_basis_desc *d, double *q)
{
  if (!d || !q) return fail(EVR_ERR_ARG, "NULL argument");
  std::vector<int64_t> ox, oB, oD;
  slot_offsets(d, ox, oB, oD);
  if (d->grid_last_fastest) {
    int64_t nq = 1;
    for (int k = 0; k < d->D; ++k) nq *= d->nq[k * d->nlev];
    for (int64_t i = 0; i < nq; ++i) {
      int64_t r = i;
      for (int k = d->D - 1; k >= 0; --k) {
        const int s = k * d->nlev;
        q[i * d->D + k] = d->x[ox[s] + r % d->nq[s]];
        r /= d->nq[s];
      }
    }
    return EVR_OK;
  }
  int64_t off = 0;
  for (int64_t t = 0; t < d->nterms; ++t) {
    int64_t nb, nq;
    term_sizes(d, t, nb, nq);
    for (int64_t i = 0; i < nq; ++i) {
      int64_t r = i;
      for (int k = 0; k < d->D; ++k) {
        const int s = k * d->nlev + d->term_levels[t * d->D + k];
        q[(off + i) * d->D + k] = d->x[ox[s] + r % d->nq[s]];
        r /= d->nq[s];
      }
    }
    off += nq;
  }
  return EVR_OK;
}

int evr_plan_create(const evr_basis_desc *bd, const evr_op_desc *op, int max_nvec,
                    evr_plan **out)
{
  if (!bd || !op || !out) return fail(EVR_ERR_ARG, "NULL argument");
  if (bd->D < 1 || bd->D > evr::kMaxD) return fail(EVR_ERR_ARG, "D must be in 1..64");
  if (op->type != 0 && op->type != 1 && op->type != 10)
    return fail(EVR_ERR_ARG, "operator type must be 0, 1 or 10");
  if ((op->type == 10 && !op->G0) || (op->type == 1 && !op->F2))
    return fail(EVR_ERR_ARG, "type 10 needs G0, type 1 needs F2");
  if (max_nvec < 1) return fail(EVR_ERR_ARG, "max_nvec must be >= 1");
  if (op->hermitian && op->type == 10) {
    const int D0 = bd->D;
    for (int i = 0; i < D0; ++i)
      for (int j = 0; j < D0; ++j) {
        if (op->G0[i + D0 * j] != op->G0[j + D0 * i])
          return fail(EVR_ERR_ARG, "hermitian form needs a symmetric G0");
        if (op->Ga)
          for (int k = 0; k < D0; ++k) {
            const double a = op->Ga[i + D0 * (j + D0 * k)];
            if (a != op->Ga[j + D0 * (i + D0 * k)] || a != op->Ga[k + D0 * (j + D0 * i)])
              return fail(EVR_ERR_ARG, "hermitian form needs a fully symmetric Ga");
          }
      }
  }
  std::unique_ptr<evr_plan> p(new evr_plan());
  const int D = bd->D, nlev = bd->nlev;
  p->D = D;
  p->D_full = D;
  p->hermitian = op->hermitian != 0;
  p->type = op->type;
  p->max_nvec = max_nvec;
  p->dp_order = bd->grid_last_fastest != 0;
  p->nterms = bd->nterms;
  p->nb_packed = bd->nb_packed;
  HIP_TRY(hipGetDevice(&p->device));
  HIP_TRY(hipDeviceGetAttribute(&p->ncu, hipDeviceAttributeMultiprocessorCount, p->device));
  HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));

  // ---- slots: 1D matrices and basis-space KEO factors
  const int ns = D * nlev;
  std::vector<int64_t> ox, oB, oD;
  slot_offsets(bd, ox, oB, oD);
  std::vector<double> bw, keo, xs, bwdiag, d1, d2;
  std::vector<char> slot_fold(ns, 0);  // one function with vanishing P, Q, R
  p->slots.resize(ns);
  for (int s = 0; s < ns; ++s) {
    const int nb = bd->nb[s], nq = bd->nq[s];
    if (nb < 1 || nq < nb || nq > evr::kMaxNq) return fail(EVR_ERR_ARG, "each slot needs 1 <= nb <= nq <= 64");
    // above 31 functions (the term kernels' 5-bit basis digits) a term runs streamed
    if (nb > evr::kMaxNb && D > evr::kMaxDev)
      return fail(EVR_ERR_ARG, "with more than 12 modes each slot needs nb <= 31");
    evr::SlotDev &sd = p->slots[s];
    sd.nb = nb;
    sd.nq = nq;
    sd.offB = (int64_t)bw.size();
    bw.insert(bw.end(), bd->B + oB[s], bd->B + oB[s + 1]);
    sd.offW = (int64_t)bw.size();
    bw.insert(bw.end(), bd->W + oB[s], bd->W + oB[s + 1]);
    // pool for the plan diagonal: B, and W o B^T in place of W
    bwdiag.insert(bwdiag.end(), bd->B + oB[s], bd->B + oB[s + 1]);
    for (int q = 0; q < nq; ++q)
      for (int b = 0; b < nb; ++b)
        bwdiag.push_back(bd->W[oB[s] + b + (int64_t)nb * q] * bd->B[oB[s] + q + (int64_t)nq * b]);
    sd.offX = (int64_t)xs.size();
    xs.insert(xs.end(), bd->x + ox[s], bd->x + ox[s + 1]);
    sd.offD1 = (int64_t)d1.size();
    d1.insert(d1.end(), bd->D1 + oD[s], bd->D1 + oD[s + 1]);
    d2.insert(d2.end(), bd->D2 + oD[s], bd->D2 + oD[s + 1]);
    evr::Slot1D s1;
    s1.nb = nb;
    s1.nq = nq;
    s1.x.assign(bd->x + ox[s], bd->x + ox[s + 1]);
    s1.B.assign(bd->B + oB[s], bd->B + oB[s + 1]);
    s1.W.assign(bd->W + oB[s], bd->W + oB[s + 1]);
    s1.D1.assign(bd->D1 + oD[s], bd->D1 + oD[s + 1]);
    s1.D2.assign(bd->D2 + oD[s], bd->D2 + oD[s + 1]);
    std::vector<double> f;
    evr::keo_factors(s1, f);
    if (op->hermitian) evr::symmetrize_factors(nb, f);
    sd.offK = (int64_t)keo.size();
    keo.insert(keo.end(), f.begin(), f.end());
    if (nb == 1) {
      // <0|d|0>, <0|x|0> and <0|d x|0> are integrals of odd functions: zero on
      // a symmetric rule up to rounding; such a mode folds into scalars
      double scale = 1.0;
      for (int k : {evr::kT, evr::kS, evr::kU, evr::kT2}) scale = std::max(scale, std::fabs(f[k]));
      const double tiny = 1e-13 * scale;
      slot_fold[s] = std::fabs(f[evr::kP]) <= tiny && std::fabs(f[evr::kQ]) <= tiny &&
                     std::fabs(f[evr::kR]) <= tiny;
    }
  }

  // ---- operator coefficients in the kernel's (i, j, k) convention
  p->hasV = op->V != nullptr;
  p->hasKeo = op->type != 0;
  p->diagT2 = op->type == 1;
  std::vector<double> cC((size_t)D * D, 0.0), cA, cF1;
  if (op->type == 10) {
    // reference: phi_i = sum_j G_ji d_j psi; H = -1/2 sum_i d_i phi_i
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) cC[i * D + j] = -0.5 * op->G0[j + D * i];
    if (op->Ga) {
      p->hasA = true;
      cA.assign((size_t)D * D * D, 0.0);
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j)
          for (int k = 0; k < D; ++k) cA[(i * D + j) * D + k] = -0.5 * op->Ga[j + D * (i + D * k)];
    }
  } else if (op->type == 1) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) cC[i * D + j] = op->F2[i + D * j];
    if (op->F1 && !op->hermitian) {  // the symmetric part of F1_i d_i is zero
      p->hasF1 = true;
      cF1.assign(op->F1, op->F1 + D);
    }
  }

  // ---- kernel mode positions of every term.  Up to kMaxDev modes the
  // positions are the modes.  Beyond (SG4 bases of many modes at a low level,
  // where most modes of a term have one function on one grid point), a term
  // keeps its live modes in mode order, merges its other modes into one
  // composite slot and pads with a unit slot.  A merged mode contributes
  // scalars only: its B and W (1 x 1) multiply into the composite's, its
  // P, Q, R vanish (no cross terms), and its diagonal KEO c_dd Td_d sums into
  // the composite's Td with coefficient 1; the term grid and basis orders are
  // unchanged (a merged mode has extent 1).  The KEO coefficients then come
  // per position -> mode assignment (coefficient sets).
  const int K = evr::kMaxDev;
  const bool compact = D > K;
  int Dk = D;
  std::vector<int> tslot((size_t)p->nterms * K, 0), tmode((size_t)p->nterms * K, -2);
  if (!compact) {
    for (int64_t t = 0; t < p->nterms; ++t)
      for (int d = 0; d < D; ++d) {
        const int lv = bd->term_levels[t * D + d];
        if (lv < 0 || lv >= nlev) return fail(EVR_ERR_ARG, "term level out of range");
        tslot[t * K + d] = d * nlev + lv;
        tmode[t * K + d] = d;
      }
  } else {
    if (p->dp_order) return fail(EVR_ERR_ARG, "a direct-product basis needs D <= 12");
    if (p->hasA) return fail(EVR_ERR_ARG, "G(q) = G0 + Ga q needs D <= 12 (G0 alone is fine)");
    const int Td = p->diagT2 ? evr::kT2 : evr::kT;
    auto merged = [&](int s) { return bd->nb[s] == 1 && bd->nq[s] == 1 && slot_fold[s]; };
    Dk = 1;
    for (int64_t t = 0; t < p->nterms; ++t) {
      int live = 0, dead = 0;
      for (int d = 0; d < D; ++d) {
        const int lv = bd->term_levels[t * D + d];
        if (lv < 0 || lv >= nlev) return fail(EVR_ERR_ARG, "term level out of range");
        (merged(d * nlev + lv) ? dead : live)++;
      }
      Dk = std::max(Dk, live + (dead > 0));
    }
    if (Dk > K)
      return fail(EVR_ERR_SIZE, "with more than 12 modes a term may have at most 11 modes with "
                                "more than one function or grid point");
    // a slot of one function on one point: B, W, the W o B^T of the plan
    // diagonal, and the diagonal KEO factor Td (the other factors zero)
    auto add_slot = [&](double B, double W, double fTd) -> int {
      evr::SlotDev sd{};
      sd.nb = sd.nq = 1;
      sd.offB = (int64_t)bw.size();
      bw.push_back(B);
      bwdiag.push_back(B);
      sd.offW = (int64_t)bw.size();
      bw.push_back(W);
      bwdiag.push_back(W * B);
      sd.offX = (int64_t)xs.size();
      xs.push_back(0.0);
      sd.offD1 = (int64_t)d1.size();
      d1.push_back(0.0);
      d2.push_back(0.0);
      sd.offK = (int64_t)keo.size();
      keo.insert(keo.end(), evr::kNumFactors, 0.0);
      keo[sd.offK + Td] = fTd;
      p->slots.push_back(sd);
      slot_fold.push_back(1);
      return (int)p->slots.size() - 1;
    };
    const int unit = add_slot(1.0, 1.0, 0.0);
    std::map<std::vector<int>, int> composite;  // merged slots -> composite slot
    for (int64_t t = 0; t < p->nterms; ++t) {
      std::vector<int> dead;
      int k = 0;
      for (int d = 0; d < D; ++d) {
        const int s = d * nlev + bd->term_levels[t * D + d];
        if (merged(s)) {
          dead.push_back(s);
        } else {
          tslot[t * K + k] = s;
          tmode[t * K + k++] = d;
        }
      }
      if (!dead.empty()) {
        auto it = composite.find(dead);
        int cs;
        if (it != composite.end()) {
          cs = it->second;
        } else {
          double B = 1.0, W = 1.0, f = 0.0;
          for (int s : dead) {
            const evr::SlotDev &sd = p->slots[s];
            const int d = s / nlev;
            B *= bw[sd.offB];
            W *= bw[sd.offW];
            f += cC[(size_t)d * D + d] * keo[sd.offK + Td];
          }
          cs = add_slot(B, W, f);
          composite.emplace(dead, cs);
        }
        tslot[t * K + k] = cs;
        tmode[t * K + k++] = -1;
      }
      for (; k < Dk; ++k) tslot[t * K + k] = unit;  // tmode stays -2
    }
    p->D = Dk;
  }
  // coefficient sets: one per distinct position -> mode assignment
  std::map<std::vector<int>, int> set_of;
  std::vector<std::vector<int>> sets;
  std::vector<int> cix(p->nterms, 0);
  for (int64_t t = 0; t < p->nterms; ++t) {
    std::vector<int> key(tmode.begin() + t * K, tmode.begin() + t * K + Dk);
    auto it = set_of.find(key);
    if (it == set_of.end()) {
      it = set_of.emplace(key, (int)sets.size()).first;
      sets.push_back(key);
    }
    cix[t] = it->second;
  }

  // ---- terms
  p->terms.resize(p->nterms);
  int64_t onb = 0, onq = 0, onvq = 0;
  std::vector<int> cls(p->nterms);
  p->keo_layout = p->hasKeo;
  for (int64_t t = 0; t < p->nterms; ++t) {
    evr::TermDev &td = p->terms[t];
    int nbd[evr::kMaxDev], nqd[evr::kMaxDev];
    int64_t nbt = 1, nqt = 1;
    int mats = 0;
    const int D = Dk;  // the kernel's positions from here on
    td.cix = cix[t];
    for (int d = 0; d < D; ++d) {
      const int s = tslot[t * K + d];
      td.slot[d] = s;
      nbd[d] = p->slots[s].nb;
      nqd[d] = p->slots[s].nq;
      nbt *= nbd[d];
      nqt *= nqd[d];
      mats += 2 * nbd[d] * nqd[d] + (p->hasKeo ? evr::kKeoFactors * nbd[d] * nbd[d] : 0);
    }
    if (nqt > INT32_MAX) return fail(EVR_ERR_SIZE, "term grid exceeds int32");
    td.off_nb = onb;
    td.off_nq = onq;
    td.nbt = (int)nbt;
    td.nqt = (int)nqt;
    td.w = bd->weights[t];
    td.lds_mats = mats + (mats & 1);  // keep the data arrays 16-byte aligned
    int c = 0;
    while (c < kNumClasses && kClasses[c].max_nbt() < nbt) ++c;
    bool wide = false;  // a slot beyond the term kernels' 5-bit basis digits
    for (int d = 0; d < D; ++d) wide = wide || nbd[d] > evr::kMaxNb;
    if (c == kNumClasses || wide) {
      // beyond one workgroup: the streamed passes over global memory
      if (compact) return fail(EVR_ERR_SIZE, "with more than 12 modes a term may have at most 8192 functions");
      c = -1;
      p->big.push_back(t);
      p->big_nq = std::max<int64_t>(p->big_nq, nqt);
    }
    cls[t] = c;
    // mode lists: all modes (basis <-> grid transforms), the modes with more
    // than one function (the H|psi> V part, one-function modes folded into
    // V_eff), and the modes whose P, Q, R do not vanish (the kinetic part)
    td.Dv = 0;
    td.na = 0;
    int64_t qv = 1;
    for (int d = 0; d < D; ++d) {
      if (nbd[d] > 1) {
        td.vmode[td.Dv++] = d;
        qv *= nqd[d];
      }
      if (!slot_fold[td.slot[d]]) td.amode[td.na++] = d;
    }
    for (int d = td.Dv; d < evr::kMaxDev; ++d) td.vmode[d] = 0;
    for (int d = td.na; d < evr::kMaxDev; ++d) td.amode[d] = 0;
    td.off_vq = onvq;
    td.off_asm = p->asm_elems;
    p->asm_elems += (int64_t)((nbt + 15) / 16) * evr::asm_ksteps(nbt) * 64;
    p->asm_flops_per_vec += 2.0 * (double)nbt * (double)nbt;
    onb += nbt;
    onq += nqt;
    onvq += qv;
    p->max_nb = std::max(p->max_nb, nbt);
    p->max_nq = std::max(p->max_nq, nqt);
  }
  p->sum_nb = onb;
  p->sum_nq = onq;
  p->sum_vq = onvq;
  if (onb > INT32_MAX) return fail(EVR_ERR_SIZE, "sum of term basis sizes exceeds int32");

  // ---- a term whose slabs do not fit one workgroup's LDS in its class (a
  // long grid next to many functions) runs streamed as well
  {
    std::vector<evr::StepDev> probe;
    for (int64_t t = 0; t < p->nterms; ++t) {
      if (cls[t] < 0) continue;
      size_t words = 0;
      probe.clear();
      evr::TermDev td = p->terms[t];
      if (layout_term(p.get(), td, cls[t], probe, words)) continue;
      if (compact) return fail(EVR_ERR_SIZE, "term does not fit in LDS");
      cls[t] = -1;
      p->big.push_back(t);
      p->big_nq = std::max<int64_t>(p->big_nq, td.nqt);
    }
    std::sort(p->big.begin(), p->big.end());
  }

  // ---- slab layouts for the classes (the static rule: the smallest class
  // that holds the term) and the per-class work lists
  p->tclass = cls;
  p->hC = cC;
  p->hF1 = cF1;
  if (p->big.empty() && p->nterms == 1 && !compact && p->terms[0].nbt >= kStreamOneMin) {
    p->one.push_back(0);
    p->big_nq = p->terms[0].nqt;
  }
  if (!p->big.empty() || !p->one.empty()) {
    // vectors per streamed pass: the (7 + D) work tensors within 4 GiB
    const double per_vec = 8.0 * (7 + Dk) * (double)p->big_nq;
    p->big_k = (int)std::max<int64_t>(1, std::min<int64_t>(max_nvec, (int64_t)(4294967296.0 / per_vec)));
    HIP_TRY(p->d_big.alloc((size_t)(7 + Dk) * p->big_nq * p->big_k));
  }
  {
    const int rc = apply_classes(p.get());
    if (rc) return rc;
  }

  // ---- deterministic scatter: the term rows of packed function pk, in
  // increasing term order, are rows cnt[pk] .. cnt[pk+1]-1 of R; rowpos maps a
  // term row to that position (-1 for pruned functions)
  std::vector<int> cnt(p->nb_packed + 1, 0), rowpos(onb, -1);
  for (int64_t r = 0; r < onb; ++r) {
    const int mp = bd->map[r];
    if (mp >= p->nb_packed) return fail(EVR_ERR_ARG, "map entry beyond nb_packed");
    if (mp >= 0) cnt[mp + 1]++;
  }
  for (int64_t k = 0; k < p->nb_packed; ++k) cnt[k + 1] += cnt[k];
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t r = 0; r < onb; ++r) {
      const int mp = bd->map[r];
      if (mp >= 0) rowpos[r] = fill[mp]++;
    }
  }

  // ---- DP grid permutation (reference DP order <-> term grid order)
  if (p->dp_order) {
    p->dp_perm.resize(onq);
    int n[evr::kMaxDev] = {0};
    for (int64_t i = 0; i < onq; ++i) {
      int64_t g = 0;
      for (int d = 0; d < D; ++d) g = g * bd->nq[d * nlev] + n[d];
      p->dp_perm[i] = (int)g;
      for (int d = 0; d < D; ++d) {
        if (++n[d] < bd->nq[d * nlev]) break;
        n[d] = 0;
      }
    }
  }

  // ---- uploads
  std::vector<float> bw32(bw.begin(), bw.end()), keo32(keo.begin(), keo.end());
  HIP_TRY(p->d_slots.upload(p->slots.data(), p->slots.size(), p->stream));
  HIP_TRY(p->d_bw64.upload(bw.data(), bw.size(), p->stream));
  HIP_TRY(p->d_bwdiag.upload(bwdiag.data(), bwdiag.size(), p->stream));
  HIP_TRY(p->d_keo64.upload(keo.data(), keo.size(), p->stream));
  HIP_TRY(p->d_bw32.upload(bw32.data(), bw32.size(), p->stream));
  HIP_TRY(p->d_keo32.upload(keo32.data(), keo32.size(), p->stream));
  HIP_TRY(p->d_x.upload(xs.data(), xs.size(), p->stream));
  HIP_TRY(p->d_D1.upload(d1.data(), d1.size(), p->stream));
  HIP_TRY(p->d_D2.upload(d2.data(), d2.size(), p->stream));
  HIP_TRY(p->d_map.upload(bd->map, onb, p->stream));
  if (p->dp_order) HIP_TRY(p->d_dp_perm.upload(p->dp_perm.data(), p->dp_perm.size(), p->stream));
  {
    std::vector<int> seg, fseg(p->nb_packed + 1);
    seg.reserve(p->nb_packed + p->nb_packed / 8 + 1);
    for (int64_t k = 0; k < p->nb_packed; ++k) {
      fseg[k] = (int)seg.size();
      for (int r = cnt[k]; r < cnt[k + 1]; r += evr::kReduceSeg) seg.push_back(r);
    }
    fseg[p->nb_packed] = (int)seg.size();
    p->nseg = (int64_t)seg.size();
    p->nrows = cnt[p->nb_packed];
    seg.push_back(cnt[p->nb_packed]);
    HIP_TRY(p->d_seg.upload(seg.data(), seg.size(), p->stream));
    HIP_TRY(p->d_fseg.upload(fseg.data(), fseg.size(), p->stream));
    HIP_TRY(p->d_P.alloc((size_t)std::max<int64_t>(p->nseg, 1) * max_nvec));
  }
  HIP_TRY(p->d_rowpos.upload(rowpos.data(), rowpos.size(), p->stream));
  {
    // per coefficient set (positions a, b with modes m[a], m[b]): c_ab from
    // the operator, 1 on the diagonal of a composite position, 0 for padding;
    // Ga only without compaction (positions = modes).  Kernel layout of the
    // pre-zeroed strips of one set:
    //  S[j*Dp+i] = a_ijj, N[j*Dp+i] = a_iji, O[k*Dp+i] = a_iik, C[j*Dp+i] = c_ij
    //  (i != j); Cd[i] = c_ii; U[i] = a_iii; F[i] = F1_i;
    //  A[(j*Dp+k)*Dp+i] = a_ijk for i, j, k distinct (with Ga only)
    const int Dp = evr::kDp;
    const size_t nset = sets.size();
    auto cpos = [&](const std::vector<int> &m, int a, int b) -> double {
      if (m[a] >= 0 && m[b] >= 0) return cC[(size_t)m[a] * D + m[b]];
      return (a == b && m[a] == -1) ? 1.0 : 0.0;
    };
    auto apos = [&](int i, int j, int k) { return cA.empty() ? 0.0 : cA[((size_t)i * D + j) * D + k]; };
    auto fpos = [&](const std::vector<int> &m, int a) {
      return (!cF1.empty() && m[a] >= 0) ? cF1[m[a]] : 0.0;
    };
    p->keoc_stride = 4 * Dp * Dp + 3 * Dp + (cA.empty() ? 0 : Dp * Dp * Dp);
    std::vector<double> kc(nset * p->keoc_stride, 0.0), pc(nset * Dk * Dk),
        pf(cF1.empty() ? 0 : nset * Dk);
    for (size_t c = 0; c < nset; ++c) {
      const std::vector<int> &m = sets[c];
      double *kS = kc.data() + c * p->keoc_stride, *kN = kS + Dp * Dp, *kO = kN + Dp * Dp,
             *kC = kO + Dp * Dp, *kCd = kC + Dp * Dp, *kU = kCd + Dp, *kF = kU + Dp, *kA = kF + Dp;
      for (int i = 0; i < Dk; ++i) {
        for (int j = 0; j < Dk; ++j) pc[(c * Dk + i) * Dk + j] = cpos(m, i, j);
        if (!cF1.empty()) pf[c * Dk + i] = fpos(m, i);
        kCd[i] = cpos(m, i, i);
        kU[i] = apos(i, i, i);
        kF[i] = fpos(m, i);
        for (int j = 0; j < Dk; ++j) {
          if (j == i) continue;
          kC[j * Dp + i] = cpos(m, i, j);
          kS[j * Dp + i] = apos(i, j, j);
          kN[j * Dp + i] = apos(i, j, i);
          kO[j * Dp + i] = apos(i, i, j);
          if (!cA.empty())
            for (int k = 0; k < Dk; ++k)
              if (k != i && k != j) kA[(j * Dp + k) * Dp + i] = apos(i, j, k);
        }
      }
    }
    HIP_TRY(p->d_coefC.upload(pc.data(), pc.size(), p->stream));
    if (!cA.empty()) HIP_TRY(p->d_coefA.upload(cA.data(), cA.size(), p->stream));
    if (!pf.empty()) HIP_TRY(p->d_coefF1.upload(pf.data(), pf.size(), p->stream));
    HIP_TRY(p->d_keoc.upload(kc.data(), kc.size(), p->stream));
  }
  HIP_TRY(p->d_V.alloc(onq));
  HIP_TRY(p->d_Veff.alloc(onvq));
  HIP_TRY(p->d_X.alloc((size_t)p->nb_packed * max_nvec));
  HIP_TRY(p->d_R.alloc((size_t)onb * max_nvec));
  HIP_TRY(p->d_psi.alloc((size_t)p->nb_packed * max_nvec));
  HIP_TRY(p->d_hpsi.alloc((size_t)p->nb_packed * max_nvec));
  if (op->V) {
    int rc = evr_plan_set_potential(p.get(), op->V);
    if (rc) return rc;
  } else {
    HIP_TRY(zero_fill(p->d_V.ptr, onq, p->stream));
    HIP_TRY(zero_fill(p->d_Veff.ptr, onvq, p->stream));
  }
  HIP_TRY(hipStreamSynchronize(p->stream));
  *out = p.release();
  return EVR_OK;
}

int evr_plan_destroy(evr_plan *p)
{
  delete p;  // ~evr_plan releases the device buffers, events and stream
  return EVR_OK;
}

int evr_plan_set_potential(evr_plan *p, const double *V)
{
  if (!p || !V) return fail(EVR_ERR_ARG, "NULL argument");
  std::vector<double> tmp;
  const double *src = V;
  if (p->dp_order) {
    tmp.resize(p->sum_nq);
    for (int64_t i = 0; i < p->sum_nq; ++i) tmp[i] = V[p->dp_perm[i]];
    src = tmp.data();
  }
  HIP_TRY(hipMemcpyAsync(p->d_V.ptr, src, sizeof(double) * p->sum_nq, hipMemcpyHostToDevice,
                         p->stream));
  const int rc = fold_v(p);
  if (rc) return rc;
  // an assembled operator holds the old V: drop it (evr_plan_assemble again)
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->d_A64.release();
  p->d_A32.release();
  HIP_TRY(hipStreamSynchronize(p->stream));
  const bool had = p->hasV;
  p->hasV = true;
  p->V32_valid = false;
  if (!had) recount_flops(p);
  return EVR_OK;
}

int evr_plan_gen_potential_poly(evr_plan *p, const double *v2, const double *c3,
                                const double *d22, const double *e4)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  if (p->D != p->D_full)
    return fail(EVR_ERR_ARG, "the polynomial generator needs D <= 12 (set V on the grid instead)");
  const int D = p->D;
  std::vector<double> coef((size_t)D + D * D * D + D * D + D, 0.0);
  if (v2) std::copy(v2, v2 + D, coef.begin());
  if (c3) std::copy(c3, c3 + D * D * D, coef.begin() + D);
  if (d22) std::copy(d22, d22 + D * D, coef.begin() + D + D * D * D);
  if (e4) std::copy(e4, e4 + D, coef.begin() + D + D * D * D + D * D);
  DevBuf<double> dcoef;
  HIP_TRY(dcoef.upload(coef.data(), coef.size(), p->stream));
  hipLaunchKernelGGL(evr::k_gen_poly, dim3((unsigned)p->nterms), dim3(256), 0, p->stream, D,
                     p->d_terms.ptr, p->d_slots.ptr, p->d_x.ptr, dcoef.ptr, p->d_V.ptr);
  HIP_TRY(hipGetLastError());
  const int rc = fold_v(p);
  if (rc) return rc;
  // an assembled operator holds the old V: drop it (evr_plan_assemble again)
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->d_A64.release();
  p->d_A32.release();
  HIP_TRY(hipStreamSynchronize(p->stream));
  dcoef.release();
  p->hasV = true;
  p->V32_valid = false;
  recount_flops(p);
  return EVR_OK;
}

// quadrature weight of every term-grid point: W[0, q] / B[q, 0] per 1D rule
// (W = B^T diag(w)), multiplied over the modes of the term (first mode fastest)
static int quad_weights(evr_plan *p, std::vector<double> &wq)
{
  std::vector<double> bw(p->d_bw64.n);
  HIP_TRY(hipMemcpy(bw.data(), p->d_bw64.ptr, sizeof(double) * bw.size(), hipMemcpyDeviceToHost));
  wq.assign(p->sum_nq, 0.0);
  for (int64_t t = 0; t < p->nterms; ++t) {
    const evr::TermDev &td = p->terms[t];
    for (int64_t q = 0; q < td.nqt; ++q) {
      int64_t rem = q;
      double w = 1.0;
      for (int d = 0; d < p->D; ++d) {
        const evr::SlotDev &sd = p->slots[td.slot[d]];
        const int64_t iq = rem % sd.nq;
        rem /= sd.nq;
        w *= bw[sd.offW + (int64_t)sd.nb * iq] / bw[sd.offB + iq];
      }
      wq[td.off_nq + q] = w;
    }
  }
  return EVR_OK;
}

int evr_plan_set_keo_grid(evr_plan *p, const double *G, const double *Jac, const double *sqRhoOVERJac)
{
  if (!p || !G) return fail(EVR_ERR_ARG, "NULL argument");
  if (p->type != 10) return fail(EVR_ERR_ARG, "G per grid point is a type-10 operator");
  if (p->D != p->D_full) return fail(EVR_ERR_ARG, "G per grid point needs D <= 12");
  if (p->dp_order) return fail(EVR_ERR_ARG, "G per grid point is for SG4 bases");
  if (!p->big.empty()) return fail(EVR_ERR_ARG, "G per grid point needs terms of at most 8192 functions");
  const int D = p->D;
  const int64_t nq = p->sum_nq;
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->d_A64.release();
  p->d_A32.release();
  HIP_TRY(p->d_GG.upload(G, (size_t)D * D * nq, p->stream));
  std::vector<double> one;
  if (!Jac || !sqRhoOVERJac) one.assign(nq, 1.0);
  HIP_TRY(p->d_Jg.upload(Jac ? Jac : one.data(), nq, p->stream));
  HIP_TRY(p->d_Sg.upload(sqRhoOVERJac ? sqRhoOVERJac : one.data(), nq, p->stream));
  if (p->hermitian) {
    // w / (J s) per point for the transposed sequence
    std::vector<double> wq, as(nq);
    int rcw = quad_weights(p, wq);
    if (rcw) return rcw;
    for (int64_t g = 0; g < nq; ++g)
      as[g] = wq[g] / ((Jac ? Jac[g] : 1.0) * (sqRhoOVERJac ? sqRhoOVERJac[g] : 1.0));
    HIP_TRY(p->d_Wq.upload(wq.data(), nq, p->stream));
    HIP_TRY(p->d_As.upload(as.data(), nq, p->stream));
    if (!p->d_R2.ptr) HIP_TRY(p->d_R2.alloc((size_t)p->nrows * std::max(1, p->max_nvec)));
  }
  int rc = grid_form_buffers(p);
  if (rc) return rc;
  // the term kernels keep the B->G and G->B transforms; the operator itself
  // is the grid form
  p->hasKeo = false;
  p->hasA = false;
  p->hasF1 = false;
  p->keo_grid = true;
  p->grid_type = 10;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

int evr_plan_set_type1_grid(evr_plan *p, const double *F2, const double *F1)
{
  if (!p || !F2) return fail(EVR_ERR_ARG, "NULL argument");
  if (p->type != 1) return fail(EVR_ERR_ARG, "operator grids F2, F1 per point belong to a type-1 plan");
  if (p->D != p->D_full) return fail(EVR_ERR_ARG, "operator grids per point need D <= 12");
  if (p->dp_order) return fail(EVR_ERR_ARG, "operator grids per point are for SG4 bases");
  if (!p->big.empty()) return fail(EVR_ERR_ARG, "operator grids per point need terms of at most 8192 functions");
  const int D = p->D;
  const int64_t nq = p->sum_nq;
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->d_A64.release();
  p->d_A32.release();
  HIP_TRY(p->d_F2g.upload(F2, (size_t)D * D * nq, p->stream));
  if (F1) {
    HIP_TRY(p->d_F1g.upload(F1, (size_t)D * nq, p->stream));
  } else {
    p->d_F1g.release();
  }
  if (p->hermitian) {
    // the weighted tables w F2, w F1 of the transposed sequence
    std::vector<double> wq;
    int rcw = quad_weights(p, wq);
    if (rcw) return rcw;
    std::vector<double> f2w((size_t)D * D * nq), f1w(F1 ? (size_t)D * nq : 0);
    for (size_t c = 0; c < (size_t)D * D; ++c)
      for (int64_t g = 0; g < nq; ++g) f2w[c * nq + g] = wq[g] * F2[c * nq + g];
    for (size_t c = 0; c < f1w.size() / std::max<int64_t>(nq, 1); ++c)
      for (int64_t g = 0; g < nq; ++g) f1w[c * nq + g] = wq[g] * F1[c * nq + g];
    HIP_TRY(p->d_Wq.upload(wq.data(), nq, p->stream));
    HIP_TRY(p->d_F2w.upload(f2w.data(), f2w.size(), p->stream));
    if (F1) {
      HIP_TRY(p->d_F1w.upload(f1w.data(), f1w.size(), p->stream));
    } else {
      p->d_F1w.release();
    }
    if (!p->d_R2.ptr) HIP_TRY(p->d_R2.alloc((size_t)p->nrows * std::max(1, p->max_nvec)));
    HIP_TRY(hipStreamSynchronize(p->stream));  // the host tables above are freed on return
  }
  int rc = grid_form_buffers(p);
  if (rc) return rc;
  p->hasKeo = false;
  p->hasA = false;
  p->hasF1 = false;
  p->keo_grid = true;
  p->grid_type = 1;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

int evr_plan_sizes(const evr_plan *p, int64_t *out)
{
  if (!p || !out) return fail(EVR_ERR_ARG, "NULL argument");
  out[0] = p->nb_packed; out[1] = p->nterms; out[2] = p->sum_nb; out[3] = p->sum_nq;
  out[4] = p->max_nb; out[5] = p->max_nq; out[6] = p->D_full;
  return EVR_OK;
}

void *evr_plan_stream(evr_plan *p) { return p ? (void *)p->stream : nullptr; }

int evr_plan_max_nvec(const evr_plan *p) { return p ? p->max_nvec : fail(EVR_ERR_ARG, "NULL plan"); }

int evr_plan_assemble(evr_plan *p, int which)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  if (which < 1 || which > (EVR_ASM_FP64 | EVR_ASM_FP32))
    return fail(EVR_ERR_ARG, "which must be EVR_ASM_FP64 and/or EVR_ASM_FP32");
  if (p->keo_grid)
    return fail(EVR_ERR_ARG, "a plan with operator grids per point runs the grid form (not assembled)");
  if (!p->big.empty())
    return fail(EVR_ERR_ARG, "terms above 8192 functions run streamed over global memory (not assembled)");
  const size_t n = (size_t)p->asm_elems;
  if ((which & EVR_ASM_FP64) && !p->d_A64.ptr) {
    if (p->d_A64.alloc(n) != hipSuccess) {
      p->d_A64.ptr = nullptr;
      return fail(EVR_ERR_SIZE, "not enough device memory for the FP64 assembled operator (" +
                                    std::to_string(8.0 * n / 1e9) + " GB)");
    }
    HIP_TRY(zero_fill(p->d_A64.ptr, (int64_t)n, p->stream));
  }
  if ((which & EVR_ASM_FP32) && !p->d_A32.ptr) {
    if (p->d_A32.alloc(n) != hipSuccess) {
      p->d_A32.ptr = nullptr;
      return fail(EVR_ERR_SIZE, "not enough device memory for the FP32 assembled operator (" +
                                    std::to_string(4.0 * n / 1e9) + " GB)");
    }
    HIP_TRY(zero_fill(p->d_A32.ptr, (int64_t)n, p->stream));
  }
  // columns: the FP64 term kernel applied to each term's own unit vectors
  for (int c = 0; c < kNumClasses; ++c) {
    std::vector<int2> items;
    for (int64_t t = 0; t < p->nterms; ++t) {
      const int nbt = p->terms[t].nbt;
      if (p->tclass[t] != c) continue;
      for (int c0 = 0; c0 < nbt; c0 += kClasses[c].nvw) items.push_back(int2{(int)t, c0});
    }
    if (items.empty()) continue;
    DevBuf<int2> ditems;
    HIP_TRY(ditems.upload(items.data(), items.size(), p->stream));
    evr::TermParams<double> prm = base_params<double>(p, evr::kModeAssemble, 1);
    prm.bw = p->d_bw64.ptr;
    prm.keo = p->d_keo64.ptr;
    prm.V = p->d_Veff.ptr;
    prm.items = ditems.ptr;
    prm.A64 = (which & EVR_ASM_FP64) ? p->d_A64.ptr : nullptr;
    prm.A32 = (which & EVR_ASM_FP32) ? p->d_A32.ptr : nullptr;
    HIP_TRY(launch_items(c, prm, (int64_t)items.size(), p->lds_bytes64[c], p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  if (!p->d_asm_items.ptr) {
    std::vector<int> nbt(p->nterms);
    for (int64_t t = 0; t < p->nterms; ++t) nbt[t] = p->terms[t].nbt;
    const std::vector<int2> items = asm_items(nbt);
    if ((int64_t)items.size() > (int64_t)0xFFFFFFFFu / 256)
      return fail(EVR_ERR_SIZE, "too many assembled work items for one launch");
    HIP_TRY(p->d_asm_items.upload(items.data(), items.size(), p->stream));
    p->n_asm_items = (int64_t)items.size();
  }
  p->asm_bytes64 = 8.0 * (double)n;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

int evr_plan_assembled(const evr_plan *p, double *bytes)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  int m = 0;
  double b = 0.0;
  if (p->d_A64.ptr) { m |= EVR_ASM_FP64; b += 8.0 * p->asm_elems; }
  if (p->d_A32.ptr) { m |= EVR_ASM_FP32; b += 4.0 * p->asm_elems; }
  if (bytes) *bytes = b;
  return m;
}

int evr_plan_release_assembled(evr_plan *p)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  (void)hipStreamSynchronize(p->stream);
  p->d_A64.release();
  p->d_A32.release();
  return EVR_OK;
}

static void recount_flops(evr_plan *p)
{
  for (int c = 0; c < kNumClasses; ++c) p->flops_class[c] = 0;
  for (int64_t t = 0; t < p->nterms; ++t)
    if (p->tclass[t] >= 0) p->flops_class[p->tclass[t]] += term_flops(p, p->terms[t]);
}

int evr_plan_autotune(evr_plan *p, int nvec, int flags, int reps, int *chosen, double *ms_static,
                      double *ms_tuned)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  if (nvec < 1 || nvec > p->max_nvec) return fail(EVR_ERR_ARG, "nvec must be in 1..max_nvec");
  if (p->keo_grid) return fail(EVR_ERR_ARG, "plans with operator grids per point are not tuned");
  if (reps < 1) reps = 3;
  const bool fp32 = (flags & EVR_FP32) != 0;
  const int64_t n = p->nb_packed * nvec;
  DevBuf<double> X, Y;
  HIP_TRY(X.alloc((size_t)n));
  HIP_TRY(Y.alloc((size_t)n));
  hipLaunchKernelGGL(k_unit_cols, dim3(blocks_for(n)), dim3(256), 0, p->stream, p->nb_packed, nvec, (int64_t)0,
                     X.ptr);
  HIP_TRY(hipGetLastError());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  struct EvFree {
    hipEvent_t a, b;
    ~EvFree()
    {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } evfree{e0, e1};
  const bool prof = p->profiling;
  p->profiling = false;
  const int fl = EVR_DEVICE_PTRS | EVR_MATRIX_FREE | (fp32 ? EVR_FP32 : EVR_FP64);
  // median of reps timed sum-factorised H|psi> after one warm-up call
  auto time_layout = [&](double &ms) -> int {
    int rc = evr_oppsi(p, nvec, X.ptr, Y.ptr, fl);
    if (rc) return rc;
    std::vector<double> t(reps);
    for (int r = 0; r < reps; ++r) {
      HIP_TRY(hipEventRecord(e0, p->stream));
      rc = evr_oppsi(p, nvec, X.ptr, Y.ptr, fl);
      if (rc) return rc;
      HIP_TRY(hipEventRecord(e1, p->stream));
      HIP_TRY(hipEventSynchronize(e1));
      float f = 0.f;
      HIP_TRY(hipEventElapsedTime(&f, e0, e1));
      t[r] = f;
    }
    std::sort(t.begin(), t.end());
    ms = t[reps / 2];
    return EVR_OK;
  };
  // buckets: terms that share a class under the static rule (smallest class
  // that holds nb_t); each bucket in turn tries the larger classes that hold
  // it, the other buckets keeping their current choice
  std::vector<int> nat(p->nterms);
  for (int64_t t = 0; t < p->nterms; ++t) {
    int c = 0;
    while (c < kNumClasses && kClasses[c].max_nbt() < p->terms[t].nbt) ++c;
    nat[t] = c < kNumClasses && p->tclass[t] >= 0 ? c : -1;  // -1: streamed, not tuned
  }
  p->tclass = nat;
  int rc = apply_classes(p);
  if (rc) return rc;
  double best = 0.0;
  rc = time_layout(best);
  if (rc) return rc;
  const double t_static = best;
  int pick[kNumClasses];
  for (int b = 0; b < kNumClasses; ++b) pick[b] = b;
  for (int b = 0; b < kNumClasses; ++b) {
    bool present = false;
    for (int64_t t = 0; t < p->nterms && !present; ++t) present = nat[t] == b;
    if (!present) continue;
    for (int c = b + 1; c < kNumClasses && c <= b + 2; ++c) {
      for (int64_t t = 0; t < p->nterms; ++t)
        if (nat[t] == b) p->tclass[t] = c;
      double ms = 0.0;
      rc = apply_classes(p);
      if (rc == EVR_OK) rc = time_layout(ms);
      if (rc == EVR_ERR_SIZE) {  // this class cannot hold the bucket's slabs
        ms = 1e300;
        rc = EVR_OK;
      }
      if (rc) return rc;
      // keep a larger class only when it is faster by more than the noise
      if (ms < 0.97 * best) {
        best = ms;
        pick[b] = c;
      } else {
        break;  // a slower class up: larger ones are not tried (the reference's thread scan stops likewise)
      }
    }
    for (int64_t t = 0; t < p->nterms; ++t)
      if (nat[t] == b) p->tclass[t] = pick[b];
    rc = apply_classes(p);
    if (rc) return rc;
  }
  rc = time_layout(best);
  if (rc) return rc;
  recount_flops(p);
  p->profiling = prof;
  if (chosen)
    for (int b = 0; b < kNumClasses; ++b) chosen[b] = pick[b];
  if (ms_static) *ms_static = t_static;
  if (ms_tuned) *ms_tuned = best;
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

int evr_plan_set_profiling(evr_plan *p, int on)
{
  if (!p) return fail(EVR_ERR_ARG, "NULL plan");
  if (on && !p->ev[0])
    for (auto &e : p->ev) HIP_TRY(hipEventCreate(&e));
  p->profiling = on != 0;
  for (int c = 0; c <= kNumClasses; ++c) p->ms_acc[c] = p->flops_acc[c] = 0, p->launches[c] = 0;
  return EVR_OK;
}

int evr_plan_kernel_stats(evr_plan *p, int max, char *names, double *ms, double *flops,
                          int64_t *launches, int *n)
{
  if (!p || !n) return fail(EVR_ERR_ARG, "NULL argument");
  int k = 0;
  for (int c = 0; c < kNumClasses && k < max; ++c) {
    if (p->nunits[c] == 0) continue;
    if (names)
      snprintf(names + 64 * k, 64, "k_term<%d,%d,%d>", kClasses[c].nt, kClasses[c].ept,
               kClasses[c].nvw);
    if (ms) ms[k] = p->ms_acc[c];
    if (flops) flops[k] = p->flops_acc[c];
    if (launches) launches[k] = p->launches[c];
    ++k;
  }
  if (p->launches[kNumClasses] > 0 && k < max) {
    if (names) snprintf(names + 64 * k, 64, "k_apply_asm");
    if (ms) ms[k] = p->ms_acc[kNumClasses];
    if (flops) flops[k] = p->flops_acc[kNumClasses];
    if (launches) launches[k] = p->launches[kNumClasses];
    ++k;
  }
  *n = k;
  return EVR_OK;
}

static int ensure_v32(evr_plan *p)
{
  if (p->V32_valid) return EVR_OK;
  if (!p->d_Veff32.ptr) HIP_TRY(p->d_Veff32.alloc(std::max<int64_t>(p->sum_vq, 1)));
  hipLaunchKernelGGL(evr::k_to_float, dim3(blocks_for(p->sum_vq)), dim3(256), 0, p->stream,
                     p->d_Veff.ptr, p->d_Veff32.ptr, p->sum_vq);
  HIP_TRY(hipGetLastError());
  p->V32_valid = true;
  return EVR_OK;
}

// The assembled operator streams its whole pool once per call whatever the
// number of vectors (about 6 TB/s on MI355X); the sum-factorised term kernels
// cost per vector.  They are bound by latency and barrier waits, so a class
// takes its number of waves of workgroups (units x ceil(nvec / nvw) over the
// 1024 / nt workgroups each CU holds, fewer when LDS is short) times one
// workgroup's time, and a workgroup always does nvw vectors' work, padded or
// not.  The per-workgroup rates below are the classes' whole-device TFLOP/s
// measured at D=9 SG4 L=8 with 32 vectors, wave-count corrected; the model
// then gives 8.7 / 10.4 / 17.0 / 31.2 / 119.7 ms for 1 / 2 / 4 / 8 / 32
// vectors there (measured 9.1 / 10.8 / 17.2 / 31.6 / 120.5) and overestimates
// D=12 SG4 L=6 by about 30%.  With EVR_AUTO_PATH the cheaper path by this
// model is taken: at D=9 L=8 FP64 the sum-factorised one up to 5 vectors (17.3
// against 22.5 ms at 4, 23.3 against 22.8 at 5), at D=12 L=6 the assembled
// one from 1 vector (2.49 against 2.79 ms).
static bool matrix_free_cheaper(const evr_plan *p, int nvec, bool fp32)
{
  constexpr double kClassTflops[] = {1.23, 1.82, 2.56, 2.88, 3.83, 5.17, 4.70, 1.99};
  static_assert(sizeof(kClassTflops) / sizeof(double) == kNumClasses, "one rate per term class");
  double t_mf = 8e-6 * 2;  // gather-free pack + reduction
  for (int c = 0; c < kNumClasses; ++c) {
    if (!p->nterms_class[c]) continue;
    const KClass &k = kClasses[c];
    const size_t lds = fp32 ? p->lds_bytes32[c] : p->lds_bytes64[c];
    int per_cu = 1024 / k.nt;
    if (lds) per_cu = std::max(1, std::min<int>(per_cu, (int)(160 * 1024 / lds)));
    const double cap = (double)p->ncu * per_cu;
    const double nblk = (double)p->nunits[c] * ((nvec + k.nvw - 1) / k.nvw);
    const double wg_flops = p->flops_class[c] / (double)p->nterms_class[c] * k.nvw;
    const double rate = kClassTflops[c] * 1e12 * (fp32 ? 1.06 : 1.0);
    t_mf += std::ceil(nblk / cap) * cap * wg_flops / rate + 8e-6;
  }
  const double t_asm = (fp32 ? 0.5 : 1.0) * p->asm_bytes64 / 6.0e12 + 8e-6 * 3;
  return t_mf < t_asm;
}

}  // extern "C"

// H|psi> of nvec device columns on the plan stream with the work buffers
// given: X the packed block (nb_packed x nvec, FP64 words; FP32 uses the
// first half), R the term rows in scatter order (nrows x nvec words), P the
// segment sums (nseg x nvec doubles).  Every word of X, R and P that a call
// reads is written earlier in the same call.
static int oppsi_core(evr_plan *p, int nvec, const double *dpsi, double *dh, int flags, double *Xbuf,
                      double *Rbuf, double *Pbuf)
{
  const bool fp32 = (flags & EVR_FP32) != 0;
  const int64_t n = p->nb_packed * nvec;
  int rc;
  const bool use_asm = !(flags & EVR_MATRIX_FREE) &&
                       (fp32 ? p->d_A32.ptr != nullptr : p->d_A64.ptr != nullptr) &&
                       !((flags & EVR_AUTO_PATH) && matrix_free_cheaper(p, nvec, fp32));
  if (use_asm) {
    if (fp32) {
      float *X = reinterpret_cast<float *>(Xbuf), *R = reinterpret_cast<float *>(Rbuf);
      hipLaunchKernelGGL(evr::k_pack_in<float>, dim3(blocks_for(n)), dim3(256), 0, p->stream, dpsi,
                         p->nb_packed, nvec, X);
      HIP_TRY(hipGetLastError());
      HIP_TRY(apply_assembled<float>(p, p->d_A32.ptr, X, nvec, R));
      HIP_TRY(launch_reduce<float>(p, R, nvec, dh, Pbuf));
    } else {
      double *X = Xbuf, *R = Rbuf;
      hipLaunchKernelGGL(evr::k_pack_in<double>, dim3(blocks_for(n)), dim3(256), 0, p->stream, dpsi,
                         p->nb_packed, nvec, X);
      HIP_TRY(hipGetLastError());
      HIP_TRY(apply_assembled<double>(p, p->d_A64.ptr, X, nvec, R));
      HIP_TRY(launch_reduce<double>(p, R, nvec, dh, Pbuf));
    }
  } else if (!fp32) {
    if (p->keo_grid) {
      rc = keo_grid_hpsi<double>(p, nvec, dpsi, dh);
      if (rc) return rc;
    } else {
      double *X = Xbuf, *R = Rbuf;
      hipLaunchKernelGGL(evr::k_pack_in<double>, dim3(blocks_for(n)), dim3(256), 0, p->stream, dpsi,
                         p->nb_packed, nvec, X);
      HIP_TRY(hipGetLastError());
      rc = run_terms<double>(p, evr::kModeHpsi, nvec, X, R, nullptr, nullptr, p->d_bw64.ptr,
                             p->d_keo64.ptr, p->d_Veff.ptr);
      if (rc) return rc;
      HIP_TRY(launch_reduce<double>(p, R, nvec, dh, Pbuf));
    }
  } else {
    if (p->keo_grid) return keo_grid_hpsi<float>(p, nvec, dpsi, dh);
    rc = ensure_v32(p);
    if (rc) return rc;
    float *X = reinterpret_cast<float *>(Xbuf), *R = reinterpret_cast<float *>(Rbuf);
    hipLaunchKernelGGL(evr::k_pack_in<float>, dim3(blocks_for(n)), dim3(256), 0, p->stream, dpsi,
                       p->nb_packed, nvec, X);
    HIP_TRY(hipGetLastError());
    rc = run_terms<float>(p, evr::kModeHpsi, nvec, X, R, nullptr, nullptr, p->d_bw32.ptr,
                          p->d_keo32.ptr, p->d_Veff32.ptr);
    if (rc) return rc;
    HIP_TRY(launch_reduce<float>(p, R, nvec, dh, Pbuf));
  }
  return EVR_OK;
}

extern "C" {

int evr_oppsi(evr_plan *p, int nvec, const double *psi, double *hpsi, int flags)
{
  if (!p || !psi || !hpsi) return fail(EVR_ERR_ARG, "NULL argument");
  if (nvec < 1 || nvec > p->max_nvec) return fail(EVR_ERR_ARG, "nvec must be in 1..max_nvec");
  const bool dev = (flags & EVR_DEVICE_PTRS) != 0;
  const int64_t n = p->nb_packed * nvec;
  const double *dpsi = psi;
  double *dh = hpsi;
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(p->d_psi.ptr, psi, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
    dpsi = p->d_psi.ptr;
    dh = p->d_hpsi.ptr;
  }
  const int rc = oppsi_core(p, nvec, dpsi, dh, flags, p->d_X.ptr, p->d_R.ptr, p->d_P.ptr);
  if (rc) return rc;
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(hpsi, dh, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  return EVR_OK;
}

}  // extern "C"

// ---------------------------------------------------- allocation-free H|psi> --
// Work memory of evr_oppsi_noalloc: one file-scope device array (each GPU
// holds its own copy from module load to process end).  Host code never uses
// the array's name as an address: its device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch_ptr.
namespace {
constexpr size_t kScratchBytes = (size_t)EVR_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_device(int dev, double **out)
{
  if (dev < 0 || dev >= kMaxDevices) return fail(EVR_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    double *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(EVR_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return EVR_OK;
}

// 256-byte aligned word offsets of X, R and P inside the scratch array
int64_t scratch_words(int64_t words) { return (words + 31) & ~(int64_t)31; }
}  // namespace

extern "C" {

int evr_oppsi_noalloc(evr_plan *p, int nvec, const double *psi, double *hpsi, int flags)
{
  if (!p || !psi || !hpsi) return fail(EVR_ERR_ARG, "NULL argument");
  if (!(flags & EVR_DEVICE_PTRS))
    return fail(EVR_ERR_ARG, "evr_oppsi_noalloc takes device pointers only (EVR_DEVICE_PTRS)");
  if (nvec < 1 || nvec > p->max_nvec) return fail(EVR_ERR_ARG, "nvec must be in 1..max_nvec");
  if (p->keo_grid) return fail(EVR_ERR_ARG, "evr_oppsi_noalloc: plans with operator grids per point use evr_oppsi");
  if ((flags & EVR_FP32) && !p->V32_valid && !p->d_A32.ptr)
    return fail(EVR_ERR_STATE, "evr_oppsi_noalloc: the FP32 potential copy is made by a first FP32 evr_oppsi call");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != p->device) return fail(EVR_ERR_ARG, "evr_oppsi_noalloc: the plan's device is not the current device");
  const int64_t wx = scratch_words(p->nb_packed * nvec);
  const int64_t wr = scratch_words(std::max<int64_t>(p->nrows, 1) * nvec);
  const int64_t wp = scratch_words(std::max<int64_t>(p->nseg, 1) * nvec);
  const int64_t need = 8 * (wx + wr + wp);
  if (need > (int64_t)kScratchBytes)
    return fail(EVR_ERR_SIZE, "evr_oppsi_noalloc: " + std::to_string(need) + " bytes of work memory for " +
                                  std::to_string(nvec) + " vectors exceed the " +
                                  std::to_string(kScratchBytes) + "-byte scratch (use evr_oppsi)");
  double *base = nullptr;
  const int rc = scratch_for_device(dev, &base);
  if (rc) return rc;
  return oppsi_core(p, nvec, psi, hpsi, flags, base, base + wx, base + wx + wr);
}

int evr_basis_to_grid(evr_plan *p, const double *psiB, double *psiG)
{
  if (!p || !psiB || !psiG) return fail(EVR_ERR_ARG, "NULL argument");
  if (!p->d_G.ptr) HIP_TRY(p->d_G.alloc(p->sum_nq));
  HIP_TRY(hipMemcpyAsync(p->d_X.ptr, psiB, sizeof(double) * p->nb_packed, hipMemcpyHostToDevice,
                         p->stream));
  int rc = run_terms<double>(p, evr::kModeB2G, 1, p->d_X.ptr, nullptr, p->d_G.ptr, nullptr,
                             p->d_bw64.ptr, p->d_keo64.ptr, p->d_V.ptr);
  if (rc) return rc;
  std::vector<double> tmp;
  double *dst = psiG;
  if (p->dp_order) {
    tmp.resize(p->sum_nq);
    dst = tmp.data();
  }
  HIP_TRY(hipMemcpyAsync(dst, p->d_G.ptr, sizeof(double) * p->sum_nq, hipMemcpyDeviceToHost,
                         p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->dp_order)
    for (int64_t i = 0; i < p->sum_nq; ++i) psiG[p->dp_perm[i]] = tmp[i];
  return EVR_OK;
}

int evr_plan_diagonal(evr_plan *p, double *diag, int flags)
{
  if (!p || !diag) return fail(EVR_ERR_ARG, "NULL argument");
  if (p->keo_grid || !p->big.empty()) {
    // G, J, s per grid point: the per-point factors do not separate over the
    // modes, so the diagonal is read off H applied to blocks of unit vectors
    // (nb_packed / max_nvec applications: exact, and costly for large bases);
    // likewise with streamed terms, whose passes form no kinetic diagonal
    const int64_t nb = p->nb_packed;
    const int nv = std::max(1, p->max_nvec);
    DevBuf<double> X, Y, out;
    hipError_t e = X.alloc((size_t)nb * nv);
    if (e == hipSuccess) e = Y.alloc((size_t)nb * nv);
    if (e == hipSuccess && !(flags & EVR_DEVICE_PTRS)) e = out.alloc((size_t)nb);
    int rc = EVR_OK;
    if (e != hipSuccess) rc = fail(EVR_ERR_HIP, std::string("diagonal buffers: ") + hipGetErrorString(e));
    double *dst = (flags & EVR_DEVICE_PTRS) ? diag : out.ptr;
    for (int64_t b0 = 0; rc == EVR_OK && b0 < nb; b0 += nv) {
      const int nc = (int)std::min<int64_t>(nv, nb - b0);
      hipLaunchKernelGGL(k_unit_cols, dim3(blocks_for(nb * nc)), dim3(256), 0, p->stream, nb, nc, b0, X.ptr);
      rc = evr_oppsi(p, nc, X.ptr, Y.ptr, EVR_DEVICE_PTRS);
      if (rc == EVR_OK) {
        hipLaunchKernelGGL(k_take_diag, dim3((nc + 255) / 256), dim3(256), 0, p->stream, nb, nc, b0, Y.ptr, dst);
        e = hipGetLastError();
        if (e != hipSuccess) rc = fail(EVR_ERR_HIP, std::string("diagonal: ") + hipGetErrorString(e));
      }
    }
    if (rc == EVR_OK && !(flags & EVR_DEVICE_PTRS)) {
      e = hipMemcpyAsync(diag, dst, sizeof(double) * nb, hipMemcpyDeviceToHost, p->stream);
      if (e != hipSuccess) rc = fail(EVR_ERR_HIP, std::string("diagonal: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(p->stream);
    X.release(); Y.release(); out.release();
    return rc;
  }
  // V part: GridTOBasis of V with W o B^T as the grid->basis matrices
  const double *Gin = p->hasV ? p->d_V.ptr : nullptr;
  if (!Gin) {
    if (!p->d_G.ptr) HIP_TRY(p->d_G.alloc(p->sum_nq));
    HIP_TRY(zero_fill(p->d_G.ptr, p->sum_nq, p->stream));
    Gin = p->d_G.ptr;
  }
  int rc = run_terms<double>(p, evr::kModeG2B, 1, nullptr, p->d_R.ptr, nullptr, Gin,
                             p->d_bwdiag.ptr, p->d_keo64.ptr, p->d_V.ptr, 1);
  if (rc) return rc;
  double *dst = (flags & EVR_DEVICE_PTRS) ? diag : p->d_hpsi.ptr;
  HIP_TRY(launch_reduce<double>(p, p->d_R.ptr, 1, dst));
  if (!(flags & EVR_DEVICE_PTRS))
    HIP_TRY(hipMemcpyAsync(diag, dst, sizeof(double) * p->nb_packed, hipMemcpyDeviceToHost,
                           p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

// ncol columns through the B->G mode, the weighted G->B mode or a grid
// derivative, device pointers on the plan stream (DP plans: the reference's
// grid order through k_perm_cols and a work grid).
static int transform_cols(evr_plan *p, int what, int ncol, const double *in, double *out, int di, int dj)
{
  const int64_t nq = p->sum_nq, nb = p->nb_packed;
  DevBuf<double> tmp;
  if (p->dp_order && what != 0) HIP_TRY(tmp.alloc((size_t)nq * std::min(ncol, p->max_nvec)));
  for (int c0 = 0; c0 < ncol; c0 += p->max_nvec) {
    const int k = std::min(p->max_nvec, ncol - c0);
    const unsigned gb = blocks_for(nq * k);
    if (what == 0) {  // basis -> grid
      double *G = out + nq * c0;
      int rc = evr::plan_b2g_device(p, k, in + nb * c0, p->dp_order ? p->d_G.ptr : G);
      if (rc) return rc;
      if (p->dp_order) {
        hipLaunchKernelGGL(evr::k_perm_cols, dim3(gb), dim3(256), 0, p->stream, nq, k, p->d_dp_perm.ptr,
                           p->d_G.ptr, G, 1);
        HIP_TRY(hipGetLastError());
      }
      continue;
    }
    const double *Gin = in + nq * c0;
    if (p->dp_order) {
      hipLaunchKernelGGL(evr::k_perm_cols, dim3(gb), dim3(256), 0, p->stream, nq, k, p->d_dp_perm.ptr, Gin,
                         tmp.ptr, 0);
      HIP_TRY(hipGetLastError());
      Gin = tmp.ptr;
    }
    if (what == 1) {  // grid -> basis, weighted Smolyak sum
      int rc = run_terms<double>(p, evr::kModeG2B, k, nullptr, p->d_R.ptr, nullptr, Gin, p->d_bw64.ptr,
                                 p->d_keo64.ptr, p->d_V.ptr);
      if (rc) return rc;
      HIP_TRY(launch_reduce<double>(p, p->d_R.ptr, k, out + nb * c0));
    } else {  // derivative on the grids
      double *dst = p->dp_order ? p->d_G.ptr : out + nq * c0;
      hipLaunchKernelGGL(evr::k_grid_deriv, dim3((unsigned)p->n_gitems), dim3(256), 0, p->stream,
                         p->d_terms.ptr, p->d_slots.ptr, p->d_D1.ptr, p->d_D2.ptr, p->d_gitems.ptr, di, dj,
                         Gin, nq, k, dst);
      HIP_TRY(hipGetLastError());
      if (p->dp_order) {
        hipLaunchKernelGGL(evr::k_perm_cols, dim3(gb), dim3(256), 0, p->stream, nq, k, p->d_dp_perm.ptr,
                           p->d_G.ptr, out + nq * c0, 1);
        HIP_TRY(hipGetLastError());
      }
    }
  }
  HIP_TRY(hipStreamSynchronize(p->stream));  // tmp is released on return
  return EVR_OK;
}

// Host or device columns for transform_cols: host arrays go through device
// copies made for the call.
static int transform_entry(evr_plan *p, int what, int ncol, const double *in, double *out, int flags, int di,
                           int dj)
{
  if (!p || !in || !out) return fail(EVR_ERR_ARG, "NULL argument");
  if (ncol < 1) return fail(EVR_ERR_ARG, "ncol must be >= 1");
  if (p->D != p->D_full && what == 2) return fail(EVR_ERR_ARG, "grid derivatives need D <= 12");
  const int64_t nq = p->sum_nq, nb = p->nb_packed;
  if (p->dp_order) {
    const int64_t perm_cols = std::min(ncol, p->max_nvec);
    if (!p->d_G.ptr || p->d_G.n < (size_t)(nq * perm_cols)) HIP_TRY(p->d_G.alloc((size_t)(nq * perm_cols)));
  }
  if (what == 2) {
    int rc = grid_items(p);
    if (rc) return rc;
  }
  const int64_t nin = (what == 0 ? nb : nq) * ncol, nout = (what == 1 ? nb : nq) * ncol;
  if (flags & EVR_DEVICE_PTRS) return transform_cols(p, what, ncol, in, out, di, dj);
  DevBuf<double> din, dout;
  HIP_TRY(din.alloc((size_t)nin));
  HIP_TRY(dout.alloc((size_t)nout));
  HIP_TRY(hipMemcpyAsync(din.ptr, in, sizeof(double) * nin, hipMemcpyHostToDevice, p->stream));
  int rc = transform_cols(p, what, ncol, din.ptr, dout.ptr, di, dj);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, dout.ptr, sizeof(double) * nout, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

int evr_basis_to_grid_n(evr_plan *p, int ncol, const double *psiB, double *psiG, int flags)
{
  return transform_entry(p, 0, ncol, psiB, psiG, flags, -1, -1);
}

int evr_grid_to_basis_n(evr_plan *p, int ncol, const double *psiG, double *psiB, int flags)
{
  return transform_entry(p, 1, ncol, psiG, psiB, flags, -1, -1);
}

int evr_grid_deriv(evr_plan *p, int ncol, const int *tab_der, const double *psiG, double *out, int flags)
{
  if (!tab_der) return fail(EVR_ERR_ARG, "NULL argument");
  const int D = p ? p->D_full : 0;
  int d[2];
  for (int k = 0; k < 2; ++k) {
    if (tab_der[k] > D) return fail(EVR_ERR_ARG, "tab_der entries must be <= D");
    d[k] = tab_der[k] > 0 ? tab_der[k] - 1 : -1;  // tab_der < 0 counts as 0 (BG:1250)
  }
  return transform_entry(p, 2, ncol, psiG, out, flags, d[0], d[1]);
}

int evr_grid_to_basis(evr_plan *p, const double *psiG, double *psiB)
{
  if (!p || !psiB || !psiG) return fail(EVR_ERR_ARG, "NULL argument");
  if (!p->d_G.ptr) HIP_TRY(p->d_G.alloc(p->sum_nq));
  std::vector<double> tmp;
  const double *src = psiG;
  if (p->dp_order) {
    tmp.resize(p->sum_nq);
    for (int64_t i = 0; i < p->sum_nq; ++i) tmp[i] = psiG[p->dp_perm[i]];
    src = tmp.data();
  }
  HIP_TRY(hipMemcpyAsync(p->d_G.ptr, src, sizeof(double) * p->sum_nq, hipMemcpyHostToDevice,
                         p->stream));
  int rc = run_terms<double>(p, evr::kModeG2B, 1, nullptr, p->d_R.ptr, nullptr, p->d_G.ptr,
                             p->d_bw64.ptr, p->d_keo64.ptr, p->d_V.ptr);
  if (rc) return rc;
  HIP_TRY(launch_reduce<double>(p, p->d_R.ptr, 1, p->d_hpsi.ptr));
  HIP_TRY(hipMemcpyAsync(psiB, p->d_hpsi.ptr, sizeof(double) * p->nb_packed, hipMemcpyDeviceToHost,
                         p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return EVR_OK;
}

}  // extern "C"

// ------------------------------------------------------ multi-system operator --
// n assembled plans as one block-diagonal operator: the per-term tables are
// concatenated with their offsets shifted (term rows, packed functions, rows
// of R, reduction segments, pool words), so the assembled apply and the
// scatter run over every system's terms in one launch each.
struct evr_multi {
  int n = 0, max_nvec = 0;
  std::vector<int64_t> nb_off;  // packed offsets, n + 1
  int64_t nterms = 0, sum_nb = 0, nrows = 0, nseg = 0, nwords = 0, n_items = 0;
  DevBuf<evr::TermDev> d_terms;
  DevBuf<double> d_A64;
  DevBuf<int> d_map, d_rowpos, d_seg, d_fseg;
  DevBuf<int2> d_items;
  DevBuf<double> d_X, d_R, d_P, d_psi, d_hpsi;
  hipStream_t stream = nullptr;
  void release()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    d_terms.release(); d_A64.release(); d_map.release(); d_rowpos.release(); d_seg.release();
    d_fseg.release(); d_items.release(); d_X.release(); d_R.release(); d_P.release();
    d_psi.release(); d_hpsi.release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  evr_multi() = default;
  evr_multi(const evr_multi &) = delete;
  evr_multi &operator=(const evr_multi &) = delete;
  ~evr_multi() { release(); }  // also on a failed evr_multi_create
};

namespace {

AsmOp multi_op(const evr_multi *m)
{
  return AsmOp{m->d_terms.ptr, m->d_items.ptr, m->n_items, m->d_map.ptr, m->d_rowpos.ptr, m->d_seg.ptr,
               m->d_fseg.ptr, m->nseg, m->nb_off.back(), m->d_P.ptr, m->stream};
}

// dst[0..n) = src[0..n) + add (entries < 0 stay -1)
hipError_t offset_copy(const int *src, int64_t n, int64_t add, int *dst, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(evr::k_offset_ints, dim3(blocks_for(n)), dim3(256), 0, s, src, n, add, dst);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int evr_multi_create(evr_plan *const *plans, int n, int max_nvec, evr_multi **out)
{
  if (!plans || !out || n < 1 || max_nvec < 1) return fail(EVR_ERR_ARG, "bad argument");
  *out = nullptr;
  for (int k = 0; k < n; ++k) {
    if (!plans[k]) return fail(EVR_ERR_ARG, "NULL plan " + std::to_string(k));
    if (!plans[k]->d_A64.ptr)
      return fail(EVR_ERR_STATE, "plan " + std::to_string(k) + " has no FP64 assembled operator");
  }
  auto m = std::make_unique<evr_multi>();
  m->n = n;
  m->max_nvec = max_nvec;
  m->nb_off.assign(n + 1, 0);
  std::vector<int64_t> row_off(n + 1, 0), nb_rows(n + 1, 0), seg_off(n + 1, 0), word_off(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    const evr_plan *p = plans[k];
    m->nb_off[k + 1] = m->nb_off[k] + p->nb_packed;
    nb_rows[k + 1] = nb_rows[k] + p->sum_nb;
    row_off[k + 1] = row_off[k] + p->nrows;
    seg_off[k + 1] = seg_off[k] + p->nseg;
    word_off[k + 1] = word_off[k] + p->asm_elems;
    m->nterms += p->nterms;
  }
  m->sum_nb = nb_rows[n];
  m->nrows = row_off[n];
  m->nseg = seg_off[n];
  m->nwords = word_off[n];
  if (m->nb_off[n] >= INT32_MAX || m->sum_nb >= INT32_MAX || m->nseg >= INT32_MAX)
    return fail(EVR_ERR_SIZE, "systems too large to combine (32-bit tables)");
  HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  hipStream_t s = m->stream;
  // terms with shifted offsets, and the work items over all of them
  std::vector<evr::TermDev> terms;
  std::vector<int> nbt;
  terms.reserve(m->nterms);
  nbt.reserve(m->nterms);
  for (int k = 0; k < n; ++k)
    for (evr::TermDev td : plans[k]->terms) {
      td.off_nb += nb_rows[k];
      td.off_asm += word_off[k];
      terms.push_back(td);
      nbt.push_back(td.nbt);
    }
  const std::vector<int2> items = asm_items(nbt);
  if ((int64_t)items.size() > (int64_t)0xFFFFFFFFu / 256)
    return fail(EVR_ERR_SIZE, "too many assembled work items for one launch");
  HIP_TRY(m->d_terms.upload(terms.data(), terms.size(), s));
  HIP_TRY(m->d_items.upload(items.data(), items.size(), s));
  m->n_items = (int64_t)items.size();
  // pools and tables
  if (m->d_A64.alloc((size_t)m->nwords) != hipSuccess) {
    m->d_A64.ptr = nullptr;
    return fail(EVR_ERR_SIZE, "not enough device memory for the combined assembled operators (" +
                                  std::to_string(8.0 * m->nwords / 1e9) + " GB)");
  }
  HIP_TRY(m->d_map.alloc((size_t)m->sum_nb));
  HIP_TRY(m->d_rowpos.alloc((size_t)m->sum_nb));
  HIP_TRY(m->d_seg.alloc((size_t)m->nseg + 1));
  HIP_TRY(m->d_fseg.alloc((size_t)m->nb_off[n] + 1));
  for (int k = 0; k < n; ++k) {
    evr_plan *p = plans[k];
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpyAsync(m->d_A64.ptr + word_off[k], p->d_A64.ptr, sizeof(double) * p->asm_elems,
                           hipMemcpyDeviceToDevice, s));
    HIP_TRY(offset_copy(p->d_map.ptr, p->sum_nb, m->nb_off[k], m->d_map.ptr + nb_rows[k], s));
    HIP_TRY(offset_copy(p->d_rowpos.ptr, p->sum_nb, row_off[k], m->d_rowpos.ptr + nb_rows[k], s));
    HIP_TRY(offset_copy(p->d_seg.ptr, p->nseg, row_off[k], m->d_seg.ptr + seg_off[k], s));
    HIP_TRY(offset_copy(p->d_fseg.ptr, p->nb_packed, seg_off[k], m->d_fseg.ptr + m->nb_off[k], s));
  }
  const int seg_end = (int)m->nrows, fseg_end = (int)m->nseg;
  HIP_TRY(hipMemcpyAsync(m->d_seg.ptr + m->nseg, &seg_end, sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(m->d_fseg.ptr + m->nb_off[n], &fseg_end, sizeof(int), hipMemcpyHostToDevice, s));
  // work buffers
  HIP_TRY(m->d_X.alloc((size_t)m->nb_off[n] * max_nvec));
  HIP_TRY(m->d_R.alloc((size_t)std::max<int64_t>(m->nrows, 1) * max_nvec));
  HIP_TRY(m->d_P.alloc((size_t)std::max<int64_t>(m->nseg, 1) * max_nvec));
  HIP_TRY(m->d_psi.alloc((size_t)m->nb_off[n] * max_nvec));
  HIP_TRY(m->d_hpsi.alloc((size_t)m->nb_off[n] * max_nvec));
  HIP_TRY(hipStreamSynchronize(s));
  *out = m.release();
  return EVR_OK;
}

int evr_multi_destroy(evr_multi *m)
{
  delete m;  // releases its buffers and stream
  return EVR_OK;
}

int evr_multi_offsets(const evr_multi *m, int64_t *nb_offsets)
{
  if (!m || !nb_offsets) return fail(EVR_ERR_ARG, "NULL argument");
  std::copy(m->nb_off.begin(), m->nb_off.end(), nb_offsets);
  return EVR_OK;
}

void *evr_multi_stream(evr_multi *m) { return m ? (void *)m->stream : nullptr; }

int evr_multi_count(const evr_multi *m) { return m ? m->n : fail(EVR_ERR_ARG, "NULL argument"); }

int evr_multi_oppsi(evr_multi *m, int nvec, const double *psi, double *hpsi, int flags)
{
  if (!m || !psi || !hpsi) return fail(EVR_ERR_ARG, "NULL argument");
  if (nvec < 1 || nvec > m->max_nvec) return fail(EVR_ERR_ARG, "nvec must be in 1..max_nvec");
  const bool dev = (flags & EVR_DEVICE_PTRS) != 0;
  const int64_t nbp = m->nb_off.back(), n = nbp * nvec;
  const double *dpsi = psi;
  double *dh = hpsi;
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(m->d_psi.ptr, psi, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
    dpsi = m->d_psi.ptr;
    dh = m->d_hpsi.ptr;
  }
  hipLaunchKernelGGL(evr::k_pack_in<double>, dim3(blocks_for(n)), dim3(256), 0, m->stream, dpsi, nbp, nvec,
                     m->d_X.ptr);
  HIP_TRY(hipGetLastError());
  const AsmOp o = multi_op(m);
  HIP_TRY(apply_op<double>(o, m->d_A64.ptr, m->d_X.ptr, nvec, m->d_R.ptr));
  HIP_TRY(reduce_op<double>(o, m->d_R.ptr, nvec, dh));
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(hpsi, dh, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
  }
  return EVR_OK;
}

}  // extern "C"

// ------------------------------------------------------ coupled states (nb0) --
// H_ij blocks over one packed basis: (H psi)_i = sum_j H_ij psi_j, each block
// a plan of its own (the diagonal ones with the kinetic part and V_ii, the
// off-diagonal ones type 0 with V_ij), applied through its fastest path.
struct evr_coupled {
  int nb0 = 0, max_nvec = 0;
  int64_t nb = 0;
  std::vector<evr_plan *> blocks;  // nb0 x nb0, row-major, NULL = zero block
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  DevBuf<double> xin, tmp, acc, hin, hout;
  evr_coupled() = default;
  evr_coupled(const evr_coupled &) = delete;
  evr_coupled &operator=(const evr_coupled &) = delete;
  ~evr_coupled()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// dst (nb x nvec, column stride ldd) = src (nb x nvec, column stride lds) on s
hipError_t copy_cols(double *dst, int64_t ldd, const double *src, int64_t lds, int64_t nb, int nvec,
                     hipStream_t s)
{
  return hipMemcpy2DAsync(dst, sizeof(double) * ldd, src, sizeof(double) * lds, sizeof(double) * nb, nvec,
                          hipMemcpyDeviceToDevice, s);
}

}  // namespace

extern "C" {

int evr_coupled_create(evr_plan *const *blocks, int nb0, evr_coupled **out)
{
  if (!blocks || !out || nb0 < 1) return fail(EVR_ERR_ARG, "bad argument");
  *out = nullptr;
  auto c = std::make_unique<evr_coupled>();
  c->nb0 = nb0;
  c->blocks.assign(blocks, blocks + (size_t)nb0 * nb0);
  for (int i = 0; i < nb0; ++i) {
    bool row = false;
    for (int j = 0; j < nb0; ++j) {
      const evr_plan *p = c->blocks[(size_t)i * nb0 + j];
      if (!p) continue;
      row = true;
      if (c->nb == 0) {
        c->nb = p->nb_packed;
        c->max_nvec = p->max_nvec;
      }
      if (p->nb_packed != c->nb) return fail(EVR_ERR_ARG, "blocks on different basis sizes");
      c->max_nvec = std::min(c->max_nvec, p->max_nvec);
    }
    if (!row) return fail(EVR_ERR_ARG, "state " + std::to_string(i) + " has no block");
  }
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
  const size_t blk = (size_t)c->nb * c->max_nvec;
  HIP_TRY(c->xin.alloc(blk));
  HIP_TRY(c->tmp.alloc(blk));
  HIP_TRY(c->acc.alloc(blk));
  HIP_TRY(c->hin.alloc(blk * nb0));
  HIP_TRY(c->hout.alloc(blk * nb0));
  *out = c.release();
  return EVR_OK;
}

int evr_coupled_destroy(evr_coupled *c)
{
  delete c;
  return EVR_OK;
}

int evr_coupled_sizes(const evr_coupled *c, int64_t *nb, int *nb0, int *max_nvec)
{
  if (!c) return fail(EVR_ERR_ARG, "NULL argument");
  if (nb) *nb = c->nb;
  if (nb0) *nb0 = c->nb0;
  if (max_nvec) *max_nvec = c->max_nvec;
  return EVR_OK;
}

int evr_coupled_diagonal(evr_coupled *c, double *diag, int flags)
{
  if (!c || !diag) return fail(EVR_ERR_ARG, "NULL argument");
  const bool dev = (flags & EVR_DEVICE_PTRS) != 0;
  for (int i = 0; i < c->nb0; ++i) {
    evr_plan *b = c->blocks[(size_t)i * c->nb0 + i];
    double *di = diag + (int64_t)i * c->nb;
    if (b) {
      const int rc = evr_plan_diagonal(b, di, dev ? EVR_DEVICE_PTRS : 0);
      if (rc) return rc;
    } else if (dev) {
      HIP_TRY(zero_fill(di, c->nb, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
      std::fill(di, di + c->nb, 0.0);
    }
  }
  return EVR_OK;
}

void *evr_coupled_stream(evr_coupled *c) { return c ? (void *)c->stream : nullptr; }

int evr_coupled_oppsi(evr_coupled *c, int nvec, const double *psi, double *hpsi, int flags)
{
  if (!c || !psi || !hpsi) return fail(EVR_ERR_ARG, "NULL argument");
  if (nvec < 1 || nvec > c->max_nvec) return fail(EVR_ERR_ARG, "nvec must be in 1..max_nvec");
  const bool dev = (flags & EVR_DEVICE_PTRS) != 0;
  const int nb0 = c->nb0;
  const int64_t nb = c->nb, ld = nb * nb0, n = ld * nvec;
  const double *dpsi = psi;
  double *dh = hpsi;
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(c->hin.ptr, psi, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    dpsi = c->hin.ptr;
    dh = c->hout.ptr;
  }
  const int bflags = (flags & (EVR_FP32 | EVR_MATRIX_FREE | EVR_AUTO_PATH)) | EVR_DEVICE_PTRS;
  for (int i = 0; i < nb0; ++i) {
    bool first = true;
    for (int j = 0; j < nb0; ++j) {
      evr_plan *p = c->blocks[(size_t)i * nb0 + j];
      if (!p) continue;
      // psi_j (rows j nb .. (j+1) nb - 1 of every column) -> contiguous
      HIP_TRY(copy_cols(c->xin.ptr, nb, dpsi + j * nb, ld, nb, nvec, c->stream));
      HIP_TRY(hipEventRecord(c->fork, c->stream));
      HIP_TRY(hipStreamWaitEvent(p->stream, c->fork, 0));
      const int rc = evr_oppsi(p, nvec, c->xin.ptr, first ? c->acc.ptr : c->tmp.ptr, bflags);
      if (rc) return rc;
      HIP_TRY(hipEventRecord(c->join, p->stream));
      HIP_TRY(hipStreamWaitEvent(c->stream, c->join, 0));
      if (!first) {
        hipLaunchKernelGGL(evr::k_axpy, dim3(blocks_for(nb * nvec)), dim3(256), 0, c->stream, nb * nvec,
                           1.0, c->tmp.ptr, c->acc.ptr);
        HIP_TRY(hipGetLastError());
      }
      first = false;
    }
    HIP_TRY(copy_cols(dh + i * nb, ld, c->acc.ptr, nb, nb, nvec, c->stream));
  }
  if (!dev) {
    HIP_TRY(hipMemcpyAsync(hpsi, dh, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return EVR_OK;
}

}  // extern "C"

// ---------------------------------------------- internal (driver) helpers --

int evr::plan_b2g_device(evr_plan *p, int ncol, const double *X, double *G)
{
  if (!p || ncol < 0 || (ncol && (!X || !G))) return fail(EVR_ERR_ARG, "bad argument");
  for (int c0 = 0; c0 < ncol; c0 += p->max_nvec) {
    const int k = std::min(p->max_nvec, ncol - c0);
    const int64_t n = p->nb_packed * k;
    hipLaunchKernelGGL(evr::k_pack_in<double>, dim3(blocks_for(n)), dim3(256), 0, p->stream,
                       X + p->nb_packed * c0, p->nb_packed, k, p->d_X.ptr);
    HIP_TRY(hipGetLastError());
    int rc = run_terms<double>(p, evr::kModeB2G, k, p->d_X.ptr, nullptr, G + p->sum_nq * c0, nullptr,
                               p->d_bw64.ptr, p->d_keo64.ptr, p->d_V.ptr);
    if (rc) return rc;
  }
  return EVR_OK;
}

int evr::plan_grid_weights(evr_plan *p, const double **out)
{
  if (!p || !out) return fail(EVR_ERR_ARG, "bad argument");
  if (!p->d_gw.ptr) {
    // quadrature weight of 1D point q: W[0, q] / B[q, 0] (W = B^T diag(w), B[q, 0] > 0)
    std::vector<double> bw(p->d_bw64.n);
    HIP_TRY(hipMemcpy(bw.data(), p->d_bw64.ptr, sizeof(double) * bw.size(), hipMemcpyDeviceToHost));
    std::vector<std::vector<double>> wq(p->slots.size());
    for (size_t s = 0; s < p->slots.size(); ++s) {
      const evr::SlotDev &sd = p->slots[s];
      wq[s].resize(sd.nq);
      for (int q = 0; q < sd.nq; ++q) wq[s][q] = bw[sd.offW + (int64_t)sd.nb * q] / bw[sd.offB + q];
    }
    std::vector<double> gw(p->sum_nq);
    for (const evr::TermDev &td : p->terms) {
      double *g = gw.data() + td.off_nq;
      g[0] = td.w;
      int64_t len = 1;  // first mode fastest: extend the tensor product mode by mode
      for (int d = 0; d < p->D; ++d) {
        const std::vector<double> &w = wq[td.slot[d]];
        for (int q = (int)w.size() - 1; q >= 0; --q)
          for (int64_t i = 0; i < len; ++i) g[q * len + i] = g[i] * w[q];
        len *= (int64_t)w.size();
      }
    }
    HIP_TRY(p->d_gw.upload(gw.data(), gw.size(), p->stream));
  }
  *out = p->d_gw.ptr;
  return EVR_OK;
}

int evr::plan_grid_to_user(const evr_plan *p, const double *term_grid, double *user)
{
  if (!p) return fail(EVR_ERR_ARG, "bad argument");
  if (p->dp_order)
    for (int64_t i = 0; i < p->sum_nq; ++i) user[p->dp_perm[i]] = term_grid[i];
  else
    std::copy(term_grid, term_grid + p->sum_nq, user);
  return EVR_OK;
}

