// batch.hip -- heterogeneous batches (include/matfun_hip.h, mf_hip_*), host side: the plan (class-homogeneous tiles
// of up to 64 matrices and their row tables), the launch strategies and their tuner, the Philox perturbation of the
// Newton driver and graph-captured Newton sequences.  The batch kernels are in batch_kernels.hpp.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/matfun_hip.h"
#include "batch_kernels.hpp"
#include "gmf_sorted.hpp"
#include "philox.hpp"

namespace mf {

// A <- A o (1 + sigma Z) for the matrices of tiles t0 .. t1 - 1 (one matrix per lane)
__global__ void __launch_bounds__(WAVE) perturb_kernel(BatchArgs a, double* __restrict__ A, double sigma,
                                                       uint64_t seed, uint32_t iteration) {
  const int lane = threadIdx.x;
  for (int64_t t = a.t0 + blockIdx.x; t < a.t1; t += gridDim.x) {
    const BatchTile tl = a.tiles[t];
    if (lane < tl.nm) {
      const int64_t m = a.rm[tl.off + lane];
      double f[9];
      perturb_factors(seed, iteration, (uint64_t)m, cls_fam(tl.cls) == 0, sigma, f);
#pragma unroll
      for (int k = 0; k < 9; k++) A[9 * m + k] *= f[k];
    }
  }
}

}  // namespace mf

using namespace mf;

struct mf_hip_plan {
  int64_t N = 0, NJ = 0, NH = 0, ntiles = 0;
  int ncls_present = 0;
  int needJ = 0, needH = 0;
  int cus = 256;
  int has_pow = 0;
  std::vector<BatchTile> tiles;           // class-major (host copy)
  int64_t cls_t0[NCLS] = {}, cls_t1[NCLS] = {};
  BatchTile* d_sorted = nullptr;          // class-major
  BatchTile* d_mixed = nullptr;           // interleaved in proportion to the class sizes
  int64_t *d_rm = nullptr, *d_rj = nullptr, *d_rh = nullptr;
  double* d_alpha = nullptr;              // the groups' exponents per matrix (pow groups; 1 elsewhere)
  BatchTables tab;                        // (host copy: the gmf series tables of the staged classes)
  // mode 3: the workspace of the staged general-matrix classes (one class at a time on the caller's stream),
  // allocated when the mode is first set
  void* d_work = nullptr;
  size_t work_bytes = 0;
  int mode = 0, per_cu = 6, order = 1;    // launch strategy
  // mode 2: the classes on the caller's stream and up to NSIDE side streams (default priority), forked and joined
  // with events; the classes' assignment balanced by estimated cost
  static constexpr int NSIDE = 3;
  hipStream_t side[NSIDE] = {};
  hipEvent_t fork = nullptr, join[NSIDE] = {};
  int lane_of_cls[NCLS] = {};             // 0: the caller's stream, k: side[k - 1]
  // sequences captured from this plan and not yet destroyed: their graphs hold the plan's device tables and side
  // streams, so the plan refuses to be destroyed while any is alive (mutable: sequences take a const plan)
  mutable std::atomic<int64_t> live_sequences{0};
};

struct mf_hip_sequence {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const mf_hip_plan* plan = nullptr;      // (its live_sequences counts this sequence)
};

namespace {

template <class T>
int upload(T** d, const T* h, size_t count) {
  if (count == 0) count = 1;
  if (hipMalloc((void**)d, count * sizeof(T)) != hipSuccess) return -2;
  if (h && hipMemcpy(*d, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return -2;
  return 0;
}

void free_plan(mf_hip_plan* p) {
  if (!p) return;
  void* ptrs[] = {p->d_sorted, p->d_mixed, p->d_rm, p->d_rj, p->d_rh, p->d_alpha, p->d_work};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (int k = 0; k < mf_hip_plan::NSIDE; k++) {
    if (p->side[k]) (void)hipStreamDestroy(p->side[k]);
    if (p->join[k]) (void)hipEventDestroy(p->join[k]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  delete p;
}

// estimated time of one tile of a class, relative (order-2 smf = 1; from the homogeneous kernels' rates)
double tile_cost(int c) {
  const int o = cls_order(c);
  return cls_fam(c) == 0 ? (o == 2 ? 1.0 : (o == 1 ? 0.15 : 0.05)) : (o == 2 ? 1.05 : (o == 1 ? 0.45 : 0.3));
}

// Aw: with the perturbation pt applied by the evaluation (A itself, writable), else null
BatchArgs args_of(const mf_hip_plan* p, const BatchTile* tiles, int64_t t0, int64_t t1, const double* A,
                  const double* alpha, double* Aw = nullptr, const BatchPerturb* pt = nullptr) {
  BatchArgs a;
  a.tiles = tiles;
  a.t0 = t0;
  a.t1 = t1;
  a.rm = p->d_rm;
  a.rj = p->d_rj;
  a.rh = p->d_rh;
  a.A = A;
  a.alv = alpha ? alpha : p->d_alpha;
  a.perturb = Aw && pt ? 1 : 0;
  a.Aw = Aw;
  a.pt = pt ? *pt : BatchPerturb{0.0, 0, 0};
  // (read by the class launches, whose t0 .. t1 are one class's tiles in the class-major order of p->tiles)
  a.coff = t1 > t0 ? p->tiles[(size_t)t0].off : 0;
  a.cn = t1 > t0 ? (t1 - t0 - 1) * WAVE + p->tiles[(size_t)(t1 - 1)].nm : 0;
  return a;
}

unsigned grid_for(const mf_hip_plan* p, int64_t ntiles, int per_cu) {
  int64_t g = per_cu > 0 ? (int64_t)per_cu * p->cus : ntiles;
  g = std::min<int64_t>(g, ntiles);
  g = std::min<int64_t>(g, (int64_t)INT32_MAX);
  return (unsigned)std::max<int64_t>(g, 1);
}

// the matrices of class c (its tiles are full but the last)
int64_t cls_matrices(const mf_hip_plan* p, int c) {
  const int64_t t = p->cls_t1[c] - p->cls_t0[c];
  return t == 0 ? 0 : (t - 1) * WAVE + p->tiles[(size_t)(p->cls_t1[c] - 1)].nm;
}

// the workspace bytes of mode 3: the largest stageable class's
size_t staged_work_bytes(const mf_hip_plan* p) {
  size_t b = 0;
  for (int c = 0; c < NCLS; c++)
    if (batch_gmf_stageable(c)) b = std::max(b, gmf_staged_list_bytes(cls_matrices(p, c), cls_order(c)));
  return b;
}

// mode 0: the kernel over all classes; mode 1: one kernel per class; mode 3: as mode 1, the gmf classes of order >= 1
// other than exp through the staged evaluation (gmf_staged_list, then their tiles finished from the staged data).
// per_cu > 0: persistent, per_cu workgroups per compute unit; 0: one workgroup per tile.  order (mode 0): 0
// class-major tiles, 1 interleaved.  pt (with Aw = A): each matrix perturbed first by the kernel that evaluates it
// (or by the staged nodes pass), its perturbed inputs written back over A -- mf_hip_perturb then the evaluation,
// bitwise, in one pass over A.
template <class OT>
int batch_eval(const mf_hip_plan* p, const double* A, const double* alpha, OT* F, OT* J, OT* H, hipStream_t st,
               int mode, int per_cu, int order, const BatchPerturb* pt = nullptr, double* Aw = nullptr) {
  if (!p || (p->N > 0 && (!A || !F)) || (p->needJ && !J) || (p->needH && !H)) return -1;
  if (p->ntiles == 0) return 0;
  if (mode == 3) {
    if (!p->d_work && staged_work_bytes(p) > 0) return -1;   // (mf_hip_plan_set_strategy allocates it)
    for (int c = 0; c < NCLS; c++) {
      const int64_t n = p->cls_t1[c] - p->cls_t0[c];
      if (n == 0) continue;
      const BatchArgs a = args_of(p, p->d_sorted, p->cls_t0[c], p->cls_t1[c], A, alpha, Aw, pt);
      if (batch_gmf_stageable(c)) {
        const int slot = cls_slot(c), ord = cls_order(c);
        const int64_t off0 = p->tiles[(size_t)p->cls_t0[c]].off;
        const GmfCoef& bt = slot == SLOT_POW ? p->tab.recips : p->tab.gmf[slot];
        const GmfStaged sg = gmf_staged_list(st, slot == 1 ? LOG : POW, ord, slot == SLOT_POW, cls_matrices(p, c),
                                             p->d_rm + off0, A, slot == 2 ? 0.5 : 1.0, a.alv, bt, p->d_work,
                                             a.perturb ? Aw : nullptr, a.pt.sigma, a.pt.seed, a.pt.iteration);
        BatchArgs af = a;                       // (the finishing pass reads the inputs the nodes pass perturbed)
        af.perturb = 0;
        batch_launch_gmf_staged(c, st, (unsigned)n, af, off0, sg, F, J, H);
      } else if (cls_fam(c) == 0) {
        batch_launch_smf(c, st, grid_for(p, n, per_cu), per_cu > 0, a, F, J, H);
      } else {
        batch_launch_gmf(c, st, grid_for(p, n, per_cu), per_cu > 0, a, F, J, H);
      }
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  if (mode == 0) {
    const BatchArgs a = args_of(p, order == 1 ? p->d_mixed : p->d_sorted, 0, p->ntiles, A, alpha, Aw, pt);
    batch_launch_fused(st, grid_for(p, p->ntiles, per_cu), per_cu > 0, a, F, J, H);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  bool used[mf_hip_plan::NSIDE] = {};
  if (mode == 2 && hipEventRecord(p->fork, st) != hipSuccess) return -2;
  for (int c = 0; c < NCLS; c++) {
    const int64_t n = p->cls_t1[c] - p->cls_t0[c];
    if (n == 0) continue;
    hipStream_t s = st;
    if (mode == 2 && p->lane_of_cls[c] > 0) {
      const int k = p->lane_of_cls[c] - 1;
      if (!used[k] && hipStreamWaitEvent(p->side[k], p->fork, 0) != hipSuccess) return -2;
      used[k] = true;
      s = p->side[k];
    }
    const BatchArgs a = args_of(p, p->d_sorted, p->cls_t0[c], p->cls_t1[c], A, alpha, Aw, pt);
    if (cls_fam(c) == 0) batch_launch_smf(c, s, grid_for(p, n, per_cu), per_cu > 0, a, F, J, H);
    else batch_launch_gmf(c, s, grid_for(p, n, per_cu), per_cu > 0, a, F, J, H);
  }
  const bool launched_ok = hipGetLastError() == hipSuccess;
  for (int k = 0; k < mf_hip_plan::NSIDE; k++)     // (joined even after a launch error: no stream left forked)
    if (used[k] && (hipEventRecord(p->join[k], p->side[k]) != hipSuccess ||
                    hipStreamWaitEvent(st, p->join[k], 0) != hipSuccess))
      return -2;
  return launched_ok ? 0 : -2;
}

int slot_of(int family, int kind) {
  if (kind == POW) return SLOT_POW;
  if (family == 0) return kind == EXP ? 0 : (kind == LOG ? 1 : 2);
  return kind == EXP ? 0 : (kind == LOG ? 1 : 2);
}

}  // namespace

extern "C" int mf_hip_plan_create(const mf_hip_group* groups, int64_t ngroups, mf_hip_plan** out) {
  if (!out || ngroups < 0 || (ngroups > 0 && !groups)) return -1;
  *out = nullptr;
  // ---- validate every group first (fewer than 2^31 matrices in all: the batch kernels may keep input rows as
  //      32-bit values), then the rows of each matrix in the caller's layout, matrices per class in group order
  int64_t total = 0;
  for (int64_t g = 0; g < ngroups; g++) {
    const mf_hip_group& G = groups[g];
    if (G.family < 0 || G.family > 1 || G.kind < EXP || G.kind > SQRT || G.order < 0 || G.order > 2 ||
        G.count < 0 || (G.kind == POW && !std::isfinite(G.alpha)))
      return -1;
    total += G.count;
    if (total > INT32_MAX) return -1;
  }
  std::vector<std::vector<int64_t>> cm(NCLS), cj(NCLS), ch(NCLS);
  std::vector<double> galpha;
  mf_hip_plan* p = new (std::nothrow) mf_hip_plan;
  if (!p) return -2;
  for (int64_t g = 0; g < ngroups; g++) {
    const mf_hip_group& G = groups[g];
    const int c = cls_of(G.family, slot_of(G.family, G.kind), G.order);
    for (int32_t i = 0; i < G.count; i++) {
      cm[c].push_back(p->N + i);
      cj[c].push_back(G.order >= 1 ? p->NJ + i : 0);
      ch[c].push_back(G.order >= 2 ? p->NH + i : 0);
    }
    galpha.resize((size_t)(p->N + G.count), 1.0);
    if (G.kind == POW)
      for (int32_t i = 0; i < G.count; i++) galpha[(size_t)(p->N + i)] = G.alpha;
    p->N += G.count;
    if (G.order >= 1) p->NJ += G.count;
    if (G.order >= 2) p->NH += G.count;
    if (G.count > 0 && G.order >= 1) p->needJ = 1;
    if (G.count > 0 && G.order >= 2) p->needH = 1;
    if (G.count > 0 && G.kind == POW) p->has_pow = 1;
  }
  // ---- tiles (class-major) and the row tables in tile order
  std::vector<int64_t> rm, rj, rh;
  rm.reserve((size_t)p->N);
  rj.reserve((size_t)p->N);
  rh.reserve((size_t)p->N);
  for (int c = 0; c < NCLS; c++) {
    p->cls_t0[c] = (int64_t)p->tiles.size();
    const int64_t n = (int64_t)cm[c].size();
    if (n > 0) p->ncls_present++;
    for (int64_t lo = 0; lo < n; lo += WAVE) {
      const int nm = (int)std::min<int64_t>(WAVE, n - lo);
      p->tiles.push_back(BatchTile{c, nm, (int64_t)rm.size()});
      for (int k = 0; k < nm; k++) {
        rm.push_back(cm[c][(size_t)(lo + k)]);
        rj.push_back(cj[c][(size_t)(lo + k)]);
        rh.push_back(ch[c][(size_t)(lo + k)]);
      }
    }
    p->cls_t1[c] = (int64_t)p->tiles.size();
  }
  p->ntiles = (int64_t)p->tiles.size();
  // interleaved order: tile i of a class with T tiles at key (i + 1/2) / T, stable by class
  std::vector<std::pair<double, int64_t>> key;
  key.reserve(p->tiles.size());
  for (int c = 0; c < NCLS; c++) {
    const int64_t T = p->cls_t1[c] - p->cls_t0[c];
    for (int64_t i = 0; i < T; i++) key.emplace_back((i + 0.5) / (double)T, p->cls_t0[c] + i);
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<double, int64_t>& x, const std::pair<double, int64_t>& y) { return x.first < y.first; });
  std::vector<BatchTile> mixed;
  mixed.reserve(p->tiles.size());
  for (auto& kv : key) mixed.push_back(p->tiles[(size_t)kv.second]);
  // ---- mode 2: classes to streams, the largest estimated cost first onto the least loaded of 1 + NSIDE streams
  {
    std::vector<std::pair<double, int>> cost;
    for (int c = 0; c < NCLS; c++)
      if (p->cls_t1[c] > p->cls_t0[c]) cost.emplace_back(tile_cost(c) * (double)(p->cls_t1[c] - p->cls_t0[c]), c);
    std::sort(cost.begin(), cost.end(), [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
      return x.first > y.first;
    });
    double load[1 + mf_hip_plan::NSIDE] = {};
    for (auto& cc : cost) {
      int best = 0;
      for (int k = 1; k <= mf_hip_plan::NSIDE; k++)
        if (load[k] < load[best]) best = k;
      load[best] += cc.first;
      p->lane_of_cls[cc.second] = best;
    }
  }
  // ---- series tables
  BatchTables tab;
  for (int s = 0; s < 3; s++) {
    SmfCoef c;
    smf_coefs(s == 0 ? EXP : (s == 1 ? LOG : SQRT), s == 2 ? 0.5 : 0.0, c);
    for (int k = 0; k < SMF_NTERMS; k++) tab.smf_c2[s][k] = c.c[k];
  }
  gmf_coefs(EXP, 0.0, tab.gmf[0]);
  gmf_coefs(LOG, 0.0, tab.gmf[1]);
  gmf_coefs(POW, 0.5, tab.gmf[2]);
  gmf_recips(tab.recips);
  p->tab = tab;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess) {
    free_plan(p);
    return -2;
  }
  if (hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess) {
    free_plan(p);
    return -2;
  }
  for (int k = 0; k < mf_hip_plan::NSIDE; k++)
    if (hipStreamCreateWithFlags(&p->side[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming) != hipSuccess) {
      free_plan(p);
      return -2;
    }
  if (upload(&p->d_sorted, p->tiles.data(), p->tiles.size()) || upload(&p->d_mixed, mixed.data(), mixed.size()) ||
      upload(&p->d_rm, rm.data(), rm.size()) || upload(&p->d_rj, rj.data(), rj.size()) ||
      upload(&p->d_rh, rh.data(), rh.size()) || upload(&p->d_alpha, galpha.data(), galpha.size()) ||
      batch_upload_fused(tab) || batch_upload_smf(tab) || batch_upload_gmf(tab)) {
    free_plan(p);
    return -2;
  }
  *out = p;
  return 0;
}

extern "C" int mf_hip_plan_destroy(mf_hip_plan* p) {
  if (p && p->live_sequences.load() > 0) return -1;     // (a sequence's graph still points at the plan's tables)
  free_plan(p);
  return 0;
}

extern "C" int64_t mf_hip_plan_size(const mf_hip_plan* p, int what) {
  if (!p) return -1;
  switch (what) {
    case 0: return p->N;
    case 1: return p->NJ;
    case 2: return p->NH;
    case 3: return p->ntiles;
    case 4: return p->ncls_present;
    default: return -1;
  }
}

extern "C" int mf_hip_batch_eval(const mf_hip_plan* p, const double* A, const double* alpha, double* F, double* J,
                                 double* H, void* stream) {
  return p ? batch_eval(p, A, alpha, F, J, H, (hipStream_t)stream, p->mode, p->per_cu, p->order) : -1;
}

extern "C" int mf_hip_batch_eval_mixed(const mf_hip_plan* p, const double* A, const double* alpha, float* F, float* J,
                                       float* H, void* stream) {
  return p ? batch_eval(p, A, alpha, F, J, H, (hipStream_t)stream, p->mode, p->per_cu, p->order) : -1;
}

// the workspace of mode 3, allocated once per plan (0, or -2 when the allocation fails)
static int alloc_staged_work(mf_hip_plan* p) {
  if (p->d_work) return 0;
  const size_t b = staged_work_bytes(p);
  if (b == 0) return 0;
  if (hipMalloc(&p->d_work, b) != hipSuccess) {
    p->d_work = nullptr;
    return -2;
  }
  p->work_bytes = b;
  return 0;
}

extern "C" int mf_hip_plan_set_strategy(mf_hip_plan* p, int mode, int per_cu, int order) {
  if (!p || mode < 0 || mode > 3 || per_cu < 0 || per_cu > 64 || order < 0 || order > 1) return -1;
  if (mode == 3 && alloc_staged_work(p) != 0) return -2;
  p->mode = mode;
  p->per_cu = per_cu;
  p->order = order;
  return 0;
}

extern "C" int mf_hip_plan_tune(mf_hip_plan* p, const double* A, const double* alpha, void* F, void* J, void* H,
                                int mixed, void* stream, double* result) {
  if (!p) return -1;
  struct Cand {
    int mode, per_cu, order;
  };
  // (the fused kernel, mode 0, is left out: measured on MI355X on the C4 batch 14.1 ms with one workgroup per tile and
  // 17.5-30.6 ms persistent, against 12.35 ms for the split kernels -- every tile runs with the registers and LDS of
  // the largest class, gmf order 2, two wavefronts per SIMD)
  std::vector<Cand> cands;
  for (int k : {0, 2, 4, 8}) cands.push_back({1, k, 0});
  for (int k : {0, 2}) cands.push_back({2, k, 0});
  if (staged_work_bytes(p) > 0 && alloc_staged_work(p) == 0) cands.push_back({3, 0, 0});
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -2;
  double best = 1e300;
  Cand bc = {p->mode, p->per_cu, p->order};
  int rc = 0;
  for (const Cand& c : cands) {
    auto run = [&]() {
      return mixed ? batch_eval(p, A, alpha, (float*)F, (float*)J, (float*)H, st, c.mode, c.per_cu, c.order)
                   : batch_eval(p, A, alpha, (double*)F, (double*)J, (double*)H, st, c.mode, c.per_cu, c.order);
    };
    if ((rc = run()) != 0) break;                      // (warm-up)
    const int reps = 3;
    if (hipEventRecord(e0, st) != hipSuccess) { rc = -2; break; }
    for (int r = 0; r < reps && rc == 0; r++) rc = run();
    if (rc != 0 || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = rc ? rc : -2; break; }
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = -2; break; }
    const double us = 1e3 * ms / reps;
    if (us < best) { best = us; bc = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != 0) return rc;
  p->mode = bc.mode;
  p->per_cu = bc.per_cu;
  p->order = bc.order;
  if (result) {
    result[0] = bc.mode;
    result[1] = bc.per_cu;
    result[2] = bc.order;
    result[3] = best;
  }
  return 0;
}

extern "C" int mf_hip_perturb(const mf_hip_plan* p, double* A, double rel_sigma, uint64_t seed, int iteration,
                              void* stream) {
  if (!p || (p->N > 0 && !A) || iteration < 0 || !std::isfinite(rel_sigma)) return -1;
  if (p->ntiles == 0) return 0;
  const BatchArgs a = args_of(p, p->d_sorted, 0, p->ntiles, A, nullptr);
  hipLaunchKernelGGL(perturb_kernel, dim3(grid_for(p, p->ntiles, 8)), dim3(WAVE), 0, (hipStream_t)stream, a, A,
                     rel_sigma, seed, (uint32_t)iteration);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int mf_hip_sequence_create(const mf_hip_plan* p, const double* A0, double* A, const double* alpha, void* F,
                                      void* J, void* H, int mixed, int iterations, double rel_sigma, uint64_t seed,
                                      mf_hip_sequence** out) {
  if (!p || !out || iterations < 0 || !std::isfinite(rel_sigma) || (p->N > 0 && (!A0 || !A || !F)) ||
      (p->needJ && !J) || (p->needH && !H))
    return -1;
  *out = nullptr;
  mf_hip_sequence* s = new (std::nothrow) mf_hip_sequence;
  if (!s) return -2;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return -2;
  }
  int rc = 0;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = -2;
  if (rc == 0 && p->N > 0 &&
      hipMemcpyAsync(A, A0, (size_t)p->N * 9 * sizeof(double), hipMemcpyDeviceToDevice, cs) != hipSuccess)
    rc = -2;
  for (int it = 1; rc == 0 && it <= iterations; it++) {
    // (the perturbation of iteration it applied by the evaluation's kernels: mf_hip_perturb then the evaluation,
    // bitwise, with one pass over A)
    const BatchPerturb pt{rel_sigma, seed, (uint32_t)it};
    rc = mixed ? batch_eval(p, A, alpha, (float*)F, (float*)J, (float*)H, cs, p->mode, p->per_cu, p->order, &pt, A)
               : batch_eval(p, A, alpha, (double*)F, (double*)J, (double*)H, cs, p->mode, p->per_cu, p->order, &pt,
                            A);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(cs, &g);
  if (rc == 0 && (ec != hipSuccess || !g)) rc = -2;
  if (rc == 0 && hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0) != hipSuccess) rc = -2;
  s->graph = g;
  (void)hipStreamDestroy(cs);
  if (rc != 0) {
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    delete s;
    return rc;
  }
  s->plan = p;
  p->live_sequences.fetch_add(1);
  *out = s;
  return 0;
}

extern "C" int mf_hip_sequence_launch(mf_hip_sequence* s, void* stream) {
  if (!s || !s->exec) return -1;
  return hipGraphLaunch(s->exec, (hipStream_t)stream) == hipSuccess ? 0 : -2;
}

extern "C" int mf_hip_sequence_destroy(mf_hip_sequence* s) {
  if (!s) return 0;
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->plan) s->plan->live_sequences.fetch_sub(1);
  delete s;
  return 0;
}
