// batch.hip -- many independent grids advanced together on one GPU.
//
// Every instance keeps its own state in HBM (state w, sweep output b,
// predictor w1), its own device scalars (dt, t, it, dtcourant, smallp,
// nerror, stop condition) and its own selectors.  One batched step runs the
// TVDLF or TVD-MUSCL step (one x and one y sweep, src/vactvdlf.t:5-299) of
// every active instance: each launch walks a work list of (instance, tile)
// entries, so grids of any size share a launch.  Instances are grouped by kernel instance (physics, arithmetic
// tier, limiter) and by tile shape, which the autotuner picks per size
// class.  The sweep order of an instance follows its own step parity
// (src/vac.t:310-324): its tiles sit in the x-first or the y-first list.
//
// The per-instance reductions ride on the sweeps: the last block of an
// instance to finish a fused sweep takes the instance's Courant maximum
// (src/vactvdlf.t:184-201) and, after the second sweep, ends the step (it,
// t) and computes the next dt and the stop condition (getdt_courant with
// lagged dtcourant, src/vac.t:664-722, and src/vac.t:73-83).  The step is
// then a handful of sweep launches, captured once per membership into a
// hipGraph and replayed.  The first dt of an instance (pressure floor and
// cell-centre Courant maxima) is computed at admission, on the copy stream.
//
// Admission, removal, status and result copies do not wait for the steps:
// states are staged through pinned buffers and uploaded on a copy stream
// that the step stream waits for (by event) before the new instance's first
// step; removed instances keep their memory and slot until every step and
// copy enqueued before the removal has completed.
//
// The per-cell arithmetic is the single-grid one (sweep.hpp), so an instance
// advanced in a batch is bitwise identical to the same grid advanced alone.
#include "../../include/vacgpu.h"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Work {
  int inst;
  int bx, by, bz;
  int pidx;  // index of this tile in the partial-reduction arrays
};

struct GWork {  // ghost-cell work item: 256 ghost cells of one instance
  int inst;
  int start;
};

struct InstDev {
  Geo geo;
  BC bc;
  Sel sel;
  double* w;   // state between steps
  double* b;   // output of the first sweep
  double* w1;  // predictor
  long long nbound, slab0, slab1;
  int pbase2, ntiles2;  // sweep blocks (fused) or corrector tiles and their partials
  int ntx, ntm;         // TVD-MUSCL: corrector tiles of the x and the y sweep (partials from pbase2)
  int mhd;
  int seglen;           // fused sweeps: rows per marching segment
};

struct BScal {
  Scal s;
  int itmax;
  int done;
  int needdt;  // the next dt needs cell-centre Courant maxima (a dtcourant stayed unset)
  double smallpcoeff;
};

// dt of the next step from the lagged dtcourant and, where one is missing,
// the cell-centre maxima cmd[], then the stop condition (src/vac.t:73) and
// the tmaxexact clip (src/vac.t:83); thread 0 only
__device__ void batch_getdt_finish(BScal& S, const InstDev& I, const double* dtc,
                                   const double* cmd) {
  const int nd = I.geo.nd;
  double dt = kBig, cm = 0.0;
  for (int d = 0; d < nd; ++d) {
    if (dtc[d] < kBig) dt = dtc[d] < dt ? dtc[d] : dt;
    else cm = cm < cmd[d] ? cmd[d] : cm;
  }
  for (int d = 0; d < nd; ++d) S.s.dtcourant[d] = kBig;
  if (cm > kSmall) { double c = I.sel.courantpar / cm; dt = c < dt ? c : dt; }
  if (S.s.it >= S.itmax || S.s.t >= S.s.tmax || dt < 1.0e-99) {
    S.done = 1;
  } else {
    if (S.s.tmaxexact) { double c = S.s.tmax - S.s.t + kSmall; dt = c < dt ? c : dt; }
    S.s.dt = dt;
  }
}

// direction of sweep k in the instance's current step (src/vac.t:310-324)
__device__ __forceinline__ int sweep_dir(const BScal& S, const InstDev& I, int k) {
  const bool fwd = (S.s.it / 2) * 2 == S.s.it;
  return fwd ? k : I.geo.nd - 1 - k;
}

// after the last sweep of a step: the end of the step (src/vac.t:97-98) and
// the next dt; thread 0 only
__device__ void batch_step_tail(BScal& S, const InstDev& I, int k) {
  const int nd = I.geo.nd;
  if (k != nd - 1) return;
  S.s.it += 1;
  S.s.t = S.s.t + S.s.dt;
  double dtc[3] = {kBig, kBig, kBig}, cmd[3] = {0.0, 0.0, 0.0};
  bool lagged = true;
  for (int d = 0; d < nd; ++d) {
    dtc[d] = S.s.dtcourant[d];
    lagged &= dtc[d] < kBig;
  }
  if (lagged) batch_getdt_finish(S, I, dtc, cmd);
  else S.needdt = 1;
}

// after the TVDLF corrector of sweep k: the dtcourant update
// (src/vactvdlf.t:184-201) from the instance's Courant maximum cm, then
// batch_step_tail; thread 0 only
__device__ void batch_sweep_end(BScal& S, const InstDev& I, int k, double cm) {
  if (I.sel.courantpar > 0.0) {
    const int d = sweep_dir(S, I, k);
    if (S.s.dt * cm > 1.0) S.s.nerror[2] += 1;
    if (cm > kSmall) {
      double c = I.sel.courantpar / cm;
      S.s.dtcourant[d] = c < S.s.dtcourant[d] ? c : S.s.dtcourant[d];
    }
  }
  batch_step_tail(S, I, k);
}

// after the TVD-MUSCL corrector of sweep k: the dtcourant update of
// tvdlimit2 for the two extreme waves' Courant numbers (src/vactvd.t:145-157),
// then batch_step_tail; thread 0 only
__device__ void batch_sweep_end_tvd(BScal& S, const InstDev& I, int k, double c0, double c1) {
  if (I.sel.courantpar > 0.0) {
    const int d = sweep_dir(S, I, k);
    const double cm[2] = {c0, c1};
    for (int q = 0; q < 2; ++q) {
      if (cm[q] > 1.0) S.s.nerror[2] += 1;
      if (cm[q] > kSmall) {
        const double c = S.s.dt * I.sel.courantpar / cm[q];
        S.s.dtcourant[d] = c < S.s.dtcourant[d] ? c : S.s.dtcourant[d];
      }
    }
  }
  batch_step_tail(S, I, k);
}

// block-uniform early exit for finished instances
// A work list holds each instance's tiles together, neighbours adjacent;
// block L takes item xcd_order(L), so every XCD works through a contiguous
// stretch of the list and the halo rows its tiles share stay in its L2.
#define BATCH_PROLOGUE                                     \
  const Work wk = work[xcd_order(blockIdx.x, gridDim.x)];  \
  const InstDev& I = inst[wk.inst];           \
  const BScal& S = bs[wk.inst];               \
  if (S.done) return;

// sweep k (0: x-first instances do x, 1: ...) -- D is the direction this
// launch works on; src/dst roles: sweep 0 reads w and writes b, sweep 1
// reads b and writes w
template <int ND, bool MHD, class R, int D>
__global__ __launch_bounds__(512) void kb_pred(const InstDev* __restrict__ inst,
                                              const BScal* __restrict__ bs,
                                              const Work* __restrict__ work, int k) {
  BATCH_PROLOGUE
  const double* src = (k == 0) ? I.w : I.b;
  tvdlf_pred_body<ND, MHD, R, D>(I.geo, I.sel, src, I.w1, S.s.dt, S.s.smallp, wk.bx, wk.by, wk.bz);
}

template <int ND, bool MHD, class R, int D>
__global__ __launch_bounds__(512) void kb_corr(const InstDev* __restrict__ inst,
                                              const BScal* __restrict__ bs,
                                              const Work* __restrict__ work, int k,
                                              double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  const double* src = (k == 0) ? I.w : I.b;
  double* dst = (k == 0) ? I.b : I.w;
  const double cm = tvdlf_corr_body<ND, MHD, R, D>(I.geo, I.sel, src, I.w1, dst, S.s.dt,
                                                   S.s.smallp, wk.bx, wk.by, wk.bz);
  const double b = block_reduce<true>(cm, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + wk.pidx] = b;
}

// fused 2D TVDLF sweep of instances whose faces have one boundary type for all
// variables (the single-grid kernel body): work item (strip, block of four
// segments); sweep k reads w and writes b (k = 0) or reads b and writes w
// (k = 1); ghost cells of the output are written by the sweep itself.  With
// fold set, the last of an instance's blocks to finish (fold_courant_last
// on the instance's two words of folds, zero between launches) runs
// batch_sweep_end; the autotuner's timing launches pass fold = 0.
template <bool MHD, class R, int D, int LIM>
__global__ __launch_bounds__(256) void kb_fused(const InstDev* __restrict__ inst,
                                               BScal* __restrict__ bs,
                                               const Work* __restrict__ work, int k,
                                               unsigned long long* __restrict__ folds, int fold) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  const double* src = (k == 0) ? I.w : I.b;
  double* dst = (k == 0) ? I.b : I.w;
  const int seg = wk.by * 4 + (int)threadIdx.y;
  double cmx;
  if constexpr (sizeof(R) == 4)  // FP32-bulk
    cmx = D == 0 ? bulk_fx_body<MHD, LIM>(I.geo, I.sel, I.bc, src, dst, S.s.dt, S.s.smallp, wk.bx,
                                          seg, I.seglen)
                 : bulk_fy_body<MHD, LIM>(I.geo, I.sel, I.bc, src, dst, S.s.dt, S.s.smallp, wk.bx,
                                          seg, I.seglen);
  else if constexpr (D == 0)
    cmx = tvdlf_fx_body<MHD, R, LIM>(I.geo, I.sel, I.bc, src, dst, S.s.dt, S.s.smallp, wk.bx, seg,
                                     I.seglen);
  else
    cmx = tvdlf_fy_body<MHD, R, LIM>(I.geo, I.sel, I.bc, src, dst, S.s.dt, S.s.smallp, wk.bx, seg,
                                     I.seglen);
  const double bm = block_reduce<true>(cmx, sm);
  if (!fold || threadIdx.x != 0 || threadIdx.y != 0) return;
  double cm;
  if (fold_courant_last(folds + 2 * wk.inst, bm, (unsigned long long)I.ntiles2, &cm))
    batch_sweep_end(bs[wk.inst], I, k, cm);
}

// ghost cells of one array of every listed instance: which = 0 w, 1 b, 2 w1
template <int NW>
__global__ __launch_bounds__(256) void kb_bound(const InstDev* __restrict__ inst,
                                               const BScal* __restrict__ bs,
                                               const GWork* __restrict__ work, int which) {
  const GWork gw = work[xcd_order(blockIdx.x, gridDim.x)];
  const InstDev& I = inst[gw.inst];
  if (bs[gw.inst].done) return;
  const long long t = (long long)gw.start + threadIdx.x;
  if (t >= I.nbound) return;
  double* w = (which == 0) ? I.w : (which == 1 ? I.b : I.w1);
  bound_body<NW>(I.geo, I.bc, w, t, I.slab0, I.slab1);
}

__device__ double fold_range(const double* p, int n, bool is_max, double* sm) {
  double v = is_max ? -INFINITY : INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double u = p[i];
    v = is_max ? (v < u ? u : v) : (u < v ? u : v);
  }
  return is_max ? block_reduce<true>(v, sm) : block_reduce<false>(v, sm);
}

// predictor + corrector instances: one block per listed instance folds the
// corrector's partials of sweep k and runs batch_sweep_end
__global__ void kb_fin_courant(const InstDev* __restrict__ inst, BScal* __restrict__ bs,
                               const int* __restrict__ ids, const double* __restrict__ part,
                               long long pstride, int k) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done) return;
  const InstDev& I = inst[i];
  const double cm = fold_range(part + SLOT_C0 * pstride + I.pbase2, I.ntiles2, true, sm);
  if (threadIdx.x == 0) batch_sweep_end(S, I, k, cm);
}

// TVD-MUSCL corrector tiles of 2D instances (the single-grid bodies): x
// sweep strips of 62 faces by 4 rows, y sweep 64 columns by 4 segments of
// seglen rows; sweep k reads w (k = 0) or b with the predictor w1 and
// writes b or w (mesh cells; the ghost cells follow in a pass), and the two
// extreme waves' Courant numbers go to the partials
template <bool MHD, class R, int LIM>
__global__ __launch_bounds__(256) void kb_tvdmu_x(const InstDev* __restrict__ inst,
                                                 const BScal* __restrict__ bs,
                                                 const Work* __restrict__ work, int k,
                                                 double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  const double* src = (k == 0) ? I.w : I.b;
  double* dst = (k == 0) ? I.b : I.w;
  double c0, c1;
  tvdmu_x_body<2, R, LIM, MHD>(I.geo, I.sel, I.bc, src, I.w1, dst, 0, src, S.s.dt, S.s.smallp,
                               Tile{wk.bx, wk.by, 0}, c0, c1);
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + wk.pidx] = b0;
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + wk.pidx] = b1;
}

template <bool MHD, class R, int LIM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void kb_tvdmu_m(
    const InstDev* __restrict__ inst, const BScal* __restrict__ bs, const Work* __restrict__ work,
    int k, double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  const double* src = (k == 0) ? I.w : I.b;
  double* dst = (k == 0) ? I.b : I.w;
  double c0, c1;
  tvdmu_m_body<2, R, 1, LIM, MHD>(I.geo, I.sel, I.bc, src, I.w1, dst, I.seglen, 0, src, S.s.dt,
                                  S.s.smallp, Tile{wk.bx, wk.by, 0}, c0, c1);
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + wk.pidx] = b0;
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + wk.pidx] = b1;
}

// TVD-MUSCL instances: one block per listed instance folds the two partial
// slots of sweep k and runs batch_sweep_end_tvd
__global__ void kb_fin_courant_tvd(const InstDev* __restrict__ inst, BScal* __restrict__ bs,
                                   const int* __restrict__ ids, const double* __restrict__ part,
                                   long long pstride, int k) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done) return;
  const InstDev& I = inst[i];
  const int n = sweep_dir(S, I, k) == 0 ? I.ntx : I.ntm;
  const double c0 = fold_range(part + SLOT_C0 * pstride + I.pbase2, n, true, sm);
  __syncthreads();
  const double c1 = fold_range(part + SLOT_C1 * pstride + I.pbase2, n, true, sm);
  if (threadIdx.x == 0) batch_sweep_end_tvd(S, I, k, c0, c1);
}

// The next dt of instances whose sweeps left a dtcourant unset (a Courant
// maximum at or below smalldouble): the cell-centre maxima over the mesh
// (src/vac.t:691-700) in one block per instance, then batch_getdt_finish.
// Returns at once for every other instance.
template <bool MHD>
__device__ void getdt_rare_body(BScal& S, const InstDev& I, double* sm) {
  constexpr int ND = 2;
  const Geo& g = I.geo;
  const long long ncell = (long long)g.n[0] * g.n[1];
  double dtc[3] = {kBig, kBig, kBig}, cmd[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ND; ++d) dtc[d] = S.s.dtcourant[d];
  for (int d = 0; d < ND; ++d) {
    if (dtc[d] < kBig) continue;  // block-uniform
    double v = -INFINITY;
    for (long long t = threadIdx.x; t < ncell; t += blockDim.x) {
      const int c0 = 2 + (int)(t % g.n[0]), c1 = 2 + (int)(t / g.n[0]);
      const int cc[2] = {c0, c1};
      double u[Layout<ND, MHD>::nw];
      load<Layout<ND, MHD>::nw>(I.w, g.ps, g.off + c0 + (long long)c1 * g.s[1], u);
      const double q = pdiv(cmax<ND, MHD, double>(u, d, I.sel.gamma, S.s.smallp), g.dx[d][cc[d]]);
      v = v < q ? q : v;
    }
    cmd[d] = block_reduce<true>(v, sm);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    batch_getdt_finish(S, I, dtc, cmd);
    S.needdt = 0;
  }
}

__global__ __launch_bounds__(256) void kb_getdt_rare(const InstDev* __restrict__ inst,
                                                    BScal* __restrict__ bs,
                                                    const int* __restrict__ ids) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done || !S.needdt) return;
  if (inst[i].mhd) getdt_rare_body<true>(S, inst[i], sm);
  else getdt_rare_body<false>(S, inst[i], sm);
}

// first dt of a newly admitted instance, after k_pminmax / k_fin_smallp /
// k_cmaxdx on its own partials: fold the cell-centre maxima and run
// batch_getdt_finish (every dtcourant is unset at admission)
__global__ void kb_first_getdt(InstDev I, BScal* __restrict__ S, const double* __restrict__ part,
                               long long pstride, int nb) {
  __shared__ double sm[8];
  double dtc[3] = {kBig, kBig, kBig}, cmd[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < I.geo.nd; ++d) {
    cmd[d] = fold_range(part + (SLOT_DX + d) * pstride, nb, true, sm);
    __syncthreads();
  }
  if (threadIdx.x == 0) batch_getdt_finish(*S, I, dtc, cmd);
}

// ---------------------------------------------------------------- host
enum { NCLASS = 4, NSNAP = 3 };
const int kTileRows[] = {1, 2, 4, 8};
const int kSegLens[] = {8, 16, 32, 64};

// kinds of work list: fused TVDLF sweep blocks; TVDLF predictor + corrector
// tiles; TVD-MUSCL predictor tiles, x-corrector strips, y-corrector columns
enum { KIND_FUSED = 0, KIND_TILE = 1, KIND_MUPRED = 2, KIND_MUX = 3, KIND_MUM = 4 };
constexpr int MU_SEGLEN = 16;  // rows per marching segment of the TVD-MUSCL y corrector

// kernel instance of a work list: physics, tier, kind and (fused sweeps and
// TVD-MUSCL correctors) the limiter type
struct Key {
  int mhd, f32, lim, kind;
  bool operator<(const Key& o) const {
    if (mhd != o.mhd) return mhd < o.mhd;
    if (f32 != o.f32) return f32 < o.f32;
    if (lim != o.lim) return lim < o.lim;
    return kind < o.kind;
  }
};

struct HostInst {
  bool used = false;
  vacgpu_config cfg{};
  vacgpu_method m{};
  InstDev dev{};
  long long bytes = 0;
  int klass = 0;
  bool joined = false;  // in the work lists (from the rebuild after admission)
  int phase_class = 0;  // (it - batch step) mod 2 when it joined
  long long join_seq = 0;  // first status snapshot that shows this instance
  int it0 = 0;
  double t0 = 0.0;
  dim3 tiles;           // predictor / corrector tile grid of an unfused instance
  bool fused = false;   // sweeps by kb_fused (else predictor + corrector tiles)
  int lim = -1;         // the common limiter type of a fused or TVD-MUSCL instance
  bool tvdmu = false;   // TVD-MUSCL (else TVDLF)
};

struct Launch {          // one work list on the device
  Work* d = nullptr;
  int n = 0;
  int cap = 0;
};

// pinned host buffer for state uploads and result copies; free for reuse
// when not held by a result and its last copy (event) has completed
struct Stage {
  double* p = nullptr;
  size_t bytes = 0;
  hipEvent_t ev = nullptr;
  bool recorded = false;
  bool held = false;
};

// memory and slots of removed instances, and superseded device buffers and
// graphs, released once the event (recorded after every step and copy
// enqueued before the removal) has completed
struct Retired {
  hipEvent_t ev = nullptr, evc = nullptr;
  int slot = -1;
  long long bytes = 0;
  std::vector<std::pair<size_t, double*>> blocks;
  std::vector<void*> frees;
  std::vector<hipGraphExec_t> graphs;
};

struct Fetch {
  bool used = false;
  int stage = -1;
  int nw = 0, ng0 = 0, ng1 = 0;
};

// Host copies into and out of the pinned stages (states of up to hundreds of
// MB) split over a few persistent threads and the caller: one thread's
// memcpy, not the PCIe link, would otherwise bound admission and result
// delivery.  One job at a time (the batch is not thread-safe anyway).
class CopyPool {
 public:
  explicit CopyPool(int workers) {
    for (int i = 0; i < workers; ++i) th_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // n rows of row bytes, source rows spitch apart, destination rows dpitch apart
  void copy(char* dst, size_t dpitch, const char* src, size_t spitch, size_t row, size_t n) {
    const size_t total = row * n;
    if (total < (size_t(4) << 20) || th_.empty()) {
      copy_rows(dst, dpitch, src, spitch, row, 0, n);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = dst; src_ = src; dp_ = dpitch; sp_ = spitch; row_ = row; n_ = n;
      parts_ = (int)th_.size() + 1;
      next_ = 0;
      left_ = parts_;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return left_ == 0; });
  }

 private:
  static void copy_rows(char* dst, size_t dp, const char* src, size_t sp, size_t row, size_t r0,
                        size_t r1) {
    if (dp == row && sp == row) {
      std::memcpy(dst + r0 * row, src + r0 * row, (r1 - r0) * row);
      return;
    }
    for (size_t r = r0; r < r1; ++r) std::memcpy(dst + r * dp, src + r * sp, row);
  }
  void work() {
    for (;;) {
      char* d;
      const char* s;
      size_t dp, sp, row, r0, r1;
      {
        std::lock_guard<std::mutex> lk(m_);
        if (next_ >= parts_) return;
        const int k = next_++;
        d = dst_; s = src_; dp = dp_; sp = sp_; row = row_;
        if (dp == row && sp == row) {  // contiguous: split bytes, in whole rows of 64 KiB blocks
          const size_t tot = row * n_, per = (tot / parts_ + 65535) / 65536 * 65536;
          r0 = std::min(tot, per * k);
          r1 = std::min(tot, per * (k + 1));
          row = 1;
          dp = sp = 1;
        } else {
          const size_t per = (n_ + parts_ - 1) / parts_;
          r0 = std::min(n_, per * k);
          r1 = std::min(n_, per * (k + 1));
        }
      }
      if (r1 > r0) copy_rows(d, dp, s, sp, row, r0, r1);
      std::lock_guard<std::mutex> lk(m_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void loop() {
    unsigned long long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t dp_ = 0, sp_ = 0, row_ = 0, n_ = 0;
  int parts_ = 0, next_ = 0, left_ = 0;
  unsigned long long gen_ = 0;
  bool stop_ = false;
};

}  // namespace

struct vacgpu_batch {
  long long budget = 0, used_bytes = 0;
  std::vector<HostInst> inst;
  std::vector<char> slot_busy;  // removed, not yet released
  InstDev* d_inst = nullptr;
  BScal* d_bs = nullptr;
  unsigned long long* d_fold = nullptr;  // per-instance Courant fold words of the fused sweeps
  int cap = 0;
  double* part = nullptr;
  long long pstride = 0;
  int rows[NCLASS] = {4, 4, 4, 4};       // tile rows per size class (autotuned)
  int seglen[NCLASS] = {16, 16, 16, 16};  // fused sweeps: rows per segment per size class
  bool use_fused = true;
  long long step_count = 0;
  bool dirty = true;
  // work lists per (kernel key, tile rows or segment length) and phase class
  struct Lists {
    Launch ph[2];
  };
  std::map<std::pair<Key, int>, Lists> lists;
  struct { GWork* d = nullptr; int n = 0, cap = 0; } gw4, gw6;  // ghost work per nw
  int* d_ids = nullptr;    // every joined instance
  int* d_idsu = nullptr;   // joined TVDLF instances swept by predictor + corrector tiles
  int* d_idst = nullptr;   // joined TVD-MUSCL instances
  int nids = 0, nidsu = 0, nidst = 0, ids_cap = 0;
  std::multimap<size_t, double*> pool;  // reusable device blocks by size
  hipStream_t stream = nullptr;         // steps
  hipStream_t cstream = nullptr;        // admissions and result copies
  bool use_graph = false;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // per step parity
  // admissions since the last rebuild: the step stream waits for ev_admit
  hipEvent_t ev_admit = nullptr;
  bool admit_pending = false;
  std::vector<Stage> stages;
  // rebuild uploads: two pinned arenas used in turn
  struct Arena { char* h = nullptr; size_t cap = 0; hipEvent_t ev = nullptr; bool recorded = false; } arena[2];
  int arena_cur = 0;
  // status snapshots
  BScal* h_snap[NSNAP] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_snap[NSNAP] = {nullptr, nullptr, nullptr};
  long long snap_seqno[NSNAP] = {-1, -1, -1};
  std::deque<int> snap_queue;  // mirrors with a copy in flight, oldest first
  int snap_cur = -1;
  long long snap_next = 0;
  std::deque<Retired> retired;
  std::vector<Fetch> fetches;
  std::vector<hipEvent_t> evpool;
  std::vector<InstDev> hdev;  // host image of d_inst
  CopyPool* copier = nullptr;
  std::string err;
};

namespace {

int bfail(vacgpu_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}

#define BHIP(b, call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail(b, VACGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device blocks are kept in size classes (64 KiB steps up to 1 MiB, then
// four classes per power of two) so that freed blocks serve later instances
// of similar size; the class size is what an instance occupies.
size_t pool_round(size_t bytes) {
  const size_t k64 = size_t(1) << 16;
  if (bytes <= (size_t(1) << 20)) return (bytes + k64 - 1) / k64 * k64;
  int lg = 63 - __builtin_clzll((unsigned long long)bytes);
  const size_t q = size_t(1) << (lg - 2);
  return (bytes + q - 1) / q * q;
}

double* pool_get(vacgpu_batch* b, size_t bytes) {
  auto it = b->pool.find(bytes);
  if (it != b->pool.end()) {
    double* p = it->second;
    b->pool.erase(it);
    return p;
  }
  double* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void pool_put(vacgpu_batch* b, size_t bytes, double* p) {
  if (p) b->pool.emplace(bytes, p);
}

int size_class(const vacgpu_config& c) {
  long long cells = (long long)c.nx[0] * c.nx[1];
  if (cells < 16384) return 0;
  if (cells < 65536) return 1;
  if (cells < 262144) return 2;
  return 3;
}

Geo make_geo(const vacgpu_config& c) {
  Geo geo{};
  const int nd = c.ndim;
  geo.nd = nd;
  for (int d = 0; d < 3; ++d) {
    geo.n[d] = d < nd ? c.nx[d] : 1;
    geo.ng[d] = d < nd ? c.nx[d] + 4 : 1;
  }
  const long long px = ((geo.ng[0] + 15) / 16) * 16 + 16;
  geo.s[0] = 1;
  geo.s[1] = px;
  geo.s[2] = 0;
  geo.off = 14;
  geo.ps = ((geo.off + px * geo.ng[1] + 31) / 32) * 32;
  return geo;
}

size_t plane_bytes_of(const Geo& g, int nw) { return pool_round(sizeof(double) * (size_t)g.ps * nw); }
size_t dx_bytes_of(const Geo& g) { return pool_round(sizeof(double) * (size_t)(g.ng[0] + g.ng[1])); }

long long footprint(const vacgpu_config& c) {
  const Geo g = make_geo(c);
  const int nw = c.physics == VACGPU_MHD ? 2 + 2 * c.ndim : 2 + c.ndim;
  return (long long)(3 * plane_bytes_of(g, nw) + dx_bytes_of(g));
}

int ev_get(vacgpu_batch* b, hipEvent_t* e) {
  if (!b->evpool.empty()) {
    *e = b->evpool.back();
    b->evpool.pop_back();
    return 0;
  }
  BHIP(b, hipEventCreateWithFlags(e, hipEventDisableTiming));
  return 0;
}

bool ev_done(hipEvent_t e) { return hipEventQuery(e) != hipErrorNotReady; }

// release what retired entries hold, oldest first, as far as their events
// have completed (or all of them, waiting, with wait set)
int reap(vacgpu_batch* b, bool wait = false) {
  while (!b->retired.empty()) {
    Retired& r = b->retired.front();
    if (wait) BHIP(b, hipEventSynchronize(r.ev));
    else if (!ev_done(r.ev)) break;
    for (auto& kv : r.blocks) pool_put(b, kv.first, kv.second);
    for (void* p : r.frees) (void)hipFree(p);
    for (hipGraphExec_t g : r.graphs) (void)hipGraphExecDestroy(g);
    if (r.slot >= 0) b->slot_busy[r.slot] = 0;
    b->used_bytes -= r.bytes;
    b->evpool.push_back(r.ev);
    b->evpool.push_back(r.evc);
    b->retired.pop_front();
  }
  return 0;
}

// queue r for release after every step and copy enqueued so far
int retire(vacgpu_batch* b, Retired&& r) {
  int rc;
  if ((rc = ev_get(b, &r.evc)) || (rc = ev_get(b, &r.ev))) return rc;
  BHIP(b, hipEventRecord(r.evc, b->stream));
  BHIP(b, hipStreamWaitEvent(b->cstream, r.evc, 0));
  BHIP(b, hipEventRecord(r.ev, b->cstream));
  b->retired.push_back(std::move(r));
  return 0;
}

// a free pinned stage of at least bytes (the smallest that fits), else a new one
int stage_get(vacgpu_batch* b, size_t bytes, int* idx) {
  int best = -1;
  for (int i = 0; i < (int)b->stages.size(); ++i) {
    Stage& S = b->stages[i];
    if (S.held || S.bytes < bytes) continue;
    if (S.recorded && !ev_done(S.ev)) continue;
    if (best < 0 || S.bytes < b->stages[best].bytes) best = i;
  }
  if (best < 0) {
    Stage S;
    size_t sz = size_t(1) << 20;
    while (sz < bytes) sz <<= 1;
    BHIP(b, hipHostMalloc(&S.p, sz, hipHostMallocDefault));
    S.bytes = sz;
    int rc;
    if ((rc = ev_get(b, &S.ev))) { (void)hipHostFree(S.p); return rc; }
    b->stages.push_back(S);
    best = (int)b->stages.size() - 1;
  }
  *idx = best;
  return 0;
}

int ensure_dev(vacgpu_batch* b, void** d, int* cap, int n, size_t elem, std::vector<void*>& old) {
  if (n <= *cap) return 0;
  if (*d) old.push_back(*d);
  *d = nullptr;
  *cap = n * 2 + 16;
  BHIP(b, hipMalloc(d, elem * (size_t)*cap));
  return 0;
}

// Rebuild work lists, partial ranges and the id lists after a change of
// membership, tile shapes or kernels.  Nothing here waits for the device:
// uploads go through a pinned arena on the step stream (the arena used two
// rebuilds ago is reused once its copies have completed), superseded device
// buffers and graphs are retired.
int rebuild(vacgpu_batch* b) {
  int rc;
  if ((rc = reap(b))) return rc;
  if (b->admit_pending) {
    BHIP(b, hipStreamWaitEvent(b->stream, b->ev_admit, 0));
    b->admit_pending = false;
  }
  std::map<std::pair<Key, int>, std::vector<Work>> ph0, ph1;
  std::vector<GWork> g4, g6;
  std::vector<int> ids, idsu, idst;
  int pbase2 = 0;
  int maxid = -1;
  for (int i = 0; i < (int)b->inst.size(); ++i) {
    HostInst& h = b->inst[i];
    if (!h.used) continue;
    if (!h.joined) {
      h.joined = true;
      h.phase_class = (int)(((h.it0 - b->step_count) % 2 + 2) % 2);
      h.join_seq = b->snap_next;
    }
    const Geo& g = h.dev.geo;
    const bool fused = h.fused && b->use_fused;
    const Key kr{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK, -1, KIND_TILE};
    auto& phl = h.phase_class ? ph1 : ph0;
    h.dev.pbase2 = pbase2;
    if (h.tvdmu) {
      // predictor tiles (64 x rows), x-corrector strips (62 faces x 4 rows),
      // y-corrector columns (64 x 4 segments of MU_SEGLEN rows)
      const int rows = b->rows[h.klass];
      const Key kp{kr.mhd, kr.f32, -1, KIND_MUPRED};
      auto& vp = phl[std::make_pair(kp, rows)];
      const int px = (g.n[0] + BX - 1) / BX, py = (g.n[1] + rows - 1) / rows;
      for (int by = 0; by < py; ++by)
        for (int bx = 0; bx < px; ++bx) vp.push_back(Work{i, bx, by, 0, 0});
      const Key kx{kr.mhd, kr.f32, h.lim, KIND_MUX};
      auto& vx = phl[std::make_pair(kx, BY)];
      const int xx = (g.n[0] + CX_STRIDE - 1) / CX_STRIDE, xy = (g.n[1] + BY - 1) / BY;
      int t = 0;
      for (int by = 0; by < xy; ++by)
        for (int bx = 0; bx < xx; ++bx, ++t) vx.push_back(Work{i, bx, by, 0, pbase2 + t});
      h.dev.ntx = t;
      const Key km{kr.mhd, kr.f32, h.lim, KIND_MUM};
      auto& vm = phl[std::make_pair(km, MU_SEGLEN)];
      const int mx = (g.n[0] + BX - 1) / BX;
      const int my = ((g.n[1] + MU_SEGLEN - 1) / MU_SEGLEN + BY - 1) / BY;
      t = 0;
      for (int by = 0; by < my; ++by)
        for (int bx = 0; bx < mx; ++bx, ++t) vm.push_back(Work{i, bx, by, 0, pbase2 + t});
      h.dev.ntm = t;
      h.dev.seglen = MU_SEGLEN;
      h.dev.ntiles2 = h.dev.ntx > h.dev.ntm ? h.dev.ntx : h.dev.ntm;
      for (long long q = 0; q < h.dev.nbound; q += 256) (kr.mhd ? g6 : g4).push_back(GWork{i, (int)q});
      idst.push_back(i);
      pbase2 += h.dev.ntiles2;
    } else if (fused) {
      // sweep blocks: strips of 62 x cells, blocks of four segments
      const int sl = b->seglen[h.klass];
      const int strips = (g.n[0] + FS_STRIDE - 1) / FS_STRIDE;
      const int segs = (g.n[1] + sl - 1) / sl;
      const int sblocks = (segs + 3) / 4;
      h.dev.seglen = sl;
      h.dev.ntiles2 = strips * sblocks;
      const Key kf{kr.mhd, kr.f32, h.lim, KIND_FUSED};
      auto& v = phl[std::make_pair(kf, sl)];
      for (int sb = 0; sb < sblocks; ++sb)
        for (int st = 0; st < strips; ++st) v.push_back(Work{i, st, sb, 0, 0});
    } else {
      const int rows = b->rows[h.klass];
      h.tiles = dim3((g.n[0] + BX - 1) / BX, (g.n[1] + rows - 1) / rows, 1);
      h.dev.seglen = 0;
      h.dev.ntiles2 = (int)(h.tiles.x * h.tiles.y);
      auto& v = phl[std::make_pair(kr, rows)];
      int t = 0;
      for (unsigned by = 0; by < h.tiles.y; ++by)
        for (unsigned bx = 0; bx < h.tiles.x; ++bx, ++t) v.push_back(Work{i, (int)bx, (int)by, 0, pbase2 + t});
      // ghost cells are filled by separate launches only for unfused instances
      for (long long q = 0; q < h.dev.nbound; q += 256) (kr.mhd ? g6 : g4).push_back(GWork{i, (int)q});
      idsu.push_back(i);
      pbase2 += h.dev.ntiles2;
    }
    ids.push_back(i);
    b->hdev[i] = h.dev;
    maxid = i;
  }
  // device buffers (grown geometrically; old ones retired)
  Retired old;
  if (pbase2 > b->pstride) {
    if (b->part) old.frees.push_back(b->part);
    b->part = nullptr;
    b->pstride = (long long)pbase2 * 2 + 64;
    BHIP(b, hipMalloc(&b->part, sizeof(double) * 2 * b->pstride));  // slots C0, C1
  }
  std::set<std::pair<Key, int>> keys;
  for (auto& kv : ph0) keys.insert(kv.first);
  for (auto& kv : ph1) keys.insert(kv.first);
  for (auto& kv : b->lists) { kv.second.ph[0].n = 0; kv.second.ph[1].n = 0; }
  size_t total = sizeof(InstDev) * (size_t)(maxid + 1) + 256;
  for (const auto& k : keys) {
    auto& L = b->lists[k];
    const std::vector<Work>* src[2] = {&ph0[k], &ph1[k]};
    for (int c = 0; c < 2; ++c) {
      void* d = L.ph[c].d;
      if ((rc = ensure_dev(b, &d, &L.ph[c].cap, (int)src[c]->size(), sizeof(Work), old.frees))) return rc;
      L.ph[c].d = static_cast<Work*>(d);
      L.ph[c].n = (int)src[c]->size();
      total += sizeof(Work) * src[c]->size() + 256;
    }
  }
  {
    void* d = b->gw4.d;
    if ((rc = ensure_dev(b, &d, &b->gw4.cap, (int)g4.size(), sizeof(GWork), old.frees))) return rc;
    b->gw4.d = static_cast<GWork*>(d);
    d = b->gw6.d;
    if ((rc = ensure_dev(b, &d, &b->gw6.cap, (int)g6.size(), sizeof(GWork), old.frees))) return rc;
    b->gw6.d = static_cast<GWork*>(d);
    b->gw4.n = (int)g4.size();
    b->gw6.n = (int)g6.size();
    total += sizeof(GWork) * (g4.size() + g6.size()) + 512;
  }
  if ((int)ids.size() > b->ids_cap) {
    if (b->d_ids) old.frees.push_back(b->d_ids);
    if (b->d_idsu) old.frees.push_back(b->d_idsu);
    if (b->d_idst) old.frees.push_back(b->d_idst);
    b->d_ids = b->d_idsu = b->d_idst = nullptr;
    b->ids_cap = (int)ids.size() * 2 + 16;
    BHIP(b, hipMalloc(&b->d_ids, sizeof(int) * b->ids_cap));
    BHIP(b, hipMalloc(&b->d_idsu, sizeof(int) * b->ids_cap));
    BHIP(b, hipMalloc(&b->d_idst, sizeof(int) * b->ids_cap));
  }
  b->nids = (int)ids.size();
  b->nidsu = (int)idsu.size();
  b->nidst = (int)idst.size();
  total += sizeof(int) * (ids.size() + idsu.size() + idst.size()) + 768;
  // uploads through the arena
  vacgpu_batch::Arena& A = b->arena[b->arena_cur];
  b->arena_cur ^= 1;
  if (A.recorded) BHIP(b, hipEventSynchronize(A.ev));
  A.recorded = false;
  if (A.cap < total) {
    if (A.h) BHIP(b, hipHostFree(A.h));
    A.h = nullptr;
    A.cap = total * 2;
    BHIP(b, hipHostMalloc(reinterpret_cast<void**>(&A.h), A.cap, hipHostMallocDefault));
  }
  if (!A.ev && (rc = ev_get(b, &A.ev))) return rc;
  size_t off = 0;
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!bytes) return 0;
    std::memcpy(A.h + off, src, bytes);
    BHIP(b, hipMemcpyAsync(dst, A.h + off, bytes, hipMemcpyHostToDevice, b->stream));
    off += (bytes + 255) / 256 * 256;
    return 0;
  };
  if (maxid >= 0 && (rc = up(b->d_inst, b->hdev.data(), sizeof(InstDev) * (size_t)(maxid + 1)))) return rc;
  for (const auto& k : keys) {
    auto& L = b->lists[k];
    if ((rc = up(L.ph[0].d, ph0[k].data(), sizeof(Work) * ph0[k].size()))) return rc;
    if ((rc = up(L.ph[1].d, ph1[k].data(), sizeof(Work) * ph1[k].size()))) return rc;
  }
  if ((rc = up(b->gw4.d, g4.data(), sizeof(GWork) * g4.size()))) return rc;
  if ((rc = up(b->gw6.d, g6.data(), sizeof(GWork) * g6.size()))) return rc;
  if ((rc = up(b->d_ids, ids.data(), sizeof(int) * ids.size()))) return rc;
  if ((rc = up(b->d_idsu, idsu.data(), sizeof(int) * idsu.size()))) return rc;
  if ((rc = up(b->d_idst, idst.data(), sizeof(int) * idst.size()))) return rc;
  BHIP(b, hipEventRecord(A.ev, b->stream));
  A.recorded = true;
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) { old.graphs.push_back(b->graph[p]); b->graph[p] = nullptr; }
  if (!old.frees.empty() || !old.graphs.empty())
    if ((rc = retire(b, std::move(old)))) return rc;
  b->dirty = false;
  return 0;
}

#define BDISPATCH(MHDV, F32V, ...)                                                  \
  do {                                                                              \
    if (MHDV) { constexpr bool MHD = true;                                          \
      if (F32V) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
    else { constexpr bool MHD = false;                                              \
      if (F32V) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
  } while (0)

int launch_err(vacgpu_batch* b, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, VACGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

int bound_all(vacgpu_batch* b, int which) {
  if (b->gw4.n) hipLaunchKernelGGL(kb_bound<4>, dim3(b->gw4.n), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->gw4.d, which);
  if (b->gw6.n) hipLaunchKernelGGL(kb_bound<6>, dim3(b->gw6.n), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->gw6.d, which);
  return launch_err(b, "kb_bound");
}

// enqueue one batched step; parity = step_count % 2 selects which phase class
// sweeps x first
int enqueue_step(vacgpu_batch* b, int parity) {
  constexpr int ND = 2;
  int rc;
  // dt of instances whose last sweeps left a dtcourant unset (returns at
  // once for the others; every other dt was set by the previous step's end)
  if (b->nids) hipLaunchKernelGGL(kb_getdt_rare, dim3(b->nids), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_ids);
  if ((rc = launch_err(b, "kb_getdt_rare"))) return rc;
  // sweeps: phase class c has even it when (c + parity) is even
  for (int k = 0; k < ND; ++k) {
    for (int pass = 0; pass < 2; ++pass) {  // 0 predictors, 1 correctors
      for (auto& kv : b->lists) {
        const auto& L = kv.second;
        const Key& key = kv.first.first;
        const int shape = kv.first.second;
        const int lim = key.lim;
        for (int c = 0; c < 2; ++c) {
          const Launch& W = L.ph[c];
          if (!W.n) continue;
          const bool fwd = ((c + parity) % 2) == 0;
          const int D = fwd ? k : ND - 1 - k;
          const dim3 fb(64, 4);
          if (key.kind == KIND_FUSED) {  // one launch, in the corrector pass
            if (pass == 1) {
#define KBF(DV, LV) hipLaunchKernelGGL((kb_fused<MHD, R, DV, LV>), dim3(W.n), fb, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->d_fold, 1)
#define KBFL(DV) do { if (lim == LIM_MINMOD) KBF(DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) KBF(DV, LIM_WOODWARD); else KBF(DV, LIM_SUPERBEE); } while (0)
              BDISPATCH(key.mhd, key.f32, {
                if (D == 0) KBFL(0); else KBFL(1);
              });
#undef KBFL
#undef KBF
            }
          } else if (key.kind == KIND_TILE || key.kind == KIND_MUPRED) {
            const dim3 blk(BX, shape);
            BDISPATCH(key.mhd, key.f32, {
              if (pass == 0) {
                if (D == 0) hipLaunchKernelGGL((kb_pred<ND, MHD, R, 0>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k);
                else hipLaunchKernelGGL((kb_pred<ND, MHD, R, 1>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k);
              } else if (key.kind == KIND_TILE) {
                if (D == 0) hipLaunchKernelGGL((kb_corr<ND, MHD, R, 0>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride);
                else hipLaunchKernelGGL((kb_corr<ND, MHD, R, 1>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride);
              }
            });
          } else if (pass == 1 && ((key.kind == KIND_MUX && D == 0) || (key.kind == KIND_MUM && D == 1))) {
#define KMU(KN, LV) hipLaunchKernelGGL((KN<MHD, R, LV>), dim3(W.n), fb, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride)
#define KMUL(KN) do { if (lim == LIM_MINMOD) KMU(KN, LIM_MINMOD); else if (lim == LIM_WOODWARD) KMU(KN, LIM_WOODWARD); else KMU(KN, LIM_SUPERBEE); } while (0)
            BDISPATCH(key.mhd, key.f32, {
              if (key.kind == KIND_MUX) KMUL(kb_tvdmu_x); else KMUL(kb_tvdmu_m);
            });
#undef KMUL
#undef KMU
          }
        }
      }
      if ((rc = launch_err(b, pass ? "batch correctors" : "batch predictors"))) return rc;
      if (pass == 0) {
        if ((rc = bound_all(b, 2))) return rc;
      } else {
        if (b->nidsu) hipLaunchKernelGGL(kb_fin_courant, dim3(b->nidsu), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_idsu, b->part, b->pstride, k);
        if (b->nidst) hipLaunchKernelGGL(kb_fin_courant_tvd, dim3(b->nidst), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_idst, b->part, b->pstride, k);
        if ((rc = launch_err(b, "kb_fin_courant"))) return rc;
        if ((rc = bound_all(b, k == 0 ? 1 : 0))) return rc;
      }
    }
  }
  return 0;
}

// wait for the snapshot copy of mirror m and make it current
int snap_take(vacgpu_batch* b) {
  const int m = b->snap_queue.front();
  BHIP(b, hipEventSynchronize(b->ev_snap[m]));
  b->snap_queue.pop_front();
  b->snap_cur = m;
  return 0;
}

int sync_all(vacgpu_batch* b) {
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipStreamSynchronize(b->cstream));
  return reap(b, true);
}

}  // namespace

extern "C" {

long long vacgpu_batch_footprint(const vacgpu_config* cfg) {
  if (!cfg) return -1;
  return footprint(*cfg);
}

int vacgpu_batch_create(long long hbm_budget_bytes, int capacity, vacgpu_batch** out) {
  if (!out || capacity <= 0) return VACGPU_E_ARG;
  *out = nullptr;
  vacgpu_batch* b = new vacgpu_batch();
  b->budget = hbm_budget_bytes;
  b->cap = capacity;
  b->inst.resize(capacity);
  b->slot_busy.assign(capacity, 0);
  b->hdev.resize(capacity);
  {
    const unsigned hc = std::thread::hardware_concurrency();
    b->copier = new CopyPool(hc >= 16 ? 7 : (hc > 2 ? (int)hc / 2 - 1 : 0));
  }
  hipError_t e = hipMalloc(&b->d_inst, sizeof(InstDev) * capacity);
  if (e == hipSuccess) e = hipMalloc(&b->d_bs, sizeof(BScal) * capacity);
  if (e == hipSuccess) e = hipMalloc(&b->d_fold, 2 * sizeof(unsigned long long) * capacity);
  for (int m = 0; m < NSNAP && e == hipSuccess; ++m) {
    e = hipHostMalloc(&b->h_snap[m], sizeof(BScal) * capacity, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_snap[m], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_admit, hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    std::vector<BScal> init(capacity);
    for (auto& x : init) { std::memset(&x, 0, sizeof(BScal)); x.done = 1; }
    e = hipMemcpy(b->d_bs, init.data(), sizeof(BScal) * capacity, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMemset(b->d_fold, 0, 2 * sizeof(unsigned long long) * capacity);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    fprintf(stderr, "vacgpu_batch_create: %s\n", hipGetErrorString(e));
    vacgpu_batch_destroy(b);
    return VACGPU_E_HIP;
  }
  *out = b;
  return 0;
}

}  // extern "C"

namespace {

// admission; mesh = 1: w is the mesh only, (nw, ny, nx) tight, and the
// ghost cells come from the device fill alone
int batch_add(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
              const double* dx0, const double* dx1, const double* w, const int* ld, int mesh,
              int it, double t, int itmax, double tmax, int tmaxexact, int* slot) {
  if (!b || !cfg || !m || !dx0 || !dx1 || !w || !slot) return VACGPU_E_ARG;
  if (cfg->ndim != 2) return bfail(b, VACGPU_E_ARG, "the batch engine runs 2D grids");
  if (m->typefull != VACGPU_TVDLF && m->typefull != VACGPU_TVDMU)
    return bfail(b, VACGPU_E_ARG, "the batch engine runs typefull='tvdlf' and 'tvdmu'");
  if (m->dimsplit_xy) return bfail(b, VACGPU_E_ARG, "the batch engine runs typedimsplit='xyyx'");
  if (m->typelimited != VACGPU_LIM_ORIGINAL) return bfail(b, VACGPU_E_ARG, "the batch engine runs typelimited='original'");
  if (m->useprimitive) return bfail(b, VACGPU_E_ARG, "the batch engine limits the conserved variables (useprimitive=F)");
  if (m->unsplit) return bfail(b, VACGPU_E_ARG, "the batch engine runs dimensionally split sweeps");
  if (m->dtcantgrow || m->slowsteps)
    return bfail(b, VACGPU_E_ARG, "the batch engine's time step has no dtcantgrow / slowsteps");
  if (m->courantpar <= 0.0) return bfail(b, VACGPU_E_ARG, "the batch engine needs courantpar > 0");
  if (vacgpu_supported(cfg, m) != 0) return bfail(b, VACGPU_E_ARG, "unsupported selectors (see vacgpu_supported)");
  const int nwv = cfg->physics == VACGPU_MHD ? 6 : 4;
  if (m->typefull == VACGPU_TVDMU) {
    for (int iw = 0; iw < nwv; ++iw) {
      if (m->typeentropy[iw] != VACGPU_ENT_NUL)
        return bfail(b, VACGPU_E_ARG, "the batch engine's TVD-MUSCL has no entropy fix");
      if (m->typelimiter[iw] != m->typelimiter[0])
        return bfail(b, VACGPU_E_ARG, "the batch engine's TVD-MUSCL takes one limiter for every wave");
    }
    if (m->typelimiter[0] != LIM_MINMOD && m->typelimiter[0] != LIM_WOODWARD && m->typelimiter[0] != LIM_SUPERBEE)
      return bfail(b, VACGPU_E_ARG, "the batch engine's TVD-MUSCL takes minmod, woodward or superbee");
  }
  for (int f = 0; f < 4; ++f)
    for (int iw = 0; iw < 8; ++iw)
      if (cfg->typeB[f][iw] != VACGPU_CONT && cfg->typeB[f][iw] != VACGPU_PERIODIC)
        return bfail(b, VACGPU_E_ARG, "the batch engine runs 'cont' and 'periodic' boundaries");
  int rc;
  if ((rc = reap(b))) return rc;
  const long long need = footprint(*cfg);
  if (b->used_bytes + need > b->budget) return bfail(b, VACGPU_E_NOMEM, "HBM budget exhausted");
  int s = -1;
  for (int i = 0; i < b->cap; ++i)
    if (!b->inst[i].used && !b->slot_busy[i]) { s = i; break; }
  if (s < 0) return bfail(b, VACGPU_E_NOMEM, "no free slot");
  HostInst h;
  h.cfg = *cfg;
  h.m = *m;
  h.bytes = need;
  h.klass = size_class(*cfg);
  h.it0 = it;
  h.t0 = t;
  InstDev& I = h.dev;
  I.geo = make_geo(*cfg);
  const int nw = cfg->physics == VACGPU_MHD ? 6 : 4;
  I.mhd = cfg->physics == VACGPU_MHD;
  for (int f = 0; f < 6; ++f)
    for (int iw = 0; iw < 8; ++iw) {
      I.bc.t[f][iw] = (f < 4 && iw < nw) ? (unsigned char)(cfg->typeB[f][iw] == VACGPU_PERIODIC) : 0;
      I.bc.k[f][iw] = BK_COPY;
    }
  for (int i = 0; i < 8; ++i) I.sel.lim[i] = m->typelimiter[i];
  I.sel.divbfix = m->divbfix;
  I.sel.divbwave = m->divbwave;
  I.sel.sourceunsplit = m->sourceunsplit;
  I.sel.srcmul = 1;  // split sweeps: each carries its own share of the source
  I.sel.courantpar = m->courantpar;
  I.sel.gamma = cfg->eqpar[0];
  // fused sweeps: one boundary type per face for all variables, one limiter,
  // n >= 4, planes below 2^29 elements, MHD with the Powell source on
  {
    bool ok = cfg->nx[0] >= 4 && cfg->nx[1] >= 4 && (unsigned long long)I.geo.ps * 8ULL < (1ULL << 32);
    for (int f = 0; f < 4; ++f)
      for (int iw = 1; iw < nw; ++iw) ok &= I.bc.t[f][iw] == I.bc.t[f][0];
    for (int iw = 1; iw < nw; ++iw) ok &= m->typelimiter[iw] == m->typelimiter[0];
    if (I.mhd) ok &= m->divbfix && m->sourceunsplit;
    h.tvdmu = m->typefull == VACGPU_TVDMU;
    h.fused = ok && !h.tvdmu;
    h.lim = (ok || h.tvdmu) ? m->typelimiter[0] : -1;
  }
  long long slab[3];
  I.nbound = ghost_slabs(I.geo, slab);
  I.slab0 = slab[0];
  I.slab1 = slab[1];
  // first-dt partials live in w1 (unused until the first sweep): six slots
  // of one value per 64 x 4 tile
  const dim3 ftile((I.geo.n[0] + BX - 1) / BX, (I.geo.n[1] + BY - 1) / BY, 1);
  const long long nb = (long long)ftile.x * ftile.y;
  if (NSLOT * nb > I.geo.ps * nw) return bfail(b, VACGPU_E_ARG, "grid too small for the first-dt partials");
  const size_t plane_bytes = plane_bytes_of(I.geo, nw);
  const size_t dx_bytes = dx_bytes_of(I.geo);
  const int l0 = mesh ? I.geo.n[0] : (ld ? ld[0] : I.geo.ng[0]);
  const int l1 = mesh ? I.geo.n[1] : (ld ? ld[1] : I.geo.ng[1]);
  if (!mesh && (l0 < I.geo.ng[0] || l1 < I.geo.ng[1]))
    return bfail(b, VACGPU_E_ARG, "leading dimensions below ixG");
  const size_t wbytes = sizeof(double) * (size_t)nw * l0 * l1;
  const size_t dbytes = sizeof(double) * (size_t)(I.geo.ng[0] + I.geo.ng[1]);
  int si;
  if ((rc = stage_get(b, wbytes + dbytes + sizeof(BScal) + 512, &si))) return rc;
  I.w = pool_get(b, plane_bytes);
  I.b = pool_get(b, plane_bytes);
  I.w1 = pool_get(b, plane_bytes);
  double* dxd = pool_get(b, dx_bytes);
  if (!I.w || !I.b || !I.w1 || !dxd) {
    pool_put(b, plane_bytes, I.w);
    pool_put(b, plane_bytes, I.b);
    pool_put(b, plane_bytes, I.w1);
    pool_put(b, dx_bytes, dxd);
    return bfail(b, VACGPU_E_NOMEM, "device allocation failed");
  }
  I.geo.dx[0] = dxd;
  I.geo.dx[1] = dxd + I.geo.ng[0];
  I.geo.dx[2] = dxd;
  // host copies into the pinned stage; the caller may release its arrays on return
  Stage& St = b->stages[si];
  char* sp = reinterpret_cast<char*>(St.p);
  b->copier->copy(sp, wbytes, reinterpret_cast<const char*>(w), wbytes, wbytes, 1);
  double* sdx = reinterpret_cast<double*>(sp + (wbytes + 255) / 256 * 256);
  std::memcpy(sdx, dx0, sizeof(double) * I.geo.ng[0]);
  std::memcpy(sdx + I.geo.ng[0], dx1, sizeof(double) * I.geo.ng[1]);
  BScal* sbs = reinterpret_cast<BScal*>(reinterpret_cast<char*>(sdx) + (dbytes + 255) / 256 * 256);
  std::memset(sbs, 0, sizeof(BScal));
  sbs->s.dt = 0.0;
  sbs->s.t = t;
  sbs->s.smallp = -1.0;
  for (int d = 0; d < 3; ++d) sbs->s.dtcourant[d] = kBig;
  sbs->s.tmax = tmax;
  sbs->s.tmaxexact = tmaxexact;
  sbs->s.it = it;
  sbs->itmax = itmax;
  sbs->done = 0;
  sbs->needdt = 0;
  sbs->smallpcoeff = m->smallpcoeff;
  // uploads, the initial ghost cells (src/vac.t:54) and the first dt, all on
  // the copy stream
  hipStream_t cs = b->cstream;
  BHIP(b, hipMemcpyAsync(dxd, sdx, dbytes, hipMemcpyHostToDevice, cs));
  if (mesh) {
    BHIP(b, hipMemsetAsync(I.w, 0, plane_bytes, cs));
    for (int iw = 0; iw < nw; ++iw)
      BHIP(b, hipMemcpy2DAsync(I.w + iw * I.geo.ps + I.geo.off + 2 + 2 * I.geo.s[1],
                               sizeof(double) * I.geo.s[1], St.p + (size_t)iw * l0 * l1,
                               sizeof(double) * l0, sizeof(double) * I.geo.n[0], I.geo.n[1],
                               hipMemcpyHostToDevice, cs));
  } else {
    for (int iw = 0; iw < nw; ++iw)
      BHIP(b, hipMemcpy2DAsync(I.w + iw * I.geo.ps + I.geo.off, sizeof(double) * I.geo.s[1],
                               St.p + (size_t)iw * l0 * l1, sizeof(double) * l0,
                               sizeof(double) * I.geo.ng[0], I.geo.ng[1], hipMemcpyHostToDevice, cs));
  }
  BHIP(b, hipMemsetAsync(I.b, 0, plane_bytes, cs));
  BHIP(b, hipMemsetAsync(I.w1, 0, plane_bytes, cs));
  BHIP(b, hipMemcpyAsync(b->d_bs + s, sbs, sizeof(BScal), hipMemcpyHostToDevice, cs));
  BHIP(b, hipEventRecord(St.ev, cs));
  St.recorded = true;
  Scal* sc = &(b->d_bs + s)->s;
  double* fp = I.w1;
  const dim3 tb(BX, BY);
  constexpr int ND = 2;
  BDISPATCH(I.mhd, false, {
    (void)sizeof(R);
    constexpr int NWV = Layout<ND, MHD>::nw;
    hipLaunchKernelGGL((k_bound<NWV>), dim3((unsigned)((I.nbound + 255) / 256)), dim3(256), 0, cs, I.geo, I.bc,
                       I.w, I.nbound, I.slab0, I.slab1, 0, NWV, (const double*)nullptr);
    hipLaunchKernelGGL((k_pminmax<ND, MHD>), ftile, tb, 0, cs, I.geo, I.w, I.sel.gamma, sc, fp, nb);
    hipLaunchKernelGGL(k_fin_smallp, dim3(1), dim3(256), 0, cs, sc, fp, nb, nb, m->smallpcoeff, (int)MHD);
    for (int d = 0; d < ND; ++d)
      hipLaunchKernelGGL((k_cmaxdx<ND, MHD>), ftile, tb, 0, cs, I.geo, I.w, I.sel.gamma, sc, fp, nb, d);
  });
  hipLaunchKernelGGL(kb_first_getdt, dim3(1), dim3(256), 0, cs, I, b->d_bs + s, fp, nb, (int)nb);
  if ((rc = launch_err(b, "first dt"))) return rc;
  BHIP(b, hipEventRecord(b->ev_admit, cs));
  b->admit_pending = true;
  h.used = true;
  b->inst[s] = h;
  b->used_bytes += need;
  b->dirty = true;
  *slot = s;
  return 0;
}

}  // namespace

extern "C" {

int vacgpu_batch_add(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
                     const double* dx0, const double* dx1, const double* w, const int* ld,
                     int it, double t, int itmax, double tmax, int tmaxexact, int* slot) {
  return batch_add(b, cfg, m, dx0, dx1, w, ld, 0, it, t, itmax, tmax, tmaxexact, slot);
}

int vacgpu_batch_add_mesh(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
                          const double* dx0, const double* dx1, const double* wmesh, int it,
                          double t, int itmax, double tmax, int tmaxexact, int* slot) {
  return batch_add(b, cfg, m, dx0, dx1, wmesh, nullptr, 1, it, t, itmax, tmax, tmaxexact, slot);
}

int vacgpu_batch_remove(vacgpu_batch* b, int slot) {
  if (!b || slot < 0 || slot >= b->cap || !b->inst[slot].used) return VACGPU_E_ARG;
  HostInst& h = b->inst[slot];
  const int nw = h.dev.mhd ? 6 : 4;
  const size_t plane_bytes = plane_bytes_of(h.dev.geo, nw);
  const size_t dx_bytes = dx_bytes_of(h.dev.geo);
  Retired r;
  r.slot = slot;
  r.bytes = h.bytes;
  r.blocks = {{plane_bytes, h.dev.w}, {plane_bytes, h.dev.b}, {plane_bytes, h.dev.w1},
              {dx_bytes, const_cast<double*>(h.dev.geo.dx[0])}};
  b->slot_busy[slot] = 1;
  h = HostInst();
  b->dirty = true;
  return retire(b, std::move(r));
}

int vacgpu_batch_step(vacgpu_batch* b, int nsteps) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  for (int n = 0; n < nsteps; ++n) {
    const int parity = (int)(b->step_count % 2);
    if (b->use_graph) {
      if (!b->graph[parity]) {
        hipGraph_t gr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(b, parity);
        hipError_t e = hipStreamEndCapture(b->stream, &gr);
        if (rc) return rc;
        if (e != hipSuccess) return bfail(b, VACGPU_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        BHIP(b, hipGraphInstantiate(&b->graph[parity], gr, nullptr, nullptr, 0));
        BHIP(b, hipGraphDestroy(gr));
      }
      BHIP(b, hipGraphLaunch(b->graph[parity], b->stream));
    } else {
      if ((rc = enqueue_step(b, parity))) return rc;
    }
    b->step_count += 1;
  }
  return 0;
}

int vacgpu_batch_snapshot(vacgpu_batch* b) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if ((int)b->snap_queue.size() == NSNAP - 1 && (rc = snap_take(b))) return rc;
  int m = 0;
  for (; m < NSNAP; ++m) {
    bool busy = m == b->snap_cur;
    for (int q : b->snap_queue) busy |= q == m;
    if (!busy) break;
  }
  BHIP(b, hipMemcpyAsync(b->h_snap[m], b->d_bs, sizeof(BScal) * b->cap, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipEventRecord(b->ev_snap[m], b->stream));
  b->snap_seqno[m] = b->snap_next++;
  b->snap_queue.push_back(m);
  return 0;
}

int vacgpu_batch_snapshot_wait(vacgpu_batch* b, int wait) {
  if (!b) return VACGPU_E_ARG;
  if (b->snap_queue.empty()) return 0;
  if (!wait && !ev_done(b->ev_snap[b->snap_queue.front()])) return 0;
  int rc;
  if ((rc = snap_take(b))) return rc;
  return 1;
}

int vacgpu_batch_sync(vacgpu_batch* b) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  if ((rc = vacgpu_batch_snapshot(b))) return rc;
  while (!b->snap_queue.empty())
    if ((rc = snap_take(b))) return rc;
  return sync_all(b);
}

int vacgpu_batch_status(vacgpu_batch* b, int slot, int* it, double* t, double* dt, int* done,
                        int* nerror) {
  if (!b || slot < 0 || slot >= b->cap || !b->inst[slot].used) return VACGPU_E_ARG;
  const HostInst& h = b->inst[slot];
  if (b->snap_cur < 0 || !h.joined || b->snap_seqno[b->snap_cur] < h.join_seq) {
    // no snapshot taken since the instance joined: its admission values
    if (it) *it = h.it0;
    if (t) *t = h.t0;
    if (dt) *dt = 0.0;
    if (done) *done = 0;
    if (nerror)
      for (int i = 0; i < 4; ++i) nerror[i] = 0;
    return 0;
  }
  const BScal& S = b->h_snap[b->snap_cur][slot];
  if (it) *it = S.s.it;
  if (t) *t = S.s.t;
  if (dt) *dt = S.s.dt;
  if (done) *done = S.done;
  if (nerror)
    for (int i = 0; i < 4; ++i) nerror[i] = S.s.nerror[i];
  return 0;
}

int vacgpu_batch_finished(vacgpu_batch* b, int* slots, int cap) {
  if (!b || (!slots && cap > 0)) return VACGPU_E_ARG;
  if (b->snap_cur < 0) return 0;
  const BScal* S = b->h_snap[b->snap_cur];
  const long long seq = b->snap_seqno[b->snap_cur];
  int n = 0;
  for (int i = 0; i < b->cap; ++i) {
    const HostInst& h = b->inst[i];
    if (!h.used || !h.joined || seq < h.join_seq || !S[i].done) continue;
    if (n < cap) slots[n] = i;
    ++n;
  }
  return n;
}

int vacgpu_batch_fetch_start(vacgpu_batch* b, int slot, int mesh, int* ticket) {
  if (!b || !ticket || slot < 0 || slot >= b->cap || !b->inst[slot].used) return VACGPU_E_ARG;
  int rc;
  // the copy runs on the copy stream after the instance's own upload and
  // after the steps enqueued so far (an instance that has not joined yet
  // has taken no step: its admitted state is copied)
  const InstDev& I = b->inst[slot].dev;
  const int nw = I.mhd ? 6 : 4;
  const int e0 = mesh ? I.geo.n[0] : I.geo.ng[0], e1 = mesh ? I.geo.n[1] : I.geo.ng[1];
  const long long o0 = mesh ? 2 + 2 * I.geo.s[1] : 0;
  const size_t bytes = sizeof(double) * (size_t)nw * e0 * e1;
  int si;
  if ((rc = stage_get(b, bytes, &si))) return rc;
  Stage& St = b->stages[si];
  hipEvent_t ec;
  if ((rc = ev_get(b, &ec))) return rc;
  BHIP(b, hipEventRecord(ec, b->stream));
  BHIP(b, hipStreamWaitEvent(b->cstream, ec, 0));
  b->evpool.push_back(ec);
  for (int iw = 0; iw < nw; ++iw)
    BHIP(b, hipMemcpy2DAsync(St.p + (size_t)iw * e0 * e1, sizeof(double) * e0,
                             I.w + iw * I.geo.ps + I.geo.off + o0, sizeof(double) * I.geo.s[1],
                             sizeof(double) * e0, e1, hipMemcpyDeviceToHost, b->cstream));
  BHIP(b, hipEventRecord(St.ev, b->cstream));
  St.recorded = true;
  St.held = true;
  int k = 0;
  while (k < (int)b->fetches.size() && b->fetches[k].used) ++k;
  if (k == (int)b->fetches.size()) b->fetches.emplace_back();
  b->fetches[k] = Fetch{true, si, nw, e0, e1};
  *ticket = k;
  return 0;
}

int vacgpu_batch_fetch_ready(vacgpu_batch* b, int ticket, double* w, const int* ld, int wait) {
  if (!b || ticket < 0 || ticket >= (int)b->fetches.size() || !b->fetches[ticket].used) return VACGPU_E_ARG;
  Fetch& F = b->fetches[ticket];
  Stage& St = b->stages[F.stage];
  if (!wait && !ev_done(St.ev)) return 0;
  BHIP(b, hipEventSynchronize(St.ev));
  const int l0 = ld ? ld[0] : F.ng0, l1 = ld ? ld[1] : F.ng1;
  if (w) {
    if (l0 < F.ng0 || l1 < F.ng1) return bfail(b, VACGPU_E_ARG, "leading dimensions below the fetched extent");
    // rows beyond the fetched extent in the caller's layout are left as the caller had them
    if (l0 == F.ng0 && l1 == F.ng1) {
      const size_t bytes = sizeof(double) * (size_t)F.nw * F.ng0 * F.ng1;
      b->copier->copy(reinterpret_cast<char*>(w), bytes, reinterpret_cast<const char*>(St.p), bytes,
                      bytes, 1);
    } else {
      for (int iw = 0; iw < F.nw; ++iw)
        b->copier->copy(reinterpret_cast<char*>(w + (size_t)iw * l0 * l1), sizeof(double) * l0,
                        reinterpret_cast<const char*>(St.p + (size_t)iw * F.ng0 * F.ng1),
                        sizeof(double) * F.ng0, sizeof(double) * F.ng0, F.ng1);
    }
  }
  St.held = false;
  F = Fetch{};
  return 1;
}

int vacgpu_batch_get_state(vacgpu_batch* b, int slot, double* w, const int* ld) {
  if (!b || !w) return VACGPU_E_ARG;
  int ticket, rc;
  if ((rc = vacgpu_batch_fetch_start(b, slot, 0, &ticket))) return rc;
  rc = vacgpu_batch_fetch_ready(b, ticket, w, ld, 1);
  return rc == 1 ? 0 : rc;
}

int vacgpu_batch_graph(vacgpu_batch* b, int enable) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if ((rc = sync_all(b))) return rc;
  b->use_graph = enable != 0;
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) { (void)hipGraphExecDestroy(b->graph[p]); b->graph[p] = nullptr; }
  return 0;
}

// Autotuner: for every size class present, time the predictor + corrector
// launches of both directions for each candidate tile shape (64 x rows) and
// keep the fastest.  The timed launches write only the scratch arrays b and
// w1 and the partials (the fused sweeps without their fold), so the state
// is unchanged.
int vacgpu_batch_tune(vacgpu_batch* b, int* rows_out) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  if ((rc = sync_all(b))) return rc;
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  Work* tmp = nullptr;
  int tmp_cap = 0;
  double* tpart = nullptr;
  BHIP(b, hipMalloc(&tpart, sizeof(double) * 64));
  for (int c = 0; c < NCLASS; ++c) {
    // tiles of class-c instances per physics/tier key, for each candidate shape
    bool present = false;
    for (const auto& h : b->inst) present |= (h.used && h.klass == c && !h.tvdmu && !(h.fused && b->use_fused));
    if (!present) continue;
    float best = 1e30f;
    int best_rows = b->rows[c];
    for (int r : kTileRows) {
      std::map<Key, std::vector<Work>> lists;
      for (int i = 0; i < (int)b->inst.size(); ++i) {
        const HostInst& h = b->inst[i];
        if (!h.used || h.klass != c || h.tvdmu) continue;
        Key k{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK, -1, KIND_TILE};
        const Geo& g = h.dev.geo;
        const int tx = (g.n[0] + BX - 1) / BX, ty = (g.n[1] + r - 1) / r;
        for (int by = 0; by < ty; ++by)
          for (int bx = 0; bx < tx; ++bx) lists[k].push_back(Work{i, bx, by, 0, 0});
      }
      size_t total = 0;
      for (auto& kv : lists) total += kv.second.size();
      if ((int)total > tmp_cap) {
        if (tmp) BHIP(b, hipFree(tmp));
        tmp_cap = (int)total;
        BHIP(b, hipMalloc(&tmp, sizeof(Work) * tmp_cap));
      }
      size_t off = 0;
      std::vector<std::pair<Key, std::pair<size_t, int>>> segs;
      for (auto& kv : lists) {
        BHIP(b, hipMemcpy(tmp + off, kv.second.data(), sizeof(Work) * kv.second.size(), hipMemcpyHostToDevice));
        segs.push_back({kv.first, {off, (int)kv.second.size()}});
        off += kv.second.size();
      }
      float tot = 0.f;
      for (int rep = 0; rep < 3; ++rep) {
        BHIP(b, hipEventRecord(e0, b->stream));
        for (auto& sg : segs) {
          const Work* W = tmp + sg.second.first;
          const int n = sg.second.second;
          const dim3 blk(BX, r);
          constexpr int ND = 2;
          BDISPATCH(sg.first.mhd, sg.first.f32, {
            hipLaunchKernelGGL((kb_pred<ND, MHD, R, 0>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0);
            hipLaunchKernelGGL((kb_corr<ND, MHD, R, 0>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0, tpart, (long long)0);
            hipLaunchKernelGGL((kb_pred<ND, MHD, R, 1>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0);
            hipLaunchKernelGGL((kb_corr<ND, MHD, R, 1>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0, tpart, (long long)0);
          });
        }
        BHIP(b, hipEventRecord(e1, b->stream));
        BHIP(b, hipEventSynchronize(e1));
        float ms = 0.f;
        BHIP(b, hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) tot += ms;  // the first repetition warms up
      }
      if ((rc = launch_err(b, "tune"))) return rc;
      if (tot < best) { best = tot; best_rows = r; }
    }
    b->rows[c] = best_rows;
  }
  // fused instances: segment length per size class, timing both sweeps
  for (int c = 0; c < NCLASS; ++c) {
    bool present = false;
    for (const auto& h : b->inst) present |= (h.used && h.klass == c && h.fused && b->use_fused);
    if (!present) continue;
    float best = 1e30f;
    int best_sl = b->seglen[c];
    for (int sl : kSegLens) {
      std::map<Key, std::vector<Work>> lists;
      for (int i = 0; i < (int)b->inst.size(); ++i) {
        const HostInst& h = b->inst[i];
        if (!h.used || h.klass != c || !h.fused) continue;
        b->hdev[i].seglen = sl;
        const Geo& g = h.dev.geo;
        const int strips = (g.n[0] + FS_STRIDE - 1) / FS_STRIDE;
        const int sblocks = ((g.n[1] + sl - 1) / sl + 3) / 4;
        Key k{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK, h.lim, KIND_FUSED};
        for (int sb = 0; sb < sblocks; ++sb)
          for (int st = 0; st < strips; ++st) lists[k].push_back(Work{i, st, sb, 0, 0});
      }
      BHIP(b, hipMemcpy(b->d_inst, b->hdev.data(), sizeof(InstDev) * b->inst.size(), hipMemcpyHostToDevice));
      size_t total = 0;
      for (auto& kv : lists) total += kv.second.size();
      if ((int)total > tmp_cap) {
        if (tmp) BHIP(b, hipFree(tmp));
        tmp_cap = (int)total;
        BHIP(b, hipMalloc(&tmp, sizeof(Work) * tmp_cap));
      }
      size_t off = 0;
      std::vector<std::pair<Key, std::pair<size_t, int>>> segs;
      for (auto& kv : lists) {
        BHIP(b, hipMemcpy(tmp + off, kv.second.data(), sizeof(Work) * kv.second.size(), hipMemcpyHostToDevice));
        segs.push_back({kv.first, {off, (int)kv.second.size()}});
        off += kv.second.size();
      }
      float tot = 0.f;
      for (int rep = 0; rep < 3; ++rep) {
        BHIP(b, hipEventRecord(e0, b->stream));
        for (auto& sg : segs) {
          const Work* W = tmp + sg.second.first;
          const int n = sg.second.second;
          const int lim = sg.first.lim;
          const dim3 fb(64, 4);
#define KBT(DV, LV) hipLaunchKernelGGL((kb_fused<MHD, R, DV, LV>), dim3(n), fb, 0, b->stream, b->d_inst, b->d_bs, W, 0, b->d_fold, 0)
#define KBTL(DV) do { if (lim == LIM_MINMOD) KBT(DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) KBT(DV, LIM_WOODWARD); else KBT(DV, LIM_SUPERBEE); } while (0)
          BDISPATCH(sg.first.mhd, sg.first.f32, { KBTL(0); KBTL(1); });
#undef KBTL
#undef KBT
        }
        BHIP(b, hipEventRecord(e1, b->stream));
        BHIP(b, hipEventSynchronize(e1));
        float ms = 0.f;
        BHIP(b, hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) tot += ms;
      }
      if ((rc = launch_err(b, "tune"))) return rc;
      if (tot < best) { best = tot; best_sl = sl; }
    }
    b->seglen[c] = best_sl;
  }
  if (tmp) BHIP(b, hipFree(tmp));
  BHIP(b, hipFree(tpart));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->dirty = true;
  if ((rc = rebuild(b))) return rc;
  if (rows_out)
    for (int c = 0; c < NCLASS; ++c) rows_out[c] = b->rows[c];
  return 0;
}

int vacgpu_batch_set_tiles(vacgpu_batch* b, const int* rows) {
  if (!b || !rows) return VACGPU_E_ARG;
  for (int c = 0; c < NCLASS; ++c) {
    bool ok = false;
    for (int r : kTileRows) ok |= (rows[c] == r);
    if (!ok) return bfail(b, VACGPU_E_ARG, "tile rows must be 1, 2, 4 or 8");
    b->rows[c] = rows[c];
  }
  b->dirty = true;
  return 0;
}

int vacgpu_batch_set_kernels(vacgpu_batch* b, int fused, const int* seglen) {
  if (!b) return VACGPU_E_ARG;
  if (seglen)
    for (int c = 0; c < NCLASS; ++c)
      if (seglen[c] < 4 || seglen[c] > 256) return bfail(b, VACGPU_E_ARG, "segment length must be in [4, 256]");
  b->use_fused = fused != 0;
  if (seglen)
    for (int c = 0; c < NCLASS; ++c) b->seglen[c] = seglen[c];
  b->dirty = true;
  return 0;
}

int vacgpu_batch_get_tiles(const vacgpu_batch* b, int* rows, int* seglen) {
  if (!b) return VACGPU_E_ARG;
  for (int c = 0; c < NCLASS; ++c) {
    if (rows) rows[c] = b->rows[c];
    if (seglen) seglen[c] = b->seglen[c];
  }
  return 0;
}

long long vacgpu_batch_bytes_used(const vacgpu_batch* b) { return b ? b->used_bytes : -1; }

const char* vacgpu_batch_last_error(const vacgpu_batch* b) { return b ? b->err.c_str() : "null batch"; }

void vacgpu_batch_destroy(vacgpu_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->cstream) (void)hipStreamSynchronize(b->cstream);
  (void)reap(b, true);
  for (auto& h : b->inst)
    if (h.used) {
      (void)hipFree(h.dev.w);
      (void)hipFree(h.dev.b);
      (void)hipFree(h.dev.w1);
      (void)hipFree(const_cast<double*>(h.dev.geo.dx[0]));
    }
  for (auto& kv : b->lists)
    for (int c = 0; c < 2; ++c)
      if (kv.second.ph[c].d) (void)hipFree(kv.second.ph[c].d);
  for (auto& kv : b->pool) (void)hipFree(kv.second);
  if (b->gw4.d) (void)hipFree(b->gw4.d);
  if (b->gw6.d) (void)hipFree(b->gw6.d);
  if (b->d_ids) (void)hipFree(b->d_ids);
  if (b->d_idsu) (void)hipFree(b->d_idsu);
  if (b->d_idst) (void)hipFree(b->d_idst);
  if (b->part) (void)hipFree(b->part);
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) (void)hipGraphExecDestroy(b->graph[p]);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_bs) (void)hipFree(b->d_bs);
  if (b->d_fold) (void)hipFree(b->d_fold);
  for (int m = 0; m < NSNAP; ++m) {
    if (b->h_snap[m]) (void)hipHostFree(b->h_snap[m]);
    if (b->ev_snap[m]) (void)hipEventDestroy(b->ev_snap[m]);
  }
  for (auto& S : b->stages) {
    if (S.p) (void)hipHostFree(S.p);
    if (S.ev) (void)hipEventDestroy(S.ev);
  }
  for (int a = 0; a < 2; ++a) {
    if (b->arena[a].h) (void)hipHostFree(b->arena[a].h);
    if (b->arena[a].ev) (void)hipEventDestroy(b->arena[a].ev);
  }
  for (hipEvent_t e : b->evpool) (void)hipEventDestroy(e);
  if (b->ev_admit) (void)hipEventDestroy(b->ev_admit);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->cstream) (void)hipStreamDestroy(b->cstream);
  delete b->copier;
  delete b;
}

}  // extern "C"
