// batch.hip -- many independent grids advanced together on one GPU.
//
// Every instance keeps its own state in HBM (state w, sweep output b,
// predictor w1), its own device scalars (dt, t, it, dtcourant, smallp,
// nerror, stop condition) and its own selectors.  One batched step runs the
// TVDLF step (Courant dt, then one x and one y sweep, each predictor +
// corrector + ghost cells) of every active instance with ~20 launches in
// total: each launch walks a work list of (instance, tile) entries, so grids
// of any size share a launch.  Instances are grouped by kernel instance
// (physics, arithmetic tier) and by tile shape, which the autotuner picks
// per size class.  The sweep order of an instance follows its own step
// parity (src/vac.t:310-324): its tiles sit in the x-first or the y-first
// list.  Stop conditions (itmax, tmax, tmaxexact) are evaluated on the
// device, so a finished instance drops out without the host.  With fixed
// membership a step is captured once into a hipGraph and replayed.
//
// The per-cell arithmetic is the single-grid one (sweep.hpp), so an instance
// advanced in a batch is bitwise identical to the same grid advanced alone.
#include "../../include/vacgpu.h"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
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
  int pbase, ntiles;    // partials of the reduction tiles (pressure floor, cell-centre dt)
  int pbase2, ntiles2;  // partials of the sweep launches (Courant maximum)
  int mhd;
  int seglen;           // fused sweeps: rows per marching segment
};

struct BScal {
  Scal s;
  int itmax;
  int done;
  double smallpcoeff;
};

// block-uniform early exit for finished instances
#define BATCH_PROLOGUE                        \
  const Work wk = work[blockIdx.x];           \
  const InstDev& I = inst[wk.inst];           \
  const BScal& S = bs[wk.inst];               \
  if (S.done) return;

template <int ND, bool MHD>
__global__ __launch_bounds__(512) void kb_pminmax(const InstDev* __restrict__ inst,
                                                 const BScal* __restrict__ bs,
                                                 const Work* __restrict__ work,
                                                 double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  if (S.s.smallp >= 0.0) return;
  double mn, mx;
  pminmax_body<ND, MHD>(I.geo, I.w, I.sel.gamma, wk.bx, wk.by, wk.bz, mn, mx);
  const double bmn = block_reduce<false>(mn, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMIN * pstride + wk.pidx] = bmn;
  const double bmx = block_reduce<true>(mx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMAX * pstride + wk.pidx] = bmx;
}

template <int ND, bool MHD>
__global__ __launch_bounds__(512) void kb_cmaxdx(const InstDev* __restrict__ inst,
                                                const BScal* __restrict__ bs,
                                                const Work* __restrict__ work,
                                                double* __restrict__ part, long long pstride, int d) {
  __shared__ double sm[8];
  BATCH_PROLOGUE
  if (S.s.dtcourant[d] < kBig) return;
  const double v = cmaxdx_body<ND, MHD>(I.geo, I.w, I.sel.gamma, S.s.smallp, d, wk.bx, wk.by, wk.bz);
  const double b = block_reduce<true>(v, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[(SLOT_DX + d) * pstride + wk.pidx] = b;
}

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
// (k = 1); ghost cells of the output are written by the sweep itself
template <bool MHD, class R, int D, int LIM>
__global__ __launch_bounds__(256) void kb_fused(const InstDev* __restrict__ inst,
                                               const BScal* __restrict__ bs,
                                               const Work* __restrict__ work, int k,
                                               double* __restrict__ part, long long pstride) {
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
  const double b = block_reduce<true>(cmx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + wk.pidx] = b;
}

// ghost cells of one array of every listed instance: which = 0 w, 1 b, 2 w1
template <int NW>
__global__ __launch_bounds__(256) void kb_bound(const InstDev* __restrict__ inst,
                                               const BScal* __restrict__ bs,
                                               const GWork* __restrict__ work, int which) {
  const GWork gw = work[blockIdx.x];
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

// one block per listed instance: pressure floor, then getdt_courant
// (src/vac.t:664-722) and the stop condition of the main loop (src/vac.t:73)
__global__ void kb_fin_smallp(const InstDev* __restrict__ inst, BScal* __restrict__ bs,
                              const int* __restrict__ ids, const double* __restrict__ part,
                              long long pstride) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done || S.s.smallp >= 0.0) return;
  const InstDev& I = inst[i];
  const double mn = fold_range(part + SLOT_PMIN * pstride + I.pbase, I.ntiles, false, sm);
  __syncthreads();
  const double mx = fold_range(part + SLOT_PMAX * pstride + I.pbase, I.ntiles, true, sm);
  if (threadIdx.x == 0) {
    double sp = mn * S.smallpcoeff;
    if (I.mhd && sp < 0.0) sp = mx * S.smallpcoeff;
    S.s.smallp = sp;
  }
}

__global__ void kb_fin_getdt(const InstDev* __restrict__ inst, BScal* __restrict__ bs,
                             const int* __restrict__ ids, const double* __restrict__ part,
                             long long pstride) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done) return;
  const InstDev& I = inst[i];
  const int nd = I.geo.nd;
  double dtc[3] = {kBig, kBig, kBig};
  for (int d = 0; d < nd; ++d) dtc[d] = S.s.dtcourant[d];
  double cmd[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < nd; ++d)
    if (dtc[d] >= kBig) {
      cmd[d] = fold_range(part + (SLOT_DX + d) * pstride + I.pbase, I.ntiles, true, sm);
      __syncthreads();
    }
  __syncthreads();
  if (threadIdx.x == 0) {
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
}

// dtcourant update after the corrector of sweep k (src/vactvdlf.t:184-201)
__global__ void kb_fin_courant(const InstDev* __restrict__ inst, BScal* __restrict__ bs,
                               const int* __restrict__ ids, const double* __restrict__ part,
                               long long pstride, int k) {
  __shared__ double sm[8];
  const int i = ids[blockIdx.x];
  BScal& S = bs[i];
  if (S.done) return;
  const InstDev& I = inst[i];
  const double cm = fold_range(part + SLOT_C0 * pstride + I.pbase2, I.ntiles2, true, sm);
  if (threadIdx.x == 0 && I.sel.courantpar > 0.0) {
    const int nd = I.geo.nd;
    const bool fwd = (S.s.it / 2) * 2 == S.s.it;
    const int d = fwd ? k : nd - 1 - k;
    if (S.s.dt * cm > 1.0) S.s.nerror[2] += 1;
    if (cm > kSmall) {
      double c = I.sel.courantpar / cm;
      S.s.dtcourant[d] = c < S.s.dtcourant[d] ? c : S.s.dtcourant[d];
    }
  }
}

__global__ void kb_step_end(BScal* __restrict__ bs, const int* __restrict__ ids, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  BScal& S = bs[ids[j]];
  if (S.done) return;
  S.s.it += 1;
  S.s.t = S.s.t + S.s.dt;
}

// ---------------------------------------------------------------- host
enum { NCLASS = 4 };
const int kTileRows[] = {1, 2, 4, 8};
const int kSegLens[] = {8, 16, 32, 64};

// kernel instance of a work list: physics, tier, and for the fused sweeps
// the limiter type (lim < 0: predictor / corrector tiles)
struct Key {
  int mhd, f32, lim;
  bool operator<(const Key& o) const {
    if (mhd != o.mhd) return mhd < o.mhd;
    if (f32 != o.f32) return f32 < o.f32;
    return lim < o.lim;
  }
};

struct HostInst {
  bool used = false;
  vacgpu_config cfg{};
  vacgpu_method m{};
  InstDev dev{};
  long long bytes = 0;
  int klass = 0;
  int phase_class = 0;  // (it at admission - batch step at admission) mod 2
  int ntiles = 0;
  dim3 tiles;           // tile grid for the class tile shape
  bool fused = false;   // sweeps by kb_fused (else predictor + corrector tiles)
  int lim = -1;         // the common limiter type of a fused instance
};

struct Launch {          // one work list on the device
  Work* d = nullptr;
  int n = 0;
  int cap = 0;
};

}  // namespace

struct vacgpu_batch {
  long long budget = 0, used_bytes = 0;
  std::vector<HostInst> inst;
  InstDev* d_inst = nullptr;
  InstDev* h_inst = nullptr;  // pinned mirror, uploaded whole by rebuild
  double* stage = nullptr;    // pinned staging for state uploads / downloads
  size_t stage_bytes = 0;
  BScal* d_bs = nullptr;
  BScal* h_bs = nullptr;  // pinned mirror
  int cap = 0;
  double* part = nullptr;
  long long pstride = 0;
  int rows[NCLASS] = {4, 4, 4, 4};       // tile rows per size class (autotuned)
  int seglen[NCLASS] = {16, 16, 16, 16};  // fused sweeps: rows per segment per size class
  bool use_fused = true;
  long long step_count = 0;
  bool dirty = true;
  // work lists per (physics/tier key, tile rows): every tile, and the tiles
  // of each phase class c (0/1)
  struct Lists {
    Launch all;          // every tile of the group/class
    Launch ph[2];        // tiles of instances with phase class c
  };
  std::map<std::pair<Key, int>, Lists> lists;
  std::vector<GWork> hg[3];
  struct { GWork* d = nullptr; int n = 0, cap = 0; } gw4, gw6;  // ghost work per nw
  int* d_ids = nullptr;
  int nids = 0, ids_cap = 0;
  std::vector<int> fresh;                     // slots whose ghost cells need a first fill
  std::multimap<size_t, double*> pool;        // reusable device blocks by size
  hipStream_t stream = nullptr;
  bool use_graph = false;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // per step parity
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

int ensure_stage(vacgpu_batch* b, size_t bytes) {
  if (bytes <= b->stage_bytes) return 0;
  if (b->stage) BHIP(b, hipHostFree(b->stage));
  b->stage = nullptr;
  b->stage_bytes = 0;
  BHIP(b, hipHostMalloc(&b->stage, bytes, hipHostMallocDefault));
  b->stage_bytes = bytes;
  return 0;
}

int ensure_launch(vacgpu_batch* b, Launch& L, const std::vector<Work>& v) {
  if ((int)v.size() > L.cap) {
    if (L.d) BHIP(b, hipFree(L.d));
    L.cap = (int)v.size() * 2 + 16;
    BHIP(b, hipMalloc(&L.d, sizeof(Work) * L.cap));
  }
  L.n = (int)v.size();
  if (L.n) BHIP(b, hipMemcpyAsync(L.d, v.data(), sizeof(Work) * L.n, hipMemcpyHostToDevice, b->stream));
  return 0;
}

// Rebuild work lists, partial ranges and the active-id list after a change
// of membership or tile shapes.
int rebuild(vacgpu_batch* b) {
  std::map<std::pair<Key, int>, std::vector<Work>> all, ph0, ph1;
  std::vector<GWork> g4, g6;
  std::vector<int> ids;
  int pbase = 0, pbase2 = 0;
  for (int i = 0; i < (int)b->inst.size(); ++i) {
    HostInst& h = b->inst[i];
    if (!h.used) continue;
    const int rows = b->rows[h.klass];
    const Geo& g = h.dev.geo;
    // reduction tiles (64 x rows), also the sweep tiles of unfused instances
    h.tiles = dim3((g.n[0] + BX - 1) / BX, (g.n[1] + rows - 1) / rows, 1);
    h.ntiles = (int)(h.tiles.x * h.tiles.y);
    h.dev.pbase = pbase;
    h.dev.ntiles = h.ntiles;
    const bool fused = h.fused && b->use_fused;
    const Key kr{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK, -1};
    auto key = std::make_pair(kr, rows);
    auto& phl = h.phase_class ? ph1 : ph0;
    int t = 0;
    for (unsigned by = 0; by < h.tiles.y; ++by)
      for (unsigned bx = 0; bx < h.tiles.x; ++bx, ++t) {
        all[key].push_back(Work{i, (int)bx, (int)by, 0, pbase + t});
        if (!fused) phl[key].push_back(Work{i, (int)bx, (int)by, 0, pbase2 + t});
      }
    pbase += h.ntiles;
    if (fused) {
      // sweep blocks: strips of 62 x cells, blocks of four segments
      const int sl = b->seglen[h.klass];
      const int strips = (g.n[0] + FS_STRIDE - 1) / FS_STRIDE;
      const int segs = (g.n[1] + sl - 1) / sl;
      const int sblocks = (segs + 3) / 4;
      h.dev.seglen = sl;
      h.dev.pbase2 = pbase2;
      h.dev.ntiles2 = strips * sblocks;
      const Key kf{kr.mhd, kr.f32, h.lim};
      auto keyf = std::make_pair(kf, sl);
      all[keyf];  // the list exists (its reduction part stays empty)
      int u = 0;
      for (int sb = 0; sb < sblocks; ++sb)
        for (int st = 0; st < strips; ++st, ++u) phl[keyf].push_back(Work{i, st, sb, 0, pbase2 + u});
      pbase2 += h.dev.ntiles2;
    } else {
      h.dev.seglen = 0;
      h.dev.pbase2 = pbase2;
      h.dev.ntiles2 = h.ntiles;
      pbase2 += h.ntiles;
      // ghost cells are filled by separate launches only for unfused instances
      for (long long q = 0; q < h.dev.nbound; q += 256) (kr.mhd ? g6 : g4).push_back(GWork{i, (int)q});
    }
    ids.push_back(i);
    b->h_inst[i] = h.dev;
  }
  if (!ids.empty())
    BHIP(b, hipMemcpyAsync(b->d_inst, b->h_inst, sizeof(InstDev) * (ids.back() + 1), hipMemcpyHostToDevice, b->stream));
  // partial arrays
  const int pneed = pbase > pbase2 ? pbase : pbase2;
  if (pneed > b->pstride) {
    if (b->part) BHIP(b, hipFree(b->part));
    b->pstride = (long long)pneed * 2 + 64;
    BHIP(b, hipMalloc(&b->part, sizeof(double) * NSLOT * b->pstride));
  }
  for (auto& kv : b->lists) { kv.second.all.n = 0; kv.second.ph[0].n = 0; kv.second.ph[1].n = 0; }
  int rc;
  for (auto& kv : all) {
    auto& L = b->lists[kv.first];
    if ((rc = ensure_launch(b, L.all, kv.second))) return rc;
    if ((rc = ensure_launch(b, L.ph[0], ph0[kv.first]))) return rc;
    if ((rc = ensure_launch(b, L.ph[1], ph1[kv.first]))) return rc;
  }
  auto upload_g = [&](auto& G, const std::vector<GWork>& v) -> int {
    if ((int)v.size() > G.cap) {
      if (G.d) BHIP(b, hipFree(G.d));
      G.cap = (int)v.size() * 2 + 16;
      BHIP(b, hipMalloc(&G.d, sizeof(GWork) * G.cap));
    }
    G.n = (int)v.size();
    if (G.n) BHIP(b, hipMemcpyAsync(G.d, v.data(), sizeof(GWork) * G.n, hipMemcpyHostToDevice, b->stream));
    return 0;
  };
  if ((rc = upload_g(b->gw4, g4))) return rc;
  if ((rc = upload_g(b->gw6, g6))) return rc;
  if ((int)ids.size() > b->ids_cap) {
    if (b->d_ids) BHIP(b, hipFree(b->d_ids));
    b->ids_cap = (int)ids.size() * 2 + 16;
    BHIP(b, hipMalloc(&b->d_ids, sizeof(int) * b->ids_cap));
  }
  b->nids = (int)ids.size();
  if (b->nids) BHIP(b, hipMemcpyAsync(b->d_ids, ids.data(), sizeof(int) * b->nids, hipMemcpyHostToDevice, b->stream));
  // first ghost-cell fill of newly admitted states (src/vac.t:54)
  if (!b->fresh.empty()) {
    std::vector<GWork> f4, f6;
    for (int sl : b->fresh) {
      const HostInst& h = b->inst[sl];
      if (!h.used) continue;
      for (long long q = 0; q < h.dev.nbound; q += 256) (h.dev.mhd ? f6 : f4).push_back(GWork{sl, (int)q});
    }
    GWork* tmp = nullptr;
    const size_t n = f4.size() + f6.size();
    if (n) {
      BHIP(b, hipMalloc(&tmp, sizeof(GWork) * n));
      if (!f4.empty()) BHIP(b, hipMemcpyAsync(tmp, f4.data(), sizeof(GWork) * f4.size(), hipMemcpyHostToDevice, b->stream));
      if (!f6.empty()) BHIP(b, hipMemcpyAsync(tmp + f4.size(), f6.data(), sizeof(GWork) * f6.size(), hipMemcpyHostToDevice, b->stream));
      if (!f4.empty()) hipLaunchKernelGGL(kb_bound<4>, dim3((unsigned)f4.size()), dim3(256), 0, b->stream, b->d_inst, b->d_bs, tmp, 0);
      if (!f6.empty()) hipLaunchKernelGGL(kb_bound<6>, dim3((unsigned)f6.size()), dim3(256), 0, b->stream, b->d_inst, b->d_bs, tmp + f4.size(), 0);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return bfail(b, VACGPU_E_HIP, std::string("kb_bound: ") + hipGetErrorString(e));
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    if (tmp) BHIP(b, hipFree(tmp));
    b->fresh.clear();
  }
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) { (void)hipGraphExecDestroy(b->graph[p]); b->graph[p] = nullptr; }
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
  // Courant dt
  for (auto& kv : b->lists) {
    const auto& L = kv.second;
    if (!L.all.n) continue;
    const dim3 blk(BX, kv.first.second);
    BDISPATCH(kv.first.first.mhd, kv.first.first.f32, {
      (void)sizeof(R);
      hipLaunchKernelGGL((kb_pminmax<ND, MHD>), dim3(L.all.n), blk, 0, b->stream, b->d_inst, b->d_bs, L.all.d, b->part, b->pstride);
    });
  }
  if (b->nids) hipLaunchKernelGGL(kb_fin_smallp, dim3(b->nids), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_ids, b->part, b->pstride);
  for (int d = 0; d < ND; ++d)
    for (auto& kv : b->lists) {
      const auto& L = kv.second;
      if (!L.all.n) continue;
      const dim3 blk(BX, kv.first.second);
      BDISPATCH(kv.first.first.mhd, kv.first.first.f32, {
        (void)sizeof(R);
        hipLaunchKernelGGL((kb_cmaxdx<ND, MHD>), dim3(L.all.n), blk, 0, b->stream, b->d_inst, b->d_bs, L.all.d, b->part, b->pstride, d);
      });
    }
  if (b->nids) hipLaunchKernelGGL(kb_fin_getdt, dim3(b->nids), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_ids, b->part, b->pstride);
  if ((rc = launch_err(b, "getdt"))) return rc;
  // sweeps: phase class c has even it when (c + parity) is even
  for (int k = 0; k < ND; ++k) {
    for (int pass = 0; pass < 2; ++pass) {  // 0 predictor, 1 corrector
      for (auto& kv : b->lists) {
        const auto& L = kv.second;
        const dim3 blk(BX, kv.first.second);
        const int lim = kv.first.first.lim;
        for (int c = 0; c < 2; ++c) {
          const Launch& W = L.ph[c];
          if (!W.n) continue;
          const bool fwd = ((c + parity) % 2) == 0;
          const int D = fwd ? k : ND - 1 - k;
          if (lim >= 0) {  // fused sweep: one launch, in the corrector pass
            if (pass == 1) {
              const dim3 fb(64, 4);
#define KBF(DV, LV) hipLaunchKernelGGL((kb_fused<MHD, R, DV, LV>), dim3(W.n), fb, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride)
#define KBFL(DV) do { if (lim == LIM_MINMOD) KBF(DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) KBF(DV, LIM_WOODWARD); else KBF(DV, LIM_SUPERBEE); } while (0)
              BDISPATCH(kv.first.first.mhd, kv.first.first.f32, {
                if (D == 0) KBFL(0); else KBFL(1);
              });
#undef KBFL
#undef KBF
            }
            continue;
          }
          BDISPATCH(kv.first.first.mhd, kv.first.first.f32, {
            if (pass == 0) {
              if (D == 0) hipLaunchKernelGGL((kb_pred<ND, MHD, R, 0>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k);
              else hipLaunchKernelGGL((kb_pred<ND, MHD, R, 1>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k);
            } else {
              if (D == 0) hipLaunchKernelGGL((kb_corr<ND, MHD, R, 0>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride);
              else hipLaunchKernelGGL((kb_corr<ND, MHD, R, 1>), dim3(W.n), blk, 0, b->stream, b->d_inst, b->d_bs, W.d, k, b->part, b->pstride);
            }
          });
        }
      }
      if ((rc = launch_err(b, pass ? "kb_corr" : "kb_pred"))) return rc;
      if (pass == 0) {
        if ((rc = bound_all(b, 2))) return rc;
      } else {
        if (b->nids) hipLaunchKernelGGL(kb_fin_courant, dim3(b->nids), dim3(256), 0, b->stream, b->d_inst, b->d_bs, b->d_ids, b->part, b->pstride, k);
        if ((rc = bound_all(b, k == 0 ? 1 : 0))) return rc;
      }
    }
  }
  if (b->nids) hipLaunchKernelGGL(kb_step_end, dim3((b->nids + 255) / 256), dim3(256), 0, b->stream, b->d_bs, b->d_ids, b->nids);
  return launch_err(b, "kb_step_end");
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
  hipError_t e = hipMalloc(&b->d_inst, sizeof(InstDev) * capacity);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_inst, sizeof(InstDev) * capacity, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->d_bs, sizeof(BScal) * capacity);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_bs, sizeof(BScal) * capacity, hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fprintf(stderr, "vacgpu_batch_create: %s\n", hipGetErrorString(e));
    vacgpu_batch_destroy(b);
    return VACGPU_E_HIP;
  }
  memset(b->h_bs, 0, sizeof(BScal) * capacity);
  for (int i = 0; i < capacity; ++i) b->h_bs[i].done = 1;
  if (hipMemcpy(b->d_bs, b->h_bs, sizeof(BScal) * capacity, hipMemcpyHostToDevice) != hipSuccess) {
    vacgpu_batch_destroy(b);
    return VACGPU_E_HIP;
  }
  *out = b;
  return 0;
}

int vacgpu_batch_add(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
                     const double* dx0, const double* dx1, const double* w, const int* ld,
                     int it, double t, int itmax, double tmax, int tmaxexact, int* slot) {
  if (!b || !cfg || !m || !dx0 || !dx1 || !w || !slot) return VACGPU_E_ARG;
  if (cfg->ndim != 2) return bfail(b, VACGPU_E_ARG, "the batch engine runs 2D grids");
  if (m->typefull != VACGPU_TVDLF) return bfail(b, VACGPU_E_ARG, "the batch engine runs typefull='tvdlf'");
  if (m->dimsplit_xy) return bfail(b, VACGPU_E_ARG, "the batch engine runs typedimsplit='xyyx'");
  if (m->typelimited != VACGPU_LIM_ORIGINAL) return bfail(b, VACGPU_E_ARG, "the batch engine runs typelimited='original'");
  if (m->useprimitive) return bfail(b, VACGPU_E_ARG, "the batch engine limits the conserved variables (useprimitive=F)");
  if (m->unsplit) return bfail(b, VACGPU_E_ARG, "the batch engine runs dimensionally split sweeps");
  if (m->dtcantgrow || m->slowsteps)
    return bfail(b, VACGPU_E_ARG, "the batch engine's time step has no dtcantgrow / slowsteps");
  if (m->courantpar <= 0.0) return bfail(b, VACGPU_E_ARG, "the batch engine needs courantpar > 0");
  if (vacgpu_supported(cfg, m) != 0) return bfail(b, VACGPU_E_ARG, "unsupported selectors (see vacgpu_supported)");
  for (int f = 0; f < 4; ++f)
    for (int iw = 0; iw < 8; ++iw)
      if (cfg->typeB[f][iw] != VACGPU_CONT && cfg->typeB[f][iw] != VACGPU_PERIODIC)
        return bfail(b, VACGPU_E_ARG, "the batch engine runs 'cont' and 'periodic' boundaries");
  const long long need = footprint(*cfg);
  if (b->used_bytes + need > b->budget) return bfail(b, VACGPU_E_NOMEM, "HBM budget exhausted");
  int s = -1;
  for (int i = 0; i < b->cap; ++i)
    if (!b->inst[i].used) { s = i; break; }
  if (s < 0) return bfail(b, VACGPU_E_NOMEM, "no free slot");
  HostInst& h = b->inst[s];
  h = HostInst();
  h.cfg = *cfg;
  h.m = *m;
  h.bytes = need;
  h.klass = size_class(*cfg);
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
    h.fused = ok;
    h.lim = ok ? m->typelimiter[0] : -1;
  }
  long long slab[3];
  I.nbound = ghost_slabs(I.geo, slab);
  I.slab0 = slab[0];
  I.slab1 = slab[1];
  const size_t plane_bytes = plane_bytes_of(I.geo, nw);
  const size_t dx_bytes = dx_bytes_of(I.geo);
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
  BHIP(b, hipMemcpyAsync(dxd, dx0, sizeof(double) * I.geo.ng[0], hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(dxd + I.geo.ng[0], dx1, sizeof(double) * I.geo.ng[1], hipMemcpyHostToDevice, b->stream));
  // state (full ixG with caller leading dimensions), through pinned staging:
  // one host copy, then pitched DMA per variable
  const int l0 = ld ? ld[0] : I.geo.ng[0], l1 = ld ? ld[1] : I.geo.ng[1];
  const size_t hbytes = sizeof(double) * (size_t)nw * l0 * l1;
  int rc;
  if ((rc = ensure_stage(b, hbytes))) {
    pool_put(b, plane_bytes, I.w);
    pool_put(b, plane_bytes, I.b);
    pool_put(b, plane_bytes, I.w1);
    pool_put(b, dx_bytes, dxd);
    return rc;
  }
  BHIP(b, hipStreamSynchronize(b->stream));  // the stage may still feed an earlier copy
  memcpy(b->stage, w, hbytes);
  for (int iw = 0; iw < nw; ++iw)
    BHIP(b, hipMemcpy2DAsync(I.w + iw * I.geo.ps + I.geo.off, sizeof(double) * I.geo.s[1],
                             b->stage + (size_t)iw * l0 * l1, sizeof(double) * l0,
                             sizeof(double) * I.geo.ng[0], I.geo.ng[1], hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemsetAsync(I.b, 0, plane_bytes, b->stream));
  BHIP(b, hipMemsetAsync(I.w1, 0, plane_bytes, b->stream));
  BScal bs{};
  bs.s.dt = 0.0;
  bs.s.t = t;
  bs.s.smallp = -1.0;
  for (int d = 0; d < 3; ++d) bs.s.dtcourant[d] = kBig;
  bs.s.tmax = tmax;
  bs.s.tmaxexact = tmaxexact;
  bs.s.it = it;
  bs.itmax = itmax;
  bs.done = 0;
  bs.smallpcoeff = m->smallpcoeff;
  b->h_bs[s] = bs;
  BHIP(b, hipMemcpyAsync(b->d_bs + s, &b->h_bs[s], sizeof(BScal), hipMemcpyHostToDevice, b->stream));
  h.used = true;
  // which phase list: parity of it at the next step is (it + 0) while the
  // batch is at step_count, so class = (it - step_count) mod 2
  h.phase_class = (int)(((it - b->step_count) % 2 + 2) % 2);
  b->used_bytes += need;
  b->dirty = true;
  b->fresh.push_back(s);
  *slot = s;
  // the caller may release its host arrays on return
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int vacgpu_batch_remove(vacgpu_batch* b, int slot) {
  if (!b || slot < 0 || slot >= b->cap || !b->inst[slot].used) return VACGPU_E_ARG;
  // nothing in flight may still use the buffers when they return to the pool
  BHIP(b, hipStreamSynchronize(b->stream));
  HostInst& h = b->inst[slot];
  const int nw = h.dev.mhd ? 6 : 4;
  const size_t plane_bytes = plane_bytes_of(h.dev.geo, nw);
  const size_t dx_bytes = dx_bytes_of(h.dev.geo);
  pool_put(b, plane_bytes, h.dev.w);
  pool_put(b, plane_bytes, h.dev.b);
  pool_put(b, plane_bytes, h.dev.w1);
  pool_put(b, dx_bytes, const_cast<double*>(h.dev.geo.dx[0]));
  b->used_bytes -= h.bytes;
  h = HostInst();
  b->h_bs[slot].done = 1;
  BHIP(b, hipMemcpyAsync(b->d_bs + slot, &b->h_bs[slot], sizeof(BScal), hipMemcpyHostToDevice, b->stream));
  b->dirty = true;
  return 0;
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

int vacgpu_batch_sync(vacgpu_batch* b) {
  if (!b) return VACGPU_E_ARG;
  BHIP(b, hipMemcpyAsync(b->h_bs, b->d_bs, sizeof(BScal) * b->cap, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int vacgpu_batch_status(vacgpu_batch* b, int slot, int* it, double* t, double* dt, int* done,
                        int* nerror) {
  if (!b || slot < 0 || slot >= b->cap) return VACGPU_E_ARG;
  const BScal& S = b->h_bs[slot];
  if (it) *it = S.s.it;
  if (t) *t = S.s.t;
  if (dt) *dt = S.s.dt;
  if (done) *done = S.done;
  if (nerror)
    for (int i = 0; i < 4; ++i) nerror[i] = S.s.nerror[i];
  return 0;
}

int vacgpu_batch_get_state(vacgpu_batch* b, int slot, double* w, const int* ld) {
  if (!b || !w || slot < 0 || slot >= b->cap || !b->inst[slot].used) return VACGPU_E_ARG;
  const HostInst& h = b->inst[slot];
  const InstDev& I = h.dev;
  const int nw = I.mhd ? 6 : 4;
  const int l0 = ld ? ld[0] : I.geo.ng[0], l1 = ld ? ld[1] : I.geo.ng[1];
  const size_t hbytes = sizeof(double) * (size_t)nw * l0 * l1;
  int rc;
  BHIP(b, hipStreamSynchronize(b->stream));
  if ((rc = ensure_stage(b, hbytes))) return rc;
  for (int iw = 0; iw < nw; ++iw)
    BHIP(b, hipMemcpy2DAsync(b->stage + (size_t)iw * l0 * l1, sizeof(double) * l0,
                             I.w + iw * I.geo.ps + I.geo.off, sizeof(double) * I.geo.s[1],
                             sizeof(double) * I.geo.ng[0], I.geo.ng[1], hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  // rows beyond ixG in the caller's layout are left as the caller had them
  for (int iw = 0; iw < nw; ++iw)
    for (int y = 0; y < I.geo.ng[1]; ++y)
      memcpy(w + (size_t)iw * l0 * l1 + (size_t)y * l0, b->stage + (size_t)iw * l0 * l1 + (size_t)y * l0,
             sizeof(double) * I.geo.ng[0]);
  return 0;
}

int vacgpu_batch_graph(vacgpu_batch* b, int enable) {
  if (!b) return VACGPU_E_ARG;
  BHIP(b, hipStreamSynchronize(b->stream));
  b->use_graph = enable != 0;
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) { (void)hipGraphExecDestroy(b->graph[p]); b->graph[p] = nullptr; }
  return 0;
}

// Autotuner: for every size class present, time the predictor + corrector
// launches of both directions for each candidate tile shape (64 x rows) and
// keep the fastest.  The timed launches write only the scratch arrays b and
// w1, so the state is unchanged.
int vacgpu_batch_tune(vacgpu_batch* b, int* rows_out) {
  if (!b) return VACGPU_E_ARG;
  int rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  Work* tmp = nullptr;
  int tmp_cap = 0;
  for (int c = 0; c < NCLASS; ++c) {
    // tiles of class-c instances per physics/tier key, for each candidate shape
    bool present = false;
    for (const auto& h : b->inst) present |= (h.used && h.klass == c && !(h.fused && b->use_fused));
    if (!present) continue;
    float best = 1e30f;
    int best_rows = b->rows[c];
    for (int r : kTileRows) {
      std::map<Key, std::vector<Work>> lists;
      for (int i = 0; i < (int)b->inst.size(); ++i) {
        const HostInst& h = b->inst[i];
        if (!h.used || h.klass != c) continue;
        Key k{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK};
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
            hipLaunchKernelGGL((kb_corr<ND, MHD, R, 0>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0, b->part, b->pstride);
            hipLaunchKernelGGL((kb_pred<ND, MHD, R, 1>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0);
            hipLaunchKernelGGL((kb_corr<ND, MHD, R, 1>), dim3(n), blk, 0, b->stream, b->d_inst, b->d_bs, W, 0, b->part, b->pstride);
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
        b->h_inst[i].seglen = sl;
        const Geo& g = h.dev.geo;
        const int strips = (g.n[0] + FS_STRIDE - 1) / FS_STRIDE;
        const int sblocks = ((g.n[1] + sl - 1) / sl + 3) / 4;
        Key k{h.cfg.physics == VACGPU_MHD, h.cfg.precision == VACGPU_FP32BULK, h.lim};
        for (int sb = 0; sb < sblocks; ++sb)
          for (int st = 0; st < strips; ++st) lists[k].push_back(Work{i, st, sb, 0, 0});
      }
      BHIP(b, hipMemcpyAsync(b->d_inst, b->h_inst, sizeof(InstDev) * b->inst.size(), hipMemcpyHostToDevice, b->stream));
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
#define KBT(DV, LV) hipLaunchKernelGGL((kb_fused<MHD, R, DV, LV>), dim3(n), fb, 0, b->stream, b->d_inst, b->d_bs, W, 0, b->part, b->pstride)
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
  BHIP(b, hipStreamSynchronize(b->stream));
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
  for (auto& h : b->inst)
    if (h.used) {
      (void)hipFree(h.dev.w);
      (void)hipFree(h.dev.b);
      (void)hipFree(h.dev.w1);
      (void)hipFree(const_cast<double*>(h.dev.geo.dx[0]));
    }
  for (auto& kv : b->lists) {
    if (kv.second.all.d) (void)hipFree(kv.second.all.d);
    for (int c = 0; c < 2; ++c)
      if (kv.second.ph[c].d) (void)hipFree(kv.second.ph[c].d);
  }
  for (auto& kv : b->pool) (void)hipFree(kv.second);
  if (b->gw4.d) (void)hipFree(b->gw4.d);
  if (b->gw6.d) (void)hipFree(b->gw6.d);
  if (b->d_ids) (void)hipFree(b->d_ids);
  if (b->part) (void)hipFree(b->part);
  for (int p = 0; p < 2; ++p)
    if (b->graph[p]) (void)hipGraphExecDestroy(b->graph[p]);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_bs) (void)hipFree(b->d_bs);
  if (b->h_bs) (void)hipHostFree(b->h_bs);
  if (b->h_inst) (void)hipHostFree(b->h_inst);
  if (b->stage) (void)hipHostFree(b->stage);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
