// vacgpu.hip -- MI355X (gfx950) kernels and C ABI for VAC's explicit
// TVDLF / TVD (Roe) advect step, ideal MHD and HD, 2D and 3D.
//
// Data layout in HBM: structure of arrays, one plane per conserved variable,
// x fastest, the full ixG box (mesh + 2 ghost layers per side).  Rows are
// padded to a multiple of 16 doubles and each plane is shifted so that the
// first mesh cell of every row (x index 2) starts a 128-byte line.
//
// Kernels compute one mesh cell per thread with the 64 lanes of a wavefront
// along x (coalesced loads for every sweep direction); the interface values
// a cell needs are recomputed from the neighbouring states rather than
// staged through HBM.  Ghost cells are refreshed by a separate kernel after
// every sub-step, as VAC's advect1 does (src/vac.t:567).  Time step, Courant
// limits, error counters and the pressure floor live in device memory, so a
// sequence of steps runs without host synchronisation.  Reductions write one
// partial per block and a one-block kernel folds them (no atomics).
#include "../../include/vacgpu.h"
#include "physics.hpp"
#include "sweep.hpp"
#include "project.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace vacgpu;


// ======================================================================
// host side
// ======================================================================

struct vacgpu_grid {
  vacgpu_config cfg{};
  Geo geo{};
  BC bc{};
  int nw = 0;
  double* buf[3] = {nullptr, nullptr, nullptr};
  // a fourth state buffer, allocated on first need: the unsplit corrector's
  // intermediate when wold (kept from before a projection) holds the third
  double* xbuf = nullptr;
  double* w1 = nullptr;
  double* dxdev = nullptr;
  Scal* sc = nullptr;
  Scal* sc_host = nullptr;  // pinned
  double* part = nullptr;
  long long pstride = 0;
  dim3 grid, block;
  dim3 grid_tx;              // k_tvd_x
  dim3 grid_tm[3];           // k_tvd_m per sweep direction
  int seglen = 16;
  // fused 2D TVDLF sweeps: eligible when every face has one boundary type
  // for all variables; use_fused selects them (default on)
  bool fused_ok = false, use_fused = true;
  bool bc_uniform = false;  // one boundary type per face for all variables, n >= 4
  // 'fixed', 'fixed1' and 'grad1' faces: a copy of the state after the first
  // ghost fill, whose ghost values (differences to the edge for 'grad1')
  // those faces keep (null until the first vacgpu_put_state)
  bool any_fixed = false;
  double* fixw = nullptr;
  // 'symm0' faces: the variables with zero flux through each face
  unsigned char nf[6] = {0, 0, 0, 0, 0, 0};
  bool nfany = false;
  int fs_seglen[2] = {16, 16};  // fused TVDLF sweeps, per direction (x, y)
  dim3 grid_fs[2], block_fs;
  int ry_seglen = 16;  // k_ryu_ab_y (2D one-pass Ryu step)
  dim3 grid_ry;
  double* stage = nullptr;  // pinned staging of put_state / get_state
  unsigned long long* fold = nullptr;  // Courant fold words of the fused sweeps (zero between launches)
  // Courant words of the TVD sweeps (cs_max; 6 rows of CS_N, zero between folds)
  unsigned long long* cslot = nullptr;
  // device loop without a per-step host sync: the TVD sweeps' Courant folds
  // (and the step's it / t) are left to the next step's k_getdt_lagged
  bool defer_folds = false, pend_step_end = false;
  int pend_mask = 0;
  double pend_courantpar = 0.0;
  // device loop: the last fused sweep of a step also advances it and t
  bool fold_step_end = false, step_end_done = false;
  size_t stage_bytes = 0;
  long long nbound = 0, slab[3] = {0, 0, 0};
  int cur = 0, old = 0;
  // a state kept as the next step's wold (the state before a projection at
  // the start of a step, as VAC stores wold before advance), or -1
  int old_pending = -1;
  bool mid_step = false;  // between a step's first and last sweep
  // the state's ghost cells are not all copies of mesh cells (after a
  // projection: B's second ghost layer keeps its old values, and 'nul'
  // Poisson faces give the first layer values of its own); the next sweep
  // then reads ghost cells from memory instead of mapping them to the mesh
  bool ghosts_dirty = false;
  // projection of B (process / projectb): work planes, CG scalars, the
  // values projectb saves across calls, and the processing schedule
  struct Proj {
    double* rhs = nullptr;
    double* phi = nullptr;
    double* z = nullptr;
    double* az = nullptr;
    double* z2 = nullptr;   // the search direction alternates between z and z2
    double* rdx[3] = {nullptr, nullptr, nullptr};  // pdiv's reciprocals of dx
    Cg* cg = nullptr;
    Cg* cg_host = nullptr;  // pinned
    double divbmax = 0.0;
    int last_iters = 0;  // CG steps of the previous solve (the first chunk's length)
    int nproc1 = 0, itmin = 0;
    double procpar[3] = {-1.0, -1.0, -1.0};
  } pr;
  bool have_state = false;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // asynchronous snapshots: a device copy taken in stream order, then copied
  // to pinned host memory on copy_stream while the steps go on
  struct Snap {
    double* dev = nullptr;
    double* host = nullptr;
    Scal* dsc = nullptr;
    Scal* hsc = nullptr;
    hipEvent_t copied = nullptr, done = nullptr;
    bool pending = false, valid = false;
  } snap[VACGPU_SNAP_SLOTS];
  hipStream_t copy_stream = nullptr;
  std::string err;
  // profiling
  bool prof = false;
  struct Ev { int k; hipEvent_t a, b; };
  std::vector<Ev> evs;
  double tsum[KF_N] = {};
  long tcnt[KF_N] = {};
};

namespace {

int fail(vacgpu_grid* g, int code, const std::string& msg) {
  if (g) g->err = msg;
  return code;
}

#define HIPCHK(g, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(g, VACGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

struct Timed {
  vacgpu_grid* g;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(vacgpu_grid* g_, int k_) : g(g_), k(k_) {
    if (g->prof) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, g->stream);
    }
  }
  ~Timed() {
    if (g->prof) {
      (void)hipEventRecord(b, g->stream);
      g->evs.push_back({k, a, b});
    }
  }
};

void harvest(vacgpu_grid* g) {
  for (auto& e : g->evs) {
    float ms = 0.f;
    if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      g->tsum[e.k] += ms;
      g->tcnt[e.k] += 1;
    }
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  g->evs.clear();
}

int check_config(const vacgpu_config* c, std::string& why) {
  if (!c) { why = "null config"; return VACGPU_E_ARG; }
  if (c->ndim != 2 && c->ndim != 3) { why = "ndim must be 2 or 3 (setvac -d=22 or -d=33)"; return VACGPU_E_ARG; }
  if (c->physics != VACGPU_MHD && c->physics != VACGPU_HD) { why = "physics must be mhd or hd"; return VACGPU_E_ARG; }
  for (int d = 0; d < c->ndim; ++d)
    if (c->nx[d] < 2) { why = "nx must be >= 2 in every dimension"; return VACGPU_E_ARG; }
  if (!(c->eqpar[0] > 0.0)) { why = "eqpar(gamma_) must be > 0 (ideal gas)"; return VACGPU_E_ARG; }
  if (c->physics == VACGPU_MHD && c->eqpar[1] != 0.0) { why = "eqpar(eta_) must be 0: resistive MHD is not on the GPU path"; return VACGPU_E_ARG; }
  if (c->precision != VACGPU_FP64 && c->precision != VACGPU_FP32BULK) { why = "precision must be FP64 or FP32BULK"; return VACGPU_E_ARG; }
  int nw = c->physics == VACGPU_MHD ? 2 + 2 * c->ndim : 2 + c->ndim;
  for (int f = 0; f < 2 * c->ndim; ++f)
    for (int iw = 0; iw < nw; ++iw)
      if (c->typeB[f][iw] < VACGPU_CONT || c->typeB[f][iw] > VACGPU_SYMM0) {
        why = "typeB must be 'cont', 'periodic', 'symm', 'symm0', 'asymm', 'fixed', 'cont1', 'fixed1' or 'grad1'";
        return VACGPU_E_ARG;
      }
  for (int d = 0; d < c->ndim; ++d)
    for (int iw = 0; iw < nw; ++iw)
      if ((c->typeB[2 * d][iw] == VACGPU_PERIODIC) != (c->typeB[2 * d + 1][iw] == VACGPU_PERIODIC)) {
        why = "periodic boundaries need a periodic pair (src/vacgrid.t:107-120)";
        return VACGPU_E_ARG;
      }
  return 0;
}

int check_method(const vacgpu_config* c, const vacgpu_method* m, std::string& why) {
  if (!m) { why = "null method"; return VACGPU_E_ARG; }
  if (m->typefull != VACGPU_TVDLF && m->typefull != VACGPU_TVD && m->typefull != VACGPU_TVDMU) { why = "typefull must be 'tvdlf', 'tvd' or 'tvdmu'"; return VACGPU_E_ARG; }
  if (m->typelimited < VACGPU_LIM_ORIGINAL || m->typelimited > VACGPU_LIM_PREDICTOR) { why = "typelimited must be 'original', 'previous' or 'predictor'"; return VACGPU_E_ARG; }
  if (m->typefull == VACGPU_TVD && m->typelimited == VACGPU_LIM_PREDICTOR) { why = "typelimited='predictor' with typefull='tvd' (limiting the centred-difference result) is not on the GPU path"; return VACGPU_E_ARG; }
  {
    const int nw = c->physics == VACGPU_MHD ? 2 + 2 * c->ndim : 2 + c->ndim;
    auto speeds = [&](int il) {
      const int t = m->typeentropy[il];
      return t == VACGPU_ENT_HARTEN || t == VACGPU_ENT_POWELL || t == VACGPU_ENT_RATIO;
    };
    for (int il = 0; il < nw; ++il) {
      const int t = m->typeentropy[il];
      if (t < VACGPU_ENT_NUL || t > VACGPU_ENT_YEE) { why = "typeentropy must be 'nul', 'harten', 'powell', 'ratio' or 'yee'"; return VACGPU_E_ARG; }
      if (t == VACGPU_ENT_RATIO && c->physics == VACGPU_HD) { why = "typeentropy='ratio' needs the MHD Roe solver"; return VACGPU_E_ARG; }
    }
    // VAC's MHD solver computes the sound speeds for the fast right wave and
    // the Alfven speeds for the right Alfven wave, and the others reuse them
    if (c->physics == VACGPU_MHD) {
      if ((speeds(1) || speeds(2) || speeds(3)) && !speeds(0)) { why = "typeentropy: the fast and slow waves take their speeds from the fast right wave's fix (give it 'harten', 'powell' or 'ratio' too)"; return VACGPU_E_ARG; }
      if (c->ndim == 3 && speeds(7) && !speeds(6)) { why = "typeentropy: the left Alfven wave takes its speeds from the right one's fix"; return VACGPU_E_ARG; }
    }
  }
  int nw = c->physics == VACGPU_MHD ? 2 + 2 * c->ndim : 2 + c->ndim;
  for (int iw = 0; iw < nw; ++iw)
    if (m->typelimiter[iw] < VACGPU_MINMOD || m->typelimiter[iw] > VACGPU_SUPERBEE) { why = "typelimiter must be minmod, woodward or superbee"; return VACGPU_E_ARG; }
  if (c->physics == VACGPU_MHD && m->divbfix && !m->sourceunsplit) { why = "divbfix=T requires sourceunsplit=T (split sources are not on the GPU path)"; return VACGPU_E_ARG; }
  if (!(m->smallpcoeff >= 0.0)) { why = "smallpcoeff must be >= 0"; return VACGPU_E_ARG; }
  if (m->useprimitive != 0 && m->useprimitive != 1) { why = "useprimitive must be 0 or 1"; return VACGPU_E_ARG; }
  if (m->unsplit != 0 && m->unsplit != 1) { why = "unsplit must be 0 or 1"; return VACGPU_E_ARG; }
  if (m->unsplit && m->typefull == VACGPU_TVD) { why = "unsplit sweeps (dimsplit=F) are on the GPU path for typefull='tvdlf' and 'tvdmu' (VAC warns against one-step TVD unsplit, src/vacio.t:508-509)"; return VACGPU_E_ARG; }
  if (m->dtcantgrow != 0 && m->dtcantgrow != 1) { why = "dtcantgrow must be 0 or 1"; return VACGPU_E_ARG; }
  if (m->slowsteps < 0) { why = "slowsteps must be >= 0"; return VACGPU_E_ARG; }
  if (m->useprimitive && m->typefull != VACGPU_TVD) {
    if (c->precision != VACGPU_FP64) { why = "useprimitive=T with TVDLF / TVD-MUSCL runs in FP64 only"; return VACGPU_E_ARG; }
  }
  return 0;
}

Sel make_sel(const vacgpu_grid* g, const vacgpu_method* m) {
  Sel s{};
  for (int i = 0; i < 8; ++i) s.lim[i] = m->typelimiter[i];
  s.divbfix = m->divbfix;
  s.divbwave = m->divbwave;
  s.sourceunsplit = m->sourceunsplit;
  s.courantpar = m->courantpar;
  s.gamma = g->cfg.eqpar[0];
  for (int i = 0; i < 8; ++i) {
    s.ent[i] = m->typefull == VACGPU_TVDLF ? ENT_NUL : m->typeentropy[i];
    s.ecoef[i] = m->entropycoef[i];
  }
  s.prim = m->useprimitive;
  s.srcmul = 1;
  // a split sweep limiting its own input ('original'): the predictor's
  // upwindLR has converted it once, the corrector's converts it again
  s.rtw = 1;
  s.rtb = 2;
  s.rtu = 0;
  s.rts = 0;
  for (int f = 0; f < 6; ++f) s.nf[f] = g->nf[f];
  return s;
}

// the round trips of a corrector with useprimitive (Sel::rtw ...) for
// direction k (0-based, in turn) of an advect over nd directions (nd = 1
// for a split sweep): the predictor's upwindLR converted the sweep's input
// once per direction; the corrector's converts its limiter's base, which
// for 'previous' is wold (round-tripped in place after each direction by
// k_roundtrip) and for 'predictor' w1
void set_prim_roundtrips(Sel& s, int typelimited, int k, int nd) {
  const int base = k == 0 ? nd : 0;  // later directions update the previous one's output
  if (typelimited == VACGPU_LIM_ORIGINAL) {
    s.rtw = 1; s.rtb = 2; s.rtu = 0; s.rts = 0;
  } else if (typelimited == VACGPU_LIM_PREVIOUS) {
    s.rtw = 0; s.rtb = base; s.rtu = 0; s.rts = 0;
  } else {
    s.rtw = k; s.rtb = base; s.rtu = k; s.rts = nd;
  }
}



// an entropy fix on any wave of the TVD / TVD-MUSCL Roe dissipation
bool entropy_on(const vacgpu_grid* g, const vacgpu_method* m) {
  if (m->typefull == VACGPU_TVDLF) return false;
  for (int i = 0; i < g->nw; ++i)
    if (m->typeentropy[i] != VACGPU_ENT_NUL) return true;
  return false;
}

// template dispatch over (ndim, physics, precision)
#define DISPATCH(g, ...)                                                               \
  do {                                                                                 \
    const bool mhd_ = (g)->cfg.physics == VACGPU_MHD;                                  \
    const bool f32_ = (g)->cfg.precision == VACGPU_FP32BULK;                           \
    if ((g)->cfg.ndim == 2) {                                                          \
      constexpr int ND = 2;                                                            \
      if (mhd_) { constexpr bool MHD = true;                                           \
        if (f32_) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
      else { constexpr bool MHD = false;                                               \
        if (f32_) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
    } else {                                                                           \
      constexpr int ND = 3;                                                            \
      if (mhd_) { constexpr bool MHD = true;                                           \
        if (f32_) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
      else { constexpr bool MHD = false;                                               \
        if (f32_) { using R = float; __VA_ARGS__ } else { using R = double; __VA_ARGS__ } } \
    }                                                                                  \
  } while (0)

int launch_check(vacgpu_grid* g, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(g, VACGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// ghost cells of variables iw0 <= iw < iw1 (all of them by default)
int bound(vacgpu_grid* g, double* w, int iw0 = 0, int iw1 = 8) {
  if (g->nbound == 0) return 0;
  Timed t(g, KF_BOUND);
  const unsigned nb = (unsigned)((g->nbound + 255) / 256);
#define BOUND_NW(NWV) hipLaunchKernelGGL(k_bound<NWV>, dim3(nb), dim3(256), 0, g->stream, g->geo, g->bc, w, g->nbound, g->slab[0], g->slab[1], iw0, iw1, g->fixw)
  switch (g->nw) {
    case 4: BOUND_NW(4); break;
    case 5: BOUND_NW(5); break;
    case 6: BOUND_NW(6); break;
    case 8: BOUND_NW(8); break;
    default: return fail(g, VACGPU_E_ARG, "unsupported nw");
  }
#undef BOUND_NW
  return launch_check(g, "k_bound");
}

int other_buffer(const vacgpu_grid* g) {
  for (int i = 0; i < 3; ++i)
    if (i != g->cur && i != g->old) return i;
  return (g->cur + 1) % 3;
}

// one getdt_courant evaluation on the device, result in sc->dt
int getdt_device(vacgpu_grid* g, const vacgpu_method* m, double dtpar, int tmaxexact) {
  double* w = g->buf[g->cur];
  const long long nb = (long long)g->grid.x * g->grid.y * g->grid.z;
  const bool mhd = g->cfg.physics == VACGPU_MHD;
  {
    Timed t(g, KF_RED);
    DISPATCH(g, {
      (void)sizeof(R);
      hipLaunchKernelGGL((k_pminmax<ND, MHD>), g->grid, g->block, 0, g->stream, g->geo, w,
                         g->cfg.eqpar[0], g->sc, g->part, g->pstride);
      hipLaunchKernelGGL(k_fin_smallp, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part,
                         g->pstride, nb, m->smallpcoeff, (int)mhd);
      if (dtpar <= 0.0)
        hipLaunchKernelGGL((k_cmaxdx_all<ND, MHD>), g->grid, g->block, 0, g->stream, g->geo, w,
                           g->cfg.eqpar[0], g->sc, g->part, g->pstride);
    });
    hipLaunchKernelGGL(k_fin_getdt, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part, g->pstride,
                       nb, g->cfg.ndim, m->courantpar, dtpar, tmaxexact);
  }
  return launch_check(g, "getdt");
}

// Blocks of 256 threads of a kernel that the device holds at once: CUs
// times the kernel's resident blocks per CU, or 0 when the runtime cannot say.
long long resident_blocks(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess) return 0;
  return per_cu > 0 && cus > 0 ? (long long)per_cu * cus : 0;
}

// Segment length for a marching launch of `strips` strips over `rows` rows,
// four segments (waves) per block, `slots` resident blocks: every wave
// marches seglen rows after a prologue worth `prologue` rows, the waves run in
// rounds of `slots` blocks, and a partly filled last round costs a whole
// wave time; the length with the least estimated time wins.
int pick_seglen(long long strips, int rows, long long slots, double prologue) {
  int best_sl = 4;
  double best = 1e300;
  for (int sl = 4; sl <= 256; ++sl) {
    const long long segs = (rows + sl - 1) / sl;
    const long long blocks = strips * ((segs + 3) / 4);
    const double cost = (double)((blocks + slots - 1) / slots) * (sl + prologue);
    if (cost < best) {
      best = cost;
      best_sl = sl;
    }
  }
  return best_sl;
}

// launch shape of the fused TVDLF sweeps: strips of 62 x cells, waves of
// one segment of seglen rows each, four waves per block.  seglen 0 picks the
// length of each direction's launch by pick_seglen (a prologue of two
// predictor rows, about 1.5 rows of work) over that kernel's resident
// blocks (the x and y kernels may hold different numbers of waves), or about
// 4096 waves when the occupancy is unknown.
void set_fused_shape(vacgpu_grid* g, int seglen) {
  const int strips = (g->geo.n[0] + FS_STRIDE - 1) / FS_STRIDE;
  const void* fk[2] = {nullptr, nullptr};
  const bool mhd = g->cfg.physics == VACGPU_MHD, f32 = g->cfg.precision == VACGPU_FP32BULK;
#define FSLOT(MV, RT)                                                     \
  do {                                                                    \
    fk[0] = reinterpret_cast<const void*>(&k_tvdlf_fused<MV, RT, 0, 0>);  \
    fk[1] = reinterpret_cast<const void*>(&k_tvdlf_fused<MV, RT, 1, 0>);  \
  } while (0)
  if (mhd) { if (f32) FSLOT(true, float); else FSLOT(true, double); }
  else { if (f32) FSLOT(false, float); else FSLOT(false, double); }
#undef FSLOT
  g->block_fs = dim3(64, 4, 1);
  for (int d = 0; d < 2; ++d) {
    int sl = seglen;
    if (sl <= 0) {
      const long long slots = resident_blocks(fk[d]);
      if (slots > 0) {
        sl = pick_seglen(strips, g->geo.n[1], slots, 1.5);
      } else {
        const int want = (4096 + strips - 1) / strips;
        sl = (g->geo.n[1] + want - 1) / want;
      }
    }
    sl = sl < 4 ? 4 : (sl > 256 ? 256 : sl);
    g->fs_seglen[d] = sl;
    const int segs = (g->geo.n[1] + sl - 1) / sl;
    g->grid_fs[d] = dim3(strips, (segs + 3) / 4, 1);
  }
}

// launch shape of k_ryu_ab_y (2D): strips of 62 x cells, four segments per
// block; the prologue is two rows of wt, about 1.3 rows of work
void set_ryu_shape(vacgpu_grid* g) {
  const int strips = (g->geo.n[0] + RY_STRIDE - 1) / RY_STRIDE;
  const long long slots = resident_blocks(reinterpret_cast<const void*>(&k_ryu_ab_y));
  g->ry_seglen = slots > 0 ? pick_seglen(strips, g->geo.n[1], slots, 1.3) : 16;
  const int segs = (g->geo.n[1] + g->ry_seglen - 1) / g->ry_seglen;
  g->grid_ry = dim3(strips, (segs + 3) / 4, 1);
}

// getdt of the steps after the first in a device loop: one block, lagged
// dtcourant (see k_getdt_lagged)
int getdt_lagged(vacgpu_grid* g, const vacgpu_method* m, double dtpar, int tmaxexact) {
  Timed t(g, KF_RED);
  DISPATCH(g, {
    (void)sizeof(R);
    hipLaunchKernelGGL((k_getdt_lagged<ND, MHD>), dim3(1), dim3(1024), 0, g->stream, g->geo,
                       g->buf[g->cur], g->cfg.eqpar[0], g->sc, m->smallpcoeff, m->courantpar, dtpar,
                       tmaxexact, g->cslot, g->pend_mask, g->pend_step_end ? 1 : 0);
  });
  g->pend_mask = 0;
  g->pend_step_end = false;
  return launch_check(g, "k_getdt_lagged");
}

// folds that a device loop left pending when it stopped before the next
// step's getdt (an error in a step): applied as the sweeps would have
void flush_folds(vacgpu_grid* g) {
  for (int d = 0; d < 3; ++d) {
    if (!((g->pend_mask >> d) & 1)) continue;
    const bool last = (g->pend_mask >> (d + 1)) == 0;
    hipLaunchKernelGGL(k_fin_courant_tvd_cs, dim3(1), dim3(384), 0, g->stream, g->sc, g->cslot, d,
                       g->pend_courantpar, (last && g->pend_step_end) ? 1 : 0);
  }
  g->pend_mask = 0;
  g->pend_step_end = false;
}

// the limiter type shared by every variable / characteristic wave, or -1
int common_limiter(const vacgpu_grid* g, const vacgpu_method* m) {
  if (g->nfany) return -1;  // 'symm0' faces: the general variants zero their fluxes
  for (int iw = 1; iw < g->nw; ++iw)
    if (m->typelimiter[iw] != m->typelimiter[0]) return -1;
  return m->typelimiter[0];
}

// one dimensional sweep (advect, src/vac.t:363-486), dt already on device
// wold converted to primitive variables and back in place, as the
// corrector's upwindLR leaves it (useprimitive with 'previous')
int roundtrip_wold(vacgpu_grid* g) {
  const long long n = (long long)g->geo.ng[0] * g->geo.ng[1] * (g->cfg.ndim == 3 ? g->geo.ng[2] : 1);
  const unsigned nb = (unsigned)((n + 255) / 256);
  DISPATCH(g, {
    (void)sizeof(R);
    hipLaunchKernelGGL((k_roundtrip<ND, MHD>), dim3(nb), dim3(256), 0, g->stream, g->geo,
                       g->buf[g->old], g->cfg.eqpar[0], g->sc);
  });
  return launch_check(g, "k_roundtrip");
}

// Hancock predictor launch (w -> w1, ghost cells of w1 with ghosts set):
// primitive-variable limiting (FP64 only, checked with the method), else
// specialised for a limiter common to every variable
int launch_pred(vacgpu_grid* g, const vacgpu_method* m, int d, const Sel& s, const double* win,
                int ghosts, bool dw = false) {
  Timed t(g, KF_PRED);
  if (dw) {
    // FP32-bulk 3D: w1 as its FP32 increment over win
    const int lim = common_limiter(g, m);
#define PLW(DV, LV) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, DV, false, LV, true>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts)
#define PLWD(DV) do { if (lim == LIM_MINMOD) PLW(DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) PLW(DV, LIM_WOODWARD); else PLW(DV, LIM_SUPERBEE); } while (0)
    DISPATCH(g, {
      if constexpr (ND == 3 && std::is_same<R, float>::value) {
        if (d == 0) PLWD(0);
        else if (d == 1) PLWD(1);
        else PLWD(2);
      }
    });
#undef PLWD
#undef PLW
  } else if (m->useprimitive) {
    DISPATCH(g, {
      if constexpr (std::is_same<R, double>::value) {
        if (d == 0) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, 0, true>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts);
        else if (d == 1) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, 1, true>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts);
        else hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, (ND == 3 ? 2 : 1), true>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts);
      }
    });
  } else {
    const int lim = common_limiter(g, m);
#define PL(DV, LV) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, DV, false, LV>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts)
#define PLD(DV) do { if (lim == LIM_MINMOD) PL(DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) PL(DV, LIM_WOODWARD); else if (lim == LIM_SUPERBEE) PL(DV, LIM_SUPERBEE); else PL(DV, -1); } while (0)
    DISPATCH(g, {
      if (d == 0) PLD(0);
      else if (d == 1) PLD(1);
      else PLD((ND == 3 ? 2 : 1));
    });
#undef PLD
#undef PL
  }
  return launch_check(g, "k_tvdlf_pred");
}

// TVD-MUSCL corrector of sweep direction d (face-parallel: x one face per
// lane, y / z marching): win -> wout with the predictor state g->w1 and the
// limiter's base wl; nbc receives the block count of its Courant partials
int launch_tvdmu_corr(vacgpu_grid* g, const vacgpu_method* m, int d, const Sel& s,
                      const double* win, double* wout, const double* wl, int ghosts,
                      long long& nbc) {
  int rc = 0;
  const bool mhd = g->cfg.physics == VACGPU_MHD;
  {
    Timed t(g, KF_CORR);
    const bool f32 = g->cfg.precision == VACGPU_FP32BULK;
    // the general variant for entropy fixes and 'symm0' faces
    const bool ent = entropy_on(g, m) || g->nfany;
    // with entropy fixes the limiter types are read at run time
    const int lim = ent ? -1 : common_limiter(g, m);
  #define LIMSW(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else if (lim == LIM_SUPERBEE) X(LIM_SUPERBEE); else X(-1); } while (0)
    if (d == 0) {
      const dim3 gr((g->geo.n[0] + CX_STRIDE - 1) / CX_STRIDE, g->grid.y, g->grid.z);
      nbc = (long long)gr.x * gr.y * gr.z;
  #define MUX(NDV, RT, LV) do { if (ent) { if (mhd) MUXE(NDV, RT, true); else MUXE(NDV, RT, false); } else if (mhd) hipLaunchKernelGGL((k_tvdmu_x<NDV, RT, LV, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, ghosts, wl); \
                             else hipLaunchKernelGGL((k_tvdmu_x<NDV, RT, LV, false>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, ghosts, wl); } while (0)
  #define MUXE(NDV, RT, MH) hipLaunchKernelGGL((k_tvdmu_x<NDV, RT, -1, MH, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, ghosts, wl)
  #define X2D(LV) MUX(2, double, LV)
  #define X2F(LV) MUX(2, float, LV)
  #define X3D(LV) MUX(3, double, LV)
  #define X3F(LV) MUX(3, float, LV)
      if (m->useprimitive) {
        DISPATCH(g, {
          if constexpr (std::is_same<R, double>::value) {
            if (ent) hipLaunchKernelGGL((k_tvdmu_x<ND, R, -1, MHD, true, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, ghosts, wl);
            else hipLaunchKernelGGL((k_tvdmu_x<ND, R, -1, MHD, false, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, ghosts, wl);
          }
        });
      } else if (g->cfg.ndim == 2) { if (f32) LIMSW(X2F); else LIMSW(X2D); }
      else { if (f32) LIMSW(X3F); else LIMSW(X3D); }
  #undef X2D
  #undef X2F
  #undef X3D
  #undef X3F
  #undef MUX
  #undef MUXE
    } else {
      const dim3 gr = g->grid_tm[d];
      nbc = (long long)gr.x * gr.y * gr.z;
  #define MUME(NDV, RT, DV, MH) hipLaunchKernelGGL((k_tvdmu_m<NDV, RT, DV, -1, MH, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, g->seglen, ghosts, wl)
  #define MUM(NDV, RT, DV, LV) do { if (ent) { if (mhd) MUME(NDV, RT, DV, true); else MUME(NDV, RT, DV, false); } else if (mhd) hipLaunchKernelGGL((k_tvdmu_m<NDV, RT, DV, LV, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, g->seglen, ghosts, wl); \
                                 else hipLaunchKernelGGL((k_tvdmu_m<NDV, RT, DV, LV, false>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, g->seglen, ghosts, wl); } while (0)
  #define M21D(LV) MUM(2, double, 1, LV)
  #define M21F(LV) MUM(2, float, 1, LV)
  #define M31D(LV) MUM(3, double, 1, LV)
  #define M31F(LV) MUM(3, float, 1, LV)
  #define M32D(LV) MUM(3, double, 2, LV)
  #define M32F(LV) MUM(3, float, 2, LV)
      if (m->useprimitive) {
  #define MUP(DV) do { if (ent) hipLaunchKernelGGL((k_tvdmu_m<ND, R, DV, -1, MHD, true, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, g->seglen, ghosts, wl); \
                   else hipLaunchKernelGGL((k_tvdmu_m<ND, R, DV, -1, MHD, false, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->cslot, g->seglen, ghosts, wl); } while (0)
        DISPATCH(g, {
          if constexpr (std::is_same<R, double>::value) {
            if (d == 1) MUP(1); else MUP((ND == 3 ? 2 : 1));
          }
        });
  #undef MUP
      } else if (g->cfg.ndim == 2) { if (f32) LIMSW(M21F); else LIMSW(M21D); }
      else if (d == 1) { if (f32) LIMSW(M31F); else LIMSW(M31D); }
      else { if (f32) LIMSW(M32F); else LIMSW(M32D); }
  #undef M21D
  #undef M21F
  #undef M31D
  #undef M31F
  #undef M32D
  #undef M32F
  #undef MUM
  #undef MUME
    }
  #undef LIMSW
    if ((rc = launch_check(g, "k_tvdmu_corr"))) return rc;
  }
  return 0;
}

int sweep_device(vacgpu_grid* g, const vacgpu_method* m, int d, int firstsweep, int lastsweep) {
  if (firstsweep) {
    g->old = g->old_pending >= 0 ? g->old_pending : g->cur;
    g->old_pending = -1;
  }
  g->mid_step = !lastsweep;
  const bool dirty = g->ghosts_dirty;
  g->ghosts_dirty = false;
  const int outi = other_buffer(g);
  double* win = g->buf[g->cur];
  double* wout = g->buf[outi];
  double* wold = g->buf[g->old];
  Sel s = make_sel(g, m);
  if (m->useprimitive) set_prim_roundtrips(s, m->typelimited, 0, 1);
  const bool rt_wold = m->useprimitive && m->typelimited == VACGPU_LIM_PREVIOUS &&
                       m->typefull != VACGPU_TVD;
  const long long nb = (long long)g->grid.x * g->grid.y * g->grid.z;
  const bool mhd = g->cfg.physics == VACGPU_MHD;
  const bool src = mhd && m->divbfix && m->sourceunsplit;
  const bool courant = m->courantpar > 0.0;
  int rc = 0;
  // fused sweeps: 2D, one boundary type per face, one limiter for all
  // variables, and for MHD the Powell source on (VAC's defaults)
  bool one_lim = true;
  for (int iw = 1; iw < g->nw; ++iw) one_lim &= m->typelimiter[iw] == m->typelimiter[0];
  // the limiter's base (typelimited, src/vactvdlf.t:150-161)
  const double* wl = m->typelimited == VACGPU_LIM_PREVIOUS ? wold
                     : (m->typelimited == VACGPU_LIM_PREDICTOR ? g->w1 : win);
  const bool fused = m->typefull == VACGPU_TVDLF && g->cfg.ndim == 2 && g->fused_ok &&
                     g->use_fused && one_lim && (!mhd || src) && !dirty &&
                     m->typelimited == VACGPU_LIM_ORIGINAL && !m->useprimitive;
  if (fused) {
    {
      Timed t(g, KF_FUSED);
      const bool f32 = g->cfg.precision == VACGPU_FP32BULK;
      const int lim = m->typelimiter[0];
      const int step_end = (g->fold_step_end && lastsweep) ? 1 : 0;
      if (step_end) g->step_end_done = true;
#define FUSED_L(MV, RT, DV, LV) hipLaunchKernelGGL((k_tvdlf_fused<MV, RT, DV, LV>), g->grid_fs[DV], g->block_fs, 0, g->stream, g->geo, s, g->bc, win, wout, g->sc, g->fs_seglen[DV], g->fold, m->courantpar, step_end)
#define FUSED(MV, RT, DV) do { if (lim == LIM_MINMOD) FUSED_L(MV, RT, DV, LIM_MINMOD); else if (lim == LIM_WOODWARD) FUSED_L(MV, RT, DV, LIM_WOODWARD); else FUSED_L(MV, RT, DV, LIM_SUPERBEE); } while (0)
      if (mhd) {
        if (f32) { if (d == 0) FUSED(true, float, 0); else FUSED(true, float, 1); }
        else { if (d == 0) FUSED(true, double, 0); else FUSED(true, double, 1); }
      } else {
        if (f32) { if (d == 0) FUSED(false, float, 0); else FUSED(false, float, 1); }
        else { if (d == 0) FUSED(false, double, 0); else FUSED(false, double, 1); }
      }
#undef FUSED
#undef FUSED_L
      if ((rc = launch_check(g, "k_tvdlf_fused"))) return rc;
    }
    // the fused kernel has folded the Courant maximum and filled wout's ghost cells
    g->cur = outi;
    return 0;
  } else if (m->typefull == VACGPU_TVDMU) {
    // TVD-MUSCL: the Hancock predictor of TVDLF, then the face-parallel
    // corrector (central flux, Roe dissipation, Powell source for MHD)
    const int ghosts = (g->bc_uniform && g->use_fused) ? 1 : 0;
    if ((rc = launch_pred(g, m, d, s, win, ghosts))) return rc;
    if (!ghosts && (rc = bound(g, g->w1))) return rc;
    long long nbc = 0;
    if ((rc = launch_tvdmu_corr(g, m, d, s, win, wout, wl, ghosts, nbc))) return rc;
    if (courant) {
      // (tvdlimit2's check; the Courant words of the TVD sweeps, folded now or
      // by the next step's getdt)
      const int step_end = (g->fold_step_end && lastsweep) ? 1 : 0;
      if (step_end) g->step_end_done = true;
      if (g->defer_folds) {
        g->pend_mask |= 1 << d;
        g->pend_courantpar = m->courantpar;
        if (step_end) g->pend_step_end = true;
      } else {
        Timed t(g, KF_RED);
        hipLaunchKernelGGL(k_fin_courant_tvd_cs, dim3(1), dim3(384), 0, g->stream, g->sc, g->cslot, d,
                           m->courantpar, step_end);
        if ((rc = launch_check(g, "k_fin_courant_tvd_cs"))) return rc;
      }
    }
    (void)nbc;
    if (rt_wold && (rc = roundtrip_wold(g))) return rc;
    if (ghosts) {
      g->cur = outi;
      return 0;
    }
  } else if (m->typefull == VACGPU_TVDLF) {
    long long corr_blocks = nb;
    // with one boundary type per face the kernels write their outputs' ghost cells
    const int ghosts = (g->bc_uniform && g->use_fused) ? 1 : 0;
    const int lim = common_limiter(g, m);
    // FP32-bulk on 3D grids with copy-type faces: the predictor stores w1 as
    // its FP32 increment over win and the corrector adds it back to win in
    // FP64 (the limiter's base must be win: typelimited 'original'), which
    // moves 4 instead of 8 bytes per w1 variable each way
    const bool dw = g->cfg.precision == VACGPU_FP32BULK && g->cfg.ndim == 3 && ghosts &&
                    lim >= 0 && m->typelimited == VACGPU_LIM_ORIGINAL && !m->useprimitive;
    if ((rc = launch_pred(g, m, d, s, win, ghosts, dw))) return rc;
    if (!ghosts && (rc = bound(g, g->w1))) return rc;
    {
      // face-parallel corrector: x sweep one face per lane, y / z sweeps marching
      Timed t(g, KF_CORR);
      long long nbc = 0;
#define LIMSW(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else if (lim == LIM_SUPERBEE) X(LIM_SUPERBEE); else X(-1); } while (0)
#define LIMSW3(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else X(LIM_SUPERBEE); } while (0)
      if (d == 0) {
        const dim3 gr((g->geo.n[0] + CX_STRIDE - 1) / CX_STRIDE, g->grid.y, g->grid.z);
        nbc = (long long)gr.x * gr.y * gr.z;
#define CX(LV) hipLaunchKernelGGL((k_tvdlf_corrx<ND, MHD, R, LV>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, ghosts, wl)
#define CXW(LV) hipLaunchKernelGGL((k_tvdlf_corrx<ND, MHD, R, LV, false, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, ghosts, wl)
        if (dw)
          DISPATCH(g, {
            if constexpr (ND == 3 && std::is_same<R, float>::value) LIMSW3(CXW);
          });
        else if (m->useprimitive)
          DISPATCH(g, {
            if constexpr (std::is_same<R, double>::value)
              hipLaunchKernelGGL((k_tvdlf_corrx<ND, MHD, R, -1, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, ghosts, wl);
          });
        else
          DISPATCH(g, { LIMSW(CX); });
#undef CXW
#undef CX
      } else {
        const dim3 gr = g->grid_tm[d];
        nbc = (long long)gr.x * gr.y * gr.z;
#define CM(DV, LV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, DV, LV>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, g->seglen, ghosts, wl)
#define CM1(LV) CM(1, LV)
#define CM2(LV) CM((ND == 3 ? 2 : 1), LV)
#define CMP(DV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, DV, -1, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, g->seglen, ghosts, wl)
#define CMW1(LV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, 1, LV, false, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, g->seglen, ghosts, wl)
#define CMW2(LV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, 2, LV, false, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, wout, g->sc, g->part, g->pstride, g->seglen, ghosts, wl)
        if (dw)
          DISPATCH(g, {
            if constexpr (ND == 3 && std::is_same<R, float>::value) {
              if (d == 1) LIMSW3(CMW1); else LIMSW3(CMW2);
            }
          });
        else if (m->useprimitive)
          DISPATCH(g, {
            if constexpr (std::is_same<R, double>::value) {
              if (d == 1) CMP(1); else CMP((ND == 3 ? 2 : 1));
            }
          });
        else
          DISPATCH(g, { if (d == 1) LIMSW(CM1); else LIMSW(CM2); });
#undef CMW2
#undef CMW1
#undef CMP
#undef CM2
#undef CM1
#undef CM
      }
#undef LIMSW3
#undef LIMSW
      if ((rc = launch_check(g, "k_tvdlf_corr"))) return rc;
      corr_blocks = nbc;
    }
    if (courant) {
      Timed t(g, KF_RED);
      const int step_end = (g->fold_step_end && lastsweep) ? 1 : 0;
      if (step_end) g->step_end_done = true;
      hipLaunchKernelGGL(k_fin_courant_tvdlf, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part,
                         g->pstride, corr_blocks, d, m->courantpar, step_end);
      if ((rc = launch_check(g, "k_fin_courant_tvdlf"))) return rc;
    }
    if (rt_wold && (rc = roundtrip_wold(g))) return rc;
    if (ghosts) {
      g->cur = outi;
      return 0;
    }
  } else {
    // 2D grids with one boundary type per face fold the half source into the
    // first sweep and the first half of the Ryu step into the last one
    // (one limiter for all characteristic waves is fixed at compile time)
    const int lim = common_limiter(g, m);
    // entropy fixes, and useprimitive's dp for the MHD Roe solver, run in
    // the general kernels
    const bool ent = entropy_on(g, m) || (m->useprimitive && mhd) || g->nfany;
    const bool fold = g->cfg.ndim == 2 && g->bc_uniform && src && g->use_fused && lim >= 0 && !dirty &&
                      !ent;
    // (the Ryu half folds into the x sweep only: in the marching y sweep it
    // would cost occupancy)
    const int srcmode = (fold && firstsweep && !lastsweep) ? 1
                        : ((fold && lastsweep && !firstsweep && d == 0) ? 2 : 0);
    // with one boundary type per face the kernels write the ghost cells of
    // what they produce, and no separate boundary pass runs
    const int ghosts = (g->bc_uniform && g->use_fused) ? 1 : 0;
    const double* wct = win;
    if (firstsweep && src && srcmode != 1) {
      {
        Timed t(g, KF_RYU);
        if (g->cfg.ndim == 2)
          hipLaunchKernelGGL(k_tvd_halfsrc<2>, g->grid, g->block, 0, g->stream, g->geo, win, g->w1, g->sc);
        else
          hipLaunchKernelGGL(k_tvd_halfsrc<3>, g->grid, g->block, 0, g->stream, g->geo, win, g->w1, g->sc);
        if ((rc = launch_check(g, "k_tvd_halfsrc"))) return rc;
      }
      if ((rc = bound(g, g->w1))) return rc;
      wct = g->w1;
    }
    long long nbt;
    {
      Timed t(g, KF_TVD);
      const bool f32 = g->cfg.precision == VACGPU_FP32BULK;
#define LIMSW(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else if (lim == LIM_SUPERBEE) X(LIM_SUPERBEE); else X(-1); } while (0)
#define LIMSW3(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else X(LIM_SUPERBEE); } while (0)
      if (ent) {
        // entropy fixes: limiter types read at run time, source steps unfolded
        const dim3 gr = d == 0 ? g->grid_tx : g->grid_tm[d];
        nbt = (long long)gr.x * gr.y * gr.z;
#define EX(NDV, RT, MH) hipLaunchKernelGGL((k_tvd_x<NDV, RT, -1, 0, MH, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, wold, g->w1, ghosts)
#define EM(NDV, RT, DV, MH) hipLaunchKernelGGL((k_tvd_m<NDV, RT, DV, -1, 0, MH, true>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, g->seglen, wold, g->w1, ghosts)
#define EPH(X2, X3) do { if (g->cfg.ndim == 2) { if (f32) X2(float); else X2(double); } else { if (f32) X3(float); else X3(double); } } while (0)
#define EX2M(RT) EX(2, RT, true)
#define EX3M(RT) EX(3, RT, true)
#define EX2H(RT) EX(2, RT, false)
#define EX3H(RT) EX(3, RT, false)
#define EM2M(RT) EM(2, RT, 1, true)
#define EM3M(RT) do { if (d == 1) EM(3, RT, 1, true); else EM(3, RT, 2, true); } while (0)
#define EM2H(RT) EM(2, RT, 1, false)
#define EM3H(RT) do { if (d == 1) EM(3, RT, 1, false); else EM(3, RT, 2, false); } while (0)
        if (d == 0) { if (mhd) EPH(EX2M, EX3M); else EPH(EX2H, EX3H); }
        else { if (mhd) EPH(EM2M, EM3M); else EPH(EM2H, EM3H); }
#undef EX2M
#undef EX3M
#undef EX2H
#undef EX3H
#undef EM2M
#undef EM3M
#undef EM2H
#undef EM3H
#undef EPH
#undef EX
#undef EM
      } else if (!mhd) {
        // HD: the HD Roe solver, no source steps
        const dim3 gr = d == 0 ? g->grid_tx : g->grid_tm[d];
        nbt = (long long)gr.x * gr.y * gr.z;
#define HX(NDV, RT, LV) hipLaunchKernelGGL((k_tvd_x<NDV, RT, LV, 0, false>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, wold, g->w1, ghosts)
#define HM(NDV, RT, DV, LV) hipLaunchKernelGGL((k_tvd_m<NDV, RT, DV, LV, 0, false>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, g->seglen, wold, g->w1, ghosts)
#define HX2D(LV) HX(2, double, LV)
#define HX2F(LV) HX(2, float, LV)
#define HX3D(LV) HX(3, double, LV)
#define HX3F(LV) HX(3, float, LV)
#define HM21D(LV) HM(2, double, 1, LV)
#define HM21F(LV) HM(2, float, 1, LV)
#define HM31D(LV) HM(3, double, 1, LV)
#define HM31F(LV) HM(3, float, 1, LV)
#define HM32D(LV) HM(3, double, 2, LV)
#define HM32F(LV) HM(3, float, 2, LV)
        if (d == 0) {
          if (g->cfg.ndim == 2) { if (f32) LIMSW(HX2F); else LIMSW(HX2D); }
          else { if (f32) LIMSW(HX3F); else LIMSW(HX3D); }
        } else if (g->cfg.ndim == 2) { if (f32) LIMSW(HM21F); else LIMSW(HM21D); }
        else if (d == 1) { if (f32) LIMSW(HM31F); else LIMSW(HM31D); }
        else { if (f32) LIMSW(HM32F); else LIMSW(HM32D); }
#undef HX2D
#undef HX2F
#undef HX3D
#undef HX3F
#undef HM21D
#undef HM21F
#undef HM31D
#undef HM31F
#undef HM32D
#undef HM32F
#undef HX
#undef HM
      } else if (d == 0) {
        const dim3 gr = g->grid_tx;
        nbt = (long long)gr.x * gr.y * gr.z;
#define TVDX_L(NDV, RT, LV, SV) hipLaunchKernelGGL((k_tvd_x<NDV, RT, LV, SV>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, wold, g->w1, ghosts)
#define X2D(LV) TVDX_L(2, double, LV, 0)
#define X2F(LV) TVDX_L(2, float, LV, 0)
#define X3D(LV) TVDX_L(3, double, LV, 0)
#define X3F(LV) TVDX_L(3, float, LV, 0)
#define X2D1(LV) TVDX_L(2, double, LV, 1)
#define X2F1(LV) TVDX_L(2, float, LV, 1)
#define X2D2(LV) TVDX_L(2, double, LV, 2)
#define X2F2(LV) TVDX_L(2, float, LV, 2)
        if (g->cfg.ndim == 2) {
          if (srcmode == 1) { if (f32) LIMSW3(X2F1); else LIMSW3(X2D1); }
          else if (srcmode == 2) { if (f32) LIMSW3(X2F2); else LIMSW3(X2D2); }
          else { if (f32) LIMSW(X2F); else LIMSW(X2D); }
        } else { if (f32) LIMSW(X3F); else LIMSW(X3D); }
#undef X2D
#undef X2F
#undef X3D
#undef X3F
#undef X2D1
#undef X2F1
#undef X2D2
#undef X2F2
#undef TVDX_L
      } else {
        const dim3 gr = g->grid_tm[d];
        nbt = (long long)gr.x * gr.y * gr.z;
#define TVDM_L(NDV, RT, DV, LV, SV) hipLaunchKernelGGL((k_tvd_m<NDV, RT, DV, LV, SV>), gr, g->block, 0, g->stream, g->geo, s, g->bc, wct, win, wout, g->sc, g->cslot, g->seglen, wold, g->w1, ghosts)
#define M21D(LV) TVDM_L(2, double, 1, LV, 0)
#define M21F(LV) TVDM_L(2, float, 1, LV, 0)
#define M21D1(LV) TVDM_L(2, double, 1, LV, 1)
#define M21F1(LV) TVDM_L(2, float, 1, LV, 1)
#define M31D(LV) TVDM_L(3, double, 1, LV, 0)
#define M31F(LV) TVDM_L(3, float, 1, LV, 0)
#define M32D(LV) TVDM_L(3, double, 2, LV, 0)
#define M32F(LV) TVDM_L(3, float, 2, LV, 0)
        if (g->cfg.ndim == 2) {
          if (srcmode == 1) { if (f32) LIMSW3(M21F1); else LIMSW3(M21D1); }
          else { if (f32) LIMSW(M21F); else LIMSW(M21D); }
        }
        else if (d == 1) { if (f32) LIMSW(M31F); else LIMSW(M31D); }
        else { if (f32) LIMSW(M32F); else LIMSW(M32D); }
#undef M21D
#undef M21F
#undef M21D1
#undef M21F1
#undef M31D
#undef M31F
#undef M32D
#undef M32F
#undef TVDM_L
      }
#undef LIMSW
#undef LIMSW3
      if ((rc = launch_check(g, "k_tvd_sweep"))) return rc;
    }
    if (courant) {
      // (the Ryu step that may follow reads dt only)
      const int step_end = (g->fold_step_end && lastsweep) ? 1 : 0;
      if (step_end) g->step_end_done = true;
      if (g->defer_folds) {
        // the next step's k_getdt_lagged folds the words and ends this step
        g->pend_mask |= 1 << d;
        g->pend_courantpar = m->courantpar;
        if (step_end) g->pend_step_end = true;
      } else {
        Timed t(g, KF_RED);
        hipLaunchKernelGGL(k_fin_courant_tvd_cs, dim3(1), dim3(384), 0, g->stream, g->sc, g->cslot, d,
                           m->courantpar, step_end);
        if ((rc = launch_check(g, "k_fin_courant_tvd_cs"))) return rc;
      }
    }
    (void)nbt;
    const bool ryu_folded = srcmode == 2;
    if (lastsweep && src && ghosts && !ryu_folded && !firstsweep) {
      // both Ryu halves in one pass, into the buffer of this sweep's input
      // (dead now); that buffer becomes the step's output
      Timed t(g, KF_RYU);
      if (g->cfg.ndim == 2) {
        hipLaunchKernelGGL(k_ryu_ab_y, g->grid_ry, dim3(64, 4), 0, g->stream, g->geo, g->bc, wold,
                           wout, win, g->sc, g->ry_seglen);
      } else {
        hipLaunchKernelGGL(k_ryu_ab<3>, g->grid, g->block, 0, g->stream, g->geo, g->bc, wold, wout, win, g->sc);
      }
      return launch_check(g, "k_ryu_ab");
    }
    if (lastsweep && src) {
      if (!ryu_folded) {
        Timed t(g, KF_RYU);
        if (g->cfg.ndim == 2)
          hipLaunchKernelGGL(k_ryu_a<2>, g->grid, g->block, 0, g->stream, g->geo, g->bc, wold, wout, g->w1, g->sc, ghosts);
        else
          hipLaunchKernelGGL(k_ryu_a<3>, g->grid, g->block, 0, g->stream, g->geo, g->bc, wold, wout, g->w1, g->sc, ghosts);
        if ((rc = launch_check(g, "k_ryu_a"))) return rc;
      }
      if (!ghosts && (rc = bound(g, g->w1))) return rc;
      {
        Timed t(g, KF_RYU);
        if (g->cfg.ndim == 2)
          hipLaunchKernelGGL(k_ryu_b<2>, g->grid, g->block, 0, g->stream, g->geo, g->bc, g->w1, wout, g->sc, ghosts);
        else
          hipLaunchKernelGGL(k_ryu_b<3>, g->grid, g->block, 0, g->stream, g->geo, g->bc, g->w1, wout, g->sc, ghosts);
        if ((rc = launch_check(g, "k_ryu_b"))) return rc;
      }
    }
    // ghost cells of wout: written by an x sweep, or by the Ryu step on the
    // last sweep; a y / z sweep that does not end the step needs the pass
    if (ghosts && (d == 0 || (lastsweep && src))) {
      g->cur = outi;
      return 0;
    }
  }
  if ((rc = bound(g, wout))) return rc;
  g->cur = outi;
  return 0;
}

// unsplit sweep (advect over idim = 1..ndim, src/vac.t:325-329) of TVDLF
// and TVD-MUSCL: the predictor adds every direction's fluxes in one pass per
// cell, the corrector runs direction by direction; 'original' limiting reads
// as 'previous' (src/vacio.t:363-364), the step's start
int sweep_unsplit_device(vacgpu_grid* g, const vacgpu_method* m) {
  g->old = g->old_pending >= 0 ? g->old_pending : g->cur;
  g->old_pending = -1;
  g->mid_step = false;
  g->ghosts_dirty = false;  // these kernels read every ghost cell from memory
  const int outi = other_buffer(g);
  const double* win = g->buf[g->cur];
  double* wout = g->buf[outi];
  const double* wold = g->buf[g->old];
  const Sel s = make_sel(g, m);
  // 'original' reads as 'previous' here
  const int tl = m->typelimited == VACGPU_LIM_PREDICTOR ? VACGPU_LIM_PREDICTOR : VACGPU_LIM_PREVIOUS;
  const int ghosts = (g->bc_uniform && g->use_fused) ? 1 : 0;
  const double* wl = m->typelimited == VACGPU_LIM_PREDICTOR ? g->w1 : wold;
  int rc = 0;
  {
    Timed t(g, KF_PRED);
    if (m->useprimitive)
      DISPATCH(g, {
        if constexpr (std::is_same<R, double>::value)
          hipLaunchKernelGGL((k_tvdlf_pred_u<ND, MHD, R, true>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts);
      });
    else
      DISPATCH(g, {
        hipLaunchKernelGGL((k_tvdlf_pred_u<ND, MHD, R>), g->grid, g->block, 0, g->stream, g->geo, s, g->bc, win, g->w1, g->sc, ghosts);
      });
    if ((rc = launch_check(g, "k_tvdlf_pred_u"))) return rc;
  }
  if (!ghosts && (rc = bound(g, g->w1))) return rc;
  // corrector: direction by direction, each launch reading the previous
  // one's result as its base (its own cell only) and writing another buffer:
  // 2D win -> tmp -> wout, 3D win -> wout -> tmp -> wout; the limiter's base
  // wl and w1 stay put.  The source comes with the last direction.
  const int nd = g->cfg.ndim;
  // the chain's intermediate must be neither the input, the output nor wold
  // (the limiter's base, read by every launch of the chain)
  double* tmp = nullptr;
  for (int i = 0; i < 3; ++i)
    if (i != g->cur && i != outi && i != g->old) tmp = g->buf[i];
  if (!tmp) {
    if (!g->xbuf)
      HIPCHK(g, hipMalloc(&g->xbuf, sizeof(double) * (size_t)g->geo.ps * g->nw));
    tmp = g->xbuf;
  }
  const int lim = common_limiter(g, m);
  for (int d = 0; d < nd; ++d) {
    const double* base = d == 0 ? win : ((nd - d) % 2 ? tmp : wout);
    double* dst = (nd - 1 - d) % 2 ? tmp : wout;
    Sel sd = s;
    if (d < nd - 1) sd.sourceunsplit = 0;
    else sd.srcmul = nd;
    if (m->useprimitive) set_prim_roundtrips(sd, tl, d, nd);
    const bool rt_wold = m->useprimitive && tl == VACGPU_LIM_PREVIOUS;
    const int gh = d == nd - 1 ? ghosts : 0;
    long long nbc = 0;
    if (m->typefull == VACGPU_TVDMU) {
      // TVD-MUSCL: each direction's central flux and Roe dissipation
      // (src/vactvdlf.t:137-283), its Courant check that of tvdlimit2
      if ((rc = launch_tvdmu_corr(g, m, d, sd, base, dst, wl, gh, nbc))) return rc;
      if (m->courantpar > 0.0) {
        Timed t(g, KF_RED);
        hipLaunchKernelGGL(k_fin_courant_tvd_cs, dim3(1), dim3(384), 0, g->stream, g->sc, g->cslot, d,
                           m->courantpar, 0);
        if ((rc = launch_check(g, "k_fin_courant_tvd_cs"))) return rc;
      }
      if (rt_wold && (rc = roundtrip_wold(g))) return rc;
      continue;
    }
    {
      Timed t(g, KF_CORR);
#define LIMSW(X) do { if (lim == LIM_MINMOD) X(LIM_MINMOD); else if (lim == LIM_WOODWARD) X(LIM_WOODWARD); else if (lim == LIM_SUPERBEE) X(LIM_SUPERBEE); else X(-1); } while (0)
      if (d == 0) {
        const dim3 gr((g->geo.n[0] + CX_STRIDE - 1) / CX_STRIDE, g->grid.y, g->grid.z);
        nbc = (long long)gr.x * gr.y * gr.z;
#define CX(LV) hipLaunchKernelGGL((k_tvdlf_corrx<ND, MHD, R, LV>), gr, g->block, 0, g->stream, g->geo, sd, g->bc, base, g->w1, dst, g->sc, g->part, g->pstride, gh, wl)
        if (m->useprimitive)
          DISPATCH(g, {
            if constexpr (std::is_same<R, double>::value)
              hipLaunchKernelGGL((k_tvdlf_corrx<ND, MHD, R, -1, true>), gr, g->block, 0, g->stream, g->geo, sd, g->bc, base, g->w1, dst, g->sc, g->part, g->pstride, gh, wl);
          });
        else
          DISPATCH(g, { LIMSW(CX); });
#undef CX
      } else {
        const dim3 gr = g->grid_tm[d];
        nbc = (long long)gr.x * gr.y * gr.z;
#define CM(DV, LV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, DV, LV>), gr, g->block, 0, g->stream, g->geo, sd, g->bc, base, g->w1, dst, g->sc, g->part, g->pstride, g->seglen, gh, wl)
#define CM1(LV) CM(1, LV)
#define CM2(LV) CM((ND == 3 ? 2 : 1), LV)
#define CMP(DV) hipLaunchKernelGGL((k_tvdlf_corrm<ND, MHD, R, DV, -1, true>), gr, g->block, 0, g->stream, g->geo, sd, g->bc, base, g->w1, dst, g->sc, g->part, g->pstride, g->seglen, gh, wl)
        if (m->useprimitive)
          DISPATCH(g, {
            if constexpr (std::is_same<R, double>::value) {
              if (d == 1) CMP(1); else CMP((ND == 3 ? 2 : 1));
            }
          });
        else
          DISPATCH(g, { if (d == 1) LIMSW(CM1); else LIMSW(CM2); });
#undef CMP
#undef CM2
#undef CM1
#undef CM
      }
#undef LIMSW
      if ((rc = launch_check(g, "k_tvdlf_corr"))) return rc;
    }
    if (m->courantpar > 0.0) {
      Timed t(g, KF_RED);
      hipLaunchKernelGGL(k_fin_courant_tvdlf, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part,
                         g->pstride, nbc, d, m->courantpar, 0);
      if ((rc = launch_check(g, "k_fin_courant_tvdlf"))) return rc;
    }
    if (rt_wold && (rc = roundtrip_wold(g))) return rc;
  }
  if (!ghosts && (rc = bound(g, wout))) return rc;
  g->cur = outi;
  return 0;
}

// boundscalar's face types from the normal B component's (typeBscalar,
// src/vacproc.projectb.t:82-101): periodic -> periodic, symm -> asymm,
// asymm -> symm, fixed -> grad0, cont -> nul (the default case)
SBC make_sbc(const vacgpu_grid* g) {
  SBC sb;
  const int nd = g->cfg.ndim;
  for (int f = 0; f < 6; ++f) {
    const int tb = f < 2 * nd ? g->cfg.typeB[f][2 + nd + f / 2] : VACGPU_CONT;
    sb.type[f] = (unsigned char)(tb == VACGPU_PERIODIC ? SB_PERIODIC
                                 : (tb == VACGPU_SYMM || tb == VACGPU_SYMM0) ? SB_ASYMM
                                 : tb == VACGPU_ASYMM ? SB_SYMM
                                 : (tb == VACGPU_FIXED || tb == VACGPU_FIXED1) ? SB_GRAD0 : SB_NUL);
  }
  return sb;
}

int ensure_proj(vacgpu_grid* g) {
  auto& P = g->pr;
  if (P.cg) return 0;
  const size_t plane = sizeof(double) * (size_t)g->geo.ps;
  double** planes[5] = {&P.rhs, &P.phi, &P.z, &P.z2, &P.az};
  for (auto q : planes) {
    HIPCHK(g, hipMalloc(q, plane));
    HIPCHK(g, hipMemset(*q, 0, plane));
  }
  for (int d = 0; d < g->cfg.ndim; ++d) {
    const int n = g->geo.ng[d];
    HIPCHK(g, hipMalloc(&P.rdx[d], sizeof(double) * n));
    hipLaunchKernelGGL(k_recip, dim3((n + 255) / 256), dim3(256), 0, g->stream, g->geo.dx[d], P.rdx[d], n);
    int rc;
    if ((rc = launch_check(g, "k_recip"))) return rc;
  }
  HIPCHK(g, hipMalloc(&P.cg, sizeof(Cg)));
  HIPCHK(g, hipMemset(P.cg, 0, sizeof(Cg)));
  HIPCHK(g, hipHostMalloc(&P.cg_host, sizeof(Cg), hipHostMallocDefault));
  return 0;
}

// projectb (src/vacproc.projectb.t:4-157) on the current state.  keep_old:
// the state before the projection becomes the next first sweep's wold.
int projectb_device(vacgpu_grid* g, const double* procpar, bool keep_old, int* info_out,
                    int* iter_out, double* resid_out) {
  if (g->cfg.physics != VACGPU_MHD) return fail(g, VACGPU_E_ARG, "the projection of B needs MHD");
  int rc;
  if ((rc = ensure_proj(g))) return rc;
  auto& P = g->pr;
  const int nd = g->cfg.ndim;
  const long long nb = (long long)g->grid.x * g->grid.y * g->grid.z;
  if (keep_old) {
    const int x = other_buffer(g);
    const size_t bytes = sizeof(double) * (size_t)g->geo.ps * g->nw;
    HIPCHK(g, hipMemcpyAsync(g->buf[x], g->buf[g->cur], bytes, hipMemcpyDeviceToDevice, g->stream));
    g->old_pending = x;
  }
  double* w = g->buf[g->cur];
  const SBC sb = make_sbc(g);
  const RDX rx = {{P.rdx[0], P.rdx[1], P.rdx[2]}};
  Timed t(g, KF_PROJ);
  if (nd == 2)
    hipLaunchKernelGGL(k_pr_divb<2>, g->grid, g->block, 0, g->stream, g->geo, rx, w, P.rhs, P.z, P.phi, g->part, g->pstride);
  else
    hipLaunchKernelGGL(k_pr_divb<3>, g->grid, g->block, 0, g->stream, g->geo, rx, w, P.rhs, P.z, P.phi, g->part, g->pstride);
  hipLaunchKernelGGL(k_pr_fold, dim3(1), dim3(1024), 0, g->stream, P.cg, g->part, g->pstride, nb);
  if ((rc = launch_check(g, "k_pr_divb"))) return rc;
  if (P.divbmax < 1.0e-99) {
    // the first call fixes the allowed maximum, the stopping rule and the
    // matvec limit from the initial div B (:29-58)
    HIPCHK(g, hipMemcpyAsync(P.cg_host, P.cg, sizeof(Cg), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(g, hipStreamSynchronize(g->stream));
    const double coeff = procpar[0] == -1.0 ? 1.0 / 10 : procpar[0];
    const double cnst = procpar[1] == -1.0 ? 0.0 : procpar[1];
    double divbmax = P.cg_host->mx;
    int typestop;
    if (coeff < 0.0 && coeff > -1.0) {
      divbmax = -coeff;
      typestop = CG_STOP_REL;
    } else {
      divbmax = divbmax * std::max(0.0, coeff) + std::max(0.0, cnst);
      typestop = CG_STOP_MAX;
      if (divbmax < 1.0e-99)
        return fail(g, VACGPU_E_ARG, "projectb: too small value for divbmax (procpar)");
    }
    int maxit = 1000;
    if (cnst < -1.0) maxit = (int)std::lround(-cnst);
    if (coeff < -1.0) maxit = (int)std::lround(-coeff);
    P.divbmax = divbmax;
    hipLaunchKernelGGL(k_cg_set, dim3(1), dim3(1), 0, g->stream, P.cg, divbmax, typestop, maxit);
    if ((rc = launch_check(g, "k_cg_set"))) return rc;
  }
  hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, g->stream, P.cg);
  if ((rc = launch_check(g, "k_cg_init"))) return rc;
  // CG steps in chunks; a step after the stopping test returns at once
  double* zb[2] = {P.z, P.z2};
  int par = 0;
  bool first = true;
  // the first chunk as long as the previous solve (successive solves of a
  // run take similar numbers of steps), then 2, 4, 8, 16 steps between polls
  int chunk = std::min(64, std::max(2, P.last_iters));
  int next = 2;
  long long queued = 0;
  for (;;) {
    for (int k = 0; k < chunk; ++k) {
      if (nd == 2)
        hipLaunchKernelGGL(k_cg_matvec<2>, g->grid, g->block, 0, g->stream, g->geo, sb, rx, P.rhs, zb[par], zb[1 - par], P.az, g->part, g->pstride, P.cg, first ? 1 : 0);
      else
        hipLaunchKernelGGL(k_cg_matvec<3>, g->grid, g->block, 0, g->stream, g->geo, sb, rx, P.rhs, zb[par], zb[1 - par], P.az, g->part, g->pstride, P.cg, first ? 1 : 0);
      hipLaunchKernelGGL(k_cg_alf, dim3(1), dim3(1024), 0, g->stream, P.cg, g->part, g->pstride, nb);
      if (nd == 2)
        hipLaunchKernelGGL(k_cg_update<2>, g->grid, g->block, 0, g->stream, g->geo, P.phi, P.rhs, zb[1 - par], P.az, g->part, g->pstride, P.cg);
      else
        hipLaunchKernelGGL(k_cg_update<3>, g->grid, g->block, 0, g->stream, g->geo, P.phi, P.rhs, zb[1 - par], P.az, g->part, g->pstride, P.cg);
      hipLaunchKernelGGL(k_cg_res, dim3(1), dim3(1024), 0, g->stream, P.cg, g->part, g->pstride, nb);
      par = 1 - par;
      first = false;
    }
    if ((rc = launch_check(g, "k_cg_step"))) return rc;
    queued += chunk;
    HIPCHK(g, hipMemcpyAsync(P.cg_host, P.cg, sizeof(Cg), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(g, hipStreamSynchronize(g->stream));
    if (P.cg_host->done) break;
    if (queued > (long long)P.cg_host->maxit + 2)
      return fail(g, VACGPU_E_STATE, "projectb: the CG iteration did not stop");
    chunk = next;
    next = next < 16 ? 2 * next : 16;
  }
  P.last_iters = P.cg_host->itr;
  hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(1), 0, g->stream, P.cg, g->sc);
  if ((rc = launch_check(g, "k_cg_final"))) return rc;
  {
    const dim3 gr((g->geo.n[0] + 2 + 63) / 64, (g->geo.n[1] + 2 + 3) / 4, nd == 3 ? g->geo.n[2] + 2 : 1);
    const int eb = (int)std::lround(std::fabs(procpar[2]));
    if (nd == 2)
      hipLaunchKernelGGL(k_pr_apply<2>, gr, dim3(64, 4), 0, g->stream, g->geo, sb, rx, P.phi, w, P.cg, eb, g->cfg.eqpar[0]);
    else
      hipLaunchKernelGGL(k_pr_apply<3>, gr, dim3(64, 4), 0, g->stream, g->geo, sb, rx, P.phi, w, P.cg, eb, g->cfg.eqpar[0]);
    if ((rc = launch_check(g, "k_pr_apply"))) return rc;
  }
  // B's ghost cells when procpar(divbbound_) > 0 (:143-144)
  if (procpar[2] > 0.0 && (rc = bound(g, w, 2 + nd, 2 + 2 * nd))) return rc;
  // what the solve returned (the sign rule of :570-573, as k_cg_final applies it)
  const Cg& c = *P.cg_host;
  int info = c.info;
  if ((c.typestop == CG_STOP_REL && c.res > 1.0) || (c.typestop != CG_STOP_REL && c.res > c.res0))
    info = -info;
  if (info != 3 && info >= 0) g->ghosts_dirty = true;
  if (info_out) *info_out = info;
  if (iter_out) *iter_out = c.itr;
  if (resid_out) *resid_out = c.res;
  return 0;
}

int advance_device(vacgpu_grid* g, const vacgpu_method* m, int it) {
  const int nd = g->cfg.ndim;
  const int np = g->pr.nproc1;
  const int di = it - g->pr.itmin;
  // wold is the state from before the step's projection (VAC copies it
  // before advance, src/vac.t:85): TVD's Ryu step reads it, and so does a
  // corrector that limits the step's start ('previous', and 'original' of
  // unsplit sweeps, src/vacio.t:363-364)
  const bool needs_old = (m->typefull == VACGPU_TVD && g->cfg.physics == VACGPU_MHD &&
                          m->divbfix && m->sourceunsplit) ||
                         m->typelimited == VACGPU_LIM_PREVIOUS ||
                         (m->unsplit && m->typelimited != VACGPU_LIM_PREDICTOR);
  auto sweep = [&](int d, bool first, bool last) {
    // process(1) before the sweep when due (advect, src/vac.t:392-399)
    if (np != 0 && (np < 0 || (first && di == (di / np) * np))) {
      int rc = projectb_device(g, g->pr.procpar, first && needs_old, nullptr, nullptr, nullptr);
      if (rc) return rc;
    }
    return d < 0 ? sweep_unsplit_device(g, m) : sweep_device(g, m, d, first, last);
  };
  int rc;
  if (m->unsplit) return sweep(-1, true, true);  // all directions at once (src/vac.t:325-329)
  if ((it / 2) * 2 == it || m->dimsplit_xy) {
    for (int d = 0; d < nd; ++d)
      if ((rc = sweep(d, d == 0, d == nd - 1))) return rc;
  } else {
    for (int d = nd - 1; d >= 0; --d)
      if ((rc = sweep(d, d == nd - 1, d == 0))) return rc;
  }
  return 0;
}

int read_scalars(vacgpu_grid* g) {
  HIPCHK(g, hipMemcpyAsync(g->sc_host, g->sc, sizeof(Scal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return 0;
}

int write_scalars(vacgpu_grid* g) {
  HIPCHK(g, hipMemcpyAsync(g->sc, g->sc_host, sizeof(Scal), hipMemcpyHostToDevice, g->stream));
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return 0;
}

}  // namespace

// pdiv, the quotient of the sweep kernels, beside the compiler's division
__global__ void k_check_quotients(const double* __restrict__ a, const double* __restrict__ b,
                                  long long n, double* __restrict__ qk, double* __restrict__ qi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qk[i] = pdiv(a[i], b[i]);
  qi[i] = a[i] / b[i];
}

// psqrt, the square root of the sweep kernels, beside the compiler's sqrt
__global__ void k_check_roots(const double* __restrict__ x, long long n, double* __restrict__ rk,
                              double* __restrict__ ri) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rk[i] = psqrt(x[i]);
  ri[i] = sqrt(x[i]);
}

extern "C" {

int vacgpu_supported(const vacgpu_config* cfg, const vacgpu_method* m) {
  std::string why;
  int rc = check_config(cfg, why);
  if (rc) return rc;
  if (m) return check_method(cfg, m, why);
  return 0;
}

int vacgpu_create(const vacgpu_config* cfg, const double* dx0, const double* dx1,
                  const double* dx2, vacgpu_grid** out) {
  if (!out) return VACGPU_E_ARG;
  *out = nullptr;
  std::string why;
  if (check_config(cfg, why)) {
    fprintf(stderr, "vacgpu_create: %s\n", why.c_str());
    return VACGPU_E_ARG;
  }
  const double* dxs[3] = {dx0, dx1, dx2};
  for (int d = 0; d < cfg->ndim; ++d)
    if (!dxs[d]) { fprintf(stderr, "vacgpu_create: missing dx array\n"); return VACGPU_E_ARG; }
  vacgpu_grid* g = new vacgpu_grid();
  g->cfg = *cfg;
  const int nd = cfg->ndim;
  g->nw = cfg->physics == VACGPU_MHD ? 2 + 2 * nd : 2 + nd;
  Geo& geo = g->geo;
  geo.nd = nd;
  for (int d = 0; d < 3; ++d) {
    geo.n[d] = d < nd ? cfg->nx[d] : 1;
    geo.ng[d] = d < nd ? cfg->nx[d] + 4 : 1;
  }
  const long long px = ((geo.ng[0] + 15) / 16) * 16 + 16;  // room for the alignment shift
  geo.s[0] = 1;
  geo.s[1] = px;
  geo.s[2] = (nd == 3) ? px * geo.ng[1] : 0;
  geo.off = 14;  // (off + 2) % 16 == 0: mesh rows start on a 128-byte line
  const long long rows = (long long)geo.ng[1] * (nd == 3 ? geo.ng[2] : 1);
  geo.ps = ((geo.off + px * rows + 31) / 32) * 32;
  for (int f = 0; f < 6; ++f)
    for (int iw = 0; iw < 8; ++iw)
    {
      const int tb = (f < 2 * nd && iw < g->nw) ? cfg->typeB[f][iw] : VACGPU_CONT;
      g->bc.t[f][iw] = (unsigned char)(tb == VACGPU_PERIODIC);
      g->bc.k[f][iw] = (unsigned char)((tb == VACGPU_SYMM || tb == VACGPU_SYMM0) ? BK_SYMM
                                       : tb == VACGPU_ASYMM ? BK_ASYMM
                                       : tb == VACGPU_FIXED ? BK_FIXED
                                       : tb == VACGPU_CONT1 ? BK_CONT1
                                       : tb == VACGPU_FIXED1 ? BK_FIXED1
                                       : tb == VACGPU_GRAD1 ? BK_GRAD1 : BK_COPY);
      if (tb == VACGPU_FIXED || tb == VACGPU_FIXED1 || tb == VACGPU_GRAD1) g->any_fixed = true;
      if (tb == VACGPU_SYMM0) {
        g->nf[f] |= (unsigned char)(1u << iw);
        g->nfany = true;
      }
    }
  // ghost-cell enumeration: slab d has ghost coordinate d, mesh coordinates
  // below d and full ixG coordinates above d
  long long tot = 0;
  for (int d = 0; d < nd; ++d) {
    long long v = 4;
    for (int k = 0; k < nd; ++k)
      if (k < d) v *= geo.n[k];
      else if (k > d) v *= geo.ng[k];
    g->slab[d] = v;
    tot += v;
  }
  g->nbound = tot;
  g->block = dim3(BX, BY, 1);
  g->grid = dim3((geo.n[0] + BX - 1) / BX, (geo.n[1] + BY - 1) / BY, nd == 3 ? geo.n[2] : 1);
  g->grid_tx = dim3((geo.n[0] + TX_STRIDE - 1) / TX_STRIDE, g->grid.y, g->grid.z);
  {
    // segment length of the marching sweeps: enough threads to fill the chip
    long long cells = (long long)geo.n[0] * geo.n[1] * (nd == 3 ? geo.n[2] : 1);
    long long sl = cells / 262144;
    g->seglen = (int)(sl < 4 ? 4 : (sl > 32 ? 32 : sl));
  }
  for (int d = 1; d < nd; ++d) {
    const int nseg = (geo.n[d] + g->seglen - 1) / g->seglen;
    const int other = (d == 1) ? 2 : 1;
    g->grid_tm[d] = dim3(g->grid.x, (nseg + BY - 1) / BY, nd == 3 ? geo.n[other] : 1);
  }
  // fused TVDLF: one boundary type per face for all variables
  g->bc_uniform = true;
  for (int d = 0; d < nd; ++d) g->bc_uniform &= geo.n[d] >= 4;
  for (int f = 0; f < 2 * nd; ++f)
    for (int iw = 0; iw < g->nw; ++iw) {
      if (g->bc.t[f][iw] != g->bc.t[f][0]) g->bc_uniform = false;
      // kernels that write ghost cells as copies of mesh cells, or read them
      // through that mapping, need copy-type faces
      if (g->bc.k[f][iw] != BK_COPY) g->bc_uniform = false;
    }
  g->fused_ok = nd == 2 && g->bc_uniform && (unsigned long long)geo.ps * 8ULL < (1ULL << 32);
  set_fused_shape(g, 0);
  if (nd == 2) set_ryu_shape(g);
  auto count = [](dim3 v) { return (long long)v.x * v.y * v.z; };
  g->pstride = count(g->grid);
  {
    // partials of the fused sweeps at the shortest segment length
    const long long strips = (geo.n[0] + FS_STRIDE - 1) / FS_STRIDE;
    const long long segs = (geo.n[1] + 3) / 4;
    const long long nbmax = strips * ((segs + 3) / 4);
    if (nbmax > g->pstride) g->pstride = nbmax;
  }
  if (count(g->grid_tx) > g->pstride) g->pstride = count(g->grid_tx);
  for (int d = 1; d < nd; ++d)
    if (count(g->grid_tm[d]) > g->pstride) g->pstride = count(g->grid_tm[d]);
  const size_t bytes = sizeof(double) * (size_t)geo.ps * g->nw;
  hipError_t e = hipSuccess;
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipMalloc(&g->buf[i], bytes);
  if (e == hipSuccess) e = hipMalloc(&g->w1, bytes);
  size_t dxn = 0;
  for (int d = 0; d < nd; ++d) dxn += geo.ng[d];
  if (e == hipSuccess) e = hipMalloc(&g->dxdev, sizeof(double) * dxn);
  if (e == hipSuccess) e = hipMalloc(&g->sc, sizeof(Scal));
  if (e == hipSuccess) e = hipHostMalloc(&g->sc_host, sizeof(Scal), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&g->part, sizeof(double) * NSLOT * g->pstride);
  if (e == hipSuccess) e = hipMalloc(&g->fold, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(g->fold, 0, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&g->cslot, 6 * CS_N * CS_PAD * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(g->cslot, 0, 6 * CS_N * CS_PAD * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fprintf(stderr, "vacgpu_create: %s\n", hipGetErrorString(e));
    vacgpu_destroy(g);
    return e == hipErrorOutOfMemory ? VACGPU_E_NOMEM : VACGPU_E_HIP;
  }
  g->own_stream = true;
  size_t o = 0;
  for (int d = 0; d < nd; ++d) {
    e = hipMemcpy(g->dxdev + o, dxs[d], sizeof(double) * geo.ng[d], hipMemcpyHostToDevice);
    geo.dx[d] = g->dxdev + o;
    o += geo.ng[d];
  }
  for (int d = nd; d < 3; ++d) geo.dx[d] = g->dxdev;
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipMemset(g->buf[i], 0, bytes);
  if (e == hipSuccess) e = hipMemset(g->w1, 0, bytes);
  if (e != hipSuccess) {
    fprintf(stderr, "vacgpu_create: %s\n", hipGetErrorString(e));
    vacgpu_destroy(g);
    return VACGPU_E_HIP;
  }
  Scal s0{};
  s0.dt = 0.0;
  s0.t = 0.0;
  s0.smallp = -1.0;
  for (int d = 0; d < 3; ++d) s0.dtcourant[d] = kBig;
  s0.tmax = kBig;
  *g->sc_host = s0;
  if (write_scalars(g)) {
    fprintf(stderr, "vacgpu_create: %s\n", g->err.c_str());
    vacgpu_destroy(g);
    return VACGPU_E_HIP;
  }
  *out = g;
  return 0;
}

int vacgpu_set_stream(vacgpu_grid* g, void* s) {
  if (!g) return VACGPU_E_ARG;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  if (g->own_stream) HIPCHK(g, hipStreamDestroy(g->stream));
  g->stream = (hipStream_t)s;
  g->own_stream = false;
  return 0;
}

static int copy_state(vacgpu_grid* g, double* host, const int* ld, bool put) {
  const Geo& geo = g->geo;
  const int nd = g->cfg.ndim;
  int l[3] = {geo.ng[0], geo.ng[1], nd == 3 ? geo.ng[2] : 1};
  if (ld)
    for (int d = 0; d < nd; ++d) {
      if (ld[d] < geo.ng[d]) return fail(g, VACGPU_E_ARG, "leading dimension smaller than ixG");
      l[d] = ld[d];
    }
  const size_t hpitch = sizeof(double) * l[0];
  const size_t dpitch = sizeof(double) * geo.s[1];
  const size_t width = sizeof(double) * geo.ng[0];
  const int nz = nd == 3 ? geo.ng[2] : 1;
  double* dev = g->buf[g->cur];
  // through a pinned staging copy of the caller's array: one host copy, then
  // pitched DMA per plane
  const size_t plane = (size_t)l[0] * l[1];
  const size_t hbytes = sizeof(double) * plane * (nd == 3 ? l[2] : 1) * g->nw;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  if (hbytes > g->stage_bytes) {
    if (g->stage) HIPCHK(g, hipHostFree(g->stage));
    g->stage = nullptr;
    g->stage_bytes = 0;
    HIPCHK(g, hipHostMalloc(&g->stage, hbytes, hipHostMallocDefault));
    g->stage_bytes = hbytes;
  }
  if (put) memcpy(g->stage, host, hbytes);
  for (int iw = 0; iw < g->nw; ++iw)
    for (int z = 0; z < nz; ++z) {
      const size_t ho = (size_t)iw * plane * (nd == 3 ? l[2] : 1) + (size_t)z * plane;
      double* h = g->stage + ho;
      double* dv = dev + iw * geo.ps + geo.off + z * geo.s[2];
      if (put)
        HIPCHK(g, hipMemcpy2DAsync(dv, dpitch, h, hpitch, width, geo.ng[1], hipMemcpyHostToDevice, g->stream));
      else
        HIPCHK(g, hipMemcpy2DAsync(h, hpitch, dv, dpitch, width, geo.ng[1], hipMemcpyDeviceToHost, g->stream));
    }
  HIPCHK(g, hipStreamSynchronize(g->stream));
  if (!put)  // only the ixG part of each row: the caller's padding stays as it was
    for (int iw = 0; iw < g->nw; ++iw)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < geo.ng[1]; ++y) {
          const size_t ho = (size_t)iw * plane * (nd == 3 ? l[2] : 1) + (size_t)z * plane + (size_t)y * l[0];
          memcpy(host + ho, g->stage + ho, width);
        }
  return 0;
}

int vacgpu_put_state(vacgpu_grid* g, const double* w, const int* ld) {
  if (!g || !w) return VACGPU_E_ARG;
  int rc = copy_state(g, const_cast<double*>(w), ld, true);
  if (rc) return rc;
  g->old = g->cur;
  g->have_state = true;
  // VAC fills the ghost cells once after reading the initial state (src/vac.t:54)
  if ((rc = bound(g, g->buf[g->cur]))) return rc;
  if (g->any_fixed && !g->fixw) {
    // that first fill copies 'fixed' faces like 'cont'; their ghost values
    // stay what it left (src/vacgrid.t:372-404)
    const size_t bytes = sizeof(double) * (size_t)g->geo.ps * g->nw;
    HIPCHK(g, hipMalloc(&g->fixw, bytes));
    HIPCHK(g, hipMemcpyAsync(g->fixw, g->buf[g->cur], bytes, hipMemcpyDeviceToDevice, g->stream));
  }
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return 0;
}

int vacgpu_put_state_ghosts(vacgpu_grid* g, const double* w, const int* ld) {
  if (!g || !w) return VACGPU_E_ARG;
  int rc = copy_state(g, const_cast<double*>(w), ld, true);
  if (rc) return rc;
  g->old = g->cur;
  g->have_state = true;
  // the caller's ghost cells need not be copies of mesh cells (a projection
  // leaves its own there): the next sweep reads them from memory
  g->ghosts_dirty = true;
  if (g->any_fixed && !g->fixw) {
    // the values of 'fixed' faces are those of the first state put here
    const size_t bytes = sizeof(double) * (size_t)g->geo.ps * g->nw;
    HIPCHK(g, hipMalloc(&g->fixw, bytes));
    HIPCHK(g, hipMemcpyAsync(g->fixw, g->buf[g->cur], bytes, hipMemcpyDeviceToDevice, g->stream));
  }
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return 0;
}

int vacgpu_get_state(vacgpu_grid* g, double* w, const int* ld) {
  if (!g || !w) return VACGPU_E_ARG;
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  return copy_state(g, w, ld, false);
}

int vacgpu_getdt_courant(vacgpu_grid* g, const vacgpu_method* m, double* dtcourant, double* smallp,
                         double* dt) {
  if (!g) return VACGPU_E_ARG;
  std::string why;
  if (check_method(&g->cfg, m, why)) return fail(g, VACGPU_E_ARG, why);
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  if (m->courantpar <= 0.0) return fail(g, VACGPU_E_ARG, "getdt_courant needs courantpar > 0");
  int rc;
  if ((rc = read_scalars(g))) return rc;
  if (dtcourant)
    for (int d = 0; d < g->cfg.ndim; ++d) g->sc_host->dtcourant[d] = dtcourant[d];
  if (smallp) g->sc_host->smallp = *smallp;
  // getdt_courant alone: the main loop's dtcantgrow / slowsteps are the host's
  g->sc_host->dtcantgrow = 0;
  g->sc_host->slowsteps = 0;
  if ((rc = write_scalars(g))) return rc;
  if ((rc = getdt_device(g, m, -1.0, 0))) return rc;
  if ((rc = read_scalars(g))) return rc;
  if (dtcourant)
    for (int d = 0; d < g->cfg.ndim; ++d) dtcourant[d] = g->sc_host->dtcourant[d];
  if (smallp) *smallp = g->sc_host->smallp;
  if (dt) *dt = g->sc_host->dt;
  return 0;
}

int vacgpu_advect(vacgpu_grid* g, const vacgpu_method* m, int idimmin, int idimmax, double dt,
                  int firstsweep, int lastsweep, double* dtcourant, int* nerror, double* smallp) {
  if (!g) return VACGPU_E_ARG;
  std::string why;
  if (check_method(&g->cfg, m, why)) return fail(g, VACGPU_E_ARG, why);
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  if (idimmin < 1 || idimmin > g->cfg.ndim || idimmax < idimmin || idimmax > g->cfg.ndim)
    return fail(g, VACGPU_E_ARG, "idim out of range");
  const bool unsplit = idimmin != idimmax;
  if (unsplit != (m->unsplit != 0) || (unsplit && (idimmin != 1 || idimmax != g->cfg.ndim)))
    return fail(g, VACGPU_E_ARG, "an unsplit sweep covers idim = 1..ndim and needs method.unsplit = 1; a split one needs idimmin == idimmax and method.unsplit = 0");
  int rc;
  if (dtcourant || nerror || smallp) {
    if ((rc = read_scalars(g))) return rc;
    if (dtcourant)
      for (int d = 0; d < g->cfg.ndim; ++d) g->sc_host->dtcourant[d] = dtcourant[d];
    if (nerror)
      for (int i = 0; i < 4; ++i) g->sc_host->nerror[i] = nerror[i];
    if (smallp && *smallp >= 0.0) g->sc_host->smallp = *smallp;
    if ((rc = write_scalars(g))) return rc;
  }
  if (!smallp) {
    if ((rc = read_scalars(g))) return rc;
  }
  if (g->sc_host->smallp < 0.0)
    return fail(g, VACGPU_E_STATE, "smallp is not set: call vacgpu_getdt_courant first or pass smallp >= 0");
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, g->stream, g->sc, dt);
  if ((rc = launch_check(g, "k_set_dt"))) return rc;
  if ((rc = unsplit ? sweep_unsplit_device(g, m) : sweep_device(g, m, idimmin - 1, firstsweep, lastsweep)))
    return rc;
  if (dtcourant || nerror || smallp) {
    if ((rc = read_scalars(g))) return rc;
    if (dtcourant)
      for (int d = 0; d < g->cfg.ndim; ++d) dtcourant[d] = g->sc_host->dtcourant[d];
    if (nerror)
      for (int i = 0; i < 4; ++i) nerror[i] = g->sc_host->nerror[i];
    if (smallp) *smallp = g->sc_host->smallp;
  }
  return 0;
}

int vacgpu_set_time(vacgpu_grid* g, int it, double t) {
  if (!g) return VACGPU_E_ARG;
  int rc;
  if ((rc = read_scalars(g))) return rc;
  g->sc_host->it = it;
  g->sc_host->t = t;
  g->sc_host->itmin = it;  // the run's first step (VAC's itmin, src/vacio.t:160-170)
  return write_scalars(g);
}

int vacgpu_get_time(vacgpu_grid* g, int* it, double* t) {
  if (!g) return VACGPU_E_ARG;
  int rc;
  if ((rc = read_scalars(g))) return rc;
  if (it) *it = g->sc_host->it;
  if (t) *t = g->sc_host->t;
  return 0;
}

int vacgpu_run(vacgpu_grid* g, const vacgpu_method* m, int nsteps, int itmax, double tmax,
               int tmaxexact, double dtpar, double* t_hist) {
  if (!g) return VACGPU_E_ARG;
  std::string why;
  if (check_method(&g->cfg, m, why)) return fail(g, VACGPU_E_ARG, why);
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  if (dtpar <= 0.0 && m->courantpar <= 0.0) return fail(g, VACGPU_E_ARG, "need courantpar > 0 or dtpar > 0");
  int rc;
  if ((rc = read_scalars(g))) return rc;
  if (dtpar > 0.0 && g->sc_host->smallp < 0.0)
    return fail(g, VACGPU_E_ARG, "dtpar > 0 needs an explicit smallp >= 0 (no Courant pass sets it)");
  g->sc_host->tmax = tmax;
  g->sc_host->tmaxexact = tmaxexact;
  g->sc_host->dtcantgrow = m->dtcantgrow;
  g->sc_host->slowsteps = m->slowsteps;
  if ((rc = write_scalars(g))) return rc;
  int it = g->sc_host->it;
  const int itmin = g->sc_host->itmin;
  double t = g->sc_host->t;
  const bool sync_each = (tmax < kBig) || t_hist;
  // without a host sync per step, the TVD sweeps' Courant folds wait for
  // the next step's getdt; whatever way the loop ends, none is left behind
  struct Defer {
    vacgpu_grid* g;
    ~Defer() {
      g->defer_folds = false;
      if (g->pend_mask) flush_folds(g);
    }
  } defer{g};
  g->defer_folds = !sync_each;
  int done = 0;
  for (;;) {
    if (t_hist) t_hist[done] = t;
    // slowsteps = 1 takes dtpar for the run's first step only (src/vac.t:75)
    const double dtp = (m->slowsteps == 1 && it > itmin) ? -1.0 : dtpar;
    if (done == 0) rc = getdt_device(g, m, dtp, tmaxexact);
    else rc = getdt_lagged(g, m, dtp, tmaxexact);
    if (rc) return rc;
    double dt = 1.0;
    if (sync_each) {
      if ((rc = read_scalars(g))) return rc;
      dt = g->sc_host->dtchk;  // VAC tests dt < dtmin before slowsteps and tmaxexact
    }
    if (it >= itmax || t >= tmax || dt < 1.0e-99 || done >= nsteps) break;
    g->fold_step_end = true;
    g->step_end_done = false;
    rc = advance_device(g, m, it);
    g->fold_step_end = false;
    if (rc) return rc;
    if (!g->step_end_done) {
      hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, g->stream, g->sc);
      if ((rc = launch_check(g, "k_step_end"))) return rc;
    }
    it += 1;
    done += 1;
    if (sync_each) {
      if ((rc = read_scalars(g))) return rc;
      t = g->sc_host->t;
    }
  }
  HIPCHK(g, hipStreamSynchronize(g->stream));
  if (g->prof) harvest(g);
  return done;
}

int vacgpu_get_scalars(vacgpu_grid* g, double* dtcourant, int* nerror, double* smallp, double* dt) {
  if (!g) return VACGPU_E_ARG;
  int rc;
  if ((rc = read_scalars(g))) return rc;
  if (dtcourant)
    for (int d = 0; d < g->cfg.ndim; ++d) dtcourant[d] = g->sc_host->dtcourant[d];
  if (nerror)
    for (int i = 0; i < 4; ++i) nerror[i] = g->sc_host->nerror[i];
  if (smallp) *smallp = g->sc_host->smallp;
  if (dt) *dt = g->sc_host->dt;
  return 0;
}

int vacgpu_profile(vacgpu_grid* g, int enable) {
  if (!g) return VACGPU_E_ARG;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  harvest(g);
  g->prof = enable != 0;
  for (int k = 0; k < KF_N; ++k) { g->tsum[k] = 0.0; g->tcnt[k] = 0; }
  return 0;
}

int vacgpu_set_kernels(vacgpu_grid* g, int fused, int seglen) {
  if (!g) return VACGPU_E_ARG;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  g->use_fused = fused != 0;
  set_fused_shape(g, seglen);
  return 0;
}

int vacgpu_kernel_time(vacgpu_grid* g, int k, double* avg_ms, long* count) {
  if (!g || k < 0 || k >= KF_N) return VACGPU_E_ARG;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  harvest(g);
  if (avg_ms) *avg_ms = g->tcnt[k] ? g->tsum[k] / g->tcnt[k] : 0.0;
  if (count) *count = g->tcnt[k];
  return 0;
}

int vacgpu_check_quotients(const double* a, const double* b, long long n, double* q_kernel,
                           double* q_ieee) {
  if (!a || !b || !q_kernel || !q_ieee || n < 0) return VACGPU_E_ARG;
  if (n == 0) return 0;
  const size_t bytes = sizeof(double) * (size_t)n;
  double* d = nullptr;
  if (hipMalloc(&d, 4 * bytes) != hipSuccess) return VACGPU_E_HIP;
  hipError_t e = hipMemcpy(d, a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_quotients, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr,
                       d, d + n, n, d + 2 * n, d + 3 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(q_kernel, d + 2 * n, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(q_ieee, d + 3 * n, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : VACGPU_E_HIP;
}

int vacgpu_check_roots(const double* x, long long n, double* r_kernel, double* r_ieee) {
  if (!x || !r_kernel || !r_ieee || n < 0) return VACGPU_E_ARG;
  if (n == 0) return 0;
  const size_t bytes = sizeof(double) * (size_t)n;
  double* d = nullptr;
  if (hipMalloc(&d, 3 * bytes) != hipSuccess) return VACGPU_E_HIP;
  hipError_t e = hipMemcpy(d, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_roots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, d,
                       n, d + n, d + 2 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(r_kernel, d + n, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(r_ieee, d + 2 * n, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : VACGPU_E_HIP;
}

int vacgpu_snapshot_start(vacgpu_grid* g, int slot) {
  if (!g || slot < 0 || slot >= VACGPU_SNAP_SLOTS) return VACGPU_E_ARG;
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  auto& S = g->snap[slot];
  const size_t bytes = sizeof(double) * (size_t)g->geo.ps * g->nw;
  if (!g->copy_stream) HIPCHK(g, hipStreamCreateWithFlags(&g->copy_stream, hipStreamNonBlocking));
  if (S.pending) HIPCHK(g, hipEventSynchronize(S.done));  // the slot's last copy is out
  if (!S.dev) {
    HIPCHK(g, hipMalloc(&S.dev, bytes));
    HIPCHK(g, hipHostMalloc(&S.host, bytes, hipHostMallocDefault));
    HIPCHK(g, hipMalloc(&S.dsc, sizeof(Scal)));
    HIPCHK(g, hipHostMalloc(&S.hsc, sizeof(Scal), hipHostMallocDefault));
    HIPCHK(g, hipEventCreateWithFlags(&S.copied, hipEventDisableTiming));
    HIPCHK(g, hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
  }
  HIPCHK(g, hipMemcpyAsync(S.dev, g->buf[g->cur], bytes, hipMemcpyDeviceToDevice, g->stream));
  HIPCHK(g, hipMemcpyAsync(S.dsc, g->sc, sizeof(Scal), hipMemcpyDeviceToDevice, g->stream));
  HIPCHK(g, hipEventRecord(S.copied, g->stream));
  HIPCHK(g, hipStreamWaitEvent(g->copy_stream, S.copied, 0));
  HIPCHK(g, hipMemcpyAsync(S.host, S.dev, bytes, hipMemcpyDeviceToHost, g->copy_stream));
  HIPCHK(g, hipMemcpyAsync(S.hsc, S.dsc, sizeof(Scal), hipMemcpyDeviceToHost, g->copy_stream));
  HIPCHK(g, hipEventRecord(S.done, g->copy_stream));
  S.pending = S.valid = true;
  return 0;
}

// wait for snapshot `slot` and copy its ixG rows [lo, lo + n) of every
// variable into w with row pitch l[0] and plane pitch l[0] * l[1]
static int snapshot_rows(vacgpu_grid* g, int slot, double* w, const int* l, int lo, int* it,
                         double* t) {
  if (!g || !w || slot < 0 || slot >= VACGPU_SNAP_SLOTS) return VACGPU_E_ARG;
  auto& S = g->snap[slot];
  if (!S.valid) return fail(g, VACGPU_E_STATE, "no snapshot in this slot (call vacgpu_snapshot_start)");
  const Geo& geo = g->geo;
  const int nd = g->cfg.ndim;
  HIPCHK(g, hipEventSynchronize(S.done));
  S.pending = false;
  const int n0 = geo.ng[0] - 2 * lo, n1 = geo.ng[1] - 2 * lo;
  const int nz = nd == 3 ? geo.ng[2] - 2 * lo : 1;
  const int z0 = nd == 3 ? lo : 0;
  const size_t plane = (size_t)l[0] * l[1];
  const size_t var = plane * (nd == 3 ? l[2] : 1);
  for (int iw = 0; iw < g->nw; ++iw)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < n1; ++y)
        memcpy(w + (size_t)iw * var + (size_t)z * plane + (size_t)y * l[0],
               S.host + (size_t)iw * geo.ps + geo.off + (size_t)(z + z0) * geo.s[2] +
                   (size_t)(y + lo) * geo.s[1] + lo,
               sizeof(double) * n0);
  if (it) *it = S.hsc->it;
  if (t) *t = S.hsc->t;
  return 0;
}

int vacgpu_snapshot_fetch(vacgpu_grid* g, int slot, double* w, const int* ld, int* it, double* t) {
  if (!g || !w) return VACGPU_E_ARG;
  const Geo& geo = g->geo;
  const int nd = g->cfg.ndim;
  int l[3] = {geo.ng[0], geo.ng[1], nd == 3 ? geo.ng[2] : 1};
  if (ld)
    for (int d = 0; d < nd; ++d) {
      if (ld[d] < geo.ng[d]) return fail(g, VACGPU_E_ARG, "leading dimension smaller than ixG");
      l[d] = ld[d];
    }
  return snapshot_rows(g, slot, w, l, 0, it, t);
}

int vacgpu_snapshot_fetch_mesh(vacgpu_grid* g, int slot, double* w, int* it, double* t) {
  if (!g || !w) return VACGPU_E_ARG;
  const int l[3] = {g->geo.n[0], g->geo.n[1], g->cfg.ndim == 3 ? g->geo.n[2] : 1};
  return snapshot_rows(g, slot, w, l, 2, it, t);
}

int vacgpu_set_process(vacgpu_grid* g, int nproc1, const double* procpar, int itmin) {
  if (!g) return VACGPU_E_ARG;
  if (nproc1 != 0 && g->cfg.physics != VACGPU_MHD)
    return fail(g, VACGPU_E_ARG, "processing (the projection of B) needs MHD");
  g->pr.nproc1 = nproc1;
  g->pr.itmin = itmin;
  for (int i = 0; i < 3; ++i) g->pr.procpar[i] = procpar ? procpar[i] : -1.0;
  g->pr.divbmax = 0.0;  // projectb's saved values start over
  return 0;
}

int vacgpu_process(vacgpu_grid* g, int count, const double* procpar, int* info, int* iter,
                   double* resid) {
  if (!g) return VACGPU_E_ARG;
  if (!g->have_state) return fail(g, VACGPU_E_STATE, "no state on the device (call vacgpu_put_state)");
  if (count != 1 && count < 3)
    return fail(g, VACGPU_E_ARG, "process: count must be 1 (before a step or sweep) or ifile+2 (before a save)");
  const double* pp = procpar ? procpar : g->pr.procpar;
  return projectb_device(g, pp, count == 1 && !g->mid_step, info, iter, resid);
}

const char* vacgpu_last_error(const vacgpu_grid* g) {
  if (!g) return "null handle";
  return g->err.c_str();
}

void vacgpu_destroy(vacgpu_grid* g) {
  if (!g) return;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  harvest(g);
  for (int i = 0; i < 3; ++i)
    if (g->buf[i]) (void)hipFree(g->buf[i]);
  if (g->xbuf) (void)hipFree(g->xbuf);
  if (g->w1) (void)hipFree(g->w1);
  if (g->dxdev) (void)hipFree(g->dxdev);
  if (g->sc) (void)hipFree(g->sc);
  if (g->sc_host) (void)hipHostFree(g->sc_host);
  if (g->stage) (void)hipHostFree(g->stage);
  if (g->part) (void)hipFree(g->part);
  if (g->fold) (void)hipFree(g->fold);
  if (g->cslot) (void)hipFree(g->cslot);
  if (g->fixw) (void)hipFree(g->fixw);
  {
    auto& P = g->pr;
    double* planes[8] = {P.rhs, P.phi, P.z, P.z2, P.az, P.rdx[0], P.rdx[1], P.rdx[2]};
    for (double* q : planes)
      if (q) (void)hipFree(q);
    if (P.cg) (void)hipFree(P.cg);
    if (P.cg_host) (void)hipHostFree(P.cg_host);
  }
  for (auto& S : g->snap) {
    if (S.pending) (void)hipEventSynchronize(S.done);
    if (S.dev) (void)hipFree(S.dev);
    if (S.host) (void)hipHostFree(S.host);
    if (S.dsc) (void)hipFree(S.dsc);
    if (S.hsc) (void)hipHostFree(S.hsc);
    if (S.copied) (void)hipEventDestroy(S.copied);
    if (S.done) (void)hipEventDestroy(S.done);
  }
  if (g->copy_stream) (void)hipStreamDestroy(g->copy_stream);
  if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // extern "C"
