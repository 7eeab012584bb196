// sweep.hpp -- device-side kernels of the advect path, shared by the
// single-grid entry points (vacgpu.hip) and the batched ones (batch.hip).
//
// Kernel bodies are written as __device__ functions of the tile they work on
// (block coordinates bx, by, bz), so a single-grid launch maps blockIdx to
// the tile while a batched launch reads (instance, tile) from a work list.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "physics.hpp"

using namespace vacgpu;

namespace {

constexpr int BX = 64, BY = 4;      // cell kernels: 64 lanes along x, 4 rows

#ifndef VACGPU_NT
#define VACGPU_NT 1
#endif
// A sweep's output (and the predictor's w1) is read again only by a later
// launch, long after the 4 MB L2 of an XCD has turned over, so its stores
// are nontemporal: they do not displace the input rows and planes that
// neighbouring tiles re-read from the L2.  The stored values are the same.
template <class T>
__device__ __forceinline__ void st_out(T* p, T v) {
  if constexpr (VACGPU_NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
constexpr int NSLOT = 8;            // partial-reduction slots
enum { SLOT_C0 = 0, SLOT_C1 = 1, SLOT_PMIN = 2, SLOT_PMAX = 3, SLOT_DX = 4 };
enum { KF_TVD = 0, KF_PRED = 1, KF_CORR = 2, KF_RYU = 3, KF_BOUND = 4, KF_RED = 5, KF_FUSED = 6,
       KF_PROJ = 7, KF_N = 8 };

struct Geo {
  int nd;
  int n[3], ng[3];
  long long s[3];  // element strides per dimension
  long long ps;    // plane stride
  long long off;   // element offset of ixG cell (0,0,0)
  const double* dx[3];
};

struct Sel {
  int lim[8];
  int divbfix, divbwave, sourceunsplit;
  double courantpar, gamma;
  int ent[8];        // typeentropy per wave (ENT_*)
  double ecoef[8];   // entropycoef per wave ('yee')
  int prim;          // useprimitive
  int srcmul;        // idimmax-idimmin+1 of the sweep carrying the source (TVDLF correctors)
  // useprimitive: every upwindLR call converts its base state to primitive
  // variables and back in place (src/vactvdlf.t:328-364), so a corrector
  // reads states that earlier calls of the same advect have round-tripped:
  // the limiter's base rtw times, the update's base (w at the cell) rtb
  // times, the w1 of the face states rtu times and the w1 of the source
  // terms rts times
  int rtw, rtb, rtu, rts;
  // 'symm0' faces: bit iw of nf[2*d + upper] set when variable iw's flux
  // through that face is zero (setnoflux, src/vacgrid.t:416-449); read by
  // the general kernel variants only (LIM < 0 or ENT), which the host
  // selects when any face has it
  unsigned char nf[6];
};

// the variables whose flux through face J (between cells J and J+1) along d
// is zero: J = 1 is the lower mesh boundary face, n+1 the upper one
__device__ __forceinline__ unsigned noflux(const Sel& s, const Geo& g, int d, int J) {
  return (J == 1 ? unsigned(s.nf[2 * d]) : 0u) | (J == g.n[d] + 1 ? unsigned(s.nf[2 * d + 1]) : 0u);
}
template <int NW>
__device__ __forceinline__ void zero_flux(double* F, unsigned mask) {
#pragma unroll
  for (int iw = 0; iw < NW; ++iw)
    if ((mask >> iw) & 1u) F[iw] = 0.0;
}

// boundary types per (face, variable): t = 1 for 'periodic'; k the kind of a
// non-periodic face: a copy of the edge ('cont'), a mirror image ('symm', or
// 'asymm' with the sign flipped) or the values the first fill left ('fixed')
enum { BK_COPY = 0, BK_SYMM = 1, BK_ASYMM = 2, BK_FIXED = 3, BK_CONT1 = 4, BK_FIXED1 = 5, BK_GRAD1 = 6 };
struct BC {
  unsigned char t[6][8];
  unsigned char k[6][8];
};

// device-resident scalars
struct Scal {
  double dt, t, smallp;
  double dtcourant[3];
  double tmax;
  int nerror[4];
  int it, tmaxexact;
  // main-loop time-step controls (src/vac.t:60-81): the previous Courant
  // step (dtcantgrow), the run's first it (slowsteps), and dt before the
  // slowsteps and tmaxexact factors, which VAC's dt < dtmin test reads
  double dtold, dtchk;
  int itmin, dtcantgrow, slowsteps;
};

struct CellIx {
  int c[3];
  long long idx;
  bool ok;
};

__device__ __forceinline__ CellIx cell_at(const Geo& g, int bx, int by, int bz) {
  CellIx r;
  r.c[0] = 2 + (int)(bx * blockDim.x + threadIdx.x);
  r.c[1] = 2 + (int)(by * blockDim.y + threadIdx.y);
  r.c[2] = (g.nd == 3) ? 2 + bz : 0;
  r.ok = (r.c[0] <= g.n[0] + 1) && (r.c[1] <= g.n[1] + 1);
  r.idx = g.off + r.c[0] + (long long)r.c[1] * g.s[1] + (long long)r.c[2] * g.s[2];
  return r;
}

// Tile of a single-grid launch, XCD-aware.  The hardware hands the blocks of
// a launch to the eight XCDs in turn (block L to XCD L mod 8), each with its
// own L2; tiles that share halo rows, planes or cache lines should run on
// the same XCD at about the same time.  So XCD j takes the j-th contiguous
// share of a tile order: block L, the (L / 8)-th block on its XCD, gets
// order position start(L mod 8) + L / 8.  The order runs x fastest, then y
// within a group of 8 tile rows, then z, then the groups: neighbours along
// x are adjacent, along y 1 x-row apart (except across groups), along z one
// group layer (8 x-rows) apart, so a tile's halos were loaded by recent
// tiles of its XCD.  A bijection for any grid size, so every tile is
// computed exactly once whatever the dispatch order really is.
struct Tile {
  int x, y, z;
};
// Position of block L of an n-block launch in an XCD-contiguous order: XCD
// j (the blocks L with L mod 8 = j, in turn) takes the j-th contiguous share
// of the order (n/8 positions, one more for the first n mod 8 XCDs).  A
// bijection of [0, n).
__device__ __forceinline__ unsigned xcd_order(unsigned L, unsigned n) {
  const unsigned j = L & 7u, p = L >> 3, q = n >> 3, r = n & 7u;
  return j < r ? j * (q + 1) + p : r * (q + 1) + (j - r) * q + p;
}
__device__ __forceinline__ Tile xcd_tile() {
  constexpr unsigned YG = 8;  // tile rows per group
  const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const unsigned n = gx * gy * gz;
  const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  unsigned t = xcd_order(L, n);
  const unsigned full = gy / YG, per = gx * YG * gz;
  unsigned g = t / per, rows = YG;
  if (g >= full) {  // the last, partial group
    g = full;
    rows = gy - full * YG;
  }
  t -= g * per;
  const unsigned bx = t % gx;
  t /= gx;
  return Tile{(int)bx, (int)(g * YG + t % rows), (int)(t / rows)};
}

__device__ __forceinline__ CellIx cell_of(const Geo& g) {
  const Tile b = xcd_tile();
  return cell_at(g, b.x, b.y, b.z);
}

__device__ __forceinline__ long long block_linear() {
  return blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

template <int NW>
__device__ __forceinline__ void load(const double* __restrict__ w, long long ps, long long idx,
                                     double* u) {
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) u[iw] = w[iw * ps + idx];
}

// block-wide max / min of one value per thread (all threads must call)
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o, 64);
    v = v < u ? u : v;
  }
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}

template <bool IsMax>
__device__ __forceinline__ double block_reduce(double v, double* sm) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
  v = IsMax ? wave_max(v) : wave_min(v);
  __syncthreads();
  if ((tid & 63) == 0) sm[tid >> 6] = v;
  __syncthreads();
  if (tid < 64) {
    double u = (tid < (nthr + 63) / 64) ? sm[tid] : (IsMax ? -INFINITY : INFINITY);
    v = IsMax ? wave_max(u) : wave_min(u);
  }
  return v;  // valid in thread 0
}

// divB at a cell (src/vacphys.mhd0.t:119-133, src/vacgrid.t:822): sum over
// dims of 0.5*(B_d(i+1) - B_d(i-1))/dx_d(i), accumulated from 0.
template <int ND>
__device__ __forceinline__ double divb_at(const Geo& g, const double* __restrict__ w, long long idx,
                                          const int* c) {
  using L = Layout<ND, true>;
  double dv = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double* b = w + L::b(d) * g.ps;
    double gq = pdiv(0.5 * (b[idx + g.s[d]] - b[idx - g.s[d]]), g.dx[d][c[d]]);
    dv = dv + gq;
  }
  return dv;
}

// ------------------------------------------------------------ ghost cells
// getboundary (src/vacgrid.t:195-413) for 'cont', 'periodic', 'symm',
// 'asymm' and 'fixed' on slab grids, one thread per ghost cell.  VAC fills
// face regions in the order x-low, x-high, y-low, ... over the full
// transverse ixG range, so a ghost cell holds what its last face (the highest
// dimension along which it is a ghost) puts there: a copy of another cell --
// the edge ('cont'), the periodic image, the mirror image ('symm'; times -1
// for 'asymm', :317-352) -- which, if still a ghost along a lower dimension,
// got its value from that dimension's face earlier in the same fill; or, for
// a 'fixed' face, the value the first fill left there (fix, a copy of the
// state after that fill, :372-404; null during the first fill itself, when
// 'fixed' copies like 'cont').  For copies alone this is the mesh cell
// reached by mapping each coordinate on its own.
// The value getboundary leaves in the cell c of variable iw (q = its plane,
// fq = the plane of the first fill's state, or null during the first fill):
// the face of the highest dimension K along which c is a ghost decides, from
// values at positions that are mesh cells along K and are worked out the same
// way along the lower dimensions.  The extrapolating faces ('cont1', and
// 'fixed1' / 'grad1' on the first fill) combine two such values; every other
// kind takes one, so the recursion reads at most 2^ndim mesh cells.
template <int K>
__device__ __forceinline__ double ghost_value(const Geo& g, const BC& bc, const double* __restrict__ q,
                                              const double* __restrict__ fq, int iw, const int* c) {
  if constexpr (K < 0) {
    return q[g.off + c[0] + c[1] * g.s[1] + c[2] * g.s[2]];
  } else {
    if (K >= g.nd) return ghost_value<K - 1>(g, bc, q, fq, iw, c);
    const int ck = c[K], n = g.n[K];
    int f;
    if (ck < 2) f = 2 * K;
    else if (ck > n + 1) f = 2 * K + 1;
    else return ghost_value<K - 1>(g, bc, q, fq, iw, c);
    const int kind = bc.k[f][iw];
    int cc[3] = {c[0], c[1], c[2]};
    if (bc.t[f][iw]) {
      cc[K] = (f & 1) ? ck - n : ck + n;
      return ghost_value<K - 1>(g, bc, q, fq, iw, cc);
    }
    if ((kind == BK_FIXED || kind == BK_FIXED1) && fq)
      return fq[g.off + c[0] + c[1] * g.s[1] + c[2] * g.s[2]];
    if (kind == BK_SYMM || kind == BK_ASYMM) {
      cc[K] = (f & 1) ? 2 * n + 3 - ck : 3 - ck;
      const double v = ghost_value<K - 1>(g, bc, q, fq, iw, cc);
      return kind == BK_ASYMM ? -v : v;
    }
    const int e = (f & 1) ? n + 1 : 2;  // the edge cell along K
    cc[K] = e;
    if (kind == BK_GRAD1 && fq) {
      // fixB = w(c) - w(e) of the first fill, added to the current edge
      const long long oc = g.off + c[0] + c[1] * g.s[1] + c[2] * g.s[2];
      const long long oe = g.off + cc[0] + cc[1] * g.s[1] + cc[2] * g.s[2];
      return (fq[oc] - fq[oe]) + ghost_value<K - 1>(g, bc, q, fq, iw, cc);
    }
    if (kind == BK_CONT1 || kind == BK_FIXED1 || kind == BK_GRAD1) {
      // (|c-e|+1)*w(e) - |c-e|*w(f), f the next cell inward
      const int a = ck > e ? ck - e : e - ck;
      const double we = ghost_value<K - 1>(g, bc, q, fq, iw, cc);
      cc[K] = (f & 1) ? n : 3;
      const double wf = ghost_value<K - 1>(g, bc, q, fq, iw, cc);
      return double(a + 1) * we - double(a) * wf;
    }
    return ghost_value<K - 1>(g, bc, q, fq, iw, cc);  // 'cont', 'fixed' on the first fill
  }
}

template <int NW>
__device__ __forceinline__ void bound_body(const Geo& g, const BC& bc, double* __restrict__ w,
                                           long long t, long long slab0, long long slab1,
                                           int iw0 = 0, int iw1 = NW,
                                           const double* __restrict__ fix = nullptr) {
  int d;
  if (t < slab0) d = 0;
  else if (t < slab0 + slab1) { d = 1; t -= slab0; }
  else { d = 2; t -= slab0 + slab1; }
  int c[3] = {0, 0, 0};
  for (int k = 0; k < g.nd; ++k) {
    int ext = (k < d) ? g.n[k] : (k == d ? 4 : g.ng[k]);
    int v = (int)(t % ext);
    t /= ext;
    if (k < d) c[k] = 2 + v;
    else if (k == d) c[k] = (v < 2) ? v : g.n[k] + v;
    else c[k] = v;
  }
  const long long dst = g.off + c[0] + c[1] * g.s[1] + c[2] * g.s[2];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    if (iw < iw0 || iw >= iw1) continue;
    w[iw * g.ps + dst] = ghost_value<2>(g, bc, w + iw * g.ps, fix ? fix + iw * g.ps : nullptr, iw, c);
  }
}

// number of ghost cells and the slab sizes of bound_body's enumeration: slab d
// has the ghost coordinate along d, mesh coordinates below d and full ixG
// coordinates above d
inline long long ghost_slabs(const Geo& g, long long* slab) {
  long long tot = 0;
  for (int d = 0; d < 3; ++d) slab[d] = 0;
  for (int d = 0; d < g.nd; ++d) {
    long long v = 4;
    for (int k = 0; k < g.nd; ++k)
      if (k < d) v *= g.n[k];
      else if (k > d) v *= g.ng[k];
    slab[d] = v;
    tot += v;
  }
  return tot;
}

template <int NW>
__global__ __launch_bounds__(256) void k_bound(Geo g, BC bc, double* __restrict__ w,
                                              long long total, long long slab0, long long slab1,
                                              int iw0, int iw1, const double* __restrict__ fix) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  bound_body<NW>(g, bc, w, t, slab0, slab1, iw0, iw1, fix);
}

// ------------------------------------------------------------ TVDLF
// Hancock predictor (src/vactvdlf.t:5-89) with qdt = dt/2 at one cell, from
// its gathered inputs: uc = w(i), um / up = w(i -/+ e_D) for all variables,
// bn[k] = {B_k(i - e_k), B_k(i + e_k)} for the divergence along k != D (along
// D the B values come from um / up) and dxk[k] = dx_k(i):
// w1 = w - qdt/dx*(F(w+l/2) - F(w-l/2)) - (qdt/ndim)*S_Powell(w).
// The limited slopes l along D are returned in lout when it is not null.
// LIM: limiter type fixed at compile time (>= 0) or per variable from s
// (< 0); PWM: Powell source on (1), off (0) or as s selects (2).
// PRIM (useprimitive): the limiter and the w -+ l/2 states in primitive
// variables (upwindLR, src/vactvdlf.t:328-364); the source then sees the
// state after its conversion to primitive variables and back.
template <int ND, bool MHD, class R, int D, int LIM, int PWM, bool PRIM = false>
__device__ __forceinline__ void pred_compute(const Sel& s, const double* uc, const double* um,
                                             const double* up, const double (*bn)[2],
                                             const double* dxk, double dt, double smallp,
                                             double* out, double* lout, unsigned nfl = 0u,
                                             unsigned nfr = 0u) {
  constexpr int d = D;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const double qdt = dt / 2;
  double wl[NW], wr[NW];
  [[maybe_unused]] double ucr[NW];
  if constexpr (PRIM) {
    double qc[NW], qm[NW], qp[NW];
    to_prim<ND, MHD>(uc, s.gamma, smallp, qc);
    to_prim<ND, MHD>(um, s.gamma, smallp, qm);
    to_prim<ND, MHD>(up, s.gamma, smallp, qp);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R l = limiter_t<LIM, R>(s.lim[iw], R(qp[iw] - qc[iw]), R(qc[iw] - qm[iw]));
      wr[iw] = qc[iw] - 0.5 * double(l);
      wl[iw] = qc[iw] + 0.5 * double(l);
      if (lout) lout[iw] = double(l);
      ucr[iw] = qc[iw];
    }
    to_cons<ND, MHD>(wr, s.gamma);
    to_cons<ND, MHD>(wl, s.gamma);
    to_cons<ND, MHD>(ucr, s.gamma);
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R l = limiter_t<LIM, R>(s.lim[iw], R(up[iw] - uc[iw]), R(uc[iw] - um[iw]));
      wr[iw] = uc[iw] - 0.5 * double(l);
      wl[iw] = uc[iw] + 0.5 * double(l);
      if (lout) lout[iw] = double(l);
    }
  }
  double fL[NW], fR[NW];
  fluxes<ND, MHD, R>(wl, d, s.gamma, smallp, fL);
  fluxes<ND, MHD, R>(wr, d, s.gamma, smallp, fR);
  // 'symm0': fL is the flux through the cell's right face, fR its left one's
  if (nfl | nfr) {
    zero_flux<NW>(fR, nfl);
    zero_flux<NW>(fL, nfr);
  }
  const double qdx = pdiv(qdt, dxk[d]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) out[iw] = uc[iw] - qdx * (fL[iw] - fR[iw]);
  if constexpr (MHD && PWM != 0) {
    if (PWM == 1 || (s.sourceunsplit && s.divbfix)) {
      // divB (src/vacphys.mhd0.t:119-133, src/vacgrid.t:822), summed from 0
      double dv = 0.0;
#pragma unroll
      for (int k = 0; k < ND; ++k) {
        const double bm = (k == d) ? um[L::b(k)] : bn[k][0];
        const double bp = (k == d) ? up[L::b(k)] : bn[k][1];
        const double gq = pdiv(0.5 * (bp - bm), dxk[k]);
        dv = dv + gq;
      }
      powell<ND>(PRIM ? ucr : uc, dv, qdt * 1.0 / ND, out);
    }
  }
}

// the predictor at cell idx / c of a grid in HBM
template <int ND, bool MHD, class R, int D, int LIM = -1, int PWM = 2, bool PRIM = false>
__device__ __forceinline__ void pred_point(const Geo& g, const Sel& s, const double* __restrict__ w,
                                           long long idx, const int* c, double dt, double smallp,
                                           double* out) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const long long sd = g.s[D];
  double uc[NW], um[NW], up[NW], bn[ND][2], dxk[ND];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    uc[iw] = w[iw * g.ps + idx];
    um[iw] = w[iw * g.ps + idx - sd];
    up[iw] = w[iw * g.ps + idx + sd];
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    dxk[k] = g.dx[k][c[k]];
    bn[k][0] = bn[k][1] = 0.0;
    if constexpr (MHD && PWM != 0) {
      if (k != D && (PWM == 1 || (s.sourceunsplit && s.divbfix))) {
        bn[k][0] = w[L::b(k) * g.ps + idx - g.s[k]];
        bn[k][1] = w[L::b(k) * g.ps + idx + g.s[k]];
      }
    }
  }
  // 'symm0' faces at the mesh boundary (general variant only)
  const unsigned nfl = LIM < 0 ? noflux(s, g, D, c[D] - 1) : 0u;
  const unsigned nfr = LIM < 0 ? noflux(s, g, D, c[D]) : 0u;
  pred_compute<ND, MHD, R, D, LIM, PWM, PRIM>(s, uc, um, up, bn, dxk, dt, smallp, out, nullptr,
                                              nfl, nfr);
}

template <int ND, bool MHD, class R, int D>
__device__ __forceinline__ void tvdlf_pred_body(const Geo& g, const Sel& s,
                                                const double* __restrict__ w,
                                                double* __restrict__ w1, double dt, double smallp,
                                                int bx, int by, int bz) {
  constexpr int NW = Layout<ND, MHD>::nw;
  const CellIx cl = cell_at(g, bx, by, bz);
  if (!cl.ok) return;
  double out[NW];
  pred_point<ND, MHD, R, D>(g, s, w, cl.idx, cl.c, dt, smallp, out);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) st_out(w1 + iw * g.ps + cl.idx, out[iw]);
}


// the MUSCL states of the face J+1/2 from w1(J) = pl, w1(J+1) = pr and the
// limited slopes ll, lr: wL = pl + ll/2, wR = pr - lr/2, formed in primitive
// variables and converted back with PRIM (upwindLR, src/vactvdlf.t:328-364)
template <int ND, bool MHD, bool PRIM>
__device__ __forceinline__ void muscl_states(const Sel& s, const double* pl, const double* pr,
                                             const double* ll, const double* lr, double smallp,
                                             double* wlf, double* wrf) {
  constexpr int NW = Layout<ND, MHD>::nw;
  if constexpr (PRIM) {
    to_prim<ND, MHD>(pr, s.gamma, smallp, wrf);
    to_prim<ND, MHD>(pl, s.gamma, smallp, wlf);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      wrf[iw] = wrf[iw] - 0.5 * lr[iw];
      wlf[iw] = wlf[iw] + 0.5 * ll[iw];
    }
    to_cons<ND, MHD>(wrf, s.gamma);
    to_cons<ND, MHD>(wlf, s.gamma);
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      wrf[iw] = pr[iw] - 0.5 * lr[iw];
      wlf[iw] = pl[iw] + 0.5 * ll[iw];
    }
  }
}

// TVDLF corrector flux at the face J+1/2 (src/vactvdlf.t:147-254): MUSCL
// states wR = w1(J+1) - l(J+1)/2, wL = w1(J) + l(J)/2 from the predictor
// states pl = w1(J), pr = w1(J+1) and the limited slopes ll, lr of the
// start-of-sweep w (typelimited='original'); cmax of the mean state; the
// wLC = 2*mean - wRC restore (:205); F = (F(wL) + F(wR))/2 - cmax*(wR - wL)/2.
// cv = 2*cmax/(dx(J) + dx(J+1)) is the face's Courant value (:190).
// PRIM (useprimitive): the MUSCL states are formed in primitive variables.
template <int ND, bool MHD, class R, int D, bool PRIM = false>
__device__ __forceinline__ void corr_face(const Sel& s, const double* pl, const double* pr,
                                          const double* ll, const double* lr, double dxsum,
                                          double smallp, double* F, double& cv) {
  constexpr int d = D;
  constexpr int NW = Layout<ND, MHD>::nw;
  double wlf[NW], wrf[NW], mean[NW];
  muscl_states<ND, MHD, PRIM>(s, pl, pr, ll, lr, smallp, wlf, wrf);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) mean[iw] = 0.5 * (wlf[iw] + wrf[iw]);
  const R cm = cmax<ND, MHD, R>(mean, d, s.gamma, smallp);
  cv = pdiv(double(cm) * 2.0, dxsum);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) wlf[iw] = 2.0 * mean[iw] - wrf[iw];
  double fL[NW], fR[NW];
  fluxes<ND, MHD, R>(wlf, d, s.gamma, smallp, fL);
  fluxes<ND, MHD, R>(wrf, d, s.gamma, smallp, fR);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const double fm = 0.5 * (fL[iw] + fR[iw]);
    const R diss = -(cm * R(0.5) * R(wrf[iw] - wlf[iw]));
    F[iw] = fm + double(diss);
  }
}

// limited slope of w at a cell along d from w(i-1), w(i), w(i+1)
template <class R, int LIM = -1>
__device__ __forceinline__ double slope(int lim, double a, double b, double e) {
  return double(limiter_t<LIM, R>(lim, R(e - b), R(b - a)));
}

// TVDLF corrector (tvdmusclf, src/vactvdlf.t:92-299) for the cell i from its
// faces i-1/2 and i+1/2, then the Powell source of w1 with dt/ndim.
// Returns the max face Courant value (-inf outside the mesh).
template <int ND, bool MHD, class R, int D>
__device__ __forceinline__ double tvdlf_corr_body(const Geo& g, const Sel& s,
                                                  const double* __restrict__ w,
                                                  const double* __restrict__ w1,
                                                  double* __restrict__ wout, double dt,
                                                  double smallp, int bx, int by, int bz) {
  constexpr int d = D;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_at(g, bx, by, bz);
  double cmx = -INFINITY;
  if (cl.ok) {
    const long long sd = g.s[d];
    const int ci = cl.c[d];
    // limited slopes l(i-1), l(i), l(i+1) of w
    double lm[NW], l0[NW], lp[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double* p = w + iw * g.ps + cl.idx;
      const double a = p[-2 * sd], b = p[-sd], c = p[0], e = p[sd], f = p[2 * sd];
      lm[iw] = slope<R>(s.lim[iw], a, b, c);
      l0[iw] = slope<R>(s.lim[iw], b, c, e);
      lp[iw] = slope<R>(s.lim[iw], c, e, f);
    }
    double um[NW], u0[NW], up[NW];
    load<NW>(w1, g.ps, cl.idx - sd, um);
    load<NW>(w1, g.ps, cl.idx, u0);
    load<NW>(w1, g.ps, cl.idx + sd, up);
    double F0[NW], F1[NW], cv;
    corr_face<ND, MHD, R, D>(s, um, u0, lm, l0, g.dx[d][ci - 1] + g.dx[d][ci], smallp, F0, cv);
    cmx = cmx < cv ? cv : cmx;
    corr_face<ND, MHD, R, D>(s, u0, up, l0, lp, g.dx[d][ci] + g.dx[d][ci + 1], smallp, F1, cv);
    cmx = cmx < cv ? cv : cmx;
    double out[NW];
    const double qdx = pdiv(dt, g.dx[d][ci]);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw)
      out[iw] = w[iw * g.ps + cl.idx] - qdx * (F1[iw] - F0[iw]);
    if constexpr (MHD) {
      if (s.sourceunsplit && s.divbfix) {
        const double dv = divb_at<ND>(g, w1, cl.idx, cl.c);
        powell<ND>(u0, dv, dt * 1.0 / ND, out);
      }
    }
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) st_out(wout + iw * g.ps + cl.idx, out[iw]);
  }
  return cmx;
}

template <int ND, bool MHD, class R, int D>
__global__ __launch_bounds__(256) void k_tvdlf_corr(Geo g, Sel s, const double* __restrict__ w,
                                                   const double* __restrict__ w1,
                                                   double* __restrict__ wout,
                                                   const Scal* __restrict__ sc,
                                                   double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  const Tile tb = xcd_tile();
  const double cmx = tvdlf_corr_body<ND, MHD, R, D>(g, s, w, w1, wout, sc->dt, sc->smallp, tb.x,
                                                    tb.y, tb.z);
  const double bm = block_reduce<true>(cmx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bm;
}

// ------------------------------------------------------------ fused TVDLF (2D)
// One launch per sweep: predictor and corrector together, without the w1
// array.  A wave covers 64 consecutive x cells (cell 1 + 62*strip + lane;
// lanes 1..62 produce output) and marches along y through a segment of rows,
// keeping the predictor states of rows j-1, j, j+1 in registers.  Each face
// is evaluated once: along x by the lane left of it (the neighbour's values
// arrive by lane shuffles), along y once per marching step.  The predictor's
// limited slopes are the corrector's (typelimited='original': both limit the
// start-of-sweep w), so they are computed once.
//
// The predictor state of a ghost cell is the one of the mesh cell that
// getboundary copies into it (src/vacgrid.t:234-273), computed there directly.
// The corrector slope at a ghost cell is that image's slope for 'periodic'
// (the ghost values are copies of the image's neighbourhood) and exactly 0
// for 'cont' (both differences vanish).  Output cells are also written to the
// ghost cells that take their value, so the sweep leaves wout's ghost cells
// filled.  The Courant maximum of the sweep is folded by the last block to
// finish.  All of this needs one boundary type per face for all variables,
// at least 4 cells per dimension and planes below 2^29 elements, which the
// host checks before choosing these kernels.
constexpr int FS_STRIDE = 62;

__device__ __forceinline__ int ghost_image(int c, int n, bool per_lo, bool per_hi) {
  const int lo = per_lo ? c + n : 2, hi = per_hi ? c - n : n + 1;
  return c < 2 ? lo : (c > n + 1 ? hi : c);
}

// whether mesh cell c (n >= 4) is the image of some ghost cell: the cheap
// test that lets interior cells skip ghost_preimages
__device__ __forceinline__ bool has_preimage(int c, int n, bool per_lo, bool per_hi) {
  return (per_lo ? c >= n : c == 2) || (per_hi ? c <= 3 : c == n + 1);
}

// ghost cells whose image is mesh cell c (n >= 4), written to t; returns the count
__device__ __forceinline__ int ghost_preimages(int c, int n, bool per_lo, bool per_hi, int* t) {
  int k = 0;
  if (per_lo) { if (c >= n) t[k++] = c - n; }
  else if (c == 2) { t[k++] = 0; t[k++] = 1; }
  if (per_hi) { if (c <= 3) t[k++] = c + n; }
  else if (c == n + 1) { t[k++] = n + 2; t[k++] = n + 3; }
  return k;
}

// Whole-wave shifts by one lane through DPP (wave_shl:1 / wave_shr:1, one
// v_mov_b32_dpp per 32-bit half, no LDS round trip).  lane_next gives lane i
// the value of lane i+1, lane_prev that of lane i-1; the lane with no source
// (63, resp. 0) reads 0 (bound_ctrl), so no copy of the old value is needed.
// Every caller uses lanes 0 and 63 as halo lanes whose shifted-in values
// reach no output.
template <int CTRL, class T>
__device__ __forceinline__ T dpp_shift(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const unsigned rlo = (unsigned)__builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, true);
    const unsigned rhi = (unsigned)__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(T, ((unsigned long long)rhi << 32) | rlo);
  } else {
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true));
  }
}
template <class T>
__device__ __forceinline__ T lane_next(T v) { return dpp_shift<0x130>(v); }
template <class T>
__device__ __forceinline__ T lane_prev(T v) { return dpp_shift<0x138>(v); }

// store out at mesh cell c (index idx) and, when ghosts is set, at every
// ghost cell whose image it is (one boundary type per face for all
// variables, n >= 4): the sweep then leaves the output's ghost cells filled
// NT = false: plain stores (k_ryu_ab_y, whose waves write 62-cell rows one
// marching step at a time, measured 8% slower with nontemporal ones)
template <int ND, int NW, class T = double, bool NT = true>
__device__ __forceinline__ void store_cell(const Geo& g, const BC& bc, T* __restrict__ wout,
                                           long long idx, const int* c, const T* out,
                                           int ghosts) {
  auto st = [](T* p, T v) {
    if constexpr (NT) st_out(p, v);
    else *p = v;
  };
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) st(wout + iw * g.ps + idx, out[iw]);
  if (!ghosts) return;
  bool edge = false;
#pragma unroll
  for (int k = 0; k < ND; ++k) edge |= has_preimage(c[k], g.n[k], bc.t[2 * k][0], bc.t[2 * k + 1][0]);
  if (!edge) return;  // an interior cell
  int t[3][5], nt[3] = {1, 1, 1};
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    t[k][0] = c[k];
    nt[k] = 1 + ghost_preimages(c[k], g.n[k], bc.t[2 * k][0], bc.t[2 * k + 1][0], &t[k][1]);
  }
  if (nt[0] + nt[1] + nt[2] == 3) return;  // an interior cell
  for (int a = 0; a < nt[0]; ++a)
    for (int b = 0; b < nt[1]; ++b)
      for (int e = 0; e < (ND == 3 ? nt[2] : 1); ++e) {
        if (a == 0 && b == 0 && e == 0) continue;
        long long o = g.off + t[0][a] + t[1][b] * g.s[1];
        if (ND == 3) o += t[2][e] * g.s[2];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) st(wout + iw * g.ps + o, out[iw]);
      }
}

// predictor launch: one cell per thread; with ghosts set the ghost cells of
// w1 that take a cell's value are written too (one boundary type per face).
// LIM >= 0: one limiter type for every variable and no 'symm0' face.
// DW (FP32-bulk, copy-type faces): w1 is stored as its FP32 increment over
// the sweep's input, w1 - w, in the first half of the w1 buffer (planes of
// ps floats); the corrector adds it back to w in FP64.  A ghost cell's
// increment is its image's, as its w is its image's.
template <int ND, bool MHD, class R, int D, bool PRIM = false, int LIM = -1, bool DW = false>
__global__ __launch_bounds__(256) void k_tvdlf_pred(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                   double* __restrict__ w1,
                                                   const Scal* __restrict__ sc, int ghosts) {
  constexpr int NW = Layout<ND, MHD>::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double out[NW];
  pred_point<ND, MHD, R, D, LIM, 2, PRIM>(g, s, w, cl.idx, cl.c, sc->dt, sc->smallp, out);
  if constexpr (DW) {
    float dw[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) dw[iw] = float(out[iw] - w[iw * g.ps + cl.idx]);
    store_cell<ND, NW, float>(g, bc, reinterpret_cast<float*>(w1), cl.idx, cl.c, dw, ghosts);
  } else {
    store_cell<ND, NW>(g, bc, w1, cl.idx, cl.c, out, ghosts);
  }
}

// w1 at idx in the increment form: w + the stored FP32 increment
template <int NW>
__device__ __forceinline__ void load_dw(const double* __restrict__ w, const double* __restrict__ w1,
                                        long long ps, long long idx, double* u) {
  const float* d = reinterpret_cast<const float*>(w1);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) u[iw] = w[iw * ps + idx] + double(d[iw * ps + idx]);
}
__device__ __forceinline__ double ld_dw(const double* __restrict__ w, const double* __restrict__ w1,
                                        long long ps, int iw, long long idx) {
  return w[iw * ps + idx] + double(reinterpret_cast<const float*>(w1)[iw * ps + idx]);
}


// ---- face-parallel TVDLF corrector (two-launch form, 2D and 3D): each face
// is evaluated once.  x sweep: lane L of a wave owns the face c+1/2 of cell
// c = x0 - 1 + L and updates c for L in [1, 62] (the left face arrives by a
// lane shuffle); y / z sweeps: a thread marches along the sweep direction,
// carrying the face below and a window of w and w1.  The arithmetic of a
// face and of a cell update is tvdlf_corr_body's.
constexpr int CX_STRIDE = 62;

template <int ND, bool MHD, class R, int LIM, bool PRIM = false, bool DW = false>
__global__ __launch_bounds__(256) void k_tvdlf_corrx(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                    const double* __restrict__ w1,
                                                    double* __restrict__ wout,
                                                    const Scal* __restrict__ sc,
                                                    double* __restrict__ part, long long pstride,
                                                    int ghosts, const double* __restrict__ wl) {
  constexpr int D = 0;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const int lane = (int)threadIdx.x;
  const Tile tb = xcd_tile();
  const int c = 1 + tb.x * CX_STRIDE + lane;
  const int y = 2 + (int)(tb.y * BY + threadIdx.y);
  const int z = (ND == 3) ? 2 + tb.z : 0;
  const bool rowok = y <= g.n[1] + 1;
  const bool useful = rowok && lane >= 1 && lane <= CX_STRIDE && c <= g.n[0] + 1;
  const bool face_ok = rowok && lane <= CX_STRIDE && c <= g.n[0] + 1;
  const int cc = c < g.n[0] + 2 ? c : g.n[0] + 2;  // idle lanes: in-bounds, unused
  const int yc = rowok ? y : g.n[1] + 1;
  const long long idx = g.off + cc + (long long)yc * g.s[1] + (long long)z * g.s[2];
  double l[NW], lr[NW], u0[NW], u1[NW];
  [[maybe_unused]] double ub[NW];
  if constexpr (PRIM) {
    // useprimitive: limit the primitive variables of wl after its round
    // trips, update w after its own, face states from w1 after its own
    double qa[NW];
    load<NW>(wl, g.ps, idx - 1, qa);
    load<NW>(wl, g.ps, idx, ub);
    load<NW>(wl, g.ps, idx + 1, u1);
    prim_after<ND, MHD>(qa, s.rtw, s.gamma, smallp);
    prim_after<ND, MHD>(ub, s.rtw, s.gamma, smallp);
    prim_after<ND, MHD>(u1, s.rtw, s.gamma, smallp);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) l[iw] = slope<R, LIM>(s.lim[iw], qa[iw], ub[iw], u1[iw]);
    load<NW>(w, g.ps, idx, ub);
    roundtrips<ND, MHD>(ub, s.rtb, s.gamma, smallp);
    load<NW>(w1, g.ps, idx, u0);
    load<NW>(w1, g.ps, idx + 1, u1);
    roundtrips<ND, MHD>(u0, s.rtu, s.gamma, smallp);
    roundtrips<ND, MHD>(u1, s.rtu, s.gamma, smallp);
  } else if constexpr (DW) {
    // increment form (the limiter's base is w): w1 at this lane's cell, at
    // the next cell from the next lane
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double* p = w + iw * g.ps + idx;
      l[iw] = slope<R, LIM>(s.lim[iw], p[-1], p[0], p[1]);
    }
    load_dw<NW>(w, w1, g.ps, idx, u0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) u1[iw] = lane_next(u0[iw]);
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double* p = wl + iw * g.ps + idx;  // the limiter's base (typelimited)
      l[iw] = slope<R, LIM>(s.lim[iw], p[-1], p[0], p[1]);
      u0[iw] = w1[iw * g.ps + idx];
      u1[iw] = w1[iw * g.ps + idx + 1];
    }
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) lr[iw] = lane_next(l[iw]);
  double F[NW], cv;
  corr_face<ND, MHD, R, D, PRIM>(s, u0, u1, l, lr, g.dx[0][cc] + g.dx[0][cc + 1], smallp, F, cv);
  if constexpr (LIM < 0) zero_flux<NW>(F, noflux(s, g, 0, cc));
  double cmx = face_ok ? cv : -INFINITY;
  double out[NW];
  const double qdx = pdiv(dt, g.dx[0][cc]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const double Fm = lane_prev(F[iw]);
    out[iw] = (PRIM ? ub[iw] : w[iw * g.ps + idx]) - qdx * (F[iw] - Fm);
  }
  if constexpr (MHD) {
    if (s.sourceunsplit && s.divbfix) {
      const int c3[3] = {cc, yc, z};
      double dv;
      if constexpr (DW) {
        // divb_at's arithmetic on w1: B_x of the x neighbours from the next
        // and previous lanes (lanes 1..62 store), the rest loaded
        using LB = Layout<ND, true>;
        const double bxp = lane_next(u0[LB::b(0)]), bxm = lane_prev(u0[LB::b(0)]);
        dv = 0.0;
        const double g0 = pdiv(0.5 * (bxp - bxm), g.dx[0][cc]);
        dv = dv + g0;
#pragma unroll
        for (int d = 1; d < ND; ++d) {
          const double bp = ld_dw(w, w1, g.ps, LB::b(d), idx + g.s[d]);
          const double bm = ld_dw(w, w1, g.ps, LB::b(d), idx - g.s[d]);
          const double gq = pdiv(0.5 * (bp - bm), g.dx[d][c3[d]]);
          dv = dv + gq;
        }
      } else {
        dv = divb_at<ND>(g, w1, idx, c3);
      }
      if constexpr (PRIM) {
        double us[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) us[iw] = u0[iw];
        roundtrips<ND, MHD>(us, s.rts - s.rtu, s.gamma, smallp);
        powell<ND>(us, dv, dt * double(s.srcmul) / ND, out);
      } else {
        powell<ND>(u0, dv, dt * double(s.srcmul) / ND, out);
      }
    }
  }
  if (useful) {
    const int c3[3] = {cc, yc, z};
    store_cell<ND, NW>(g, bc, wout, idx, c3, out, ghosts);
  }
  const double bm = block_reduce<true>(cmx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bm;
}

template <int ND, bool MHD, class R, int D, int LIM, bool PRIM = false, bool DW = false>
__global__ __launch_bounds__(256) void k_tvdlf_corrm(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                    const double* __restrict__ w1,
                                                    double* __restrict__ wout,
                                                    const Scal* __restrict__ sc,
                                                    double* __restrict__ part, long long pstride,
                                                    int seglen, int ghosts,
                                                    const double* __restrict__ wl) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const Tile tb = xcd_tile();
  const int x = 2 + (int)(tb.x * BX + threadIdx.x);
  const int seg = (int)(tb.y * BY + threadIdx.y);
  const int s0 = 2 + seg * seglen;
  const int t = (ND == 3) ? 2 + tb.z : 0;  // the coordinate that is neither x nor D
  const int other = (D == 1) ? 2 : 1;
  const int nd_ = g.n[D];
  double cmx = -INFINITY;
  if (x <= g.n[0] + 1 && s0 <= nd_ + 1) {
    const int s1 = (s0 + seglen - 1 < nd_ + 1) ? s0 + seglen - 1 : nd_ + 1;
    const long long sd = g.s[D];
    const long long base = g.off + x + (ND == 3 ? (long long)t * g.s[other] : 0);
    const double* dxd = g.dx[D];
    int c3[3] = {x, 0, 0};
    if (ND == 3) c3[other] = t;
    // window: the limiter's base wl (typelimited) at J-1, J, J+1 (for the
    // slope at J) and w1 at J; slope at J.  The update's base is w, the
    // window itself when wl is w.
    // With PRIM (useprimitive) the window holds the primitive variables of
    // wl after its round trips (s.rtw), the face states come from w1 after
    // its own (s.rtu) and the update's base is w after its own (s.rtb).
    const bool own = wl == w;
    auto lwin = [&](long long at, double* v) {
      load<NW>(wl, g.ps, at, v);
      if constexpr (PRIM) prim_after<ND, MHD>(v, s.rtw, s.gamma, smallp);
    };
    auto lw1 = [&](long long at, double* v) {
      load<NW>(w1, g.ps, at, v);
      if constexpr (PRIM) roundtrips<ND, MHD>(v, s.rtu, s.gamma, smallp);
    };
    // increment form (DW, the limiter's base is w): w1 at J is the window's
    // w at J plus the stored increment
    auto dw1 = [&](long long at, const double* wj, double* v) {
      const float* d = reinterpret_cast<const float*>(w1);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) v[iw] = wj[iw] + double(d[iw * g.ps + at]);
    };
    double wa[NW], wb[NW], wc[NW], u0[NW], lc[NW], Fp[NW];
    [[maybe_unused]] double bdm = 0.0;  // B_D of w1 one cell below (DW)
    lwin(base + (long long)(s0 - 2) * sd, wa);
    lwin(base + (long long)(s0 - 1) * sd, wb);
    lwin(base + (long long)s0 * sd, wc);
    if constexpr (DW) dw1(base + (long long)(s0 - 1) * sd, wb, u0);
    else lw1(base + (long long)(s0 - 1) * sd, u0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) lc[iw] = slope<R, LIM>(s.lim[iw], wa[iw], wb[iw], wc[iw]);
    // face s0-1/2
    {
      double u1[NW], ln[NW], wd[NW];
      if constexpr (DW) dw1(base + (long long)s0 * sd, wc, u1);
      else lw1(base + (long long)s0 * sd, u1);
      lwin(base + (long long)(s0 + 1) * sd, wd);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) ln[iw] = slope<R, LIM>(s.lim[iw], wb[iw], wc[iw], wd[iw]);
      double cv;
      corr_face<ND, MHD, R, D, PRIM>(s, u0, u1, lc, ln, dxd[s0 - 1] + dxd[s0], smallp, Fp, cv);
      if constexpr (LIM < 0) zero_flux<NW>(Fp, noflux(s, g, D, s0 - 1));
      cmx = cmx < cv ? cv : cmx;
      if constexpr (DW && MHD) bdm = u0[Layout<ND, true>::b(D)];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        u0[iw] = u1[iw];
        lc[iw] = ln[iw];
        wa[iw] = wb[iw];
        wb[iw] = wc[iw];
        wc[iw] = wd[iw];
      }
    }
    // now: u0 = w1(i), lc = slope(i), wb = w(i), wc = w(i+1)
    for (int i = s0; i <= s1; ++i) {
      double u1[NW], ln[NW], wd[NW], F[NW], cv;
      if constexpr (DW) dw1(base + (long long)(i + 1) * sd, wc, u1);
      else lw1(base + (long long)(i + 1) * sd, u1);
      lwin(base + (long long)(i + 2) * sd, wd);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) ln[iw] = slope<R, LIM>(s.lim[iw], wb[iw], wc[iw], wd[iw]);
      corr_face<ND, MHD, R, D, PRIM>(s, u0, u1, lc, ln, dxd[i] + dxd[i + 1], smallp, F, cv);
      if constexpr (LIM < 0) zero_flux<NW>(F, noflux(s, g, D, i));
      cmx = cmx < cv ? cv : cmx;
      const long long idx = base + (long long)i * sd;
      const double qdx = pdiv(dt, dxd[i]);
      double out[NW];
      [[maybe_unused]] double ub[NW];
      if constexpr (PRIM) {
        load<NW>(w, g.ps, idx, ub);
        roundtrips<ND, MHD>(ub, s.rtb, s.gamma, smallp);
      }
#pragma unroll
      for (int iw = 0; iw < NW; ++iw)
        out[iw] = (PRIM ? ub[iw] : (own ? wb[iw] : w[iw * g.ps + idx])) - qdx * (F[iw] - Fp[iw]);
      if constexpr (MHD) {
        if (s.sourceunsplit && s.divbfix) {
          c3[D] = i;
          double dv;
          if constexpr (DW) {
            // divb_at's arithmetic on w1: B_D below and above from the window
            using LB = Layout<ND, true>;
            dv = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
              const double bp = d == D ? u1[LB::b(d)] : ld_dw(w, w1, g.ps, LB::b(d), idx + g.s[d]);
              const double bm = d == D ? bdm : ld_dw(w, w1, g.ps, LB::b(d), idx - g.s[d]);
              const double gq = pdiv(0.5 * (bp - bm), g.dx[d][c3[d]]);
              dv = dv + gq;
            }
          } else {
            dv = divb_at<ND>(g, w1, idx, c3);
          }
          if constexpr (PRIM) {
            double us[NW];
#pragma unroll
            for (int iw = 0; iw < NW; ++iw) us[iw] = u0[iw];
            roundtrips<ND, MHD>(us, s.rts - s.rtu, s.gamma, smallp);
            powell<ND>(us, dv, dt * double(s.srcmul) / ND, out);
          } else {
            powell<ND>(u0, dv, dt * double(s.srcmul) / ND, out);
          }
        }
      }
      c3[D] = i;
      store_cell<ND, NW>(g, bc, wout, idx, c3, out, ghosts);
      if constexpr (DW && MHD) bdm = u0[Layout<ND, true>::b(D)];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        Fp[iw] = F[iw];
        u0[iw] = u1[iw];
        lc[iw] = ln[iw];
        wb[iw] = wc[iw];
        wc[iw] = wd[iw];
      }
    }
  }
  const double bm = block_reduce<true>(cmx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bm;
}


// ---- unsplit TVDLF (dimsplit=F, src/vac.t:325-329): one advect over all
// directions.  hancock and tvdmusclf loop over idim inside one call
// (src/vactvdlf.t:33-81, :137-287), adding each direction's flux
// differences to wnew in turn, and add the unsplit source once after the
// last direction with qdt*(idimmax-idimmin+1)/ndim (:85-86, :294-295).
// The predictor: one thread per cell, every direction in order.  The
// corrector: the face-parallel correctors of the split sweeps, one launch
// per direction, each adding its direction's differences to the previous
// one's result.

// the predictor's flux difference of direction D at a cell, subtracted from out.
// PRIM (useprimitive): hancock's upwindLR converts its base in place for
// every direction (src/vactvdlf.t:33-45), so direction D (the D-th in turn)
// limits and reconstructs the primitive variables of w after D round trips.
template <int ND, bool MHD, class R, int D, bool PRIM = false>
__device__ __forceinline__ void pred_dir(const Geo& g, const Sel& s, const double* __restrict__ w,
                                         long long idx, const int* c, const double* uc, double qdt,
                                         double smallp, double* out, double (*bn)[2]) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  double um[NW], up[NW], wl[NW], wr[NW];
  load<NW>(w, g.ps, idx - g.s[D], um);
  load<NW>(w, g.ps, idx + g.s[D], up);
  if constexpr (MHD) {
    bn[D][0] = um[L::b(D)];
    bn[D][1] = up[L::b(D)];
  }
  if constexpr (PRIM) {
    double qc[NW], qm[NW], qp[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) qc[iw] = uc[iw];
    prim_after<ND, MHD>(qc, D, s.gamma, smallp);
    prim_after<ND, MHD>(um, D, s.gamma, smallp);
    prim_after<ND, MHD>(up, D, s.gamma, smallp);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      qm[iw] = um[iw];
      qp[iw] = up[iw];
    }
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R l = limiter_t<-1, R>(s.lim[iw], R(qp[iw] - qc[iw]), R(qc[iw] - qm[iw]));
      wr[iw] = qc[iw] - 0.5 * double(l);
      wl[iw] = qc[iw] + 0.5 * double(l);
    }
    to_cons<ND, MHD>(wr, s.gamma);
    to_cons<ND, MHD>(wl, s.gamma);
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R l = limiter_t<-1, R>(s.lim[iw], R(up[iw] - uc[iw]), R(uc[iw] - um[iw]));
      wr[iw] = uc[iw] - 0.5 * double(l);
      wl[iw] = uc[iw] + 0.5 * double(l);
    }
  }
  double fL[NW], fR[NW];
  fluxes<ND, MHD, R>(wl, D, s.gamma, smallp, fL);
  fluxes<ND, MHD, R>(wr, D, s.gamma, smallp, fR);
  zero_flux<NW>(fR, noflux(s, g, D, c[D] - 1));  // 'symm0' faces
  zero_flux<NW>(fL, noflux(s, g, D, c[D]));
  const double qdx = pdiv(qdt, g.dx[D][c[D]]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) out[iw] = out[iw] - qdx * (fL[iw] - fR[iw]);
}

template <int ND, bool MHD, class R, bool PRIM = false>
__global__ __launch_bounds__(256) void k_tvdlf_pred_u(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                     double* __restrict__ w1,
                                                     const Scal* __restrict__ sc, int ghosts) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  const double qdt = sc->dt / 2, smallp = sc->smallp;
  double uc[NW], out[NW], bn[3][2];
  load<NW>(w, g.ps, cl.idx, uc);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) out[iw] = uc[iw];
  pred_dir<ND, MHD, R, 0, PRIM>(g, s, w, cl.idx, cl.c, uc, qdt, smallp, out, bn);
  pred_dir<ND, MHD, R, 1, PRIM>(g, s, w, cl.idx, cl.c, uc, qdt, smallp, out, bn);
  if constexpr (ND == 3) pred_dir<ND, MHD, R, 2, PRIM>(g, s, w, cl.idx, cl.c, uc, qdt, smallp, out, bn);
  if constexpr (MHD) {
    if (s.sourceunsplit && s.divbfix) {
      double dv = 0.0;
#pragma unroll
      for (int k = 0; k < ND; ++k) dv = dv + pdiv(0.5 * (bn[k][1] - bn[k][0]), g.dx[k][cl.c[k]]);
      if constexpr (PRIM) {
        // the source sees the state after every direction's round trip
        double us[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) us[iw] = uc[iw];
        roundtrips<ND, MHD>(us, ND, s.gamma, smallp);
        powell<ND>(us, dv, qdt * double(ND) / ND, out);
      } else {
        powell<ND>(uc, dv, qdt * double(ND) / ND, out);
      }
    }
  }
  store_cell<ND, NW>(g, bc, w1, cl.idx, cl.c, out, ghosts);
}

// useprimitive with the limiter's base wold: upwindLR converts wold to
// primitive variables and back in place over ixG every time it limits it
// (src/vactvdlf.t:328-332, :360-364); one thread per ixG cell
template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_roundtrip(Geo g, double* __restrict__ w, double gamma,
                                                  const Scal* __restrict__ sc) {
  constexpr int NW = Layout<ND, MHD>::nw;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)g.ng[0] * g.ng[1] * (ND == 3 ? g.ng[2] : 1);
  if (t >= n) return;
  const int x = (int)(t % g.ng[0]);
  const int y = (int)((t / g.ng[0]) % g.ng[1]);
  const int z = ND == 3 ? (int)(t / ((long long)g.ng[0] * g.ng[1])) : 0;
  const long long idx = g.off + x + (long long)y * g.s[1] + (long long)z * g.s[2];
  double u[NW];
  load<NW>(w, g.ps, idx, u);
  roundtrips<ND, MHD>(u, 1, gamma, sc->smallp);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) w[iw * g.ps + idx] = u[iw];
}

// Wait for every vector memory load in flight (vmcnt(0); expcnt and lgkmcnt
// left at their maxima).  The fused y sweep calls it after a step's
// arithmetic and before its stores: its prefetch for the next step has had the whole step
// to arrive, and the next step's first use of the prefetched registers then
// needs no wait.  Without it that use must wait with a count that cannot
// tell the step's stores (a data-dependent number) from the loads, i.e. for
// the stores just issued as well, since memory operations complete in order.
__device__ __forceinline__ void prefetch_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// plane iw at byte offset o (planes are addressed with 32-bit byte offsets)
__device__ __forceinline__ double ldp(const double* __restrict__ w, long long ps, int iw, unsigned o) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(w + iw * ps) + o);
}
__device__ __forceinline__ void stp(double* __restrict__ w, long long ps, int iw, unsigned o, double v) {
  st_out(reinterpret_cast<double*>(reinterpret_cast<char*>(w + iw * ps) + o), v);
}

struct FusedGeo {  // per-wave constants of a fused sweep
  int n0, n1;
  unsigned sy8;     // row stride in bytes
  unsigned base8;   // byte offset of ixG cell (0, 0)
  bool pxlo, pxhi, pylo, pyhi;
  __device__ FusedGeo(const Geo& g, const BC& bc)
      : n0(g.n[0]), n1(g.n[1]), sy8((unsigned)g.s[1] * 8u), base8((unsigned)g.off * 8u),
        pxlo(bc.t[0][0]), pxhi(bc.t[1][0]), pylo(bc.t[2][0]), pyhi(bc.t[3][0]) {}
  __device__ unsigned at(int x, int y) const { return base8 + (unsigned)x * 8u + (unsigned)y * sy8; }
};

// inputs of the predictor at (px, py) loaded from HBM: w at the cell and its
// two neighbours along D, B across D, and the widths
template <bool MHD>
struct PredIn {
  static constexpr int NW = Layout<2, MHD>::nw;
  double uc[NW], um[NW], up[NW];
  double bn[2][2];
  double dxk[2];
};

template <bool MHD, int D>
__device__ __forceinline__ void pred_load(const Geo& g, const FusedGeo& f,
                                          const double* __restrict__ w, int px, int py,
                                          PredIn<MHD>& in) {
  using L = Layout<2, MHD>;
  constexpr int NW = L::nw;
  const unsigned o = f.at(px, py);
  const unsigned sd = D == 0 ? 8u : f.sy8, st = D == 0 ? f.sy8 : 8u;
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    in.uc[iw] = ldp(w, g.ps, iw, o);
    in.um[iw] = ldp(w, g.ps, iw, o - sd);
    in.up[iw] = ldp(w, g.ps, iw, o + sd);
  }
  in.bn[0][0] = in.bn[0][1] = in.bn[1][0] = in.bn[1][1] = 0.0;
  if constexpr (MHD) {
    in.bn[1 - D][0] = ldp(w, g.ps, L::b(1 - D), o - st);
    in.bn[1 - D][1] = ldp(w, g.ps, L::b(1 - D), o + st);
  }
  in.dxk[0] = g.dx[0][px];
  in.dxk[1] = g.dx[1][py];
}

template <bool MHD, class R, int D, int LIM>
__device__ __forceinline__ void pred_run(const Sel& s, const PredIn<MHD>& in, double dt,
                                         double smallp, double* out, double* lout) {
  pred_compute<2, MHD, R, D, LIM, MHD ? 1 : 0>(s, in.uc, in.um, in.up, in.bn, in.dxk, dt, smallp,
                                               out, lout);
}

// predictor (and its slopes) at (px, py), every input loaded from HBM
template <bool MHD, class R, int D, int LIM>
__device__ __forceinline__ void pred_gather(const Geo& g, const FusedGeo& f, const Sel& s,
                                            const double* __restrict__ w, int px, int py, double dt,
                                            double smallp, double* out, double* lout, double* uout) {
  PredIn<MHD> in;
  pred_load<MHD, D>(g, f, w, px, py, in);
  pred_run<MHD, R, D, LIM>(s, in, dt, smallp, out, lout);
#pragma unroll
  for (int iw = 0; iw < PredIn<MHD>::NW; ++iw) uout[iw] = in.uc[iw];
}

// store out at mesh cell (c, j) and at the ghost cells taking its value
template <int NW>
__device__ __forceinline__ void store_with_ghosts(const Geo& g, const FusedGeo& f,
                                                  double* __restrict__ wout, int c, int j,
                                                  const double* out) {
  int tx[5], ty[5];
  tx[0] = c;
  ty[0] = j;
  const int nx = 1 + ghost_preimages(c, f.n0, f.pxlo, f.pxhi, tx + 1);
  const int ny = 1 + ghost_preimages(j, f.n1, f.pylo, f.pyhi, ty + 1);
  for (int a = 0; a < nx; ++a)
    for (int b = 0; b < ny; ++b) {
      const unsigned o = f.at(tx[a], ty[b]);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) stp(wout, g.ps, iw, o, out[iw]);
    }
}

// x sweep
template <bool MHD, class R, int LIM>
__device__ __forceinline__ double tvdlf_fx_body(const Geo& g, const Sel& s, const BC& bc,
                                                const double* __restrict__ w,
                                                double* __restrict__ wout, double dt,
                                                double smallp, int strip, int seg, int seglen) {
  constexpr int ND = 2;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const FusedGeo f(g, bc);
  const int lane = (int)threadIdx.x;
  int c = 1 + strip * FS_STRIDE + lane;
  const bool useful = lane >= 1 && lane <= FS_STRIDE && c <= f.n0 + 1;
  const bool face_ok = lane <= FS_STRIDE && c <= f.n0 + 1;  // face c+1/2 with c in [1, n0+1]
  c = c < f.n0 + 2 ? c : f.n0 + 2;                          // idle lanes: in-bounds, unused
  const int px = ghost_image(c, f.n0, f.pxlo, f.pxhi);
  const bool zero_slope = (c < 2 && !f.pxlo) || (c > f.n0 + 1 && !f.pxhi);
  const int s0 = 2 + seg * seglen;
  double cmx = -INFINITY;
  if (s0 > f.n1 + 1) return cmx;  // wave-uniform
  const int s1 = (s0 + seglen - 1 < f.n1 + 1) ? s0 + seglen - 1 : f.n1 + 1;
  const double* dx0 = g.dx[0];
  const double* dx1 = g.dx[1];
  double Pc[NW], Pn[NW], Lc[NW], Ln[NW], Uc[NW], Un[NW];
  double byp = 0.0;  // B_y of the predictor state one row below
  pred_gather<MHD, R, 0, LIM>(g, f, s, w, px, ghost_image(s0 - 1, f.n1, f.pylo, f.pyhi), dt,
                              smallp, Pn, Ln, Un);
  if constexpr (MHD) byp = Pn[L::b(1)];
  pred_gather<MHD, R, 0, LIM>(g, f, s, w, px, s0, dt, smallp, Pc, Lc, Uc);
  const double dxs = dx0[c] + dx0[c + 1];
  const double qdx = pdiv(dt, dx0[c]);
  // inputs of the next row's predictor are loaded one marching step ahead;
  // dy(j) comes from the previous step's inputs, so no load is consumed in
  // the step that issues it (a fresh load waited for would also wait for the
  // prefetch: loads complete in order)
  PredIn<MHD> in;
  pred_load<MHD, 0>(g, f, w, px, ghost_image(s0 + 1, f.n1, f.pylo, f.pyhi), in);
  double dyj = dx1[s0];
#pragma unroll 2
  for (int j = s0; j <= s1; ++j) {
    pred_run<MHD, R, 0, LIM>(s, in, dt, smallp, Pn, Ln);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) Un[iw] = in.uc[iw];
    const double dyn = in.dxk[1];  // dy(j + 1)
    if (j < s1) pred_load<MHD, 0>(g, f, w, px, ghost_image(j + 2, f.n1, f.pylo, f.pyhi), in);
    double lw[NW], pr[NW], lr[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      lw[iw] = zero_slope ? 0.0 : Lc[iw];
      pr[iw] = lane_next(Pc[iw]);
      lr[iw] = lane_next(lw[iw]);
    }
    double F[NW], cv;
    corr_face<ND, MHD, R, 0>(s, Pc, pr, lw, lr, dxs, smallp, F, cv);
    if (face_ok) cmx = cmx < cv ? cv : cmx;
    double out[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double Fm = lane_prev(F[iw]);
      out[iw] = Uc[iw] - qdx * (F[iw] - Fm);
    }
    if constexpr (MHD) {
      const double bxp = lane_next(Pc[L::b(0)]), bxm = lane_prev(Pc[L::b(0)]);
      double dv = 0.0;
      const double g0 = pdiv(0.5 * (bxp - bxm), dx0[c]);
      dv = dv + g0;
      const double g1 = pdiv(0.5 * (Pn[L::b(1)] - byp), dyj);
      dv = dv + g1;
      powell<ND>(Pc, dv, dt * 1.0 / ND, out);
      byp = Pc[L::b(1)];
    }
    if (useful) store_with_ghosts<NW>(g, f, wout, c, j, out);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      Pc[iw] = Pn[iw];
      Lc[iw] = Ln[iw];
      Uc[iw] = Un[iw];
    }
    dyj = dyn;
  }
  return cmx;
}

// y sweep: raw rows j, j+1, j+2 of w at column px are kept in registers, so
// a marching step loads one new row
template <bool MHD, class R, int LIM>
__device__ __forceinline__ double tvdlf_fy_body(const Geo& g, const Sel& s, const BC& bc,
                                                const double* __restrict__ w,
                                                double* __restrict__ wout, double dt,
                                                double smallp, int strip, int seg, int seglen) {
  constexpr int ND = 2;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const FusedGeo f(g, bc);
  const int lane = (int)threadIdx.x;
  int c = 1 + strip * FS_STRIDE + lane;
  const bool useful = lane >= 1 && lane <= FS_STRIDE && c <= f.n0 + 1;
  c = c < f.n0 + 2 ? c : f.n0 + 2;
  const int px = ghost_image(c, f.n0, f.pxlo, f.pxhi);
  const int s0 = 2 + seg * seglen;
  double cmx = -INFINITY;
  if (s0 > f.n1 + 1) return cmx;  // wave-uniform
  const int s1 = (s0 + seglen - 1 < f.n1 + 1) ? s0 + seglen - 1 : f.n1 + 1;
  const double* dx0 = g.dx[0];
  const double* dx1 = g.dx[1];
  const double dxp = dx0[px];
  // B_x at (px -/+ 1, r) for the predictor's divergence
  auto load_bx = [&](int r, double* bx) {
    bx[0] = bx[1] = 0.0;
    if constexpr (MHD) {
      const unsigned o = f.at(px, r);
      bx[0] = ldp(w, g.ps, L::b(0), o - 8u);
      bx[1] = ldp(w, g.ps, L::b(0), o + 8u);
    }
  };
  // predictor at mesh row r from raw rows r-1, r, r+1 and B_x in registers
  auto pred_rowx = [&](double dyr, const double* wm, const double* w0, const double* wp,
                       const double* bx, double* P, double* l) {
    const double bn[2][2] = {{bx[0], bx[1]}, {0.0, 0.0}};
    const double dxk[2] = {dxp, dyr};
    pred_compute<2, MHD, R, 1, LIM, MHD ? 1 : 0>(s, w0, wm, wp, bn, dxk, dt, smallp, P, l);
  };
  auto pred_row = [&](int r, const double* wm, const double* w0, const double* wp, double* P,
                      double* l) {
    double bx[2];
    load_bx(r, bx);
    pred_rowx(dx1[r], wm, w0, wp, bx, P, l);
  };
  // predictor at a ghost row: at its image, with that image's slopes ('periodic')
  // or zero slopes ('cont')
  auto pred_ghost = [&](int r, double* P, double* l) {
    double u[NW];
    pred_gather<MHD, R, 1, LIM>(g, f, s, w, px, ghost_image(r, f.n1, f.pylo, f.pyhi), dt, smallp,
                                P, l, u);
    const bool zs = (r < 2 && !f.pylo) || (r > f.n1 + 1 && !f.pyhi);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) l[iw] = zs ? 0.0 : l[iw];
  };
  auto load_row = [&](int r, double* u) {
    const unsigned o = f.at(px, r);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) u[iw] = ldp(w, g.ps, iw, o);
  };
  double Pc[NW], Pn[NW], lc[NW], ln[NW], Fp[NW], Fn[NW], W0[NW], W1[NW], W2[NW], cv;
  double byp = 0.0;
  // face s0-1/2 from rows s0-1 and s0
  load_row(s0 - 1, W0);
  load_row(s0, W1);
  load_row(s0 + 1, W2);
  if (s0 - 1 >= 2) {
    double Wm[NW];
    load_row(s0 - 2, Wm);
    pred_row(s0 - 1, Wm, W0, W1, Pn, ln);
  } else {
    pred_ghost(s0 - 1, Pn, ln);
  }
  pred_row(s0, W0, W1, W2, Pc, lc);
  corr_face<ND, MHD, R, 1>(s, Pn, Pc, ln, lc, dx1[s0 - 1] + dx1[s0], smallp, Fp, cv);
  if (useful) cmx = cmx < cv ? cv : cmx;
  if constexpr (MHD) byp = Pn[L::b(1)];
  // window: W0 = row j, W1 = row j+1, W2 = row j+2, bx = B_x around row j+1
  // and the widths dy0 = dy(j), dy1 = dy(j+1), all loaded one marching step
  // ahead
  double bx[2] = {0.0, 0.0}, W3[NW] = {}, bxn[2] = {0.0, 0.0};
  double dy0 = dx1[s0], dy1 = dx1[s0 + 1], dy2 = 0.0;
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) W0[iw] = W1[iw];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) W1[iw] = W2[iw];
  if (s0 + 1 <= f.n1 + 1) {
    load_row(s0 + 2, W2);
    load_bx(s0 + 1, bx);
  }
#pragma unroll 2
  for (int j = s0; j <= s1; ++j) {
    const bool ahead = j + 1 <= s1 && j + 2 <= f.n1 + 1;
    if (ahead) {
      load_row(j + 3, W3);
      load_bx(j + 2, bxn);
    }
    dy2 = dx1[j + 2];
    if (j + 1 <= f.n1 + 1) {
      pred_rowx(dy1, W0, W1, W2, bx, Pn, ln);
    } else {
      pred_ghost(j + 1, Pn, ln);
    }
    corr_face<ND, MHD, R, 1>(s, Pc, Pn, lc, ln, dy0 + dy1, smallp, Fn, cv);
    if (useful) cmx = cmx < cv ? cv : cmx;
    const double qdx = pdiv(dt, dy0);
    double out[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) out[iw] = W0[iw] - qdx * (Fn[iw] - Fp[iw]);
    if constexpr (MHD) {
      const double bxp = lane_next(Pc[L::b(0)]), bxm = lane_prev(Pc[L::b(0)]);
      double dv = 0.0;
      const double g0 = pdiv(0.5 * (bxp - bxm), dx0[c]);
      dv = dv + g0;
      const double g1 = pdiv(0.5 * (Pn[L::b(1)] - byp), dy0);
      dv = dv + g1;
      powell<ND>(Pc, dv, dt * 1.0 / ND, out);
      byp = Pc[L::b(1)];
    }
    prefetch_done();
    if (useful) store_with_ghosts<NW>(g, f, wout, c, j, out);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      Pc[iw] = Pn[iw];
      lc[iw] = ln[iw];
      Fp[iw] = Fn[iw];
      W0[iw] = W1[iw];
      W1[iw] = W2[iw];
      W2[iw] = W3[iw];
    }
    bx[0] = bxn[0];
    bx[1] = bxn[1];
    dy0 = dy1;
    dy1 = dy2;
  }
  return cmx;
}

#include "bulk32.hpp"

// Cross-block Courant maximum of a launch without a memory fence (a
// device-scope fence writes back and invalidates the whole L2 of the XCD,
// once per block).  fold[0] counts finished blocks, fold[1] holds the
// maximum as the bit pattern of a non-negative double (those order as the
// values; a block maximum that is not above zero -- -inf for a block
// without faces, -0, NaN -- enters as +0, which no use of the maximum tells
// apart: dt*cm > 1 and cm > smalldouble are both false either way).  Each
// block takes its maximum into fold[1] with an atomic max, then counts
// itself; the count's address depends on the max's returned value (which
// is never all ones), so the count is issued only after the max has been
// performed.  Device-scope atomic read-modify-writes are performed at one
// coherence point, so the block whose count completes the launch reads
// every block's max with its exchange, and resets both words for the next
// launch.  Thread 0 only; returns true in that last block, with *cm set.
__device__ __forceinline__ bool fold_courant_last(unsigned long long* fold, double bm,
                                                  unsigned long long nblocks, double* cm) {
  const unsigned long long bits = bm > 0.0 ? (unsigned long long)__double_as_longlong(bm) : 0ull;
  const unsigned long long old =
      __hip_atomic_fetch_max(fold + 1, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long n = __hip_atomic_fetch_add(fold + (old == ~0ull ? 1 : 0), 1ull,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (n != nblocks - 1) return false;
  *cm = __longlong_as_double(
      (long long)__hip_atomic_exchange(fold + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  __hip_atomic_store(fold, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

#ifndef VACGPU_FUSED_WAVES
#define VACGPU_FUSED_WAVES 2  // at most 256 VGPRs: two waves per SIMD
#endif
// blockDim = (64, 4): wave threadIdx.y takes segment 4*tile.y + y of strip
// tile.x (xcd_tile).  The last block to finish (fold_courant_last on fold[2],
// zero between launches) applies the dtcourant update of the TVDLF
// corrector (src/vactvdlf.t:184-201) when courantpar > 0 and, on the last
// sweep of a device-loop step, advances it and t.
template <bool MHD, class R, int D, int LIM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VACGPU_FUSED_WAVES))) void k_tvdlf_fused(
    Geo g, Sel s, BC bc, const double* __restrict__ w, double* __restrict__ wout,
    Scal* __restrict__ sc, int seglen, unsigned long long* __restrict__ fold, double courantpar,
    int step_end) {
  __shared__ double sm[8];
  const Tile tb = xcd_tile();
  const int seg = (int)(tb.y * blockDim.y + threadIdx.y);
  const double dt = sc->dt, smallp = sc->smallp;
  double cmx;
  if constexpr (sizeof(R) == 4)  // FP32-bulk: increments over FP64 bases
    cmx = D == 0 ? bulk_fx_body<MHD, LIM>(g, s, bc, w, wout, dt, smallp, tb.x, seg, seglen)
                 : bulk_fy_body<MHD, LIM>(g, s, bc, w, wout, dt, smallp, tb.x, seg, seglen);
  else if constexpr (D == 0)
    cmx = tvdlf_fx_body<MHD, R, LIM>(g, s, bc, w, wout, dt, smallp, tb.x, seg, seglen);
  else
    cmx = tvdlf_fy_body<MHD, R, LIM>(g, s, bc, w, wout, dt, smallp, tb.x, seg, seglen);
  const double bm = block_reduce<true>(cmx, sm);
  if (threadIdx.x != 0 || threadIdx.y != 0) return;
  double cm;
  if (!fold_courant_last(fold, bm, (unsigned long long)gridDim.x * gridDim.y, &cm)) return;
  if (courantpar > 0.0) {
    if (dt * cm > 1.0) sc->nerror[2] += 1;
    if (cm > kSmall) {
      const double cc = courantpar / cm;
      sc->dtcourant[D] = cc < sc->dtcourant[D] ? cc : sc->dtcourant[D];
    }
  }
  if (step_end) {  // the last sweep of a device-loop step: it and t advance (src/vac.t:97-98)
    sc->it += 1;
    sc->t = sc->t + dt;
  }
}

// ------------------------------------------------------------ TVD (Roe)
// First-sweep half source of the TVD pre-processing (src/vaccd.t:44-68):
// w1 = w - (dt/2)*S_Powell(w) on the mesh.
template <int ND>
__global__ __launch_bounds__(256) void k_tvd_halfsrc(Geo g, const double* __restrict__ w,
                                                    double* __restrict__ w1,
                                                    const Scal* __restrict__ sc) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double u[NW], out[NW];
  load<NW>(w, g.ps, cl.idx, u);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) out[iw] = u[iw];
  powell<ND>(u, divb_at<ND>(g, w, cl.idx, cl.c), sc->dt / 2, out);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) st_out(w1 + iw * g.ps + cl.idx, out[iw]);
}

// Sources folded into the TVD sweeps (2D, one boundary type per face for
// all variables).  SRC = 1, first sweep: the sweep reads wCT = w + (dt/2)
// S_Powell(w) computed on the fly at the mesh cell whose value the ghost rule
// copies into the cell (what k_tvd_halfsrc + getboundary produce).  SRC = 2,
// last sweep: the sweep also writes the first half of Ryu's source step,
// wt = (wold + wnew + dt S_Powell(wold))/2, at every cell it updates (what
// k_ryu_a produces).
__device__ __forceinline__ int image_of(const Geo& g, const BC& bc, int k, int c) {
  const int n = g.n[k];
  const int lo = bc.t[2 * k][0] ? c + n : 2, hi = bc.t[2 * k + 1][0] ? c - n : n + 1;
  return c < 2 ? lo : (c > n + 1 ? hi : c);
}

template <int ND>
__device__ __forceinline__ void wct_halfsrc(const Geo& g, const BC& bc, const double* __restrict__ w,
                                            const int* craw, double dt, double* u) {
  constexpr int NW = Layout<ND, true>::nw;
  int c[3] = {0, 0, 0};
  long long idx = g.off;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    c[k] = image_of(g, bc, k, craw[k]);
    idx += c[k] * g.s[k];
  }
  double uc[NW];
  load<NW>(w, g.ps, idx, uc);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) u[iw] = uc[iw];
  powell<ND>(uc, divb_at<ND>(g, w, idx, c), dt / 2, u);
}

template <int ND>
__device__ __forceinline__ void ryu_first_half(const Geo& g, const BC& bc,
                                               const double* __restrict__ wold,
                                               double* __restrict__ wt, long long idx, const int* c,
                                               double dt, const double* wnew, int ghosts) {
  constexpr int NW = Layout<ND, true>::nw;
  double uo[NW], t[NW];
  load<NW>(wold, g.ps, idx, uo);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) t[iw] = wnew[iw];
  powell<ND>(uo, divb_at<ND>(g, wold, idx, c), dt, t);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) t[iw] = 0.5 * (uo[iw] + t[iw]);
  store_cell<ND, NW>(g, bc, wt, idx, c, t, ghosts);
}

// ---- TVD sweep.  For cell i: the centred flux difference of wCT (centdiff,
// src/vaccd.t:78-136) followed by the Roe-type dissipation of tvdlimit2
// (src/vactvd.t:86-206, getphi 'roe' :244-302) through the faces i-1/2 and
// i+1/2, whose limiters also need the jumps at faces i-3/2 and i+3/2.  Face
// quantities are computed once per face: along x each lane of a wavefront
// owns one face and neighbours are read with cross-lane shuffles; along y
// and z each thread marches through a segment of cells, keeping a sliding
// window of faces in registers.  Both kernels write the partial maxima of
// |a dt 2/(dx_J+dx_J+1)| of the two fast waves over the faces they compute
// (covering ixIC, src/vactvd.t:42,145-157) to part[SLOT_C0/C1].

// getphi for 'tvd' at a face, with the entropy fix when ENT (sa: the face's smalla)
template <int LIM, class R, bool ENT>
__device__ __forceinline__ R tvd_phi_f(const Sel& s, int il, R jm, R j0, R jp, R av, const R* sa) {
  if constexpr (ENT) return tvd_phi_ph<LIM, R>(s.lim[il], jm, j0, jp, av, ent_tvd(av, sa[il], s.ent[il]));
  else return tvd_phi_t<LIM, R>(s.lim[il], jm, j0, jp, av);
}

// Everything one face contributes to the update of its two cells.
template <int ND, class R, bool MHD = true, bool ENT = false>
struct TvdFace {
  R jump[Layout<ND, MHD>::nw];
  R adt[Layout<ND, MHD>::nw];
  R sa[ENT ? Layout<ND, MHD>::nw : 1];  // smalla of the entropy fix (scaled like adt)
  typename RoeOf<ND, R, MHD>::Rt rt;
};

// PDP: the Roe solver's dp = pR - pL of useprimitive (src/vacphys.mhdroe.t:
// 42-44) fixed at compile time; the general kernels (ENT) read it from s.
template <int ND, class R, bool MHD = true, bool ENT = false, bool PDP = false>
__device__ __forceinline__ void tvd_face(const double* uL, const double* uR, int d, const Sel& s,
                                         double smallp, double dt, double dxsum, TvdFace<ND, R, MHD, ENT>& f,
                                         double& c0, double& c1) {
  constexpr int NW = Layout<ND, MHD>::nw;
  typename RoeOf<ND, R, MHD>::St r;
  roe_setup(uL, uR, d, s.gamma, smallp, r, (PDP || (ENT && s.prim)) ? 1 : 0);
  const R sdx = R(dxsum);
  [[maybe_unused]] R c2L = R(0), c2R = R(0), caL = R(0), caR = R(0);
#pragma unroll
  for (int il = 0; il < NW; ++il) {
    R a, jp;
    roe_wave(r, uL, uR, il, d, s.divbwave, a, jp);
    if constexpr (ENT) {
      // entropyfix (src/vactvd.t:422-449), smalla scaled like a (:130-141)
      const int et = s.ent[il];
      R sa = R(0);
      if (et == ENT_YEE) {
        sa = R(s.ecoef[il]);
      } else if (et != ENT_NUL) {
        R aL, aR;
        ent_speeds<ND, R, MHD>(uL, uR, il, d, s.gamma, smallp, aL, aR, c2L, c2R, caL, caR);
        if (et == ENT_HARTEN) sa = rmax(rmax(R(0), a - aL), aR - a);
        else if (et == ENT_POWELL) sa = rmax(R(0), R(2) * (aR - aL));
        else if (aL < R(0) && aR > R(0)) a = a - pdiv(R(2) * aR * aL, aR - aL);
      }
      if (et == ENT_HARTEN || et == ENT_POWELL) sa = pdiv(sa * R(dt) * R(2), sdx);
      f.sa[il] = sa;
    }
    const R ad = pdiv(a * R(dt) * R(2), sdx);
    f.jump[il] = jp;
    f.adt[il] = ad;
  }
  c0 = fabs(double(f.adt[0]));
  c1 = fabs(double(f.adt[1]));
  f.rt = roe_rt(r);
}

// out[iw] += sum over il (in order) of (R.phi(J) q(J) - R.phi(J-1) q(J-1)) / dV(i)
// nfm / nf0: the variables whose dissipative flux through face J-1 / J is
// zero ('symm0' faces)
template <int ND, class R, bool MHD = true>
__device__ __forceinline__ void tvd_dissipate(const typename RoeOf<ND, R, MHD>::Rt& rm, const R* phim,
                                              const typename RoeOf<ND, R, MHD>::Rt& r0, const R* phi0, int d,
                                              const Sel& s, R dvm, R dv0, double dvi, double* out,
                                              unsigned nfm = 0u, unsigned nf0 = 0u) {
  constexpr int NW = Layout<ND, MHD>::nw;
#pragma unroll
  for (int il = 0; il < NW; ++il) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      R a0 = roe_rtimes(rm, phim[il], iw, il, d, s.divbwave) * R(0.25) * dvm;
      R a1 = roe_rtimes(r0, phi0[il], iw, il, d, s.divbwave) * R(0.25) * dv0;
      if ((nfm >> iw) & 1u) a0 = R(0);
      if ((nf0 >> iw) & 1u) a1 = R(0);
      out[iw] = out[iw] + pdiv(double(a1) - double(a0), dvi);
    }
  }
}

// tvd_dissipate for the x sweep, where lane L owns face J and the face J-1
// term of cell J is lane L-1's face term: a0(il, iw) of this lane is a1(il, iw)
// of the previous lane (the same R.phi product, scaled by the same sum
// dvol(J) + dvol(J-1)), so it is taken from there instead of recomputed.
// nf0: the variables whose dissipative flux through this lane's face is zero
// ('symm0' faces; the next lane takes the zero as its face J-1 term)
template <int ND, class R, bool MHD = true>
__device__ __forceinline__ void tvd_dissipate_x(const typename RoeOf<ND, R, MHD>::Rt& r0, const R* phi0, const Sel& s,
                                                R dv0, double dvi, double* out, unsigned nf0 = 0u) {
  constexpr int NW = Layout<ND, MHD>::nw;
#pragma unroll
  for (int il = 0; il < NW; ++il) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      R a1 = roe_rtimes(r0, phi0[il], iw, il, 0, s.divbwave) * R(0.25) * dv0;
      if ((nf0 >> iw) & 1u) a1 = R(0);
      const R a0 = lane_prev(a1);
      out[iw] = out[iw] + pdiv(double(a1) - double(a0), dvi);
    }
  }
}

template <class T>
__device__ __forceinline__ T shfl_up1(T v) { return lane_prev(v); }
template <class T>
__device__ __forceinline__ T shfl_down1(T v) { return lane_next(v); }

template <class S, class R>
__device__ __forceinline__ S shfl_up_struct(const S& v) {
  static_assert(sizeof(S) % sizeof(R) == 0, "struct of R");
  S o;
  const R* a = reinterpret_cast<const R*>(&v);
  R* b = reinterpret_cast<R*>(&o);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(S) / sizeof(R)); ++k) b[k] = shfl_up1(a[k]);
  return o;
}

// TVD-MUSCL corrector (typefull='tvdmu', src/vactvdlf.t:92-299 with
// tvdlimit2 at :279-283, src/vactvd.t:86-206), face-parallel like the TVDLF
// corrector (k_tvdlf_corrx / k_tvdlf_corrm: the same face states from the
// predictor w1 and the limited slopes of w, typelimited='original').  A face
// carries the central flux (F(wL) + F(wR))/2 and the Roe decomposition of its
// states with phi = |a dt 2/(dx_J + dx_J+1)| jump (getphi for 'tvdmu',
// src/vactvd.t:227-240); a cell takes the flux difference, the dissipation of
// its two faces and the Powell source of w1 as in the TVDLF corrector.  The
// two extreme waves' |a dt 2/(dx_J + dx_J+1)| over the faces ixC feed the
// Courant fold (src/vactvd.t:145-157).
template <int ND, class R, int D, bool MHD = true, bool ENT = false, bool PRIM = false>
__device__ __forceinline__ void mu_face(const Sel& s, const double* u0, const double* u1,
                                        const double* ll, const double* lr, double dt,
                                        double smallp, double dxsum, double* F,
                                        TvdFace<ND, R, MHD, ENT>& f, R* phi, double& c0, double& c1) {
  constexpr int NW = Layout<ND, MHD>::nw;
  double uL[NW], uR[NW];
  muscl_states<ND, MHD, PRIM>(s, u0, u1, ll, lr, smallp, uL, uR);
  {
    double fL[NW], fR[NW];
    fluxes<ND, MHD, R>(uL, D, s.gamma, smallp, fL);
    fluxes<ND, MHD, R>(uR, D, s.gamma, smallp, fR);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) F[iw] = 0.5 * (fL[iw] + fR[iw]);
  }
  tvd_face<ND, R, MHD, ENT, PRIM>(uL, uR, D, s, smallp, dt, dxsum, f, c0, c1);
#pragma unroll
  for (int il = 0; il < NW; ++il) {
    if constexpr (ENT) phi[il] = ent_abs(f.adt[il], f.sa[il], s.ent[il]) * f.jump[il];
    else phi[il] = R(fabs(f.adt[il])) * f.jump[il];
  }
}

// x sweep: lane L owns face c+1/2 of cell c = x0 - 1 + L and updates c for
// L in [1, 62]; the left face's flux and dissipation term come from lane L-1.
template <int ND, class R, int LIM, bool MHD = true, bool ENT = false, bool PRIM = false>
__device__ __forceinline__ void tvdmu_x_body(const Geo& g, const Sel& s, const BC& bc,
                                             const double* __restrict__ w,
                                             const double* __restrict__ w1,
                                             double* __restrict__ wout, int ghosts,
                                             const double* __restrict__ wl, double dt,
                                             double smallp, Tile tb, double& c0o, double& c1o) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const int lane = (int)threadIdx.x;
  const int c = 1 + tb.x * CX_STRIDE + lane;
  const int y = 2 + (int)(tb.y * BY + threadIdx.y);
  const int z = (ND == 3) ? 2 + tb.z : 0;
  const bool rowok = y <= g.n[1] + 1;
  const bool useful = rowok && lane >= 1 && lane <= CX_STRIDE && c <= g.n[0] + 1;
  const bool face_ok = rowok && lane <= CX_STRIDE && c <= g.n[0] + 1;
  const int cc = c < g.n[0] + 2 ? c : g.n[0] + 2;  // idle lanes: in-bounds, unused
  const int yc = rowok ? y : g.n[1] + 1;
  const long long idx = g.off + cc + (long long)yc * g.s[1] + (long long)z * g.s[2];
  double l[NW], lr[NW], u0[NW], u1[NW], wc[NW];
  const bool own = wl == w;  // the limiter's base (typelimited) is the sweep's input
  if constexpr (PRIM) {
    // useprimitive: limit the primitive variables of wl after its round
    // trips, update w after its own, face states from w1 after its own
    double qa[NW];
    load<NW>(wl, g.ps, idx - 1, qa);
    load<NW>(wl, g.ps, idx, wc);
    load<NW>(wl, g.ps, idx + 1, u1);
    prim_after<ND, MHD>(qa, s.rtw, s.gamma, smallp);
    prim_after<ND, MHD>(wc, s.rtw, s.gamma, smallp);
    prim_after<ND, MHD>(u1, s.rtw, s.gamma, smallp);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) l[iw] = slope<R, LIM>(s.lim[iw], qa[iw], wc[iw], u1[iw]);
    load<NW>(w, g.ps, idx, wc);
    roundtrips<ND, MHD>(wc, s.rtb, s.gamma, smallp);
    load<NW>(w1, g.ps, idx, u0);
    load<NW>(w1, g.ps, idx + 1, u1);
    roundtrips<ND, MHD>(u0, s.rtu, s.gamma, smallp);
    roundtrips<ND, MHD>(u1, s.rtu, s.gamma, smallp);
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double* p = wl + iw * g.ps + idx;
      const double pc = p[0];
      wc[iw] = own ? pc : w[iw * g.ps + idx];
      l[iw] = slope<R, LIM>(s.lim[iw], p[-1], pc, p[1]);
      u0[iw] = w1[iw * g.ps + idx];
      u1[iw] = w1[iw * g.ps + idx + 1];
    }
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) lr[iw] = lane_next(l[iw]);
  double F[NW], c0, c1;
  TvdFace<ND, R, MHD, ENT> f;
  R phi[NW];
  mu_face<ND, R, 0, MHD, ENT, PRIM>(s, u0, u1, l, lr, dt, smallp, g.dx[0][cc] + g.dx[0][cc + 1], F, f, phi, c0, c1);
  const unsigned nff = ENT ? noflux(s, g, 0, cc) : 0u;  // 'symm0' faces (general variant)
  zero_flux<NW>(F, nff);
  if (!face_ok) { c0 = -INFINITY; c1 = -INFINITY; }
  double out[NW];
  const double qdx = pdiv(dt, g.dx[0][cc]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const double Fm = lane_prev(F[iw]);
    out[iw] = wc[iw] - qdx * (F[iw] - Fm);
  }
  // dvolume = (dx1*dx2)*dx3 evaluated left to right
  auto dvol = [&](int xi) {
    double v = g.dx[0][xi];
    v = v * g.dx[1][yc];
    if (ND == 3) v = v * g.dx[2][z];
    return v;
  };
  const double d0 = dvol(cc), dp = dvol(cc + 1);
  tvd_dissipate_x<ND, R, MHD>(f.rt, phi, s, R(dp + d0), d0, out, nff);
  const int c3[3] = {cc, yc, z};
  if constexpr (MHD) {
    if (s.sourceunsplit && s.divbfix) {
      const double dv = divb_at<ND>(g, w1, idx, c3);
      if constexpr (PRIM) {
        double us[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) us[iw] = u0[iw];
        roundtrips<ND, MHD>(us, s.rts - s.rtu, s.gamma, smallp);
        powell<ND>(us, dv, dt * double(s.srcmul) / ND, out);
      } else {
        powell<ND>(u0, dv, dt * double(s.srcmul) / ND, out);
      }
    }
  }
  if (useful) store_cell<ND, NW>(g, bc, wout, idx, c3, out, ghosts);
  c0o = c0;
  c1o = c1;
}

// Courant maxima of the TVD sweeps (tvdlimit2's two extreme waves): each
// block takes its two maxima into one of CS_N words per wave family and
// sweep direction (rows 2d and 2d + 1 of the word array, words 128 bytes
// apart, chosen by block number) with a device-scope atomic max on the bit
// pattern (non-negative doubles order as their bit patterns; a maximum not
// above zero, -inf or NaN enters as +0, which no use of the maximum tells
// apart: cm > 1 and cm > smalldouble are false either way).  A fold then
// reads 2 x CS_N words instead of one partial per block, and clears them.
constexpr int CS_N = 64, CS_PAD = 16;  // words per row, u64 per 128-byte line
__device__ __forceinline__ void cs_max(unsigned long long* __restrict__ cs, int row, double v) {
  const unsigned long long bits = v > 0.0 ? (unsigned long long)__double_as_longlong(v) : 0ull;
  __hip_atomic_fetch_max(cs + ((long long)row * CS_N + (block_linear() & (CS_N - 1))) * CS_PAD, bits,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The maxima of the rows k < nrows with bit k of rowmask set, one wave per
// row (CS_N is the wave size): wave k reads and clears its row's words, one
// per lane, and leaves the maximum in sm[k]; a one-dimensional block of at
// least nrows waves, all threads calling.
__device__ __forceinline__ void cs_fold_rows(unsigned long long* __restrict__ cs, int nrows,
                                             unsigned rowmask, double* sm) {
  static_assert(CS_N == 64, "one word per lane of a wave");
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  if (wave < nrows && ((rowmask >> wave) & 1u)) {  // wave-uniform
    double v = __longlong_as_double((long long)__hip_atomic_exchange(
        cs + ((long long)wave * CS_N + lane) * CS_PAD, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    v = wave_max(v);
    if (lane == 0) sm[wave] = v;
  }
  __syncthreads();
}

template <int ND, class R, int LIM, bool MHD = true, bool ENT = false, bool PRIM = false>
__global__ __launch_bounds__(256) void k_tvdmu_x(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                const double* __restrict__ w1,
                                                double* __restrict__ wout,
                                                const Scal* __restrict__ sc,
                                                unsigned long long* __restrict__ cs,
                                                int ghosts, const double* __restrict__ wl) {
  __shared__ double sm[8];
  double c0, c1;
  tvdmu_x_body<ND, R, LIM, MHD, ENT, PRIM>(g, s, bc, w, w1, wout, ghosts, wl, sc->dt, sc->smallp, xcd_tile(), c0, c1);
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 0, b0);
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 0 + 1, b1);
}

// y / z sweep: thread (x, segment) marches through cells s0..s1 along D,
// carrying the face below (flux, Roe decomposition, phi) and a window of w and w1.
template <int ND, class R, int D, int LIM, bool MHD = true, bool ENT = false, bool PRIM = false>
__device__ __forceinline__ void tvdmu_m_body(const Geo& g, const Sel& s, const BC& bc,
                                             const double* __restrict__ w,
                                             const double* __restrict__ w1,
                                             double* __restrict__ wout, int seglen, int ghosts,
                                             const double* __restrict__ wl, double dt,
                                             double smallp, Tile tb, double& c0o, double& c1o) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const int x = 2 + (int)(tb.x * BX + threadIdx.x);
  const int seg = (int)(tb.y * BY + threadIdx.y);
  const int s0 = 2 + seg * seglen;
  const int t = (ND == 3) ? 2 + tb.z : 0;  // the coordinate that is neither x nor D
  const int other = (D == 1) ? 2 : 1;
  const int nd_ = g.n[D];
  double c0 = -INFINITY, c1 = -INFINITY;
  if (x <= g.n[0] + 1 && s0 <= nd_ + 1) {
    const int s1 = (s0 + seglen - 1 < nd_ + 1) ? s0 + seglen - 1 : nd_ + 1;
    const long long sd = g.s[D];
    const long long base = g.off + x + (ND == 3 ? (long long)t * g.s[other] : 0);
    const double* dxd = g.dx[D];
    const double dxt = (ND == 3) ? g.dx[other][t] : 1.0;
    // volume of cell c along D: (dx1*dx2)*dx3 in dimension order
    auto dvol = [&](int cd) {
      double v = g.dx[0][x];
      if (ND == 2) return v * dxd[cd];
      if (D == 1) return (v * dxd[cd]) * dxt;
      return (v * dxt) * dxd[cd];
    };
    int c3[3] = {x, 0, 0};
    if (ND == 3) c3[other] = t;
    double a, b;
    // window: the limiter's base wl (typelimited) at J, J+1, w1 at J, slope
    // at J; the face J-1/2 below.  The update's base is w, the window itself
    // when wl is w.
    // With PRIM (useprimitive) the window holds the primitive variables of
    // wl after its round trips (s.rtw), the face states come from w1 after
    // its own (s.rtu) and the update's base is w after its own (s.rtb).
    const bool own = wl == w;
    auto lwin = [&](long long at, double* v) {
      load<NW>(wl, g.ps, at, v);
      if constexpr (PRIM) prim_after<ND, MHD>(v, s.rtw, s.gamma, smallp);
    };
    auto lw1 = [&](long long at, double* v) {
      load<NW>(w1, g.ps, at, v);
      if constexpr (PRIM) roundtrips<ND, MHD>(v, s.rtu, s.gamma, smallp);
    };
    double wb[NW], wc[NW], u0[NW], lc[NW], Fp[NW];
    TvdFace<ND, R, MHD, ENT> fp;
    R phip[NW];
    unsigned nfp = 0u;  // 'symm0' variables of the face below
    {
      double wa[NW], u1[NW], ln[NW], wd[NW];
      lwin(base + (long long)(s0 - 2) * sd, wa);
      lwin(base + (long long)(s0 - 1) * sd, wb);
      lwin(base + (long long)s0 * sd, wc);
      lw1(base + (long long)(s0 - 1) * sd, u0);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) lc[iw] = slope<R, LIM>(s.lim[iw], wa[iw], wb[iw], wc[iw]);
      // face s0-1/2
      lw1(base + (long long)s0 * sd, u1);
      lwin(base + (long long)(s0 + 1) * sd, wd);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) ln[iw] = slope<R, LIM>(s.lim[iw], wb[iw], wc[iw], wd[iw]);
      mu_face<ND, R, D, MHD, ENT, PRIM>(s, u0, u1, lc, ln, dt, smallp, dxd[s0 - 1] + dxd[s0], Fp, fp, phip, a, b);
      nfp = ENT ? noflux(s, g, D, s0 - 1) : 0u;
      zero_flux<NW>(Fp, nfp);
      c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        u0[iw] = u1[iw];
        lc[iw] = ln[iw];
        wb[iw] = wc[iw];
        wc[iw] = wd[iw];
      }
    }
    // now: u0 = w1(i), lc = slope(i), wb = w(i), wc = w(i+1)
    for (int i = s0; i <= s1; ++i) {
      double u1[NW], ln[NW], wd[NW], F[NW];
      TvdFace<ND, R, MHD, ENT> f;
      R phi[NW];
      lw1(base + (long long)(i + 1) * sd, u1);
      lwin(base + (long long)(i + 2) * sd, wd);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) ln[iw] = slope<R, LIM>(s.lim[iw], wb[iw], wc[iw], wd[iw]);
      mu_face<ND, R, D, MHD, ENT, PRIM>(s, u0, u1, lc, ln, dt, smallp, dxd[i] + dxd[i + 1], F, f, phi, a, b);
      const unsigned nff = ENT ? noflux(s, g, D, i) : 0u;
      zero_flux<NW>(F, nff);
      c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
      const long long idx = base + (long long)i * sd;
      const double qdx = pdiv(dt, dxd[i]);
      double out[NW];
      [[maybe_unused]] double ub[NW];
      if constexpr (PRIM) {
        load<NW>(w, g.ps, idx, ub);
        roundtrips<ND, MHD>(ub, s.rtb, s.gamma, smallp);
      }
#pragma unroll
      for (int iw = 0; iw < NW; ++iw)
        out[iw] = (PRIM ? ub[iw] : (own ? wb[iw] : w[iw * g.ps + idx])) - qdx * (F[iw] - Fp[iw]);
      const double dm = dvol(i - 1), d0 = dvol(i), dp = dvol(i + 1);
      tvd_dissipate<ND, R, MHD>(fp.rt, phip, f.rt, phi, D, s, R(d0 + dm), R(dp + d0), d0, out,
                                nfp, nff);
      c3[D] = i;
      if constexpr (MHD) {
        if (s.sourceunsplit && s.divbfix) {
          const double dv = divb_at<ND>(g, w1, idx, c3);
          if constexpr (PRIM) {
            double us[NW];
#pragma unroll
            for (int iw = 0; iw < NW; ++iw) us[iw] = u0[iw];
            roundtrips<ND, MHD>(us, s.rts - s.rtu, s.gamma, smallp);
            powell<ND>(us, dv, dt * double(s.srcmul) / ND, out);
          } else {
            powell<ND>(u0, dv, dt * double(s.srcmul) / ND, out);
          }
        }
      }
      store_cell<ND, NW>(g, bc, wout, idx, c3, out, ghosts);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        Fp[iw] = F[iw];
        u0[iw] = u1[iw];
        lc[iw] = ln[iw];
        wb[iw] = wc[iw];
        wc[iw] = wd[iw];
        phip[iw] = phi[iw];
      }
      fp = f;
      nfp = nff;
    }
  }
  c0o = c0;
  c1o = c1;
}

// (the plain variant is held to 256 VGPRs, two waves per SIMD, at the cost
// of a few spills; the entropy-fix and primitive variants would spill far
// more and keep one wave)
template <int ND, class R, int D, int LIM, bool MHD = true, bool ENT = false, bool PRIM = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((ENT || PRIM) ? 1 : 2))) void k_tvdmu_m(Geo g, Sel s, BC bc, const double* __restrict__ w,
                                                const double* __restrict__ w1,
                                                double* __restrict__ wout,
                                                const Scal* __restrict__ sc,
                                                unsigned long long* __restrict__ cs,
                                                int seglen, int ghosts,
                                                const double* __restrict__ wl) {
  __shared__ double sm[8];
  double c0, c1;
  tvdmu_m_body<ND, R, D, LIM, MHD, ENT, PRIM>(g, s, bc, w, w1, wout, seglen, ghosts, wl, sc->dt, sc->smallp, xcd_tile(), c0, c1);
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 2 * D, b0);
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 2 * D + 1, b1);
}

constexpr int TX_STRIDE = 61;  // cells updated per 64-lane wavefront along x

// x sweep: lane L of a wavefront owns face J = x0 - 2 + L (between cells J and
// J+1) and updates cell J for L in [2, 62].

template <int ND, class R, int LIM = -1, int SRC = 0, bool MHD = true, bool ENT = false>
__global__ __launch_bounds__(256) void k_tvd_x(Geo g, Sel s, BC bc, const double* __restrict__ wct,
                                              const double* __restrict__ win,
                                              double* __restrict__ wout,
                                              const Scal* __restrict__ sc,
                                              unsigned long long* __restrict__ cs,
                                              const double* __restrict__ wold,
                                              double* __restrict__ wt, int ghosts) {
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const int lane = threadIdx.x;
  const Tile tb = xcd_tile();
  const int x0 = 2 + tb.x * TX_STRIDE;
  const int J = x0 - 2 + lane;
  const int y = 2 + (int)(tb.y * BY + threadIdx.y);
  const int z = (g.nd == 3) ? 2 + tb.z : 0;
  const bool rowok = y <= g.n[1] + 1;
  const int yc = rowok ? y : g.n[1] + 1;
  const int Jc = J < g.n[0] + 2 ? J : g.n[0] + 2;
  const long long base = g.off + (long long)yc * g.s[1] + (long long)z * g.s[2];
  double uL[NW], uR[NW];
  if constexpr (SRC == 1) {
    const int cl[3] = {Jc, yc, z}, cr[3] = {Jc + 1, yc, z};
    wct_halfsrc<ND>(g, bc, win, cl, dt, uL);
    wct_halfsrc<ND>(g, bc, win, cr, dt, uR);
  } else {
    load<NW>(wct, g.ps, base + Jc, uL);
    load<NW>(wct, g.ps, base + Jc + 1, uR);
  }
  TvdFace<ND, R, MHD, ENT> f;
  double c0, c1;
  tvd_face<ND, R, MHD, ENT>(uL, uR, 0, s, smallp, dt, g.dx[0][Jc] + g.dx[0][Jc + 1], f, c0, c1);
  const bool faceok = rowok && J <= g.n[0] + 2;
  if (!faceok) { c0 = -INFINITY; c1 = -INFINITY; }
  // centdiff: flux at cell J, face average fC(J) and fC(J-1)
  double F[NW], fC[NW], fCm[NW];
  fluxes<ND, MHD, R>(uL, 0, s.gamma, smallp, F);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) fC[iw] = 0.5 * (F[iw] + shfl_down1(F[iw]));
  // 'symm0' faces (general variant): this lane's face J
  const unsigned nfJ = ENT ? noflux(s, g, 0, J) : 0u;
  zero_flux<NW>(fC, nfJ);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) fCm[iw] = shfl_up1(fC[iw]);
  // phi at face J from the jumps at J-1, J, J+1
  R phi[NW];
#pragma unroll
  for (int il = 0; il < NW; ++il) {
    const R jm = shfl_up1(f.jump[il]), jp = shfl_down1(f.jump[il]);
    if constexpr (ENT)
      phi[il] = tvd_phi_ph<LIM, R>(s.lim[il], jm, f.jump[il], jp, f.adt[il],
                                   ent_tvd(f.adt[il], f.sa[il], s.ent[il]));
    else
      phi[il] = tvd_phi_t<LIM, R>(s.lim[il], jm, f.jump[il], jp, f.adt[il]);
  }
  const int i = J;
  // the dissipation needs every lane's face term (lane_prev inside), so it is
  // evaluated by all lanes and only the owners of a cell store it
  double out[NW];
  {
    // lanes 1..61 feed the next lane exactly (cells 1 .. n+1 as they are);
    // the rest are clamped into ixG and their values go nowhere
    const int ic = i < g.n[0] + 2 ? (i > 1 ? i : 1) : g.n[0] + 2;
    const long long idx = base + ic;
    const double qdx = pdiv(dt, g.dx[0][ic]);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) out[iw] = win[iw * g.ps + idx] - qdx * (fC[iw] - fCm[iw]);
    // dvolume = (dx1*dx2)*dx3 evaluated left to right
    auto dvol = [&](int xi) {
      double v = g.dx[0][xi];
      v = v * g.dx[1][yc];
      if (ND == 3) v = v * g.dx[2][z];
      return v;
    };
    const double d0 = dvol(ic), dp = dvol(ic + 1);
    tvd_dissipate_x<ND, R, MHD>(f.rt, phi, s, R(dp + d0), d0, out, nfJ);
  }
  if (rowok && lane >= 2 && lane <= 62 && i <= g.n[0] + 1) {
    const long long idx = base + i;
    const int c[3] = {i, yc, z};
    // (the last sweep's output is final only after the Ryu step, which fills its ghosts)
    store_cell<ND, NW>(g, bc, wout, idx, c, out, SRC == 2 ? 0 : ghosts);
    if constexpr (SRC == 2) ryu_first_half<ND>(g, bc, wold, wt, idx, c, dt, out, ghosts);
  }
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 0, b0);
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 0 + 1, b1);
}

// y / z sweep: thread (x, segment) marches through cells s0..s1 along d.
template <int ND, class R, int D, int LIM = -1, int SRC = 0, bool MHD = true, bool ENT = false>
__global__ __launch_bounds__(256) void k_tvd_m(Geo g, Sel s, BC bc, const double* __restrict__ wct,
                                              const double* __restrict__ win,
                                              double* __restrict__ wout,
                                              const Scal* __restrict__ sc,
                                              unsigned long long* __restrict__ cs,
                                              int seglen, const double* __restrict__ wold,
                                              double* __restrict__ wt, int ghosts) {
  // (ghost cells of the output are left to a boundary pass: writing them here
  // costs the marching kernel occupancy; the Ryu half is folded into x sweeps only)
  static_assert(SRC != 2, "the Ryu half is folded into x sweeps only");
  (void)ghosts;
  constexpr int d = D;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const Tile tb = xcd_tile();
  const int x = 2 + (int)(tb.x * BX + threadIdx.x);
  const int seg = (int)(tb.y * BY + threadIdx.y);
  const int s0 = 2 + seg * seglen;
  // the transverse coordinate that is neither x nor d (3D only)
  const int t = (g.nd == 3) ? 2 + tb.z : 0;
  const int nd_ = g.n[d];
  const bool active = x <= g.n[0] + 1 && s0 <= nd_ + 1;
  double c0 = -INFINITY, c1 = -INFINITY;
  if (active) {
    const int s1 = (s0 + seglen - 1 < nd_ + 1) ? s0 + seglen - 1 : nd_ + 1;
    const int other = (d == 1) ? 2 : 1;  // dimension of t in 3D
    const long long sd = g.s[d];
    const long long base = g.off + x + (ND == 3 ? (long long)t * g.s[other] : 0);
    const double* dxd = g.dx[d];
    // volume of cell c along d: (dx1*dx2)*dx3 in dimension order
    const double dxt = (ND == 3) ? g.dx[other][t] : 1.0;
    auto dvol = [&](int c) {
      double v = g.dx[0][x];
      if (ND == 2) return v * dxd[c];
      // d == 1: (dx0*dx1[c])*dx2[t];  d == 2: (dx0*dx1[t])*dx2[c]
      if (d == 1) return (v * dxd[c]) * dxt;
      return (v * dxt) * dxd[c];
    };
    // wCT at cell c along d (through the half source on the first sweep)
    auto get_ct = [&](int c, double* u) {
      if constexpr (SRC == 1) {
        int cc[3] = {x, 0, 0};
        cc[d] = c;
        if (ND == 3) cc[other] = t;
        wct_halfsrc<ND>(g, bc, win, cc, dt, u);
      } else {
        load<NW>(wct, g.ps, base + (long long)c * sd, u);
      }
    };
    double uL[NW], uR[NW];
    TvdFace<ND, R, MHD, ENT> fm, f0, fp;  // faces i-1, i, i+1
    R phim[NW], phi0[NW];
    double Fm[NW], F0[NW], fCm[NW];
    double a, b;
    // prologue: faces s0-2, s0-1, s0; centred fluxes at cells s0-1, s0
    get_ct(s0 - 2, uL);
    get_ct(s0 - 1, uR);
    tvd_face<ND, R, MHD, ENT>(uL, uR, d, s, smallp, dt, dxd[s0 - 2] + dxd[s0 - 1], fm, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    fluxes<ND, MHD, R>(uR, d, s.gamma, smallp, Fm);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
    get_ct(s0, uR);
    tvd_face<ND, R, MHD, ENT>(uL, uR, d, s, smallp, dt, dxd[s0 - 1] + dxd[s0], f0, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    fluxes<ND, MHD, R>(uR, d, s.gamma, smallp, F0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) fCm[iw] = 0.5 * (Fm[iw] + F0[iw]);  // face s0-1
    // 'symm0' faces (general variant): of the face below and of face i
    unsigned nfm = ENT ? noflux(s, g, d, s0 - 1) : 0u;
    zero_flux<NW>(fCm, nfm);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
    get_ct(s0 + 1, uR);
    tvd_face<ND, R, MHD, ENT>(uL, uR, d, s, smallp, dt, dxd[s0] + dxd[s0 + 1], fp, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    // phi at face s0-1 from faces s0-2, s0-1, s0
#pragma unroll
    for (int il = 0; il < NW; ++il) phim[il] = tvd_phi_f<LIM, R, ENT>(s, il, fm.jump[il], f0.jump[il], fp.jump[il], f0.adt[il], f0.sa);
    // window now: fm <- face s0-1, f0 <- face s0 (fp holds face s0)
    fm = f0;
    f0 = fp;
    for (int i = s0;; ++i) {
      // f0 = face i, fm = face i-1 (with phim), uL = w(i), uR = w(i+1)
      fluxes<ND, MHD, R>(uR, d, s.gamma, smallp, Fm);  // F(i+1), reuse Fm as scratch
      double fC[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) fC[iw] = 0.5 * (F0[iw] + Fm[iw]);  // face i
      const unsigned nf0 = ENT ? noflux(s, g, d, i) : 0u;
      zero_flux<NW>(fC, nf0);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) F0[iw] = Fm[iw];
      // face i+1 from w(i+1), w(i+2)
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
      get_ct(i + 2, uR);
      tvd_face<ND, R, MHD, ENT>(uL, uR, d, s, smallp, dt, dxd[i + 1] + dxd[i + 2], fp, a, b);
      c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
#pragma unroll
      for (int il = 0; il < NW; ++il) phi0[il] = tvd_phi_f<LIM, R, ENT>(s, il, fm.jump[il], f0.jump[il], fp.jump[il], f0.adt[il], f0.sa);
      // update cell i
      const long long idx = base + (long long)i * sd;
      double out[NW];
      const double qdx = pdiv(dt, dxd[i]);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) out[iw] = win[iw * g.ps + idx] - qdx * (fC[iw] - fCm[iw]);
      const double dm = dvol(i - 1), d0 = dvol(i), dp = dvol(i + 1);
      tvd_dissipate<ND, R, MHD>(fm.rt, phim, f0.rt, phi0, d, s, R(d0 + dm), R(dp + d0), d0, out,
                                nfm, nf0);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) st_out(wout + iw * g.ps + idx, out[iw]);
      if (i == s1) break;
      // shift the window
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) fCm[iw] = fC[iw];
      nfm = nf0;
#pragma unroll
      for (int il = 0; il < NW; ++il) phim[il] = phi0[il];
      fm = f0;
      f0 = fp;
    }
  }
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 2 * D, b0);
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs_max(cs, 2 * D + 1, b1);
}

// k_ryu_ab on 2D grids as a march along y: lanes cover a strip of 62 x
// cells (lanes 0 and 63 are the neighbours on either side), each marching
// step evaluates wt once at one new row (at the image of a ghost row or
// column), keeps B_y of the rows above and below in registers and takes B_x
// of the x neighbours from the next / previous lane.  Per cell: wold and wnew
// read once, the result written once.  Same arithmetic as k_ryu_ab.
constexpr int RY_STRIDE = 62;
__global__ __launch_bounds__(256) void k_ryu_ab_y(Geo g, BC bc, const double* __restrict__ wold,
                                                 const double* __restrict__ wnew,
                                                 double* __restrict__ wout,
                                                 const Scal* __restrict__ sc, int seglen) {
  using L = Layout<2, true>;
  constexpr int NW = L::nw;
  const int lane = (int)threadIdx.x;
  const int n0 = g.n[0], n1 = g.n[1];
  const Tile tb = xcd_tile();
  int c = 1 + tb.x * RY_STRIDE + lane;
  const bool useful = lane >= 1 && lane <= RY_STRIDE && c <= n0 + 1;
  c = c < n0 + 2 ? c : n0 + 2;  // idle lanes: in-bounds, unused
  const int px = image_of(g, bc, 0, c);
  const int s0 = 2 + (tb.y * (int)blockDim.y + (int)threadIdx.y) * seglen;
  if (s0 > n1 + 1) return;  // wave-uniform
  const int s1 = (s0 + seglen - 1 < n1 + 1) ? s0 + seglen - 1 : n1 + 1;
  const double dt = sc->dt;
  const long long sy = g.s[1];
  const double dxp = g.dx[0][px];
  // inputs of wt at (px, image of row r), loaded one marching step ahead
  struct In {
    double uo[NW], un[NW], bxm, bxp, bym, byp, dyp;
  } in;
  auto load_row = [&](int r) {
    const int py = image_of(g, bc, 1, r);
    const long long k = g.off + px + (long long)py * sy;
    load<NW>(wold, g.ps, k, in.uo);
    load<NW>(wnew, g.ps, k, in.un);
    in.bxm = wold[L::b(0) * g.ps + k - 1];
    in.bxp = wold[L::b(0) * g.ps + k + 1];
    in.bym = wold[L::b(1) * g.ps + k - sy];
    in.byp = wold[L::b(1) * g.ps + k + sy];
    in.dyp = g.dx[1][py];
  };
  // wt = (wold + wnew - dt S(wold))/2 and wnew there (divb_at's arithmetic)
  auto wt_of = [&](double* t, double* un) {
    double dv = 0.0;
    const double g0 = pdiv(0.5 * (in.bxp - in.bxm), dxp);
    dv = dv + g0;
    const double g1 = pdiv(0.5 * (in.byp - in.bym), in.dyp);
    dv = dv + g1;
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) t[iw] = in.un[iw];
    powell<2>(in.uo, dv, dt, t);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      t[iw] = 0.5 * (in.uo[iw] + t[iw]);
      un[iw] = in.un[iw];
    }
  };
  double tc[NW], uc[NW], tn[NW], un[NW];
  double bym;  // B_y of wt one row below
  load_row(s0 - 1);
  wt_of(tn, un);
  bym = tn[L::b(1)];
  load_row(s0);
  wt_of(tc, uc);
  load_row(s0 + 1);
  const double dxc = g.dx[0][c];
  for (int j = s0; j <= s1; ++j) {
    wt_of(tn, un);  // row j + 1
    if (j < s1) load_row(j + 2);
    const double bxp = lane_next(tc[L::b(0)]), bxm = lane_prev(tc[L::b(0)]);
    double dv = 0.0;
    const double g0 = pdiv(0.5 * (bxp - bxm), dxc);
    dv = dv + g0;
    const double g1 = pdiv(0.5 * (tn[L::b(1)] - bym), g.dx[1][j]);
    dv = dv + g1;
    double o[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) o[iw] = uc[iw];
    powell<2>(tc, dv, dt, o);
    if (useful) {
      const int c3[3] = {c, j, 0};
      store_cell<2, NW, double, false>(g, bc, wout, g.off + c + (long long)j * sy, c3, o, 1);
    }
    bym = tc[L::b(1)];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      tc[iw] = tn[iw];
      uc[iw] = un[iw];
    }
  }
}

// Ryu source step on the last sweep (src/vactvd.t:60-80):
// wt = (wold + wnew - dt*S(wold))/2 on the mesh (ghosts refreshed after) ...
template <int ND>
__global__ __launch_bounds__(256) void k_ryu_a(Geo g, BC bc, const double* __restrict__ wold,
                                              const double* __restrict__ wnew,
                                              double* __restrict__ wt,
                                              const Scal* __restrict__ sc, int ghosts) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double un[NW];
  load<NW>(wnew, g.ps, cl.idx, un);
  ryu_first_half<ND>(g, bc, wold, wt, cl.idx, cl.c, sc->dt, un, ghosts);
}

// ... then wnew = wnew - dt*S(wt).
// Both halves of Ryu's source step in one pass, for a last sweep that did
// not fold the first half (one boundary type per face for all variables):
// wnew - dt S(wt) with wt = (wold + wnew - dt S(wold))/2 evaluated where it
// is needed, at the cell and, for the B component that divB(wt) takes, at
// the images of its neighbours (the ghost values getboundary would give wt),
// with the arithmetic of k_ryu_a and k_ryu_b.  wnew is read only; the result
// goes to wout (a buffer the step no longer needs) with its ghost cells.
template <int ND>
__device__ __forceinline__ long long cell_index(const Geo& g, const int* c) {
  return g.off + c[0] + (long long)c[1] * g.s[1] + (ND == 3 ? (long long)c[2] * g.s[2] : 0);
}
template <int ND>
__global__ __launch_bounds__(256) void k_ryu_ab(Geo g, BC bc, const double* __restrict__ wold,
                                               const double* __restrict__ wnew,
                                               double* __restrict__ wout,
                                               const Scal* __restrict__ sc) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  const double dt = sc->dt;
  double ut[NW], o[NW];
  {
    double uo[NW];
    load<NW>(wold, g.ps, cl.idx, uo);
    load<NW>(wnew, g.ps, cl.idx, ut);
    powell<ND>(uo, divb_at<ND>(g, wold, cl.idx, cl.c), dt, ut);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) ut[iw] = 0.5 * (uo[iw] + ut[iw]);
  }
  // B_d of wt at a mesh cell: the b(d) row of the block above
  auto wt_b = [&](int d, const int* c) {
    const long long k = cell_index<ND>(g, c);
    const double divb = dt * divb_at<ND>(g, wold, k, c);
    const double t = wnew[L::b(d) * g.ps + k] - pdiv(wold[L::m(d) * g.ps + k], wold[k]) * divb;
    return 0.5 * (wold[L::b(d) * g.ps + k] + t);
  };
  double dv = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    int cm[3] = {cl.c[0], cl.c[1], cl.c[2]}, cp[3] = {cl.c[0], cl.c[1], cl.c[2]};
    cm[d] = image_of(g, bc, d, cl.c[d] - 1);
    cp[d] = image_of(g, bc, d, cl.c[d] + 1);
    const double gq = pdiv(0.5 * (wt_b(d, cp) - wt_b(d, cm)), g.dx[d][cl.c[d]]);
    dv = dv + gq;
  }
  load<NW>(wnew, g.ps, cl.idx, o);
  powell<ND>(ut, dv, dt, o);
  store_cell<ND, NW>(g, bc, wout, cl.idx, cl.c, o, 1);
}

template <int ND>
__global__ __launch_bounds__(256) void k_ryu_b(Geo g, BC bc, const double* __restrict__ wt,
                                              double* __restrict__ wnew,
                                              const Scal* __restrict__ sc, int ghosts) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double ut[NW], o[NW];
  load<NW>(wt, g.ps, cl.idx, ut);
  load<NW>(wnew, g.ps, cl.idx, o);
  powell<ND>(ut, divb_at<ND>(g, wt, cl.idx, cl.c), sc->dt, o);
  store_cell<ND, NW>(g, bc, wnew, cl.idx, cl.c, o, ghosts);
}

// ------------------------------------------------------------ time step
// Partial min / max of the unclipped pressure over the mesh, for the
// pressure floor set on the first getpthermal call (src/vacphys0.t.mhd:202-214).
template <int ND, bool MHD>
__device__ __forceinline__ void pminmax_body(const Geo& g, const double* __restrict__ w,
                                             double gamma, int bx, int by, int bz, double& mn,
                                             double& mx) {
  using L = Layout<ND, MHD>;
  const CellIx cl = cell_at(g, bx, by, bz);
  mn = INFINITY;
  mx = -INFINITY;
  if (cl.ok) {
    double u[L::nw];
    load<L::nw>(w, g.ps, cl.idx, u);
    mn = mx = pthermal_raw<ND, MHD>(u, gamma);
  }
}

template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_pminmax(Geo g, const double* __restrict__ w, double gamma,
                                                const Scal* __restrict__ sc,
                                                double* __restrict__ part, long long pstride) {
  __shared__ double sm[8];
  if (sc->smallp >= 0.0) return;  // uniform: already set
  double mn, mx;
  pminmax_body<ND, MHD>(g, w, gamma, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, mn, mx);
  const double bmn = block_reduce<false>(mn, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMIN * pstride + block_linear()] = bmn;
  const double bmx = block_reduce<true>(mx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMAX * pstride + block_linear()] = bmx;
}

// cmax/dx at cell centres for dimension d (src/vac.t:697-701)
template <int ND, bool MHD>
__device__ __forceinline__ double cmaxdx_body(const Geo& g, const double* __restrict__ w,
                                              double gamma, double smallp, int d, int bx, int by,
                                              int bz) {
  using L = Layout<ND, MHD>;
  const CellIx cl = cell_at(g, bx, by, bz);
  double v = -INFINITY;
  if (cl.ok) {
    double u[L::nw];
    load<L::nw>(w, g.ps, cl.idx, u);
    v = pdiv(cmax<ND, MHD, double>(u, d, gamma, smallp), g.dx[d][cl.c[d]]);
  }
  return v;
}

// Partial max of cmax/dx for dimension d, skipped when a lagged dtcourant(d)
// exists.
template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_cmaxdx(Geo g, const double* __restrict__ w, double gamma,
                                               const Scal* __restrict__ sc,
                                               double* __restrict__ part, long long pstride, int d) {
  __shared__ double sm[8];
  if (sc->dtcourant[d] < kBig) return;  // uniform
  const double v = cmaxdx_body<ND, MHD>(g, w, gamma, sc->smallp, d, (int)blockIdx.x,
                                        (int)blockIdx.y, (int)blockIdx.z);
  const double b = block_reduce<true>(v, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[(SLOT_DX + d) * pstride + block_linear()] = b;
}

// k_cmaxdx for every dimension without a lagged dtcourant in one pass over
// w (the first getdt of a run): the same per-cell values and block partials
template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_cmaxdx_all(Geo g, const double* __restrict__ w, double gamma,
                                                   const Scal* __restrict__ sc,
                                                   double* __restrict__ part, long long pstride) {
  using L = Layout<ND, MHD>;
  __shared__ double sm[8];
  const CellIx cl = cell_at(g, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
  double u[L::nw];
  if (cl.ok) load<L::nw>(w, g.ps, cl.idx, u);
  const double smallp = sc->smallp;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    if (sc->dtcourant[d] < kBig) continue;  // uniform
    const double v = cl.ok ? pdiv(cmax<ND, MHD, double>(u, d, gamma, smallp), g.dx[d][cl.c[d]])
                           : -INFINITY;
    const double b = block_reduce<true>(v, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[(SLOT_DX + d) * pstride + block_linear()] = b;
  }
}

// Block-wide max / min of p[0..n).  Each thread keeps four running values
// and so four loads in flight (a one-block fold is bound by load latency);
// max and min do not depend on the order of the values.
__device__ double fold(const double* p, long long n, bool is_max, double* sm) {
  auto pick = [is_max](double v, double u) { return is_max ? (v < u ? u : v) : (u < v ? u : v); };
  const double id = is_max ? -INFINITY : INFINITY;
  double v0 = id, v1 = id, v2 = id, v3 = id;
  const long long st = blockDim.x;
  long long i = threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    const double a = p[i], b = p[i + st], c = p[i + 2 * st], e = p[i + 3 * st];
    v0 = pick(v0, a);
    v1 = pick(v1, b);
    v2 = pick(v2, c);
    v3 = pick(v3, e);
  }
  for (; i < n; i += st) v0 = pick(v0, p[i]);
  const double v = pick(pick(v0, v1), pick(v2, v3));
  return is_max ? block_reduce<true>(v, sm) : block_reduce<false>(v, sm);
}

__global__ void k_fin_smallp(Scal* sc, const double* part, long long pstride, long long nb,
                             double coeff, int mhd) {
  __shared__ double sm[16];
  if (sc->smallp >= 0.0) return;
  const double mn = fold(part + SLOT_PMIN * pstride, nb, false, sm);
  __syncthreads();
  const double mx = fold(part + SLOT_PMAX * pstride, nb, true, sm);
  if (threadIdx.x == 0) {
    double sp = mn * coeff;
    if (mhd && sp < 0.0) sp = mx * coeff;
    sc->smallp = sp;
  }
}

// getdt_courant's fold (src/vac.t:690-719) and the main loop's dtcantgrow,
// slowsteps and tmaxexact (src/vac.t:68-83); dtpar > 0 replaces the
// Courant step.
// thread 0: dt from the lagged dtcourant and, where it is missing, the
// cell-centre maxima cmd[] (src/vac.t:690-719), then the tmaxexact clip
__device__ __forceinline__ void getdt_finish(Scal* sc, const double* dtc, const double* cmd, int nd,
                                             double courantpar, double dtpar, int tmaxexact) {
  {
    double dt;
    if (dtpar > 0.0) {
      dt = dtpar;
    } else {
      dt = kBig;
      double cm = 0.0;
      for (int d = 0; d < nd; ++d) {
        if (dtc[d] < kBig) dt = dtc[d] < dt ? dtc[d] : dt;
        else cm = cm < cmd[d] ? cmd[d] : cm;
      }
      for (int d = 0; d < nd; ++d) sc->dtcourant[d] = kBig;
      if (cm > kSmall) { double c = courantpar / cm; dt = c < dt ? c : dt; }
      // dtcantgrow (src/vac.t:68-69)
      if (sc->dtcantgrow && sc->it > sc->itmin) dt = sc->dtold < dt ? sc->dtold : dt;
      sc->dtold = dt;
    }
    sc->dtchk = dt;
    // slowsteps > 1 (src/vac.t:78)
    if (sc->slowsteps > sc->it - sc->itmin + 1) {
      const double q = ((double)(sc->it - sc->itmin) + 1.0) / (double)sc->slowsteps;
      const double a = 1.0 - q;
      dt = dt * (1.0 - a * a);
    }
    if (tmaxexact) {
      double c = sc->tmax - sc->t + kSmall;
      dt = c < dt ? c : dt;
    }
    sc->dt = dt;
  }
}

__global__ void k_fin_getdt(Scal* sc, const double* part, long long pstride, long long nb, int nd,
                            double courantpar, double dtpar, int tmaxexact) {
  __shared__ double sm[16];
  double dtc[3] = {kBig, kBig, kBig};
  for (int d = 0; d < nd; ++d) dtc[d] = sc->dtcourant[d];
  double cmd[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < nd; ++d) {
    if (dtc[d] >= kBig && dtpar <= 0.0) {
      cmd[d] = fold(part + (SLOT_DX + d) * pstride, nb, true, sm);
      __syncthreads();
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) getdt_finish(sc, dtc, cmd, nd, courantpar, dtpar, tmaxexact);
}

// getdt_courant in one block for the steps after the first of a device loop,
// where the pressure floor is set and every dimension has a lagged
// dtcourant: then this is k_fin_getdt without partials.  Should either be
// missing (a pressure floor that stays negative, a sweep whose Courant
// maximum was below smalldouble), the block computes it itself over the mesh
// -- slow, but exactly what the multi-block passes compute.
// tvdlimit2's dtcourant update (src/vactvd.t:145-157) from one Courant
// maximum of a TVD sweep in direction d (thread 0)
__device__ __forceinline__ void tvd_courant_apply(Scal* sc, int d, double courantpar, double cm) {
  if (cm > 1.0) sc->nerror[2] += 1;
  if (cm > kSmall) {
    double c = sc->dt * courantpar / cm;
    sc->dtcourant[d] = c < sc->dtcourant[d] ? c : sc->dtcourant[d];
  }
}

// With pend_mask set, the TVD sweeps of the previous device-loop step left
// their Courant maxima in the words cs (directions in pend_mask): they are
// folded first, as k_fin_courant_tvd_cs would have after each sweep, and
// with pend_step_end the previous step's it and t advance; nothing between
// those sweeps and this launch reads dtcourant, nerror, it or t.
template <int ND, bool MHD>
__global__ __launch_bounds__(1024) void k_getdt_lagged(Geo g, const double* __restrict__ w, double gamma,
                                                      Scal* sc, double smallpcoeff, double courantpar,
                                                      double dtpar, int tmaxexact,
                                                      unsigned long long* cs, int pend_mask,
                                                      int pend_step_end) {
  using L = Layout<ND, MHD>;
  __shared__ double sm[16];
  if (pend_mask) {  // uniform
    unsigned rows = 0u;
    for (int d = 0; d < ND; ++d)
      if ((pend_mask >> d) & 1) rows |= 3u << (2 * d);
    cs_fold_rows(cs, 2 * ND, rows, sm);
    if (threadIdx.x == 0) {
      for (int d = 0; d < ND; ++d) {
        if (!((pend_mask >> d) & 1)) continue;
        tvd_courant_apply(sc, d, courantpar, sm[2 * d]);
        tvd_courant_apply(sc, d, courantpar, sm[2 * d + 1]);
      }
    }
    if (threadIdx.x == 0 && pend_step_end) {
      sc->it += 1;
      sc->t = sc->t + sc->dt;
    }
    __syncthreads();
  }
  const long long ncell = (long long)g.n[0] * g.n[1] * (ND == 3 ? g.n[2] : 1);
  auto cell = [&](long long t, int* c) {
    c[0] = 2 + (int)(t % g.n[0]);
    c[1] = 2 + (int)((t / g.n[0]) % g.n[1]);
    c[2] = ND == 3 ? 2 + (int)(t / ((long long)g.n[0] * g.n[1])) : 0;
    return g.off + c[0] + c[1] * g.s[1] + c[2] * g.s[2];
  };
  if (sc->smallp < 0.0) {  // block-uniform
    double mn = INFINITY, mx = -INFINITY;
    for (long long t = threadIdx.x; t < ncell; t += blockDim.x) {
      int c[3];
      double u[L::nw];
      load<L::nw>(w, g.ps, cell(t, c), u);
      const double p = pthermal_raw<ND, MHD>(u, gamma);
      mn = p < mn ? p : mn;
      mx = mx < p ? p : mx;
    }
    mn = block_reduce<false>(mn, sm);
    __syncthreads();
    mx = block_reduce<true>(mx, sm);
    if (threadIdx.x == 0) {
      double sp = mn * smallpcoeff;
      if (MHD && sp < 0.0) sp = mx * smallpcoeff;
      sc->smallp = sp;
    }
    __syncthreads();
  }
  const double smallp = sc->smallp;
  double dtc[3] = {kBig, kBig, kBig};
  for (int d = 0; d < ND; ++d) dtc[d] = sc->dtcourant[d];
  double cmd[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ND; ++d) {
    if (dtc[d] >= kBig && dtpar <= 0.0) {  // block-uniform
      double v = -INFINITY;
      for (long long t = threadIdx.x; t < ncell; t += blockDim.x) {
        int c[3];
        double u[L::nw];
        load<L::nw>(w, g.ps, cell(t, c), u);
        const double q = pdiv(cmax<ND, MHD, double>(u, d, gamma, smallp), g.dx[d][c[d]]);
        v = v < q ? q : v;
      }
      cmd[d] = block_reduce<true>(v, sm);
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) getdt_finish(sc, dtc, cmd, ND, courantpar, dtpar, tmaxexact);
}

// dtcourant update of the TVDLF corrector (src/vactvdlf.t:184-201)
// step_end: the last sweep of a device-loop step also advances it and t
// (src/vac.t:97-98); what follows in the step reads dt only
__global__ void k_fin_courant_tvdlf(Scal* sc, const double* part, long long pstride, long long nb,
                                    int d, double courantpar, int step_end) {
  __shared__ double sm[16];
  const double cm = fold(part + SLOT_C0 * pstride, nb, true, sm);
  if (threadIdx.x == 0) {
    if (sc->dt * cm > 1.0) sc->nerror[2] += 1;
    if (cm > kSmall) {
      double c = courantpar / cm;
      sc->dtcourant[d] = c < sc->dtcourant[d] ? c : sc->dtcourant[d];
    }
    if (step_end) {
      sc->it += 1;
      sc->t = sc->t + sc->dt;
    }
  }
}


// k_fin_courant_tvd for a TVD sweep's Courant words (cs_max): both extreme
// waves of direction d in VAC's order, and on the last sweep of a
// device-loop step the step's it and t
__global__ __launch_bounds__(384) void k_fin_courant_tvd_cs(Scal* sc, unsigned long long* cs, int d,
                                                           double courantpar, int step_end) {
  __shared__ double sm[6];
  cs_fold_rows(cs, 6, 3u << (2 * d), sm);
  if (threadIdx.x == 0) {
    tvd_courant_apply(sc, d, courantpar, sm[2 * d]);
    tvd_courant_apply(sc, d, courantpar, sm[2 * d + 1]);
    if (step_end) {
      sc->it += 1;
      sc->t = sc->t + sc->dt;
    }
  }
}

__global__ void k_set_dt(Scal* sc, double dt) { sc->dt = dt; }
__global__ void k_step_end(Scal* sc) {
  sc->it += 1;
  sc->t = sc->t + sc->dt;
}

}  // namespace
