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

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace vacgpu;

namespace {

constexpr int BX = 64, BY = 4;      // cell kernels: 64 lanes along x, 4 rows
constexpr int NSLOT = 8;            // partial-reduction slots
enum { SLOT_C0 = 0, SLOT_C1 = 1, SLOT_PMIN = 2, SLOT_PMAX = 3, SLOT_DX = 4 };
enum { KF_TVD = 0, KF_PRED = 1, KF_CORR = 2, KF_RYU = 3, KF_BOUND = 4, KF_RED = 5, KF_N = 6 };

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
};

struct BC {
  unsigned char t[6][8];
};

// device-resident scalars
struct Scal {
  double dt, t, smallp;
  double dtcourant[3];
  double tmax;
  int nerror[4];
  int it, tmaxexact;
};

struct CellIx {
  int c[3];
  long long idx;
  bool ok;
};

__device__ __forceinline__ CellIx cell_of(const Geo& g) {
  CellIx r;
  r.c[0] = 2 + (int)(blockIdx.x * BX + threadIdx.x);
  r.c[1] = 2 + (int)(blockIdx.y * BY + threadIdx.y);
  r.c[2] = (g.nd == 3) ? 2 + (int)blockIdx.z : 0;
  r.ok = (r.c[0] <= g.n[0] + 1) && (r.c[1] <= g.n[1] + 1);
  r.idx = g.off + r.c[0] + (long long)r.c[1] * g.s[1] + (long long)r.c[2] * g.s[2];
  return r;
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
    double gq = 0.5 * (b[idx + g.s[d]] - b[idx - g.s[d]]) / g.dx[d][c[d]];
    dv = dv + gq;
  }
  return dv;
}

// ------------------------------------------------------------ ghost cells
// getboundary for 'periodic' and 'cont' (src/vacgrid.t:195-413).  VAC fills
// face regions in the order x-low, x-high, y-low, ... over the full
// transverse ixG range; since both types copy, every ghost cell ends up
// equal to the mesh cell obtained by mapping each coordinate on its own,
// which is what this kernel writes (one thread per ghost cell).
template <int NW>
__global__ __launch_bounds__(256) void k_bound(Geo g, BC bc, double* __restrict__ w,
                                              long long total, long long slab0, long long slab1) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
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
    long long src = g.off;
    for (int k = 0; k < g.nd; ++k) {
      int s = c[k];
      if (c[k] < 2) s = bc.t[2 * k][iw] ? c[k] + g.n[k] : 2;
      else if (c[k] > g.n[k] + 1) s = bc.t[2 * k + 1][iw] ? c[k] - g.n[k] : g.n[k] + 1;
      src += s * g.s[k];
    }
    w[iw * g.ps + dst] = w[iw * g.ps + src];
  }
}

// ------------------------------------------------------------ TVDLF
// Hancock predictor (src/vactvdlf.t:5-89) with qdt = dt/2:
// w1 = w - qdt/dx*(F(w+l/2) - F(w-l/2)) - (qdt/ndim)*S_Powell(w)
template <int ND, bool MHD, class R, int D>
__global__ __launch_bounds__(256) void k_tvdlf_pred(Geo g, Sel s, const double* __restrict__ w,
                                                   double* __restrict__ w1,
                                                   const Scal* __restrict__ sc) {
  constexpr int d = D;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  const double qdt = sc->dt / 2;
  const double smallp = sc->smallp;
  const long long sd = g.s[d];
  double uc[NW], wl[NW], wr[NW];
  load<NW>(w, g.ps, cl.idx, uc);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const double um = w[iw * g.ps + cl.idx - sd], up = w[iw * g.ps + cl.idx + sd];
    const R l = limiter<R>(s.lim[iw], R(up - uc[iw]), R(uc[iw] - um));
    wr[iw] = uc[iw] - 0.5 * double(l);
    wl[iw] = uc[iw] + 0.5 * double(l);
  }
  R fL[NW], fR[NW];
  fluxes<ND, MHD, R>(wl, d, s.gamma, smallp, fL);
  fluxes<ND, MHD, R>(wr, d, s.gamma, smallp, fR);
  const double qdx = qdt / g.dx[d][cl.c[d]];
  double out[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) out[iw] = uc[iw] - qdx * double(fL[iw] - fR[iw]);
  if constexpr (MHD) {
    if (s.sourceunsplit && s.divbfix) {
      const double dv = divb_at<ND>(g, w, cl.idx, cl.c);
      powell<ND>(uc, dv, qdt * 1.0 / ND, out);
    }
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) w1[iw * g.ps + cl.idx] = out[iw];
}

// TVDLF corrector (tvdmusclf, src/vactvdlf.t:92-299) for the faces i-1/2 and
// i+1/2 of cell i: MUSCL states of w1 with limited slopes of w
// (typelimited='original'), local Lax-Friedrichs flux with cmax of the mean
// state, Powell source of w1 with dt/ndim.  Writes the partial max of
// 2*cmax/(dx_j+dx_j+1) (src/vactvdlf.t:190) to part[SLOT_C0].
template <int ND, bool MHD, class R, int D>
__global__ __launch_bounds__(256) void k_tvdlf_corr(Geo g, Sel s, const double* __restrict__ w,
                                                   const double* __restrict__ w1,
                                                   double* __restrict__ wout,
                                                   const Scal* __restrict__ sc,
                                                   double* __restrict__ part, long long pstride) {
  constexpr int d = D;
  using L = Layout<ND, MHD>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const CellIx cl = cell_of(g);
  double cmx = -INFINITY;
  if (cl.ok) {
    const double dt = sc->dt;
    const double smallp = sc->smallp;
    const long long sd = g.s[d];
    const int ci = cl.c[d];
    // limited slopes l(i-1), l(i), l(i+1) of w
    double lm[NW], l0[NW], lp[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const double* p = w + iw * g.ps + cl.idx;
      const double a = p[-2 * sd], b = p[-sd], c = p[0], e = p[sd], f = p[2 * sd];
      lm[iw] = double(limiter<R>(s.lim[iw], R(c - b), R(b - a)));
      l0[iw] = double(limiter<R>(s.lim[iw], R(e - c), R(c - b)));
      lp[iw] = double(limiter<R>(s.lim[iw], R(f - e), R(e - c)));
    }
    R F[2][NW];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      // face J+1/2 with J = i-1 (side 0) or J = i (side 1)
      const long long jl = cl.idx + (side - 1) * sd;
      double wlf[NW], wrf[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        const double* p = w1 + iw * g.ps + jl;
        wrf[iw] = p[sd] - 0.5 * (side ? lp[iw] : l0[iw]);
        wlf[iw] = p[0] + 0.5 * (side ? l0[iw] : lm[iw]);
      }
      double mean[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) mean[iw] = 0.5 * (wlf[iw] + wrf[iw]);
      const R cm = cmax<ND, MHD, R>(mean, d, s.gamma, smallp);
      const int J = ci + side - 1;
      const double cv = double(cm) * 2.0 / (g.dx[d][J] + g.dx[d][J + 1]);
      cmx = cmx < cv ? cv : cmx;
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) wlf[iw] = 2.0 * mean[iw] - wrf[iw];
      R fL[NW], fR[NW];
      fluxes<ND, MHD, R>(wlf, d, s.gamma, smallp, fL);
      fluxes<ND, MHD, R>(wrf, d, s.gamma, smallp, fR);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        const R fm = R(0.5) * (fL[iw] + fR[iw]);
        const R diss = -(cm * R(0.5) * R(wrf[iw] - wlf[iw]));
        F[side][iw] = fm + diss;
      }
    }
    double out[NW];
    const double qdx = dt / g.dx[d][ci];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw)
      out[iw] = w[iw * g.ps + cl.idx] - qdx * double(F[1][iw] - F[0][iw]);
    if constexpr (MHD) {
      if (s.sourceunsplit && s.divbfix) {
        double u1[NW];
        load<NW>(w1, g.ps, cl.idx, u1);
        const double dv = divb_at<ND>(g, w1, cl.idx, cl.c);
        powell<ND>(u1, dv, dt * 1.0 / ND, out);
      }
    }
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) wout[iw * g.ps + cl.idx] = out[iw];
  }
  const double bm = block_reduce<true>(cmx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bm;
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
  for (int iw = 0; iw < NW; ++iw) w1[iw * g.ps + cl.idx] = out[iw];
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

// Everything one face contributes to the update of its two cells.
template <int ND, class R>
struct TvdFace {
  R jump[Layout<ND, true>::nw];
  R adt[Layout<ND, true>::nw];
  RoeRt<ND, R> rt;
};

template <int ND, class R>
__device__ __forceinline__ void tvd_face(const double* uL, const double* uR, int d, const Sel& s,
                                         double smallp, double dt, double dxsum, TvdFace<ND, R>& f,
                                         double& c0, double& c1) {
  constexpr int NW = Layout<ND, true>::nw;
  Roe<ND, R> r;
  roe_setup<ND, R>(uL, uR, d, s.gamma, smallp, r);
  const R sdx = R(dxsum);
#pragma unroll
  for (int il = 0; il < NW; ++il) {
    R a, jp;
    roe_wave<ND, R>(r, uL, uR, il, d, s.divbwave, a, jp);
    const R ad = a * R(dt) * R(2) / sdx;
    f.jump[il] = jp;
    f.adt[il] = ad;
  }
  c0 = fabs(double(f.adt[0]));
  c1 = fabs(double(f.adt[1]));
  f.rt = roe_rt<ND, R>(r);
}

// out[iw] += sum over il (in order) of (R.phi(J) q(J) - R.phi(J-1) q(J-1)) / dV(i)
template <int ND, class R>
__device__ __forceinline__ void tvd_dissipate(const RoeRt<ND, R>& rm, const R* phim,
                                              const RoeRt<ND, R>& r0, const R* phi0, int d,
                                              const Sel& s, R dvm, R dv0, double dvi, double* out) {
  constexpr int NW = Layout<ND, true>::nw;
#pragma unroll
  for (int il = 0; il < NW; ++il) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R a0 = roe_rtimes<ND, R>(rm, phim[il], iw, il, d, s.divbwave) * R(0.25) * dvm;
      const R a1 = roe_rtimes<ND, R>(r0, phi0[il], iw, il, d, s.divbwave) * R(0.25) * dv0;
      out[iw] = out[iw] + double(a1 - a0) / dvi;
    }
  }
}

template <class T>
__device__ __forceinline__ T shfl_up1(T v) { return __shfl_up(v, 1, 64); }
template <class T>
__device__ __forceinline__ T shfl_down1(T v) { return __shfl_down(v, 1, 64); }

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

constexpr int TX_STRIDE = 61;  // cells updated per 64-lane wavefront along x

// x sweep: lane L of a wavefront owns face J = x0 - 2 + L (between cells J and
// J+1) and updates cell J for L in [2, 62].
template <int ND, class R>
__global__ __launch_bounds__(256) void k_tvd_x(Geo g, Sel s, const double* __restrict__ wct,
                                              const double* __restrict__ win,
                                              double* __restrict__ wout,
                                              const Scal* __restrict__ sc,
                                              double* __restrict__ part, long long pstride) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const int lane = threadIdx.x;
  const int x0 = 2 + (int)blockIdx.x * TX_STRIDE;
  const int J = x0 - 2 + lane;
  const int y = 2 + (int)(blockIdx.y * BY + threadIdx.y);
  const int z = (g.nd == 3) ? 2 + (int)blockIdx.z : 0;
  const bool rowok = y <= g.n[1] + 1;
  const int yc = rowok ? y : g.n[1] + 1;
  const int Jc = J < g.n[0] + 2 ? J : g.n[0] + 2;
  const long long base = g.off + (long long)yc * g.s[1] + (long long)z * g.s[2];
  double uL[NW], uR[NW];
  load<NW>(wct, g.ps, base + Jc, uL);
  load<NW>(wct, g.ps, base + Jc + 1, uR);
  TvdFace<ND, R> f;
  double c0, c1;
  tvd_face<ND, R>(uL, uR, 0, s, smallp, dt, g.dx[0][Jc] + g.dx[0][Jc + 1], f, c0, c1);
  const bool faceok = rowok && J <= g.n[0] + 2;
  if (!faceok) { c0 = -INFINITY; c1 = -INFINITY; }
  // centdiff: flux at cell J, face average fC(J) and fC(J-1)
  R F[NW], fC[NW], fCm[NW];
  fluxes<ND, true, R>(uL, 0, s.gamma, smallp, F);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) fC[iw] = R(0.5) * (F[iw] + shfl_down1(F[iw]));
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) fCm[iw] = shfl_up1(fC[iw]);
  // phi at face J from the jumps at J-1, J, J+1
  R phi[NW];
#pragma unroll
  for (int il = 0; il < NW; ++il) {
    const R jm = shfl_up1(f.jump[il]), jp = shfl_down1(f.jump[il]);
    phi[il] = tvd_phi<R>(s.lim[il], jm, f.jump[il], jp, f.adt[il]);
  }
  R phim[NW];
#pragma unroll
  for (int il = 0; il < NW; ++il) phim[il] = shfl_up1(phi[il]);
  const RoeRt<ND, R> rtm = shfl_up_struct<RoeRt<ND, R>, R>(f.rt);
  const int i = J;
  if (rowok && lane >= 2 && lane <= 62 && i <= g.n[0] + 1) {
    const long long idx = base + i;
    double out[NW];
    const double qdx = dt / g.dx[0][i];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) out[iw] = win[iw * g.ps + idx] - qdx * double(fC[iw] - fCm[iw]);
    // dvolume = (dx1*dx2)*dx3 evaluated left to right
    auto dvol = [&](int xi) {
      double v = g.dx[0][xi];
      v = v * g.dx[1][yc];
      if (ND == 3) v = v * g.dx[2][z];
      return v;
    };
    const double dm = dvol(i - 1), d0 = dvol(i), dp = dvol(i + 1);
    tvd_dissipate<ND, R>(rtm, phim, f.rt, phi, 0, s, R(d0 + dm), R(dp + d0), d0, out);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) wout[iw * g.ps + idx] = out[iw];
  }
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = b0;
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + block_linear()] = b1;
}

// y / z sweep: thread (x, segment) marches through cells s0..s1 along d.
template <int ND, class R, int D>
__global__ __launch_bounds__(256) void k_tvd_m(Geo g, Sel s, const double* __restrict__ wct,
                                              const double* __restrict__ win,
                                              double* __restrict__ wout,
                                              const Scal* __restrict__ sc,
                                              double* __restrict__ part, long long pstride,
                                              int seglen) {
  constexpr int d = D;
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  __shared__ double sm[8];
  const double dt = sc->dt, smallp = sc->smallp;
  const int x = 2 + (int)(blockIdx.x * BX + threadIdx.x);
  const int seg = (int)(blockIdx.y * BY + threadIdx.y);
  const int s0 = 2 + seg * seglen;
  // the transverse coordinate that is neither x nor d (3D only)
  const int t = (g.nd == 3) ? 2 + (int)blockIdx.z : 0;
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
    double uL[NW], uR[NW];
    TvdFace<ND, R> fm, f0, fp;  // faces i-1, i, i+1
    R phim[NW], phi0[NW];
    R Fm[NW], F0[NW], fCm[NW];
    double a, b;
    // prologue: faces s0-2, s0-1, s0; centred fluxes at cells s0-1, s0
    load<NW>(wct, g.ps, base + (long long)(s0 - 2) * sd, uL);
    load<NW>(wct, g.ps, base + (long long)(s0 - 1) * sd, uR);
    tvd_face<ND, R>(uL, uR, d, s, smallp, dt, dxd[s0 - 2] + dxd[s0 - 1], fm, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    fluxes<ND, true, R>(uR, d, s.gamma, smallp, Fm);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
    load<NW>(wct, g.ps, base + (long long)s0 * sd, uR);
    tvd_face<ND, R>(uL, uR, d, s, smallp, dt, dxd[s0 - 1] + dxd[s0], f0, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    fluxes<ND, true, R>(uR, d, s.gamma, smallp, F0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) fCm[iw] = R(0.5) * (Fm[iw] + F0[iw]);  // face s0-1
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
    load<NW>(wct, g.ps, base + (long long)(s0 + 1) * sd, uR);
    tvd_face<ND, R>(uL, uR, d, s, smallp, dt, dxd[s0] + dxd[s0 + 1], fp, a, b);
    c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
    // phi at face s0-1 from faces s0-2, s0-1, s0
#pragma unroll
    for (int il = 0; il < NW; ++il) phim[il] = tvd_phi<R>(s.lim[il], fm.jump[il], f0.jump[il], fp.jump[il], f0.adt[il]);
    // window now: fm <- face s0-1, f0 <- face s0 (fp holds face s0)
    fm = f0;
    f0 = fp;
    for (int i = s0;; ++i) {
      // f0 = face i, fm = face i-1 (with phim), uL = w(i), uR = w(i+1)
      fluxes<ND, true, R>(uR, d, s.gamma, smallp, Fm);  // F(i+1), reuse Fm as scratch
      R fC[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) fC[iw] = R(0.5) * (F0[iw] + Fm[iw]);  // face i
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) F0[iw] = Fm[iw];
      // face i+1 from w(i+1), w(i+2)
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) uL[iw] = uR[iw];
      load<NW>(wct, g.ps, base + (long long)(i + 2) * sd, uR);
      tvd_face<ND, R>(uL, uR, d, s, smallp, dt, dxd[i + 1] + dxd[i + 2], fp, a, b);
      c0 = c0 < a ? a : c0; c1 = c1 < b ? b : c1;
#pragma unroll
      for (int il = 0; il < NW; ++il) phi0[il] = tvd_phi<R>(s.lim[il], fm.jump[il], f0.jump[il], fp.jump[il], f0.adt[il]);
      // update cell i
      const long long idx = base + (long long)i * sd;
      double out[NW];
      const double qdx = dt / dxd[i];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) out[iw] = win[iw * g.ps + idx] - qdx * double(fC[iw] - fCm[iw]);
      const double dm = dvol(i - 1), d0 = dvol(i), dp = dvol(i + 1);
      tvd_dissipate<ND, R>(fm.rt, phim, f0.rt, phi0, d, s, R(d0 + dm), R(dp + d0), d0, out);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) wout[iw * g.ps + idx] = out[iw];
      if (i == s1) break;
      // shift the window
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) fCm[iw] = fC[iw];
#pragma unroll
      for (int il = 0; il < NW; ++il) phim[il] = phi0[il];
      fm = f0;
      f0 = fp;
    }
  }
  const double b0 = block_reduce<true>(c0, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = b0;
  const double b1 = block_reduce<true>(c1, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + block_linear()] = b1;
}

// Ryu source step on the last sweep (src/vactvd.t:60-80):
// wt = (wold + wnew - dt*S(wold))/2 on the mesh (ghosts refreshed after) ...
template <int ND>
__global__ __launch_bounds__(256) void k_ryu_a(Geo g, const double* __restrict__ wold,
                                              const double* __restrict__ wnew,
                                              double* __restrict__ wt,
                                              const Scal* __restrict__ sc) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double uo[NW], t[NW];
  load<NW>(wold, g.ps, cl.idx, uo);
  load<NW>(wnew, g.ps, cl.idx, t);
  powell<ND>(uo, divb_at<ND>(g, wold, cl.idx, cl.c), sc->dt, t);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) wt[iw * g.ps + cl.idx] = 0.5 * (uo[iw] + t[iw]);
}

// ... then wnew = wnew - dt*S(wt).
template <int ND>
__global__ __launch_bounds__(256) void k_ryu_b(Geo g, const double* __restrict__ wt,
                                              double* __restrict__ wnew,
                                              const Scal* __restrict__ sc) {
  using L = Layout<ND, true>;
  constexpr int NW = L::nw;
  const CellIx cl = cell_of(g);
  if (!cl.ok) return;
  double ut[NW], o[NW];
  load<NW>(wt, g.ps, cl.idx, ut);
  load<NW>(wnew, g.ps, cl.idx, o);
  powell<ND>(ut, divb_at<ND>(g, wt, cl.idx, cl.c), sc->dt, o);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) wnew[iw * g.ps + cl.idx] = o[iw];
}

// ------------------------------------------------------------ time step
// Partial min / max of the unclipped pressure over the mesh, for the
// pressure floor set on the first getpthermal call (src/vacphys0.t.mhd:202-214).
template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_pminmax(Geo g, const double* __restrict__ w, double gamma,
                                                const Scal* __restrict__ sc,
                                                double* __restrict__ part, long long pstride) {
  using L = Layout<ND, MHD>;
  __shared__ double sm[8];
  if (sc->smallp >= 0.0) return;  // uniform: already set
  const CellIx cl = cell_of(g);
  double mn = INFINITY, mx = -INFINITY;
  if (cl.ok) {
    double u[L::nw];
    load<L::nw>(w, g.ps, cl.idx, u);
    mn = mx = pthermal_raw<ND, MHD>(u, gamma);
  }
  const double bmn = block_reduce<false>(mn, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMIN * pstride + block_linear()] = bmn;
  const double bmx = block_reduce<true>(mx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_PMAX * pstride + block_linear()] = bmx;
}

// Partial max of cmax/dx at cell centres for dimension d (src/vac.t:697-701),
// skipped when a lagged dtcourant(d) exists.
template <int ND, bool MHD>
__global__ __launch_bounds__(256) void k_cmaxdx(Geo g, const double* __restrict__ w, double gamma,
                                               const Scal* __restrict__ sc,
                                               double* __restrict__ part, long long pstride, int d) {
  using L = Layout<ND, MHD>;
  __shared__ double sm[8];
  if (sc->dtcourant[d] < kBig) return;  // uniform
  const double smallp = sc->smallp;
  const CellIx cl = cell_of(g);
  double v = -INFINITY;
  if (cl.ok) {
    double u[L::nw];
    load<L::nw>(w, g.ps, cl.idx, u);
    v = cmax<ND, MHD, double>(u, d, gamma, smallp) / g.dx[d][cl.c[d]];
  }
  const double b = block_reduce<true>(v, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[(SLOT_DX + d) * pstride + block_linear()] = b;
}

__device__ double fold(const double* p, long long n, bool is_max, double* sm) {
  double v = is_max ? -INFINITY : INFINITY;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    double u = p[i];
    v = is_max ? (v < u ? u : v) : (u < v ? u : v);
  }
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

// getdt_courant's fold (src/vac.t:690-719) and the main loop's tmaxexact
// clip (src/vac.t:83); dtpar > 0 replaces the Courant step.
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
  if (threadIdx.x == 0) {
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
    }
    if (tmaxexact) {
      double c = sc->tmax - sc->t + kSmall;
      dt = c < dt ? c : dt;
    }
    sc->dt = dt;
  }
}

// dtcourant update of the TVDLF corrector (src/vactvdlf.t:184-201)
__global__ void k_fin_courant_tvdlf(Scal* sc, const double* part, long long pstride, long long nb,
                                    int d, double courantpar) {
  __shared__ double sm[16];
  const double cm = fold(part + SLOT_C0 * pstride, nb, true, sm);
  if (threadIdx.x == 0) {
    if (sc->dt * cm > 1.0) sc->nerror[2] += 1;
    if (cm > kSmall) {
      double c = courantpar / cm;
      sc->dtcourant[d] = c < sc->dtcourant[d] ? c : sc->dtcourant[d];
    }
  }
}

// dtcourant update of tvdlimit2 for the two extreme waves (src/vactvd.t:145-157)
__global__ void k_fin_courant_tvd(Scal* sc, const double* part, long long pstride, long long nb,
                                  int d, double courantpar) {
  __shared__ double sm[16];
  for (int slot = SLOT_C0; slot <= SLOT_C1; ++slot) {
    const double cm = fold(part + slot * pstride, nb, true, sm);
    if (threadIdx.x == 0) {
      if (cm > 1.0) sc->nerror[2] += 1;
      if (cm > kSmall) {
        double c = sc->dt * courantpar / cm;
        sc->dtcourant[d] = c < sc->dtcourant[d] ? c : sc->dtcourant[d];
      }
    }
    __syncthreads();
  }
}

__global__ void k_set_dt(Scal* sc, double dt) { sc->dt = dt; }
__global__ void k_step_end(Scal* sc) {
  sc->it += 1;
  sc->t = sc->t + sc->dt;
}

}  // namespace

// ======================================================================
// host side
// ======================================================================

struct vacgpu_grid {
  vacgpu_config cfg{};
  Geo geo{};
  BC bc{};
  int nw = 0;
  double* buf[3] = {nullptr, nullptr, nullptr};
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
  long long nbound = 0, slab[3] = {0, 0, 0};
  int cur = 0, old = 0;
  bool have_state = false;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  // profiling
  bool prof = false;
  struct Ev { int k; hipEvent_t a, b; };
  std::vector<Ev> evs;
  double tsum[KF_N] = {0, 0, 0, 0, 0, 0};
  long tcnt[KF_N] = {0, 0, 0, 0, 0, 0};
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
      if (c->typeB[f][iw] != VACGPU_CONT && c->typeB[f][iw] != VACGPU_PERIODIC) {
        why = "typeB must be 'cont' or 'periodic'";
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
  if (m->typefull != VACGPU_TVDLF && m->typefull != VACGPU_TVD) { why = "typefull must be 'tvdlf' or 'tvd'"; return VACGPU_E_ARG; }
  if (m->typefull == VACGPU_TVD && c->physics == VACGPU_HD) { why = "typefull='tvd' needs the HD Roe solver, which is not on the GPU path"; return VACGPU_E_ARG; }
  int nw = c->physics == VACGPU_MHD ? 2 + 2 * c->ndim : 2 + c->ndim;
  for (int iw = 0; iw < nw; ++iw)
    if (m->typelimiter[iw] < VACGPU_MINMOD || m->typelimiter[iw] > VACGPU_SUPERBEE) { why = "typelimiter must be minmod, woodward or superbee"; return VACGPU_E_ARG; }
  if (c->physics == VACGPU_MHD && m->divbfix && !m->sourceunsplit) { why = "divbfix=T requires sourceunsplit=T (split sources are not on the GPU path)"; return VACGPU_E_ARG; }
  if (!(m->smallpcoeff >= 0.0)) { why = "smallpcoeff must be >= 0"; return VACGPU_E_ARG; }
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
  return s;
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

int bound(vacgpu_grid* g, double* w) {
  if (g->nbound == 0) return 0;
  Timed t(g, KF_BOUND);
  const unsigned nb = (unsigned)((g->nbound + 255) / 256);
#define BOUND_NW(NWV) hipLaunchKernelGGL(k_bound<NWV>, dim3(nb), dim3(256), 0, g->stream, g->geo, g->bc, w, g->nbound, g->slab[0], g->slab[1])
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
        for (int d = 0; d < ND; ++d)
          hipLaunchKernelGGL((k_cmaxdx<ND, MHD>), g->grid, g->block, 0, g->stream, g->geo, w,
                             g->cfg.eqpar[0], g->sc, g->part, g->pstride, d);
    });
    hipLaunchKernelGGL(k_fin_getdt, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part, g->pstride,
                       nb, g->cfg.ndim, m->courantpar, dtpar, tmaxexact);
  }
  return launch_check(g, "getdt");
}

// one dimensional sweep (advect, src/vac.t:363-486), dt already on device
int sweep_device(vacgpu_grid* g, const vacgpu_method* m, int d, int firstsweep, int lastsweep) {
  if (firstsweep) g->old = g->cur;
  const int outi = other_buffer(g);
  double* win = g->buf[g->cur];
  double* wout = g->buf[outi];
  double* wold = g->buf[g->old];
  const Sel s = make_sel(g, m);
  const long long nb = (long long)g->grid.x * g->grid.y * g->grid.z;
  const bool mhd = g->cfg.physics == VACGPU_MHD;
  const bool src = mhd && m->divbfix && m->sourceunsplit;
  const bool courant = m->courantpar > 0.0;
  int rc = 0;
  if (m->typefull == VACGPU_TVDLF) {
    {
      Timed t(g, KF_PRED);
      DISPATCH(g, {
        if (d == 0) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, 0>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, g->sc);
        else if (d == 1) hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, 1>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, g->sc);
        else hipLaunchKernelGGL((k_tvdlf_pred<ND, MHD, R, (ND == 3 ? 2 : 1)>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, g->sc);
      });
      if ((rc = launch_check(g, "k_tvdlf_pred"))) return rc;
    }
    if ((rc = bound(g, g->w1))) return rc;
    {
      Timed t(g, KF_CORR);
      DISPATCH(g, {
        if (d == 0) hipLaunchKernelGGL((k_tvdlf_corr<ND, MHD, R, 0>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, wout, g->sc, g->part, g->pstride);
        else if (d == 1) hipLaunchKernelGGL((k_tvdlf_corr<ND, MHD, R, 1>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, wout, g->sc, g->part, g->pstride);
        else hipLaunchKernelGGL((k_tvdlf_corr<ND, MHD, R, (ND == 3 ? 2 : 1)>), g->grid, g->block, 0, g->stream, g->geo, s, win, g->w1, wout, g->sc, g->part, g->pstride);
      });
      if ((rc = launch_check(g, "k_tvdlf_corr"))) return rc;
    }
    if (courant) {
      Timed t(g, KF_RED);
      hipLaunchKernelGGL(k_fin_courant_tvdlf, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part,
                         g->pstride, nb, d, m->courantpar);
      if ((rc = launch_check(g, "k_fin_courant_tvdlf"))) return rc;
    }
  } else {
    const double* wct = win;
    if (firstsweep && src) {
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
      if (d == 0) {
        const dim3 gr = g->grid_tx;
        nbt = (long long)gr.x * gr.y * gr.z;
#define TVDX(NDV, RT) hipLaunchKernelGGL((k_tvd_x<NDV, RT>), gr, g->block, 0, g->stream, g->geo, s, wct, win, wout, g->sc, g->part, g->pstride)
        if (g->cfg.ndim == 2) { if (f32) TVDX(2, float); else TVDX(2, double); }
        else { if (f32) TVDX(3, float); else TVDX(3, double); }
#undef TVDX
      } else {
        const dim3 gr = g->grid_tm[d];
        nbt = (long long)gr.x * gr.y * gr.z;
#define TVDM(NDV, RT, DV) hipLaunchKernelGGL((k_tvd_m<NDV, RT, DV>), gr, g->block, 0, g->stream, g->geo, s, wct, win, wout, g->sc, g->part, g->pstride, g->seglen)
        if (g->cfg.ndim == 2) { if (f32) TVDM(2, float, 1); else TVDM(2, double, 1); }
        else if (d == 1) { if (f32) TVDM(3, float, 1); else TVDM(3, double, 1); }
        else { if (f32) TVDM(3, float, 2); else TVDM(3, double, 2); }
#undef TVDM
      }
      if ((rc = launch_check(g, "k_tvd_sweep"))) return rc;
    }
    if (courant) {
      Timed t(g, KF_RED);
      hipLaunchKernelGGL(k_fin_courant_tvd, dim3(1), dim3(1024), 0, g->stream, g->sc, g->part,
                         g->pstride, nbt, d, m->courantpar);
      if ((rc = launch_check(g, "k_fin_courant_tvd"))) return rc;
    }
    if (lastsweep && src) {
      {
        Timed t(g, KF_RYU);
        if (g->cfg.ndim == 2)
          hipLaunchKernelGGL(k_ryu_a<2>, g->grid, g->block, 0, g->stream, g->geo, wold, wout, g->w1, g->sc);
        else
          hipLaunchKernelGGL(k_ryu_a<3>, g->grid, g->block, 0, g->stream, g->geo, wold, wout, g->w1, g->sc);
        if ((rc = launch_check(g, "k_ryu_a"))) return rc;
      }
      if ((rc = bound(g, g->w1))) return rc;
      {
        Timed t(g, KF_RYU);
        if (g->cfg.ndim == 2)
          hipLaunchKernelGGL(k_ryu_b<2>, g->grid, g->block, 0, g->stream, g->geo, g->w1, wout, g->sc);
        else
          hipLaunchKernelGGL(k_ryu_b<3>, g->grid, g->block, 0, g->stream, g->geo, g->w1, wout, g->sc);
        if ((rc = launch_check(g, "k_ryu_b"))) return rc;
      }
    }
  }
  if ((rc = bound(g, wout))) return rc;
  g->cur = outi;
  return 0;
}

int advance_device(vacgpu_grid* g, const vacgpu_method* m, int it) {
  const int nd = g->cfg.ndim;
  int rc;
  if ((it / 2) * 2 == it || m->dimsplit_xy) {
    for (int d = 0; d < nd; ++d)
      if ((rc = sweep_device(g, m, d, d == 0, d == nd - 1))) return rc;
  } else {
    for (int d = nd - 1; d >= 0; --d)
      if ((rc = sweep_device(g, m, d, d == nd - 1, d == 0))) return rc;
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
      g->bc.t[f][iw] = (f < 2 * nd && iw < g->nw) ? (unsigned char)(cfg->typeB[f][iw] == VACGPU_PERIODIC) : 0;
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
  auto count = [](dim3 v) { return (long long)v.x * v.y * v.z; };
  g->pstride = count(g->grid);
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
  for (int iw = 0; iw < g->nw; ++iw)
    for (int z = 0; z < nz; ++z) {
      double* h = host + (size_t)iw * l[0] * l[1] * (nd == 3 ? l[2] : 1) + (size_t)z * l[0] * l[1];
      double* dv = dev + iw * geo.ps + geo.off + z * geo.s[2];
      if (put)
        HIPCHK(g, hipMemcpy2DAsync(dv, dpitch, h, hpitch, width, geo.ng[1], hipMemcpyHostToDevice, g->stream));
      else
        HIPCHK(g, hipMemcpy2DAsync(h, hpitch, dv, dpitch, width, geo.ng[1], hipMemcpyDeviceToHost, g->stream));
    }
  HIPCHK(g, hipStreamSynchronize(g->stream));
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
  if (dtcourant || smallp) {
    if ((rc = read_scalars(g))) return rc;
    if (dtcourant)
      for (int d = 0; d < g->cfg.ndim; ++d) g->sc_host->dtcourant[d] = dtcourant[d];
    if (smallp) g->sc_host->smallp = *smallp;
    if ((rc = write_scalars(g))) return rc;
  }
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
  if (idimmin != idimmax) return fail(g, VACGPU_E_ARG, "only dimensionally split sweeps (idimmin == idimmax) are on the GPU path");
  if (idimmin < 1 || idimmin > g->cfg.ndim) return fail(g, VACGPU_E_ARG, "idim out of range");
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
  if ((rc = sweep_device(g, m, idimmin - 1, firstsweep, lastsweep))) return rc;
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
  if ((rc = write_scalars(g))) return rc;
  int it = g->sc_host->it;
  double t = g->sc_host->t;
  const bool sync_each = (tmax < kBig) || t_hist;
  int done = 0;
  for (;;) {
    if (t_hist) t_hist[done] = t;
    if ((rc = getdt_device(g, m, dtpar, tmaxexact))) return rc;
    double dt = 1.0;
    if (sync_each) {
      if ((rc = read_scalars(g))) return rc;
      dt = g->sc_host->dt;
    }
    if (it >= itmax || t >= tmax || dt < 1.0e-99 || done >= nsteps) break;
    if ((rc = advance_device(g, m, it))) return rc;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, g->stream, g->sc);
    if ((rc = launch_check(g, "k_step_end"))) return rc;
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

int vacgpu_kernel_time(vacgpu_grid* g, int k, double* avg_ms, long* count) {
  if (!g || k < 0 || k >= KF_N) return VACGPU_E_ARG;
  HIPCHK(g, hipStreamSynchronize(g->stream));
  harvest(g);
  if (avg_ms) *avg_ms = g->tcnt[k] ? g->tsum[k] / g->tcnt[k] : 0.0;
  if (count) *count = g->tcnt[k];
  return 0;
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
  if (g->w1) (void)hipFree(g->w1);
  if (g->dxdev) (void)hipFree(g->dxdev);
  if (g->sc) (void)hipFree(g->sc);
  if (g->sc_host) (void)hipHostFree(g->sc_host);
  if (g->part) (void)hipFree(g->part);
  if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // extern "C"
