// bulk32.hpp -- the FP32-bulk form of the fused 2D TVDLF sweeps.
//
// Every state the sweep evaluates (the Hancock half-states, the predictor
// state, the corrector face states and their mean) is written as an FP64
// base -- the start-of-sweep w of one cell, read from HBM -- plus an FP32
// increment.  The fluxes, the limiters, cmax and the Powell source run in
// FP32 on base + increment.  What stays in FP64:
//   * the pressure recovery: p at the base cell is recovered in FP64 once per
//     cell, and the pressure of a state is that FP64 value plus the FP32
//     change dp of the pressure formula over the increment (computed from
//     the algebraic difference, so no cancellation is left to FP32);
//   * the pressure part of the normal-momentum flux, and the flux
//     difference across a cell of that component;
//   * differences of base states between cells (then rounded to FP32);
//   * the accumulation into the conserved variables;
//   * the Courant value handed to the dt reduction.
// Included by sweep.hpp inside its anonymous namespace.
#pragma once

// pressure increments of state b + q over state b (FP32, b rounded base):
// total pressure (MHD getptotal, src/vacphys0.t.mhd:244-252) and thermal
// pressure (getpthermal before clipping, :191-199; HD src/vacphys0.t.hd:118)
template <int ND, bool MHD>
__device__ __forceinline__ void dpress(const float* b, const float* q, float gm1, float& dptot,
                                       float& dpth) {
  using L = Layout<ND, MHD>;
  float dm2 = (2.f * b[L::m(0)] + q[L::m(0)]) * q[L::m(0)];
  float m2 = b[L::m(0)] * b[L::m(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) {
    dm2 += (2.f * b[L::m(c)] + q[L::m(c)]) * q[L::m(c)];
    m2 += b[L::m(c)] * b[L::m(c)];
  }
  const float rho = b[L::rho], rho1 = b[L::rho] + q[L::rho];
  const float dkin = pdiv(rho * dm2 - q[L::rho] * m2, rho * rho1);  // d(m^2/rho)
  if constexpr (MHD) {
    float db2 = (2.f * b[L::b(0)] + q[L::b(0)]) * q[L::b(0)];
#pragma unroll
    for (int c = 1; c < ND; ++c) db2 += (2.f * b[L::b(c)] + q[L::b(c)]) * q[L::b(c)];
    dpth = gm1 * (q[L::e] - 0.5f * dkin - 0.5f * db2);
    dptot = dpth + 0.5f * db2;
  } else {
    dpth = gm1 * (q[L::e] - 0.5f * dkin);
    dptot = dpth;
  }
}

// FP64 pressures of a base state: total (MHD flux), thermal (cmax, HD flux)
template <int ND, bool MHD>
__device__ __forceinline__ void base_press(const double* u, double g, double& ptot, double& pth) {
  pth = pthermal_raw<ND, MHD>(u, g);
  if constexpr (MHD) {
    using L = Layout<ND, true>;
    double b2 = u[L::b(0)] * u[L::b(0)];
#pragma unroll
    for (int c = 1; c < ND; ++c) b2 = b2 + u[L::b(c)] * u[L::b(c)];
    ptot = pth + 0.5 * b2;
  } else {
    ptot = pth;
  }
}

// flux along d of the FP32 state s whose normal-momentum flux pressure part
// is pmd (FP64: the total pressure for MHD, the clipped thermal pressure for
// HD) and whose energy-flux pressure is pf (FP32).  f[] gets every flux in
// FP32 with f[m_d] excluding the pressure; fmd = pmd + that rest, in FP64.
template <int ND, bool MHD, int D>
__device__ __forceinline__ void flux32(const float* s, double pmd, float pf, float* f, double& fmd) {
  using L = Layout<ND, MHD>;
  constexpr int d = D;
  const float rho = s[L::rho];
  const float v = pdiv(s[L::m(d)], rho);
  f[L::rho] = v * rho;
  if constexpr (MHD) {
    const float bd = s[L::b(d)];
#pragma unroll
    for (int c = 0; c < ND; ++c) f[L::m(c)] = -(s[L::b(c)] * bd) + v * s[L::m(c)];
    float bm = s[L::b(0)] * s[L::m(0)];
#pragma unroll
    for (int c = 1; c < ND; ++c) bm += s[L::b(c)] * s[L::m(c)];
    f[L::e] = pdiv(s[L::m(d)] * pf - bd * bm, rho) + v * s[L::e];
#pragma unroll
    for (int c = 0; c < ND; ++c) f[L::b(c)] = (c == d) ? 0.f : -(pdiv(s[L::m(c)], rho) * bd) + v * s[L::b(c)];
  } else {
#pragma unroll
    for (int c = 0; c < ND; ++c) f[L::m(c)] = v * s[L::m(c)];
    f[L::e] = v * pf + v * s[L::e];
  }
  fmd = pmd + double(f[L::m(d)]);
}

// one state's flux from its increment q over the rounded base b with FP64
// base pressures ptot / pth
template <int ND, bool MHD, int D>
__device__ __forceinline__ void flux_of(const float* b, const float* q, double ptot, double pth,
                                        float gm1, double smallp, float* f, double& fmd,
                                        float* s_out = nullptr) {
  constexpr int NW = Layout<ND, MHD>::nw;
  float s[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) s[iw] = b[iw] + q[iw];
  float dpt, dph;
  dpress<ND, MHD>(b, q, gm1, dpt, dph);
  double pmd;
  if constexpr (MHD) {
    pmd = ptot + double(dpt);
  } else {
    pmd = pth + double(dph);
    pmd = pmd < smallp ? smallp : pmd;  // getpthermal clips in the HD flux
  }
  flux32<ND, MHD, D>(s, pmd, float(pmd), f, fmd);
  if (s_out)
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) s_out[iw] = s[iw];
}

// cmax along d of the state b + q (thermal pressure clipped with smallp)
template <int ND, bool MHD, int D>
__device__ __forceinline__ float cmax_of(const float* b, const float* q, double pth, float g,
                                         float gm1, double smallp) {
  using L = Layout<ND, MHD>;
  float s[L::nw];
#pragma unroll
  for (int iw = 0; iw < L::nw; ++iw) s[iw] = b[iw] + q[iw];
  float dpt, dph;
  dpress<ND, MHD>(b, q, gm1, dpt, dph);
  double p = pth + double(dph);
  p = p < smallp ? smallp : p;
  const float rho = s[L::rho];
  const float c2 = pdiv(g * float(p), rho);
  if constexpr (MHD) {
    float bsq = s[L::b(0)] * s[L::b(0)];
#pragma unroll
    for (int c = 1; c < ND; ++c) bsq += s[L::b(c)] * s[L::b(c)];
    const float cf2 = pdiv(bsq, rho) + c2;
    const float bi = s[L::b(D)];
    // the discriminant is >= 0 in exact arithmetic; FP32 rounding can take it
    // a hair below 0 where slow and fast speeds meet
    const float disc = fmaxf(cf2 * cf2 - pdiv(4.f * c2 * (bi * bi), rho), 0.f);
    return sqrtf(0.5f * (cf2 + sqrtf(disc))) + fabsf(pdiv(s[L::m(D)], rho));
  } else {
    return sqrtf(c2) + fabsf(pdiv(s[L::m(D)], rho));
  }
}

// Powell source increment of state s (FP32) with divB already scaled by the
// substep (src/vacphys.mhd0.t:263-276): inc[m_c] = -B_c divb,
// inc[B_c] = -(m_c/rho) divb, inc[e] = -(m.B/rho) divb
template <int ND>
__device__ __forceinline__ void powell32(const float* s, float divb, float* inc) {
  using L = Layout<ND, true>;
  const float rho = s[L::rho];
  float mb = s[L::m(0)] * s[L::b(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) mb += s[L::m(c)] * s[L::b(c)];
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    inc[L::m(c)] -= s[L::b(c)] * divb;
    inc[L::b(c)] -= pdiv(s[L::m(c)], rho) * divb;
  }
  inc[L::e] -= pdiv(mb, rho) * divb;
}

// The predictor in increments: dP = P - U at a cell with base U (FP64) and
// its neighbours um / up along D, returning the FP32 slopes too.  bdiff[k]
// = B_k(i + e_k) - B_k(i - e_k) for k != D (FP32 of an FP64 difference);
// dxk = widths at the cell.
template <bool MHD, int D, int LIM>
__device__ __forceinline__ void pred32(const Sel& s, const double* uc, const double* um,
                                       const double* up, const float* bdiff, const double* dxk,
                                       double ptot, double pth, double dt, double smallp,
                                       const float* b, float* dP, float* l) {
  using L = Layout<2, MHD>;
  constexpr int NW = L::nw;
  const float gm1 = float(s.gamma - 1.0);
  float ql[NW], qr[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    l[iw] = limiter_t<LIM, float>(s.lim[iw], float(up[iw] - uc[iw]), float(uc[iw] - um[iw]));
    ql[iw] = 0.5f * l[iw];
    qr[iw] = -0.5f * l[iw];
  }
  float fL[NW], fR[NW];
  double fLm, fRm;
  flux_of<2, MHD, D>(b, ql, ptot, pth, gm1, smallp, fL, fLm);
  flux_of<2, MHD, D>(b, qr, ptot, pth, gm1, smallp, fR, fRm);
  const float qdx = float(dt / 2 / dxk[D]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) dP[iw] = -qdx * (fL[iw] - fR[iw]);
  dP[L::m(D)] = -qdx * float(fLm - fRm);
  if constexpr (MHD) {
    // divB of w with dt/2/ndim
    float dv = 0.f;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const float bd = (k == D) ? float(up[L::b(k)] - um[L::b(k)]) : bdiff[k];
      dv += pdiv(0.5f * bd, float(dxk[k]));
    }
    powell32<2>(b, dv * float(dt / 2 / 2), dP);
  }
}

// The corrector flux at a face from the increments dL, dR of its left and
// right states over the left cell's base b (FP64 pressures ptot, pth).
// Returns F in FP32 (F[m_d] unused), Fmd in FP64 and the face's cmax.
template <bool MHD, int D>
__device__ __forceinline__ float face32(const Sel& s, const float* b, const float* dL,
                                        const float* dR, double ptot, double pth, double smallp,
                                        float* F, double& Fmd) {
  using L = Layout<2, MHD>;
  constexpr int NW = L::nw;
  const float g = float(s.gamma), gm1 = float(s.gamma - 1.0);
  float dM[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) dM[iw] = 0.5f * (dL[iw] + dR[iw]);
  const float cm = cmax_of<2, MHD, D>(b, dM, pth, g, gm1, smallp);
  float fL[NW], fR[NW];
  double fLm, fRm;
  flux_of<2, MHD, D>(b, dL, ptot, pth, gm1, smallp, fL, fLm);
  flux_of<2, MHD, D>(b, dR, ptot, pth, gm1, smallp, fR, fRm);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) F[iw] = 0.5f * (fL[iw] + fR[iw]) - cm * 0.5f * (dR[iw] - dL[iw]);
  Fmd = 0.5 * (fLm + fRm) + double(-cm * 0.5f * (dR[L::m(D)] - dL[L::m(D)]));
  return cm;
}

// The predictor of one row at column px (image coordinates): base U, rounded
// base b, FP64 pressures, increment dP and slopes l.  make() loads its
// inputs from HBM; build() takes them from registers (uc, its neighbours um
// and up along D, the rounded difference bd of the B component across D
// and the two widths).
template <bool MHD, int D, int LIM>
struct BulkRow {
  static constexpr int NW = Layout<2, MHD>::nw;
  double U[NW];
  float dP[NW], l[NW];
  double ptot, pth;
  __device__ __forceinline__ void build(const Sel& s, const double* uc, const double* um,
                                        const double* up, float bd, double dx0, double dx1,
                                        double dt, double smallp) {
    float b[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      U[iw] = uc[iw];
      b[iw] = float(U[iw]);
    }
    float bdiff[2] = {0.f, 0.f};
    if constexpr (MHD) bdiff[1 - D] = bd;
    const double dxk[2] = {dx0, dx1};
    base_press<2, MHD>(U, s.gamma, ptot, pth);
    pred32<MHD, D, LIM>(s, U, um, up, bdiff, dxk, ptot, pth, dt, smallp, b, dP, l);
  }
  __device__ __forceinline__ void make(const Geo& g, const FusedGeo& f, const Sel& s,
                                       const double* __restrict__ w, int px, int py, double dt,
                                       double smallp) {
    using L = Layout<2, MHD>;
    const unsigned o = f.at(px, py);
    const unsigned sd = D == 0 ? 8u : f.sy8, st = D == 0 ? f.sy8 : 8u;
    double um[NW], up[NW];
    float b[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      U[iw] = ldp(w, g.ps, iw, o);
      um[iw] = ldp(w, g.ps, iw, o - sd);
      up[iw] = ldp(w, g.ps, iw, o + sd);
      b[iw] = float(U[iw]);
    }
    float bdiff[2] = {0.f, 0.f};
    if constexpr (MHD)
      bdiff[1 - D] = float(ldp(w, g.ps, L::b(1 - D), o + st) - ldp(w, g.ps, L::b(1 - D), o - st));
    const double dxk[2] = {g.dx[0][px], g.dx[1][py]};
    base_press<2, MHD>(U, s.gamma, ptot, pth);
    pred32<MHD, D, LIM>(s, U, um, up, bdiff, dxk, ptot, pth, dt, smallp, b, dP, l);
  }
};

// x sweep, FP32-bulk.  Lane layout and marching as in tvdlf_fx_body; the
// faces of a row are evaluated right after its predictor, so a marching step
// carries only the row's base, predictor state and face fluxes.
template <bool MHD, int LIM>
__device__ __forceinline__ double bulk_fx_body(const Geo& g, const Sel& s, const BC& bc,
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
  const bool face_ok = lane <= FS_STRIDE && c <= f.n0 + 1;
  c = c < f.n0 + 2 ? c : f.n0 + 2;
  const int px = ghost_image(c, f.n0, f.pxlo, f.pxhi);
  const bool zero_slope = (c < 2 && !f.pxlo) || (c > f.n0 + 1 && !f.pxhi);
  const int s0 = 2 + seg * seglen;
  double cmx = -INFINITY;
  if (s0 > f.n1 + 1) return cmx;
  const int s1 = (s0 + seglen - 1 < f.n1 + 1) ? s0 + seglen - 1 : f.n1 + 1;
  const double* dx0 = g.dx[0];
  const double* dx1 = g.dx[1];
  const double dxs = dx0[c] + dx0[c + 1];
  const double qdx = pdiv(dt, dx0[c]);
  // the faces c+1/2 of a row: F (FP32, m_x part excluded) and Fmd (FP64)
  auto faces = [&](const BulkRow<MHD, 0, LIM>& R, float* F, double& Fmd, double& cv) {
    float dL[NW], dR[NW], b[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const float lc = zero_slope ? 0.f : R.l[iw];
      const double Ur = lane_next(R.U[iw]);
      const float dPr = lane_next(R.dP[iw]);
      const float lr = lane_next(lc);
      b[iw] = float(R.U[iw]);
      dL[iw] = R.dP[iw] + 0.5f * lc;
      dR[iw] = float(Ur - R.U[iw]) + dPr - 0.5f * lr;
    }
    const float cm = face32<MHD, 0>(s, b, dL, dR, R.ptot, R.pth, smallp, F, Fmd);
    cv = pdiv(double(cm) * 2.0, dxs);
  };
  // carried from row j: base, predictor B_x increment, rounded predictor
  // state, face fluxes; and B_y (base, increment) of rows j-1, j
  double U0[NW], Fmd0, byU_m = 0.0, byU_0 = 0.0;
  float sP0[NW], F0[NW], dPbx0 = 0.f, byd_m = 0.f, byd_0 = 0.f;
  {
    BulkRow<MHD, 0, LIM> R;
    R.make(g, f, s, w, px, ghost_image(s0 - 1, f.n1, f.pylo, f.pyhi), dt, smallp);
    if constexpr (MHD) { byU_m = R.U[L::b(1)]; byd_m = R.dP[L::b(1)]; }
    R.make(g, f, s, w, px, s0, dt, smallp);
    double cv;
    faces(R, F0, Fmd0, cv);
    if (face_ok) cmx = cmx < cv ? cv : cmx;
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      U0[iw] = R.U[iw];
      sP0[iw] = float(R.U[iw]) + R.dP[iw];
    }
    if constexpr (MHD) { dPbx0 = R.dP[L::b(0)]; byU_0 = R.U[L::b(1)]; byd_0 = R.dP[L::b(1)]; }
  }
  for (int j = s0; j <= s1; ++j) {
    BulkRow<MHD, 0, LIM> R;  // row j+1
    R.make(g, f, s, w, px, ghost_image(j + 1, f.n1, f.pylo, f.pyhi), dt, smallp);
    float F1[NW];
    double Fmd1 = 0.0;
    if (j + 1 <= s1) {  // wave-uniform
      double cv;
      faces(R, F1, Fmd1, cv);
      if (face_ok) cmx = cmx < cv ? cv : cmx;
    }
    // update row j
    float inc[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) inc[iw] = 0.f;
    if constexpr (MHD) {
      const double bxp = lane_next(U0[L::b(0)]), bxm = lane_prev(U0[L::b(0)]);
      const float dxp_ = lane_next(dPbx0), dxm_ = lane_prev(dPbx0);
      const float gx = pdiv(0.5f * (float(bxp - bxm) + (dxp_ - dxm_)), float(dx0[c]));
      const float gy = pdiv(0.5f * (float(R.U[L::b(1)] - byU_m) + (R.dP[L::b(1)] - byd_m)), float(dx1[j]));
      powell32<ND>(sP0, (gx + gy) * float(dt / ND), inc);
    }
    double out[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const float Fm = lane_prev(F0[iw]);
      out[iw] = U0[iw] - qdx * double(F0[iw] - Fm) + double(inc[iw]);
    }
    {
      const double Fmm = lane_prev(Fmd0);
      out[L::m(0)] = U0[L::m(0)] - qdx * (Fmd0 - Fmm) + double(inc[L::m(0)]);
    }
    if (useful) store_with_ghosts<NW>(g, f, wout, c, j, out);
    // shift to row j+1
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      U0[iw] = R.U[iw];
      sP0[iw] = float(R.U[iw]) + R.dP[iw];
      F0[iw] = F1[iw];
    }
    Fmd0 = Fmd1;
    if constexpr (MHD) {
      dPbx0 = R.dP[L::b(0)];
      byU_m = byU_0; byd_m = byd_0;
      byU_0 = R.U[L::b(1)]; byd_0 = R.dP[L::b(1)];
    }
  }
  return cmx;
}

// y sweep, FP32-bulk: a marching step carries row j's base, its face-left
// increment, rounded predictor state and pressures, and the face j-1/2 flux.
// The raw rows j, j+1, j+2 of w at column px stay in registers, so a step
// loads one new row (and B_x beside the next one), one step ahead.
template <bool MHD, int LIM>
__device__ __forceinline__ double bulk_fy_body(const Geo& g, const Sel& s, const BC& bc,
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
  if (s0 > f.n1 + 1) return cmx;
  const int s1 = (s0 + seglen - 1 < f.n1 + 1) ? s0 + seglen - 1 : f.n1 + 1;
  const double* dx0 = g.dx[0];
  const double* dx1 = g.dx[1];
  const double dxp = dx0[px];
  // row r's predictor gathered from HBM ('cont' ghost rows take zero
  // corrector slopes)
  auto make = [&](int r, BulkRow<MHD, 1, LIM>& R) {
    R.make(g, f, s, w, px, ghost_image(r, f.n1, f.pylo, f.pyhi), dt, smallp);
    const bool zs = (r < 2 && !f.pylo) || (r > f.n1 + 1 && !f.pyhi);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) R.l[iw] = zs ? 0.f : R.l[iw];
  };
  auto load_row = [&](int r, double* u) {
    const unsigned o = f.at(px, r);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) u[iw] = ldp(w, g.ps, iw, o);
  };
  // B_x at (px -/+ 1, r), for the predictor's divergence
  auto load_bx = [&](int r, double* bx) {
    bx[0] = bx[1] = 0.0;
    if constexpr (MHD) {
      const unsigned o = f.at(px, r);
      bx[0] = ldp(w, g.ps, L::b(0), o - 8u);
      bx[1] = ldp(w, g.ps, L::b(0), o + 8u);
    }
  };
  // carried row state (the lower row of the next face)
  double U0[NW], pt0, ph0, Fpm, byU_m = 0.0;
  float dL0[NW], sP0[NW], Fp[NW], dPbx0 = 0.f, dPby0 = 0.f, byd_m = 0.f;
  auto carry = [&](const BulkRow<MHD, 1, LIM>& R) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      U0[iw] = R.U[iw];
      dL0[iw] = R.dP[iw] + 0.5f * R.l[iw];
      sP0[iw] = float(R.U[iw]) + R.dP[iw];
    }
    pt0 = R.ptot;
    ph0 = R.pth;
    if constexpr (MHD) {
      dPbx0 = R.dP[L::b(0)];
      dPby0 = R.dP[L::b(1)];
    }
  };
  // face between the carried row and row R above it
  auto face = [&](const BulkRow<MHD, 1, LIM>& R, double dxs, float* F, double& Fmd) {
    float dR[NW], b[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      b[iw] = float(U0[iw]);
      dR[iw] = float(R.U[iw] - U0[iw]) + R.dP[iw] - 0.5f * R.l[iw];
    }
    const float cm = face32<MHD, 1>(s, b, dL0, dR, pt0, ph0, smallp, F, Fmd);
    return pdiv(double(cm) * 2.0, dxs);
  };
  {
    BulkRow<MHD, 1, LIM> R;
    make(s0 - 1, R);
    carry(R);
    if constexpr (MHD) { byU_m = R.U[L::b(1)]; byd_m = R.dP[L::b(1)]; }
    make(s0, R);
    const double cv = face(R, dx1[s0 - 1] + dx1[s0], Fp, Fpm);
    if (useful) cmx = cmx < cv ? cv : cmx;
    carry(R);
  }
  // window: W0 = row j, W1 = row j+1, W2 = row j+2, bx beside row j+1 and
  // the widths dy0 = dy(j), dy1 = dy(j+1), all loaded one marching step
  // ahead, so no load is consumed in the step that issues it (a load waited
  // for would also wait for the prefetch: loads complete in order)
  double W0[NW], W1[NW], W2[NW], W3[NW] = {}, bx[2] = {0.0, 0.0}, bxn[2] = {0.0, 0.0};
  load_row(s0, W0);
  load_row(s0 + 1, W1);
  load_row(s0 + 2, W2);
  if (s0 + 1 <= f.n1 + 1) load_bx(s0 + 1, bx);
  double dy0 = dx1[s0], dy1 = dx1[s0 + 1];
  const double dxc = dx0[c];
  for (int j = s0; j <= s1; ++j) {
    if (j + 1 <= s1 && j + 2 <= f.n1 + 1) {  // the next step's row is a mesh row
      load_row(j + 3, W3);
      load_bx(j + 2, bxn);
    }
    const double dy2 = dx1[j + 2];
    BulkRow<MHD, 1, LIM> R;  // row j+1
    if (j + 1 <= f.n1 + 1) {
      float bd = 0.f;
      if constexpr (MHD) bd = float(bx[1] - bx[0]);
      R.build(s, W1, W0, W2, bd, dxp, dy1, dt, smallp);
    } else {
      make(j + 1, R);
    }
    float Fn[NW];
    double Fnm;
    const double cv = face(R, dy0 + dy1, Fn, Fnm);
    if (useful) cmx = cmx < cv ? cv : cmx;
    float inc[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) inc[iw] = 0.f;
    if constexpr (MHD) {
      const double bxp = lane_next(U0[L::b(0)]), bxm = lane_prev(U0[L::b(0)]);
      const float dxp_ = lane_next(dPbx0), dxm_ = lane_prev(dPbx0);
      const float gx = pdiv(0.5f * (float(bxp - bxm) + (dxp_ - dxm_)), float(dxc));
      const float gy = pdiv(0.5f * (float(R.U[L::b(1)] - byU_m) + (R.dP[L::b(1)] - byd_m)), float(dy0));
      powell32<ND>(sP0, (gx + gy) * float(dt / ND), inc);
      byU_m = U0[L::b(1)];
      byd_m = dPby0;
    }
    const double qdx = pdiv(dt, dy0);
    double out[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) out[iw] = U0[iw] - qdx * double(Fn[iw] - Fp[iw]) + double(inc[iw]);
    out[L::m(1)] = U0[L::m(1)] - qdx * (Fnm - Fpm) + double(inc[L::m(1)]);
    prefetch_done();
    if (useful) store_with_ghosts<NW>(g, f, wout, c, j, out);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) Fp[iw] = Fn[iw];
    Fpm = Fnm;
    carry(R);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
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
