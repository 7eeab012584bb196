// Device-side ideal MHD / HD physics for the advect path, one cell or one
// interface at a time.  Every expression keeps the evaluation order of the
// VAC statement it restates (cited per function), so that the FP64 build
// (compiled with -ffp-contract=off) reproduces VAC bit for bit.
//
// R is the arithmetic type of the bulk flux / limiter / eigen work (double,
// or float for the FP32-bulk tier).  Thermal and total pressure recovery,
// source terms and every update of a conserved variable stay in double.
#pragma once
#include <hip/hip_runtime.h>

namespace vacgpu {

constexpr double kBig = 1.0e99;     // bigdouble (src/vacdef.t:51)
constexpr double kSmall = 1.0e-99;  // smalldouble (src/vacdef.t:51)

enum { LIM_MINMOD = 0, LIM_WOODWARD = 1, LIM_SUPERBEE = 2 };

template <int ND, bool MHD>
struct Layout {
  static constexpr int nd = ND;           // ndim == ndir
  static constexpr int nw = MHD ? 2 + 2 * ND : 2 + ND;
  static constexpr int rho = 0;
  static constexpr int e = 1 + ND;
  __host__ __device__ static constexpr int m(int c) { return 1 + c; }
  __host__ __device__ static constexpr int b(int c) { return 2 + ND + c; }
};

template <class R>
__device__ __forceinline__ R rmax(R a, R b) { return a < b ? b : a; }
template <class R>
__device__ __forceinline__ R rmin(R a, R b) { return b < a ? b : a; }
// Fortran SIGN(a, b)
template <class R>
__device__ __forceinline__ R fsign(R a, R b) { return copysign(fabs(a), b); }

// a / b for a denominator b > 0.
//
// double: the hardware reciprocal estimate of b refined by two Newton-Raphson
// steps, then q0 = a*y and one Markstein residual correction
// q = q0 + (a - b*q0)*y (P. Markstein, IBM J. Res. Dev. 34(1), 1990).  These
// are the operations of the device's correctly rounded double division when
// that division needs neither operand scaling nor a special-value fix-up,
// i.e. for b and 1/b normal, a zero or of magnitude >= 2^-969 and a/b
// normal, so there the result is the same double.  The reciprocal depends on
// b alone, so quotients with one denominator share it (8 instructions for the
// first, 3 for each further one, against 11 for every division).  The
// residual is formed negated, which keeps the sign of a zero numerator.
// float (FP32-bulk tier only): a times the hardware reciprocal of b (about
// 1 ulp), shared like the double one; the tier's 1e-5 error budget allows it.
__device__ __forceinline__ double pdiv(double a, double b) {
  const double y0 = __builtin_amdgcn_rcp(b);
  const double y1 = __builtin_fma(y0, __builtin_fma(-b, y0, 1.0), y0);
  const double y = __builtin_fma(y1, __builtin_fma(-b, y1, 1.0), y1);
  const double q0 = a * y;
  return __builtin_fma(-__builtin_fma(b, q0, -a), y, q0);
}
__device__ __forceinline__ float pdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// sqrt(x), the same double as the compiler's correctly rounded square root.
// That root (LLVM's f64 sqrt lowering for AMDGPU) scales an input below
// 2^-767 by 2^256, takes the hardware reciprocal square root y, refines the
// root s0 = x*y and the half-reciprocal h0 = y/2 by one Goldschmidt step,
// corrects the root twice by its residual x - s^2, undoes the scaling and
// passes 0 and +inf through.  For x in [2^-767, DBL_MAX] the scaling and
// the pass-through do nothing, so the ten operations in between are the
// whole computation and give the same double; every other input (0, tiny,
// subnormal, +inf, negative, NaN) takes sqrt() itself.  The test is one
// compare pair against nine scaling and fix-up instructions saved.  Used by
// the Roe solvers (about ten roots per face); cmax keeps sqrt(): with it
// there the fused TVDLF sweeps measured about 1% slower.
// float (FP32-bulk tier only): the hardware square root, as sqrt() is there.
__device__ __forceinline__ double psqrt(double x) {
  if (__builtin_expect(x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023, 1)) {
    const double y = __builtin_amdgcn_rsq(x);
    const double s0 = x * y;
    const double h0 = y * 0.5;
    const double r0 = __builtin_fma(-h0, s0, 0.5);
    const double h1 = __builtin_fma(h0, r0, h0);
    const double s1 = __builtin_fma(s0, r0, s0);
    const double d0 = __builtin_fma(-s1, s1, x);
    const double s2 = __builtin_fma(d0, h1, s1);
    const double d1 = __builtin_fma(-s2, s2, x);
    return __builtin_fma(d1, h1, s2);
  }
  return sqrt(x);
}
__device__ __forceinline__ float psqrt(float x) { return sqrt(x); }

// dwlimiter2 (src/vacgrid.t:938-996): a = dwC(i), b = dwC(i-1).  LIM >= 0
// fixes the type at compile time (no branches); LIM < 0 takes it from type.
// max / min are the hardware v_max / v_min (one instruction, against a compare
// and two selects for rmax / rmin).  They differ from max(x, y) = x < y ? y : x
// only in the sign of a zero result when the operands are +0 and -0, and every
// such inner min feeds a max whose other operand is +0 (or a value >= +0 that
// is itself such a max), which returns +0 either way: the limited slope is the
// same double.  (NaN operands do not occur on a valid state.)
template <int LIM, class R>
__device__ __forceinline__ R limiter_t(int type, R a, R b) {
  const int t = LIM >= 0 ? LIM : type;
  const R one = R(1), zero = R(0);
  R s = copysign(one, a);
  if (t == LIM_MINMOD) return s * fmax(zero, fmin(R(fabs(a)), s * b));
  if (t == LIM_WOODWARD)
    return R(2) * s * fmax(zero, fmin(fmin(R(fabs(a)), s * b), s * R(0.25) * (b + a)));
  return s * fmax(fmax(zero, fmin(R(2) * R(fabs(a)), s * b)), fmin(R(fabs(a)), R(2) * s * b));
}

template <class R>
__device__ __forceinline__ R limiter(int type, R a, R b) {
  return limiter_t<-1, R>(type, a, b);
}

// Unclipped thermal pressure (MHD src/vacphys0.t.mhd:191-199,
// HD src/vacphys0.t.hd:118-119), always in double.
template <int ND, bool MHD>
__device__ __forceinline__ double pthermal_raw(const double* u, double g) {
  using L = Layout<ND, MHD>;
  double msq = u[L::m(0)] * u[L::m(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) msq = msq + u[L::m(c)] * u[L::m(c)];
  if constexpr (MHD) {
    double p = pdiv(msq, u[L::rho]);
#pragma unroll
    for (int c = 0; c < ND; ++c) p = p + u[L::b(c)] * u[L::b(c)];
    return (g - 1.0) * (u[L::e] - 0.5 * p);
  } else {
    return (g - 1.0) * (u[L::e] - pdiv(0.5 * msq, u[L::rho]));
  }
}

// Clipped thermal pressure (getpthermal with clipping=.true.)
template <int ND, bool MHD>
__device__ __forceinline__ double pthermal(const double* u, double g, double smallp) {
  double p = pthermal_raw<ND, MHD>(u, g);
  return p < smallp ? smallp : p;
}

// primitive (MHD src/vacphys0.t.mhd:31-52, HD src/vacphys0.t.hd:30-51):
// e -> the clipped thermal pressure, m -> v = m/rho; the rest is copied.
template <int ND, bool MHD>
__device__ __forceinline__ void to_prim(const double* u, double g, double smallp, double* q) {
  using L = Layout<ND, MHD>;
#pragma unroll
  for (int iw = 0; iw < L::nw; ++iw) q[iw] = u[iw];
  q[L::e] = pthermal<ND, MHD>(u, g, smallp);
#pragma unroll
  for (int c = 0; c < ND; ++c) q[L::m(c)] = pdiv(u[L::m(c)], u[L::rho]);
}

// conserve (MHD src/vacphys0.t.mhd:5-28, HD src/vacphys0.t.hd:5-27):
// e = p/(g-1) + (rho*v^2 + B^2)/2 (HD: rho*v^2/2), then m = rho*v; in place.
template <int ND, bool MHD>
__device__ __forceinline__ void to_cons(double* q, double g) {
  using L = Layout<ND, MHD>;
  double vsq = q[L::m(0)] * q[L::m(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) vsq = vsq + q[L::m(c)] * q[L::m(c)];
  if constexpr (MHD) {
    double bsq = q[L::b(0)] * q[L::b(0)];
#pragma unroll
    for (int c = 1; c < ND; ++c) bsq = bsq + q[L::b(c)] * q[L::b(c)];
    q[L::e] = pdiv(q[L::e], g - 1.0) + 0.5 * (q[L::rho] * vsq + bsq);
  } else {
    q[L::e] = pdiv(q[L::e], g - 1.0) + 0.5 * q[L::rho] * vsq;
  }
#pragma unroll
  for (int c = 0; c < ND; ++c) q[L::m(c)] = q[L::rho] * q[L::m(c)];
}

// With useprimitive every upwindLR call converts its base state to primitive
// variables and back in place (src/vactvdlf.t:328-332, :360-364).  The
// TVDLF / TVD-MUSCL corrector limits a state the predictor of the same sweep
// has already converted once, so its limiter sees prim(cons(prim(u))) of the
// sweep's input u (computed here in place), and the update's base is
// cons of that.
template <int ND, bool MHD>
__device__ __forceinline__ void prim_of_roundtrip(double* u, double g, double smallp) {
  double q[Layout<ND, MHD>::nw];
  to_prim<ND, MHD>(u, g, smallp, q);
  to_cons<ND, MHD>(q, g);
  to_prim<ND, MHD>(q, g, smallp, u);
}

// n of those round trips in place: a state after n upwindLR calls on it
template <int ND, bool MHD>
__device__ __forceinline__ void roundtrips(double* u, int n, double g, double smallp) {
  constexpr int NW = Layout<ND, MHD>::nw;
  for (int k = 0; k < n; ++k) {
    double q[NW];
    to_prim<ND, MHD>(u, g, smallp, q);
    to_cons<ND, MHD>(q, g);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) u[iw] = q[iw];
  }
}

// the primitive variables of a state after n round trips, in place (n = 1
// is prim_of_roundtrip)
template <int ND, bool MHD>
__device__ __forceinline__ void prim_after(double* u, int n, double g, double smallp) {
  constexpr int NW = Layout<ND, MHD>::nw;
  roundtrips<ND, MHD>(u, n, g, smallp);
  double q[NW];
  to_prim<ND, MHD>(u, g, smallp, q);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) u[iw] = q[iw];
}

// Total pressure (src/vacphys0.t.mhd:244-252), unclipped, in double.
template <int ND>
__device__ __forceinline__ double ptotal(const double* u, double g) {
  using L = Layout<ND, true>;
  double bsq = u[L::b(0)] * u[L::b(0)], msq = u[L::m(0)] * u[L::m(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) {
    bsq = bsq + u[L::b(c)] * u[L::b(c)];
    msq = msq + u[L::m(c)] * u[L::m(c)];
  }
  double p = (1.0 - 0.5 * g) * bsq;
  p = p + (g - 1.0) * u[L::e];
  p = p - pdiv(0.5 * (g - 1.0) * msq, u[L::rho]);
  return p;
}

// Physical flux plus transport flux of every variable along d
// (MHD getflux src/vacphys.mhd0.t:149-205, HD src/vacphys.t.hd:49-75;
// transport f + v*w with v = m_d/rho as in src/vactvdlf.t:67-70).
// Results are stored in double.  The flux of the normal momentum m_d, which
// carries the total (MHD) or thermal (HD) pressure, is evaluated in double
// for every R; the other fluxes in R.  For R = double every expression is
// VAC's.
template <int ND, bool MHD, class R>
__device__ __forceinline__ void fluxes(const double* u, int d, double g, double smallp, double* f) {
  using L = Layout<ND, MHD>;
  const R rho = R(u[L::rho]);
  const R v = pdiv(R(u[L::m(d)]), rho);
  const double vD = pdiv(u[L::m(d)], u[L::rho]);
  if constexpr (MHD) {
    const double ptD = ptotal<ND>(u, g);
    const R pt = R(ptD);
    const R bd = R(u[L::b(d)]);
    f[L::rho] = double(R(0) + v * rho);
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      if (c == d) {
        f[L::m(c)] = ptD - u[L::b(c)] * u[L::b(d)] + vD * u[L::m(c)];
      } else {
        const R bc = R(u[L::b(c)]);
        f[L::m(c)] = double(-(bc * bd) + v * R(u[L::m(c)]));
      }
    }
    R bm = R(u[L::b(0)]) * R(u[L::m(0)]);
#pragma unroll
    for (int c = 1; c < ND; ++c) bm = bm + R(u[L::b(c)]) * R(u[L::m(c)]);
    f[L::e] = double(pdiv(R(u[L::m(d)]) * pt - bd * bm, rho) + v * R(u[L::e]));
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      if (c == d) f[L::b(c)] = 0.0;
      else f[L::b(c)] = double(-(pdiv(R(u[L::m(c)]), rho) * bd) + v * R(u[L::b(c)]));
    }
  } else {
    const double pD = pthermal<ND, false>(u, g, smallp);
    const R p = R(pD);
    f[L::rho] = double(R(0) + v * rho);
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      if (c == d) f[L::m(c)] = pD + vD * u[L::m(c)];
      else f[L::m(c)] = double(R(0) + v * R(u[L::m(c)]));
    }
    f[L::e] = double(pdiv(R(u[L::m(d)]), rho) * p + v * R(u[L::e]));
  }
}

// cmax_d = fast (or sound) speed + |v_d| (MHD src/vacphys0.t.mhd:79-116
// with getcsound2 :140-162; HD src/vacphys0.t.hd:72-100).
template <int ND, bool MHD, class R>
__device__ __forceinline__ R cmax(const double* u, int d, double g, double smallp) {
  using L = Layout<ND, MHD>;
  const double p = pthermal<ND, MHD>(u, g, smallp);
  const R rho = R(u[L::rho]);
  if constexpr (MHD) {
    const R c2 = pdiv(R(g) * R(p), rho);
    R bsq = R(u[L::b(0)]) * R(u[L::b(0)]);
#pragma unroll
    for (int c = 1; c < ND; ++c) bsq = bsq + R(u[L::b(c)]) * R(u[L::b(c)]);
    const R cf2 = pdiv(bsq, rho) + c2;
    const R bi = R(u[L::b(d)]);
    return R(sqrt(R(0.5) * (cf2 + R(sqrt(cf2 * cf2 - pdiv(R(4) * c2 * (bi * bi), rho)))))) +
           R(fabs(pdiv(R(u[L::m(d)]), rho)));
  } else {
    const R cs = R(sqrt(pdiv(R(g) * R(p), rho)));
    return cs + R(fabs(pdiv(R(u[L::m(d)]), rho)));
  }
}

// Powell divB source (src/vacphys.mhd0.t:246-280): subtracts qdt*divB*(...)
// from out, with divB from the slab centred gradient (src/vacgrid.t:816-823)
// of the state at the cell (uc) and its +-1 neighbours along each dim.
template <int ND>
__device__ __forceinline__ void powell(const double* uc, double divb_raw, double qdt, double* out) {
  using L = Layout<ND, true>;
  const double divb = qdt * divb_raw;
#pragma unroll
  for (int c = 0; c < ND; ++c) out[L::m(c)] = out[L::m(c)] - uc[L::b(c)] * divb;
#pragma unroll
  for (int c = 0; c < ND; ++c)
    out[L::b(c)] = out[L::b(c)] - pdiv(uc[L::m(c)], uc[L::rho]) * divb;
  double mb = uc[L::m(0)] * uc[L::b(0)];
#pragma unroll
  for (int c = 1; c < ND; ++c) mb = mb + uc[L::m(c)] * uc[L::b(c)];
  out[L::e] = out[L::e] - pdiv(mb, uc[L::rho]) * divb;
}

// ---------------------------------------------------------------- Roe (MHD)
// Interface state of the eight-wave MHD Roe solver (src/vacphys.mhdroe.t).
template <int ND, class R>
struct Roe {
  R sq;          // sqrt of the arithmetic mean density
  R v[ND];       // mean velocity
  R bw[ND];      // B_d (clamped) at slot d, beta at the transverse slots
  R cs2, cf, cs, af, as;
  R dp, rhodv;
  R bdvf, bdbf;  // fast/slow-wave projections (il == fastRW)
  R bdva, bdba;  // Alfven-wave projections (ND == 3)
  R v2a2, bvf, bva = R(0);  // rtimes helpers for the energy row
};

// The part of the interface state that rtimes needs (sqcs2 = sqrt(cs2) as
// computed inline at src/vacphys.mhdroe.t:430-434).  All members have type R,
// so a RoeRt can be moved between lanes member by member.
template <int ND, class R>
struct RoeRt {
  R sq;
  R v[ND];
  R bw[ND];
  R cs2, sqcs2, cf, cs, af, as;
  R v2a2, bvf, bva;
};

// average + the parts of geteigenjump that do not depend on il
// (src/vacphys.mhdroe.t:4-122 and :156-186), plus the rtimes helpers
// (:320-345).  uL, uR are the conserved states left and right of the face.
// prim: useprimitive, where dp = pR - pL (src/vacphys.mhdroe.t:42-44).
template <int ND, class R>
__device__ __forceinline__ void roe_setup(const double* uL, const double* uR, int d, double g,
                                          double smallp, Roe<ND, R>& r, int prim = 0) {
  using L = Layout<ND, true>;
  const R rL = R(uL[L::rho]), rR = R(uR[L::rho]);
  const R rho = R(0.5) * (rL + rR);
  R B[ND];
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    r.v[c] = R(0.5) * (pdiv(R(uL[L::m(c)]), rL) + pdiv(R(uR[L::m(c)]), rR));
    B[c] = R(0.5) * (R(uL[L::b(c)]) + R(uR[L::b(c)]));
  }
  const double pL = pthermal<ND, true>(uL, g, smallp);
  const double pR = pthermal<ND, true>(uR, g, smallp);
  const R pbar = R(0.5 * (pL + pR));
  R s1 = r.v[0] * (R(uR[L::m(0)]) - R(uL[L::m(0)]));
  R s2 = r.v[0] * r.v[0];
  R s3 = R(uR[L::b(0)]) * R(uR[L::b(0)]) - R(uL[L::b(0)]) * R(uL[L::b(0)]);
#pragma unroll
  for (int c = 1; c < ND; ++c) {
    s1 = s1 + r.v[c] * (R(uR[L::m(c)]) - R(uL[L::m(c)]));
    s2 = s2 + r.v[c] * r.v[c];
    s3 = s3 + R(uR[L::b(c)]) * R(uR[L::b(c)]) - R(uL[L::b(c)]) * R(uL[L::b(c)]);
  }
  const R drho = rR - rL;
  r.dp = prim ? R(pR - pL)
              : (R(g) - R(1)) * (R(uR[L::e]) - R(uL[L::e]) - s1 + R(0.5) * s2 * drho - R(0.5) * s3);
  r.rhodv = R(uR[L::m(d)]) - R(uL[L::m(d)]) - r.v[d] * drho;
  const R c2 = pdiv(R(g) * pbar, rho);
  r.cs2 = c2;
  R bsq = B[0] * B[0];
#pragma unroll
  for (int c = 1; c < ND; ++c) bsq = bsq + B[c] * B[c];
  const R bi = B[d];
  R cf = pdiv(bsq, rho) + c2;
  R cs = rmax(R(0), R(0.5) * (cf - R(psqrt(cf * cf - pdiv(R(4) * c2 * (bi * bi), rho)))));
  cf = cf - cs;
  R af = (c2 - cs) / (cf - cs);
  af = rmin(R(1), rmax(af, R(0)));
  r.as = R(psqrt(R(1) - af));
  r.af = R(psqrt(af));
  r.cf = R(psqrt(cf));
  r.cs = R(psqrt(cs));
  r.sq = R(psqrt(rho));
  r.bw[d] = (R(fabs(bi)) < R(kSmall)) ? R(kSmall) : bi;
  const int idir = (d + 1) % ND;
  const int jdir = (idir + 1) % ND;
  if constexpr (ND == 2) {
    r.bw[idir] = (B[idir] >= R(0)) ? R(1) : R(-1);
  } else {
    const R t = R(psqrt(B[idir] * B[idir] + B[jdir] * B[jdir]));
    if (t > R(kSmall)) {
      r.bw[idir] = pdiv(B[idir], t);
      r.bw[jdir] = pdiv(B[jdir], t);
    } else {
      r.bw[idir] = R(sqrt(R(0.5)));
      r.bw[jdir] = R(sqrt(R(0.5)));
    }
  }
  // fast/slow projections (il == fastRW_, :159-168)
  {
    R t = r.bw[idir] * (pdiv(R(uR[L::m(idir)]), rR) - pdiv(R(uL[L::m(idir)]), rL));
    if constexpr (ND == 3) t = t + r.bw[jdir] * (pdiv(R(uR[L::m(jdir)]), rR) - pdiv(R(uL[L::m(jdir)]), rL));
    r.bdvf = t * fsign(r.sq * r.sq, r.bw[d]);
    R u = r.bw[idir] * (R(uR[L::b(idir)]) - R(uL[L::b(idir)]));
    if constexpr (ND == 3) u = u + r.bw[jdir] * (R(uR[L::b(jdir)]) - R(uL[L::b(jdir)]));
    r.bdbf = u * R(psqrt(r.cs2)) * r.sq;
  }
  if constexpr (ND == 3) {  // Alfven projections (il == alfvRW_, :176-183)
    const R bI = r.bw[idir], bJ = r.bw[jdir];
    R t = bJ * (pdiv(R(uR[L::m(idir)]), rR) - pdiv(R(uL[L::m(idir)]), rL)) -
          bI * (pdiv(R(uR[L::m(jdir)]), rR) - pdiv(R(uL[L::m(jdir)]), rL));
    R u = bJ * (R(uR[L::b(idir)]) - R(uL[L::b(idir)])) - bI * (R(uR[L::b(jdir)]) - R(uL[L::b(jdir)]));
    r.bdva = t * R(0.5) * (r.sq * r.sq);
    r.bdba = u * R(0.5) * fsign(r.sq, r.bw[d]);
    r.bva = bJ * r.v[idir] - bI * r.v[jdir];
  }
  // rtimes helpers for the energy row (:324-336)
  R vs = r.v[0] * r.v[0];
#pragma unroll
  for (int c = 1; c < ND; ++c) vs = vs + r.v[c] * r.v[c];
  r.v2a2 = R(0.5) * vs + pdiv(R(2) - R(g), R(g) - R(1)) * r.cs2;
  {
    R t = r.bw[idir] * r.v[idir];
    if constexpr (ND == 3) t = t + r.bw[jdir] * r.v[jdir];
    r.bvf = t * copysign(R(1), r.bw[d]);
  }
}

template <int ND, class R>
__device__ __forceinline__ RoeRt<ND, R> roe_rt(const Roe<ND, R>& r) {
  RoeRt<ND, R> t;
  t.sq = r.sq;
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    t.v[c] = r.v[c];
    t.bw[c] = r.bw[c];
  }
  t.cs2 = r.cs2;
  t.sqcs2 = R(psqrt(r.cs2));
  t.cf = r.cf;
  t.cs = r.cs;
  t.af = r.af;
  t.as = r.as;
  t.v2a2 = r.v2a2;
  t.bvf = r.bvf;
  t.bva = (ND == 3) ? r.bva : R(0);
  return t;
}

// Wave speed a and characteristic jump of wave il (geteigenjump
// src/vacphys.mhdroe.t:188-226, entropy fix 'nul').
template <int ND, class R>
__device__ __forceinline__ void roe_wave(const Roe<ND, R>& r, const double* uL, const double* uR,
                                         int il, int d, int divbwave, R& a, R& jump) {
  using L = Layout<ND, true>;
  const R vd = r.v[d];
  switch (il) {
    case 0:
      a = vd + r.cf;
      jump = pdiv(R(0.5), r.cs2) * (r.af * (+r.cf * r.rhodv + r.dp) + r.as * (-r.cs * r.bdvf + r.bdbf));
      break;
    case 1:
      a = vd - r.cf;
      jump = pdiv(R(0.5), r.cs2) * (r.af * (-r.cf * r.rhodv + r.dp) + r.as * (+r.cs * r.bdvf + r.bdbf));
      break;
    case 2:
      a = vd + r.cs;
      jump = pdiv(R(0.5), r.cs2) * (r.as * (+r.cs * r.rhodv + r.dp) + r.af * (+r.cf * r.bdvf - r.bdbf));
      break;
    case 3:
      a = vd - r.cs;
      jump = pdiv(R(0.5), r.cs2) * (r.as * (-r.cs * r.rhodv + r.dp) + r.af * (-r.cf * r.bdvf - r.bdbf));
      break;
    case 4:
      a = vd;
      jump = R(uR[L::rho]) - R(uL[L::rho]) - pdiv(r.dp, r.cs2);
      break;
    case 5:
      if (divbwave) {
        a = vd;
        jump = R(uR[L::b(d)]) - R(uL[L::b(d)]);
      } else {
        a = R(0);
        jump = R(0);
      }
      break;
    case 6:
      a = vd + pdiv(R(fabs(r.bw[d])), r.sq);
      jump = +r.bdva - r.bdba;
      break;
    default:
      a = vd - pdiv(R(fabs(r.bw[d])), r.sq);
      jump = -r.bdva - r.bdba;
      break;
  }
}

// R(iw, il) * q (rtimes, src/vacphys.mhdroe.t:287-450)
template <int ND, class R>
__device__ __forceinline__ R roe_rtimes(const RoeRt<ND, R>& r, R q, int iw, int il, int d, int divbwave) {
  using L = Layout<ND, true>;
  const int idir = (d + 1) % ND;
  const int jdir = (idir + 1) % ND;
  const R cf = r.cf, cs = r.cs, af = r.af, as = r.as;
  if (iw == L::rho) {
    if (il <= 1) return q * af;
    if (il <= 3) return q * as;
    if (il == 4) return q;
    return R(0);
  }
  if (iw == L::e) {
    const R vd = r.v[d];
    switch (il) {
      case 0: return q * (-as * cs * r.bvf + af * (r.v2a2 + cf * (cf + vd)));
      case 1: return q * (+as * cs * r.bvf + af * (r.v2a2 + cf * (cf - vd)));
      case 2: return q * (+af * cf * r.bvf + as * (r.v2a2 + cs * (cs + vd)));
      case 3: return q * (-af * cf * r.bvf + as * (r.v2a2 + cs * (cs - vd)));
      case 4: {
        R vs = r.v[0] * r.v[0];
#pragma unroll
        for (int c = 1; c < ND; ++c) vs = vs + r.v[c] * r.v[c];
        return q * R(0.5) * vs;
      }
      case 5: return divbwave ? q * r.bw[d] : R(0);
      case 6: return +q * r.bva;
      default: return -q * r.bva;
    }
  }
  if (iw < L::e) {  // momentum
    const int c = iw - 1;
    const R vw = r.v[c];
    if (c == d) {
      switch (il) {
        case 0: return q * af * (vw + cf);
        case 1: return q * af * (vw - cf);
        case 2: return q * as * (vw + cs);
        case 3: return q * as * (vw - cs);
        case 4: return q * vw;
        default: return R(0);
      }
    }
    const R beta = r.bw[c];
    const R sg = copysign(R(1), r.bw[d]);
    switch (il) {
      case 0: return q * (af * vw - as * cs * beta * sg);
      case 1: return q * (af * vw + as * cs * beta * sg);
      case 2: return q * (as * vw + af * cf * beta * sg);
      case 3: return q * (as * vw - af * cf * beta * sg);
      case 4: return q * vw;
      case 5: return R(0);
      case 6: return (c == idir) ? +q * r.bw[jdir] : -q * r.bw[idir];
      default: return (c == idir) ? -q * r.bw[jdir] : +q * r.bw[idir];
    }
  }
  // magnetic field
  const int c = iw - (2 + ND);
  if (c == d) return (il == 5 && divbwave) ? q : R(0);
  const R beta = r.bw[c];
  switch (il) {
    case 0:
    case 1: return +pdiv(q * as * r.sqcs2 * beta, r.sq);
    case 2:
    case 3: return -pdiv(q * af * r.sqcs2 * beta, r.sq);
    case 4:
    case 5: return R(0);
    default:
      // x / SIGN(sq, bw_d) == +-(x / sq): IEEE division is odd in its divisor
      if (c == idir) return -(signbit(r.bw[d]) ? -pdiv(q * r.bw[jdir], r.sq) : pdiv(q * r.bw[jdir], r.sq));
      return signbit(r.bw[d]) ? -pdiv(q * r.bw[idir], r.sq) : pdiv(q * r.bw[idir], r.sq);
  }
}

// ---- HD Roe solver (src/vacphys.hdroe.t) for TVD and TVD-MUSCL.
// Waves: 0 sound right, 1 sound left, 2 entropy, 3.. shear (direction il-3,
// the sweep direction skipped).
template <int ND, class R>
struct RoeH {
  R rho, v[ND], h, c, dpc, dvc;  // sqrt(rhoL*rhoR), v, h, c, (pR-pL)/c^2, (vR-vL)/c
};
template <int ND, class R>
struct RoeRtH {
  R v[ND];
  R h, c, vsq;
};

// average (:4-46) and the il == 1 part of geteigenjump (:66-92).  uL, uR are
// the conserved states left and right of the face.
template <int ND, class R>
__device__ __forceinline__ void roe_setup(const double* uL, const double* uR, int d, double g,
                                          double smallp, RoeH<ND, R>& r, int /*prim: no effect*/ = 0) {
  using L = Layout<ND, false>;
  const R rL = R(uL[L::rho]), rR = R(uR[L::rho]);
  R t = R(psqrt(rL));
  const R t2 = R(psqrt(rR));
  r.rho = t * t2;
  t = pdiv(t, t2);  // sqrt(rhoL/rhoR)
  const R w1 = R(1) + t;
#pragma unroll
  for (int c = 0; c < ND; ++c)
    r.v[c] = pdiv(pdiv(R(uL[L::m(c)]), rL) * t + pdiv(R(uR[L::m(c)]), rR), w1);
  const R pL = R(pthermal<ND, false>(uL, g, smallp));
  const R pR = R(pthermal<ND, false>(uR, g, smallp));
  const R hL = pdiv(pL + R(uL[L::e]), rL);
  const R hR = pdiv(pR + R(uR[L::e]), rR);
  r.h = pdiv(hL * t + hR, w1);
  R vs = r.v[0] * r.v[0];
#pragma unroll
  for (int c = 1; c < ND; ++c) vs = vs + r.v[c] * r.v[c];
  R c2 = (R(g) - R(1)) * (r.h - R(0.5) * vs);
  c2 = rmax(c2, pdiv(R(g) * R(smallp), r.rho));
  r.dpc = pdiv(pR - pL, c2);
  r.c = R(psqrt(c2));
  r.dvc = pdiv(pdiv(R(uR[L::m(d)]), rR) - pdiv(R(uL[L::m(d)]), rL), r.c);
}

template <int ND, class R>
__device__ __forceinline__ RoeRtH<ND, R> roe_rt(const RoeH<ND, R>& r) {
  RoeRtH<ND, R> t;
#pragma unroll
  for (int c = 0; c < ND; ++c) t.v[c] = r.v[c];
  t.h = r.h;
  t.c = r.c;
  R vs = r.v[0] * r.v[0];
#pragma unroll
  for (int c = 1; c < ND; ++c) vs = vs + r.v[c] * r.v[c];
  t.vsq = vs;
  return t;
}

// the direction of shear wave il (il >= 3) for a sweep along d
__device__ __forceinline__ int shear_dir(int il, int d) {
  const int k = il - 3;
  return k >= d ? k + 1 : k;
}

// wave speed a and characteristic jump of wave il (geteigenjump :94-118,
// entropy fix 'nul')
template <int ND, class R>
__device__ __forceinline__ void roe_wave(const RoeH<ND, R>& r, const double* uL, const double* uR,
                                         int il, int d, int divbwave, R& a, R& jump) {
  using L = Layout<ND, false>;
  (void)divbwave;
  const R vd = r.v[d];
  switch (il) {
    case 0:
      a = vd + r.c;
      jump = R(0.5) * (r.dpc + r.rho * r.dvc);
      break;
    case 1:
      a = vd - r.c;
      jump = R(0.5) * (r.dpc - r.rho * r.dvc);
      break;
    case 2:
      a = vd;
      jump = -r.dpc + R(uR[L::rho]) - R(uL[L::rho]);
      break;
    default: {
      const int k = shear_dir(il, d);
      a = vd;
      jump = r.rho * (pdiv(R(uR[L::m(k)]), R(uR[L::rho])) - pdiv(R(uL[L::m(k)]), R(uL[L::rho])));
    }
  }
}

// R(iw, il) * q (rtimes :152-215)
template <int ND, class R>
__device__ __forceinline__ R roe_rtimes(const RoeRtH<ND, R>& r, R q, int iw, int il, int d,
                                        int divbwave) {
  using L = Layout<ND, false>;
  (void)divbwave;
  const bool shear = il >= 3;
  const int k = shear ? shear_dir(il, d) : 0;
  if (iw == L::rho) return shear ? R(0) : q;
  const R vd = r.v[d];
  if (iw == L::e) {
    switch (il) {
      case 0: return q * (r.h + vd * r.c);
      case 1: return q * (r.h - vd * r.c);
      case 2: return q * R(0.5) * r.vsq;
      default: return q * r.v[k];
    }
  }
  const int c = iw - 1;  // momentum component
  if (c == d) {
    switch (il) {
      case 0: return q * (vd + r.c);
      case 1: return q * (vd - r.c);
      case 2: return q * vd;
      default: return R(0);
    }
  }
  if (shear) return (c == k) ? q : R(0);
  return q * r.v[c];
}

// the Roe state types of the MHD and HD solvers
template <int ND, class R, bool MHD>
struct RoeOf {
  using St = Roe<ND, R>;
  using Rt = RoeRt<ND, R>;
};
template <int ND, class R>
struct RoeOf<ND, R, false> {
  using St = RoeH<ND, R>;
  using Rt = RoeRtH<ND, R>;
};

// getphi for method 'tvd', typetvd 'roe', entropy fix 'nul'
// (src/vactvd.t:246-247, :294-302): jm, j0, jp are the jumps of one wave at
// faces J-1, J, J+1 and av = a*dt*2/(dx_J+dx_J+1) at face J.
template <int LIM, class R>
__device__ __forceinline__ R tvd_phi_ph(int lim, R jm, R j0, R jp, R av, R ph) {
  const R ljl = limiter_t<LIM, R>(lim, j0, jm);  // ljump(J)
  const R ljr = limiter_t<LIM, R>(lim, jp, j0);  // ljump(J+1)
  if (av <= R(0)) ph = ph * (j0 - ljr);
  else ph = ph * (j0 - ljl);
  return ph + av * av * j0;
}
template <int LIM, class R>
__device__ __forceinline__ R tvd_phi_t(int lim, R jm, R j0, R jp, R av) {
  return tvd_phi_ph<LIM, R>(lim, jm, j0, jp, av, R(fabs(av)) - av * av);
}

// ---- entropy fixes (typeentropy, src/vactvd.t:422-449)
enum { ENT_NUL = 0, ENT_HARTEN = 1, ENT_POWELL = 2, ENT_RATIO = 3, ENT_YEE = 4 };

// getphi's first factors with the fix (src/vactvd.t:227-256): where |a| is
// below smalla, |a| becomes (smalla + a^2/smalla)/2 ('tvdmu') and |a| - a^2
// becomes smalla/2 + (1/(2 smalla) - 1) a^2 ('tvd'); 'nul' and 'ratio' keep them
template <class R>
__device__ __forceinline__ R ent_abs(R av, R sa, int type) {
  if (type != ENT_NUL && type != ENT_RATIO && !(R(fabs(av)) >= sa))
    return R(0.5) * (sa + pdiv(av * av, sa));
  return R(fabs(av));
}
template <class R>
__device__ __forceinline__ R ent_tvd(R av, R sa, int type) {
  if (type != ENT_NUL && type != ENT_RATIO && !(R(fabs(av)) >= sa))
    return R(0.5) * sa + (pdiv(R(0.5), sa) - R(1)) * (av * av);
  return R(fabs(av)) - av * av;
}

// The left and right speeds of wave il for the entropy fix: the velocity
// along d plus or minus the fast, slow or Alfven speed of each state (MHD,
// src/vacphys.mhdroe.t:236-279) or the sound speed (HD,
// src/vacphys.hdroe.t:124-145).  The fast right wave (il 0) sets the sound
// speed quantities the other fast and slow waves reuse (c2 = the
// discriminant, ca = cs^2 + ca^2); the right Alfven wave (il 6) stores the
// Alfven speeds in ca for the left one, as VAC's saved arrays do.
template <int ND, class R, bool MHD>
__device__ __forceinline__ void ent_speeds(const double* uL, const double* uR, int il, int d,
                                           double g, double smallp, R& aL, R& aR, R& c2L,
                                           R& c2R, R& caL, R& caR) {
  using L = Layout<ND, MHD>;
  const R rL = R(uL[L::rho]), rR = R(uR[L::rho]);
  aL = pdiv(R(uL[L::m(d)]), rL);
  aR = pdiv(R(uR[L::m(d)]), rR);
  if constexpr (MHD) {
    if (il == 0) {
      c2L = pdiv(R(g) * R(pthermal<ND, true>(uL, g, smallp)), rL);
      c2R = pdiv(R(g) * R(pthermal<ND, true>(uR, g, smallp)), rR);
      R bl = R(uL[L::b(0)]) * R(uL[L::b(0)]), br = R(uR[L::b(0)]) * R(uR[L::b(0)]);
#pragma unroll
      for (int c = 1; c < ND; ++c) {
        bl = bl + R(uL[L::b(c)]) * R(uL[L::b(c)]);
        br = br + R(uR[L::b(c)]) * R(uR[L::b(c)]);
      }
      caL = c2L + pdiv(bl, rL);
      caR = c2R + pdiv(br, rR);
      const R bdL = R(uL[L::b(d)]), bdR = R(uR[L::b(d)]);
      c2L = R(psqrt(caL * caL - pdiv(R(4) * c2L * (bdL * bdL), rL)));
      c2R = R(psqrt(caR * caR - pdiv(R(4) * c2R * (bdR * bdR), rR)));
    }
    if (il == 6) {
      caL = pdiv(R(fabs(R(uL[L::b(d)]))), R(psqrt(rL)));
      caR = pdiv(R(fabs(R(uR[L::b(d)]))), R(psqrt(rR)));
    }
    switch (il) {
      case 0: aL = aL + R(psqrt(R(0.5) * (caL + c2L))); aR = aR + R(psqrt(R(0.5) * (caR + c2R))); break;
      case 1: aL = aL - R(psqrt(R(0.5) * (caL + c2L))); aR = aR - R(psqrt(R(0.5) * (caR + c2R))); break;
      case 2: aL = aL + R(psqrt(R(0.5) * (caL - c2L))); aR = aR + R(psqrt(R(0.5) * (caR - c2R))); break;
      case 3: aL = aL - R(psqrt(R(0.5) * (caL - c2L))); aR = aR - R(psqrt(R(0.5) * (caR - c2R))); break;
      case 6: aL = aL + caL; aR = aR + caR; break;
      case 7: aL = aL - caL; aR = aR - caR; break;
      default: break;
    }
  } else {
    (void)c2L; (void)c2R; (void)caL; (void)caR;
    if (il <= 1) {
      const R cl = R(psqrt(pdiv(R(g) * R(pthermal<ND, false>(uL, g, smallp)), rL)));
      const R cr = R(psqrt(pdiv(R(g) * R(pthermal<ND, false>(uR, g, smallp)), rR)));
      if (il == 0) { aL = aL + cl; aR = aR + cr; }
      else { aL = aL - cl; aR = aR - cr; }
    }
  }
}

template <class R>
__device__ __forceinline__ R tvd_phi(int lim, R jm, R j0, R jp, R av) {
  return tvd_phi_t<-1, R>(lim, jm, j0, jp, av);
}

}  // namespace vacgpu
