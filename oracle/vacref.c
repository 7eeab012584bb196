/*
 * oracle/vacref.c -- CPU restatement of VAC 4.52's explicit advect path
 * (TVDLF with the Hancock predictor, and one-step TVD with the Roe solver)
 * for ideal MHD and HD on slab Cartesian grids.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker the parity tests hold the
 * HIP path against; the product library never links or calls it.
 *
 * It is written as VAC is: every routine is a sequence of passes over index
 * regions of full-grid arrays, and every arithmetic expression keeps the
 * left-to-right evaluation order of the Fortran statement it restates, so
 * that with FMA contraction off (-ffp-contract=off on both sides) it agrees
 * bitwise with a VAC build on the same input.  Each function names the
 * reference lines it follows.  The equalities that pin it are listed in
 * tests/test_oracle.py (reference VAC fixtures, roetest identity, limiter
 * closed forms, invariants).
 *
 * Index convention: 0-based; ixG = [0, nx+3], mesh ixM = [2, nx+1].
 */
#include "vacref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BIGD 1.0e99     /* bigdouble, src/vacdef.t:51 */
#define SMALLD 1.0e-99  /* smalldouble, src/vacdef.t:51 */

typedef struct {
  const vacref_params* p;
  vacref_state* s;
  int ndim, ndir, nw, mhd;
  int nG[3];
  long st[3], nc;
  const double* dx[3];
  int mlo[3], mhi[3];  /* mesh */
  int glo[3], ghi[3];  /* grid */
  int istep, nstep;
  int firstsweep, lastsweep;
  /* work arrays (nw planes unless noted) */
  double *w1, *wold, *wLC, *wRC, *wroe, *tmpw;
  double *f, *fC, *sc1, *sc2, *sc3, *sc4, *sc5; /* 1 plane each */
  double *jump, *adtdx, *phi, *rphi, *ljump;    /* 1 plane each */
  /* Roe common block (src/vacphys.mhdroe.t:17-18) and saved arrays */
  double *cfast, *cslow, *afast, *aslow, *csound2, *dp, *rhodv;
  double *bdv, *bdb, *bv, *v2a2;
  /* projectb / Poisson work planes: divb (the residual), phi, CG's Z and A.Z,
   * gradients */
  double *pr_rhs, *pr_phi, *pr_z, *pr_az, *pr_qdx, *pr_qddx;
  /* entropy fixes: smalla, the left / right wave speeds, and the saved
   * sound speeds of geteigenjump (src/vacphys.mhdroe.t:147-148) */
  double *smalla, *eaL, *eaR, *cs2L, *cs2R, *cs2ca2L, *cs2ca2R;
  /* 'fixed' boundaries: the state the first getboundary call left */
  double* fixw;
  int fix_ready;
} ctx_t;

/* variable indices (0-based), src/vacpar.t.mhd:9-11, src/vacpar.t.hd:9-10 */
#define RHO 0
#define M(c) (1 + (c))
#define E_(x) (1 + (x)->ndir)
#define B(x, c) (2 + (x)->ndir + (c))

#define PL(x, w, iw) ((w) + (long)(iw) * (x)->nc)

#define LOOP(lo, hi)                               \
  for (int k = (lo)[2]; k <= (hi)[2]; ++k)         \
    for (int j = (lo)[1]; j <= (hi)[1]; ++j)       \
      for (int i = (lo)[0]; i <= (hi)[0]; ++i)
#define IDX(x) ((long)i + (long)j * (x)->st[1] + (long)k * (x)->st[2])
/* coordinate of the current cell along dimension d */
#define CRD(d) ((d) == 0 ? i : ((d) == 1 ? j : k))

static inline double dmax(double a, double b) { return a < b ? b : a; }
static inline double dmin(double a, double b) { return b < a ? b : a; }
/* Fortran SIGN(a,b) */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

static inline double dvol(const ctx_t* x, long idx) {
  /* dvolume = dx1*dx2[*dx3] (src/vacgrid.t:574-575) */
  long r = idx;
  int c[3] = {0, 0, 0};
  for (int d = x->ndim - 1; d >= 0; --d) { c[d] = (int)(r / x->st[d]); r %= x->st[d]; }
  double v = x->dx[0][c[0]];
  for (int d = 1; d < x->ndim; ++d) v = v * x->dx[d][c[d]];
  return v;
}

static inline double dxat(const ctx_t* x, int d, long idx) {
  long r = idx;
  for (int dd = x->ndim - 1; dd > d; --dd) r %= x->st[dd];
  return x->dx[d][(int)(r / x->st[d])];
}

/* ---------------------------------------------------------------- limiter */
/* dwlimiter2 (src/vacgrid.t:938-996) on a = dwC(i), b = dwC(i-1) */
double vacref_limit(int type, double a, double b) {
  double s = copysign(1.0, a);  /* tmp=sign(one,dwC) */
  switch (type) {
    case VACREF_MINMOD:  /* :968-971 */
      return s * dmax(0.0, dmin(fabs(a), s * b));
    case VACREF_WOODWARD: /* :972-976 */
      return 2.0 * s * dmax(0.0, dmin(dmin(fabs(a), s * b), s * 0.25 * (b + a)));
    case VACREF_SUPERBEE: /* :977-981 */
      return s * dmax(dmax(0.0, dmin(2.0 * fabs(a), s * b)), dmin(fabs(a), 2.0 * s * b));
  }
  return 0.0;
}

/* ------------------------------------------------------------ boundaries */
/* getboundary for 'cont', 'periodic', 'symm', 'asymm' and 'fixed'
 * (src/vacgrid.t:195-413, slab grids): regions x-low, x-high, y-low, ...
 * each spanning the full transverse ixG range.  'symm' / 'asymm' reflect the
 * mesh into the ghost layers (times -1 for 'asymm', :317-352); 'fixed' copies
 * like 'cont' on the first call and keeps the values that call left ever
 * after (fixB, :249-257 and :372-404).  'cont1' extrapolates linearly from
 * the edge cell e and the next one f, (|c-e|+1)*w(e) - |c-e|*w(f)
 * (:274-309); 'fixed1' and 'grad1' extrapolate so on the first call, then
 * 'fixed1' keeps that call's values and 'grad1' its differences to the edge,
 * w(e) + (fix(c) - fix(e)) (fixB holds w(c) - w(e) of the first call,
 * :388-398). */
static void getboundary_vars(ctx_t* x, double* w, int iw0, int iw1) {
  for (int d = 0; d < x->ndim; ++d) {
    for (int up = 0; up < 2; ++up) {
      int iB = 2 * d + up;
      int lo[3], hi[3];
      for (int dd = 0; dd < 3; ++dd) { lo[dd] = x->glo[dd]; hi[dd] = x->ghi[dd]; }
      if (up) { lo[d] = x->mhi[d] + 1; hi[d] = x->ghi[d]; }
      else { lo[d] = x->glo[d]; hi[d] = x->mlo[d] - 1; }
      int n = x->p->nx[d];
      for (int iw = iw0; iw < iw1; ++iw) {
        double* pw = PL(x, w, iw);
        int bc = x->p->bc[iB][iw];
        const double* fx = PL(x, x->fixw, iw);
        LOOP(lo, hi) {
          int c = CRD(d), src;
          long idx = IDX(x);
          if ((bc == VACREF_FIXED || bc == VACREF_FIXED1) && x->fix_ready) { pw[idx] = fx[idx]; continue; }
          if (bc == VACREF_CONT1 || bc == VACREF_FIXED1 || bc == VACREF_GRAD1) {
            /* ixe, ixf: the edge cell and the next one inward (:281-286) */
            const int e = up ? x->mhi[d] : x->mlo[d], f = up ? x->mhi[d] - 1 : x->mlo[d] + 1;
            const long eidx = idx + (long)(e - c) * x->st[d], fidx = idx + (long)(f - c) * x->st[d];
            if (bc == VACREF_GRAD1 && x->fix_ready) {
              pw[idx] = (fx[idx] - fx[eidx]) + pw[eidx];
            } else {
              const int a = c > e ? c - e : e - c;
              pw[idx] = (double)(a + 1) * pw[eidx] - (double)a * pw[fidx];
            }
            continue;
          }
          double sg = 1.0;
          if (bc == VACREF_PERIODIC) src = up ? c - n : c + n;   /* :311-313, pairs :107-120 */
          else if (bc == VACREF_SYMM || bc == VACREF_SYMM0 || bc == VACREF_ASYMM) {
            /* ixe = 2*ixmin-1 (upper) or 2*ixmax+1 (lower), :329-333 */
            src = up ? 2 * (x->mhi[d] + 1) - 1 - c : 2 * (x->mlo[d] - 1) + 1 - c;
            if (bc == VACREF_ASYMM) sg = -1.0;
          } else src = up ? x->mhi[d] : x->mlo[d];                /* 'cont', 'fixed' :251-273 */
          long sidx = idx + (long)(src - c) * x->st[d];
          pw[idx] = sg == 1.0 ? pw[sidx] : -pw[sidx];
        }
      }
    }
  }
  /* the first call stores what it left for the 'fixed' regions (:372-404) */
  if (x->fixw && !x->fix_ready) {
    memcpy(x->fixw, w, sizeof(double) * x->nc * x->nw);
    x->fix_ready = 1;
  }
}

static void getboundary(ctx_t* x, double* w) { getboundary_vars(x, w, 0, x->nw); }

/* ------------------------------------------------------------- physics */
/* getpthermal (MHD src/vacphys0.t.mhd:165-222, HD src/vacphys0.t.hd:103-134) */
static void getpthermal(ctx_t* x, int clip, const double* w, const int* lo,
                        const int* hi, double* p) {
  const double g = x->p->gamma;
  const double* rho = PL(x, w, RHO);
  const double* e = PL(x, w, E_(x));
  LOOP(lo, hi) {
    long idx = IDX(x);
    double msq = 0.0;
    for (int c = 0; c < x->ndir; ++c) {
      double m = PL(x, w, M(c))[idx];
      msq = (c == 0) ? m * m : msq + m * m;
    }
    double pv;
    if (x->mhd) {
      pv = msq / rho[idx];
      for (int c = 0; c < x->ndir; ++c) {
        double b = PL(x, w, B(x, c))[idx];
        pv = pv + b * b;
      }
      pv = (g - 1.0) * (e[idx] - 0.5 * pv);
    } else {
      pv = (g - 1.0) * (e[idx] - 0.5 * msq / rho[idx]);
    }
    p[idx] = pv;
  }
  if (x->s->smallp < 0.0) {
    double mn = BIGD, mx = -BIGD;
    LOOP(lo, hi) { long idx = IDX(x); mn = dmin(mn, p[idx]); mx = dmax(mx, p[idx]); }
    x->s->smallp = mn * x->p->smallpcoeff;
    if (x->mhd && x->s->smallp < 0.0) x->s->smallp = mx * x->p->smallpcoeff;
  }
  if (clip) {
    const double sp = x->s->smallp;
    LOOP(lo, hi) { long idx = IDX(x); p[idx] = dmax(p[idx], sp); }
  }
}

/* primitive (MHD src/vacphys0.t.mhd:31-52, HD src/vacphys0.t.hd:30-51):
 * e -> p (clipped getpthermal), m -> v = m/rho, in place over lo..hi */
static void primitive(ctx_t* x, double* w, const int* lo, const int* hi) {
  double* p = x->sc5;
  getpthermal(x, 1, w, lo, hi, p);
  double* e = PL(x, w, E_(x));
  const double* rho = PL(x, w, RHO);
  LOOP(lo, hi) { long idx = IDX(x); e[idx] = p[idx]; }
  for (int c = 0; c < x->ndir; ++c) {
    double* m = PL(x, w, M(c));
    LOOP(lo, hi) { long idx = IDX(x); m[idx] = m[idx] / rho[idx]; }
  }
}

/* conserve (MHD src/vacphys0.t.mhd:5-28, HD src/vacphys0.t.hd:5-27):
 * p -> e = p/(g-1) + (rho v^2 + B^2)/2 (HD rho v^2/2), v -> m = rho v */
static void conserve(ctx_t* x, double* w, const int* lo, const int* hi) {
  const double g = x->p->gamma;
  double* e = PL(x, w, E_(x));
  const double* rho = PL(x, w, RHO);
  LOOP(lo, hi) {
    long idx = IDX(x);
    double vsq = 0.0, bsq = 0.0;
    for (int c = 0; c < x->ndir; ++c) {
      double v = PL(x, w, M(c))[idx];
      vsq = (c == 0) ? v * v : vsq + v * v;
      if (x->mhd) {
        double b = PL(x, w, B(x, c))[idx];
        bsq = (c == 0) ? b * b : bsq + b * b;
      }
    }
    if (x->mhd)
      e[idx] = e[idx] / (g - 1.0) + 0.5 * (rho[idx] * vsq + bsq);
    else
      e[idx] = e[idx] / (g - 1.0) + 0.5 * rho[idx] * vsq;
  }
  for (int c = 0; c < x->ndir; ++c) {
    double* m = PL(x, w, M(c));
    LOOP(lo, hi) { long idx = IDX(x); m[idx] = rho[idx] * m[idx]; }
  }
}

/* getcmax (MHD src/vacphys0.t.mhd:79-116 with getcsound2 :140-162;
 * HD src/vacphys0.t.hd:72-100) */
static void getcmax(ctx_t* x, const double* w, const int* lo, const int* hi,
                    int idim, double* cmax) {
  const double g = x->p->gamma;
  const double* rho = PL(x, w, RHO);
  double* cs2 = x->sc5;
  getpthermal(x, 1, w, lo, hi, cs2);
  if (x->mhd) {
    LOOP(lo, hi) {
      long idx = IDX(x);
      double c2 = g * cs2[idx] / rho[idx];
      double bsq = 0.0;
      for (int c = 0; c < x->ndir; ++c) {
        double b = PL(x, w, B(x, c))[idx];
        bsq = (c == 0) ? b * b : bsq + b * b;
      }
      double cf2 = bsq / rho[idx] + c2;
      double bi = PL(x, w, B(x, idim))[idx];
      double mi = PL(x, w, M(idim))[idx];
      cmax[idx] = sqrt(0.5 * (cf2 + sqrt(cf2 * cf2 - 4.0 * c2 * (bi * bi) / rho[idx])))
                  + fabs(mi / rho[idx]);
    }
  } else {
    LOOP(lo, hi) {
      long idx = IDX(x);
      double cs = sqrt(g * cs2[idx] / rho[idx]);
      cmax[idx] = cs + fabs(PL(x, w, M(idim))[idx] / rho[idx]);
    }
  }
}

/* getflux + transport flux for one variable over a region.
 * MHD src/vacphys.mhd0.t:149-205 with getptotal src/vacphys0.t.mhd:225-257;
 * HD src/vacphys.t.hd:49-75.  Transport added as in e.g.
 * src/vactvdlf.t:67-70: f = f + v*w(iw) with v = m_idim/rho (getv). */
static void getflux_full(ctx_t* x, const double* w, const int* lo, const int* hi,
                         int iw, int idim, double* f) {
  const double g = x->p->gamma;
  const double* rho = PL(x, w, RHO);
  const double* e = PL(x, w, E_(x));
  int transport = 1;
  if (!x->mhd) {
    if (iw == M(idim) || iw == E_(x)) getpthermal(x, 1, w, lo, hi, f);
    LOOP(lo, hi) {
      long idx = IDX(x);
      if (iw == M(idim)) { /* f = p */ }
      else if (iw == E_(x)) f[idx] = PL(x, w, M(idim))[idx] / rho[idx] * f[idx];
      else f[idx] = 0.0;
    }
  } else {
    int bdim = B(x, idim);
    LOOP(lo, hi) {
      long idx = IDX(x);
      double r = rho[idx];
      double val;
      if (iw == RHO) {
        val = 0.0;
      } else if (iw >= M(0) && iw < M(x->ndir)) {
        int c = iw - M(0);
        double bc = PL(x, w, B(x, c))[idx], bi = PL(x, w, bdim)[idx];
        if (c == idim) {
          /* getptotal */
          double bsq = 0.0, msq = 0.0;
          for (int cc = 0; cc < x->ndir; ++cc) {
            double b = PL(x, w, B(x, cc))[idx], m = PL(x, w, M(cc))[idx];
            bsq = (cc == 0) ? b * b : bsq + b * b;
            msq = (cc == 0) ? m * m : msq + m * m;
          }
          double pt = (1.0 - 0.5 * g) * bsq;
          pt = pt + (g - 1.0) * e[idx];
          pt = pt - 0.5 * (g - 1.0) * msq / r;
          val = pt - bc * bi;
        } else {
          val = -(bc * bi);
        }
      } else if (iw == E_(x)) {
        double bsq = 0.0, msq = 0.0, bm = 0.0;
        for (int cc = 0; cc < x->ndir; ++cc) {
          double b = PL(x, w, B(x, cc))[idx], m = PL(x, w, M(cc))[idx];
          bsq = (cc == 0) ? b * b : bsq + b * b;
          msq = (cc == 0) ? m * m : msq + m * m;
          bm = (cc == 0) ? b * m : bm + b * m;
        }
        double pt = (1.0 - 0.5 * g) * bsq;
        pt = pt + (g - 1.0) * e[idx];
        pt = pt - 0.5 * (g - 1.0) * msq / r;
        val = (PL(x, w, M(idim))[idx] * pt - PL(x, w, bdim)[idx] * bm) / r;
      } else {
        int c = iw - B(x, 0);
        if (c == idim) { val = 0.0; }
        else val = -(PL(x, w, M(c))[idx] / r * PL(x, w, bdim)[idx]);
      }
      f[idx] = val;
    }
    if (iw == bdim) transport = 0;
  }
  if (transport) {
    const double* wi = PL(x, w, iw);
    const double* mi = PL(x, w, M(idim));
    LOOP(lo, hi) {
      long idx = IDX(x);
      double v = mi[idx] / rho[idx];
      f[idx] = f[idx] + v * wi[idx];
    }
  }
}

/* addsource_divb (src/vacphys.mhd0.t:246-280) with getdivb (:106-146) and
 * the slab gradient (src/vacgrid.t:816-823).  ensurebound (src/vacgrid.t:163-192)
 * refills the ghost cells of wCT when ixI does not reach ixO +- 1. */
static void addsource(ctx_t* x, double qdt, int ixI_full, const double* wCTc,
                      double* wnew) {
  if (!x->mhd) return;                  /* HD: vacnul.addsource.t */
  if (!x->p->divbfix) return;
  double* wCT = (double*)wCTc;
  if (!ixI_full) getboundary(x, wCT);
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double* divb = x->sc4;
  LOOP(lo, hi) {
    long idx = IDX(x);
    double dv = 0.0;
    for (int d = 0; d < x->ndim; ++d) {
      const double* b = PL(x, wCT, B(x, d));
      double gq = 0.5 * (b[idx + x->st[d]] - b[idx - x->st[d]]) / x->dx[d][CRD(d)];
      dv = dv + gq;
    }
    divb[idx] = qdt * dv;
  }
  const double* rho = PL(x, wCT, RHO);
  for (int iw = 0; iw < x->nw; ++iw) {
    double* wn = PL(x, wnew, iw);
    if (iw >= M(0) && iw < M(x->ndir)) {
      const double* b = PL(x, wCT, B(x, iw - M(0)));
      LOOP(lo, hi) { long idx = IDX(x); wn[idx] = wn[idx] - b[idx] * divb[idx]; }
    } else if (iw >= B(x, 0) && iw < B(x, x->ndir)) {
      const double* m = PL(x, wCT, M(iw - B(x, 0)));
      LOOP(lo, hi) { long idx = IDX(x); wn[idx] = wn[idx] - m[idx] / rho[idx] * divb[idx]; }
    } else if (iw == E_(x)) {
      LOOP(lo, hi) {
        long idx = IDX(x);
        double mb = 0.0;
        for (int c = 0; c < x->ndir; ++c) {
          double t = PL(x, wCT, M(c))[idx] * PL(x, wCT, B(x, c))[idx];
          mb = (c == 0) ? t : mb + t;
        }
        wn[idx] = wn[idx] - mb / rho[idx] * divb[idx];
      }
    }
  }
}

/* ---------------------------------------------------------------- TVDLF */
/* upwindLR (src/vactvdlf.t:302-367) for the common case ixR = ixL - shift
 * (hancock) and ixR = ixL (tvdmusclf): the limited slope ldw(i) is
 * dwlimiter2 of dwC(i) = w(i+1)-w(i) and dwC(i-1). */
static double ldw_at(ctx_t* x, const double* w, int iw, int d, long idx) {
  const double* pw = PL(x, w, iw);
  long s = x->st[d];
  double a = pw[idx + s] - pw[idx];
  double b = pw[idx] - pw[idx - s];
  return vacref_limit(x->p->limiter[iw], a, b);
}

/* setnoflux (src/vacgrid.t:416-449), called by addflux (:639-640) for a
 * variable whose typeB is 'symm0' on a face of dimension d: the flux f
 * through that face is zero.  f holds the flux of face idx+1/2 at idx
 * (lo_at = mlo-1 is the lower boundary face, hi_at = mhi the upper); a
 * predictor's per-cell fluxes pass the cell indices they are kept at, and
 * -1 for a face they do not touch. */
static void setnoflux(ctx_t* x, int iw, int d, double* f, int lo_at, int hi_at) {
  for (int up = 0; up < 2; ++up) {
    if (x->p->bc[2 * d + up][iw] != VACREF_SYMM0 || (up ? hi_at : lo_at) < 0) continue;
    int lo[3], hi[3];
    for (int dd = 0; dd < 3; ++dd) { lo[dd] = x->glo[dd]; hi[dd] = x->ghi[dd]; }
    lo[d] = hi[d] = up ? hi_at : lo_at;
    LOOP(lo, hi) f[IDX(x)] = 0.0;
  }
}

/* hancock (src/vactvdlf.t:5-89): predictor wnew += -qdt/dx*(F(wL)-F(wR)) */
static void hancock(ctx_t* x, double qdt, int d, const double* wCT, double* wnew, int srcmul) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double* wL = x->wLC;  /* wLC(ixO) */
  double* wR = x->wRC;  /* wRC(hxO) stored at the cell index (ixO) */
  if (x->p->useprimitive) {
    /* upwindLR with useprimitive: wLC, wRC are copies of wCT made before
     * the conversion; wCT itself is converted over ixG and back, in place */
    for (int iw = 0; iw < x->nw; ++iw) {
      const double* pc = PL(x, wCT, iw);
      double* pl = PL(x, wL, iw);
      double* pr = PL(x, wR, iw);
      LOOP(lo, hi) { long idx = IDX(x); pl[idx] = pc[idx]; pr[idx] = pc[idx]; }
    }
    primitive(x, (double*)wCT, x->glo, x->ghi);
    primitive(x, wL, lo, hi);
    primitive(x, wR, lo, hi);
    for (int iw = 0; iw < x->nw; ++iw) {
      double* pl = PL(x, wL, iw);
      double* pr = PL(x, wR, iw);
      LOOP(lo, hi) {
        long idx = IDX(x);
        double l = ldw_at(x, wCT, iw, d, idx);
        pr[idx] = pr[idx] - 0.5 * l;
        pl[idx] = pl[idx] + 0.5 * l;
      }
    }
    conserve(x, (double*)wCT, x->glo, x->ghi);
    conserve(x, wL, lo, hi);
    conserve(x, wR, lo, hi);
  } else
  for (int iw = 0; iw < x->nw; ++iw) {
    const double* pc = PL(x, wCT, iw);
    double* pl = PL(x, wL, iw);
    double* pr = PL(x, wR, iw);
    LOOP(lo, hi) {
      long idx = IDX(x);
      double l = ldw_at(x, wCT, iw, d, idx);
      pr[idx] = pc[idx] - 0.5 * l;
      pl[idx] = pc[idx] + 0.5 * l;
    }
  }
  double* fL = x->f;
  double* fR = x->fC;
  for (int iw = 0; iw < x->nw; ++iw) {
    getflux_full(x, wR, lo, hi, iw, d, fR);
    getflux_full(x, wL, lo, hi, iw, d, fL);
    /* addflux(qdt,ixO,iw,idim,fLC,ixO,fRC,hxO,wnew): fLC is the right face's
     * flux of a cell, fRC (kept here at the cell) its left face's */
    setnoflux(x, iw, d, fR, x->mlo[d], -1);
    setnoflux(x, iw, d, fL, -1, x->mhi[d]);
    double* wn = PL(x, wnew, iw);
    LOOP(lo, hi) {  /* addflux (src/vacgrid.t:676-677) */
      long idx = IDX(x);
      wn[idx] = wn[idx] - qdt / x->dx[d][CRD(d)] * (fL[idx] - fR[idx]);
    }
  }
  /* after the last direction: qdt*(idimmax-idimmin+1)/ndim (src/vactvdlf.t:85-86) */
  if (x->p->sourceunsplit && srcmul) addsource(x, qdt * (double)srcmul / x->ndim, 1, wCT, wnew);
}

static void tvdlimit2(ctx_t* x, double qdt, int d, const double* wL, const double* wR,
                      double* wnew, int mu);

/* tvdmusclf for method='tvdlf' (mu = 0) or 'tvdmu' (mu = 1)
 * (src/vactvdlf.t:92-299); wnew is the base of the limiter
 * (typelimited='original') and receives the update.  'tvdmu' has no Lax-
 * Friedrichs dissipation (and no cmax Courant check): the central flux is
 * added, then tvdlimit2 adds the Roe dissipation of the reconstructed face
 * states (src/vactvdlf.t:279-283). */
static void tvdlf_corrector(ctx_t* x, double qdt, int d, const double* wCT, double* wnew, int mu,
                            int srcmul) {
  int lo[3], hi[3];
  for (int dd = 0; dd < 3; ++dd) { lo[dd] = x->mlo[dd]; hi[dd] = x->mhi[dd]; }
  lo[d] -= 1;  /* ixC = [ixOmin-1, ixOmax] */
  long s = x->st[d];
  double* wL = x->wLC;
  double* wR = x->wRC;
  /* the limiter's base, typelimited (src/vactvdlf.t:150-161) */
  /* unsplit sweeps limit 'previous' where 'original' is asked (src/vacio.t:363-364) */
  const int tl = (x->p->unsplit && x->p->typelimited == 0) ? 1 : x->p->typelimited;
  const double* wlim = tl == 1 ? x->wold : (tl == 2 ? wCT : wnew);
  if (x->p->useprimitive) {
    /* upwindLR with useprimitive: the copies of wCT and the limiter's base
     * (over ixG, in place and back) in primitive variables */
    for (int iw = 0; iw < x->nw; ++iw) {
      const double* pc = PL(x, wCT, iw);
      double* pl = PL(x, wL, iw);
      double* pr = PL(x, wR, iw);
      LOOP(lo, hi) { long idx = IDX(x); pr[idx] = pc[idx + s]; pl[idx] = pc[idx]; }
    }
    primitive(x, (double*)wlim, x->glo, x->ghi);
    primitive(x, wL, lo, hi);
    primitive(x, wR, lo, hi);
    for (int iw = 0; iw < x->nw; ++iw) {
      double* pl = PL(x, wL, iw);
      double* pr = PL(x, wR, iw);
      LOOP(lo, hi) {
        long idx = IDX(x);
        double lj = ldw_at(x, wlim, iw, d, idx);
        double lj1 = ldw_at(x, wlim, iw, d, idx + s);
        pr[idx] = pr[idx] - 0.5 * lj1;
        pl[idx] = pl[idx] + 0.5 * lj;
      }
    }
    conserve(x, (double*)wlim, x->glo, x->ghi);
    conserve(x, wL, lo, hi);
    conserve(x, wR, lo, hi);
  } else
  for (int iw = 0; iw < x->nw; ++iw) {
    const double* pc = PL(x, wCT, iw);
    double* pl = PL(x, wL, iw);
    double* pr = PL(x, wR, iw);
    LOOP(lo, hi) {
      long idx = IDX(x);
      double lj = ldw_at(x, wlim, iw, d, idx);
      double lj1 = ldw_at(x, wlim, iw, d, idx + s);
      pr[idx] = pc[idx + s] - 0.5 * lj1;
      pl[idx] = pc[idx] + 0.5 * lj;
    }
  }
  double* cmaxC = x->sc1;
  if (!mu) {
  /* mean state in wLC, cmax of it */
  for (int iw = 0; iw < x->nw; ++iw) {
    double* pl = PL(x, wL, iw);
    const double* pr = PL(x, wR, iw);
    LOOP(lo, hi) { long idx = IDX(x); pl[idx] = 0.5 * (pl[idx] + pr[idx]); }
  }
  getcmax(x, wL, lo, hi, d, cmaxC);
  if (x->p->courantpar > 0.0 && x->istep == x->nstep) {
    double cm = -BIGD;
    LOOP(lo, hi) {
      long idx = IDX(x);
      int c = CRD(d);
      double v = cmaxC[idx] * 2.0 / (x->dx[d][c] + x->dx[d][c + 1]);
      cm = dmax(cm, v);
    }
    if (qdt * cm > 1.0) x->s->nerror[2]++;
    if (cm > SMALLD) x->s->dtcourant[d] = dmin(x->s->dtcourant[d], x->p->courantpar / cm);
  }
  for (int iw = 0; iw < x->nw; ++iw) {
    double* pl = PL(x, wL, iw);
    const double* pr = PL(x, wR, iw);
    LOOP(lo, hi) { long idx = IDX(x); pl[idx] = 2.0 * pl[idx] - pr[idx]; }
  }
  }
  double* fL = x->f;
  double* fR = x->fC;
  int olo[3], ohi[3];
  for (int dd = 0; dd < 3; ++dd) { olo[dd] = x->mlo[dd]; ohi[dd] = x->mhi[dd]; }
  for (int iw = 0; iw < x->nw; ++iw) {
    getflux_full(x, wL, lo, hi, iw, d, fL);
    getflux_full(x, wR, lo, hi, iw, d, fR);
    const double* pl = PL(x, wL, iw);
    const double* pr = PL(x, wR, iw);
    LOOP(lo, hi) {
      long idx = IDX(x);
      double fm = 0.5 * (fL[idx] + fR[idx]);
      if (mu) { fL[idx] = fm; continue; }
      double diss = -(cmaxC[idx] * 0.5 * (pr[idx] - pl[idx]));
      fL[idx] = fm + diss;
    }
    setnoflux(x, iw, d, fL, x->mlo[d] - 1, x->mhi[d]);
    double* wn = PL(x, wnew, iw);
    LOOP(olo, ohi) {
      long idx = IDX(x);
      wn[idx] = wn[idx] - qdt / x->dx[d][CRD(d)] * (fL[idx] - fL[idx - s]);
    }
  }
  if (mu) tvdlimit2(x, qdt, d, wL, wR, wnew, 1);
  /* ixI = ixO expanded by 2 in the idim range: for dimensionally split
   * sweeps in idim only, so ensurebound refills wCT ghosts */
  if (x->p->sourceunsplit && srcmul)
    addsource(x, qdt * (double)srcmul / x->ndim, x->ndim == 1 || x->p->unsplit, wCT, wnew);
}

/* ------------------------------------------------------------ MHD Roe */
/* average (src/vacphys.mhdroe.t:4-122) over an interface region; wroe
 * receives sqrt(rho), v, p, B_idim (clamped) and beta_idir[,beta_jdir]. */
static void roe_average(ctx_t* x, const double* wL, const double* wR,
                        const int* lo, const int* hi, int idim, double* wroe) {
  const double g = x->p->gamma;
  const int nd = x->ndir, eI = E_(x);
  LOOP(lo, hi) {
    long idx = IDX(x);
    PL(x, wroe, RHO)[idx] = 0.5 * (PL(x, wL, RHO)[idx] + PL(x, wR, RHO)[idx]);
    for (int c = 0; c < nd; ++c) {
      PL(x, wroe, M(c))[idx] = 0.5 * (PL(x, wL, M(c))[idx] / PL(x, wL, RHO)[idx] +
                                      PL(x, wR, M(c))[idx] / PL(x, wR, RHO)[idx]);
      PL(x, wroe, B(x, c))[idx] = 0.5 * (PL(x, wL, B(x, c))[idx] + PL(x, wR, B(x, c))[idx]);
    }
  }
  getpthermal(x, 1, wL, lo, hi, x->afast);
  getpthermal(x, 1, wR, lo, hi, x->aslow);
  LOOP(lo, hi) {
    long idx = IDX(x);
    double* r = PL(x, wroe, 0);
    (void)r;
    PL(x, wroe, eI)[idx] = 0.5 * (x->afast[idx] + x->aslow[idx]);
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int c = 0; c < nd; ++c) {
      double v = PL(x, wroe, M(c))[idx];
      double t = v * (PL(x, wR, M(c))[idx] - PL(x, wL, M(c))[idx]);
      s1 = (c == 0) ? t : s1 + t;
      s2 = (c == 0) ? v * v : s2 + v * v;
      double bR = PL(x, wR, B(x, c))[idx], bL = PL(x, wL, B(x, c))[idx];
      s3 = (c == 0) ? bR * bR - bL * bL : s3 + bR * bR - bL * bL;
    }
    double drho = PL(x, wR, RHO)[idx] - PL(x, wL, RHO)[idx];
    double t = PL(x, wR, eI)[idx] - PL(x, wL, eI)[idx] - s1 + 0.5 * s2 * drho - 0.5 * s3;
    /* useprimitive: dp = pR - pL (src/vacphys.mhdroe.t:42-44) */
    x->dp[idx] = x->p->useprimitive ? x->aslow[idx] - x->afast[idx] : (g - 1.0) * t;
    x->rhodv[idx] = PL(x, wR, M(idim))[idx] - PL(x, wL, M(idim))[idx] -
                    PL(x, wroe, M(idim))[idx] * drho;
    /* getcsound2prim (src/vacphys0.t.mhd:119-137) */
    double c2 = g * PL(x, wroe, eI)[idx] / PL(x, wroe, RHO)[idx];
    x->csound2[idx] = c2;
    double bsq = 0.0;
    for (int c = 0; c < nd; ++c) {
      double b = PL(x, wroe, B(x, c))[idx];
      bsq = (c == 0) ? b * b : bsq + b * b;
    }
    double rr = PL(x, wroe, RHO)[idx];
    double bi = PL(x, wroe, B(x, idim))[idx];
    double cf = bsq / rr + c2;
    double cs = dmax(0.0, 0.5 * (cf - sqrt(cf * cf - 4.0 * c2 * (bi * bi) / rr)));
    cf = cf - cs;
    double af = (c2 - cs) / (cf - cs);
    af = dmin(1.0, dmax(af, 0.0));
    x->aslow[idx] = sqrt(1.0 - af);
    x->afast[idx] = sqrt(af);
    x->cfast[idx] = sqrt(cf);
    x->cslow[idx] = sqrt(cs);
    PL(x, wroe, RHO)[idx] = sqrt(rr);
    if (fabs(bi) < SMALLD) PL(x, wroe, B(x, idim))[idx] = SMALLD;
    int idir = (idim + 1) % nd;
    if (nd == 2) {
      double* bb = PL(x, wroe, B(x, idir));
      bb[idx] = (bb[idx] >= 0.0) ? 1.0 : -1.0;
    } else {
      int jdir = (idir + 1) % nd;
      double bI = PL(x, wroe, B(x, idir))[idx], bJ = PL(x, wroe, B(x, jdir))[idx];
      double tt = sqrt(bI * bI + bJ * bJ);
      if (tt > SMALLD) {
        PL(x, wroe, B(x, idir))[idx] = bI / tt;
        PL(x, wroe, B(x, jdir))[idx] = bJ / tt;
      } else {
        PL(x, wroe, B(x, idir))[idx] = sqrt(0.5);
        PL(x, wroe, B(x, jdir))[idx] = sqrt(0.5);
      }
    }
  }
}

/* geteigenjump (src/vacphys.mhdroe.t:125-284), entropy fix 'nul' */
static void roe_eigenjump(ctx_t* x, const double* wL, const double* wR,
                          const double* wroe, const int* lo, const int* hi,
                          int il, int idim, double* a, double* jump) {
  const int nd = x->ndir;
  int idir = (idim + 1) % nd, jdir = (idir + 1) % nd;
  const double* sq = PL(x, wroe, RHO);
  const double* vd = PL(x, wroe, M(idim));
  const double* bd = PL(x, wroe, B(x, idim));
  if (il == 0) {  /* fastRW_ */
    LOOP(lo, hi) {
      long idx = IDX(x);
      double t = PL(x, wroe, B(x, idir))[idx] *
                 (PL(x, wR, M(idir))[idx] / PL(x, wR, RHO)[idx] -
                  PL(x, wL, M(idir))[idx] / PL(x, wL, RHO)[idx]);
      if (nd == 3)
        t = t + PL(x, wroe, B(x, jdir))[idx] *
                    (PL(x, wR, M(jdir))[idx] / PL(x, wR, RHO)[idx] -
                     PL(x, wL, M(jdir))[idx] / PL(x, wL, RHO)[idx]);
      x->bdv[idx] = t * fsign(sq[idx] * sq[idx], bd[idx]);
      double u = PL(x, wroe, B(x, idir))[idx] *
                 (PL(x, wR, B(x, idir))[idx] - PL(x, wL, B(x, idir))[idx]);
      if (nd == 3)
        u = u + PL(x, wroe, B(x, jdir))[idx] *
                    (PL(x, wR, B(x, jdir))[idx] - PL(x, wL, B(x, jdir))[idx]);
      x->bdb[idx] = u * sqrt(x->csound2[idx]) * sq[idx];
    }
  }
  if (il == 6) {  /* alfvRW_ */
    LOOP(lo, hi) {
      long idx = IDX(x);
      double bI = PL(x, wroe, B(x, idir))[idx], bJ = PL(x, wroe, B(x, jdir))[idx];
      double t = bJ * (PL(x, wR, M(idir))[idx] / PL(x, wR, RHO)[idx] -
                       PL(x, wL, M(idir))[idx] / PL(x, wL, RHO)[idx]) -
                 bI * (PL(x, wR, M(jdir))[idx] / PL(x, wR, RHO)[idx] -
                       PL(x, wL, M(jdir))[idx] / PL(x, wL, RHO)[idx]);
      double u = bJ * (PL(x, wR, B(x, idir))[idx] - PL(x, wL, B(x, idir))[idx]) -
                 bI * (PL(x, wR, B(x, jdir))[idx] - PL(x, wL, B(x, jdir))[idx]);
      x->bdv[idx] = t * 0.5 * (sq[idx] * sq[idx]);
      x->bdb[idx] = u * 0.5 * fsign(sq[idx], bd[idx]);
    }
  }
  LOOP(lo, hi) {
    long idx = IDX(x);
    double cf = x->cfast[idx], cs = x->cslow[idx], af = x->afast[idx], as = x->aslow[idx];
    double c2 = x->csound2[idx], dpv = x->dp[idx], rdv = x->rhodv[idx];
    double bdv = x->bdv[idx], bdb = x->bdb[idx];
    switch (il) {
      case 0:
        a[idx] = vd[idx] + cf;
        jump[idx] = 0.5 / c2 * (af * (+cf * rdv + dpv) + as * (-cs * bdv + bdb));
        break;
      case 1:
        a[idx] = vd[idx] - cf;
        jump[idx] = 0.5 / c2 * (af * (-cf * rdv + dpv) + as * (+cs * bdv + bdb));
        break;
      case 2:
        a[idx] = vd[idx] + cs;
        jump[idx] = 0.5 / c2 * (as * (+cs * rdv + dpv) + af * (+cf * bdv - bdb));
        break;
      case 3:
        a[idx] = vd[idx] - cs;
        jump[idx] = 0.5 / c2 * (as * (-cs * rdv + dpv) + af * (-cf * bdv - bdb));
        break;
      case 4:
        a[idx] = vd[idx];
        jump[idx] = PL(x, wR, RHO)[idx] - PL(x, wL, RHO)[idx] - dpv / c2;
        break;
      case 5:
        if (x->p->divbwave) {
          a[idx] = vd[idx];
          jump[idx] = PL(x, wR, B(x, idim))[idx] - PL(x, wL, B(x, idim))[idx];
        } else {
          a[idx] = 0.0;
          jump[idx] = 0.0;
        }
        break;
      case 6:
        a[idx] = vd[idx] + fabs(bd[idx]) / sq[idx];
        jump[idx] = +bdv - bdb;
        break;
      case 7:
        a[idx] = vd[idx] - fabs(bd[idx]) / sq[idx];
        jump[idx] = -bdv - bdb;
        break;
    }
  }
}

/* rtimes (src/vacphys.mhdroe.t:287-450): rq = R(iw,il) * q */
static void roe_rtimes(ctx_t* x, const double* q, const double* wroe, const int* lo,
                       const int* hi, int iw, int il, int idim, double* rq) {
  const int nd = x->ndir;
  const double g = x->p->gamma;
  int idir = (idim + 1) % nd, jdir = (idir + 1) % nd;
  const double* sq = PL(x, wroe, RHO);
  const double* vd = PL(x, wroe, M(idim));
  const double* bd = PL(x, wroe, B(x, idim));
  const int eI = E_(x);
  if (iw == eI) {
    if (il == 0) {
      LOOP(lo, hi) {
        long idx = IDX(x);
        double vs = 0.0;
        for (int c = 0; c < nd; ++c) {
          double v = PL(x, wroe, M(c))[idx];
          vs = (c == 0) ? v * v : vs + v * v;
        }
        x->v2a2[idx] = 0.5 * vs + (2.0 - g) / (g - 1.0) * x->csound2[idx];
        double t = PL(x, wroe, B(x, idir))[idx] * PL(x, wroe, M(idir))[idx];
        if (nd == 3) t = t + PL(x, wroe, B(x, jdir))[idx] * PL(x, wroe, M(jdir))[idx];
        x->bv[idx] = t * copysign(1.0, bd[idx]);
      }
    } else if (il == 6) {
      LOOP(lo, hi) {
        long idx = IDX(x);
        x->bv[idx] = PL(x, wroe, B(x, jdir))[idx] * PL(x, wroe, M(idir))[idx] -
                     PL(x, wroe, B(x, idir))[idx] * PL(x, wroe, M(jdir))[idx];
      }
    }
  }
  LOOP(lo, hi) {
    long idx = IDX(x);
    double qq = q[idx];
    double cf = x->cfast[idx], cs = x->cslow[idx], af = x->afast[idx], as = x->aslow[idx];
    double r = 0.0;
    if (iw == RHO) {
      switch (il) {
        case 0: case 1: r = qq * af; break;
        case 2: case 3: r = qq * as; break;
        case 4: r = qq; break;
        default: r = 0.0;
      }
    } else if (iw == eI) {
      double bv = x->bv[idx], v2a2 = x->v2a2[idx];
      switch (il) {
        case 0: r = qq * (-as * cs * bv + af * (v2a2 + cf * (cf + vd[idx]))); break;
        case 1: r = qq * (+as * cs * bv + af * (v2a2 + cf * (cf - vd[idx]))); break;
        case 2: r = qq * (+af * cf * bv + as * (v2a2 + cs * (cs + vd[idx]))); break;
        case 3: r = qq * (-af * cf * bv + as * (v2a2 + cs * (cs - vd[idx]))); break;
        case 4: {
          double vs = 0.0;
          for (int c = 0; c < nd; ++c) {
            double v = PL(x, wroe, M(c))[idx];
            vs = (c == 0) ? v * v : vs + v * v;
          }
          r = qq * 0.5 * vs;
          break;
        }
        case 5: r = x->p->divbwave ? qq * bd[idx] : 0.0; break;
        case 6: r = +qq * bv; break;
        case 7: r = -qq * bv; break;
      }
    } else if (iw >= M(0) && iw < M(nd)) {
      int c = iw - M(0);
      double vw = PL(x, wroe, iw)[idx];
      if (c == idim) {
        switch (il) {
          case 0: r = qq * af * (vw + cf); break;
          case 1: r = qq * af * (vw - cf); break;
          case 2: r = qq * as * (vw + cs); break;
          case 3: r = qq * as * (vw - cs); break;
          case 4: r = qq * vw; break;
          default: r = 0.0;
        }
      } else {
        double beta = PL(x, wroe, B(x, c))[idx];
        double sg = copysign(1.0, bd[idx]);
        switch (il) {
          case 0: r = qq * (af * vw - as * cs * beta * sg); break;
          case 1: r = qq * (af * vw + as * cs * beta * sg); break;
          case 2: r = qq * (as * vw + af * cf * beta * sg); break;
          case 3: r = qq * (as * vw - af * cf * beta * sg); break;
          case 4: r = qq * vw; break;
          case 5: r = 0.0; break;
          case 6:
            r = (c == idir) ? +qq * PL(x, wroe, B(x, jdir))[idx]
                            : -qq * PL(x, wroe, B(x, idir))[idx];
            break;
          case 7:
            r = (c == idir) ? -qq * PL(x, wroe, B(x, jdir))[idx]
                            : +qq * PL(x, wroe, B(x, idir))[idx];
            break;
        }
      }
    } else {
      int c = iw - B(x, 0);
      if (c == idim) {
        r = (il == 5 && x->p->divbwave) ? qq : 0.0;
      } else {
        double beta = PL(x, wroe, iw)[idx];
        switch (il) {
          case 0: case 1: r = +qq * as * sqrt(x->csound2[idx]) * beta / sq[idx]; break;
          case 2: case 3: r = -(qq * af * sqrt(x->csound2[idx]) * beta / sq[idx]); break;
          case 4: case 5: r = 0.0; break;
          case 6: case 7:
            if (c == idir) r = -(qq * PL(x, wroe, B(x, jdir))[idx] / fsign(sq[idx], bd[idx]));
            else r = +qq * PL(x, wroe, B(x, idir))[idx] / fsign(sq[idx], bd[idx]);
            break;
        }
      }
    }
    rq[idx] = r;
  }
}

/* ------------------------------------------------------------ HD Roe */
/* average (src/vacphys.hdroe.t:4-46) over an interface region; wroe receives
 * sqrt(rhoL*rhoR), the Roe-averaged velocity and the Roe-averaged enthalpy
 * h (in the energy slot). */
static void roe_average_hd(ctx_t* x, const double* wL, const double* wR,
                           const int* lo, const int* hi, double* wroe) {
  const int nd = x->ndir, eI = E_(x);
  double* t1 = x->sc2;
  double* t2 = x->sc3;
  LOOP(lo, hi) {
    long idx = IDX(x);
    t1[idx] = sqrt(PL(x, wL, RHO)[idx]);
    t2[idx] = sqrt(PL(x, wR, RHO)[idx]);
    PL(x, wroe, RHO)[idx] = t1[idx] * t2[idx];
    t1[idx] = t1[idx] / t2[idx];  /* sqrt(rhoL/rhoR) */
    for (int c = 0; c < nd; ++c)
      PL(x, wroe, M(c))[idx] = (PL(x, wL, M(c))[idx] / PL(x, wL, RHO)[idx] * t1[idx] +
                                PL(x, wR, M(c))[idx] / PL(x, wR, RHO)[idx]) / (1 + t1[idx]);
  }
  getpthermal(x, 1, wL, lo, hi, t2);
  LOOP(lo, hi) {
    long idx = IDX(x);
    PL(x, wroe, eI)[idx] = (t2[idx] + PL(x, wL, eI)[idx]) / PL(x, wL, RHO)[idx];
  }
  getpthermal(x, 1, wR, lo, hi, t2);
  LOOP(lo, hi) {
    long idx = IDX(x);
    t2[idx] = (t2[idx] + PL(x, wR, eI)[idx]) / PL(x, wR, RHO)[idx];
    PL(x, wroe, eI)[idx] = (PL(x, wroe, eI)[idx] * t1[idx] + t2[idx]) / (1 + t1[idx]);
  }
}

/* geteigenjump (src/vacphys.hdroe.t:48-150), entropy fix 'nul'.  The first
 * wave sets the sound speed (common /roe/ csound), (pR-pL)/c^2 and
 * (vR-vL)/c along idim, which the others reuse.  Waves: 0 sound right,
 * 1 sound left, 2 entropy, 3.. shear (direction il-3, idim skipped). */
static void roe_eigenjump_hd(ctx_t* x, const double* wL, const double* wR,
                             const double* wroe, const int* lo, const int* hi,
                             int il, int idim, double* a, double* jump) {
  const int nd = x->ndir, eI = E_(x);
  const double g = x->p->gamma;
  double* cs = x->cfast;
  double* dpc = x->dp;
  double* dvc = x->rhodv;
  const double* rr = PL(x, wroe, RHO);
  const double* vd = PL(x, wroe, M(idim));
  if (il == 0) {
    LOOP(lo, hi) {
      long idx = IDX(x);
      double vs = 0.0;
      for (int c = 0; c < nd; ++c) {
        double v = PL(x, wroe, M(c))[idx];
        vs = (c == 0) ? v * v : vs + v * v;
      }
      cs[idx] = (g - 1) * (PL(x, wroe, eI)[idx] - 0.5 * vs);
      cs[idx] = dmax(g * x->s->smallp / rr[idx], cs[idx]);
    }
    getpthermal(x, 1, wL, lo, hi, x->sc2);
    getpthermal(x, 1, wR, lo, hi, x->sc3);
    LOOP(lo, hi) {
      long idx = IDX(x);
      dpc[idx] = (x->sc3[idx] - x->sc2[idx]) / cs[idx];
      cs[idx] = sqrt(cs[idx]);
      dvc[idx] = (PL(x, wR, M(idim))[idx] / PL(x, wR, RHO)[idx] -
                  PL(x, wL, M(idim))[idx] / PL(x, wL, RHO)[idx]) / cs[idx];
    }
  }
  LOOP(lo, hi) {
    long idx = IDX(x);
    switch (il) {
      case 0: a[idx] = vd[idx] + cs[idx]; jump[idx] = 0.5 * (dpc[idx] + rr[idx] * dvc[idx]); break;
      case 1: a[idx] = vd[idx] - cs[idx]; jump[idx] = 0.5 * (dpc[idx] - rr[idx] * dvc[idx]); break;
      case 2:
        a[idx] = vd[idx];
        jump[idx] = -dpc[idx] + PL(x, wR, RHO)[idx] - PL(x, wL, RHO)[idx];
        break;
      default: {
        int idir = il - 3;
        if (idir >= idim) idir += 1;
        a[idx] = vd[idx];
        jump[idx] = rr[idx] * (PL(x, wR, M(idir))[idx] / PL(x, wR, RHO)[idx] -
                               PL(x, wL, M(idir))[idx] / PL(x, wL, RHO)[idx]);
      }
    }
  }
}

/* rtimes (src/vacphys.hdroe.t:152-215): rq = R(iw, il) * q */
static void roe_rtimes_hd(ctx_t* x, const double* q, const double* wroe, const int* lo,
                          const int* hi, int iw, int il, int idim, double* rq) {
  const int nd = x->ndir, eI = E_(x);
  const double* cs = x->cfast;
  const double* vd = PL(x, wroe, M(idim));
  const double* h = PL(x, wroe, eI);
  const int shear = il >= 3;
  int idir = il - 3;
  if (shear && idir >= idim) idir += 1;
  LOOP(lo, hi) {
    long idx = IDX(x);
    double r;
    if (iw == RHO) {
      r = shear ? 0.0 : q[idx];
    } else if (iw == eI) {
      switch (il) {
        case 0: r = q[idx] * (h[idx] + vd[idx] * cs[idx]); break;
        case 1: r = q[idx] * (h[idx] - vd[idx] * cs[idx]); break;
        case 2: {
          double vs = 0.0;
          for (int c = 0; c < nd; ++c) {
            double v = PL(x, wroe, M(c))[idx];
            vs = (c == 0) ? v * v : vs + v * v;
          }
          r = q[idx] * 0.5 * vs;
          break;
        }
        default: r = q[idx] * PL(x, wroe, M(idir))[idx];
      }
    } else if (iw == M(idim)) {
      switch (il) {
        case 0: r = q[idx] * (vd[idx] + cs[idx]); break;
        case 1: r = q[idx] * (vd[idx] - cs[idx]); break;
        case 2: r = q[idx] * vd[idx]; break;
        default: r = 0.0;
      }
    } else if (shear) {
      r = (iw == M(idir)) ? q[idx] : 0.0;
    } else {
      r = q[idx] * PL(x, wroe, iw)[idx];
    }
    rq[idx] = r;
  }
}

/* ---------------------------------------------------------- entropy fix */
/* The end of geteigenjump (src/vacphys.mhdroe.t:228-281,
 * src/vacphys.hdroe.t:120-147) and entropyfix (src/vactvd.t:422-449) for
 * wave il over [lo, hi]: 'yee' sets smalla = entropycoef(il); 'harten',
 * 'powell' and (MHD) 'ratio' take the wave's speed at the left and right
 * states -- the velocity, plus or minus the fast, slow or Alfven speed (the
 * sound speed for HD) -- and set smalla, or modify a ('ratio').  The fast
 * wave (il 0) sets the saved sound speeds the slow waves use; the right
 * Alfven wave saves the Alfven speeds for the left one. */
static void entropy_fix(ctx_t* x, const double* wL, const double* wR, const int* lo,
                        const int* hi, int il, int idim, double* a, double* smalla) {
  const int type = x->p->typeentropy[il];
  if (type == VACREF_ENT_NUL) return;
  if (type == VACREF_ENT_YEE) {
    LOOP(lo, hi) smalla[IDX(x)] = x->p->entropycoef[il];
    return;
  }
  const double g = x->p->gamma;
  double* aL = x->eaL;
  double* aR = x->eaR;
  LOOP(lo, hi) {
    long idx = IDX(x);
    aL[idx] = PL(x, wL, M(idim))[idx] / PL(x, wL, RHO)[idx];
    aR[idx] = PL(x, wR, M(idim))[idx] / PL(x, wR, RHO)[idx];
  }
  if (x->mhd) {
    double *c2L = x->cs2L, *c2R = x->cs2R, *caL = x->cs2ca2L, *caR = x->cs2ca2R;
    if (il == 0) {
      getpthermal(x, 1, wL, lo, hi, c2L);
      getpthermal(x, 1, wR, lo, hi, c2R);
      LOOP(lo, hi) {
        long idx = IDX(x);
        const double rL = PL(x, wL, RHO)[idx], rR = PL(x, wR, RHO)[idx];
        c2L[idx] = g * c2L[idx] / rL;  /* getcsound2, src/vacphys0.t.mhd:140-162 */
        c2R[idx] = g * c2R[idx] / rR;
        double bl = 0.0, br = 0.0;
        for (int c = 0; c < x->ndir; ++c) {
          double u = PL(x, wL, B(x, c))[idx], v = PL(x, wR, B(x, c))[idx];
          bl = (c == 0) ? u * u : bl + u * u;
          br = (c == 0) ? v * v : br + v * v;
        }
        caL[idx] = c2L[idx] + bl / rL;
        caR[idx] = c2R[idx] + br / rR;
        const double bdL = PL(x, wL, B(x, idim))[idx], bdR = PL(x, wR, B(x, idim))[idx];
        c2L[idx] = sqrt(caL[idx] * caL[idx] - 4 * c2L[idx] * (bdL * bdL) / rL);
        c2R[idx] = sqrt(caR[idx] * caR[idx] - 4 * c2R[idx] * (bdR * bdR) / rR);
      }
    }
    if (il == 6) {
      LOOP(lo, hi) {
        long idx = IDX(x);
        caL[idx] = fabs(PL(x, wL, B(x, idim))[idx]) / sqrt(PL(x, wL, RHO)[idx]);
        caR[idx] = fabs(PL(x, wR, B(x, idim))[idx]) / sqrt(PL(x, wR, RHO)[idx]);
      }
    }
    LOOP(lo, hi) {
      long idx = IDX(x);
      switch (il) {
        case 0: aL[idx] = aL[idx] + sqrt(0.5 * (caL[idx] + c2L[idx]));
                aR[idx] = aR[idx] + sqrt(0.5 * (caR[idx] + c2R[idx])); break;
        case 1: aL[idx] = aL[idx] - sqrt(0.5 * (caL[idx] + c2L[idx]));
                aR[idx] = aR[idx] - sqrt(0.5 * (caR[idx] + c2R[idx])); break;
        case 2: aL[idx] = aL[idx] + sqrt(0.5 * (caL[idx] - c2L[idx]));
                aR[idx] = aR[idx] + sqrt(0.5 * (caR[idx] - c2R[idx])); break;
        case 3: aL[idx] = aL[idx] - sqrt(0.5 * (caL[idx] - c2L[idx]));
                aR[idx] = aR[idx] - sqrt(0.5 * (caR[idx] - c2R[idx])); break;
        case 6: aL[idx] = aL[idx] + caL[idx]; aR[idx] = aR[idx] + caR[idx]; break;
        case 7: aL[idx] = aL[idx] - caL[idx]; aR[idx] = aR[idx] - caR[idx]; break;
        default: break;
      }
    }
  } else if (il <= 1) {
    getpthermal(x, 1, wL, lo, hi, x->cs2L);
    getpthermal(x, 1, wR, lo, hi, x->cs2R);
    const double sg = il == 0 ? 1.0 : -1.0;
    LOOP(lo, hi) {
      long idx = IDX(x);
      const double cl = sqrt(g * x->cs2L[idx] / PL(x, wL, RHO)[idx]);
      const double cr = sqrt(g * x->cs2R[idx] / PL(x, wR, RHO)[idx]);
      aL[idx] = sg > 0 ? aL[idx] + cl : aL[idx] - cl;
      aR[idx] = sg > 0 ? aR[idx] + cr : aR[idx] - cr;
    }
  }
  LOOP(lo, hi) {
    long idx = IDX(x);
    switch (type) {
      case VACREF_ENT_HARTEN:
        smalla[idx] = dmax(dmax(0.0, a[idx] - aL[idx]), aR[idx] - a[idx]);
        break;
      case VACREF_ENT_POWELL:
        smalla[idx] = dmax(0.0, 2.0 * (aR[idx] - aL[idx]));
        break;
      case VACREF_ENT_RATIO:
        if (aL[idx] < 0.0 && aR[idx] > 0.0)
          a[idx] = a[idx] - 2 * aR[idx] * aL[idx] / (aR[idx] - aL[idx]);
        break;
    }
  }
}

/* tvdlimit2 (src/vactvd.t:86-206) with getphi for typetvd='roe' (:209-302),
 * typeentropy='nul', acmcoef<0, acmexpo=0.  mu = 0: method 'tvd', called by
 * tvdlimit over ixIC = ixC widened by one face below (the limiter of phi
 * needs the neighbour jumps); mu = 1: method 'tvdmu', called by tvdmusclf
 * (src/vactvdlf.t:279-283) with ixIC = ixC, where getphi returns
 * phi = |a dt/dx| jump (src/vactvd.t:227-240). */
static void tvdlimit2(ctx_t* x, double qdt, int d, const double* wL, const double* wR,
                      double* wnew, int mu) {
  int icl[3], ich[3], cl[3], ch[3];
  for (int dd = 0; dd < 3; ++dd) {
    icl[dd] = cl[dd] = x->mlo[dd];
    ich[dd] = ch[dd] = x->mhi[dd];
  }
  icl[d] = x->mlo[d] - 2; ich[d] = x->mhi[d] + 1;  /* ixIC */
  cl[d] = x->mlo[d] - 1;                            /* ixC */
  if (mu) { icl[d] = cl[d]; ich[d] = ch[d]; }
  long s = x->st[d];
  if (x->mhd) roe_average(x, wL, wR, icl, ich, d, x->wroe);
  else roe_average_hd(x, wL, wR, icl, ich, x->wroe);
  for (int il = 0; il < x->nw; ++il) {
    if (x->mhd) roe_eigenjump(x, wL, wR, x->wroe, icl, ich, il, d, x->adtdx, x->jump);
    else roe_eigenjump_hd(x, wL, wR, x->wroe, icl, ich, il, d, x->adtdx, x->jump);
    entropy_fix(x, wL, wR, icl, ich, il, d, x->adtdx, x->smalla);
    const int et = x->p->typeentropy[il];
    const int fixed = et != VACREF_ENT_NUL && et != VACREF_ENT_RATIO;
    LOOP(icl, ich) {
      long idx = IDX(x);
      int c = CRD(d);
      x->adtdx[idx] = x->adtdx[idx] * qdt * 2.0 / (x->dx[d][c] + x->dx[d][c + 1]);
      if (et == VACREF_ENT_HARTEN || et == VACREF_ENT_POWELL)  /* :130-141 */
        x->smalla[idx] = x->smalla[idx] * qdt * 2.0 / (x->dx[d][c] + x->dx[d][c + 1]);
    }
    if ((il == 0 || il == 1) && x->p->courantpar > 0.0 && x->istep == x->nstep) {
      double cm = 0.0;
      int first = 1;
      LOOP(icl, ich) {
        long idx = IDX(x);
        double v = fabs(x->adtdx[idx]);
        cm = first ? v : dmax(cm, v);
        first = 0;
      }
      if (cm > 1.0) x->s->nerror[2]++;
      if (cm > SMALLD) x->s->dtcourant[d] = dmin(x->s->dtcourant[d], qdt * x->p->courantpar / cm);
    }
    if (mu) {  /* getphi, method 'tvdmu', with the entropy fix (:227-240) */
      LOOP(cl, ch) {
        long idx = IDX(x);
        double av = x->adtdx[idx], sa = x->smalla[idx];
        if (!fixed || fabs(av) >= sa) x->phi[idx] = fabs(av) * x->jump[idx];
        else x->phi[idx] = 0.5 * (sa + av * av / sa) * x->jump[idx];
      }
    } else {
    /* getphi, method 'tvd', with the entropy fix (:244-256) */
    LOOP(icl, ich) {
      long idx = IDX(x);
      double av = x->adtdx[idx], sa = x->smalla[idx];
      if (!fixed || fabs(av) >= sa) x->phi[idx] = fabs(av) - av * av;
      else x->phi[idx] = 0.5 * sa + (0.5 / sa - 1.0) * (av * av);
    }
    int ll[3];
    for (int dd = 0; dd < 3; ++dd) ll[dd] = icl[dd];
    ll[d] = icl[d] + 1;
    LOOP(ll, ich) {
      long idx = IDX(x);
      x->ljump[idx] = vacref_limit(x->p->limiter[il], x->jump[idx], x->jump[idx - s]);
    }
    LOOP(cl, ch) {
      long idx = IDX(x);
      double av = x->adtdx[idx];
      if (av <= 0.0) x->phi[idx] = x->phi[idx] * (x->jump[idx] - x->ljump[idx + s]);
      else x->phi[idx] = x->phi[idx] * (x->jump[idx] - x->ljump[idx]);
      x->phi[idx] = x->phi[idx] + av * av * x->jump[idx];
    }
    }
    for (int iw = 0; iw < x->nw; ++iw) {
      if (x->mhd) roe_rtimes(x, x->phi, x->wroe, cl, ch, iw, il, d, x->rphi);
      else roe_rtimes_hd(x, x->phi, x->wroe, cl, ch, iw, il, d, x->rphi);
      LOOP(cl, ch) {
        long idx = IDX(x);
        x->rphi[idx] = x->rphi[idx] * 0.25 * (dvol(x, idx + s) + dvol(x, idx));
      }
      setnoflux(x, iw, d, x->rphi, x->mlo[d] - 1, x->mhi[d]);
      double* wn = PL(x, wnew, iw);
      LOOP(x->mlo, x->mhi) {  /* addflux with qdt<0 (src/vacgrid.t:716) */
        long idx = IDX(x);
        wn[idx] = wn[idx] + (x->rphi[idx] - x->rphi[idx - s]) / dvol(x, idx);
      }
    }
  }
}

/* centdiff (src/vaccd.t:5-147), slab, the TVD pre-processing branch */
static void centdiff(ctx_t* x, double qdt, int d, double* wCT, double* w) {
  int tvdpreproc = (x->p->typefull == VACREF_TVD) && x->p->sourceunsplit;
  if (tvdpreproc && x->firstsweep) {
    if (x->p->sourceunsplit) addsource(x, qdt / 2, 1, w, wCT);
    getboundary(x, wCT);
  }
  int lo[3], hi[3];
  for (int dd = 0; dd < 3; ++dd) { lo[dd] = x->mlo[dd]; hi[dd] = x->mhi[dd]; }
  lo[d] -= 1; hi[d] += 1;  /* ix = ixO +- 1 */
  long s = x->st[d];
  for (int iw = 0; iw < x->nw; ++iw) {
    getflux_full(x, wCT, lo, hi, iw, d, x->f);
    int cl[3], ch[3];
    for (int dd = 0; dd < 3; ++dd) { cl[dd] = x->mlo[dd]; ch[dd] = x->mhi[dd]; }
    cl[d] -= 1;
    LOOP(cl, ch) {
      long idx = IDX(x);
      x->fC[idx] = 0.5 * (x->f[idx] + x->f[idx + s]);
    }
    setnoflux(x, iw, d, x->fC, x->mlo[d] - 1, x->mhi[d]);
    double* wn = PL(x, w, iw);
    LOOP(x->mlo, x->mhi) {
      long idx = IDX(x);
      wn[idx] = wn[idx] - qdt / x->dx[d][CRD(d)] * (x->fC[idx] - x->fC[idx - s]);
    }
  }
}

/* tvdlimit (src/vactvd.t:5-83) including the Ryu source step */
static void tvdlimit(ctx_t* x, double qdt, int d, double* wCT, double* wnew) {
  /* wL = wCT, wR = wCT shifted by one cell along d */
  long s = x->st[d];
  for (int iw = 0; iw < x->nw; ++iw) {
    const double* pc = PL(x, wCT, iw);
    double* pr = PL(x, x->wRC, iw);
    for (long idx = 0; idx + s < x->nc; ++idx) pr[idx] = pc[idx + s];
  }
  tvdlimit2(x, qdt, d, wCT, x->wRC, wnew, 0);
  if (x->lastsweep && x->p->sourceunsplit) {
    /* Ryu et al. (ApJ 452, 364): whalf=(wold+whydro+S)/2, wnew=whydro+S(whalf) */
    double* wt = wCT;  /* VAC overwrites its w argument (the wCT array) */
    for (int iw = 0; iw < x->nw; ++iw) {
      double* a = PL(x, wt, iw);
      const double* b = PL(x, wnew, iw);
      LOOP(x->mlo, x->mhi) { long idx = IDX(x); a[idx] = b[idx]; }
    }
    addsource(x, qdt, 1, x->wold, wt);
    for (int iw = 0; iw < x->nw; ++iw) {
      double* a = PL(x, wt, iw);
      const double* o = PL(x, x->wold, iw);
      LOOP(x->mlo, x->mhi) { long idx = IDX(x); a[idx] = 0.5 * (o[idx] + a[idx]); }
    }
    addsource(x, qdt, 0, wt, wnew);
  }
}

/* ------------------------------------------------------ divB projection */
enum { STOP_MAX = 0, STOP_REL = 1 };

/* the region ixM^L^LADD1: the mesh widened by one cell in every dimension */
static void widen1(const ctx_t* x, int* lo, int* hi) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = x->mlo[d]; hi[d] = x->mhi[d];
    if (d < x->ndim) { lo[d] -= 1; hi[d] += 1; }
  }
}

/* gradient (src/vacgrid.t:775-835), slab branch :815-821:
 * gradq = half*(q(i+1) - q(i-1))/dx(i) along d over [lo, hi] */
static void gradient(ctx_t* x, const double* q, const int* lo, const int* hi, int d, double* g) {
  long s = x->st[d];
  LOOP(lo, hi) {
    long idx = IDX(x);
    g[idx] = 0.5 * (q[idx + s] - q[idx - s]) / x->dx[d][CRD(d)];
  }
}

/* getdivb (src/vacphys.mhd0.t:106-146), second order, over ixM */
static void getdivb(ctx_t* x, const double* w, double* divb) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  LOOP(lo, hi) divb[IDX(x)] = 0.0;
  for (int d = 0; d < x->ndim; ++d) {
    gradient(x, PL(x, w, B(x, d)), lo, hi, d, x->pr_qdx);
    LOOP(lo, hi) { long idx = IDX(x); divb[idx] = divb[idx] + x->pr_qdx[idx]; }
  }
}

/* typeBscalar of face iB from the type of the normal B component there
 * (src/vacproc.projectb.t:82-101): periodic -> periodic, symm -> asymm,
 * asymm -> symm, fixed -> grad0, anything else ('cont') -> nul. */
enum { SB_NUL = 0, SB_PERIODIC = 1, SB_SYMM = 2, SB_ASYMM = 3, SB_GRAD0 = 4 };
static int bscalar_type(const ctx_t* x, int iB) {
  switch (x->p->bc[iB][B(x, iB / 2)]) {
    case VACREF_PERIODIC: return SB_PERIODIC;
    case VACREF_SYMM: case VACREF_SYMM0: return SB_ASYMM;  /* src/vacproc.projectb.t:74 */
    case VACREF_ASYMM: return SB_SYMM;
    case VACREF_FIXED: return SB_GRAD0;
    case VACREF_FIXED1: return SB_GRAD0;
  }
  return SB_NUL;
}

/* boundscalar for 'periodic', 'symm', 'asymm', 'grad0' and 'nul'
 * (src/vacpoisson.t:158-300, slab grids), the faces in VAC's order, each
 * region spanning the full transverse ixG range */
static void boundscalar(ctx_t* x, double* q) {
  for (int d = 0; d < x->ndim; ++d) {
    for (int up = 0; up < 2; ++up) {
      int iB = 2 * d + up;
      int lo[3], hi[3];
      for (int dd = 0; dd < 3; ++dd) { lo[dd] = x->glo[dd]; hi[dd] = x->ghi[dd]; }
      if (up) { lo[d] = x->mhi[d] + 1; hi[d] = x->ghi[d]; }
      else { lo[d] = x->glo[d]; hi[d] = x->mlo[d] - 1; }
      int n = x->p->nx[d];
      int type = bscalar_type(x, iB);
      LOOP(lo, hi) {
        long idx = IDX(x);
        int c = CRD(d);
        int src = c;
        double sg = 1.0;
        switch (type) {
          case SB_PERIODIC: src = up ? c - n : c + n; break;  /* pair face shifted by dixB, :183-194 */
          case SB_SYMM: case SB_ASYMM:                        /* :226-253 */
            src = up ? 2 * (x->mhi[d] + 1) - 1 - c : 2 * (x->mlo[d] - 1) + 1 - c;
            if (type == SB_ASYMM) sg = -1.0;
            break;
          case SB_GRAD0:  /* 0 in the first ghost layer, the edge value in the second, :254-285 */
            if (c == (up ? x->mhi[d] + 1 : x->mlo[d] - 1)) { q[idx] = 0.0; continue; }
            src = up ? x->mhi[d] : x->mlo[d];
            break;
          default: q[idx] = 0.0; continue;                    /* 'nul', :286-287 */
        }
        double v = q[idx + (long)(src - c) * x->st[d]];
        q[idx] = sg == 1.0 ? v : -v;
      }
    }
  }
}

/* matvec_poisson (src/vacpoisson.t:100-155), second order:
 * qy = sum_d d/dx_d (d qx/dx_d), the gradient over ixM+1, its divergence over ixM */
static void matvec_poisson(ctx_t* x, double* qx, double* qy) {
  boundscalar(x, qx);
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  int lo1[3], hi1[3];
  widen1(x, lo1, hi1);
  LOOP(lo, hi) qy[IDX(x)] = 0.0;
  for (int d = 0; d < x->ndim; ++d) {
    gradient(x, qx, lo1, hi1, d, x->pr_qdx);
    gradient(x, x->pr_qdx, lo, hi, d, x->pr_qddx);
    LOOP(lo, hi) { long idx = IDX(x); qy[idx] = qy[idx] + x->pr_qddx[idx]; }
  }
}

/* sum(q(ixM)**2), maxval(abs(q(ixM))), sum(a(ixM)*b(ixM)): Fortran array
 * reductions in array element order */
static double sum_sq(ctx_t* x, const double* q) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double r = 0.0;
  LOOP(lo, hi) { long idx = IDX(x); r = r + q[idx] * q[idx]; }
  return r;
}
static double max_abs(ctx_t* x, const double* q) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double r = -INFINITY;
  LOOP(lo, hi) r = dmax(r, fabs(q[IDX(x)]));
  return r;
}
static double sum_prod(ctx_t* x, const double* a, const double* b) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double r = 0.0;
  LOOP(lo, hi) { long idx = IDX(x); r = r + a[idx] * b[idx]; }
  return r;
}

/* cgscalar (src/vacpoisson.t:391-595) with a zero initial guess
 * (nonzero=.false.), typestop 'max' or 'rel'.  rhs: right-hand side in,
 * residual out; qx: solution; iter: matvec limit in, iterations out; tol:
 * tolerance in, achieved residual out. */
static void cgscalar(ctx_t* x, double* rhs, double* qx, int* iter, double* tol, int typestop,
                     int* info) {
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  double* z = x->pr_z;
  double* az = x->pr_az;
  double assumedzero = 1.0e-16;
  int itr = 0;
  double rho = sqrt(sum_sq(x, rhs));
  double res0 = rho;
  if (typestop == STOP_MAX) res0 = max_abs(x, rhs);
  double res = res0;
  assumedzero = assumedzero * res0;
  if (res0 < SMALLD || (typestop != STOP_REL && res0 <= *tol)) {
    *info = 3;
  } else {
    rho = rho * rho;
    LOOP(lo, hi) { long idx = IDX(x); z[idx] = rhs[idx]; }
    for (;;) {
      matvec_poisson(x, z, az);
      double alf = sum_prod(x, z, az);
      if (fabs(alf) <= assumedzero * assumedzero) { *info = 1; break; }
      alf = rho / alf;
      LOOP(lo, hi) { long idx = IDX(x); qx[idx] = qx[idx] + alf * z[idx]; }
      LOOP(lo, hi) { long idx = IDX(x); rhs[idx] = rhs[idx] - alf * az[idx]; }
      double rhonew = sqrt(sum_sq(x, rhs));
      if (typestop == STOP_MAX) res = max_abs(x, rhs);
      else res = rhonew / res0;
      rhonew = rhonew * rhonew;
      itr = itr + 1;
      if (res <= *tol) { *info = 0; break; }
      if (itr >= *iter) { *info = 2; break; }
      if (rho <= assumedzero * assumedzero) { *info = 1; break; }
      double bet = rhonew / rho;
      LOOP(lo, hi) { long idx = IDX(x); z[idx] = rhs[idx] + bet * z[idx]; }
      rho = rhonew;
    }
  }
  *iter = itr;
  *tol = res;
  if ((typestop == STOP_REL && res > 1.0) || (typestop != STOP_REL && res > res0)) *info = -*info;
}

/* projectb (src/vacproc.projectb.t:4-157) with the Poisson driver's CG
 * branch (src/vacpoisson.t:6-97; slab grids select 'cg').  Returns -4 where
 * VAC dies (a zero allowed maximum of div B). */
static int projectb(ctx_t* x, double* w) {
  const vacref_params* p = x->p;
  vacref_state* s = x->s;
  double* divb = x->pr_rhs;
  double* phi = x->pr_phi;
  const int* lo = x->mlo;
  const int* hi = x->mhi;
  getdivb(x, w, divb);
  if (s->divbmax < SMALLD) {
    /* the first call fixes divbmax, typestop and matvecmax, :29-58 */
    s->divbmax = max_abs(x, divb);
    double coeff = p->procpar[0] == -1.0 ? 1.0 / 10 : p->procpar[0];
    double cnst = p->procpar[1] == -1.0 ? 0.0 : p->procpar[1];
    if (coeff < 0.0 && coeff > -1.0) {
      s->divbmax = -coeff;
      s->typestop = STOP_REL;
    } else {
      s->divbmax = s->divbmax * dmax(0.0, coeff) + dmax(0.0, cnst);
      s->typestop = STOP_MAX;
      if (s->divbmax < SMALLD) return -4;
    }
    s->matvecmax = 1000;
    if (cnst < -1.0) s->matvecmax = (int)lround(-cnst);
    if (coeff < -1.0) s->matvecmax = (int)lround(-coeff);
  }
  LOOP(lo, hi) phi[IDX(x)] = 0.0;
  int iter = s->matvecmax, info = 0;
  double resid = s->divbmax;
  cgscalar(x, divb, phi, &iter, &resid, s->typestop, &info);
  if (info != 0 && info != 3) s->nerror[3] += 1;  /* nerror(poissonerr_), :79-81 */
  s->proj_info = info;
  s->proj_iter = iter;
  s->proj_resid = resid;
  s->nproj += 1;
  if (info == 3 || info < 0) return 0;
  boundscalar(x, phi);
  int lo1[3], hi1[3];
  widen1(x, lo1, hi1);
  double* g = x->pr_qdx;
  const int bnd = (int)lround(fabs(p->procpar[2]));
  const double gam = p->gamma;
  for (int d = 0; d < x->ndim; ++d) {
    gradient(x, phi, lo1, hi1, d, g);
    double* pe = PL(x, w, E_(x));
    double* pb = PL(x, w, B(x, d));
    if (bnd == 2) {
      /* keep the thermal pressure, :129-131 */
      LOOP(lo1, hi1) { long idx = IDX(x); pe[idx] = pe[idx] - pb[idx] * g[idx] + g[idx] * g[idx] / 2; }
    } else if (bnd == 3 && gam != 2.0 && gam != 1.0) {
      /* keep the total pressure, :132-135 */
      LOOP(lo1, hi1) {
        long idx = IDX(x);
        pe[idx] = pe[idx] + (gam - 2.0) / (gam - 1.0) * (-(pb[idx] * g[idx]) + g[idx] * g[idx] / 2);
      }
    }
    LOOP(lo1, hi1) { long idx = IDX(x); pb[idx] = pb[idx] - g[idx]; }
  }
  if (p->procpar[2] > 0.0) getboundary_vars(x, w, B(x, 0), B(x, 0) + x->ndim);
  return 0;
}

/* -------------------------------------------------------- time stepping */
/* advect1 (src/vac.t:489-572) */
static void advect1(ctx_t* x, int method, double qdt, int d0, int d1, double* wCT, double* w) {
  x->istep += 1;
  /* the directions d0..d1 in turn, the unsplit source after the last */
  for (int d = d0; d <= d1; ++d) {
    const int sm = d == d1 ? d1 - d0 + 1 : 0;
    switch (method) {
      case 0: hancock(x, qdt, d, wCT, w, sm); break;              /* 'hancock' */
      case 1: tvdlf_corrector(x, qdt, d, wCT, w, 0, sm); break;   /* 'tvdlf' */
      case 3: tvdlf_corrector(x, qdt, d, wCT, w, 1, sm); break;   /* 'tvdmu' */
      case 2: centdiff(x, qdt, d, wCT, w); tvdlimit(x, qdt, d, wCT, w); break;  /* 'tvd' */
    }
  }
  getboundary(x, w);
}

/* advect (src/vac.t:363-486) */
static int advect(ctx_t* x, int d0, int d1, double* w) {
  /* process(1) when due, :392-399 */
  const int np = x->p->nproc1;
  if (np != 0) {
    const int di = x->s->it - x->s->itmin;
    if (np < 0 || (x->firstsweep && di == (di / np) * np)) {
      int rc = projectb(x, w);
      if (rc) return rc;
    }
  }
  memcpy(x->w1, w, sizeof(double) * x->nc * x->nw);
  x->istep = 0;
  double dt = x->s->dt;
  if (x->p->typefull == VACREF_TVDLF || x->p->typefull == VACREF_TVDMU) {
    advect1(x, 0, dt / 2, d0, d1, w, x->w1);
    advect1(x, x->p->typefull == VACREF_TVDMU ? 3 : 1, dt, d0, d1, x->w1, w);
  } else {
    advect1(x, 2, dt, d0, d1, x->w1, w);
  }
  x->firstsweep = 0;
  return 0;
}

/* advance_expl with dimsplit (src/vac.t:283-360) */
static int advance(ctx_t* x, double* w) {
  x->firstsweep = 1;
  int it = x->s->it, rc = 0;
  if (x->p->unsplit) {
    /* fluxes from all directions at once (src/vac.t:325-329) */
    x->lastsweep = 1;
    return advect(x, 0, x->ndim - 1, w);
  }
  if ((it / 2) * 2 == it || x->p->dimsplit_xy) {
    for (int d = 0; d < x->ndim && !rc; ++d) { x->lastsweep = (d == x->ndim - 1); rc = advect(x, d, d, w); }
  } else {
    for (int d = x->ndim - 1; d >= 0 && !rc; --d) { x->lastsweep = (d == 0); rc = advect(x, d, d, w); }
  }
  return rc;
}

/* getdt_courant (src/vac.t:664-722) */
static void getdt_courant(ctx_t* x, const double* w) {
  double dt = BIGD, cm = 0.0;
  for (int d = 0; d < x->ndim; ++d) {
    if (x->s->dtcourant[d] < BIGD) {
      dt = dmin(dt, x->s->dtcourant[d]);
    } else {
      getcmax(x, w, x->mlo, x->mhi, d, x->sc1);
      double mx = -BIGD;
      const int* lo = x->mlo;
      const int* hi = x->mhi;
      LOOP(lo, hi) {
        long idx = IDX(x);
        mx = dmax(mx, x->sc1[idx] / x->dx[d][CRD(d)]);
      }
      cm = dmax(cm, mx);
    }
  }
  for (int d = 0; d < x->ndim; ++d) x->s->dtcourant[d] = BIGD;
  if (cm > SMALLD) dt = dmin(dt, x->p->courantpar / cm);
  x->s->dt = dt;
}

static int ctx_init(ctx_t* x, const vacref_params* p, vacref_state* s,
                    const double* dx0, const double* dx1, const double* dx2) {
  memset(x, 0, sizeof(*x));
  if (p->ndim < 1 || p->ndim > 3) return -1;
  x->p = p;
  x->s = s;
  x->ndim = p->ndim;
  x->ndir = p->ndim;
  x->mhd = (p->phys == VACREF_MHD);
  x->nw = x->mhd ? 2 + 2 * x->ndir : 2 + x->ndim;

  const double* dxs[3] = {dx0, dx1, dx2};
  long st = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < x->ndim) {
      if (p->nx[d] < 2) return -1;
      x->nG[d] = p->nx[d] + 4;
      x->mlo[d] = 2; x->mhi[d] = p->nx[d] + 1;
      x->glo[d] = 0; x->ghi[d] = p->nx[d] + 3;
      x->dx[d] = dxs[d];
      x->st[d] = st;
      st *= x->nG[d];
    } else {
      x->nG[d] = 1; x->mlo[d] = x->mhi[d] = x->glo[d] = x->ghi[d] = 0;
      x->st[d] = 0;
    }
  }
  x->nc = st;
  x->nstep = (p->typefull == VACREF_TVD) ? 1 : 2;
  size_t pl = sizeof(double) * (size_t)x->nc;
  x->w1 = calloc(x->nw, pl);
  x->wold = calloc(x->nw, pl);
  x->wLC = calloc(x->nw, pl);
  x->wRC = calloc(x->nw, pl);
  x->wroe = calloc(x->nw, pl);
  x->tmpw = calloc(x->nw, pl);
  double** one[] = {&x->f, &x->fC, &x->sc1, &x->sc2, &x->sc3, &x->sc4, &x->sc5,
                    &x->jump, &x->adtdx, &x->phi, &x->rphi, &x->ljump,
                    &x->cfast, &x->cslow, &x->afast, &x->aslow, &x->csound2, &x->dp,
                    &x->rhodv, &x->bdv, &x->bdb, &x->bv, &x->v2a2,
                    &x->pr_rhs, &x->pr_phi, &x->pr_z, &x->pr_az, &x->pr_qdx, &x->pr_qddx,
                    &x->smalla, &x->eaL, &x->eaR, &x->cs2L, &x->cs2R, &x->cs2ca2L, &x->cs2ca2R};
  for (size_t q = 0; q < sizeof(one) / sizeof(one[0]); ++q) *one[q] = calloc(1, pl);
  int anyfixed = 0;
  for (int f = 0; f < 2 * x->ndim; ++f)
    for (int iw = 0; iw < x->nw; ++iw)
      anyfixed |= p->bc[f][iw] == VACREF_FIXED || p->bc[f][iw] == VACREF_FIXED1 ||
                  p->bc[f][iw] == VACREF_GRAD1;
  x->fixw = anyfixed ? calloc(x->nw, pl) : NULL;
  x->fix_ready = 0;
  return 0;
}

static void ctx_free(ctx_t* x) {
  double* all[] = {x->w1, x->wold, x->wLC, x->wRC, x->wroe, x->tmpw, x->f, x->fC,
                   x->sc1, x->sc2, x->sc3, x->sc4, x->sc5, x->jump, x->adtdx, x->phi,
                   x->rphi, x->ljump, x->cfast, x->cslow, x->afast, x->aslow,
                   x->csound2, x->dp, x->rhodv, x->bdv, x->bdb, x->bv, x->v2a2,
                   x->pr_rhs, x->pr_phi, x->pr_z, x->pr_az, x->pr_qdx, x->pr_qddx, x->fixw,
                   x->smalla, x->eaL, x->eaR, x->cs2L, x->cs2R, x->cs2ca2L, x->cs2ca2R};
  for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
}

int vacref_getboundary(const vacref_params* p, double* w) {
  ctx_t x;
  vacref_state s;
  memset(&s, 0, sizeof(s));
  double one = 1.0;
  int rc = ctx_init(&x, p, &s, &one, &one, &one);
  if (rc) return rc;
  getboundary(&x, w);
  ctx_free(&x);
  return 0;
}

int vacref_projectb(const vacref_params* p, vacref_state* s, double* w,
                    const double* dx0, const double* dx1, const double* dx2) {
  ctx_t x;
  int rc = ctx_init(&x, p, s, dx0, dx1, dx2);
  if (rc) return rc;
  if (!x.mhd || x.ndim < 2) { ctx_free(&x); return -5; }
  rc = projectb(&x, w);
  ctx_free(&x);
  return rc;
}

int vacref_getdt(const vacref_params* p, vacref_state* s, const double* w,
                 const double* dx0, const double* dx1, const double* dx2) {
  ctx_t x;
  int rc = ctx_init(&x, p, s, dx0, dx1, dx2);
  if (rc) return rc;
  getdt_courant(&x, w);
  ctx_free(&x);
  return 0;
}

/* main loop (src/vac.t:54-106) */
int vacref_run(const vacref_params* p, vacref_state* s, double* w,
               const double* dx0, const double* dx1, const double* dx2,
               int nsteps, double* t_hist) {
  ctx_t x;
  int rc = ctx_init(&x, p, s, dx0, dx1, dx2);
  if (rc) return rc;
  if (p->dtpar > 0.0 && s->smallp < 0.0) { ctx_free(&x); return -3; }
  if (p->nproc1 != 0 && (!x.mhd || x.ndim < 2)) { ctx_free(&x); return -5; }  /* ^NOONED, MHD */
  /* 'ratio' needs the MHD wave speeds: the HD solver leaves them unset */
  if (!x.mhd)
    for (int il = 0; il < x.nw; ++il)
      if (p->typeentropy[il] == VACREF_ENT_RATIO) { ctx_free(&x); return -7; }
  /* TVD with typelimited='predictor' limits the centdiff result (src/vactvd.t:30-36) */
  if (p->typefull == VACREF_TVD && p->typelimited == 2) { ctx_free(&x); return -6; }
  /* unsplit sweeps are restated for TVDLF and TVD-MUSCL (one-step TVD
   * unsplit is a warning in VAC, src/vacio.t:508-509) */
  if (p->typefull == VACREF_TVD && p->unsplit) { ctx_free(&x); return -8; }
  getboundary(&x, w);
  int done = 0;
  double dtpar = p->dtpar, dtold = 0.0;
  for (;;) {
    if (t_hist) t_hist[done] = s->t;
    /* the time step (src/vac.t:60-81) */
    if (dtpar > 0.0) {
      s->dt = dtpar;
    } else {
      if (p->courantpar > 0.0) getdt_courant(&x, w);
      if (p->dtcantgrow && s->it > s->itmin) s->dt = dmin(s->dt, dtold);
      dtold = s->dt;
    }
    if (s->it >= p->itmax || s->t >= p->tmax || s->dt < 1.0e-99 || done >= nsteps) break;
    if (p->slowsteps == 1 && s->it == s->itmin) dtpar = -1.0;
    if (p->slowsteps > s->it - s->itmin + 1) {
      const double q = ((double)(s->it - s->itmin) + 1.0) / (double)p->slowsteps;
      const double a = 1.0 - q;
      s->dt = s->dt * (1.0 - a * a);
    }
    if (p->tmaxexact) s->dt = dmin(s->dt, p->tmax - s->t + SMALLD);
    memcpy(x.wold, w, sizeof(double) * x.nc * x.nw);
    if ((rc = advance(&x, w))) { ctx_free(&x); return rc; }
    s->it += 1;
    s->t = s->t + s->dt;
    done += 1;
  }
  ctx_free(&x);
  return done;
}

/* roetest (src/roetest.t:38-57) on a single interface */
int vacref_roe_identity(int ndim, double gamma, int idim, int divbwave, double smallp,
                        const double* wL, const double* wR, double* out, double* a) {
  vacref_params p;
  memset(&p, 0, sizeof p);
  p.ndim = ndim; p.phys = VACREF_MHD; p.gamma = gamma; p.typefull = VACREF_TVD;
  p.divbwave = divbwave; p.smallpcoeff = 1e-6;
  for (int d = 0; d < ndim; ++d) p.nx[d] = 2;
  vacref_state s;
  memset(&s, 0, sizeof s);
  s.smallp = smallp;
  double dxa[6] = {1, 1, 1, 1, 1, 1};
  ctx_t x;
  if (ctx_init(&x, &p, &s, dxa, dxa, dxa)) return -1;
  /* one interface at cell 0 */
  double* L = x.tmpw;
  double* R = x.wRC;
  for (int iw = 0; iw < x.nw; ++iw) { PL(&x, L, iw)[0] = wL[iw]; PL(&x, R, iw)[0] = wR[iw]; }
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  roe_average(&x, L, R, lo, hi, idim, x.wroe);
  for (int iw = 0; iw < x.nw; ++iw) out[iw] = 0.0;
  for (int il = 0; il < x.nw; ++il) {
    roe_eigenjump(&x, L, R, x.wroe, lo, hi, il, idim, x.adtdx, x.jump);
    a[il] = x.adtdx[0];
    for (int iw = 0; iw < x.nw; ++iw) {
      roe_rtimes(&x, x.jump, x.wroe, lo, hi, iw, il, idim, x.rphi);
      out[iw] += x.rphi[0];
    }
  }
  ctx_free(&x);
  return 0;
}
