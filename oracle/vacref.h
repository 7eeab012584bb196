/*
 * oracle/vacref.h -- CPU restatement of VAC's explicit TVDLF / TVD / TVD-MUSCL (Roe)
 * advect path for ideal MHD and HD.  TEST INFRASTRUCTURE ONLY: tests, the
 * smoke check and the CPU-baseline leg of bench.py use it as the checker;
 * the product (libvacgpu) never links or calls it.
 *
 * The state layout is VAC's w(ixG^T, nw) with the leading dimensions equal
 * to the actual grid: nw planes of prod(nx[d]+4) doubles, x fastest.
 * Mesh cells are indices [2, nx[d]+1] (0-based); two ghost layers per side.
 */
#ifndef VACREF_H
#define VACREF_H

#ifdef __cplusplus
extern "C" {
#endif

enum { VACREF_HD = 0, VACREF_MHD = 1 };
enum { VACREF_TVDLF = 0, VACREF_TVD = 1, VACREF_TVDMU = 2 };
enum { VACREF_MINMOD = 0, VACREF_WOODWARD = 1, VACREF_SUPERBEE = 2 };
/* typeB (src/vacgrid.t:195-413); 'symm0' fills ghost cells as 'symm' does
 * and zeroes the flux through the face (setnoflux, src/vacgrid.t:416-449) */
enum { VACREF_CONT = 0, VACREF_PERIODIC = 1, VACREF_SYMM = 2, VACREF_ASYMM = 3, VACREF_FIXED = 4,
       VACREF_CONT1 = 5, VACREF_FIXED1 = 6, VACREF_GRAD1 = 7, VACREF_SYMM0 = 8 };

typedef struct {
  int ndim;              /* 1..3; ndir == ndim (setvac -d=11/22/33) */
  int phys;              /* VACREF_HD / VACREF_MHD */
  int nx[3];             /* mesh cells per dimension */
  double gamma;          /* eqpar(gamma_) */
  int typefull;          /* VACREF_TVDLF / VACREF_TVDMU (twostep, hancock), VACREF_TVD (onestep) */
  int limiter[8];        /* typelimiter(1:nw) */
  int bc[6][8];          /* typeB(iw, iB), iB = 2*d + upper */
  int divbfix, divbwave, sourceunsplit;
  int dimsplit_xy;       /* typedimsplit='xy' (else 'xyyx') */
  double courantpar;
  double smallpcoeff;
  double dtpar;          /* > 0: fixed time step */
  double tmax;
  int tmaxexact;
  int itmax;
  /* processing before the step (process(1), src/vac.t:392-399): nproc(1);
   * 0 = none, n > 0 = every n-th step on the first sweep, n < 0 = every
   * sweep.  The only processing of an MHD grid is the projection of B
   * (projectb, src/vacproc.projectb.t) with CG (src/vacpoisson.t:391-595). */
  int nproc1;
  double procpar[3];     /* procpar(divbcoeff_, divbconst_, divbbound_); -1 = unset */
  /* typelimited: the state the corrector of TVDLF / TVD-MUSCL limits
   * (src/vactvdlf.t:150-161): 0 'original' (the sweep's input), 1 'previous'
   * (wold, the step's start), 2 'predictor' (wCT); for TVD 'previous' limits
   * wCT as 'original' does (src/vactvd.t:37-38) */
  int typelimited;
  /* typeentropy(il) per characteristic wave (src/vactvd.t:422-449, the wave
   * speeds of src/vacphys.mhdroe.t:228-281 and src/vacphys.hdroe.t:120-147):
   * VACREF_ENT_NUL, _HARTEN, _POWELL, _RATIO (MHD), _YEE; entropycoef(il) is
   * 'yee''s lower limit of |a dt/dx| */
  int typeentropy[8];
  double entropycoef[8];
  /* useprimitive: upwindLR limits the primitive variables (src/vactvdlf.t:
   * 328-332, :360-364), converting its base state in place and back; the MHD
   * Roe average takes dp = pR - pL (src/vacphys.mhdroe.t:42-44) */
  int useprimitive;
  /* unsplit: dimsplit=F, one advect over all directions per step
   * (src/vac.t:325-329) with 'original' limiting read as 'previous'
   * (src/vacio.t:363-364); TVDLF and TVD-MUSCL */
  int unsplit;
  /* &paramlist time-step controls of the main loop (src/vac.t:68-80):
   * dtcantgrow = dt may not exceed the previous step's Courant dt;
   * slowsteps > 1 scales dt by 1-(1-(it-itmin+1)/slowsteps)^2 in the first
   * steps, slowsteps = 1 takes dtpar for the first step only */
  int dtcantgrow;
  int slowsteps;
} vacref_params;

enum { VACREF_ENT_NUL = 0, VACREF_ENT_HARTEN = 1, VACREF_ENT_POWELL = 2, VACREF_ENT_RATIO = 3,
       VACREF_ENT_YEE = 4 };

typedef struct {
  int it, itmin;
  double t, dt;
  double smallp;         /* < 0: set on the first pressure evaluation */
  double dtcourant[3];
  int nerror[4];
  /* projectb's saved variables (src/vacproc.projectb.t:16-17): divbmax = 0
   * until the first call sets it, with matvecmax and typestop (0 'max',
   * 1 'rel'); and what the last Poisson solve returned */
  double divbmax;
  int matvecmax, typestop;
  int proj_info, proj_iter;
  double proj_resid;
  int nproj;             /* projections done */
} vacref_state;

/* Main loop of src/vac.t:55-106 for up to nsteps steps (and the itmax/tmax
 * stop conditions).  w holds the full ixG state (ghost cells are filled on
 * entry, as VAC does at src/vac.t:54).  dx[d] are the nx[d]+4 cell widths.
 * t_hist (optional, nsteps+1 doubles) receives t before each step and after
 * the last one.  Returns the number of steps taken, or -1 on bad input. */
int vacref_run(const vacref_params* p, vacref_state* s, double* w,
               const double* dx0, const double* dx1, const double* dx2,
               int nsteps, double* t_hist);

/* getboundary (src/vacgrid.t:195-413) of every variable of w. */
int vacref_getboundary(const vacref_params* p, double* w);

/* One projectb call (src/vacproc.projectb.t:4-157) on the full ixG state w,
 * with p->procpar and the saved variables in s.  Returns 0, or < 0 on bad
 * input (HD, ndim < 2). */
int vacref_projectb(const vacref_params* p, vacref_state* s, double* w,
                    const double* dx0, const double* dx1, const double* dx2);

/* One pass of VAC's getdt_courant (src/vac.t:664-722): sets s->dt. */
int vacref_getdt(const vacref_params* p, vacref_state* s, const double* w,
                 const double* dx0, const double* dx1, const double* dx2);

/* dwlimiter2 on one pair (src/vacgrid.t:938-996): a = dw(i), b = dw(i-1). */
double vacref_limit(int type, double a, double b);

/* roetest identity (src/roetest.t:38-57): out[iw] = sum_il R(iw,il) *
 * jump(il) for the MHD Roe matrices at the arithmetic average of wL, wR
 * (src/vacphys.mhdroe.t).  ndim in {2,3} selects nw = 6 / 8.  Also returns
 * the wave speeds a[il].  smallp is the pressure floor used by the average. */
int vacref_roe_identity(int ndim, double gamma, int idim, int divbwave,
                        double smallp, const double* wL, const double* wR,
                        double* out, double* a);

#ifdef __cplusplus
}
#endif
#endif
