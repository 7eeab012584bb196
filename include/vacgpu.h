/*
 * vacgpu.h -- C ABI of the MI355X (gfx950) advect path for VAC.
 *
 * One handle holds one grid whose state stays resident in HBM between calls.
 * Each entry point below replaces a piece of the reference's host loop for
 * the TVDLF / TVD (Roe) step of the ideal MHD and HD modules; the citations
 * are to the reference sources (VAC 4.52, src/).
 *
 * All functions return 0 on success and a negative VACGPU_E* code on
 * failure; vacgpu_last_error() then returns a message, which a Fortran host
 * passes to its die() (src/vacio.t:1108-1116).  Nothing falls back to a CPU
 * path: an unsupported selector is an error.
 *
 * Arrays are passed as plain pointers in VAC's layout: w(ixG^T, nw) column
 * major, x fastest, with caller-chosen leading dimensions ld[] (VAC's
 * compile-time ixGhi^D).  Index 0 of each dimension is VAC's ixGlo = 1.
 */
#ifndef VACGPU_H
#define VACGPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VACGPU_OK = 0,
  VACGPU_E_ARG = -1,         /* bad argument or unsupported selector */
  VACGPU_E_HIP = -2,         /* HIP runtime error */
  VACGPU_E_STATE = -3,       /* call out of order (e.g. advect before put_state) */
  VACGPU_E_NOMEM = -4
};

enum { VACGPU_HD = 0, VACGPU_MHD = 1 };                 /* setvac -p= */
enum { VACGPU_TVDLF = 0, VACGPU_TVD = 1, VACGPU_TVDMU = 2 };  /* typefull */
enum { VACGPU_MINMOD = 0, VACGPU_WOODWARD = 1, VACGPU_SUPERBEE = 2 };  /* typelimiter */
enum { VACGPU_CONT = 0, VACGPU_PERIODIC = 1, VACGPU_SYMM = 2, VACGPU_ASYMM = 3,
       VACGPU_FIXED = 4, VACGPU_CONT1 = 5, VACGPU_FIXED1 = 6,
       VACGPU_GRAD1 = 7,                               /* typeB (src/vacgrid.t:234-404) */
       VACGPU_SYMM0 = 8 };                             /* 'symm', with zero flux through the face
                                                          (setnoflux, src/vacgrid.t:416-449) */
enum { VACGPU_FP64 = 0, VACGPU_FP32BULK = 1 };          /* arithmetic tier */
enum { VACGPU_LIM_ORIGINAL = 0, VACGPU_LIM_PREVIOUS = 1, VACGPU_LIM_PREDICTOR = 2 };  /* typelimited */
enum { VACGPU_ENT_NUL = 0, VACGPU_ENT_HARTEN = 1, VACGPU_ENT_POWELL = 2, VACGPU_ENT_RATIO = 3,
       VACGPU_ENT_YEE = 4 };                           /* typeentropy */

typedef struct vacgpu_grid vacgpu_grid;

/* Grid and physics: what VAC fixes at setvac time and in boundsetup /
 * gridsetup1 (src/vacgrid.t:9-160, :457-593). */
typedef struct {
  int ndim;             /* 2 or 3; ndir == ndim (setvac -d=22 / -d=33) */
  int physics;          /* VACGPU_MHD or VACGPU_HD */
  int nx[3];            /* mesh cells per dimension (the ixM extent) */
  double eqpar[2];      /* eqpar(gamma_), eqpar(eta_); eta must be 0 */
  int typeB[6][8];      /* typeB(iw, iB) with iB = 2*(idim-1) + upper */
  int precision;        /* VACGPU_FP64 or VACGPU_FP32BULK */
} vacgpu_config;

/* Method selectors read from the par file (&methodlist, &paramlist). */
typedef struct {
  int typefull;         /* VACGPU_TVDLF or VACGPU_TVDMU (twostep, hancock), VACGPU_TVD (onestep) */
  int typelimiter[8];   /* per variable (TVDLF) or per characteristic wave (TVD) */
  int divbfix;          /* Powell source terms (MHD) */
  int divbwave;         /* divergence wave in the Roe solver (MHD) */
  int sourceunsplit;
  int dimsplit_xy;      /* 1: typedimsplit='xy'; 0: 'xyyx' (sweep order flips with it) */
  double courantpar;
  double smallpcoeff;
  int typelimited;      /* the state the TVDLF / TVD-MUSCL corrector limits
                           (src/vactvdlf.t:150-161): VACGPU_LIM_ORIGINAL (the
                           sweep's input), VACGPU_LIM_PREVIOUS (the step's
                           start, wold), VACGPU_LIM_PREDICTOR (wCT); TVD takes
                           'original' or 'previous', which limit alike */
  int typeentropy[8];   /* entropy fix per characteristic wave (TVD, TVD-MUSCL;
                           src/vactvd.t:422-449): VACGPU_ENT_NUL, _HARTEN,
                           _POWELL, _RATIO (MHD only), _YEE */
  double entropycoef[8];/* 'yee''s lower limit per wave (VAC's physini
                           defaults: 0.2 fast/slow/sound, 0.4 Alfven, -1 else,
                           stored from single-precision literals) */
  int useprimitive;     /* 1: TVDLF / TVD-MUSCL limit the primitive variables
                           (upwindLR, src/vactvdlf.t:328-332, :360-364; FP64,
                           typelimited 'original'), and the MHD Roe solver of
                           TVD / TVD-MUSCL takes dp = pR - pL
                           (src/vacphys.mhdroe.t:42-44) */
  int unsplit;          /* 1: dimsplit=F, one advect over idim = 1..ndim per
                           step (src/vac.t:325-329), TVDLF and TVD-MUSCL;
                           'original' limiting reads as 'previous'
                           (src/vacio.t:363-364) */
  int dtcantgrow;       /* vacgpu_run's time step may not exceed the previous
                           Courant step (src/vac.t:68-69) */
  int slowsteps;        /* vacgpu_run: > 1 scales dt by 1-(1-(it-itmin+1)/
                           slowsteps)^2 over the first steps, 1 takes dtpar
                           for the first step only (src/vac.t:75-78); itmin
                           is the it of the last vacgpu_set_time */
} vacgpu_method;

/* Replaces the geometry hand-off after gridsetup1 (src/vacgrid.t:457-593):
 * dxd[i] are the cell widths along dimension d over ixG (nx[d]+4 values;
 * VAC's dx(:, ..., d) restricted to one line -- the grid is slab Cartesian). */
int vacgpu_create(const vacgpu_config* cfg, const double* dx0, const double* dx1,
                  const double* dx2, vacgpu_grid** out);

/* Optional: run on a caller-provided hipStream_t (default: a private stream). */
int vacgpu_set_stream(vacgpu_grid* g, void* hip_stream);

/* Copy the full ixG state in / out (src/vacio.t readfileini / savefile hand-off).
 * ld[d] is the leading dimension of the caller's array along d (VAC: ixGhi^D). */
int vacgpu_put_state(vacgpu_grid* g, const double* w, const int* ld);
/* As vacgpu_put_state, but the caller's ghost cells are kept as they are
 * (VAC's w always carries them, src/vac.t:54; after a projection they need
 * not be copies of mesh cells, src/vacproc.projectb.t:111-144), and the next
 * sweep reads them from memory.  The drop-in's transfer of VAC's w. */
int vacgpu_put_state_ghosts(vacgpu_grid* g, const double* w, const int* ld);
int vacgpu_get_state(vacgpu_grid* g, double* w, const int* ld);

/* getdt_courant (src/vac.t:664-722).  dtcourant (ndim, in/out) and smallp
 * (in/out, <0 means "set from the initial pressure minimum" as in
 * src/vacphys0.t.mhd:202-214) may be NULL to use the device-resident values.
 * On return *dt holds the Courant time step. */
int vacgpu_getdt_courant(vacgpu_grid* g, const vacgpu_method* m, double* dtcourant,
                         double* smallp, double* dt);

/* advect (src/vac.t:363-486) for one dimensional sweep idimmin == idimmax
 * (1-based, dimsplit), including advect1's getboundary (src/vac.t:567).
 * firstsweep / lastsweep are VAC's common/first/ flags (src/vac.t:298-299).
 * The first sweep of a step also records the state as wold (src/vac.t:85).
 * dtcourant (ndim) and nerror (4) are min-accumulated / incremented as in
 * src/vactvdlf.t:184-201 and src/vactvd.t:145-157; NULL keeps them on the
 * device (no host synchronisation).  smallp (in/out, may be NULL) is VAC's
 * pressure floor: a host that ran getdt_courant itself passes its value. */
int vacgpu_advect(vacgpu_grid* g, const vacgpu_method* m, int idimmin, int idimmax,
                  double dt, int firstsweep, int lastsweep, double* dtcourant, int* nerror,
                  double* smallp);

/* The time loop of src/vac.t:55-106 on the device for up to nsteps steps:
 * dt from getdt_courant (or dtpar > 0), tmaxexact, advance with it-parity
 * sweep order, it += 1, t += dt.  Stops at itmax or t >= tmax.  t_hist
 * (optional, nsteps+1) receives t before every step and after the last.
 * Synchronises with the host only when tmax is finite or t_hist is given. */
int vacgpu_run(vacgpu_grid* g, const vacgpu_method* m, int nsteps, int itmax, double tmax,
               int tmaxexact, double dtpar, double* t_hist);

/* Time and step counter used by vacgpu_run (VAC's t, it). */
int vacgpu_set_time(vacgpu_grid* g, int it, double t);
int vacgpu_get_time(vacgpu_grid* g, int* it, double* t);

/* Device-resident scalars: dtcourant[ndim], nerror[4], smallp, dt. Any may be NULL. */
int vacgpu_get_scalars(vacgpu_grid* g, double* dtcourant, int* nerror, double* smallp,
                       double* dt);

/* Per-kernel timing with HIP events on the grid's stream (0 = off). */
int vacgpu_profile(vacgpu_grid* g, int enable);
/* Average duration (ms) and launch count of kernel family k since enabling:
 * k = 0 TVD sweep, 1 TVDLF predictor, 2 TVDLF corrector, 3 Ryu source,
 * 4 ghost cells, 5 reductions, 6 fused TVDLF sweep, 7 projection of B (a
 * whole projectb call: its kernels and the host's polls of the CG state). */
int vacgpu_kernel_time(vacgpu_grid* g, int k, double* avg_ms, long* count);

/* Kernel choice for 2D TVDLF sweeps (no counterpart in VAC; results are
 * bitwise the same either way).  fused = 1 (the default) runs predictor and
 * corrector of a sweep in one launch when every boundary face has one type
 * for all variables; fused = 0 runs them as two launches with the predictor
 * state stored in between.  seglen = rows per marching segment of the fused
 * kernels (0 = automatic, else clamped to [4, 256]). */
int vacgpu_set_kernels(vacgpu_grid* g, int fused, int seglen);

/* Self-test of the quotient the device kernels use for every division by a
 * positive quantity (densities, cell widths, squared sound speeds; no
 * counterpart in VAC, whose divisions are plain IEEE divisions): computes
 * q_kernel[i] with that quotient and q_ieee[i] = a[i] / b[i] with the
 * compiler's correctly rounded division, on the current device.  The two
 * agree bit for bit for b > 0 with b and 1/b normal, a zero or of magnitude
 * >= 2^-969, and a/b normal.  Host arrays of n doubles. */
int vacgpu_check_quotients(const double* a, const double* b, long long n, double* q_kernel,
                           double* q_ieee);

/* Self-test of the square root the device kernels use (sound and
 * magnetosonic speeds, Roe-mean sqrt(rho), eigenvector weights; no
 * counterpart in VAC, whose SQRT is the compiler's): computes r_kernel[i]
 * with that root and r_ieee[i] = sqrt(x[i]) with the compiler's correctly
 * rounded square root, on the current device.  The two agree bit for bit
 * for every x.  Host arrays of n doubles. */
int vacgpu_check_roots(const double* x, long long n, double* r_kernel, double* r_ieee);

/* Asynchronous snapshots (no counterpart in VAC, whose savefile,
 * src/vacio.t:1008-1047, writes the host copy of w and so needs the state
 * downloaded first).  vacgpu_snapshot_start copies the current state (ixG,
 * every variable) and its it / t into a device buffer of slot `slot` in
 * stream order, then to pinned host memory on a second stream, and returns
 * at once: later steps may be enqueued and run while the copy to the host is
 * in flight.  vacgpu_snapshot_fetch waits for that copy and writes the state
 * into w in VAC's layout (as vacgpu_get_state does, leading dimensions ld or
 * ixG) with the it and t it was taken at; a host then writes the .out record
 * (vac_amd/vacfile.py) while the device keeps stepping.  Starting a slot
 * again first waits for its previous copy. */
#define VACGPU_SNAP_SLOTS 2
int vacgpu_snapshot_start(vacgpu_grid* g, int slot);
int vacgpu_snapshot_fetch(vacgpu_grid* g, int slot, double* w, const int* ld, int* it, double* t);
/* The same, mesh cells only (ixM, no ghost cells), packed: w holds nw planes
 * of nx[0] * nx[1] (* nx[2]) doubles, x fastest -- the layout of the w
 * records of VAC's files. */
int vacgpu_snapshot_fetch_mesh(vacgpu_grid* g, int slot, double* w, int* it, double* t);

/* Processing of the state: VAC's process() for MHD (src/vacphys.mhd0.t:35-80),
 * which projects B (projectb, src/vacproc.projectb.t:4-157: div B by centred
 * differences, Laplace(phi) = div B solved by the CG scheme of
 * src/vacpoisson.t:391-595 from phi = 0, then B = B - grad(phi) over the mesh
 * and its first ghost layer, with the energy fixes of |procpar(3)| = 2 / 3 and
 * B's ghost cells refilled when procpar(3) > 0).  The Poisson boundary of a
 * face follows the normal B component there: periodic stays periodic, cont
 * gives phi = 0 ghosts ('nul').  The allowed maximum of div B, the stopping
 * rule ('max' or 'rel') and the matvec limit are fixed on the first call from
 * procpar and the div B found then, and kept for later calls, as projectb's
 * saved variables are.  A solve that stops without meeting its tolerance
 * counts nerror(4) (poissonerr_) as VAC's poisson() does.
 *
 * vacgpu_set_process replaces the nproc(1)/procpar/itmin part of readparameters
 * (src/vacio.t:88-97): vacgpu_run then projects where VAC's advect calls
 * process(1) (src/vac.t:392-399) -- before the first sweep of every nproc1-th
 * step counted from itmin, or before every sweep when nproc1 < 0.  procpar may
 * be NULL (all -1: VAC's defaults 0.1, 0, -1).  It resets the saved values.
 *
 * vacgpu_process replaces one process(count, ...) call (src/vac.t:392-399 for
 * count 1; src/vacio.t:910, count ifile+2, before a save).  With count 1 at a
 * step boundary the state before the projection is kept as the next step's
 * wold (VAC stores wold before advect projects).  procpar NULL: the values of
 * vacgpu_set_process.  info / iter / resid (optional) receive what the solve
 * returned (VAC's info codes; info 3 or < 0 leaves B unchanged).  Count 2
 * (processing between a two-step method's halves) is not on the GPU path. */
int vacgpu_set_process(vacgpu_grid* g, int nproc1, const double* procpar, int itmin);
int vacgpu_process(vacgpu_grid* g, int count, const double* procpar, int* info, int* iter,
                   double* resid);

/* getdt_courant (src/vac.t:664-722) on a state the caller holds in device
 * memory, without a grid handle and without allocating anything.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  It runs the
 * kernels of vacgpu_getdt_courant, with their block partials and scalars in
 * a fixed scratch array of the library's device image (VACGPU_NOALLOC_
 * SCRATCH_DOUBLES doubles, 4 MiB per GPU, held until the process ends).
 *   w      device array w(ixG^T, nw) in VAC's layout, x fastest, leading
 *          dimensions ld[d] >= nx[d] + 4; only mesh cells are read;
 *   dxd    device arrays of the cell widths along d over ixG (nx[d] + 4);
 *   dtcourant (ndim, host, may be NULL = none): lagged limits, in; reset to
 *          1e99 on return as VAC does;
 *   smallp (host, in/out): < 0 sets the pressure floor from the state's
 *          pressure minimum times m->smallpcoeff (src/vacphys0.t.mhd:202-214);
 *   dt     (host, out): the Courant time step;
 *   hip_stream: the stream to run on (NULL: the null stream); the call
 *          returns after synchronising it;
 *   err    (may be NULL): receives a message on failure.
 * Grids whose mesh needs more than VACGPU_NOALLOC_SCRATCH_DOUBLES - 32
 * partials ((4 + ndim) per block of 64 x 4 cells: about 22M mesh cells in
 * 2D, 19M in 3D) are rejected with VACGPU_E_ARG.  Every call on a GPU shares
 * that GPU's scratch: calls that overlap on one GPU race. */
#define VACGPU_NOALLOC_SCRATCH_DOUBLES 524288
int vacgpu_getdt_courant_noalloc(const vacgpu_config* cfg, const vacgpu_method* m,
                                 const double* w, const int* ld, const double* dx0,
                                 const double* dx1, const double* dx2, double* dtcourant,
                                 double* smallp, double* dt, void* hip_stream, char* err,
                                 int errlen);

/* Start-up check of a selector set: 0 if supported, else VACGPU_E_ARG. */
int vacgpu_supported(const vacgpu_config* cfg, const vacgpu_method* m);

const char* vacgpu_last_error(const vacgpu_grid* g);
void vacgpu_destroy(vacgpu_grid* g);

/* ---------------------------------------------------------------- batches
 * Many independent 2D TVDLF grids advanced together (no counterpart in the
 * reference, which runs one grid per process: each instance follows VAC's
 * time loop, src/vac.t:55-106, with its own dt, t, it, dtcourant, smallp and
 * stop condition).  An instance in a batch computes exactly the same bits as
 * the same grid advanced alone. */
typedef struct vacgpu_batch vacgpu_batch;

/* Device bytes one instance of cfg occupies (admission control). */
long long vacgpu_batch_footprint(const vacgpu_config* cfg);

/* A batch that admits instances while their footprints fit hbm_budget_bytes,
 * with at most capacity instances at a time. */
int vacgpu_batch_create(long long hbm_budget_bytes, int capacity, vacgpu_batch** out);

/* Admit one 2D TVDLF instance with its initial ixG state w (leading dims ld,
 * NULL = tight), step counter it, time t and stop condition (itmax, tmax,
 * tmaxexact, src/vac.t:73-83).  VACGPU_E_NOMEM (nothing changed) when the
 * budget or the slots are full.  *slot receives the instance id. */
int vacgpu_batch_add(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
                     const double* dx0, const double* dx1, const double* w, const int* ld,
                     int it, double t, int itmax, double tmax, int tmaxexact, int* slot);
/* The same with the initial state given as the mesh only, (nw, ny, nx)
 * tight: the ghost cells are filled on the device, as vacgpu_batch_add
 * refills them (src/vac.t:54). */
int vacgpu_batch_add_mesh(vacgpu_batch* b, const vacgpu_config* cfg, const vacgpu_method* m,
                          const double* dx0, const double* dx1, const double* wmesh, int it,
                          double t, int itmax, double tmax, int tmaxexact, int* slot);
int vacgpu_batch_remove(vacgpu_batch* b, int slot);

/* Admission and removal do not wait for the device: the state is staged in
 * pinned memory (the caller may release w on return) and uploaded, with its
 * first dt, on a copy stream that the next step waits for; a removed
 * instance's memory and slot are reused only after every step and copy
 * enqueued before the removal has completed. */

/* Enqueue nsteps time steps of every instance that has not reached its stop
 * condition (asynchronous). */
int vacgpu_batch_step(vacgpu_batch* b, int nsteps);
/* Status snapshots: vacgpu_batch_snapshot enqueues a copy of every
 * instance's scalars after the steps enqueued so far and returns at once
 * (at most two in flight: a third waits for the oldest);
 * vacgpu_batch_snapshot_wait makes the oldest one in flight current --
 * returns 1 when it did, 0 when none is in flight or (wait = 0) the oldest
 * has not arrived yet.  vacgpu_batch_sync takes a snapshot and waits for it
 * and for every copy. */
int vacgpu_batch_snapshot(vacgpu_batch* b);
int vacgpu_batch_snapshot_wait(vacgpu_batch* b, int wait);
int vacgpu_batch_sync(vacgpu_batch* b);
/* Status in the current snapshot; an instance admitted after it reports its
 * admission it and t, done = 0.  Any output may be NULL. */
int vacgpu_batch_status(vacgpu_batch* b, int slot, int* it, double* t, double* dt, int* done,
                        int* nerror);
/* Instances done in the current snapshot: returns their number and writes up
 * to cap of their slots. */
int vacgpu_batch_finished(vacgpu_batch* b, int* slots, int cap);
/* Result copies: vacgpu_batch_fetch_start enqueues a copy of the instance's
 * state -- ixG, or with mesh = 1 the mesh only, (nw, ny, nx) tight -- after
 * the steps enqueued so far (the instance may be removed right after) and
 * returns a ticket; vacgpu_batch_fetch_ready copies it into w (leading dims
 * ld of that extent, NULL = tight; w may be NULL to discard) and releases
 * the ticket -- returns 1, or 0 when (wait = 0) it has not arrived yet.
 * vacgpu_batch_get_state fetches the ixG state, waiting. */
int vacgpu_batch_fetch_start(vacgpu_batch* b, int slot, int mesh, int* ticket);
int vacgpu_batch_fetch_ready(vacgpu_batch* b, int ticket, double* w, const int* ld, int wait);
int vacgpu_batch_get_state(vacgpu_batch* b, int slot, double* w, const int* ld);

/* Replay each step from a captured hipGraph (re-captured after membership changes). */
int vacgpu_batch_graph(vacgpu_batch* b, int enable);
/* Autotuner: per size class of the current instances (cells < 128^2, < 256^2,
 * < 512^2, larger) time the candidate tile shapes and keep the fastest:
 * 64 x rows tiles (rows in {1,2,4,8}) for predictor + corrector instances,
 * segment lengths {8,16,32,64} for fused-sweep instances.  rows_out[4] (may
 * be NULL) receives the tile rows; vacgpu_batch_get_tiles returns both. */
int vacgpu_batch_tune(vacgpu_batch* b, int* rows_out);
int vacgpu_batch_set_tiles(vacgpu_batch* b, const int* rows);
/* Sweep kernels of the batch: fused = 1 (the default) advances instances
 * whose faces have one boundary type for all variables, that use one
 * limiter for all variables and (MHD) the Powell source, with the fused
 * predictor-corrector sweep (ghost cells written by the sweep); fused = 0
 * uses predictor + corrector tiles for every instance.  Results are bitwise
 * the same.  seglen[4] (may be NULL) sets the fused sweeps' rows per
 * marching segment per size class, each in [4, 256]; vacgpu_batch_tune
 * chooses them from {8, 16, 32, 64}. */
int vacgpu_batch_set_kernels(vacgpu_batch* b, int fused, const int* seglen);
/* Current tile rows and fused segment lengths per size class (either may be NULL). */
int vacgpu_batch_get_tiles(const vacgpu_batch* b, int* rows, int* seglen);
long long vacgpu_batch_bytes_used(const vacgpu_batch* b);
const char* vacgpu_batch_last_error(const vacgpu_batch* b);
void vacgpu_batch_destroy(vacgpu_batch* b);

#ifdef __cplusplus
}
#endif
#endif
