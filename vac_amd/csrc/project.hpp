// Projection of B on the device: VAC's process() for MHD (src/vacphys.mhd0.t:35-80)
// calls projectb (src/vacproc.projectb.t:4-157), which solves the Poisson
// equation  Laplace(phi) = div B  on the mesh with the conjugate-gradient
// solver of src/vacpoisson.t:391-595 (slab grids select 'cg',
// src/vacpoisson.t:45-52), then sets B' = B - grad(phi) over the mesh and its
// first ghost layer.
//
// Every array here is one plane of the state's layout (same strides and
// offset).  Each CG step is four launches: the new search direction with the
// Laplacian and the first dot product, a one-block fold, the vector update
// with the residual norms, a one-block fold with the stopping tests.  (A fold
// by the last block to finish, through one atomic counter, costs more: the
// blocks' increments of one address serialise.)  The CG scalars stay in device memory (Cg);
// once a stopping test fires, every later launch of the solve returns at once
// (Cg::done), so the host can queue several steps before it looks.
//
// Arithmetic is VAC's statement by statement (divisions as pdiv forms them,
// square roots correctly rounded); the dot products are sums in a fixed tree
// order, so a solve is deterministic but its sums are not VAC's left-to-right
// ones.
#pragma once

#include "sweep.hpp"

namespace {

constexpr int CG_STOP_MAX = 0, CG_STOP_REL = 1;

// device-resident CG state (cgscalar's locals, src/vacpoisson.t:455-458)
struct Cg {
  double s2, mx;               // folded sum of squares / max |.| of the last pass
  double rho, res0, res, tol, assumedzero, alf, bet;
  int itr, maxit, info, done, typestop;
};

// boundscalar's type per face (src/vacpoisson.t:158-300)
enum { SB_NUL = 0, SB_PERIODIC = 1, SB_SYMM = 2, SB_ASYMM = 3, SB_GRAD0 = 4 };
struct SBC {
  unsigned char type[6];
};

// a / b with y the refined reciprocal of b that pdiv computes (rdx below):
// the same operations as pdiv after its reciprocal, so the same double
__device__ __forceinline__ double pdiv_y(double a, double b, double y) {
  const double q0 = a * y;
  return __builtin_fma(-__builtin_fma(b, q0, -a), y, q0);
}

// pdiv's reciprocal of every cell width along one dimension
__global__ void k_recip(const double* __restrict__ dx, double* __restrict__ r, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const double b = dx[i];
  const double y0 = __builtin_amdgcn_rcp(b);
  const double y1 = __builtin_fma(y0, __builtin_fma(-b, y0, 1.0), y0);
  r[i] = __builtin_fma(y1, __builtin_fma(-b, y1, 1.0), y1);
}

struct RDX {
  const double* r[3];
};

// the ixG coordinate whose value boundscalar leaves at coordinate c along k:
// c itself on the mesh, the periodic image, the mirror image ('symm', and
// 'asymm' with neg toggled), the edge cell ('grad0', second ghost layer), or
// -1 for a zero ('nul'; 'grad0', first ghost layer).  These are copies, sign
// changes and zeros, so mapping each coordinate on its own gives what VAC's
// face-by-face fill gives (every face region spans the full transverse ixG
// range).
__device__ __forceinline__ int smap(const Geo& g, const SBC& sb, int k, int c, bool& neg) {
  const int n = g.n[k];
  int f;
  if (c < 2) f = 2 * k;
  else if (c > n + 1) f = 2 * k + 1;
  else return c;
  const bool up = f & 1;
  switch (sb.type[f]) {
    case SB_PERIODIC: return up ? c - n : c + n;
    case SB_ASYMM: neg = !neg; return up ? 2 * n + 3 - c : 3 - c;
    case SB_SYMM: return up ? 2 * n + 3 - c : 3 - c;
    case SB_GRAD0:
      if (c == (up ? n + 2 : 1)) return -1;
      return up ? n + 1 : 2;
    default: return -1;
  }
}

template <int ND>
__device__ __forceinline__ double sval(const Geo& g, const SBC& sb, const double* __restrict__ q,
                                       const int* c) {
  long long idx = g.off;
  bool neg = false;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const int m = smap(g, sb, k, c[k], neg);
    if (m < 0) return 0.0;
    idx += (long long)m * g.s[k];
  }
  return neg ? -q[idx] : q[idx];
}

// sum over the block (all threads call; valid in thread 0), fixed order
__device__ __forceinline__ double block_sum(double v, double* sm) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  __syncthreads();
  if ((tid & 63) == 0) sm[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    double r = sm[0];
    for (int i = 1; i < (nthr + 63) / 64; ++i) r = r + sm[i];
    v = r;
  }
  return v;
}

// fold nb block partials of slot SLOT_C0 (sums) and, if want_max, SLOT_C1
// (maxima) with one block of 1024 threads; results valid in thread 0
__device__ __forceinline__ void fold_partials(const double* __restrict__ part, long long pstride,
                                              long long nb, bool want_max, double* sm,
                                              double& s, double& m) {
  double a = 0.0, b = -INFINITY;
  for (long long i = threadIdx.x; i < nb; i += blockDim.x) {
    a = a + part[SLOT_C0 * pstride + i];
    if (want_max) {
      const double u = part[SLOT_C1 * pstride + i];
      b = b < u ? u : b;
    }
  }
  s = block_sum(a, sm);
  if (want_max) m = block_reduce<true>(b, sm);
}

// getdivb (src/vacphys.mhd0.t:106-146, gradient at src/vacgrid.t:815-821):
// rhs = z = div B over the mesh, phi = 0; partial sum of rhs^2 and max |rhs|
template <int ND>
__global__ __launch_bounds__(256) void k_pr_divb(Geo g, RDX rx, const double* __restrict__ w,
                                                double* __restrict__ rhs, double* __restrict__ z,
                                                double* __restrict__ phi, double* __restrict__ part,
                                                long long pstride) {
  using L = Layout<ND, true>;
  __shared__ double sm[16];
  const CellIx cl = cell_of(g);
  double sq = 0.0, mx = -INFINITY;
  if (cl.ok) {
    double dv = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double* b = w + L::b(d) * g.ps;
      const double gq = pdiv_y(0.5 * (b[cl.idx + g.s[d]] - b[cl.idx - g.s[d]]), g.dx[d][cl.c[d]],
                               rx.r[d][cl.c[d]]);
      dv = dv + gq;
    }
    rhs[cl.idx] = dv;
    z[cl.idx] = dv;
    phi[cl.idx] = 0.0;
    sq = dv * dv;
    mx = fabs(dv);
  }
  const double bs = block_sum(sq, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bs;
  const double bm = block_reduce<true>(mx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + block_linear()] = bm;
}

// the folded sum of squares and maximum of k_pr_divb into Cg
__global__ __launch_bounds__(1024) void k_pr_fold(Cg* cg, const double* __restrict__ part,
                                                 long long pstride, long long nb) {
  __shared__ double sm[16];
  double s, m;
  fold_partials(part, pstride, nb, true, sm, s, m);
  if (threadIdx.x == 0) {
    cg->s2 = s;
    cg->mx = m;
  }
}

// the saved stopping rule of projectb (src/vacproc.projectb.t:29-58)
__global__ void k_cg_set(Cg* cg, double tol, int typestop, int maxit) {
  cg->tol = tol;
  cg->typestop = typestop;
  cg->maxit = maxit;
}

// cgscalar's set-up with a zero initial guess (src/vacpoisson.t:466-507)
__global__ void k_cg_init(Cg* cg) {
  double rho = sqrt(cg->s2);
  double res0 = rho;
  if (cg->typestop == CG_STOP_MAX) res0 = cg->mx;
  cg->res0 = res0;
  cg->res = res0;
  cg->assumedzero = 1.0e-16 * res0;
  cg->itr = 0;
  cg->alf = 0.0;
  cg->bet = 0.0;
  if (res0 < 1.0e-99 || (cg->typestop != CG_STOP_REL && res0 <= cg->tol)) {
    cg->info = 3;
    cg->done = 1;
  } else {
    cg->rho = rho * rho;
    cg->info = 0;
    cg->done = 0;
  }
}

// One CG step's first half (src/vacpoisson.t:509-513 and :562): Z = rhs +
// bet*Z (the previous step's new search direction; the first step uses Z as
// set up), AZ = Laplace(Z) as matvec_poisson (src/vacpoisson.t:100-155)
// computes it -- the divergence of the gradient, each a centred difference
// half*(q(i+1) - q(i-1))/dx(i), Z's ghost values as boundscalar sets them --
// and the block's partial sum of Z*AZ.  Z is read from zo and written to zn
// (the new direction is evaluated at the stencil's points).
template <int ND>
__global__ __launch_bounds__(256) void k_cg_matvec(Geo g, SBC sb, RDX rx,
                                                  const double* __restrict__ rhs,
                                                  const double* __restrict__ zo,
                                                  double* __restrict__ zn,
                                                  double* __restrict__ az,
                                                  double* __restrict__ part, long long pstride,
                                                  const Cg* __restrict__ cg, int first) {
  if (cg->done) return;  // uniform: every thread reads the same flag
  __shared__ double sm[16];
  const double bet = cg->bet;
  // the search direction at ixG point c (mapped as boundscalar does)
  auto zdir = [&](const int* c) -> double {
    long long idx = g.off;
    bool neg = false;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      const int m = smap(g, sb, k, c[k], neg);
      if (m < 0) return 0.0;
      idx += (long long)m * g.s[k];
    }
    const double v = first ? zo[idx] : rhs[idx] + bet * zo[idx];
    return neg ? -v : v;
  };
  const CellIx cl = cell_of(g);
  double pz = 0.0;
  if (cl.ok) {
    const double zc = first ? zo[cl.idx] : rhs[cl.idx] + bet * zo[cl.idx];
    st_out(zn + cl.idx, zc);
    double qy = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double* dxd = g.dx[d];
      const double* rd = rx.r[d];
      const int cd = cl.c[d];
      double zmm, zpp;
      if (cd >= 4 && cd <= g.n[d] - 1) {
        const long long s2 = 2 * g.s[d];
        zmm = first ? zo[cl.idx - s2] : rhs[cl.idx - s2] + bet * zo[cl.idx - s2];
        zpp = first ? zo[cl.idx + s2] : rhs[cl.idx + s2] + bet * zo[cl.idx + s2];
      } else {
        int c2[3] = {cl.c[0], cl.c[1], cl.c[2]};
        c2[d] = cd - 2;
        zmm = zdir(c2);
        c2[d] = cd + 2;
        zpp = zdir(c2);
      }
      const double qdm = pdiv_y(0.5 * (zc - zmm), dxd[cd - 1], rd[cd - 1]);
      const double qdp = pdiv_y(0.5 * (zpp - zc), dxd[cd + 1], rd[cd + 1]);
      const double qdd = pdiv_y(0.5 * (qdp - qdm), dxd[cd], rd[cd]);
      qy = qy + qdd;
    }
    st_out(az + cl.idx, qy);
    pz = zc * qy;
  }
  const double bs = block_sum(pz, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bs;
}

// alf (src/vacpoisson.t:512-522)
__global__ __launch_bounds__(1024) void k_cg_alf(Cg* cg, const double* __restrict__ part,
                                                long long pstride, long long nb) {
  if (cg->done) return;
  __shared__ double sm[16];
  double s, m;
  fold_partials(part, pstride, nb, false, sm, s, m);
  if (threadIdx.x == 0) {
    const double az = cg->assumedzero;
    if (fabs(s) <= az * az) {
      cg->info = 1;
      cg->done = 1;
    } else {
      cg->alf = cg->rho / s;
    }
  }
}

// qx = qx + alf*Z, rhs = rhs - alf*AZ (:524-525); partial sum of rhs^2 and max |rhs|
template <int ND>
__global__ __launch_bounds__(256) void k_cg_update(Geo g, double* __restrict__ phi,
                                                  double* __restrict__ rhs,
                                                  const double* __restrict__ z,
                                                  const double* __restrict__ az,
                                                  double* __restrict__ part, long long pstride,
                                                  const Cg* __restrict__ cg) {
  if (cg->done) return;
  __shared__ double sm[16];
  const CellIx cl = cell_of(g);
  const double alf = cg->alf;
  double sq = 0.0, mx = -INFINITY;
  if (cl.ok) {
    st_out(phi + cl.idx, phi[cl.idx] + alf * z[cl.idx]);
    const double r = rhs[cl.idx] - alf * az[cl.idx];
    rhs[cl.idx] = r;
    sq = r * r;
    mx = fabs(r);
  }
  const double bs = block_sum(sq, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C0 * pstride + block_linear()] = bs;
  const double bm = block_reduce<true>(mx, sm);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[SLOT_C1 * pstride + block_linear()] = bm;
}

// the residual norm and the stopping tests (:527-560), then bet
__global__ __launch_bounds__(1024) void k_cg_res(Cg* cg, const double* __restrict__ part,
                                                long long pstride, long long nb) {
  if (cg->done) return;
  __shared__ double sm[16];
  double s, m;
  fold_partials(part, pstride, nb, true, sm, s, m);
  if (threadIdx.x == 0) {
    double rhonew = sqrt(s);
    double res;
    if (cg->typestop == CG_STOP_MAX) res = m;
    else res = rhonew / cg->res0;
    rhonew = rhonew * rhonew;
    cg->itr = cg->itr + 1;
    cg->res = res;
    const double az = cg->assumedzero;
    if (res <= cg->tol) {
      cg->info = 0;
      cg->done = 1;
    } else if (cg->itr >= cg->maxit) {
      cg->info = 2;
      cg->done = 1;
    } else if (cg->rho <= az * az) {
      cg->info = 1;
      cg->done = 1;
    } else {
      cg->bet = rhonew / cg->rho;
      cg->rho = rhonew;
    }
  }
}

// the returned info's sign (:570-573) and nerror(poissonerr_) (:79-81)
__global__ void k_cg_final(Cg* cg, Scal* sc) {
  int info = cg->info;
  const double res = cg->res;
  if ((cg->typestop == CG_STOP_REL && res > 1.0) || (cg->typestop != CG_STOP_REL && res > cg->res0))
    info = -info;
  cg->info = info;
  if (info != 0 && info != 3) sc->nerror[3] += 1;
}

// B' = B - grad(phi) over the mesh widened by one cell (src/vacproc.projectb.t:111-141),
// phi's ghost values as boundscalar sets them; |procpar(divbbound_)| = 2 / 3
// corrects e to keep the thermal / total pressure (:128-136).  Nothing is done
// when the solve returned info 3 or info < 0 (:104-108).
template <int ND>
__global__ __launch_bounds__(256) void k_pr_apply(Geo g, SBC sb, RDX rx, const double* __restrict__ phi,
                                                 double* __restrict__ w,
                                                 const Cg* __restrict__ cg, int ebound,
                                                 double gamma) {
  const int info = cg->info;
  if (info == 3 || info < 0) return;
  using L = Layout<ND, true>;
  int c[3];
  c[0] = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  c[1] = 1 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  c[2] = (ND == 3) ? 1 + (int)blockIdx.z : 0;
  if (c[0] > g.n[0] + 2 || c[1] > g.n[1] + 2) return;
  const long long idx = g.off + c[0] + (long long)c[1] * g.s[1] + (long long)c[2] * g.s[2];
  double* pe = w + L::e * g.ps;
  double e = pe[idx];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    int cc[3] = {c[0], c[1], c[2]};
    cc[d] = c[d] + 1;
    const double pp = sval<ND>(g, sb, phi, cc);
    cc[d] = c[d] - 1;
    const double pm = sval<ND>(g, sb, phi, cc);
    const double gq = pdiv_y(0.5 * (pp - pm), g.dx[d][c[d]], rx.r[d][c[d]]);
    double* pb = w + L::b(d) * g.ps;
    const double b = pb[idx];
    if (ebound == 2) {
      e = e - b * gq + gq * gq / 2;
    } else if (ebound == 3 && gamma != 2.0 && gamma != 1.0) {
      e = e + (gamma - 2.0) / (gamma - 1.0) * (-(b * gq) + gq * gq / 2);
    }
    pb[idx] = b - gq;
  }
  if (ebound == 2 || ebound == 3) pe[idx] = e;
}

}  // namespace
