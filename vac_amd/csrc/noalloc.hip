// noalloc.hip -- getdt_courant (src/vac.t:664-722) on a state the caller
// holds in device memory, with no grid handle and no allocation: the block
// partials of the reductions and the device scalars live in `scratch`, a
// fixed array in the library's device image, instead of the memory
// vacgpu_create allocates.  For callers that cannot allocate (code under
// HIP graph capture, tight serving loops): see vacgpu_getdt_courant_noalloc
// in include/vacgpu.h.
//
// The kernels are those of vacgpu_getdt_courant (sweep.hpp): k_pminmax and
// k_fin_smallp set the pressure floor from the state's pressure minimum
// when none is given (src/vacphys0.t.mhd:202-214), k_cmaxdx and k_fin_getdt
// take the maximum of cmax/dx over the mesh for every direction without a
// lagged dtcourant and fold dt = courantpar / that maximum.  They read mesh
// cells only, through per-dimension strides, so the caller's array is in
// VAC's own layout w(ixG^T, nw) with leading dimensions ld.
//
// One scratch array per GPU, shared by every call on that GPU; each call
// rewrites the scalars and partials it reads, so nothing carries over from
// the last call, but calls that overlap on one GPU (two host threads, or two
// streams) race for it.
#include "../../include/vacgpu.h"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

using namespace vacgpu;

namespace {

// Scal, rounded up to 32 doubles (256 bytes), then the partials: slots
// SLOT_PMIN, SLOT_PMAX and SLOT_DX + d, one double per block each
constexpr long long kScalDoubles = 32;
static_assert(sizeof(Scal) <= kScalDoubles * sizeof(double), "Scal outgrew its scratch slot");

}  // namespace

// the fixed scratch: VACGPU_NOALLOC_SCRATCH_DOUBLES doubles (4 MiB)
static __device__ __align__(256) double scratch[VACGPU_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};  // the array's device address, per device

// the scalars of one call: a fill kernel instead of a copy from the host
__global__ void k_noalloc_scal(Scal* sc, double smallp, double dtc0, double dtc1, double dtc2) {
  if (threadIdx.x != 0) return;
  sc->dt = 0.0;
  sc->t = 0.0;
  sc->smallp = smallp;
  sc->dtcourant[0] = dtc0;
  sc->dtcourant[1] = dtc1;
  sc->dtcourant[2] = dtc2;
  sc->tmax = kBig;
  for (int i = 0; i < 4; ++i) sc->nerror[i] = 0;
  sc->it = 0;
  sc->tmaxexact = 0;
  sc->dtold = 0.0;
  sc->dtchk = 0.0;
  sc->itmin = 0;
  sc->dtcantgrow = 0;
  sc->slowsteps = 0;
}

int say(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) {
    strncpy(err, msg, (size_t)errlen - 1);
    err[errlen - 1] = '\0';
  }
  return VACGPU_E_ARG;
}

int say_hip(char* err, int errlen, const char* what, hipError_t e) {
  if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s: %s", what, hipGetErrorString(e));
  return VACGPU_E_HIP;
}

// the device address of `scratch` on the current device, looked up once
hipError_t scratch_on_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = scratch_dev[dev];
  return hipSuccess;
}

template <int ND, bool MHD>
void launch_getdt(const Geo& geo, const double* w, double gamma, Scal* sc, double* part,
                  long long nb, dim3 grid, const vacgpu_method* m, hipStream_t st) {
  hipLaunchKernelGGL((k_pminmax<ND, MHD>), grid, dim3(BX, BY, 1), 0, st, geo, w, gamma, sc, part,
                     nb);
  hipLaunchKernelGGL(k_fin_smallp, dim3(1), dim3(1024), 0, st, sc, part, nb, nb, m->smallpcoeff,
                     (int)MHD);
  for (int d = 0; d < ND; ++d)
    hipLaunchKernelGGL((k_cmaxdx<ND, MHD>), grid, dim3(BX, BY, 1), 0, st, geo, w, gamma, sc, part,
                       nb, d);
  hipLaunchKernelGGL(k_fin_getdt, dim3(1), dim3(1024), 0, st, sc, part, nb, nb, geo.nd,
                     m->courantpar, -1.0, 0);
}

}  // namespace

extern "C" int vacgpu_getdt_courant_noalloc(const vacgpu_config* cfg, const vacgpu_method* m,
                                            const double* w, const int* ld, const double* dx0,
                                            const double* dx1, const double* dx2,
                                            double* dtcourant, double* smallp, double* dt,
                                            void* hip_stream, char* err, int errlen) {
  if (!cfg || !m || !w || !ld || !smallp || !dt) return say(err, errlen, "null argument");
  const int nd = cfg->ndim;
  if (nd != 2 && nd != 3) return say(err, errlen, "ndim must be 2 or 3");
  if (cfg->physics != VACGPU_MHD && cfg->physics != VACGPU_HD)
    return say(err, errlen, "physics must be mhd or hd");
  if (!(cfg->eqpar[0] > 0.0)) return say(err, errlen, "eqpar(gamma_) must be > 0");
  if (!(m->courantpar > 0.0)) return say(err, errlen, "getdt_courant needs courantpar > 0");
  if (!(m->smallpcoeff >= 0.0)) return say(err, errlen, "smallpcoeff must be >= 0");
  const double* dxs[3] = {dx0, dx1, dx2};
  Geo geo{};
  geo.nd = nd;
  long long stride = 1;
  for (int d = 0; d < 3; ++d) {
    const bool on = d < nd;
    geo.n[d] = on ? cfg->nx[d] : 1;
    geo.ng[d] = on ? cfg->nx[d] + 4 : 1;
    if (on && cfg->nx[d] < 2) return say(err, errlen, "nx must be >= 2 in every dimension");
    if (on && ld[d] < geo.ng[d]) return say(err, errlen, "ld[d] must be >= nx[d] + 4");
    if (on && !dxs[d]) return say(err, errlen, "missing dx array");
    geo.s[d] = on ? stride : 0;
    geo.dx[d] = on ? dxs[d] : nullptr;
    if (on) stride *= ld[d];
  }
  geo.ps = stride;
  geo.off = 0;
  const dim3 grid((geo.n[0] + BX - 1) / BX, (geo.n[1] + BY - 1) / BY, nd == 3 ? geo.n[2] : 1);
  const long long nb = (long long)grid.x * grid.y * grid.z;
  // slots SLOT_PMIN, SLOT_PMAX, then SLOT_DX + d for d < nd
  const long long need = kScalDoubles + (long long)(SLOT_DX + nd) * nb;
  if (need > VACGPU_NOALLOC_SCRATCH_DOUBLES) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "grid too large for the fixed scratch: %lld partials of %d needed, the limit is "
             "%lld mesh cells in blocks of %d",
             need, VACGPU_NOALLOC_SCRATCH_DOUBLES,
             (long long)((VACGPU_NOALLOC_SCRATCH_DOUBLES - kScalDoubles) / (SLOT_DX + nd)) * BX * BY,
             BX * BY);
    return say(err, errlen, msg);
  }
  double* base = nullptr;
  hipError_t e = scratch_on_device(&base);
  if (e != hipSuccess) return say_hip(err, errlen, "hipGetSymbolAddress(scratch)", e);
  Scal* sc = reinterpret_cast<Scal*>(base);
  double* part = base + kScalDoubles;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  double dtc[3] = {kBig, kBig, kBig};
  if (dtcourant)
    for (int d = 0; d < nd; ++d) dtc[d] = dtcourant[d];
  hipLaunchKernelGGL(k_noalloc_scal, dim3(1), dim3(64), 0, st, sc, *smallp < 0.0 ? -1.0 : *smallp,
                     dtc[0], dtc[1], dtc[2]);
  const double gamma = cfg->eqpar[0];
  const bool mhd = cfg->physics == VACGPU_MHD;
  if (nd == 2) {
    if (mhd) launch_getdt<2, true>(geo, w, gamma, sc, part, nb, grid, m, st);
    else launch_getdt<2, false>(geo, w, gamma, sc, part, nb, grid, m, st);
  } else {
    if (mhd) launch_getdt<3, true>(geo, w, gamma, sc, part, nb, grid, m, st);
    else launch_getdt<3, false>(geo, w, gamma, sc, part, nb, grid, m, st);
  }
  if ((e = hipGetLastError()) != hipSuccess) return say_hip(err, errlen, "getdt kernels", e);
  Scal h;
  if ((e = hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return say_hip(err, errlen, "hipMemcpyAsync", e);
  if ((e = hipStreamSynchronize(st)) != hipSuccess)
    return say_hip(err, errlen, "hipStreamSynchronize", e);
  *dt = h.dt;
  *smallp = h.smallp;
  if (dtcourant)
    for (int d = 0; d < nd; ++d) dtcourant[d] = h.dtcourant[d];
  if (err && errlen > 0) err[0] = '\0';
  return VACGPU_OK;
}
