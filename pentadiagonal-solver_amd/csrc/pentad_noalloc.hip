// pentad_noalloc.hip -- the batched FP64 solve for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc.
//
// pentadsolver_D_gpsv_batch_noalloc runs exactly the solve of
// pentadsolver_D_gpsv_batch (same planner, kernels and errors) with its
// scratch -- the one-lane kernel's [3][len][n_sys] intermediates, the
// shared factorisation's factor table, the two-workgroup kernel's ticket
// counters, flags and trade words -- in `scratch` below, a fixed array
// in the library's own device memory, instead of a caller buffer.  Every
// GPU the library runs on holds its own copy of the array until the process
// ends.  A solve whose scratch need exceeds PENTAD_NOALLOC_SCRATCH_BYTES is
// refused before anything is launched.  Every kernel writes the scratch it
// reads before reading it, so nothing in the array has to start at zero.
//
// In host code `scratch` names a host-side stand-in, not a device address:
// its device address is looked up once per device with hipGetSymbolAddress
// (at pentadsolver_create for the current device, else at the first call on
// a device) and only that pointer is handed to kernels.
#include "pentad_api_internal.hpp"

#include <cstdio>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[PENTAD_NOALLOC_SCRATCH_BYTES];

namespace pentad {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

} // namespace

// Device address of `scratch` on device dev, looked up once per device with
// dev current (the previous current device is restored).
hipError_t noalloc_scratch(int dev, void **out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

} // namespace pentad

using namespace pentad;

extern "C" pentadStatus_t pentadsolver_D_gpsv_batch_noalloc(
    pentadsolver_handle_t h, const double *ds, const double *dl, const double *d,
    const double *du, const double *dw, double *x, const int *t_dims, int t_ndim,
    int t_solvedim) {
  Context *c = ctx_of(h);
  // the handle's device (the null-handle context follows the current device)
  int dev = c->device;
  if (h == nullptr && hipGetDevice(&dev) != hipSuccess) dev = -1;
  void *p = nullptr;
  const hipError_t e = noalloc_scratch(dev, &p);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] noalloc: scratch array of device %d unavailable: %s\n", dev,
                 hipGetErrorString(e));
    set_status(c, PENTAD_STATUS_EXECUTION_FAILED);
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  // (a plan whose lines are shared by two workgroups runs on one workgroup
  // per line when its trade does not fit the array)
  const size_t need = noalloc_need(c, ds, dl, d, du, dw, x, t_dims, t_ndim, t_solvedim,
                                   PENTAD_NOALLOC_SCRATCH_BYTES);
  if (need > PENTAD_NOALLOC_SCRATCH_BYTES) {
    std::fprintf(stderr,
                 "[penta] noalloc: the solve along solvedim %d needs %zu bytes of scratch, more "
                 "than the library's fixed %zu-byte array; call pentadsolver_D_gpsv_batch with a "
                 "buffer instead\n",
                 t_solvedim, need, (size_t)PENTAD_NOALLOC_SCRATCH_BYTES);
    set_status(c, PENTAD_STATUS_INVALID_VALUE);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const pentadStatus_t s =
      gpsv_batch_pitched<double>(c, ds, dl, d, du, dw, x, t_dims, nullptr, t_ndim, t_solvedim, p,
                                 PENTAD_NOALLOC_SCRATCH_BYTES);
  set_status(c, s);
  return s;
}
