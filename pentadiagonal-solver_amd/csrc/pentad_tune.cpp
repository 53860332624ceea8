// pentad_tune.cpp -- project-written autotuner and bandwidth probe.
//
// The autotuner enumerates every kernel configuration that can solve a
// shape -- rows per lane and lanes per system (x), rows per lane and systems
// per workgroup (y/z), the one-lane-per-system kernel when scratch is
// given -- times each on the handle's stream with HIP events, and stores
// the fastest in the handle's table under (precision, direction, length).
// Later solves of that length use it.  The reference picks its kernel by
// fixed rules and a PENTA_PCR_LANES override
// (src/cuda/singlenode/pentad_cuda.cu:613-697); this replaces those rules.
//
// The probe is a stream kernel with the solve's access pattern (six arrays
// read, one written, optional FMAs per element), the counterpart of the
// reference's apps/roofline_probe_cuda.cu: it measures the attainable HBM
// bandwidth the solve kernels are compared with.
#include "pentad_api_internal.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace pentad {

hipError_t launch_probe(const void *const in[6], void *out, size_t n,
                        int elem_bytes, int fmas, hipStream_t st);
hipError_t launch_probe_ir_mix(int streams, const float *const coef[5], const double *b,
                               double *x, size_t n, hipStream_t st);
hipError_t launch_probe_strided_io(int split, const double *const a[5], double *x, int64_t lo,
                                   int64_t n_sys, hipStream_t st);
hipError_t launch_probe_ir_layout(int layout, int rounds, const float *coef, const double *b,
                                  double *x, size_t n, hipStream_t st);

namespace {

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    if (hipEventCreate(&a) != hipSuccess) a = nullptr;
    if (hipEventCreate(&b) != hipSuccess) b = nullptr;
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

std::vector<Plan> candidates(const Lines &ln, int elem_bytes, bool ir,
                             bool aligned) {
  std::vector<Plan> out;
  static const int rows_plain[] = {4, 6, 8, 10, 12, 16};
  static const int rows_ir[] = {4, 6, 8, 10};
  const int *rows = ir ? rows_ir : rows_plain;
  const int nrows = ir ? 4 : 6;
  for (int k = 0; k < nrows; ++k) {
    const int r = rows[k];
    // the plain strided kernels pad a partial last chunk; the others need
    // whole chunks
    if (ln.len % r && (ir || ln.dir == 0)) continue;
    if (ln.dir == 0) {
      if (!aligned) continue;
      const int64_t lanes = ln.len / r;
      const bool ok = ir ? ir_x_supported(r, ln.len)
                         : ((lanes <= 64 && part_x_supported(elem_bytes, r, (int)lanes)) ||
                            part_x_wide_supported(elem_bytes, r, (int)lanes));
      if (ok) {
        Plan p;
        p.kind = ir ? Plan::IrX : Plan::PartX;
        p.rows = r;
        p.lanes = (int)lanes;
        out.push_back(p);
      }
    } else {
      for (int S = 4; S <= 64; S <<= 1) {
        const bool ok = ir ? ir_yz_supported(r, S, ln.len)
                           : part_yz_supported(elem_bytes, r, S, ln.len);
        if (ok) {
          Plan p;
          p.kind = ir ? Plan::IrYZ : Plan::PartYZ;
          p.rows = r;
          p.sys_per_wg = S;
          out.push_back(p);
        }
      }
    }
  }
  if (ir && ln.dir != 0) { // the FP64 solve of the widened coefficients
    for (int r : {4, 6, 8, 12, 16})
      for (int S = 4; S <= 64; S <<= 1)
        if (part_yz_supported(8, r, S, ln.len)) {
          Plan p;
          p.kind = Plan::IrWideYZ;
          p.rows = r;
          p.sys_per_wg = S;
          out.push_back(p);
        }
  }
  if (!ir && ln.dir == 0 && ln.len > 16 * 64) { // long lines: the workgroup-wide kernel
    for (int r : {8, 12, 16})
      for (int S = 1; S <= 8; S <<= 1)
        if (part_yz_supported(elem_bytes, r, S, ln.len)) {
          Plan p;
          p.kind = Plan::PartYZ;
          p.rows = r;
          p.sys_per_wg = S;
          out.push_back(p);
        }
  }
  if (!ir && ln.dir == 0) { // LDS-staged solve ("transpose"), any length
    for (int r : {8, 16}) {
      int lanes = 1;
      while ((int64_t)lanes * r < ln.len) lanes <<= 1;
      if (tr_x_supported(r, lanes, ln.len)) {
        Plan p;
        p.kind = Plan::TrX;
        p.rows = r;
        p.lanes = lanes;
        out.push_back(p);
      }
    }
  }
  return out;
}

bool all_aligned(const void *const p[6]) {
  for (int i = 0; i < 6; ++i)
    if (reinterpret_cast<uintptr_t>(p[i]) & 15u) return false;
  return true;
}

// Generic driver: `run(plan)` launches one solve; returns best plan.
template <typename Run>
int tune_loop(Context *c, const std::vector<Plan> &cands, int reps, Run run,
              Plan *best, std::string *report) {
  Timer t;
  if (!t.a || !t.b) return -1;
  float best_ms = 1e30f;
  int timed = 0;
  for (const Plan &p : cands) {
    if (run(p) != hipSuccess) { // warm-up; a declined config is skipped
      (void)hipGetLastError();
      continue;
    }
    if (hipEventRecord(t.a, c->stream) != hipSuccess) return -1;
    hipError_t e = hipSuccess;
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = run(p);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    if (hipEventRecord(t.b, c->stream) != hipSuccess) return -1;
    if (hipEventSynchronize(t.b) != hipSuccess) return -1;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) return -1;
    ms /= reps;
    ++timed;
    if (report) {
      char line[160];
      std::snprintf(line, sizeof line, "%s %.4f ms\n", plan_config(p).c_str(), ms);
      *report += line;
    }
    if (ms < best_ms) {
      best_ms = ms;
      *best = p;
    }
  }
  return timed;
}

void copy_report(const std::string &rep, char *out, size_t len) {
  if (!out || len == 0) return;
  const size_t n = rep.size() < len - 1 ? rep.size() : len - 1;
  std::memcpy(out, rep.data(), n);
  out[n] = '\0';
}

template <typename T>
int autotune(Context *c, const T *ds, const T *dl, const T *d, const T *du,
             const T *dw, T *x, const int *dims, int ndim, int solvedim,
             void *buffer, size_t buffer_bytes, int reps, char *report,
             size_t report_len) {
  Lines ln;
  if (decompose(dims, ndim, solvedim, &ln) != 0 || ln.len < 3 || reps < 1)
    return -1;
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  std::vector<Plan> cands = candidates(ln, (int)sizeof(T), false, all_aligned(ptrs));
  if (buffer && buffer_bytes >= plan_scratch_bytes(Plan{}, ln, (int)sizeof(T)))
    cands.push_back(Plan{});
  // keep x: solve on a copy restored afterwards
  const size_t nbytes = (size_t)ln.n_sys * (size_t)ln.len * sizeof(T);
  void *save = nullptr;
  if (hipMalloc(&save, nbytes) != hipSuccess) return -1;
  if (hipMemcpyAsync(save, x, nbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
    (void)hipFree(save);
    return -1;
  }
  const Diags<T> a{ds, dl, d, du, dw};
  Plan best;
  std::string rep;
  const int n = tune_loop(
      c, cands, reps,
      [&](const Plan &p) { return launch_plan<T>(p, a, x, ln, buffer, c->stream); },
      &best, &rep);
  const hipError_t e1 = hipMemcpyAsync(x, save, nbytes, hipMemcpyDeviceToDevice, c->stream);
  const hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(save);
  if (n > 0 && e1 == hipSuccess && e2 == hipSuccess) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune[TuneKey{(int)sizeof(T), ln.dir, ln.len}] = best;
  }
  copy_report(rep, report, report_len);
  return (e1 == hipSuccess && e2 == hipSuccess) ? n : -1;
}

int autotune_ir(Context *c, const float *ds, const float *dl, const float *d,
                const float *du, const float *dw, double *x, const int *dims,
                int ndim, int solvedim, int max_refine, double tol, void *buffer,
                size_t buffer_bytes, int reps, char *report, size_t report_len) {
  Lines ln;
  if (decompose(dims, ndim, solvedim, &ln) != 0 || ln.len < 3 || reps < 1)
    return -1;
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  std::vector<Plan> cands = candidates(ln, 4, true, all_aligned(ptrs));
  Plan lane;
  lane.kind = Plan::IrLane;
  if (buffer && buffer_bytes >= ir_lane_scratch_bytes(ln)) cands.push_back(lane);
  const size_t nbytes = (size_t)ln.n_sys * (size_t)ln.len * sizeof(double);
  void *save = nullptr;
  if (hipMalloc(&save, nbytes) != hipSuccess) return -1;
  if (hipMemcpyAsync(save, x, nbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
    (void)hipFree(save);
    return -1;
  }
  const Diags<float> a{ds, dl, d, du, dw};
  Plan best;
  std::string rep;
  // every timed solve starts from the original right-hand side, so the
  // refinement runs its real number of steps
  const int n = tune_loop(
      c, cands, reps,
      [&](const Plan &p) {
        hipError_t e = hipMemcpyAsync(x, save, nbytes, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) return e;
        return launch_plan_ir(p, a, x, ln, buffer, max_refine, tol, nullptr, c->stream);
      },
      &best, &rep);
  const hipError_t e1 = hipMemcpyAsync(x, save, nbytes, hipMemcpyDeviceToDevice, c->stream);
  const hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(save);
  if (n > 0 && e1 == hipSuccess && e2 == hipSuccess) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune[TuneKey{TUNE_IR, ln.dir, ln.len}] = best;
  }
  copy_report(rep, report, report_len);
  return (e1 == hipSuccess && e2 == hipSuccess) ? n : -1;
}

} // namespace

// First solve of a key under PENTA_AUTOTUNE=1: tune that (precision,
// direction, length) on the caller's arrays unless the handle already holds
// an entry for it (x is left as it was).  Returns true when it tuned.  A key
// with no candidate (unaligned arrays, for example) gets the one-lane entry,
// which the planner ignores, so that it is not tried again.
template <typename T>
bool autotune_on_first_use(Context *c, const T *ds, const T *dl, const T *d, const T *du,
                           const T *dw, T *x, const int *dims, int ndim, int solvedim) {
  Lines ln;
  if (decompose(dims, ndim, solvedim, &ln) != 0) return false;
  const TuneKey key{(int)sizeof(T), ln.dir, ln.len};
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->tune.count(key)) return false;
  }
  const int n =
      autotune<T>(c, ds, dl, d, du, dw, x, dims, ndim, solvedim, nullptr, 0, 3, nullptr, 0);
  if (n <= 0) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune.emplace(key, Plan{});
  }
  return n > 0;
}
template bool autotune_on_first_use<float>(Context *, const float *, const float *,
                                           const float *, const float *, const float *,
                                           float *, const int *, int, int);
template bool autotune_on_first_use<double>(Context *, const double *, const double *,
                                            const double *, const double *, const double *,
                                            double *, const int *, int, int);

} // namespace pentad

using namespace pentad;

extern "C" {

int pentadsolver_D_autotune(pentadsolver_handle_t h, const double *ds,
                            const double *dl, const double *d, const double *du,
                            const double *dw, double *x, const int *t_dims,
                            int t_ndim, int t_solvedim, void *t_buffer,
                            size_t buffer_bytes, int reps, char *report,
                            size_t report_len) {
  return autotune<double>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim,
                          t_solvedim, t_buffer, buffer_bytes, reps, report,
                          report_len);
}

int pentadsolver_S_autotune(pentadsolver_handle_t h, const float *ds,
                            const float *dl, const float *d, const float *du,
                            const float *dw, float *x, const int *t_dims,
                            int t_ndim, int t_solvedim, void *t_buffer,
                            size_t buffer_bytes, int reps, char *report,
                            size_t report_len) {
  return autotune<float>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim,
                         t_solvedim, t_buffer, buffer_bytes, reps, report,
                         report_len);
}

int pentadsolver_DS_autotune_ir(pentadsolver_handle_t h, const float *ds,
                                const float *dl, const float *d,
                                const float *du, const float *dw, double *x,
                                const int *t_dims, int t_ndim, int t_solvedim,
                                int max_refine, double tol, void *t_buffer,
                                size_t buffer_bytes, int reps, char *report,
                                size_t report_len) {
  return autotune_ir(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim,
                     t_solvedim, max_refine, tol, t_buffer, buffer_bytes, reps,
                     report, report_len);
}

void pentadsolver_clear_tuning(pentadsolver_handle_t h) {
  Context *c = ctx_of(h);
  std::lock_guard<std::mutex> lk(c->mu);
  c->tune.clear();
}

pentadStatus_t pentadsolver_probe_stream(pentadsolver_handle_t h,
                                         const void *const in[6], void *out,
                                         size_t n, int elem_bytes, int fmas) {
  if (!in || !out || (elem_bytes != 4 && elem_bytes != 8) || fmas < 0)
    return PENTAD_STATUS_INVALID_VALUE;
  for (int i = 0; i < 6; ++i)
    if (!in[i]) return PENTAD_STATUS_INVALID_VALUE;
  const hipError_t e = launch_probe(in, out, n, elem_bytes, fmas, ctx_of(h)->stream);
  return e == hipSuccess ? PENTAD_STATUS_SUCCESS : PENTAD_STATUS_EXECUTION_FAILED;
}

pentadStatus_t pentadsolver_probe_ir_mix(pentadsolver_handle_t h, int streams,
                                         const float *const coef[5], const double *b,
                                         double *x, size_t n) {
  if (!coef || !b || !x || n == 0 || n % 4 || (streams != 7 && streams != 3 && streams != 2))
    return PENTAD_STATUS_INVALID_VALUE;
  const int need = streams == 7 ? 5 : 1;
  for (int i = 0; i < need; ++i)
    if (!coef[i]) return PENTAD_STATUS_INVALID_VALUE;
  const hipError_t e = launch_probe_ir_mix(streams, coef, b, x, n, ctx_of(h)->stream);
  return e == hipSuccess ? PENTAD_STATUS_SUCCESS : PENTAD_STATUS_EXECUTION_FAILED;
}

pentadStatus_t pentadsolver_probe_ir_layout(pentadsolver_handle_t h, int layout, int rounds,
                                            const float *coef, const double *b, double *x,
                                            size_t n) {
  if (!coef || !b || !x || n == 0 || rounds < 0 || (layout != 0 && layout != 1))
    return PENTAD_STATUS_INVALID_VALUE;
  if (layout == 0 ? (n % 2048 != 0) : (n != (size_t)512 * 512 * 512))
    return PENTAD_STATUS_INVALID_VALUE;
  const hipError_t e = launch_probe_ir_layout(layout, rounds, coef, b, x, n, ctx_of(h)->stream);
  return e == hipSuccess ? PENTAD_STATUS_SUCCESS : PENTAD_STATUS_EXECUTION_FAILED;
}

pentadStatus_t pentadsolver_probe_strided_io(pentadsolver_handle_t h, int split,
                                             const double *const coef[5], double *x,
                                             const int *t_dims, int t_solvedim) {
  if (!coef || !x || !t_dims || (t_solvedim != 1 && t_solvedim != 2) ||
      (split != 1 && split != 2 && split != 4))
    return PENTAD_STATUS_INVALID_VALUE;
  for (int i = 0; i < 5; ++i)
    if (!coef[i]) return PENTAD_STATUS_INVALID_VALUE;
  for (int i = 0; i < 3; ++i)
    if (t_dims[i] < 1) return PENTAD_STATUS_INVALID_VALUE;
  if (t_dims[t_solvedim] != 512) return PENTAD_STATUS_INVALID_VALUE;
  const int64_t lo = t_solvedim == 1 ? (int64_t)t_dims[0] : (int64_t)t_dims[0] * t_dims[1];
  const int64_t n_sys = (int64_t)t_dims[0] * t_dims[1] * t_dims[2] / 512;
  if (n_sys % 64 != 0) return PENTAD_STATUS_INVALID_VALUE;
  const hipError_t e = launch_probe_strided_io(split, coef, x, lo, n_sys, ctx_of(h)->stream);
  return e == hipSuccess ? PENTAD_STATUS_SUCCESS : PENTAD_STATUS_EXECUTION_FAILED;
}

} // extern "C"
