// pentad_varlen.cpp -- host side of the variable-length batch.
//
// pentadsolver_varlen_create turns the caller's offsets into a device
// schedule once (sorted permutation + work items, see kernels_varlen.hip);
// the solves then launch one persistent kernel each (after a one-int fill
// of the plan's work counter), with no host work and no allocation, so they
// can be captured in a hipGraph.  Solves of one plan share that counter and
// must not overlap (one stream per plan).
//
// Lanes per system follow the length bucket: a system of c = ceil(len /
// VL_ROWS) chunks is in bucket b = ceil(log2 c) and gets 2^(b + spread[b])
// lanes (at most VL_BLOCK).  spread is 0 by default; the autotuner times
// spreads 0, 1 and 2 for every bucket on the caller's data, and keeps the
// fastest per bucket when the whole schedule with them beats the default.
// No reference counterpart (the reference solves batches of equal length).
#include "pentad_api_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace pentad {

constexpr int VL_BUCKETS = 10; // chunk counts 2^0 .. 2^9 (VL_BLOCK = 512 lanes)

struct Varlen {
  Context *c = nullptr;
  int64_t n_sys = 0;
  int n_items = 0;
  int64_t *offs = nullptr; // device, n_sys + 1
  int *perm = nullptr;     // device, n_sys
  int2 *items = nullptr;   // device, room for n_sys items
  int *next = nullptr;     // device, the solves' work counter
  void *mem = nullptr;
  int64_t end = 0;          // offsets[n_sys]: x spans elements [0, end)
  std::vector<int64_t> len; // host, per system
  std::vector<int> order;   // host, longest first
  int spread[VL_BUCKETS] = {};
  bool refine = false; // FP32 + refinement kernel instead of the widened FP64 solve
};

namespace {

int bucket_of(int64_t len) {
  const int64_t chunks = (len + VL_ROWS - 1) / VL_ROWS;
  int b = 0;
  while ((int64_t(1) << b) < chunks) ++b;
  return b;
}

// work items of the longest-first list: consecutive systems of one bucket,
// VL_BLOCK / lanes of them per item; only_bucket >= 0 keeps that bucket.
// Claim order: the longest quarter of the items alternates with the
// shortest items (shortest first), then the rest follow longest first.  In
// a pure longest-first order every workgroup starts on an item of the same
// size, so the CUs load, solve and store in step for the first rounds (the
// 8192-system batch: its 1035 longest items fill four rounds of 256 CUs)
// and HBM idles while they all solve; items of mixed sizes put them out of
// step from the start, and the end of the list still holds items no longer
// than the middle ones.  Measured on the 8192-system batch (lengths
// log-uniform in 32..8192), against the longest-first order: FP64 0.290 ->
// 0.272 ms, FP32 0.200 -> 0.176, FP32 coefficients widened to FP64 0.233 ->
// 0.218, FP32 solve + FP64 refinement 0.368 -> 0.315 (a random order: 0.287
// / 0.187 / 0.225 / 0.331; alternating over the whole list: 0.287 / 0.186 /
// 0.229 / 0.326).
std::vector<int2> claim_order(const std::vector<int2> &lpt) {
  const size_t n = lpt.size(), q = n / 4;
  std::vector<int2> out;
  out.reserve(n);
  for (size_t k = 0; k < q; ++k) {
    out.push_back(lpt[k]);         // k-th longest
    out.push_back(lpt[n - 1 - k]); // k-th shortest
  }
  for (size_t k = q; k < n - q; ++k) out.push_back(lpt[k]);
  return out;
}

std::vector<int2> build_items(const Varlen &V, const int spread[VL_BUCKETS],
                              int only_bucket = -1) {
  std::vector<int2> items;
  const int64_t n = V.n_sys;
  for (int64_t i = 0; i < n;) {
    const int b = bucket_of(V.len[V.order[i]]);
    int64_t j = i;
    while (j < n && bucket_of(V.len[V.order[j]]) == b) ++j; // the bucket's run
    if (only_bucket < 0 || b == only_bucket) {
      const int pshift = std::min(b + spread[b], 9);
      const int64_t S = VL_BLOCK >> pshift;
      for (int64_t k = i; k < j; k += S)
        items.push_back(int2{(int)k, (int)std::min<int64_t>(S, j - k) | (pshift << 16)});
    }
    i = j;
  }
  return claim_order(items);
}

pentadStatus_t upload_items(Varlen *V, const std::vector<int2> &items) {
  if (hipStreamSynchronize(V->c->stream) != hipSuccess ||
      hipMemcpy(V->items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice) !=
          hipSuccess)
    return PENTAD_STATUS_EXECUTION_FAILED;
  V->n_items = (int)items.size();
  return PENTAD_STATUS_SUCCESS;
}

// mode 0 FP64, 1 FP32, 2 FP32 + FP64 refinement
hipError_t launch_mode(Varlen *V, int mode, const void *const c[5], void *x, const int2 *items,
                       int n_items) {
  if (mode == 0)
    return launch_varlen<double>(
        Diags<double>{(const double *)c[0], (const double *)c[1], (const double *)c[2],
                      (const double *)c[3], (const double *)c[4]},
        (double *)x, V->offs, V->perm, items, n_items, V->next, V->c->stream);
  const Diags<float> a{(const float *)c[0], (const float *)c[1], (const float *)c[2],
                       (const float *)c[3], (const float *)c[4]};
  if (mode == 1)
    return launch_varlen<float>(a, (float *)x, V->offs, V->perm, items, n_items, V->next,
                                V->c->stream);
  if (!V->refine)
    return launch_varlen_widened(a, (double *)x, V->offs, V->perm, items, n_items, V->next,
                                 V->c->stream);
  return launch_varlen_ir(a, (double *)x, V->offs, V->perm, items, n_items, V->next, 5, 1e-13,
                          nullptr, V->c->stream);
}

// times spreads 0..2 of every bucket (items of that bucket only, on a copy
// of x restored before each launch), then the whole schedule with the
// fastest spread per bucket against the whole default one, and keeps the
// faster of the two
int autotune(Varlen *V, int mode, const void *const c[5], const void *x, int reps,
             char *report, size_t report_len) {
  if (!V || reps < 1 || !x) return -1;
  for (int k = 0; k < 5; ++k)
    if (!c[k]) return -1;
  const size_t xbytes = (size_t)V->end * (mode == 1 ? 4 : 8);
  void *tmp = nullptr;
  int2 *titems = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string rep;
  int status = 0;
  const hipStream_t st = V->c->stream;
  if (hipMalloc(&tmp, xbytes) != hipSuccess ||
      hipMalloc(&titems, (size_t)V->n_sys * sizeof(int2)) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    status = -2;
  }
  int best[VL_BUCKETS] = {};
  for (int b = 0; b < VL_BUCKETS && status == 0; ++b) {
    float best_ms = -1.f;
    for (int s = 0; s <= 2 && b + s <= 9 && status == 0; ++s) {
      int trial[VL_BUCKETS] = {};
      trial[b] = s;
      const std::vector<int2> items = build_items(*V, trial, b);
      if (items.empty()) break; // no system in this bucket
      if (hipMemcpy(titems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice) !=
          hipSuccess) {
        status = -3;
        break;
      }
      float total = 0.f;
      for (int r = 0; r <= reps && status == 0; ++r) { // r = 0: warm-up
        float ms = 0.f;
        if (hipMemcpyAsync(tmp, x, xbytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipEventRecord(e0, st) != hipSuccess ||
            launch_mode(V, mode, c, tmp, titems, (int)items.size()) != hipSuccess ||
            hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          status = -4;
          break;
        }
        if (r > 0) total += ms;
      }
      if (status) break;
      const float ms = total / reps;
      char line[128];
      std::snprintf(line, sizeof line, "bucket %d (chunks <= %d) spread %d %.6f ms\n", b, 1 << b,
                    s, ms);
      rep += line;
      if (best_ms < 0.f || ms < best_ms) {
        best_ms = ms;
        best[b] = s;
      }
    }
  }
  // the buckets overlap in the one launch: keep the per-bucket choice only if
  // the whole schedule with it beats the whole default schedule
  if (status == 0) {
    const int zero[VL_BUCKETS] = {};
    float whole[2] = {0.f, 0.f};
    for (int k = 0; k < 2 && status == 0; ++k) {
      const std::vector<int2> items = build_items(*V, k == 0 ? zero : best);
      if (hipMemcpy(titems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice) !=
          hipSuccess) {
        status = -3;
        break;
      }
      for (int r = 0; r <= reps && status == 0; ++r) {
        float ms = 0.f;
        if (hipMemcpyAsync(tmp, x, xbytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipEventRecord(e0, st) != hipSuccess ||
            launch_mode(V, mode, c, tmp, titems, (int)items.size()) != hipSuccess ||
            hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          status = -4;
          break;
        }
        if (r > 0) whole[k] += ms / reps;
      }
      char line[128];
      std::snprintf(line, sizeof line, "whole %s %.6f ms\n", k == 0 ? "default" : "tuned",
                    whole[k]);
      rep += line;
    }
    if (status == 0) {
      std::memcpy(V->spread, whole[1] < whole[0] ? best : zero, sizeof best);
      if (upload_items(V, build_items(*V, V->spread)) != PENTAD_STATUS_SUCCESS) status = -5;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (titems) (void)hipFree(titems);
  if (tmp) (void)hipFree(tmp);
  if (report && report_len) {
    std::snprintf(report, report_len, "%s", rep.c_str());
  }
  return status;
}

} // namespace
} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_varlen_create(pentadsolver_handle_t handle,
                                          pentadsolver_varlen_t *plan,
                                          const int64_t *offsets, int64_t n_sys) {
  if (!plan || !offsets || n_sys < 1 || n_sys >= (int64_t(1) << 31))
    return PENTAD_STATUS_INVALID_VALUE;
  const int64_t max_len = (int64_t)VL_BLOCK * VL_ROWS;
  Varlen *V = new Varlen();
  V->c = ctx_of(handle);
  V->n_sys = n_sys;
  V->end = offsets[n_sys];
  V->len.resize(n_sys);
  for (int64_t k = 0; k < n_sys; ++k) {
    V->len[k] = offsets[k + 1] - offsets[k];
    if (V->len[k] < 3 || V->len[k] > max_len) {
      std::fprintf(stderr, "[penta] varlen: system %lld has length %lld (allowed 3..%lld)\n",
                   (long long)k, (long long)V->len[k], (long long)max_len);
      delete V;
      return PENTAD_STATUS_INVALID_VALUE;
    }
  }
  V->order.resize(n_sys);
  std::iota(V->order.begin(), V->order.end(), 0);
  std::stable_sort(V->order.begin(), V->order.end(),
                   [&](int a, int b) { return V->len[a] > V->len[b]; });
  const std::vector<int2> items = build_items(*V, V->spread);
  const size_t b_off = (size_t)(n_sys + 1) * sizeof(int64_t);
  const size_t b_perm = ((size_t)n_sys * sizeof(int) + 15) & ~(size_t)15;
  const size_t b_items = (size_t)n_sys * sizeof(int2); // room for any spread
  if (hipMalloc(&V->mem, b_off + b_perm + b_items + sizeof(int)) != hipSuccess) {
    delete V;
    return PENTAD_STATUS_ALLOC_FAILED;
  }
  V->offs = static_cast<int64_t *>(V->mem);
  V->perm = reinterpret_cast<int *>(static_cast<char *>(V->mem) + b_off);
  V->items = reinterpret_cast<int2 *>(static_cast<char *>(V->mem) + b_off + b_perm);
  V->next = reinterpret_cast<int *>(static_cast<char *>(V->mem) + b_off + b_perm + b_items);
  if (hipMemcpy(V->offs, offsets, b_off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(V->perm, V->order.data(), (size_t)n_sys * sizeof(int),
                hipMemcpyHostToDevice) != hipSuccess ||
      upload_items(V, items) != PENTAD_STATUS_SUCCESS) {
    (void)hipFree(V->mem);
    delete V;
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  *plan = V;
  return PENTAD_STATUS_SUCCESS;
}

void pentadsolver_varlen_destroy(pentadsolver_varlen_t *plan) {
  if (!plan || !*plan) return;
  Varlen *V = static_cast<Varlen *>(*plan);
  (void)hipFree(V->mem);
  delete V;
  *plan = nullptr;
}

int pentadsolver_varlen_items(pentadsolver_varlen_t plan) {
  return plan ? static_cast<Varlen *>(plan)->n_items : -1;
}

pentadStatus_t pentadsolver_varlen_set_spread(pentadsolver_varlen_t plan, const int *spread) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || !spread) return PENTAD_STATUS_INVALID_VALUE;
  for (int b = 0; b < VL_BUCKETS; ++b)
    if (spread[b] < 0 || spread[b] > 9) return PENTAD_STATUS_INVALID_VALUE;
  std::memcpy(V->spread, spread, sizeof V->spread);
  return upload_items(V, build_items(*V, V->spread));
}

pentadStatus_t pentadsolver_varlen_set_refine(pentadsolver_varlen_t plan, int refine) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || refine < 0 || refine > 1) return PENTAD_STATUS_INVALID_VALUE;
  V->refine = refine != 0;
  return PENTAD_STATUS_SUCCESS;
}

int pentadsolver_varlen_get_refine(pentadsolver_varlen_t plan) {
  return plan ? (static_cast<Varlen *>(plan)->refine ? 1 : 0) : -1;
}

pentadStatus_t pentadsolver_varlen_get_spread(pentadsolver_varlen_t plan, int *spread) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || !spread) return PENTAD_STATUS_INVALID_VALUE;
  std::memcpy(spread, V->spread, sizeof V->spread);
  return PENTAD_STATUS_SUCCESS;
}

int pentadsolver_D_varlen_autotune(pentadsolver_varlen_t plan, const double *ds,
                                   const double *dl, const double *d, const double *du,
                                   const double *dw, const double *x, int reps, char *report,
                                   size_t report_len) {
  const void *c[5] = {ds, dl, d, du, dw};
  return autotune(static_cast<Varlen *>(plan), 0, c, x, reps, report, report_len);
}

int pentadsolver_S_varlen_autotune(pentadsolver_varlen_t plan, const float *ds,
                                   const float *dl, const float *d, const float *du,
                                   const float *dw, const float *x, int reps, char *report,
                                   size_t report_len) {
  const void *c[5] = {ds, dl, d, du, dw};
  return autotune(static_cast<Varlen *>(plan), 1, c, x, reps, report, report_len);
}

int pentadsolver_DS_varlen_autotune_ir(pentadsolver_varlen_t plan, const float *ds,
                                       const float *dl, const float *d, const float *du,
                                       const float *dw, const double *x, int reps,
                                       char *report, size_t report_len) {
  const void *c[5] = {ds, dl, d, du, dw};
  return autotune(static_cast<Varlen *>(plan), 2, c, x, reps, report, report_len);
}

static pentadStatus_t finish(Varlen *V, hipError_t e, const char *what,
                             const char *kind = "") {
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] %s launch FAILED: %s -- output is invalid\n", what,
                 hipGetErrorString(e));
    set_status(V->c, PENTAD_STATUS_EXECUTION_FAILED);
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  char cfg[64];
  std::snprintf(cfg, sizeof cfg, "varlen%s rows=%d items=%d", kind, VL_ROWS, V->n_items);
  note_kernel(V->c, 0, "varlen", cfg);
  set_status(V->c, PENTAD_STATUS_SUCCESS);
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_D_gpsv_varlen_batch(pentadsolver_varlen_t plan, const double *ds,
                                                const double *dl, const double *d,
                                                const double *du, const double *dw,
                                                double *x) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || !ds || !dl || !d || !du || !dw || !x) return PENTAD_STATUS_INVALID_VALUE;
  return finish(V,
                launch_varlen<double>(Diags<double>{ds, dl, d, du, dw}, x, V->offs, V->perm,
                                      V->items, V->n_items, V->next, V->c->stream),
                "varlen FP64");
}

pentadStatus_t pentadsolver_S_gpsv_varlen_batch(pentadsolver_varlen_t plan, const float *ds,
                                                const float *dl, const float *d,
                                                const float *du, const float *dw, float *x) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || !ds || !dl || !d || !du || !dw || !x) return PENTAD_STATUS_INVALID_VALUE;
  return finish(V,
                launch_varlen<float>(Diags<float>{ds, dl, d, du, dw}, x, V->offs, V->perm,
                                     V->items, V->n_items, V->next, V->c->stream),
                "varlen FP32");
}

pentadStatus_t pentadsolver_DS_gpsv_varlen_batch_ir(pentadsolver_varlen_t plan,
                                                    const float *ds, const float *dl,
                                                    const float *d, const float *du,
                                                    const float *dw, double *x,
                                                    int max_refine, double tol,
                                                    int *d_refine_steps) {
  Varlen *V = static_cast<Varlen *>(plan);
  if (!V || !ds || !dl || !d || !du || !dw || !x || max_refine < 0)
    return PENTAD_STATUS_INVALID_VALUE;
  const Diags<float> a{ds, dl, d, du, dw};
  if (!V->refine) // no refinement sweep: *d_refine_steps is left as it is
    return finish(V,
                  launch_varlen_widened(a, x, V->offs, V->perm, V->items, V->n_items,
                                        V->next, V->c->stream),
                  "varlen FP32 widened to FP64", " widened");
  return finish(V,
                launch_varlen_ir(a, x, V->offs, V->perm, V->items, V->n_items, V->next,
                                 max_refine, tol, d_refine_steps, V->c->stream),
                "varlen FP32+FP64 refinement", " ir");
}

} // extern "C"
