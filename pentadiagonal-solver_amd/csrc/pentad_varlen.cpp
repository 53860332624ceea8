// pentad_varlen.cpp -- host side of the variable-length batch.
//
// pentadsolver_varlen_create turns the caller's offsets into a device
// schedule once (sorted permutation + work items, see kernels_varlen.hip);
// the solves then launch one persistent kernel each (after a one-int fill
// of the plan's work counter), with no host work and no allocation, so they
// can be captured in a hipGraph.  Solves of one plan share that counter and
// must not overlap (one stream per plan).
// No reference counterpart (the reference solves batches of equal length).
#include "pentad_api_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>

namespace pentad {

struct Varlen {
  Context *c = nullptr;
  int64_t n_sys = 0;
  int n_items = 0;
  int64_t *offs = nullptr; // device, n_sys + 1
  int *perm = nullptr;     // device, n_sys
  int2 *items = nullptr;   // device, n_items
  int *next = nullptr;     // device, the solves' work counter
  void *mem = nullptr;
};

} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_varlen_create(pentadsolver_handle_t handle,
                                          pentadsolver_varlen_t *plan,
                                          const int64_t *offsets, int64_t n_sys) {
  if (!plan || !offsets || n_sys < 1 || n_sys >= (int64_t(1) << 31))
    return PENTAD_STATUS_INVALID_VALUE;
  const int64_t max_len = (int64_t)VL_BLOCK * VL_ROWS;
  std::vector<int64_t> len(n_sys);
  for (int64_t k = 0; k < n_sys; ++k) {
    len[k] = offsets[k + 1] - offsets[k];
    if (len[k] < 3 || len[k] > max_len) {
      std::fprintf(stderr, "[penta] varlen: system %lld has length %lld (allowed 3..%lld)\n",
                   (long long)k, (long long)len[k], (long long)max_len);
      return PENTAD_STATUS_INVALID_VALUE;
    }
  }
  std::vector<int> perm(n_sys);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int a, int b) { return len[a] > len[b]; });
  // items: consecutive systems of the longest-first list sharing a chunk count
  std::vector<int2> items;
  for (int64_t i = 0; i < n_sys;) {
    const int64_t chunks = (len[perm[i]] + VL_ROWS - 1) / VL_ROWS;
    int pshift = 0;
    while ((int64_t(1) << pshift) < chunks) ++pshift;
    const int64_t S = VL_BLOCK >> pshift;
    const int count = (int)std::min<int64_t>(S, n_sys - i);
    items.push_back(int2{(int)i, count | (pshift << 16)});
    i += count;
  }
  Varlen *V = new Varlen();
  V->c = ctx_of(handle);
  V->n_sys = n_sys;
  V->n_items = (int)items.size();
  const size_t b_off = (size_t)(n_sys + 1) * sizeof(int64_t);
  const size_t b_perm = ((size_t)n_sys * sizeof(int) + 15) & ~(size_t)15;
  const size_t b_items = items.size() * sizeof(int2);
  if (hipMalloc(&V->mem, b_off + b_perm + b_items + sizeof(int)) != hipSuccess) {
    delete V;
    return PENTAD_STATUS_ALLOC_FAILED;
  }
  V->offs = static_cast<int64_t *>(V->mem);
  V->perm = reinterpret_cast<int *>(static_cast<char *>(V->mem) + b_off);
  V->items = reinterpret_cast<int2 *>(static_cast<char *>(V->mem) + b_off + b_perm);
  V->next = reinterpret_cast<int *>(static_cast<char *>(V->mem) + b_off + b_perm + b_items);
  if (hipMemcpy(V->offs, offsets, b_off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(V->perm, perm.data(), (size_t)n_sys * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(V->items, items.data(), b_items, hipMemcpyHostToDevice) != hipSuccess) {
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

static pentadStatus_t finish(Varlen *V, hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] %s launch FAILED: %s -- output is invalid\n", what,
                 hipGetErrorString(e));
    set_status(V->c, PENTAD_STATUS_EXECUTION_FAILED);
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  char cfg[64];
  std::snprintf(cfg, sizeof cfg, "varlen rows=%d items=%d", VL_ROWS, V->n_items);
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
  return finish(V,
                launch_varlen_ir(Diags<float>{ds, dl, d, du, dw}, x, V->offs, V->perm,
                                 V->items, V->n_items, V->next, max_refine, tol,
                                 d_refine_steps, V->c->stream),
                "varlen FP32+FP64 refinement");
}

} // extern "C"
