// pentad_adi.cpp -- ADI sweep driver: x, then y, then z on one field that
// stays resident in HBM, with the steps captured as a hipGraph.
//
// Counterpart of the reference's apps/adi_cuda.cu (run_one_iteration
// :167-176 calls pentadsolver_gpsv_batch for solvedim 0, 1, 2 in place on
// the same field).  Here the driver owns the data layout.  Every direction
// is solved as a batch of contiguous lines with the register-resident
// partitioned kernel, and writes its solution in the layout the next
// direction reads contiguously.  A layout is an order of the axes, fastest
// first; for 3-D the field cycles through
//
//     x reads (i,k,j) -> writes (j,i,k) -> y reads it, writes (k,j,i)
//     -> z reads it, writes (i,k,j) -> next step ...
//
// (the caller's layout (i,j,k) is read by the first x-solve and written by
// the last z-solve).  A workgroup owns consecutive systems along the next
// direction's axis, so its solutions go out as contiguous runs whose rows
// sit one line length apart -- no strided pass over the field and no
// separate transposes.  Coefficients are copied once, per direction, into
// the layout that direction reads.  Lengths the partitioned kernels do not
// cover fall back to the in-place solves of the drop-in API.
//
// Modes: 0 = FP64, 1 = FP32, 2 = FP32 coefficients with an FP64 field and
// on-chip iterative refinement (pentadsolver_DS_gpsv_batch_ir semantics),
// 3 / 4 = shared factorisation in FP64 / FP32: each direction's system 0 is
// factored once into a table and the steps stream only the field;
// 5 / 6 = structured coefficients in FP64 / FP32 (pentadsolver_adi_create_
// stencil): every line's diagonals are id * I + s * w generated inside the
// kernels from one scale per line, so the steps stream only the field.
#include "pentad_api_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

namespace pentad {

hipError_t launch_permute3(const void *in, int in_bytes, void *out, int out_bytes,
                           const int64_t dims[3], const int64_t ostride[3],
                           hipStream_t st);

pentadStatus_t gpsv_batch_ir(Context *c, const float *ds, const float *dl,
                             const float *d, const float *du, const float *dw,
                             double *x, const int *dims, int ndim, int solvedim,
                             int max_steps, double tol, int *steps, void *buffer);
template <typename T>
pentadStatus_t gpsv_batch(Context *c, const T *ds, const T *dl, const T *d,
                          const T *du, const T *dw, T *x, const int *dims,
                          int ndim, int solvedim, void *buffer);

namespace {

struct Order { // axes, fastest first
  int ax[3] = {0, 1, 2};
};

// element strides of each axis for a layout
void strides_of(const Order &o, const int64_t dims[3], int ndim, int64_t st[3]) {
  int64_t s = 1;
  for (int q = 0; q < 3; ++q) st[q] = 0;
  for (int q = 0; q < ndim; ++q) {
    st[o.ax[q]] = s;
    s *= dims[o.ax[q]];
  }
}

} // namespace

struct Adi {
  Context *c = nullptr;
  int mode = 0;
  int ndim = 0;
  int64_t dims[3] = {1, 1, 1};
  int64_t n = 0;
  int cbytes = 8; // coefficient element size used by the kernels
  int fbytes = 8; // field element size
  bool rotating = true;
  void *coef[3][5] = {};
  // mode 2: direction d's five FP32 arrays interleaved per il[d] rows (4 or
  // 2; 0: five arrays) at coef[d][0] (one coefficient stream instead of five)
  int il[3] = {0, 0, 0};
  void *coef_mem = nullptr;
  void *tab[3] = {}; // shared-fact factor tables (modes 3, 4), in coef_mem
  void *scale[3] = {}; // structured coefficients (modes 5, 6, 7), in coef_mem
  double sw[3][5] = {};
  double sid[3] = {1.0, 1.0, 1.0};
  Order corder[3]; // coefficient layout per direction
  void *field[2] = {nullptr, nullptr};
  void *field_mem = nullptr;
  void *scratch = nullptr; // in-place fallback only
  size_t scratch_bytes = 0;
  Plan plan[3];
  int max_refine = 5;
  double tol = 1e-13;
  int *d_steps = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string cfg[3];
};

namespace {

// Axis order of the field read by direction d's solve.  first_step: the x
// solve of the first step reads the caller's layout.
Order in_order(const Adi *A, int d, bool first_step) {
  Order o;
  if (A->ndim == 2) {
    o.ax[0] = d;
    o.ax[1] = 1 - d;
    return o;
  }
  if (d == 0) {
    o.ax[0] = 0;
    o.ax[1] = first_step ? 1 : 2;
    o.ax[2] = first_step ? 2 : 1;
  } else if (d == 1) {
    o.ax[0] = 1; o.ax[1] = 0; o.ax[2] = 2;
  } else {
    o.ax[0] = 2; o.ax[1] = 1; o.ax[2] = 0;
  }
  return o;
}

// Direction d solves lines along axis d; its workgroups run along the next
// direction's axis X; Y is the remaining axis.  last_step: the last
// direction of the last step writes the caller's layout.
XMap make_map(const Adi *A, int d, bool first_step, bool last_step) {
  const int ndim = A->ndim;
  const int X = (d + 1) % ndim;
  const int Y = (ndim == 3) ? (d + 2) % 3 : -1;
  Order O;
  if (last_step && d == ndim - 1) {
    O = Order(); // caller's layout
  } else if (ndim == 2) {
    O.ax[0] = X;
    O.ax[1] = d;
  } else {
    O.ax[0] = X;
    O.ax[1] = d;
    O.ax[2] = Y;
  }
  const Order I = in_order(A, d, first_step);
  int64_t si[3], so[3], sc[3];
  strides_of(I, A->dims, ndim, si);
  strides_of(O, A->dims, ndim, so);
  strides_of(A->corder[d], A->dims, ndim, sc);
  XMap m;
  m.S1 = A->dims[X];
  m.in_lo = si[X];
  m.c_lo = sc[X];
  m.out_lo = so[X];
  m.out_row = so[d];
  if (Y >= 0) {
    m.in_hi = si[Y];
    m.c_hi = sc[Y];
    m.out_hi = so[Y];
  }
  return m;
}

hipError_t solve_dir(Adi *A, int d, const void *in, void *out, bool first_step,
                     bool last_step, hipStream_t st) {
  const Plan &pl = A->plan[d];
  const int64_t len = A->dims[d], nsys = A->n / len;
  const XMap map = make_map(A, d, first_step, last_step);
  if (A->mode >= 5) {
    // scale of system (lo, hi): the caller lays out direction d's scales over
    // the remaining axes in the standard order (fastest first)
    const int ndim = A->ndim, X = (d + 1) % ndim, Y = (ndim == 3) ? (d + 2) % 3 : -1;
    int64_t sa = 1, sb = 0;
    if (Y >= 0) {
      if (X < Y) sb = A->dims[X];
      else { sa = A->dims[Y]; sb = 1; }
    }
    if (A->mode == 7) { // FP32 solve + FP64 refinement against the FP64 operator
      Stencil<double> stc;
      stc.scale = (const double *)A->scale[d];
      stc.sa = sa;
      stc.sb = sb;
      for (int k = 0; k < 5; ++k) stc.w[k] = A->sw[d][k];
      stc.id = A->sid[d];
      return launch_stencil_ir_x_map(stc, (const double *)in, (double *)out, nsys, len, map,
                                     pl.rows, A->max_refine, A->tol, A->d_steps, st);
    }
    if (A->mode == 5) {
      Stencil<double> stc;
      stc.scale = (const double *)A->scale[d];
      stc.sa = sa;
      stc.sb = sb;
      for (int k = 0; k < 5; ++k) stc.w[k] = A->sw[d][k];
      stc.id = A->sid[d];
      return launch_stencil_x_map<double>(stc, (const double *)in, (double *)out, true, nsys,
                                          len, map, pl.rows, pl.lanes, st);
    }
    Stencil<float> stc;
    stc.scale = (const float *)A->scale[d];
    stc.sa = sa;
    stc.sb = sb;
    for (int k = 0; k < 5; ++k) stc.w[k] = (float)A->sw[d][k];
    stc.id = (float)A->sid[d];
    return launch_stencil_x_map<float>(stc, (const float *)in, (float *)out, true, nsys, len,
                                       map, pl.rows, pl.lanes, st);
  }
  if (A->mode == 3)
    return launch_sf_x_map<double>((const double *)A->tab[d], (const double *)in,
                                   (double *)out, true, nsys, len, map, pl.rows, st);
  if (A->mode == 4)
    return launch_sf_x_map<float>((const float *)A->tab[d], (const float *)in, (float *)out,
                                  true, nsys, len, map, pl.rows, st);
  if (A->mode == 2 && pl.kind == Plan::IrWideX)
    return launch_part_x_map_widened(
        Diags<float>{(const float *)A->coef[d][0], (const float *)A->coef[d][1],
                     (const float *)A->coef[d][2], (const float *)A->coef[d][3],
                     (const float *)A->coef[d][4]},
        (const double *)in, (double *)out, true, nsys, len, map, pl.rows, pl.lanes, st);
  if (A->mode == 2 && A->il[d])
    return launch_ir_x_map_il((const float *)A->coef[d][0], (const double *)in, (double *)out,
                              nsys, len, map, pl.rows, A->max_refine, A->tol, A->d_steps, st);
  if (A->mode == 2) {
    const Diags<float> a{(const float *)A->coef[d][0], (const float *)A->coef[d][1],
                         (const float *)A->coef[d][2], (const float *)A->coef[d][3],
                         (const float *)A->coef[d][4]};
    return launch_ir_x_map(a, (const double *)in, (double *)out, true, nsys, len, map,
                           pl.rows, A->max_refine, A->tol, A->d_steps, st);
  }
  if (A->mode == 0) {
    const Diags<double> a{(const double *)A->coef[d][0], (const double *)A->coef[d][1],
                          (const double *)A->coef[d][2], (const double *)A->coef[d][3],
                          (const double *)A->coef[d][4]};
    return launch_part_x_map<double>(a, (const double *)in, (double *)out, true, nsys,
                                     len, map, pl.rows, pl.lanes, st);
  }
  const Diags<float> a{(const float *)A->coef[d][0], (const float *)A->coef[d][1],
                       (const float *)A->coef[d][2], (const float *)A->coef[d][3],
                       (const float *)A->coef[d][4]};
  return launch_part_x_map<float>(a, (const float *)in, (float *)out, true, nsys, len,
                                  map, pl.rows, pl.lanes, st);
}

// In-place solve along direction d with the drop-in dispatcher (fallback
// when a length has no partitioned kernel; coefficients in the caller's
// layout), on stream st.
hipError_t inplace_dir(Adi *A, int d, void *x, hipStream_t st) {
  int dims[3] = {(int)A->dims[0], (int)A->dims[1], (int)A->dims[2]};
  hipStream_t keep = A->c->stream;
  A->c->stream = st;
  void *const *cf = A->coef[d];
  pentadStatus_t s;
  if (A->mode == 2)
    s = gpsv_batch_ir(A->c, (const float *)cf[0], (const float *)cf[1],
                      (const float *)cf[2], (const float *)cf[3], (const float *)cf[4],
                      (double *)x, dims, A->ndim, d, A->max_refine, A->tol, A->d_steps,
                      A->scratch);
  else if (A->mode == 0)
    s = gpsv_batch<double>(A->c, (const double *)cf[0], (const double *)cf[1],
                           (const double *)cf[2], (const double *)cf[3],
                           (const double *)cf[4], (double *)x, dims, A->ndim, d,
                           A->scratch);
  else
    s = gpsv_batch<float>(A->c, (const float *)cf[0], (const float *)cf[1],
                          (const float *)cf[2], (const float *)cf[3], (const float *)cf[4],
                          (float *)x, dims, A->ndim, d, A->scratch);
  A->c->stream = keep;
  return s == PENTAD_STATUS_SUCCESS ? hipSuccess : hipErrorLaunchFailure;
}

// One direction of step `s` of `nsteps`; `in` is the field it reads,
// returns (through *next) the field the next direction reads.
hipError_t one_dir(Adi *A, int d, int s, int nsteps, void *x, const void *in,
                   const void **next, hipStream_t st) {
  if (!A->rotating) {
    *next = x;
    return inplace_dir(A, d, x, st);
  }
  // the last direction writes back into x (its input is a scratch field)
  void *out = (d == A->ndim - 1) ? x : A->field[d % 2];
  *next = out;
  return solve_dir(A, d, in, out, s == 0, s == nsteps - 1, st);
}

// nsteps ADI steps on x (caller's layout in and out).
hipError_t run_steps(Adi *A, void *x, int nsteps, hipStream_t st) {
  for (int s = 0; s < nsteps; ++s) {
    const void *in = x;
    for (int d = 0; d < A->ndim; ++d) {
      const hipError_t e = one_dir(A, d, s, nsteps, x, in, &in, st);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

void free_graph(Adi *A) {
  if (A->exec) (void)hipGraphExecDestroy(A->exec);
  if (A->graph) (void)hipGraphDestroy(A->graph);
  A->exec = nullptr;
  A->graph = nullptr;
}

void free_adi(Adi *A) {
  free_graph(A);
  if (A->cap_stream) (void)hipStreamDestroy(A->cap_stream);
  if (A->coef_mem) (void)hipFree(A->coef_mem);
  if (A->field_mem) (void)hipFree(A->field_mem);
  if (A->scratch) (void)hipFree(A->scratch);
  if (A->d_steps) (void)hipFree(A->d_steps);
  delete A;
}

} // namespace
} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_adi_create(pentadsolver_handle_t h,
                                       pentadsolver_adi_t *out, int mode,
                                       int ndim, const int *dims,
                                       const void *const *coeffs,
                                       int coeff_sets, int coeff_elem_bytes,
                                       int max_refine, double tol) {
  const bool shared = (mode == 3 || mode == 4);
  if (shared && coeff_elem_bytes != (mode == 3 ? 8 : 4)) return PENTAD_STATUS_INVALID_VALUE;
  if (!out || !dims || !coeffs || mode < 0 || mode > 4 || ndim < 2 ||
      ndim > 3 || (coeff_sets != 1 && coeff_sets != ndim) ||
      (coeff_elem_bytes != 4 && coeff_elem_bytes != 8) || max_refine < 0)
    return PENTAD_STATUS_INVALID_VALUE;
  for (int i = 0; i < ndim; ++i)
    if (dims[i] < 3) return PENTAD_STATUS_INVALID_VALUE;
  Context *c = ctx_of(h);
  Adi *A = new Adi();
  A->c = c;
  A->mode = mode;
  A->ndim = ndim;
  A->max_refine = max_refine;
  A->tol = tol;
  A->n = 1;
  for (int i = 0; i < ndim; ++i) {
    A->dims[i] = dims[i];
    A->n *= dims[i];
  }
  A->cbytes = (mode == 0 || mode == 3) ? 8 : 4;
  A->fbytes = (mode == 1 || mode == 4) ? 4 : 8;
  hipStream_t st = c->stream;
  auto fail = [&](pentadStatus_t s) {
    free_adi(A);
    return s;
  };

  // kernel per direction: contiguous lines of length dims[d]; driver-owned
  // buffers are 256-byte aligned (the caller's x is only read/written at
  // mapped positions by these kernels)
  A->rotating = true;
  for (int d = 0; d < ndim; ++d) {
    Lines ln;
    ln.dir = 0;
    ln.len = A->dims[d];
    ln.n_sys = A->n / ln.len;
    const void *aligned[6] = {};
    if (shared) {
      static const int sf64[] = {8, 16, 6, 4}, sf32[] = {16, 8, 6, 4};
      for (int k = 0; k < 4 && A->plan[d].kind != Plan::SfX; ++k) {
        const int r = (mode == 3 ? sf64 : sf32)[k];
        if (sf_x_supported(A->cbytes, r, ln.len)) {
          A->plan[d].kind = Plan::SfX;
          A->plan[d].rows = r;
          A->plan[d].lanes = (int)(ln.len / r);
        }
      }
      if (A->plan[d].kind != Plan::SfX) {
        std::fprintf(stderr, "[penta] adi: no shared-fact kernel for length %lld\n",
                     (long long)ln.len);
        return fail(PENTAD_STATUS_INVALID_VALUE);
      }
      continue;
    }
    A->plan[d] = (mode == 2) ? plan_ir(c, ln, aligned) : plan_solve(c, ln, A->cbytes, aligned);
    // refinement mode: lengths no refinement kernel covers run the FP64
    // solve of the widened coefficients with the same rotated stores
    if (A->plan[d].kind != Plan::PartX && A->plan[d].kind != Plan::IrX &&
        A->plan[d].kind != Plan::IrWideX)
      A->rotating = false;
  }
  // the caller's x must be 16-byte aligned for the vector loads of the first
  // x-solve; that is checked when stepping
  for (int d = 0; d < ndim; ++d) A->corder[d] = A->rotating ? in_order(A, d, false) : Order();
  const size_t csz = (size_t)A->n * A->cbytes;
  const size_t fsz = (size_t)A->n * A->fbytes;
  if (shared) {
    size_t off[3] = {0, 0, 0}, total = 0;
    for (int d = 0; d < ndim; ++d) {
      off[d] = total;
      total += (sf_table_bytes(A->cbytes, A->plan[d].rows, A->dims[d]) + 255) & ~(size_t)255;
    }
    if (hipMalloc(&A->coef_mem, total) != hipSuccess ||
        hipMalloc(&A->field_mem, fsz * 2) != hipSuccess ||
        hipMalloc((void **)&A->d_steps, sizeof(int)) != hipSuccess)
      return fail(PENTAD_STATUS_ALLOC_FAILED);
    A->field[0] = A->field_mem;
    A->field[1] = (char *)A->field_mem + fsz;
    if (launch_fill_int(A->d_steps, 0, st) != hipSuccess)
      return fail(PENTAD_STATUS_EXECUTION_FAILED);
    int64_t stride = 1;
    for (int d = 0; d < ndim; ++d) {
      A->tab[d] = (char *)A->coef_mem + off[d];
      const void *const *cs = coeffs + (coeff_sets == 1 ? 0 : d) * 5;
      for (int k = 0; k < 5; ++k)
        if (!cs[k]) return fail(PENTAD_STATUS_INVALID_VALUE);
      Lines ln; // system 0 along axis d of the caller's layout
      ln.dir = 0;
      ln.len = A->dims[d];
      ln.n_sys = A->n / ln.len;
      ln.stride = stride;
      stride *= A->dims[d];
      const hipError_t e =
          mode == 3
              ? launch_sf_factor<double>(Diags<double>{(const double *)cs[0], (const double *)cs[1],
                                                       (const double *)cs[2], (const double *)cs[3],
                                                       (const double *)cs[4]},
                                         ln, A->plan[d].rows, (double *)A->tab[d], st)
              : launch_sf_factor<float>(Diags<float>{(const float *)cs[0], (const float *)cs[1],
                                                     (const float *)cs[2], (const float *)cs[3],
                                                     (const float *)cs[4]},
                                        ln, A->plan[d].rows, (float *)A->tab[d], st);
      if (e != hipSuccess) return fail(PENTAD_STATUS_EXECUTION_FAILED);
      A->cfg[d] = plan_config(A->plan[d]);
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(PENTAD_STATUS_EXECUTION_FAILED);
    *out = A;
    return PENTAD_STATUS_SUCCESS;
  }
  if (hipMalloc(&A->coef_mem, csz * 5 * ndim) != hipSuccess ||
      hipMalloc((void **)&A->d_steps, sizeof(int)) != hipSuccess)
    return fail(PENTAD_STATUS_ALLOC_FAILED);
  if (A->rotating) {
    if (hipMalloc(&A->field_mem, fsz * 2) != hipSuccess) return fail(PENTAD_STATUS_ALLOC_FAILED);
    A->field[0] = A->field_mem;
    A->field[1] = (char *)A->field_mem + fsz;
  }
  if (launch_fill_int(A->d_steps, 0, st) != hipSuccess)
    return fail(PENTAD_STATUS_EXECUTION_FAILED);
  for (int d = 0; d < ndim; ++d) {
    int64_t ostr[3];
    strides_of(A->corder[d], A->dims, ndim, ostr);
    for (int k = 0; k < 5; ++k) {
      A->coef[d][k] = (char *)A->coef_mem + (size_t)(d * 5 + k) * csz;
      const void *src = coeffs[(coeff_sets == 1 ? 0 : d) * 5 + k];
      if (!src) return fail(PENTAD_STATUS_INVALID_VALUE);
      if (launch_permute3(src, coeff_elem_bytes, A->coef[d][k], A->cbytes, A->dims, ostr,
                          st) != hipSuccess)
        return fail(PENTAD_STATUS_EXECUTION_FAILED);
    }
  }
  // refinement mode on rotated layouts: the five permuted arrays of a
  // direction whose kernel has an interleaved form are rewritten interleaved
  // per group of rows into the same 5 * n floats (the 3-stream byte mix
  // streams about 2% faster than five separate coefficient streams)
  if (mode == 2 && A->rotating) {
    void *tmp = nullptr;
    for (int d = 0; d < ndim; ++d) {
      const int G = A->plan[d].kind == Plan::IrX
                        ? ir_x_interleave_group(A->plan[d].rows, A->dims[d])
                        : 0;
      if (G == 0 || A->n % G) continue;
      if (!tmp && hipMalloc(&tmp, csz * 5) != hipSuccess) return fail(PENTAD_STATUS_ALLOC_FAILED);
      const float *src[5];
      for (int k = 0; k < 5; ++k) src[k] = (const float *)A->coef[d][k];
      if (launch_interleave(src, (float *)tmp, A->n, G, st) != hipSuccess ||
          hipMemcpyAsync(A->coef[d][0], tmp, csz * 5, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
        return fail(PENTAD_STATUS_EXECUTION_FAILED);
      }
      A->il[d] = G;
    }
    if (tmp) {
      const hipError_t e = hipStreamSynchronize(st);
      (void)hipFree(tmp);
      if (e != hipSuccess) return fail(PENTAD_STATUS_EXECUTION_FAILED);
    }
  }
  if (!A->rotating) {
    // scratch for the in-place fallback kernels, enough for any of them
    size_t need = 0;
    for (int d = 0; d < ndim; ++d) {
      Lines ln;
      if (decompose(dims, ndim, d, &ln) != 0) return fail(PENTAD_STATUS_INVALID_VALUE);
      need = std::max(need, mode == 2 ? ir_lane_scratch_bytes(ln)
                                      : plan_scratch_bytes(Plan{}, ln, A->cbytes));
    }
    if (need && hipMalloc(&A->scratch, need) != hipSuccess)
      return fail(PENTAD_STATUS_ALLOC_FAILED);
    A->scratch_bytes = need;
  }
  for (int d = 0; d < ndim; ++d) {
    A->cfg[d] = A->rotating ? plan_config(A->plan[d]) : std::string("in place");
    if (A->il[d]) A->cfg[d].insert(A->cfg[d].find(' ') + 1, "interleaved ");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(PENTAD_STATUS_EXECUTION_FAILED);
  *out = A;
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_adi_create_stencil(pentadsolver_handle_t h,
                                               pentadsolver_adi_t *out, int mode, int ndim,
                                               const int *dims, const void *const *scales,
                                               const double *weights,
                                               const double *id_weights) {
  if (!out || !dims || !scales || !weights || !id_weights || mode < 0 || mode > 2 ||
      ndim < 2 || ndim > 3)
    return PENTAD_STATUS_INVALID_VALUE;
  for (int i = 0; i < ndim; ++i)
    if (dims[i] < 3) return PENTAD_STATUS_INVALID_VALUE;
  for (int d = 0; d < ndim; ++d)
    if (!scales[d]) return PENTAD_STATUS_INVALID_VALUE;
  Context *c = ctx_of(h);
  Adi *A = new Adi();
  A->c = c;
  A->mode = mode == 0 ? 5 : (mode == 1 ? 6 : 7);
  if (mode == 2) { // refine every solve to the limit of FP64 by default
    A->max_refine = 5;
    A->tol = 1e-15;
  }
  A->ndim = ndim;
  A->n = 1;
  for (int i = 0; i < ndim; ++i) {
    A->dims[i] = dims[i];
    A->n *= dims[i];
  }
  A->cbytes = A->fbytes = (mode == 1) ? 4 : 8;
  A->rotating = true;
  hipStream_t st = c->stream;
  auto fail = [&](pentadStatus_t s) {
    free_adi(A);
    return s;
  };
  // every direction on the x-like partitioned kernel (no in-place fallback)
  for (int d = 0; d < ndim; ++d) {
    Lines ln;
    ln.dir = 0;
    ln.len = A->dims[d];
    ln.n_sys = A->n / ln.len;
    bool planned = false;
    if (A->mode == 7) { // the refinement kernels' chunk sizes
      for (int r : {8, 6, 4, 10})
        if (!planned && ir_x_supported(r, ln.len)) {
          A->plan[d].kind = Plan::IrX;
          A->plan[d].rows = r;
          A->plan[d].lanes = (int)(ln.len / r);
          planned = true;
        }
    } else {
      planned = stencil_plan_x(A->fbytes, ln.len, A->plan[d]);
    }
    if (!planned) {
      std::fprintf(stderr, "[penta] adi stencil: no partitioned kernel for length %lld\n",
                   (long long)ln.len);
      return fail(PENTAD_STATUS_INVALID_VALUE);
    }
    for (int k = 0; k < 5; ++k) A->sw[d][k] = weights[d * 5 + k];
    A->sid[d] = id_weights[d];
    A->corder[d] = in_order(A, d, false);
  }
  // the scales: one per line of each direction, copied into the driver
  size_t off[3] = {0, 0, 0}, total = 0;
  for (int d = 0; d < ndim; ++d) {
    off[d] = total;
    total += ((size_t)(A->n / A->dims[d]) * A->fbytes + 255) & ~(size_t)255;
  }
  const size_t fsz = (size_t)A->n * A->fbytes;
  if (hipMalloc(&A->coef_mem, total) != hipSuccess ||
      hipMalloc(&A->field_mem, fsz * 2) != hipSuccess ||
      hipMalloc((void **)&A->d_steps, sizeof(int)) != hipSuccess)
    return fail(PENTAD_STATUS_ALLOC_FAILED);
  A->field[0] = A->field_mem;
  A->field[1] = (char *)A->field_mem + fsz;
  if (launch_fill_int(A->d_steps, 0, st) != hipSuccess)
    return fail(PENTAD_STATUS_EXECUTION_FAILED);
  for (int d = 0; d < ndim; ++d) {
    A->scale[d] = (char *)A->coef_mem + off[d];
    if (hipMemcpyAsync(A->scale[d], scales[d], (size_t)(A->n / A->dims[d]) * A->fbytes,
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(PENTAD_STATUS_EXECUTION_FAILED);
    A->cfg[d] = "stencil " + plan_config(A->plan[d]); // "stencil ir ..." in mode 2
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(PENTAD_STATUS_EXECUTION_FAILED);
  *out = A;
  return PENTAD_STATUS_SUCCESS;
}

static bool x_ok(const Adi *A, const void *x) {
  return x && (!A->rotating || (reinterpret_cast<uintptr_t>(x) & 15u) == 0);
}

pentadStatus_t pentadsolver_adi_step(pentadsolver_adi_t adi, void *x, int nsteps) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || !x_ok(A, x) || nsteps < 0) return PENTAD_STATUS_INVALID_VALUE;
  const hipError_t e = run_steps(A, x, nsteps, A->c->stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] adi step launch FAILED: %s\n", hipGetErrorString(e));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  if (A->rotating)
    for (int d = 0; d < A->ndim; ++d)
      note_kernel(A->c, d,
                  A->mode == 2 ? "ir"
                               : (A->mode >= 5 ? (A->mode == 7 ? "stencil ir" : "stencil")
                                               : (A->mode >= 3 ? "shared-fact" : "thomas-pcr")),
                  A->cfg[d]);
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_adi_graph_capture(pentadsolver_adi_t adi, void *x,
                                              int nsteps) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || !x_ok(A, x) || nsteps < 1) return PENTAD_STATUS_INVALID_VALUE;
  free_graph(A);
  hipStream_t st = A->c->stream;
  if (st == nullptr) { // the legacy null stream cannot be captured
    if (!A->cap_stream &&
        hipStreamCreateWithFlags(&A->cap_stream, hipStreamNonBlocking) != hipSuccess)
      return PENTAD_STATUS_EXECUTION_FAILED;
    st = A->cap_stream;
  }
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return PENTAD_STATUS_EXECUTION_FAILED;
  const hipError_t e = run_steps(A, x, nsteps, st);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(st, &g);
  if (e != hipSuccess || e2 != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  A->graph = g;
  if (hipGraphInstantiate(&A->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    free_graph(A);
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_adi_graph_launch(pentadsolver_adi_t adi, int reps) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || !A->exec || reps < 0) return PENTAD_STATUS_INVALID_VALUE;
  for (int s = 0; s < reps; ++s)
    if (hipGraphLaunch(A->exec, A->c->stream) != hipSuccess)
      return PENTAD_STATUS_EXECUTION_FAILED;
  return PENTAD_STATUS_SUCCESS;
}

// Pass B of the reference's driver (apps/adi_cuda.cu:216-256): HIP events
// around each direction's launch on the handle's stream, in steady-state
// steps (the first and last step of `reps + 2` are left out of the means).
// ms[d] receives the mean duration of direction d.
pentadStatus_t pentadsolver_adi_direction_ms(pentadsolver_adi_t adi, void *x,
                                             int reps, float *ms) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || !x_ok(A, x) || !ms || reps < 1) return PENTAD_STATUS_INVALID_VALUE;
  hipStream_t st = A->c->stream;
  hipEvent_t ev[4] = {};
  for (int i = 0; i < 4; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) {
      for (int j = 0; j < i; ++j) (void)hipEventDestroy(ev[j]);
      return PENTAD_STATUS_EXECUTION_FAILED;
    }
  const int nsteps = reps + 2;
  double acc[3] = {0, 0, 0};
  pentadStatus_t status = PENTAD_STATUS_SUCCESS;
  for (int s = 0; s < nsteps && status == PENTAD_STATUS_SUCCESS; ++s) {
    const void *in = x;
    if (hipEventRecord(ev[0], st) != hipSuccess) status = PENTAD_STATUS_EXECUTION_FAILED;
    for (int d = 0; d < A->ndim && status == PENTAD_STATUS_SUCCESS; ++d) {
      if (one_dir(A, d, s, nsteps, x, in, &in, st) != hipSuccess ||
          hipEventRecord(ev[d + 1], st) != hipSuccess)
        status = PENTAD_STATUS_EXECUTION_FAILED;
    }
    if (status != PENTAD_STATUS_SUCCESS) break;
    if (hipEventSynchronize(ev[A->ndim]) != hipSuccess) {
      status = PENTAD_STATUS_EXECUTION_FAILED;
      break;
    }
    if (s == 0 || s == nsteps - 1) continue;
    for (int d = 0; d < A->ndim; ++d) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev[d], ev[d + 1]) != hipSuccess)
        status = PENTAD_STATUS_EXECUTION_FAILED;
      acc[d] += t;
    }
  }
  for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
  for (int d = 0; d < 3; ++d) ms[d] = (d < A->ndim) ? (float)(acc[d] / reps) : 0.f;
  return status;
}

pentadStatus_t pentadsolver_adi_set_refine(pentadsolver_adi_t adi, int max_refine,
                                           double tol) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || max_refine < 0 || !(tol >= 0.0)) return PENTAD_STATUS_INVALID_VALUE;
  A->max_refine = max_refine;
  A->tol = tol;
  return PENTAD_STATUS_SUCCESS;
}

int pentadsolver_adi_refine_steps(pentadsolver_adi_t adi) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A) return -1;
  int v = -1;
  if (hipMemcpyAsync(&v, A->d_steps, sizeof(int), hipMemcpyDeviceToHost, A->c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(A->c->stream) != hipSuccess)
    return -1;
  if (launch_fill_int(A->d_steps, 0, A->c->stream) != hipSuccess) return -1;
  return v;
}

const char *pentadsolver_adi_config(pentadsolver_adi_t adi, int dir) {
  Adi *A = static_cast<Adi *>(adi);
  if (!A || dir < 0 || dir >= A->ndim) return "none";
  return A->cfg[dir].c_str();
}

void pentadsolver_adi_destroy(pentadsolver_adi_t *adi) {
  if (!adi || !*adi) return;
  free_adi(static_cast<Adi *>(*adi));
  *adi = nullptr;
}

} // extern "C"
