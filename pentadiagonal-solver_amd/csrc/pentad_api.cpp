// pentad_api.cpp -- C ABI, handle and dispatch of libpentadsolver.
//
// Replaces the host half of src/cuda/singlenode/pentad_cuda.cu (:1129-1652):
// direction routing, line decomposition, per-direction algorithm selection
// with the PENTA_*ALGO environment names, the executed-kernel record and the
// buffer sizing.  Differences by design: state lives in a handle (stream,
// overrides, tuning, record) instead of process globals, dispatch reports a
// status, and the buffer size is the true need of the kernel that will run.
#include "pentad_api_internal.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

namespace pentad {

int decompose(const int *dims, int ndim, int solvedim, Lines *out) {
  if (dims == nullptr || ndim < 1 || solvedim < 0 || solvedim >= ndim)
    return -1;
  int64_t n_in = 1, n_out = 1;
  for (int i = 0; i < ndim; ++i)
    if (dims[i] < 1) return -1;
  for (int i = 0; i < solvedim; ++i) n_in *= dims[i];
  for (int i = solvedim + 1; i < ndim; ++i) n_out *= dims[i];
  Lines ln;
  ln.len = dims[solvedim];
  ln.n_sys = n_in * n_out;
  if (solvedim == 0) {
    ln.dir = 0;
    ln.stride = 1;
    ln.n_in = 1;
  } else {
    ln.dir = (solvedim == ndim - 1) ? 2 : 1;
    ln.stride = n_in;
    ln.n_in = n_in;
  }
  *out = ln;
  return 0;
}

int decompose_padded(const int *dims, const int *pads, int ndim, int solvedim, Lines *out) {
  if (decompose(dims, ndim, solvedim, out) != 0) return -1;
  if (pads == nullptr) return 0;
  if (ndim > 3) return -1; // the reference's MAXDIM (include/pentad_common.hpp:17)
  int64_t st[3] = {1, 1, 1};
  bool padded = false;
  for (int i = 0; i < ndim; ++i) {
    if (pads[i] < dims[i]) return -1;
    if (i > 0) st[i] = st[i - 1] * pads[i - 1];
    // padding of the last axis moves no element
    if (i + 1 < ndim && pads[i] != dims[i]) padded = true;
  }
  if (!padded) return 0;
  Lines &ln = *out;
  ln.stride = st[solvedim];
  // the (at most two) other axes: the lower one indexes lines within a group
  int other[2] = {-1, -1}, no = 0;
  for (int i = 0; i < ndim; ++i)
    if (i != solvedim) other[no++] = i;
  ln.n_in = no > 0 ? dims[other[0]] : 1;
  ln.lstep = no > 0 ? st[other[0]] : 1;
  ln.gstep = no > 1 ? st[other[1]] : 0;
  return 0;
}

// ---------------------------------------------------------------------------
// contexts
// ---------------------------------------------------------------------------
namespace {
std::mutex g_ran_mu;
const char *g_ran[3] = {"none", "none", "none"};
Context g_default;
} // namespace

Context *ctx_of(pentadsolver_handle_t h) {
  return h ? static_cast<Context *>(h) : &g_default;
}

void note_kernel(Context *c, int dir, const char *name, const std::string &cfg) {
  if (dir < 0 || dir > 2) return;
  {
    std::lock_guard<std::mutex> lk(g_ran_mu);
    g_ran[dir] = name;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  c->ran[dir] = name;
  c->ran_cfg[dir] = cfg;
}

const char *ran_global(int dir) {
  std::lock_guard<std::mutex> lk(g_ran_mu);
  return (dir >= 0 && dir < 3) ? g_ran[dir] : "none";
}

void set_status(Context *c, pentadStatus_t s) { c->last_status = s; }

void destroy_context_resources(Context * /*c*/) {}

// Algorithm requested for a direction: handle override, then the
// environment with the reference's precedence (PENTA_YALGO / PENTA_ZALGO
// over PENTA_YZALGO, src/cuda/singlenode/pentad_cuda.cu:1163-1170), then
// "auto".
std::string requested_algo(Context *c, int dir) {
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->algo[dir].empty()) return c->algo[dir];
    if (c->env_frozen) return c->env_algo[dir];
  }
  return env_algo(dir);
}

std::string env_algo(int dir) {
  const char *v = nullptr;
  if (dir == 0) {
    v = std::getenv("PENTA_XALGO");
  } else {
    v = std::getenv(dir == 1 ? "PENTA_YALGO" : "PENTA_ZALGO");
    if (v == nullptr || v[0] == '\0') v = std::getenv("PENTA_YZALGO");
  }
  if (v == nullptr || v[0] == '\0') return "auto";
  std::string s(v);
  if (s == "legacy") s = "naive";
  return s;
}

bool autotune_env() {
  const char *v = std::getenv("PENTA_AUTOTUNE");
  return v != nullptr && v[0] == '1';
}

bool debug_launch() {
  const char *v = std::getenv("PENTA_DEBUG_LAUNCH");
  return v != nullptr && v[0] == '1';
}

// ---------------------------------------------------------------------------
// plan: which kernel a solve will launch
// ---------------------------------------------------------------------------
static bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

// x lines for the vector-load kernels: every array 16-byte aligned, and in a
// padded layout the line and group pitches keep each line start aligned
static bool x_aligned(const Lines &ln, int elem_bytes, const void *const ptrs[6]) {
  for (int i = 0; i < 6; ++i)
    if (!aligned16(ptrs[i])) return false;
  return ln.dense() || ((ln.lstep * elem_bytes) % 16 == 0 && (ln.gstep * elem_bytes) % 16 == 0);
}

// long x lines on the strided partitioned kernel: 16- or 12-row chunks (a
// padded last chunk is allowed), the widest workgroup of whole systems
bool plan_x_long(const Lines &ln, int elem_bytes, Plan &pl) {
  for (int S = 8; S >= 1; S >>= 1)
    for (int r : {16, 12, 8}) {
      if (!part_yz_supported(elem_bytes, r, S, ln.len)) continue;
      pl.kind = Plan::PartYZ;
      pl.rows = r;
      pl.sys_per_wg = S;
      return true;
    }
  return false;
}

// Lines shared by two workgroups (kernels_split.hip): chunks of 16 rows (8
// for FP64 where 16 do not halve the line), S = 512 / P systems per line set
// (32 or 64).  Automatic choice only for the outermost direction when a row
// step is at least 1 MB (every row of a line in its own page) and there are
// enough line sets to fill the chip; "thomas-pcr-split" by name wherever the
// kernel applies.  Measured at 512^3 along z (one workgroup per line -> two):
// FP64 2.10 -> 1.81 ms, FP32 1.30 -> 1.01, FP32 coefficients widened 1.43 ->
// 1.34; along y, and for lines of 256 rows or fewer (whose one-workgroup
// kernel already holds 32 systems), the split measured slower.
bool plan_split(const Lines &ln, int elem_bytes, bool forced, Plan &pl) {
  if (ln.dir == 0) return false;
  if (!forced && (ln.dir != 2 || ln.len <= 384 || ln.stride * elem_bytes < (1 << 18) ||
                  ln.n_sys < 8192))
    return false;
  for (int r : {16, 8}) {
    if (ln.len % (2 * r) != 0) continue;
    const int64_t P = ln.len / (2 * r);
    if (P < 2 || P > 16) continue;
    const int S = P <= 8 ? 64 : 32;
    if (!part_yz_split_supported(elem_bytes, r, S, ln.len)) continue;
    pl.kind = Plan::PartYZ;
    pl.rows = r;
    pl.sys_per_wg = S;
    pl.split = true;
    return true;
  }
  return false;
}

Plan plan_solve(Context *c, const Lines &ln, int elem_bytes, const void *const ptrs[6],
                bool allow_split) {
  Plan pl;
  std::string want = requested_algo(c, ln.dir);
  // the two-workgroup strided kernel by name (an extension of the
  // reference's names); where it does not apply, "thomas-pcr"
  const bool want_split = want == "thomas-pcr-split";
  if (want_split) want = "thomas-pcr";
  // x: an unrecognised name runs the automatic choice, as the reference's
  // PENTA_XALGO does (pentad_cuda.cu:1463-1488); along y / z any name other
  // than auto, thomas-pcr and shared-fact runs the one-lane kernel, as the
  // reference's strided dispatch does (its "legacy")
  if (ln.dir == 0 && want != "naive" && want != "transpose" && want != "thomas-pcr" &&
      want != "shared-fact")
    want = "auto";
  const bool is64 = (elem_bytes == 8);
  const TuneKey key{elem_bytes, ln.dir, ln.len};
  TuneVal tuned;
  bool have_tuned = false;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tune.find(key);
    if (it != c->tune.end()) {
      tuned = it->second;
      have_tuned = true;
    }
    if (c->force_rows[ln.dir] > 0) { // explicit parameters win over tuning
      tuned.kind = ln.dir == 0 ? Plan::PartX : Plan::PartYZ;
      tuned.rows = c->force_rows[ln.dir];
      tuned.lanes = c->force_width[ln.dir];
      tuned.sys_per_wg = c->force_width[ln.dir];
      have_tuned = true;
    }
  }

  if (want == "naive") {
    pl.kind = Plan::Naive;
    return pl;
  }

  if (want == "shared-fact") {
    // opt-in only (every system is solved with system 0's coefficients, as
    // in src/cuda/singlenode/pentad_cuda.cu:962-963); a length the kernels
    // do not cover falls back as the reference's does: x to the automatic
    // choice (:1480-1488), y / z to the one-lane kernel (:1330-1359)
    static const int sf64[] = {8, 16, 6, 4};
    static const int sf32[] = {16, 8, 6, 4};
    const int *pref = is64 ? sf64 : sf32;
    if (ln.dir == 0) {
      const bool al = x_aligned(ln, elem_bytes, ptrs);
      for (int k = 0; al && k < 4; ++k) {
        if (sf_x_supported(elem_bytes, pref[k], ln.len)) {
          pl.kind = Plan::SfX;
          pl.rows = pref[k];
          pl.lanes = (int)(ln.len / pref[k]);
          return pl;
        }
      }
      want = "auto"; // x: the reference falls through to its automatic choice
    } else {
      for (int k = 0; k < 4; ++k) {
        for (int S = 64; S >= 1; S >>= 1) {
          if (sf_yz_supported(elem_bytes, pref[k], S, ln.len)) {
            pl.kind = Plan::SfYZ;
            pl.rows = pref[k];
            pl.sys_per_wg = S;
            return pl;
          }
        }
      }
    }
  }

  // LDS-staged x solve for any length (the "transpose" capability): asked
  // for by name, or the automatic choice when no partitioned shape fits
  auto plan_tr = [&](Plan &q) {
    if (!ln.dense()) return false; // stages spans of adjacent systems
    for (int r : {8, 16}) {
      int lanes = 1;
      while ((int64_t)lanes * r < ln.len) lanes <<= 1;
      if (tr_x_supported(r, lanes, ln.len)) {
        q.kind = Plan::TrX;
        q.rows = r;
        q.lanes = lanes;
        return true;
      }
    }
    return false;
  };
  if (ln.dir == 0 && want == "transpose" && plan_tr(pl)) return pl;

  if (ln.dir == 0) {
    const bool al = x_aligned(ln, elem_bytes, ptrs);
    if (al && (want == "auto" || want == "thomas-pcr")) {
      if (have_tuned && tuned.kind == Plan::PartX &&
          (part_x_supported(elem_bytes, tuned.rows, tuned.lanes) ||
           part_x_wide_supported(elem_bytes, tuned.rows, tuned.lanes)) &&
          (int64_t)tuned.rows * tuned.lanes == ln.len) {
        pl.kind = Plan::PartX;
        pl.rows = tuned.rows;
        pl.lanes = tuned.lanes;
        return pl;
      }
      if (have_tuned && tuned.kind == Plan::TrX && ln.dense() &&
          tr_x_supported(tuned.rows, tuned.lanes, ln.len)) {
        pl = tuned;
        return pl;
      }
      if (have_tuned && tuned.kind == Plan::PartYZ &&
          part_yz_supported(elem_bytes, tuned.rows, tuned.sys_per_wg, ln.len)) {
        pl = tuned;
        return pl;
      }
      // preferred rows per lane: 8 (FP64) / 16 (FP32), then the others
      // (FP32 lines of 10 * 2^k rows on the LDS-staged kernel instead: 640
      // rows 1.07 ms against 0.36 ms with 10 rows per lane)
      static const int pref64[] = {8, 6, 10, 12, 4, 16};
      static const int pref32[] = {16, 12, 8, 10, 6, 4};
      const int *pref = is64 ? pref64 : pref32;
      const int npref = 6;
      // lines of 128 or 256 chunks of 8 rows (FP64) / 16 rows (FP32): the
      // wide kernel, a system's lanes across wavefronts (in place, 2^25
      // rows: FP64 1024-row lines 0.39 ms against 1.08 on the LDS-staged
      // kernel and 1.22 with 16 x 64; 2048 rows 0.37 against 1.03; FP32
      // 2048 / 4096 rows 0.21 against 0.50)
      {
        const int r = is64 ? 8 : 16;
        const int64_t lanes = ln.len / r;
        if (ln.len % r == 0 && lanes <= 256 && part_x_wide_supported(elem_bytes, r, (int)lanes)) {
          pl.kind = Plan::PartX;
          pl.rows = r;
          pl.lanes = (int)lanes;
          return pl;
        }
      }
      for (int k = 0; k < npref; ++k) {
        const int r = pref[k];
        if (ln.len % r) continue;
        const int64_t lanes = ln.len / r;
        if (lanes > 64 || lanes < 2 || (lanes & (lanes - 1))) continue;
        if (!part_x_supported(elem_bytes, r, (int)lanes)) continue;
        pl.kind = Plan::PartX;
        pl.rows = r;
        pl.lanes = (int)lanes;
        return pl;
      }
    }
    // lines longer than one wavefront's chunks cover (above 64 x 16 rows):
    // "thomas-pcr" by name, and the automatic choice where the LDS-staged
    // kernel cannot stage (padded layouts), take the workgroup-wide
    // partitioned kernel of the strided directions, whose chunks are
    // contiguous rows here (stride 1), separators exchanged through LDS; up
    // to 8192 rows (FP64) in one workgroup, no scratch
    if (want == "thomas-pcr" && ln.len > 16 * 64 && plan_x_long(ln, elem_bytes, pl)) return pl;
    if ((want == "auto" || want == "thomas-pcr") && plan_tr(pl)) return pl;
    if ((want == "auto" || want == "thomas-pcr") && plan_x_long(ln, elem_bytes, pl)) return pl;
    pl.kind = Plan::Naive;
    return pl;
  }

  // strided directions
  if (want == "auto" || want == "thomas-pcr") {
    if (allow_split && !(have_tuned && c->force_rows[ln.dir] > 0) &&
        plan_split(ln, elem_bytes, want_split, pl))
      return pl;
    if (have_tuned && tuned.kind == Plan::PartYZ &&
        part_yz_supported(elem_bytes, tuned.rows, tuned.sys_per_wg, ln.len)) {
      pl.kind = Plan::PartYZ;
      pl.rows = tuned.rows;
      pl.sys_per_wg = tuned.sys_per_wg;
      return pl;
    }
    // lines of 32..384 rows: large chunks -- 16 or 12 rows, whichever
    // splits the line into a power-of-two number of whole chunks, else
    // FP64 16 rows, FP32 16 rows up to 128 and 12 above (its kernel runs
    // 1024-thread workgroups) -- and 32 systems per workgroup, 16 when
    // that leaves fewer than 2048 workgroups.  Measured with 2^26 rows
    // along z, against the whole-chunks-first choice below: FP64 256 rows
    // 0.86 vs 1.22 ms, 128 rows 0.79 vs 0.93, 100 rows 0.81 vs 0.93; FP32
    // 384 rows 0.41 vs 0.49; within 5% of the best configuration the
    // autotuner finds at the lengths measured (40 .. 384).
    if (ln.len >= 32 && ln.len <= 384) {
      auto pow2_chunks = [&](int r) {
        const int64_t P = ln.len / r;
        return ln.len % r == 0 && (P & (P - 1)) == 0;
      };
      const int r = pow2_chunks(16)   ? 16
                    : pow2_chunks(12) ? 12
                    : (is64 || ln.len <= 128) ? 16 : 12;
      int S = 32;
      while (S > 8 && !part_yz_supported(elem_bytes, r, S, ln.len)) S >>= 1;
      if (S > 16 && ln.n_sys / S < 2048) S = 16;
      if (is64 && ln.dir == 1 && S > 16) S = 16; // middle axis: 0.69 vs 0.81 ms, 192 rows
      if (part_yz_supported(elem_bytes, r, S, ln.len)) {
        pl.kind = Plan::PartYZ;
        pl.rows = r;
        pl.sys_per_wg = S;
        return pl;
      }
    }
    static const int pref64[] = {8, 6, 12, 4, 16};
    static const int pref32[] = {16, 12, 8, 6, 4};
    const int *pref = is64 ? pref64 : pref32;
    // lengths that split into whole chunks first, then padded last chunks
    for (int k = 0; k < 10; ++k) {
      const int r = pref[k % 5];
      if ((k < 5) != (ln.len % r == 0)) continue;
      for (int S = 64; S >= 16; S >>= 1) {
        if (part_yz_supported(elem_bytes, r, S, ln.len)) {
          pl.kind = Plan::PartYZ;
          pl.rows = r;
          pl.sys_per_wg = S;
          return pl;
        }
      }
    }
    // long lines (no 16 systems of them fit one workgroup): the widest
    // workgroup that fits, before the one-lane kernel with its scratch
    // passes
    for (int S = 8; S >= 1; S >>= 1) {
      for (int k = 0; k < 10; ++k) {
        const int r = pref[k % 5];
        if ((k < 5) != (ln.len % r == 0)) continue;
        if (part_yz_supported(elem_bytes, r, S, ln.len)) {
          pl.kind = Plan::PartYZ;
          pl.rows = r;
          pl.sys_per_wg = S;
          return pl;
        }
      }
    }
  }
  pl.kind = Plan::Naive;
  return pl;
}

// FP32 + refinement: partitioned kernels when the length splits into
// supported chunks, else the one-lane-per-system kernel with scratch.
Plan plan_ir(Context *c, const Lines &ln, const void *const ptrs[6], bool allow_split) {
  Plan pl;
  pl.kind = Plan::IrLane;
  const std::string want = requested_algo(c, ln.dir);
  if (want == "naive") return pl;
  int frows = 0, fwidth = 0;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    frows = c->force_rows[ln.dir];
    fwidth = c->force_width[ln.dir];
  }
  {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tune.find(TuneKey{TUNE_IR, ln.dir, ln.len});
    if (frows <= 0 && it != c->tune.end() && it->second.kind == Plan::IrWideYZ &&
        ln.dir != 0 && part_yz_supported(8, it->second.rows, it->second.sys_per_wg, ln.len))
      return it->second;
    if (frows <= 0 && it != c->tune.end()) {
      frows = it->second.rows;
      fwidth = ln.dir == 0 ? it->second.lanes : it->second.sys_per_wg;
    }
  }
  if (ln.dir == 0) {
    const bool al = x_aligned(ln, 4, ptrs); // (FP32 pitches aligned => FP64 ones too)
    // lengths (or alignments) without the vector-load kernel: LDS-staged
    auto staged = [&](Plan &q) {
      if (!ln.dense() || !ir_tr_x_supported(8, ln.len)) return false;
      q.kind = Plan::IrTrX;
      q.rows = 8;
      int lanes = 1;
      while ((int64_t)lanes * 8 < ln.len) lanes <<= 1;
      q.lanes = lanes;
      return true;
    };
    if (want == "transpose" && staged(pl)) return pl;
    if (!al) {
      staged(pl);
      return pl;
    }
    if (frows > 0 && ir_x_supported(frows, ln.len)) {
      pl.kind = Plan::IrX;
      pl.rows = frows;
      pl.lanes = (int)(ln.len / frows);
      return pl;
    }
    static const int pref[] = {8, 6, 4, 10};
    for (int r : pref) {
      if (ir_x_supported(r, ln.len)) {
        pl.kind = Plan::IrX;
        pl.rows = r;
        pl.lanes = (int)(ln.len / r);
        return pl;
      }
    }
    if (staged(pl)) return pl;
    // lengths no refinement kernel covers (past 64 chunks of 10 rows, or
    // no power-of-two split): one FP64 solve of the coefficients widened on
    // load -- the system the refinement converges to, in one pass over the
    // same 36 B per row -- with the FP64 x kernel of whole chunks, or for
    // lines past 64 x 16 rows the workgroup-wide kernel (stride 1).  These
    // lengths took the one-lane kernel with its scratch passes before
    // (768-row lines: 12 vs 1.7 ms for the FP64 solve of the same step).
    // Lines of 128 or 256 chunks of 8 rows take the wide x kernel first.
    if (ln.len % 8 == 0 && part_x_wide_supported(8, 8, (int)(ln.len / 8))) {
      pl.kind = Plan::IrWideX;
      pl.rows = 8;
      pl.lanes = (int)(ln.len / 8);
      return pl;
    }
    for (int r : {8, 6, 10, 12, 16, 4}) {
      const int64_t lanes = ln.len / r;
      if (ln.len % r == 0 && part_x_supported(8, r, (int)lanes)) {
        pl.kind = Plan::IrWideX;
        pl.rows = r;
        pl.lanes = (int)lanes;
        return pl;
      }
    }
    Plan q;
    if (plan_x_long(ln, 8, q)) {
      q.kind = Plan::IrWideYZ;
      return q;
    }
    return pl;
  }
  if (frows > 0 && ir_yz_supported(frows, fwidth, ln.len)) {
    pl.kind = Plan::IrYZ;
    pl.rows = frows;
    pl.sys_per_wg = fwidth;
    return pl;
  }
  // strided lines, automatic choice: one FP64 solve of the coefficients
  // widened on load.  The refinement converges to the widened system's
  // solution to FP64 rounding; the direct solve gets there in one pass over
  // the same 36 B per row, where every refinement sweep along y / z
  // exchanges separators through LDS.  Measured at 512^3: y 2.81 -> 1.15 ms,
  // z 3.54 -> 1.44 ms, solutions 4e-15 apart.  "thomas-pcr" asked for by
  // name keeps the refinement kernel.
  if (want == "auto" || want == "thomas-pcr-split") {
    Plan q = plan_solve(c, ln, 8, ptrs, allow_split);
    if (q.kind == Plan::PartYZ) {
      q.kind = Plan::IrWideYZ;
      return q;
    }
  }
  // 6-row chunks when they, and not 8-row ones, give a power-of-two chunk
  // count (measured with 2^26 rows: 384 rows along z 1.43 vs 1.66 ms, 192
  // rows 1.14 vs 1.39 ms)
  auto pow2_chunks = [&](int r) {
    const int64_t P = ln.len / r;
    return ln.len % r == 0 && (P & (P - 1)) == 0;
  };
  static const int pref8[] = {8, 6, 4, 10}, pref6[] = {6, 8, 4, 10};
  const int *pref = (!pow2_chunks(8) && pow2_chunks(6)) ? pref6 : pref8;
  for (int k = 0; k < 8; ++k) { // whole chunks first, then a padded last chunk
    const int r = pref[k % 4];
    if ((k < 4) != (ln.len % r == 0)) continue;
    for (int S = 64; S >= 4; S >>= 1) {
      if (ir_yz_supported(r, S, ln.len)) {
        pl.kind = Plan::IrYZ;
        pl.rows = r;
        pl.sys_per_wg = S;
        return pl;
      }
    }
  }
  return pl;
}

const char *plan_algo_name(const Plan &pl) {
  switch (pl.kind) {
  case Plan::PartX:
  case Plan::PartYZ: return "thomas-pcr";
  case Plan::TrX: return "transpose";
  case Plan::SfX:
  case Plan::SfYZ: return "shared-fact";
  default: return "naive";
  }
}

size_t plan_scratch_bytes(const Plan &pl, const Lines &ln, int elem_bytes) {
  if (pl.kind == Plan::IrLane) return ir_lane_scratch_bytes(ln);
  if (pl.split && pl.kind == Plan::PartYZ)
    return part_yz_split_scratch_bytes(ln, pl.sys_per_wg, elem_bytes);
  if (pl.split && pl.kind == Plan::IrWideYZ) // an FP64 solve
    return part_yz_split_scratch_bytes(ln, pl.sys_per_wg, 8);
  if (pl.kind == Plan::Naive)
    return (size_t)3 * (size_t)ln.n_sys * (size_t)ln.len * (size_t)elem_bytes;
  if (pl.kind == Plan::SfX || pl.kind == Plan::SfYZ)
    return sf_table_bytes(elem_bytes, pl.rows, ln.len);
  return 0;
}

std::string plan_config(const Plan &pl) {
  char buf[96];
  switch (pl.kind) {
  case Plan::PartX:
    std::snprintf(buf, sizeof buf, "thomas-pcr rows=%d lanes=%d", pl.rows,
                  pl.lanes);
    return buf;
  case Plan::PartYZ:
    std::snprintf(buf, sizeof buf, "thomas-pcr%s rows=%d sys_per_wg=%d",
                  pl.split ? " split" : "", pl.rows, pl.sys_per_wg);
    return buf;
  case Plan::IrX:
    std::snprintf(buf, sizeof buf, "ir thomas-pcr rows=%d lanes=%d", pl.rows,
                  pl.lanes);
    return buf;
  case Plan::IrYZ:
    std::snprintf(buf, sizeof buf, "ir thomas-pcr rows=%d sys_per_wg=%d",
                  pl.rows, pl.sys_per_wg);
    return buf;
  case Plan::IrLane:
    return "ir naive";
  case Plan::IrTrX:
    std::snprintf(buf, sizeof buf, "ir transpose rows=%d lanes=%d", pl.rows, pl.lanes);
    return buf;
  case Plan::IrWideYZ:
    std::snprintf(buf, sizeof buf, "ir widened%s rows=%d sys_per_wg=%d",
                  pl.split ? " split" : "", pl.rows, pl.sys_per_wg);
    return buf;
  case Plan::IrWideX:
    std::snprintf(buf, sizeof buf, "ir widened rows=%d lanes=%d", pl.rows, pl.lanes);
    return buf;
  case Plan::SfX:
    std::snprintf(buf, sizeof buf, "shared-fact rows=%d lanes=%d", pl.rows, pl.lanes);
    return buf;
  case Plan::TrX:
    std::snprintf(buf, sizeof buf, "transpose rows=%d lanes=%d", pl.rows, pl.lanes);
    return buf;
  case Plan::SfYZ:
    std::snprintf(buf, sizeof buf, "shared-fact rows=%d sys_per_wg=%d", pl.rows,
                  pl.sys_per_wg);
    return buf;
  default:
    return "naive";
  }
}

// ---------------------------------------------------------------------------
// launch one plan (no fallback); used by the solve and by the autotuner
// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_plan(const Plan &pl, const Diags<T> &a, T *x, const Lines &ln,
                       void *buffer, hipStream_t st) {
  switch (pl.kind) {
  case Plan::PartX: return launch_part_x<T>(a, x, ln, pl.rows, pl.lanes, st);
  case Plan::PartYZ:
    if (pl.split) return launch_part_yz_split<T>(a, x, ln, pl.rows, pl.sys_per_wg, buffer, st);
    return launch_part_yz<T>(a, x, ln, pl.rows, pl.sys_per_wg, st);
  case Plan::TrX: return launch_tr_x<T>(a, x, ln, pl.rows, pl.lanes, st);
  case Plan::Naive:
    if (buffer == nullptr) return hipErrorInvalidValue;
    return launch_thomas_lane<T>(a, x, ln, static_cast<T *>(buffer), st);
  case Plan::SfX:
  case Plan::SfYZ: {
    if (buffer == nullptr) return hipErrorInvalidValue;
    T *tab = static_cast<T *>(buffer);
    const bool ok = pl.kind == Plan::SfX
                        ? sf_x_supported((int)sizeof(T), pl.rows, ln.len)
                        : sf_yz_supported((int)sizeof(T), pl.rows, pl.sys_per_wg, ln.len);
    if (!ok) return hipErrorInvalidValue; // decline before anything is launched
    hipError_t e = launch_sf_factor<T>(a, ln, pl.rows, tab, st);
    if (e != hipSuccess) return e;
    if (pl.kind == Plan::SfX)
      return launch_sf_x_map<T>(tab, x, x, false, ln.n_sys, ln.len, xmap_inplace(ln),
                                pl.rows, st);
    return launch_sf_yz<T>(tab, x, ln, pl.rows, pl.sys_per_wg, st);
  }
  default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_plan<float>(const Plan &, const Diags<float> &, float *,
                                       const Lines &, void *, hipStream_t);
template hipError_t launch_plan<double>(const Plan &, const Diags<double> &,
                                        double *, const Lines &, void *,
                                        hipStream_t);

hipError_t launch_plan_ir(const Plan &pl, const Diags<float> &a, double *x,
                          const Lines &ln, void *buffer, int max_steps,
                          double tol, int *steps, hipStream_t st) {
  switch (pl.kind) {
  case Plan::IrX:
    return launch_ir_x(a, x, ln, pl.rows, max_steps, tol, steps, st);
  case Plan::IrYZ:
    return launch_ir_yz(a, x, ln, pl.rows, pl.sys_per_wg, max_steps, tol, steps,
                        st);
  case Plan::IrTrX:
    return launch_ir_tr_x(a, x, ln, pl.rows, max_steps, tol, steps, st);
  case Plan::IrWideYZ: // no refinement sweep: *steps is left as it is
    if (pl.split)
      return launch_part_yz_split_widened(a, x, ln, pl.rows, pl.sys_per_wg, buffer, st);
    return launch_part_yz_widened(a, x, ln, pl.rows, pl.sys_per_wg, st);
  case Plan::IrWideX: // the same along x, in place
    return launch_part_x_map_widened(a, x, x, false, ln.n_sys, ln.len, xmap_inplace(ln),
                                     pl.rows, pl.lanes, st);
  case Plan::IrLane:
    if (buffer == nullptr) return hipErrorInvalidValue;
    return launch_ir_lane(a, x, ln, buffer, max_steps, tol, steps, st);
  default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// the batched solve
// ---------------------------------------------------------------------------
// pads == nullptr: the dense layout; else the padded layout of
// decompose_padded (extents pads[k] >= dims[k] in memory).  buffer_cap: bytes
// the buffer holds (SIZE_MAX: the caller sized it with the matching
// _buffer_size query); a kernel needing more is not launched.
template <typename T>
pentadStatus_t gpsv_batch_pitched(Context *c, const T *ds, const T *dl, const T *d,
                                  const T *du, const T *dw, T *x, const int *dims,
                                  const int *pads, int ndim, int solvedim, void *buffer,
                                  size_t buffer_cap) {
  Lines ln;
  if (decompose_padded(dims, pads, ndim, solvedim, &ln) != 0 || ln.len < 3) {
    std::fprintf(stderr, "[penta] invalid shape (ndim %d, solvedim %d%s)\n",
                 ndim, solvedim, pads ? ", padded" : "");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!ds || !dl || !d || !du || !dw || !x) {
    std::fprintf(stderr, "[penta] null array pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  Plan pl = plan_solve(c, ln, (int)sizeof(T), ptrs, true);
  const Diags<T> a{ds, dl, d, du, dw};
  hipStream_t st = c->stream;
  hipError_t err = hipSuccess;

  // an algorithm asked for by name that has no kernel for this shape: the
  // automatic choice runs instead, which PENTA_DEBUG_LAUNCH=1 reports (the
  // reference's launch_succeeded diagnostic, pentad_cuda.cu:17-37)
  if (debug_launch()) {
    const std::string want = requested_algo(c, ln.dir);
    const char *got = plan_algo_name(pl);
    if (want != "auto" && want != got)
      std::fprintf(stderr,
                   "[penta] %s declined along solvedim %d (length %lld) -- falling back to %s\n",
                   want.c_str(), solvedim, (long long)ln.len, got);
  }

  // the scratch a plan needs is there, or nothing is launched
  auto scratch_ok = [&](const Plan &q) {
    const size_t need = plan_scratch_bytes(q, ln, (int)sizeof(T));
    if (need == 0 || (buffer != nullptr && need <= buffer_cap)) return true;
    if (buffer == nullptr)
      std::fprintf(stderr,
                   "[penta] %s along solvedim %d needs %zu bytes of scratch "
                   "(pentadsolver_*_gpsv_batch_buffer_size_ext); got NULL\n",
                   plan_algo_name(q), solvedim, need);
    else
      std::fprintf(stderr,
                   "[penta] %s along solvedim %d needs %zu bytes of scratch, more than the "
                   "%zu bytes available\n",
                   plan_algo_name(q), solvedim, need, buffer_cap);
    return false;
  };

  // the two-workgroup kernel without room for its trade: the one-workgroup
  // plan (no scratch)
  if (pl.split && !(buffer != nullptr && plan_scratch_bytes(pl, ln, (int)sizeof(T)) <= buffer_cap))
    pl = plan_solve(c, ln, (int)sizeof(T), ptrs, false);
  const bool sf = pl.kind == Plan::SfX || pl.kind == Plan::SfYZ;
  if (!scratch_ok(pl)) return PENTAD_STATUS_INVALID_VALUE;
  if (pl.kind != Plan::Naive) err = launch_plan<T>(pl, a, x, ln, buffer, st);
  // (a declined shared-fact launch does not fall back: the caller's buffer
  // is sized for the factor table, not for the one-lane kernel's scratch)
  if (pl.kind != Plan::Naive && !sf && err != hipSuccess) {
    // an opt-in kernel that declines falls back, as the reference does
    // (src/cuda/singlenode/pentad_cuda.cu:17-37)
    if (debug_launch())
      std::fprintf(stderr, "[penta] %s declined: %s -- falling back\n",
                   plan_config(pl).c_str(), hipGetErrorString(err));
    pl = Plan{};
    if (!scratch_ok(pl)) return PENTAD_STATUS_INVALID_VALUE;
  }
  if (pl.kind == Plan::Naive)
    err = launch_thomas_lane<T>(a, x, ln, static_cast<T *>(buffer), st);
  if (err != hipSuccess) {
    std::fprintf(stderr, "[penta] solvedim %d launch FAILED: %s -- output is invalid\n",
                 solvedim, hipGetErrorString(err));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  note_kernel(c, ln.dir, plan_algo_name(pl), plan_config(pl));
  return PENTAD_STATUS_SUCCESS;
}
template pentadStatus_t gpsv_batch_pitched<float>(Context *, const float *, const float *,
                                                  const float *, const float *, const float *,
                                                  float *, const int *, const int *, int, int,
                                                  void *, size_t);
template pentadStatus_t gpsv_batch_pitched<double>(Context *, const double *, const double *,
                                                   const double *, const double *,
                                                   const double *, double *, const int *,
                                                   const int *, int, int, void *, size_t);

pentadStatus_t gpsv_batch_ir(Context *c, const float *ds, const float *dl,
                             const float *d, const float *du, const float *dw,
                             double *x, const int *dims, int ndim, int solvedim,
                             int max_steps, double tol, int *steps,
                             void *buffer) {
  Lines ln;
  if (decompose(dims, ndim, solvedim, &ln) != 0 || ln.len < 3 ||
      max_steps < 0) {
    std::fprintf(stderr, "[penta] ir: invalid arguments (ndim %d, solvedim %d)\n",
                 ndim, solvedim);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!ds || !dl || !d || !du || !dw || !x) {
    std::fprintf(stderr, "[penta] ir: null array pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  Plan pl = plan_ir(c, ln, ptrs, true);
  // the two-workgroup kernel without room for its trade: the one-workgroup plan
  if (pl.split && buffer == nullptr) pl = plan_ir(c, ln, ptrs, false);
  const Diags<float> a{ds, dl, d, du, dw};
  hipError_t err = hipSuccess;
  if (pl.kind != Plan::IrLane)
    err = launch_plan_ir(pl, a, x, ln, pl.split ? buffer : nullptr, max_steps, tol, steps,
                         c->stream);
  if (pl.kind != Plan::IrLane && err != hipSuccess) {
    if (debug_launch())
      std::fprintf(stderr, "[penta] %s declined: %s -- falling back\n",
                   plan_config(pl).c_str(), hipGetErrorString(err));
    pl = Plan{};
    pl.kind = Plan::IrLane;
  }
  if (pl.kind == Plan::IrLane) {
    if (buffer == nullptr) {
      std::fprintf(stderr,
                   "[penta] ir solvedim %d needs %zu bytes of scratch; got NULL\n",
                   solvedim, ir_lane_scratch_bytes(ln));
      return PENTAD_STATUS_INVALID_VALUE;
    }
    err = launch_plan_ir(pl, a, x, ln, buffer, max_steps, tol, steps, c->stream);
  }
  if (err != hipSuccess) {
    std::fprintf(stderr, "[penta] ir solvedim %d launch FAILED: %s -- output is invalid\n",
                 solvedim, hipGetErrorString(err));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  note_kernel(c, ln.dir, "ir", plan_config(pl));
  return PENTAD_STATUS_SUCCESS;
}

template <typename T>
pentadStatus_t gpsv_batch(Context *c, const T *ds, const T *dl, const T *d,
                          const T *du, const T *dw, T *x, const int *dims,
                          int ndim, int solvedim, void *buffer) {
  // opt-in: the automatic choice of a new (precision, direction, length)
  // tuned on these arrays first (synchronises and allocates a copy of x once
  // per key; the solve below then runs the winner)
  Lines ln;
  if (autotune_env() && ds && dl && d && du && dw && x &&
      decompose(dims, ndim, solvedim, &ln) == 0 && requested_algo(c, ln.dir) == "auto")
    (void)autotune_on_first_use<T>(c, ds, dl, d, du, dw, x, dims, ndim, solvedim);
  return gpsv_batch_pitched<T>(c, ds, dl, d, du, dw, x, dims, nullptr, ndim, solvedim, buffer,
                               SIZE_MAX);
}

template pentadStatus_t gpsv_batch<float>(Context *, const float *, const float *,
                                          const float *, const float *, const float *,
                                          float *, const int *, int, int, void *);
template pentadStatus_t gpsv_batch<double>(Context *, const double *, const double *,
                                           const double *, const double *,
                                           const double *, double *, const int *, int,
                                           int, void *);

// cap: the scratch the caller will pass, when it is fixed (the two-workgroup
// kernel's trade does not fit: the one-workgroup plan's need)
template <typename T>
size_t buffer_size(Context *c, const T *ds, const T *dl, const T *d,
                   const T *du, const T *dw, const T *x, const int *dims,
                   int ndim, int solvedim, const int *pads = nullptr,
                   size_t cap = SIZE_MAX) {
  Lines ln;
  if (decompose_padded(dims, pads, ndim, solvedim, &ln) != 0) return 0;
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  Plan pl = plan_solve(c, ln, (int)sizeof(T), ptrs, true);
  if (pl.split && plan_scratch_bytes(pl, ln, (int)sizeof(T)) > cap)
    pl = plan_solve(c, ln, (int)sizeof(T), ptrs, false);
  return plan_scratch_bytes(pl, ln, (int)sizeof(T));
}

size_t noalloc_need(Context *c, const double *ds, const double *dl, const double *d,
                    const double *du, const double *dw, const double *x, const int *dims,
                    int ndim, int solvedim, size_t cap) {
  return buffer_size<double>(c, ds, dl, d, du, dw, x, dims, ndim, solvedim, nullptr, cap);
}

} // namespace pentad

using namespace pentad;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void pentadsolver_create(pentadsolver_handle_t *handle, void * /*comm*/,
                         int /*ndims*/, const int * /*num_procs*/) {
  if (handle == nullptr) return;
  Context *c = new Context();
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  // set-up of the allocation-free solve on this device (a failure here is
  // reported by pentadsolver_D_gpsv_batch_noalloc itself)
  void *scratch_ptr = nullptr;
  (void)noalloc_scratch(c->device, &scratch_ptr);
  *handle = c;
}

void pentadsolver_destroy(pentadsolver_handle_t *handle) {
  if (handle == nullptr || *handle == nullptr) return;
  Context *c = static_cast<Context *>(*handle);
  destroy_context_resources(c);
  delete c;
  *handle = nullptr;
}

size_t pentadsolver_D_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t h, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, const double *x,
    const int *t_dims, int t_ndim, int t_solvedim) {
  return buffer_size<double>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim,
                             t_solvedim);
}

size_t pentadsolver_S_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t h, const float *ds, const float *dl, const float *d,
    const float *du, const float *dw, const float *x, const int *t_dims,
    int t_ndim, int t_solvedim) {
  return buffer_size<float>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim,
                            t_solvedim);
}

void pentadsolver_D_gpsv_batch(pentadsolver_handle_t h, const double *ds,
                               const double *dl, const double *d,
                               const double *du, const double *dw, double *x,
                               const int *t_dims, int t_ndim, int t_solvedim,
                               void *t_buffer) {
  Context *c = ctx_of(h);
  set_status(c, gpsv_batch<double>(c, ds, dl, d, du, dw, x, t_dims, t_ndim,
                                   t_solvedim, t_buffer));
}

void pentadsolver_S_gpsv_batch(pentadsolver_handle_t h, const float *ds,
                               const float *dl, const float *d,
                               const float *du, const float *dw, float *x,
                               const int *t_dims, int t_ndim, int t_solvedim,
                               void *t_buffer) {
  Context *c = ctx_of(h);
  set_status(c, gpsv_batch<float>(c, ds, dl, d, du, dw, x, t_dims, t_ndim,
                                  t_solvedim, t_buffer));
}

size_t pentadsolver_D_gpsv_batch_padded_buffer_size(
    pentadsolver_handle_t h, const double *ds, const double *dl, const double *d,
    const double *du, const double *dw, const double *x, const int *t_dims,
    const int *t_pads, int t_ndim, int t_solvedim) {
  return buffer_size<double>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim, t_solvedim,
                             t_pads);
}

size_t pentadsolver_S_gpsv_batch_padded_buffer_size(
    pentadsolver_handle_t h, const float *ds, const float *dl, const float *d,
    const float *du, const float *dw, const float *x, const int *t_dims,
    const int *t_pads, int t_ndim, int t_solvedim) {
  return buffer_size<float>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_ndim, t_solvedim,
                            t_pads);
}

pentadStatus_t pentadsolver_D_gpsv_batch_padded(pentadsolver_handle_t h, const double *ds,
                                                const double *dl, const double *d,
                                                const double *du, const double *dw,
                                                double *x, const int *t_dims,
                                                const int *t_pads, int t_ndim,
                                                int t_solvedim, void *t_buffer) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = gpsv_batch_pitched<double>(c, ds, dl, d, du, dw, x, t_dims, t_pads,
                                                      t_ndim, t_solvedim, t_buffer, SIZE_MAX);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_S_gpsv_batch_padded(pentadsolver_handle_t h, const float *ds,
                                                const float *dl, const float *d,
                                                const float *du, const float *dw, float *x,
                                                const int *t_dims, const int *t_pads,
                                                int t_ndim, int t_solvedim, void *t_buffer) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = gpsv_batch_pitched<float>(c, ds, dl, d, du, dw, x, t_dims, t_pads,
                                                     t_ndim, t_solvedim, t_buffer, SIZE_MAX);
  set_status(c, s);
  return s;
}

const char *pentadsolver_kernel_that_ran(int dir) { return ran_global(dir); }

size_t pentadsolver_DS_gpsv_batch_ir_buffer_size_ext(
    pentadsolver_handle_t h, const float *ds, const float *dl, const float *d,
    const float *du, const float *dw, const double *x, const int *t_dims,
    int t_ndim, int t_solvedim) {
  Lines ln;
  if (decompose(t_dims, t_ndim, t_solvedim, &ln) != 0) return 0;
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  return plan_scratch_bytes(plan_ir(ctx_of(h), ln, ptrs, true), ln, 4);
}

void pentadsolver_DS_gpsv_batch_ir(pentadsolver_handle_t h, const float *ds,
                                   const float *dl, const float *d,
                                   const float *du, const float *dw, double *x,
                                   const int *t_dims, int t_ndim,
                                   int t_solvedim, int max_refine, double tol,
                                   int *d_refine_steps, void *t_buffer) {
  Context *c = ctx_of(h);
  set_status(c, gpsv_batch_ir(c, ds, dl, d, du, dw, x, t_dims, t_ndim,
                              t_solvedim, max_refine, tol, d_refine_steps,
                              t_buffer));
}

pentadStatus_t pentadsolver_get_last_status(pentadsolver_handle_t h) {
  return ctx_of(h)->last_status;
}

pentadStatus_t pentadsolver_set_stream(pentadsolver_handle_t h, void *stream) {
  Context *c = ctx_of(h);
  c->stream = static_cast<hipStream_t>(stream);
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_set_algo(pentadsolver_handle_t h, int dir,
                                     const char *algo) {
  if (dir < 0 || dir > 2) return PENTAD_STATUS_INVALID_VALUE;
  Context *c = ctx_of(h);
  std::string s = algo ? algo : "";
  if (s == "legacy") s = "naive";
  if (!s.empty() && s != "auto" && s != "naive" && s != "transpose" &&
      s != "thomas-pcr" && s != "thomas-pcr-split" && s != "shared-fact")
    return PENTAD_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(c->mu);
  c->algo[dir] = s;
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_freeze_env(pentadsolver_handle_t h, int frozen) {
  Context *c = ctx_of(h);
  std::string e[3] = {env_algo(0), env_algo(1), env_algo(2)};
  std::lock_guard<std::mutex> lk(c->mu);
  c->env_frozen = frozen != 0;
  for (int i = 0; i < 3; ++i) c->env_algo[i] = e[i];
  return PENTAD_STATUS_SUCCESS;
}

pentadStatus_t pentadsolver_set_kernel_params(pentadsolver_handle_t h, int dir,
                                              int rows, int width) {
  if (dir < 0 || dir > 2 || rows < 0 || width < 0)
    return PENTAD_STATUS_INVALID_VALUE;
  Context *c = ctx_of(h);
  std::lock_guard<std::mutex> lk(c->mu);
  c->force_rows[dir] = rows;
  c->force_width[dir] = width;
  return PENTAD_STATUS_SUCCESS;
}

const char *pentadsolver_kernel_config(pentadsolver_handle_t h, int dir) {
  if (dir < 0 || dir > 2) return "none";
  Context *c = ctx_of(h);
  std::lock_guard<std::mutex> lk(c->mu);
  return c->ran_cfg[dir].c_str();
}

} // extern "C"

// ===========================================================================
// C++ overloads (reference include/pentadsolver.hpp:75-84, :120-131)
// ===========================================================================
size_t pentadsolver_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t h, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, const double *x,
    const int *t_dims, int t_ndim, int t_solvedim) {
  return pentadsolver_D_gpsv_batch_buffer_size_ext(h, ds, dl, d, du, dw, x,
                                                   t_dims, t_ndim, t_solvedim);
}

size_t pentadsolver_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t h, const float *ds, const float *dl, const float *d,
    const float *du, const float *dw, const float *x, const int *t_dims,
    int t_ndim, int t_solvedim) {
  return pentadsolver_S_gpsv_batch_buffer_size_ext(h, ds, dl, d, du, dw, x,
                                                   t_dims, t_ndim, t_solvedim);
}

void pentadsolver_gpsv_batch(pentadsolver_handle_t h, const double *ds,
                             const double *dl, const double *d,
                             const double *du, const double *dw, double *x,
                             const int *t_dims, int t_ndim, int t_solvedim,
                             void *t_buffer) {
  pentadsolver_D_gpsv_batch(h, ds, dl, d, du, dw, x, t_dims, t_ndim,
                            t_solvedim, t_buffer);
}

void pentadsolver_gpsv_batch(pentadsolver_handle_t h, const float *ds,
                             const float *dl, const float *d, const float *du,
                             const float *dw, float *x, const int *t_dims,
                             int t_ndim, int t_solvedim, void *t_buffer) {
  pentadsolver_S_gpsv_batch(h, ds, dl, d, du, dw, x, t_dims, t_ndim,
                            t_solvedim, t_buffer);
}
