// C ABI of the FIT time-stepping library (declarations: include/wakis_fit.h).
//
// One context = one solve on one HIP device with one HIP stream (default
// priority).  The context owns E, H, J (3N each), the per-cell material ids,
// the material table, the metric tables and the wake-probe history.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wakis_fit.h"
#include "fit_aux.hpp"
#include "fit_fused.hpp"
#include "fit_fused2.hpp"
#include "fit_sweep.hpp"
#include "fit_split.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(FIT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

}  // namespace

struct fit_ctx {
  int nx = 0, ny = 0, nz = 0;
  int64_t N = 0;
  int dtype = FIT_F64;
  size_t esz = 8;
  int device = 0;
  hipStream_t stream = nullptr;
  void *E = nullptr, *H = nullptr, *J = nullptr;
  void *E_alt = nullptr, *H_alt = nullptr;  // second buffers of the fused sweep
  uint8_t *mat = nullptr;
  bool mat_uniform = true;
  // wide material ids (fit_set_materials_wide): one int32 per cell, the table
  // in global memory; stepped by the split kernels only
  int32_t *mat32 = nullptr;
  bool wide = false;
  void *mtab = nullptr;
  int nmat = 0;
  std::vector<double> mtab_host;  // nmat * 9 FP64 (1/eps, 1/mu, sigma)
  void *geo = nullptr;
  int64_t geo_count = 0;
  bool geo_set = false, mat_set = false, dt_set = false;
  uint32_t pec = 0, pmc = 0;
  double dt = 0;
  int jmode = FIT_J_ZERO;
  int beam_mode = FIT_BEAM_OFF, ixs = 0, iys = 0;
  void *jline = nullptr;
  double *zc = nullptr;
  double bp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t step = 0;
  void *hist = nullptr;
  int64_t hist_nt = 0;
  int ixt = 0, iyt = 0, hw = 0;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  int kernel = FIT_KERNEL_AUTO;
  bool chunk_rsrc = false;  // fit_set_chunk_rsrc
  int fused_cfg = -1;  // tile configuration of the fused sweep (-1: default)
  int f2_cfg = -1;     // tile shape of the two-step sweep (fit_set_tile2; -1: default)
  int f2_def = -1;     // the modelled default shape, computed for f2_def_nmat rows
  int f2_def_nmat = -1;
  void *btab = nullptr;  // beam line table of the steps of one fit_step call
  int *flag = nullptr;   // watchdog flag (fit_field_finite), 256 bytes
  int64_t btab_steps = 0;
  // field monitor: complex128 accumulators (3 * mon_nf * N) and the staged
  // phases (mon_ph_steps rows of mon_nf complex values) of the next fit_step
  void *mon = nullptr;
  int mon_nf = 0;
  double *mon_ph = nullptr;
  int64_t mon_ph_cap = 0;  // rows
  int64_t mon_ph_steps = 0;
  // launches of the step kernels so far: [0] one-step (fused or split pair),
  // [1] two-step fused sweeps
  int64_t launches[2] = {0, 0};
  // device bytes (fit_device_bytes): sizes of the buffers above that are not
  // fixed by the grid, and the high-water mark of everything the context has
  // held at once, the scratch of fit_wake_potential / fit_impedance included
  size_t mtab_bytes = 0, hist_bytes = 0, mon_bytes = 0, mon_ph_bytes = 0;
  int64_t peak_bytes = 0;
};

namespace {

int set_device(fit_ctx *c) {
  HIPCHK(hipSetDevice(c->device));
  return 0;
}

// Device bytes the context holds now.
int64_t resident_bytes(const fit_ctx *c) {
  const int64_t fb = 3 * c->N * (int64_t)c->esz;
  int64_t b = 0;
  for (const void *p : {c->E, c->H, c->J, c->E_alt, c->H_alt})
    if (p) b += fb;
  if (c->mat) b += c->N;
  if (c->mat32) b += 4 * c->N;
  if (c->geo) b += c->geo_count * (int64_t)c->esz;
  if (c->jline) b += (int64_t)c->nz * c->esz;
  if (c->zc) b += (int64_t)c->nz * 8;
  if (c->btab) b += c->btab_steps * c->nz * (int64_t)c->esz;
  if (c->flag) b += 256;
  b += (int64_t)(c->mtab_bytes + c->hist_bytes + c->staging_bytes + c->mon_bytes + c->mon_ph_bytes);
  return b;
}

// Raise the high-water mark to the resident bytes plus `extra` held for the
// duration of a call.
void note_peak(fit_ctx *c, int64_t extra = 0) {
  c->peak_bytes = std::max(c->peak_bytes, resident_bytes(c) + extra);
}

int ensure_staging(fit_ctx *c, size_t bytes) {
  if (c->staging_bytes >= bytes) return 0;
  if (c->staging) HIPCHK(hipFree(c->staging));
  c->staging = nullptr;
  c->staging_bytes = 0;
  if (hipMalloc(&c->staging, bytes) != hipSuccess)
    return fail(FIT_ENOMEM, "staging allocation failed");
  c->staging_bytes = bytes;
  note_peak(c);
  return 0;
}

// Material table in the compute type: dt/mu, dt/eps, sigma.
int upload_mtab(fit_ctx *c) {
  if (!c->mat_set) return 0;
  const int nm = c->nmat;
  std::vector<double> t((size_t)nm * fit::MAT_STRIDE);
  for (size_t m = 0; m < (size_t)nm; ++m) {
    const double *r = &c->mtab_host[m * 9];
    for (int d = 0; d < 3; ++d) {
      t[m * 9 + d] = c->dt * r[3 + d];      // dt/mu
      t[m * 9 + 3 + d] = c->dt * r[d];      // dt/eps
      t[m * 9 + 6 + d] = r[6 + d];          // sigma
    }
  }
  if (c->mtab) HIPCHK(hipFree(c->mtab));
  c->mtab = nullptr;
  c->mtab_bytes = 0;
  if (hipMalloc(&c->mtab, t.size() * c->esz) != hipSuccess)
    return fail(FIT_ENOMEM, "material table allocation failed");
  c->mtab_bytes = t.size() * c->esz;
  note_peak(c);
  if (c->dtype == FIT_F64) {
    HIPCHK(hipMemcpy(c->mtab, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(t.begin(), t.end());
    HIPCHK(hipMemcpy(c->mtab, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return 0;
}

template <typename T>
fit::StepArgs<T> make_args(fit_ctx *c) {
  fit::StepArgs<T> a{};
  a.E = (T *)c->E;
  a.H = (T *)c->H;
  a.J = (const T *)c->J;
  a.mat = (c->mat_uniform || c->wide) ? nullptr : c->mat;
  a.mat32 = c->wide ? c->mat32 : nullptr;
  a.mtab = (const T *)c->mtab;
  a.nmat = c->nmat;
  {
    const int64_t sz = (int64_t)sizeof(T);
    const int64_t geo = 2 * ((int64_t)c->nx + c->ny + c->nz) +
                        2 * ((int64_t)c->ny * c->nz + (int64_t)c->nx * c->nz + (int64_t)c->nx * c->ny);
    a.geo_bytes = geo * sz;
    a.buf32 = 3 * c->N * sz < ((int64_t)1 << 32) - 1 && a.geo_bytes < ((int64_t)1 << 32) - 1;
    // non-temporal field stores when a step moves far more than the 256 MB
    // last-level cache (measured: +2.6 % on a 2.4 GB step; -9 % on batched
    // serving of mid-size grids, which keeps ordinary stores)
    a.nt = 12 * c->N * sz >= ((int64_t)1 << 30);
    a.pcr = c->chunk_rsrc;
  }
  const T *g = (const T *)c->geo;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  a.g.Lx = g;
  a.g.Ly = g + nx;
  a.g.Lz = g + nx + ny;
  const T *t = g + nx + ny + nz;
  a.g.tLx = t;
  a.g.tLy = t + nx;
  a.g.tLz = t + nx + ny;
  const T *ia = t + nx + ny + nz;
  a.g.iAx = ia;
  a.g.iAy = ia + (int64_t)ny * nz;
  a.g.iAz = ia + (int64_t)ny * nz + (int64_t)nx * nz;
  const T *ita = ia + (int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny;
  a.g.itAx = ita;
  a.g.itAy = ita + (int64_t)ny * nz;
  a.g.itAz = ita + (int64_t)ny * nz + (int64_t)nx * nz;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.N = c->N;
  a.pec = c->pec;
  a.pmc = c->pmc;
  a.jmode = c->jmode;
  a.beam_mode = c->beam_mode;
  a.ixs = c->ixs;
  a.iys = c->iys;
  a.jline = (const T *)c->jline;
  a.zc = c->zc;
  a.b_qv = c->bp[0];
  a.b_A = c->bp[1];
  a.b_B = c->bp[2];
  a.b_s0 = c->bp[3];
  a.b_v = c->bp[4];
  a.b_dx = c->bp[5];
  a.b_dy = c->bp[6];
  a.b_f32time = c->bp[7] != 0.0;
  a.dt64 = c->dt;
  a.dt32 = (float)c->dt;
  a.step = c->step;
  a.hist = (T *)c->hist;
  a.hist_nt = c->hist_nt;
  a.ixt = c->ixt;
  a.iyt = c->iyt;
  a.hw = c->hw;
  return a;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// The fused sweep writes new buffers; allocate them on first use.  With
// FIT_KERNEL_AUTO a failed allocation selects the split kernels instead.
int want_fused(fit_ctx *c, bool *fused) {
  *fused = false;
  if (c->wide) {
    // the fused sweeps keep the material table in LDS (at most 256 rows)
    if (c->kernel == FIT_KERNEL_FUSED || c->kernel == FIT_KERNEL_FUSED2)
      return fail(FIT_ESTATE, "more than 256 material rows: the fused sweeps do not apply "
                              "(use FIT_KERNEL_AUTO or FIT_KERNEL_SPLIT)");
    return 0;
  }
  if (!fit::fused_supported(c->kernel, c->nx, c->ny, c->nz)) return 0;
  if (!c->E_alt || !c->H_alt) {
    const size_t fb = (size_t)3 * c->N * c->esz;
    if (!c->E_alt && hipMalloc(&c->E_alt, fb) != hipSuccess) c->E_alt = nullptr;
    if (!c->H_alt && hipMalloc(&c->H_alt, fb) != hipSuccess) c->H_alt = nullptr;
    if (!c->E_alt || !c->H_alt) {
      (void)hipGetLastError();
      if (c->kernel == FIT_KERNEL_FUSED)
        return fail(FIT_ENOMEM, "fused sweep: second field buffers do not fit");
      note_peak(c);
      return 0;
    }
    note_peak(c);
  }
  *fused = true;
  return 0;
}

// One field-monitor update of the current E with nf phases at `phase`
// (device memory).
template <typename T>
void monitor_launch(fit_ctx *c, const double *phase) {
  hipLaunchKernelGGL(fit::monitor_accumulate<T>, dim3(nblocks(c->N, 256)), dim3(256), 0,
                     c->stream, (const T *)c->E, c->N, c->mon_nf, phase, (double2 *)c->mon);
}

// Tile shape of the two-step sweep for these step arguments: the one set by
// fit_set_tile2 when it applies, else the default shape (-1: none applies).
template <typename T>
int f2_cfg_of(fit_ctx *c, const fit::StepArgs<T> &a) {
  if (c->f2_cfg >= 0 && fit::fused2_supported<T>(a, c->f2_cfg)) return c->f2_cfg;
  if (c->f2_def >= 0 && c->f2_def_nmat == c->nmat && fit::fused2_supported<T>(a, c->f2_def))
    return c->f2_def;
  const int d = fit::fused2_default_cfg<T>(a);
  if (d >= 0) {
    c->f2_def = d;
    c->f2_def_nmat = c->nmat;
  }
  return d;
}

// ev (optional): 3 events per step [before first kernel, between, after].
template <typename T>
int step_impl(fit_ctx *c, int nsteps, int conductive, hipEvent_t *ev = nullptr) {
  bool fused = false;
  if (int rc = want_fused(c, &fused)) return rc;
  // the Gaussian bunch of all steps of this call, evaluated once
  const bool gauss = c->beam_mode == FIT_BEAM_GAUSS;
  if (gauss && nsteps > 0) {
    if (c->btab_steps < nsteps) {
      if (c->btab) HIPCHK(hipFree(c->btab));
      c->btab = nullptr;
      c->btab_steps = 0;
      if (hipMalloc(&c->btab, (size_t)nsteps * c->nz * c->esz) != hipSuccess)
        return fail(FIT_ENOMEM, "beam table allocation failed");
      c->btab_steps = nsteps;
      note_peak(c);
    }
    fit::StepArgs<T> a = make_args<T>(c);
    const int64_t cnt = (int64_t)nsteps * c->nz;
    hipLaunchKernelGGL(fit::beam_table<T>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream, a,
                       c->step, nsteps, (T *)c->btab);
    HIPCHK(hipGetLastError());
  }
  // two-step sweeps for pairs of steps where they apply: the fused kernel
  // selected (or AUTO), FP32 full tiles, J zero or sigma*E, and no field
  // monitor update staged between the two steps
  const bool pairs = fused && (c->kernel == FIT_KERNEL_AUTO || c->kernel == FIT_KERNEL_FUSED2);
  for (int s = 0; s < nsteps;) {
    fit::StepArgs<T> a = make_args<T>(c);
    const int f2 = pairs && s + 1 < nsteps && !(c->mon && s < c->mon_ph_steps)
                       ? f2_cfg_of<T>(c, a)
                       : -1;
    if (f2 >= 0) {
      const T *jl = (const T *)c->jline;
      const T *beam1 = nullptr, *beam2 = nullptr;
      if (gauss) {
        beam1 = (const T *)c->btab + (int64_t)s * c->nz;
        beam2 = beam1 + c->nz;
      } else if (c->beam_mode == FIT_BEAM_VALUES) {
        beam1 = jl;
        beam2 = conductive ? nullptr : jl;  // a conductive step consumes the line
      }
      const bool js1 = c->jmode == FIT_J_SIGMA_E;
      const bool js2 = conductive ? true : js1;
      if (ev) HIPCHK(hipEventRecord(ev[3 * s], c->stream));
      if (!fit::launch_fused2<T>(a, f2, (T *)c->E_alt, (T *)c->H_alt, beam1, beam2, js1, js2,
                                 c->stream))
        return fail(FIT_ESTATE, "two-step sweep not launched");
      std::swap(c->E, c->E_alt);
      std::swap(c->H, c->H_alt);
      if (ev) {
        HIPCHK(hipEventRecord(ev[3 * s + 1], c->stream));
        for (int q = 3 * s + 2; q < 3 * s + 6; ++q) HIPCHK(hipEventRecord(ev[q], c->stream));
      }
      HIPCHK(hipGetLastError());
      c->launches[1] += 1;
      c->step += 2;
      if (conductive) {
        c->jmode = FIT_J_SIGMA_E;
        if (c->beam_mode == FIT_BEAM_VALUES) c->beam_mode = FIT_BEAM_OFF;
      }
      s += 2;
      continue;
    }
    if (gauss) {
      a.beam_mode = FIT_BEAM_VALUES;
      a.jline = (const T *)c->btab + (int64_t)s * c->nz;
    }
    if (ev) HIPCHK(hipEventRecord(ev[3 * s], c->stream));
    const bool mon_step = c->mon && s < c->mon_ph_steps;
    bool mon_fused = false;
    if (fused) {
      const int cfg = c->fused_cfg >= 0 ? c->fused_cfg : fit::sweep_default(c->nx, c->ny, c->nz, c->dtype == FIT_F64);
      // the field monitor update of this step runs inside the sweep (its
      // general tile path) when the sweep is a fused_sweep tile
      if (mon_step && cfg < fit::FUSED_NCFG) {
        a.mon = (double2 *)c->mon;
        a.mph = c->mon_ph + (int64_t)s * 2 * c->mon_nf;
        a.mnf = c->mon_nf;
        mon_fused = true;
      }
      if (!fit::launch_sweep<T>(a, (T *)c->E_alt, (T *)c->H_alt, c->stream, cfg))
        return fail(FIT_ESTATE, "fused sweep tile configuration " + std::to_string(cfg) +
                                    " spills to scratch memory in this build; not launched");
      std::swap(c->E, c->E_alt);
      std::swap(c->H, c->H_alt);
      if (ev) HIPCHK(hipEventRecord(ev[3 * s + 1], c->stream));
    } else {
      const dim3 grid(nblocks(c->nx, fit::SPLIT_BX), nblocks(c->ny, fit::SPLIT_BY), c->nz);
      const dim3 block(fit::SPLIT_BX, fit::SPLIT_BY);
      hipLaunchKernelGGL(fit::h_update_split<T>, grid, block, 0, c->stream, a);
      if (ev) HIPCHK(hipEventRecord(ev[3 * s + 1], c->stream));
      hipLaunchKernelGGL(fit::e_update_split<T>, grid, block, 0, c->stream, a);
    }
    if (ev) HIPCHK(hipEventRecord(ev[3 * s + 2], c->stream));
    if (mon_step && !mon_fused) monitor_launch<T>(c, c->mon_ph + (int64_t)s * 2 * c->mon_nf);
    HIPCHK(hipGetLastError());
    c->launches[0] += 1;
    c->step += 1;
    if (conductive) {
      c->jmode = FIT_J_SIGMA_E;
      if (c->beam_mode == FIT_BEAM_VALUES) c->beam_mode = FIT_BEAM_OFF;
    }
    s += 1;
  }
  c->mon_ph_steps = 0;  // the staged monitor phases belong to this call
  return 0;
}

int check_ready(fit_ctx *c) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (!c->geo_set) return fail(FIT_ESTATE, "geometry not set");
  if (!c->mat_set) return fail(FIT_ESTATE, "materials not set");
  if (!c->dt_set) return fail(FIT_ESTATE, "dt not set");
  return 0;
}

void *field_ptr(fit_ctx *c, int which) {
  return which == FIT_FIELD_E ? c->E : (which == FIT_FIELD_H ? c->H : c->J);
}

int check_box(fit_ctx *c, int which, int comp, const int32_t *b) {
  if (which < 0 || which > 2) return fail(FIT_EINVAL, "bad field id");
  if (comp < 0 || comp > 2) return fail(FIT_EINVAL, "bad component");
  if (!b) return fail(FIT_EINVAL, "null box");
  if (b[0] < 0 || b[1] > c->nx || b[0] >= b[1] || b[2] < 0 || b[3] > c->ny ||
      b[2] >= b[3] || b[4] < 0 || b[5] > c->nz || b[4] >= b[5])
    return fail(FIT_EINVAL, "box out of range or empty");
  return 0;
}

template <typename T>
int tune_impl(fit_ctx *c, int reps, double *ms_out) {
  // The sweep reads E, H and writes the second buffers, so timing launches
  // without swapping leaves the solver state unchanged.
  fit::StepArgs<T> a = make_args<T>(c);
  a.hist = nullptr;  // no probe writes from tuning launches
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int best = -1;
  double best_ms = 1e300;
  int rc = 0;
  if (a.beam_mode == FIT_BEAM_GAUSS) a.beam_mode = FIT_BEAM_OFF;  // timing only
  for (int cfg = 0; cfg < fit::SWEEP_NCFG && !rc; ++cfg) {
    if (!fit::sweep_cfg_ok<T>(cfg) || !fit::sweep_cfg_runs<T>(a, cfg)) {
      if (ms_out) ms_out[cfg] = -1.0;
      continue;
    }
    fit::launch_sweep<T>(a, (T *)c->E_alt, (T *)c->H_alt, c->stream, cfg);
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(FIT_EHIP, "event");
    for (int r = 0; r < reps && !rc; ++r)
      fit::launch_sweep<T>(a, (T *)c->E_alt, (T *)c->H_alt, c->stream, cfg);
    float t = 0;
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = fail(FIT_EHIP, "tuning timing");
    if (hipGetLastError() != hipSuccess && !rc) rc = fail(FIT_EHIP, "tuning launch");
    const double ms = (double)t / reps;
    if (ms_out) ms_out[cfg] = ms;
    if (!rc && ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) c->fused_cfg = best;
  return rc;
}

// Times each tile shape of the two-step sweep (reps launches, written into
// the second buffers without swapping, no probe and no beam: the state is
// unchanged) and keeps the fastest.
template <typename T>
int tune2_impl(fit_ctx *c, int reps, double *ms_out) {
  fit::StepArgs<T> a = make_args<T>(c);
  a.hist = nullptr;
  const bool js = c->jmode == FIT_J_SIGMA_E;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int best = -1;
  double best_ms = 1e300;
  int rc = 0;
  for (int cfg = 0; cfg < fit::F2_NCFG && !rc; ++cfg) {
    if (!fit::fused2_supported<T>(a, cfg)) {
      if (ms_out) ms_out[cfg] = -1.0;
      continue;
    }
    fit::launch_fused2<T>(a, cfg, (T *)c->E_alt, (T *)c->H_alt, nullptr, nullptr, js, js,
                          c->stream);
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(FIT_EHIP, "event");
    for (int r = 0; r < reps && !rc; ++r)
      fit::launch_fused2<T>(a, cfg, (T *)c->E_alt, (T *)c->H_alt, nullptr, nullptr, js, js,
                            c->stream);
    float t = 0;
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = fail(FIT_EHIP, "tuning timing");
    if (hipGetLastError() != hipSuccess && !rc) rc = fail(FIT_EHIP, "tuning launch");
    const double ms = (double)t / reps;
    if (ms_out) ms_out[cfg] = ms;
    if (!rc && ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc && best >= 0) c->f2_cfg = best;
  return rc;
}

}  // namespace

// ---------------- allocation-free impedance DFT ----------------
//
// fit_impedance_noalloc stages its inputs and outputs in this file-scope
// device array instead of hipMalloc'd buffers: x (NOALLOC_MAX_LEN doubles),
// lam (NOALLOC_MAX_LEN doubles), then the six output arrays of
// NOALLOC_MAX_BINS doubles each.  Every device holds its own copy for the
// life of the process.  Host code never passes `scratch` itself to a kernel
// or a copy (in host code the name is only a stand-in): it uses the device
// address that hipGetSymbolAddress returns, resolved once per device with
// that device current.
constexpr int64_t NOALLOC_MAX_LEN = FIT_NOALLOC_MAX_LEN;
constexpr int64_t NOALLOC_MAX_BINS = FIT_NOALLOC_MAX_BINS;
constexpr int64_t NOALLOC_WORDS = 2 * NOALLOC_MAX_LEN + 6 * NOALLOC_MAX_BINS;
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on device `dev` (must be current); resolved on
// the first call per device.
int noalloc_scratch(int dev, double **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(FIT_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FIT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

const char *fit_last_error(void) { return g_err.c_str(); }
const char *fit_version(void) { return "wakis_amd-fit 0.1 gfx950"; }

int fit_create(const fit_desc *d, fit_ctx **out) {
  if (!d || !out) return fail(FIT_EINVAL, "null argument");
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nz < 2) return fail(FIT_EINVAL, "need nx, ny, nz >= 2");
  if (d->dtype != FIT_F64 && d->dtype != FIT_F32) return fail(FIT_EINVAL, "bad dtype");
  const int64_t N = (int64_t)d->nx * d->ny * d->nz;
  if (N >= (int64_t)1 << 31) return fail(FIT_EINVAL, "grid must have fewer than 2^31 cells");
  if (d->nz > 65535) return fail(FIT_EINVAL, "nz must be <= 65535");
  fit_ctx *c = new fit_ctx();
  c->nx = d->nx;
  c->ny = d->ny;
  c->nz = d->nz;
  c->N = N;
  c->dtype = d->dtype;
  c->esz = d->dtype == FIT_F64 ? 8 : 4;
  c->device = d->device;
  int rc = set_device(c);
  if (rc) {
    delete c;
    return rc;
  }
  {
    // set-up of this device's allocation-free scratch (fit_impedance_noalloc)
    double *sp = nullptr;
    if ((rc = noalloc_scratch(c->device, &sp))) {
      delete c;
      return rc;
    }
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(FIT_EHIP, "stream creation failed");
  }
  const size_t fb = (size_t)3 * N * c->esz;
  if (hipMalloc(&c->E, fb) != hipSuccess || hipMalloc(&c->H, fb) != hipSuccess ||
      hipMalloc(&c->J, fb) != hipSuccess || hipMalloc((void **)&c->mat, (size_t)N) != hipSuccess ||
      hipMalloc(&c->jline, (size_t)c->nz * c->esz) != hipSuccess ||
      hipMalloc((void **)&c->zc, (size_t)c->nz * 8) != hipSuccess) {
    fit_destroy(c);
    return fail(FIT_ENOMEM, "field allocation failed (" + std::to_string(3 * fb) + " bytes)");
  }
  if (hipMemsetAsync(c->E, 0, fb, c->stream) != hipSuccess ||
      hipMemsetAsync(c->H, 0, fb, c->stream) != hipSuccess ||
      hipMemsetAsync(c->J, 0, fb, c->stream) != hipSuccess ||
      hipMemsetAsync(c->mat, 0, (size_t)N, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    fit_destroy(c);
    return fail(FIT_EHIP, "field initialisation failed");
  }
  note_peak(c);
  *out = c;
  return 0;
}

int fit_destroy(fit_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *ps[] = {c->E,   c->H,     c->J,  c->E_alt, c->H_alt,   c->mat,  c->mtab,
                c->geo, c->jline, c->zc, c->hist,  c->staging, c->btab, c->mon,
                c->mon_ph, c->mat32, c->flag};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int fit_set_geometry(fit_ctx *c, const double *packed, int64_t count) {
  if (!c || !packed) return fail(FIT_EINVAL, "null argument");
  const int64_t nx = c->nx, ny = c->ny, nz = c->nz;
  const int64_t want = 2 * (nx + ny + nz) + 2 * (ny * nz + nx * nz + nx * ny);
  if (count != want)
    return fail(FIT_EINVAL, "geometry count " + std::to_string(count) + " != " + std::to_string(want));
  if (int rc = set_device(c)) return rc;
  if (c->geo) HIPCHK(hipFree(c->geo));
  c->geo = nullptr;
  c->geo_count = 0;
  if (hipMalloc(&c->geo, (size_t)count * c->esz) != hipSuccess)
    return fail(FIT_ENOMEM, "geometry allocation failed");
  c->geo_count = count;
  if (c->dtype == FIT_F64) {
    HIPCHK(hipMemcpy(c->geo, packed, (size_t)count * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(packed, packed + count);
    HIPCHK(hipMemcpy(c->geo, f.data(), (size_t)count * 4, hipMemcpyHostToDevice));
  }
  c->geo_set = true;
  note_peak(c);
  return 0;
}

int fit_set_materials(fit_ctx *c, const uint8_t *ids, int32_t nmat, const double *table) {
  if (!c || !table) return fail(FIT_EINVAL, "null argument");
  if (nmat < 1 || nmat > 256) return fail(FIT_EINVAL, "nmat must be in [1, 256]");
  if (int rc = set_device(c)) return rc;
  if (ids) {
    for (int64_t n = 0; n < c->N; ++n)
      if (ids[n] >= nmat) return fail(FIT_EINVAL, "material id out of range");
  }
  c->mtab_host.assign(table, table + (size_t)nmat * 9);
  c->nmat = nmat;
  if (ids) {
    HIPCHK(hipMemcpy(c->mat, ids, (size_t)c->N, hipMemcpyHostToDevice));
    c->mat_uniform = false;
  } else {
    HIPCHK(hipMemset(c->mat, 0, (size_t)c->N));
    c->mat_uniform = true;
  }
  if (c->mat32) HIPCHK(hipFree(c->mat32));
  c->mat32 = nullptr;
  c->wide = false;
  c->mat_set = true;
  return upload_mtab(c);
}

int fit_set_materials_wide(fit_ctx *c, const int32_t *ids, int32_t nmat, const double *table) {
  if (!c || !ids || !table) return fail(FIT_EINVAL, "null argument");
  if (nmat < 1 || (int64_t)nmat > c->N) return fail(FIT_EINVAL, "nmat must be in [1, N]");
  for (int64_t n = 0; n < c->N; ++n)
    if (ids[n] < 0 || ids[n] >= nmat) return fail(FIT_EINVAL, "material id out of range");
  if (int rc = set_device(c)) return rc;
  if (!c->mat32) {
    if (hipMalloc((void **)&c->mat32, (size_t)c->N * 4) != hipSuccess) {
      c->mat32 = nullptr;
      return fail(FIT_ENOMEM, "material id allocation failed");
    }
  }
  HIPCHK(hipMemcpy(c->mat32, ids, (size_t)c->N * 4, hipMemcpyHostToDevice));
  c->mtab_host.assign(table, table + (size_t)nmat * 9);
  c->nmat = nmat;
  c->wide = true;
  c->mat_uniform = false;
  c->mat_set = true;
  note_peak(c);
  return upload_mtab(c);
}

int fit_set_bc(fit_ctx *c, uint32_t pec_cols, uint32_t pmc_rows) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if ((pec_cols | pmc_rows) & ~63u) return fail(FIT_EINVAL, "face mask uses bits above 5");
  c->pec = pec_cols;
  c->pmc = pmc_rows;
  return 0;
}

int fit_set_dt(fit_ctx *c, double dt) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (!(dt > 0)) return fail(FIT_EINVAL, "dt must be > 0");
  if (int rc = set_device(c)) return rc;
  c->dt = dt;
  c->dt_set = true;
  return upload_mtab(c);
}

int fit_set_jmode(fit_ctx *c, int32_t jmode) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (jmode < FIT_J_ZERO || jmode > FIT_J_BUFFER) return fail(FIT_EINVAL, "bad J mode");
  c->jmode = jmode;
  return 0;
}

int fit_get_state(fit_ctx *c, int32_t *jmode, int32_t *beam_mode) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (jmode) *jmode = c->jmode;
  if (beam_mode) *beam_mode = c->beam_mode;
  return 0;
}

int fit_set_beam(fit_ctx *c, int32_t mode, int32_t ixs, int32_t iys, const double *values,
                 const double *params, const double *zc) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (mode == FIT_BEAM_OFF) {
    c->beam_mode = mode;
    return 0;
  }
  if (ixs < 0 || ixs >= c->nx || iys < 0 || iys >= c->ny)
    return fail(FIT_EINVAL, "beam line outside the grid");
  if (int rc = set_device(c)) return rc;
  if (mode == FIT_BEAM_VALUES) {
    if (!values) return fail(FIT_EINVAL, "null beam values");
    if (c->dtype == FIT_F64) {
      HIPCHK(hipMemcpyAsync(c->jline, values, (size_t)c->nz * 8, hipMemcpyHostToDevice, c->stream));
    } else {
      std::vector<float> f(values, values + c->nz);
      HIPCHK(hipMemcpyAsync(c->jline, f.data(), (size_t)c->nz * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  } else if (mode == FIT_BEAM_GAUSS) {
    if (!params || !zc) return fail(FIT_EINVAL, "null beam parameters");
    std::memcpy(c->bp, params, sizeof(c->bp));
    HIPCHK(hipMemcpyAsync(c->zc, zc, (size_t)c->nz * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    return fail(FIT_EINVAL, "bad beam mode");
  }
  c->beam_mode = mode;
  c->ixs = ixs;
  c->iys = iys;
  return 0;
}

int fit_field_read(fit_ctx *c, int32_t which, int32_t comp, const int32_t box[6], void *host) {
  if (!c || !host) return fail(FIT_EINVAL, "null argument");
  if (int rc = check_box(c, which, comp, box)) return rc;
  if (int rc = set_device(c)) return rc;
  const int bi = box[1] - box[0], bj = box[3] - box[2], bk = box[5] - box[4];
  const int64_t cnt = (int64_t)bi * bj * bk;
  char *src = (char *)field_ptr(c, which) + (size_t)comp * c->N * c->esz;
  if (bi == c->nx && bj == c->ny) {
    // whole planes: contiguous
    HIPCHK(hipMemcpyAsync(host, src + (size_t)box[4] * c->nx * c->ny * c->esz, (size_t)cnt * c->esz,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  if (int rc = ensure_staging(c, (size_t)cnt * c->esz)) return rc;
  if (c->dtype == FIT_F64)
    hipLaunchKernelGGL(fit::box_pack<double>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (const double *)src, (double *)c->staging, c->nx, c->ny, box[0], bi, box[2],
                       bj, box[4], cnt);
  else
    hipLaunchKernelGGL(fit::box_pack<float>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (const float *)src, (float *)c->staging, c->nx, c->ny, box[0], bi, box[2], bj,
                       box[4], cnt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, c->staging, (size_t)cnt * c->esz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_field_write(fit_ctx *c, int32_t which, int32_t comp, const int32_t box[6],
                    const void *host) {
  if (!c || !host) return fail(FIT_EINVAL, "null argument");
  if (int rc = check_box(c, which, comp, box)) return rc;
  if (int rc = set_device(c)) return rc;
  const int bi = box[1] - box[0], bj = box[3] - box[2], bk = box[5] - box[4];
  const int64_t cnt = (int64_t)bi * bj * bk;
  char *dst = (char *)field_ptr(c, which) + (size_t)comp * c->N * c->esz;
  if (bi == c->nx && bj == c->ny) {
    HIPCHK(hipMemcpyAsync(dst + (size_t)box[4] * c->nx * c->ny * c->esz, host, (size_t)cnt * c->esz,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  if (int rc = ensure_staging(c, (size_t)cnt * c->esz)) return rc;
  HIPCHK(hipMemcpyAsync(c->staging, host, (size_t)cnt * c->esz, hipMemcpyHostToDevice, c->stream));
  if (c->dtype == FIT_F64)
    hipLaunchKernelGGL(fit::box_unpack<double>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (const double *)c->staging, (double *)dst, c->nx, c->ny, box[0], bi, box[2],
                       bj, box[4], cnt);
  else
    hipLaunchKernelGGL(fit::box_unpack<float>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (const float *)c->staging, (float *)dst, c->nx, c->ny, box[0], bi, box[2], bj,
                       box[4], cnt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_field_zero(fit_ctx *c, int32_t which) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (which < 0 || which > 2) return fail(FIT_EINVAL, "bad field id");
  if (int rc = set_device(c)) return rc;
  const int64_t cnt = 3 * c->N;
  if (c->dtype == FIT_F64)
    hipLaunchKernelGGL(fit::fill_zero<double>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (double *)field_ptr(c, which), cnt);
  else
    hipLaunchKernelGGL(fit::fill_zero<float>, dim3(nblocks(cnt, 256)), dim3(256), 0, c->stream,
                       (float *)field_ptr(c, which), cnt);
  HIPCHK(hipGetLastError());
  return 0;
}

int fit_field_finite(fit_ctx *c, int32_t which, int32_t *finite) {
  if (!c || !finite) return fail(FIT_EINVAL, "null argument");
  if (which < 0 || which > 2) return fail(FIT_EINVAL, "bad field id");
  if (int rc = set_device(c)) return rc;
  if (!c->flag) {
    if (hipMalloc((void **)&c->flag, 256) != hipSuccess) {
      c->flag = nullptr;
      (void)hipGetLastError();
      return fail(FIT_ENOMEM, "watchdog flag allocation failed");
    }
    note_peak(c);
  }
  if (which == FIT_FIELD_J && c->jmode != FIT_J_BUFFER) {
    // J is implicit (0 or sigma*E): it is finite whenever E is
    which = FIT_FIELD_E;
  }
  HIPCHK(hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
  const int64_t cnt = 3 * c->N;
  const int nb = (int)std::min<int64_t>(nblocks(cnt, 256), 4096);
  if (c->dtype == FIT_F64)
    hipLaunchKernelGGL(fit::nonfinite_flag<double>, dim3(nb), dim3(256), 0, c->stream,
                       (const double *)field_ptr(c, which), cnt, c->flag);
  else
    hipLaunchKernelGGL(fit::nonfinite_flag<float>, dim3(nb), dim3(256), 0, c->stream,
                       (const float *)field_ptr(c, which), cnt, c->flag);
  HIPCHK(hipGetLastError());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *finite = h ? 0 : 1;
  return 0;
}

int fit_materialize_j(fit_ctx *c) {
  if (int rc = check_ready(c)) return rc;
  if (int rc = set_device(c)) return rc;
  if (c->jmode == FIT_J_BUFFER) {
    // buffer already holds J; only the beam line may need applying
    if (c->beam_mode == FIT_BEAM_OFF) return 0;
  }
  if (c->dtype == FIT_F64) {
    auto a = make_args<double>(c);
    if (c->jmode == FIT_J_BUFFER) a.jmode = -1;
    hipLaunchKernelGGL(fit::materialize_j<double>, dim3(nblocks(c->N, 256)), dim3(256), 0,
                       c->stream, a);
  } else {
    auto a = make_args<float>(c);
    if (c->jmode == FIT_J_BUFFER) a.jmode = -1;
    hipLaunchKernelGGL(fit::materialize_j<float>, dim3(nblocks(c->N, 256)), dim3(256), 0,
                       c->stream, a);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_zero_ghosts(fit_ctx *c) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (int rc = set_device(c)) return rc;
  if (c->dtype == FIT_F64)
    hipLaunchKernelGGL(fit::zero_ghosts<double>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream,
                       (double *)c->E, (double *)c->H, c->nx, c->ny, c->nz, c->N);
  else
    hipLaunchKernelGGL(fit::zero_ghosts<float>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream,
                       (float *)c->E, (float *)c->H, c->nx, c->ny, c->nz, c->N);
  HIPCHK(hipGetLastError());
  return 0;
}

int fit_apply_abc(fit_ctx *c, uint32_t mask) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (mask & ~7u) return fail(FIT_EINVAL, "abc mask uses bits above 2");
  if (int rc = set_device(c)) return rc;
  for (int ax = 0; ax < 3; ++ax) {
    if (!(mask & (1u << ax))) continue;
    const int64_t plane = ax == 0 ? (int64_t)c->ny * c->nz
                                  : (ax == 1 ? (int64_t)c->nx * c->nz : (int64_t)c->nx * c->ny);
    if (c->dtype == FIT_F64)
      hipLaunchKernelGGL(fit::abc_low<double>, dim3(nblocks(plane, 256)), dim3(256), 0, c->stream,
                         (double *)c->E, (double *)c->H, ax, c->nx, c->ny, c->nz, c->N);
    else
      hipLaunchKernelGGL(fit::abc_low<float>, dim3(nblocks(plane, 256)), dim3(256), 0, c->stream,
                         (float *)c->E, (float *)c->H, ax, c->nx, c->ny, c->nz, c->N);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int fit_wake_probe(fit_ctx *c, int32_t ixt, int32_t iyt, int32_t h, int64_t nt_cap) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (int rc = set_device(c)) return rc;
  if (c->hist) HIPCHK(hipFree(c->hist));
  c->hist = nullptr;
  c->hist_nt = 0;
  c->hist_bytes = 0;
  if (nt_cap <= 0) return 0;
  if (h < 0 || h > 1) return fail(FIT_EINVAL, "window half-width must be 0 or 1");
  if (ixt - h < 0 || ixt + h >= c->nx || iyt - h < 0 || iyt + h >= c->ny)
    return fail(FIT_EINVAL, "probe window outside the grid");
  const int nw = (2 * h + 1) * (2 * h + 1);
  const size_t bytes = (size_t)nw * c->nz * nt_cap * c->esz;
  if (hipMalloc(&c->hist, bytes) != hipSuccess)
    return fail(FIT_ENOMEM, "probe history allocation failed");
  c->hist_bytes = bytes;
  note_peak(c);
  HIPCHK(hipMemsetAsync(c->hist, 0, bytes, c->stream));
  c->hist_nt = nt_cap;
  c->ixt = ixt;
  c->iyt = iyt;
  c->hw = h;
  return 0;
}

int fit_wake_history(fit_ctx *c, void *host, int64_t count) {
  if (!c || !host) return fail(FIT_EINVAL, "null argument");
  if (!c->hist) return fail(FIT_ESTATE, "no probe history");
  const int nw = (2 * c->hw + 1) * (2 * c->hw + 1);
  if (count != (int64_t)nw * c->nz * c->hist_nt) return fail(FIT_EINVAL, "history count mismatch");
  if (int rc = set_device(c)) return rc;
  HIPCHK(hipMemcpyAsync(host, c->hist, (size_t)count * c->esz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_wake_history_load(fit_ctx *c, int32_t ixt, int32_t iyt, int32_t h, int64_t nt,
                          const void *host) {
  if (!c || !host) return fail(FIT_EINVAL, "null argument");
  if (nt <= 0) return fail(FIT_EINVAL, "history needs at least one time column");
  if (int rc = fit_wake_probe(c, ixt, iyt, h, nt)) return rc;
  const size_t bytes = c->hist_bytes;
  HIPCHK(hipMemcpyAsync(c->hist, host, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_monitor_set(fit_ctx *c, int32_t nf) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (nf < 0) return fail(FIT_EINVAL, "nf < 0");
  if (int rc = set_device(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->mon) HIPCHK(hipFree(c->mon));
  c->mon = nullptr;
  c->mon_nf = 0;
  c->mon_bytes = 0;
  c->mon_ph_steps = 0;
  if (nf == 0) return 0;
  const size_t bytes = (size_t)3 * nf * c->N * 16;
  if (hipMalloc(&c->mon, bytes) != hipSuccess) {
    c->mon = nullptr;
    (void)hipGetLastError();
    return fail(FIT_ENOMEM, "field monitor accumulators do not fit");
  }
  c->mon_bytes = bytes;
  note_peak(c);
  HIPCHK(hipMemsetAsync(c->mon, 0, bytes, c->stream));
  c->mon_nf = nf;
  return 0;
}

namespace {
int monitor_stage(fit_ctx *c, const double *rows, int64_t nrows) {
  // the table is sized in bytes: a later fit_monitor_set may change nf
  const size_t need = (size_t)nrows * 2 * c->mon_nf * 8;
  if (c->mon_ph_bytes < need) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->mon_ph) HIPCHK(hipFree(c->mon_ph));
    c->mon_ph = nullptr;
    c->mon_ph_cap = 0;
    c->mon_ph_bytes = 0;
    if (hipMalloc((void **)&c->mon_ph, need) != hipSuccess) {
      c->mon_ph = nullptr;
      (void)hipGetLastError();
      return fail(FIT_ENOMEM, "field monitor phase table allocation failed");
    }
    c->mon_ph_bytes = need;
    note_peak(c);
  }
  c->mon_ph_cap = (int64_t)(c->mon_ph_bytes / ((size_t)2 * c->mon_nf * 8));
  // the previous use of the table (launches of an earlier call) must be done
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->mon_ph, rows, need, hipMemcpyHostToDevice));
  return 0;
}
}  // namespace

int fit_monitor_accumulate(fit_ctx *c, const double *phase) {
  if (!c || !phase) return fail(FIT_EINVAL, "null argument");
  if (!c->mon) return fail(FIT_ESTATE, "no field monitor");
  if (int rc = set_device(c)) return rc;
  if (int rc = monitor_stage(c, phase, 1)) return rc;
  c->mon_ph_steps = 0;
  if (c->dtype == FIT_F64) monitor_launch<double>(c, c->mon_ph);
  else monitor_launch<float>(c, c->mon_ph);
  HIPCHK(hipGetLastError());
  return 0;
}

int fit_monitor_phases(fit_ctx *c, const double *phases, int64_t nsteps) {
  if (!c || (!phases && nsteps > 0)) return fail(FIT_EINVAL, "null argument");
  if (!c->mon) return fail(FIT_ESTATE, "no field monitor");
  if (nsteps < 0) return fail(FIT_EINVAL, "nsteps < 0");
  if (int rc = set_device(c)) return rc;
  if (nsteps > 0)
    if (int rc = monitor_stage(c, phases, nsteps)) return rc;
  c->mon_ph_steps = nsteps;
  return 0;
}

int fit_monitor_read(fit_ctx *c, int32_t f, int32_t comp, double *host) {
  if (!c || !host) return fail(FIT_EINVAL, "null argument");
  if (!c->mon) return fail(FIT_ESTATE, "no field monitor");
  if (f < 0 || f >= c->mon_nf || comp < 0 || comp > 2) return fail(FIT_EINVAL, "bad monitor index");
  if (int rc = set_device(c)) return rc;
  const size_t off = ((size_t)f * 3 + comp) * c->N * 16;
  HIPCHK(hipMemcpyAsync(host, (char *)c->mon + off, (size_t)c->N * 16, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_step(fit_ctx *c, int32_t nsteps, int32_t conductive) {
  if (int rc = check_ready(c)) return rc;
  if (nsteps < 0) return fail(FIT_EINVAL, "nsteps < 0");
  if (c->jmode == FIT_J_BUFFER && !c->J) return fail(FIT_ESTATE, "no J buffer");
  if (int rc = set_device(c)) return rc;
  return c->dtype == FIT_F64 ? step_impl<double>(c, nsteps, conductive)
                             : step_impl<float>(c, nsteps, conductive);
}

int fit_step_timed(fit_ctx *c, int32_t nsteps, int32_t conductive, int32_t per_kernel,
                   double *ms) {
  if (int rc = check_ready(c)) return rc;
  if (!ms) return fail(FIT_EINVAL, "null output");
  if (nsteps < 1) return fail(FIT_EINVAL, "nsteps < 1");
  if (int rc = set_device(c)) return rc;
  ms[0] = ms[1] = ms[2] = 0.0;
  if (!per_kernel) {
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = 0;
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(FIT_EHIP, "event record");
    if (!rc) rc = c->dtype == FIT_F64 ? step_impl<double>(c, nsteps, conductive)
                                      : step_impl<float>(c, nsteps, conductive);
    if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(FIT_EHIP, "event record");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(FIT_EHIP, "event sync");
    float t = 0;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(FIT_EHIP, "elapsed");
    ms[0] = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  // one more event before the call: ms[0] spans the whole launch sequence
  // (beam table included), the same window as per_kernel = 0
  std::vector<hipEvent_t> ev((size_t)3 * nsteps + 1);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  int rc = 0;
  if (hipEventRecord(ev.back(), c->stream) != hipSuccess) rc = fail(FIT_EHIP, "event record");
  if (!rc)
    rc = c->dtype == FIT_F64 ? step_impl<double>(c, nsteps, conductive, ev.data())
                             : step_impl<float>(c, nsteps, conductive, ev.data());
  if (!rc && hipEventSynchronize(ev[3 * nsteps - 1]) != hipSuccess) rc = fail(FIT_EHIP, "event sync");
  for (int s = 0; !rc && s < nsteps; ++s) {
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, ev[3 * s], ev[3 * s + 1]) != hipSuccess ||
        hipEventElapsedTime(&b, ev[3 * s + 1], ev[3 * s + 2]) != hipSuccess) {
      rc = fail(FIT_EHIP, "elapsed");
      break;
    }
    ms[1] += a;
    ms[2] += b;
  }
  float tot = 0;
  if (!rc && hipEventElapsedTime(&tot, ev.back(), ev[3 * nsteps - 1]) != hipSuccess)
    rc = fail(FIT_EHIP, "elapsed");
  ms[0] = tot;
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

extern "C++" {
template <typename T>
static int copy_tiled_impl(int nx, int ny, int nz, int cfg, int tx, int ty, int kc, int reps,
                           double *ms, int device) {
  const int64_t N = (int64_t)nx * ny * nz;
  // unused dynamic LDS that leaves room for as many blocks per CU as the
  // sweep itself keeps resident, so that both run the same rounds of tiles,
  // and the sweep's choice of planes per chunk for this grid
  const int occ = fit::fused_blocks_per_cu<T>(cfg);
  int lds_cu = 0, ncu = 0;
  if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                            device) != hipSuccess)
    lds_cu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    ncu = 256;
  const int ncol = ((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
  kc = fit::pick_kc(kc, ncol, nz, ncu * (occ > 0 ? occ : 1));
  const int64_t blocks = (int64_t)ncol * ((nz + kc - 1) / kc);
  const dim3 grid((unsigned)blocks), block((unsigned)(tx * ty));
  const size_t dyn = (occ > 0 && lds_cu > 2048) ? ((size_t)(lds_cu / occ - 1024) & ~(size_t)255) : 0;
  fit::Copy6<T> p{};
  std::vector<T *> bufs;
  int rc = 0;
  for (int c = 0; c < 12 && !rc; ++c) {
    T *q = nullptr;
    if (hipMalloc(&q, N * sizeof(T)) != hipSuccess) {
      rc = fail(FIT_ENOMEM, "tiled copy benchmark buffers");
      break;
    }
    bufs.push_back(q);
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    rc = fail(FIT_EHIP, "stream");
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(FIT_EHIP, "events");
  if (!rc) {
    for (int c = 0; c < 6; ++c) {
      p.a[c] = bufs[c];
      p.b[c] = bufs[6 + c];
      (void)hipMemsetAsync(bufs[c], 0, N * sizeof(T), st);
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL(fit::copy_zmarch<T>, grid, block, dyn, st, p, nx, ny, nz, tx, ty, kc);
    if (hipGetLastError() != hipSuccess) rc = fail(FIT_EHIP, "tiled copy launch");
  }
  if (!rc) {
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(fit::copy_zmarch<T>, grid, block, dyn, st, p, nx, ny, nz, tx, ty, kc);
    (void)hipEventRecord(e1, st);
    float t = 0;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      rc = fail(FIT_EHIP, "tiled copy benchmark timing");
    *ms = t / reps;
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  for (T *q : bufs) (void)hipFree(q);
  return rc;
}
}  // extern "C++"

int fit_copy_bench_tiled(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t dtype,
                         int32_t cfg, int32_t reps, double *ms) {
  if (!ms || nx < 1 || ny < 1 || nz < 1 || reps < 1 || cfg < 0 || cfg >= fit::FUSED_NCFG ||
      (dtype != FIT_F32 && dtype != FIT_F64))
    return fail(FIT_EINVAL, "bad tiled copy benchmark arguments");
  if ((int64_t)nx * ny * nz >= ((int64_t)1 << 31))
    return fail(FIT_EINVAL, "tiled copy benchmark grid too large");
  int tx, ty, kc;
  fit::fused_dims(cfg, tx, ty, kc);
  if (tx * ty > 1024) return fail(FIT_EINVAL, "tile too large");
  HIPCHK(hipSetDevice(device));
  return dtype == FIT_F64 ? copy_tiled_impl<double>(nx, ny, nz, cfg, tx, ty, kc, reps, ms, device)
                          : copy_tiled_impl<float>(nx, ny, nz, cfg, tx, ty, kc, reps, ms, device);
}

int fit_copy_bench(int32_t device, int64_t nbytes, int32_t reps, double *ms) {
  if (!ms || nbytes < 16 || reps < 1) return fail(FIT_EINVAL, "bad copy benchmark arguments");
  HIPCHK(hipSetDevice(device));
  const int64_t n4 = nbytes / 16;
  if (n4 > ((int64_t)1 << 31) - 256) return fail(FIT_EINVAL, "copy benchmark size too large");
  float4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n4 * 16) != hipSuccess || hipMalloc(&dst, n4 * 16) != hipSuccess) {
    (void)hipFree(src);
    return fail(FIT_ENOMEM, "copy benchmark buffers");
  }
  hipStream_t st;
  hipEvent_t e0, e1;
  int rc = 0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = fail(FIT_EHIP, "stream");
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(FIT_EHIP, "events");
  if (!rc) {
    const unsigned blocks = (unsigned)((n4 + 255) / 256);
    (void)hipMemsetAsync(src, 0, n4 * 16, st);
    hipLaunchKernelGGL(fit::stream_copy, dim3(blocks), dim3(256), 0, st, src, dst, n4);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(fit::stream_copy, dim3(blocks), dim3(256), 0, st, src, dst, n4);
    (void)hipEventRecord(e1, st);
    float t = 0;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      rc = fail(FIT_EHIP, "copy benchmark timing");
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

int fit_set_step_index(fit_ctx *c, int64_t step) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (step < 0) return fail(FIT_EINVAL, "step < 0");
  c->step = step;
  return 0;
}

int fit_get_step_index(fit_ctx *c, int64_t *step) {
  if (!c || !step) return fail(FIT_EINVAL, "null argument");
  *step = c->step;
  return 0;
}

int fit_set_kernel(fit_ctx *c, int32_t kernel) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (kernel < FIT_KERNEL_AUTO || kernel > FIT_KERNEL_FUSED2) return fail(FIT_EINVAL, "bad kernel id");
  c->kernel = kernel;
  if (kernel == FIT_KERNEL_SPLIT) {
    // the second buffers are no longer needed
    if (int rc = set_device(c)) return rc;
    if (c->E_alt) HIPCHK(hipFree(c->E_alt));
    if (c->H_alt) HIPCHK(hipFree(c->H_alt));
    c->E_alt = c->H_alt = nullptr;
  }
  return 0;
}

int fit_set_tile(fit_ctx *c, int32_t cfg) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (cfg < -1 || cfg >= fit::SWEEP_NCFG) return fail(FIT_EINVAL, "bad tile configuration");
  if (cfg >= 0 && !(c->dtype == FIT_F64 ? fit::sweep_cfg_ok<double>(cfg) : fit::sweep_cfg_ok<float>(cfg)))
    return fail(FIT_EINVAL, "tile configuration does not fit the LDS for this dtype");
  c->fused_cfg = cfg;
  return 0;
}

int fit_set_chunk_rsrc(fit_ctx *c, int32_t on) {
  if (!c) return fail(FIT_EINVAL, "null context");
  c->chunk_rsrc = on != 0;
  c->f2_def = -1;  // the two-step default shape depends on it
  return 0;
}

int fit_get_tile(fit_ctx *c, int32_t *cfg, int32_t *tx, int32_t *ty, int32_t *kc) {
  if (!c) return fail(FIT_EINVAL, "null context");
  const int k = c->fused_cfg >= 0 ? c->fused_cfg : fit::sweep_default(c->nx, c->ny, c->nz, c->dtype == FIT_F64);
  int a = 0, b = 0, d = 0;
  fit::sweep_dims(k, a, b, d);
  if (cfg) *cfg = k;
  if (tx) *tx = a;
  if (ty) *ty = b;
  if (kc) *kc = d;
  return 0;
}

int fit_tune(fit_ctx *c, int32_t reps, double *ms) {
  if (int rc = check_ready(c)) return rc;
  if (reps < 1) return fail(FIT_EINVAL, "reps < 1");
  bool fused = false;
  if (int rc = want_fused(c, &fused)) return rc;
  if (!fused) return fail(FIT_ESTATE, "the fused sweep is not in use");
  if (int rc = set_device(c)) return rc;
  return c->dtype == FIT_F64 ? tune_impl<double>(c, reps, ms) : tune_impl<float>(c, reps, ms);
}

int fit_set_tile2(fit_ctx *c, int32_t cfg) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (cfg < -1 || cfg >= fit::F2_NCFG) return fail(FIT_EINVAL, "bad two-step tile shape");
  c->f2_cfg = cfg;
  return 0;
}

int fit_get_tile2(fit_ctx *c, int32_t *cfg, int32_t *tx, int32_t *ty) {
  if (!c || !cfg) return fail(FIT_EINVAL, "null argument");
  *cfg = -1;
  if (c->geo_set && c->mat_set && c->dt_set) {
    if (int rc = set_device(c)) return rc;
    *cfg = c->dtype == FIT_F32 ? f2_cfg_of<float>(c, make_args<float>(c))
                               : f2_cfg_of<double>(c, make_args<double>(c));
  }
  int a = 0, b = 0;
  if (*cfg >= 0) fit::f2_dims(*cfg, a, b);
  if (tx) *tx = a;
  if (ty) *ty = b;
  return 0;
}

int fit_tune2(fit_ctx *c, int32_t reps, double *ms) {
  if (int rc = check_ready(c)) return rc;
  if (reps < 1) return fail(FIT_EINVAL, "reps < 1");
  bool fused = false;
  if (int rc = want_fused(c, &fused)) return rc;
  if (!fused) return fail(FIT_ESTATE, "the fused sweeps are not in use");
  if (int rc = set_device(c)) return rc;
  return c->dtype == FIT_F32 ? tune2_impl<float>(c, reps, ms) : tune2_impl<double>(c, reps, ms);
}

int fit_kernel_in_use(fit_ctx *c, int32_t *kernel) {
  if (!c || !kernel) return fail(FIT_EINVAL, "null argument");
  bool fused = false;
  if (int rc = want_fused(c, &fused)) return rc;
  *kernel = fused ? FIT_KERNEL_FUSED : FIT_KERNEL_SPLIT;
  if (fused && (c->kernel == FIT_KERNEL_AUTO || c->kernel == FIT_KERNEL_FUSED2) &&
      c->geo_set && c->mat_set && c->dt_set) {
    if (int rc = set_device(c)) return rc;
    const bool two = c->dtype == FIT_F32 ? fit::fused2_supported<float>(make_args<float>(c))
                                         : fit::fused2_supported<double>(make_args<double>(c));
    if (two) *kernel = FIT_KERNEL_FUSED2;
  }
  return 0;
}

int fit_device_bytes(fit_ctx *c, int64_t *now, int64_t *peak) {
  if (!c || !now || !peak) return fail(FIT_EINVAL, "null argument");
  *now = resident_bytes(c);
  *peak = c->peak_bytes;
  return 0;
}

int fit_launch_counts(fit_ctx *c, int64_t *counts) {
  if (!c || !counts) return fail(FIT_EINVAL, "null argument");
  counts[0] = c->launches[0];
  counts[1] = c->launches[1];
  return 0;
}

int fit_sync(fit_ctx *c) {
  if (!c) return fail(FIT_EINVAL, "null context");
  if (int rc = set_device(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int fit_wake_potential(fit_ctx *c, int32_t nwin, const int32_t *wids, int32_t carry, int32_t ns,
                       int32_t nzw, int32_t k0, int32_t zrev, const int32_t *cols,
                       const double *dz, double qnorm, double *wp) {
  if (!c || !wids || !cols || !dz || !wp) return fail(FIT_EINVAL, "null argument");
  if (!c->hist) return fail(FIT_ESTATE, "no probe history");
  const int nwh = (2 * c->hw + 1) * (2 * c->hw + 1);
  if (nwin < 1 || ns < 1 || nzw < 0 || k0 < 0 || k0 + nzw > c->nz)
    return fail(FIT_EINVAL, "bad wake potential sizes");
  for (int w = 0; w < nwin; ++w)
    if (wids[w] < 0 || wids[w] >= nwh) return fail(FIT_EINVAL, "bad window id");
  for (int64_t q = 0; q < (int64_t)nzw * ns; ++q)
    if (cols[q] >= c->hist_nt) return fail(FIT_EINVAL, "time column beyond the probe history");
  if (int rc = set_device(c)) return rc;
  int32_t *dw = nullptr, *dcols = nullptr;
  double *ddz = nullptr, *dwp = nullptr;
  const size_t cb = (size_t)std::max<int64_t>((int64_t)nzw * ns, 1) * 4;
  if (hipMalloc(&dw, nwin * 4) != hipSuccess || hipMalloc(&dcols, cb) != hipSuccess ||
      hipMalloc(&ddz, std::max(nzw, 1) * 8) != hipSuccess ||
      hipMalloc(&dwp, (size_t)nwin * ns * 8) != hipSuccess) {
    (void)hipFree(dw);
    (void)hipFree(dcols);
    (void)hipFree(ddz);
    (void)hipFree(dwp);
    return fail(FIT_ENOMEM, "wake potential scratch allocation failed");
  }
  note_peak(c, (int64_t)nwin * 4 + (int64_t)cb + (int64_t)std::max(nzw, 1) * 8 +
                   (int64_t)nwin * ns * 8);
  int rc = 0;
  hipError_t e = hipMemcpyAsync(dw, wids, nwin * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && nzw > 0)
    e = hipMemcpyAsync(dcols, cols, (size_t)nzw * ns * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && nzw > 0)
    e = hipMemcpyAsync(ddz, dz, (size_t)nzw * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    if (c->dtype == FIT_F64)
      hipLaunchKernelGGL(fit::wake_potential<double>, dim3(nblocks(ns, 256)), dim3(256), 0,
                         c->stream, (const double *)c->hist, c->hist_nt, c->nz, dw, nwin, carry,
                         ns, nzw, k0, zrev != 0, dcols, ddz, qnorm, dwp);
    else
      hipLaunchKernelGGL(fit::wake_potential<float>, dim3(nblocks(ns, 256)), dim3(256), 0,
                         c->stream, (const float *)c->hist, c->hist_nt, c->nz, dw, nwin, carry,
                         ns, nzw, k0, zrev != 0, dcols, ddz, qnorm, dwp);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(wp, dwp, (size_t)nwin * ns * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) rc = fail(FIT_EHIP, std::string("wake potential: ") + hipGetErrorString(e));
  (void)hipFree(dw);
  (void)hipFree(dcols);
  (void)hipFree(ddz);
  (void)hipFree(dwp);
  return rc;
}

int fit_impedance(fit_ctx *c, const double *x, const double *lam, int64_t len, int64_t N,
                  int32_t K, double ds, double *z_re, double *z_im, double *x_re, double *x_im,
                  double *l_re, double *l_im) {
  if (!c || !x || !lam || !z_re || !z_im || !x_re || !x_im || !l_re || !l_im)
    return fail(FIT_EINVAL, "null argument");
  if (len < 1 || N < 1 || K < 1 || K > N) return fail(FIT_EINVAL, "bad DFT sizes");
  if ((int64_t)K * N > ((int64_t)1 << 62)) return fail(FIT_EINVAL, "DFT index overflow");
  if (int rc = set_device(c)) return rc;
  const int64_t M = std::min(len, N);
  double *dx = nullptr, *dl = nullptr, *dout = nullptr;
  if (hipMalloc(&dx, M * 8) != hipSuccess || hipMalloc(&dl, M * 8) != hipSuccess ||
      hipMalloc(&dout, (size_t)K * 6 * 8) != hipSuccess) {
    (void)hipFree(dx);
    (void)hipFree(dl);
    (void)hipFree(dout);
    return fail(FIT_ENOMEM, "impedance scratch allocation failed");
  }
  note_peak(c, 16 * M + (int64_t)K * 48);
  std::vector<double> out((size_t)K * 6);
  hipError_t e = hipMemcpyAsync(dx, x, M * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, lam, M * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fit::impedance_dft, dim3(K), dim3(256), 0, c->stream, dx, dl, M, N, ds, dout,
                       (int64_t)6, (int64_t)1);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), dout, (size_t)K * 6 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dx);
  (void)hipFree(dl);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(FIT_EHIP, std::string("impedance: ") + hipGetErrorString(e));
  for (int j = 0; j < K; ++j) {
    z_re[j] = out[j * 6 + 0];
    z_im[j] = out[j * 6 + 1];
    x_re[j] = out[j * 6 + 2];
    x_im[j] = out[j * 6 + 3];
    l_re[j] = out[j * 6 + 4];
    l_im[j] = out[j * 6 + 5];
  }
  return 0;
}

int fit_impedance_noalloc(fit_ctx *c, const double *x, const double *lam, int64_t len,
                          int64_t N, int32_t K, double ds, double *z_re, double *z_im,
                          double *x_re, double *x_im, double *l_re, double *l_im) {
  if (!c || !x || !lam || !z_re || !z_im || !x_re || !x_im || !l_re || !l_im)
    return fail(FIT_EINVAL, "null argument");
  if (len < 1 || N < 1 || K < 1 || K > N) return fail(FIT_EINVAL, "bad DFT sizes");
  if ((int64_t)K * N > ((int64_t)1 << 62)) return fail(FIT_EINVAL, "DFT index overflow");
  const int64_t M = std::min(len, N);
  if (M > NOALLOC_MAX_LEN)
    return fail(FIT_EINVAL, "fit_impedance_noalloc: min(len, N) = " + std::to_string(M) +
                                " exceeds the scratch limit of " +
                                std::to_string(NOALLOC_MAX_LEN) + " samples");
  if (K > NOALLOC_MAX_BINS)
    return fail(FIT_EINVAL, "fit_impedance_noalloc: K = " + std::to_string(K) +
                                " exceeds the scratch limit of " +
                                std::to_string(NOALLOC_MAX_BINS) + " bins");
  if (int rc = set_device(c)) return rc;
  double *sp = nullptr;
  if (int rc = noalloc_scratch(c->device, &sp)) return rc;
  double *dx = sp, *dl = sp + NOALLOC_MAX_LEN, *dout = sp + 2 * NOALLOC_MAX_LEN;
  // every output word the copies below read is written by the kernel, and the
  // inputs are written before it, so nothing needs clearing between calls
  hipError_t e = hipMemcpyAsync(dx, x, M * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, lam, M * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fit::impedance_dft, dim3(K), dim3(256), 0, c->stream, dx, dl, M, N, ds,
                       dout, (int64_t)1, (int64_t)K);
    e = hipGetLastError();
  }
  double *const outs[6] = {z_re, z_im, x_re, x_im, l_re, l_im};
  for (int q = 0; q < 6 && e == hipSuccess; ++q)
    e = hipMemcpyAsync(outs[q], dout + (int64_t)q * K, (size_t)K * 8, hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess)
    return fail(FIT_EHIP, std::string("impedance (no alloc): ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"

#include "fit_batch.inc"
