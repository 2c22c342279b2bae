/*
 * wakis_fit.h - C ABI of the MI355X (gfx950) FIT time-stepping library.
 *
 * Plain C types only (pointers, sizes, ints, doubles): any FFI (ctypes,
 * cffi, cgo, JNI) can bind it.  Every entry point returns 0 on success and a
 * negative FIT_E* code on failure; fit_last_error() then returns a message
 * (per host thread).  A context owns all device memory of one solve on one
 * HIP device and one HIP stream; host pointers are only borrowed for the
 * duration of a call.  Calls on one context are not thread-safe.
 *
 * Field layout (identical to the reference's Field, wakis/field.py:19-91):
 * a field is 3 blocks (x, y, z) of N = nx*ny*nz values each, every block in
 * Fortran order: n = i + nx*j + nx*ny*k.
 *
 * The reference is pure Python; its drop-in slot for this path is the
 * 0-argument callable SolverFIT3D.one_step bound at construction
 * (wakis/solverFIT3D.py:143, 504-525) plus the host loops emsolve/wakesolve
 * (wakis/routines.py:17-229, 231-485) and the WakeSolver post-processing
 * (wakis/wakeSolver.py:423-551, 628-720).  Each entry point below names the
 * reference code it replaces.
 */
#ifndef WAKIS_FIT_H
#define WAKIS_FIT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fit_ctx fit_ctx;

/* storage / compute precision of the field updates */
enum { FIT_F64 = 0, FIT_F32 = 1 };
/* field ids */
enum { FIT_FIELD_E = 0, FIT_FIELD_H = 1, FIT_FIELD_J = 2 };
/* source of the current density J used by the E update */
enum {
  FIT_J_ZERO = 0,   /* J == 0 everywhere (except the beam line) */
  FIT_J_SIGMA_E = 1,/* J == sigma*E (state after a conductive step) */
  FIT_J_BUFFER = 2  /* J read from the J field buffer */
};
/* beam line source of J_z at (ixs, iys, :) */
enum { FIT_BEAM_OFF = 0, FIT_BEAM_VALUES = 1, FIT_BEAM_GAUSS = 2 };
/* step kernels: AUTO = the fused sweeps when their second field buffers fit
 * (two-step sweeps for pairs of steps where they apply, see fit_step; the
 * one-step sweep for the other steps), else the split pair; FUSED = the
 * one-step fused sweep only; FUSED2 = as AUTO, but never the split pair */
enum { FIT_KERNEL_AUTO = 0, FIT_KERNEL_SPLIT = 1, FIT_KERNEL_FUSED = 2, FIT_KERNEL_FUSED2 = 3 };

/* error codes */
enum {
  FIT_OK = 0,
  FIT_EINVAL = -1,  /* bad argument */
  FIT_EHIP = -2,    /* HIP runtime error */
  FIT_ENOMEM = -3,  /* device allocation failed */
  FIT_ESTATE = -4   /* call not valid in the current context state */
};

typedef struct fit_desc {
  int32_t nx, ny, nz; /* cells per axis */
  int32_t dtype;      /* FIT_F64 or FIT_F32 */
  int32_t device;     /* HIP device ordinal */
  int32_t reserved[3];
} fit_desc;

/* Message of the last failed call on this host thread ("" if none). */
const char *fit_last_error(void);
/* Library version string. */
const char *fit_version(void);

/* Allocate a context: E, H, J (3N each, zeroed), per-cell material ids.
 * Replaces the Field allocations of SolverFIT3D.__init__
 * (wakis/solverFIT3D.py:182-190). */
int fit_create(const fit_desc *desc, fit_ctx **out);
int fit_destroy(fit_ctx *ctx);

/* FIT metric tables, FP64, replacing the diagonal matrices Ds, tDs, iDa, itDa
 * (wakis/solverFIT3D.py:212-217) after the boundary fix-ups of
 * wakis/boundaries.py:36-70, 153-181.  Packed layout, in this order:
 *   L    [nx | ny | nz]     primal lengths   L_x(i), L_y(j), L_z(k)
 *   tL   [nx | ny | nz]     dual lengths     tL_x(i), tL_y(j), tL_z(k)
 *   iA_x [ny*nz] (j + ny*k), iA_y [nx*nz] (i + nx*k), iA_z [nx*ny] (i + nx*j)
 *   itA_x, itA_y, itA_z     same 2-D layouts
 * total 2*(nx+ny+nz) + 2*(ny*nz + nx*nz + nx*ny) doubles. */
int fit_set_geometry(fit_ctx *ctx, const double *packed, int64_t count);

/* Materials: ids[N] (uint8, Fortran order; NULL = all cells id 0) and a
 * table of nmat rows x 9 FP64 values (1/eps x,y,z, 1/mu x,y,z, sigma x,y,z).
 * Replaces iDeps/iDmu/Dsigma and the operator products
 * (wakis/solverFIT3D.py:337-344, update_tensors 378-436). */
int fit_set_materials(fit_ctx *ctx, const uint8_t *ids, int32_t nmat,
                      const double *table);
/* Materials with more than 256 distinct rows (per-cell ieps/imu/sigma Fields
 * of any content, re-derived by update_tensors, wakis/solverFIT3D.py:378-436):
 * ids[N] int32 in [0, nmat), nmat <= N, same table layout.  The table stays
 * in global memory and every step of such a context runs the split kernels
 * (with FIT_KERNEL_AUTO; FIT_KERNEL_FUSED/FUSED2 then fail with FIT_ESTATE),
 * which read the cell's row per update: 4 bytes of id per cell and update
 * beyond fit_set_materials.  fit_set_materials switches back. */
int fit_set_materials_wide(fit_ctx *ctx, const int32_t *ids, int32_t nmat,
                           const double *table);

/* Boundary masks, bit f set = face f masked, f = xlo,xhi,ylo,yhi,zlo,zhi
 * (bits 0..5).  pec_cols: tangential-E column mask of C (PEC/PML faces,
 * wakis/boundaries.py:73-110); pmc_rows: tangential-H row mask of C
 * (wakis/boundaries.py:113-150, 0 when no PMC block is active). */
int fit_set_bc(fit_ctx *ctx, uint32_t pec_cols, uint32_t pmc_rows);

/* Time step (already cast to the solver dtype by the caller,
 * wakis/solverFIT3D.py:291-318). */
int fit_set_dt(fit_ctx *ctx, double dt);

/* How the E update obtains J (see FIT_J_*). */
int fit_set_jmode(fit_ctx *ctx, int32_t jmode);

/* Beam line: J_z(ixs, iys, k) for the next steps.
 * mode FIT_BEAM_VALUES: values[nz] FP64 (already evaluated by the host, as
 *   sources.Beam.update does, wakis/sources.py:76-110);
 * mode FIT_BEAM_GAUSS: the Gaussian of wakis/sources.py:96-110 evaluated in
 *   the kernel each step at t = step*dt; params (FP64) =
 *   {q*v, 1/sqrt(2*pi*sigmaz^2), 2*sigmaz^2, s0 = zmin - v*ti, v, dx[ixs],
 *    dy[iys], f32time} and zc[nz] the cell-centre z coordinates; f32time =
 *   1.0 evaluates t = step*dt and v*t as float32 products (what numpy does
 *   for a float32 dt), 0.0 in FP64;
 * mode FIT_BEAM_OFF: no override. */
int fit_set_beam(fit_ctx *ctx, int32_t mode, int32_t ixs, int32_t iys,
                 const double *values, const double *params,
                 const double *zc);

/* Host <-> device copies of a box [i0,i1) x [j0,j1) x [k0,k1) of one
 * component (0,1,2) of field `which`; host data is dense, Fortran order,
 * in the context dtype.  Replaces Field.__getitem__/__setitem__ on device
 * arrays (wakis/field.py:239-249, 313-335).  Reads synchronise. */
int fit_field_read(fit_ctx *ctx, int32_t which, int32_t comp,
                   const int32_t box[6], void *host);
int fit_field_write(fit_ctx *ctx, int32_t which, int32_t comp,
                    const int32_t box[6], const void *host);
/* Fill a whole field with zeros. */
int fit_field_zero(fit_ctx *ctx, int32_t which);
/* Watchdog (no counterpart in the reference, whose solver never checks):
 * *finite = 1 if every value of field `which` is finite, 0 if any is NaN or
 * infinite; one pass over the 3N values on the context's stream (a 256-byte
 * flag is allocated on the first call and kept).  Synchronous. */
int fit_field_finite(fit_ctx *ctx, int32_t which, int32_t *finite);
/* Write the J the next step would use (per the J mode: 0, sigma*E or the
 * buffer, then the beam line) into the J buffer, so that it can be read. */
int fit_materialize_j(fit_ctx *ctx);
/* Current J mode and beam mode (they change after a conductive step:
 * J mode becomes FIT_J_SIGMA_E and a FIT_BEAM_VALUES line is consumed,
 * because the reference overwrites all of J with sigma*E). */
int fit_get_state(fit_ctx *ctx, int32_t *jmode, int32_t *beam_mode);

/* First-step ghost zeroing (wakis/solverFIT3D.py:842-862). */
int fit_zero_ghosts(fit_ctx *ctx);

/* Absorbing boundary, CPU semantics of the reference (wakis/boundaries.py:
 * 307-404 with the view aliasing of wakis/field.py:250-256): for every
 * low face f (bit 0 = x, 1 = y, 2 = z) set in abc_low_mask, in x, y, z
 * order, plane 0 := plane 1 for E and H, all three components. */
int fit_apply_abc(fit_ctx *ctx, uint32_t abc_low_mask);

/* Wake probe: record E_z on a (2h+1)x(2h+1)xnz window centred at (ixt, iyt)
 * before every step (wakis/routines.py:438-448), into a device history of
 * nt_cap columns (column = step counter).  h = 0 or 1.  nt_cap = 0 turns the
 * probe off. */
int fit_wake_probe(fit_ctx *ctx, int32_t ixt, int32_t iyt, int32_t h,
                   int64_t nt_cap);
/* Copy the probe history (window-major, then z, then time; ctx dtype). */
int fit_wake_history(fit_ctx *ctx, void *host, int64_t count);
/* Load a recorded E_z history from the host, for wake post-processing
 * without a time loop (WakeSolver.solve on an existing Ez.h5 read by
 * read_Ez, wakis/wakeSolver.py:211-255, 1317-1364): as fit_wake_probe(ctx,
 * ixt, iyt, h, nt), then the (2h+1)^2 x nz x nt values of host (same layout
 * as fit_wake_history: window w = (di+h)*(2h+1) + (dj+h), then z, then time;
 * ctx dtype) are copied into the history.  Synchronous. */
int fit_wake_history_load(fit_ctx *ctx, int32_t ixt, int32_t iyt, int32_t h,
                          int64_t nt, const void *host);

/* Frequency-domain field monitor (wakis/field_monitors.py:14-131: the
 * FieldMonitor that wakesolve updates after every step, wakis/routines.py:
 * 450-451).  The context holds complex128 accumulators acc[f][c][n] (nf
 * frequencies x 3 components x N cells, flat Field order, re/im
 * interleaved); one update adds E_c * phase_f to each, with the real field
 * promoted to (E, 0) and numpy's complex product rounding.
 *   fit_monitor_set:        allocate nf frequencies' accumulators, zeroed
 *                           (nf = 0 frees them).
 *   fit_monitor_accumulate: one update with the current E; phase = nf
 *                           complex values (re, im interleaved).
 *   fit_monitor_phases:     stage nsteps rows of nf phases: the next
 *                           fit_step call updates after each of its first
 *                           nsteps steps with row s (the stage is consumed by
 *                           that call).
 *   fit_monitor_read:       copy acc[f][comp] (2N doubles) to the host. */
int fit_monitor_set(fit_ctx *ctx, int32_t nf);
int fit_monitor_accumulate(fit_ctx *ctx, const double *phase);
int fit_monitor_phases(fit_ctx *ctx, const double *phases, int64_t nsteps);
int fit_monitor_read(fit_ctx *ctx, int32_t f, int32_t comp, double *host);

/* Advance nsteps leapfrog steps (wakis/solverFIT3D.py:504-525):
 *   H -= dt*M_H E;  E += dt*(M_E H - ieps*J);  J = sigma*E if conductive.
 * conductive != 0 selects J := sigma*E after each step.  Asynchronous.
 * With FIT_KERNEL_AUTO or FIT_KERNEL_FUSED2, pairs of steps run as ONE
 * two-step sweep (E and H read and written once per two steps) on
 * contexts with nx >= 16 and ny >= 16 (FP32; FP64 with fields below 2^32
 * bytes), when J is
 * 0 or sigma*E
 * (not FIT_J_BUFFER) and no field-monitor update is staged for the pair;
 * every other step runs the one-step sweep.  Each step's beam value, J and
 * E_z probe sample are those of the one-step sequence. */
int fit_step(fit_ctx *ctx, int32_t nsteps, int32_t conductive);
/* Launches of the step kernels since the context was created: counts[0]
 * one-step launches (a fused sweep, or a split H/E pair), counts[1] two-step
 * sweeps.  No counterpart in the reference. */
int fit_launch_counts(fit_ctx *ctx, int64_t *counts);
/* Device memory of the context: *now the bytes of the buffers it holds
 * (fields, second field buffers, material ids and table, metric tables,
 * beam tables, probe history, staging, monitor), *peak the most it has held
 * at once since fit_create, the scratch that fit_wake_potential and
 * fit_impedance hold during the call included.  A device admission budget
 * is checked against these.  No counterpart in the reference. */
int fit_device_bytes(fit_ctx *ctx, int64_t *now, int64_t *peak);
/* fit_step timed with HIP events on the context stream (synchronises).
 * per_kernel = 0: ms[0] = elapsed time of the whole launch sequence.
 * per_kernel = 1: events around every launch; ms[1] = summed time of the
 * first kernel of a step (H update, or the fused sweep, one- or two-step),
 * ms[2] = summed time of the second (E update; 0 for the fused sweeps);
 * ms[0] = the whole launch sequence, as with per_kernel = 0.
 * No counterpart in the reference (it has no timing facility beyond
 * time.time() deltas, wakis/routines.py:437, 484). */
int fit_step_timed(fit_ctx *ctx, int32_t nsteps, int32_t conductive, int32_t per_kernel,
                   double *ms);
/* Plain device-to-device streaming copy of nbytes (16-byte vectors), timed
 * over reps launches: *ms = average milliseconds per launch.  Calibrates the
 * achievable HBM bandwidth and the rocprof FETCH/WRITE counters.  No
 * counterpart in the reference. */
int fit_copy_bench(int32_t device, int64_t nbytes, int32_t reps, double *ms);
/* Copy of six (nx, ny, nz) arrays of the given dtype into six others, with
 * the traversal of fused tile configuration `cfg` (a block per tile of
 * columns marching through its planes), timed over reps launches: *ms =
 * average milliseconds per launch.  The achievable rate of the sweep's own
 * access pattern; needs 12 N elements of device memory.  No counterpart in
 * the reference. */
int fit_copy_bench_tiled(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t dtype,
                         int32_t cfg, int32_t reps, double *ms);
/* Step counter used by the beam time and the probe column. */
int fit_set_step_index(fit_ctx *ctx, int64_t step);
int fit_get_step_index(fit_ctx *ctx, int64_t *step);
/* Select the step kernel (FIT_KERNEL_*). */
int fit_set_kernel(fit_ctx *ctx, int32_t kernel);
/* Tile configuration of the fused sweep (0 .. FIT_NTILE-1; -1 = default
 * for the grid); fit_get_tile reports the one in use and its tile sizes. */
enum { FIT_NTILE = 12 };
int fit_set_tile(fit_ctx *ctx, int32_t cfg);
/* Addressing of the fused sweeps (no reference counterpart: a layout choice
 * of this library).  0 (default): one buffer resource per field while a
 * field spans less than 2^32 bytes, else one per field component based at
 * each chunk (any grid fit_create accepts: fewer than 2^31 cells): in FP32
 * in the one-step sweep's lean path and in the two-step sweep, whose 64 x 8
 * shape is then the only one; in FP64 in the two-step sweep, whose 32 x 8
 * and 28 x 12 shapes are then the only ones (the FP64 one-step sweep keeps 64-bit
 * addresses); 1: per component and chunk on every grid (the same results;
 * lets small grids exercise that path). */
int fit_set_chunk_rsrc(fit_ctx *ctx, int32_t on);
int fit_get_tile(fit_ctx *ctx, int32_t *cfg, int32_t *tx, int32_t *ty, int32_t *kc);
/* Autotuner: time every tile configuration of the fused sweep on this
 * context's grid and data (reps launches each; the state is not changed)
 * and select the fastest; ms[FIT_NTILE] receives the average times. */
int fit_tune(fit_ctx *ctx, int32_t reps, double *ms);
/* Tile shape of the two-step sweep (no reference counterpart: a launch
 * choice of this library): FIT_TILE2_* columns per workgroup, -1 = the
 * default (the shape of least modelled cost: window cells per tile cell,
 * tile overhang over the grid and measured time per window cell).  A shape
 * the grid cannot take (nx < tile width, ny < max(16, tile height)) falls
 * back to the default, and so does one not built for the context's
 * precision (FP32: 64 x 8, 32 x 16, 32 x 8, 16 x 16, 60 x 12; FP64: 32 x 8,
 * 16 x 16, 28 x 12).  Every shape
 * evaluates the same per-cell expressions.  fit_get_tile2
 * reports the shape the next two-step sweep will use (-1: none, 0 x 0);
 * fit_tune2 times every shape that applies on this context's grid and data
 * (reps two-step launches each; the state is not changed), selects the
 * fastest and writes the average times to ms[FIT_NTILE2] (-1 where a shape
 * does not apply). */
enum {
  FIT_TILE2_64x8 = 0,
  FIT_TILE2_32x16 = 1,
  FIT_TILE2_32x8 = 2,
  FIT_TILE2_16x16 = 3,
  FIT_TILE2_60x12 = 4,
  FIT_TILE2_28x12 = 5,
  FIT_NTILE2 = 6
};
int fit_set_tile2(fit_ctx *ctx, int32_t cfg);
int fit_get_tile2(fit_ctx *ctx, int32_t *cfg, int32_t *tx, int32_t *ty);
int fit_tune2(fit_ctx *ctx, int32_t reps, double *ms);
/* The step kernel the next fit_step will use: FIT_KERNEL_SPLIT,
 * FIT_KERNEL_FUSED, or FIT_KERNEL_FUSED2 when its pairs of steps run the
 * two-step sweep (FP32 or FP64; see fit_step).  AUTO resolves to a fused
 * sweep when its second field buffers fit in device memory. */
int fit_kernel_in_use(fit_ctx *ctx, int32_t *kernel);
int fit_sync(fit_ctx *ctx);

/* Longitudinal wake potential on the device (wakis/wakeSolver.py:423-551,
 * 311-421).  For every window w < nwin (taken in order from window_ids) and
 * every sample n < ns:
 *   acc = carry ? acc : 0;  for k < nzw: c = cols[k*ns + n];
 *     if c >= 0: acc = acc + hist[w][r(k)][c] * dz[k];
 *   acc = acc / qnorm;  wp[w*ns + n] = acc,
 * with the history row r(k) = k0 + k, or k0 + nzw - 1 - k when zrev != 0
 * (the counter-moving integral of wakis/wakeSolver.py:500-513, which walks z
 * from the top; the caller passes dz[k] negated, as the reference weights
 * that branch's terms by -dz[k]).
 * cols[nzw*ns] holds the time column of (n, k) (-1 = no contribution),
 * computed by the host exactly as the reference's int((...)/dt). */
int fit_wake_potential(fit_ctx *ctx, int32_t nwin, const int32_t *window_ids,
                       int32_t carry, int32_t ns, int32_t nzw, int32_t k0,
                       int32_t zrev, const int32_t *cols, const double *dz,
                       double qnorm, double *wp);

/* Impedance DFT (wakis/wakeSolver.py:676-703): bins j < K of the N-point
 * zero-padded (or truncated) DFT of x[len] and of lam[len], both FP64,
 * then Z_j = -(X_j*ds)/(Lam_j*ds).  Outputs re/im of Z, X*ds and Lam*ds.
 * Runs on the context's device and stream. */
int fit_impedance(fit_ctx *ctx, const double *x, const double *lam,
                  int64_t len, int64_t N, int32_t K, double ds,
                  double *z_re, double *z_im, double *x_re, double *x_im,
                  double *l_re, double *l_im);

/* fit_impedance without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernel and results (bit for bit) as
 * fit_impedance; the device staging lives in a fixed per-device scratch
 * array of the library (4.375 MiB per device, held until the process ends),
 * so min(len, N) must be <= FIT_NOALLOC_MAX_LEN and K <= FIT_NOALLOC_MAX_BINS
 * (else FIT_EINVAL).  Every call on one device shares that array: calls that
 * overlap on one device (two host threads, or two contexts' streams) race. */
enum { FIT_NOALLOC_MAX_LEN = 262144, FIT_NOALLOC_MAX_BINS = 8192 };
int fit_impedance_noalloc(fit_ctx *ctx, const double *x, const double *lam,
                          int64_t len, int64_t N, int32_t K, double ds,
                          double *z_re, double *z_im, double *x_re,
                          double *x_im, double *l_re, double *l_im);

/* ---------------- batches (no counterpart in the reference) ----------------
 * The reference steps one solver at a time (wakis/routines.py:231-485).  A
 * batch steps many independent contexts (same device and dtype, fused
 * sweep) with ONE launch per step (a block per tile of every member); each member's arithmetic is
 * that of its own fit_step with the same tile configuration, so results are
 * the same alone and in a batch.  FP32 batches whose members all suit the
 * two-step sweep (nx >= 32, ny >= 16, J zero or sigma*E) step pairs of steps
 * with it instead (fit_batch_set_two_step): one launch per pair over the
 * lean chunks of all members, then their face chunks, E and H read and
 * written once per two steps, in tiles of 32 x 16 or (ny >= 24) 24 x 24
 * columns; an odd last step of a call takes the one-step launch.
 * Such a member's results are those of the two-step sweep, which agree with
 * the one-step sweep to FP32 rounding.  Members keep their own state (fields,
 * beam, wake probe); after fit_batch_step every member is as if it had run
 * fit_step(nsteps, conductive[i]).  Members may be replaced between calls
 * (continuous batching).  Members with host beam values (FIT_BEAM_VALUES)
 * are rejected. */
typedef struct fit_batch fit_batch;
int fit_batch_create(fit_ctx *const *members, int32_t n, const int32_t *conductive,
                     fit_batch **out);
int fit_batch_set_members(fit_batch *batch, fit_ctx *const *members, int32_t n,
                          const int32_t *conductive);
int fit_batch_destroy(fit_batch *batch);
/* Step every member nsteps; ms (optional) receives the elapsed time. */
int fit_batch_step(fit_batch *batch, int64_t nsteps, double *ms);
/* Device bytes of the batch's own tables (member table, tile offsets, step
 * counter, beam tables: group x nz values per member with a Gaussian beam);
 * they grow to the largest member set seen and are freed by
 * fit_batch_destroy. */
int fit_batch_device_bytes(fit_batch *batch, int64_t *bytes);
/* Tile configuration of the batched launch (0 .. FIT_NTILE-1 of the fused
 * sweep family; -1: default for the widest member). */
int fit_batch_set_tile(fit_batch *batch, int32_t cfg);
int fit_batch_get_tile(fit_batch *batch, int32_t *cfg, int32_t *nblocks, int32_t *ntiles);
/* Autotuner over the batch (state unchanged); ms[FIT_NTILE], -1 = not run. */
int fit_batch_tune(fit_batch *batch, int32_t reps, double *ms);
/* Launch groups of `group` steps (beam table, group sweeps, counter
 * advance); with on != 0 a group is captured once into a HIP graph and
 * replayed. */
int fit_batch_set_graph(fit_batch *batch, int32_t on, int32_t group);
/* Two-step pairs on (default) or off for this batch, their tile
 * configuration (FIT_BATCH2_32x16 or FIT_BATCH2_24x24 in FP32,
 * FIT_BATCH2_28x12 in FP64, or -1) and planes per z chunk (8 .. 4096, or
 * 0).  -1 and 0 pick by the launch model: the least rounds x (kc + 8) x
 * waves per block over all members' blocks. */
enum { FIT_BATCH2_32x16 = 0, FIT_BATCH2_24x24 = 1, FIT_BATCH2_28x12 = 2, FIT_BATCH2_NTUNE = 9 };
int fit_batch_set_two_step(fit_batch *batch, int32_t on, int32_t cfg, int32_t kc);
/* Whether the next fit_batch_step steps pairs with the two-step sweep (*on),
 * with which tile configuration (*cfg, -1 if not) and planes per chunk. */
int fit_batch_two_step(fit_batch *batch, int32_t *on, int32_t *cfg, int32_t *kc);
/* Autotuner of the two-step batch (state unchanged): times a launch for
 * each tile configuration every member suits and chunks of 48, 64 and 96
 * planes, keeps the fastest (*cfg, *kc; -1, 0 if none applies).
 * ms[cfg * 3 + i] (optional, FIT_BATCH2_NTUNE values): ms per launch of
 * chunk i, -1 = not run. */
int fit_batch_tune_two_step(fit_batch *batch, int32_t reps, int32_t *cfg, int32_t *kc,
                            double *ms);

#ifdef __cplusplus
}
#endif
#endif /* WAKIS_FIT_H */
