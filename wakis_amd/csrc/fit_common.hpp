// Shared device-side definitions of the FIT step kernels (gfx950).
//
// Field layout: three component blocks of N = nx*ny*nz values, Fortran order
// inside a block (n = i + nx*j + nx*ny*k), as wakis/field.py:19-91.
//
// The discrete curl of the reference is a 3N x 3N sparse matrix built from
// flat-index shift matrices P_d = diags([-1, 1], [0, s_d]) with
// s = (1, nx, nx*ny) (wakis/solverFIT3D.py:207-227).  Its neighbour n + s_d
// exists iff n + s_d < N, so an x or y neighbour past the end of a row wraps
// "helically" into the next row/plane (i = nx-1 -> (0, j+1, k)), and a z
// neighbour past the last plane is dropped.  The kernels reproduce that rule
// exactly: nb_plus/nb_minus step (i, j, k) with carries/borrows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <unordered_map>

namespace fit {

// A sweep variant is launched only if it keeps every value in registers:
// variants the compiler spills to scratch memory are never launched (they
// are slow, and their results have not been validated).  Cached per kernel,
// behind a lock.
inline bool kernel_spill_free(const void *fn) {
  static std::mutex mu;
  static std::unordered_map<const void *, bool> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(fn);
  if (it != cache.end()) return it->second;
  hipFuncAttributes at{};
  const bool ok = hipFuncGetAttributes(&at, fn) == hipSuccess && at.localSizeBytes == 0;
  cache[fn] = ok;
  return ok;
}

// Face bits: xlo, xhi, ylo, yhi, zlo, zhi.
constexpr uint32_t XLO = 1u, XHI = 2u, YLO = 4u, YHI = 8u, ZLO = 16u, ZHI = 32u;

enum { J_ZERO = 0, J_SIGMA_E = 1, J_BUFFER = 2 };
enum { BEAM_OFF = 0, BEAM_VALUES = 1, BEAM_GAUSS = 2 };

// T in the constant address space: a load through it is a scalar load when
// the address is wave-uniform, and does not wait behind the vector loads in
// flight.  Only for tables that no kernel writes while a step runs.
template <typename T>
using CT = __attribute__((address_space(4))) T;

// T in the global address space (vector memory instructions, not flat ones,
// also for a pointer that was read from memory).
template <typename T>
using GT = __attribute__((address_space(1))) T;

// Material table row: dt/mu x,y,z | dt/eps x,y,z | sigma x,y,z.
constexpr int MAT_STRIDE = 9;
constexpr int MAX_MATERIALS = 256;

template <typename T>
struct Geo {
  const T *Lx, *Ly, *Lz;        // primal lengths, 1-D per axis
  const T *tLx, *tLy, *tLz;     // dual lengths, 1-D per axis
  const T *iAx, *iAy, *iAz;     // inverse primal areas, 2-D (j,k) (i,k) (i,j)
  const T *itAx, *itAy, *itAz;  // inverse dual areas, same 2-D layouts
};

template <typename T>
struct StepArgs {
  T *E;            // 3N
  T *H;            // 3N
  const T *J;      // 3N (J_BUFFER only)
  const uint8_t *mat;  // N material ids (nullptr: all 0)
  const int32_t *mat32;  // N wide material ids (more than 256 rows), else nullptr
  const T *mtab;   // nmat * MAT_STRIDE
  int nmat;
  Geo<T> g;
  int nx, ny, nz;
  int64_t N;
  bool buf32;      // a field (3N values) and the metric tables span < 2^32 bytes
  bool nt;         // field stores of the step kernels non-temporal
  bool pcr;        // lean path: buffer resources per field component and chunk
  int64_t geo_bytes;  // bytes of the packed metric tables (g.Lx first)
  uint32_t pec;    // tangential-E column mask faces
  uint32_t pmc;    // tangential-H row mask faces
  int jmode;
  // beam line
  int beam_mode, ixs, iys;
  const T *jline;        // BEAM_VALUES: nz values
  const double *zc;      // BEAM_GAUSS: cell-centre z
  double b_qv, b_A, b_B, b_s0, b_v, b_dx, b_dy;
  bool b_f32time;  // t = step*dt and v*t as float32 products
  double dt64;
  float dt32;
  int64_t step;
  // wake probe
  T *hist;
  int64_t hist_nt;
  int ixt, iyt, hw;
  // field monitor update fused into the step (one-step fused sweep): the
  // complex128 accumulators acc[f][c][N] and this step's nf phases, or null
  double2 *mon;
  const double *mph;
  int mnf;
};

// Boundary mask of component `comp` at (i, j, k).  The reference builds it as
// a Field of ones and assigns each face's flag to the tangential components on
// that face, in the order xlo, xhi, ylo, yhi, zlo, zhi
// (wakis/boundaries.py:90-100, 130-140): on an edge the LAST face written
// wins, so an unmasked face (flag 1) written after a masked one unmasks the
// edge.  m holds the masked faces (flag 0).
__device__ __forceinline__ bool masked(uint32_t m, int comp, int i, int j, int k,
                                       int nx, int ny, int nz) {
  bool r = false;
  if (comp != 0) {
    if (i == 0) r = (m & XLO) != 0;
    if (i == nx - 1) r = (m & XHI) != 0;
  }
  if (comp != 1) {
    if (j == 0) r = (m & YLO) != 0;
    if (j == ny - 1) r = (m & YHI) != 0;
  }
  if (comp != 2) {
    if (k == 0) r = (m & ZLO) != 0;
    if (k == nz - 1) r = (m & ZHI) != 0;
  }
  return r;
}

// Flat-index neighbour n + s_axis, as (i, j, k) with carries; false if >= N.
__device__ __forceinline__ bool nb_plus(int axis, int &i, int &j, int &k, int nx,
                                        int ny, int nz) {
  if (axis == 0) {
    if (++i < nx) return true;
    i = 0;
    if (++j < ny) return true;
    j = 0;
    return ++k < nz;
  }
  if (axis == 1) {
    if (++j < ny) return true;
    j = 0;
    return ++k < nz;
  }
  return ++k < nz;
}

// Flat-index neighbour n - s_axis, as (i, j, k) with borrows; false if < 0.
__device__ __forceinline__ bool nb_minus(int axis, int &i, int &j, int &k, int nx,
                                         int ny, int nz) {
  if (axis == 0) {
    if (--i >= 0) return true;
    i = nx - 1;
    if (--j >= 0) return true;
    j = ny - 1;
    return --k >= 0;
  }
  if (axis == 1) {
    if (--j >= 0) return true;
    j = ny - 1;
    return --k >= 0;
  }
  return --k >= 0;
}

__device__ __forceinline__ int64_t flat(int i, int j, int k, int nx, int ny) {
  return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k);
}

template <typename T>
__device__ __forceinline__ T Lof(const Geo<T> &g, int comp, int i, int j, int k) {
  return comp == 0 ? g.Lx[i] : (comp == 1 ? g.Ly[j] : g.Lz[k]);
}
template <typename T>
__device__ __forceinline__ T tLof(const Geo<T> &g, int comp, int i, int j, int k) {
  return comp == 0 ? g.tLx[i] : (comp == 1 ? g.tLy[j] : g.tLz[k]);
}

// Column-scaled E of the H update: u_d = L_d * pecmask_d * E_d  (0 if absent).
template <typename T>
__device__ __forceinline__ T u_at(const StepArgs<T> &a, int comp, int i, int j, int k) {
  if (masked(a.pec, comp, i, j, k, a.nx, a.ny, a.nz)) return T(0);
  return Lof(a.g, comp, i, j, k) * a.E[comp * a.N + flat(i, j, k, a.nx, a.ny)];
}

// Column-scaled H of the E update: w_d = tL_d * pmcmask_d * H_d.
template <typename T>
__device__ __forceinline__ T w_at(const StepArgs<T> &a, int comp, int i, int j, int k) {
  if (masked(a.pmc, comp, i, j, k, a.nx, a.ny, a.nz)) return T(0);
  return tLof(a.g, comp, i, j, k) * a.H[comp * a.N + flat(i, j, k, a.nx, a.ny)];
}

// Raw buffer access (CDNA buffer instructions): a 4-SGPR resource per array,
// a per-lane 32-bit byte offset and a wave-uniform 32-bit byte offset, so an
// address costs no vector arithmetic.  Used only where every offset is in
// bounds and the array spans less than 2^32 bytes.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc make_rsrc(const void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)(uint32_t)bytes,
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(Rsrc r, uint32_t v, uint32_t s);
template <>
__device__ __forceinline__ float bload<float>(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)s, 0));
}
template <>
__device__ __forceinline__ double bload<double>(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)v, (int)s, 0));
}
// nt: non-temporal store (cache policy bit 1), for a step's output that is
// read only by the next step, after far more data than the caches hold.
// (nt is a compile-time constant at every call site)
__device__ __forceinline__ void bstore(Rsrc r, uint32_t v, uint32_t s, float x, const bool nt) {
  const unsigned u = __builtin_bit_cast(unsigned, x);
  if (nt) __builtin_amdgcn_raw_buffer_store_b32(u, r, (int)v, (int)s, 2);
  else __builtin_amdgcn_raw_buffer_store_b32(u, r, (int)v, (int)s, 0);
}
__device__ __forceinline__ void bstore(Rsrc r, uint32_t v, uint32_t s, double x, const bool nt) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 u = __builtin_bit_cast(u32x2, x);
  if (nt) __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)v, (int)s, 2);
  else __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)v, (int)s, 0);
}
template <typename P, typename V>
__device__ __forceinline__ void gstore(P *p, V x, const bool nt) {
  if (nt) __builtin_nontemporal_store(x, p);
  else *p = x;
}
__device__ __forceinline__ int bload_u8(Rsrc r, uint32_t v, uint32_t s) {
  return (int)__builtin_amdgcn_raw_buffer_load_b8(r, (int)v, (int)s, 0);
}

// Material row of cell n: the wide ids when set (split kernels only), else
// the 8-bit ids, else row 0.
template <typename T>
__device__ __forceinline__ const T *mat_row(const StepArgs<T> &a, int64_t n) {
  const int64_t m = a.mat32 ? (int64_t)a.mat32[n] : (a.mat ? (int64_t)a.mat[n] : 0);
  return a.mtab + m * MAT_STRIDE;
}

// FieldMonitor update of cell n with the step's new E (wakis/field_monitors.py:
// 96-102): acc[f][c][n] += E_c * phase_f in complex128, E promoted to (E, 0)
// and the product rounded as numpy's complex multiply, (e*pr - 0*pi,
// e*pi + 0*pr), then the two adds, no contraction into an FMA (the same
// arithmetic as the monitor_accumulate kernel).
template <typename T>
__device__ __forceinline__ void monitor_cell(const StepArgs<T> &a, int64_t n, T ex, T ey, T ez) {
#pragma clang fp contract(off)
  const T ev[3] = {ex, ey, ez};
  for (int f = 0; f < a.mnf; ++f) {
    const double pr = a.mph[2 * f], pi = a.mph[2 * f + 1];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double e = (double)ev[c];
      const double re = e * pr - 0.0 * pi;
      const double im = e * pi + 0.0 * pr;
      double2 *p = a.mon + ((int64_t)f * 3 + c) * a.N + n;
      double2 v = *p;
      v.x = v.x + re;
      v.y = v.y + im;
      *p = v;
    }
  }
}

// Wake probe: E_z of the state before the update of this step
// (wakis/routines.py:438-448) into history column `step`.
template <typename T>
__device__ __forceinline__ void probe(const StepArgs<T> &a, int i, int j, int k, T ez) {
  if (a.hist && a.step < a.hist_nt) {
    const int di = i - a.ixt, dj = j - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw) {
      const int w = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
      ((GT<T> *)a.hist)[((int64_t)w * a.nz + k) * a.hist_nt + a.step] = ez;
    }
  }
}

// Beam current J_z on the beam line at plane k (wakis/sources.py:96-110).
// With a float32 dt, numpy evaluates t = n*dt and v*t as float32 products
// (python scalars are cast to float32 first); the rest is FP64; the value is
// then stored in the solver dtype.
template <typename T>
__device__ __forceinline__ T beam_gauss(const StepArgs<T> &a, int k) {
  double vt;
  if (a.b_f32time) {
    const float t = (float)a.step * a.dt32;
    vt = (double)((float)a.b_v * t);
  } else {
    const double t = (double)a.step * a.dt64;
    vt = a.b_v * t;
  }
  const double s = a.zc[k] - vt;
  const double d = s - a.b_s0;
  const double prof = a.b_A * exp(-(d * d) / a.b_B);
  return (T)(a.b_qv * prof / a.b_dx / a.b_dy);
}

}  // namespace fit
