// Auxiliary device kernels: box copies, ghost zeroing, ABC planes, J
// materialisation, wake potential and impedance DFT.
#pragma once

#include "fit_common.hpp"

namespace fit {

// Box [i0,i1)x[j0,j1)x[k0,k1) of one component <-> dense Fortran-order buffer.
template <typename T>
__global__ void box_pack(const T *__restrict__ src, T *__restrict__ dst, int nx, int ny,
                         int i0, int bi, int j0, int bj, int k0, int64_t cnt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  int64_t r = t;
  const int di = (int)(r % bi);
  r /= bi;
  const int dj = (int)(r % bj);
  const int dk = (int)(r / bj);
  dst[t] = src[flat(i0 + di, j0 + dj, k0 + dk, nx, ny)];
}

template <typename T>
__global__ void box_unpack(const T *__restrict__ src, T *__restrict__ dst, int nx, int ny,
                           int i0, int bi, int j0, int bj, int k0, int64_t cnt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  int64_t r = t;
  const int di = (int)(r % bi);
  r /= bi;
  const int dj = (int)(r % bj);
  const int dk = (int)(r / bj);
  dst[flat(i0 + di, j0 + dj, k0 + dk, nx, ny)] = src[t];
}

template <typename T>
__global__ void fill_zero(T *p, int64_t cnt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < cnt) p[t] = T(0);
}

// wakis/solverFIT3D.py:842-862: H tangential and E normal at the last index.
template <typename T>
__global__ void zero_ghosts(T *E, T *H, int nx, int ny, int nz, int64_t N) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t nxy = (int64_t)nx * ny;
  const int k = (int)(n / nxy);
  const int64_t r = n - (int64_t)k * nxy;
  const int j = (int)(r / nx);
  const int i = (int)(r - (int64_t)j * nx);
  const bool li = i == nx - 1, lj = j == ny - 1, lk = k == nz - 1;
  if (lj || lk) H[n] = T(0);
  if (li || lk) H[N + n] = T(0);
  if (li || lj) H[2 * N + n] = T(0);
  if (li) E[n] = T(0);
  if (lj) E[N + n] = T(0);
  if (lk) E[2 * N + n] = T(0);
}

// ABC, low face of `axis`: plane 0 := plane 1 for E and H, 3 components.
template <typename T>
__global__ void abc_low(T *E, T *H, int axis, int nx, int ny, int nz, int64_t N) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a, b;
  int64_t plane;
  if (axis == 0) {
    plane = (int64_t)ny * nz;
    if (t >= plane) return;
    a = (int)(t % ny);
    b = (int)(t / ny);
  } else if (axis == 1) {
    plane = (int64_t)nx * nz;
    if (t >= plane) return;
    a = (int)(t % nx);
    b = (int)(t / nx);
  } else {
    plane = (int64_t)nx * ny;
    if (t >= plane) return;
    a = (int)(t % nx);
    b = (int)(t / nx);
  }
  int64_t d0, d1;
  if (axis == 0) {
    d0 = flat(0, a, b, nx, ny);
    d1 = flat(1, a, b, nx, ny);
  } else if (axis == 1) {
    d0 = flat(a, 0, b, nx, ny);
    d1 = flat(a, 1, b, nx, ny);
  } else {
    d0 = flat(a, b, 0, nx, ny);
    d1 = flat(a, b, 1, nx, ny);
  }
  for (int c = 0; c < 3; ++c) {
    E[c * N + d0] = E[c * N + d1];
    H[c * N + d0] = H[c * N + d1];
  }
}

// Beam line of steps step0 .. step0+nsteps-1 (wakis/sources.py:96-110),
// evaluated once per fit_step call so the step kernels only read a table.
template <typename T>
__global__ void beam_table(StepArgs<T> a, int64_t step0, int nsteps, T *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nsteps * a.nz) return;
  const int s = (int)(t / a.nz);
  const int k = (int)(t - (int64_t)s * a.nz);
  a.step = step0 + s;
  out[t] = beam_gauss<T>(a, k);
}

// J := sigma*E (jmode J_SIGMA_E) or 0 (J_ZERO), or kept as it is (jmode < 0:
// J lives in its buffer), then the beam line.
template <typename T>
__global__ void materialize_j(StepArgs<T> a) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const T *mt = mat_row(a, n);
  T *J = const_cast<T *>(a.J);
  // jmode < 0: J is in its buffer already (only the beam line is applied)
  if (a.jmode >= 0) {
    for (int c = 0; c < 3; ++c) {
      T v = T(0);
      if (a.jmode == J_SIGMA_E) v = mt[6 + c] * a.E[c * a.N + n];
      J[c * a.N + n] = v;
    }
  }
  if (a.beam_mode != BEAM_OFF) {
    const int64_t nxy = (int64_t)a.nx * a.ny;
    const int k = (int)(n / nxy);
    const int64_t r = n - (int64_t)k * nxy;
    const int j = (int)(r / a.nx);
    const int i = (int)(r - (int64_t)j * a.nx);
    if (i == a.ixs && j == a.iys)
      J[2 * a.N + n] = (a.beam_mode == BEAM_VALUES) ? a.jline[k] : beam_gauss<T>(a, k);
  }
}

// Watchdog: flag := 1 if any of the n values is NaN or infinite (every
// writer stores the same value; the host clears the flag before the launch).
template <typename T>
__global__ void nonfinite_flag(const T *__restrict__ a, int64_t n, int *__restrict__ flag) {
  bool bad = false;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    bad |= !__builtin_isfinite(a[q]);
  if (bad) *flag = 1;
}

// FieldMonitor update (wakis/field_monitors.py:96-102): acc[f][c][n] +=
// E_c[n] * phase_f in complex128, E promoted to (E, 0) and the product
// rounded as numpy's complex multiply, (e*pr - 0*pi, e*pi + 0*pr), then
// the two adds (no contraction into an FMA), so the sums match numpy's.
template <typename T>
__global__ void monitor_accumulate(const T *__restrict__ E, int64_t N, int nf,
                                   const double *__restrict__ phase,
                                   double2 *__restrict__ acc) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  for (int c = 0; c < 3; ++c) {
    const double e = (double)E[c * N + n];
    for (int f = 0; f < nf; ++f) {
      const double pr = phase[2 * f], pi = phase[2 * f + 1];
      const double re = e * pr - 0.0 * pi;
      const double im = e * pi + 0.0 * pr;
      double2 *p = acc + ((int64_t)f * 3 + c) * N + n;
      double2 v = *p;
      v.x = v.x + re;
      v.y = v.y + im;
      *p = v;
    }
  }
}

// Longitudinal wake potential (wakis/wakeSolver.py:492-534): thread n owns
// sample n and sums over k in order, with separate multiply and add roundings
// (no contraction into an FMA) exactly as the reference's WP[n] = WP[n] +
// Ezt*dz, so that the same history gives the same bits.
template <typename T>
__global__ void wake_potential(const T *__restrict__ hist, int64_t hist_nt, int nzh,
                               const int32_t *__restrict__ wids, int nwin, int carry,
                               int ns, int nzw, int k0, int zrev,
                               const int32_t *__restrict__ cols,
                               const double *__restrict__ dz, double qnorm,
                               double *__restrict__ wp) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ns) return;
  // history row of term k: k0 + k, or walking z down from the top row of the
  // window (k0 + nzw - 1 - k) for the counter-moving integral
  const int r0 = zrev ? k0 + nzw - 1 : k0;
  const int rs = zrev ? -1 : 1;
  double acc = 0.0;
  for (int w = 0; w < nwin; ++w) {
    if (!carry) acc = 0.0;
    const T *hw = hist + (int64_t)wids[w] * nzh * hist_nt;
    for (int k = 0; k < nzw; ++k) {
      const int c = cols[(int64_t)k * ns + n];
      if (c >= 0) {
        const double e = (double)hw[(int64_t)(r0 + rs * k) * hist_nt + c];
        const double prod = e * dz[k];
        acc = acc + prod;
      }
    }
    acc = acc / qnorm;
    wp[(int64_t)w * ns + n] = acc;
  }
}

// Direct DFT bins j < K of x and lam (length len, zero-padded or truncated to
// N): X_j = sum_n x[n] exp(-2 pi i j n / N).  One workgroup per bin; exact
// integer argument reduction (j*n mod N) and sincospi for the twiddle.
__global__ void __launch_bounds__(256) impedance_dft(const double *__restrict__ x,
                                                     const double *__restrict__ lam,
                                                     int64_t len, int64_t N, double ds,
                                                     double *__restrict__ out, int64_t oj,
                                                     int64_t oq) {
  const int jb = blockIdx.x;
  const int64_t M = len < N ? len : N;
  double xr = 0, xi = 0, lr = 0, li = 0;
  for (int64_t n = threadIdx.x; n < M; n += blockDim.x) {
    const int64_t r = ((int64_t)jb * n) % N;
    double s, c;
    sincospi(2.0 * (double)r / (double)N, &s, &c);
    xr += x[n] * c;
    xi -= x[n] * s;
    lr += lam[n] * c;
    li -= lam[n] * s;
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = xr;
  red[1][threadIdx.x] = xi;
  red[2][threadIdx.x] = lr;
  red[3][threadIdx.x] = li;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double Xr = red[0][0] * ds, Xi = red[1][0] * ds;
    const double Lr = red[2][0] * ds, Li = red[3][0] * ds;
    // Z = -X / Lam
    const double den = Lr * Lr + Li * Li;
    const double zr = -(Xr * Lr + Xi * Li) / den;
    const double zi = -(Xi * Lr - Xr * Li) / den;
    // output q of bin jb at out[jb * oj + q * oq]: interleaved (oj = 6,
    // oq = 1) or one array per quantity (oj = 1, oq = K)
    double *o = out + (int64_t)jb * oj;
    o[0 * oq] = zr;
    o[1 * oq] = zi;
    o[2 * oq] = Xr;
    o[3 * oq] = Xi;
    o[4 * oq] = Lr;
    o[5 * oq] = Li;
  }
}

}  // namespace fit
