// Split step: one H-update kernel, then one E-update kernel, in place.
//
// H update (rows of M_H = iDa*iDmu*C*Ds, wakis/solverFIT3D.py:343, 510-512):
//   H_x -= dt/mu_x * iA_x * pmc_x * ( u_y[n] - u_y[n+sz] - u_z[n] + u_z[n+sy] )
//   H_y -= dt/mu_y * iA_y * pmc_y * ( u_z[n] - u_z[n+1]  - u_x[n] + u_x[n+sz] )
//   H_z -= dt/mu_z * iA_z * pmc_z * ( u_x[n] - u_x[n+sy] - u_y[n] + u_y[n+1]  )
//   with u_d = L_d * pec_d * E_d.
// E update (rows of M_E = iDeps*itDa*C^T*tDs, wakis/solverFIT3D.py:344, 514-521):
//   E_x += dt/eps_x * ( pec_x * itA_x * ( w_z[n] - w_z[n-sy] - w_y[n] + w_y[n-sz] ) - J_x )
//   E_y += dt/eps_y * ( pec_y * itA_y * ( w_x[n] - w_x[n-sz] - w_z[n] + w_z[n-1]  ) - J_y )
//   E_z += dt/eps_z * ( pec_z * itA_z * ( w_y[n] - w_y[n-1]  - w_x[n] + w_x[n-sy] ) - J_z )
//   with w_d = tL_d * pmc_d * H_d and J per the J mode (J_SIGMA_E: J = sigma*E
//   of the state before this update, which is what the reference's
//   J = sigma*E (solverFIT3D.py:524-525) holds at that point).
// These kernels handle every cell with the generic flat-index neighbour rule;
// they are the reference implementation of the operator on the device and
// serve any grid size.
#pragma once

#include "fit_common.hpp"

namespace fit {

// Launch: block (SPLIT_BX, SPLIT_BY), grid (ceil(nx/BX), ceil(ny/BY), nz).
constexpr int SPLIT_BX = 64;
constexpr int SPLIT_BY = 4;

template <typename T>
__global__ void __launch_bounds__(SPLIT_BX *SPLIT_BY) h_update_split(StepArgs<T> a) {
  const int i = blockIdx.x * SPLIT_BX + threadIdx.x;
  const int j = blockIdx.y * SPLIT_BY + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  if (i >= nx || j >= ny) return;
  const int64_t n = flat(i, j, k, nx, ny);

  const T ux0 = u_at(a, 0, i, j, k), uy0 = u_at(a, 1, i, j, k), uz0 = u_at(a, 2, i, j, k);
  T uy_px = 0, uz_px = 0, ux_py = 0, uz_py = 0, ux_pz = 0, uy_pz = 0;
  {
    int ii = i, jj = j, kk = k;
    if (nb_plus(0, ii, jj, kk, nx, ny, nz)) {
      uy_px = u_at(a, 1, ii, jj, kk);
      uz_px = u_at(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_plus(1, ii, jj, kk, nx, ny, nz)) {
      ux_py = u_at(a, 0, ii, jj, kk);
      uz_py = u_at(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_plus(2, ii, jj, kk, nx, ny, nz)) {
      ux_pz = u_at(a, 0, ii, jj, kk);
      uy_pz = u_at(a, 1, ii, jj, kk);
    }
  }
  const T *mt = mat_row(a, n);
  const T cx = masked(a.pmc, 0, i, j, k, nx, ny, nz) ? T(0) : mt[0] * a.g.iAx[j + (int64_t)ny * k];
  const T cy = masked(a.pmc, 1, i, j, k, nx, ny, nz) ? T(0) : mt[1] * a.g.iAy[i + (int64_t)nx * k];
  const T cz = masked(a.pmc, 2, i, j, k, nx, ny, nz) ? T(0) : mt[2] * a.g.iAz[i + (int64_t)nx * j];
  const T curl_x = ((uy0 - uy_pz) - uz0) + uz_py;
  const T curl_y = ((uz0 - uz_px) - ux0) + ux_pz;
  const T curl_z = ((ux0 - ux_py) - uy0) + uy_px;
  T *H = a.H;
  H[n] = H[n] - cx * curl_x;
  H[a.N + n] = H[a.N + n] - cy * curl_y;
  H[2 * a.N + n] = H[2 * a.N + n] - cz * curl_z;
}

template <typename T>
__global__ void __launch_bounds__(SPLIT_BX *SPLIT_BY) e_update_split(StepArgs<T> a) {
  const int i = blockIdx.x * SPLIT_BX + threadIdx.x;
  const int j = blockIdx.y * SPLIT_BY + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  if (i >= nx || j >= ny) return;
  const int64_t n = flat(i, j, k, nx, ny);

  const T wx0 = w_at(a, 0, i, j, k), wy0 = w_at(a, 1, i, j, k), wz0 = w_at(a, 2, i, j, k);
  T wy_mx = 0, wz_mx = 0, wx_my = 0, wz_my = 0, wx_mz = 0, wy_mz = 0;
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(0, ii, jj, kk, nx, ny, nz)) {
      wy_mx = w_at(a, 1, ii, jj, kk);
      wz_mx = w_at(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(1, ii, jj, kk, nx, ny, nz)) {
      wx_my = w_at(a, 0, ii, jj, kk);
      wz_my = w_at(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(2, ii, jj, kk, nx, ny, nz)) {
      wx_mz = w_at(a, 0, ii, jj, kk);
      wy_mz = w_at(a, 1, ii, jj, kk);
    }
  }
  const T *mt = mat_row(a, n);
  T *E = a.E;
  const T ex = E[n], ey = E[a.N + n], ez = E[2 * a.N + n];

  // wake probe: E_z of the state before this update (wakis/routines.py:443)
  if (a.hist && a.step < a.hist_nt) {
    const int di = i - a.ixt, dj = j - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw) {
      const int w = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
      a.hist[((int64_t)w * nz + k) * a.hist_nt + a.step] = ez;
    }
  }

  T jx = 0, jy = 0, jz = 0;
  if (a.jmode == J_SIGMA_E) {
    jx = mt[6] * ex;
    jy = mt[7] * ey;
    jz = mt[8] * ez;
  } else if (a.jmode == J_BUFFER) {
    jx = a.J[n];
    jy = a.J[a.N + n];
    jz = a.J[2 * a.N + n];
  }
  if (a.beam_mode != BEAM_OFF && i == a.ixs && j == a.iys) jz = a.jline[k];

  const T ax = masked(a.pec, 0, i, j, k, nx, ny, nz) ? T(0) : a.g.itAx[j + (int64_t)ny * k];
  const T ay = masked(a.pec, 1, i, j, k, nx, ny, nz) ? T(0) : a.g.itAy[i + (int64_t)nx * k];
  const T az = masked(a.pec, 2, i, j, k, nx, ny, nz) ? T(0) : a.g.itAz[i + (int64_t)nx * j];
  const T cx = ((wz0 - wz_my) - wy0) + wy_mz;
  const T cy = ((wx0 - wx_mz) - wz0) + wz_mx;
  const T cz = ((wy0 - wy_mx) - wx0) + wx_my;
  E[n] = ex + mt[3] * (ax * cx - jx);
  E[a.N + n] = ey + mt[4] * (ay * cy - jy);
  E[2 * a.N + n] = ez + mt[5] * (az * cz - jz);
}

}  // namespace fit

namespace fit {

// Plain streaming copy (calibration of the achievable HBM bandwidth and of
// the rocprof byte counters): 16-byte vector loads and stores.
__global__ void __launch_bounds__(256) stream_copy(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, int64_t n4) {
  // One 16-byte word per thread and no grid-stride loop: on gfx950 this form
  // reaches about 6.3 TB/s, a grid-stride loop only 4.6-5.5 TB/s.
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n4) dst[t] = src[t];
}

// Copy of six arrays into six others (the twelve field streams of a step)
// with the fused sweep's traversal: a block per (tx x ty columns, kc planes)
// tile, marching in z.  Calibrates what a plain copy reaches with the same
// access pattern as the sweep.
template <typename T>
struct Copy6 {
  const T *a[6];
  T *b[6];
};

template <typename T>
__global__ void __launch_bounds__(1024) copy_zmarch(Copy6<T> p, int nx, int ny, int nz, int tx,
                                                    int ty, int kc) {
  const int ntx = (nx + tx - 1) / tx, nty = (ny + ty - 1) / ty;
  const int t = blockIdx.x;
  const int bx = t % ntx, r = t / ntx, by = r % nty, bz = r / nty;
  const int i = bx * tx + (int)threadIdx.x % tx, j = by * ty + (int)threadIdx.x / tx;
  if (i >= nx || j >= ny) return;
  const int64_t plane = (int64_t)nx * ny;
  const int64_t col = i + (int64_t)nx * j;
  const int k1 = min(nz, (bz + 1) * kc);
  for (int k = bz * kc; k < k1; ++k) {
    const int64_t n = col + plane * k;
    T v[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = p.a[c][n];
#pragma unroll
    for (int c = 0; c < 6; ++c) p.b[c][n] = v[c];
  }
}

}  // namespace fit
