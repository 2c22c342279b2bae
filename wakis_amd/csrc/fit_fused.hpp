// Fused single-sweep leapfrog step: one launch reads E, H (old buffers) and
// writes E, H (new buffers) once per cell.
//
// Interior cells (1 <= i <= nx-2, 1 <= j <= ny-2, 1 <= k <= nz-2) are swept
// by z-marching workgroups: a workgroup owns a TX x TY column of cells and a
// chunk of KC planes; per plane it stages E_old of the plane above into LDS
// (tile plus a one-cell halo), computes H_new for the tile and for its low
// x/y halo (so no other workgroup's result is needed), keeps H_new of the
// previous plane in registers, and finishes E_new of the plane.  Neighbours
// of interior cells never wrap, so the interior path uses plain 3-D indexing;
// the face masks and the metric tables are still looked up per cell.
//
// Shell cells (any index on the first or last plane of an axis) take the
// generic path: H_new and the E_new of the cell are recomputed from the old
// buffers with the flat-index neighbour rule (helical wrap, dropped z
// neighbour), recomputing the H_new of the n-1, n-nx, n-nx*ny neighbours
// that the E update reads.  Both paths read only old buffers and write only
// new buffers, so they need no ordering between them.
//
// Blocks are remapped so that consecutive tiles land on the same XCD (the
// hardware dispatches block b to XCD b % 8), keeping halo re-reads in one L2.
// The interior sweep and the shell are two launches on the same stream (the
// shell's generic recomputation needs many more registers; a separate kernel
// keeps the interior sweep's occupancy).
#pragma once

#include "fit_common.hpp"

namespace fit {

constexpr int FUSED_TX = 64;
constexpr int FUSED_TY = 8;
constexpr int FUSED_KC = 32;
constexpr int FUSED_THREADS = FUSED_TX * FUSED_TY;
constexpr int SHELL_THREADS = 256;

template <typename T>
struct FusedArgs {
  StepArgs<T> a;   // a.E, a.H: old buffers
  T *En, *Hn;      // new buffers
  int ntx, nty, nkc;
  int nb_interior;
  int64_t nshell;
};

// ---- generic (flat-index) path -------------------------------------------

template <typename T>
__device__ __forceinline__ T h_new_generic(const StepArgs<T> &a, int comp, int i, int j, int k) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  // per component: the two E components and the shift axes of C's row
  //   x: u_y[n] - u_y[n+sz] - u_z[n] + u_z[n+sy]
  //   y: u_z[n] - u_z[n+1]  - u_x[n] + u_x[n+sz]
  //   z: u_x[n] - u_x[n+sy] - u_y[n] + u_y[n+1]
  const int ca = comp == 0 ? 1 : (comp == 1 ? 2 : 0);
  const int cb = comp == 0 ? 2 : (comp == 1 ? 0 : 1);
  const int sa = comp == 0 ? 2 : (comp == 1 ? 0 : 1);  // shift axis of ca
  const int sb = comp == 0 ? 1 : (comp == 1 ? 2 : 0);  // shift axis of cb
  const T a0 = u_at(a, ca, i, j, k);
  const T b0 = u_at(a, cb, i, j, k);
  T a1 = 0, b1 = 0;
  {
    int ii = i, jj = j, kk = k;
    if (nb_plus(sa, ii, jj, kk, nx, ny, nz)) a1 = u_at(a, ca, ii, jj, kk);
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_plus(sb, ii, jj, kk, nx, ny, nz)) b1 = u_at(a, cb, ii, jj, kk);
  }
  const int64_t n = flat(i, j, k, nx, ny);
  const int m = a.mat ? a.mat[n] : 0;
  T c = T(0);
  if (!masked(a.pmc, comp, i, j, k, nx, ny, nz)) {
    const T area = comp == 0 ? a.g.iAx[j + (int64_t)ny * k]
                             : (comp == 1 ? a.g.iAy[i + (int64_t)nx * k] : a.g.iAz[i + (int64_t)nx * j]);
    c = a.mtab[m * MAT_STRIDE + comp] * area;
  }
  const T curl = ((a0 - a1) - b0) + b1;
  return a.H[comp * a.N + n] - c * curl;
}

// w_d of the E update from a freshly computed H_new.
template <typename T>
__device__ __forceinline__ T w_new_generic(const StepArgs<T> &a, int comp, int i, int j, int k) {
  if (masked(a.pmc, comp, i, j, k, a.nx, a.ny, a.nz)) return T(0);
  return tLof(a.g, comp, i, j, k) * h_new_generic(a, comp, i, j, k);
}

template <typename T>
__device__ __forceinline__ T j_of(const StepArgs<T> &a, const T *mt, int comp, int64_t n, T e) {
  if (a.jmode == J_SIGMA_E) return mt[6 + comp] * e;
  if (a.jmode == J_BUFFER) return a.J[comp * a.N + n];
  return T(0);
}

template <typename T>
__device__ __forceinline__ void probe(const StepArgs<T> &a, int i, int j, int k, T ez) {
  if (a.hist && a.step < a.hist_nt) {
    const int di = i - a.ixt, dj = j - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw) {
      const int w = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
      a.hist[((int64_t)w * a.nz + k) * a.hist_nt + a.step] = ez;
    }
  }
}

template <typename T>
__device__ void shell_cell(const FusedArgs<T> &f, int i, int j, int k) {
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t n = flat(i, j, k, nx, ny);
  const T hx = h_new_generic(a, 0, i, j, k);
  const T hy = h_new_generic(a, 1, i, j, k);
  const T hz = h_new_generic(a, 2, i, j, k);
  f.Hn[n] = hx;
  f.Hn[a.N + n] = hy;
  f.Hn[2 * a.N + n] = hz;

  const T wx0 = masked(a.pmc, 0, i, j, k, nx, ny, nz) ? T(0) : a.g.tLx[i] * hx;
  const T wy0 = masked(a.pmc, 1, i, j, k, nx, ny, nz) ? T(0) : a.g.tLy[j] * hy;
  const T wz0 = masked(a.pmc, 2, i, j, k, nx, ny, nz) ? T(0) : a.g.tLz[k] * hz;
  T wy_mx = 0, wz_mx = 0, wx_my = 0, wz_my = 0, wx_mz = 0, wy_mz = 0;
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(0, ii, jj, kk, nx, ny, nz)) {
      wy_mx = w_new_generic(a, 1, ii, jj, kk);
      wz_mx = w_new_generic(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(1, ii, jj, kk, nx, ny, nz)) {
      wx_my = w_new_generic(a, 0, ii, jj, kk);
      wz_my = w_new_generic(a, 2, ii, jj, kk);
    }
  }
  {
    int ii = i, jj = j, kk = k;
    if (nb_minus(2, ii, jj, kk, nx, ny, nz)) {
      wx_mz = w_new_generic(a, 0, ii, jj, kk);
      wy_mz = w_new_generic(a, 1, ii, jj, kk);
    }
  }
  const int m = a.mat ? a.mat[n] : 0;
  const T *mt = a.mtab + m * MAT_STRIDE;
  const T ex = a.E[n], ey = a.E[a.N + n], ez = a.E[2 * a.N + n];
  probe(a, i, j, k, ez);
  const T jx = j_of(a, mt, 0, n, ex), jy = j_of(a, mt, 1, n, ey);
  T jz = j_of(a, mt, 2, n, ez);
  if (a.beam_mode != BEAM_OFF && i == a.ixs && j == a.iys)
    jz = (a.beam_mode == BEAM_VALUES) ? a.jline[k] : beam_gauss<T>(a, k);
  const T ax = masked(a.pec, 0, i, j, k, nx, ny, nz) ? T(0) : a.g.itAx[j + (int64_t)ny * k];
  const T ay = masked(a.pec, 1, i, j, k, nx, ny, nz) ? T(0) : a.g.itAy[i + (int64_t)nx * k];
  const T az = masked(a.pec, 2, i, j, k, nx, ny, nz) ? T(0) : a.g.itAz[i + (int64_t)nx * j];
  const T cx = ((wz0 - wz_my) - wy0) + wy_mz;
  const T cy = ((wx0 - wx_mz) - wz0) + wz_mx;
  const T cz = ((wy0 - wy_mx) - wx0) + wx_my;
  f.En[n] = ex + mt[3] * (ax * cx - jx);
  f.En[a.N + n] = ey + mt[4] * (ay * cy - jy);
  f.En[2 * a.N + n] = ez + mt[5] * (az * cz - jz);
}

// shell index -> (i, j, k): the planes k = 0 and k = nz-1, then for every
// inner plane the ring j = 0, j = ny-1, i = 0, i = nx-1.
__device__ __forceinline__ bool shell_ijk(int64_t t, int nx, int ny, int nz, int &i, int &j,
                                          int &k) {
  const int64_t plane = (int64_t)nx * ny;
  if (t < 2 * plane) {
    k = t < plane ? 0 : nz - 1;
    const int64_t r = t < plane ? t : t - plane;
    j = (int)(r / nx);
    i = (int)(r - (int64_t)j * nx);
    return true;
  }
  const int64_t ring = 2 * (int64_t)nx + 2 * (int64_t)(ny - 2);
  const int64_t r = t - 2 * plane;
  k = 1 + (int)(r / ring);
  if (k > nz - 2) return false;
  const int64_t p = r % ring;
  if (p < nx) {
    i = (int)p;
    j = 0;
  } else if (p < 2 * nx) {
    i = (int)(p - nx);
    j = ny - 1;
  } else {
    const int64_t q = p - 2 * nx;
    if (q < ny - 2) {
      i = 0;
      j = 1 + (int)q;
    } else {
      i = nx - 1;
      j = 1 + (int)(q - (ny - 2));
    }
  }
  return true;
}

// ---- interior (z-marching) path -----------------------------------------

template <typename T>
struct FusedSmem {
  T e[2][3][FUSED_TY + 2][FUSED_TX + 2];  // E_old planes (ring of 2), tile + halo
  T h[3][FUSED_TY + 1][FUSED_TX + 1];     // H_new of the plane, tile + low halo
};

template <typename T>
__device__ __forceinline__ void load_plane(const T *__restrict__ Ex, const T *__restrict__ Ey,
                                           const T *__restrict__ Ez, int nx, int ny,
                                           T (*dst)[FUSED_TY + 2][FUSED_TX + 2], int x0, int y0,
                                           int k) {
  constexpr int W = FUSED_TX + 2, Hh = FUSED_TY + 2;
  const int base = nx * ny * k;
#pragma unroll
  for (int r0 = 0; r0 < W * Hh; r0 += FUSED_THREADS) {
    const int r = r0 + threadIdx.x;
    if (r < W * Hh) {
      const int rj = r / W, ri = r - rj * W;
      const int gi = x0 - 1 + ri, gj = y0 - 1 + rj;
      T vx = 0, vy = 0, vz = 0;
      if (gi < nx && gj < ny) {
        const int o = base + gi + nx * gj;
        vx = Ex[o];
        vy = Ey[o];
        vz = Ez[o];
      }
      dst[0][rj][ri] = vx;
      dst[1][rj][ri] = vy;
      dst[2][rj][ri] = vz;
    }
  }
}

// u_d = L_d * pecmask_d * E_d from a staged plane at tile-local (li, lj); the
// generic form, used for the low-halo cells.
template <typename T>
__device__ __forceinline__ T u_s(const StepArgs<T> &a, const T (*pl)[FUSED_TY + 2][FUSED_TX + 2],
                                 int comp, int i, int j, int k, int li, int lj) {
  if (masked(a.pec, comp, i, j, k, a.nx, a.ny, a.nz)) return T(0);
  return Lof(a.g, comp, i, j, k) * pl[comp][lj][li];
}

// Generic H_new of a low-halo cell (i, j, k) from staged planes k, k+1.
template <typename T>
__device__ __forceinline__ T h_new_staged(const StepArgs<T> &a,
                                          const T (*pk)[FUSED_TY + 2][FUSED_TX + 2],
                                          const T (*pk1)[FUSED_TY + 2][FUSED_TX + 2], int comp,
                                          int i, int j, int k, int li, int lj, int m) {
  T curl;
  if (comp == 0) {
    const T uy0 = u_s(a, pk, 1, i, j, k, li, lj), uy1 = u_s(a, pk1, 1, i, j, k + 1, li, lj);
    const T uz0 = u_s(a, pk, 2, i, j, k, li, lj), uz1 = u_s(a, pk, 2, i, j + 1, k, li, lj + 1);
    curl = ((uy0 - uy1) - uz0) + uz1;
  } else if (comp == 1) {
    const T uz0 = u_s(a, pk, 2, i, j, k, li, lj), uz1 = u_s(a, pk, 2, i + 1, j, k, li + 1, lj);
    const T ux0 = u_s(a, pk, 0, i, j, k, li, lj), ux1 = u_s(a, pk1, 0, i, j, k + 1, li, lj);
    curl = ((uz0 - uz1) - ux0) + ux1;
  } else {
    const T ux0 = u_s(a, pk, 0, i, j, k, li, lj), ux1 = u_s(a, pk, 0, i, j + 1, k, li, lj + 1);
    const T uy0 = u_s(a, pk, 1, i, j, k, li, lj), uy1 = u_s(a, pk, 1, i + 1, j, k, li + 1, lj);
    curl = ((ux0 - ux1) - uy0) + uy1;
  }
  T c = T(0);
  if (!masked(a.pmc, comp, i, j, k, a.nx, a.ny, a.nz)) {
    const T area = comp == 0 ? a.g.iAx[j + a.ny * k]
                             : (comp == 1 ? a.g.iAy[i + a.nx * k] : a.g.iAz[i + a.nx * j]);
    c = a.mtab[m * MAT_STRIDE + comp] * area;
  }
  return a.H[(int64_t)comp * a.N + (i + a.nx * (j + a.ny * k))] - c * curl;
}

// The interior cell (i, j, k) has no face of its own, so its row masks and
// its own column masks are 1; only neighbours on the last/first plane of an
// axis can be masked.  With the reference's last-writer rule the relevant
// flags reduce to the per-thread booleans below.
template <typename T>
__device__ void interior_tile(const FusedArgs<T> &f, int tile, FusedSmem<T> &sm) {
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int N = (int)a.N;
  const int tx = tile % f.ntx;
  const int r = tile / f.ntx;
  const int ty = r % f.nty;
  const int kc = r / f.nty;
  const int x0 = 1 + tx * FUSED_TX, y0 = 1 + ty * FUSED_TY;
  const int k0 = 1 + kc * FUSED_KC;
  const int k1 = min(k0 + FUSED_KC, nz - 1);
  const int lx = threadIdx.x % FUSED_TX, ly = threadIdx.x / FUSED_TX;
  const int i = x0 + lx, j = y0 + ly;
  const bool own = i <= nx - 2 && j <= ny - 2;
  const T *__restrict__ Ex = a.E;
  const T *__restrict__ Ey = a.E + N;
  const T *__restrict__ Ez = a.E + 2 * (int64_t)N;
  const T *__restrict__ Hx = a.H;
  const T *__restrict__ Hy = a.H + N;
  const T *__restrict__ Hz = a.H + 2 * (int64_t)N;
  T *__restrict__ Enx = f.En;
  T *__restrict__ Eny = f.En + N;
  T *__restrict__ Enz = f.En + 2 * (int64_t)N;
  T *__restrict__ Hnx = f.Hn;
  T *__restrict__ Hny = f.Hn + N;
  T *__restrict__ Hnz = f.Hn + 2 * (int64_t)N;
  const uint8_t *__restrict__ mat = a.mat;
  const T *__restrict__ mtab = a.mtab;

  // low-halo work: threads 0..TX-1 -> row (x0+t, y0-1); TX..TX+TY-1 ->
  // column (x0-1, y0+t-TX)
  int hi = -1, hj = -1;
  if (threadIdx.x < FUSED_TX) {
    hi = x0 + threadIdx.x;
    hj = y0 - 1;
    if (hi > nx - 2) hi = -1;
  } else if (threadIdx.x < FUSED_TX + FUSED_TY) {
    hi = x0 - 1;
    hj = y0 + (threadIdx.x - FUSED_TX);
    if (hj > ny - 2) hi = -1;
  }
  const int hli = hi - x0 + 1, hlj = hj - y0 + 1;
  const int li = lx + 1, lj = ly + 1;

  // per-thread constants (clamped indices keep non-owned lanes in range)
  const int ic = min(i, nx - 2), jc = min(j, ny - 2);
  const T Lx0 = a.g.Lx[ic], Ly0 = a.g.Ly[jc];
  const bool mxh = (ic + 1 == nx - 1) && (a.pec & XHI);  // E at i+1 on the xhi face
  const bool myh = (jc + 1 == ny - 1) && (a.pec & YHI);  // E at j+1 on the yhi face
  const T uLy_x1 = mxh ? T(0) : Ly0;  // u_y(i+1) = L_y(j) * E_y(i+1)
  const T uLx_y1 = myh ? T(0) : Lx0;  // u_x(j+1) = L_x(i) * E_x(j+1)
  const bool pxl = (ic - 1 == 0) && (a.pmc & XLO);  // H at i-1 on the xlo face
  const bool pyl = (jc - 1 == 0) && (a.pmc & YLO);  // H at j-1 on the ylo face
  const T tLx0 = a.g.tLx[ic], tLy0 = a.g.tLy[jc];
  const T iAz0 = a.g.iAz[ic + nx * jc];
  const T itAz0 = a.g.itAz[ic + nx * jc];

  load_plane(Ex, Ey, Ez, nx, ny, sm.e[0], x0, y0, k0 - 1);
  load_plane(Ex, Ey, Ez, nx, ny, sm.e[1], x0, y0, k0);
  __syncthreads();
  T hx_prev = 0, hy_prev = 0;
  if (own) {
    const int m = mat ? mat[i + nx * (j + ny * (k0 - 1))] : 0;
    hx_prev = h_new_staged(a, sm.e[0], sm.e[1], 0, i, j, k0 - 1, li, lj, m);
    hy_prev = h_new_staged(a, sm.e[0], sm.e[1], 1, i, j, k0 - 1, li, lj, m);
  }
  int cur = 1;
  for (int k = k0; k < k1; ++k) {
    const int nxt = cur ^ 1;
    __syncthreads();  // plane k-1 and sm.h of plane k-1 no longer read
    load_plane(Ex, Ey, Ez, nx, ny, sm.e[nxt], x0, y0, k + 1);
    const int n = ic + nx * (jc + ny * k);
    const int m = (own && mat) ? mat[n] : 0;
    const T *mt = mtab + m * MAT_STRIDE;
    const T hox = own ? Hx[n] : T(0), hoy = own ? Hy[n] : T(0), hoz = own ? Hz[n] : T(0);
    const T Lz0 = a.g.Lz[k];
    const bool mzh = (k + 1 == nz - 1) && (a.pec & ZHI);  // E at k+1 on the zhi face
    const bool pzl = (k - 1 == 0) && (a.pmc & ZLO);       // H at k-1 on the zlo face
    const T iAx0 = a.g.iAx[jc + ny * k], iAy0 = a.g.iAy[ic + nx * k];
    const T itAx0 = a.g.itAx[jc + ny * k], itAy0 = a.g.itAy[ic + nx * k];
    const T tLz0 = a.g.tLz[k];
    __syncthreads();
    const T(*pk)[FUSED_TY + 2][FUSED_TX + 2] = sm.e[cur];
    const T(*pk1)[FUSED_TY + 2][FUSED_TX + 2] = sm.e[nxt];
    T hx = 0, hy = 0, hz = 0;
    if (own) {
      const T ex = pk[0][lj][li], ey = pk[1][lj][li], ez = pk[2][lj][li];
      const T ux0 = Lx0 * ex, uy0 = Ly0 * ey, uz0 = Lz0 * ez;
      const T uy_pz = mzh ? T(0) : Ly0 * pk1[1][lj][li];
      const T ux_pz = mzh ? T(0) : Lx0 * pk1[0][lj][li];
      const T uz_py = myh ? T(0) : Lz0 * pk[2][lj + 1][li];
      const T uz_px = mxh ? T(0) : Lz0 * pk[2][lj][li + 1];
      const T ux_py = uLx_y1 * pk[0][lj + 1][li];
      const T uy_px = uLy_x1 * pk[1][lj][li + 1];
      hx = hox - (mt[0] * iAx0) * (((uy0 - uy_pz) - uz0) + uz_py);
      hy = hoy - (mt[1] * iAy0) * (((uz0 - uz_px) - ux0) + ux_pz);
      hz = hoz - (mt[2] * iAz0) * (((ux0 - ux_py) - uy0) + uy_px);
      Hnx[n] = hx;
      Hny[n] = hy;
      Hnz[n] = hz;
    }
    sm.h[0][lj][li] = hx;
    sm.h[1][lj][li] = hy;
    sm.h[2][lj][li] = hz;
    if (hi >= 0) {
      const int hn = hi + nx * (hj + ny * k);
      const int hm = mat ? mat[hn] : 0;
      if (threadIdx.x < FUSED_TX) {  // row y0-1 feeds w_x and w_z
        sm.h[0][hlj][hli] = h_new_staged(a, pk, pk1, 0, hi, hj, k, hli, hlj, hm);
        sm.h[2][hlj][hli] = h_new_staged(a, pk, pk1, 2, hi, hj, k, hli, hlj, hm);
      } else {  // column x0-1 feeds w_y and w_z
        sm.h[1][hlj][hli] = h_new_staged(a, pk, pk1, 1, hi, hj, k, hli, hlj, hm);
        sm.h[2][hlj][hli] = h_new_staged(a, pk, pk1, 2, hi, hj, k, hli, hlj, hm);
      }
    }
    __syncthreads();
    if (own) {
      const T wx0 = tLx0 * hx, wy0 = tLy0 * hy, wz0 = tLz0 * hz;
      const T wy_mx = pxl ? T(0) : tLy0 * sm.h[1][lj][li - 1];
      const T wz_mx = pxl ? T(0) : tLz0 * sm.h[2][lj][li - 1];
      const T wx_my = pyl ? T(0) : tLx0 * sm.h[0][lj - 1][li];
      const T wz_my = pyl ? T(0) : tLz0 * sm.h[2][lj - 1][li];
      const T wx_mz = pzl ? T(0) : tLx0 * hx_prev;
      const T wy_mz = pzl ? T(0) : tLy0 * hy_prev;
      const T ex = pk[0][lj][li], ey = pk[1][lj][li], ez = pk[2][lj][li];
      probe(a, i, j, k, ez);
      T jx = 0, jy = 0, jz = 0;
      if (a.jmode == J_SIGMA_E) {
        jx = mt[6] * ex;
        jy = mt[7] * ey;
        jz = mt[8] * ez;
      } else if (a.jmode == J_BUFFER) {
        jx = a.J[n];
        jy = (a.J + N)[n];
        jz = (a.J + 2 * (int64_t)N)[n];
      }
      if (a.beam_mode != BEAM_OFF && i == a.ixs && j == a.iys)
        jz = (a.beam_mode == BEAM_VALUES) ? a.jline[k] : beam_gauss<T>(a, k);
      const T cx = ((wz0 - wz_my) - wy0) + wy_mz;
      const T cy = ((wx0 - wx_mz) - wz0) + wz_mx;
      const T cz = ((wy0 - wy_mx) - wx0) + wx_my;
      Enx[n] = ex + mt[3] * (itAx0 * cx - jx);
      Eny[n] = ey + mt[4] * (itAy0 * cy - jy);
      Enz[n] = ez + mt[5] * (itAz0 * cz - jz);
    }
    hx_prev = hx;
    hy_prev = hy;
    cur = nxt;
  }
}

template <typename T>
__global__ void __launch_bounds__(FUSED_THREADS) fused_sweep(FusedArgs<T> f) {
  __shared__ FusedSmem<T> sm;
  const int b = blockIdx.x;
  // XCD-aware remap: blocks b, b+8, b+16, ... share an XCD; give them
  // consecutive tiles
  const int q = f.nb_interior / 8;
  const int tile = (b < 8 * q) ? (b % 8) * q + b / 8 : b;
  interior_tile(f, tile, sm);
}

template <typename T>
__global__ void __launch_bounds__(SHELL_THREADS) fused_shell(FusedArgs<T> f) {
  const int64_t t = (int64_t)blockIdx.x * SHELL_THREADS + threadIdx.x;
  if (t >= f.nshell) return;
  int i, j, k;
  if (shell_ijk(t, f.a.nx, f.a.ny, f.a.nz, i, j, k)) shell_cell(f, i, j, k);
}

inline bool fused_supported(int kernel, int nx, int ny, int nz) {
  (void)nx;
  (void)ny;
  (void)nz;
  return kernel != 1;  // FIT_KERNEL_SPLIT selects the split kernels
}

inline int64_t shell_count(int nx, int ny, int nz) {
  const int64_t plane = (int64_t)nx * ny;
  const int64_t inner = nz > 2 ? (int64_t)(nz - 2) * (2 * (int64_t)nx + 2 * (int64_t)(ny - 2)) : 0;
  return (nz >= 2 ? 2 * plane : plane) + inner;
}

template <typename T>
inline void launch_fused(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s,
                         hipEvent_t shell_ev = nullptr) {
  FusedArgs<T> f;
  f.a = a;
  f.En = En;
  f.Hn = Hn;
  const int ix = a.nx - 2, iy = a.ny - 2, iz = a.nz - 2;
  if (ix > 0 && iy > 0 && iz > 0) {
    f.ntx = (ix + FUSED_TX - 1) / FUSED_TX;
    f.nty = (iy + FUSED_TY - 1) / FUSED_TY;
    f.nkc = (iz + FUSED_KC - 1) / FUSED_KC;
  } else {
    f.ntx = f.nty = f.nkc = 0;
  }
  f.nb_interior = f.ntx * f.nty * f.nkc;
  f.nshell = shell_count(a.nx, a.ny, a.nz);
  const int nb_shell = (int)((f.nshell + SHELL_THREADS - 1) / SHELL_THREADS);
  if (f.nb_interior > 0)
    hipLaunchKernelGGL(fused_sweep<T>, dim3(f.nb_interior), dim3(FUSED_THREADS), 0, s, f);
  if (shell_ev) (void)hipEventRecord(shell_ev, s);
  hipLaunchKernelGGL(fused_shell<T>, dim3(nb_shell), dim3(SHELL_THREADS), 0, s, f);
}

}  // namespace fit
