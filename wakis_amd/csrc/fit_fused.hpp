// Fused single-sweep leapfrog step: one launch reads E, H (old buffers) and
// writes E, H (new buffers), every field word once per cell.
//
// A workgroup owns a TX x TY tile of (i, j) columns and a chunk of KC planes
// and marches up in z.  It stages a window of E_old in LDS: window cell
// (gi, gj) of plane k, gi in [x0-1, x0+TX], gj in [y0-1, y0+TY], holds the
// field at FLAT index gi + nx*gj + nx*ny*k (0 outside [0, N)).  Because the
// reference's curl is built from flat-index shift matrices
// (wakis/solverFIT3D.py:207-227), every stencil neighbour n+-1, n+-nx,
// n+-nx*ny of a cell is the adjacent window cell: the helical x/y wrap and the
// dropped neighbours beyond the array ends come out exactly, for every cell
// of the grid, boundary or not.  Each staged value is stored pre-scaled with
// the metric and boundary mask of the cell it really is (u = L*pec*E), so the
// stencil arithmetic is uniform.
//
// Per plane k (one barrier per plane):
//   1. scale and store the prefetched E_old plane k+2 into the LDS ring
//      (4 slots; the slot overwritten held plane k-2, last read before the
//      previous iteration's barrier);
//   2. issue the global loads of plane k+3 and of H_old / material / E_old
//      of the next planes (consumed one iteration later);
//   3. barrier;
//   4. E_new(k) from E_old(k), w = tL*pmc*H_new of plane k (LDS: tile plus
//      the low x/y halo, which the workgroup recomputes itself) and of plane
//      k-1 (registers);
//   5. H_new(k+1) of the tile and its low halo into the other w buffer.
// Blocks are remapped so that consecutive tiles share an XCD (the hardware
// dispatches block b to XCD b % 8), keeping halo re-reads in one L2.
#pragma once

#include "fit_common.hpp"

namespace fit {

template <typename T>
struct FusedArgs {
  StepArgs<T> a;  // a.E, a.H: old buffers
  T *En, *Hn;     // new buffers
  int ntx, nty, nkc;
  int nb;
};

// Boundary mask of component c at true cell coordinates, split into the part
// fixed by (i, j) and the z faces (last writer wins: z faces are written
// last, then y, then x; wakis/boundaries.py:90-100).
__device__ __forceinline__ bool xy_mask(uint32_t m, int c, int i, int j, int nx, int ny) {
  if (c != 1) {
    if (j == ny - 1) return (m & YHI) != 0;
    if (j == 0) return (m & YLO) != 0;
  }
  if (c != 0) {
    if (i == nx - 1) return (m & XHI) != 0;
    if (i == 0) return (m & XLO) != 0;
  }
  return false;
}

__device__ __forceinline__ bool full_mask(uint32_t m, int c, bool xy, int k, int nz) {
  if (c != 2) {
    if (k == nz - 1) return (m & ZHI) != 0;
    if (k == 0) return (m & ZLO) != 0;
  }
  return xy;
}

template <int TX, int TY>
struct Tile {
  static constexpr int NT = TX * TY;
  static constexpr int RW = TX + 2;
  static constexpr int RH = TY + 2;
  static constexpr int RC = RW * RH;
  static constexpr int RL = (RC + NT - 1) / NT;
  static constexpr int NRING = 4;
};

template <typename T, int TX, int TY>
struct FusedSmem {
  using G = Tile<TX, TY>;
  T u[G::NRING][3][G::RH][G::RW];   // u = L*pec*E_old, window planes (ring)
  T w[2][3][TY + 1][TX + 1];        // w = tL*pmc*H_new, tile + low halo
  T mt[MAX_MATERIALS * MAT_STRIDE];  // material table (dt/mu, dt/eps, sigma)
};

// A window cell's true coordinates, fixed over the march (dk: plane offset
// from the y carry).
struct WinCell {
  int i, j, dk;
  bool ok;
};

__device__ __forceinline__ WinCell win_cell(int gi, int gj, int nx, int ny) {
  WinCell w;
  const int cx = gi >= 0 ? gi / nx : -((-gi + nx - 1) / nx);
  w.i = gi - cx * nx;
  const int jj = gj + cx;
  const int cy = jj >= 0 ? jj / ny : -((-jj + ny - 1) / ny);
  w.j = jj - cy * ny;
  w.dk = cy;
  w.ok = true;
  return w;
}

// After the barrier of a plane the kernel reads only LDS and registers: every
// global load it needs there (own and halo H_old/E_old/material, the 2-D
// metric entries that vary along x, the staged plane and its L_z) is issued
// one plane earlier, so the in-order completion of vector loads never makes
// the compute of a plane wait for the prefetch of the next one.  The material
// table lives in LDS; metric entries that are uniform over a wave (one row j
// per wave, one plane k per iteration) are scalar loads.
// One tile (TX x TY columns, KC planes) of one step; `tile` is the tile
// index in this grid's tile order.  Shared by the single-grid launch and the
// batched launch, so a grid's arithmetic is the same alone and in a batch.
template <typename T, int TX, int TY, int KC>
__device__ __forceinline__ void fused_tile(const FusedArgs<T> &f, const int tile,
                                           FusedSmem<T, TX, TY> &sm) {
  using G = Tile<TX, TY>;
  static_assert(TX % 64 == 0, "a wave must cover one tile row");
  constexpr int NT = G::NT, RW = G::RW, RH = G::RH, RC = G::RC, RL = G::RL, NR = G::NRING;
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int N = (int)a.N;
  const int plane = nx * ny;

  const int tx = tile % f.ntx;
  const int tr = tile / f.ntx;
  const int ty = tr % f.nty;
  const int kc = tr / f.nty;
  const int x0 = tx * TX, y0 = ty * TY;
  const int k0 = kc * KC;
  const int k1 = min(k0 + KC, nz);  // planes k0 .. k1-1

  // Field and table pointers typed as global memory (address space 1): in
  // the batched launch they are read from memory, and a generic pointer
  // would make every access a flat instruction.
  const GT<T> *__restrict__ Ex = (const GT<T> *)a.E;
  const GT<T> *__restrict__ Ey = Ex + N;
  const GT<T> *__restrict__ Ez = Ex + 2 * (int64_t)N;
  const GT<T> *__restrict__ Hx = (const GT<T> *)a.H;
  const GT<T> *__restrict__ Hy = Hx + N;
  const GT<T> *__restrict__ Hz = Hx + 2 * (int64_t)N;
  GT<T> *__restrict__ Enx = (GT<T> *)f.En;
  GT<T> *__restrict__ Eny = Enx + N;
  GT<T> *__restrict__ Enz = Enx + 2 * (int64_t)N;
  GT<T> *__restrict__ Hnx = (GT<T> *)f.Hn;
  GT<T> *__restrict__ Hny = Hnx + N;
  GT<T> *__restrict__ Hnz = Hnx + 2 * (int64_t)N;
  const GT<T> *__restrict__ Jx = (const GT<T> *)a.J;
  const GT<uint8_t> *__restrict__ mat = (const GT<uint8_t> *)a.mat;
  const Geo<T> &g = a.g;
  // read-only tables in the constant address space: wave-uniform entries
  // (one row j per wave, one plane per iteration) are scalar loads
  const CT<T> *gLz = (const CT<T> *)g.Lz;
  const CT<T> *gtLz = (const CT<T> *)g.tLz;
  const CT<T> *giAx = (const CT<T> *)g.iAx;
  const GT<T> *__restrict__ giAy = (const GT<T> *)g.iAy;
  const CT<T> *gitAx = (const CT<T> *)g.itAx;
  const GT<T> *__restrict__ gitAy = (const GT<T> *)g.itAy;
  const CT<T> *jline = (const CT<T> *)a.jline;
  const CT<T> *mtab = (const CT<T> *)a.mtab;
  const uint32_t pec = a.pec, pmc = a.pmc;

  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += NT) sm.mt[t] = mtab[t];

  // ---- staging cells of this thread (window slots q) ----
  // Only window cells that an active lane reads are loaded: columns up to
  // gi = nx, rows up to gj = ny; their plane carry is -1, 0 or 1.  The others
  // stage 0 and point at the cell (x0, y0), inside the grid on every plane.
  WinCell sc[RL];
  bool smx[RL][3];  // (i, j) part of the pec mask of each staged cell
  T sL[RL][2];      // L_x(i'), L_y(j')
  const int ri_max = min(TX + 1, nx - x0 + 1), rj_max = min(TY + 1, ny - y0 + 1);
#pragma unroll
  for (int q = 0; q < RL; ++q) {
    const int c = threadIdx.x + q * NT;
    const int rj = c / RW, ri = c - rj * RW;
    const bool need = c < RC && ri <= ri_max && rj <= rj_max;
    sc[q] = need ? win_cell(x0 - 1 + ri, y0 - 1 + rj, nx, ny) : WinCell{x0, y0, 0, false};
    sc[q].ok = need;
#pragma unroll
    for (int d = 0; d < 3; ++d) smx[q][d] = xy_mask(pec, d, sc[q].i, sc[q].j, nx, ny);
    sL[q][0] = g.Lx[sc[q].i];
    sL[q][1] = g.Ly[sc[q].j];
  }
  // raw E of a window plane plus L_z of each staged cell.  On planes 1 ..
  // nz-2 every address is inside the grid and the loads are unconditional
  // (no per-lane branch, so the compiler never waits for the loads in
  // flight to reuse their registers); near the z ends they are checked.
  auto fetch = [&](int k, T(&v)[RL][4]) {
    const bool inner = k >= 1 && k <= nz - 2;
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int kk = k + sc[q].dk;
      if (inner) {
        const int o = sc[q].i + nx * sc[q].j + plane * kk;
        v[q][0] = Ex[o];
        v[q][1] = Ey[o];
        v[q][2] = Ez[o];
        v[q][3] = gLz[kk];
      } else {
        T vx = 0, vy = 0, vz = 0, lz = 0;
        if (sc[q].ok && kk >= 0 && kk < nz) {
          const int o = sc[q].i + nx * sc[q].j + plane * kk;
          vx = Ex[o];
          vy = Ey[o];
          vz = Ez[o];
          lz = gLz[kk];
        }
        v[q][0] = vx;
        v[q][1] = vy;
        v[q][2] = vz;
        v[q][3] = lz;
      }
    }
  };
  auto stage = [&](int k, const T(&v)[RL][4]) {
    T(*dst)[RH][RW] = sm.u[k & (NR - 1)];
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int c = threadIdx.x + q * NT;
      if (c < RC) {
        const int rj = c / RW, ri = c - rj * RW;
        const int kk = k + sc[q].dk;
        const bool z = !sc[q].ok;
        dst[0][rj][ri] = (z || full_mask(pec, 0, smx[q][0], kk, nz)) ? T(0) : sL[q][0] * v[q][0];
        dst[1][rj][ri] = (z || full_mask(pec, 1, smx[q][1], kk, nz)) ? T(0) : sL[q][1] * v[q][1];
        dst[2][rj][ri] = (z || full_mask(pec, 2, smx[q][2], kk, nz)) ? T(0) : v[q][3] * v[q][2];
      }
    }
  };

  // ---- own cell ----
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int i = x0 + lx, j = y0 + ly;
  const bool own = i < nx && j < ny;
  const int ic = min(i, nx - 1);
  // one tile row per wave: j is wave-uniform, so row-indexed metrics are
  // scalar loads
  const int jc = __builtin_amdgcn_readfirstlane(min(j, ny - 1));
  const int li = lx + 1, lj = ly + 1;
  const int col = ic + nx * jc;
  bool opec[3], opmc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    opec[d] = xy_mask(pec, d, ic, jc, nx, ny);
    opmc[d] = xy_mask(pmc, d, ic, jc, nx, ny);
  }
  const T tLx0 = g.tLx[ic], tLy0 = g.tLy[jc];
  const T iAz0 = g.iAz[ic + nx * jc];
  const T itAz0 = g.itAz[ic + nx * jc];
  const bool beam_col = a.beam_mode != BEAM_OFF && i == a.ixs && j == a.iys;

  // ---- low-halo cell of this thread: row (x0+t, y0-1) -> H_x, H_z;
  //      column (x0-1, y0+t-TX) -> H_y, H_z ----
  const bool hrow = threadIdx.x < TX;
  const bool hcolm = !hrow && threadIdx.x < TX + TY;
  int hli = 0, hlj = 0;
  WinCell hc{0, 0, 0, false};
  if (hrow) {
    hli = threadIdx.x + 1;
    hlj = 0;
    hc = win_cell(x0 + threadIdx.x, y0 - 1, nx, ny);
    hc.ok = x0 + (int)threadIdx.x < nx;  // needed by an own cell only then
  } else if (hcolm) {
    hli = 0;
    hlj = threadIdx.x - TX + 1;
    hc = win_cell(x0 - 1, y0 + ((int)threadIdx.x - TX), nx, ny);
    hc.ok = y0 + ((int)threadIdx.x - TX) < ny;
  }
  const int hc0 = hrow ? 0 : 1;  // first halo component (second is z)
  const int hcolo = hc.i + nx * hc.j;
  bool hpmc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) hpmc[d] = xy_mask(pmc, d, hc.i, hc.j, nx, ny);
  const T h_tla = hrow ? g.tLx[hc.i] : g.tLy[hc.j];  // tL of the first halo component
  const T h_iAz = g.iAz[hc.i + nx * hc.j];

  // Halo prefetch packet of plane kk: H_old (2 comps), material, area of the
  // first component, tL_z.
  struct HaloPk {
    T ha, hb, area_a, tlz;
    int m;
    bool in;
  };
  // The waves holding halo threads (rows: waves 0 .. TX/64-1, column: wave
  // TX/64) load for every lane, at a clamped plane; p.in marks the values
  // that are used.  No per-lane branch around the loads.
  const bool hwave = (int)(threadIdx.x >> 6) <= TX / 64;  // wave-uniform
  auto halo_load = [&](int kk, HaloPk &p) {
    const int hk = kk + hc.dk;
    p.in = (hrow || hcolm) && hc.ok && hk >= 0 && hk < nz;
    if (hwave) {
      const int hkc = min(max(hk, 0), nz - 1);
      const int hn = hcolo + plane * hkc;
      p.hb = Hz[hn];
      p.m = mat ? mat[hn] : 0;
      p.tlz = gtLz[hkc];
      if (hrow) {  // wave-uniform (TX is a multiple of 64)
        p.ha = Hx[hn];
        p.area_a = giAx[hc.j + ny * hkc];
      } else {
        p.ha = Hy[hn];
        p.area_a = giAy[hc.i + nx * hkc];
      }
    } else {
      p.ha = p.hb = p.area_a = p.tlz = 0;
      p.m = 0;
    }
  };
  // curl of H component d at window (wi, wj) from slots p0 (plane) / p1 (above)
  auto curl_u = [&](int d, int p0, int p1, int wi, int wj) -> T {
    const T(*u0)[RH][RW] = sm.u[p0];
    const T(*u1)[RH][RW] = sm.u[p1];
    if (d == 0) return ((u0[1][wj][wi] - u1[1][wj][wi]) - u0[2][wj][wi]) + u0[2][wj + 1][wi];
    if (d == 1) return ((u0[2][wj][wi] - u0[2][wj][wi + 1]) - u0[0][wj][wi]) + u1[0][wj][wi];
    return ((u0[0][wj][wi] - u0[0][wj + 1][wi]) - u0[1][wj][wi]) + u0[1][wj][wi + 1];
  };
  auto halo_w = [&](int kk, int p0, int p1, int s, const HaloPk &p) {
    T wa = 0, wb = 0;
    if (p.in) {
      const int hk = kk + hc.dk;
      const bool ra = full_mask(pmc, hc0, hpmc[hc0], hk, nz);
      const bool rb = full_mask(pmc, 2, hpmc[2], hk, nz);
      const T *mt = sm.mt + p.m * MAT_STRIDE;
      const T ha = p.ha - (ra ? T(0) : mt[hc0] * p.area_a) * curl_u(hc0, p0, p1, hli, hlj);
      const T hb = p.hb - (rb ? T(0) : mt[2] * h_iAz) * curl_u(2, p0, p1, hli, hlj);
      wa = ra ? T(0) : h_tla * ha;
      wb = rb ? T(0) : p.tlz * hb;
    }
    sm.w[s][hc0][hlj][hli] = wa;
    sm.w[s][2][hlj][hli] = wb;
  };

  // ---- prologue: planes k0-1, k0, k0+1 staged; w of planes k0-1 and k0 ----
  T pE[RL][4];
  {
    fetch(k0 - 1, pE);
    stage(k0 - 1, pE);
    fetch(k0, pE);
    stage(k0, pE);
    fetch(k0 + 1, pE);
    stage(k0 + 1, pE);
  }
  if (k0 + 2 <= k1) fetch(k0 + 2, pE);
  T eox = 0, eoy = 0, eoz = 0, itAy_k = 0;
  int mcur = 0;
  T hx0o = 0, hy0o = 0, hz0o = 0, iAy0 = 0;
  T hxmo = 0, hymo = 0, iAym = 0;
  int mm = 0;
  if (own) {
    const int n0 = col + plane * k0;
    eox = Ex[n0];
    eoy = Ey[n0];
    eoz = Ez[n0];
    itAy_k = gitAy[ic + nx * k0];
    mcur = mat ? mat[n0] : 0;
    hx0o = Hx[n0];
    hy0o = Hy[n0];
    hz0o = Hz[n0];
    iAy0 = giAy[ic + nx * k0];
    if (k0 >= 1) {
      const int nm = n0 - plane;
      hxmo = Hx[nm];
      hymo = Hy[nm];
      mm = mat ? mat[nm] : 0;
      iAym = giAy[ic + nx * (k0 - 1)];
    }
  }
  HaloPk hp0;
  halo_load(k0, hp0);
  __syncthreads();
  T wxp = 0, wyp = 0;        // own w of plane k-1 (x, y)
  T wx = 0, wy = 0, wz = 0;  // own w of plane k
  {
    const int s = k0 & 1;
    const int pm = (k0 - 1) & (NR - 1), p0 = k0 & (NR - 1), p1 = (k0 + 1) & (NR - 1);
    if (own && k0 >= 1) {
      const int km = k0 - 1;
      const bool r0 = full_mask(pmc, 0, opmc[0], km, nz);
      const bool r1 = full_mask(pmc, 1, opmc[1], km, nz);
      const T *mt = sm.mt + mm * MAT_STRIDE;
      const T hxm = hxmo - (r0 ? T(0) : mt[0] * giAx[jc + ny * km]) * curl_u(0, pm, p0, li, lj);
      const T hym = hymo - (r1 ? T(0) : mt[1] * iAym) * curl_u(1, pm, p0, li, lj);
      wxp = r0 ? T(0) : tLx0 * hxm;
      wyp = r1 ? T(0) : tLy0 * hym;
    }
    if (own) {
      const int n0 = col + plane * k0;
      const bool r0 = full_mask(pmc, 0, opmc[0], k0, nz);
      const bool r1 = full_mask(pmc, 1, opmc[1], k0, nz);
      const bool r2 = full_mask(pmc, 2, opmc[2], k0, nz);
      const T *mt = sm.mt + mcur * MAT_STRIDE;
      const T hx = hx0o - (r0 ? T(0) : mt[0] * giAx[jc + ny * k0]) * curl_u(0, p0, p1, li, lj);
      const T hy = hy0o - (r1 ? T(0) : mt[1] * iAy0) * curl_u(1, p0, p1, li, lj);
      const T hz = hz0o - (r2 ? T(0) : mt[2] * iAz0) * curl_u(2, p0, p1, li, lj);
      Hnx[n0] = hx;
      Hny[n0] = hy;
      Hnz[n0] = hz;
      wx = r0 ? T(0) : tLx0 * hx;
      wy = r1 ? T(0) : tLy0 * hy;
      wz = r2 ? T(0) : gtLz[k0] * hz;
    }
    sm.w[s][0][lj][li] = wx;
    sm.w[s][1][lj][li] = wy;
    sm.w[s][2][lj][li] = wz;
    if (hrow || hcolm) halo_w(k0, p0, p1, s, hp0);
  }
  // packet of plane k0+1 for the H update at iteration k0 and the E update at
  // iteration k0+1
  T pHx = 0, pHy = 0, pHz = 0, pEx = 0, pEy = 0, pEz = 0, piAy = 0, pitAy = 0;
  int pM = 0;
  HaloPk php;
  php.in = false;
  php.ha = php.hb = php.area_a = php.tlz = 0;
  php.m = 0;
  T pJ = 0;
  if (k0 + 1 < k1) {
    if (own) {
      const int n1 = col + plane * (k0 + 1);
      pHx = Hx[n1];
      pHy = Hy[n1];
      pHz = Hz[n1];
      pEx = Ex[n1];
      pEy = Ey[n1];
      pEz = Ez[n1];
      pM = mat ? mat[n1] : 0;
      piAy = giAy[ic + nx * (k0 + 1)];
      pitAy = gitAy[ic + nx * (k0 + 1)];
    }
    halo_load(k0 + 1, php);
  }
  T jcur = beam_col ? jline[k0] : T(0);
  if (beam_col && k0 + 1 < k1) pJ = jline[k0 + 1];

  for (int k = k0; k < k1; ++k) {
    // 1. stage plane k+2 (its loads were issued one iteration ago)
    if (k + 2 <= k1) stage(k + 2, pE);
    // 2. prefetch plane k+3 and the own/halo packet of plane k+2
    if (k + 3 <= k1) fetch(k + 3, pE);
    // own packet of plane k+2 for every lane (clamped column and plane: a
    // value past the chunk or the grid is never used)
    const int k2 = min(k + 2, nz - 1);
    const int n2 = col + plane * k2;
    const T qHx = Hx[n2];
    const T qHy = Hy[n2];
    const T qHz = Hz[n2];
    const T qEx = Ex[n2];
    const T qEy = Ey[n2];
    const T qEz = Ez[n2];
    const int qM = mat ? mat[n2] : 0;
    const T qiAy = giAy[ic + nx * k2];
    const T qitAy = gitAy[ic + nx * k2];
    HaloPk qhp;
    halo_load(k + 2, qhp);
    const T qJ = beam_col ? jline[k2] : T(0);
    const int n = col + plane * k;
    const T itAx_k = gitAx[jc + ny * k];                       // scalar
    const T iAx_k1 = k + 1 < nz ? giAx[jc + ny * (k + 1)] : T(0);  // scalar
    const T tLz_k1 = k + 1 < nz ? gtLz[k + 1] : T(0);            // scalar
    // 3.
    __syncthreads();
    // 4. E_new(k)
    if (own) {
      const T(*ww)[TY + 1][TX + 1] = sm.w[k & 1];
      const T *mt = sm.mt + mcur * MAT_STRIDE;
      probe(a, i, j, k, eoz);
      const T ax = full_mask(pec, 0, opec[0], k, nz) ? T(0) : itAx_k;
      const T ay = full_mask(pec, 1, opec[1], k, nz) ? T(0) : itAy_k;
      const T az = full_mask(pec, 2, opec[2], k, nz) ? T(0) : itAz0;
      const T cx = ((wz - ww[2][lj - 1][li]) - wy) + wyp;
      const T cy = ((wx - wxp) - wz) + ww[2][lj][li - 1];
      const T cz = ((wy - ww[1][lj][li - 1]) - wx) + ww[0][lj - 1][li];
      // Each J mode finishes the update in its own branch: the J loads of the
      // buffer mode never share registers with the other modes, so their
      // wait does not hold back the loads in flight for the next planes.
      if (a.jmode == J_BUFFER) {
        const T jx = Jx[n];
        const T jy = (Jx + N)[n];
        const T jz = beam_col ? jcur : (Jx + 2 * (int64_t)N)[n];
        Enx[n] = eox + mt[3] * (ax * cx - jx);
        Eny[n] = eoy + mt[4] * (ay * cy - jy);
        Enz[n] = eoz + mt[5] * (az * cz - jz);
      } else {
        const T js = a.jmode == J_SIGMA_E ? T(1) : T(0);
        const T jx = js * (mt[6] * eox);
        const T jy = js * (mt[7] * eoy);
        const T jz = beam_col ? jcur : js * (mt[8] * eoz);
        Enx[n] = eox + mt[3] * (ax * cx - jx);
        Eny[n] = eoy + mt[4] * (ay * cy - jy);
        Enz[n] = eoz + mt[5] * (az * cz - jz);
      }
    }
    wxp = wx;
    wyp = wy;
    // 5. H_new(k+1)
    if (k + 1 < k1) {
      const int kk = k + 1;
      const int s = kk & 1;
      const int p0 = kk & (NR - 1), p1 = (kk + 1) & (NR - 1);
      if (own) {
        const int n1 = n + plane;
        const bool r0 = full_mask(pmc, 0, opmc[0], kk, nz);
        const bool r1 = full_mask(pmc, 1, opmc[1], kk, nz);
        const bool r2 = full_mask(pmc, 2, opmc[2], kk, nz);
        const T *mt = sm.mt + pM * MAT_STRIDE;
        const T hx = pHx - (r0 ? T(0) : mt[0] * iAx_k1) * curl_u(0, p0, p1, li, lj);
        const T hy = pHy - (r1 ? T(0) : mt[1] * piAy) * curl_u(1, p0, p1, li, lj);
        const T hz = pHz - (r2 ? T(0) : mt[2] * iAz0) * curl_u(2, p0, p1, li, lj);
        Hnx[n1] = hx;
        Hny[n1] = hy;
        Hnz[n1] = hz;
        wx = r0 ? T(0) : tLx0 * hx;
        wy = r1 ? T(0) : tLy0 * hy;
        wz = r2 ? T(0) : tLz_k1 * hz;
      }
      sm.w[s][0][lj][li] = wx;
      sm.w[s][1][lj][li] = wy;
      sm.w[s][2][lj][li] = wz;
      if (hrow || hcolm) halo_w(kk, p0, p1, s, php);
    }
    // rotate the packets
    mcur = pM;
    eox = pEx;
    eoy = pEy;
    eoz = pEz;
    itAy_k = pitAy;
    jcur = pJ;
    pHx = qHx;
    pHy = qHy;
    pHz = qHz;
    pEx = qEx;
    pEy = qEy;
    pEz = qEz;
    piAy = qiAy;
    pitAy = qitAy;
    pM = qM;
    pJ = qJ;
    php = qhp;
  }
}

// Tile order of block b among nb blocks: the hardware dispatches block b to
// XCD b % 8, so consecutive tiles (which share halo planes in L2) are given
// to blocks of one XCD.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q8 = nb / 8;
  return (b < 8 * q8) ? (b % 8) * q8 + b / 8 : b;
}

template <typename T, int TX, int TY, int KC>
__global__ void __launch_bounds__(TX *TY) fused_sweep(FusedArgs<T> f) {
  __shared__ FusedSmem<T, TX, TY> sm;
  fused_tile<T, TX, TY, KC>(f, xcd_tile(blockIdx.x, f.nb), sm);
}

// ---------------- batched launch ----------------

// One member grid of a batch: its step arguments for both buffer parities
// (parity p: E, H read from f[p].a, written to f[p].En/Hn), its step index
// when the batch counter is 0, and its beam table (nz values per step of the
// current launch group, or null).
template <typename T>
struct BatchMember {
  FusedArgs<T> f[2];
  int64_t step0;
  const T *btab;
};

// One step of every member grid in one persistent launch: block b takes the
// batch's tiles b, b + gridDim.x, ...; tstart[m] is the first tile of member
// m (ascending, tstart[nmem] = ntiles).  Step = *dstep + g of the batch.
template <typename T, int TX, int TY, int KC>
__global__ void __launch_bounds__(TX *TY)
    fused_batch(const BatchMember<T> *__restrict__ mem, const int *__restrict__ tstart, int nmem,
                int ntiles, const int64_t *__restrict__ dstep, int g) {
  __shared__ FusedSmem<T, TX, TY> sm;
  const int64_t s = *dstep + g;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchMember<T> &m = mem[lo];
    FusedArgs<T> f = m.f[s & 1];
    f.a.step = m.step0 + s;
    if (m.btab) f.a.jline = m.btab + (int64_t)g * f.a.nz;
    fused_tile<T, TX, TY, KC>(f, xcd_tile(t - tstart[lo], f.nb), sm);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// Beam tables of a launch group: member m, step g (0 .. G-1), plane k.
template <typename T>
__global__ void batch_beam_table(const BatchMember<T> *__restrict__ mem, int nmem,
                                 const int64_t *__restrict__ dstep, int G, T *const *btabs) {
  const int m = blockIdx.y;
  if (m >= nmem || !btabs[m]) return;
  const StepArgs<T> &a0 = mem[m].f[0].a;
  const int64_t cnt = (int64_t)G * a0.nz;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(t / a0.nz), k = (int)(t % a0.nz);
    StepArgs<T> a = a0;
    a.step = mem[m].step0 + *dstep + g;
    btabs[m][t] = beam_gauss(a, k);
  }
}

__global__ void batch_advance(int64_t *dstep, int G) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dstep += G;
}

inline bool fused_supported(int kernel, int nx, int ny, int nz) {
  (void)nx;
  (void)ny;
  (void)nz;
  return kernel != 1;  // FIT_KERNEL_SPLIT selects the split kernels
}

// Tile configurations (TX, TY, KC) selectable per context.
enum {
  FUSED_CFG_256x2 = 0,
  FUSED_CFG_128x4 = 1,
  FUSED_CFG_64x8 = 2,
  FUSED_CFG_256x4 = 3,
  FUSED_CFG_64x4 = 4,
  FUSED_CFG_256x2_K64 = 5,
  FUSED_NCFG = 6
};

inline void fused_dims(int cfg, int &tx, int &ty, int &kc) {
  switch (cfg) {
    case FUSED_CFG_128x4: tx = 128; ty = 4; kc = 32; break;
    case FUSED_CFG_64x8: tx = 64; ty = 8; kc = 32; break;
    case FUSED_CFG_256x4: tx = 256; ty = 4; kc = 32; break;
    case FUSED_CFG_64x4: tx = 64; ty = 4; kc = 32; break;
    case FUSED_CFG_256x2_K64: tx = 256; ty = 2; kc = 64; break;
    default: tx = 256; ty = 2; kc = 32; break;
  }
}

// Cost model of a tile configuration for a grid: cells the tiles cover
// (partial tiles padded) times the configuration's time per covered cell,
// measured on the 256 x 256 x 768 FP32 pillbox (ms per step per 5.03e7
// cells: 256x4 0.645, 64x8 0.657, 128x4 0.687, 64x4 0.689, 256x2 K64 0.78,
// 256x2 1.30).  Used when no tuning was done.
inline double fused_cost(int cfg, int nx, int ny, int nz) {
  int tx, ty, kc;
  fused_dims(cfg, tx, ty, kc);
  static const double w[FUSED_NCFG] = {1.30, 0.687, 0.657, 0.645, 0.689, 0.78};
  const double cx = (double)((nx + tx - 1) / tx) * tx;
  const double cy = (double)((ny + ty - 1) / ty) * ty;
  const double cz = (double)((nz + kc - 1) / kc) * kc;
  return cx * cy * cz * w[cfg];
}

// Default when no tuning was done: full 256-wide rows when the grid has them
// (256 x 4 measured fastest in FP32 on the 256 x 256 x 768 pillbox).
inline int fused_default(int nx, bool fp64) {
  if (nx >= 192 && !fp64) return FUSED_CFG_256x4;
  if (nx >= 96) return FUSED_CFG_128x4;
  return FUSED_CFG_64x8;
}

template <typename T, int TX, int TY>
constexpr bool cfg_fits() {
  return sizeof(FusedSmem<T, TX, TY>) <= 160 * 1024;
}

// Whether a tile configuration fits the LDS of one workgroup for T.
template <typename T>
inline bool fused_cfg_ok(int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: return cfg_fits<T, 128, 4>();
    case FUSED_CFG_64x8: return cfg_fits<T, 64, 8>();
    case FUSED_CFG_256x4: return cfg_fits<T, 256, 4>();
    case FUSED_CFG_64x4: return cfg_fits<T, 64, 4>();
    default: return cfg_fits<T, 256, 2>();
  }
}

// The configuration of least modelled cost over n grids (one grid, or the
// members of a batch) among those that fit the LDS for T.
template <typename T>
inline int fused_pick(const int *nx, const int *ny, const int *nz, int n) {
  int best = FUSED_CFG_128x4;
  double bc = 1e300;
  for (int cfg = 0; cfg < FUSED_NCFG; ++cfg) {
    if (!fused_cfg_ok<T>(cfg)) continue;
    double c = 0;
    for (int i = 0; i < n; ++i) c += fused_cost(cfg, nx[i], ny[i], nz[i]);
    if (c < bc) {
      bc = c;
      best = cfg;
    }
  }
  return best;
}

template <typename T, int TX, int TY, int KC>
inline void launch_cfg(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s) {
  if constexpr (!cfg_fits<T, TX, TY>()) {
    (void)a; (void)En; (void)Hn; (void)s;
    return;
  } else {
  FusedArgs<T> f;
  f.a = a;
  f.En = En;
  f.Hn = Hn;
  f.ntx = (a.nx + TX - 1) / TX;
  f.nty = (a.ny + TY - 1) / TY;
  f.nkc = (a.nz + KC - 1) / KC;
  f.nb = f.ntx * f.nty * f.nkc;
  hipLaunchKernelGGL((fused_sweep<T, TX, TY, KC>), dim3(f.nb), dim3(TX * TY), 0, s, f);
  }
}

template <typename T>
inline void launch_fused(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s, int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: launch_cfg<T, 128, 4, 32>(a, En, Hn, s); break;
    case FUSED_CFG_64x8: launch_cfg<T, 64, 8, 32>(a, En, Hn, s); break;
    case FUSED_CFG_256x4: launch_cfg<T, 256, 4, 32>(a, En, Hn, s); break;
    case FUSED_CFG_64x4: launch_cfg<T, 64, 4, 32>(a, En, Hn, s); break;
    case FUSED_CFG_256x2_K64: launch_cfg<T, 256, 2, 64>(a, En, Hn, s); break;
    default: launch_cfg<T, 256, 2, 32>(a, En, Hn, s); break;
  }
}

}  // namespace fit
