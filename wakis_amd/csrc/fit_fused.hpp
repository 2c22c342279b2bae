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
  int kc = 0;                    // planes per z chunk
  int kh = 0, kt = 0, nmid = 0;  // z chunk layout (chunk_span)
  int nremap = 0;                // blocks under the XCD remap
};

// Planes [k0, k1) of z chunk c of a tile column (chunks of f.kc planes).
// With kh > 0 the first nmid chunks take kc planes each from plane kh, the
// last of them ending at nz - kt, chunk nmid is [0, kh) and chunk nmid + 1
// is [nz - kt, nz): every plane but the kh + kt at the z faces lies in a
// chunk the lean path takes, and the blocks of the two short chunks come
// last (outside the XCD remap, so every XCD gets the same number of full
// chunks).  With kh = 0 chunk c is [c kc, (c + 1) kc).
template <typename T>
__host__ __device__ inline void chunk_span(const FusedArgs<T> &f, int c, int &k0, int &k1) {
  const int nz = f.a.nz, kc = f.kc;
  if (f.kh == 0) {
    k0 = c * kc;
    k1 = k0 + kc < nz ? k0 + kc : nz;
  } else if (c < f.nmid) {
    k0 = f.kh + c * kc;
    k1 = k0 + kc < nz - f.kt ? k0 + kc : nz - f.kt;
  } else if (c == f.nmid) {
    k0 = 0;
    k1 = f.kh;
  } else {
    k0 = nz - f.kt;
    k1 = nz;
  }
}

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
  static constexpr int NRING = 3;
};

// Static LDS of a workgroup; the material table (dt/mu, dt/eps, sigma rows)
// follows in dynamic LDS, sized to the grid's table.
template <typename T, int TX, int TY>
struct FusedSmem {
  using G = Tile<TX, TY>;
  T u[G::NRING][3][G::RH][G::RW];   // u = L*pec*E_old, window planes (ring)
  T w[2][3][TY + 1][TX + 1];        // w = tL*pmc*H_new, tile + low halo
};

// Dynamic LDS bytes of the material table of `nmat` rows.
template <typename T>
inline size_t fused_mt_bytes(int nmat) {
  return (size_t)(nmat > 1 ? nmat : 1) * MAT_STRIDE * sizeof(T);
}

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
// JB: the J buffer mode (J read from memory) or not (J = 0 or sigma*E); the
// two are separate code, so the J loads never meet the other mode's values
// in a register (a join there would make every wave wait for all its loads).
template <typename T, int TX, int TY, int KC, bool JB, bool NTS, bool MON = true>
__device__ __forceinline__ void fused_tile(const FusedArgs<T> &f, const int tile,
                                           FusedSmem<T, TX, TY> &sm, T *__restrict__ smt) {
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
  const int x0 = tx * TX, y0 = ty * TY;
  int k0, k1;  // planes k0 .. k1-1 of the tile's z chunk
  if constexpr (sizeof(T) == 8) {
    // FP64 grids keep uniform chunks of KC planes (no lean path to serve);
    // the compile-time length keeps this path as fast as it was
    k0 = (tr / f.nty) * KC;
    k1 = min(k0 + KC, nz);
  } else {
    chunk_span(f, tr / f.nty, k0, k1);
  }

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

  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += NT) smt[t] = mtab[t];

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
  // ring slot of window plane k (k >= -1; k is wave-uniform)
  auto slot = [](int k) { return (k + 3) % NR; };
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
    T(*dst)[RH][RW] = sm.u[slot(k)];
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
  const bool beam_on = a.beam_mode != BEAM_OFF;  // wave-uniform
  const bool beam_col = beam_on && i == a.ixs && j == a.iys;

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
      const T *mt = smt + p.m * MAT_STRIDE;
      const T ha = p.ha - (ra ? T(0) : mt[hc0] * p.area_a) * curl_u(hc0, p0, p1, hli, hlj);
      const T hb = p.hb - (rb ? T(0) : mt[2] * h_iAz) * curl_u(2, p0, p1, hli, hlj);
      wa = ra ? T(0) : h_tla * ha;
      wb = rb ? T(0) : p.tlz * hb;
    }
    sm.w[s][hc0][hlj][hli] = wa;
    sm.w[s][2][hlj][hli] = wb;
  };

  // ---- prologue: planes k0-1, k0, k0+1 staged; w of planes k0-1 and k0 ----
  // Packet consumed at iteration k, loaded one iteration earlier: the E
  // update data of plane k and the H update data (own and halo) of plane
  // k+1.  Two packets alternate (the loop body is written out twice), so no
  // register of a load still in flight is ever copied (a copy would make the
  // wave wait for the load there).
  struct Pk {
    T ex, ey, ez, itay;     // plane k: E_old, itA_y
    T hx, hy, hz, iay;      // plane k+1: H_old, iA_y
    int m;                  // material of plane k+1
    HaloPk h;               // halo packet of plane k+1
  };
  // every lane loads (clamped column and plane: a value past the chunk or the
  // grid is never used)
  auto load_pk = [&](int k, Pk &p) {
    const int ke = min(k, nz - 1);
    const int ne = col + plane * ke;
    p.ex = Ex[ne];
    p.ey = Ey[ne];
    p.ez = Ez[ne];
    p.itay = gitAy[ic + nx * ke];
    const int kh = min(k + 1, nz - 1);
    const int nh = col + plane * kh;
    p.hx = Hx[nh];
    p.hy = Hy[nh];
    p.hz = Hz[nh];
    p.iay = giAy[ic + nx * kh];
    p.m = mat ? mat[nh] : 0;
    halo_load(k + 1, p.h);
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
  int mcur = 0;  // material of the plane of the E update
  T hx0o = 0, hy0o = 0, hz0o = 0, iAy0 = 0;
  T hxmo = 0, hymo = 0, iAym = 0;
  int mm = 0;
  if (own) {
    const int n0 = col + plane * k0;
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
  Pk A, B;
  load_pk(k0, A);
  __syncthreads();
  T wxp = 0, wyp = 0;        // own w of plane k-1 (x, y)
  T wx = 0, wy = 0, wz = 0;  // own w of plane k
  {
    const int s = k0 & 1;
    const int pm = slot(k0 - 1), p0 = slot(k0), p1 = slot(k0 + 1);
    if (own && k0 >= 1) {
      const int km = k0 - 1;
      const bool r0 = full_mask(pmc, 0, opmc[0], km, nz);
      const bool r1 = full_mask(pmc, 1, opmc[1], km, nz);
      const T *mt = smt + mm * MAT_STRIDE;
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
      const T *mt = smt + mcur * MAT_STRIDE;
      const T hx = hx0o - (r0 ? T(0) : mt[0] * giAx[jc + ny * k0]) * curl_u(0, p0, p1, li, lj);
      const T hy = hy0o - (r1 ? T(0) : mt[1] * iAy0) * curl_u(1, p0, p1, li, lj);
      const T hz = hz0o - (r2 ? T(0) : mt[2] * iAz0) * curl_u(2, p0, p1, li, lj);
      gstore(&Hnx[n0], hx, NTS);
      gstore(&Hny[n0], hy, NTS);
      gstore(&Hnz[n0], hz, NTS);
      wx = r0 ? T(0) : tLx0 * hx;
      wy = r1 ? T(0) : tLy0 * hy;
      wz = r2 ? T(0) : gtLz[k0] * hz;
    }
    sm.w[s][0][lj][li] = wx;
    sm.w[s][1][lj][li] = wy;
    sm.w[s][2][lj][li] = wz;
    if (hrow || hcolm) halo_w(k0, p0, p1, s, hp0);
  }
  // the first iteration stages plane k0+2 into the slot of plane k0-1
  __syncthreads();

  // One plane: E_new(k) and H_new(k+1) from `cur`; loads `nxt` for plane k+1.
  auto iter = [&](const int k, const Pk &cur, Pk &nxt) {
    // 1. stage plane k+2 (its loads were issued one iteration ago)
    if (k + 2 <= k1) stage(k + 2, pE);
    // 2. prefetch window plane k+3 and the packet of the next iteration
    if (k + 3 <= k1) fetch(k + 3, pE);
    load_pk(k + 1, nxt);
    const int n = col + plane * k;
    const T itAx_k = gitAx[jc + ny * k];                       // scalar
    const T iAx_k1 = k + 1 < nz ? giAx[jc + ny * (k + 1)] : T(0);  // scalar
    const T tLz_k1 = k + 1 < nz ? gtLz[k + 1] : T(0);            // scalar
    const T jk = beam_on ? jline[k] : T(0);                     // scalar
    // 3.
    __syncthreads();
    // 4. E_new(k)
    if (own) {
      const T(*ww)[TY + 1][TX + 1] = sm.w[k & 1];
      const T *mt = smt + mcur * MAT_STRIDE;
      const T eox = cur.ex, eoy = cur.ey, eoz = cur.ez;
      probe(a, i, j, k, eoz);
      const T ax = full_mask(pec, 0, opec[0], k, nz) ? T(0) : itAx_k;
      const T ay = full_mask(pec, 1, opec[1], k, nz) ? T(0) : cur.itay;
      const T az = full_mask(pec, 2, opec[2], k, nz) ? T(0) : itAz0;
      const T cx = ((wz - ww[2][lj - 1][li]) - wy) + wyp;
      const T cy = ((wx - wxp) - wz) + ww[2][lj][li - 1];
      const T cz = ((wy - ww[1][lj][li - 1]) - wx) + ww[0][lj - 1][li];
      T enx, eny, enz;
      if (JB) {
        const T jx = Jx[n];
        const T jy = (Jx + N)[n];
        const T jz = beam_col ? jk : (Jx + 2 * (int64_t)N)[n];
        enx = eox + mt[3] * (ax * cx - jx);
        eny = eoy + mt[4] * (ay * cy - jy);
        enz = eoz + mt[5] * (az * cz - jz);
      } else {
        const T js = a.jmode == J_SIGMA_E ? T(1) : T(0);
        const T jx = js * (mt[6] * eox);
        const T jy = js * (mt[7] * eoy);
        const T jz = beam_col ? jk : js * (mt[8] * eoz);
        enx = eox + mt[3] * (ax * cx - jx);
        eny = eoy + mt[4] * (ay * cy - jy);
        enz = eoz + mt[5] * (az * cz - jz);
      }
      gstore(&Enx[n], enx, NTS);
      gstore(&Eny[n], eny, NTS);
      gstore(&Enz[n], enz, NTS);
      // field monitor fused into the step: accumulate the new E here instead
      // of reading it back in a separate pass
      if (MON && a.mon) monitor_cell(a, n, enx, eny, enz);
    }
    wxp = wx;
    wyp = wy;
    // 5. H_new(k+1)
    if (k + 1 < k1) {
      const int kk = k + 1;
      const int s = kk & 1;
      const int p0 = slot(kk), p1 = slot(kk + 1);
      if (own) {
        const int n1 = n + plane;
        const bool r0 = full_mask(pmc, 0, opmc[0], kk, nz);
        const bool r1 = full_mask(pmc, 1, opmc[1], kk, nz);
        const bool r2 = full_mask(pmc, 2, opmc[2], kk, nz);
        const T *mt = smt + cur.m * MAT_STRIDE;
        const T hx = cur.hx - (r0 ? T(0) : mt[0] * iAx_k1) * curl_u(0, p0, p1, li, lj);
        const T hy = cur.hy - (r1 ? T(0) : mt[1] * cur.iay) * curl_u(1, p0, p1, li, lj);
        const T hz = cur.hz - (r2 ? T(0) : mt[2] * iAz0) * curl_u(2, p0, p1, li, lj);
        gstore(&Hnx[n1], hx, NTS);
        gstore(&Hny[n1], hy, NTS);
        gstore(&Hnz[n1], hz, NTS);
        wx = r0 ? T(0) : tLx0 * hx;
        wy = r1 ? T(0) : tLy0 * hy;
        wz = r2 ? T(0) : tLz_k1 * hz;
      }
      sm.w[s][0][lj][li] = wx;
      sm.w[s][1][lj][li] = wy;
      sm.w[s][2][lj][li] = wz;
      if (hrow || hcolm) halo_w(kk, p0, p1, s, cur.h);
    }
    mcur = cur.m;
  };
  for (int k = k0; k < k1; k += 2) {
    iter(k, A, B);
    if (k + 1 < k1) iter(k + 1, B, A);
  }
}

// ---------------- lean path for chunks away from the z ends ----------------
//
// The same tile step as fused_tile for a chunk whose planes (window and halo
// included) all lie in 1 .. nz-2: no z-face masks and no clamps apply, the
// x/y masks are folded into per-lane factors fixed over the march (a masked
// term is a product with 0 instead of a selected 0: the same value for
// finite fields), and every global access is a buffer instruction: a
// per-lane byte offset fixed over the march plus a wave-uniform plane
// offset, so no address costs vector arithmetic.  Needs the metric tables
// and the material map each to span less than 2^32 bytes, and either the
// fields too (PCR false: one resource per field, component c at byte offset
// c N sizeof(T)) or the planes one chunk touches (PCR true: one resource per
// field component, based at the chunk's lowest plane k0-2).
template <typename T, int TX, int TY, int KC, bool JB, bool NTS, bool PCR>
__device__ __forceinline__ void fused_tile_inner(const FusedArgs<T> &f, const int tile,
                                                 FusedSmem<T, TX, TY> &sm, T *__restrict__ smt) {
  using G = Tile<TX, TY>;
  static_assert(TX % 64 == 0, "a wave must cover one tile row");
  constexpr int NT = G::NT, RW = G::RW, RH = G::RH, RC = G::RC, RL = G::RL, NR = G::NRING;
  constexpr uint32_t SZ = sizeof(T);
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int plane = nx * ny;
  const int tx = tile % f.ntx;
  const int tr = tile / f.ntx;
  const int ty = tr % f.nty;
  const int x0 = tx * TX, y0 = ty * TY;
  int k0, k1;
  chunk_span(f, tr / f.nty, k0, k1);
  const uint32_t PB = (uint32_t)plane * SZ;    // bytes per plane
  const uint32_t NB = (uint32_t)a.N * SZ;      // bytes per component
  const int64_t FB = (int64_t)3 * a.N * SZ;    // bytes per field
  const Geo<T> &g = a.g;
  // field component c: resource rX<c>, byte offset co<c> + cell offset +
  // (k - kb) PB for plane k
  const int kb = PCR ? k0 - 2 : 0;
  const int64_t eb = (int64_t)kb * plane;  // elements below the base plane
  const int64_t cspan = PCR ? ((int64_t)a.N - eb) * SZ : FB;
  const int64_t rspan = cspan < 0xFFFFFFFFll ? cspan : 0xFFFFFFFFll;
  auto comp_rsrc = [&](const T *p, int c) __attribute__((always_inline)) -> Rsrc {
    return PCR ? make_rsrc(p + (int64_t)c * a.N + eb, rspan) : make_rsrc(p, FB);
  };
  const uint32_t co1 = PCR ? 0u : NB, co2 = PCR ? 0u : 2 * NB;
  const Rsrc rE0 = comp_rsrc(a.E, 0), rE1 = comp_rsrc(a.E, 1), rE2 = comp_rsrc(a.E, 2);
  const Rsrc rH0 = comp_rsrc(a.H, 0), rH1 = comp_rsrc(a.H, 1), rH2 = comp_rsrc(a.H, 2);
  const Rsrc rEn0 = comp_rsrc(f.En, 0), rEn1 = comp_rsrc(f.En, 1), rEn2 = comp_rsrc(f.En, 2);
  const Rsrc rHn0 = comp_rsrc(f.Hn, 0), rHn1 = comp_rsrc(f.Hn, 1), rHn2 = comp_rsrc(f.Hn, 2);
  auto pofs = [&](int k) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)(k - kb) * PB;
  };
  const Rsrc rG = make_rsrc(g.Lx, a.geo_bytes);
  const bool hasmat = a.mat != nullptr;
  const Rsrc rM = make_rsrc(hasmat ? (const void *)a.mat : (const void *)g.Lx, hasmat ? a.N : 1);
  auto gofs = [&](const T *t) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)((const char *)t - (const char *)g.Lx);
  };
  const uint32_t oLz = gofs(g.Lz), otLz = gofs(g.tLz), oiAy = gofs(g.iAy), oitAy = gofs(g.itAy);
  const uint32_t rowb = (uint32_t)nx * SZ;  // row stride of iA_y, itA_y
  const CT<T> *gtLz = (const CT<T> *)g.tLz;
  const CT<T> *giAx = (const CT<T> *)g.iAx;
  const CT<T> *gitAx = (const CT<T> *)g.itAx;
  const CT<T> *jline = (const CT<T> *)a.jline;
  const CT<T> *mtab = (const CT<T> *)a.mtab;
  const uint32_t pec = a.pec, pmc = a.pmc;

  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += NT) smt[t] = mtab[t];

  // ---- window slots: LDS index, byte offsets (per component) from plane
  //      kk-1, mask factors ----
  uint32_t wld[RL], wob[RL][3], wlz[RL];
  T fx[RL], fy[RL], fz[RL];
  const int ri_max = min(TX + 1, nx - x0 + 1), rj_max = min(TY + 1, ny - y0 + 1);
#pragma unroll
  for (int q = 0; q < RL; ++q) {
    const int c = threadIdx.x + q * NT;
    const int rj = c / RW, ri = c - rj * RW;
    const bool need = c < RC && ri <= ri_max && rj <= rj_max;
    const WinCell w = need ? win_cell(x0 - 1 + ri, y0 - 1 + rj, nx, ny) : WinCell{x0, y0, 0, false};
    wld[q] = c < RC ? (uint32_t)c : 0u;
    wob[q][0] = (uint32_t)(w.i + nx * w.j + plane * (w.dk + 1)) * SZ;
    wob[q][1] = wob[q][0] + co1;
    wob[q][2] = wob[q][0] + co2;
    wlz[q] = oLz + (uint32_t)(w.dk + 1) * SZ;
    fx[q] = (need && !xy_mask(pec, 0, w.i, w.j, nx, ny)) ? g.Lx[w.i] : T(0);
    fy[q] = (need && !xy_mask(pec, 1, w.i, w.j, nx, ny)) ? g.Ly[w.j] : T(0);
    fz[q] = (need && !xy_mask(pec, 2, w.i, w.j, nx, ny)) ? T(1) : T(0);
  }

  // ---- own cell ----
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int i = x0 + lx, j = y0 + ly;
  const bool own = i < nx && j < ny;
  const int ic = min(i, nx - 1);
  const int jc = __builtin_amdgcn_readfirstlane(min(j, ny - 1));  // one row per wave
  const int li = lx + 1, lj = ly + 1;
  const int col = ic + nx * jc;
  // own cell byte offsets: field components; itA_y of row k (from row k);
  // iA_y of row k+1 (from row k)
  const uint32_t cb0 = (uint32_t)col * SZ, cb1 = cb0 + co1, cb2 = cb0 + co2;
  const uint32_t itb = oitAy + (uint32_t)ic * SZ, iab = oiAy + (uint32_t)ic * SZ + rowb;
  const uint32_t colm = (uint32_t)(col + plane);  // material of plane k+1 from plane k
  const T fe0 = xy_mask(pec, 0, ic, jc, nx, ny) ? T(0) : T(1);
  const T fe1 = xy_mask(pec, 1, ic, jc, nx, ny) ? T(0) : T(1);
  const T az0 = xy_mask(pec, 2, ic, jc, nx, ny) ? T(0) : g.itAz[col];
  const bool m0 = xy_mask(pmc, 0, ic, jc, nx, ny), m1 = xy_mask(pmc, 1, ic, jc, nx, ny);
  const bool m2 = xy_mask(pmc, 2, ic, jc, nx, ny);
  const T fm0 = m0 ? T(0) : T(1), fm1 = m1 ? T(0) : T(1), fm2 = m2 ? T(0) : T(1);
  const T tLx0m = m0 ? T(0) : g.tLx[ic], tLy0m = m1 ? T(0) : g.tLy[jc];
  const T iAz0m = m2 ? T(0) : g.iAz[col];
  const bool beam_on = a.beam_mode != BEAM_OFF;  // wave-uniform
  const bool beam_col = beam_on && i == a.ixs && j == a.iys;
  // wake probe: E_z history column `step` of this lane's window, or none
  int pw = -1;
  GT<T> *phist = nullptr;
  if (a.hist && a.step < a.hist_nt) {
    const int di = i - a.ixt, dj = j - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw) {
      pw = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
      phist = (GT<T> *)a.hist + (int64_t)pw * nz * a.hist_nt + a.step;
    }
  }

  // ---- low-halo cell: row (x0+t, y0-1) -> H_x, H_z; column (x0-1, y0+t-TX)
  //      -> H_y, H_z (rows fill waves 0 .. TX/64-1, the column wave TX/64) ----
  const bool hrow = threadIdx.x < TX;  // wave-uniform
  const bool hcolm = !hrow && threadIdx.x < TX + TY;
  const bool hwave = (int)(threadIdx.x >> 6) <= TX / 64;  // wave-uniform
  int hli = 0, hlj = 0;
  WinCell hc{0, 0, 0, false};
  if (hrow) {
    hli = threadIdx.x + 1;
    hc = win_cell(x0 + threadIdx.x, y0 - 1, nx, ny);
    hc.ok = x0 + (int)threadIdx.x < nx;
  } else if (hcolm) {
    hlj = threadIdx.x - TX + 1;
    hc = win_cell(x0 - 1, y0 + ((int)threadIdx.x - TX), nx, ny);
    hc.ok = y0 + ((int)threadIdx.x - TX) < ny;
  }
  const int hc0 = hrow ? 0 : 1;  // first halo component (second is z)
  const int hci = hc.i, hcj = hc.j, hdk = hc.dk;
  const bool hin = (hrow || hcolm) && hc.ok;
  const int hcolo = hci + nx * hcj;
  const bool hma = xy_mask(pmc, hc0, hci, hcj, nx, ny), hmb = xy_mask(pmc, 2, hci, hcj, nx, ny);
  // halo offsets from plane kk-1 (kk: the halo's plane)
  const uint32_t hom = (uint32_t)(hcolo + plane * (hdk + 1));
  const uint32_t hoa = hom * SZ + (hrow ? 0u : co1);  // H_x (rows) or H_y (column)
  const uint32_t hob = hom * SZ + co2;                 // H_z
  const Rsrc rHa = hrow ? rH0 : rH1;
  const uint32_t has = (uint32_t)__builtin_amdgcn_readfirstlane(hrow ? ny : nx) * SZ;
  const uint32_t oiAx = gofs(g.iAx);
  const uint32_t hab = (hrow ? oiAx : oiAy) +
                       (uint32_t)(hrow ? hcj + ny * (hdk + 1) : hci + nx * (hdk + 1)) * SZ;
  const uint32_t hlz = otLz + (uint32_t)(hdk + 2) * SZ;  // tL_z of plane kk+dk from row kk-2
  const T fha = (hin && !hma) ? T(1) : T(0);
  const T fhb = (hin && !hmb) ? T(1) : T(0);
  const T tlx_h = g.tLx[hci], tly_h = g.tLy[hcj];
  const T h_tlam = (hin && !hma) ? (hrow ? tlx_h : tly_h) : T(0);
  const T h_iAzm = (hin && !hmb) ? g.iAz[hcolo] : T(0);

  auto slot = [](int k) __attribute__((always_inline)) { return (k + 3) % NR; };
  auto fetch = [&](int kk, T(&v)[RL][4]) __attribute__((always_inline)) {
    const uint32_t s0 = pofs(kk - 1), sz0 = (uint32_t)(kk - 1) * SZ;
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      v[q][0] = bload<T>(rE0, wob[q][0], s0);
      v[q][1] = bload<T>(rE1, wob[q][1], s0);
      v[q][2] = bload<T>(rE2, wob[q][2], s0);
      v[q][3] = bload<T>(rG, wlz[q], sz0);
    }
  };
  auto stage = [&](int kk, const T(&v)[RL][4]) __attribute__((always_inline)) {
    T *dst = &sm.u[slot(kk)][0][0][0];
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      if (q < RL - 1 || threadIdx.x + q * NT < RC) {
        dst[wld[q]] = fx[q] * v[q][0];
        dst[RC + wld[q]] = fy[q] * v[q][1];
        dst[2 * RC + wld[q]] = fz[q] * (v[q][3] * v[q][2]);
      }
    }
  };
  // curl of H component d at window (wi, wj) from slots p0 (plane) / p1 (above)
  auto curl_u = [&](int d, int p0, int p1, int wi, int wj) __attribute__((always_inline)) -> T {
    const T(*u0)[RH][RW] = sm.u[p0];
    const T(*u1)[RH][RW] = sm.u[p1];
    if (d == 0) return ((u0[1][wj][wi] - u1[1][wj][wi]) - u0[2][wj][wi]) + u0[2][wj + 1][wi];
    if (d == 1) return ((u0[2][wj][wi] - u0[2][wj][wi + 1]) - u0[0][wj][wi]) + u1[0][wj][wi];
    return ((u0[0][wj][wi] - u0[0][wj + 1][wi]) - u0[1][wj][wi]) + u0[1][wj][wi + 1];
  };
  struct HaloPk {
    T ha, hb, area_a, tlz;
    int m;
  };
  // halo cell of plane kk (plane kk+dk: base kk-1, offset dk+1); the halo
  // waves load for every lane (lanes without a halo cell read a valid cell)
  auto halo_load = [&](int kk, HaloPk &p) __attribute__((always_inline)) {
    // (fields assigned once, after the branch: stores of different fields
    // in the two arms would be merged through a pointer select, which keeps
    // the packet in scratch memory)
    T ha = 0, hb = 0, tlz = 0, area = 0;
    int m = 0;
    if (hwave) {
      const uint32_t s0 = pofs(kk - 1);
      ha = bload<T>(rHa, hoa, s0);
      hb = bload<T>(rH2, hob, s0);
      m = hasmat ? bload_u8(rM, hom, (uint32_t)(kk - 1) * (uint32_t)plane) : 0;
      tlz = bload<T>(rG, hlz, (uint32_t)(kk - 2) * SZ);
      area = bload<T>(rG, hab, (uint32_t)(kk - 1) * has);
    }
    p.ha = ha;
    p.hb = hb;
    p.area_a = area;
    p.tlz = tlz;
    p.m = m;
  };
  auto halo_w = [&](int s, int p0, int p1, const HaloPk &p) __attribute__((always_inline)) {
    const T *mt = smt + p.m * MAT_STRIDE;
    const T ha = p.ha - (mt[hc0] * (fha * p.area_a)) * curl_u(hc0, p0, p1, hli, hlj);
    const T hb = p.hb - (mt[2] * h_iAzm) * curl_u(2, p0, p1, hli, hlj);
    sm.w[s][hc0][hlj][hli] = h_tlam * ha;
    sm.w[s][2][hlj][hli] = (fhb * p.tlz) * hb;
  };
  // Packet consumed at iteration k: E update data of plane k, H update data
  // (own and halo) of plane k+1.  It is loaded at the end of iteration k-2,
  // after that iteration's use of the same registers, so two packets (and
  // two window buffers) alternate with the loop body written out twice: the
  // loads have about two iterations to arrive, and no register of a load in
  // flight is ever copied.
  struct Pk {
    T ex, ey, ez, itay;  // plane k
    T hx, hy, hz, iay;   // plane k+1
    int m;               // material of plane k+1
    HaloPk h;            // halo of plane k+1
  };
  auto load_pk = [&](int k, Pk &p) __attribute__((always_inline)) {
    const uint32_t se = pofs(k), sh = se + PB, sr = (uint32_t)k * rowb;
    p.ex = bload<T>(rE0, cb0, se);
    p.ey = bload<T>(rE1, cb1, se);
    p.ez = bload<T>(rE2, cb2, se);
    p.itay = bload<T>(rG, itb, sr);
    p.hx = bload<T>(rH0, cb0, sh);
    p.hy = bload<T>(rH1, cb1, sh);
    p.hz = bload<T>(rH2, cb2, sh);
    p.iay = bload<T>(rG, iab, sr);
    p.m = hasmat ? bload_u8(rM, colm, (uint32_t)k * (uint32_t)plane) : 0;
    halo_load(k + 1, p.h);
  };

  // ---- prologue: planes k0-1, k0, k0+1 staged; w of planes k0-1 and k0 ----
  T pEa[RL][4], pEb[RL][4];
  fetch(k0 - 1, pEa);
  stage(k0 - 1, pEa);
  fetch(k0, pEa);
  stage(k0, pEa);
  fetch(k0 + 1, pEa);
  stage(k0 + 1, pEa);
  if (k0 + 2 <= k1) fetch(k0 + 2, pEa);
  if (k0 + 3 <= k1) fetch(k0 + 3, pEb);
  T hxm, hym, iAym, hx0, hy0, hz0, iAy0;
  int mm, mcur;
  {
    const uint32_t sm1 = pofs(k0 - 1), s0 = sm1 + PB;
    hxm = bload<T>(rH0, cb0, sm1);
    hym = bload<T>(rH1, cb1, sm1);
    iAym = bload<T>(rG, iab, (uint32_t)(k0 - 2) * rowb);
    mm = hasmat ? bload_u8(rM, colm, (uint32_t)(k0 - 2) * (uint32_t)plane) : 0;
    hx0 = bload<T>(rH0, cb0, s0);
    hy0 = bload<T>(rH1, cb1, s0);
    hz0 = bload<T>(rH2, cb2, s0);
    iAy0 = bload<T>(rG, iab, (uint32_t)(k0 - 1) * rowb);
    mcur = hasmat ? bload_u8(rM, colm, (uint32_t)(k0 - 1) * (uint32_t)plane) : 0;
  }
  HaloPk hp0;
  halo_load(k0, hp0);
  Pk A, B;
  load_pk(k0, A);
  if (k0 + 1 < k1) load_pk(k0 + 1, B);
  __syncthreads();
  T wxp, wyp, wx, wy, wz;
  {
    const int s = k0 & 1;
    const int pm = slot(k0 - 1), p0 = slot(k0), p1 = slot(k0 + 1);
    const T *mtm = smt + mm * MAT_STRIDE;
    const T hxn = hxm - (mtm[0] * (fm0 * giAx[jc + ny * (k0 - 1)])) * curl_u(0, pm, p0, li, lj);
    const T hyn = hym - (mtm[1] * (fm1 * iAym)) * curl_u(1, pm, p0, li, lj);
    wxp = tLx0m * hxn;
    wyp = tLy0m * hyn;
    const T *mt = smt + mcur * MAT_STRIDE;
    const T hx = hx0 - (mt[0] * (fm0 * giAx[jc + ny * k0])) * curl_u(0, p0, p1, li, lj);
    const T hy = hy0 - (mt[1] * (fm1 * iAy0)) * curl_u(1, p0, p1, li, lj);
    const T hz = hz0 - (mt[2] * iAz0m) * curl_u(2, p0, p1, li, lj);
    if (own) {
      const uint32_t s0 = pofs(k0);
      bstore(rHn0, cb0, s0, hx, NTS);
      bstore(rHn1, cb1, s0, hy, NTS);
      bstore(rHn2, cb2, s0, hz, NTS);
    }
    wx = own ? tLx0m * hx : T(0);
    wy = own ? tLy0m * hy : T(0);
    wz = own ? (fm2 * gtLz[k0]) * hz : T(0);
    if (!own) wxp = wyp = T(0);
    sm.w[s][0][lj][li] = wx;
    sm.w[s][1][lj][li] = wy;
    sm.w[s][2][lj][li] = wz;
    if (hrow || hcolm) halo_w(s, p0, p1, hp0);
  }
  // the first iteration stages plane k0+2 into the slot of plane k0-1
  __syncthreads();

  // One plane: E_new(k) and H_new(k+1) from `cur`, then the loads of
  // iteration k+2 into `cur` and pE.
  auto iter = [&](const int k, Pk &cur, T(&pE)[RL][4]) __attribute__((always_inline)) {
    // 1. stage plane k+2 (its loads were issued two iterations ago)
    if (k + 2 <= k1) stage(k + 2, pE);
    const T itAx_k = gitAx[jc + ny * k];       // scalar
    const T iAx_k1 = giAx[jc + ny * (k + 1)];  // scalar
    const T tLz_k1 = gtLz[k + 1];              // scalar
    const T jk = beam_on ? jline[k] : T(0);    // scalar
    // 3.
    __syncthreads();
    // 4. E_new(k)
    const uint32_t se = pofs(k);
    if (own) {
      const T(*ww)[TY + 1][TX + 1] = sm.w[k & 1];
      const T *mt = smt + mcur * MAT_STRIDE;
      const T eox = cur.ex, eoy = cur.ey, eoz = cur.ez;
      if (pw >= 0) phist[(int64_t)k * a.hist_nt] = eoz;
      const T ax = fe0 * itAx_k;
      const T ay = fe1 * cur.itay;
      const T cx = ((wz - ww[2][lj - 1][li]) - wy) + wyp;
      const T cy = ((wx - wxp) - wz) + ww[2][lj][li - 1];
      const T cz = ((wy - ww[1][lj][li - 1]) - wx) + ww[0][lj - 1][li];
      T jx, jy, jz;
      if (JB) {
        jx = bload<T>(comp_rsrc(a.J, 0), cb0, se);
        jy = bload<T>(comp_rsrc(a.J, 1), cb1, se);
        jz = beam_col ? jk : bload<T>(comp_rsrc(a.J, 2), cb2, se);
      } else {
        const T js = a.jmode == J_SIGMA_E ? T(1) : T(0);
        jx = js * (mt[6] * eox);
        jy = js * (mt[7] * eoy);
        jz = beam_col ? jk : js * (mt[8] * eoz);
      }
      bstore(rEn0, cb0, se, eox + mt[3] * (ax * cx - jx), NTS);
      bstore(rEn1, cb1, se, eoy + mt[4] * (ay * cy - jy), NTS);
      bstore(rEn2, cb2, se, eoz + mt[5] * (az0 * cz - jz), NTS);
    }
    wxp = wx;
    wyp = wy;
    // 5. H_new(k+1)
    if (k + 1 < k1) {
      const int kk = k + 1;
      const int s = kk & 1;
      const int p0 = slot(kk), p1 = slot(kk + 1);
      if (own) {
        const uint32_t sh = se + PB;
        const T *mt = smt + cur.m * MAT_STRIDE;
        const T hx = cur.hx - (mt[0] * (fm0 * iAx_k1)) * curl_u(0, p0, p1, li, lj);
        const T hy = cur.hy - (mt[1] * (fm1 * cur.iay)) * curl_u(1, p0, p1, li, lj);
        const T hz = cur.hz - (mt[2] * iAz0m) * curl_u(2, p0, p1, li, lj);
        bstore(rHn0, cb0, sh, hx, NTS);
        bstore(rHn1, cb1, sh, hy, NTS);
        bstore(rHn2, cb2, sh, hz, NTS);
        wx = tLx0m * hx;
        wy = tLy0m * hy;
        wz = (fm2 * tLz_k1) * hz;
      }
      sm.w[s][0][lj][li] = wx;
      sm.w[s][1][lj][li] = wy;
      sm.w[s][2][lj][li] = wz;
      if (hrow || hcolm) halo_w(s, p0, p1, cur.h);
    }
    mcur = cur.m;
    // 6. prefetch window plane k+4 and the packet of iteration k+2
    if (k + 4 <= k1) fetch(k + 4, pE);
    if (k + 2 < k1) load_pk(k + 2, cur);
  };
  for (int k = k0; k < k1; k += 2) {
    iter(k, A, pEa);
    if (k + 1 < k1) iter(k + 1, B, pEb);
  }
}

// Host: the chunk layout of the grid for chunks of kc planes (f.kc, f.kh,
// f.kt, f.nmid, f.nkc, f.nb and f.nremap, the blocks under the XCD
// remap).  With `faces`, a 4-plane chunk at each z face when the grid has
// room for at least one full chunk between them (the batched launch, whose
// members are mostly small, keeps uniform chunks: their face chunks' extra
// blocks cost more there than the planes they move to the lean path gain).
template <typename T>
inline void chunk_layout(FusedArgs<T> &f, int nz, int kc, bool faces = true) {
  constexpr int kh = 4, kt = 4;
  const int ncol = f.ntx * f.nty;
  f.kc = kc;
  if (faces && nz - kh - kt >= kc) {
    f.kh = kh;
    f.kt = kt;
    f.nmid = (nz - kh - kt + kc - 1) / kc;
    f.nkc = f.nmid + 2;
    f.nremap = ncol * f.nmid;
  } else {
    f.kh = 0;
    f.kt = 0;
    f.nmid = 0;
    f.nkc = (nz + kc - 1) / kc;
    f.nremap = ncol * f.nkc;
  }
  f.nb = ncol * f.nkc;
}

// Planes per chunk of a single-grid sweep over ncol tile columns and nz
// planes, with `slots` blocks resident on the device at once: the blocks run
// in rounds (the XCD remap gives each of the 8 XCDs an eighth of them and of
// the slots), a round lasting as long as a chunk; modelled time rounds x
// (kc + 4 planes of prologue) over kc = KC, 1.5, 2, 3, 4 and 6 KC (the first
// least).  On the 256 x 256 x 768 pillbox with 256 x 4 tiles: 192 planes,
// one round of 256 blocks (0.546 ms per step against 0.577 for 64).  When
// the tile's own KC already needs more than 8 rounds the rounding matters
// little and KC stays (the 1024^3 grid: 13.74 ms per step with 128 planes
// against 13.43 with 32).
inline int pick_kc(int KC, int ncol, int nz, int slots) {
  const int mult[6] = {2, 3, 4, 6, 8, 12};  // x KC / 2
  int best = KC;
  double bt = 1e300;
  const int xslots = slots / 8 > 0 ? slots / 8 : 1;
  {
    const int nmid = (nz + KC - 1) / KC;
    const long long per_xcd = ((long long)ncol * nmid + 7) / 8;
    if ((per_xcd + xslots - 1) / xslots > 8) return KC;
  }
  for (int q = 0; q < 6; ++q) {
    const int kc = KC * mult[q] / 2;
    if (q > 0 && kc > nz - 8) break;
    const int nmid = nz - 8 >= kc ? (nz - 8 + kc - 1) / kc : (nz + kc - 1) / kc;
    const long long blocks = (long long)ncol * nmid;
    const long long per_xcd = (blocks + 7) / 8;
    const long long rounds = (per_xcd + xslots - 1) / xslots;
    const double t = (double)rounds * (kc + 4);
    if (t < bt) {
      bt = t;
      best = kc;
    }
  }
  return best;
}

// One tile with the code path for its J mode and z position: the lean path
// when every plane the chunk touches (window and halo included) lies in
// 1 .. nz-2 and its buffer offsets fit 32 bits: a field and the metric
// tables span less than 2^32 bytes, or (PCR, set by the host only when it
// holds) the planes of every chunk do, per field component.  (In FP64 the
// lean tile holds 209 VGPRs against the general tile's 189: two waves per
// SIMD either way.)  NTS: non-temporal field stores.
// MON: the general path may carry a fused field-monitor update (a.mon set:
// every chunk then takes the general path); false for the batched launch,
// whose members have no monitor.
template <typename T, int TX, int TY, int KC, bool NTS, bool PCR = false, bool MON = true>
__device__ __forceinline__ void fused_dispatch(const FusedArgs<T> &f, int tile,
                                               FusedSmem<T, TX, TY> &sm, T *smt) {
  {
    int k0, k1;
    chunk_span(f, tile / (f.ntx * f.nty), k0, k1);
    if ((PCR || f.a.buf32) && k0 >= 3 && k1 + 4 <= f.a.nz && !(MON && f.a.mon)) {
      if (f.a.jmode == J_BUFFER) fused_tile_inner<T, TX, TY, KC, true, NTS, PCR>(f, tile, sm, smt);
      else fused_tile_inner<T, TX, TY, KC, false, NTS, PCR>(f, tile, sm, smt);
      return;
    }
  }
  if (f.a.jmode == J_BUFFER) fused_tile<T, TX, TY, KC, true, NTS, MON>(f, tile, sm, smt);
  else fused_tile<T, TX, TY, KC, false, NTS, MON>(f, tile, sm, smt);
}

// Tile order of block b among nb blocks: the hardware dispatches block b to
// XCD b % 8, so consecutive tiles (which share halo planes in L2) are given
// to blocks of one XCD.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q8 = nb / 8;
  return (b < 8 * q8) ? (b % 8) * q8 + b / 8 : b;
}

extern __shared__ double fit_dyn_lds[];

template <typename T, int TX, int TY, int KC, bool NTS, bool PCR = false>
__global__ void __launch_bounds__(TX *TY)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) fused_sweep(FusedArgs<T> f) {
  __shared__ FusedSmem<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  const int b = blockIdx.x;
  fused_dispatch<T, TX, TY, KC, NTS, PCR>(f, b < f.nremap ? xcd_tile(b, f.nremap) : b, sm, smt);
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

// One step of every member grid in one launch of one block per tile: block
// t takes tile t of the batch, of the member m with tstart[m] <= t <
// tstart[m+1] (tstart ascending, tstart[nmem] = ntiles).  Step = *dstep + g
// of the batch.  (A persistent grid walking the tiles keeps the walk's state
// live across the tile body and made the compiler spill it to scratch.)
template <typename T, int TX, int TY, int KC>
__global__ void __launch_bounds__(TX *TY)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) fused_batch(const BatchMember<T> *__restrict__ mem, const int *__restrict__ tstart, int nmem,
                int ntiles, const int64_t *__restrict__ dstep, int g) {
  __shared__ FusedSmem<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  const int64_t s = *dstep + g;
  const int t = blockIdx.x;
  if (t < ntiles) {
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
    const int b = t - tstart[lo];
    fused_dispatch<T, TX, TY, KC, false, false, false>(f, b < f.nremap ? xcd_tile(b, f.nremap) : b,
                                                       sm, smt);
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
  FUSED_CFG_320x3_K64 = 0,
  FUSED_CFG_128x4 = 1,
  FUSED_CFG_64x8 = 2,
  FUSED_CFG_256x4 = 3,
  FUSED_CFG_64x4 = 4,
  FUSED_CFG_256x4_K64 = 5,
  FUSED_NCFG = 6
};

inline void fused_dims(int cfg, int &tx, int &ty, int &kc) {
  switch (cfg) {
    case FUSED_CFG_128x4: tx = 128; ty = 4; kc = 32; break;
    case FUSED_CFG_64x8: tx = 64; ty = 8; kc = 32; break;
    case FUSED_CFG_256x4: tx = 256; ty = 4; kc = 32; break;
    case FUSED_CFG_64x4: tx = 64; ty = 4; kc = 32; break;
    case FUSED_CFG_256x4_K64: tx = 256; ty = 4; kc = 64; break;
    default: tx = 320; ty = 3; kc = 64; break;
  }
}

// Cost model of a tile configuration for a grid: cells the tiles cover
// (partial tiles padded) times the configuration's time per covered cell,
// measured on the 256 x 256 x 768 FP32 pillbox (ms per step per 5.03e7
// cells: 256x4 K64 0.603, 256x4 0.611, 64x8 0.637, 128x4 0.640, 64x4 0.752;
// 320x3 K64 on the same pillbox 320 cells wide: 0.752).  Used when no
// tuning was done.
inline double fused_cost(int cfg, int nx, int ny, int nz) {
  int tx, ty, kc;
  fused_dims(cfg, tx, ty, kc);
  static const double w[FUSED_NCFG] = {0.752, 0.640, 0.637, 0.611, 0.752, 0.603};
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

// (with a full 256-row material table)
template <typename T, int TX, int TY>
constexpr bool cfg_fits() {
  return sizeof(FusedSmem<T, TX, TY>) + MAX_MATERIALS * MAT_STRIDE * sizeof(T) <= 160 * 1024;
}

// Whether a tile configuration fits the LDS of one workgroup for T.
template <typename T>
inline bool fused_cfg_ok(int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: return cfg_fits<T, 128, 4>();
    case FUSED_CFG_64x8: return cfg_fits<T, 64, 8>();
    case FUSED_CFG_256x4: return cfg_fits<T, 256, 4>();
    case FUSED_CFG_64x4: return cfg_fits<T, 64, 4>();
    case FUSED_CFG_256x4_K64: return cfg_fits<T, 256, 4>();
    default: return cfg_fits<T, 320, 3>();
  }
}

// The configuration of least modelled cost over n grids (one grid, or the
// members of a batch) among those that fit the LDS for T (and, when given,
// that allowed[cfg] admits).
template <typename T>
inline int fused_pick(const int *nx, const int *ny, const int *nz, int n,
                      const bool *allowed = nullptr) {
  int best = FUSED_CFG_128x4;
  double bc = 1e300;
  for (int cfg = 0; cfg < FUSED_NCFG; ++cfg) {
    if (!fused_cfg_ok<T>(cfg) || (allowed && !allowed[cfg])) continue;
    double c = 0;
    for (int i = 0; i < n; ++i) c += fused_cost(cfg, nx[i], ny[i], nz[i]);
    if (c < bc) {
      bc = c;
      best = cfg;
    }
  }
  return best;
}

// Resident blocks of kernel fn on device dev with `lds` bytes of dynamic
// LDS per block (CUs x blocks per CU), cached per (device, kernel, LDS bytes)
// behind a lock.
inline int resident_slots(int dev, const void *fn, int nthreads, size_t lds) {
  struct Key {
    int dev;
    const void *fn;
    size_t lds;
    bool operator==(const Key &o) const { return dev == o.dev && fn == o.fn && lds == o.lds; }
  };
  struct KeyHash {
    size_t operator()(const Key &k) const {
      return std::hash<const void *>()(k.fn) ^ (k.lds * 1000003u) ^ ((size_t)k.dev << 48);
    }
  };
  static std::mutex mu;
  static std::unordered_map<Key, int, KeyHash> cache;
  std::lock_guard<std::mutex> lk(mu);
  const Key key{dev, fn, lds};
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int ncu = 0, occ = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, nthreads, lds);
  const int slots = (ncu > 0 ? ncu : 256) * (occ > 0 ? occ : 1);
  cache[key] = slots;
  return slots;
}

// Every single-grid instantiation of a tile configuration (stores ordinary
// and non-temporal; FP32 also the per-component resources) is spill-free.
template <typename T, int TX, int TY, int KC>
inline bool fused_cfg_spill_free() {
  if constexpr (!cfg_fits<T, TX, TY>()) {
    return false;
  } else {
    bool ok = kernel_spill_free((const void *)fused_sweep<T, TX, TY, KC, false>) &&
              kernel_spill_free((const void *)fused_sweep<T, TX, TY, KC, true>);
    if constexpr (sizeof(T) == 4)
      ok = ok && kernel_spill_free((const void *)fused_sweep<T, TX, TY, KC, false, true>) &&
           kernel_spill_free((const void *)fused_sweep<T, TX, TY, KC, true, true>);
    return ok;
  }
}

// Launches one step of the fused sweep; false (nothing launched) when the
// configuration does not fit the LDS or a variant of it spills.
template <typename T, int TX, int TY, int KC>
inline bool launch_cfg(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s) {
  if constexpr (!cfg_fits<T, TX, TY>()) {
    (void)a; (void)En; (void)Hn; (void)s;
    return false;
  } else {
  if (!fused_cfg_spill_free<T, TX, TY, KC>()) return false;
  FusedArgs<T> f;
  f.a = a;
  f.En = En;
  f.Hn = Hn;
  f.ntx = (a.nx + TX - 1) / TX;
  f.nty = (a.ny + TY - 1) / TY;
  bool pcr = false;
  {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int slots = resident_slots(dev, (const void *)fused_sweep<T, TX, TY, KC, false>,
                                     TX * TY, fused_mt_bytes<T>(a.nmat));
    // the face chunks and other chunk lengths serve the FP32 lean path;
    // FP64 keeps uniform chunks of KC planes (its general tile decodes the
    // chunk at compile time)
    const int kc = sizeof(T) == 4 ? pick_kc(KC, f.ntx * f.nty, a.nz, slots) : KC;
    // FP32 fields of 2^32 bytes or more: buffer resources per field
    // component based at each chunk, when the planes a chunk touches (at
    // most kc+4 from k0-2), the material map and the metric tables fit
    // 32-bit offsets.  (In FP64 that form was slower than the general path
    // on 1024^3: 28.50 against 26.96 ms per step with 64 x 4 tiles.)
    constexpr int64_t LIM = 0xFFFFFFFFll;
    pcr = sizeof(T) == 4 && (!a.buf32 || a.pcr) && a.geo_bytes < LIM && a.N < LIM &&
          (int64_t)(kc + 8) * a.nx * a.ny * (int64_t)sizeof(T) < LIM;
    chunk_layout(f, a.nz, kc, sizeof(T) == 4 && (a.buf32 || pcr));
  }
  if constexpr (sizeof(T) == 4) {
    if (pcr) {
      if (a.nt)
        hipLaunchKernelGGL((fused_sweep<T, TX, TY, KC, true, true>), dim3(f.nb), dim3(TX * TY),
                           fused_mt_bytes<T>(a.nmat), s, f);
      else
        hipLaunchKernelGGL((fused_sweep<T, TX, TY, KC, false, true>), dim3(f.nb), dim3(TX * TY),
                           fused_mt_bytes<T>(a.nmat), s, f);
      return true;
    }
  }
  if (a.nt)
    hipLaunchKernelGGL((fused_sweep<T, TX, TY, KC, true>), dim3(f.nb), dim3(TX * TY),
                       fused_mt_bytes<T>(a.nmat), s, f);
  else
    hipLaunchKernelGGL((fused_sweep<T, TX, TY, KC, false>), dim3(f.nb), dim3(TX * TY),
                       fused_mt_bytes<T>(a.nmat), s, f);
  return true;
  }
}

// Resident blocks per CU of the single-grid sweep of a tile configuration
// (0 if the configuration does not fit this dtype).
template <typename T, int TX, int TY, int KC>
inline int occ_cfg() {
  if constexpr (!cfg_fits<T, TX, TY>()) {
    return 0;
  } else {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_sweep<T, TX, TY, KC, false>,
                                                     TX * TY, fused_mt_bytes<T>(1)) != hipSuccess)
      return 0;
    return n;
  }
}

template <typename T>
inline int fused_blocks_per_cu(int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: return occ_cfg<T, 128, 4, 32>();
    case FUSED_CFG_64x8: return occ_cfg<T, 64, 8, 32>();
    case FUSED_CFG_256x4: return occ_cfg<T, 256, 4, 32>();
    case FUSED_CFG_64x4: return occ_cfg<T, 64, 4, 32>();
    case FUSED_CFG_256x4_K64: return occ_cfg<T, 256, 4, 64>();
    default: return occ_cfg<T, 320, 3, 64>();
  }
}

// false: nothing launched (see launch_cfg)
template <typename T>
inline bool launch_fused(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s, int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: return launch_cfg<T, 128, 4, 32>(a, En, Hn, s);
    case FUSED_CFG_64x8: return launch_cfg<T, 64, 8, 32>(a, En, Hn, s);
    case FUSED_CFG_256x4: return launch_cfg<T, 256, 4, 32>(a, En, Hn, s);
    case FUSED_CFG_64x4: return launch_cfg<T, 64, 4, 32>(a, En, Hn, s);
    case FUSED_CFG_256x4_K64: return launch_cfg<T, 256, 4, 64>(a, En, Hn, s);
    default: return launch_cfg<T, 320, 3, 64>(a, En, Hn, s);
  }
}

template <typename T>
inline bool fused_spill_free(int cfg) {
  switch (cfg) {
    case FUSED_CFG_128x4: return fused_cfg_spill_free<T, 128, 4, 32>();
    case FUSED_CFG_64x8: return fused_cfg_spill_free<T, 64, 8, 32>();
    case FUSED_CFG_256x4: return fused_cfg_spill_free<T, 256, 4, 32>();
    case FUSED_CFG_64x4: return fused_cfg_spill_free<T, 64, 4, 32>();
    case FUSED_CFG_256x4_K64: return fused_cfg_spill_free<T, 256, 4, 64>();
    default: return fused_cfg_spill_free<T, 320, 3, 64>();
  }
}

}  // namespace fit
