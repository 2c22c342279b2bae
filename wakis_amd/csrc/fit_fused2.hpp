// Two-step fused sweep: one launch advances E and H by TWO leapfrog steps,
// reading E, H (old buffers) once and writing E, H (new buffers) once per
// cell -- 12 field words per cell per launch for 2 cell-steps, against 12 per
// cell-step for the one-step sweep (fit_fused.hpp).
//
// A workgroup owns a TX x TY tile of (i, j) columns and a chunk of planes and
// marches up in z.  It holds a flat window of (TX+4) x (TY+4) cells per plane
// in LDS: window cell (wi, wj) of plane k is the cell at FLAT index
//   (x0 - 2 + wi) + nx * (y0 - 2 + wj) + nx * ny * k
// (zero when outside [0, N)), so, as in the one-step sweep, every stencil
// neighbour n +- 1, n +- nx, n +- nx*ny of a window cell is the adjacent
// window cell and the reference's helical x/y wrap and dropped neighbours
// (wakis/solverFIT3D.py:207-227) come out exactly, however deep the halo.
// Per plane the window goes through both steps with a one-plane skew:
//   H1(k)   on the window minus its high row/column  (reads u0 of k, k+1)
//   E1(k)   on that minus the low row/column          (reads w1 of k, k-1)
//   H2(k-1) on the window minus two high and one low  (reads u1 of k-1, k)
//   E2(k-1) on the TX x TY tile                       (reads w2 of k-1, k-2)
// with u = L*pec*E and w = tL*pmc*H pre-scaled by the metric and mask of the
// cell each value really is, and the same per-cell arithmetic, in the same
// order, as the one-step kernels.  The first step's J is sigma*E_old (or 0),
// the second's sigma*E1 (or 0); the beam line overrides J_z on its column
// with each step's value; the E_z probe records E_old and E1 of the two
// steps.  Every value a tile stores comes from window cells the tile
// computed itself, so no workgroup waits for another.
//
// One thread per window cell ((TX+4)(TY+4) threads): every thread runs the
// same code on its cell, the phases its cell takes part in predicated by its
// position, so each thread holds the state of one cell only.  LDS: four
// 2-plane rings (u0, w1, u1, w2) of 3 x (TX+4)(TY+4) values; the material
// table follows in dynamic LDS.
#pragma once

#include <type_traits>

#include "fit_fused.hpp"

namespace fit {

template <int TX, int TY>
struct Win2 {
  static constexpr int WW = TX + 4;   // window row length
  static constexpr int WH = TY + 4;   // window rows
  static constexpr int WC = WW * WH;  // window cells per plane = threads
};

template <typename T, int TX, int TY>
struct Smem2 {
  using W = Win2<TX, TY>;
  T u0[2][3][W::WC];  // u of E_old, planes k and k+1
  T w1[2][3][W::WC];  // w of H1, planes k and k-1
  T u1[2][3][W::WC];  // u of E1, planes k and k-1
  T w2[2][3][W::WC];  // w of H2, planes k-1 and k-2
};

// Waves per SIMD the lean kernels are compiled for: 4 in FP32 (at most 128
// VGPRs), 2 in FP64 (at most 256; a block of at most 8 waves).
#define F2_WAVES(T) (sizeof(T) == 4 ? 4 : 2)

template <typename T>
struct Fused2Args {
  StepArgs<T> a;           // a.E, a.H: buffers read; a.step: index of the first step
  T *En, *Hn;              // buffers written
  const T *beam1, *beam2;  // J_z line (nz values) of step 1 / step 2, or null
  T js1, js2;              // 1: J = sigma * (the step's input E); 0: J = 0
  int ntx, nty, kc, nchunk, nb, nremap;
  int kh = 0, nmid = 0;  // z face chunks of kh planes (chunk2_span)
};

// Planes [k0, k1) of z chunk c: with kh > 0 the first nmid chunks take kc
// planes each from plane kh (the last ending at nz - kh), chunk nmid is
// [0, kh) and chunk nmid + 1 is [nz - kh, nz); with kh = 0 chunk c is
// [c kc, (c + 1) kc).
template <typename T>
__host__ __device__ inline void chunk2_span(const Fused2Args<T> &f, int c, int &k0, int &k1) {
  const int nz = f.a.nz, kc = f.kc, kh = f.kh;
  if (kh == 0) {
    k0 = c * kc;
    k1 = k0 + kc < nz ? k0 + kc : nz;
  } else if (c < f.nmid) {
    k0 = kh + c * kc;
    k1 = k0 + kc < nz - kh ? k0 + kc : nz - kh;
  } else if (c == f.nmid) {
    k0 = 0;
    k1 = kh;
  } else {
    k0 = nz - kh;
    k1 = nz;
  }
}

// One tile of a chunk.  LEAN: every plane the chunk's outputs depend on (its
// window cells' true planes) lies in 1 .. nz-2, so validity is constant and
// the boundary masks reduce to per-cell 0/1 factors fixed over the march
// (the same expressions, with fewer instructions per plane).
// PCR: fields addressed with one buffer resource per component, based at
// the chunk's lowest plane (fields of 2^32 bytes or more), as the one-step
// sweep's lean path does; else one resource per field.
template <typename T, int TX, int TY, bool NTS, bool LEAN, bool PCR = false>
__device__ __forceinline__ void fused2_tile(const Fused2Args<T> &f, const int tile,
                                            Smem2<T, TX, TY> &sm, T *__restrict__ smt) {
  using W = Win2<TX, TY>;
  constexpr int WW = W::WW, WC = W::WC;
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int plane = nx * ny;
  const int tx = tile % f.ntx;
  const int tr = tile / f.ntx;
  const int ty = tr % f.nty;
  const int ch = tr / f.nty;
  const int x0 = tx * TX, y0 = ty * TY;
  int k0, k1;
  chunk2_span(f, ch, k0, k1);

  const Geo<T> &g = a.g;
  const CT<T> *mtab = (const CT<T> *)a.mtab;
  const uint32_t pec = a.pec, pmc = a.pmc;
  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += WC) smt[t] = mtab[t];

  auto clampk = [&](int kk) __attribute__((always_inline)) { return min(max(kk, 0), nz - 1); };
  auto fac = [](bool masked) __attribute__((always_inline)) { return masked ? T(0) : T(1); };

  // ---- this thread's window cell ----
  const int cl = threadIdx.x;  // LDS index wj * WW + wi
  const int wj = cl / WW, wi = cl - wj * WW;
  const WinCell wc = win_cell(x0 - 2 + wi, y0 - 2 + wj, nx, ny);
  const int ci = wc.i, cj = wc.j, cdk = wc.dk;
  const int ccol = ci + nx * cj;
  // roles: H1 and E1 of step 1, H2 of step 2, output (the tile)
  const bool rh1 = wi <= TX + 2 && wj <= TY + 2;
  const bool re1 = rh1 && wi >= 1 && wj >= 1;
  const bool rh2 = wi >= 1 && wj >= 1 && wi <= TX + 1 && wj <= TY + 1;
  // (a tile may overhang the grid's high x/y edges: its cells there are the
  // flat continuation, computed like any other, but stored by their own tile)
  const bool rout = wi >= 2 && wj >= 2 && wi < TX + 2 && wj < TY + 2 && x0 - 2 + wi < nx &&
                    y0 - 2 + wj < ny;
  bool mpe[3], mpm[3];  // (i, j) part of the pec / pmc masks
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mpe[d] = xy_mask(pec, d, ci, cj, nx, ny);
    mpm[d] = xy_mask(pmc, d, ci, cj, nx, ny);
  }
  const T cLx = g.Lx[ci], cLy = g.Ly[cj], ctLx = g.tLx[ci], ctLy = g.tLy[cj];
  const T ciAz = g.iAz[ccol], citAz = g.itAz[ccol];
  const bool cbeam = a.beam_mode != BEAM_OFF && ci == a.ixs && cj == a.iys;
  // wake probe window of this column (tile cells only), or -1
  int pw = -1;
  if (a.hist && rout) {
    const int di = ci - a.ixt, dj = cj - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw)
      pw = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
  }
  const bool probe1 = pw >= 0 && a.step < a.hist_nt;
  const bool probe2 = pw >= 0 && a.step + 1 < a.hist_nt;
  GT<T> *phist = pw >= 0 ? (GT<T> *)a.hist + (int64_t)pw * nz * a.hist_nt + a.step : nullptr;

  // boundary-mask factors (pec: E columns, pmc: H rows) at true plane kk
  const T peF[3] = {fac(mpe[0]), fac(mpe[1]), fac(mpe[2])};
  const T pmF[3] = {fac(mpm[0]), fac(mpm[1]), fac(mpm[2])};
  auto pe = [&](int c, int kk) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return peF[c];
    else return fac(full_mask(pec, c, mpe[c], kk, nz));
  };
  auto pm = [&](int c, int kk) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return pmF[c];
    else return fac(full_mask(pmc, c, mpm[c], kk, nz));
  };
  auto valid = [&](int kk) __attribute__((always_inline)) -> bool {
    if constexpr (LEAN) return true;
    else return kk >= 0 && kk < nz;
  };

  // ---- loads: CDNA buffer instructions ----
  // One resource per array (the fields span < 2^32 bytes), a per-lane byte
  // offset and the component offset c N sizeof(T) as the uniform one; or
  // (PCR) one resource per field component based at plane kbr, the lowest
  // the chunk touches, with plane offsets (k - kbr) PB.  Planes are clamped
  // to the grid, so every access is in range; values at clamped planes are
  // never used (their u and w are 0).
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t NB = (uint32_t)a.N * SZ;
  const uint32_t PB = (uint32_t)plane * SZ;
  const int64_t FB = 3 * a.N * (int64_t)SZ;
  const int kbr = PCR ? max(k0 - 3, 0) : 0;
  const int64_t ebr = (int64_t)kbr * plane;
  const int64_t cspan = PCR ? ((int64_t)a.N - ebr) * SZ : FB;
  const int64_t rspan = cspan < 0xFFFFFFFFll ? cspan : 0xFFFFFFFFll;
  auto comp_rsrc = [&](const T *p, int c) __attribute__((always_inline)) -> Rsrc {
    return PCR ? make_rsrc(p + (int64_t)c * a.N + ebr, rspan) : make_rsrc(p, FB);
  };
  const uint32_t co1 = PCR ? 0u : NB, co2 = PCR ? 0u : 2 * NB;
  const Rsrc rE0 = comp_rsrc(a.E, 0), rE1 = comp_rsrc(a.E, 1), rE2 = comp_rsrc(a.E, 2);
  const Rsrc rH0 = comp_rsrc(a.H, 0), rH1 = comp_rsrc(a.H, 1), rH2 = comp_rsrc(a.H, 2);
  const Rsrc rEn0 = comp_rsrc(f.En, 0), rEn1 = comp_rsrc(f.En, 1), rEn2 = comp_rsrc(f.En, 2);
  const Rsrc rHn0 = comp_rsrc(f.Hn, 0), rHn1 = comp_rsrc(f.Hn, 1), rHn2 = comp_rsrc(f.Hn, 2);
  const bool hasmat = a.mat != nullptr;
  const Rsrc rM = make_rsrc(hasmat ? (const void *)a.mat : (const void *)g.Lx, hasmat ? a.N : 1);
  const Rsrc rG = make_rsrc(g.Lx, a.geo_bytes);
  auto gofs = [&](const T *t) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)((const char *)t - (const char *)g.Lx);
  };
  // beam lines (a null line reads the L_z table instead: valid memory, value
  // unused)
  const bool b1on = f.beam1 != nullptr, b2on = f.beam2 != nullptr;
  const Rsrc rB1 = make_rsrc(b1on ? f.beam1 : g.Lz, (int64_t)nz * SZ);
  const Rsrc rB2 = make_rsrc(b2on ? f.beam2 : g.Lz, (int64_t)nz * SZ);
  const uint32_t cb = (uint32_t)ccol * SZ;
  const uint32_t oLz = gofs(g.Lz), otLz = gofs(g.tLz);
  const uint32_t oiAx = gofs(g.iAx) + (uint32_t)cj * SZ, oitAx = gofs(g.itAx) + (uint32_t)cj * SZ;
  const uint32_t oiAy = gofs(g.iAy) + (uint32_t)ci * SZ, oitAy = gofs(g.itAy) + (uint32_t)ci * SZ;
  const uint32_t rowx = (uint32_t)ny * SZ, rowy = (uint32_t)nx * SZ;  // (j,k) / (i,k) tables

  // staging data of one plane: E_old and L_z
  struct SPk {
    T e[3], lz;
  };
  // update data of one plane: H_old, material, areas, tL_z, beam values
  struct CPk {
    T h[3], iax, iay, itax, itay, tlz, b1, b2;
    int m;
  };
  auto load_S = [&](int k, SPk &p) __attribute__((always_inline)) {
    const int kr = clampk(k + cdk);
    const uint32_t v = cb + (uint32_t)(kr - kbr) * PB;
    p.e[0] = bload<T>(rE0, v, 0);
    p.e[1] = bload<T>(rE1, v, co1);
    p.e[2] = bload<T>(rE2, v, co2);
    p.lz = bload<T>(rG, oLz + (uint32_t)kr * SZ, 0);
  };
  auto load_C = [&](int k, CPk &p) __attribute__((always_inline)) {
    const int kr = clampk(k + cdk);
    const uint32_t v = cb + (uint32_t)(kr - kbr) * PB;
    p.h[0] = bload<T>(rH0, v, 0);
    p.h[1] = bload<T>(rH1, v, co1);
    p.h[2] = bload<T>(rH2, v, co2);
    p.m = hasmat ? bload_u8(rM, (uint32_t)ccol + (uint32_t)kr * (uint32_t)plane, 0) : 0;
    p.iax = bload<T>(rG, oiAx + (uint32_t)kr * rowx, 0);
    p.iay = bload<T>(rG, oiAy + (uint32_t)kr * rowy, 0);
    p.itax = bload<T>(rG, oitAx + (uint32_t)kr * rowx, 0);
    p.itay = bload<T>(rG, oitAy + (uint32_t)kr * rowy, 0);
    p.tlz = bload<T>(rG, otLz + (uint32_t)kr * SZ, 0);
    p.b1 = bload<T>(rB1, (uint32_t)kr * SZ, 0);
    p.b2 = bload<T>(rB2, (uint32_t)kr * SZ, 0);
  };

  auto curl_u = [&](const T (*u)[WC], const T (*un)[WC], int d)
      __attribute__((always_inline)) -> T {
    const int l = cl;
    if (d == 0) return ((u[1][l] - un[1][l]) - u[2][l]) + u[2][l + WW];
    if (d == 1) return ((u[2][l] - u[2][l + 1]) - u[0][l]) + un[0][l];
    return ((u[0][l] - u[0][l + WW]) - u[1][l]) + u[1][l + 1];
  };
  auto curl_w = [&](const T (*w)[WC], const T (*wm)[WC], int d)
      __attribute__((always_inline)) -> T {
    const int l = cl;
    if (d == 0) return ((w[2][l] - w[2][l - WW]) - w[1][l]) + wm[1][l];
    if (d == 1) return ((w[0][l] - wm[0][l]) - w[2][l]) + w[2][l - 1];
    return ((w[1][l] - w[1][l - 1]) - w[0][l]) + w[0][l - WW];
  };
  // u = L pec E of this cell at window plane k into ring slot u
  auto put_u = [&](T (*u)[WC], int k, const T (&e)[3], T lz) __attribute__((always_inline)) {
    const int kr = k + cdk;
    const bool v = valid(kr);
    u[0][cl] = v ? (cLx * pe(0, kr)) * e[0] : T(0);
    u[1][cl] = v ? (cLy * pe(1, kr)) * e[1] : T(0);
    u[2][cl] = v ? pe(2, kr) * (lz * e[2]) : T(0);
  };

  // ---- pipeline state ----
  SPk S;                   // E_old of the plane staged next (in flight)
  T E0c[3], E0n[3];        // raw E_old of planes k / k+1
  T lzc = 0, lzn = 0;      // L_z of planes k / k+1
  T H1p[3], E1p[3];        // raw H1 and E1 of plane k-1
  int mp = 0;              // material of plane k-1
  T iaxp = 0, iayp = 0, itaxp = 0, itayp = 0, tlzp = 0, b2p = 0;  // plane k-1
#pragma unroll
  for (int d = 0; d < 3; ++d) E0c[d] = E0n[d] = H1p[d] = E1p[d] = T(0);

  // prologue: u0 of plane k0-2 staged; S = E_old of plane k0-1; C of k0-2
  const int kb = k0 - 2;
  load_S(kb, S);
  put_u(sm.u0[0], kb, S.e, S.lz);
  load_S(kb + 1, S);
  CPk CA, CB;
  load_C(kb, CA);

  // One plane: stage u0(k+1); H1(k); E1(k); H2(k-1); E2(k-1).  Cc: update
  // data of plane k (loaded one iteration earlier); Cn receives plane k+1's.
  // PH: phases run besides staging and H1 (1 = E1, 2 = H2, 4 = E2); the three
  // pipeline-fill iterations before the loop run fewer.  PA: LDS ring slot of
  // plane k, (k - kb) & 1, a compile-time constant in every unrolled copy.
  auto iter = [&](const int k, const CPk &Cc, CPk &Cn, auto ph, auto pa)
      __attribute__((always_inline)) {
    constexpr int PH = decltype(ph)::value;
    constexpr int PA = decltype(pa)::value, PB1 = PA ^ 1;
    // A. stage u0(k+1); keep its raw E and L_z for E1(k+1)
    put_u(sm.u0[PB1], k + 1, S.e, S.lz);
#pragma unroll
    for (int d = 0; d < 3; ++d) E0n[d] = S.e[d];
    lzn = S.lz;
    // B. loads of the next iteration (unconditional, at clamped planes: a
    // conditional load would make the compiler merge and copy registers of
    // loads in flight)
    load_S(k + 2, S);
    load_C(k + 1, Cn);
    __syncthreads();

    // C. H1(k) -> w1 slot PA
    T H1n[3] = {T(0), T(0), T(0)};
    if (rh1) {
      const int kr = k + cdk;
      const T *mt = smt + Cc.m * MAT_STRIDE;
      const T fm0 = pm(0, kr), fm1 = pm(1, kr), fm2 = pm(2, kr);
      const T(*u)[WC] = sm.u0[PA];
      const T(*un)[WC] = sm.u0[PB1];
      H1n[0] = Cc.h[0] - (mt[0] * (fm0 * Cc.iax)) * curl_u(u, un, 0);
      H1n[1] = Cc.h[1] - (mt[1] * (fm1 * Cc.iay)) * curl_u(u, un, 1);
      H1n[2] = Cc.h[2] - (mt[2] * (fm2 * ciAz)) * curl_u(u, un, 2);
      const T vf = valid(kr) ? T(1) : T(0);
      T(*w)[WC] = sm.w1[PA];
      w[0][cl] = (vf * (fm0 * ctLx)) * H1n[0];
      w[1][cl] = (vf * (fm1 * ctLy)) * H1n[1];
      w[2][cl] = (vf * (fm2 * Cc.tlz)) * H1n[2];
    }
    __syncthreads();

    // D. E1(k) -> u1 slot PA; probe both steps (tile cells)
    T E1n[3] = {T(0), T(0), T(0)};
    if constexpr ((PH & 1) != 0) {
      if (re1) {
        const int kr = k + cdk;
        const T *mt = smt + Cc.m * MAT_STRIDE;
        const T(*w)[WC] = sm.w1[PA];
        const T(*wm)[WC] = sm.w1[PB1];
        const T ax = pe(0, kr) * Cc.itax;
        const T ay = pe(1, kr) * Cc.itay;
        const T az = pe(2, kr) * citAz;
        const T jx = f.js1 * (mt[6] * E0c[0]);
        const T jy = f.js1 * (mt[7] * E0c[1]);
        const T jz = (cbeam && b1on) ? Cc.b1 : f.js1 * (mt[8] * E0c[2]);
        E1n[0] = E0c[0] + mt[3] * (ax * curl_w(w, wm, 0) - jx);
        E1n[1] = E0c[1] + mt[4] * (ay * curl_w(w, wm, 1) - jy);
        E1n[2] = E0c[2] + mt[5] * (az * curl_w(w, wm, 2) - jz);
        put_u(sm.u1[PA], k, E1n, lzc);
        if (k >= k0 && k < k1) {
          if (probe1) phist[(int64_t)k * a.hist_nt] = E0c[2];
          if (probe2) phist[(int64_t)k * a.hist_nt + 1] = E1n[2];
        }
      }
    }
    __syncthreads();

    // E. H2(k-1) -> w2 slot PB1; store H2(k-1) (tile cells)
    if constexpr ((PH & 2) != 0) {
      if (rh2) {
        const int km = k - 1;
        const int kr = km + cdk;
        const T *mt = smt + mp * MAT_STRIDE;
        const T fm0 = pm(0, kr), fm1 = pm(1, kr), fm2 = pm(2, kr);
        const T(*u)[WC] = sm.u1[PB1];
        const T(*un)[WC] = sm.u1[PA];
        const T h0 = H1p[0] - (mt[0] * (fm0 * iaxp)) * curl_u(u, un, 0);
        const T h1 = H1p[1] - (mt[1] * (fm1 * iayp)) * curl_u(u, un, 1);
        const T h2 = H1p[2] - (mt[2] * (fm2 * ciAz)) * curl_u(u, un, 2);
        const T vf = valid(kr) ? T(1) : T(0);
        T(*w)[WC] = sm.w2[PB1];
        w[0][cl] = (vf * (fm0 * ctLx)) * h0;
        w[1][cl] = (vf * (fm1 * ctLy)) * h1;
        w[2][cl] = (vf * (fm2 * tlzp)) * h2;
        if (rout && km >= k0 && km < k1) {
          const uint32_t v = cb + (uint32_t)(km - kbr) * PB;
          bstore(rHn0, v, 0, h0, NTS);
          bstore(rHn1, v, co1, h1, NTS);
          bstore(rHn2, v, co2, h2, NTS);
        }
      }
    }
    __syncthreads();

    // F. E2(k-1) of the tile; store it
    if constexpr ((PH & 4) != 0) {
      const int km = k - 1;
      if (rout && km < k1) {
        const T *mt = smt + mp * MAT_STRIDE;
        const T(*w)[WC] = sm.w2[PB1];
        const T(*wm)[WC] = sm.w2[PA];  // plane k-2
        const T ax = pe(0, km) * itaxp;
        const T ay = pe(1, km) * itayp;
        const T az = pe(2, km) * citAz;
        const T jx = f.js2 * (mt[6] * E1p[0]);
        const T jy = f.js2 * (mt[7] * E1p[1]);
        const T jz = (cbeam && b2on) ? b2p : f.js2 * (mt[8] * E1p[2]);
        const uint32_t v = cb + (uint32_t)(km - kbr) * PB;
        bstore(rEn0, v, 0, E1p[0] + mt[3] * (ax * curl_w(w, wm, 0) - jx), NTS);
        bstore(rEn1, v, co1, E1p[1] + mt[4] * (ay * curl_w(w, wm, 1) - jy), NTS);
        bstore(rEn2, v, co2, E1p[2] + mt[5] * (az * curl_w(w, wm, 2) - jz), NTS);
      }
    }

    // plane k becomes the previous plane
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      H1p[d] = H1n[d];
      E1p[d] = E1n[d];
      E0c[d] = E0n[d];
    }
    lzc = lzn;
    mp = Cc.m;
    iaxp = Cc.iax;
    iayp = Cc.iay;
    itaxp = Cc.itax;
    itayp = Cc.itay;
    tlzp = Cc.tlz;
    b2p = Cc.b2;
  };

  __syncthreads();  // u0(k0-2) staged, material table loaded
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P3 = std::integral_constant<int, 3>;
  using P7 = std::integral_constant<int, 7>;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  iter(kb, CA, CB, P0{}, S0{});      // H1(k0-2)
  iter(kb + 1, CB, CA, P1{}, S1{});  // H1, E1 of k0-1
  iter(kb + 2, CA, CB, P3{}, S0{});  // H1, E1 of k0; H2(k0-1)
  // iterations k0+1 .. k1, plus one more when their number is odd (it
  // computes planes beyond the chunk and stores nothing), so that the loop
  // body, written out twice, has no conditional part
  const int kend = ((k1 - k0) & 1) ? k1 + 1 : k1;
  for (int k = k0 + 1; k <= kend; k += 2) {
    iter(k, CB, CA, P7{}, S1{});
    iter(k + 1, CA, CB, P7{}, S0{});
  }
}

// The same tile step with two barriers per plane instead of four, for the
// lean chunks: all H updates of a plane, then all E updates,
//   P(k): H1(k)   (reads u0 of k, k+1)    and H2(k-2) (reads u1 of k-2, k-1)
//   Q(k): E1(k)   (reads w1 of k, k-1)    and E2(k-2) (reads w2 of k-2, k-3),
//         staging of u0(k+2)
// so the second step lags two planes; every ring keeps 2 planes (a plane's
// slot is its parity) and each thread carries its cell's H1, E1 and update
// data of the two previous planes.  Same per-cell expressions as
// fused2_tile.
template <typename T, int TX, int TY, bool NTS, bool PCR = false>
__device__ __forceinline__ void fused2_tile_lean2b(const Fused2Args<T> &f, const int tile,
                                                   Smem2<T, TX, TY> &sm, T *__restrict__ smt) {
  using W = Win2<TX, TY>;
  constexpr int WW = W::WW, WC = W::WC;
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int plane = nx * ny;
  const int tx = tile % f.ntx;
  const int tr = tile / f.ntx;
  const int ty = tr % f.nty;
  const int ch = tr / f.nty;
  const int x0 = tx * TX, y0 = ty * TY;
  int k0, k1;
  chunk2_span(f, ch, k0, k1);

  const Geo<T> &g = a.g;
  const CT<T> *mtab = (const CT<T> *)a.mtab;
  const uint32_t pec = a.pec, pmc = a.pmc;
  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += WC) smt[t] = mtab[t];
  auto fac = [](bool masked) __attribute__((always_inline)) { return masked ? T(0) : T(1); };

  // ---- this thread's window cell ----
  const int cl = threadIdx.x;
  const int wj = cl / WW, wi = cl - wj * WW;
  const WinCell wc = win_cell(x0 - 2 + wi, y0 - 2 + wj, nx, ny);
  const int ci = wc.i, cj = wc.j, cdk = wc.dk;
  const int ccol = ci + nx * cj;
  const bool rh1 = wi <= TX + 2 && wj <= TY + 2;
  const bool re1 = rh1 && wi >= 1 && wj >= 1;
  const bool rh2 = wi >= 1 && wj >= 1 && wi <= TX + 1 && wj <= TY + 1;
  const bool rout = wi >= 2 && wj >= 2 && wi < TX + 2 && wj < TY + 2 && x0 - 2 + wi < nx &&
                    y0 - 2 + wj < ny;
  // lean: interior planes only, so the masks are their (i, j) parts
  const T pe0 = fac(xy_mask(pec, 0, ci, cj, nx, ny)), pe1 = fac(xy_mask(pec, 1, ci, cj, nx, ny));
  const T pe2 = fac(xy_mask(pec, 2, ci, cj, nx, ny));
  const T pm0 = fac(xy_mask(pmc, 0, ci, cj, nx, ny)), pm1 = fac(xy_mask(pmc, 1, ci, cj, nx, ny));
  const T pm2 = fac(xy_mask(pmc, 2, ci, cj, nx, ny));
  const T uxF = g.Lx[ci] * pe0, uyF = g.Ly[cj] * pe1;
  const T wxF = pm0 * g.tLx[ci], wyF = pm1 * g.tLy[cj];
  const T iAz2 = pm2 * g.iAz[ccol];
  const T itAz2 = pe2 * g.itAz[ccol];
  const bool cbeam = a.beam_mode != BEAM_OFF && ci == a.ixs && cj == a.iys;
  int pw = -1;
  if (a.hist && rout) {
    const int di = ci - a.ixt, dj = cj - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw)
      pw = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
  }
  const bool probe1 = pw >= 0 && a.step < a.hist_nt;
  const bool probe2 = pw >= 0 && a.step + 1 < a.hist_nt;
  GT<T> *phist = pw >= 0 ? (GT<T> *)a.hist + (int64_t)pw * nz * a.hist_nt + a.step : nullptr;

  // ---- loads (as fused2_tile; every plane read lies inside the grid) ----
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t NB = (uint32_t)a.N * SZ;
  const uint32_t PB = (uint32_t)plane * SZ;
  const int64_t FB = 3 * a.N * (int64_t)SZ;
  // fields: as fused2_tile (PCR: per-component resources based at plane
  // kbr = k0 - 3, the lowest a lean chunk touches, which is >= 1)
  const int kbr = PCR ? k0 - 3 : 0;
  const int64_t ebr = (int64_t)kbr * plane;
  const int64_t cspan = PCR ? ((int64_t)a.N - ebr) * SZ : FB;
  const int64_t rspan = cspan < 0xFFFFFFFFll ? cspan : 0xFFFFFFFFll;
  auto comp_rsrc = [&](const T *p, int c) __attribute__((always_inline)) -> Rsrc {
    return PCR ? make_rsrc(p + (int64_t)c * a.N + ebr, rspan) : make_rsrc(p, FB);
  };
  const uint32_t co1 = PCR ? 0u : NB, co2 = PCR ? 0u : 2 * NB;
  const Rsrc rE0 = comp_rsrc(a.E, 0), rE1 = comp_rsrc(a.E, 1), rE2 = comp_rsrc(a.E, 2);
  const Rsrc rH0 = comp_rsrc(a.H, 0), rH1 = comp_rsrc(a.H, 1), rH2 = comp_rsrc(a.H, 2);
  const Rsrc rEn0 = comp_rsrc(f.En, 0), rEn1 = comp_rsrc(f.En, 1), rEn2 = comp_rsrc(f.En, 2);
  const Rsrc rHn0 = comp_rsrc(f.Hn, 0), rHn1 = comp_rsrc(f.Hn, 1), rHn2 = comp_rsrc(f.Hn, 2);
  const bool hasmat = a.mat != nullptr;
  const Rsrc rM = make_rsrc(hasmat ? (const void *)a.mat : (const void *)g.Lx, hasmat ? a.N : 1);
  const Rsrc rG = make_rsrc(g.Lx, a.geo_bytes);
  auto gofs = [&](const T *t) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)((const char *)t - (const char *)g.Lx);
  };
  const bool b1on = f.beam1 != nullptr, b2on = f.beam2 != nullptr;
  const Rsrc rB1 = make_rsrc(b1on ? f.beam1 : g.Lz, (int64_t)nz * SZ);
  const Rsrc rB2 = make_rsrc(b2on ? f.beam2 : g.Lz, (int64_t)nz * SZ);
  const uint32_t cb = (uint32_t)ccol * SZ;
  // plane offset of the carry, relative to the resources' base plane
  const uint32_t cbk = cb + (uint32_t)(cdk - kbr) * PB;
  const uint32_t cmk = (uint32_t)ccol + (uint32_t)cdk * (uint32_t)plane;
  // metric tables and beam lines: one lane offset per table layout (its
  // plane part added per plane), the table's offset in the geometry block
  // in the wave-uniform SGPR operand
  const uint32_t vZ = (uint32_t)cdk * SZ;
  const uint32_t rowx = (uint32_t)ny * SZ, rowy = (uint32_t)nx * SZ;
  const uint32_t vAx = (uint32_t)cj * SZ + (uint32_t)cdk * rowx;
  const uint32_t vAy = (uint32_t)ci * SZ + (uint32_t)cdk * rowy;
  const uint32_t oLz = gofs(g.Lz), otLz = gofs(g.tLz);
  const uint32_t oiAx = gofs(g.iAx), oitAx = gofs(g.itAx);
  const uint32_t oiAy = gofs(g.iAy), oitAy = gofs(g.itAy);

  struct SPk {
    T e[3], lz;
  };
  struct CPk {
    T h[3], iax, iay, itax, itay, tlz, b1, b2;
    int m;
  };
  // window plane k (true plane k + cdk, inside the grid for a lean chunk)
  auto load_S = [&](int k, SPk &p) __attribute__((always_inline)) {
    const uint32_t v = cbk + (uint32_t)k * PB;
    p.e[0] = bload<T>(rE0, v, 0);
    p.e[1] = bload<T>(rE1, v, co1);
    p.e[2] = bload<T>(rE2, v, co2);
    p.lz = bload<T>(rG, vZ + (uint32_t)k * SZ, oLz);
  };
  auto load_C = [&](int k, CPk &p) __attribute__((always_inline)) {
    const uint32_t v = cbk + (uint32_t)k * PB;
    p.h[0] = bload<T>(rH0, v, 0);
    p.h[1] = bload<T>(rH1, v, co1);
    p.h[2] = bload<T>(rH2, v, co2);
    p.m = hasmat ? bload_u8(rM, cmk + (uint32_t)k * (uint32_t)plane, 0) : 0;
    const uint32_t vax = vAx + (uint32_t)k * rowx, vay = vAy + (uint32_t)k * rowy;
    const uint32_t vz = vZ + (uint32_t)k * SZ;
    p.iax = bload<T>(rG, vax, oiAx);
    p.iay = bload<T>(rG, vay, oiAy);
    p.itax = bload<T>(rG, vax, oitAx);
    p.itay = bload<T>(rG, vay, oitAy);
    p.tlz = bload<T>(rG, vz, otLz);
    p.b1 = bload<T>(rB1, vz, 0);
    p.b2 = bload<T>(rB2, vz, 0);
  };
  auto curl_u = [&](const T (*u)[WC], const T (*un)[WC], int d)
      __attribute__((always_inline)) -> T {
    const int l = cl;
    if (d == 0) return ((u[1][l] - un[1][l]) - u[2][l]) + u[2][l + WW];
    if (d == 1) return ((u[2][l] - u[2][l + 1]) - u[0][l]) + un[0][l];
    return ((u[0][l] - u[0][l + WW]) - u[1][l]) + u[1][l + 1];
  };
  auto curl_w = [&](const T (*w)[WC], const T (*wm)[WC], int d)
      __attribute__((always_inline)) -> T {
    const int l = cl;
    if (d == 0) return ((w[2][l] - w[2][l - WW]) - w[1][l]) + wm[1][l];
    if (d == 1) return ((w[0][l] - wm[0][l]) - w[2][l]) + w[2][l - 1];
    return ((w[1][l] - w[1][l - 1]) - w[0][l]) + w[0][l - WW];
  };
  auto put_u = [&](T (*u)[WC], const T (&e)[3], T lz) __attribute__((always_inline)) {
    u[0][cl] = uxF * e[0];
    u[1][cl] = uyF * e[1];
    u[2][cl] = pe2 * (lz * e[2]);
  };

  // ---- pipeline state: plane k-1 ("p") and k-2 ("q") of this cell ----
  struct Prev {
    T H1[3], E1[3];
    T iax, iay, itax, itay, tlz, b2;
    int m;
  };
  SPk S;                         // E_old of the plane staged next (in flight)
  T E0a[3], E0b[3];              // raw E_old of planes k and k+1
  T lza = 0, lzb = 0;            // L_z of planes k and k+1
  Prev Pp{}, Pq{};
#pragma unroll
  for (int d = 0; d < 3; ++d) E0a[d] = E0b[d] = T(0);

  // prologue: u0 of planes k0-2 and k0-1 staged; S = E_old of plane k0;
  // C of plane k0-2 loaded
  const int kb = k0 - 2;
  load_S(kb, S);
  put_u(sm.u0[0], S.e, S.lz);
#pragma unroll
  for (int d = 0; d < 3; ++d) E0a[d] = S.e[d];
  lza = S.lz;
  load_S(kb + 1, S);
  put_u(sm.u0[1], S.e, S.lz);
#pragma unroll
  for (int d = 0; d < 3; ++d) E0b[d] = S.e[d];
  lzb = S.lz;
  load_S(kb + 2, S);
  // update data (H, material, metric, beam) is loaded CD planes ahead: two
  // in the FP64 shapes of 7 waves per CU (32 x 8, 16 x 16), which keep fewer
  // loads in flight than FP32's 13 (pillbox, 32 x 8: 1.837 against 1.925 ms
  // per launch); one elsewhere (FP64 28 x 12, at 256 VGPRs with two planes
  // ahead, was 3 % slower on 512 x 512 x 200)
  constexpr int CD = (sizeof(T) == 8 && TX * TY <= 256) ? 2 : 1;
  CPk CA, CB, CC;
  load_C(kb, CA);
  if constexpr (CD == 2) load_C(kb + 1, CB);

  // One plane k.  Cc: update data of plane k (loaded CD iterations
  // earlier), Cn receives plane k+CD's.  PH: 1 = E1(k), 2 = H2 and E2 of plane k-2 (the
  // pipeline-fill iterations before the loop run fewer).  PA: ring slot of
  // plane k, (k - kb) & 1, a compile-time constant in every unrolled copy.
  auto iter = [&](const int k, const CPk &Cc, CPk &Cn, auto ph, auto pa)
      __attribute__((always_inline)) {
    constexpr int PH = decltype(ph)::value;
    constexpr int PA = decltype(pa)::value, PB1 = PA ^ 1;
    load_C(k + CD, Cn);
    // P. H1(k) -> w1[PA]; H2(k-2) -> w2[PA]
    T H1n[3] = {T(0), T(0), T(0)};
    if (rh1) {
      const T *mt = smt + Cc.m * MAT_STRIDE;
      const T(*u)[WC] = sm.u0[PA];
      const T(*un)[WC] = sm.u0[PB1];
      H1n[0] = Cc.h[0] - (mt[0] * (pm0 * Cc.iax)) * curl_u(u, un, 0);
      H1n[1] = Cc.h[1] - (mt[1] * (pm1 * Cc.iay)) * curl_u(u, un, 1);
      H1n[2] = Cc.h[2] - (mt[2] * iAz2) * curl_u(u, un, 2);
      T(*w)[WC] = sm.w1[PA];
      w[0][cl] = wxF * H1n[0];
      w[1][cl] = wyF * H1n[1];
      w[2][cl] = (pm2 * Cc.tlz) * H1n[2];
    }
    if constexpr ((PH & 2) != 0) {
      if (rh2) {
        const int km = k - 2;
        const T *mt = smt + Pq.m * MAT_STRIDE;
        const T(*u)[WC] = sm.u1[PA];    // plane k-2
        const T(*un)[WC] = sm.u1[PB1];  // plane k-1
        const T h0 = Pq.H1[0] - (mt[0] * (pm0 * Pq.iax)) * curl_u(u, un, 0);
        const T h1 = Pq.H1[1] - (mt[1] * (pm1 * Pq.iay)) * curl_u(u, un, 1);
        const T h2 = Pq.H1[2] - (mt[2] * iAz2) * curl_u(u, un, 2);
        T(*w)[WC] = sm.w2[PA];
        w[0][cl] = wxF * h0;
        w[1][cl] = wyF * h1;
        w[2][cl] = (pm2 * Pq.tlz) * h2;
        if (rout && km >= k0 && km < k1) {
          const uint32_t v = cb + (uint32_t)(km - kbr) * PB;
          bstore(rHn0, v, 0, h0, NTS);
          bstore(rHn1, v, co1, h1, NTS);
          bstore(rHn2, v, co2, h2, NTS);
        }
      }
    }
    __syncthreads();

    // Q. E1(k) -> u1[PA]; E2(k-2); stage u0(k+2) -> u0[PA]
    T E1n[3] = {T(0), T(0), T(0)};
    if constexpr ((PH & 1) != 0) {
      if (re1) {
        const T *mt = smt + Cc.m * MAT_STRIDE;
        const T(*w)[WC] = sm.w1[PA];
        const T(*wm)[WC] = sm.w1[PB1];
        const T ax = pe0 * Cc.itax;
        const T ay = pe1 * Cc.itay;
        const T jx = f.js1 * (mt[6] * E0a[0]);
        const T jy = f.js1 * (mt[7] * E0a[1]);
        const T jz = (cbeam && b1on) ? Cc.b1 : f.js1 * (mt[8] * E0a[2]);
        E1n[0] = E0a[0] + mt[3] * (ax * curl_w(w, wm, 0) - jx);
        E1n[1] = E0a[1] + mt[4] * (ay * curl_w(w, wm, 1) - jy);
        E1n[2] = E0a[2] + mt[5] * (itAz2 * curl_w(w, wm, 2) - jz);
        put_u(sm.u1[PA], E1n, lza);
        if (k >= k0 && k < k1) {
          if (probe1) phist[(int64_t)k * a.hist_nt] = E0a[2];
          if (probe2) phist[(int64_t)k * a.hist_nt + 1] = E1n[2];
        }
      }
    }
    if constexpr ((PH & 2) != 0) {
      const int km = k - 2;
      if (rout && km >= k0 && km < k1) {
        const T *mt = smt + Pq.m * MAT_STRIDE;
        const T(*w)[WC] = sm.w2[PA];    // plane k-2
        const T(*wm)[WC] = sm.w2[PB1];  // plane k-3
        const T ax = pe0 * Pq.itax;
        const T ay = pe1 * Pq.itay;
        const T jx = f.js2 * (mt[6] * Pq.E1[0]);
        const T jy = f.js2 * (mt[7] * Pq.E1[1]);
        const T jz = (cbeam && b2on) ? Pq.b2 : f.js2 * (mt[8] * Pq.E1[2]);
        const uint32_t v = cb + (uint32_t)(km - kbr) * PB;
        bstore(rEn0, v, 0, Pq.E1[0] + mt[3] * (ax * curl_w(w, wm, 0) - jx), NTS);
        bstore(rEn1, v, co1, Pq.E1[1] + mt[4] * (ay * curl_w(w, wm, 1) - jy), NTS);
        bstore(rEn2, v, co2, Pq.E1[2] + mt[5] * (itAz2 * curl_w(w, wm, 2) - jz), NTS);
      }
    }
    put_u(sm.u0[PA], S.e, S.lz);  // plane k+2 (u0 of plane k is no longer read)
    // rotate: planes k+1, k+2 become "k, k+1"; k, k-1 become "k-1, k-2"
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      E0a[d] = E0b[d];
      E0b[d] = S.e[d];
    }
    lza = lzb;
    lzb = S.lz;
    load_S(k + 3, S);
    Pq = Pp;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Pp.H1[d] = H1n[d];
      Pp.E1[d] = E1n[d];
    }
    Pp.m = Cc.m;
    Pp.iax = Cc.iax;
    Pp.iay = Cc.iay;
    Pp.itax = Cc.itax;
    Pp.itay = Cc.itay;
    Pp.tlz = Cc.tlz;
    Pp.b2 = Cc.b2;
    __syncthreads();
  };

  __syncthreads();  // u0 of k0-2, k0-1 staged, material table loaded
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P3 = std::integral_constant<int, 3>;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  if constexpr (CD == 1) {
    iter(kb, CA, CB, P0{}, S0{});      // H1(k0-2)
    iter(kb + 1, CB, CA, P1{}, S1{});  // H1, E1 of k0-1
    // iterations k0 .. k1+1 (H2/E2 of plane k-2 from k0-2 on; only planes
    // in [k0, k1) are stored), plus one more when their number is odd, so
    // that the loop body, written out twice, has no conditional part
    const int kend = ((k1 - k0) & 1) ? k1 + 3 : k1 + 2;
    for (int k = k0; k < kend; k += 2) {
      iter(k, CA, CB, P3{}, S0{});
      iter(k + 1, CB, CA, P3{}, S1{});
    }
  } else {
    // plane p's update data in packet (p - kb) mod 3 (CA, CB, CC); the
    // packet loaded in iteration k (plane k+2) is plane k-1's, consumed;
    // iterations k0 .. k1+1 rounded up to a multiple of 6 (ring parity x
    // packet), the extra ones storing nothing
    iter(kb, CA, CC, P0{}, S0{});
    iter(kb + 1, CB, CA, P1{}, S1{});
    const int kend = k0 + 6 * ((k1 + 2 - k0 + 5) / 6);
    for (int k = k0; k < kend; k += 6) {
      iter(k, CC, CB, P3{}, S0{});
      iter(k + 1, CA, CC, P3{}, S1{});
      iter(k + 2, CB, CA, P3{}, S0{});
      iter(k + 3, CC, CB, P3{}, S1{});
      iter(k + 4, CA, CC, P3{}, S0{});
      iter(k + 5, CB, CA, P3{}, S1{});
    }
  }
}

extern __shared__ double fit_dyn_lds[];

// The lean chunks (between the 4-plane z-face chunks) and the others are two
// kernels, so that each is compiled for its own register need: block b takes
// chunk b / (ntx nty) of the lean chunks (nremap blocks, under the XCD
// remap), or of the others, which follow them in the chunk order.
template <typename T, int TX, int TY, bool NTS, bool PCR = false>
__global__ void __launch_bounds__((TX + 4) * (TY + 4), F2_WAVES(T)) fused2_sweep(Fused2Args<T> f) {
  __shared__ Smem2<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  fused2_tile_lean2b<T, TX, TY, NTS, PCR>(f, xcd_tile(blockIdx.x, f.nremap), sm, smt);
}

template <typename T, int TX, int TY, bool NTS, bool PCR = false>
__global__ void __launch_bounds__((TX + 4) * (TY + 4)) fused2_sweep_faces(Fused2Args<T> f) {
  __shared__ Smem2<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  fused2_tile<T, TX, TY, NTS, false, PCR>(f, f.nremap + blockIdx.x, sm, smt);
}

// Both kinds of chunk in one launch: blocks below nremap run the lean
// chunks (XCD-remapped), the rest the z-face chunks, so that face blocks
// take the CUs the lean blocks leave idle instead of waiting for a second
// launch.  Used where it compiles spill-free (else the two kernels above).
template <typename T, int TX, int TY, bool NTS, bool PCR = false>
__global__ void __launch_bounds__((TX + 4) * (TY + 4), F2_WAVES(T)) fused2_sweep_all(Fused2Args<T> f) {
  __shared__ Smem2<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  if ((int)blockIdx.x < f.nremap)
    fused2_tile_lean2b<T, TX, TY, NTS, PCR>(f, xcd_tile(blockIdx.x, f.nremap), sm, smt);
  else
    fused2_tile<T, TX, TY, NTS, false, PCR>(f, blockIdx.x, sm, smt);
}

// Tile of the two-step sweep: 64 x 8 columns, a (68 x 12)-cell window, one
// thread per window cell (816 threads).  (64 x 7 balances the waves over the
// SIMDs, 12 against 13, and ran 10 % faster per plane, but 8 % slower per
// cell.)
constexpr int F2_TX = 64, F2_TY = 8;

// Shapes built per precision: in FP32 all but 28 x 12 (slower than one of
// the others on most grids); in FP64 those of at most 512 window cells
// (32 x 8, 16 x 16, 28 x 12), whose block of at most 8 waves runs at two
// waves per SIMD with up to 256 VGPRs (the FP64 body holds 189-214) and
// whose rings (77-98 KB) leave room for the material table in the LDS.
template <typename T, int TX, int TY>
constexpr bool cfg2_fits() {
  constexpr bool shape = sizeof(T) == 4 ? !(TX == 28 && TY == 12)
                                        : sizeof(T) == 8 && (TX + 4) * (TY + 4) <= 512;
  return shape &&
         sizeof(Smem2<T, TX, TY>) + MAX_MATERIALS * MAT_STRIDE * sizeof(T) <= 160 * 1024;
}

// Planes the z chunks of a two-step launch split: with 4-plane chunks at
// the z faces (when at least one 16-plane chunk fits between them) the
// planes between, else all nz planes.
inline int f2_span(int nz) { return nz - 8 >= 16 ? nz - 8 : nz; }

// Planes per chunk: blocks run in rounds of `slots` (an eighth per XCD); a
// round lasts about a chunk plus its 3 planes of pipeline fill; the least
// modelled time rounds x (kc + 3) among a few lengths and the equal splits
// of the span (m chunks of ceil(nz / m) planes, so that no short last chunk
// leaves its round's other blocks running alone), 16 to 384 planes; ties go
// to the longer chunk.
inline int pick_kc2(int ncol, int nz, int slots) {
  static const int cand[] = {16, 24, 32, 48, 64, 96, 128, 192, 256, 384};
  int best = 16;
  double bt = 1e300;
  const int xslots = slots / 8 > 0 ? slots / 8 : 1;
  auto consider = [&](int kc) {
    const long long blocks = (long long)ncol * ((nz + kc - 1) / kc);
    const long long per_xcd = (blocks + 7) / 8;
    const long long rounds = (per_xcd + xslots - 1) / xslots;
    const double t = (double)rounds * (kc + 3);
    if (t < bt || (t == bt && kc > best)) {
      bt = t;
      best = kc;
    }
  };
  for (int kc : cand)
    if (kc <= nz || kc == cand[0]) consider(kc);
  for (int m = 1; m <= 64; ++m) {
    const int kc = (nz + m - 1) / m;
    if (kc < 16) break;
    if (kc <= 384) consider(kc);
  }
  return best;
}

// ---- batched two-step sweep --------------------------------------------
// Every member of a batch advances two steps per launch: blocks [0, ntl)
// take the lean chunks of all members, blocks [ntl, ntl + ntf) their face
// chunks (or every chunk of a member too short for a lean chunk), so the
// short face blocks fill the tail of the long lean ones.  Block t of a part
// takes tile t - tstart[m] of the member m with tstart[m] <= t < tstart[m+1]
// (one table per part); the member's arguments of the current buffer parity
// come from a device table, with the step index and beam rows of pair g of
// the launch group taken from the device step counter, so a group can be
// captured and replayed.  The tile bodies are those of the single-grid
// launch.
template <typename T>
struct Batch2Member {
  Fused2Args<T> f[2];  // f[p]: buffers after p (mod 2) two-step launches
  int64_t step0;       // the member's step index when the table was built
  const T *btab;       // beam rows of the group's steps (G x nz), or null
};

// Tiles of the batched launch: 32 x 16 columns (a 36 x 20 window, 720
// threads) or 24 x 24 (28 x 28, 784 threads).  Against the 64 x 8 tile of the
// single-grid launch both leave less overhang on the narrow grids of a batch
// and have fewer window cells per tile cell (1.41 and 1.36 against 1.59);
// 24 x 24 needs ny >= 24.  Both are FP32 only (12-13 waves per block); FP64
// batches take 28 x 12 (32 x 16 window, 512 threads, 8 waves: two per SIMD,
// as the single-grid FP64 shapes).
enum { F2B_CFG_32x16 = 0, F2B_CFG_24x24 = 1, F2B_CFG_28x12 = 2, F2B_NCFG = 3 };

inline void f2b_dims(int cfg, int &tx, int &ty) {
  tx = cfg == F2B_CFG_24x24 ? 24 : cfg == F2B_CFG_28x12 ? 28 : 32;
  ty = cfg == F2B_CFG_24x24 ? 24 : cfg == F2B_CFG_28x12 ? 12 : 16;
}

template <typename T, int TX, int TY>
__global__ void __launch_bounds__((TX + 4) * (TY + 4))
    fused2_batch(const Batch2Member<T> *__restrict__ mem, const int *__restrict__ tl,
                 const int *__restrict__ tf, int nmem, int ntl, int ntf,
                 const int64_t *__restrict__ dstep, int g) {
  __shared__ Smem2<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  const bool lean = (int)blockIdx.x < ntl;
  // each XCD takes a contiguous run of the part's tiles (xcd_tile over the
  // whole part), so neighbouring tiles of a member share that XCD's L2
  const int t = lean ? xcd_tile((int)blockIdx.x, ntl) : xcd_tile((int)blockIdx.x - ntl, ntf);
  const int *tstart = lean ? tl : tf;
  if (t < (lean ? ntl : ntf)) {
    const int64_t s = *dstep + 2 * g;  // first step of the pair (even)
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const Batch2Member<T> &m = mem[lo];
    Fused2Args<T> f = m.f[(s >> 1) & 1];
    f.a.step = m.step0 + s;
    if (m.btab) {
      f.beam1 = m.btab + (int64_t)2 * g * f.a.nz;
      f.beam2 = f.beam1 + f.a.nz;
    }
    const int b = t - tstart[lo];
    if (lean) fused2_tile_lean2b<T, TX, TY, false>(f, b, sm, smt);
    else fused2_tile<T, TX, TY, false, false>(f, f.nremap + b, sm, smt);
  }
}

// Beam rows of a launch group of G steps: member m, step q (0 .. G-1), plane k.
template <typename T>
__global__ void batch2_beam_table(const Batch2Member<T> *__restrict__ mem, int nmem,
                                  const int64_t *__restrict__ dstep, int G, T *const *btabs) {
  const int m = blockIdx.y;
  if (m >= nmem || !btabs[m]) return;
  const StepArgs<T> &a0 = mem[m].f[0].a;
  const int64_t cnt = (int64_t)G * a0.nz;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(t / a0.nz), k = (int)(t % a0.nz);
    StepArgs<T> a = a0;
    a.step = mem[m].step0 + *dstep + q;
    btabs[m][t] = beam_gauss(a, k);
  }
}

// Modelled time of a batched two-step launch with tiles TX x TY and chunks
// of kc planes: rounds of resident blocks (an eighth per XCD) x (kc + 8) x
// the block's waves (a block's set-up and pipeline fill cost about 8 planes:
// on 64 cavities of 120-320 planes, chunks of 64 and 96 planes ran 3 %
// faster than chunks of 32 and 48, and 128, which leaves the shorter
// members without a lean chunk, 13 % slower).
inline double batch2_cost(const int *nx, const int *ny, const int *nz, int n, int tx, int ty,
                          int kc, int slots) {
  long long blocks = 0;
  for (int m = 0; m < n; ++m) {
    const long long ncol = (long long)((nx[m] + tx - 1) / tx) * ((ny[m] + ty - 1) / ty);
    const int mid = nz[m] - 8;
    blocks += ncol * (mid >= kc ? (mid + kc - 1) / kc : (nz[m] + kc - 1) / kc);
  }
  const int xslots = slots / 8 > 0 ? slots / 8 : 1;
  const long long rounds = ((blocks + 7) / 8 + xslots - 1) / xslots;
  const int waves = ((tx + 4) * (ty + 4) + 63) / 64;
  return (double)rounds * (kc + 8) * waves;
}

// Planes per chunk for a tile shape: the least batch2_cost.
inline int pick_kc2_batch(const int *nx, const int *ny, const int *nz, int n, int tx, int ty,
                          int slots, double *cost = nullptr) {
  static const int cand[] = {32, 48, 64, 96, 128, 192, 256, 384};
  int nzmax = 0;
  for (int m = 0; m < n; ++m) nzmax = nz[m] > nzmax ? nz[m] : nzmax;
  int best = 64;
  double bt = 1e300;
  for (int kc : cand) {
    if (kc > nzmax && kc != cand[0]) break;
    const double t = batch2_cost(nx, ny, nz, n, tx, ty, kc, slots);
    if (t < bt) {
      bt = t;
      best = kc;
    }
  }
  if (cost) *cost = bt;
  return best;
}

// Chunking of one member of a batch: kc planes per chunk, 4-plane face
// chunks when a lean chunk fits (as launch_fused2 chunks a single grid).
template <typename T, int TX, int TY>
inline void chunk2_layout(Fused2Args<T> &f, int kc) {
  const StepArgs<T> &a = f.a;
  f.ntx = (a.nx + TX - 1) / TX;
  f.nty = (a.ny + TY - 1) / TY;
  const int ncol = f.ntx * f.nty;
  f.kc = kc;
  constexpr int KH = 4;
  if (a.nz - 2 * KH >= kc) {
    f.kh = KH;
    f.nmid = (a.nz - 2 * KH + kc - 1) / kc;
    f.nchunk = f.nmid + 2;
    f.nremap = ncol * f.nmid;
  } else {
    f.kh = 0;
    f.nmid = 0;
    f.nchunk = (a.nz + kc - 1) / kc;
    f.nremap = 0;
  }
  f.nb = ncol * f.nchunk;
}

// Single-grid tile shapes of the two-step sweep: 64 x 8 columns (a 68 x 12
// window, 816 threads), 32 x 16 (36 x 20, 720), 32 x 8 (36 x 12, 432),
// 16 x 16 (20 x 20, 400), 60 x 12 (64 x 16, 1024) or, in FP64, 28 x 12
// (32 x 16, 512).  64 x 8 has the longest
// contiguous rows and, with 60 x 12, the least time per window cell; 60 x 12
// has the fewest window cells per column (one 64-cell window row per wave,
// 16 waves, four per SIMD) but overhangs grids whose width is not near a
// multiple of 60; 32-wide shapes leave less overhang on grids whose width is
// not a multiple of 64, and the two small ones give more blocks (and two
// resident per CU) on small grids.  (In FP32, 24 x 24, 28 x 12 and 60 x 8
// were slower than one of these on most grids measured, 48^2 to 512^2
// cells.)  FP64 builds the shapes of at most 512 window cells: 32 x 8,
// 16 x 16 and 28 x 12 (32 x 16 window, 512 threads), 28 x 12 having the
// fewest window cells per tile cell of them (1.52 against 1.69 for 32 x 8)
// and 32 x 8 the least overhang on widths that are multiples of 32.
enum {
  F2_CFG_64x8 = 0,
  F2_CFG_32x16 = 1,
  F2_CFG_32x8 = 2,
  F2_CFG_16x16 = 3,
  F2_CFG_60x12 = 4,
  F2_CFG_28x12 = 5,
  F2_NCFG = 6
};

inline void f2_dims(int cfg, int &tx, int &ty) {
  static const int d[F2_NCFG][2] = {{64, 8}, {32, 16}, {32, 8}, {16, 16}, {60, 12}, {28, 12}};
  tx = d[cfg][0];
  ty = d[cfg][1];
}

// Whether the two-step sweep with tiles TX x TY serves these step
// arguments: FP32, nx >= TX and ny >= max(16, TY) (a window cell's flat
// carry then stays within one row and one plane), fields of less than 2^32
// bytes, J zero or sigma*E (no J buffer), the kernels spill-free.
// Whether the two-step sweep addresses the fields with per-component
// resources: fields of 2^32 bytes or more, or forced (fit_set_chunk_rsrc).
// The shapes built with per-component resources: 64 x 8 in FP32, 32 x 8 and
// 28 x 12 in FP64 (each precision's fastest shapes on large grids).
template <typename T, int TX, int TY>
constexpr bool f2_pcr_shape() {
  return sizeof(T) == 4 ? (TX == 64 && TY == 8)
                        : (TX == 32 && TY == 8) || (TX == 28 && TY == 12);
}

template <typename T>
inline bool fused2_pcr(const StepArgs<T> &a) {
  return !a.buf32 || a.pcr;
}

template <typename T, int TX, int TY>
inline bool fused2_supported_t(const StepArgs<T> &a) {
  if constexpr (!cfg2_fits<T, TX, TY>()) {
    (void)a;
    return false;
  } else {
    if (a.jmode == J_BUFFER || a.nx < TX || a.ny < 16 || a.ny < TY || a.nz < 2) return false;
    if (!fused2_pcr(a))
      return kernel_spill_free((const void *)fused2_sweep<T, TX, TY, false>) &&
             kernel_spill_free((const void *)fused2_sweep<T, TX, TY, true>) &&
             kernel_spill_free((const void *)fused2_sweep_faces<T, TX, TY, false>) &&
             kernel_spill_free((const void *)fused2_sweep_faces<T, TX, TY, true>);
    // per-component resources: one shape per precision (the one large grids
    // take), when the planes of the longest chunk (384) and its halo, the
    // material map and the metric tables fit 32-bit offsets
    if constexpr (f2_pcr_shape<T, TX, TY>()) {
      constexpr int64_t LIM = 0xFFFFFFFFll;
      const int64_t pbytes = (int64_t)a.nx * a.ny * (int64_t)sizeof(T);
      return a.geo_bytes < LIM && a.N < LIM && (int64_t)(384 + 16) * pbytes < LIM &&
             kernel_spill_free((const void *)fused2_sweep<T, TX, TY, false, true>) &&
             kernel_spill_free((const void *)fused2_sweep<T, TX, TY, true, true>) &&
             kernel_spill_free((const void *)fused2_sweep_faces<T, TX, TY, false, true>) &&
             kernel_spill_free((const void *)fused2_sweep_faces<T, TX, TY, true, true>);
    } else {
      return false;
    }
  }
}

template <typename T>
inline bool fused2_supported(const StepArgs<T> &a, int cfg) {
  switch (cfg) {
    case F2_CFG_64x8: return fused2_supported_t<T, 64, 8>(a);
    case F2_CFG_32x16: return fused2_supported_t<T, 32, 16>(a);
    case F2_CFG_32x8: return fused2_supported_t<T, 32, 8>(a);
    case F2_CFG_16x16: return fused2_supported_t<T, 16, 16>(a);
    case F2_CFG_60x12: return fused2_supported_t<T, 60, 12>(a);
    case F2_CFG_28x12: return fused2_supported_t<T, 28, 12>(a);
    default: return false;
  }
}

// Whether any tile shape of the two-step sweep serves these arguments.
template <typename T>
inline bool fused2_supported(const StepArgs<T> &a) {
  for (int c = 0; c < F2_NCFG; ++c)
    if (fused2_supported<T>(a, c)) return true;
  return false;
}

// Whether a batch member's step arguments suit the batched two-step sweep
// with tiles TX x TY (the conditions of fused2_supported otherwise).
template <typename T, int TX, int TY>
inline bool fused2_batch_supported_t(const StepArgs<T> &a) {
  if constexpr (!cfg2_fits<T, TX, TY>()) {
    (void)a;
    return false;
  } else {
    return a.jmode != J_BUFFER && a.buf32 && a.nx >= TX && a.ny >= 16 && a.ny >= TY &&
           a.nz >= 2 && kernel_spill_free((const void *)fused2_batch<T, TX, TY>);
  }
}

template <typename T>
inline bool fused2_batch_supported(const StepArgs<T> &a, int cfg) {
  switch (cfg) {
    case F2B_CFG_32x16: return fused2_batch_supported_t<T, 32, 16>(a);
    case F2B_CFG_24x24: return fused2_batch_supported_t<T, 24, 24>(a);
    case F2B_CFG_28x12: return fused2_batch_supported_t<T, 28, 12>(a);
    default: return false;
  }
}

// Launches two steps with tiles TX x TY; false (nothing launched) when not
// supported.
template <typename T, int TX, int TY>
inline bool launch_fused2_t(const StepArgs<T> &a, T *En, T *Hn, const T *beam1, const T *beam2,
                            bool js1, bool js2, hipStream_t s) {
  if constexpr (!cfg2_fits<T, TX, TY>()) {
    (void)a; (void)En; (void)Hn; (void)beam1; (void)beam2; (void)js1; (void)js2; (void)s;
    return false;
  } else {
    if (!fused2_supported_t<T, TX, TY>(a)) return false;
    constexpr int NT = (TX + 4) * (TY + 4);
    Fused2Args<T> f;
    f.a = a;
    f.En = En;
    f.Hn = Hn;
    f.beam1 = beam1;
    f.beam2 = beam2;
    f.js1 = js1 ? T(1) : T(0);
    f.js2 = js2 ? T(1) : T(0);
    f.ntx = (a.nx + TX - 1) / TX;
    f.nty = (a.ny + TY - 1) / TY;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t lds = fused_mt_bytes<T>(a.nmat);
    const int slots = resident_slots(dev, (const void *)fused2_sweep<T, TX, TY, false>, NT, lds);
    const int ncol = f.ntx * f.nty;
    f.kc = pick_kc2(ncol, f2_span(a.nz), slots);
    // 4-plane chunks at the z faces take the general path (fused2_sweep_faces),
    // the chunks between (all lean: every plane their outputs depend on lies
    // in 1 .. nz-2) the two-barrier lean path (fused2_sweep); a grid too short
    // for a lean chunk takes the general path throughout
    constexpr int KH = 4;
    if (a.nz - 2 * KH >= f.kc) {
      f.kh = KH;
      f.nmid = (a.nz - 2 * KH + f.kc - 1) / f.kc;
      f.nchunk = f.nmid + 2;
      f.nremap = ncol * f.nmid;
    } else {
      f.kh = 0;
      f.nmid = 0;
      f.nchunk = (a.nz + f.kc - 1) / f.kc;
      f.nremap = 0;
    }
    f.nb = ncol * f.nchunk;
    const int nfaces = f.nb - f.nremap;
    if constexpr (f2_pcr_shape<T, TX, TY>()) {
      if (fused2_pcr(a)) {
        if (f.nremap > 0) {
          if (a.nt)
            hipLaunchKernelGGL((fused2_sweep<T, TX, TY, true, true>), dim3(f.nremap), dim3(NT),
                               lds, s, f);
          else
            hipLaunchKernelGGL((fused2_sweep<T, TX, TY, false, true>), dim3(f.nremap), dim3(NT),
                               lds, s, f);
        }
        if (a.nt)
          hipLaunchKernelGGL((fused2_sweep_faces<T, TX, TY, true, true>), dim3(nfaces), dim3(NT),
                             lds, s, f);
        else
          hipLaunchKernelGGL((fused2_sweep_faces<T, TX, TY, false, true>), dim3(nfaces),
                             dim3(NT), lds, s, f);
        return true;
      }
    }
    if (kernel_spill_free((const void *)fused2_sweep_all<T, TX, TY, true>) &&
        kernel_spill_free((const void *)fused2_sweep_all<T, TX, TY, false>)) {
      if (a.nt)
        hipLaunchKernelGGL((fused2_sweep_all<T, TX, TY, true>), dim3(f.nb), dim3(NT), lds, s, f);
      else
        hipLaunchKernelGGL((fused2_sweep_all<T, TX, TY, false>), dim3(f.nb), dim3(NT), lds, s, f);
      return true;
    }
    if (f.nremap > 0) {
      if (a.nt)
        hipLaunchKernelGGL((fused2_sweep<T, TX, TY, true>), dim3(f.nremap), dim3(NT), lds, s, f);
      else
        hipLaunchKernelGGL((fused2_sweep<T, TX, TY, false>), dim3(f.nremap), dim3(NT), lds, s, f);
    }
    if (a.nt)
      hipLaunchKernelGGL((fused2_sweep_faces<T, TX, TY, true>), dim3(nfaces), dim3(NT), lds, s, f);
    else
      hipLaunchKernelGGL((fused2_sweep_faces<T, TX, TY, false>), dim3(nfaces), dim3(NT), lds, s,
                         f);
    return true;
  }
}

template <typename T>
inline bool launch_fused2(const StepArgs<T> &a, int cfg, T *En, T *Hn, const T *beam1,
                          const T *beam2, bool js1, bool js2, hipStream_t s) {
  switch (cfg) {
    case F2_CFG_64x8: return launch_fused2_t<T, 64, 8>(a, En, Hn, beam1, beam2, js1, js2, s);
    case F2_CFG_32x16: return launch_fused2_t<T, 32, 16>(a, En, Hn, beam1, beam2, js1, js2, s);
    case F2_CFG_32x8: return launch_fused2_t<T, 32, 8>(a, En, Hn, beam1, beam2, js1, js2, s);
    case F2_CFG_16x16: return launch_fused2_t<T, 16, 16>(a, En, Hn, beam1, beam2, js1, js2, s);
    case F2_CFG_60x12: return launch_fused2_t<T, 60, 12>(a, En, Hn, beam1, beam2, js1, js2, s);
    case F2_CFG_28x12: return launch_fused2_t<T, 28, 12>(a, En, Hn, beam1, beam2, js1, js2, s);
    default: return false;
  }
}

// Modelled time of a two-step launch with a shape: the blocks (columns x
// chunks of the kc that pick_kc2 takes) run in rounds of the resident
// blocks, a round lasting kc + 3 planes of one block, and a block's plane
// time is its window cells x the shape's measured time per window cell
// relative to 64 x 8 (launches on a 256 x 256 x 300 grid, no overhang:
// 64 x 8 0.378 ms, 32 x 16 0.390, 32 x 8 0.432, 16 x 16 0.472; 60 x 12
// from grids it tiles well; FP64, relative to 32 x 8, fitted over nine
// grids from 48 x 48 x 300 to 256 x 256 x 768: 16 x 16 1.105, 28 x 12
// 0.948) x the blocks that share a CU in a round (the
// resident blocks per CU, or fewer when a round holds fewer blocks than
// there are CUs).  On
// large grids this is the window cells over the covered (overhanging)
// columns; on small ones the rounds and chunk lengths decide.
template <typename T, int TX, int TY>
inline double fused2_time_t(const StepArgs<T> &a, int dev, double rel) {
  if constexpr (!cfg2_fits<T, TX, TY>()) {
    (void)a; (void)dev; (void)rel;
    return 1e300;
  } else {
  constexpr int NT = (TX + 4) * (TY + 4);
  int ncu = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (ncu <= 0) ncu = 256;
  const int slots = resident_slots(dev, (const void *)fused2_sweep<T, TX, TY, false>, NT,
                                   fused_mt_bytes<T>(a.nmat));
  const int bpc = slots / ncu > 0 ? slots / ncu : 1;
  const int ncol = ((a.nx + TX - 1) / TX) * ((a.ny + TY - 1) / TY);
  const int span = f2_span(a.nz);
  const int kc = pick_kc2(ncol, span, slots);
  const long long blocks = (long long)ncol * ((span + kc - 1) / kc);
  const int xslots = slots / 8 > 0 ? slots / 8 : 1;
  const long long rounds = ((blocks + 7) / 8 + xslots - 1) / xslots;
  const long long per_round = blocks < slots ? blocks : slots;
  const long long share = std::min<long long>(bpc, (per_round + ncu - 1) / ncu);
  return (double)rounds * (kc + 3) * NT * rel * (double)share;
  }
}

template <typename T>
inline double fused2_time(const StepArgs<T> &a, int cfg, int dev) {
  constexpr bool fp64 = sizeof(T) == 8;
  switch (cfg) {
    case F2_CFG_64x8: return fused2_time_t<T, 64, 8>(a, dev, 1.0);
    case F2_CFG_32x16: return fused2_time_t<T, 32, 16>(a, dev, 1.17);
    case F2_CFG_32x8: return fused2_time_t<T, 32, 8>(a, dev, fp64 ? 1.0 : 1.08);
    case F2_CFG_16x16: return fused2_time_t<T, 16, 16>(a, dev, fp64 ? 1.105 : 1.27);
    case F2_CFG_60x12: return fused2_time_t<T, 60, 12>(a, dev, 0.97);
    case F2_CFG_28x12: return fused2_time_t<T, 28, 12>(a, dev, 0.948);
    default: return 1e300;
  }
}

// The tile shape used when none is set: the least fused2_time among the
// shapes that apply (-1: none).
template <typename T>
inline int fused2_default_cfg(const StepArgs<T> &a) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  int best = -1;
  double bc = 1e300;
  for (int c = 0; c < F2_NCFG; ++c) {
    if (!fused2_supported<T>(a, c)) continue;
    const double t = fused2_time<T>(a, c, dev);
    if (t < bc) {
      bc = t;
      best = c;
    }
  }
  return best;
}

}  // namespace fit
