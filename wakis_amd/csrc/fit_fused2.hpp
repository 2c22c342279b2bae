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
// Each thread owns one tile cell (window (2 + t % TX, 2 + t / TX)); threads
// t < NR also own one ring cell of the window's outer two rows/columns.
// TX = 64: a tile row is one wave, so row-indexed metrics are scalar loads.
// LDS: four 2-plane rings (u0, w1, u1, w2) of 3 x (TX+4)(TY+4) values; the
// material table follows in dynamic LDS.
#pragma once

#include <type_traits>

#include "fit_fused.hpp"

namespace fit {

template <int TX, int TY>
struct Win2 {
  static constexpr int WW = TX + 4;           // window row length
  static constexpr int WH = TY + 4;           // window rows
  static constexpr int WC = WW * WH;          // window cells per plane
  static constexpr int NT = TX * TY;          // threads (one tile cell each)
  static constexpr int NR = 4 * WW + 4 * TY;  // ring cells (WC - NT)
};

template <typename T, int TX, int TY>
struct Smem2 {
  using W = Win2<TX, TY>;
  T u0[2][3][W::WC];  // u of E_old, planes k and k+1
  T w1[2][3][W::WC];  // w of H1, planes k and k-1
  T u1[2][3][W::WC];  // u of E1, planes k and k-1
  T w2[2][3][W::WC];  // w of H2, planes k-1 and k-2
};

template <typename T>
struct Fused2Args {
  StepArgs<T> a;           // a.E, a.H: buffers read; a.step: index of the first step
  T *En, *Hn;              // buffers written
  const T *beam1, *beam2;  // J_z line (nz values) of step 1 / step 2, or null
  T js1, js2;              // 1: J = sigma * (the step's input E); 0: J = 0
  int ntx, nty, kc, nchunk, nb, nremap;
  int kh = 0, nmid = 0;  // z face chunks of kh planes (chunk2_span)
};

// Window coordinates of ring cell r (rows 0, 1, TY+2, TY+3, then columns
// 0, 1, TX+2, TX+3 of rows 2 .. TY+1).
template <int TX, int TY>
__device__ __forceinline__ void ring_cell(int r, int &wi, int &wj) {
  constexpr int WW = TX + 4;
  if (r < 4 * WW) {
    const int rr = r / WW;
    wi = r - rr * WW;
    wj = rr < 2 ? rr : TY + rr;
  } else {
    const int u = r - 4 * WW;
    const int cc = u / TY;
    wj = 2 + (u - cc * TY);
    wi = cc < 2 ? cc : TX + cc;
  }
}

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

// One tile of a chunk.  LEAN: every plane the chunk touches (its window
// cells' true planes included) lies inside the grid and away from the z
// faces, so validity is constant and the boundary masks reduce to per-cell
// 0/1 factors fixed over the march (the same expressions, with fewer
// instructions per plane).
template <typename T, int TX, int TY, bool NTS, bool LEAN>
__device__ __forceinline__ void fused2_tile(const Fused2Args<T> &f, const int tile,
                                            Smem2<T, TX, TY> &sm, T *__restrict__ smt) {
  using W = Win2<TX, TY>;
  static_assert(TX == 64, "a wave covers one tile row");
  constexpr int WW = W::WW, WC = W::WC, NT = W::NT, NR = W::NR;
  static_assert(NR <= NT, "every ring cell needs a thread");
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
  const CT<T> *gLz = (const CT<T> *)g.Lz;
  const CT<T> *gtLz = (const CT<T> *)g.tLz;
  const CT<T> *giAx = (const CT<T> *)g.iAx;
  const CT<T> *gitAx = (const CT<T> *)g.itAx;
  const CT<T> *mtab = (const CT<T> *)a.mtab;
  const uint32_t pec = a.pec, pmc = a.pmc;
  // beam lines: the own cell's plane index is wave-uniform (scalar loads);
  // the ring cell's value travels in its update packet.  A null line reads
  // the L_z table instead (valid memory, value unused).
  const bool b1on = f.beam1 != nullptr, b2on = f.beam2 != nullptr;
  const CT<T> *cb1 = (const CT<T> *)(b1on ? f.beam1 : g.Lz);
  const CT<T> *cb2 = (const CT<T> *)(b2on ? f.beam2 : g.Lz);
  const T *gb1 = b1on ? f.beam1 : g.Lz;

  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += NT) smt[t] = mtab[t];

  auto clampk = [&](int kk) __attribute__((always_inline)) { return min(max(kk, 0), nz - 1); };

  // ---- own cell (tile cell of this thread; never wraps: full tiles) ----
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  const int oi = x0 + lx;
  const int oj = __builtin_amdgcn_readfirstlane(y0 + ly);  // wave-uniform
  const int ol = (ly + 2) * WW + (lx + 2);
  const int ocol = oi + nx * oj;
  bool opec[3], opmc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    opec[d] = xy_mask(pec, d, oi, oj, nx, ny);
    opmc[d] = xy_mask(pmc, d, oi, oj, nx, ny);
  }
  const T oLx = g.Lx[oi], oLy = g.Ly[oj], otLx = g.tLx[oi], otLy = g.tLy[oj];
  const T oiAz = g.iAz[ocol], oitAz = g.itAz[ocol];
  const bool obeam = a.beam_mode != BEAM_OFF && oi == a.ixs && oj == a.iys;
  // wake probe window of this column (or -1)
  int pw = -1;
  if (a.hist) {
    const int di = oi - a.ixt, dj = oj - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw)
      pw = (di + a.hw) * (2 * a.hw + 1) + (dj + a.hw);
  }

  // ---- ring cell of this thread (threads < NR; waves 0 .. NR/64) ----
  const bool ring = (int)threadIdx.x < NR;
  int rwi = 0, rwj = 0;
  if (ring) ring_cell<TX, TY>(threadIdx.x, rwi, rwj);
  const WinCell rc = win_cell(x0 - 2 + rwi, y0 - 2 + rwj, nx, ny);
  const int rl = rwj * WW + rwi;
  const int rcol = rc.i + nx * rc.j;
  const int rdk = rc.dk;
  const bool rh1 = ring && rwi <= TX + 2 && rwj <= TY + 2;
  const bool re1 = rh1 && rwi >= 1 && rwj >= 1;
  const bool rh2 = ring && rwi >= 1 && rwj >= 1 && rwi <= TX + 1 && rwj <= TY + 1;
  bool rpec[3], rpmc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    rpec[d] = xy_mask(pec, d, rc.i, rc.j, nx, ny);
    rpmc[d] = xy_mask(pmc, d, rc.i, rc.j, nx, ny);
  }
  const T rLx = g.Lx[rc.i], rLy = g.Ly[rc.j], rtLx = g.tLx[rc.i], rtLy = g.tLy[rc.j];
  const T riAz = g.iAz[rcol], ritAz = g.itAz[rcol];
  const bool rbeam = a.beam_mode != BEAM_OFF && rc.i == a.ixs && rc.j == a.iys;

  auto fac = [](bool masked) __attribute__((always_inline)) { return masked ? T(0) : T(1); };
  // boundary-mask factors of the own (o) and ring (r) cell at true plane k,
  // pec (column mask, E) and pmc (row mask, H); validity of a true plane
  const T opeF[3] = {fac(opec[0]), fac(opec[1]), fac(opec[2])};
  const T opmF[3] = {fac(opmc[0]), fac(opmc[1]), fac(opmc[2])};
  const T rpeF[3] = {fac(rpec[0]), fac(rpec[1]), fac(rpec[2])};
  const T rpmF[3] = {fac(rpmc[0]), fac(rpmc[1]), fac(rpmc[2])};
  auto ope = [&](int c, int k) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return opeF[c];
    else return fac(full_mask(pec, c, opec[c], k, nz));
  };
  auto opm = [&](int c, int k) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return opmF[c];
    else return fac(full_mask(pmc, c, opmc[c], k, nz));
  };
  auto rpe = [&](int c, int k) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return rpeF[c];
    else return fac(full_mask(pec, c, rpec[c], k, nz));
  };
  auto rpm = [&](int c, int k) __attribute__((always_inline)) -> T {
    if constexpr (LEAN) return rpmF[c];
    else return fac(full_mask(pmc, c, rpmc[c], k, nz));
  };
  auto valid = [&](int k) __attribute__((always_inline)) -> bool {
    if constexpr (LEAN) return true;
    else return k >= 0 && k < nz;
  };

  // ---- packets ----
  // staging data of one plane: E_old of the own and the ring cell, L_z of the
  // ring cell (the own cell's is a scalar)
  struct SPk {
    T oe[3], re[3], rlz;
  };
  // update data of one plane: H_old, material, the per-lane areas
  struct CPk {
    T oh[3], oiay, oitay;
    int om;
    T rh[3], riax, riay, ritax, ritay, rtlz, rb1;
    int rm;
  };
  // Field, material and metric loads are CDNA buffer instructions: one
  // resource per array (the fields span < 2^32 bytes: fused2_supported), a
  // per-lane 32-bit byte offset and the component offset c N sizeof(T) as the
  // uniform one.  Planes are clamped to the grid, so every access is in
  // range; values of cells outside it are never used (their u and w are 0).
  // (Packet fields are assigned once, after a branch: stores of different
  // fields in the two arms would be merged through memory, which keeps the
  // packet in scratch and makes every load wait.)
  constexpr uint32_t SZ = sizeof(T);
  const uint32_t NB = (uint32_t)a.N * SZ, NB2 = 2 * NB;
  const uint32_t PB = (uint32_t)plane * SZ;
  const int64_t FB = 3 * a.N * (int64_t)SZ;
  const Rsrc rE = make_rsrc(a.E, FB), rH = make_rsrc(a.H, FB);
  const Rsrc rEn = make_rsrc(f.En, FB), rHn = make_rsrc(f.Hn, FB);
  const bool hasmat = a.mat != nullptr;
  const Rsrc rM = make_rsrc(hasmat ? (const void *)a.mat : (const void *)g.Lx, hasmat ? a.N : 1);
  const Rsrc rG = make_rsrc(g.Lx, a.geo_bytes);
  auto gofs = [&](const T *t) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)((const char *)t - (const char *)g.Lx);
  };
  const uint32_t oLzT = gofs(g.Lz), otLzT = gofs(g.tLz), oiAxT = gofs(g.iAx), oiAyT = gofs(g.iAy);
  const uint32_t oitAxT = gofs(g.itAx), oitAyT = gofs(g.itAy);
  const uint32_t ocb = (uint32_t)ocol * SZ;
  const uint32_t rcb = (uint32_t)rcol * SZ;
  const uint32_t orow = (uint32_t)nx * SZ, rrow = (uint32_t)ny * SZ;  // (i,k) / (j,k) tables
  // ring beam line through the geometry resource when on (it is another
  // allocation: read it with its own resource)
  const Rsrc rB1 = make_rsrc(gb1, (int64_t)nz * SZ);

  auto load_S = [&](int k, SPk &p) __attribute__((always_inline)) {
    const uint32_t so = (uint32_t)clampk(k) * PB;  // wave-uniform
    p.oe[0] = bload<T>(rE, ocb, so);
    p.oe[1] = bload<T>(rE, ocb, so + NB);
    p.oe[2] = bload<T>(rE, ocb, so + NB2);
    T e0 = T(0), e1 = T(0), e2 = T(0), lz = T(0);
    if (ring) {
      const int kr = clampk(k + rdk);
      const uint32_t v = rcb + (uint32_t)kr * PB;
      e0 = bload<T>(rE, v, 0);
      e1 = bload<T>(rE, v, NB);
      e2 = bload<T>(rE, v, NB2);
      lz = bload<T>(rG, oLzT + (uint32_t)kr * SZ, 0);
    }
    p.re[0] = e0;
    p.re[1] = e1;
    p.re[2] = e2;
    p.rlz = lz;
  };
  auto load_C = [&](int k, CPk &p) __attribute__((always_inline)) {
    const int ko = clampk(k);
    const uint32_t so = (uint32_t)ko * PB;
    p.oh[0] = bload<T>(rH, ocb, so);
    p.oh[1] = bload<T>(rH, ocb, so + NB);
    p.oh[2] = bload<T>(rH, ocb, so + NB2);
    p.om = hasmat ? bload_u8(rM, (uint32_t)ocol, (uint32_t)ko * (uint32_t)plane) : 0;
    p.oiay = bload<T>(rG, oiAyT + (uint32_t)oi * SZ, (uint32_t)ko * orow);
    p.oitay = bload<T>(rG, oitAyT + (uint32_t)oi * SZ, (uint32_t)ko * orow);
    T h0 = T(0), h1 = T(0), h2 = T(0), iax = T(0), iay = T(0), itax = T(0), itay = T(0),
      tlz = T(0), b1 = T(0);
    int m = 0;
    if (ring) {
      const int kr = clampk(k + rdk);
      const uint32_t v = rcb + (uint32_t)kr * PB;
      h0 = bload<T>(rH, v, 0);
      h1 = bload<T>(rH, v, NB);
      h2 = bload<T>(rH, v, NB2);
      m = hasmat ? bload_u8(rM, (uint32_t)rcol + (uint32_t)kr * (uint32_t)plane, 0) : 0;
      const uint32_t jk = (uint32_t)(rc.j + ny * kr) * SZ, ik = (uint32_t)(rc.i + nx * kr) * SZ;
      iax = bload<T>(rG, oiAxT + jk, 0);
      iay = bload<T>(rG, oiAyT + ik, 0);
      itax = bload<T>(rG, oitAxT + jk, 0);
      itay = bload<T>(rG, oitAyT + ik, 0);
      tlz = bload<T>(rG, otLzT + (uint32_t)kr * SZ, 0);
      b1 = bload<T>(rB1, (uint32_t)kr * SZ, 0);
    }
    p.rb1 = b1;
    p.rh[0] = h0;
    p.rh[1] = h1;
    p.rh[2] = h2;
    p.rm = m;
    p.riax = iax;
    p.riay = iay;
    p.ritax = itax;
    p.ritay = itay;
    p.rtlz = tlz;
  };

  auto curl_u = [&](const T (*u)[WC], const T (*un)[WC], int d, int l)
      __attribute__((always_inline)) -> T {
    if (d == 0) return ((u[1][l] - un[1][l]) - u[2][l]) + u[2][l + WW];
    if (d == 1) return ((u[2][l] - u[2][l + 1]) - u[0][l]) + un[0][l];
    return ((u[0][l] - u[0][l + WW]) - u[1][l]) + u[1][l + 1];
  };
  auto curl_w = [&](const T (*w)[WC], const T (*wm)[WC], int d, int l)
      __attribute__((always_inline)) -> T {
    if (d == 0) return ((w[2][l] - w[2][l - WW]) - w[1][l]) + wm[1][l];
    if (d == 1) return ((w[0][l] - wm[0][l]) - w[2][l]) + w[2][l - 1];
    return ((w[1][l] - w[1][l - 1]) - w[0][l]) + w[0][l - WW];
  };

  // stage u0 of window plane k (own and ring cell) from the packet
  auto stage_u0 = [&](int k, int slot, const SPk &p) __attribute__((always_inline)) {
    T(*u)[WC] = sm.u0[slot];
    {
      const bool v = valid(k);  // wave-uniform
      const T lz = gLz[clampk(k)];
      u[0][ol] = v ? (oLx * ope(0, k)) * p.oe[0] : T(0);
      u[1][ol] = v ? (oLy * ope(1, k)) * p.oe[1] : T(0);
      u[2][ol] = v ? ope(2, k) * (lz * p.oe[2]) : T(0);
    }
    if (ring) {
      const int kr = k + rdk;
      const bool v = valid(kr);
      u[0][rl] = v ? (rLx * rpe(0, kr)) * p.re[0] : T(0);
      u[1][rl] = v ? (rLy * rpe(1, kr)) * p.re[1] : T(0);
      u[2][rl] = v ? rpe(2, kr) * (p.rlz * p.re[2]) : T(0);
    }
  };

  // ---- pipeline state ----
  SPk S;          // E_old of the plane staged next (in flight)
  T oE0c[3], oE0n[3], rE0c[3], rE0n[3];  // raw E_old of planes k / k+1
  T rLzc = 0, rLzn = 0;                  // ring L_z of planes k / k+1
  T oH1p[3], rH1p[3];                    // raw H1 of plane k-1
  T oE1p[3];                             // raw E1 of plane k-1 (own)
  int omp = 0, rmp = 0;                  // materials of plane k-1
  T oiayp = 0, oitayp = 0, riaxp = 0, riayp = 0, rtlzp = 0;  // areas of plane k-1
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    oE0c[d] = oE0n[d] = rE0c[d] = rE0n[d] = T(0);
    oH1p[d] = rH1p[d] = oE1p[d] = T(0);
  }

  // prologue: u0 of plane k0-2 staged; S = E_old of plane k0-1; C of k0-2
  const int kb = k0 - 2;
  load_S(kb, S);
  stage_u0(kb, 0, S);
  load_S(kb + 1, S);
  CPk CA, CB;
  load_C(kb, CA);

  const bool probe1 = pw >= 0 && a.step < a.hist_nt;
  const bool probe2 = pw >= 0 && a.step + 1 < a.hist_nt;
  GT<T> *phist = pw >= 0 ? (GT<T> *)a.hist + (int64_t)pw * nz * a.hist_nt + a.step : nullptr;

  // One plane: stage u0(k+1); H1(k); E1(k); H2(k-1); E2(k-1).  Cc: update
  // data of plane k (loaded one iteration earlier); Cn receives plane k+1's.
  // PH: phases run (compile time): 1 = E1, 2 = H2, 4 = E2 (staging and H1
  // always); the three pipeline-fill iterations before the loop run fewer
  // PA: LDS ring slot of plane k, (k - kb) & 1, a compile-time constant in
  // every unrolled copy (so LDS addresses are immediate offsets)
  auto iter = [&](const int k, const CPk &Cc, CPk &Cn, auto ph, auto pa)
      __attribute__((always_inline)) {
    constexpr int PH = decltype(ph)::value;
    constexpr int PA = decltype(pa)::value, PB1 = PA ^ 1;
    // A. stage u0(k+1); keep its raw E for E1(k+1)
    stage_u0(k + 1, PB1, S);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      oE0n[d] = S.oe[d];
      rE0n[d] = S.re[d];
    }
    rLzn = S.rlz;
    // B. loads of the next iteration (unconditional, at clamped planes: a
    // conditional load would make the compiler merge and copy registers of
    // loads in flight at the loop's back edge)
    load_S(k + 2, S);
    load_C(k + 1, Cn);
    __syncthreads();

    // C. H1(k) -> w1 slot PA
    T oH1n[3], rH1n[3];
    {
      const T(*u)[WC] = sm.u0[PA];
      const T(*un)[WC] = sm.u0[PB1];
      T(*w)[WC] = sm.w1[PA];
      {
        const bool v = valid(k);
        const int kc = clampk(k);
        const T *mt = smt + Cc.om * MAT_STRIDE;
        const T fm0 = opm(0, k), fm1 = opm(1, k), fm2 = opm(2, k);
        const T iax = giAx[oj + ny * kc];
        oH1n[0] = Cc.oh[0] - (mt[0] * (fm0 * iax)) * curl_u(u, un, 0, ol);
        oH1n[1] = Cc.oh[1] - (mt[1] * (fm1 * Cc.oiay)) * curl_u(u, un, 1, ol);
        oH1n[2] = Cc.oh[2] - (mt[2] * (fm2 * oiAz)) * curl_u(u, un, 2, ol);
        const T vf = v ? T(1) : T(0);
        w[0][ol] = (vf * (fm0 * otLx)) * oH1n[0];
        w[1][ol] = (vf * (fm1 * otLy)) * oH1n[1];
        w[2][ol] = (vf * (fm2 * gtLz[kc])) * oH1n[2];
      }
      if (rh1) {
        const int kr = k + rdk;
        const bool v = valid(kr);
        const T *mt = smt + Cc.rm * MAT_STRIDE;
        const T fm0 = rpm(0, kr), fm1 = rpm(1, kr), fm2 = rpm(2, kr);
        rH1n[0] = Cc.rh[0] - (mt[0] * (fm0 * Cc.riax)) * curl_u(u, un, 0, rl);
        rH1n[1] = Cc.rh[1] - (mt[1] * (fm1 * Cc.riay)) * curl_u(u, un, 1, rl);
        rH1n[2] = Cc.rh[2] - (mt[2] * (fm2 * riAz)) * curl_u(u, un, 2, rl);
        const T vf = v ? T(1) : T(0);
        w[0][rl] = (vf * (fm0 * rtLx)) * rH1n[0];
        w[1][rl] = (vf * (fm1 * rtLy)) * rH1n[1];
        w[2][rl] = (vf * (fm2 * Cc.rtlz)) * rH1n[2];
      } else {
        rH1n[0] = rH1n[1] = rH1n[2] = T(0);
      }
    }
    __syncthreads();

    // D. E1(k) -> u1 slot PA; probe both steps
    T oE1n[3] = {T(0), T(0), T(0)};
    if constexpr ((PH & 1) != 0) {
      const T(*w)[WC] = sm.w1[PA];
      const T(*wm)[WC] = sm.w1[PB1];
      T(*u)[WC] = sm.u1[PA];
      {
        const bool v = valid(k);
        const int kc = clampk(k);
        const T *mt = smt + Cc.om * MAT_STRIDE;
        const T ax = ope(0, k) * gitAx[oj + ny * kc];
        const T ay = ope(1, k) * Cc.oitay;
        const T az = ope(2, k) * oitAz;
        const T jx = f.js1 * (mt[6] * oE0c[0]);
        const T jy = f.js1 * (mt[7] * oE0c[1]);
        const T jz = (obeam && b1on) ? cb1[kc] : f.js1 * (mt[8] * oE0c[2]);
        oE1n[0] = oE0c[0] + mt[3] * (ax * curl_w(w, wm, 0, ol) - jx);
        oE1n[1] = oE0c[1] + mt[4] * (ay * curl_w(w, wm, 1, ol) - jy);
        oE1n[2] = oE0c[2] + mt[5] * (az * curl_w(w, wm, 2, ol) - jz);
        u[0][ol] = v ? (oLx * ope(0, k)) * oE1n[0] : T(0);
        u[1][ol] = v ? (oLy * ope(1, k)) * oE1n[1] : T(0);
        u[2][ol] = v ? ope(2, k) * (gLz[kc] * oE1n[2]) : T(0);
        if (k >= k0 && k < k1) {
          if (probe1) phist[(int64_t)k * a.hist_nt] = oE0c[2];
          if (probe2) phist[(int64_t)k * a.hist_nt + 1] = oE1n[2];
        }
      }
      if (re1) {
        const int kr = k + rdk;
        const bool v = valid(kr);
        const T *mt = smt + Cc.rm * MAT_STRIDE;
        const T ax = rpe(0, kr) * Cc.ritax;
        const T ay = rpe(1, kr) * Cc.ritay;
        const T az = rpe(2, kr) * ritAz;
        const T jx = f.js1 * (mt[6] * rE0c[0]);
        const T jy = f.js1 * (mt[7] * rE0c[1]);
        const T jz = (rbeam && b1on) ? Cc.rb1 : f.js1 * (mt[8] * rE0c[2]);
        const T e0 = rE0c[0] + mt[3] * (ax * curl_w(w, wm, 0, rl) - jx);
        const T e1 = rE0c[1] + mt[4] * (ay * curl_w(w, wm, 1, rl) - jy);
        const T e2 = rE0c[2] + mt[5] * (az * curl_w(w, wm, 2, rl) - jz);
        u[0][rl] = v ? (rLx * rpe(0, kr)) * e0 : T(0);
        u[1][rl] = v ? (rLy * rpe(1, kr)) * e1 : T(0);
        u[2][rl] = v ? rpe(2, kr) * (rLzc * e2) : T(0);
      }
    }
    __syncthreads();

    // E. H2(k-1) -> w2 slot PB1; store H2(k-1)
    if constexpr ((PH & 2) != 0) {
      const int km = k - 1;
      const T(*u)[WC] = sm.u1[PB1];
      const T(*un)[WC] = sm.u1[PA];
      T(*w)[WC] = sm.w2[PB1];
      {
        const bool v = valid(km);
        const int kc = clampk(km);
        const T *mt = smt + omp * MAT_STRIDE;
        const T fm0 = opm(0, km), fm1 = opm(1, km), fm2 = opm(2, km);
        const T iax = giAx[oj + ny * kc];
        const T h0 = oH1p[0] - (mt[0] * (fm0 * iax)) * curl_u(u, un, 0, ol);
        const T h1 = oH1p[1] - (mt[1] * (fm1 * oiayp)) * curl_u(u, un, 1, ol);
        const T h2 = oH1p[2] - (mt[2] * (fm2 * oiAz)) * curl_u(u, un, 2, ol);
        const T vf = v ? T(1) : T(0);
        w[0][ol] = (vf * (fm0 * otLx)) * h0;
        w[1][ol] = (vf * (fm1 * otLy)) * h1;
        w[2][ol] = (vf * (fm2 * gtLz[kc])) * h2;
        if (km >= k0 && km < k1) {
          const uint32_t so = (uint32_t)km * PB;
          bstore(rHn, ocb, so, h0, NTS);
          bstore(rHn, ocb, so + NB, h1, NTS);
          bstore(rHn, ocb, so + NB2, h2, NTS);
        }
      }
      if (rh2) {
        const int kr = km + rdk;
        const bool v = valid(kr);
        const T *mt = smt + rmp * MAT_STRIDE;
        const T fm0 = rpm(0, kr), fm1 = rpm(1, kr), fm2 = rpm(2, kr);
        const T h0 = rH1p[0] - (mt[0] * (fm0 * riaxp)) * curl_u(u, un, 0, rl);
        const T h1 = rH1p[1] - (mt[1] * (fm1 * riayp)) * curl_u(u, un, 1, rl);
        const T h2 = rH1p[2] - (mt[2] * (fm2 * riAz)) * curl_u(u, un, 2, rl);
        const T vf = v ? T(1) : T(0);
        w[0][rl] = (vf * (fm0 * rtLx)) * h0;
        w[1][rl] = (vf * (fm1 * rtLy)) * h1;
        w[2][rl] = (vf * (fm2 * rtlzp)) * h2;
      }
    }
    __syncthreads();

    // F. E2(k-1) of the tile; store it
    if constexpr ((PH & 4) != 0) {
      const int km = k - 1;
      const T(*w)[WC] = sm.w2[PB1];
      const T(*wm)[WC] = sm.w2[PA];  // plane k-2
      const int kc = clampk(km);
      const T *mt = smt + omp * MAT_STRIDE;
      const T ax = ope(0, km) * gitAx[oj + ny * kc];
      const T ay = ope(1, km) * oitayp;
      const T az = ope(2, km) * oitAz;
      const T jx = f.js2 * (mt[6] * oE1p[0]);
      const T jy = f.js2 * (mt[7] * oE1p[1]);
      const T jz = (obeam && b2on) ? cb2[kc] : f.js2 * (mt[8] * oE1p[2]);
      if (km < k1) {
        const uint32_t so = (uint32_t)km * PB;
        bstore(rEn, ocb, so, oE1p[0] + mt[3] * (ax * curl_w(w, wm, 0, ol) - jx), NTS);
        bstore(rEn, ocb, so + NB, oE1p[1] + mt[4] * (ay * curl_w(w, wm, 1, ol) - jy), NTS);
        bstore(rEn, ocb, so + NB2, oE1p[2] + mt[5] * (az * curl_w(w, wm, 2, ol) - jz), NTS);
      }
    }

    // rotate plane k into the "previous plane" state
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      oH1p[d] = oH1n[d];
      rH1p[d] = rH1n[d];
      oE1p[d] = oE1n[d];
      oE0c[d] = oE0n[d];
      rE0c[d] = rE0n[d];
    }
    rLzc = rLzn;
    omp = Cc.om;
    rmp = Cc.rm;
    oiayp = Cc.oiay;
    oitayp = Cc.oitay;
    riaxp = Cc.riax;
    riayp = Cc.riay;
    rtlzp = Cc.rtlz;
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

extern __shared__ double fit_dyn_lds[];

template <typename T, int TX, int TY, bool NTS>
__global__ void __launch_bounds__(TX *TY) fused2_sweep(Fused2Args<T> f) {
  __shared__ Smem2<T, TX, TY> sm;
  T *smt = (T *)fit_dyn_lds;
  const int b = blockIdx.x;
  const int tile = b < f.nremap ? xcd_tile(b, f.nremap) : b;
  int k0, k1;
  chunk2_span(f, tile / (f.ntx * f.nty), k0, k1);
  // the true planes that feed the chunk's outputs span k0-3 .. k1+2 (window
  // planes k0-2 .. k1+1, plane carry -1 .. 1); lean when all lie in
  // 1 .. nz-2 (no z face, no plane outside the grid)
  if (k0 >= 4 && k1 + 4 <= f.a.nz) fused2_tile<T, TX, TY, NTS, true>(f, tile, sm, smt);
  else fused2_tile<T, TX, TY, NTS, false>(f, tile, sm, smt);
}

// Tile of the two-step sweep: 64 x 8 columns, 512 threads (a 64 x 16 tile's
// 1024 threads leave 128 registers per lane, which this body exceeds).
// FP32 only: an FP64 window does not fit the LDS.
constexpr int F2_TX = 64, F2_TY = 8;

template <typename T, int TX, int TY>
constexpr bool cfg2_fits() {
  return sizeof(T) == 4 &&
         sizeof(Smem2<T, TX, TY>) + MAX_MATERIALS * MAT_STRIDE * sizeof(T) <= 160 * 1024;
}

// Planes per chunk: blocks run in rounds of `slots` (an eighth per XCD); a
// round lasts about a chunk plus its 3 planes of pipeline fill; the least
// modelled time rounds x (kc + 3) among a few lengths.
inline int pick_kc2(int ncol, int nz, int slots) {
  static const int cand[] = {32, 48, 64, 96, 128, 192, 256, 384};
  int best = 64;
  double bt = 1e300;
  const int xslots = slots / 8 > 0 ? slots / 8 : 1;
  for (int kc : cand) {
    if (kc > nz && kc != cand[0]) break;
    const long long blocks = (long long)ncol * ((nz + kc - 1) / kc);
    const long long per_xcd = (blocks + 7) / 8;
    const long long rounds = (per_xcd + xslots - 1) / xslots;
    const double t = (double)rounds * (kc + 3);
    if (t < bt) {
      bt = t;
      best = kc;
    }
  }
  return best;
}

// Whether the two-step sweep serves these step arguments: FP32, full tiles,
// fields of less than 2^32 bytes, J zero or sigma*E (no J buffer), the
// kernels spill-free.
template <typename T>
inline bool fused2_supported(const StepArgs<T> &a) {
  if constexpr (!cfg2_fits<T, F2_TX, F2_TY>()) {
    (void)a;
    return false;
  } else {
    return a.jmode != J_BUFFER && a.buf32 && a.nx % F2_TX == 0 && a.ny % F2_TY == 0 &&
           a.nz >= 2 && kernel_spill_free((const void *)fused2_sweep<T, F2_TX, F2_TY, false>) &&
           kernel_spill_free((const void *)fused2_sweep<T, F2_TX, F2_TY, true>);
  }
}

// Launches two steps; false (nothing launched) when not supported.
template <typename T>
inline bool launch_fused2(const StepArgs<T> &a, T *En, T *Hn, const T *beam1, const T *beam2,
                          bool js1, bool js2, hipStream_t s) {
  if constexpr (!cfg2_fits<T, F2_TX, F2_TY>()) {
    (void)a; (void)En; (void)Hn; (void)beam1; (void)beam2; (void)js1; (void)js2; (void)s;
    return false;
  } else {
    if (!fused2_supported<T>(a)) return false;
    constexpr int TX = F2_TX, TY = F2_TY;
    Fused2Args<T> f;
    f.a = a;
    f.En = En;
    f.Hn = Hn;
    f.beam1 = beam1;
    f.beam2 = beam2;
    f.js1 = js1 ? T(1) : T(0);
    f.js2 = js2 ? T(1) : T(0);
    f.ntx = a.nx / TX;
    f.nty = a.ny / TY;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t lds = fused_mt_bytes<T>(a.nmat);
    const int slots = resident_slots(dev, (const void *)fused2_sweep<T, TX, TY, false>, TX * TY,
                                     lds);
    const int ncol = f.ntx * f.nty;
    f.kc = pick_kc2(ncol, a.nz, slots);
    // 4-plane chunks at the z faces take the general path, the chunks
    // between (all lean) come first, under the XCD remap
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
      f.nremap = ncol * f.nchunk;
    }
    f.nb = ncol * f.nchunk;
    if (a.nt)
      hipLaunchKernelGGL((fused2_sweep<T, TX, TY, true>), dim3(f.nb), dim3(TX * TY), lds, s, f);
    else
      hipLaunchKernelGGL((fused2_sweep<T, TX, TY, false>), dim3(f.nb), dim3(TX * TY), lds, s, f);
    return true;
  }
}

}  // namespace fit
