// Row-wave fused leapfrog sweep: one launch reads E, H (old buffers) and
// writes E, H (new buffers), every field word once per cell.
//
// Same flat-window scheme as fused_sweep (fit_fused.hpp): a workgroup owns a
// TX x TY tile of (i, j) columns and KC planes and marches up in z; window
// cell (gi, gj) of plane p, gi in [x0-1, x0+TX], gj in [y0-1, y0+TY], holds
// the value at FLAT index gi + nx*gj + nx*ny*p (0 outside [0, N)), so every
// stencil neighbour n+-1, n+-nx, n+-nx*ny of the reference's flat-index curl
// (wakis/solverFIT3D.py:207-227) is the adjacent window cell, for every cell
// of the grid.  Staged values are u = L*pec*E_old of the cell they really are.
//
// Organisation:
//  * The workgroup is TY+1 rows of TX threads.  Thread row 0 computes H_new of
//    the low halo row (y0-1) with the same code the TY own rows use for their
//    own cells, so no own wave carries halo work.  A wave covers 64 cells of
//    one row: the row (j, plane offset) of a wave is uniform, so per-row
//    metrics are scalar loads and every field load is a wave-uniform 64-bit
//    row base plus a constant 32-bit lane offset.
//  * The low halo column (x0-1) is recomputed by the first TY lanes of thread
//    row 0.  When one tile spans the whole x extent it is not needed: flat
//    index n-1 of (0, j) is (nx-1, j-1), a cell of the tile's own w.
//  * Loads run two planes ahead (three register packets); the staged window
//    is a three-plane LDS ring and w = tL*pmc*H_new is triple buffered.  The
//    plane loop is unrolled by three, so every ring, buffer and packet index
//    is static.  One barrier per plane.
//  * Boundary masks are per-lane 0/1 factors with a wave-uniform override on
//    the z faces.  Window planes away from the z ends stage with one or two
//    multiplies per word; the planes next to the z ends take a general path.
#pragma once

#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "fit_fused.hpp"

namespace fit {

template <typename T, int TX, int TY>
struct RowSmem {
  static constexpr int RW = TX + 2, RH = TY + 2;
  T u[3][3][RH][RW];          // u = L*pec*E_old of 3 window planes
  T w[2][3][TY + 1][TX + 1];  // w = tL*pmc*H_new of 2 planes (row 0: y0-1, col 0: x0-1)
  T mt[MAX_MATERIALS * MAT_STRIDE];
};

template <typename T, int TX, int TY>
constexpr bool row_fits() {
  return sizeof(RowSmem<T, TX, TY>) <= 160 * 1024 && TX * (TY + 1) <= 1024;
}

template <typename T>
__device__ __forceinline__ T ldg(const T *base, uint32_t byte_off) {
  return *(const T *)((const char *)base + byte_off);
}
template <typename T>
__device__ __forceinline__ void stg(T *base, uint32_t byte_off, T v) {
  *(T *)((char *)base + byte_off) = v;
}

// z-face part of a tangential boundary mask (components x, y) at plane k:
// true if plane k is a z face, with that face's flag in `masked` (the z faces
// are written last, so they override the (i, j) part: wakis/boundaries.py:90-100).
__device__ __forceinline__ bool zface(uint32_t m, int k, int nz, bool &masked) {
  if (k == nz - 1) {
    masked = (m & ZHI) != 0;
    return true;
  }
  if (k == 0) {
    masked = (m & ZLO) != 0;
    return true;
  }
  masked = false;
  return false;
}

template <typename T, int RL>
struct RowPk {
  T we[RL][3];  // raw E of window plane p+2 (staging)
  T eo[3];      // E_old of the own cell at plane p
  T ho[3];      // H_old of the row cell at window plane p+1
  T iay;        // iA_y of the row cell at window plane p+1
  T itay;       // itA_y of the own cell at plane p
  int m1;       // material of the row cell at window plane p+1
  T ca, cb;     // halo column: H_old y, z at window plane p+1
  T ciay, ctlz;
  int cm;
  T jb[3];  // J of the own cell at plane p (J_BUFFER)
};

// Packed window-cell descriptor: true i (14 bits), j (14 bits), plane carry
// + 1 (2 bits), loaded (1 bit: a cell some active lane reads; its carry is
// then -1, 0 or 1).
__device__ __forceinline__ int wc_i(int d) { return d & 0x3fff; }
__device__ __forceinline__ int wc_j(int d) { return (d >> 14) & 0x3fff; }
__device__ __forceinline__ int wc_dk(int d) { return ((d >> 28) & 3) - 1; }
__device__ __forceinline__ bool wc_ok(int d) { return (d >> 30) & 1; }

template <typename T, int TX, int TY, int KC, bool FULLX, bool JB>
__global__ void __launch_bounds__(TX *(TY + 1)) row_sweep(FusedArgs<T> f) {
  static_assert(TX % 64 == 0, "a wave must cover 64 cells of one row");
  constexpr int NT = TX * (TY + 1);
  constexpr int RW = TX + 2, RH = TY + 2, RC = RW * RH;
  constexpr int RL = (RC + NT - 1) / NT;
  constexpr int WW = TX + 1;
  constexpr int UP = RH * RW;        // u: one component of one slot
  constexpr int WP = (TY + 1) * WW;  // w: one component of one slot
  constexpr uint32_t SZ = sizeof(T);
  __shared__ RowSmem<T, TX, TY> sm;
  const StepArgs<T> &a = f.a;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int plane = nx * ny;
  const int64_t NN = a.N;

  const int b = blockIdx.x;
  const int q8 = f.nb / 8;
  const int tile = (b < 8 * q8) ? (b % 8) * q8 + b / 8 : b;
  const int tx = tile % f.ntx;
  const int tr = tile / f.ntx;
  const int ty = tr % f.nty;
  const int kc = tr / f.nty;
  const int x0 = tx * TX, y0 = ty * TY;
  const int k0 = kc * KC;
  const int k1 = min(k0 + KC, nz);
  // Addresses are one 64-bit base per field at plane kb = k0-2 plus a 32-bit
  // byte offset (component * NC + plane and cell part): every accessed plane
  // is >= kb, and 2*NC + (KC+6) planes < 2^32 (checked by the launcher).
  const int kb = k0 - 2;
  const int64_t pb = (int64_t)plane * kb;
  const uint32_t PB = (uint32_t)plane * SZ;  // bytes per plane
  const uint32_t NC = (uint32_t)(NN * SZ);   // bytes per component

  // Read-only tables in the constant address space: a wave-uniform entry is
  // a scalar load, which does not wait behind the vector-load pipeline (the
  // tables are not written while a step runs).
  const CT<T> *gLx = (const CT<T> *)a.g.Lx;
  const CT<T> *gLy = (const CT<T> *)a.g.Ly;
  const CT<T> *gLz = (const CT<T> *)a.g.Lz;
  const CT<T> *gtLx = (const CT<T> *)a.g.tLx;
  const CT<T> *gtLy = (const CT<T> *)a.g.tLy;
  const CT<T> *gtLz = (const CT<T> *)a.g.tLz;
  const CT<T> *giAx = (const CT<T> *)a.g.iAx;
  const CT<T> *giAz = (const CT<T> *)a.g.iAz;
  const CT<T> *gitAx = (const CT<T> *)a.g.itAx;
  const CT<T> *gitAz = (const CT<T> *)a.g.itAz;
  const CT<T> *jline = (const CT<T> *)a.jline;
  const T *__restrict__ Eb = a.E + pb;
  const T *__restrict__ Jb = JB ? a.J + pb : nullptr;
  // J_SIGMA_E: J = sigma*E_old; J_ZERO: no J (J_BUFFER is the JB variant)
  const T jsig = a.jmode == J_SIGMA_E ? T(1) : T(0);
  const T *__restrict__ Hb = a.H + pb;
  T *__restrict__ Enb = f.En + pb;
  T *__restrict__ Hnb = f.Hn + pb;
  const uint8_t *__restrict__ matb = a.mat ? a.mat + pb : nullptr;
  const T *__restrict__ iAyb = a.g.iAy + (int64_t)nx * kb;
  const T *__restrict__ itAyb = a.g.itAy + (int64_t)nx * kb;
  const uint32_t pec = a.pec, pmc = a.pmc;

  for (int t = threadIdx.x; t < a.nmat * MAT_STRIDE; t += NT) sm.mt[t] = a.mtab[t];

  // ---------------- window staging cells of this thread ----------------
  // window (0, 0) of plane p is flat index p*plane + fx0, i.e. byte offset
  // (p - kb)*PB + fx0*SZ from the block bases
  const int fx0 = (x0 - 1) + nx * (y0 - 1);
  uint32_t woff[RL];  // byte offset of the cell from window (0, 0) of plane kb
  int wlds[RL];       // LDS element offset in one slot/component (-1: none)
  int wd[RL];         // packed descriptor
  T wfx[RL], wfy[RL], wfz[RL];  // (i, j)-masked L_x, L_y and 0/1 for E_z
  bool carry = false;
  // window cells read by an active lane: columns up to gi = nx, rows up to
  // gj = ny (cells past those only feed lanes whose results are discarded)
  const int ri_max = min(TX + 1, nx - x0 + 1), rj_max = min(TY + 1, ny - y0 + 1);
#pragma unroll
  for (int q = 0; q < RL; ++q) {
    const int c = threadIdx.x + q * NT;
    const int rj = c / RW, ri = c - rj * RW;
    const bool ok = c < RC && ri <= ri_max && rj <= rj_max;
    WinCell wc = win_cell(x0 - 1 + ri, y0 - 1 + rj, nx, ny);
    if (!ok) wc = WinCell{0, 0, 0, false};
    wd[q] = wc.i | (wc.j << 14) | ((wc.dk + 1) << 28) | ((ok ? 1 : 0) << 30);
    // A cell not loaded reads window cell (1, 1) (the cell (x0, y0), inside
    // the grid on every plane of the fast path) with zero factors: the fast
    // path loads without per-lane branches.
    woff[q] = (uint32_t)(fx0 + (ok ? ri + nx * rj : 1 + nx)) * SZ;  // fx0 < 0 wraps mod 2^32
    wlds[q] = c < RC ? rj * RW + ri : -1;
    wfx[q] = !ok || xy_mask(pec, 0, wc.i, wc.j, nx, ny) ? T(0) : gLx[wc.i];
    wfy[q] = !ok || xy_mask(pec, 1, wc.i, wc.j, nx, ny) ? T(0) : gLy[wc.j];
    wfz[q] = !ok || xy_mask(pec, 2, wc.i, wc.j, nx, ny) ? T(0) : T(1);
    if (ok && wc.dk != 0) carry = true;
  }
  // a window with a plane carry needs L_z of the carried plane, and the fast
  // path must keep every carried plane inside the grid
  const bool gen = __syncthreads_or(carry) != 0;
  const int pfast = gen ? nz - 3 : nz - 2;  // last window plane of the fast path

  // ---------------- the row cell of this thread ----------------
  const int lx = threadIdx.x % TX;
  const int ly = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
  const bool hrow = ly == 0;
  const int gi = x0 + lx;
  const int jr = __builtin_amdgcn_readfirstlane(hrow ? (y0 == 0 ? ny - 1 : y0 - 1) : y0 + ly - 1);
  const int dkr = (hrow && y0 == 0) ? -1 : 0;  // plane offset of the row's true cells
  const bool row_ok = hrow || jr < ny;
  const bool act = row_ok && gi < nx;
  const bool own = !hrow && act;
  const int ic = min(gi, nx - 1);
  const int jc = __builtin_amdgcn_readfirstlane(min(jr, ny - 1));
  const uint32_t voff = (uint32_t)(ic + nx * jc) * SZ;  // row cell in a plane
  const uint32_t vmo = (uint32_t)(ic + nx * jc);        // same, bytes of a uint8 array
  const uint32_t vxo = (uint32_t)ic * SZ;               // i in an (i, k) metric table
  const int li = lx + 1, lj = ly;
  // (i, j) parts of the masks as 0/1 factors (0: masked)
  const T fhx = xy_mask(pmc, 0, ic, jc, nx, ny) ? T(0) : T(1);
  const T fhy = xy_mask(pmc, 1, ic, jc, nx, ny) ? T(0) : T(1);
  const T fex = xy_mask(pec, 0, ic, jc, nx, ny) ? T(0) : T(1);
  const T fey = xy_mask(pec, 1, ic, jc, nx, ny) ? T(0) : T(1);
  const T fact = act ? T(1) : T(0);
  const T tLx = gtLx[ic];
  const T tLy = gtLy[jc];
  const T ciAz = xy_mask(pmc, 2, ic, jc, nx, ny) ? T(0) : giAz[ic + nx * jc];
  const T ctLz = xy_mask(pmc, 2, ic, jc, nx, ny) || !act ? T(0) : T(1);  // w_z factor
  const T caz = xy_mask(pec, 2, ic, jc, nx, ny) ? T(0) : gitAz[ic + nx * jc];
  const bool beam_col = own && a.beam_mode != BEAM_OFF && gi == a.ixs && jr == a.iys;
  // wake probe: byte offset of the own cell's history row (column = step)
  const bool probe_on = a.hist != nullptr && a.step < a.hist_nt;
  T *__restrict__ histb = probe_on ? a.hist + a.step : nullptr;
  const uint32_t HK = (uint32_t)a.hist_nt * SZ;  // bytes per plane of a history row
  bool pon = false;
  uint32_t hoff = 0;
  if (own && probe_on) {
    const int di = gi - a.ixt, dj = jr - a.iyt;
    if (di >= -a.hw && di <= a.hw && dj >= -a.hw && dj <= a.hw) {
      pon = true;
      hoff = (uint32_t)((di + a.hw) * (2 * a.hw + 1) + (dj + a.hw)) * (uint32_t)nz * HK;
    }
  }
  // LDS element offsets in a w slot/component
  const int wo_c = lj * WW + li;
  const int wo_s = (lj > 0 ? lj - 1 : 0) * WW + li;                                  // (i, j-1)
  const int wo_w = (FULLX && lx == 0 && lj > 0) ? (lj - 1) * WW + nx : lj * WW + li - 1;  // n-1
  const int uo = lj * RW + li;

  // ---------------- halo column (x0-1, y0+t), t = lx < TY, in thread row 0 ----------------
  const bool hcol = !FULLX && hrow && lx < TY;
  int hd = 0;  // packed descriptor of the halo-column cell (+ bit 31: pmc_y (i, j) mask)
  T h_tly = 0, h_iaz = 0;
  if (hcol) {
    const WinCell hc = win_cell(x0 - 1, y0 + lx, nx, ny);
    const bool ok = y0 + lx < ny;
    hd = hc.i | (hc.j << 14) | ((hc.dk + 1) << 28) | ((ok ? 1 : 0) << 30);
    if (xy_mask(pmc, 1, hc.i, hc.j, nx, ny)) hd |= 1 << 31;
    h_tly = gtLy[hc.j];
    h_iaz = xy_mask(pmc, 2, hc.i, hc.j, nx, ny) ? T(0) : giAz[hc.i + nx * hc.j];
  }

  // ---------------- loads ----------------
  // raw E of window plane p (0 outside the grid).  FAST: p in [1, pfast],
  // where every staged cell is inside the grid.
  auto load_win = [&](auto fastc, int p, T (&we)[RL][3]) {
    constexpr bool FAST = decltype(fastc)::value;
    const uint32_t po = (uint32_t)(p - kb) * PB;
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const uint32_t o = woff[q] + po;
      if constexpr (FAST) {
        // unconditional: every address is inside the grid on a fast plane
        we[q][0] = ldg(Eb, o);
        we[q][1] = ldg(Eb, o + NC);
        we[q][2] = ldg(Eb, o + 2 * NC);
      } else {
        const int kk = p + wc_dk(wd[q]);
        T vx = 0, vy = 0, vz = 0;
        if (wc_ok(wd[q]) && kk >= 0 && kk < nz) {
          vx = ldg(Eb, o);
          vy = ldg(Eb, o + NC);
          vz = ldg(Eb, o + 2 * NC);
        }
        we[q][0] = vx;
        we[q][1] = vy;
        we[q][2] = vz;
      }
    }
  };
  // row-cell data of window plane pw1 (true plane pw1+dkr), clamped into the
  // grid (a clamped value is never used)
  auto load_row = [&](int pw1, RowPk<T, RL> &pk) {
    const int kt = min(max(pw1 + dkr, 0), nz - 1);
    const uint32_t o = (uint32_t)(kt - kb) * PB + voff;
    pk.ho[0] = ldg(Hb, o);
    pk.ho[1] = ldg(Hb, o + NC);
    pk.ho[2] = ldg(Hb, o + 2 * NC);
    pk.iay = ldg(iAyb, (uint32_t)((kt - kb) * nx) * SZ + vxo);
    pk.m1 = matb ? (int)matb[(uint32_t)(kt - kb) * (uint32_t)plane + vmo] : 0;
    if constexpr (!FULLX) {
      pk.ca = pk.cb = pk.ciay = pk.ctlz = T(0);
      pk.cm = 0;
      if (hcol && wc_ok(hd)) {
        const int hk = pw1 + wc_dk(hd);
        if (hk >= 0 && hk < nz) {
          const uint32_t hn = (uint32_t)(hk - kb) * (uint32_t)plane + (uint32_t)(wc_i(hd) + nx * wc_j(hd));
          pk.ca = ldg(Hb, hn * SZ + NC);
          pk.cb = ldg(Hb, hn * SZ + 2 * NC);
          pk.cm = matb ? (int)matb[hn] : 0;
          pk.ciay = ldg(iAyb, (uint32_t)((hk - kb) * nx + wc_i(hd)) * SZ);
          pk.ctlz = gtLz[hk];
        }
      }
    }
  };
  // packet p: window plane p+2, own E_old(p) and itA_y(p) (p < nz), row-cell
  // data of window plane p+1
  auto load_pk = [&](auto fastc, int p, RowPk<T, RL> &pk) {
    load_win(fastc, p + 2, pk.we);
    if (!hrow) {
      const uint32_t o = (uint32_t)(p - kb) * PB + voff;
      pk.eo[0] = ldg(Eb, o);
      pk.eo[1] = ldg(Eb, o + NC);
      pk.eo[2] = ldg(Eb, o + 2 * NC);
      pk.itay = ldg(itAyb, (uint32_t)((p - kb) * nx) * SZ + vxo);
      if constexpr (JB) {
        pk.jb[0] = ldg(Jb, o);
        pk.jb[1] = ldg(Jb, o + NC);
        pk.jb[2] = ldg(Jb, o + 2 * NC);
      }
    } else {
      pk.eo[0] = pk.eo[1] = pk.eo[2] = T(0);
      pk.itay = T(0);
      if constexpr (JB) pk.jb[0] = pk.jb[1] = pk.jb[2] = T(0);
    }
    load_row(p + 1, pk);
  };

  // ---------------- staging of window plane p into ring slot s ----------------
  auto stage = [&](auto fastc, int p, int s, const T (&we)[RL][3]) {
    constexpr bool FAST = decltype(fastc)::value;
    T *u = &sm.u[s][0][0][0];
    if constexpr (FAST) {
      // every staged cell is at plane p + dk in [1, nz-2]: no z face
      const T lz0 = gLz[p];
      T lzm = lz0, lzp = lz0;
      if (gen) {
        lzm = gLz[p - 1];
        lzp = gLz[p + 1];
      }
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        if (wlds[q] >= 0) {
          T lz = lz0;
          if (gen) {
            const int dk = wc_dk(wd[q]);
            lz = dk < 0 ? lzm : (dk > 0 ? lzp : lz0);
          }
          u[wlds[q]] = wfx[q] * we[q][0];
          u[UP + wlds[q]] = wfy[q] * we[q][1];
          u[2 * UP + wlds[q]] = wfz[q] * (lz * we[q][2]);
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        if (wlds[q] >= 0) {
          const int kk = p + wc_dk(wd[q]);
          const int ci = wc_i(wd[q]), cj = wc_j(wd[q]);
          T vx = 0, vy = 0, vz = 0;
          if (wc_ok(wd[q]) && kk >= 0 && kk < nz) {
            if (!full_mask(pec, 0, xy_mask(pec, 0, ci, cj, nx, ny), kk, nz)) vx = gLx[ci] * we[q][0];
            if (!full_mask(pec, 1, xy_mask(pec, 1, ci, cj, nx, ny), kk, nz)) vy = gLy[cj] * we[q][1];
            vz = wfz[q] * (gLz[kk] * we[q][2]);
          }
          u[wlds[q]] = vx;
          u[UP + wlds[q]] = vy;
          u[2 * UP + wlds[q]] = vz;
        }
      }
    }
  };

  // ---------------- H_new of the row cell at window plane p ----------------
  // u slots sa (plane p) and sb (plane p+1).  Masks are 0/1 factors: the
  // (i, j) part per lane, replaced on a z face by that face's flag.
  auto h_row = [&](int p, int sa, int sb, const RowPk<T, RL> &pk, T (&h)[3], T (&w)[3]) {
    const int kt = p + dkr;
    const int ktc = min(max(kt, 0), nz - 1);
    const T fin = (kt >= 0 && kt < nz) ? T(1) : T(0);  // uniform
    const T *u0 = &sm.u[sa][0][0][0] + uo;
    const T *u1 = &sm.u[sb][0][0][0] + uo;
    const T cux = ((u0[UP] - u1[UP]) - u0[2 * UP]) + u0[2 * UP + RW];
    const T cuy = ((u0[2 * UP] - u0[2 * UP + 1]) - u0[0]) + u1[0];
    const T cuz = ((u0[0] - u0[RW]) - u0[UP]) + u0[UP + 1];
    const T *mt = sm.mt + pk.m1 * MAT_STRIDE;
    const bool zf = ktc == 0 || ktc == nz - 1;
    const T zv = (ktc == nz - 1 ? (pmc & ZHI) : (pmc & ZLO)) ? T(0) : T(1);
    const T mx = zf ? zv : fhx;
    const T my = zf ? zv : fhy;
    const T iax = giAx[jc + ny * ktc];  // scalar
    const T tlz = gtLz[ktc];            // scalar
    h[0] = pk.ho[0] - (mx * (mt[0] * iax)) * cux;
    h[1] = pk.ho[1] - (my * (mt[1] * pk.iay)) * cuy;
    h[2] = pk.ho[2] - (mt[2] * ciAz) * cuz;
    const T fv = fact * fin;
    w[0] = (mx * fv) * (tLx * h[0]);
    w[1] = (my * fv) * (tLy * h[1]);
    w[2] = (ctLz * fin) * (tlz * h[2]);
  };
  // halo column cell at window (0, lx+1), plane p: w_y, w_z into w slot sw
  auto h_col = [&](int p, int sa, int sb, int sw, const RowPk<T, RL> &pk) {
    if constexpr (!FULLX) {
      const int hk = p + wc_dk(hd);
      T wa = 0, wb = 0;
      if (wc_ok(hd) && hk >= 0 && hk < nz) {
        const int huo = (lx + 1) * RW;
        const T *u0 = &sm.u[sa][0][0][0] + huo;
        const T *u1 = &sm.u[sb][0][0][0] + huo;
        const T cuy = ((u0[2 * UP] - u0[2 * UP + 1]) - u0[0]) + u1[0];
        const T cuz = ((u0[0] - u0[RW]) - u0[UP]) + u0[UP + 1];
        const bool r1 = full_mask(pmc, 1, (hd >> 31) & 1, hk, nz);
        const T *mt = sm.mt + pk.cm * MAT_STRIDE;
        const T ha = pk.ca - (r1 ? T(0) : mt[1] * pk.ciay) * cuy;
        const T hb = pk.cb - (mt[2] * h_iaz) * cuz;
        wa = r1 ? T(0) : h_tly * ha;
        wb = (h_iaz == T(0)) ? T(0) : pk.ctlz * hb;
      }
      const int hwo = (lx + 1) * WW;
      T *wv = &sm.w[sw][0][0][0];
      wv[WP + hwo] = wa;
      wv[2 * WP + hwo] = wb;
    }
  };

  const std::integral_constant<bool, true> FASTC;
  const std::integral_constant<bool, false> SLOWC;

  // ---------------- prologue ----------------
  {
    T we[RL][3];
    load_win(SLOWC, k0 - 1, we);
    stage(SLOWC, k0 - 1, 0, we);
    load_win(SLOWC, k0, we);
    stage(SLOWC, k0, 1, we);
    load_win(SLOWC, k0 + 1, we);
    stage(SLOWC, k0 + 1, 2, we);
  }
  // Fast iterations k (window plane k+2 staged, k+3 loaded) while k+3 <= pfast;
  // an even number of them, so the packet pipeline ends in pkA.
  int kf = min(k1, pfast - 2);
  if (kf < k0) kf = k0;
  kf = k0 + ((kf - k0) & ~1);
  RowPk<T, RL> pkA, pkB;
  if (kf > k0) load_pk(FASTC, k0, pkA);
  RowPk<T, RL> pro;
  load_row(k0, pro);
  T hm0 = 0, hm1 = 0, iaym = 0;
  int mm = 0, mcur = 0;
  if (!hrow && k0 >= 1) {
    const uint32_t o = (uint32_t)(k0 - 1 - kb) * PB + voff;
    hm0 = ldg(Hb, o);
    hm1 = ldg(Hb, o + NC);
    iaym = ldg(iAyb, (uint32_t)((k0 - 1 - kb) * nx) * SZ + vxo);
    mm = matb ? (int)matb[(uint32_t)(k0 - 1 - kb) * (uint32_t)plane + vmo] : 0;
  }
  if (!hrow) mcur = matb ? (int)matb[(uint32_t)(k0 - kb) * (uint32_t)plane + vmo] : 0;
  __syncthreads();
  T wxp = 0, wyp = 0;
  if (!hrow && k0 >= 1) {
    // w_x, w_y of the own cell at plane k0-1 (slots 0: k0-1, 1: k0)
    const T *u0 = &sm.u[0][0][0][0] + uo;
    const T *u1 = &sm.u[1][0][0][0] + uo;
    const T cux = ((u0[UP] - u1[UP]) - u0[2 * UP]) + u0[2 * UP + RW];
    const T cuy = ((u0[2 * UP] - u0[2 * UP + 1]) - u0[0]) + u1[0];
    const int km = k0 - 1;
    const bool zf = km == 0 || km == nz - 1;
    const T zv = (km == nz - 1 ? (pmc & ZHI) : (pmc & ZLO)) ? T(0) : T(1);
    const T mx = zf ? zv : fhx;
    const T my = zf ? zv : fhy;
    const T *mt = sm.mt + mm * MAT_STRIDE;
    const T hx = hm0 - (mx * (mt[0] * giAx[jc + ny * km])) * cux;
    const T hy = hm1 - (my * (mt[1] * iaym)) * cuy;
    wxp = (mx * fact) * (tLx * hx);
    wyp = (my * fact) * (tLy * hy);
  }
  T wx, wy, wz;
  {
    T h[3], w[3];
    h_row(k0, 1, 2, pro, h, w);
    if (own) {
      const uint32_t o = (uint32_t)(k0 - kb) * PB + voff;
      stg(Hnb, o, h[0]);
      stg(Hnb, o + NC, h[1]);
      stg(Hnb, o + 2 * NC, h[2]);
    }
    T *wv = &sm.w[0][0][0][0];
    wv[wo_c] = w[0];
    wv[WP + wo_c] = w[1];
    wv[2 * WP + wo_c] = w[2];
    wx = w[0];
    wy = w[1];
    wz = w[2];
    if (hcol) h_col(k0, 1, 2, 0, pro);
  }
  __syncthreads();  // the loop's first staging reuses slot 0

  // ---------------- one plane: E_new(k), H_new(k+1) ----------------
  int rs = 0;  // ring slot of window plane k+2
  // ws: w slot of plane k; cur: packet of plane k (loaded); nxt: packet k+1
  // is loaded into it when LOADNEXT.
  auto iter = [&](auto fastc, auto loadc, int k, int ws, const RowPk<T, RL> &cur,
                  RowPk<T, RL> &nxt) {
    constexpr bool LOADNEXT = decltype(loadc)::value;
    const int r1 = rs == 0 ? 2 : rs - 1;  // ring slot of window plane k+1
    // issue the next packet's loads first: staging then waits only for this
    // packet's window loads
    if constexpr (LOADNEXT) load_pk(fastc, k + 1, nxt);
    if (k + 2 <= k1) stage(fastc, k + 2, rs, cur.we);
    const T jl = (a.beam_mode != BEAM_OFF) ? jline[k] : T(0);
    const T itax = gitAx[jc + ny * k];  // scalar
    __syncthreads();
    // E_new(k)
    if (own) {
      const uint32_t o = (uint32_t)(k - kb) * PB + voff;
      const T *w0 = &sm.w[ws][0][0][0];
      const T *mt = sm.mt + mcur * MAT_STRIDE;
      const T ex = cur.eo[0], ey = cur.eo[1], ez = cur.eo[2];
      if (pon) stg(histb, hoff + (uint32_t)k * HK, ez);
      T jx, jy, jz;
      if constexpr (JB) {
        jx = cur.jb[0];
        jy = cur.jb[1];
        jz = cur.jb[2];
      } else {
        jx = jsig * (mt[6] * ex);
        jy = jsig * (mt[7] * ey);
        jz = jsig * (mt[8] * ez);
      }
      if (beam_col) jz = jl;
      const bool zf = k == 0 || k == nz - 1;
      const T zv = (k == nz - 1 ? (pec & ZHI) : (pec & ZLO)) ? T(0) : T(1);
      const T ax = (zf ? zv : fex) * itax;
      const T ay = (zf ? zv : fey) * cur.itay;
      const T cx = ((wz - w0[2 * WP + wo_s]) - wy) + wyp;
      const T cy = ((wx - wxp) - wz) + w0[2 * WP + wo_w];
      const T cz = ((wy - w0[WP + wo_w]) - wx) + w0[wo_s];
      stg(Enb, o, ex + mt[3] * (ax * cx - jx));
      stg(Enb, o + NC, ey + mt[4] * (ay * cy - jy));
      stg(Enb, o + 2 * NC, ez + mt[5] * (caz * cz - jz));
    }
    wxp = wx;
    wyp = wy;
    // H_new(k+1)
    if (k + 1 < k1) {
      T h[3], w[3];
      h_row(k + 1, r1, rs, cur, h, w);
      if (own) {
        const uint32_t o = (uint32_t)(k + 1 - kb) * PB + voff;
        stg(Hnb, o, h[0]);
        stg(Hnb, o + NC, h[1]);
        stg(Hnb, o + 2 * NC, h[2]);
      }
      T *wv = &sm.w[ws ^ 1][0][0][0];
      wv[wo_c] = w[0];
      wv[WP + wo_c] = w[1];
      wv[2 * WP + wo_c] = w[2];
      wx = w[0];
      wy = w[1];
      wz = w[2];
      if (hcol) h_col(k + 1, r1, rs, ws ^ 1, cur);
      mcur = cur.m1;
    }
    rs = rs == 2 ? 0 : rs + 1;
  };
  const std::integral_constant<bool, true> LOAD;
  const std::integral_constant<bool, false> NOLOAD;
  int k = k0;
  // fast pipelined iterations: packet k+1 is loaded while plane k is computed
  for (; k < kf; k += 2) {
    iter(FASTC, LOAD, k, 0, pkA, pkB);
    if (k + 2 < kf) iter(FASTC, LOAD, k + 1, 1, pkB, pkA);
    else iter(FASTC, NOLOAD, k + 1, 1, pkB, pkA);
  }
  // remaining planes (near the z end of the grid): load, then compute
  for (; k < k1; ++k) {
    load_pk(SLOWC, k, pkA);
    iter(SLOWC, NOLOAD, k, (k - k0) & 1, pkA, pkB);
  }
}

// Tile configurations of the row-wave sweep (TX, TY, KC).
enum {
  ROW_CFG_256x3 = 0,
  ROW_CFG_128x7 = 1,
  ROW_CFG_64x7 = 2,
  ROW_CFG_64x15 = 3,
  ROW_CFG_128x3 = 4,
  ROW_CFG_64x7_K64 = 5,
  ROW_NCFG = 6
};

inline void row_dims(int cfg, int &tx, int &ty, int &kc) {
  switch (cfg) {
    case ROW_CFG_256x3: tx = 256; ty = 3; kc = 32; break;
    case ROW_CFG_128x7: tx = 128; ty = 7; kc = 32; break;
    case ROW_CFG_64x15: tx = 64; ty = 15; kc = 32; break;
    case ROW_CFG_128x3: tx = 128; ty = 3; kc = 32; break;
    case ROW_CFG_64x7_K64: tx = 64; ty = 7; kc = 64; break;
    default: tx = 64; ty = 7; kc = 32; break;
  }
}

template <typename T>
inline bool row_cfg_ok(int cfg) {
  switch (cfg) {
    case ROW_CFG_256x3: return row_fits<T, 256, 3>();
    case ROW_CFG_128x7: return row_fits<T, 128, 7>();
    case ROW_CFG_64x15: return row_fits<T, 64, 15>();
    case ROW_CFG_128x3: return row_fits<T, 128, 3>();
    default: return row_fits<T, 64, 7>();
  }
}

// The row sweep addresses fields with 32-bit byte offsets from a block base
// (2 components + (KC+6) planes) and the probe history with 32-bit offsets;
// window descriptors pack i and j in 14 bits each.
template <typename T>
inline bool row_offsets_fit(const StepArgs<T> &a, int KC) {
  const double nc = (double)a.N * sizeof(T);
  const double pl = (double)a.nx * a.ny * sizeof(T);
  const double hb = a.hist ? (double)(2 * a.hw + 1) * (2 * a.hw + 1) * a.nz * a.hist_nt * sizeof(T) : 0.0;
  return 2.0 * nc + (KC + 6) * pl < 4294967296.0 && hb < 4294967296.0 && a.nx < 16384 &&
         a.ny < 16384;
}

// Launches the row sweep for these args (or only checks, when dry); false if
// the variant it needs cannot run (LDS, 32-bit offsets, scratch spills).
template <typename T, int TX, int TY, int KC>
inline bool launch_row(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s, bool dry) {
  if constexpr (!row_fits<T, TX, TY>()) {
    (void)a; (void)En; (void)Hn; (void)s; (void)dry;
    return false;
  } else {
    if (!row_offsets_fit<T>(a, KC)) return false;
    FusedArgs<T> f;
    f.a = a;
    f.En = En;
    f.Hn = Hn;
    f.ntx = (a.nx + TX - 1) / TX;
    f.nty = (a.ny + TY - 1) / TY;
    f.nkc = (a.nz + KC - 1) / KC;
    f.nb = f.ntx * f.nty * f.nkc;
    const dim3 grid(f.nb), block(TX * (TY + 1));
    const bool jb = a.jmode == J_BUFFER;
    void (*fn)(FusedArgs<T>);
    if (f.ntx == 1) fn = jb ? row_sweep<T, TX, TY, KC, true, true> : row_sweep<T, TX, TY, KC, true, false>;
    else fn = jb ? row_sweep<T, TX, TY, KC, false, true> : row_sweep<T, TX, TY, KC, false, false>;
    if (!kernel_spill_free((const void *)fn)) return false;
    if (!dry) hipLaunchKernelGGL(fn, grid, block, 0, s, f);
    return true;
  }
}

template <typename T>
inline bool launch_row_cfg(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s, int cfg, bool dry) {
  switch (cfg) {
    case ROW_CFG_256x3: return launch_row<T, 256, 3, 32>(a, En, Hn, s, dry);
    case ROW_CFG_128x7: return launch_row<T, 128, 7, 32>(a, En, Hn, s, dry);
    case ROW_CFG_64x15: return launch_row<T, 64, 15, 32>(a, En, Hn, s, dry);
    case ROW_CFG_128x3: return launch_row<T, 128, 3, 32>(a, En, Hn, s, dry);
    case ROW_CFG_64x7_K64: return launch_row<T, 64, 7, 64>(a, En, Hn, s, dry);
    default: return launch_row<T, 64, 7, 32>(a, En, Hn, s, dry);
  }
}

// ---------------- all sweep configurations ----------------
// ids 0 .. FUSED_NCFG-1: fused_sweep tiles; then the row-wave tiles.
constexpr int SWEEP_NCFG = FUSED_NCFG + ROW_NCFG;

// dims: tile x, tile y (own rows), z chunk
inline void sweep_dims(int cfg, int &tx, int &ty, int &kc) {
  if (cfg < FUSED_NCFG) fused_dims(cfg, tx, ty, kc);
  else row_dims(cfg - FUSED_NCFG, tx, ty, kc);
}

template <typename T>
inline bool sweep_cfg_ok(int cfg) {
  if (cfg < 0 || cfg >= SWEEP_NCFG) return false;
  return cfg < FUSED_NCFG ? fused_cfg_ok<T>(cfg) : row_cfg_ok<T>(cfg - FUSED_NCFG);
}

// Default when no tuning was done: the fused_sweep tile of least modelled
// cost (fused_cost; the fused_sweep family measured faster than the row-wave
// tiles on the 256 x 256 x 768 pillbox); the autotuner considers both.
inline int sweep_default(int nx, int ny, int nz, bool fp64) {
  return fp64 ? fused_pick<double>(&nx, &ny, &nz, 1) : fused_pick<float>(&nx, &ny, &nz, 1);
}

// Whether configuration cfg runs as itself for these args (a row-sweep
// configuration that cannot run falls back to the fused_sweep default).
template <typename T>
inline bool sweep_cfg_runs(const StepArgs<T> &a, int cfg) {
  if (cfg < FUSED_NCFG) return fused_cfg_ok<T>(cfg) && fused_spill_free<T>(cfg);
  return launch_row_cfg<T>(a, nullptr, nullptr, nullptr, cfg - FUSED_NCFG, true);
}

// false: nothing launched (the fused_sweep configuration needed spills)
template <typename T>
inline bool launch_sweep(const StepArgs<T> &a, T *En, T *Hn, hipStream_t s, int cfg) {
  if (cfg >= FUSED_NCFG) {
    if (launch_row_cfg<T>(a, En, Hn, s, cfg - FUSED_NCFG, false)) return true;
    cfg = fused_default(a.nx, sizeof(T) == 8);
  }
  return launch_fused<T>(a, En, Hn, s, cfg);
}

}  // namespace fit
