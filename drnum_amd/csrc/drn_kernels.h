// drn_kernels.h -- gfx950 kernels of the per-patch explicit update.
//
//   k_stage_cell   one thread per cell; residual from the six faces of the cell
//                  (each interior face evaluated by both neighbours), fused with
//                  the RK stage update.  Residual summation order per cell follows
//                  CartesianIterator::compute (cartesianiterator.h:132-308).
//   k_stage_march  2.5-D form: a workgroup owns a chunk of the (j,k) plane of a
//                  patch and marches along i; each face flux is computed once
//                  (x faces in registers, y/z faces through LDS).
//   k_exchange_*   gather form of Patch::accessDonorDataDirect (patch.cpp:573-653).
//   k_ls_*         CPU_LevelSetObjectBC<LSOBCCompressibleEulerOp> (cpu_levelsetobjectbc.h:58-90).
//   k_sums_*       FP64 core sums (deterministic two-pass reduction).
#pragma once

#include "drn_physics.h"

namespace drn {

// cells [i0,i1) x [j0,j1) x [k0,k1) of one patch
struct CellBox {
  int32_t i0, i1, j0, j1, k0, k1;
};

// each range test as one unsigned compare: (unsigned)(x - lo) < (unsigned)(hi - lo)
__device__ __forceinline__ bool in_box(const CellBox& b, int i, int j, int k) {
  return (uint32_t)(i - b.i0) < (uint32_t)(b.i1 - b.i0) && (uint32_t)(j - b.j0) < (uint32_t)(b.j1 - b.j0) &&
         (uint32_t)(k - b.k0) < (uint32_t)(b.k1 - b.k0);
}

struct PatchDev {
  int32_t ni, nj, nk, njk;
  uint32_t off;   // first global cell of the patch in a state buffer
  int32_t pad0;
  double dx, dy, dz;
  int32_t bc[6];
  int32_t seek[6];
  double par[6][16];
  // Exchange copies of the z-face seek layers done by the stage kernel itself
  // (used only when a launch asks for it): a cell in push[s] also writes its
  // new value, as the exchange would copy it, to global cell
  // push_base[s] + i*push_si[s] + j*push_sj[s] + k of the receiving patch, and
  // a cell in recv[s] is such a receiver and does not store its own value.
  CellBox push[2];
  int64_t push_base[2];
  uint32_t push_si[2], push_sj[2];
  CellBox recv[2];
  int32_t npush, nrecv;
};

// work item of the cell-per-thread stage kernel: patch, plane i, first (j,k) cell
struct TileCell {
  int32_t patch, i, c2, pad;
};

// work item of the marching stage kernel: patch, i range, (j,k) chunk start
struct TileMarch {
  int32_t patch, i0, i1, c0;
};

// work item of the wave-tile stage kernel: patch, i range [i0, i1), and the
// first row / column (j0, k0) of the wavefront's 8x8 lane tile
struct TileWave {
  int32_t patch;
  int16_t i0, i1, j0, k0;
  int32_t pad;
};

template <typename T>
__device__ __forceinline__ void zero5(T* f) {
#pragma unroll
  for (int v = 0; v < 5; ++v) f[v] = T(0);
}

// face between cells cm1 (minus side) and cp1 (plus side) along DIR; cm2 and cp2
// are the next cells outwards, valid only when has_m2 / has_p2 (upwind2.h:38-50).
// The cells are 32-bit byte offsets into the variable arrays sv[0..4] (uniform
// bases), so every load is a scalar-base + 32-bit vector-offset access.
template <typename T>
__device__ __forceinline__ T ldv(const char* base, uint32_t boff) {
  return *reinterpret_cast<const T*>(base + boff);
}

template <typename T, int RECON, int LIM, bool FAST, int DIR>
__device__ __forceinline__ void face_flux_g(const char* const* sv, uint32_t cm2, uint32_t cm1, uint32_t cp1,
                                            uint32_t cp2, bool has_m2, bool has_p2, T A, T* f) {
  T ua[5], ub[5], vl[5], vr[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    ua[v] = ldv<T>(sv[v], cm1);
    ub[v] = ldv<T>(sv[v], cp1);
    vl[v] = ua[v];
    vr[v] = ub[v];
  }
  if constexpr (RECON == RECON_UPWIND2) {
    if (has_m2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T m = ldv<T>(sv[v], cm2);
        vl[v] = ua[v] + T(0.5) * limited<T, LIM, FAST>(ua[v] - m, ub[v] - ua[v]);
      }
    }
    if (has_p2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T q = ldv<T>(sv[v], cp2);
        vr[v] = ub[v] + T(0.5) * limited<T, LIM, FAST>(ub[v] - q, ua[v] - ub[v]);
      }
    }
  }
  ausm_flux<T, FAST, DIR>(vl, vr, A, f);
}

// the same face from values already in registers: a = cell on the minus side,
// b = cell on the plus side, m / q the next cells outwards (valid when hm / hp)
template <typename T, int RECON, int LIM, bool FAST, int DIR>
__device__ __forceinline__ void face_flux_r(const T* m, const T* a, const T* b, const T* q, bool hm, bool hp, T A,
                                            T* f) {
  T vl[5], vr[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    vl[v] = a[v];
    vr[v] = b[v];
  }
  if constexpr (RECON == RECON_UPWIND2) {
    if (hm) {
#pragma unroll
      for (int v = 0; v < 5; ++v) vl[v] = a[v] + T(0.5) * limited<T, LIM, FAST>(a[v] - m[v], b[v] - a[v]);
    }
    if (hp) {
#pragma unroll
      for (int v = 0; v < 5; ++v) vr[v] = b[v] + T(0.5) * limited<T, LIM, FAST>(b[v] - q[v], a[v] - b[v]);
    }
  }
  ausm_flux_set<T, FAST, DIR>(vl, vr, A, f);  // every caller passes a zeroed f
}

// getXGrad / getYGrad / getZGrad (cartesianpatch_common.h:188-300) of one cell
// along AX from its two neighbours lo and hi on that axis: central difference,
// one sided at the patch ends with D doubled per clamped side (there lo or hi
// is the cell itself); the y gradient is 0 when sizeJ() <= 2.  `dh` is
// (real)0.5/delta along AX, idx/n the cell's index and count along AX.
template <typename T, int AX>
__device__ __forceinline__ void cell_grad_v(const T* lo, const T* hi, T dh, int idx, int n, T* g) {
  if (AX == 1 && n <= 2) {
    zero5(g);
    return;
  }
  T D = dh;
  if (idx < 1) D *= T(2);
  if (idx + 1 >= n) D *= T(2);
#pragma unroll
  for (int v = 0; v < 5; ++v) g[v] = (hi[v] - lo[v]) * D;
}

// Split faces (Patch::isSplitFace, patch_common.h:500-503; Upwind2 falls back
// to first order there, upwind2.h:41): `ins` holds the inside bits of the cells
// idx-2 .. idx+2 of the axis in bits 0..4 (a bit of a cell outside the patch is
// never read).  A side's slope enters a face only when its outer cell i0 and
// its cell i1 are on the same side of the inside boundary.
__device__ __forceinline__ bool split_bits(uint32_t ins, int a, int b) { return ((ins >> a) ^ (ins >> b)) & 1u; }

// the same gradient of the cell at byte offset cc from its neighbours in
// memory (byte stride st along AX)
template <typename T, int AX>
__device__ __forceinline__ void cell_grad_ld(const char* const* sv, uint32_t cc, uint32_t st, T dh, int idx, int n,
                                             T* g) {
  if (AX == 1 && n <= 2) {
    zero5(g);
    return;
  }
  const uint32_t lo = idx < 1 ? cc : cc - st, hi = idx + 1 >= n ? cc : cc + st;
  T a[5], b[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    a[v] = ldv<T>(sv[v], lo);
    b[v] = ldv<T>(sv[v], hi);
  }
  cell_grad_v<T, AX>(a, b, dh, idx, n, g);
}

// viscous part of the face between the cells at byte offsets cl (minus side)
// and cr along DIR, their tangential gradients loaded from memory (both cells
// share the indices idx[] on the tangential axes; st[], dh[], n[] per axis)
template <typename T, bool FAST, int DIR>
__device__ __forceinline__ void visc_face_ld(const char* const* sv, uint32_t cl, uint32_t cr, const T* vl,
                                             const T* vr, const uint32_t* st, const T* dh, const int* idx,
                                             const int* n, T dn, T A, T* f) {
  constexpr int TA = DIR == 0 ? 1 : 0, TB = DIR == 2 ? 1 : 2;
  T gla[5], glb[5], gra[5], grb[5];
  cell_grad_ld<T, TA>(sv, cl, st[TA], dh[TA], idx[TA], n[TA], gla);
  cell_grad_ld<T, TB>(sv, cl, st[TB], dh[TB], idx[TB], n[TB], glb);
  cell_grad_ld<T, TA>(sv, cr, st[TA], dh[TA], idx[TA], n[TA], gra);
  cell_grad_ld<T, TB>(sv, cr, st[TB], dh[TB], idx[TB], n[TB], grb);
  visc_face<T, FAST, DIR>(vl, vr, gla, gra, glb, grb, dn, A, f);
}

// The two faces of a cell along DIR (index idx of n along that axis, byte
// stride st): axis_load loads the four neighbours once (the cell itself stands
// in for one that does not exist; NEAR: the next two also when n <= 2, for the
// viscous cell gradients), axis_eval adds the minus face to res and returns
// the plus face in fp (zero when the face does not exist).  vm(f) / vp(f) add
// the viscous part to the minus / plus face's Euler flux f.
// MAYSKIP, minus = false: the minus face is left out (another wavefront
// evaluates it as its plus face), and with it the outer minus neighbour's load
// and slope
template <typename T, bool MAYSKIP = false, bool NEAR = false>
__device__ __forceinline__ void axis_load(const char* const* sv, uint32_t c, uint32_t st, int idx, int n, const T* u0,
                                          T* um2, T* um1, T* up1, T* up2, bool minus = true) {
  const bool have = n > 2;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    um2[v] = um1[v] = up1[v] = up2[v] = u0[v];
  }
  if (have && idx >= 2 && (!MAYSKIP || minus)) {
#pragma unroll
    for (int v = 0; v < 5; ++v) um2[v] = ldv<T>(sv[v], c - 2 * st);
  }
  if ((NEAR || have) && idx >= 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) um1[v] = ldv<T>(sv[v], c - st);
  }
  if ((NEAR || have) && idx + 1 < n) {
#pragma unroll
    for (int v = 0; v < 5; ++v) up1[v] = ldv<T>(sv[v], c + st);
  }
  if (have && idx + 2 < n) {
#pragma unroll
    for (int v = 0; v < 5; ++v) up2[v] = ldv<T>(sv[v], c + 2 * st);
  }
}

struct no_visc {
  template <typename T>
  __device__ __forceinline__ void operator()(T*) const {}
};

template <typename T, int RECON, int LIM, bool FAST, int DIR, bool FLG = false, bool MAYSKIP = false,
          class VM = no_visc, class VP = no_visc>
__device__ __forceinline__ void axis_eval(const T* um2, const T* um1, const T* up1, const T* up2, int idx, int n,
                                          const T* u0, T A, T* res, T* fp, uint32_t ins = 0, bool minus = true,
                                          VM vm = VM(), VP vp = VP()) {
  const bool have = n > 2;
  zero5(fp);
  if constexpr (FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD)) {
    // Fast arithmetic: the limited slope term of a cell, lim(dL, dR)*dL with
    // dL = U_c - U_c-1, dR = U_c+1 - U_c, is symmetric in (dL, dR) for van
    // Albada and minmod, and the plus-side state of a face uses its negative
    // (upwind2.h:44 with the cell as i1 and the face neighbour as i2).  So the
    // three slopes of the cells c-1, c, c+1 serve both faces of the axis.
    if (!have) return;
    T sm[5], s0[5], sp[5];
    bool okm = idx >= 2, ok0 = idx >= 1 && idx + 1 < n, okp = idx + 2 < n;
    // with split faces the cell's own slope may enter one of its faces only
    bool ok0m = true, ok0p = true;
    if constexpr (FLG) {
      okm = okm && !split_bits(ins, 0, 1);
      okp = okp && !split_bits(ins, 4, 3);
      ok0m = !split_bits(ins, 3, 2);
      ok0p = !split_bits(ins, 1, 2);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if constexpr (!MAYSKIP) sm[v] = okm ? limited<T, LIM, true>(um1[v] - um2[v], u0[v] - um1[v]) : T(0);
      s0[v] = ok0 ? limited<T, LIM, true>(u0[v] - um1[v], up1[v] - u0[v]) : T(0);
      sp[v] = okp ? limited<T, LIM, true>(up1[v] - u0[v], up2[v] - up1[v]) : T(0);
    }
    if (idx > 0 && minus) {
      T vl[5], vr[5], f[5];
      if constexpr (MAYSKIP) {
#pragma unroll
        for (int v = 0; v < 5; ++v) sm[v] = okm ? limited<T, LIM, true>(um1[v] - um2[v], u0[v] - um1[v]) : T(0);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = um1[v] + T(0.5) * sm[v];
        vr[v] = u0[v] - T(0.5) * (ok0m ? s0[v] : T(0));
      }
      zero5(f);
      ausm_flux_set<T, FAST, DIR>(vl, vr, A, f);
      vm(f);
#pragma unroll
      for (int v = 0; v < 5; ++v) res[v] += f[v];
    }
    if (idx < n - 1) {
      T vl[5], vr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = u0[v] + T(0.5) * (ok0p ? s0[v] : T(0));
        vr[v] = up1[v] - T(0.5) * sp[v];
      }
      ausm_flux_set<T, FAST, DIR>(vl, vr, A, fp);  // fp was zeroed above
      vp(fp);
    }
  } else {
    if (have && idx > 0 && minus) {
      T f[5];
      zero5(f);
      face_flux_r<T, RECON, LIM, FAST, DIR>(um2, um1, u0, up1, idx >= 2 && !(FLG && split_bits(ins, 0, 1)),
                                            idx + 1 < n && !(FLG && split_bits(ins, 3, 2)), A, f);
      vm(f);
#pragma unroll
      for (int v = 0; v < 5; ++v) res[v] += f[v];
    }
    if (have && idx < n - 1) {
      face_flux_r<T, RECON, LIM, FAST, DIR>(um1, u0, up1, up2, idx >= 1 && !(FLG && split_bits(ins, 1, 2)),
                                            idx + 2 < n && !(FLG && split_bits(ins, 4, 3)), A, fp);
      vp(fp);
    }
  }
}

template <typename T, int RECON, int LIM, bool FAST, int DIR, bool FLG = false, bool MAYSKIP = false>
__device__ __forceinline__ void axis_faces_r(const char* const* sv, uint32_t c, uint32_t st, int idx, int n,
                                             const T* u0, T A, T* res, T* fp, uint32_t ins = 0, bool minus = true) {
  T um2[5], um1[5], up1[5], up2[5];
  axis_load<T, MAYSKIP>(sv, c, st, idx, n, u0, um2, um1, up1, up2, minus);
  axis_eval<T, RECON, LIM, FAST, DIR, FLG, MAYSKIP>(um2, um1, up1, up2, idx, n, u0, A, res, fp, ins, minus);
}

// ZS = false: one thread per cell, BS cells per work item.
// ZS = true:  each wavefront owns 63 consecutive cells of the (j,k) plane (a work
//             item covers 252); lane 63 only evaluates the z face k-1/2 of the
//             next cell.  Every lane's z face k-1/2 is the plus z face of the
//             lane below it, handed over by a lane shuffle, so z faces are
//             evaluated once: 5 face evaluations per cell instead of 6.  The z
//             neighbours themselves are the neighbouring lanes' cells, also
//             taken by lane shuffles (15 global loads fewer per cell).
// FLG: per-cell flags (DRN_CELL_INACTIVE = 1, DRN_CELL_INSIDE = 2) at flg[global
// cell]: split faces reconstruct first order, inactive cells keep their state.
// the fast symmetric-slope form of axis_faces_r from the four neighbours'
// values already loaded (um2, um1, up1, up2: the cells at idx-2, idx-1, idx+1,
// idx+2, or the nearest cell of the patch where one does not exist; the slopes
// that involve a missing cell are selected to zero as in axis_faces_r)
template <typename T, int LIM, int DIR>
__device__ __forceinline__ void axis_faces_v(const T* um2, const T* um1, const T* up1, const T* up2, int idx, int n,
                                             const T* u0, T A, T* res, T* fp) {
  zero5(fp);
  if (n <= 2) return;
  const bool okm = idx >= 2, ok0 = idx >= 1 && idx + 1 < n, okp = idx + 2 < n;
  T sm[5], s0[5], sp[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    sm[v] = okm ? limited<T, LIM, true>(um1[v] - um2[v], u0[v] - um1[v]) : T(0);
    s0[v] = ok0 ? limited<T, LIM, true>(u0[v] - um1[v], up1[v] - u0[v]) : T(0);
    sp[v] = okp ? limited<T, LIM, true>(up1[v] - u0[v], up2[v] - up1[v]) : T(0);
  }
  if (idx > 0) {
    T vl[5], vr[5], f[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      vl[v] = um1[v] + T(0.5) * sm[v];
      vr[v] = u0[v] - T(0.5) * s0[v];
    }
    zero5(f);
    ausm_flux_set<T, true, DIR>(vl, vr, A, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (idx < n - 1) {
    T vl[5], vr[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      vl[v] = u0[v] + T(0.5) * s0[v];
      vr[v] = up1[v] - T(0.5) * sp[v];
    }
    ausm_flux_set<T, true, DIR>(vl, vr, A, fp);  // fp was zeroed above
  }
}

// XP (z-shared form only): planes per work item.  The four wavefronts of an
// item cover 4/XP x 63 cells of each of XP consecutive planes i0 .. i0+XP-1 at
// the same (j,k) cells; a wavefront above the first plane takes its x face
// i-1/2 from the wavefront below, which evaluated it as its plus face and left
// it in xbuf (LDS, 4 x 5 x 64 values), so a cell evaluates 5.02 - (XP-1)/XP
// faces.  The face is the same function of the same four cells either way.
template <typename T, int RECON, int LIM, bool FAST, int BS, bool ZS, bool VISC = false, bool FLG = false,
          int XP = 1>
__device__ __forceinline__ void stage_cell_item(const PatchDev* __restrict__ pds, const TileCell t,
                                                const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                T factor, int push, const uint8_t* __restrict__ flg = nullptr,
                                                T* xbuf = nullptr) {
  static_assert(XP == 1 || (ZS && !VISC && !FLG && (XP == 2 || XP == 4)), "planes per item: 1, 2 or 4");
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int lane = (int)(threadIdx.x & 63);
  constexpr int WPP = 4 / XP;  // wavefronts per plane
  const int wv = (int)(threadIdx.x >> 6), pw = wv / WPP;
  const int c2 = ZS ? t.c2 + 63 * (wv % WPP) + lane : t.c2 + (int)threadIdx.x;
  if constexpr (!ZS) {
    if (c2 >= njk) return;
  }
  const int i = t.i + (XP > 1 ? pw : 0);
  const bool in_plane = c2 < njk && (XP == 1 || i < ni);  // a cell of the patch
  const bool own = in_plane && (!ZS || lane < 63);        // this lane updates its cell
  const int ii = XP > 1 && i >= ni ? ni - 1 : i;          // a plane of the patch for the addresses
  const int cc2 = c2 < njk ? c2 : njk - 1;
  const int j = cc2 / nk;
  const int k = cc2 - j * nk;
  // byte offsets within one variable array (the layout keeps them below 2^32)
  constexpr uint32_t W = sizeof(T);
  const uint32_t c = (P.off + (uint32_t)ii * (uint32_t)njk + (uint32_t)cc2) * W;
  const uint32_t sx = (uint32_t)njk * W, sy = (uint32_t)nk * W;
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  T res[5], f[5];
  zero5(res);
  // Residual in the reference's order Fx- + Fy- + Fz- - Fz+ - Fy+ - Fx+: each
  // axis loads its four neighbours once, adds its minus face and keeps its
  // plus face until the end.
  T u0[5], fxp[5], fyp[5], fzp[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u0[v] = ldv<T>(sv[v], c);
  static_assert(!VISC || ZS, "the viscous flux runs the z-shared cell kernel");
  // viscous flux: (real)0.5/delta of the cell gradients, 1/delta of the face
  // gradients.  FP32 shares the cell gradients between faces and lanes (VSH);
  // FP64 loads them per face, which needs fewer registers (3 waves per SIMD
  // instead of 2, measured 4.5 % faster).  VSH keeps the cell's own x and y
  // gradients for its z face.
  constexpr bool VSH = VISC && sizeof(T) == 4;
  const T dhx = T(0.5) / dx, dhy = T(0.5) / dy, dhz = T(0.5) / dz;
  const T dnx = T(1.0) / dx, dny = T(1.0) / dy, dnz = T(1.0) / dz;
  const uint32_t gst[3] = {sx, sy, W};
  const T gdh[3] = {dhx, dhy, dhz};
  const int gidx[3] = {i, j, k}, gn[3] = {ni, nj, nk};
  T gxc[5], gyc[5];
  // the cell's flags, and the inside bits of the x and y neighbours
  uint32_t fo = 0, insx = 0, insy = 0;
  const uint32_t gc = P.off + (uint32_t)ii * (uint32_t)njk + (uint32_t)cc2;  // the cell's global index
  if constexpr (FLG) {
    fo = flg[gc];
    auto ins_axis = [&](uint32_t stc, int idx, int n) {
      uint32_t b = ((fo >> 1) & 1u) << 2;
      if (idx >= 2) b |= ((uint32_t)(flg[gc - 2 * stc] >> 1) & 1u) << 0;
      if (idx >= 1) b |= ((uint32_t)(flg[gc - stc] >> 1) & 1u) << 1;
      if (idx + 1 < n) b |= ((uint32_t)(flg[gc + stc] >> 1) & 1u) << 3;
      if (idx + 2 < n) b |= ((uint32_t)(flg[gc + 2 * stc] >> 1) & 1u) << 4;
      return b;
    };
    if (own) {
      insx = ins_axis((uint32_t)njk, i, ni);
      insy = ins_axis((uint32_t)nk, j, nj);
    }
  }
  if constexpr (XP > 1) {
    if (own) axis_faces_r<T, RECON, LIM, FAST, 0, false, true>(sv, c, sx, i, ni, u0, Ax, res, fxp, 0, pw == 0);
    // fast symmetric form: the y neighbours are loaded before the barrier, so
    // that their latency passes while the wavefronts wait for each other (a
    // missing neighbour is replaced by the nearest cell of the patch)
    constexpr bool YPRE = FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD);
    T ym[4][5];
    if constexpr (YPRE) {
      if (own) {
        const uint32_t om1 = j >= 1 ? c - sy : c, op1 = j + 1 < nj ? c + sy : c;
        const uint32_t om2 = j >= 2 ? c - 2 * sy : om1, op2 = j + 2 < nj ? c + 2 * sy : op1;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          ym[0][v] = ldv<T>(sv[v], om2);
          ym[1][v] = ldv<T>(sv[v], om1);
          ym[2][v] = ldv<T>(sv[v], op1);
          ym[3][v] = ldv<T>(sv[v], op2);
        }
      }
    }
    if (own) {
#pragma unroll
      for (int v = 0; v < 5; ++v) xbuf[(wv * 5 + v) * 64 + lane] = fxp[v];
    }
    __syncthreads();
    if (own && pw > 0 && i > 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v) res[v] += xbuf[((wv - WPP) * 5 + v) * 64 + lane];
    }
    if constexpr (YPRE) {
      if (own) axis_faces_v<T, LIM, 1>(ym[0], ym[1], ym[2], ym[3], j, nj, u0, Ay, res, fyp);
    } else {
      if (own) axis_faces_r<T, RECON, LIM, FAST, 1>(sv, c, sy, j, nj, u0, Ay, res, fyp);
    }
  } else if constexpr (VSH) {
    // Viscous faces (CompressibleViscFlux after AUSM+): a face takes the
    // tangential cell gradients of its two cells (cell_grad_v).  The cell's own
    // gradients serve all its faces.  The z differences of its x and y
    // neighbours come from the neighbouring lanes, which hold those neighbours
    // at k -+ 1.  The four cells (i -+ 1, j -+ 1) give the y gradients of the x
    // neighbours and the x gradients of the y neighbours.  Every lane loads its
    // neighbours (lane 63's feed lane 62); lane 0 loads the cells at k - 1.
    T xm2[5], xm1[5], xp1[5], xp2[5], ym1[5], yp1[5], gzc[5];
    {
      T ym2[5], yp2[5];
      axis_load<T, false, true>(sv, c, sy, j, nj, u0, ym2, ym1, yp1, yp2);
    }
    axis_load<T, false, true>(sv, c, sx, i, ni, u0, xm2, xm1, xp1, xp2);
    cell_grad_v<T, 0>(xm1, xp1, dhx, i, ni, gxc);
    cell_grad_v<T, 1>(ym1, yp1, dhy, j, nj, gyc);
    // z gradient of the cell in this lane whose value is a: the values at
    // k -+ 1 from lanes L -+ 1 (at k - 1 lane 0 loads it from byte offset
    // off - W when ok0), the cell itself at the patch ends
    auto zgrad = [&](const T* a, uint32_t off, bool ok0, T* g) {
      T lo[5], hi[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        lo[v] = __shfl_up(a[v], 1, 64);
        hi[v] = __shfl_down(a[v], 1, 64);
      }
      if (lane == 0 && in_plane && k >= 1 && ok0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) lo[v] = ldv<T>(sv[v], off - W);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (k < 1) lo[v] = a[v];
        if (k + 1 >= nk) hi[v] = a[v];
      }
      cell_grad_v<T, 2>(lo, hi, dhz, k, nk, g);
    };
    zgrad(u0, c, true, gzc);
    // y gradient of the cell i + di (x neighbour a), or x gradient of the cell
    // j + dj (y neighbour a), from the cells (i + di, j -+ 1) / (i -+ 1, j + dj):
    // the neighbour itself stands in where one does not exist
    auto diag_grad = [&](const T* a, uint32_t off, bool ok, uint32_t st, int idx, int n, T dh, bool ygrad, T* g) {
      T lo[5], hi[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) lo[v] = hi[v] = a[v];
      if (ok && idx >= 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) lo[v] = ldv<T>(sv[v], off - st);
      }
      if (ok && idx + 1 < n) {
#pragma unroll
        for (int v = 0; v < 5; ++v) hi[v] = ldv<T>(sv[v], off + st);
      }
      if (ygrad)
        cell_grad_v<T, 1>(lo, hi, dh, idx, n, g);
      else
        cell_grad_v<T, 0>(lo, hi, dh, idx, n, g);
    };
    {
      // the x faces: z gradients of the cells i -+ 1 (shuffles), their y
      // gradients where a face needs them
      T gzm[5], gzp[5];
      zgrad(xm1, c - sx, i >= 1, gzm);
      zgrad(xp1, c + sx, i + 1 < ni, gzp);
      if (own)
        axis_eval<T, RECON, LIM, FAST, 0, FLG>(
            xm2, xm1, xp1, xp2, i, ni, u0, Ax, res, fxp, insx, true,
            [&](T* fl) {
              T g[5];
              diag_grad(xm1, c - sx, i >= 1, sy, j, nj, dhy, true, g);
              visc_face<T, FAST, 0>(xm1, u0, g, gyc, gzm, gzc, dnx, Ax, fl);
            },
            [&](T* fl) {
              T g[5];
              diag_grad(xp1, c + sx, i + 1 < ni, sy, j, nj, dhy, true, g);
              visc_face<T, FAST, 0>(u0, xp1, gyc, g, gzc, gzp, dnx, Ax, fl);
            });
    }
    {
      // the y faces: z gradients of the cells j -+ 1, their x gradients
      T ym2[5], yp2[5], gzm[5], gzp[5];
      zgrad(ym1, c - sy, j >= 1, gzm);
      zgrad(yp1, c + sy, j + 1 < nj, gzp);
      {
        T t1[5], t2[5];
        axis_load<T, false, true>(sv, c, sy, j, nj, u0, ym2, t1, t2, yp2);
      }
      if (own)
        axis_eval<T, RECON, LIM, FAST, 1, FLG>(
            ym2, ym1, yp1, yp2, j, nj, u0, Ay, res, fyp, insy, true,
            [&](T* fl) {
              T g[5];
              diag_grad(ym1, c - sy, j >= 1, sx, i, ni, dhx, false, g);
              visc_face<T, FAST, 1>(ym1, u0, g, gxc, gzm, gzc, dny, Ay, fl);
            },
            [&](T* fl) {
              T g[5];
              diag_grad(yp1, c + sy, j + 1 < nj, sx, i, ni, dhx, false, g);
              visc_face<T, FAST, 1>(u0, yp1, gxc, g, gzc, gzp, dny, Ay, fl);
            });
    }
  } else if (own) {
    if constexpr (VISC) {
      T m2[5], m1[5], p1[5], p2[5];
      axis_load<T>(sv, c, sx, i, ni, u0, m2, m1, p1, p2);
      axis_eval<T, RECON, LIM, FAST, 0, FLG>(
          m2, m1, p1, p2, i, ni, u0, Ax, res, fxp, insx, true,
          [&](T* fl) { visc_face_ld<T, FAST, 0>(sv, c - sx, c, m1, u0, gst, gdh, gidx, gn, dnx, Ax, fl); },
          [&](T* fl) { visc_face_ld<T, FAST, 0>(sv, c, c + sx, u0, p1, gst, gdh, gidx, gn, dnx, Ax, fl); });
      axis_load<T>(sv, c, sy, j, nj, u0, m2, m1, p1, p2);
      axis_eval<T, RECON, LIM, FAST, 1, FLG>(
          m2, m1, p1, p2, j, nj, u0, Ay, res, fyp, insy, true,
          [&](T* fl) { visc_face_ld<T, FAST, 1>(sv, c - sy, c, m1, u0, gst, gdh, gidx, gn, dny, Ay, fl); },
          [&](T* fl) { visc_face_ld<T, FAST, 1>(sv, c, c + sy, u0, p1, gst, gdh, gidx, gn, dny, Ay, fl); });
    } else {
      axis_faces_r<T, RECON, LIM, FAST, 0, FLG>(sv, c, sx, i, ni, u0, Ax, res, fxp, insx);
      axis_faces_r<T, RECON, LIM, FAST, 1, FLG>(sv, c, sy, j, nj, u0, Ay, res, fyp, insy);
    }
  }
  if constexpr (ZS) {
    // z face k-1/2 of every lane's cell (lane 63 included), then the plus face
    // from the lane above when it lies in the same row
    T fzm[5];
    zero5(fzm);
    // the z neighbours k-1, k-2, k+1 are the cells of lanes L-1, L-2, L+1 (the
    // wavefront holds consecutive cells of the flattened (j,k) plane); lanes
    // 0, 1 and 63 load the ones outside the wavefront.  face_flux_r reads k-2
    // and k+1 only where they exist.
    T um2[5], um1[5], up1[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      um1[v] = __shfl_up(u0[v], 1, 64);
      um2[v] = __shfl_up(u0[v], 2, 64);
      up1[v] = __shfl_down(u0[v], 1, 64);
    }
    if (lane == 0 && in_plane && k >= 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) um1[v] = ldv<T>(sv[v], c - W);
    }
    if (lane <= 1 && in_plane && k >= 2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) um2[v] = ldv<T>(sv[v], c - 2 * W);
    }
    if (lane == 63 && in_plane && k + 1 < nk) {
#pragma unroll
      for (int v = 0; v < 5; ++v) up1[v] = ldv<T>(sv[v], c + W);
    }
    bool zm_ok = k >= 2, zp_ok = k + 1 < nk;
    if constexpr (FLG) {
      // inside bits of k-2, k-1, k+1 from the neighbouring lanes (edge lanes load them)
      const uint32_t io = (fo >> 1) & 1u;
      uint32_t im1 = (uint32_t)__shfl_up((int)io, 1, 64), im2 = (uint32_t)__shfl_up((int)io, 2, 64);
      uint32_t ip1 = (uint32_t)__shfl_down((int)io, 1, 64);
      if (lane == 0 && in_plane && k >= 1) im1 = (flg[gc - 1] >> 1) & 1u;
      if (lane <= 1 && in_plane && k >= 2) im2 = (flg[gc - 2] >> 1) & 1u;
      if (lane == 63 && in_plane && k + 1 < nk) ip1 = (flg[gc + 1] >> 1) & 1u;
      zm_ok = zm_ok && im2 == im1;
      zp_ok = zp_ok && ip1 == io;
    }
    if (in_plane && nk > 2 && k >= 1) {
      face_flux_r<T, RECON, LIM, FAST, 2>(um2, um1, u0, up1, zm_ok, zp_ok, Az, fzm);
    }
    if constexpr (VISC && !VSH) {
      if (in_plane && nk > 2 && k >= 1)
        visc_face_ld<T, FAST, 2>(sv, c - W, c, um1, u0, gst, gdh, gidx, gn, dnz, Az, fzm);
    } else if constexpr (VSH) {
      // x and y gradients of the cell k-1: lane L-1's own, lane 0 loads its
      // neighbours (the cell k-1 itself stands in at the patch ends)
      T gxm[5], gym[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        gxm[v] = __shfl_up(gxc[v], 1, 64);
        gym[v] = __shfl_up(gyc[v], 1, 64);
      }
      if (lane == 0 && in_plane && k >= 1) {
        T lo[5], hi[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          lo[v] = ldv<T>(sv[v], i >= 1 ? c - W - sx : c - W);
          hi[v] = ldv<T>(sv[v], i + 1 < ni ? c - W + sx : c - W);
        }
        cell_grad_v<T, 0>(lo, hi, dhx, i, ni, gxm);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          lo[v] = ldv<T>(sv[v], j >= 1 ? c - W - sy : c - W);
          hi[v] = ldv<T>(sv[v], j + 1 < nj ? c - W + sy : c - W);
        }
        cell_grad_v<T, 1>(lo, hi, dhy, j, nj, gym);
      }
      if (in_plane && nk > 2 && k >= 1) visc_face<T, FAST, 2>(um1, u0, gxm, gxc, gym, gyc, dnz, Az, fzm);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      res[v] += fzm[v];
      const T up = __shfl_down(fzm[v], 1, 64);
      fzp[v] = (nk > 2 && k + 1 < nk) ? up : T(0);
    }
    if (!own) return;
  } else {
    static_assert(!FLG, "cell flags run the z-shared cell kernel");
    axis_faces_r<T, RECON, LIM, FAST, 2>(sv, c, W, k, nk, u0, Az, res, fzp);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    res[v] -= fzp[v];
    res[v] -= fyp[v];
    res[v] -= fxp[v];
  }
  // patch-face functors (cartesianiterator.h:190-308); faces marked BC_DEAD
  // (all boundary cells exchange receivers, in the RK step) are skipped
  const bool lx0 = P.bc[0] != BC_DEAD, lx1 = P.bc[1] != BC_DEAD, ly0 = P.bc[2] != BC_DEAD,
             ly1 = P.bc[3] != BC_DEAD, lz0 = P.bc[4] != BC_DEAD, lz1 = P.bc[5] != BC_DEAD;
  const bool wx = ni > 2 && ((i == 0 && lx0) || (i == ni - 1 && lx1));
  const bool wy = nj > 2 && ((j == 0 && ly0) || (j == nj - 1 && ly1));
  const bool wz = nk > 2 && ((k == 0 && lz0) || (k == nk - 1 && lz1));
  if (wx || wy || wz) {
    const T* var = u0;
    const T xc = T(0.5) * dx + T(i) * dx, yc = T(0.5) * dy + T(j) * dy, zc = T(0.5) * dz + T(k) * dz;
    if (ni > 2) {
      if (i == 0 && lx0) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[0], P.par[0], var, 0, false, yc, zc, j, k, Ax, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] += f[v];
      }
      if (i == ni - 1 && lx1) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[1], P.par[1], var, 0, true, yc, zc, j, k, Ax, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] -= f[v];
      }
    }
    if (nj > 2) {
      if (j == 0 && ly0) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[2], P.par[2], var, 1, false, xc, zc, i, k, Ay, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] += f[v];
      }
      if (j == nj - 1 && ly1) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[3], P.par[3], var, 1, true, xc, zc, i, k, Ay, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] -= f[v];
      }
    }
    if (nk > 2) {
      if (k == 0 && lz0) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[4], P.par[4], var, 2, false, xc, yc, i, j, Az, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] += f[v];
      }
      if (k == nk - 1 && lz1) {
        zero5(f);
        wall_flux<T, FAST>(P.bc[5], P.par[5], var, 2, true, xc, yc, i, j, Az, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) res[v] -= f[v];
      }
    }
  }
  // cartesianiterator.h:310-322
  const T pf = factor / (dx * dy * dz);
  T nv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
    nv[v] = ldv<T>(bv, c) + pf * res[v];
  }
  if constexpr (FLG) {
    // an inactive cell keeps its state (Patch::isActive, cartesianiterator.h:315)
    if (fo & 1u) {
#pragma unroll
      for (int v = 0; v < 5; ++v) nv[v] = u0[v];
    }
  }
  bool store = true;
  if (push) {
    // the exchange of the z-face seek layers fused into the stage: a receiver
    // keeps the value its donor writes, a donor also writes 0 + value (the
    // exchange's unit-weight sum, which turns -0 into +0) into its receiver
    for (int s = 0; s < P.nrecv; ++s) store = store && !in_box(P.recv[s], i, j, k);
    for (int s = 0; s < P.npush; ++s) {
      if (in_box(P.push[s], i, j, k)) {
        // a receiver cell of this layout, whose byte offset is below 2^32: the
        // 32-bit sum, exact modulo 2^32, is that offset
        const uint32_t tb =
            ((uint32_t)P.push_base[s] + (uint32_t)i * P.push_si[s] + (uint32_t)j * P.push_sj[s] + (uint32_t)k) * W;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          *reinterpret_cast<T*>(reinterpret_cast<char*>(dst + (uint64_t)v * vs) + tb) = T(0) + nv[v];
      }
    }
  }
  if (store) {
#pragma unroll
    for (int v = 0; v < 5; ++v) *reinterpret_cast<T*>(reinterpret_cast<char*>(dst + (uint64_t)v * vs) + c) = nv[v];
  }
}

template <typename T, int RECON, int LIM, bool FAST, int BS, bool ZS = false, bool VISC = false>
__global__ __launch_bounds__(BS) void k_stage_cell(const PatchDev* __restrict__ pds, const TileCell* __restrict__ tiles,
                                                   const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                   T factor, int push) {
  stage_cell_item<T, RECON, LIM, FAST, BS, ZS, VISC>(pds, tiles[blockIdx.x], src, base, dst, vs, factor, push);
}

// the z-shared cell kernel over XP planes per work item, x faces shared between
// the planes through LDS
template <typename T, int RECON, int LIM, bool FAST, int XP>
__global__ __launch_bounds__(256) void k_stage_cell_xs(const PatchDev* __restrict__ pds,
                                                       const TileCell* __restrict__ tiles, const T* __restrict__ src,
                                                       const T* base, T* dst, uint64_t vs, T factor, int push) {
  __shared__ T xbuf[4 * 5 * 64];
  stage_cell_item<T, RECON, LIM, FAST, 256, true, false, false, XP>(pds, tiles[blockIdx.x], src, base, dst, vs,
                                                                    factor, push, nullptr, xbuf);
}

// the z-shared cell kernel with per-cell flags (inactive cells, split faces)
template <typename T, int RECON, int LIM, bool FAST, bool VISC>
__global__ __launch_bounds__(256) void k_stage_cell_flags(const PatchDev* __restrict__ pds,
                                                          const TileCell* __restrict__ tiles, const T* __restrict__ src,
                                                          const T* base, T* dst, uint64_t vs, T factor, int push,
                                                          const uint8_t* __restrict__ flg) {
  stage_cell_item<T, RECON, LIM, FAST, 256, true, VISC, true>(pds, tiles[blockIdx.x], src, base, dst, vs, factor,
                                                              push, flg);
}

// The FP32 viscous form of k_stage_cell<..., ZS = true>, with the cell
// gradients shared between faces and lanes: 111 VGPRs, 4 waves per SIMD.
// Capped at 5 waves it spills and runs 28 % slower on the viscous shear layer.
// (The per-face form it replaced ran best capped at 5 waves, and the shared
// form runs 5 % faster than that.)  The FP64 form keeps k_stage_cell.
template <int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_cell_visc_f32(
    const PatchDev* __restrict__ pds, const TileCell* __restrict__ tiles, const float* __restrict__ src,
    const float* base, float* dst, uint64_t vs, float factor, int push) {
  stage_cell_item<float, RECON, LIM, FAST, 256, true, true>(pds, tiles[blockIdx.x], src, base, dst, vs, factor, push);
}

// Persistent form of k_stage_cell<..., ZS = true>: a grid sized to the resident
// capacity of the device walks the work list with a stride of gridDim.x
// (a multiple of 8, so each workgroup stays on one XCD and keeps taking items
// of that XCD's segment).  Every workgroup ends after its last item.
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_cell_persist(const PatchDev* __restrict__ pds,
                                                            const TileCell* __restrict__ tiles, int n_items,
                                                            const T* __restrict__ src, const T* base, T* dst,
                                                            uint64_t vs, T factor, int push) {
  for (int it = (int)blockIdx.x; it < n_items; it += (int)gridDim.x)
    stage_cell_item<T, RECON, LIM, FAST, 256, true>(pds, tiles[it], src, base, dst, vs, factor, push);
}

// ---------------------------------------------------------------------------
// i-marching z-shared cell kernel
// ---------------------------------------------------------------------------

// As k_stage_cell<..., ZS = true> (63 cells of the flattened (j,k) plane per
// wavefront, z neighbours and z faces handed between lanes), but a work item
// marches its 252 (j,k) cells over the planes [i0, i1) of the patch.  Each x
// face i+1/2 is evaluated once, by the cell i: the next plane's residual
// starts from it, and the MUSCL data of the face's left side (the limited
// slope of cell i in the fast symmetric form, else the reconstructed left
// state) is carried in registers: 4 face evaluations per cell instead of 5.
// The per-lane work of the (j,k) cell (division by NK, the (j,k) part of the
// fused-exchange boxes, the wall classification) is done once per item.
//
// The patch-face functors of a cell (cartesianiterator.h:190-308) are summed
// first, in the reference's order xM, xP, yM, yP, zM, zP, while only the cell
// and the carried x data are live; the residual is then that sum + Fx- + Fy-
// + Fz- - Fz+ - Fy+ - Fx+.  The reference adds the functors after the faces:
// for a cell on one patch face the two orders add the same terms in another
// association only.  Requires NI, NJ, NK >= 3 (thinner patches run
// k_stage_cell).
// LC: the values carried from plane to plane besides U(i) -- the x slope (or
// left state) xs and the x face flux fxm -- live in LDS, `carry[10][256]`
// (variable-major, each lane its own column: no barrier, no bank conflict),
// instead of 10 VGPRs; the kernel then fits 64 VGPRs (8 waves per SIMD).
template <typename T, int RECON, int LIM, bool FAST, bool LC = false>
__device__ __forceinline__ void stage_xm_item(const PatchDev* __restrict__ pds, const TileMarch* __restrict__ items,
                                              const T* __restrict__ src, const T* base, T* dst, uint64_t vs, T factor,
                                              int push, T* carry = nullptr) {
  // the fast symmetric form carries the slope of cell i, the general form the
  // left MUSCL state of face i+1/2
  constexpr bool SYM = FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD);
  const TileMarch t = items[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int lane = (int)(threadIdx.x & 63);
  const int c2 = t.c0 + 63 * (int)(threadIdx.x >> 6) + lane;
  const bool in_plane = c2 < njk;
  const bool own = in_plane && lane < 63;
  const int cc2 = in_plane ? c2 : njk - 1;
  const int j = cc2 / nk;
  const int k = cc2 - j * nk;
  constexpr uint32_t W = sizeof(T);
  const uint32_t sx = (uint32_t)njk * W, sy = (uint32_t)nk * W;
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);
  const int i0 = t.i0, i1 = t.i1;
  // fused exchange: the (j,k) part of the boxes per lane, the i part per plane
  auto jk_in = [&](const CellBox& b) {
    return (uint32_t)(j - b.j0) < (uint32_t)(b.j1 - b.j0) && (uint32_t)(k - b.k0) < (uint32_t)(b.k1 - b.k0);
  };
  const bool pjk0 = push && P.npush > 0 && jk_in(P.push[0]), pjk1 = push && P.npush > 1 && jk_in(P.push[1]);
  const bool rjk0 = push && P.nrecv > 0 && jk_in(P.recv[0]), rjk1 = push && P.nrecv > 1 && jk_in(P.recv[1]);
  // boundary cells of live faces (BC_DEAD faces are skipped, see stage_cell_item)
  const bool wyz = own && ((j == 0 && P.bc[2] != BC_DEAD) || (j == nj - 1 && P.bc[3] != BC_DEAD) ||
                           (k == 0 && P.bc[4] != BC_DEAD) || (k == nk - 1 && P.bc[5] != BC_DEAD));
  const bool lx0 = P.bc[0] != BC_DEAD, lx1 = P.bc[1] != BC_DEAD;

  // x data entering plane i: u0 = U(i), xs = slope of cell i (SYM) or the
  // left state of face i+1/2, fxm = F(i-1/2) (zero at i = 0)
  uint32_t c = (P.off + (uint32_t)i0 * (uint32_t)njk + (uint32_t)cc2) * W;
  T u0[5], xs[5], fxm[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u0[v] = ldv<T>(sv[v], c);
  zero5(fxm);
  {
    // U(i0-2), U(i0-1), U(i0+1): the slope of cell i0 and the face i0-1/2
    T um1[5], um2[5], up1[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) um1[v] = um2[v] = up1[v] = u0[v];
    if (i0 >= 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) um1[v] = ldv<T>(sv[v], c - sx);
    }
    if (i0 >= 2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) um2[v] = ldv<T>(sv[v], c - 2 * sx);
    }
    if (i0 + 1 < ni) {
#pragma unroll
      for (int v = 0; v < 5; ++v) up1[v] = ldv<T>(sv[v], c + sx);
    }
    if constexpr (SYM) {
      const bool okm = i0 >= 2, ok0 = i0 >= 1 && i0 + 1 < ni;
#pragma unroll
      for (int v = 0; v < 5; ++v) xs[v] = ok0 ? limited<T, LIM, true>(u0[v] - um1[v], up1[v] - u0[v]) : T(0);
      if (i0 >= 1) {
        T vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T sm = okm ? limited<T, LIM, true>(um1[v] - um2[v], u0[v] - um1[v]) : T(0);
          vl[v] = um1[v] + T(0.5) * sm;
          vr[v] = u0[v] - T(0.5) * xs[v];
        }
        ausm_flux_set<T, FAST, 0>(vl, vr, Ax, fxm);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) xs[v] = u0[v];
      if constexpr (RECON == RECON_UPWIND2) {
        if (i0 >= 1 && i0 + 1 < ni) {
#pragma unroll
          for (int v = 0; v < 5; ++v) xs[v] = u0[v] + T(0.5) * limited<T, LIM, FAST>(u0[v] - um1[v], up1[v] - u0[v]);
        }
      }
      if (i0 >= 1) face_flux_r<T, RECON, LIM, FAST, 0>(um2, um1, u0, up1, i0 >= 2, i0 + 1 < ni, Ax, fxm);
    }
  }
  const int tid = (int)threadIdx.x;
  if constexpr (LC) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      carry[v * 256 + tid] = xs[v];
      carry[(5 + v) * 256 + tid] = fxm[v];
    }
  }

  for (int i = i0; i < i1; ++i, c += sx) {
    T res[5];
    zero5(res);
    // patch-face functors: one call site, the faces in the reference's order
    if (wyz || (i == 0 && lx0) || (i == ni - 1 && lx1)) {
#pragma unroll 1
      for (int f = 0; f < 6; ++f) {
        const int d = f >> 1;
        const bool plus = (f & 1) != 0;
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int n = d == 0 ? ni : (d == 1 ? nj : nk);
        if (P.bc[f] != BC_DEAD && own && idx == (plus ? n - 1 : 0)) {
          const T xc = T(0.5) * dx + T(i) * dx, yc = T(0.5) * dy + T(j) * dy, zc = T(0.5) * dz + T(k) * dz;
          const T ta = d == 0 ? yc : xc, tb = d == 2 ? yc : zc;
          const int la = d == 0 ? j : i, lb = d == 2 ? j : k;
          const T A = d == 0 ? Ax : (d == 1 ? Ay : Az);
          T g[5];
          zero5(g);
          wall_flux<T, FAST>(P.bc[f], P.par[f], u0, d, plus, ta, tb, la, lb, A, g);
#pragma unroll
          for (int v = 0; v < 5; ++v) res[v] = plus ? res[v] - g[v] : res[v] + g[v];
        }
      }
    }
    if constexpr (LC) {
#pragma unroll
      for (int v = 0; v < 5; ++v) res[v] += carry[(5 + v) * 256 + tid];
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) res[v] += fxm[v];
    }
    T fyp[5], fzp[5], f[5];
    if (own) axis_faces_r<T, RECON, LIM, FAST, 1>(sv, c, sy, j, nj, u0, Ay, res, fyp);
    {
      // z faces by lane shuffles, as in stage_cell_item
      T fzm[5];
      zero5(fzm);
      T um2[5], um1[5], up1[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        um1[v] = __shfl_up(u0[v], 1, 64);
        um2[v] = __shfl_up(u0[v], 2, 64);
        up1[v] = __shfl_down(u0[v], 1, 64);
      }
      if (lane == 0 && in_plane && k >= 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) um1[v] = ldv<T>(sv[v], c - W);
      }
      if (lane <= 1 && in_plane && k >= 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v) um2[v] = ldv<T>(sv[v], c - 2 * W);
      }
      if (lane == 63 && in_plane && k + 1 < nk) {
#pragma unroll
        for (int v = 0; v < 5; ++v) up1[v] = ldv<T>(sv[v], c + W);
      }
      if (in_plane && k >= 1) face_flux_r<T, RECON, LIM, FAST, 2>(um2, um1, u0, up1, k >= 2, k + 1 < nk, Az, fzm);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        res[v] += fzm[v];
        const T up = __shfl_down(fzm[v], 1, 64);
        fzp[v] = k + 1 < nk ? up : T(0);
      }
    }
    // the x face i+1/2 from U(i), U(i+1), U(i+2)
    T up1[5];
    zero5(f);
#pragma unroll
    for (int v = 0; v < 5; ++v) up1[v] = u0[v];
    if (i + 1 < ni) {
      T up2[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        up1[v] = ldv<T>(sv[v], c + sx);
        up2[v] = up1[v];
      }
      if (i + 2 < ni) {
#pragma unroll
        for (int v = 0; v < 5; ++v) up2[v] = ldv<T>(sv[v], c + 2 * sx);
      }
      if (own) {
        T vl[5], vr[5], xn[5];
        if constexpr (LC) {
#pragma unroll
          for (int v = 0; v < 5; ++v) xs[v] = carry[v * 256 + tid];
        }
        if constexpr (SYM) {
          const bool okp = i + 2 < ni;
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            xn[v] = okp ? limited<T, LIM, true>(up1[v] - u0[v], up2[v] - up1[v]) : T(0);
            vl[v] = u0[v] + T(0.5) * xs[v];
            vr[v] = up1[v] - T(0.5) * xn[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            vl[v] = xs[v];
            vr[v] = up1[v];
            xn[v] = up1[v];
          }
          if constexpr (RECON == RECON_UPWIND2) {
            if (i + 2 < ni) {
#pragma unroll
              for (int v = 0; v < 5; ++v) {
                vr[v] = up1[v] + T(0.5) * limited<T, LIM, FAST>(up1[v] - up2[v], u0[v] - up1[v]);
                xn[v] = up1[v] + T(0.5) * limited<T, LIM, FAST>(up1[v] - u0[v], up2[v] - up1[v]);
              }
            }
          }
        }
        ausm_flux_set<T, FAST, 0>(vl, vr, Ax, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) xs[v] = xn[v];
        if constexpr (LC) {
#pragma unroll
          for (int v = 0; v < 5; ++v) carry[v * 256 + tid] = xn[v];
        }
      }
    }
    if (own) {
      // cartesianiterator.h:310-322
      T nv[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        res[v] -= fzp[v];
        res[v] -= fyp[v];
        res[v] -= f[v];
        const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
        nv[v] = ldv<T>(bv, c) + pf * res[v];
      }
      bool store = true;
      if (push) {
        auto i_in = [&](const CellBox& b) { return (uint32_t)(i - b.i0) < (uint32_t)(b.i1 - b.i0); };
        store = !(rjk0 && i_in(P.recv[0])) && !(rjk1 && i_in(P.recv[1]));
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if ((s == 0 ? pjk0 : pjk1) && i_in(P.push[s])) {
            const uint32_t tb =
                ((uint32_t)P.push_base[s] + (uint32_t)i * P.push_si[s] + (uint32_t)j * P.push_sj[s] + (uint32_t)k) * W;
#pragma unroll
            for (int v = 0; v < 5; ++v)
              *reinterpret_cast<T*>(reinterpret_cast<char*>(dst + (uint64_t)v * vs) + tb) = T(0) + nv[v];
          }
        }
      }
      if (store) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
          *reinterpret_cast<T*>(reinterpret_cast<char*>(dst + (uint64_t)v * vs) + c) = nv[v];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if constexpr (LC)
        carry[(5 + v) * 256 + tid] = f[v];
      else
        fxm[v] = f[v];
      u0[v] = up1[v];
    }
  }
}

// The marching kernel as the compiler allocates it, and capped at 8 and at 7
// waves per SIMD (at most 64 and 72 VGPRs; the 8-wave form spills a few
// values per plane).  drn_autotune picks among them.
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_xm(const PatchDev* __restrict__ pds, const TileMarch* __restrict__ items,
                                                  const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                  T factor, int push) {
  stage_xm_item<T, RECON, LIM, FAST>(pds, items, src, base, dst, vs, factor, push);
}
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_stage_xm8(
    const PatchDev* __restrict__ pds, const TileMarch* __restrict__ items, const T* __restrict__ src, const T* base,
    T* dst, uint64_t vs, T factor, int push) {
  stage_xm_item<T, RECON, LIM, FAST>(pds, items, src, base, dst, vs, factor, push);
}
// the same with xs and the x face flux carried in LDS
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_xml(const PatchDev* __restrict__ pds,
                                                   const TileMarch* __restrict__ items, const T* __restrict__ src,
                                                   const T* base, T* dst, uint64_t vs, T factor, int push) {
  __shared__ T carry[10 * 256];
  stage_xm_item<T, RECON, LIM, FAST, true>(pds, items, src, base, dst, vs, factor, push, carry);
}
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 7))) void k_stage_xm7(
    const PatchDev* __restrict__ pds, const TileMarch* __restrict__ items, const T* __restrict__ src, const T* base,
    T* dst, uint64_t vs, T factor, int push) {
  stage_xm_item<T, RECON, LIM, FAST>(pds, items, src, base, dst, vs, factor, push);
}

// ---------------------------------------------------------------------------
// cell kernel for two planes per thread
// ---------------------------------------------------------------------------

// face functors at the patch faces of cell (i,j,k) (cartesianiterator.h:190-308)
template <typename T, bool FAST>
__device__ __forceinline__ void add_walls(const PatchDev& P, int i, int j, int k, const T* u0, T dx, T dy, T dz,
                                          T* res) {
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const bool wx = ni > 2 && (i == 0 || i == ni - 1);
  const bool wy = nj > 2 && (j == 0 || j == nj - 1);
  const bool wz = nk > 2 && (k == 0 || k == nk - 1);
  if (!(wx || wy || wz)) return;
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T xc = T(0.5) * dx + T(i) * dx, yc = T(0.5) * dy + T(j) * dy, zc = T(0.5) * dz + T(k) * dz;
  T f[5];
  if (ni > 2 && i == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[0], P.par[0], u0, 0, false, yc, zc, j, k, Ax, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (ni > 2 && i == ni - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[1], P.par[1], u0, 0, true, yc, zc, j, k, Ax, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
  if (nj > 2 && j == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[2], P.par[2], u0, 1, false, xc, zc, i, k, Ay, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (nj > 2 && j == nj - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[3], P.par[3], u0, 1, true, xc, zc, i, k, Ay, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
  if (nk > 2 && k == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[4], P.par[4], u0, 2, false, xc, yc, i, j, Az, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (nk > 2 && k == nk - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[5], P.par[5], u0, 2, true, xc, yc, i, j, Az, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
}

// One thread per column of two cells (i0, i0+1): the x face between them is
// evaluated once and the column i0-2 .. i0+3 is loaded once; 63 cells per
// wavefront with z faces shared by lane shuffle as in k_stage_cell<..., ZS>.
// About 4.5 face evaluations per cell.  Each cell's residual keeps the order
// Fx- + Fy- + Fz- - Fz+ - Fy+ - Fx+.  Work items start at even planes.
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_cell_x2(const PatchDev* __restrict__ pds,
                                                       const TileCell* __restrict__ tiles, const T* __restrict__ src,
                                                       const T* base, T* dst, uint64_t vs, T factor) {
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int lane = (int)(threadIdx.x & 63);
  const int c2 = t.c2 + 63 * (int)(threadIdx.x >> 6) + lane;
  const bool in_plane = c2 < njk;
  const bool own = in_plane && lane < 63;
  const int i0 = t.i;
  const bool has_b = i0 + 1 < ni;
  const int cc2 = in_plane ? c2 : njk - 1;
  const int j = cc2 / nk;
  const int k = cc2 - j * nk;
  constexpr uint32_t W = sizeof(T);
  const uint32_t ca = (P.off + (uint32_t)i0 * (uint32_t)njk + (uint32_t)cc2) * W;
  const uint32_t sx = (uint32_t)njk * W, sy = (uint32_t)nk * W;
  const uint32_t cb = ca + (has_b ? sx : 0u);
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);

  // column planes i0-2 .. i0+3 (entries outside the patch repeat a valid one)
  T col[6][5];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int ip = i0 - 2 + q;
    const int ic = ip < 0 ? 0 : (ip >= ni ? ni - 1 : ip);
    const uint32_t off = (P.off + (uint32_t)ic * (uint32_t)njk + (uint32_t)cc2) * W;
#pragma unroll
    for (int v = 0; v < 5; ++v) col[q][v] = ldv<T>(sv[v], off);
  }
  T ua[5], ub[5], resa[5], resb[5], fx1[5], fx2[5];
  zero5(resa);
  zero5(resb);
  zero5(fx1);
  zero5(fx2);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    ua[v] = col[2][v];
    ub[v] = col[3][v];
  }
  if (own && ni > 2) {
    const bool f0 = i0 > 0, f1 = i0 < ni - 1, f2 = has_b && i0 + 1 < ni - 1;
    if constexpr (FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD)) {
      T sl[4][5];  // limited slopes of planes i0-1 .. i0+2
#pragma unroll
      for (int q = 1; q <= 4; ++q) {
        const int ip = i0 - 2 + q;
        const bool ok = ip >= 1 && ip + 1 < ni;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          sl[q - 1][v] = ok ? limited<T, LIM, true>(col[q][v] - col[q - 1][v], col[q + 1][v] - col[q][v]) : T(0);
      }
      if (f0) {
        T vl[5], vr[5], f[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          vl[v] = col[1][v] + T(0.5) * sl[0][v];
          vr[v] = col[2][v] - T(0.5) * sl[1][v];
        }
        zero5(f);
        ausm_flux<T, FAST, 0>(vl, vr, Ax, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) resa[v] += f[v];
      }
      if (f1) {
        T vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          vl[v] = col[2][v] + T(0.5) * sl[1][v];
          vr[v] = col[3][v] - T(0.5) * sl[2][v];
        }
        ausm_flux<T, FAST, 0>(vl, vr, Ax, fx1);
      }
      if (f2) {
        T vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          vl[v] = col[3][v] + T(0.5) * sl[2][v];
          vr[v] = col[4][v] - T(0.5) * sl[3][v];
        }
        ausm_flux<T, FAST, 0>(vl, vr, Ax, fx2);
      }
    } else {
      if (f0) {
        T f[5];
        zero5(f);
        face_flux_r<T, RECON, LIM, FAST, 0>(col[0], col[1], col[2], col[3], i0 >= 2, i0 + 1 < ni, Ax, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) resa[v] += f[v];
      }
      if (f1) face_flux_r<T, RECON, LIM, FAST, 0>(col[1], col[2], col[3], col[4], i0 >= 1, i0 + 2 < ni, Ax, fx1);
      if (f2) face_flux_r<T, RECON, LIM, FAST, 0>(col[2], col[3], col[4], col[5], true, i0 + 3 < ni, Ax, fx2);
    }
    // the face between the two cells is the minus x face of cell b
#pragma unroll
    for (int v = 0; v < 5; ++v) resb[v] += fx1[v];
  }
  // the two cells one after the other: y faces, z faces (shared), walls, update
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const bool live = h == 0 || has_b;
    const int i = i0 + h;
    const uint32_t c = h == 0 ? ca : cb;
    T* u0 = h == 0 ? ua : ub;
    T* res = h == 0 ? resa : resb;
    T fyp[5], fzm[5], fzp[5];
    zero5(fyp);
    if (own && live) axis_faces_r<T, RECON, LIM, FAST, 1>(sv, c, sy, j, nj, u0, Ay, res, fyp);
    zero5(fzm);
    if (in_plane && live && nk > 2 && k >= 1) {
      T um2[5], um1[5], up1[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        um1[v] = ldv<T>(sv[v], c - W);
        um2[v] = k >= 2 ? ldv<T>(sv[v], c - 2 * W) : um1[v];
        up1[v] = k + 1 < nk ? ldv<T>(sv[v], c + W) : u0[v];
      }
      face_flux_r<T, RECON, LIM, FAST, 2>(um2, um1, u0, up1, k >= 2, k + 1 < nk, Az, fzm);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      res[v] += fzm[v];
      const T up = __shfl_down(fzm[v], 1, 64);
      fzp[v] = (nk > 2 && k + 1 < nk) ? up : T(0);
    }
    if (own && live) {
      const T* fxp = h == 0 ? fx1 : fx2;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        res[v] -= fzp[v];
        res[v] -= fyp[v];
        res[v] -= fxp[v];
      }
      add_walls<T, FAST>(P, i, j, k, u0, dx, dy, dz, res);
      // cartesianiterator.h:310-322
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
        char* dv = reinterpret_cast<char*>(dst + (uint64_t)v * vs);
        *reinterpret_cast<T*>(dv + c) = ldv<T>(bv, c) + pf * res[v];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// cell kernel with y and z faces shared inside the workgroup
// ---------------------------------------------------------------------------

// 8 wavefronts own 8 x 63 = 504 consecutive cells of one (j,k) plane.  As in
// k_stage_cell<..., ZS = true> the z faces are evaluated once and handed to the
// neighbouring lane.  Each cell also evaluates its y face j-1/2 and publishes
// it in LDS; after one barrier a cell takes its plus y face j+1/2 from the cell
// NK further on when that cell belongs to the workgroup, and evaluates it only
// otherwise (the last NK cells of the item).  About 4.1 face evaluations per
// cell.  The residual keeps the order Fx- + Fy- + Fz- - Fz+ - Fy+ - Fx+.
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(512) void k_stage_cell_yz(const PatchDev* __restrict__ pds,
                                                       const TileCell* __restrict__ tiles, const T* __restrict__ src,
                                                       const T* base, T* dst, uint64_t vs, T factor) {
  constexpr int NC = 63 * 8;
  __shared__ T sfy[5][NC];
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int lane = (int)(threadIdx.x & 63);
  const int q = 63 * (int)(threadIdx.x >> 6) + lane;  // cell of the item; lane 63 shadows the next wave's first
  const int c2 = t.c2 + q;
  const bool in_plane = c2 < njk;
  const bool own = in_plane && lane < 63;
  const int i = t.i;
  const int cc2 = in_plane ? c2 : njk - 1;
  const int j = cc2 / nk;
  const int k = cc2 - j * nk;
  constexpr uint32_t W = sizeof(T);
  const uint32_t c = (P.off + (uint32_t)i * (uint32_t)njk + (uint32_t)cc2) * W;
  const uint32_t sx = (uint32_t)njk * W, sy = (uint32_t)nk * W;
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  T res[5], f[5], u0[5], fxp[5];
  zero5(res);
  zero5(fxp);
#pragma unroll
  for (int v = 0; v < 5; ++v) u0[v] = ldv<T>(sv[v], c);
  if (own) axis_faces_r<T, RECON, LIM, FAST, 0>(sv, c, sx, i, ni, u0, Ax, res, fxp);
  // y face j-1/2, published for the cell below
  T fym[5];
  zero5(fym);
  if (own && nj > 2 && j >= 1) {
    T um2[5], um1[5], up1[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      um1[v] = ldv<T>(sv[v], c - sy);
      um2[v] = j >= 2 ? ldv<T>(sv[v], c - 2 * sy) : um1[v];
      up1[v] = j + 1 < nj ? ldv<T>(sv[v], c + sy) : u0[v];
    }
    face_flux_r<T, RECON, LIM, FAST, 1>(um2, um1, u0, up1, j >= 2, j + 1 < nj, Ay, fym);
  }
  if (own) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      res[v] += fym[v];
      sfy[v][q] = fym[v];
    }
  }
  // z face k-1/2 of every lane, the plus z face from the lane above
  T fzm[5], fzp[5];
  zero5(fzm);
  if (in_plane && nk > 2 && k >= 1) {
    T um2[5], um1[5], up1[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      um1[v] = ldv<T>(sv[v], c - W);
      um2[v] = k >= 2 ? ldv<T>(sv[v], c - 2 * W) : um1[v];
      up1[v] = k + 1 < nk ? ldv<T>(sv[v], c + W) : u0[v];
    }
    face_flux_r<T, RECON, LIM, FAST, 2>(um2, um1, u0, up1, k >= 2, k + 1 < nk, Az, fzm);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    res[v] += fzm[v];
    const T up = __shfl_down(fzm[v], 1, 64);
    fzp[v] = (nk > 2 && k + 1 < nk) ? up : T(0);
  }
  __syncthreads();
  if (!own) return;
  // plus y face j+1/2: from the cell NK further on, or evaluated here
  T fyp[5];
  zero5(fyp);
  if (nj > 2 && j + 1 < nj) {
    if (q + nk < NC) {
#pragma unroll
      for (int v = 0; v < 5; ++v) fyp[v] = sfy[v][q + nk];
    } else {
      T um1[5], up1[5], up2[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        up1[v] = ldv<T>(sv[v], c + sy);
        um1[v] = j >= 1 ? ldv<T>(sv[v], c - sy) : u0[v];
        up2[v] = j + 2 < nj ? ldv<T>(sv[v], c + 2 * sy) : up1[v];
      }
      face_flux_r<T, RECON, LIM, FAST, 1>(um1, u0, up1, up2, j >= 1, j + 2 < nj, Ay, fyp);
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    res[v] -= fzp[v];
    res[v] -= fyp[v];
    res[v] -= fxp[v];
  }
  // patch-face functors (cartesianiterator.h:190-308)
  const T xc = T(0.5) * dx + T(i) * dx, yc = T(0.5) * dy + T(j) * dy, zc = T(0.5) * dz + T(k) * dz;
  if (ni > 2 && i == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[0], P.par[0], u0, 0, false, yc, zc, j, k, Ax, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (ni > 2 && i == ni - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[1], P.par[1], u0, 0, true, yc, zc, j, k, Ax, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
  if (nj > 2 && j == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[2], P.par[2], u0, 1, false, xc, zc, i, k, Ay, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (nj > 2 && j == nj - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[3], P.par[3], u0, 1, true, xc, zc, i, k, Ay, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
  if (nk > 2 && k == 0) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[4], P.par[4], u0, 2, false, xc, yc, i, j, Az, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] += f[v];
  }
  if (nk > 2 && k == nk - 1) {
    zero5(f);
    wall_flux<T, FAST>(P.bc[5], P.par[5], u0, 2, true, xc, yc, i, j, Az, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) res[v] -= f[v];
  }
  // cartesianiterator.h:310-322
  const T pf = factor / (dx * dy * dz);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
    char* dv = reinterpret_cast<char*>(dst + (uint64_t)v * vs);
    *reinterpret_cast<T*>(dv + c) = ldv<T>(bv, c) + pf * res[v];
  }
}

// ---------------------------------------------------------------------------
// marching stage kernel
// ---------------------------------------------------------------------------

// MUSCL state at a face from the three cells on one side (upwind2.h:29-52):
// a = cell next to the face, m = next cell outwards (valid when hm), b = cell across
template <typename T, int RECON, int LIM, bool FAST>
__device__ __forceinline__ T recon1(T a, T m, T b, bool hm) {
  if constexpr (RECON == RECON_UPWIND2) {
    if (hm) return a + T(0.5) * limited<T, LIM, FAST>(a - m, b - a);
  }
  return a;
}

// AUSM+ flux with a run-time normal direction (1 = y, 2 = z) for the LDS tasks
template <typename T, bool FAST>
__device__ __forceinline__ void ausm_flux_rt(const T* vl, const T* vr, int dir, T A, T* f) {
  if (dir == 1)
    ausm_flux<T, FAST, 1>(vl, vr, A, f);
  else
    ausm_flux<T, FAST, 2>(vl, vr, A, f);
}

// A workgroup of BS threads owns BS consecutive cells c0.. of the (j,k) plane of
// one patch (plane index c = j*NK + k) and marches over the planes [i0, i1).
// Per plane i:
//   x faces: each thread keeps U(i-1..i+2) of its column in registers and
//            computes F(i+1/2); F(i-1/2) is carried from the previous plane.
//   y, z faces: the plane with a 2*NK halo is staged in LDS; the faces needed
//            by the chunk (BS + NK y faces, BS + rows + 1 z faces) are computed
//            once each and written to LDS.
//   update:  dst = base + factor/dV * (Fx- - Fx+ + Fy- - Fy+ + Fz- - Fz+).
// Patch-face functors are evaluated as the fluxes of the boundary faces.
// Requires NI, NJ, NK >= 3 (the host routes thinner patches to k_stage_cell).
template <typename T, int RECON, int LIM, bool FAST, int BS>
__global__ __launch_bounds__(BS) void k_stage_march(const PatchDev* __restrict__ pds,
                                                    const TileMarch* __restrict__ tiles, const T* __restrict__ src,
                                                    const T* base, T* dst, uint64_t vs, T factor, int lds_cells,
                                                    int lds_faces) {
  extern __shared__ __attribute__((aligned(16))) unsigned char drn_smem[];
  T* sU = reinterpret_cast<T*>(drn_smem);  // [5][lds_cells]
  T* sF = sU + 5 * lds_cells;               // [5][lds_faces]
  const TileMarch t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int c0 = t.c0;
  const int cnt = min(BS, njk - c0);
  const int tid = threadIdx.x;
  const bool own = tid < cnt;
  const int c = c0 + tid;
  const int j = own ? c / nk : 0;
  const int k = own ? c - j * nk : 0;
  const int h0 = max(0, c0 - 2 * nk), h1 = min(njk, c0 + cnt + 2 * nk), hn = h1 - h0;
  const int nY = cnt + nk;
  const int jfirst = c0 / nk, jlast = (c0 + cnt - 1) / nk;
  const int zf0 = c0 + jfirst;
  const int nZ = cnt + 1 + (jlast - jfirst);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);
  const uint64_t pb = (uint64_t)P.off;
  const uint64_t col = pb + (uint64_t)c;
  const uint64_t sx = (uint64_t)njk;
  const int i0 = t.i0, i1 = t.i1;

  // register window of the own column: W[q] = U(i - 1 + q)
  T W[4][5];
  auto loadW = [&](int q, int ip) {
    if (own && ip >= 0 && ip < ni) {
#pragma unroll
      for (int v = 0; v < 5; ++v) W[q][v] = src[v * vs + col + (uint64_t)ip * sx];
    }
  };
  // x face between planes ia-1 and ia of the own column, window holds ia-2..ia+1 in W[0..3]
  auto xface = [&](int ia, T* f) {
    zero5(f);
    if (ia == 0) {
      wall_flux<T, FAST>(P.bc[0], P.par[0], W[2], 0, false, T(0.5) * dy + T(j) * dy, T(0.5) * dz + T(k) * dz, j, k, Ax, f);
    } else if (ia == ni) {
      wall_flux<T, FAST>(P.bc[1], P.par[1], W[1], 0, true, T(0.5) * dy + T(j) * dy, T(0.5) * dz + T(k) * dz, j, k, Ax, f);
    } else {
      const bool hm = ia >= 2, hp = ia + 1 < ni;
      T vl[5], vr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = recon1<T, RECON, LIM, FAST>(W[1][v], W[0][v], W[2][v], hm);
        vr[v] = recon1<T, RECON, LIM, FAST>(W[2][v], W[3][v], W[1][v], hp);
      }
      ausm_flux<T, FAST, 0>(vl, vr, Ax, f);
    }
  };

  T fm[5], fp[5];
  loadW(0, i0 - 2);
  loadW(1, i0 - 1);
  loadW(2, i0);
  loadW(3, i0 + 1);
  if (own) xface(i0, fm);
  // shift to W = U(i0-1 .. i0+2)
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    W[0][v] = W[1][v];
    W[1][v] = W[2][v];
    W[2][v] = W[3][v];
  }
  loadW(3, i0 + 2);

  for (int i = i0; i < i1; ++i) {
    __syncthreads();  // previous plane's LDS readers are done
    const uint64_t prow = pb + (uint64_t)i * sx + (uint64_t)h0;
    for (int q = tid; q < hn; q += BS) {
#pragma unroll
      for (int v = 0; v < 5; ++v) sU[v * lds_cells + q] = src[v * vs + prow + q];
    }
    __syncthreads();
    const T xi = T(0.5) * dx + T(i) * dx;
    // y and z faces of the chunk
    for (int tk = tid; tk < nY + nZ; tk += BS) {
      T f[5];
      zero5(f);
      int dir, cm1, cp1, face_wall = -1;  // face_wall: functor face id, cell in cp1 (minus walls) / cm1 (plus)
      bool hm = false, hp = false;
      int cm2 = 0, cp2 = 0;
      int ja = 0, kb = 0;
      if (tk < nY) {
        dir = 1;
        const int fy = c0 + tk;
        const int jf = fy / nk;
        cp1 = fy;
        cm1 = fy - nk;
        if (jf == 0) {
          face_wall = 2;
          ja = jf;
          kb = fy - jf * nk;
        } else if (jf == nj) {
          face_wall = 3;
          ja = jf - 1;
          kb = fy - jf * nk;
        } else {
          hm = jf >= 2;
          hp = jf + 1 < nj;
          cm2 = cm1 - nk;
          cp2 = cp1 + nk;
        }
      } else {
        dir = 2;
        const int zf = zf0 + (tk - nY);
        const int jz = zf / (nk + 1);
        const int kz = zf - jz * (nk + 1);
        cm1 = jz * nk + kz - 1;
        cp1 = cm1 + 1;
        if (kz == 0) {
          face_wall = 4;
          ja = jz;
          kb = 0;
        } else if (kz == nk) {
          face_wall = 5;
          ja = jz;
          kb = nk - 1;
        } else {
          hm = kz >= 2;
          hp = kz + 1 < nk;
          cm2 = cm1 - 1;
          cp2 = cp1 + 1;
        }
      }
      if (face_wall >= 0) {
        const bool plus = (face_wall & 1) != 0;
        const int cell = (plus ? cm1 : cp1) - h0;
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = sU[v * lds_cells + cell];
        if (dir == 1)
          wall_flux<T, FAST>(P.bc[face_wall], P.par[face_wall], var, 1, plus, xi, T(0.5) * dz + T(kb) * dz, i, kb, Ay, f);
        else
          wall_flux<T, FAST>(P.bc[face_wall], P.par[face_wall], var, 2, plus, xi, T(0.5) * dy + T(ja) * dy, i, ja, Az, f);
      } else {
        T vl[5], vr[5];
        const int qm1 = cm1 - h0, qp1 = cp1 - h0, qm2 = hm ? cm2 - h0 : qm1, qp2 = hp ? cp2 - h0 : qp1;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T a = sU[v * lds_cells + qm1], b = sU[v * lds_cells + qp1];
          const T m = sU[v * lds_cells + qm2], q = sU[v * lds_cells + qp2];
          vl[v] = recon1<T, RECON, LIM, FAST>(a, m, b, hm);
          vr[v] = recon1<T, RECON, LIM, FAST>(b, q, a, hp);
        }
        ausm_flux_rt<T, FAST>(vl, vr, dir, dir == 1 ? Ay : Az, f);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) sF[v * lds_faces + tk] = f[v];
    }
    // x face i+1/2 of the own column
    if (own) xface(i + 1, fp);
    __syncthreads();
    if (own) {
      const int qy = c - c0, qz = nY + (c + j) - zf0;
      const uint64_t g = col + (uint64_t)i * sx;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T r = (fm[v] - fp[v]) + (sF[v * lds_faces + qy] - sF[v * lds_faces + qy + nk]) +
                    (sF[v * lds_faces + qz] - sF[v * lds_faces + qz + 1]);
        dst[v * vs + g] = base[v * vs + g] + pf * r;
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        fm[v] = fp[v];
        W[0][v] = W[1][v];
        W[1][v] = W[2][v];
        W[2][v] = W[3][v];
      }
    }
    loadW(3, i + 3);
  }
}

// ---------------------------------------------------------------------------
// slope-form marching stage kernel (fast arithmetic, van Albada / minmod)
// ---------------------------------------------------------------------------

// Limited MUSCL slope of a cell from its backward and forward differences.
// For van Albada and minmod the reference's left-face state
//   U_i + 1/2 lim(dL, dR) dL   (upwind2.h:44 with d01 = dL, d12 = dR)
// and right-face state U_i + 1/2 lim(-dR, -dL)(-dR) are U_i + s/2 and U_i - s/2
// with the symmetric slope s below (van Albada: lim(r)*d01 = dL*dR*(dL+dR)/(dL^2+dR^2);
// minmod: lim(r)*d01 = minmod(dL, dR)).  The reference's nonZero(d01, 1e-10)
// guard is replaced by a 1e-30 term in the denominator; the two differ only for
// |dL| < 1e-10, where the slope itself is below 1e-10.
template <typename T, int LIM>
__device__ __forceinline__ T sym_slope(T dl, T dr) {
  if constexpr (LIM == LIM_VANALBADA) {
    const T num = dl * dr * (dl + dr);
    const T den = dl * dl + dr * dr + T(1e-30);
    return ddiv<T, true>(num, den);
  } else {
    // minmod
    const T al = dabs(dl), ar = dabs(dr);
    const T m = al < ar ? dl : dr;
    return (dl * dr > T(0)) ? m : T(0);
  }
}

__device__ __forceinline__ int fdiv_idx(int a, float rinv) {
  // exact floor(a / n) for 0 <= a < 2^20 and n < 2^10, with rinv = 1/n
  return (int)(((float)a + 0.5f) * rinv);
}

template <typename T, int LIM, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_stage_slope(const PatchDev* __restrict__ pds,
                                                    const TileMarch* __restrict__ tiles, const T* __restrict__ src,
                                                    const T* base, T* dst, uint64_t vs, T factor, int lds_cells,
                                                    int lds_faces, int lds_sy, int lds_sz) {
  extern __shared__ __attribute__((aligned(16))) unsigned char drn_smem[];
  T* sU0 = reinterpret_cast<T*>(drn_smem);  // [2][5][lds_cells]
  T* sSy = sU0 + 10 * lds_cells;            // [5][lds_sy]
  T* sSz = sSy + 5 * lds_sy;                // [5][lds_sz]
  T* sF = sSz + 5 * lds_sz;                 // [5][lds_faces]
  const TileMarch t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const float rnk = 1.0f / (float)nk, rnk1 = 1.0f / (float)(nk + 1);
  const int c0 = t.c0;
  const int cnt = min(BS, njk - c0);
  const int tid = threadIdx.x;
  const bool own = tid < cnt;
  const int c = c0 + tid;
  const int j = own ? fdiv_idx(c, rnk) : 0;
  const int k = own ? c - j * nk : 0;
  const int h0 = max(0, c0 - 2 * nk), h1 = min(njk, c0 + cnt + 2 * nk), hn = h1 - h0;
  const int sy0 = max(0, c0 - nk), sy1 = min(njk, c0 + cnt + nk), nsy = sy1 - sy0;
  const int sz0 = max(0, c0 - 1), sz1 = min(njk, c0 + cnt + 1), nsz = sz1 - sz0;
  const int nY = cnt + nk;
  const int jfirst = fdiv_idx(c0, rnk), jlast = fdiv_idx(c0 + cnt - 1, rnk);
  const int zf0 = c0 + jfirst;
  const int nZ = cnt + 1 + (jlast - jfirst);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);
  const uint64_t pb = (uint64_t)P.off;
  const uint64_t col = pb + (uint64_t)c;
  const uint64_t sx = (uint64_t)njk;
  const int i0 = t.i0, i1 = t.i1;
  const T half = T(0.5);

  // x direction in registers: W[q] = U(i-1+q) of the own column, slopes of planes i and i+1
  T W[4][5], sxc[5], sxn[5], fm[5], fp[5];
  auto loadW = [&](int q, int ip) {
    if (own && ip >= 0 && ip < ni) {
#pragma unroll
      for (int v = 0; v < 5; ++v) W[q][v] = src[v * vs + col + (uint64_t)ip * sx];
    }
  };
  // slope of plane ip from W[a] (ip-1), W[a+1] (ip), W[a+2] (ip+1)
  auto xslope = [&](int ip, int a, T* sl) {
    const bool ok = ip >= 1 && ip + 1 < ni;
#pragma unroll
    for (int v = 0; v < 5; ++v) sl[v] = ok ? sym_slope<T, LIM>(W[a + 1][v] - W[a][v], W[a + 2][v] - W[a + 1][v]) : T(0);
  };
  // x face between planes ia-1 (W[b], slope sl) and ia (W[b+1], slope sr)
  auto xface = [&](int ia, int b, const T* sl, const T* sr, T* f) {
    zero5(f);
    const T yc = half * dy + T(j) * dy, zc = half * dz + T(k) * dz;
    if (ia == 0) {
      wall_flux<T, true>(P.bc[0], P.par[0], W[b + 1], 0, false, yc, zc, j, k, Ax, f);
    } else if (ia == ni) {
      wall_flux<T, true>(P.bc[1], P.par[1], W[b], 0, true, yc, zc, j, k, Ax, f);
    } else {
      T vl[5], vr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = W[b][v] + half * sl[v];
        vr[v] = W[b + 1][v] - half * sr[v];
      }
      ausm_flux<T, true, 0>(vl, vr, Ax, f);
    }
  };

  // prologue: x face i0-1/2 from U(i0-2 .. i0+1)
  loadW(0, i0 - 2);
  loadW(1, i0 - 1);
  loadW(2, i0);
  loadW(3, i0 + 1);
  if (own) {
    T sa[5];
    xslope(i0 - 1, 0, sa);
    xslope(i0, 1, sxc);
    xface(i0, 1, sa, sxc, fm);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      W[0][v] = W[1][v];
      W[1][v] = W[2][v];
      W[2][v] = W[3][v];
    }
  }
  loadW(3, i0 + 2);
  // plane i0 straight into LDS buffer 0, plane i0+1 into the prefetch registers
  constexpr int PRE = 2;  // the first PRE*BS halo cells of a plane go through registers
  T pre[PRE][5];
  auto prefetch = [&](int ip) {
    if (ip < i1) {
      const uint64_t prow = pb + (uint64_t)ip * sx + (uint64_t)h0;
#pragma unroll
      for (int r = 0; r < PRE; ++r) {
        const int q = tid + r * BS;
        if (q < hn) {
#pragma unroll
          for (int v = 0; v < 5; ++v) pre[r][v] = src[v * vs + prow + q];
        }
      }
    }
  };
  {
    const uint64_t prow = pb + (uint64_t)i0 * sx + (uint64_t)h0;
    for (int q = tid; q < hn; q += BS) {
#pragma unroll
      for (int v = 0; v < 5; ++v) sU0[v * lds_cells + q] = src[v * vs + prow + q];
    }
  }
  prefetch(i0 + 1);
  T rx[5];  // x contribution Fx(i-1/2) - Fx(i+1/2) of the plane awaiting its update
  // residual and update of plane ip from rx and the y/z fluxes in sF
  auto update = [&](int ip) {
    if (own) {
      const int qy = c - c0, qz = nY + (c + j) - zf0;
      const uint64_t g = col + (uint64_t)ip * sx;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T r = rx[v] + (sF[v * lds_faces + qy] - sF[v * lds_faces + qy + nk]) +
                    (sF[v * lds_faces + qz] - sF[v * lds_faces + qz + 1]);
        dst[v * vs + g] = base[v * vs + g] + pf * r;
      }
    }
  };

  for (int i = i0; i < i1; ++i) {
    T* sU = sU0 + (size_t)((i - i0) & 1) * 5 * lds_cells;
    T* sUn = sU0 + (size_t)((i - i0 + 1) & 1) * 5 * lds_cells;
    __syncthreads();
    // phase A: y and z slopes of plane i; x slope of plane i+1; update of plane i-1
    for (int tk = tid; tk < nsy + nsz; tk += BS) {
      if (tk < nsy) {
        const int cc = sy0 + tk;
        const int jj = fdiv_idx(cc, rnk);
        const int q = cc - h0;
        const bool ok = jj >= 1 && jj + 1 < nj;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T um = sU[v * lds_cells + (ok ? q - nk : q)], u0 = sU[v * lds_cells + q],
                  up = sU[v * lds_cells + (ok ? q + nk : q)];
          sSy[v * lds_sy + tk] = ok ? sym_slope<T, LIM>(u0 - um, up - u0) : T(0);
        }
      } else {
        const int r = tk - nsy;
        const int cc = sz0 + r;
        const int kk = cc - fdiv_idx(cc, rnk) * nk;
        const int q = cc - h0;
        const bool ok = kk >= 1 && kk + 1 < nk;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T um = sU[v * lds_cells + (ok ? q - 1 : q)], u0 = sU[v * lds_cells + q],
                  up = sU[v * lds_cells + (ok ? q + 1 : q)];
          sSz[v * lds_sz + r] = ok ? sym_slope<T, LIM>(u0 - um, up - u0) : T(0);
        }
      }
    }
    if (i > i0) update(i - 1);
    if (own) xslope(i + 1, 1, sxn);
    __syncthreads();
    // phase B: y and z face fluxes of plane i, x face i+1/2; stage plane i+1
    const T xi = half * dx + T(i) * dx;
    for (int tk = tid; tk < nY + nZ; tk += BS) {
      T f[5];
      zero5(f);
      int cm1, cp1, wall = -1, dir, ja = 0, kb = 0;
      if (tk < nY) {
        dir = 1;
        const int fy = c0 + tk;
        const int jf = fdiv_idx(fy, rnk);
        cp1 = fy;
        cm1 = fy - nk;
        if (jf == 0) {
          wall = 2;
          kb = fy;
        } else if (jf == nj) {
          wall = 3;
          kb = fy - jf * nk;
        }
      } else {
        dir = 2;
        const int zf = zf0 + (tk - nY);
        const int jz = fdiv_idx(zf, rnk1);
        const int kz = zf - jz * (nk + 1);
        cm1 = jz * nk + kz - 1;
        cp1 = cm1 + 1;
        ja = jz;
        if (kz == 0)
          wall = 4;
        else if (kz == nk)
          wall = 5;
      }
      if (wall >= 0) {
        const bool plus = (wall & 1) != 0;
        const int cell = (plus ? cm1 : cp1) - h0;
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = sU[v * lds_cells + cell];
        if (dir == 1)
          wall_flux<T, true>(P.bc[wall], P.par[wall], var, 1, plus, xi, half * dz + T(kb) * dz, i, kb, Ay, f);
        else
          wall_flux<T, true>(P.bc[wall], P.par[wall], var, 2, plus, xi, half * dy + T(ja) * dy, i, ja, Az, f);
      } else {
        T vl[5], vr[5];
        const int qm = cm1 - h0, qp = cp1 - h0;
        const T* sS = dir == 1 ? sSy : sSz;
        const int ls = dir == 1 ? lds_sy : lds_sz;
        const int s0 = dir == 1 ? sy0 : sz0;
        const int sm = cm1 - s0, sp = cp1 - s0;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          vl[v] = sU[v * lds_cells + qm] + half * sS[v * ls + sm];
          vr[v] = sU[v * lds_cells + qp] - half * sS[v * ls + sp];
        }
        ausm_flux_rt<T, true>(vl, vr, dir, dir == 1 ? Ay : Az, f);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) sF[v * lds_faces + tk] = f[v];
    }
    if (own) {
      xface(i + 1, 1, sxc, sxn, fp);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        rx[v] = fm[v] - fp[v];
        fm[v] = fp[v];
        sxc[v] = sxn[v];
        W[0][v] = W[1][v];
        W[1][v] = W[2][v];
        W[2][v] = W[3][v];
      }
    }
    if (i + 1 < i1) {
#pragma unroll
      for (int r = 0; r < PRE; ++r) {
        const int q = tid + r * BS;
        if (q < hn) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sUn[v * lds_cells + q] = pre[r][v];
        }
      }
      // wide planes: the rest of the halo is loaded directly
      const uint64_t prow = pb + (uint64_t)(i + 1) * sx + (uint64_t)h0;
      for (int q = PRE * BS + tid; q < hn; q += BS) {
#pragma unroll
        for (int v = 0; v < 5; ++v) sUn[v * lds_cells + q] = src[v * vs + prow + q];
      }
      prefetch(i + 2);
    }
    loadW(3, i + 3);
  }
  __syncthreads();
  update(i1 - 1);
}

// ---------------------------------------------------------------------------
// column stage kernel (fast arithmetic, van Albada / minmod): no LDS, no barriers
// ---------------------------------------------------------------------------

// Each wavefront owns 63 consecutive cells of the (j,k) plane of one patch
// (lane l < 63 owns cell base + l) and marches them along i over [i0, i1).
//   x: every owned column keeps U(i-1..i+2) and the slope of plane i in
//      registers, so each x face and each x slope is evaluated once.
//   z: lane l evaluates the z face between cells base+l-1 and base+l, i.e. the
//      minus face of its own cell; the plus face is the next lane's value
//      (lane 63 exists only to evaluate the plus face of cell base+62).  The
//      patch-face functors at k = 0 / NK-1 are the faces of those cells.
//   y: both y faces of the owned cell are evaluated from the slopes of cells
//      j-1, j, j+1 (the reference's two MUSCL states of each face).
template <typename T, int LIM>
__global__ __launch_bounds__(256) void k_stage_column(const PatchDev* __restrict__ pds,
                                                      const TileMarch* __restrict__ tiles,
                                                      const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                      T factor) {
  const TileMarch t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const float rnk = 1.0f / (float)nk;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wbase = t.c0 + 63 * wv;
  const int c = wbase + lane;                      // cell whose minus z face this lane evaluates
  const bool in_patch = c < njk;
  const bool own = lane < 63 && in_patch;          // lane owns cell c
  const int cc = in_patch ? c : njk - 1;            // clamped for index math of idle lanes
  const int j = fdiv_idx(cc, rnk);
  const int k = cc - j * nk;
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);
  const T half = T(0.5);
  const uint64_t sx = (uint64_t)njk, sy = (uint64_t)nk;
  const uint64_t col = (uint64_t)P.off + (uint64_t)cc;
  const int i0 = t.i0, i1 = t.i1;
  const T yc = half * dy + T(j) * dy, zc = half * dz + T(k) * dz;
  // slope validity of the cells j-1, j, j+1 and k-1, k (first order next to the patch faces)
  const bool okj_m = j - 1 >= 1 && j - 1 + 1 < nj, okj_0 = j >= 1 && j + 1 < nj, okj_p = j + 1 >= 1 && j + 2 < nj;
  const bool okk_m = k - 1 >= 1 && k < nk, okk_0 = k >= 1 && k + 1 < nk;
  const bool has_jm2 = j >= 2, has_jp2 = j + 2 < nj, has_km2 = k >= 2, has_kp1 = k + 1 < nk;

  auto ld = [&](int v, uint64_t g) { return src[v * vs + g]; };

  T W[4][5], sxc[5], sxn[5], fm[5], fp[5];
  auto loadW = [&](int q, int ip) {
    if (own && ip >= 0 && ip < ni) {
#pragma unroll
      for (int v = 0; v < 5; ++v) W[q][v] = ld(v, col + (uint64_t)ip * sx);
    }
  };
  auto xslope = [&](int ip, int a, T* sl) {
    const bool ok = ip >= 1 && ip + 1 < ni;
#pragma unroll
    for (int v = 0; v < 5; ++v) sl[v] = ok ? sym_slope<T, LIM>(W[a + 1][v] - W[a][v], W[a + 2][v] - W[a + 1][v]) : T(0);
  };
  auto xface = [&](int ia, int b, const T* sl, const T* sr, T* f) {
    zero5(f);
    if (ia == 0) {
      wall_flux<T, true>(P.bc[0], P.par[0], W[b + 1], 0, false, yc, zc, j, k, Ax, f);
    } else if (ia == ni) {
      wall_flux<T, true>(P.bc[1], P.par[1], W[b], 0, true, yc, zc, j, k, Ax, f);
    } else {
      T vl[5], vr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = W[b][v] + half * sl[v];
        vr[v] = W[b + 1][v] - half * sr[v];
      }
      ausm_flux<T, true, 0>(vl, vr, Ax, f);
    }
  };

  loadW(0, i0 - 2);
  loadW(1, i0 - 1);
  loadW(2, i0);
  loadW(3, i0 + 1);
  if (own) {
    T sa[5];
    xslope(i0 - 1, 0, sa);
    xslope(i0, 1, sxc);
    xface(i0, 1, sa, sxc, fm);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      W[0][v] = W[1][v];
      W[1][v] = W[2][v];
      W[2][v] = W[3][v];
    }
  }
  loadW(3, i0 + 2);

  for (int i = i0; i < i1; ++i) {
    const uint64_t g = col + (uint64_t)i * sx;
    const T xi = half * dx + T(i) * dx;
    T U0[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) U0[v] = own ? W[1][v] : ld(v, g);
    // ---- z: minus face of cell c (every lane in the patch)
    T fz[5];
    zero5(fz);
    if (in_patch) {
      if (k == 0) {
        wall_flux<T, true>(P.bc[4], P.par[4], U0, 2, false, xi, yc, i, j, Az, fz);
      } else {
        T Um[5], Umm[5], Up[5], vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          Um[v] = ld(v, g - 1);
          Umm[v] = has_km2 ? ld(v, g - 2) : Um[v];
          Up[v] = has_kp1 ? ld(v, g + 1) : U0[v];
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T sm = okk_m ? sym_slope<T, LIM>(Um[v] - Umm[v], U0[v] - Um[v]) : T(0);
          const T s0 = okk_0 ? sym_slope<T, LIM>(U0[v] - Um[v], Up[v] - U0[v]) : T(0);
          vl[v] = Um[v] + half * sm;
          vr[v] = U0[v] - half * s0;
        }
        ausm_flux<T, true, 2>(vl, vr, Az, fz);
      }
    }
    // plus z face of the owned cell: the next lane's minus face, or the wall at k = NK-1
    T fzp[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) fzp[v] = __shfl_down(fz[v], 1, 64);
    if (own && k == nk - 1) {
      zero5(fzp);
      wall_flux<T, true>(P.bc[5], P.par[5], U0, 2, true, xi, yc, i, j, Az, fzp);
    }
    if (own) {
      // ---- y: both faces of the owned cell
      T fym[5], fyp[5];
      zero5(fym);
      zero5(fyp);
      {
        T Ujm[5], Ujmm[5], Ujp[5], Ujpp[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          Ujm[v] = j >= 1 ? ld(v, g - sy) : U0[v];
          Ujmm[v] = has_jm2 ? ld(v, g - 2 * sy) : Ujm[v];
          Ujp[v] = j + 1 < nj ? ld(v, g + sy) : U0[v];
          Ujpp[v] = has_jp2 ? ld(v, g + 2 * sy) : Ujp[v];
        }
        T sjm[5], sj0[5], sjp[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          sjm[v] = okj_m ? sym_slope<T, LIM>(Ujm[v] - Ujmm[v], U0[v] - Ujm[v]) : T(0);
          sj0[v] = okj_0 ? sym_slope<T, LIM>(U0[v] - Ujm[v], Ujp[v] - U0[v]) : T(0);
          sjp[v] = okj_p ? sym_slope<T, LIM>(Ujp[v] - U0[v], Ujpp[v] - Ujp[v]) : T(0);
        }
        if (j == 0) {
          wall_flux<T, true>(P.bc[2], P.par[2], U0, 1, false, xi, zc, i, k, Ay, fym);
        } else {
          T vl[5], vr[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            vl[v] = Ujm[v] + half * sjm[v];
            vr[v] = U0[v] - half * sj0[v];
          }
          ausm_flux<T, true, 1>(vl, vr, Ay, fym);
        }
        if (j == nj - 1) {
          wall_flux<T, true>(P.bc[3], P.par[3], U0, 1, true, xi, zc, i, k, Ay, fyp);
        } else {
          T vl[5], vr[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            vl[v] = U0[v] + half * sj0[v];
            vr[v] = Ujp[v] - half * sjp[v];
          }
          ausm_flux<T, true, 1>(vl, vr, Ay, fyp);
        }
      }
      // ---- x: face i+1/2 from the register window
      xslope(i + 1, 1, sxn);
      xface(i + 1, 1, sxc, sxn, fp);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const T r = (fm[v] - fp[v]) + (fym[v] - fyp[v]) + (fz[v] - fzp[v]);
        dst[v * vs + g] = base[v * vs + g] + pf * r;
        fm[v] = fp[v];
        sxc[v] = sxn[v];
        W[0][v] = W[1][v];
        W[1][v] = W[2][v];
        W[2][v] = W[3][v];
      }
    }
    loadW(3, i + 3);
  }
}

// ---------------------------------------------------------------------------
// cell-per-thread stage kernel, slope form (fast arithmetic, van Albada / minmod)
// ---------------------------------------------------------------------------

// One thread per cell as in k_stage_cell, but along each axis the limited
// slopes of the cells pos-1, pos, pos+1 are evaluated once (sym_slope) and the
// four MUSCL states of the two faces are U -/+ s/2; the own cell's slope serves
// both faces.  Nine slope evaluations per cell instead of twelve reconstructions.
template <typename T, int LIM, int DIR>
__device__ __forceinline__ void axis_faces(const PatchDev& P, const char* const* sv, uint32_t g, uint32_t st, int pos,
                                           int n, const T* U0, T A, T ta, T tb, int la, int lb, T* res) {
  const T half = T(0.5);
  if (n <= 2) return;  // the reference skips an axis with at most two cells
  T Um[5], Umm[5], Up[5], Upp[5];
  const bool hm = pos >= 1, hmm = pos >= 2, hp = pos + 1 < n, hpp = pos + 2 < n;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    Um[v] = hm ? ldv<T>(sv[v], g - st) : U0[v];
    Umm[v] = hmm ? ldv<T>(sv[v], g - 2 * st) : Um[v];
    Up[v] = hp ? ldv<T>(sv[v], g + st) : U0[v];
    Upp[v] = hpp ? ldv<T>(sv[v], g + 2 * st) : Up[v];
  }
  T s0[5];
  const bool ok0 = hm && hp;
#pragma unroll
  for (int v = 0; v < 5; ++v) s0[v] = ok0 ? sym_slope<T, LIM>(U0[v] - Um[v], Up[v] - U0[v]) : T(0);
  T f[5];
  // minus face
  zero5(f);
  if (!hm) {
    wall_flux<T, true>(P.bc[2 * DIR], P.par[2 * DIR], U0, DIR, false, ta, tb, la, lb, A, f);
  } else {
    T vl[5], vr[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const T sm = hmm ? sym_slope<T, LIM>(Um[v] - Umm[v], U0[v] - Um[v]) : T(0);
      vl[v] = Um[v] + half * sm;
      vr[v] = U0[v] - half * s0[v];
    }
    ausm_flux<T, true, DIR>(vl, vr, A, f);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) res[v] += f[v];
  // plus face
  zero5(f);
  if (!hp) {
    wall_flux<T, true>(P.bc[2 * DIR + 1], P.par[2 * DIR + 1], U0, DIR, true, ta, tb, la, lb, A, f);
  } else {
    T vl[5], vr[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const T sp = hpp ? sym_slope<T, LIM>(Up[v] - U0[v], Upp[v] - Up[v]) : T(0);
      vl[v] = U0[v] + half * s0[v];
      vr[v] = Up[v] - half * sp;
    }
    ausm_flux<T, true, DIR>(vl, vr, A, f);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) res[v] -= f[v];
}

template <typename T, int LIM, int BS>
__global__ __launch_bounds__(BS) void k_stage_cell_slope(const PatchDev* __restrict__ pds,
                                                         const TileCell* __restrict__ tiles,
                                                         const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                         T factor) {
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int nk = P.nk, njk = P.njk;
  const int c2 = t.c2 + (int)threadIdx.x;
  if (c2 >= njk) return;
  const int i = t.i;
  const int j = fdiv_idx(c2, 1.0f / (float)nk);
  const int k = c2 - j * nk;
  constexpr uint32_t W = sizeof(T);
  const uint32_t g = (P.off + (uint32_t)i * (uint32_t)njk + (uint32_t)c2) * W;
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T half = T(0.5);
  const T xc = half * dx + T(i) * dx, yc = half * dy + T(j) * dy, zc = half * dz + T(k) * dz;
  T U0[5], res[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    U0[v] = ldv<T>(sv[v], g);
    res[v] = T(0);
  }
  axis_faces<T, LIM, 0>(P, sv, g, (uint32_t)njk * W, i, P.ni, U0, dy * dz, yc, zc, j, k, res);
  axis_faces<T, LIM, 1>(P, sv, g, (uint32_t)nk * W, j, P.nj, U0, dx * dz, xc, zc, i, k, res);
  axis_faces<T, LIM, 2>(P, sv, g, W, k, nk, U0, dx * dy, xc, yc, i, j, res);
  const T pf = factor / (dx * dy * dz);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
    char* dv = reinterpret_cast<char*>(dst + (uint64_t)v * vs);
    *reinterpret_cast<T*>(dv + g) = ldv<T>(bv, g) + pf * res[v];
  }
}

// ---------------------------------------------------------------------------
// wave-tile marching stage kernel
// ---------------------------------------------------------------------------

// Every wavefront of a 256-thread block works alone (no block barrier): its 64
// lanes form an 8x8 tile of (j, k) columns starting at (j0, k0) and march over
// the planes [i0, i1).  Lane (jl, kl) owns column (j0 + jl, k0 + kl) and per
// plane computes the three faces on the minus side of its cell:
//   x face i+1/2   from a register window of its own column (the i-1/2 face is
//                  carried from the previous plane),
//   y face j-1/2   and z face k-1/2 from the plane's 11x11 halo tile, staged in
//                  a wave-private, double-buffered LDS tile,
// where a face on a patch boundary (j == 0 or NJ, k == 0 or NK, i == 0 or NI)
// is the face functor of that boundary.  The plus-side y and z faces of a cell
// are the minus-side faces of the lanes (jl+1, kl) and (jl, kl+1), fetched by
// lane shuffles; lanes with jl == 7 or kl == 7 only supply faces, so one tile
// advances 7x7 columns.  Each face is evaluated once (plus the seam faces of
// the tile and of the i range), 3 face evaluations per lane per plane.
// The residual of a cell is summed as Fx- + Fy- + Fz- - Fz+ - Fy+ - Fx+
// (cartesianiterator.h:132-308); a boundary face functor takes the place of its
// face in that sum.  Requires NI, NJ, NK >= 3 (thinner patches use k_stage_cell).
template <typename T, int RECON, int LIM, bool FAST>
__global__ __launch_bounds__(256) void k_stage_wtile(const PatchDev* __restrict__ pds,
                                                     const TileWave* __restrict__ items, int n_items,
                                                     const T* __restrict__ src, const T* base, T* dst, uint64_t vs,
                                                     T factor) {
  constexpr int TR = 11;      // tile rows / columns staged: j0-2 .. j0+8
  constexpr int RS = 12;      // LDS row stride
  constexpr int TS = TR * RS; // values per variable and buffer
  __shared__ T sm[4][2][5][TS];
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int item = (int)blockIdx.x * 4 + wave;
  if (item >= n_items) return;  // the whole wavefront leaves; no block-level barrier follows
  const TileWave t = items[item];
  const PatchDev& P = pds[t.patch];
  const int ni = P.ni, nj = P.nj, nk = P.nk, njk = P.njk;
  const int i0 = t.i0, i1 = t.i1, j0 = t.j0, k0 = t.k0;
  const int jl = lane >> 3, kl = lane & 7;
  const int j = j0 + jl, k = k0 + kl;
  const bool in_col = j < nj && k < nk;
  const bool upd = in_col && jl < 7 && kl < 7;
  const bool has_y = j <= nj && k < nk;  // a y face at j-1/2 (functor face when j == 0 or j == nj)
  const bool has_z = j < nj && k <= nk;  // a z face at k-1/2
  const T dx = T(P.dx), dy = T(P.dy), dz = T(P.dz);
  const T Ax = dy * dz, Ay = dx * dz, Az = dx * dy;
  const T pf = factor / (dx * dy * dz);
  constexpr uint32_t W = sizeof(T);
  const char* sv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) sv[v] = reinterpret_cast<const char*>(src + (uint64_t)v * vs);
  const uint32_t pb = P.off;  // first cell of the patch
  const uint32_t col = in_col ? (uint32_t)(j * nk + k) : 0u;

  // staging of plane ip: lane loads tile entries q = lane and lane + 64 (q < 121)
  T stage[2][5];
  auto load_plane = [&](int ip) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = lane + 64 * h;
      const int r = q / TR, c = q - r * TR;
      const int jj = j0 - 2 + r, kk = k0 - 2 + c;
      if (q < TR * TR && jj >= 0 && jj < nj && kk >= 0 && kk < nk) {
        const uint32_t off = (pb + (uint32_t)ip * (uint32_t)njk + (uint32_t)(jj * nk + kk)) * W;
#pragma unroll
        for (int v = 0; v < 5; ++v) stage[h][v] = ldv<T>(sv[v], off);
      }
    }
  };
  auto store_plane = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = lane + 64 * h;
      const int r = q / TR, c = q - r * TR;
      const int jj = j0 - 2 + r, kk = k0 - 2 + c;
      if (q < TR * TR && jj >= 0 && jj < nj && kk >= 0 && kk < nk) {
#pragma unroll
        for (int v = 0; v < 5; ++v) sm[wave][buf][v][r * RS + c] = stage[h][v];
      }
    }
    // the wavefront's own LDS writes complete before its reads of them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  // register window of the own column: Wn[q] = U(i - 1 + q)
  T Wn[4][5];
  auto load_col = [&](int q, int ip) {
    if (in_col && ip >= 0 && ip < ni) {
      const uint32_t off = (pb + (uint32_t)ip * (uint32_t)njk + col) * W;
#pragma unroll
      for (int v = 0; v < 5; ++v) Wn[q][v] = ldv<T>(sv[v], off);
    }
  };
  // x face between planes ia-1 and ia, window holding planes ia-2 .. ia+1
  auto xface = [&](int ia, T* f) {
    zero5(f);
    if (!in_col) return;
    const T yc = T(0.5) * dy + T(j) * dy, zc = T(0.5) * dz + T(k) * dz;
    if (ia == 0) {
      wall_flux<T, FAST>(P.bc[0], P.par[0], Wn[2], 0, false, yc, zc, j, k, Ax, f);
    } else if (ia == ni) {
      wall_flux<T, FAST>(P.bc[1], P.par[1], Wn[1], 0, true, yc, zc, j, k, Ax, f);
    } else {
      const bool hm = ia >= 2, hp = ia + 1 < ni;
      T vl[5], vr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        vl[v] = recon1<T, RECON, LIM, FAST>(Wn[1][v], Wn[0][v], Wn[2][v], hm);
        vr[v] = recon1<T, RECON, LIM, FAST>(Wn[2][v], Wn[3][v], Wn[1][v], hp);
      }
      ausm_flux<T, FAST, 0>(vl, vr, Ax, f);
    }
  };

  T fxm[5], fxp[5];
  load_col(0, i0 - 2);
  load_col(1, i0 - 1);
  load_col(2, i0);
  load_col(3, i0 + 1);
  xface(i0, fxm);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    Wn[0][v] = Wn[1][v];
    Wn[1][v] = Wn[2][v];
    Wn[2][v] = Wn[3][v];
  }
  load_col(3, i0 + 2);
  load_plane(i0);
  store_plane(0);

  for (int i = i0; i < i1; ++i) {
    const int cur = (i - i0) & 1;
    if (i + 1 < i1) load_plane(i + 1);  // in flight while this plane is computed
    const T xc = T(0.5) * dx + T(i) * dx;
    const T(*S)[TS] = sm[wave][cur];
    // y face j-1/2
    T fym[5];
    zero5(fym);
    if (has_y) {
      const int cc = kl + 2;
      if (j == 0) {
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = S[v][(jl + 2) * RS + cc];
        wall_flux<T, FAST>(P.bc[2], P.par[2], var, 1, false, xc, T(0.5) * dz + T(k) * dz, i, k, Ay, fym);
      } else if (j == nj) {
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = S[v][(jl + 1) * RS + cc];
        wall_flux<T, FAST>(P.bc[3], P.par[3], var, 1, true, xc, T(0.5) * dz + T(k) * dz, i, k, Ay, fym);
      } else {
        const bool hm = j >= 2, hp = j + 1 < nj;
        T vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T a = S[v][(jl + 1) * RS + cc], b = S[v][(jl + 2) * RS + cc];
          const T m = S[v][(hm ? jl : jl + 1) * RS + cc], q = S[v][(hp ? jl + 3 : jl + 2) * RS + cc];
          vl[v] = recon1<T, RECON, LIM, FAST>(a, m, b, hm);
          vr[v] = recon1<T, RECON, LIM, FAST>(b, q, a, hp);
        }
        ausm_flux<T, FAST, 1>(vl, vr, Ay, fym);
      }
    }
    // z face k-1/2
    T fzm[5];
    zero5(fzm);
    if (has_z) {
      const int rr = (jl + 2) * RS;
      if (k == 0) {
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = S[v][rr + kl + 2];
        wall_flux<T, FAST>(P.bc[4], P.par[4], var, 2, false, xc, T(0.5) * dy + T(j) * dy, i, j, Az, fzm);
      } else if (k == nk) {
        T var[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) var[v] = S[v][rr + kl + 1];
        wall_flux<T, FAST>(P.bc[5], P.par[5], var, 2, true, xc, T(0.5) * dy + T(j) * dy, i, j, Az, fzm);
      } else {
        const bool hm = k >= 2, hp = k + 1 < nk;
        T vl[5], vr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const T a = S[v][rr + kl + 1], b = S[v][rr + kl + 2];
          const T m = S[v][rr + (hm ? kl : kl + 1)], q = S[v][rr + (hp ? kl + 3 : kl + 2)];
          vl[v] = recon1<T, RECON, LIM, FAST>(a, m, b, hm);
          vr[v] = recon1<T, RECON, LIM, FAST>(b, q, a, hp);
        }
        ausm_flux<T, FAST, 2>(vl, vr, Az, fzm);
      }
    }
    // x face i+1/2 of the own column
    xface(i + 1, fxp);
    // plus-side y and z faces from the neighbouring lanes (all lanes take part)
    T fyp[5], fzp[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      fzp[v] = __shfl_down(fzm[v], 1, 64);
      fyp[v] = __shfl_down(fym[v], 8, 64);
    }
    if (upd) {
      const uint32_t g = (pb + (uint32_t)i * (uint32_t)njk + col) * W;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        T r = T(0);
        r += fxm[v];
        r += fym[v];
        r += fzm[v];
        r -= fzp[v];
        r -= fyp[v];
        r -= fxp[v];
        const char* bv = reinterpret_cast<const char*>(base + (uint64_t)v * vs);
        char* dv = reinterpret_cast<char*>(dst + (uint64_t)v * vs);
        *reinterpret_cast<T*>(dv + g) = ldv<T>(bv, g) + pf * r;
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      fxm[v] = fxp[v];
      Wn[0][v] = Wn[1][v];
      Wn[1][v] = Wn[2][v];
      Wn[2][v] = Wn[3][v];
    }
    load_col(3, i + 3);
    if (i + 1 < i1) store_plane(cur ^ 1);
  }
}

// plain vectorised copy (PatchIterator::copyField); also the STREAM-style
// reference kernel whose bytes are known exactly (16 B per lane per iteration)
__global__ __launch_bounds__(256) void k_copy(uint64_t n16, const uint4* __restrict__ src, uint4* __restrict__ dst) {
  for (uint64_t q = blockIdx.x * 256ull + threadIdx.x; q < n16; q += (uint64_t)gridDim.x * 256ull) dst[q] = src[q];
}

// ---------------------------------------------------------------------------
// exchange (gather): receivers are seek cells, donors are donor-zone cells of
// other patches, so the two sets never intersect and all receivers of all
// patches can be filled in one launch without ordering.
// ---------------------------------------------------------------------------

// receivers served by exactly one donor patch with `stride` contributions
template <typename T>
__global__ __launch_bounds__(256) void k_exchange_single(uint32_t n, int stride, const uint32_t* __restrict__ rec,
                                                         const uint32_t* __restrict__ idx,
                                                         const T* __restrict__ w, T* U, uint64_t vs) {
#pragma clang fp contract(off)
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  T acc[5];
  zero5(acc);
  for (int s = 0; s < stride; ++s) {
    const uint64_t d = idx[(uint64_t)r * stride + s];
    const T ww = w[(uint64_t)r * stride + s];
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[v] += U[v * vs + d] * ww;
  }
  const uint64_t g = rec[r];
#pragma unroll
  for (int v = 0; v < 5; ++v) U[v * vs + g] = T(0) + acc[v];
}

// receivers served by several donor patches, or by a turned one (and, where
// the host puts them here, single-donor receivers): donor sums added in donor
// order.  The receivers are sorted by their number of (donor cell, weight)
// entries in bands of 8, then by their first donor cell, and each wavefront's
// 64 receivers keep their entries interleaved: entry q of lane L at
// tbase[r / 64] + 64 q + L, so that a lane loops over its own entries only and
// every entry load is one coalesced access.  An entry is the donor cell index with bit 31 set where it closes
// the sum of one donor patch.  TURN: at such an entry, et = t > 0 turns the
// momentum of that donor sum into the receiver's frame with the row-major
// matrix t - 1 of `turn` before it is added (patch.cpp:617-645: inter_vars,
// turned, added).
template <typename T, bool TURN>
__global__ __launch_bounds__(256) void k_exchange_multi(uint32_t n, const uint32_t* __restrict__ rec,
                                                        const uint32_t* __restrict__ cnt,
                                                        const uint32_t* __restrict__ tbase,
                                                        const uint32_t* __restrict__ ent, const T* __restrict__ w,
                                                        T* U, uint64_t vs, const uint32_t* __restrict__ et,
                                                        const T* __restrict__ turn) {
#pragma clang fp contract(off)
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t e0 = tbase[r >> 6] + (r & 63u), ne = cnt[r];
  // uniform bases of the five variable arrays and 32-bit byte offsets (the
  // layout keeps a variable array below 2^32 bytes): one address register per
  // donor cell instead of five 64-bit ones
  constexpr uint32_t W = sizeof(T);
  const char* ub[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) ub[v] = reinterpret_cast<const char*>(U + (uint64_t)v * vs);
  T tot[5], acc[5];
  zero5(tot);
  zero5(acc);
  // one entry in order: accumulate, and at the end of a donor patch's entries
  // (turn and) add the donor sum
  auto add = [&](uint32_t e, uint32_t x, T ww, const T* u) {
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[v] += u[v] * ww;
    if (x >> 31) {
      if constexpr (TURN) {
        const uint32_t t = et[e];
        if (t > 0) {
          const T* a = turn + 9 * (uint64_t)(t - 1);
          const T m1 = acc[1], m2 = acc[2], m3 = acc[3];
          acc[1] = a[0] * m1 + a[1] * m2 + a[2] * m3;
          acc[2] = a[3] * m1 + a[4] * m2 + a[5] * m3;
          acc[3] = a[6] * m1 + a[7] * m2 + a[8] * m3;
        }
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        tot[v] += acc[v];
        acc[v] = T(0);
      }
    }
  };
  uint32_t q = 0;
  // four entries at a time: their index, weight and donor loads are issued
  // together, the sums are still formed entry after entry
  for (; q + 4 <= ne; q += 4) {
    uint32_t x[4];
    T ww[4], u[4][5];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      x[m] = ent[e0 + 64u * (q + m)];
      ww[m] = w[e0 + 64u * (q + m)];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const uint32_t b = (x[m] & 0x7fffffffu) * W;
#pragma unroll
      for (int v = 0; v < 5; ++v) u[m][v] = ldv<T>(ub[v], b);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) add(e0 + 64u * (q + m), x[m], ww[m], u[m]);
  }
  for (; q < ne; ++q) {
    const uint32_t e = e0 + 64u * q;
    const uint32_t x = ent[e];
    const uint32_t b = (x & 0x7fffffffu) * W;
    T u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = ldv<T>(ub[v], b);
    add(e, x, w[e], u);
  }
  const uint32_t g = rec[r] * W;
#pragma unroll
  for (int v = 0; v < 5; ++v) *reinterpret_cast<T*>(const_cast<char*>(ub[v]) + g) = tot[v];
}

// ---------------------------------------------------------------------------
// level-set fill: mirror gather into scratch, then the slip reflection
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_ls_access(uint32_t n, const uint32_t* __restrict__ donors,
                                                   const T* __restrict__ w, const T* __restrict__ U, uint64_t vs,
                                                   T* __restrict__ acc) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  T a[5];
  zero5(a);
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const uint64_t d = donors[(uint64_t)e * 8 + s];
    const T ww = w[(uint64_t)e * 8 + s];
#pragma unroll
    for (int v = 0; v < 5; ++v) a[v] += ww * U[v * vs + d];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) acc[(uint64_t)v * n + e] = a[v];
}

// lsobccompressibleeulerop.h:83-144 with the reference's mixed precision
// (the double literals promote the reflection and the blend to double)
template <typename T>
__global__ __launch_bounds__(256) void k_ls_operate(uint32_t n, const uint32_t* __restrict__ cells,
                                                    const T* __restrict__ nrm, const T* __restrict__ acc, T* U,
                                                    uint64_t vs, T relax) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint64_t c = cells[e];
  const T g[3] = {nrm[(uint64_t)e * 3 + 0], nrm[(uint64_t)e * 3 + 1], nrm[(uint64_t)e * 3 + 2]};
  const T a1 = acc[1 * (uint64_t)n + e], a2 = acc[2 * (uint64_t)n + e], a3 = acc[3 * (uint64_t)n + e];
  const T run = a1 * g[0] + a2 * g[1] + a3 * g[2];
  const T eps = T(1.e-6);
  const T av[3] = {a1, a2, a3};
  T out[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T hard = T(double(av[q]) - 2. * double(run) * double(g[q]));
    const T cur = U[(1 + q) * vs + c];
    const T diff = cur - hard;
    T damp = dabs(diff) / (dabs(cur) + eps);
    if (double(damp) > (1. - double(relax))) damp = T(1. - double(relax));
    out[q] = T((1. - double(damp)) * double(hard) + double(damp * cur));
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) U[(1 + q) * vs + c] = out[q];
  U[0 * vs + c] = acc[0 * (uint64_t)n + e];
  U[4 * vs + c] = acc[4 * (uint64_t)n + e];
}

// lsobccompressibleswallop.h:82-116: density and total energy from the mirror
// sum, the momentum reversed and relaxed.  In the reference's mixed precision
// relax * (-acc) is a product in the state type and (1. - relax) * rhou a
// double, whose sum is stored in the state type.
template <typename T>
__global__ __launch_bounds__(256) void k_ls_operate_swall(uint32_t n, const uint32_t* __restrict__ cells,
                                                          const T* __restrict__ acc, T* U, uint64_t vs, T relax) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint64_t c = cells[e];
  U[0 * vs + c] = acc[0 * (uint64_t)n + e];
#pragma unroll
  for (int q = 1; q <= 3; ++q) {
    const T a = acc[(uint64_t)q * n + e];
    const T r = relax * (-a);
    U[q * vs + c] = T(double(r) + (1. - double(relax)) * double(U[q * vs + c]));
  }
  U[4 * vs + c] = acc[4 * (uint64_t)n + e];
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// post-processing variables
// ---------------------------------------------------------------------------

// CompressibleVariables<PerfectGas>::getScalar / getVector
// (compressiblevariables.h:81-105) for the cells [0, n) of one patch starting
// at global cell `first`: q = 0 Ma, 1 p, 2 T, 3 rho, 4 S (entropy against
// ref_p, ref_T), 5..7 u, v, w.  The selected quantities (bit q of `mask`) are
// written in increasing q to out[slot * n + cell].
template <typename T>
__global__ __launch_bounds__(256) void k_post(uint32_t n, uint64_t first, const T* __restrict__ U, uint64_t vs,
                                              int mask, T ref_p, T ref_T, T* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  T var[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) var[v] = U[(uint64_t)v * vs + first + c];
  T p, Tt, u, v, w;
  cons_to_prim<T>(var, p, Tt, u, v, w);
  T q[8];
  q[0] = dsqrt<T>((u * u + v * v + w * w) / (Gas<T>::GAMMA * Gas<T>::R * Tt));
  q[1] = p;
  q[2] = Tt;
  q[3] = var[0];
  q[4] = Gas<T>::CP * log(Tt / ref_T) - Gas<T>::R * log(p / ref_p);
  q[5] = u;
  q[6] = v;
  q[7] = w;
  int slot = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (mask & (1 << k)) out[(uint64_t)(slot++) * n + c] = q[k];
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
  return x;
}

// per-block partial sums of the core cells of one plane chunk (256 threads)
template <typename T>
__global__ __launch_bounds__(256) void k_sums_partial(const PatchDev* __restrict__ pds, const TileCell* __restrict__ tiles,
                                                      const T* __restrict__ U, uint64_t vs, double* __restrict__ part) {
  __shared__ double red[4][5];
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int c2 = t.c2 + (int)threadIdx.x;
  double s[5] = {0, 0, 0, 0, 0};
  if (c2 < P.njk) {
    const int j = c2 / P.nk, k = c2 - j * P.nk, i = t.i;
    const bool core = i >= P.seek[0] && i < P.ni - P.seek[1] && j >= P.seek[2] && j < P.nj - P.seek[3] &&
                      k >= P.seek[4] && k < P.nk - P.seek[5];
    if (core) {
      const uint64_t c = (uint64_t)P.off + (uint64_t)i * P.njk + c2;
      const double dV = P.dx * P.dy * P.dz;
#pragma unroll
      for (int v = 0; v < 5; ++v) s[v] = double(U[v * vs + c]) * dV;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double x = wave_sum(s[v]);
    if (lane == 0) red[wv][v] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int v = threadIdx.x;
    part[(uint64_t)blockIdx.x * 5 + v] = ((red[0][v] + red[1][v]) + red[2][v]) + red[3][v];
  }
}

// fixed-order final sum of the partials (one block of 256)
__global__ __launch_bounds__(256) void k_sums_final(uint32_t nblk, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  __shared__ double red[4][5];
  double s[5] = {0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nblk; b += 256) {
#pragma unroll
    for (int v = 0; v < 5; ++v) s[v] += part[(uint64_t)b * 5 + v];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double x = wave_sum(s[v]);
    if (lane == 0) red[wv][v] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int v = threadIdx.x;
    out[v] = ((red[0][v] + red[1][v]) + red[2][v]) + red[3][v];
  }
}

// The CFL number of one cell as the jet demo computes it
// (applications/drnumJetDemo/main.h:196-205): PerfectGas::conservativeToPrimitive
// (perfectgas.h:84-92) and a = sqrt(gamma R T) in the state type, with the
// reference's double-literal promotions (1.0/rho, 0.5*|u|^2, /cv), then
// max(|u|, |u+a|, |u-a|) * dt / dx with the x velocity and spacing only.
template <typename T>
__device__ __forceinline__ T cfl_x_reference(const T* var, T dt, T dx) {
#pragma clang fp contract(off)
  const T ir = T(1.0 / double(var[0]));
  const T u = var[1] * ir, v = var[2] * ir, w = var[3] * ir;
  const T Tt = T((double(var[4] * ir) - 0.5 * double(u * u + v * v + w * w)) / double(Gas<T>::CV));
  const T a = dsqrt(Gas<T>::GAMMA * Gas<T>::R * Tt);
  T c = dabs(u) * dt / dx;
  c = fmax(c, dabs(u + a) * dt / dx);
  c = fmax(c, dabs(u - a) * dt / dx);
  return c;
}

// diagnostics partials per tile: [0] the reference CFL (x axis only), [1] min
// rho, [2] max rho, [3] count of non-finite values, [4] the CFL over all three
// axes, max_d (|u_d| + a) dt / d_d, evaluated in FP64
template <typename T>
__global__ __launch_bounds__(256) void k_diag_partial(const PatchDev* __restrict__ pds, const TileCell* __restrict__ tiles,
                                                      const T* __restrict__ U, uint64_t vs, double dt,
                                                      double* __restrict__ part) {
  __shared__ double red[4][5];
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int c2 = t.c2 + (int)threadIdx.x;
  double cflx = 0, cfl3 = 0, rmin = 1e300, rmax = -1e300, bad = 0;
  if (c2 < P.njk) {
    const uint64_t c = (uint64_t)P.off + (uint64_t)t.i * P.njk + c2;
    T tv[5];
    double var[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      tv[v] = U[v * vs + c];
      var[v] = double(tv[v]);
      if (!isfinite(var[v])) bad += 1;
    }
    cflx = double(cfl_x_reference<T>(tv, T(dt), T(P.dx)));
    if (!isfinite(cflx)) cflx = 0;
    double ir = 1.0 / var[0];
    double u = var[1] * ir, v = var[2] * ir, w = var[3] * ir;
    double Tt = (var[4] * ir - 0.5 * (u * u + v * v + w * w)) / 717.5;
    double a = sqrt(1.4 * 287.2 * Tt);
    cfl3 = fmax(fmax((fabs(u) + a) * dt / P.dx, (fabs(v) + a) * dt / P.dy), (fabs(w) + a) * dt / P.dz);
    if (!isfinite(cfl3)) cfl3 = 0;
    rmin = var[0];
    rmax = var[0];
  }
  cflx = wave_max(cflx);
  cfl3 = wave_max(cfl3);
  rmin = wave_min(rmin);
  rmax = wave_max(rmax);
  bad = wave_sum(bad);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv][0] = cflx;
    red[wv][1] = rmin;
    red[wv][2] = rmax;
    red[wv][3] = bad;
    red[wv][4] = cfl3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double o[5] = {red[0][0], red[0][1], red[0][2], red[0][3], red[0][4]};
    for (int q = 1; q < 4; ++q) {
      o[0] = fmax(o[0], red[q][0]);
      o[1] = fmin(o[1], red[q][1]);
      o[2] = fmax(o[2], red[q][2]);
      o[3] += red[q][3];
      o[4] = fmax(o[4], red[q][4]);
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) part[(uint64_t)blockIdx.x * 5 + q] = o[q];
  }
}

// Patch::computeVariableDifference (patch.cpp:733-746) over the cells of one
// patch's tiles: per cell (b - a)^2 in the state type; per block its maximum
// (exactly a state-type value) and its FP64 sum.
template <typename T>
__global__ __launch_bounds__(256) void k_vardiff_partial(const PatchDev* __restrict__ pds,
                                                         const TileCell* __restrict__ tiles, const T* __restrict__ a,
                                                         const T* __restrict__ b, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[4][2];
  const TileCell t = tiles[blockIdx.x];
  const PatchDev& P = pds[t.patch];
  const int c2 = t.c2 + (int)threadIdx.x;
  double mx = 0, sum = 0;
  if (c2 < P.njk) {
    const uint64_t c = (uint64_t)P.off + (uint64_t)t.i * P.njk + c2;
    const T d = b[c] - a[c];
    const T sq = d * d;
    mx = double(sq);
    sum = double(sq);
  }
  mx = wave_max(mx);
  sum = wave_sum(sum);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv][0] = mx;
    red[wv][1] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double o0 = red[0][0], o1 = red[0][1];
    for (int q = 1; q < 4; ++q) {
      o0 = fmax(o0, red[q][0]);
      o1 += red[q][1];
    }
    part[(uint64_t)blockIdx.x * 2 + 0] = o0;
    part[(uint64_t)blockIdx.x * 2 + 1] = o1;
  }
}

}  // namespace drn
