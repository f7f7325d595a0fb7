// elsim_sweep.h -- the 7-point operator sweep shared by the CG drivers
// (elsim_stencil.hip) and the multigrid preconditioner (elsim_mg.hip): the
// homogeneous ghost rules of each operator, the register-tiled stencil walk
// and the flat-loop macros of the pointwise kernels.
#pragma once

#include "elsim_solver.h"

namespace elsim {

template <int OP>
struct Rule;
template <>
struct Rule<ELSIM_OP_CNE> {
  static constexpr bool lo_zero = false, hi_zero = false, diag = true;
};
template <>
struct Rule<ELSIM_OP_PHP> {
  static constexpr bool lo_zero = false, hi_zero = true, diag = false;
};
template <>
struct Rule<ELSIM_OP_PHE> {
  static constexpr bool lo_zero = true, hi_zero = false, diag = false;
};
template <>
struct Rule<ELSIM_OP_POT_DD> {
  static constexpr bool lo_zero = true, hi_zero = true, diag = false;
};

// A p over one stencil tile (256 d1 voxels x SJ rows x up to SKC planes);
// epi(c_ghosted, n_interior, p_centre, value) for every interior voxel of the
// tile.  Every lane of the block must call it (the d1 neighbours are exchanged
// between lanes).  Rows past n2 + 1 and columns past n1 + 1 read the last
// ghost row/column, so the voxel next to an edge finds the ghost value in the
// neighbouring register or lane, as the non-homogeneous operator needs.
// Even d3 tiles march upward, odd ones downward: a tile then reads its halo
// plane at the same time as the neighbouring tile that owns it (in the same
// L2 or the shared last-level cache), instead of one tile-time apart, which
// cut the operator's fetched bytes from 1.13 to 1.07 times its algorithmic
// bytes at 256^3.  Per-voxel arithmetic does not depend on the direction.
// The values of p come from ld(q) (q a ghosted index), so that a caller can
// form p on the fly from other arrays (the fused CG direction update).
template <typename T, int OP, bool HOM, class Ld, class Epi>
__device__ __forceinline__ void sweep_ld(const Dims& d, const StIdx& s, Ld&& ld, const T* __restrict__ kap,
                                         const T* __restrict__ pse, T scale, Epi&& epi) {
  constexpr int SJ = Rows<T>::SJ;
  const int64_t sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  const int i = s.i;
  const bool col_ok = i <= d.n1;
  const bool down = ((s.k0 - 1) / SKC) & 1;
  const int kstart = down ? s.k1 : s.k0;
  const int64_t step = down ? -sz : sz, nstep = down ? -nz : nz;
  int64_t rc[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) rc[t] = d.gi(s.ic, min(s.j0 + t, d.n2 + 1), kstart);
  const int64_t clo = d.gi(s.ic, min(s.j0 - 1, d.n2 + 1), kstart),
                chi = d.gi(s.ic, min(s.j0 + SJ, d.n2 + 1), kstart);
  const bool ld_left = s.lane == 0 && (i > 1 || !HOM);
  const bool ld_right = s.lane == 63 && (i < d.n1 || (!HOM && i == d.n1));
  const T half = T(0.5);
  T pc[SJ], kc[SJ], pb[SJ], kb[SJ];  // plane k, the plane behind (k - direction)
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    pc[t] = ld(rc[t]);
    kc[t] = kap[rc[t]];
    kb[t] = kap[rc[t] - step];
    pb[t] = ld(rc[t] - step);
  }
  int64_t off = 0, noff = (int64_t)(i - 1) + (int64_t)d.n1 * ((s.j0 - 1) + (int64_t)d.n2 * (kstart - 1));
  const int nk = s.k1 - s.k0 + 1;
  for (int kk = 0; kk < nk; kk++, off += step, noff += nstep) {
    const int k = down ? s.k1 - kk : s.k0 + kk;
    T pa[SJ], ka[SJ], ps[SJ];  // plane ahead (k + direction)
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pa[t] = ld(rc[t] + off + step);
      ka[t] = kap[rc[t] + off + step];
      if (Rule<OP>::diag) ps[t] = pse[rc[t] + off];
    }
    const T pjlo = ld(clo + off), kjlo = kap[clo + off], pjhi = ld(chi + off), kjhi = kap[chi + off];
    T pim[SJ], pip[SJ], kim[SJ], kip[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pim[t] = lane_shift<true>(pc[t]);
      pip[t] = lane_shift<false>(pc[t]);
      kim[t] = lane_shift<true>(kc[t]);
      kip[t] = lane_shift<false>(kc[t]);
    }
    if (ld_left) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        pim[t] = ld(rc[t] + off - 1);
        kim[t] = kap[rc[t] + off - 1];
      }
    }
    if (ld_right) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        pip[t] = ld(rc[t] + off + 1);
        kip[t] = kap[rc[t] + off + 1];
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      const int j = s.j0 + t;
      const T x = pc[t];
      T pz = down ? pb[t] : pa[t], pm = down ? pa[t] : pb[t];
      const T kz = down ? kb[t] : ka[t], km = down ? ka[t] : kb[t];
      T pl = pim[t], pr = pip[t];
      T pjm = t == 0 ? pjlo : pc[t - 1];
      T pjp = t == SJ - 1 ? pjhi : pc[t + 1];
      const T kjm = t == 0 ? kjlo : kc[t - 1];
      const T kjp = t == SJ - 1 ? kjhi : kc[t + 1];
      if (HOM) {
        if (k == 1) pm = x;
        if (k == d.n3) pz = x;
        if (i == 1) pl = x;
        if (i == d.n1) pr = x;
        if (j == 1) pjm = Rule<OP>::lo_zero ? T(0) : x;
        if (j == d.n2) pjp = Rule<OP>::hi_zero ? T(0) : x;
      }
      const T k0v = kc[t];
      const T f1 = (kim[t] + k0v) * half, f2 = (kip[t] + k0v) * half;
      const T f3 = (kjm + k0v) * half, f4 = (kjp + k0v) * half;
      const T f5 = (km + k0v) * half, f6 = (kz + k0v) * half;
      T sv = f2 * (pr - x) - f1 * (x - pl);
      sv = sv + f4 * (pjp - x);
      sv = sv - f3 * (x - pjm);
      sv = sv + f6 * (pz - x);
      sv = sv - f5 * (x - pm);
      T v;
      if (Rule<OP>::diag)
        v = ps[t] * x - scale * sv;
      else
        v = scale * sv;
      if (col_ok && j <= d.n2) epi(rc[t] + off, noff + (int64_t)t * d.n1, x, v);
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pb[t] = pc[t];
      pc[t] = pa[t];
      kb[t] = kc[t];
      kc[t] = ka[t];
    }
  }
}

template <typename T, int OP, bool HOM, class Epi>
__device__ __forceinline__ void sweep(const Dims& d, const StIdx& s, const T* __restrict__ p,
                                      const T* __restrict__ kap, const T* __restrict__ pse, T scale, Epi&& epi) {
  sweep_ld<T, OP, HOM>(
      d, s, [p](int64_t q) { return p[q]; }, kap, pse, scale, epi);
}

// Pointwise walk over a block's columns (interior only).
template <class F>
__device__ __forceinline__ void for_col(const Dims& d, const ColIdx& ci, F&& f) {
  if (!ci.active) return;
  const int64_t sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  int64_t c = d.gi(ci.i, ci.j, ci.k0), n = d.ni(ci.i, ci.j, ci.k0);
  for (int k = ci.k0; k <= ci.k1; k++, c += sz, n += nz) f(c, n, k);
}

// Same with U items per thread (fewer blocks, so fewer reduction partials).
#define FLAT_LOOP_U(U, n, LOADS, BODY)                                                  \
  {                                                                                      \
    const int64_t base_ = (int64_t)blockIdx.x * (FLAT_T * (U)) + threadIdx.x;           \
    _Pragma("unroll") for (int u = 0; u < (U); u++) {                                    \
      const int64_t q = base_ + u * FLAT_T;                                              \
      if (q < (n)) { LOADS }                                                             \
    }                                                                                    \
    _Pragma("unroll") for (int u = 0; u < (U); u++) {                                    \
      const int64_t q = base_ + u * FLAT_T;                                              \
      if (q < (n)) { BODY }                                                              \
    }                                                                                    \
  }

// Flat loop over [0, n): block b owns the contiguous tile
// [b*FLAT_T*FLAT_U, (b+1)*FLAT_T*FLAT_U); a load phase then a compute/store
// phase, so the FLAT_U loads of a thread are in flight together.
#define FLAT_LOOP(n, LOADS, BODY)                                                        \
  {                                                                                      \
    const int64_t base_ = (int64_t)blockIdx.x * (FLAT_T * FLAT_U) + threadIdx.x;        \
    _Pragma("unroll") for (int u = 0; u < FLAT_U; u++) {                                 \
      const int64_t q = base_ + u * FLAT_T;                                              \
      if (q < (n)) { LOADS }                                                             \
    }                                                                                    \
    _Pragma("unroll") for (int u = 0; u < FLAT_U; u++) {                                 \
      const int64_t q = base_ + u * FLAT_T;                                              \
      if (q < (n)) { BODY }                                                              \
    }                                                                                    \
  }

}  // namespace elsim
