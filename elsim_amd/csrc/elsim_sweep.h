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
// DUAL (with HOM false): epi(c, n, p_centre, value, value_hom) also gets
// the homogeneous rule's value, computed from the same registers: equal to
// value away from the faces, and at a face voxel the arithmetic of the HOM
// sweep on the substituted neighbours -- the values of a second, HOM sweep
// over the same tile, without its loads.
template <typename T, int OP, bool HOM, class Ld, class Epi, bool DUAL = false, typename P = T>
__device__ __forceinline__ void sweep_ld(const Dims& d, const StIdx& s, Ld&& ld, const T* __restrict__ kap,
                                         const T* __restrict__ pse, T scale, Epi&& epi) {
  static_assert(!(DUAL && HOM), "the dual sweep reads the stored ghosts");
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
  P pc[SJ], pb[SJ];  // plane k, the plane behind (k - direction)
  T kc[SJ], kb[SJ];
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
    P pa[SJ];  // plane ahead (k + direction)
    T ka[SJ], ps[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pa[t] = ld(rc[t] + off + step);
      ka[t] = kap[rc[t] + off + step];
      if (Rule<OP>::diag) ps[t] = pse[rc[t] + off];
    }
    const P pjlo = ld(clo + off), pjhi = ld(chi + off);
    const T kjlo = kap[clo + off], kjhi = kap[chi + off];
    P pim[SJ], pip[SJ];
    T kim[SJ], kip[SJ];
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
      const P x = pc[t];
      P pz = down ? pb[t] : pa[t], pm = down ? pa[t] : pb[t];
      const T kz = down ? kb[t] : ka[t], km = down ? ka[t] : kb[t];
      P pl = pim[t], pr = pip[t];
      P pjm = t == 0 ? pjlo : pc[t - 1];
      P pjp = t == SJ - 1 ? pjhi : pc[t + 1];
      const T kjm = t == 0 ? kjlo : kc[t - 1];
      const T kjp = t == SJ - 1 ? kjhi : kc[t + 1];
      if (HOM) {
        if (k == 1) pm = x;
        if (k == d.n3) pz = x;
        if (i == 1) pl = x;
        if (i == d.n1) pr = x;
        if (j == 1) pjm = Rule<OP>::lo_zero ? P(T(0)) : x;
        if (j == d.n2) pjp = Rule<OP>::hi_zero ? P(T(0)) : x;
      }
      const T k0v = kc[t];
      const T f1 = (kim[t] + k0v) * half, f2 = (kip[t] + k0v) * half;
      const T f3 = (kjm + k0v) * half, f4 = (kjp + k0v) * half;
      const T f5 = (km + k0v) * half, f6 = (kz + k0v) * half;
      P sv = f2 * (pr - x) - f1 * (x - pl);
      sv = sv + f4 * (pjp - x);
      sv = sv - f3 * (x - pjm);
      sv = sv + f6 * (pz - x);
      sv = sv - f5 * (x - pm);
      P v;
      if (Rule<OP>::diag)
        v = ps[t] * x - scale * sv;
      else
        v = scale * sv;
      if constexpr (DUAL) {
        P vh = v;
        if (k == 1 || k == d.n3 || i == 1 || i == d.n1 || j == 1 || j == d.n2) {
          P hm = pm, hz = pz, hl = pl, hr = pr, hjm = pjm, hjp = pjp;
          if (k == 1) hm = x;
          if (k == d.n3) hz = x;
          if (i == 1) hl = x;
          if (i == d.n1) hr = x;
          if (j == 1) hjm = Rule<OP>::lo_zero ? P(T(0)) : x;
          if (j == d.n2) hjp = Rule<OP>::hi_zero ? P(T(0)) : x;
          P sh = f2 * (hr - x) - f1 * (x - hl);
          sh = sh + f4 * (hjp - x);
          sh = sh - f3 * (x - hjm);
          sh = sh + f6 * (hz - x);
          sh = sh - f5 * (x - hm);
          if (Rule<OP>::diag)
            vh = ps[t] * x - scale * sh;
          else
            vh = scale * sh;
        }
        if (col_ok && j <= d.n2) epi(rc[t] + off, noff + (int64_t)t * d.n1, x, v, vh);
      } else {
        if (col_ok && j <= d.n2) epi(rc[t] + off, noff + (int64_t)t * d.n1, x, v);
      }
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
// Two arrays walked together over one coefficient: the per-voxel arithmetic
// of sweep_ld applied to each component (the same operations in the same
// order as a walk over that array alone).
struct D2 {
  double a, b;
  D2() = default;
  __device__ explicit D2(double v) : a(v), b(v) {}
  __device__ D2(double x, double y) : a(x), b(y) {}
};
__device__ __forceinline__ D2 operator+(D2 p, D2 q) { return D2(p.a + q.a, p.b + q.b); }
__device__ __forceinline__ D2 operator-(D2 p, D2 q) { return D2(p.a - q.a, p.b - q.b); }
__device__ __forceinline__ D2 operator*(double f, D2 p) { return D2(f * p.a, f * p.b); }
template <bool FROM_LEFT>
__device__ __forceinline__ D2 lane_shift(D2 v) {
  return D2(lane_shift<FROM_LEFT>(v.a), lane_shift<FROM_LEFT>(v.b));
}

// sweep with the stored ghosts, epi(c, n, p_centre, value, value_hom).
template <typename T, int OP, class Epi>
__device__ __forceinline__ void sweep_dual(const Dims& d, const StIdx& s, const T* __restrict__ p,
                                           const T* __restrict__ kap, const T* __restrict__ pse, T scale,
                                           Epi&& epi) {
  auto ld = [p](int64_t q) { return p[q]; };
  sweep_ld<T, OP, false, decltype(ld)&, Epi&, true>(d, s, ld, kap, pse, scale, epi);
}

// x with the stored ghosts and y with the homogeneous rule, in one walk:
// epi(c, n, x_c, (A x)_c, y_c, (A_hom y)_c) -- the values of a sweep over x
// and a HOM sweep over y.
template <int OP, class Epi>
__device__ __forceinline__ void sweep_pair(const Dims& d, const StIdx& s, const double* __restrict__ x,
                                           const double* __restrict__ y, const double* __restrict__ kap,
                                           const double* __restrict__ pse, double scale, Epi&& epi) {
  auto ld = [x, y](int64_t q) { return D2(x[q], y[q]); };
  auto e2 = [&](int64_t c, int64_t n, D2 pc, D2 v, D2 vh) { epi(c, n, pc.a, v.a, pc.b, vh.b); };
  sweep_ld<double, OP, false, decltype(ld)&, decltype(e2)&, true, D2>(d, s, ld, kap, pse, scale, e2);
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

// ------------------------------------------ FP32 operator, 4 voxels a lane --
//
// The FP32 operator of the CG loop with four consecutive d1 voxels per lane:
// one 16-byte load per row and array (rows start at odd ghosted indices, so
// these loads are 4-byte aligned, which the hardware serves as one access).
// A d1 row segment of TW = 64, 128 or 256 voxels takes TW/4 lanes, so a wave
// covers 256/TW segments, each with its own SJ4 rows; a lane keeps its rows
// of p and kappa for the planes k and k+1 in registers and marches SKC4
// planes.  Every face flux f (p_b - p_a) is formed once and used by both
// voxels it separates: the face coefficient and the difference are the same
// IEEE operations as in sweep() (the addition commutes) and the six fluxes
// are combined in the same order, so the results are bit-identical to it.
// At 256^3 (phP, data from HBM) this takes 44.7 us against 56 us for sweep(),
// which issues 32 memory instructions per wave and plane (16 of them
// single-lane d1 edge loads) where this kernel issues 16.
//
// mv4_tile walks one tile with the values of p supplied by the caller:
// ld(q, i, j, k) returns the four values at ghosted index q, the voxels
// i..i+3 of row j, plane k (j and k may be ghost rows/planes, whose values
// the homogeneous rule discards); ld1(q, i, j, k) one value (the d1
// neighbour across a tile edge).  For every interior row of four
// voxels epi(t, j, k, q, p, Ap) receives the row slot t (0..SJ4-1), its row
// and plane, its ghosted index and the four values of p and A p.  EDGE: the
// box has more than one d1 tile, so the first and last lane of a row segment
// may load their d1 neighbour from the next tile (n1 <= 256 leaves that code,
// and its registers, out).  Every lane of a 256-thread block must call it,
// and the same number of times (a block barrier per plane: the row groups'
// boundary rows go through LDS to the neighbouring groups, so only the
// block's outer halo rows are loaded -- the FP32 down-sweep fetched 1.96
// times its algorithmic bytes when every group loaded its own).
// SJ4 = 2 rows a lane (with the row exchange): twice the blocks of 4 rows,
// 13.98 -> 13.86 ms per 256^3 step (the operator 45.5 -> 42.4 us).
constexpr int SJ4 = 2, SKC4 = 8;
__host__ __device__ inline int mv4_tw(const Dims& d) { return d.n1 <= 64 ? 64 : (d.n1 <= 128 ? 128 : 256); }
// usable when n1 is a multiple of 4 and ghosted indices fit 32 bits
inline bool mv4_ok(const Dims& d) { return d.n1 % 4 == 0 && d.G() < ((int64_t)1 << 32); }
__host__ __device__ inline TileGrid mv4_tiles(const Dims& d) {
  const int tw = mv4_tw(d), rows = 4 * (256 / tw) * SJ4;  // d2 rows per block
  return TileGrid{(d.n1 + tw - 1) / tw, (d.n2 + rows - 1) / rows, (d.n3 + SKC4 - 1) / SKC4};
}
__host__ __device__ inline int mv4_ntiles(const Dims& d) {
  const TileGrid t = mv4_tiles(d);
  return t.gx * t.gy * t.gz;
}
__device__ __forceinline__ float4 ld4(const float* __restrict__ a, uint32_t q) {
  float4 v;
  __builtin_memcpy(&v, a + q, 16);
  return v;
}
__device__ __forceinline__ void st4(float* __restrict__ a, uint32_t q, float4 v) { __builtin_memcpy(a + q, &v, 16); }
__device__ __forceinline__ void unpack4(const float4& v, float (&o)[4]) {
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
  o[3] = v.w;
}


template <int OP, bool EDGE, class LD, class LD1, class EPI>
__device__ __forceinline__ void mv4_tile(const Dims& d, int tile, LD&& ld, LD1&& ld1, const float* __restrict__ kap,
                                         const float* __restrict__ pse, float scale, EPI&& epi) {
  constexpr int SJ = SJ4;
  __shared__ float4 xch[2][256][4];  // row exchange: first / last row of p and kappa per lane
  const TileGrid tg = mv4_tiles(d);
  const int nt = tg.gx * tg.gy * tg.gz;
  const int tw = mv4_tw(d), lpr = tw / 4, rpw = 64 / lpr;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int grp = lane / lpr, x4 = lane % lpr;
  const float half = 0.5f;
  const uint32_t sz = (uint32_t)d.g12();
  const int per = nt / 8;
  int b = tile;
  if (b < per * 8) b = (b % 8) * per + b / 8;
  const int tj = b % tg.gy, ti = (b / tg.gy) % tg.gx, tk = b / (tg.gx * tg.gy);
  const int i0 = ti * tw + 4 * x4 + 1;
  const bool active = i0 <= d.n1;
  const int ia = active ? i0 : d.n1 - 3;  // idle lanes read an interior column
  const bool left_b = i0 == 1, right_b = i0 + 3 == d.n1;
  const bool ld_l = EDGE && active && x4 == 0 && i0 > 1, ld_r = EDGE && active && x4 == lpr - 1 && i0 + 3 < d.n1;
  const int j0 = ((tj * 4 + wv) * rpw + grp) * SJ + 1;
  const int k0 = tk * SKC4 + 1, k1 = min(k0 + SKC4 - 1, d.n3);
  const bool down = tk & 1;  // odd tiles march downward (see sweep_ld)
  const int kstart = down ? k1 : k0, kdir = down ? -1 : 1;
  const uint32_t step = down ? 0u - sz : sz;
  int jr[SJ];
  uint32_t rc[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    jr[t] = min(j0 + t, d.n2 + 1);
    rc[t] = (uint32_t)d.gi(ia, jr[t], kstart);
  }
  const int jlo = min(j0 - 1, d.n2 + 1), jhi = min(j0 + SJ, d.n2 + 1);
  const uint32_t clo = (uint32_t)d.gi(ia, jlo, kstart);
  const uint32_t chi = (uint32_t)d.gi(ia, jhi, kstart);
  float4 pc[SJ], kc[SJ];
  float fb[SJ][4];  // flux through the face between plane k and the plane behind it
  {
    const bool bnd = down ? kstart == d.n3 : kstart == 1;
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pc[t] = ld(rc[t], ia, jr[t], kstart);
      kc[t] = ld4(kap, rc[t]);
      float pcv[4], kcv[4], pbv[4], kbv[4];
      unpack4(pc[t], pcv);
      unpack4(kc[t], kcv);
      unpack4(ld(rc[t] - step, ia, jr[t], kstart - kdir), pbv);
      unpack4(ld4(kap, rc[t] - step), kbv);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        // upward: f5 (x - pm); downward: f6 (pz - x); pm / pz = x on the boundary plane
        const float f = (kbv[q] + kcv[q]) * half;
        const float nb = bnd ? pcv[q] : pbv[q];
        fb[t][q] = down ? f * (nb - pcv[q]) : f * (pcv[q] - nb);
      }
    }
  }
  uint32_t off = 0;
  const int nk = k1 - k0 + 1;
  for (int kk = 0; kk < nk; kk++, off += step) {
    const int k = down ? k1 - kk : k0 + kk;
    float4 pa[SJ], ka[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pa[t] = ld(rc[t] + off + step, ia, jr[t], k + kdir);
      ka[t] = ld4(kap, rc[t] + off + step);
    }
    // d2 halo rows of this plane: the last row of the row group below and
    // the first row of the group above, which those groups' lanes hold in
    // registers -- exchanged through LDS (double-buffered, one barrier per
    // plane); only the block's first and last groups load them.
    const int tid = threadIdx.x, bsel = kk & 1;
    xch[bsel][tid][0] = pc[0];
    xch[bsel][tid][1] = kc[0];
    xch[bsel][tid][2] = pc[SJ - 1];
    xch[bsel][tid][3] = kc[SJ - 1];
    const bool glo = tid < lpr, ghi = tid >= 256 - lpr;
    float4 hlp, hlk, hhp, hhk;
    if (glo) {
      hlp = ld(clo + off, ia, jlo, k);
      hlk = ld4(kap, clo + off);
    }
    if (ghi) {
      hhp = ld(chi + off, ia, jhi, k);
      hhk = ld4(kap, chi + off);
    }
    __syncthreads();
    if (!glo) {
      hlp = xch[bsel][tid - lpr][2];
      hlk = xch[bsel][tid - lpr][3];
    }
    if (!ghi) {
      hhp = xch[bsel][tid + lpr][0];
      hhk = xch[bsel][tid + lpr][1];
    }
    float psv[SJ][4];
    if (Rule<OP>::diag) {
#pragma unroll
      for (int t = 0; t < SJ; t++) unpack4(ld4(pse, rc[t] + off), psv[t]);
    }
    float elp[SJ], elk[SJ], erp[SJ], erk[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) elp[t] = elk[t] = erp[t] = erk[t] = 0.0f;
    if (EDGE && ld_l) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        elp[t] = ld1(rc[t] + off - 1, ia - 1, jr[t], k);
        elk[t] = kap[rc[t] + off - 1];
      }
    }
    if (EDGE && ld_r) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        erp[t] = ld1(rc[t] + off + 4, ia + 4, jr[t], k);
        erk[t] = kap[rc[t] + off + 4];
      }
    }
    const bool kbnd = down ? k == 1 : k == d.n3;  // the face ahead is a boundary face
    // d2 faces: jf[t] lies between rows j0 + t - 1 and j0 + t
    float jf[SJ + 1][4];
#pragma unroll
    for (int t = 0; t <= SJ; t++) {
      float plv[4], klv[4], puv[4], kuv[4];
      unpack4(t == 0 ? hlp : pc[t - 1], plv);
      unpack4(t == 0 ? hlk : kc[t - 1], klv);
      unpack4(t == SJ ? hhp : pc[t], puv);
      unpack4(t == SJ ? hhk : kc[t], kuv);
      const int jl = j0 + t - 1;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float f = (klv[q] + kuv[q]) * half;
        float lo = plv[q], hi = puv[q];
        if (jl == 0) lo = Rule<OP>::lo_zero ? 0.0f : hi;     // below row 1
        if (jl == d.n2) hi = Rule<OP>::hi_zero ? 0.0f : lo;  // above row n2
        jf[t][q] = f * (hi - lo);
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      const int j = j0 + t;
      float pv[4], kv[4], pav[4], kav[4];
      unpack4(pc[t], pv);
      unpack4(kc[t], kv);
      unpack4(pa[t], pav);
      unpack4(ka[t], kav);
      float pL = lane_shift<true>(pv[3]), kL = lane_shift<true>(kv[3]);
      float pR = lane_shift<false>(pv[0]), kR = lane_shift<false>(kv[0]);
      if (EDGE && ld_l) {
        pL = elp[t];
        kL = elk[t];
      }
      if (EDGE && ld_r) {
        pR = erp[t];
        kR = erk[t];
      }
      // d1 faces: xf[q] lies between voxels q - 1 and q
      float xf[5];
#pragma unroll
      for (int q = 0; q <= 4; q++) {
        const float kl = q == 0 ? kL : kv[q - 1], kr = q == 4 ? kR : kv[q];
        float pl = q == 0 ? pL : pv[q - 1], pr = q == 4 ? pR : pv[q];
        if (q == 0 && left_b) pl = pr;
        if (q == 4 && right_b) pr = pl;
        xf[q] = ((kl + kr) * half) * (pr - pl);
      }
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float xv = pv[q];
        const float f = (kav[q] + kv[q]) * half;
        const float nb = kbnd ? xv : pav[q];
        const float fa = down ? f * (xv - nb) : f * (nb - xv);
        const float fup = down ? fb[t][q] : fa, flo = down ? fa : fb[t][q];
        float sv = xf[q + 1] - xf[q];
        sv = sv + jf[t + 1][q];
        sv = sv - jf[t][q];
        sv = sv + fup;
        sv = sv - flo;
        float v;
        if (Rule<OP>::diag)
          v = psv[t][q] * xv - scale * sv;
        else
          v = scale * sv;
        o[q] = v;
        fb[t][q] = fa;
      }
      if (active && j <= d.n2) epi(t, j, k, rc[t] + off, pv, o);
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pc[t] = pa[t];
      kc[t] = ka[t];
    }
  }
}

// ------------------------------------------ FP64 operator, 2 voxels a lane --
//
// The FP64 7-point sweep with two consecutive d1 voxels per lane (one 16-byte
// load per row and array, 8-byte aligned) and the tile walk of mv4_tile:
// row groups exchange their boundary rows through LDS, a lane marches SKC2
// planes with its SJ2 rows of p and kappa for planes k and k+1 in registers,
// and every face flux is formed once.  HOM: the homogeneous ghost rule of
// sweep<HOM = true>; otherwise the stored ghosts, as sweep<HOM = false>
// (lanes at the box's d1 edges load the ghost column).  Per voxel the same
// IEEE operations in the same order as sweep(), so the same values.
// epi(c, n, x, v) for every interior voxel (ghosted and interior index, p and
// A p there), as sweep's.  A segment of TW = 64 or 128 d1 voxels takes TW/2
// lanes.  Launch: mv2_ntiles(d) blocks of 256 threads, every lane calling it
// the same number of times (a block barrier per plane).
constexpr int SJ2 = 2, SKC2 = 8;
__host__ __device__ inline int mv2_tw(const Dims& d) { return d.n1 <= 64 ? 64 : 128; }
// usable when n1 is even and ghosted indices fit 32 bits
__host__ __device__ inline bool mv2_ok(const Dims& d) { return d.n1 % 2 == 0 && d.G() < ((int64_t)1 << 32); }
__host__ __device__ inline TileGrid mv2_tiles(const Dims& d) {
  const int tw = mv2_tw(d), rows = 4 * (128 / tw) * SJ2;  // d2 rows per block
  return TileGrid{(d.n1 + tw - 1) / tw, (d.n2 + rows - 1) / rows, (d.n3 + SKC2 - 1) / SKC2};
}
__host__ __device__ inline int mv2_ntiles(const Dims& d) {
  const TileGrid t = mv2_tiles(d);
  return t.gx * t.gy * t.gz;
}
__device__ __forceinline__ void ld2(const double* __restrict__ a, uint32_t q, double (&o)[2]) {
  __builtin_memcpy(o, a + q, 16);
}
// The row-exchange buffer of mv2_tile: one array for all of a kernel's walks
// (a kernel may walk with and without the homogeneous rule, or both edge
// variants), 2 x 256 lanes x (p, kappa of the first and the last row).
typedef double Xch64[256][8];
__device__ __forceinline__ Xch64* xch64() {
  __shared__ double buf[2][256][8];
  return buf;
}

template <int OP, bool HOM, bool EDGE, class EPI>
__device__ __forceinline__ void mv2_tile(const Dims& d, int tile, const double* __restrict__ p,
                                         const double* __restrict__ kap, const double* __restrict__ pse,
                                         double scale, EPI&& epi) {
  constexpr int SJ = SJ2;
  Xch64* xch = xch64();
  const TileGrid tg = mv2_tiles(d);
  const int nt = tg.gx * tg.gy * tg.gz;
  const int tw = mv2_tw(d), lpr = tw / 2, rpw = 64 / lpr;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int grp = lane / lpr, x2 = lane % lpr;
  const double half = 0.5;
  const uint32_t sz = (uint32_t)d.g12();
  const int per = nt / 8;
  int b = tile;
  if (b < per * 8) b = (b % 8) * per + b / 8;
  const int tj = b % tg.gy, ti = (b / tg.gy) % tg.gx, tk = b / (tg.gx * tg.gy);
  const int i0 = ti * tw + 2 * x2 + 1;
  const bool active = i0 <= d.n1;
  const int ia = active ? i0 : d.n1 - 1;  // idle lanes read an interior column
  const bool left_b = i0 == 1, right_b = i0 + 1 == d.n1;
  // the d1 neighbour across the lane's ends: from the next tile (EDGE) or,
  // with stored ghosts, the ghost column; a lane shift otherwise
  const bool ld_l = active && ((EDGE && x2 == 0 && i0 > 1) || (!HOM && left_b));
  const bool ld_r = active && ((EDGE && x2 == lpr - 1 && i0 + 1 < d.n1) || (!HOM && right_b));
  const int j0 = ((tj * 4 + wv) * rpw + grp) * SJ + 1;
  const int k0 = tk * SKC2 + 1, k1 = min(k0 + SKC2 - 1, d.n3);
  const bool down = tk & 1;  // odd tiles march downward (see sweep_ld)
  const int kstart = down ? k1 : k0, kdir = down ? -1 : 1;
  const uint32_t step = down ? 0u - sz : sz;
  int jr[SJ];
  uint32_t rc[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    jr[t] = min(j0 + t, d.n2 + 1);
    rc[t] = (uint32_t)d.gi(ia, jr[t], kstart);
  }
  const int jlo = min(j0 - 1, d.n2 + 1), jhi = min(j0 + SJ, d.n2 + 1);
  const uint32_t clo = (uint32_t)d.gi(ia, jlo, kstart);
  const uint32_t chi = (uint32_t)d.gi(ia, jhi, kstart);
  double pc[SJ][2], kc[SJ][2];
  double fb[SJ][2];  // flux through the face between plane k and the plane behind it
  {
    const bool bnd = down ? kstart == d.n3 : kstart == 1;
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      ld2(p, rc[t], pc[t]);
      ld2(kap, rc[t], kc[t]);
      double pbv[2], kbv[2];
      ld2(p, rc[t] - step, pbv);
      ld2(kap, rc[t] - step, kbv);
#pragma unroll
      for (int q = 0; q < 2; q++) {
        // upward: f5 (x - pm); downward: f6 (pz - x); pm / pz = x on the boundary plane (HOM)
        const double f = (kbv[q] + kc[t][q]) * half;
        const double nb = (HOM && bnd) ? pc[t][q] : pbv[q];
        fb[t][q] = down ? f * (nb - pc[t][q]) : f * (pc[t][q] - nb);
      }
    }
  }
  uint32_t off = 0;
  const int64_t nz = (int64_t)d.n1 * d.n2;
  const int nk = k1 - k0 + 1;
  for (int kk = 0; kk < nk; kk++, off += step) {
    const int k = down ? k1 - kk : k0 + kk;
    double pa[SJ][2], ka[SJ][2];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      ld2(p, rc[t] + off + step, pa[t]);
      ld2(kap, rc[t] + off + step, ka[t]);
    }
    // d2 halo rows: the neighbouring row groups' boundary rows through LDS,
    // the block's outer ones from memory
    const int bsel = kk & 1;
    xch[bsel][tid][0] = pc[0][0];
    xch[bsel][tid][1] = pc[0][1];
    xch[bsel][tid][2] = kc[0][0];
    xch[bsel][tid][3] = kc[0][1];
    xch[bsel][tid][4] = pc[SJ - 1][0];
    xch[bsel][tid][5] = pc[SJ - 1][1];
    xch[bsel][tid][6] = kc[SJ - 1][0];
    xch[bsel][tid][7] = kc[SJ - 1][1];
    const bool glo = tid < lpr, ghi = tid >= 256 - lpr;
    double hlp[2], hlk[2], hhp[2], hhk[2];
    if (glo) {
      ld2(p, clo + off, hlp);
      ld2(kap, clo + off, hlk);
    }
    if (ghi) {
      ld2(p, chi + off, hhp);
      ld2(kap, chi + off, hhk);
    }
    double psv[SJ][2];
    if (Rule<OP>::diag) {
#pragma unroll
      for (int t = 0; t < SJ; t++) ld2(pse, rc[t] + off, psv[t]);
    }
    double elp[SJ], elk[SJ], erp[SJ], erk[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) elp[t] = elk[t] = erp[t] = erk[t] = 0.0;
    if (ld_l) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        elp[t] = p[rc[t] + off - 1];
        elk[t] = kap[rc[t] + off - 1];
      }
    }
    if (ld_r) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        erp[t] = p[rc[t] + off + 2];
        erk[t] = kap[rc[t] + off + 2];
      }
    }
    __syncthreads();
    if (!glo) {
      hlp[0] = xch[bsel][tid - lpr][4];
      hlp[1] = xch[bsel][tid - lpr][5];
      hlk[0] = xch[bsel][tid - lpr][6];
      hlk[1] = xch[bsel][tid - lpr][7];
    }
    if (!ghi) {
      hhp[0] = xch[bsel][tid + lpr][0];
      hhp[1] = xch[bsel][tid + lpr][1];
      hhk[0] = xch[bsel][tid + lpr][2];
      hhk[1] = xch[bsel][tid + lpr][3];
    }
    const bool kbnd = down ? k == 1 : k == d.n3;  // the face ahead is a boundary face
    // d2 faces: jf[t] lies between rows j0 + t - 1 and j0 + t
    double jf[SJ + 1][2];
#pragma unroll
    for (int t = 0; t <= SJ; t++) {
      const int jl = j0 + t - 1;
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const double plq = t == 0 ? hlp[q] : pc[t - 1][q], klq = t == 0 ? hlk[q] : kc[t - 1][q];
        const double puq = t == SJ ? hhp[q] : pc[t][q], kuq = t == SJ ? hhk[q] : kc[t][q];
        const double f = (klq + kuq) * half;
        double lo = plq, hi = puq;
        if (HOM && jl == 0) lo = Rule<OP>::lo_zero ? 0.0 : hi;     // below row 1
        if (HOM && jl == d.n2) hi = Rule<OP>::hi_zero ? 0.0 : lo;  // above row n2
        jf[t][q] = f * (hi - lo);
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      const int j = j0 + t;
      double pL = lane_shift<true>(pc[t][1]), kL = lane_shift<true>(kc[t][1]);
      double pR = lane_shift<false>(pc[t][0]), kR = lane_shift<false>(kc[t][0]);
      if (ld_l) {
        pL = elp[t];
        kL = elk[t];
      }
      if (ld_r) {
        pR = erp[t];
        kR = erk[t];
      }
      // d1 faces: xf[q] lies between voxels q - 1 and q
      double xf[3];
#pragma unroll
      for (int q = 0; q <= 2; q++) {
        const double kl = q == 0 ? kL : kc[t][q - 1], kr = q == 2 ? kR : kc[t][q];
        double pl = q == 0 ? pL : pc[t][q - 1], pr = q == 2 ? pR : pc[t][q];
        if (HOM && q == 0 && left_b) pl = pr;
        if (HOM && q == 2 && right_b) pr = pl;
        xf[q] = ((kl + kr) * half) * (pr - pl);
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const double xv = pc[t][q];
        const double f = (ka[t][q] + kc[t][q]) * half;
        const double nb = (HOM && kbnd) ? xv : pa[t][q];
        const double fa = down ? f * (xv - nb) : f * (nb - xv);
        const double fup = down ? fb[t][q] : fa, flo = down ? fa : fb[t][q];
        double sv = xf[q + 1] - xf[q];
        sv = sv + jf[t + 1][q];
        sv = sv - jf[t][q];
        sv = sv + fup;
        sv = sv - flo;
        double v;
        if (Rule<OP>::diag)
          v = psv[t][q] * xv - scale * sv;
        else
          v = scale * sv;
        fb[t][q] = fa;
        if (active && j <= d.n2)
          epi((int64_t)(rc[t] + off) + q, (int64_t)(ia - 1 + q) + (int64_t)d.n1 * (j - 1) + nz * (k - 1), xv, v);
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++)
#pragma unroll
      for (int q = 0; q < 2; q++) {
        pc[t][q] = pa[t][q];
        kc[t][q] = ka[t][q];
      }
  }
}

// Tiles of the FP64 operator (k_matvec<double>): the two-voxels-a-lane walk
// where the box allows it, the one-voxel sweep() otherwise.  ntiles64(d)
// blocks of 256 threads; tile64 calls epi(c, n, x, v) per interior voxel
// either way.  The FP64 kernels of the fast path's set-up keep sweep(): the
// walk is 8-18 % faster there too, but it sums their reductions in another
// order, and the coupled, controlled step amplifies such last-bit changes
// into another trajectory (at 256^3, seed 1, 48 steps: 11.3 phP iterations a
// step against 9.7; seeds 2 and 3 unchanged), which a benchmark window of
// fixed length cannot average out.
__host__ __device__ inline int ntiles64(const Dims& d) {
  return mv2_ok(d) ? mv2_ntiles(d) : stencil_ntiles<double>(d);
}
template <int OP, bool HOM, class EPI>
__device__ __forceinline__ void tile64(const Dims& d, int tile, const double* __restrict__ p,
                                       const double* __restrict__ kap, const double* __restrict__ pse, double scale,
                                       EPI&& epi) {
  if (mv2_ok(d)) {
    if (mv2_tiles(d).gx > 1)
      mv2_tile<OP, HOM, true>(d, tile, p, kap, pse, scale, epi);
    else
      mv2_tile<OP, HOM, false>(d, tile, p, kap, pse, scale, epi);
  } else {
    sweep<double, OP, HOM>(d, stencil_idx<Rows<double>::SJ>(d, tile), p, kap, pse, scale, epi);
  }
}

// dinv = 1 / diag(A_hom) (FP64), stored as T (uint16_t: BF16).  Mirror
// faces do not contribute; zero-Dirichlet faces do.  Interior of a ghosted
// array.  kap32 / pse32 (optional): FP32 copies of the interior of kap / pse,
// read here anyway (their ghosts: k_ghost_f32).
template <int OP, typename T>
__global__ __launch_bounds__(256) void k_dinv(Dims d, const double* __restrict__ kap, const double* __restrict__ pse,
                                              double scale, T* __restrict__ dinv,
                                              float* __restrict__ kap32 = nullptr,
                                              float* __restrict__ pse32 = nullptr) {
  const ColIdx ci = col_idx(d);
  if (!ci.active) return;
  const int64_t sx = d.g1(), sz = d.g12();
  // the column's d3 neighbours roll through registers
  int64_t c = d.gi(ci.i, ci.j, ci.k0);
  double kb = kap[c - sz], kc = kap[c];
  for (int k = ci.k0; k <= ci.k1; k++, c += sz) {
    const double ka = kap[c + sz];
    if (kap32) kap32[c] = (float)kc;
    if (pse32) pse32[c] = (float)pse[c];
    double s = 0.0;
    if (ci.i > 1) s += (kap[c - 1] + kc) * 0.5;
    if (ci.i < d.n1) s += (kap[c + 1] + kc) * 0.5;
    if (ci.j > 1 || Rule<OP>::lo_zero) s += (kap[c - sx] + kc) * 0.5;
    if (ci.j < d.n2 || Rule<OP>::hi_zero) s += (kap[c + sx] + kc) * 0.5;
    if (k > 1) s += (kb + kc) * 0.5;
    if (k < d.n3) s += (ka + kc) * 0.5;
    const double dg = Rule<OP>::diag ? pse[c] + scale * s : -scale * s;
    if constexpr (std::is_same<T, uint16_t>::value)
      dinv[c] = f32_to_bf16((float)(1.0 / dg));
    else
      dinv[c] = (T)(1.0 / dg);
    kb = kc;
    kc = ka;
  }
}

}  // namespace elsim
