// elsim_connect.hip -- phase connectivity on the GPU.
//
// 1. Accelerated diffusion (AccDiff_3D.f90:83-243), FP64, bit-identical to the
//    reference: Jacobi update where psi >= 0.4f, the every-5th-iteration snap
//    to 1 where Cn > 1e-4f and psi > 0.4f, Neumann ghosts and the Dirichlet
//    d2 face (value 0 in the first iteration, when the reference's ghost is
//    still unset, then 1).  The reference's quirk that the electrolyte pass
//    replaces only the interior psi (AD:175) -- its ghosts keep the solid
//    values -- is reproduced: one ghosted psi array serves both phases, the
//    electrolyte interior being evaluated as 1 - psi on the fly.
//    24 B/voxel/iteration: Cn in (stencil), psi in (stencil), Cn out.
// 2. Connected-component labelling of {psi >= 0.4f} (6-connectivity) by
//    parallel union-find with atomicMin links from the larger root to the
//    smaller (Playne & Hawick, IEEE TPDS 2018), so every component's label is
//    its smallest voxel index whatever the schedule; components touching the
//    Dirichlet face are the connected set that AD converges to.
#include "elsim_common.h"

namespace elsim {
namespace {

constexpr double kThrUpd = (double)4.0e-1f;   // AD:118
constexpr double kThrSnap = (double)1.0e-4f;  // AD:133, RM:80

// psi = 0.5(1+tanh(d/0.75)) (AD:84-97)
__device__ __forceinline__ double ad_psi(double rd) { return 0.5 * (1.0 + tanh(rd / 0.75)); }

// psi over the ghosted box: interior from d, mirrored ghosts (AD:84-97)
__global__ void k_ad_psi(Dims d, const double* __restrict__ rdin, double* __restrict__ psi) {
  const int64_t G = d.G();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < G; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % d.g1()), j = (int)((q / d.g1()) % (d.n2 + 2)), k = (int)(q / d.g12());
    const int ic = min(max(i, 1), d.n1), jc = min(max(j, 1), d.n2), kc = min(max(k, 1), d.n3);
    psi[q] = ad_psi(rdin[d.ni(ic, jc, kc)]);
  }
}

template <int PHASE>
__device__ __forceinline__ double ps_int(const double* psi, int64_t q) {
  return PHASE == 1 ? 1.0 - psi[q] : psi[q];
}

// One AD iteration: cn (interior) -> cn_new (interior), on stencil tiles of
// 256 d1 voxels x AD_SJ rows x up to 16 planes per block:
// each thread keeps its rows of psi and Cn for the planes k-1, k, k+1 in
// registers, d1 neighbours come from the adjacent lanes (DPP), and only the
// rows beside the tile and the wave-edge lanes load neighbours from memory.
// Registers hold psi as stored; a neighbour inside the box is taken as this
// phase's value (1 - psi for the electrolyte), one in the ghost layer as
// stored.  Cn has no ghost layer: its loads are clamped into the box, and a
// neighbour across a face is replaced by the face rule (mirror, or bval on
// the Dirichlet face) before use.  Per-voxel arithmetic and its order are
// those of the reference loop body (AD:103-148, 175-222).
template <int PHASE>
__device__ __forceinline__ double ph(double raw) {
  return PHASE == 1 ? 1.0 - raw : raw;
}

// 4 rows per thread: 877 us per iteration at 512^3 against 946 with 2 rows,
// 892 with 3 and 888 with 8; marching odd plane tiles downward (as the CG
// operator does) measured slower here (915 us with 4 rows).
constexpr int AD_SJ = 4;

// Skipping unchanged tiles.  Each block records in a flag word per tile
// whether any value of its tile changed (bit 0) and whether any value could
// still be snapped, psi > 0.4f and 1e-4f < Cn != 1 (bit 1).  When the
// previous iteration did not snap and had the Dirichlet value of this one, a
// tile whose values and whose six neighbour tiles' values did not change in
// the previous iteration gets the same update values again -- the map is
// applied to the same inputs -- and, if this iteration snaps, keeps them
// unless one can be snapped.  The output buffer already holds those values
// (it was the previous iteration's input), so such a block only rewrites its
// flag word: unchanged, snappable as before.
__device__ __forceinline__ int ad_tile_id(const TileGrid& t, int ti, int tj, int tk) {
  return tj + t.gy * (ti + t.gx * tk);
}

template <int PHASE>
__global__ __launch_bounds__(256) void k_ad_iter(Dims d, const double* __restrict__ psi, const double* __restrict__ cn,
                                                 double* __restrict__ cnn, double bval, int snap,
                                                 const int* __restrict__ chg_r, int* __restrict__ chg_w,
                                                 int may_skip) {
  constexpr int CHANGED = 1, SNAPPABLE = 2;
  constexpr int SJ = AD_SJ;
  const StIdx st = stencil_idx<SJ>(d, blockIdx.x);
  // this block's tile in the tile grid
  const TileGrid tgr = stencil_tiles(d, SJ);
  const TileXyz tc = stencil_tile(d, SJ, blockIdx.x);
  const int ti = tc.ti, tj = tc.tj, tk = tc.tk;
  const int my_tile = ad_tile_id(tgr, ti, tj, tk);
  if (may_skip) {
    const int own = chg_r[my_tile];
    int any = own & CHANGED;
    if (ti > 0) any |= chg_r[ad_tile_id(tgr, ti - 1, tj, tk)] & CHANGED;
    if (ti + 1 < tgr.gx) any |= chg_r[ad_tile_id(tgr, ti + 1, tj, tk)] & CHANGED;
    if (tj > 0) any |= chg_r[ad_tile_id(tgr, ti, tj - 1, tk)] & CHANGED;
    if (tj + 1 < tgr.gy) any |= chg_r[ad_tile_id(tgr, ti, tj + 1, tk)] & CHANGED;
    if (tk > 0) any |= chg_r[ad_tile_id(tgr, ti, tj, tk - 1)] & CHANGED;
    if (tk + 1 < tgr.gz) any |= chg_r[ad_tile_id(tgr, ti, tj, tk + 1)] & CHANGED;
    if (snap) any |= own & SNAPPABLE;
    if (!any) {
      if (threadIdx.x == 0) chg_w[my_tile] = own & SNAPPABLE;
      return;
    }
  }
  int changed = 0, snappable = 0;
  const int i = st.i;
  const bool col_ok = i <= d.n1;
  const int icn = min(st.ic, d.n1);  // Cn column (lanes past the edge read the last one)
  const int kstart = st.k0;
  const int64_t sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  int64_t rq[SJ], rn[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    rq[t] = d.gi(st.ic, min(st.j0 + t, d.n2 + 1), kstart);
    rn[t] = d.ni(icn, min(st.j0 + t, d.n2), kstart);
  }
  const int64_t qlo = d.gi(st.ic, min(st.j0 - 1, d.n2 + 1), kstart),
                qhi = d.gi(st.ic, min(st.j0 + SJ, d.n2 + 1), kstart);
  const int64_t nlo = d.ni(icn, min(max(st.j0 - 1, 1), d.n2), kstart), nhi = d.ni(icn, min(st.j0 + SJ, d.n2), kstart);
  const bool ld_left = st.lane == 0 && col_ok, ld_right = st.lane == 63 && col_ok;
  const bool has_behind = kstart > 1;  // plane k0 - 1 is in the box
  double pc[SJ], pb[SJ], cc[SJ], cb[SJ];  // planes k and k - 1
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    pc[t] = psi[rq[t]];
    pb[t] = psi[rq[t] - sz];
    cc[t] = cn[rn[t]];
    cb[t] = has_behind ? cn[rn[t] - nz] : cc[t];
  }
  const double dt = 0.0625, idh2 = 1.0;
  int64_t oq = 0, on = 0;
  for (int k = st.k0; k <= st.k1; k++, oq += sz, on += nz) {
    double pa[SJ], ca[SJ];  // plane k + 1
    const bool has_ahead = k < d.n3;
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pa[t] = psi[rq[t] + oq + sz];
      ca[t] = has_ahead ? cn[rn[t] + on + nz] : cc[t];
    }
    const double pjlo = psi[qlo + oq], pjhi = psi[qhi + oq], cjlo = cn[nlo + on], cjhi = cn[nhi + on];
    double pl[SJ], pr[SJ], cl[SJ], cr[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pl[t] = lane_shift<true>(pc[t]);
      pr[t] = lane_shift<false>(pc[t]);
      cl[t] = lane_shift<true>(cc[t]);
      cr[t] = lane_shift<false>(cc[t]);
    }
    if (ld_left) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        pl[t] = psi[rq[t] + oq - 1];
        cl[t] = i > 1 ? cn[rn[t] + on - 1] : cc[t];
      }
    }
    if (ld_right) {
#pragma unroll
      for (int t = 0; t < SJ; t++) {
        pr[t] = psi[rq[t] + oq + 1];
        cr[t] = i < d.n1 ? cn[rn[t] + on + 1] : cc[t];
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      const int j = st.j0 + t;
      if (!col_ok || j > d.n2) continue;
      const double p = ph<PHASE>(pc[t]);
      const double c = cc[t];
      double v = c;
      if (p >= kThrUpd) {
        const double pyp = i < d.n1 ? ph<PHASE>(pr[t]) : pr[t];
        const double pym = i > 1 ? ph<PHASE>(pl[t]) : pl[t];
        const double pxr = t == SJ - 1 ? pjhi : pc[t + 1], pxl = t == 0 ? pjlo : pc[t - 1];
        const double pxp = j < d.n2 ? ph<PHASE>(pxr) : pxr;
        const double pxm = j > 1 ? ph<PHASE>(pxl) : pxl;
        const double pzp = k < d.n3 ? ph<PHASE>(pa[t]) : pa[t];
        const double pzm = k > 1 ? ph<PHASE>(pb[t]) : pb[t];
        const double cyp = i < d.n1 ? cr[t] : c;
        const double cym = i > 1 ? cl[t] : c;
        const double cxr = t == SJ - 1 ? cjhi : cc[t + 1], cxl = t == 0 ? cjlo : cc[t - 1];
        const double cxp = j < d.n2 ? cxr : (PHASE == 0 ? bval : c);
        const double cxm = j > 1 ? cxl : (PHASE == 1 ? bval : c);
        const double czp = k < d.n3 ? ca[t] : c;
        const double czm = k > 1 ? cb[t] : c;
        double s = (pyp + p) * (cyp - c);
        s = s - (p + pym) * (c - cym);
        s = s + (pxp + p) * (cxp - c);
        s = s - (p + pxm) * (c - cxm);
        s = s + (pzp + p) * (czp - c);
        s = s - (p + pzm) * (c - czm);
        const double div = 0.5 * idh2 / p * s;
        v = c + dt * div;
      }
      if (snap && v > kThrSnap && p > kThrUpd) v = 1.0;
      changed |= __double_as_longlong(v) != __double_as_longlong(c);
      snappable |= p > kThrUpd && v > kThrSnap && v != 1.0;
      cnn[rn[t] + on] = v;
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      pb[t] = pc[t];
      pc[t] = pa[t];
      cb[t] = cc[t];
      cc[t] = ca[t];
    }
  }
  changed = __syncthreads_or(changed);
  snappable = __syncthreads_or(snappable);
  if (threadIdx.x == 0) chg_w[my_tile] = (changed ? CHANGED : 0) | (snappable ? SNAPPABLE : 0);
}

// Connected-component labels (the threshold step of RM_UCNCT_3D, RM:79-82,
// labels instead of an AD front): every in-phase voxel gets the smallest
// interior index of its 6-connected component.  Three passes:
//  1. a one-byte phase mask (psi read once);
//  2. union-find inside 64 x 8 x 8 tiles in LDS: every component's tile-local
//     root is its smallest local index, which is also its smallest interior
//     index (both orders are (k, j, i) lexicographic), and every voxel's
//     label is set to that root;
//  3. unions across the three lower faces of every tile on the global
//     labels (roots are only ever linked to smaller roots, so the smallest
//     voxel of a component ends as its root, whatever the order of the
//     atomics), then every label is replaced by its root.
// At 512^3 the single-pass form (a global union for every in-phase voxel and
// each of its three lower neighbours, phase read three times in FP64 per
// voxel) took 27 ms per phase.
constexpr int LTX = 64, LTY = 8, LTZ = 8, LTV = LTX * LTY * LTZ;

// phase mask over the interior, straight from the distance (the same psi as
// k_ad_psi's interior, without storing it)
template <int PHASE>
__global__ __launch_bounds__(256) void k_lab_mask(int64_t N, const double* __restrict__ rdin, uint8_t* __restrict__ m) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x) {
    const double p = ad_psi(rdin[q]);
    m[q] = (PHASE == 1 ? 1.0 - p : p) >= kThrUpd;
  }
}

template <int SCOPE>
__device__ __forceinline__ int uf_find_s(int* L, int x) {
  int p = __hip_atomic_load(&L[x], __ATOMIC_RELAXED, SCOPE);
  while (p != x) {
    x = p;
    p = __hip_atomic_load(&L[x], __ATOMIC_RELAXED, SCOPE);
  }
  return x;
}

// link the roots of a and b, the larger to the smaller
template <int SCOPE>
__device__ __forceinline__ void uf_union(int* L, int a, int b) {
  bool done = false;
  while (!done) {
    a = uf_find_s<SCOPE>(L, a);
    b = uf_find_s<SCOPE>(L, b);
    if (a < b) {
      const int old = atomicMin(&L[b], a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(&L[a], b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  }
}

struct LabTiles {
  int gx, gy, gz;
};
__host__ __device__ inline LabTiles lab_tiles(const Dims& d) {
  return LabTiles{(d.n1 + LTX - 1) / LTX, (d.n2 + LTY - 1) / LTY, (d.n3 + LTZ - 1) / LTZ};
}

// A wave owns whole 64-voxel tile rows: the d1 runs of a row come from one
// ballot (every voxel starts at the first voxel of its run, so no d1 unions
// are needed), and a d2 or d3 union is skipped when the voxel's d1
// predecessor and that predecessor's neighbour are both in phase (the same
// two runs are joined by the predecessor's union).
__global__ __launch_bounds__(256) void k_lab_local(Dims d, const uint8_t* __restrict__ m, int* __restrict__ L) {
  __shared__ int lab[LTV];
  const LabTiles t = lab_tiles(d);
  const int tx = blockIdx.x % t.gx, ty = (blockIdx.x / t.gx) % t.gy, tz = blockIdx.x / (t.gx * t.gy);
  const int i0 = tx * LTX, j0 = ty * LTY, k0 = tz * LTZ;  // 0-based interior coordinates
  const int64_t nx = d.n1, nxy = (int64_t)d.n1 * d.n2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int i = i0 + lane;
  const uint64_t below = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
  for (int r = wv; r < LTY * LTZ; r += nw) {
    const int j = j0 + (r & (LTY - 1)), k = k0 + r / LTY;
    const bool in = i < d.n1 && j < d.n2 && k < d.n3 && m[i + nx * j + nxy * k];
    const uint64_t bits = __ballot(in);
    const uint64_t starts = bits & ~(bits << 1);
    const uint64_t sb = starts & below;  // non-zero for an in-phase lane
    lab[r * LTX + lane] = in && sb ? r * LTX + 63 - __builtin_clzll(sb) : -1;
  }
  __syncthreads();
  for (int r = wv; r < LTY * LTZ; r += nw) {
    const int v = r * LTX + lane, y = r & (LTY - 1), z = r / LTY;
    if (lab[v] < 0) continue;
    const bool prev = lane > 0 && lab[v - 1] >= 0;
    if (y > 0 && lab[v - LTX] >= 0 && !(prev && lab[v - 1 - LTX] >= 0))
      uf_union<__HIP_MEMORY_SCOPE_WORKGROUP>(lab, v, v - LTX);
    if (z > 0 && lab[v - LTX * LTY] >= 0 && !(prev && lab[v - 1 - LTX * LTY] >= 0))
      uf_union<__HIP_MEMORY_SCOPE_WORKGROUP>(lab, v, v - LTX * LTY);
  }
  __syncthreads();
  for (int r = wv; r < LTY * LTZ; r += nw) {
    const int v = r * LTX + lane;
    const int j = j0 + (r & (LTY - 1)), k = k0 + r / LTY;
    if (i >= d.n1 || j >= d.n2 || k >= d.n3) continue;
    int out = -1;
    if (lab[v] >= 0) {
      const int rt = uf_find_s<__HIP_MEMORY_SCOPE_WORKGROUP>(lab, v);
      out = (int)((i0 + (rt & (LTX - 1))) + nx * (j0 + ((rt / LTX) & (LTY - 1))) + nxy * (k0 + rt / (LTX * LTY)));
    }
    L[i + nx * j + nxy * k] = out;
  }
}

// unions across the lower d1, d2 and d3 faces of every tile
__global__ __launch_bounds__(256) void k_lab_link(Dims d, const uint8_t* __restrict__ m, int* L) {
  const LabTiles t = lab_tiles(d);
  const int tx = blockIdx.x % t.gx, ty = (blockIdx.x / t.gx) % t.gy, tz = blockIdx.x / (t.gx * t.gy);
  const int i0 = tx * LTX, j0 = ty * LTY, k0 = tz * LTZ;
  const int64_t nx = d.n1, nxy = (int64_t)d.n1 * d.n2;
  constexpr int FX = LTY * LTZ, FY = LTX * LTZ, FZ = LTX * LTY;
  for (int f = threadIdx.x; f < FX + FY + FZ; f += blockDim.x) {
    int i, j, k;
    int64_t step;
    if (f < FX) {
      if (i0 == 0) continue;
      i = i0, j = j0 + f % LTY, k = k0 + f / LTY, step = 1;
    } else if (f < FX + FY) {
      if (j0 == 0) continue;
      const int g = f - FX;
      i = i0 + g % LTX, j = j0, k = k0 + g / LTX, step = nx;
    } else {
      if (k0 == 0) continue;
      const int g = f - FX - FY;
      i = i0 + g % LTX, j = j0 + g / LTX, k = k0, step = nxy;
    }
    if (i >= d.n1 || j >= d.n2 || k >= d.n3) continue;
    const int64_t q = i + nx * j + nxy * k;
    if (!(m[q] && m[q - step])) continue;
    // the same two tile components are joined by the face predecessor's
    // union when it and its neighbour are in phase too (d1 predecessor on
    // the d2 and d3 faces, d2 predecessor on the d1 face)
    const int64_t pre = step == 1 ? nx : 1;
    const bool has_pre = step == 1 ? j > j0 : i > i0;
    if (has_pre && m[q - pre] && m[q - pre - step]) continue;
    uf_union<__HIP_MEMORY_SCOPE_AGENT>(L, (int)q, (int)(q - step));
  }
}

__global__ void k_lab_compress(int64_t N, int* L) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x)
    if (L[q] >= 0) L[q] = uf_find_s<__HIP_MEMORY_SCOPE_AGENT>(L, (int)q);
}

// flag roots of components that touch the Dirichlet d2 plane jf
__global__ void k_lab_face(Dims d, const int* __restrict__ L, int jf, int* __restrict__ flag) {
  const int64_t M = (int64_t)d.n1 * d.n3;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < M; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % d.n1) + 1, k = (int)(q / d.n1) + 1;
    const int l = L[d.ni(i, jf, k)];
    if (l >= 0) flag[l] = 1;
  }
}

__global__ void k_lab_conn(int64_t N, const int* __restrict__ L, const int* __restrict__ flag, int* __restrict__ conn) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x) {
    const int l = L[q];
    conn[q] = l >= 0 ? flag[l] : 0;
  }
}

__global__ void k_rm(int64_t N, const double* __restrict__ cn, const double* __restrict__ pin, double* __restrict__ pout) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x)
    pout[q] = cn[q] < kThrSnap ? 0.0 : pin[q];
}

int accdiff(const Dims& d, const double* rdin, int phase, int term, double* cn) {
  ELSIM_TRY(ensure_init());
  hipStream_t s = stream();
  double* psi = (double*)scratch(5, (size_t)d.G() * 8);
  double* tmp = (double*)scratch(6, (size_t)d.N() * 8);
  if (!psi || !tmp) return ELSIM_ERR_ALLOC;
  k_ad_psi<<<flat_blocks(d.G()), FLAT_T, 0, s>>>(d, rdin, psi);
  ELSIM_HIP(hipMemsetAsync(cn, 0, (size_t)d.N() * 8, s));
  const TileGrid tg = stencil_tiles(d, AD_SJ);
  const int nt = tg.gx * tg.gy * tg.gz;
  int* chg = (int*)scratch(7, (size_t)2 * nt * sizeof(int));
  if (!chg) return ELSIM_ERR_ALLOC;
  double* a = cn;
  double* b = tmp;
  for (int it = 2; it <= term; it++) {
    const double bval = it == 2 ? 0.0 : 1.0;
    const int snap = (it % 5) == 1;
    // the previous iteration did not snap and had this Dirichlet value
    // (iterations 3 and on use 1)
    const int may_skip = it >= 4 && (it % 5) != 2;
    const int* chg_r = chg + ((it + 1) & 1) * nt;  // written by iteration it - 1
    int* chg_w = chg + (it & 1) * nt;
    if (phase == 0)
      k_ad_iter<0><<<nt, 256, 0, s>>>(d, psi, a, b, bval, snap, chg_r, chg_w, may_skip);
    else
      k_ad_iter<1><<<nt, 256, 0, s>>>(d, psi, a, b, bval, snap, chg_r, chg_w, may_skip);
    double* t = a;
    a = b;
    b = t;
  }
  ELSIM_CHECK_LAUNCH();
  if (a != cn) ELSIM_HIP(hipMemcpyAsync(cn, a, (size_t)d.N() * 8, hipMemcpyDeviceToDevice, s));
  return ELSIM_OK;
}

int label(const Dims& d, const double* rdin, int phase, int32_t* L, int32_t* conn) {
  ELSIM_TRY(ensure_init());
  hipStream_t s = stream();
  const int64_t N = d.N();
  // the component flags (N ints), then the phase mask (N bytes)
  int* flag = (int*)scratch(7, (size_t)N * 5);
  if (!flag) return ELSIM_ERR_ALLOC;
  uint8_t* mask = (uint8_t*)(flag + N);
  ELSIM_HIP(hipMemsetAsync(flag, 0, (size_t)N * 4, s));
  const int nb = flat_blocks(N);
  if (phase == 0)
    k_lab_mask<0><<<nb, FLAT_T, 0, s>>>(N, rdin, mask);
  else
    k_lab_mask<1><<<nb, FLAT_T, 0, s>>>(N, rdin, mask);
  const LabTiles lt = lab_tiles(d);
  const int ntl = lt.gx * lt.gy * lt.gz;
  k_lab_local<<<ntl, 256, 0, s>>>(d, mask, L);
  k_lab_link<<<ntl, 256, 0, s>>>(d, mask, L);
  k_lab_compress<<<nb, FLAT_T, 0, s>>>(N, L);
  k_lab_face<<<flat_blocks((int64_t)d.n1 * d.n3), FLAT_T, 0, s>>>(d, L, phase == 0 ? d.n2 : 1, flag);
  k_lab_conn<<<nb, FLAT_T, 0, s>>>(N, L, flag, conn);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

// RM_UCNCT_3D's sign block for one phase (RM:64-96 solid, RM:115-136
// electrolyte): psA = 0.5(1+tanh(RdIn/1.0)) (1 - psA for the electrolyte),
// zeroed where the connectivity Cn < 1e-4f, NoGrdBc3D, then sgF = +1 where
// the value > 0.5, else -1, over the ghosted box: a ghost takes the sign of
// its clamped interior cell.
template <int PHASE>
__global__ __launch_bounds__(256) void k_rm_sign(Dims d, const double* __restrict__ rdin,
                                                 const double* __restrict__ cn, double* __restrict__ sgf) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.G()) return;
  const int i = (int)(q % d.g1()), j = (int)((q / d.g1()) % (d.n2 + 2)), k = (int)(q / d.g12());
  const int64_t n = d.ni(min(max(i, 1), d.n1), min(max(j, 1), d.n2), min(max(k, 1), d.n3));
  double ps = 0.5 * (1.0 + tanh(rdin[n] / 1.0));
  if (PHASE == 1) ps = 1.0 - ps;
  if (cn[n] < kThrSnap) ps = 0.0;
  sgf[q] = ps > 0.5 ? 1.0 : -1.0;
}

}  // namespace
}  // namespace elsim

using namespace elsim;

extern "C" {

int elsim_dev_accdiff_phase(int py, int px, int pz, const double* rdin, int phase, int term, double* cn) {
  if (!dims_ok(py, px, pz) || !rdin || !cn || (phase != 0 && phase != 1) || term < 1) return ELSIM_ERR_ARG;
  return accdiff(Dims{py, px, pz}, rdin, phase, term, cn);
}

int elsim_accdiff_phase(const int* py, const int* px, const int* pz, const double* rdin, const int* phase,
                        const int* term, double* cn) {
  if (!py || !px || !pz || !rdin || !phase || !term || !cn) return ELSIM_ERR_ARG;
  if (!dims_ok(*py, *px, *pz) || (*phase != 0 && *phase != 1) || *term < 1) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t N = (size_t)d.N() * 8;
  double* dr = (double*)scratch(3, N);
  double* dc = (double*)scratch(4, N);
  if (!dr || !dc) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dr, rdin, N, hipMemcpyHostToDevice, s));
  ELSIM_TRY(accdiff(d, dr, *phase, *term, dc));
  ELSIM_HIP(hipMemcpyAsync(cn, dc, N, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_dev_label_connected(int py, int px, int pz, const double* rdin, int phase, int32_t* labels,
                              int32_t* connected) {
  if (!dims_ok(py, px, pz) || !rdin || !labels || !connected || (phase != 0 && phase != 1)) return ELSIM_ERR_ARG;
  return label(Dims{py, px, pz}, rdin, phase, labels, connected);
}

int elsim_label_connected(const int* py, const int* px, const int* pz, const double* rdin, const int* phase,
                          int32_t* labels, int32_t* connected) {
  if (!py || !px || !pz || !rdin || !phase || !labels || !connected) return ELSIM_ERR_ARG;
  if (!dims_ok(*py, *px, *pz) || (*phase != 0 && *phase != 1)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t N = (size_t)d.N();
  double* dr = (double*)scratch(3, N * 8);
  int32_t* dl = (int32_t*)scratch(4, N * 4);
  int32_t* dc = (int32_t*)scratch(6, N * 4);
  if (!dr || !dl || !dc) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dr, rdin, N * 8, hipMemcpyHostToDevice, s));
  ELSIM_TRY(label(d, dr, *phase, dl, dc));
  ELSIM_HIP(hipMemcpyAsync(labels, dl, N * 4, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipMemcpyAsync(connected, dc, N * 4, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_rm_threshold(const int* py, const int* px, const int* pz, const double* cn, const double* psi_in,
                       double* psi_out) {
  if (!py || !px || !pz || !cn || !psi_in || !psi_out || !dims_ok(*py, *px, *pz)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const int64_t N = (int64_t)*py * *px * *pz;
  double* a = (double*)scratch(3, N * 8);
  double* b = (double*)scratch(4, N * 8);
  double* c = (double*)scratch(6, N * 8);
  if (!a || !b || !c) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(a, cn, N * 8, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(b, psi_in, N * 8, hipMemcpyHostToDevice, s));
  k_rm<<<flat_blocks(N), FLAT_T, 0, s>>>(N, a, b, c);
  ELSIM_CHECK_LAUNCH();
  ELSIM_HIP(hipMemcpyAsync(psi_out, c, N * 8, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_dev_rm_phase(int py, int px, int pz, const double* rdin, const double* cn, int phase, double* sgf) {
  if (!dims_ok(py, px, pz) || !rdin || !cn || !sgf || (phase != 0 && phase != 1)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{py, px, pz};
  const int nb = (int)((d.G() + 255) / 256);
  if (phase == 0)
    k_rm_sign<0><<<nb, 256, 0, stream()>>>(d, rdin, cn, sgf);
  else
    k_rm_sign<1><<<nb, 256, 0, stream()>>>(d, rdin, cn, sgf);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int elsim_rm_phase(const int* py, const int* px, const int* pz, const double* rdin, const double* cn,
                   const int* phase, double* sgf) {
  if (!py || !px || !pz || !rdin || !cn || !phase || !sgf || !dims_ok(*py, *px, *pz)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t N = (size_t)d.N() * 8, G = (size_t)d.G() * 8;
  double* a = (double*)scratch(3, N);
  double* b = (double*)scratch(4, N);
  double* c = (double*)scratch(6, G);
  if (!a || !b || !c) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(a, rdin, N, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(b, cn, N, hipMemcpyHostToDevice, s));
  ELSIM_TRY(elsim_dev_rm_phase(*py, *px, *pz, a, b, *phase, c));
  ELSIM_HIP(hipMemcpyAsync(sgf, c, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

}  // extern "C"
