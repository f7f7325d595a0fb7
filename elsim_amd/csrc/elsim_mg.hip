// elsim_mg.hip -- geometric multigrid preconditioner of the potential solves.
//
// The potentials' operators (SBM:670-676, 751-757) are variable-coefficient
// diffusion operators whose coefficient spans seven orders of magnitude
// (psi-weighted: ~1e-7 in the other phase, O(1) in the own one).  Jacobi-PCG
// needs thousands of iterations on them at 256^3, because the error that
// survives is smooth along the particle network.  ELSIM_MODE_MIXED_MG
// preconditions the same FP32 CG (FP64 reliable updates unchanged) with one
// V-cycle of a geometric multigrid instead:
//
//   * Coarsening: 2 x 2 x 2 blocks of cells (ceil(n/2) per axis), piecewise
//     constant prolongation P and the Galerkin operator P^T A P.  For a
//     diffusion operator this is again a 7-point operator: the conductance
//     between two coarse cells is the sum of the fine face conductances
//     between their children, and the boundary (Dirichlet) and reaction
//     (psE for CnE) terms add up.  Everything is kept as sums of positive
//     terms (conductances and masses, the diagonal formed from them), so no
//     cancellation occurs however large the contrast; a neck of solid that
//     conducts keeps its conductance on every level.
//   * Smoother: one damped-Jacobi sweep before and one after the coarse
//     correction (weight omega), coarse correction weighted by oc.  With the
//     pre-smoother started from zero the cycle is
//       M^-1 = 2S - SAS + oc (I - SA) P A_c^-1 P^T (I - AS),  S = omega D^-1,
//     symmetric, and positive definite for omega < 2 / lambda_max(D^-1 A)
//     (lambda_max <= 2 for these diagonally dominant operators).
//   * Coarsest level (at most 64 cells): a dense inverse, formed once per
//     hierarchy by Gauss-Jordan elimination in FP64 in one block.
//
// The fine level is the CG's own FP32 operator.  The down-sweep forms
// r - A x1 (x1 = omega D^-1 r, the pre-smoothed iterate) with the
// four-voxels-a-lane tile walk of the CG operator (mv4_tile) and sums it
// over 2 x 2 x 2 blocks in registers (a lane's four d1 voxels, its two rows
// and the tile's eight planes hold whole blocks), so the fine residual is
// never stored; the up-sweep forms x2 = x1 + oc P x_c in the loader and
// writes z = x2 + omega D^-1 (r - A x2).  On boxes whose n1 is a multiple of
// 4 the CG's residual update r -= alpha w runs inside the down-sweep and x1
// is formed in both sweeps' loaders, never stored.  Levels 1 .. lb-1 run one
// kernel per level and direction (down: a block per row of 32 coarse cells,
// the children along the lanes; up: a thread per cell); the levels of at
// most kMgBottomMax cells run in one single-block kernel.
//
// The operators are negative definite for the potentials (A = -A_hat) and
// positive for CnE; the coarse levels hold A_hat = sigma A, sigma = -1 / +1,
// and the fine level feeds them sigma times its residual, so the cycle's
// output approximates A^-1 r.
//
// There is no counterpart in the reference (its CG is unpreconditioned,
// SBM:657-714, 737-795); the solve still stops on the FP64 true residual of
// the reference's system, so the preconditioner changes the iteration count,
// not the result's tolerance.
#include <cmath>

#include "elsim_mg.h"
#include "elsim_sweep.h"

namespace elsim {

MgCfg& mg_cfg() {
  static MgCfg c{0.7f, 1.5f, 1.0e-3};
  return c;
}

namespace {

__device__ __forceinline__ bool mg_skip(const CgDev* cg) { return cg && (cg->done || cg->need_reliable); }

__host__ __device__ inline int64_t lvl_n(const MgLevel& L) { return (int64_t)L.m1 * L.m2 * L.m3; }
__device__ __forceinline__ int64_t lidx(const MgLevel& L, int i, int j, int k) {
  return i + (int64_t)L.m1 * (j + (int64_t)L.m2 * k);
}

struct Ijk {
  int i, j, k;
};
__device__ __forceinline__ Ijk unflat(const MgLevel& L, int64_t q) {
  const int64_t r = q / L.m1;
  return Ijk{(int)(q - r * L.m1), (int)(r % L.m2), (int)(r / L.m2)};
}

// A_hat x at cell c = (i, j, k) (flat q) of level L; xc = x there, xn(q', i',
// j', k') = x at a neighbour.
template <class X>
__device__ __forceinline__ float apply_lvl(const MgLevel& L, const Ijk& c, int64_t q, float xc, X&& xn) {
  const int64_t sy = L.m1, sz = (int64_t)L.m1 * L.m2;
  float s = L.dg[q] * xc;
  if (c.i + 1 < L.m1) s -= L.cx[q] * xn(q + 1, c.i + 1, c.j, c.k);
  if (c.i > 0) s -= L.cx[q - 1] * xn(q - 1, c.i - 1, c.j, c.k);
  if (c.j + 1 < L.m2) s -= L.cy[q] * xn(q + sy, c.i, c.j + 1, c.k);
  if (c.j > 0) s -= L.cy[q - sy] * xn(q - sy, c.i, c.j - 1, c.k);
  if (c.k + 1 < L.m3) s -= L.cz[q] * xn(q + sz, c.i, c.j, c.k + 1);
  if (c.k > 0) s -= L.cz[q - sz] * xn(q - sz, c.i, c.j, c.k - 1);
  return s;
}

// ----------------------------------------------------------- hierarchy --

__global__ __launch_bounds__(256) void k_mg_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) b[q] = (float)a[q];
}

__device__ __forceinline__ void store_level(MgLevel& c, int64_t q, double cxp, double cxm, double cyp, double cym,
                                            double czp, double czm, double ms, float omega) {
  const double dg = ms + ((cxp + cxm) + (cyp + cym) + (czp + czm));
  c.cx[q] = (float)cxp;
  c.cy[q] = (float)cyp;
  c.cz[q] = (float)czp;
  c.ms[q] = (float)ms;
  c.dg[q] = (float)dg;
  c.wd[q] = (float)(omega / dg);
}

// Level 1 from the fine operator: thread per level-1 cell.  Fine face
// conductance between ghosted cells a and b: scale * 0.5 (kap_a + kap_b), the
// face coefficient of the CG operator.  Mass: psE (CnE) and, for a face
// whose homogeneous ghost rule is zero (Dirichlet data), that face's
// conductance.
template <int OP>
__global__ __launch_bounds__(256) void k_mg_build1(Dims d, const double* __restrict__ kap,
                                                   const double* __restrict__ pse, double scale, MgLevel c,
                                                   float omega) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= lvl_n(c)) return;
  const Ijk C = unflat(c, q);
  // children (ghosted, 1-based): [2I+1, min(2I+2, n)]
  const int i0 = 2 * C.i + 1, i1 = min(2 * C.i + 2, d.n1);
  const int j0 = 2 * C.j + 1, j1 = min(2 * C.j + 2, d.n2);
  const int k0 = 2 * C.k + 1, k1 = min(2 * C.k + 2, d.n3);
  const int64_t sx = d.g1(), sz = d.g12();
  auto f = [&](int64_t a, int64_t b) { return scale * ((kap[a] + kap[b]) * 0.5); };
  double cxp = 0.0, cxm = 0.0, cyp = 0.0, cym = 0.0, czp = 0.0, czm = 0.0, ms = 0.0;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++) {
      if (C.i + 1 < c.m1) cxp += f(d.gi(i1, j, k), d.gi(i1, j, k) + 1);
      if (C.i > 0) cxm += f(d.gi(i0, j, k), d.gi(i0, j, k) - 1);
    }
  for (int k = k0; k <= k1; k++)
    for (int i = i0; i <= i1; i++) {
      if (C.j + 1 < c.m2) cyp += f(d.gi(i, j1, k), d.gi(i, j1, k) + sx);
      if (C.j > 0) cym += f(d.gi(i, j0, k), d.gi(i, j0, k) - sx);
      if (Rule<OP>::lo_zero && j0 == 1) ms += f(d.gi(i, 1, k), d.gi(i, 0, k));
      if (Rule<OP>::hi_zero && j1 == d.n2) ms += f(d.gi(i, d.n2, k), d.gi(i, d.n2 + 1, k));
    }
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      if (C.k + 1 < c.m3) czp += f(d.gi(i, j, k1), d.gi(i, j, k1) + sz);
      if (C.k > 0) czm += f(d.gi(i, j, k0), d.gi(i, j, k0) - sz);
    }
  if (Rule<OP>::diag)
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) ms += pse[d.gi(i, j, k)];
  store_level(c, q, cxp, cxm, cyp, cym, czp, czm, ms, omega);
}

// Level l+1 from level l (f): thread per coarse cell.
__global__ __launch_bounds__(256) void k_mg_build(MgLevel f, MgLevel c, float omega) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= lvl_n(c)) return;
  const Ijk C = unflat(c, q);
  const int i0 = 2 * C.i, i1 = min(2 * C.i + 1, f.m1 - 1);
  const int j0 = 2 * C.j, j1 = min(2 * C.j + 1, f.m2 - 1);
  const int k0 = 2 * C.k, k1 = min(2 * C.k + 1, f.m3 - 1);
  double cxp = 0.0, cxm = 0.0, cyp = 0.0, cym = 0.0, czp = 0.0, czm = 0.0, ms = 0.0;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++) {
      if (C.i + 1 < c.m1) cxp += f.cx[lidx(f, i1, j, k)];
      if (C.i > 0) cxm += f.cx[lidx(f, i0 - 1, j, k)];
    }
  for (int k = k0; k <= k1; k++)
    for (int i = i0; i <= i1; i++) {
      if (C.j + 1 < c.m2) cyp += f.cy[lidx(f, i, j1, k)];
      if (C.j > 0) cym += f.cy[lidx(f, i, j0 - 1, k)];
    }
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      if (C.k + 1 < c.m3) czp += f.cz[lidx(f, i, j, k1)];
      if (C.k > 0) czm += f.cz[lidx(f, i, j, k0 - 1)];
    }
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) ms += f.ms[lidx(f, i, j, k)];
  store_level(c, q, cxp, cxm, cyp, cym, czp, czm, ms, omega);
}

// Dense inverse of the coarsest operator (n <= kMgCoarseMax = 64 cells):
// in-place Gauss-Jordan elimination in FP64 in LDS, one block of 1024
// threads, rows padded to 64.  The operator is symmetric positive definite
// (every level keeps at least one Dirichlet or reaction term), so no pivoting
// is needed.
static_assert(kMgCoarseMax == 64, "k_mg_coarse_inv indexes rows of 64");
__global__ __launch_bounds__(1024) void k_mg_coarse_inv(MgLevel c, float* __restrict__ ainv) {
  __shared__ double a[64 * 64];
  __shared__ double fac[64];
  __shared__ double piv;
  const int n = (int)lvl_n(c), tid = threadIdx.x;
  for (int e = tid; e < 64 * 64; e += 1024) a[e] = 0.0;
  __syncthreads();
  if (tid < n) {
    const int q = tid;
    const Ijk C = unflat(c, q);
    const int sy = c.m1, sz = c.m1 * c.m2;
    a[q * 64 + q] = c.dg[q];
    if (C.i + 1 < c.m1) a[q * 64 + q + 1] = -(double)c.cx[q];
    if (C.i > 0) a[q * 64 + q - 1] = -(double)c.cx[q - 1];
    if (C.j + 1 < c.m2) a[q * 64 + q + sy] = -(double)c.cy[q];
    if (C.j > 0) a[q * 64 + q - sy] = -(double)c.cy[q - sy];
    if (C.k + 1 < c.m3) a[q * 64 + q + sz] = -(double)c.cz[q];
    if (C.k > 0) a[q * 64 + q - sz] = -(double)c.cz[q - sz];
  }
  __syncthreads();
  for (int p = 0; p < n; p++) {
    if (tid == 0) piv = a[p * 64 + p];
    if (tid < n) fac[tid] = a[tid * 64 + p];
    __syncthreads();
    if (tid == 0) a[p * 64 + p] = 1.0;
    __syncthreads();
    if (tid < n) a[p * 64 + tid] /= piv;
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 1024) {
      const int r = e >> 6, j = e & 63;
      if (r < n && j < n && r != p) a[e] = (j == p ? 0.0 : a[e]) - fac[r] * a[p * 64 + j];
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += 1024) ainv[e] = (float)a[(e / n) * 64 + e % n];
}

// ------------------------------------------------- coarse-level cycle --

// Down-pass from level f to c for the child-thread index g (8 per coarse
// cell, children in adjacent lanes): c.b = restriction of f.b - A_hat x1,
// x1 = wd f.b.  Every lane of the wave must call it (butterfly shuffles).
__device__ __forceinline__ void down_child(const MgLevel& f, const MgLevel& c, int64_t g) {
  const int64_t Q = g >> 3;
  const int ch = (int)(g & 7);
  float v = 0.0f;
  if (Q < lvl_n(c)) {
    const Ijk C = unflat(c, Q);
    const Ijk F{2 * C.i + (ch & 1), 2 * C.j + ((ch >> 1) & 1), 2 * C.k + (ch >> 2)};
    if (F.i < f.m1 && F.j < f.m2 && F.k < f.m3) {
      const int64_t p = lidx(f, F.i, F.j, F.k);
      auto x1 = [&](int64_t q, int, int, int) { return f.wd[q] * f.b[q]; };
      v = f.b[p] - apply_lvl(f, F, p, x1(p, 0, 0, 0), x1);
    }
  }
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  if (ch == 0 && Q < lvl_n(c)) c.b[Q] = v;
}

// Up-pass at level f (cell p): x2 = wd b + oc P x_c, f.x = x2 + wd (b - A_hat x2).
__device__ __forceinline__ void up_cell(const MgLevel& f, const MgLevel& c, int64_t p, float oc) {
  const Ijk F = unflat(f, p);
  auto x2 = [&](int64_t q, int i, int j, int k) { return f.wd[q] * f.b[q] + oc * c.x[lidx(c, i >> 1, j >> 1, k >> 1)]; };
  const float xc = x2(p, F.i, F.j, F.k);
  const float ax = apply_lvl(f, F, p, xc, x2);
  f.x[p] = xc + f.wd[p] * (f.b[p] - ax);
}

__global__ __launch_bounds__(256) void k_mg_down(MgLevel f, MgLevel c, const CgDev* cg) {
  if (mg_skip(cg)) return;
  down_child(f, c, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

// The same pass with the children laid out along the lanes: a block is a
// row of 32 coarse cells (64 fine cells in d1, one lane each) and its four
// waves the four (d2, d3) child rows, so that every load of the level's
// arrays is one contiguous run per wave.  Children are summed in pairs along
// d1 (shuffle), then over the four rows (LDS).
__global__ __launch_bounds__(256) void k_mg_down_rows(MgLevel f, MgLevel c, const CgDev* cg) {
  if (mg_skip(cg)) return;
  __shared__ float part[4][32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nbx = (c.m1 + 31) / 32;
  const int bx = blockIdx.x % nbx, cj = (blockIdx.x / nbx) % c.m2, ck = blockIdx.x / (nbx * c.m2);
  const Ijk F{64 * bx + lane, 2 * cj + (w & 1), 2 * ck + (w >> 1)};
  float v = 0.0f;
  if (F.i < f.m1 && F.j < f.m2 && F.k < f.m3) {
    const int64_t p = lidx(f, F.i, F.j, F.k);
    auto x1 = [&](int64_t q, int, int, int) { return f.wd[q] * f.b[q]; };
    v = f.b[p] - apply_lvl(f, F, p, x1(p, 0, 0, 0), x1);
  }
  v += __shfl_xor(v, 1, 64);
  if ((lane & 1) == 0) part[w][lane >> 1] = v;
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 32 && 32 * bx + t < c.m1) c.b[lidx(c, 32 * bx + t, cj, ck)] = (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
}

__global__ __launch_bounds__(256) void k_mg_up(MgLevel f, MgLevel c, float oc, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < lvl_n(f)) up_cell(f, c, p, oc);
}

// Levels lb .. nlev in one block: down-passes, the dense coarse solve,
// up-passes, with a block barrier between passes.  The levels' arrays and
// the coarsest inverse are staged in LDS first (at most kMgBottomCells cells
// together), so the passes' dependent loads are LDS accesses, not L2 round
// trips (13.6 us a V-cycle with the arrays in global memory); the block
// writes back only level lb's iterate, which the up-pass above reads.
// lv: the level table in device memory.
__global__ __launch_bounds__(1024) void k_mg_bottom(const MgLevel* __restrict__ lv, int lb, int nlev,
                                                    const float* __restrict__ ainv, float oc, const CgDev* cg) {
  if (mg_skip(cg)) return;
  __shared__ float sm[7 * kMgBottomCells];
  __shared__ float sinv[kMgCoarseMax * kMgCoarseMax];
  __shared__ MgLevel sl[kMgMaxLev + 1];
  const int tid = threadIdx.x;
  if (tid == 0) {
    int off = 0;
    for (int l = lb; l <= nlev; l++) {
      MgLevel v = lv[l];
      const int n = (int)lvl_n(v);
      v.cx = sm + off;
      v.cy = v.cx + n;
      v.cz = v.cy + n;
      v.dg = v.cz + n;
      v.wd = v.dg + n;
      v.b = v.wd + n;
      v.x = v.b + n;
      v.ms = nullptr;
      off += 7 * n;
      sl[l] = v;
    }
  }
  __syncthreads();
  for (int l = lb; l <= nlev; l++) {
    const MgLevel &g = lv[l], &v = sl[l];
    const int n = (int)lvl_n(v);
    for (int e = tid; e < n; e += 1024) {
      v.cx[e] = g.cx[e];
      v.cy[e] = g.cy[e];
      v.cz[e] = g.cz[e];
      v.dg[e] = g.dg[e];
      v.wd[e] = g.wd[e];
      if (l == lb) v.b[e] = g.b[e];
    }
  }
  const int nc = (int)lvl_n(sl[nlev]);
  for (int e = tid; e < nc * nc; e += 1024) sinv[e] = ainv[e];
  __syncthreads();
  for (int l = lb; l < nlev; l++) {
    const int64_t ng = 8 * lvl_n(sl[l + 1]);
    for (int64_t g0 = 0; g0 < ng; g0 += 1024) down_child(sl[l], sl[l + 1], g0 + tid);
    __syncthreads();
  }
  {
    const MgLevel& c = sl[nlev];
    if (tid < nc) {
      float s = 0.0f;
      for (int j = 0; j < nc; j++) s += sinv[tid * nc + j] * c.b[j];
      c.x[tid] = s;
    }
    __syncthreads();
  }
  for (int l = nlev - 1; l >= lb; l--) {
    const int64_t nf = lvl_n(sl[l]);
    for (int64_t p = tid; p < nf; p += 1024) up_cell(sl[l], sl[l + 1], p, oc);
    __syncthreads();
  }
  const int nb = (int)lvl_n(sl[lb]);
  for (int e = tid; e < nb; e += 1024) lv[lb].x[e] = sl[lb].x[e];
}

// --------------------------------------------------------- fine level --

constexpr int MG_P = 4;  // element pairs per thread of the flat kernels

// r -= alpha w; x1 = omega dinv r.
__global__ __launch_bounds__(256) void k_mg_update(int64_t G2, float2* __restrict__ r32,
                                                   const float2* __restrict__ w32, const float2* __restrict__ dinv,
                                                   float2* __restrict__ x1, float omega, const CgDev* cg) {
  if (cg->done || cg->need_reliable) return;
  const float alp = (float)cg->alpha;
  float2 rv[MG_P], wv[MG_P], dv[MG_P];
  FLAT_LOOP_U(MG_P, G2, { rv[u] = r32[q]; wv[u] = w32[q]; dv[u] = dinv[q]; }, {
    float2 rn;
    float2 xn;
    rn.x = rv[u].x - alp * wv[u].x;
    rn.y = rv[u].y - alp * wv[u].y;
    xn.x = omega * dv[u].x * rn.x;
    xn.y = omega * dv[u].y * rn.y;
    r32[q] = rn;
    x1[q] = xn;
  })
}

// x1 = omega dinv r.
__global__ __launch_bounds__(256) void k_mg_presmooth(int64_t G2, const float2* __restrict__ r32,
                                                      const float2* __restrict__ dinv, float2* __restrict__ x1,
                                                      float omega, const CgDev* cg) {
  if (mg_skip(cg)) return;
  float2 rv[MG_P], dv[MG_P];
  FLAT_LOOP_U(MG_P, G2, { rv[u] = r32[q]; dv[u] = dinv[q]; }, {
    float2 xn;
    xn.x = omega * dv[u].x * rv[u].x;
    xn.y = omega * dv[u].y * rv[u].y;
    x1[q] = xn;
  })
}

// Down-sweep, four voxels a lane: c.b = restriction of sigma (r - A x1) to
// level 1, summed in registers; partials of r.r (slot 0).  A lane's four
// voxels are two level-1 cells in d1, its rows (SJ4, even, from an odd row)
// whole level-1 rows in d2, and the tile's eight planes (aligned to 8) four
// in d3, so every level-1 cell is summed by one lane, over its planes in the
// tile's marching order.
template <int OP, bool EDGE>
__global__ __launch_bounds__(256) void k_mg_down4(Dims d, const float* __restrict__ x1, const float* __restrict__ r,
                                                  const float* __restrict__ kap, const float* __restrict__ pse,
                                                  float scale, MgLevel c, const CgDev* cg, double* partials) {
  if (mg_skip(cg)) return;
  constexpr float sg = Rule<OP>::diag ? 1.0f : -1.0f;
  double rr = 0.0;
  const int nt = mv4_ntiles(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    bool used[2] = {false, false};
    int jc[2] = {0, 0}, ic = 0, kc = -1;
    auto flush = [&]() {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (used[h]) {
          const int64_t q = lidx(c, ic, jc[h], kc);
          c.b[q] = acc[h][0];
          c.b[q + 1] = acc[h][1];
        }
        acc[h][0] = acc[h][1] = 0.f;
        used[h] = false;
      }
    };
    mv4_tile<OP, EDGE>(
        d, tile, [&](uint32_t q, int, int, int) { return ld4(x1, q); },
        [&](uint32_t q, int, int, int) { return x1[q]; }, kap, pse, scale,
        [&](int t, int j, int k, uint32_t q, const float (&)[4], const float (&o)[4]) {
          float rv[4];
          unpack4(ld4(r, q), rv);
          float tq[4];
#pragma unroll
          for (int e = 0; e < 4; e++) {
            tq[e] = sg * (rv[e] - o[e]);
            rr += (double)rv[e] * (double)rv[e];
          }
          const int K = (k - 1) >> 1;
          if (K != kc) {
            flush();
            kc = K;
          }
          const int h = t >> 1;
          acc[h][0] += tq[0] + tq[1];
          acc[h][1] += tq[2] + tq[3];
          used[h] = true;
          jc[h] = (j - 1) >> 1;
          ic = ((int)(q % (uint32_t)d.g1()) - 1) >> 1;
        });
    flush();
  }
  double v[1] = {rr};
  block_partials<1>(v, partials);
}

// Up-sweep, four voxels a lane: x2 = x1 + oc P x_c formed in the loader,
// z = x2 + omega dinv (r - A x2); partials of r.z (slot 1).
template <int OP, bool EDGE>
__global__ __launch_bounds__(256) void k_mg_up4(Dims d, const float* __restrict__ x1, const float* __restrict__ r,
                                                const float* __restrict__ dinv, const float* __restrict__ kap,
                                                const float* __restrict__ pse, float scale, float omega, float oc,
                                                MgLevel c, float* __restrict__ z, const CgDev* cg,
                                                double* partials) {
  if (mg_skip(cg)) return;
  double rz = 0.0;
  const int nt = mv4_ntiles(d);
  // level-1 row / plane of a fine row / plane, ghosts clamped to the edge
  auto cj = [&](int j) { return min(max((j - 1) >> 1, 0), c.m2 - 1); };
  auto ck = [&](int k) { return min(max((k - 1) >> 1, 0), c.m3 - 1); };
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    mv4_tile<OP, EDGE>(
        d, tile,
        [&](uint32_t q, int i, int j, int k) {
          float xv[4];
          unpack4(ld4(x1, q), xv);
          float2 xc;
          __builtin_memcpy(&xc, c.x + lidx(c, (i - 1) >> 1, cj(j), ck(k)), 8);
          return make_float4(xv[0] + oc * xc.x, xv[1] + oc * xc.x, xv[2] + oc * xc.y, xv[3] + oc * xc.y);
        },
        [&](uint32_t q, int i, int j, int k) {
          const int ci = min(max((i - 1) >> 1, 0), c.m1 - 1);
          return x1[q] + oc * c.x[lidx(c, ci, cj(j), ck(k))];
        },
        kap, pse, scale,
        [&](int, int, int, uint32_t q, const float (&xv)[4], const float (&o)[4]) {
          float rv[4], dv[4], zv[4];
          unpack4(ld4(r, q), rv);
          unpack4(ld4(dinv, q), dv);
#pragma unroll
          for (int e = 0; e < 4; e++) {
            zv[e] = xv[e] + omega * dv[e] * (rv[e] - o[e]);
            rz += (double)rv[e] * (double)zv[e];
          }
          st4(z, q, make_float4(zv[0], zv[1], zv[2], zv[3]));
        });
  double v[1] = {rz};
  block_sum<1>(v);
  if (linear_tid() == 0) partials[kMaxRedBlocks + linear_bid()] = v[0];
}

// The same sweeps with the CG's residual update folded in, so that the
// pre-smoothed iterate is never stored: the down-sweep forms
// r = r_in - alpha w and x1 = omega dinv r in its loaders (alpha from cg,
// or 0 at a start or restart: r = r_in) and writes r to r_out (a second
// buffer: its neighbours still read r_in), the up-sweep forms x1 from r and
// dinv again.  Per voxel and iteration this moves 16 bytes less than the
// separate update (r, w, dinv in, r, x1 out) and the sweeps reading x1.
template <int OP, bool EDGE>
__global__ __launch_bounds__(256) void k_mg_down4f(Dims d, const float* __restrict__ r_in,
                                                   const float* __restrict__ w, const float* __restrict__ dinv,
                                                   int use_alpha, float omega, float* __restrict__ r_out,
                                                   const float* __restrict__ kap, const float* __restrict__ pse,
                                                   float scale, MgLevel c, const CgDev* cg, double* partials) {
  if (mg_skip(cg)) return;
  constexpr float sg = Rule<OP>::diag ? 1.0f : -1.0f;
  const float alp = use_alpha ? (float)cg->alpha : 0.0f;
  double rr = 0.0;
  const int nt = mv4_ntiles(d);
  auto rnew4 = [&](uint32_t q, float (&rv)[4]) {
    float a[4], b[4];
    unpack4(ld4(r_in, q), a);
    unpack4(ld4(w, q), b);
#pragma unroll
    for (int e = 0; e < 4; e++) rv[e] = a[e] - alp * b[e];
  };
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    bool used[2] = {false, false};
    int jc[2] = {0, 0}, ic = 0, kc = -1;
    auto flush = [&]() {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (used[h]) {
          const int64_t q = lidx(c, ic, jc[h], kc);
          c.b[q] = acc[h][0];
          c.b[q + 1] = acc[h][1];
        }
        acc[h][0] = acc[h][1] = 0.f;
        used[h] = false;
      }
    };
    mv4_tile<OP, EDGE>(
        d, tile,
        [&](uint32_t q, int, int, int) {
          float rv[4], dv[4];
          rnew4(q, rv);
          unpack4(ld4(dinv, q), dv);
          return make_float4(omega * dv[0] * rv[0], omega * dv[1] * rv[1], omega * dv[2] * rv[2],
                             omega * dv[3] * rv[3]);
        },
        [&](uint32_t q, int, int, int) { return omega * dinv[q] * (r_in[q] - alp * w[q]); }, kap, pse, scale,
        [&](int t, int j, int k, uint32_t q, const float (&)[4], const float (&o)[4]) {
          float rv[4];
          rnew4(q, rv);
          if (r_out) st4(r_out, q, make_float4(rv[0], rv[1], rv[2], rv[3]));
          float tq[4];
#pragma unroll
          for (int e = 0; e < 4; e++) {
            tq[e] = sg * (rv[e] - o[e]);
            rr += (double)rv[e] * (double)rv[e];
          }
          const int K = (k - 1) >> 1;
          if (K != kc) {
            flush();
            kc = K;
          }
          const int h = t >> 1;
          acc[h][0] += tq[0] + tq[1];
          acc[h][1] += tq[2] + tq[3];
          used[h] = true;
          jc[h] = (j - 1) >> 1;
          ic = ((int)(q % (uint32_t)d.g1()) - 1) >> 1;
        });
    flush();
  }
  double v[1] = {rr};
  block_partials<1>(v, partials);
}

template <int OP, bool EDGE>
__global__ __launch_bounds__(256) void k_mg_up4f(Dims d, const float* __restrict__ r, const float* __restrict__ dinv,
                                                 const float* __restrict__ kap, const float* __restrict__ pse,
                                                 float scale, float omega, float oc, MgLevel c,
                                                 float* __restrict__ z, const CgDev* cg, double* partials) {
  if (mg_skip(cg)) return;
  double rz = 0.0;
  const int nt = mv4_ntiles(d);
  auto cj = [&](int j) { return min(max((j - 1) >> 1, 0), c.m2 - 1); };
  auto ck = [&](int k) { return min(max((k - 1) >> 1, 0), c.m3 - 1); };
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    mv4_tile<OP, EDGE>(
        d, tile,
        [&](uint32_t q, int i, int j, int k) {
          float rv[4], dv[4];
          unpack4(ld4(r, q), rv);
          unpack4(ld4(dinv, q), dv);
          float2 xc;
          __builtin_memcpy(&xc, c.x + lidx(c, (i - 1) >> 1, cj(j), ck(k)), 8);
          return make_float4(omega * dv[0] * rv[0] + oc * xc.x, omega * dv[1] * rv[1] + oc * xc.x,
                             omega * dv[2] * rv[2] + oc * xc.y, omega * dv[3] * rv[3] + oc * xc.y);
        },
        [&](uint32_t q, int i, int j, int k) {
          const int ci = min(max((i - 1) >> 1, 0), c.m1 - 1);
          return omega * dinv[q] * r[q] + oc * c.x[lidx(c, ci, cj(j), ck(k))];
        },
        kap, pse, scale,
        [&](int, int, int, uint32_t q, const float (&xv)[4], const float (&o)[4]) {
          float rv[4], dv[4], zv[4];
          unpack4(ld4(r, q), rv);
          unpack4(ld4(dinv, q), dv);
#pragma unroll
          for (int e = 0; e < 4; e++) {
            zv[e] = xv[e] + omega * dv[e] * (rv[e] - o[e]);
            rz += (double)rv[e] * (double)zv[e];
          }
          st4(z, q, make_float4(zv[0], zv[1], zv[2], zv[3]));
        });
  double v[1] = {rz};
  block_sum<1>(v);
  if (linear_tid() == 0) partials[kMaxRedBlocks + linear_bid()] = v[0];
}

// One voxel a lane, for boxes whose n1 is not a multiple of 4.
template <int OP>
__global__ __launch_bounds__(256) void k_mg_down1(Dims d, const float* __restrict__ x1, const float* __restrict__ r,
                                                  const float* __restrict__ kap, const float* __restrict__ pse,
                                                  float scale, float* __restrict__ t, const CgDev* cg,
                                                  double* partials) {
  if (mg_skip(cg)) return;
  constexpr float sg = Rule<OP>::diag ? 1.0f : -1.0f;
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<float>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<float, OP, true>(d, stencil_idx<Rows<float>::SJ>(d, tile), x1, kap, pse, scale,
                           [&](int64_t q, int64_t, float, float v) {
                             const float rc = r[q];
                             t[q] = sg * (rc - v);
                             acc[0] += (double)rc * (double)rc;
                           });
  block_partials<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_mg_restrict1(Dims d, const float* __restrict__ t, MgLevel c,
                                                      const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= lvl_n(c)) return;
  const Ijk C = unflat(c, q);
  const int i1 = min(2 * C.i + 2, d.n1), j1 = min(2 * C.j + 2, d.n2), k1 = min(2 * C.k + 2, d.n3);
  float s = 0.0f;
  for (int k = 2 * C.k + 1; k <= k1; k++)
    for (int j = 2 * C.j + 1; j <= j1; j++)
      for (int i = 2 * C.i + 1; i <= i1; i++) s += t[d.gi(i, j, k)];
  c.b[q] = s;
}

// t = x1 + oc P x_c over the interior (ghosts of t stay zero).
__global__ __launch_bounds__(256) void k_mg_prolong1(Dims d, const float* __restrict__ x1, MgLevel c, float oc,
                                                     float* __restrict__ t, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const ColIdx ci = col_idx(d);
  if (!ci.active) return;
  const int64_t base = lidx(c, (ci.i - 1) >> 1, (ci.j - 1) >> 1, 0), sz = (int64_t)c.m1 * c.m2;
  for_col(d, ci, [&](int64_t q, int64_t, int k) { t[q] = x1[q] + oc * c.x[base + sz * ((k - 1) >> 1)]; });
}

template <int OP>
__global__ __launch_bounds__(256) void k_mg_up1(Dims d, const float* __restrict__ t, const float* __restrict__ r,
                                                const float* __restrict__ dinv, const float* __restrict__ kap,
                                                const float* __restrict__ pse, float scale, float omega,
                                                float* __restrict__ z, const CgDev* cg, double* partials) {
  if (mg_skip(cg)) return;
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<float>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<float, OP, true>(d, stencil_idx<Rows<float>::SJ>(d, tile), t, kap, pse, scale,
                           [&](int64_t q, int64_t, float x, float v) {
                             const float rc = r[q];
                             const float zc = x + omega * dinv[q] * (rc - v);
                             z[q] = zc;
                             acc[0] += (double)rc * (double)zc;
                           });
  block_sum<1>(acc);
  if (linear_tid() == 0) partials[kMaxRedBlocks + linear_bid()] = acc[0];
}

inline int blocks256(int64_t n) { return (int)((n + 255) / 256); }

// down-passes from level 1 to level lb (one kernel each)
inline void down_levels(MgWork& m, const CgDev* cg, hipStream_t s) {
  MgLevel* lv = m.L.lv;
  for (int l = 1; l < m.lb; l++) {
    const MgLevel& c = lv[l + 1];
    k_mg_down_rows<<<((c.m1 + 31) / 32) * c.m2 * c.m3, 256, 0, s>>>(lv[l], c, cg);
  }
}

template <int OP>
int setup_t(MgWork& m, const Dims& d, const double* kap, const double* pse, double scale) {
  hipStream_t s = stream();
  const float omega = mg_cfg().omega;
  const int64_t G = d.G();
  k_mg_to_f32<<<blocks256(G), 256, 0, s>>>(G, kap, m.kap32);
  if (Rule<OP>::diag) k_mg_to_f32<<<blocks256(G), 256, 0, s>>>(G, pse, m.pse32);
  k_dinv<OP, float><<<col_grid(d), dim3(TX, TY, 1), 0, s>>>(d, kap, pse, scale, m.dinv);
  k_mg_build1<OP><<<blocks256(lvl_n(m.L.lv[1])), 256, 0, s>>>(d, kap, pse, scale, m.L.lv[1], omega);
  for (int l = 1; l < m.nlev; l++)
    k_mg_build<<<blocks256(lvl_n(m.L.lv[l + 1])), 256, 0, s>>>(m.L.lv[l], m.L.lv[l + 1], omega);
  k_mg_coarse_inv<<<1, 1024, 0, s>>>(m.L.lv[m.nlev], m.ainv);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

template <int OP>
int vcycle_t(MgWork& m, const Dims& d, const float* r32, float* z, const CgDev* cg, double* partials, bool tm) {
  hipStream_t s = stream();
  const MgCfg& g = mg_cfg();
  const float sc = (float)m.scale;
  MgLevel* lv = m.L.lv;
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_DOWN, 0));
  if (mv4_ok(d)) {
    const int nt = mv4_ntiles(d);
    if (mv4_tiles(d).gx > 1)
      k_mg_down4<OP, true><<<nt, 256, 0, s>>>(d, m.x1, r32, m.kap32, m.pse32, sc, lv[1], cg, partials);
    else
      k_mg_down4<OP, false><<<nt, 256, 0, s>>>(d, m.x1, r32, m.kap32, m.pse32, sc, lv[1], cg, partials);
  } else {
    k_mg_down1<OP><<<stencil_ntiles<float>(d), 256, 0, s>>>(d, m.x1, r32, m.kap32, m.pse32, sc, m.t, cg, partials);
    k_mg_restrict1<<<blocks256(lvl_n(lv[1])), 256, 0, s>>>(d, m.t, lv[1], cg);
  }
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_DOWN, 1));
  down_levels(m, cg, s);
  k_mg_bottom<<<1, 1024, 0, s>>>(m.dlv, m.lb, m.nlev, m.ainv, g.oc, cg);
  for (int l = m.lb - 1; l >= 1; l--) k_mg_up<<<blocks256(lvl_n(lv[l])), 256, 0, s>>>(lv[l], lv[l + 1], g.oc, cg);
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_UP, 0));
  if (mv4_ok(d)) {
    const int nt = mv4_ntiles(d);
    if (mv4_tiles(d).gx > 1)
      k_mg_up4<OP, true><<<nt, 256, 0, s>>>(d, m.x1, r32, m.dinv, m.kap32, m.pse32, sc, m.omega, g.oc, lv[1], z, cg,
                                            partials);
    else
      k_mg_up4<OP, false><<<nt, 256, 0, s>>>(d, m.x1, r32, m.dinv, m.kap32, m.pse32, sc, m.omega, g.oc, lv[1], z,
                                             cg, partials);
  } else {
    k_mg_prolong1<<<col_grid(d), dim3(TX, TY, 1), 0, s>>>(d, m.x1, lv[1], g.oc, m.t, cg);
    k_mg_up1<OP><<<stencil_ntiles<float>(d), 256, 0, s>>>(d, m.t, r32, m.dinv, m.kap32, m.pse32, sc, m.omega, z, cg,
                                                          partials);
  }
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_UP, 1));
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

template <int OP>
int vcycle_fused_t(MgWork& m, const Dims& d, const float* r_in, const float* w32, bool use_alpha, float* r_out,
                   float* z, const CgDev* cg, double* partials, bool tm) {
  hipStream_t s = stream();
  const MgCfg& g = mg_cfg();
  const float sc = (float)m.scale;
  MgLevel* lv = m.L.lv;
  const int nt = mv4_ntiles(d);
  const bool edge = mv4_tiles(d).gx > 1;
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_DOWN, 0));
  if (edge)
    k_mg_down4f<OP, true><<<nt, 256, 0, s>>>(d, r_in, w32, m.dinv, use_alpha, m.omega, r_out, m.kap32, m.pse32, sc,
                                             lv[1], cg, partials);
  else
    k_mg_down4f<OP, false><<<nt, 256, 0, s>>>(d, r_in, w32, m.dinv, use_alpha, m.omega, r_out, m.kap32, m.pse32,
                                              sc, lv[1], cg, partials);
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_DOWN, 1));
  down_levels(m, cg, s);
  k_mg_bottom<<<1, 1024, 0, s>>>(m.dlv, m.lb, m.nlev, m.ainv, g.oc, cg);
  for (int l = m.lb - 1; l >= 1; l--) k_mg_up<<<blocks256(lvl_n(lv[l])), 256, 0, s>>>(lv[l], lv[l + 1], g.oc, cg);
  const float* r = r_out ? r_out : r_in;
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_UP, 0));
  if (edge)
    k_mg_up4f<OP, true><<<nt, 256, 0, s>>>(d, r, m.dinv, m.kap32, m.pse32, sc, m.omega, g.oc, lv[1], z, cg,
                                           partials);
  else
    k_mg_up4f<OP, false><<<nt, 256, 0, s>>>(d, r, m.dinv, m.kap32, m.pse32, sc, m.omega, g.oc, lv[1], z, cg,
                                            partials);
  if (tm) ELSIM_TRY(prof_event(ELSIM_PROF_MG_UP, 1));
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

}  // namespace

int MgWork::alloc(const Dims& dd) {
  if (d.n1 == dd.n1 && d.n2 == dd.n2 && d.n3 == dd.n3 && pool.p) return ELSIM_OK;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  MgLevels Ls{};
  int m1 = dd.n1, m2 = dd.n2, m3 = dd.n3, nl = 0;
  size_t bytes = 0;
  do {
    if (nl == kMgMaxLev) {
      set_error("multigrid: more than %d levels", kMgMaxLev);
      return ELSIM_ERR_ARG;
    }
    m1 = (m1 + 1) / 2;
    m2 = (m2 + 1) / 2;
    m3 = (m3 + 1) / 2;
    nl++;
    Ls.lv[nl].m1 = m1;
    Ls.lv[nl].m2 = m2;
    Ls.lv[nl].m3 = m3;
    bytes += 8 * up((size_t)m1 * m2 * m3 * sizeof(float));
  } while ((int64_t)m1 * m2 * m3 > kMgCoarseMax);
  const size_t Ge = ((size_t)dd.G() + 1) & ~(size_t)1;
  bytes += 5 * up(Ge * sizeof(float)) + up((size_t)kMgCoarseMax * kMgCoarseMax * sizeof(float)) +
           up(sizeof(MgLevels));
  ELSIM_TRY(dev_alloc(pool, bytes));
  unsigned char* p = (unsigned char*)pool.p;
  for (int l = 1; l <= nl; l++) {
    const size_t b = up((size_t)lvl_n(Ls.lv[l]) * sizeof(float));
    float** arr[] = {&Ls.lv[l].cx, &Ls.lv[l].cy, &Ls.lv[l].cz, &Ls.lv[l].dg, &Ls.lv[l].wd,
                     &Ls.lv[l].ms, &Ls.lv[l].b, &Ls.lv[l].x};
    for (float** a : arr) {
      *a = (float*)p;
      p += b;
    }
  }
  float** fine[] = {&kap32, &pse32, &dinv, &x1, &t};
  for (float** a : fine) {
    *a = (float*)p;
    p += up(Ge * sizeof(float));
  }
  ainv = (float*)p;
  p += up((size_t)kMgCoarseMax * kMgCoarseMax * sizeof(float));
  dlv = (MgLevel*)p;
  // every fine vector keeps zero ghosts (only interiors are written)
  ELSIM_HIP(hipMemsetAsync(pool.p, 0, bytes, stream()));
  L = Ls;
  ELSIM_HIP(hipMemcpyAsync(dlv, L.lv, sizeof(MgLevels), hipMemcpyHostToDevice, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  nlev = nl;
  lb = nl;
  int64_t bottom_cells = lvl_n(L.lv[nl]);
  while (lb > 1 && lvl_n(L.lv[lb - 1]) <= kMgBottomMax && bottom_cells + lvl_n(L.lv[lb - 1]) <= kMgBottomCells) {
    lb--;
    bottom_cells += lvl_n(L.lv[lb]);
  }
  d = dd;
  op = -1;
  coef_id = 0;
  return ELSIM_OK;
}

void MgWork::release() {
  dev_free(pool);
  *this = MgWork{};
}

int mg_partial_blocks(const Dims& d) { return mv4_ok(d) ? mv4_ntiles(d) : stencil_ntiles<float>(d); }

int mg_setup(MgWork& m, int op, const Dims& d, const double* kap, const double* pse, double scale,
             uint64_t coef_id) {
  if (mg_partial_blocks(d) > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  ELSIM_TRY(m.alloc(d));
  const float omega = mg_cfg().omega;
  if (coef_id != 0 && coef_id == m.coef_id && op == m.op && scale == m.scale && omega == m.omega) return ELSIM_OK;
  switch (op) {
    case ELSIM_OP_CNE: ELSIM_TRY(setup_t<ELSIM_OP_CNE>(m, d, kap, pse, scale)); break;
    case ELSIM_OP_PHP: ELSIM_TRY(setup_t<ELSIM_OP_PHP>(m, d, kap, pse, scale)); break;
    case ELSIM_OP_PHE: ELSIM_TRY(setup_t<ELSIM_OP_PHE>(m, d, kap, pse, scale)); break;
    case ELSIM_OP_POT_DD: ELSIM_TRY(setup_t<ELSIM_OP_POT_DD>(m, d, kap, pse, scale)); break;
    default: set_error("unknown operator %d", op); return ELSIM_ERR_ARG;
  }
  m.op = op;
  m.coef_id = coef_id;
  m.scale = scale;
  m.omega = omega;
  return ELSIM_OK;
}

int mg_update(MgWork& m, const Dims& d, float* r32, const float* w32, const CgDev* cg) {
  const int64_t G2 = (d.G() + 1) / 2;
  const int nb = (int)((G2 + (int64_t)FLAT_T * MG_P - 1) / ((int64_t)FLAT_T * MG_P));
  k_mg_update<<<nb, FLAT_T, 0, stream()>>>(G2, (float2*)r32, (const float2*)w32, (const float2*)m.dinv,
                                           (float2*)m.x1, m.omega, cg);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int mg_presmooth(MgWork& m, const Dims& d, const float* r32, const CgDev* cg) {
  const int64_t G2 = (d.G() + 1) / 2;
  const int nb = (int)((G2 + (int64_t)FLAT_T * MG_P - 1) / ((int64_t)FLAT_T * MG_P));
  k_mg_presmooth<<<nb, FLAT_T, 0, stream()>>>(G2, (const float2*)r32, (const float2*)m.dinv, (float2*)m.x1, m.omega,
                                              cg);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int mg_vcycle(MgWork& m, int op, const Dims& d, const float* r32, float* z, const CgDev* cg, double* partials,
              bool timed) {
  if (m.d.n1 != d.n1 || m.d.n2 != d.n2 || m.d.n3 != d.n3 || !m.pool.p || m.op != op) {
    set_error("multigrid hierarchy not set up for this grid and operator");
    return ELSIM_ERR_ARG;
  }
  switch (op) {
    case ELSIM_OP_CNE: return vcycle_t<ELSIM_OP_CNE>(m, d, r32, z, cg, partials, timed);
    case ELSIM_OP_PHP: return vcycle_t<ELSIM_OP_PHP>(m, d, r32, z, cg, partials, timed);
    case ELSIM_OP_PHE: return vcycle_t<ELSIM_OP_PHE>(m, d, r32, z, cg, partials, timed);
    case ELSIM_OP_POT_DD: return vcycle_t<ELSIM_OP_POT_DD>(m, d, r32, z, cg, partials, timed);
  }
  set_error("unknown operator %d", op);
  return ELSIM_ERR_ARG;
}

bool mg_fused_ok(const Dims& d) { return mv4_ok(d); }

int mg_vcycle_fused(MgWork& m, int op, const Dims& d, const float* r_in, const float* w32, bool use_alpha,
                    float* r_out, float* z, const CgDev* cg, double* partials, bool timed) {
  if (m.d.n1 != d.n1 || m.d.n2 != d.n2 || m.d.n3 != d.n3 || !m.pool.p || m.op != op || !mv4_ok(d)) {
    set_error("multigrid hierarchy not set up for this grid and operator");
    return ELSIM_ERR_ARG;
  }
  switch (op) {
    case ELSIM_OP_CNE: return vcycle_fused_t<ELSIM_OP_CNE>(m, d, r_in, w32, use_alpha, r_out, z, cg, partials, timed);
    case ELSIM_OP_PHP: return vcycle_fused_t<ELSIM_OP_PHP>(m, d, r_in, w32, use_alpha, r_out, z, cg, partials, timed);
    case ELSIM_OP_PHE: return vcycle_fused_t<ELSIM_OP_PHE>(m, d, r_in, w32, use_alpha, r_out, z, cg, partials, timed);
    case ELSIM_OP_POT_DD:
      return vcycle_fused_t<ELSIM_OP_POT_DD>(m, d, r_in, w32, use_alpha, r_out, z, cg, partials, timed);
  }
  set_error("unknown operator %d", op);
  return ELSIM_ERR_ARG;
}

}  // namespace elsim

int elsim_mg_config(double omega, double oc, double delta) {
  if (!(omega > 0.0 && omega < 1.0) || !(oc > 0.0 && oc <= 4.0) || !(delta > 0.0 && delta < 1.0)) {
    elsim::set_error("elsim_mg_config: omega in (0, 1), oc in (0, 4], delta in (0, 1)");
    return ELSIM_ERR_ARG;
  }
  elsim::MgCfg& c = elsim::mg_cfg();
  c.omega = (float)omega;
  c.oc = (float)oc;
  c.delta = delta;
  return ELSIM_OK;
}

int elsim_mg_config_get(double* omega, double* oc, double* delta) {
  const elsim::MgCfg& c = elsim::mg_cfg();
  if (omega) *omega = c.omega;
  if (oc) *oc = c.oc;
  if (delta) *delta = c.delta;
  return ELSIM_OK;
}
