// elsim_stencil.hip -- the variable-coefficient 7-point operators of the SBM
// solves and the conjugate-gradient drivers built on them.
//
// Operators (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90):
//   CnE  A_c p = psE p - dt/dx^2 S(peN, p)   SBM:515-521
//   phP  L p   = dx^-2 S(pkN, p)              SBM:670-676
//   phE  L p   = dx^-2 S(pcN, p)              SBM:751-757
//   S(k, p) = +f2(p_i+1 - p) - f1(p - p_i-1) + f4(...) - f3(...) + f6(...) - f5(...)
// evaluated in that order, with face coefficients f = 0.5(k_nb + k_c)
// averaged on the fly from ONE ghosted cell-centred array (the reference
// stores six face arrays, SBM:326-331, 479-484, 605-610; the average is the
// same IEEE operation, so results are bit-identical).
//
// Layout: ghosted (n1+2)(n2+2)(n3+2) arrays, d1 contiguous.  A stencil
// block owns 256 d1 voxels x SJ rows x SKC planes (SJ = 4 in FP32, 2 in FP64);
// each thread holds its SJ rows of p and kappa for the planes k-1, k, k+1 in
// registers, d1 neighbours come from the adjacent lanes (DPP wave shifts) and
// only the rows above and below the tile and the wave-edge lanes load from
// memory.  Measured at 256^3 from HBM (data not in the 256 MB last-level
// cache), this takes the FP32 operator from 70 us (a thread per column, 14
// loads per voxel) to 54 us.  The search direction's homogeneous ghost rule
// (SBM:511-513, 666-668, 747-749) is applied in registers, so ghost memory of
// p is never relied on.
//
// Every CG work vector lives in the same ghosted layout with its ghost cells
// held at zero, so the pointwise kernels are flat over the whole box, one
// contiguous tile per block: fully coalesced, and adding alpha*p (zero
// ghosts) to the unknown leaves its boundary ghosts untouched.
//
// Reductions are deterministic: every block stores its partial sums, then a
// one-block finalising kernel adds them in index order and updates the CG
// scalars (alpha, beta, stop tests) on the device, so the host only polls a
// done flag every few iterations.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "elsim_mg.h"
#include "elsim_solver.h"
#include "elsim_sweep.h"

namespace elsim {

// ------------------------------------------------------------- kernels --

// out = A p (hom or stored ghosts); out in the ghosted layout (OUTG) or
// interior layout; optional dot = p.out into *dot or cg->pw with
// alpha = (num_is_rz ? rz : rr) / pw.
template <typename T, int OP, bool HOM, bool OUTG>
__global__ __launch_bounds__(256) void k_matvec(Dims d, const T* __restrict__ p, const T* __restrict__ kap,
                                                const T* __restrict__ pse, T scale, T* __restrict__ out,
                                                double* dot, CgDev* cg, int num_is_rz, double* partials) {
  if (cg && (cg->done || cg->need_reliable)) return;
  double acc[1] = {0.0};
  auto epi = [&](int64_t c, int64_t n, T pc, T v) {
    out[OUTG ? c : n] = v;
    acc[0] += (double)pc * (double)v;
  };
  if constexpr (std::is_same<T, double>::value) {
    const int nt = ntiles64(d);
    for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) tile64<OP, HOM>(d, tile, p, kap, pse, scale, epi);
  } else {
    const int nt = stencil_ntiles<T>(d);
    for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
      sweep<T, OP, HOM>(d, stencil_idx<Rows<T>::SJ>(d, tile), p, kap, pse, scale, epi);
  }
  if (dot || cg) block_partials<1>(acc, partials);
}

// out = A p with the homogeneous ghost rule; out ghosted (OUTG) or interior;
// partials of p.out.  Launch: mv4_ntiles(d) blocks of 256 threads.  EDGE:
// the box has more than one d1 tile, so the first and last lane of a row
// segment may load their d1 neighbour from the next tile (n1 <= 256 leaves
// that code, and its registers, out).
template <int OP, bool OUTG, bool EDGE>
__global__ __launch_bounds__(256) void k_matvec4(Dims d, const float* __restrict__ p, const float* __restrict__ kap,
                                                 const float* __restrict__ pse, float scale, float* __restrict__ out,
                                                 CgDev* cg, double* partials) {
  if (cg && (cg->done || cg->need_reliable)) return;
  double acc = 0.0;
  const int nt = mv4_ntiles(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    mv4_tile<OP, EDGE>(
        d, tile, [&](uint32_t q, int, int, int) { return ld4(p, q); }, [&](uint32_t q, int, int, int) { return p[q]; },
        kap, pse, scale, [&](int, int j, int k, uint32_t q, const float (&pv)[4], const float (&o)[4]) {
#pragma unroll
          for (int c = 0; c < 4; c++) acc += (double)pv[c] * (double)o[c];
          const uint32_t dst = OUTG ? q : (uint32_t)d.ni((int)(q % (uint32_t)d.g1()), j, k);
          st4(out, dst, make_float4(o[0], o[1], o[2], o[3]));
        });
  if (partials) {
    double v[1] = {acc};
    block_partials<1>(v, partials);
  }
}

// Initial residual with x's ghosts as stored (SBM:492-499, 649-655, 729-735):
// res = r0 = rhs - A x; prj(interior) = res; x0 = x(interior); cg->rr = sum res^2.
// res and prj may be null (the mixed solvers start from r0 alone), and x0
// (no copy of the start wanted).
template <int OP>
__global__ __launch_bounds__(256) void k_resid64(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                                 const double* __restrict__ pse, double scale,
                                                 const double* __restrict__ rhs, double* __restrict__ res,
                                                 double* __restrict__ r0, double* __restrict__ prj,
                                                 double* __restrict__ x0, CgDev* cg, double* partials) {
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<double, OP, false>(d, stencil_idx<Rows<double>::SJ>(d, tile), x, kap, pse, scale,
                             [&](int64_t c, int64_t n, double xc, double v) {
      const double r = rhs[n] - v;
      if (res) res[c] = r;
      r0[c] = r;
      if (prj) prj[c] = r;
      if (x0) x0[c] = xc;
      acc[0] += r * r;
    });
  block_partials<1>(acc, partials);
}

// k_resid64 and ||b_eff||^2 in one kernel, b_eff = r0 + A_hom x (the
// residual of a zero interior with the boundary data of x): one sweep that
// yields both A x (stored ghosts) and A_hom x, so each voxel's r and b_eff
// are the values and summation order of a residual sweep followed by a
// homogeneous sweep of the same tile.
// Partials: slot 0 r.r, slot 1 ||b_eff||^2 (FIN_BEFF then sets tol2 and, for
// the relative rule, done if the start already meets it).
template <int OP>
__global__ __launch_bounds__(256) void k_resid64_beff(Dims d, const double* __restrict__ x,
                                                      const double* __restrict__ kap, const double* __restrict__ pse,
                                                      double scale, const double* __restrict__ rhs,
                                                      double* __restrict__ res, double* __restrict__ r0,
                                                      double* __restrict__ prj, double* __restrict__ x0,
                                                      double* partials) {
  double acc[2] = {0.0, 0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const StIdx si = stencil_idx<Rows<double>::SJ>(d, tile);
    sweep_dual<double, OP>(d, si, x, kap, pse, scale, [&](int64_t c, int64_t n, double xc, double v, double vh) {
      const double r = rhs[n] - v;
      if (res) res[c] = r;
      r0[c] = r;
      if (prj) prj[c] = r;
      if (x0) x0[c] = xc;
      acc[0] += r * r;
      const double b = r + vh;
      acc[1] += b * b;
    });
  }
  block_partials<2>(acc, partials);
}

// FP64 reference CG update (SBM:533-547, 688-705, 769-786), flat over the box.
template <int OP>
__global__ __launch_bounds__(256) void k_cg64_update(int64_t G, double* __restrict__ x, const double* __restrict__ prj,
                                                     double* __restrict__ res, const double* __restrict__ wmx,
                                                     const double* __restrict__ w8, CgDev* cg, double* partials) {
  if (cg->done) return;
  const double alp = cg->alpha;
  double acc[2] = {0.0, 0.0};
  double xo[FLAT_U], pv[FLAT_U], rv[FLAT_U], wv[FLAT_U], ww[FLAT_U];
  FLAT_LOOP(G, { xo[u] = x[q]; pv[u] = prj[q]; rv[u] = res[q]; wv[u] = wmx[q]; ww[u] = w8[q]; }, {
    const double xn = xo[u] + alp * pv[u];
    x[q] = xn;
    const double r = rv[u] - alp * wv[u];
    res[q] = r;
    acc[0] += r * r;
    if (OP == ELSIM_OP_CNE) {
      acc[1] += r * r * ww[u];
    } else {
      const double t = xn - xo[u];
      acc[1] += t * t * ww[u];
    }
  })
  block_partials<2>(acc, partials);
}

// PRJ = RES + bet*PRJ (SBM:542, 697, 778).
__global__ __launch_bounds__(256) void k_cg64_dir(int64_t G, double* __restrict__ prj, const double* __restrict__ res,
                                                  const CgDev* cg) {
  if (cg->done) return;
  const double bet = cg->beta;
  double rv[FLAT_U], pv[FLAT_U];
  FLAT_LOOP(G, { rv[u] = res[q]; pv[u] = prj[q]; }, { prj[q] = rv[u] + bet * pv[u]; })
}

// e = x - x0 (interior of a ghosted array).
__global__ __launch_bounds__(256) void k_diff(Dims d, const double* __restrict__ x, const double* __restrict__ x0,
                                              double* __restrict__ e) {
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t, int) { e[c] = x[c] - x0[c]; });
}

// FP64 true residual of the correction system: r = r0 - A_hom e.
// FINAL: only record rr_true.  Otherwise (mixed reliable update): r32 = r,
// and partials of r.r (FP64) and r32.(dinv r32), from which the FP32 Krylov
// recurrence restarts.
template <int OP, bool FINAL>
__global__ __launch_bounds__(256) void k_true_resid(Dims d, const double* __restrict__ e, const double* __restrict__ kap,
                                                    const double* __restrict__ pse, double scale,
                                                    const double* __restrict__ r0, float* __restrict__ r32,
                                                    const uint16_t* __restrict__ dinv, CgDev* cg, double* partials) {
  if (!FINAL && (cg->done || !cg->need_reliable)) return;
  double acc[2] = {0.0, 0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<double, OP, true>(d, stencil_idx<Rows<double>::SJ>(d, tile), e, kap, pse, scale,
                            [&](int64_t c, int64_t, double, double v) {
      const double r = r0[c] - v;
      acc[0] += r * r;
      if (!FINAL) {
        const float rf = (float)r;
        r32[c] = rf;
        if (dinv) acc[1] += (double)rf * (double)(bf16_to_f32(dinv[c]) * rf);
      }
    });
  block_partials<2>(acc, partials);
}

// sum over the interior of (rhs - A x)^2 with every ghost of x as stored.
template <int OP>
__global__ __launch_bounds__(256) void k_resid_norm(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                                    const double* __restrict__ pse, double scale,
                                                    const double* __restrict__ rhs, double* partials) {
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<double, OP, false>(d, stencil_idx<Rows<double>::SJ>(d, tile), x, kap, pse, scale,
                             [&](int64_t, int64_t n, double, double v) {
                               const double r = rhs[n] - v;
                               acc[0] += r * r;
                             });
  block_partials<1>(acc, partials);
}

// r(interior) = rhs - A x (x's ghosts as stored) and the sum of r^2.
template <int OP>
__global__ __launch_bounds__(256) void k_resid_store(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                                     const double* __restrict__ pse, double scale,
                                                     const double* __restrict__ rhs, double* __restrict__ r,
                                                     double* partials) {
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<double, OP, false>(d, stencil_idx<Rows<double>::SJ>(d, tile), x, kap, pse, scale,
                             [&](int64_t, int64_t n, double, double v) {
                               const double q = rhs[n] - v;
                               r[n] = q;
                               acc[0] += q * q;
                             });
  block_partials<1>(acc, partials);
}

// sum over the interior of dv * (rhs - A x), x's ghosts as stored.
template <int OP>
__global__ __launch_bounds__(256) void k_dot_resid(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                                   const double* __restrict__ pse, double scale,
                                                   const double* __restrict__ rhs, const double* __restrict__ dv,
                                                   double* partials) {
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<double, OP, false>(d, stencil_idx<Rows<double>::SJ>(d, tile), x, kap, pse, scale,
                             [&](int64_t c, int64_t n, double, double v) { acc[0] += dv[c] * (rhs[n] - v); });
  block_partials<1>(acc, partials);
}

// num = dv.(rhs - A x) (x's ghosts as stored) and den = dv.(A dv) (the
// homogeneous rule): the two dot products of a Galerkin step along dv, in
// one kernel -- one walk over x and dv together, sharing the coefficient's
// loads -- instead of k_dot_resid and a k_matvec whose output only fed its
// dot product.  Same per-tile sums as those two kernels.
// Partials: slot 0 num, slot 1 den.
template <int OP>
__global__ __launch_bounds__(256) void k_galerkin_dots(Dims d, const double* __restrict__ x,
                                                       const double* __restrict__ kap, double scale,
                                                       const double* __restrict__ rhs, const double* __restrict__ dv,
                                                       double* partials) {
  double acc[2] = {0.0, 0.0};
  const int nt = stencil_ntiles<double>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const StIdx si = stencil_idx<Rows<double>::SJ>(d, tile);
    sweep_pair<OP>(d, si, x, dv, kap, nullptr, scale,
                   [&](int64_t, int64_t n, double, double v, double pc, double vh) {
                     acc[0] += pc * (rhs[n] - v);
                     acc[1] += pc * vh;
                   });
  }
  block_partials<2>(acc, partials);
}

// ------------------------------------------------------------- Jacobi --



// ------------------------------------------------------ FP64 Jacobi-PCG --

// p = dinv r0, rz = r.z, e = 0
__global__ __launch_bounds__(256) void k_pcg64_start(int64_t G, const double* __restrict__ res,
                                                     const double* __restrict__ dinv, double* __restrict__ prj,
                                                     double* __restrict__ e, CgDev* cg, double* partials) {
  double acc[1] = {0.0};
  double rv[FLAT_U], dv[FLAT_U];
  FLAT_LOOP(G, { rv[u] = res[q]; dv[u] = dinv[q]; }, {
    const double z = dv[u] * rv[u];
    prj[q] = z;
    e[q] = 0.0;
    acc[0] += rv[u] * z;
  })
  block_partials<1>(acc, partials);
}

// e += alpha p; r -= alpha w; rz' = r.(dinv r), rr = r.r
__global__ __launch_bounds__(256) void k_pcg64_update(int64_t G, double* __restrict__ x, const double* __restrict__ prj,
                                                      double* __restrict__ res, const double* __restrict__ wmx,
                                                      const double* __restrict__ dinv, CgDev* cg, double* partials) {
  if (cg->done) return;
  const double alp = cg->alpha;
  double acc[2] = {0.0, 0.0};
  double xv[FLAT_U], pv[FLAT_U], rv[FLAT_U], wv[FLAT_U], dv[FLAT_U];
  FLAT_LOOP(G, { xv[u] = x[q]; pv[u] = prj[q]; rv[u] = res[q]; wv[u] = wmx[q]; dv[u] = dinv[q]; }, {
    x[q] = xv[u] + alp * pv[u];
    const double r = rv[u] - alp * wv[u];
    res[q] = r;
    acc[0] += r * (dv[u] * r);
    acc[1] += r * r;
  })
  block_partials<2>(acc, partials);
}

// p = dinv r + beta p
__global__ __launch_bounds__(256) void k_pcg64_dir(int64_t G, double* __restrict__ prj, const double* __restrict__ res,
                                                   const double* __restrict__ dinv, const CgDev* cg) {
  if (cg->done) return;
  const double bet = cg->beta;
  double rv[FLAT_U], pv[FLAT_U], dv[FLAT_U];
  FLAT_LOOP(G, { rv[u] = res[q]; pv[u] = prj[q]; dv[u] = dinv[q]; }, { prj[q] = dv[u] * rv[u] + bet * pv[u]; })
}

// ------------------------------------------------ mixed-precision kernels --

// b = (float) a on the ghost cells of the box (the interior: k_dinv).
__global__ __launch_bounds__(256) void k_ghost_f32(Dims d, const double* __restrict__ a, float* __restrict__ b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ghost_cells(d)) return;
  int i, j, k;
  ghost_cell(d, q, i, j, k);
  const int64_t c = d.gi(i, j, k);
  b[c] = (float)a[c];
}

// The vector kernels of the mixed solver work on element pairs: a lane
// loads 8 B of each FP32 vector and one 4 B word holding two BF16 inverse
// diagonals (16-bit loads measured slower).  The FP32 work vectors are
// allocated with an even length; the padding element is zero and stays zero.
struct BfPair {
  float a, b;
};
__device__ __forceinline__ BfPair bf16x2(uint32_t w) {
  return BfPair{bf16_to_f32((uint16_t)(w & 0xffffu)), bf16_to_f32((uint16_t)(w >> 16))};
}

// UPD_P element pairs per thread in the update and start kernels (4096
// elements per block: a quarter of the blocks of the other flat kernels, so
// the finalising kernel that follows reads a quarter of the partials; 2, 4,
// 8 and 16 pairs measured at 256^3: 8 is the fastest).
constexpr int UPD_P = 8;

// Mixed start: r32 = r0, p32 = dinv r32 (the first direction), y32 = 0,
// e64 = 0; partials of r.r and r.z.
__global__ __launch_bounds__(256) void k_mx_start(int64_t G2, const double2* __restrict__ r0,
                                                  const uint32_t* __restrict__ dinv, float2* __restrict__ r32,
                                                  float2* __restrict__ p32, float2* __restrict__ y32,
                                                  double2* __restrict__ e64, double* partials) {
  double acc[2] = {0.0, 0.0};
  double2 rv[UPD_P];
  uint32_t dv[UPD_P];
  FLAT_LOOP_U(UPD_P, G2, { rv[u] = r0[q]; dv[u] = dinv[q]; }, {
    const BfPair dd = bf16x2(dv[u]);
    float2 r;
    float2 z;
    r.x = (float)rv[u].x;
    r.y = (float)rv[u].y;
    z.x = dd.a * r.x;
    z.y = dd.b * r.y;
    r32[q] = r;
    p32[q] = z;
    y32[q] = make_float2(0.f, 0.f);
    e64[q] = make_double2(0.0, 0.0);
    acc[0] += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
    acc[1] += (double)r.x * (double)z.x + (double)r.y * (double)z.y;
  })
  block_partials<2>(acc, partials);
}

// One iteration of the mixed solver is the FP32 operator (with the p.Ap
// partials), then the two flat kernels below, each followed by a one-block
// finalising kernel.  (Forming alpha inside the update kernel from the
// operator's partials instead, to save the first finaliser, made the update
// kernel 7 us slower at 256^3, more than the 4.7 us the finaliser takes;
// finalising in the last block of the producer needs device-scope fences,
// which write back the L2 of every XCD: 384 us per iteration.)
// r -= alpha w; partials of r.z and r.r.  The solution update of the same
// iteration is left to k_mx_dir3, which reads p anyway.
__global__ __launch_bounds__(256) void k_mx_upd3(int64_t G2, float2* __restrict__ r32, const float2* __restrict__ w32,
                                                 const uint32_t* __restrict__ dinv, CgDev* cg, double* partials) {
  if (cg->done || cg->need_reliable) return;
  const float alp = (float)cg->alpha;
  double acc[2] = {0.0, 0.0};
  float2 rv[UPD_P], wv[UPD_P];
  uint32_t dv[UPD_P];
  FLAT_LOOP_U(UPD_P, G2, { rv[u] = r32[q]; wv[u] = w32[q]; dv[u] = dinv[q]; }, {
    const BfPair dd = bf16x2(dv[u]);
    float2 rn;
    rn.x = rv[u].x - alp * wv[u].x;
    rn.y = rv[u].y - alp * wv[u].y;
    r32[q] = rn;
    acc[0] += (double)rn.x * (double)(dd.a * rn.x) + (double)rn.y * (double)(dd.b * rn.y);
    acc[1] += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
  })
  block_partials<2>(acc, partials);
}

// y += alpha p (the solution update of the iteration whose residual update
// ran, flagged by pending_y), then p = dinv r + beta p unless the iteration
// ended the batch.  DIR_P element pairs per thread.
constexpr int DIR_P = 2;
__global__ __launch_bounds__(256) void k_mx_dir3(int64_t G2, float2* __restrict__ y32, float2* __restrict__ p32,
                                                 const float2* __restrict__ r32, const uint32_t* __restrict__ dinv,
                                                 const CgDev* cg) {
  const bool upd_y = cg->pending_y != 0, upd_p = !(cg->done || cg->need_reliable);
  const float alp = (float)cg->alpha, bet = (float)cg->beta;
  if (upd_y && upd_p) {
    float2 yv[DIR_P], rv[DIR_P], pv[DIR_P];
    uint32_t dv[DIR_P];
    FLAT_LOOP_U(DIR_P, G2, { yv[u] = y32[q]; rv[u] = r32[q]; pv[u] = p32[q]; dv[u] = dinv[q]; }, {
      const BfPair dd = bf16x2(dv[u]);
      float2 yn;
      float2 pn;
      yn.x = yv[u].x + alp * pv[u].x;
      yn.y = yv[u].y + alp * pv[u].y;
      pn.x = dd.a * rv[u].x + bet * pv[u].x;
      pn.y = dd.b * rv[u].y + bet * pv[u].y;
      y32[q] = yn;
      p32[q] = pn;
    })
  } else if (upd_y) {
    float2 yv[DIR_P], pv[DIR_P];
    FLAT_LOOP_U(DIR_P, G2, { yv[u] = y32[q]; pv[u] = p32[q]; }, {
      float2 yn;
      yn.x = yv[u].x + alp * pv[u].x;
      yn.y = yv[u].y + alp * pv[u].y;
      y32[q] = yn;
    })
  } else if (upd_p) {
    float2 rv[DIR_P], pv[DIR_P];
    uint32_t dv[DIR_P];
    FLAT_LOOP_U(DIR_P, G2, { rv[u] = r32[q]; pv[u] = p32[q]; dv[u] = dinv[q]; }, {
      const BfPair dd = bf16x2(dv[u]);
      float2 pn;
      pn.x = dd.a * rv[u].x + bet * pv[u].x;
      pn.y = dd.b * rv[u].y + bet * pv[u].y;
      p32[q] = pn;
    })
  }
}

// ------------------------------------- multigrid-preconditioned mixed CG --
// The same FP32 recurrence with z = M^-1 r from one V-cycle (elsim_mg.hip)
// written into w32 once the update kernel has consumed A p there.

// Start: r32 = r0, y32 = p32 = 0 (the first direction is z itself), e64 = 0.
__global__ __launch_bounds__(256) void k_mg_start(int64_t G2, const double2* __restrict__ r0, float2* __restrict__ r32,
                                                  float2* __restrict__ p32, float2* __restrict__ y32,
                                                  double2* __restrict__ e64) {
  double2 rv[UPD_P];
  FLAT_LOOP_U(UPD_P, G2, { rv[u] = r0[q]; }, {
    r32[q] = make_float2((float)rv[u].x, (float)rv[u].y);
    p32[q] = make_float2(0.f, 0.f);
    y32[q] = make_float2(0.f, 0.f);
    e64[q] = make_double2(0.0, 0.0);
  })
}

// y += alpha p (pending_y), then p = z + beta p unless the iteration ended
// the batch.
__global__ __launch_bounds__(256) void k_mg_dir3(int64_t G2, float2* __restrict__ y32, float2* __restrict__ p32,
                                                 const float2* __restrict__ z32, const CgDev* cg) {
  const bool upd_y = cg->pending_y != 0, upd_p = !(cg->done || cg->need_reliable);
  const float alp = (float)cg->alpha, bet = (float)cg->beta;
  if (!upd_y && !upd_p) return;
  float2 yv[DIR_P], zv[DIR_P], pv[DIR_P];
  FLAT_LOOP_U(DIR_P, G2, {
    pv[u] = p32[q];
    if (upd_y) yv[u] = y32[q];
    if (upd_p) zv[u] = z32[q];
  }, {
    if (upd_y) {
      float2 yn;
      yn.x = yv[u].x + alp * pv[u].x;
      yn.y = yv[u].y + alp * pv[u].y;
      y32[q] = yn;
    }
    if (upd_p) {
      float2 pn;
      pn.x = zv[u].x + bet * pv[u].x;
      pn.y = zv[u].y + bet * pv[u].y;
      p32[q] = pn;
    }
  })
}

// Reliable update, part 1: e64 += y32, y32 = 0.
__global__ __launch_bounds__(256) void k_mx_fold(int64_t G, double* __restrict__ e64, float* __restrict__ y32,
                                                 const CgDev* cg) {
  if (cg->done || !cg->need_reliable) return;
  double ev[FLAT_U];
  float yv[FLAT_U];
  FLAT_LOOP(G, { ev[u] = e64[q]; yv[u] = y32[q]; }, {
    e64[q] = ev[u] + (double)yv[u];
    y32[q] = 0.0f;
  })
}

// x += e (ghosts of e are zero).
__global__ __launch_bounds__(256) void k_add_e(int64_t G, double* __restrict__ x, const double* __restrict__ e) {
  double xv[FLAT_U], ev[FLAT_U];
  FLAT_LOOP(G, { xv[u] = x[q]; ev[u] = e[q]; }, { x[q] = xv[u] + ev[u]; })
}

__global__ __launch_bounds__(256) void k_copy_g2n(Dims d, const double* __restrict__ g, double* __restrict__ nn) {
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n, int) { nn[n] = g[c]; });
}

// ---------------------------------------------------------- finalising --

enum FinOp {
  FIN_DOT = 0,        // *dot = sum
  FIN_ALPHA_RR,       // pw = sum, alpha = rr / pw
  FIN_ALPHA_RZ,       // pw = sum, alpha = rz / pw
  FIN_RESID,          // rr = rmax2 = sum
  FIN_BEFF,           // bn2, tol2, relative-rule early exit
  FIN_CG64,           // reference CG: beta, stop rules
  FIN_TRUE_FINAL,     // rr_true
  FIN_TRUE_RELIABLE,  // mixed reliable update: done, cap, or restart from the true residual
  FIN_RZ,             // rz = sum
  FIN_PCG64,          // FP64 Jacobi-PCG: beta, stop
  FIN_MX_UPDATE,      // three-kernel mixed path: beta or request a reliable update
  FIN_MX_START,       // three-kernel mixed path: rz of the start (second value)
  FIN_MG_UPDATE,      // multigrid path: r.r (first value) and r.z (second) after the V-cycle
  FIN_MG_START        // multigrid path: rz of the start or restart, beta = 0
};

// The skip test repeats the producer's, evaluated on the same (unchanged)
// scalars, so a skipped producer is never finalised.
template <int NV>
__global__ __launch_bounds__(FIN_T) void k_fin(int op, int nb, const double* __restrict__ partials, CgDev* cg,
                                               double* dot, double rtol) {
  bool skip = false;
  if (cg) {
    if (op == FIN_ALPHA_RR || op == FIN_ALPHA_RZ || op == FIN_MX_UPDATE || op == FIN_MG_UPDATE)
      skip = cg->done || cg->need_reliable;
    if (op == FIN_CG64 || op == FIN_PCG64 || op == FIN_MG_START) skip = cg->done;
    if (op == FIN_TRUE_RELIABLE) skip = cg->done || !cg->need_reliable;
  }
  // three-kernel mixed path: the iteration whose alpha this is will update
  // r, and its y += alpha p falls to k_mx_dir3
  if (cg && op == FIN_ALPHA_RZ && linear_tid() == 0) cg->pending_y = skip ? 0 : 1;
  if (skip) return;
  // the operations of the two-value reductions read tot[1]; with NV = 1 it
  // stays zero (those operations are only launched with NV = 2)
  double tot[2] = {0.0, 0.0};
  {
    double t[NV];
    sum_partials<NV>(partials, nb, t);
#pragma unroll
    for (int q = 0; q < NV; q++) tot[q] = t[q];
  }
  if (linear_tid() != 0) return;
  switch (op) {
    case FIN_DOT: *dot = tot[0]; break;
    case FIN_ALPHA_RR:
    case FIN_ALPHA_RZ:
      cg->pw = tot[0];
      cg->alpha = (op == FIN_ALPHA_RZ ? cg->rz : cg->rr) / tot[0];
      break;
    case FIN_RESID:
      cg->rr = tot[0];
      cg->rmax2 = tot[0];
      cg->rr0 = tot[0];
      break;
    case FIN_BEFF:
      cg->bn2 = cg->rel_r0 ? cg->rr : tot[0];
      cg->tol2 = rtol * rtol * cg->bn2;
      cg->rr_true = cg->rr;
      if (!isfinite(cg->rr) || !isfinite(tot[0]))
        cg->done = 3;  // a non-finite start or right-hand side is an error, never "converged"
      else if (cg->stop_rule == ELSIM_STOP_RELRES && !(cg->rr > cg->tol2))
        cg->done = 1;
      break;
    case FIN_CG64: {
      const double alp = cg->alpha, dmy = cg->rr, rrn = tot[0];
      const double bet = rrn / dmy;
      cg->beta = bet;
      cg->rr = rrn;
      cg->iters += 1;
      if (!isfinite(alp) || !isfinite(bet)) {
        cg->done = 3;
      } else if (cg->stop_rule == ELSIM_STOP_REFERENCE) {
        const double st = sqrt(tot[1] / cg->tw);
        cg->stop = st;
        if (!(st > cg->crt)) cg->done = 1;
      } else if (!(rrn > cg->tol2)) {
        cg->done = 1;
      }
      if (!cg->done && cg->iters > cg->maxit) cg->done = 2;
      break;
    }
    case FIN_TRUE_FINAL: cg->rr_true = tot[0]; break;
    case FIN_TRUE_RELIABLE:
      cg->rr_true = tot[0];
      cg->n_reliable += 1;
      cg->need_reliable = 0;
      if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        cg->done = 3;
      } else if (!(tot[0] > cg->tol2)) {
        cg->done = 1;
      } else if (cg->iters >= cg->maxit) {
        cg->done = 2;
      } else {
        // Restart the FP32 recurrence from the true residual: the next
        // k_mx_dir3 forms p = dinv r (beta = 0).  Against continuing the old
        // direction with the replaced residual (beta = r.z ratio), and
        // against restarts with delta = 0.1, the mean over six 256^3 packs x
        // 48 warm-started steps was 0.159 s/step (restart, delta 0.5), 0.168
        // and 0.165 s/step: within the spread between packs (0.09-0.24 s
        // for one setting), which the current control's sign changes set.
        cg->pending_y = 0;
        cg->beta = 0.0;
        cg->rz = tot[1];
        cg->rmax2 = tot[0];
      }
      break;
    case FIN_RZ: cg->rz = tot[0]; break;
    case FIN_MX_START: cg->rz = tot[1]; break;
    case FIN_MX_UPDATE:
      cg->iters += 1;
      cg->rr = tot[1];
      if (!isfinite(cg->alpha) || !isfinite(tot[0])) {
        cg->done = 3;
      } else if (!(tot[1] > cg->delta2 * cg->rmax2) || !(tot[1] > cg->tol2) || cg->iters >= cg->maxit) {
        cg->need_reliable = 1;
      } else {
        cg->beta = tot[0] / cg->rz;
        cg->rz = tot[0];
      }
      break;
    case FIN_MG_UPDATE:
      cg->iters += 1;
      cg->rr = tot[0];
      if (!isfinite(cg->alpha) || !isfinite(tot[0]) || !isfinite(tot[1])) {
        cg->done = 3;
      } else if (!(tot[0] > cg->delta2 * cg->rmax2) || !(tot[0] > cg->tol2) || cg->iters >= cg->maxit) {
        cg->need_reliable = 1;
      } else {
        cg->beta = tot[1] / cg->rz;
        cg->rz = tot[1];
      }
      break;
    case FIN_MG_START:
      if (!isfinite(tot[1])) {
        cg->done = 3;
      } else {
        cg->rz = tot[1];
        cg->beta = 0.0;
      }
      break;
    case FIN_PCG64: {
      const double alp = cg->alpha;
      cg->iters += 1;
      cg->rr = tot[1];
      const double bet = tot[0] / cg->rz;
      cg->beta = bet;
      cg->rz = tot[0];
      if (!isfinite(alp) || !isfinite(bet))
        cg->done = 3;
      else if (!(tot[1] > cg->tol2))
        cg->done = 1;
      else if (cg->iters >= cg->maxit)
        cg->done = 2;
      break;
    }
  }
}

// ------------------------------------------------------------------ host --

// Zero fill of n bytes (n a multiple of 8): a kernel rather than a memset,
// so that the allocation-free path records only kernels.
__global__ void k_zero8(int64_t n8, uint64_t* __restrict__ p) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n8; q += (int64_t)gridDim.x * blockDim.x) p[q] = 0;
}

size_t CgWork::bound_bytes(const Dims& dd) {
  const size_t G = (size_t)dd.G(), Ge = (G + 1) & ~(size_t)1;
  const size_t v64 = (Ge * 8 + 255) & ~(size_t)255, v32 = (Ge * 4 + 255) & ~(size_t)255;
  const size_t part = 2 * (size_t)kMaxRedBlocks * 8;  // two-value reductions (slots 0 and 1)
  return 7 * v64 + 7 * v32 + part + 256;
}

int CgWork::bind(const Dims& dd, unsigned char* base, size_t bytes, CgDev* hostp) {
  if (!base || !hostp || bound_bytes(dd) > bytes) {
    set_error("work space of %zu bytes does not fit %zu bytes", bytes, bound_bytes(dd));
    return ELSIM_ERR_ARG;
  }
  if (!bound) release();
  bound = true;
  d = dd;
  const size_t G = (size_t)d.G(), Ge = (G + 1) & ~(size_t)1;
  const size_t v64 = (Ge * 8 + 255) & ~(size_t)255, v32 = (Ge * 4 + 255) & ~(size_t)255;
  unsigned char* q = base;
  DevBuf* g64[] = {&res, &wmx, &prj, &x0, &r0, &e64, &dinv64};
  DevBuf* g32[] = {&y32, &r32, &w32, &p32, &kap32, &pse32, &dinv32};
  for (DevBuf* b : g64) {
    b->p = q;
    b->bytes = v64;
    q += v64;
  }
  for (DevBuf* b : g32) {
    b->p = q;
    b->bytes = v32;
    q += v32;
  }
  partials = (double*)q;
  q += 2 * (size_t)kMaxRedBlocks * 8;
  cg.p = q;
  cg.bytes = 256;
  host = hostp;
  // every work vector starts with zero ghosts (the last call left its own
  // layout here); the partials and scalars are written before they are read
  const int64_t n8 = (int64_t)((7 * v64 + 7 * v32) / 8);
  k_zero8<<<(int)std::min<int64_t>((n8 + 255) / 256, 65536), 256, 0, stream()>>>(n8, (uint64_t*)base);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int CgWork::alloc(const Dims& dd) {
  if (bound) {
    set_error("work space bound to caller memory");
    return ELSIM_ERR_ARG;
  }
  if (d.n1 == dd.n1 && d.n2 == dd.n2 && d.n3 == dd.n3 && cg.p) return ELSIM_OK;
  d = dd;
  const size_t G = (size_t)d.G(), Ge = (G + 1) & ~(size_t)1;  // even length: pair kernels
  DevBuf* g64[] = {&res, &wmx, &prj, &x0, &r0, &e64, &dinv64};
  DevBuf* g32[] = {&y32, &r32, &w32, &p32, &kap32, &pse32, &dinv32};
  for (DevBuf* b : g64) ELSIM_TRY(dev_alloc(*b, Ge * 8));
  for (DevBuf* b : g32) ELSIM_TRY(dev_alloc(*b, Ge * 4));
  ELSIM_TRY(dev_alloc(cg, sizeof(CgDev)));
  if (!host) ELSIM_HIP(hipHostMalloc((void**)&host, sizeof(CgDev), hipHostMallocDefault));
  // every work vector keeps zero ghosts (the kernels only write interiors
  // or add zero ghosts to zero ghosts)
  for (DevBuf* b : g64) ELSIM_HIP(hipMemsetAsync(b->p, 0, Ge * 8, stream()));
  for (DevBuf* b : g32) ELSIM_HIP(hipMemsetAsync(b->p, 0, Ge * 4, stream()));
  return ELSIM_OK;
}

void CgWork::release() {
  if (bound) {  // caller memory: forget the layout only
    *this = CgWork{};
    return;
  }
  DevBuf* bs[] = {&res, &wmx, &prj, &x0, &r0, &e64, &y32, &r32, &w32, &p32, &kap32, &pse32, &dinv32, &dinv64,
                  &cg};
  for (DevBuf* b : bs) dev_free(*b);
  for (MgWork*& m : mg)
    if (m) {
      m->release();
      delete m;
      m = nullptr;
    }
  if (host) (void)hipHostFree(host);
  host = nullptr;
  d = Dims{0, 0, 0};
}

namespace {

// Optional event brackets around the first CG iteration of each batch.
struct Prof {
  bool on = false, ready = false;
  int op = -1;  // time the solves of this operator only (-1: every solve)
  bool rec[ELSIM_PROF_N] = {};  // slots bracketed since the last collection
  hipEvent_t ev[2 * ELSIM_PROF_N];
  elsim_kernel_prof acc{};
};
Prof g_prof;

template <class F>
int timed(bool on, int slot, hipStream_t s, F&& launch) {
  if (on) ELSIM_HIP(hipEventRecord(g_prof.ev[2 * slot], s));
  launch();
  if (on) {
    ELSIM_HIP(hipEventRecord(g_prof.ev[2 * slot + 1], s));
    g_prof.rec[slot] = true;
  }
  return ELSIM_OK;
}

// After a synchronisation: add the bracketed durations of the slots.
int prof_collect(int first, int last) {
  for (int sl = first; sl <= last; sl++) {
    if (!g_prof.rec[sl]) continue;
    float ms = 0.f;
    ELSIM_HIP(hipEventElapsedTime(&ms, g_prof.ev[2 * sl], g_prof.ev[2 * sl + 1]));
    g_prof.acc.count[sl] += 1;
    g_prof.acc.ms_total[sl] += ms;
    g_prof.rec[sl] = false;
  }
  return ELSIM_OK;
}

int poll(CgWork& w, hipStream_t s) {
  ELSIM_HIP(hipMemcpyAsync(w.host, w.cg.p, sizeof(CgDev), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

// Iterations to launch before the next poll of a mixed solver: as many as
// the FP32 residual needs to reach its next threshold (the tolerance or the
// reliable-update trigger) at the per-iteration reduction rate seen so far,
// 1 .. kMaxBatch.  Fewer cost extra host round trips, more cost launches
// that return at once.
constexpr int kMaxBatch = 8;
struct BatchPlan {
  double rate = -1.0;  // per-iteration factor of r.r (unknown: < 0)
  double rr_base = -1.0;
  int it_base = 0;
  // after a poll: learn the rate from the iterations since the last one
  // (also when the batch ended in a reliable-update request: h.rr is then
  // the FP32 residual of the iteration that made it, h.iters counts it)
  // (the first poll measures from the start: r.r in rmax2, iteration 0)
  void observe(const CgDev& h) {
    const double base = rr_base > 0.0 ? rr_base : h.rmax2;
    const int it0 = rr_base > 0.0 ? it_base : 0;
    if (base > 0.0 && h.iters > it0 && h.rr > 0.0) rate = std::pow(h.rr / base, 1.0 / (h.iters - it0));
  }
  // the next batch from the state after the poll (and a reliable update,
  // which leaves the true residual in rmax2)
  int next(const CgDev& h, bool restarted, int dflt) {
    const double rr = restarted ? h.rmax2 : h.rr;
    rr_base = rr;
    it_base = h.iters;
    if (!(rate > 0.0) || !(rate < 0.999) || !(rr > 0.0)) return dflt;
    const double thr = std::max(h.tol2, h.delta2 * h.rmax2);
    if (!(rr > thr)) return 1;
    const double need = std::ceil(std::log(thr / rr) / std::log(rate));
    return (int)std::min<double>(kMaxBatch, std::max(1.0, need));
  }
};

template <int OP>
int solve_op(CgWork& w, const SolveArgs& a, int mode, SolveResult* out) {
  const Dims& d = a.d;
  double* P = w.partials ? w.partials : workspace().partials;
  hipStream_t s = stream();
  const dim3 gc = col_grid(d), blk(TX, TY, 1);
  const int nb64 = stencil_ntiles<double>(d), nb32 = stencil_ntiles<float>(d);
  const int nbo64 = ntiles64(d);  // the FP64 operator's walk
  const int64_t G = d.G();
  const int fb = flat_blocks(G);
  const int64_t G2 = (G + 1) / 2;  // element pairs of the FP32 pair kernels
  const int fbu = (int)((G2 + (int64_t)FLAT_T * UPD_P - 1) / ((int64_t)FLAT_T * UPD_P));
  if (nb64 > kMaxRedBlocks || nbo64 > kMaxRedBlocks || nb32 > kMaxRedBlocks || fb > kMaxRedBlocks ||
      mv4_ntiles(d) > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  CgDev* cg = (CgDev*)w.cg.p;
  CgDev h{};
  h.crt = a.crt;
  h.tw = a.tw;
  h.maxit = a.maxit;
  h.stop_rule = mode == ELSIM_MODE_FP64 ? a.stop_rule : ELSIM_STOP_RELRES;
  const double delta = mode == ELSIM_MODE_MIXED_MG ? mg_cfg().delta : a.delta;
  h.delta2 = delta * delta;
  h.rel_r0 = a.rel_r0;

  ELSIM_HIP(hipMemcpyAsync(cg, &h, sizeof(CgDev), hipMemcpyHostToDevice, s));
  double* res = (double*)w.res.p;
  double* wmx = (double*)w.wmx.p;
  double* prj = (double*)w.prj.p;
  double* x0 = (double*)w.x0.p;
  double* r0 = (double*)w.r0.p;
  double* e64 = (double*)w.e64.p;
  auto fin1 = [&](int op, int nb) { k_fin<1><<<1, FIN_T, 0, s>>>(op, nb, P, cg, nullptr, a.rtol); };
  auto fin2 = [&](int op, int nb) { k_fin<2><<<1, FIN_T, 0, s>>>(op, nb, P, cg, nullptr, a.rtol); };
  const bool fp32_inner = mode == ELSIM_MODE_MIXED || mode == ELSIM_MODE_MIXED_MG;
  bool e_pending = false;  // defer_add: the correction is left in e64 for the caller
  // the start's copy: the FP64 paths form their correction from it; the
  // mixed paths, which keep x fixed and accumulate the correction, store it
  // only on request
  const bool store_x0 = !fp32_inner || a.keep_x0;
  double* x0s = store_x0 ? x0 : nullptr;
  // b_eff's norm is not used when the rule is relative to ||r0|| (CnE's
  // increment rule): the finaliser then only sets the tolerance
  if (a.rel_r0) {
    k_resid64<OP><<<nb64, blk, 0, s>>>(d, a.x, a.kap, a.pse, a.scale, a.rhs, fp32_inner ? nullptr : res, r0,
                                       fp32_inner ? nullptr : prj, x0s, cg, P);
    fin1(FIN_RESID, nb64);
    fin1(FIN_BEFF, 0);
  } else {
    k_resid64_beff<OP><<<nb64, blk, 0, s>>>(d, a.x, a.kap, a.pse, a.scale, a.rhs, fp32_inner ? nullptr : res, r0,
                                            fp32_inner ? nullptr : prj, x0s, P);
    fin1(FIN_RESID, nb64);
    k_fin<1><<<1, FIN_T, 0, s>>>(FIN_BEFF, nb64, P + kMaxRedBlocks, cg, nullptr, a.rtol);  // slot 1
  }
  ELSIM_CHECK_LAUNCH();
  const int batch = kMaxBatch;
  // Multigrid path: a start that already meets the tolerance (phE in most
  // steps of a warm-started run) needs no hierarchy, so one poll decides
  // before the coefficient's set-up and the first V-cycle.
  bool mg_start_done = false;
  if (mode == ELSIM_MODE_MIXED_MG) {
    ELSIM_TRY(poll(w, s));
    mg_start_done = w.host->done != 0;
  }
  if (mode == ELSIM_MODE_FP64_PCG) {
    double* dinv64 = (double*)w.dinv64.p;
    k_dinv<OP, double><<<gc, blk, 0, s>>>(d, a.kap, a.pse, a.scale, dinv64);
    // The iterate is the correction e (x = x0 + e at the end), so the true
    // residual r0 - A e is limited by the rounding of e, not of x: with the
    // CnE increment rule the tolerance is relative to the step's change,
    // which is many orders below x, and iterating on x itself left a true
    // residual above it (relres 2.4e-7 at 128^3 where the recurred one met
    // 1e-8).
    k_pcg64_start<<<fb, FLAT_T, 0, s>>>(G, res, dinv64, prj, e64, cg, P);
    fin1(FIN_RZ, fb);
    int it_prev = 0;
    for (;;) {
      const bool pr = g_prof.on && (g_prof.op < 0 || g_prof.op == OP);
      for (int b = 0; b < batch; b++) {
        const bool t = pr && b == 0;
        ELSIM_TRY(timed(t, ELSIM_PROF_OP64, s, [&] {
          k_matvec<double, OP, true, true><<<nbo64, blk, 0, s>>>(d, prj, a.kap, a.pse, a.scale, wmx, nullptr, cg, 1,
                                                                 P);
        }));
        fin1(FIN_ALPHA_RZ, nbo64);
        ELSIM_TRY(timed(t, ELSIM_PROF_UPD64, s, [&] {
          k_pcg64_update<<<fb, FLAT_T, 0, s>>>(G, e64, prj, res, wmx, dinv64, cg, P);
        }));
        fin2(FIN_PCG64, fb);
        ELSIM_TRY(timed(t, ELSIM_PROF_DIR64, s, [&] { k_pcg64_dir<<<fb, FLAT_T, 0, s>>>(G, prj, res, dinv64, cg); }));
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_TRY(poll(w, s));
      // the bracketed first iteration ran if the counter moved (its direction
      // update ran if the next iteration started)
      if (pr && w.host->iters >= it_prev + 2) ELSIM_TRY(prof_collect(ELSIM_PROF_OP64, ELSIM_PROF_DIR64));
      it_prev = w.host->iters;
      if (w.host->done) break;
    }
    k_true_resid<OP, true><<<nb64, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, nullptr, nullptr, cg, P);
    fin2(FIN_TRUE_FINAL, nb64);
    k_add_e<<<fb, FLAT_T, 0, s>>>(G, a.x, e64);
  } else if (mode == ELSIM_MODE_FP64) {
    for (;;) {
      for (int b = 0; b < batch; b++) {
        k_matvec<double, OP, true, true><<<nbo64, blk, 0, s>>>(d, prj, a.kap, a.pse, a.scale, wmx, nullptr, cg, 0, P);
        fin1(FIN_ALPHA_RR, nbo64);
        k_cg64_update<OP><<<fb, FLAT_T, 0, s>>>(G, a.x, prj, res, wmx, a.w8, cg, P);
        fin2(FIN_CG64, fb);
        k_cg64_dir<<<fb, FLAT_T, 0, s>>>(G, prj, res, cg);
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_TRY(poll(w, s));
      if (w.host->done) break;
    }
    // true residual of the correction system for the report
    k_diff<<<gc, blk, 0, s>>>(d, a.x, x0, e64);
    k_true_resid<OP, true><<<nb64, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, nullptr, nullptr, cg, P);
    fin2(FIN_TRUE_FINAL, nb64);
  } else if (mode == ELSIM_MODE_MIXED_MG && mg_start_done) {
    // nothing to correct: x stays as it came in
  } else if (mode == ELSIM_MODE_MIXED_MG) {
    // FP32 CG preconditioned by one multigrid V-cycle, FP64 reliable updates
    // as in the Jacobi-preconditioned mixed path below.  Per iteration: the
    // operator (p.Ap) -> alpha; r -= alpha w with the pre-smoothed iterate
    // omega dinv r; the V-cycle z = M^-1 r (r.r and r.z partials) -> beta or a
    // reliable-update request; y += alpha p, p = z + beta p.  The operator's
    // FP32 coefficient and the hierarchy live in the operator's multigrid
    // work space and are reused while coef_id names the same coefficient.
    if (!w.mg[OP]) w.mg[OP] = new MgWork();
    MgWork& m = *w.mg[OP];
    float *r32 = (float*)w.r32.p, *w32 = (float*)w.w32.p, *y32 = (float*)w.y32.p, *p32 = (float*)w.p32.p;
    ELSIM_TRY(mg_setup(m, OP, d, a.kap, a.pse, a.scale, a.coef_id));
    const float* k32 = m.kap32;
    const float* s32 = m.pse32;
    const float sc32 = (float)a.scale;
    const int nbv = mg_partial_blocks(d);
    const bool mv4 = mv4_ok(d), mv4_edge = mv4_tiles(d).gx > 1;
    const int nbm = mv4 ? mv4_ntiles(d) : nb32;
    const int fbd = (int)((G2 + (int64_t)FLAT_T * DIR_P - 1) / ((int64_t)FLAT_T * DIR_P));
    // Fused form (n1 a multiple of 4): the residual update runs inside the
    // V-cycle's down-sweep, which writes the new residual to the other of two
    // buffers (r32, and the work space's pre-smoothing buffer, unused then);
    // par names the current one.  A restart writes r32 and resets par.  An
    // iteration that the device skips (solve done, or a reliable update
    // due) leaves par stale only until that exit or restart.
    const bool fused = mg_fused_ok(d);
    float* rb[2] = {r32, m.x1};
    int par = 0;
    auto precondition = [&]() {
      par = 0;
      if (fused) return mg_vcycle_fused(m, OP, d, r32, r32, false, nullptr, w32, cg, P, false);
      ELSIM_TRY(mg_presmooth(m, d, r32, cg));
      return mg_vcycle(m, OP, d, r32, w32, cg, P, false);
    };
    k_mg_start<<<fbu, FLAT_T, 0, s>>>(G2, (const double2*)r0, (float2*)r32, (float2*)p32, (float2*)y32,
                                      (double2*)e64);
    ELSIM_TRY(precondition());
    fin2(FIN_MG_START, nbv);
    k_mg_dir3<<<fbd, FLAT_T, 0, s>>>(G2, (float2*)y32, (float2*)p32, (const float2*)w32, cg);
    ELSIM_CHECK_LAUNCH();
    int it_prev = 0;
    BatchPlan plan;
    int nrun = 2;  // a few iterations before the first poll: warm-started solves often need only these
    for (;;) {
      const bool pr = g_prof.on && (g_prof.op < 0 || g_prof.op == OP);
      for (int b = 0; b < nrun; b++) {
        const bool t = pr && b == 0;
        ELSIM_TRY(timed(t, ELSIM_PROF_OP32, s, [&] {
          if (mv4 && mv4_edge)
            k_matvec4<OP, true, true><<<nbm, 256, 0, s>>>(d, p32, k32, s32, sc32, w32, cg, P);
          else if (mv4)
            k_matvec4<OP, true, false><<<nbm, 256, 0, s>>>(d, p32, k32, s32, sc32, w32, cg, P);
          else
            k_matvec<float, OP, true, true><<<nb32, blk, 0, s>>>(d, p32, k32, s32, sc32, w32, nullptr, cg, 1, P);
        }));
        fin1(FIN_ALPHA_RZ, nbm);
        int rc = ELSIM_OK;
        if (fused) {
          ELSIM_TRY(timed(t, ELSIM_PROF_MG, s, [&] {
            rc = mg_vcycle_fused(m, OP, d, rb[par], w32, true, rb[par ^ 1], w32, cg, P, t);
          }));
          par ^= 1;
        } else {
          ELSIM_TRY(timed(t, ELSIM_PROF_UPD32, s, [&] { rc = mg_update(m, d, r32, w32, cg); }));
          ELSIM_TRY(rc);
          ELSIM_TRY(timed(t, ELSIM_PROF_MG, s, [&] { rc = mg_vcycle(m, OP, d, r32, w32, cg, P, t); }));
        }
        ELSIM_TRY(rc);
        fin2(FIN_MG_UPDATE, nbv);
        ELSIM_TRY(timed(t, ELSIM_PROF_DIR32, s, [&] {
          k_mg_dir3<<<fbd, FLAT_T, 0, s>>>(G2, (float2*)y32, (float2*)p32, (const float2*)w32, cg);
        }));
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_TRY(poll(w, s));
      if (pr && w.host->iters >= it_prev + 2) {
        ELSIM_TRY(prof_collect(ELSIM_PROF_OP32, ELSIM_PROF_DIR32));
        ELSIM_TRY(prof_collect(ELSIM_PROF_MG, ELSIM_PROF_MG_UP));
      }
      it_prev = w.host->iters;
      if (w.host->done) break;
      plan.observe(*w.host);
      const bool restart = w.host->need_reliable != 0;
      if (restart) {
        // fold y into the FP64 correction, FP64 true residual, restart from
        // it with z = M^-1 r
        k_mx_fold<<<fb, FLAT_T, 0, s>>>(G, e64, y32, cg);
        k_true_resid<OP, false><<<nb64, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, r32, nullptr, cg, P);
        fin2(FIN_TRUE_RELIABLE, nb64);
        ELSIM_TRY(precondition());
        fin2(FIN_MG_START, nbv);
        k_mg_dir3<<<fbd, FLAT_T, 0, s>>>(G2, (float2*)y32, (float2*)p32, (const float2*)w32, cg);
        ELSIM_CHECK_LAUNCH();
        ELSIM_TRY(poll(w, s));
        if (w.host->done) break;
      }
      nrun = plan.next(*w.host, restart, 2);
    }
    if (a.defer_add)
      e_pending = true;
    else
      k_add_e<<<fb, FLAT_T, 0, s>>>(G, a.x, e64);
  } else {
    uint16_t* dinv = (uint16_t*)w.dinv32.p;  // BF16
    float *r32 = (float*)w.r32.p, *w32 = (float*)w.w32.p, *y32 = (float*)w.y32.p, *p32 = (float*)w.p32.p;
    // FP32 copies of the coefficients: interior in the Jacobi set-up, which
    // reads them anyway, ghosts by a face kernel
    const int64_t ng = ghost_cells(d);
    if (ng >= ((int64_t)1 << 31)) {
      set_error("boundary faces of %d x %d x %d exceed 2^31 cells", d.n1, d.n2, d.n3);
      return ELSIM_ERR_ARG;
    }
    const unsigned gb = (unsigned)((ng + 255) / 256);
    k_ghost_f32<<<gb, 256, 0, s>>>(d, a.kap, (float*)w.kap32.p);
    if (OP == ELSIM_OP_CNE) k_ghost_f32<<<gb, 256, 0, s>>>(d, a.pse, (float*)w.pse32.p);
    k_dinv<OP, uint16_t><<<gc, blk, 0, s>>>(d, a.kap, a.pse, a.scale, dinv, (float*)w.kap32.p,
                                            OP == ELSIM_OP_CNE ? (float*)w.pse32.p : nullptr);
    k_mx_start<<<fbu, FLAT_T, 0, s>>>(G2, (const double2*)r0, (const uint32_t*)dinv, (float2*)r32, (float2*)p32,
                                      (float2*)y32, (double2*)e64, P);
    fin2(FIN_MX_START, fbu);
    ELSIM_CHECK_LAUNCH();
    const float sc32 = (float)a.scale;
    const bool mv4 = mv4_ok(d), mv4_edge = mv4_tiles(d).gx > 1;
    const int nbm = mv4 ? mv4_ntiles(d) : nb32;
    const int fbd = (int)((G2 + (int64_t)FLAT_T * DIR_P - 1) / ((int64_t)FLAT_T * DIR_P));
    int it_prev = 0;
    BatchPlan plan;
    // two iterations before the first poll, then batches sized from the
    // observed rate: the CnE solve converges a segment in about two
    // iterations, and a batch sized past that runs only skipped launches
    int nrun = 2;
    for (;;) {
      const bool pr = g_prof.on && (g_prof.op < 0 || g_prof.op == OP);
      for (int b = 0; b < nrun; b++) {
        const bool t = pr && b == 0;
        ELSIM_TRY(timed(t, ELSIM_PROF_OP32, s, [&] {
          if (mv4 && mv4_edge)
            k_matvec4<OP, true, true><<<nbm, 256, 0, s>>>(d, p32, (const float*)w.kap32.p, (const float*)w.pse32.p,
                                                          sc32, w32, cg, P);
          else if (mv4)
            k_matvec4<OP, true, false><<<nbm, 256, 0, s>>>(d, p32, (const float*)w.kap32.p, (const float*)w.pse32.p,
                                                           sc32, w32, cg, P);
          else
            k_matvec<float, OP, true, true><<<nb32, blk, 0, s>>>(d, p32, (const float*)w.kap32.p,
                                                                 (const float*)w.pse32.p, sc32, w32, nullptr, cg, 1, P);
        }));
        fin1(FIN_ALPHA_RZ, nbm);
        ELSIM_TRY(timed(t, ELSIM_PROF_UPD32, s, [&] {
          k_mx_upd3<<<fbu, FLAT_T, 0, s>>>(G2, (float2*)r32, (const float2*)w32, (const uint32_t*)dinv, cg, P);
        }));
        fin2(FIN_MX_UPDATE, fbu);
        ELSIM_TRY(timed(t, ELSIM_PROF_DIR32, s, [&] {
          k_mx_dir3<<<fbd, FLAT_T, 0, s>>>(G2, (float2*)y32, (float2*)p32, (const float2*)r32,
                                           (const uint32_t*)dinv, cg);
        }));
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_TRY(poll(w, s));
      if (pr && w.host->iters >= it_prev + 2) ELSIM_TRY(prof_collect(ELSIM_PROF_OP32, ELSIM_PROF_DIR32));
      it_prev = w.host->iters;
      if (w.host->done) break;
      plan.observe(*w.host);
      const bool restart = w.host->need_reliable != 0;
      // Once the FP32 residual has dropped by delta (or meets the tolerance,
      // or the cap is reached) the device sets need_reliable and the rest of
      // the batch is skipped; the reliable update runs here, between
      // batches: fold y into the FP64 correction, FP64 true residual, and a
      // restart of the FP32 recurrence from it.
      if (restart) {
        k_mx_fold<<<fb, FLAT_T, 0, s>>>(G, e64, y32, cg);
        k_true_resid<OP, false><<<nb64, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, r32, dinv, cg, P);
        fin2(FIN_TRUE_RELIABLE, nb64);
        k_mx_dir3<<<fbd, FLAT_T, 0, s>>>(G2, (float2*)y32, (float2*)p32, (const float2*)r32, (const uint32_t*)dinv,
                                         cg);
        ELSIM_CHECK_LAUNCH();
        ELSIM_TRY(poll(w, s));
        if (w.host->done) break;
      }
      nrun = plan.next(*w.host, restart, batch);
    }
    if (a.defer_add)
      e_pending = true;
    else
      k_add_e<<<fb, FLAT_T, 0, s>>>(G, a.x, e64);
  }
  ELSIM_CHECK_LAUNCH();
  ELSIM_TRY(poll(w, s));
  const CgDev& r = *w.host;
  out->e_pending = e_pending;
  out->x0_stored = store_x0;
  out->iters = r.iters;
  out->n_reliable = r.n_reliable;
  out->relres = r.bn2 > 0 ? std::sqrt(r.rr_true / r.bn2) : 0.0;
  out->res0 = std::sqrt(r.rr0);
  out->status = r.done == 1 ? ELSIM_OK : (r.done == 2 ? ELSIM_NOT_CONVERGED : ELSIM_ERR_NAN);
  if (out->status == ELSIM_ERR_NAN) set_error("CG produced a non-finite value (op %d)", OP);
  return ELSIM_OK;
}

}  // namespace

int prof_event(int slot, int end) {
  if (g_prof.on) {
    ELSIM_HIP(hipEventRecord(g_prof.ev[2 * slot + (end ? 1 : 0)], stream()));
    if (end) g_prof.rec[slot] = true;
  }
  return ELSIM_OK;
}

int launch_resid_norm2(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                       const double* rhs, double* out) {
  const int nb = stencil_ntiles<double>(d);
  if (nb > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  double* P = workspace().partials;
  hipStream_t s = stream();
  switch (op) {
    case ELSIM_OP_CNE: k_resid_norm<ELSIM_OP_CNE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, P); break;
    case ELSIM_OP_PHP: k_resid_norm<ELSIM_OP_PHP><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, P); break;
    case ELSIM_OP_PHE: k_resid_norm<ELSIM_OP_PHE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, P); break;
    default: set_error("unknown operator %d", op); return ELSIM_ERR_ARG;
  }
  ELSIM_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ELSIM_HIP(hipMemcpyAsync(h.data(), P, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : h) t += v;
  *out = t;
  return ELSIM_OK;
}

int launch_resid_store(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                       const double* rhs, double* r, double* out) {
  const int nb = stencil_ntiles<double>(d);
  if (nb > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  double* P = workspace().partials;
  hipStream_t s = stream();
  switch (op) {
    case ELSIM_OP_CNE: k_resid_store<ELSIM_OP_CNE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, r, P); break;
    case ELSIM_OP_PHP: k_resid_store<ELSIM_OP_PHP><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, r, P); break;
    case ELSIM_OP_PHE: k_resid_store<ELSIM_OP_PHE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, r, P); break;
    default: set_error("unknown operator %d", op); return ELSIM_ERR_ARG;
  }
  ELSIM_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ELSIM_HIP(hipMemcpyAsync(h.data(), P, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : h) t += v;
  *out = t;
  return ELSIM_OK;
}

int launch_dot_resid(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                     const double* rhs, const double* dv, double* out) {
  const int nb = stencil_ntiles<double>(d);
  if (nb > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  double* P = workspace().partials;
  hipStream_t s = stream();
  switch (op) {
    case ELSIM_OP_CNE: k_dot_resid<ELSIM_OP_CNE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, dv, P); break;
    case ELSIM_OP_PHP: k_dot_resid<ELSIM_OP_PHP><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, dv, P); break;
    case ELSIM_OP_PHE: k_dot_resid<ELSIM_OP_PHE><<<nb, TX, 0, s>>>(d, x, kap, pse, scale, rhs, dv, P); break;
    default: set_error("unknown operator %d", op); return ELSIM_ERR_ARG;
  }
  ELSIM_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ELSIM_HIP(hipMemcpyAsync(h.data(), P, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : h) t += v;
  *out = t;
  return ELSIM_OK;
}

int launch_galerkin_dots(int op, const Dims& d, const double* x, const double* kap, double scale, const double* rhs,
                         const double* dv, double* num, double* den) {
  const int nb = stencil_ntiles<double>(d);
  if (nb >= kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  double* P = workspace().partials;
  hipStream_t s = stream();
  switch (op) {
    case ELSIM_OP_PHP: k_galerkin_dots<ELSIM_OP_PHP><<<nb, TX, 0, s>>>(d, x, kap, scale, rhs, dv, P); break;
    case ELSIM_OP_PHE: k_galerkin_dots<ELSIM_OP_PHE><<<nb, TX, 0, s>>>(d, x, kap, scale, rhs, dv, P); break;
    default: set_error("Galerkin step: operator %d (phP or phE)", op); return ELSIM_ERR_ARG;
  }
  // den summed as the operator's dot product is (one-block finaliser, into
  // the unused end of slot 0), num on the host in block order, as before
  double* dden = P + kMaxRedBlocks - 1;
  k_fin<1><<<1, FIN_T, 0, s>>>(FIN_DOT, nb, P + kMaxRedBlocks, nullptr, dden, 0.0);
  ELSIM_CHECK_LAUNCH();
  std::vector<double> h(nb);
  ELSIM_HIP(hipMemcpyAsync(h.data(), P, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipMemcpyAsync(den, dden, sizeof(double), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : h) t += v;
  *num = t;
  return ELSIM_OK;
}

int cg_solve(CgWork& w, const SolveArgs& a, int mode, SolveResult* res) {
  if (!w.bound) ELSIM_TRY(w.alloc(a.d));
  if (w.bound && (w.d.n1 != a.d.n1 || w.d.n2 != a.d.n2 || w.d.n3 != a.d.n3)) {
    set_error("work space laid out for another grid");
    return ELSIM_ERR_ARG;
  }
  if (mode != ELSIM_MODE_FP64 && a.stop_rule != ELSIM_STOP_RELRES) {
    set_error("the preconditioned solvers use the relative-residual stop rule only");
    return ELSIM_ERR_ARG;
  }
  // CnE's reaction term keeps it well conditioned: Jacobi-PCG needs a few
  // iterations there, fewer than a V-cycle costs
  // (with the multigrid path's reliable-update threshold: CnE converges in a
  // few iterations, each reducing its residual by more than half)
  if (mode == ELSIM_MODE_MIXED_MG && a.op == ELSIM_OP_CNE) {
    SolveArgs b = a;
    b.delta = mg_cfg().delta;
    return solve_op<ELSIM_OP_CNE>(w, b, ELSIM_MODE_MIXED, res);
  }
  if (mode == ELSIM_MODE_MIXED_MG && w.bound) {
    set_error("the multigrid solver needs library-allocated work space");
    return ELSIM_ERR_ARG;
  }
  switch (a.op) {
    case ELSIM_OP_CNE: return solve_op<ELSIM_OP_CNE>(w, a, mode, res);
    case ELSIM_OP_PHP: return solve_op<ELSIM_OP_PHP>(w, a, mode, res);
    case ELSIM_OP_PHE: return solve_op<ELSIM_OP_PHE>(w, a, mode, res);
    case ELSIM_OP_POT_DD: return solve_op<ELSIM_OP_POT_DD>(w, a, mode, res);
  }
  set_error("unknown operator %d", a.op);
  return ELSIM_ERR_ARG;
}

template <typename T>
static int launch_mv(int op, const Dims& d, const T* p, const T* kap, const T* pse, T scale, T* out, double* dot,
                     bool hom) {
  ELSIM_TRY(ensure_init());
  double* P = workspace().partials;
  const int grid = std::is_same<T, double>::value ? ntiles64(d) : stencil_ntiles<T>(d);
  const dim3 blk(TX, TY, 1);
  if (grid > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  hipStream_t s = stream();
  if constexpr (std::is_same<T, float>::value) {
   if (hom && mv4_ok(d)) {
    // the operator of the FP32 CG loop
    const int g4 = mv4_ntiles(d);
    if (g4 > kMaxRedBlocks) {
      set_error("grid too large for the reduction workspace");
      return ELSIM_ERR_ARG;
    }
    const float *pf = p, *kf = kap, *sf = pse;
    float* of = out;
    const float scf = scale;
    double* Pd = dot ? P : nullptr;
    const bool edge = mv4_tiles(d).gx > 1;
#define MV4(OPV)                                                                               \
  do {                                                                                         \
    if (edge)                                                                                  \
      k_matvec4<OPV, false, true><<<g4, 256, 0, s>>>(d, pf, kf, sf, scf, of, nullptr, Pd);     \
    else                                                                                       \
      k_matvec4<OPV, false, false><<<g4, 256, 0, s>>>(d, pf, kf, sf, scf, of, nullptr, Pd);    \
  } while (0)
    switch (op) {
      case ELSIM_OP_CNE: MV4(ELSIM_OP_CNE); break;
      case ELSIM_OP_PHP: MV4(ELSIM_OP_PHP); break;
      case ELSIM_OP_PHE: MV4(ELSIM_OP_PHE); break;
      case ELSIM_OP_POT_DD: MV4(ELSIM_OP_POT_DD); break;
      default: set_error("unknown operator %d", op); return ELSIM_ERR_ARG;
    }
#undef MV4
    if (dot) k_fin<1><<<1, FIN_T, 0, s>>>(FIN_DOT, g4, P, nullptr, dot, 0.0);
    ELSIM_CHECK_LAUNCH();
    return ELSIM_OK;
   }
  }
#define MV(OPV, H) \
  k_matvec<T, OPV, H, false><<<grid, blk, 0, s>>>(d, p, kap, pse, scale, out, dot, nullptr, 0, P)
  if (op == ELSIM_OP_CNE) {
    if (hom) MV(ELSIM_OP_CNE, true); else MV(ELSIM_OP_CNE, false);
  } else if (op == ELSIM_OP_PHP) {
    if (hom) MV(ELSIM_OP_PHP, true); else MV(ELSIM_OP_PHP, false);
  } else if (op == ELSIM_OP_PHE) {
    if (hom) MV(ELSIM_OP_PHE, true); else MV(ELSIM_OP_PHE, false);
  } else if (op == ELSIM_OP_POT_DD) {
    if (hom) MV(ELSIM_OP_POT_DD, true); else MV(ELSIM_OP_POT_DD, false);
  } else {
    set_error("unknown operator %d", op);
    return ELSIM_ERR_ARG;
  }
#undef MV
  if (dot) k_fin<1><<<1, FIN_T, 0, s>>>(FIN_DOT, blocks_of(grid), P, nullptr, dot, 0.0);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int launch_matvec_f64(int op, const Dims& d, const double* p, const double* kap, const double* pse, double scale,
                      double* out, double* dot, bool hom) {
  return launch_mv<double>(op, d, p, kap, pse, scale, out, dot, hom);
}
int launch_matvec_f32(int op, const Dims& d, const float* p, const float* kap, const float* pse, float scale,
                      float* out, double* dot, bool hom) {
  return launch_mv<float>(op, d, p, kap, pse, scale, out, dot, hom);
}

int launch_copy_interior_g2n(const Dims& d, const double* g, double* n) {
  k_copy_g2n<<<col_grid(d), dim3(TX, TY, 1), 0, stream()>>>(d, g, n);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

}  // namespace elsim

using namespace elsim;

extern "C" {

int elsim_dev_matvec_f64(int op, int n1, int n2, int n3, const double* p, const double* kap, const double* pse,
                         double scale, double* out, double* dot) {
  if (!dims_ok(n1, n2, n3) || !p || !kap || !out || (op == ELSIM_OP_CNE && !pse)) return ELSIM_ERR_ARG;
  return launch_matvec_f64(op, Dims{n1, n2, n3}, p, kap, pse, scale, out, dot, true);
}

int elsim_dev_matvec_f32(int op, int n1, int n2, int n3, const float* p, const float* kap, const float* pse,
                         float scale, float* out, double* dot) {
  if (!dims_ok(n1, n2, n3) || !p || !kap || !out || (op == ELSIM_OP_CNE && !pse)) return ELSIM_ERR_ARG;
  return launch_matvec_f32(op, Dims{n1, n2, n3}, p, kap, pse, scale, out, dot, true);
}

int elsim_dev_mg_apply(int op, int n1, int n2, int n3, const double* kap, const double* pse, double scale,
                       const float* r, float* z) {
  if (!dims_ok(n1, n2, n3) || !kap || !r || !z || (op == ELSIM_OP_CNE && !pse) || op < ELSIM_OP_CNE ||
      op > ELSIM_OP_POT_DD) {
    set_error("elsim_dev_mg_apply: bad arguments");
    return ELSIM_ERR_ARG;
  }
  ELSIM_TRY(ensure_init());
  static MgWork mg;
  const Dims d{n1, n2, n3};
  hipStream_t s = stream();
  ELSIM_HIP(hipMemsetAsync(z, 0, (size_t)d.G() * 4, s));
  ELSIM_TRY(mg_setup(mg, op, d, kap, pse, scale, 0));
  ELSIM_TRY(mg_presmooth(mg, d, r, nullptr));
  ELSIM_TRY(mg_vcycle(mg, op, d, r, z, nullptr, workspace().partials, false));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_prof_enable(int on) {
  ELSIM_TRY(ensure_init());
  if (on && !g_prof.ready) {
    for (hipEvent_t& e : g_prof.ev) ELSIM_HIP(hipEventCreate(&e));
    g_prof.ready = true;
  }
  if (on < 0 || on > 2 + ELSIM_OP_POT_DD) return ELSIM_ERR_ARG;
  g_prof.on = on != 0;
  g_prof.op = on >= 2 ? on - 2 : -1;
  return ELSIM_OK;
}

int elsim_prof_read(elsim_kernel_prof* out) {
  if (!out) return ELSIM_ERR_ARG;
  *out = g_prof.acc;
  g_prof.acc = elsim_kernel_prof{};
  return ELSIM_OK;
}

int elsim_dev_matvec_stored_f64(int op, int n1, int n2, int n3, const double* p, const double* kap,
                                const double* pse, double scale, double* out) {
  if (!dims_ok(n1, n2, n3) || !p || !kap || !out || (op == ELSIM_OP_CNE && !pse)) return ELSIM_ERR_ARG;
  return launch_matvec_f64(op, Dims{n1, n2, n3}, p, kap, pse, scale, out, nullptr, false);
}

}  // extern "C"
