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
// The fine level is the CG's own FP32 operator (FP32 coefficient and the
// BF16 Jacobi inverse), swept with the register-tiled stencil walk, with the
// pre-smoothed iterate formed on the fly from the residual.  The operators
// are negative definite for the potentials (A = -A_hat) and positive for
// CnE; the coarse levels hold A_hat = sigma A, sigma = -1 / +1, and the
// V-cycle's input is sigma r, so that its output approximates A^-1 r.
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

// Flat index q of level L -> (i, j, k).
struct Ijk {
  int i, j, k;
};
__device__ __forceinline__ Ijk unflat(const MgLevel& L, int64_t q) {
  const int64_t r = q / L.m1;
  return Ijk{(int)(q - r * L.m1), (int)(r % L.m2), (int)(r / L.m2)};
}

// A_hat x at cell q = (i, j, k) of level L, xc = x at the cell, xn(q') = x at
// a neighbour.
template <class X>
__device__ __forceinline__ float apply_lvl(const MgLevel& L, const Ijk& c, int64_t q, float xc, X&& xn) {
  const int64_t sy = L.m1, sz = (int64_t)L.m1 * L.m2;
  float s = L.dg[q] * xc;
  if (c.i + 1 < L.m1) s -= L.cx[q] * xn(q + 1);
  if (c.i > 0) s -= L.cx[q - 1] * xn(q - 1);
  if (c.j + 1 < L.m2) s -= L.cy[q] * xn(q + sy);
  if (c.j > 0) s -= L.cy[q - sy] * xn(q - sy);
  if (c.k + 1 < L.m3) s -= L.cz[q] * xn(q + sz);
  if (c.k > 0) s -= L.cz[q - sz] * xn(q - sz);
  return s;
}

// ----------------------------------------------------------- hierarchy --

// Level 1 from the fine operator: thread per level-1 cell.  Fine face
// conductance between ghosted cells a and b: scale * 0.5 (kap_a + kap_b), the
// face coefficient of the CG operator.  Mass: psE (CnE) and, for a face
// whose homogeneous ghost rule is zero (Dirichlet data), that face's
// conductance.
template <int OP>
__global__ __launch_bounds__(256) void k_mg_build1(Dims d, const double* __restrict__ kap,
                                                   const double* __restrict__ pse, double scale, MgLevel c) {
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
  c.cx[q] = (float)cxp;
  c.cy[q] = (float)cyp;
  c.cz[q] = (float)czp;
  c.ms[q] = (float)ms;
  c.dg[q] = (float)(ms + ((cxp + cxm) + (cyp + cym) + (czp + czm)));
}

// Level l+1 from level l (f): thread per coarse cell.
__global__ __launch_bounds__(256) void k_mg_build(MgLevel f, MgLevel c) {
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
  c.cx[q] = (float)cxp;
  c.cy[q] = (float)cyp;
  c.cz[q] = (float)czp;
  c.ms[q] = (float)ms;
  c.dg[q] = (float)(ms + ((cxp + cxm) + (cyp + cym) + (czp + czm)));
}

// Dense inverse of the coarsest operator (n <= kMgCoarseMax cells):
// in-place Gauss-Jordan elimination in FP64 in LDS, one block.  The operator
// is symmetric positive definite (every level keeps at least one Dirichlet
// or reaction term), so no pivoting is needed.
__global__ __launch_bounds__(256) void k_mg_coarse_inv(MgLevel c, float* __restrict__ ainv) {
  __shared__ double a[kMgCoarseMax][kMgCoarseMax];
  __shared__ double fac[kMgCoarseMax];
  __shared__ double piv;
  const int n = (int)lvl_n(c), tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) a[e / n][e % n] = 0.0;
  __syncthreads();
  for (int q = tid; q < n; q += nt) {
    const Ijk C = unflat(c, q);
    const int sy = c.m1, sz = c.m1 * c.m2;
    a[q][q] = c.dg[q];
    if (C.i + 1 < c.m1) a[q][q + 1] = -(double)c.cx[q];
    if (C.i > 0) a[q][q - 1] = -(double)c.cx[q - 1];
    if (C.j + 1 < c.m2) a[q][q + sy] = -(double)c.cy[q];
    if (C.j > 0) a[q][q - sy] = -(double)c.cy[q - sy];
    if (C.k + 1 < c.m3) a[q][q + sz] = -(double)c.cz[q];
    if (C.k > 0) a[q][q - sz] = -(double)c.cz[q - sz];
  }
  __syncthreads();
  for (int p = 0; p < n; p++) {
    if (tid == 0) piv = a[p][p];
    for (int r = tid; r < n; r += nt) fac[r] = a[r][p];
    __syncthreads();
    if (tid == 0) a[p][p] = 1.0;
    __syncthreads();
    for (int j = tid; j < n; j += nt) a[p][j] /= piv;
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) {
      const int r = e / n, j = e % n;
      if (r == p) continue;
      a[r][j] = (j == p ? 0.0 : a[r][j]) - fac[r] * a[p][j];
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += nt) ainv[e] = (float)a[e / n][e % n];
}

// ------------------------------------------------------------- V-cycle --

// Fine level, down: x1 = omega dinv r (on the fly), t = sigma (r - A x1)
// over the interior; partials of r.r (slot 0).
template <int OP>
__global__ __launch_bounds__(256) void k_mg_down0(Dims d, const float* __restrict__ r, const uint16_t* __restrict__ dinv,
                                                  const float* __restrict__ kap, const float* __restrict__ pse,
                                                  float scale, float omega, float* __restrict__ t, const CgDev* cg,
                                                  double* partials) {
  if (mg_skip(cg)) return;
  constexpr float sg = Rule<OP>::diag ? 1.0f : -1.0f;
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<float>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep_ld<float, OP, true>(
        d, stencil_idx<Rows<float>::SJ>(d, tile),
        [&](int64_t q) { return omega * bf16_to_f32(dinv[q]) * r[q]; }, kap, pse, scale,
        [&](int64_t c, int64_t, float, float v) {
          const float rc = r[c];
          t[c] = sg * (rc - v);
          acc[0] += (double)rc * (double)rc;
        });
  block_partials<1>(acc, partials);
}

// Restriction of the fine residual t to level 1 (sum over the children).
__global__ __launch_bounds__(256) void k_mg_restrict0(Dims d, const float* __restrict__ t, MgLevel c,
                                                      const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= lvl_n(c)) return;
  const Ijk C = unflat(c, q);
  const int i0 = 2 * C.i + 1, i1 = min(2 * C.i + 2, d.n1);
  const int j0 = 2 * C.j + 1, j1 = min(2 * C.j + 2, d.n2);
  const int k0 = 2 * C.k + 1, k1 = min(2 * C.k + 2, d.n3);
  float s = 0.0f;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) s += t[d.gi(i, j, k)];
  c.b[q] = s;
}

// Coarse level l (f), down: c.b = restriction of f.b - A_hat (S f.b), thread
// per level-(l+1) cell, the pre-smoothed iterate formed on the fly.
__global__ __launch_bounds__(256) void k_mg_down(MgLevel f, MgLevel c, float omega, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= lvl_n(c)) return;
  const Ijk C = unflat(c, q);
  const int i1 = min(2 * C.i + 1, f.m1 - 1), j1 = min(2 * C.j + 1, f.m2 - 1), k1 = min(2 * C.k + 1, f.m3 - 1);
  auto x1 = [&](int64_t p) { return omega * f.b[p] / f.dg[p]; };
  float s = 0.0f;
  for (int k = 2 * C.k; k <= k1; k++)
    for (int j = 2 * C.j; j <= j1; j++)
      for (int i = 2 * C.i; i <= i1; i++) {
        const int64_t p = lidx(f, i, j, k);
        s += f.b[p] - apply_lvl(f, Ijk{i, j, k}, p, x1(p), x1);
      }
  c.b[q] = s;
}

// Coarsest level: x = A_hat^-1 b (one block).
__global__ __launch_bounds__(64) void k_mg_solve(MgLevel c, const float* __restrict__ ainv, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int n = (int)lvl_n(c), i = threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int j = 0; j < n; j++) s += ainv[i * n + j] * c.b[j];
  c.x[i] = s;
}

// Coarse level l (f), up: x2 = S b + oc P x_c (on the fly), then
// x = x2 + S (b - A_hat x2); thread per level-l cell.
__global__ __launch_bounds__(256) void k_mg_up(MgLevel f, MgLevel c, float omega, float oc, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= lvl_n(f)) return;
  const Ijk F = unflat(f, p);
  auto x2 = [&](int64_t q) {
    const Ijk Q = unflat(f, q);
    return omega * f.b[q] / f.dg[q] + oc * c.x[lidx(c, Q.i >> 1, Q.j >> 1, Q.k >> 1)];
  };
  const float xc = omega * f.b[p] / f.dg[p] + oc * c.x[lidx(c, F.i >> 1, F.j >> 1, F.k >> 1)];
  const float ax = apply_lvl(f, F, p, xc, x2);
  f.x[p] = xc + omega * (f.b[p] - ax) / f.dg[p];
}

// Fine level, up, part 1: t = omega dinv r + oc P x_1 over the interior
// (ghosts of t stay zero).
__global__ __launch_bounds__(256) void k_mg_prolong0(Dims d, const float* __restrict__ r,
                                                     const uint16_t* __restrict__ dinv, MgLevel c, float omega,
                                                     float oc, float* __restrict__ t, const CgDev* cg) {
  if (mg_skip(cg)) return;
  const ColIdx ci = col_idx(d);
  if (!ci.active) return;
  const int64_t base = lidx(c, (ci.i - 1) >> 1, (ci.j - 1) >> 1, 0), sz = (int64_t)c.m1 * c.m2;
  for_col(d, ci, [&](int64_t q, int64_t, int k) {
    t[q] = omega * bf16_to_f32(dinv[q]) * r[q] + oc * c.x[base + sz * ((k - 1) >> 1)];
  });
}

// Fine level, up, part 2: z = t + omega dinv (r - A t) over the interior;
// partials of r.z (slot 1).
template <int OP>
__global__ __launch_bounds__(256) void k_mg_post0(Dims d, const float* __restrict__ t, const float* __restrict__ r,
                                                  const uint16_t* __restrict__ dinv, const float* __restrict__ kap,
                                                  const float* __restrict__ pse, float scale, float omega,
                                                  float* __restrict__ z, const CgDev* cg, double* partials) {
  if (mg_skip(cg)) return;
  double acc[1] = {0.0};
  const int nt = stencil_ntiles<float>(d);
  for (int tile = blockIdx.x; tile < nt; tile += gridDim.x)
    sweep<float, OP, true>(d, stencil_idx<Rows<float>::SJ>(d, tile), t, kap, pse, scale,
                           [&](int64_t c, int64_t, float x, float v) {
                             const float rc = r[c];
                             const float zc = x + omega * bf16_to_f32(dinv[c]) * (rc - v);
                             z[c] = zc;
                             acc[0] += (double)rc * (double)zc;
                           });
  block_sum<1>(acc);
  if (linear_tid() == 0) partials[kMaxRedBlocks + linear_bid()] = acc[0];
}

inline int blocks256(int64_t n) { return (int)((n + 255) / 256); }

template <int OP>
int setup_t(MgWork& m, const Dims& d, const double* kap, const double* pse, double scale) {
  hipStream_t s = stream();
  k_mg_build1<OP><<<blocks256(lvl_n(m.lv[1])), 256, 0, s>>>(d, kap, pse, scale, m.lv[1]);
  for (int l = 1; l < m.nlev; l++) k_mg_build<<<blocks256(lvl_n(m.lv[l + 1])), 256, 0, s>>>(m.lv[l], m.lv[l + 1]);
  k_mg_coarse_inv<<<1, 256, 0, s>>>(m.lv[m.nlev], m.ainv);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

template <int OP>
int vcycle_t(MgWork& m, const Dims& d, const float* kap32, const float* pse32, const uint16_t* dinv, float scale,
             const float* r32, float* z, const CgDev* cg, double* partials) {
  hipStream_t s = stream();
  const MgCfg& g = mg_cfg();
  const int nt = stencil_ntiles<float>(d);
  k_mg_down0<OP><<<nt, 256, 0, s>>>(d, r32, dinv, kap32, pse32, scale, g.omega, m.t, cg, partials);
  k_mg_restrict0<<<blocks256(lvl_n(m.lv[1])), 256, 0, s>>>(d, m.t, m.lv[1], cg);
  for (int l = 1; l < m.nlev; l++)
    k_mg_down<<<blocks256(lvl_n(m.lv[l + 1])), 256, 0, s>>>(m.lv[l], m.lv[l + 1], g.omega, cg);
  k_mg_solve<<<1, 64, 0, s>>>(m.lv[m.nlev], m.ainv, cg);
  for (int l = m.nlev - 1; l >= 1; l--)
    k_mg_up<<<blocks256(lvl_n(m.lv[l])), 256, 0, s>>>(m.lv[l], m.lv[l + 1], g.omega, g.oc, cg);
  k_mg_prolong0<<<col_grid(d), dim3(TX, TY, 1), 0, s>>>(d, r32, dinv, m.lv[1], g.omega, g.oc, m.t, cg);
  k_mg_post0<OP><<<nt, 256, 0, s>>>(d, m.t, r32, dinv, kap32, pse32, scale, g.omega, z, cg, partials);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

}  // namespace

int MgWork::alloc(const Dims& dd) {
  if (d.n1 == dd.n1 && d.n2 == dd.n2 && d.n3 == dd.n3 && pool.p) return ELSIM_OK;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  MgLevel L[kMgMaxLev + 1]{};
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
    L[nl].m1 = m1;
    L[nl].m2 = m2;
    L[nl].m3 = m3;
    bytes += 7 * up((size_t)m1 * m2 * m3 * sizeof(float));
  } while ((int64_t)m1 * m2 * m3 > kMgCoarseMax);
  const size_t Ge = ((size_t)dd.G() + 1) & ~(size_t)1;
  bytes += up(Ge * sizeof(float)) + up((size_t)kMgCoarseMax * kMgCoarseMax * sizeof(float));
  ELSIM_TRY(dev_alloc(pool, bytes));
  unsigned char* p = (unsigned char*)pool.p;
  for (int l = 1; l <= nl; l++) {
    const size_t b = up((size_t)lvl_n(L[l]) * sizeof(float));
    float** arr[] = {&L[l].cx, &L[l].cy, &L[l].cz, &L[l].dg, &L[l].ms, &L[l].b, &L[l].x};
    for (float** a : arr) {
      *a = (float*)p;
      p += b;
    }
  }
  t = (float*)p;
  p += up(Ge * sizeof(float));
  ainv = (float*)p;
  // zero ghosts of t (only its interior is ever written)
  ELSIM_HIP(hipMemsetAsync(pool.p, 0, bytes, stream()));
  for (int l = 0; l <= kMgMaxLev; l++) lv[l] = L[l];
  nlev = nl;
  d = dd;
  return ELSIM_OK;
}

void MgWork::release() {
  dev_free(pool);
  *this = MgWork{};
}

int mg_partial_blocks(const Dims& d) { return stencil_ntiles<float>(d); }

int mg_setup(MgWork& m, int op, const Dims& d, const double* kap, const double* pse, double scale) {
  ELSIM_TRY(m.alloc(d));
  switch (op) {
    case ELSIM_OP_CNE: return setup_t<ELSIM_OP_CNE>(m, d, kap, pse, scale);
    case ELSIM_OP_PHP: return setup_t<ELSIM_OP_PHP>(m, d, kap, pse, scale);
    case ELSIM_OP_PHE: return setup_t<ELSIM_OP_PHE>(m, d, kap, pse, scale);
    case ELSIM_OP_POT_DD: return setup_t<ELSIM_OP_POT_DD>(m, d, kap, pse, scale);
  }
  set_error("unknown operator %d", op);
  return ELSIM_ERR_ARG;
}

int mg_vcycle(MgWork& m, int op, const Dims& d, const float* kap32, const float* pse32, const uint16_t* dinv,
              float scale, const float* r32, float* z, const CgDev* cg, double* partials) {
  if (m.d.n1 != d.n1 || m.d.n2 != d.n2 || m.d.n3 != d.n3 || !m.pool.p) {
    set_error("multigrid hierarchy not set up for this grid");
    return ELSIM_ERR_ARG;
  }
  switch (op) {
    case ELSIM_OP_CNE: return vcycle_t<ELSIM_OP_CNE>(m, d, kap32, pse32, dinv, scale, r32, z, cg, partials);
    case ELSIM_OP_PHP: return vcycle_t<ELSIM_OP_PHP>(m, d, kap32, pse32, dinv, scale, r32, z, cg, partials);
    case ELSIM_OP_PHE: return vcycle_t<ELSIM_OP_PHE>(m, d, kap32, pse32, dinv, scale, r32, z, cg, partials);
    case ELSIM_OP_POT_DD: return vcycle_t<ELSIM_OP_POT_DD>(m, d, kap32, pse32, dinv, scale, r32, z, cg, partials);
  }
  set_error("unknown operator %d", op);
  return ELSIM_ERR_ARG;
}

}  // namespace elsim
