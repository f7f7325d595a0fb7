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
// Layout: ghosted (n1+2)(n2+2)(n3+2) arrays, d1 contiguous.  A block of
// 64 x 4 threads owns 64 x 4 (d1,d2) columns and marches KC planes along d3,
// keeping p and k of planes k-1, k, k+1 in registers; the d1/d2 neighbours
// come from L1/L2 (each 64-wide row load is one coalesced 512 B / 256 B
// transaction).  The search direction's homogeneous ghost rule
// (SBM:511-513, 666-668, 747-749) is applied in registers, so ghost memory
// of p is never relied on.
//
// Reductions are deterministic: per-block partial sums, then the last block
// to finish adds the partials in index order.  All CG scalars (alpha, beta,
// stop tests) are computed on the device, so the host only polls a done flag
// every few iterations.
#include <cmath>

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

struct ColIdx {
  int i, j, k0, k1;
  bool active;
};
__device__ __forceinline__ ColIdx col_idx(const Dims& d) {
  ColIdx c;
  c.i = blockIdx.x * TX + threadIdx.x + 1;
  c.j = blockIdx.y * TY + threadIdx.y + 1;
  c.k0 = blockIdx.z * KC + 1;
  c.k1 = min(c.k0 + KC - 1, d.n3);
  c.active = c.i <= d.n1 && c.j <= d.n2;
  return c;
}

// A p along one column; epi(c_ghosted, n_interior, p_centre, value).
template <typename T, int OP, bool HOM, class Epi>
__device__ __forceinline__ void column(const Dims& d, const ColIdx& ci, const T* __restrict__ p,
                                       const T* __restrict__ kap, const T* __restrict__ pse, T scale,
                                       Epi&& epi) {
  const int64_t sx = d.g1(), sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  const int i = ci.i, j = ci.j;
  int64_t c = d.gi(i, j, ci.k0);
  int64_t n = d.ni(i, j, ci.k0);
  const T half = T(0.5);
  T pc = p[c], kc = kap[c], km = kap[c - sz];
  T pm = (HOM && ci.k0 == 1) ? pc : p[c - sz];
  for (int k = ci.k0; k <= ci.k1; k++) {
    const T kp = kap[c + sz];
    T pp = p[c + sz];
    T pim = p[c - 1], pip = p[c + 1], pjm = p[c - sx], pjp = p[c + sx];
    if (HOM) {
      if (k == d.n3) pp = pc;
      if (i == 1) pim = pc;
      if (i == d.n1) pip = pc;
      if (j == 1) pjm = Rule<OP>::lo_zero ? T(0) : pc;
      if (j == d.n2) pjp = Rule<OP>::hi_zero ? T(0) : pc;
    }
    const T f1 = (kap[c - 1] + kc) * half, f2 = (kap[c + 1] + kc) * half;
    const T f3 = (kap[c - sx] + kc) * half, f4 = (kap[c + sx] + kc) * half;
    const T f5 = (km + kc) * half, f6 = (kp + kc) * half;
    T s = f2 * (pip - pc) - f1 * (pc - pim);
    s = s + f4 * (pjp - pc);
    s = s - f3 * (pc - pjm);
    s = s + f6 * (pp - pc);
    s = s - f5 * (pc - pm);
    T v;
    if (Rule<OP>::diag)
      v = pse[c] * pc - scale * s;
    else
      v = scale * s;
    epi(c, n, pc, v);
    pm = pc;
    pc = pp;
    km = kc;
    kc = kp;
    c += sz;
    n += nz;
  }
}

// Pointwise walk over a block's columns.
template <class F>
__device__ __forceinline__ void for_col(const Dims& d, const ColIdx& ci, F&& f) {
  if (!ci.active) return;
  const int64_t sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  int64_t c = d.gi(ci.i, ci.j, ci.k0), n = d.ni(ci.i, ci.j, ci.k0);
  for (int k = ci.k0; k <= ci.k1; k++, c += sz, n += nz) f(c, n);
}

// ------------------------------------------------------------- kernels --

// out = A p (hom or stored ghosts); optional dot = p.out into *dot or cg->pw
// with alpha = (num_is_rz ? rz : rr) / pw.
template <typename T, int OP, bool HOM>
__global__ __launch_bounds__(256) void k_matvec(Dims d, const T* __restrict__ p, const T* __restrict__ kap,
                                                const T* __restrict__ pse, T scale, T* __restrict__ out,
                                                double* dot, CgDev* cg, int num_is_rz, double* partials,
                                                unsigned* counter) {
  if (cg && (cg->done || cg->need_reliable)) return;
  const ColIdx ci = col_idx(d);
  double acc[1] = {0.0};
  if (ci.active)
    column<T, OP, HOM>(d, ci, p, kap, pse, scale, [&](int64_t, int64_t n, T pc, T v) {
      out[n] = v;
      acc[0] += (double)pc * (double)v;
    });
  if (dot || cg) {
    double tot[1];
    if (reduce_last_block<1>(acc, partials, counter, tot) && linear_tid() == 0) {
      if (dot) *dot = tot[0];
      if (cg) {
        cg->pw = tot[0];
        cg->alpha = (num_is_rz ? cg->rz : cg->rr) / tot[0];
      }
    }
  }
}

// Initial residual with x's ghosts as stored (SBM:492-499, 649-655, 729-735):
// res = r0 = rhs - A x; prj(interior) = res; x0 = x(interior); cg->rr = sum res^2.
template <int OP>
__global__ __launch_bounds__(256) void k_resid64(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                                 const double* __restrict__ pse, double scale,
                                                 const double* __restrict__ rhs, double* __restrict__ res,
                                                 double* __restrict__ r0, double* __restrict__ prj,
                                                 double* __restrict__ x0, CgDev* cg, double* partials,
                                                 unsigned* counter) {
  const ColIdx ci = col_idx(d);
  double acc[1] = {0.0};
  if (ci.active)
    column<double, OP, false>(d, ci, x, kap, pse, scale, [&](int64_t c, int64_t n, double xc, double v) {
      const double r = rhs[n] - v;
      res[n] = r;
      r0[n] = r;
      if (prj) prj[c] = r;
      x0[n] = xc;
      acc[0] += r * r;
    });
  double tot[1];
  if (reduce_last_block<1>(acc, partials, counter, tot) && linear_tid() == 0) {
    cg->rr = tot[0];
    cg->rmax2 = tot[0];
  }
}

// ||b_eff||^2 with b_eff = r0 + A_hom x (the residual of a zero interior with
// the boundary data of x); sets tol2 and, for the relative rule, done if the
// initial residual already meets it.
template <int OP>
__global__ __launch_bounds__(256) void k_beff(Dims d, const double* __restrict__ x, const double* __restrict__ kap,
                                              const double* __restrict__ pse, double scale,
                                              const double* __restrict__ r0, double rtol, CgDev* cg,
                                              double* partials, unsigned* counter) {
  const ColIdx ci = col_idx(d);
  double acc[1] = {0.0};
  if (ci.active)
    column<double, OP, true>(d, ci, x, kap, pse, scale, [&](int64_t, int64_t n, double, double v) {
      const double b = r0[n] + v;
      acc[0] += b * b;
    });
  double tot[1];
  if (reduce_last_block<1>(acc, partials, counter, tot) && linear_tid() == 0) {
    cg->bn2 = tot[0];
    cg->tol2 = rtol * rtol * tot[0];
    cg->rr_true = cg->rr;
    if (cg->stop_rule == ELSIM_STOP_RELRES && !(cg->rr > cg->tol2)) cg->done = 1;
  }
}

// FP64 reference CG update (SBM:533-547, 688-705, 769-786).
template <int OP>
__global__ __launch_bounds__(256) void k_cg64_update(Dims d, double* __restrict__ x, const double* __restrict__ prj,
                                                     double* __restrict__ res, const double* __restrict__ wmx,
                                                     const double* __restrict__ w8, CgDev* cg, double* partials,
                                                     unsigned* counter) {
  if (cg->done) return;
  const double alp = cg->alpha;
  const ColIdx ci = col_idx(d);
  double acc[2] = {0.0, 0.0};
  for_col(d, ci, [&](int64_t c, int64_t n) {
    const double xo = x[c];
    const double xn = xo + alp * prj[c];
    x[c] = xn;
    const double r = res[n] - alp * wmx[n];
    res[n] = r;
    acc[0] += r * r;
    if (OP == ELSIM_OP_CNE) {
      acc[1] += r * r * w8[c];
    } else {
      const double t = xn - xo;
      acc[1] += t * t * w8[c];
    }
  });
  double tot[2];
  if (reduce_last_block<2>(acc, partials, counter, tot) && linear_tid() == 0) {
    const double dmy = cg->rr, rrn = tot[0];
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
  }
}

// PRJ = RES + bet*PRJ (SBM:542, 697, 778).
__global__ __launch_bounds__(256) void k_cg64_dir(Dims d, double* __restrict__ prj, const double* __restrict__ res,
                                                  const CgDev* cg) {
  if (cg->done) return;
  const double bet = cg->beta;
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n) { prj[c] = res[n] + bet * prj[c]; });
}

// e = x - x0 (interior of a ghosted array).
__global__ __launch_bounds__(256) void k_diff(Dims d, const double* __restrict__ x, const double* __restrict__ x0,
                                              double* __restrict__ e) {
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n) { e[c] = x[c] - x0[n]; });
}

// FP64 true residual of the correction system: r = r0 - A_hom e.
// FINAL: only record rr_true.  Otherwise (mixed reliable update): r32 = r,
// and the Krylov scalars restart from the true residual.
template <int OP, bool FINAL>
__global__ __launch_bounds__(256) void k_true_resid(Dims d, const double* __restrict__ e, const double* __restrict__ kap,
                                                    const double* __restrict__ pse, double scale,
                                                    const double* __restrict__ r0, float* __restrict__ r32,
                                                    const float* __restrict__ dinv, CgDev* cg, double* partials,
                                                    unsigned* counter) {
  if (!FINAL && (cg->done || !cg->need_reliable)) return;
  const ColIdx ci = col_idx(d);
  double acc[2] = {0.0, 0.0};
  if (ci.active)
    column<double, OP, true>(d, ci, e, kap, pse, scale, [&](int64_t, int64_t n, double, double v) {
      const double r = r0[n] - v;
      acc[0] += r * r;
      if (!FINAL) {
        const float rf = (float)r;
        r32[n] = rf;
        acc[1] += (double)rf * (double)(dinv[n] * rf);
      }
    });
  double tot[2];
  if (reduce_last_block<2>(acc, partials, counter, tot) && linear_tid() == 0) {
    cg->rr_true = tot[0];
    if (!FINAL) {
      cg->n_reliable += 1;
      cg->need_reliable = 0;
      if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        cg->done = 3;
      } else if (!(tot[0] > cg->tol2)) {
        cg->done = 1;
      } else if (cg->iters >= cg->maxit) {
        cg->done = 2;
      } else {
        cg->beta = tot[1] / cg->rz;
        cg->rz = tot[1];
        cg->rmax2 = tot[0];
      }
    }
  }
}

// ------------------------------------------------ mixed-precision kernels --

// FP32 copies of the coefficient (and psE) over the ghosted box.
__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// Jacobi: dinv = 1 / diag(A_hom) (FP64, stored FP32).  Mirror faces do not
// contribute; zero-Dirichlet faces do.
template <int OP>
__global__ __launch_bounds__(256) void k_dinv(Dims d, const double* __restrict__ kap, const double* __restrict__ pse,
                                              double scale, float* __restrict__ dinv) {
  const ColIdx ci = col_idx(d);
  const int64_t sx = d.g1(), sz = d.g12();
  for_col(d, ci, [&](int64_t c, int64_t n) {
    const int k = ci.k0 + (int)((c - d.gi(ci.i, ci.j, ci.k0)) / sz);
    const double kc = kap[c];
    double s = 0.0;
    if (ci.i > 1) s += (kap[c - 1] + kc) * 0.5;
    if (ci.i < d.n1) s += (kap[c + 1] + kc) * 0.5;
    if (ci.j > 1 || Rule<OP>::lo_zero) s += (kap[c - sx] + kc) * 0.5;
    if (ci.j < d.n2 || Rule<OP>::hi_zero) s += (kap[c + sx] + kc) * 0.5;
    if (k > 1) s += (kap[c - sz] + kc) * 0.5;
    if (k < d.n3) s += (kap[c + sz] + kc) * 0.5;
    const double dg = Rule<OP>::diag ? pse[c] + scale * s : -scale * s;
    dinv[n] = (float)(1.0 / dg);
  });
}

// Mixed start: r32 = r0, p32 = dinv r32, y32 = 0, e64 = 0; rz = r.z.
__global__ __launch_bounds__(256) void k_mx_start(Dims d, const double* __restrict__ r0, const float* __restrict__ dinv,
                                                  float* __restrict__ r32, float* __restrict__ p32,
                                                  float* __restrict__ y32, double* __restrict__ e64, CgDev* cg,
                                                  double* partials, unsigned* counter) {
  const ColIdx ci = col_idx(d);
  double acc[1] = {0.0};
  for_col(d, ci, [&](int64_t c, int64_t n) {
    const float r = (float)r0[n];
    const float z = dinv[n] * r;
    r32[n] = r;
    p32[c] = z;
    y32[n] = 0.0f;
    e64[c] = 0.0;
    acc[0] += (double)r * (double)z;
  });
  double tot[1];
  if (reduce_last_block<1>(acc, partials, counter, tot) && linear_tid() == 0) cg->rz = tot[0];
}

// y += alpha p; r -= alpha w; rz' = r.(dinv r); decide on a reliable update.
__global__ __launch_bounds__(256) void k_mx_update(Dims d, float* __restrict__ y32, const float* __restrict__ p32,
                                                   float* __restrict__ r32, const float* __restrict__ w32,
                                                   const float* __restrict__ dinv, CgDev* cg, double* partials,
                                                   unsigned* counter) {
  if (cg->done || cg->need_reliable) return;
  const float alp = (float)cg->alpha;
  const ColIdx ci = col_idx(d);
  double acc[2] = {0.0, 0.0};
  for_col(d, ci, [&](int64_t c, int64_t n) {
    y32[n] = y32[n] + alp * p32[c];
    const float r = r32[n] - alp * w32[n];
    r32[n] = r;
    acc[0] += (double)r * (double)(dinv[n] * r);
    acc[1] += (double)r * (double)r;
  });
  double tot[2];
  if (reduce_last_block<2>(acc, partials, counter, tot) && linear_tid() == 0) {
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
  }
}

// Reliable update, part 1: e64 += y32, y32 = 0.
__global__ __launch_bounds__(256) void k_mx_fold(Dims d, double* __restrict__ e64, float* __restrict__ y32,
                                                 const CgDev* cg) {
  if (cg->done || !cg->need_reliable) return;
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n) {
    e64[c] = e64[c] + (double)y32[n];
    y32[n] = 0.0f;
  });
}

// p = dinv r + beta p.
__global__ __launch_bounds__(256) void k_mx_dir(Dims d, float* __restrict__ p32, const float* __restrict__ r32,
                                                const float* __restrict__ dinv, const CgDev* cg) {
  if (cg->done || cg->need_reliable) return;
  const float bet = (float)cg->beta;
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n) { p32[c] = dinv[n] * r32[n] + bet * p32[c]; });
}

// x += e (interior).
__global__ __launch_bounds__(256) void k_add_e(Dims d, double* __restrict__ x, const double* __restrict__ e) {
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t) { x[c] = x[c] + e[c]; });
}

__global__ __launch_bounds__(256) void k_copy_g2n(Dims d, const double* __restrict__ g, double* __restrict__ nn) {
  const ColIdx ci = col_idx(d);
  for_col(d, ci, [&](int64_t c, int64_t n) { nn[n] = g[c]; });
}

// ------------------------------------------------------------------ host --

int CgWork::alloc(const Dims& dd) {
  if (d.n1 == dd.n1 && d.n2 == dd.n2 && d.n3 == dd.n3 && cg.p) return ELSIM_OK;
  d = dd;
  const size_t N = (size_t)d.N(), G = (size_t)d.G();
  ELSIM_TRY(dev_alloc(res, N * 8));
  ELSIM_TRY(dev_alloc(wmx, N * 8));
  ELSIM_TRY(dev_alloc(prj, G * 8));
  ELSIM_TRY(dev_alloc(x0, N * 8));
  ELSIM_TRY(dev_alloc(r0, N * 8));
  ELSIM_TRY(dev_alloc(e64, G * 8));
  ELSIM_TRY(dev_alloc(y32, N * 4));
  ELSIM_TRY(dev_alloc(r32, N * 4));
  ELSIM_TRY(dev_alloc(w32, N * 4));
  ELSIM_TRY(dev_alloc(p32, G * 4));
  ELSIM_TRY(dev_alloc(kap32, G * 4));
  ELSIM_TRY(dev_alloc(pse32, G * 4));
  ELSIM_TRY(dev_alloc(dinv32, N * 4));
  ELSIM_TRY(dev_alloc(cg, sizeof(CgDev)));
  if (!host) ELSIM_HIP(hipHostMalloc((void**)&host, sizeof(CgDev), hipHostMallocDefault));
  // ghost memory of the search directions is read (then replaced in
  // registers): give it defined contents once
  ELSIM_HIP(hipMemsetAsync(prj.p, 0, G * 8, stream()));
  ELSIM_HIP(hipMemsetAsync(p32.p, 0, G * 4, stream()));
  ELSIM_HIP(hipMemsetAsync(e64.p, 0, G * 8, stream()));
  return ELSIM_OK;
}

void CgWork::release() {
  DevBuf* bs[] = {&res, &wmx, &prj, &x0, &r0, &e64, &y32, &r32, &w32, &p32, &kap32, &pse32, &dinv32, &cg};
  for (DevBuf* b : bs) dev_free(*b);
  if (host) (void)hipHostFree(host);
  host = nullptr;
  d = Dims{0, 0, 0};
}

namespace {

template <int OP>
int solve_op(CgWork& w, const SolveArgs& a, int mode, SolveResult* out) {
  const Dims& d = a.d;
  Workspace& ws = workspace();
  hipStream_t s = stream();
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  if (blocks_of(grid) > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  CgDev* cg = (CgDev*)w.cg.p;
  CgDev h{};
  h.crt = a.crt;
  h.tw = a.tw;
  h.maxit = a.maxit;
  h.stop_rule = mode == ELSIM_MODE_MIXED ? ELSIM_STOP_RELRES : a.stop_rule;
  h.delta2 = a.delta * a.delta;
  ELSIM_HIP(hipMemcpyAsync(cg, &h, sizeof(CgDev), hipMemcpyHostToDevice, s));
  double* res = (double*)w.res.p;
  double* wmx = (double*)w.wmx.p;
  double* prj = (double*)w.prj.p;
  double* x0 = (double*)w.x0.p;
  double* r0 = (double*)w.r0.p;
  double* e64 = (double*)w.e64.p;
  float* dinv = (float*)w.dinv32.p;
  k_resid64<OP><<<grid, blk, 0, s>>>(d, a.x, a.kap, a.pse, a.scale, a.rhs, res, r0, mode == ELSIM_MODE_FP64 ? prj : nullptr,
                                     x0, cg, ws.partials, ws.counter);
  k_beff<OP><<<grid, blk, 0, s>>>(d, a.x, a.kap, a.pse, a.scale, r0, a.rtol, cg, ws.partials, ws.counter);
  ELSIM_CHECK_LAUNCH();
  const int batch = 8;
  if (mode == ELSIM_MODE_FP64) {
    for (;;) {
      for (int b = 0; b < batch; b++) {
        k_matvec<double, OP, true><<<grid, blk, 0, s>>>(d, prj, a.kap, a.pse, a.scale, wmx, nullptr, cg, 0, ws.partials,
                                                        ws.counter);
        k_cg64_update<OP><<<grid, blk, 0, s>>>(d, a.x, prj, res, wmx, a.w8, cg, ws.partials, ws.counter);
        k_cg64_dir<<<grid, blk, 0, s>>>(d, prj, res, cg);
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_HIP(hipMemcpyAsync(w.host, cg, sizeof(CgDev), hipMemcpyDeviceToHost, s));
      ELSIM_HIP(hipStreamSynchronize(s));
      if (w.host->done) break;
    }
    // true residual of the correction system for the report
    k_diff<<<grid, blk, 0, s>>>(d, a.x, x0, e64);
    k_true_resid<OP, true><<<grid, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, nullptr, nullptr, cg, ws.partials,
                                                ws.counter);
  } else {
    const int64_t G = d.G();
    k_to_f32<<<flat_blocks(G), FLAT_T, 0, s>>>(G, a.kap, (float*)w.kap32.p);
    if (OP == ELSIM_OP_CNE) k_to_f32<<<flat_blocks(G), FLAT_T, 0, s>>>(G, a.pse, (float*)w.pse32.p);
    k_dinv<OP><<<grid, blk, 0, s>>>(d, a.kap, a.pse, a.scale, dinv);
    k_mx_start<<<grid, blk, 0, s>>>(d, r0, dinv, (float*)w.r32.p, (float*)w.p32.p, (float*)w.y32.p, e64, cg,
                                    ws.partials, ws.counter);
    ELSIM_CHECK_LAUNCH();
    const float sc32 = (float)a.scale;
    // the first iteration compares against the initial residual norm
    for (;;) {
      for (int b = 0; b < batch; b++) {
        k_matvec<float, OP, true><<<grid, blk, 0, s>>>(d, (const float*)w.p32.p, (const float*)w.kap32.p,
                                                       (const float*)w.pse32.p, sc32, (float*)w.w32.p, nullptr, cg, 1,
                                                       ws.partials, ws.counter);
        k_mx_update<<<grid, blk, 0, s>>>(d, (float*)w.y32.p, (const float*)w.p32.p, (float*)w.r32.p,
                                         (const float*)w.w32.p, dinv, cg, ws.partials, ws.counter);
        k_mx_fold<<<grid, blk, 0, s>>>(d, e64, (float*)w.y32.p, cg);
        k_true_resid<OP, false><<<grid, blk, 0, s>>>(d, e64, a.kap, a.pse, a.scale, r0, (float*)w.r32.p, dinv, cg,
                                                     ws.partials, ws.counter);
        k_mx_dir<<<grid, blk, 0, s>>>(d, (float*)w.p32.p, (const float*)w.r32.p, dinv, cg);
      }
      ELSIM_CHECK_LAUNCH();
      ELSIM_HIP(hipMemcpyAsync(w.host, cg, sizeof(CgDev), hipMemcpyDeviceToHost, s));
      ELSIM_HIP(hipStreamSynchronize(s));
      if (w.host->done) break;
    }
    k_add_e<<<grid, blk, 0, s>>>(d, a.x, e64);
  }
  ELSIM_CHECK_LAUNCH();
  ELSIM_HIP(hipMemcpyAsync(w.host, cg, sizeof(CgDev), hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  const CgDev& r = *w.host;
  out->iters = r.iters;
  out->n_reliable = r.n_reliable;
  out->relres = r.bn2 > 0 ? std::sqrt(r.rr_true / r.bn2) : 0.0;
  out->status = r.done == 1 ? ELSIM_OK : (r.done == 2 ? ELSIM_NOT_CONVERGED : ELSIM_ERR_NAN);
  if (out->status == ELSIM_ERR_NAN) set_error("CG produced a non-finite value (op %d)", OP);
  return ELSIM_OK;
}

}  // namespace

int cg_solve(CgWork& w, const SolveArgs& a, int mode, SolveResult* res) {
  ELSIM_TRY(w.alloc(a.d));
  if (mode == ELSIM_MODE_MIXED && a.stop_rule != ELSIM_STOP_RELRES) {
    set_error("mixed precision requires the relative-residual stop rule");
    return ELSIM_ERR_ARG;
  }
  switch (a.op) {
    case ELSIM_OP_CNE: return solve_op<ELSIM_OP_CNE>(w, a, mode, res);
    case ELSIM_OP_PHP: return solve_op<ELSIM_OP_PHP>(w, a, mode, res);
    case ELSIM_OP_PHE: return solve_op<ELSIM_OP_PHE>(w, a, mode, res);
  }
  set_error("unknown operator %d", a.op);
  return ELSIM_ERR_ARG;
}

template <typename T>
static int launch_mv(int op, const Dims& d, const T* p, const T* kap, const T* pse, T scale, T* out, double* dot,
                     bool hom) {
  ELSIM_TRY(ensure_init());
  Workspace& ws = workspace();
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  if (blocks_of(grid) > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  hipStream_t s = stream();
#define MV(OPV, H) \
  k_matvec<T, OPV, H><<<grid, blk, 0, s>>>(d, p, kap, pse, scale, out, dot, nullptr, 0, ws.partials, ws.counter)
  if (op == ELSIM_OP_CNE) {
    if (hom) MV(ELSIM_OP_CNE, true); else MV(ELSIM_OP_CNE, false);
  } else if (op == ELSIM_OP_PHP) {
    if (hom) MV(ELSIM_OP_PHP, true); else MV(ELSIM_OP_PHP, false);
  } else if (op == ELSIM_OP_PHE) {
    if (hom) MV(ELSIM_OP_PHE, true); else MV(ELSIM_OP_PHE, false);
  } else {
    set_error("unknown operator %d", op);
    return ELSIM_ERR_ARG;
  }
#undef MV
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

}  // extern "C"
