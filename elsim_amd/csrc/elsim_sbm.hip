// elsim_sbm.hip -- device-resident SBM electrochemical step
// (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:143-840 on one GPU).
//
// The whole domain lives in HBM: ghosted FP64 fields (psP, psE, CnP, CnE,
// phP, phE and the cell-centred coefficients) plus interior work arrays.  The
// MPI halo exchange BcMpiYZ (SBM:1031-1090) becomes in-kernel boundary rules:
// on one rank its outer boundary is a Neumann mirror, and the d2 faces carry
// the per-field Dirichlet / flux values (SBM:557-560, 647, 717-719, 798-800,
// 833-835).  Every kernel restates the reference expression in its
// evaluation order (compiled without FMA contraction), so the arithmetic
// kernels are bit-identical to the FP64 CPU restatement; only reductions
// (tree order) and exp/tanh (device libm) differ at the rounding level.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "elsim_nn.h"
#include "elsim_solver.h"

namespace elsim {

struct SbmScal {
  double tpp, tpe, xrf, infx, bve, scrnt, dv, avcne, cvgp, cvgl, red;
  double shift;  // the current control's uniform shift of phi_E applied at the end of the last step (dv dt)
};

// Generic reduction ops finished by the last block.
enum RedOp { R_TPP = 0, R_TPE, R_XRF, R_INFX, R_AVCNE, R_CVGP, R_CVGL, R_SCRNT, R_RAW };

struct BcArgs {
  int lo, hi;          // 0 mirror, 1 value, 2 CnE flux (low face only), 3 keep the stored value
  double lo_v, hi_v;   // values (kind 1) unless lo_vp / hi_vp given
  const double* lo_vp; // device scalar (e.g. BvE)
  const double* dab01; // CnE flux: Dab on planes j=0 and j=1, (n1+2)(n3+2) each
  const double* infx;  // CnE flux: device scalar
  double tm, dx;       // CnE flux: t_minus, dx
};

namespace {

__device__ __forceinline__ double table(double c, const double* __restrict__ cord, const double* __restrict__ vals,
                                        int sz) {
  // Tbl_Chk_Pln_3DA (SBM:998-1021)
  const double ditv = cord[1] - cord[0];
  if (c < 0.0) c = 1.0e-6;
  if (c > 1.0) c = 1.0;
  int lb = (int)floor(c / ditv);
  if (lb == 0)
    lb = 1;
  else if (lb == sz)
    lb = sz - 1;
  // a well-formed table (uniform grid over [0,1]) never needs this clamp;
  // it keeps malformed tables and NaN inside the arrays
  if (lb < 1 || lb > sz - 1) lb = lb < 1 ? 1 : sz - 1;
  return vals[lb - 1] + (c - cord[lb - 1]) / ditv * (vals[lb] - vals[lb - 1]);
}

using Col = ColIdx;
__device__ __forceinline__ Col colx(const Dims& d) { return col_idx(d); }
template <class F>
__device__ __forceinline__ void each(const Dims& d, const Col& ci, F&& f) {
  if (!ci.active) return;
  const int64_t sz = d.g12(), nz = (int64_t)d.n1 * d.n2;
  int64_t c = d.gi(ci.i, ci.j, ci.k0), n = d.ni(ci.i, ci.j, ci.k0);
  for (int k = ci.k0; k <= ci.k1; k++, c += sz, n += nz) f(c, n, k);
}

__device__ __forceinline__ void finish_red(int op, double tot, SbmScal* s, const elsim_sbm_params& P, const Dims& d,
                                           double* raw) {
  switch (op) {
    case R_TPP: s->tpp = tot; break;
    case R_TPE: s->tpe = tot; break;
    case R_XRF: s->xrf = tot / s->tpp; break;
    case R_INFX: {
      const double lsm = tot * (P.dx * P.dx * P.dx);
      s->infx = lsm / ((double)(d.n1 * d.n3) * (P.dx * P.dx));
      break;
    }
    case R_AVCNE: s->avcne = tot / s->tpe; break;
    case R_CVGP: s->cvgp = sqrt(tot / s->tpp); break;
    case R_CVGL: s->cvgl = sqrt(tot / s->tpe); break;
    case R_SCRNT: {
      s->scrnt = tot;
      const double dv = (P.trgi - tot) >= 0 ? fabs(P.vsr) : -fabs(P.vsr);
      s->dv = dv;
      s->shift = dv * P.dt;
      s->bve = s->bve + dv * P.dt;
      break;
    }
    default: s->red = tot; break;
  }
  if (raw) *raw = tot;
}

// one-block finalisation of an SBM reduction
__global__ __launch_bounds__(FIN_T) void k_fin_sbm(int op, int nb, const double* __restrict__ partials, SbmScal* s,
                                                   elsim_sbm_params P, Dims d) {
  double tot[1];
  sum_partials<1>(partials, nb, tot);
  if (linear_tid() == 0) finish_red(op, tot[0], s, P, d, nullptr);
}

// sum over the interior of a(G) * w(G) (w may be null)
__global__ __launch_bounds__(256) void k_red_gw(Dims d, const double* __restrict__ a, const double* __restrict__ w,
                                                int op, SbmScal* s, elsim_sbm_params P, double* partials) {
  const Col ci = colx(d);
  double acc[1] = {0.0};
  each(d, ci, [&](int64_t c, int64_t, int) { acc[0] += w ? a[c] * w[c] : a[c]; });
  block_partials<1>(acc, partials);
}

// sum over the interior of (x - x0)^2 w (all ghosted); dx (optional): the
// interior of x - x0 as well (the iteration's correction, for the next
// coupling iteration's Galerkin step).  e (optional): x still holds the
// solve's start x0 and the solve's correction e is applied here first
// (x <- x + e on the interior; the same values as a separate pass).
__global__ __launch_bounds__(256) void k_red_cvg(Dims d, double* __restrict__ x, const double* __restrict__ x0,
                                                 const double* __restrict__ w, int op, SbmScal* s, elsim_sbm_params P,
                                                 double* partials, double* __restrict__ dx,
                                                 const double* __restrict__ e) {
  const Col ci = colx(d);
  double acc[1] = {0.0};
  each(d, ci, [&](int64_t c, int64_t, int) {
    double t;
    if (e) {  // x still holds the solve's start (x0): apply its correction here
      const double xo = x[c], xn = xo + e[c];
      x[c] = xn;
      t = xn - xo;
    } else {
      t = x[c] - x0[c];
    }
    if (dx) dx[c] = t;
    acc[0] += t * t * w[c];
  });
  block_partials<1>(acc, partials);
}

// sum over the interior of rxn(N) * (avp(N) / dx)
__global__ __launch_bounds__(256) void k_red_rxn(Dims d, const double* __restrict__ rxn, const double* __restrict__ avp,
                                                 int op, SbmScal* s, elsim_sbm_params P, double* partials) {
  const Col ci = colx(d);
  double acc[1] = {0.0};
  each(d, ci, [&](int64_t, int64_t n, int) { acc[0] += rxn[n] * (avp[n] / P.dx); });
  block_partials<1>(acc, partials);
}

// psi = clamp(0.5(1+tanh(d/zeta))) + off on the interior (SBM:154-161, 247/258)
__global__ __launch_bounds__(256) void k_psi(Dims d, const float* __restrict__ sdf, double zeta, double off,
                                             double* __restrict__ psi) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t c, int64_t n, int) {
    double g = (double)sdf[n];
    g = 0.5 * (1.0 + tanh(g / zeta));
    if (g < 0.0) g = 0.0;
    if (g > 1.0) g = 1.0;
    psi[c] = g + off;
  });
}

// Ghost cells of a ghosted array.  Every ghost takes the value of its clamped
// interior neighbour (the one-rank BcMpiYZ outer boundary plus the d2 mirror,
// SBM:163-176), except the d2 planes whose rule is a value or the CnE flux.
// One thread per ghost cell (ghost_cell's numbering), 32-bit index
// arithmetic: the launch is latency-bound, so no thread walks several cells.
__global__ __launch_bounds__(256) void k_bc(Dims d, double* __restrict__ f, BcArgs b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ghost_cells(d)) return;
  int i, j, k;
  ghost_cell(d, q, i, j, k);
  const int ic = min(max(i, 1), d.n1), jc = min(max(j, 1), d.n2), kc = min(max(k, 1), d.n3);
  double v = f[d.gi(ic, jc, kc)];
  if (j == 0 && b.lo == 1) {
    v = b.lo_vp ? *b.lo_vp : b.lo_v;
  } else if (j == 0 && b.lo == 2) {
    // SBM:558: infx*t_minus*dx/(0.5*(Dab(:,0,:)+Dab(:,1,:))) + CnE(:,1,:)
    const int64_t pl = (int64_t)(d.n1 + 2) * (d.n3 + 2);
    const int64_t o = i + (int64_t)(d.n1 + 2) * k;
    v = *b.infx * b.tm * b.dx / (0.5 * (b.dab01[o] + b.dab01[pl + o])) + v;
  } else if (j == d.n2 + 1 && b.hi == 1) {
    v = b.hi_v;
  }
  if ((j == 0 && b.lo == 3) || (j == d.n2 + 1 && b.hi == 3)) v = f[d.gi(i, j, k)];
  f[d.gi(i, j, k)] = v;
}

// AvP (SBM:274-289), interior
__global__ __launch_bounds__(256) void k_avp(Dims d, const double* __restrict__ psp, const double* __restrict__ pse,
                                             double amin, double* __restrict__ avp) {
  const Col ci = colx(d);
  const int64_t sx = d.g1(), sz = d.g12();
  each(d, ci, [&](int64_t c, int64_t n, int) {
    double a = (psp[c + 1] - psp[c - 1]) / 2;
    double b = (psp[c + sx] - psp[c - sx]) / 2;
    double e = (psp[c + sz] - psp[c - sz]) / 2;
    const double vp = sqrt(a * a + b * b + e * e);
    a = (pse[c + 1] - pse[c - 1]) / 2;
    b = (pse[c + sx] - pse[c - sx]) / 2;
    e = (pse[c + sz] - pse[c - sz]) / 2;
    const double ve = sqrt(a * a + b * b + e * e);
    double v = sqrt(vp * ve);
    if (v < amin) v = 0.0;
    avp[n] = v;
  });
}

__global__ void k_scale_g(int64_t n, const double* __restrict__ a, double s, double* __restrict__ b) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    b[q] = s * a[q];
}

// CH part 1 (SBM:398-419): mu = T_mu(c) - eps Lap c, pm = psP T_M(c), interior.
__global__ __launch_bounds__(256) void k_ch_mu(Dims d, const double* __restrict__ c, const double* __restrict__ psp,
                                               const double* __restrict__ tab, int sz, double eps, double dx,
                                               double* __restrict__ mu, double* __restrict__ pm) {
  // the column's d3 neighbours roll through registers (planes k - 1, k, k + 1)
  const Col ci = colx(d);
  if (!ci.active) return;
  const int64_t sx = d.g1(), sz3 = d.g12(), nz = (int64_t)d.n1 * d.n2;
  int64_t q = d.gi(ci.i, ci.j, ci.k0), n = d.ni(ci.i, ci.j, ci.k0);
  double cb = c[q - sz3], cc = c[q];
  for (int k = ci.k0; k <= ci.k1; k++, q += sz3, n += nz) {
    const double ca = c[q + sz3];
    double v = c[q - sx] - 2 * cc + c[q + sx];
    v = v + c[q - 1] - 2 * cc + c[q + 1];
    v = v + cb - 2 * cc + ca;
    const double lap = v / (dx * dx);
    mu[n] = table(cc, tab, tab + sz, sz) - eps * lap;
    pm[n] = psp[q] * table(cc, tab + 6 * sz, tab + 7 * sz, sz);
    cb = cc;
    cc = ca;
  }
}

// CH part 2 (SBM:422-443): DvC with mirrored mu/pm, update where psP > min,
// Xrf partial.
__global__ __launch_bounds__(256) void k_ch_update(Dims d, double* __restrict__ c, const double* __restrict__ mu,
                                                   const double* __restrict__ pm, const double* __restrict__ psp,
                                                   const double* __restrict__ rxn, const double* __restrict__ avp,
                                                   elsim_sbm_params P, SbmScal* s, double* partials) {
  // the column's d3 neighbours of mu and pm roll through registers (the
  // mirrored value at the box's end planes, as the offsets zp, zm = 0 give)
  const Col ci = colx(d);
  const int64_t nx = d.n1, nz = (int64_t)d.n1 * d.n2;
  double acc[1] = {0.0};
  if (ci.active) {
    const int i = ci.i, j = ci.j;
    const int64_t yp = i < d.n1 ? 1 : 0, ym = i > 1 ? 1 : 0;
    const int64_t xp = j < d.n2 ? nx : 0, xm = j > 1 ? nx : 0;
    const int64_t sz3 = d.g12();
    int64_t q = d.gi(i, j, ci.k0), n = d.ni(i, j, ci.k0);
    double m0 = mu[n], p0 = pm[n];
    double mb = ci.k0 > 1 ? mu[n - nz] : m0, pb = ci.k0 > 1 ? pm[n - nz] : p0;
    for (int k = ci.k0; k <= ci.k1; k++, q += sz3, n += nz) {
      const double ma = k < d.n3 ? mu[n + nz] : m0, pa = k < d.n3 ? pm[n + nz] : p0;
      double v = 0.5 * (pm[n + yp] + p0) * (mu[n + yp] - m0) - 0.5 * (p0 + pm[n - ym]) * (m0 - mu[n - ym]);
      v = v + 1.0 * 0.5 * (pm[n + xp] + p0) * (mu[n + xp] - m0);
      v = v - 1.0 * 0.5 * (p0 + pm[n - xm]) * (m0 - mu[n - xm]);
      v = v + 0.5 * (pa + p0) * (ma - m0);
      v = v - 0.5 * (p0 + pb) * (m0 - mb);
      const double dvc = v / (P.dx * P.dx);
      double cn = c[q];
      const double ps = psp[q];
      if (ps > P.ch_psi_min) {
        cn = cn + P.dt * (dvc + (rxn[n] / P.rho) * (avp[n] / P.dx)) / ps;
        c[q] = cn;
      }
      acc[0] += cn * ps;
      mb = m0;
      pb = p0;
      m0 = ma;
      p0 = pa;
    }
  }
  block_partials<1>(acc, partials);
}

// CnE right-hand side (SBM:461-462) and the influx partial (SBM:465-467).
__global__ __launch_bounds__(256) void k_cne_rhs(Dims d, const double* __restrict__ cne, const double* __restrict__ pse,
                                                 const double* __restrict__ rxn, const double* __restrict__ avp,
                                                 elsim_sbm_params P, double* __restrict__ rhs, SbmScal* s,
                                                 double* partials) {
  const Col ci = colx(d);
  double acc[1] = {0.0};
  each(d, ci, [&](int64_t q, int64_t n, int) {
    const double ax = avp[n] / P.dx;
    rhs[n] = pse[q] * cne[q] - P.dt * (rxn[n]) * ax * P.t_minus;
    acc[0] += rxn[n] * ax;
  });
  block_partials<1>(acc, partials);
}

// Cell coefficient over the ghosted box: out = psE * [c *] (scale*D0)*exp(c0 - c1 c + c2 (c3 c)^2)
// (SBM:474+478, 579+583, 601+604).  dab01 (optional): D on planes j = 0, 1.
__global__ void k_coef(Dims d, const double* __restrict__ cne, const double* __restrict__ pse, double sd0,
                       elsim_sbm_params P, int times_c, double* __restrict__ out, double* __restrict__ dab01) {
  const int64_t G = d.G(), g1 = d.g1(), g12 = d.g12();
  const int64_t pl = (int64_t)(d.n1 + 2) * (d.n3 + 2);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < G; q += (int64_t)gridDim.x * blockDim.x) {
    const double c = cne[q];
    const double e = P.dab_c3 * c;
    const double dab = sd0 * exp(P.dab_c0 - P.dab_c1 * c + P.dab_c2 * (e * e));
    out[q] = times_c ? pse[q] * c * dab : pse[q] * dab;
    if (dab01) {
      const int64_t j = (q / g1) % (d.n2 + 2);
      if (j <= 1) {
        const int64_t i = q % g1, k = q / g12;
        dab01[j * pl + i + (int64_t)(d.n1 + 2) * k] = dab;
      }
    }
  }
}

// Both potential coefficients in one pass over c_E and psi_E (SBM:579-583,
// 601-604): pce = psE (sd0 D), pcl = psE c (sd1 D), D = exp(...) of k_coef,
// evaluated once (the same values as two k_coef passes).
__global__ __launch_bounds__(256) void k_coef_pair(int64_t G, const double* __restrict__ cne,
                                                   const double* __restrict__ pse, double sd0, double sd1,
                                                   elsim_sbm_params P, double* __restrict__ pce,
                                                   double* __restrict__ pcl) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= G) return;
  const double c = cne[q], w = pse[q];
  const double e = P.dab_c3 * c;
  const double ex = exp(P.dab_c0 - P.dab_c1 * c + P.dab_c2 * (e * e));
  pce[q] = w * (sd0 * ex);
  pcl[q] = w * c * (sd1 * ex);
}

// interior += delta (device scalar expression)
__global__ __launch_bounds__(256) void k_shift(Dims d, double* __restrict__ f, const SbmScal* s, int which,
                                               double target, double dt) {
  const Col ci = colx(d);
  // which 0: CnE mass correction (SBM:567); 1: phE current control (SBM:831)
  const double del = which == 0 ? (s->avcne - target) : s->dv * dt;
  each(d, ci, [&](int64_t q, int64_t, int) { f[q] = which == 0 ? f[q] - del : f[q] + del; });
}

__global__ void k_table_n(int64_t n, const double* __restrict__ cord, const double* __restrict__ vals, int sz,
                          const double* __restrict__ c, double* __restrict__ f) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    f[q] = table(c[q], cord, vals, sz);
}

// Butler-Volmer (SBM:631-641) and both potential right-hand sides
// (SBM:645, 727).  The rate constants Kfw, Kbw (SBM:614-622) depend on
// c_P alone, which the coupling loop does not change, so they are formed
// here from c_P, where it is read anyway, instead of being stored once and
// read back by every evaluation.  WS: the step's first evaluation also
// extrapolates both potentials first (phi_P += dir, phi_E += dir_e - shift,
// interior), the values k_ws_axpy and k_ws_axpy_shift would store.
template <bool WS>
__global__ __launch_bounds__(256) void k_bv(Dims d, double* __restrict__ php, double* __restrict__ phe,
                                            const double* __restrict__ cne, const double* __restrict__ cnp,
                                            const double* __restrict__ tab, int sz, const double* __restrict__ avp,
                                            const double* __restrict__ rhe, elsim_sbm_params P,
                                            double* __restrict__ rxn, double* __restrict__ rxc,
                                            double* __restrict__ rhsp, double* __restrict__ rhse,
                                            const double* __restrict__ dir, const double* __restrict__ dir_e,
                                            double shift) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t q, int64_t n, int) {
    double pp, pe;
    if constexpr (WS) {
      pp = php[q] + dir[q];
      pe = phe[q] + (dir_e[q] - shift);
      php[q] = pp;
      phe[q] = pe;
    }
    const double a = avp[n];
    if (a > P.avp_min) {
      if constexpr (!WS) {
        pp = php[q];
        pe = phe[q];
      }
      const double cp = cnp[q];
      const double ioc = table(cp, tab + 2 * sz, tab + 3 * sz, sz) * P.i0_scale;
      const double ocv = table(cp, tab + 4 * sz, tab + 5 * sz, sz);
      const double kfw = ioc / (P.frd * P.kfw_cref) * exp(0.5 * P.cst1 * ocv);
      const double kbw = ioc / (P.frd * cp) * exp(-(0.5 * P.cst1 * ocv));
      const double dph = pp - pe;
      const double rx = kfw * cne[q] * exp(-(0.5 * P.cst1 * dph)) - kbw * cp * exp(0.5 * P.cst1 * dph);
      rxn[n] = rx;
      rxc[n] = rx * P.frd;
    }
    const double ax = a / P.dx;
    rhsp[n] = -(rxc[n] * ax);
    rhse[n] = (rxn[n]) * ax + rhe[n];
  });
}

}  // namespace

// ------------------------------------------------------------------ state --

// ------------------------------------------------ learned warm start --

// Network input x[voxel][16 or 8] (bf16, full or not) at the first phP
// solve of a step:
// psi_P, the FP64 residual of the extrapolated start in units of its rms,
// the through-thickness coordinate (d2, the axis of the boundary
// conditions) in (-1, 1), log10 of the phP coefficient over 7 (its range is
// 7 decades) and the last step's change of phi_P in units of its rms; the
// other channels are zero.
__global__ __launch_bounds__(256) void k_feat_pack(Dims d, const double* __restrict__ psp,
                                                   const double* __restrict__ r0, const double* __restrict__ kap,
                                                   const double* __restrict__ dphi, double ir, double idp,
                                                   int full, uint16_t* __restrict__ x) {
  const Col ci = colx(d);
  const float depth = (float)(2.0 * (ci.j - 0.5) / d.n2 - 1.0);
  each(d, ci, [&](int64_t c, int64_t n, int) {
    uint16_t h[16];
    h[0] = f2bf((float)psp[c]);
    h[1] = f2bf((float)(r0[n] * ir));
    h[2] = f2bf(depth);
    h[3] = f2bf((float)(log10(fmax(kap[c], 1e-30)) / 7.0));
    h[4] = f2bf((float)(dphi[c] * idp));
    for (int q = 5; q < 16; q++) h[q] = 0;
    uint4* o = (uint4*)(x + n * (full ? 16 : 8));
    o[0] = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16, h[4] | (uint32_t)h[5] << 16,
                      h[6] | (uint32_t)h[7] << 16);
    if (full)
      o[1] = make_uint4(h[8] | (uint32_t)h[9] << 16, h[10] | (uint32_t)h[11] << 16, h[12] | (uint32_t)h[13] << 16,
                        h[14] | (uint32_t)h[15] << 16);
  });
}

// sum over the interior of a^2 (ghosted) -- the rms of the last change.
__global__ __launch_bounds__(256) void k_sum_sq_g(Dims d, const double* __restrict__ a, double* partials) {
  const Col ci = colx(d);
  double acc[1] = {0.0};
  each(d, ci, [&](int64_t c, int64_t, int) { acc[0] += a[c] * a[c]; });
  block_partials<1>(acc, partials);
}

// interior copy out[n] = a[c].
__global__ __launch_bounds__(256) void k_interior(Dims d, const double* __restrict__ a, double* __restrict__ out) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t c, int64_t n, int) { out[n] = a[c]; });
}

// out = a - prev, then prev = a, over the whole ghosted box (the last
// step's change of a field and the copy kept for the next one, in one pass).
__global__ __launch_bounds__(256) void k_ws_diff_keep(int64_t G, const double* __restrict__ a,
                                                      double* __restrict__ prev, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * (FLAT_T * FLAT_U) + threadIdx.x;
  double av[FLAT_U], bv[FLAT_U];
#pragma unroll
  for (int u = 0; u < FLAT_U; u++) {
    const int64_t q = e + u * FLAT_T;
    if (q < G) {
      av[u] = a[q];
      bv[u] = prev[q];
    }
  }
#pragma unroll
  for (int u = 0; u < FLAT_U; u++) {
    const int64_t q = e + u * FLAT_T;
    if (q < G) {
      out[q] = av[u] - bv[u];
      prev[q] = av[u];
    }
  }
}

// out(interior) = a - b (ghosts of out untouched).
__global__ __launch_bounds__(256) void k_diff_interior(Dims d, const double* __restrict__ a,
                                                       const double* __restrict__ b, double* __restrict__ out) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t c, int64_t, int) { out[c] = a[c] - b[c]; });
}

// out(interior) = a - 2 b + c (ghosts of out untouched).
__global__ __launch_bounds__(256) void k_diff3_interior(Dims d, const double* __restrict__ a,
                                                        const double* __restrict__ b, const double* __restrict__ c,
                                                        double* __restrict__ out) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t i, int64_t, int) { out[i] = (a[i] - 2.0 * b[i]) + c[i]; });
}

// x(interior) += alpha d.
__global__ __launch_bounds__(256) void k_ws_axpy(Dims d, double* __restrict__ x, const double* __restrict__ dv,
                                                 double alpha) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t c, int64_t, int) { x[c] += alpha * dv[c]; });
}

// x(interior) += dv - shift
__global__ __launch_bounds__(256) void k_ws_axpy_shift(Dims d, double* __restrict__ x, const double* __restrict__ dv,
                                                       double shift) {
  const Col ci = colx(d);
  each(d, ci, [&](int64_t c, int64_t, int) { x[c] += dv[c] - shift; });
}

struct SbmState {
  Dims d{0, 0, 0};
  elsim_sbm_params P{};
  int sz = 0;
  int iter = 1;
  DevBuf tab, psp, pse, cnp, cne, php, phe, kapp, pce, pcl;  // ghosted (tab: 8 sz)
  DevBuf avp, rxn, rxc, rhsp, rhse, rhsc, rhe, mu, pm;  // interior
  DevBuf dab01, scal;
  CgWork cg;
  SbmScal* hs = nullptr;  // pinned mirror
  // learned warm start: phi_P of the previous step, the network, its input,
  // output and the candidate field
  DevBuf php_prev, phe_prev, nn_x, ws_cand, ws_dir, ws_dir_e, ws_r0, outer_dp, outer_de;
  // phi_P's changes over the two steps before the last one (ws_p_hist of
  // them valid)
  DevBuf ws_dir2, ws_dir3;
  int ws_p_hist = 0;
  uint16_t* rec_x = nullptr;  // armed recording of the next step's network input
  double* rec_start = nullptr;
  int nprev = 0;
  // names the current contents of kapp for the solver's cached multigrid
  // hierarchy: bumped whenever a field may have been changed from outside
  uint64_t kap_version = 1;
  NnNet* ws_net = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // step timer, created once
  ~SbmState() {
    DevBuf* bs[] = {&tab, &psp, &pse, &cnp, &cne, &php, &phe, &kapp, &pce, &pcl, &avp, &rxn, &rxc, &rhsp,
                    &rhse, &rhsc, &rhe, &mu, &pm, &dab01, &scal, &php_prev, &phe_prev, &nn_x, &ws_cand, &ws_dir, &ws_dir_e, &ws_r0, &outer_dp,
                    &outer_de, &ws_dir2, &ws_dir3};
    for (DevBuf* b : bs) dev_free(*b);
    cg.release();
    if (hs) (void)hipHostFree(hs);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
  double* g(DevBuf& b) { return (double*)b.p; }
  SbmScal* S() { return (SbmScal*)scal.p; }
};

namespace {
std::mutex g_sbm_mu;
std::map<int64_t, std::unique_ptr<SbmState>> g_sbm;
int64_t g_next = 1;

SbmState* lookup(const int64_t* h) {
  if (!h) return nullptr;
  std::lock_guard<std::mutex> lk(g_sbm_mu);
  auto it = g_sbm.find(*h);
  if (it == g_sbm.end()) {
    set_error("unknown SBM handle %lld", (long long)*h);
    return nullptr;
  }
  return it->second.get();
}

int launch_bc(const Dims& d, double* f, const BcArgs& b) {
  const int64_t tot = ghost_cells(d);
  if (tot >= ((int64_t)1 << 31)) {
    set_error("boundary faces of %d x %d x %d exceed 2^31 cells", d.n1, d.n2, d.n3);
    return ELSIM_ERR_ARG;
  }
  k_bc<<<(unsigned)((tot + 255) / 256), 256, 0, stream()>>>(d, f, b);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}
int bc(SbmState& s, double* f, const BcArgs& b) { return launch_bc(s.d, f, b); }
BcArgs mirror_bc() { return BcArgs{0, 0, 0.0, 0.0, nullptr, nullptr, nullptr, 0.0, 0.0}; }

int sync_scal(SbmState& s) {
  ELSIM_HIP(hipMemcpyAsync(s.hs, s.scal.p, sizeof(SbmScal), hipMemcpyDeviceToHost, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  return ELSIM_OK;
}

int field_info(SbmState& s, int f, double** p, size_t* n) {
  const size_t G = (size_t)s.d.G(), N = (size_t)s.d.N();
  switch (f) {
    case ELSIM_F_PSP: *p = s.g(s.psp); *n = G; return 0;
    case ELSIM_F_PSE: *p = s.g(s.pse); *n = G; return 0;
    case ELSIM_F_AVP: *p = s.g(s.avp); *n = N; return 0;
    case ELSIM_F_CNP: *p = s.g(s.cnp); *n = G; return 0;
    case ELSIM_F_CNE: *p = s.g(s.cne); *n = G; return 0;
    case ELSIM_F_PHP: *p = s.g(s.php); *n = G; return 0;
    case ELSIM_F_PHE: *p = s.g(s.phe); *n = G; return 0;
    case ELSIM_F_RXN: *p = s.g(s.rxn); *n = N; return 0;
    case ELSIM_F_KAPP: *p = s.g(s.kapp); *n = G; return 0;
  }
  set_error("unknown field id %d", f);
  return ELSIM_ERR_ARG;
}

// A named host range over a phase of the step, for rocprofv3 --marker-trace
// (the solves poll the device, so a range spans its phase's GPU work).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

// One CG solve of the state's current system.
int solve(SbmState& s, int op, int mode, SolveResult* r) {
  const Range range(op == ELSIM_OP_CNE ? "cne_solve" : (op == ELSIM_OP_PHP ? "php_solve" : "phe_solve"));
  const elsim_sbm_params& P = s.P;
  SolveArgs a{};
  a.d = s.d;
  a.op = op;
  a.stop_rule = P.stop_rule;
  a.rtol = P.rtol;
  a.delta = P.inner_rtol;
  // the mixed solvers keep x fixed until their correction is applied, so
  // the convergence measures need no copy of the start (k_red_cvg)
  a.keep_x0 = false;
  if (op == ELSIM_OP_CNE) {
    a.x = s.g(s.cne);
    a.kap = s.g(s.pce);
    a.pse = s.g(s.pse);
    a.rhs = s.g(s.rhsc);
    a.scale = P.dt / (P.dx * P.dx);
    a.crt = P.crt_e;
    a.maxit = P.max_it_e;
    a.w8 = s.g(s.pse);
    a.tw = s.hs->tpe;
    a.rel_r0 = P.cne_rule == 1;
  } else if (op == ELSIM_OP_PHP) {
    a.defer_add = true;  // applied with the convergence measure
    a.x = s.g(s.php);
    a.kap = s.g(s.kapp);
    a.pse = s.g(s.pse);
    a.rhs = s.g(s.rhsp);
    a.scale = 1.0 / (P.dx * P.dx);
    a.coef_id = s.kap_version;  // kapp = kap_scale psi_P, fixed over the run
    a.crt = P.crt_p;
    a.maxit = P.max_it_p;
    a.w8 = s.g(s.psp);
    a.tw = s.hs->tpp;
  } else {
    a.defer_add = true;
    a.x = s.g(s.phe);
    a.kap = s.g(s.pcl);
    a.pse = s.g(s.pse);
    a.rhs = s.g(s.rhse);
    a.scale = 1.0 / (P.dx * P.dx);
    a.crt = P.crt_l;
    a.maxit = P.max_it_l;
    a.w8 = s.g(s.pse);
    a.tw = s.hs->tpe;
  }
  return cg_solve(s.cg, a, mode, r);
}

int fin_sbm(SbmState& s, int op) {
  k_fin_sbm<<<1, FIN_T, 0, stream()>>>(op, blocks_of(col_grid(s.d)), workspace().partials, s.S(), s.P, s.d);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int red_gw(SbmState& s, const double* a, const double* w, int op) {
  Workspace& ws = workspace();
  k_red_gw<<<col_grid(s.d), dim3(TX, TY, 1), 0, stream()>>>(s.d, a, w, op, s.S(), s.P, ws.partials);
  ELSIM_CHECK_LAUNCH();
  return fin_sbm(s, op);
}

int ch_update(SbmState& s) {
  const Dims& d = s.d;
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  Workspace& ws = workspace();
  hipStream_t st = stream();
  k_ch_mu<<<grid, blk, 0, st>>>(d, s.g(s.cnp), s.g(s.psp), s.g(s.tab), s.sz, s.P.eps, s.P.dx, s.g(s.mu), s.g(s.pm));
  k_ch_update<<<grid, blk, 0, st>>>(d, s.g(s.cnp), s.g(s.mu), s.g(s.pm), s.g(s.psp), s.g(s.rxn), s.g(s.avp), s.P,
                                    s.S(), ws.partials);
  ELSIM_CHECK_LAUNCH();
  ELSIM_TRY(fin_sbm(s, R_XRF));
  return bc(s, s.g(s.cnp), mirror_bc());
}

int coef(SbmState& s, double sd0, int times_c, double* out, double* dab01) {
  const int64_t G = s.d.G();
  k_coef<<<flat_blocks(G), FLAT_T, 0, stream()>>>(s.d, s.g(s.cne), s.g(s.pse), sd0, s.P, times_c, out, dab01);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

BcArgs cne_bc_args(SbmState& s) {
  BcArgs b = mirror_bc();
  b.lo = 2;
  b.dab01 = s.g(s.dab01);
  b.infx = &s.S()->infx;
  b.tm = s.P.t_minus;
  b.dx = s.P.dx;
  return b;
}

double host_sum(const double* partials, int nb) {
  std::vector<double> h(nb);
  if (hipMemcpyAsync(h.data(), partials, nb * sizeof(double), hipMemcpyDeviceToHost, stream()) != hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess)
    return NAN;
  double t = 0.0;
  for (double v : h) t += v;
  return t;
}

// The network input at the first phP solve: the start's residual r0
// (interior, rms rms_r) and the last change of phi_P (ws_dir, ghosted).
// full: rows of all 16 channels (a recorded sample); else rows of channels
// 0..7, the network's input with channels 8..15 zero.
int ws_features(SbmState& s, const double* r0, double rms_r, uint16_t* x, bool full) {
  const Dims& d = s.d;
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  k_sum_sq_g<<<grid, blk, 0, stream()>>>(d, s.g(s.ws_dir), workspace().partials);
  ELSIM_CHECK_LAUNCH();
  const double sd = host_sum(workspace().partials, blocks_of(grid));
  const double rms_d = std::sqrt(sd / (double)d.N());
  k_feat_pack<<<grid, blk, 0, stream()>>>(d, s.g(s.psp), r0, s.g(s.kapp), s.g(s.ws_dir),
                                          rms_r > 0 ? 1.0 / rms_r : 0.0, rms_d > 0 ? 1.0 / rms_d : 0.0, full ? 1 : 0,
                                          x);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

// x(interior) += alpha dv with alpha = dv.(b - A x) / dv.(-A_hom) dv, the
// A-norm-optimal step along dv (A = -L is SPD on corrections with the search
// direction's homogeneous ghosts); ghosts re-applied by `bcx`.  *alpha_out
// = 0 when dv carries no information.
int galerkin_step(SbmState& s, int op, double* x, const double* kap, const double* rhs, const double* dv,
                  const BcArgs& bcx, double* alpha_out) {
  const Dims& d = s.d;
  const double sc = 1.0 / (s.P.dx * s.P.dx);
  double num = 0.0, den = 0.0;
  ELSIM_TRY(launch_galerkin_dots(op, d, x, kap, sc, rhs, dv, &num, &den));
  const double alpha = (den < 0.0 && std::isfinite(num / den)) ? num / den : 0.0;
  if (alpha != 0.0) {
    k_ws_axpy<<<col_grid(d), dim3(TX, TY, 1), 0, stream()>>>(d, x, dv, alpha);
    ELSIM_CHECK_LAUNCH();
    ELSIM_TRY(bc(s, x, bcx));
  }
  *alpha_out = alpha;
  return ELSIM_OK;
}

int ensure_ws_buffers(SbmState& s) {
  const size_t N = (size_t)s.d.N(), G = (size_t)s.d.G();
  ELSIM_TRY(dev_alloc(s.nn_x, N * NN_IN_CH * sizeof(uint16_t)));
  ELSIM_TRY(dev_alloc(s.ws_cand, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.ws_dir, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.ws_dir2, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.ws_dir3, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.ws_dir_e, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.ws_r0, N * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.outer_dp, G * sizeof(double)));
  ELSIM_TRY(dev_alloc(s.outer_de, G * sizeof(double)));
  return ELSIM_OK;
}

int step_impl(SbmState& s, int mode, elsim_step_stats* st) {
  const Dims& d = s.d;
  const elsim_sbm_params& P = s.P;
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  Workspace& ws = workspace();
  hipStream_t stv = stream();
  elsim_step_stats z{};
  int status = 0;
  SolveResult r;
  if (!s.ev0) ELSIM_HIP(hipEventCreate(&s.ev0));
  if (!s.ev1) ELSIM_HIP(hipEventCreate(&s.ev1));
  ELSIM_HIP(hipEventRecord(s.ev0, stv));

  // 0. warm start of the coupling loop: keep the potentials' change over
  //    the last step (from the third step on: the first starts from the
  //    initial fields, the second's predecessor change is that start-up
  //    transient)
  const bool ws_learned = P.ws_mode == ELSIM_WS_LEARNED && (s.ws_net != nullptr || s.rec_x != nullptr);
  const bool ws_on = (P.ws_mode == ELSIM_WS_EXTRAP || ws_learned) && s.nprev >= 2;
  int ws_hist = 0;  // earlier changes of phi_P known besides the last one
  if (ws_on) {
    ELSIM_TRY(ensure_ws_buffers(s));
    std::swap(s.ws_dir3, s.ws_dir2);
    std::swap(s.ws_dir2, s.ws_dir);
    k_ws_diff_keep<<<flat_blocks(d.G()), FLAT_T, 0, stv>>>(d.G(), s.g(s.php), s.g(s.php_prev), s.g(s.ws_dir));
    k_ws_diff_keep<<<flat_blocks(d.G()), FLAT_T, 0, stv>>>(d.G(), s.g(s.phe), s.g(s.phe_prev), s.g(s.ws_dir_e));
    ELSIM_CHECK_LAUNCH();
    ws_hist = s.ws_p_hist;
  } else {
    ELSIM_HIP(hipMemcpyAsync(s.php_prev.p, s.php.p, (size_t)d.G() * sizeof(double), hipMemcpyDeviceToDevice, stv));
    ELSIM_HIP(hipMemcpyAsync(s.phe_prev.p, s.phe.p, (size_t)d.G() * sizeof(double), hipMemcpyDeviceToDevice, stv));
  }
  s.ws_p_hist = ws_on ? std::min(ws_hist + 1, 2) : 0;
  s.nprev++;

  // 1. Cahn-Hilliard (SBM:397-443)
  ELSIM_TRY(ch_update(s));

  // 2. CnE (SBM:461-572)
  k_cne_rhs<<<grid, blk, 0, stv>>>(d, s.g(s.cne), s.g(s.pse), s.g(s.rxn), s.g(s.avp), P, s.g(s.rhsc), s.S(),
                                   ws.partials);
  ELSIM_CHECK_LAUNCH();
  ELSIM_TRY(fin_sbm(s, R_INFX));
  ELSIM_TRY(coef(s, P.d0, 0, s.g(s.pce), s.g(s.dab01)));
  // the host mirror of the scalars is current from the last step's end; the
  // first step reads the totals of psi_P and psi_E (the reference stop
  // rule's weights) here
  if (s.iter <= 1) ELSIM_TRY(sync_scal(s));
  if (!(P.skip_cne_first && s.iter == 1)) {
    ELSIM_TRY(solve(s, ELSIM_OP_CNE, mode, &r));
    z.it_cne = r.iters;
    z.relres_cne = r.relres;
    z.n_reliable += r.n_reliable;
    if (r.status) status = r.status;
  }
  ELSIM_TRY(bc(s, s.g(s.cne), cne_bc_args(s)));
  if (P.mass_corr_period > 0 && s.iter % P.mass_corr_period == 1) {
    ELSIM_TRY(red_gw(s, s.g(s.pse), s.g(s.cne), R_AVCNE));
    k_shift<<<grid, blk, 0, stv>>>(d, s.g(s.cne), s.S(), 0, P.cne_target, P.dt);
    ELSIM_CHECK_LAUNCH();
    ELSIM_TRY(bc(s, s.g(s.cne), cne_bc_args(s)));
  }

  // 3. potential coefficients (SBM:579-622)
  k_coef_pair<<<(unsigned)((d.G() + 255) / 256), 256, 0, stv>>>(d.G(), s.g(s.cne), s.g(s.pse), P.tc1 * P.d0,
                                                                  P.tc2 * P.d0, P, s.g(s.pce), s.g(s.pcl));
  ELSIM_CHECK_LAUNCH();
  ELSIM_TRY(launch_matvec_f64(ELSIM_OP_PHP, d, s.g(s.cne), s.g(s.pce), nullptr, 1.0 / (P.dx * P.dx), s.g(s.rhe),
                              nullptr, false));

  // 4. warm start: both potentials extrapolated linearly in time before the
  //    first Butler-Volmer evaluation, so the coupling loop starts near the
  //    step's solution (fewer outer iterations, shorter solves)
  const double sc = 1.0 / (P.dx * P.dx);
  BcArgs pbc0 = mirror_bc();
  pbc0.hi = 1;
  pbc0.hi_v = P.bvp;
  auto butler_volmer = [&](bool extrapolate) -> int {
    const Range range("butler_volmer");
    if (extrapolate)
      k_bv<true><<<grid, blk, 0, stv>>>(d, s.g(s.php), s.g(s.phe), s.g(s.cne), s.g(s.cnp), s.g(s.tab), s.sz,
                                        s.g(s.avp), s.g(s.rhe), P, s.g(s.rxn), s.g(s.rxc), s.g(s.rhsp),
                                        s.g(s.rhse), s.g(s.ws_dir), s.g(s.ws_dir_e), s.hs->shift);
    else
      k_bv<false><<<grid, blk, 0, stv>>>(d, s.g(s.php), s.g(s.phe), s.g(s.cne), s.g(s.cnp), s.g(s.tab), s.sz,
                                         s.g(s.avp), s.g(s.rhe), P, s.g(s.rxn), s.g(s.rxc), s.g(s.rhsp),
                                         s.g(s.rhse), nullptr, nullptr, 0.0);
    ELSIM_CHECK_LAUNCH();
    return ELSIM_OK;
  };
  // the first Butler-Volmer evaluation of the coupling loop, done here when
  // it also applies the extrapolation
  bool bv_done = false;
  if (ws_on) {
    // phi_P += its last change; phi_E += only the solves' part of its last
    // change.  The rest of that change is the current control's uniform
    // shift (SBM:831), which the previous step applied at its end (its dv
    // dt, kept in the scalars with that step's own dt) and which this step
    // applies again with its own sign.  The reaction reads interior values
    // only, so the ghosts follow.
    if (P.max_outer > 0 && 1.0 > P.outer_tol) {  // the coupling loop runs at least once
      ELSIM_TRY(butler_volmer(true));
      bv_done = true;
    } else {
      k_ws_axpy<<<grid, blk, 0, stv>>>(d, s.g(s.php), s.g(s.ws_dir), 1.0);
      k_ws_axpy_shift<<<grid, blk, 0, stv>>>(d, s.g(s.phe), s.g(s.ws_dir_e), s.hs->shift);
      ELSIM_CHECK_LAUNCH();
    }
    ELSIM_TRY(bc(s, s.g(s.php), pbc0));
    BcArgs ebc0 = mirror_bc();
    ebc0.lo = 1;
    ebc0.lo_vp = &s.S()->bve;
    ELSIM_TRY(bc(s, s.g(s.phe), ebc0));
    z.ws_used = 1;
  }

  // 5. outer loop (SBM:624-812)
  double cvgp = 1.0, cvgl = 1.0;
  int nlp = 0;
  while (cvgp > P.outer_tol && cvgl > P.outer_tol) {
    if (nlp >= P.max_outer) {
      if (!status) status = ELSIM_NOT_CONVERGED;
      break;
    }
    if (!bv_done) ELSIM_TRY(butler_volmer(false));
    bv_done = false;
    bool moved = false;  // a start correction below moved phi_P or phi_E
    {
      // SBM:647: phP(:, nx+1, :) = BvP only
      BcArgs only_hi = mirror_bc();
      only_hi.hi = 1;
      only_hi.hi_v = P.bvp;
      // the mirror ghosts of phP are already exact (last BC application),
      // so re-mirroring them changes no value
      ELSIM_TRY(bc(s, s.g(s.php), only_hi));
    }
    if (nlp == 0 && ws_on && ws_hist >= 1 && P.ws_order != 1) {
      // second-order term: a Galerkin step along the second difference
      // d1 - d2 of phi_P's last two changes (step 1 is quadratic
      // extrapolation, 0 keeps the linear one), A-norm optimal, so never
      // worse than the start it corrects.  Before the learned correction,
      // which then sees the residual that this step leaves
      k_diff_interior<<<grid, blk, 0, stv>>>(d, s.g(s.ws_dir), s.g(s.ws_dir2), s.g(s.ws_cand));
      ELSIM_CHECK_LAUNCH();
      double a2 = 0.0;
      ELSIM_TRY(galerkin_step(s, ELSIM_OP_PHP, s.g(s.php), s.g(s.kapp), s.g(s.rhsp), s.g(s.ws_cand), pbc0, &a2));
      moved = moved || a2 != 0.0;
      if (ws_hist >= 2 && P.ws_order != 2) {
        // third-order term, likewise: the third difference d1 - 2 d2 + d3
        k_diff3_interior<<<grid, blk, 0, stv>>>(d, s.g(s.ws_dir), s.g(s.ws_dir2), s.g(s.ws_dir3), s.g(s.ws_cand));
        ELSIM_CHECK_LAUNCH();
        double a3 = 0.0;
        ELSIM_TRY(galerkin_step(s, ELSIM_OP_PHP, s.g(s.php), s.g(s.kapp), s.g(s.rhsp), s.g(s.ws_cand), pbc0, &a3));
        moved = moved || a3 != 0.0;
      }
    }
    if (nlp == 0 && ws_learned && ws_on) {
      // learned correction of the extrapolated phi_P: the network maps
      // (psi_P, r0 = b - A x, the coefficient, the last change) to the shape
      // p of the remaining error; the Galerkin step alpha = p.r0 / p.(-A) p
      // along it minimises the A-norm error (A = -L is SPD on corrections
      // with the search direction's homogeneous ghosts), so it never makes
      // the start worse in the norm CG minimises.
      double* r0 = s.g(s.ws_r0);
      double rr = 0.0;
      ELSIM_TRY(launch_resid_store(ELSIM_OP_PHP, d, s.g(s.php), s.g(s.kapp), nullptr, sc, s.g(s.rhsp), r0, &rr));
      const double rms_r = std::sqrt(rr / (double)d.N());
      z.ws_res_plain = std::sqrt(rr);
      z.ws_res_nn = z.ws_res_plain;
      uint16_t* xin = s.rec_x ? s.rec_x : (uint16_t*)s.nn_x.p;
      ELSIM_TRY(ws_features(s, r0, rms_r, xin, s.rec_x != nullptr));
      if (s.rec_start) {
        k_interior<<<grid, blk, 0, stv>>>(d, s.g(s.php), s.rec_start);
        ELSIM_CHECK_LAUNCH();
      }
      if (s.ws_net && rms_r > 0.0) {
        {
          const Range range("ws_network");
          // the correction direction, written by the network's last kernel
          // as FP64 into the interior; its ghosts are never read (the
          // Galerkin step's operator on it uses the homogeneous rule, and
          // the update is interior)
          ELSIM_TRY(nn_forward_g(s.ws_net, d.n1, d.n2, d.n3, xin, s.g(s.ws_cand), s.rec_x ? 16 : 8));
        }
        double* pd = s.g(s.ws_cand);
        double num = 0.0, den = 0.0;
        ELSIM_TRY(launch_galerkin_dots(ELSIM_OP_PHP, d, s.g(s.php), s.g(s.kapp), sc, s.g(s.rhsp), pd, &num, &den));
        const double alpha = (den < 0.0 && std::isfinite(num / den)) ? num / den : 0.0;  // den = p.L p < 0
        if (alpha != 0.0) {
          k_ws_axpy<<<grid, blk, 0, stv>>>(d, s.g(s.php), pd, alpha);
          ELSIM_CHECK_LAUNCH();
          ELSIM_TRY(bc(s, s.g(s.php), pbc0));
          z.ws_used = 2;
          z.ws_alpha = alpha;
          z.ws_energy_gain = num * num / -den;  // decrease of the squared A-norm error
          moved = true;
        }
      }
      s.rec_x = nullptr;
      s.rec_start = nullptr;
    }
    BcArgs ebc = mirror_bc();
    ebc.lo = 1;
    ebc.lo_vp = &s.S()->bve;
    const bool outer_acc = P.ws_mode != ELSIM_WS_PLAIN && nlp >= 1;
    if (outer_acc) {
      // later coupling iterations: Galerkin steps along the last phP and
      // phE corrections (successive corrections of the linearly converging
      // coupling loop are nearly parallel)
      double a_p = 0.0, a_e = 0.0;
      ELSIM_TRY(galerkin_step(s, ELSIM_OP_PHP, s.g(s.php), s.g(s.kapp), s.g(s.rhsp), s.g(s.outer_dp), pbc0, &a_p));
      ELSIM_TRY(galerkin_step(s, ELSIM_OP_PHE, s.g(s.phe), s.g(s.pcl), s.g(s.rhse), s.g(s.outer_de), ebc, &a_e));
      moved = moved || a_p != 0.0 || a_e != 0.0;
    }
    // The reference evaluates Butler-Volmer at the fields each coupling
    // iteration starts from and measures CvgP / CvgL from them (ppO, peO,
    // SBM:631-644, 726, 721, 802).  When a start correction moved a field
    // after that evaluation, the reaction is evaluated again at the corrected
    // start, so that both solves start where the reaction was evaluated and
    // CvgP / CvgL remain the coupling iteration's fixed-point residual.
    if (moved) ELSIM_TRY(butler_volmer(false));
    ELSIM_TRY(solve(s, ELSIM_OP_PHP, mode, &r));
    // the residual the corrected start leaves, at the reaction re-evaluated
    // there (the first phP solve's start)
    if (nlp == 0 && z.ws_used == 2) z.ws_res_nn = r.res0;
    z.it_php += r.iters;
    z.relres_php = r.relres;
    z.n_reliable += r.n_reliable;
    if (r.status) status = r.status;
    const bool keep_dx = P.ws_mode != ELSIM_WS_PLAIN;  // the corrections, for the next iteration's steps
    if (keep_dx) ELSIM_TRY(ensure_ws_buffers(s));
    // CvgP (SBM:723) and the change, in the pass that applies the solve's
    // correction when the solve left it; the ghosts follow from the interior
    // without a stored start and without a pending correction the solve
    // left x as it came in (its start met the tolerance): x is its own start
    const double* e64 = (const double*)s.cg.e64.p;
    auto start_of = [&](const SolveResult& q, const double* x) {
      return q.x0_stored ? (const double*)s.cg.x0.p : x;
    };
    k_red_cvg<<<grid, blk, 0, stv>>>(d, s.g(s.php), start_of(r, s.g(s.php)), s.g(s.psp), R_CVGP, s.S(), P,
                                     ws.partials, keep_dx ? s.g(s.outer_dp) : nullptr, r.e_pending ? e64 : nullptr);
    ELSIM_CHECK_LAUNCH();
    ELSIM_TRY(fin_sbm(s, R_CVGP));
    BcArgs pbc = mirror_bc();
    pbc.hi = 1;
    pbc.hi_v = P.bvp;
    ELSIM_TRY(bc(s, s.g(s.php), pbc));
    ELSIM_TRY(solve(s, ELSIM_OP_PHE, mode, &r));
    z.it_phe += r.iters;
    z.relres_phe = r.relres;
    z.n_reliable += r.n_reliable;
    if (r.status) status = r.status;
    k_red_cvg<<<grid, blk, 0, stv>>>(d, s.g(s.phe), start_of(r, s.g(s.phe)), s.g(s.pse), R_CVGL, s.S(), P,
                                     ws.partials, keep_dx ? s.g(s.outer_de) : nullptr, r.e_pending ? e64 : nullptr);
    ELSIM_CHECK_LAUNCH();
    ELSIM_TRY(fin_sbm(s, R_CVGL));
    ELSIM_TRY(bc(s, s.g(s.phe), ebc));
    ELSIM_TRY(sync_scal(s));
    cvgp = s.hs->cvgp;
    cvgl = s.hs->cvgl;
    nlp++;
  }
  z.n_outer = nlp;
  z.cvgp = cvgp;
  z.cvgl = cvgl;

  // 6. current control (SBM:826-835)
  k_red_rxn<<<grid, blk, 0, stv>>>(d, s.g(s.rxn), s.g(s.avp), R_SCRNT, s.S(), P, ws.partials);
  ELSIM_TRY(fin_sbm(s, R_SCRNT));
  k_shift<<<grid, blk, 0, stv>>>(d, s.g(s.phe), s.S(), 1, 0.0, P.dt);
  ELSIM_CHECK_LAUNCH();
  BcArgs ebc = mirror_bc();
  ebc.lo = 1;
  ebc.lo_vp = &s.S()->bve;
  ELSIM_TRY(bc(s, s.g(s.phe), ebc));
  ELSIM_HIP(hipEventRecord(s.ev1, stv));
  ELSIM_TRY(sync_scal(s));
  float ms = 0.f;
  ELSIM_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
  z.xrf = s.hs->xrf;
  z.scrnt = s.hs->scrnt;
  z.status = status;
  z.seconds = ms * 1e-3;
  s.iter++;
  if (st) *st = z;
  return status;
}

// A step that fails part-way has already recorded its start as the
// previous fields and counted itself: the warm-start history is then
// inconsistent, so it is dropped and the next step starts plain.
int step(SbmState& s, int mode, elsim_step_stats* st) {
  const Range range("sbm_step");
  const int rc = step_impl(s, mode, st);
  if (rc < 0) {
    s.nprev = 0;
    s.ws_p_hist = 0;
  }
  return rc;
}

}  // namespace

namespace {
CgWork g_pot_work;  // work space of the standalone potential solves (grows with the grid)
}  // namespace

int pot_args_check(int n1, int n2, int n3, double dx, const void* kap, const void* rhs, const void* phi, int bc_lo,
                int bc_hi, int mode, double rtol, int maxit) {
  if (!dims_ok(n1, n2, n3)) {
    set_error("bad grid %d x %d x %d", n1, n2, n3);
    return ELSIM_ERR_ARG;
  }
  if (!kap || !rhs || !phi) {
    set_error("null array argument");
    return ELSIM_ERR_ARG;
  }
  if ((bc_lo != 0 && bc_lo != 1) || (bc_hi != 0 && bc_hi != 1) || (bc_lo == 0 && bc_hi == 0)) {
    set_error("d2 face conditions (%d, %d): each 0 (Neumann) or 1 (Dirichlet), at least one Dirichlet", bc_lo,
              bc_hi);
    return ELSIM_ERR_ARG;
  }
  if (mode != ELSIM_MODE_FP64 && mode != ELSIM_MODE_MIXED && mode != ELSIM_MODE_FP64_PCG &&
      mode != ELSIM_MODE_MIXED_MG) {
    set_error("unknown mode %d", mode);
    return ELSIM_ERR_ARG;
  }
  if (!(rtol > 0.0) || maxit < 1 || !(dx > 0.0)) {
    set_error("rtol %g, maxit %d, dx %g: need rtol > 0, maxit >= 1, dx > 0", rtol, maxit, dx);
    return ELSIM_ERR_ARG;
  }
  return ELSIM_OK;
}

namespace {
// Ghost planes of the unknown: mirrors on the d1 and d3 faces and on a
// Neumann d2 face; a Dirichlet d2 plane keeps the value the caller stored.
BcArgs pot_bc(int bc_lo, int bc_hi) {
  BcArgs b = mirror_bc();
  b.lo = bc_lo ? 3 : 0;
  b.hi = bc_hi ? 3 : 0;
  return b;
}

}  // namespace

int solve_pot_with(CgWork& w, int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi,
                   int bc_lo, int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres) {
  const Dims d{n1, n2, n3};
  const int op = bc_lo && bc_hi ? ELSIM_OP_POT_DD : (bc_hi ? ELSIM_OP_PHP : ELSIM_OP_PHE);
  ELSIM_TRY(launch_bc(d, phi, pot_bc(bc_lo, bc_hi)));
  SolveArgs a{};
  a.d = d;
  a.op = op;
  a.x = phi;
  a.kap = kap;
  a.pse = nullptr;
  a.rhs = rhs;
  a.scale = 1.0 / (dx * dx);
  a.stop_rule = ELSIM_STOP_RELRES;
  a.rtol = rtol;
  a.crt = 0.0;
  a.maxit = maxit;
  a.w8 = kap;
  a.tw = 1.0;
  a.delta = 0.5;
  SolveResult r;
  ELSIM_TRY(cg_solve(w, a, mode, &r));
  ELSIM_TRY(launch_bc(d, phi, pot_bc(bc_lo, bc_hi)));  // SBM:716-719
  ELSIM_HIP(hipStreamSynchronize(stream()));
  if (iters) *iters = r.iters;
  if (relres) *relres = r.relres;
  return r.status;
}

namespace {
int solve_pot_dev(int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi, int bc_lo,
                  int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres) {
  return solve_pot_with(g_pot_work, n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit, iters, relres);
}
}  // namespace

void sbm_forget_net(NnNet* net) {
  std::lock_guard<std::mutex> lk(g_sbm_mu);
  for (auto& kv : g_sbm)
    if (kv.second->ws_net == net) kv.second->ws_net = nullptr;
}

}  // namespace elsim

using namespace elsim;

extern "C" {

int elsim_sbm_create(const int* n1, const int* n2, const int* n3, const elsim_sbm_params* prm, const float* sdf_s,
                     const float* sdf_l, const int* sz, const double* tables, int64_t* handle) {
  if (!n1 || !n2 || !n3 || !prm || !sdf_s || !sdf_l || !sz || !tables || !handle) return ELSIM_ERR_ARG;
  if (!dims_ok(*n1, *n2, *n3) || *sz < 3) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  auto s = std::make_unique<SbmState>();
  s->d = Dims{*n1, *n2, *n3};
  s->P = *prm;
  s->sz = *sz;
  const Dims& d = s->d;
  if (blocks_of(col_grid(d)) > kMaxRedBlocks) {
    set_error("grid too large for the reduction workspace");
    return ELSIM_ERR_ARG;
  }
  const size_t G = (size_t)d.G() * 8, N = (size_t)d.N() * 8;
  ELSIM_TRY(dev_alloc(s->tab, 8 * (size_t)*sz * 8));
  for (DevBuf* b : {&s->psp, &s->pse, &s->cnp, &s->cne, &s->php, &s->phe, &s->kapp, &s->pce, &s->pcl, &s->php_prev,
                    &s->phe_prev})
    ELSIM_TRY(dev_alloc(*b, G));
  for (DevBuf* b : {&s->avp, &s->rxn, &s->rxc, &s->rhsp, &s->rhse, &s->rhsc, &s->rhe, &s->mu, &s->pm})
    ELSIM_TRY(dev_alloc(*b, N));
  ELSIM_TRY(dev_alloc(s->dab01, 2 * (size_t)(d.n1 + 2) * (d.n3 + 2) * 8));
  ELSIM_TRY(dev_alloc(s->scal, sizeof(SbmScal)));
  ELSIM_HIP(hipHostMalloc((void**)&s->hs, sizeof(SbmScal), hipHostMallocDefault));
  hipStream_t st = stream();
  ELSIM_HIP(hipMemcpyAsync(s->tab.p, tables, 8 * (size_t)*sz * 8, hipMemcpyHostToDevice, st));
  ELSIM_HIP(hipMemsetAsync(s->rxn.p, 0, N, st));
  ELSIM_HIP(hipMemsetAsync(s->rxc.p, 0, N, st));
  ELSIM_HIP(hipMemsetAsync(s->scal.p, 0, sizeof(SbmScal), st));
  // psi fields (SBM:149-207, 247, 258)
  float* dsdf = (float*)scratch(0, (size_t)d.N() * 4);
  if (!dsdf) return ELSIM_ERR_ALLOC;
  const dim3 grid = col_grid(d), blk(TX, TY, 1);
  ELSIM_HIP(hipMemcpyAsync(dsdf, sdf_s, (size_t)d.N() * 4, hipMemcpyHostToDevice, st));
  k_psi<<<grid, blk, 0, st>>>(d, dsdf, prm->zeta, prm->psp_off, s->g(s->psp));
  ELSIM_TRY(bc(*s, s->g(s->psp), mirror_bc()));
  ELSIM_HIP(hipMemcpyAsync(dsdf, sdf_l, (size_t)d.N() * 4, hipMemcpyHostToDevice, st));
  k_psi<<<grid, blk, 0, st>>>(d, dsdf, prm->zeta, prm->pse_off, s->g(s->pse));
  ELSIM_TRY(bc(*s, s->g(s->pse), mirror_bc()));
  ELSIM_CHECK_LAUNCH();
  ELSIM_TRY(red_gw(*s, s->g(s->psp), nullptr, R_TPP));
  ELSIM_TRY(red_gw(*s, s->g(s->pse), nullptr, R_TPE));
  k_avp<<<grid, blk, 0, st>>>(d, s->g(s->psp), s->g(s->pse), prm->avp_min, s->g(s->avp));
  k_scale_g<<<flat_blocks(d.G()), FLAT_T, 0, st>>>(d.G(), s->g(s->psp), prm->kap_scale, s->g(s->kapp));
  ELSIM_CHECK_LAUNCH();
  // initial fields (SBM:297-337)
  std::vector<double> h((size_t)d.G());
  auto fill = [&](DevBuf& b, double v) -> int {
    for (auto& x : h) x = v;
    ELSIM_HIP(hipMemcpyAsync(b.p, h.data(), G, hipMemcpyHostToDevice, st));
    ELSIM_HIP(hipStreamSynchronize(st));
    return ELSIM_OK;
  };
  ELSIM_TRY(fill(s->cnp, 2.02e-2));
  ELSIM_TRY(fill(s->cne, prm->cne_target));
  ELSIM_TRY(fill(s->php, prm->bvp));
  ELSIM_TRY(fill(s->phe, prm->bve));
  ELSIM_TRY(red_gw(*s, s->g(s->cnp), s->g(s->psp), R_XRF));
  ELSIM_TRY(sync_scal(*s));
  s->hs->bve = prm->bve;
  ELSIM_HIP(hipMemcpyAsync(s->scal.p, s->hs, sizeof(SbmScal), hipMemcpyHostToDevice, st));
  ELSIM_TRY(s->cg.alloc(d));
  ELSIM_HIP(hipStreamSynchronize(st));
  std::lock_guard<std::mutex> lk(g_sbm_mu);
  int64_t id = g_next++;
  g_sbm[id] = std::move(s);
  *handle = id;
  return ELSIM_OK;
}

int elsim_sbm_destroy(const int64_t* handle) {
  if (!handle) return ELSIM_ERR_ARG;
  (void)hipStreamSynchronize(stream());
  std::lock_guard<std::mutex> lk(g_sbm_mu);
  return g_sbm.erase(*handle) ? ELSIM_OK : ELSIM_ERR_ARG;
}

int elsim_sbm_get(const int64_t* handle, const int* field, double* out) {
  SbmState* s = lookup(handle);
  if (!s || !field || !out) return ELSIM_ERR_ARG;
  double* p;
  size_t n;
  ELSIM_TRY(field_info(*s, *field, &p, &n));
  ELSIM_HIP(hipMemcpyAsync(out, p, n * 8, hipMemcpyDeviceToHost, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  return ELSIM_OK;
}

int elsim_sbm_set(const int64_t* handle, const int* field, const double* in) {
  SbmState* s = lookup(handle);
  if (!s || !field || !in) return ELSIM_ERR_ARG;
  double* p;
  size_t n;
  ELSIM_TRY(field_info(*s, *field, &p, &n));
  ELSIM_HIP(hipMemcpyAsync(p, in, n * 8, hipMemcpyHostToDevice, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  s->kap_version++;
  if (*field == ELSIM_F_CNP) {
    ELSIM_TRY(red_gw(*s, s->g(s->cnp), s->g(s->psp), R_XRF));
    ELSIM_TRY(sync_scal(*s));
  }
  // a field set from outside (e.g. a restart) breaks the time history that
  // the warm start extrapolates: it starts over from the next step
  if (*field == ELSIM_F_CNP || *field == ELSIM_F_CNE || *field == ELSIM_F_PHP || *field == ELSIM_F_PHE) {
    s->nprev = 0;
    s->ws_p_hist = 0;
  }
  return ELSIM_OK;
}

int elsim_sbm_field_ptr(const int64_t* handle, const int* field, double** ptr) {
  SbmState* s = lookup(handle);
  if (!s || !field || !ptr) return ELSIM_ERR_ARG;
  size_t n;
  s->kap_version++;  // the caller may write through the pointer
  return field_info(*s, *field, ptr, &n);
}

int elsim_sbm_scalars(const int64_t* handle, double* tpp, double* tpe, double* xrf, double* bve, int* iter) {
  SbmState* s = lookup(handle);
  if (!s) return ELSIM_ERR_ARG;
  ELSIM_TRY(sync_scal(*s));
  if (tpp) *tpp = s->hs->tpp;
  if (tpe) *tpe = s->hs->tpe;
  if (xrf) *xrf = s->hs->xrf;
  if (bve) *bve = s->hs->bve;
  if (iter) *iter = s->iter;
  return ELSIM_OK;
}

int elsim_sbm_set_params(const int64_t* handle, const elsim_sbm_params* prm) {
  SbmState* s = lookup(handle);
  if (!s || !prm) return ELSIM_ERR_ARG;
  s->P = *prm;
  s->kap_version++;
  return ELSIM_OK;
}

int elsim_sbm_step(const int64_t* handle, const int* mode, elsim_step_stats* stats) {
  SbmState* s = lookup(handle);
  if (!s || !mode ||
      (*mode != ELSIM_MODE_FP64 && *mode != ELSIM_MODE_MIXED && *mode != ELSIM_MODE_FP64_PCG &&
       *mode != ELSIM_MODE_MIXED_MG))
    return ELSIM_ERR_ARG;
  return step(*s, *mode, stats);
}

int elsim_sbm_ch_update(const int64_t* handle) {
  SbmState* s = lookup(handle);
  if (!s) return ELSIM_ERR_ARG;
  ELSIM_TRY(ch_update(*s));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  return ELSIM_OK;
}

int elsim_sbm_set_warm_start(const int64_t* handle, const int64_t* nn) {
  SbmState* s = lookup(handle);
  if (!s) return ELSIM_ERR_ARG;
  NnNet* net = nullptr;
  if (nn && *nn) {
    net = nn_lookup(*nn);
    if (!net) {
      set_error("unknown network handle %lld", (long long)*nn);
      return ELSIM_ERR_ARG;
    }
  }
  if (net && (s->d.n1 % 8 || s->d.n2 % 8 || s->d.n3 % 8)) {
    set_error("the warm-start network needs grid dimensions that are multiples of 8 (%dx%dx%d)", s->d.n1, s->d.n2,
              s->d.n3);
    return ELSIM_ERR_ARG;
  }
  s->ws_net = net;
  return ELSIM_OK;
}

int elsim_sbm_ws_record(const int64_t* handle, uint16_t* x_dev, double* start_dev) {
  SbmState* s = lookup(handle);
  if (!s) return ELSIM_ERR_ARG;
  s->rec_x = x_dev;
  s->rec_start = x_dev ? start_dev : nullptr;
  return ELSIM_OK;
}

int elsim_sbm_matvec(const int64_t* handle, const int* op, const double* p, double* out) {
  SbmState* s = lookup(handle);
  if (!s || !op || !p || !out) return ELSIM_ERR_ARG;
  const elsim_sbm_params& P = s->P;
  if (*op == ELSIM_OP_CNE)
    return launch_matvec_f64(*op, s->d, p, s->g(s->pce), s->g(s->pse), P.dt / (P.dx * P.dx), out, nullptr, true);
  const double* kap = *op == ELSIM_OP_PHP ? s->g(s->kapp) : s->g(s->pcl);
  return launch_matvec_f64(*op, s->d, p, kap, nullptr, 1.0 / (P.dx * P.dx), out, nullptr, true);
}

int elsim_dev_solve_pot(int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi,
                        int bc_lo, int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres) {
  ELSIM_TRY(pot_args_check(n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit));
  ELSIM_TRY(ensure_init());
  return solve_pot_dev(n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit, iters, relres);
}

int elsim_solve_pot(const int* n1, const int* n2, const int* n3, const double* dx, const double* kap,
                    const double* rhs, double* phi, const int* bc_lo, const int* bc_hi, const int* mode,
                    const double* rtol, const int* maxit, int* iters, double* relres) {
  if (!n1 || !n2 || !n3 || !dx || !bc_lo || !bc_hi || !mode || !rtol || !maxit) {
    set_error("null scalar argument");
    return ELSIM_ERR_ARG;
  }
  ELSIM_TRY(pot_args_check(*n1, *n2, *n3, *dx, kap, rhs, phi, *bc_lo, *bc_hi, *mode, *rtol, *maxit));
  ELSIM_TRY(ensure_init());
  const Dims d{*n1, *n2, *n3};
  const size_t G = (size_t)d.G() * 8, N = (size_t)d.N() * 8;
  double* dk = (double*)scratch(5, G);
  double* dr = (double*)scratch(6, N);
  double* dp = (double*)scratch(7, G);
  if (!dk || !dr || !dp) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dk, kap, G, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(dr, rhs, N, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(dp, phi, G, hipMemcpyHostToDevice, s));
  const int rc = solve_pot_dev(*n1, *n2, *n3, *dx, dk, dr, dp, *bc_lo, *bc_hi, *mode, *rtol, *maxit, iters, relres);
  if (rc < 0) return rc;
  ELSIM_HIP(hipMemcpyAsync(phi, dp, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return rc;
}

int elsim_tbl_chk_pln_3da(const int* py, const int* px, const int* pz, const int* sz, const double* cord,
                          const double* vals, const double* cn, double* fn) {
  if (!py || !px || !pz || !sz || !cord || !vals || !cn || !fn || *sz < 2 || !dims_ok(*py, *px, *pz))
    return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const int64_t N = (int64_t)*py * *px * *pz;
  double* t = (double*)scratch(2, 2 * (size_t)*sz * 8);
  double* c = (double*)scratch(3, (size_t)N * 8);
  double* f = (double*)scratch(4, (size_t)N * 8);
  if (!t || !c || !f) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(t, cord, (size_t)*sz * 8, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(t + *sz, vals, (size_t)*sz * 8, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(c, cn, (size_t)N * 8, hipMemcpyHostToDevice, s));
  k_table_n<<<flat_blocks(N), FLAT_T, 0, s>>>(N, t, t + *sz, *sz, c, f);
  ELSIM_CHECK_LAUNCH();
  ELSIM_HIP(hipMemcpyAsync(fn, f, (size_t)N * 8, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

}  // extern "C"
