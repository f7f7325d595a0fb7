/*
 * elsim_oracle.c -- FP64 CPU restatement of the reference SBM solver core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (elsim_amd/, libelsim_hip.so) never links or calls it.
 *
 * It restates, in the reference's operation order and with its literal kinds
 * (REAL(4) literals promoted to the double the reference computes with), the
 * single-rank semantics of:
 *   SBM:143-291  psi / interface-density setup        (or_psi_from_sdf, or_avp)
 *   SBM:324-331, 473-484, 578-610  face coefficients  (cell-centred; face = 0.5(k_nb + k_c))
 *   SBM:492-499, 515-521, 591-596, 649-655, 670-676, 729-735, 751-757  operators
 *   SBM:501-555, 657-714, 737-795  CG loops           (or_cg)
 *   SBM:397-445  Cahn-Hilliard update                 (or_ch_update)
 *   SBM:556-572  CnE boundary + mass correction
 *   SBM:613-645  Butler-Volmer                        (or_bv)
 *   SBM:624-838  outer loop + current control         (or_sbm_step)
 *   SBM:987-1028 Tbl_Chk_Pln_3DA                      (or_table)
 *   SBM:1031-1090 BcMpiYZ boundary semantics on one rank (or_bc_yz: mirror)
 *   SBM:358-383  time-loop control: dt, Vsr            (or_time_control)
 *   RM:64-136    sign block of one RM phase            (or_rm_phase)
 *                the 2x refinement to the SBM grid (not in the reference) (or_refine2)
 *   LS:84-144    geometry front-end: reorient, tunnels, stacking, Allen-Cahn, sign
 *                                                     (or_ls_frontend)
 *   LS:164-322   NoGrdBc3D, ReDist3D_XYZ, AbsGrd3DUW_XYZ
 *   AD:83-243    accelerated-diffusion connectivity   (or_accdiff_phase)
 *   plus a breadth-first 6-connectivity labeller      (or_label_connected)
 * Build: oracle/Makefile with gcc -O2 -ffp-contract=off (no FMA contraction,
 * no reassociation), so every product and sum is rounded as written.
 *
 * Parity pins (tests/test_oracle_*.py): the reinit family is compared bit for
 * bit with the reference's own subroutines compiled from LevelSet_3D.f90 and
 * RM_UCNCT_3D.f90 (oracle/_ref, fixtures in tests/golden/); the operators
 * against closed forms (exact on linear fields), symmetry and a dense
 * assembly; CG against a dense direct solve; the table against hand-computed
 * values; CH against mass balance; AD against breadth-first reachability.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/elsim.h"

/* ghosted index, i,j,k in 0..n+1 */
#define GI(i, j, k) ((size_t)(i) + (size_t)(n1 + 2) * ((size_t)(j) + (size_t)(n2 + 2) * (size_t)(k)))
/* interior index, i,j,k in 1..n */
#define NI(i, j, k) ((size_t)((i)-1) + (size_t)n1 * ((size_t)((j)-1) + (size_t)n2 * (size_t)((k)-1)))

#define R4(x) ((double)(float)(x))

/* ======================================================================= */
/* Boundary semantics                                                      */
/* ======================================================================= */

/* BcMpiYZ on one rank (SBM:1031-1090): the d3 ghost planes copy the first /
 * last interior plane over the interior d1 x d2 range (SBM:1043-1064), then
 * the d1 ghost lines copy over interior d2 and all d3 (SBM:1067-1088). */
void or_bc_yz(int n1, int n2, int n3, double* u) {
  for (int j = 1; j <= n2; j++)
    for (int i = 1; i <= n1; i++) {
      u[GI(i, j, 0)] = u[GI(i, j, 1)];
      u[GI(i, j, n3 + 1)] = u[GI(i, j, n3)];
    }
  for (int k = 0; k <= n3 + 1; k++)
    for (int j = 1; j <= n2; j++) {
      u[GI(0, j, k)] = u[GI(1, j, k)];
      u[GI(n1 + 1, j, k)] = u[GI(n1, j, k)];
    }
}

/* d2 ghost planes over (0:n1+1, 0:n3+1), e.g. SBM:175-176. lo/hi: 0 = mirror,
 * 1 = set to value. */
static void bc_x(int n1, int n2, int n3, double* u, int lo_set, double lo_v, int hi_set, double hi_v) {
  for (int k = 0; k <= n3 + 1; k++)
    for (int i = 0; i <= n1 + 1; i++) {
      u[GI(i, 0, k)] = lo_set ? lo_v : u[GI(i, 1, k)];
      u[GI(i, n2 + 1, k)] = hi_set ? hi_v : u[GI(i, n2, k)];
    }
}

/* Full mirror (clamp) of all ghosts: BcMpiYZ + d2 mirror (SBM:173-176). */
void or_ghost_mirror(int n1, int n2, int n3, double* u) {
  or_bc_yz(n1, n2, n3, u);
  bc_x(n1, n2, n3, u, 0, 0.0, 0, 0.0);
}

/* NoGrdBc3D (LS:164-190). */
void or_nogrdbc3d(int n1, int n2, int n3, double* f) {
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++) {
      f[GI(0, j, k)] = f[GI(1, j, k)];
      f[GI(n1 + 1, j, k)] = f[GI(n1, j, k)];
    }
  for (int k = 1; k <= n3; k++)
    for (int i = 0; i <= n1 + 1; i++) {
      f[GI(i, 0, k)] = f[GI(i, 1, k)];
      f[GI(i, n2 + 1, k)] = f[GI(i, n2, k)];
    }
  for (int j = 0; j <= n2 + 1; j++)
    for (int i = 0; i <= n1 + 1; i++) {
      f[GI(i, j, 0)] = f[GI(i, j, 1)];
      f[GI(i, j, n3 + 1)] = f[GI(i, j, n3)];
    }
}

/* ======================================================================= */
/* Level set (LS:194-322)                                                  */
/* ======================================================================= */

/* AbsGrd3DUW_XYZ (LS:265-322): gr interior. */
void or_absgrd3duw(int n1, int n2, int n3, const double* sg, const double* f, double* gr) {
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double c = f[GI(i, j, k)];
        double ag = c - f[GI(i - 1, j, k)];
        double bg = f[GI(i + 1, j, k)] - c;
        double cg = c - f[GI(i, j - 1, k)];
        double dg = f[GI(i, j + 1, k)] - c;
        double eg = c - f[GI(i, j, k - 1)];
        double fg = f[GI(i, j, k + 1)] - c;
        double s = sg[GI(i, j, k)], g;
        if (s == 0.0) {
          g = 0.0;
        } else if (s > 0.0) {
          double a1 = fmax(ag, 0.0), a2 = fmax(-bg, 0.0);
          double b1 = fmax(cg, 0.0), b2 = fmax(-dg, 0.0);
          double c1 = fmax(eg, 0.0), c2 = fmax(-fg, 0.0);
          g = sqrt(fmax(a1 * a1, a2 * a2) + fmax(b1 * b1, b2 * b2) + fmax(c1 * c1, c2 * c2));
        } else {
          double a1 = fmax(-ag, 0.0), a2 = fmax(bg, 0.0);
          double b1 = fmax(-cg, 0.0), b2 = fmax(dg, 0.0);
          double c1 = fmax(-eg, 0.0), c2 = fmax(fg, 0.0);
          g = sqrt(fmax(a1 * a1, a2 * a2) + fmax(b1 * b1, b2 * b2) + fmax(c1 * c1, c2 * c2));
        }
        gr[NI(i, j, k)] = g;
      }
}

/* ReDist3D_XYZ (LS:194-261; RM:209-276 with dt = REAL(4) 0.1). */
void or_redist3d(int lit, int n1, int n2, int n3, const double* sg, double* f, double dt) {
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  double* gf = (double*)calloc(G, sizeof(double));
  double* gi = (double*)malloc((size_t)n1 * n2 * n3 * sizeof(double));
  double* nf = (double*)malloc(G * sizeof(double));
  for (int m = 1; m <= lit; m++) {
    or_absgrd3duw(n1, n2, n3, sg, f, gi);
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) gf[GI(i, j, k)] = gi[NI(i, j, k)];
    or_nogrdbc3d(n1, n2, n3, gf);
    for (size_t q = 0; q < G; q++) {
      double v = f[q] + sg[q] * (1.0 - gf[q]) * dt;
      if (!(v * f[q] > 0.0)) v = f[q] + dt * 1.0e-7 * sg[q];
      nf[q] = v;
    }
    memcpy(f, nf, G * sizeof(double));
  }
  free(gf);
  free(gi);
  free(nf);
}

/* FP32 variant with the same operation order (for the FP32 GPU error
 * statement); sg as int8. */
void or_redist3d_f32(int lit, int n1, int n2, int n3, const int8_t* sg, float* f, float dt) {
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  float* gf = (float*)calloc(G, sizeof(float));
  float* nf = (float*)malloc(G * sizeof(float));
  for (int m = 1; m <= lit; m++) {
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          float c = f[GI(i, j, k)];
          float ag = c - f[GI(i - 1, j, k)], bg = f[GI(i + 1, j, k)] - c;
          float cg = c - f[GI(i, j - 1, k)], dg = f[GI(i, j + 1, k)] - c;
          float eg = c - f[GI(i, j, k - 1)], fg = f[GI(i, j, k + 1)] - c;
          int s = sg[GI(i, j, k)];
          float g = 0.0f;
          if (s != 0) {
            float ss = (float)s;
            float a1 = fmaxf(ss * ag, 0.0f), a2 = fmaxf(-ss * bg, 0.0f);
            float b1 = fmaxf(ss * cg, 0.0f), b2 = fmaxf(-ss * dg, 0.0f);
            float c1 = fmaxf(ss * eg, 0.0f), c2 = fmaxf(-ss * fg, 0.0f);
            g = sqrtf(fmaxf(a1 * a1, a2 * a2) + fmaxf(b1 * b1, b2 * b2) + fmaxf(c1 * c1, c2 * c2));
          }
          gf[GI(i, j, k)] = g;
        }
    /* NoGrdBc3D on gf */
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++) {
        gf[GI(0, j, k)] = gf[GI(1, j, k)];
        gf[GI(n1 + 1, j, k)] = gf[GI(n1, j, k)];
      }
    for (int k = 1; k <= n3; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        gf[GI(i, 0, k)] = gf[GI(i, 1, k)];
        gf[GI(i, n2 + 1, k)] = gf[GI(i, n2, k)];
      }
    for (int j = 0; j <= n2 + 1; j++)
      for (int i = 0; i <= n1 + 1; i++) {
        gf[GI(i, j, 0)] = gf[GI(i, j, 1)];
        gf[GI(i, j, n3 + 1)] = gf[GI(i, j, n3)];
      }
    for (size_t q = 0; q < G; q++) {
      float s = (float)sg[q];
      float v = f[q] + s * (1.0f - gf[q]) * dt;
      if (!(v * f[q] > 0.0f)) v = f[q] + dt * 1.0e-7f * s;
      nf[q] = v;
    }
    memcpy(f, nf, G * sizeof(float));
  }
  free(gf);
  free(nf);
}

/* ======================================================================= */
/* Accelerated-diffusion connectivity (AD:83-243) and labelling             */
/* ======================================================================= */
/* Time-loop control (SBM:358-383)                                          */
/* ======================================================================= */

/* Vsr by the current error, then dt and a further factor by Xrf; default-REAL
 * literals (0.07, 0.20, 0.40, 0.75, 0.12, 0.20 and the dt product) in REAL(4). */
void or_time_control(double xrf, double scrnt, double trgi, double vsr0, double* dt, double* vsr) {
  const double dcrnt = fabs(trgi - scrnt);
  double v;
  if (dcrnt < R4(0.07) * fabs(trgi))
    v = vsr0 * 0.5;
  else if (dcrnt < R4(0.20) * fabs(trgi))
    v = vsr0;
  else if (dcrnt < R4(0.40) * fabs(trgi))
    v = vsr0 * 3.0;
  else if (dcrnt < R4(0.75) * fabs(trgi))
    v = vsr0 * 5.0;
  else
    v = vsr0 * 8.0;
  volatile float a = 1.05625e-02f, b = 0.125f, c = 0.75f, h = 0.5f;
  if (xrf < R4(0.12)) {
    *dt = (double)((float)(a * b) * c);
    v = v * 2;
  } else if (xrf < R4(0.20)) {
    *dt = (double)((float)(a * b) * h);
    v = v * 10;
  } else {
    *dt = (double)((float)(a * b) * h);
    v = v * 15;
  }
  *vsr = v;
}

/* ======================================================================= */
/* RM_UCNCT_3D sign block (RM:64-96, 115-136) and the 2x refinement         */
/* ======================================================================= */

/* One phase of RM_UCNCT_3D before its reinitialisation: psA = 0.5(1 +
 * tanh(RdIn/1.0)) (RM:66-67; 1 - psA for the electrolyte, RM:115), psB = 0
 * where Cn < 1e-4 (REAL(4), RM:80, 128), dsF = psB, NoGrdBc3D, sgF = +-1 by
 * dsF > 0.5 (RM:92-96, 133-137).  sgf ghosted. */
void or_rm_phase(int n1, int n2, int n3, const double* rdin, const double* cn, int phase, double* sgf) {
  double* ds = (double*)calloc((size_t)(n1 + 2) * (n2 + 2) * (n3 + 2), sizeof(double));
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        const size_t n = NI(i, j, k);
        double ps = 0.5 * (1.0 + tanh(rdin[n] / 1.0));
        if (phase == 1) ps = 1.0 - ps;
        if (cn[n] < R4(1.0e-4)) ps = 0.0;
        ds[GI(i, j, k)] = ps;
      }
  or_nogrdbc3d(n1, n2, n3, ds);
  const size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  for (size_t q = 0; q < G; q++) sgf[q] = ds[q] > 0.5 ? 1.0 : -1.0;
  free(ds);
}

/* The 2x refinement between RM_UCNCT_3D and the SBM program (not in the
 * reference): fine cell I of 2n at coarse coordinate (I + 0.5)/2, trilinear
 * from the eight surrounding coarse cells (weights 1/4, 3/4, indices clamped
 * to 1..n), interpolated along d1, then d2, then d3, times 2, to FP32. */
static void refine_axis(int I, int n, int* c0, int* c1, double* w0, double* w1) {
  const int c = (I + 1) / 2, lo = I & 1;
  int a = lo ? c - 1 : c, b = a + 1;
  *w0 = lo ? 0.25 : 0.75;
  *w1 = lo ? 0.75 : 0.25;
  *c0 = a < 1 ? 1 : (a > n ? n : a);
  *c1 = b < 1 ? 1 : (b > n ? n : b);
}
void or_refine2(int n1, int n2, int n3, const double* sdf, float* out) {
  const int m1 = 2 * n1, m2 = 2 * n2, m3 = 2 * n3;
  for (int K = 1; K <= m3; K++)
    for (int J = 1; J <= m2; J++)
      for (int I = 1; I <= m1; I++) {
        int i0, i1, j0, j1, k0, k1;
        double a0, a1, b0, b1, c0, c1;
        refine_axis(I, n1, &i0, &i1, &a0, &a1);
        refine_axis(J, n2, &j0, &j1, &b0, &b1);
        refine_axis(K, n3, &k0, &k1, &c0, &c1);
        const double x00 = a0 * sdf[NI(i0, j0, k0)] + a1 * sdf[NI(i1, j0, k0)];
        const double x10 = a0 * sdf[NI(i0, j1, k0)] + a1 * sdf[NI(i1, j1, k0)];
        const double x01 = a0 * sdf[NI(i0, j0, k1)] + a1 * sdf[NI(i1, j0, k1)];
        const double x11 = a0 * sdf[NI(i0, j1, k1)] + a1 * sdf[NI(i1, j1, k1)];
        const double y0 = b0 * x00 + b1 * x10, y1 = b0 * x01 + b1 * x11;
        out[(size_t)(I - 1) + (size_t)m1 * ((size_t)(J - 1) + (size_t)m2 * (size_t)(K - 1))] =
            (float)(2.0 * (c0 * y0 + c1 * y1));
      }
}

/* ======================================================================= */
/* Geometry front-end of LevelSet_3D (LS:84-144)                           */
/* ======================================================================= */

/* One frame of the LevelSet_3D main loop up to the reinitialisation, in the
 * reference's order: reorient the voxel image TmpA(yo,xo,zo) into
 * TmpB(yo,zo,xo) (LS:84-87), zero two cylindrical tunnels (LS:89-108), stack
 * the thickness copies behind a separator of sep voxels (LS:76, 110-121),
 * NoGrdBc3D, n_ac Allen-Cahn iterations (LS:123-135) and the sign function
 * (LS:138-143).  The reference's grid is py = 93, px = 50 + 4*170, pz = 160
 * and its tunnel axes sit at (i,k) = (1,160) and (93,1), i.e. (1,pz) and
 * (py,1).  The tunnel loop runs j over 1..px although TmpB's second extent is
 * zo; Fortran does not check bounds, so the element it zeroes is the one at
 * the column-major offset (i-1) + yo*(j-1) + yo*zo*(k-1), which this
 * restatement writes explicitly.  sgf, dsf: ghosted (py+2)(px+2)(pz+2).
 * Returns 0, or -1 if the tunnel loop would leave TmpB or the stacking does
 * not fit the grid. */
int or_ls_frontend(int yo, int xo, int zo, const double* tmpa, int py, int px, int pz, int sep, double tr,
                   int n_ac, double* sgf, double* dsf) {
  const int n1 = py, n2 = px, n3 = pz;
  const size_t ntb = (size_t)yo * zo * xo;
#define TA(i, j, k) tmpa[(size_t)((i)-1) + (size_t)yo * ((size_t)((j)-1) + (size_t)xo * (size_t)((k)-1))]
#define TBO(i, j, k) ((size_t)((i)-1) + (size_t)yo * ((size_t)((j)-1) + (size_t)zo * (size_t)((k)-1)))
  if (py > yo || pz > xo || sep < 0 || px <= sep || (px - sep) % zo != 0 || TBO(py, px, pz) >= ntb) return -1;
  double* tb = (double*)malloc(ntb * sizeof(double));
  if (!tb) return -1;
  /* LS:84-87 */
  for (int k = 1; k <= zo; k++)
    for (int j = 1; j <= xo; j++)
      for (int i = 1; i <= yo; i++) tb[TBO(i, k, j)] = TA(i, j, zo - k + 1);
  /* LS:89-108 */
  for (int k = 1; k <= pz; k++)
    for (int j = 1; j <= px; j++)
      for (int i = 1; i <= py; i++) {
        const double a = (double)i - 1.0, b = (double)k - (double)pz;
        double rad = sqrt(a * a + b * b);
        if (rad < tr) tb[TBO(i, j, k)] = 0.0;
        const double c = (double)i - (double)py, e = (double)k - 1.0;
        rad = sqrt(c * c + e * e);
        if (rad < tr) tb[TBO(i, j, k)] = 0.0;
      }
  /* LS:76, 110-121: block 0 as is, block 1 mirrored in d2, blocks 2, 3 copies of 0, 1 */
  memset(dsf, 0, (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2) * sizeof(double));
  for (int k = 1; k <= pz; k++)
    for (int j = sep + 1; j <= px; j++) {
      const int r = (j - sep - 1) % zo, blk = (j - sep - 1) / zo;
      const int jb = (blk % 2 == 0) ? r + 1 : zo - r;
      for (int i = 1; i <= py; i++) dsf[GI(i, j, k)] = tb[TBO(i, jb, k)];
    }
  free(tb);
#undef TA
#undef TBO
  or_nogrdbc3d(py, px, pz, dsf);
  /* LS:123-135: sgF is the reference's work array of the smoothing */
  const double c64 = R4(0.64), c05 = R4(0.05);
  for (int it = 0; it < n_ac; it++) {
    for (int k = 1; k <= pz; k++)
      for (int j = 1; j <= px; j++)
        for (int i = 1; i <= py; i++) {
          const size_t q = GI(i, j, k);
          const double d = dsf[q];
          const double lap = dsf[GI(i + 1, j, k)] + dsf[GI(i - 1, j, k)] + dsf[GI(i, j + 1, k)] +
                             dsf[GI(i, j - 1, k)] + dsf[GI(i, j, k + 1)] + dsf[GI(i, j, k - 1)] - 6 * d;
          sgf[q] = 2 * d * (1.0 - d) * (1.0 - 2 * d) - c64 * lap;
        }
    for (int k = 1; k <= pz; k++)
      for (int j = 1; j <= px; j++)
        for (int i = 1; i <= py; i++) {
          const size_t q = GI(i, j, k);
          dsf[q] = dsf[q] - c05 * sgf[q];
        }
    or_nogrdbc3d(py, px, pz, dsf);
  }
  /* LS:138-143 */
  const size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  for (size_t q = 0; q < G; q++) sgf[q] = dsf[q] > 0.5 ? 1.0 : -1.0;
  return 0;
}

/* ======================================================================= */

/* One phase of AccDiff_3D.  rdin interior FP64; cn interior FP64 output. */
void or_accdiff_phase(int n1, int n2, int n3, const double* rdin, int phase, int term, double* cn_out) {
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  double* psi = (double*)calloc(G, sizeof(double));
  double* cn = (double*)calloc(G, sizeof(double));
  double* nw = (double*)malloc((size_t)n1 * n2 * n3 * sizeof(double));
  /* AD:83-97 */
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) psi[GI(i, j, k)] = 0.5 * (1.0 + tanh(rdin[NI(i, j, k)] / 0.75));
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++) {
      psi[GI(0, j, k)] = psi[GI(1, j, k)];
      psi[GI(n1 + 1, j, k)] = psi[GI(n1, j, k)];
    }
  for (int k = 1; k <= n3; k++)
    for (int i = 0; i <= n1 + 1; i++) {
      psi[GI(i, 0, k)] = psi[GI(i, 1, k)];
      psi[GI(i, n2 + 1, k)] = psi[GI(i, n2, k)];
    }
  for (int j = 0; j <= n2 + 1; j++)
    for (int i = 0; i <= n1 + 1; i++) {
      psi[GI(i, j, 0)] = psi[GI(i, j, 1)];
      psi[GI(i, j, n3 + 1)] = psi[GI(i, j, n3)];
    }
  /* AD:175: electrolyte psi replaces the interior only; ghosts keep the
   * solid values. */
  if (phase == 1)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) psi[GI(i, j, k)] = 1.0 - psi[GI(i, j, k)];
  const double dt = R4(0.125f * 0.5f); /* AD:51 */
  const double idh2 = 1.0;               /* AD:49 */
  const double thr_upd = R4(4.0e-1f), thr_snap = R4(1.0e-4f);
  for (int iter = 2; iter <= term; iter++) {
    /* AD:118-128 (WHERE: right-hand side evaluated before assignment) */
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          double p = psi[GI(i, j, k)];
          double c = cn[GI(i, j, k)];
          if (p >= thr_upd) {
            double s = (psi[GI(i + 1, j, k)] + p) * (cn[GI(i + 1, j, k)] - c);
            s = s - (p + psi[GI(i - 1, j, k)]) * (c - cn[GI(i - 1, j, k)]);
            s = s + (psi[GI(i, j + 1, k)] + p) * (cn[GI(i, j + 1, k)] - c);
            s = s - (p + psi[GI(i, j - 1, k)]) * (c - cn[GI(i, j - 1, k)]);
            s = s + (psi[GI(i, j, k + 1)] + p) * (cn[GI(i, j, k + 1)] - c);
            s = s - (p + psi[GI(i, j, k - 1)]) * (c - cn[GI(i, j, k - 1)]);
            double div = 0.5 * idh2 / p * s;
            nw[NI(i, j, k)] = c + dt * div;
          } else {
            nw[NI(i, j, k)] = c;
          }
        }
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) cn[GI(i, j, k)] = nw[NI(i, j, k)];
    /* AD:132-134 */
    if (iter % 5 == 1)
      for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++)
          for (int i = 1; i <= n1; i++)
            if (cn[GI(i, j, k)] > thr_snap && psi[GI(i, j, k)] > thr_upd) cn[GI(i, j, k)] = 1.0;
    /* AD:138-148 (solid) / AD:212-222 (electrolyte) */
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++) {
        cn[GI(0, j, k)] = cn[GI(1, j, k)];
        cn[GI(n1 + 1, j, k)] = cn[GI(n1, j, k)];
      }
    for (int k = 1; k <= n3; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        if (phase == 0) {
          cn[GI(i, 0, k)] = cn[GI(i, 1, k)];
          cn[GI(i, n2 + 1, k)] = 1.0;
        } else {
          cn[GI(i, 0, k)] = 1.0;
          cn[GI(i, n2 + 1, k)] = cn[GI(i, n2, k)];
        }
      }
    for (int j = 0; j <= n2 + 1; j++)
      for (int i = 0; i <= n1 + 1; i++) {
        cn[GI(i, j, 0)] = cn[GI(i, j, 1)];
        cn[GI(i, j, n3 + 1)] = cn[GI(i, j, n3)];
      }
  }
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) cn_out[NI(i, j, k)] = cn[GI(i, j, k)];
  free(psi);
  free(cn);
  free(nw);
}

/* psi of the AD phase at an interior voxel (AD:84, AD:175). */
static double ad_psi(double d, int phase) {
  double p = 0.5 * (1.0 + tanh(d / 0.75));
  return phase == 1 ? 1.0 - p : p;
}

/* Breadth-first 6-connectivity labelling of {psi >= 0.4f} and reachability
 * from the Dirichlet face (d2 = n2 for the solid, d2 = 1 for the
 * electrolyte).  labels: min 0-based interior index of the component, -1
 * outside the phase. */
void or_label_connected(int n1, int n2, int n3, const double* rdin, int phase, int32_t* labels,
                        int32_t* connected) {
  size_t N = (size_t)n1 * n2 * n3;
  const double thr = R4(4.0e-1f);
  uint8_t* in = (uint8_t*)malloc(N);
  for (size_t q = 0; q < N; q++) {
    in[q] = ad_psi(rdin[q], phase) >= thr;
    labels[q] = -1;
    connected[q] = 0;
  }
  int64_t* queue = (int64_t*)malloc(N * sizeof(int64_t));
  /* components in increasing index order: the seed is the smallest index */
  for (size_t s = 0; s < N; s++) {
    if (!in[s] || labels[s] >= 0) continue;
    size_t head = 0, tail = 0;
    queue[tail++] = (int64_t)s;
    labels[s] = (int32_t)s;
    int touches = 0;
    size_t comp_start = 0;
    while (head < tail) {
      int64_t q = queue[head++];
      int i = (int)(q % n1) + 1;
      int j = (int)((q / n1) % n2) + 1;
      int k = (int)(q / ((int64_t)n1 * n2)) + 1;
      if ((phase == 0 && j == n2) || (phase == 1 && j == 1)) touches = 1;
      int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
      for (int t = 0; t < 6; t++) {
        int a = nb[t][0], b = nb[t][1], c = nb[t][2];
        if (a < 1 || a > n1 || b < 1 || b > n2 || c < 1 || c > n3) continue;
        size_t r = NI(a, b, c);
        if (in[r] && labels[r] < 0) {
          labels[r] = (int32_t)s;
          queue[tail++] = (int64_t)r;
        }
      }
    }
    if (touches)
      for (size_t t = comp_start; t < tail; t++) connected[queue[t]] = 1;
  }
  free(in);
  free(queue);
}

/* ======================================================================= */
/* Tbl_Chk_Pln_3DA (SBM:987-1028)                                          */
/* ======================================================================= */
double or_table(double c, const double* cord, const double* vals, int sz) {
  double ditv = cord[1] - cord[0];
  if (c < 0.0) c = 1.0e-6;
  if (c > 1.0) c = 1.0;
  int lb = (int)floor(c / ditv);
  if (lb == 0)
    lb = 1;
  else if (lb == sz)
    lb = sz - 1;
  if (lb < 1 || lb > sz - 1) lb = lb < 1 ? 1 : sz - 1; /* malformed tables / NaN only */
  return vals[lb - 1] + (c - cord[lb - 1]) / ditv * (vals[lb] - vals[lb - 1]);
}

void or_tbl_chk_pln_3da(size_t n, int sz, const double* cord, const double* vals, const double* cn,
                        double* fn) {
  for (size_t q = 0; q < n; q++) fn[q] = or_table(cn[q], cord, vals, sz);
}

/* ======================================================================= */
/* SBM state                                                               */
/* ======================================================================= */
typedef struct OrSbm {
  int n1, n2, n3;
  elsim_sbm_params p;
  int sz;
  double* tab; /* [4][2][sz]: mu, i0, OCV, mobility (scaled) */
  double tpp, tpe, xrf, bve;
  int iter;
  /* ghosted */
  double *psp, *pse, *cnp, *cne, *php, *phe, *kapp, *dab, *pce, *pcl, *mb, *mu, *pmn, *prj;
  /* interior */
  double *avp, *rxn, *rxc, *rhs, *res, *wmx, *tmp, *ceo, *ppo, *peo, *rhe, *kfw, *kbw, *lap, *dvc;
} OrSbm;

static double* galloc(const OrSbm* s) {
  return (double*)calloc((size_t)(s->n1 + 2) * (s->n2 + 2) * (s->n3 + 2), sizeof(double));
}
static double* nalloc(const OrSbm* s) {
  return (double*)calloc((size_t)s->n1 * s->n2 * s->n3, sizeof(double));
}

/* psi from an FP32 SDF (SBM:154-176). */
void or_psi_from_sdf(int n1, int n2, int n3, const float* sdf, double zeta, double* psi) {
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double g = (double)sdf[NI(i, j, k)];
        g = 0.5 * (1.0 + tanh(g / zeta));
        if (g < 0.0) g = 0.0;
        if (g > 1.0) g = 1.0;
        psi[GI(i, j, k)] = g;
      }
  or_ghost_mirror(n1, n2, n3, psi);
}

/* |grad psi| product (SBM:274-291); avp interior. */
void or_avp(int n1, int n2, int n3, const double* psp, const double* pse, double avp_min, double* avp) {
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double a = (psp[GI(i + 1, j, k)] - psp[GI(i - 1, j, k)]) / 2;
        double b = (psp[GI(i, j + 1, k)] - psp[GI(i, j - 1, k)]) / 2;
        double c = (psp[GI(i, j, k + 1)] - psp[GI(i, j, k - 1)]) / 2;
        double vp = sqrt(a * a + b * b + c * c);
        a = (pse[GI(i + 1, j, k)] - pse[GI(i - 1, j, k)]) / 2;
        b = (pse[GI(i, j + 1, k)] - pse[GI(i, j - 1, k)]) / 2;
        c = (pse[GI(i, j, k + 1)] - pse[GI(i, j, k - 1)]) / 2;
        double ve = sqrt(a * a + b * b + c * c);
        double v = sqrt(vp * ve);
        if (v < avp_min) v = 0.0;
        avp[NI(i, j, k)] = v;
      }
}

static double sum_n(const OrSbm* s, const double* a) {
  size_t N = (size_t)s->n1 * s->n2 * s->n3;
  double t = 0.0;
  for (size_t q = 0; q < N; q++) t += a[q];
  return t;
}

/* Interior sum of a ghosted array times a ghosted weight. */
static double sum_gw(const OrSbm* s, const double* a, const double* w) {
  int n1 = s->n1, n2 = s->n2, n3 = s->n3;
  double t = 0.0;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) t += a[GI(i, j, k)] * w[GI(i, j, k)];
  return t;
}

/* The 7-point stencil sum S = +f2 -f1 +f4 -f3 +f6 -f5 with face coefficient
 * 0.5(k_nb + k_c) (SBM:326-331 pattern) in the reference's evaluation order. */
static inline double stencil_sum(int n1, int n2, const double* kap, const double* p, int i, int j, int k) {
  size_t c = GI(i, j, k);
  double kc = kap[c], pc = p[c];
  size_t sy = 1, sx = (size_t)(n1 + 2), sz = (size_t)(n1 + 2) * (n2 + 2);
  double f1 = (kap[c - sy] + kc) * 0.5, f2 = (kap[c + sy] + kc) * 0.5;
  double f3 = (kap[c - sx] + kc) * 0.5, f4 = (kap[c + sx] + kc) * 0.5;
  double f5 = (kap[c - sz] + kc) * 0.5, f6 = (kap[c + sz] + kc) * 0.5;
  double s = f2 * (p[c + sy] - pc) - f1 * (pc - p[c - sy]);
  s = s + f4 * (p[c + sx] - pc);
  s = s - f3 * (pc - p[c - sx]);
  s = s + f6 * (p[c + sz] - pc);
  s = s - f5 * (pc - p[c - sz]);
  return s;
}

/* Operator on a ghosted p (ghosts as stored); out interior. */
void or_apply_op(int op, int n1, int n2, int n3, const double* kap, const double* pse, const double* p,
                 double dt, double dx, double* out) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double s = stencil_sum(n1, n2, kap, p, i, j, k);
        double v;
        if (op == ELSIM_OP_CNE)
          v = pse[GI(i, j, k)] * p[GI(i, j, k)] - dt / (dx * dx) * s;
        else
          v = 1.0 / (dx * dx) * s;
        out[NI(i, j, k)] = v;
      }
}

/* Homogeneous ghost rule of the CG search direction. */
void or_prj_bc(int op, int n1, int n2, int n3, double* p) {
  or_bc_yz(n1, n2, n3, p);
  if (op == ELSIM_OP_CNE) {
    /* SBM:511-513 */
    for (int k = 0; k <= n3 + 1; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        p[GI(i, n2 + 1, k)] = p[GI(i, n2, k)];
        p[GI(i, 0, k)] = p[GI(i, 1, k)];
      }
  } else if (op == ELSIM_OP_PHP) {
    /* SBM:666-668 */
    for (int k = 0; k <= n3 + 1; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        p[GI(i, n2 + 1, k)] = 0.0;
        p[GI(i, 0, k)] = p[GI(i, 1, k)];
      }
  } else if (op == ELSIM_OP_PHE) {
    /* SBM:747-749 */
    for (int k = 0; k <= n3 + 1; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        p[GI(i, n2 + 1, k)] = p[GI(i, n2, k)];
        p[GI(i, 0, k)] = 0.0;
      }
  } else {
    /* Dirichlet data on both d2 faces: zero search-direction ghosts there */
    for (int k = 0; k <= n3 + 1; k++)
      for (int i = 0; i <= n1 + 1; i++) {
        p[GI(i, n2 + 1, k)] = 0.0;
        p[GI(i, 0, k)] = 0.0;
      }
  }
}

/* Homogeneous operator A_hom e: e interior given (ghosted array whose ghosts
 * are overwritten with the search-direction rule of `op`). */
static void apply_hom(OrSbm* s, int op, const double* kap, double* e, double* out) {
  or_prj_bc(op, s->n1, s->n2, s->n3, e);
  or_apply_op(op, s->n1, s->n2, s->n3, kap, s->pse, e, s->p.dt, s->p.dx, out);
}

/* CG solve (SBM:492-555, 649-714, 729-795).  x ghosted (solution, interior
 * updated), kap ghosted coefficient, s->rhs interior.  Returns iterations;
 * *relres = true relative residual at exit (rel. to b_eff). */
static int or_cg(OrSbm* s, int op, double* x, const double* kap, int* status, double* relres) {
  int n1 = s->n1, n2 = s->n2, n3 = s->n3;
  size_t N = (size_t)n1 * n2 * n3;
  const elsim_sbm_params* P = &s->p;
  const double dt = P->dt, dx = P->dx;
  double crt = op == ELSIM_OP_CNE ? P->crt_e : (op == ELSIM_OP_PHP ? P->crt_p : P->crt_l);
  int maxit = op == ELSIM_OP_CNE ? P->max_it_e : (op == ELSIM_OP_PHP ? P->max_it_p : P->max_it_l);
  const double* w8 = op == ELSIM_OP_PHP ? s->psp : s->pse;
  double tw = op == ELSIM_OP_PHP ? s->tpp : s->tpe;
  /* The solve is for the correction e = x - x0 with homogeneous ghosts:
   * A_hom e = r0.  b_eff = b - A(zero interior, boundary ghosts)
   * = r0 + A_hom x0. */
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  double* ev = (double*)calloc(G, sizeof(double));
  double* r0 = (double*)malloc(N * sizeof(double));
  double* x0 = (double*)malloc(N * sizeof(double));
  /* initial residual, ghosts of x as stored (SBM:492-499, 649-655, 729-735) */
  or_apply_op(op, n1, n2, n3, kap, s->pse, x, dt, dx, s->wmx);
  for (size_t q = 0; q < N; q++) s->res[q] = s->rhs[q] - s->wmx[q];
  memcpy(r0, s->res, N * sizeof(double));
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        x0[NI(i, j, k)] = x[GI(i, j, k)];
        ev[GI(i, j, k)] = x[GI(i, j, k)];
        s->prj[GI(i, j, k)] = s->res[NI(i, j, k)];
      }
  apply_hom(s, op, kap, ev, s->wmx);
  double bn = 0.0;
  for (size_t q = 0; q < N; q++) {
    double b = r0[q] + s->wmx[q];
    bn += b * b;
  }
  bn = sqrt(bn);
  double rr = 0.0;
  for (size_t q = 0; q < N; q++) rr += s->res[q] * s->res[q];
  if (op == ELSIM_OP_CNE && P->cne_rule == 1) bn = sqrt(rr); /* relative to the step's change */
  int it = 0;
  *status = 0;
  if (P->stop_rule == ELSIM_STOP_RELRES && !(sqrt(rr) > P->rtol * bn)) {
    *relres = bn > 0 ? sqrt(rr) / bn : 0.0;
    free(ev);
    free(r0);
    free(x0);
    return 0;
  }
  double rms = 1.0;
  int itlp = 1;
  for (;;) {
    if (P->stop_rule == ELSIM_STOP_REFERENCE && !(rms > crt)) break;
    if (op != ELSIM_OP_CNE)
      for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++)
          for (int i = 1; i <= n1; i++) s->tmp[NI(i, j, k)] = x[GI(i, j, k)];
    or_prj_bc(op, n1, n2, n3, s->prj);
    or_apply_op(op, n1, n2, n3, kap, s->pse, s->prj, dt, dx, s->wmx);
    double dmy = 0.0, btm = 0.0;
    for (size_t q = 0; q < N; q++) dmy += s->res[q] * s->res[q];
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) btm += s->prj[GI(i, j, k)] * s->wmx[NI(i, j, k)];
    double alp = dmy / btm;
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) x[GI(i, j, k)] = x[GI(i, j, k)] + alp * s->prj[GI(i, j, k)];
    for (size_t q = 0; q < N; q++) s->res[q] = s->res[q] - alp * s->wmx[q];
    btm = 0.0;
    for (size_t q = 0; q < N; q++) btm += s->res[q] * s->res[q];
    double bet = btm / dmy;
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++)
          s->prj[GI(i, j, k)] = s->res[NI(i, j, k)] + bet * s->prj[GI(i, j, k)];
    it++;
    if (!isfinite(alp) || !isfinite(bet)) {
      *status = ELSIM_ERR_NAN;
      break;
    }
    if (P->stop_rule == ELSIM_STOP_REFERENCE) {
      double lsm = 0.0;
      if (op == ELSIM_OP_CNE) {
        for (int k = 1; k <= n3; k++)
          for (int j = 1; j <= n2; j++)
            for (int i = 1; i <= n1; i++) {
              double r = s->res[NI(i, j, k)];
              lsm += r * r * s->pse[GI(i, j, k)];
            }
      } else {
        for (int k = 1; k <= n3; k++)
          for (int j = 1; j <= n2; j++)
            for (int i = 1; i <= n1; i++) {
              double t = x[GI(i, j, k)] - s->tmp[NI(i, j, k)];
              lsm += t * t * w8[GI(i, j, k)];
            }
      }
      rms = sqrt(lsm / tw);
      if (!(rms > crt)) break;
    } else {
      if (!(sqrt(btm) > P->rtol * bn)) break;
    }
    if (itlp > maxit) {
      *status = ELSIM_NOT_CONVERGED;
      break;
    }
    itlp++;
  }
  /* true residual of the correction system: r0 - A_hom (x - x0) */
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) ev[GI(i, j, k)] = x[GI(i, j, k)] - x0[NI(i, j, k)];
  apply_hom(s, op, kap, ev, s->wmx);
  double t = 0.0;
  for (size_t q = 0; q < N; q++) {
    double r = r0[q] - s->wmx[q];
    t += r * r;
  }
  free(ev);
  free(r0);
  free(x0);
  *relres = bn > 0 ? sqrt(t) / bn : 0.0;
  if (getenv("ELSIM_ORACLE_TRACE"))
    fprintf(stderr, "or_cg op=%d it=%d status=%d bnorm=%.3e relres=%.3e\n", op, it, *status, bn, *relres);
  return it;
}

/* Dab = scale*D0*exp(c0 - c1*c + c2*(c3*c)**2) over the ghosted box (SBM:474,579,601). */
static inline double dab_of(const elsim_sbm_params* P, double scale_d0, double c) {
  double e = P->dab_c3 * c;
  return scale_d0 * exp(P->dab_c0 - P->dab_c1 * c + P->dab_c2 * (e * e));
}

/* Cell-centred coefficients of the step's operators over the ghosted box,
 * from the electrolyte concentration cne and psi_E (the face arrays of the
 * reference are their face averages):
 *   which 0: CnE solve   pce = psE * D0 exp(...)            (SBM:474, 478)
 *   which 1: RHE         pce = psE * tc1 D0 exp(...)        (SBM:579, 583)
 *   which 2: phE solve   pcN = psE * cne * tc2 D0 exp(...)  (SBM:601, 604)
 * dab (optional, ghosted) receives Dab itself. */
void or_coef(const elsim_sbm_params* P, int n1, int n2, int n3, int which, const double* cne, const double* pse,
             double* out, double* dab) {
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  double sd0 = which == 0 ? P->d0 : (which == 1 ? P->tc1 * P->d0 : P->tc2 * P->d0);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < G; q++) {
    double d = dab_of(P, sd0, cne[q]);
    if (dab) dab[q] = d;
    out[q] = which == 2 ? pse[q] * cne[q] * d : pse[q] * d;
  }
}

/* RHE = dx^-2 S(pce, CnE) over the interior, with CnE's ghosts as stored
 * (SBM:591-596); pce is the which-1 coefficient of or_coef. */
void or_rhe(int n1, int n2, int n3, const double* pce, const double* cne, double dx, double* rhe) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) rhe[NI(i, j, k)] = 1.0 / (dx * dx) * stencil_sum(n1, n2, pce, cne, i, j, k);
}

void* or_sbm_create(int n1, int n2, int n3, const elsim_sbm_params* prm, const float* sdf_s,
                    const float* sdf_l, int sz, const double* tables) {
  OrSbm* s = (OrSbm*)calloc(1, sizeof(OrSbm));
  s->n1 = n1;
  s->n2 = n2;
  s->n3 = n3;
  s->p = *prm;
  s->sz = sz;
  s->tab = (double*)malloc(sizeof(double) * 8 * sz);
  /* tables are used as given; the mobility scaling of SBM:242 is the caller's */
  memcpy(s->tab, tables, sizeof(double) * 8 * sz);
  s->psp = galloc(s); s->pse = galloc(s); s->cnp = galloc(s); s->cne = galloc(s);
  s->php = galloc(s); s->phe = galloc(s); s->kapp = galloc(s); s->dab = galloc(s);
  s->pce = galloc(s); s->pcl = galloc(s); s->mb = galloc(s); s->mu = galloc(s);
  s->pmn = galloc(s); s->prj = galloc(s);
  s->avp = nalloc(s); s->rxn = nalloc(s); s->rxc = nalloc(s); s->rhs = nalloc(s);
  s->res = nalloc(s); s->wmx = nalloc(s); s->tmp = nalloc(s); s->ceo = nalloc(s);
  s->ppo = nalloc(s); s->peo = nalloc(s); s->rhe = nalloc(s); s->kfw = nalloc(s);
  s->kbw = nalloc(s); s->lap = nalloc(s); s->dvc = nalloc(s);
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  or_psi_from_sdf(n1, n2, n3, sdf_s, prm->zeta, s->psp);
  or_psi_from_sdf(n1, n2, n3, sdf_l, prm->zeta, s->pse);
  for (size_t q = 0; q < G; q++) s->psp[q] = s->psp[q] + prm->psp_off; /* SBM:247 */
  for (size_t q = 0; q < G; q++) s->pse[q] = s->pse[q] + prm->pse_off; /* SBM:258 */
  s->tpp = 0.0;
  s->tpe = 0.0;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) s->tpp += s->psp[GI(i, j, k)];
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) s->tpe += s->pse[GI(i, j, k)];
  or_avp(n1, n2, n3, s->psp, s->pse, prm->avp_min, s->avp);
  /* SBM:297-341 initial fields */
  for (size_t q = 0; q < G; q++) {
    s->cnp[q] = 2.02e-2;
    s->cne[q] = prm->cne_target;
    s->php[q] = prm->bvp;
    s->phe[q] = prm->bve;
    s->kapp[q] = prm->kap_scale * s->psp[q]; /* SBM:324 */
  }
  s->bve = prm->bve;
  s->xrf = sum_gw(s, s->cnp, s->psp) / s->tpp;
  s->iter = 1;
  return s;
}

void or_sbm_destroy(void* h) {
  OrSbm* s = (OrSbm*)h;
  double* ptrs[] = {s->psp, s->pse, s->cnp, s->cne, s->php, s->phe, s->kapp, s->dab, s->pce, s->pcl,
                    s->mb,  s->mu,  s->pmn, s->prj, s->avp, s->rxn, s->rxc, s->rhs, s->res, s->wmx,
                    s->tmp, s->ceo, s->ppo, s->peo, s->rhe, s->kfw, s->kbw, s->lap, s->dvc, s->tab};
  for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++) free(ptrs[q]);
  free(s);
}

static double* field_ptr(OrSbm* s, int f, int* ghosted) {
  *ghosted = 1;
  switch (f) {
    case ELSIM_F_PSP: return s->psp;
    case ELSIM_F_PSE: return s->pse;
    case ELSIM_F_AVP: *ghosted = 0; return s->avp;
    case ELSIM_F_CNP: return s->cnp;
    case ELSIM_F_CNE: return s->cne;
    case ELSIM_F_PHP: return s->php;
    case ELSIM_F_PHE: return s->phe;
    case ELSIM_F_RXN: *ghosted = 0; return s->rxn;
    case ELSIM_F_KAPP: return s->kapp;
  }
  return NULL;
}

int or_sbm_get(void* h, int f, double* out) {
  OrSbm* s = (OrSbm*)h;
  int g;
  double* p = field_ptr(s, f, &g);
  if (!p) return -1;
  size_t n = g ? (size_t)(s->n1 + 2) * (s->n2 + 2) * (s->n3 + 2) : (size_t)s->n1 * s->n2 * s->n3;
  memcpy(out, p, n * sizeof(double));
  return 0;
}

int or_sbm_set(void* h, int f, const double* in) {
  OrSbm* s = (OrSbm*)h;
  int g;
  double* p = field_ptr(s, f, &g);
  if (!p) return -1;
  size_t n = g ? (size_t)(s->n1 + 2) * (s->n2 + 2) * (s->n3 + 2) : (size_t)s->n1 * s->n2 * s->n3;
  memcpy(p, in, n * sizeof(double));
  if (f == ELSIM_F_CNP) s->xrf = sum_gw(s, s->cnp, s->psp) / s->tpp; /* SBM:300-303 */
  return 0;
}

void or_sbm_scalars(void* h, double* tpp, double* tpe, double* xrf, double* bve, int* iter) {
  OrSbm* s = (OrSbm*)h;
  *tpp = s->tpp;
  *tpe = s->tpe;
  *xrf = s->xrf;
  *bve = s->bve;
  *iter = s->iter;
}

void or_sbm_set_params(void* h, const elsim_sbm_params* prm) {
  OrSbm* s = (OrSbm*)h;
  double bve = s->bve;
  s->p = *prm;
  s->bve = bve;
}

/* Cahn-Hilliard update (SBM:397-443). */
void or_ch_update(void* h) {
  OrSbm* s = (OrSbm*)h;
  int n1 = s->n1, n2 = s->n2, n3 = s->n3, sz = s->sz;
  const elsim_sbm_params* P = &s->p;
  double dx = P->dx;
  double* c = s->cnp;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double cc = c[GI(i, j, k)];
        double v = c[GI(i, j - 1, k)] - 2 * cc + c[GI(i, j + 1, k)];
        v = v + c[GI(i - 1, j, k)] - 2 * cc + c[GI(i + 1, j, k)];
        v = v + c[GI(i, j, k - 1)] - 2 * cc + c[GI(i, j, k + 1)];
        s->lap[NI(i, j, k)] = v / (dx * dx);
      }
  const double* T = s->tab;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        double cc = c[GI(i, j, k)];
        s->mu[GI(i, j, k)] = or_table(cc, T + 0 * sz, T + 1 * sz, sz);
        s->mb[GI(i, j, k)] = or_table(cc, T + 6 * sz, T + 7 * sz, sz);
      }
  or_ghost_mirror(n1, n2, n3, s->mb);
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) s->mu[GI(i, j, k)] = s->mu[GI(i, j, k)] - P->eps * s->lap[NI(i, j, k)];
  or_ghost_mirror(n1, n2, n3, s->mu);
  size_t G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  for (size_t q = 0; q < G; q++) s->pmn[q] = s->psp[q] * s->mb[q];
  const double* pm = s->pmn;
  const double* mu = s->mu;
  size_t sy = 1, sxx = (size_t)(n1 + 2), szz = (size_t)(n1 + 2) * (n2 + 2);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t q = GI(i, j, k);
        double v = 0.5 * (pm[q + sy] + pm[q]) * (mu[q + sy] - mu[q]) - 0.5 * (pm[q] + pm[q - sy]) * (mu[q] - mu[q - sy]);
        v = v + 1.0 * 0.5 * (pm[q + sxx] + pm[q]) * (mu[q + sxx] - mu[q]);
        v = v - 1.0 * 0.5 * (pm[q] + pm[q - sxx]) * (mu[q] - mu[q - sxx]);
        v = v + 0.5 * (pm[q + szz] + pm[q]) * (mu[q + szz] - mu[q]);
        v = v - 0.5 * (pm[q] + pm[q - szz]) * (mu[q] - mu[q - szz]);
        s->dvc[NI(i, j, k)] = v / (dx * dx);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t q = GI(i, j, k), r = NI(i, j, k);
        if (s->psp[q] > P->ch_psi_min)
          c[q] = c[q] + P->dt * (s->dvc[r] + (s->rxn[r] / P->rho) * (s->avp[r] / dx)) / s->psp[q];
      }
  or_ghost_mirror(n1, n2, n3, c);
  s->xrf = sum_gw(s, c, s->psp) / s->tpp;
}

/* CnE flux / Neumann ghosts (SBM:557-560, 569-571). */
static void cne_bc(OrSbm* s, double infx) {
  int n1 = s->n1, n2 = s->n2, n3 = s->n3;
  const elsim_sbm_params* P = &s->p;
  or_bc_yz(n1, n2, n3, s->cne);
  for (int k = 0; k <= n3 + 1; k++)
    for (int i = 0; i <= n1 + 1; i++)
      s->cne[GI(i, 0, k)] =
          infx * P->t_minus * P->dx / (0.5 * (s->dab[GI(i, 0, k)] + s->dab[GI(i, 1, k)])) + s->cne[GI(i, 1, k)];
  for (int k = 0; k <= n3 + 1; k++)
    for (int i = 0; i <= n1 + 1; i++) s->cne[GI(i, n2 + 1, k)] = s->cne[GI(i, n2, k)];
}

/* Butler-Volmer (SBM:631-641). */
void or_bv(void* h) {
  OrSbm* s = (OrSbm*)h;
  int n1 = s->n1, n2 = s->n2, n3 = s->n3;
  const elsim_sbm_params* P = &s->p;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t q = GI(i, j, k), r = NI(i, j, k);
        if (s->avp[r] > P->avp_min) {
          double dph = s->php[q] - s->phe[q];
          double rx = s->kfw[r] * s->cne[q] * exp(-(0.5 * P->cst1 * dph)) -
                      s->kbw[r] * s->cnp[q] * exp(0.5 * P->cst1 * dph);
          s->rxn[r] = rx;
          s->rxc[r] = rx * P->frd;
        }
      }
}

/* One electrochemical step (SBM:397-840) with the plain FP64 CG. */
int or_sbm_step(void* h, elsim_step_stats* st) {
  OrSbm* s = (OrSbm*)h;
  int n1 = s->n1, n2 = s->n2, n3 = s->n3, sz = s->sz;
  const elsim_sbm_params* P = &s->p;
  size_t N = (size_t)n1 * n2 * n3;
  double dx = P->dx, dt = P->dt;
  memset(st, 0, sizeof(*st));
  int status = 0, stt;
  double rr;

  /* 1. CH (SBM:397-443) */
  or_ch_update(s);
  st->xrf = s->xrf;

  /* 2. CnE (SBM:461-572) */
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t q = GI(i, j, k), r = NI(i, j, k);
        s->ceo[r] = s->cne[q];
        s->rhs[r] = s->pse[q] * s->ceo[r] - dt * (s->rxn[r]) * (s->avp[r] / dx) * P->t_minus;
      }
  double lsm = 0.0;
  for (size_t r = 0; r < N; r++) lsm += s->rxn[r] * (s->avp[r] / dx);
  lsm = lsm * (dx * dx * dx);
  double infx = lsm / ((double)(n1 * n3) * (dx * dx));
  or_coef(P, n1, n2, n3, 0, s->cne, s->pse, s->pce, s->dab);
  if (!(P->skip_cne_first && s->iter == 1)) {
    st->it_cne = or_cg(s, ELSIM_OP_CNE, s->cne, s->pce, &stt, &rr);
    st->relres_cne = rr;
    if (stt) status = stt;
  }
  cne_bc(s, infx);
  if (P->mass_corr_period > 0 && s->iter % P->mass_corr_period == 1) {
    double av = sum_gw(s, s->pse, s->cne) / s->tpe;
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) s->cne[GI(i, j, k)] = s->cne[GI(i, j, k)] - (av - P->cne_target);
    cne_bc(s, infx);
  }

  /* 3. potential coefficients (SBM:579-622) */
  or_coef(P, n1, n2, n3, 1, s->cne, s->pse, s->pce, NULL);
  or_rhe(n1, n2, n3, s->pce, s->cne, dx, s->rhe);
  or_coef(P, n1, n2, n3, 2, s->cne, s->pse, s->pcl, NULL);
  const double* T = s->tab;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t q = GI(i, j, k), r = NI(i, j, k);
        double cp = s->cnp[q];
        double ioc = or_table(cp, T + 2 * sz, T + 3 * sz, sz) * P->i0_scale;
        double ocv = or_table(cp, T + 4 * sz, T + 5 * sz, sz);
        s->kfw[r] = ioc / (P->frd * P->kfw_cref) * exp(0.5 * P->cst1 * ocv);
        s->kbw[r] = ioc / (P->frd * cp) * exp(-(0.5 * P->cst1 * ocv));
      }

  /* 4. outer loop (SBM:624-812) */
  double cvgp = 1.0, cvgl = 1.0;
  int nlp = 0;
  while (cvgp > P->outer_tol && cvgl > P->outer_tol) {
    if (nlp >= P->max_outer) {
      if (!status) status = ELSIM_NOT_CONVERGED;
      break;
    }
    or_bv(s);
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          size_t r = NI(i, j, k);
          s->ppo[r] = s->php[GI(i, j, k)];
          s->rhs[r] = -(s->rxc[r] * (s->avp[r] / dx));
        }
    for (int k = 0; k <= n3 + 1; k++)
      for (int i = 0; i <= n1 + 1; i++) s->php[GI(i, n2 + 1, k)] = P->bvp; /* SBM:647 */
    st->it_php += or_cg(s, ELSIM_OP_PHP, s->php, s->kapp, &stt, &rr);
    st->relres_php = rr;
    if (stt) status = stt;
    or_bc_yz(n1, n2, n3, s->php);
    bc_x(n1, n2, n3, s->php, 0, 0.0, 1, P->bvp); /* SBM:717-719 */
    double l = 0.0;
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          double d = s->php[GI(i, j, k)] - s->ppo[NI(i, j, k)];
          l += d * d * s->psp[GI(i, j, k)];
        }
    cvgp = sqrt(l / s->tpp);

    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          size_t r = NI(i, j, k);
          s->peo[r] = s->phe[GI(i, j, k)];
          s->rhs[r] = (s->rxn[r]) * (s->avp[r] / dx) + s->rhe[r];
        }
    st->it_phe += or_cg(s, ELSIM_OP_PHE, s->phe, s->pcl, &stt, &rr);
    st->relres_phe = rr;
    if (stt) status = stt;
    or_bc_yz(n1, n2, n3, s->phe);
    bc_x(n1, n2, n3, s->phe, 1, s->bve, 0, 0.0); /* SBM:798-800 */
    l = 0.0;
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          double d = s->phe[GI(i, j, k)] - s->peo[NI(i, j, k)];
          l += d * d * s->pse[GI(i, j, k)];
        }
    cvgl = sqrt(l / s->tpe);
    nlp++;
  }
  st->n_outer = nlp;
  st->cvgp = cvgp;
  st->cvgl = cvgl;

  /* 5. current control (SBM:826-835) */
  double sc = 0.0;
  for (size_t r = 0; r < N; r++) sc += s->rxn[r] * (s->avp[r] / dx);
  st->scrnt = sc;
  double dv = (P->trgi - sc) >= 0 ? fabs(P->vsr) : -fabs(P->vsr);
  s->bve = s->bve + dv * dt;
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) s->phe[GI(i, j, k)] = s->phe[GI(i, j, k)] + dv * dt;
  or_bc_yz(n1, n2, n3, s->phe);
  for (int k = 1; k <= n3; k++)
    for (int i = 0; i <= n1 + 1; i++) {
      s->phe[GI(i, 0, k)] = s->bve;
      s->phe[GI(i, n2 + 1, k)] = s->phe[GI(i, n2, k)];
    }
  st->status = status;
  s->iter++;
  return status;
}

/* ======================================================================= */
/* Standalone pieces for the oracle's own pins                             */
/* ======================================================================= */

/* One CG solve (same code path as the step) on caller-supplied arrays:
 * kap, pse, w8, x ghosted; rhs interior.  Returns the status. */
int or_solve_system(int op, int n1, int n2, int n3, const double* kap, const double* pse, double* x,
                    const double* rhs, double dt, double dx, int stop_rule, double rtol, double crt, int maxit,
                    const double* w8, double tw, int* iters, double* relres) {
  OrSbm s;
  memset(&s, 0, sizeof(s));
  s.n1 = n1;
  s.n2 = n2;
  s.n3 = n3;
  s.p.dt = dt;
  s.p.dx = dx;
  s.p.stop_rule = stop_rule;
  s.p.rtol = rtol;
  s.p.crt_e = s.p.crt_p = s.p.crt_l = crt;
  s.p.max_it_e = s.p.max_it_p = s.p.max_it_l = maxit;
  /* or_cg weights the reference stop norm with psp (phP) or pse (CnE, phE);
   * for phE the weight array is w8 */
  s.psp = (double*)w8;
  s.pse = (double*)(op == ELSIM_OP_PHE ? w8 : pse);
  s.tpp = s.tpe = tw;
  s.res = nalloc(&s);
  s.wmx = nalloc(&s);
  s.tmp = nalloc(&s);
  s.rhs = nalloc(&s);
  s.prj = galloc(&s);
  memcpy(s.rhs, rhs, sizeof(double) * (size_t)n1 * n2 * n3);
  int st = 0;
  *iters = or_cg(&s, op, x, kap, &st, relres);
  free(s.res);
  free(s.wmx);
  free(s.tmp);
  free(s.rhs);
  free(s.prj);
  return st;
}

/* Kfw, Kbw (SBM:614-622) and the Butler-Volmer rate (SBM:636-637) of one
 * voxel, with the step's code. */
double or_bv_voxel(const elsim_sbm_params* P, const double* tab, int sz, double cp, double ce, double dph) {
  double ioc = or_table(cp, tab + 2 * sz, tab + 3 * sz, sz) * P->i0_scale;
  double ocv = or_table(cp, tab + 4 * sz, tab + 5 * sz, sz);
  double kfw = ioc / (P->frd * P->kfw_cref) * exp(0.5 * P->cst1 * ocv);
  double kbw = ioc / (P->frd * cp) * exp(-(0.5 * P->cst1 * ocv));
  return kfw * ce * exp(-(0.5 * P->cst1 * dph)) - kbw * cp * exp(0.5 * P->cst1 * dph);
}

/* ======================================================================= */
/* CPU baseline timing (bench.py only)                                     */
/* ======================================================================= */
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads the build runs its parallel loops on (1 without OpenMP). */
int or_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Seconds taken by `niter` iterations of FP64 Jacobi-preconditioned CG on the
 * system of `op` (the algorithm of the FP64 GPU baseline: p = z + beta p,
 * w = A_hom p, alpha = r.z / p.w, x += alpha p, r -= alpha w, z = dinv r),
 * started from r = rhs, x = 0; kap ghosted, rhs interior.  The iterates are
 * not used: this times the CPU cost per iteration of the same workload. */
double or_pcg_time(int op, int n1, int n2, int n3, const double* kap, const double* pse, const double* rhs,
                   double dt, double dx, int niter) {
  size_t N = (size_t)n1 * n2 * n3, G = (size_t)(n1 + 2) * (n2 + 2) * (n3 + 2);
  double *x = (double*)calloc(N, sizeof(double)), *r = (double*)malloc(N * sizeof(double));
  double *z = (double*)malloc(N * sizeof(double)), *w = (double*)malloc(N * sizeof(double));
  double *dinv = (double*)malloc(N * sizeof(double)), *p = (double*)calloc(G, sizeof(double));
  const double sc = op == ELSIM_OP_CNE ? dt / (dx * dx) : 1.0 / (dx * dx);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        size_t c = GI(i, j, k), sx = (size_t)(n1 + 2), sz = sx * (n2 + 2);
        double kc = kap[c], sum = (kap[c - 1] + kc) * 0.5 + (kap[c + 1] + kc) * 0.5 + (kap[c - sx] + kc) * 0.5 +
                                  (kap[c + sx] + kc) * 0.5 + (kap[c - sz] + kc) * 0.5 + (kap[c + sz] + kc) * 0.5;
        double dg = op == ELSIM_OP_CNE ? pse[c] + sc * sum : -sc * sum;
        dinv[NI(i, j, k)] = 1.0 / dg;
      }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  double rz = 0.0;
#pragma omp parallel for reduction(+ : rz) schedule(static)
  for (size_t q = 0; q < N; q++) {
    r[q] = rhs[q];
    z[q] = dinv[q] * r[q];
    rz += r[q] * z[q];
  }
  double beta = 0.0;
  for (int it = 0; it < niter; it++) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) p[GI(i, j, k)] = z[NI(i, j, k)] + beta * p[GI(i, j, k)];
    or_prj_bc(op, n1, n2, n3, p);
    or_apply_op(op, n1, n2, n3, kap, pse, p, dt, dx, w);
    double pw = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : pw) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) pw += p[GI(i, j, k)] * w[NI(i, j, k)];
    const double alpha = rz / pw;
    double rzn = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : rzn) schedule(static)
    for (int k = 1; k <= n3; k++)
      for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
          size_t q = NI(i, j, k);
          x[q] += alpha * p[GI(i, j, k)];
          r[q] -= alpha * w[q];
          z[q] = dinv[q] * r[q];
          rzn += r[q] * z[q];
        }
    beta = rzn / rz;
    rz = rzn;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(x); free(r); free(z); free(w); free(dinv); free(p);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
