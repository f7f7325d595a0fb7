// elsim_mg.h -- geometric multigrid preconditioner of the potential solves
// (ELSIM_MODE_MIXED_MG): the level hierarchy and the V-cycle entry points
// used by the CG driver.
#pragma once

#include "elsim_solver.h"

namespace elsim {

// Coarse level: an operator A_hat x = dg x - sum_nb c_nb x_nb on an
// m1 x m2 x m3 box (interior only, index i + m1 (j + m2 k)), stored as the
// conductance to the +1 neighbour along each axis (cx, cy, cz; zero past the
// last cell), the diagonal dg and the mass ms (boundary and reaction terms,
// kept for building the next level), plus the level's right-hand side b and
// iterate x.  All FP32.
struct MgLevel {
  int m1, m2, m3;
  float *cx, *cy, *cz, *dg, *ms, *b, *x;
};

constexpr int kMgMaxLev = 24;
constexpr int kMgCoarseMax = 64;  // cells of the directly solved level

struct MgWork {
  Dims d{0, 0, 0};
  int nlev = 0;                // coarse levels 1 .. nlev; level nlev is solved directly
  MgLevel lv[kMgMaxLev + 1]{};  // lv[0] unused: the fine level is the CG's own operator
  float* t = nullptr;          // fine-level temporary, ghosted FP32 with zero ghosts
  float* ainv = nullptr;       // dense inverse of level nlev, row-major FP32
  DevBuf pool;
  int alloc(const Dims& d);
  void release();
};

// Smoother weight, coarse-correction weight and the mixed solver's FP32
// residual reduction between FP64 reliable updates when it runs with the
// multigrid preconditioner.
struct MgCfg {
  float omega;
  float oc;
  double delta;
};
MgCfg& mg_cfg();

// Galerkin hierarchy of operator op on grid d from its FP64 ghosted
// cell-centred coefficient (and psE for ELSIM_OP_CNE), with the scale of the
// CG operator (dt/dx^2 or 1/dx^2).
int mg_setup(MgWork& m, int op, const Dims& d, const double* kap, const double* pse, double scale);

// z = M^-1 r by one V-cycle.  r and z are ghosted FP32 vectors with zero
// ghosts; kap32 / pse32 / dinv are the FP32 coefficient, FP32 psE and the
// BF16 Jacobi inverse of the CG operator.  Reduction partials: slot 0 = r.r,
// slot 1 = r.z, over mg_partial_blocks(d) blocks.  Every kernel returns at
// once when cg says the solve is done or waits for a reliable update.
int mg_vcycle(MgWork& m, int op, const Dims& d, const float* kap32, const float* pse32, const uint16_t* dinv,
              float scale, const float* r32, float* z, const CgDev* cg, double* partials);
int mg_partial_blocks(const Dims& d);

}  // namespace elsim
