// elsim_mg.h -- geometric multigrid preconditioner of the potential solves
// (ELSIM_MODE_MIXED_MG): the level hierarchy and the entry points used by the
// CG driver.
#pragma once

#include "elsim_solver.h"

namespace elsim {

// Coarse level: an operator A_hat x = dg x - sum_nb c_nb x_nb on an
// m1 x m2 x m3 box (interior only, index i + m1 (j + m2 k)), stored as the
// conductance to the +1 neighbour along each axis (cx, cy, cz; zero past the
// last cell), the diagonal dg, the smoother weight wd = omega / dg and the
// mass ms (boundary and reaction terms, kept for building the next level),
// plus the level's right-hand side b and iterate x.  All FP32.
struct MgLevel {
  int m1, m2, m3;
  float *cx, *cy, *cz, *dg, *wd, *ms, *b, *x;
};

constexpr int kMgMaxLev = 24;
constexpr int kMgCoarseMax = 64;    // cells of the directly solved level
constexpr int kMgBottomMax = 512;   // levels this small run in one single-block kernel
constexpr int kMgBottomCells = 640; // ... with at most this many cells together (staged in LDS)

struct MgLevels {
  MgLevel lv[kMgMaxLev + 1];
};

struct MgWork {
  Dims d{0, 0, 0};
  int nlev = 0;  // coarse levels 1 .. nlev; level nlev is solved directly
  int lb = 0;    // levels lb .. nlev run in the single-block kernel
  MgLevels L{};  // L.lv[0] unused: the fine level is the CG's own operator
  // fine level, ghosted FP32 with zero ghosts: the CG operator's coefficient
  // (and psE), its Jacobi inverse, the pre-smoothed iterate omega dinv r and
  // a temporary of the one-voxel-a-lane path
  float *kap32 = nullptr, *pse32 = nullptr, *dinv = nullptr, *x1 = nullptr, *t = nullptr;
  float* ainv = nullptr;    // dense inverse of level nlev, row-major FP32
  MgLevel* dlv = nullptr;  // the level table in device memory (the single-block kernel reads it)
  DevBuf pool;
  // what the set-up was built for (a nonzero coef_id names a coefficient
  // that has not changed since; 0 always rebuilds)
  int op = -1;
  uint64_t coef_id = 0;
  double scale = 0.0;
  float omega = 0.0f;
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

// Fine-level FP32 copies and Galerkin hierarchy of operator op on grid d from
// its FP64 ghosted cell-centred coefficient (and psE for ELSIM_OP_CNE), with
// the scale of the CG operator (dt/dx^2 or 1/dx^2).  Skipped when coef_id is
// nonzero and names what the work space holds.
int mg_setup(MgWork& m, int op, const Dims& d, const double* kap, const double* pse, double scale,
             uint64_t coef_id);

// r -= alpha w and m.x1 = omega dinv r (the CG's residual update and the
// V-cycle's pre-smoothing), or m.x1 alone (start and restarts).  r and w are
// ghosted FP32 vectors with zero ghosts.  Both return at once when cg says
// the solve is done or waits for a reliable update.
int mg_update(MgWork& m, const Dims& d, float* r32, const float* w32, const CgDev* cg);
int mg_presmooth(MgWork& m, const Dims& d, const float* r32, const CgDev* cg);

// z = M^-1 r by one V-cycle, m.x1 already holding omega dinv r.  Reduction
// partials: slot 0 = r.r, slot 1 = r.z, over mg_partial_blocks(d) blocks.
// timed: bracket the fine-level sweeps for the in-loop kernel timing.
int mg_vcycle(MgWork& m, int op, const Dims& d, const float* r32, float* z, const CgDev* cg, double* partials,
              bool timed);
int mg_partial_blocks(const Dims& d);

// The CG iteration's residual update and the V-cycle in one pass over the
// fine level (boxes whose n1 is a multiple of 4): r_out = r_in - alpha w
// (alpha from cg; use_alpha false: r = r_in, r_out may be null, as at a start
// or restart), z = M^-1 r, the pre-smoothed iterate formed where it is used.
// r_out must not alias r_in or w; its ghosts stay zero.  Partials as
// mg_vcycle (r.r of the new residual, r.z).
bool mg_fused_ok(const Dims& d);
int mg_vcycle_fused(MgWork& m, int op, const Dims& d, const float* r_in, const float* w32, bool use_alpha,
                    float* r_out, float* z, const CgDev* cg, double* partials, bool timed);

}  // namespace elsim
