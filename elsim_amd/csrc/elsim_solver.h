// elsim_solver.h -- the conjugate-gradient solver interface shared by the SBM
// step driver and the standalone entry points.
#pragma once

#include "elsim_common.h"

namespace elsim {

// Device-resident CG scalars; kernels read alpha/beta from here so that a
// batch of iterations runs without host round trips.
struct CgDev {
  double rr;       // FP64 path: sum r^2 (the reference's DMY); mixed: sum r32^2
  double rz;       // mixed: sum r dinv r
  double pw;       // p . A p
  double alpha, beta;
  double bn2;      // ||b_eff||^2
  double tol2;     // (rtol ||b_eff||)^2
  double crt;      // reference stop tolerance
  double tw;       // reference stop weight total (tPP or tPE)
  double stop;     // last reference stop norm
  double rmax2;    // mixed: ||r||^2 at the last reliable update
  double delta2;   // mixed: reliable-update trigger (relative, squared)
  double rr_true;  // last FP64 true residual norm^2
  double rr0;      // FP64 residual norm^2 of the solve's start
  int iters;
  int maxit;
  int done;        // 0 running, 1 converged, 2 iteration cap, 3 not finite
  int need_reliable;
  int n_reliable;
  int stop_rule;
  int rel_r0;      // the relative rule's norm is ||r0|| instead of ||b_eff||
  int pending_y;   // mixed: this iteration's y += alpha p is still to be applied (by k_mx_dir3)
};

struct MgWork;

// Device buffers of one grid's solves (allocated once per grid).
struct CgWork {
  Dims d{0, 0, 0};
  DevBuf res, wmx, prj, x0, r0;      // FP64, G each (x0: the solve's start)
  DevBuf e64;                        // FP64 G (correction, ghosted for the stencil)
  DevBuf y32, r32, w32, p32;         // FP32, ghosted (even length: pair kernels)
  DevBuf kap32, pse32, dinv32;       // FP32: G, G; BF16 Jacobi inverse: G
  DevBuf dinv64;                     // FP64: N (ELSIM_MODE_FP64_PCG)
  DevBuf cg;                         // CgDev
  CgDev* host = nullptr;             // pinned mirror
  double* partials = nullptr;        // reduction partials (two values); null: the library workspace
  bool bound = false;                // buffers live in memory the caller provided (never freed here)
  MgWork* mg[4] = {};                // multigrid work space per operator (ELSIM_MODE_MIXED_MG), made on first use
  int alloc(const Dims& d);
  // Lay the buffers of grid d out in [base, base + bytes) (no allocation),
  // with `host` as the scalars' host mirror, and clear them with a fill
  // kernel.  ELSIM_ERR_ARG if they do not fit.
  int bind(const Dims& d, unsigned char* base, size_t bytes, CgDev* host);
  static size_t bound_bytes(const Dims& d);
  void release();
};

struct SolveArgs {
  Dims d;
  int op;               // ELSIM_OP_*
  double* x;            // ghosted FP64 unknown (ghosts = boundary data as stored)
  const double* kap;    // ghosted cell-centred coefficient
  const double* pse;    // ghosted psi_E (ELSIM_OP_CNE diagonal; may be null otherwise)
  const double* rhs;    // interior right-hand side
  double scale;         // dt/dx^2 (CNE) or 1/dx^2
  int stop_rule;        // ELSIM_STOP_*
  double rtol;
  double crt;           // reference tolerance (CrtE/CrtP/CrtL)
  int maxit;
  const double* w8;     // ghosted weight of the reference stop norm (psE or psP)
  double tw;            // its interior total
  double delta;         // mixed: reliable-update trigger
  int rel_r0 = 0;       // relative rule against ||r0|| (the start's residual) instead of ||b_eff||
  uint64_t coef_id = 0;  // nonzero: names kap (and pse); the multigrid set-up is reused while it is unchanged
  bool defer_add = false;  // mixed modes: x stays as it came in, the correction is left in work.e64 (e_pending)
  bool keep_x0 = true;     // mixed modes: store the start's copy in work.x0 (the FP64 modes always do)
};

struct SolveResult {
  int iters = 0;
  int status = 0;
  int n_reliable = 0;
  double relres = 0.0;  // FP64 true relative residual of the correction system
  double res0 = 0.0;    // ||b - A x|| at the solve's start (FP64)
  bool e_pending = false;  // defer_add: x += work.e64 (interior) is the caller's
  bool x0_stored = false;  // work.x0 holds the start's copy
};

// One CG solve of the correction system A_hom e = b - A x (x's ghosts as
// stored), x <- x + e.  On return x0 (interior copy of x at entry) is in
// work.x0 when SolveResult.x0_stored.  mode = ELSIM_MODE_FP64 or ELSIM_MODE_MIXED.
int cg_solve(CgWork& w, const SolveArgs& a, int mode, SolveResult* res);

// Single operator application (parity tests / roofline).  hom: in-kernel
// homogeneous ghost rule; otherwise ghosts are read from memory.
int launch_matvec_f64(int op, const Dims& d, const double* p, const double* kap, const double* pse,
                      double scale, double* out, double* dot_dev, bool hom);
int launch_matvec_f32(int op, const Dims& d, const float* p, const float* kap, const float* pse,
                      float scale, float* out, double* dot_dev, bool hom);

// ||rhs - A x||^2 over the interior, x's ghosts as stored (host result).
int launch_resid_norm2(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                       const double* rhs, double* out);

// r = rhs - A x (interior), x's ghosts as stored; *out = sum r^2 (host).
int launch_resid_store(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                       const double* rhs, double* r, double* out);
// sum over the interior of dv * (rhs - A x), x's ghosts as stored (host result).
int launch_dot_resid(int op, const Dims& d, const double* x, const double* kap, const double* pse, double scale,
                     const double* rhs, const double* dv, double* out);
// The Galerkin step's dot products along dv (phP or phE operator, host
// results): num = dv.(rhs - A x), x's ghosts as stored; den = dv.(A dv) with
// the homogeneous rule.
int launch_galerkin_dots(int op, const Dims& d, const double* x, const double* kap, double scale, const double* rhs,
                         const double* dv, double* num, double* den);

static_assert(sizeof(CgDev) <= 256, "CgDev must fit the 256 bytes a bound work space reserves for it");

// The standalone potential solve (elsim_solve_pot): argument checks (no
// device call) and the solve on work space w.
int pot_args_check(int n1, int n2, int n3, double dx, const void* kap, const void* rhs, const void* phi, int bc_lo,
                   int bc_hi, int mode, double rtol, int maxit);
int solve_pot_with(CgWork& w, int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi,
                   int bc_lo, int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres);
// True once elsim_init has run (the allocation-free entry points do not
// initialise the library themselves, since that allocates).
bool initialized();
// Set-up of the allocation-free entry points on the current device (called
// by elsim_init).
int noalloc_init(int device);

// In-loop kernel timing (elsim_prof_enable): when on, records the event at
// the start (end = 0) or end (end = 1) of slot's bracket on the library
// stream.
int prof_event(int slot, int end);

// Flat kernels shared with the SBM driver.
int launch_copy_interior_g2n(const Dims& d, const double* g, double* n);  // n = g(interior)

}  // namespace elsim
