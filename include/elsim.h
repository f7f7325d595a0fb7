/*
 * elsim.h -- C-ABI of the MI355X-native SBM solver core (libelsim_hip.so).
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and is
 * callable from Fortran through `bind(C)` interfaces with arguments passed by
 * reference (see elsim_amd/fortran/elsim_gpu.f90).  Return value: 0 = ok,
 * > 0 = did not converge (iteration cap hit), < 0 = error (see
 * elsim_last_error()).
 *
 * Array conventions follow the reference Fortran program exactly:
 *   - column-major, Fortran dim 1 ("y", i) contiguous, dim 2 ("x", j) is the
 *     through-thickness axis that carries every Dirichlet/flux condition,
 *     dim 3 ("z", k) slowest;
 *   - "ghosted" arrays are (0:n1+1, 0:n2+1, 0:n3+1), "interior" arrays are
 *     (1:n1, 1:n2, 1:n3);
 *   - ghost values are boundary conditions (Dirichlet values sit at the
 *     ghost-cell centre, SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:647).
 * Host-array entry points (no `_dev_` in the name) copy their inputs to the
 * GPU, run, and copy results back, so data crosses PCIe once per call.
 * `_dev_` entry points take device pointers (from elsim_malloc) and leave the
 * data resident in HBM.
 *
 * Reference file abbreviations used in the citations below:
 *   SBM = src/frtrn/SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90
 *   LS  = src/frtrn/LevelSet_3D.f90
 *   RM  = src/frtrn/RM_UCNCT_3D.f90
 *   AD  = src/frtrn/AccDiff_3D.f90
 */
#ifndef ELSIM_H
#define ELSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
#define ELSIM_OK            0
#define ELSIM_NOT_CONVERGED 1
#define ELSIM_ERR_ARG      -1
#define ELSIM_ERR_HIP      -2
#define ELSIM_ERR_NAN      -3
#define ELSIM_ERR_ALLOC    -4

/* Stop rules of the conjugate-gradient solves. */
#define ELSIM_STOP_REFERENCE 0 /* SBM:545-547 (CnE residual RMS), SBM:699-705, 780-786 (update RMS) */
#define ELSIM_STOP_RELRES    1 /* ||r||_2 <= rtol * ||b_eff||_2, b_eff = residual of a zero interior */

/* Precision / warm-start modes of an electrochemical step. */
#define ELSIM_MODE_FP64      0 /* reference CG (unpreconditioned) in FP64 */
#define ELSIM_MODE_MIXED     1 /* FP32 Jacobi-PCG inner iterations, FP64 reliable updates */
#define ELSIM_MODE_FP64_PCG  2 /* Jacobi-PCG in FP64 (same preconditioner as MIXED, all FP64) */
#define ELSIM_MODE_MIXED_MG  3 /* FP32 CG preconditioned by a geometric-multigrid V-cycle, FP64 reliable updates
                                  (the potential solves; CnE runs as ELSIM_MODE_MIXED) */

/* Which linear system (operator + homogeneous ghost rule of the search
 * direction) a CG solve works on. */
#define ELSIM_OP_CNE 0 /* A_c p = psE p - dt/dx^2 sum_f peN_f (p_nb - p); SBM:515-521; PRJ ghosts SBM:511-513 */
#define ELSIM_OP_PHP 1 /* L p = dx^-2 sum_f pkN_f (p_nb - p); SBM:670-676; PRJ ghosts SBM:666-668 */
#define ELSIM_OP_PHE 2 /* L p = dx^-2 sum_f pcN_f (p_nb - p); SBM:751-757; PRJ ghosts SBM:747-749 */
#define ELSIM_OP_POT_DD 3 /* the potential operator with Dirichlet data on both d2 faces: zero search-direction
                             ghosts there (elsim_solve_pot with bc_lo = bc_hi = 1) */

/* Parameters of one SBM run.  Defaults (elsim_sbm_default_params) are the
 * reference's literal values with their REAL(4)/REAL(8) kinds resolved to the
 * double the reference computes with.  Layout is mirrored by
 * elsim_amd/sbm.py (ctypes) and elsim_amd/fortran/elsim_gpu.f90 (bind(C)). */
typedef struct elsim_sbm_params {
  double dx;          /* SBM:90 */
  double dt;          /* SBM:91; per step from elsim_sbm_time_control in the reference's loop (SBM:374-383) */
  double zeta;        /* SBM:94  psi = 0.5(1+tanh(d/zeta)) */
  double psp_off;     /* SBM:247 */
  double pse_off;     /* SBM:258 */
  double avp_min;     /* SBM:287 */
  double eps;         /* SBM:136 gradient-energy coefficient */
  double rho;         /* SBM:100 */
  double ch_psi_min;  /* SBM:430 */
  double frd;         /* SBM:102 Faraday constant */
  double cst1;        /* SBM:101 F/RT */
  double t_minus;     /* SBM:121 */
  double d0;          /* SBM:124 */
  double tc1;         /* SBM:125 */
  double tc2;         /* SBM:126 */
  double dab_c0;      /* SBM:474  D = D0*exp(c0 - c1*c + c2*(c3*c)**2) */
  double dab_c1;
  double dab_c2;
  double dab_c3;
  double kap_scale;   /* SBM:324 */
  double i0_scale;    /* SBM:617 */
  double kfw_cref;    /* SBM:621 */
  double bvp;         /* SBM:106 */
  double bve;         /* SBM:107 (evolves: SBM:830) */
  double vsr;         /* SBM:830 voltage scan rate (fixed) */
  double trgi;        /* SBM:341 target current */
  double cne_target;  /* SBM:567 */
  double rtol;        /* relative-residual tolerance (ELSIM_STOP_RELRES) */
  double crt_e;       /* SBM:96 */
  double crt_p;       /* SBM:97 */
  double crt_l;       /* SBM:98 */
  double outer_tol;   /* SBM:627 */
  double inner_rtol;  /* mixed mode: FP32 residual reduction between FP64 reliable updates */
  int32_t mass_corr_period; /* SBM:562 */
  int32_t skip_cne_first;   /* SBM:505 */
  int32_t stop_rule;        /* ELSIM_STOP_* */
  int32_t max_it_e;         /* SBM:551 */
  int32_t max_it_p;         /* SBM:710 */
  int32_t max_it_l;         /* SBM:791 */
  int32_t max_outer;        /* cap on SBM:627 (the reference has none) */
  int32_t ws_mode;          /* warm start of the first phP solve: ELSIM_WS_* */
  int32_t cne_rule;         /* ELSIM_STOP_RELRES only: 0 = the CnE solve stops at rtol ||b_eff|| like the
                               potentials; 1 = at rtol ||r0||, r0 = the residual of the last step's CnE,
                               i.e. the step's change is resolved to rtol (CnE has no warm start, so this
                               does not depend on a start policy) */
  int32_t ws_order;         /* terms of the extrapolated start of phi_P: 0 = all (linear extrapolation and
                               the Galerkin steps along the second and third differences of its last
                               changes), 1 = linear only, 2 = up to the second difference */
} elsim_sbm_params;

/* Warm start of a step's coupling loop (SBM:624-812), from the third step
 * on: phi_P and phi_E are extrapolated linearly in time (phi + its change
 * over the last step) before the first Butler-Volmer evaluation, and with a
 * network attached the first phP solve starts from the extrapolated phi_P plus
 * a Galerkin step along the network's prediction of the remaining error. */
enum {
  ELSIM_WS_PLAIN = 0,    /* the last step's fields (as the reference) */
  ELSIM_WS_EXTRAP = 1,   /* linear extrapolation in time */
  ELSIM_WS_LEARNED = 2   /* extrapolation + learned correction of phi_P */
};

/* Per-step statistics returned by the step entry points. */
typedef struct elsim_step_stats {
  int32_t it_cne;      /* CG iterations of the CnE solve */
  int32_t it_php;      /* CG iterations summed over outer iterations */
  int32_t it_phe;
  int32_t n_outer;     /* outer (BV/phP/phE) iterations */
  int32_t n_reliable;  /* FP64 reliable updates (mixed mode) */
  int32_t status;      /* worst solve status */
  double  xrf;         /* SBM:443 */
  double  scrnt;       /* SBM:828 */
  double  cvgp, cvgl;  /* SBM:723, 804 at exit */
  double  relres_cne, relres_php, relres_phe; /* true FP64 relative residual at exit of the last solve */
  double  seconds;     /* device time of the step (hipEvent) */
  int32_t ws_used;     /* 1: extrapolated start, 2: plus the learned correction */
  int32_t ws_pad_;
  double  ws_res_plain, ws_res_nn; /* ||b - A x|| of the first phP solve: at the extrapolated start (before
                                      the correction), and at the corrected start with the reaction
                                      re-evaluated there (the residual that solve starts from) */
  double  ws_alpha;     /* Galerkin step length along the prediction */
  double  ws_energy_gain; /* decrease of the squared A-norm error it gave */
} elsim_step_stats;

void elsim_sbm_default_params(elsim_sbm_params* p);

/* ------------------------------------------------------------------------ */
/* Runtime                                                                  */
/* ------------------------------------------------------------------------ */
int         elsim_init(int device);
const char* elsim_last_error(void);
int         elsim_device_count(void);
int         elsim_synchronize(void);
int         elsim_malloc(void** ptr, size_t bytes);
int         elsim_free(void* ptr);
int         elsim_memcpy_h2d(void* dst, const void* src, size_t bytes);
int         elsim_memcpy_d2h(void* dst, const void* src, size_t bytes);
int         elsim_memcpy_d2d(void* dst, const void* src, size_t bytes);
int         elsim_memset(void* dst, int value, size_t bytes);
/* Free and total device memory of the current device (bytes). */
int         elsim_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Event timer on the library's stream: start, stop -> milliseconds. */
int         elsim_timer_start(void);
int         elsim_timer_stop(float* ms);

/* ------------------------------------------------------------------------ */
/* Level-set family (LS:164-322, RM:180-337)                                */
/* ------------------------------------------------------------------------ */
/* Replaces NoGrdBc3D(py,px,pz,f) (LS:164-190): Neumann ghost copy, in place. */
int elsim_nogrdbc3d(const int* py, const int* px, const int* pz, double* f);
/* Replaces AbsGrd3DUW_XYZ(py,px,pz,sg,f,gr) (LS:265-322): Godunov upwind
 * |grad f|; sg, f ghosted; gr interior. */
int elsim_absgrd3duw_xyz(const int* py, const int* px, const int* pz,
                         const double* sg, const double* f, double* gr);
/* Replaces ReDist3D_XYZ(lit,py,px,pz,sg,f) (LS:194-261; RM:209-276 with
 * dt = REAL(4) 0.1): lit sweeps of f <- f + sg(1-|grad f|)dt with the
 * sign-flip guard, over the ghosted box.  dt is 0.1d0 for LS, and
 * 0.10000000149011612 for RM. */
int elsim_redist3d_xyz(const int* lit, const int* py, const int* px, const int* pz,
                       const double* sg, double* f, const double* dt);
/* Same sweeps in FP32 on device: sg as int8 (+1/0/-1), f ghosted FP32. */
int elsim_dev_redist3d_f32(int lit, int py, int px, int pz, const int8_t* sg,
                           float* f, float dt);
int elsim_dev_redist3d_f64(int lit, int py, int px, int pz, const double* sg,
                           double* f, double dt);
/* The step the reference does not contain between RM_UCNCT_3D (distance
 * functions on py x px x pz, FP64, RM:99-107) and the SBM program (FP32
 * distance functions on a grid twice as fine, SBM:138-152, 181-190): no
 * reference interface exists.  out (2py x 2px x 2pz, FP32, interior,
 * column-major) = 2 x the trilinear interpolation of sdf (interior FP64) at
 * the fine cell centres (the two fine cells of coarse cell c at c -+ 0.25,
 * weights 1/4, 3/4; coarse indices clamped to the box, the Neumann rule),
 * i.e. the distance in fine voxels. */
int elsim_refine2_sdf(const int* py, const int* px, const int* pz, const double* sdf, float* out);
int elsim_dev_refine2_sdf(int py, int px, int pz, const double* sdf, float* out);
/* Replaces the geometry front-end of LevelSet_3D's frame loop (LS:84-144),
 * the inline block between READ(20) of the voxel image and the call of
 * ReDist3D_XYZ: reorient tmpa(yo,xo,zo) into TmpB(yo,zo,xo) (LS:84-87), zero
 * the two cylindrical tunnels of radius tr around the d2 lines at
 * (i,k) = (1,pz) and (py,1) (LS:89-108; the loop's TmpB(i,j,k) with j up to
 * px beyond TmpB's second extent zo is resolved, as Fortran does without
 * bounds checks, by column-major offset), stack (px-sep)/zo thickness copies
 * behind a separator of sep voxels, alternately as is and mirrored in d2
 * (LS:76, 110-121), NoGrdBc3D, n_ac Allen-Cahn smoothing iterations
 * (LS:123-135) and the sign function (LS:138-143).  The reference: yo, xo,
 * zo = 500, 520, 170; py, px, pz = 93, 50 + 4*170, 160; sep = 50; n_ac = 50;
 * tr = TuRd(n) as stored (REAL(4) literals).  Outputs ghosted FP64:
 * sgf = +-1, dsf = the smoothed field (the reference then zeroes dsF and calls
 * ReDist3D_XYZ(175, ..., sgF, dsF)).  ELSIM_ERR_ARG unless py <= yo,
 * pz <= xo, (px - sep) % zo == 0 and the tunnel loop stays inside TmpB. */
int elsim_ls_frontend(const int* yo, const int* xo, const int* zo, const double* tmpa,
                      const int* py, const int* px, const int* pz, const int* sep,
                      const double* tr, const int* n_ac, double* sgf, double* dsf);
int elsim_dev_ls_frontend(int yo, int xo, int zo, const double* tmpa, int py, int px,
                          int pz, int sep, double tr, int n_ac, double* sgf, double* dsf);

/* ------------------------------------------------------------------------ */
/* Connectivity (AD:68-245, RM:79-82,128-130)                               */
/* ------------------------------------------------------------------------ */
/* One phase of AccDiff_3D (AD:103-169 solid, AD:174-243 electrolyte):
 * rdin = signed distance (interior, FP64); phase 0 = solid (Dirichlet at the
 * high d2 face), 1 = electrolyte (psi -> 1-psi, Dirichlet at the low d2
 * face).  term = last iteration index (reference 1501).  cn = interior FP64
 * connectivity field written by AD:166 / AD:240. */
int elsim_accdiff_phase(const int* py, const int* px, const int* pz, const double* rdin,
                        const int* phase, const int* term, double* cn);
int elsim_dev_accdiff_phase(int py, int px, int pz, const double* rdin, int phase,
                            int term, double* cn);
/* Connected-component labelling (6-connectivity) of {psi >= thr} with
 * psi = 0.5(1+tanh(sign*rdin/width)), and of which components touch the
 * Dirichlet face of `phase` (as in AccDiff_3D).  labels: int32 interior,
 * -1 outside the phase, otherwise the smallest 0-based interior linear index
 * in the component.  connected: 1 where the component touches the face.
 * This is the exact, iteration-free form of the AD mask `Cn >= 1e-4`
 * (RM:80,128) once the AD front has covered every reachable voxel. */
int elsim_label_connected(const int* py, const int* px, const int* pz, const double* rdin,
                          const int* phase, int32_t* labels, int32_t* connected);
int elsim_dev_label_connected(int py, int px, int pz, const double* rdin, int phase,
                              int32_t* labels, int32_t* connected);
/* RM:79-82 / RM:128-130: psi_out = (cn < 1e-4f) ? 0 : psi_in (interior). */
int elsim_rm_threshold(const int* py, const int* px, const int* pz, const double* cn,
                       const double* psi_in, double* psi_out);
/* Replaces RM_UCNCT_3D's sign block of one phase (RM:64-96 solid, phase 0;
 * RM:115-136 electrolyte, phase 1): psA = 0.5(1+tanh(rdin/1.0)) (1 - psA for
 * the electrolyte), zeroed where cn < 1e-4f, NoGrdBc3D, sgf = +1 where > 0.5
 * else -1 over the ghosted box.  rdin, cn interior FP64 (the L24Z distance
 * function and the AccDiff connectivity of that phase); sgf ghosted FP64.
 * The program then sets dsF = 0 and calls ReDist3D_XYZ(175, ..., sgF, dsF)
 * with its REAL(4) time step to write GRP_LS / POR_LS. */
int elsim_rm_phase(const int* py, const int* px, const int* pz, const double* rdin,
                   const double* cn, const int* phase, double* sgf);
int elsim_dev_rm_phase(int py, int px, int pz, const double* rdin, const double* cn, int phase,
                       double* sgf);

/* ------------------------------------------------------------------------ */
/* SBM solver core (SBM:143-838)                                             */
/* ------------------------------------------------------------------------ */
/* Replaces Tbl_Chk_Pln_3DA(py,px,pz,sz,Cord,Vals,Cn,fn) (SBM:987-1028). */
int elsim_tbl_chk_pln_3da(const int* py, const int* px, const int* pz, const int* sz,
                          const double* cord, const double* vals, const double* cn, double* fn);

/* The SBM state lives on the device; the handle is an opaque integer.
 * tables: 4 tables x 2 rows x sz, row 0 = concentration grid, row 1 = values,
 * in the order mu (C_Li_M6), i0 (C_Li_J2), OCV (C_Li_O3), mobility (C_Li_Mb5
 * after the caller's SBM:242 scaling).
 * sdf_s, sdf_l: interior FP32 signed distances of the particle and
 * electrolyte phases (SBM:149-152, 181-184). */
int elsim_sbm_create(const int* n1, const int* n2, const int* n3, const elsim_sbm_params* prm,
                     const float* sdf_s, const float* sdf_l, const int* sz,
                     const double* tables, int64_t* handle);
int elsim_sbm_destroy(const int64_t* handle);
/* Field ids for get/set: interior or ghosted FP64 arrays. */
#define ELSIM_F_PSP   0  /* ghosted */
#define ELSIM_F_PSE   1  /* ghosted */
#define ELSIM_F_AVP   2  /* interior */
#define ELSIM_F_CNP   3  /* ghosted */
#define ELSIM_F_CNE   4  /* ghosted */
#define ELSIM_F_PHP   5  /* ghosted */
#define ELSIM_F_PHE   6  /* ghosted */
#define ELSIM_F_RXN   7  /* interior */
#define ELSIM_F_KAPP  8  /* ghosted, cell-centred kappa of the phP operator */
#define ELSIM_F_COUNT 9
int elsim_sbm_get(const int64_t* handle, const int* field, double* out);
int elsim_sbm_set(const int64_t* handle, const int* field, const double* in);
/* Device pointer of a state field (for device-resident callers). */
int elsim_sbm_field_ptr(const int64_t* handle, const int* field, double** ptr);
/* Scalars: tPP, tPE, Xrf, BvE, iteration counter. */
int elsim_sbm_scalars(const int64_t* handle, double* tpp, double* tpe, double* xrf,
                      double* bve, int* iter);
int elsim_sbm_set_params(const int64_t* handle, const elsim_sbm_params* prm);
/* One electrochemical step SBM:397-840 (CH -> CnE solve -> outer
 * {BV -> phP -> phE} -> current control).  mode = ELSIM_MODE_*. */
int elsim_sbm_step(const int64_t* handle, const int* mode, elsim_step_stats* stats);
/* The time-loop control of the SBM program before each step (SBM:358-383):
 * from the lithiation fraction xrf, the last step's current scrnt, the target
 * current trgi and the base scan rate vsr0 (the reference's Vsr0, SBM:112),
 * the time step dt and scan rate vsr of the next step, with the reference's
 * REAL(4) literals.  Host logic; the caller puts dt and vsr into the step's
 * parameters (elsim_sbm_set_params). */
int elsim_sbm_time_control(const double* xrf, const double* scrnt, const double* trgi, const double* vsr0,
                           double* dt, double* vsr);
/* One potential solve on its own: the reference's inline block of the
 * particle potential (SBM:643-723) and electrolyte potential (SBM:725-804)
 * solves as a subroutine.  Solves dx^-2 S(kap, phi) = rhs (S the 7-point sum
 * of SBM:670-676 with face coefficients averaged from the cell-centred kap)
 * for phi, starting from phi as given, to ||rhs - A phi|| <= rtol ||b_eff||
 * (b_eff = the residual of a zero interior with phi's boundary data).
 * Boundary conditions: Neumann (mirror ghosts) on the d1 and d3 faces; on
 * each d2 face bc = 0 Neumann or 1 Dirichlet at the ghost-cell centre, the
 * value phi holds in that ghost plane (at least one face Dirichlet; the
 * reference's phP is (0, 1) with BvP, SBM:647, its phE (1, 0) with BvE,
 * SBM:798-800).  kap, phi ghosted (0:n1+1, 0:n2+1, 0:n3+1), rhs interior.
 * mode = ELSIM_MODE_* (FP64: the reference CG, SBM:661-714).  Returns 0,
 * ELSIM_NOT_CONVERGED after maxit iterations, ELSIM_ERR_NAN on a non-finite
 * value; argument errors are checked before any device call. */
int elsim_solve_pot(const int* n1, const int* n2, const int* n3, const double* dx, const double* kap,
                    const double* rhs, double* phi, const int* bc_lo, const int* bc_hi, const int* mode,
                    const double* rtol, const int* maxit, int* iters, double* relres);
int elsim_dev_solve_pot(int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi,
                        int bc_lo, int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): elsim_dev_solve_pot with its CG work space in a fixed
 * device array of the library instead of allocated buffers.  Same arguments
 * and results (bit-identical).  Grids up to ELSIM_NOALLOC_MAX_G ghosted
 * elements (n1+2)(n2+2)(n3+2), i.e. 64^3 interior; larger grids are
 * ELSIM_ERR_ARG, and so is ELSIM_MODE_MIXED_MG (its hierarchy is built in
 * allocated memory).  elsim_init must have run (it sets the array up; this entry
 * point allocates nothing and does not initialise the library).  The array is
 * one per GPU and shared by all calls on it: calls that overlap on one GPU
 * race. */
#define ELSIM_NOALLOC_MAX_G 287496
int elsim_dev_solve_pot_noalloc(int n1, int n2, int n3, double dx, const double* kap, const double* rhs, double* phi,
                                int bc_lo, int bc_hi, int mode, double rtol, int maxit, int* iters, double* relres);
/* The Cahn-Hilliard stage alone (SBM:397-443). */
int elsim_sbm_ch_update(const int64_t* handle);
/* Homogeneous operator of the state's current coefficients applied to a
 * device ghosted field (interior output on device). */
int elsim_sbm_matvec(const int64_t* handle, const int* op, const double* p_dev_ghosted,
                     double* out_dev_interior);

/* Standalone operator application on device arrays (one launch), used for
 * parity tests and roofline measurement.  p: ghosted input whose ghost rule
 * is the homogeneous one of `op` (applied in-kernel, ghost memory not read
 * except the ghost coefficients).  kap: ghosted cell-centred coefficient.
 * pse: ghosted psi_E (only for ELSIM_OP_CNE).  out: interior.  scale:
 * dt/dx^2 (CNE) or 1/dx^2 (potentials).  dot (device, 1 double): p.out. */
int elsim_dev_matvec_f64(int op, int n1, int n2, int n3, const double* p, const double* kap,
                         const double* pse, double scale, double* out, double* dot);
int elsim_dev_matvec_f32(int op, int n1, int n2, int n3, const float* p, const float* kap,
                         const float* pse, float scale, float* out, double* dot);
/* Settings of ELSIM_MODE_MIXED_MG, library-wide: the damped-Jacobi weight
 * omega of its smoother (0 < omega < 1), the weight oc of its coarse
 * correction (0 < oc <= 4) and delta, the reduction of the FP32 residual
 * between FP64 reliable updates (0 < delta < 1; also used for the CnE solve
 * of a step in that mode).  Defaults 0.7, 1.5, 1e-3; a hierarchy is rebuilt
 * when omega changes.  No counterpart in the reference (its CG is
 * unpreconditioned). */
int elsim_mg_config(double omega, double oc, double delta);
int elsim_mg_config_get(double* omega, double* oc, double* delta);
/* One V-cycle of the ELSIM_MODE_MIXED_MG preconditioner, z = M^-1 r, on
 * device arrays: kap (and pse for ELSIM_OP_CNE) FP64 ghosted, scale the
 * operator's dt/dx^2 or 1/dx^2; r and z FP32 ghosted (r's ghosts zero, z's are
 * written zero).  The hierarchy is built on every call.  No counterpart in
 * the reference (its CG is unpreconditioned, SBM:657-714, 737-795): it
 * exposes the preconditioner to tests. */
int elsim_dev_mg_apply(int op, int n1, int n2, int n3, const double* kap, const double* pse, double scale,
                       const float* r, float* z);
/* Kernel timing inside the CG loops.  When enabled, the first iteration of
 * every batch of CG iterations is bracketed by HIP events on the library
 * stream (operator, update, direction kernels, V-cycle); read-out averages them over
 * the batches whose iteration ran, then resets.  Slots: ELSIM_PROF_*.
 * elsim_prof_enable(on): 0 off, 1 every solve, 2 + op the solves of
 * operator op (ELSIM_OP_*) only. */
enum {
  ELSIM_PROF_OP32 = 0, ELSIM_PROF_UPD32 = 1, ELSIM_PROF_DIR32 = 2,
  ELSIM_PROF_OP64 = 3, ELSIM_PROF_UPD64 = 4, ELSIM_PROF_DIR64 = 5,
  ELSIM_PROF_MG = 6,      /* the V-cycle of ELSIM_MODE_MIXED_MG, all its kernels */
  ELSIM_PROF_MG_DOWN = 7, /* its fine-level down-sweep (r - A x1, restricted; with the
                             residual update when n1 is a multiple of 4) */
  ELSIM_PROF_MG_UP = 8,   /* its fine-level up-sweep (prolongation, post-smoothing) */
  ELSIM_PROF_N = 9
};
typedef struct elsim_kernel_prof {
  int32_t count[ELSIM_PROF_N];
  double  ms_total[ELSIM_PROF_N];
} elsim_kernel_prof;
int elsim_prof_enable(int on);
int elsim_prof_read(elsim_kernel_prof* out);

/* The same operator with every ghost of p read as stored (the form of the
 * initial residual b - A x, SBM:492-499, 649-655, 729-735).  out: interior. */
int elsim_dev_matvec_stored_f64(int op, int n1, int n2, int n3, const double* p, const double* kap,
                                const double* pse, double scale, double* out);

/* ------------------------------------------------------------------------ */
/* Learned CG warm start (no counterpart in the reference).                 */
/* A 4-level 3-D convolutional encoder-decoder (3x3x3 convolutions,         */
/* channels 16/32/64/64, average pooling, nearest upsampling, concatenated  */
/* skips, ReLU, 1x1x1 output) in bfloat16 with FP32 accumulation on MFMA.   */
/* Tensors on device, channels-last: x[voxel][16] (bf16 bits), voxel =      */
/* i + n1 (j + n2 k), grid dims multiples of 8.  Parameters (FP32, flat):   */
/* for the layers e0 e1 e2 e3 d2 d1 d0 (cin->cout: 16->16, 16->32, 32->64,  */
/* 64->64, 128->64, 96->32, 48->16) W[cout][27][cin] (tap = dx + 3 (dy + 3 dz),*/
/* d1 fastest) then b[cout]; then the output layer w[16], b.                */
/* ------------------------------------------------------------------------ */
int elsim_nn_param_count(void);
int elsim_nn_create(int64_t* handle);
int elsim_nn_destroy(const int64_t* handle);
int elsim_nn_set_params(const int64_t* handle, const float* host_params);
int elsim_nn_get_params(const int64_t* handle, float* host_params);
int elsim_nn_get_grads(const int64_t* handle, float* host_grads);
/* pred[voxel] (FP32, device) of the input features x (device). */
int elsim_nn_forward(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, float* pred);
/* The same with x[voxel][in_used] (in_used 8 or 16), the network's input
 * channels in_used..15 being zero: with 8 (the SBM warm start's features
 * occupy channels 0..4) the first layer runs on 8-channel slabs. */
int elsim_nn_forward_used(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, int in_used,
                          float* pred);
/* Forward, loss = mean((pred - target)^2) (host double) and the gradients of
 * every parameter (kept on device, read with elsim_nn_get_grads). */
int elsim_nn_loss_grad(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, const float* target,
                       float* pred, double* loss);
/* Training with another loss: forward keeping the activations, a loss that
 * leaves dL/dpred on the device, then the backward pass of the last forward.
 * energy_loss: L = e.(-A e)/N with e = pred - target and A the homogeneous
 * phP operator (SBM:670-676 with unit scale) of the ghosted FP32 coefficient
 * kap -- the squared A-norm error that CG minimises. */
int elsim_nn_train_forward(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, float* pred);
int elsim_nn_energy_loss(const int64_t* handle, const float* pred, const float* target, const float* kap,
                         double* loss);
int elsim_nn_backward(const int64_t* handle, const uint16_t* x);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the last gradients. */
int elsim_nn_adam(const int64_t* handle, float lr);
/* Attach a network to an SBM state (nn NULL or 0: detach) for
 * ELSIM_WS_LEARNED: CG then solves to the same tolerance from a start that is
 * never worse in the A-norm than the extrapolation, so the network can only
 * change the iteration count.  An unknown network handle is ELSIM_ERR_ARG;
 * destroying an attached network detaches it from every state. */
int elsim_sbm_set_warm_start(const int64_t* sbm_handle, const int64_t* nn_handle);
/* Training data: during the next step (ELSIM_WS_LEARNED, from the third
 * step on) write the network input (bf16 [voxel][16]: psi_P, the start's
 * FP64 residual / its rms, the through-thickness coordinate, log10(kappa)/7,
 * the last change / its rms) to x_dev and the extrapolated start's interior
 * (FP64 [voxel]) to start_dev; the target is the converged phi_P minus that
 * start.  x_dev NULL disarms. */
int elsim_sbm_ws_record(const int64_t* sbm_handle, uint16_t* x_dev, double* start_dev);
/* Single layers on device data (parity tests).  conv3: out[v][co] =
 * bf16(relu?(b[co] + sum_{tap,ci} wt[co][tap cin + ci] x[v + o_tap][ci]));
 * bwd_data: dx[v][ci] (FP32) = sum_{tap,co} wb[ci][tap cout + co] dz[v + o_tap][co]
 * with wb the tap-flipped, transposed weights; bwd_weight: dw[co][tap][ci]
 * (FP32) = sum_v dz[v][co] x[v + o_tap][ci].  Zero padding, stride 1.
 * conv3: cin 8 (then cout <= 32) or a multiple of 16, cout <= 64. */
int elsim_nn_conv3(int n1, int n2, int n3, const uint16_t* x, int cin, const uint16_t* wt, const float* bias,
                   int cout, int relu, uint16_t* out);
int elsim_nn_conv3_bwd_data(int n1, int n2, int n3, const uint16_t* dz, int cout, const uint16_t* wb, int cin,
                            float* dx);
int elsim_nn_conv3_bwd_weight(int n1, int n2, int n3, const uint16_t* x, int cin, const uint16_t* dz, int cout,
                              float* dw);

/* ------------------------------------------------------------------------ */
/* Synthetic inputs (host only; deterministic for a seed)                   */
/* ------------------------------------------------------------------------ */
/* Random spheres, radius ~ U[rmin,rmax], centres ~ U[0,n)^3, with a hard
 * core: centres at least hardcore*(r1+r2) apart (1 = non-overlapping), added
 * until the voxel solid fraction reaches `target` or attempts run out.
 * sdf (interior FP32) = max_s (r_s - |x - c_s|), clamped below at -dcut.
 * connect != 0 drops particles not 6-connected to the high d2 face and fills
 * electrolyte pockets not connected to the low d2 face. */
int elsim_gen_sphere_pack(int n1, int n2, int n3, uint64_t seed, double rmin, double rmax,
                          double target, double hardcore, int connect, double dcut, float* sdf,
                          int* nspheres, double* frac);
/* Random ellipsoids, semi-axes ~ U[amin,amax], major axis within `cone_deg`
 * of the d2 axis (the through-thickness axis), added until the solid
 * fraction reaches `target`; a shape is rejected when more than `overlap` of
 * its voxels are already solid (0: voxel-level non-overlap, 1: free
 * overlap).  Approximate SDF (1 - normalised radius) * smallest semi-axis,
 * exact for spheres, max over shapes, clamped below at -dcut. */
int elsim_gen_ellipsoid_pack(int n1, int n2, int n3, uint64_t seed, double amin, double amax,
                             double cone_deg, double target, double overlap, double dcut, float* sdf,
                             int* nshapes, double* frac);
/* Carve a random-walk tunnel of radius `radius` through the solid from the
 * low to the high d2 face: sdf <- min(sdf, dist(path) - radius). */
int elsim_gen_tunnel(int n1, int n2, int n3, uint64_t seed, double radius, float* sdf);
/* Synthetic 101-point property tables (mu, i0, OCV, mobility) on [0,1]. */
int elsim_gen_tables(int sz, double* tables);

#ifdef __cplusplus
}
#endif
#endif /* ELSIM_H */
