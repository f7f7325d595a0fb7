!! elsim_gpu -- iso_c_binding interfaces of libelsim_hip.so (include/elsim.h).
!! All arguments are passed by reference unless marked VALUE, so Fortran
!! arrays in the reference layout (0:n1+1, 0:n2+1, 0:n3+1) go straight through.
MODULE elsim_gpu
   USE, INTRINSIC :: iso_c_binding
   IMPLICIT NONE

   INTEGER(c_int), PARAMETER :: ELSIM_OK = 0, ELSIM_NOT_CONVERGED = 1
   INTEGER(c_int), PARAMETER :: ELSIM_STOP_REFERENCE = 0, ELSIM_STOP_RELRES = 1
   INTEGER(c_int), PARAMETER :: ELSIM_MODE_FP64 = 0, ELSIM_MODE_MIXED = 1, ELSIM_MODE_FP64_PCG = 2, &
      ELSIM_MODE_MIXED_MG = 3
   INTEGER(c_int), PARAMETER :: ELSIM_F_PSP = 0, ELSIM_F_PSE = 1, ELSIM_F_AVP = 2, ELSIM_F_CNP = 3, &
      ELSIM_F_CNE = 4, ELSIM_F_PHP = 5, ELSIM_F_PHE = 6, ELSIM_F_RXN = 7, ELSIM_F_KAPP = 8

   TYPE, BIND(C) :: elsim_sbm_params
      REAL(c_double) :: dx, dt, zeta, psp_off, pse_off, avp_min, eps, rho, ch_psi_min, frd, cst1, t_minus, &
         d0, tc1, tc2, dab_c0, dab_c1, dab_c2, dab_c3, kap_scale, i0_scale, kfw_cref, bvp, bve, vsr, trgi, &
         cne_target, rtol, crt_e, crt_p, crt_l, outer_tol, inner_rtol
      INTEGER(c_int32_t) :: mass_corr_period, skip_cne_first, stop_rule, max_it_e, max_it_p, max_it_l, &
         max_outer, ws_mode, cne_rule, ws_order
   END TYPE elsim_sbm_params

   TYPE, BIND(C) :: elsim_step_stats
      INTEGER(c_int32_t) :: it_cne, it_php, it_phe, n_outer, n_reliable, status
      REAL(c_double) :: xrf, scrnt, cvgp, cvgl, relres_cne, relres_php, relres_phe, seconds
      INTEGER(c_int32_t) :: ws_used, ws_pad_
      REAL(c_double) :: ws_res_plain, ws_res_nn, ws_alpha, ws_energy_gain
   END TYPE elsim_step_stats

   INTERFACE
      INTEGER(c_int) FUNCTION elsim_init(device) BIND(C, name="elsim_init")
         IMPORT :: c_int
         INTEGER(c_int), VALUE :: device
      END FUNCTION
      TYPE(c_ptr) FUNCTION elsim_last_error() BIND(C, name="elsim_last_error")
         IMPORT :: c_ptr
      END FUNCTION
      SUBROUTINE elsim_sbm_default_params(p) BIND(C, name="elsim_sbm_default_params")
         IMPORT :: elsim_sbm_params
         TYPE(elsim_sbm_params), INTENT(OUT) :: p
      END SUBROUTINE
      INTEGER(c_int) FUNCTION elsim_nogrdbc3d(py, px, pz, f) BIND(C, name="elsim_nogrdbc3d")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz
         REAL(c_double), INTENT(INOUT) :: f(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_absgrd3duw_xyz(py, px, pz, sg, f, gr) BIND(C, name="elsim_absgrd3duw_xyz")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz
         REAL(c_double), INTENT(IN) :: sg(*), f(*)
         REAL(c_double), INTENT(OUT) :: gr(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_redist3d_xyz(lit, py, px, pz, sg, f, dt) BIND(C, name="elsim_redist3d_xyz")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: lit, py, px, pz
         REAL(c_double), INTENT(IN) :: sg(*), dt
         REAL(c_double), INTENT(INOUT) :: f(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_ls_frontend(yo, xo, zo, tmpa, py, px, pz, sep, tr, n_ac, sgf, dsf) &
         BIND(C, name="elsim_ls_frontend")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: yo, xo, zo, py, px, pz, sep, n_ac
         REAL(c_double), INTENT(IN) :: tmpa(*), tr
         REAL(c_double), INTENT(OUT) :: sgf(*), dsf(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_rm_phase(py, px, pz, rdin, cn, phase, sgf) BIND(C, name="elsim_rm_phase")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz, phase
         REAL(c_double), INTENT(IN) :: rdin(*), cn(*)
         REAL(c_double), INTENT(OUT) :: sgf(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_refine2_sdf(py, px, pz, sdf, out) BIND(C, name="elsim_refine2_sdf")
         IMPORT :: c_int, c_double, c_float
         INTEGER(c_int), INTENT(IN) :: py, px, pz
         REAL(c_double), INTENT(IN) :: sdf(*)
         REAL(c_float), INTENT(OUT) :: out(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_tbl_chk_pln_3da(py, px, pz, sz, cord, vals, cn, fn) &
         BIND(C, name="elsim_tbl_chk_pln_3da")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz, sz
         REAL(c_double), INTENT(IN) :: cord(*), vals(*), cn(*)
         REAL(c_double), INTENT(OUT) :: fn(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_accdiff_phase(py, px, pz, rdin, phase, term, cn) &
         BIND(C, name="elsim_accdiff_phase")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz, phase, term
         REAL(c_double), INTENT(IN) :: rdin(*)
         REAL(c_double), INTENT(OUT) :: cn(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_label_connected(py, px, pz, rdin, phase, labels, connected) &
         BIND(C, name="elsim_label_connected")
         IMPORT :: c_int, c_double, c_int32_t
         INTEGER(c_int), INTENT(IN) :: py, px, pz, phase
         REAL(c_double), INTENT(IN) :: rdin(*)
         INTEGER(c_int32_t), INTENT(OUT) :: labels(*), connected(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_rm_threshold(py, px, pz, cn, psi_in, psi_out) BIND(C, name="elsim_rm_threshold")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: py, px, pz
         REAL(c_double), INTENT(IN) :: cn(*), psi_in(*)
         REAL(c_double), INTENT(OUT) :: psi_out(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_create(n1, n2, n3, prm, sdf_s, sdf_l, sz, tables, handle) &
         BIND(C, name="elsim_sbm_create")
         IMPORT :: c_int, c_float, c_double, c_int64_t, elsim_sbm_params
         INTEGER(c_int), INTENT(IN) :: n1, n2, n3, sz
         TYPE(elsim_sbm_params), INTENT(IN) :: prm
         REAL(c_float), INTENT(IN) :: sdf_s(*), sdf_l(*)
         REAL(c_double), INTENT(IN) :: tables(*)
         INTEGER(c_int64_t), INTENT(OUT) :: handle
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_destroy(handle) BIND(C, name="elsim_sbm_destroy")
         IMPORT :: c_int, c_int64_t
         INTEGER(c_int64_t), INTENT(IN) :: handle
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_get(handle, field, out) BIND(C, name="elsim_sbm_get")
         IMPORT :: c_int, c_int64_t, c_double
         INTEGER(c_int64_t), INTENT(IN) :: handle
         INTEGER(c_int), INTENT(IN) :: field
         REAL(c_double), INTENT(OUT) :: out(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_set(handle, field, in) BIND(C, name="elsim_sbm_set")
         IMPORT :: c_int, c_int64_t, c_double
         INTEGER(c_int64_t), INTENT(IN) :: handle
         INTEGER(c_int), INTENT(IN) :: field
         REAL(c_double), INTENT(IN) :: in(*)
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_step(handle, mode, stats) BIND(C, name="elsim_sbm_step")
         IMPORT :: c_int, c_int64_t, elsim_step_stats
         INTEGER(c_int64_t), INTENT(IN) :: handle
         INTEGER(c_int), INTENT(IN) :: mode
         TYPE(elsim_step_stats), INTENT(OUT) :: stats
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_solve_pot(n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit, &
         iters, relres) BIND(C, name="elsim_solve_pot")
         IMPORT :: c_int, c_double
         INTEGER(c_int), INTENT(IN) :: n1, n2, n3, bc_lo, bc_hi, mode, maxit
         REAL(c_double), INTENT(IN) :: dx, rtol, kap(*), rhs(*)
         REAL(c_double), INTENT(INOUT) :: phi(*)
         INTEGER(c_int), INTENT(OUT) :: iters
         REAL(c_double), INTENT(OUT) :: relres
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_scalars(handle, tpp, tpe, xrf, bve, iter) BIND(C, name="elsim_sbm_scalars")
         IMPORT :: c_int, c_int64_t, c_double
         INTEGER(c_int64_t), INTENT(IN) :: handle
         REAL(c_double), INTENT(OUT) :: tpp, tpe, xrf, bve
         INTEGER(c_int), INTENT(OUT) :: iter
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_set_params(handle, prm) BIND(C, name="elsim_sbm_set_params")
         IMPORT :: c_int, c_int64_t, elsim_sbm_params
         INTEGER(c_int64_t), INTENT(IN) :: handle
         TYPE(elsim_sbm_params), INTENT(IN) :: prm
      END FUNCTION
      INTEGER(c_int) FUNCTION elsim_sbm_time_control(xrf, scrnt, trgi, vsr0, dt, vsr) &
         BIND(C, name="elsim_sbm_time_control")
         IMPORT :: c_int, c_double
         REAL(c_double), INTENT(IN) :: xrf, scrnt, trgi, vsr0
         REAL(c_double), INTENT(OUT) :: dt, vsr
      END FUNCTION
   END INTERFACE

CONTAINS

   !! The reference has no error convention (subroutines return nothing and
   !! MPI error codes are ignored); a failing GPU call stops the program.
   SUBROUTINE elsim_check(rc, what)
      INTEGER(c_int), INTENT(IN) :: rc
      CHARACTER(LEN=*), INTENT(IN) :: what
      CHARACTER(KIND=c_char), POINTER :: msg(:)
      INTEGER :: n
      IF (rc >= 0) RETURN
      CALL c_f_pointer(elsim_last_error(), msg, [1024])
      n = 0
      DO WHILE (n < 1024)
         IF (msg(n + 1) == c_null_char) EXIT
         n = n + 1
      END DO
      WRITE (0, '(A,A,I0,A,1024A)') what, ' failed with status ', rc, ': ', msg(1:n)
      ERROR STOP 2
   END SUBROUTINE elsim_check

   !! As elsim_check, and a solve that stopped at its iteration cap also
   !! stops the program (the reference leaves such a loop silently,
   !! SBM:551, 710, 791).
   SUBROUTINE elsim_check_converged(rc, what)
      INTEGER(c_int), INTENT(IN) :: rc
      CHARACTER(LEN=*), INTENT(IN) :: what
      CALL elsim_check(rc, what)
      IF (rc == ELSIM_NOT_CONVERGED) THEN
         WRITE (0, '(A,A)') what, ' did not converge within its iteration cap'
         ERROR STOP 3
      END IF
   END SUBROUTINE elsim_check_converged

END MODULE elsim_gpu
