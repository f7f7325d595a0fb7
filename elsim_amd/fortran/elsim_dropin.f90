!! Drop-in replacements for the reference's subroutines, same names and
!! argument lists, running on the GPU through libelsim_hip.so:
!!   NoGrdBc3D       LevelSet_3D.f90:164-190  (RM_UCNCT_3D.f90:180-205)
!!   AbsGrd3DUW_XYZ  LevelSet_3D.f90:265-322  (RM_UCNCT_3D.f90:280-337)
!!   ReDist3D_XYZ    LevelSet_3D.f90:194-261  (dt = 0.1d0; RM_UCNCT_3D.f90:209-276
!!                   uses REAL(4) 0.1: link ReDist3D_XYZ_R4 there)
!!   Tbl_Chk_Pln_3DA SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:987-1028
!! and the subroutines carved out of the inline blocks of the SBM and AD
!! programs (SBM:397-840, AccDiff_3D.f90:103-243, RM_UCNCT_3D.f90:79-82):
!!   SBM_Init / SBM_Step / SBM_Get / SBM_Set / SBM_Free, AccDiff_Phase,
!!   Label_Connected, RM_Threshold; SBM_Solve_Pot, one potential solve of the
!!   SBM program's inline blocks (SBM:643-723, 725-804); LS_FrontEnd, the inline geometry
!!   front-end of LevelSet_3D's frame loop (LevelSet_3D.f90:84-144); RM_Phase,
!!   the sign block of one phase of RM_UCNCT_3D (RM_UCNCT_3D.f90:64-96,
!!   115-136); and Refine2_SDF, the 2x refinement to FP32 that the reference
!!   runs between RM_UCNCT_3D and the SBM program but does not contain.
!! A program that linked the reference subroutines links these instead.

SUBROUTINE NoGrdBc3D(py, px, pz, f)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: f
   CALL elsim_check(elsim_nogrdbc3d(py, px, pz, f), 'NoGrdBc3D')
END SUBROUTINE NoGrdBc3D

SUBROUTINE AbsGrd3DUW_XYZ(py, px, pz, sg, f, gr)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(IN), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sg, f
   REAL(KIND=8), INTENT(OUT), DIMENSION(py,px,pz) :: gr
   CALL elsim_check(elsim_absgrd3duw_xyz(py, px, pz, sg, f, gr), 'AbsGrd3DUW_XYZ')
END SUBROUTINE AbsGrd3DUW_XYZ

SUBROUTINE ReDist3D_XYZ(lit, py, px, pz, sg, f)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: lit, py, px, pz
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sg, f
   CALL elsim_check(elsim_redist3d_xyz(lit, py, px, pz, sg, f, 1.0d-1), 'ReDist3D_XYZ')
END SUBROUTINE ReDist3D_XYZ

SUBROUTINE ReDist3D_XYZ_R4(lit, py, px, pz, sg, f)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: lit, py, px, pz
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sg, f
   REAL(KIND=8) :: dt
   dt = 1.0e-1
   CALL elsim_check(elsim_redist3d_xyz(lit, py, px, pz, sg, f, dt), 'ReDist3D_XYZ')
END SUBROUTINE ReDist3D_XYZ_R4

!! LevelSet_3D.f90:84-144 for one tunnel radius tr: sgF and the smoothed dsF
!! on the ghosted grid; the frame loop then sets dsF = 0 and calls
!! ReDist3D_XYZ(175, py, px, pz, sgF, dsF).
SUBROUTINE LS_FrontEnd(yo, xo, zo, TmpA, py, px, pz, sep, tr, nac, sgF, dsF)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: yo, xo, zo, py, px, pz, sep, nac
   REAL(KIND=8), INTENT(IN), DIMENSION(yo,xo,zo) :: TmpA
   REAL(KIND=8), INTENT(IN) :: tr
   REAL(KIND=8), INTENT(OUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sgF, dsF
   CALL elsim_check(elsim_ls_frontend(yo, xo, zo, TmpA, py, px, pz, sep, tr, nac, sgF, dsF), 'LS_FrontEnd')
END SUBROUTINE LS_FrontEnd

!! RM_UCNCT_3D.f90:64-96 (phase 0, solid) / 115-136 (phase 1, electrolyte):
!! sgF from the L24Z distance function and that phase's connectivity; the
!! program then sets dsF = 0 and calls ReDist3D_XYZ_R4(175, ...).
SUBROUTINE RM_Phase(py, px, pz, RdIn, Cn, phase, sgF)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz, phase
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: RdIn, Cn
   REAL(KIND=8), INTENT(OUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sgF
   CALL elsim_check(elsim_rm_phase(py, px, pz, RdIn, Cn, phase, sgF), 'RM_Phase')
END SUBROUTINE RM_Phase

!! GRP_LS / POR_LS (py x px x pz, FP64) -> the FP32 distance function on the
!! 2py x 2px x 2pz grid that the SBM program reads (SBM:138-152).
SUBROUTINE Refine2_SDF(py, px, pz, Sdf, Out)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: Sdf
   REAL(KIND=4), INTENT(OUT), DIMENSION(2*py,2*px,2*pz) :: Out
   CALL elsim_check(elsim_refine2_sdf(py, px, pz, Sdf, Out), 'Refine2_SDF')
END SUBROUTINE Refine2_SDF

SUBROUTINE Tbl_Chk_Pln_3DA(py, px, pz, sz, Cord, Vals, Cn, fn)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: px, py, pz, sz
   REAL(KIND=8), INTENT(IN), DIMENSION(sz) :: Cord, Vals
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: Cn
   REAL(KIND=8), INTENT(OUT), DIMENSION(py,px,pz) :: fn
   CALL elsim_check(elsim_tbl_chk_pln_3da(py, px, pz, sz, Cord, Vals, Cn, fn), 'Tbl_Chk_Pln_3DA')
END SUBROUTINE Tbl_Chk_Pln_3DA

!! One phase of AccDiff_3D: phase 0 = solid, 1 = electrolyte; term = 1501 in
!! the reference.  Cn is the FP64 field the reference writes (AD:166, AD:240).
SUBROUTINE AccDiff_Phase(py, px, pz, RdIn, phase, term, Cn)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz, phase, term
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: RdIn
   REAL(KIND=8), INTENT(OUT), DIMENSION(py,px,pz) :: Cn
   CALL elsim_check(elsim_accdiff_phase(py, px, pz, RdIn, phase, term, Cn), 'AccDiff_Phase')
END SUBROUTINE AccDiff_Phase

SUBROUTINE Label_Connected(py, px, pz, RdIn, phase, Labels, Conn)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz, phase
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: RdIn
   INTEGER(KIND=4), INTENT(OUT), DIMENSION(py,px,pz) :: Labels, Conn
   CALL elsim_check(elsim_label_connected(py, px, pz, RdIn, phase, Labels, Conn), 'Label_Connected')
END SUBROUTINE Label_Connected

!! RM_UCNCT_3D.f90:79-82: psB = psA where Cn >= 1e-4, else 0.
SUBROUTINE RM_Threshold(py, px, pz, Cn, psA, psB)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: Cn, psA
   REAL(KIND=8), INTENT(OUT), DIMENSION(py,px,pz) :: psB
   CALL elsim_check(elsim_rm_threshold(py, px, pz, Cn, psA, psB), 'RM_Threshold')
END SUBROUTINE RM_Threshold

!! SBM set-up (SBM:143-341): psi from the FP32 SDFs, interface density,
!! initial fields; the state stays on the GPU under `handle`.
SUBROUTINE SBM_Init(ny, nx, nz, prm, DmyS, DmyL, sz, Tbls, handle)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: ny, nx, nz, sz
   TYPE(elsim_sbm_params), INTENT(IN) :: prm
   REAL(KIND=4), INTENT(IN), DIMENSION(ny,nx,nz) :: DmyS, DmyL
   REAL(KIND=8), INTENT(IN), DIMENSION(sz,8) :: Tbls
   INTEGER(KIND=8), INTENT(OUT) :: handle
   CALL elsim_check(elsim_init(0_c_int), 'elsim_init')
   CALL elsim_check(elsim_sbm_create(ny, nx, nz, prm, DmyS, DmyL, sz, Tbls, handle), 'SBM_Init')
END SUBROUTINE SBM_Init

!! One electrochemical step SBM:397-840; mode = ELSIM_MODE_FP64 / _MIXED / _FP64_PCG / _MIXED_MG.
SUBROUTINE SBM_Step(handle, mode, stats)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER(KIND=8), INTENT(IN) :: handle
   INTEGER, INTENT(IN) :: mode
   TYPE(elsim_step_stats), INTENT(OUT) :: stats
   INTEGER(c_int) :: rc
   rc = elsim_sbm_step(handle, mode, stats)
   CALL elsim_check(rc, 'SBM_Step')
END SUBROUTINE SBM_Step

!! Copy a ghosted (0:ny+1,0:nx+1,0:nz+1) or interior field back to the host.
SUBROUTINE SBM_Get(handle, field, arr)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER(KIND=8), INTENT(IN) :: handle
   INTEGER, INTENT(IN) :: field
   REAL(KIND=8), INTENT(OUT) :: arr(*)
   CALL elsim_check(elsim_sbm_get(handle, field, arr), 'SBM_Get')
END SUBROUTINE SBM_Get

SUBROUTINE SBM_Set(handle, field, arr)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER(KIND=8), INTENT(IN) :: handle
   INTEGER, INTENT(IN) :: field
   REAL(KIND=8), INTENT(IN) :: arr(*)
   CALL elsim_check(elsim_sbm_set(handle, field, arr), 'SBM_Set')
END SUBROUTINE SBM_Set

!! One potential solve (SBM:657-714 for phP, 737-795 for phE): kap, phi
!! ghosted (0:py+1,0:px+1,0:pz+1), rhs interior; bclo / bchi = 0 Neumann, 1
!! Dirichlet on the low / high d2 ghost plane (the value phi holds there).
!! Errors and a solve that hits maxit end the program.
SUBROUTINE SBM_Solve_Pot(py, px, pz, dx, kap, rhs, phi, bclo, bchi, mode, rtol, maxit, iters, relres)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz, bclo, bchi, mode, maxit
   REAL(KIND=8), INTENT(IN) :: dx, rtol
   REAL(KIND=8), INTENT(IN), DIMENSION(0:py+1,0:px+1,0:pz+1) :: kap
   REAL(KIND=8), INTENT(IN), DIMENSION(py,px,pz) :: rhs
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: phi
   INTEGER, INTENT(OUT) :: iters
   REAL(KIND=8), INTENT(OUT) :: relres
   CALL elsim_check_converged(elsim_solve_pot(py, px, pz, dx, kap, rhs, phi, bclo, bchi, mode, rtol, maxit, &
      iters, relres), 'SBM_Solve_Pot')
END SUBROUTINE SBM_Solve_Pot

SUBROUTINE SBM_Free(handle)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER(KIND=8), INTENT(IN) :: handle
   CALL elsim_check(elsim_sbm_destroy(handle), 'SBM_Free')
END SUBROUTINE SBM_Free
