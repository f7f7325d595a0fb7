!! Drop-in replacements for the level-set subroutines of RM_UCNCT_3D.f90,
!! same names and argument lists as that program's own copies, running on the
!! GPU through libelsim_hip.so.  Built into libelsim_rm.so, apart from
!! libelsim_fortran.so, because RM_UCNCT_3D's ReDist3D_XYZ differs from
!! LevelSet_3D's in one literal: its pseudo-time step is the default-REAL
!! 0.1 (RM_UCNCT_3D.f90:236, 0.10000000149011612 as REAL(8)) where
!! LevelSet_3D.f90:221 has 0.1d0.  RM's calls (RM_UCNCT_3D.f90:100, 148) link
!! against this library unchanged.
!!   NoGrdBc3D       RM_UCNCT_3D.f90:180-205
!!   ReDist3D_XYZ    RM_UCNCT_3D.f90:209-276
!!   AbsGrd3DUW_XYZ  RM_UCNCT_3D.f90:280-337

SUBROUTINE NoGrdBc3D(py, px, pz, f)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: f
   CALL elsim_check(elsim_nogrdbc3d(py, px, pz, f), 'NoGrdBc3D')
END SUBROUTINE NoGrdBc3D

SUBROUTINE ReDist3D_XYZ(lit, py, px, pz, sg, f)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: lit, py, px, pz
   REAL(KIND=8), INTENT(INOUT), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sg, f
   REAL(KIND=8) :: dt
   dt = 1.0e-1
   CALL elsim_check(elsim_redist3d_xyz(lit, py, px, pz, sg, f, dt), 'ReDist3D_XYZ')
END SUBROUTINE ReDist3D_XYZ

SUBROUTINE AbsGrd3DUW_XYZ(py, px, pz, sg, f, gr)
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER, INTENT(IN) :: py, px, pz
   REAL(KIND=8), INTENT(IN), DIMENSION(0:py+1,0:px+1,0:pz+1) :: sg, f
   REAL(KIND=8), INTENT(OUT), DIMENSION(py,px,pz) :: gr
   CALL elsim_check(elsim_absgrd3duw_xyz(py, px, pz, sg, f, gr), 'AbsGrd3DUW_XYZ')
END SUBROUTINE AbsGrd3DUW_XYZ
