!! geom_run -- Fortran host driver of the geometry programs on one GPU, one
!! frame: LevelSet_3D (LevelSet_3D.f90:59-151), AccDiff_3D for both phases
!! (AccDiff_3D.f90:68-243), RM_UCNCT_3D for both phases (RM_UCNCT_3D.f90:
!! 48-168), and the 2x refinement to FP32 that produces the SBM program's
!! inputs (SBM:138-152, not in the reference), through the drop-in
!! subroutines, with the reference's file formats and names.
!!
!! Usage: geom_run <namelist file>
!!   &geom image='img.txt', yo=, xo=, zo=, py=, px=, pz=, sep=, nac=50,
!!         fr=5, tr=4.5, lit=175, term=1501, outdir='.' /
!! image: the voxel image as list-directed text, yo*xo*zo values in column
!! order (LevelSet_3D.f90:60-63).  Outputs, FP64 unformatted records unless
!! noted (<g> = py x px x pz, <f> = 100 + fr):
!!   L24Z<1000+fr>.dat               LevelSet_3D, OutPut3D AB=2 (LS:150, 326-374)
!!   SOD_3D_CN_<g>_Z_II_<f>.dat      AccDiff_3D solid connectivity (AD:159-166)
!!   LIQ_3D_CN_<g>_Z_II_<f>.dat      AccDiff_3D electrolyte connectivity (AD:235-242)
!!   GRP_LS_<g>_Z_II_<f>.dat         RM_UCNCT_3D solid distance function (RM:101-106)
!!   POR_LS_<g>_Z_II_<f>.dat         RM_UCNCT_3D electrolyte distance function, in the
!!                                   unformatted form of RM:152-157 (the reference
!!                                   writes it as text, RM:160-165)
!!   GrMic_II_DS_<2g>_B<f>.dat,      the refined FP32 solid / electrolyte distance
!!   GrMic_II_DL_<2g>_B<f>.dat       functions, the SBM program's input records
PROGRAM geom_run
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER :: yo, xo, zo, py, px, pz, sep, nac, fr, lit, term, u, ph
   REAL(KIND=8) :: tr
   CHARACTER(LEN=512) :: image, outdir, nml
   CHARACTER(LEN=64) :: g, g2
   CHARACTER(LEN=4) :: ser
   CHARACTER(LEN=3) :: frm
   REAL(KIND=8), ALLOCATABLE :: TmpA(:,:,:), sgF(:,:,:), dsF(:,:,:), RdIn(:,:,:), CnS(:,:,:), CnL(:,:,:), &
      Ls(:,:,:)
   REAL(KIND=4), ALLOCATABLE :: Fine(:,:,:)
   NAMELIST /geom/ image, yo, xo, zo, py, px, pz, sep, nac, fr, tr, lit, term, outdir

   nac = 50; fr = 5; tr = 4.5d0; lit = 175; term = 1501; outdir = '.'; sep = 0
   IF (COMMAND_ARGUMENT_COUNT() < 1) THEN
      WRITE (0, *) 'usage: geom_run <namelist file>'
      ERROR STOP 1
   END IF
   CALL GET_COMMAND_ARGUMENT(1, nml)
   OPEN (NEWUNIT=u, FILE=TRIM(nml), STATUS='OLD', ACTION='READ')
   READ (u, NML=geom)
   CLOSE (u)
   WRITE (g, "(I0,'x',I0,'x',I0)") py, px, pz
   WRITE (g2, "(I0,'x',I0,'x',I0)") 2*py, 2*px, 2*pz
   WRITE (ser, "(I4)") 1000 + fr
   WRITE (frm, "(I3)") 100 + fr

   ALLOCATE (TmpA(yo,xo,zo), sgF(0:py+1,0:px+1,0:pz+1), dsF(0:py+1,0:px+1,0:pz+1), RdIn(py,px,pz), &
      CnS(py,px,pz), CnL(py,px,pz), Ls(py,px,pz), Fine(2*py,2*px,2*pz))
   OPEN (NEWUNIT=u, FILE=TRIM(image), FORM='FORMATTED', STATUS='OLD', ACTION='READ')
   READ (u, *) TmpA
   CLOSE (u)

   !! LevelSet_3D, one frame (LS:76-151)
   CALL LS_FrontEnd(yo, xo, zo, TmpA, py, px, pz, sep, tr, nac, sgF, dsF)
   dsF = 0.0
   CALL ReDist3D_XYZ(lit, py, px, pz, sgF, dsF)
   RdIn = dsF(1:py,1:px,1:pz)
   CALL put(TRIM(outdir)//'/L24Z'//ser//'.dat', RdIn)
   WRITE (*, '(A)') 'LevelSet_3D done'

   !! AccDiff_3D, both phases (AD:103-243)
   CALL AccDiff_Phase(py, px, pz, RdIn, 0, term, CnS)
   CALL put(TRIM(outdir)//'/SOD_3D_CN_'//TRIM(g)//'_Z_II_'//frm//'.dat', CnS)
   CALL AccDiff_Phase(py, px, pz, RdIn, 1, term, CnL)
   CALL put(TRIM(outdir)//'/LIQ_3D_CN_'//TRIM(g)//'_Z_II_'//frm//'.dat', CnL)
   WRITE (*, '(A)') 'AccDiff_3D done'

   !! RM_UCNCT_3D, both phases (RM:56-165), then the refinement
   DO ph = 0, 1
      IF (ph == 0) THEN
         CALL RM_Phase(py, px, pz, RdIn, CnS, 0, sgF)
      ELSE
         CALL RM_Phase(py, px, pz, RdIn, CnL, 1, sgF)
      END IF
      dsF = 0.0
      CALL ReDist3D_XYZ_R4(lit, py, px, pz, sgF, dsF)
      Ls = dsF(1:py,1:px,1:pz)
      CALL Refine2_SDF(py, px, pz, Ls, Fine)
      IF (ph == 0) THEN
         CALL put(TRIM(outdir)//'/GRP_LS_'//TRIM(g)//'_Z_II_'//frm//'.dat', Ls)
         CALL put4(TRIM(outdir)//'/GrMic_II_DS_'//TRIM(g2)//'_B'//frm//'.dat', Fine)
      ELSE
         CALL put(TRIM(outdir)//'/POR_LS_'//TRIM(g)//'_Z_II_'//frm//'.dat', Ls)
         CALL put4(TRIM(outdir)//'/GrMic_II_DL_'//TRIM(g2)//'_B'//frm//'.dat', Fine)
      END IF
   END DO
   WRITE (*, '(A)') 'RM_UCNCT_3D and refinement done'

CONTAINS

   SUBROUTINE put(name, a)
      CHARACTER(LEN=*), INTENT(IN) :: name
      REAL(KIND=8), INTENT(IN) :: a(:,:,:)
      INTEGER :: v
      OPEN (NEWUNIT=v, FILE=name, FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
      WRITE (v) a
      CLOSE (v)
   END SUBROUTINE put

   SUBROUTINE put4(name, a)
      CHARACTER(LEN=*), INTENT(IN) :: name
      REAL(KIND=4), INTENT(IN) :: a(:,:,:)
      INTEGER :: v
      OPEN (NEWUNIT=v, FILE=name, FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
      WRITE (v) a
      CLOSE (v)
   END SUBROUTINE put4

END PROGRAM geom_run
