!! rm_check -- calls ReDist3D_XYZ as RM_UCNCT_3D does (RM_UCNCT_3D.f90:100,
!! 148), linked against libelsim_rm.so, on the records sg.dat and f.dat in
!! <dir> (ghosted); writes f_out.dat.
!! Usage: rm_check <dir> lit py px pz
PROGRAM rm_check
   IMPLICIT NONE
   CHARACTER(LEN=512) :: d, arg
   INTEGER :: lit, py, px, pz, v
   REAL(KIND=8), ALLOCATABLE :: sg(:,:,:), f(:,:,:)
   CALL GET_COMMAND_ARGUMENT(1, d)
   CALL GET_COMMAND_ARGUMENT(2, arg); READ (arg, *) lit
   CALL GET_COMMAND_ARGUMENT(3, arg); READ (arg, *) py
   CALL GET_COMMAND_ARGUMENT(4, arg); READ (arg, *) px
   CALL GET_COMMAND_ARGUMENT(5, arg); READ (arg, *) pz
   ALLOCATE (sg(0:py+1,0:px+1,0:pz+1), f(0:py+1,0:px+1,0:pz+1))
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/sg.dat', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (v) sg
   CLOSE (v)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/f.dat', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (v) f
   CLOSE (v)
   CALL ReDist3D_XYZ(lit, py, px, pz, sg, f)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/f_out.dat', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
   WRITE (v) f
   CLOSE (v)
END PROGRAM rm_check
