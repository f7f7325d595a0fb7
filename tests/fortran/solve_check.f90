!! solve_check -- one potential solve through the Fortran drop-in
!! SBM_Solve_Pot on the records kap.dat (ghosted), rhs.dat (interior) and
!! phi.dat (ghosted, Dirichlet values in the d2 ghost planes) in <dir>;
!! writes phi_out.dat and a text line "iters relres".  An iteration cap or
!! an error ends the program through ERROR STOP (non-zero exit status).
!! Usage: solve_check <dir> py px pz bclo bchi mode maxit
PROGRAM solve_check
   IMPLICIT NONE
   CHARACTER(LEN=512) :: d, arg
   INTEGER :: py, px, pz, bclo, bchi, mode, maxit, iters, v
   REAL(KIND=8) :: relres
   REAL(KIND=8), ALLOCATABLE :: kap(:,:,:), rhs(:,:,:), phi(:,:,:)
   CALL GET_COMMAND_ARGUMENT(1, d)
   CALL GET_COMMAND_ARGUMENT(2, arg); READ (arg, *) py
   CALL GET_COMMAND_ARGUMENT(3, arg); READ (arg, *) px
   CALL GET_COMMAND_ARGUMENT(4, arg); READ (arg, *) pz
   CALL GET_COMMAND_ARGUMENT(5, arg); READ (arg, *) bclo
   CALL GET_COMMAND_ARGUMENT(6, arg); READ (arg, *) bchi
   CALL GET_COMMAND_ARGUMENT(7, arg); READ (arg, *) mode
   CALL GET_COMMAND_ARGUMENT(8, arg); READ (arg, *) maxit
   ALLOCATE (kap(0:py+1,0:px+1,0:pz+1), rhs(py,px,pz), phi(0:py+1,0:px+1,0:pz+1))
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/kap.dat', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (v) kap
   CLOSE (v)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/rhs.dat', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (v) rhs
   CLOSE (v)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/phi.dat', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (v) phi
   CLOSE (v)
   CALL SBM_Solve_Pot(py, px, pz, 1.0d0, kap, rhs, phi, bclo, bchi, mode, 1.0d-8, maxit, iters, relres)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/phi_out.dat', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
   WRITE (v) phi
   CLOSE (v)
   OPEN (NEWUNIT=v, FILE=TRIM(d)//'/solve.txt', STATUS='REPLACE', ACTION='WRITE')
   WRITE (v, *) iters, relres
   CLOSE (v)
END PROGRAM solve_check
