!! dropin_check -- calls the drop-in subroutines that the driver programs do
!! not (NoGrdBc3D, AbsGrd3DUW_XYZ, Tbl_Chk_Pln_3DA, Label_Connected,
!! RM_Threshold) on the records in <dir> and writes their results there, for
!! the comparison with the oracle in tests/test_gpu_fortran.py.
!! Usage: dropin_check <dir> py px pz sz phase
PROGRAM dropin_check
   IMPLICIT NONE
   CHARACTER(LEN=512) :: d, arg
   INTEGER :: py, px, pz, sz, phase
   REAL(KIND=8), ALLOCATABLE :: f(:,:,:), sg(:,:,:), gr(:,:,:), c(:,:,:), fn(:,:,:), cord(:), vals(:), psa(:,:,:), &
      psb(:,:,:), rd(:,:,:)
   INTEGER(KIND=4), ALLOCATABLE :: lab(:,:,:), con(:,:,:)
   CALL GET_COMMAND_ARGUMENT(1, d)
   CALL GET_COMMAND_ARGUMENT(2, arg); READ (arg, *) py
   CALL GET_COMMAND_ARGUMENT(3, arg); READ (arg, *) px
   CALL GET_COMMAND_ARGUMENT(4, arg); READ (arg, *) pz
   CALL GET_COMMAND_ARGUMENT(5, arg); READ (arg, *) sz
   CALL GET_COMMAND_ARGUMENT(6, arg); READ (arg, *) phase
   ALLOCATE (f(0:py+1,0:px+1,0:pz+1), sg(0:py+1,0:px+1,0:pz+1), gr(py,px,pz), c(py,px,pz), fn(py,px,pz), &
      cord(sz), vals(sz), psa(py,px,pz), psb(py,px,pz), rd(py,px,pz), lab(py,px,pz), con(py,px,pz))
   CALL get(TRIM(d)//'/f.dat', f)
   CALL get(TRIM(d)//'/sg.dat', sg)
   CALL get(TRIM(d)//'/c.dat', c)
   CALL get(TRIM(d)//'/psa.dat', psa)
   CALL get(TRIM(d)//'/rd.dat', rd)
   CALL get(TRIM(d)//'/cord.dat', cord)
   CALL get(TRIM(d)//'/vals.dat', vals)
   CALL AbsGrd3DUW_XYZ(py, px, pz, sg, f, gr)
   CALL NoGrdBc3D(py, px, pz, f)
   CALL Tbl_Chk_Pln_3DA(py, px, pz, sz, cord, vals, c, fn)
   CALL RM_Threshold(py, px, pz, c, psa, psb)
   CALL Label_Connected(py, px, pz, rd, phase, lab, con)
   CALL put(TRIM(d)//'/gr_out.dat', gr)
   CALL put(TRIM(d)//'/f_out.dat', f)
   CALL put(TRIM(d)//'/fn_out.dat', fn)
   CALL put(TRIM(d)//'/psb_out.dat', psb)
   CALL puti(TRIM(d)//'/lab_out.dat', lab)
   CALL puti(TRIM(d)//'/con_out.dat', con)
CONTAINS
   SUBROUTINE get(name, a)
      CHARACTER(LEN=*), INTENT(IN) :: name
      REAL(KIND=8), INTENT(OUT) :: a(..)
      INTEGER :: v
      OPEN (NEWUNIT=v, FILE=name, FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
      SELECT RANK (a)
      RANK (1)
         READ (v) a
      RANK (3)
         READ (v) a
      END SELECT
      CLOSE (v)
   END SUBROUTINE get
   SUBROUTINE put(name, a)
      CHARACTER(LEN=*), INTENT(IN) :: name
      REAL(KIND=8), INTENT(IN) :: a(:,:,:)
      INTEGER :: v
      OPEN (NEWUNIT=v, FILE=name, FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
      WRITE (v) a
      CLOSE (v)
   END SUBROUTINE put
   SUBROUTINE puti(name, a)
      CHARACTER(LEN=*), INTENT(IN) :: name
      INTEGER(KIND=4), INTENT(IN) :: a(:,:,:)
      INTEGER :: v
      OPEN (NEWUNIT=v, FILE=name, FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
      WRITE (v) a
      CLOSE (v)
   END SUBROUTINE puti
END PROGRAM dropin_check
