!! sbm_asmb -- the rank-output stitcher (SBMES_3D_Cmbn_MPI_v2.f90:3-137): reads
!! the per-rank FP32 records <FF><1000+rank><1000+frame>.dat that an MPI run of
!! the SBM program writes, places each rank's box in the global array and
!! writes <FF>ASMB<1000+frame>.dat, the files the MATLAB viewers read.
!! A one-GPU run of sbm_run already holds the global array and can write the
!! ASMB names itself (asmb = T); this program is for outputs of a decomposed run.
!!
!! Usage: sbm_asmb <namelist file>
!!   &asmb gy=186, gx=1460, gz=320, rnb=12, cnb=16, fields='CP','CE','PE','PP',
!!         fb=135, fe=135, itv=10, indir='.', outdir='.' /
!! (fields: up to 16 two-letter prefixes; none given means CP, CE, PE, PP)
!! The decomposition is the SBM program's (MyDmnCmpn_YZ, SBM:947-982) as CB
!! re-derives it (CB:26-72): rank -> column ddC = rank/rnb + 1 and row ddR = MOD(rank,rnb) + 1;
!! d3 split into rnb and d1 into cnb near-equal parts, the first MOD(g,parts)
!! parts one cell longer; d2 whole.
PROGRAM sbm_asmb
   IMPLICIT NONE
   INTEGER :: gy, gx, gz, rnb, cnb, fb, fe, itv, np, rank, tfrm, u, f
   INTEGER, ALLOCATABLE :: box(:,:)
   REAL(KIND=4), ALLOCATABLE :: glob(:,:,:), part(:,:,:)
   CHARACTER(LEN=512) :: indir, outdir, nml
   CHARACTER(LEN=2) :: fields(16)
   CHARACTER(LEN=4) :: frk, fsn
   NAMELIST /asmb/ gy, gx, gz, rnb, cnb, fields, fb, fe, itv, indir, outdir

   gy = 186; gx = 1460; gz = 320; rnb = 12; cnb = 16
   fields = '  '
   fb = 1; fe = 1; itv = 1; indir = '.'; outdir = '.'
   IF (COMMAND_ARGUMENT_COUNT() < 1) THEN
      WRITE (0, *) 'usage: sbm_asmb <namelist file>'
      ERROR STOP 1
   END IF
   CALL GET_COMMAND_ARGUMENT(1, nml)
   OPEN (NEWUNIT=u, FILE=TRIM(nml), STATUS='OLD', ACTION='READ')
   READ (u, NML=asmb)
   CLOSE (u)
   IF (ALL(fields == '  ')) fields(1:4) = ['CP', 'CE', 'PE', 'PP']  ! CB:81-91, St = 2..5
   IF (gy < cnb .OR. gz < rnb .OR. gx < 1 .OR. rnb < 1 .OR. cnb < 1 .OR. itv < 1) THEN
      WRITE (0, *) 'sbm_asmb: every rank needs at least one cell'
      ERROR STOP 2
   END IF

   np = rnb*cnb
   ALLOCATE (box(0:np-1,6), glob(gy,gx,gz))
   DO rank = 0, np-1
      CALL split(gz, rnb, MOD(rank, rnb), box(rank,4), box(rank,6))
      CALL split(gy, cnb, rank/rnb, box(rank,1), box(rank,3))
      box(rank,2) = box(rank,1) + box(rank,3) - 1
      box(rank,5) = box(rank,4) + box(rank,6) - 1
   END DO

   DO f = 1, SIZE(fields)
      IF (fields(f) == '  ') CYCLE
      DO tfrm = fb, fe, itv
         WRITE (fsn, "(I4)") 1000 + tfrm
         DO rank = 0, np-1
            WRITE (frk, "(I4)") 1000 + rank
            ALLOCATE (part(box(rank,3), gx, box(rank,6)))
            OPEN (NEWUNIT=u, FILE=TRIM(indir)//'/'//fields(f)//frk//fsn//'.dat', FORM='UNFORMATTED', STATUS='OLD', &
               ACTION='READ')
            READ (u) part
            CLOSE (u)
            glob(box(rank,1):box(rank,2), :, box(rank,4):box(rank,5)) = part
            DEALLOCATE (part)
         END DO
         OPEN (NEWUNIT=u, FILE=TRIM(outdir)//'/'//fields(f)//'ASMB'//fsn//'.dat', FORM='UNFORMATTED', &
            STATUS='REPLACE', ACTION='WRITE')
         WRITE (u) glob
         CLOSE (u)
         WRITE (*, '(A,A,I6,I6)') 'assembled ', fields(f), tfrm, np
      END DO
   END DO

CONTAINS

   !! Part p (0-based) of g cells in n parts: the first MOD(g,n) parts hold
   !! g/n + 1 cells, the rest g/n; lo is its first cell (1-based).
   SUBROUTINE split(g, n, p, lo, len)
      INTEGER, INTENT(IN) :: g, n, p
      INTEGER, INTENT(OUT) :: lo, len
      INTEGER :: b, r
      b = g/n
      r = MOD(g, n)
      IF (p < r) THEN
         len = b + 1
         lo = p*len + 1
      ELSE
         len = b
         lo = p*len + 1 + r
      END IF
   END SUBROUTINE split

END PROGRAM sbm_asmb
