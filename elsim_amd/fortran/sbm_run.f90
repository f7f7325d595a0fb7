!! sbm_run -- Fortran host driver of the GPU SBM step: the time loop of
!! SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:350-877 on one GPU instead of
!! 192 MPI ranks.
!!
!! Usage: sbm_run <namelist file>
!!   &run n1=, n2=, n3=, nsteps=, mode=, sdf_s='', sdf_l='', cnp_file='',
!!        tables_file='', outdir='.', outpf='GBQ105', adaptive=F, cr=6.0, asmb=F /
!!   &params prm%dx=..., prm%dt=..., ... /   (defaults: the reference values)
!! adaptive = F: nsteps steps at the fixed prm%dt and prm%vsr, then one output
!!   frame and one TmFx line.
!! adaptive = T: the reference's loop -- target current trgI from tPP (SBM:341),
!!   dt and Vsr from the time-loop control before every step (SBM:358-383, with
!!   prm%vsr as Vsr0), output frames and TmFx lines on the lithiation cadence
!!   (SBM:842-857) and the stop test (SBM:861-872); nsteps caps the iterations
!!   (the reference has no cap), and a run that reaches the cap writes its last
!!   frame and TmFx line as the stop test does.
!! Inputs are Fortran unformatted records as the reference reads them:
!! FP32 SDFs (SBM:149-152, 181-184), the FP64 ghosted CnP (optional) and the
!! tables as one FP64 (101,8) record.  Outputs: one FP32 record per field
!! named as OutPut3D_A does (SBM:892-919) -- the volume fractions PS and PL
!! once before the loop (SBM:347-348), CP, CE, PP, PE per frame -- and the TmFx
!! text file of ReC_TmFX (SBM:922-943).  Names are <FF>1000<1000+frame>.dat,
!! rank 0's of the MPI program; with asmb = T they are <FF>ASMB<1000+frame>.dat,
!! the names the rank stitcher writes (SBMES_3D_Cmbn_MPI_v2.f90:113-129) and the
!! MATLAB viewers read: one GPU holds the global array, so no stitching is needed.
PROGRAM sbm_run
   USE elsim_gpu
   IMPLICIT NONE
   INTEGER :: n1, n2, n3, nsteps, mode, sz, k, u, ios
   CHARACTER(LEN=512) :: sdf_s, sdf_l, cnp_file, tables_file, outdir, nml
   CHARACTER(LEN=6) :: outpf
   TYPE(elsim_sbm_params) :: prm
   TYPE(elsim_step_stats) :: st
   REAL(KIND=4), ALLOCATABLE :: DmyS(:,:,:), DmyL(:,:,:)
   REAL(KIND=8), ALLOCATABLE :: Tbls(:,:), Fld(:,:,:)
   INTEGER(KIND=8) :: h
   REAL(KIND=8) :: tpp, tpe, xrf, bve, tm, cr, vsr0, scrnt
   INTEGER(c_int) :: it
   INTEGER :: cnt
   LOGICAL :: adaptive, asmb
   CHARACTER(LEN=4) :: rk
   NAMELIST /run/ n1, n2, n3, nsteps, mode, sdf_s, sdf_l, cnp_file, tables_file, outdir, outpf, adaptive, cr, asmb
   NAMELIST /params/ prm

   CALL elsim_sbm_default_params(prm)
   n1 = 0; n2 = 0; n3 = 0; nsteps = 1; mode = ELSIM_MODE_FP64
   sdf_s = ''; sdf_l = ''; cnp_file = ''; tables_file = ''; outdir = '.'; outpf = 'GBQ105'
   adaptive = .FALSE.; cr = 6.0; asmb = .FALSE.
   IF (COMMAND_ARGUMENT_COUNT() < 1) THEN
      WRITE (0, *) 'usage: sbm_run <namelist file>'
      ERROR STOP 1
   END IF
   CALL GET_COMMAND_ARGUMENT(1, nml)
   OPEN (NEWUNIT=u, FILE=TRIM(nml), STATUS='OLD', ACTION='READ')
   READ (u, NML=run)
   READ (u, NML=params, IOSTAT=ios)
   CLOSE (u)

   ALLOCATE (DmyS(n1,n2,n3), DmyL(n1,n2,n3), Fld(0:n1+1,0:n2+1,0:n3+1))
   OPEN (NEWUNIT=u, FILE=TRIM(sdf_s), FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (u) DmyS
   CLOSE (u)
   OPEN (NEWUNIT=u, FILE=TRIM(sdf_l), FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (u) DmyL
   CLOSE (u)
   sz = 101
   ALLOCATE (Tbls(sz,8))
   OPEN (NEWUNIT=u, FILE=TRIM(tables_file), FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
   READ (u) Tbls
   CLOSE (u)

   CALL SBM_Init(n1, n2, n3, prm, DmyS, DmyL, sz, Tbls, h)
   IF (LEN_TRIM(cnp_file) > 0) THEN
      OPEN (NEWUNIT=u, FILE=TRIM(cnp_file), FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
      READ (u) Fld
      CLOSE (u)
      CALL SBM_Set(h, ELSIM_F_CNP, Fld)
   END IF

   IF (asmb) THEN
      rk = 'ASMB'
   ELSE
      WRITE (rk, "(I4)") 1000
   END IF
   CALL write_field(ELSIM_F_PSP, 'PS', 1)
   CALL write_field(ELSIM_F_PSE, 'PL', 1)
   tm = 0.0d0
   IF (adaptive) THEN
      CALL time_loop()
   ELSE
      DO k = 1, nsteps
         CALL SBM_Step(h, mode, st)
         tm = tm + prm%dt
         CALL report(k)
      END DO
      CALL elsim_check(elsim_sbm_scalars(h, tpp, tpe, xrf, bve, it), 'scalars')
      CALL write_frame(1)
      CALL ReC_TmFX(1, outpf, nsteps, tm, xrf, st%scrnt, prm%trgi, bve)
   END IF
   CALL SBM_Free(h)

CONTAINS

   !! The reference's time loop (SBM:350-877).  Literals are default REAL as
   !! there: 0.95 (SBM:341), 0.971, 0.02 + (cnt-1)*0.005, 0.97 and 0.005 are
   !! REAL(4) values promoted in the comparisons.
   SUBROUTINE time_loop()
      CALL elsim_check(elsim_sbm_scalars(h, tpp, tpe, xrf, bve, it), 'scalars')
      prm%trgi = 0.95*tpp*prm%rho/(3600/cr)
      vsr0 = prm%vsr
      scrnt = prm%trgi
      cnt = 1
      k = 0
      DO WHILE (xrf < 0.971)
         IF (k >= nsteps) THEN
            CALL ReC_TmFX(cnt, outpf, k, tm, xrf, scrnt, prm%trgi, bve)
            CALL write_frame(cnt)
            EXIT
         END IF
         k = k + 1
         CALL elsim_check(elsim_sbm_time_control(xrf, scrnt, prm%trgi, vsr0, prm%dt, prm%vsr), 'time control')
         CALL elsim_check(elsim_sbm_set_params(h, prm), 'set params')
         CALL SBM_Step(h, mode, st)
         CALL elsim_check(elsim_sbm_scalars(h, tpp, tpe, xrf, bve, it), 'scalars')
         scrnt = st%scrnt
         tm = tm + prm%dt
         CALL report(k)
         IF (xrf > 0.02 + (cnt-1)*0.005) THEN
            IF (MOD(cnt, 10) == 1) CALL write_frame(cnt)
            CALL ReC_TmFX(cnt, outpf, k, tm, xrf, scrnt, prm%trgi, bve)
            cnt = cnt + 1
         END IF
         IF (bve >= prm%bvp + 0.005 .OR. xrf > 0.97) THEN
            CALL ReC_TmFX(cnt, outpf, k, tm, xrf, scrnt, prm%trgi, bve)
            CALL write_frame(cnt)
            EXIT
         END IF
      END DO
   END SUBROUTINE time_loop

   SUBROUTINE report(kk)
      INTEGER, INTENT(IN) :: kk
      WRITE (*, '(A,I6,A,I6,A,I7,A,I7,A,I4,A,ES12.5,A,ES10.3,A,ES10.3,A,ES10.3)') 'step', kk, ' it_cne', &
         st%it_cne, ' it_php', st%it_php, ' it_phe', st%it_phe, ' outer', st%n_outer, ' xrf', st%xrf, &
         ' relres_php', st%relres_php, ' dt', prm%dt, ' s', st%seconds
      FLUSH (6)
   END SUBROUTINE report

   SUBROUTINE write_frame(fr)
      INTEGER, INTENT(IN) :: fr
      CALL write_field(ELSIM_F_CNP, 'CP', fr)
      CALL write_field(ELSIM_F_CNE, 'CE', fr)
      CALL write_field(ELSIM_F_PHP, 'PP', fr)
      CALL write_field(ELSIM_F_PHE, 'PE', fr)
   END SUBROUTINE write_frame

   !! OutPut3D_A(fr, ..., prfn=FF//rk, AB=2): SNGL record named
   !! FF//rk//I4(1000+frame)//'.dat', rk = I4(1000+rank) or 'ASMB'
   SUBROUTINE write_field(fid, ff, fr)
      INTEGER, INTENT(IN) :: fid, fr
      CHARACTER(LEN=2), INTENT(IN) :: ff
      CHARACTER(LEN=4) :: fs
      INTEGER :: v
      CALL SBM_Get(h, fid, Fld)
      WRITE (fs, "(I4)") 1000 + fr
      OPEN (NEWUNIT=v, FILE=TRIM(outdir)//'/'//ff//rk//fs//'.dat', FORM='UNFORMATTED', STATUS='REPLACE', &
         ACTION='WRITE')
      WRITE (v) SNGL(Fld(1:n1,1:n2,1:n3))
      CLOSE (v)
   END SUBROUTINE write_field

   !! ReC_TmFX (SBM:922-943)
   SUBROUTINE ReC_TmFX(fr, prfn, iter, tmv, frx, crnt, Tg, BvEv)
      INTEGER, INTENT(IN) :: fr, iter
      REAL(KIND=8), INTENT(IN) :: tmv, frx, crnt, Tg, BvEv
      CHARACTER(LEN=6), INTENT(IN) :: prfn
      INTEGER :: v
      IF (fr == 1) THEN
         OPEN (NEWUNIT=v, FILE=TRIM(outdir)//'/'//prfn//'TmFx.txt', STATUS='REPLACE', ACTION='WRITE')
      ELSE
         OPEN (NEWUNIT=v, FILE=TRIM(outdir)//'/'//prfn//'TmFx.txt', POSITION='APPEND', ACTION='WRITE')
      END IF
      WRITE (v, *) fr, iter, tmv, frx, crnt, Tg, BvEv
      CLOSE (v)
   END SUBROUTINE ReC_TmFX

END PROGRAM sbm_run
