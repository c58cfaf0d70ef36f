! main.f90 -- Fortran driver of the GPU engine for ultracold-plasma MD.
! Same program flow, input files and output files as the reference's
! PROGRAM OCTREE_HYBRID (main.f90:44-187) on one domain: control.prm and
! input.xyz in, ener###.dat (TEMPORAL, sample.f90:55-96) and res###.xyz
! (Restart, print.f90:5-109) out.  Each step is VVerletNotemp1, Force,
! VVerletNotemp2; Force builds both trees and runs LocalCoulombSum on the GPU.
! The walk is the reference-order FP64 walk unless TREEMI_WALK=mixed.
PROGRAM OCTREE_GPU
  USE ISO_C_BINDING
  USE DATASTR
  USE TREEMI_CAPI
  USE TREEMI_FORCE
  IMPLICIT NONE
  TYPE(PT), ALLOCATABLE :: qel(:), qion(:)
  TYPE(TM) :: ttime
  TYPE(PM) :: param
  TYPE(ST) :: sys
  TYPE(NM) :: NS
  REAL(C_DOUBLE) :: dt
  INTEGER :: Nloop_max, Nfreq, c0, c1, crate
  INTEGER(C_INT32_T) :: walk
  CHARACTER(LEN=20) :: ENERFILE
  CHARACTER(LEN=16) :: wenv

  CALL SYSTEM_CLOCK(c0, crate)
  ENERFILE = "ener000.dat"
  OPEN(UNIT=25, FILE=ENERFILE)
  WRITE(25, 200)
200 FORMAT("# time(ns)-   T_el.  -  T_ion  -  T_el_radial  -  T_ion_radial &
         & -  T_el_circumf  -  T_ion_circumf  -  Potential Energy  - &
         & Total energy")

  CALL Init(NS, qel, qion, ttime, param, sys, dt)
  walk = TB_WALK_REF64
  CALL GET_ENVIRONMENT_VARIABLE("TREEMI_WALK", wenv)
  IF (TRIM(wenv) == "mixed") walk = TB_WALK_MIXED
  CALL Engine_Init(param, walk)
  sys%rs = param%rs
  IF (param%rest == 'OFF') THEN
     CALL Vinit(NS, qel, qion, param, ttime, sys, dt)
  ELSE
     CALL Vinit_REST(NS, qel, qion, param, ttime, sys, dt)
  END IF

  ! the RDF accumulates from here (main.f90:118-120): the FP64 walk's leaf
  ! pairs; the mixed walk's histograms hold every pair in their range
  CALL Engine_RDF_Init(NS)
  ttime%Ndump = NINT(ttime%tdump/dt)
  ttime%Nrest = NINT(ttime%trest/dt)
  ttime%Nsamp = NINT(ttime%tsamp/dt)
  ttime%Nloop = 0
  ttime%tnow = 0.0D0
  ttime%tmax = ttime%tmax + 0.5D0*dt
  Nloop_max = NINT(ttime%tmax/dt) - 1

  DO WHILE (ttime%Nloop <= Nloop_max)
     ttime%tnow = ttime%tnow + dt
     ttime%Nloop = ttime%Nloop + 1
     CALL VVerletNotemp1(NS, qel, qion, param, dt, ttime)
     CALL Force(NS, qel, qion, param, ttime)
     CALL VVerletNotemp2(NS, qel, qion, param, sys, dt, ttime)
     IF (MOD(ttime%Nloop, ttime%Ndump) == 0) CALL RDF(NS, ttime)
     sys%Epot = sys%Epot_e + sys%Epot_i
     IF (MOD(ttime%Nloop, ttime%Nsamp) == 0) CALL TEMPORAL(NS, sys, ttime)
     IF (MOD(ttime%Nloop, ttime%Nrest) == 0) THEN
        CALL Engine_Sync(NS, qel, qion)
        CALL Restart(NS, qel, qion, ttime)
        CLOSE(25)
        Nfreq = INT(ttime%Nloop/ttime%Nrest)
        WRITE(ENERFILE, 480) Nfreq
        OPEN(UNIT=25, FILE=ENERFILE)
        WRITE(25, 200)
     END IF
  END DO
480 FORMAT("ener", I3.3, ".dat")
  CLOSE(25)
  CALL Engine_Sync(NS, qel, qion)
  CALL Engine_Done()
  CALL SYSTEM_CLOCK(c1)
  PRINT '("Wall time is", ES11.3, " s for ", I0, " steps")', DBLE(c1 - c0)/DBLE(crate), ttime%Nloop

CONTAINS

  ! control.prm and input.xyz, positional as in the reference (main.f90:222-383)
  SUBROUTINE Init(NS, qel, qion, ttime, param, sys, dt)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(PT), ALLOCATABLE, INTENT(INOUT) :: qel(:), qion(:)
    TYPE(TM), INTENT(INOUT) :: ttime
    TYPE(PM), INTENT(INOUT) :: param
    TYPE(ST), INTENT(INOUT) :: sys
    REAL(C_DOUBLE), INTENT(INOUT) :: dt
    CHARACTER(LEN=5) :: dummy
    CHARACTER(LEN=256) :: line
    INTEGER :: i, nid, ios, nhead
    REAL(C_DOUBLE) :: x, y, z, xv, yv, zv
    OPEN(UNIT=15, FILE="control.prm", STATUS="OLD", IOSTAT=ios)
    IF (ios /= 0) STOP "control.prm: cannot open"
    READ(15, *, IOSTAT=ios) dummy
    READ(15, *, IOSTAT=ios) dummy
    DO i = 1, Nparam
       IF (ios == 0) READ(15, *, IOSTAT=ios) param%q(i)
    END DO
    CALL prm_check(ios, "charges")
    READ(15, *, IOSTAT=ios) dummy
    DO i = 1, Nparam
       IF (ios == 0) READ(15, *, IOSTAT=ios) param%xm(i)
    END DO
    CALL prm_check(ios, "masses")
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) NS%NTel
    IF (ios == 0) READ(15, *, IOSTAT=ios) NS%NTion
    CALL prm_check(ios, "particle numbers")
    IF (NS%NTel < 1 .OR. NS%NTion < 1) STOP "control.prm: particle numbers must be positive"
    NS%NTpt = NS%NTel + NS%NTion
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) param%Te
    IF (ios == 0) READ(15, *, IOSTAT=ios) param%Ti
    CALL prm_check(ios, "temperatures")
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) ttime%tmax, ttime%trest, ttime%tdump, ttime%tsamp, dt
    CALL prm_check(ios, "time control")
    IF (.NOT. (dt > 0.D0)) STOP "control.prm: the time step must be positive"
    ttime%tmax = ttime%tmax/tps
    ttime%trest = ttime%trest/tps
    ttime%tdump = ttime%tdump/tps
    ttime%tsamp = ttime%tsamp/tps
    dt = dt/tps
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) sys%kelbg, param%rs
    CALL prm_check(ios, "Kelbg flag and length")
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) param%rest
    CALL prm_check(ios, "restart flag")
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) NS%dx_ee, NS%dx_ei, NS%dx_ii
    CALL prm_check(ios, "RDF bin widths")
    READ(15, *, IOSTAT=ios) dummy
    IF (ios == 0) READ(15, *, IOSTAT=ios) param%sd_ratio
    CALL prm_check(ios, "opening angle")
    CLOSE(15)

    NS%Nel = NS%NTel
    NS%Nion = NS%NTion
    NS%Npt = NS%NTpt
    ALLOCATE(qel(NS%Nel), qion(NS%Nion))
    OPEN(UNIT=11, FILE="input.xyz", STATUS="OLD", IOSTAT=ios)
    IF (ios /= 0) STOP "input.xyz: cannot open"
    ! the first line holds the particle count (as Restart writes it): it must
    ! match control.prm, as the reference's count check does (main.f90:415-421)
    READ(11, '(A)', IOSTAT=ios) line
    IF (ios /= 0) STOP "input.xyz: empty file"
    READ(line, *, IOSTAT=ios) nhead
    IF (ios == 0 .AND. nhead /= NS%NTpt) THEN
       PRINT *, "parsing error", nhead, "while given number is", NS%NTpt
       STOP "input.xyz: particle count differs from control.prm"
    END IF
    READ(11, *, IOSTAT=ios) dummy
    IF (ios /= 0) STOP "input.xyz: missing second header line"
    sys%mv2el = 0.D0
    sys%mv2ion = 0.D0
    DO i = 1, NS%Nel
       IF (param%rest == 'OFF') THEN
          READ(11, *, IOSTAT=ios) dummy, x, y, z, nid
          xv = 0.D0; yv = 0.D0; zv = 0.D0
       ELSE
          READ(11, *, IOSTAT=ios) dummy, x, y, z, xv, yv, zv, nid
       END IF
       CALL xyz_check(ios, i, NS%NTpt)
       qel(i)%xx = (/x, y, z/); qel(i)%xv = (/xv, yv, zv/); qel(i)%ff = 0.D0
       qel(i)%id = nid; qel(i)%pad = 0
       sys%mv2el = sys%mv2el + param%xm(1)*(xv**2 + yv**2 + zv**2)
    END DO
    DO i = 1, NS%Nion
       IF (param%rest == 'OFF') THEN
          READ(11, *, IOSTAT=ios) dummy, x, y, z, nid
          xv = 0.D0; yv = 0.D0; zv = 0.D0
       ELSE
          READ(11, *, IOSTAT=ios) dummy, x, y, z, xv, yv, zv, nid
       END IF
       CALL xyz_check(ios, NS%Nel + i, NS%NTpt)
       qion(i)%xx = (/x, y, z/); qion(i)%xv = (/xv, yv, zv/); qion(i)%ff = 0.D0
       qion(i)%id = nid; qion(i)%pad = 0
       sys%mv2ion = sys%mv2ion + param%xm(2)*(xv**2 + yv**2 + zv**2)
    END DO
    CLOSE(11)
    ! temperature normalisations with Nref removed degrees of freedom (main.f90:393-399)
    sys%kbm = kboltz*DBLE(3*NS%NTel + 3*NS%NTion - Nref)
    sys%kbm_e = kboltz*DBLE(3*NS%NTel - Nref)
    sys%kbm_i = kboltz*DBLE(3*NS%NTion - Nref)
    sys%kbm_e_rad = kboltz*DBLE(NS%NTel - Nref)
    sys%kbm_e_cir = kboltz*DBLE(2*NS%NTel - Nref)
    sys%kbm_i_rad = kboltz*DBLE(NS%NTion - Nref)
    sys%kbm_i_cir = kboltz*DBLE(2*NS%NTion - Nref)
    IF (param%rest /= 'OFF') sys%temp = (sys%mv2el + sys%mv2ion)/sys%kbm
  END SUBROUTINE Init

  ! a failed control.prm read STOPs with the field's name
  SUBROUTINE prm_check(ios, what)
    INTEGER, INTENT(IN) :: ios
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (ios /= 0) THEN
       PRINT *, "control.prm: cannot read ", what
       STOP "control.prm: malformed"
    END IF
  END SUBROUTINE prm_check

  ! a short or malformed input.xyz STOPs with the particle it stopped at
  SUBROUTINE xyz_check(ios, k, n)
    INTEGER, INTENT(IN) :: ios, k, n
    IF (ios /= 0) THEN
       PRINT *, "parsing error at particle", k, "of", n
       STOP "input.xyz: short or malformed"
    END IF
  END SUBROUTINE xyz_check

  ! Maxwell-Boltzmann velocities at exactly Te / Ti (main.f90:709-790), from
  ! the engine's seeded generator, then the initial force
  SUBROUTINE Vinit(NS, qel, qion, param, ttime, sys, dt)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(:), qion(:)
    TYPE(PM), INTENT(IN) :: param
    TYPE(TM), INTENT(INOUT) :: ttime
    TYPE(ST), INTENT(INOUT) :: sys
    REAL(C_DOUBLE), INTENT(IN) :: dt
    TYPE(PT), ALLOCATABLE :: ve(:), vi(:)
    INTEGER :: i
    ALLOCATE(ve(NS%Nel), vi(NS%Nion))
    CALL tb_check(tb_gen_plasma(1_C_INT64_T, INT(NS%Nel, C_INT32_T), INT(NS%Nion, C_INT32_T), 1.D0, &
         param%Te, param%Ti, param%xm(1), param%xm(2), ve, vi), "Vinit")
    DO i = 1, NS%Nel
       qel(i)%xv = ve(i)%xv
    END DO
    DO i = 1, NS%Nion
       qion(i)%xv = vi(i)%xv
    END DO
    sys%Te = param%Te
    sys%Ti = param%Ti
    sys%temp = (sys%kbm_e*param%Te + sys%kbm_i*param%Ti)/sys%kbm
    CALL Engine_Upload(NS, qel, qion)
    CALL Force(NS, qel, qion, param, ttime)
  END SUBROUTINE Vinit

  SUBROUTINE Vinit_REST(NS, qel, qion, param, ttime, sys, dt)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(:), qion(:)
    TYPE(PM), INTENT(IN) :: param
    TYPE(TM), INTENT(INOUT) :: ttime
    TYPE(ST), INTENT(INOUT) :: sys
    REAL(C_DOUBLE), INTENT(IN) :: dt
    CALL Engine_Upload(NS, qel, qion)
    CALL Force(NS, qel, qion, param, ttime)
  END SUBROUTINE Vinit_REST

  ! main.f90:506-541 on one domain: gather, build both trees, then the local sum
  SUBROUTINE Force(NS, qel, qion, param, ttime)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(NS%Nel), qion(NS%Nion)
    TYPE(PM), INTENT(IN) :: param
    TYPE(TM), INTENT(INOUT) :: ttime
    TYPE(TREE_REF) :: Roote, Rooti
    TYPE(GS), ALLOCATABLE, SAVE :: gel(:), gion(:)
    CALL Engine_Gather(NS, gel, gion)
    CALL BuildLocalTree(NS, gel, gion, Roote, Rooti, ttime)
    CALL LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)
  END SUBROUTINE Force

  ! sample.f90:5-50 on one domain: rdf###.dat, counts averaged over the Ndump
  ! steps since the last print, then cleared
  SUBROUTINE RDF(NS, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(TM), INTENT(IN) :: ttime
    INTEGER(C_INT64_T) :: Nei(Ndist + 1), Nii(Ndist + 1)
    REAL(C_DOUBLE) :: Ntau
    INTEGER :: i, Nfreq
    CHARACTER(LEN=20) :: FILENAME
    CALL Engine_RDF_Take(Nei, Nii)
    Ntau = REAL(ttime%Ndump, C_DOUBLE)
    Nfreq = INT(ttime%Nloop/ttime%Ndump)
    WRITE(FILENAME, 50) Nfreq
    OPEN(UNIT=30, FILE=FILENAME)
    WRITE(30, 100)
    DO i = 1, Ndist + 1
       WRITE(30, 200) NS%dx_ei*REAL(i, C_DOUBLE), REAL(Nei(i), C_DOUBLE)/Ntau, &
            NS%dx_ii*REAL(i, C_DOUBLE), REAL(Nii(i), C_DOUBLE)/Ntau
    END DO
    CLOSE(30)
50  FORMAT("rdf", I3.3, ".dat")
100 FORMAT("# R_ei,   RDF_ei       R_ii     RDF_ii     scaled RDF_ii")
200 FORMAT(5(ES14.6, 1X))
  END SUBROUTINE RDF

  ! sample.f90:55-96 on one domain
  SUBROUTINE TEMPORAL(NS, sys, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(ST), INTENT(INOUT) :: sys
    TYPE(TM), INTENT(IN) :: ttime
    REAL(C_DOUBLE) :: Tel, Tion, Te_rad, Ti_rad, Te_cir, Ti_cir, Epot
    Epot = sys%Epot_e + sys%Epot_i
    Tel = sys%mv2el
    Tion = sys%mv2ion
    sys%Te = Tel/sys%kbm_e
    sys%Ti = Tion/sys%kbm_i
    Te_cir = (Tel - sys%Te_rad)/sys%kbm_e_cir
    Ti_cir = (Tion - sys%Ti_rad)/sys%kbm_i_cir
    Te_rad = sys%Te_rad/sys%kbm_e_rad
    Ti_rad = sys%Ti_rad/sys%kbm_i_rad
    WRITE(25, 500) ttime%tnow*tps, sys%Te, sys%Ti, Te_rad, Ti_rad, Te_cir, Ti_cir, Epot, &
         0.5D0*(Tel + Tion) + Epot
500 FORMAT(9(ES14.6, 1X))
  END SUBROUTINE TEMPORAL

  ! print.f90:5-109 on one domain: res###.xyz, P = electron, O = ion
  SUBROUTINE Restart(NS, qel, qion, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(PT), INTENT(IN) :: qel(NS%Nel), qion(NS%Nion)
    TYPE(TM), INTENT(IN) :: ttime
    CHARACTER(LEN=20) :: FILENAME
    INTEGER :: i, Nf
    Nf = INT(ttime%Nloop/ttime%Nrest)
    WRITE(FILENAME, 55) Nf
55  FORMAT("res", I3.3, ".xyz")
    OPEN(UNIT=30, FILE=FILENAME)
    WRITE(30, *) NS%NTpt
    WRITE(30, *) "frame = ", Nf, "  energy = 0"
    DO i = 1, NS%Nel
       WRITE(30, 80) qel(i)%xx, qel(i)%xv, qel(i)%id
    END DO
    DO i = 1, NS%Nion
       WRITE(30, 90) qion(i)%xx, qion(i)%xv, qion(i)%id
    END DO
    CLOSE(30)
80  FORMAT("P  ", 6(ES14.6, 1X), I8)
90  FORMAT("O  ", 6(ES14.6, 1X), I8)
  END SUBROUTINE Restart
END PROGRAM OCTREE_GPU
