! ref_harness.f90 -- drives the reference's own BINTREE module (tree.f90,
! compiled from /root/reference by oracle/Makefile) on one whole domain, to
! produce golden vectors for the C oracle.  TEST INFRASTRUCTURE ONLY.
!
! Calls, with no MPI and one domain:
!   BuildLocalTree   tree.f90:39
!   LocalCoulombSum  tree.f90:662
!   VVerletNotemp1/2 vverlet.f90:9, 49
!   Search_Tree_ee/_ei/_ie/_ii  tree.f90:721-1133 (timing mode, a target subset)
!
! Input  (stream file, argv(1)): int32 mode, ne, ni, nsteps, stride;
!        real64 theta, rs, xm_e, xm_i, dt; real64 xe(3,ne), ve(3,ne), xi(3,ni), vi(3,ni)
! Output (stream file, argv(2)):
!   mode 1: for each species (e then i): int32 nnodes, then per node in
!           preorder (children 1..8): int32 depth, ileaf, pt, next;
!           real64 xx(3), lx(3), qeff, dx, D(3), Q(3), QQ(3).
!           Then real64 fe(3,ne), fi(3,ni), Epot_e, Epot_i; then the
!           reference's RDF histograms int32 RDF_ei(Ndist+1), RDF_ii(Ndist+1)
!           (bin width 0.01 um; run single-threaded: Search_Tree_ie adds into
!           RDF_ei outside the OpenMP reduction, tree.f90:918).
!   mode 2: nsteps of VV1 / build+LocalCoulombSum / VV2 (sampling every step);
!           per step real64 Epot_e, Epot_i, mv2el, mv2ion, Te_rad, Ti_rad;
!           then final xe, ve, fe, xi, vi, fi.
!   mode 3: timing; real64 t_build, t_walk, int32 n_walked.
PROGRAM REF_HARNESS
  USE DATASTR
  USE BINTREE
  IMPLICIT NONE
  TYPE(PT), ALLOCATABLE :: qel(:), qion(:)
  TYPE(GS), ALLOCATABLE :: gel(:), gion(:)
  TYPE(NM) :: NS
  TYPE(PM) :: param
  TYPE(ST) :: sys
  TYPE(TM) :: ttime
  TYPE(BT), POINTER :: Roote, Rooti
  INTEGER :: mode, ne, ni, nsteps, stride, i, istep, nw
  REAL*8 :: theta, rs, xme, xmi, dt, t0, t1, t2, Epot
  REAL*8, ALLOCATABLE :: buf(:,:)
  CHARACTER(LEN=512) :: fin, fout
  REAL*8, EXTERNAL :: omp_get_wtime

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(UNIT=11, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(11) mode, ne, ni, nsteps, stride
  READ(11) theta, rs, xme, xmi, dt
  ALLOCATE(qel(ne), qion(ni), gel(ne), gion(ni), buf(3, MAX(ne, ni)))
  READ(11) buf(:, 1:ne)
  DO i = 1, ne
     qel(i)%xx(:) = buf(:, i)
  END DO
  READ(11) buf(:, 1:ne)
  DO i = 1, ne
     qel(i)%xv(:) = buf(:, i)
     qel(i)%ff(:) = 0.D0
     qel(i)%id = i
  END DO
  READ(11) buf(:, 1:ni)
  DO i = 1, ni
     qion(i)%xx(:) = buf(:, i)
  END DO
  READ(11) buf(:, 1:ni)
  DO i = 1, ni
     qion(i)%xv(:) = buf(:, i)
     qion(i)%ff(:) = 0.D0
     qion(i)%id = i
  END DO
  CLOSE(11)

  NS%Nel = ne; NS%Nion = ni; NS%NTel = ne; NS%NTion = ni
  NS%NTpt = ne + ni; NS%Npt = ne + ni
  NS%dx_ee = 0.01D0; NS%dx_ei = 0.01D0; NS%dx_ii = 0.01D0
  NS%RDF_ei = 0; NS%RDF_ii = 0
  param%xm(1) = xme; param%xm(2) = xmi
  param%rs = rs; param%sd_ratio = theta
  param%q(1) = -1.D0; param%q(2) = 1.D0
  ttime%Nloop = 0; ttime%Nsamp = 1

  OPEN(UNIT=12, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  IF (mode == 1) THEN
     CALL build_trees()
     CALL dump_tree(Roote)
     CALL dump_tree(Rooti)
     CALL LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)
     DO i = 1, ne
        WRITE(12) qel(i)%ff(:)
     END DO
     DO i = 1, ni
        WRITE(12) qion(i)%ff(:)
     END DO
     WRITE(12) sys%Epot_e, sys%Epot_i
     WRITE(12) NS%RDF_ei(:), NS%RDF_ii(:)
  ELSE IF (mode == 2) THEN
     DO istep = 1, nsteps
        ttime%Nloop = istep
        IF (istep == 1) THEN
           ! initial force, as Vinit does before the loop (main.f90:777)
           CALL build_trees()
           CALL LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)
           CALL DEALLOC_TREE(Roote); CALL DEALLOC_TREE(Rooti)
        END IF
        CALL VVerletNotemp1(NS, qel, qion, param, dt, ttime)
        CALL build_trees()
        CALL LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)
        CALL DEALLOC_TREE(Roote); CALL DEALLOC_TREE(Rooti)
        CALL VVerletNotemp2(NS, qel, qion, param, sys, dt, ttime)
        WRITE(12) sys%Epot_e, sys%Epot_i, sys%mv2el, sys%mv2ion, sys%Te_rad, sys%Ti_rad
     END DO
     DO i = 1, ne
        WRITE(12) qel(i)%xx(:), qel(i)%xv(:), qel(i)%ff(:)
     END DO
     DO i = 1, ni
        WRITE(12) qion(i)%xx(:), qion(i)%xv(:), qion(i)%ff(:)
     END DO
  ELSE IF (mode == 3) THEN
     t0 = omp_get_wtime()
     CALL build_trees()
     t1 = omp_get_wtime()
     Epot = 0.D0
     nw = 0
     !$OMP PARALLEL DO SCHEDULE(STATIC) REDUCTION(+:Epot,nw)
     DO i = 1, ne, stride
        CALL Search_Tree_ee(NS, Roote, qel, i, param, Epot)
        CALL Search_Tree_ei(NS, Rooti, qel, i, param, Epot, NS%RDF_ei)
        nw = nw + 1
     END DO
     !$OMP END PARALLEL DO
     !$OMP PARALLEL DO SCHEDULE(STATIC) REDUCTION(+:Epot,nw)
     DO i = 1, ni, stride
        CALL Search_Tree_ie(NS, Roote, qion, i, param, Epot, NS%RDF_ei)
        CALL Search_Tree_ii(NS, Rooti, qion, i, param, Epot, NS%RDF_ii)
        nw = nw + 1
     END DO
     !$OMP END PARALLEL DO
     t2 = omp_get_wtime()
     WRITE(12) t1 - t0, t2 - t1, nw, Epot
  END IF
  CLOSE(12)

CONTAINS

  SUBROUTINE build_trees()
    ! Force's local part, main.f90:532-541: fresh roots, gather, build
    DO i = 1, ne
       gel(i)%xx(:) = qel(i)%xx(:)
    END DO
    DO i = 1, ni
       gion(i)%xx(:) = qion(i)%xx(:)
    END DO
    NULLIFY(Roote)
    ALLOCATE(Roote)
    Roote%qeff = 0.D0
    NULLIFY(Rooti)
    ALLOCATE(Rooti)
    Rooti%qeff = 0.D0
    CALL BuildLocalTree(NS, gel, gion, Roote, Rooti, ttime)
  END SUBROUTINE build_trees

  SUBROUTINE dump_tree(R)
    TYPE(BT), POINTER :: R
    INTEGER :: cnt
    cnt = 0
    CALL count_nodes(R, cnt)
    WRITE(12) cnt
    cnt = 0
    CALL dump_rec(R, 0, cnt)
  END SUBROUTINE dump_tree

  RECURSIVE SUBROUTINE count_nodes(P, cnt)
    TYPE(BT), POINTER :: P
    INTEGER :: cnt
    cnt = cnt + 1
    IF (ASSOCIATED(P%Child1)) CALL count_nodes(P%Child1, cnt)
    IF (ASSOCIATED(P%Child2)) CALL count_nodes(P%Child2, cnt)
    IF (ASSOCIATED(P%Child3)) CALL count_nodes(P%Child3, cnt)
    IF (ASSOCIATED(P%Child4)) CALL count_nodes(P%Child4, cnt)
    IF (ASSOCIATED(P%Child5)) CALL count_nodes(P%Child5, cnt)
    IF (ASSOCIATED(P%Child6)) CALL count_nodes(P%Child6, cnt)
    IF (ASSOCIATED(P%Child7)) CALL count_nodes(P%Child7, cnt)
    IF (ASSOCIATED(P%Child8)) CALL count_nodes(P%Child8, cnt)
  END SUBROUTINE count_nodes

  RECURSIVE SUBROUTINE dump_rec(P, depth, cnt)
    TYPE(BT), POINTER :: P
    INTEGER :: depth, cnt, nxt, sub
    sub = 0
    CALL count_nodes(P, sub)
    cnt = cnt + 1
    nxt = cnt - 1 + sub
    WRITE(12) depth, P%ileaf, P%pt, nxt
    WRITE(12) P%xx(:), P%lx(:), P%qeff, P%dx, P%D(:), P%Q(:), P%QQ(:)
    IF (ASSOCIATED(P%Child1)) CALL dump_rec(P%Child1, depth + 1, cnt)
    IF (ASSOCIATED(P%Child2)) CALL dump_rec(P%Child2, depth + 1, cnt)
    IF (ASSOCIATED(P%Child3)) CALL dump_rec(P%Child3, depth + 1, cnt)
    IF (ASSOCIATED(P%Child4)) CALL dump_rec(P%Child4, depth + 1, cnt)
    IF (ASSOCIATED(P%Child5)) CALL dump_rec(P%Child5, depth + 1, cnt)
    IF (ASSOCIATED(P%Child6)) CALL dump_rec(P%Child6, depth + 1, cnt)
    IF (ASSOCIATED(P%Child7)) CALL dump_rec(P%Child7, depth + 1, cnt)
    IF (ASSOCIATED(P%Child8)) CALL dump_rec(P%Child8, depth + 1, cnt)
  END SUBROUTINE dump_rec

END PROGRAM REF_HARNESS
