! treemi_force.f90 -- the reference's tree and integrator subroutines, by name
! and argument list, over the GPU engine:
!   BuildLocalTree(NS, gel, gion, Roote, Rooti, ttime)               tree.f90:39
!   LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)  tree.f90:662
!   VVerletNotemp1(NS, qel, qion, param, dt, ttime)                  vverlet.f90:9
!   VVerletNotemp2(NS, qel, qion, param, sys, dt, ttime)             vverlet.f90:49
! The particle state lives on the GPU between calls; qel/qion are host
! mirrors refreshed by Engine_Sync (Restart and the end of the run call it).
! Roote/Rooti are opaque references to the device trees (the pointer trees of
! TYPE BT cannot cross the C boundary).
MODULE TREEMI_FORCE
  USE ISO_C_BINDING
  USE DATASTR
  USE TREEMI_CAPI
  IMPLICIT NONE

  TYPE TREE_REF
     INTEGER :: species = -1     ! 0 electrons, 1 ions
  END TYPE TREE_REF

  TYPE(C_PTR), SAVE :: hnd = C_NULL_PTR

CONTAINS

  SUBROUTINE Engine_Init(param, walk)
    TYPE(PM), INTENT(IN) :: param
    INTEGER(C_INT32_T), INTENT(IN) :: walk
    TYPE(TB_CONFIG) :: cfg
    cfg%xm = param%xm
    cfg%rs = param%rs
    cfg%theta = param%sd_ratio
    cfg%theta_mixed = param%sd_ratio
    cfg%theta_lo = 0.D0
    cfg%err_target = 0.D0
    cfg%walk = walk
    cfg%mac = TB_MAC_FIXED
    cfg%group = 0
    cfg%device = 0
    cfg%mem_cap = 0
    cfg%far_ratio = 0.0_C_DOUBLE
    cfg%far_levels = 0
    cfg%cost_weight = 0.0_C_FLOAT
    CALL tb_check(tb_create(hnd, cfg), "Engine_Init")
  END SUBROUTINE Engine_Init

  SUBROUTINE Engine_Upload(NS, qel, qion)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(PT), INTENT(IN) :: qel(NS%Nel), qion(NS%Nion)
    CALL tb_check(tb_set_particles(hnd, INT(NS%Nel, C_INT32_T), qel, INT(NS%Nion, C_INT32_T), qion), &
         "Engine_Upload")
  END SUBROUTINE Engine_Upload

  SUBROUTINE Engine_Sync(NS, qel, qion)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(NS%Nel), qion(NS%Nion)
    CALL tb_check(tb_get_particles(hnd, qel, qion), "Engine_Sync")
  END SUBROUTINE Engine_Sync

  ! the RDF histograms (NS%RDF_ei / RDF_ii, data.f90:139), accumulated by the
  ! FP64 walk in every force from here on (main.f90:118-120 zeroes them here)
  SUBROUTINE Engine_RDF_Init(NS)
    TYPE(NM), INTENT(IN) :: NS
    CALL tb_check(tb_rdf_enable(hnd, NS%dx_ei, NS%dx_ii, INT(Ndist + 1, C_INT32_T)), "Engine_RDF_Init")
  END SUBROUTINE Engine_RDF_Init

  ! the histograms accumulated since the last call, which clears them
  SUBROUTINE Engine_RDF_Take(rdf_ei, rdf_ii)
    INTEGER(C_INT64_T), INTENT(OUT) :: rdf_ei(Ndist + 1), rdf_ii(Ndist + 1)
    CALL tb_check(tb_rdf_get(hnd, INT(Ndist + 1, C_INT32_T), rdf_ei, rdf_ii, 1_C_INT32_T), "Engine_RDF_Take")
  END SUBROUTINE Engine_RDF_Take

  ! PTEgathering2 / PTIgathering2 (main.f90:430-502) on one domain: every
  ! particle is local, so the totals NS%NTel / NS%NTion are the local counts;
  ! the position copies are sized for BuildLocalTree's argument list but not
  ! filled (the positions live on the GPU)
  SUBROUTINE Engine_Gather(NS, gel, gion)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(GS), ALLOCATABLE, INTENT(INOUT) :: gel(:), gion(:)
    NS%NTel = NS%Nel
    NS%NTion = NS%Nion
    IF (ALLOCATED(gel)) THEN
       IF (SIZE(gel) /= NS%NTel) DEALLOCATE(gel)
    END IF
    IF (ALLOCATED(gion)) THEN
       IF (SIZE(gion) /= NS%NTion) DEALLOCATE(gion)
    END IF
    IF (.NOT. ALLOCATED(gel)) ALLOCATE(gel(NS%NTel))
    IF (.NOT. ALLOCATED(gion)) ALLOCATE(gion(NS%NTion))
  END SUBROUTINE Engine_Gather

  SUBROUTINE Engine_Done()
    INTEGER(C_INT) :: rc
    rc = tb_destroy(hnd)
    hnd = C_NULL_PTR
  END SUBROUTINE Engine_Done

  ! The reference's argument list (tree.f90:39-43): gel / gion are the
  ! position copies PTEgathering2 / PTIgathering2 make.  Here the positions
  ! are already resident on the GPU (the last VVerletNotemp1 moved them
  ! there), so the trees are built from the device state and gel / gion are
  ! not read.
  SUBROUTINE BuildLocalTree(NS, gel, gion, Roote, Rooti, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(GS), INTENT(IN) :: gel(NS%NTel), gion(NS%NTion)
    TYPE(TREE_REF), INTENT(OUT) :: Roote, Rooti
    TYPE(TM), INTENT(IN) :: ttime
    CALL tb_check(tb_build(hnd), "BuildLocalTree")
    Roote%species = 0
    Rooti%species = 1
  END SUBROUTINE BuildLocalTree

  SUBROUTINE LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, param, ttime)
    TYPE(NM), INTENT(INOUT) :: NS
    TYPE(TREE_REF), INTENT(IN) :: Roote, Rooti
    TYPE(PT), INTENT(INOUT) :: qel(NS%Nel), qion(NS%Nion)  ! forces stay on the GPU until Engine_Sync
    TYPE(ST), INTENT(INOUT) :: sys
    TYPE(PM), INTENT(IN) :: param
    TYPE(TM), INTENT(IN) :: ttime
    REAL(C_DOUBLE) :: ee, ei
    IF (Roote%species /= 0 .OR. Rooti%species /= 1) STOP "LocalCoulombSum: trees not built"
    CALL tb_check(tb_force(hnd, ee, ei), "LocalCoulombSum")
    sys%Epot_e = ee
    sys%Epot_i = ei
  END SUBROUTINE LocalCoulombSum

  SUBROUTINE VVerletNotemp1(NS, qel, qion, param, dt, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(NS%Nel), qion(NS%Nion)
    TYPE(PM), INTENT(IN) :: param
    REAL(C_DOUBLE), INTENT(IN) :: dt
    TYPE(TM), INTENT(IN) :: ttime
    CALL tb_check(tb_vv1(hnd, dt), "VVerletNotemp1")
  END SUBROUTINE VVerletNotemp1

  SUBROUTINE VVerletNotemp2(NS, qel, qion, param, sys, dt, ttime)
    TYPE(NM), INTENT(IN) :: NS
    TYPE(PT), INTENT(INOUT) :: qel(NS%Nel), qion(NS%Nion)
    TYPE(PM), INTENT(IN) :: param
    TYPE(ST), INTENT(INOUT) :: sys
    REAL(C_DOUBLE), INTENT(IN) :: dt
    TYPE(TM), INTENT(IN) :: ttime
    REAL(C_DOUBLE) :: sums(4)
    INTEGER(C_INT32_T) :: samp
    samp = 0
    IF (MOD(ttime%Nloop, ttime%Nsamp) == 0) samp = 1     ! sums only on sampling steps (vverlet.f90:81)
    CALL tb_check(tb_vv2(hnd, dt, samp, sums), "VVerletNotemp2")
    IF (samp == 1) THEN
       sys%mv2el = sums(1)
       sys%mv2ion = sums(2)
       sys%Te_rad = sums(3)
       sys%Ti_rad = sums(4)
    END IF
    sys%mv2 = sys%mv2ion + sys%mv2el
  END SUBROUTINE VVerletNotemp2
END MODULE TREEMI_FORCE
