! treemi_capi.f90 -- ISO_C_BINDING interfaces to libtreemi.so (include/treemi.h).
MODULE TREEMI_CAPI
  USE ISO_C_BINDING
  USE DATASTR, ONLY: PT
  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: TB_WALK_REF64 = 0, TB_WALK_MIXED = 1
  INTEGER(C_INT32_T), PARAMETER :: TB_MAC_FIXED = 0, TB_MAC_LEARNED = 1

  TYPE, BIND(C) :: TB_CONFIG
     REAL(C_DOUBLE) :: xm(2), rs, theta, theta_mixed, theta_lo, err_target
     INTEGER(C_INT32_T) :: walk, mac, group, device
     INTEGER(C_INT64_T) :: mem_cap
     REAL(C_DOUBLE) :: far_ratio
     INTEGER(C_INT32_T) :: far_levels
     REAL(C_FLOAT) :: cost_weight
  END TYPE TB_CONFIG

  INTERFACE
     INTEGER(C_INT) FUNCTION tb_create(h, cfg) BIND(C, NAME='tb_create')
       IMPORT :: C_INT, C_PTR, TB_CONFIG
       TYPE(C_PTR), INTENT(OUT) :: h
       TYPE(TB_CONFIG), INTENT(IN) :: cfg
     END FUNCTION tb_create
     INTEGER(C_INT) FUNCTION tb_destroy(h) BIND(C, NAME='tb_destroy')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: h
     END FUNCTION tb_destroy
     INTEGER(C_INT) FUNCTION tb_set_particles(h, ne, e, ni, i) BIND(C, NAME='tb_set_particles')
       IMPORT :: C_INT, C_INT32_T, C_PTR, PT
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT32_T), VALUE :: ne, ni
       TYPE(PT), INTENT(IN) :: e(*), i(*)
     END FUNCTION tb_set_particles
     INTEGER(C_INT) FUNCTION tb_get_particles(h, e, i) BIND(C, NAME='tb_get_particles')
       IMPORT :: C_INT, C_PTR, PT
       TYPE(C_PTR), VALUE :: h
       TYPE(PT), INTENT(INOUT) :: e(*), i(*)
     END FUNCTION tb_get_particles
     INTEGER(C_INT) FUNCTION tb_build(h) BIND(C, NAME='tb_build')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: h
     END FUNCTION tb_build
     INTEGER(C_INT) FUNCTION tb_force(h, epot_e, epot_i) BIND(C, NAME='tb_force')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), INTENT(OUT) :: epot_e, epot_i
     END FUNCTION tb_force
     INTEGER(C_INT) FUNCTION tb_vv1(h, dt) BIND(C, NAME='tb_vv1')
       IMPORT :: C_INT, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), VALUE :: dt
     END FUNCTION tb_vv1
     INTEGER(C_INT) FUNCTION tb_vv2(h, dt, sample, sums) BIND(C, NAME='tb_vv2')
       IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), VALUE :: dt
       INTEGER(C_INT32_T), VALUE :: sample
       REAL(C_DOUBLE), INTENT(OUT) :: sums(4)
     END FUNCTION tb_vv2
     INTEGER(C_INT) FUNCTION tb_rdf_enable(h, dx_ei, dx_ii, nbin) BIND(C, NAME='tb_rdf_enable')
       IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), VALUE :: dx_ei, dx_ii
       INTEGER(C_INT32_T), VALUE :: nbin
     END FUNCTION tb_rdf_enable
     INTEGER(C_INT) FUNCTION tb_rdf_get(h, nbin, rdf_ei, rdf_ii, reset) BIND(C, NAME='tb_rdf_get')
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT32_T), VALUE :: nbin, reset
       INTEGER(C_INT64_T), INTENT(OUT) :: rdf_ei(*), rdf_ii(*)
     END FUNCTION tb_rdf_get
     TYPE(C_PTR) FUNCTION tb_last_error() BIND(C, NAME='tb_last_error')
       IMPORT :: C_PTR
     END FUNCTION tb_last_error
     REAL(C_DOUBLE) FUNCTION tb_kelbg_length(te, xme, xmi) BIND(C, NAME='tb_kelbg_length')
       IMPORT :: C_DOUBLE
       REAL(C_DOUBLE), VALUE :: te, xme, xmi
     END FUNCTION tb_kelbg_length
     INTEGER(C_INT) FUNCTION tb_gen_plasma(seed, ne, ni, sigma, te, ti, xme, xmi, e, i) &
          BIND(C, NAME='tb_gen_plasma')
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, PT
       INTEGER(C_INT64_T), VALUE :: seed
       INTEGER(C_INT32_T), VALUE :: ne, ni
       REAL(C_DOUBLE), VALUE :: sigma, te, ti, xme, xmi
       TYPE(PT), INTENT(INOUT) :: e(*), i(*)
     END FUNCTION tb_gen_plasma
  END INTERFACE

CONTAINS

  ! the reference STOPs on errors; so do the bindings, with the library's message
  SUBROUTINE tb_check(rc, where)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: k, n
    IF (rc == 0) RETURN
    CALL C_F_POINTER(tb_last_error(), msg, [1024])
    n = 0
    DO k = 1, 1024
       IF (msg(k) == C_NULL_CHAR) EXIT
       n = k
    END DO
    PRINT '(A, A, I0, A, 1024A)', where, ": tb error ", rc, ": ", msg(1:n)
    STOP 1
  END SUBROUTINE tb_check
END MODULE TREEMI_CAPI
