! datastr.f90 -- data types of the Fortran driver (this project's own module).
! Field names and meanings follow the reference's MODULE DATASTR
! (data.f90:29-143) so that driver code reads the same; PT is BIND(C) so that
! arrays of it pass straight to the C-ABI as tb_pt (include/treemi.h).
MODULE DATASTR
  USE ISO_C_BINDING
  IMPLICIT NONE
  INTEGER, PARAMETER :: Nparam = 2, Nref = 6, Ndist = 10000
  REAL(C_DOUBLE), PARAMETER :: tps = 1.0180505305371923D-1   ! ns per time unit
  REAL(C_DOUBLE), PARAMETER :: eps = 0.0014399835661145824D0 ! eV*um
  REAL(C_DOUBLE), PARAMETER :: kboltz = 8.617343D-5          ! eV/K

  ! particle: position, velocity, force, id (+ padding to 80 bytes)
  TYPE, BIND(C) :: PT
     REAL(C_DOUBLE) :: xx(3), xv(3), ff(3)
     INTEGER(C_INT32_T) :: id, pad
  END TYPE PT

  ! position-only record (data.f90:54-56): what PTEgathering2 / PTIgathering2
  ! hand to BuildLocalTree
  TYPE GS
     REAL(C_DOUBLE) :: xx(3)
  END TYPE GS

  ! time control
  TYPE TM
     INTEGER :: Nloop, Nrest, Ndump, Nsamp
     REAL(C_DOUBLE) :: tmax, trest, tdump, tsamp, tnow, self, comm, other
  END TYPE TM

  ! parameters: charges (read, unused as in the reference), temperatures,
  ! masses, Kelbg length, opening angle, restart flag
  TYPE PM
     REAL(C_DOUBLE) :: q(Nparam), Te, Ti, xm(Nparam), rs, sd_ratio
     CHARACTER(LEN=3) :: rest
  END TYPE PM

  ! system sums and temperature normalisations
  TYPE ST
     REAL(C_DOUBLE) :: temp, Te, Ti, mv2, mv2el, mv2ion, Epot, Epot_e, Epot_i, &
          kbm, kbm_e, kbm_i, kbm_e_rad, kbm_e_cir, kbm_i_rad, kbm_i_cir, rs, Te_rad, Ti_rad
     CHARACTER(LEN=5) :: kelbg
  END TYPE ST

  ! counts and RDF bins
  TYPE NM
     INTEGER :: Nion, Nel, NTion, NTel, NTpt, Npt
     REAL(C_DOUBLE) :: dx_ee, dx_ei, dx_ii
  END TYPE NM
END MODULE DATASTR
