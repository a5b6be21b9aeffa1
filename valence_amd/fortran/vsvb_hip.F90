! ISO_C_BINDING interface to libvsvb_hip.so (include/vsvb_hip.h) for the
! Fortran host of VALENCE.  The types mirror vsvb_system and vsvb_options
! field by field; every function returns 0 or a negative VSVB_ERR_* code, and
! vsvb_error_text() gives the library's message for the last failure.
! The reference passes its module state (src/integrals_module.F90,
! src/density_module.F90) to these calls; see INTEGRATION.md for the call sites.
module vsvb_hip
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: VSVB_OK = 0, VSVB_ERR_ARG = -1, VSVB_ERR_DEVICE = -2, &
                               VSVB_ERR_NUMERIC = -3, VSVB_ERR_UNSUPPORTED = -4

  type, bind(C) :: vsvb_system_t
     integer(c_int32_t) :: natom, ntype, nshell_type, nprim, norbs, mxctr, totlen
     integer(c_int32_t) :: npair, nunpd, ndocc, ndf, nspinc, ntol_d, ntol_i
     type(c_ptr) :: atom_type, coords, nuc_charge, type_shell_start, shell_l, shell_prim_start
     type(c_ptr) :: prim_exp, prim_coef, orb_natom, orb_atoms, orb_start, xpset, coeff
     type(c_ptr) :: pair_sc, coeff_sc
  end type vsvb_system_t

  type, bind(C) :: vsvb_options_t
     real(c_double)     :: quartet_tau = 0.0_c_double   ! <= 0: library default 1e-17
     integer(c_int32_t) :: device = 0
     integer(c_int32_t) :: pair_chunk = 0             ! 0: claim width from the system size
     integer(c_int32_t) :: grid_inverse_min = 0       ! 0: overlap blocks from 129 rows up inverted grid-wide
     integer(c_int32_t) :: beta_only_last = 0         ! 1: last docc orbital beta only (first-order step)
  end type vsvb_options_t

  interface
     integer(c_int) function vsvb_context_create(sys, opt, ctx) bind(C, name="vsvb_context_create")
       import :: c_int, c_ptr, vsvb_system_t, vsvb_options_t
       type(vsvb_system_t), intent(in) :: sys
       type(vsvb_options_t), intent(in) :: opt
       type(c_ptr), intent(out) :: ctx
     end function vsvb_context_create
     subroutine vsvb_context_destroy(ctx) bind(C, name="vsvb_context_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine vsvb_context_destroy
     integer(c_int) function vsvb_energy(ctx, coords, coeff, energy, parts) bind(C, name="vsvb_energy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, coords, coeff, parts
       real(c_double), intent(out) :: energy
     end function vsvb_energy
     integer(c_int) function vsvb_energy_parts(ctx, coeff, num, wfnorm) bind(C, name="vsvb_energy_parts")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, coeff
       real(c_double), intent(out) :: num, wfnorm
     end function vsvb_energy_parts
     integer(c_int) function vsvb_energy_grad(ctx, coords, coeff, energy, grad) bind(C, name="vsvb_energy_grad")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, coords, coeff
       real(c_double), intent(out) :: energy
       real(c_double), intent(out) :: grad(*)
     end function vsvb_energy_grad
     integer(c_int) function vsvb_spin_matrices(ctx, coeff, ham, ovl) bind(C, name="vsvb_spin_matrices")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, coeff
       real(c_double), intent(out) :: ham(*), ovl(*)
     end function vsvb_spin_matrices
     integer(c_int) function vsvb_spin_opt(ctx, coeff, energy, coeff_sc) bind(C, name="vsvb_spin_opt")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, coeff
       real(c_double), intent(out) :: energy, coeff_sc(*)
     end function vsvb_spin_opt
     integer(c_int) function vsvb_set_spin_coupling(ctx, coeff_sc) bind(C, name="vsvb_set_spin_coupling")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: coeff_sc(*)
     end function vsvb_set_spin_coupling
     integer(c_int) function vsvb_surface(ctx, x, v) bind(C, name="vsvb_surface")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, x
       real(c_double), intent(out) :: v
     end function vsvb_surface
     type(c_ptr) function vsvb_last_error() bind(C, name="vsvb_last_error")
       import :: c_ptr
     end function vsvb_last_error
     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! The library's message for its last failure, as a Fortran string.
  function vsvb_error_text() result(text)
    character(len=:), allocatable :: text
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = vsvb_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [max(n, 1)])
    allocate(character(len=n) :: text)
    do i = 1, n
       text(i:i) = chars(i)
    end do
  end function vsvb_error_text

end module vsvb_hip
