! Test driver of the Fortran binding (valence_amd/fortran/vsvb_hip.F90).
!   vsvb_drive --layout           sizes and component offsets of the bind(C) types
!   vsvb_drive <system file>      energy and gradient of the system through the C ABI
! The system file is the vsvb_system as raw stream data: the 14 int32 sizes,
! then every array in the order of the struct (see tests/fortran/dump.py).
! Prints "energy", "energy_grad" and "grad_norm" lines; writes the gradient
! to <system file>.grad as raw float64.
program vsvb_drive
  use iso_c_binding
  use vsvb_hip
  implicit none
  integer(c_int32_t) :: hdr(14)
  integer(c_int32_t), allocatable, target :: atom_type(:), tss(:), shell_l(:), sps(:), orb_natom(:), &
       orb_atoms(:), orb_start(:), xpset(:), pair_sc(:)
  real(c_double), allocatable, target :: coords(:), nuc(:), pexp(:), pcoef(:), coeff(:), coeff_sc(:)
  real(c_double), allocatable :: grad(:)
  real(c_double), target :: parts(4)
  type(vsvb_system_t), target :: sys
  type(vsvb_options_t), target :: opt
  type(c_ptr) :: ctx
  real(c_double) :: e, eg
  character(len=1024) :: arg
  integer :: u, nsc, natom, ntype, nsht, nprim, norbs, mxctr, totlen, npair
  integer(c_int) :: rc

  call get_command_argument(1, arg)
  if (trim(arg) == '--layout') then
     print '(a,i0)', 'sizeof vsvb_system ', c_sizeof(sys)
     print '(a,i0)', 'offset coords ', off(c_loc(sys%coords))
     print '(a,i0)', 'offset xpset ', off(c_loc(sys%xpset))
     print '(a,i0)', 'offset coeff_sc ', off(c_loc(sys%coeff_sc))
     print '(a,i0)', 'sizeof vsvb_options ', c_sizeof(opt)
     print '(a,i0)', 'offset grid_inverse_min ', &
          transfer(c_loc(opt%grid_inverse_min), 0_c_intptr_t) - transfer(c_loc(opt), 0_c_intptr_t)
     print '(a,i0)', 'offset beta_only_last ', &
          transfer(c_loc(opt%beta_only_last), 0_c_intptr_t) - transfer(c_loc(opt), 0_c_intptr_t)
     stop
  end if

  open(newunit=u, file=trim(arg), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr
  natom = hdr(1); ntype = hdr(2); nsht = hdr(3); nprim = hdr(4); norbs = hdr(5); mxctr = hdr(6)
  totlen = hdr(7); npair = hdr(8); nsc = max(hdr(12), 1)
  allocate(atom_type(natom), coords(3 * natom), nuc(ntype), tss(ntype + 1), shell_l(nsht), sps(nsht + 1), &
       pexp(nprim), pcoef(nprim), orb_natom(norbs), orb_atoms(norbs * mxctr), orb_start(norbs + 1), &
       xpset(totlen), coeff(totlen), pair_sc(max(nsc * 2 * npair, 1)), coeff_sc(nsc))
  read(u) atom_type, coords, nuc, tss, shell_l, sps, pexp, pcoef, orb_natom, orb_atoms, orb_start, xpset, coeff
  if (npair > 0) read(u) pair_sc(1:nsc * 2 * npair)
  read(u) coeff_sc
  close(u)

  sys = vsvb_system_t(hdr(1), hdr(2), hdr(3), hdr(4), hdr(5), hdr(6), hdr(7), hdr(8), hdr(9), hdr(10), &
       hdr(11), hdr(12), hdr(13), hdr(14), c_loc(atom_type), c_loc(coords), c_loc(nuc), c_loc(tss), &
       c_loc(shell_l), c_loc(sps), c_loc(pexp), c_loc(pcoef), c_loc(orb_natom), c_loc(orb_atoms), &
       c_loc(orb_start), c_loc(xpset), c_loc(coeff), c_loc(pair_sc), c_loc(coeff_sc))
  rc = vsvb_context_create(sys, opt, ctx)
  call check(rc, 'vsvb_context_create')
  rc = vsvb_energy(ctx, c_null_ptr, c_null_ptr, e, c_loc(parts))
  call check(rc, 'vsvb_energy')
  allocate(grad(totlen))
  rc = vsvb_energy_grad(ctx, c_null_ptr, c_null_ptr, eg, grad)
  call check(rc, 'vsvb_energy_grad')
  call vsvb_context_destroy(ctx)
  print '(a,es25.17)', 'energy ', e
  print '(a,es25.17)', 'energy_grad ', eg
  print '(a,es25.17)', 'grad_norm ', sqrt(sum(grad**2))
  open(newunit=u, file=trim(arg)//'.grad', access='stream', form='unformatted', status='replace', action='write')
  write(u) grad
  close(u)

contains

  integer function off(p)
    type(c_ptr), intent(in) :: p
    off = int(transfer(p, 0_c_intptr_t) - transfer(c_loc(sys), 0_c_intptr_t))
  end function off

  subroutine check(code, where)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: where
    if (code == VSVB_OK) return
    print '(a,a,a,i0,a,a)', 'error: ', where, ' returned ', code, ': ', vsvb_error_text()
    error stop 1
  end subroutine check

end program vsvb_drive
