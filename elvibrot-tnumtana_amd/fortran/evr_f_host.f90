!===============================================================================
! A small Fortran host over evr_hip_f: what a reference-side caller does.
! D=4 Hermite SG4 basis (LG = LB = 3), type-1 Henon-Heiles operator
! (F2 = -1/2 I, V = 1/2 sum q_i^2 + lambda sum_i (q_i^2 q_{i+1} - q_{i+1}^3/3),
! lambda = 0.111803, the 6D/21D working tests' potential), then:
!   1. H psi for 4 fixed real vectors and 2 complex ones,
!   2. the 8 lowest levels by Davidson,
!   3. 5 Chebyshev steps of a complex wavepacket,
!   4. H psi again on the assembled per-term operators (evr_plan_assemble),
!   5. BasisTOGrid of the 4 vectors (evr_BasisRep_TO_GridRep_n) and the grid
!      derivative d_1 d_2 of those grids (evr_DerivOp_TO_RVecG).
! Results go to the text file named by the first argument.
!===============================================================================
program evr_f_host
  use, intrinsic :: iso_c_binding
  use evr_hip_f
  implicit none
  integer, parameter :: D = 4, nvec = 4, nev = 8
  real(c_double), parameter :: lam = 0.111803d0
  type(c_ptr) :: desc, plan
  type(evr_basis_desc), pointer :: bd
  type(evr_op_desc) :: op
  integer(c_int64_t) :: sz(4)
  real(c_double), allocatable, target :: q(:, :), V(:), psi(:, :), Hpsi(:, :), Vec(:, :), Hasm(:, :)
  real(c_double), allocatable, target :: psiG(:, :), dG(:, :)
  real(c_double) :: abytes
  real(c_double), target :: F2(D, D)
  real(c_double) :: Ene(nev)
  complex(c_double), allocatable :: zpsi(:, :), zH(:, :)
  type(evr_control_opts) :: copts
  integer, parameter :: nfc = 2, nsc = 10
  real(c_double), target :: gE(nfc, nsc), Eold(nfc, nsc), Ezr(nfc, nsc), Ekr(nfc, nsc)
  type(evr_davidson_info) :: dinfo
  type(evr_propa_info) :: pinfo
  integer :: nb, npts, i, k, u
  character(len=512) :: out

  call get_command_argument(1, out)
  call evr_check(evr_sg4_hermite_desc(D, 3, 3, 1, 2, 2, 2, desc), 'evr_sg4_hermite_desc')
  call c_f_pointer(desc, bd)
  nb = int(bd%nb_packed)
  call evr_check(evr_basis_desc_sizes(desc, sz), 'evr_basis_desc_sizes')
  npts = int(sz(2))
  allocate(q(D, npts), V(npts))
  call evr_check(evr_basis_desc_grid(desc, q), 'evr_basis_desc_grid')
  do i = 1, npts
    V(i) = 0.5d0 * sum(q(:, i)**2)
    do k = 1, D - 1
      V(i) = V(i) + lam * (q(k, i)**2 * q(k + 1, i) - q(k + 1, i)**3 / 3.0d0)
    end do
  end do
  F2 = 0.0d0
  do k = 1, D
    F2(k, k) = -0.5d0
  end do
  op%type = 1
  op%V = c_loc(V)
  op%F2 = c_loc(F2)
  call evr_check(evr_plan_create(desc, op, 8, plan), 'evr_plan_create')

  allocate(psi(nb, nvec), Hpsi(nb, nvec), zpsi(nb, 2), zH(nb, 2), Vec(nb, nev), Hasm(nb, nvec))
  do k = 1, nvec
    do i = 1, nb
      psi(i, k) = sin(0.37d0 * i + 1.3d0 * k) / (1.0d0 + 0.01d0 * i)
    end do
  end do
  call evr_sub_TabOpPsi(plan, psi, Hpsi)
  zpsi(:, 1) = cmplx(psi(:, 1), psi(:, 2), kind=c_double)
  zpsi(:, 2) = cmplx(psi(:, 3), -psi(:, 4), kind=c_double)
  call evr_sub_TabOpPsi_cplx(plan, zpsi, zH)
  call evr_propagation_Davidson(plan, nev, 8, 1.0d-10, 1.0d-12, Ene, Vec, dinfo)
  zpsi(:, 1) = zpsi(:, 1) / sqrt(sum(abs(zpsi(:, 1))**2))
  call evr_march_cheby(plan, zpsi(:, 1:1), 0.5d0, 5, 200, 1.0d-14, pinfo)
  call evr_check(evr_plan_assemble(plan, EVR_ASM_FP64), 'evr_plan_assemble')
  if (evr_plan_assembled(plan, abytes) /= EVR_ASM_FP64) stop 'evr_plan_assembled'
  call evr_sub_TabOpPsi(plan, psi, Hasm)
  call evr_check(evr_plan_release_assembled(plan), 'evr_plan_release_assembled')
  allocate(psiG(npts, nvec), dG(npts, nvec))
  call evr_BasisRep_TO_GridRep_n(plan, psi, psiG)
  call evr_DerivOp_TO_RVecG(plan, psiG, dG, [1, 2])

  open(newunit=u, file=trim(out), status='replace', action='write')
  write(u, '(i0,1x,i0)') nb, npts
  do k = 1, nvec
    do i = 1, nb
      write(u, '(es25.17)') Hpsi(i, k)
    end do
  end do
  do k = 1, 2
    do i = 1, nb
      write(u, '(2es25.17)') real(zH(i, k)), aimag(zH(i, k))
    end do
  end do
  do k = 1, nev
    write(u, '(es25.17)') Ene(k)
  end do
  do i = 1, nb
    write(u, '(2es25.17)') real(zpsi(i, 1)), aimag(zpsi(i, 1))
  end do
  do k = 1, nvec
    do i = 1, nb
      write(u, '(es25.17)') Hasm(i, k)
    end do
  end do
  do k = 1, nvec
    do i = 1, npts
      write(u, '(2es25.17)') psiG(i, k), dG(i, k)
    end do
  end do
  close(u)
  ! the control field update (build_field) on a given dJ/dE, Zhu-Rabitz and
  ! Krotov, written to <out>.field
  copts%T = 100.0d0
  copts%nsteps = nsc
  do k = 1, nsc
    do i = 1, nfc
      gE(i, k) = sin(0.3d0 * k + i)
      Eold(i, k) = 0.01d0 * cos(0.7d0 * k - i)
    end do
  end do
  call evr_check(evr_control_build_field(copts, nfc, gE, 2.0d0, 100.0d0, 0, c_null_ptr, Ezr), &
                 'evr_control_build_field')
  call evr_check(evr_control_build_field(copts, nfc, gE, 2.0d0, 100.0d0, 1, c_loc(Eold), Ekr), &
                 'evr_control_build_field')
  open(newunit=u, file=trim(out)//'.field', status='replace', action='write')
  do k = 1, nsc
    do i = 1, nfc
      write(u, '(4es25.17)') gE(i, k), Eold(i, k), Ezr(i, k), Ekr(i, k)
    end do
  end do
  close(u)
  write(*, '(a,i0,a,i0,a)') 'evr_f_host: Davidson ', dinfo%iterations, ' iterations, Chebyshev order ', &
      pinfo%max_order, ', done'
  call evr_check(evr_plan_destroy(plan), 'evr_plan_destroy')
  call evr_basis_desc_free(desc)
end program evr_f_host
