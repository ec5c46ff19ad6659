!===============================================================================
! Fortran host over evr_hip_f for two coupled electronic states: the 12-mode
! pyrazine model of Working_tests/MPI_tests/12D_propagation_S0
! (sub_system_pyrazine.f:29-63, calc_f2_f1Q.f90:98-108):
!   T   = sum_i -w_i/2 d^2/dQ_i^2
!   V11 = sum_i w_i Q_i^2/2 - delta + k1.Q,  V22 = ... + delta + k2.Q,
!   V12 = V21 = lambda Q_1                    (eV, times 1/auTOeV)
! on the SG4 basis LB = LG = 1 with Hermite nb_d(l) = nq_d(l) = 1 + B_d l,
! B = (3,3,2,...,2).  The ground vibrational function on state 2 is propagated
! by march_cheby (npoly 300, poly_tol 1e-10) for 60 steps of 0.1 fs; the
! autocorrelation <psi0|psi(t)> is written to the file named by the first
! argument as lines "t Re Im |.|", as the reference's benchmark file.
! The final wavepacket goes to <file>.psi in the reference's save-file format.
!===============================================================================
program evr_f_pyrazine
  use, intrinsic :: iso_c_binding
  use evr_hip_f
  implicit none
  integer, parameter :: D = 12, L = 1, nsteps = 60
  ! CODATA 2006 Eh/e and 0.1 fs in atomic units of time, as the reference's
  ! constants module computes them
  real(c_double), parameter :: auTOeV = 27.211383865564702d0, dt = 4.1341373350251231d0
  real(c_double), parameter :: w(D) = [0.09357d0, 0.0740d0, 0.1273d0, 0.1568d0, 0.1347d0, 0.3431d0, &
                                       0.1157d0, 0.3242d0, 0.3621d0, 0.2673d0, 0.3052d0, 0.0968d0]
  real(c_double), parameter :: k1(D) = [0.0d0, -0.0964d0, 0.0470d0, 0.1594d0, 0.0308d0, 0.0782d0, &
                                        0.0261d0, 0.0717d0, 0.0560d0, 0.0625d0, 0.0780d0, 0.0188d0]
  real(c_double), parameter :: k2(D) = [0.0d0, 0.1194d0, 0.2012d0, 0.0484d0, -0.0308d0, -0.0782d0, &
                                        -0.0261d0, -0.0717d0, -0.0560d0, -0.0625d0, -0.0780d0, -0.0188d0]
  real(c_double), parameter :: delta = 0.46165d0, lambda = 0.1825d0
  integer(c_int) :: lawA(D), lawB(D)
  type(c_ptr) :: desc, p11, p22, p12, hc
  type(c_ptr), target :: blocks(4)
  type(evr_basis_desc), pointer :: bd
  type(evr_op_desc) :: op
  type(evr_propa_opts) :: o
  type(evr_propa_info) :: info
  integer(c_int64_t) :: sz(4), nbc
  integer(c_int) :: nb0, mv
  real(c_double), allocatable, target :: q(:, :), V11(:), V22(:), V12(:), psi(:, :), ac(:, :)
  real(c_double), target :: F2(D, D)
  real(c_double) :: hq
  integer :: nb, npts, i, n, u
  character(len=512) :: out

  call get_command_argument(1, out)
  lawA = 1
  lawB = [3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2]
  call evr_check(evr_sg4_hermite_desc_modes(D, L, L, lawA, lawB, lawA, lawB, desc), 'evr_sg4_hermite_desc_modes')
  call c_f_pointer(desc, bd)
  nb = int(bd%nb_packed)
  call evr_check(evr_basis_desc_sizes(desc, sz), 'evr_basis_desc_sizes')
  npts = int(sz(2))
  allocate(q(D, npts), V11(npts), V22(npts), V12(npts))
  call evr_check(evr_basis_desc_grid(desc, q), 'evr_basis_desc_grid')
  do i = 1, npts
    hq = 0.5d0 * sum(w * q(:, i)**2)
    V11(i) = (hq - delta + sum(k1 * q(:, i))) / auTOeV
    V22(i) = (hq + delta + sum(k2 * q(:, i))) / auTOeV
    V12(i) = lambda * q(1, i) / auTOeV
  end do
  F2 = 0.0d0
  do i = 1, D
    F2(i, i) = -0.5d0 * w(i) / auTOeV
  end do
  op%type = 1
  op%F2 = c_loc(F2)
  op%V = c_loc(V11)
  call evr_check(evr_plan_create(desc, op, 2, p11), 'evr_plan_create H11')
  op%V = c_loc(V22)
  call evr_check(evr_plan_create(desc, op, 2, p22), 'evr_plan_create H22')
  op%type = 0
  op%F2 = c_null_ptr
  op%V = c_loc(V12)
  call evr_check(evr_plan_create(desc, op, 2, p12), 'evr_plan_create V12')
  blocks = [p11, p12, p12, p22]     ! row-major H_ij
  call evr_check(evr_coupled_create(blocks, 2, hc), 'evr_coupled_create')
  call evr_check(evr_coupled_sizes(hc, nbc, nb0, mv), 'evr_coupled_sizes')
  if (nbc /= nb .or. nb0 /= 2) stop 'evr_coupled_sizes'

  allocate(psi(2 * nb, 2), ac(2, nsteps))
  psi = 0.0d0
  psi(nb + 1, 1) = 1.0d0           ! packed function 1 (ground) of state 2
  o%method = EVR_PROPA_CHEBY
  o%dt = dt
  o%nsteps = nsteps
  o%npoly = 300
  o%poly_tol = 1.0d-10
  call evr_check(evr_coupled_propagate(hc, o, 1, c_loc(psi), c_loc(ac), info), 'evr_coupled_propagate')

  open(newunit=u, file=trim(out), status='replace', action='write')
  write(u, '(4es25.17)') 0.0d0, 1.0d0, 0.0d0, 1.0d0
  do n = 1, nsteps
    write(u, '(4es25.17)') n * dt, ac(1, n), ac(2, n), sqrt(ac(1, n)**2 + ac(2, n)**2)
  end do
  close(u)
  ! the final wavepacket in the reference's save-file format (version 0)
  call evr_check(evr_psi_save(desc, 2, 1, 1, c_loc(psi), 0, 1, trim(out)//'.psi'//c_null_char), &
                 'evr_psi_save')
  write(*, '(a,i0,a,i0,a,i0,a)') 'evr_f_pyrazine: ', nb, ' functions x 2 states, ', info%hpsi, &
      ' H psi, max Chebyshev order ', info%max_order, ', done'
  call evr_check(evr_coupled_destroy(hc), 'evr_coupled_destroy')
  call evr_check(evr_plan_destroy(p11), 'evr_plan_destroy')
  call evr_check(evr_plan_destroy(p22), 'evr_plan_destroy')
  call evr_check(evr_plan_destroy(p12), 'evr_plan_destroy')
  call evr_basis_desc_free(desc)
end program evr_f_pyrazine
