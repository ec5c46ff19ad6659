!===============================================================================
! evr_hip_f -- ISO_C_BINDING layer of libevrhip.so for a Fortran host.
!
! Every C entry point of include/evr_hip.h has an interface here (same name,
! same argument order); the bind(C) derived types mirror the C structs field
! by field.  The module procedures evr_sub_TabOpPsi, evr_BasisRep_TO_GridRep,
! evr_GridRep_TO_BasisRep, evr_propagation_Davidson and evr_march_cheby wrap
! them with the argument conventions of the reference's routines:
!   sub_TabOpPsi               Source_ElVibRot/sub_Operator/sub_OpPsi.f90:701
!   sub_PsiBasisRep_TO_GridRep Source_ElVibRot/sub_WP/sub_module_psi_B_TO_G.f90:60
!   sub_PsiGridRep_TO_BasisRep Source_ElVibRot/sub_WP/sub_module_psi_B_TO_G.f90:178
!   sub_propagation_Davidson   Source_ElVibRot/sub_propagation/sub_module_Davidson.f90:54
!   march_cheby                Source_ElVibRot/sub_propagation/sub_module_propa_march.f90:4142
! A nonzero status is fatal in the wrappers (message, then STOP), as the
! reference STOPs on the same conditions; the raw interfaces return it.
!===============================================================================
module evr_hip_f
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: EVR_OK = 0, EVR_FP64 = 0, EVR_FP32 = 1, EVR_DEVICE_PTRS = 2
  integer(c_int), parameter :: EVR_PROPA_CHEBY = 1, EVR_PROPA_SIL = 2, EVR_PROPA_TAYLOR = 3, EVR_PROPA_RK4 = 4
  integer(c_int), parameter :: EVR_MATRIX_FREE = 4, EVR_AUTO_PATH = 8, EVR_ASM_FP64 = 1, EVR_ASM_FP32 = 2

  type, bind(C) :: evr_basis_desc
    integer(c_int)     :: D, nlev
    integer(c_int64_t) :: nterms
    type(c_ptr)        :: term_levels, weights, nb, nq, x, B, W, D1, D2
    integer(c_int64_t) :: nb_packed
    type(c_ptr)        :: map
    integer(c_int)     :: grid_last_fastest
  end type evr_basis_desc

  type, bind(C) :: evr_op_desc
    integer(c_int) :: type = 10
    type(c_ptr)    :: V = c_null_ptr, G0 = c_null_ptr, Ga = c_null_ptr
    type(c_ptr)    :: F2 = c_null_ptr, F1 = c_null_ptr
    integer(c_int) :: hermitian = 0
  end type evr_op_desc

  type, bind(C) :: evr_davidson_opts
    integer(c_int) :: nev = 1, block = 16, max_sub = 0, max_it = 0
    real(c_double) :: conv_resi = 1.0d-8, conv_ene = 1.0d-9
    integer(c_int) :: precond = 1
    type(c_ptr)    :: diag = c_null_ptr, guess = c_null_ptr
    integer(c_int) :: nguess = 0, precision = EVR_FP64
    real(c_double) :: refine_resi = 1.0d-3
    integer(c_int) :: max_nvec_hint = 0, verbose = 0
    type(c_ptr)    :: basis = c_null_ptr, guess_file = c_null_ptr, save_file = c_null_ptr
    integer(c_int) :: graph = 0
  end type evr_davidson_opts

  type, bind(C) :: evr_davidson_info
    integer(c_int)     :: iterations, converged, nconv
    integer(c_int64_t) :: hpsi_fp64, hpsi_fp32
    real(c_double)     :: max_resi, seconds
  end type evr_davidson_info

  type, bind(C) :: evr_propa_opts
    integer(c_int) :: method = EVR_PROPA_CHEBY
    real(c_double) :: dt = 0.0d0
    integer(c_int) :: nsteps = 1, npoly = 200
    real(c_double) :: poly_tol = 1.0d-12, Hmin = 0.0d0, Hmax = 0.0d0
    type(c_ptr)    :: field = c_null_ptr, mu = c_null_ptr, psi0 = c_null_ptr
    integer(c_int) :: precision = EVR_FP64, device_ptrs = 0, graph = 0
    type(c_ptr)    :: field_sub = c_null_ptr
    integer(c_int) :: nsub = 0
  end type evr_propa_opts

  type, bind(C) :: evr_propa_info
    integer(c_int64_t) :: hpsi
    integer(c_int)     :: max_order, retries
    real(c_double)     :: Hmin, Hmax, seconds
  end type evr_propa_info

  type, bind(C) :: evr_control_opts
    real(c_double) :: T = 0.0d0
    integer(c_int) :: nsteps = 1, nk = 1, checkpoint = 0, npoly = 200
    real(c_double) :: poly_tol = 1.0d-14, Hmin = 0.0d0, Hmax = 0.0d0
    integer(c_int) :: precision = EVR_FP64, graph = 0
  end type evr_control_opts

  type, bind(C) :: evr_control_info
    integer(c_int64_t) :: hpsi
    integer(c_int)     :: max_order, retries
    real(c_double)     :: Hmin, Hmax, seconds_forward, seconds_backward
  end type evr_control_info

  interface
    function evr_last_error() bind(C, name="evr_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function evr_sg4_hermite_desc(D, LG, LB, nbA, nbB, nqA, nqB, out) bind(C, name="evr_sg4_hermite_desc")
      import :: c_int, c_ptr
      integer(c_int), value :: D, LG, LB, nbA, nbB, nqA, nqB
      type(c_ptr), intent(out) :: out
      integer(c_int) :: evr_sg4_hermite_desc
    end function
    function evr_sg4_hermite_desc_modes(D, LG, LB, nbA, nbB, nqA, nqB, out) &
        bind(C, name="evr_sg4_hermite_desc_modes")
      import :: c_int, c_ptr
      integer(c_int), value :: D, LG, LB
      integer(c_int), intent(in) :: nbA(*), nbB(*), nqA(*), nqB(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: evr_sg4_hermite_desc_modes
    end function
    ! wavepacket files (sub_save_psi / sub_read_psi0 formats); path is a
    ! C string: trim(name)//c_null_char
    function evr_psi_save(basis, nb0, nwp, cplx, psi, version, formatted, path) &
        bind(C, name="evr_psi_save")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: basis, psi
      integer(c_int), value :: nb0, nwp, cplx, version, formatted
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: evr_psi_save
    end function
    function evr_psi_read(basis, nb0, cplx, max_wp, psi, nwp, path) bind(C, name="evr_psi_read")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: basis, psi
      integer(c_int), value :: nb0, cplx, max_wp
      integer(c_int), intent(out) :: nwp
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: evr_psi_read
    end function
    function evr_dp_hermite_desc(D, nb, nq, out) bind(C, name="evr_dp_hermite_desc")
      import :: c_int, c_ptr
      integer(c_int), value :: D
      integer(c_int), intent(in) :: nb(*), nq(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: evr_dp_hermite_desc
    end function
    subroutine evr_basis_desc_free(d) bind(C, name="evr_basis_desc_free")
      import :: c_ptr
      type(c_ptr), value :: d
    end subroutine
    function evr_basis_desc_sizes(d, out) bind(C, name="evr_basis_desc_sizes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: d
      integer(c_int64_t), intent(out) :: out(4)
      integer(c_int) :: evr_basis_desc_sizes
    end function
    function evr_basis_desc_grid(d, q) bind(C, name="evr_basis_desc_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: d
      real(c_double), intent(out) :: q(*)
      integer(c_int) :: evr_basis_desc_grid
    end function
    function evr_plan_create(basis, op, max_nvec, out) bind(C, name="evr_plan_create")
      import :: c_int, c_ptr, evr_op_desc
      type(c_ptr), value :: basis
      type(evr_op_desc), intent(in) :: op
      integer(c_int), value :: max_nvec
      type(c_ptr), intent(out) :: out
      integer(c_int) :: evr_plan_create
    end function
    function evr_plan_destroy(p) bind(C, name="evr_plan_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: evr_plan_destroy
    end function
    function evr_plan_set_potential(p, V) bind(C, name="evr_plan_set_potential")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(in) :: V(*)
      integer(c_int) :: evr_plan_set_potential
    end function
    ! G(sum_nq, D, D) = GGiq per grid point; Jac, sqRhoOVERJac: c_loc of
    ! sum_nq arrays or c_null_ptr (= 1)
    function evr_plan_set_keo_grid(p, G, Jac, sqRhoOVERJac) bind(C, name="evr_plan_set_keo_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p, Jac, sqRhoOVERJac
      real(c_double), intent(in) :: G(*)
      integer(c_int) :: evr_plan_set_keo_grid
    end function
    function evr_plan_gen_potential_poly(p, v2, c3, d22, e4) bind(C, name="evr_plan_gen_potential_poly")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(in) :: v2(*), c3(*), d22(*), e4(*)
      integer(c_int) :: evr_plan_gen_potential_poly
    end function
    function evr_plan_sizes(p, out) bind(C, name="evr_plan_sizes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: p
      integer(c_int64_t), intent(out) :: out(7)
      integer(c_int) :: evr_plan_sizes
    end function
    function evr_plan_stream(p) bind(C, name="evr_plan_stream") result(s)
      import :: c_ptr
      type(c_ptr), value :: p
      type(c_ptr) :: s
    end function
    function evr_plan_max_nvec(p) bind(C, name="evr_plan_max_nvec")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: evr_plan_max_nvec
    end function
    function evr_plan_set_profiling(p, on) bind(C, name="evr_plan_set_profiling")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int), value :: on
      integer(c_int) :: evr_plan_set_profiling
    end function
    function evr_plan_kernel_stats(p, max, names, ms, flops, launches, n) bind(C, name="evr_plan_kernel_stats")
      import :: c_int, c_ptr, c_double, c_int64_t, c_char
      type(c_ptr), value :: p
      integer(c_int), value :: max
      character(kind=c_char), intent(out) :: names(*)
      real(c_double), intent(out) :: ms(*), flops(*)
      integer(c_int64_t), intent(out) :: launches(*)
      integer(c_int), intent(out) :: n
      integer(c_int) :: evr_plan_kernel_stats
    end function
    function evr_oppsi(p, nvec, psi, hpsi, flags) bind(C, name="evr_oppsi")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, psi, hpsi
      integer(c_int), value :: nvec, flags
      integer(c_int) :: evr_oppsi
    end function
    function evr_plan_diagonal(p, diag, flags) bind(C, name="evr_plan_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, diag
      integer(c_int), value :: flags
      integer(c_int) :: evr_plan_diagonal
    end function
    function evr_basis_to_grid(p, psiB, psiG) bind(C, name="evr_basis_to_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(in) :: psiB(*)
      real(c_double), intent(out) :: psiG(*)
      integer(c_int) :: evr_basis_to_grid
    end function
    function evr_grid_to_basis(p, psiG, psiB) bind(C, name="evr_grid_to_basis")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(in) :: psiG(*)
      real(c_double), intent(out) :: psiB(*)
      integer(c_int) :: evr_grid_to_basis
    end function
    function evr_basis_to_grid_n(p, ncol, psiB, psiG, flags) bind(C, name="evr_basis_to_grid_n")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, psiB, psiG
      integer(c_int), value :: ncol, flags
      integer(c_int) :: evr_basis_to_grid_n
    end function
    function evr_grid_to_basis_n(p, ncol, psiG, psiB, flags) bind(C, name="evr_grid_to_basis_n")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, psiG, psiB
      integer(c_int), value :: ncol, flags
      integer(c_int) :: evr_grid_to_basis_n
    end function
    function evr_grid_deriv(p, ncol, tab_der, psiG, dpsiG, flags) bind(C, name="evr_grid_deriv")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, psiG, dpsiG
      integer(c_int), value :: ncol, flags
      integer(c_int), intent(in) :: tab_der(2)
      integer(c_int) :: evr_grid_deriv
    end function
    function evr_plan_set_type1_grid(p, F2, F1) bind(C, name="evr_plan_set_type1_grid")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, F2, F1
      integer(c_int) :: evr_plan_set_type1_grid
    end function
    function evr_propa_save_restart(T, nwp, size_wp, psi, path) bind(C, name="evr_propa_save_restart")
      import :: c_int, c_int64_t, c_double, c_ptr, c_char
      real(c_double), value :: T
      integer(c_int), value :: nwp
      integer(c_int64_t), value :: size_wp
      type(c_ptr), value :: psi
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: evr_propa_save_restart
    end function
    function evr_propa_read_restart(path, nwp, size_wp, psi, T, found) bind(C, name="evr_propa_read_restart")
      import :: c_int, c_int64_t, c_double, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nwp
      integer(c_int64_t), value :: size_wp
      type(c_ptr), value :: psi
      real(c_double), intent(out) :: T
      integer(c_int), intent(out) :: found
      integer(c_int) :: evr_propa_read_restart
    end function
    function evr_control_field_sub(opts, nf, amp, method, nsub, field_sub) bind(C, name="evr_control_field_sub")
      import :: c_int, c_double, evr_control_opts
      type(evr_control_opts), intent(in) :: opts
      integer(c_int), value :: nf, method, nsub
      real(c_double), intent(in) :: amp(*)
      real(c_double), intent(out) :: field_sub(*)
      integer(c_int) :: evr_control_field_sub
    end function
    function evr_plan_autotune(p, nvec, flags, reps, chosen, ms_static, ms_tuned) bind(C, name="evr_plan_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      integer(c_int), value :: nvec, flags, reps
      integer(c_int), intent(out) :: chosen(*)
      real(c_double), intent(out) :: ms_static, ms_tuned
      integer(c_int) :: evr_plan_autotune
    end function
    function evr_davidson(p, opts, evals, evecs, info) bind(C, name="evr_davidson")
      import :: c_int, c_ptr, c_double, evr_davidson_opts, evr_davidson_info
      type(c_ptr), value :: p, evecs
      type(evr_davidson_opts), intent(in) :: opts
      real(c_double), intent(out) :: evals(*)
      type(evr_davidson_info), intent(out) :: info
      integer(c_int) :: evr_davidson
    end function
    function evr_propagate(p, opts, nwp, psi, autocorr, info) bind(C, name="evr_propagate")
      import :: c_int, c_ptr, evr_propa_opts, evr_propa_info
      type(c_ptr), value :: p, psi, autocorr
      type(evr_propa_opts), intent(in) :: opts
      integer(c_int), value :: nwp
      type(evr_propa_info), intent(out) :: info
      integer(c_int) :: evr_propagate
    end function
    function evr_plan_assemble(p, which) bind(C, name="evr_plan_assemble")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int), value :: which
      integer(c_int) :: evr_plan_assemble
    end function
    function evr_plan_assembled(p, bytes) bind(C, name="evr_plan_assembled")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(out) :: bytes
      integer(c_int) :: evr_plan_assembled
    end function
    function evr_plan_release_assembled(p) bind(C, name="evr_plan_release_assembled")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: evr_plan_release_assembled
    end function
    function evr_multi_create(plans, n, max_nvec, m) bind(C, name="evr_multi_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: plans(*)
      integer(c_int), value :: n, max_nvec
      type(c_ptr), intent(out) :: m
      integer(c_int) :: evr_multi_create
    end function
    function evr_multi_destroy(m) bind(C, name="evr_multi_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int) :: evr_multi_destroy
    end function
    function evr_multi_offsets(m, nb_offsets) bind(C, name="evr_multi_offsets")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: m
      integer(c_int64_t), intent(out) :: nb_offsets(*)
      integer(c_int) :: evr_multi_offsets
    end function
    function evr_multi_count(m) bind(C, name="evr_multi_count")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int) :: evr_multi_count
    end function
    function evr_multi_stream(m) bind(C, name="evr_multi_stream")
      import :: c_ptr
      type(c_ptr), value :: m
      type(c_ptr) :: evr_multi_stream
    end function
    function evr_multi_oppsi(m, nvec, psi, hpsi, flags) bind(C, name="evr_multi_oppsi")
      import :: c_int, c_ptr
      type(c_ptr), value :: m, psi, hpsi
      integer(c_int), value :: nvec, flags
      integer(c_int) :: evr_multi_oppsi
    end function
    function evr_multi_propagate(m, dt, nsteps, npoly, poly_tol, Hmin, Hmax, psi, device_ptrs, info) &
        bind(C, name="evr_multi_propagate")
      import :: c_int, c_ptr, c_double, evr_propa_info
      type(c_ptr), value :: m, psi
      real(c_double), value :: dt, poly_tol
      integer(c_int), value :: nsteps, npoly, device_ptrs
      real(c_double), intent(in) :: Hmin(*), Hmax(*)
      type(evr_propa_info), intent(out) :: info
      integer(c_int) :: evr_multi_propagate
    end function
    ! nb0 coupled electronic states.  blocks is row-major: blocks((i-1)*nb0 + j)
    ! holds the plan of H_ij (C_NULL_PTR for a zero block); from a Fortran
    ! matrix B(i, j) of plans pass reshape(transpose(B), [nb0*nb0])
    function evr_coupled_create(blocks, nb0, c) bind(C, name="evr_coupled_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: blocks(*)
      integer(c_int), value :: nb0
      type(c_ptr), intent(out) :: c
      integer(c_int) :: evr_coupled_create
    end function
    function evr_coupled_destroy(c) bind(C, name="evr_coupled_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: evr_coupled_destroy
    end function
    function evr_coupled_sizes(c, nb, nb0, max_nvec) bind(C, name="evr_coupled_sizes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: c
      integer(c_int64_t), intent(out) :: nb
      integer(c_int), intent(out) :: nb0, max_nvec
      integer(c_int) :: evr_coupled_sizes
    end function
    function evr_coupled_stream(c) bind(C, name="evr_coupled_stream")
      import :: c_ptr
      type(c_ptr), value :: c
      type(c_ptr) :: evr_coupled_stream
    end function
    function evr_coupled_oppsi(c, nvec, psi, hpsi, flags) bind(C, name="evr_coupled_oppsi")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, psi, hpsi
      integer(c_int), value :: nvec, flags
      integer(c_int) :: evr_coupled_oppsi
    end function
    function evr_coupled_propagate(c, opts, nwp, psi, autocorr, info) bind(C, name="evr_coupled_propagate")
      import :: c_int, c_ptr, evr_propa_opts, evr_propa_info
      type(c_ptr), value :: c, psi, autocorr
      type(evr_propa_opts), intent(in) :: opts
      integer(c_int), value :: nwp
      type(evr_propa_info), intent(out) :: info
      integer(c_int) :: evr_coupled_propagate
    end function
    function evr_coupled_diagonal(c, diag, flags) bind(C, name="evr_coupled_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, diag
      integer(c_int), value :: flags
      integer(c_int) :: evr_coupled_diagonal
    end function
    function evr_coupled_davidson(c, opts, evals, evecs, info) bind(C, name="evr_coupled_davidson")
      import :: c_int, c_ptr, evr_davidson_opts, evr_davidson_info
      type(c_ptr), value :: c, evals, evecs
      type(evr_davidson_opts), intent(in) :: opts
      type(evr_davidson_info), intent(out) :: info
      integer(c_int) :: evr_coupled_davidson
    end function
    function evr_coupled_propagate_field(c, mu, opts, nwp, psi, autocorr, info) &
        bind(C, name="evr_coupled_propagate_field")
      import :: c_int, c_ptr, evr_propa_opts, evr_propa_info
      type(c_ptr), value :: c, mu, psi, autocorr
      type(evr_propa_opts), intent(in) :: opts
      integer(c_int), value :: nwp
      type(evr_propa_info), intent(out) :: info
      integer(c_int) :: evr_coupled_propagate_field
    end function
    function evr_plan_bounds(p, steps, Hmin, Hmax) bind(C, name="evr_plan_bounds")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      integer(c_int), value :: steps
      real(c_double), intent(out) :: Hmin, Hmax
      integer(c_int) :: evr_plan_bounds
    end function
    function evr_bessel_j(n, r, out) bind(C, name="evr_bessel_j")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), value :: r
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: evr_bessel_j
    end function
    function evr_control_field(opts, nf, amp, field) bind(C, name="evr_control_field")
      import :: c_int, c_double, evr_control_opts
      type(evr_control_opts), intent(in) :: opts
      integer(c_int), value :: nf
      real(c_double), intent(in) :: amp(*)
      real(c_double), intent(out) :: field(*)
      integer(c_int) :: evr_control_field
    end function
    ! the reference's field update (build_field, CTRL:902-971): field_old is
    ! c_loc of the old field (krotov = 1) or c_null_ptr
    function evr_control_build_field(opts, nf, grad_E, alpha, Tenv, krotov, field_old, field_new) &
        bind(C, name="evr_control_build_field")
      import :: c_int, c_double, c_ptr, evr_control_opts
      type(evr_control_opts), intent(in) :: opts
      integer(c_int), value :: nf, krotov
      real(c_double), intent(in) :: grad_E(*)
      real(c_double), value :: alpha, Tenv
      type(c_ptr), value :: field_old
      real(c_double), intent(out) :: field_new(*)
      integer(c_int) :: evr_control_build_field
    end function
    function evr_control_gradient(H, mu, opts, nf, amp, psi0, phi1, J, grad_a, grad_E, grad_V, info) &
        bind(C, name="evr_control_gradient")
      import :: c_int, c_ptr, c_double, evr_control_opts, evr_control_info
      type(c_ptr), value :: H, mu
      type(evr_control_opts), intent(in) :: opts
      integer(c_int), value :: nf
      real(c_double), intent(in) :: amp(*), psi0(*), phi1(*)
      real(c_double), intent(out) :: J(*)
      type(c_ptr), value :: grad_a, grad_E, grad_V
      type(evr_control_info), intent(out) :: info
      integer(c_int) :: evr_control_gradient
    end function
    function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  function evr_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i, n
    p = evr_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function evr_error_message

  subroutine evr_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= EVR_OK) then
      write(*, '(a,a,a,i0,a,a)') 'ERROR in ', where, ' (status ', status, '): ', evr_error_message()
      stop 1
    end if
  end subroutine evr_check

  ! Hpsi(:, k) = H psi(:, k) for real vectors (nb_packed x nvec), in chunks of
  ! the plan's max_nvec: sub_TabOpPsi(Psi, OpPsi, para_Op).
  subroutine evr_sub_TabOpPsi(plan, psi, Hpsi)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in), target, contiguous :: psi(:, :)
    real(c_double), intent(inout), target, contiguous :: Hpsi(:, :)
    integer :: k0, k, mv
    mv = evr_plan_max_nvec(plan)
    do k0 = 1, size(psi, 2), mv
      k = min(mv, size(psi, 2) - k0 + 1)
      call evr_check(evr_oppsi(plan, int(k, c_int), c_loc(psi(1, k0)), c_loc(Hpsi(1, k0)), EVR_FP64), &
                     'evr_sub_TabOpPsi')
    end do
  end subroutine evr_sub_TabOpPsi

  ! Complex vectors: each column is passed as its (Re, Im) column pair.
  subroutine evr_sub_TabOpPsi_cplx(plan, psi, Hpsi)
    type(c_ptr), intent(in) :: plan
    complex(c_double), intent(in) :: psi(:, :)
    complex(c_double), intent(inout) :: Hpsi(:, :)
    real(c_double), allocatable :: r(:, :), hr(:, :)
    integer :: k
    allocate(r(size(psi, 1), 2 * size(psi, 2)), hr(size(psi, 1), 2 * size(psi, 2)))
    do k = 1, size(psi, 2)
      r(:, 2 * k - 1) = real(psi(:, k), c_double)
      r(:, 2 * k) = aimag(psi(:, k))
    end do
    call evr_sub_TabOpPsi(plan, r, hr)
    do k = 1, size(psi, 2)
      Hpsi(:, k) = cmplx(hr(:, 2 * k - 1), hr(:, 2 * k), kind=c_double)
    end do
  end subroutine evr_sub_TabOpPsi_cplx

  subroutine evr_BasisRep_TO_GridRep(plan, psiB, psiG)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in) :: psiB(:)
    real(c_double), intent(out) :: psiG(:)
    call evr_check(evr_basis_to_grid(plan, psiB, psiG), 'evr_BasisRep_TO_GridRep')
  end subroutine evr_BasisRep_TO_GridRep

  subroutine evr_GridRep_TO_BasisRep(plan, psiG, psiB)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in) :: psiG(:)
    real(c_double), intent(out) :: psiB(:)
    call evr_check(evr_grid_to_basis(plan, psiG, psiB), 'evr_GridRep_TO_BasisRep')
  end subroutine evr_GridRep_TO_BasisRep

  ! BasisTOGrid / GridTOBasis of the columns of a table of psi (host arrays):
  ! psiB(nb_packed, ncol) <-> psiG(sum_nq, ncol).
  subroutine evr_BasisRep_TO_GridRep_n(plan, psiB, psiG)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in), target, contiguous :: psiB(:, :)
    real(c_double), intent(out), target, contiguous :: psiG(:, :)
    call evr_check(evr_basis_to_grid_n(plan, int(size(psiB, 2), c_int), c_loc(psiB), c_loc(psiG), 0_c_int), &
                   'evr_BasisRep_TO_GridRep_n')
  end subroutine evr_BasisRep_TO_GridRep_n

  subroutine evr_GridRep_TO_BasisRep_n(plan, psiG, psiB)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in), target, contiguous :: psiG(:, :)
    real(c_double), intent(out), target, contiguous :: psiB(:, :)
    call evr_check(evr_grid_to_basis_n(plan, int(size(psiG, 2), c_int), c_loc(psiG), c_loc(psiB), 0_c_int), &
                   'evr_GridRep_TO_BasisRep_n')
  end subroutine evr_GridRep_TO_BasisRep_n

  ! DerivOp_TO_RVecG(RVecG, nq, basis, tab_der) on the columns of psiG
  ! (tab_der: 1-based modes, 0 for none).
  subroutine evr_DerivOp_TO_RVecG(plan, psiG, dpsiG, tab_der)
    type(c_ptr), intent(in) :: plan
    real(c_double), intent(in), target, contiguous :: psiG(:, :)
    real(c_double), intent(out), target, contiguous :: dpsiG(:, :)
    integer, intent(in) :: tab_der(2)
    integer(c_int) :: td(2)
    td = int(tab_der, c_int)
    call evr_check(evr_grid_deriv(plan, int(size(psiG, 2), c_int), td, c_loc(psiG), c_loc(dpsiG), 0_c_int), &
                   'evr_DerivOp_TO_RVecG')
  end subroutine evr_DerivOp_TO_RVecG

  ! nev lowest levels (and vectors, nb_packed x nev) by block Davidson.
  subroutine evr_propagation_Davidson(plan, nev, block, conv_resi, conv_ene, Ene, Vec, info)
    type(c_ptr), intent(in) :: plan
    integer, intent(in) :: nev, block
    real(c_double), intent(in) :: conv_resi, conv_ene
    real(c_double), intent(out) :: Ene(:)
    real(c_double), intent(out), target, contiguous :: Vec(:, :)
    type(evr_davidson_info), intent(out) :: info
    type(evr_davidson_opts) :: o
    o%nev = int(nev, c_int)
    o%block = int(block, c_int)
    o%conv_resi = conv_resi
    o%conv_ene = conv_ene
    call evr_check(evr_davidson(plan, o, Ene, c_loc(Vec), info), 'evr_propagation_Davidson')
  end subroutine evr_propagation_Davidson

  ! nsteps Chebyshev steps of dt for the complex wavepackets psi(:, 1:nwp).
  subroutine evr_march_cheby(plan, psi, dt, nsteps, npoly, poly_tol, info)
    type(c_ptr), intent(in) :: plan
    complex(c_double), intent(inout) :: psi(:, :)
    real(c_double), intent(in) :: dt, poly_tol
    integer, intent(in) :: nsteps, npoly
    type(evr_propa_info), intent(out) :: info
    type(evr_propa_opts) :: o
    real(c_double), allocatable, target :: r(:, :)
    integer :: k
    allocate(r(size(psi, 1), 2 * size(psi, 2)))
    do k = 1, size(psi, 2)
      r(:, 2 * k - 1) = real(psi(:, k), c_double)
      r(:, 2 * k) = aimag(psi(:, k))
    end do
    o%method = EVR_PROPA_CHEBY
    o%dt = dt
    o%nsteps = int(nsteps, c_int)
    o%npoly = int(npoly, c_int)
    o%poly_tol = poly_tol
    call evr_check(evr_propagate(plan, o, int(size(psi, 2), c_int), c_loc(r), c_null_ptr, info), &
                   'evr_march_cheby')
    do k = 1, size(psi, 2)
      psi(:, k) = cmplx(r(:, 2 * k - 1), r(:, 2 * k), kind=c_double)
    end do
  end subroutine evr_march_cheby

end module evr_hip_f
