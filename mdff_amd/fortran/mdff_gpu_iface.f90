! mdff_gpu_iface.f90 -- iso_c_binding interface to libmdff_gpu (include/mdff_gpu.h).
!
! The derived types mirror the C structs member by member (bind(C) gives them
! the C layout); the interfaces bind the extern "C" entry points by name, so
! Fortran host code can call the gfx950 engforce library directly.
!
! Replaces, at the call sites of the reference, MODULE engforce_driver
! (src/engforce_driver.f90:46-190) and the setup of the force-field modules
! (src/non_bonded.f90:455-644, src/coulomb.f90:363-493, src/pim.f90:246-315).
! Errors come back as status codes; mdff_check prints the message the
! reference prints for the same condition and STOPs, as the reference does
! (src/tools.f90:337-340, src/moment_from_pola.f90:176-216).
module mdff_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: dp = c_double
  integer(c_int32_t), parameter, public :: MDFF_NTYPEMAX = 16
  integer(c_int32_t), parameter, public :: MDFF_ABI_VERSION = 3

  integer(c_int), parameter, public :: MDFF_OK = 0, MDFF_ERR_BAD_ARG = 1, MDFF_ERR_LIST_OVERFLOW = 2, &
       MDFF_ERR_SCF_NOT_CONVERGED = 3, MDFF_ERR_POLAR_CATASTROPHE = 4, MDFF_ERR_HIP = 5, &
       MDFF_ERR_NO_DEVICE = 6, MDFF_ERR_UNSUPPORTED = 7
  integer(c_int32_t), parameter, public :: MDFF_PREC_FP64 = 0, MDFF_PREC_MIXED = 1
  integer(c_int32_t), parameter, public :: MDFF_TRUNC_NONE = 0, MDFF_TRUNC_LINEAR = 1, &
       MDFF_TRUNC_QUADRATIC = 2

  ! struct mdff_desc
  type, bind(C), public :: mdff_desc
    integer(c_int32_t) :: abi_version, natm, ntype, precision
    real(c_double)     :: cell(9)
    integer(c_int32_t) :: lnmlj, lbmhft, lbmhftd, lmorse, lcoulomb, lsurf, lvnlist
    integer(c_int32_t) :: calc_md
    real(c_double)     :: cutshortrange, cutlongrange, skindiff
    real(c_double)     :: coul_unit, press_unit
    integer(c_int32_t) :: trunc, pad0
    real(c_double)     :: epslj(256), sigmalj(256), plj(256), qlj(256)
    real(c_double)     :: epsmor(256), sigmamor(256), rhomor(256)
    real(c_double)     :: Abmhftd(256), Bbmhftd(256), Cbmhftd(256), Dbmhftd(256), BDbmhftd(256)
    real(c_double)     :: alphaES
    integer(c_int32_t) :: kES(3)
    integer(c_int32_t) :: doefield
    real(c_double)     :: mass(16), qch(16), dip(48)
    integer(c_int32_t) :: ldip_polar(16)
    real(c_double)     :: poldip(144)
    integer(c_int32_t) :: ldip_damping(4096)
    real(c_double)     :: pol_damp_b(4096), pol_damp_c(4096)
    integer(c_int32_t) :: pol_damp_k(4096)
    real(c_double)     :: conv_tol_ind
    integer(c_int32_t) :: min_scf_pol_iter, max_scf_pol_iter, extrapolate_order
    integer(c_int32_t) :: algo_moment_from_pola
    real(c_double)     :: omegakO
    integer(c_int32_t) :: doefg, pad1
  end type mdff_desc

  ! struct mdff_terms
  type, bind(C), public :: mdff_terms
    real(c_double)     :: u_lj, u_bmhft, u_morse
    real(c_double)     :: vir_lj, vir_bmhft, vir_morse
    real(c_double)     :: u_coul, u_pol
    real(c_double)     :: u_dir, u_rec, u_self, u_surf
    real(c_double)     :: pvirial_coul
    real(c_double)     :: u_tail, p_tail
    real(c_double)     :: tau(9), tau_nonb(9), tau_coul(9)
    integer(c_int32_t) :: n_scf, list_rebuilt
    real(c_double)     :: rmsd
  end type mdff_terms

  public :: mdff_desc_defaults, mdff_ctx_create, mdff_ctx_destroy, mdff_last_error_c, mdff_version_c
  public :: mdff_set_positions, mdff_get_positions, mdff_set_velocities, mdff_get_velocities
  public :: mdff_get_forces, mdff_set_cell, mdff_vnlist_build, mdff_vnlistcheck, mdff_get_pairs
  public :: mdff_engforce, mdff_get_dipoles, mdff_get_efg, mdff_set_itime, mdff_md_vv, mdff_md_nhcn, mdff_last_terms
  public :: mdff_csvr_rescale, mdff_calc_temp, mdff_last_timings, mdff_rescale_velocities
  public :: mdff_check, mdff_error_string, mdff_version_string

  interface
    subroutine mdff_desc_defaults(d) bind(C, name="mdff_desc_defaults")
      import :: mdff_desc
      type(mdff_desc), intent(inout) :: d
    end subroutine

    integer(c_int) function mdff_ctx_create(d, itype, ctx) bind(C, name="mdff_ctx_create")
      import :: mdff_desc, c_int, c_int32_t, c_ptr
      type(mdff_desc), intent(in) :: d
      integer(c_int32_t), intent(in) :: itype(*)
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function mdff_ctx_destroy(ctx) bind(C, name="mdff_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function mdff_last_error_c(ctx) bind(C, name="mdff_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function mdff_version_c() bind(C, name="mdff_version")
      import :: c_ptr
    end function

    integer(c_int) function mdff_set_positions(ctx, rx, ry, rz) bind(C, name="mdff_set_positions")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rx(*), ry(*), rz(*)
    end function

    integer(c_int) function mdff_get_positions(ctx, rx, ry, rz) bind(C, name="mdff_get_positions")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rx(*), ry(*), rz(*)
    end function

    integer(c_int) function mdff_set_velocities(ctx, vx, vy, vz) bind(C, name="mdff_set_velocities")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vx(*), vy(*), vz(*)
    end function

    integer(c_int) function mdff_get_velocities(ctx, vx, vy, vz) bind(C, name="mdff_get_velocities")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: vx(*), vy(*), vz(*)
    end function

    integer(c_int) function mdff_get_forces(ctx, fx, fy, fz) bind(C, name="mdff_get_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: fx(*), fy(*), fz(*)
    end function

    integer(c_int) function mdff_set_cell(ctx, cell) bind(C, name="mdff_set_cell")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: cell(9)
    end function

    integer(c_int) function mdff_vnlist_build(ctx) bind(C, name="mdff_vnlist_build")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function mdff_vnlistcheck(ctx, rebuilt) bind(C, name="mdff_vnlistcheck")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: rebuilt
    end function

    integer(c_int) function mdff_get_pairs(ctx, which, point, list, cap, n) bind(C, name="mdff_get_pairs")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      integer(c_int32_t), intent(out) :: point(*), list(*)
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: n
    end function

    integer(c_int) function mdff_engforce(ctx, fx, fy, fz, terms) bind(C, name="mdff_engforce")
      import :: c_int, c_ptr, c_double, mdff_terms
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: fx(*), fy(*), fz(*)
      type(mdff_terms), intent(out) :: terms
    end function

    integer(c_int) function mdff_last_terms(ctx, terms) bind(C, name="mdff_last_terms")
      import :: c_int, c_ptr, mdff_terms
      type(c_ptr), value :: ctx
      type(mdff_terms), intent(out) :: terms
    end function

    integer(c_int) function mdff_get_efg(ctx, efg) bind(C, name="mdff_get_efg")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: efg(3, 3, *)
    end function

    integer(c_int) function mdff_get_dipoles(ctx, mu, ef) bind(C, name="mdff_get_dipoles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mu(3, *), ef(3, *)
    end function

    integer(c_int) function mdff_set_itime(ctx, itime) bind(C, name="mdff_set_itime")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: itime
    end function

    integer(c_int) function mdff_md_vv(ctx, nsteps, dt, ekin, upot) bind(C, name="mdff_md_vv")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: dt
      real(c_double), intent(out) :: ekin(*), upot(*)
    end function

    integer(c_int) function mdff_md_nhcn(ctx, nsteps, dt, temp, timesca_thermo, nhc_n, yosh_order, &
                                         nhc_mults, xi, vxi, ekin, upot, e_nvt) bind(C, name="mdff_md_nhcn")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, nhc_n, yosh_order, nhc_mults
      real(c_double), value :: dt, temp, timesca_thermo
      real(c_double), intent(inout) :: xi(*), vxi(*)
      real(c_double), intent(out) :: ekin(*), upot(*), e_nvt(*)
    end function

    integer(c_int) function mdff_csvr_rescale(ctx, temp, taut, r1, sumnoise, conint) &
        bind(C, name="mdff_csvr_rescale")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: temp, taut, r1, sumnoise
      real(c_double), intent(inout) :: conint
    end function

    integer(c_int) function mdff_rescale_velocities(ctx, lambda) bind(C, name="mdff_rescale_velocities")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lambda
    end function

    integer(c_int) function mdff_calc_temp(ctx, boltz_unit, temp, ekin) bind(C, name="mdff_calc_temp")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: boltz_unit
      real(c_double), intent(out) :: temp, ekin
    end function

    integer(c_int) function mdff_last_timings(ctx, ms, n) bind(C, name="mdff_last_timings")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(8)
      integer(c_int32_t), intent(out) :: n
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! copy a NUL-terminated C string into a Fortran string
  function c_to_f(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function c_to_f

  function mdff_error_string(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: s
    s = c_to_f(mdff_last_error_c(ctx))
  end function mdff_error_string

  function mdff_version_string() result(s)
    character(len=:), allocatable :: s
    s = c_to_f(mdff_version_c())
  end function mdff_version_string

  ! status check: the reference prints and STOPs on these conditions
  subroutine mdff_check(status, ctx, where)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: where
    if (status == MDFF_OK) return
    if (c_associated(ctx)) then
      write (*, '(a,a,a,i0,a,a)') 'ERROR in ', where, ' (status ', status, ') : ', mdff_error_string(ctx)
    else
      write (*, '(a,a,a,i0,a)') 'ERROR in ', where, ' (status ', status, ')'
    end if
    stop 1
  end subroutine mdff_check

end module mdff_gpu
