! mdff_gpu_driver.f90 -- Fortran MD driver on the gfx950 engforce library.
!
! Usage:  mdff_gpu_driver <control file> [fp64|mixed]
!
! Reads the reference's input files and runs them through the C-ABI:
!   - the namelists &controltag &mdtag &non_bondedtag &coulombtag &fieldtag
!     &pimtag with the reference's variable names and defaults
!     (src/control.f90:128-245, src/md.f90:160-306, src/non_bonded.f90:130-297,
!     src/coulomb.f90:120-238, src/field.f90:113-188, src/pim.f90:63-240);
!     every variable of those groups is accepted, the ones outside the engforce
!     path are read and ignored;
!   - the configuration file POSFF of the working directory
!     (src/read_pos.f90:62-108; posff_data rnn / rvn / rvf, Direct or Cartesian).
! Runs one engforce at t = 0 (src/runmd.f90:175-202), then npas steps of
! nve-vv or nvt-nhcn with the skin check after each step and the CSVR / NVE
! equilibration rescale of src/runmd.f90:294-330, printing the energies of the
! first force call and a thermo line every nprint steps.  Writes CONTFF
! (src/config.f90:208-252) and FORCES (natm lines fx fy fz of the t = 0 call).
program mdff_gpu_driver
  use, intrinsic :: iso_c_binding
  use mdff_gpu
  implicit none

  integer, parameter :: ntypemax = 16
  ! constants (src/constants.f90:88-93)
  real(dp) :: press_unit = 0.00624150964712042_dp
  real(dp) :: boltz_unit = 8.61734229648141e-05_dp
  real(dp) :: time_unit = 98.2269514139276_dp
  real(dp) :: coul_unit = 14.3996441494161_dp
  real(dp), parameter :: pi = 3.14159265358979323846_dp

  ! ---- &controltag
  logical :: lnmlj, lcoulomb, lmorse, lbmhft, lbmhftd, lsurf, lcsvr, lharm, ltraj, lvnlist, lstatic, &
             lreduced, lreducedN, ltest, lmsd, lvacf, lwrite_restart, full_restart
  real(dp) :: cutlongrange, cutshortrange, skindiff
  character(len=60) :: calc, dgauss, longrange
  integer :: itraj_start, itraj_period, itraj_format, iscff_format, iefall_format, iefgall_format, &
             idipall_format
  character(len=3) :: trajff_data, iscff_data, posff_data
  namelist /controltag/ lnmlj, lcoulomb, lmorse, lbmhft, lbmhftd, lsurf, lcsvr, lharm, ltraj, lvnlist, &
       lstatic, lreduced, lreducedN, ltest, lmsd, lvacf, lwrite_restart, full_restart, cutlongrange, &
       cutshortrange, calc, dgauss, longrange, itraj_start, itraj_period, itraj_format, trajff_data, &
       iscff_format, iscff_data, iefall_format, iefgall_format, idipall_format, posff_data, skindiff

  ! ---- &mdtag
  character(len=60) :: integrator, setvel
  integer :: npas, nequil, nequil_period, npropr, npropr_start, nprint, fprint, spas
  integer :: nhc_yosh_order, nhc_mults, nhc_n
  real(dp) :: annealing, dt, temp, press, nuandersen, taucsvr, tauTberendsen, tauPberendsen
  real(dp) :: timesca_thermo, timesca_baro
  namelist /mdtag/ integrator, setvel, npas, nequil, nequil_period, annealing, npropr, npropr_start, &
       nprint, fprint, spas, dt, temp, press, nuandersen, taucsvr, tauTberendsen, tauPberendsen, &
       nhc_yosh_order, nhc_mults, nhc_n, timesca_thermo, timesca_baro

  ! ---- &non_bondedtag
  logical :: lKA, symmetric_pot
  character(len=60) :: ctrunc
  real(dp), dimension(ntypemax, ntypemax) :: qlj, plj, sigmalj, epslj, sigmamor, epsmor, rhomor, &
       Abmhftd, Bbmhftd, Cbmhftd, Dbmhftd, BDbmhftd
  namelist /non_bondedtag/ lKA, ctrunc, symmetric_pot, qlj, plj, sigmalj, epslj, sigmamor, epsmor, &
       rhomor, Abmhftd, Bbmhftd, Cbmhftd, Dbmhftd, BDbmhftd

  ! ---- &coulombtag
  logical :: lautoES, thole_functions, doefield, doefg
  real(dp) :: epsw, alphaES, thole_param(ntypemax, ntypemax)
  integer :: kES(3)
  character(len=6) :: thole_function_type
  namelist /coulombtag/ lautoES, epsw, alphaES, kES, thole_functions, thole_param, thole_function_type, &
       symmetric_pot, doefield, doefg

  ! ---- &fieldtag
  real(dp) :: mass(ntypemax), qch(ntypemax), quad_nuc(ntypemax), dip(ntypemax, 3), quad(ntypemax, 3, 3)
  namelist /fieldtag/ mass, qch, quad_nuc, dip, quad

  ! ---- &pimtag
  real(dp) :: poldip(ntypemax, 3, 3), poldip_iso(ntypemax), polquad(ntypemax, 3, 3, 3), polquad_iso(ntypemax)
  real(dp) :: pol_damp_b(ntypemax, ntypemax, ntypemax), pol_damp_c(ntypemax, ntypemax, ntypemax)
  integer :: pol_damp_k(ntypemax, ntypemax, ntypemax)
  integer :: extrapolate_order, min_scf_pol_iter, max_scf_pol_iter, lwfc(ntypemax)
  real(dp) :: conv_tol_ind, omegakO, rcut_wfc
  character(len=11) :: algo_moment_from_pola
  character(len=4) :: algo_ext_dipole
  logical :: lwrite_dip_wfc, ldip_wfc, ldip_polar(ntypemax), lquad_polar(ntypemax)
  logical :: ldip_damping(ntypemax, ntypemax, ntypemax)
  namelist /pimtag/ symmetric_pot, poldip, poldip_iso, polquad, polquad_iso, pol_damp_b, pol_damp_c, &
       pol_damp_k, extrapolate_order, conv_tol_ind, min_scf_pol_iter, max_scf_pol_iter, &
       algo_moment_from_pola, algo_ext_dipole, omegakO, lwfc, lwrite_dip_wfc, ldip_wfc, rcut_wfc, &
       ldip_polar, ldip_damping, lquad_polar

  ! ---- configuration
  integer :: natm, ntype, ia, it, it2, natmi(ntypemax)
  character(len=60) :: system_name
  character(len=3) :: atypei(ntypemax)
  character(len=60) :: cpos
  character(len=3), allocatable :: atype(:)
  integer(c_int32_t), allocatable :: itype(:)
  real(dp), allocatable :: rx(:), ry(:), rz(:), vx(:), vy(:), vz(:), fx(:), fy(:), fz(:)
  real(dp) :: A(3, 3), sx, sy, sz

  ! ---- run
  character(len=256) :: filename, precarg
  type(mdff_desc) :: d
  type(mdff_terms) :: t
  type(c_ptr) :: ctx
  integer(c_int) :: st
  integer :: ios, u, itime, nchunk, k
  logical :: sym_nb, sym_pim
  real(dp) :: ekin, tempi, upot, etot, conint, r1, sumnoise, kT
  real(dp), allocatable :: ek(:), up(:), env(:), xi(:), vxi(:)
  ! ---- thermodynamic (src/thermo.f90:37-110, calc_thermo :122-211)
  real(dp) :: B(3, 3), omega_c, anorm(3)
  real(dp) :: e_kin, temp_r, e_nvt, u_tot, e_tot, h_tot, u_vdw, pvirial_vdw, pvirial_coul, pressure_tot
  real(dp) :: pvirial_tot
  integer :: oszcall, itime0, itime1, u_osz, u_traj
  ! ---- RESTART (src/restart.f90:34-478)
  integer, parameter :: vnlmax = 2000
  integer :: itime_rst = 1
  real(dp), allocatable :: xi_rst(:), vxi_rst(:)

  ctx = c_null_ptr
  if (command_argument_count() < 1) then
    write (*, '(a)') 'usage: mdff_gpu_driver <control file> [fp64|mixed]'
    stop 1
  end if
  call get_command_argument(1, filename)
  precarg = 'fp64'
  if (command_argument_count() >= 2) call get_command_argument(2, precarg)

  call default_tags()
  call read_tags()
  if (full_restart) then
    call read_restart()
  else
    call read_posff()
  end if
  call fill_desc()

  st = mdff_ctx_create(d, itype, ctx)
  call mdff_check(st, ctx, 'mdff_ctx_create')
  write (*, '(a,a)') 'engforce library : ', mdff_version_string()
  st = mdff_set_positions(ctx, rx, ry, rz)
  call mdff_check(st, ctx, 'mdff_set_positions')
  if (.not. full_restart) call init_velocities()
  st = mdff_set_velocities(ctx, vx, vy, vz)
  call mdff_check(st, ctx, 'mdff_set_velocities')

  ! ---- t = 0 force call
  st = mdff_engforce(ctx, fx, fy, fz, t)
  if (st == MDFF_ERR_POLAR_CATASTROPHE) call write_dipff()
  call mdff_check(st, ctx, 'engforce')
  call print_terms(t)
  open (newunit=u, file='FORCES', status='replace')
  do ia = 1, natm
    write (u, '(3es25.16e3)') fx(ia), fy(ia), fz(ia)
  end do
  close (u)

  ! ---- MD (md_run, src/runmd.f90:63-504)
  if (trim(calc) == 'md') call md_run()

  st = mdff_ctx_destroy(ctx)

contains

  ! defaults of the reference (src/control.f90:200-245, src/md.f90:180-210,
  ! src/non_bonded.f90:199-224, src/coulomb.f90:156-177, src/field.f90:68-81,
  ! src/pim.f90:111-136)
  subroutine default_tags()
    lnmlj = .false.; lbmhft = .false.; lbmhftd = .false.; lmorse = .false.; lcoulomb = .false.
    lsurf = .false.; lcsvr = .false.; lharm = .false.; ltraj = .false.; lvnlist = .true.
    lstatic = .false.; lreduced = .false.; lreducedN = .false.; ltest = .false.; lmsd = .false.
    lvacf = .false.; lwrite_restart = .false.; full_restart = .false.
    calc = 'md'; dgauss = 'boxmuller_basic'; longrange = 'ewald'
    skindiff = 0.15_dp; cutshortrange = 0.0_dp; cutlongrange = 0.0_dp
    itraj_start = 1; itraj_period = 10000; itraj_format = 1; iscff_format = 1
    iefall_format = 1; iefgall_format = 1; idipall_format = 1
    trajff_data = 'rnn'; iscff_data = 'rnn'; posff_data = 'rnn'

    integrator = 'nve-vv'; setvel = 'MaxwBoltz'; npas = 10; nequil = 0; nequil_period = 1
    nprint = 1; fprint = 1; spas = 1000; dt = 0.0_dp; temp = 1.0_dp; press = 0.0_dp
    nuandersen = 0.0_dp
    taucsvr = 0.0_dp; tauTberendsen = 0.0_dp; tauPberendsen = 0.0_dp
    nhc_yosh_order = 1; nhc_mults = 1; nhc_n = 2; annealing = 1.0_dp; npropr = 1; npropr_start = 0
    timesca_thermo = 1.0_dp; timesca_baro = 1.0_dp

    ctrunc = 'notrunc'; symmetric_pot = .true.; lKA = .false.
    epslj = 0.0_dp; sigmalj = 0.0_dp; qlj = 12.0_dp; plj = 6.0_dp
    epsmor = 0.0_dp; sigmamor = 0.0_dp; rhomor = 0.0_dp
    Abmhftd = 0.0_dp; Bbmhftd = 0.0_dp; Cbmhftd = 0.0_dp; Dbmhftd = 0.0_dp; BDbmhftd = 0.0_dp

    kES = 10; alphaES = 1.0_dp; epsw = 1e-6_dp; lautoES = .false.; doefield = .false.; doefg = .false.
    thole_functions = .false.; thole_function_type = 'linear'; thole_param = 1.662_dp

    mass = 1.0_dp; qch = 0.0_dp; dip = 0.0_dp; quad = 0.0_dp; quad_nuc = 0.0_dp

    ldip_polar = .false.; lquad_polar = .false.; poldip = 0.0_dp; poldip_iso = 0.0_dp
    polquad = 0.0_dp; polquad_iso = 0.0_dp
    pol_damp_b = 0.0_dp; pol_damp_c = 0.0_dp; pol_damp_k = 0; ldip_damping = .false.
    conv_tol_ind = 1e-6_dp; min_scf_pol_iter = 3; max_scf_pol_iter = 100; extrapolate_order = 0
    algo_ext_dipole = 'aspc'; algo_moment_from_pola = 'scf'; omegakO = 0.7_dp
    lwfc = 0; ldip_wfc = .true.; rcut_wfc = 0.5_dp; lwrite_dip_wfc = .false.
  end subroutine default_tags

  ! each group is read from the top of the control file, as each reference
  ! module does; controltag and fieldtag are required, the others optional
  subroutine read_tags()
    integer :: uu
    open (newunit=uu, file=trim(filename), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write (*, '(a,a)') 'ERROR cannot open control file ', trim(filename)
      stop 1
    end if
    read (uu, nml=controltag, iostat=ios)
    call tag_status('controltag', .true.)
    rewind (uu)
    read (uu, nml=mdtag, iostat=ios)
    call tag_status('mdtag', .false.)
    rewind (uu)
    symmetric_pot = .true.
    read (uu, nml=non_bondedtag, iostat=ios)
    call tag_status('non_bondedtag', .false.)
    sym_nb = symmetric_pot
    rewind (uu)
    read (uu, nml=coulombtag, iostat=ios)
    call tag_status('coulombtag', .false.)
    rewind (uu)
    read (uu, nml=fieldtag, iostat=ios)
    call tag_status('fieldtag', .true.)
    rewind (uu)
    symmetric_pot = .true.
    read (uu, nml=pimtag, iostat=ios)
    call tag_status('pimtag', .false.)
    sym_pim = symmetric_pot
    close (uu)
  end subroutine read_tags

  subroutine tag_status(name, required)
    character(len=*), intent(in) :: name
    logical, intent(in) :: required
    if (ios < 0 .and. required) then
      write (*, '(a,a,a)') 'ERROR reading input_file : ', name, ' section is absent'
      stop 1
    else if (ios > 0) then
      write (*, '(a,a,a,i8)') 'ERROR reading input_file : ', name, ' wrong tag', ios
      stop 1
    end if
  end subroutine tag_status

  ! POSFF (src/read_pos.f90:62-108)
  subroutine read_posff()
    integer :: uu, j
    open (newunit=uu, file='POSFF', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write (*, '(a)') 'ERROR cannot open POSFF'
      stop 1
    end if
    read (uu, *) natm
    read (uu, *) system_name
    read (uu, *) A(1, 1), A(2, 1), A(3, 1)
    read (uu, *) A(1, 2), A(2, 2), A(3, 2)
    read (uu, *) A(1, 3), A(2, 3), A(3, 3)
    read (uu, *) ntype
    if (ntype > ntypemax) then
      write (*, '(a)') 'ERROR too many types in POSFF'
      stop 1
    end if
    read (uu, *) (atypei(it), it=1, ntype)
    read (uu, *) (natmi(it), it=1, ntype)
    read (uu, *) cpos
    allocate (atype(natm), itype(natm), rx(natm), ry(natm), rz(natm), vx(natm), vy(natm), vz(natm), &
              fx(natm), fy(natm), fz(natm))
    vx = 0.0_dp; vy = 0.0_dp; vz = 0.0_dp
    if (posff_data == 'rnn') then
      read (uu, *) (atype(ia), rx(ia), ry(ia), rz(ia), ia=1, natm)
    else if (posff_data == 'rvn') then
      read (uu, *) (atype(ia), rx(ia), ry(ia), rz(ia), vx(ia), vy(ia), vz(ia), ia=1, natm)
    else if (posff_data == 'rvf') then
      read (uu, *) (atype(ia), rx(ia), ry(ia), rz(ia), vx(ia), vy(ia), vz(ia), fx(ia), fy(ia), fz(ia), &
                    ia=1, natm)
    else
      write (*, '(a,a)') 'ERROR posff_data not allowed: ', posff_data
      stop 1
    end if
    close (uu)
    if (trim(cpos) == 'Direct' .or. trim(cpos) == 'D') then
      ! dirkar: r = A s (src/lattice.f90:249-256)
      do ia = 1, natm
        sx = rx(ia); sy = ry(ia); sz = rz(ia)
        rx(ia) = A(1, 1) * sx + A(1, 2) * sy + A(1, 3) * sz
        ry(ia) = A(2, 1) * sx + A(2, 2) * sy + A(2, 3) * sz
        rz(ia) = A(3, 1) * sx + A(3, 2) * sy + A(3, 3) * sz
      end do
    else if (trim(cpos) /= 'Cartesian' .and. trim(cpos) /= 'C') then
      write (*, '(a,a)') 'ERROR coordinate format not allowed in POSFF: ', trim(cpos)
      stop 1
    end if
    do ia = 1, natm
      itype(ia) = 0
      do j = 1, ntype
        if (atype(ia) == atypei(j)) itype(ia) = j
      end do
      if (itype(ia) == 0) then
        write (*, '(a,i0)') 'ERROR unknown type in POSFF at ion ', ia
        stop 1
      end if
    end do
  end subroutine read_posff

  ! C-ABI descriptor from the namelists, with the reference's input
  ! post-processing: Kob-Andersen set, symmetric tables (src/non_bonded.f90:
  ! 258-290, src/pim.f90:186-196), isotropic polarisabilities (src/pim.f90:
  ! 221-234), automatic Ewald parameters (src/coulomb.f90:363-410)
  subroutine fill_desc()
    real(dp) :: rc_half, eps, tol, tol1, alpha, w(3), an(3), cr(3), omega
    integer :: ka, kb, kc
    call mdff_desc_defaults(d)
    if (lKA) then
      if (ntype /= 2) then
        write (*, '(a)') 'ERROR non_bondedtag lKA should be used with 2 differents types'
        stop 1
      end if
      sigmalj(1, 1) = 1.0_dp; sigmalj(2, 2) = 0.88_dp; sigmalj(1, 2) = 0.8_dp; sigmalj(2, 1) = 0.8_dp
      epslj(1, 1) = 1.0_dp; epslj(2, 2) = 0.5_dp; epslj(1, 2) = 1.5_dp; epslj(2, 1) = 1.5_dp
    end if
    if (sym_nb) then
      do it = 1, ntype
        epslj(:, it) = epslj(it, :); sigmalj(:, it) = sigmalj(it, :)
        plj(:, it) = plj(it, :); qlj(:, it) = qlj(it, :)
        Abmhftd(:, it) = Abmhftd(it, :); Bbmhftd(:, it) = Bbmhftd(it, :)
        Cbmhftd(:, it) = Cbmhftd(it, :); Dbmhftd(:, it) = Dbmhftd(it, :)
        BDbmhftd(:, it) = BDbmhftd(it, :)
      end do
    end if
    if (sym_pim) then
      do it = 1, ntype
        do it2 = 1, ntype
          ldip_damping(it2, :, it) = ldip_damping(it2, it, :)
          pol_damp_b(it2, :, it) = pol_damp_b(it2, it, :)
          pol_damp_c(it2, :, it) = pol_damp_c(it2, it, :)
          pol_damp_k(it2, :, it) = pol_damp_k(it2, it, :)
        end do
      end do
    end if
    do it = 1, ntype
      if (poldip_iso(it) /= 0.0_dp) then
        poldip(it, :, :) = 0.0_dp
        do ka = 1, 3
          poldip(it, ka, ka) = poldip_iso(it)
        end do
      end if
    end do
    if (any(lquad_polar) .or. any(quad /= 0.0_dp) .or. thole_functions .or. lwfc(1) /= 0 .or. &
        (trim(algo_moment_from_pola) /= 'scf' .and. trim(algo_moment_from_pola) /= 'scf_kO_v2' .and. &
         trim(algo_moment_from_pola) /= 'scf_kO_v1') .or. &
        (trim(algo_moment_from_pola) /= 'scf' .and. trim(algo_ext_dipole) /= 'aspc') .or. &
        lmorse .and. (lnmlj .or. lbmhft .or. lbmhftd)) then
      write (*, '(a)') 'ERROR option outside the engforce path of this library ' // &
                       '(quadrupoles, Thole, Wannier centres, SCF variants other than scf, scf_kO_v1, scf_kO_v2)'
      stop 1
    end if
    if (lreduced) then
      coul_unit = 1.0_dp; press_unit = 1.0_dp; boltz_unit = 1.0_dp; time_unit = 1.0_dp
    end if
    ! widths of the cell and the automatic Ewald parameters (dl_poly form)
    omega = A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) - A(1, 2) * (A(2, 1) * A(3, 3) - &
            A(3, 1) * A(2, 3)) + A(1, 3) * (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2))
    do ka = 1, 3
      kb = mod(ka, 3) + 1
      kc = mod(ka + 1, 3) + 1
      cr = [A(2, kb) * A(3, kc) - A(3, kb) * A(2, kc), A(3, kb) * A(1, kc) - A(1, kb) * A(3, kc), &
            A(1, kb) * A(2, kc) - A(2, kb) * A(1, kc)]
      w(ka) = abs(omega) / sqrt(sum(cr * cr))
      an(ka) = sqrt(sum(A(:, ka) * A(:, ka)))
    end do
    if (lcoulomb .and. lautoES) then
      rc_half = 0.5_dp * minval(w)
      cutlongrange = min(cutlongrange, rc_half)
      eps = min(abs(epsw), 0.5_dp)
      tol = sqrt(abs(log(eps * cutlongrange)))
      alpha = sqrt(abs(log(eps * cutlongrange * tol))) / cutlongrange
      tol1 = sqrt(-log(eps * cutlongrange * (2.0_dp * tol * alpha)**2))
      alphaES = alpha
      do ka = 1, 3
        kES(ka) = nint(0.25_dp + an(ka) * alpha * tol1 / pi)
      end do
    end if

    d%abi_version = MDFF_ABI_VERSION
    d%natm = natm
    d%ntype = ntype
    d%precision = MDFF_PREC_FP64
    if (trim(precarg) == 'mixed') d%precision = MDFF_PREC_MIXED
    d%cell = reshape(A, [9])
    d%lnmlj = merge(1, 0, lnmlj); d%lbmhft = merge(1, 0, lbmhft); d%lbmhftd = merge(1, 0, lbmhftd)
    d%lmorse = merge(1, 0, lmorse); d%lcoulomb = merge(1, 0, lcoulomb); d%lsurf = merge(1, 0, lsurf)
    d%lvnlist = merge(1, 0, lvnlist); d%calc_md = merge(1, 0, trim(calc) == 'md')
    d%cutshortrange = cutshortrange; d%cutlongrange = cutlongrange; d%skindiff = skindiff
    d%coul_unit = coul_unit; d%press_unit = press_unit
    d%trunc = MDFF_TRUNC_NONE
    if (trim(ctrunc) == 'linear') d%trunc = MDFF_TRUNC_LINEAR
    if (trim(ctrunc) == 'quadratic') d%trunc = MDFF_TRUNC_QUADRATIC
    d%epslj = reshape(epslj, [256]); d%sigmalj = reshape(sigmalj, [256])
    d%plj = reshape(plj, [256]); d%qlj = reshape(qlj, [256])
    d%epsmor = reshape(epsmor, [256]); d%sigmamor = reshape(sigmamor, [256]); d%rhomor = reshape(rhomor, [256])
    d%Abmhftd = reshape(Abmhftd, [256]); d%Bbmhftd = reshape(Bbmhftd, [256])
    d%Cbmhftd = reshape(Cbmhftd, [256]); d%Dbmhftd = reshape(Dbmhftd, [256])
    d%BDbmhftd = reshape(BDbmhftd, [256])
    d%alphaES = alphaES
    d%kES = kES
    d%doefield = merge(1, 0, doefield)
    d%doefg = merge(1, 0, doefg)
    d%mass = mass; d%qch = qch; d%dip = reshape(dip, [48])
    d%ldip_polar = merge(1, 0, ldip_polar)
    d%poldip = reshape(poldip, [144])
    d%ldip_damping = reshape(merge(1, 0, ldip_damping), [4096])
    d%pol_damp_b = reshape(pol_damp_b, [4096]); d%pol_damp_c = reshape(pol_damp_c, [4096])
    d%pol_damp_k = reshape(pol_damp_k, [4096])
    d%conv_tol_ind = conv_tol_ind
    d%min_scf_pol_iter = min_scf_pol_iter; d%max_scf_pol_iter = max_scf_pol_iter
    d%extrapolate_order = extrapolate_order
    d%algo_moment_from_pola = 0
    if (trim(algo_moment_from_pola) == 'scf_kO_v2') d%algo_moment_from_pola = 1
    if (trim(algo_moment_from_pola) == 'scf_kO_v1') d%algo_moment_from_pola = 2
    d%omegakO = omegakO
  end subroutine fill_desc

  ! velocities: from POSFF (rvn, rvf), else Maxwell-Boltzmann at temp with the
  ! centre-of-mass velocity removed and the temperature set exactly (the
  ! compiler's RANDOM_NUMBER stream, as in src/rand.f90)
  subroutine init_velocities()
    real(dp) :: g, m, vcm(3), mt, e, tcur
    integer :: nseed
    integer, allocatable :: seed(:)
    if (posff_data /= 'rnn' .or. temp <= 0.0_dp .or. trim(setvel) /= 'MaxwBoltz') return
    call random_seed(size=nseed)
    allocate (seed(nseed))
    seed = 20160101 + [(k, k=1, nseed)]
    call random_seed(put=seed)
    do ia = 1, natm
      m = mass(itype(ia))
      call gauss(g); vx(ia) = g * sqrt(temp * boltz_unit / m)
      call gauss(g); vy(ia) = g * sqrt(temp * boltz_unit / m)
      call gauss(g); vz(ia) = g * sqrt(temp * boltz_unit / m)
    end do
    vcm = 0.0_dp; mt = 0.0_dp
    do ia = 1, natm
      m = mass(itype(ia))
      vcm = vcm + m * [vx(ia), vy(ia), vz(ia)]
      mt = mt + m
    end do
    vcm = vcm / mt
    vx = vx - vcm(1); vy = vy - vcm(2); vz = vz - vcm(3)
    e = 0.0_dp
    do ia = 1, natm
      e = e + 0.5_dp * mass(itype(ia)) * (vx(ia)**2 + vy(ia)**2 + vz(ia)**2)
    end do
    tcur = 2.0_dp * e / (3.0_dp * natm * boltz_unit)
    if (tcur > 0.0_dp) then
      vx = vx * sqrt(temp / tcur); vy = vy * sqrt(temp / tcur); vz = vz * sqrt(temp / tcur)
    end if
  end subroutine init_velocities

  ! N(0,1) deviate (Box-Muller)
  subroutine gauss(g)
    real(dp), intent(out) :: g
    real(dp) :: u1, u2
    call random_number(u1)
    call random_number(u2)
    u1 = max(u1, 1e-300_dp)
    g = sqrt(-2.0_dp * log(u1)) * cos(2.0_dp * pi * u2)
  end subroutine gauss

  ! sum of n squared N(0,1) deviates (the sumnoises of src/kinetic.f90:590-635)
  function sum_noises(n) result(s)
    integer, intent(in) :: n
    real(dp) :: s, g
    integer :: j
    s = 0.0_dp
    do j = 1, n
      call gauss(g)
      s = s + g * g
    end do
  end function sum_noises

  subroutine print_terms(tt)
    type(mdff_terms), intent(in) :: tt
    write (*, '(a,es25.16e3)') 'u_lj         = ', tt%u_lj
    write (*, '(a,es25.16e3)') 'u_bmhft      = ', tt%u_bmhft
    write (*, '(a,es25.16e3)') 'u_morse      = ', tt%u_morse
    write (*, '(a,es25.16e3)') 'u_coul       = ', tt%u_coul
    write (*, '(a,es25.16e3)') 'u_pol        = ', tt%u_pol
    write (*, '(a,es25.16e3)') 'u_dir        = ', tt%u_dir
    write (*, '(a,es25.16e3)') 'u_rec        = ', tt%u_rec
    write (*, '(a,es25.16e3)') 'u_self       = ', tt%u_self
    write (*, '(a,es25.16e3)') 'u_surf       = ', tt%u_surf
    write (*, '(a,es25.16e3)') 'vir_bmhft    = ', tt%vir_bmhft
    write (*, '(a,es25.16e3)') 'vir_lj       = ', tt%vir_lj
    write (*, '(a,es25.16e3)') 'pvirial_coul = ', tt%pvirial_coul
    write (*, '(a,9es25.16e3)') 'tau_nonb     = ', tt%tau_nonb
    write (*, '(a,9es25.16e3)') 'tau_coul     = ', tt%tau_coul
    write (*, '(a,i0)') 'n_scf        = ', tt%n_scf
  end subroutine print_terms

  ! lattice (src/lattice.f90:60-168): B = reciprocal basis / omega, the
  ! cell volume and the lattice vector lengths
  subroutine lattice_b()
    integer :: j
    B(1, 1) = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)
    B(2, 1) = A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3)
    B(3, 1) = A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3)
    B(1, 2) = A(2, 3) * A(3, 1) - A(3, 3) * A(2, 1)
    B(2, 2) = A(3, 3) * A(1, 1) - A(1, 3) * A(3, 1)
    B(3, 2) = A(1, 3) * A(2, 1) - A(2, 3) * A(1, 1)
    B(1, 3) = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2)
    B(2, 3) = A(3, 1) * A(1, 2) - A(1, 1) * A(3, 2)
    B(3, 3) = A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)
    omega_c = B(1, 1) * A(1, 1) + B(2, 1) * A(2, 1) + B(3, 1) * A(3, 1)
    B = B / omega_c
    do j = 1, 3
      anorm(j) = sqrt(sum(A(:, j) * A(:, j)))
    end do
  end subroutine lattice_b

  ! calc_thermo (src/thermo.f90:122-211) from the terms of the last force
  ! evaluation and the kinetic energy of the step
  subroutine calc_thermo(tt)
    type(mdff_terms), intent(in) :: tt
    real(dp) :: pv_lj, pv_morse, pv_bmhft
    temp_r = 2.0_dp * e_kin / (3.0_dp * real(natm, dp) * boltz_unit)
    u_tot = tt%u_lj + tt%u_coul + tt%u_morse + tt%u_bmhft + tt%u_tail
    e_tot = u_tot + e_kin
    u_vdw = tt%u_lj + tt%u_bmhft + tt%u_morse
    pv_lj = tt%vir_lj / omega_c
    pv_morse = tt%vir_morse / omega_c
    pv_bmhft = tt%vir_bmhft / omega_c
    pvirial_coul = tt%pvirial_coul
    pvirial_tot = pv_lj + pvirial_coul + pv_morse + pv_bmhft
    pvirial_vdw = pv_lj + pv_morse + pv_bmhft
    pressure_tot = (pvirial_tot + real(natm, dp) * temp_r * boltz_unit / omega_c) / press_unit
    if (trim(integrator) == 'nvt-nhcn') then
      h_tot = e_tot + e_nvt
    else
      h_tot = e_tot
      if (lcsvr) h_tot = e_tot + conint
    end if
  end subroutine calc_thermo

  ! write_thermo (src/thermo.f90:340-424): 'osz' lines to OSZIFF (a header
  ! every 10 calls), 'std' block to the standard output
  subroutine write_thermo(step, key)
    integer, intent(in) :: step
    character(len=3), intent(in) :: key
    real(dp) :: time, density
    time = real(step, dp) * dt
    density = real(natm, dp) / omega_c
    if (key == 'osz') then
      if (mod(oszcall, 10) == 0) then
        write (u_osz, '(a)') repeat('-', 110)
        write (u_osz, '(a)') '     step      Time/Temp   Etot/Pressure    Ekin/Pvir_nb  Utot/Pvir_coul' // &
                             '      U_nb/Omega     U_coul/Htot'
        write (u_osz, '(a)') repeat('-', 110)
      end if
      write (u_osz, '(I9,2X,6E16.8)') step, time, e_tot, e_kin, u_tot, u_vdw, t%u_coul
      write (u_osz, '(I9,2X,6E16.8)') step, temp_r, pressure_tot, pvirial_vdw, pvirial_coul, omega_c, h_tot
      oszcall = oszcall + 1
    else
      write (*, 300) step, time, e_kin, temp_r, u_tot, u_vdw, t%u_coul, pressure_tot, pvirial_vdw, &
                     pvirial_coul, pvirial_tot, omega_c, anorm(1), anorm(2), anorm(3), density, e_tot, h_tot
    end if
300 format(/ '  Thermodynamic information ' / '  ---------------------------------------------' / &
           '  step                  = ', I9 / '  time                  = ', E19.12 / &
           '  Ekin                  = ', E19.12 / '  Temp                  = ', E19.12 / &
           '  Utot                  = ', E19.12 / '  U_vdw                 = ', E19.12 / &
           '  U_coul                = ', E19.12 / '  Pressure              = ', E19.12 / &
           '  Pvir_vdw              = ', E19.12 / '  Pvir_coul             = ', E19.12 / &
           '  Pvir_tot              = ', E19.12 / '  volume                = ', E19.12 / &
           '  a cell                = ', E19.12 / '  b cell                = ', E19.12 / &
           '  c cell                = ', E19.12 / '  density               = ', E19.12 / &
           '  ---------------------------------------------' / '  Etot                  = ', E19.12 / &
           '  Htot                  = ', E19.12)
  end subroutine write_thermo

  ! write_trajff_xyz (src/config.f90:543-630): positions folded into the
  ! cell (kardir, periodicbc with nint, dirkar), formatted or unformatted
  subroutine write_trajff()
    real(dp), allocatable :: xx(:), yy(:), zz(:)
    real(dp) :: v1, v2, v3
    integer :: i, j
    st = mdff_get_positions(ctx, rx, ry, rz)
    call mdff_check(st, ctx, 'get_positions')
    st = mdff_get_velocities(ctx, vx, vy, vz)
    call mdff_check(st, ctx, 'get_velocities')
    st = mdff_get_forces(ctx, fx, fy, fz)
    call mdff_check(st, ctx, 'get_forces')
    allocate (xx(natm), yy(natm), zz(natm))
    do ia = 1, natm
      v1 = rx(ia) * B(1, 1) + ry(ia) * B(2, 1) + rz(ia) * B(3, 1)
      v2 = rx(ia) * B(1, 2) + ry(ia) * B(2, 2) + rz(ia) * B(3, 2)
      v3 = rx(ia) * B(1, 3) + ry(ia) * B(2, 3) + rz(ia) * B(3, 3)
      v1 = v1 - nint(v1); v2 = v2 - nint(v2); v3 = v3 - nint(v3)
      xx(ia) = v1 * A(1, 1) + v2 * A(1, 2) + v3 * A(1, 3)
      yy(ia) = v1 * A(2, 1) + v2 * A(2, 2) + v3 * A(2, 3)
      zz(ia) = v1 * A(3, 1) + v2 * A(3, 2) + v3 * A(3, 3)
    end do
    if (itraj_format /= 0) then
      write (u_traj, '(i6)') natm
      write (u_traj, '(a)') system_name
      do i = 1, 3
        write (u_traj, '(3f20.12)') (A(i, j), j=1, 3)
      end do
      write (u_traj, '(i4)') ntype
      write (u_traj, *) (atypei(it), it=1, ntype)
      write (u_traj, *) (natmi(it), it=1, ntype)
      write (u_traj, '(A)') 'Cartesian'
      do ia = 1, natm
        select case (trajff_data)
        case ('rvf')
          write (u_traj, '(a,9e20.12)') atype(ia), xx(ia), yy(ia), zz(ia), vx(ia), vy(ia), vz(ia), &
                                        fx(ia), fy(ia), fz(ia)
        case ('rvn')
          write (u_traj, '(a,9e20.12)') atype(ia), xx(ia), yy(ia), zz(ia), vx(ia), vy(ia), vz(ia)
        case ('rnf')
          write (u_traj, '(a,9e20.12)') atype(ia), xx(ia), yy(ia), zz(ia), fx(ia), fy(ia), fz(ia)
        case ('rnn')
          write (u_traj, '(a,9e20.12)') atype(ia), xx(ia), yy(ia), zz(ia)
        end select
      end do
    else
      write (u_traj) natm
      write (u_traj) system_name
      do i = 1, 3
        write (u_traj) (A(i, j), j=1, 3)
      end do
      write (u_traj) ntype
      write (u_traj) (atypei(it), it=1, ntype)
      write (u_traj) (natmi(it), it=1, ntype)
      write (u_traj) 'C'
      select case (trajff_data)
      case ('rvf')
        write (u_traj) xx, yy, zz, vx, vy, vz, fx, fy, fz
      case ('rvn')
        write (u_traj) xx, yy, zz, vx, vy, vz
      case ('rnf')
        write (u_traj) xx, yy, zz, fx, fy, fz
      case ('rnn')
        write (u_traj) xx, yy, zz
      end select
    end if
    deallocate (xx, yy, zz)
  end subroutine write_trajff

  ! write_DIPFF (src/field.f90:300-340): the total dipoles mu_t
  subroutine write_dipff()
    integer :: uu
    real(dp), allocatable :: mu(:, :), ef(:, :)
    integer(c_int) :: s2
    allocate (mu(3, natm), ef(3, natm))
    s2 = mdff_get_dipoles(ctx, mu, ef)
    open (newunit=uu, file='DIPFF', status='unknown')
    write (uu, *) natm
    write (uu, *) system_name
    write (uu, *) A(1, 1), A(2, 1), A(3, 1)
    write (uu, *) A(1, 2), A(2, 2), A(3, 2)
    write (uu, *) A(1, 3), A(2, 3), A(3, 3)
    write (uu, *) ntype
    write (uu, *) (atypei(it), it=1, ntype)
    write (uu, *) (natmi(it), it=1, ntype)
    write (uu, '(a)') '      ia type                   mux                  muy                 muz'
    do ia = 1, natm
      write (uu, '(i8,2x,a3,6e24.16)') ia, atype(ia), mu(1, ia), mu(2, ia), mu(3, ia)
    end do
    close (uu)
    deallocate (mu, ef)
  end subroutine write_dipff

  ! write_EFALL (src/field.f90:455-497): the electric field ef_t
  subroutine write_efall()
    integer :: uu
    real(dp), allocatable :: mu(:, :), ef(:, :)
    allocate (mu(3, natm), ef(3, natm))
    st = mdff_get_dipoles(ctx, mu, ef)
    call mdff_check(st, ctx, 'get_dipoles')
    open (newunit=uu, file='EFALL', status='unknown')
    write (uu, *) natm
    write (uu, *) system_name
    write (uu, *) A(1, 1), A(2, 1), A(3, 1)
    write (uu, *) A(1, 2), A(2, 2), A(3, 2)
    write (uu, *) A(1, 3), A(2, 3), A(3, 3)
    write (uu, *) ntype
    write (uu, *) (atypei(it), it=1, ntype)
    write (uu, *) (natmi(it), it=1, ntype)
    write (uu, '(a)') '      ia type                   Ex                      Ey                      Ez'
    do ia = 1, natm
      write (uu, '(i8,2x,a3,3e24.16)') ia, atype(ia), ef(1, ia), ef(2, ia), ef(3, ia)
    end do
    close (uu)
    deallocate (mu, ef)
  end subroutine write_efall

  ! write_EFGALL (src/field.f90:392-448): the field gradient efg_t, formatted
  ! (upper triangle and diagonal) unless iefgall_format = 0 (the whole array)
  subroutine write_efgall()
    integer :: uu
    real(dp), allocatable :: efg(:, :, :)
    allocate (efg(3, 3, natm))
    st = mdff_get_efg(ctx, efg)
    call mdff_check(st, ctx, 'get_efg')
    if (iefgall_format /= 0) then
      open (newunit=uu, file='EFGALL', status='unknown')
      write (uu, *) natm
      write (uu, *) system_name
      write (uu, *) A(1, 1), A(2, 1), A(3, 1)
      write (uu, *) A(1, 2), A(2, 2), A(3, 2)
      write (uu, *) A(1, 3), A(2, 3), A(3, 3)
      write (uu, *) ntype
      write (uu, *) (atypei(it), it=1, ntype)
      write (uu, *) (natmi(it), it=1, ntype)
      write (uu, '(a)') '      ia type                   vxx                     vyy                     vzz' // &
                        '                     vxy                     vxz                     vyz'
      do ia = 1, natm
        write (uu, '(i8,2x,a3,6e24.16)') ia, atype(ia), efg(1, 1, ia), efg(2, 2, ia), efg(3, 3, ia), &
                                         efg(1, 2, ia), efg(1, 3, ia), efg(2, 3, ia)
      end do
    else
      open (newunit=uu, file='EFGALL', status='unknown', form='unformatted')
      write (uu) natm
      write (uu) system_name
      write (uu) A(1, 1), A(2, 1), A(3, 1)
      write (uu) A(1, 2), A(2, 2), A(3, 2)
      write (uu) A(1, 3), A(2, 3), A(3, 3)
      write (uu) ntype
      write (uu) (atypei(it), it=1, ntype)
      write (uu) (natmi(it), it=1, ntype)
      write (uu) efg
    end if
    close (uu)
    deallocate (efg)
  end subroutine write_efgall

  ! the next step after `now` at which the host acts: the equilibration
  ! rescale, a trajectory frame, thermo output, CONTFF, or the last step
  integer function next_event(now) result(nx)
    integer, intent(in) :: now
    nx = now
    do
      nx = nx + 1
      if (nx >= itime1) exit
      if (nx <= nequil .and. mod(nx, max(nequil_period, 1)) == 0) exit
      if (ltraj .and. nx > itraj_start .and. mod(nx, max(itraj_period, 1)) == 0) exit
      if (mod(nx, max(nprint, 1)) == 0 .or. mod(nx, max(fprint, 1)) == 0 .or. mod(nx, max(spas, 1)) == 0) exit
    end do
  end function next_event

  ! md_run (src/runmd.f90:63-504) on the device: steps between host actions
  ! run in one call of the library's MD entry point; at each step the host
  ! acts on, in the reference's order: the step (with its skin check), the
  ! equilibration rescale, the trajectory frame, calc_thermo and the
  ! standard-output / OSZIFF / CONTFF writes
  subroutine md_run()
    integer :: nchunk, nmax
    real(dp) :: lambda
    if (trim(integrator) /= 'nve-vv' .and. trim(integrator) /= 'nvt-nhcn') then
      write (*, '(a,a)') 'ERROR integrator not available on this path: ', trim(integrator)
      stop 1
    end if
    call lattice_b()
    ! md_init (src/md.f90:308-309): itime0 = itime = 1, itime1 = itime0 - 1 +
    ! npas; after a full restart restart_init sets itime0 = itime (the step
    ! after the last one written), itime1 = itime0 + npas
    ! (src/restart.f90:196-197)
    if (full_restart) then
      itime0 = itime_rst
      itime1 = itime0 + npas
    else
      itime0 = 1
      itime1 = itime0 - 1 + npas
    end if
    oszcall = 0
    open (newunit=u_osz, file='OSZIFF', status='unknown')
    if (itraj_format /= 0) open (newunit=u_traj, file='TRAJFF')
    if (itraj_format == 0) open (newunit=u_traj, file='TRAJFF', form='unformatted')
    allocate (xi(nhc_n), vxi(nhc_n))
    xi = 0.0_dp
    vxi = 0.0_dp
    if (full_restart .and. allocated(xi_rst)) then
      if (size(xi_rst) == nhc_n) then
        xi = xi_rst
        vxi = vxi_rst
      end if
    end if
    conint = 0.0_dp
    e_nvt = 0.0_dp
    kT = temp * boltz_unit
    ! thermo of the configuration at t = 0 (src/runmd.f90:160-183)
    st = mdff_calc_temp(ctx, boltz_unit, tempi, e_kin)
    call mdff_check(st, ctx, 'calc_temp')
    call calc_thermo(t)
    call write_thermo(itime0 - 1, 'std')
    call write_thermo(itime0 - 1, 'osz')
    nmax = max(1, itime1 - itime0 + 1)
    allocate (ek(nmax), up(nmax), env(nmax))
    env = 0.0_dp
    itime = itime0 - 1
    do while (itime < itime1)
      nchunk = next_event(itime) - itime
      if (trim(integrator) == 'nve-vv') then
        st = mdff_md_vv(ctx, int(nchunk, c_int32_t), dt * time_unit, ek, up)
      else
        st = mdff_md_nhcn(ctx, int(nchunk, c_int32_t), dt * time_unit, kT, timesca_thermo * time_unit, &
                          int(nhc_n, c_int32_t), int(nhc_yosh_order, c_int32_t), &
                          int(nhc_mults, c_int32_t), xi, vxi, ek, up, env)
      end if
      if (st == MDFF_ERR_POLAR_CATASTROPHE) call write_dipff()
      call mdff_check(st, ctx, 'md step')
      itime = itime + nchunk
      e_kin = ek(nchunk)
      e_nvt = env(nchunk)
      st = mdff_last_terms(ctx, t)
      call mdff_check(st, ctx, 'last_terms')
      ! equilibration rescale of the NVE integrators (src/runmd.f90:316-330,
      ! rescale_velocities src/kinetic.f90:122-198)
      if (trim(integrator) == 'nve-vv' .and. itime <= nequil .and. mod(itime, max(nequil_period, 1)) == 0 .and. &
          itime /= itime0 .and. itime /= itime1) then
        if (lcsvr) then
          call gauss(r1)
          sumnoise = sum_noises(3 * natm - 4)
          ! the deviates of this rescale, so that a run can be replayed
          write (*, '(a,i10,2es26.17)') ' csvr noise', itime, r1, sumnoise
          st = mdff_csvr_rescale(ctx, kT, taucsvr * time_unit, r1, sumnoise, conint)
          call mdff_check(st, ctx, 'csvr rescale')
        else
          st = mdff_calc_temp(ctx, boltz_unit, tempi, ekin)
          call mdff_check(st, ctx, 'calc_temp')
          lambda = (1.0_dp + (dt / tauTberendsen) * ((kT / tempi / boltz_unit) - 1.0_dp))**0.5_dp
          if (annealing /= 1.0_dp) lambda = annealing
          st = mdff_rescale_velocities(ctx, lambda)
          call mdff_check(st, ctx, 'rescale velocities')
        end if
      end if
      if (ltraj .and. itime > itraj_start .and. mod(itime, max(itraj_period, 1)) == 0) call write_trajff()
      call calc_thermo(t)
      if (mod(itime, max(nprint, 1)) == 0) call write_thermo(itime, 'std')
      if (mod(itime, max(fprint, 1)) == 0) call write_thermo(itime, 'osz')
      if (mod(itime, max(spas, 1)) == 0) call write_contff()
    end do
    close (u_osz)
    close (u_traj)
    ! final configuration (src/mdff.f90:337-345); the loop index of the
    ! reference ends one past itime1, the step a restart starts from
    call write_contff()
    if (doefield) call write_efall()
    if (doefg) call write_efgall()
    if (lwrite_restart) call write_restart(itime1 + 1)
  end subroutine md_run

  ! write_RESTART (src/restart.f90:295-478): the reference's unformatted
  ! record sequence, with the reference's types and sizes (character
  ! lengths, ntypemax-sized tables, natm*vnlmax list arrays); md times and
  ! temperatures in internal units as the reference holds them.  Quantities
  ! outside this path (leap-frog / Beeman / barostat state, quadrupoles,
  ! Wannier flags) carry their reference defaults.
  subroutine write_restart(itime_next)
    integer, intent(in) :: itime_next
    integer :: uu, j
    character(len=3) :: atypei0(0:ntypemax)
    integer :: natmi0(0:ntypemax)
    integer, allocatable :: lst(:), pnt(:)
    integer(c_int64_t) :: nl
    real(dp), allocatable :: zero1(:), sfx(:), sfy(:), sfz(:), dipia(:, :), q0(:), m0(:), pold(:, :, :), &
                             ipold(:, :, :), mu_t(:, :), ef_t(:, :), quad3(:, :, :), quad4(:, :, :, :)
    logical, allocatable :: ipolar(:)
    logical :: task_coul(6), lfalse
    real(dp) :: vxib(max(nhc_n, 1)), xib(max(nhc_n, 1)), xiw(max(nhc_n, 1)), vxiw(max(nhc_n, 1)), det
    integer :: nth
    st = mdff_get_positions(ctx, rx, ry, rz)
    call mdff_check(st, ctx, 'get_positions')
    st = mdff_get_velocities(ctx, vx, vy, vz)
    call mdff_check(st, ctx, 'get_velocities')
    st = mdff_get_forces(ctx, fx, fy, fz)
    call mdff_check(st, ctx, 'get_forces')
    atypei0 = '   '
    atypei0(1:ntype) = atypei(1:ntype)
    atypei0(0) = 'ALL'
    natmi0 = 0
    natmi0(1:ntype) = natmi(1:ntype)
    natmi0(0) = natm
    allocate (zero1(natm), sfx(natm), sfy(natm), sfz(natm), dipia(3, natm), q0(natm), m0(natm), &
              pold(3, 3, natm), ipold(3, 3, natm), mu_t(3, natm), ef_t(3, natm), quad3(3, 3, natm), &
              quad4(3, 3, 3, natm), ipolar(natm), lst(natm * vnlmax), pnt(natm + 1))
    zero1 = 0.0_dp
    quad3 = 0.0_dp
    quad4 = 0.0_dp
    pold = 0.0_dp
    ipold = 0.0_dp
    do ia = 1, natm
      it = itype(ia)
      q0(ia) = qch(it)
      m0(ia) = mass(it)
      dipia(:, ia) = dip(it, :)
      ipolar(ia) = ldip_polar(it)
      if (ldip_polar(it)) then
        pold(:, :, ia) = poldip(it, :, :)
        ! 3x3 inverse (typeinfo_init's DGETRF/DGETRI, src/read_pos.f90:203-212)
        det = pold(1, 1, ia) * (pold(2, 2, ia) * pold(3, 3, ia) - pold(2, 3, ia) * pold(3, 2, ia)) - &
              pold(1, 2, ia) * (pold(2, 1, ia) * pold(3, 3, ia) - pold(2, 3, ia) * pold(3, 1, ia)) + &
              pold(1, 3, ia) * (pold(2, 1, ia) * pold(3, 2, ia) - pold(2, 2, ia) * pold(3, 1, ia))
        ipold(1, 1, ia) = (pold(2, 2, ia) * pold(3, 3, ia) - pold(2, 3, ia) * pold(3, 2, ia)) / det
        ipold(1, 2, ia) = (pold(1, 3, ia) * pold(3, 2, ia) - pold(1, 2, ia) * pold(3, 3, ia)) / det
        ipold(1, 3, ia) = (pold(1, 2, ia) * pold(2, 3, ia) - pold(1, 3, ia) * pold(2, 2, ia)) / det
        ipold(2, 1, ia) = (pold(2, 3, ia) * pold(3, 1, ia) - pold(2, 1, ia) * pold(3, 3, ia)) / det
        ipold(2, 2, ia) = (pold(1, 1, ia) * pold(3, 3, ia) - pold(1, 3, ia) * pold(3, 1, ia)) / det
        ipold(2, 3, ia) = (pold(1, 3, ia) * pold(2, 1, ia) - pold(1, 1, ia) * pold(2, 3, ia)) / det
        ipold(3, 1, ia) = (pold(2, 1, ia) * pold(3, 2, ia) - pold(2, 2, ia) * pold(3, 1, ia)) / det
        ipold(3, 2, ia) = (pold(1, 2, ia) * pold(3, 1, ia) - pold(1, 1, ia) * pold(3, 2, ia)) / det
        ipold(3, 3, ia) = (pold(1, 1, ia) * pold(2, 2, ia) - pold(1, 2, ia) * pold(2, 1, ia)) / det
      end if
      ! fractional coordinates of the current positions (the saved list
      ! positions of vnlistcheck after a rebuild here)
      sfx(ia) = rx(ia) * B(1, 1) + ry(ia) * B(2, 1) + rz(ia) * B(3, 1)
      sfy(ia) = rx(ia) * B(1, 2) + ry(ia) * B(2, 2) + rz(ia) * B(3, 2)
      sfz(ia) = rx(ia) * B(1, 3) + ry(ia) * B(2, 3) + rz(ia) * B(3, 3)
    end do
    mu_t = 0.0_dp
    if (lcoulomb) then
      st = mdff_get_dipoles(ctx, mu_t, ef_t)
      call mdff_check(st, ctx, 'get_dipoles')
    end if
    task_coul = .false.
    lfalse = .false.
    nth = max(nhc_n, 1)
    if (trim(integrator) /= 'nvt-nhcn') nth = 1
    xiw = 0.0_dp
    vxiw = 0.0_dp
    vxib = 0.0_dp
    xib = 0.0_dp
    if (trim(integrator) == 'nvt-nhcn') then
      xiw(1:nhc_n) = xi(1:nhc_n)
      vxiw(1:nhc_n) = vxi(1:nhc_n)
    end if
    open (newunit=uu, file='RESTART', form='unformatted', status='replace')
    write (uu) lnmlj
    write (uu) lcoulomb
    write (uu) lmorse
    write (uu) lbmhft
    write (uu) lbmhftd
    write (uu) lsurf
    write (uu) lcsvr
    write (uu) lharm
    write (uu) ltraj
    write (uu) lvnlist
    write (uu) lstatic
    write (uu) lreduced
    write (uu) lreducedN
    write (uu) ltest
    write (uu) lmsd
    write (uu) lvacf
    write (uu) lwrite_restart
    write (uu) cutlongrange
    write (uu) cutshortrange
    write (uu) calc
    write (uu) dgauss
    write (uu) longrange
    write (uu) itraj_start
    write (uu) itraj_period
    write (uu) itraj_format
    write (uu) trajff_data
    write (uu) iscff_format
    write (uu) iscff_data
    write (uu) iefall_format
    write (uu) iefgall_format
    write (uu) idipall_format
    write (uu) posff_data
    write (uu) skindiff
    write (uu) natm
    write (uu) system_name
    write (uu) A
    write (uu) ntype
    write (uu) ntype
    write (uu) atypei0
    write (uu) natmi0
    write (uu) rx, ry, rz
    write (uu) vx, vy, vz
    write (uu) fx, fy, fz
    write (uu) zero1, zero1, zero1
    write (uu) zero1, zero1, zero1
    write (uu) sfx, sfy, sfz
    write (uu) atype
    write (uu) itype
    do j = 0, 1
      lst = 0
      pnt = 0
      if (lvnlist) then
        st = mdff_get_pairs(ctx, int(j, c_int32_t), pnt, lst, int(size(lst), c_int64_t), nl)
        call mdff_check(st, ctx, 'get_pairs')
      end if
      write (uu) lst
      write (uu) pnt
    end do
    write (uu) cutshortrange
    write (uu) cutlongrange
    write (uu) q0
    write (uu) m0
    write (uu) zero1
    write (uu) dipia
    write (uu) quad3
    write (uu) pold
    write (uu) quad4
    write (uu) ipold
    write (uu) ipolar
    write (uu) dt * time_unit
    write (uu) npas
    write (uu) itime_next
    write (uu) nequil
    write (uu) nequil_period
    write (uu) spas
    write (uu) npropr
    write (uu) npropr_start
    write (uu) nprint
    write (uu) fprint
    write (uu) 0
    write (uu) setvel
    write (uu) temp * boltz_unit
    write (uu) press * press_unit
    write (uu) timesca_thermo * time_unit
    write (uu) timesca_baro * time_unit
    write (uu) tauTberendsen * time_unit
    write (uu) tauPberendsen * time_unit
    write (uu) taucsvr * time_unit
    write (uu) nuandersen
    write (uu) annealing
    write (uu) .true.
    write (uu) integrator
    write (uu) nhc_n
    write (uu) nhc_yosh_order
    write (uu) nhc_mults
    write (uu) vxiw(1:nth)
    write (uu) xiw(1:nth)
    write (uu) vxib(1:nth)
    write (uu) xib(1:nth)
    write (uu) 0.0_dp
    write (uu) 0.0_dp
    write (uu) 0.0_dp
    write (uu) lKA
    write (uu) ctrunc
    write (uu) 2
    write (uu) kES
    write (uu) alphaES
    write (uu) qlj
    write (uu) plj
    write (uu) sigmalj
    write (uu) epslj
    write (uu) sigmamor
    write (uu) epsmor
    write (uu) rhomor
    write (uu) mass
    write (uu) doefield
    write (uu) doefg
    write (uu) qch
    write (uu) quad_nuc
    write (uu) dip
    write (uu) quad
    write (uu) poldip
    write (uu) poldip_iso
    write (uu) polquad
    write (uu) polquad_iso
    write (uu) pol_damp_b
    write (uu) pol_damp_c
    write (uu) pol_damp_k
    write (uu) extrapolate_order
    write (uu) conv_tol_ind
    write (uu) min_scf_pol_iter
    write (uu) max_scf_pol_iter
    write (uu) algo_ext_dipole
    write (uu) thole_functions
    write (uu) thole_function_type
    write (uu) thole_param
    write (uu) omegakO
    write (uu) epsw
    write (uu) lautoES
    write (uu) lwrite_dip_wfc
    write (uu) lfalse
    write (uu) lfalse
    write (uu) lfalse
    write (uu) lfalse
    write (uu) ldip_wfc
    write (uu) rcut_wfc
    write (uu) ldip_polar
    write (uu) ldip_damping
    write (uu) lquad_polar
    write (uu) Abmhftd
    write (uu) Bbmhftd
    write (uu) Cbmhftd
    write (uu) Dbmhftd
    write (uu) BDbmhftd
    write (uu) task_coul
    write (uu) algo_moment_from_pola
    write (uu) mu_t
    close (uu)
  end subroutine write_restart

  ! restart_init (src/restart.f90:34-292): every record in the order
  ! written; the control, md and force-field variables it holds replace
  ! those of the control file, times and temperatures converted back to
  ! the input units the rest of the driver uses
  subroutine read_restart()
    integer :: uu, j, ntype2, nth, itmp, ncd
    character(len=3) :: atypei0(0:ntypemax)
    integer :: natmi0(0:ntypemax)
    integer, allocatable :: lst(:), pnt(:)
    real(dp), allocatable :: r1(:), r2(:), r3(:), d2(:, :), d3(:, :, :), d4(:, :, :, :)
    logical, allocatable :: l1(:)
    logical :: ltmp, task_coul(6)
    real(dp) :: rtmp, vc
    open (newunit=uu, file='RESTART', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write (*, '(a)') 'ERROR full_restart: cannot open RESTART'
      stop 1
    end if
    read (uu) lnmlj
    read (uu) lcoulomb
    read (uu) lmorse
    read (uu) lbmhft
    read (uu) lbmhftd
    read (uu) lsurf
    read (uu) lcsvr
    read (uu) lharm
    read (uu) ltraj
    read (uu) lvnlist
    read (uu) lstatic
    read (uu) lreduced
    read (uu) lreducedN
    read (uu) ltest
    read (uu) lmsd
    read (uu) lvacf
    read (uu) lwrite_restart
    read (uu) cutlongrange
    read (uu) cutshortrange
    read (uu) calc
    read (uu) dgauss
    read (uu) longrange
    read (uu) itraj_start
    read (uu) itraj_period
    read (uu) itraj_format
    read (uu) trajff_data
    read (uu) iscff_format
    read (uu) iscff_data
    read (uu) iefall_format
    read (uu) iefgall_format
    read (uu) idipall_format
    read (uu) posff_data
    read (uu) skindiff
    read (uu) natm
    read (uu) system_name
    read (uu) A
    read (uu) ntype
    read (uu) ntype2
    read (uu) atypei0
    read (uu) natmi0
    atypei(1:ntype) = atypei0(1:ntype)
    natmi(1:ntype) = natmi0(1:ntype)
    allocate (atype(natm), itype(natm), rx(natm), ry(natm), rz(natm), vx(natm), vy(natm), vz(natm), &
              fx(natm), fy(natm), fz(natm), r1(natm), r2(natm), r3(natm), lst(natm * vnlmax), pnt(natm + 1), &
              d2(3, natm), d3(3, 3, natm), d4(3, 3, 3, natm), l1(natm))
    read (uu) rx, ry, rz
    read (uu) vx, vy, vz
    read (uu) fx, fy, fz
    read (uu) r1, r2, r3
    read (uu) r1, r2, r3
    read (uu) r1, r2, r3
    read (uu) atype
    read (uu) itype
    do j = 0, 1
      read (uu) lst
      read (uu) pnt
    end do
    read (uu) rtmp
    read (uu) rtmp
    read (uu) r1
    read (uu) r1
    read (uu) r1
    read (uu) d2
    read (uu) d3
    read (uu) d3
    read (uu) d4
    read (uu) d3
    read (uu) l1
    read (uu) dt
    dt = dt / time_unit
    read (uu) npas
    read (uu) itime_rst
    read (uu) nequil
    read (uu) nequil_period
    read (uu) spas
    read (uu) npropr
    read (uu) npropr_start
    read (uu) nprint
    read (uu) fprint
    read (uu) itmp
    read (uu) setvel
    read (uu) temp
    temp = temp / boltz_unit
    read (uu) press
    press = press / press_unit
    read (uu) timesca_thermo
    timesca_thermo = timesca_thermo / time_unit
    read (uu) timesca_baro
    timesca_baro = timesca_baro / time_unit
    read (uu) tauTberendsen
    tauTberendsen = tauTberendsen / time_unit
    read (uu) tauPberendsen
    tauPberendsen = tauPberendsen / time_unit
    read (uu) taucsvr
    taucsvr = taucsvr / time_unit
    read (uu) nuandersen
    read (uu) annealing
    read (uu) ltmp
    read (uu) integrator
    read (uu) nhc_n
    read (uu) nhc_yosh_order
    read (uu) nhc_mults
    nth = 1
    if (trim(integrator) == 'nvt-nhcn' .or. trim(integrator) == 'npt-nhcnp') nth = nhc_n
    allocate (xi_rst(nth), vxi_rst(nth))
    read (uu) vxi_rst
    read (uu) xi_rst
    read (uu) r1(1:nth)
    read (uu) r1(1:nth)
    read (uu) vc
    read (uu) vc
    read (uu) vc
    read (uu) lKA
    read (uu) ctrunc
    read (uu) ncd
    read (uu) kES
    read (uu) alphaES
    read (uu) qlj
    read (uu) plj
    read (uu) sigmalj
    read (uu) epslj
    read (uu) sigmamor
    read (uu) epsmor
    read (uu) rhomor
    read (uu) mass
    read (uu) doefield
    read (uu) doefg
    read (uu) qch
    read (uu) quad_nuc
    read (uu) dip
    read (uu) quad
    read (uu) poldip
    read (uu) poldip_iso
    read (uu) polquad
    read (uu) polquad_iso
    read (uu) pol_damp_b
    read (uu) pol_damp_c
    read (uu) pol_damp_k
    read (uu) extrapolate_order
    read (uu) conv_tol_ind
    read (uu) min_scf_pol_iter
    read (uu) max_scf_pol_iter
    read (uu) algo_ext_dipole
    read (uu) thole_functions
    read (uu) thole_function_type
    read (uu) thole_param
    read (uu) omegakO
    read (uu) epsw
    read (uu) lautoES
    read (uu) lwrite_dip_wfc
    read (uu) ltmp
    read (uu) ltmp
    read (uu) ltmp
    read (uu) ltmp
    read (uu) ldip_wfc
    read (uu) rcut_wfc
    read (uu) ldip_polar
    read (uu) ldip_damping
    read (uu) lquad_polar
    read (uu) Abmhftd
    read (uu) Bbmhftd
    read (uu) Cbmhftd
    read (uu) Dbmhftd
    read (uu) BDbmhftd
    read (uu) task_coul
    read (uu) algo_moment_from_pola
    read (uu) d2
    close (uu)
    ! the tables were stored in full: no symmetrisation or automatic Ewald
    ! parameters again (restart_init calls initialize_param_non_bonded /
    ! initialize_coulomb on the stored values)
    sym_nb = .false.
    sym_pim = .false.
    lautoES = .false.
    poldip_iso = 0.0_dp
  end subroutine read_restart

  ! CONTFF (src/config.f90:208-252): Cartesian positions, velocities, forces
  subroutine write_contff()
    integer :: uu
    st = mdff_get_positions(ctx, rx, ry, rz)
    call mdff_check(st, ctx, 'get_positions')
    st = mdff_get_velocities(ctx, vx, vy, vz)
    call mdff_check(st, ctx, 'get_velocities')
    st = mdff_get_forces(ctx, fx, fy, fz)
    call mdff_check(st, ctx, 'get_forces')
    open (newunit=uu, file='CONTFF', status='replace')
    write (uu, '(i8)') natm
    write (uu, '(a)') trim(system_name)
    write (uu, '(3f20.12)') A(1, 1), A(2, 1), A(3, 1)
    write (uu, '(3f20.12)') A(1, 2), A(2, 2), A(3, 2)
    write (uu, '(3f20.12)') A(1, 3), A(2, 3), A(3, 3)
    write (uu, '(i4)') ntype
    write (uu, *) (atypei(it), it=1, ntype)
    write (uu, *) (natmi(it), it=1, ntype)
    write (uu, '(a)') 'Cartesian'
    do ia = 1, natm
      write (uu, '(a,9e24.16)') atype(ia), rx(ia), ry(ia), rz(ia), vx(ia), vy(ia), vz(ia), fx(ia), fy(ia), fz(ia)
    end do
    close (uu)
  end subroutine write_contff

end program mdff_gpu_driver
