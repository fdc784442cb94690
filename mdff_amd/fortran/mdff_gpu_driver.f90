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
  call read_posff()
  call fill_desc()

  st = mdff_ctx_create(d, itype, ctx)
  call mdff_check(st, ctx, 'mdff_ctx_create')
  write (*, '(a,a)') 'engforce library : ', mdff_version_string()
  st = mdff_set_positions(ctx, rx, ry, rz)
  call mdff_check(st, ctx, 'mdff_set_positions')
  call init_velocities()
  st = mdff_set_velocities(ctx, vx, vy, vz)
  call mdff_check(st, ctx, 'mdff_set_velocities')

  ! ---- t = 0 force call
  st = mdff_engforce(ctx, fx, fy, fz, t)
  call mdff_check(st, ctx, 'engforce')
  call print_terms(t)
  open (newunit=u, file='FORCES', status='replace')
  do ia = 1, natm
    write (u, '(3es25.16e3)') fx(ia), fy(ia), fz(ia)
  end do
  close (u)

  ! ---- MD
  if (npas > 0 .and. trim(calc) == 'md') then
    if (trim(integrator) /= 'nve-vv' .and. trim(integrator) /= 'nvt-nhcn') then
      write (*, '(a,a)') 'ERROR integrator not available on this path: ', trim(integrator)
      stop 1
    end if
    if (nequil > 0 .and. trim(integrator) == 'nve-vv' .and. .not. lcsvr) then
      write (*, '(a)') 'ERROR Berendsen equilibration rescale not available on this path (use lcsvr)'
      stop 1
    end if
    write (*, '(a)') '      step        time(ps)              etot              ekin              upot' // &
                     '              temp'
    allocate (ek(max(nprint, 1)), up(max(nprint, 1)), env(max(nprint, 1)))
    allocate (xi(nhc_n), vxi(nhc_n))
    xi = 0.0_dp
    vxi = 0.0_dp
    conint = 0.0_dp
    kT = temp * boltz_unit
    itime = 0
    do while (itime < npas)
      ! one step at a time while the equilibration rescale may apply, else a
      ! chunk up to the next print
      if (lcsvr .and. itime + 1 <= nequil) then
        nchunk = 1
      else
        nchunk = min(max(nprint, 1) - mod(itime, max(nprint, 1)), npas - itime)
      end if
      if (trim(integrator) == 'nve-vv') then
        st = mdff_md_vv(ctx, int(nchunk, c_int32_t), dt * time_unit, ek, up)
      else
        st = mdff_md_nhcn(ctx, int(nchunk, c_int32_t), dt * time_unit, kT, timesca_thermo * time_unit, &
                          int(nhc_n, c_int32_t), int(nhc_yosh_order, c_int32_t), &
                          int(nhc_mults, c_int32_t), xi, vxi, ek, up, env)
      end if
      call mdff_check(st, ctx, 'md step')
      itime = itime + nchunk
      ! equilibration rescale after the step and its skin check (src/runmd.f90:316-330)
      if (lcsvr .and. itime <= nequil .and. mod(itime, max(nequil_period, 1)) == 0 .and. &
          itime /= 1 .and. itime /= npas) then
        call gauss(r1)
        sumnoise = sum_noises(3 * natm - 4)
        st = mdff_csvr_rescale(ctx, kT, taucsvr * time_unit, r1, sumnoise, conint)
        call mdff_check(st, ctx, 'csvr rescale')
      end if
      if (mod(itime, max(nprint, 1)) == 0 .or. itime == npas) then
        st = mdff_calc_temp(ctx, boltz_unit, tempi, ekin)
        call mdff_check(st, ctx, 'calc_temp')
        upot = up(nchunk)
        etot = ekin + upot
        write (*, '(i10,f16.6,4es18.9)') itime, itime * dt, etot, ekin, upot, tempi
      end if
    end do
    call write_contff()
  end if

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
    if (any(lquad_polar) .or. any(quad /= 0.0_dp) .or. doefg .or. thole_functions .or. lwfc(1) /= 0 .or. &
        (trim(algo_moment_from_pola) /= 'scf' .and. trim(algo_moment_from_pola) /= 'scf_kO_v2') .or. &
        lmorse .and. (lnmlj .or. lbmhft .or. lbmhftd)) then
      write (*, '(a)') 'ERROR option outside the engforce path of this library ' // &
                       '(quadrupoles, EFG, Thole, Wannier centres, SCF variants other than scf and scf_kO_v2)'
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
    d%mass = mass; d%qch = qch; d%dip = reshape(dip, [48])
    d%ldip_polar = merge(1, 0, ldip_polar)
    d%poldip = reshape(poldip, [144])
    d%ldip_damping = reshape(merge(1, 0, ldip_damping), [4096])
    d%pol_damp_b = reshape(pol_damp_b, [4096]); d%pol_damp_c = reshape(pol_damp_c, [4096])
    d%pol_damp_k = reshape(pol_damp_k, [4096])
    d%conv_tol_ind = conv_tol_ind
    d%min_scf_pol_iter = min_scf_pol_iter; d%max_scf_pol_iter = max_scf_pol_iter
    d%extrapolate_order = extrapolate_order
    d%algo_moment_from_pola = merge(1, 0, trim(algo_moment_from_pola) == 'scf_kO_v2')
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
