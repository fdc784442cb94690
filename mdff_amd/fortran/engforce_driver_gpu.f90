! engforce_driver_gpu.f90 -- drop-in replacement of the reference's
! MODULE engforce_driver (src/engforce_driver.f90:46-190) for a reference
! build that links libmdff_gpu: same module name, same public type tables,
! same argument-less SUBROUTINE engforce.  The first call creates the GPU
! context from the reference's module globals; every call copies the
! positions in, runs mdff_engforce and copies forces, energies, virials,
! stress tensors, dipoles, fields and field gradients back into the module
! globals the rest
! of the reference reads.  Build it in place of src/engforce_driver.f90,
! together with mdff_gpu_iface.f90, and link -lmdff_gpu.
MODULE engforce_driver
  USE, INTRINSIC :: iso_c_binding
  USE constants, ONLY : dp, coul_unit, press_unit
  USE config,    ONLY : ntypemax
  USE mdff_gpu,  gpu_dp => dp
  implicit none
  ! the reference's public type tables (src/engforce_driver.f90:64-75)
  real(kind=dp) :: mass(ntypemax), qch(ntypemax), dip(ntypemax,3), quad(ntypemax,3,3)
  real(kind=dp) :: poldip(ntypemax,3,3), poldip_iso(ntypemax)
  real(kind=dp) :: polquad(ntypemax,3,3,3), polquad_iso(ntypemax), quad_nuc(ntypemax)
  logical       :: ldip_polar(ntypemax)
  type(c_ptr), save :: gpu_ctx = c_null_ptr
CONTAINS

SUBROUTINE engforce_gpu_init
  USE config,     ONLY : natm, ntype, itype, simu_cell
  USE control,    ONLY : lnmlj, lbmhft, lbmhftd, lmorse, lcoulomb, lsurf, lvnlist, calc, &
                         cutshortrange, cutlongrange, skindiff
  USE non_bonded, ONLY : epslj, sigmalj, plj, qlj, epsmor, sigmamor, rhomor, &
                         Abmhftd, Bbmhftd, Cbmhftd, Dbmhftd, BDbmhftd, trunc
  USE coulomb,    ONLY : alphaES, kES, doefield, doefg
  USE pim,        ONLY : ldip_damping, pol_damp_b, pol_damp_c, pol_damp_k, conv_tol_ind, &
                         min_scf_pol_iter, max_scf_pol_iter, extrapolate_order, &
                         algo_moment_from_pola, omegakO
  type(mdff_desc) :: d
  integer(c_int)  :: st
  call mdff_desc_defaults(d)
  d%natm = natm ; d%ntype = ntype ; d%precision = MDFF_PREC_FP64
  d%cell = reshape(simu_cell%A, [9])
  d%lnmlj = merge(1,0,lnmlj) ; d%lbmhft = merge(1,0,lbmhft) ; d%lbmhftd = merge(1,0,lbmhftd)
  d%lmorse = merge(1,0,lmorse) ; d%lcoulomb = merge(1,0,lcoulomb) ; d%lsurf = merge(1,0,lsurf)
  d%lvnlist = merge(1,0,lvnlist) ; d%calc_md = merge(1,0,calc .eq. 'md')
  d%cutshortrange = cutshortrange ; d%cutlongrange = cutlongrange ; d%skindiff = skindiff
  d%coul_unit = coul_unit ; d%press_unit = press_unit ; d%trunc = trunc
  d%epslj = reshape(epslj,[256]) ; d%sigmalj = reshape(sigmalj,[256])
  d%plj = reshape(plj,[256]) ; d%qlj = reshape(qlj,[256])
  d%epsmor = reshape(epsmor,[256]) ; d%sigmamor = reshape(sigmamor,[256]) ; d%rhomor = reshape(rhomor,[256])
  d%Abmhftd = reshape(Abmhftd,[256]) ; d%Bbmhftd = reshape(Bbmhftd,[256])
  d%Cbmhftd = reshape(Cbmhftd,[256]) ; d%Dbmhftd = reshape(Dbmhftd,[256]) ; d%BDbmhftd = reshape(BDbmhftd,[256])
  d%alphaES = alphaES ; d%kES = kES ; d%doefield = merge(1,0,doefield) ; d%doefg = merge(1,0,doefg)
  d%mass = mass ; d%qch = qch ; d%dip = reshape(dip,[48])
  d%ldip_polar = merge(1,0,ldip_polar) ; d%poldip = reshape(poldip,[144])
  d%ldip_damping = reshape(merge(1,0,ldip_damping),[4096])
  d%pol_damp_b = reshape(pol_damp_b,[4096]) ; d%pol_damp_c = reshape(pol_damp_c,[4096])
  d%pol_damp_k = reshape(pol_damp_k,[4096])
  d%conv_tol_ind = conv_tol_ind ; d%min_scf_pol_iter = min_scf_pol_iter
  d%max_scf_pol_iter = max_scf_pol_iter ; d%extrapolate_order = extrapolate_order
  d%algo_moment_from_pola = 0 ; d%omegakO = omegakO
  if (algo_moment_from_pola .eq. 'scf_kO_v2') d%algo_moment_from_pola = 1
  if (algo_moment_from_pola .eq. 'scf_kO_v1') d%algo_moment_from_pola = 2
  st = mdff_ctx_create(d, itype, gpu_ctx)
  call mdff_check(st, gpu_ctx, 'engforce_gpu_init')
END SUBROUTINE engforce_gpu_init

SUBROUTINE engforce
  USE config,        ONLY : natm, rx, ry, rz, fx, fy, fz, tau, tau_nonb, tau_coul
  USE control,       ONLY : lcoulomb
  USE coulomb,       ONLY : mu_t, ef_t, efg_t
  USE md,            ONLY : itime
  USE thermodynamic, ONLY : u_lj, u_bmhft, u_morse, u_coul, u_pol, u_tail, &
                            vir_lj, vir_bmhft, vir_morse, pvirial_coul
  type(mdff_terms) :: t
  integer(c_int)   :: st
  if (.not. c_associated(gpu_ctx)) call engforce_gpu_init
  st = mdff_set_itime(gpu_ctx, int(itime, c_int32_t))
  st = mdff_set_positions(gpu_ctx, rx, ry, rz)
  call mdff_check(st, gpu_ctx, 'engforce')
  st = mdff_engforce(gpu_ctx, fx, fy, fz, t)
  call mdff_check(st, gpu_ctx, 'engforce')
  u_lj = t%u_lj ; u_bmhft = t%u_bmhft ; u_morse = t%u_morse ; u_tail = t%u_tail
  u_coul = t%u_coul ; u_pol = t%u_pol
  vir_lj = t%vir_lj ; vir_bmhft = t%vir_bmhft ; vir_morse = t%vir_morse
  pvirial_coul = t%pvirial_coul
  tau      = reshape(t%tau, [3,3])
  tau_nonb = reshape(t%tau_nonb, [3,3])
  tau_coul = reshape(t%tau_coul, [3,3])
  if (lcoulomb) then
    st = mdff_get_dipoles(gpu_ctx, mu_t, ef_t)
    call mdff_check(st, gpu_ctx, 'engforce')
    st = mdff_get_efg(gpu_ctx, efg_t)
    call mdff_check(st, gpu_ctx, 'engforce')
  endif
END SUBROUTINE engforce

END MODULE engforce_driver
