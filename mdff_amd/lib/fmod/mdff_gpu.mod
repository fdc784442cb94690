﻿!mod$ v1 sum:37af09ba10df4b1c
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module mdff_gpu
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::dp=8_4
integer(4),parameter::mdff_ntypemax=16_4
integer(4),parameter::mdff_abi_version=2_4
integer(4),parameter::mdff_ok=0_4
integer(4),parameter::mdff_err_bad_arg=1_4
integer(4),parameter::mdff_err_list_overflow=2_4
integer(4),parameter::mdff_err_scf_not_converged=3_4
integer(4),parameter::mdff_err_polar_catastrophe=4_4
integer(4),parameter::mdff_err_hip=5_4
integer(4),parameter::mdff_err_no_device=6_4
integer(4),parameter::mdff_err_unsupported=7_4
integer(4),parameter::mdff_prec_fp64=0_4
integer(4),parameter::mdff_prec_mixed=1_4
integer(4),parameter::mdff_trunc_none=0_4
integer(4),parameter::mdff_trunc_linear=1_4
integer(4),parameter::mdff_trunc_quadratic=2_4
type,bind(c)::mdff_desc
integer(4)::abi_version
integer(4)::natm
integer(4)::ntype
integer(4)::precision
real(8)::cell(1_8:9_8)
integer(4)::lnmlj
integer(4)::lbmhft
integer(4)::lbmhftd
integer(4)::lmorse
integer(4)::lcoulomb
integer(4)::lsurf
integer(4)::lvnlist
integer(4)::calc_md
real(8)::cutshortrange
real(8)::cutlongrange
real(8)::skindiff
real(8)::coul_unit
real(8)::press_unit
integer(4)::trunc
integer(4)::pad0
real(8)::epslj(1_8:256_8)
real(8)::sigmalj(1_8:256_8)
real(8)::plj(1_8:256_8)
real(8)::qlj(1_8:256_8)
real(8)::epsmor(1_8:256_8)
real(8)::sigmamor(1_8:256_8)
real(8)::rhomor(1_8:256_8)
real(8)::abmhftd(1_8:256_8)
real(8)::bbmhftd(1_8:256_8)
real(8)::cbmhftd(1_8:256_8)
real(8)::dbmhftd(1_8:256_8)
real(8)::bdbmhftd(1_8:256_8)
real(8)::alphaes
integer(4)::kes(1_8:3_8)
integer(4)::doefield
real(8)::mass(1_8:16_8)
real(8)::qch(1_8:16_8)
real(8)::dip(1_8:48_8)
integer(4)::ldip_polar(1_8:16_8)
real(8)::poldip(1_8:144_8)
integer(4)::ldip_damping(1_8:4096_8)
real(8)::pol_damp_b(1_8:4096_8)
real(8)::pol_damp_c(1_8:4096_8)
integer(4)::pol_damp_k(1_8:4096_8)
real(8)::conv_tol_ind
integer(4)::min_scf_pol_iter
integer(4)::max_scf_pol_iter
integer(4)::extrapolate_order
integer(4)::algo_moment_from_pola
real(8)::omegako
end type
type,bind(c)::mdff_terms
real(8)::u_lj
real(8)::u_bmhft
real(8)::u_morse
real(8)::vir_lj
real(8)::vir_bmhft
real(8)::vir_morse
real(8)::u_coul
real(8)::u_pol
real(8)::u_dir
real(8)::u_rec
real(8)::u_self
real(8)::u_surf
real(8)::pvirial_coul
real(8)::u_tail
real(8)::p_tail
real(8)::tau(1_8:9_8)
real(8)::tau_nonb(1_8:9_8)
real(8)::tau_coul(1_8:9_8)
integer(4)::n_scf
integer(4)::list_rebuilt
real(8)::rmsd
end type
interface
subroutine mdff_desc_defaults(d) bind(c,name="mdff_desc_defaults")
import::mdff_desc
type(mdff_desc),intent(inout)::d
end
end interface
interface
function mdff_ctx_create(d,itype,ctx) bind(c,name="mdff_ctx_create")
import::c_ptr
import::mdff_desc
type(mdff_desc),intent(in)::d
integer(4),intent(in)::itype(1_8:*)
type(c_ptr),intent(out)::ctx
integer(4)::mdff_ctx_create
end
end interface
interface
function mdff_ctx_destroy(ctx) bind(c,name="mdff_ctx_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::mdff_ctx_destroy
end
end interface
interface
function mdff_last_error_c(ctx) bind(c,name="mdff_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::mdff_last_error_c
end
end interface
interface
function mdff_version_c() bind(c,name="mdff_version")
import::c_ptr
type(c_ptr)::mdff_version_c
end
end interface
interface
function mdff_set_positions(ctx,rx,ry,rz) bind(c,name="mdff_set_positions")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::rx(1_8:*)
real(8),intent(in)::ry(1_8:*)
real(8),intent(in)::rz(1_8:*)
integer(4)::mdff_set_positions
end
end interface
interface
function mdff_get_positions(ctx,rx,ry,rz) bind(c,name="mdff_get_positions")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::rx(1_8:*)
real(8),intent(out)::ry(1_8:*)
real(8),intent(out)::rz(1_8:*)
integer(4)::mdff_get_positions
end
end interface
interface
function mdff_set_velocities(ctx,vx,vy,vz) bind(c,name="mdff_set_velocities")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::vx(1_8:*)
real(8),intent(in)::vy(1_8:*)
real(8),intent(in)::vz(1_8:*)
integer(4)::mdff_set_velocities
end
end interface
interface
function mdff_get_velocities(ctx,vx,vy,vz) bind(c,name="mdff_get_velocities")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::vx(1_8:*)
real(8),intent(out)::vy(1_8:*)
real(8),intent(out)::vz(1_8:*)
integer(4)::mdff_get_velocities
end
end interface
interface
function mdff_get_forces(ctx,fx,fy,fz) bind(c,name="mdff_get_forces")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::fx(1_8:*)
real(8),intent(out)::fy(1_8:*)
real(8),intent(out)::fz(1_8:*)
integer(4)::mdff_get_forces
end
end interface
interface
function mdff_set_cell(ctx,cell) bind(c,name="mdff_set_cell")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::cell(1_8:9_8)
integer(4)::mdff_set_cell
end
end interface
interface
function mdff_vnlist_build(ctx) bind(c,name="mdff_vnlist_build")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::mdff_vnlist_build
end
end interface
interface
function mdff_vnlistcheck(ctx,rebuilt) bind(c,name="mdff_vnlistcheck")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(out)::rebuilt
integer(4)::mdff_vnlistcheck
end
end interface
interface
function mdff_get_pairs(ctx,which,point,list,cap,n) bind(c,name="mdff_get_pairs")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::which
integer(4),intent(out)::point(1_8:*)
integer(4),intent(out)::list(1_8:*)
integer(8),value::cap
integer(8),intent(out)::n
integer(4)::mdff_get_pairs
end
end interface
interface
function mdff_engforce(ctx,fx,fy,fz,terms) bind(c,name="mdff_engforce")
import::c_ptr
import::mdff_terms
type(c_ptr),value::ctx
real(8),intent(out)::fx(1_8:*)
real(8),intent(out)::fy(1_8:*)
real(8),intent(out)::fz(1_8:*)
type(mdff_terms),intent(out)::terms
integer(4)::mdff_engforce
end
end interface
interface
function mdff_get_dipoles(ctx,mu,ef) bind(c,name="mdff_get_dipoles")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::mu(1_8:3_8,1_8:*)
real(8),intent(out)::ef(1_8:3_8,1_8:*)
integer(4)::mdff_get_dipoles
end
end interface
interface
function mdff_set_itime(ctx,itime) bind(c,name="mdff_set_itime")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::itime
integer(4)::mdff_set_itime
end
end interface
interface
function mdff_md_vv(ctx,nsteps,dt,ekin,upot) bind(c,name="mdff_md_vv")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nsteps
real(8),value::dt
real(8),intent(out)::ekin(1_8:*)
real(8),intent(out)::upot(1_8:*)
integer(4)::mdff_md_vv
end
end interface
interface
function mdff_md_nhcn(ctx,nsteps,dt,temp,timesca_thermo,nhc_n,yosh_order,nhc_mults,xi,vxi,ekin,upot,e_nvt) bind(c,name="mdff_md_nhcn")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nsteps
real(8),value::dt
real(8),value::temp
real(8),value::timesca_thermo
integer(4),value::nhc_n
integer(4),value::yosh_order
integer(4),value::nhc_mults
real(8),intent(inout)::xi(1_8:*)
real(8),intent(inout)::vxi(1_8:*)
real(8),intent(out)::ekin(1_8:*)
real(8),intent(out)::upot(1_8:*)
real(8),intent(out)::e_nvt(1_8:*)
integer(4)::mdff_md_nhcn
end
end interface
interface
function mdff_csvr_rescale(ctx,temp,taut,r1,sumnoise,conint) bind(c,name="mdff_csvr_rescale")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::temp
real(8),value::taut
real(8),value::r1
real(8),value::sumnoise
real(8),intent(inout)::conint
integer(4)::mdff_csvr_rescale
end
end interface
interface
function mdff_calc_temp(ctx,boltz_unit,temp,ekin) bind(c,name="mdff_calc_temp")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::boltz_unit
real(8),intent(out)::temp
real(8),intent(out)::ekin
integer(4)::mdff_calc_temp
end
end interface
interface
function mdff_last_timings(ctx,ms,n) bind(c,name="mdff_last_timings")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::ms(1_8:8_8)
integer(4),intent(out)::n
integer(4)::mdff_last_timings
end
end interface
private::c_strlen
interface
function c_strlen(s) bind(c,name="strlen")
import::c_ptr
type(c_ptr),value::s
integer(8)::c_strlen
end
end interface
private::c_to_f
contains
function c_to_f(p) result(s)
type(c_ptr),intent(in)::p
character(:,1),allocatable::s
end
function mdff_error_string(ctx) result(s)
type(c_ptr),intent(in)::ctx
character(:,1),allocatable::s
end
function mdff_version_string() result(s)
character(:,1),allocatable::s
end
subroutine mdff_check(status,ctx,where)
integer(4),intent(in)::status
type(c_ptr),intent(in)::ctx
character(*,1),intent(in)::where
end
end
