﻿!mod$ v1 sum:4f1e5ecd77cdb00c
!need$ 0bde2ac47243ead2 i iso_c_binding
module nft_hip
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
integer(4),parameter::npoinx=800000_4
integer(4),parameter::nedgex=4000000_4
integer(4),parameter::nfacesx=110000_4
type,bind(c)::nft_mesh_t
integer(4)::npoin
integer(4)::nedge
integer(4)::nbface
integer(4)::nboun
integer(4)::ld_poin
integer(4)::ld_edge
integer(4)::ld_face
type(c_ptr)::iedge
type(c_ptr)::sn
type(c_ptr)::kbface
type(c_ptr)::sb
type(c_ptr)::ibpoin
type(c_ptr)::sbb
type(c_ptr)::anp
type(c_ptr)::coord
type(c_ptr)::vol
integer(4)::ncol
type(c_ptr)::col_start
end type
type,bind(c)::nft_params_t
real(8)::dt
real(8)::u0
real(8)::v0
real(8)::w0
real(8)::strat
real(8)::theta0
integer(4)::lord
integer(4)::itr
integer(4)::exit_mode
real(8)::eps0
integer(4)::precon
integer(4)::sm_flag
integer(4)::itrp1
real(8)::omega
real(8)::dxabl
real(8)::dxabr
real(8)::dyabd
real(8)::dyabu
real(8)::dzabd
real(8)::dzabu
real(8)::towxl
real(8)::towxr
real(8)::towyd
real(8)::towyu
real(8)::towzd
real(8)::towzu
integer(4)::irl
real(8)::dom(1_8:6_8)
integer(4)::precision
integer(4)::itraj
end type
type,bind(c)::nft_gcr_stats_t
integer(4)::niter
integer(4)::nitsm
integer(4)::icount
integer(4)::converged
real(8)::eer
real(8)::eem
real(8)::res_max
real(8)::tol
end type
type,bind(c)::nft_run_stats_t
integer(4)::steps
integer(4)::gcr_solves
integer(4)::gcr_iters
integer(4)::gcr_unconverged
real(8)::ms_total
real(8)::ms_gcr
real(8)::ms_mpdata
end type
type,bind(c)::nft_diag_t
real(8)::div(1_8:4_8)
real(8)::eer
real(8)::eem
integer(4)::niter
integer(4)::nitav
integer(4)::crnic
real(8)::cour(1_8:4_8)
real(8)::u(1_8:4_8)
real(8)::v(1_8:4_8)
real(8)::w(1_8:4_8)
real(8)::t(1_8:4_8)
integer(8)::step
end type
type,bind(c)::nft_multigrid_t
integer(4)::mg_flag
integer(4)::itrp2
integer(4)::itrp_c1
integer(4)::itrp_c2
integer(4)::itrp_cc
integer(4)::transfer
real(8)::omega_c
end type
type(c_ptr),save::nft_context
real(8),save::nft_dt
interface
function nft_init(mesh,params,ctx) bind(c,name="nft_init")
import::c_ptr
import::nft_mesh_t
import::nft_params_t
type(nft_mesh_t),intent(in)::mesh
type(nft_params_t),intent(in)::params
type(c_ptr),intent(out)::ctx
integer(4)::nft_init
end
end interface
interface
function nft_set_multigrid(ctx,m1,match1,m2,match2,mg) bind(c,name="nft_set_multigrid")
import::c_ptr
import::nft_mesh_t
import::nft_multigrid_t
type(c_ptr),value::ctx
type(nft_mesh_t),intent(in)::m1
type(c_ptr),value::match1
type(nft_mesh_t),intent(in)::m2
type(c_ptr),value::match2
type(nft_multigrid_t),intent(in)::mg
integer(4)::nft_set_multigrid
end
end interface
interface
function nft_finalize(ctx) bind(c,name="nft_finalize")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::nft_finalize
end
end interface
interface
function nft_last_error() bind(c,name="nft_last_error")
import::c_ptr
type(c_ptr)::nft_last_error
end
end interface
interface
function nft_derivative(ctx,unk,dpdxa,idx) bind(c,name="nft_derivative")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::unk(1_8:*)
real(8),intent(out)::dpdxa(1_8:*)
integer(4),value::idx
integer(4)::nft_derivative
end
end interface
interface
function nft_laplc(ctx,press,alap,impfl) bind(c,name="nft_laplc")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::press(1_8:*)
real(8),intent(out)::alap(1_8:*)
integer(4),value::impfl
integer(4)::nft_laplc
end
end interface
interface
function nft_prfr(ctx,press,ustr,vstr,wstr,uout,vout,wout,impfl) bind(c,name="nft_prfr")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::press(1_8:*)
real(8),intent(in)::ustr(1_8:*)
real(8),intent(in)::vstr(1_8:*)
real(8),intent(in)::wstr(1_8:*)
real(8),intent(out)::uout(1_8:*)
real(8),intent(out)::vout(1_8:*)
real(8),intent(out)::wout(1_8:*)
integer(4),value::impfl
integer(4)::nft_prfr
end
end interface
interface
function nft_bcv(ctx,u,v,w,flag) bind(c,name="nft_bcv")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(inout)::u(1_8:*)
real(8),intent(inout)::v(1_8:*)
real(8),intent(inout)::w(1_8:*)
real(8),value::flag
integer(4)::nft_bcv
end
end interface
interface
function nft_div(ctx,ua,va,wa,rhs,iflg) bind(c,name="nft_div")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::ua(1_8:*)
real(8),intent(in)::va(1_8:*)
real(8),intent(in)::wa(1_8:*)
real(8),intent(out)::rhs(1_8:*)
integer(4),value::iflg
integer(4)::nft_div
end
end interface
interface
function nft_precon(ctx,rhs,p,impfl) bind(c,name="nft_precon")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::rhs(1_8:*)
real(8),intent(out)::p(1_8:*)
integer(4),value::impfl
integer(4)::nft_precon
end
end interface
interface
function nft_gcrk(ctx,p,u1,v1,w1,u2,v2,w2,impfl,stats) bind(c,name="nft_gcrk")
import::c_ptr
import::nft_gcr_stats_t
type(c_ptr),value::ctx
real(8),intent(inout)::p(1_8:*)
real(8),intent(in)::u1(1_8:*)
real(8),intent(in)::v1(1_8:*)
real(8),intent(in)::w1(1_8:*)
real(8),intent(out)::u2(1_8:*)
real(8),intent(out)::v2(1_8:*)
real(8),intent(out)::w2(1_8:*)
integer(4),value::impfl
type(nft_gcr_stats_t),intent(out)::stats
integer(4)::nft_gcrk
end
end interface
interface
function nft_mpdatam(ctx,uuu,ld,unk,iflag) bind(c,name="nft_mpdatam")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::uuu(1_8:*)
integer(4),value::ld
real(8),intent(inout)::unk(1_8:*)
integer(4),value::iflag
integer(4)::nft_mpdatam
end
end interface
interface
function nft_state_init(ctx) bind(c,name="nft_state_init")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::nft_state_init
end
end interface
interface
function nft_run_steps(ctx,nsteps,stats) bind(c,name="nft_run_steps")
import::c_ptr
import::nft_run_stats_t
type(c_ptr),value::ctx
integer(4),value::nsteps
type(nft_run_stats_t),intent(out)::stats
integer(4)::nft_run_steps
end
end interface
interface
function nft_save_state(ctx,path) bind(c,name="nft_save_state")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(in)::path(1_8:*)
integer(4)::nft_save_state
end
end interface
interface
function nft_load_state(ctx,path) bind(c,name="nft_load_state")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(in)::path(1_8:*)
integer(4)::nft_load_state
end
end interface
interface
function nft_get_state(ctx,name,host,ld) bind(c,name="nft_get_state")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(in)::name(1_8:*)
real(8),intent(out)::host(1_8:*)
integer(4),value::ld
integer(4)::nft_get_state
end
end interface
interface
function nft_set_state(ctx,name,host,ld) bind(c,name="nft_set_state")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(in)::name(1_8:*)
real(8),intent(in)::host(1_8:*)
integer(4),value::ld
integer(4)::nft_set_state
end
end interface
interface
function nft_diagnostics(ctx,courant_flag,d) bind(c,name="nft_diagnostics")
import::c_ptr
import::nft_diag_t
type(c_ptr),value::ctx
integer(4),value::courant_flag
type(nft_diag_t),intent(out)::d
integer(4)::nft_diagnostics
end
end interface
interface
function nft_get_field(ctx,name,host) bind(c,name="nft_get_field")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(in)::name(1_8:*)
real(8),intent(out)::host(1_8:*)
integer(4)::nft_get_field
end
end interface
interface
function nft_get_gcr_stats(ctx,stats) bind(c,name="nft_get_gcr_stats")
import::c_ptr
import::nft_gcr_stats_t
type(c_ptr),value::ctx
type(nft_gcr_stats_t),intent(out)::stats
integer(4)::nft_get_gcr_stats
end
end interface
contains
subroutine nft_print_diag(d,parts)
type(nft_diag_t),intent(in)::d
integer(4),intent(in)::parts
end
subroutine nft_check(rc,what)
integer(4),intent(in)::rc
character(*,1),intent(in)::what
end
subroutine nft_attach(ncol,col_start,precon_kind,precision,lord,strat,theta0)
integer(4),intent(in)::ncol
integer(4),intent(in),target::col_start(1_8:*)
integer(4),intent(in)::precon_kind
integer(4),intent(in)::precision
integer(4),intent(in)::lord
real(8),intent(in)::strat
real(8),intent(in)::theta0
end
subroutine nft_attach_multigrid(itrp2,itrp_c1,itrp_c2,itrp_cc)
integer(4),intent(in)::itrp2
integer(4),intent(in)::itrp_c1
integer(4),intent(in)::itrp_c2
integer(4),intent(in)::itrp_cc
end
end
