﻿!mod$ v1 sum:17c341a273ea9f76
!need$ 0bde2ac47243ead2 i iso_c_binding
module gatt_hip
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
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
private::__builtin_c_ptr
integer(4),parameter::gatt_deriv=0_4
integer(4),parameter::gatt_filter=1_4
integer(4),parameter::gatt_dist_lid_driven=1_4
integer(4),parameter::gatt_dist_hold_inlet=2_4
integer(4),parameter::gatt_forcing_turbulence=1_4
type,bind(c)::gatt_lineop
integer(4)::n=0_4
integer(4)::ntypes=0_4
integer(4)::nr=0_4
integer(4)::periodic=0_4
type(c_ptr)::a=__builtin_c_ptr(__address=0_8)
type(c_ptr)::b=__builtin_c_ptr(__address=0_8)
type(c_ptr)::c=__builtin_c_ptr(__address=0_8)
type(c_ptr)::d=__builtin_c_ptr(__address=0_8)
type(c_ptr)::r=__builtin_c_ptr(__address=0_8)
end type
type,bind(c)::gatt_box_list
integer(4)::count=0_4
type(c_ptr)::box=__builtin_c_ptr(__address=0_8)
type(c_ptr)::value=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dir=__builtin_c_ptr(__address=0_8)
end type
type,bind(c)::gatt_disturbance
integer(4)::kind=0_4
integer(4)::box(1_8:6_8)=[INTEGER(4)::0_4,0_4,0_4,0_4,0_4,0_4]
type(c_ptr)::x=__builtin_c_ptr(__address=0_8)
real(8)::par(1_8:4_8)=[REAL(8)::0._8,0._8,0._8,0._8]
end type
type,bind(c)::gatt_forcing
integer(4)::kind=0_4
integer(4)::box(1_8:6_8)=[INTEGER(4)::0_4,0_4,0_4,0_4,0_4,0_4]
real(8)::par(1_8:4_8)=[REAL(8)::0._8,0._8,0._8,0._8]
end type
type,bind(c)::gatt_bc_desc
type(gatt_box_list)::dir(1_8:5_8)
type(gatt_box_list)::neu(1_8:5_8)
type(gatt_box_list)::sec(1_8:5_8)
integer(4)::ncorners=0_4
type(c_ptr)::cl=__builtin_c_ptr(__address=0_8)
type(c_ptr)::cd=__builtin_c_ptr(__address=0_8)
type(c_ptr)::cadiabatic=__builtin_c_ptr(__address=0_8)
integer(4)::neumannlength=0_4
integer(4)::neumann2length=0_4
type(c_ptr)::neumanncoeffs=__builtin_c_ptr(__address=0_8)
type(c_ptr)::neumann2coeffs=__builtin_c_ptr(__address=0_8)
integer(4)::ndisturb=0_4
type(c_ptr)::disturb=__builtin_c_ptr(__address=0_8)
end type
type,bind(c)::gatt_problem_desc
integer(4)::nx=0_4
integer(4)::ny=0_4
integer(4)::nz=0_4
real(8)::re=0._8
real(8)::ma=0._8
real(8)::pr=0._8
real(8)::gamma=0._8
real(8)::t0=0._8
type(gatt_lineop)::deriv(1_8:3_8)
type(gatt_lineop)::filter(1_8:3_8)
type(c_ptr)::types(1_8:3_8)=[__builtin_c_ptr::__builtin_c_ptr(__address=0_8),__builtin_c_ptr(__address=0_8),__builtin_c_ptr(__address=0_8)]
integer(4)::filterdir(1_8:3_8)=[INTEGER(4)::0_4,0_4,0_4]
real(8)::filterchartime=-1._8
type(gatt_bc_desc)::bc
integer(4)::nforcing=0_4
type(c_ptr)::forcing=__builtin_c_ptr(__address=0_8)
integer(4)::sfdtype=0_4
real(8)::sfddelta=0._8
real(8)::sfdxval=0._8
type(c_ptr)::sfdx=__builtin_c_ptr(__address=0_8)
real(8)::dxmin(1_8:3_8)=[REAL(8)::0._8,0._8,0._8]
end type
interface
function gatt_last_error() bind(c,name="gatt_last_error")
import::c_ptr
type(c_ptr)::gatt_last_error
end
end interface
interface
function gatt_create(desc,batch,ctx) bind(c,name="gatt_create")
import::c_ptr
import::gatt_problem_desc
type(gatt_problem_desc),intent(in)::desc
integer(4),value::batch
type(c_ptr),intent(out)::ctx
integer(4)::gatt_create
end
end interface
interface
function gatt_destroy(ctx) bind(c,name="gatt_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::gatt_destroy
end
end interface
interface
function gatt_set_state(ctx,member,q) bind(c,name="gatt_set_state")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::member
type(c_ptr),intent(in)::q(1_8:5_8)
integer(4)::gatt_set_state
end
end interface
interface
function gatt_get_state(ctx,member,q) bind(c,name="gatt_get_state")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::member
type(c_ptr),intent(in)::q(1_8:5_8)
integer(4)::gatt_get_state
end
end interface
interface
function gatt_rk4(ctx,t,dt) bind(c,name="gatt_rk4")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::t
real(8),value::dt
integer(4)::gatt_rk4
end
end interface
interface
function gatt_apply_filters(ctx,theta) bind(c,name="gatt_apply_filters")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::theta
integer(4)::gatt_apply_filters
end
end interface
interface
function gatt_apply_bound_cond(ctx,t) bind(c,name="gatt_apply_bound_cond")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::t
integer(4)::gatt_apply_bound_cond
end
end interface
interface
function gatt_sfd(ctx,dt) bind(c,name="gatt_sfd")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::dt
integer(4)::gatt_sfd
end
end interface
interface
function gatt_step(ctx,t,dt,umax,maxchange,diverged) bind(c,name="gatt_step")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::t
real(8),value::dt
real(8),intent(out)::umax(1_8:3_8)
real(8),intent(out)::maxchange(1_8:5_8)
integer(4),intent(out)::diverged
integer(4)::gatt_step
end
end interface
interface
function gatt_advance(ctx,nsteps,t0,dt,maxchange,diverged) bind(c,name="gatt_advance")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nsteps
real(8),value::t0
real(8),value::dt
real(8),intent(out)::maxchange(1_8:5_8)
integer(4),intent(out)::diverged
integer(4)::gatt_advance
end
end interface
contains
function gatt_error_message() result(msg)
character(:,1),allocatable::msg
end
end
