! hormone_hip_mod.f90 -- iso_c_binding interface to the C ABI declared in
! include/hormone_hip.h (libhormone_hip.so).  A Fortran host (the reference's
! main program, or hg_driver.f90 here) uses this module to hand its
! Cartesian hydro sweep to the GPU: hg_timestep replaces timestep
! (src/main/timestep.f90:15-105), hg_gravity replaces gravity for the
! hyperbolic solver (src/main/gravity.f90:18-64), hg_hydro_step replaces
! hydro_step (src/main/hydro.f90:15-52) and hg_phidamp replaces phidamp
! (src/main/source.f90:239-260); hg_run runs the hydro-only loop on the
! device, hg_run_gravity the loop with hyperbolic gravity.  Field arrays are
! passed by reference (the reference's allocatable module arrays need no
! TARGET attribute); an array section such as u(:,:,:,n) or spc(n,:,:,:)
! is passed as a contiguous copy where it is not contiguous itself.
module hormone_hip_mod
 use iso_c_binding
 implicit none

 integer(c_int),parameter:: HG_OK=0, HG_ERR_ARG=1, HG_ERR_UNSUPPORTED=2, &
                            HG_ERR_DEVICE=3, HG_ERR_FACE_EOS=4, HG_ERR_NEG_EINT=5, &
                            HG_ERR_NOMEM=6, HG_ERR_STATE=7
 integer(c_int),parameter:: HG_LIMITER_FLAT=0, HG_LIMITER_MINMOD=1, &
                            HG_LIMITER_MODIFIED_MC=2
 integer(c_int),parameter:: HG_PREC_FP64=0, HG_PREC_FP32MIX=1
 integer(c_int),parameter:: HG_F_D=0, HG_F_V1=1, HG_F_V2=2, HG_F_V3=3, HG_F_E=4, &
                            HG_F_B1=5, HG_F_B2=6, HG_F_B3=7, HG_F_PHI=8, &
                            HG_F_EINT=9, HG_F_P=10, HG_F_GRVPHI=11, HG_F_GRVPSI=12, &
                            HG_F_U=16, HG_F_SPC=32

 ! mirrors struct hg_desc field by field
 type,bind(C):: hg_desc
  integer(c_int32_t):: is, ie, js, je, ks, ke
  type(c_ptr):: x1, xi1, dx1, dxi1, idx1, idxi1
  type(c_ptr):: x2, xi2, dx2, dxi2, idx2, idxi2
  type(c_ptr):: x3, xi3, dx3, dxi3, idx3, idxi3
  integer(c_int32_t):: crdnt, eostype, compswitch, mag_on, flux_limiter, rktype
  integer(c_int32_t):: bc(12)
  real(c_double):: gamma, courant, alpha9wave
  integer(c_int32_t):: gravswitch, grktype
  real(c_double):: HGfac, hgcfl, alphagrv, G
  integer(c_int32_t):: maxtngrv, precision, device, spn
 end type hg_desc

 interface
  function hg_create(desc, ctx) bind(C, name='hg_create') result(st)
   import:: c_int, c_ptr, hg_desc
   type(hg_desc),intent(in):: desc
   type(c_ptr),intent(out):: ctx
   integer(c_int):: st
  end function hg_create

  function hg_destroy(ctx) bind(C, name='hg_destroy') result(st)
   import:: c_int, c_ptr
   type(c_ptr),value:: ctx
   integer(c_int):: st
  end function hg_destroy

  function hg_ufnmax(ctx) bind(C, name='hg_ufnmax') result(n)
   import:: c_int, c_ptr
   type(c_ptr),value:: ctx
   integer(c_int):: n
  end function hg_ufnmax

  function hg_upload(ctx, field, padded) bind(C, name='hg_upload') result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: field
   real(c_double),intent(in):: padded(*)
   integer(c_int):: st
  end function hg_upload

  function hg_download(ctx, field, padded) bind(C, name='hg_download') result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: field
   real(c_double),intent(inout):: padded(*)
   integer(c_int):: st
  end function hg_download

  function hg_timestep(ctx, time, t_out, t_end, outstyle, endstyle, fixed_dt, &
                       dt, ch, cgrav, dtgrav) bind(C, name='hg_timestep') result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),value:: time, t_out, t_end, fixed_dt
   integer(c_int32_t),value:: outstyle, endstyle
   real(c_double),intent(out):: dt, ch, cgrav, dtgrav
   integer(c_int):: st
  end function hg_timestep

  function hg_hydro_step(ctx, dt, ch) bind(C, name='hg_hydro_step') result(st)
   import:: c_int, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),value:: dt, ch
   integer(c_int):: st
  end function hg_hydro_step

  function hg_phidamp(ctx, factor) bind(C, name='hg_phidamp') result(st)
   import:: c_int, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),value:: factor
   integer(c_int):: st
  end function hg_phidamp

  ! replaces gravity (gravity.f90:18-64 -> gravity_hyperbolic.f90:19-50) for
  ! gravswitch 3 at tn > 0; cgrav and dtgrav come from hg_timestep
  function hg_gravity(ctx, time, dt, grvtime, cgrav, cgrav_old, dtgrav, nsub) &
                      bind(C, name='hg_gravity') result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),value:: time, dt, cgrav
   real(c_double),intent(inout):: grvtime, cgrav_old, dtgrav
   integer(c_int32_t),intent(out):: nsub
   integer(c_int):: st
  end function hg_gravity

  ! the main loop with hyperbolic gravity on the device (hormone.f90:124-179
  ! with gravswitch 3); grvtime and cgrav_old are the gravmod values
  function hg_run_gravity(ctx, time, tn, t_out, dt_out, t_end, outstyle, endstyle, tnlim, &
                          tn_out, max_steps, damp_factor, grvtime, cgrav_old, nsub_total, &
                          stop_reason) bind(C, name='hg_run_gravity') result(st)
   import:: c_int, c_int32_t, c_int64_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),intent(inout):: time, t_out, grvtime, cgrav_old
   integer(c_int64_t),intent(inout):: tn
   real(c_double),value:: dt_out, t_end, damp_factor
   integer(c_int32_t),value:: outstyle, endstyle
   integer(c_int64_t),value:: tnlim, tn_out, max_steps
   integer(c_int64_t),intent(out):: nsub_total
   integer(c_int32_t),intent(out):: stop_reason
   integer(c_int):: st
  end function hg_run_gravity

  ! hydro-only main loop on the device (hormone.f90:124-179 without gravity)
  function hg_run(ctx, time, tn, t_out, dt_out, t_end, outstyle, endstyle, tnlim, &
                  tn_out, max_steps, damp_factor, stop_reason) &
                  bind(C, name='hg_run') result(st)
   import:: c_int, c_int32_t, c_int64_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   real(c_double),intent(inout):: time, t_out
   integer(c_int64_t),intent(inout):: tn
   real(c_double),value:: dt_out, t_end, damp_factor
   integer(c_int32_t),value:: outstyle, endstyle
   integer(c_int64_t),value:: tnlim, tn_out, max_steps
   integer(c_int32_t),intent(out):: stop_reason
   integer(c_int):: st
  end function hg_run

  ! sweep launch shape (wavefronts per block: 1, 2, 4, 8 or 16)
  function hg_set_launch(ctx, sweep_block_waves) bind(C, name='hg_set_launch') result(st)
   import:: c_int, c_int32_t, c_ptr
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: sweep_block_waves
   integer(c_int):: st
  end function hg_set_launch

  ! one sweep direction fused with the RK update: 0 off, 1 the last solved
  ! direction, 2 x (bit-identical; a launch-shape choice)
  function hg_set_fused(ctx, on) bind(C, name='hg_set_fused') result(st)
   import:: c_int, c_int32_t, c_ptr
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: on
   integer(c_int):: st
  end function hg_set_fused

  ! ghost values of the faces with codes 4, 5, 9, 10 (for Dirichlet faces:
  ! the d0, p0, v10.. arrays, after each dirichletbound)
  function hg_set_ghosts(ctx, field, padded) bind(C, name='hg_set_ghosts') result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: field
   real(c_double),intent(in):: padded(*)
   integer(c_int):: st
  end function hg_set_ghosts

  ! flux of a flux-boundary face (face 0..5 = x1i x1o x2i x2o x3i x3o;
  ! ufnmax values in conserved-variable order)
  function hg_set_boundary_flux(ctx, face, flux) bind(C, name='hg_set_boundary_flux') &
           result(st)
   import:: c_int, c_int32_t, c_ptr, c_double
   type(c_ptr),value:: ctx
   integer(c_int32_t),value:: face
   real(c_double),intent(in):: flux(*)
   integer(c_int):: st
  end function hg_set_boundary_flux

  function hg_last_error(ctx, i, j, k, dir) bind(C, name='hg_last_error') result(st)
   import:: c_int, c_int32_t, c_ptr
   type(c_ptr),value:: ctx
   integer(c_int32_t),intent(out):: i, j, k, dir
   integer(c_int):: st
  end function hg_last_error
 end interface

contains

 ! hg_create for the Cartesian grid module's coordinate arrays: the caller
 ! passes x1(is-2), xi1(is-2), ... (or the whole arrays, which start there)
 ! by reference, so its arrays need no TARGET attribute; the library copies
 ! them during hg_create and keeps no pointer to them.
 function hg_create_cartesian(desc, ctx, x1, xi1, dx1, dxi1, idx1, idxi1, &
                              x2, xi2, dx2, dxi2, idx2, idxi2, &
                              x3, xi3, dx3, dxi3, idx3, idxi3) result(st)
  type(hg_desc),intent(inout):: desc
  type(c_ptr),intent(out):: ctx
  real(c_double),target,intent(in):: x1(*), xi1(*), dx1(*), dxi1(*), idx1(*), idxi1(*)
  real(c_double),target,intent(in):: x2(*), xi2(*), dx2(*), dxi2(*), idx2(*), idxi2(*)
  real(c_double),target,intent(in):: x3(*), xi3(*), dx3(*), dxi3(*), idx3(*), idxi3(*)
  integer(c_int):: st
  desc%x1 = c_loc(x1) ; desc%xi1 = c_loc(xi1) ; desc%dx1 = c_loc(dx1)
  desc%dxi1 = c_loc(dxi1) ; desc%idx1 = c_loc(idx1) ; desc%idxi1 = c_loc(idxi1)
  desc%x2 = c_loc(x2) ; desc%xi2 = c_loc(xi2) ; desc%dx2 = c_loc(dx2)
  desc%dxi2 = c_loc(dxi2) ; desc%idx2 = c_loc(idx2) ; desc%idxi2 = c_loc(idxi2)
  desc%x3 = c_loc(x3) ; desc%xi3 = c_loc(xi3) ; desc%dx3 = c_loc(dx3)
  desc%dxi3 = c_loc(dxi3) ; desc%idx3 = c_loc(idx3) ; desc%idxi3 = c_loc(idxi3)
  st = hg_create(desc, ctx)
  ! the addresses are not valid past this call
  desc%x1 = c_null_ptr ; desc%xi1 = c_null_ptr ; desc%dx1 = c_null_ptr
  desc%dxi1 = c_null_ptr ; desc%idx1 = c_null_ptr ; desc%idxi1 = c_null_ptr
  desc%x2 = c_null_ptr ; desc%xi2 = c_null_ptr ; desc%dx2 = c_null_ptr
  desc%dxi2 = c_null_ptr ; desc%idx2 = c_null_ptr ; desc%idxi2 = c_null_ptr
  desc%x3 = c_null_ptr ; desc%xi3 = c_null_ptr ; desc%dx3 = c_null_ptr
  desc%dxi3 = c_null_ptr ; desc%idx3 = c_null_ptr ; desc%idxi3 = c_null_ptr
 end function hg_create_cartesian

 ! Print the reference's message for a hot-path error and stop, as the
 ! reference does at numflux.f90:579-593 and eos.f90:833-839.
 subroutine hg_check(ctx, st, where)
  type(c_ptr),intent(in):: ctx
  integer(c_int),intent(in):: st
  character(len=*),intent(in):: where
  integer(c_int32_t):: i, j, k, dir
  integer(c_int):: s2
  if(st==HG_OK)return
  s2 = hg_last_error(ctx, i, j, k, dir)
  select case(st)
  case(HG_ERR_FACE_EOS)
   print*,'Error in calculating pressure'
   print'("i=",i5,", j=",i5,", k=",i5)',i,j,k
   print'("In numflux",i2)',dir
  case(HG_ERR_NEG_EINT)
   print*,'Error in internalenergy, Negative internal energy'
   print'("i=",i4,", j=",i4,", k=",i4)',i,j,k
  case default
   print*,'hormone_hip error ',st,' in ',where
  end select
  stop 1
 end subroutine hg_check

end module hormone_hip_mod
