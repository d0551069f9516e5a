! gpu_bridge.f90 -- the module a maintainer adds to HORMONE to run its
! Cartesian hydro step and hyperbolic gravity through libhormone_hip.so.
!
! The reference's main program (src/main/hormone.f90) changes in five
! places: `use gpu_bridge`; `call gpu_init` after the initial `pressure`
! (hormone.f90:100); gpu_timestep, gpu_gravity, gpu_hydro_step and
! gpu_phidamp in place of timestep, gravity, hydro_step and phidamp in the
! main loop (hormone.f90:126,130,136,140); gpu_output in place of output and
! `call gpu_sync` before `call test` (hormone.f90:193).  Everything the
! bridge reads and writes is the reference's own module state (settings,
! grid, physval, gravmod), passed by reference.
!
! The switch is a namelist group of its own in the parameter file,
!   &gpucon use_gpu=.true. fused=2 /
! read like setup.f90:154-165 reads each group (iostat, so a file without
! it leaves the defaults).  With use_gpu = .false. every gpu_* call is the
! reference's own routine, unchanged.
module gpu_bridge
 use iso_c_binding
 use hormone_hip_mod
 implicit none
 private
 public:: gpu_init, gpu_timestep, gpu_gravity, gpu_hydro_step, gpu_phidamp, gpu_output, &
          gpu_sync, use_gpu

 logical,save:: use_gpu = .false.
 integer,save:: fused = 0, sweep_waves = 0, mixed = 0, device = 0
 type(c_ptr),save:: hctx = c_null_ptr

 namelist /gpucon/ use_gpu, fused, sweep_waves, mixed, device

contains

 ! After the initial state is complete (initialcondition, boundarycondition,
 ! conserve, pressure): create the device context and upload the state.
 subroutine gpu_init
  use settings
  use grid
  use physval
  use gravmod,only:grvphi,grvpsi
  use constants,only:G
  type(hg_desc):: desc
  integer:: ui, ios, n
  integer(c_int):: st
  real(8),allocatable:: work(:,:,:)

  open(newunit=ui,file=parafile,status='old',iostat=ios)
  if(ios==0)then
   read(ui,NML=gpucon,iostat=ios)
   close(ui)
  end if
  if(.not.use_gpu)return

  desc%is = is ; desc%ie = ie ; desc%js = js ; desc%je = je ; desc%ks = ks ; desc%ke = ke
  desc%crdnt = crdnt ; desc%eostype = eostype ; desc%compswitch = compswitch
  desc%mag_on = merge(1,0,mag_on) ; desc%rktype = rktype
  select case(trim(flux_limiter))
  case('flat')  ; desc%flux_limiter = HG_LIMITER_FLAT
  case('minmod'); desc%flux_limiter = HG_LIMITER_MINMOD
  case default  ; desc%flux_limiter = HG_LIMITER_MODIFIED_MC
  end select
  desc%bc = [bc1is,bc1os,bc2is,bc2os,bc3is,bc3os,bc1iv,bc1ov,bc2iv,bc2ov,bc3iv,bc3ov]
  desc%gamma = gamma ; desc%courant = courant ; desc%alpha9wave = alpha9wave
  desc%gravswitch = gravswitch ; desc%grktype = grktype
  desc%HGfac = HGfac ; desc%hgcfl = hgcfl ; desc%alphagrv = alphagrv ; desc%G = G
  desc%maxtngrv = maxtngrv ; desc%device = device ; desc%spn = spn
  desc%precision = merge(HG_PREC_FP32MIX, HG_PREC_FP64, mixed/=0)
  st = hg_create_cartesian(desc, hctx, x1, xi1, dx1, dxi1, idx1, idxi1, &
                           x2, xi2, dx2, dxi2, idx2, idxi2, x3, xi3, dx3, dxi3, idx3, idxi3)
  call hg_check(hctx, st, 'hg_create')
  if(fused/=0)call hg_check(hctx, hg_set_fused(hctx, int(fused,c_int32_t)), 'hg_set_fused')
  if(sweep_waves/=0) &
   call hg_check(hctx, hg_set_launch(hctx, int(sweep_waves,c_int32_t)), 'hg_set_launch')

  call hg_check(hctx, hg_upload(hctx, HG_F_D, d), 'upload d')
  call hg_check(hctx, hg_upload(hctx, HG_F_V1, v1), 'upload v1')
  call hg_check(hctx, hg_upload(hctx, HG_F_V2, v2), 'upload v2')
  call hg_check(hctx, hg_upload(hctx, HG_F_V3, v3), 'upload v3')
  call hg_check(hctx, hg_upload(hctx, HG_F_E, e), 'upload e')
  if(mag_on)then
   call hg_check(hctx, hg_upload(hctx, HG_F_B1, b1), 'upload b1')
   call hg_check(hctx, hg_upload(hctx, HG_F_B2, b2), 'upload b2')
   call hg_check(hctx, hg_upload(hctx, HG_F_B3, b3), 'upload b3')
   call hg_check(hctx, hg_upload(hctx, HG_F_PHI, phi), 'upload phi')
  end if
  do n = 1, ufnmax
   call hg_check(hctx, hg_upload(hctx, HG_F_U+n-1, u(:,:,:,n)), 'upload u')
  end do
  do n = 1, merge(spn, 0, compswitch==2)
   call hg_check(hctx, hg_upload(hctx, HG_F_SPC+n-1, spc(n,:,:,:)), 'upload spc')
  end do
  if(gravswitch==3)then
   ! the potential with its boundary layer; psi is stored without ghosts
   call hg_check(hctx, hg_upload(hctx, HG_F_GRVPHI, grvphi), 'upload grvphi')
   allocate(work(is-2:ie+2,js-2:je+2,ks-2:ke+2))
   work = 0d0
   work(is:ie,js:je,ks:ke) = grvpsi
   call hg_check(hctx, hg_upload(hctx, HG_F_GRVPSI, work), 'upload grvpsi')
  end if
 end subroutine gpu_init

 ! timestep (timestep.f90:15-105): dt, ch, cgrav and dtgrav
 subroutine gpu_timestep
  use timestep_mod,only:timestep
  use settings,only:outstyle,endstyle,t_end
  use grid,only:time,t_out,fixed_dt,dt
  use physval,only:ch
  use gravmod,only:cgrav,dtgrav
  if(.not.use_gpu)then
   call timestep
   return
  end if
  call hg_check(hctx, hg_timestep(hctx, time, t_out, t_end, int(outstyle,c_int32_t), &
                                  int(endstyle,c_int32_t), fixed_dt, dt, ch, cgrav, &
                                  dtgrav), 'timestep')
 end subroutine gpu_timestep

 ! gravity (gravity.f90:18-64) for the hyperbolic solver: at tn = 0 the
 ! potential is the initial one and grvtime starts at 0; afterwards
 ! gravity_hyperbolic's sub-steps
 subroutine gpu_gravity
  use gravity_mod,only:gravity
  use settings,only:gravswitch
  use grid,only:time,dt,tn
  use gravmod,only:grvtime,cgrav,cgrav_old,dtgrav
  integer(c_int32_t):: nsub
  if(.not.use_gpu.or.gravswitch/=3)then
   call gravity
   return
  end if
  if(tn==0)then
   grvtime = 0d0
   return
  end if
  call hg_check(hctx, hg_gravity(hctx, time, dt, grvtime, cgrav, cgrav_old, dtgrav, nsub), &
                'gravity')
 end subroutine gpu_gravity

 ! hydro_step (hydro.f90:15-52)
 subroutine gpu_hydro_step
  use hydro_mod,only:hydro_step
  use grid,only:dt
  use physval,only:ch
  if(.not.use_gpu)then
   call hydro_step
   return
  end if
  call hg_check(hctx, hg_hydro_step(hctx, dt, ch), 'hydro_step')
 end subroutine gpu_hydro_step

 ! phidamp (source.f90:239-260); the factor as source.f90:253 evaluates it
 subroutine gpu_phidamp
  use source_mod,only:phidamp
  use settings,only:alpha9wave,courant
  if(.not.use_gpu)then
   call phidamp
   return
  end if
  call hg_check(hctx, hg_phidamp(hctx, exp(-alpha9wave*courant)), 'phidamp')
 end subroutine gpu_phidamp

 ! The device state into the module arrays that output and test read.
 subroutine gpu_sync
  use settings,only:mag_on,gravswitch,compswitch,spn
  use grid,only:is,ie,js,je,ks,ke
  use physval
  use gravmod,only:grvphi,grvpsi
  integer:: n
  real(8),allocatable:: work(:,:,:)
  if(.not.use_gpu)return
  call hg_check(hctx, hg_download(hctx, HG_F_D, d), 'download d')
  call hg_check(hctx, hg_download(hctx, HG_F_V1, v1), 'download v1')
  call hg_check(hctx, hg_download(hctx, HG_F_V2, v2), 'download v2')
  call hg_check(hctx, hg_download(hctx, HG_F_V3, v3), 'download v3')
  call hg_check(hctx, hg_download(hctx, HG_F_E, e), 'download e')
  call hg_check(hctx, hg_download(hctx, HG_F_EINT, eint), 'download eint')
  call hg_check(hctx, hg_download(hctx, HG_F_P, p), 'download p')
  if(mag_on)then
   call hg_check(hctx, hg_download(hctx, HG_F_B1, b1), 'download b1')
   call hg_check(hctx, hg_download(hctx, HG_F_B2, b2), 'download b2')
   call hg_check(hctx, hg_download(hctx, HG_F_B3, b3), 'download b3')
   call hg_check(hctx, hg_download(hctx, HG_F_PHI, phi), 'download phi')
  end if
  do n = 1, merge(spn, 0, compswitch==2)
   call hg_check(hctx, hg_download(hctx, HG_F_SPC+n-1, spc(n,:,:,:)), 'download spc')
  end do
  if(gravswitch==3)then
   call hg_check(hctx, hg_download(hctx, HG_F_GRVPHI, grvphi), 'download grvphi')
   allocate(work(is-2:ie+2,js-2:je+2,ks-2:ke+2))
   call hg_check(hctx, hg_download(hctx, HG_F_GRVPSI, work), 'download grvpsi')
   grvpsi = work(is:ie,js:je,ks:ke)
  end if
 end subroutine gpu_sync

 ! output (output.f90:20-60) of the device state
 subroutine gpu_output
  use output_mod,only:output
  call gpu_sync
  call output
 end subroutine gpu_output

end module gpu_bridge
