! hg_driver.f90 -- a small Fortran host for the MI355X hydro path.
!
! It reads the reference's namelist groups (&gridcon &out_con &eos_con
! &simucon &bouncon, same variable names and meanings as src/main/setup.f90,
! plus its own &gpucon group), builds the uniform Cartesian grid the way
! gridset does (src/main/gridset.f90:43-158), sets the shock-tube or
! Orszag-Tang initial state (src/init/shocktube.f90, src/init/orszagtang.f90)
! or reads an initial state from a file (&gpucon initfile), and runs the main
! loop of hormone.f90:124-179 through the C ABI (hormone_hip_mod): timestep,
! the hyperbolic self-gravity solver when &gravcon gravswitch=3, the hydro
! step and phidamp.  The final state is written in the reference's binary
! snapshot layout (output.f90:697-795), so it can be compared with the
! reference's golden files directly.
!
! usage: hg_driver <parameter file>
program hg_driver
 use iso_c_binding
 use hormone_hip_mod
 implicit none

 ! &gridcon
 real(8):: xi1s=0d0, xi1e=1d0, xi2s=0d0, xi2e=1d0, xi3s=0d0, xi3e=1d0
 real(8):: x1min=0d0, x2min=0d0, x3min=0d0
 integer:: is=1, ie=10, js=1, je=1, ks=1, ke=1, imesh=0, jmesh=0, kmesh=0
 integer:: fmr_max=0, fmr_lvl(1:20)=0
 ! &out_con
 character(len=30):: outdir='data', dt_unit='s'
 integer:: outstyle=1, endstyle=1, tnlim=0, tn_out=1, tn_evo=10, sigfig=7, outres=1
 real(8):: t_end=0d0, dt_out=1d0
 logical:: output_ascii=.true., write_other_vel=.false., write_shock=.false., &
           write_evo=.false., write_other_slice=.false., write_temp=.false., &
           write_mc=.false.
 ! &eos_con
 integer:: eostype=0, compswitch=0, spn=0
 real(8):: eoserr=1d-15, muconst=0.62d0, gamma=1.6666666666666667d0
 real(8):: X_uniform=0.7d0, Y_uniform=0.28d0, Z_uniform=0.02d0
 ! &simucon
 integer:: crdnt=0, rktype=3, frame=0, smeartype=1
 real(8):: courant=0.9d0, alpha9wave=0.1d0, fixed_dt=-1d0
 logical:: mag_on=.false., include_cooling=.false., include_extforce=.false., &
           solve_hydro=.true., solve_grav=.true.
 character(len=30):: flux_limiter='modified_mc', extrasfile='extras'
 ! &bouncon
 integer:: bc1is=0, bc1os=0, bc2is=0, bc2os=0, bc3is=0, bc3os=0
 integer:: bc1iv=0, bc1ov=0, bc2iv=0, bc2ov=0, bc3iv=0, bc3ov=0
 logical:: eq_sym=.false.
 ! &gravcon (the variables this path reads; the rest are accepted and ignored)
 integer:: gravswitch=0, grvsrctype=0, maxtngrv=-1, gbtype=0, grktype=2
 integer:: gis=1, gie=1, gjs=1, gje=1, gks=1, gke=1
 real(8):: grverr=1d-15, cgerr=1d-15, HGfac=5d0, hgcfl=0.9d0, alphagrv=0.1d0
 logical:: grav_init_other=.false., grav_init_relax=.false., include_extgrv=.false.
 ! &gpucon (this driver only)
 character(len=40):: simtype='orszagtang_xy'
 character(len=256):: outfile='final.bin', initfile=''
 logical:: is_test=.true.
 integer:: device=0
 integer:: fused=0             ! hg_set_fused: 0 off, 1 last direction, 2 x
 real(8):: gconst=6.67430d-8   ! G (a parameter of the reference, G=1 for test set-ups)

 namelist /gridcon/ xi1s,xi1e,xi2s,xi2e,xi3s,xi3e,is,ie,js,je,ks,ke, &
                    imesh,jmesh,kmesh,x1min,x2min,x3min,fmr_max,fmr_lvl
 namelist /out_con/ outdir,outstyle,endstyle,tnlim,t_end,dt_out,tn_out,tn_evo, &
                    dt_unit,sigfig,outres,output_ascii,write_other_vel,write_shock, &
                    write_evo,write_other_slice,write_temp,write_mc
 namelist /eos_con/ eostype,eoserr,compswitch,muconst,spn,gamma, &
                    X_uniform,Y_uniform,Z_uniform
 namelist /simucon/ crdnt,courant,rktype,mag_on,flux_limiter,alpha9wave, &
                    include_cooling,include_extforce,frame,extrasfile, &
                    solve_hydro,solve_grav,fixed_dt,smeartype
 namelist /bouncon/ bc1is,bc1os,bc2is,bc2os,bc3is,bc3os, &
                    bc1iv,bc1ov,bc2iv,bc2ov,bc3iv,bc3ov,eq_sym
 namelist /gravcon/ gravswitch,grvsrctype,grverr,cgerr,HGfac,hgcfl,maxtngrv,gbtype, &
                    grktype,alphagrv,grav_init_other,grav_init_relax,include_extgrv, &
                    gis,gie,gjs,gje,gks,gke
 namelist /gpucon/ simtype,outfile,initfile,is_test,device,gconst,fused

 real(8),allocatable:: x1(:),xi1(:),dx1(:),dxi1(:),idx1(:),idxi1(:)
 real(8),allocatable:: x2(:),xi2(:),dx2(:),dxi2(:),idx2(:),idxi2(:)
 real(8),allocatable:: x3(:),xi3(:),dx3(:),dxi3(:),idx3(:),idxi3(:)
 real(8),allocatable:: d(:,:,:),p(:,:,:),v1(:,:,:),v2(:,:,:),v3(:,:,:), &
                       e(:,:,:),b1(:,:,:),b2(:,:,:),b3(:,:,:),phi(:,:,:), &
                       eint(:,:,:),u(:,:,:,:),grvphi(:,:,:),grvpsi(:,:,:)
 type(hg_desc):: desc
 type(c_ptr):: ctx
 integer(c_int):: st
 integer:: i,j,k,n,ufnmax,dim,tn,ui
 logical:: solve_i,solve_j,solve_k
 real(8):: time,dt,ch,cgrav,dtgrav,t_out,unit_sec,damp,pi,grvtime,cgrav_old
 integer(c_int32_t):: nsub
 character(len=256):: parafile

 call get_command_argument(1, parafile)
 if(len_trim(parafile)==0) stop 'usage: hg_driver <parameter file>'
 call read_groups(trim(parafile))

 ! checksetup.f90:41-75,126-136,181-208 (the parts this path needs)
 solve_i = ie>is ; solve_j = je>js ; solve_k = ke>ks
 dim = count([solve_i,solve_j,solve_k])
 if(.not.solve_i)then ; bc1is=2 ; bc1os=2 ; bc1iv=2 ; bc1ov=2 ; end if
 if(.not.solve_j)then ; bc2is=3 ; bc2os=3 ; bc2iv=3 ; bc2ov=3 ; end if
 if(.not.solve_k)then ; bc3is=3 ; bc3os=3 ; bc3iv=3 ; bc3ov=3 ; end if
 ufnmax = 5
 if(mag_on)then
  ufnmax = 8
  if(dim>1) ufnmax = 9
 end if
 select case(trim(dt_unit))
 case('yr') ; unit_sec = 3600d0*24d0*365.25d0
 case('d')  ; unit_sec = 3600*24d0
 case('hr') ; unit_sec = 3600d0
 case('ks') ; unit_sec = 1000d0
 case('min'); unit_sec = 60d0
 case('s')  ; unit_sec = 1d0
 case('ms') ; unit_sec = 1d-3
 case('ns') ; unit_sec = 1d-9
 case default ; stop 'unknown dt_unit'
 end select
 dt_out = dt_out*unit_sec
 t_end = t_end*unit_sec
 t_out = min(dt_out,t_end)
 if(is_test)then
  t_out = t_end
  tn_out = tnlim
 end if

 call make_axis(xi1s,xi1e,is,ie,x1,xi1,dx1,dxi1,idx1,idxi1)
 call make_axis(xi2s,xi2e,js,je,x2,xi2,dx2,dxi2,idx2,idxi2)
 call make_axis(xi3s,xi3e,ks,ke,x3,xi3,dx3,dxi3,idx3,idxi3)

 allocate(d(is-2:ie+2,js-2:je+2,ks-2:ke+2))
 allocate(p,v1,v2,v3,e,b1,b2,b3,phi,eint,mold=d)
 allocate(u(is-2:ie+2,js-2:je+2,ks-2:ke+2,1:ufnmax))
 d=0d0; p=0d0; v1=0d0; v2=0d0; v3=0d0; e=0d0; b1=0d0; b2=0d0; b3=0d0; phi=0d0
 u=0d0
 allocate(grvphi,grvpsi,mold=d)
 grvphi=0d0; grvpsi=0d0

 pi = acos(-1d0)
 if(len_trim(initfile)>0)then
  call read_state(trim(initfile))
 else
  call initial_state
 end if
 ! initialcondition.f90:133-150, eostype 0
 do k=ks,ke ; do j=js,je ; do i=is,ie
  eint(i,j,k) = p(i,j,k)/(gamma-1d0)
  e(i,j,k) = eint(i,j,k) &
           + d(i,j,k)*0.5d0*( v1(i,j,k)**2 + v2(i,j,k)**2 + v3(i,j,k)**2 ) &
           +          0.5d0*( b1(i,j,k)**2 + b2(i,j,k)**2 + b3(i,j,k)**2 )
 end do ; end do ; end do
 ! conserve.f90:13-49
 u(:,:,:,1) = d ; u(:,:,:,2) = d*v1 ; u(:,:,:,3) = d*v2 ; u(:,:,:,4) = d*v3
 u(:,:,:,5) = e
 if(mag_on)then
  u(:,:,:,6) = b1 ; u(:,:,:,7) = b2 ; u(:,:,:,8) = b3
  if(dim>1) u(:,:,:,9) = phi
 end if

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
 desc%HGfac = HGfac ; desc%hgcfl = hgcfl ; desc%alphagrv = alphagrv ; desc%G = gconst
 desc%maxtngrv = maxtngrv ; desc%precision = HG_PREC_FP64 ; desc%device = device
 desc%spn = spn

 st = hg_create_cartesian(desc, ctx, x1, xi1, dx1, dxi1, idx1, idxi1, &
                          x2, xi2, dx2, dxi2, idx2, idxi2, x3, xi3, dx3, dxi3, idx3, idxi3)
 call hg_check(ctx, st, 'hg_create')
 if(fused/=0)then
  st = hg_set_fused(ctx, int(fused, c_int32_t)) ; call hg_check(ctx, st, 'hg_set_fused')
 end if
 st = hg_upload(ctx, HG_F_D, d)   ; call hg_check(ctx, st, 'upload')
 st = hg_upload(ctx, HG_F_V1, v1) ; call hg_check(ctx, st, 'upload')
 st = hg_upload(ctx, HG_F_V2, v2) ; call hg_check(ctx, st, 'upload')
 st = hg_upload(ctx, HG_F_V3, v3) ; call hg_check(ctx, st, 'upload')
 st = hg_upload(ctx, HG_F_E, e)   ; call hg_check(ctx, st, 'upload')
 if(mag_on)then
  st = hg_upload(ctx, HG_F_B1, b1) ; call hg_check(ctx, st, 'upload')
  st = hg_upload(ctx, HG_F_B2, b2) ; call hg_check(ctx, st, 'upload')
  st = hg_upload(ctx, HG_F_B3, b3) ; call hg_check(ctx, st, 'upload')
  st = hg_upload(ctx, HG_F_PHI, phi) ; call hg_check(ctx, st, 'upload')
 end if
 do n = 1, ufnmax
  st = hg_upload(ctx, HG_F_U+n-1, u(:,:,:,n))
  call hg_check(ctx, st, 'upload')
 end do
 if(gravswitch==3)then
  st = hg_upload(ctx, HG_F_GRVPHI, grvphi) ; call hg_check(ctx, st, 'upload')
  st = hg_upload(ctx, HG_F_GRVPSI, grvpsi) ; call hg_check(ctx, st, 'upload')
 end if

 ! main loop, hormone.f90:124-179 (hydro only)
 damp = exp(-alpha9wave*courant)
 time = 0d0 ; tn = 0
 grvtime = 0d0 ; cgrav_old = 0d0   ! gravmod values at start-up
 main_loop: do
  st = hg_timestep(ctx, time, t_out, t_end, outstyle, endstyle, fixed_dt, &
                   dt, ch, cgrav, dtgrav)
  call hg_check(ctx, st, 'hg_timestep')
  ! gravity.f90:41-60: the initial potential at tn = 0, the hyperbolic
  ! solver afterwards
  if(gravswitch==3.and.tn>0)then
   st = hg_gravity(ctx, time, dt, grvtime, cgrav, cgrav_old, dtgrav, nsub)
   call hg_check(ctx, st, 'hg_gravity')
  end if
  if(solve_hydro)then
   st = hg_hydro_step(ctx, dt, ch)
   call hg_check(ctx, st, 'hg_hydro_step')
  end if
  if(mag_on.and.dim>1)then
   st = hg_phidamp(ctx, damp)
   call hg_check(ctx, st, 'hg_phidamp')
  end if
  time = time + dt ; tn = tn + 1
  select case (endstyle)
  case(1)
   if(time>=t_end)exit main_loop
  case(2)
   if(tn>=tnlim)exit main_loop
  end select
  select case(outstyle)
  case(1)
   if(time>=t_out) t_out = t_out + dt_out
  end select
 end do main_loop

 st = hg_download(ctx, HG_F_D, d)   ; call hg_check(ctx, st, 'download')
 st = hg_download(ctx, HG_F_V1, v1) ; call hg_check(ctx, st, 'download')
 st = hg_download(ctx, HG_F_V2, v2) ; call hg_check(ctx, st, 'download')
 st = hg_download(ctx, HG_F_V3, v3) ; call hg_check(ctx, st, 'download')
 st = hg_download(ctx, HG_F_E, e)   ; call hg_check(ctx, st, 'download')
 if(mag_on)then
  st = hg_download(ctx, HG_F_B1, b1) ; call hg_check(ctx, st, 'download')
  st = hg_download(ctx, HG_F_B2, b2) ; call hg_check(ctx, st, 'download')
  st = hg_download(ctx, HG_F_B3, b3) ; call hg_check(ctx, st, 'download')
  st = hg_download(ctx, HG_F_PHI, phi) ; call hg_check(ctx, st, 'download')
 end if
 st = hg_destroy(ctx)

 ! snapshot in the reference's layout (legacy record markers)
 open(newunit=ui, file=trim(outfile), form='unformatted', access='stream', status='replace')
 write(ui) -1, tn, time, -1
 write(ui) -1, d(is:ie,js:je,ks:ke), v1(is:ie,js:je,ks:ke), v2(is:ie,js:je,ks:ke), &
           v3(is:ie,js:je,ks:ke), e(is:ie,js:je,ks:ke), -1
 if(mag_on) write(ui) -1, b1(is:ie,js:je,ks:ke), b2(is:ie,js:je,ks:ke), &
                      b3(is:ie,js:je,ks:ke), phi(is:ie,js:je,ks:ke), -1
 close(ui)
 print'(a,i8,a,es23.16)', 'hg_driver: tn =', tn, '  time =', time

contains

 subroutine read_groups(fname)
  character(len=*),intent(in):: fname
  integer:: u_, ios
  open(newunit=u_, file=fname, status='old', action='read')
  read(u_, nml=gridcon, iostat=ios) ; rewind(u_)
  read(u_, nml=out_con, iostat=ios) ; rewind(u_)
  read(u_, nml=eos_con, iostat=ios) ; rewind(u_)
  read(u_, nml=simucon, iostat=ios) ; rewind(u_)
  read(u_, nml=bouncon, iostat=ios) ; rewind(u_)
  read(u_, nml=gravcon, iostat=ios) ; rewind(u_)
  read(u_, nml=gpucon, iostat=ios)
  close(u_)
 end subroutine read_groups

 ! Initial state from a stream file: d, p, v1, v2, v3, b1, b2, b3 over the
 ! interior (x fastest), then, for gravswitch 3, the potential over the
 ! padded range is-2:ie+2 (its outer layer is the boundary value).
 subroutine read_state(fname)
  character(len=*),intent(in):: fname
  integer:: u_
  open(newunit=u_, file=fname, form='unformatted', access='stream', status='old', &
       action='read')
  read(u_) d(is:ie,js:je,ks:ke), p(is:ie,js:je,ks:ke), v1(is:ie,js:je,ks:ke), &
           v2(is:ie,js:je,ks:ke), v3(is:ie,js:je,ks:ke), b1(is:ie,js:je,ks:ke), &
           b2(is:ie,js:je,ks:ke), b3(is:ie,js:je,ks:ke)
  if(gravswitch==3) read(u_) grvphi
  close(u_)
 end subroutine read_state

 ! uniform Cartesian axis with running sums, gridset.f90:50-66
 subroutine make_axis(xs,xe,s,e_,x,xi,dx,dxi,idx,idxi)
  real(8),intent(in):: xs,xe
  integer,intent(in):: s,e_
  real(8),allocatable,intent(out):: x(:),xi(:),dx(:),dxi(:),idx(:),idxi(:)
  integer:: q
  allocate(x(s-2:e_+2)) ; allocate(xi,dx,dxi,idx,idxi,mold=x)
  dxi = (xe-xs)/dble(e_-s+1)
  dx(s-2) = dxi(s-2) ; idxi(s-2) = 1d0/dxi(s-2) ; idx(s-2) = 1d0/dx(s-2)
  do q = s-1, e_+2
   dx(q)   = 0.5d0*(dxi(q-1)+dxi(q))
   idxi(q) = 1d0/dxi(q)
   idx(q)  = 1d0/dx(q)
  end do
  x(s-1)  = xs - dxi(s-1)*0.5d0
  x(s-2)  = x(s-1) - dx(s-1)
  xi(s-1) = xs
  xi(s-2) = xs - dxi(s-1)
  do q = s, e_+2
   x(q)  = x(q-1)  + dx(q)
   xi(q) = xi(q-1) + dxi(q)
  end do
 end subroutine make_axis

 subroutine initial_state
  integer:: dirn, l
  real(8):: dl_,dr_,pl_,pr_,b(6)
  l = len_trim(simtype)
  select case(simtype(1:l-2))
  case('sodshock','briowushock')
   dl_ = 1d0 ; dr_ = 0.125d0 ; pl_ = 1d0 ; pr_ = 0.1d0 ; b = 0d0
   select case(simtype(l:l))
   case('x') ; dirn = 1
   case('y') ; dirn = 2
   case default ; dirn = 3
   end select
   if(simtype(1:l-2)=='briowushock')then
    select case(dirn)
    case(1) ; b = [0.75d0,0.75d0,1d0,-1d0,0d0,0d0]
    case(2) ; b = [0d0,0d0,0.75d0,0.75d0,1d0,-1d0]
    case(3) ; b = [1d0,-1d0,0d0,0d0,0.75d0,0.75d0]
    end select
   end if
   do k=ks,ke ; do j=js,je ; do i=is,ie
    if(left(dirn,i,j,k))then
     d(i,j,k)=dl_ ; p(i,j,k)=pl_ ; b1(i,j,k)=b(1) ; b2(i,j,k)=b(3) ; b3(i,j,k)=b(5)
    else
     d(i,j,k)=dr_ ; p(i,j,k)=pr_ ; b1(i,j,k)=b(2) ; b2(i,j,k)=b(4) ; b3(i,j,k)=b(6)
    end if
   end do ; end do ; end do
  case default
   if(simtype(1:11)/='orszagtang_') stop 'hg_driver: unknown simtype'
   select case(simtype(12:13))
   case('xy')
    do j=js,je ; do i=is,ie
     d(i,j,ks) = gamma**2/(4.d0*pi) ; p(i,j,ks) = gamma/(4.d0*pi)
     v1(i,j,ks) = -sin(2.d0*pi*x2(j)) ; v2(i,j,ks) = sin(2.d0*pi*x1(i))
     b1(i,j,ks) = -sin(2.d0*pi*x2(j))/sqrt(4.d0*pi)
     b2(i,j,ks) =  sin(4.d0*pi*x1(i))/sqrt(4.d0*pi)
    end do ; end do
   case('xz')
    do k=ks,ke ; do i=is,ie
     d(i,js,k) = gamma**2/(4.d0*pi) ; p(i,js,k) = gamma/(4.d0*pi)
     v1(i,js,k) = -sin(2.d0*pi*x3(k)) ; v3(i,js,k) = sin(2.d0*pi*x1(i))
     b1(i,js,k) = -sin(2.d0*pi*x3(k))/sqrt(4.d0*pi)
     b3(i,js,k) =  sin(4.d0*pi*x1(i))/sqrt(4.d0*pi)
    end do ; end do
   case default
    do k=ks,ke ; do j=js,je
     d(is,j,k) = gamma**2/(4.d0*pi) ; p(is,j,k) = gamma/(4.d0*pi)
     v2(is,j,k) = -sin(2.d0*pi*x3(k)) ; v3(is,j,k) = sin(2.d0*pi*x2(j))
     b2(is,j,k) = -sin(2.d0*pi*x3(k))/sqrt(4.d0*pi)
     b3(is,j,k) =  sin(4.d0*pi*x2(j))/sqrt(4.d0*pi)
    end do ; end do
   end select
  end select
 end subroutine initial_state

 logical function left(dirn,i,j,k)
  integer,intent(in):: dirn,i,j,k
  select case(dirn)
  case(1) ; left = i <= (is+ie)/2
  case(2) ; left = j <= (js+je)/2
  case default ; left = k <= (ks+ke)/2
  end select
 end function left

end program hg_driver
