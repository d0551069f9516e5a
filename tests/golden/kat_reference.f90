! Known-answer vectors from the reference's own compiled routines.
!
! Linked against the object files of the reference build (oracle/build_ref.py
! -> oracle/_ref/obj) by tests/golden/make_kat.py, which writes the inputs,
! runs this program once per routine and stores inputs and outputs as the
! fixture tests/golden/kat_reference.npz.  Only the build container runs it.
!
! usage: kat_reference KIND IN OUT
!   IN : stream file, int32 n, int32 nin, then n*nin real(8) (record-major)
!   OUT: stream file, int32 n, int32 nout, then n*nout real(8)
! KIND (routine, inputs -> outputs):
!   hlld    hlldflux (hlldflux.f90:19-179), its 25 arguments in order
!           cfl..ch,erl,err -> fflux(1:10)
!   cf      get_cf (eos.f90:351-365): d cs b1 b2 b3 -> cf
!   eos     eos_p_cs with erad= (eos.f90:195-242, eostype 0):
!           d eint gamma erad -> p cs ierr   (p, cs undefined when ierr=1)
!   eint    get_eint (eos.f90:846-865): etot d v1 v2 v3 b1 b2 b3 mag_on
!           -> eint ierr
!   mc      fluxlimiter with flux_limiter='modified_mc' (fluxlimiter.f90:
!           18-38,58-75): u(1:3) x(1:3) xi(1:2) -> du
!   minmod  fluxlimiter with flux_limiter='minmod' (fluxlimiter.f90:40-56):
!           u(1:3) dx(1:2) -> du
program kat_reference
 use settings,only:eostype,mag_on,flux_limiter
 use physval,only:gamma
 use hllflux_mod,only:hlldflux
 use pressure_mod,only:eos_p_cs,get_cf,get_eint
 use fluxlimiter_mod,only:fluxlimiter
 implicit none
 character(len=16):: kind
 character(len=512):: fin,fout
 integer(4):: n,nin,nout,q,ierr
 real(8),allocatable:: a(:,:),o(:,:)
 real(8):: fl(10),T,imu,p,cs,du,uu(3),dx(2),x(3),xi(2)

 call get_command_argument(1,kind)
 call get_command_argument(2,fin)
 call get_command_argument(3,fout)
 open(10,file=trim(fin),access='stream',form='unformatted',status='old')
 read(10) n,nin
 allocate(a(nin,n))
 read(10) a
 close(10)
 eostype = 0
 select case(trim(kind))
 case('hlld');   nout = 10
 case('cf');     nout = 1
 case('eos');    nout = 3
 case('eint');   nout = 2
 case('mc');     nout = 1
 case('minmod'); nout = 1
 case default
  stop 'unknown kind'
 end select
 allocate(o(nout,n))
 o = 0d0
 do q = 1, n
  select case(trim(kind))
  case('hlld')
   fl = 0d0
   call hlldflux(fl,a(1,q),a(2,q),a(3,q),a(4,q),a(5,q),a(6,q),a(7,q),a(8,q),&
                 a(9,q),a(10,q),a(11,q),a(12,q),a(13,q),a(14,q),a(15,q),a(16,q),&
                 a(17,q),a(18,q),a(19,q),a(20,q),a(21,q),a(22,q),a(23,q),a(24,q),&
                 a(25,q))
   o(1:10,q) = fl
  case('cf')
   o(1,q) = get_cf(a(1,q),a(2,q),a(3,q),a(4,q),a(5,q))
  case('eos')
   gamma = a(3,q)
   T = 1d0
   imu = 1d0
   p = 0d0
   cs = 0d0
   call eos_p_cs(a(1,q),a(2,q),T,imu,p,cs,erad=a(4,q),ierr=ierr)
   o(1,q) = p
   o(2,q) = cs
   o(3,q) = dble(ierr)
  case('eint')
   mag_on = a(9,q) > 0.5d0
   o(1,q) = get_eint(a(1,q),a(2,q),a(3,q),a(4,q),a(5,q),a(6,q),a(7,q),a(8,q),ierr)
   o(2,q) = dble(ierr)
  case('mc')
   flux_limiter = 'modified_mc'
   uu = a(1:3,q)
   x = a(4:6,q)
   xi = a(7:8,q)
   dx = 0d0
   call fluxlimiter(uu,dx,x,xi,du)
   o(1,q) = du
  case('minmod')
   flux_limiter = 'minmod'
   uu = a(1:3,q)
   dx = a(4:5,q)
   x = 0d0
   xi = 0d0
   call fluxlimiter(uu,dx,x,xi,du)
   o(1,q) = du
  end select
 end do
 open(11,file=trim(fout),access='stream',form='unformatted',status='replace')
 write(11) n,nout
 write(11) o
 close(11)
end program kat_reference
