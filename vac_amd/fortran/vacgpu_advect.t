!##############################################################################
! module vacgpu_advect -- MI355X drop-in for VAC's advect (src/vac.t:363-486)
!
! Linked into VAC in place of the advect subroutine of vac.t.  Every call
! advances one dimensionally split sweep on the GPU through libvacgpu (the C
! ABI of include/vacgpu.h, bound in vacgpu_mod.f90).  The selectors come from
! VAC's own common block, so the par file keeps choosing typefull,
! typelimiter, courantpar, divbfix, typeB, ...  A selector the GPU path does
! not implement stops the run through die(); nothing falls back to the CPU.
!
! The build also routes the main loop's wold=w (src/vac.t:85) through
! vacgpu_store_wold below, which copies only while the host's w is current.
!
! State transfer: the state stays on the device between steps.  VAC's w is
! copied there at the first sweep of the run (and after the host changed it)
! and fetched back only where VAC reads it on the host:
!  * before every savefile (src/vacio.t:1008-1105), through the savefile
!    below, which the drop-in build links in place of VAC's (renamed
!    savefile_cpu);
!  * every step, when a host routine reads or changes w between steps: the
!    residual (residmin>0 or residmax<1d99, src/vac.t:90-101), split sources
!    (sourcesplit, src/vac.t:236-272) or a filter (typefilter, src/vac.t:
!    332-355); the state then goes back to the device at the next step;
!  * after a step that left some direction without a lagged dtcourant, since
!    getdt_courant then takes cmax from w (src/vac.t:691-700).
! dtcourant, nerror and smallp are exchanged with the common block on every
! call, as the CPU advect would update them.
!
! User hooks: specialsource is called by VAC inside advect, in every substep
! (addsource2, src/vac.t:592).  The device path does not call host code
! inside a sweep, so the first advect of a run calls the linked
! specialsource once on copies of the state and stops the run if it changes
! anything (vacusr=nul passes).  specialbound (typeB='special') is rejected
! with the other boundary types the GPU does not implement.
!=============================================================================
subroutine advect(method,ix^L,iws,idim^LIM,w1,w)

use iso_c_binding
use vacgpu_mod

include 'vacdef.f'

character*^LENTYPE :: method
integer:: ix^L,iws(niw_),idim^LIM
double precision:: w1(ixG^T,nw),w(ixG^T,nw)

logical:: firstsweep,lastsweep
common/first/firstsweep,lastsweep

type(c_ptr) :: grid
type(vacgpu_method) :: m
integer(c_int) :: rc
integer(c_int), target :: ld(3), ne(4)
integer:: iw, code, idim
real(c_double), target :: dtc(3), sp
logical:: hostsync, fetch
logical, save :: probed = .false.
integer(8) :: clk, clkrate
!-----------------------------------------------------------------------------

if(ix^L/=ixG^L|.or.|.or.) call die('vacgpu advect: subgrids are not supported')
if(idimmin/=idimmax .and. (idimmin/=1 .or. idimmax/=ndim)) &
   call die('vacgpu advect: a sweep covers one direction or all of them')
if(nproc(2)/=0) call die('vacgpu advect: nproc(2) processing is not on the GPU path')
if(typeaxial/='slab' .or. gencoord) call die('vacgpu advect: slab Cartesian grids only')
if(smallfix .or. extraB) call die('vacgpu advect: smallfix/extraB are not on the GPU path')
if(acmexpo>zero) call die('vacgpu advect: acmexpo>0 is not on the GPU path')
if(ndir/=ndim) call die('vacgpu advect: ndir must equal ndim')
do iw=1,nw
   if(acmcoef(iw)>=zero) call die('vacgpu advect: acmcoef is not on the GPU path')
end do
select case(method)
case('tvdlf')
   if(typeadvance/='twostep' .or. typepred1/='hancock') &
      call die('vacgpu advect: tvdlf needs typeadvance=twostep, typepred=hancock')
   m%typefull=VACGPU_TVDLF
case('tvdmu')
   if(typeadvance/='twostep' .or. typepred1/='hancock') &
      call die('vacgpu advect: tvdmu needs typeadvance=twostep, typepred=hancock')
   if(typetvd/='roe') call die('vacgpu advect: typetvd must be roe')
   m%typefull=VACGPU_TVDMU
case('tvd')
   if(typeadvance/='onestep') call die('vacgpu advect: tvd needs typeadvance=onestep')
   if(typetvd/='roe') call die('vacgpu advect: typetvd must be roe')
   m%typefull=VACGPU_TVD
case default
   call die('vacgpu advect: typefull='//method//' is not on the GPU path')
end select

^D&ld(^D)=ixGhi^D;
if(ndim<3) ld(3)=1
if(ndim<2) ld(2)=1

call vacgpu_dropin_init
grid=vacgpu_dropin_grid

m%typelimiter=VACGPU_MINMOD
do iw=1,nw
   code=vacgpu_limiter_code(typelimiter(iw))
   if(code<0) call die('vacgpu advect: typelimiter='//typelimiter(iw)//' is not on the GPU path')
   m%typelimiter(iw)=code
end do
m%useprimitive=merge(1,0,useprimitive)
! VAC's own main loop applies dtcantgrow and slowsteps (src/vac.t:68-78)
m%dtcantgrow=0
m%slowsteps=0
m%unsplit=merge(1,0,idimmin/=idimmax)
m%divbfix=merge(1,0,divbfix)
m%divbwave=merge(1,0,divbwave)
m%sourceunsplit=merge(1,0,sourceunsplit)
m%dimsplit_xy=merge(1,0,typedimsplit=='xy')
m%courantpar=courantpar
if(implpar>=zero) m%courantpar=-one
m%smallpcoeff=smallpcoeff
m%typeentropy=VACGPU_ENT_NUL
m%entropycoef=-one
do iw=1,nw
   select case(typeentropy(iw))
   case('nul')
      m%typeentropy(iw)=VACGPU_ENT_NUL
   case('harten')
      m%typeentropy(iw)=VACGPU_ENT_HARTEN
   case('powell')
      m%typeentropy(iw)=VACGPU_ENT_POWELL
   case('ratio')
      m%typeentropy(iw)=VACGPU_ENT_RATIO
   case('yee')
      m%typeentropy(iw)=VACGPU_ENT_YEE
   case default
      call die('vacgpu advect: typeentropy='//typeentropy(iw)//' is not on the GPU path')
   end select
   m%entropycoef(iw)=entropycoef(iw)
end do
select case(typelimited)
case('original')
   m%typelimited=VACGPU_LIM_ORIGINAL
case('previous')
   m%typelimited=VACGPU_LIM_PREVIOUS
case('predictor')
   m%typelimited=VACGPU_LIM_PREDICTOR
case default
   call die('vacgpu advect: typelimited='//typelimited//' is not on the GPU path')
end select

! host routines that read or change w between steps
hostsync = residmin>zero .or. residmax<bigdouble .or. sourcesplit .or. typefilter1/='nul'

if(firstsweep)then
   ! with the trace on, the wall clock at the start of every step: the
   ! difference of two such lines is one step of VAC's loop, advect included
   if(vacgpu_trace)then
      call system_clock(clk,clkrate)
      write(*,'(a,i8,a,f18.6)')'vacgpu: step at it=',it,' wall_s=',dble(clk)/dble(clkrate)
   endif
   if(.not.probed)then
      probed=.true.
      call vacgpu_probe_source(iws,w,w1)
   endif
   if(.not.vacgpu_dev_current) call vacgpu_dropin_put(w)
endif

! process(1) when due, as VAC's advect does (src/vac.t:392-399).  VAC's
! process calls projectb, which a build with the Poisson module replaces by
! the device projection (vacgpu_projectb.t); inside advect it works on the
! device state put there above.
if(nproc(1)/=0.and.implpar<=zero)then
   if(nproc(1)<0.or.(firstsweep.and.it-itmin==((it-itmin)/nproc(1))*nproc(1)))then
      vacgpu_in_advect=.true.
      call process(1,idim^LIM,w)
      vacgpu_in_advect=.false.
   endif
endif

dtc=bigdouble
dtc(1:ndim)=dtcourant(1:ndim)
ne=nerror(1:4)
sp=smallp
rc=vacgpu_advect(grid,m,idimmin,idimmax,dt,merge(1,0,firstsweep),merge(1,0,lastsweep), &
   c_loc(dtc),c_loc(ne),c_loc(sp))
if(rc/=0) call die('vacgpu advect: '//vacgpu_errmsg(grid))
dtcourant(1:ndim)=dtc(1:ndim)
nerror(1:4)=ne
smallp=sp
vacgpu_host_current=.false.

if(lastsweep)then
   fetch=hostsync
   if(dtpar<=zero .and. courantpar>zero)then
      do idim=1,ndim
         if(dtcourant(idim)>=bigdouble) fetch=.true.
      end do
   endif
   if(fetch) call vacgpu_dropin_get(w)
   ! the host may change w before the next step (split sources, filters)
   if(hostsync) vacgpu_dev_current=.false.
endif

firstsweep=.false.

return
end

!=============================================================================
subroutine savefile(ifile,w)

! Replaces VAC's savefile (src/vacio.t:1008-1105, renamed savefile_cpu in
! the drop-in build): fetches the device state into w when w lags behind it,
! then saves as VAC does.  A projection before the save (nproc(ifile+2),
! src/vacio.t:1017) runs on the device through vacgpu_projectb.t and leaves
! both copies current.

use vacgpu_mod

include 'vacdef.f'

integer:: ifile
double precision:: w(ixG^T,nw)
!-----------------------------------------------------------------------------

if(.not.vacgpu_host_current) call vacgpu_dropin_get(w)
call savefile_cpu(ifile,w)

return
end

!=============================================================================
subroutine vacgpu_store_wold(w)

! The main loop's wold=w (src/vac.t:85), which the drop-in build routes
! here.  On the host only the residual (src/vac.t:90-101) reads wold; the
! device keeps its own copy of the step's start for the limiters and the
! source steps.  Between steps the host's w lags the device state unless a
! host routine made it current (a save, the residual, split sources, a
! filter), so the copy is made only then -- for the residual it always is.

use vacgpu_mod

include 'vacdef.f'

double precision:: w(ixG^T,nw)
!-----------------------------------------------------------------------------

if(vacgpu_host_current) wold(ixG^S,1:nw)=w(ixG^S,1:nw)

return
end

!=============================================================================
subroutine vacgpu_dropin_put(w)

! VAC's w (ghost cells as they are) -> the device state

use iso_c_binding
use vacgpu_mod

include 'vacdef.f'

double precision:: w(ixG^T,nw)
integer(c_int) :: rc
integer(c_int), target :: ld(3)
!-----------------------------------------------------------------------------

call vacgpu_dropin_init
ld=1
^D&ld(^D)=ixGhi^D;
rc=vacgpu_put_state_ghosts(vacgpu_dropin_grid,w,ld)
if(rc/=0) call die('vacgpu: '//vacgpu_errmsg(vacgpu_dropin_grid))
vacgpu_dev_current=.true.
vacgpu_n_put=vacgpu_n_put+1
if(vacgpu_trace) write(*,'(a,i8)')'vacgpu: state to device at it=',it

return
end

!=============================================================================
subroutine vacgpu_dropin_get(w)

! the device state -> VAC's w

use iso_c_binding
use vacgpu_mod

include 'vacdef.f'

double precision:: w(ixG^T,nw)
integer(c_int) :: rc
integer(c_int), target :: ld(3)
!-----------------------------------------------------------------------------

ld=1
^D&ld(^D)=ixGhi^D;
rc=vacgpu_get_state(vacgpu_dropin_grid,w,ld)
if(rc/=0) call die('vacgpu: '//vacgpu_errmsg(vacgpu_dropin_grid))
vacgpu_host_current=.true.
vacgpu_n_get=vacgpu_n_get+1
if(vacgpu_trace) write(*,'(a,i8)')'vacgpu: state to host at it=',it

return
end

!=============================================================================
subroutine vacgpu_probe_source(iws,w,w1)

! Calls the linked specialsource once, with qdt = dt, on copies of the
! state (wCT and w both equal to VAC's w; w1 is advect's scratch array) and
! stops the run if it changed anything: user source terms would be added by
! VAC inside every substep of advect (addsource2, src/vac.t:592), which the
! device path does not do.

include 'vacdef.f'

integer:: iws(niw_)
double precision:: w(ixG^T,nw),w1(ixG^T,nw),wct(ixG^T,nw)
!-----------------------------------------------------------------------------

wct(ixG^S,1:nw)=w(ixG^S,1:nw)
w1(ixG^S,1:nw)=w(ixG^S,1:nw)
call specialsource(dt,ixG^L,ixM^L,iws,t,wct,t,w1)
if(any(w1(ixG^S,1:nw)/=w(ixG^S,1:nw))) call die( &
   'vacgpu advect: the linked specialsource changes w; user source terms '// &
   'are not on the GPU path (build VAC with -u=nul)')

return
end

!=============================================================================
subroutine vacgpu_dropin_init

! Creates the device grid from VAC's common block (ixG, ixM, dx, typeB,
! eqpar) on the first call: from the first advect, or from the first
! projection when a save comes first.

use iso_c_binding
use vacgpu_mod

include 'vacdef.f'

type(vacgpu_config), save :: cfg
type(c_ptr) :: grid
integer(c_int) :: rc
integer(c_int) :: ld(3)
integer:: iw, iB, code, ngx(3), mlo(3), mhi(3)
double precision, save, target :: dxl(ixGhimax,3)
character(len=8) :: trace
!-----------------------------------------------------------------------------

if(c_associated(vacgpu_dropin_grid)) return
call get_environment_variable('VACGPU_DROPIN_TRACE',trace)
vacgpu_trace = trace/=' ' .and. trace/='0'
^D&ld(^D)=ixGhi^D;
if(ndim<3) ld(3)=1
if(ndim<2) ld(2)=1
cfg%ndim=ndim
select case(typephys)
case('mhd')
   cfg%physics=VACGPU_MHD
case('hd')
   cfg%physics=VACGPU_HD
case default
   call die('vacgpu advect: physics '//typephys//' is not on the GPU path')
end select
cfg%nx=1
^D&cfg%nx(^D)=ixMmax^D-ixMmin^D+1;
cfg%eqpar=zero
cfg%eqpar(1:min(2,neqpar))=eqpar(1:min(2,neqpar))
if(nB/=2*ndim) call die('vacgpu advect: only the default boundary regions are supported')
cfg%typeB=VACGPU_CONT
do iB=1,nB
   do iw=1,nw
      code=vacgpu_bc_code(typeB(iw,iB))
      if(code<0) call die('vacgpu advect: typeB='//typeB(iw,iB)//' is not on the GPU path')
      ! 'fixed' keeps the first fill's copies; fullgridini=T would keep
      ! the file's ghost values instead (src/vacgrid.t:249-257)
      if((code==VACGPU_FIXED .or. code==VACGPU_FIXED1 .or. code==VACGPU_GRAD1) .and. fullgridini) &
         call die('vacgpu advect: typeB=fixed with fullgridini=T is not on the GPU path')
      cfg%typeB(iw,iB)=code
   end do
end do
cfg%precision=VACGPU_FP64
ngx=1
mlo=1
mhi=1
^D&ngx(^D)=ixGmax^D;
^D&mlo(^D)=ixMmin^D;
^D&mhi(^D)=ixMmax^D;
call vacgpu_dxlines(dx,ld,ndim,ngx,mlo,mhi,dxl,ixGhimax)
rc=vacgpu_create(cfg,c_loc(dxl(1,1)),c_loc(dxl(1,2)),c_loc(dxl(1,3)),grid)
if(rc/=0) call die('vacgpu advect: vacgpu_create failed (GPU missing or unsupported grid)')
vacgpu_dropin_grid=grid

return
end

!=============================================================================
subroutine vacgpu_dxlines(dxa,ld,nd,ngx,mlo,mhi,dxl,nmax)

! Cell widths along each dimension over ixG, taken on the line through the
! first mesh cell, after checking that dx(:,...,d) does not vary across the
! mesh (a slab Cartesian grid, src/vacgrid.t:527-530).

integer:: ld(3),nd,ngx(3),mlo(3),mhi(3),nmax
double precision:: dxa(*),dxl(nmax,3)
integer:: d,i,j,k,ic(3),lin
dxl=1.d0
do d=1,nd
   do i=1,ngx(d)
      ic=mlo
      if(nd<3) ic(3)=1
      if(nd<2) ic(2)=1
      ic(d)=i
      lin=ic(1)+(ic(2)-1)*ld(1)+(ic(3)-1)*ld(1)*ld(2)+(d-1)*ld(1)*ld(2)*ld(3)
      dxl(i,d)=dxa(lin)
   end do
   do k=mlo(3),mhi(3)
      do j=mlo(2),mhi(2)
         do i=mlo(1),mhi(1)
            ic(1)=i; ic(2)=j; ic(3)=k
            if(nd<3) ic(3)=1
            lin=ic(1)+(ic(2)-1)*ld(1)+(ic(3)-1)*ld(1)*ld(2)+(d-1)*ld(1)*ld(2)*ld(3)
            if(dxa(lin)/=dxl(ic(d),d)) call die('vacgpu advect: dx varies across the mesh')
         end do
      end do
   end do
end do

return
end
