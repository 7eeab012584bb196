!##############################################################################
! module vacgpu_projectb -- MI355X drop-in for VAC's projectb
! (src/vacproc.projectb.t:4-157), for builds with the Poisson module on.
!
! VAC's process() calls projectb(w) before a step (from advect, src/vac.t:
! 392-399) and before a save (src/vacio.t:910).  This replacement projects B
! on the device (vacgpu_process: div B, the CG Poisson solve of
! src/vacpoisson.t:391-595, B - grad(phi)), so every projection of the run
! shares the device's saved allowed maximum of div B, as VAC's calls share
! theirs.  Inside advect the device already holds w; before a save w is
! copied there and back.  nerror(poissonerr_) counts as poisson() counts it
! (src/vacpoisson.t:79-81).
!=============================================================================
subroutine projectb(w)

use iso_c_binding
use vacgpu_mod

include 'vacdef.f'

double precision:: w(ixG^T,nw)

integer(c_int) :: rc
integer(c_int), target :: ld(3), pinfo, piter
real(c_double), target :: pp(3), pres
!-----------------------------------------------------------------------------

if(typephys/='mhd' .or. ndim<2) &
   call die('vacgpu projectb: the projection of B needs 2D or 3D MHD')
call vacgpu_dropin_init
^D&ld(^D)=ixGhi^D;
if(ndim<3) ld(3)=1
pp=-one
pp(1:min(3,nprocpar))=procpar(1:min(3,nprocpar))
if(vacgpu_in_advect)then
   rc=vacgpu_process(vacgpu_dropin_grid,1,c_loc(pp),c_loc(pinfo),c_loc(piter),c_loc(pres))
   if(rc/=0) call die('vacgpu projectb: '//vacgpu_errmsg(vacgpu_dropin_grid))
else
   ! before a save: the drop-in's savefile has made w current; project it on
   ! the device and bring the result back, so both copies are current
   if(.not.vacgpu_host_current) call die('vacgpu projectb: the host state lags the device')
   call vacgpu_dropin_put(w)
   rc=vacgpu_process(vacgpu_dropin_grid,3,c_loc(pp),c_loc(pinfo),c_loc(piter),c_loc(pres))
   if(rc/=0) call die('vacgpu projectb: '//vacgpu_errmsg(vacgpu_dropin_grid))
   call vacgpu_dropin_get(w)
endif
if(pinfo/=0.and.pinfo/=3) nerror(poissonerr_)=nerror(poissonerr_)+1

return
end

!=============================================================================
! end module vacgpu_projectb
!##############################################################################
