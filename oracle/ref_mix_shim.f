! ref_mix_shim.f -- C-callable entry points into the reference's visc
! (trunk/mixing/src/laps.f:535), compiled where it lies under the
! reference tree (see oracle/Makefile.ref).  TEST INFRASTRUCTURE ONLY:
! used by CPU tests to pin the oracle's visc at the mixing code's
! compile-time grid (trunk/mixing/src/ctes3D: my = 256).
!
! Written for this project: it only fills the common blocks that visc
! reads (/d2y/, /cfdiff/, /fis/) with matrices the caller passes, and
! forwards the arguments.
      subroutine ref_mix_dims(d) bind(C, name="ref_mix_dims")
      use iso_c_binding
      implicit none
      include "ctes3D"
      integer(c_int) :: d(1)
      d(1) = my
      end subroutine

! prem1: LU of the d/deta LHS (bandec7 storage); dt12: its RHS band;
! dt21, dt22: the d2/dy2 bands; fm: fmap.  All (nb,my) column-major.
      subroutine ref_mix_set(p1, t12, t21, t22, fm)
     &   bind(C, name="ref_mix_set")
      use iso_c_binding
      implicit none
      include "ctes3D"
      real(c_double) :: p1(7,my), t12(7,my), t21(5,my), t22(5,my)
      real(c_double) :: fm(my)
      real*8 prem3,dt21,dt22
      common /d2y/ prem3(5,my),dt21(5,my),dt22(5,my)
      real*8 prem1,dt12,endder
      common /cfdiff/ prem1(7,my),dt12(7,my),endder(my)
      real*8 Re,alp,bet,a0,y,hy,fmap,trp,mss
      common /fis/ Re,alp,bet,a0,y(my),hy(my),fmap(my),trp(my),mss(my)
      prem1 = p1
      dt12 = t12
      dt21 = t21
      dt22 = t22
      prem3 = t21
      fmap = fm
      end subroutine

      subroutine ref_mix_visc(vorwk, vor, phi, v, dvdy, phiwk, rk1,
     &   rk2, dalbe, dtri, dtxi) bind(C, name="ref_mix_visc")
      use iso_c_binding
      implicit none
      include "ctes3D"
      real(c_float) :: vorwk(2,my), vor(2,my), phi(2,my), v(2,my)
      real(c_float) :: dvdy(2,my), phiwk(2,my)
      real(c_double), value :: rk1, rk2, dalbe, dtri, dtxi
      real(8) :: a1, a2, a3, a4, a5
      integer :: flag
      a1 = rk1
      a2 = rk2
      a3 = dalbe
      a4 = dtri
      a5 = dtxi
      flag = 0
      call visc(vorwk, vor, phi, v, dvdy, phiwk, a1, a2, a3, a4, a5,
     &          flag)
      end subroutine
