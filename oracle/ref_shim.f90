! ref_shim.f90 -- C-callable entry points into the reference's own Fortran
! routines, compiled where they lie under the reference tree (see
! oracle/Makefile.ref).  TEST INFRASTRUCTURE ONLY: used by CPU tests to pin
! the FP64 oracle against the real reference (trunk/lomacte3/src) at the
! reference's compile-time grid (ctes3D: mgalx=192, mgalz=96, my=453).
!
! This file is written for this project; it only sets the reference's module
! state (fis, matrices, point) and forwards arguments.
module ref_shim
  use iso_c_binding
  use fis
  use matrices
  use point
  implicit none
  integer, parameter :: NYREF = 453, MGXREF = 192, MGZREF = 96
contains

  subroutine ref_dims(d) bind(C, name="ref_dims")
    integer(c_int), intent(out) :: d(3)
    d(1) = MGXREF; d(2) = NYREF; d(3) = MGZREF
  end subroutine

  ! derivadas (cfdiff.f:869) + trapezoid weights (main.f:689-696)
  subroutine ref_setup(yin) bind(C, name="ref_setup")
    real(c_double), intent(in) :: yin(NYREF)
    real(8), allocatable :: d11(:,:), d12(:,:), d21(:,:), d22(:,:)
    integer :: j
    y(1:NYREF) = yin(1:NYREF)
    allocate(d11(NYREF,7), d12(NYREF,7), d21(NYREF,5), d22(NYREF,5))
    call derivadas(d11, d12, d21, d22)
    deallocate(d11, d12, d21, d22)
    do j = 2, NYREF-1
      trp(j) = 0.25d0*(y(j+1)-y(j-1))
      hy(j) = (y(j+1)-y(j-1))/2d0
    enddo
    trp(1) = 0.25d0*(y(2)-y(1)); trp(NYREF) = 0.25d0*(y(NYREF)-y(NYREF-1))
    hy(1) = (y(2)-y(1))/2d0;     hy(NYREF) = (y(NYREF)-y(NYREF-1))/2d0
    pb = 1; pe = 1
  end subroutine

  subroutine ref_matrices(p1, t11, t12, p3, t21, t22, fm, ed) bind(C, name="ref_matrices")
    real(c_double), intent(out) :: p1(7,NYREF), t11(7,NYREF), t12(7,NYREF)
    real(c_double), intent(out) :: p3(5,NYREF), t21(5,NYREF), t22(5,NYREF)
    real(c_double), intent(out) :: fm(NYREF), ed(NYREF)
    p1 = prem1; t11 = dt11; t12 = dt12
    p3 = prem3; t21 = dt21; t22 = dt22
    fm = fmap; ed = endder
  end subroutine

  subroutine ref_deryr2(u, du, m) bind(C, name="ref_deryr2")
    integer(c_int), value :: m
    real(c_float) :: u(2,NYREF,m), du(2,NYREF,m)
    call deryr2(u, du, m)
  end subroutine

  subroutine ref_deryyr2(u, du, m) bind(C, name="ref_deryyr2")
    integer(c_int), value :: m
    real(c_float) :: u(2,NYREF,m), du(2,NYREF,m)
    call deryyr2(u, du, m)
  end subroutine

  subroutine ref_deryr(u, du) bind(C, name="ref_deryr")
    real(c_double) :: u(NYREF), du(NYREF)
    call deryr(u, du)
  end subroutine

  subroutine ref_deryyr(u, du) bind(C, name="ref_deryyr")
    real(c_double) :: u(NYREF), du(NYREF)
    call deryyr(u, du)
  end subroutine

  subroutine ref_inty8(du, u) bind(C, name="ref_inty8")
    real(c_double) :: du(NYREF), u(NYREF)
    call inty8(du, u)
  end subroutine

  subroutine ref_lapvdvhom(phi, v, dvdy, rk) bind(C, name="ref_lapvdvhom")
    real(c_float) :: phi(2,NYREF), v(2,NYREF), dvdy(2,NYREF)
    real(c_double), value :: rk
    call Lapvdvhom(phi, v, dvdy, rk)
  end subroutine

  subroutine ref_solvepsi(phi, v, dvdy, rk) bind(C, name="ref_solvepsi")
    real(c_float) :: phi(2,NYREF), v(2,NYREF), dvdy(2,NYREF)
    real(c_double), value :: rk
    call solvepsi(phi, v, dvdy, rk)
  end subroutine

  subroutine ref_bandec(a, n) bind(C, name="ref_bandec")
    integer(c_int), value :: n
    real(c_double) :: a(5,n)
    call bandec(a, n)
  end subroutine

  subroutine ref_banbks(a, n, b) bind(C, name="ref_banbks")
    integer(c_int), value :: n
    real(c_double) :: a(5,n), b(n)
    call banbks(a, n, b)
  end subroutine

  subroutine ref_bandec7(a, n) bind(C, name="ref_bandec7")
    integer(c_int), value :: n
    real(c_double) :: a(7,n)
    call bandec7(a, n)
  end subroutine

  subroutine ref_banbks7(a, n, b) bind(C, name="ref_banbks7")
    integer(c_int), value :: n
    real(c_double) :: a(7,n), b(n)
    call banbks7(a, n, b)
  end subroutine

  ! FFTPACK single precision (cftsingle.f / rftsingle.f) through the
  ! reference's own fou3D.f wrappers
  subroutine ref_fft_init() bind(C, name="ref_fft_init")
    call cfti(MGZREF)
    call rfti(MGXREF)
  end subroutine

  ! fourz on one plane: phys(2*my*mgalz) real(4), z fastest per y
  subroutine ref_fourz(phys, iopt) bind(C, name="ref_fourz")
    real(c_float) :: phys(2*NYREF*MGZREF)
    integer(c_int), value :: iopt
    call fourz(phys, iopt)
  end subroutine

  ! fourx on one line: phys(mgalx+2) real(4) -> phys8 real(8)
  subroutine ref_fourx(phys, phys8, iopt) bind(C, name="ref_fourx")
    real(c_float) :: phys(MGXREF+2)
    real(c_double) :: phys8(MGXREF+2)
    integer(c_int), value :: iopt
    call fourx(phys, phys8, iopt)
  end subroutine

  ! local transposes of one plane (fou3D.f:103-189)
  subroutine ref_localyz2zy(f, chwk) bind(C, name="ref_localyz2zy")
    complex(c_float_complex) :: f(0:NYREF-1, 0:MGZREF-1)
    complex(c_float_complex) :: chwk(0:NYREF*MGZREF-1)
    call localyz2zy(f, f, chwk)
  end subroutine

  subroutine ref_localzy2yz(f, chwk) bind(C, name="ref_localzy2yz")
    complex(c_float_complex) :: f(0:NYREF-1, 0:MGZREF-1)
    complex(c_float_complex) :: chwk(0:NYREF*MGZREF-1)
    call localzy2yz(f, f, chwk)
  end subroutine
end module ref_shim
