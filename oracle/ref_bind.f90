! ref_bind.f90 -- bind(C) entry points onto reference routines, for tests.
! TEST INFRASTRUCTURE ONLY: compiled together with the reference's own
! HACApK_lib.f90 / m_ppohBEM_*.f90 by oracle/ref_build.sh.  Each routine
! only forwards its C arguments to the reference routine named in its
! comment; nothing here reimplements reference behaviour.
module ref_bind
  use iso_c_binding
  use m_ppohBEM_user_func, only: face_integral
  implicit none
  interface
    real*8 function HACApK_unrm_d(nd, za)
      real*8 :: za(:)
      integer :: nd
    end function
    subroutine HACApK_maxabsvalloc_d(za, zz, il, nd)
      real*8 :: za(:)
      real*8 :: zz
      integer :: il, nd
    end subroutine
    subroutine HACApK_minabsvalloc_d(za, zz, il, nd)
      real*8 :: za(:)
      real*8 :: zz
      integer :: il, nd
    end subroutine
    subroutine HACApK_maxabsvallocm_d(za, zz, il, nd, lmask)
      real*8 :: za(:)
      real*8 :: zz
      integer :: il, nd
      integer :: lmask(:)
    end subroutine
  end interface
contains
  ! HACApK_lib.f90:31-46
  real(c_double) function ref_unrm_d(nd, za) bind(C, name="ref_unrm_d")
    integer(c_int), value :: nd
    real(c_double), intent(in) :: za(nd)
    ref_unrm_d = HACApK_unrm_d(int(nd), za)
  end function
  ! HACApK_lib.f90:48-57
  subroutine ref_maxabsvalloc_d(nd, za, zz, il) bind(C, name="ref_maxabsvalloc_d")
    integer(c_int), value :: nd
    real(c_double), intent(in) :: za(nd)
    real(c_double), intent(out) :: zz
    integer(c_int), intent(out) :: il
    integer :: ilf
    real*8 :: zzf
    call HACApK_maxabsvalloc_d(za, zzf, ilf, int(nd))
    zz = zzf; il = ilf
  end subroutine
  ! HACApK_lib.f90:71-81
  subroutine ref_minabsvalloc_d(nd, za, zz, il) bind(C, name="ref_minabsvalloc_d")
    integer(c_int), value :: nd
    real(c_double), intent(in) :: za(nd)
    real(c_double), intent(out) :: zz
    integer(c_int), intent(out) :: il
    integer :: ilf
    real*8 :: zzf
    call HACApK_minabsvalloc_d(za, zzf, ilf, int(nd))
    zz = zzf; il = ilf
  end subroutine
  ! m_ppohBEM_user_func.f90:49-100
  real(c_double) function ref_face_integral(xs, ys, zs, x, y, z) bind(C, name="ref_face_integral")
    real(c_double), intent(in) :: xs(3), ys(3), zs(3)
    real(c_double), value :: x, y, z
    ref_face_integral = face_integral(xs, ys, zs, x, y, z)
  end function
end module ref_bind
