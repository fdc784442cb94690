! ref_shim.f90 -- C-callable entry points into the reference's own scalar
! routines, compiled from the reference sources (test infrastructure only).
! errfc / errf: src/functions.f90:28-50.  TT_damping_functions, get_TT_damp:
! src/tt_damp.f90:42-97.  The reference sources are compiled where they lie;
! nothing of them is copied here.
module mdff_ref_shim
  use iso_c_binding
  use tt_damp, only : get_TT_damp, TT_damping_functions
  implicit none
contains
  function ref_errfc(x) bind(C, name="ref_errfc") result(r)
    real(c_double), value :: x
    real(c_double) :: r
    real(kind=8), external :: errfc
    real(kind=8) :: xx
    xx = x
    r = errfc(xx)
  end function ref_errfc

  subroutine ref_tt_damp(b, c, r, order, f, fd) bind(C, name="ref_tt_damp")
    real(c_double), value :: b, c, r
    integer(c_int), value :: order
    real(c_double) :: f, fd
    real(kind=8) :: bb, cc, rr, ff, ffd
    integer :: o
    call get_TT_damp
    bb = b ; cc = c ; rr = r ; o = order
    call TT_damping_functions(bb, cc, rr, ff, ffd, o)
    f = ff ; fd = ffd
  end subroutine ref_tt_damp
end module mdff_ref_shim
