! ref_bind.f90 -- C binding for the reference's own chlld/minmod/linear,
! compiled from SET_Fortran_Cuda_MHD/Solvers.f90 where it lies (see
! ref_build.sh).  Test infrastructure: it only makes the unmodified
! reference routines callable from the oracle's tests.
module ref_bind
  use iso_c_binding
  use ref_chlld_mod
  implicit none
contains
  subroutine ref_chlld_c(n_state, al, be, ge, w, q1, q2, dsl, dsp, dsc, qqq, n_disc) &
      bind(C, name="ref_chlld")
    integer(c_int), value :: n_state, n_disc
    real(c_double), value :: al, be, ge, w
    real(c_double), intent(in) :: q1(8), q2(8)
    real(c_double), intent(out) :: dsl, dsp, dsc, qqq(8)
    real(8) :: a1(8), a2(8), out(8)
    a1 = q1
    a2 = q2
    out = 0.0_8
    call chlld(n_state, al, be, ge, w, a1, a2, dsl, dsp, dsc, out, n_disc)
    qqq = out
  end subroutine ref_chlld_c

  real(c_double) function ref_minmod_c(x, y) bind(C, name="ref_minmod")
    real(c_double), value :: x, y
    ref_minmod_c = minmod(x, y)
  end function ref_minmod_c

  real(c_double) function ref_linear_c(x1, t1, x2, t2, x3, t3, y) bind(C, name="ref_linear")
    real(c_double), value :: x1, t1, x2, t2, x3, t3, y
    ref_linear_c = linear(x1, t1, x2, t2, x3, t3, y)
  end function ref_linear_c
end module ref_bind
