! fd_ref_shim -- C entry points over the reference's own finite-difference
! primitives avg_old / diff_old / diff2_old (FD_functions.f90:8-219), so
! the oracle's restatement of avg / diff / diff2 can be checked against the
! reference's code.  TEST INFRASTRUCTURE ONLY: built by `make -C oracle ref`
! from the reference's source where it lies, into oracle/_ref/, and loaded
! only by tests/test_oracle_ref.py.
!
! Each entry takes A(n1,n2,n3) and an output buffer B(m1,m2,m3) of the
! shape the caller expects; it returns 0, or 1 if the reference's result has
! another shape.
module fd_ref_shim
    use, intrinsic :: iso_c_binding
    use FD_functions
    implicit none
contains

    integer(c_int) function put(r, b, m1, m2, m3)
        real(8), intent(in) :: r(:,:,:)
        integer(c_int), intent(in) :: m1, m2, m3
        real(c_double), intent(out) :: b(m1, m2, m3)
        if (size(r, 1) /= m1 .or. size(r, 2) /= m2 .or. size(r, 3) /= m3) then
            put = 1
            return
        end if
        b = r
        put = 0
    end function put

    integer(c_int) function fdref_avg_old(a, n1, n2, n3, k, bc, b, m1, m2, m3) bind(C, name="fdref_avg_old")
        integer(c_int), value :: n1, n2, n3, k, bc, m1, m2, m3
        real(c_double), intent(in) :: a(n1, n2, n3)
        real(c_double), intent(out) :: b(m1, m2, m3)
        fdref_avg_old = put(avg_old(a, k, bc), b, m1, m2, m3)
    end function fdref_avg_old

    integer(c_int) function fdref_diff_old(a, n1, n2, n3, k, b, m1, m2, m3) bind(C, name="fdref_diff_old")
        integer(c_int), value :: n1, n2, n3, k, m1, m2, m3
        real(c_double), intent(in) :: a(n1, n2, n3)
        real(c_double), intent(out) :: b(m1, m2, m3)
        fdref_diff_old = put(diff_old(a, 1, k), b, m1, m2, m3)
    end function fdref_diff_old

    integer(c_int) function fdref_diff2_old(a, n1, n2, n3, k, bc, b, m1, m2, m3) bind(C, name="fdref_diff2_old")
        integer(c_int), value :: n1, n2, n3, k, bc, m1, m2, m3
        real(c_double), intent(in) :: a(n1, n2, n3)
        real(c_double), intent(out) :: b(m1, m2, m3)
        fdref_diff2_old = put(diff2_old(a, k, bc), b, m1, m2, m3)
    end function fdref_diff2_old

end module fd_ref_shim
