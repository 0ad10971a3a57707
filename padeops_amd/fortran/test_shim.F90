! Driver of the Fortran shim on one GPU: cd06 ddx of sin(x + phase) on device
! arrays, compared with the scheme's closed-form modified wavenumber
! (k' dx = (14/9 sin w + 1/18 sin 2w) / (1 + 2/3 cos w)), the cf90 filter of a
! constant (T(0) = 1) and the reference's init error codes.  Prints PASS or
! stops with a nonzero code.
program test_shim
    use iso_c_binding
    use padeops_hip_mod
    implicit none
    integer, parameter :: n = 64, na = 3, nb = 2
    real(c_double), parameter :: pi = 3.14159265358979323846d0
    real(c_double), allocatable, target :: h(:,:,:), r(:,:,:)
    real(c_double), pointer, contiguous :: fd(:,:,:), dfd(:,:,:)
    type(c_ptr) :: pf, pdf
    type(cd06) :: d
    type(cf90) :: flt
    integer(c_size_t) :: bytes
    integer :: i, j, k, ierr
    real(c_double) :: dx, w, kp, err, ph

    dx = 2 * pi / n
    allocate(h(n, na, nb), r(n, na, nb))
    do k = 1, nb
        do j = 1, na
            ph = 0.3d0 * j + 0.7d0 * k
            do i = 1, n
                h(i, j, k) = sin((i - 1) * dx + ph)
            end do
        end do
    end do
    bytes = int(8, c_size_t) * n * na * nb
    pf = pdo_alloc(bytes)
    pdf = pdo_alloc(bytes)
    call c_f_pointer(pf, fd, [n, na, nb])
    call c_f_pointer(pdf, dfd, [n, na, nb])
    call pdo_to_device(pf, c_loc(h), bytes)

    ierr = d%init(5, dx, .true., 0, 0)
    if (ierr /= 3) then
        print *, "FAIL: cd06 periodic n=5 ierr", ierr
        error stop 2
    end if
    ierr = d%init(n, dx, .true., 0, 0)
    if (ierr /= 0) error stop 3
    call d%dd1(fd, dfd, na, nb)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(r), pdf, bytes)
    w = dx
    kp = (14.d0 / 9.d0 * sin(w) + 1.d0 / 18.d0 * sin(2 * w)) / (1.d0 + 2.d0 / 3.d0 * cos(w)) / dx
    err = 0
    do k = 1, nb
        do j = 1, na
            ph = 0.3d0 * j + 0.7d0 * k
            do i = 1, n
                err = max(err, abs(r(i, j, k) - kp * cos((i - 1) * dx + ph)))
            end do
        end do
    end do
    print '(a,es10.3)', "cd06 dd1 max error vs closed form: ", err
    if (err > 1d-12) error stop 4

    h = 2.5d0
    call pdo_to_device(pf, c_loc(h), bytes)
    ierr = flt%init(9, .true.)
    if (ierr /= 7) error stop 5
    ierr = flt%init(n, .true.)
    call flt%filter1(fd, dfd, na, nb)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(r), pdf, bytes)
    err = maxval(abs(r - 2.5d0))
    print '(a,es10.3)', "cf90 filter1 of a constant, max error: ", err
    if (err > 1d-13) error stop 6

    call d%destroy()
    call flt%destroy()
    call pdo_release(pf)
    call pdo_release(pdf)
    print '(a)', "PASS"
end program
