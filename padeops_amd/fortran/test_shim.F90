! Driver of the Fortran shim on one GPU.  Every check compares device results
! with a closed form that a wrong binding, argument order or callback would
! miss:
!   cd06 ddx of sin(x + phase)      collocated modified wavenumber
!                                   (14/9 sin w + 1/18 sin 2w) / (1 + 2/3 cos w)
!   cf90 filter1 of a constant      T(0) = 1; init error codes 3 and 7
!   cd06stagg (both generics)       periodic ddz_E2C of sin(z) on edges, real
!                                   and complex: staggered modified wavenumber
!                                   (2a sin(w/2) + 2/3 b sin(3w/2)) / (1 + 2 al cos w),
!                                   a = 63/62, b = 17/62, al = 9/62; non-periodic
!                                   init (even ends) keeps a constant constant
!   PoissonPeriodic%init            Get_ModKz callback (staggered cd06) applied:
!                                   the solve of -k'^2 sin(x + 2y + 3z) returns
!                                   sin(...)
!   derivatives / filters           "four" ddz of sin(z) = cos(z); "cd06" ddx;
!                                   getMethod*; cf90 filterx of a constant
!   spectral / Pade6stagg /         fft -> mTimes_ik1 -> ifft of sin(x) = cos(x);
!   padepoisson                     Pade6stagg ddz_C2E (real and complex) with
!                                   the staggered wavenumber; PressureProjection
!                                   of sin(x) + sin(y) leaves sin(y), with zero
!                                   DivergenceCheck; the same with walls in z
!                                   (w zero on both walls); getPressure and
!                                   getPressureAndUpdateRHS (walls, Stokes) of
!                                   (sin x, 0, 0) give -cos x and u = 0
!   cgrid                           getRHS of an entropy wave (density and
!                                   momentum equations in closed form); filter
!                                   and the Gaussian filter keep a constant
!   padeops_hip_decomp              decomp_info sizes, transpose_x_to_y copy
! Prints PASS or stops with a nonzero code.
module shim_callbacks
    use iso_c_binding
    implicit none
contains
    ! Get_ModK callback: staggered cd06 modified wavenumber of k*dz (nonzero
    ! at the Nyquist mode, unlike a collocated one, whose zero the 1e-20
    ! guard of the reference's solve would turn into a 1e20 gain)
    subroutine modk_stagg(k)
        real(c_double), dimension(:), intent(inout) :: k
        integer :: i
        do i = 1, size(k)
            k(i) = kstagg(k(i))
        end do
    end subroutine

    pure real(c_double) function kstagg(w)
        real(c_double), intent(in) :: w
        kstagg = (2 * (63.d0 / 62.d0) * sin(w / 2) + (2.d0 / 3.d0) * (17.d0 / 62.d0) * sin(1.5d0 * w)) / &
                 (1.d0 + 2 * (9.d0 / 62.d0) * cos(w))
    end function
end module

program test_shim
    use iso_c_binding
    use padeops_hip_mod
    use padeops_hip_decomp
    use shim_callbacks
    implicit none
    integer, parameter :: n = 64, na = 3, nb = 2, m = 16, mh = m / 2 + 1
    real(c_double), parameter :: pi = 3.14159265358979323846d0
    real(c_double), allocatable, target :: h(:,:,:), r(:,:,:), g(:,:,:), gE(:,:,:), gr(:,:,:)
    complex(c_double), allocatable, target :: zc(:,:,:), zE(:,:,:), zr(:,:,:)
    real(c_double), pointer, contiguous :: fd(:,:,:), dfd(:,:,:), ad(:,:,:), bd(:,:,:), ed(:,:,:)
    complex(c_double), pointer, contiguous :: cu(:,:,:), cv(:,:,:), cw(:,:,:), ct(:,:,:)
    real(c_double), allocatable, target :: q(:,:,:,:), qr(:,:,:,:)
    real(c_double), pointer, contiguous :: Wd(:,:,:,:), Rd(:,:,:,:), Fq(:,:,:)
    type(c_ptr) :: pf, pdf, pa, pb, pe, pcu, pcv, pcw, pct, pW, pR
    type(cd06) :: d
    type(cf90) :: flt
    type(cd06stagg) :: stg, stgnp
    type(poissonperiodic) :: poiss
    type(derivatives) :: der
    type(filters) :: fil
    type(spectral) :: spC, spE, spW
    type(Pade6stagg) :: p6
    type(padepoisson) :: pp
    type(cgrid) :: cg
    type(cgrid_params) :: cgp
    type(decomp_info) :: gpC, gpE
    integer(c_size_t) :: bytes, rb, rbE, cb, cbE
    integer :: i, j, k, ierr
    real(c_double) :: dx, w, kp, err, ph, dz, x, y, z, ks

    ! ---------------------------------------------------------------- cd06, cf90
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

    ! ------------------------------------------------ m^3 boxes for the rest
    dx = 2 * pi / m
    dz = dx
    rb = int(8, c_size_t) * m * m * m
    rbE = int(8, c_size_t) * m * m * (m + 1)
    cb = int(16, c_size_t) * mh * m * m
    cbE = int(16, c_size_t) * mh * m * (m + 1)
    allocate(g(m, m, m), gr(m, m, m), gE(m, m, m + 1), zc(mh, m, m), zE(mh, m, m + 1), zr(mh, m, m + 1))
    pa = pdo_alloc(rbE); pb = pdo_alloc(rbE); pe = pdo_alloc(rbE)
    pcu = pdo_alloc(cbE); pcv = pdo_alloc(cbE); pcw = pdo_alloc(cbE); pct = pdo_alloc(cbE)

    ! ---------------------------------------------- decomp_info, transpose
    call decomp_2d_init(m, m, m, 0, 0)
    call decomp_info_init(m, m, m, gpC)
    call decomp_info_init(m, m, m + 1, gpE)
    if (any(gpC%ysz /= [m, m, m]) .or. gpE%zsz(3) /= m + 1 .or. gpC%xst(1) /= 1) error stop 40
    do k = 1, m
        do j = 1, m
            do i = 1, m
                g(i, j, k) = i + 100 * j + 10000 * k
            end do
        end do
    end do
    call c_f_pointer(pa, ad, [m, m, m])
    call c_f_pointer(pb, bd, [m, m, m])
    call pdo_to_device(pa, c_loc(g), rb)
    call transpose_x_to_y(ad, bd, gpC)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    if (any(gr /= g)) error stop 41
    print '(a)', "decomp_info / transpose_x_to_y: ok"

    ! --------------------------------------------------------------- cd06stagg
    ! periodic ddz_E2C, real: sin(z) on the m+1 edges -> k' cos(z) on cells
    do k = 1, m + 1
        gE(:, :, k) = sin((k - 1) * dz)
    end do
    call c_f_pointer(pa, ed, [m, m, m + 1])
    call c_f_pointer(pb, bd, [m, m, m])
    call pdo_to_device(pa, c_loc(gE), rbE)
    call stg%init(m, dz)
    call stg%ddz_E2C(ed, bd, m, m)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    ks = kstagg(dz) / dz
    err = 0
    do k = 1, m
        err = max(err, maxval(abs(gr(:, :, k) - ks * cos((k - 0.5d0) * dz))))
    end do
    print '(a,es10.3)', "cd06stagg ddz_E2C real vs closed form: ", err
    if (err > 1d-12) error stop 10
    ! complex generic: (sin z, 2 cos z) -> k' (cos z, -2 sin z)
    do k = 1, m + 1
        zE(:, :, k) = cmplx(sin((k - 1) * dz), 2 * cos((k - 1) * dz), c_double)
    end do
    call c_f_pointer(pcu, cw, [mh, m, m + 1])
    call c_f_pointer(pcv, ct, [mh, m, m])
    call pdo_to_device(pcu, c_loc(zE), cbE)
    call stg%ddz_E2C(cw, ct, mh, m)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(zc), pcv, cb)
    err = 0
    do k = 1, m
        z = (k - 0.5d0) * dz
        err = max(err, maxval(abs(zc(:, :, k) - ks * cmplx(cos(z), -2 * sin(z), c_double))))
    end do
    print '(a,es10.3)', "cd06stagg ddz_E2C complex vs closed form: ", err
    if (err > 1d-12) error stop 11
    call stg%destroy()
    ! non-periodic generic, even ends: interpolation keeps a constant
    g = 1.75d0
    call c_f_pointer(pa, ad, [m, m, m])
    call c_f_pointer(pb, ed, [m, m, m + 1])
    call pdo_to_device(pa, c_loc(g), rb)
    call stgnp%init(m, dz, .true., .true.)
    call stgnp%InterpZ_C2E(ad, ed, m, m)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gE), pb, rbE)
    err = maxval(abs(gE - 1.75d0))
    print '(a,es10.3)', "cd06stagg non-periodic InterpZ_C2E of a constant: ", err
    if (err > 1d-13) error stop 12
    call stgnp%destroy()

    ! ------------------------------------- PoissonPeriodic with Get_ModKz
    ! rhs = -(1 + 4 + kz'(3)^2) sin(x + 2y + 3z); the solve returns sin(...)
    ks = kstagg(3 * dz) / dz
    do k = 1, m
        do j = 1, m
            do i = 1, m
                x = (i - 1) * dx; y = (j - 1) * dx; z = (k - 1) * dz
                g(i, j, k) = sin(x + 2 * y + 3 * z)
            end do
        end do
    end do
    gr = -(1.d0 + 4.d0 + ks * ks) * g
    call c_f_pointer(pa, ad, [m, m, m])
    call c_f_pointer(pb, bd, [m, m, m])
    call pdo_to_device(pa, c_loc(gr), rb)
    call poiss%init(dx, dx, dz, gpC, 1, Get_ModKz=modk_stagg)
    if (abs(poiss%kz(4) - ks) > 1d-13 * ks) error stop 20
    call poiss%poisson_solve(ad, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = maxval(abs(gr - g))
    print '(a,es10.3)', "PoissonPeriodic (Get_ModKz = cd06stagg) single mode: ", err
    if (err > 1d-12) error stop 21
    call poiss%poisson_solve(ad)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pa, rb)
    err = maxval(abs(gr - g))
    if (err > 1d-12) error stop 22
    call poiss%destroy()

    ! ----------------------------------------------- derivatives, filters
    do k = 1, m
        g(:, :, k) = sin((k - 1) * dz)
    end do
    call pdo_to_device(pa, c_loc(g), rb)
    call der%init(m, m, m, dx, dx, dz, .true., .true., .true., "cd06", "cd10", "four")
    if (der%getMethodx() /= 2 .or. der%getMethody() /= 1 .or. der%getMethodz() /= 3) error stop 30
    call der%ddz(ad, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = 0
    do k = 1, m
        err = max(err, maxval(abs(gr(:, :, k) - cos((k - 1) * dz))))
    end do
    print '(a,es10.3)', "derivatives four ddz of sin(z): ", err
    if (err > 1d-12) error stop 31
    do i = 1, m
        g(i, :, :) = sin((i - 1) * dx)
    end do
    call pdo_to_device(pa, c_loc(g), rb)
    call der%ddx(ad, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    kp = (14.d0 / 9.d0 * sin(dx) + 1.d0 / 18.d0 * sin(2 * dx)) / (1.d0 + 2.d0 / 3.d0 * cos(dx)) / dx
    err = 0
    do i = 1, m
        err = max(err, maxval(abs(gr(i, :, :) - kp * cos((i - 1) * dx))))
    end do
    print '(a,es10.3)', "derivatives cd06 ddx of sin(x): ", err
    if (err > 1d-12) error stop 32
    call der%destroy()
    g = -0.5d0
    call pdo_to_device(pa, c_loc(g), rb)
    call fil%init(m, m, m, .true., .true., .true., "cf90", "cf90", "cf90")
    call fil%filterx(ad, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = maxval(abs(gr + 0.5d0))
    print '(a,es10.3)', "filters cf90 filterx of a constant: ", err
    if (err > 1d-13) error stop 33
    call fil%destroy()

    ! -------------------------------------- spectral, Pade6stagg, padepoisson
    call spC%init("x", m, m, m, dx, dx, dz, "four", "none", 2, fixOddball=.false., init_periodicInZ=.true.)
    call spE%init("x", m, m, m + 1, dx, dx, dz, "four", "none", 2, fixOddball=.false., init_periodicInZ=.false.)
    if (any(spC%spectdecomp%ysz /= [mh, m, m])) error stop 50
    do i = 1, m
        g(i, :, :) = sin((i - 1) * dx)
    end do
    call c_f_pointer(pcu, cu, [mh, m, m])
    call c_f_pointer(pcv, cv, [mh, m, m])
    call pdo_to_device(pa, c_loc(g), rb)
    call spC%fft(ad, cu)
    call spC%mTimes_ik1_oop(cu, cv)
    call spC%ifft(cv, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = 0
    do i = 1, m
        err = max(err, maxval(abs(gr(i, :, :) - cos((i - 1) * dx))))
    end do
    print '(a,es10.3)', "spectral fft -> ik1 -> ifft of sin(x): ", err
    if (err > 1d-12) error stop 51

    call p6%init(spC%spectdecomp, spC%spectdecomp, spE%spectdecomp, spE%spectdecomp, dz, cd06_scheme, .true.)
    call p6%destroy()
    call p6%init(gpC, spC%spectdecomp, gpE, spE%spectdecomp, dz, cd06_scheme, .true.)
    do k = 1, m
        g(:, :, k) = cos((k - 0.5d0) * dz)
    end do
    call c_f_pointer(pb, ed, [m, m, m + 1])
    call pdo_to_device(pa, c_loc(g), rb)
    call p6%ddz_C2E(ad, ed, 0, 0)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gE), pb, rbE)
    ks = kstagg(dz) / dz
    err = 0
    do k = 1, m + 1
        err = max(err, maxval(abs(gE(:, :, k) + ks * sin((k - 1) * dz))))
    end do
    print '(a,es10.3)', "Pade6stagg ddz_C2E real vs closed form: ", err
    if (err > 1d-12) error stop 52
    do k = 1, m
        zc(:, :, k) = cmplx(cos((k - 0.5d0) * dz), 0.d0, c_double)
    end do
    call c_f_pointer(pcw, cw, [mh, m, m + 1])
    call pdo_to_device(pcu, c_loc(zc), cb)
    call p6%ddz_C2E(cu, cw, 0, 0)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(zE), pcw, cbE)
    err = 0
    do k = 1, m + 1
        err = max(err, maxval(abs(zE(:, :, k) + ks * sin((k - 1) * dz))))
    end do
    print '(a,es10.3)', "Pade6stagg ddz_C2E complex vs closed form: ", err
    if (err > 1d-12) error stop 53

    ! projection of (sin x + sin y, 0, 0) keeps sin y and is divergence free
    do j = 1, m
        do i = 1, m
            g(i, j, :) = sin((i - 1) * dx) + sin((j - 1) * dx)
        end do
    end do
    gE = 0
    call c_f_pointer(pe, ed, [m, m, m + 1])
    call pdo_to_device(pa, c_loc(g), rb)
    call pdo_to_device(pe, c_loc(gE), rbE)
    call spC%fft(ad, cu)
    call pdo_to_device(pa, c_loc(gE), rb)
    call spC%fft(ad, cv)
    call spE%fft(ed, cw)
    call pp%init(dx, dx, dz, spC, spE, .false., 2 * pi, .false., gpC, p6, .true.)
    call pp%PressureProjection(cu, cv, cw)
    call spC%ifft(cu, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = 0
    do j = 1, m
        err = max(err, maxval(abs(gr(:, j, :) - sin((j - 1) * dx))))
    end do
    print '(a,es10.3)', "padepoisson PressureProjection of sin(x) + sin(y): ", err
    if (err > 1d-12) error stop 54
    call pp%DivergenceCheck(cu, cv, cw, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = maxval(abs(gr))
    print '(a,es10.3)', "padepoisson DivergenceCheck after projection: ", err
    if (err > 1d-12) error stop 55
    call pp%destroy()

    ! walls in z (PeriodicInZ false): sin(y) survives again, w =
    ! cos(2y) sin(pi z / Lz) (zero on the walls, as the extension assumes)
    ! only adds divergence in the (0, 2) modes; w stays zero on both walls and
    ! DivergenceCheck (odd-ended ddz_E2C) vanishes
    call spW%init("x", m, m, m, dx, dx, dz, "four", "none", 2, fixOddball=.false., init_periodicInZ=.false.)
    do j = 1, m
        do i = 1, m
            g(i, j, :) = sin((i - 1) * dx) + sin((j - 1) * dx)
        end do
    end do
    do k = 1, m + 1
        do j = 1, m
            gE(:, j, k) = cos(2 * (j - 1) * dx) * sin(pi * (k - 1) / m)
        end do
    end do
    gr = 0
    call pdo_to_device(pa, c_loc(g), rb)
    call spW%fft(ad, cu)
    call pdo_to_device(pa, c_loc(gr), rb)
    call spW%fft(ad, cv)
    call pdo_to_device(pe, c_loc(gE), rbE)
    call spE%fft(ed, cw)
    call pp%init(dx, dx, dz, spW, spE, .false., 2 * pi, .false., gpC, p6, .false.)
    call pp%PressureProjection(cu, cv, cw)
    call spW%ifft(cu, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    call pdo_to_host(c_loc(zE), pcw, cbE)
    err = 0
    do j = 1, m
        err = max(err, maxval(abs(gr(:, j, :) - sin((j - 1) * dx))))
    end do
    print '(a,es10.3)', "padepoisson with walls: PressureProjection of sin(x) + sin(y): ", err
    if (err > 1d-12) error stop 56
    err = max(maxval(abs(zE(:, :, 1))), maxval(abs(zE(:, :, m + 1))))
    print '(a,es10.3)', "padepoisson with walls: w on the walls: ", err
    if (err /= 0) error stop 57
    call pp%DivergenceCheck(cu, cv, cw, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = maxval(abs(gr))
    print '(a,es10.3)', "padepoisson with walls: DivergenceCheck after projection: ", err
    if (err > 1d-11) error stop 58
    call pp%destroy()

    ! getPressure / getPressureAndUpdateRHS with walls and the Stokes
    ! pressure: (sin x, 0, 0) is the gradient of -cos x, so both pressures are
    ! -cos x and the update leaves u = 0 (no wall velocity: no Stokes part)
    call pp%init(dx, dx, dz, spW, spE, .true., m * dz, .false., gpC, p6, .false.)
    do i = 1, m
        g(i, :, :) = sin((i - 1) * dx)
    end do
    gr = 0
    gE = 0
    call pdo_to_device(pa, c_loc(g), rb)
    call spW%fft(ad, cu)
    call pdo_to_device(pa, c_loc(gr), rb)
    call spW%fft(ad, cv)
    call pdo_to_device(pe, c_loc(gE), rbE)
    call spE%fft(ed, cw)
    call pp%getPressure(cu, cv, cw, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = 0
    do i = 1, m
        err = max(err, maxval(abs(gr(i, :, :) + cos((i - 1) * dx))))
    end do
    print '(a,es10.3)', "padepoisson with walls + Stokes: getPressure of (sin x, 0, 0) vs -cos x: ", err
    if (err > 1d-12) error stop 59
    call pp%getPressureAndUpdateRHS(cu, cv, cw, bd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(gr), pb, rb)
    err = 0
    do i = 1, m
        err = max(err, maxval(abs(gr(i, :, :) + cos((i - 1) * dx))))
    end do
    call pdo_to_host(c_loc(zr), pcu, cb)
    err = max(err, maxval(abs(zr(:, :, 1:m))))
    print '(a,es10.3)', "padepoisson with walls + Stokes: getPressureAndUpdateRHS (pressure, u): ", err
    if (err > 1d-12) error stop 60
    call pp%destroy()
    call spW%destroy()
    call p6%destroy()
    call spC%destroy()
    call spE%destroy()

    ! ---------------------------------------------------------------- cgrid
    ! entropy wave rho = 1 + 0.1 sin x, u = 0.5, p = 1 on a periodic
    ! 32 x 16 x 16 grid (cd10, LAD on, inviscid): d(rho)/dt = -0.05 cos x,
    ! d(rho u)/dt = -0.025 cos x, the v and w equations vanish (the artificial
    ! conductivity acts on the energy equation only)
    cgp%nx = 2 * m; cgp%ny = m; cgp%nz = m
    cgp%dx = 2 * pi / (2 * m); cgp%dy = 2 * pi / m; cgp%dz = 2 * pi / m
    ierr = cg%init(cgp)
    if (ierr /= 0) error stop 61
    allocate(q(2 * m, m, m, 5), qr(2 * m, m, m, 5))
    do i = 1, 2 * m
        x = (i - 1) * cgp%dx
        q(i, :, :, 1) = 1.d0 + 0.1d0 * sin(x)
        q(i, :, :, 2) = 0.5d0 * q(i, :, :, 1)
        q(i, :, :, 3) = 0
        q(i, :, :, 4) = 0
        q(i, :, :, 5) = 1.d0 / (cgp%gam - 1.d0) + 0.125d0 * q(i, :, :, 1)
    end do
    bytes = int(8, c_size_t) * size(q)
    pW = pdo_alloc(bytes)
    pR = pdo_alloc(bytes)
    call c_f_pointer(pW, Wd, [2 * m, m, m, 5])
    call c_f_pointer(pR, Rd, [2 * m, m, m, 5])
    call pdo_to_device(pW, c_loc(q), bytes)
    call cg%getRHS(Wd, Rd)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(qr), pR, bytes)
    err = 0
    do i = 1, 2 * m
        x = (i - 1) * cgp%dx
        err = max(err, maxval(abs(qr(i, :, :, 1) + 0.05d0 * cos(x))), maxval(abs(qr(i, :, :, 2) + 0.025d0 * cos(x))))
    end do
    err = max(err, maxval(abs(qr(:, :, :, 3:4))))
    print '(a,es10.3)', "cgrid getRHS of an entropy wave vs -0.05 cos x, -0.025 cos x, 0, 0: ", err
    if (err > 1d-9) error stop 62
    ! filter and the Gaussian filter keep a constant
    q(:, :, :, 1) = 2.5d0
    call pdo_to_device(pW, c_loc(q), int(8, c_size_t) * 2 * m * m * m)
    call c_f_pointer(pW, Fq, [2 * m, m, m])
    call cg%filter(Fq)
    call cg%filter(Fq, numtimes=2, gaussian=.true.)
    call pdo_synchronize(c_null_ptr)
    call pdo_to_host(c_loc(qr), pW, int(8, c_size_t) * 2 * m * m * m)
    err = maxval(abs(qr(:, :, :, 1) - 2.5d0))
    print '(a,es10.3)', "cgrid filter (fil, then gfil twice) of a constant: ", err
    if (err > 1d-13) error stop 63
    call cg%destroy()
    call pdo_release(pW); call pdo_release(pR)

    call pdo_release(pa); call pdo_release(pb); call pdo_release(pe)
    call pdo_release(pcu); call pdo_release(pcv); call pdo_release(pcw); call pdo_release(pct)
    print '(a)', "PASS"
end program
