! ISO_C_BINDING shim: the reference's Fortran derivative / filter / Poisson
! types over the C ABI of include/padeops_hip.h.  A solver switches with
!     use padeops_hip_mod, only: cd06, cd10, cf90, cd06stagg, poissonperiodic
! in place of the modules of src/derivatives/cd06.F90, cd10.F90,
! cd06stagg.F90, src/filters/cf90.F90 and src/utilities/PoissonPeriodic.F90.
! Field arguments are device arrays (allocated with pdo_alloc below and
! viewed through c_f_pointer); the shim passes their addresses, it never
! reads them.  init returns the reference's ierr; a bad boundary flag stops
! with code 324 like the reference's GracefulExit.
module padeops_hip_mod
    use iso_c_binding
    implicit none
    private

    integer, parameter, public :: PDO_F64 = 0, PDO_F32_REFINE = 1, PDO_F32_STORE = 2, PDO_F32 = 3
    integer, parameter :: rkind = c_double

    public :: pdo_alloc, pdo_release, pdo_to_device, pdo_to_host, pdo_synchronize

    interface
        function c_last_error() bind(C, name="pdo_last_error")
            import :: c_ptr
            type(c_ptr) :: c_last_error
        end function
        integer(c_int) function c_malloc(p, bytes) bind(C, name="pdo_malloc")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr) :: p
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function c_free(p) bind(C, name="pdo_free")
            import :: c_ptr, c_int
            type(c_ptr), value :: p
        end function
        integer(c_int) function c_h2d(dst, src, bytes) bind(C, name="pdo_memcpy_h2d")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function c_d2h(dst, src, bytes) bind(C, name="pdo_memcpy_d2h")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function c_sync(stream) bind(C, name="pdo_sync")
            import :: c_ptr, c_int
            type(c_ptr), value :: stream
        end function
        integer(c_int) function c_op_destroy(h) bind(C, name="pdo_op_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function
        ! cd06 / cd10: init(h, n, dx, periodic, bc1, bcn)
        integer(c_int) function c_cd06_init(h, n, dx, periodic, bc1, bcn) bind(C, name="pdo_cd06_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n, periodic, bc1, bcn
            real(c_double), value :: dx
        end function
        integer(c_int) function c_cd10_init(h, n, dx, periodic, bc1, bcn) bind(C, name="pdo_cd10_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n, periodic, bc1, bcn
            real(c_double), value :: dx
        end function
        integer(c_int) function c_cf90_init(h, n, periodic) bind(C, name="pdo_cf90_init")
            import :: c_ptr, c_int
            type(c_ptr) :: h
            integer(c_int), value :: n, periodic
        end function
        integer(c_int) function c_stagg_init(h, n, dz) bind(C, name="pdo_cd06stagg_init_periodic")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n
            real(c_double), value :: dz
        end function
        integer(c_int) function c_poisson_init(h, nx, ny, nz, dx, dy, dz, mx, my, mz, zp) &
                bind(C, name="pdo_poisson_periodic_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: nx, ny, nz, mx, my, mz, zp
            real(c_double), value :: dx, dy, dz
        end function
        integer(c_int) function c_poisson_solve(h, prec, rhs, f, stream) bind(C, name="pdo_poisson_periodic_solve")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, rhs, f, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_poisson_destroy(h) bind(C, name="pdo_poisson_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function
    end interface

    type, public :: cd06
        type(c_ptr) :: h = c_null_ptr
        integer :: n = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => cd06_init
        procedure :: dd1 => cd06_dd1
        procedure :: dd2 => cd06_dd2
        procedure :: dd3 => cd06_dd3
        procedure :: d2d1 => cd06_d2d1
        procedure :: d2d2 => cd06_d2d2
        procedure :: d2d3 => cd06_d2d3
        procedure :: destroy => cd06_destroy
    end type

    type, public :: cd10
        type(c_ptr) :: h = c_null_ptr
        integer :: n = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => cd10_init
        procedure :: dd1 => cd10_dd1
        procedure :: dd2 => cd10_dd2
        procedure :: dd3 => cd10_dd3
        procedure :: d2d1 => cd10_d2d1
        procedure :: d2d2 => cd10_d2d2
        procedure :: d2d3 => cd10_d2d3
        procedure :: destroy => cd10_destroy
    end type

    type, public :: cf90
        type(c_ptr) :: h = c_null_ptr
        integer :: n = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => cf90_init
        procedure :: filter1 => cf90_filter1
        procedure :: filter2 => cf90_filter2
        procedure :: filter3 => cf90_filter3
        procedure :: destroy => cf90_destroy
    end type

    type, public :: cd06stagg
        type(c_ptr) :: h = c_null_ptr
        integer :: n = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => stagg_init
        procedure :: ddz_E2C => stagg_ddz_E2C
        procedure :: ddz_C2E => stagg_ddz_C2E
        procedure :: d2dz2_C2C => stagg_d2dz2_C2C
        procedure :: d2dz2_E2E => stagg_d2dz2_E2E
        procedure :: InterpZ_E2C => stagg_interpz_E2C
        procedure :: InterpZ_C2E => stagg_interpz_C2E
        procedure :: destroy => stagg_destroy
    end type

    type, public :: poissonperiodic
        type(c_ptr) :: h = c_null_ptr
        integer :: nx = 0, ny = 0, nz = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => poisson_init
        procedure :: poisson_solve => poisson_solve_oop
        procedure :: poisson_solve_inplace
        procedure :: destroy => poisson_destroy
    end type

    interface
        integer(c_int) function pdo_cd06_dd1(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_dd1")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd06_dd2(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_dd2")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd06_dd3(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_dd3")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd06_d2d1(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_d2d1")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd06_d2d2(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_d2d2")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd06_d2d3(h, prec, f, df, na, nb, stream) bind(C, name="pdo_cd06_d2d3")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb
        end function
        integer(c_int) function pdo_cd10_dd1(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_dd1")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd10_dd2(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_dd2")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd10_dd3(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_dd3")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd10_d2d1(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_d2d1")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd10_d2d2(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_d2d2")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd10_d2d3(h, prec, f, df, na, nb, bc1, bcn, stream) bind(C, name="pdo_cd10_d2d3")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cf90_filter1(h, prec, f, df, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_cf90_filter1")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cf90_filter2(h, prec, f, df, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_cf90_filter2")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cf90_filter3(h, prec, f, df, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_cf90_filter3")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_cd06stagg_ddz_E2C(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_ddz_E2C")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_ddz_C2E(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_ddz_C2E")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_d2dz2_C2C(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_d2dz2_C2C")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_d2dz2_E2E(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_d2dz2_E2E")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_interpz_E2C(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_interpz_E2C")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_interpz_C2E(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_interpz_C2E")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        ! explicit filters, derivatives / filters dispatch, operators module
        integer(c_int) function pdo_gaussian_init(h, n, periodic) &
                bind(C, name="pdo_gaussian_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n, periodic
        end function
        integer(c_int) function pdo_lstsq_init(h, n, periodic) &
                bind(C, name="pdo_lstsq_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n, periodic
        end function
        integer(c_int) function pdo_xfilter_filter1(h, prec, f, fil, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_xfilter_filter1")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, fil, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_xfilter_filter2(h, prec, f, fil, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_xfilter_filter2")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, fil, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_xfilter_filter3(h, prec, f, fil, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_xfilter_filter3")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, fil, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_xfilter_filter3_cmplx(h, prec, f, fil, na, nb, bc1, bcn, stream) &
                bind(C, name="pdo_xfilter_filter3_cmplx")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, fil, stream
            integer(c_int), value :: prec, na, nb, bc1, bcn
        end function
        integer(c_int) function pdo_xfilter_destroy(h) &
                bind(C, name="pdo_xfilter_destroy")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
        end function
        integer(c_int) function pdo_derivatives_init(h, nx, ny, nz, dx, dy, dz, periodic, method) &
                bind(C, name="pdo_derivatives_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: nx, ny, nz
            real(c_double), value :: dx, dy, dz
            integer(c_int), dimension(3) :: periodic, method
        end function
        integer(c_int) function pdo_derivatives_ddx(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_ddx")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_ddy(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_ddy")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_ddz(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_ddz")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_d2dx2(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_d2dx2")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_d2dy2(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_d2dy2")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_d2dz2(h, prec, f, df, bc1, bcn, stream) &
                bind(C, name="pdo_derivatives_d2dz2")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_derivatives_destroy(h) &
                bind(C, name="pdo_derivatives_destroy")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
        end function
        integer(c_int) function pdo_filters_init(h, nx, ny, nz, periodic, method) &
                bind(C, name="pdo_filters_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: nx, ny, nz
            integer(c_int), dimension(3) :: periodic, method
        end function
        integer(c_int) function pdo_filters_filterx(h, prec, f, ff, bc1, bcn, stream) &
                bind(C, name="pdo_filters_filterx")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, ff, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_filters_filtery(h, prec, f, ff, bc1, bcn, stream) &
                bind(C, name="pdo_filters_filtery")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, ff, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_filters_filterz(h, prec, f, ff, bc1, bcn, stream) &
                bind(C, name="pdo_filters_filterz")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h, f, ff, stream
            integer(c_int), value :: prec, bc1, bcn
        end function
        integer(c_int) function pdo_filters_destroy(h) &
                bind(C, name="pdo_filters_destroy")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
        end function
        integer(c_int) function pdo_operators_gradient(der, prec, f, dfdx, dfdy, dfdz, bc, stream) &
                bind(C, name="pdo_operators_gradient")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: der, f, dfdx, dfdy, dfdz, stream
            integer(c_int), value :: prec
            integer(c_int), dimension(6) :: bc
        end function
        integer(c_int) function pdo_operators_curl(der, prec, u, v, w, curl, bc, stream) &
                bind(C, name="pdo_operators_curl")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: der, u, v, w, curl, stream
            integer(c_int), value :: prec
            integer(c_int), dimension(6) :: bc
        end function
        integer(c_int) function pdo_operators_divergence(der, prec, u, v, w, div, bc, stream) &
                bind(C, name="pdo_operators_divergence")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: der, u, v, w, div, stream
            integer(c_int), value :: prec
            integer(c_int), dimension(6) :: bc
        end function
        integer(c_int) function pdo_operators_filter3d(fil, prec, arr, numtimes, bc, stream) &
                bind(C, name="pdo_operators_filter3d")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: fil, arr, stream
            integer(c_int), value :: prec, numtimes
            integer(c_int), dimension(6) :: bc
        end function
    end interface

contains

    ! ---- errors ---------------------------------------------------------------
    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        character(len=256), pointer :: msg
        type(c_ptr) :: p
        integer :: k
        if (rc == 0) return
        p = c_last_error()
        call c_f_pointer(p, msg)
        k = index(msg, c_null_char) - 1
        if (k < 0) k = len(msg)
        if (rc == 324) then
            ! GracefulExit(..., 324) of the reference (exits.F90:28-35)
            print '(a)', "ERROR: Incorrect boundary specification: " // msg(1:k)
            error stop 324
        end if
        print '(a,i0,a)', "padeops_hip error ", rc, ": " // msg(1:k)
        error stop 1
    end subroutine

    ! ---- device memory --------------------------------------------------------
    type(c_ptr) function pdo_alloc(bytes) result(p)
        integer(c_size_t), intent(in) :: bytes
        p = c_null_ptr
        call check(c_malloc(p, bytes))
    end function

    subroutine pdo_release(p)
        type(c_ptr), intent(in) :: p
        call check(c_free(p))
    end subroutine

    subroutine pdo_to_device(dst, src, bytes)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: bytes
        call check(c_h2d(dst, src, bytes))
    end subroutine

    subroutine pdo_to_host(dst, src, bytes)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: bytes
        call check(c_d2h(dst, src, bytes))
    end subroutine

    subroutine pdo_synchronize(stream)
        type(c_ptr), intent(in) :: stream
        call check(c_sync(stream))
    end subroutine

    integer function flag(l)
        logical, intent(in) :: l
        flag = merge(1, 0, l)
    end function

    ! ---- cd06 (cd06.F90:129-839) ---------------------------------------------
    function cd06_init(this, n, dx, periodic, bc1, bcn) result(ierr)
        class(cd06), intent(inout) :: this
        integer, intent(in) :: n, bc1, bcn
        real(rkind), intent(in) :: dx
        logical, intent(in) :: periodic
        integer :: ierr
        this%n = n
        ierr = c_cd06_init(this%h, n, dx, flag(periodic), bc1, bcn)
        if (ierr < 0) call check(ierr)
    end function

    subroutine cd06_dd1(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in), target :: f
        real(rkind), dimension(this%n, na, nb), intent(out), target :: df
        call check(pdo_cd06_dd1(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_dd2(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in), target :: f
        real(rkind), dimension(na, this%n, nb), intent(out), target :: df
        call check(pdo_cd06_dd2(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_dd3(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in), target :: f
        real(rkind), dimension(na, nb, this%n), intent(out), target :: df
        call check(pdo_cd06_dd3(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_d2d1(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in), target :: f
        real(rkind), dimension(this%n, na, nb), intent(out), target :: df
        call check(pdo_cd06_d2d1(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_d2d2(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in), target :: f
        real(rkind), dimension(na, this%n, nb), intent(out), target :: df
        call check(pdo_cd06_d2d2(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_d2d3(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in), target :: f
        real(rkind), dimension(na, nb, this%n), intent(out), target :: df
        call check(pdo_cd06_d2d3(this%h, this%prec, c_loc(f), c_loc(df), na, nb, this%stream))
    end subroutine

    subroutine cd06_destroy(this)
        class(cd06), intent(inout) :: this
        if (c_associated(this%h)) call check(c_op_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---- cd10 (cd10.F90:195-2447) --------------------------------------------
    function cd10_init(this, n, dx, periodic, bc1, bcn) result(ierr)
        class(cd10), intent(inout) :: this
        integer, intent(in) :: n, bc1, bcn
        real(rkind), intent(in) :: dx
        logical, intent(in) :: periodic
        integer :: ierr
        this%n = n
        ierr = c_cd10_init(this%h, n, dx, flag(periodic), bc1, bcn)
        if (ierr < 0) call check(ierr)
    end function

    subroutine cd10_dd1(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in), target :: f
        real(rkind), dimension(this%n, na, nb), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_dd1(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_dd2(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in), target :: f
        real(rkind), dimension(na, this%n, nb), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_dd2(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_dd3(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in), target :: f
        real(rkind), dimension(na, nb, this%n), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_dd3(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_d2d1(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in), target :: f
        real(rkind), dimension(this%n, na, nb), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_d2d1(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_d2d2(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in), target :: f
        real(rkind), dimension(na, this%n, nb), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_d2d2(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_d2d3(this, f, df, na, nb, bc1, bcn)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in), target :: f
        real(rkind), dimension(na, nb, this%n), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cd10_d2d3(this%h, this%prec, c_loc(f), c_loc(df), na, nb, opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine cd10_destroy(this)
        class(cd10), intent(inout) :: this
        if (c_associated(this%h)) call check(c_op_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    integer(c_int) function opt0(v)
        integer, optional, intent(in) :: v
        opt0 = 0
        if (present(v)) opt0 = v
    end function

    ! ---- cf90 (cf90.F90:107-1251) --------------------------------------------
    function cf90_init(this, n, periodic) result(ierr)
        class(cf90), intent(inout) :: this
        integer, intent(in) :: n
        logical, intent(in) :: periodic
        integer :: ierr
        this%n = n
        ierr = c_cf90_init(this%h, n, flag(periodic))
        if (ierr < 0) call check(ierr)
    end function

    subroutine cf90_filter1(this, f, fil, na, nb, bc1, bcn)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in), target :: f
        real(rkind), dimension(this%n, na, nb), intent(out), target :: fil
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cf90_filter1(this%h, this%prec, c_loc(f), c_loc(fil), na, nb, opt0(bc1), opt0(bcn), &
                                    this%stream))
    end subroutine

    subroutine cf90_filter2(this, f, fil, na, nb, bc1, bcn)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in), target :: f
        real(rkind), dimension(na, this%n, nb), intent(out), target :: fil
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cf90_filter2(this%h, this%prec, c_loc(f), c_loc(fil), na, nb, opt0(bc1), opt0(bcn), &
                                    this%stream))
    end subroutine

    subroutine cf90_filter3(this, f, fil, na, nb, bc1, bcn)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in), target :: f
        real(rkind), dimension(na, nb, this%n), intent(out), target :: fil
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_cf90_filter3(this%h, this%prec, c_loc(f), c_loc(fil), na, nb, opt0(bc1), opt0(bcn), &
                                    this%stream))
    end subroutine

    subroutine cf90_destroy(this)
        class(cf90), intent(inout) :: this
        if (c_associated(this%h)) call check(c_op_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---- cd06stagg, periodic (cd06stagg.F90:170-195, 820-1059) ---------------
    ! real fields; complex fields are passed as real arrays of twice the first
    ! extent with n1 the complex extent (is_complex = 1 in the C ABI)
    function stagg_init(this, n, dz) result(ierr)
        class(cd06stagg), intent(inout) :: this
        integer, intent(in) :: n
        real(rkind), intent(in) :: dz
        integer :: ierr
        this%n = n
        ierr = c_stagg_init(this%h, n, dz)
        if (ierr < 0) call check(ierr)
    end function

    subroutine stagg_ddz_E2C(this, fE, dfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n + 1), intent(in), target :: fE
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: dfC
        call check(pdo_cd06stagg_ddz_E2C(this%h, this%prec, c_loc(fE), c_loc(dfC), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_ddz_C2E(this, fC, dfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fC
        real(rkind), dimension(n1, n2, this%n + 1), intent(out), target :: dfE
        call check(pdo_cd06stagg_ddz_C2E(this%h, this%prec, c_loc(fC), c_loc(dfE), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_d2dz2_C2C(this, fC, d2fC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fC
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: d2fC
        call check(pdo_cd06stagg_d2dz2_C2C(this%h, this%prec, c_loc(fC), c_loc(d2fC), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_d2dz2_E2E(this, fE, d2fE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n + 1), intent(in), target :: fE
        real(rkind), dimension(n1, n2, this%n + 1), intent(out), target :: d2fE
        call check(pdo_cd06stagg_d2dz2_E2E(this%h, this%prec, c_loc(fE), c_loc(d2fE), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_interpz_E2C(this, fE, ifC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n + 1), intent(in), target :: fE
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: ifC
        call check(pdo_cd06stagg_interpz_E2C(this%h, this%prec, c_loc(fE), c_loc(ifC), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_interpz_C2E(this, fC, ifE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fC
        real(rkind), dimension(n1, n2, this%n + 1), intent(out), target :: ifE
        call check(pdo_cd06stagg_interpz_C2E(this%h, this%prec, c_loc(fC), c_loc(ifE), n1, n2, 0, this%stream))
    end subroutine

    subroutine stagg_destroy(this)
        class(cd06stagg), intent(inout) :: this
        if (c_associated(this%h)) call check(c_op_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---- PoissonPeriodic (PoissonPeriodic.F90:37-212), x-decomposed ---------
    ! modk: 0 Fourier, 1 cd06, 2 cd06stagg, 3 cd10 (the Get_ModK callbacks)
    function poisson_init(this, dx, dy, dz, nx, ny, nz, modk, zero_policy) result(ierr)
        class(poissonperiodic), intent(inout) :: this
        real(rkind), intent(in) :: dx, dy, dz
        integer, intent(in) :: nx, ny, nz
        integer, dimension(3), optional, intent(in) :: modk
        integer, optional, intent(in) :: zero_policy
        integer :: ierr, mk(3), zp
        mk = 0
        if (present(modk)) mk = modk
        zp = 0
        if (present(zero_policy)) zp = zero_policy
        this%nx = nx; this%ny = ny; this%nz = nz
        ierr = c_poisson_init(this%h, nx, ny, nz, dx, dy, dz, mk(1), mk(2), mk(3), zp)
        if (ierr < 0) call check(ierr)
    end function

    subroutine poisson_solve_oop(this, rhs, f)
        class(poissonperiodic), intent(in) :: this
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(in), target :: rhs
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(out), target :: f
        call check(c_poisson_solve(this%h, this%prec, c_loc(rhs), c_loc(f), this%stream))
    end subroutine

    subroutine poisson_solve_inplace(this, rhs)
        class(poissonperiodic), intent(in) :: this
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(inout), target :: rhs
        call check(c_poisson_solve(this%h, this%prec, c_loc(rhs), c_loc(rhs), this%stream))
    end subroutine

    subroutine poisson_destroy(this)
        class(poissonperiodic), intent(inout) :: this
        if (c_associated(this%h)) call check(c_poisson_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

end module padeops_hip_mod
