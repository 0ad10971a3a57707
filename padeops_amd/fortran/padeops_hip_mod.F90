! ISO_C_BINDING shim: the reference's Fortran derivative / filter / Poisson
! types over the C ABI of include/padeops_hip.h.  A solver switches with
!     use padeops_hip_mod, only: cd06, cd10, cf90, cd06stagg, poissonperiodic, &
!                                derivatives, filters, spectral, Pade6stagg, padepoisson, &
!                                cgrid, cgrid_params
! in place of the modules of src/derivatives/cd06.F90, cd10.F90,
! cd06stagg.F90, derivatives.F90, src/filters/cf90.F90, filters.F90,
! src/utilities/PoissonPeriodic.F90 and src/incompressible/spectral.F90,
! PadeDerOps.F90, PadePoisson.F90 (and cgrid's right-hand side and filter,
! src/compressible/cgrid.F90), and uses padeops_hip_decomp for the
! decomp_info / transpose_* of 2DECOMP on one GPU.
! Field arguments are device arrays (allocated with pdo_alloc below and
! viewed through c_f_pointer); the shim passes their addresses, it never
! reads them.  init returns the reference's ierr; a bad boundary flag stops
! with code 324 like the reference's GracefulExit.
module padeops_hip_mod
    use iso_c_binding
    use padeops_hip_decomp, only: decomp_info
    implicit none
    private

    integer, parameter, public :: PDO_F64 = 0, PDO_F32_REFINE = 1, PDO_F32_STORE = 2, PDO_F32 = 3
    integer, parameter :: rkind = c_double

    public :: pdo_alloc, pdo_release, pdo_to_device, pdo_to_host, pdo_synchronize
    ! Pade6stagg schemes (PadeDerOps.F90:16-18: fd02 = 0, cd06 = 1,
    ! fourierColl = 2; the name cd06 is taken by the derivative type here)
    integer, parameter, public :: fd02 = 0, cd06_scheme = 1, fourierColl = 2

    ! Get_ModK callback of PoissonPeriodic%init (PoissonPeriodic.F90:11-16)
    abstract interface
        subroutine GetModifiedWavenum(kinout)
            import :: c_double
            real(c_double), dimension(:), intent(inout) :: kinout
        end subroutine
    end interface
    public :: GetModifiedWavenum

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
        integer(c_int) function c_stagg_init_np(h, n, dz, te, be, ts, bs) bind(C, name="pdo_cd06stagg_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: n, te, be, ts, bs
            real(c_double), value :: dz
        end function
        integer(c_int) function c_poisson_init_k(h, nx, ny, nz, kx, ky, kz, zp) &
                bind(C, name="pdo_poisson_periodic_init_k")
            import :: c_ptr, c_int
            type(c_ptr) :: h
            integer(c_int), value :: nx, ny, nz, zp
            type(c_ptr), value :: kx, ky, kz
        end function
        ! spectral, Pade6stagg, padepoisson
        integer(c_int) function c_sp_init(h, nx, ny, nz, dx, dy, dz, fixodd, dealias, perz) &
                bind(C, name="pdo_spectral_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: nx, ny, nz, fixodd, perz
            real(c_double), value :: dx, dy, dz, dealias
        end function
        integer(c_int) function c_sp_fft(h, prec, fin, fout, stream) bind(C, name="pdo_spectral_fft")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, fin, fout, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_sp_ifft(h, prec, fin, fout, oddball, stream) bind(C, name="pdo_spectral_ifft")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, fin, fout, stream
            integer(c_int), value :: prec, oddball
        end function
        integer(c_int) function c_sp_ik(h, prec, dir, fin, fout, stream) bind(C, name="pdo_spectral_mtimes_ik")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, fin, fout, stream
            integer(c_int), value :: prec, dir
        end function
        integer(c_int) function c_sp_dealias(h, prec, f, stream) bind(C, name="pdo_spectral_dealias")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_sp_dealias_edge(h, prec, f, stream) bind(C, name="pdo_spectral_dealias_edge")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_sp_destroy(h) bind(C, name="pdo_spectral_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_p6_init(h, nz, dz, periodic) bind(C, name="pdo_pade6stagg_init")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            integer(c_int), value :: nz, periodic
            real(c_double), value :: dz
        end function
        integer(c_int) function c_p6_apply(h, kind, bot, top, prec, fin, fout, n1, n2, is_complex, stream) &
                bind(C, name="pdo_pade6stagg_apply")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, fin, fout, stream
            integer(c_int), value :: kind, bot, top, prec, n1, n2, is_complex
        end function
        integer(c_int) function c_p6_destroy(h) bind(C, name="pdo_pade6stagg_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_pp_init(h, sp, dz) bind(C, name="pdo_padepoisson_init_periodic")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            type(c_ptr), value :: sp
            real(c_double), value :: dz
        end function
        integer(c_int) function c_pp_init_np(h, sp, dz) bind(C, name="pdo_padepoisson_init_nonperiodic")
            import :: c_ptr, c_int, c_double
            type(c_ptr) :: h
            type(c_ptr), value :: sp
            real(c_double), value :: dz
        end function
        integer(c_int) function c_pp_project(h, prec, uh, vh, wh, stream) bind(C, name="pdo_padepoisson_projection")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, uh, vh, wh, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_pp_div(h, prec, uh, vh, wh, div, stream) bind(C, name="pdo_padepoisson_divergence")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, uh, vh, wh, div, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_pp_set_stokes(h, Lz) bind(C, name="pdo_padepoisson_set_stokes")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            real(c_double), value :: Lz
        end function
        integer(c_int) function c_pp_pressure(h, prec, uh, vh, wh, p, upd, stream) bind(C, name="pdo_padepoisson_pressure")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, uh, vh, wh, p, stream
            integer(c_int), value :: prec, upd
        end function
        integer(c_int) function c_pp_destroy(h) bind(C, name="pdo_padepoisson_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
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
        integer(c_int) function c_cg_init(h, params, prec) bind(C, name="pdo_cgrid_init")
            import :: c_ptr, c_int
            type(c_ptr) :: h
            type(c_ptr), value :: params
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_cg_rhs(h, prec, W, rhs, stream) bind(C, name="pdo_cgrid_rhs")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, W, rhs, stream
            integer(c_int), value :: prec
        end function
        integer(c_int) function c_cg_filter(h, prec, arr, numtimes, gaussian, stream) bind(C, name="pdo_cgrid_filter")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, arr, stream
            integer(c_int), value :: prec, numtimes, gaussian
        end function
        integer(c_int) function c_cg_destroy(h) bind(C, name="pdo_cgrid_destroy")
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
        integer :: n = 0, nE = 0
        logical :: AmIPeriodic = .false.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure, private :: init_periodic => stagg_init_periodic
        procedure, private :: init_nonperiodic => stagg_init_nonperiodic
        generic :: init => init_periodic, init_nonperiodic
        procedure, private :: ddz_E2C_REAL => stagg_ddz_E2C_r
        procedure, private :: ddz_C2E_REAL => stagg_ddz_C2E_r
        procedure, private :: ddz_C2C_REAL => stagg_ddz_C2C_r
        procedure, private :: ddz_E2E_REAL => stagg_ddz_E2E_r
        procedure, private :: d2dz2_C2C_REAL => stagg_d2dz2_C2C_r
        procedure, private :: d2dz2_E2E_REAL => stagg_d2dz2_E2E_r
        procedure, private :: InterpZ_E2C_REAL => stagg_interpz_E2C_r
        procedure, private :: InterpZ_C2E_REAL => stagg_interpz_C2E_r
        procedure, private :: ddz_E2C_CMPLX => stagg_ddz_E2C_c
        procedure, private :: ddz_C2E_CMPLX => stagg_ddz_C2E_c
        procedure, private :: ddz_C2C_CMPLX => stagg_ddz_C2C_c
        procedure, private :: ddz_E2E_CMPLX => stagg_ddz_E2E_c
        procedure, private :: d2dz2_C2C_CMPLX => stagg_d2dz2_C2C_c
        procedure, private :: d2dz2_E2E_CMPLX => stagg_d2dz2_E2E_c
        procedure, private :: InterpZ_E2C_CMPLX => stagg_interpz_E2C_c
        procedure, private :: InterpZ_C2E_CMPLX => stagg_interpz_C2E_c
        generic :: ddz_E2C => ddz_E2C_REAL, ddz_E2C_CMPLX
        generic :: ddz_C2E => ddz_C2E_REAL, ddz_C2E_CMPLX
        generic :: ddz_C2C => ddz_C2C_REAL, ddz_C2C_CMPLX
        generic :: ddz_E2E => ddz_E2E_REAL, ddz_E2E_CMPLX
        generic :: d2dz2_C2C => d2dz2_C2C_REAL, d2dz2_C2C_CMPLX
        generic :: d2dz2_E2E => d2dz2_E2E_REAL, d2dz2_E2E_CMPLX
        generic :: InterpZ_E2C => InterpZ_E2C_REAL, InterpZ_E2C_CMPLX
        generic :: InterpZ_C2E => InterpZ_C2E_REAL, InterpZ_C2E_CMPLX
        procedure :: destroy => stagg_destroy
    end type

    type, public :: poissonperiodic
        type(c_ptr) :: h = c_null_ptr
        integer :: dir_id = 1, nx_g = 0, ny_g = 0, nz_g = 0
        integer :: nx_in = 0, ny_in = 0, nz_in = 0
        integer :: nx_hat = 0, ny_hat = 0, nz_hat = 0
        real(rkind), dimension(:), allocatable :: kx, ky, kz
        logical :: doIhaveZeroWaveNum = .true.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => poisson_init
        procedure :: init_modk => poisson_init_modk
        procedure, private :: poisson_solve_inplace
        procedure, private :: poisson_solve_outofplace
        generic :: poisson_solve => poisson_solve_inplace, poisson_solve_outofplace
        procedure :: destroy => poisson_destroy
    end type

    ! derivatives (derivatives.F90:14-56): one GPU holds the whole box, so
    ! every pencil size xsz = ysz = zsz = (nx, ny, nz)
    type, public :: derivatives
        type(c_ptr) :: h = c_null_ptr
        integer, dimension(3) :: xsz = 0, ysz = 0, zsz = 0
        integer :: xmethod = 0, ymethod = 0, zmethod = 0
        logical :: initialized = .false.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure, private :: init_parallel => der_init_parallel
        procedure, private :: init_serial => der_init_serial
        generic :: init => init_parallel, init_serial
        procedure :: destroy => der_destroy
        procedure :: ddx => der_ddx
        procedure :: ddy => der_ddy
        procedure :: ddz => der_ddz
        procedure :: d2dx2 => der_d2dx2
        procedure :: d2dy2 => der_d2dy2
        procedure :: d2dz2 => der_d2dz2
        procedure :: getMethodx => der_getMethodx
        procedure :: getMethody => der_getMethody
        procedure :: getMethodz => der_getMethodz
    end type

    ! filters (filters.F90:14-38)
    type, public :: filters
        type(c_ptr) :: h = c_null_ptr
        integer, dimension(3) :: xsz = 0, ysz = 0, zsz = 0
        integer :: xmethod = 0, ymethod = 0, zmethod = 0
        logical :: initialized = .false.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure, private :: init_parallel => fil_init_parallel
        procedure, private :: init_serial => fil_init_serial
        generic :: init => init_parallel, init_serial
        procedure :: destroy => fil_destroy
        procedure :: filterx => fil_filterx
        procedure :: filtery => fil_filtery
        procedure :: filterz => fil_filterz
        procedure :: getMethodx => fil_getMethodx
        procedure :: getMethody => fil_getMethody
        procedure :: getMethodz => fil_getMethodz
    end type

    ! spectral as igrid uses it (spectral.F90:867-931 with dimTransform = 2,
    ! x pencil): physical (nx_r, ny_r, nz_r), spectral (nx_c, ny_c, nz_c) =
    ! (nx/2+1, ny, nz), spectdecomp their decomposition
    type, public :: spectral
        type(c_ptr) :: h = c_null_ptr
        integer :: nx_g = 0, ny_g = 0, nz_g = 0
        integer :: nx_r = 0, ny_r = 0, nz_r = 0
        integer :: nx_c = 0, ny_c = 0, nz_c = 0
        integer, dimension(3) :: fft_size = 0
        real(rkind) :: dx = 0, dy = 0, dz = 0, normfact = 0
        real(rkind) :: dealiasFact = 2.d0 / 3.d0
        logical :: fixOddball = .true., init_periodicInZ = .false., isInitialized = .false.
        type(decomp_info) :: spectdecomp
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => sp_init
        procedure :: destroy => sp_destroy
        procedure :: fft => sp_fft
        procedure :: ifft => sp_ifft
        procedure :: dealias => sp_dealias
        procedure :: dealias_edgeField => sp_dealias_edge
        procedure :: mTimes_ik1_oop => sp_ik1_oop
        procedure :: mTimes_ik1_ip => sp_ik1_ip
        procedure :: mTimes_ik2_oop => sp_ik2_oop
        procedure :: mTimes_ik2_ip => sp_ik2_ip
    end type

    ! Pade6stagg (PadeDerOps.F90:20-54), scheme cd06
    type, public :: Pade6stagg
        type(c_ptr) :: h = c_null_ptr
        type(decomp_info) :: gp, sp_gp
        integer :: scheme = cd06_scheme
        real(rkind) :: dz = 0
        logical :: isPeriodic = .false.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => p6_init
        procedure :: destroy => p6_destroy
        procedure, private :: ddz_C2E_real => p6_ddz_C2E_r
        procedure, private :: ddz_C2E_cmplx => p6_ddz_C2E_c
        procedure, private :: ddz_E2C_real => p6_ddz_E2C_r
        procedure, private :: ddz_E2C_cmplx => p6_ddz_E2C_c
        procedure, private :: interpz_C2E_real => p6_interpz_C2E_r
        procedure, private :: interpz_C2E_cmplx => p6_interpz_C2E_c
        procedure, private :: interpz_E2C_real => p6_interpz_E2C_r
        procedure, private :: interpz_E2C_cmplx => p6_interpz_E2C_c
        procedure, private :: d2dz2_C2C_real => p6_d2dz2_C2C_r
        procedure, private :: d2dz2_C2C_cmplx => p6_d2dz2_C2C_c
        procedure, private :: d2dz2_E2E_real => p6_d2dz2_E2E_r
        procedure, private :: d2dz2_E2E_cmplx => p6_d2dz2_E2E_c
        generic :: ddz_C2E => ddz_C2E_real, ddz_C2E_cmplx
        generic :: ddz_E2C => ddz_E2C_real, ddz_E2C_cmplx
        generic :: d2dz2_C2C => d2dz2_C2C_real, d2dz2_C2C_cmplx
        generic :: d2dz2_E2E => d2dz2_E2E_real, d2dz2_E2E_cmplx
        generic :: interpz_C2E => interpz_C2E_real, interpz_C2E_cmplx
        generic :: interpz_E2C => interpz_E2C_real, interpz_E2C_cmplx
    end type

    ! padepoisson (PadePoisson.F90:20-72), periodic in z or with walls
    type, public :: padepoisson
        type(c_ptr) :: h = c_null_ptr
        integer :: nx_in = 0, ny_in = 0, nz_in = 0, nxE_in = 0, nyE_in = 0, nzE_in = 0, nx = 0, ny = 0, nz = 0
        logical :: PeriodicInZ = .true.
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => pp_init
        procedure :: PressureProjection => pp_projection
        procedure :: getPressure => pp_get_pressure
        procedure :: getPressureAndUpdateRHS => pp_get_pressure_update
        procedure :: DivergenceCheck => pp_divergence
        procedure :: destroy => pp_destroy
    end type

    ! The settings of cgrid%init (cgrid.F90:113-760) that its right-hand side
    ! reads, one species of ideal gas, laid out as pdo_cgrid_params: the
    ! caller fills it from the values its input file gave (derivative 1 = cd10,
    ! 2 = cd06; filter 1 = cf90; periodic 0/1; x_bc, y_bc, z_bc the (bc1, bcn)
    ! of each direction; inviscid 1 = no physical viscosity, else the power law
    ! mu_ref (T/T_ref)^visc_n with Prandtl number Pr).
    type, bind(C), public :: cgrid_params
        integer(c_int) :: nx = 0, ny = 0, nz = 0
        real(c_double) :: dx = 0, dy = 0, dz = 0
        integer(c_int) :: periodic(3) = 1
        integer(c_int) :: x_bc(2) = 0, y_bc(2) = 0, z_bc(2) = 0
        integer(c_int) :: derivative(3) = 1
        integer(c_int) :: filter(3) = 1
        real(c_double) :: gam = 1.4d0, Rgas = 1.d0
        real(c_double) :: Cmu = 0.002d0, Cbeta = 1.75d0, Ckap = 0.01d0
        integer(c_int) :: inviscid = 1
        real(c_double) :: mu_ref = 1.d0, T_ref = 1.d0, visc_n = 0.75d0, Pr = 0.75d0
    end type

    ! cgrid (cgrid.F90:26-111): the right-hand side of the conserved variables
    ! (getRHS :822-894 with get_primitive, getRHS_x/y/z, getLAD, get_tau and
    ! get_q) and filter (:1259-1336).  The reference's getRHS is private to its
    ! advance_RK45; a solver calls this one there, on device arrays
    ! W, rhs (nxp, nyp, nzp, 5) = rho, rho u, rho v, rho w, TE.
    type, public :: cgrid
        type(c_ptr) :: h = c_null_ptr
        integer :: nxp = 0, nyp = 0, nzp = 0
        integer :: prec = PDO_F64
        type(c_ptr) :: stream = c_null_ptr
    contains
        procedure :: init => cg_init
        procedure :: getRHS => cg_get_rhs
        procedure :: filter => cg_filter
        procedure :: destroy => cg_destroy
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
        integer(c_int) function pdo_cd06stagg_ddz_C2C(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_ddz_C2C")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_cd06stagg_ddz_E2E(h, prec, f, df, n1, n2, is_complex, stream) &
                bind(C, name="pdo_cd06stagg_ddz_E2E")
            import :: c_ptr, c_int
            type(c_ptr), value :: h, f, df, stream
            integer(c_int), value :: prec, n1, n2, is_complex
        end function
        integer(c_int) function pdo_derivatives_method(h, dir, method) bind(C, name="pdo_derivatives_method")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
            integer(c_int), value :: dir
            integer(c_int) :: method
        end function
        integer(c_int) function pdo_filters_method(h, dir, method) bind(C, name="pdo_filters_method")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
            integer(c_int), value :: dir
            integer(c_int) :: method
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

    ! ---- cd06stagg (cd06stagg.F90:170-265 init, 820-1059 generics) -----------
    ! init_periodic / init_nonperiodic are subroutines like the reference's;
    ! a bad size stops through check().  Complex fields use is_complex = 1.
    subroutine stagg_init_periodic(this, nx, dx)
        class(cd06stagg), intent(inout) :: this
        integer, intent(in) :: nx
        real(rkind), intent(in) :: dx
        call stagg_destroy(this)
        this%n = nx; this%nE = nx + 1; this%AmIPeriodic = .true.
        call check(c_stagg_init(this%h, nx, dx))
    end subroutine

    subroutine stagg_init_nonperiodic(this, nx, dx, isTopEven, isBotEven, isTopSided, isBotSided)
        class(cd06stagg), intent(inout) :: this
        integer, intent(in) :: nx
        real(rkind), intent(in) :: dx
        logical, intent(in) :: isTopEven, isBotEven
        logical, intent(in), optional :: isTopSided, isBotSided
        logical :: ts, bs
        ts = .false.; bs = .false.
        if (present(isTopSided)) ts = isTopSided
        if (present(isBotSided)) bs = isBotSided
        call stagg_destroy(this)
        this%n = nx; this%nE = nx + 1; this%AmIPeriodic = .false.
        call check(c_stagg_init_np(this%h, nx, dx, flag(isTopEven), flag(isBotEven), flag(ts), flag(bs)))
    end subroutine

    subroutine stagg_go(this, which, fin, fout, n1, n2, cplx)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: which, n1, n2, cplx
        type(c_ptr), intent(in) :: fin, fout
        integer(c_int) :: rc
        select case (which)
        case (1); rc = pdo_cd06stagg_ddz_E2C(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (2); rc = pdo_cd06stagg_ddz_C2E(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (3); rc = pdo_cd06stagg_ddz_C2C(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (4); rc = pdo_cd06stagg_ddz_E2E(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (5); rc = pdo_cd06stagg_d2dz2_C2C(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (6); rc = pdo_cd06stagg_d2dz2_E2E(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case (7); rc = pdo_cd06stagg_interpz_E2C(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        case default; rc = pdo_cd06stagg_interpz_C2E(this%h, this%prec, fin, fout, n1, n2, cplx, this%stream)
        end select
        call check(rc)
    end subroutine

    subroutine stagg_ddz_E2C_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 1, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_ddz_E2C_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 1, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_ddz_C2E_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 2, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_ddz_C2E_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 2, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_ddz_C2C_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 3, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_ddz_C2C_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 3, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_ddz_E2E_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 4, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_ddz_E2E_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 4, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_d2dz2_C2C_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 5, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_d2dz2_C2C_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 5, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_d2dz2_E2E_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 6, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_d2dz2_E2E_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 6, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_interpz_E2C_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 7, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_interpz_E2C_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%n), intent(out), target :: fout
        call stagg_go(this, 7, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine
    subroutine stagg_interpz_C2E_r(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        real(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 8, c_loc(fin), c_loc(fout), n1, n2, 0)
    end subroutine
    subroutine stagg_interpz_C2E_c(this, fin, fout, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in), target :: fin
        complex(rkind), dimension(n1, n2, this%nE), intent(out), target :: fout
        call stagg_go(this, 8, c_loc(fin), c_loc(fout), n1, n2, 1)
    end subroutine

    subroutine stagg_destroy(this)
        class(cd06stagg), intent(inout) :: this
        if (c_associated(this%h)) call check(c_op_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---- PoissonPeriodic (PoissonPeriodic.F90:37-261) ------------------------
    ! init follows :114-212: wavenumbers of GetWaveNums, the (nx/2+1, ny, nz)
    ! spectral block of the x-decomposed r2c transform, each Get_ModK
    ! callback applied to k*d and divided back by d.  On one GPU every pencil
    ! holds the whole box, so dir_id 1, 2 and 3 give the same solve.
    pure function GetWaveNums(nx, dx) result(k)
        integer, intent(in) :: nx
        real(rkind), intent(in) :: dx
        real(rkind), dimension(nx) :: k
        real(rkind), parameter :: pi = 3.14159265358979323846264338327950288d0
        real(rkind), dimension(nx) :: t
        integer :: i, half
        do i = 1, nx
            t(i) = (-pi + (i - 1) * 2 * pi / real(nx - mod(nx, 2), rkind)) / dx
        end do
        half = nx / 2
        k = cshift(t, half)
    end function

    subroutine poisson_init(this, dx, dy, dz, gp, dir_id, useExhaustiveFFT, Get_ModKx, Get_ModKy, Get_ModKz)
        class(poissonperiodic), intent(inout) :: this
        real(rkind), intent(in) :: dx, dy, dz
        class(decomp_info), intent(in), target :: gp
        integer, intent(in) :: dir_id
        logical, intent(in), optional :: useExhaustiveFFT
        procedure(GetModifiedWavenum), optional :: Get_ModKx, Get_ModKy, Get_ModKz
        real(rkind), dimension(:), allocatable :: ktmp
        integer :: nx, ny, nz
        if (dir_id < 1 .or. dir_id > 3) then
            print '(a)', "ERROR: Incorrect option for DIR_ID"
            error stop 31243
        end if
        if (present(useExhaustiveFFT)) continue
        call poisson_destroy(this)
        nx = gp%xsz(1); ny = gp%ysz(2); nz = gp%zsz(3)
        this%dir_id = dir_id
        this%nx_g = nx; this%ny_g = ny; this%nz_g = nz
        this%nx_in = nx; this%ny_in = ny; this%nz_in = nz
        this%nx_hat = nx / 2 + 1; this%ny_hat = ny; this%nz_hat = nz
        if (allocated(this%kx)) deallocate(this%kx, this%ky, this%kz)
        allocate(this%kx(this%nx_hat), this%ky(ny), this%kz(nz))
        allocate(ktmp(nx)); ktmp = GetWaveNums(nx, dx); this%kx = ktmp(1:this%nx_hat); deallocate(ktmp)
        this%ky = GetWaveNums(ny, dy)
        this%kz = GetWaveNums(nz, dz)
        if (present(Get_ModKx)) then
            this%kx = this%kx * dx
            call Get_ModKx(this%kx)
            this%kx = this%kx / dx
        end if
        if (present(Get_ModKy)) then
            this%ky = this%ky * dy
            call Get_ModKy(this%ky)
            this%ky = this%ky / dy
        end if
        if (present(Get_ModKz)) then
            this%kz = this%kz * dz
            call Get_ModKz(this%kz)
            this%kz = this%kz / dz
        end if
        this%doIhaveZeroWaveNum = .true.
        call poisson_upload(this, 0)
    end subroutine

    subroutine poisson_upload(this, zp)
        class(poissonperiodic), intent(inout), target :: this
        integer, intent(in) :: zp
        call check(c_poisson_init_k(this%h, this%nx_g, this%ny_g, this%nz_g, c_loc(this%kx), c_loc(this%ky), &
                                    c_loc(this%kz), zp))
    end subroutine

    ! the same object with a built-in modified wavenumber per direction
    ! (0 Fourier, 1 cd06, 2 cd06stagg, 3 cd10) and the zero-mode policy
    ! (0 the reference's, 1 every |k'|^2 <= 1e-14 mode zeroed)
    function poisson_init_modk(this, dx, dy, dz, nx, ny, nz, modk, zero_policy) result(ierr)
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
        call poisson_destroy(this)
        this%nx_g = nx; this%ny_g = ny; this%nz_g = nz
        this%nx_in = nx; this%ny_in = ny; this%nz_in = nz
        this%nx_hat = nx / 2 + 1; this%ny_hat = ny; this%nz_hat = nz
        ierr = c_poisson_init(this%h, nx, ny, nz, dx, dy, dz, mk(1), mk(2), mk(3), zp)
        if (ierr < 0) call check(ierr)
    end function

    subroutine poisson_solve_outofplace(this, rhs, f)
        class(poissonperiodic), intent(inout) :: this
        real(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(in), target :: rhs
        real(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(out), target :: f
        call check(c_poisson_solve(this%h, this%prec, c_loc(rhs), c_loc(f), this%stream))
    end subroutine

    subroutine poisson_solve_inplace(this, rhs)
        class(poissonperiodic), intent(inout) :: this
        real(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(inout), target :: rhs
        call check(c_poisson_solve(this%h, this%prec, c_loc(rhs), c_loc(rhs), this%stream))
    end subroutine

    subroutine poisson_destroy(this)
        class(poissonperiodic), intent(inout) :: this
        if (c_associated(this%h)) call check(c_poisson_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---- derivatives (derivatives.F90:140-569) -------------------------------
    ! method strings as the reference's init_procedures (:276-390): "cd10" 1,
    ! "cd06" 2, "four" 3, "cheb" 4; anything else stops like its GracefulExit.
    ! The bc arguments reach cd10 only when both are present (:447-569).
    integer function der_method_id(m)
        character(len=*), intent(in) :: m
        select case (trim(adjustl(m)))
        case ("cd10"); der_method_id = 1
        case ("cd06"); der_method_id = 2
        case ("four"); der_method_id = 3
        case ("cheb"); der_method_id = 4
        case default
            print '(a)', "ERROR: Invalid derivative method " // trim(m)
            error stop 1
        end select
    end function

    subroutine der_init_serial(this, nx, ny, nz, dx, dy, dz, periodic_x, periodic_y, periodic_z, &
                               method_x, method_y, method_z, xmetric, ymetric, zmetric, curvilinear)
        class(derivatives), intent(inout) :: this
        integer, intent(in) :: nx, ny, nz
        real(rkind), intent(in) :: dx, dy, dz
        logical, intent(in) :: periodic_x, periodic_y, periodic_z
        character(len=*), intent(in) :: method_x, method_y, method_z
        logical, intent(in), optional :: xmetric, ymetric, zmetric, curvilinear
        integer(c_int) :: per(3), met(3)
        if (present(curvilinear)) then
            if (curvilinear) error stop "derivatives: curvilinear metrics are not supported"
        end if
        if (present(xmetric) .or. present(ymetric) .or. present(zmetric)) continue
        if (this%initialized) call der_destroy(this)
        this%xsz = [nx, ny, nz]; this%ysz = this%xsz; this%zsz = this%xsz
        per = [flag(periodic_x), flag(periodic_y), flag(periodic_z)]
        met = [der_method_id(method_x), der_method_id(method_y), der_method_id(method_z)]
        this%xmethod = met(1); this%ymethod = met(2); this%zmethod = met(3)
        call check(pdo_derivatives_init(this%h, nx, ny, nz, dx, dy, dz, per, met))
        this%initialized = .true.
    end subroutine

    subroutine der_init_parallel(this, gp, dx, dy, dz, periodic_x, periodic_y, periodic_z, &
                                 method_x, method_y, method_z, xmetric, ymetric, zmetric, curvilinear)
        class(derivatives), intent(inout) :: this
        class(decomp_info), intent(in) :: gp
        real(rkind), intent(in) :: dx, dy, dz
        logical, intent(in) :: periodic_x, periodic_y, periodic_z
        character(len=*), intent(in) :: method_x, method_y, method_z
        logical, intent(in), optional :: xmetric, ymetric, zmetric, curvilinear
        call der_init_serial(this, gp%xsz(1), gp%ysz(2), gp%zsz(3), dx, dy, dz, periodic_x, periodic_y, &
                             periodic_z, method_x, method_y, method_z, xmetric, ymetric, zmetric, curvilinear)
    end subroutine

    subroutine der_ddx(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(in), target :: f
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_ddx(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    subroutine der_ddy(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(in), target :: f
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_ddy(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    subroutine der_ddz(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(in), target :: f
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_ddz(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    subroutine der_d2dx2(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(in), target :: f
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_d2dx2(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    subroutine der_d2dy2(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(in), target :: f
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_d2dy2(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    subroutine der_d2dz2(this, f, df, bc1, bcn)
        class(derivatives), intent(in) :: this
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(in), target :: f
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(out), target :: df
        integer, optional, intent(in) :: bc1, bcn
        integer(c_int) :: b1, bn
        b1 = 0; bn = 0
        if (present(bc1) .and. present(bcn)) then
            b1 = bc1; bn = bcn
        end if
        call check(pdo_derivatives_d2dz2(this%h, this%prec, c_loc(f), c_loc(df), b1, bn, this%stream))
    end subroutine

    pure integer function der_getMethodx(this)
        class(derivatives), intent(in) :: this
        der_getMethodx = this%xmethod
    end function

    pure integer function der_getMethody(this)
        class(derivatives), intent(in) :: this
        der_getMethody = this%ymethod
    end function

    pure integer function der_getMethodz(this)
        class(derivatives), intent(in) :: this
        der_getMethodz = this%zmethod
    end function

    subroutine der_destroy(this)
        class(derivatives), intent(inout) :: this
        if (c_associated(this%h)) call check(pdo_derivatives_destroy(this%h))
        this%h = c_null_ptr
        this%initialized = .false.
    end subroutine

    ! ---- filters (filters.F90:82-269) ----------------------------------------
    integer function fil_method_id(m)
        character(len=*), intent(in) :: m
        select case (trim(adjustl(m)))
        case ("cf90"); fil_method_id = 1
        case ("gaussian"); fil_method_id = 2
        case ("lstsq"); fil_method_id = 3
        case default
            print '(a)', "ERROR: Invalid filter method " // trim(m)
            error stop 1
        end select
    end function

    subroutine fil_init_serial(this, nx, ny, nz, periodicx, periodicy, periodicz, methodx, methody, methodz)
        class(filters), intent(inout) :: this
        integer, intent(in) :: nx, ny, nz
        logical, intent(in) :: periodicx, periodicy, periodicz
        character(len=*), intent(in) :: methodx, methody, methodz
        integer(c_int) :: per(3), met(3)
        if (this%initialized) call fil_destroy(this)
        this%xsz = [nx, ny, nz]; this%ysz = this%xsz; this%zsz = this%xsz
        per = [flag(periodicx), flag(periodicy), flag(periodicz)]
        met = [fil_method_id(methodx), fil_method_id(methody), fil_method_id(methodz)]
        this%xmethod = met(1); this%ymethod = met(2); this%zmethod = met(3)
        call check(pdo_filters_init(this%h, nx, ny, nz, per, met))
        this%initialized = .true.
    end subroutine

    subroutine fil_init_parallel(this, gp, periodicx, periodicy, periodicz, methodx, methody, methodz)
        class(filters), intent(inout) :: this
        class(decomp_info), intent(in) :: gp
        logical, intent(in) :: periodicx, periodicy, periodicz
        character(len=*), intent(in) :: methodx, methody, methodz
        call fil_init_serial(this, gp%xsz(1), gp%ysz(2), gp%zsz(3), periodicx, periodicy, periodicz, &
                             methodx, methody, methodz)
    end subroutine

    subroutine fil_filterx(this, f, ff, bc1, bcn)
        class(filters), intent(in) :: this
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(in), target :: f
        real(rkind), dimension(this%xsz(1), this%xsz(2), this%xsz(3)), intent(out), target :: ff
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_filters_filterx(this%h, this%prec, c_loc(f), c_loc(ff), opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine fil_filtery(this, f, ff, bc1, bcn)
        class(filters), intent(in) :: this
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(in), target :: f
        real(rkind), dimension(this%ysz(1), this%ysz(2), this%ysz(3)), intent(out), target :: ff
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_filters_filtery(this%h, this%prec, c_loc(f), c_loc(ff), opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    subroutine fil_filterz(this, f, ff, bc1, bcn)
        class(filters), intent(in) :: this
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(in), target :: f
        real(rkind), dimension(this%zsz(1), this%zsz(2), this%zsz(3)), intent(out), target :: ff
        integer, optional, intent(in) :: bc1, bcn
        call check(pdo_filters_filterz(this%h, this%prec, c_loc(f), c_loc(ff), opt0(bc1), opt0(bcn), this%stream))
    end subroutine

    pure integer function fil_getMethodx(this)
        class(filters), intent(in) :: this
        fil_getMethodx = this%xmethod
    end function

    pure integer function fil_getMethody(this)
        class(filters), intent(in) :: this
        fil_getMethody = this%ymethod
    end function

    pure integer function fil_getMethodz(this)
        class(filters), intent(in) :: this
        fil_getMethodz = this%zmethod
    end function

    subroutine fil_destroy(this)
        class(filters), intent(inout) :: this
        if (c_associated(this%h)) call check(pdo_filters_destroy(this%h))
        this%h = c_null_ptr
        this%initialized = .false.
    end subroutine

    ! ---- spectral (spectral.F90:235-363, 867-931, 1413-1453) ----------------
    ! igrid's use: pencil "x", scheme "four", dimTransform 2.  Other choices
    ! stop with the reference's code (102 for dimTransform, 111 on re-init).
    subroutine sp_init(this, pencil, nx_g, ny_g, nz_g, dx, dy, dz, scheme, filt, dimTransform, fixOddball, &
                       use2decompFFT, useConsrvD2, createK, exhaustiveFFT, init_periodicInZ, dealiasF)
        class(spectral), intent(inout) :: this
        character(len=1), intent(in) :: pencil
        integer, intent(in) :: nx_g, ny_g, nz_g
        real(rkind), intent(in) :: dx, dy, dz
        character(len=*), intent(in) :: scheme, filt
        integer, intent(in), optional :: dimTransform
        logical, optional :: fixOddball, use2decompFFT, useConsrvD2
        logical, intent(in), optional :: createK, exhaustiveFFT, init_periodicInZ
        real(rkind), intent(in), optional :: dealiasF
        if (this%isInitialized) then
            print '(a)', "ERROR: You are trying to reinitialize SPECTRAL derived type. This is not allowed"
            error stop 111
        end if
        if (present(dimTransform)) then
            if (dimTransform /= 2 .and. dimTransform /= 3) then
                print '(a)', "ERROR: Incorrect choice for DIMTRANSFORM while initializing SPECTRAL derived type."
                error stop 102
            end if
            if (dimTransform == 3) error stop "spectral: only dimTransform = 2 (igrid's transforms) is supported"
        else
            error stop "spectral: only dimTransform = 2 (igrid's transforms) is supported"
        end if
        if (pencil /= "x") error stop "spectral: only the x pencil is supported"
        if (trim(scheme) /= "four" .and. trim(scheme) /= "FOUR") error stop "spectral: only scheme four is supported"
        if (len(filt) < 0) continue
        if (present(use2decompFFT) .or. present(useConsrvD2) .or. present(createK) .or. present(exhaustiveFFT)) continue
        if (present(dealiasF)) this%dealiasFact = dealiasF
        if (present(fixOddball)) this%fixOddball = fixOddball
        this%init_periodicInZ = .false.
        if (present(init_periodicInZ)) this%init_periodicInZ = init_periodicInZ
        this%nx_g = nx_g; this%ny_g = ny_g; this%nz_g = nz_g
        this%dx = dx; this%dy = dy; this%dz = dz
        this%nx_r = nx_g; this%ny_r = ny_g; this%nz_r = nz_g
        this%nx_c = nx_g / 2 + 1; this%ny_c = ny_g; this%nz_c = nz_g
        this%fft_size = [this%nx_c, this%ny_c, this%nz_c]
        this%spectdecomp%xst = 1; this%spectdecomp%yst = 1; this%spectdecomp%zst = 1
        this%spectdecomp%xsz = this%fft_size; this%spectdecomp%ysz = this%fft_size
        this%spectdecomp%zsz = this%fft_size
        this%spectdecomp%xen = this%fft_size; this%spectdecomp%yen = this%fft_size
        this%spectdecomp%zen = this%fft_size
        this%normfact = 1.d0 / real(nx_g, rkind) / real(ny_g, rkind) / real(nz_g, rkind)
        call check(c_sp_init(this%h, nx_g, ny_g, nz_g, dx, dy, dz, flag(this%fixOddball), this%dealiasFact, &
                             flag(this%init_periodicInZ)))
        this%isInitialized = .true.
    end subroutine

    subroutine sp_destroy(this)
        class(spectral), intent(inout) :: this
        if (c_associated(this%h)) call check(c_sp_destroy(this%h))
        this%h = c_null_ptr
        this%isInitialized = .false.
    end subroutine

    subroutine sp_fft(this, arr_in, arr_out)
        class(spectral), intent(inout) :: this
        real(rkind), dimension(this%nx_r, this%ny_r, this%nz_r), intent(in), target :: arr_in
        complex(rkind), dimension(this%nx_c, this%ny_c, this%nz_c), intent(out), target :: arr_out
        call check(c_sp_fft(this%h, this%prec, c_loc(arr_in), c_loc(arr_out), this%stream))
    end subroutine

    subroutine sp_ifft(this, arr_in, arr_out, setOddball)
        class(spectral), intent(inout) :: this
        complex(rkind), dimension(this%nx_c, this%ny_c, this%nz_c), intent(in), target :: arr_in
        real(rkind), dimension(this%nx_r, this%ny_r, this%nz_r), intent(out), target :: arr_out
        logical, intent(in), optional :: setOddball
        integer(c_int) :: odd
        odd = 0
        if (present(setOddball)) odd = flag(setOddball)
        call check(c_sp_ifft(this%h, this%prec, c_loc(arr_in), c_loc(arr_out), odd, this%stream))
    end subroutine

    subroutine sp_dealias(this, fhat)
        class(spectral), intent(inout) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(inout), target :: fhat
        call check(c_sp_dealias(this%h, this%prec, c_loc(fhat), this%stream))
    end subroutine

    subroutine sp_dealias_edge(this, fhat)
        class(spectral), intent(inout) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3) + 1), intent(inout), &
            target :: fhat
        call check(c_sp_dealias_edge(this%h, this%prec, c_loc(fhat), this%stream))
    end subroutine

    subroutine sp_ik1_oop(this, fin, fout)
        class(spectral), intent(in) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(in), target :: fin
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(out), target :: fout
        call check(c_sp_ik(this%h, this%prec, 1, c_loc(fin), c_loc(fout), this%stream))
    end subroutine

    subroutine sp_ik1_ip(this, fin)
        class(spectral), intent(in) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(inout), target :: fin
        call check(c_sp_ik(this%h, this%prec, 1, c_loc(fin), c_loc(fin), this%stream))
    end subroutine

    subroutine sp_ik2_oop(this, fin, fout)
        class(spectral), intent(in) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(in), target :: fin
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(out), target :: fout
        call check(c_sp_ik(this%h, this%prec, 2, c_loc(fin), c_loc(fout), this%stream))
    end subroutine

    subroutine sp_ik2_ip(this, fin)
        class(spectral), intent(in) :: this
        complex(rkind), dimension(this%fft_size(1), this%fft_size(2), this%fft_size(3)), intent(inout), target :: fin
        call check(c_sp_ik(this%h, this%prec, 2, c_loc(fin), c_loc(fin), this%stream))
    end subroutine

    ! ---- Pade6stagg (PadeDerOps.F90:57-990), scheme cd06 ----------------------
    ! the cell-centred decompositions give the sizes (gpC / sp_gpC real /
    ! complex, z extents nz and nz+1); bot, top in {-1 odd, 0 one-sided,
    ! 1 even} select the boundary object as the reference does
    subroutine p6_init(this, gpC, sp_gpC, gpE, sp_gpE, dz, scheme, isPeriodic, spectC)
        class(Pade6stagg), intent(inout) :: this
        type(decomp_info), intent(in), target :: gpC, sp_gpC, gpE, sp_gpE
        integer, intent(in) :: scheme
        real(rkind), intent(in) :: dz
        logical, intent(in) :: isPeriodic
        type(spectral), intent(in), target, optional :: spectC
        if (scheme < fd02 .or. scheme > fourierColl) then
            print '(a)', "ERROR: Invalid choice of numerical scheme in vertical"
            error stop 434
        end if
        if (scheme /= cd06_scheme) error stop "Pade6stagg: only the cd06 scheme is supported"
        if (present(spectC) .or. gpE%zsz(3) < 0 .or. sp_gpE%zsz(3) < 0) continue
        call p6_destroy(this)
        this%gp = gpC; this%sp_gp = sp_gpC
        this%scheme = scheme; this%dz = dz; this%isPeriodic = isPeriodic
        call check(c_p6_init(this%h, gpC%zsz(3), dz, flag(isPeriodic)))
    end subroutine

    subroutine p6_destroy(this)
        class(Pade6stagg), intent(inout) :: this
        if (c_associated(this%h)) call check(c_p6_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    subroutine p6_ddz_E2C_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 10, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_ddz_E2C_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 10, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine
    subroutine p6_ddz_C2E_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 11, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_ddz_C2E_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 11, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine
    subroutine p6_d2dz2_C2C_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 12, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_d2dz2_C2C_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 12, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine
    subroutine p6_d2dz2_E2E_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 13, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_d2dz2_E2E_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 13, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine
    subroutine p6_interpz_E2C_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 14, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_interpz_E2C_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 14, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine
    subroutine p6_interpz_C2E_r(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3)), intent(in), target :: input
        real(rkind), dimension(this%gp%zsz(1), this%gp%zsz(2), this%gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 15, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%gp%zsz(1), this%gp%zsz(2), 0, this%stream))
    end subroutine
    subroutine p6_interpz_C2E_c(this, input, output, bot, top)
        class(Pade6stagg), intent(in) :: this
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3)), intent(in), target :: input
        complex(rkind), dimension(this%sp_gp%zsz(1), this%sp_gp%zsz(2), this%sp_gp%zsz(3) + 1), intent(out), target :: output
        integer, intent(in) :: bot, top
        call check(c_p6_apply(this%h, 15, bot, top, this%prec, c_loc(input), c_loc(output), &
                              this%sp_gp%zsz(1), this%sp_gp%zsz(2), 1, this%stream))
    end subroutine

    ! ---- padepoisson (PadePoisson.F90:130-1244): periodic in z, or with walls
    ! and optionally the Stokes pressure
    subroutine pp_init(this, dx, dy, dz, sp, spE, computeStokesPressure, Lz, storePressure, gpC, derivZ, PeriodicInZ)
        class(padepoisson), intent(inout) :: this
        real(rkind), intent(in) :: dx, dy, dz
        class(spectral), intent(in), target :: sp, spE
        logical, intent(in) :: computeStokesPressure, storePressure
        real(rkind), intent(in), optional :: Lz
        type(decomp_info), intent(in), target :: gpC
        type(Pade6stagg), intent(in), target :: derivZ
        logical, intent(in) :: PeriodicInZ
        if (derivZ%scheme /= cd06_scheme) error stop "padepoisson: derivZ must use the cd06 scheme"
        if (present(Lz) .or. storePressure .or. dx < 0 .or. dy < 0 .or. gpC%xsz(1) < 0) continue
        call pp_destroy(this)
        this%PeriodicInZ = PeriodicInZ
        this%nx = sp%nx_g; this%ny = sp%ny_g; this%nz = sp%nz_g
        this%nx_in = sp%nx_c; this%ny_in = sp%ny_c; this%nz_in = sp%nz_c
        this%nxE_in = spE%nx_c; this%nyE_in = spE%ny_c; this%nzE_in = spE%nz_c
        if (PeriodicInZ) then
            call check(c_pp_init(this%h, sp%h, dz))
        else
            call check(c_pp_init_np(this%h, sp%h, dz))
            if (computeStokesPressure) then
                if (.not. present(Lz)) error stop "padepoisson: the Stokes pressure needs Lz"
                call check(c_pp_set_stokes(this%h, Lz))
            end if
        end if
    end subroutine

    subroutine pp_projection(this, uhat, vhat, what)
        class(padepoisson), intent(inout) :: this
        complex(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(inout), target :: uhat, vhat
        complex(rkind), dimension(this%nxE_in, this%nyE_in, this%nzE_in), intent(inout), target :: what
        call check(c_pp_project(this%h, this%prec, c_loc(uhat), c_loc(vhat), c_loc(what), this%stream))
    end subroutine

    ! getPressure (PadePoisson.F90:752-898): the velocities stay as they are
    subroutine pp_get_pressure(this, uhat, vhat, what, pressure)
        class(padepoisson), intent(inout) :: this
        complex(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(in), target :: uhat, vhat
        complex(rkind), dimension(this%nxE_in, this%nyE_in, this%nzE_in), intent(in), target :: what
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(inout), target :: pressure
        call check(c_pp_pressure(this%h, this%prec, c_loc(uhat), c_loc(vhat), c_loc(what), c_loc(pressure), 0, &
                                 this%stream))
    end subroutine

    ! getPressureAndUpdateRHS (PadePoisson.F90:952-1163): the projection and
    ! its pressure
    subroutine pp_get_pressure_update(this, uhat, vhat, what, pressure)
        class(padepoisson), intent(inout) :: this
        complex(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(inout), target :: uhat, vhat
        complex(rkind), dimension(this%nxE_in, this%nyE_in, this%nzE_in), intent(inout), target :: what
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(inout), target :: pressure
        call check(c_pp_pressure(this%h, this%prec, c_loc(uhat), c_loc(vhat), c_loc(what), c_loc(pressure), 1, &
                                 this%stream))
    end subroutine

    ! steps 1-3 of DivergenceCheck: the divergence field (the reference then
    ! prints its maximum and may fix it, which a caller does from this field)
    subroutine pp_divergence(this, uhat, vhat, what, divergence, fixDiv)
        class(padepoisson), intent(inout) :: this
        complex(rkind), dimension(this%nx_in, this%ny_in, this%nz_in), intent(inout), target :: uhat, vhat
        complex(rkind), dimension(this%nxE_in, this%nyE_in, this%nzE_in), intent(inout), target :: what
        real(rkind), dimension(this%nx, this%ny, this%nz), intent(out), target :: divergence
        logical, intent(in), optional :: fixDiv
        if (present(fixDiv)) then
            if (fixDiv) error stop "padepoisson: fixDiv is not supported"
        end if
        call check(c_pp_div(this%h, this%prec, c_loc(uhat), c_loc(vhat), c_loc(what), c_loc(divergence), &
                            this%stream))
    end subroutine

    subroutine pp_destroy(this)
        class(padepoisson), intent(inout) :: this
        if (c_associated(this%h)) call check(c_pp_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

    ! ---------------------------------------------------------------- cgrid
    function cg_init(this, params, prec) result(ierr)
        class(cgrid), intent(inout) :: this
        type(cgrid_params), intent(in), target :: params
        integer, optional, intent(in) :: prec
        integer :: ierr
        if (present(prec)) this%prec = prec
        if (c_associated(this%h)) call check(c_cg_destroy(this%h))
        this%h = c_null_ptr
        ierr = c_cg_init(this%h, c_loc(params), this%prec)
        if (ierr /= 0) return
        this%nxp = params%nx
        this%nyp = params%ny
        this%nzp = params%nz
    end function

    subroutine cg_get_rhs(this, W, rhs)
        class(cgrid), intent(inout) :: this
        real(rkind), dimension(this%nxp, this%nyp, this%nzp, 5), intent(in), target :: W
        real(rkind), dimension(this%nxp, this%nyp, this%nzp, 5), intent(out), target :: rhs
        call check(c_cg_rhs(this%h, this%prec, c_loc(W), c_loc(rhs), this%stream))
    end subroutine

    ! cgrid%filter(arr, myfil, numtimes): numtimes passes (default 1) along
    ! y, x and z with the conserved-variable filter fil, or, with
    ! gaussian = .true., the Gaussian gfil (the reference passes this%gfil as
    ! myfil for that); the boundary flags are the init's
    subroutine cg_filter(this, arr, numtimes, gaussian)
        class(cgrid), intent(inout) :: this
        real(rkind), dimension(this%nxp, this%nyp, this%nzp), intent(inout), target :: arr
        integer, optional, intent(in) :: numtimes
        logical, optional, intent(in) :: gaussian
        integer(c_int) :: nt, gs
        nt = 1
        if (present(numtimes)) nt = numtimes
        gs = 0
        if (present(gaussian)) then
            if (gaussian) gs = 1
        end if
        call check(c_cg_filter(this%h, this%prec, c_loc(arr), nt, gs, this%stream))
    end subroutine

    subroutine cg_destroy(this)
        class(cgrid), intent(inout) :: this
        if (c_associated(this%h)) call check(c_cg_destroy(this%h))
        this%h = c_null_ptr
    end subroutine

end module padeops_hip_mod
