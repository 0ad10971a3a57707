! Single-GPU replacement of the 2DECOMP&FFT pieces the PadeOps solvers call
! around the hot path (vendored 2decomp_fft: decomp_info, decomp_info_init,
! transpose_x_to_y / y_to_z / z_to_y / y_to_x).  On one GPU every pencil of a
! grid holds the whole (nx, ny, nz) box in the same (i, j, k) order, so a
! transpose is an order-preserving copy of device memory (none at all when
! input and output are the same array).  Arrays are device arrays (pdo_alloc
! and c_f_pointer, as with padeops_hip_mod); the copies are asynchronous on
! the given stream (default: the null stream).
module padeops_hip_decomp
    use iso_c_binding
    implicit none
    private

    integer, parameter :: rkind = c_double
    integer, public :: nrank = 0, nproc = 1

    ! the fields of decomp_info the solvers read (xst/xen/xsz, y..., z...)
    type, public :: decomp_info
        integer, dimension(3) :: xst = 1, xen = 0, xsz = 0
        integer, dimension(3) :: yst = 1, yen = 0, ysz = 0
        integer, dimension(3) :: zst = 1, zen = 0, zsz = 0
    end type

    public :: decomp_2d_init, decomp_info_init, decomp_info_finalize
    public :: transpose_x_to_y, transpose_y_to_z, transpose_z_to_y, transpose_y_to_x

    interface
        integer(c_int) function c_d2d(dst, src, bytes, stream) bind(C, name="pdo_memcpy_d2d")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr), value :: dst, src, stream
            integer(c_size_t), value :: bytes
        end function
    end interface

    interface transpose_x_to_y
        module procedure tr_real, tr_cmplx
    end interface
    interface transpose_y_to_z
        module procedure tr_real, tr_cmplx
    end interface
    interface transpose_z_to_y
        module procedure tr_real, tr_cmplx
    end interface
    interface transpose_y_to_x
        module procedure tr_real, tr_cmplx
    end interface

contains

    ! one process, one pencil: p_row = p_col = 1 whatever is asked
    subroutine decomp_2d_init(nx, ny, nz, p_row, p_col)
        integer, intent(in) :: nx, ny, nz, p_row, p_col
        if (nx < 1 .or. ny < 1 .or. nz < 1 .or. p_row < 0 .or. p_col < 0) error stop "decomp_2d_init: bad sizes"
    end subroutine

    subroutine decomp_info_init(nx, ny, nz, decomp)
        integer, intent(in) :: nx, ny, nz
        type(decomp_info), intent(inout) :: decomp
        decomp%xst = 1; decomp%yst = 1; decomp%zst = 1
        decomp%xsz = [nx, ny, nz]; decomp%ysz = decomp%xsz; decomp%zsz = decomp%xsz
        decomp%xen = decomp%xsz; decomp%yen = decomp%ysz; decomp%zen = decomp%zsz
    end subroutine

    subroutine decomp_info_finalize(decomp)
        type(decomp_info), intent(inout) :: decomp
        decomp%xsz = 0; decomp%ysz = 0; decomp%zsz = 0
        decomp%xen = 0; decomp%yen = 0; decomp%zen = 0
    end subroutine

    subroutine copy(dst, src, bytes, stream)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: bytes
        type(c_ptr), optional, intent(in) :: stream
        type(c_ptr) :: s
        s = c_null_ptr
        if (present(stream)) s = stream
        if (c_associated(dst, src)) return
        if (c_d2d(dst, src, bytes, s) /= 0) error stop "padeops_hip_decomp: device copy failed"
    end subroutine

    subroutine tr_real(src, dst, decomp, stream)
        real(rkind), dimension(:,:,:), intent(in), target, contiguous :: src
        real(rkind), dimension(:,:,:), intent(inout), target, contiguous :: dst
        type(decomp_info), optional, intent(in) :: decomp
        type(c_ptr), optional, intent(in) :: stream
        if (size(src) /= size(dst)) error stop "padeops_hip_decomp: transpose size mismatch"
        if (present(decomp)) then
            if (any(shape(src) /= decomp%xsz)) error stop "padeops_hip_decomp: pencil shape mismatch"
        end if
        call copy(c_loc(dst), c_loc(src), int(8, c_size_t) * size(src, kind=c_size_t), stream)
    end subroutine

    subroutine tr_cmplx(src, dst, decomp, stream)
        complex(rkind), dimension(:,:,:), intent(in), target, contiguous :: src
        complex(rkind), dimension(:,:,:), intent(inout), target, contiguous :: dst
        type(decomp_info), optional, intent(in) :: decomp
        type(c_ptr), optional, intent(in) :: stream
        if (size(src) /= size(dst)) error stop "padeops_hip_decomp: transpose size mismatch"
        if (present(decomp)) then
            if (any(shape(src) /= decomp%xsz)) error stop "padeops_hip_decomp: pencil shape mismatch"
        end if
        call copy(c_loc(dst), c_loc(src), int(16, c_size_t) * size(src, kind=c_size_t), stream)
    end subroutine

end module padeops_hip_decomp
