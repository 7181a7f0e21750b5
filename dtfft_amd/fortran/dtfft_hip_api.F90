!------------------------------------------------------------------------------------------------
! dtfft_hip_api -- iso_c_binding interface to libdtfft_hip (include/dtfft_hip.h).
!
! This is the thin layer through which dtFFT's Fortran plugins call the gfx950 kernels:
!   * kernel_hip (extends abstract_kernel, src/dtfft_abstract_kernel.F90:106-163) calls
!     dtfft_hip_kernel_create / _execute / _destroy;
!   * executor_hip (extends abstract_executor, src/dtfft_abstract_executor.F90:47-112)
!     calls dtfft_hip_fft_create / _execute / _destroy and dtfft_hip_mem_alloc / _free.
! The module depends on iso_c_binding only, so it builds without dtFFT or MPI.
!------------------------------------------------------------------------------------------------
module dtfft_hip_api
use iso_c_binding
implicit none
private

  integer(c_int32_t), parameter, public :: DTFFT_HIP_SUCCESS = 0
  ! kernel_type_t values, src/dtfft_abstract_kernel.F90:59-98
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PACK = 1
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_COPY_PIPELINED = 2
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_UNPACK = 3
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_COPY = 4
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_UNPACK_PIPELINED = 5
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PACK_PIPELINED = 6
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PERMUTE_FORWARD = 7
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PERMUTE_BACKWARD = 8
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_START = 9
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END = 10
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PERMUTE_BACKWARD_END_PIPELINED = 11
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PACK_FORWARD = 12
  integer(c_int32_t), parameter, public :: DTFFT_HIP_KERNEL_PACK_BACKWARD = 13
  ! abstract_executor constants, src/dtfft_abstract_executor.F90:35-45
  integer(c_int8_t),  parameter, public :: DTFFT_HIP_FFT_C2C = 0_c_int8_t
  integer(c_int8_t),  parameter, public :: DTFFT_HIP_FFT_R2C = 1_c_int8_t
  integer(c_int8_t),  parameter, public :: DTFFT_HIP_FFT_FORWARD = -1_c_int8_t
  integer(c_int8_t),  parameter, public :: DTFFT_HIP_FFT_BACKWARD = 1_c_int8_t

  public :: dtfft_hip_kernel_create, dtfft_hip_kernel_execute, dtfft_hip_kernel_destroy
  public :: dtfft_hip_fft_create, dtfft_hip_fft_execute, dtfft_hip_fft_destroy
  public :: dtfft_hip_mem_alloc, dtfft_hip_mem_free
  public :: dtfft_hip_memcpy_h2d, dtfft_hip_memcpy_d2h, dtfft_hip_device_synchronize

  interface
    integer(c_int32_t) function dtfft_hip_kernel_create(kernel_type, ndims, dims, base_storage, neighbor_data, &
                                                        n_peers, effort, force_effort, handle) bind(C)
    import
      integer(c_int32_t), value         :: kernel_type, ndims
      integer(c_int32_t), intent(in)    :: dims(*)
      integer(c_int64_t), value         :: base_storage
      type(c_ptr),        value         :: neighbor_data   ! c_loc of neighbor_data(5, P), or c_null_ptr
      integer(c_int32_t), value         :: n_peers, effort, force_effort
      type(c_ptr),        intent(out)   :: handle
    end function dtfft_hip_kernel_create

    integer(c_int32_t) function dtfft_hip_kernel_execute(handle, in, out, stream, neighbor, sync) bind(C)
    import
      type(c_ptr),        value :: handle, in, out, stream
      integer(c_int32_t), value :: neighbor, sync
    end function dtfft_hip_kernel_execute

    integer(c_int32_t) function dtfft_hip_kernel_destroy(handle) bind(C)
    import
      type(c_ptr), value :: handle
    end function dtfft_hip_kernel_destroy

    integer(c_int32_t) function dtfft_hip_fft_create(fft_rank, fft_type, precision, idist, odist, how_many, &
                                                     fft_sizes, inembed, onembed, r2r_kinds, effort, stream, &
                                                     plan) bind(C)
    import
      integer(c_int8_t),  value         :: fft_rank, fft_type
      integer(c_int32_t), value         :: precision, idist, odist, how_many
      integer(c_int32_t), intent(in)    :: fft_sizes(*), inembed(*), onembed(*)
      type(c_ptr),        value         :: r2r_kinds
      integer(c_int32_t), value         :: effort
      type(c_ptr),        value         :: stream
      type(c_ptr),        intent(out)   :: plan
    end function dtfft_hip_fft_create

    integer(c_int32_t) function dtfft_hip_fft_execute(plan, in, out, sign) bind(C)
    import
      type(c_ptr),       value :: plan, in, out
      integer(c_int8_t), value :: sign
    end function dtfft_hip_fft_execute

    integer(c_int32_t) function dtfft_hip_fft_destroy(plan) bind(C)
    import
      type(c_ptr), value :: plan
    end function dtfft_hip_fft_destroy

    integer(c_int32_t) function dtfft_hip_mem_alloc(bytes, ptr) bind(C)
    import
      integer(c_size_t), value       :: bytes
      type(c_ptr),       intent(out) :: ptr
    end function dtfft_hip_mem_alloc

    integer(c_int32_t) function dtfft_hip_mem_free(ptr) bind(C)
    import
      type(c_ptr), value :: ptr
    end function dtfft_hip_mem_free

    integer(c_int32_t) function dtfft_hip_memcpy_h2d(dst, src, bytes) bind(C)
    import
      type(c_ptr),       value :: dst, src
      integer(c_size_t), value :: bytes
    end function dtfft_hip_memcpy_h2d

    integer(c_int32_t) function dtfft_hip_memcpy_d2h(dst, src, bytes) bind(C)
    import
      type(c_ptr),       value :: dst, src
      integer(c_size_t), value :: bytes
    end function dtfft_hip_memcpy_d2h

    integer(c_int32_t) function dtfft_hip_device_synchronize() bind(C)
    import
    end function dtfft_hip_device_synchronize
  end interface
end module dtfft_hip_api
