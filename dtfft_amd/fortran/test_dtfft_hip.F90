!------------------------------------------------------------------------------------------------
! Fortran driver of the C-ABI through dtfft_hip_api (what dtFFT's kernel_hip and
! executor_hip plugins do):
!   1. KERNEL_PERMUTE_FORWARD on a [33,77,21] complex128 array, compared bit for bit with
!      the host loop of permute_forward_write (src/include/_dtfft_kernel_host_routines.inc:
!      140-196);
!   2. a batched 1D C2C executor (n = 60, 97 lines, FP64), forward then backward, the
!      round trip divided by n compared with the input to 1e-13;
!   3. per-peer KERNEL_PACK_PIPELINED with Fortran-built neighbor_data(5, P).
! Prints PASS and exits 0 on success.
!------------------------------------------------------------------------------------------------
program test_dtfft_hip
use iso_c_binding
use dtfft_hip_api
implicit none
  integer(c_int32_t), parameter :: nx = 33, ny = 77, nz = 21
  complex(c_double_complex), allocatable, target :: a(:), b(:), gold(:)
  integer(c_int32_t), target :: dims(3), nd(5, 3)
  type(c_ptr) :: d_a, d_b, k, plan
  integer(c_int32_t) :: ierr, x, y, z, n, i, nl
  integer(c_size_t) :: bytes
  real(c_double) :: err
  integer(c_int32_t), target :: sizes(1)

  allocate(a(nx * ny * nz), b(nx * ny * nz), gold(nx * ny * nz))
  do i = 1, size(a)
    a(i) = cmplx(real(i, c_double), -0.5_c_double * i, c_double)
  enddo
  bytes = int(size(a), c_size_t) * 16_c_size_t
  call check(dtfft_hip_mem_alloc(bytes, d_a), "mem_alloc")
  call check(dtfft_hip_mem_alloc(bytes, d_b), "mem_alloc")
  call check(dtfft_hip_memcpy_h2d(d_a, c_loc(a), bytes), "h2d")

  ! 1. permute forward vs the host loop
  dims = [nx, ny, nz]
  call check(dtfft_hip_kernel_create(DTFFT_HIP_KERNEL_PERMUTE_FORWARD, 3, dims, 16_c_int64_t, c_null_ptr, 0, 0, 0, k), &
             "kernel_create")
  call check(dtfft_hip_kernel_execute(k, d_a, d_b, c_null_ptr, 0, 1), "kernel_execute")
  call check(dtfft_hip_memcpy_d2h(c_loc(b), d_b, bytes), "d2h")
  do x = 0, nx - 1
    do z = 0, nz - 1
      do y = 0, ny - 1
        gold(x * ny * nz + z * ny + y + 1) = a(z * nx * ny + x + y * nx + 1)
      enddo
    enddo
  enddo
  if ( any(transfer(b, [0_c_int64_t]) /= transfer(gold, [0_c_int64_t])) ) error stop "permute_forward mismatch"
  call check(dtfft_hip_kernel_destroy(k), "kernel_destroy")

  ! 2. executor round trip
  n = 60; nl = 97
  sizes(1) = n
  call check(dtfft_hip_fft_create(1_c_int8_t, DTFFT_HIP_FFT_C2C, 1, n, n, nl, sizes, sizes, sizes, c_null_ptr, 0, &
                                  c_null_ptr, plan), "fft_create")
  call check(dtfft_hip_fft_execute(plan, d_a, d_b, DTFFT_HIP_FFT_FORWARD), "fft fwd")
  call check(dtfft_hip_fft_execute(plan, d_b, d_b, DTFFT_HIP_FFT_BACKWARD), "fft bwd")
  call check(dtfft_hip_memcpy_d2h(c_loc(b), d_b, int(n * nl, c_size_t) * 16_c_size_t), "d2h")
  err = maxval(abs(b(1:n * nl) / n - a(1:n * nl))) / maxval(abs(a(1:n * nl)))
  if ( err > 1e-13_c_double ) error stop "executor round trip error"
  call check(dtfft_hip_fft_destroy(plan), "fft_destroy")

  ! 3. pack per peer: three x-slabs of [nx, ny, nz] into contiguous chunks
  nd(:, 1) = [11, ny, nz, 0, 0]
  nd(:, 2) = [11, ny, nz, 11, 11 * ny * nz]
  nd(:, 3) = [11, ny, nz, 22, 22 * ny * nz]
  call check(dtfft_hip_kernel_create(DTFFT_HIP_KERNEL_PACK_PIPELINED, 3, dims, 16_c_int64_t, c_loc(nd), 3, 0, 0, k), &
             "pack create")
  do i = 1, 3
    call check(dtfft_hip_kernel_execute(k, d_a, d_b, c_null_ptr, i, 0), "pack execute")
  enddo
  call check(dtfft_hip_device_synchronize(), "sync")
  call check(dtfft_hip_memcpy_d2h(c_loc(b), d_b, bytes), "d2h")
  do i = 0, 2
    do z = 0, nz - 1
      do y = 0, ny - 1
        do x = 0, 10
          gold(i * 11 * ny * nz + z * 11 * ny + y * 11 + x + 1) = a(z * nx * ny + y * nx + 11 * i + x + 1)
        enddo
      enddo
    enddo
  enddo
  if ( any(transfer(b, [0_c_int64_t]) /= transfer(gold, [0_c_int64_t])) ) error stop "pack mismatch"
  call check(dtfft_hip_kernel_destroy(k), "kernel_destroy")

  call check(dtfft_hip_mem_free(d_a), "free")
  call check(dtfft_hip_mem_free(d_b), "free")
  print '(a)', "PASS"
contains
  subroutine check(code, what)
    integer(c_int32_t), intent(in) :: code
    character(len=*),   intent(in) :: what
    if ( code /= DTFFT_HIP_SUCCESS ) then
      print '(a,a,i0)', what, " failed with code ", code
      error stop 1
    endif
  end subroutine check
end program test_dtfft_hip
