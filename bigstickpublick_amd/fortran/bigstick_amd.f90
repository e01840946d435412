! iso_c_binding interface to libbigstick_amd.so (include/bigstick_amd.h), the
! MI355X engine for the M-scheme Hamiltonian and Lanczos driver.  A Fortran
! host (e.g. BIGSTICK's lanczos_p, src/blanczosmain.f90:1413) calls these in
! place of applyHbundled_omp / the br_* reorthogonalization sequence.
module bigstick_amd
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: BS_F32 = 4, BS_F64 = 8

  type, bind(C) :: bs_lanczos_params
    integer(c_int) :: nkeep, niter, thick, max_restarts, dtype
    real(c_double) :: tol
    integer(c_int) :: polish
    integer(c_int) :: conv_test = 0, ncheck_ex = 0
    integer(c_int) :: target = 0
    real(c_double) :: etarget = 0.0d0
  end type bs_lanczos_params

  interface
    function bs_init(sps, intf, z, n, jz2, spscale, a0, a, x, sys) bind(C, name='bs_init') result(rc)
      import :: c_char, c_int, c_float, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: sps, intf
      integer(c_int), value :: z, n, jz2
      real(c_float), value :: spscale, a0, a, x
      type(c_ptr), intent(out) :: sys
      integer(c_int) :: rc
    end function bs_init

    function bs_init_host(sps, intf, z, n, jz2, spscale, a0, a, x, sys) bind(C, name='bs_init_host') result(rc)
      import :: c_char, c_int, c_float, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: sps, intf
      integer(c_int), value :: z, n, jz2
      real(c_float), value :: spscale, a0, a, x
      type(c_ptr), intent(out) :: sys
      integer(c_int) :: rc
    end function bs_init_host

    function bs_upload(sys) bind(C, name='bs_upload') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int) :: rc
    end function bs_upload

    function bs_finalize(sys) bind(C, name='bs_finalize') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int) :: rc
    end function bs_finalize

    function bs_last_error() bind(C, name='bs_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function bs_last_error

    function bs_dim(sys, dim) bind(C, name='bs_dim') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: sys
      integer(c_int64_t), intent(out) :: dim
      integer(c_int) :: rc
    end function bs_dim

    function bs_num_sd(sys, it, nsd) bind(C, name='bs_num_sd') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: sys
      integer(c_int), value :: it
      integer(c_int64_t), intent(out) :: nsd
      integer(c_int) :: rc
    end function bs_num_sd

    function bs_basis(sys, it, mask, start) bind(C, name='bs_basis') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int), value :: it
      type(c_ptr), value :: mask, start
      integer(c_int) :: rc
    end function bs_basis

    function bs_apply_h(sys, x, y, dtype, accumulate, stream) bind(C, name='bs_apply_h') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys, x, y, stream
      integer(c_int), value :: dtype, accumulate
      integer(c_int) :: rc
    end function bs_apply_h

    function bs_apply_h_host(sys, x, y, dtype, accumulate) bind(C, name='bs_apply_h_host') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys, x, y
      integer(c_int), value :: dtype, accumulate
      integer(c_int) :: rc
    end function bs_apply_h_host

    function bs_autotune(sys, dtype, variant) bind(C, name='bs_autotune') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int), value :: dtype
      integer(c_int), intent(out) :: variant
      integer(c_int) :: rc
    end function bs_autotune

    function bs_lanczos(sys, p, alpha, beta, evals, evecs, resid, iters) bind(C, name='bs_lanczos') result(rc)
      import :: c_ptr, c_int, bs_lanczos_params
      type(c_ptr), value :: sys
      type(bs_lanczos_params), intent(in) :: p
      type(c_ptr), value :: alpha, beta, evals, evecs, resid
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function bs_lanczos

    function bs_lanczos_stats(sys, random_restarts, second_passes, steps, thick_restarts) &
        bind(C, name='bs_lanczos_stats') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int), intent(out) :: random_restarts, second_passes, steps, thick_restarts
      integer(c_int) :: rc
    end function bs_lanczos_stats

    function bs_parity(sys, parity) bind(C, name='bs_parity') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys
      integer(c_int), intent(out) :: parity
      integer(c_int) :: rc
    end function bs_parity

    function bs_exact(sys, nkeep, dtype, evals, evecs, resid, iters) bind(C, name='bs_exact') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: sys, evals, evecs, resid
      integer(c_int), value :: nkeep, dtype
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function bs_exact

    function bs_block_lanczos(sys, p, block, evals, evecs, resid, iters) bind(C, name='bs_block_lanczos') result(rc)
      import :: c_ptr, c_int, bs_lanczos_params
      type(c_ptr), value :: sys, evals, evecs, resid
      type(bs_lanczos_params), intent(in) :: p
      integer(c_int), value :: block
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function bs_block_lanczos

    function bs_block_strength(sys, p, block, pivots, pivot_dtype, pnorm2, strength, evals, evecs, resid, iters) &
        bind(C, name='bs_block_strength') result(rc)
      import :: c_ptr, c_int, bs_lanczos_params
      type(c_ptr), value :: sys, pivots, pnorm2, strength, evals, evecs, resid
      type(bs_lanczos_params), intent(in) :: p
      integer(c_int), value :: block, pivot_dtype
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function bs_block_strength

    function bs_strength(sys, p, pivot, pivot_norm, ritz, strength, alpha, beta, evals, evecs, resid, iters) &
        bind(C, name='bs_strength') result(rc)
      import :: c_ptr, c_int, c_double, bs_lanczos_params
      type(c_ptr), value :: sys, pivot, ritz, strength, alpha, beta, evals, evecs, resid
      type(bs_lanczos_params), intent(in) :: p
      real(c_double), intent(out) :: pivot_norm
      integer(c_int), intent(out) :: iters
      integer(c_int) :: rc
    end function bs_strength

    function bs_observable(sys, which, psi, dtype, val) bind(C, name='bs_observable') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: sys, psi
      integer(c_int), value :: which, dtype
      real(c_double), intent(out) :: val
      integer(c_int) :: rc
    end function bs_observable

    function bs_write_wfn(sys, path, nkeep, energy, xj, xt2, vecs, dtype, on_device) &
        bind(C, name='bs_write_wfn') result(rc)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: sys, energy, xj, xt2, vecs
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: nkeep, dtype, on_device
      integer(c_int) :: rc
    end function bs_write_wfn

    function bs_init_wfn(wfn, intf, spscale, a0, a, x, sys) bind(C, name='bs_init_wfn') result(rc)
      import :: c_char, c_int, c_float, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: wfn, intf
      real(c_float), value :: spscale, a0, a, x
      type(c_ptr), intent(out) :: sys
      integer(c_int) :: rc
    end function bs_init_wfn

    ! intf = c_null_ptr: the basis only (H = 0)
    function bs_init_wfn_host(wfn, intf, spscale, a0, a, x, sys) bind(C, name='bs_init_wfn_host') result(rc)
      import :: c_char, c_int, c_float, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: wfn
      type(c_ptr), value :: intf
      real(c_float), value :: spscale, a0, a, x
      type(c_ptr), intent(out) :: sys
      integer(c_int) :: rc
    end function bs_init_wfn_host

    function bs_apply_onebody(sys, opme, x, y, nvec, dtype, jop, top) bind(C, name='bs_apply_onebody') result(rc)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: sys, x, y
      character(kind=c_char), dimension(*), intent(in) :: opme
      integer(c_int), value :: nvec, dtype
      integer(c_int), intent(out) :: jop, top
      integer(c_int) :: rc
    end function bs_apply_onebody

    function bs_read_wfn(sys, path, first, count, energy, xj, xt2, vecs, dtype, on_device, nkeep) &
        bind(C, name='bs_read_wfn') result(rc)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: sys, energy, xj, xt2, vecs
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: first, count, dtype, on_device
      integer(c_int), intent(out) :: nkeep
      integer(c_int) :: rc
    end function bs_read_wfn

    ! intf = c_null_ptr applies the system's own Hamiltonian
    function bs_operator_matrix(sys, intf, spscale, a0, a, x, vecs, nvec, dtype, out) &
        bind(C, name='bs_operator_matrix') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: sys, intf, vecs, out
      real(c_float), value :: spscale, a0, a, x
      integer(c_int), value :: nvec, dtype
      integer(c_int) :: rc
    end function bs_operator_matrix

    ! HIP runtime, for device buffers owned by the Fortran host
    function hipMalloc(ptr, bytes) bind(C, name='hipMalloc') result(rc)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), intent(out) :: ptr
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function hipMalloc

    ! kind 1 = host to device, 2 = device to host
    function hipMemcpy(dst, src, bytes, kind) bind(C, name='hipMemcpy') result(rc)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
      integer(c_int) :: rc
    end function hipMemcpy

    function hipFree(ptr) bind(C, name='hipFree') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ptr
      integer(c_int) :: rc
    end function hipFree

    function c_strlen(s) bind(C, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  ! Fortran string -> NUL-terminated C string
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function cstr

  ! message of the last failing call
  function bs_error_message() result(msg)
    character(len=512) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    msg = ' '
    p = bs_last_error()
    if (.not. c_associated(p)) return
    n = int(min(c_strlen(p), int(len(msg), c_size_t)))
    call c_f_pointer(p, chars, [n])
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function bs_error_message

  ! stop with the engine's message on a nonzero return code (the reference's
  ! own error handling is `stop`, src/bapplyhlibOMP.f90:941-946)
  subroutine bs_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(6, '(a,": ",a)') what, trim(bs_error_message())
      stop 1
    end if
  end subroutine bs_check

  ! device address offset by `bytes`
  function ptr_offset(p, bytes) result(q)
    type(c_ptr), intent(in) :: p
    integer(c_int64_t), intent(in) :: bytes
    type(c_ptr) :: q
    integer(c_intptr_t) :: a
    a = transfer(p, a)
    a = a + int(bytes, c_intptr_t)
    q = transfer(a, q)
  end function ptr_offset
end module bigstick_amd
