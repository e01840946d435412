! bs_driver: a Fortran host for the MI355X engine that reads BIGSTICK's own
! stdin menu answers for a normal run (menu option n, src/bmenu.f90:11;
! interaction menu, src/binput_tbme.f90:178-573; Lanczos options lf/ld/tf/td,
! src/blanczosmain.f90:380-745) and writes <name>.res (energy table in the
! format of src/bparticleocc.f90:256) and <name>.lcoef (alpha, beta per
! iteration in the format of src/blanczosmain.f90:1685).
!
! Precision: FP32 vectors with FP64 coefficients and an FP64 Rayleigh-quotient
! polish (the default), or all FP64 with environment BS_DTYPE=f64.
! Environment BS_HOST_ONLY=1 builds the tables, prints the dimension and stops
! (no device needed).
program bs_driver
  use iso_c_binding
  use bigstick_amd
  implicit none
  character(len=256) :: opt, outname, spsname, intname, line, lopt
  integer :: z, n, jz2, nkeep, niter, nthick, maxrest, ios, k, iu
  real(c_float) :: spscale, a0, a, x
  type(c_ptr) :: sys, dvec
  integer(c_int64_t) :: dim, nbytes, esize
  integer(c_int) :: rc, iters, dtype, variant
  type(bs_lanczos_params) :: p
  real(c_double), allocatable, target :: alpha(:), beta(:), evals(:), resid(:)
  real(c_double) :: j2, t2
  real :: xj, xt
  real(c_float), allocatable, target :: we(:), wj(:), wt2(:)
  logical :: write_wfn
  character(len=16) :: env

  read(5, '(a)') opt
  opt = adjustl(opt)
  if (opt(1:1) /= 'n' .and. opt(1:1) /= 'N') then
    write(6, *) ' bs_driver: only the normal diagonalization run (menu n / ns) is supported'
    stop 1
  end if
  ! n writes <name>.wfn, ns suppresses it (src/bmenu.f90:680-687)
  write_wfn = .not. (opt(2:2) == 's' .or. opt(2:2) == 'S')
  read(5, '(a)') outname
  read(5, '(a)') spsname
  read(5, *) z, n
  read(5, *) jz2
  read(5, '(a)') intname
  read(5, *) spscale, a0, a, x
  read(5, '(a)') line
  if (adjustl(line) /= 'end') then
    write(6, *) ' bs_driver: one interaction file is supported'
    stop 1
  end if
  call get_environment_variable('BS_HOST_ONLY', env)
  if (trim(env) == '1') then
    rc = bs_init_host(cstr(trim(first_word(spsname)) // '.sps'), cstr(trim(first_word(intname)) // '.int'), &
                      z, n, jz2, spscale, a0, a, x, sys)
    call bs_check(rc, 'bs_init_host')
    rc = bs_dim(sys, dim)
    write(6, '(a,i0)') ' dimension = ', dim
    rc = bs_finalize(sys)
    stop
  end if
  read(5, '(a)') lopt
  lopt = adjustl(lopt)
  read(5, *) nkeep, niter
  nthick = 0
  maxrest = 0
  if (lopt(1:1) == 't') then
    read(5, *) nthick
    read(5, *, iostat=ios) maxrest
    if (ios /= 0) maxrest = 20
  end if

  rc = bs_init(cstr(trim(first_word(spsname)) // '.sps'), cstr(trim(first_word(intname)) // '.int'), &
               z, n, jz2, spscale, a0, a, x, sys)
  call bs_check(rc, 'bs_init')
  rc = bs_dim(sys, dim)
  call get_environment_variable('BS_DTYPE', env)
  dtype = BS_F32
  if (trim(env) == 'f64') dtype = BS_F64
  esize = 4
  if (dtype == BS_F64) esize = 8
  rc = bs_autotune(sys, dtype, variant)
  call bs_check(rc, 'bs_autotune')

  niter = int(min(int(niter, c_int64_t), dim))
  nkeep = min(nkeep, niter)
  allocate(alpha(niter), beta(niter), evals(nkeep), resid(nkeep))
  p%nkeep = nkeep
  p%niter = niter
  p%thick = nthick
  p%max_restarts = maxrest
  p%dtype = dtype
  p%tol = 0.0d0
  if (lopt(2:2) == 'd' .or. lopt(2:2) == 'c') p%tol = 1.0d-6
  p%polish = 1
  nbytes = int(nkeep, c_int64_t) * dim * esize
  rc = hipMalloc(dvec, int(nbytes, c_size_t))
  if (rc /= 0) then
    write(6, *) ' hipMalloc failed ', rc
    stop 1
  end if
  rc = bs_lanczos(sys, p, c_loc(alpha), c_loc(beta), c_loc(evals), dvec, c_loc(resid), iters)
  call bs_check(rc, 'bs_lanczos')

  open(newunit=iu, file=trim(first_word(outname)) // '.res', status='replace')
  write(iu, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
  write(iu, '(a,a)') '  single-particle file = ', trim(first_word(spsname))
  write(iu, '(1x,i0,1x,i0)') z, n
  write(iu, '(1x,i0,a)') jz2, ' +'
  write(iu, '(1x,i0,a)') iters, '  iterations '
  write(iu, '(a)') '  '
  write(iu, '(a)') '  State      E        Ex         J       T      par'
  allocate(we(nkeep), wj(nkeep), wt2(nkeep))
  do k = 1, nkeep
    rc = bs_observable(sys, 0, ptr_offset(dvec, int(k - 1, c_int64_t) * dim * esize), dtype, j2)
    call bs_check(rc, 'bs_observable J^2')
    rc = bs_observable(sys, 1, ptr_offset(dvec, int(k - 1, c_int64_t) * dim * esize), dtype, t2)
    call bs_check(rc, 'bs_observable T^2')
    xj = real(-0.5d0 + sqrt(0.25d0 + max(j2, 0.0d0)))
    xt = real(-0.5d0 + sqrt(0.25d0 + max(t2, 0.0d0)))
    write(iu, 100) k, evals(k), evals(k) - evals(1), xj, xt, 1
    ! .wfn carries E, J and T(T+1) in single precision (src/blanczoslib.f90:1330-1350)
    we(k) = real(evals(k), c_float)
    wj(k) = xj
    wt2(k) = real(t2, c_float)
  end do
100 format(i5,3x,2f10.5,2x,2f8.3,3x,i2)
  close(iu)
  if (write_wfn) then
    rc = bs_write_wfn(sys, cstr(trim(first_word(outname)) // '.wfn'), nkeep, c_loc(we), c_loc(wj), c_loc(wt2), &
                      dvec, dtype, 1_c_int)
    call bs_check(rc, 'bs_write_wfn')
  end if
  open(newunit=iu, file=trim(first_word(outname)) // '.lcoef', status='replace')
  do k = 1, niter
    write(iu, '(i10,2(2x,f20.15))') k, alpha(k), beta(k)
  end do
  close(iu)
  write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', dim, ', ', iters, ' matvecs'
  do k = 1, nkeep
    write(6, '(i5,3x,f12.6,2x,es10.2)') k, evals(k), resid(k)
  end do
  rc = hipFree(dvec)
  rc = bs_finalize(sys)

contains
  ! the reference's menu lines may carry a trailing '! comment'
  function first_word(s) result(w)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: w
    integer :: i
    w = adjustl(s)
    i = index(w, ' ')
    if (i > 0) w = w(1:i - 1)
    i = index(w, '!')
    if (i > 0) w = w(1:i - 1)
  end function first_word
end program bs_driver
