! bs_driver: a Fortran host for the MI355X engine that reads BIGSTICK's own
! stdin menu answers for a normal run (menu option n, src/bmenu.f90:11;
! menu s, strength function from a pivot in a previous .wfn; menu x,
! expectation values in previous states; menu o, a one-body operator applied
! to previous states, below;
! interaction menu, src/binput_tbme.f90:178-573; Lanczos options lf/ld/tf/td,
! the block options bf/bd and exact diagonalization ex,
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
  character(len=256) :: opt

  read(5, '(a)') opt
  opt = adjustl(opt)
  ! a second letter s suppresses the .wfn (menus ns, ss: src/bmenu.f90:572-580, :680-687)
  select case (opt(1:1))
  case ('n', 'N')
    if (opt(2:2) == 'p' .or. opt(2:2) == 'P') then
      ! np: the spectrum from a pivot read from a previous .wfn (src/bmenu.f90:700-707)
      call pivot_run()
    else
      call diagonalize_run(.not. (opt(2:2) == 's' .or. opt(2:2) == 'S'))
    end if
  case ('s', 'S')
    if (opt(2:2) == 'b' .or. opt(2:2) == 'B') then
      ! sb, sbs: block strength function (src/bmenu.f90:612-627)
      call block_strength_run(opt(3:3) == 's' .or. opt(3:3) == 'S')
    else
      call strength_run(opt(2:2) == 's' .or. opt(2:2) == 'S')
    end if
  case ('x', 'X')
    call expectation_run(.false.)
  case ('h', 'H')
    call expectation_run(.true.)
  case ('o', 'O')
    ! oo runs the same applicator1b as o (src/bmenu.f90:465-478)
    if (len_trim(opt) > 2 .or. (len_trim(opt) == 2 .and. opt(2:2) /= 'o' .and. opt(2:2) /= 'O')) then
      write(6, *) ' bs_driver: of the o options only o and oo are supported'
      stop 1
    end if
    call onebody_run()
  case default
    write(6, *) ' bs_driver: supported menu options are n, ns, np (diagonalization), s, ss, sb, sbs (strength function),', &
                ' x (expectation values), h (operator matrix) and o, oo (one-body operator applied to states)'
    stop 1
  end select

contains
  ! Menu n / ns: Lanczos from the uniform pivot.
  subroutine diagonalize_run(write_wfn)
    logical, intent(in) :: write_wfn
    character(len=256) :: outname, spsname, intname, line, lopt
    integer :: z, n, jz2, nkeep, niter, nthick, maxrest, k, iu, dimblock, nbiter, ncheck, ctest, tgt
    integer(c_int) :: nrr, nsp, steps, nthr
    integer(8) :: t0, t1, t2, trate
    real(c_double) :: ctol, etarget
    real(c_float) :: spscale, a0, a, x
    type(c_ptr) :: sys, dvec
    integer(c_int64_t) :: dim, nbytes, esize
    integer(c_int) :: rc, iters, dtype, variant
    type(bs_lanczos_params) :: p
    real(c_double), allocatable, target :: alpha(:), beta(:), evals(:), resid(:)
    character(len=16) :: env

    read(5, '(a)') outname
    read(5, '(a)') spsname
    read(5, *) z, n
    read(5, *) jz2
    read(5, '(a)') intname
    spscale = 1.0
    a0 = 1.0
    a = 1.0
    x = 1.0
    ! an autoscaled .int (negative count) carries its own scaling; no line
    ! is read for it (src/binput_tbme.f90:1015)
    if (.not. int_autoscaled(trim(first_word(intname)) // '.int')) read(5, *) spscale, a0, a, x
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
      return
    end if
    read(5, '(a)') lopt
    lopt = lower(adjustl(lopt))
    select case (lopt(1:2))
    case ('ex', 'ld', 'lf', 'lc', 'td', 'tf', 'tc', 'tx', 'tb', 'bd', 'bf', 'bc', 'sk', 'li')
    case default
      write(6, *) ' bs_driver: diagonalization option ', trim(lopt), ' is not supported', &
                  ' (supported: ex, ld, lf, lc, td, tf, tc, tx, tb, bd, bf, bc, sk, li)'
      stop 1
    end select
    dimblock = 0
    if (lopt(1:2) == 'sk') then
      ! sk: set up only, to time it (src/blanczosmain.f90:1175-1179): tables
      ! built, uploaded and the kernel variant tuned; no diagonalization
      call system_clock(t0, trate)
      rc = bs_init(cstr(trim(first_word(spsname)) // '.sps'), cstr(trim(first_word(intname)) // '.int'), &
                   z, n, jz2, spscale, a0, a, x, sys)
      call bs_check(rc, 'bs_init')
      call system_clock(t1)
      rc = bs_autotune(sys, dtype_from_env(), variant)
      call bs_check(rc, 'bs_autotune')
      call system_clock(t2)
      rc = bs_dim(sys, dim)
      open(newunit=iu, file=trim(first_word(outname)) // '.res', status='replace')
      write(iu, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
      write(iu, '(1x,i0,1x,i0)') z, n
      write(iu, '(1x,i0,a)') jz2, ' +'
      write(iu, '(a,i0)') '  Total basis dimension: ', dim
      write(iu, '(a,f10.3)') '  Time to build and upload the tables (s): ', real(t1 - t0) / real(trate)
      write(iu, '(a,f10.3)') '  Time to tune the kernel (s): ', real(t2 - t1) / real(trate)
      close(iu)
      write(6, '(a,i0,a,f8.3,a)') ' bs_driver: dimension ', dim, ', setup ', real(t1 - t0) / real(trate), ' s'
      rc = bs_finalize(sys)
      return
    end if
    if (lopt(1:2) == 'ex') then
      ! exact diagonalization: nkeep only (src/blanczosmain.f90:754-763)
      read(5, *) nkeep
      niter = nkeep
    else if (lopt(1:2) == 'li') then
      ! li: Lanczos coefficients only, # iterations (src/blanczosmain.f90:787-795);
      ! the .lcoef and the .res header are written, no eigenstates
      read(5, *) niter
      nkeep = 1
    else if (lopt(1:1) == 'b' .or. lopt(1:2) == 'tb') then
      ! block Lanczos bf / bd: nkeep, block size, block iterations; the basis
      ! holds (iterations + 1) blocks (src/blanczosmain.f90:849-869); tb
      ! restarts there (src/blanczosmain.f90:1179-1212)
      read(5, *) nkeep, dimblock, nbiter
      niter = nbiter * dimblock + dimblock
    else
      read(5, *) nkeep, niter
    end if
    nthick = 0
    maxrest = 0
    tgt = 0
    etarget = 0.0d0
    if (lopt(1:2) == 'tb') then
      ! tb: niter raised to at least max((nkeep + 5) * 2, 10), then the kept
      ! vectors (at least dimblock + 2) and max restarts; test 0 at 1e-3 over
      ! nkeep + 2 energies (src/blanczosmain.f90:1197-1212)
      niter = max(niter, max((nkeep + 5) * 2, 10))
      read(5, *) nthick
      if (nthick < dimblock + 2) nthick = dimblock + 2
      read(5, *) maxrest
    else if (lopt(1:2) == 'tx') then
      ! tx: thick restart keeping the states nearest a target energy
      ! (src/blanczosmain.f90:1022-1053): niter raised to at least
      ! max((nkeep + 5) * 2, 10), kept vectors at least nkeep, restarts,
      ! target energy; test 0 with tolerance 1e-3 over the nkeep energies
      niter = max(niter, max((nkeep + 5) * 2, 10))
      read(5, *) nthick
      if (nthick < nkeep) nthick = nkeep
      read(5, *) maxrest
      read(5, *) etarget
      tgt = 1
    else if (lopt(1:2) == 'td') then
      ! td: no count of kept vectors is asked (src/blanczosmain.f90:993-1019):
      ! niter raised to at least max((nkeep + 5) * 2, 10), max(nkeep + 5,
      ! 3 nkeep) vectors kept but at most niter - 5; then max restarts
      niter = max(niter, max((nkeep + 5) * 2, 10))
      nthick = min(max(nkeep + 5, 3 * nkeep), niter - 5)
      read(5, *) maxrest
    else if (lopt(1:1) == 't') then
      ! tf, tc: kept vectors (tf: at least nkeep), restarts
      ! (src/blanczosmain.f90:1091-1145)
      read(5, *) nthick
      if (lopt(1:2) == 'tf' .and. nthick < nkeep) nthick = nkeep
      read(5, *) maxrest
    end if
    ! default convergence (ld, td): the reference's test 0 with tolerance
    ! converge_tol0_def = 1e-3 over nkeep + ncheck_ex energies, ncheck_ex =
    ! 5 for ld (src/blanczosmain.f90:800-812, :1000-1020, bmodules_flags.f90:38)
    ctol = 1.0d-3
    ctest = 0
    ncheck = 0
    if (lopt(1:2) == 'ld') ncheck = 5
    if (lopt(1:2) == 'tb') ncheck = 2
    if (lopt(1:1) == 'b') ctol = 1.0d-6
    if (lopt(2:2) == 'c' .or. lopt(2:2) == 'C') then
      ! lc / bc / tc: [additional states for the test,] test choice, tolerance
      ! (src/blanczosmain.f90:943-985, :873-921, :1117-1175): the energy tests
      ! 0 (mean change) and 1 (largest change), and for lc / tc the Ritz-vector
      ! tests 2 (mean change) and 3 (largest change)
      if (lopt(1:1) /= 't') read(5, *) ncheck
      read(5, *) ctest
      if (ctest < 0 .or. ctest > 3 .or. (lopt(1:1) == 'b' .and. ctest > 1)) then
        write(6, *) ' bs_driver: convergence test not supported here (0-3 for lc and tc, 0-1 for bc)'
        stop 1
      end if
      read(5, *) ctol
      if (.not. (ctol > 0.0d0)) then
        write(6, *) ' bs_driver: the tolerance must be positive'
        stop 1
      end if
    end if

    rc = bs_init(cstr(trim(first_word(spsname)) // '.sps'), cstr(trim(first_word(intname)) // '.int'), &
                 z, n, jz2, spscale, a0, a, x, sys)
    call bs_check(rc, 'bs_init')
    rc = bs_dim(sys, dim)
    dtype = dtype_from_env()
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
    if (lopt(2:2) == 'd' .or. lopt(2:2) == 'c' .or. lopt(2:2) == 'x' .or. lopt(1:2) == 'tb') p%tol = ctol
    p%target = tgt
    p%etarget = etarget
    p%conv_test = ctest
    p%ncheck_ex = ncheck
    if (lopt(1:1) == 'b') p%ncheck_ex = 0
    p%polish = 1
    nbytes = int(nkeep, c_int64_t) * dim * esize
    dvec = device_alloc(nbytes)
    alpha = 0.0d0
    beta = 0.0d0
    if (lopt(1:2) == 'ex') then
      rc = bs_exact(sys, nkeep, dtype, c_loc(evals), dvec, c_loc(resid), iters)
      call bs_check(rc, 'bs_exact')
    else if (dimblock > 0) then
      rc = bs_block_lanczos(sys, p, dimblock, c_loc(evals), dvec, c_loc(resid), iters)
      call bs_check(rc, 'bs_block_lanczos')
    else if (lopt(1:2) == 'li') then
      p%polish = 0
      resid = 0.0d0
      rc = bs_lanczos(sys, p, c_loc(alpha), c_loc(beta), c_loc(evals), c_null_ptr, c_null_ptr, iters)
      call bs_check(rc, 'bs_lanczos')
    else
      rc = bs_lanczos(sys, p, c_loc(alpha), c_loc(beta), c_loc(evals), dvec, c_loc(resid), iters)
      call bs_check(rc, 'bs_lanczos')
    end if
    steps = 0
    nthr = 0
    if (dimblock == 0 .and. lopt(1:2) /= 'ex') then
      rc = bs_lanczos_stats(sys, nrr, nsp, steps, nthr)
      call bs_check(rc, 'bs_lanczos_stats')
    end if

    open(newunit=iu, file=trim(first_word(outname)) // '.res', status='replace')
    write(iu, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(iu, '(a,a)') '  single-particle file = ', trim(first_word(spsname))
    write(iu, '(1x,i0,1x,i0)') z, n
    write(iu, '(1x,i0,a)') jz2, ' +'
    if (steps > 0) then
      ! the size of the final projected matrix and, after thick restarts, the
      ! reference's total iter + niter * restarts (src/blanczosmain.f90:1762-1770)
      write(iu, '(1x,i0,a)') steps, '  iterations '
      write(iu, '(a)') '  '
      if (nthr > 0) write(iu, '(a,i0,a)') '  ( total =  ', steps + niter * nthr, '  ) '
    else
      write(iu, '(1x,i0,a)') iters, '  iterations '
      write(iu, '(a)') '  '
    end if
    if (lopt(1:2) /= 'li') &
      call state_table(iu, sys, dvec, nkeep, evals, dtype, dim, write_wfn, trim(first_word(outname)) // '.wfn')
    close(iu)
    if (dimblock == 0 .and. lopt(1:2) /= 'ex') then
      ! a plain run that converged stopped after `steps` iterations, where
      ! the reference's .lcoef ends too; after a thick restart the file holds
      ! the kept Ritz values and their couplings as lines -1 .. -nthick, then
      ! the steps of the last cycle (src/blanczoslib.f90:1750-1757)
      open(newunit=iu, file=trim(first_word(outname)) // '.lcoef', status='replace')
      do k = 1, min(niter, int(steps))
        if (nthr > 0 .and. k <= nthick) then
          write(iu, '(i10,2(2x,f20.15))') -k, alpha(k), beta(k)
        else
          write(iu, '(i10,2(2x,f20.15))') k, alpha(k), beta(k)
        end if
      end do
      close(iu)
    end if
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', dim, ', ', iters, ' matvecs'
    do k = 1, nkeep
      write(6, '(i5,3x,f12.6,2x,es10.2)') k, evals(k), resid(k)
    end do
    rc = hipFree(dvec)
    rc = bs_finalize(sys)
  end subroutine diagonalize_run

  ! Menu s / ss: strength function.  stdin: s, .wfn holding the pivot, output
  ! name, .int name, scaling line, end, nkeep and the number of (fixed)
  ! iterations (src/blanczosmain.f90:230-250), the pivot's index in the
  ! .wfn (readpivot, src/bwfnlib.f90:298-366).  <name>.res gets the total
  ! strength, the Energy/Strength table of every Ritz value
  ! (src/blanczoslib.f90:1122-1142) and, for s, the state table and
  ! <name>.wfn of the normalized Ritz vectors; ss writes neither
  ! (src/bmenu.f90:572-580).
  subroutine strength_run(suppress)
    logical, intent(in) :: suppress
    character(len=256) :: wfnname, oname, iname, l
    character(len=:), allocatable :: wpath
    real(c_float) :: s1, s2, s3, s4
    type(c_ptr) :: sy, dpiv, dvec
    integer(c_int64_t) :: d, esize
    integer(c_int) :: r, nk, nfile, ipiv, niter, iters, dtype, variant
    type(bs_lanczos_params) :: p
    real(c_float), target :: pe(1), pj(1), pt(1)
    real(c_double), allocatable, target :: ritz(:), stren(:), alpha(:), beta(:), evals(:), resid(:)
    real(c_double), target :: pnorm
    integer :: j, u

    read(5, '(a)') wfnname
    read(5, '(a)') oname
    read(5, '(a)') iname
    s1 = 1.0
    s2 = 1.0
    s3 = 1.0
    s4 = 1.0
    if (.not. int_autoscaled(trim(first_word(iname)) // '.int')) read(5, *) s1, s2, s3, s4
    read(5, '(a)') l
    if (adjustl(l) /= 'end') then
      write(6, *) ' bs_driver: one interaction file is supported'
      stop 1
    end if
    read(5, *) nk, niter
    read(5, *) ipiv
    wpath = trim(first_word(wfnname)) // '.wfn'
    r = bs_init_wfn(cstr(wpath), cstr(trim(first_word(iname)) // '.int'), s1, s2, s3, s4, sy)
    call bs_check(r, 'bs_init_wfn')
    r = bs_dim(sy, d)
    r = bs_read_wfn(sy, cstr(wpath), 0, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, BS_F32, 0, nfile)
    call bs_check(r, 'bs_read_wfn')
    if (ipiv < 1 .or. ipiv > nfile) then
      write(6, '(a,i0)') ' Selected vector is not in range 1:', nfile
      stop 1
    end if
    dtype = dtype_from_env()
    esize = 4
    if (dtype == BS_F64) esize = 8
    dpiv = device_alloc(d * esize)
    r = bs_read_wfn(sy, cstr(wpath), ipiv - 1, 1, c_loc(pe), c_loc(pj), c_loc(pt), dpiv, dtype, 1, nfile)
    call bs_check(r, 'bs_read_wfn')
    r = bs_autotune(sy, dtype, variant)
    call bs_check(r, 'bs_autotune')

    niter = int(min(int(niter, c_int64_t), d))
    nk = min(nk, niter)
    allocate(ritz(niter), stren(niter), alpha(niter), beta(niter), evals(nk), resid(nk))
    p%nkeep = nk
    p%niter = niter
    p%thick = 0
    p%max_restarts = 0
    p%dtype = dtype
    p%tol = 0.0d0
    p%polish = 0     ! the table prints Ritz values, as the reference does
    dvec = device_alloc(int(nk, c_int64_t) * d * esize)
    r = bs_strength(sy, p, dpiv, pnorm, c_loc(ritz), c_loc(stren), c_loc(alpha), c_loc(beta), c_loc(evals), &
                    dvec, c_loc(resid), iters)
    call bs_check(r, 'bs_strength')

    open(newunit=u, file=trim(first_word(oname)) // '.res', status='replace')
    write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(u, '(a,a,a,i0)') '  pivot read from ', wpath, ', state ', ipiv
    write(u, *) pnorm * pnorm, ' = total strength '
    write(u, '(1x,i0,a)') niter, '  iterations '
    write(u, '(a)') '  '
    write(u, *) ' Energy   Strength    '
    write(u, *) ' ______   ________  '
    do j = 1, niter
      write(u, 88) ritz(j), stren(j)
    end do
88  format(3f12.6)
    write(u, *) '_______   ________   '
    write(u, *) pnorm**2, '= total strength, included above '
    write(u, *) ' '
    if (suppress) then
      write(u, *) ' Note: not writing wfns to file, nor computing J, T '
    else
      call state_table(u, sy, dvec, nk, evals, dtype, d, .true., trim(first_word(oname)) // '.wfn')
    end if
    close(u)
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', iters, ' matvecs'
    r = hipFree(dvec)
    r = hipFree(dpiv)
    r = bs_finalize(sy)
  end subroutine strength_run

  ! Menu np with option lf: Lanczos from one state of a previous .wfn as the
  ! pivot (readpivot, src/bwfnlib.f90:298-366) for a fixed number of
  ! iterations, then the usual .res state table, .wfn and .lcoef.
  ! stdin: np, .wfn name, output name, .int name, scaling line, end, lf,
  ! nkeep niter, which state is the pivot.
  subroutine pivot_run()
    character(len=256) :: wfnname, oname, iname, l, lopt
    character(len=:), allocatable :: wpath
    real(c_float) :: s1, s2, s3, s4
    type(c_ptr) :: sy, dpiv, dvec
    integer(c_int64_t) :: d, esize
    integer(c_int) :: r, nk, nfile, ipiv, niter, iters, dtype, variant
    type(bs_lanczos_params) :: p
    real(c_float), target :: pe(1), pj(1), pt(1)
    real(c_double), allocatable, target :: ritz(:), stren(:), alpha(:), beta(:), evals(:), resid(:)
    real(c_double), target :: pnorm
    integer :: j, u

    read(5, '(a)') wfnname
    read(5, '(a)') oname
    read(5, '(a)') iname
    s1 = 1.0
    s2 = 1.0
    s3 = 1.0
    s4 = 1.0
    if (.not. int_autoscaled(trim(first_word(iname)) // '.int')) read(5, *) s1, s2, s3, s4
    read(5, '(a)') l
    if (adjustl(l) /= 'end') then
      write(6, *) ' bs_driver: one interaction file is supported'
      stop 1
    end if
    read(5, '(a)') lopt
    lopt = adjustl(lopt)
    if (lopt(1:2) /= 'lf' .and. lopt(1:2) /= 'LF' .and. lopt(1:2) /= 'Lf') then
      write(6, *) ' bs_driver: menu np supports the fixed-iteration option lf'
      stop 1
    end if
    read(5, *) nk, niter
    read(5, *) ipiv
    wpath = trim(first_word(wfnname)) // '.wfn'
    r = bs_init_wfn(cstr(wpath), cstr(trim(first_word(iname)) // '.int'), s1, s2, s3, s4, sy)
    call bs_check(r, 'bs_init_wfn')
    r = bs_dim(sy, d)
    r = bs_read_wfn(sy, cstr(wpath), 0, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, BS_F32, 0, nfile)
    call bs_check(r, 'bs_read_wfn')
    if (ipiv < 1 .or. ipiv > nfile) then
      write(6, '(a,i0)') ' Selected vector is not in range 1:', nfile
      stop 1
    end if
    dtype = dtype_from_env()
    esize = 4
    if (dtype == BS_F64) esize = 8
    dpiv = device_alloc(d * esize)
    r = bs_read_wfn(sy, cstr(wpath), ipiv - 1, 1, c_loc(pe), c_loc(pj), c_loc(pt), dpiv, dtype, 1, nfile)
    call bs_check(r, 'bs_read_wfn')
    r = bs_autotune(sy, dtype, variant)
    call bs_check(r, 'bs_autotune')
    niter = int(min(int(niter, c_int64_t), d))
    nk = min(nk, niter)
    allocate(ritz(niter), stren(niter), alpha(niter), beta(niter), evals(nk), resid(nk))
    p%nkeep = nk
    p%niter = niter
    p%thick = 0
    p%max_restarts = 0
    p%dtype = dtype
    p%tol = 0.0d0
    p%polish = 1
    alpha = 0.0d0
    beta = 0.0d0
    dvec = device_alloc(int(nk, c_int64_t) * d * esize)
    r = bs_strength(sy, p, dpiv, pnorm, c_loc(ritz), c_loc(stren), c_loc(alpha), c_loc(beta), c_loc(evals), &
                    dvec, c_loc(resid), iters)
    call bs_check(r, 'bs_strength')
    open(newunit=u, file=trim(first_word(oname)) // '.res', status='replace')
    write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(u, '(a,a,a,i0)') '  pivot read from ', wpath, ', state ', ipiv
    write(u, '(1x,i0,a)') niter, '  iterations '
    write(u, '(a)') '  '
    call state_table(u, sy, dvec, nk, evals, dtype, d, .true., trim(first_word(oname)) // '.wfn')
    close(u)
    open(newunit=u, file=trim(first_word(oname)) // '.lcoef', status='replace')
    do j = 1, niter
      write(u, '(i10,2(2x,f20.15))') j, alpha(j), beta(j)
    end do
    close(u)
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', iters, ' matvecs'
    r = hipFree(dvec)
    r = hipFree(dpiv)
    r = bs_finalize(sy)
  end subroutine pivot_run

  ! Menus sb / sbs: block Lanczos from a pivot block of states of a previous
  ! .wfn (read_pivot_block_choose, src/blanczoslib.f90:522-707), fixed block
  ! iterations; per pivot, its total strength and its Energy/Strength table
  ! over every Ritz value (src/blanczoslib.f90:1146-1181), then (sb) the
  ! state table and the .wfn of the Ritz vectors.
  ! stdin: sb, .wfn name, output name, .int name, scaling line, end,
  ! nkeep dimblock iterations (nkeep raised to dimblock*(iterations+1)),
  ! c + first last (a contiguous list) or s + dimblock state numbers.
  subroutine block_strength_run(suppress)
    logical, intent(in) :: suppress
    character(len=256) :: wfnname, oname, iname, l
    character(len=:), allocatable :: wpath
    real(c_float) :: s1, s2, s3, s4
    type(c_ptr) :: sy, dpiv, dvec
    integer(c_int64_t) :: d, esize
    integer(c_int) :: r, nk, nfile, kb, nbiter, iters, dtype, variant, nf
    integer(c_int), allocatable :: choice(:)
    type(bs_lanczos_params) :: p
    real(c_float), target :: pe(1), pj(1), pt(1)
    real(c_double), allocatable, target :: stren(:), evals(:), resid(:), pn2(:)
    character(len=16) :: listchar
    integer :: i, j, u, c0, c1

    read(5, '(a)') wfnname
    read(5, '(a)') oname
    read(5, '(a)') iname
    s1 = 1.0
    s2 = 1.0
    s3 = 1.0
    s4 = 1.0
    if (.not. int_autoscaled(trim(first_word(iname)) // '.int')) read(5, *) s1, s2, s3, s4
    read(5, '(a)') l
    if (adjustl(l) /= 'end') then
      write(6, *) ' bs_driver: one interaction file is supported'
      stop 1
    end if
    read(5, *) nk, kb, nbiter
    if (kb < 1 .or. kb > 8 .or. nbiter < 0) then
      write(6, *) ' bs_driver: block dimension must be 1..8 and iterations >= 0'
      stop 1
    end if
    nk = max(nk, kb * (nbiter + 1))
    wpath = trim(first_word(wfnname)) // '.wfn'
    r = bs_init_wfn(cstr(wpath), cstr(trim(first_word(iname)) // '.int'), s1, s2, s3, s4, sy)
    call bs_check(r, 'bs_init_wfn')
    r = bs_dim(sy, d)
    r = bs_read_wfn(sy, cstr(wpath), 0, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, BS_F32, 0, nfile)
    call bs_check(r, 'bs_read_wfn')
    allocate(choice(kb))
    read(5, '(a)') listchar
    select case (adjustl(listchar))
    case ('c', 'C')
      read(5, *) c0, c1
      if (c1 - c0 + 1 /= kb .or. c0 < 1 .or. c1 > nfile) then
        write(6, '(a,i0,a,i0,a)') ' bs_driver: the pivot list must name ', kb, ' of the ', nfile, &
                                  ' states (random fill is not supported)'
        stop 1
      end if
      choice = [(c0 + i - 1, i = 1, kb)]
    case ('s', 'S')
      read(5, *) choice
      if (any(choice < 1) .or. any(choice > nfile)) then
        write(6, '(a,i0,a)') ' bs_driver: pivot states must lie in 1:', nfile, ' (random fill is not supported)'
        stop 1
      end if
    case default
      write(6, *) ' bs_driver: pivot list must be c (contiguous) or s (specified)'
      stop 1
    end select
    dtype = dtype_from_env()
    esize = 4
    if (dtype == BS_F64) esize = 8
    dpiv = device_alloc(int(kb, c_int64_t) * d * esize)
    do i = 1, kb
      r = bs_read_wfn(sy, cstr(wpath), choice(i) - 1, 1, c_loc(pe), c_loc(pj), c_loc(pt), &
                      ptr_offset(dpiv, int(i - 1, c_int64_t) * d * esize), dtype, 1, nf)
      call bs_check(r, 'bs_read_wfn')
    end do
    r = bs_autotune(sy, dtype, variant)
    call bs_check(r, 'bs_autotune')
    ! basis size dimblock*(iterations+1), capped by the dimension
    nk = int(min(int(nk, c_int64_t), (d / kb) * kb))
    allocate(stren(int(kb) * nk), evals(nk), resid(nk), pn2(kb))
    p%nkeep = nk
    p%niter = kb * (nbiter + 1)
    p%thick = 0
    p%max_restarts = 0
    p%dtype = dtype
    p%tol = 0.0d0
    p%polish = 0     ! the tables print Ritz values, as the reference does
    dvec = device_alloc(int(nk, c_int64_t) * d * esize)
    r = bs_block_strength(sy, p, kb, dpiv, dtype, c_loc(pn2), c_loc(stren), c_loc(evals), dvec, c_loc(resid), iters)
    call bs_check(r, 'bs_block_strength')

    open(newunit=u, file=trim(first_word(oname)) // '.res', status='replace')
    write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(u, '(a,a)') '  pivot block read from ', wpath
    write(u, '(1x,i0,a)') nk, '  iterations '
    write(u, '(a)') '  '
    do i = 1, kb
      write(u, *) '###################'
      write(u, *) ' For pivot state ', choice(i), ' total strength = ', pn2(i)
      write(u, *) ' Energy   Strength  '
      write(u, *) ' ______   ________  '
      do j = 1, nk
        write(u, 88) evals(j), stren((i - 1) * nk + j)
      end do
    end do
88  format(3f12.6)
    if (suppress) then
      write(u, *) ' Note: not writing wfns to file, nor computing J, T '
    else
      call state_table(u, sy, dvec, nk, evals, dtype, d, .true., trim(first_word(oname)) // '.wfn')
    end if
    close(u)
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', iters, ' matvecs'
    r = hipFree(dvec)
    r = hipFree(dpiv)
    r = bs_finalize(sy)
  end subroutine block_strength_run

  ! Menu o: the one-body operator of an .opme file applied to every state of
  ! a previous .wfn; <name>.wfn gets O|state>, not normalized, with each
  ! state's E, J, T(T+1) (src/bmenu.f90:423-434, src/bdenslib1.f90:1645-1693).
  ! stdin: o, .wfn name, output name, .opme name.
  subroutine onebody_run()
    character(len=256) :: wfnname, oname, opname
    character(len=:), allocatable :: wpath
    type(c_ptr) :: sy, dx, dy
    integer(c_int64_t) :: d, esize
    integer(c_int) :: r, nk, nk2, dtype, jop, top
    real(c_float), allocatable, target :: e(:), ej(:), et(:)
    integer :: u

    read(5, '(a)') wfnname
    read(5, '(a)') oname
    read(5, '(a)') opname
    wpath = trim(first_word(wfnname)) // '.wfn'
    r = bs_init_wfn_host(cstr(wpath), c_null_ptr, 1.0, 1.0, 1.0, 1.0, sy)
    call bs_check(r, 'bs_init_wfn_host')
    r = bs_dim(sy, d)
    r = bs_read_wfn(sy, cstr(wpath), 0, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, BS_F32, 0, nk)
    call bs_check(r, 'bs_read_wfn')
    if (nk < 1) then
      write(6, *) ' bs_driver: no states in ', trim(wpath)
      stop 1
    end if
    dtype = dtype_from_env()
    esize = 4
    if (dtype == BS_F64) esize = 8
    allocate(e(nk), ej(nk), et(nk))
    dx = device_alloc(int(nk, c_int64_t) * d * esize)
    dy = device_alloc(int(nk, c_int64_t) * d * esize)
    r = bs_read_wfn(sy, cstr(wpath), 0, nk, c_loc(e), c_loc(ej), c_loc(et), dx, dtype, 1, nk2)
    call bs_check(r, 'bs_read_wfn')
    r = bs_apply_onebody(sy, cstr(trim(first_word(opname)) // '.opme'), dx, dy, nk, dtype, jop, top)
    call bs_check(r, 'bs_apply_onebody')
    r = bs_write_wfn(sy, cstr(trim(first_word(oname)) // '.wfn'), nk, c_loc(e), c_loc(ej), c_loc(et), dy, dtype, 1)
    call bs_check(r, 'bs_write_wfn')
    open(newunit=u, file=trim(first_word(oname)) // '.res', status='replace')
    write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(u, '(a,a)') '  states read from ', wpath
    write(u, '(a,a,a,i0,a,i0)') '  one-body operator ', trim(first_word(opname)) // '.opme', ', J = ', jop, &
                                ', T = ', top
    write(u, '(1x,i0,a)') nk, '  states '
    close(u)
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', nk, ' states'
    r = hipFree(dy)
    r = hipFree(dx)
    r = bs_finalize(sy)
  end subroutine onebody_run

  ! The State/E/Ex/J/T/par table (src/bparticleocc.f90:256) of nk device
  ! vectors with J and T from <J^2>, <T^2>, and optionally their .wfn.
  subroutine state_table(iu, sy, dv, nk, ev, dtype, d, write_wfn, wfnpath)
    integer, intent(in) :: iu, nk
    type(c_ptr), intent(in) :: sy, dv
    real(c_double), intent(in) :: ev(nk)
    integer(c_int), intent(in) :: dtype
    integer(c_int64_t), intent(in) :: d
    logical, intent(in) :: write_wfn
    character(len=*), intent(in) :: wfnpath
    real(c_float), allocatable, target :: we(:), wj(:), wt2(:)
    real(c_double) :: j2, t2
    real :: xj, xt
    integer(c_int64_t) :: esize
    integer(c_int) :: rc, ipar
    integer :: k

    esize = 4
    if (dtype == BS_F64) esize = 8
    rc = bs_parity(sy, ipar)
    write(iu, '(a)') '  State      E        Ex         J       T      par'
    allocate(we(nk), wj(nk), wt2(nk))
    do k = 1, nk
      rc = bs_observable(sy, 0, ptr_offset(dv, int(k - 1, c_int64_t) * d * esize), dtype, j2)
      call bs_check(rc, 'bs_observable J^2')
      rc = bs_observable(sy, 1, ptr_offset(dv, int(k - 1, c_int64_t) * d * esize), dtype, t2)
      call bs_check(rc, 'bs_observable T^2')
      xj = real(-0.5d0 + sqrt(0.25d0 + max(j2, 0.0d0)))
      xt = real(-0.5d0 + sqrt(0.25d0 + max(t2, 0.0d0)))
      write(iu, 100) k, ev(k), ev(k) - ev(1), xj, xt, ipar
      ! .wfn carries E, J and T(T+1) in single precision (src/blanczoslib.f90:1330-1350)
      we(k) = real(ev(k), c_float)
      wj(k) = xj
      wt2(k) = real(t2, c_float)
    end do
100 format(i5,3x,2f10.5,2x,2f8.3,3x,i2)
    if (write_wfn) then
      rc = bs_write_wfn(sy, cstr(wfnpath), nk, c_loc(we), c_loc(wj), c_loc(wt2), dv, dtype, 1_c_int)
      call bs_check(rc, 'bs_write_wfn')
    end if
  end subroutine state_table

  ! vector precision: FP32 (default) or FP64 with environment BS_DTYPE=f64
  function dtype_from_env() result(dt)
    integer(c_int) :: dt
    character(len=16) :: env
    call get_environment_variable('BS_DTYPE', env)
    dt = BS_F32
    if (trim(env) == 'f64') dt = BS_F64
  end function dtype_from_env

  function device_alloc(nbytes) result(ptr)
    integer(c_int64_t), intent(in) :: nbytes
    type(c_ptr) :: ptr
    integer(c_int) :: rc
    rc = hipMalloc(ptr, int(nbytes, c_size_t))
    if (rc /= 0) then
      write(6, *) ' hipMalloc failed ', rc
      stop 1
    end if
  end function device_alloc

  ! Menu x: expectation value of a scalar two-body operator (an .int file)
  ! in each state of a previous .wfn (src/bmenu.f90:520-526,
  ! src/bexpect.f90:16-74).  stdin: x, .wfn name, output name, .int name,
  ! scaling line, end.  Each state is normalized first (dnormvec_p,
  ! src/bvectorlib.f90:200-264); <name>.res gets the reference's table.
  !
  ! Menu h (matrix): the matrix <f|O|i> of the raw stored states instead
  ! (hmatrixelements_boss, src/bexpect.f90:84-180): <name>.xme gets the state
  ! table and, for f <= i with equal J, every element above 1e-6, then -1 -1 0.
  subroutine expectation_run(matrix)
    logical, intent(in) :: matrix
    character(len=256) :: wfnname, oname, iname, l
    character(len=:), allocatable :: wpath
    real(c_float) :: s1, s2, s3, s4
    type(c_ptr) :: sy, dv
    integer(c_int64_t) :: d
    integer(c_int) :: r, nk, nk2
    real(c_float), allocatable, target :: e(:), ej(:), et(:)
    real(c_double), allocatable, target :: hv(:, :), hme(:, :)
    real(c_double), allocatable :: dn(:)
    real :: h4
    integer(c_int) :: ipar
    integer :: i, f, u

    read(5, '(a)') wfnname
    read(5, '(a)') oname
    read(5, '(a)') iname
    s1 = 1.0
    s2 = 1.0
    s3 = 1.0
    s4 = 1.0
    if (.not. int_autoscaled(trim(first_word(iname)) // '.int')) read(5, *) s1, s2, s3, s4
    read(5, '(a)') l
    if (adjustl(l) /= 'end') then
      write(6, *) ' bs_driver: one interaction file is supported'
      stop 1
    end if
    wpath = trim(first_word(wfnname)) // '.wfn'
    r = bs_init_wfn(cstr(wpath), cstr(trim(first_word(iname)) // '.int'), s1, s2, s3, s4, sy)
    call bs_check(r, 'bs_init_wfn')
    r = bs_dim(sy, d)
    r = bs_read_wfn(sy, cstr(wpath), 0, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, BS_F64, 0, nk)
    call bs_check(r, 'bs_read_wfn')
    if (nk < 1) then
      write(6, *) ' bs_driver: no states in ', trim(wpath)
      stop 1
    end if
    allocate(e(nk), ej(nk), et(nk), hv(d, nk), hme(nk, nk), dn(nk))
    r = bs_read_wfn(sy, cstr(wpath), 0, nk, c_loc(e), c_loc(ej), c_loc(et), c_loc(hv), BS_F64, 0, nk2)
    call bs_check(r, 'bs_read_wfn')
    do i = 1, nk
      dn(i) = sqrt(sum(hv(:, i)**2))
      if (dn(i) > 0.0d0 .and. .not. matrix) hv(:, i) = hv(:, i) / dn(i)
    end do
    dv = device_alloc(8 * d * nk)
    r = hipMemcpy(dv, c_loc(hv), int(8 * d * nk, c_size_t), 1_c_int)
    if (r /= 0) then
      write(6, *) ' hipMemcpy failed ', r
      stop 1
    end if
    ! hme(f, i) = <v_f | O | v_i>; the diagonal is the expectation value
    r = bs_operator_matrix(sy, c_null_ptr, 1.0, 1.0, 1.0, 1.0, dv, nk, BS_F64, c_loc(hme))
    call bs_check(r, 'bs_operator_matrix')
    if (matrix) then
      open(newunit=u, file=trim(first_word(oname)) // '.xme', status='replace')
      write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
      write(u, '(a,a)') '  states read from ', wpath
      write(u, *) ' '
      write(u, '(a)') '  State      E        Ex         J       T      par'
      r = bs_parity(sy, ipar)
      do i = 1, nk
        write(u, 111) i, e(i), e(i) - e(1), ej(i), et(i), ipar
      end do
111   format(i5,3x,2f10.5,2x,2f8.3,3x,i2)
      write(u, *) ' '
      do i = 1, nk
        do f = 1, i
          if (abs(ej(i) - ej(f)) > 0.0001) cycle
          h4 = real(hme(f, i))
          if (abs(h4) > 1e-6) write(u, *) i, f, h4
        end do
      end do
      write(u, *) -1, -1, 0.0
      close(u)
      write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', nk, ' states'
      r = hipFree(dv)
      r = bs_finalize(sy)
      return
    end if
    open(newunit=u, file=trim(first_word(oname)) // '.res', status='replace')
    write(u, '(a)') '  bigstick_amd MI355X engine (BIGSTICK v8.0.0 input surface)'
    write(u, '(a,a)') '  states read from ', wpath
    write(u, '(''nkeep ='',1x,i10)') nk
    write(u, '(a)') '   STATE      E            J           T^2           <H >      (norm)'
    do i = 1, nk
      write(u, 101) i, e(i), ej(i), et(i), hme(i, i), dn(i)
    end do
101 format(i5,2x,2(1x,f11.4),2(1x,f13.6),1x,f10.5)
    close(u)
    write(6, '(a,i0,a,i0,a)') ' bs_driver: dimension ', d, ', ', nk, ' states'
    r = hipFree(dv)
    r = bs_finalize(sy)
  end subroutine expectation_run

  ! .int files whose first data line starts with a negative count are
  ! autoscaled (src/binput_tbme.f90:1005-1011)
  logical function int_autoscaled(path)
    character(len=*), intent(in) :: path
    character(len=512) :: l
    integer :: u, ios, nme
    int_autoscaled = .false.
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do
      read(u, '(a)', iostat=ios) l
      if (ios /= 0) exit
      l = adjustl(l)
      if (len_trim(l) == 0 .or. l(1:1) == '!' .or. l(1:1) == '#') cycle
      read(l, *, iostat=ios) nme
      if (ios == 0) int_autoscaled = nme < 0
      exit
    end do
    close(u)
  end function int_autoscaled

  ! the reference's menu lines may carry a trailing '! comment'
  ! ASCII lower case (the reference accepts LF, Lf and lf alike)
  function lower(s) result(w)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: w
    integer :: i, c
    w = s
    do i = 1, len(s)
      c = ichar(s(i:i))
      if (c >= ichar('A') .and. c <= ichar('Z')) w(i:i) = char(c + 32)
    end do
  end function lower

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
