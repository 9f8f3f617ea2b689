!  Fortran host driving the gfx950 sweep engine through alfq_mod: the Hubbard
!  model on a 4-site ring (t = 1, U = 4, beta = 2, dtau = 0.1, Nwrap = 5, N_SUN = 2,
!  density-channel four-state HS field), with exp(-dtau T) in closed form.
!
!  It follows the reference's sequential sweep at the granularity of the shims
!  for the first up pass slices (alfq_WRAPGRUP_shim with G on the host, as
!  main.F90 calls WRAPGRUP) and then runs whole sweeps on the device.
!
!  With a third file name it runs the ring again with the checkerboard hopping
!  (two families of bonds on disjoint sites, E_b = exp(dtau t_b sigma_x), with
!  unequal t_b so that the families do not commute and the bond order matters)
!  set through alfq_set_checkerboard, the measurement geometry of the ring, and
!  TAU_M after each of two sweeps, as main.F90 does with Ltau = 1.
!
!  usage: alfq_ring_demo <tape file> <output file> [<checkerboard output file>]
!    tape file: count, then that many uniforms (one per line)
!    output:    final fields, then G (re im per line, column-major), PHASE,
!               and the NC_up / ACC_up counters
!    checkerboard output: the same, then N, NCG_tau, the sign sum, XMAX_tau and
!               the time-displaced sums (Green, SpinZ, Den) x (tau = 0..Ltrot)
!               x (distance 0..3), then Den0, re im per line
program alfq_ring_demo
  use iso_c_binding
  use alfq_mod
  implicit none
  integer, parameter :: n = 4, ltrot = 20, nwrap = 5, n_op = 4
  real(c_double), target :: expT(n, n), expTi(n, n), opE(n_op)
  integer(c_int32_t), target :: opP(n_op), optype(n_op), f(n_op, ltrot)
  complex(c_double_complex), target :: opg(n_op), opalpha(n_op)
  complex(kind=8) :: GR(n, n), phase, z
  real(c_double), allocatable :: tape(:)
  type(alfq_model_desc) :: desc
  type(alfq_control) :: ctl
  type(c_ptr) :: ctx
  ! checkerboard bonds (1,2) (3,4) | (2,3) (4,1) and the ring's measurement tables
  integer(c_int32_t), target :: famlen(2), bp(2, n), invlist(n), imj(n, n), orow(2*n), ocol(2*n)
  complex(c_double_complex), target :: expo(2, 2, n), expoi(2, 2, n), oval(2*n)
  real(c_double), target :: tlatt(2, n, 0:ltrot, 3), tden0(2)
  type(alfq_checkerboard_desc) :: cbd
  type(alfq_obs_desc) :: od
  type(alfq_tau_sums) :: ts
  integer :: i, j, d, st, nt, ntape, unit_in, unit_out, b, q
  real(kind=8) :: a, dtau, U
  character(len=512) :: tape_file, out_file, cb_file

  call get_command_argument(1, tape_file)
  call get_command_argument(2, out_file)
  cb_file = ''
  if (command_argument_count() >= 3) call get_command_argument(3, cb_file)
  dtau = 0.1d0
  U = 4.d0
  ! exp(-dtau T) for T = -t (S + S^T) on the ring: entries depend on the distance d only
  do j = 1, n
    do i = 1, n
      d = modulo(i - j, n)
      a = dtau
      expT(i, j)  = (exp(2*a) + 2*cos(acos(-1.d0)*d/2) + exp(-2*a)*(-1)**d) / 4.d0
      expTi(i, j) = (exp(-2*a) + 2*cos(acos(-1.d0)*d/2) + exp(2*a)*(-1)**d) / 4.d0
    enddo
  enddo
  do i = 1, n_op
    opP(i) = i
    optype(i) = 2
    opE(i) = 1.d0
    opg(i) = sqrt(cmplx(-dtau*U/2.d0, 0.d0, kind=8))
    opalpha(i) = cmplx(-0.5d0, 0.d0, kind=8)
  enddo
  do nt = 1, ltrot
    do i = 1, n_op
      f(i, nt) = merge(1, -1, modulo(i + nt, 2) == 0) * merge(2, 1, modulo(i * nt, 3) == 0)
    enddo
  enddo
  desc%ndim = n; desc%ltrot = ltrot; desc%nwrap = nwrap; desc%n_sun = 2; desc%n_fl = 1
  desc%n_op = n_op; desc%t_is_real = 1
  desc%expT = c_loc(expT); desc%expT_inv = c_loc(expTi)
  desc%op_P = c_loc(opP); desc%op_type = c_loc(optype); desc%op_E = c_loc(opE)
  desc%op_g = c_loc(opg); desc%op_alpha = c_loc(opalpha)

  call check(alfq_ctx_create(desc, ctx))
  call check(alfq_fields_set(ctx, f, int(n_op, c_int32_t), int(ltrot, c_int32_t)))
  open(newunit=unit_in, file=trim(tape_file), status='old', action='read')
  read(unit_in, *) ntape
  allocate(tape(ntape))
  read(unit_in, *) tape
  close(unit_in)
  call check(alfq_rng_set_tape(ctx, tape, int(ntape, c_int64_t)))
  call check(alfq_init(ctx))
  call check(alfq_green_get(ctx, GR, int(n, c_int32_t)))
  call check(alfq_phase_get(ctx, phase))
  ! the first slices the way main.F90 calls WRAPGRUP (G on the host between calls)
  call check(alfq_udv_reset(ctx, ALFQ_UDVR, 'R'))
  ! main.F90:729 passes Propose_S0, Nt_sequential_start/end and N_Global_tau; the engine takes
  ! the full sequential pass without global moves and refuses anything else before any work
  if (alfq_WRAPGRUP_shim(ctx, GR, 0, phase, .false., 1, n_op, 1) /= 6) stop 'N_Global_tau > 0 not refused'
  if (alfq_WRAPGRUP_shim(ctx, GR, 0, phase, .false., 2, n_op, 0) /= 6) stop 'partial sequential pass not refused'
  do nt = 0, nwrap - 1
    call check(alfq_WRAPGRUP_shim(ctx, GR, nt, phase, .false., 1, n_op, 0))
  enddo
  ! the rest of this sweep is not run: reset to tau = 0 from a fresh init, then whole sweeps
  call check(alfq_ctx_destroy(ctx))
  call check(alfq_ctx_create(desc, ctx))
  call check(alfq_fields_set(ctx, f, int(n_op, c_int32_t), int(ltrot, c_int32_t)))
  call check(alfq_rng_set_tape(ctx, tape, int(ntape, c_int64_t)))
  call check(alfq_init(ctx))
  call check(alfq_sweep(ctx, 2_c_int32_t))
  call check(alfq_fields_get(ctx, f, int(n_op, c_int32_t), int(ltrot, c_int32_t)))
  call check(alfq_green_get(ctx, GR, int(n, c_int32_t)))
  call check(alfq_phase_get(ctx, phase))
  call check(alfq_control_get(ctx, ctl))
  call check(alfq_ctx_destroy(ctx))

  open(newunit=unit_out, file=trim(out_file), status='replace', action='write')
  do nt = 1, ltrot
    do i = 1, n_op
      write(unit_out, '(I3)') f(i, nt)
    enddo
  enddo
  do j = 1, n
    do i = 1, n
      write(unit_out, '(2ES26.17)') real(GR(i, j)), aimag(GR(i, j))
    enddo
  enddo
  write(unit_out, '(2ES26.17)') real(phase), aimag(phase)
  write(unit_out, '(2I12)') ctl%nc_up, ctl%acc_up
  close(unit_out)
  print '(A,ES10.2,A,I8)', 'alfq_ring_demo: Precision Green max ', ctl%xmax_g, '  accepted ', ctl%acc_up
  if (len_trim(cb_file) > 0) call checkerboard_tau_run()

contains

  ! The ring with the checkerboard Op_T and TAU_M: bonds b = 1..4 join sites bp(:, b)
  ! with amplitude tb(b); family 1 holds (1,2) and (3,4), family 2 holds (2,3) and (4,1).
  ! H_T = c^+ O c with O = -t_b on the bonds, so exp(-dtau O_b) = [[ch, sh], [sh, ch]] of
  ! dtau t_b.
  subroutine checkerboard_tau_run()
    complex(c_double_complex) :: P(n, n), Pi(n, n), Eb(n, n)
    real(kind=8), parameter :: tb(n) = [1.0d0, 0.8d0, 1.2d0, 0.9d0]
    real(kind=8) :: ch, sh
    famlen = [2, 2]
    bp(:, 1) = [1, 2]; bp(:, 2) = [3, 4]; bp(:, 3) = [2, 3]; bp(:, 4) = [4, 1]
    do b = 1, n
      ch = cosh(dtau*tb(b)); sh = sinh(dtau*tb(b))
      expo(:, :, b) = reshape([ch, sh, sh, ch], [2, 2])
      expoi(:, :, b) = reshape([ch, -sh, -sh, ch], [2, 2])
    enddo
    ! the descriptor's dense exp(-dtau T) is P = E_1 E_2 E_3 E_4 and its inverse
    P = 0; Pi = 0
    do i = 1, n
      P(i, i) = 1; Pi(i, i) = 1
    enddo
    do b = 1, n
      call embed(expo(:, :, b), bp(:, b), Eb)
      P = matmul(P, Eb)
      call embed(expoi(:, :, n + 1 - b), bp(:, n + 1 - b), Eb)
      Pi = matmul(Pi, Eb)
    enddo
    expT = real(P); expTi = real(Pi)
    cbd%nbond = n; cbd%nfam = 2
    cbd%fam_len = c_loc(famlen); cbd%p = c_loc(bp); cbd%expo = c_loc(expo); cbd%expo_inv = c_loc(expoi)
    ! one orbital per cell; Latt%imj(I, J) = cell of R_I - R_J on the ring
    do i = 1, n
      invlist(i) = i
      do j = 1, n
        imj(i, j) = modulo(i - j, n) + 1
      enddo
      orow(2*i - 1) = bp(1, i); ocol(2*i - 1) = bp(2, i); oval(2*i - 1) = cmplx(-tb(i), 0.d0, kind=8)
      orow(2*i) = bp(2, i); ocol(2*i) = bp(1, i); oval(2*i) = cmplx(-tb(i), 0.d0, kind=8)
    enddo
    od%ncell = n; od%norb = 1; od%nnz = 2*n; od%lobs_st = 1; od%lobs_en = ltrot
    od%invlist = c_loc(invlist); od%imj = c_loc(imj); od%o_row = c_loc(orow); od%o_col = c_loc(ocol)
    od%o_val = c_loc(oval); od%ham_u = U

    call check(alfq_ctx_create(desc, ctx))
    call check(alfq_set_checkerboard(ctx, cbd))
    call check(alfq_obs_setup(ctx, od))
    call check(alfq_tau_enable(ctx, 1_c_int32_t))
    do nt = 1, ltrot
      do i = 1, n_op
        f(i, nt) = merge(1, -1, modulo(i + nt, 2) == 0) * merge(2, 1, modulo(i * nt, 3) == 0)
      enddo
    enddo
    call check(alfq_fields_set(ctx, f, int(n_op, c_int32_t), int(ltrot, c_int32_t)))
    call check(alfq_rng_set_tape(ctx, tape, int(ntape, c_int64_t)))
    call check(alfq_init(ctx))
    call check(alfq_sweep(ctx, 2_c_int32_t))
    call check(alfq_fields_get(ctx, f, int(n_op, c_int32_t), int(ltrot, c_int32_t)))
    call check(alfq_green_get(ctx, GR, int(n, c_int32_t)))
    call check(alfq_phase_get(ctx, phase))
    call check(alfq_control_get(ctx, ctl))
    call check(alfq_tau_get(ctx, ts, c_loc(tlatt), c_loc(tden0)))
    call check(alfq_ctx_destroy(ctx))

    open(newunit=unit_out, file=trim(cb_file), status='replace', action='write')
    do nt = 1, ltrot
      do i = 1, n_op
        write(unit_out, '(I3)') f(i, nt)
      enddo
    enddo
    do j = 1, n
      do i = 1, n
        write(unit_out, '(2ES26.17)') real(GR(i, j)), aimag(GR(i, j))
      enddo
    enddo
    write(unit_out, '(2ES26.17)') real(phase), aimag(phase)
    write(unit_out, '(2I12)') ctl%nc_up, ctl%acc_up
    write(unit_out, '(2I12)') ts%n, ts%ncg_tau
    write(unit_out, '(2ES26.17)') ts%sign, ts%xmax_tau
    do q = 1, 3
      do nt = 0, ltrot
        do d = 1, n
          write(unit_out, '(2ES26.17)') tlatt(1, d, nt, q), tlatt(2, d, nt, q)
        enddo
      enddo
    enddo
    write(unit_out, '(2ES26.17)') tden0(1), tden0(2)
    close(unit_out)
    print '(A,I4,A,ES10.2)', 'alfq_ring_demo: checkerboard TAU_M calls ', ts%n, '  Precision tau max ', ts%xmax_tau
  end subroutine

  ! the 2 x 2 bond matrix e on sites s(1), s(2) of the n x n identity
  subroutine embed(e, s, M)
    complex(c_double_complex), intent(in) :: e(2, 2)
    integer(c_int32_t), intent(in) :: s(2)
    complex(c_double_complex), intent(out) :: M(n, n)
    integer :: k
    M = 0
    do k = 1, n
      M(k, k) = 1
    enddo
    M(s(1), s(1)) = e(1, 1); M(s(2), s(1)) = e(2, 1)
    M(s(1), s(2)) = e(1, 2); M(s(2), s(2)) = e(2, 2)
  end subroutine

  subroutine check(status)
    integer(c_int), intent(in) :: status
    if (status /= 0) then
      write(*, '(A,I3,2A)') 'alfq error ', status, ': ', alfq_last_error_string()
      error stop 1
    endif
  end subroutine

end program alfq_ring_demo
