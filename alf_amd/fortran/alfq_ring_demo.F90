!  Fortran host driving the gfx950 sweep engine through alfq_mod: the Hubbard
!  model on a 4-site ring (t = 1, U = 4, beta = 2, dtau = 0.1, Nwrap = 5, N_SUN = 2,
!  density-channel four-state HS field), with exp(-dtau T) in closed form.
!
!  It follows the reference's sequential sweep at the granularity of the shims
!  for the first up pass slices (alfq_WRAPGRUP_shim with G on the host, as
!  main.F90 calls WRAPGRUP) and then runs whole sweeps on the device.
!
!  usage: alfq_ring_demo <tape file> <output file>
!    tape file: count, then that many uniforms (one per line)
!    output:    final fields, then G (re im per line, column-major), PHASE,
!               and the NC_up / ACC_up counters
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
  integer :: i, j, d, st, nt, ntape, unit_in, unit_out
  real(kind=8) :: a, dtau, U
  character(len=512) :: tape_file, out_file

  call get_command_argument(1, tape_file)
  call get_command_argument(2, out_file)
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
  do nt = 0, nwrap - 1
    call check(alfq_WRAPGRUP_shim(ctx, GR, nt, phase))
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

contains

  subroutine check(status)
    integer(c_int), intent(in) :: status
    if (status /= 0) then
      write(*, '(A,I3,2A)') 'alfq error ', status, ': ', alfq_last_error_string()
      error stop 1
    endif
  end subroutine

end program alfq_ring_demo
