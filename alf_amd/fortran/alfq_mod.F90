!  ISO_C_BINDING interface of the gfx950 DQMC sweep engine (include/alfq.h) and
!  drop-in shims with the argument lists of the reference routines they replace:
!
!    alfq_WRAPGRUP(GR, NTAU, PHASE[, Propose_S0, Nt_sequential_start, Nt_sequential_end,
!                  N_Global_tau])         <->  WRAPGRUP  (Prog/Wrapgr_mod.F90:81)
!    alfq_WRAPGRDO(GR, NTAU, PHASE[, ...]) <->  WRAPGRDO  (Prog/Wrapgr_mod.F90:160)
!    alfq_WRAPUR(NTAU, NTAU1)             <->  WRAPUR    (Prog/wrapur_mod.F90:37)
!    alfq_WRAPUL(NTAU1, NTAU)             <->  WRAPUL    (Prog/wrapul_mod.F90:36)
!    alfq_CGR(Z, NVAR, GRUP)              <->  CGR       (Prog/cgr1_mod.F90:36)
!
!  GR is the host copy of the Green's function (COMPLEX(8), Ndim x Ndim); the shims
!  upload it before and download it after the device work, so a caller that keeps
!  the reference's host-side control flow sees the same in/out semantics.  A
!  caller that does not need G on the host between slices uses alfq_sweep.
!  Non-zero status codes are the reference's runtime_error_mod values and are
!  passed to the caller's error handler (Terminate_on_error in ALF).
module alfq_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: ALFQ_UDVR = -1, ALFQ_UDVL = -2

  type, bind(C), public :: alfq_model_desc
    integer(c_int32_t) :: ndim, ltrot, nwrap, n_sun, n_fl, n_op, t_is_real
    type(c_ptr) :: expT, expT_inv, op_P, op_type, op_E, op_g, op_alpha
    ! Symm = .T.: mat_1D2 and invmat_1D2 of Op_T; c_null_ptr for Symm = .F.
    type(c_ptr) :: expT_half = c_null_ptr, expT_half_inv = c_null_ptr
  end type alfq_model_desc

  type, bind(C), public :: alfq_control
    integer(c_int64_t) :: nc_up, acc_up, ncg
    real(c_double) :: xmean_g_sum, xmax_g, xmax_p
  end type alfq_control

  ! equal-time measurements (ham%Obser of the Hubbard plugin on the device)
  type, bind(C), public :: alfq_obs_desc
    integer(c_int32_t) :: ncell, norb, nnz, lobs_st, lobs_en
    type(c_ptr) :: invlist, imj, o_row, o_col, o_val
    real(c_double) :: ham_u
  end type alfq_obs_desc

  ! checkerboard Op_T: 2 x 2 bond exponentials in families on disjoint sites
  type, bind(C), public :: alfq_checkerboard_desc
    integer(c_int32_t) :: nbond, nfam
    type(c_ptr) :: fam_len, p, expo, expo_inv
    ! Symm = .T.: the bonds' mat_1D2 and invmat_1D2; c_null_ptr for Symm = .F.
    type(c_ptr) :: expo_half = c_null_ptr, expo_half_inv = c_null_ptr
  end type alfq_checkerboard_desc

  ! time-displaced sums (Ltau = 1) and Control_Precision_tau
  type, bind(C), public :: alfq_tau_sums
    real(c_double) :: sign
    integer(c_int64_t) :: n, ncg_tau
    real(c_double) :: xmax_tau, xmean_tau_sum
  end type alfq_tau_sums

  type, bind(C), public :: alfq_obs_sums
    real(c_double) :: scal(8)
    real(c_double) :: sign_scal
    integer(c_int64_t) :: n_scal
    real(c_double) :: sign_eq
    integer(c_int64_t) :: n_eq
  end type alfq_obs_sums

  public :: alfq_ctx_create, alfq_ctx_destroy, alfq_set_delay, alfq_set_precision, alfq_fields_set, alfq_fields_get, &
            alfq_green_set, alfq_green_get, alfq_phase_set, alfq_phase_get, alfq_udv_reset, &
            alfq_udv_assign, alfq_rng_set_tape, alfq_rng_set_philox, alfq_wrapgrup, alfq_wrapgrdo, &
            alfq_wrapur, alfq_wrapul, alfq_cgr, alfq_init, alfq_sweep, alfq_sync, alfq_control_get, &
            alfq_last_error_string, alfq_obs_setup, alfq_obs_enable, alfq_obs_reset, alfq_obs_get, &
            alfq_set_checkerboard, alfq_tau_enable, alfq_tau_m, alfq_tau_get, alfq_check_wrap_args
  public :: alfq_WRAPGRUP_shim, alfq_WRAPGRDO_shim, alfq_WRAPUR_shim, alfq_WRAPUL_shim, alfq_CGR_shim

  interface
    integer(c_int) function alfq_ctx_create(desc, ctx) bind(C, name="alfq_ctx_create")
      import :: c_int, c_ptr, alfq_model_desc
      type(alfq_model_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function alfq_ctx_destroy(ctx) bind(C, name="alfq_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function alfq_set_delay(ctx, k) bind(C, name="alfq_set_delay")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k
    end function
    integer(c_int) function alfq_set_precision(ctx, mode) bind(C, name="alfq_set_precision")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function alfq_set_checkerboard(ctx, desc) bind(C, name="alfq_set_checkerboard")
      import :: c_int, c_ptr, alfq_checkerboard_desc
      type(c_ptr), value :: ctx
      type(alfq_checkerboard_desc), intent(in) :: desc
    end function
    integer(c_int) function alfq_tau_enable(ctx, on) bind(C, name="alfq_tau_enable")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function alfq_tau_m(ctx) bind(C, name="alfq_tau_m")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function alfq_tau_get(ctx, sums, latt, den0) bind(C, name="alfq_tau_get")
      import :: c_int, c_ptr, alfq_tau_sums
      type(c_ptr), value :: ctx
      type(alfq_tau_sums), intent(out) :: sums
      type(c_ptr), value :: latt, den0
    end function
    integer(c_int) function alfq_obs_setup(ctx, desc) bind(C, name="alfq_obs_setup")
      import :: c_int, c_ptr, alfq_obs_desc
      type(c_ptr), value :: ctx
      type(alfq_obs_desc), intent(in) :: desc
    end function
    integer(c_int) function alfq_obs_enable(ctx, on) bind(C, name="alfq_obs_enable")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function alfq_obs_reset(ctx) bind(C, name="alfq_obs_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function alfq_obs_get(ctx, sums, latt, den0) bind(C, name="alfq_obs_get")
      import :: c_int, c_ptr, c_double, alfq_obs_sums
      type(c_ptr), value :: ctx
      type(alfq_obs_sums), intent(out) :: sums
      real(c_double), intent(out) :: latt(*), den0(*)
    end function
    integer(c_int) function alfq_fields_set(ctx, f, n_op, ltrot) bind(C, name="alfq_fields_set")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: f(*)
      integer(c_int32_t), value :: n_op, ltrot
    end function
    integer(c_int) function alfq_fields_get(ctx, f, n_op, ltrot) bind(C, name="alfq_fields_get")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: f(*)
      integer(c_int32_t), value :: n_op, ltrot
    end function
    integer(c_int) function alfq_green_set(ctx, G, ld) bind(C, name="alfq_green_set")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: G(*)
      integer(c_int32_t), value :: ld
    end function
    integer(c_int) function alfq_green_get(ctx, G, ld) bind(C, name="alfq_green_get")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: G(*)
      integer(c_int32_t), value :: ld
    end function
    integer(c_int) function alfq_phase_set(ctx, z) bind(C, name="alfq_phase_set")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: z
    end function
    integer(c_int) function alfq_phase_get(ctx, z) bind(C, name="alfq_phase_get")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: z
    end function
    integer(c_int) function alfq_udv_reset(ctx, var, side) bind(C, name="alfq_udv_reset")
      import :: c_int, c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: var
      character(kind=c_char), value :: side
    end function
    integer(c_int) function alfq_udv_assign(ctx, dst, src) bind(C, name="alfq_udv_assign")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dst, src
    end function
    integer(c_int) function alfq_rng_set_tape(ctx, u, n) bind(C, name="alfq_rng_set_tape")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function alfq_rng_set_philox(ctx, seed, stream) bind(C, name="alfq_rng_set_philox")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed
      integer(c_int32_t), value :: stream
    end function
    integer(c_int) function alfq_wrapgrup(ctx, ntau, phase) bind(C, name="alfq_wrapgrup")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntau
      complex(c_double_complex), intent(inout) :: phase
    end function
    integer(c_int) function alfq_wrapgrdo(ctx, ntau, phase) bind(C, name="alfq_wrapgrdo")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntau
      complex(c_double_complex), intent(inout) :: phase
    end function
    integer(c_int) function alfq_check_wrap_args(ctx, propose_s0, nt_start, nt_end, n_global_tau) &
        bind(C, name="alfq_check_wrap_args")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: propose_s0, nt_start, nt_end, n_global_tau
    end function
    integer(c_int) function alfq_wrapur(ctx, ntau, ntau1) bind(C, name="alfq_wrapur")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntau, ntau1
    end function
    integer(c_int) function alfq_wrapul(ctx, ntau1, ntau) bind(C, name="alfq_wrapul")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntau1, ntau
    end function
    integer(c_int) function alfq_cgr(ctx, nvar, var_r, var_l, check, z) bind(C, name="alfq_cgr")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nvar, var_r, var_l, check
      complex(c_double_complex), intent(out) :: z
    end function
    integer(c_int) function alfq_init(ctx) bind(C, name="alfq_init")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function alfq_sweep(ctx, nsweep) bind(C, name="alfq_sweep")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsweep
    end function
    integer(c_int) function alfq_ctx_create_batch(descs, nchain, ctx) bind(C, name="alfq_ctx_create_batch")
      import :: c_int, c_ptr, c_int32_t, alfq_model_desc
      type(alfq_model_desc), intent(in) :: descs(*)
      integer(c_int32_t), value :: nchain
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function alfq_set_chain(ctx, chain) bind(C, name="alfq_set_chain")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: chain
    end function
    integer(c_int) function alfq_sweep_graphs(ctx, enable) bind(C, name="alfq_sweep_graphs")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function alfq_sync(ctx) bind(C, name="alfq_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function alfq_control_get(ctx, out) bind(C, name="alfq_control_get")
      import :: c_int, c_ptr, alfq_control
      type(c_ptr), value :: ctx
      type(alfq_control), intent(out) :: out
    end function
    type(c_ptr) function c_alfq_last_error() bind(C, name="alfq_last_error")
      import :: c_ptr
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! the library's last error message as a Fortran string
  function alfq_last_error_string() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = c_alfq_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    enddo
  end function

  ! The optional trailing arguments of WRAPGRUP/WRAPGRDO (Propose_S0, Nt_sequential_start,
  ! Nt_sequential_end, N_Global_tau, Prog/Wrapgr_mod.F90:81, 160), when the caller passes
  ! them, checked against what the engine does (alfq_check_wrap_args); all four or none
  integer function wrap_args_status(ctx, Propose_S0, Nt_sequential_start, Nt_sequential_end, &
                                    N_Global_tau) result(st)
    type(c_ptr), intent(in) :: ctx
    logical, intent(in), optional :: Propose_S0
    integer, intent(in), optional :: Nt_sequential_start, Nt_sequential_end, N_Global_tau
    integer(c_int32_t) :: ps0
    st = 0
    if (.not. (present(Propose_S0) .or. present(Nt_sequential_start) .or. present(Nt_sequential_end) &
               .or. present(N_Global_tau))) return
    if (.not. (present(Propose_S0) .and. present(Nt_sequential_start) .and. present(Nt_sequential_end) &
               .and. present(N_Global_tau))) then
      st = 1
      return
    endif
    ps0 = 0
    if (Propose_S0) ps0 = 1
    st = alfq_check_wrap_args(ctx, ps0, int(Nt_sequential_start, c_int32_t), int(Nt_sequential_end, c_int32_t), &
                              int(N_Global_tau, c_int32_t))
  end function

  ! WRAPGRUP(GR, NTAU, PHASE): G(NTAU) -> G(NTAU+1) with the updates of slice NTAU+1
  integer function alfq_WRAPGRUP_shim(ctx, GR, NTAU, PHASE, Propose_S0, Nt_sequential_start, &
                                      Nt_sequential_end, N_Global_tau) result(st)
    type(c_ptr), intent(in) :: ctx
    complex(kind=8), intent(inout) :: GR(:,:)
    integer, intent(in) :: NTAU
    complex(kind=8), intent(inout) :: PHASE
    logical, intent(in), optional :: Propose_S0
    integer, intent(in), optional :: Nt_sequential_start, Nt_sequential_end, N_Global_tau
    st = wrap_args_status(ctx, Propose_S0, Nt_sequential_start, Nt_sequential_end, N_Global_tau)
    if (st /= 0) return
    st = alfq_green_set(ctx, GR, int(size(GR, 1), c_int32_t))
    if (st /= 0) return
    st = alfq_wrapgrup(ctx, int(NTAU, c_int32_t), PHASE)
    if (st /= 0) return
    st = alfq_green_get(ctx, GR, int(size(GR, 1), c_int32_t))
  end function

  ! WRAPGRDO(GR, NTAU, PHASE): updates of slice NTAU, then G(NTAU) -> G(NTAU-1)
  integer function alfq_WRAPGRDO_shim(ctx, GR, NTAU, PHASE, Propose_S0, Nt_sequential_start, &
                                      Nt_sequential_end, N_Global_tau) result(st)
    type(c_ptr), intent(in) :: ctx
    complex(kind=8), intent(inout) :: GR(:,:)
    integer, intent(in) :: NTAU
    complex(kind=8), intent(inout) :: PHASE
    logical, intent(in), optional :: Propose_S0
    integer, intent(in), optional :: Nt_sequential_start, Nt_sequential_end, N_Global_tau
    st = wrap_args_status(ctx, Propose_S0, Nt_sequential_start, Nt_sequential_end, N_Global_tau)
    if (st /= 0) return
    st = alfq_green_set(ctx, GR, int(size(GR, 1), c_int32_t))
    if (st /= 0) return
    st = alfq_wrapgrdo(ctx, int(NTAU, c_int32_t), PHASE)
    if (st /= 0) return
    st = alfq_green_get(ctx, GR, int(size(GR, 1), c_int32_t))
  end function

  ! WRAPUR(NTAU, NTAU1, udvr) on the device-resident udvr
  integer function alfq_WRAPUR_shim(ctx, NTAU, NTAU1) result(st)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: NTAU, NTAU1
    st = alfq_wrapur(ctx, int(NTAU, c_int32_t), int(NTAU1, c_int32_t))
  end function

  ! WRAPUL(NTAU1, NTAU, udvl) on the device-resident udvl
  integer function alfq_WRAPUL_shim(ctx, NTAU1, NTAU) result(st)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: NTAU1, NTAU
    st = alfq_wrapul(ctx, int(NTAU1, c_int32_t), int(NTAU, c_int32_t))
  end function

  ! CGR(Z, NVAR, GRUP, udvr, udvl): fresh G from the device-resident udvr, udvl
  integer function alfq_CGR_shim(ctx, Z, NVAR, GRUP) result(st)
    type(c_ptr), intent(in) :: ctx
    complex(kind=8), intent(out) :: Z
    integer, intent(in) :: NVAR
    complex(kind=8), intent(inout) :: GRUP(:,:)
    st = alfq_cgr(ctx, int(NVAR, c_int32_t), ALFQ_UDVR, ALFQ_UDVL, 0_c_int32_t, Z)
    if (st /= 0) return
    st = alfq_green_get(ctx, GRUP, int(size(GRUP, 1), c_int32_t))
  end function

end module alfq_mod
