! ------------------------------------------------------------
! ISO_C_BINDING interface to libhugks.so (include/hugks.h).
!
! The derived types mirror hugks_desc and hugks_step_info component for
! component (bind(c) gives them the C layout); the interfaces bind the
! C entry points by name.  A Fortran main program that used to call
! timestep/interpolation/evolution/update (2d/src/main.f90:78-81,
! 1d/src/main.f90:86-89) calls hugks_step or hugks_advance_until instead.
! ------------------------------------------------------------
module hugks_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: HUGKS_ABI_VERSION = 1

  integer(c_int), parameter :: HUGKS_OK = 0, HUGKS_E_ARG = -1, HUGKS_E_NOMEM = -2, &
       HUGKS_E_DEVICE = -3, HUGKS_E_UNSUPPORTED = -4, HUGKS_E_NUMERIC = -5, &
       HUGKS_E_WALL = -6, HUGKS_E_NOHIP = -7

  integer(c_int32_t), parameter :: HUGKS_FP64 = 0, HUGKS_MIXED_FP32 = 1, HUGKS_MIXED_BF16X3 = 2
  integer(c_int32_t), parameter :: HUGKS_BC_PERIODIC = 0, HUGKS_BC_SPECULAR = 1, &
       HUGKS_BC_DIFFUSE = 2
  integer(c_int32_t), parameter :: HUGKS_TAU_MU_OVER_RHO = 0, HUGKS_TAU_VHS = 1
  integer(c_int32_t), parameter :: HUGKS_DT_2D_GENERAL = 0, HUGKS_DT_2D_WALL_NORMAL = 1, &
       HUGKS_DT_1D = 2
  integer(c_int32_t), parameter :: HUGKS_FORCE_HERMITE_SYMMETRIC = 0, &
       HUGKS_FORCE_HERMITE_SYMMETRIC_B = 1, HUGKS_FORCE_HERMITE_LEGACY = 2, &
       HUGKS_FORCE_FINITE_DIFFERENCE = 3
  integer(c_int32_t), parameter :: HUGKS_STOP_NONE = 0, HUGKS_STOP_MAXTIME = 1, &
       HUGKS_STOP_RESIDUAL = 2, HUGKS_STOP_TIME = 3, HUGKS_STOP_STEADY = 4, HUGKS_STOP_ERROR = 5
  integer(c_int32_t), parameter :: HUGKS_FLUX_FUSED = 0, HUGKS_FLUX_SPLIT = 1, HUGKS_FLUX_STENCIL = 2, &
       HUGKS_FLUX_CELLSUM = 3, HUGKS_UPDATE_LDS = 0, HUGKS_UPDATE_MFMA = 1, &
       HUGKS_UPDATE_MFMA_NODE = 2, HUGKS_UPDATE_MFMA_FACE = 3, HUGKS_BATCH_AUTOTUNE = 1

  type, bind(c) :: hugks_desc
    integer(c_int32_t) :: abi_version = HUGKS_ABI_VERSION
    integer(c_int32_t) :: dim = 2
    integer(c_int32_t) :: nx = 0, ny = 0
    integer(c_int32_t) :: nv = 0
    integer(c_int32_t) :: nm = 0
    integer(c_int32_t) :: precision = HUGKS_FP64
    integer(c_int32_t) :: tau_model = HUGKS_TAU_VHS
    integer(c_int32_t) :: dt_rule = HUGKS_DT_2D_GENERAL
    integer(c_int32_t) :: clamp_dt = 0
    integer(c_int32_t) :: stop_on_residual = 0
    integer(c_int32_t) :: bc(4) = 0
    integer(c_int32_t) :: force_variant = HUGKS_FORCE_HERMITE_SYMMETRIC
    integer(c_int32_t) :: steady_steps = 0
    integer(c_int32_t) :: reserved_i(1) = 0
    real(c_double) :: dx = 0, dy = 0
    real(c_double) :: pi = 0, smv = 0, innerk = 0, gamma = 0, prandtl = 0
    real(c_double) :: shakhov_coef = 0, muref = 0, omega = 0
    real(c_double) :: cfl = 0, maxtime = 0, eps_res = 0
    real(c_double) :: umax_dt = 0, vmax_dt = 0, force_rate = 0
    real(c_double) :: tw(4) = 1
    real(c_double) :: tref = 1, phi = 0
    real(c_double) :: steady_tol = 0, mass_tol = 0
    real(c_double) :: fd_span = 0
    real(c_double) :: reserved_d(1) = 0
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, weight = c_null_ptr, prim0 = c_null_ptr
    type(c_ptr) :: accel_cell = c_null_ptr, accel_vface = c_null_ptr, accel_hface = c_null_ptr
  end type hugks_desc

  type, bind(c) :: hugks_step_info
    integer(c_int32_t) :: steps, iter, stop_reason, device_flags
    real(c_double) :: simtime, dt
    real(c_double) :: res(4)
    real(c_double) :: steady_res(3)
    real(c_double) :: mass_rel_change
  end type hugks_step_info

  interface
    integer(c_int) function hugks_create(desc, handle) bind(c, name='hugks_create')
      import :: c_int, c_ptr, hugks_desc
      type(hugks_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: handle
    end function hugks_create

    integer(c_int) function hugks_destroy(handle) bind(c, name='hugks_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function hugks_destroy

    type(c_ptr) function hugks_last_error_c(handle) bind(c, name='hugks_last_error')
      import :: c_ptr
      type(c_ptr), value :: handle
    end function hugks_last_error_c

    integer(c_int) function hugks_step(handle, info) bind(c, name='hugks_step')
      import :: c_int, c_ptr, hugks_step_info
      type(c_ptr), value :: handle
      type(hugks_step_info), intent(out) :: info
    end function hugks_step

    integer(c_int) function hugks_advance(handle, max_steps, info) bind(c, name='hugks_advance')
      import :: c_int, c_int32_t, c_ptr, hugks_step_info
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: max_steps
      type(hugks_step_info), intent(out) :: info
    end function hugks_advance

    integer(c_int) function hugks_advance_until(handle, max_steps, t_stop, info) &
         bind(c, name='hugks_advance_until')
      import :: c_int, c_int32_t, c_double, c_ptr, hugks_step_info
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: max_steps
      real(c_double), value :: t_stop
      type(hugks_step_info), intent(out) :: info
    end function hugks_advance_until

    integer(c_int) function hugks_get_macro(handle, w, q) bind(c, name='hugks_get_macro')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      type(c_ptr), value :: w, q
    end function hugks_get_macro

    ! macro*.dat files written by a writer thread of the library, off the
    ! step loop; path and header are NUL-terminated
    integer(c_int) function hugks_write_blocks_async(handle, path, header, per_line, nblock, &
         block_len, values) bind(c, name='hugks_write_blocks_async')
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), dimension(*), intent(in) :: path, header
      integer(c_int32_t), value :: per_line, nblock
      type(c_ptr), value :: block_len, values
    end function hugks_write_blocks_async

    integer(c_int) function hugks_write_wait(handle) bind(c, name='hugks_write_wait')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function hugks_write_wait

    integer(c_int) function hugks_get_f(handle, hf, bf) bind(c, name='hugks_get_f')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      type(c_ptr), value :: hf, bf
    end function hugks_get_f

    integer(c_int) function hugks_set_f(handle, hf, bf, w) bind(c, name='hugks_set_f')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      type(c_ptr), value :: hf, bf, w
    end function hugks_set_f

    integer(c_int) function hugks_get_time(handle, simtime, iter) bind(c, name='hugks_get_time')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: simtime
      integer(c_int32_t), intent(out) :: iter
    end function hugks_get_time

    integer(c_int) function hugks_checkpoint_write_c(handle, path) &
         bind(c, name='hugks_checkpoint_write')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: path(*)
    end function hugks_checkpoint_write_c

    integer(c_int) function hugks_checkpoint_read_c(handle, path) &
         bind(c, name='hugks_checkpoint_read')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: path(*)
    end function hugks_checkpoint_read_c

    type(c_ptr) function hugks_device_info_c() bind(c, name='hugks_device_info')
      import :: c_ptr
    end function hugks_device_info_c

    integer(c_int) function hugks_set_schedule(handle, flux_kernel, update_kernel) &
         bind(c, name='hugks_set_schedule')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: flux_kernel, update_kernel
    end function hugks_set_schedule

    integer(c_int) function hugks_tune(desc, flux_kernel, update_kernel, ms_per_step) &
         bind(c, name='hugks_tune')
      import :: c_int, c_int32_t, c_double, hugks_desc
      type(hugks_desc), intent(in) :: desc
      integer(c_int32_t), intent(out) :: flux_kernel, update_kernel
      real(c_double), intent(out) :: ms_per_step
    end function hugks_tune

    integer(c_int) function hugks_batch_create(n, descs, flags, batch) &
         bind(c, name='hugks_batch_create')
      import :: c_int, c_int32_t, c_ptr, hugks_desc
      integer(c_int32_t), value :: n
      type(hugks_desc), intent(in) :: descs(*)
      integer(c_int32_t), value :: flags
      type(c_ptr), intent(out) :: batch
    end function hugks_batch_create

    integer(c_int) function hugks_batch_destroy(batch) bind(c, name='hugks_batch_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function hugks_batch_destroy

    integer(c_int32_t) function hugks_batch_grouped(batch) bind(c, name='hugks_batch_grouped')
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
    end function hugks_batch_grouped

    integer(c_int) function hugks_batch_handle(batch, i, handle) bind(c, name='hugks_batch_handle')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: i
      type(c_ptr), intent(out) :: handle
    end function hugks_batch_handle

    integer(c_int) function hugks_batch_advance(batch, max_steps, infos) &
         bind(c, name='hugks_batch_advance')
      import :: c_int, c_int32_t, c_ptr, hugks_step_info
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: max_steps
      type(hugks_step_info), intent(out) :: infos(*)
    end function hugks_batch_advance
  end interface

contains

  ! Message of the last failing call on handle (c_null_ptr: last create).
  function hugks_last_error(handle) result(msg)
    type(c_ptr), intent(in) :: handle
    character(len=:), allocatable :: msg
    msg = c_string(hugks_last_error_c(handle))
  end function hugks_last_error

  ! Whole solver state to / from a file (C strings built here).
  integer(c_int) function hugks_checkpoint_write(handle, path)
    type(c_ptr), intent(in) :: handle
    character(len=*), intent(in) :: path
    hugks_checkpoint_write = hugks_checkpoint_write_c(handle, trim(path) // c_null_char)
  end function hugks_checkpoint_write

  integer(c_int) function hugks_checkpoint_read(handle, path)
    type(c_ptr), intent(in) :: handle
    character(len=*), intent(in) :: path
    hugks_checkpoint_read = hugks_checkpoint_read_c(handle, trim(path) // c_null_char)
  end function hugks_checkpoint_read

  function hugks_device_info() result(msg)
    character(len=:), allocatable :: msg
    msg = c_string(hugks_device_info_c())
  end function hugks_device_info

  ! Stops with the library's message on a nonzero status, as the
  ! reference stops on its own error conditions (error stop).
  subroutine hugks_check(rc, handle, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: handle
    character(len=*), intent(in) :: what
    if (rc /= HUGKS_OK) then
      write(*, '(A,A,I0,A,A)') trim(what), ': hugks error ', rc, ': ', hugks_last_error(handle)
      error stop 1
    end if
  end subroutine hugks_check

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(c, name='strlen')
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
      end function c_strlen
    end interface
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function c_string

end module hugks_capi
