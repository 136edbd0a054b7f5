! Fortran binding of the C ABI in fdfk.h (ISO_C_BINDING), for drivers that
! keep the simulation resident on the GPU instead of calling the drop-in
! updatedispl_ every iteration.  The derived types have the C structs' layout.
module fdfk
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FDFK_OK = 0, FDFK_EINVAL = -22, FDFK_ENOMEM = -12, &
                                       FDFK_EDEVICE = -5, FDFK_ERANGE = -34, FDFK_ENOSYS = -38
  integer(c_int32_t), parameter, public :: FDFK_FP32 = 0, FDFK_BF16 = 1, FDFK_FP16 = 2, FDFK_FP64 = 3
  integer(c_int32_t), parameter, public :: FDFK_PSV = 0, FDFK_SH = 1

  ! one simulation: the scalars of get_parameter plus the arrays of read_data
  type, bind(C), public :: fdfk_job
    integer(c_int32_t) :: nx, nz, npml, norder   ! norder = FD order / 2
    integer(c_int32_t) :: is_pml_top             ! 1: absorbing top layer, rows from izt = 1-npml
    integer(c_int32_t) :: npower                 ! PML profile power (2)
    real(c_float)      :: x0, z0, dx, dz
    real(c_double)     :: dt
    real(c_double)     :: R0                     ! PML reflection coefficient
    real(c_float)      :: vpml                   ! <= 0: max Vp of c
    integer(c_int32_t) :: nt                     ! iterations it = 0..nt
    integer(c_int32_t) :: nstp                   ! record every nstp iterations
    integer(c_int32_t) :: nrcv
    type(c_ptr)        :: rx, rz                 ! receiver coordinates (m), real(c_float)
    type(c_ptr)        :: c                      ! rho, lambda, mu (reference layout), real(c_float)
    integer(c_int32_t) :: precision              ! FDFK_FP32, FDFK_FP16 (materials) or FDFK_FP64
    integer(c_int32_t) :: wave                   ! FDFK_PSV (in-plane) or FDFK_SH (transverse)
  end type

  ! FK background: the inputs of comput_background_wavefield
  type, bind(C), public :: fdfk_fk
    integer(c_int32_t) :: nlayer_left, nlayer_right
    type(c_ptr)        :: fk_left, fk_right      ! (4, nlayer): ztop, vp, vs, rho, real(c_float)
    integer(c_int32_t) :: src_is_p               ! P (1) or SV (0) incidence
    real(c_float)      :: ds, f0                 ! source strength (m) and frequency (Hz)
    real(c_double)     :: rayp                   ! s/m
    type(c_ptr)        :: bazs                   ! back-azimuth per padded column (rad), real(c_double)
    integer(c_int32_t) :: block_steps            ! 0: store every step; B > 0: blocks of B steps
  end type

  public :: fdfk_create, fdfk_compute_background, fdfk_run, fdfk_seismograms, fdfk_destroy, &
            fdfk_iteration, fdfk_sync, fdfk_last_error, fdfk_error_message

  interface
    integer(c_int) function fdfk_create(job, h) bind(C, name="fdfk_create")
      import :: c_int, c_ptr, fdfk_job
      type(fdfk_job), intent(in) :: job
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function fdfk_compute_background(h, fk) bind(C, name="fdfk_compute_background")
      import :: c_int, c_ptr, fdfk_fk
      type(c_ptr), value :: h
      type(fdfk_fk), intent(in) :: fk
    end function
    integer(c_int) function fdfk_run(h, nsteps) bind(C, name="fdfk_run")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
    end function
    ! seismograms (nt/nstp + 1 samples per receiver, time fastest)
    integer(c_int) function fdfk_seismograms(h, seisx, seisz) bind(C, name="fdfk_seismograms")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: seisx(*), seisz(*)
    end function
    integer(c_int) function fdfk_iteration(h) bind(C, name="fdfk_iteration")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fdfk_sync(h) bind(C, name="fdfk_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    subroutine fdfk_destroy(h) bind(C, name="fdfk_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    type(c_ptr) function fdfk_last_error() bind(C, name="fdfk_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! the last error of this thread as a Fortran string
  function fdfk_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(fdfk_last_error(), p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module fdfk
