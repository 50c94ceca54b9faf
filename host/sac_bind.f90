!=============================================================================
! sac_bind.f90 -- ISO_C_BINDING interfaces to the C ABI of libsac_amd.so
! (include/sac_amd.h).  Every function returns 0 on success, a positive
! SAC_E_* code otherwise; sac_error_message() returns the library's text.
!=============================================================================
module sac_bind
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SAC_BC_CONT = 1, SAC_BC_FIXED = 2, SAC_BC_PERIODIC = 3
  integer(c_int), parameter :: SAC_BC_SYMM = 4, SAC_BC_SYMM0 = 5, SAC_BC_ASYMM = 6
  integer(c_int), parameter :: SAC_BC_CONT1 = 7, SAC_BC_FIXED1 = 8, SAC_BC_GRAD1 = 9
  integer(c_int), parameter :: SAC_BC_CONTCD4 = 10
  integer(c_int), parameter :: SAC_BC_ZERO = 11
  integer(c_int), parameter :: SAC_ADV_ONESTEP = 1, SAC_ADV_TWOSTEP = 2, SAC_ADV_THREESTEP = 3, &
                                SAC_ADV_FOURSTEP = 4, SAC_ADV_STERCK = 5, SAC_ADV_JAMESON = 6
  integer(c_int), parameter :: SAC_PREC_FP64 = 0

  interface
    integer(c_int) function sac_dev_init(ndim, ndir, nx, nw, iws, niws, eqpar, neqpar, &
         grav0_, nu_, dx, typeB, fullgridini, typeadvance, precision, handle) &
         bind(C, name='sac_dev_init')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ndim, ndir, nw, niws, neqpar, grav0_, nu_
      integer(c_int), value :: fullgridini, typeadvance, precision
      integer(c_int), intent(in) :: nx(*), iws(*), typeB(*)
      real(c_double), intent(in) :: eqpar(*), dx(*)
      type(c_ptr), intent(out) :: handle
    end function sac_dev_init

    integer(c_int) function sac_dev_free(handle) bind(C, name='sac_dev_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function sac_dev_free

    integer(c_int) function sac_dev_upload(handle, w) bind(C, name='sac_dev_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: w(*)
    end function sac_dev_upload

    integer(c_int) function sac_dev_download(handle, w) bind(C, name='sac_dev_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: w(*)
    end function sac_dev_download

    integer(c_int) function sac_dev_save_begin(handle) bind(C, name='sac_dev_save_begin')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function sac_dev_save_begin

    integer(c_int) function sac_dev_save_ready(handle, ready) bind(C, name='sac_dev_save_ready')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), intent(out) :: ready
    end function sac_dev_save_ready

    integer(c_int) function sac_dev_save_end(handle, w) bind(C, name='sac_dev_save_end')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: w(*)
    end function sac_dev_save_end

    integer(c_int) function sac_dev_getboundary(handle) bind(C, name='sac_dev_getboundary')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function sac_dev_getboundary

    integer(c_int) function sac_dev_getdt_courant(handle, courantpar, dt, courantmax) &
         bind(C, name='sac_dev_getdt_courant')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), value :: courantpar
      real(c_double), intent(out) :: dt, courantmax
    end function sac_dev_getdt_courant

    integer(c_int) function sac_dev_getdt_special(handle, dt) bind(C, name='sac_dev_getdt_special')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(inout) :: dt
    end function sac_dev_getdt_special

    integer(c_int) function sac_dev_advance(handle, dt) bind(C, name='sac_dev_advance')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), value :: dt
    end function sac_dev_advance

    integer(c_int) function sac_dev_set_divbfix(handle, on) bind(C, name='sac_dev_set_divbfix')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: on
    end function sac_dev_set_divbfix

    integer(c_int) function sac_dev_set_resistive(handle, compactres, dtdiffpar) &
         bind(C, name='sac_dev_set_resistive')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: compactres
      real(c_double), value :: dtdiffpar
    end function sac_dev_set_resistive

    type(c_ptr) function sac_dev_last_error() bind(C, name='sac_dev_last_error')
      import :: c_ptr
    end function sac_dev_last_error
  end interface

contains

  ! text of the library's last error as a Fortran string
  function sac_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(sac_dev_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function sac_error_message

end module sac_bind
