!> iso_c_binding layer over libhydft (include/hydft.h) for the Fortran host.
!>
!> hydft_hip_spectral provides hip_spec_t, a drop-in for the reference's
!> spec_t (src/core/spectral.f90:19-36) and fft_t (src/core/fft.f90:20-32):
!> same method names and argument meaning on host (nx, ny, nz) arrays, with
!> the transforms and operators executed by libhydft on the GPU.  A nonzero
!> status from the library is turned into fatal(), i.e. error stop 1, as the
!> reference does (src/core/kinds.f90:26-30).
module hydft_hip_c
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: HY_OK = 0, HY_PREC_FP64 = 0, HY_PREC_MIXED = 1

  interface
    integer(c_int) function hy_init(device) bind(C, name='hy_init')
      import :: c_int
      integer(c_int), value :: device
    end function
    type(c_ptr) function hy_last_error() bind(C, name='hy_last_error')
      import :: c_ptr
    end function
    integer(c_int) function hy_device_alloc(bytes, d) bind(C, name='hy_device_alloc')
      import :: c_int, c_long_long, c_ptr
      integer(c_long_long), value :: bytes
      type(c_ptr) :: d
    end function
    integer(c_int) function hy_device_free(d) bind(C, name='hy_device_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: d
    end function
    integer(c_int) function hy_copy_to_device(d, h, bytes) bind(C, name='hy_copy_to_device')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: d, h
      integer(c_long_long), value :: bytes
    end function
    integer(c_int) function hy_copy_to_host(h, d, bytes) bind(C, name='hy_copy_to_host')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: h, d
      integer(c_long_long), value :: bytes
    end function
    integer(c_int) function hy_grid_create(nx, ny, nz, lx, ly, lz, dealias, precision, g) &
        bind(C, name='hy_grid_create')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, dealias, precision
      real(c_double), value :: lx, ly, lz
      type(c_ptr) :: g
    end function
    integer(c_int) function hy_grid_destroy(g) bind(C, name='hy_grid_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function
    integer(c_int) function hy_grid_query(g, nkx, ndim, kept3, volume, kmax) bind(C, name='hy_grid_query')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: g
      integer(c_int) :: nkx, ndim, kept3(3)
      real(c_double) :: volume, kmax
    end function
    integer(c_int) function hy_fft_r2c(g, du, duhat, stream) bind(C, name='hy_fft_r2c')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, du, duhat, stream
    end function
    integer(c_int) function hy_fft_c2r(g, duhat, du, stream) bind(C, name='hy_fft_c2r')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, duhat, du, stream
    end function
    integer(c_int) function hy_spec_grad_hat(g, duhat, idir, ddu, stream) bind(C, name='hy_spec_grad_hat')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, duhat, ddu, stream
      integer(c_int), value :: idir
    end function
    integer(c_int) function hy_spec_grad(g, du, ddu3, stream) bind(C, name='hy_spec_grad')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, du, ddu3, stream
    end function
    integer(c_int) function hy_spec_lap(g, du, dlu, stream) bind(C, name='hy_spec_lap')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, du, dlu, stream
    end function
    integer(c_int) function hy_spec_div(g, dv3, ddv, stream) bind(C, name='hy_spec_div')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, dv3, ddv, stream
    end function
    integer(c_int) function hy_spec_convolve(g, dkern, duhat, dcu, stream) bind(C, name='hy_spec_convolve')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, dkern, duhat, dcu, stream
    end function
    integer(c_int) function hy_spec_dealias(g, du, stream) bind(C, name='hy_spec_dealias')
      import :: c_int, c_ptr
      type(c_ptr), value :: g, du, stream
    end function
    integer(c_int) function hy_spec_mean(g, du, mean, stream) bind(C, name='hy_spec_mean')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: g, du, stream
      real(c_double) :: mean
    end function
    integer(c_int) function hy_spec_mode(g, du, mx, my, mz, reim, stream) bind(C, name='hy_spec_mode')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: g, du, stream
      integer(c_int), value :: mx, my, mz
      real(c_double) :: reim(2)
    end function
  end interface

contains

  !> Message of the last failed call.
  function hy_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(C, name='strlen')
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
      end function
    end interface
    p = hy_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function hy_error_message

  !> fatal() on a nonzero status (kinds.f90:26-30 semantics).
  subroutine hy_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= HY_OK) then
      write(0, '(a)') 'hydft: fatal: '//where//': '//hy_error_message()
      error stop 1
    end if
  end subroutine hy_check

end module hydft_hip_c


module hydft_hip_spectral
  use, intrinsic :: iso_c_binding
  use hydft_hip_c
  implicit none
  private
  public :: hip_spec_t

  integer, parameter :: rp = c_double

  !> Drop-in for spec_t: grid + GPU transforms and operators on host arrays.
  type :: hip_spec_t
    type(c_ptr) :: g = c_null_ptr
    integer :: nx = 0, ny = 0, nz = 0, nkx = 0, ndim = 0
    real(rp) :: volume = 0.0_rp, kmax_dealiased = 0.0_rp
    type(c_ptr) :: dr(4) = c_null_ptr     !< device real fields (1 and 3-vector)
    type(c_ptr) :: dc = c_null_ptr        !< device half spectrum
    type(c_ptr) :: dk = c_null_ptr        !< device kernel table
  contains
    procedure :: init => spec_init
    procedure :: destroy => spec_destroy
    procedure :: fwd => spec_fwd
    procedure :: bwd => spec_bwd
    procedure :: grad_hat => spec_grad_hat
    procedure :: grad => spec_grad
    procedure :: lap => spec_lap
    procedure :: div => spec_div
    procedure :: convolve => spec_convolve
    procedure :: dealias => spec_dealias
    procedure :: mean => spec_mean
    procedure :: mode => spec_mode
  end type hip_spec_t

contains

  subroutine spec_init(s, nx, ny, nz, lx, ly, lz, dealias)
    class(hip_spec_t), intent(inout) :: s
    integer, intent(in) :: nx, ny, nz
    real(rp), intent(in) :: lx, ly, lz
    logical, intent(in), optional :: dealias
    integer(c_int) :: nkx, ndim, kept(3), dl
    integer(c_long_long) :: nr, nc
    dl = 1
    if (present(dealias)) then
      if (.not. dealias) dl = 0
    end if
    call s%destroy()
    call hy_check(hy_grid_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), lx, ly, lz, dl, &
                                 HY_PREC_FP64, s%g), 'grid_init')
    call hy_check(hy_grid_query(s%g, nkx, ndim, kept, s%volume, s%kmax_dealiased), 'grid_query')
    s%nx = nx; s%ny = ny; s%nz = nz; s%nkx = nkx; s%ndim = ndim
    nr = int(nx, c_long_long)*ny*nz*8
    nc = int(nkx, c_long_long)*ny*nz*16
    call hy_check(hy_device_alloc(nr, s%dr(1)), 'alloc')
    call hy_check(hy_device_alloc(3*nr, s%dr(2)), 'alloc')
    call hy_check(hy_device_alloc(nc, s%dc), 'alloc')
    call hy_check(hy_device_alloc(nc/2, s%dk), 'alloc')
  end subroutine spec_init

  subroutine spec_destroy(s)
    class(hip_spec_t), intent(inout) :: s
    integer :: i
    do i = 1, 4
      if (c_associated(s%dr(i))) call hy_check(hy_device_free(s%dr(i)), 'free')
      s%dr(i) = c_null_ptr
    end do
    if (c_associated(s%dc)) call hy_check(hy_device_free(s%dc), 'free')
    if (c_associated(s%dk)) call hy_check(hy_device_free(s%dk), 'free')
    if (c_associated(s%g)) call hy_check(hy_grid_destroy(s%g), 'grid_destroy')
    s%dc = c_null_ptr; s%dk = c_null_ptr; s%g = c_null_ptr
  end subroutine spec_destroy

  integer(c_long_long) function nbr(s)
    class(hip_spec_t), intent(in) :: s
    nbr = int(s%nx, c_long_long)*s%ny*s%nz*8
  end function nbr

  integer(c_long_long) function nbc(s)
    class(hip_spec_t), intent(in) :: s
    nbc = int(s%nkx, c_long_long)*s%ny*s%nz*16
  end function nbc

  subroutine spec_fwd(s, u, uhat)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: u(:,:,:)
    complex(rp), intent(out), target :: uhat(:,:,:)
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'fwd')
    call hy_check(hy_fft_r2c(s%g, s%dr(1), s%dc, c_null_ptr), 'fwd')
    call hy_check(hy_copy_to_host(c_loc(uhat), s%dc, nbc(s)), 'fwd')
  end subroutine spec_fwd

  subroutine spec_bwd(s, uhat, u)
    class(hip_spec_t), intent(inout) :: s
    complex(rp), intent(in), target :: uhat(:,:,:)
    real(rp), intent(out), target :: u(:,:,:)
    call hy_check(hy_copy_to_device(s%dc, c_loc(uhat), nbc(s)), 'bwd')
    call hy_check(hy_fft_c2r(s%g, s%dc, s%dr(1), c_null_ptr), 'bwd')
    call hy_check(hy_copy_to_host(c_loc(u), s%dr(1), nbr(s)), 'bwd')
  end subroutine spec_bwd

  subroutine spec_grad_hat(s, uhat, idir, du)
    class(hip_spec_t), intent(inout) :: s
    complex(rp), intent(in), target :: uhat(:,:,:)
    integer, intent(in) :: idir
    real(rp), intent(out), target :: du(:,:,:)
    call hy_check(hy_copy_to_device(s%dc, c_loc(uhat), nbc(s)), 'grad_hat')
    call hy_check(hy_spec_grad_hat(s%g, s%dc, int(idir, c_int), s%dr(1), c_null_ptr), 'grad_hat')
    call hy_check(hy_copy_to_host(c_loc(du), s%dr(1), nbr(s)), 'grad_hat')
  end subroutine spec_grad_hat

  subroutine spec_grad(s, u, du)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: u(:,:,:)
    real(rp), intent(out), target :: du(:,:,:,:)
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'grad')
    call hy_check(hy_spec_grad(s%g, s%dr(1), s%dr(2), c_null_ptr), 'grad')
    call hy_check(hy_copy_to_host(c_loc(du), s%dr(2), 3*nbr(s)), 'grad')
  end subroutine spec_grad

  subroutine spec_lap(s, u, lu)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: u(:,:,:)
    real(rp), intent(out), target :: lu(:,:,:)
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'lap')
    call hy_check(hy_spec_lap(s%g, s%dr(1), s%dr(2), c_null_ptr), 'lap')
    call hy_check(hy_copy_to_host(c_loc(lu), s%dr(2), nbr(s)), 'lap')
  end subroutine spec_lap

  subroutine spec_div(s, v, dv)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: v(:,:,:,:)
    real(rp), intent(out), target :: dv(:,:,:)
    call hy_check(hy_copy_to_device(s%dr(2), c_loc(v), 3*nbr(s)), 'div')
    call hy_check(hy_spec_div(s%g, s%dr(2), s%dr(1), c_null_ptr), 'div')
    call hy_check(hy_copy_to_host(c_loc(dv), s%dr(1), nbr(s)), 'div')
  end subroutine spec_div

  subroutine spec_convolve(s, kernel, uhat, cu)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: kernel(:,:,:)
    complex(rp), intent(in), target :: uhat(:,:,:)
    real(rp), intent(out), target :: cu(:,:,:)
    call hy_check(hy_copy_to_device(s%dk, c_loc(kernel), nbc(s)/2), 'convolve')
    call hy_check(hy_copy_to_device(s%dc, c_loc(uhat), nbc(s)), 'convolve')
    call hy_check(hy_spec_convolve(s%g, s%dk, s%dc, s%dr(1), c_null_ptr), 'convolve')
    call hy_check(hy_copy_to_host(c_loc(cu), s%dr(1), nbr(s)), 'convolve')
  end subroutine spec_convolve

  subroutine spec_dealias(s, u)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(inout), target :: u(:,:,:)
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'dealias')
    call hy_check(hy_spec_dealias(s%g, s%dr(1), c_null_ptr), 'dealias')
    call hy_check(hy_copy_to_host(c_loc(u), s%dr(1), nbr(s)), 'dealias')
  end subroutine spec_dealias

  real(rp) function spec_mean(s, u)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: u(:,:,:)
    real(c_double) :: m
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'mean')
    call hy_check(hy_spec_mean(s%g, s%dr(1), m, c_null_ptr), 'mean')
    spec_mean = m
  end function spec_mean

  complex(rp) function spec_mode(s, u, mx, my, mz)
    class(hip_spec_t), intent(inout) :: s
    real(rp), intent(in), target :: u(:,:,:)
    integer, intent(in) :: mx, my, mz
    real(c_double) :: z(2)
    call hy_check(hy_copy_to_device(s%dr(1), c_loc(u), nbr(s)), 'mode')
    call hy_check(hy_spec_mode(s%g, s%dr(1), int(mx, c_int), int(my, c_int), int(mz, c_int), z, c_null_ptr), &
                  'mode')
    spec_mode = cmplx(z(1), z(2), rp)
  end function spec_mode

end module hydft_hip_spectral
