!> test_fft.f90 and test_ops.f90 of the reference, run through hip_spec_t
!> (the GPU drop-in for spec_t).  Same grids, same checks, same tolerances;
!> prints pass/FAIL lines and ends with error stop 1 on any failure.
program drv_spec
  use, intrinsic :: iso_c_binding
  use hydft_hip_c
  use hydft_hip_spectral
  implicit none
  integer, parameter :: rp = c_double
  real(rp), parameter :: pi = 3.14159265358979323846264338327950288_rp, twopi = 2.0_rp*pi
  complex(rp), parameter :: i_unit = (0.0_rp, 1.0_rp)
  type(hip_spec_t) :: s
  real(rp), allocatable :: u(:,:,:), v(:,:,:), du(:,:,:,:), lu(:,:,:), ex(:,:,:), vv(:,:,:,:), dv(:,:,:)
  real(rp), allocatable :: cu(:,:,:), kern(:,:,:)
  complex(rp), allocatable :: uhat(:,:,:)
  integer :: i, j, k, nfail
  real(rp) :: err, x, y, z, ly, kx, kyv
  complex(rp) :: a

  nfail = 0
  call hy_check(hy_init(0_c_int), 'init')

  ! ---- test_fft.f90:20-36
  call s%init(64, 1, 1, twopi, twopi, twopi)
  allocate(u(64,1,1), v(64,1,1), uhat(s%nkx,1,1))
  do i = 1, 64
    x = (i-1)*twopi/64
    u(i,1,1) = 1.5_rp + 0.3_rp*cos(3.0_rp*x) + 0.2_rp*sin(5.0_rp*x)
  end do
  call s%fwd(u, uhat)
  call s%bwd(uhat, v)
  err = maxval(abs(u - v)); call check(err < 1.0e-13_rp, '1-D round trip', err)
  call check(abs(uhat(1,1,1) - 1.5_rp) < 1.0e-13_rp, '1-D mean', abs(uhat(1,1,1) - 1.5_rp))
  call check(abs(uhat(4,1,1) - 0.15_rp) < 1.0e-13_rp, '1-D cos mode', abs(uhat(4,1,1) - 0.15_rp))
  call check(abs(uhat(6,1,1) + 0.1_rp*i_unit) < 1.0e-13_rp, '1-D sin mode', abs(uhat(6,1,1) + 0.1_rp*i_unit))
  a = s%mode(u, 3, 0, 0)
  call check(abs(a - 0.15_rp) < 1.0e-13_rp, '1-D mode()', abs(a - 0.15_rp))
  deallocate(u, v, uhat)

  ! ---- test_fft.f90:38-52
  call s%init(16, 12, 8, twopi, 4.0_rp, 1.0_rp)
  allocate(u(16,12,8), v(16,12,8), uhat(s%nkx,12,8))
  do k = 1, 8; do j = 1, 12; do i = 1, 16
    x = (i-1)*twopi/16; y = (j-1)*4.0_rp/12; z = (k-1)*1.0_rp/8
    u(i,j,k) = 0.7_rp + cos(2.0_rp*x)*sin(twopi*y/4.0_rp) + 0.4_rp*cos(twopi*2.0_rp*z)
  end do; end do; end do
  call s%fwd(u, uhat)
  call s%bwd(uhat, v)
  err = maxval(abs(u - v)); call check(err < 1.0e-13_rp, '3-D round trip', err)
  a = s%mode(u, 0, 0, 2)
  call check(abs(a - 0.2_rp) < 1.0e-13_rp, '3-D z mode', abs(a - 0.2_rp))
  a = s%mode(u, 2, 1, 0)
  call check(abs(a + 0.25_rp*i_unit) < 1.0e-13_rp, '3-D xy mode', abs(a + 0.25_rp*i_unit))
  deallocate(u, v, uhat)

  ! ---- test_ops.f90:30-71
  ly = 4.0_rp
  call s%init(32, 16, 1, twopi, ly, 1.0_rp)
  allocate(u(32,16,1), du(32,16,1,3), lu(32,16,1), ex(32,16,1), vv(32,16,1,3), dv(32,16,1), cu(32,16,1))
  allocate(kern(s%nkx,16,1), uhat(s%nkx,16,1))
  do j = 1, 16; do i = 1, 32
    x = (i-1)*twopi/32; y = (j-1)*ly/16
    u(i,j,1) = sin(2.0_rp*x)*cos(twopi*y/ly)
  end do; end do
  call s%grad(u, du)
  do j = 1, 16; do i = 1, 32
    x = (i-1)*twopi/32; y = (j-1)*ly/16
    ex(i,j,1) = 2.0_rp*cos(2.0_rp*x)*cos(twopi*y/ly)
  end do; end do
  err = maxval(abs(du(:,:,:,1) - ex)); call check(err < 1e-12_rp, 'd/dx', err)
  do j = 1, 16; do i = 1, 32
    x = (i-1)*twopi/32; y = (j-1)*ly/16
    ex(i,j,1) = -(twopi/ly)*sin(2.0_rp*x)*sin(twopi*y/ly)
  end do; end do
  err = maxval(abs(du(:,:,:,2) - ex)); call check(err < 1e-12_rp, 'd/dy', err)
  err = maxval(abs(du(:,:,:,3))); call check(err < 1e-14_rp, 'd/dz = 0 in 2-D', err)
  call s%lap(u, lu)
  ex = -(4.0_rp + (twopi/ly)**2)*u
  err = maxval(abs(lu - ex)); call check(err < 1e-11_rp, 'laplacian', err)
  vv(:,:,:,1) = du(:,:,:,1); vv(:,:,:,2) = du(:,:,:,2); vv(:,:,:,3) = 0.0_rp
  call s%div(vv, dv)
  err = maxval(abs(dv - ex)); call check(err < 1e-11_rp, 'div grad = lap', err)
  do j = 1, 16; do i = 1, s%nkx
    kx = twopi*(i-1)/twopi
    kyv = twopi*merge(j-1, j-1-16, j-1 <= 8)/ly
    kern(i,j,1) = 1.0_rp/(1.0_rp + kx*kx + kyv*kyv)
  end do; end do
  call s%fwd(u, uhat)
  call s%convolve(kern, uhat, cu)
  ex = u/(1.0_rp + 4.0_rp + (twopi/ly)**2)
  err = maxval(abs(cu - ex)); call check(err < 1e-13_rp, 'convolution', err)
  do j = 1, 16; do i = 1, 32
    x = (i-1)*twopi/32
    u(i,j,1) = cos(12.0_rp*x) + cos(10.0_rp*x)
    ex(i,j,1) = cos(10.0_rp*x)
  end do; end do
  call s%dealias(u)
  err = maxval(abs(u - ex)); call check(err < 1e-13_rp, 'dealias removes k > 2/3 kmax', err)
  err = abs(s%kmax_dealiased - sqrt(100.0_rp + (twopi*5.0_rp/ly)**2)); call check(err < 1e-12_rp, 'kmax_dealiased', err)
  u = 3.0_rp + du(:,:,:,1)
  err = abs(s%mean(u) - 3.0_rp); call check(err < 1e-13_rp, 'mean', err)
  call s%destroy()

  if (nfail > 0) then
    write(*,'(a,i0,a)') 'drv_spec: ', nfail, ' failure(s)'
    error stop 1
  end if
  write(*,'(a)') 'drv_spec: all passed'

contains
  subroutine check(ok, name, val)
    logical, intent(in) :: ok
    character(len=*), intent(in) :: name
    real(rp), intent(in) :: val
    if (ok) then
      write(*,'(a,a,es12.3)') '  pass  ', name, val
    else
      write(*,'(a,a,es12.3)') '  FAIL  ', name, val
      nfail = nfail + 1
    end if
  end subroutine check
end program drv_spec
