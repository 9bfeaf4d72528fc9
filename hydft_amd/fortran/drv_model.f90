!> The model entry points through the Fortran binding (hydft_hip_model):
!> a single-mode run of classical ions in the manner of the reference's
!> test_mode_newtonian.f90 (mode 2, nx = 32, lx = 20, Gamma = 2, kappa = 0.1,
!> amplitude 1e-4, compute_dt every 20 steps at cfl 0.1), here with the ideal
!> and Hartree terms and no viscosity, so that linear theory is the closed
!> form omega = q sqrt(n0 K(q) / m), K(q) = 1/(beta n0) + 4 pi e2/(q^2 + kappa^2).
!> Checks: frequency within 1 % of linear theory, mass conserved to 1e-12
!> (test_conservation.f90), the kept 2/3-rule box, batched and captured
!> steps equal to hy_model_advance bitwise, and a group of a batch whose runs
!> have different step limits.
program drv_model
  use, intrinsic :: iso_c_binding
  use hydft_hip_c
  use hydft_hip_model
  implicit none
  integer, parameter :: nx = 32, mode = 2
  real(c_double), parameter :: pi = 3.14159265358979323846264338327950288_c_double
  real(c_double), parameter :: gamma = 2.0_c_double, kappa = 0.1_c_double, lx = 20.0_c_double
  type(hy_term_desc), target :: terms(2)
  type(hy_model_desc) :: desc
  type(c_ptr) :: m, ms(2), ma, mg, b, g
  type(c_ptr), target :: pair(2), grp(1)
  integer(c_int), target :: probe3(3), stops(2)
  real(c_double), allocatable, target :: rows(:,:), rows_a(:,:), rows_b(:,:), rows_g(:,:)
  real(c_double), target :: n(nx), mom(nx, 3)
  real(c_double) :: n0, beta, e2, q, kq, om_lin, om, t, dt, bud0(6), bud1(6), err
  real(c_double) :: tz(2000)
  integer(c_int) :: taken, step, idle, kept(3)
  integer :: i, nz, nfail, nsteps

  nfail = 0
  call hy_check(hy_init(0_c_int), 'init')
  n0 = 3.0_c_double / (4.0_c_double * pi)     ! units.f90: Wigner-Seitz units
  beta = 3.0_c_double * gamma
  e2 = 1.0_c_double / 3.0_c_double
  terms(1)%kind = HY_TERM_IDEAL
  terms(2)%kind = HY_TERM_HARTREE
  do i = 1, 2
    terms(i)%n0 = n0; terms(i)%beta = beta; terms(i)%e2 = e2; terms(i)%kappa = kappa
  end do
  desc%nx = nx; desc%lx = lx
  desc%mass = 1.0_c_double; desc%n0 = n0
  desc%nterms = 2; desc%terms = c_loc(terms)
  desc%closure = HY_CLOSURE_NEWTONIAN
  desc%cfl = 0.1_c_double
  probe3 = [mode, 0, 0]

  ! ---- single mode vs linear theory
  call hy_check(hy_model_create(desc, m), 'model_create')
  ! kept 2/3-rule box: |kx| < (2/3) (2 pi / lx) (nx / 2) keeps kx indices 0..10 of nx = 32
  call hy_check(hy_model_kept(m, kept), 'model_kept')
  call check(all(kept == [11, 1, 1]), 'kept box', real(kept(1), c_double))
  call set_mode(m, 1.0e-4_c_double)
  call hy_check(hy_model_budgets(m, bud0), 'budgets')
  nsteps = 1000
  allocate(rows(6, nsteps))
  call hy_check(hy_model_advance(m, nsteps, 20_c_int, 100.0_c_double, c_loc(probe3), c_loc(rows), taken), &
                'advance')
  call hy_check(hy_model_get_time(m, t, dt, step), 'get_time')
  call hy_check(hy_model_budgets(m, bud1), 'budgets')
  q = 2.0_c_double * pi * mode / lx
  kq = 1.0_c_double / (beta * n0) + 4.0_c_double * pi * e2 / (q * q + kappa * kappa)
  om_lin = q * sqrt(n0 * kq)
  ! zero crossings of Re(dn_k / n0) after t = 5
  nz = 0
  do i = 2, taken
    if (rows(1, i - 1) >= 5.0_c_double .and. rows(2, i - 1) * rows(2, i) < 0.0_c_double) then
      nz = nz + 1
      tz(nz) = rows(1, i - 1) - rows(2, i - 1) * (rows(1, i) - rows(1, i - 1)) / (rows(2, i) - rows(2, i - 1))
    end if
  end do
  om = pi * (nz - 1) / (tz(nz) - tz(1))
  write(*, '(a,i0,a,f10.4,a,es10.2,a,f10.6,a,f10.6,a,i0)') '  steps = ', taken, '  t =', t, '  dt =', dt, &
      '  linear omega =', om_lin, '  simulation =', om, '  zeros = ', nz
  err = abs(om / om_lin - 1.0_c_double); call check(err < 0.01_c_double .and. nz > 10, 'frequency', err)
  err = abs(bud1(1) / bud0(1) - 1.0_c_double); call check(err < 1.0e-12_c_double, 'mass', err)
  call hy_check(hy_model_destroy(m), 'destroy')

  ! ---- batched and captured steps equal hy_model_advance
  allocate(rows_a(6, 40), rows_b(6, 40), rows_g(6, 40))
  call hy_check(hy_model_create(desc, ms(1)), 'model_create')
  call hy_check(hy_model_create(desc, ms(2)), 'model_create')
  call hy_check(hy_model_create(desc, ma), 'model_create')
  call hy_check(hy_model_create(desc, mg), 'model_create')
  call set_mode(ms(1), 1.0e-4_c_double)
  call set_mode(ms(2), 2.0e-4_c_double)
  call set_mode(ma, 1.0e-4_c_double)
  call set_mode(mg, 2.0e-4_c_double)
  call hy_check(hy_model_advance(ma, 40_c_int, 20_c_int, 0.0_c_double, c_loc(probe3), c_loc(rows_a), taken), &
                'advance')
  pair = ms
  call hy_check(hy_batch_create(c_loc(pair), 2_c_int, 20_c_int, 20_c_int, c_loc(probe3), 40_c_int, b), 'batch_create')
  call hy_check(hy_batch_launch(b, 2_c_int), 'batch_launch')
  call hy_check(hy_batch_sync(b), 'batch_sync')
  call hy_check(hy_model_probe_ring(ms(1), 40_c_int, c_loc(rows_b)), 'probe_ring')
  call check(all(rows_b == rows_a), 'batch = advance', maxval(abs(rows_b - rows_a)))
  call hy_check(hy_model_graph_build(mg, 20_c_int, 20_c_int, c_loc(probe3), 40_c_int), 'graph_build')
  call hy_check(hy_model_graph_launch(mg, 2_c_int), 'graph_launch')
  idle = 0
  do while (idle == 0)
    call hy_check(hy_model_idle(mg, idle), 'idle')
  end do
  call hy_check(hy_model_graph_probe(mg, c_loc(rows_g)), 'graph_probe')
  call hy_check(hy_model_probe_ring(ms(2), 40_c_int, c_loc(rows_b)), 'probe_ring')
  call check(all(rows_b == rows_g), 'batch = graph', maxval(abs(rows_b - rows_g)))
  call hy_check(hy_batch_destroy(b), 'batch_destroy')
  do i = 1, 2
    call hy_check(hy_model_destroy(ms(i)), 'destroy')
  end do

  ! ---- a group launch of a batch whose runs stop at 20 and 40 steps
  call hy_check(hy_model_create(desc, ms(1)), 'model_create')
  call hy_check(hy_model_create(desc, ms(2)), 'model_create')
  call set_mode(ms(1), 1.0e-4_c_double)
  call set_mode(ms(2), 1.0e-4_c_double)
  pair = ms
  stops = [20_c_int, 40_c_int]
  call hy_check(hy_batch_create(c_loc(pair), 2_c_int, 20_c_int, 20_c_int, c_loc(probe3), 40_c_int, b), 'batch_create')
  call hy_check(hy_batch_set_stops(b, c_loc(stops)), 'batch_set_stops')
  grp(1) = b
  call hy_check(hy_group_create(c_loc(grp), 1_c_int, g), 'group_create')
  call hy_check(hy_group_launch(g, 2_c_int), 'group_launch')
  call hy_check(hy_group_sync(g), 'group_sync')
  call hy_check(hy_group_idle(g, idle), 'group_idle')
  call check(idle == 1, 'group idle after sync', real(idle, c_double))
  call hy_check(hy_group_destroy(g), 'group_destroy')
  call hy_check(hy_model_probe_ring(ms(2), 40_c_int, c_loc(rows_b)), 'probe_ring')
  call check(all(rows_b == rows_a), 'group = advance', maxval(abs(rows_b - rows_a)))
  call hy_check(hy_model_probe_ring(ms(1), 40_c_int, c_loc(rows_b)), 'probe_ring')
  call check(all(rows_b(:, 1:20) == rows_a(:, 1:20)), 'group, 20-step run', &
             maxval(abs(rows_b(:, 1:20) - rows_a(:, 1:20))))
  call hy_check(hy_model_get_time(ms(1), t, dt, step), 'get_time')
  call check(step == 20, 'run stopped at its limit', real(step, c_double))
  call hy_check(hy_batch_destroy(b), 'batch_destroy')
  do i = 1, 2
    call hy_check(hy_model_destroy(ms(i)), 'destroy')
  end do
  call hy_check(hy_model_destroy(ma), 'destroy')
  call hy_check(hy_model_destroy(mg), 'destroy')

  if (nfail > 0) then
    write(*, '(a,i0,a)') 'drv_model: ', nfail, ' failure(s)'; error stop 1
  end if
  write(*, '(a)') 'drv_model: all passed'

contains

  !> n = n0 (1 + a cos(q x)), mom = 0 (initial.f90 'mode')
  subroutine set_mode(mm, a)
    type(c_ptr), intent(in) :: mm
    real(c_double), intent(in) :: a
    integer :: j
    do j = 1, nx
      n(j) = n0 * (1.0_c_double + a * cos(2.0_c_double * pi * mode * (j - 1) / nx))
    end do
    mom = 0.0_c_double
    call hy_check(hy_model_set_state(mm, c_loc(n), c_loc(mom), c_null_ptr), 'set_state')
  end subroutine set_mode

  subroutine check(ok, name, val)
    logical, intent(in) :: ok
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: val
    if (ok) then
      write(*, '(a,a,es12.3)') '  pass  ', name, val
    else
      write(*, '(a,a,es12.3)') '  FAIL  ', name, val
      nfail = nfail + 1
    end if
  end subroutine check

end program drv_model
