!> iso_c_binding layer over the model entry points of libhydft
!> (include/hydft.h): the device-resident model_t (src/model/model.f90:29-49)
!> with its functional terms and closure, whole RK3 steps with the driver's
!> compute_dt cadence (src/hydft.f90:41-51), budgets and probes
!> (src/model/diagnostics.f90:41-77), captured steps and batches of models.
!>
!> hy_term_desc and hy_model_desc are bind(C) mirrors of the C structs: the
!> host fills them from its species_t / params_t (species.f90:97-153) and
!> hands them to hy_model_create.  Arrays cross as c_loc of host arrays in
!> the reference layout (x fastest, vector fields component-major).
module hydft_hip_model
  use, intrinsic :: iso_c_binding
  use hydft_hip_c
  implicit none
  public

  integer(c_int), parameter :: HY_TERM_IDEAL = 0, HY_TERM_TF = 1, HY_TERM_GRADIENT = 2, &
                               HY_TERM_HARTREE = 3, HY_TERM_RY = 4
  integer(c_int), parameter :: HY_CLOSURE_NEWTONIAN = 0, HY_CLOSURE_MAXWELL = 1
  integer(c_int), parameter :: HY_VEXT_NONE = 0, HY_VEXT_MODE = 1, HY_VEXT_DRIVEN = 2, &
                               HY_VEXT_GAUSSIAN = 3

  !> fe_term_t scalars (free_energy.f90:23-25, units.f90:31, species.f90:127-136)
  type, bind(C) :: hy_term_desc
    integer(c_int) :: kind = 0
    real(c_double) :: n0 = 0, beta = 0
    real(c_double) :: e2 = 0, kappa = 0
    real(c_double) :: cn = 0, alpha0 = 0, cg = 0
    integer(c_int) :: subtract_mean_field = 0
    integer(c_int) :: ntab = 0
    type(c_ptr) :: k_tab = c_null_ptr, c_tab = c_null_ptr
  end type hy_term_desc

  !> grid, species, closure, v_ext and time-step parameters of a model
  type, bind(C) :: hy_model_desc
    integer(c_int) :: nx = 1, ny = 1, nz = 1
    real(c_double) :: lx = 1, ly = 1, lz = 1
    integer(c_int) :: dealias = 1
    integer(c_int) :: precision = 0
    real(c_double) :: mass = 1, n0 = 0
    integer(c_int) :: nterms = 0
    type(c_ptr) :: terms = c_null_ptr
    integer(c_int) :: closure = 0
    real(c_double) :: eta = 0, xi = 0, tau = 0
    real(c_double) :: eta_l = 0
    integer(c_int) :: vext_kind = 0
    real(c_double) :: vext_amp = 0
    integer(c_int) :: vext_mode(3) = 0
    real(c_double) :: vext_omega = 0, vext_width = 0, vext_ramp = 0
    real(c_double) :: dt = 0, cfl = 0
  end type hy_model_desc

  interface
    integer(c_int) function hy_model_create(desc, m) bind(C, name='hy_model_create')
      import :: c_int, c_ptr, hy_model_desc
      type(hy_model_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function hy_model_destroy(m) bind(C, name='hy_model_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function hy_model_set_state(m, n, mom3, pi6) bind(C, name='hy_model_set_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, n, mom3, pi6
    end function
    integer(c_int) function hy_model_get_state(m, n, mom3, pi6) bind(C, name='hy_model_get_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, n, mom3, pi6
    end function
    integer(c_int) function hy_model_get_time(m, t, dt, step) bind(C, name='hy_model_get_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: t, dt
      integer(c_int), intent(out) :: step
    end function
    integer(c_int) function hy_model_compute_dt(m, dt) bind(C, name='hy_model_compute_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function hy_model_advance(m, nsteps, dt_every, tend, probe3, probe, taken) &
        bind(C, name='hy_model_advance')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      integer(c_int), value :: nsteps, dt_every
      real(c_double), value :: tend
      type(c_ptr), value :: probe3, probe
      integer(c_int), intent(out) :: taken
    end function
    integer(c_int) function hy_model_rhs(m, t, dn, dmom3, dpi6) bind(C, name='hy_model_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), value :: t
      type(c_ptr), value :: dn, dmom3, dpi6
    end function
    integer(c_int) function hy_model_budgets(m, out6) bind(C, name='hy_model_budgets')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: out6(6)
    end function
    integer(c_int) function hy_model_probe(m, mx, my, mz, out4) bind(C, name='hy_model_probe')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      integer(c_int), value :: mx, my, mz
      real(c_double), intent(out) :: out4(4)
    end function
    integer(c_int) function hy_model_kernel(m, k, kk) bind(C, name='hy_model_kernel')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), value :: k
      real(c_double), intent(out) :: kk
    end function
    ! captured steps (no counterpart in the reference)
    integer(c_int) function hy_model_graph_build(m, nsteps, dt_every, probe3, ring) &
        bind(C, name='hy_model_graph_build')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, probe3
      integer(c_int), value :: nsteps, dt_every, ring
    end function
    integer(c_int) function hy_model_graph_launch(m, nlaunch) bind(C, name='hy_model_graph_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nlaunch
    end function
    integer(c_int) function hy_model_graph_probe(m, probe) bind(C, name='hy_model_graph_probe')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, probe
    end function
    integer(c_int) function hy_model_idle(m, idle) bind(C, name='hy_model_idle')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), intent(out) :: idle
    end function
    ! batches of same-shape models (no counterpart in the reference)
    integer(c_int) function hy_batch_create(models, n, nsteps, dt_every, probe3, ring, b) &
        bind(C, name='hy_batch_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: models, probe3
      integer(c_int), value :: n, nsteps, dt_every, ring
      type(c_ptr), intent(out) :: b
    end function
    integer(c_int) function hy_batch_launch(b, nlaunch) bind(C, name='hy_batch_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nlaunch
    end function
    integer(c_int) function hy_batch_sync(b) bind(C, name='hy_batch_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function hy_batch_destroy(b) bind(C, name='hy_batch_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    !> per-run step limits of a batch (stops(i): steps of run i, a multiple of the
    !> batch's steps per launch; 0: none)
    integer(c_int) function hy_batch_set_stops(b, stops) bind(C, name='hy_batch_set_stops')
      import :: c_int, c_ptr
      type(c_ptr), value :: b, stops
    end function
    ! groups of batches on different grids: one graph launch per pass
    integer(c_int) function hy_group_create(batches, n, g) bind(C, name='hy_group_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: batches
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: g
    end function
    integer(c_int) function hy_group_launch(g, nlaunch) bind(C, name='hy_group_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
      integer(c_int), value :: nlaunch
    end function
    integer(c_int) function hy_group_idle(g, idle) bind(C, name='hy_group_idle')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
      integer(c_int), intent(out) :: idle
    end function
    integer(c_int) function hy_group_sync(g) bind(C, name='hy_group_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function
    integer(c_int) function hy_group_destroy(g) bind(C, name='hy_group_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function
    integer(c_int) function hy_model_probe_ring(m, ring, probe) bind(C, name='hy_model_probe_ring')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, probe
      integer(c_int), value :: ring
    end function
    !> kept 2/3-rule box (Kx, Ky, Kz) of a model: models batch together only
    !> when their boxes agree
    integer(c_int) function hy_model_kept(m, kept3) bind(C, name='hy_model_kept')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), intent(out) :: kept3(3)
    end function
    ! batched OZ-HNC direct solve on the GPU (oz_hnc.f90:111-296 for many state points)
    integer(c_int) function hy_hnc_iterate(nsolve, nr, dr, dk, n0, bvs, bvlk, gs0, tol, mix, maxiter, gs, status, &
        niter, resid) bind(C, name='hy_hnc_iterate')
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nsolve, nr, maxiter
      real(c_double), value :: dr, dk, tol, mix
      type(c_ptr), value :: n0, bvs, bvlk, gs0, gs, status, niter, resid
    end function
  end interface

end module hydft_hip_model
