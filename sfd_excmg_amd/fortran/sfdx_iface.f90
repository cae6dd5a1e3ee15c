! sfdx_iface.f90 -- ISO_C_BINDING interfaces to libsfdx.so (include/sfdx.h).
module sfdx_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SFDX_SOLVER_BICGSTAB = 1, SFDX_SOLVER_GPBICG = 3, SFDX_SOLVER_COCR = 5
  integer(c_int), parameter :: SFDX_PREC_FP64 = 0, SFDX_PREC_MIXED = 1, SFDX_PREC_MIXED_IR = 2, &
                                SFDX_PREC_MIXED_SPLIT = 3
  integer(c_int), parameter :: SFDX_PRECOND_SSOR = 0, SFDX_PRECOND_HYBRID = 1, SFDX_PRECOND_MG = 2, &
                                SFDX_PRECOND_AUTO = 3
  integer(c_int), parameter :: SFDX_NORM_REFERENCE = 0, SFDX_NORM_INTERIOR = 1

  type, bind(C) :: sfdx_grid_t
    integer(c_int) :: nx, ny, nz, nair
    type(c_ptr) :: a, b, c
  end type

  type, bind(C) :: sfdx_opts_t
    integer(c_int) :: level, use_excmg, solver
    real(c_double) :: eps
    integer(c_int) :: corr_intv
    real(c_double) :: corr_tol, ssor_omega
    integer(c_int) :: precision, norm, maxit_plain, maxit_excmg
    real(c_double) :: inner_tol
    integer(c_int) :: max_refine, verbose, dc_restart, dc_at_start, inner_maxit, precond
    real(c_double) :: sigma_floor32
    integer(c_int) :: mg_levels, mg_coarse_pairs
    real(c_double) :: restart_drop
    integer(c_int) :: use_graph, nodal_smoother, coarse_solver, full_background, spmv_tile, ssor_cross_mass
    real(c_double) :: surface_sigma
    integer(c_int) :: pol_streams
    real(c_double) :: mg_merge_ratio
  end type

  type, bind(C) :: sfdx_stats_t
    integer(c_int) :: status
    real(c_double) :: total_seconds, fine_seconds, fine_bytes
    integer(c_int) :: iters(2, 8)
    real(c_double) :: relres(2, 8)
    integer(c_int) :: dc_calls(2, 8), pcg_iters(2, 8), refine_steps(2), fine_iters_total, fallback_fp64, &
                      solver_switches, spmv_tile, mg_freqs
  end type

  interface
    integer(c_int) function sfdx_create(device, ctx) bind(C, name="sfdx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function sfdx_destroy(ctx) bind(C, name="sfdx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function sfdx_last_error() bind(C, name="sfdx_last_error")
      import :: c_ptr
    end function
    subroutine sfdx_default_opts(o) bind(C, name="sfdx_default_opts")
      import :: sfdx_opts_t
      type(sfdx_opts_t) :: o
    end subroutine
    integer(c_int) function sfdx_fwd_solve(ctx, g, sigma, nf, freq, o, ex, st) bind(C, name="sfdx_fwd_solve")
      import :: c_int, c_ptr, c_double, c_double_complex, sfdx_grid_t, sfdx_opts_t, sfdx_stats_t
      type(c_ptr), value :: ctx
      type(sfdx_grid_t), intent(in) :: g
      real(c_double), intent(in) :: sigma(*)
      integer(c_int), value :: nf
      real(c_double), intent(in) :: freq(*)
      type(sfdx_opts_t), intent(in) :: o
      complex(c_double_complex) :: ex(*)
      type(sfdx_stats_t) :: st
    end function
    integer(c_int) function sfdx_post_process(g, nf, freq, nrec, recv, ex, out_type, rd, results, tipper) &
        bind(C, name="sfdx_post_process")
      import :: c_int, c_double, c_double_complex, sfdx_grid_t
      type(sfdx_grid_t), intent(in) :: g
      integer(c_int), value :: nf, nrec, out_type, rd
      real(c_double), intent(in) :: freq(*), recv(*)
      complex(c_double_complex), intent(in) :: ex(*)
      real(c_double) :: results(*), tipper(*)
    end function
    integer(c_int) function sfdx_post_process_topo(g, surf, nf, freq, nrec, recv, ex, out_type, rd, results, &
        tipper) bind(C, name="sfdx_post_process_topo")
      import :: c_int, c_double, c_double_complex, sfdx_grid_t
      type(sfdx_grid_t), intent(in) :: g
      integer(c_int), intent(in) :: surf(*)
      integer(c_int), value :: nf, nrec, out_type, rd
      real(c_double), intent(in) :: freq(*), recv(*)
      complex(c_double_complex), intent(in) :: ex(*)
      real(c_double) :: results(*), tipper(*)
    end function
  end interface

contains

  function sfdx_error_text() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    type(c_ptr) :: cp
    cp = sfdx_last_error()
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    do while (n > 0)
      msg(n:n) = p(n)
      n = n - 1
    end do
  end function

end module
