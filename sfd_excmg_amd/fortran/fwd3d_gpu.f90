! fwd3d_gpu.f90 -- module fwd3d with the reference's entry points, backed by
! libsfdx.so on the GPU.
!
!   fwd_solver(modelname,a,b,c,nair,level,slv,eps,sigma,freq,ex)
!       same signature as sfd_excmg/src/fdfe_main.f90:1565; allocates
!       ex(2*nl,nf) (xy then yx per frequency) and fills it on the device.
!   post_process(modelname,a,b,c,nair,abr,freq,recv,u,out_type[,topo])
!       same signature as sfd_excmg/src/fdfe_main.f90:549; with topo (the
!       elevations read_topo returns) the surface of each column follows the
!       ground; writes <modelname>.res.
!
! Run-time switches (the reference's compile-time parameters in
! global_para.f90:5-19): SFDX_FWD_ACC (0 = single level, the reference
! default; 1 = EXCMG), SFDX_CORR_INTV, SFDX_CORR_TOL, SFDX_PRECISION
! (fp64|mixed), SFDX_NORM (reference|interior), SFDX_SOLVER (overrides slv:
! 1 BiCGStab, 3 GPBiCG, 5 COCR), SFDX_PRECOND (0 SSOR, 1 hybrid, 2 multigrid,
! 3 = the library default: multigrid or hybrid per frequency and level),
! SFDX_COARSE (0 direct solves of EXCMG levels 1-2, 1 device Krylov solves),
! SFDX_FULL_BG (1: full-tensor 1-D boundary with cross-polar components),
! SFDX_MG_LEVELS and SFDX_MG_PAIRS (multigrid grids and coarsest smoothing
! pairs), SFDX_RESTART_DROP (mixed precision: restart factor), SFDX_SPMV_TILE,
! SFDX_VERBOSE, SFDX_RES_FULL (also write <name>.resx with 17 significant digits).
module fwd3d
  use iso_c_binding
  use sfdx_iface
  use file_operation
  implicit none
  type(c_ptr), save :: sfdx_handle = c_null_ptr

contains

  integer function env_int(name, default)
    character(len=*), intent(in) :: name
    integer, intent(in) :: default
    character(len=64) :: v
    integer :: st
    env_int = default
    call get_environment_variable(name, v, status=st)
    if (st == 0 .and. len_trim(v) > 0) read(v, *) env_int
  end function

  real(8) function env_real(name, default)
    character(len=*), intent(in) :: name
    real(8), intent(in) :: default
    character(len=64) :: v
    integer :: st
    env_real = default
    call get_environment_variable(name, v, status=st)
    if (st == 0 .and. len_trim(v) > 0) read(v, *) env_real
  end function

  logical function env_is(name, value)
    character(len=*), intent(in) :: name, value
    character(len=64) :: v
    integer :: st
    call get_environment_variable(name, v, status=st)
    env_is = (st == 0 .and. trim(v) == value)
  end function

  subroutine fwd_solver(modelname, a, b, c, nair, level, slv, eps, sigma, freq, ex)
    character(len=*) :: modelname
    real(8), allocatable, target :: a(:), b(:), c(:), sigma(:,:), freq(:)
    integer :: nair, level, slv
    real(8) :: eps
    complex(8), allocatable :: ex(:,:)
    type(sfdx_grid_t) :: g
    type(sfdx_opts_t) :: o
    type(sfdx_stats_t) :: st
    integer :: nx, ny, nz, nf, nl, rc
    integer(c_int) :: rcc

    nx = size(a); ny = size(b); nz = size(c); nf = size(freq)
    nl = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    allocate(ex(2 * nl, nf))
    if (.not. c_associated(sfdx_handle)) then
      rcc = sfdx_create(int(env_int('SFDX_DEVICE', 0), c_int), sfdx_handle)
      if (rcc /= 0) then
        write(*,*) 'sfdx_create failed: ', sfdx_error_text()
        stop 1
      end if
    end if
    g%nx = nx; g%ny = ny; g%nz = nz; g%nair = nair
    g%a = c_loc(a); g%b = c_loc(b); g%c = c_loc(c)
    call sfdx_default_opts(o)
    o%level = level
    o%use_excmg = env_int('SFDX_FWD_ACC', 0)
    o%solver = env_int('SFDX_SOLVER', slv)
    o%eps = eps
    o%corr_intv = env_int('SFDX_CORR_INTV', 200)
    o%corr_tol = env_real('SFDX_CORR_TOL', 1.d-3)
    if (env_is('SFDX_PRECISION', 'mixed')) o%precision = SFDX_PREC_MIXED
    if (env_is('SFDX_NORM', 'interior')) o%norm = SFDX_NORM_INTERIOR
    o%verbose = env_int('SFDX_VERBOSE', 0)
    o%precond = env_int('SFDX_PRECOND', int(o%precond))
    o%coarse_solver = env_int('SFDX_COARSE', 0)
    o%full_background = env_int('SFDX_FULL_BG', 0)
    o%spmv_tile = env_int('SFDX_SPMV_TILE', 0)
    o%mg_levels = env_int('SFDX_MG_LEVELS', int(o%mg_levels))
    o%mg_coarse_pairs = env_int('SFDX_MG_PAIRS', int(o%mg_coarse_pairs))
    o%restart_drop = env_real('SFDX_RESTART_DROP', o%restart_drop)
    rc = sfdx_fwd_solve(sfdx_handle, g, sigma, int(nf, c_int), freq, o, ex, st)
    if (rc > 0) then
      write(*,*) 'Error in the GPU forward solve: ', sfdx_error_text()
      stop 1
    else if (rc < 0) then
      write(*,*) 'Warning: maximum iterations reached, last iterate kept'
    end if
    write(*, '(a, f10.3, a, f10.3, a, i8)') ' GPU solve: total', st%total_seconds, ' s, fine level', &
        st%fine_seconds, ' s, fine iterations', st%fine_iters_total
  end subroutine

  subroutine post_process(modelname, a, b, c, nair, abr, freq, recv, u, out_type, topo)
    character(len=*) :: modelname
    real(8), intent(in), target :: a(:), b(:), c(:), freq(:), recv(:,:)
    integer, intent(in) :: nair, abr(:), out_type
    complex(8), allocatable, intent(in) :: u(:,:)
    real(8), allocatable, optional :: topo(:,:)
    real(8), allocatable, target :: a_(:), b_(:), c_(:)
    real(8), allocatable :: rt(:,:), res(:,:), tip(:,:), arho(:,:), tipper(:,:)
    integer(c_int), allocatable :: surf(:,:)
    type(sfdx_grid_t) :: g
    integer :: nrec, nf, rc, k, u2

    nrec = size(recv, 1); nf = size(freq)
    allocate(a_(size(a)), b_(size(b)), c_(size(c)))
    a_ = a; b_ = b; c_ = c
    g%nx = size(a); g%ny = size(b); g%nz = size(c); g%nair = nair
    g%a = c_loc(a_); g%b = c_loc(b_); g%c = c_loc(c_)
    allocate(rt(3, nrec), res(4, nrec * nf), tip(4, nrec * nf))
    rt = transpose(recv)
    if (present(topo)) then
      if (allocated(topo)) then
        ! topography branch (fdfe_main.f90:604-614, 656-666)
        allocate(surf(size(a), size(b)))
        call surface_planes_of(c, nair, topo, surf)
        rc = sfdx_post_process_topo(g, surf, int(nf, c_int), freq, int(nrec, c_int), rt, u, int(out_type, c_int), &
                                    2_c_int, res, tip)
      else
        rc = sfdx_post_process(g, int(nf, c_int), freq, int(nrec, c_int), rt, u, int(out_type, c_int), 2_c_int, &
                               res, tip)
      end if
    else
      rc = sfdx_post_process(g, int(nf, c_int), freq, int(nrec, c_int), rt, u, int(out_type, c_int), 2_c_int, res, tip)
    end if
    if (rc /= 0) then
      write(*,*) 'Error in post_process: ', sfdx_error_text()
      stop 1
    end if
    allocate(arho(nrec * nf, 4), tipper(nrec * nf, 4))
    arho = transpose(res)
    tipper = transpose(tip)
    call output_data(modelname, freq, recv, out_type, arho, tipper, .false.)
    if (env_int('SFDX_RES_FULL', 0) /= 0) then
      open(newunit=u2, file=trim(adjustl(modelname))//'.resx', status='replace', action='write')
      do k = 1, nrec * nf
        write(u2, '(4es25.16e3)') arho(k, :)
      end do
      close(u2)
    end if
    if (abr(1) < 0) write(*,*) abr(1)
  end subroutine

end module
