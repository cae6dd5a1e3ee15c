! file_operation.f90 -- the reference's on-disk formats, read and written by
! the drop-in driver: <name>.fwd / <name>.dat in, <name>.res out
! (sfd_excmg/src/file_opt.f90:8-95, 131-158).
module file_operation
  use iso_c_binding, only: c_double
  implicit none
  ! air conductivity written into the air layers after reading
  ! (sfd_excmg/src/global_para.f90:12 uses 1e-9; SFDX_SIGMA_AIR overrides)
  real(8) :: sigma_air = 1.d-9

contains

  subroutine read_input(model_name, a, b, c, nair, level, eps, solver, sigma, freq, recv, out_type, abu, topo)
    character(len=*), intent(in) :: model_name
    real(8), allocatable :: a(:), b(:), c(:), sigma(:,:), freq(:), recv(:,:)
    real(8), allocatable, optional :: topo(:,:)
    integer :: nair, level, solver, out_type, abu(6)
    real(8) :: eps
    integer :: nx, ny, nz, topo_indic, conduc_type, i, j, k, nf, nrec, ie, u, ios
    real(8) :: ext(2), s1, s3(3), s6(6)
    character(len=256) :: line, env

    call get_environment_variable("SFDX_SIGMA_AIR", env, status=ios)
    if (ios == 0) read(env, *) sigma_air
    open(newunit=u, file=trim(adjustl(model_name))//'.fwd', status='old', action='read')
    read(u, '(a)') line
    read(u, *) level, eps, solver, out_type
    if (level < 3) stop 'Error: the grid level should be larger than 3!'
    if (eps >= 1e-8) write(*,*) 'Warning: Too big tolerance, the SFD results might be wrong!'
    if (solver < 0 .or. solver > 5) stop 'Error: Unsupported solver type!'
    if (out_type < 0 .or. out_type >= 3) stop 'Error: Unsupported ouput type!'
    read(u, *) nx, ny, nz, nair, topo_indic
    if (nx / 2**level < 2) stop 'Error: the grid level is too big, please check'
    allocate(a(nx), b(ny), c(nz + nair))
    read(u, *) a
    read(u, *) b
    read(u, *) c(nair + 1:nair + nz)
    read(u, *) ext(1), ext(2)
    read(u, *) abu
    c(1:nair) = c(nz + nair:nz + 1:-1)
    read(u, *) conduc_type
    allocate(sigma(nx * ny * (nz + nair), 6))
    sigma = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          ie = (nair + k - 1) * nx * ny + (j - 1) * nx + i
          select case (conduc_type)
          case (1)
            read(u, *) s1
            sigma(ie, 1:3) = s1
          case (2)
            read(u, *) s3
            sigma(ie, 1:3) = s3
          case default
            read(u, *) s6
            sigma(ie, :) = s6
          end select
        end do
      end do
    end do
    sigma(1:nair * nx * ny, 1:3) = sigma_air
    sigma(1:nair * nx * ny, 4:6) = 0
    close(u)
    if (topo_indic /= 0) then
      if (present(topo)) then
        call read_topo(model_name, a, b, c, nair, sigma, topo)
      else
        write(*,*) 'Warning: topography flag set but no topo argument; the .topo file is not read'
      end if
    end if

    open(newunit=u, file=trim(adjustl(model_name))//'.dat', status='old', action='read')
    read(u, *) nf
    allocate(freq(nf))
    do i = 1, nf
      read(u, *) freq(i)
    end do
    read(u, *) nrec
    allocate(recv(nrec, 3))
    do i = 1, nrec
      read(u, *) recv(i, :)
    end do
    close(u)
  end subroutine

  ! <name>.topo (the reference's read_topo, file_opt.f90:97-128, corrected):
  ! line 1 "n1 n2", the square window of cell columns i, j in n1..n2 (1-based);
  ! then (n2-n1+1)**2 ground elevations in metres above the flat datum (the
  ! air/earth interface), i fastest.  topo(nx, ny) returns the elevation of
  ! every column, 0 outside the window.  Every cell whose centre lies above the
  ! ground becomes air (sigma_air on the diagonal, off-diagonals 0); every air
  ! cell whose centre lies below it becomes earth with the tensor of the
  ! column's first earth cell (nair+1).  The reference sets columns 1, 4, 6
  ! (xx, xy, yz) of the raised cells to sigma_hhs, allocates topo with
  ! (n2-n1)**2 entries and cannot lower the ground below the datum.
  subroutine read_topo(model_name, a, b, c, nair, sigma, topo)
    character(len=*), intent(in) :: model_name
    real(8), intent(in) :: a(:), b(:), c(:)
    integer, intent(in) :: nair
    real(8) :: sigma(:,:)
    real(8), allocatable :: topo(:,:)
    real(8), allocatable :: ccz(:)
    real(8) :: earth(6)
    integer :: n1, n2, i, j, k, nx, ny, nz, u, ie, iearth
    nx = size(a); ny = size(b); nz = size(c)
    allocate(topo(nx, ny), ccz(nz))
    topo = 0
    ! cell-centre depths below the datum (fdfe_main.f90:592-594)
    do k = 1, nz
      ccz(k) = -sum(c(1:nair)) + sum(c(1:k)) - c(k) / 2
    end do
    open(newunit=u, file=trim(adjustl(model_name))//'.topo', status='old', action='read')
    read(u, *) n1, n2
    if (n1 < 1 .or. n2 > min(nx, ny) .or. n2 < n1) stop 'Error: topography window outside the grid'
    do j = n1, n2
      do i = n1, n2
        read(u, *) topo(i, j)
      end do
    end do
    close(u)
    do j = n1, n2
      do i = n1, n2
        iearth = nair * nx * ny + (j - 1) * nx + i
        earth = sigma(iearth, :)
        do k = 1, nz
          ie = (k - 1) * nx * ny + (j - 1) * nx + i
          if (ccz(k) < -topo(i, j)) then
            sigma(ie, 1:3) = sigma_air
            sigma(ie, 4:6) = 0
          else if (k <= nair) then
            sigma(ie, :) = earth
          end if
        end do
      end do
    end do
  end subroutine

  ! Surface node plane of every cell column for the topography branch of
  ! post_process (fdfe_main.f90:604-614): one less than the first cell k1 whose
  ! centre lies below the ground, i.e. the top node plane of that cell
  ! (0-based); nair on the flat datum.
  subroutine surface_planes_of(c, nair, topo, surf)
    real(8), intent(in) :: c(:), topo(:,:)
    integer, intent(in) :: nair
    integer, intent(out) :: surf(:,:)
    integer :: i, j, k, nz
    real(8) :: zc
    nz = size(c)
    do j = 1, size(topo, 2)
      do i = 1, size(topo, 1)
        surf(i, j) = nz - 1
        do k = 1, nz
          zc = -sum(c(1:nair)) + sum(c(1:k)) - c(k) / 2
          if (zc >= -topo(i, j)) then
            surf(i, j) = k - 1
            exit
          end if
        end do
      end do
    end do
  end subroutine

  ! <name>.res: one line per (frequency, station): x y z f type v1..v4
  ! (the reference prints an uninitialised integer in the type column,
  ! file_opt.f90:143; out_type is written there instead)
  subroutine output_data(model_name, freq, recv, out_type, arho, tipper, with_tipper)
    character(len=*), intent(in) :: model_name
    real(8), intent(in) :: freq(:), recv(:,:), arho(:,:), tipper(:,:)
    integer, intent(in) :: out_type
    logical, intent(in) :: with_tipper
    integer :: i, k, nf, nrec, u
    nf = size(freq)
    nrec = size(recv, 1)
    open(newunit=u, file=trim(adjustl(model_name))//'.res', status='replace', action='write')
    do k = 1, nf
      do i = 1, nrec
        write(u, 145) recv(i, :), freq(k), out_type, arho((k - 1) * nrec + i, :)
      end do
    end do
    if (with_tipper) then
      do k = 1, nf
        do i = 1, nrec
          write(u, 145) recv(i, :), freq(k), 3, tipper((k - 1) * nrec + i, :)
        end do
      end do
    end if
    close(u)
145 format(4(e12.5, 2x), i1, 2x, 4(e12.5, 2x))
  end subroutine

end module
