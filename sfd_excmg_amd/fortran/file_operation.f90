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

  subroutine read_input(model_name, a, b, c, nair, level, eps, solver, sigma, freq, recv, out_type, abu)
    character(len=*), intent(in) :: model_name
    real(8), allocatable :: a(:), b(:), c(:), sigma(:,:), freq(:), recv(:,:)
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
    if (topo_indic /= 0) write(*,*) 'Warning: .topo files are not read; carve topography into sigma instead'

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
