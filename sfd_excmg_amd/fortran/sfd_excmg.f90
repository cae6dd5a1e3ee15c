! sfd_excmg.f90 -- drop-in driver with the reference's command line
! (sfd_excmg/src/sfd_excmg.f90:20-58):  sfd_excmg_gpu <path/name> [level eps]
program sfd_excmg
  use file_operation
  use fwd3d
  implicit none
  real(8), allocatable :: a(:), b(:), c(:), sigma(:,:), freq(:), recv(:,:), topo(:,:)
  complex(8), allocatable :: ex(:,:)
  integer :: out_type, nair, level, slv, abu(6), n
  real(8) :: eps
  character(len=256) :: name, str

  n = command_argument_count()
  call get_command_argument(1, name)
  call read_input(name, a, b, c, nair, level, eps, slv, sigma, freq, recv, out_type, abu, topo)
  if (n > 1) then
    call get_command_argument(2, str)
    read(str, *) level
    call get_command_argument(3, str)
    read(str, *) eps
  end if
  call fwd_solver(name, a, b, c, nair, level, slv, eps, sigma, freq, ex)
  call post_process(name, a, b, c, nair, abu, freq, recv, ex, out_type, topo)
end program
