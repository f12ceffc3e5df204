! fortran_np.f90 -- a Fortran driver of the device step through feram_gpu, as
! the replacement units of a feram build call it: parameters (the bind(C)
! fhip_params record) and u (feram's padded layout) from files in the
! directory given as the first argument, fhip_md_prologue, then five
! one-step Nose-Poincare calls; prints the energies after each call and the
! thermostat scalars after each step, and writes u and p to out.bin.
program fortran_np
  use, intrinsic :: iso_c_binding
  use feram_hip_binding
  use feram_gpu
  implicit none
  type(fhip_params) :: params
  real(c_double), allocatable :: u(:), pb(:)
  real(c_double) :: e(13), s, pi, H0, strain(6)
  integer :: n, k, unit
  character(len=1024) :: dir
  call get_command_argument(1, dir)
  open(newunit=unit, file=trim(dir)//'/params.bin', access='stream', form='unformatted', status='old')
  read(unit) params
  close(unit)
  n = 3 * 2 * (params%Lx / 2 + 1) * (params%Ly + params%padding_y) * params%Lz
  allocate(u(n), pb(n))
  open(newunit=unit, file=trim(dir)//'/u.bin', access='stream', form='unformatted', status='old')
  read(unit) u
  close(unit)
  call gpu_open(params)
  call gpu_check(fhip_put(gpu, FHIP_DIPO_R, u), 'fhip_put')
  call gpu_md_prologue(e)
  write(*, '(a,13es25.16e3)') 'E ', e
  do k = 1, 5
     call gpu_nose_poincare(e)
     call gpu_scalars(s, pi, H0, strain)
     write(*, '(a,13es25.16e3)') 'E ', e
     write(*, '(a,3es25.16e3)') 'S ', s, pi, H0
  end do
  call gpu_check(fhip_get(gpu, FHIP_DIPO_R, u), 'fhip_get')
  call gpu_check(fhip_get(gpu, FHIP_DIPO_P, pb), 'fhip_get')
  open(newunit=unit, file=trim(dir)//'/out.bin', access='stream', form='unformatted', status='replace')
  write(unit) u, pb
  close(unit)
  call gpu_close()
end program fortran_np
