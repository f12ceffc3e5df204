! leapfrog.F (GPU build) -- replaces the reference's leapfrog
! (leapfrog.F:9-157): one leapfrog step on the device, heat-source layers
! included when fhip_set_heat_source was called at set-up.
subroutine leapfrog(p,c,e,i_step)
  use Param_module
  use Coord_module
  use Energy_module
  use feram_gpu
  implicit none
  type(Param_type),  intent(inout) :: p
  type(Coord_type),  intent(inout) :: c
  type(Energy_type), intent(out)   :: e
  integer i_step
  real*8 g(13), s, pi, H0
  call gpu_leapfrog(i_step, g)
  call gpu_to_energy(g, e)
  call gpu_scalars(s, pi, H0, c%strain)
end subroutine leapfrog
