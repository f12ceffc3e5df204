! velocity_scaling.F (GPU build) -- replaces the reference's velocity_scaling
! (velocity_scaling.F:9-103): one velocity-scaling step on the device.
subroutine velocity_scaling(p,c,e)
  use Param_module
  use Coord_module
  use Energy_module
  use feram_gpu
  implicit none
  type(Param_type),  intent(inout) :: p
  type(Coord_type),  intent(inout) :: c
  type(Energy_type), intent(out)   :: e
  real*8 g(13), s, pi, H0
  call gpu_velocity_scaling(g)
  call gpu_to_energy(g, e)
  call gpu_scalars(s, pi, H0, c%strain)
end subroutine velocity_scaling
