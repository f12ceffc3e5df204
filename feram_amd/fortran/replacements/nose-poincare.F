! nose-poincare.F (GPU build) -- replaces the reference's nose_poincare
! (nose-poincare.F:10-78): one Nose-Poincare step on the device.
subroutine nose_poincare(p,c,e)
  use Param_module
  use Coord_module
  use Energy_module
  use feram_gpu
  implicit none
  type(Param_type),  intent(inout) :: p
  type(Coord_type),  intent(inout) :: c
  type(Energy_type), intent(out)   :: e
  real*8 g(13)
  call gpu_nose_poincare(g)
  call gpu_to_energy(g, e)
  call gpu_scalars(c%s_Nose, c%pi_Nose, c%H0, c%strain)
end subroutine nose_poincare
