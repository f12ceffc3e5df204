! mirror_force_mirror.F (GPU build) -- replaces the reference's
! mirror_force_mirror (mirror_force_mirror.F:6-89): the force on the device's
! u; callers that read c%dVddi fetch it with fhip_get(gpu, FHIP_DVDDI, c%dVddi).
subroutine mirror_force_mirror(p,c,e)
  use Param_module
  use Coord_module
  use Energy_module
  use feram_gpu
  implicit none
  type(Param_type),  intent(inout) :: p
  type(Coord_type),  intent(inout) :: c
  type(Energy_type), intent(out)   :: e
  real*8 g(13), s, pi, H0
  call gpu_force(g)
  call gpu_to_energy(g, e)
  call gpu_scalars(s, pi, H0, c%strain)
end subroutine mirror_force_mirror
