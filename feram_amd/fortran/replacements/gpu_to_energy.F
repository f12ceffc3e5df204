! gpu_to_energy.F (GPU build) -- the 13 energies of a device call, in
! Energy_type member order (energy_module.F:9-22), into an Energy_type.
subroutine gpu_to_energy(g, e)
  use Energy_module
  implicit none
  real*8, intent(in) :: g(13)
  type(Energy_type), intent(out) :: e
  e%acou_kinetic = g(1);    e%dipo_kinetic = g(2)
  e%short_range = g(3);     e%long_range = g(4)
  e%dipole_E_field = g(5);  e%unharmonic = g(6)
  e%homo_strain = g(7);     e%homo_coupling = g(8)
  e%inho_strain = g(9);     e%inho_coupling = g(10)
  e%inho_modulation = g(11); e%total = g(12)
  e%H_Nose_Poincare = g(13)
end subroutine gpu_to_energy
