! feram_gpu.f90 -- the device context of a feram run and the step calls on it.
!
! The context is kept here as an opaque C pointer, as feram keeps its FFTW
! plans in Coord_type (coord_module.F:55-63), and lives across input files
! (feram_common.F:74-81, continue = .true.).  Every call checks the status
! and stops with the library's message, in the reference's style of fatal
! errors (elastic.F:198-210).  The step routines take and return plain
! arrays: energies in Energy_type member order (energy_module.F:9-22), the
! Nose-Poincare scalars s_Nose, pi_Nose, H0 (coord_module.F:53) and
! c%strain(6), so the unit that replaces a reference subroutine only copies
! them into the reference's types (replacements/).
module feram_gpu
  use, intrinsic :: iso_c_binding
  use feram_hip_binding
  implicit none
  private
  public :: gpu, gpu_check, gpu_open, gpu_close, gpu_energy_array, gpu_scalars, &
       &    gpu_nose_poincare, gpu_leapfrog, gpu_velocity_scaling, gpu_force, gpu_md_prologue

  type(c_ptr), save :: gpu = c_null_ptr

contains

  subroutine gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       write(0, '(a,a,a,i0,a,a)') 'FATAL: ', what, ' returned ', rc, ': ', fhip_error_message(gpu)
       stop 1
    end if
  end subroutine gpu_check

  ! fhip_create once per process (feram_common.F:82-83, after elastic(p,c))
  subroutine gpu_open(params)
    type(fhip_params), intent(in) :: params
    if (c_associated(gpu)) return
    call gpu_check(fhip_create(params, gpu), 'fhip_create')
  end subroutine gpu_open

  subroutine gpu_close()
    if (c_associated(gpu)) call fhip_destroy(gpu)
    gpu = c_null_ptr
  end subroutine gpu_close

  ! fhip_energy -> the 13 Energy_type members in their declaration order
  subroutine gpu_energy_array(g, e)
    type(fhip_energy), intent(in) :: g
    real(c_double), intent(out) :: e(13)
    e = [g%acou_kinetic, g%dipo_kinetic, g%short_range, g%long_range, g%dipole_E_field, &
         g%unharmonic, g%homo_strain, g%homo_coupling, g%inho_strain, g%inho_coupling, &
         g%inho_modulation, g%total, g%H_Nose_Poincare]
  end subroutine gpu_energy_array

  ! thermostat scalars and homogeneous strain after a call
  subroutine gpu_scalars(s_Nose, pi_Nose, H0, strain)
    real(c_double), intent(out) :: s_Nose, pi_Nose, H0, strain(6)
    call gpu_check(fhip_get_nose(gpu, s_Nose, pi_Nose, H0), 'fhip_get_nose')
    call gpu_check(fhip_get_strain(gpu, strain), 'fhip_get_strain')
  end subroutine gpu_scalars

  ! one step of each method, as molecular-dynamics.F:113-120 calls them
  subroutine gpu_nose_poincare(e)
    real(c_double), intent(out) :: e(13)
    type(fhip_energy) :: g
    call gpu_check(fhip_nose_poincare(gpu, 1, g), 'fhip_nose_poincare')
    call gpu_energy_array(g, e)
  end subroutine gpu_nose_poincare

  subroutine gpu_leapfrog(i_step, e)
    integer, intent(in) :: i_step
    real(c_double), intent(out) :: e(13)
    type(fhip_energy) :: g
    call gpu_check(fhip_leapfrog(gpu, int(i_step, c_int), 1, g), 'fhip_leapfrog')
    call gpu_energy_array(g, e)
  end subroutine gpu_leapfrog

  subroutine gpu_velocity_scaling(e)
    real(c_double), intent(out) :: e(13)
    type(fhip_energy) :: g
    call gpu_check(fhip_velocity_scaling(gpu, 1, g), 'fhip_velocity_scaling')
    call gpu_energy_array(g, e)
  end subroutine gpu_velocity_scaling

  ! mirror_force_mirror (mirror_force_mirror.F:6-89); dV/du stays on the device
  subroutine gpu_force(e)
    real(c_double), intent(out) :: e(13)
    type(fhip_energy) :: g
    call gpu_check(fhip_force(gpu, g), 'fhip_force')
    call gpu_energy_array(g, e)
  end subroutine gpu_force

  ! initialize-dipoR.F:59-69 and the H0 set-up of molecular-dynamics.F:44-55
  subroutine gpu_md_prologue(e)
    real(c_double), intent(out) :: e(13)
    type(fhip_energy) :: g
    call gpu_check(fhip_md_prologue(gpu, g), 'fhip_md_prologue')
    call gpu_energy_array(g, e)
  end subroutine gpu_md_prologue
end module feram_gpu
