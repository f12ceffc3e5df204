! feram_hip_binding.f90 -- ISO_C_BINDING interface to libferam_hip.so.
!
! The module declares the C ABI of include/feram_hip.h for Fortran callers
! (interfaces only).  feram_gpu.f90 holds the device context and the step
! calls on top of it, and replacements/ the four units a feram build links in
! place of nose-poincare.F, leapfrog.F, velocity_scaling.F and
! mirror_force_mirror.F (file:line citations are into feram's src/).
module feram_hip_binding
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: fhip_params, fhip_energy, fhip_average
  public :: fhip_create, fhip_destroy, fhip_last_error, fhip_set_long_range, fhip_get_long_range, &
       &    fhip_set_modulation, fhip_set_cells, fhip_put, fhip_get, fhip_get_nose, fhip_set_nose, &
       &    fhip_get_strain, fhip_set_E_field_add, fhip_force, fhip_md_prologue, &
       &    fhip_nose_poincare, fhip_leapfrog, fhip_velocity_scaling, fhip_synchronize, &
       &    fhip_set_kelvin, fhip_set_heat_source, fhip_get_layer_energies, fhip_set_E_schedule, fhip_get_step, fhip_avg_reset, fhip_avg_add, fhip_avg_auto, fhip_avg_get, &
       &    fhip_batch_create, fhip_batch_destroy, fhip_batch_step
  public :: FHIP_DIPO_R, FHIP_DIPO_P, FHIP_DVDDI, FHIP_ACOU_R, FHIP_ACOU_P, FHIP_DVDAC, &
       &    FHIP_E_LOCAL, FHIP_ACOU_R_OPT, FHIP_DIPO_R_AVG, fhip_error_message

  integer(c_int), parameter :: FHIP_DIPO_R = 0, FHIP_DIPO_P = 1, FHIP_DVDDI = 2, FHIP_ACOU_R = 3, &
       &                       FHIP_ACOU_P = 4, FHIP_DVDAC = 5, FHIP_E_LOCAL = 6, FHIP_ACOU_R_OPT = 7, &
       &                       FHIP_DIPO_R_AVG = 8

  ! include/feram_hip.h: struct fhip_params (same member order)
  type, bind(C) :: fhip_params
     integer(c_int32_t) :: Lx, Ly, Lz, padding_y
     integer(c_int32_t) :: method, bulk_or_film, precision, E_field_always_zero
     real(c_double) :: dt, kelvin, Q_Nose, a0, epi_strain
     real(c_double) :: GPa(3)
     real(c_double) :: mass_dipo(3), mass_acou(3)
     real(c_double) :: P_kappa2(3), P_alpha(3), P_gamma(3), P_k1(3), P_k2(3), P_k3(3), P_k4(3)
     real(c_double) :: j(7)
     real(c_double) :: B11, B12, B44, B1xx, B1yy, B4yz
     real(c_double) :: modulation_constant
     real(c_double) :: Z_star(3)
     real(c_double) :: epsilon_inf, kappa
     real(c_double) :: external_E_field(3)
     integer(c_int32_t) :: gap_id, reserved
  end type fhip_params

  ! Energy_type twin (energy_module.F:7-23)
  type, bind(C) :: fhip_energy
     real(c_double) :: acou_kinetic, dipo_kinetic, short_range, long_range, dipole_E_field, &
          &            unharmonic, homo_strain, homo_coupling, inho_strain, inho_coupling, &
          &            inho_modulation, total, H_Nose_Poincare
  end type fhip_energy

  ! Average_type sums (average_module.F:17-30), as add_to_Average builds them
  type, bind(C) :: fhip_average
     type(fhip_energy) :: energy
     real(c_double) :: e2, acou_kinetic_true, dipo_kinetic_true
     real(c_double) :: strain(6)
     real(c_double) :: u(3), uu(6), p(3), pp(6)
     real(c_double) :: u_tmp(3), uu_tmp(3), p_tmp(3), pp_tmp(3)
     integer(c_int64_t) :: n_add
  end type fhip_average

  interface
     integer(c_int) function fhip_create(p, ctx) bind(C, name='fhip_create')
       import :: c_int, c_ptr, fhip_params
       type(fhip_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function fhip_create
     subroutine fhip_destroy(ctx) bind(C, name='fhip_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine fhip_destroy
     type(c_ptr) function fhip_last_error(ctx) bind(C, name='fhip_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function fhip_last_error
     integer(c_int) function fhip_set_long_range(ctx, lr) bind(C, name='fhip_set_long_range')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: lr(*)
     end function fhip_set_long_range
     integer(c_int) function fhip_get_long_range(ctx, lr) bind(C, name='fhip_get_long_range')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lr(*)
     end function fhip_get_long_range
     integer(c_int) function fhip_set_modulation(ctx, m) bind(C, name='fhip_set_modulation')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: m(*)
     end function fhip_set_modulation
     integer(c_int) function fhip_set_cells(ctx, relaxor, sum_p, minv) bind(C, name='fhip_set_cells')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, relaxor, sum_p, minv
     end function fhip_set_cells
     integer(c_int) function fhip_put(ctx, field, host) bind(C, name='fhip_put')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(in) :: host(*)
     end function fhip_put
     integer(c_int) function fhip_get(ctx, field, host) bind(C, name='fhip_get')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(inout) :: host(*)
     end function fhip_get
     integer(c_int) function fhip_get_nose(ctx, s, pi, H0) bind(C, name='fhip_get_nose')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: s, pi, H0
     end function fhip_get_nose
     integer(c_int) function fhip_set_nose(ctx, s, pi, H0) bind(C, name='fhip_set_nose')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: s, pi, H0
     end function fhip_set_nose
     integer(c_int) function fhip_get_strain(ctx, strain) bind(C, name='fhip_get_strain')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: strain(6)
     end function fhip_get_strain
     integer(c_int) function fhip_set_E_field_add(ctx, E) bind(C, name='fhip_set_E_field_add')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: E(3)
     end function fhip_set_E_field_add
     integer(c_int) function fhip_force(ctx, e) bind(C, name='fhip_force')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: ctx
       type(fhip_energy), intent(out) :: e
     end function fhip_force
     integer(c_int) function fhip_md_prologue(ctx, e) bind(C, name='fhip_md_prologue')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: ctx
       type(fhip_energy), intent(out) :: e
     end function fhip_md_prologue
     integer(c_int) function fhip_nose_poincare(ctx, nsteps, e) bind(C, name='fhip_nose_poincare')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       type(fhip_energy), intent(out) :: e
     end function fhip_nose_poincare
     integer(c_int) function fhip_leapfrog(ctx, i_step, nsteps, e) bind(C, name='fhip_leapfrog')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: ctx
       integer(c_int), value :: i_step, nsteps
       type(fhip_energy), intent(out) :: e
     end function fhip_leapfrog
     integer(c_int) function fhip_velocity_scaling(ctx, nsteps, e) bind(C, name='fhip_velocity_scaling')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       type(fhip_energy), intent(out) :: e
     end function fhip_velocity_scaling
     ! reset_Average / add_to_Average / the averaging loop / the sums (average_module.F:32-163)
     integer(c_int) function fhip_avg_reset(ctx, options) bind(C, name='fhip_avg_reset')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: options
     end function fhip_avg_reset
     integer(c_int) function fhip_avg_add(ctx) bind(C, name='fhip_avg_add')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fhip_avg_add
     integer(c_int) function fhip_avg_auto(ctx, enable) bind(C, name='fhip_avg_auto')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function fhip_avg_auto
     integer(c_int) function fhip_avg_get(ctx, avg, distribution, distribution3d) bind(C, name='fhip_avg_get')
       import :: c_int, c_ptr, fhip_average
       type(c_ptr), value :: ctx
       type(fhip_average), intent(out) :: avg
       type(c_ptr), value :: distribution, distribution3d
     end function fhip_avg_get
     integer(c_int) function fhip_set_E_schedule(ctx, type, amplitude, period, n_thermalize, n_average, &
          &                                      i_step) bind(C, name='fhip_set_E_schedule')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: type, period, n_thermalize, n_average
       real(c_double), intent(in) :: amplitude(3)
       integer(c_int64_t), value :: i_step
     end function fhip_set_E_schedule
     integer(c_int) function fhip_get_step(ctx, i_step, E_add) bind(C, name='fhip_get_step')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: i_step
       real(c_double), intent(out) :: E_add(3)
     end function fhip_get_step
     integer(c_int) function fhip_set_heat_source(ctx, delta_e, delta_e_acou) bind(C, name='fhip_set_heat_source')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: delta_e, delta_e_acou
     end function fhip_set_heat_source
     integer(c_int) function fhip_get_layer_energies(ctx, e_dipo, e_acou, averaged) &
          &                                          bind(C, name='fhip_get_layer_energies')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, e_dipo, e_acou
       integer(c_int), value :: averaged
     end function fhip_get_layer_energies
     integer(c_int) function fhip_set_kelvin(ctx, kelvin) bind(C, name='fhip_set_kelvin')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: kelvin
     end function fhip_set_kelvin
     ! Replica batches: ctxs(1:n) are the contexts of a temperature / field sweep.
     integer(c_int) function fhip_batch_create(ctxs, n, batch) bind(C, name='fhip_batch_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(in) :: ctxs(*)
       integer(c_int), value :: n
       type(c_ptr), intent(out) :: batch
     end function fhip_batch_create
     subroutine fhip_batch_destroy(batch) bind(C, name='fhip_batch_destroy')
       import :: c_ptr
       type(c_ptr), value :: batch
     end subroutine fhip_batch_destroy
     integer(c_int) function fhip_batch_step(batch, nsteps, e) bind(C, name='fhip_batch_step')
       import :: c_int, c_ptr, fhip_energy
       type(c_ptr), value :: batch
       integer(c_int), value :: nsteps
       type(fhip_energy), intent(out) :: e(*)
     end function fhip_batch_step
     integer(c_int) function fhip_synchronize(ctx) bind(C, name='fhip_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fhip_synchronize
     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! fhip_last_error() as a Fortran string, for msg(UNIT_LOG, ..., 'FATAL', ...).
  function fhip_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: cp
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    cp = fhip_last_error(ctx)
    n = int(c_strlen(cp))
    call c_f_pointer(cp, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function fhip_error_message
end module feram_hip_binding
