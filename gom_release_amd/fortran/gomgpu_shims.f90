! gomgpu_shims.f90 - replacements of the reference's hot-path subroutines that run each one
! as a stage of the GPU context bound by module gomgpu_bridge. They keep the reference's
! external names and argument-less interfaces, so GOM's main.f90 links against them
! unchanged in place of these objects:
!   solve_nonlinear_advection, solve_momentum_equation, solve_free_surface_equation,
!   solve_velocities, update_vertical_layers_1, calculate_velocity_at_node,
!   solve_transport_equation_v19, solve_transport_equation_ELM_v8, calculate_density_full,
!   calculate_density_linear, update_variables.
! Calls before the time loop (it = 0: prepare_gom.f90:135, :153-155) run the stage from the
! module state and bring the state back; calls in the loop leave it on the device until
! update_variables (gomgpu_bridge's gg_step_end).

! main.f90:167-176 -> solve_nonlinear_advection.f90:14-204, with the bottom friction of
! main.f90:95 (calculate_bottom_friction.f90) recomputed on the device first.
subroutine solve_nonlinear_advection
  use gomgpu_iface, only: GOMGPU_STAGE_FRICTION, GOMGPU_STAGE_ELM
  use gomgpu_bridge
  implicit none
  call gg_stage(GOMGPU_STAGE_FRICTION, 'calculate_bottom_friction')
  call gg_stage(GOMGPU_STAGE_ELM, 'solve_nonlinear_advection')
end subroutine solve_nonlinear_advection

! main.f90:180 -> solve_momentum_equation.f90:15-671. With advection off main.f90:167-172
! copies un_face into un_ELM on the host; the device makes that copy (and the friction)
! itself before the momentum stage.
subroutine solve_momentum_equation
  use gomgpu_iface, only: GOMGPU_STAGE_FRICTION, GOMGPU_STAGE_ELM, GOMGPU_STAGE_MOMENTUM
  use gomgpu_bridge
  use mod_global_variables, only: advection_flag
  implicit none
  if (advection_flag == 0) then
    call gg_stage(GOMGPU_STAGE_FRICTION, 'calculate_bottom_friction')
    call gg_stage(GOMGPU_STAGE_ELM, 'un_ELM = un_face')
  end if
  call gg_stage(GOMGPU_STAGE_MOMENTUM, 'solve_momentum_equation')
end subroutine solve_momentum_equation

subroutine solve_free_surface_equation
  use gomgpu_iface, only: GOMGPU_STAGE_FREESURF
  use gomgpu_bridge
  implicit none
  call gg_stage(GOMGPU_STAGE_FREESURF, 'solve_free_surface_equation')
end subroutine solve_free_surface_equation

subroutine solve_velocities
  use gomgpu_iface, only: GOMGPU_STAGE_VELOCITY
  use gomgpu_bridge
  implicit none
  call gg_stage(GOMGPU_STAGE_VELOCITY, 'solve_velocities')
end subroutine solve_velocities

subroutine update_vertical_layers_1
  use gomgpu_iface, only: GOMGPU_STAGE_LAYERS
  use gomgpu_bridge
  implicit none
  call gg_stage(GOMGPU_STAGE_LAYERS, 'update_vertical_layers_1')
end subroutine update_vertical_layers_1

subroutine calculate_velocity_at_node
  use gomgpu_iface, only: GOMGPU_STAGE_NODEVEL
  use gomgpu_bridge
  use mod_global_variables, only: it
  implicit none
  if (it == 0) then
    call gg_startup_stage(GOMGPU_STAGE_NODEVEL, 'calculate_velocity_at_node')
  else
    call gg_stage(GOMGPU_STAGE_NODEVEL, 'calculate_velocity_at_node')
    call gg_after_nodevel()
  end if
end subroutine calculate_velocity_at_node

subroutine solve_transport_equation_v19
  use gomgpu_iface, only: GOMGPU_STAGE_TRANSPORT
  use gomgpu_bridge
  implicit none
  call gg_tvd_boundary()
  call gg_stage(GOMGPU_STAGE_TRANSPORT, 'solve_transport_equation_v19')
end subroutine solve_transport_equation_v19

subroutine solve_transport_equation_ELM_v8
  use gomgpu_iface, only: GOMGPU_STAGE_TRANSPORT
  use gomgpu_bridge
  implicit none
  call gg_elm_boundary()
  call gg_stage(GOMGPU_STAGE_TRANSPORT, 'solve_transport_equation_ELM_v8')
end subroutine solve_transport_equation_ELM_v8

! The equation of state: the stage evaluates the one ana_density selects
! (calculate_density_full.f90 for 0, calculate_density_linear.f90 for 1).
subroutine calculate_density_full
  use gomgpu_iface, only: GOMGPU_STAGE_EOS
  use gomgpu_bridge
  use mod_global_variables, only: it
  implicit none
  if (it == 0) then
    call gg_startup_stage(GOMGPU_STAGE_EOS, 'calculate_density_full')
  else
    call gg_stage(GOMGPU_STAGE_EOS, 'calculate_density_full')
  end if
end subroutine calculate_density_full

subroutine calculate_density_linear
  use gomgpu_iface, only: GOMGPU_STAGE_EOS
  use gomgpu_bridge
  use mod_global_variables, only: it
  implicit none
  if (it == 0) then
    call gg_startup_stage(GOMGPU_STAGE_EOS, 'calculate_density_linear')
  else
    call gg_stage(GOMGPU_STAGE_EOS, 'calculate_density_linear')
  end if
end subroutine calculate_density_linear

subroutine update_variables
  use gomgpu_iface, only: GOMGPU_STAGE_UPDATE
  use gomgpu_bridge
  implicit none
  call gg_stage(GOMGPU_STAGE_UPDATE, 'update_variables')
  call gg_step_end()
end subroutine update_variables
