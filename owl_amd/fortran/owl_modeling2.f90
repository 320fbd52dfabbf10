!
! OWL's 2D forward-modelling program (src/forward/main_forward.f90:30-72 with
! forward_modeling, src/forward/2d/module_modeling.f90:35-242) over the
! drop-in solver modules, one shot at a time on the GPU:
!   owl_modeling2 <parameter file> [--dry-run]
! read_parameters, set_regular_space, load_geometry and prepare_model as OWL
! runs them (owl_caller); then per shot set_adaptive_range and the solver of
! which_medium ('acoustic-iso' or 'elastic-tti') with OWL's components,
! writing dir_synthetic/shot_<id>_seismogram_<p | x, z>.su.  --dry-run
! prints the parsed set-up (windows, live receivers, models) and stops
! before any GPU call.
!
program owl_modeling2

    use owl_caller
    use acoustic_iso_2d
    use elastic_tti_2d

    implicit none

    type(wave_solver_acoustic_iso_2d) :: solver_acoustic_iso
    type(wave_solver_elastic_tti_2d) :: solver_elastic_tti
    character(len=64) :: arg
    logical :: dry_run
    integer :: ishot

    dry_run = .false.
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        dry_run = trim(arg) == '--dry-run'
    end if

    call read_parameters('modeling')
    call load_geometry
    call prepare_model

    if (dry_run) then
        call print_setup
    else
        call run
    end if

contains

    ! the driver proper, after the set-up
    subroutine run

        select case (tidy(which_medium))
            case ('acoustic-iso', 'elastic-tti')
            case default
                call stop_with('Error: which_medium = '//tidy(which_medium)//' is not provided by owl_modeling2')
        end select
        call execute_command_line('mkdir -p '//tidy(dir_synthetic))

        do ishot = 1, ns

            call set_adaptive_range(gmtr(ishot))

            select case (tidy(which_medium))

                case ('acoustic-iso')
                    solver_acoustic_iso%nx = shot_nx
                    solver_acoustic_iso%nz = shot_nz
                    solver_acoustic_iso%dx = dx
                    solver_acoustic_iso%dz = dz
                    solver_acoustic_iso%ox = shot_xbeg
                    solver_acoustic_iso%oz = shot_zbeg
                    solver_acoustic_iso%dt = dt
                    solver_acoustic_iso%tmax = tmax
                    solver_acoustic_iso%data_dt = data_dt
                    solver_acoustic_iso%data_tmax = data_tmax
                    solver_acoustic_iso%dir_synthetic = tidy(dir_synthetic)
                    solver_acoustic_iso%pml = pml
                    solver_acoustic_iso%free_surface = yn_free_surface
                    solver_acoustic_iso%verbose = verbose
                    solver_acoustic_iso%gmtr = gmtr(ishot)
                    solver_acoustic_iso%vp = get_model('vp')
                    solver_acoustic_iso%rho = get_model('rho', 1.0)

                    call solver_acoustic_iso%forward

                case ('elastic-tti')
                    solver_elastic_tti%nx = shot_nx
                    solver_elastic_tti%nz = shot_nz
                    solver_elastic_tti%dx = dx
                    solver_elastic_tti%dz = dz
                    solver_elastic_tti%ox = shot_xbeg
                    solver_elastic_tti%oz = shot_zbeg
                    solver_elastic_tti%dt = dt
                    solver_elastic_tti%tmax = tmax
                    solver_elastic_tti%data_dt = data_dt
                    solver_elastic_tti%data_tmax = data_tmax
                    solver_elastic_tti%dir_synthetic = tidy(dir_synthetic)
                    solver_elastic_tti%pml = pml
                    solver_elastic_tti%free_surface_dz_refine = free_surface_dz_refine
                    solver_elastic_tti%dz_max = dz_max
                    solver_elastic_tti%free_surface = yn_free_surface
                    solver_elastic_tti%file_topo = tidy(file_topo)
                    solver_elastic_tti%topo_interp = tidy(topo_interp)
                    solver_elastic_tti%measure_source_depth_from_surface = measure_source_depth_from_surface
                    solver_elastic_tti%measure_receiver_depth_from_surface = measure_receiver_depth_from_surface
                    solver_elastic_tti%verbose = verbose
                    solver_elastic_tti%gmtr = gmtr(ishot)
                    solver_elastic_tti%compx = yn_compx
                    solver_elastic_tti%compz = yn_compz
                    call set_tti_medium(solver_elastic_tti)

                    call solver_elastic_tti%forward

            end select

            write (*, '(a)') ' Shot '//itoa(gmtr(ishot)%id)//' forward modeling completed.'

        end do

    end subroutine

end program owl_modeling2
