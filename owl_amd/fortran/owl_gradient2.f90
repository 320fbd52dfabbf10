!
! One FWI gradient of OWL's 2D inversion program (compute_gradient_shots and
! output_gradient, src/inverse/2d/module_gradient.f90:59-472) over the drop-in
! solver modules, one shot at a time on the GPU:
!   owl_gradient2 <parameter file> [--dry-run]
! Per shot: set_adaptive_range; forward with reconstruct = .true. into
! dir_working/iteration_1/synthetic; the misfit and adjoint source of
! misfit_type from the synthetic and dir_record/shot_<id>_seismogram_<d>.su
! for every data_name d (compute_adjoint_source through owl_adjoint_source,
! receiver weights and trace_discard_threshold as OWL applies them) into
! dir_working/iteration_1/adjoint_source; adjoint, writing the shot's
! gradients to dir_working/scratch/shot_<id>_grad_<m>.grd; each added into
! the model-size gradient on the shot's window (process_model_single_shot
! without gradient processing).  Then dir_working/iteration_1/model/
! grad_<m>.bin (float32 (nz, nx)) for every model_update name, and the summed
! misfit in dir_working/data_misfit.txt.  Data and gradient processing
! (mutes, filters, masks, smoothing) are not applied.
!
program owl_gradient2

    use iso_c_binding
    use owl_amd_c
    use owl_caller
    use owl_files, only: read_su_gather, write_su_gather, read_grd_2d
    use acoustic_iso_2d
    use elastic_tti_2d

    implicit none

    type(wave_solver_acoustic_iso_2d) :: solver_acoustic_iso
    type(wave_solver_elastic_tti_2d) :: solver_elastic_tti
    type(meta_array2_real), allocatable :: model_grad(:)
    real(c_double), allocatable :: step_misfit(:)
    character(len=1024) :: dir_scratch, dir_adjoint, dir_model, shot_prefix
    character(len=64) :: arg
    logical :: dry_run
    integer :: ishot, i, u, mtype

    dry_run = .false.
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        dry_run = trim(arg) == '--dry-run'
    end if

    call read_parameters('gradient')
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
                call stop_with('Error: which_medium = '//tidy(which_medium)//' is not provided by owl_gradient2')
        end select
        select case (tidy(misfit_type))
            case ('waveform')
                mtype = OWL_MISFIT_WAVEFORM
            case ('corr')
                mtype = OWL_MISFIT_CORR
            case ('envelope')
                mtype = OWL_MISFIT_ENVELOPE
            case ('phase')
                mtype = OWL_MISFIT_PHASE
            case ('adaptive')
                mtype = OWL_MISFIT_ADAPTIVE
            case ('local-adaptive')
                mtype = OWL_MISFIT_LOCAL_ADAPTIVE
            case ('adaptive-spacetime')
                mtype = OWL_MISFIT_ADAPTIVE_SPACETIME
            case default
                call stop_with('Error: misfit_type = '//tidy(misfit_type)//' is not provided by owl_gradient2')
        end select

        dir_scratch = tidy(dir_working)//'/scratch'
        dir_synthetic = tidy(dir_working)//'/iteration_1/synthetic'
        dir_adjoint = tidy(dir_working)//'/iteration_1/adjoint_source'
        dir_model = tidy(dir_working)//'/iteration_1/model'
        call execute_command_line('mkdir -p '//tidy(dir_scratch)//' '//tidy(dir_synthetic)//' '// &
            tidy(dir_adjoint)//' '//tidy(dir_model))

        allocate (model_grad(nmodel), step_misfit(ns))
        do i = 1, nmodel
            model_grad(i)%name = model_name(i)
            allocate (model_grad(i)%array(nz, nx))
            model_grad(i)%array = 0.0
        end do
        step_misfit = 0.0d0

        do ishot = 1, ns

            shot_prefix = tidy(dir_scratch)//'/shot_'//itoa(gmtr(ishot)%id)
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
                    solver_acoustic_iso%dir_working = tidy(dir_scratch)
                    solver_acoustic_iso%pml = pml
                    solver_acoustic_iso%free_surface = yn_free_surface
                    solver_acoustic_iso%verbose = verbose
                    solver_acoustic_iso%gmtr = gmtr(ishot)
                    solver_acoustic_iso%reconstruct = .true.
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
                    solver_elastic_tti%dir_working = tidy(dir_scratch)
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
                    solver_elastic_tti%reconstruct = .true.
                    solver_elastic_tti%compx = yn_compx
                    solver_elastic_tti%compz = yn_compz
                    call set_tti_medium(solver_elastic_tti)
                    call solver_elastic_tti%forward
            end select
            write (*, '(a)') ' Shot '//itoa(gmtr(ishot)%id)//' forward modeling completed. '

            do i = 1, ndata
                call adjoint_source(gmtr(ishot), tidy(data_name(i)), step_misfit(ishot))
            end do
            write (*, '(a,es16.8)') ' >> Shot '//itoa(gmtr(ishot)%id)//' misfit = ', step_misfit(ishot)

            select case (tidy(which_medium))
                case ('acoustic-iso')
                    solver_acoustic_iso%dir_adjoint = tidy(dir_adjoint)
                    solver_acoustic_iso%cc_step_interval = cc_step_interval
                    solver_acoustic_iso%energy_precond = yn_energy_precond
                    call solver_acoustic_iso%adjoint
                case ('elastic-tti')
                    solver_elastic_tti%dir_adjoint = tidy(dir_adjoint)
                    solver_elastic_tti%cc_step_interval = cc_step_interval
                    solver_elastic_tti%energy_precond = yn_energy_precond
                    call solver_elastic_tti%adjoint
            end select
            write (*, '(a)') ' >> Shot '//itoa(gmtr(ishot)%id)//' gradient computation completed. '

            do i = 1, nmodel
                call merge_shot(model_grad(i)%array, tidy(shot_prefix)//'_grad_'//tidy(model_name(i))//'.grd')
            end do

        end do

        do i = 1, nmodel
            open (newunit=u, file=tidy(dir_model)//'/grad_'//tidy(model_name(i))//'.bin', access='stream', &
                form='unformatted', status='replace')
            write (u) model_grad(i)%array
            close (u)
        end do
        open (newunit=u, file=tidy(dir_working)//'/data_misfit.txt', status='replace')
        write (u, '(i0,1x,es24.16)') 1, sum(step_misfit)
        close (u)
        write (*, '(a)') ' >>>>>>>>>> Gradient saved. '

    end subroutine

    ! compute_adjoint_source (module_inversion_adjoint_source.f90:985-1160)
    ! for one data component of one shot
    subroutine adjoint_source(g, d, misfit)
        type(source_receiver_geometry), intent(in) :: g
        character(len=*), intent(in) :: d
        real(c_double), intent(inout) :: misfit
        type(su) :: obs, syn, adj
        real(c_float), allocatable :: o(:), s(:), a(:), w(:)
        real(c_double), allocatable :: m(:)
        character(len=:), allocatable :: name
        integer :: ir, nt, nr
        integer(c_int) :: st
        type(owl_misfit_params_t) :: prm
        name = '/shot_'//itoa(g%id)//'_seismogram_'//d//'.su'
        call read_su_gather(tidy(dir_synthetic)//name, syn)
        call read_su_gather(tidy(dir_record)//name, obs)
        nt = syn%nt
        nr = syn%nr
        if (obs%nt /= nt .or. obs%nr /= nr) &
            call stop_with('Error: '//tidy(dir_record)//name//' does not match the synthetic gather')
        if (nr /= size(g%recr)) call stop_with('Error: the gather does not hold every receiver of the shot')
        allocate (o(nt*nr), s(nt*nr), a(nt*nr), w(nr), m(nr))
        do ir = 1, nr
            o((ir - 1)*nt + 1:ir*nt) = obs%trace(ir)%data(1:nt)
            s((ir - 1)*nt + 1:ir*nt) = syn%trace(ir)%data(1:nt)
            w(ir) = g%recr(ir)%weight
        end do
        ! the adaptive misfit's lags are in the traces' own sampling
        ! (compute_adjoint_source, module_inversion_adjoint_source.f90:1131)
        prm%dt = real(syn%dt, c_float)
        prm%tlag_max = real(tlag_max, c_float)
        select case (tidy(penalty_method))
            case ('power')
                prm%penalty = OWL_PENALTY_POWER
            case ('gaussian')
                prm%penalty = OWL_PENALTY_GAUSSIAN
            case ('exp')
                prm%penalty = OWL_PENALTY_EXP
            case default
                prm%penalty = OWL_PENALTY_LINEAR
        end select
        prm%penalty_power = real(penalty_power, c_float)
        prm%deconv_eps = real(deconv_eps, c_float)
        prm%lawi_sigma = real(lawi_sigma, c_float)
        prm%adaptive_hw = int(adaptive_hw, c_int)
        st = owl_adjoint_source_ex(int(mtype, c_int), prm, int(nr, c_int), int(nt, c_int), o, s, w, &
            real(trace_discard_threshold, c_float), a, m)
        if (st /= OWL_OK) call stop_with('<owl_adjoint_source> '//owl_error_message())
        misfit = misfit + sum(m)
        adj = syn
        do ir = 1, nr
            adj%trace(ir)%data(1:nt) = a((ir - 1)*nt + 1:ir*nt)
        end do
        call write_su_gather(tidy(dir_adjoint)//name, adj)
    end subroutine

    ! the shot's window gradient added into the model-size gradient
    subroutine merge_shot(w, file_w)
        real, intent(inout) :: w(:, :)
        character(len=*), intent(in) :: file_w
        real, allocatable :: gs(:, :)
        allocate (gs(shot_nz, shot_nx))
        call read_grd_2d(file_w, gs)
        w(shot_nzbeg:shot_nzend, shot_nxbeg:shot_nxend) = w(shot_nzbeg:shot_nzend, shot_nxbeg:shot_nxend) + gs
    end subroutine


end program owl_gradient2
