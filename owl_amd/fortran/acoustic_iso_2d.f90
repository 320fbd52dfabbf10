!
! Drop-in for OWL's module acoustic_iso_2d (src/solver/acoustic_iso_2d/
! interface.f90:25-72): type(wave_solver_acoustic_iso_2d) with the same
! public components and type-bound forward/adjoint, running on the GPU
! through the C ABI (owl_amd_c).  A caller fills it exactly as OWL's callers
! do (src/forward/2d/module_modeling.f90:72-95, src/inverse/2d/
! module_gradient.f90:114-135, 342-346): one shot (gmtr) on its window
! (nx, nz, ox, oz, vp, rho of the window).
!
! forward (compute_forward, submodule_forward_modeling.f90:31-132) fills
! seis_p and writes dir_synthetic/shot_<id>_seismogram_p.su; with
! reconstruct = .true. it keeps the boundary records on the GPU for adjoint.
! adjoint (compute_adjoint, submodule_fwi_gradient.f90:31-199) reads
! dir_adjoint/shot_<id>_seismogram_p.su into seis_pr and writes
! dir_working/shot_<id>_grad_{vp,rho}.grd, with energy_precond applied as
! OWL does.  Traces are resampled to data_dt, data_tmax after the forward
! and the adjoint source back to dt, tmax (su%resamp, owl_resample_traces).
! Not provided: wavefield snapshots (snaps is accepted and reported).
!
module acoustic_iso_2d

    use iso_c_binding
    use owl_amd_c
    use owl_geometry, only: source_receiver_geometry, su
    use owl_files, only: write_su_gather, read_su_gather, write_grd_file, itoa, resamp_gather

    implicit none

    type wave_solver_acoustic_iso_2d

        integer :: nx = 0, nz = 0
        real :: dx = 0, dz = 0
        real :: ox = 0, oz = 0
        real :: tmax = 1.0
        real :: dt = 0
        real :: data_tmax = 1.0
        real :: data_dt = 0
        integer :: pml = 15
        logical :: free_surface = .false.
        real, allocatable, dimension(:, :) :: vp, rho
        character(len=1024) :: dir_synthetic = './'
        character(len=1024) :: dir_adjoint = './'
        character(len=1024) :: dir_snapshot = './'
        character(len=1024) :: dir_working = './'
        integer :: cc_step_interval = 1
        logical :: verbose = .false.
        real, allocatable, dimension(:) :: snaps
        type(source_receiver_geometry) :: gmtr
        logical :: energy_precond = .false.
        logical :: reconstruct = .false.
        type(su) :: seis_p, seis_pr, seis_pa
        ! half-width R of the stencil (OWL fixes it at build time, -D_fdorder<2R>_)
        integer :: fdhalf = 4
        type(c_ptr), private :: batch = c_null_ptr

    contains

        procedure, public :: forward => compute_forward
        procedure, public :: adjoint => compute_adjoint
        procedure, public :: release => release_batch

    end type

    private
    public :: wave_solver_acoustic_iso_2d
    public :: to_c_sources, to_c_receivers, check_status, data_samples

contains

    ! data_nt = nint(data_tmax/data_dt + 1) (module_vars.f90:162-164); a
    ! data_dt left at 0 means dt
    subroutine data_samples(dt, tmax, data_dt, data_tmax, ddt, dnt)
        real, intent(in) :: dt, tmax, data_dt, data_tmax
        real, intent(out) :: ddt
        integer, intent(out) :: dnt
        if (data_dt > 0) then
            ddt = data_dt
            dnt = nint(data_tmax/data_dt + 1)
        else
            ddt = dt
            dnt = nint(tmax/dt + 1)
        end if
    end subroutine

    subroutine check_status(st, what)
        integer(c_int), intent(in) :: st
        character(len=*), intent(in) :: what
        if (st /= OWL_OK) then
            write (*, '(a)') ' <owl_amd> '//what//' failed: '//owl_error_message()
            error stop 1
        end if
    end subroutine

    ! OWL's source and receiver records as the C ABI's structs
    ! (mechanism and wavelet names of module_source_receiver.f90:76-80,
    ! 1036-1061)
    subroutine to_c_sources(g, out)
        type(source_receiver_geometry), intent(in) :: g
        type(owl_source_t), allocatable, intent(out) :: out(:)
        integer :: k
        allocate (out(size(g%srcr)))
        do k = 1, size(g%srcr)
            out(k)%x = g%srcr(k)%x
            out(k)%y = g%srcr(k)%y
            out(k)%z = g%srcr(k)%z
            out(k)%f0 = g%srcr(k)%f0
            out(k)%amp = g%srcr(k)%amp
            out(k)%t0 = g%srcr(k)%t0
            out(k)%polar = g%srcr(k)%polar
            out(k)%azimuth = g%srcr(k)%azimuth
            select case (trim(g%srcr(k)%mechanism))
                case ('explosion')
                    out(k)%mechanism = OWL_EXPLOSION
                case ('force')
                    out(k)%mechanism = OWL_FORCE
                case default
                    write (*, '(a)') ' <owl_amd> source mechanism '//trim(g%srcr(k)%mechanism)// &
                        ' is not provided'
                    error stop 1
            end select
            select case (trim(g%srcr(k)%wavelet))
                case ('ricker', 'gaussian_deriv_deriv', 'gaussian_deriv2')
                    out(k)%wavelet = 0
                case ('gaussian')
                    out(k)%wavelet = 1
                case ('gaussian_deriv', 'gaussian_deriv1')
                    out(k)%wavelet = 2
                case ('ricker_deriv', 'ricker_deriv1', 'gaussian_deriv3', 'gaussian_deriv_deriv_deriv')
                    out(k)%wavelet = 3
                case default
                    write (*, '(a)') ' <owl_amd> wavelet '//trim(g%srcr(k)%wavelet)//' is not provided'
                    error stop 1
            end select
        end do
    end subroutine

    subroutine to_c_receivers(g, out)
        type(source_receiver_geometry), intent(in) :: g
        type(owl_receiver_t), allocatable, intent(out) :: out(:)
        integer :: k
        allocate (out(size(g%recr)))
        do k = 1, size(g%recr)
            out(k)%x = g%recr(k)%x
            out(k)%y = g%recr(k)%y
            out(k)%z = g%recr(k)%z
            out(k)%weight = merge(g%recr(k)%weight, 0.0, g%recr(k)%valid)
        end do
    end subroutine

    subroutine release_batch(this)
        class(wave_solver_acoustic_iso_2d), intent(inout) :: this
        if (c_associated(this%batch)) call owl_acoustic2d_destroy(this%batch)
        this%batch = c_null_ptr
    end subroutine

    subroutine compute_forward(this)

        class(wave_solver_acoustic_iso_2d), intent(inout), target :: this

        type(owl_acoustic_config_t) :: cfg
        type(owl_shot_t) :: shot(1)
        type(owl_source_t), allocatable, target :: cs(:)
        type(owl_receiver_t), allocatable, target :: cr(:)
        real, allocatable :: rho(:, :), tr(:)
        integer :: nt, nr, ir, dnt
        real :: ddt
        integer(c_int) :: st

        call this%release()
        if (allocated(this%snaps)) then
            if (size(this%snaps) > 0 .and. this%verbose) &
                write (*, '(a)') ' <acoustic_iso_2d> wavefield snapshots are not written'
        end if
        cfg%nx = this%nx
        cfg%nz = this%nz
        cfg%dx = this%dx
        cfg%dz = this%dz
        cfg%ox = this%ox
        cfg%oz = this%oz
        cfg%dt = this%dt
        cfg%tmax = this%tmax
        cfg%data_dt = this%dt
        cfg%data_tmax = this%tmax
        cfg%pml = this%pml
        cfg%fdhalf = this%fdhalf
        cfg%free_surface = merge(1, 0, this%free_surface)
        cfg%cc_step_interval = this%cc_step_interval
        cfg%f0factor = this%gmtr%f0factor

        if (allocated(this%rho)) then
            rho = this%rho
        else
            allocate (rho(this%nz, this%nx))
            rho = 1.0
        end if

        call to_c_sources(this%gmtr, cs)
        call to_c_receivers(this%gmtr, cr)
        nr = size(cr)
        shot(1)%id = this%gmtr%id
        shot(1)%ns = size(cs)
        shot(1)%sources = c_loc(cs(1))
        shot(1)%nr = nr
        shot(1)%receivers = c_loc(cr(1))

        st = owl_acoustic2d_create(cfg, this%vp, rho, 1_c_int, shot, this%batch)
        call check_status(st, 'owl_acoustic2d_create')
        st = owl_acoustic2d_set_energy_precond(this%batch, merge(1_c_int, 0_c_int, this%energy_precond))
        call check_status(st, 'owl_acoustic2d_set_energy_precond')
        nt = owl_acoustic2d_nt(this%batch)
        allocate (tr(nt*nr))
        if (this%verbose) write (*, '(a)') ' <acoustic_iso_2d> forward modelling of shot '//itoa(this%gmtr%id)
        st = owl_acoustic2d_forward(this%batch, tr, merge(1_c_int, 0_c_int, this%reconstruct))
        call check_status(st, 'owl_acoustic2d_forward')

        ! receiver-major [nr][nt] traces into the gather
        this%seis_p%nr = nr
        this%seis_p%nt = nt
        this%seis_p%dt = this%dt
        this%seis_p%ot = 0.0
        if (allocated(this%seis_p%trace)) deallocate (this%seis_p%trace)
        allocate (this%seis_p%trace(nr))
        do ir = 1, nr
            this%seis_p%trace(ir)%data = tr((ir - 1)*nt + 1:ir*nt)
        end do
        ! resample if necessary (submodule_forward_modeling.f90:124-125)
        call data_samples(this%dt, this%tmax, this%data_dt, this%data_tmax, ddt, dnt)
        call resamp_gather(this%seis_p, dnt, ddt)
        call write_su_gather(trim(this%dir_synthetic)//'/shot_'//itoa(this%gmtr%id)//'_seismogram_p.su', &
            this%seis_p)

        if (.not. this%reconstruct) call this%release()

    end subroutine

    subroutine compute_adjoint(this)

        class(wave_solver_acoustic_iso_2d), intent(inout) :: this

        real, allocatable :: adj(:), gv(:), gr(:)
        integer :: ir, nt
        integer(c_int) :: st

        if (.not. c_associated(this%batch)) then
            write (*, '(a)') ' <acoustic_iso_2d> adjoint needs forward with reconstruct = .true.'
            error stop 1
        end if
        call read_su_gather(trim(this%dir_adjoint)//'/shot_'//itoa(this%gmtr%id)//'_seismogram_p.su', &
            this%seis_pr)
        nt = owl_acoustic2d_nt(this%batch)
        ! back on the solver's time axis (submodule_fwi_gradient.f90:46)
        call resamp_gather(this%seis_pr, nt, this%dt)
        if (this%seis_pr%nt /= nt .or. this%seis_pr%nr /= this%seis_p%nr) then
            write (*, '(a)') ' <acoustic_iso_2d> the adjoint source does not match the forward traces'
            error stop 1
        end if
        allocate (adj(nt*this%seis_pr%nr))
        do ir = 1, this%seis_pr%nr
            adj((ir - 1)*nt + 1:ir*nt) = this%seis_pr%trace(ir)%data
        end do
        st = owl_acoustic2d_set_energy_precond(this%batch, merge(1_c_int, 0_c_int, this%energy_precond))
        call check_status(st, 'owl_acoustic2d_set_energy_precond')
        allocate (gv(this%nx*this%nz), gr(this%nx*this%nz))
        if (this%verbose) write (*, '(a)') ' <acoustic_iso_2d> adjoint of shot '//itoa(this%gmtr%id)
        st = owl_acoustic2d_adjoint(this%batch, adj, gv, gr)
        call check_status(st, 'owl_acoustic2d_adjoint')
        ! (nz, nx) column-major, the .grd array
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%gmtr%id)//'_grad_vp.grd', &
            reshape(gv, [this%nz, this%nx]), this%dz, this%dx, this%oz, this%ox)
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%gmtr%id)//'_grad_rho.grd', &
            reshape(gr, [this%nz, this%nx]), this%dz, this%dx, this%oz, this%ox)
        call this%release()

    end subroutine

end module acoustic_iso_2d
