!
! Drop-in for OWL's module acoustic_iso_3d (src/solver/acoustic_iso_3d/
! interface.f90:25-79): type(wave_solver_acoustic_iso_3d) with the same
! public components and type-bound forward/adjoint on the GPU through the C
! ABI, one shot (gmtr) on its window.  forward (submodule_forward_modeling
! .f90:31-183) fills seis_p and writes dir_synthetic/shot_<id>_seismogram_p.su;
! adjoint (submodule_fwi_gradient.f90:33-274) reads
! dir_adjoint/shot_<id>_seismogram_p.su into seis_pr and writes
! dir_working/shot_<id>_grad_{vp,rho}.grd as (nz, ny, nx) grids.  Traces
! are resampled to data_dt, data_tmax and the adjoint source back to dt, tmax
! (su%resamp).  Not provided: energy_precond (refused), snapshots.
!
module acoustic_iso_3d

    use iso_c_binding
    use owl_amd_c
    use owl_geometry, only: source_receiver_geometry, su
    use owl_files, only: write_su_gather, read_su_gather, write_grd_file, itoa, resamp_gather
    use acoustic_iso_2d, only: to_c_sources, to_c_receivers, check_status, data_samples

    implicit none

    type wave_solver_acoustic_iso_3d

        integer :: nx = 0, ny = 0, nz = 0
        real :: dx = 0, dy = 0, dz = 0
        real :: ox = 0, oy = 0, oz = 0
        real :: tmax = 1.0
        real :: dt = 0
        real :: data_tmax = 1.0
        real :: data_dt = 0
        integer :: pml = 15
        logical :: free_surface = .false.
        real, allocatable, dimension(:, :, :) :: vp, rho
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
        type(su) :: seis_p, seis_pr
        integer :: fdhalf = 4
        type(c_ptr), private :: batch = c_null_ptr

    contains

        procedure, public :: forward => compute_forward
        procedure, public :: adjoint => compute_adjoint
        procedure, public :: release => release_batch

    end type

    private
    public :: wave_solver_acoustic_iso_3d

contains

    subroutine release_batch(this)
        class(wave_solver_acoustic_iso_3d), intent(inout) :: this
        if (c_associated(this%batch)) call owl_acoustic3d_destroy(this%batch)
        this%batch = c_null_ptr
    end subroutine

    subroutine compute_forward(this)

        class(wave_solver_acoustic_iso_3d), intent(inout), target :: this

        type(owl_acoustic_config_t) :: cfg
        type(owl_shot_t) :: shot(1)
        type(owl_source_t), allocatable, target :: cs(:)
        type(owl_receiver_t), allocatable, target :: cr(:)
        real, allocatable :: rho(:, :, :), tr(:)
        integer :: nt, nr, ir, dnt
        real :: ddt
        integer(c_int) :: st

        call this%release()
        cfg%nx = this%nx
        cfg%ny = this%ny
        cfg%nz = this%nz
        cfg%dx = this%dx
        cfg%dy = this%dy
        cfg%dz = this%dz
        cfg%ox = this%ox
        cfg%oy = this%oy
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
            allocate (rho(this%nz, this%ny, this%nx))
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

        st = owl_acoustic3d_create(cfg, this%vp, rho, 1_c_int, shot, this%batch)
        call check_status(st, 'owl_acoustic3d_create')
        nt = owl_acoustic3d_nt(this%batch)
        allocate (tr(nt*nr))
        if (this%verbose) write (*, '(a)') ' <acoustic_iso_3d> forward modelling of shot '//itoa(this%gmtr%id)
        st = owl_acoustic3d_forward(this%batch, tr, merge(1_c_int, 0_c_int, this%reconstruct))
        call check_status(st, 'owl_acoustic3d_forward')
        this%seis_p%nr = nr
        this%seis_p%nt = nt
        this%seis_p%dt = this%dt
        this%seis_p%ot = 0.0
        if (allocated(this%seis_p%trace)) deallocate (this%seis_p%trace)
        allocate (this%seis_p%trace(nr))
        do ir = 1, nr
            this%seis_p%trace(ir)%data = tr((ir - 1)*nt + 1:ir*nt)
        end do
        ! resample if necessary (submodule_forward_modeling.f90:173)
        call data_samples(this%dt, this%tmax, this%data_dt, this%data_tmax, ddt, dnt)
        call resamp_gather(this%seis_p, dnt, ddt)
        call write_su_gather(trim(this%dir_synthetic)//'/shot_'//itoa(this%gmtr%id)//'_seismogram_p.su', &
            this%seis_p)
        if (.not. this%reconstruct) call this%release()

    end subroutine

    subroutine compute_adjoint(this)

        class(wave_solver_acoustic_iso_3d), intent(inout) :: this

        real, allocatable :: adj(:), gv(:), gr(:)
        integer :: ir, nt
        integer(c_int) :: st

        if (.not. c_associated(this%batch)) then
            write (*, '(a)') ' <acoustic_iso_3d> adjoint needs forward with reconstruct = .true.'
            error stop 1
        end if
        if (this%energy_precond) then
            write (*, '(a)') ' <acoustic_iso_3d> energy_precond is not provided in 3D'
            error stop 1
        end if
        call read_su_gather(trim(this%dir_adjoint)//'/shot_'//itoa(this%gmtr%id)//'_seismogram_p.su', &
            this%seis_pr)
        nt = owl_acoustic3d_nt(this%batch)
        ! back on the solver's time axis (submodule_fwi_gradient.f90:51)
        call resamp_gather(this%seis_pr, nt, this%dt)
        if (this%seis_pr%nt /= nt .or. this%seis_pr%nr /= this%seis_p%nr) then
            write (*, '(a)') ' <acoustic_iso_3d> the adjoint source does not match the forward traces'
            error stop 1
        end if
        allocate (adj(nt*this%seis_pr%nr))
        do ir = 1, this%seis_pr%nr
            adj((ir - 1)*nt + 1:ir*nt) = this%seis_pr%trace(ir)%data
        end do
        allocate (gv(this%nx*this%ny*this%nz), gr(this%nx*this%ny*this%nz))
        st = owl_acoustic3d_adjoint(this%batch, adj, gv, gr)
        call check_status(st, 'owl_acoustic3d_adjoint')
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%gmtr%id)//'_grad_vp.grd', &
            reshape(gv, [this%nz, this%ny, this%nx]), this%dz, this%dy, this%dx, this%oz, this%oy, this%ox)
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%gmtr%id)//'_grad_rho.grd', &
            reshape(gr, [this%nz, this%ny, this%nx]), this%dz, this%dy, this%dx, this%oz, this%oy, this%ox)
        call this%release()

    end subroutine

end module acoustic_iso_3d
