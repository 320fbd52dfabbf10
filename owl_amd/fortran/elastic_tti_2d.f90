!
! Drop-in for OWL's module elastic_tti_2d (src/solver/elastic_tti_2d/
! interface.f90:25-95): type(wave_solver_elastic_tti_2d) with the same
! public components and type-bound forward/adjoint, running the Lebedev
! fully staggered grid on the GPU through the C ABI, one shot (gmtr) on its
! window.  The medium is converted from anisotropy_type 'iso', 'thomsen',
! 'a-t' or 'cij' as prepare_modeling does (module_vars.f90:284-340).
! forward (submodule_forward_modeling.f90:30-300) fills seis_vx, seis_vz and
! writes dir_synthetic/shot_<id>_seismogram_{x,z}.su (compx, compz).
! adjoint (submodule_fwi_gradient.f90:40-560) reads the adjoint sources
! dir_adjoint/shot_<id>_seismogram_{x,z}.su into seis_vxr, seis_vzr and
! writes dir_working/shot_<id>_grad_<m>.grd for the parameters of
! anisotropy_type (c11 .. c55; or vp, vs [, epsilon, delta | eta]) and rho.
! free_surface runs OWL's topography-following free surface on the
! depth-refined mesh (module_vars.f90:360-519): file_topo (two columns, x
! and elevation), topo_interp, free_surface_dz_refine, dz_max (0: 1.5 dz)
! and measure_source/receiver_depth_from_surface as OWL reads them.
! Traces are resampled to data_dt, data_tmax and the adjoint sources back
! to dt, tmax (su%resamp).  Not provided (refused): energy_precond,
! moment-tensor and source-update options, wavefield snapshots.
!
module elastic_tti_2d

    use iso_c_binding
    use owl_amd_c
    use owl_geometry, only: source_receiver_geometry, su
    use owl_files, only: write_su_gather, read_su_gather, write_grd_file, itoa, resamp_gather
    use acoustic_iso_2d, only: to_c_sources, to_c_receivers, check_status, data_samples

    implicit none

    type wave_solver_elastic_tti_2d

        integer :: nx = 0, nz = 0
        real :: dx = 0, dz = 0
        real :: ox = 0, oz = 0
        real :: tmax = 1.0
        real :: dt = 0
        real :: data_tmax = 1.0
        real :: data_dt = 0
        integer :: pml = 15
        logical :: free_surface = .false.
        real, allocatable, dimension(:, :) :: vp, vs, rho
        real, allocatable, dimension(:, :) :: tieps, tidel, tithe, tieta
        real, allocatable, dimension(:, :) :: c11, c13, c15, c33, c35, c55
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
        character(len=12) :: anisotropy_type = 'iso'
        type(su) :: seis_vx, seis_vz, seis_vxr, seis_vzr
        logical :: compx = .true.
        logical :: compz = .true.
        real :: free_surface_dz_refine = 4.0
        character(len=1024) :: file_topo = ''
        logical :: measure_source_depth_from_surface = .false.
        logical :: measure_receiver_depth_from_surface = .false.
        logical :: source_vertical_to_surface = .false.
        logical :: receiver_vertical_to_surface = .false.
        character(len=12) :: topo_interp = 'cubic'
        real :: dz_max = 0
        logical :: save_mesh = .false.
        integer :: nc_mt = 6
        real, allocatable, dimension(:) :: mt
        real, allocatable, dimension(:) :: stf
        logical :: yn_update_medium = .true.
        logical :: yn_update_source = .false.
        integer :: fdhalf = 4
        type(c_ptr), private :: batch = c_null_ptr
        real, allocatable, private :: cw(:, :, :)

    contains

        procedure, public :: forward => compute_forward
        procedure, public :: adjoint => compute_adjoint
        procedure, public :: release => release_batch

    end type

    private
    public :: wave_solver_elastic_tti_2d

contains

    subroutine release_batch(this)
        class(wave_solver_elastic_tti_2d), intent(inout) :: this
        if (c_associated(this%batch)) call owl_fsg2d_destroy(this%batch)
        this%batch = c_null_ptr
    end subroutine

    integer function param_code(this)
        class(wave_solver_elastic_tti_2d), intent(in) :: this
        select case (trim(this%anisotropy_type))
            case ('iso')
                param_code = 0
            case ('thomsen')
                param_code = 1
            case ('a-t')
                param_code = 2
            case ('cij')
                param_code = -1
            case default
                write (*, '(a)') ' <elastic_tti_2d> anisotropy_type '//trim(this%anisotropy_type)//' is not provided'
                error stop 1
        end select
    end function

    subroutine compute_forward(this)

        class(wave_solver_elastic_tti_2d), intent(inout), target :: this

        type(owl_acoustic_config_t) :: cfg
        type(owl_shot_t) :: shot(1)
        type(owl_source_t), allocatable, target :: cs(:)
        type(owl_receiver_t), allocatable, target :: cr(:)
        real, allocatable, target :: rho(:, :), tr(:), p1(:, :), p2(:, :), th(:, :)
        type(c_ptr) :: cp(6)
        type(owl_fsg_surface_t) :: surf
        real(c_float), allocatable, target :: tx(:), tz(:)
        integer :: nt, nr, ir, k, pc, dnt
        real :: ddt
        integer(c_int) :: st

        call this%release()
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
        cfg%cc_step_interval = this%cc_step_interval
        cfg%f0factor = this%gmtr%f0factor

        rho = this%rho
        if (allocated(this%cw)) deallocate (this%cw)
        allocate (this%cw(this%nz, this%nx, 6))
        pc = param_code(this)
        if (pc < 0) then
            this%cw(:, :, 1) = this%c11
            this%cw(:, :, 2) = this%c13
            this%cw(:, :, 3) = this%c15
            this%cw(:, :, 4) = this%c33
            this%cw(:, :, 5) = this%c35
            this%cw(:, :, 6) = this%c55
        else
            allocate (p1(this%nz, this%nx), p2(this%nz, this%nx), th(this%nz, this%nx))
            p1 = 0.0
            p2 = 0.0
            th = 0.0
            if (pc >= 1) then
                p1 = this%tieps
                th = this%tithe
                if (pc == 1) p2 = this%tidel
                if (pc == 2) p2 = this%tieta
            end if
            do k = 1, 6
                cp(k) = c_loc(this%cw(1, 1, k))
            end do
            st = owl_tti2d_cij(int(pc, c_int), int(this%nx, c_long_long)*this%nz, c_loc(this%vp), c_loc(this%vs), &
                c_loc(rho), c_loc(p1), c_loc(p2), c_loc(th), cp)
            call check_status(st, 'owl_tti2d_cij')
        end if
        do k = 1, 6
            cp(k) = c_loc(this%cw(1, 1, k))
        end do

        call to_c_sources(this%gmtr, cs)
        call to_c_receivers(this%gmtr, cr)
        nr = size(cr)
        shot(1)%id = this%gmtr%id
        shot(1)%ns = size(cs)
        shot(1)%sources = c_loc(cs(1))
        shot(1)%nr = nr
        shot(1)%receivers = c_loc(cr(1))
        if (this%free_surface) then
            if (this%file_topo /= '') then
                call load_topography(trim(this%file_topo), tx, tz)
                surf%ntopo = size(tx)
                surf%topo_x = c_loc(tx(1))
                surf%topo_z = c_loc(tz(1))
            end if
            select case (trim(this%topo_interp))
                case ('linear')
                    surf%topo_interp = 0
                case ('cubic')
                    surf%topo_interp = 1
                case default
                    write (*, '(a)') ' <elastic_tti_2d> topo_interp '//trim(this%topo_interp)//' is not provided'
                    error stop 1
            end select
            surf%dz_refine = this%free_surface_dz_refine
            surf%dz_max = max(this%dz_max, 0.0)
            surf%src_depth_from_surface = merge(1_c_int, 0_c_int, this%measure_source_depth_from_surface)
            surf%rec_depth_from_surface = merge(1_c_int, 0_c_int, this%measure_receiver_depth_from_surface)
            st = owl_fsg2d_create_surface(cfg, cp, rho, merge(0_c_int, 1_c_int, pc == 0), 1_c_int, shot, surf, &
                this%batch)
            call check_status(st, 'owl_fsg2d_create_surface')
        else
            st = owl_fsg2d_create(cfg, cp, rho, merge(0_c_int, 1_c_int, pc == 0), 1_c_int, shot, this%batch)
            call check_status(st, 'owl_fsg2d_create')
        end if
        nt = owl_fsg2d_nt(this%batch)
        allocate (tr(2*nt*nr))
        if (this%verbose) write (*, '(a)') ' <elastic_tti_2d> forward modelling of shot '//itoa(this%gmtr%id)
        st = owl_fsg2d_forward(this%batch, tr, merge(1_c_int, 0_c_int, this%reconstruct))
        call check_status(st, 'owl_fsg2d_forward')
        call fill(this%seis_vx, tr(1:nt*nr), nt, nr, this%dt)
        call fill(this%seis_vz, tr(nt*nr + 1:2*nt*nr), nt, nr, this%dt)
        ! resample if necessary (submodule_forward_modeling.f90:286-291)
        call data_samples(this%dt, this%tmax, this%data_dt, this%data_tmax, ddt, dnt)
        call resamp_gather(this%seis_vx, dnt, ddt)
        call resamp_gather(this%seis_vz, dnt, ddt)
        if (this%compx) call write_su_gather(trim(this%dir_synthetic)//'/shot_'//itoa(this%gmtr%id)// &
            '_seismogram_x.su', this%seis_vx)
        if (this%compz) call write_su_gather(trim(this%dir_synthetic)//'/shot_'//itoa(this%gmtr%id)// &
            '_seismogram_z.su', this%seis_vz)
        if (.not. this%reconstruct) call this%release()

    end subroutine

    ! file_topo: x and elevation per line (load(file_topo, ntp, 2, ascii))
    subroutine load_topography(path, tx, tz)
        character(len=*), intent(in) :: path
        real(c_float), allocatable, intent(out) :: tx(:), tz(:)
        integer :: u, n, k, ios
        real :: a, b
        open (newunit=u, file=path, status='old', action='read')
        n = 0
        do
            read (u, *, iostat=ios) a, b
            if (ios /= 0) exit
            n = n + 1
        end do
        if (n < 1) then
            write (*, '(a)') ' <elastic_tti_2d> no topography points in '//path
            error stop 1
        end if
        allocate (tx(n), tz(n))
        rewind (u)
        do k = 1, n
            read (u, *) tx(k), tz(k)
        end do
        close (u)
    end subroutine

    subroutine check_adjoint(g, nt, nr)
        type(su), intent(in) :: g
        integer, intent(in) :: nt, nr
        if (g%nt /= nt .or. g%nr /= nr) then
            write (*, '(a)') ' <elastic_tti_2d> the adjoint source does not match the forward traces'
            error stop 1
        end if
    end subroutine

    subroutine fill(g, tr, nt, nr, dt)
        type(su), intent(inout) :: g
        real, intent(in) :: tr(:)
        integer, intent(in) :: nt, nr
        real, intent(in) :: dt
        integer :: ir
        g%nr = nr
        g%nt = nt
        g%dt = dt
        g%ot = 0.0
        if (allocated(g%trace)) deallocate (g%trace)
        allocate (g%trace(nr))
        do ir = 1, nr
            g%trace(ir)%data = tr((ir - 1)*nt + 1:ir*nt)
        end do
    end subroutine

    subroutine compute_adjoint(this)

        class(wave_solver_elastic_tti_2d), intent(inout), target :: this

        real, allocatable, target :: adj(:), g(:, :, :), gp(:, :, :), p1(:, :), p2(:, :), th(:, :), rho(:, :)
        type(c_ptr) :: gptr(7), cptr(6), pptr(4)
        character(len=8) :: names(4)
        integer :: ir, nt, nr, k, pc, np
        integer(c_int) :: st
        character(len=:), allocatable :: pre

        if (.not. c_associated(this%batch)) then
            write (*, '(a)') ' <elastic_tti_2d> adjoint needs forward with reconstruct = .true.'
            error stop 1
        end if
        if (this%energy_precond) then
            write (*, '(a)') ' <elastic_tti_2d> energy_precond is not provided'
            error stop 1
        end if
        nt = owl_fsg2d_nt(this%batch)
        nr = this%seis_vx%nr
        allocate (adj(2*nt*nr))
        adj = 0.0
        pre = trim(this%dir_adjoint)//'/shot_'//itoa(this%gmtr%id)
        if (this%compx) then
            call read_su_gather(pre//'_seismogram_x.su', this%seis_vxr)
            call resamp_gather(this%seis_vxr, nt, this%dt)
            call check_adjoint(this%seis_vxr, nt, nr)
            do ir = 1, nr
                adj((ir - 1)*nt + 1:ir*nt) = this%seis_vxr%trace(ir)%data
            end do
        end if
        if (this%compz) then
            call read_su_gather(pre//'_seismogram_z.su', this%seis_vzr)
            call resamp_gather(this%seis_vzr, nt, this%dt)
            call check_adjoint(this%seis_vzr, nt, nr)
            do ir = 1, nr
                adj(nt*nr + (ir - 1)*nt + 1:nt*nr + ir*nt) = this%seis_vzr%trace(ir)%data
            end do
        end if
        allocate (g(this%nz, this%nx, 7))
        do k = 1, 7
            gptr(k) = c_loc(g(1, 1, k))
        end do
        if (this%verbose) write (*, '(a)') ' <elastic_tti_2d> adjoint of shot '//itoa(this%gmtr%id)
        st = owl_fsg2d_adjoint(this%batch, adj, gptr)
        call check_status(st, 'owl_fsg2d_adjoint')
        pre = trim(this%dir_working)//'/shot_'//itoa(this%gmtr%id)
        pc = param_code(this)
        if (pc < 0) then
            names(1:4) = ['c11', 'c13', 'c15', 'c33']
            do k = 1, 4
                call write_grd_file(pre//'_grad_'//trim(names(k))//'.grd', g(:, :, k), this%dz, this%dx, this%oz, this%ox)
            end do
            call write_grd_file(pre//'_grad_c35.grd', g(:, :, 5), this%dz, this%dx, this%oz, this%ox)
            call write_grd_file(pre//'_grad_c55.grd', g(:, :, 6), this%dz, this%dx, this%oz, this%ox)
        else
            np = merge(2, 4, pc == 0)
            allocate (gp(this%nz, this%nx, 4), p1(this%nz, this%nx), p2(this%nz, this%nx), th(this%nz, this%nx))
            p1 = 0.0
            p2 = 0.0
            th = 0.0
            if (pc >= 1) then
                p1 = this%tieps
                th = this%tithe
                if (pc == 1) p2 = this%tidel
                if (pc == 2) p2 = this%tieta
            end if
            rho = this%rho
            do k = 1, 6
                cptr(k) = c_loc(g(1, 1, k))
            end do
            do k = 1, 4
                pptr(k) = c_loc(gp(1, 1, k))
            end do
            st = owl_tti2d_param_gradient(int(pc, c_int), int(this%nx, c_long_long)*this%nz, c_loc(this%vp), &
                c_loc(this%vs), c_loc(rho), c_loc(p1), c_loc(p2), c_loc(th), cptr, pptr)
            call check_status(st, 'owl_tti2d_param_gradient')
            names = ['vp      ', 'vs      ', 'epsilon ', 'delta   ']
            if (pc == 2) names(4) = 'eta'
            do k = 1, np
                call write_grd_file(pre//'_grad_'//trim(names(k))//'.grd', gp(:, :, k), this%dz, this%dx, &
                    this%oz, this%ox)
            end do
        end if
        call write_grd_file(pre//'_grad_rho.grd', g(:, :, 7), this%dz, this%dx, this%oz, this%ox)
        call this%release()

    end subroutine

end module elastic_tti_2d
