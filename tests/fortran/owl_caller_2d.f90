!
! OWL's two callers of the 2D acoustic solver, restated line for line over
! the drop-in module acoustic_iso_2d (owl_amd/fortran):
!   mode forward:  src/forward/2d/module_modeling.f90:66-95 (forward_modeling)
!   mode gradient: src/inverse/2d/module_gradient.f90:114-135 (forward with
!                  reconstruct) and :342-346 (adjoint), with the adjoint source
!                  files already in dir_adjoint.
! The state those lines read (model, geometry, shot window, parameters) comes
! from a job file through caller_env, which stands in for OWL's
! mod_parameters, mod_model and mod_utility.
!   owl_caller_2d <forward|gradient> <job.txt>
! job.txt: nx nz dx dz dt tmax pml
!          vp_file                      (float32, (nz, nx) column-major)
!          rho_file
!          yn_adpx adp_extrax
!          nshot
!          id sx sz f0 nr               (per shot), then nr lines "rx rz"
!          dir_synthetic
!          dir_scratch
!          dir_adjoint
!
module caller_env

    use iso_c_binding
    use owl_amd_c
    use owl_geometry

    implicit none

    integer :: nx, nz, pml, nshot
    real :: dx, dz, ox = 0.0, oz = 0.0, dt, tmax, data_dt, data_tmax
    logical :: yn_free_surface = .false., verbose = .false., yn_reconstruct = .true.
    logical :: yn_energy_precond = .false.
    integer :: cc_step_interval = 1
    real :: snaps(3) = 0.0
    character(len=1024) :: dir_synthetic, dir_scratch, dir_adjoint, dir_snapshot = './'
    type(source_receiver_geometry), allocatable :: gmtr(:)
    real, allocatable :: model_vp(:, :), model_rho(:, :)
    integer :: yn_adpx = 0
    real :: adp_extrax = 0.0
    ! the shot window of mod_utility (set_adaptive_range)
    integer :: shot_nx, shot_nz, shot_nxbeg, shot_nxend, shot_nzbeg, shot_nzend
    real :: shot_xbeg, shot_zbeg

contains

    ! set_adaptive_range (src/lib/module_utility.f90:410-481) through the C ABI
    subroutine set_adaptive_range(g)
        type(source_receiver_geometry), intent(inout) :: g
        type(owl_window_params_t) :: p
        type(owl_source_t) :: cs(size(g%srcr))
        type(owl_receiver_t) :: cr(size(g%recr))
        integer(c_int) :: win(6), st
        real(c_float) :: rng(6)
        integer :: k
        p%nx = nx
        p%nz = nz
        p%dx = dx
        p%dz = dz
        p%ox = ox
        p%oz = oz
        p%xmin = ox
        p%xmax = ox + (nx - 1)*dx
        p%zmin = oz
        p%zmax = oz + (nz - 1)*dz
        p%yn_adpx = yn_adpx
        p%adp_extrax = adp_extrax
        do k = 1, size(g%srcr)
            cs(k)%x = g%srcr(k)%x
            cs(k)%z = g%srcr(k)%z
        end do
        do k = 1, size(g%recr)
            cr(k)%x = g%recr(k)%x
            cr(k)%z = g%recr(k)%z
            cr(k)%weight = g%recr(k)%weight
        end do
        st = owl_set_adaptive_range(p, size(cs), cs, size(cr), cr, win, rng)
        if (st /= OWL_OK) error stop 'set_adaptive_range failed'
        do k = 1, size(g%recr)
            g%recr(k)%weight = cr(k)%weight
        end do
        shot_nxbeg = win(1)
        shot_nxend = win(2)
        shot_nzbeg = win(5)
        shot_nzend = win(6)
        shot_nx = shot_nxend - shot_nxbeg + 1
        shot_nz = shot_nzend - shot_nzbeg + 1
        shot_xbeg = rng(1)
        shot_zbeg = rng(5)
    end subroutine

    ! get_model (src/lib/module_model.f90): the model on the current window
    function get_model(name, default) result(m)
        character(len=*), intent(in) :: name
        real, intent(in), optional :: default
        real, allocatable :: m(:, :)
        select case (name)
            case ('vp')
                m = model_vp(shot_nzbeg:shot_nzend, shot_nxbeg:shot_nxend)
            case ('rho')
                m = model_rho(shot_nzbeg:shot_nzend, shot_nxbeg:shot_nxend)
            case default
                allocate (m(shot_nz, shot_nx))
                m = default
        end select
    end function

    function tidy(s) result(t)
        character(len=*), intent(in) :: s
        character(len=:), allocatable :: t
        t = trim(adjustl(s))
    end function

    function regspace(a, b, c) result(r)
        real, intent(in) :: a, b, c
        real, allocatable :: r(:)
        integer :: n, k
        n = int((b - a)/c) + 1
        allocate (r(n))
        do k = 1, n
            r(k) = a + (k - 1)*c
        end do
    end function

    subroutine read_job(path)
        character(len=*), intent(in) :: path
        character(len=1024) :: f_vp, f_rho
        integer :: u, v, k, ir, id, nr
        real :: sx, sz, f0
        open (newunit=u, file=path, status='old', action='read')
        read (u, *) nx, nz, dx, dz, dt, tmax, pml
        read (u, '(a)') f_vp
        read (u, '(a)') f_rho
        read (u, *) yn_adpx, adp_extrax
        read (u, *) nshot
        allocate (gmtr(nshot))
        do k = 1, nshot
            read (u, *) id, sx, sz, f0, nr
            gmtr(k)%id = id
            gmtr(k)%ns = 1
            allocate (gmtr(k)%srcr(1), gmtr(k)%recr(nr))
            gmtr(k)%nr = nr
            gmtr(k)%srcr(1)%x = sx
            gmtr(k)%srcr(1)%z = sz
            gmtr(k)%srcr(1)%f0 = f0
            do ir = 1, nr
                read (u, *) gmtr(k)%recr(ir)%x, gmtr(k)%recr(ir)%z
            end do
        end do
        read (u, '(a)') dir_synthetic
        read (u, '(a)') dir_scratch
        read (u, '(a)') dir_adjoint
        close (u)
        data_dt = dt
        data_tmax = tmax
        allocate (model_vp(nz, nx), model_rho(nz, nx))
        open (newunit=v, file=trim(f_vp), access='stream', form='unformatted', status='old')
        read (v) model_vp
        close (v)
        open (newunit=v, file=trim(f_rho), access='stream', form='unformatted', status='old')
        read (v) model_rho
        close (v)
    end subroutine

end module caller_env

program owl_caller_2d

    use caller_env
    use acoustic_iso_2d

    implicit none

    type(wave_solver_acoustic_iso_2d) :: solver_acoustic_iso
    character(len=1024) :: mode, job
    integer :: ishot

    call get_command_argument(1, mode)
    call get_command_argument(2, job)
    call read_job(trim(job))

    if (trim(mode) == 'forward') then
        do ishot = 1, nshot
            call set_adaptive_range(gmtr(ishot))
            ! ---- src/forward/2d/module_modeling.f90:74-95
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
            if (sum(snaps) > 0) then
                solver_acoustic_iso%dir_snapshot = tidy(dir_snapshot)
                solver_acoustic_iso%snaps = regspace(snaps(1), snaps(2), snaps(3))
            end if
            solver_acoustic_iso%pml = pml
            solver_acoustic_iso%free_surface = yn_free_surface
            solver_acoustic_iso%verbose = verbose
            solver_acoustic_iso%gmtr = gmtr(ishot)
            solver_acoustic_iso%vp = get_model('vp')
            solver_acoustic_iso%rho = get_model('rho', 1.0)

            call solver_acoustic_iso%forward
        end do
    else
        do ishot = 1, nshot
            call set_adaptive_range(gmtr(ishot))
            ! ---- src/inverse/2d/module_gradient.f90:116-135
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
            solver_acoustic_iso%reconstruct = yn_reconstruct
            solver_acoustic_iso%vp = get_model('vp')
            solver_acoustic_iso%rho = get_model('rho', 1.0)

            call solver_acoustic_iso%forward

            ! ---- src/inverse/2d/module_gradient.f90:342-346
            solver_acoustic_iso%dir_adjoint = tidy(dir_adjoint)
            solver_acoustic_iso%cc_step_interval = cc_step_interval
            solver_acoustic_iso%energy_precond = yn_energy_precond
            call solver_acoustic_iso%adjoint
        end do
    end if
    write (*, '(a,i0,a)') 'owl_caller_2d: ', nshot, ' shots done'

end program owl_caller_2d
