!
! OWL's 2D driver state and set-up routines, for the owl_modeling2 and
! owl_gradient2 programs over the drop-in solver modules of this directory:
!   read_parameters     src/lib/module_parameters.f90:316-703 (the keys the
!                       2D modelling and gradient drivers read; data
!                       processing, regularisation and MPI keys are left out)
!   set_regular_space   src/lib/module_model.f90:72-160
!   load_geometry       src/lib/module_source_receiver.f90:183-551 (2D)
!   prepare_model       src/forward/2d/module_modeling.f90:35-47 with
!                       prepare_model_single_parameter (module_model.f90:
!                       199-334): file_<name> as float32 (nz, nx)
!   get_model           module_model.f90:163-197
!   set_adaptive_range  src/lib/module_utility.f90:410-510 (owl_set_adaptive_range)
! Names of the variables follow OWL's modules, so the drivers read like
! OWL's.  Rules where FLIT (not part of the OWL tree) decides:
!   * the model is resampled to the target grid (nnx, ddx, xmin ...) by
!     linear interpolation along z, then along x, with the end values held
!     outside the original grid (FLIT's interp 'linear' taken so);
!   * topography for the free-surface checks goes through owl_ginterp
!     (FLIT's ginterp 'linear' or 'cubic', as the free-surface mesh uses).
!
module owl_caller

    use iso_c_binding
    use owl_amd_c
    use owl_geometry
    use owl_params
    use elastic_tti_2d, only: wave_solver_elastic_tti_2d

    implicit none

    real, parameter :: float_huge = huge(1.0)
    integer, parameter :: int4_huge = huge(1)
    real, parameter :: const_deg2rad = 0.017453292519943295

    character(len=1024) :: file_parameter = ''

    ! module_model.f90:72-160
    integer :: nx0 = 0, ny0 = 1, nz0 = 0, nx = 0, ny = 1, nz = 0
    real :: dx0 = 1, dy0 = 1, dz0 = 1, ox0 = 0, oy0 = 0, oz0 = 0
    real :: dx = 1, dy = 1, dz = 1, ox = 0, oy = 0, oz = 0
    real :: xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0
    logical :: require_model_interp = .false.
    integer :: pml = 15
    real :: dz_max = 0

    ! module_parameters.f90:316-703
    real :: dt = 0, tmax = 1, data_dt = 0, data_tmax = 1
    integer :: cc_step_interval = 1
    character(len=64) :: which_medium = 'acoustic-iso', aniso_param = 'iso'
    integer :: ns = 1
    character(len=1024) :: file_geometry = ''
    integer :: shot_min = 1, shot_every = 1, shot_max = 1
    integer :: rec_min = 1, rec_every = 1, rec_max = int4_huge
    integer, allocatable :: shot_index(:), rec_index(:), sid_select(:), src_select(:), sid_exclude(:), src_exclude(:)
    real :: offset_min = 0, offset_max = float_huge
    real :: sxmin, sxmax, symin, symax, szmin, szmax, rxmin, rxmax, rymin, rymax, rzmin, rzmax
    integer :: sid_min = 0, sid_max = int4_huge
    real :: f0_factor = 1.0
    real, allocatable :: src_filt_freqs(:), src_filt_coefs(:)
    character(len=1024) :: dir_working = './test', dir_record = './data', dir_synthetic = './data_synthetic'
    character(len=1024) :: dir_snapshot = './snapshot'
    real, allocatable :: snaps(:)
    logical :: yn_adpx = .false., yn_adpz = .false.
    real :: adp_extrax = 0, adp_extraz = 0
    logical :: verbose = .false.
    logical :: yn_free_surface = .false.
    real :: free_surface_dz_refine = 4.0
    character(len=1024) :: file_topo = ''
    character(len=32) :: topo_interp = 'cubic'
    logical :: measure_source_depth_from_surface = .false., measure_receiver_depth_from_surface = .false.
    character(len=64), allocatable :: model_name(:), model_name_aux(:), data_name(:)
    integer :: nmodel = 0, nmodel_aux = 0, ndata = 0
    logical :: yn_compx = .true., yn_compz = .true.
    character(len=64) :: misfit_type = 'waveform'
    real :: trace_discard_threshold = 1.0e-6
    ! the adaptive misfit's parameters (module_parameters.f90:262-265, 551-557)
    real :: tlag_max = 0.1, penalty_power = 1.0, deconv_eps = 0.1, lawi_sigma = 0.25
    character(len=32) :: penalty_method = 'linear'
    integer :: adaptive_hw = 3
    logical :: yn_energy_precond = .false.
    integer :: niter_max = 1

    ! the models of model_name on the target grid
    type meta_array2_real
        character(len=64) :: name = ''
        real, allocatable :: array(:, :)
    end type
    type(meta_array2_real), allocatable :: model_m(:)

    type(source_receiver_geometry), allocatable :: gmtr(:)

    ! the current shot's window (mod_utility)
    integer :: shot_nx = 0, shot_nz = 0, shot_nxbeg = 1, shot_nxend = 1, shot_nzbeg = 1, shot_nzend = 1
    real :: shot_xbeg = 0, shot_zbeg = 0

contains

    function tidy(s) result(t)
        character(len=*), intent(in) :: s
        character(len=:), allocatable :: t
        t = trim(adjustl(s))
    end function

    subroutine stop_with(msg)
        character(len=*), intent(in) :: msg
        write (*, '(a)') ' '//msg
        error stop 1
    end subroutine

    ! program = 'modeling' (OWL's _forward_ build) or 'gradient' (_fwi_)
    subroutine read_parameters(program)
        character(len=*), intent(in) :: program
        logical :: exists
        if (command_argument_count() == 0) call stop_with('Error: The program needs a parameter file. Exiting.')
        call get_command_argument(1, file_parameter)
        inquire (file=tidy(file_parameter), exist=exists)
        if (.not. exists) call stop_with('Error: Parameter file = '//tidy(file_parameter)//' not found. Exiting.')

        call readpar_float(file_parameter, 'dt', dt, 0.0)
        call readpar_float(file_parameter, 'tmax', tmax, 1.0)
        call readpar_float(file_parameter, 'data_dt', data_dt, dt)
        call readpar_float(file_parameter, 'data_tmax', data_tmax, tmax)
        call readpar_int(file_parameter, 'cc_step_interval', cc_step_interval, 1)
        call readpar_string(file_parameter, 'which_medium', which_medium, 'acoustic-iso')
        call readpar_string(file_parameter, 'anisotropy_type', aniso_param, 'iso')

        call readpar_int(file_parameter, 'ns', ns, 1)
        call readpar_string(file_parameter, 'file_geometry', file_geometry, '', required=.true.)
        call readpar_nint(file_parameter, 'src_index', shot_index, [1, 1, ns])
        if (size(shot_index) /= 3) call stop_with('<read_parameter> Error: src_index needs three values')
        shot_min = max(shot_index(1), 1)
        shot_every = max(shot_index(2), 1)
        shot_max = min(shot_index(3), ns)
        call readpar_nint(file_parameter, 'rec_index', rec_index, [1, 1, int4_huge])
        if (size(rec_index) /= 3) call stop_with('<read_parameter> Error: rec_index needs three values')
        rec_min = max(rec_index(1), 1)
        rec_every = max(rec_index(2), 1)
        rec_max = min(rec_index(3), int4_huge)
        call readpar_float(file_parameter, 'offset_min', offset_min, 0.0)
        call readpar_float(file_parameter, 'offset_max', offset_max, float_huge)
        call readpar_nint(file_parameter, 'sid_select', sid_select, [0])
        call readpar_nint(file_parameter, 'src_select', src_select, [0])
        call readpar_nint(file_parameter, 'sid_exclude', sid_exclude, [0])
        call readpar_nint(file_parameter, 'src_exclude', src_exclude, [0])
        call readpar_int(file_parameter, 'sid_min', sid_min, 0)
        call readpar_int(file_parameter, 'sid_max', sid_max, int4_huge)
        call readpar_float(file_parameter, 'sx_min', sxmin, -float_huge)
        call readpar_float(file_parameter, 'sx_max', sxmax, +float_huge)
        call readpar_float(file_parameter, 'sy_min', symin, -float_huge)
        call readpar_float(file_parameter, 'sy_max', symax, +float_huge)
        call readpar_float(file_parameter, 'sz_min', szmin, -float_huge)
        call readpar_float(file_parameter, 'sz_max', szmax, +float_huge)
        call readpar_float(file_parameter, 'rx_min', rxmin, -float_huge)
        call readpar_float(file_parameter, 'rx_max', rxmax, +float_huge)
        call readpar_float(file_parameter, 'ry_min', rymin, -float_huge)
        call readpar_float(file_parameter, 'ry_max', rymax, +float_huge)
        call readpar_float(file_parameter, 'rz_min', rzmin, -float_huge)
        call readpar_float(file_parameter, 'rz_max', rzmax, +float_huge)
        call readpar_float(file_parameter, 'f0_factor', f0_factor, 1.0)
        call readpar_nfloat(file_parameter, 'src_filt_freqs', src_filt_freqs, [-1.0])
        call readpar_nfloat(file_parameter, 'src_filt_coefs', src_filt_coefs, [-1.0])
        if (size(src_filt_freqs) /= size(src_filt_coefs)) &
            call stop_with('Error: size(src_filt_freqs) must = size(src_filt_coefs)')

        call readpar_string(file_parameter, 'dir_working', dir_working, './test')
        call readpar_string(file_parameter, 'dir_record', dir_record, './data')
        call readpar_string(file_parameter, 'dir_synthetic', dir_synthetic, './data_synthetic')
        call readpar_string(file_parameter, 'dir_snapshot', dir_snapshot, './snapshot')
        call readpar_nfloat(file_parameter, 'snaps', snaps, [-1.0])
        call readpar_logical(file_parameter, 'yn_adpx', yn_adpx, .false.)
        call readpar_logical(file_parameter, 'yn_adpz', yn_adpz, .false.)
        call readpar_float(file_parameter, 'adp_extrax', adp_extrax, 0.0)
        call readpar_float(file_parameter, 'adp_extraz', adp_extraz, 0.0)
        call readpar_logical(file_parameter, 'verbose', verbose, .false.)

        call readpar_logical(file_parameter, 'yn_free_surface', yn_free_surface, .false.)
        call readpar_float(file_parameter, 'free_surface_dz_refine', free_surface_dz_refine, 4.0)
        call readpar_string(file_parameter, 'file_topo', file_topo, '')
        call readpar_string(file_parameter, 'topo_interp', topo_interp, 'cubic')
        call readpar_logical(file_parameter, 'measure_source_depth_from_surface', &
            measure_source_depth_from_surface, .false.)
        call readpar_logical(file_parameter, 'measure_receiver_depth_from_surface', &
            measure_receiver_depth_from_surface, .false.)

        select case (program)
            case ('modeling')
                call readpar_nstring(file_parameter, 'model_name', model_name, [''], required=.true.)
                if (model_name(1) == '') &
                    call stop_with('<read_parameter> Error: model_name cannot be empty or start with null')
            case ('gradient')
                call readpar_nstring(file_parameter, 'model_update', model_name, ['vp'])
                if (model_name(1) == '') &
                    call stop_with('<read_parameter> Error: model_update cannot be empty or start with null.')
                call readpar_string(file_parameter, 'misfit_type', misfit_type, 'waveform')
                call readpar_float(file_parameter, 'tlag_max', tlag_max, 0.1*tmax)
                call readpar_string(file_parameter, 'penalty_method', penalty_method, 'linear')
                call readpar_float(file_parameter, 'deconv_eps', deconv_eps, 0.1)
                call readpar_float(file_parameter, 'penalty_power', penalty_power, 1.0)
                call readpar_float(file_parameter, 'lawi_sigma', lawi_sigma, 0.25)
                call readpar_int(file_parameter, 'adaptive_half_window', adaptive_hw, 3)
                call readpar_float(file_parameter, 'trace_discard_threshold', trace_discard_threshold, 1.0e-6)
                call readpar_logical(file_parameter, 'yn_energy_precond', yn_energy_precond, .false.)
                call readpar_int(file_parameter, 'niter_max', niter_max, 100)
                call readpar_nstring(file_parameter, 'model_aux', model_name_aux, [''])
                if (model_name_aux(1) /= '') nmodel_aux = size(model_name_aux)
            case default
                call stop_with('<read_parameter> Error: unknown program '//program)
        end select
        nmodel = size(model_name)

        select case (tidy(which_medium))
            case ('acoustic-iso', 'acoustic-tti')
                call readpar_nstring(file_parameter, 'data_name', data_name, ['p'])
            case ('elastic-iso', 'elastic-vhtiort', 'elastic-tti')
                call readpar_nstring(file_parameter, 'data_name', data_name, ['x', 'z'])
                yn_compx = any(data_name == 'x')
                yn_compz = any(data_name == 'z')
            case default
                call stop_with('<read_parameter> Error: which_medium '//tidy(which_medium)//' is not known')
        end select
        ndata = size(data_name)
        if (ndata < 1) call stop_with('<read_parameter> Error: data_name is empty.')
        if (.not. all(data_name == 'x' .or. data_name == 'y' .or. data_name == 'z' .or. data_name == 'p')) &
            call stop_with('<read_parameter> Error: data_name must be in x, y, z, p')

        call set_regular_space

    end subroutine

    ! module_model.f90:72-160 (2D: ny = 1)
    subroutine set_regular_space
        call readpar_int(file_parameter, 'nx', nx0, 0, required=.true.)
        call readpar_int(file_parameter, 'ny', ny0, 1)
        call readpar_int(file_parameter, 'nz', nz0, 0, required=.true.)
        call readpar_float(file_parameter, 'dx', dx0, 1.0, required=.true.)
        call readpar_float(file_parameter, 'dy', dy0, 1.0)
        call readpar_float(file_parameter, 'dz', dz0, 1.0, required=.true.)
        call readpar_float(file_parameter, 'ox', ox0, 0.0)
        call readpar_float(file_parameter, 'oy', oy0, 0.0)
        call readpar_float(file_parameter, 'oz', oz0, 0.0)
        if (ny0 /= 1) call stop_with('<set_regular_space> Error: the 2D drivers need ny = 1')

        call readpar_int(file_parameter, 'nnx', nx, nx0)
        call readpar_int(file_parameter, 'nny', ny, ny0)
        call readpar_int(file_parameter, 'nnz', nz, nz0)
        call readpar_float(file_parameter, 'ddx', dx, dx0)
        call readpar_float(file_parameter, 'ddy', dy, dy0)
        call readpar_float(file_parameter, 'ddz', dz, dz0)
        call readpar_float(file_parameter, 'oox', ox, ox0)
        call readpar_float(file_parameter, 'ooy', oy, oy0)
        call readpar_float(file_parameter, 'ooz', oz, oz0)

        call readpar_float(file_parameter, 'xmin', xmin, ox)
        call readpar_float(file_parameter, 'xmax', xmax, ox + (nx - 1)*dx)
        if (xmax < xmin) call stop_with('<set_regular_space> Error: xmax must >= xmin')
        call readpar_float(file_parameter, 'ymin', ymin, oy)
        call readpar_float(file_parameter, 'ymax', ymax, oy + (ny - 1)*dy)
        if (ymax < ymin) call stop_with('<set_regular_space> Error: ymax must >= ymin')
        call readpar_float(file_parameter, 'zmin', zmin, oz)
        call readpar_float(file_parameter, 'zmax', zmax, oz + (nz - 1)*dz)
        if (zmax < zmin) call stop_with('<set_regular_space> Error: zmax must >= zmin')

        ox = xmin
        oy = ymin
        oz = zmin
        nx = ceiling((xmax - xmin)/dx) + 1
        ny = ceiling((ymax - ymin)/dy) + 1
        nz = ceiling((zmax - zmin)/dz) + 1

        require_model_interp = nx /= nx0 .or. ny /= ny0 .or. nz /= nz0 .or. &
            dx /= dx0 .or. dy /= dy0 .or. dz /= dz0 .or. &
            ox /= ox0 .or. oy /= oy0 .or. oz /= oz0

        if (.not. (nx >= 1 .and. ny >= 1 .and. nz >= 1)) call stop_with('<set_regular_space> Error: n# must >= 1')
        if (.not. (dx > 0 .and. dy > 0 .and. dz > 0)) call stop_with('<set_regular_space> Error: d# must > 0')

        call readpar_int(file_parameter, 'npml', pml, 15)
        if (yn_free_surface) then
            call readpar_float(file_parameter, 'dz_max', dz_max, 1.5*dz)
        else
            dz_max = dz
        end if
    end subroutine

    ! FLIT's ginterp through the C ABI
    function ginterp(x, y, xi, method) result(yi)
        real, intent(in) :: x(:), y(:), xi(:)
        character(len=*), intent(in) :: method
        real, allocatable :: yi(:)
        integer(c_int) :: st
        allocate (yi(size(xi)))
        if (size(xi) == 0) return
        st = owl_ginterp(size(x), x, y, size(xi), xi, yi, merge(1_c_int, 0_c_int, tidy(method) == 'cubic'))
        if (st /= OWL_OK) call stop_with('<owl_ginterp> '//owl_error_message())
    end function

    ! file_topo: x and elevation per non-empty line
    subroutine load_topo(path, topo)
        character(len=*), intent(in) :: path
        real, allocatable, intent(out) :: topo(:, :)
        character(len=1024) :: line
        integer :: u, ios, n, k
        open (newunit=u, file=tidy(path), status='old', action='read')
        n = 0
        do
            read (u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            if (len_trim(line) > 0) n = n + 1
        end do
        allocate (topo(n, 2))
        rewind (u)
        k = 0
        do while (k < n)
            read (u, '(a)') line
            if (len_trim(line) == 0) cycle
            k = k + 1
            read (line, *) topo(k, 1), topo(k, 2)
        end do
        close (u)
    end subroutine

    function mean(a) result(m)
        real, intent(in) :: a(:)
        real :: m
        integer :: k
        m = 0.0
        do k = 1, size(a)
            m = m + a(k)
        end do
        m = m/size(a)
    end function

    ! module_source_receiver.f90:183-551, 2D
    subroutine load_geometry
        integer :: i, j, l, u, v, ios, nq
        character(len=1024) :: dir_geometry, file_source, ff
        character(len=1024), allocatable :: sq(:)
        real, allocatable :: fq(:), topo(:, :), topox(:), topo_i(:), stp(:), rtp(:)
        logical, allocatable :: qs(:), src_below(:), rec_below(:)
        logical :: exists
        real :: topo_max
        type(source_receiver_geometry), allocatable :: g(:)
        integer :: k, nk

        if (allocated(gmtr)) deallocate (gmtr)
        allocate (gmtr(1:ns))
        inquire (file=tidy(file_geometry), exist=exists)
        if (.not. exists) call stop_with('Error: Geometry file not found.')
        k = index(file_geometry, '/', back=.true.)
        dir_geometry = ''
        if (k > 0) dir_geometry = file_geometry(1:k)

        open (newunit=u, file=tidy(file_geometry), action='read', status='old')
        do i = 1, ns

            do
                read (u, '(a)', iostat=ios) file_source
                if (ios /= 0) call stop_with('<load_geometry> Error: '//tidy(file_geometry)//' lists fewer than ns files')
                if (len_trim(file_source) > 0) exit
            end do

            open (newunit=v, file=tidy(dir_geometry)//tidy(file_source), status='old', action='read')
            read (v, *) gmtr(i)%id
            read (v, *) gmtr(i)%ns
            allocate (gmtr(i)%srcr(1:gmtr(i)%ns))

            do j = 1, gmtr(i)%ns
                read (v, *) gmtr(i)%srcr(j)%x, gmtr(i)%srcr(j)%y, gmtr(i)%srcr(j)%z

                read (v, '(a)') ff
                call split_items(ff, sq)
                gmtr(i)%srcr(j)%mechanism = tidy(sq(1))
                select case (tidy(sq(1)))
                    case ('explosion')
                    case ('force')
                        if (size(sq) < 3) call stop_with('<load_geometry> Error: force needs polar and azimuth')
                        gmtr(i)%srcr(j)%polar = to_real(sq(2))*const_deg2rad
                        gmtr(i)%srcr(j)%azimuth = to_real(sq(3))*const_deg2rad
                    case ('mt')
                        ! kept by name: the solvers refuse moment tensors
                    case default
                        call stop_with('<load_geometry> Error: mechanism must be one of explosion, force, mt.')
                end select

                read (v, '(a)') ff
                call split_items(ff, sq)
                if (size(sq) < 4) call stop_with('<load_geometry> Error: wavelet line needs name, f0, amp, t0')
                gmtr(i)%srcr(j)%wavelet = tidy(sq(1))
                gmtr(i)%srcr(j)%f0 = to_real(sq(2))
                gmtr(i)%srcr(j)%amp = to_real(sq(3))
                gmtr(i)%srcr(j)%t0 = to_real(sq(4))
                if (gmtr(i)%srcr(j)%wavelet == 'custom') read (v, '(a)') ff

                ! time-function filtering: frequencies then coefficients
                read (v, '(a)') ff
                call split_items(ff, sq)
                nq = size(sq)/2
                if (nq > 1 .and. gmtr(i)%srcr(j)%wavelet /= 'ormsby') then
                    allocate (fq(size(sq)))
                    do nk = 1, size(sq)
                        fq(nk) = to_real(sq(nk))
                    end do
                    if (any(fq(1:nq) /= 0)) &
                        call stop_with('<load_geometry> Error: source time-function filtering is not provided')
                    deallocate (fq)
                end if
                if (src_filt_freqs(1) >= 0) &
                    call stop_with('<load_geometry> Error: src_filt_freqs (source filtering) is not provided')
            end do

            read (v, *) gmtr(i)%nr
            allocate (gmtr(i)%recr(1:gmtr(i)%nr))
            do j = 1, gmtr(i)%nr
                read (v, *) gmtr(i)%recr(j)%x, gmtr(i)%recr(j)%y, gmtr(i)%recr(j)%z, gmtr(i)%recr(j)%weight
            end do
            close (v)

            if (ny == 1) then
                gmtr(i)%srcr(:)%y = 0.5*(ymin + ymax)
                gmtr(i)%recr(:)%y = 0.5*(ymin + ymax)
            end if

            gmtr(i)%f0factor = f0_factor

            allocate (src_below(gmtr(i)%ns), rec_below(gmtr(i)%nr))
            if (yn_free_surface .and. tidy(file_topo) /= '') then
                call load_topo(file_topo, topo)
                allocate (topox(nx + 2*pml + 2))
                do k = 1, size(topox)
                    topox(k) = (k - 1)*dx - (pml + 1)*dx
                end do
                topo_i = ginterp(topo(:, 1), topo(:, 2), topox, topo_interp)
                if (measure_source_depth_from_surface) then
                    src_below = .true.
                else
                    stp = ginterp(topox, -topo_i, gmtr(i)%srcr(:)%x, topo_interp)
                    src_below = gmtr(i)%srcr(:)%z >= stp
                end if
                if (measure_receiver_depth_from_surface) then
                    rec_below = .true.
                else
                    rtp = ginterp(topox, -topo_i, gmtr(i)%recr(:)%x, topo_interp)
                    rec_below = gmtr(i)%recr(:)%z >= rtp
                end if
                topo_max = maxval(topo_i)
                deallocate (topox)
            else
                src_below = .true.
                rec_below = .true.
                topo_max = 0
            end if

            l = 0
            do j = 1, gmtr(i)%ns
                associate (s => gmtr(i)%srcr(j))
                    if (s%x >= xmin .and. s%x <= xmax .and. s%y >= ymin .and. s%y <= ymax .and. &
                            s%z >= zmin - topo_max .and. s%z <= zmax - topo_max .and. &
                            s%x >= sxmin .and. s%x <= sxmax .and. s%y >= symin .and. s%y <= symax .and. &
                            s%z - topo_max >= szmin .and. s%z <= szmax - topo_max .and. &
                            gmtr(i)%id >= sid_min .and. gmtr(i)%id <= sid_max .and. src_below(j)) then
                        l = l + 1
                    else
                        s%amp = 0
                    end if
                end associate
            end do
            if (l == 0) gmtr(i)%ns = 0

            l = 0
            do j = 1, gmtr(i)%nr
                associate (r => gmtr(i)%recr(j))
                    r%aoff = sqrt((r%x - mean(gmtr(i)%srcr(:)%x))**2 + (r%y - mean(gmtr(i)%srcr(:)%y))**2 &
                        + (r%z - mean(gmtr(i)%srcr(:)%z))**2)
                    if (r%x >= xmin .and. r%x <= xmax .and. r%y >= ymin .and. r%y <= ymax .and. &
                            r%z >= zmin - topo_max .and. r%z <= zmax - topo_max .and. &
                            r%x >= rxmin .and. r%x <= rxmax .and. r%y >= rymin .and. r%y <= rymax .and. &
                            r%z >= rzmin - topo_max .and. r%z <= rzmax - topo_max .and. &
                            r%aoff >= offset_min .and. r%aoff <= offset_max .and. &
                            j >= rec_min .and. j <= rec_max .and. mod(j - rec_min, rec_every) == 0 .and. &
                            r%weight /= 0 .and. rec_below(j)) then
                        l = l + 1
                    else
                        r%weight = 0.0
                    end if
                end associate
            end do
            if (l == 0) gmtr(i)%nr = 0
            deallocate (src_below, rec_below)

        end do
        close (u)

        allocate (qs(ns))
        qs = .false.
        do i = shot_min, shot_max, shot_every
            if (gmtr(i)%ns /= 0 .and. gmtr(i)%nr /= 0) qs(i) = .true.
        end do
        if (any(src_select /= 0)) then
            qs = .false.
            do i = 1, ns
                if (gmtr(i)%ns /= 0 .and. gmtr(i)%nr /= 0 .and. any(src_select == i)) qs(i) = .true.
            end do
        end if
        if (any(sid_select /= 0)) then
            qs = .false.
            do i = 1, ns
                if (gmtr(i)%ns /= 0 .and. gmtr(i)%nr /= 0 .and. any(sid_select == gmtr(i)%id)) qs(i) = .true.
            end do
        end if
        if (any(src_exclude /= 0)) then
            do i = 1, ns
                if (qs(i) .and. any(src_exclude == i)) qs(i) = .false.
            end do
        end if
        if (any(sid_exclude /= 0)) then
            do i = 1, ns
                if (qs(i) .and. any(sid_exclude == gmtr(i)%id)) qs(i) = .false.
            end do
        end if
        if (.not. any(qs)) call stop_with('Error: No shot or receiver in the model.')

        allocate (g(count(qs)))
        l = 0
        do i = 1, ns
            if (qs(i)) then
                l = l + 1
                g(l) = gmtr(i)
            end if
        end do
        call move_alloc(g, gmtr)
        ns = size(gmtr)

        do i = 1, ns
            do j = i + 1, ns
                if (gmtr(i)%id == gmtr(j)%id) call stop_with('<load_geometry> Erorr: Duplicate source ID found.')
            end do
        end do
    end subroutine

    function to_real(s) result(r)
        character(len=*), intent(in) :: s
        real :: r
        integer :: ios
        read (s, *, iostat=ios) r
        if (ios /= 0) call stop_with('<load_geometry> Error: cannot read a number from '//tidy(s))
    end function

    ! linear resampling of a (nz0, nx0) model onto the target grid
    subroutine interp_linear_2d(w0, w)
        real, intent(in) :: w0(:, :)
        real, intent(out) :: w(:, :)
        real, allocatable :: t(:, :)
        integer :: i, j, k
        real :: p, f
        allocate (t(nz, nx0))
        do j = 1, nx0
            do i = 1, nz
                p = (oz + (i - 1)*dz - oz0)/dz0
                k = max(1, min(nz0 - 1, floor(p) + 1))
                f = max(0.0, min(1.0, p - (k - 1)))
                if (nz0 == 1) then
                    t(i, j) = w0(1, j)
                else
                    t(i, j) = w0(k, j)*(1 - f) + w0(k + 1, j)*f
                end if
            end do
        end do
        do j = 1, nx
            p = (ox + (j - 1)*dx - ox0)/dx0
            k = max(1, min(nx0 - 1, floor(p) + 1))
            f = max(0.0, min(1.0, p - (k - 1)))
            if (nx0 == 1) then
                w(:, j) = t(:, 1)
            else
                w(:, j) = t(:, k)*(1 - f) + t(:, k + 1)*f
            end if
        end do
    end subroutine

    ! module_modeling.f90:35-47: file_<name> for every model_name (the
    ! gradient driver also reads the model_aux names after them)
    subroutine prepare_model
        integer :: i, u
        character(len=1024) :: fname
        real, allocatable :: wt(:, :)
        if (allocated(model_m)) deallocate (model_m)
        allocate (model_m(nmodel + nmodel_aux))
        do i = 1, nmodel + nmodel_aux
            if (i <= nmodel) then
                model_m(i)%name = model_name(i)
            else
                model_m(i)%name = model_name_aux(i - nmodel)
            end if
            call readpar_string(file_parameter, 'file_'//tidy(model_m(i)%name), fname, '')
            allocate (model_m(i)%array(nz, nx))
            model_m(i)%array = 0.0
            if (tidy(fname) /= '') then
                allocate (wt(nz0, nx0))
                open (newunit=u, file=tidy(fname), access='stream', form='unformatted', status='old', action='read')
                read (u) wt
                close (u)
                if (require_model_interp) then
                    call interp_linear_2d(wt, model_m(i)%array)
                else
                    model_m(i)%array = wt
                end if
                deallocate (wt)
            else
                write (*, '(a)') ' Warning: Moodel '//tidy(model_m(i)%name)//' is empty. '
            end if
        end do
    end subroutine

    ! module_model.f90:163-197: a model on the current window, or the
    ! default value when model_name does not hold it
    function get_model(name, default_value) result(w)
        character(len=*), intent(in) :: name
        real, intent(in), optional :: default_value
        real, allocatable :: w(:, :)
        integer :: i
        do i = 1, size(model_m)
            if (tidy(model_m(i)%name) == name) then
                w = model_m(i)%array(shot_nzbeg:shot_nzend, shot_nxbeg:shot_nxend)
                return
            end if
        end do
        if (.not. present(default_value)) &
            write (*, '(a)') ' <get_model> Warning: Default_value for '//name//' not given; set to zero. '
        allocate (w(shot_nz, shot_nx))
        w = 0.0
        if (present(default_value)) w = w + default_value
    end function

    function has_model(name) result(yes)
        character(len=*), intent(in) :: name
        logical :: yes
        integer :: i
        yes = .false.
        do i = 1, size(model_m)
            if (tidy(model_m(i)%name) == name) yes = .true.
        end do
    end function

    ! module_utility.f90:410-510 through owl_set_adaptive_range: the window
    ! of shot g; receivers outside it or outside the offset range get weight 0
    subroutine set_adaptive_range(g)
        type(source_receiver_geometry), intent(inout) :: g
        type(owl_window_params_t) :: p
        type(owl_source_t), allocatable :: cs(:)
        type(owl_receiver_t), allocatable :: cr(:)
        integer(c_int) :: win(6), st
        real(c_float) :: rng(6)
        integer :: k
        allocate (cs(size(g%srcr)), cr(size(g%recr)))
        p%nx = nx
        p%nz = nz
        p%dx = dx
        p%dz = dz
        p%ox = ox
        p%oz = oz
        p%xmin = xmin
        p%xmax = xmax
        p%zmin = zmin
        p%zmax = zmax
        p%yn_adpx = merge(1, 0, yn_adpx)
        p%yn_adpz = merge(1, 0, yn_adpz)
        p%adp_extrax = adp_extrax
        p%adp_extraz = adp_extraz
        p%offset_min = offset_min
        p%offset_max = offset_max
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
        if (st /= OWL_OK) call stop_with('<set_adaptive_range> '//owl_error_message())
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

    function itoa(i) result(s)
        integer, intent(in) :: i
        character(len=:), allocatable :: s
        character(len=32) :: buf
        write (buf, '(i0)') i
        s = trim(buf)
    end function

    ! the parsed set-up, one "key value" per line (the drivers' --dry-run)
    subroutine print_setup
        integer :: i, j, nlive
        write (*, '(a,a)') 'which_medium ', tidy(which_medium)
        write (*, '(a,a)') 'anisotropy_type ', tidy(aniso_param)
        write (*, '(a,3(1x,i0))') 'grid', nx, nz, pml
        write (*, '(a,4(1x,es16.8))') 'spacing', dx, dz, ox, oz
        write (*, '(a,4(1x,es16.8))') 'time', dt, tmax, data_dt, data_tmax
        write (*, '(a,l2,1x,es16.8,1x,es16.8)') 'free_surface', yn_free_surface, free_surface_dz_refine, dz_max
        write (*, '(a,*(1x,a))') 'models', (tidy(model_name(i)), i = 1, nmodel)
        write (*, '(a,*(1x,a))') 'data', (tidy(data_name(i)), i = 1, ndata)
        write (*, '(a,1x,i0)') 'shots', ns
        do i = 1, ns
            call set_adaptive_range(gmtr(i))
            nlive = count(gmtr(i)%recr(:)%weight /= 0)
            write (*, '(a,*(1x,i0))') 'shot', gmtr(i)%id, size(gmtr(i)%srcr), size(gmtr(i)%recr), nlive, &
                shot_nxbeg, shot_nxend, shot_nzbeg, shot_nzend
            do j = 1, size(gmtr(i)%srcr)
                write (*, '(a,1x,a,1x,a,6(1x,es16.8))') 'src', tidy(gmtr(i)%srcr(j)%mechanism), &
                    tidy(gmtr(i)%srcr(j)%wavelet), gmtr(i)%srcr(j)%x, gmtr(i)%srcr(j)%z, gmtr(i)%srcr(j)%f0, &
                    gmtr(i)%srcr(j)%amp, gmtr(i)%srcr(j)%polar, gmtr(i)%srcr(j)%t0
            end do
            do j = 1, size(gmtr(i)%recr)
                write (*, '(a,4(1x,es16.8))') 'rec', gmtr(i)%recr(j)%x, gmtr(i)%recr(j)%z, gmtr(i)%recr(j)%weight, &
                    gmtr(i)%recr(j)%aoff
            end do
        end do
        do i = 1, size(model_m)
            write (*, '(a,1x,a,2(1x,es16.8),1x,es24.16)') 'model', tidy(model_m(i)%name), minval(model_m(i)%array), &
                maxval(model_m(i)%array), sum(real(model_m(i)%array, 8))
        end do
    end subroutine

    ! the anisotropy_type branches of module_modeling.f90:192-229 and
    ! module_gradient.f90:240-275
    subroutine set_tti_medium(s)
        type(wave_solver_elastic_tti_2d), intent(inout) :: s
        select case (tidy(aniso_param))
            case ('iso')
                s%anisotropy_type = 'iso'
                s%vp = get_model('vp')
                s%vs = get_model('vs')
                s%rho = get_model('rho', 1.0)
            case ('thomsen')
                s%anisotropy_type = 'thomsen'
                s%vp = get_model('vp')
                s%vs = get_model('vs')
                s%tieps = get_model('epsilon', 0.0)
                s%tidel = get_model('delta', 0.0)
                s%tithe = get_model('theta', 0.0)
                s%rho = get_model('rho', 1.0)
            case ('a-t')
                s%anisotropy_type = 'a-t'
                s%vp = get_model('vp')
                s%vs = get_model('vs')
                s%tieps = get_model('epsilon', 0.0)
                s%tieta = get_model('eta', 0.0)
                s%tithe = get_model('theta', 0.0)
                s%rho = get_model('rho', 1.0)
            case ('cij')
                s%anisotropy_type = 'cij'
                s%c11 = get_model('c11')
                s%c13 = get_model('c13')
                s%c15 = get_model('c15')
                s%c33 = get_model('c33')
                s%c35 = get_model('c35')
                s%c55 = get_model('c55')
                s%rho = get_model('rho', 1.0)
            case default
                call stop_with('Error: anisotropy_type = '//tidy(aniso_param)//' is not known')
        end select
    end subroutine

end module owl_caller
