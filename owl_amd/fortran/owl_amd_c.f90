!
! iso_c_binding declarations of the C ABI in include/owl_amd.h: the
! interoperable derived types and the entry points used from Fortran.
!
module owl_amd_c

    use iso_c_binding

    implicit none

    integer(c_int), parameter :: OWL_OK = 0
    integer(c_int), parameter :: OWL_EXPLOSION = 0, OWL_FORCE = 1
    integer(c_int), parameter :: OWL_RICKER = 0
    integer(c_int), parameter :: OWL_GRAD_OWL = 0, OWL_GRAD_EXACT = 1
    integer(c_int), parameter :: OWL_CKPT_FULL = 0, OWL_CKPT_UNIFORM = 1, OWL_CKPT_BINOMIAL = 2
    integer(c_int), parameter :: OWL_STORE_F32 = 0, OWL_STORE_F16 = 1, OWL_STORE_BF16 = 2
    integer(c_int), parameter :: OWL_MISFIT_WAVEFORM = 0, OWL_MISFIT_CORR = 1, OWL_MISFIT_ENVELOPE = 2, &
        OWL_MISFIT_PHASE = 3, OWL_MISFIT_ADAPTIVE = 4, OWL_MISFIT_LOCAL_ADAPTIVE = 5, &
        OWL_MISFIT_ADAPTIVE_SPACETIME = 6
    integer(c_int), parameter :: OWL_PENALTY_LINEAR = 0, OWL_PENALTY_POWER = 1, OWL_PENALTY_GAUSSIAN = 2, &
        OWL_PENALTY_EXP = 3

    type, bind(c) :: owl_misfit_params_t
        real(c_float) :: dt, tlag_max
        integer(c_int) :: penalty
        real(c_float) :: penalty_power, deconv_eps, lawi_sigma
        integer(c_int) :: adaptive_hw
    end type

    type, bind(c) :: owl_source_t
        real(c_float) :: x = 0, y = 0, z = 0
        real(c_float) :: f0 = 20, amp = 1, t0 = 0
        real(c_float) :: polar = 0, azimuth = 0
        integer(c_int) :: mechanism = 0
        integer(c_int) :: wavelet = 0
    end type

    type, bind(c) :: owl_receiver_t
        real(c_float) :: x = 0, y = 0, z = 0
        real(c_float) :: weight = 1
    end type

    type, bind(c) :: owl_shot_t
        integer(c_int) :: id = 1
        integer(c_int) :: ns = 0
        type(c_ptr) :: sources = c_null_ptr
        integer(c_int) :: nr = 0
        type(c_ptr) :: receivers = c_null_ptr
        integer(c_int) :: nxbeg = 0, nxend = 0, nybeg = 0, nyend = 0, nzbeg = 0, nzend = 0
    end type

    type, bind(c) :: owl_acoustic_config_t
        integer(c_int) :: nx = 0, ny = 1, nz = 0
        real(c_float) :: dx = 0, dy = 1, dz = 0
        real(c_float) :: ox = 0, oy = 0, oz = 0
        real(c_float) :: dt = 0, tmax = 0
        real(c_float) :: data_dt = 0, data_tmax = 0
        integer(c_int) :: pml = 15
        integer(c_int) :: fdhalf = 4
        integer(c_int) :: free_surface = 0
        real(c_float) :: f0factor = 1
        integer(c_int) :: cc_step_interval = 1
        real(c_float) :: trace_discard_threshold = 1.0e-6
    end type

    ! owl_fsg_surface: OWL's free surface (file_topo points, topo_interp,
    ! free_surface_dz_refine, dz_max, depths measured from the surface)
    type, bind(c) :: owl_fsg_surface_t
        integer(c_int) :: ntopo = 0
        type(c_ptr) :: topo_x = c_null_ptr
        type(c_ptr) :: topo_z = c_null_ptr
        integer(c_int) :: topo_interp = 1
        real(c_float) :: dz_refine = 4.0
        real(c_float) :: dz_max = 0.0
        integer(c_int) :: src_depth_from_surface = 0
        integer(c_int) :: rec_depth_from_surface = 0
    end type

    type, bind(c) :: owl_window_params_t
        integer(c_int) :: nx = 1, ny = 1, nz = 1
        real(c_float) :: dx = 1, dy = 1, dz = 1
        real(c_float) :: ox = 0, oy = 0, oz = 0
        real(c_float) :: xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0
        integer(c_int) :: yn_adpx = 0, yn_adpy = 0, yn_adpz = 0
        real(c_float) :: adp_extrax = 0, adp_extray = 0, adp_extraz = 0
        real(c_float) :: offset_min = 0, offset_max = huge(1.0_c_float)
    end type

    interface

        function owl_last_error() bind(c, name='owl_last_error') result(msg)
            import :: c_ptr
            type(c_ptr) :: msg
        end function

        function owl_acoustic2d_create(cfg, vp, rho, nshot, shots, out) &
                bind(c, name='owl_acoustic2d_create') result(st)
            import :: c_int, c_float, c_ptr, owl_acoustic_config_t, owl_shot_t
            type(owl_acoustic_config_t), intent(in) :: cfg
            real(c_float), intent(in) :: vp(*), rho(*)
            integer(c_int), value :: nshot
            type(owl_shot_t), intent(in) :: shots(*)
            type(c_ptr), intent(out) :: out
            integer(c_int) :: st
        end function

        subroutine owl_acoustic2d_destroy(b) bind(c, name='owl_acoustic2d_destroy')
            import :: c_ptr
            type(c_ptr), value :: b
        end subroutine

        function owl_acoustic2d_nt(b) bind(c, name='owl_acoustic2d_nt') result(nt)
            import :: c_ptr, c_int
            type(c_ptr), value :: b
            integer(c_int) :: nt
        end function

        function owl_acoustic2d_forward(b, traces, save_boundary) &
                bind(c, name='owl_acoustic2d_forward') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(out) :: traces(*)
            integer(c_int), value :: save_boundary
            integer(c_int) :: st
        end function

        function owl_acoustic2d_adjoint(b, adj, grad_vp, grad_rho) &
                bind(c, name='owl_acoustic2d_adjoint') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(in) :: adj(*)
            real(c_float), intent(out) :: grad_vp(*), grad_rho(*)
            integer(c_int) :: st
        end function

        function owl_acoustic2d_gradient(b, obs, mode, ckpt, ckpt_param, store, misfit, &
                grad_vp, grad_rho, traces) bind(c, name='owl_acoustic2d_gradient') result(st)
            import :: c_ptr, c_int, c_float, c_double
            type(c_ptr), value :: b
            real(c_float), intent(in) :: obs(*)
            integer(c_int), value :: mode, ckpt, ckpt_param, store
            real(c_double), intent(out) :: misfit(*), grad_vp(*), grad_rho(*)
            real(c_float), intent(out) :: traces(*)
            integer(c_int) :: st
        end function

        function owl_acoustic2d_set_energy_precond(b, on) &
                bind(c, name='owl_acoustic2d_set_energy_precond') result(st)
            import :: c_ptr, c_int
            type(c_ptr), value :: b
            integer(c_int), value :: on
            integer(c_int) :: st
        end function

        function owl_acoustic3d_create(cfg, vp, rho, nshot, shots, out) &
                bind(c, name='owl_acoustic3d_create') result(st)
            import :: c_int, c_float, c_ptr, owl_acoustic_config_t, owl_shot_t
            type(owl_acoustic_config_t), intent(in) :: cfg
            real(c_float), intent(in) :: vp(*), rho(*)
            integer(c_int), value :: nshot
            type(owl_shot_t), intent(in) :: shots(*)
            type(c_ptr), intent(out) :: out
            integer(c_int) :: st
        end function

        subroutine owl_acoustic3d_destroy(b) bind(c, name='owl_acoustic3d_destroy')
            import :: c_ptr
            type(c_ptr), value :: b
        end subroutine

        function owl_acoustic3d_nt(b) bind(c, name='owl_acoustic3d_nt') result(nt)
            import :: c_ptr, c_int
            type(c_ptr), value :: b
            integer(c_int) :: nt
        end function

        function owl_acoustic3d_forward(b, traces, save_boundary) &
                bind(c, name='owl_acoustic3d_forward') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(out) :: traces(*)
            integer(c_int), value :: save_boundary
            integer(c_int) :: st
        end function

        function owl_acoustic3d_adjoint(b, adj, grad_vp, grad_rho) &
                bind(c, name='owl_acoustic3d_adjoint') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(in) :: adj(*)
            real(c_float), intent(out) :: grad_vp(*), grad_rho(*)
            integer(c_int) :: st
        end function

        function owl_set_adaptive_range(p, ns, src, nr, rec, window, range) &
                bind(c, name='owl_set_adaptive_range') result(st)
            import :: c_int, c_float, owl_window_params_t, owl_source_t, owl_receiver_t
            type(owl_window_params_t), intent(in) :: p
            integer(c_int), value :: ns
            type(owl_source_t), intent(in) :: src(*)
            integer(c_int), value :: nr
            type(owl_receiver_t), intent(inout) :: rec(*)
            integer(c_int), intent(out) :: window(6)
            real(c_float), intent(out) :: range(6)
            integer(c_int) :: st
        end function

        function owl_adjoint_source(mtype, ntr, nt, obs, syn, weight, threshold, adj, misfit) &
                bind(c, name='owl_adjoint_source') result(st)
            import :: c_int, c_float, c_double
            integer(c_int), value :: mtype, ntr, nt
            real(c_float), intent(in) :: obs(*), syn(*), weight(*)
            real(c_float), value :: threshold
            real(c_float), intent(out) :: adj(*)
            real(c_double), intent(out) :: misfit(*)
            integer(c_int) :: st
        end function

        function owl_adjoint_source_ex(mtype, prm, ntr, nt, obs, syn, weight, threshold, adj, misfit) &
                bind(c, name='owl_adjoint_source_ex') result(st)
            import :: c_int, c_float, c_double, owl_misfit_params_t
            integer(c_int), value :: mtype, ntr, nt
            type(owl_misfit_params_t), intent(in) :: prm
            real(c_float), intent(in) :: obs(*), syn(*), weight(*)
            real(c_float), value :: threshold
            real(c_float), intent(out) :: adj(*)
            real(c_double), intent(out) :: misfit(*)
            integer(c_int) :: st
        end function

        function owl_resample_traces(nr, nt, dt, tin, nnt, ddt, oot, tout) &
                bind(c, name='owl_resample_traces') result(st)
            import :: c_int, c_float
            integer(c_int), value :: nr, nt, nnt
            real(c_float), value :: dt, ddt, oot
            real(c_float), intent(in) :: tin(*)
            real(c_float), intent(out) :: tout(*)
            integer(c_int) :: st
        end function

        function owl_ginterp(n, x, y, m, xi, yi, cubic) bind(c, name='owl_ginterp') result(st)
            import :: c_int, c_float
            integer(c_int), value :: n, m, cubic
            real(c_float), intent(in) :: x(*), y(*), xi(*)
            real(c_float), intent(out) :: yi(*)
            integer(c_int) :: st
        end function

        function owl_fsg2d_create_surface(cfg, cij, rho, aniso, nshot, shots, surface, out) &
                bind(c, name='owl_fsg2d_create_surface') result(st)
            import :: c_int, c_float, c_ptr, owl_acoustic_config_t, owl_shot_t, owl_fsg_surface_t
            type(owl_acoustic_config_t), intent(in) :: cfg
            type(c_ptr), intent(in) :: cij(6)
            real(c_float), intent(in) :: rho(*)
            integer(c_int), value :: aniso, nshot
            type(owl_shot_t), intent(in) :: shots(*)
            type(owl_fsg_surface_t), intent(in) :: surface
            type(c_ptr), intent(out) :: out
            integer(c_int) :: st
        end function

        function owl_fsg2d_create(cfg, cij, rho, aniso, nshot, shots, out) &
                bind(c, name='owl_fsg2d_create') result(st)
            import :: c_int, c_float, c_ptr, owl_acoustic_config_t, owl_shot_t
            type(owl_acoustic_config_t), intent(in) :: cfg
            type(c_ptr), intent(in) :: cij(6)
            real(c_float), intent(in) :: rho(*)
            integer(c_int), value :: aniso, nshot
            type(owl_shot_t), intent(in) :: shots(*)
            type(c_ptr), intent(out) :: out
            integer(c_int) :: st
        end function

        subroutine owl_fsg2d_destroy(b) bind(c, name='owl_fsg2d_destroy')
            import :: c_ptr
            type(c_ptr), value :: b
        end subroutine

        function owl_fsg2d_nt(b) bind(c, name='owl_fsg2d_nt') result(nt)
            import :: c_ptr, c_int
            type(c_ptr), value :: b
            integer(c_int) :: nt
        end function

        function owl_fsg2d_forward(b, traces, save_boundary) bind(c, name='owl_fsg2d_forward') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(out) :: traces(*)
            integer(c_int), value :: save_boundary
            integer(c_int) :: st
        end function

        function owl_fsg2d_adjoint(b, adj, grads) bind(c, name='owl_fsg2d_adjoint') result(st)
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: b
            real(c_float), intent(in) :: adj(*)
            type(c_ptr), intent(in) :: grads(7)
            integer(c_int) :: st
        end function

        function owl_tti2d_cij(param, n, vp, vs, rho, p1, p2, theta, cij) bind(c, name='owl_tti2d_cij') result(st)
            import :: c_int, c_long_long, c_ptr
            integer(c_int), value :: param
            integer(c_long_long), value :: n
            type(c_ptr), value :: vp, vs, rho, p1, p2, theta
            type(c_ptr), intent(in) :: cij(6)
            integer(c_int) :: st
        end function

        function owl_tti2d_param_gradient(param, n, vp, vs, rho, p1, p2, theta, grad_cij, grads) &
                bind(c, name='owl_tti2d_param_gradient') result(st)
            import :: c_int, c_long_long, c_ptr
            integer(c_int), value :: param
            integer(c_long_long), value :: n
            type(c_ptr), value :: vp, vs, rho, p1, p2, theta
            type(c_ptr), intent(in) :: grad_cij(6), grads(4)
            integer(c_int) :: st
        end function

    end interface

contains

    ! Message of the last failed call, as a Fortran string.
    function owl_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        type(c_ptr) :: p
        p = owl_last_error()
        call c_f_pointer(p, s, [4096])
        n = 0
        do while (n < 4096)
            if (s(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate (character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function

end module owl_amd_c
