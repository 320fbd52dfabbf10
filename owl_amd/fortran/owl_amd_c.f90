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
