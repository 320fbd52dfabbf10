!
! Fortran host side of the 2D acoustic solver: the public components and
! type-bound procedures of OWL's type(wave_solver_acoustic_iso_2d)
! (src/solver/acoustic_iso_2d/interface.f90:30-72), with forward and
! adjoint running on the GPU through the C ABI (owl_amd_c).  Outputs are
! OWL's files: shot_<id>_seismogram_p.su in dir_synthetic, and
! shot_<id>_grad_{vp,rho}.grd in dir_working; the adjoint source is read from
! dir_adjoint/shot_<id>_seismogram_p.su.
!
! Geometry is passed as arrays of the interoperable owl_source_t /
! owl_receiver_t; inside OWL's tree a caller fills them from
! type(source_receiver_geometry).
!
module owl_acoustic_iso_2d

    use iso_c_binding
    use owl_amd_c

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
        character(len=1024) :: dir_working = './'
        integer :: cc_step_interval = 1
        logical :: reconstruct = .false.
        integer :: fdhalf = 4
        integer :: id = 1
        type(owl_source_t), allocatable, dimension(:) :: srcr
        type(owl_receiver_t), allocatable, dimension(:) :: recr
        real, allocatable, dimension(:, :) :: seis_p      ! (nt, nr)
        type(c_ptr), private :: batch = c_null_ptr

    contains

        procedure, public :: forward => compute_forward
        procedure, public :: adjoint => compute_adjoint
        procedure, public :: release => release_batch

    end type

    private
    public :: wave_solver_acoustic_iso_2d, write_su_file, read_su_file, write_grd_file

contains

    subroutine check(st, what)
        integer(c_int), intent(in) :: st
        character(len=*), intent(in) :: what
        if (st /= OWL_OK) then
            write (*, '(a)') ' <owl_acoustic_iso_2d> '//what//' failed: '//owl_error_message()
            error stop 1
        end if
    end subroutine

    subroutine release_batch(this)
        class(wave_solver_acoustic_iso_2d), intent(inout) :: this
        if (c_associated(this%batch)) call owl_acoustic2d_destroy(this%batch)
        this%batch = c_null_ptr
    end subroutine

    ! compute_forward, submodule_forward_modeling.f90:31-132
    subroutine compute_forward(this)

        class(wave_solver_acoustic_iso_2d), intent(inout), target :: this

        type(owl_acoustic_config_t) :: cfg
        type(owl_shot_t) :: shot(1)
        real, allocatable :: rho(:, :), tr(:)
        integer :: nt, nr
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
        cfg%free_surface = merge(1, 0, this%free_surface)
        cfg%cc_step_interval = this%cc_step_interval

        if (allocated(this%rho)) then
            rho = this%rho
        else
            allocate (rho(this%nz, this%nx))
            rho = 1.0
        end if

        nr = size(this%recr)
        shot(1)%id = this%id
        shot(1)%ns = size(this%srcr)
        shot(1)%sources = c_loc(this%srcr(1))
        shot(1)%nr = nr
        shot(1)%receivers = c_loc(this%recr(1))

        st = owl_acoustic2d_create(cfg, this%vp, rho, 1_c_int, shot, this%batch)
        call check(st, 'owl_acoustic2d_create')
        nt = owl_acoustic2d_nt(this%batch)
        allocate (tr(nt*nr))
        st = owl_acoustic2d_forward(this%batch, tr, merge(1_c_int, 0_c_int, this%reconstruct))
        call check(st, 'owl_acoustic2d_forward')

        ! receiver-major [nr][nt] -> seis_p(nt, nr)
        this%seis_p = reshape(tr, [nt, nr])
        call write_su_file(trim(this%dir_synthetic)//'/shot_'//itoa(this%id)//'_seismogram_p.su', &
            this%seis_p, this%dt)

        if (.not. this%reconstruct) call this%release()

    end subroutine

    ! compute_adjoint, submodule_fwi_gradient.f90:31-199
    subroutine compute_adjoint(this)

        class(wave_solver_acoustic_iso_2d), intent(inout) :: this

        real, allocatable :: adj(:, :), gv(:), gr(:)
        real :: dt
        integer(c_int) :: st

        if (.not. c_associated(this%batch)) then
            write (*, '(a)') ' <owl_acoustic_iso_2d> adjoint needs forward with reconstruct = .true.'
            error stop 1
        end if
        call read_su_file(trim(this%dir_adjoint)//'/shot_'//itoa(this%id)//'_seismogram_p.su', adj, dt)
        allocate (gv(this%nx*this%nz), gr(this%nx*this%nz))
        st = owl_acoustic2d_adjoint(this%batch, reshape(adj, [size(adj)]), gv, gr)
        call check(st, 'owl_acoustic2d_adjoint')
        ! (nz, nx) column-major, the .grd array
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%id)//'_grad_vp.grd', &
            reshape(gv, [this%nz, this%nx]), this%dz, this%dx, this%oz, this%ox)
        call write_grd_file(trim(this%dir_working)//'/shot_'//itoa(this%id)//'_grad_rho.grd', &
            reshape(gr, [this%nz, this%nx]), this%dz, this%dx, this%oz, this%ox)
        call this%release()

    end subroutine

    function etoa(x) result(s)
        real, intent(in) :: x
        character(len=:), allocatable :: s
        character(len=32) :: buf
        write (buf, '(es15.7)') x
        s = trim(adjustl(buf))
    end function

    function itoa(i) result(s)
        integer, intent(in) :: i
        character(len=:), allocatable :: s
        character(len=32) :: buf
        write (buf, '(i0)') i
        s = trim(buf)
    end function

    ! SU file: 240-byte header (ns int16 at byte 114, dt in us int16 at 116,
    ! d1 float at 180, ntr int32 at 204), then nt float32 per trace
    ! (module_su.f90:1454-1607).
    subroutine write_su_file(path, d, dt)
        character(len=*), intent(in) :: path
        real, intent(in) :: d(:, :)        ! (nt, nr)
        real, intent(in) :: dt
        integer :: u, ir, nt, nr
        integer(1) :: hdr(240)
        integer(2) :: h2
        integer(4) :: h4
        nt = size(d, 1)
        nr = size(d, 2)
        open (newunit=u, file=path, access='stream', form='unformatted', status='replace')
        do ir = 1, nr
            hdr = 0
            h2 = int(nt, 2)
            hdr(115:116) = transfer(h2, hdr(1:2))
            h2 = int(dt*1.0e6, 2)
            hdr(117:118) = transfer(h2, hdr(1:2))
            hdr(181:184) = transfer(dt, hdr(1:4))
            h4 = nr
            hdr(205:208) = transfer(h4, hdr(1:4))
            write (u) hdr
            write (u) d(:, ir)
        end do
        close (u)
    end subroutine

    subroutine read_su_file(path, d, dt)
        character(len=*), intent(in) :: path
        real, allocatable, intent(out) :: d(:, :)
        real, intent(out) :: dt
        integer :: u, ir, nt, nr
        integer(8) :: sz
        integer(1) :: hdr(240)
        integer(2) :: h2
        open (newunit=u, file=path, access='stream', form='unformatted', status='old')
        inquire (unit=u, size=sz)
        read (u) hdr
        h2 = transfer(hdr(115:116), h2)
        nt = h2
        dt = transfer(hdr(181:184), dt)
        nr = int(sz/(240 + 4*nt))
        allocate (d(nt, nr))
        rewind (u)
        do ir = 1, nr
            read (u) hdr
            read (u) d(:, ir)
        end do
        close (u)
    end subroutine

    ! 2D grid file: 50 lines of 80 ASCII bytes, then float32 from byte 4001
    ! (module_grid.f90:317-477).
    subroutine write_grd_file(path, a, d1, d2, o1, o2)
        character(len=*), intent(in) :: path
        real, intent(in) :: a(:, :)
        real, intent(in) :: d1, d2, o1, o2
        character(len=80) :: lines(50)
        integer :: u, i
        lines = ''
        lines(1) = '# Start 4000-byte ASCII header '
        lines(2) = 'n1 = '//itoa(size(a, 1))
        lines(3) = 'd1 = '//etoa(d1)
        lines(4) = 'o1 = '//etoa(o1)
        lines(5) = 'l1 = '//etoa(o1 + (size(a, 1) - 1)*d1)
        lines(6) = 'label1 = Axis 1'
        lines(7) = 'unit1 = unitless'
        lines(8) = 'n2 = '//itoa(size(a, 2))
        lines(9) = 'd2 = '//etoa(d2)
        lines(10) = 'o2 = '//etoa(o2)
        lines(11) = 'l2 = '//etoa(o2 + (size(a, 2) - 1)*d2)
        lines(12) = 'label2 = Axis 2'
        lines(13) = 'unit2 = unitless'
        lines(50) = '# End 4000-byte ASCII header'
        do i = 1, 50
            lines(i)(79:79) = achar(13)
            lines(i)(80:80) = achar(10)
        end do
        open (newunit=u, file=path, access='stream', form='unformatted', status='replace')
        write (u) lines
        write (u, pos=4001) a
        close (u)
    end subroutine

end module owl_acoustic_iso_2d
