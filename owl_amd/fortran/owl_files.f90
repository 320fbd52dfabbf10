!
! OWL's on-disk formats for the Fortran host side: SU gathers (240-byte
! trace headers with ns at byte 115, dt in microseconds at 117, d1 at 181,
! ntr at 205, then nt float32 samples; src/lib/module_su.f90:1454-1607) and
! grid files (fifty 80-byte ASCII header lines with n, d, o, l, label, unit
! per axis, then float32 from byte 4001; src/lib/module_grid.f90:317-477).
!
module owl_files

    use iso_c_binding, only: c_int
    use owl_geometry, only: su
    use owl_amd_c, only: owl_resample_traces, owl_error_message, OWL_OK

    implicit none

    private
    public :: write_su_gather, read_su_gather, write_grd_file, itoa, resamp_gather, read_grd_2d

    interface write_grd_file
        module procedure :: write_grd_2d
        module procedure :: write_grd_3d
    end interface

contains

    function itoa(i) result(s)
        integer, intent(in) :: i
        character(len=:), allocatable :: s
        character(len=32) :: buf
        write (buf, '(i0)') i
        s = trim(buf)
    end function

    function etoa(x) result(s)
        real, intent(in) :: x
        character(len=:), allocatable :: s
        character(len=32) :: buf
        write (buf, '(es15.7)') x
        s = trim(adjustl(buf))
    end function

    subroutine write_su_gather(path, g)
        character(len=*), intent(in) :: path
        type(su), intent(in) :: g
        integer :: u, ir
        integer(1) :: hdr(240)
        integer(2) :: h2
        integer(4) :: h4
        open (newunit=u, file=path, access='stream', form='unformatted', status='replace')
        do ir = 1, g%nr
            hdr = 0
            h2 = int(g%nt, 2)
            hdr(115:116) = transfer(h2, hdr(1:2))
            h2 = int(g%dt*1.0e6, 2)
            hdr(117:118) = transfer(h2, hdr(1:2))
            hdr(181:184) = transfer(g%dt, hdr(1:4))
            h4 = g%nr
            hdr(205:208) = transfer(h4, hdr(1:4))
            write (u) hdr
            write (u) g%trace(ir)%data(1:g%nt)
        end do
        close (u)
    end subroutine

    ! su%resamp (src/lib/module_su.f90:425-468) through owl_resample_traces:
    ! each trace to nnt samples at ddt (FLIT's interp 'cubic'); a gather that
    ! already has nnt samples at ddt is left as it is.
    subroutine resamp_gather(g, nnt, ddt)
        type(su), intent(inout) :: g
        integer, intent(in) :: nnt
        real, intent(in) :: ddt
        real, allocatable :: a(:), b(:)
        integer :: ir
        integer(c_int) :: st
        if (g%nt == nnt .and. g%dt == ddt) return
        allocate (a(max(1, g%nr*g%nt)), b(max(1, g%nr*nnt)))
        do ir = 1, g%nr
            a((ir - 1)*g%nt + 1:ir*g%nt) = g%trace(ir)%data(1:g%nt)
        end do
        st = owl_resample_traces(g%nr, g%nt, g%dt, a, nnt, ddt, 0.0, b)
        if (st /= OWL_OK) then
            write (*, '(a)') ' <owl_amd> owl_resample_traces failed: '//owl_error_message()
            error stop 1
        end if
        do ir = 1, g%nr
            g%trace(ir)%data = b((ir - 1)*nnt + 1:ir*nnt)
        end do
        g%nt = nnt
        g%dt = ddt
    end subroutine

    subroutine read_su_gather(path, g)
        character(len=*), intent(in) :: path
        type(su), intent(out) :: g
        integer :: u, ir
        integer(8) :: sz
        integer(1) :: hdr(240)
        integer(2) :: h2
        open (newunit=u, file=path, access='stream', form='unformatted', status='old')
        inquire (unit=u, size=sz)
        read (u) hdr
        h2 = transfer(hdr(115:116), h2)
        g%nt = h2
        g%dt = transfer(hdr(181:184), g%dt)
        g%nr = int(sz/(240 + 4*g%nt))
        allocate (g%trace(g%nr))
        rewind (u)
        do ir = 1, g%nr
            allocate (g%trace(ir)%data(g%nt))
            read (u) hdr
            read (u) g%trace(ir)%data
        end do
        close (u)
    end subroutine

    subroutine grd_header(u, n, d, o)
        integer, intent(in) :: u
        integer, intent(in) :: n(:)
        real, intent(in) :: d(:), o(:)
        character(len=80) :: lines(50)
        integer :: i, a
        character(len=1) :: c
        lines = ''
        lines(1) = '# Start 4000-byte ASCII header '
        do a = 1, size(n)
            write (c, '(i1)') a
            lines(2 + 6*(a - 1)) = 'n'//c//' = '//itoa(n(a))
            lines(3 + 6*(a - 1)) = 'd'//c//' = '//etoa(d(a))
            lines(4 + 6*(a - 1)) = 'o'//c//' = '//etoa(o(a))
            lines(5 + 6*(a - 1)) = 'l'//c//' = '//etoa(o(a) + (n(a) - 1)*d(a))
            lines(6 + 6*(a - 1)) = 'label'//c//' = Axis '//c
            lines(7 + 6*(a - 1)) = 'unit'//c//' = unitless'
        end do
        lines(50) = '# End 4000-byte ASCII header'
        do i = 1, 50
            lines(i)(79:79) = achar(13)
            lines(i)(80:80) = achar(10)
        end do
        write (u) lines
    end subroutine

    subroutine write_grd_2d(path, a, d1, d2, o1, o2)
        character(len=*), intent(in) :: path
        real, intent(in) :: a(:, :)
        real, intent(in) :: d1, d2, o1, o2
        integer :: u
        open (newunit=u, file=path, access='stream', form='unformatted', status='replace')
        call grd_header(u, shape(a), [d1, d2], [o1, o2])
        write (u, pos=4001) a
        close (u)
    end subroutine

    ! the float32 body of a 2D grid file into a, whose shape the caller knows
    subroutine read_grd_2d(path, a)
        character(len=*), intent(in) :: path
        real, intent(out) :: a(:, :)
        integer :: u
        open (newunit=u, file=path, access='stream', form='unformatted', status='old', action='read')
        read (u, pos=4001) a
        close (u)
    end subroutine

    subroutine write_grd_3d(path, a, d1, d2, d3, o1, o2, o3)
        character(len=*), intent(in) :: path
        real, intent(in) :: a(:, :, :)
        real, intent(in) :: d1, d2, d3, o1, o2, o3
        integer :: u
        open (newunit=u, file=path, access='stream', form='unformatted', status='replace')
        call grd_header(u, shape(a), [d1, d2, d3], [o1, o2, o3])
        write (u, pos=4001) a
        close (u)
    end subroutine

end module owl_files
