!
! Drives the drop-in modules acoustic_iso_3d and elastic_tti_2d
! (owl_amd/fortran) the way OWL's callers do: fill the solver's components
! for one shot, call forward (with reconstruct), then adjoint with the
! adjoint sources already in dir_adjoint.
!   owl_solvers_3d_tti 3d  <job>   job: nx ny nz d dt tmax pml
!                                       vp_file / rho_file  ((nz, ny, nx) float32)
!                                       id sx sy sz f0 / nr / nr lines "rx ry rz"
!                                       dir_synthetic / dir_adjoint / dir_working
!   owl_solvers_3d_tti tti <job>   job: nx nz d dt tmax pml (optional last line: the free surface)
!                                       anisotropy_type
!                                       vp vs rho epsilon delta theta files ((nz, nx))
!                                       id sx sz f0 / nr / nr lines "rx rz"
!                                       dir_synthetic / dir_adjoint / dir_working
!
program owl_solvers_3d_tti

    use owl_geometry
    use acoustic_iso_3d
    use elastic_tti_2d

    implicit none

    character(len=1024) :: mode, job
    call get_command_argument(1, mode)
    call get_command_argument(2, job)
    if (trim(mode) == '3d') then
        call run3d(trim(job))
    else
        call runtti(trim(job))
    end if

contains

    subroutine load3(path, a)
        character(len=*), intent(in) :: path
        real, intent(out) :: a(:, :, :)
        integer :: v
        open (newunit=v, file=path, access='stream', form='unformatted', status='old')
        read (v) a
        close (v)
    end subroutine

    subroutine load2(path, a)
        character(len=*), intent(in) :: path
        real, intent(out) :: a(:, :)
        integer :: v
        open (newunit=v, file=path, access='stream', form='unformatted', status='old')
        read (v) a
        close (v)
    end subroutine

    subroutine run3d(path)
        character(len=*), intent(in) :: path
        type(wave_solver_acoustic_iso_3d) :: s
        character(len=1024) :: fvp, frho
        integer :: u, nr, ir
        real :: d
        open (newunit=u, file=path, status='old', action='read')
        read (u, *) s%nx, s%ny, s%nz, d, s%dt, s%tmax, s%pml
        s%dx = d
        s%dy = d
        s%dz = d
        read (u, '(a)') fvp
        read (u, '(a)') frho
        allocate (s%gmtr%srcr(1))
        read (u, *) s%gmtr%id, s%gmtr%srcr(1)%x, s%gmtr%srcr(1)%y, s%gmtr%srcr(1)%z, s%gmtr%srcr(1)%f0
        read (u, *) nr
        allocate (s%gmtr%recr(nr))
        s%gmtr%nr = nr
        do ir = 1, nr
            read (u, *) s%gmtr%recr(ir)%x, s%gmtr%recr(ir)%y, s%gmtr%recr(ir)%z
        end do
        read (u, '(a)') s%dir_synthetic
        read (u, '(a)') s%dir_adjoint
        read (u, '(a)') s%dir_working
        close (u)
        allocate (s%vp(s%nz, s%ny, s%nx), s%rho(s%nz, s%ny, s%nx))
        call load3(trim(fvp), s%vp)
        call load3(trim(frho), s%rho)
        s%data_dt = s%dt
        s%data_tmax = s%tmax
        s%reconstruct = .true.
        call s%forward
        call s%adjoint
        write (*, '(a,i0,a,i0)') 'acoustic_iso_3d: nt = ', s%seis_p%nt, ', nr = ', s%seis_p%nr
    end subroutine

    subroutine runtti(path)
        character(len=*), intent(in) :: path
        type(wave_solver_elastic_tti_2d) :: s
        character(len=1024) :: f(6), topo
        character(len=2048) :: line
        integer :: u, nr, ir, k, ios, fs1, fs2
        real :: d, refine, dzmax
        open (newunit=u, file=path, status='old', action='read')
        read (u, *) s%nx, s%nz, d, s%dt, s%tmax, s%pml
        s%dx = d
        s%dz = d
        read (u, '(a)') s%anisotropy_type
        do k = 1, 6
            read (u, '(a)') f(k)
        end do
        allocate (s%gmtr%srcr(1))
        read (u, *) s%gmtr%id, s%gmtr%srcr(1)%x, s%gmtr%srcr(1)%z, s%gmtr%srcr(1)%f0
        s%gmtr%srcr(1)%mechanism = 'force'
        read (u, *) nr
        allocate (s%gmtr%recr(nr))
        s%gmtr%nr = nr
        do ir = 1, nr
            read (u, *) s%gmtr%recr(ir)%x, s%gmtr%recr(ir)%z
        end do
        read (u, '(a)') s%dir_synthetic
        read (u, '(a)') s%dir_adjoint
        read (u, '(a)') s%dir_working
        ! optional: free surface  <file_topo or '-'>  <refine>  <dz_max>  <src, rec depths from surface 0|1>
        ! (the path is the first blank-separated token; list-directed input
        ! would stop at its slashes)
        read (u, '(a)', iostat=ios) line
        if (ios == 0 .and. len_trim(line) > 0) then
            line = adjustl(line)
            k = index(line, ' ')
            topo = line(1:k - 1)
            read (line(k:), *, iostat=ios) refine, dzmax, fs1, fs2
        else
            ios = 1
        end if
        if (ios == 0) then
            s%free_surface = .true.
            if (topo /= '-') s%file_topo = topo
            s%free_surface_dz_refine = refine
            s%dz_max = dzmax
            s%measure_source_depth_from_surface = fs1 /= 0
            s%measure_receiver_depth_from_surface = fs2 /= 0
        end if
        close (u)
        allocate (s%vp(s%nz, s%nx), s%vs(s%nz, s%nx), s%rho(s%nz, s%nx))
        allocate (s%tieps(s%nz, s%nx), s%tidel(s%nz, s%nx), s%tithe(s%nz, s%nx))
        call load2(trim(f(1)), s%vp)
        call load2(trim(f(2)), s%vs)
        call load2(trim(f(3)), s%rho)
        call load2(trim(f(4)), s%tieps)
        call load2(trim(f(5)), s%tidel)
        call load2(trim(f(6)), s%tithe)
        s%data_dt = s%dt
        s%data_tmax = s%tmax
        s%reconstruct = .true.
        call s%forward
        call s%adjoint
        write (*, '(a,i0,a,i0)') 'elastic_tti_2d: nt = ', s%seis_vx%nt, ', nr = ', s%seis_vx%nr
    end subroutine

end program owl_solvers_3d_tti
