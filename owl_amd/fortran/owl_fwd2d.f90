!
! Fortran driver of the 2D acoustic solver on the GPU: reads a model and a
! one-shot geometry from a small text file, runs forward (and, with
! reconstruct, the adjoint from dir_adjoint), writing OWL's SU/grd files.
!   owl_fwd2d <job.txt>
! job.txt: nx nz dx dz dt tmax pml fdhalf reconstruct
!          vp_file                      (float32, (nz, nx) column-major)
!          rho_file
!          sx sz f0 amp
!          nr
!          rx rz                        (nr lines)
!          dir_synthetic
!          dir_adjoint                  (these two read when reconstruct = 1)
!          dir_working
! Paths take a line each (list-directed input would stop at '/').
!
program owl_fwd2d

    use acoustic_iso_2d
    use owl_amd_c

    implicit none

    type(wave_solver_acoustic_iso_2d) :: solver
    character(len=1024) :: job, f_vp, f_rho
    integer :: u, ir, nr, recon
    real :: sx, sz, f0, amp

    call get_command_argument(1, job)
    open (newunit=u, file=trim(job), status='old', action='read')
    read (u, *) solver%nx, solver%nz, solver%dx, solver%dz, solver%dt, solver%tmax, solver%pml, &
        solver%fdhalf, recon
    read (u, '(a)') f_vp
    read (u, '(a)') f_rho
    read (u, *) sx, sz, f0, amp
    read (u, *) nr
    allocate (solver%gmtr%srcr(1), solver%gmtr%recr(nr))
    solver%gmtr%id = 1
    solver%gmtr%ns = 1
    solver%gmtr%nr = nr
    solver%gmtr%srcr(1)%x = sx
    solver%gmtr%srcr(1)%z = sz
    solver%gmtr%srcr(1)%f0 = f0
    solver%gmtr%srcr(1)%amp = amp
    do ir = 1, nr
        read (u, *) solver%gmtr%recr(ir)%x, solver%gmtr%recr(ir)%z
    end do
    read (u, '(a)') solver%dir_synthetic
    if (recon /= 0) then
        read (u, '(a)') solver%dir_adjoint
        read (u, '(a)') solver%dir_working
    end if
    close (u)

    allocate (solver%vp(solver%nz, solver%nx), solver%rho(solver%nz, solver%nx))
    call load(trim(f_vp), solver%vp)
    call load(trim(f_rho), solver%rho)
    solver%data_dt = solver%dt
    solver%data_tmax = solver%tmax
    solver%reconstruct = recon /= 0

    call solver%forward
    if (recon /= 0) call solver%adjoint
    write (*, '(a,i0,a,i0)') 'owl_fwd2d: nt = ', solver%seis_p%nt, ', nr = ', solver%seis_p%nr

contains

    subroutine load(path, a)
        character(len=*), intent(in) :: path
        real, intent(out) :: a(:, :)
        integer :: v
        open (newunit=v, file=path, access='stream', form='unformatted', status='old')
        read (v) a
        close (v)
    end subroutine

end program owl_fwd2d
