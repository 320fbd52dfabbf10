!
! Acquisition geometry and trace-gather types with the component names of
! OWL's type(source), type(receiver), type(source_receiver_geometry)
! (src/lib/module_source_receiver.f90:28-142) and type(su)
! (src/lib/module_su.f90:127-169), for callers outside OWL's tree.  The
! solver modules of this directory touch only these components, so inside
! OWL's tree this module can be replaced by one that re-exports OWL's own
! types:
!     module owl_geometry
!         use mod_source_receiver, only: source, receiver, source_receiver_geometry
!         use libflit, only: su, su_trace
!     end module
!
module owl_geometry

    implicit none

    type receiver
        real :: x = 0.0, y = 0.0, z = 0.0
        real :: aoff = 0.0
        real :: weight = 1.0
        logical :: valid = .true.
    end type receiver

    type source
        real :: x = 0.0, y = 0.0, z = 0.0
        real :: f0 = 20.0
        real :: amp = 1.0
        real :: t0 = 0.0
        ! radians (load_geometry converts the file's degrees)
        real :: polar = 0.0
        real :: azimuth = 0.0
        character(len=24) :: mechanism = 'explosion'
        character(len=24) :: wavelet = 'ricker'
        logical :: valid = .true.
    end type source

    type source_receiver_geometry
        integer :: id = 1
        integer :: ns = 1
        type(source), allocatable, dimension(:) :: srcr
        integer :: nr = 1
        type(receiver), allocatable, dimension(:) :: recr
        real :: f0factor = 1.0
    end type source_receiver_geometry

    type su_trace
        real(kind=4), allocatable, dimension(:) :: data
    end type su_trace

    type su
        integer :: nr = 0, nt = 0
        real :: dt = 0, ot = 0
        type(su_trace), allocatable, dimension(:) :: trace
    end type su

end module owl_geometry
