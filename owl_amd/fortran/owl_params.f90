!
! Parameter-file reader with the names and calling convention of FLIT's
! readpar_* family, as OWL's read_parameters and set_regular_space call it
! (src/lib/module_parameters.f90:316-703, src/lib/module_model.f90:72-160):
!     call readpar_float(file_parameter, 'dt', dt, 0.0)
!     call readpar_nstring(file_parameter, 'model_name', model_name, [''], required=.true.)
! FLIT is not part of the OWL tree; the file format is taken from OWL's
! examples (example/*/param_*.rb): one "key = value" per line, lists
! separated by commas, logicals written y/n.  Rules this module applies:
!   * the key is the text before the first '=', blanks trimmed, matched
!     exactly (case-sensitive); text after a '#' is a comment;
!   * when a key appears more than once the last line wins;
!   * a missing key takes the default, or stops with a message when
!     required = .true.; a value that does not parse stops with a message;
!   * logicals: y, yes, t, true, .true., 1 (any case) are true; n, no, f,
!     false, .false., 0 are false;
!   * list items are separated by commas and/or blanks.
!
module owl_params

    implicit none

    private
    public :: readpar_string, readpar_int, readpar_float, readpar_logical
    public :: readpar_nint, readpar_nfloat, readpar_nstring, split_items, lowercase

    integer, parameter :: line_len = 4096

contains

    ! value text of key, found = .false. when absent
    subroutine find_key(file, key, val, found)
        character(len=*), intent(in) :: file, key
        character(len=:), allocatable, intent(out) :: val
        logical, intent(out) :: found
        character(len=line_len) :: line
        integer :: u, ios, k, c
        logical :: exists
        found = .false.
        val = ''
        inquire (file=trim(file), exist=exists)
        if (.not. exists) then
            write (*, '(a)') ' <readpar> Error: parameter file '//trim(file)//' not found'
            error stop 1
        end if
        open (newunit=u, file=trim(file), status='old', action='read')
        do
            read (u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            c = index(line, '#')
            if (c > 0) line(c:) = ' '
            k = index(line, '=')
            if (k <= 1) cycle
            if (trim(adjustl(line(1:k - 1))) == key) then
                val = trim(adjustl(line(k + 1:)))
                found = .true.
            end if
        end do
        close (u)
    end subroutine

    subroutine missing(key, required)
        character(len=*), intent(in) :: key
        logical, intent(in), optional :: required
        if (present(required)) then
            if (required) then
                write (*, '(a)') ' <readpar> Error: required parameter '//key//' is not given'
                error stop 1
            end if
        end if
    end subroutine

    subroutine bad_value(key, val)
        character(len=*), intent(in) :: key, val
        write (*, '(a)') ' <readpar> Error: cannot read '//key//' = '//val
        error stop 1
    end subroutine

    function lowercase(s) result(t)
        character(len=*), intent(in) :: s
        character(len=len(s)) :: t
        integer :: k, c
        t = s
        do k = 1, len(s)
            c = iachar(s(k:k))
            if (c >= iachar('A') .and. c <= iachar('Z')) t(k:k) = achar(c + 32)
        end do
    end function

    ! items of a comma- and/or blank-separated list
    subroutine split_items(val, items)
        character(len=*), intent(in) :: val
        character(len=1024), allocatable, intent(out) :: items(:)
        character(len=1024) :: buf(1024)
        integer :: k, n, b
        n = 0
        b = 0
        do k = 1, len_trim(val) + 1
            if (k > len_trim(val)) then
                call push(k)
            else if (val(k:k) == ',' .or. val(k:k) == ' ' .or. val(k:k) == achar(9)) then
                call push(k)
            else if (b == 0) then
                b = k
            end if
        end do
        allocate (items(n))
        items = buf(1:n)
    contains
        subroutine push(e)
            integer, intent(in) :: e
            if (b > 0) then
                if (n == size(buf)) then
                    write (*, '(a)') ' <readpar> Error: a list holds more than 1024 items'
                    error stop 1
                end if
                n = n + 1
                buf(n) = val(b:e - 1)
                b = 0
            end if
        end subroutine
    end subroutine

    subroutine readpar_string(file, key, var, default, required)
        character(len=*), intent(in) :: file, key, default
        character(len=*), intent(out) :: var
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        logical :: found
        call find_key(file, key, val, found)
        if (found) then
            var = val
        else
            call missing(key, required)
            var = default
        end if
    end subroutine

    subroutine readpar_int(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        integer, intent(out) :: var
        integer, intent(in) :: default
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        logical :: found
        integer :: ios
        real :: r
        call find_key(file, key, val, found)
        if (.not. found) then
            call missing(key, required)
            var = default
            return
        end if
        read (val, *, iostat=ios) var
        if (ios /= 0) then
            ! "1e3" and "100.0" as integers
            read (val, *, iostat=ios) r
            if (ios /= 0) call bad_value(key, val)
            var = nint(r)
        end if
    end subroutine

    subroutine readpar_float(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        real, intent(out) :: var
        real, intent(in) :: default
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        logical :: found
        integer :: ios
        call find_key(file, key, val, found)
        if (.not. found) then
            call missing(key, required)
            var = default
            return
        end if
        read (val, *, iostat=ios) var
        if (ios /= 0) call bad_value(key, val)
    end subroutine

    subroutine readpar_logical(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        logical, intent(out) :: var
        logical, intent(in) :: default
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        logical :: found
        call find_key(file, key, val, found)
        if (.not. found) then
            call missing(key, required)
            var = default
            return
        end if
        select case (trim(lowercase(val)))
            case ('y', 'yes', 't', 'true', '.true.', '1')
                var = .true.
            case ('n', 'no', 'f', 'false', '.false.', '0')
                var = .false.
            case default
                call bad_value(key, val)
        end select
    end subroutine

    subroutine readpar_nstring(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        character(len=*), allocatable, intent(inout) :: var(:)
        character(len=*), intent(in) :: default(:)
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        character(len=1024), allocatable :: items(:)
        logical :: found
        integer :: k
        if (allocated(var)) deallocate (var)
        call find_key(file, key, val, found)
        if (found) call split_items(val, items)
        if (.not. found) then
            call missing(key, required)
            allocate (var(size(default)))
            do k = 1, size(default)
                var(k) = default(k)
            end do
            return
        end if
        if (size(items) == 0) then
            allocate (var(1))
            var(1) = ''
            return
        end if
        allocate (var(size(items)))
        do k = 1, size(items)
            var(k) = items(k)
        end do
    end subroutine

    subroutine readpar_nfloat(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        real, allocatable, intent(inout) :: var(:)
        real, intent(in) :: default(:)
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        character(len=1024), allocatable :: items(:)
        logical :: found
        integer :: k, ios
        if (allocated(var)) deallocate (var)
        call find_key(file, key, val, found)
        if (.not. found) then
            call missing(key, required)
            var = default
            return
        end if
        call split_items(val, items)
        allocate (var(size(items)))
        do k = 1, size(items)
            read (items(k), *, iostat=ios) var(k)
            if (ios /= 0) call bad_value(key, val)
        end do
    end subroutine

    subroutine readpar_nint(file, key, var, default, required)
        character(len=*), intent(in) :: file, key
        integer, allocatable, intent(inout) :: var(:)
        integer, intent(in) :: default(:)
        logical, intent(in), optional :: required
        character(len=:), allocatable :: val
        character(len=1024), allocatable :: items(:)
        logical :: found
        integer :: k, ios
        if (allocated(var)) deallocate (var)
        call find_key(file, key, val, found)
        if (.not. found) then
            call missing(key, required)
            var = default
            return
        end if
        call split_items(val, items)
        allocate (var(size(items)))
        do k = 1, size(items)
            read (items(k), *, iostat=ios) var(k)
            if (ios /= 0) call bad_value(key, val)
        end do
    end subroutine

end module owl_params
