! Writes the values of a raw little-endian float64 file one record per value,
! the way writeSolution (src/Post_Process_mod.F90:1271-1300) writes TL and EU:
! open(form='unformatted') + one `write(unit) value` per entry.
program write_records
    implicit none
    character(len=512) :: fin, fout
    integer :: n, i, ios
    real(8), allocatable :: v(:)
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', action='read', iostat=ios)
    if (ios /= 0) stop 2
    inquire(10, size=n)
    n = n / 8
    allocate(v(n))
    read(10) v
    close(10)
    open(12, file=trim(fout), status='replace', action='write', position='rewind', form='unformatted')
    do i = 1, n
        write(12) v(i)
    end do
    close(12, status='keep')
end program write_records
