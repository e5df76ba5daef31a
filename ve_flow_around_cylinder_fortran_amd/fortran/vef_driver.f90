! vef_driver — a Fortran caller of the assembler, shaped like the reference:
! the mesh comes from a <name>_MESH.DTA file read as readMeshFile does
! (src/FEM_mod.F90:172-256), the solution as Sol%TL(Nnodes,Neq) and Sol%EU,
! and one assembly replaces initializeLSObjects + setEquationsBulk +
! setEquationsEdge.
!
! usage: vef_driver <name> <NRP|MNR>
!   reads  <name>_MESH.DTA, <name>_OUTLET.DTA ("n" then n lines "element edge",
!          1-based elements), <name>_PARAMS.DTA ("ReN BvN VeN DfN WiN0 model"),
!          <name>_STATE.bin (TL(Nnodes,6) column-major, then EU; raw float64)
!   writes <name>_B.bin (b, N doubles) and, for NRP, <name>_VALUES.bin
!          (nnz doubles) and <name>_PATTERN.bin (rowptr int64, colind int32)
program vef_driver

    use iso_c_binding
    use vef_assembly_mod

    implicit none

    character(len=512) :: name, flag_arg
    integer :: Nnodes, Nel_total, Nel, i, j, dum, etype, ierr, n_edges, u
    integer(c_int64_t) :: N, nnz
    real(c_double), allocatable :: coords(:), TL(:,:), EU(:), b(:), values(:)
    real(c_double) :: xyz(3)
    integer(c_int32_t), allocatable :: conn_all(:,:), conn(:), edges(:), colind(:)
    integer(c_int64_t), allocatable :: rowptr(:)
    integer :: row(9)
    type(vef_params) :: prm
    type(c_ptr) :: h
    integer(c_int) :: rc
    real(c_double) :: p6(6)

    call get_command_argument(1, name)
    call get_command_argument(2, flag_arg)

    open(8, file=trim(name)//'_MESH.DTA', action='read')
    read(8,*) Nnodes, Nel_total
    allocate( coords(2*Nnodes) )
    do i = 1, Nnodes
        read(8,*) dum, xyz(1), xyz(2), xyz(3)
        coords(2*i-1) = xyz(1) ; coords(2*i) = xyz(2)
    end do
    allocate( conn_all(9, Nel_total) )
    Nel = 0
    do i = 1, Nel_total
        read(8,*) dum, etype, row(1:mod(etype,100))
        if (etype/100 == 2) then
            Nel = Nel + 1
            conn_all(:,Nel) = row(:) - 1
        end if
    end do
    close(8)
    allocate( conn(9*Nel) )
    conn(:) = reshape(conn_all(:,1:Nel), [9*Nel])

    open(9, file=trim(name)//'_OUTLET.DTA', action='read')
    read(9,*) n_edges
    allocate( edges(2*max(n_edges,1)) )
    do i = 1, n_edges
        read(9,*) edges(2*i-1), edges(2*i)
        edges(2*i-1) = edges(2*i-1) - 1
    end do
    close(9)

    open(10, file=trim(name)//'_PARAMS.DTA', action='read')
    read(10,*) p6
    close(10)
    prm = vef_params(p6(1), p6(2), p6(3), p6(4), p6(5), int(p6(6), c_int32_t), 0_c_int32_t)

    rc = vef_create(int(Nnodes, c_int32_t), int(Nel, c_int32_t), coords, conn, &
        int(n_edges, c_int32_t), edges, prm, h)
    if (rc /= 0) then
        write(*,'(a)') 'vef_create: '//vef_last_error(c_null_ptr)
        stop 1
    end if

    N = vef_num_rows(h)
    nnz = vef_nnz(h)
    allocate( TL(Nnodes,6), EU(1), b(N) )
    open(newunit=u, file=trim(name)//'_STATE.bin', access='stream', form='unformatted', action='read')
    read(u) TL, EU
    close(u)

    if (trim(flag_arg) == 'NRP') then
        allocate( values(nnz) )
    else
        allocate( values(1) )
    end if
    call vef_assemble_tl(h, TL, EU, trim(flag_arg), b, values, ierr)
    if (ierr /= 0) then
        write(*,'(a)') 'vef_assemble: '//vef_last_error(h)
        stop 1
    end if

    open(newunit=u, file=trim(name)//'_B.bin', access='stream', form='unformatted', action='write', &
        status='replace')
    write(u) b
    close(u)
    if (trim(flag_arg) == 'NRP') then
        allocate( rowptr(N+1), colind(nnz) )
        rc = vef_get_pattern(h, rowptr, colind)
        open(newunit=u, file=trim(name)//'_VALUES.bin', access='stream', form='unformatted', &
            action='write', status='replace')
        write(u) values
        close(u)
        open(newunit=u, file=trim(name)//'_PATTERN.bin', access='stream', form='unformatted', &
            action='write', status='replace')
        write(u) rowptr, colind
        close(u)
    end if

    call vef_destroy(h)
    write(*,'(a,i0,a,i0)') 'vef_driver: N = ', N, ', nnz = ', nnz

end program vef_driver
