! vef_assembly_mod — Fortran 2008 iso_c_binding interface to the MI355X
! assembler (include/vef.h).  A Newton driver written like the reference's
! assemblePetscObjects (src/Newton_Raphson_mod.F90:281-400) replaces
!
!     call initializeLSObjects(Flag_NR, LS)
!     call setEquationsBulk(Problem, FE, Mesh, Boundary, GaussInt, Sol, basis_p, Flag_NR, Elem, LS)
!     call setEquationsEdge(Problem, FE, Mesh, Boundary(ibnd), Boundary, GaussInt, Sol, basis_p, Flag_NR, Elem, LS)
!
! by one call of vef_assemble_tl (host arrays) or vef_assemble (device arrays),
! and reads the CSR pattern once with vef_get_pattern to set up its matrix.
module vef_assembly_mod

    use iso_c_binding

    implicit none

    integer(c_int32_t), parameter :: VEF_MODEL_LPTT = 0, VEF_MODEL_GIESEKUS = 1
    integer(c_int32_t), parameter :: VEF_RES_ONLY = 0, VEF_RES_JAC = 1, VEF_RES_JAC_FP32 = 2

    type, bind(C) :: vef_params
        real(c_double) :: ReN, BvN, VeN, DfN, WiN0
        integer(c_int32_t) :: model
        integer(c_int32_t) :: reserved
    end type vef_params

    interface

        integer(c_int) function vef_create(n_nodes, n_elements, coords, connectivity, &
            n_outlet_edges, outlet_edges, params, handle) bind(C, name='vef_create')
            import :: c_int, c_int32_t, c_double, c_ptr, vef_params
            integer(c_int32_t), value :: n_nodes, n_elements, n_outlet_edges
            real(c_double), intent(in) :: coords(*)
            integer(c_int32_t), intent(in) :: connectivity(*), outlet_edges(*)
            type(vef_params), intent(in) :: params
            type(c_ptr), intent(out) :: handle
        end function vef_create

        integer(c_int) function vef_set_params(handle, params) bind(C, name='vef_set_params')
            import :: c_int, c_ptr, vef_params
            type(c_ptr), value :: handle
            type(vef_params), intent(in) :: params
        end function vef_set_params

        integer(c_int64_t) function vef_num_rows(handle) bind(C, name='vef_num_rows')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: handle
        end function vef_num_rows

        integer(c_int64_t) function vef_nnz(handle) bind(C, name='vef_nnz')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: handle
        end function vef_nnz

        integer(c_int) function vef_get_pattern(handle, rowptr, colind) bind(C, name='vef_get_pattern')
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: handle
            integer(c_int64_t), intent(out) :: rowptr(*)
            integer(c_int32_t), intent(out) :: colind(*)
        end function vef_get_pattern

        integer(c_int) function vef_assemble(handle, x, flag, b, values, stream) bind(C, name='vef_assemble')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: handle, x, b, values, stream
            integer(c_int32_t), value :: flag
        end function vef_assemble

        ! Linear stability (replaces assembleSlepcObjects for Cylinder_Stability_2D):
        ! A and B in the node-block pattern, device arrays; x the base state.
        integer(c_int64_t) function vef_stability_nnz(handle) bind(C, name='vef_stability_nnz')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: handle
        end function vef_stability_nnz

        integer(c_int) function vef_get_stability_pattern(handle, rowptr, colind) &
            bind(C, name='vef_get_stability_pattern')
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: handle
            integer(c_int64_t), intent(out) :: rowptr(*)
            integer(c_int32_t), intent(out) :: colind(*)
        end function vef_get_stability_pattern

        integer(c_int) function vef_set_stability_dirichlet(handle, n_rows, rows) &
            bind(C, name='vef_set_stability_dirichlet')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: handle
            integer(c_int64_t), value :: n_rows
            integer(c_int64_t), intent(in) :: rows(*)
        end function vef_set_stability_dirichlet

        integer(c_int) function vef_stability_assemble(handle, x, A, B, stream) bind(C, name='vef_stability_assemble')
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, x, A, B, stream
        end function vef_stability_assemble

        integer(c_int) function vef_assemble_host(handle, x, flag, b, values) bind(C, name='vef_assemble_host')
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: handle
            real(c_double), intent(in) :: x(*)
            integer(c_int32_t), value :: flag
            real(c_double), intent(out) :: b(*)
            real(c_double), intent(out) :: values(*)
        end function vef_assemble_host

        type(c_ptr) function vef_last_error_c(handle) bind(C, name='vef_last_error')
            import :: c_ptr
            type(c_ptr), value :: handle
        end function vef_last_error_c

        subroutine vef_destroy(handle) bind(C, name='vef_destroy')
            import :: c_ptr
            type(c_ptr), value :: handle
        end subroutine vef_destroy

    end interface

    contains

    ! Sol%TL(Nnodes, Neq) and Sol%EU(Nex) (src/Global_Variables_mod.F90:28-40)
    ! -> the global vector x(6*(g-1)+q) = TL(g,q), x(N) = EU(Nex), then one
    ! assembly.  Flag_NR is 'NRP' (residual and Jacobian), 'MNR', or 'NRP32'
    ! (the FP64 residual with the FP32 analytic Jacobian, for an inexact Newton).
    subroutine vef_assemble_tl(handle, TL, EU, Flag_NR, b, values, ierr)

        type(c_ptr), intent(in) :: handle
        real(c_double), dimension(:,:), intent(in) :: TL
        real(c_double), dimension(:), intent(in) :: EU
        character(*), intent(in) :: Flag_NR
        real(c_double), dimension(:), intent(out) :: b
        real(c_double), dimension(:), intent(inout) :: values
        integer, intent(out) :: ierr

        real(c_double), allocatable :: x(:)
        integer :: g, q, Nnodes, Neq
        integer(c_int32_t) :: flag

        Nnodes = size(TL,1) ; Neq = size(TL,2)
        allocate( x(Nnodes*Neq+1) )
        do g = 1, Nnodes
            do q = 1, Neq
                x(Neq*(g-1)+q) = TL(g,q)
            end do
        end do
        x(Nnodes*Neq+1) = EU(size(EU))

        select case (Flag_NR)
        case ('NRP')
            flag = VEF_RES_JAC
        case ('MNR')
            flag = VEF_RES_ONLY
        case ('NRP32')
            flag = VEF_RES_JAC_FP32
        case default
            ierr = -1
            return
        end select

        ierr = vef_assemble_host(handle, x, flag, b, values)

    end subroutine vef_assemble_tl

    function vef_last_error(handle) result(msg)

        type(c_ptr), intent(in) :: handle
        character(len=:), allocatable :: msg

        character(kind=c_char), pointer :: p(:)
        type(c_ptr) :: cp
        integer :: n

        cp = vef_last_error_c(handle)
        call c_f_pointer(cp, p, [4096])
        n = 0
        do while (n < 4096)
            if (p(n+1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(p(1:n), msg)

    end function vef_last_error

end module vef_assembly_mod
