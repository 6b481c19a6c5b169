!==========================================================================================
! ceos_gpu_interface: ISO_C_BINDING declarations of libceos_gpu.so (include/ceos_gpu.h)
! for the reference's Fortran driver.
!
! The element loops of the reference (SolveConstitutiveModel, InternalForce,
! TangentStiffnessMatrix, UpdateMeshCoordinates -- FEM/Modules/ModuleInterfaces.f90:132-270)
! are replaced by these calls; the Newton loop stays as it is.  Dirichlet BCs and the PARDISO
! solve either stay too, or move to the device (ceos_bc_set, ceos_residual_bc, ceos_linear_solve).
! Arrays are passed by reference with the reference's own layouts:
!   U, Fint(nDOF)           DOF = 3*(node-1)+k
!   Kg%Val(nnz), Kg%RowMap, Kg%Col   1-based upper CSR (ModGlobalSparseMatrix.f90:18-25)
!   conn(8, nElem)          hex8 node ids (ModElementHexa8.f90 node order)
!==========================================================================================
module ceos_gpu_interface
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: CEOS_OK = 0, CEOS_ERR_DETJ = -1, CEOS_ERR_LOCAL_NR = -2, &
                                 CEOS_ERR_SLOT = -3, CEOS_ERR_ARG = -4, CEOS_ERR_HIP = -5
    integer(c_int), parameter :: CEOS_K_FP64 = 0, CEOS_K_FP32_STORED = 1
    integer(c_int), parameter :: CEOS_TANGENT_ANALYTIC = 0, CEOS_TANGENT_NUMERICAL = 1
    integer(c_int), parameter :: CEOS_STATE_STRIDE = 20
    integer(c_int), parameter :: CEOS_ELEM_FULL_INTEGRATION = 1, CEOS_ELEM_MEAN_DILATATION = 2
    integer(c_int), parameter :: CEOS_NORM_L2 = 1, CEOS_NORM_MAX_ABS = 2
    integer(c_int), parameter :: CEOS_SOLVER_AUTO = 0, CEOS_SOLVER_PCG = 1, CEOS_SOLVER_MINRES = 2

    ! mirrors ClassStatus (Tools/ModStatus.f90:3-16)
    type, bind(C) :: ceos_status
        integer(c_int) :: error
        integer(c_int) :: error_id
        integer(c_int) :: n_failed
        character(kind=c_char) :: description(256)
    end type

    interface
        integer(c_int) function ceos_mesh_create(nn, X, ne, npe, conn, mat_id, h) bind(C, name="ceos_mesh_create")
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: nn, ne, npe
            real(c_double), intent(in) :: X(*)
            integer(c_int), intent(in) :: conn(*), mat_id(*)
            type(c_ptr), intent(out) :: h
        end function

        subroutine ceos_mesh_destroy(h) bind(C, name="ceos_mesh_destroy")
            import :: c_ptr
            type(c_ptr), value :: h
        end subroutine

        integer(c_int) function ceos_material_set(h, mat_id, model, props, nprops) bind(C, name="ceos_material_set")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: mat_id, model, nprops
            real(c_double), intent(in) :: props(*)
        end function

        integer(c_int) function ceos_set_analysis(h, nl_analysis, elem_tech) bind(C, name="ceos_set_analysis")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: nl_analysis, elem_tech
        end function

        integer(c_int) function ceos_set_precision(h, mode) bind(C, name="ceos_set_precision")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: mode
        end function

        integer(c_int) function ceos_csr_pattern(h, upper, n, nnz, rowmap, col) bind(C, name="ceos_csr_pattern")
            import :: c_int, c_long_long, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: upper
            integer(c_int), intent(out) :: n
            integer(c_long_long), intent(out) :: nnz
            type(c_ptr), value :: rowmap, col
        end function

        integer(c_int) function ceos_solve_constitutive(h, time, U, st) bind(C, name="ceos_solve_constitutive")
            import :: c_int, c_double, c_ptr, ceos_status
            type(c_ptr), value :: h
            real(c_double), value :: time
            real(c_double), intent(in) :: U(*)
            type(ceos_status), intent(out) :: st
        end function

        integer(c_int) function ceos_internal_force(h, Fint, st) bind(C, name="ceos_internal_force")
            import :: c_int, c_double, c_ptr, ceos_status
            type(c_ptr), value :: h
            real(c_double), intent(out) :: Fint(*)
            type(ceos_status), intent(out) :: st
        end function

        integer(c_int) function ceos_tangent(h, Kval) bind(C, name="ceos_tangent")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: Kval(*)
        end function

        integer(c_int) function ceos_tangent_f32(h, Kval) bind(C, name="ceos_tangent_f32")
            import :: c_int, c_float, c_ptr
            type(c_ptr), value :: h
            real(c_float), intent(out) :: Kval(*)
        end function

        integer(c_int) function ceos_set_tangent_mode(h, mode) bind(C, name="ceos_set_tangent_mode")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: mode
        end function

        integer(c_int) function ceos_get_state(h, which, state) bind(C, name="ceos_get_state")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), intent(out) :: state(*)
        end function

        integer(c_int) function ceos_set_state(h, which, state) bind(C, name="ceos_set_state")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), intent(in) :: state(*)
        end function

        integer(c_int) function ceos_update_mesh(h, U) bind(C, name="ceos_update_mesh")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: U(*)
        end function

        integer(c_int) function ceos_switch_converged_state(h) bind(C, name="ceos_switch_converged_state")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function ceos_failed_elements(h, ids, cap) bind(C, name="ceos_failed_elements")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), intent(out) :: ids(*)
            integer(c_int), value :: cap
        end function

        integer(c_int) function ceos_get_gp_stress(h, sig) bind(C, name="ceos_get_gp_stress")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: sig(*)
        end function

        integer(c_int) function ceos_homogenize(h, P, F) bind(C, name="ceos_homogenize")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(out) :: P(9), F(9)
        end function

        integer(c_int) function ceos_state_save(h, path) bind(C, name="ceos_state_save")
            import :: c_int, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
        end function

        integer(c_int) function ceos_state_load(h, path) bind(C, name="ceos_state_load")
            import :: c_int, c_char, c_ptr
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
        end function

        ! Newton system on the device: constrained DOFs (1-based), ApplyBoundaryConditionsNEW +
        ! the residual norm, and the linear solve that replaces the PARDISO call
        integer(c_int) function ceos_bc_set(h, n_presc, presc_dof, n_fixed, fixed_dof) bind(C, name="ceos_bc_set")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: n_presc, n_fixed
            integer(c_int), intent(in) :: presc_dof(*), fixed_dof(*)
        end function

        integer(c_int) function ceos_residual_bc(h, it, Fext, Ubar, R, norm_type, normR) bind(C, name="ceos_residual_bc")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: it, norm_type
            real(c_double), intent(in) :: Fext(*), Ubar(*)
            real(c_double), intent(out) :: R(*)
            real(c_double), intent(out) :: normR
        end function

        integer(c_int) function ceos_set_linear_solver(h, method) bind(C, name="ceos_set_linear_solver")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: method
        end function

        integer(c_int) function ceos_linear_solve(h, b, x, rtol, maxit, iters, relres, flag) &
                bind(C, name="ceos_linear_solve")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double), intent(in) :: b(*)
            real(c_double), intent(out) :: x(*)
            real(c_double), value :: rtol
            integer(c_int), value :: maxit
            integer(c_int), intent(out) :: iters, flag
            real(c_double), intent(out) :: relres
        end function
    end interface

contains

    ! Upper CSR into caller-owned arrays shaped like ClassGlobalSparseMatrix (RowMap, Col).
    integer(c_int) function ceos_csr_upper(h, RowMap, Col) result(rc)
        type(c_ptr), intent(in) :: h
        integer(c_int), allocatable, target, intent(out) :: RowMap(:), Col(:)
        integer(c_int) :: n
        integer(c_long_long) :: nnz
        rc = ceos_csr_pattern(h, 1_c_int, n, nnz, c_null_ptr, c_null_ptr)
        if (rc /= CEOS_OK) return
        allocate(RowMap(n + 1), Col(nnz))
        rc = ceos_csr_pattern(h, 1_c_int, n, nnz, c_loc(RowMap), c_loc(Col))
    end function

end module ceos_gpu_interface
