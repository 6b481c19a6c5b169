!==========================================================================================
! example_driver: one F+K evaluation of a hex8 mesh through the Fortran binding.
!
! Reads a plain-text case (nn ne / X(3,nn) / conn(8,ne) / model nprops props / time / U(3nn)),
! runs UpdateMeshCoordinates + SolveConstitutiveModel + InternalForce + TangentStiffnessMatrix
! through libceos_gpu.so and writes nnz, RowMap, Col, Fint and Kg%Val to the output file.
!
!   example_driver <case.txt> <out.txt>
!==========================================================================================
program example_driver
    use, intrinsic :: iso_c_binding
    use ceos_gpu_interface
    implicit none
    type(c_ptr) :: h
    type(ceos_status) :: st
    integer(c_int) :: nn, ne, model, nprops, rc, i
    real(c_double) :: time
    real(c_double), allocatable :: X(:), U(:), props(:), Fint(:), Kval(:)
    integer(c_int), allocatable :: conn(:), mat_id(:), RowMap(:), Col(:)
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), status='old', action='read')
    read(10, *) nn, ne
    allocate(X(3 * nn), conn(8 * ne), mat_id(ne), U(3 * nn), Fint(3 * nn))
    read(10, *) X
    read(10, *) conn
    read(10, *) model, nprops
    allocate(props(nprops))
    read(10, *) props
    read(10, *) time
    read(10, *) U
    close(10)
    mat_id = 1

    rc = ceos_mesh_create(nn, X, ne, 8_c_int, conn, mat_id, h)
    if (rc /= CEOS_OK) stop 'ceos_mesh_create'
    rc = ceos_material_set(h, 1_c_int, model, props, nprops)
    if (rc /= CEOS_OK) stop 'ceos_material_set'
    rc = ceos_csr_upper(h, RowMap, Col)
    if (rc /= CEOS_OK) stop 'ceos_csr_upper'
    allocate(Kval(size(Col)))
    rc = ceos_update_mesh(h, U)
    if (rc /= CEOS_OK) stop 'ceos_update_mesh'
    rc = ceos_solve_constitutive(h, time, U, st)
    if (rc /= CEOS_OK) stop 'ceos_solve_constitutive'
    rc = ceos_internal_force(h, Fint, st)
    if (rc /= CEOS_OK) stop 'ceos_internal_force'
    rc = ceos_tangent(h, Kval)
    if (rc /= CEOS_OK) stop 'ceos_tangent'
    call ceos_mesh_destroy(h)

    open(11, file=trim(fout), status='replace', action='write')
    write(11, '(I12)') size(Col)
    write(11, '(I12)') (RowMap(i), i = 1, size(RowMap))
    write(11, '(I12)') (Col(i), i = 1, size(Col))
    write(11, '(ES25.17)') (Fint(i), i = 1, size(Fint))
    write(11, '(ES25.17)') (Kval(i), i = 1, size(Kval))
    close(11)
end program example_driver
