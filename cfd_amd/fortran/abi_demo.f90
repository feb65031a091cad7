! Calls libcfd_amd.so from Fortran through the cfd_abi interface module: the
! neighbour tables, deriv, MASAS, laplace and biCG on a mesh read from a
! stream file, results written to another stream file.
!
!   abi_demo <in.bin> <out.bin>
! in.bin : npoin, nelem (int32); inpoel(3,nelem) (int32); x, y (npoin float64);
!          nfix (int32); fixed ids (nfix int32); fixed values (nfix float64)
! out.bin: npsup, nnz; esup1, esup2, psup1, psup2, lap_rowptr, lap_idx (int32);
!          area, dNx, dNy, HH, HHX, HHY, hmin, M, lap_sparse, lap_diag, x_solved
!          (float64); biCG iterations (int32)
program abi_demo
    use cfd_abi
    implicit none
    character(len=1024) :: fin, fout
    integer :: u
    integer(c_int) :: npoin, nelem, nfix, npsup, nnz, iters
    integer(c_int), allocatable :: inpoel(:, :), fidx(:), esup1(:), esup2(:), psup1(:), psup2(:)
    integer(c_int), allocatable :: rowptr(:), idx(:)
    real(c_double), allocatable :: x(:), y(:), xfix(:), area(:), dnx(:, :), dny(:, :), hh(:)
    real(c_double), allocatable :: hhx(:), hhy(:), mass(:), lap(:), diag(:), sol(:), b(:)
    real(c_double) :: hmin, rz
    type(c_ptr) :: m

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read (u) npoin, nelem
    allocate (inpoel(3, nelem), x(npoin), y(npoin))
    read (u) inpoel, x, y
    read (u) nfix
    allocate (fidx(nfix), xfix(nfix))
    read (u) fidx, xfix
    close (u)
    write (*, '(a,i0,a,i0,a,i0)') 'abi_demo: library ABI ', cfd_abi_version(), ', npoin ', npoin, &
        ', nelem ', nelem

    call cfd_check(cfd_mesh_create(m, nelem, npoin, inpoel), 'cfd_mesh_create')
    call cfd_check(cfd_mesh_sizes(m, npsup, nnz), 'cfd_mesh_sizes')
    allocate (esup1(3*nelem), esup2(npoin + 1), psup1(npsup), psup2(npoin + 1))
    allocate (rowptr(npoin + 1), idx(nnz))
    call cfd_check(cfd_get_esup(m, esup1, esup2), 'getEsup')
    call cfd_check(cfd_get_psup(m, psup1, psup2), 'getPsup')
    call cfd_check(cfd_lap_structure(m, rowptr, idx), 'laplace initialize')

    allocate (area(nelem), dnx(3, nelem), dny(3, nelem), hh(nelem), hhx(nelem), hhy(nelem))
    allocate (mass(npoin), lap(nnz), diag(npoin), sol(npoin), b(npoin))
    call cfd_check(cfd_deriv(m, x, y, area, dnx, dny, hh, hhx, hhy, hmin), 'deriv')
    call cfd_check(cfd_masas(m, area, mass), 'MASAS')
    call cfd_check(cfd_laplace(m, x, y, dnx, dny, lap, diag), 'laplace')
    sol = 0.d0
    b = 0.d0
    call cfd_check(cfd_bicg(npoin, lap, idx, rowptr, diag, sol, b, xfix, fidx, nfix, iters, rz), &
                   'biCG')
    call cfd_check(cfd_mesh_destroy(m), 'cfd_mesh_destroy')

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (u) npsup, nnz
    write (u) esup1, esup2, psup1, psup2, rowptr, idx
    write (u) area, dnx, dny, hh, hhx, hhy, hmin, mass, lap, diag, sol
    write (u) iters
    close (u)
    write (*, '(a,i0,a,es10.3)') 'abi_demo: done, biCG iterations ', iters, ', |r.z| ', rz
end program abi_demo
