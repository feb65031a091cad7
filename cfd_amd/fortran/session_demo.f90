! Runs the device-resident session from Fortran through the cfd_abi
! interface module: one fixed-mesh case read from a stream file, the main
! loop's time steps on the GPU (cfd_session_step), U, T and the scalars
! written to another stream file.
!
!   session_demo <in.bin> <out.bin>
! in.bin : npoin, nelem, nsteps, precision (int32); cfd_consts (11 float64);
!          inpoel(3,nelem) (int32); x, y (npoin float64);
!          nfixv, ifixv (int32), vx, vy (float64); nfixrho, ifixrho, rho;
!          nfixt, ifixt, tfix; nsetedges, iset(3,nsetedges) (int32);
!          nmove, nfix_move, ilaux (int32); xref, yref (float64)
! out.bin: U(4,npoin), T(npoin), scalars(cfd_nscalars) (float64)
program session_demo
    use cfd_abi
    implicit none
    character(len=1024) :: fin, fout
    integer :: u
    integer(c_int) :: npoin, nelem, nsteps, precision, nfixv, nfixrho, nfixt, nset, nmove, nfm
    integer(c_int), allocatable, target :: inpoel(:, :), ifixv(:), ifixrho(:), ifixt(:)
    integer(c_int), allocatable, target :: iset(:, :), ilaux(:)
    real(c_double), allocatable, target :: x(:), y(:), vx(:), vy(:), rho(:), tfix(:)
    real(c_double), allocatable :: uu(:, :), t(:)
    real(c_double) :: sc(cfd_nscalars), xref, yref
    type(cfd_case_desc) :: d
    type(c_ptr) :: m, s
    integer(c_int), target :: dummy_i(3)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read (u) npoin, nelem, nsteps, precision
    read (u) d%c
    allocate (inpoel(3, nelem), x(npoin), y(npoin))
    read (u) inpoel, x, y
    read (u) nfixv
    allocate (ifixv(nfixv), vx(nfixv), vy(nfixv))
    read (u) ifixv, vx, vy
    read (u) nfixrho
    allocate (ifixrho(nfixrho), rho(nfixrho))
    read (u) ifixrho, rho
    read (u) nfixt
    allocate (ifixt(nfixt), tfix(nfixt))
    read (u) ifixt, tfix
    read (u) nset
    allocate (iset(3, max(nset, 1)))
    read (u) iset(:, 1:nset)
    read (u) nmove, nfm
    allocate (ilaux(max(nmove + nfm, 1)))
    read (u) ilaux(1:nmove + nfm)
    read (u) xref, yref
    close (u)

    ! the case as the main program sets it up (fixed mesh, reference options)
    d%nrk = 4
    d%moving = 0
    d%itlocal = 0
    d%xref = xref
    d%yref = yref
    d%nsetedges = nset
    d%iset = c_loc(iset)
    d%nmove = nmove
    d%nfix_move = nfm
    d%ilaux = c_loc(ilaux)
    d%bc%nfixv = nfixv
    d%bc%ifixv_node = c_loc(ifixv)
    d%bc%rfixv_valuex = c_loc(vx)
    d%bc%rfixv_valuey = c_loc(vy)
    d%bc%nfixrho = nfixrho
    d%bc%ifixrho_node = c_loc(ifixrho)
    d%bc%rfixrho_value = c_loc(rho)
    d%bc%nfixt = nfixt
    d%bc%ifixt_node = c_loc(ifixt)
    d%bc%rfixt_value = c_loc(tfix)
    d%bc%nwall = 0
    d%bc%wall = c_loc(dummy_i)
    d%x = c_loc(x)
    d%y = c_loc(y)
    d%motion = 0
    d%motion_amp = 0.d0
    d%motion_omega = 0.d0
    d%pcg_rel = 0
    d%pcg_tol = 0.d0
    d%pcg_maxit = 0
    d%gcl = -1
    d%precision = precision
    d%warm_start = 0
    d%patch_elems = 0
    d%patch_block = 0
    d%renumber = 0

    call cfd_check(cfd_mesh_create(m, nelem, npoin, inpoel), 'cfd_mesh_create')
    call cfd_check(cfd_session_create(s, m, d), 'cfd_session_create')
    call cfd_check(cfd_session_init_freestream(s), 'cfd_session_init_freestream')
    call cfd_check(cfd_session_step(s, nsteps), 'cfd_session_step')
    call cfd_check(cfd_session_sync(s), 'cfd_session_sync')
    allocate (uu(4, npoin), t(npoin))
    call cfd_check(cfd_session_get_field(s, cfd_f_u, uu), 'get U')
    call cfd_check(cfd_session_get_field(s, cfd_f_t, t), 'get T')
    call cfd_check(cfd_session_get_scalars(s, sc), 'get scalars')
    call cfd_check(cfd_session_destroy(s), 'cfd_session_destroy')
    call cfd_check(cfd_mesh_destroy(m), 'cfd_mesh_destroy')

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (u) uu, t, sc
    close (u)
    write (*, '(a,i0,a,es12.5,a,es12.5)') 'session_demo: ', nsteps, ' steps, TIME ', &
        sc(cfd_s_time + 1), ', DTMIN ', sc(cfd_s_dtmin + 1)
end program session_demo
