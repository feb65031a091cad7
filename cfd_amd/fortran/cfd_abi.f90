! Fortran 2003 interface to libcfd_amd.so: the C ABI of include/cfd_abi.h
! (host-pointer entries, the device-resident session, batches) as
! iso_c_binding interfaces, plus cfd_check, which
! turns a non-zero return code into ERROR STOP with the library's message.
! Arrays follow the reference's conventions: 1-based ids, 0-based offsets,
! column-major U(4,npoin), inpoel(3,nelem), dNx(3,nelem).
module cfd_abi
    use iso_c_binding
    implicit none

    ! InputData constants the path reads (dataLoader.f90:5-8, derived :59-64)
    type, bind(c) :: cfd_consts
        real(c_double) :: gama, fr, fmu, fk, fcv
        real(c_double) :: t_inf, rho_inf, u_inf, v_inf
        real(c_double) :: cte, fsafe
    end type cfd_consts

    ! boundary lists as left by dataLoader (dataLoader.f90:130-268)
    type, bind(c) :: cfd_bc_lists
        integer(c_int) :: nfixv
        type(c_ptr) :: ifixv_node, rfixv_valuex, rfixv_valuey
        integer(c_int) :: nfixrho
        type(c_ptr) :: ifixrho_node, rfixrho_value
        integer(c_int) :: nfixt
        type(c_ptr) :: ifixt_node, rfixt_value
        integer(c_int) :: nwall
        type(c_ptr) :: wall
    end type cfd_bc_lists

    ! device-resident session: the case (cfd_case_desc of include/cfd_abi.h)
    type, bind(c) :: cfd_case_desc
        type(cfd_consts) :: c
        integer(c_int) :: nrk, moving, itlocal
        real(c_double) :: xref, yref
        integer(c_int) :: nsetedges
        type(c_ptr) :: iset
        integer(c_int) :: nmove, nfix_move
        type(c_ptr) :: ilaux
        type(cfd_bc_lists) :: bc
        type(c_ptr) :: x, y
        integer(c_int) :: motion
        real(c_double) :: motion_amp, motion_omega
        integer(c_int) :: pcg_rel
        real(c_double) :: pcg_tol
        integer(c_int) :: pcg_maxit, gcl, precision, warm_start, patch_elems, patch_block
        integer(c_int) :: renumber
    end type cfd_case_desc

    ! one case of a batch (cfd_batch_case)
    type, bind(c) :: cfd_batch_case
        integer(c_int) :: nelem, npoin
        type(c_ptr) :: inpoel, x, y
        type(cfd_consts) :: c
        type(cfd_bc_lists) :: bc
    end type cfd_batch_case

    ! field ids (CFD_F_*) and scalar slots (CFD_S_*, 0-based) of the header
    integer(c_int), parameter :: cfd_f_u = 0, cfd_f_u1 = 1, cfd_f_rhs = 2, cfd_f_t = 3, &
        cfd_f_vel_x = 4, cfd_f_vel_y = 5, cfd_f_w_x = 6, cfd_f_w_y = 7, cfd_f_x = 8, &
        cfd_f_y = 9, cfd_f_x1 = 10, cfd_f_y1 = 11, cfd_f_rho = 12, cfd_f_e = 13, cfd_f_p = 14, &
        cfd_f_rmach = 15, cfd_f_m = 16, cfd_f_gamm = 17, cfd_f_xpos = 18, cfd_f_ypos = 19, &
        cfd_f_dtl = 20, cfd_f_dt = 21, cfd_f_shoc = 22, cfd_f_t_sugn1 = 23, &
        cfd_f_t_sugn2 = 24, cfd_f_t_sugn3 = 25, cfd_f_area = 26, cfd_f_dnx = 27, &
        cfd_f_dny = 28, cfd_f_hh = 29, cfd_f_hhx = 30, cfd_f_hhy = 31, cfd_f_lap_sparse = 32, &
        cfd_f_lap_diag = 33
    integer(c_int), parameter :: cfd_s_dtmin = 0, cfd_s_time = 1, cfd_s_dtmin1 = 2, &
        cfd_s_bandera = 3, cfd_s_iter = 4, cfd_s_hmin = 5, cfd_s_fx = 6, cfd_s_fy = 7, &
        cfd_s_rm = 8, cfd_s_disn2 = 9, cfd_s_pcg_itx = 10, cfd_s_pcg_ity = 11, &
        cfd_s_pcg_rzx = 12, cfd_s_pcg_rzy = 13, cfd_s_ws_learned = 14, cfd_s_ws_solves = 15, &
        cfd_s_ws_loss = 16, cfd_s_pcg_itsum = 17, cfd_s_fvx = 18, cfd_s_fvy = 19, &
        cfd_s_pcg_unconv = 20, cfd_nscalars = 21

    ! RK's module arrays (subrutinas.f90:645-849), nodal and per element
    type, bind(c) :: cfd_rk_nodes
        type(c_ptr) :: u1, rhs, t, vel_x, vel_y, rho, e, p, rmach
        type(c_ptr) :: u, w_x, w_y, m, gamm
    end type cfd_rk_nodes
    type, bind(c) :: cfd_rk_elems
        type(c_ptr) :: dnx, dny, area
        type(c_ptr) :: shoc, t_sugn1, t_sugn2, t_sugn3
    end type cfd_rk_elems

    interface
        integer(c_int) function cfd_abi_version() bind(c, name='cfd_abi_version')
            import :: c_int
        end function
        type(c_ptr) function cfd_last_error() bind(c, name='cfd_last_error')
            import :: c_ptr
        end function
        integer(c_int) function cfd_device_count(count) bind(c, name='cfd_device_count')
            import :: c_int
            integer(c_int), intent(out) :: count
        end function

        ! mesh context (PointNeighbor / Mlaplace tables on the device)
        integer(c_int) function cfd_mesh_create(m, nelem, npoin, inpoel) &
                bind(c, name='cfd_mesh_create')
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: m
            integer(c_int), value :: nelem, npoin
            integer(c_int), intent(in) :: inpoel(*)
        end function
        integer(c_int) function cfd_mesh_destroy(m) bind(c, name='cfd_mesh_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: m
        end function
        integer(c_int) function cfd_mesh_sizes(m, npsup, nnz) bind(c, name='cfd_mesh_sizes')
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), intent(out) :: npsup, nnz
        end function
        ! PointNeighbor::getEsup (pointNeighbor.f90:5-43)
        integer(c_int) function cfd_get_esup(m, esup1, esup2) bind(c, name='cfd_get_esup')
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), intent(out) :: esup1(*), esup2(*)
        end function
        ! PointNeighbor::getPsup (pointNeighbor.f90:45-91)
        integer(c_int) function cfd_get_psup(m, psup1, psup2) bind(c, name='cfd_get_psup')
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), intent(out) :: psup1(*), psup2(*)
        end function
        ! Mlaplace initialize (mLaplace.f90:60-94)
        integer(c_int) function cfd_lap_structure(m, lap_rowptr, lap_idx) &
                bind(c, name='cfd_lap_structure')
            import :: c_int, c_ptr
            type(c_ptr), value :: m
            integer(c_int), intent(out) :: lap_rowptr(*), lap_idx(*)
        end function
        ! Mlaplace::laplace (mLaplace.f90:7-58)
        integer(c_int) function cfd_laplace(m, x, y, dnx, dny, lap_sparse, lap_diag) &
                bind(c, name='cfd_laplace')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: x(*), y(*), dnx(*), dny(*)
            real(c_double), intent(out) :: lap_sparse(*), lap_diag(*)
        end function
        ! deriv (subrutinas.f90:88-126)
        integer(c_int) function cfd_deriv(m, x, y, area, dnx, dny, hh, hhx, hhy, hmin) &
                bind(c, name='cfd_deriv')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: x(*), y(*)
            real(c_double), intent(out) :: area(*), dnx(*), dny(*), hh(*), hhx(*), hhy(*)
            real(c_double), intent(out) :: hmin
        end function
        ! MASAS (subrutinas.f90:128-153)
        integer(c_int) function cfd_masas(m, area, mass) bind(c, name='cfd_masas')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: area(*)
            real(c_double), intent(out) :: mass(*)
        end function
        ! deltat (subrutinas.f90:155-218)
        integer(c_int) function cfd_deltat(m, c, area, t, vel_x, vel_y, w_x, w_y, dt, dtmin) &
                bind(c, name='cfd_deltat')
            import :: c_int, c_ptr, c_double, cfd_consts
            type(c_ptr), value :: m
            type(cfd_consts), intent(in) :: c
            real(c_double), intent(in) :: area(*), t(*), vel_x(*), vel_y(*), w_x(*), w_y(*)
            real(c_double), intent(out) :: dt(*), dtmin
        end function
        ! ESTAB (subrutinas.f90:332-446)
        integer(c_int) function cfd_estab(m, c, u, t, gamm, vel_x, vel_y, w_x, w_y, dnx, dny, &
                dtmin, shoc, t_sugn1, t_sugn2, t_sugn3) bind(c, name='cfd_estab')
            import :: c_int, c_ptr, c_double, cfd_consts
            type(c_ptr), value :: m
            type(cfd_consts), intent(in) :: c
            real(c_double), intent(in) :: u(*), t(*), gamm(*), vel_x(*), vel_y(*), w_x(*), w_y(*)
            real(c_double), intent(in) :: dnx(*), dny(*)
            real(c_double), value :: dtmin
            real(c_double), intent(out) :: shoc(*), t_sugn1(*), t_sugn2(*), t_sugn3(*)
        end function
        ! calcRHS_mod::calcRHS (calcRHS.f90:4-154): accumulates into rhs
        integer(c_int) function cfd_calcrhs(m, c, rhs, u, theta, t, dnx, dny, area, shoc, dtl, &
                t_sugn1, t_sugn2, t_sugn3) bind(c, name='cfd_calcrhs')
            import :: c_int, c_ptr, c_double, cfd_consts
            type(c_ptr), value :: m
            type(cfd_consts), intent(in) :: c
            real(c_double), intent(inout) :: rhs(*)
            real(c_double), intent(in) :: u(*), theta(*), t(*), dnx(*), dny(*), area(*)
            real(c_double), intent(in) :: shoc(*), dtl(*), t_sugn1(*), t_sugn2(*), t_sugn3(*)
        end function
        ! FUENTE (subrutinas.f90:1036-1094): accumulates into rhs
        integer(c_int) function cfd_fuente(m, rhs, u, w_x, w_y, dnx, dny, area, dtl) &
                bind(c, name='cfd_fuente')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(inout) :: rhs(*)
            real(c_double), intent(in) :: u(*), w_x(*), w_y(*), dnx(*), dny(*), area(*), dtl(*)
        end function
        ! Mnormales::normales (subrutinas.f90:7-64): normals kept for cfd_rk
        integer(c_int) function cfd_normales(m, bc, x, y) bind(c, name='cfd_normales')
            import :: c_int, c_ptr, c_double, cfd_bc_lists
            type(c_ptr), value :: m
            type(cfd_bc_lists), intent(in) :: bc
            real(c_double), intent(in) :: x(*), y(*)
        end function
        ! Mnormales::normalvel (subrutinas.f90:66-85) with the normals of cfd_normales
        integer(c_int) function cfd_normalvel(m, vel_x, vel_y, w_x, w_y) &
                bind(c, name='cfd_normalvel')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(inout) :: vel_x(*), vel_y(*)
            real(c_double), intent(in) :: w_x(*), w_y(*)
        end function
        ! FORCE_VISC (ns2DComp.ALE.f90:819-893) for one set of edges iset(3,nedge)
        integer(c_int) function cfd_force_visc(m, nedge, iset, x, y, p, t, vel_x, vel_y, dnx, &
                                               dny, t_inf, f_vx, f_vy) &
                bind(c, name='cfd_force_visc')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            integer(c_int), value :: nedge
            integer(c_int), intent(in) :: iset(*)
            real(c_double), intent(in) :: x(*), y(*), p(*), t(*), vel_x(*), vel_y(*), dnx(*), &
                                          dny(*)
            real(c_double), value :: t_inf
            real(c_double), intent(out) :: f_vx, f_vy
        end function
        ! smoothing_mod::smoothing (smoothing.f90:21-75); dx, dy: c_null_ptr or
        ! the addresses of the optional arguments
        integer(c_int) function cfd_smoothing(m, x, y, fixed, dx, dy, order, sweeps, hmin) &
                bind(c, name='cfd_smoothing')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m, dx, dy
            real(c_double), intent(inout) :: x(*), y(*)
            integer(c_int), intent(in) :: fixed(*)
            integer(c_int), value :: order
            integer(c_int), intent(out) :: sweeps
            real(c_double), intent(out) :: hmin
        end function
        ! smoothing_mod::setTol_move (smoothing.f90:386-401)
        integer(c_int) function cfd_smoothing_set_tol_move(m, dx, dy) &
                bind(c, name='cfd_smoothing_set_tol_move')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: dx(*), dy(*)
        end function
        ! RK (subrutinas.f90:645-849, NGAS = 0)
        integer(c_int) function cfd_rk(m, c, bc, nrk, dtmin, dtl, nodes, elems) &
                bind(c, name='cfd_rk')
            import :: c_int, c_ptr, c_double, cfd_consts, cfd_bc_lists, cfd_rk_nodes, cfd_rk_elems
            type(c_ptr), value :: m
            type(cfd_consts), intent(in) :: c
            type(cfd_bc_lists), intent(in) :: bc
            integer(c_int), value :: nrk
            real(c_double), value :: dtmin
            real(c_double), intent(in) :: dtl(*)
            type(cfd_rk_nodes), intent(in) :: nodes
            type(cfd_rk_elems), intent(in) :: elems
        end function
        ! BiconjGrad::biCG (biconjGrad.f90:8-62)
        integer(c_int) function cfd_bicg(npoin, spmtx, spidx, sprowptr, diagmtx, x, b, x_fix, &
                x_fixidx, nfix, iters, rz) bind(c, name='cfd_bicg')
            import :: c_int, c_double
            integer(c_int), value :: npoin, nfix
            real(c_double), intent(in) :: spmtx(*), diagmtx(*), b(*), x_fix(*)
            integer(c_int), intent(in) :: spidx(*), sprowptr(*), x_fixidx(*)
            real(c_double), intent(inout) :: x(*)
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: rz
        end function
        ! gcl_mod (gcl.f90:8-62)
        integer(c_int) function cfd_gcl_put_w(m, w_x, w_y) bind(c, name='cfd_gcl_put_w')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: w_x(*), w_y(*)
        end function
        integer(c_int) function cfd_gcl_put_area(m, area) bind(c, name='cfd_gcl_put_area')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(in) :: area(*)
        end function
        integer(c_int) function cfd_gcl_main(m, mass, w_x, w_y, dnx, dny, area, dt, mode) &
                bind(c, name='cfd_gcl_main')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: m
            real(c_double), intent(inout) :: mass(*)
            real(c_double), intent(in) :: w_x(*), w_y(*), dnx(*), dny(*), area(*)
            real(c_double), value :: dt
            integer(c_int), value :: mode
        end function

        ! ---- device-resident session (ns2DComp.ALE.f90:138-281) ----
        integer(c_int) function cfd_session_create(s, m, d) bind(c, name='cfd_session_create')
            import :: c_int, c_ptr, cfd_case_desc
            type(c_ptr), intent(out) :: s
            type(c_ptr), value :: m
            type(cfd_case_desc), intent(in) :: d
        end function
        integer(c_int) function cfd_session_destroy(s) bind(c, name='cfd_session_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: s
        end function
        integer(c_int) function cfd_session_init_freestream(s) &
            bind(c, name='cfd_session_init_freestream')
            import :: c_int, c_ptr
            type(c_ptr), value :: s
        end function
        integer(c_int) function cfd_session_step(s, nsteps) bind(c, name='cfd_session_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: s
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function cfd_session_sync(s) bind(c, name='cfd_session_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: s
        end function
        integer(c_int) function cfd_session_field_size(s, field, count) &
            bind(c, name='cfd_session_field_size')
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: s
            integer(c_int), value :: field
            integer(c_long_long), intent(out) :: count
        end function
        integer(c_int) function cfd_session_get_field(s, field, host) &
            bind(c, name='cfd_session_get_field')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: s
            integer(c_int), value :: field
            real(c_double), intent(out) :: host(*)
        end function
        integer(c_int) function cfd_session_set_field(s, field, host) &
            bind(c, name='cfd_session_set_field')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: s
            integer(c_int), value :: field
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function cfd_session_get_scalars(s, scalars) &
            bind(c, name='cfd_session_get_scalars')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double), intent(out) :: scalars(*)
        end function
        integer(c_int) function cfd_session_set_scalars(s, scalars) &
            bind(c, name='cfd_session_set_scalars')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double), intent(in) :: scalars(*)
        end function
        ! restart file of a session (ns2DComp.ALE.f90:394-435, :898-917); path ends
        ! with c_null_char
        integer(c_int) function cfd_session_write_restart(s, path) &
            bind(c, name='cfd_session_write_restart')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: s
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function cfd_session_read_restart(s, path, keep_clock) &
            bind(c, name='cfd_session_read_restart')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: s
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: keep_clock
        end function

        ! ---- restart files, host only (PRINTREST ns2DComp.ALE.f90:898-917,
        ! RESTART :394-435); path ends with c_null_char ----
        integer(c_int) function cfd_write_restart(path, iter, time, npoin, u, t, gamm) &
            bind(c, name='cfd_write_restart')
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: iter, npoin
            real(c_double), value :: time
            real(c_double), intent(in) :: u(*), t(*), gamm(*)
        end function
        integer(c_int) function cfd_read_restart(path, npoin, iter, time, u, t, gamm) &
            bind(c, name='cfd_read_restart')
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: npoin
            integer(c_int), intent(out) :: iter
            real(c_double), intent(out) :: time
            real(c_double), intent(out) :: u(*), t(*), gamm(*)
        end function

        ! ---- batches of independent fixed-mesh cases ----
        integer(c_int) function cfd_batch_create(b, ncases, cases, precision, itlocal, &
                                                 patch_elems, patch_block) &
            bind(c, name='cfd_batch_create')
            import :: c_int, c_ptr, cfd_batch_case
            type(c_ptr), intent(out) :: b
            integer(c_int), value :: ncases, precision, itlocal, patch_elems, patch_block
            type(cfd_batch_case), intent(in) :: cases(*)
        end function
        integer(c_int) function cfd_batch_destroy(b) bind(c, name='cfd_batch_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function
        integer(c_int) function cfd_batch_step(b, nsteps) bind(c, name='cfd_batch_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function cfd_batch_get_field(b, icase, field, host) &
            bind(c, name='cfd_batch_get_field')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int), value :: icase, field
            real(c_double), intent(out) :: host(*)
        end function
        integer(c_int) function cfd_batch_get_scalars(b, icase, scalars) &
            bind(c, name='cfd_batch_get_scalars')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int), value :: icase
            real(c_double), intent(out) :: scalars(*)
        end function
    end interface

contains

    ! ERROR STOP with the library's message when code /= 0 (the reference has
    ! no status codes: it STOPs on fatal conditions).
    subroutine cfd_check(code, what)
        integer(c_int), intent(in) :: code
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (code == 0) return
        text = ''
        call c_f_pointer(cfd_last_error(), msg, [512])
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        write (*, '(a,a,a,i0,a,a)') 'cfd: ', what, ' failed (code ', code, '): ', trim(text)
        error stop 1
    end subroutine cfd_check

end module cfd_abi
