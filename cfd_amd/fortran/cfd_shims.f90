! Drop-in replacements for the reference's hot-path routines, with the
! reference's module names, procedure names and argument lists, forwarding to
! libcfd_amd.so through the cfd_abi interface module.  They replace, in the
! reference build:
!   pointNeighbor.f90 (module PointNeighbor), biconjGrad.f90 (module
!   BiconjGrad), mLaplace.f90 (module Mlaplace), calcRHS.f90 (module
!   calcRHS_mod), gcl.f90 (module gcl_mod), and from subrutinas.f90 the
!   external subroutines deriv, MASAS, deltat, ESTAB, FUENTE and RK plus
!   module Mnormales (normales, normalvel), and smoothing.f90 (module
!   smoothing_mod), which flang cannot compile as the reference wrote it.
! The hidden inputs the reference reads by USE (InputData, MeshData,
! Mvariables, MVELOCIDADES, MVARIABGEN, MESTABILIZACION) are read the same
! way here, so this file compiles only together with the reference's own
! module sources (commonModules.f90, dataLoader.f90).
! A non-zero library status becomes ERROR STOP with the library's message.

! One mesh context per run (the reference's inpoel is fixed after loading).
module cfd_shim_ctx
    use cfd_abi
    implicit none
    type(c_ptr), save :: mesh = c_null_ptr
    integer, save :: mesh_nelem = -1, mesh_npoin = -1
contains
    function shim_mesh(inpoel, nelem, npoin) result(m)
        integer, intent(in) :: nelem, npoin
        integer, intent(in) :: inpoel(3, nelem)
        type(c_ptr) :: m
        if (.not. c_associated(mesh) .or. nelem /= mesh_nelem .or. npoin /= mesh_npoin) then
            if (c_associated(mesh)) call cfd_check(cfd_mesh_destroy(mesh), 'cfd_mesh_destroy')
            call cfd_check(cfd_mesh_create(mesh, nelem, npoin, inpoel), 'cfd_mesh_create')
            mesh_nelem = nelem
            mesh_npoin = npoin
        end if
        m = mesh
    end function shim_mesh

    function shim_mesh_data() result(m)  ! the context of MeshData's mesh
        use MeshData, only: inpoel, nelem, npoin
        type(c_ptr) :: m
        m = shim_mesh(inpoel, nelem, npoin)
    end function shim_mesh_data

    function shim_consts() result(c)  ! InputData after loading (dataLoader.f90:5-8, :59-64)
        use InputData, only: GAMA, FR, FMU, FK, FCV, T_inf, RHO_inf, U_inf, V_inf, CTE, FSAFE
        type(cfd_consts) :: c
        c = cfd_consts(GAMA, FR, FMU, FK, FCV, T_inf, RHO_inf, U_inf, V_inf, CTE, FSAFE)
    end function shim_consts
end module cfd_shim_ctx

! pointNeighbor.f90:1-92
module PointNeighbor
    use cfd_shim_ctx
    implicit none
    integer, dimension(:), allocatable :: esup1, esup2
    integer, dimension(:), allocatable :: psup1, psup2, lpoin
contains
    subroutine getEsup(inpoel, nelem, npoin)
        integer nelem, npoin
        integer inpoel(3, nelem)
        allocate (esup1(3*nelem), esup2(npoin + 1))
        call cfd_check(cfd_get_esup(shim_mesh(inpoel, nelem, npoin), esup1, esup2), 'getEsup')
    end subroutine getEsup

    subroutine getPsup(inpoel, nelem, npoin)
        integer nelem, npoin
        integer inpoel(3, nelem)
        integer(c_int) :: npsup, nnz
        type(c_ptr) :: m
        if (.not. allocated(esup1)) call getEsup(inpoel, nelem, npoin)  ! :51
        m = shim_mesh(inpoel, nelem, npoin)
        call cfd_check(cfd_mesh_sizes(m, npsup, nnz), 'getPsup')
        allocate (psup1(npsup), psup2(npoin + 1))
        call cfd_check(cfd_get_psup(m, psup1, psup2), 'getPsup')
    end subroutine getPsup
end module PointNeighbor

! biconjGrad.f90:1-191; bicg_iters / bicg_rz report the last solve
module BiconjGrad
    use cfd_abi
    implicit none
    private
    public biCG
    integer, public :: bicg_iters = 0
    real(8), public :: bicg_rz = 0.d0
contains
    subroutine biCG(spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx, npoin, nfix)
        integer npoin, nfix
        integer spRowptr(npoin + 1), spIdx(spRowptr(npoin + 1))
        integer x_fixIdx(nfix)
        real(8) spMtx(spRowptr(npoin + 1)), x(npoin), b(npoin), diagMtx(npoin)
        real(8) x_fix(nfix)
        integer(c_int) :: it
        real(c_double) :: rz
        call cfd_check(cfd_bicg(npoin, spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx, &
                                nfix, it, rz), 'biCG')
        bicg_iters = it
        bicg_rz = rz
    end subroutine biCG
end module BiconjGrad

! mLaplace.f90:1-109 (initialize runs on the first call, as there)
module Mlaplace
    use cfd_shim_ctx
    implicit none
    integer, dimension(:), allocatable :: lap_idx, lap_rowptr
    real(8), dimension(:), allocatable :: lap_sparse, lap_diag
contains
    subroutine laplace(inpoel, area, dNx, dNy, nelem, npoin)
        use MeshData, only: X, Y
        integer, intent(in) :: nelem, npoin
        integer, intent(in) :: inpoel(3, nelem)
        real*8, intent(in) :: dNx(3, nelem), dNy(3, nelem), area(nelem)
        integer(c_int) :: npsup, nnz
        type(c_ptr) :: m
        m = shim_mesh(inpoel, nelem, npoin)
        if (.not. allocated(lap_rowptr)) then
            call cfd_check(cfd_mesh_sizes(m, npsup, nnz), 'laplace')
            allocate (lap_sparse(nnz), lap_idx(nnz), lap_rowptr(npoin + 1), lap_diag(npoin))
            call cfd_check(cfd_lap_structure(m, lap_rowptr, lap_idx), 'laplace initialize')
        end if
        call cfd_check(cfd_laplace(m, X, Y, dNx, dNy, lap_sparse, lap_diag), 'laplace')
    end subroutine laplace
end module Mlaplace

! calcRHS.f90:1-155
module calcRHS_mod
    use cfd_shim_ctx
    implicit none
contains
    subroutine calcRHS(rhs, U, theta, dNx, dNy, area, shoc, dtl, t_sugn1, t_sugn2, t_sugn3, &
                       inpoel, nelem, npoin)
        use Mvariables, only: T
        integer, intent(in) :: npoin, nelem, inpoel(3, nelem)
        real*8, intent(inout) :: rhs(4, npoin)
        real*8, intent(in) :: U(4, npoin), theta(4, npoin), dNx(3, nelem), dNy(3, nelem)
        real*8, intent(in), dimension(nelem) :: area, dtl, shoc, t_sugn1, t_sugn2, t_sugn3
        call cfd_check(cfd_calcrhs(shim_mesh(inpoel, nelem, npoin), shim_consts(), rhs, U, theta, &
                                   T, dNx, dNy, area, shoc, dtl, t_sugn1, t_sugn2, t_sugn3), &
                       'calcRHS')
    end subroutine calcRHS
end module calcRHS_mod

! gcl.f90:1-63; gcl_mode 0 is the file as written, 1 builds the y-divergence from W_y
module gcl_mod
    use cfd_shim_ctx
    implicit none
    private
    public :: main, putW, putArea
    integer, public :: gcl_mode = 0
contains
    subroutine main(M, W_x, W_y, dNx, dNy, area, inpoel, dt)
        real*8, intent(inout), dimension(:) :: M
        real*8, intent(in), dimension(:, :) :: dNx, dNy
        real*8, intent(in), dimension(:) :: W_x, W_y, area
        integer, intent(in), dimension(:, :) :: inpoel
        real*8, intent(in) :: dt
        call cfd_check(cfd_gcl_main(shim_mesh(inpoel, size(inpoel, 2), size(M)), M, W_x, W_y, &
                                    dNx, dNy, area, dt, gcl_mode), 'gcl main')
    end subroutine main

    subroutine putW(W_x, W_y)
        real*8, intent(in), dimension(:) :: W_x, W_y
        call cfd_check(cfd_gcl_put_w(shim_mesh_data(), W_x, W_y), 'gcl putW')
    end subroutine putW

    subroutine putArea(area)
        real*8, intent(in), dimension(:) :: area
        call cfd_check(cfd_gcl_put_area(shim_mesh_data(), area), 'gcl putArea')
    end subroutine putArea
end module gcl_mod

! subrutinas.f90:2-86: the wall normals live in the mesh context, where RK's
! NORMALVEL reads them
module Mnormales
    use cfd_shim_ctx
    implicit none
contains
    subroutine normales
        call shim_bc_call(.true.)
    end subroutine normales

    ! subrutinas.f90:66-85: the module velocities of the wall nodes the last
    ! normales found keep their tangential part relative to the mesh velocity
    ! (ADAMSB calls it; RK applies it inside the library)
    subroutine normalvel
        use MVELOCIDADES, only: VEL_X, VEL_Y, W_X, W_Y
        call cfd_check(cfd_normalvel(shim_mesh_data(), VEL_X, VEL_Y, W_X, W_Y), 'normalvel')
    end subroutine normalvel

    ! NORMALES, or RK's boundary lists only; the lists are copied into TARGET
    ! arrays so their addresses can be handed over
    subroutine shim_bc_call(with_normals)
        use MeshData, only: X, Y, nFixV, IFIXV_NODE, RFIXV_VALUEX, RFIXV_VALUEY, nFixRho, &
                            IFIXRHO_NODE, RFIXRHO_VALUE, nFixT, IFIXT_NODE, RFIXT_VALUE, &
                            nwall, wall
        logical, intent(in) :: with_normals
        type(cfd_bc_lists) :: bc
        integer(c_int), allocatable, target :: iv(:), ir(:), it(:), iw(:)
        real(c_double), allocatable, target :: vx(:), vy(:), vr(:), vt(:)
        iv = IFIXV_NODE(1:nFixV)
        vx = RFIXV_VALUEX(1:nFixV)
        vy = RFIXV_VALUEY(1:nFixV)
        ir = IFIXRHO_NODE(1:nFixRho)
        vr = RFIXRHO_VALUE(1:nFixRho)
        it = IFIXT_NODE(1:nFixT)
        vt = RFIXT_VALUE(1:nFixT)
        iw = reshape(wall(1:2, 1:nwall), [2*nwall])
        bc = cfd_bc_lists(nFixV, c_loc(iv), c_loc(vx), c_loc(vy), nFixRho, c_loc(ir), c_loc(vr), &
                          nFixT, c_loc(it), c_loc(vt), nwall, c_loc(iw))
        if (with_normals) call cfd_check(cfd_normales(shim_mesh_data(), bc, X, Y), 'normales')
    end subroutine shim_bc_call
end module Mnormales

! smoothing.f90:1-387: the start-up mesh smoothing.  smoothing_order picks the
! node order: 0 the reference's (one lane walks the nodes; meshes of at most
! 65536 nodes); 1 colour by colour, in parallel (the same per-node updates in
! another order); -1 (the default) 0 up to 65536 nodes, else 1.  The
! tol_move mode that setTol_move switches on stays on in the mesh context, as
! in the reference's module.
module smoothing_mod
    use cfd_shim_ctx
    implicit none
    private
    integer, public, save :: smoothing_order = -1
    public :: smoothing, setTol_move
contains
    subroutine smoothing(X, Y, inpoel, fixed, npoin0, nelem0, dX, dY)
        integer, intent(in) :: npoin0, nelem0, inpoel(3, nelem0)
        real*8, intent(inout) :: X(npoin0), Y(npoin0)
        logical, intent(in) :: fixed(npoin0)
        real*8, intent(in), optional :: dX(npoin0), dY(npoin0)
        integer(c_int), allocatable :: ifixed(:)
        real(c_double), allocatable, target :: tdx(:), tdy(:)
        type(c_ptr) :: pdx, pdy
        integer(c_int) :: sweeps, order
        real(c_double) :: hmin
        ifixed = merge(1, 0, fixed)
        pdx = c_null_ptr
        pdy = c_null_ptr
        if (present(dX) .and. present(dY)) then
            tdx = dX
            tdy = dY
            pdx = c_loc(tdx)
            pdy = c_loc(tdy)
        end if
        order = smoothing_order
        if (order < 0) order = merge(0, 1, npoin0 <= 65536)
        call cfd_check(cfd_smoothing(shim_mesh(inpoel, nelem0, npoin0), X, Y, ifixed, pdx, pdy, &
                                     order, sweeps, hmin), 'smoothing')
    end subroutine smoothing

    subroutine setTol_move(dX, dY)
        real*8, dimension(:), intent(in) :: dX, dY
        call cfd_check(cfd_smoothing_set_tol_move(shim_mesh_data(), dX, dY), 'setTol_move')
    end subroutine setTol_move
end module smoothing_mod

! RK's call: the boundary lists are copied into TARGET arrays and the module
! arrays are passed through TARGET dummies, so their addresses can be handed
! over for the duration of the call
module cfd_shim_rk
    use cfd_shim_ctx
    implicit none
    private
    public :: shim_rk
contains
    subroutine shim_rk(dtmin, nrk, gamm, dtl)
        use MeshData, only: M, dNx, dNy, area, nFixV, IFIXV_NODE, RFIXV_VALUEX, RFIXV_VALUEY, &
                            nFixRho, IFIXRHO_NODE, RFIXRHO_VALUE, nFixT, IFIXT_NODE, RFIXT_VALUE, &
                            nwall, wall
        use MVELOCIDADES
        use MVARIABGEN
        use MVARIABLES
        use MESTABILIZACION
        real(8), intent(in) :: dtmin
        integer, intent(in) :: nrk
        real(c_double), intent(in) :: gamm(*), dtl(*)
        type(cfd_bc_lists) :: bc
        integer(c_int), allocatable, target :: iv(:), ir(:), it(:), iw(:)
        real(c_double), allocatable, target :: vx(:), vy(:), vr(:), vt(:)
        iv = IFIXV_NODE(1:nFixV)
        vx = RFIXV_VALUEX(1:nFixV)
        vy = RFIXV_VALUEY(1:nFixV)
        ir = IFIXRHO_NODE(1:nFixRho)
        vr = RFIXRHO_VALUE(1:nFixRho)
        it = IFIXT_NODE(1:nFixT)
        vt = RFIXT_VALUE(1:nFixT)
        iw = reshape(wall(1:2, 1:nwall), [2*nwall])
        bc = cfd_bc_lists(nFixV, c_loc(iv), c_loc(vx), c_loc(vy), nFixRho, c_loc(ir), c_loc(vr), &
                          nFixT, c_loc(it), c_loc(vt), nwall, c_loc(iw))
        call rk_targets(bc, dtmin, nrk, gamm, dtl, U1, RHS, T, VEL_X, VEL_Y, RHO, E, P, RMACH, U, &
                        W_X, W_Y, M, dNx, dNy, area, SHOC, T_SUGN1, T_SUGN2, T_SUGN3)
    end subroutine shim_rk

    subroutine rk_targets(bc, dtmin, nrk, gamm, dtl, u1_, rhs_, t_, vx_, vy_, rho_, e_, p_, rm_, &
                          u_, wx_, wy_, m_, nx_, ny_, a_, sh_, s1_, s2_, s3_)
        type(cfd_bc_lists), intent(in) :: bc
        real(8), intent(in) :: dtmin
        integer, intent(in) :: nrk
        real(c_double), target :: gamm(*), dtl(*)
        real(c_double), target :: u1_(*), rhs_(*), t_(*), vx_(*), vy_(*), rho_(*), e_(*)
        real(c_double), target :: p_(*), rm_(*), u_(*), wx_(*), wy_(*), m_(*)
        real(c_double), target :: nx_(*), ny_(*), a_(*), sh_(*), s1_(*), s2_(*), s3_(*)
        type(cfd_rk_nodes) :: nd
        type(cfd_rk_elems) :: el
        nd = cfd_rk_nodes(c_loc(u1_), c_loc(rhs_), c_loc(t_), c_loc(vx_), c_loc(vy_), c_loc(rho_), &
                          c_loc(e_), c_loc(p_), c_loc(rm_), c_loc(u_), c_loc(wx_), c_loc(wy_), &
                          c_loc(m_), c_loc(gamm))
        el = cfd_rk_elems(c_loc(nx_), c_loc(ny_), c_loc(a_), c_loc(sh_), c_loc(s1_), c_loc(s2_), &
                          c_loc(s3_))
        call cfd_check(cfd_rk(shim_mesh_data(), shim_consts(), bc, nrk, dtmin, dtl, nd, el), 'RK')
    end subroutine rk_targets
end module cfd_shim_rk

! subrutinas.f90:88-126
subroutine deriv(hmin)
    use cfd_shim_ctx
    use MeshData, only: X, Y, area, HH, HHX, HHY, dNx, dNy
    implicit none
    real(8) hmin
    call cfd_check(cfd_deriv(shim_mesh_data(), X, Y, area, dNx, dNy, HH, HHX, HHY, hmin), 'deriv')
end subroutine deriv

! subrutinas.f90:128-153 (callers pass 0-2 ignored arguments; none are read)
subroutine MASAS()
    use cfd_shim_ctx
    use MeshData, only: M, area
    implicit none
    call cfd_check(cfd_masas(shim_mesh_data(), area, M), 'MASAS')
end subroutine MASAS

! subrutinas.f90:155-218
subroutine deltat(dtmin, dt)
    use cfd_shim_ctx
    use MeshData, only: nelem, area
    use MVELOCIDADES
    use MVARIABLES
    implicit none
    real(8) dt(nelem), dtmin
    call cfd_check(cfd_deltat(shim_mesh_data(), shim_consts(), area, T, VEL_X, VEL_Y, W_X, W_Y, &
                              dt, dtmin), 'deltat')
end subroutine deltat

! subrutinas.f90:332-446 (GAMA, FR, RMU, RHOINF, TINF, UINF, VINF are the
! InputData values the library already holds)
subroutine ESTAB(U, T, GAMA, FR, RMU, DTMIN, RHOINF, TINF, UINF, VINF, GAMM)
    use cfd_shim_ctx
    use MeshData, only: npoin, dNx, dNy
    use MVELOCIDADES
    use MESTABILIZACION
    implicit none
    real(8) U(4, npoin), T(npoin), GAMM(npoin)
    real(8) GAMA, FR, RMU, DTMIN, RHOINF, TINF, UINF, VINF
    call cfd_check(cfd_estab(shim_mesh_data(), shim_consts(), U, T, GAMM, VEL_X, VEL_Y, W_X, W_Y, &
                             dNx, dNy, DTMIN, SHOC, T_SUGN1, T_SUGN2, T_SUGN3), 'ESTAB')
end subroutine ESTAB

! subrutinas.f90:1036-1094: accumulates into RHS
subroutine FUENTE(dtl)
    use cfd_shim_ctx
    use MeshData, only: nelem, dNx, dNy, area
    use MVELOCIDADES
    use MVARIABGEN
    implicit none
    real(8) dtl(nelem)
    call cfd_check(cfd_fuente(shim_mesh_data(), RHS, U, W_X, W_Y, dNx, dNy, area, dtl), 'FUENTE')
end subroutine FUENTE

! subrutinas.f90:645-849 (NGAS = 0): ESTAB at stage 1, then NRK stages of
! calcRHS + FUENTE, the lumped-mass update, primitives and boundary values
subroutine RK(DTMIN, NRK, BANDERA, GAMM, dtl)
    use cfd_shim_rk
    use MeshData, only: npoin, nelem
    implicit none
    integer NRK, BANDERA
    real(8) DTMIN, GAMM(npoin), dtl(nelem)
    call shim_rk(DTMIN, NRK, GAMM, dtl)
end subroutine RK
