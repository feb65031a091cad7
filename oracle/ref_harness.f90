! Test infrastructure only: a driver for the compiled reference solver.
!
! This program is linked against objects compiled from the reference sources
! in place (commonModules, dataLoader, pointNeighbor, mLaplace, biconjGrad,
! calcRHS, subrutinas, gcl, meshMove).  It reproduces the start-up and the
! time-step calling sequence of ns2DComp.ALE.f90 (:44-136 and :138-281,
! without smoothing, printing or restart files) and exposes each hot-path
! routine as a command acting on a binary state file, so that tests can feed
! the same inputs to the reference and to the project's C restatement.
!
! Usage (run inside a directory holding EULER.DAT, <name>-1.dat, <name>.dat):
!   ref_harness tables  out.bin          neighbour / CSR tables + loaded lists
!   ref_harness init    out.bin          state after start-up
!   ref_harness run  N  in.bin out.bin   N steps from a state (or "-" = start-up),
!                                        appending the state after every step
!   ref_harness rstwrite in.bin ITER     PRINTREST (ns2DComp.ALE.f90:898-917) of a
!                                        state: writes <name>.RST
!   ref_harness rstread out.bin          RESTART with IRESTART = 1 (:394-435):
!                                        U, T, GAMM from <name>.RST, then the state
!   ref_harness time N                   N steps from the start-up; prints their
!                                        wall time and the reference's own calcRHS
!                                        timer (calcrhs_t) over them
!   ref_harness <unit>  in.bin out.bin   one routine on a state:
!       deltat estab calcrhs fuente rk deriv masas laplace bicg fluid gcl normales
program ref_harness
   use InputData
   use MeshData
   use MVELOCIDADES
   use MVARIABGEN
   use MVARIABLES
   use MESTABILIZACION
   use PointNeighbor
   use Mlaplace
   use BiconjGrad
   use calcRHS_mod
   use gcl_mod, only: gcl_main => main, putW, putArea
   use MeshMove
   use Mnormales
   use TIMERS
   implicit none

   character(len=64) :: cmd, arg2, arg3, arg4
   integer :: nsteps, i, istep
   real(8), allocatable :: GAMM(:), DTL(:), DT(:), x1(:), y1(:)
   real(8), allocatable :: pos_aux(:), bvec(:), w_x_old(:), w_y_old(:), area_old(:)
   logical, allocatable :: smooth_fix(:)
   real(8) :: DTMIN, TIME, DTMIN1, HMIN, PORC, DTFACT, BANDERA_R, ITER_R
   integer :: BANDERA, ITER, NRK

   interface
      subroutine MASAS()
      end subroutine
   end interface

   call get_command_argument(1, cmd)
   call get_command_argument(2, arg2)
   call get_command_argument(3, arg3)
   call get_command_argument(4, arg4)

   calcrhs_t = 0; cuarto_t = 0; estab_t = 0; fuente_t = 0; grad_t = 0
   forces_t = 0; transf_t = 0; normales_t = 0; deriv_t = 0; masas_t = 0; laplace_t = 0

   call readInputData
   call loadMeshData
   allocate(GAMM(npoin), DTL(nelem), DT(nelem), x1(npoin), y1(npoin), smooth_fix(npoin))
   allocate(pos_aux(npoin), bvec(npoin), w_x_old(npoin), w_y_old(npoin), area_old(nelem))
   call startup()

   select case (trim(cmd))
   case ('tables')
      call dump_tables(trim(arg2))
   case ('init')
      call save_state(trim(arg2), .false.)
   case ('time')
      read(arg2, *) nsteps
      call time_steps(nsteps)
   case ('rstwrite')
      call load_state(trim(arg2))
      read(arg3, *) i
      call PRINTREST(i, npoin, U, T, GAMM, TIME, FILENAME)
   case ('rstread')
      IRESTART = 1
      call RESTART(GAMM)
      call save_state(trim(arg2), .false.)
   case ('run')
      read(arg2, *) nsteps
      if (trim(arg3) /= '-') call load_state(trim(arg3))
      open(21, file=trim(arg4), access='stream', form='unformatted', status='replace')
      close(21)
      do istep = 1, nsteps
         call one_step()
         call save_state(trim(arg4), .true.)
      end do
   case default
      call load_state(trim(arg2))
      call unit_call(trim(cmd))
      call save_state(trim(arg3), .false.)
   end select

contains

   subroutine startup()
      ! ns2DComp.ALE.f90:59-136 (RESTART free-stream branch :404-420 inlined;
      ! smoothing is not run: the generated meshes are used as they are)
      integer :: inod
      real(8) :: rhoamb, tamb, uamb, vamb, pamb, energia
      GAMM = GAMA
      rhoamb = RHO_INF; tamb = T_INF; uamb = U_inf; vamb = V_inf; pamb = RHO_INF*FR*T_INF
      do inod = 1, npoin
         U(1, inod) = rhoamb
         U(2, inod) = rhoamb*uamb
         U(3, inod) = rhoamb*vamb
         energia = pamb/((GAMM(inod) - 1.d0)*rhoamb) + .5d0*(uamb**2.d0 + vamb**2.d0)
         U(4, inod) = energia*rhoamb
         VEL_X(inod) = uamb
         VEL_Y(inod) = vamb
         T(inod) = tamb
      end do
      smooth_fix = .false.
      x1 = 0.d0; y1 = 0.d0
      if (NGAS .ne. 1) GAMM = GAMA
      call NORMALES
      call DERIV(HMIN)
      call MASAS()
      call laplace(inpoel, area, dNx, dNy, nelem, npoin)
      TIME = 0.d0; ITER = 0; DTMIN = 0.d0
      NRK = 4
      W_X = -0.d0
      W_Y = 0.d0
      BANDERA = 1
      call setNewmarkCondition
      allocate(xpos(npoin), ypos(npoin))
      xpos = 0.d0; ypos = 0.d0
      U1 = U; RHS = 0.d0; UN = 0.d0
      RHO = U(1, :); E = U(4, :)/U(1, :); P = 0.d0; RMACH = 0.d0
      DTL = 0.d0; DT = 0.d0; SHOC = 0.d0; T_SUGN1 = 0.d0; T_SUGN2 = 0.d0; T_SUGN3 = 0.d0
      DTMIN1 = 0.d0
      pos_aux = 0.d0; bvec = 0.d0; w_x_old = 0.d0; w_y_old = 0.d0; area_old = area
   end subroutine startup

   subroutine one_step()
      ! ns2DComp.ALE.f90:140-281 without the print block
      integer :: ipoin
      ITER = ITER + 1
      call DELTAT(DTMIN, DT)
      if (BANDERA .eq. 1) then
         DTMIN1 = DTMIN
         BANDERA = 2
      end if
      PORC = dabs((DTMIN - DTMIN1)/DTMIN)
      if (100.d0*PORC .le. 1.d0) then
         DTMIN = DTMIN1
      else
         DTMIN1 = DTMIN
         BANDERA = 2
      end if
      if (ITLOCAL .ne. 0) then
         DTFACT = 1.d0 - dexp(-ITER*4.6d0/ITLOCAL)
         DTL = DTMIN*DTFACT + DT*(1.d0 - DTFACT)
      else
         DTL = DTMIN
      end if
      TIME = TIME + DTMIN
      do ipoin = 1, npoin
         U1(:, ipoin) = U(:, ipoin)
      end do
      call RK(DTMIN, NRK, BANDERA, GAMM, DTL)
      call fluidStructure(DTMIN, TIME, smooth_fix, x1, y1)
      BANDERA = BANDERA + 1
      if (MOVING .eq. 1) then
         call NORMALES
         call DERIV(HMIN)
         call MASAS()
         call laplace(inpoel, area, dNx, dNy, nelem, npoin)
      end if
      do ipoin = 1, npoin
         U(:, ipoin) = U1(:, ipoin)
      end do
   end subroutine one_step

   subroutine time_steps(n)
      ! n steps timed with system_clock, as the reference's timer macro does
      ! (the start-up has already written every array once)
      integer, intent(in) :: n
      integer(8) :: c0, c1, rate
      real(8) :: t_calc0
      t_calc0 = calcrhs_t
      call system_clock(c0, rate)
      do istep = 1, n
         call one_step()
      end do
      call system_clock(c1)
      write(*, '(A, I8, A, ES16.8, A, ES16.8)') 'timed_steps ', n, ' wall_s ', &
         dble(c1 - c0)/dble(rate), ' calcrhs_s ', dble(calcrhs_t - t_calc0)
   end subroutine time_steps

   subroutine unit_call(name)
      character(len=*), intent(in) :: name
      real(8) :: tprev
      select case (name)
      case ('deltat')
         call DELTAT(DTMIN, DT)
      case ('estab')
         call ESTAB(U, T, GAMA, FR, FMU, DTMIN, RHO_inf, T_inf, U_inf, V_inf, GAMM)
      case ('calcrhs')
         UN = 0.d0
         call calcRHS(RHS, U, UN, dNx, dNy, area, shoc, DTL, t_sugn1, t_sugn2, t_sugn3, &
                      inpoel, nelem, npoin)
      case ('fuente')
         call FUENTE(DTL)
      case ('rk')
         call RK(DTMIN, 4, BANDERA, GAMM, DTL)
      case ('deriv')
         call DERIV(HMIN)
      case ('masas')
         call MASAS()
      case ('laplace')
         call laplace(inpoel, area, dNx, dNy, nelem, npoin)
      case ('bicg')
         call biCG(lap_sparse, lap_idx, lap_rowptr, lap_diag, xpos, bvec, &
                   pos_aux(1:nnmove), ilaux(1:nnmove), npoin, nnmove)
      case ('fluid')
         ! the private previous pitch angle is set by one call at the previous
         ! time; the state is then reloaded and the step's call is made
         tprev = TIME - DTMIN
         call fluidStructure(DTMIN, tprev, smooth_fix, x1, y1)
         call load_state(trim(arg2))
         call fluidStructure(DTMIN, TIME, smooth_fix, x1, y1)
      case ('gcl')
         call putW(w_x_old, w_y_old)
         call putArea(area_old)
         call gcl_main(M, W_X, W_Y, dNx, dNy, area, inpoel, DTMIN)
      case ('normales')
         call NORMALES
         call NORMALVEL
      case default
         write (*, *) 'unknown command ', name
         stop 2
      end select
   end subroutine unit_call

   subroutine save_state(fname, append)
      character(len=*), intent(in) :: fname
      logical, intent(in) :: append
      integer(8) :: hdr(3)
      hdr = (/ int(npoin, 8), int(nelem, 8), int(lap_rowptr(npoin + 1), 8) /)
      if (append) then
         open(21, file=fname, access='stream', form='unformatted', position='append', status='old')
      else
         open(21, file=fname, access='stream', form='unformatted', status='replace')
      end if
      BANDERA_R = dble(BANDERA); ITER_R = dble(ITER)
      write(21) hdr
      write(21) DTMIN, TIME, DTMIN1, BANDERA_R, ITER_R, HMIN, fx(1), fy(1), rm(1), 0.d0
      write(21) U, U1, RHS, T, VEL_X, VEL_Y, W_X, W_Y, X, Y, x1, y1, RHO, E, P, RMACH, M, GAMM
      write(21) xpos, ypos, pos_aux, bvec, w_x_old, w_y_old
      write(21) DTL, DT, SHOC, T_SUGN1, T_SUGN2, T_SUGN3, area, dNx, dNy, HH, HHX, HHY, area_old
      write(21) lap_sparse, lap_diag
      close(21)
   end subroutine save_state

   subroutine load_state(fname)
      character(len=*), intent(in) :: fname
      integer(8) :: hdr(3)
      real(8) :: s(10)
      open(22, file=fname, access='stream', form='unformatted', status='old')
      read(22) hdr
      if (hdr(1) /= npoin .or. hdr(2) /= nelem .or. hdr(3) /= lap_rowptr(npoin + 1)) stop 'state size'
      read(22) s
      DTMIN = s(1); TIME = s(2); DTMIN1 = s(3); BANDERA = nint(s(4)); ITER = nint(s(5)); HMIN = s(6)
      read(22) U, U1, RHS, T, VEL_X, VEL_Y, W_X, W_Y, X, Y, x1, y1, RHO, E, P, RMACH, M, GAMM
      read(22) xpos, ypos, pos_aux, bvec, w_x_old, w_y_old
      read(22) DTL, DT, SHOC, T_SUGN1, T_SUGN2, T_SUGN3, area, dNx, dNy, HH, HHX, HHY, area_old
      read(22) lap_sparse, lap_diag
      close(22)
   end subroutine load_state

   subroutine dump_tables(fname)
      character(len=*), intent(in) :: fname
      integer(8) :: hdr(8)
      integer :: iset_flat(3*nsets), k, s, j
      real(8) :: consts(20)
      hdr = (/ int(npoin, 8), int(nelem, 8), int(size(esup1), 8), int(size(psup1), 8), &
               int(lap_rowptr(npoin + 1), 8), int(nfixrho, 8), int(nfixv, 8), int(nfixt, 8) /)
      open(23, file=fname, access='stream', form='unformatted', status='replace')
      write(23) hdr
      write(23) esup2, esup1, psup2, psup1, lap_rowptr, lap_idx
      write(23) ifixrho_node(1:nfixrho), rfixrho_value(1:nfixrho)
      write(23) ifixv_node(1:nfixv), rfixv_valuex(1:nfixv), rfixv_valuey(1:nfixv)
      write(23) ifixt_node(1:nfixt), rfixt_value(1:nfixt)
      write(23) int(nset_numb, 8), int(nmove, 8), int(nfix_move, 8), int(nwall, 8)
      k = 0
      do s = 1, nset_numb
         do j = 1, ielem_sets(s)
            iset_flat(k + 1) = iset(1, s, j)
            iset_flat(k + 2) = iset(2, s, j)
            iset_flat(k + 3) = iset(3, s, j)
            k = k + 3
         end do
      end do
      write(23) int(k/3, 8)
      write(23) iset_flat(1:k)
      write(23) ilaux(1:nnmove)
      consts = 0.d0
      consts(1:17) = (/ U_inf, V_inf, T_inf, RHO_inf, P_inf, C_inf, FR, FMU, FK, FCv, GAMA, &
                        FSAFE, CTE, XREF(1), YREF(1), MACH_inf, HMIN /)
      write(23) consts
      close(23)
   end subroutine dump_tables

end program ref_harness
