! mcx_driver.f90 -- a Fortran caller of libmcx.so through mcx_binding: the
! call sequence MCMax's driver makes for one Monte Carlo pass (Main.f:467-474)
! and the temperature inversion after it (Main.f:498-549), on a model handed
! over in a binary file.
!
!   mcx_driver input.bin output.bin
!   mcx_driver --sizes          (the bind(C) types' sizes; no GPU call)
!
! input.bin (native byte order, stream access): int32 nlam, ng, nr, nt,
! scat_how, randomwalk, nlanes, unused; int64 nphot, seed, stage0,
! epoch_deposits; float64 Lstar, Rstar_cm, Tstar, factRW; then the float64
! arrays of mcx_tables and mcx_grid in the order lam, nu, BB, Kabs, Ksca, F,
! IF11, IF12, Kp, Fstar, cos2phi, sin2phi, Kext, gsym, R, Theta, dens, V, w,
! thet (sizes as in include/mcx.h).
! output.bin: int32 rc of the pass; int64 nemit, ninteract, nabs, nscat,
! nescape, nstarhit, nsteps, nraar, nmaxinteract, ngeomfail; float64 EJv,
! EJv2, Eabs; int32 Ni; float64 T, dT (each [nr*nt]).
program mcx_driver
  use iso_c_binding
  use mcx_binding
  implicit none
  integer(c_int32_t) :: hi(8)
  integer(c_int64_t) :: hl(4), cnt(10)
  real(c_double) :: hd(4)
  real(c_double), allocatable, target :: lam(:), nu(:), BB(:), Kabs(:), Ksca(:), F(:), IF11(:), IF12(:), &
                                         Kp(:), Fstar(:), c2(:), s2(:), Kext(:), gsym(:), R(:), Theta(:), &
                                         dens(:), V(:), w(:), thet(:)
  real(c_double), allocatable, target :: EJv(:), EJv2(:), Eabs(:), Tbw(:), spec(:), scatspec(:), T(:), dT(:), dEJv(:)
  integer(c_int32_t), allocatable, target :: Ni(:), ispec(:), iscatspec(:), thick(:)
  type(mcx_tables) :: tab
  type(mcx_grid) :: grid
  type(mcx_run) :: run
  type(mcx_tallies) :: tal
  type(mcx_temps) :: tmp
  type(c_ptr) :: ctx
  integer :: u, nlam, ng, nr, nt, ncell, nt1
  integer(c_int) :: rc, rcp
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  if (trim(fin) == '--sizes') then
     write(*, '(a, 6(1x, i0))') 'sizes', c_sizeof(tab), c_sizeof(grid), c_sizeof(run), c_sizeof(tal), &
          c_sizeof(tmp), 0
     stop
  end if
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hi, hl, hd
  nlam = hi(1); ng = hi(2); nr = hi(3); nt = hi(4)
  ncell = nr * nt
  nt1 = MCX_TMAX + 1
  allocate(lam(nlam), nu(nlam), BB(nt1 * nlam), Kabs(ng * nlam), Ksca(ng * nlam), F(ng * nlam * 6 * 180), &
           IF11(ng * nlam), IF12(ng * nlam), Kp(ng * nt1), Fstar(nlam), c2(361), s2(361), Kext(ng * nlam), &
           gsym(ng * nlam), R(nr + 1), Theta(nt + 1), dens(ncell), V(ncell), w(ncell * ng), thet(nt + 1))
  read(u) lam, nu, BB, Kabs, Ksca, F, IF11, IF12, Kp, Fstar, c2, s2, Kext, gsym, R, Theta, dens, V, w, thet
  close(u)

  rc = mcx_create(ctx, 0)
  if (rc /= MCX_OK) stop 'mcx_create failed'
  ! the module globals lam, nu, BB, Grain(:) and D%Fstar (Modules.f:13, 59-83, 109-113)
  tab%nlam = nlam; tab%ng = ng; tab%tmax = MCX_TMAX
  tab%lam = c_loc(lam); tab%nu = c_loc(nu); tab%BB = c_loc(BB); tab%Kabs = c_loc(Kabs); tab%Ksca = c_loc(Ksca)
  tab%F = c_loc(F); tab%IF11 = c_loc(IF11); tab%IF12 = c_loc(IF12); tab%Kp = c_loc(Kp); tab%Fstar = c_loc(Fstar)
  tab%Lstar = hd(1); tab%cos2phi = c_loc(c2); tab%sin2phi = c_loc(s2); tab%scat_how = hi(5)
  tab%Kext = c_loc(Kext); tab%gsym = c_loc(gsym)
  call check(mcx_set_tables(ctx, tab), 'mcx_set_tables')
  ! D%R, D%Theta and C(i,j)%dens/%V/%w (Modules.f:85-121)
  grid%nr = nr; grid%nt = nt; grid%R = c_loc(R); grid%Theta = c_loc(Theta); grid%dens = c_loc(dens)
  grid%V = c_loc(V); grid%w = c_loc(w); grid%Rstar_cm = hd(2); grid%Tstar = hd(3); grid%thet = c_loc(thet)
  call check(mcx_set_model(ctx, grid), 'mcx_set_model')

  ! call MCRadiation(NphotTot, niter, dofastvisc) (Main.f:467-474)
  run%nphot = hl(1); run%seed = hl(2); run%maxinteract = 0; run%nlanes = hi(7); run%stage0 = hl(3)
  run%randomwalk = hi(6); run%factRW = hd(4); run%coop_lanes = -1; run%stage_ratio = 0
  run%precision = MCX_FP64; run%max_ms = 0d0; run%epoch_deposits = hl(4)
  rcp = mcx_transport(ctx, run)
  if (rcp /= MCX_OK .and. rcp /= MCX_EGEOM) call check(rcp, 'mcx_transport')

  allocate(EJv(ncell), EJv2(ncell), Eabs(ncell), Tbw(ncell), Ni(ncell), thick(ncell), T(ncell), dT(ncell), &
           dEJv(ncell), spec((MCX_NANGLE + 1) * nlam), scatspec((MCX_NANGLE + 1) * nlam), &
           ispec((MCX_NANGLE + 1) * nlam), iscatspec((MCX_NANGLE + 1) * nlam))
  tal%EJv = c_loc(EJv); tal%EJv2 = c_loc(EJv2); tal%Eabs = c_loc(Eabs); tal%Tbw = c_loc(Tbw); tal%Ni = c_loc(Ni)
  tal%spec = c_loc(spec); tal%scatspec = c_loc(scatspec); tal%ispec = c_loc(ispec)
  tal%iscatspec = c_loc(iscatspec); tal%thick = c_loc(thick)
  call check(mcx_get_tallies(ctx, tal), 'mcx_get_tallies')
  ! the temperature inversion (Main.f:498-549)
  tmp%T = c_loc(T); tmp%dT = c_loc(dT); tmp%dEJv = c_loc(dEJv)
  call check(mcx_temperature(ctx, tmp), 'mcx_temperature')
  call mcx_destroy(ctx)

  cnt = [tal%nemit, tal%ninteract, tal%nabs, tal%nscat, tal%nescape, tal%nstarhit, tal%nsteps, tal%nraar, &
         tal%nmaxinteract, tal%ngeomfail]
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) int(rcp, c_int32_t), cnt, EJv, EJv2, Eabs, Ni, T, dT
  close(u)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= MCX_OK) then
       write(*, '(a, " failed (", i0, "): ", a)') what, code, trim(mcx_error_text(ctx))
       stop 1
    end if
  end subroutine

end program mcx_driver
