! mcx_binding.f90 -- ISO_C_BINDING interface of libmcx.so (include/mcx.h)
! for a Fortran caller: the derived types mirror the C structures field by
! field (bind(C) gives them the C layout) and the interfaces bind the entry
! points.  MCMax's driver calls them where it calls MCRadiation
! (Main.f:467-474) and inverts the tallies (Main.f:498-549).
module mcx_binding
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MCX_OK = 0, MCX_EINVAL = -1, MCX_EDEVICE = -2, MCX_ENOMEM = -3, &
                               MCX_ESTATE = -4, MCX_EGEOM = -5, MCX_ETIME = -6
  integer(c_int32_t), parameter :: MCX_TMAX = 10000, MCX_NANGLE = 30
  integer(c_int32_t), parameter :: MCX_FP64 = 0, MCX_FP32 = 1

  type, bind(C) :: mcx_tables
     integer(c_int32_t) :: nlam, ng, tmax
     type(c_ptr) :: lam, nu, BB, Kabs, Ksca, F, IF11, IF12, Kp, Fstar
     real(c_double)     :: Lstar
     type(c_ptr) :: cos2phi, sin2phi
     integer(c_int32_t) :: scat_how
     type(c_ptr) :: Kext, gsym
  end type

  type, bind(C) :: mcx_grid
     integer(c_int32_t) :: nr, nt
     type(c_ptr) :: R, Theta, dens, V, w
     real(c_double) :: Rstar_cm, Tstar
     type(c_ptr) :: thet
  end type

  type, bind(C) :: mcx_run
     integer(c_int64_t) :: nphot
     integer(c_int64_t) :: seed              ! uint64 on the C side
     integer(c_int32_t) :: maxinteract, nlanes
     integer(c_int64_t) :: stage0
     integer(c_int32_t) :: randomwalk
     real(c_double)     :: factRW
     integer(c_int32_t) :: coop_lanes, stage_ratio, precision
     real(c_double)     :: max_ms
     integer(c_int64_t) :: epoch_deposits
  end type

  type, bind(C) :: mcx_tallies
     type(c_ptr) :: EJv, EJv2, Eabs, Tbw, Ni, spec, scatspec, ispec, iscatspec
     integer(c_int64_t) :: nemit, ninteract, nabs, nscat, nescape, nstarhit, &
                           nsteps, nraar, nmaxinteract, ngeomfail
     real(c_double) :: tautot, Eescape
     integer(c_int64_t) :: nrw, noverflow
     type(c_ptr) :: thick
     integer(c_int64_t) :: maxsteps
     real(c_double) :: t_done(5)
     integer(c_int64_t) :: nvisit_overflow, ntimeout, ntallyatomic
  end type

  type, bind(C) :: mcx_temps
     type(c_ptr) :: T, dT, dEJv
  end type

  type, bind(C) :: mcx_struct
     type(c_ptr)        :: R_av, theta_av
     real(c_double)     :: Mstar
     type(c_ptr)        :: shscale, gshscale, settle, gascoupled
     integer(c_int32_t) :: mpstr
     type(c_ptr)        :: FradZ
  end type

  type, bind(C) :: mcx_density
     type(c_ptr) :: dens, dens0, gasdens, w, w0
  end type

  type, bind(C) :: mcx_cellstate
     type(c_ptr) :: T, dT, dens, ddens, EJv, dEJv
  end type

  interface
     integer(c_int) function mcx_create(ctx, device) bind(C, name="mcx_create")
       import; type(c_ptr) :: ctx; integer(c_int), value :: device
     end function
     subroutine mcx_destroy(ctx) bind(C, name="mcx_destroy")
       import; type(c_ptr), value :: ctx
     end subroutine
     type(c_ptr) function mcx_last_error(ctx) bind(C, name="mcx_last_error")
       import; type(c_ptr), value :: ctx
     end function
     integer(c_int) function mcx_set_tables(ctx, t) bind(C, name="mcx_set_tables")
       import; type(c_ptr), value :: ctx; type(mcx_tables) :: t
     end function
     integer(c_int) function mcx_set_model(ctx, g) bind(C, name="mcx_set_model")
       import; type(c_ptr), value :: ctx; type(mcx_grid) :: g
     end function
     integer(c_int) function mcx_transport(ctx, r) bind(C, name="mcx_transport")
       import; type(c_ptr), value :: ctx; type(mcx_run) :: r
     end function
     integer(c_int) function mcx_get_tallies(ctx, o) bind(C, name="mcx_get_tallies")
       import; type(c_ptr), value :: ctx; type(mcx_tallies) :: o
     end function
     integer(c_int) function mcx_temperature(ctx, o) bind(C, name="mcx_temperature")
       import; type(c_ptr), value :: ctx; type(mcx_temps) :: o
     end function
     integer(c_int) function mcx_last_timing(ctx, kms, pms) bind(C, name="mcx_last_timing")
       import; type(c_ptr), value :: ctx; real(c_double) :: kms, pms
     end function
     integer(c_int) function mcx_set_temperature(ctx, T, dT, Ni) bind(C, name="mcx_set_temperature")
       import; type(c_ptr), value :: ctx, T, dT, Ni
     end function
     integer(c_int) function mcx_diffuse(ctx, nphotmin, dTmax, factRW, R_av, theta_av, &
                                         T, thick, EJv, ncells) bind(C, name="mcx_diffuse")
       import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: nphotmin
       real(c_double), value :: dTmax, factRW; type(c_ptr), value :: R_av, theta_av, T, thick, EJv
       integer(c_int32_t) :: ncells
     end function
     integer(c_int) function mcx_temp_average(ctx, fw, first, Tav) bind(C, name="mcx_temp_average")
       import; type(c_ptr), value :: ctx; real(c_double), value :: fw
       integer(c_int32_t), value :: first; type(c_ptr), value :: Tav
     end function
     integer(c_int) function mcx_disk_structure(ctx, st, Temp, d, nsteps) bind(C, name="mcx_disk_structure")
       import; type(c_ptr), value :: ctx; type(mcx_struct) :: st; type(c_ptr), value :: Temp
       type(mcx_density) :: d; integer(c_int64_t) :: nsteps
     end function
     integer(c_int) function mcx_compare_cells(ctx, prev, cur, errT, errRho, errE) &
                    bind(C, name="mcx_compare_cells")
       import; type(c_ptr), value :: ctx; type(mcx_cellstate) :: prev, cur
       real(c_double) :: errT, errRho, errE
     end function
  end interface

contains

  ! the text of the context's last error
  function mcx_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(mcx_last_error(ctx), s, [256])
    do i = 1, 256
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function

end module mcx_binding
