! kestrel_step.f90 -- iso_c_binding interface to the MI355X time-step library
! (include/kestrel_step.h).  A Kestrel build adds this module and, in
! TimeStepper.f90, replaces the body of IntegrateTo (src/TimeStepper.f90:116-277)
! with kst_integrate_to on a handle created after LoadSourceConditions.
!
! The conventions are those of the reference's own C bridges
! (src/GeoTiffRead.f90:43-100, src/utm.f90:55-100): bind(C) interface blocks,
! a bind(C) struct, int status codes mapped to FatalErrorMessage.
module kestrel_step_binding
   use, intrinsic :: iso_c_binding
   implicit none
   private
   public :: kst_config, kst_flux_source, kst_stats
   public :: kst_config_defaults, kst_config_derive, kst_create, kst_set_topography
   public :: kst_set_state, kst_integrate_to, kst_step_n, kst_get_state, kst_get_maxima
   public :: kst_last_error, kst_destroy, kst_set_grid_cap, kst_volume, kst_get_output, kst_set_maxima
   public :: KST_TILE_OFF, KST_TILE_ACTIVE, KST_TILE_GHOST

   integer(c_int8_t), parameter :: KST_TILE_OFF = 0, KST_TILE_ACTIVE = 1, KST_TILE_GHOST = 2

   type, bind(C) :: kst_flux_source
      real(c_double) :: x, y, radius
      integer(c_int32_t) :: n_series, num_cells
      type(c_ptr) :: time, flux, psi
   end type kst_flux_source

   type, bind(C) :: kst_config
      integer(c_int32_t) :: nx_tiles, ny_tiles, nx_per_tile, ny_per_tile
      real(c_double) :: dx, dy, x_size, y_size
      integer(c_int32_t) :: bc, limiter, geometric_factors, drag, erosion, deposition
      integer(c_int32_t) :: ero_transition, morpho_damping, fswitch, morphodynamics
      integer(c_int32_t) :: tile_buffer, track_maxima, precision, n_sources
      real(c_double) :: g, rhow, rhos
      real(c_double) :: chezy_co, coulomb_co, manning_co
      real(c_double) :: pouliquen_min, pouliquen_max, pouliquen_intermediate, pouliquen_beta
      real(c_double) :: edwards_betastar, edwards_kappa, edwards_gamma
      real(c_double) :: ero_rate, ero_rate_granular, ero_depth, ero_critical_height
      real(c_double) :: voellmy_switch_rate, voellmy_switch_value
      real(c_double) :: max_pack, bed_porosity, solid_diameter, eddy_viscosity
      real(c_double) :: cfl, max_dt, height_threshold
      real(c_double) :: gred, rep, ws0, nsettling, critical_shields, diffusive_time_scale
      type(c_ptr) :: sources
      real(c_double) :: bc_hn, bc_u, bc_v, bc_psi, sponge_strength
   end type kst_config

   type, bind(C) :: kst_stats
      integer(c_int64_t) :: steps, rejected, rhs_evals, cell_updates
      integer(c_int32_t) :: active_tiles, ghost_tiles
      real(c_double) :: t, last_dt, min_dt, max_dt
      integer(c_int64_t) :: redist_cells, redist_ghost
   end type kst_stats

   interface
      subroutine kst_config_defaults(cfg) bind(C, name="kst_config_defaults")
         import :: kst_config
         type(kst_config), intent(inout) :: cfg
      end subroutine
      subroutine kst_config_derive(cfg) bind(C, name="kst_config_derive")
         import :: kst_config
         type(kst_config), intent(inout) :: cfg
      end subroutine
      function kst_create(cfg, h) result(rc) bind(C, name="kst_create")
         import :: kst_config, c_ptr, c_int
         type(kst_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: h
         integer(c_int) :: rc
      end function
      function kst_set_topography(h, b0v) result(rc) bind(C, name="kst_set_topography")
         import :: c_ptr, c_double, c_int
         type(c_ptr), value :: h
         real(c_double), intent(in) :: b0v(*)
         integer(c_int) :: rc
      end function
      function kst_set_state(h, w, rhu, rhv, hnpsi, btv, tstate, t, hnmax0) result(rc) &
            bind(C, name="kst_set_state")
         import :: c_ptr, c_double, c_int, c_int8_t
         type(c_ptr), value :: h
         real(c_double), intent(in) :: w(*), rhu(*), rhv(*), hnpsi(*)
         type(c_ptr), value :: btv
         integer(c_int8_t), intent(in) :: tstate(*)
         real(c_double), value :: t
         type(c_ptr), value :: hnmax0
         integer(c_int) :: rc
      end function
      function kst_integrate_to(h, tend, stats) result(rc) bind(C, name="kst_integrate_to")
         import :: c_ptr, c_double, c_int, kst_stats
         type(c_ptr), value :: h
         real(c_double), value :: tend
         type(kst_stats), intent(out) :: stats
         integer(c_int) :: rc
      end function
      function kst_step_n(h, n, stats) result(rc) bind(C, name="kst_step_n")
         import :: c_ptr, c_int64_t, c_int, kst_stats
         type(c_ptr), value :: h
         integer(c_int64_t), value :: n
         type(kst_stats), intent(out) :: stats
         integer(c_int) :: rc
      end function
      function kst_get_state(h, w, rhu, rhv, hnpsi, hn, btv, tstate, t) result(rc) &
            bind(C, name="kst_get_state")
         import :: c_ptr, c_double, c_int, c_int8_t
         type(c_ptr), value :: h
         real(c_double), intent(out) :: w(*), rhu(*), rhv(*), hnpsi(*), hn(*)
         type(c_ptr), value :: btv
         integer(c_int8_t), intent(out) :: tstate(*)
         real(c_double), intent(out) :: t
         integer(c_int) :: rc
      end function
      function kst_get_maxima(h, maxima) result(rc) bind(C, name="kst_get_maxima")
         import :: c_ptr, c_int
         type(c_ptr), value :: h
         type(c_ptr), intent(in) :: maxima(11)
         integer(c_int) :: rc
      end function
      function kst_get_tile_bt(h, bt_tiles) result(rc) bind(C, name="kst_get_tile_bt")
         import :: c_ptr, c_double, c_int
         type(c_ptr), value :: h
         real(c_double), intent(out) :: bt_tiles(*)
         integer(c_int) :: rc
      end function
      function kst_get_cell_topography(h, b0c, btc, dbdx, dbdy) result(rc) &
            bind(C, name="kst_get_cell_topography")
         import :: c_ptr, c_int
         type(c_ptr), value :: h, b0c, btc, dbdx, dbdy
         integer(c_int) :: rc
      end function
      function kst_set_cell_topography(h, btc, dbdx, dbdy, cells) result(rc) &
            bind(C, name="kst_set_cell_topography")
         import :: c_ptr, c_int
         type(c_ptr), value :: h, btc, dbdx, dbdy, cells
         integer(c_int) :: rc
      end function
      function kst_last_error(h, buf, buflen) result(rc) bind(C, name="kst_last_error")
         import :: c_ptr, c_char, c_int
         type(c_ptr), value :: h
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: buflen
         integer(c_int) :: rc
      end function
      subroutine kst_destroy(h) bind(C, name="kst_destroy")
         import :: c_ptr
         type(c_ptr), value :: h
      end subroutine
      function kst_get_output(h, max_tiles, fields, tiles, n_tiles) bind(C, name="kst_get_output") result(rc)
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: max_tiles
         real(c_double), intent(out) :: fields(*)
         integer(c_int32_t), intent(out) :: tiles(*)
         integer(c_int32_t), intent(out) :: n_tiles
         integer(c_int) :: rc
      end function
      function kst_set_maxima(h, maxima) bind(C, name="kst_set_maxima") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: h
         type(c_ptr), intent(in) :: maxima(11)
         integer(c_int) :: rc
      end function
      function kst_volume(h, out) bind(C, name="kst_volume") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: out(6)
         integer(c_int) :: rc
      end function
      function kst_set_grid_cap(h, tiles) bind(C, name="kst_set_grid_cap") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: h
         integer(c_int), value :: tiles
         integer(c_int) :: rc
      end function
   end interface
end module kestrel_step_binding
