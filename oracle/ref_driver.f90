! ref_driver.f90 -- drives the reference's own hydraulic operator.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference's Fortran modules
! (compiled in place from src/ by oracle/Makefile) into oracle/_ref.  It
! builds a GridType from a global state, the way InitialiseGrid / AllocateTile /
! AddGhostTiles / GetHeights would have left it (src/Grid.f90:166-220,
! src/UpdateTiles.f90:120-481, src/dem.f90:360-393), and calls the reference's
! CalculateHydraulicRHS (src/HydraulicRHS.f90:64-136), CalculateMorphodynamicRHS
! (src/MorphodynamicRHS.f90:72-149) or AddToRedistList + RedistributeGrid
! (src/Redistribute.f90:69-244) on it.  The results pin the C restatement bit
! for bit.
module kestrel_ref_driver
   use, intrinsic :: iso_c_binding
   use set_precision_module, only: wp
   use runsettings_module, only: RunSet, SetImmutableRunSettings
   use grid_module, only: GridType, TileType, InitialiseGrid, GridToPhysical
   use hydraulic_rhs_module, only: CalculateHydraulicRHS, CalculateLimitedDerivsBoundary
   use morphodynamic_rhs_module, only: ComputeCellCentredTopographicData, &
      ComputeInterfacialTopographicData, CalculateMorphodynamicRHS
   use redistribute_module, only: RedistList, AddToRedistList, RedistributeGrid
   use limiters_module
   use closures_module
   use kestrel_step_binding, only: kst_config, kst_flux_source
   use topog_funcs_module
   use interpolate_2d_module, only: Bicubic
   implicit none
   private
   public :: kref_rhs, kref_morpho_rhs, kref_redistribute

   ! Module variables: RunSet holds a proj_transformer component, and a local
   ! RunSet would be finalized through utm.f90's bindings (PROJ).  Module
   ! variables are never finalized, so utm.o is not needed at link time.
   type(RunSet), target, save :: rp
   type(GridType), target, save :: grid

contains

   subroutine bind_closures(c)
      type(kst_config), intent(in) :: c
      select case (c%limiter)
      case (0); limiter => MinMod1
      case (1); limiter => MinMod2
      case (2); limiter => vanAlbada
      case (3); limiter => WENO
      case default; limiter => LimiterNone
      end select
      if (c%geometric_factors /= 0) then
         GeometricCorrectionFactor => IversonOuyangGeometricCorrectionFactor
         GeometricCorrectionFactor_gradin => IversonOuyangGeometricCorrectionFactor_gradin
      else
         GeometricCorrectionFactor => NoGeometricCorrectionFactor
         GeometricCorrectionFactor_gradin => NoGeometricCorrectionFactor_gradin
      end if
      select case (c%drag)
      case (0); DragClosure => ChezyDrag
      case (1); DragClosure => CoulombDrag
      case (2); DragClosure => VoellmyDrag
      case (3); DragClosure => PouliquenDrag
      case (4); DragClosure => Edwards2019Drag
      case (5); DragClosure => ManningDrag
      case default; DragClosure => VariableDrag
      end select
      select case (c%fswitch)
      case (0); fswitch => tanhSwitch
      case (1); fswitch => rat3Switch
      case (2); fswitch => cosSwitch
      case (3); fswitch => equalSwitch
      case (4); fswitch => zeroSwitch
      case (5); fswitch => oneSwitch
      case (6); fswitch => linearSwitch
      case default; fswitch => stepSwitch
      end select
      select case (c%erosion)
      case (1); ErosionClosure => SimpleErosion
      case (2); ErosionClosure => FluidErosion
      case (3); ErosionClosure => GranularErosion
      case (4); ErosionClosure => MixedErosion
      case default; ErosionClosure => NoErosion
      end select
      select case (c%deposition)
      case (1); DepositionClosure => SimpleHinderedSettling
      case (2); DepositionClosure => SpearmanManningHinderedSettling
      case default; DepositionClosure => NoDeposition
      end select
      select case (c%ero_transition)
      case (1); ErosionTransition => SmoothErosionTransition
      case (2); ErosionTransition => StepErosionTransition
      case default; ErosionTransition => NoErosionTransition
      end select
      select case (c%morpho_damping)
      case (1); MorphoDamping => tanhMorphoDamping
      case (2); MorphoDamping => rat3MorphoDamping
      case default; MorphoDamping => NoMorphoDamping
      end select
   end subroutine bind_closures

   subroutine fill_runset(c, rp)
      type(kst_config), intent(in) :: c
      type(RunSet), intent(inout) :: rp
      type(kst_flux_source), pointer :: srcs(:)
      real(c_double), pointer :: arr(:)
      integer :: k, n
      rp%Vars%w = 1; rp%Vars%rhoHnu = 2; rp%Vars%rhoHnv = 3; rp%Vars%Hnpsi = 4
      rp%Vars%Hn = 5; rp%Vars%u = 6; rp%Vars%v = 7; rp%Vars%psi = 8; rp%Vars%rho = 9
      rp%Vars%b0 = 10; rp%Vars%bt = 11; rp%Vars%dbdx = 12; rp%Vars%dbdy = 13
      rp%nDimensions = 13
      if (allocated(rp%iFlux)) deallocate(rp%iFlux)
      if (allocated(rp%iDesing)) deallocate(rp%iDesing)
      if (allocated(rp%ImplicitStep)) deallocate(rp%ImplicitStep)
      if (allocated(rp%FluxSources)) deallocate(rp%FluxSources)
      call SetImmutableRunSettings(rp)
      rp%nXpertile = c%nx_per_tile; rp%nYpertile = c%ny_per_tile
      rp%nXtiles = c%nx_tiles; rp%nYtiles = c%ny_tiles
      rp%nTiles = c%nx_tiles * c%ny_tiles
      rp%deltaX = c%dx; rp%deltaY = c%dy
      rp%deltaXRecip = 1.0_wp / c%dx; rp%deltaYRecip = 1.0_wp / c%dy
      rp%xSize = c%x_size; rp%ySize = c%y_size
      rp%Xtilesize = c%dx * c%nx_per_tile; rp%Ytilesize = c%dy * c%ny_per_tile
      rp%isOneD = .false.
      rp%geometric_factors = (c%geometric_factors /= 0)
      rp%g = c%g; rp%rhow = c%rhow; rp%rhos = c%rhos; rp%gred = c%gred
      rp%ChezyCo = c%chezy_co; rp%ManningCo = c%manning_co; rp%CoulombCo = c%coulomb_co
      rp%PouliquenMinSlope = c%pouliquen_min; rp%PouliquenMaxSlope = c%pouliquen_max
      rp%PouliquenIntermediateSlope = c%pouliquen_intermediate; rp%PouliquenBeta = c%pouliquen_beta
      rp%Edwards2019betastar = c%edwards_betastar; rp%Edwards2019kappa = c%edwards_kappa
      rp%Edwards2019Gamma = c%edwards_gamma
      rp%VoellmySwitchRate = c%voellmy_switch_rate; rp%VoellmySwitchValue = c%voellmy_switch_value
      rp%MorphodynamicsOn = (c%morphodynamics /= 0)
      rp%EroRate = c%ero_rate; rp%EroRateGranular = c%ero_rate_granular
      rp%CriticalShields = c%critical_shields; rp%EroDepth = c%ero_depth
      rp%EroCriticalHeight = c%ero_critical_height; rp%BedPorosity = c%bed_porosity
      rp%maxPack = c%max_pack; rp%SolidDiameter = c%solid_diameter
      rp%Rep = c%rep; rp%ws0 = c%ws0; rp%nsettling = c%nsettling
      rp%EddyViscosity = c%eddy_viscosity
      rp%heightThreshold = c%height_threshold; rp%TileBuffer = c%tile_buffer
      rp%cfl = c%cfl; rp%diffusiveTimeScale = c%diffusive_time_scale; rp%maxdt = c%max_dt
      rp%SpongeLayer = .false.
      rp%nSources = c%n_sources
      if (c%n_sources > 0) then
         call c_f_pointer(c%sources, srcs, [c%n_sources])
         allocate(rp%FluxSources(c%n_sources))
         do k = 1, c%n_sources
            n = srcs(k)%n_series
            rp%FluxSources(k)%x = srcs(k)%x; rp%FluxSources(k)%y = srcs(k)%y
            rp%FluxSources(k)%radius = srcs(k)%radius
            rp%FluxSources(k)%nFluxSeries = n
            rp%FluxSources(k)%NumCellsInSrc = srcs(k)%num_cells
            call c_f_pointer(srcs(k)%time, arr, [n]); rp%FluxSources(k)%time = arr
            call c_f_pointer(srcs(k)%flux, arr, [n]); rp%FluxSources(k)%flux = arr
            call c_f_pointer(srcs(k)%psi, arr, [n]); rp%FluxSources(k)%psi = arr
         end do
      end if
   end subroutine fill_runset

   ! Build grid%tileContainer from a global state the way InitialiseGrid /
   ! AllocateTile / AddGhostTiles / GetHeights leave it: tstate(ntiles) 0 off /
   ! 1 active / 2 ghost, b0v((NX+1)*(NY+1)), f13(13*NX*NY) fields 1..13 as
   ! given, sx/sy(NX*NY) ghost w slopes, btt(ntiles*(nx+1)*(ny+1)) per-tile
   ! vertex bt when use_btt /= 0 (else 0).
   subroutine setup_grid(c, tstate, has_src, b0v, f13, sx, sy, btt, use_btt, gridt)
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*), has_src(*)
      real(c_double), intent(in) :: b0v(*), f13(*), sx(*), sy(*), btt(*)
      integer, intent(in) :: use_btt
      real(c_double), intent(in) :: gridt

      integer :: nx, ny, ntx, nty, NXg, NYg, k, gi, gj, i, j, d, nact, ngh
      integer(c_int64_t) :: cidx, vidx, ncell
      real(wp) :: x, y

      call bind_closures(c)
      call fill_runset(c, rp)
      call InitialiseGrid(rp, grid)
      grid%t = gridt
      nx = c%nx_per_tile; ny = c%ny_per_tile; ntx = c%nx_tiles; nty = c%ny_tiles
      NXg = ntx * nx; NYg = nty * ny
      ncell = int(NXg, c_int64_t) * NYg

      nact = count(tstate(1:rp%nTiles) == 1)
      ngh = count(tstate(1:rp%nTiles) == 2)
      if (allocated(grid%ghostTiles%List)) deallocate(grid%ghostTiles%List)
      allocate(grid%activeTiles%List(nact), grid%ghostTiles%List(ngh))
      grid%activeTiles%size = 0
      grid%ghostTiles%size = 0

      do k = 1, rp%nTiles
         if (tstate(k) == 0) cycle
         associate (tl => grid%tileContainer(k))
            tl%TileOn = (tstate(k) == 1)
            tl%isGhostTile = (tstate(k) == 2)
            tl%containsSource = (has_src(k) /= 0)
            tl%neighbours(1) = k - 1; tl%neighbours(2) = k + 1
            tl%neighbours(3) = k + ntx; tl%neighbours(4) = k - ntx
            tl%cornertiles(1) = k - ntx - 1; tl%cornertiles(2) = k - ntx + 1
            tl%cornertiles(3) = k + ntx - 1; tl%cornertiles(4) = k + ntx + 1
            tl%West => tl%neighbours(1); tl%East => tl%neighbours(2)
            tl%North => tl%neighbours(3); tl%South => tl%neighbours(4)
            allocate(tl%u(13, nx, ny), tl%uLimX(13, nx, ny), tl%uLimY(13, nx, ny))
            allocate(tl%uPlusX(13, nx + 1, ny), tl%uMinusX(13, nx + 1, ny))
            allocate(tl%uPlusY(13, nx, ny + 1), tl%uMinusY(13, nx, ny + 1))
            allocate(tl%hXFlux(13, nx + 1, ny), tl%gXFlux(13, nx + 1, ny), tl%pXFlux(13, nx + 1, ny))
            allocate(tl%hYFlux(13, nx, ny + 1), tl%gYFlux(13, nx, ny + 1), tl%pYFlux(13, nx, ny + 1))
            allocate(tl%ddtExplicit(4, nx, ny), tl%ddtImplicit(4, nx, ny))
            allocate(tl%b0(nx + 1, ny + 1), tl%bt(nx + 1, ny + 1), tl%EminusD(nx, ny))
            allocate(tl%ddtExplicitBt(nx + 1, ny + 1))
            tl%ddtExplicitBt = 0
            allocate(tl%x(nx), tl%y(ny))
            tl%u = 0; tl%uLimX = 0; tl%uLimY = 0
            tl%uPlusX = 0; tl%uMinusX = 0; tl%uPlusY = 0; tl%uMinusY = 0
            tl%hXFlux = 0; tl%gXFlux = 0; tl%pXFlux = 0; tl%hYFlux = 0; tl%gYFlux = 0; tl%pYFlux = 0
            tl%ddtExplicit = 0; tl%ddtImplicit = 0; tl%bt = 0; tl%EminusD = 0
            gi = mod(k - 1, ntx) + 1; gj = (k - 1) / ntx + 1
            do i = 1, nx
               call GridToPhysical(rp, grid, gi, gj, i, 1, x, y)
               tl%x(i) = x
            end do
            do j = 1, ny
               call GridToPhysical(rp, grid, gi, gj, 1, j, x, y)
               tl%y(j) = y
            end do
            do j = 1, ny + 1
               do i = 1, nx + 1
                  vidx = int((gj - 1) * ny + j - 1, c_int64_t) * (NXg + 1) + (gi - 1) * nx + i
                  tl%b0(i, j) = b0v(vidx)
                  if (use_btt /= 0) tl%bt(i, j) = btt(int(k - 1, c_int64_t) * (nx + 1) * (ny + 1) &
                                                      + (j - 1) * (nx + 1) + i)
               end do
            end do
            ! face b0 as GetHeights leaves it (src/dem.f90:378-386)
            do j = 1, ny
               tl%uPlusX(10, :, j) = 0.5_wp * (tl%b0(:, j) + tl%b0(:, j + 1))
               tl%uMinusX(10, :, j) = 0.5_wp * (tl%b0(:, j) + tl%b0(:, j + 1))
            end do
            do i = 1, nx
               tl%uPlusY(10, i, :) = 0.5_wp * (tl%b0(i, :) + tl%b0(i + 1, :))
               tl%uMinusY(10, i, :) = 0.5_wp * (tl%b0(i, :) + tl%b0(i + 1, :))
            end do
            do j = 1, ny
               do i = 1, nx
                  cidx = int((gj - 1) * ny + j - 1, c_int64_t) * NXg + (gi - 1) * nx + i
                  do d = 1, 13
                     tl%u(d, i, j) = f13((d - 1) * ncell + cidx)
                  end do
                  if (tstate(k) == 2) then
                     tl%uLimX(1, i, j) = sx(cidx)
                     tl%uLimY(1, i, j) = sy(cidx)
                  end if
               end do
            end do
         end associate
         if (tstate(k) == 1) then
            grid%activeTiles%size = grid%activeTiles%size + 1
            grid%activeTiles%List(grid%activeTiles%size) = k
         else
            grid%ghostTiles%size = grid%ghostTiles%size + 1
            grid%ghostTiles%List(grid%ghostTiles%size) = k
         end if
      end do

   end subroutine setup_grid

   ! One hydraulic RHS of the reference on a state given in global layout
   ! (btt: per-tile vertex bt when use_btt /= 0, else bt = 0):
   !   tstate(ntiles)   0 off / 1 active / 2 ghost
   !   has_src(ntiles)  containsSource
   !   b0v((NX+1)*(NY+1)), f13(13*NX*NY) fields 1..13, sx/sy(NX*NY) ghost w slopes
   !   recompute_ghost: 1 -> rebuild the ghost w slopes with the reference's
   !   CalculateLimitedDerivsBoundary as ReconstructwAtEdges does; 0 -> use sx/sy.
   subroutine kref_rhs(c, tstate, has_src, b0v, f13, sx, sy, btt, use_btt, recompute_ghost, substep, t, &
                       gridt, ddte, ddti, advised) bind(C, name="kref_rhs")
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*), has_src(*)
      real(c_double), intent(in) :: b0v(*), f13(*), sx(*), sy(*), btt(*)
      integer(c_int32_t), value :: use_btt, recompute_ghost, substep
      real(c_double), value :: t, gridt
      real(c_double), intent(out) :: ddte(*), ddti(*), advised

      integer :: nx, ny, ntx, nty, NXg, NYg, k, gi, gj, i, j, d, nb
      integer(c_int64_t) :: cidx, ncell
      real(wp) :: adv

      call setup_grid(c, tstate, has_src, b0v, f13, sx, sy, btt, int(use_btt), gridt)
      nx = c%nx_per_tile; ny = c%ny_per_tile; ntx = c%nx_tiles; nty = c%ny_tiles
      NXg = ntx * nx; NYg = nty * ny
      ncell = int(NXg, c_int64_t) * NYg

      do k = 1, rp%nTiles
         if (tstate(k) == 0) cycle
         call ComputeCellCentredTopographicData(rp, grid, grid%tileContainer, k)
      end do

      ! Ghost w reconstructions at edges facing active tiles
      ! (assignments of ReconstructwAtEdges, src/UpdateTiles.f90:669-750).
      do k = 1, rp%nTiles
         if (tstate(k) /= 2) cycle
         associate (tl => grid%tileContainer(k))
            nb = tl%West
            if (nb >= 1 .and. nb <= rp%nTiles) then
               if (tstate(nb) == 1) then
                  if (recompute_ghost /= 0) then
                     do j = 1, ny
                        call CalculateLimitedDerivsBoundary(rp, grid, 1, 1, j, grid%tileContainer, k, 'W')
                     end do
                  end if
                  tl%uPlusX(1, 1, :) = tl%u(1, 1, :) - tl%uLimX(1, 1, :) * 0.5_wp * c%dx
                  tl%uMinusX(1, 2, :) = tl%u(1, 1, :) + tl%uLimX(1, 1, :) * 0.5_wp * c%dx
               end if
            end if
            nb = tl%East
            if (nb >= 1 .and. nb <= rp%nTiles) then
               if (tstate(nb) == 1) then
                  if (recompute_ghost /= 0) then
                     do j = 1, ny
                        call CalculateLimitedDerivsBoundary(rp, grid, 1, nx, j, grid%tileContainer, k, 'E')
                     end do
                  end if
                  tl%uPlusX(1, nx, :) = tl%u(1, nx, :) - tl%uLimX(1, nx, :) * 0.5_wp * c%dx
                  tl%uMinusX(1, nx + 1, :) = tl%u(1, nx, :) + tl%uLimX(1, nx, :) * 0.5_wp * c%dx
               end if
            end if
            nb = tl%South
            if (nb >= 1 .and. nb <= rp%nTiles) then
               if (tstate(nb) == 1) then
                  if (recompute_ghost /= 0) then
                     do i = 1, nx
                        call CalculateLimitedDerivsBoundary(rp, grid, 1, i, 1, grid%tileContainer, k, 'S')
                     end do
                  end if
                  tl%uPlusY(1, :, 1) = tl%u(1, :, 1) - tl%uLimY(1, :, 1) * 0.5_wp * c%dy
                  tl%uMinusY(1, :, 2) = tl%u(1, :, 1) + tl%uLimY(1, :, 1) * 0.5_wp * c%dy
               end if
            end if
            nb = tl%North
            if (nb >= 1 .and. nb <= rp%nTiles) then
               if (tstate(nb) == 1) then
                  if (recompute_ghost /= 0) then
                     do i = 1, nx
                        call CalculateLimitedDerivsBoundary(rp, grid, 1, i, ny, grid%tileContainer, k, 'N')
                     end do
                  end if
                  tl%uPlusY(1, :, ny) = tl%u(1, :, ny) - tl%uLimY(1, :, ny) * 0.5_wp * c%dy
                  tl%uMinusY(1, :, ny + 1) = tl%u(1, :, ny) + tl%uLimY(1, :, ny) * 0.5_wp * c%dy
               end if
            end if
         end associate
      end do

      do k = 1, rp%nTiles
         if (tstate(k) /= 1) cycle
         call ComputeInterfacialTopographicData(rp, grid, grid%tileContainer, k)
      end do

      call CalculateHydraulicRHS(rp, grid, grid%tileContainer, t, int(substep), adv)
      advised = adv

      ddte(1:4 * ncell) = 0.0_wp
      ddti(1:4 * ncell) = 0.0_wp
      do k = 1, rp%nTiles
         if (tstate(k) /= 1) cycle
         gi = mod(k - 1, ntx) + 1; gj = (k - 1) / ntx + 1
         do j = 1, ny
            do i = 1, nx
               cidx = int((gj - 1) * ny + j - 1, c_int64_t) * NXg + (gi - 1) * nx + i
               do d = 1, 4
                  ddte((d - 1) * ncell + cidx) = grid%tileContainer(k)%ddtExplicit(d, i, j)
                  ddti((d - 1) * ncell + cidx) = grid%tileContainer(k)%ddtImplicit(d, i, j)
               end do
            end do
         end do
      end do
   end subroutine kref_rhs


   ! Scatter a container's fields 1..13 into global layout and its vertex bt
   ! per tile.
   subroutine gather_container(c, tstate, tiles, f13, btt)
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*)
      type(TileType), intent(in) :: tiles(:)
      real(c_double), intent(out) :: f13(*), btt(*)
      integer :: nx, ny, ntx, NXg, NYg, k, gi, gj, i, j, d
      integer(c_int64_t) :: cidx, ncell, nv
      nx = c%nx_per_tile; ny = c%ny_per_tile; ntx = c%nx_tiles
      NXg = ntx * nx; NYg = c%ny_tiles * ny
      ncell = int(NXg, c_int64_t) * NYg
      nv = int(nx + 1, c_int64_t) * (ny + 1)
      f13(1:13 * ncell) = 0.0_wp
      btt(1:nv * rp%nTiles) = 0.0_wp
      do k = 1, rp%nTiles
         if (tstate(k) == 0) cycle
         gi = mod(k - 1, ntx) + 1; gj = (k - 1) / ntx + 1
         do j = 1, ny
            do i = 1, nx
               cidx = int((gj - 1) * ny + j - 1, c_int64_t) * NXg + (gi - 1) * nx + i
               do d = 1, 13
                  f13((d - 1) * ncell + cidx) = tiles(k)%u(d, i, j)
               end do
            end do
         end do
         do j = 1, ny + 1
            do i = 1, nx + 1
               btt((k - 1) * nv + (j - 1) * (nx + 1) + i) = tiles(k)%bt(i, j)
            end do
         end do
      end do
   end subroutine gather_container

   ! Load fields 1..13 (global layout) and per-tile vertex bt into a container.
   subroutine scatter_container(c, tstate, tiles, f13, btt)
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*)
      type(TileType), intent(inout) :: tiles(:)
      real(c_double), intent(in) :: f13(*), btt(*)
      integer :: nx, ny, ntx, NXg, NYg, k, gi, gj, i, j, d
      integer(c_int64_t) :: cidx, ncell, nv
      nx = c%nx_per_tile; ny = c%ny_per_tile; ntx = c%nx_tiles
      NXg = ntx * nx; NYg = c%ny_tiles * ny
      ncell = int(NXg, c_int64_t) * NYg
      nv = int(nx + 1, c_int64_t) * (ny + 1)
      do k = 1, rp%nTiles
         if (tstate(k) == 0) cycle
         gi = mod(k - 1, ntx) + 1; gj = (k - 1) / ntx + 1
         do j = 1, ny
            do i = 1, nx
               cidx = int((gj - 1) * ny + j - 1, c_int64_t) * NXg + (gi - 1) * nx + i
               do d = 1, 13
                  tiles(k)%u(d, i, j) = f13((d - 1) * ncell + cidx)
               end do
            end do
         end do
         do j = 1, ny + 1
            do i = 1, nx + 1
               tiles(k)%bt(i, j) = btt((k - 1) * nv + (j - 1) * (nx + 1) + i)
            end do
         end do
      end do
   end subroutine scatter_container

   ! CalculateMorphodynamicRHS of the reference on tileContainer built from
   ! fields 1..13 and per-tile vertex bt as given: returns EminusD (NX*NY,
   ! active tiles) and ddtExplicitBt per tile (ntiles*(nx+1)*(ny+1)).
   subroutine kref_morpho_rhs(c, tstate, b0v, f13, btt, emd, dtbt) bind(C, name="kref_morpho_rhs")
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*)
      real(c_double), intent(in) :: b0v(*), f13(*), btt(*)
      real(c_double), intent(out) :: emd(*), dtbt(*)
      integer(c_int8_t), allocatable :: nosrc(:)
      real(c_double), allocatable :: zero(:)
      integer :: nx, ny, ntx, NXg, k, gi, gj, i, j
      integer(c_int64_t) :: cidx, ncell, nv

      nx = c%nx_per_tile; ny = c%ny_per_tile; ntx = c%nx_tiles
      NXg = ntx * nx
      ncell = int(NXg, c_int64_t) * (c%ny_tiles * ny)
      nv = int(nx + 1, c_int64_t) * (ny + 1)
      allocate(nosrc(c%nx_tiles * c%ny_tiles), zero(ncell))
      nosrc = 0; zero = 0.0_wp
      call setup_grid(c, tstate, nosrc, b0v, f13, zero, zero, btt, 1, 0.0_wp)
      call CalculateMorphodynamicRHS(rp, grid, grid%tileContainer)
      emd(1:ncell) = 0.0_wp
      dtbt(1:nv * rp%nTiles) = 0.0_wp
      do k = 1, rp%nTiles
         if (tstate(k) /= 1) cycle
         gi = mod(k - 1, ntx) + 1; gj = (k - 1) / ntx + 1
         do j = 1, ny
            do i = 1, nx
               cidx = int((gj - 1) * ny + j - 1, c_int64_t) * NXg + (gi - 1) * nx + i
               emd(cidx) = grid%tileContainer(k)%EminusD(i, j)
            end do
         end do
         do j = 1, ny + 1
            do i = 1, nx + 1
               dtbt((k - 1) * nv + (j - 1) * (nx + 1) + i) = grid%tileContainer(k)%ddtExplicitBt(i, j)
            end do
         end do
      end do
   end subroutine kref_morpho_rhs

   ! AddToRedistList for the candidates in the given order, then
   ! RedistributeGrid, with intermed0 = (f0, bt0) and intermed3 = (f3, bt3)
   ! (fields 1..13 global, vertex bt per tile).  Returns intermed3 afterwards
   ! and the refine flag.
   subroutine kref_redistribute(c, tstate, b0v, f0, bt0, f3, bt3, ncand, ck, ci, cj, cex, &
                                fout, btout, refine) bind(C, name="kref_redistribute")
      type(kst_config), intent(in) :: c
      integer(c_int8_t), intent(in) :: tstate(*)
      real(c_double), intent(in) :: b0v(*), f0(*), bt0(*), f3(*), bt3(*), cex(*)
      integer(c_int32_t), value :: ncand
      integer(c_int32_t), intent(in) :: ck(*), ci(*), cj(*)
      real(c_double), intent(out) :: fout(*), btout(*)
      integer(c_int32_t), intent(out) :: refine
      integer(c_int8_t), allocatable :: nosrc(:)
      real(c_double), allocatable :: zero(:)
      type(RedistList) :: list
      logical :: ref
      integer :: k, q
      integer(c_int64_t) :: ncell

      ncell = int(c%nx_tiles * c%nx_per_tile, c_int64_t) * (c%ny_tiles * c%ny_per_tile)
      allocate(nosrc(c%nx_tiles * c%ny_tiles), zero(ncell))
      nosrc = 0; zero = 0.0_wp
      call setup_grid(c, tstate, nosrc, b0v, f0, zero, zero, bt0, 1, 0.0_wp)
      do k = 1, rp%nTiles
         if (tstate(k) == 0) cycle
         grid%intermed0(k) = grid%tileContainer(k)
         grid%intermed3(k) = grid%tileContainer(k)
      end do
      call scatter_container(c, tstate, grid%intermed3, f3, bt3)
      do q = 1, ncand
         call AddToRedistList(list, int(ck(q)), int(ci(q)), int(cj(q)), cex(q))
      end do
      ref = .false.
      if (ncand > 0) call RedistributeGrid(rp, grid, list, ref)
      refine = merge(1, 0, ref)
      call gather_container(c, tstate, grid%intermed3, fout, btout)
   end subroutine kref_redistribute

   ! One of the reference's TopogFunc subroutines (src/TopogFuncs.f90) on the
   ! vertex coordinates x(nxv), y(nyv) with its 'Topog params': b0(nxv, nyv).
   ! which: 1 x2slopes, 2 usgs, 3 flume, 4 channel_powerlaw, 5 channel_trapezium,
   ! 6 xbislope, 7 xtrislope, 8 xhump, 9 xtanh, 10 xyparab.
   subroutine kref_topog(which, params, nparams, x, nxv, y, nyv, b0) bind(C, name="kref_topog")
      integer(c_int), value :: which, nparams, nxv, nyv
      real(c_double), intent(in) :: params(nparams), x(nxv), y(nyv)
      real(c_double), intent(out) :: b0(nxv, nyv)
      ! the module's RunSet (a local one would need the utm module's procedures)
      if (allocated(rp%TopogFuncParams)) deallocate (rp%TopogFuncParams)
      allocate (rp%TopogFuncParams(nparams))
      rp%TopogFuncParams = params
      select case (which)
      case (1); call x2slopes(rp, x, y, b0)
      case (2); call usgs(rp, x, y, b0)
      case (3); call flume(rp, x, y, b0)
      case (4); call channel_powerlaw(rp, x, y, b0)
      case (5); call channel_trapezium(rp, x, y, b0)
      case (6); call xbislope(rp, x, y, b0)
      case (7); call xtrislope(rp, x, y, b0)
      case (8); call xhump(rp, x, y, b0)
      case (9); call xtanh(rp, x, y, b0)
      case (10); call xyparab(rp, x, y, b0)
      case default; b0 = 0.0_wp
      end select
   end subroutine kref_topog

   ! Bicubic (src/Interp2d.f90:214-282) of real data z(nzx, nzy) at the
   ! npts points (x, y), 1-based data coordinates.
   subroutine kref_bicubic(z, nzx, nzy, x, y, npts, out) bind(C, name="kref_bicubic")
      integer(c_int), value :: nzx, nzy, npts
      real(c_double), intent(in) :: z(nzx, nzy), x(npts), y(npts)
      real(c_double), intent(out) :: out(npts)
      integer :: k
      do k = 1, npts
         out(k) = Bicubic(x(k), y(k), z)
      end do
   end subroutine kref_bicubic

end module kestrel_ref_driver
