! ref_run.f90 -- runs the reference program's own Run and dumps its state.
!
! TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile (target refrun) with
! every Fortran module of the reference compiled where it lies, in place of
! its main program (src/main.f90:41-133), whose set-up sequence it repeats:
! ReadInputFile, SetImmutableRunSettings, CreateOutDir, InitialiseGrid,
! LoadDEM, LoadInitialCondition, LoadSourceConditions, Run.
!
!   ref_run <deck> <outdir> <dump-prefix> [<b0-file>]
!   ref_run <deck> <outdir> -params
!   ref_run <deck> <outdir> none [<b0-file>]     (no dumps: for timing)
!
! The second form prints, after ReadInputFile and SetImmutableRunSettings,
! the RunSet scalars the time step reads (name = value, full precision) and
! stops: the deck reader's configuration is checked against them.
!
! Run (src/TimeStepper.f90:73-113) integrates one output interval per
! IntegrateTo call.  The driver calls it once per interval k = 1..Nout with
! FirstOut = k-1 and Nout = k, so each call integrates exactly the interval
! [tstart + (k-1) DeltaT, tstart + k DeltaT] that the reference's own loop
! would, then writes <dump-prefix>_<k>.bin (unformatted stream, native
! little-endian) holding grid%t, grid%dt, the active and ghost lists and,
! per listed tile, its u(1:13,:,:), b0, bt and maxima.  Restart is set after
! the first call only to skip Run's start-of-run output.  A deck with
! Restart = on resumes, through LoadInitialCondition, from the output folder's
! RunInfo.txt and last file; the intervals then run from that file's number
! + 1 (<dump-prefix>_0.bin holds the restarted state).
!
! <b0-file> (optional) replaces the deck's analytic topography by vertex
! elevations sampled on the whole grid: int32 nxv, nyv; float64 x0, y0, dx,
! dy; then float64 b0(nxv, nyv), x fastest.  GetHeights
! (src/dem.f90:360-393) then reads b0 at the tile's vertex coordinates
! through the TopogFunc pointer (src/TopogFuncs.f90:82-92).
module ref_run_topog
   use set_precision_module, only: wp
   use runsettings_module, only: RunSet
   implicit none
   real(kind=wp), allocatable, save :: b0g(:, :)
   real(kind=wp), save :: gx0, gy0, gdx, gdy
contains
   pure subroutine sampled_topog(RunParams, x, y, b0)
      type(RunSet), intent(in) :: RunParams
      real(kind=wp), dimension(:), intent(in) :: x
      real(kind=wp), dimension(:), intent(in) :: y
      real(kind=wp), dimension(:,:), intent(out) :: b0(size(x), size(y))
      integer :: i, j, ii, jj
      do j = 1, size(y)
         jj = nint((y(j) - gy0) / gdy) + 1
         do i = 1, size(x)
            ii = nint((x(i) - gx0) / gdx) + 1
            b0(i, j) = b0g(ii, jj)
         end do
      end do
   end subroutine sampled_topog
end module ref_run_topog

program ref_run
   use, intrinsic :: iso_fortran_env, only: int32
   use set_precision_module, only: wp
   use grid_module, only: GridType, InitialiseGrid
   use runsettings_module, only: RunSet, SetImmutableRunSettings
   use read_input_file_module, only: ReadInputFile
   use set_sources_module, only: LoadSourceConditions
   use restart_module, only: LoadInitialCondition
   use dem_module, only: LoadDEM
   use timestepper_module, only: Run
   use output_module, only: CreateOutDir
   use varstring_module, only: varString
   use version_module, only: GetVersion
   use topog_funcs_module, only: TopogFunc
   use ref_run_topog
   implicit none

   type(RunSet) :: RunParams
   type(GridType), target :: grid
   character(len=:), allocatable :: version
   character(len=4096) :: outdir, prefix, b0file, fname
   integer :: nout, k, nargs, first
   integer(int32) :: nxv, nyv

   version = GetVersion()
   RunParams%version = varString(trim(version))
   RunParams%Vars%w = 1
   RunParams%Vars%rhoHnu = 2
   RunParams%Vars%rhoHnv = 3
   RunParams%Vars%Hnpsi = 4
   RunParams%Vars%Hn = 5
   RunParams%Vars%u = 6
   RunParams%Vars%v = 7
   RunParams%Vars%psi = 8
   RunParams%Vars%rho = 9
   RunParams%Vars%b0 = 10
   RunParams%Vars%bt = 11
   RunParams%Vars%dbdx = 12
   RunParams%Vars%dbdy = 13
   RunParams%nDimensions = 13

   nargs = command_argument_count()
   if (nargs < 3) then
      write (*, *) "usage: ref_run <deck> <outdir> <dump-prefix> [<b0-file>]"
      stop 2
   end if
   call get_command_argument(2, outdir)
   call get_command_argument(3, prefix)

   call ReadInputFile(RunParams)
   call SetImmutableRunSettings(RunParams)
   RunParams%out_path = varString(trim(outdir) // '/')
   if (trim(prefix) == '-params') then
      call print_params()
      stop
   end if

   if (nargs >= 4) then
      call get_command_argument(4, b0file)
      open (61, file=trim(b0file), access='stream', form='unformatted', status='old')
      read (61) nxv, nyv
      read (61) gx0, gy0, gdx, gdy
      allocate (b0g(nxv, nyv))
      read (61) b0g
      close (61)
      TopogFunc => sampled_topog
   end if

   call CreateOutDir(RunParams)
   call InitialiseGrid(RunParams, grid)
   call LoadDEM(RunParams, grid)
   call LoadInitialCondition(RunParams, grid)
   call LoadSourceConditions(RunParams, grid)
   if (trim(prefix) /= 'none') call dump(trim(prefix) // '_0.bin')

   nout = RunParams%Nout
   first = RunParams%FirstOut   ! 0, or on a restart the last file's number (src/Restart.f90:632)
   do k = first + 1, nout
      RunParams%FirstOut = k - 1
      RunParams%Nout = k
      call Run(RunParams, grid)
      RunParams%Restart = .true.
      write (fname, '(a,a,i0,a)') trim(prefix), '_', k, '.bin'
      if (trim(prefix) /= 'none') call dump(trim(fname))
      write (*, '(a,i0,a,es25.17)') 'ref_run: interval ', k, ' t = ', grid%t
   end do

contains

   subroutine pr(name, v)
      character(*), intent(in) :: name
      real(kind=wp), intent(in) :: v
      write (*, '(a,a,a,es25.17e3)') 'param ', name, ' = ', v
   end subroutine pr

   subroutine ps(name, v)
      character(*), intent(in) :: name, v
      write (*, '(a,a,a,a)') 'param ', name, ' = ', trim(v)
   end subroutine ps

   subroutine print_params()
      write (*, '(a,4i8)') 'param tiles = ', RunParams%nXtiles, RunParams%nYtiles, &
         RunParams%nXpertile, RunParams%nYpertile
      call pr('deltaX', RunParams%deltaX); call pr('deltaY', RunParams%deltaY)
      call pr('xSize', RunParams%xSize); call pr('ySize', RunParams%ySize)
      call ps('bcs', RunParams%bcs%s)
      if (RunParams%bcs%s == 'dirichlet') then
         call pr('bcsHnval', RunParams%bcsHnval); call pr('bcsuval', RunParams%bcsuval)
         call pr('bcsvval', RunParams%bcsvval); call pr('bcspsival', RunParams%bcspsival)
      end if
      call pr('SpongeStrength', RunParams%SpongeStrength)
      call ps('limiter', RunParams%limiter%s)
      call ps('DragChoice', RunParams%DragChoice%s)
      call ps('ErosionChoice', RunParams%ErosionChoice%s)
      call ps('DepositionChoice', RunParams%DepositionChoice%s)
      call ps('ErosionTransition', RunParams%ErosionTransition%s)
      call ps('MorphoDamp', RunParams%MorphoDamp%s)
      call ps('fswitch', RunParams%fswitch%s)
      write (*, '(a,l2)') 'param geometric_factors = ', RunParams%geometric_factors
      write (*, '(a,l2)') 'param MorphodynamicsOn = ', RunParams%MorphodynamicsOn
      write (*, '(a,i8)') 'param TileBuffer = ', RunParams%TileBuffer
      call pr('g', RunParams%g); call pr('rhow', RunParams%rhow); call pr('rhos', RunParams%rhos)
      call pr('gred', RunParams%gred)
      call pr('ChezyCo', RunParams%ChezyCo); call pr('ManningCo', RunParams%ManningCo)
      call pr('CoulombCo', RunParams%CoulombCo)
      call pr('PouliquenMinSlope', RunParams%PouliquenMinSlope)
      call pr('PouliquenMaxSlope', RunParams%PouliquenMaxSlope)
      call pr('PouliquenIntermediateSlope', RunParams%PouliquenIntermediateSlope)
      call pr('PouliquenBeta', RunParams%PouliquenBeta)
      call pr('Edwards2019betastar', RunParams%Edwards2019betastar)
      call pr('Edwards2019kappa', RunParams%Edwards2019kappa)
      call pr('Edwards2019Gamma', RunParams%Edwards2019Gamma)
      call pr('VoellmySwitchRate', RunParams%VoellmySwitchRate)
      call pr('VoellmySwitchValue', RunParams%VoellmySwitchValue)
      call pr('EroRate', RunParams%EroRate); call pr('EroRateGranular', RunParams%EroRateGranular)
      call pr('CriticalShields', RunParams%CriticalShields); call pr('EroDepth', RunParams%EroDepth)
      call pr('EroCriticalHeight', RunParams%EroCriticalHeight)
      call pr('BedPorosity', RunParams%BedPorosity); call pr('maxPack', RunParams%maxPack)
      call pr('SolidDiameter', RunParams%SolidDiameter); call pr('Rep', RunParams%Rep)
      call pr('ws0', RunParams%ws0); call pr('nsettling', RunParams%nsettling)
      call pr('EddyViscosity', RunParams%EddyViscosity)
      call pr('heightThreshold', RunParams%heightThreshold); call pr('cfl', RunParams%cfl)
      call pr('diffusiveTimeScale', RunParams%diffusiveTimeScale); call pr('maxdt', RunParams%maxdt)
      call pr('tstart', RunParams%tstart); call pr('tend', RunParams%tend)
   end subroutine print_params

   subroutine dump(path)
      character(*), intent(in) :: path
      integer :: q, t, na, ng, nt, nx, ny
      integer(int32), allocatable :: st(:)
      integer(int32) :: hdr(8)
      nt = grid%nTiles
      nx = RunParams%nXpertile
      ny = RunParams%nYpertile
      na = grid%activeTiles%size
      ng = grid%ghostTiles%size
      allocate (st(nt))
      do t = 1, nt
         st(t) = 0
         if (grid%tileContainer(t)%TileOn) st(t) = 1
         if (grid%tileContainer(t)%isGhostTile) st(t) = 2
      end do
      hdr = [int(1, int32), int(nt, int32), int(nx, int32), int(ny, int32), &
             int(na, int32), int(ng, int32), int(grid%nXtiles, int32), int(grid%nYtiles, int32)]
      open (62, file=path, access='stream', form='unformatted', status='replace')
      write (62) hdr
      write (62) grid%t, grid%dt
      if (na > 0) write (62) int(grid%activeTiles%List(1:na), int32)
      if (ng > 0) write (62) int(grid%ghostTiles%List(1:ng), int32)
      write (62) st
      do q = 1, na
         t = grid%activeTiles%List(q)
         write (62) grid%tileContainer(t)%u(1:13, :, :)
         write (62) grid%tileContainer(t)%b0(:, :)
         write (62) grid%tileContainer(t)%bt(:, :)
         write (62) grid%tileContainer(t)%Hnmax, grid%tileContainer(t)%umax, &
            grid%tileContainer(t)%emax, grid%tileContainer(t)%dmax, &
            grid%tileContainer(t)%psimax, grid%tileContainer(t)%tfirst
      end do
      do q = 1, ng
         t = grid%ghostTiles%List(q)
         write (62) grid%tileContainer(t)%u(1:13, :, :)
         write (62) grid%tileContainer(t)%bt(:, :)
      end do
      close (62)
   end subroutine dump

end program ref_run
