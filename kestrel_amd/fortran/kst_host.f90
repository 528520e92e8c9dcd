! kst_host.f90 -- a Fortran host driving the MI355X time step through the
! iso_c_binding module kestrel_step_binding, the way a Kestrel build whose
! IntegrateTo (src/TimeStepper.f90:116-277) calls the library would: create a
! handle, hand over the vertex topography and the initial state once, then per
! output time call kst_integrate_to and download the state (what
! OutputSolutionData would then print).  A status other than 0 stops the run
! with the library's message and exit code 1, as FatalErrorMessage does
! (src/Messages.f90:73-89).
!
!   kst_host IN OUT
!
! IN (stream, native byte order): the kst_config image (c_sizeof bytes, no
! flux sources), int64 nout, nout FP64 output times, FP64 t0, then FP64 b0 at
! the (NY+1) x (NX+1) vertices, FP64 w, rhoHnu, rhoHnv, Hnpsi at the NY x NX
! cells (row-major, x fastest) and int8 tile states.  OUT: per output time,
! the kst_stats image, t, and FP64 w, rhoHnu, rhoHnv, Hnpsi, Hn.
program kst_host
   use, intrinsic :: iso_c_binding
   use kestrel_step_binding
   implicit none

   type(kst_config), target :: cfg
   type(kst_stats), target :: stats
   type(c_ptr) :: h
   integer(c_int8_t), allocatable :: raw(:), tstate(:)
   real(c_double), allocatable :: b0v(:), w(:), rhu(:), rhv(:), hnpsi(:), hn(:), touts(:)
   real(c_double) :: t0, t
   integer(c_int64_t) :: nout
   integer(c_int) :: rc
   integer :: uin, uout, k, ncell, nvert, ntiles
   character(len=1024) :: fin, fout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   allocate(raw(c_sizeof(cfg)))
   read(uin) raw
   cfg = transfer(raw, cfg)
   cfg%sources = c_null_ptr
   read(uin) nout
   allocate(touts(nout))
   read(uin) touts
   read(uin) t0
   ncell = cfg%nx_tiles * cfg%nx_per_tile * cfg%ny_tiles * cfg%ny_per_tile
   nvert = (cfg%nx_tiles * cfg%nx_per_tile + 1) * (cfg%ny_tiles * cfg%ny_per_tile + 1)
   ntiles = cfg%nx_tiles * cfg%ny_tiles
   allocate(b0v(nvert), w(ncell), rhu(ncell), rhv(ncell), hnpsi(ncell), hn(ncell), tstate(ntiles))
   read(uin) b0v, w, rhu, rhv, hnpsi, tstate
   close(uin)

   rc = kst_create(cfg, h)
   if (rc /= 0) call fatal(rc, 'kst_create: configuration or device')
   rc = kst_set_topography(h, b0v)
   if (rc /= 0) call fatal(rc)
   rc = kst_set_state(h, w, rhu, rhv, hnpsi, c_null_ptr, tstate, t0, c_null_ptr)
   if (rc /= 0) call fatal(rc)

   open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   do k = 1, int(nout)
      rc = kst_integrate_to(h, touts(k), stats)
      if (rc /= 0) call fatal(rc)
      rc = kst_get_state(h, w, rhu, rhv, hnpsi, hn, c_null_ptr, tstate, t)
      if (rc /= 0) call fatal(rc)
      write(uout) transfer(stats, raw(1:c_sizeof(stats))), t, w, rhu, rhv, hnpsi, hn
   end do
   close(uout)
   call kst_destroy(h)

contains

   subroutine fatal(code, what)
      integer(c_int), intent(in) :: code
      character(len=*), intent(in), optional :: what
      character(kind=c_char) :: buf(512)
      character(len=512) :: msg
      integer :: i, ierr
      msg = ''
      if (present(what)) then
         msg = what
      else
         ierr = kst_last_error(h, buf, 512_c_int)
         do i = 1, 512
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
         end do
      end if
      write(0, '(a, i0, a, a)') 'kst_host: status ', code, ': ', trim(msg)
      stop 1
   end subroutine fatal

end program kst_host
