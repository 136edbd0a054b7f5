! Test driver of the Fortran binding (include/fdfk.f90): reads one case from
! a raw file written by tests/test_gpu_fortran.py, runs it resident on the
! GPU through the handle API (create, FK background, run, seismograms,
! destroy) and writes the seismograms to a raw file.
!
! input (stream, native endian): int32 nx, nz, npml, norder, nt, nstp, nrcv,
! nl_left, nl_right, src_is_p; real32 dx, dz, x0, z0, ds, f0; real64 dt, rayp;
! then c(3*NX*NZ), rx(nrcv), rz(nrcv) real32, fk_left(4*nl_left),
! fk_right(4*nl_right) real32, bazs(NX) real64.
! output: seisx then seisz, (nt/nstp + 1) samples per receiver, real32.
program handle_api_run
  use, intrinsic :: iso_c_binding
  use fdfk
  implicit none
  integer(c_int32_t) :: nx, nz, npml, norder, nt, nstp, nrcv, nll, nlr, srcp
  real(c_float) :: dx, dz, x0, z0, ds, f0
  real(c_double) :: dt, rayp
  real(c_float), allocatable, target :: c(:), rx(:), rz(:), fkl(:), fkr(:), sx(:), sz(:)
  real(c_double), allocatable, target :: bazs(:)
  type(fdfk_job) :: job
  type(fdfk_fk) :: fk
  type(c_ptr) :: h
  character(len=1024) :: fin, fout
  integer :: u, nxp, nzp, nsamp, ier

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nx, nz, npml, norder, nt, nstp, nrcv, nll, nlr, srcp
  read(u) dx, dz, x0, z0, ds, f0
  read(u) dt, rayp
  nxp = nx + 2 * npml
  nzp = nz + npml
  allocate(c(3 * nxp * nzp), rx(nrcv), rz(nrcv), fkl(4 * nll), fkr(4 * nlr), bazs(nxp))
  read(u) c, rx, rz, fkl, fkr, bazs
  close(u)

  job = fdfk_job(nx, nz, npml, norder, 0, 2, x0, z0, dx, dz, dt, 1.0d-3, 0.0, nt, nstp, nrcv, &
                 c_loc(rx), c_loc(rz), c_loc(c), FDFK_FP32, FDFK_PSV)
  ier = fdfk_create(job, h)
  if (ier /= FDFK_OK) call fail('fdfk_create')
  fk = fdfk_fk(nll, nlr, c_loc(fkl), c_loc(fkr), srcp, ds, f0, rayp, c_loc(bazs), 0)
  ier = fdfk_compute_background(h, fk)
  if (ier /= FDFK_OK) call fail('fdfk_compute_background')
  ! the loop of simulate_wavefield, resident: all iterations, recording on the device
  ier = fdfk_run(h, nt + 1)
  if (ier /= FDFK_OK) call fail('fdfk_run')
  if (fdfk_iteration(h) /= nt + 1) call fail('iteration count')
  nsamp = nt / nstp + 1
  allocate(sx(nsamp * nrcv), sz(nsamp * nrcv))
  ier = fdfk_seismograms(h, sx, sz)
  if (ier /= FDFK_OK) call fail('fdfk_seismograms')
  call fdfk_destroy(h)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) sx, sz
  close(u)

contains

  subroutine fail(what)
    character(len=*), intent(in) :: what
    write(0, '(a)') 'handle_api_run: ' // what // ' failed: ' // fdfk_error_message()
    stop 1
  end subroutine

end program handle_api_run
