! A Fortran host that calls solver_acoustic_SS through an implicit interface,
! exactly as calculate_misfit_gradient.f90:259-261 and :315-317 do.  Linked
! against the reference's compiled solver or against libfwi_hip.so, it must
! write identical files.
!
! usage: dropin_host <input.bin> <output.bin>
!   input  (stream): ny nx nt ns nr ss (int32), d dt (real32), nxs nys nxr nyr
!                    (int32), S(nt,ns) ctmp(ny,nx) (real32)
!   output (stream): Rec(nt,nr) of the forward run, kernel(ny,nx) of the
!                    adjoint run with Sadj = -Rec reversed in time,
!                    Fields(ny,nx*(nt/ss+1))
program dropin_host
  implicit none
  integer :: ny, nx, nt, ns, nr, ss, nts, j, k
  real :: d, dt
  integer, allocatable :: nxs(:), nys(:), nxr(:), nyr(:)
  real, allocatable :: S(:,:), Sadj(:,:), c(:,:), kernel(:,:), Fields(:,:), Rec(:,:)
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ny, nx, nt, ns, nr, ss
  read(10) d, dt
  allocate(nxs(ns), nys(ns), nxr(nr), nyr(nr), S(nt,ns), c(ny,nx))
  read(10) nxs, nys, nxr, nyr
  read(10) S, c
  close(10)

  nts = nt / ss
  allocate(Fields(ny, nx*(nts+1)), kernel(ny,nx), Rec(nt,nr), Sadj(nt,nr))
  Fields = 0.
  kernel = 0.
  Rec = 0.
  Sadj = 0.

  call solver_acoustic_SS(0, 1, ns, nr, nxs, nys, nxr, nyr, S, Sadj, d, dt, ny, nx, c, ss, nts, nt, &
                          kernel, Fields, Rec)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) Rec
  do j = 1, nr
     do k = 1, nt
        Sadj(k,j) = -Rec(nt-k+1,j)
     enddo
  enddo
  call solver_acoustic_SS(0, 2, ns, nr, nxs, nys, nxr, nyr, S, Sadj, d, dt, ny, nx, c, ss, nts, nt, &
                          kernel, Fields, Rec)
  write(11) kernel
  write(11) Fields
  close(11)
end program dropin_host
