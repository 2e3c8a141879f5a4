! Fortran host test: drives libmopachip through the mopac_hip module with the
! reference routines' calling surface, on a system written by the pytest side
! (stream file, see tests/test_gpu_fortran.py), and writes every result for
! comparison with the CPU oracle.
!
! usage: test_mopac_hip <system.bin> <results.bin>
program test_mopac_hip
  use iso_c_binding
  use mopac_hip
  implicit none
  integer :: numat, norbs, nclose, mpack, i, u
  integer(c_long_long) :: n2elec
  double precision :: enuclr, atheat, e
  integer, allocatable :: nfirst(:), nlast(:)
  double precision, allocatable, target :: pdiag(:)
  double precision, allocatable :: h(:), w(:), ptest(:), f(:), pa(:), c(:, :), eig(:), &
    fcopy(:), p(:), pscf(:), fscf(:), cscf(:, :), eigscf(:)
  character(len=4096) :: fin, fout
  type(mhip_scf_opts), target :: opts
  type(mhip_scf_result) :: res

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) numat, norbs, nclose
  read (u) n2elec
  read (u) enuclr, atheat
  mpack = norbs * (norbs + 1) / 2
  allocate (nfirst(numat), nlast(numat), h(mpack), w(n2elec), pdiag(norbs), ptest(mpack))
  read (u) nfirst, nlast, h, w, pdiag, ptest
  close (u)

  call hip_init(0)
  call hip_upload(numat, nfirst, nlast, norbs, n2elec, h, w)

  ! fock2 with a test density (ptot, alpha = ptot / 2), mode 2 (with fock1)
  allocate (f(mpack), pa(mpack))
  f = h
  pa = 0.5d0 * ptest
  call hip_fock2(f, ptest, pa, w, w, w, numat, nfirst, nlast, 2)
  e = hip_helect(norbs, pa, h, f)

  ! eigenvectors_LAPACK surface: xmat gets +i*1e-10 on its diagonal in place
  allocate (c(norbs, norbs), eig(norbs), fcopy(mpack), p(mpack))
  fcopy = f
  call hip_eigenvectors(c, fcopy, eig, norbs)
  call hip_densit(c, norbs, norbs, nclose, 2.d0, 0, 0.d0, p, 0)

  ! whole SCF (1SCF, SCFCRT=1e-6)
  allocate (pscf(mpack), fscf(mpack), cscf(norbs, norbs), eigscf(norbs))
  opts%nclose = nclose
  opts%scfcrt = 1.d-6
  opts%enuclr = enuclr
  opts%atheat = atheat
  opts%pdiag = c_loc(pdiag)
  call hip_iter(opts, res, pscf, fscf, cscf, eigscf)
  call hip_finalize()

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) f, e, fcopy, eig, p, res%escf, res%niter, res%converged, pscf, eigscf
  close (u)
  write (*, '(a, f16.6, a, i4)') 'hip_iter escf ', res%escf, ' niter ', res%niter
end program test_mopac_hip
