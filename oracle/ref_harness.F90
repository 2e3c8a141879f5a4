! C-callable harness around MOPAC's own SCF routines — TEST INFRASTRUCTURE.
!
! oracle/build_ref.py compiles this file together with the reference's
! sources (src/SCF/fock2.F90, fock1.F90, src/integrals/jab.F90, kab.F90,
! helect.F90, src/SCF/cnvg.F90, pulay.F90, src/matrix/diag.F90, osinv.F90 and
! the modules they use) into oracle/_ref/libmopac_ref.so.  Tests call it to
! check the C restatement (oracle_scf.c) routine by routine on identical
! inputs.  Nothing here is part of the product library.
!
! ref_setup sets the molkst_C values those routines read and starts a new
! "calculation" (numcal), which resets their saved state.
module ref_harness
  use iso_c_binding
  implicit none
contains

  subroutine ref_setup(n, nat, n2el, npul) bind(C, name="ref_setup")
    use molkst_C, only: norbs, mpack, numat, n2elec, numcal, id, keywrd, npulay, method_indo
    use iter_C, only: pulay_work1, pulay_work2, pulay_work3
    integer(c_int), value :: n, nat, n2el, npul
    norbs = n
    mpack = (n * (n + 1)) / 2
    numat = nat
    n2elec = n2el
    id = 0
    npulay = npul
    keywrd = ' 1SCF'
    method_indo = .false.
    numcal = numcal + 1
    if (allocated(pulay_work1)) deallocate (pulay_work1, pulay_work2, pulay_work3)
    allocate (pulay_work1(n, n), pulay_work2(n, n), pulay_work3(n, n))
  end subroutine ref_setup

  subroutine ref_fock2(f, ptot, p, w, nat, nfirst, nlast, mode) bind(C, name="ref_fock2")
    use molkst_C, only: mpack, n2elec
    integer(c_int), value :: nat, mode
    real(c_double) :: f(mpack), ptot(mpack), p(mpack), w(n2elec)
    integer(c_int) :: nfirst(nat), nlast(nat)
    integer :: nf(nat), nl(nat)
    nf = nfirst
    nl = nlast
    call fock2(f, ptot, p, w, w, w, nat, nf, nl, mode)
  end subroutine ref_fock2

  function ref_helect(n, p, h, f) result(e) bind(C, name="ref_helect")
    integer(c_int), value :: n
    real(c_double) :: p(*), h(*), f(*)
    real(c_double) :: e
    double precision, external :: helect
    e = helect(n, p, h, f)
  end function ref_helect

  subroutine ref_cnvg(pnew, p, p1, niter, pl) bind(C, name="ref_cnvg")
    use molkst_C, only: mpack, norbs
    integer(c_int), value :: niter
    real(c_double) :: pnew(mpack), p(mpack), p1(norbs), pl
    call cnvg(pnew, p, p1, niter, pl)
  end subroutine ref_cnvg

  subroutine ref_diag(fao, vector, nocc, eig, n) bind(C, name="ref_diag")
    use molkst_C, only: mpack
    integer(c_int), value :: nocc, n
    real(c_double) :: fao(mpack), vector(n, n), eig(n)
    call diag(fao, vector, nocc, eig, n, n)
  end subroutine ref_diag

  subroutine ref_osinv(a, n, d) bind(C, name="ref_osinv")
    integer(c_int), value :: n
    real(c_double) :: a(n * n), d
    call osinv(a, n, d)
  end subroutine ref_osinv

  ! ncalls Pulay steps from a fresh start with the arrays iter passes
  ! (pold = fppf, pold2 = fock, pold3 = emat; src/SCF/iter.F90:830-831).
  subroutine ref_pulay_seq(n, ncalls, f_seq, p_seq, pl_seq) bind(C, name="ref_pulay_seq")
    use molkst_C, only: mpack, npulay
    integer(c_int), value :: n, ncalls
    real(c_double) :: f_seq(mpack, ncalls), p_seq(mpack, ncalls), pl_seq(ncalls)
    real(c_double), allocatable :: fppf(:), fock(:), emat(:)
    integer :: lfock, nfock, k
    logical :: start
    allocate (fppf(npulay * mpack), fock(npulay * mpack), emat((npulay + 1)**2))
    fppf = 0.d0
    fock = 0.d0
    emat = 0.d0
    lfock = 0
    nfock = 0
    start = .true.
    do k = 1, ncalls
      call pulay(f_seq(1, k), p_seq(1, k), n, fppf, fock, emat, lfock, nfock, npulay * mpack, &
                 start, pl_seq(k))
    end do
    deallocate (fppf, fock, emat)
  end subroutine ref_pulay_seq

end module ref_harness
