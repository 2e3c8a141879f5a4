! C-callable harness around MOPAC's own MOZYME kernels — TEST INFRASTRUCTURE.
!
! oracle/build_ref.py compiles this file with the reference's
! src/MOZYME/diagg1.F90, diagg2.F90, density_for_MOZYME.F90, epseta.F90 and
! the modules they read (MOZYME_C, Common_arrays_C, parameters_C, ...).  The
! routines keep their LMOs and sparse matrices in module arrays; mz_setup and
! mz_set_lmos load those arrays from plain C arrays, mz_get_lmos copies them
! back, so a test can run the reference and the GPU kernels on identical data.
! Atom numbers (icocc, icvir) are 1-based, offsets (nncf, ncocc, nnce, ncvir,
! nijbo) 0-based, as in the reference.
module ref_harness_mozyme
  use iso_c_binding
  implicit none
contains

  subroutine mz_setup(numat_in, norbs_in, mpack_in, nfirst_in, nlast_in, nat_in, nijbo_in, &
                      thresh_in, shift_in) bind(C, name="mz_setup")
    use molkst_C, only: numat, norbs, mpack, numcal, keywrd, id
    use common_arrays_C, only: nfirst, nlast, nat, p, eigs
    use MOZYME_C, only: iorbs, nijbo, lijbo, thresh, shift
    integer(c_int), value :: numat_in, norbs_in, mpack_in
    integer(c_int) :: nfirst_in(numat_in), nlast_in(numat_in), nat_in(numat_in)
    integer(c_int) :: nijbo_in(numat_in, numat_in)
    real(c_double), value :: thresh_in, shift_in
    numat = numat_in
    norbs = norbs_in
    mpack = mpack_in
    id = 0
    keywrd = ' 1SCF MOZYME'
    numcal = numcal + 1
    if (allocated(nfirst)) deallocate (nfirst, nlast, nat)
    if (allocated(p)) deallocate (p)
    if (allocated(eigs)) deallocate (eigs)
    if (allocated(iorbs)) deallocate (iorbs)
    if (allocated(nijbo)) deallocate (nijbo)
    allocate (nfirst(numat), nlast(numat), nat(numat), p(mpack), eigs(norbs), iorbs(numat), &
              nijbo(numat, numat))
    nfirst = nfirst_in
    nlast = nlast_in
    nat = nat_in
    iorbs = nlast - nfirst + 1
    nijbo = nijbo_in
    lijbo = .true.
    thresh = thresh_in
    shift = shift_in
    p = 0.d0
    eigs = 0.d0
  end subroutine mz_setup

  ! LMO arrays and their capacities; eigs_occ -> common eigs(1:nocc)
  subroutine mz_set_lmos(nocc, nvir, ncf_in, nncf_in, ncocc_in, icocc_dim_in, icocc_in, &
                         cocc_dim_in, cocc_in, nce_in, nnce_in, ncvir_in, icvir_dim_in, icvir_in, &
                         cvir_dim_in, cvir_in, fmo_dim_in, eigs_occ, p_in) bind(C, name="mz_set_lmos")
    use molkst_C, only: mpack
    use common_arrays_C, only: eigs, p
    use MOZYME_C, only: ncf, nncf, ncocc, icocc, cocc, nce, nnce, ncvir, icvir, cvir, &
      icocc_dim, cocc_dim, icvir_dim, cvir_dim, fmo_dim, fmo, ifmo, nfmo, nvirtual, noccupied
    integer(c_int), value :: nocc, nvir, icocc_dim_in, cocc_dim_in, icvir_dim_in, cvir_dim_in, fmo_dim_in
    integer(c_int) :: ncf_in(nocc), nncf_in(nocc), ncocc_in(nocc), icocc_in(icocc_dim_in)
    integer(c_int) :: nce_in(nvir), nnce_in(nvir), ncvir_in(nvir), icvir_in(icvir_dim_in)
    real(c_double) :: cocc_in(cocc_dim_in), cvir_in(cvir_dim_in), eigs_occ(nocc), p_in(mpack)
    if (allocated(ncf)) deallocate (ncf, nncf, ncocc, icocc, cocc)
    if (allocated(nce)) deallocate (nce, nnce, ncvir, icvir, cvir)
    if (allocated(fmo)) deallocate (fmo, ifmo)
    if (allocated(nfmo)) deallocate (nfmo)
    noccupied = nocc
    nvirtual = nvir
    icocc_dim = icocc_dim_in
    cocc_dim = cocc_dim_in
    icvir_dim = icvir_dim_in
    cvir_dim = cvir_dim_in
    fmo_dim = fmo_dim_in
    allocate (ncf(nocc), nncf(nocc), ncocc(nocc), icocc(icocc_dim), cocc(cocc_dim))
    allocate (nce(nvir), nnce(nvir), ncvir(nvir), icvir(icvir_dim), cvir(cvir_dim))
    allocate (fmo(fmo_dim), ifmo(2, fmo_dim), nfmo(nvir))
    ncf = ncf_in
    nncf = nncf_in
    ncocc = ncocc_in
    icocc = icocc_in
    cocc = cocc_in
    nce = nce_in
    nnce = nnce_in
    ncvir = ncvir_in
    icvir = icvir_in
    cvir = cvir_in
    fmo = 0.d0
    ifmo = 0
    nfmo = 0
    eigs(1:nocc) = eigs_occ
    p = p_in
  end subroutine mz_set_lmos

  subroutine mz_get_lmos(nocc, nvir, ncf_out, icocc_out, cocc_out, nce_out, icvir_out, cvir_out, &
                         eigs_occ) bind(C, name="mz_get_lmos")
    use common_arrays_C, only: eigs
    use MOZYME_C, only: ncf, icocc, cocc, nce, icvir, cvir, icocc_dim, cocc_dim, icvir_dim, cvir_dim
    integer(c_int), value :: nocc, nvir
    integer(c_int) :: ncf_out(nocc), nce_out(nvir), icocc_out(*), icvir_out(*)
    real(c_double) :: cocc_out(*), cvir_out(*), eigs_occ(nocc)
    ncf_out = ncf
    nce_out = nce
    icocc_out(1:icocc_dim) = icocc
    icvir_out(1:icvir_dim) = icvir
    cocc_out(1:cocc_dim) = cocc
    cvir_out(1:cvir_dim) = cvir
    eigs_occ = eigs(1:nocc)
  end subroutine mz_get_lmos

  ! density_for_MOZYME(p, mode, nclose, partp)
  subroutine mz_density(p_out, mode, nclose, partp) bind(C, name="mz_density")
    use molkst_C, only: mpack
    real(c_double) :: p_out(mpack), partp(mpack)
    integer(c_int), value :: mode, nclose
    call density_for_MOZYME(p_out, mode, nclose, partp)
  end subroutine mz_density

  ! diagg1 as diagg calls it (diagg.F90:39-40), with its module results
  subroutine mz_diagg1(fao, nocc, nvir, eigv, idiagg, nij, fmo_out, ifmo_out, tiny_io, sumt_out, &
                       ovmax_out) bind(C, name="mz_diagg1")
    use molkst_C, only: mpack, numat, norbs
    use MOZYME_C, only: fmo, ifmo, fmo_dim, icocc_dim, tiny, sumt, ovmax
    real(c_double) :: fao(mpack), eigv(nvir), fmo_out(*), tiny_io, sumt_out, ovmax_out
    integer(c_int), value :: nocc, nvir, idiagg
    integer(c_int) :: nij, ifmo_out(2, *)
    logical, allocatable :: latoms(:)
    double precision, allocatable :: ws(:), avir(:), aocc(:), aov(:)
    allocate (latoms(numat), ws(norbs), avir(norbs), aocc(max(1, icocc_dim)), aov(numat))
    tiny = tiny_io
    nij = fmo_dim
    call diagg1(fao, nocc, nvir, eigv, ws, latoms, ifmo, fmo, fmo_dim, nij, idiagg, avir, aocc, aov)
    fmo_out(1:nij) = fmo(1:nij)
    ifmo_out(:, 1:nij) = ifmo(:, 1:nij)
    tiny_io = tiny
    sumt_out = sumt
    ovmax_out = ovmax
  end subroutine mz_diagg1

  ! diagg2 on the list left by mz_diagg1 (diagg.F90:51)
  subroutine mz_diagg2(nocc, nvir, eigv, nij, idiagg, tiny_io, sumb_out) bind(C, name="mz_diagg2")
    use molkst_C, only: numat, norbs
    use MOZYME_C, only: tiny, sumb
    integer(c_int), value :: nocc, nvir, nij, idiagg
    real(c_double) :: eigv(nvir), tiny_io, sumb_out
    integer, allocatable :: iused(:)
    logical, allocatable :: latoms(:)
    double precision, allocatable :: storei(:), storej(:)
    allocate (iused(numat), latoms(numat), storei(norbs), storej(norbs))
    tiny = tiny_io
    call diagg2(nocc, nvir, eigv, iused, latoms, nij, idiagg, storei, storej)
    tiny_io = tiny
    sumb_out = sumb
  end subroutine mz_diagg2

  ! helecz() on sparse P, H, F (src/MOZYME/helecz.F90:16-52)
  function mz_helecz(p_in, h_in, f_in) result(e) bind(C, name="mz_helecz")
    use molkst_C, only: mpack
    use common_arrays_C, only: p, h, f
    real(c_double) :: p_in(mpack), h_in(mpack), f_in(mpack)
    real(c_double) :: e
    double precision, external :: helecz
    if (allocated(h)) deallocate (h)
    if (allocated(f)) deallocate (f)
    allocate (h(mpack), f(mpack))
    p = p_in
    h = h_in
    f = f_in
    e = helecz()
  end function mz_helecz

  ! cnvgz(pnew, p, p1, p2, p3, niter, idiag) (src/MOZYME/cnvgz.F90:16-85)
  subroutine mz_cnvgz(pnew, p, p1, p2, p3, niter, idiag, three_point, pmax_out) bind(C, name="mz_cnvgz")
    use molkst_C, only: norbs, mpack
    use MOZYME_C, only: use_three_point_extrap, pmax
    real(c_double) :: pnew(mpack), p(mpack), p1(norbs), p2(norbs), p3(norbs), pmax_out
    integer(c_int), value :: niter, three_point
    integer(c_int) :: idiag(norbs)
    integer :: id(norbs)
    id = idiag
    use_three_point_extrap = three_point /= 0
    call cnvgz(pnew, p, p1, p2, p3, niter, id)
    pmax_out = pmax
  end subroutine mz_cnvgz

  ! buildf(f, partf, 0) with fock2z/fz2n (src/MOZYME/buildf.F90:16-52,
  ! src/MOZYME/fock2z.F90): f = h + G(p) on the sparse storage of mz_setup,
  ! discrete system (ione = 1), integrals as stored for SEMIDIRECT (direct
  ! false, semidr true), with or without PM7's smooth switch to point charges
  ! (l_feather, trunc_1, trunc_2).  Per-atom parameters go into parameters_C
  ! at each atom's element; a0 and ev into funcon_C.
  subroutine mz_fock2z(xyz, am_in, dd_in, ad_in, nw, w_in, p_in, h_in, a0_in, ev_in, feather, trunc1, trunc2, &
      f_out) bind(C, name="mz_fock2z")
    use molkst_C, only: numat, norbs, mpack, l_feather, trunc_1, trunc_2
    use common_arrays_C, only: coord, p, nat, ifact
    use MOZYME_C, only: direct, semidr, kopt
    use parameters_C, only: am, dd, ad
    use funcon_C, only: a0, ev
    use cosmo_C, only: useps
    integer(c_long_long), value :: nw
    real(c_double) :: xyz(3, numat), am_in(numat), dd_in(numat), ad_in(numat)
    real(c_double) :: w_in(nw), p_in(mpack), h_in(mpack), f_out(mpack)
    real(c_double), value :: a0_in, ev_in, trunc1, trunc2
    integer(c_int), value :: feather
    double precision, allocatable :: q(:), qe(:), ptot2(:, :)
    integer :: i
    if (allocated(coord)) deallocate (coord)
    if (allocated(ifact)) deallocate (ifact)
    if (allocated(kopt)) deallocate (kopt)
    allocate (coord(3, numat), ifact(norbs + 1), kopt(numat), q(numat), qe(numat), ptot2(numat, 81))
    coord = xyz
    do i = 1, norbs + 1
      ifact(i) = (i * (i - 1)) / 2
    end do
    kopt = 0
    do i = 1, numat
      am(nat(i)) = am_in(i)
      dd(nat(i)) = dd_in(i)
      ad(nat(i)) = ad_in(i)
    end do
    a0 = a0_in
    ev = ev_in
    direct = .false.
    semidr = .true.
    l_feather = feather /= 0
    trunc_1 = trunc1
    trunc_2 = trunc2
    useps = .false.
    p = p_in
    f_out = h_in
    call fock2z(f_out, q, qe, w_in, w_in, ptot2, 0, 1)
    deallocate (q, qe, ptot2)
  end subroutine mz_fock2z


  ! tidy (src/MOZYME/tidy.F90:16-244) on one LMO set, at a single point
  ! (step_num 1: one compression pass, then the free space spread over the
  ! LMOs).  mode 1 occupied, 2 virtual.  Needs mz_setup (numat, norbs, iorbs).
  ! err = 1 when tidy finds the storage too small (moperr).
  subroutine mz_tidy(nmo, nc, ic, n01, c, n02, nnc, ncmo, mode, thresh_in, ln, mn, err) bind(C, name="mz_tidy")
    use molkst_C, only: step_num, step_num0, moperr, use_disk
    use MOZYME_C, only: thresh
    integer(c_int), value :: nmo, n01, n02, mode
    integer(c_int) :: nc(nmo), ic(n01), nnc(nmo), ncmo(nmo), ln, mn, err
    real(c_double) :: c(n02)
    real(c_double), value :: thresh_in
    integer :: l1, m1
    step_num = 1
    step_num0 = 0
    use_disk = .false.
    moperr = .false.
    thresh = thresh_in
    call tidy(nmo, nc, ic, n01, c, n02, nnc, ncmo, l1, m1, mode)
    ln = l1
    mn = m1
    err = 0
    if (moperr) err = 1
    moperr = .false.
  end subroutine mz_tidy

  ! check (src/MOZYME/check.F90:16-113): renormalize nvec LMOs; ws_out(i) the
  ! squared norm found.  Arrays indexed to nvec+1 as check declares them.
  ! Only for sets whose summed error is <= 0.1 (above that check stops the
  ! job through mopend, which is not part of this library).
  subroutine mz_check(nvec, nnc, nc, ic, ic_dim, ncvec, c, c_dim, ws_out) bind(C, name="mz_check")
    use MOZYME_C, only: iorbs, ws
    integer(c_int), value :: nvec, ic_dim, c_dim
    integer(c_int) :: nnc(nvec + 1), nc(nvec + 1), ic(ic_dim), ncvec(nvec + 1)
    real(c_double) :: c(c_dim), ws_out(nvec)
    if (allocated(ws)) deallocate (ws)
    allocate (ws(nvec + 1))
    call check(nvec, nnc, nc, ic, ic_dim, iorbs, ncvec, c, c_dim)
    ws_out = ws(1:nvec)
  end subroutine mz_check

  ! eimp (src/MOZYME/eimp.F90:16-45) on the sparse storage of mz_setup:
  ! p_io comes back with the first slot of each stored off-diagonal block set
  ! to the sum of squares of that block of f_in.
  subroutine mz_eimp(f_in, p_io) bind(C, name="mz_eimp")
    use molkst_C, only: mpack
    use common_arrays_C, only: p, f
    real(c_double) :: f_in(mpack), p_io(mpack)
    if (allocated(f)) deallocate (f)
    allocate (f(mpack))
    f = f_in
    p = p_io
    call eimp()
    p_io = p
  end subroutine mz_eimp


  ! The start of an iter_for_MOZYME iteration on the loaded LMOs
  ! (iter_for_MOZYME.F90:221, 326, 482-484, 495): tidy on the occupied and
  ! the virtual set, check on both, then eimp on p_in with f_in, leaving the
  ! result in the module's p (what diagg1 reads).  err = 1 when tidy found
  ! the storage too small, 2 when check found an error above 0.1 (not
  ! reached here: check would stop the job).
  subroutine mz_iter_prologue(nocc, nvir, p_in, f_in, err) bind(C, name="mz_iter_prologue")
    use molkst_C, only: mpack, step_num, step_num0, moperr, use_disk
    use common_arrays_C, only: p, f
    use MOZYME_C, only: ncf, nncf, ncocc, icocc, cocc, nce, nnce, ncvir, icvir, cvir, icocc_dim, cocc_dim, &
      icvir_dim, cvir_dim, iorbs, ws
    integer(c_int), value :: nocc, nvir
    real(c_double) :: p_in(mpack), f_in(mpack)
    integer(c_int) :: err
    integer :: lno, lnv, mn
    step_num = 1
    step_num0 = 0
    use_disk = .false.
    moperr = .false.
    err = 0
    call tidy(nocc, ncf, icocc, icocc_dim, cocc, cocc_dim, nncf, ncocc, lno, mn, 1)
    if (moperr) err = 1
    call tidy(nvir, nce, icvir, icvir_dim, cvir, cvir_dim, nnce, ncvir, lnv, mn, 2)
    if (moperr) err = 1
    moperr = .false.
    if (err /= 0) return
    if (allocated(ws)) deallocate (ws)
    allocate (ws(max(nocc, nvir) + 1))
    call check(nocc, nncf, ncf, icocc, icocc_dim, iorbs, ncocc, cocc, cocc_dim)
    call check(nvir, nnce, nce, icvir, icvir_dim, iorbs, ncvir, cvir, cvir_dim)
    if (allocated(f)) deallocate (f)
    allocate (f(mpack))
    f = f_in
    p = p_in
    call eimp()
  end subroutine mz_iter_prologue

  ! The LMO offsets tidy may have moved (nncf, ncocc, nnce, ncvir)
  subroutine mz_get_offsets(nocc, nvir, nncf_out, ncocc_out, nnce_out, ncvir_out) bind(C, name="mz_get_offsets")
    use MOZYME_C, only: nncf, ncocc, nnce, ncvir
    integer(c_int), value :: nocc, nvir
    integer(c_int) :: nncf_out(nocc), ncocc_out(nocc), nnce_out(nvir), ncvir_out(nvir)
    nncf_out = nncf(1:nocc)
    ncocc_out = ncocc(1:nocc)
    nnce_out = nnce(1:nvir)
    ncvir_out = ncvir(1:nvir)
  end subroutine mz_get_offsets


  ! the level shift diagg2 reads (MOZYME_C shift)
  subroutine mz_set_shift(shift_in) bind(C, name="mz_set_shift")
    use MOZYME_C, only: shift
    real(c_double), value :: shift_in
    shift = shift_in
  end subroutine mz_set_shift

end module ref_harness_mozyme
