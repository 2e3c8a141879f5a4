! C-callable harness around MOPAC's own PM7 integral and post-SCF routines —
! TEST INFRASTRUCTURE.
!
! oracle/build_ref.py compiles this file with the reference's
! src/integrals/rotate.F90, h1elec.F90, diat.F90, set.F90, coe.F90, ccrep.F90,
! mndod.F90, src/models/calpar.F90, parameters_for_PM7_C.F90 and the
! post-SCF correction files, so the library's PM7 producer and its
! dispersion / hydrogen-bond terms can be compared with the reference on
! identical geometries.
!
!   pm7_ref_init            PM7 parameters into parameters_C as the PM7 branch
!                           of models/switch.F90:409-439 does, element data as
!                           interface/mopac_api_initialize.F90:116-131 and
!                           moldat.F90:101-111 set them, then calpar
!   pm7_ref_element         derived multipole data of one element
!   pm7_ref_pair            rotate (W block, e1b, e2a, core-core) and h1elec
!                           (resonance block) for one atom pair
!   pm7_ref_corrections     PM6_DH_Dispersion and Hydrogen_bond_corrections
!                           for a molecule, as post_scf_corrections calls them
!                           for PM7 (src/corrections/post_scf_corrections.F90:112-114)
!   pm7_ref_mozyme_hcore    fillij then hcore_for_MOZYME on a discrete molecule
!                           (SEMIDIRECT): nijbo, the sparse one-electron matrix
!                           and the core-core energy
module ref_harness_pm7
  use iso_c_binding
  implicit none
contains

  subroutine pm7_ref_init() bind(C, name="pm7_ref_init")
    use parameters_C
    use Parameters_for_PM7_C
    use molkst_C, only: method_pm7, keywrd, id, numcal, l_feather, trunc_1, trunc_2
    use funcon_C, only: fpc
    use conref_C, only: fpcref
    use overlaps_C, only: cutof1, cutof2
    use MOZYME_C, only: cutofs
    integer :: i
    method_pm7 = .true.
    l_feather = .true.
    trunc_1 = 7.0d0
    trunc_2 = 0.22d0
    id = 0
    keywrd = ' PM7 1SCF'
    numcal = numcal + 1
    fpc(:) = fpcref(1, :)
    call fbx
    call fordd
    guess1 = gues71
    guess2 = gues72
    guess3 = gues73
    zs = zs7
    zp = zp7
    zd = zd7
    zsn = zsn7
    zpn = zpn7
    zdn = zdn7
    uss = uss7
    upp = upp7
    udd = udd7
    betas = betas7
    betap = betap7
    betad = betad7
    gss = gss7
    gpp = gpp7
    gsp = gsp7
    gp2 = gp27
    hsp = hsp7
    f0sd = f0sd7
    g2sd = g2sd7
    alp = alp7
    pocord = poc_7
    polvol = polvo7
    v_par = v_par7
    call alpb_and_xfac_pm7
    f0sd_store = f0sd
    g2sd_store = g2sd
    do i = 1, 100
      alpb(:i, i) = alpb(i, :i)
      xfac(:i, i) = xfac(i, :i)
    end do
    tore = ios + iop + iod
    do i = 1, 107
      dorbs(i) = (zd(i) > 1.d-8)
      if (dorbs(i)) then
        natorb(i) = 9
      else if (zp(i) > 1.d-20) then
        natorb(i) = 4
      else if (zs(i) > 1.d-20) then
        natorb(i) = 1
      else
        natorb(i) = 0
      end if
    end do
    call calpar
    cutofs = 1.d4
    cutof2 = 1.d10
    cutof1 = 225.d0
  end subroutine pm7_ref_init

  ! out(1:21) = natorb, tore, am, ad, aq, dd, qq, eisol, eheat, po(1), po(2),
  !             po(3), po(9), ddp(3), uss, upp, gss, gsp, gpp, gp2, hsp
  subroutine pm7_ref_element(z, out) bind(C, name="pm7_ref_element")
    use parameters_C, only: natorb, tore, am, ad, aq, dd, qq, eisol, eheat, po, ddp, &
      uss, upp, gss, gsp, gpp, gp2, hsp
    integer(c_int), value :: z
    real(c_double) :: out(21)
    out(1) = natorb(z)
    out(2) = tore(z)
    out(3) = am(z)
    out(4) = ad(z)
    out(5) = aq(z)
    out(6) = dd(z)
    out(7) = qq(z)
    out(8) = eisol(z)
    out(9) = eheat(z)
    out(10) = po(1, z)
    out(11) = po(2, z)
    out(12) = po(3, z)
    out(13) = po(9, z)
    out(14) = ddp(3, z)
    out(15) = uss(z)
    out(16) = upp(z)
    out(17) = gss(z)
    out(18) = gsp(z)
    out(19) = gpp(z)
    out(20) = gp2(z)
    out(21) = hsp(z)
  end subroutine pm7_ref_element

  ! One atom pair, atom i (element ni at xi) before atom j (nj at xj) as hcore
  ! visits them (i > j).  w(1:limij*limkl) in rotate's order, e1b/e2a packed
  ! (45), enuc (eV), di(9,9) the resonance block h1elec returns (orbitals of i
  ! down, orbitals of j across).
  subroutine pm7_ref_pair(ni, nj, xi, xj, w, e1b, e2a, enuc, di) bind(C, name="pm7_ref_pair")
    integer(c_int), value :: ni, nj
    real(c_double) :: xi(3), xj(3), w(2025), e1b(45), e2a(45), enuc, di(9, 9)
    integer :: kr
    kr = 1
    w = 0.d0
    call rotate(ni, nj, xi, xj, w, kr, e1b, e2a, enuc)
    call h1elec(ni, nj, xi, xj, di)
  end subroutine pm7_ref_pair

  ! PM7 post-SCF corrections of a discrete molecule (kcal/mol): out(1) the
  ! dispersion term, out(2) the hydrogen-bond term, out(3) the hydrogen bonds
  ! counted (sum < -1 kcal/mol).  Bonding topology from set_up_dentate, as the
  ! reference sets it up before the SCF.
  subroutine pm7_ref_corrections(numat_in, nat_in, xyz, out) bind(C, name="pm7_ref_corrections")
    use molkst_C, only: numat, numcal, id, N_Hbonds
    use common_arrays_C, only: coord, nat, nbonds, ibonds, q, p, dxyz, txtatm
    use MOZYME_C, only: nijbo
    integer(c_int), value :: numat_in
    integer(c_int) :: nat_in(numat_in)
    real(c_double) :: xyz(3, numat_in), out(3)
    double precision, external :: PM6_DH_Dispersion, Hydrogen_bond_corrections
    numat = numat_in
    id = 0
    numcal = numcal + 1
    if (allocated(coord)) deallocate (coord)
    if (allocated(nat)) deallocate (nat)
    if (allocated(nbonds)) deallocate (nbonds)
    if (allocated(ibonds)) deallocate (ibonds)
    if (allocated(q)) deallocate (q)
    if (allocated(txtatm)) deallocate (txtatm)
    allocate (coord(3, numat), nat(numat), nbonds(numat), ibonds(15, numat), q(numat), txtatm(numat))
    coord = xyz
    nat = nat_in
    q = 0.d0
    txtatm = ' '
    if (allocated(dxyz)) deallocate (dxyz)
    allocate (dxyz(3*numat))
    dxyz = 0.d0
    ! a conventional (non-MOZYME) run has no nijbo; set_up_dentate would
    ! filter bonds through one left by an earlier MOZYME call
    if (allocated(nijbo)) deallocate (nijbo)
    call set_up_dentate()
    out(1) = PM6_DH_Dispersion(.false.)
    out(2) = Hydrogen_bond_corrections(.false., .false.)
    out(3) = N_Hbonds
  end subroutine pm7_ref_corrections

  ! fillij(.false.) then hcore_for_MOZYME() for a discrete molecule, as
  ! set_up_MOZYME_arrays and compfg reach them (SEMIDIRECT integrals).  The
  ! cutoffs are MOZYME's defaults unless cut2 / cut1 / cuts (Angstrom) are > 0,
  ! which go in as the CUTOF2= / CUTOF1= / CUTOFS= keywords.  nijbo_out gets
  ! fillij's nijbo (0-based offsets, -2 dipole shell, -1 point charge),
  ! mpack_out the sparse size; h_out (capacity hcap) the one-electron matrix
  ! and enuclr_out the core-core energy (eV) when mpack <= hcap.  The
  ! conventional cutoffs pm7_ref_init sets are restored on return.
  subroutine pm7_ref_mozyme_hcore(numat_in, nat_in, xyz, cut2, cut1, cuts, nijbo_out, hcap, mpack_out, h_out, &
      enuclr_out) bind(C, name="pm7_ref_mozyme_hcore")
    use molkst_C, only: numat, natoms, norbs, id, keywrd, mpack, n2elec, cutofp, ispd, numcal, enuclr
    use common_arrays_C, only: coord, nat, h, w, uspd
    use MOZYME_C, only: iorbs, nijbo, lijbo, jopt, numred, mode, cutofs
    use overlaps_C, only: cutof1, cutof2
    use parameters_C, only: natorb, uss, upp
    use cosmo_C, only: useps
    integer(c_int), value :: numat_in, hcap
    integer(c_int) :: nat_in(numat_in), nijbo_out(numat_in, numat_in), mpack_out
    real(c_double) :: xyz(3, numat_in), h_out(*), enuclr_out
    real(c_double), value :: cut2, cut1, cuts
    character(len=40) :: word
    integer :: i, k, m
    numat = numat_in
    natoms = numat_in
    id = 0
    ispd = 0
    cutofp = 1.d10
    useps = .false.
    keywrd = ' PM7 MOZYME 1SCF SEMIDIRECT'
    if (cut2 > 0.d0) then
      write (word, '(A,F0.6)') ' CUTOF2=', cut2
      keywrd = trim(keywrd) // trim(word)
    end if
    if (cut1 > 0.d0) then
      write (word, '(A,F0.6)') ' CUTOF1=', cut1
      keywrd = trim(keywrd) // trim(word)
    end if
    if (cuts > 0.d0) then
      write (word, '(A,F0.6)') ' CUTOFS=', cuts
      keywrd = trim(keywrd) // trim(word)
    end if
    numcal = numcal + 1
    if (allocated(coord)) deallocate (coord)
    if (allocated(nat)) deallocate (nat)
    if (allocated(iorbs)) deallocate (iorbs)
    if (allocated(nijbo)) deallocate (nijbo)
    if (allocated(jopt)) deallocate (jopt)
    allocate (coord(3, numat), nat(numat), iorbs(numat), jopt(numat))
    coord = xyz
    nat = nat_in
    norbs = 0
    do i = 1, numat
      iorbs(i) = natorb(nat(i))
      norbs = norbs + iorbs(i)
    end do
    jopt = 0
    numred = numat
    mode = 0
    lijbo = .true.
    call fillij(.false.)
    mpack_out = mpack
    nijbo_out = nijbo
    if (mpack <= hcap) then
      if (allocated(h)) deallocate (h)
      if (allocated(w)) deallocate (w)
      if (allocated(uspd)) deallocate (uspd)
      allocate (h(mpack), w(n2elec + 2025), uspd(norbs))
      m = 0
      do i = 1, numat
        do k = 1, iorbs(i)
          m = m + 1
          if (k == 1) then
            uspd(m) = uss(nat(i))
          else
            uspd(m) = upp(nat(i))
          end if
        end do
      end do
      call hcore_for_MOZYME()
      h_out(1:mpack) = h(1:mpack)
      enuclr_out = enuclr
    end if
    deallocate (nijbo)
    cutofs = 1.d4
    cutof2 = 1.d10
    cutof1 = 225.d0
  end subroutine pm7_ref_mozyme_hcore

end module ref_harness_pm7
