! Fortran host side of libmopachip: ISO_C_BINDING interfaces to the C ABI
! (include/mopac_hip.h) and wrappers with the calling surface of the MOPAC
! routines they replace, so a Fortran SCF driver changes only the routine
! names it calls:
!
!   hip_fock2        <- fock2(f, ptot, p, w, wj, wk, numat, nfirst, nlast, mode)
!                       (src/SCF/fock2.F90:16)
!   hip_helect       <- helect(n, p, h, f)              (src/integrals/helect.F90:16)
!   hip_eigenvectors <- eigenvectors_LAPACK(eigenvecs, xmat, eigvals, ndim)
!                       (src/matrix/eigenvectors_LAPACK.F90:16)
!   hip_densit       <- densit(c, mdim, norbs, nocc, occ, nfract, fract, p, mode),
!                       closed shell (nfract = 0)        (src/matrix/densit.F90:16)
!   hip_diag         <- diag(fao, vector, nocc, eig, mdim, n) (src/matrix/diag.F90:16)
!   hip_iter         <- the RHF loop of iter            (src/SCF/iter.F90:16)
!
! One module-level context (device 0 unless hip_init is given another), as the
! reference keeps one global set of handles.  W stays resident after
! hip_upload, so hip_fock2 ignores its w/wj/wk arguments (they must be the
! arrays that were uploaded).  Errors stop the program with the library's
! message, the way the reference's GPU paths stop on a failed call.
module mopac_hip
  use iso_c_binding
  implicit none
  private
  public :: hip_init, hip_finalize, hip_upload, hip_fock2, hip_helect, hip_eigenvectors, &
            hip_densit, hip_diag, hip_iter, mhip_scf_opts, mhip_scf_result

  type, bind(C) :: mhip_scf_opts
    integer(c_int) :: nclose = 0, pulay = 0, itrmax = 2000
    real(c_double) :: scfcrt = 1.d-4
    integer(c_int) :: last = 1, precision = 0, pseudo_diag = 0, adaptive_eigen = 0
    real(c_double) :: enuclr = 0.d0, atheat = 0.d0
    type(c_ptr) :: pdiag = c_null_ptr, p_init = c_null_ptr
    integer(c_int) :: block_start = 0
    type(c_ptr) :: fragment = c_null_ptr
    integer(c_int) :: start_from_guess = 0
  end type

  type, bind(C) :: mhip_scf_result
    real(c_double) :: escf, ee
    integer(c_int) :: niter, converged, n_full_diag, n_pseudo_diag
    real(c_double) :: seconds
    integer(c_int) :: jacobi_sweeps, n_fragment_diag, n_split_iter
  end type

  interface
    integer(c_int) function mhip_init(device, ctx) bind(C, name="mhip_init")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function mhip_finalize(ctx) bind(C, name="mhip_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mhip_last_error() bind(C, name="mhip_last_error")
      import :: c_ptr
    end function
    integer(c_int) function mhip_upload_system(ctx, numat, nfirst, nlast, norbs, n2elec, h, w) &
        bind(C, name="mhip_upload_system")
      import :: c_int, c_ptr, c_long_long, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: numat, norbs
      integer(c_int) :: nfirst(*), nlast(*)
      integer(c_long_long), value :: n2elec
      real(c_double) :: h(*), w(*)
    end function
    integer(c_int) function mhip_fock2(ctx, f, ptot, p, mode) bind(C, name="mhip_fock2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f(*), ptot(*), p(*)
      integer(c_int), value :: mode
    end function
    integer(c_int) function mhip_helect(ctx, norbs, p, h, f, e) bind(C, name="mhip_helect")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: norbs
      real(c_double) :: p(*), h(*), f(*), e
    end function
    integer(c_int) function mhip_eigen(ctx, c, fpacked, eig, ndim) bind(C, name="mhip_eigen")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: c(*), fpacked(*), eig(*)
      integer(c_int), value :: ndim
    end function
    integer(c_int) function mhip_density(ctx, c, ndubl, occ, p) bind(C, name="mhip_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: c(*), p(*)
      integer(c_int), value :: ndubl
      real(c_double), value :: occ
    end function
    integer(c_int) function mhip_pseudo_diag(ctx, fao, vector, nocc, eig, mdim, n, nrot) &
        bind(C, name="mhip_pseudo_diag")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double) :: fao(*), vector(*), eig(*)
      integer(c_int), value :: nocc, mdim, n
      integer(c_long_long) :: nrot
    end function
    integer(c_int) function mhip_scf(ctx, opts, res, p, f, c, eig) bind(C, name="mhip_scf")
      import :: c_int, c_ptr, c_double, mhip_scf_opts, mhip_scf_result
      type(c_ptr), value :: ctx
      type(mhip_scf_opts) :: opts
      type(mhip_scf_result) :: res
      real(c_double) :: p(*), f(*), c(*), eig(*)
    end function
  end interface

  type(c_ptr), save :: ctx = c_null_ptr

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(mhip_last_error(), msg, [512])
    write (*, '(3a)', advance='no') 'libmopachip: ', where, ': '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      write (*, '(a)', advance='no') msg(i)
    end do
    write (*, *)
    stop 1
  end subroutine check

  subroutine hip_init(device)
    integer, intent(in), optional :: device
    integer(c_int) :: dev
    dev = 0
    if (present(device)) dev = device
    if (c_associated(ctx)) return
    call check(mhip_init(dev, ctx), 'mhip_init')
  end subroutine hip_init

  subroutine hip_finalize()
    if (.not. c_associated(ctx)) return
    call check(mhip_finalize(ctx), 'mhip_finalize')
    ctx = c_null_ptr
  end subroutine hip_finalize

  ! Make H and W resident (once per geometry, after hcore).
  subroutine hip_upload(numat, nfirst, nlast, norbs, n2elec, h, w)
    integer, intent(in) :: numat, norbs
    integer, intent(in) :: nfirst(numat), nlast(numat)
    integer(c_long_long), intent(in) :: n2elec
    double precision, intent(in) :: h(*), w(*)
    integer(c_int), allocatable :: nf(:), nl(:)
    call hip_init()
    allocate (nf(numat), nl(numat))
    nf = nfirst
    nl = nlast
    call check(mhip_upload_system(ctx, numat, nf, nl, norbs, n2elec, h, w), 'mhip_upload_system')
  end subroutine hip_upload

  subroutine hip_fock2(f, ptot, p, w, wj, wk, numat, nfirst, nlast, mode)
    double precision, intent(inout) :: f(*)
    double precision, intent(in) :: ptot(*), p(*), w(*), wj(*), wk(*)
    integer, intent(in) :: numat, nfirst(numat), nlast(numat), mode
    call check(mhip_fock2(ctx, f, ptot, p, int(mode, c_int)), 'mhip_fock2')
  end subroutine hip_fock2

  double precision function hip_helect(n, p, h, f)
    integer, intent(in) :: n
    double precision, intent(in) :: p(*), h(*), f(*)
    real(c_double) :: e
    call check(mhip_helect(ctx, int(n, c_int), p, h, f, e), 'mhip_helect')
    hip_helect = e
  end function hip_helect

  subroutine hip_eigenvectors(eigenvecs, xmat, eigvals, ndim)
    integer, intent(in) :: ndim
    double precision, intent(out) :: eigenvecs(ndim, ndim), eigvals(ndim)
    double precision, intent(inout) :: xmat(*)
    call check(mhip_eigen(ctx, eigenvecs, xmat, eigvals, int(ndim, c_int)), 'mhip_eigen')
  end subroutine hip_eigenvectors

  subroutine hip_densit(c, mdim, norbs, nocc, occ, nfract, fract, p, mode)
    integer, intent(in) :: mdim, norbs, nocc, nfract, mode
    double precision, intent(in) :: c(mdim, mdim), occ, fract
    double precision, intent(inout) :: p(*)
    if (nfract /= 0 .or. mdim /= norbs .or. mode /= 0) then
      write (*, *) 'hip_densit: only closed-shell densities with mdim = norbs, mode = 0'
      stop 1
    end if
    call check(mhip_density(ctx, c, int(nocc, c_int), occ, p), 'mhip_density')
  end subroutine hip_densit

  subroutine hip_diag(fao, vector, nocc, eig, mdim, n)
    integer, intent(in) :: nocc, mdim, n
    double precision, intent(in) :: fao(*), eig(*)
    double precision, intent(inout) :: vector(mdim, *)
    integer(c_long_long) :: nrot
    call check(mhip_pseudo_diag(ctx, fao, vector, int(nocc, c_int), eig, int(mdim, c_int), &
               int(n, c_int), nrot), 'mhip_pseudo_diag')
  end subroutine hip_diag

  ! The whole RHF SCF resident on the device (p, f: packed; c: n x n; eig: n).
  subroutine hip_iter(opts, res, p, f, c, eig)
    type(mhip_scf_opts), intent(in) :: opts
    type(mhip_scf_result), intent(out) :: res
    double precision, intent(inout) :: p(*), f(*), c(*), eig(*)
    type(mhip_scf_opts) :: o
    o = opts
    call check(mhip_scf(ctx, o, res, p, f, c, eig), 'mhip_scf')
  end subroutine hip_iter

end module mopac_hip
