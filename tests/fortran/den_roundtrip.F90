! Density restart files through the Fortran runtime — TEST INFRASTRUCTURE.
!
! Reads a density file with the statements MOPAC's den_in_out uses
! (src/SCF/iter.F90:1133-1227): `read (iden) old_norbs, old_numat, pa`
! (unformatted) or `read (iden, *) ...` (formatted), checks the header,
! replaces pa by pa/2 + 1/8, and writes it back with the matching write
! statement, so tests/test_den_io.py can check that the library's reader and
! writer (mhip_den_read, mhip_den_write) agree with the Fortran runtime both
! ways.
!
!   den_roundtrip u|f norbs numat infile outfile
program den_roundtrip
  implicit none
  character(len=512) :: arg, fin, fout
  character(len=8) :: form
  integer :: norbs, numat, old_norbs, old_numat, io_stat, mpack
  double precision, allocatable :: pa(:)
  call get_command_argument(1, form)
  call get_command_argument(2, arg)
  read (arg, *) norbs
  call get_command_argument(3, arg)
  read (arg, *) numat
  call get_command_argument(4, fin)
  call get_command_argument(5, fout)
  mpack = (norbs*(norbs + 1))/2
  allocate (pa(mpack))
  if (form(1:1) == 'u') then
    open (unit=21, file=fin, form='UNFORMATTED', status='OLD')
    read (21, iostat=io_stat) old_norbs, old_numat, pa
  else
    open (unit=21, file=fin, status='OLD')
    read (21, *, iostat=io_stat) old_norbs, old_numat, pa
  end if
  close (21)
  if (io_stat /= 0) stop 3
  if (old_norbs /= norbs .or. old_numat /= numat) stop 4
  pa = pa*0.5d0 + 0.125d0
  if (form(1:1) == 'u') then
    open (unit=22, file=fout, form='UNFORMATTED', status='REPLACE')
    write (22) norbs, numat, pa
  else
    open (unit=22, file=fout, status='REPLACE')
    write (22, *) norbs, numat, pa
  end if
  close (22)
end program den_roundtrip
