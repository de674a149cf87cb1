! Checker of the job-archive files with the Fortran runtime's own unformatted
! I/O, as the reference reads and writes them (acescore/aces_ja_init.F,
! aces_ja_fin.F, getrec.F; 8-byte integers, JOBARC records of 128 integers).
!   ja_check read <dir> <name>   print the record's length and integers
!   ja_check copy <dir> <out>    read JAINDX and rewrite it into <out>/JAINDX
program ja_check
  implicit none
  character(len=8) :: marker(1000)
  integer :: rloc(1000), rsize(1000), nrecs, i, k, r, w
  integer :: buf(128)
  character(len=256) :: mode, dir, arg
  call get_command_argument(1, mode)
  call get_command_argument(2, dir)
  call get_command_argument(3, arg)
  open(75, file=trim(dir)//'/JAINDX', form='UNFORMATTED', status='OLD')
  rewind(75)
  read(75) marker, rloc, rsize, nrecs
  close(75)
  if (trim(mode) == 'copy') then
    open(76, file=trim(arg)//'/JAINDX', form='UNFORMATTED', status='UNKNOWN')
    rewind(76)
    write(76) marker, rloc, rsize, nrecs
    close(76)
    stop
  end if
  open(75, file=trim(dir)//'/JOBARC', form='UNFORMATTED', access='DIRECT', status='OLD', recl=1024)
  do i = 1, 1000
    if (marker(i) == arg(1:8)) then
      print '(i0)', rsize(i)
      do k = 0, rsize(i) - 1
        w = rloc(i) - 1 + k
        r = 1 + w / 128
        read(75, rec=r) buf
        print '(i0)', buf(mod(w, 128) + 1)
      end do
    end if
  end do
  close(75)
end program ja_check
