! Closed-shell CCSD with the call sequence of ACES II's CC loop (vcc/vcc.F:798-1437),
! linked against the MI355X library.  Compile with -fdefault-integer-8 like the
! reference so INTEGER arguments are 8 bytes.
!
! Usage: xvcc_amd <input> [cc_conv] [cc_maxcyc] [cc_exporder] [mode]
! <input> a file in the stream format below, or "jobdir:<dir>" for a job
!         directory holding JAINDX, JOBARC and the list files (MOINTS, MOABCD,
!         ...) as the reference's xintprc leaves them: the archive and lists are
!         read (ACES_JA_INIT / ACES_IO_INIT), SCFEVALA comes from the archive, and
!         on convergence TOTENERG = SCFENEG + the CC correlation energy is
!         stored (vcc/finish.F:140) and the archive and lists written back
!         (ACES_IO_FIN / ACES_JA_FIN).
! mode 'lambda': after convergence move the amplitudes into the Lambda start lists
!               with DRMOVE(...,100,SING1) (vcc/vcc.F:1138) and print the sums of
!               squares of lists 44, 46, 90 and 144, 146, 190;
! mode 'xlambda_diis': call DODIIS0 with xlambda's list set (lambda/vlamcc.F:812-814),
!               which the library does not implement (it must end with status 1).
! Input (stream, little endian, int64 integers), C1: o, v, nlists; per list:
! list, typel, typer, nrows, ncols and nrows*ncols doubles (column major); then
! the n = o + v orbital energies (record SCFEVALA).  Symmetry-packed lists:
! -nirrep, pop(1:nirrep), vrt(1:nirrep), nlists; per list: list, typel, typer,
! ngroups and per group: group, nrows, ncols and the doubles; then the orbital
! energies (occupied then virtual, irrep-sorted).
program xvcc_amd
  use aces2_amd
  implicit none
  integer :: icore(1), maxcor, iuhf, inttyp, icycle, ncycle, iconvg, icontl, iorder
  integer :: o, v, nl, il, lst, tl, tr, nr, nc, n, nirrep, ng, ig, g, h
  integer :: popi(8), vrti(8)
  logical :: sing1, do_hbar
  double precision :: ecorr(3), etot, etott2, edummy, damp, escf, etotal
  double precision, allocatable :: buf(:)
  character(len=256) :: fname, arg, mode, jobdir
  integer(c_int), target :: pop(8), vrt(8)

  call get_command_argument(1, fname)
  icontl = 7
  ncycle = 50
  iorder = 6
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) icontl
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) ncycle
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) iorder
  end if
  mode = ''
  if (command_argument_count() >= 5) call get_command_argument(5, mode)
  maxcor = 1
  iuhf = 0
  inttyp = 2
  sing1 = .true.
  do_hbar = .false.
  damp = 1.0d0

  call aces_check(aces_init(c_null_ptr), 'aces_init')
  jobdir = ''
  if (len_trim(fname) > 7) then
    if (fname(1:7) == 'jobdir:') jobdir = fname(8:)
  end if
  if (len_trim(jobdir) > 0) then
    call aces_check(aces_jobarc_read(trim(jobdir)//c_null_char), 'ACES_JA_INIT')
    call aces_check(aces_lists_read(trim(jobdir)//c_null_char), 'ACES_IO_INIT')
    call getrec(1, 'JOBARC', 'COMPNIRR', 1, nirrep)
    call getrec(1, 'JOBARC', 'SYMPOPOA', nirrep, popi)
    call getrec(1, 'JOBARC', 'SYMPOPVA', nirrep, vrti)
    o = sum(popi(1:nirrep))
    v = sum(vrti(1:nirrep))
  else
  open(10, file=trim(fname), access='stream', form='unformatted', status='old')
  read(10) o
  pop = 0
  vrt = 0
  if (o < 0) then
    nirrep = -o
    read(10) (popi(h), h = 1, nirrep), (vrti(h), h = 1, nirrep), nl
    pop(1:nirrep) = int(popi(1:nirrep), c_int)
    vrt(1:nirrep) = int(vrti(1:nirrep), c_int)
    o = sum(popi(1:nirrep))
    v = sum(vrti(1:nirrep))
  else
    nirrep = 1
    read(10) v, nl
    pop(1) = int(o, c_int)
    vrt(1) = int(v, c_int)
  end if
  call aces_check(aces_set_symmetry(int(nirrep, c_int), c_loc(pop), c_null_ptr, c_loc(vrt), c_null_ptr), &
                  'symmetry')
  do il = 1, nl
    if (nirrep == 1) then
      read(10) lst, tl, tr, nr, nc
      allocate(buf(nr * nc))
      read(10) buf
      call inipck(1, tl, tr, lst, 0, 0, 1)
      call putlst(buf, 1, nc, 1, 1, lst)
      deallocate(buf)
    else
      read(10) lst, tl, tr, ng
      call inipck(1, tl, tr, lst, 0, 0, 1)
      do ig = 1, ng
        read(10) g, nr, nc
        if (nr * nc > 0) then
          allocate(buf(nr * nc))
          read(10) buf
          call putlst(buf, 1, nc, 1, g, lst)
          deallocate(buf)
        end if
      end do
    end if
  end do
  n = o + v
  allocate(buf(n))
  read(10) buf
  close(10)
  call putrec(20, 'JOBARC', 'SCFEVALA', n, buf)
  deallocate(buf)
  end if

  ! MP1 start (xintprc) and DIIS lists (DIISLST)
  call aces_check(aces_vcc_mp1(), 'MP1')
  call aces_check(aces_vcc_set_diis_order(int(iorder, c_int)), 'DIISLST')
  call cmpeng(icore, maxcor, 43, 0, ecorr, etot, etott2, iuhf, 1)

  iconvg = 1
  do icycle = 1, ncycle
    call initin(icore, maxcor, iuhf)
    call initsn(icore, maxcor, iuhf)
    call genint(icore, maxcor, iuhf, inttyp, do_hbar)
    call feacont(icore, maxcor, iuhf)
    call fmicont(icore, maxcor, iuhf)
    call fmecont(icore, maxcor, iuhf, 1)
    call t1int2a(icore, maxcor, iuhf)
    call t1int2b(icore, maxcor, iuhf)
    call t1int1(icore, maxcor, iuhf, 1)
    call dre3en(icore, maxcor, iuhf, 0)
    call e4s(icore, maxcor, iuhf, edummy)
    call newt2(icore, maxcor, iuhf)
    call cmpeng(icore, maxcor, 60, 2, ecorr, etot, etott2, iuhf, 1)
    if (trim(mode) == 'xlambda_diis') then
      call dodiis0(icore, maxcor, iuhf, 1, icycle, iconvg, icontl, sing1, 144, 61, 190, 0, 90, 2, 70, &
                   '     ', damp)
    end if
    call dodiis0(icore, maxcor, iuhf, 1, icycle, iconvg, icontl, sing1, 44, 61, 90, 0, 90, 2, 70, &
                 '     ', damp)
    if (iconvg /= 0) then
      call drmove(icore, maxcor, iuhf, 0, sing1)
      call rnabij(icore, maxcor, iuhf, 'T')
    else
      call cmpeng(icore, maxcor, 43, 0, ecorr, etot, etott2, iuhf, 1)
      write(6, '(a)') '  The CC iterations have converged.'
      if (trim(mode) == 'lambda' .and. nirrep == 1) then
        call drmove(icore, maxcor, iuhf, 100, sing1)
        call print_norms()
      end if
      exit
    end if
  end do
  if (len_trim(jobdir) > 0 .and. iconvg == 0) then
    escf = 0.0d0
    call getrec(-1, 'JOBARC', 'SCFENEG ', 1, escf)
    etotal = etot + escf
    call putrec(1, 'JOBARC', 'TOTENERG', 1, etotal)
    call aces_check(aces_lists_write(trim(jobdir)//c_null_char), 'ACES_IO_FIN')
    call aces_check(aces_jobarc_write(trim(jobdir)//c_null_char), 'ACES_JA_FIN')
  end if
  call aces_check(aces_fin(), 'aces_fin')
  if (iconvg /= 0) then
    write(6, '(a)') '@VCC: The Coupled-Cluster equations did not converge!'
    stop 1
  end if

contains

  ! Sum of squares of one list (group 1) read back through GETLST.
  subroutine list_norm(lst, nrow, ncol)
    integer, intent(in) :: lst, nrow, ncol
    double precision, allocatable :: w(:)
    allocate(w(nrow * ncol))
    call getlst(w, 1, ncol, 1, 1, lst)
    write(6, '(a,i4,a,es24.16)') '  List ', lst, ' sum of squares ', sum(w * w)
    deallocate(w)
  end subroutine list_norm

  subroutine print_norms()
    call list_norm(44, v * (v - 1) / 2, o * (o - 1) / 2)
    call list_norm(46, v * v, o * o)
    call list_norm(90, o * v, 1)
    call list_norm(144, v * (v - 1) / 2, o * (o - 1) / 2)
    call list_norm(146, v * v, o * o)
    call list_norm(190, o * v, 1)
  end subroutine print_norms
end program xvcc_amd
