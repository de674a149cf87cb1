! ISO_C_BINDING interface to the MI355X CCSD amplitude library (include/aces2_amd.h).
!
! A Fortran host (ACES II is built with -fdefault-integer-8) can either call
! the reference-named entry points directly (getlst, putlst, genint, dre3en,
! e4s, newt2, cmpeng, dodiis0, drmove, ... -- exported with the reference's
! argument lists), or use this module's C-named interfaces.
module aces2_amd
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: ACES_PREC_F64 = 0, ACES_PREC_F32 = 1, ACES_PREC_BF16X3 = 2

  interface
    integer(c_int) function aces_init(stream) bind(C, name='aces_init')
      import :: c_int, c_ptr
      type(c_ptr), value :: stream
    end function
    integer(c_int) function aces_fin() bind(C, name='aces_fin')
      import :: c_int
    end function
    integer(c_int) function aces_set_symmetry(nirrep, popa, popb, vrta, vrtb) bind(C, name='aces_set_symmetry')
      import :: c_int, c_ptr
      integer(c_int), value :: nirrep
      type(c_ptr), value :: popa, popb, vrta, vrtb
    end function
    integer(c_long) function aces_irpdpd(irrep, typ) bind(C, name='aces_irpdpd')
      import :: c_int, c_long
      integer(c_int), value :: irrep, typ
    end function
    integer(c_int) function aces_inipck(irrepx, typel, typer, list) bind(C, name='aces_inipck')
      import :: c_int
      integer(c_int), value :: irrepx, typel, typer, list
    end function
    integer(c_int) function aces_putlst(src, ifirst, ncols, group, list, on_dev) bind(C, name='aces_putlst')
      import :: c_int, c_long, c_double
      real(c_double), intent(in) :: src(*)
      integer(c_long), value :: ifirst, ncols
      integer(c_int), value :: group, list, on_dev
    end function
    integer(c_int) function aces_getlst(dest, ifirst, ncols, group, list, on_dev) bind(C, name='aces_getlst')
      import :: c_int, c_long, c_double
      real(c_double), intent(out) :: dest(*)
      integer(c_long), value :: ifirst, ncols
      integer(c_int), value :: group, list, on_dev
    end function
    integer(c_int) function aces_putrec(name, nwords, src) bind(C, name='aces_putrec')
      import :: c_int, c_long, c_char, c_double
      character(kind=c_char), intent(in) :: name(*)
      integer(c_long), value :: nwords
      real(c_double), intent(in) :: src(*)
    end function
    integer(c_int) function aces_vcc_mp1() bind(C, name='aces_vcc_mp1')
      import :: c_int
    end function
    integer(c_int) function aces_vcc_set_diis_order(order) bind(C, name='aces_vcc_set_diis_order')
      import :: c_int
      integer(c_int), value :: order
    end function
    integer(c_int) function aces_vcc_set_precision(prec) bind(C, name='aces_vcc_set_precision')
      import :: c_int
      integer(c_int), value :: prec
    end function
    integer(c_int) function aces_jobarc_read(dir) bind(C, name='aces_jobarc_read')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: dir(*)
    end function
    integer(c_int) function aces_jobarc_write(dir) bind(C, name='aces_jobarc_write')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: dir(*)
    end function
    integer(c_int) function aces_lists_read(dir) bind(C, name='aces_lists_read')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: dir(*)
    end function
    integer(c_int) function aces_lists_write(dir) bind(C, name='aces_lists_write')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: dir(*)
    end function
    type(c_ptr) function aces_last_error() bind(C, name='aces_last_error')
      import :: c_ptr
    end function
  end interface

contains

  ! Fatal path of the reference (aces_exit(1)) for a nonzero status.
  subroutine aces_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= 0) then
      write(0, '(a,a,i4)') what, ': failed with status ', status
      stop 1
    end if
  end subroutine aces_check

end module aces2_amd
