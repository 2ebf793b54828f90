! ref_layout.f90 -- byte layout that flang gives the reference's own
! declarations of st_HACApK_leafmtx and st_HACApK_leafmtxp.
! TEST INFRASTRUCTURE ONLY: oracle/ref_build.sh preprocesses
! m_HACApK_base.f90 where it lies (with or without -DHAVE_MAGMA
! -DHAVE_MAGMA_BATCH -DISO_C_BINDING), puts the two type declarations into
! a temporary ref_types.inc and compiles this program against it; the
! printed offsets go to oracle/_ref/ref_layout_{magma,plain}.txt, which the
! CPU tests compare with include/hmv.h.
!
! What is printed (one "name value" per line):
!   leaf_*      offsets inside one st_HACApK_leafmtx record, as C reads it
!               (HACApK_MAGMA.h:12-20): the six ints, a1size, and the offset
!               of the word holding a1's base address (the first word of its
!               pointer descriptor, which C reads as double *a1)
!   leaf_stride loc(st_lf(2)) - loc(st_lf(1)), as m_HACApK_solve.f90:632-636
!               measures it
!   <field>     offsets of the container's fields; st_lf = offset of the word
!               holding st_lf's base address (C reads it as the pointer)
!   size        storage size of the container in bytes
module ref_layout_types
  use iso_c_binding
  implicit none
  include 'ref_types.inc'
end module ref_layout_types

program ref_layout
  use iso_c_binding
  use ref_layout_types
  implicit none
  type(st_HACApK_leafmtxp), target :: h
  integer(c_int64_t) :: base, lbase, want
  integer :: nwords

  base = addr_of(c_loc(h))
  allocate(h%st_lf(3))
  allocate(h%st_lf(1)%a1(5, 2), h%st_lf(1)%a2(4, 2))

  lbase = addr_of(c_loc(h%st_lf(1)))
  call show('leaf_ltmtx', addr_of(c_loc(h%st_lf(1)%ltmtx)) - lbase)
  call show('leaf_kt', addr_of(c_loc(h%st_lf(1)%kt)) - lbase)
  call show('leaf_nstrtl', addr_of(c_loc(h%st_lf(1)%nstrtl)) - lbase)
  call show('leaf_ndl', addr_of(c_loc(h%st_lf(1)%ndl)) - lbase)
  call show('leaf_nstrtt', addr_of(c_loc(h%st_lf(1)%nstrtt)) - lbase)
  call show('leaf_ndt', addr_of(c_loc(h%st_lf(1)%ndt)) - lbase)
  call show('leaf_a1size', addr_of(c_loc(h%st_lf(1)%a1size)) - lbase)
  call show('leaf_a1', word_holding(lbase, 64, addr_of(c_loc(h%st_lf(1)%a1(1, 1)))))
  call show('leaf_stride', addr_of(c_loc(h%st_lf(2))) - lbase)

  call show('nd', addr_of(c_loc(h%nd)) - base)
  call show('nlf', addr_of(c_loc(h%nlf)) - base)
  call show('nlfkt', addr_of(c_loc(h%nlfkt)) - base)
  call show('ktmax', addr_of(c_loc(h%ktmax)) - base)
  call show('st_lf_stride', addr_of(c_loc(h%st_lf_stride)) - base)
#ifdef REF_MAGMA
  call show('m', addr_of(c_loc(h%m)) - base)
  call show('max_block', addr_of(c_loc(h%max_block)) - base)
  call show('gflops', addr_of(c_loc(h%gflops)) - base)
  call show('mtx1_gpu', addr_of(c_loc(h%mtx1_gpu)) - base)
  call show('zbu_gpu', addr_of(c_loc(h%zbu_gpu)) - base)
  call show('num_batch', addr_of(c_loc(h%num_batch)) - base)
  call show('transA', addr_of(c_loc(h%transA)) - base)
  call show('d_A_array', addr_of(c_loc(h%d_A_array)) - base)
  call show('batch_size', addr_of(c_loc(h%batch_size)) - base)
  call show('num_streamed', addr_of(c_loc(h%num_streamed)) - base)
  call show('num_streamed_t', addr_of(c_loc(h%num_streamed_t)) - base)
  call show('h_A_array_streamed', addr_of(c_loc(h%h_A_array_streamed)) - base)
  call show('h_lda_streamed', addr_of(c_loc(h%h_lda_streamed)) - base)
  call show('d_A_mgpu', addr_of(c_loc(h%d_A_mgpu)) - base)
  call show('iwork', addr_of(c_loc(h%iwork)) - base)
  call show('mpi_rank', addr_of(c_loc(h%mpi_rank)) - base)
#endif
  nwords = int(storage_size(h) / 64)
  want = addr_of(c_loc(h%st_lf(1)))
  call show('st_lf', word_holding(base, nwords, want))
  call show('size', int(storage_size(h) / 8, c_int64_t))

contains

  integer(c_int64_t) function addr_of(p)
    type(c_ptr), intent(in) :: p
    addr_of = transfer(p, addr_of)
  end function addr_of

  ! Byte offset (from `from`) of the first of nwords 8-byte words that holds
  ! the value `want`, or -1.
  integer(c_int64_t) function word_holding(from, nwords, want)
    integer(c_int64_t), intent(in) :: from, want
    integer, intent(in) :: nwords
    integer(c_int64_t), pointer :: w(:)
    integer :: i
    call c_f_pointer(transfer(from, c_null_ptr), w, [nwords])
    word_holding = -1
    do i = 1, nwords
      if (w(i) == want) then
        word_holding = 8 * (i - 1)
        return
      end if
    end do
  end function word_holding

  subroutine show(name, v)
    character(*), intent(in) :: name
    integer(c_int64_t), intent(in) :: v
    write(*, '(a,1x,i0)') name, v
  end subroutine show
end program ref_layout
