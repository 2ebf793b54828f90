! ref_driver.f90 -- a Fortran caller of libhmv.so through the reference's
! own calling sequence, with the reference's own declarations of
! st_HACApK_leafmtxp / st_HACApK_leafmtx / st_HACApK_lcontrol as a
! HAVE_MAGMA + HAVE_MAGMA_BATCH + ISO_C_BINDING build compiles them.
! TEST INFRASTRUCTURE ONLY: oracle/ref_build.sh compiles it against the
! type declarations it extracts from m_HACApK_base.f90 (temporary
! ref_types.inc) and links it with libhmv.so; tests/test_gpu_ref_driver.py
! runs it.
!
! Steps, each as the reference's Fortran does it:
!   magma_init                                 m_HACApK_use.f90:61
!   leaves allocated one by one, a1(ndt,kt) and a2(ndl,kt) or a1(ndt,ndl)
!                                              m_HACApK_base.f90:1021,1031
!   st_lf_stride and every a1size measured     m_HACApK_solve.f90:632-650
!   c_HACApK_adot_body_lfcpy_batch_sorted      m_HACApK_solve.f90:711
!   c_HACApK_adot_body_lfmtx_batch             m_HACApK_solve.f90:712
!   c_HACApK_bicgstab_cax_lfmtx_gpu            m_HACApK_use.f90:443
!   c_HACApK_adot_body_lfdel_batch             m_HACApK_use.f90:464
!   c_HACApK_adot_body_lfmtx (binds its own arena on first use)
!                                              m_HACApK_solve.f90:664
!   c_HACApK_adot_body_lfdel_batch, magma_finalize   m_HACApK_use.f90:74
! The MAGMA block of the container is filled with sentinel values first;
! the driver reports whether anything but mpi_rank changed.
!
! Input file (stream, little endian): int32 nd, nlf, mstep; real64 eps,
! zau0; per leaf int32 ltmtx, kt, nstrtl, ndl, nstrtt, ndt, then a1 (and
! a2) column-major; then zu(nd), b(nd).
! Output file: real64 zau(nd) (zau0 + H zu, lfmtx_batch), zau2(nd) (zau0 +
! H zu, lfmtx), u(nd) (BiCGSTAB from u = 0); int32 nstp, lrtrn, mpi_rank,
! sentinels_kept (1/0), stride, 0.
module ref_driver_types
  use iso_c_binding
  implicit none
  include 'ref_types.inc'
end module ref_driver_types

program ref_driver
  use iso_c_binding
  use ref_driver_types
  implicit none
  type(st_HACApK_leafmtxp), target :: st_leafmtxp
  type(st_HACApK_lcontrol), target :: st_ctl
  integer(c_int32_t) :: nd, nlf, mstep, hdr(6), nstp, lrtrn, kept
  real(c_double) :: eps, zau0, tb, ts, tc
  real(c_double), allocatable, target :: zu(:), b(:), zau(:), zau2(:), u(:), wws(:)
  integer :: ip, narg
  character(len=4096) :: fin, fout
  integer(c_int64_t) :: a1pt, a2pt
  integer(c_int64_t), parameter :: sentinel = int(z'5EA7BEEF0BADF00D', c_int64_t)

  narg = command_argument_count()
  if (narg /= 2) then
    write(*, '(a)') 'usage: ref_driver <case.bin> <out.bin>'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  call magma_init()

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nd, nlf, mstep, eps, zau0
  st_leafmtxp%nd = nd
  st_leafmtxp%nlf = nlf
  st_leafmtxp%nlfkt = 0
  st_leafmtxp%ktmax = 0
  allocate(st_leafmtxp%st_lf(nlf))
  do ip = 1, nlf
    read(10) hdr
    st_leafmtxp%st_lf(ip)%ltmtx = hdr(1)
    st_leafmtxp%st_lf(ip)%kt = hdr(2)
    st_leafmtxp%st_lf(ip)%nstrtl = hdr(3)
    st_leafmtxp%st_lf(ip)%ndl = hdr(4)
    st_leafmtxp%st_lf(ip)%nstrtt = hdr(5)
    st_leafmtxp%st_lf(ip)%ndt = hdr(6)
    if (hdr(1) == 1) then
      allocate(st_leafmtxp%st_lf(ip)%a1(hdr(6), hdr(2)), st_leafmtxp%st_lf(ip)%a2(hdr(4), hdr(2)))
      read(10) st_leafmtxp%st_lf(ip)%a1, st_leafmtxp%st_lf(ip)%a2
      st_leafmtxp%nlfkt = st_leafmtxp%nlfkt + 1
      st_leafmtxp%ktmax = max(st_leafmtxp%ktmax, hdr(2))
    else
      allocate(st_leafmtxp%st_lf(ip)%a1(hdr(6), hdr(4)))
      read(10) st_leafmtxp%st_lf(ip)%a1
    end if
  end do
  allocate(zu(nd), b(nd), zau(nd), zau2(nd), u(nd), wws(nd))
  read(10) zu, b
  close(10)

  ! st_lf_stride and a1size as HACApK_measurez_time_ax_FPGA_lfmtx sets them
  st_leafmtxp%st_lf_stride = int(addr_of(c_loc(st_leafmtxp%st_lf(2))) - &
                                 addr_of(c_loc(st_leafmtxp%st_lf(1))), c_int)
  do ip = 1, nlf
    if (st_leafmtxp%st_lf(ip)%ltmtx == 1) then
      a1pt = addr_of(c_loc(st_leafmtxp%st_lf(ip)%a1))
      a2pt = addr_of(c_loc(st_leafmtxp%st_lf(ip)%a2))
      st_leafmtxp%st_lf(ip)%a1size = a2pt - a1pt
    end if
  end do

  call fill_magma_block(st_leafmtxp)

  allocate(st_ctl%param(100), st_ctl%lpmd(30))
  st_ctl%param(:) = 0.0d0
  st_ctl%param(83) = mstep
  st_ctl%param(91) = eps
  st_ctl%lpmd(:) = 0
  st_ctl%lpmd(2) = 1
  st_ctl%lpmd_offset = int(addr_of(c_loc(st_ctl%lpmd(1))) - addr_of(c_loc(st_ctl%param(1))), c_int)

  call c_HACApK_adot_body_lfcpy_batch_sorted(nd, st_leafmtxp)
  zau(:) = zau0
  tb = 0; ts = 0; tc = 0
  call c_HACApK_adot_body_lfmtx_batch(zau, st_leafmtxp, zu, wws, tb, ts, tc)
  u(:) = 0.0d0
  nstp = -1
  lrtrn = -99
  call c_HACApK_bicgstab_cax_lfmtx_gpu(st_leafmtxp, st_ctl, u, b, st_ctl%param, nd, nstp, lrtrn)
  call c_HACApK_adot_body_lfdel_batch(st_leafmtxp)
  zau2(:) = zau0
  call c_HACApK_adot_body_lfmtx(zau2, st_leafmtxp, zu, wws)
  call c_HACApK_adot_body_lfdel_batch(st_leafmtxp)
  kept = 0
  if (magma_block_kept(st_leafmtxp)) kept = 1
  call magma_finalize()

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) zau, zau2, u
  write(11) nstp, lrtrn, int(st_leafmtxp%mpi_rank, c_int32_t), kept, st_leafmtxp%st_lf_stride, 0_c_int32_t
  close(11)
  write(*, '(a,i0,a,i0,a,i0,a,i0,a,i0)') 'ref_driver: nd=', nd, ' nlf=', nlf, ' stride=', &
      st_leafmtxp%st_lf_stride, ' nstp=', nstp, ' lrtrn=', lrtrn

contains

  integer(c_int64_t) function addr_of(p)
    type(c_ptr), intent(in) :: p
    addr_of = transfer(p, addr_of)
  end function addr_of

  type(c_ptr) function sentinel_ptr(i)
    integer, intent(in) :: i
    sentinel_ptr = transfer(sentinel + i, c_null_ptr)
  end function sentinel_ptr

  ! Every field of the MAGMA block gets a value the engine must not touch
  ! (mpi_rank -7: the engine writes 0 there, as the reference's lfcpy does).
  subroutine fill_magma_block(h)
    type(st_HACApK_leafmtxp), intent(inout) :: h
    h%m = -101; h%n = -102; h%gn = -103; h%max_block = -104
    h%gflops = sentinel
    h%mtx1_gpu = sentinel_ptr(1); h%mtx2_gpu = sentinel_ptr(2); h%zu_gpu = sentinel_ptr(3)
    h%zau_gpu = sentinel_ptr(4); h%zau_pin = sentinel_ptr(5); h%zbu_gpu = sentinel_ptr(6)
    h%num_batch = -105; h%total_size_y = -106; h%transA = -107
    h%d_A_array = sentinel_ptr(7); h%d_X_array = sentinel_ptr(8); h%d_Y_array = sentinel_ptr(9)
    h%d_M = sentinel_ptr(10); h%d_N = sentinel_ptr(11); h%d_lda = sentinel_ptr(12)
    h%d_inc = sentinel_ptr(13); h%batch_order = sentinel_ptr(14); h%h_A_array = sentinel_ptr(15)
    h%h_X_array = sentinel_ptr(16); h%h_Y_array = sentinel_ptr(17); h%h_type = sentinel_ptr(18)
    h%h_I = sentinel_ptr(19); h%h_J = sentinel_ptr(20); h%h_M = sentinel_ptr(21)
    h%h_N = sentinel_ptr(22); h%h_lda = sentinel_ptr(23); h%max_M = sentinel_ptr(24)
    h%max_N = sentinel_ptr(25); h%batch_size = sentinel_ptr(26)
    h%num_streamed = -108; h%num_streamed_t = -109
    h%h_A_array_streamed = sentinel_ptr(27); h%h_X_array_streamed = sentinel_ptr(28)
    h%h_Y_array_streamed = sentinel_ptr(29); h%h_M_streamed = sentinel_ptr(30)
    h%h_N_streamed = sentinel_ptr(31); h%h_lda_streamed = sentinel_ptr(32)
    h%d_A_mgpu = sentinel_ptr(33); h%d_X_mgpu = sentinel_ptr(34); h%d_Y_mgpu = sentinel_ptr(35)
    h%d_M_mgpu = sentinel_ptr(36); h%d_N_mgpu = sentinel_ptr(37); h%d_lda_mgpu = sentinel_ptr(38)
    h%d_inc_mgpu = sentinel_ptr(39); h%zu_mgpu = sentinel_ptr(40); h%zau_mgpu = sentinel_ptr(41)
    h%zbu_mgpu = sentinel_ptr(42); h%h_A_mgpu = sentinel_ptr(43); h%h_X_mgpu = sentinel_ptr(44)
    h%h_Y_mgpu = sentinel_ptr(45); h%h_type_mgpu = sentinel_ptr(46); h%h_I_mgpu = sentinel_ptr(47)
    h%h_J_mgpu = sentinel_ptr(48); h%h_M_mgpu = sentinel_ptr(49); h%h_N_mgpu = sentinel_ptr(50)
    h%h_lda_mgpu = sentinel_ptr(51); h%max_M_mgpu = sentinel_ptr(52)
    h%max_N_mgpu = sentinel_ptr(53); h%nlf_mgpu = sentinel_ptr(54)
    h%num_batch_mgpu = sentinel_ptr(55); h%total_size_y_mgpu = sentinel_ptr(56)
    h%iwork = sentinel_ptr(57)
    h%mpi_rank = -7
  end subroutine fill_magma_block

  logical function same_ptr(p, i)
    type(c_ptr), intent(in) :: p
    integer, intent(in) :: i
    same_ptr = addr_of(p) == sentinel + i
  end function same_ptr

  logical function magma_block_kept(h)
    type(st_HACApK_leafmtxp), intent(in) :: h
    magma_block_kept = h%m == -101 .and. h%n == -102 .and. h%gn == -103 .and. &
        h%max_block == -104 .and. h%gflops == sentinel .and. &
        h%num_batch == -105 .and. h%total_size_y == -106 .and. h%transA == -107 .and. &
        h%num_streamed == -108 .and. h%num_streamed_t == -109 .and. &
        same_ptr(h%mtx1_gpu, 1) .and. same_ptr(h%mtx2_gpu, 2) .and. same_ptr(h%zu_gpu, 3) .and. &
        same_ptr(h%zau_gpu, 4) .and. same_ptr(h%zau_pin, 5) .and. same_ptr(h%zbu_gpu, 6) .and. &
        same_ptr(h%d_A_array, 7) .and. same_ptr(h%d_X_array, 8) .and. same_ptr(h%d_Y_array, 9) .and. &
        same_ptr(h%d_M, 10) .and. same_ptr(h%d_N, 11) .and. same_ptr(h%d_lda, 12) .and. &
        same_ptr(h%d_inc, 13) .and. same_ptr(h%batch_order, 14) .and. same_ptr(h%h_A_array, 15) .and. &
        same_ptr(h%h_X_array, 16) .and. same_ptr(h%h_Y_array, 17) .and. same_ptr(h%h_type, 18) .and. &
        same_ptr(h%h_I, 19) .and. same_ptr(h%h_J, 20) .and. same_ptr(h%h_M, 21) .and. &
        same_ptr(h%h_N, 22) .and. same_ptr(h%h_lda, 23) .and. same_ptr(h%max_M, 24) .and. &
        same_ptr(h%max_N, 25) .and. same_ptr(h%batch_size, 26) .and. &
        same_ptr(h%h_A_array_streamed, 27) .and. same_ptr(h%h_X_array_streamed, 28) .and. &
        same_ptr(h%h_Y_array_streamed, 29) .and. same_ptr(h%h_M_streamed, 30) .and. &
        same_ptr(h%h_N_streamed, 31) .and. same_ptr(h%h_lda_streamed, 32) .and. &
        same_ptr(h%d_A_mgpu, 33) .and. same_ptr(h%d_X_mgpu, 34) .and. same_ptr(h%d_Y_mgpu, 35) .and. &
        same_ptr(h%d_M_mgpu, 36) .and. same_ptr(h%d_N_mgpu, 37) .and. same_ptr(h%d_lda_mgpu, 38) .and. &
        same_ptr(h%d_inc_mgpu, 39) .and. same_ptr(h%zu_mgpu, 40) .and. same_ptr(h%zau_mgpu, 41) .and. &
        same_ptr(h%zbu_mgpu, 42) .and. same_ptr(h%h_A_mgpu, 43) .and. same_ptr(h%h_X_mgpu, 44) .and. &
        same_ptr(h%h_Y_mgpu, 45) .and. same_ptr(h%h_type_mgpu, 46) .and. same_ptr(h%h_I_mgpu, 47) .and. &
        same_ptr(h%h_J_mgpu, 48) .and. same_ptr(h%h_M_mgpu, 49) .and. same_ptr(h%h_N_mgpu, 50) .and. &
        same_ptr(h%h_lda_mgpu, 51) .and. same_ptr(h%max_M_mgpu, 52) .and. &
        same_ptr(h%max_N_mgpu, 53) .and. same_ptr(h%nlf_mgpu, 54) .and. &
        same_ptr(h%num_batch_mgpu, 55) .and. same_ptr(h%total_size_y_mgpu, 56) .and. &
        same_ptr(h%iwork, 57)
  end function magma_block_kept
end program ref_driver
