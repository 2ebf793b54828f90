! ref_aca_bind.f90 -- bind(C) entry point onto the reference's ACA+, for tests.
! TEST INFRASTRUCTURE ONLY: compiled by oracle/ref_build.sh together with the
! reference's own m_ppohBEM_*.f90, m_HACApK_calc_entry_ij.f90, HACApK_lib.f90
! and the module ref_aca_mod, which ref_build.sh cuts out of the preprocessed
! m_HACApK_base.f90 into a temporary directory: the module's header and
! interface block, HACApK_calc_vec (m_HACApK_base.f90:677-699) and
! HACApK_acaplus (:783-975), unmodified.  The routine below only sets up the
! entry data the way the reference's driver hands a mesh to HACApK
! (bem-bb-fw-HACApK-0.4.2.f90:229-231: node coordinates, 0-based face2node,
! three nodes per face) and the ACA arguments the way HACApK_fill_leafmtx_hyp
! passes them (m_HACApK_base.f90:987, 1008: eps = param(71), ACA_EPS =
! param(72)*eps, kmax = param(63), znrmmat unused with param(61) = 1), then
! forwards to HACApK_acaplus; nothing here reimplements reference behaviour.
module ref_aca_bind
  use iso_c_binding
  use m_ppohBEM_matrix_element_ij, only: coordinate
  use m_HACApK_calc_entry_ij, only: st_HACApK_calc_entry
  use ref_aca_mod, only: HACApK_acaplus
  implicit none
contains
  ! ACA+ of the block rows nstrtl..nstrtl+ndl-1 x columns nstrtt..nstrtt+ndt-1
  ! (1-based, cluster order; lod(i) = original 1-based face of position i) of
  ! the collocation matrix of faces(3, nface) over nodes(3, nnode).
  ! zaa(ndl, kmax) and zab(ndt, kmax) receive the factors (the leaf's a2 and
  ! a1); returns the rank kt.
  integer(c_int) function ref_acaplus(nnode, nodes, nface, faces, lod, ndl, ndt, nstrtl, nstrtt, &
                                      param, zaa, zab) bind(C, name="ref_acaplus")
    integer(c_int), value :: nnode, nface, ndl, ndt, nstrtl, nstrtt
    real(c_double), intent(in) :: nodes(3, nnode)
    integer(c_int), intent(in) :: faces(3, nface)
    integer(c_int), intent(in) :: lod(nface)
    real(c_double), intent(in) :: param(100)
    real(c_double), intent(out) :: zaa(ndl, *), zab(ndt, *)
    type(st_HACApK_calc_entry) :: st
    real*8 :: prm(100), eps, aca_eps
    integer*4 :: lodf(nface), kmax
    integer :: i
    allocate(st%np(nnode), st%face2node(3, nface), st%int_para_fc(1, nface), &
             st%dble_para_fc(1, nface))
    do i = 1, nnode
      st%np(i)%x = nodes(1, i)
      st%np(i)%y = nodes(2, i)
      st%np(i)%z = nodes(3, i)
    end do
    st%face2node = faces
    st%int_para_fc = 0
    st%dble_para_fc = 0d0
    st%nond = nnode
    st%nofc = nface
    st%nond_on_face = 3
    st%nint_para_fc = 1
    st%ndble_para_fc = 1
    st%nd = nface
    st%lp61 = 0
    prm = param
    lodf = lod
    eps = prm(71)
    aca_eps = prm(72) * eps
    kmax = int(prm(63))
    ref_acaplus = aca(zaa, zab)
    deallocate(st%np, st%face2node, st%int_para_fc, st%dble_para_fc)
  contains
    integer function aca(za, zb)
      real(c_double) :: za(ndl, kmax), zb(ndt, kmax)
      aca = HACApK_acaplus(za, zb, prm, ndl, ndt, nstrtl, nstrtt, lodf, st, kmax, eps, 0d0, aca_eps)
    end function
  end function
end module ref_aca_bind
