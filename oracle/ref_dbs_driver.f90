!======================================================================
!  ref_dbs_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
!  C-callable entry points over the reference's own Dirac B-spline
!  routines (src/libs/DBS), which ref.mk compiles from the reference
!  checkout into oracle/_ref/libdbs_ref.so.  This file holds no reference
!  code: it sets the grid module (DBS_grid) and calls alloc_DBS_gauss,
!  the moment routines and mrk_pppp / mrk_qqqq / mrk_pqpq / mrk_qpqp, so
!  the tests can compare the C restatement with them.
!======================================================================

subroutine ref_dbs_init(ks_, ns_, ksp_, ksq_, tin) bind(C, name="ref_dbs_init")
  use iso_c_binding
  use DBS_grid
  use DBS_gauss
  use DBS_integrals
  implicit none
  integer(c_int), value :: ks_, ns_, ksp_, ksq_
  real(c_double) :: tin(ns_+ks_)
  ks = ks_; ns = ns_; ksp = ksp_; ksq = ksq_
  nv = ns - ks + 1; ms = ns + ns; ml = 0; me = 0
  nsp = nv + ksp - 1; nsq = nv + ksq - 1
  if (allocated(t)) deallocate(t)
  allocate(t(ns+ks)); t = tin
  call alloc_DBS_gauss
  call alloc_Rk_integral(ns, ks)
end subroutine

subroutine ref_dbs_tables(gr_o, grw_o, pb_o, qb_o) bind(C, name="ref_dbs_tables")
  use iso_c_binding
  use DBS_grid
  use DBS_gauss
  implicit none
  real(c_double) :: gr_o(nv,ks), grw_o(nv,ks), pb_o(nv+1,ks,ksp), qb_o(nv+1,ks,ksq)
  gr_o = gr; grw_o = grw
  pb_o = pbsp(:,:,1:ksp); qb_o = qbsp(:,:,1:ksq)
end subroutine

subroutine ref_dbs_moments(q, k, rkm) bind(C, name="ref_dbs_moments")
  use iso_c_binding
  use DBS_grid
  implicit none
  integer(c_int), value :: q, k
  real(c_double) :: rkm(ks*ks, nv)
  real(8), allocatable :: r(:,:)
  allocate(r(ks*ks, nv))
  if (q == 0) then
    call moments_pp(k, ks*ks, nv, r)
  else
    call moments_qq(k, ks*ks, nv, r)
  end if
  rkm = r
end subroutine

subroutine ref_dbs_mrk(family, k, rkb_o) bind(C, name="ref_dbs_mrk")
  use iso_c_binding
  use DBS_grid
  use DBS_integrals
  implicit none
  integer(c_int), value :: family, k
  real(c_double) :: rkb_o(ns, ns, ks, ks)
  select case (family)
  case (0); call mrk_pppp(k)
  case (1); call mrk_qqqq(k)
  case (2); call mrk_pqpq(k)
  case default; call mrk_qpqp(k)
  end select
  rkb_o = rkb(1:ns, 1:ns, 1:ks, 1:ks)
end subroutine

subroutine ref_dbs_convol(family, k, icase, d, a) bind(C, name="ref_dbs_convol")
  use iso_c_binding
  use DBS_grid
  use DBS_integrals
  implicit none
  integer(c_int), value :: family, k, icase
  real(c_double) :: d(ns, *), a(ns, *)
  real(8), allocatable :: dd(:,:), aa(:,:)
  integer :: nc
  select case (family)
  case (0); call mrk_pppp(k)
  case (1); call mrk_qqqq(k)
  case (2); call mrk_pqpq(k)
  case default; call mrk_qpqp(k)
  end select
  nc = ks; if (icase > 2) nc = ns
  allocate(dd(ns, nc), aa(ns, nc))
  dd = d(1:ns, 1:nc)
  call convol(ns, ks, aa, dd, icase, 's', 's')
  a(1:ns, 1:nc) = aa
end subroutine
