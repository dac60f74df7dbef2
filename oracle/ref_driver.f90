!======================================================================
!  ref_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
!  C-callable entry points over the reference's own Fortran routines,
!  which ref.mk compiles from the reference checkout into
!  oracle/_ref/libbsr_ref.so.  This file holds no reference code: it
!  only sets the reference modules (spline_param, spline_grid,
!  spline_galerkin, spline_orbitals, bsr_matrix) and calls the reference
!  routines, so that the tests can compare the C restatement with them.
!  The grid tables are filled the way initvb does it
!  (src/libs/BS/define_spline.f90:67-129) because define_spline itself
!  also factorises sb with LAPACK, which this image does not have.
!======================================================================

subroutine ref_mkgrid(kind, ks_, z, h_, hmax_, rmax_, cap, tout, ns_o, ml_o, me_o, h_o) &
     bind(C, name="ref_mkgrid")
  use iso_c_binding
  use spline_param
  use spline_grid
  implicit none
  integer(c_int), value :: kind, ks_, cap
  real(c_double), value :: z, h_, hmax_, rmax_
  real(c_double) :: tout(cap)
  integer(c_int) :: ns_o, ml_o, me_o
  real(c_double) :: h_o
  real(8) :: zz
  ks = ks_; h = h_; hmax = hmax_; rmax = rmax_; zz = z
  if (allocated(t)) deallocate(t)
  if (kind == 0) then
    call mkgrid(zz)
  else
    call mkgrid1(zz)
  end if
  ns_o = ns; ml_o = ml; me_o = me; h_o = h
  tout(1:min(cap, ns+ks)) = t(1:min(cap, ns+ks))
end subroutine

subroutine ref_gauss(k, x, w) bind(C, name="ref_gauss")
  use iso_c_binding
  implicit none
  integer(c_int), value :: k
  real(c_double) :: x(k), w(k)
  real(8) :: xx(k), ww(k)
  call gauss(k, xx, ww)
  x = xx; w = ww
end subroutine

subroutine ref_init_grid(ks_, ns_, ml_, me_, h_, tin) bind(C, name="ref_init_grid")
  use iso_c_binding
  use spline_param
  use spline_grid
  use spline_galerkin
  implicit none
  integer(c_int), value :: ks_, ns_, ml_, me_
  real(c_double), value :: h_
  real(c_double) :: tin(ns_+ks_)
  real(8) :: gx(ks_), gw(ks_)
  real(8), allocatable :: dbiatx(:,:,:)
  integer :: m, i
  ks = ks_; ns = ns_; ml = ml_; me = me_; h = h_; nv = ns - ks + 1
  if (allocated(t)) deallocate(t)
  allocate(t(ns+ks)); t = tin
  call dealloc_moments
  call dealloc_integrals
  call gauss(ks, gx, gw)
  call allocate_grid
  allocate(dbiatx(nv,ks,ks)); dbiatx = 0.d0
  do m = 1, ks
    gr(1:nv,m) = (t(1+ks:nv+ks) - t(ks:nv+ks-1))*gx(m) + t(ks:nv+ks-1)
    grm(1:nv,m) = 1.d0/gr(1:nv,m)
    grw(1:nv,m) = (t(1+ks:nv+ks) - t(ks:nv+ks-1))*gw(m)
    call vbsplvd(t, ks, nv, gr(1,m), 3, dbiatx)
    bsp(1:nv,m,1:ks) = dbiatx(1:nv,1:ks,1)
    bspd(1:nv,m,1:ks,1) = dbiatx(1:nv,1:ks,2)
    bspd(1:nv,m,1:ks,2) = dbiatx(1:nv,1:ks,3)
    do i = 1, ks
      bsq(1:nv,m,i) = bspd(1:nv,m,i,1) - grm(1:nv,m)*bsp(1:nv,m,i)
    end do
  end do
  deallocate(dbiatx)
  call allocate_galerkin
  call mrm(0, sb)
  call mrm(1, r1)
  call mrm(-1, rm1)
  call mrm(-2, rm2)
end subroutine

subroutine ref_tables(gro, grwo, bspo, sbo) bind(C, name="ref_tables")
  use iso_c_binding
  use spline_param
  use spline_grid
  use spline_galerkin
  implicit none
  real(c_double) :: gro(nv,ks), grwo(nv,ks), bspo(nv,ks,ks), sbo(ns,ks)
  gro = gr; grwo = grw; bspo = bsp(1:nv,:,:); sbo = sb
end subroutine

subroutine ref_moments(k, rkm) bind(C, name="ref_moments")
  use iso_c_binding
  use spline_param
  implicit none
  integer(c_int), value :: k
  real(c_double) :: rkm(ks*ks, nv)
  real(8) :: r(ks*ks, nv)
  r = 0.d0
  call moments(k, r, 's', 'b')
  rkm = r
end subroutine

subroutine ref_rk_moments(k, r1o, r2o, rdo) bind(C, name="ref_rk_moments")
  use iso_c_binding
  use spline_param
  use spline_moments
  implicit none
  integer(c_int), value :: k
  real(c_double) :: r1o(ks*(ks+1)/2, nv), r2o(ks*(ks+1)/2, nv)
  real(c_double) :: rdo(ks*(ks+1)/2, ks*(ks+1)/2, nv)
  integer :: ik
  ik = ks*(ks+1)/2
  call rk_moments(k)
  r1o = rkd1(1:ik, 1:nv); r2o = rkd2(1:ik, 1:nv); rdo = rkd(1:ik, 1:ik, 1:nv)
end subroutine

subroutine ref_mrk_cell(k, rkbo) bind(C, name="ref_mrk_cell")
  use iso_c_binding
  use spline_param
  use spline_integrals
  implicit none
  integer(c_int), value :: k
  real(c_double) :: rkbo(ns, ns, ks, ks)
  call mrk_cell(k)
  rkbo = rkb(1:ns, 1:ns, 1:ks, 1:ks)
end subroutine

subroutine ref_density(ns_, ks_, d, p1, p2, typ) bind(C, name="ref_density")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ns_, ks_
  real(c_double) :: d(ns_, *), p1(ns_), p2(ns_)
  character(kind=c_char), value :: typ
  character(1) :: ty
  ty = typ
  call density(ns_, ks_, d, p1, p2, ty)
end subroutine

subroutine ref_convol(k, ns_, ks_, a, d, icase) bind(C, name="ref_convol")
  use iso_c_binding
  implicit none
  integer(c_int), value :: k, ns_, ks_, icase
  real(c_double) :: a(ns_, *), d(ns_, *)
  call mrk_cell(k)
  call convol(ns_, ks_, a, d, icase, 's', 's')
end subroutine

real(c_double) function ref_sum_amb(ns_, ks_, a, b, sym) bind(C, name="ref_sum_amb")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ns_, ks_
  real(c_double) :: a(ns_, *), b(ns_, *)
  character(kind=c_char), value :: sym
  real(8), external :: sum_amb
  character(1) :: sy
  sy = sym
  ref_sum_amb = sum_amb(ns_, ks_, a, b, sy)
end function

subroutine ref_bav(ns_, ks_, b, v, y, typ, side) bind(C, name="ref_bav")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ns_, ks_
  real(c_double) :: b(ns_, *), v(ns_), y(ns_)
  character(kind=c_char), value :: typ, side
  character(1) :: ty, si
  ty = typ; si = side
  call bav(ns_, ks_, b, v, y, ty, si)
end subroutine

subroutine ref_bxv(k, n, b, v, y) bind(C, name="ref_bxv")
  use iso_c_binding
  implicit none
  integer(c_int), value :: k, n
  real(c_double) :: b(n, k), v(n), y(n)
  call bxv(k, n, b, v, y)
end subroutine

real(c_double) function ref_bvmv(n, k, a, x, y) bind(C, name="ref_bvmv")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n, k
  real(c_double) :: a(n, *), x(n), y(n)
  real(8), external :: bvmv
  ref_bvmv = bvmv(n, k, a, 's', x, y)
end function

subroutine ref_set_orbitals(nbf_, pin) bind(C, name="ref_set_orbitals")
  use iso_c_binding
  use spline_param
  use spline_orbitals
  implicit none
  integer(c_int), value :: nbf_
  real(c_double) :: pin(ns, nbf_)
  call allocate_bsorb(0)
  call allocate_bsorb(nbf_)
  nbf = nbf_
  pbs(1:ns, 1:nbf_) = pin
end subroutine

real(c_double) function ref_quadr(i, j, m) bind(C, name="ref_quadr")
  use iso_c_binding
  implicit none
  integer(c_int), value :: i, j, m
  real(8), external :: quadr
  ref_quadr = quadr(i, j, m)
end function

subroutine ref_alloc_matrix(nch, ncp, mk) bind(C, name="ref_alloc_matrix")
  use iso_c_binding
  use spline_param
  implicit none
  integer(c_int), value :: nch, ncp, mk
  call allocate_matrix(nch, ns, ncp, mk, 0)
end subroutine

subroutine ref_update_hx(ich, jch, d, sym) bind(C, name="ref_update_hx")
  use iso_c_binding
  use spline_param
  implicit none
  integer(c_int), value :: ich, jch
  real(c_double) :: d(ns, *)
  character(kind=c_char), value :: sym
  character(1) :: sy
  sy = sym
  call update_hx(ich, jch, ns, ks, d, sy)
end subroutine

subroutine ref_update_hl(ich, jch, c) bind(C, name="ref_update_hl")
  use iso_c_binding
  use spline_param
  use spline_galerkin
  implicit none
  integer(c_int), value :: ich, jch
  real(c_double), value :: c
  real(8) :: cc
  cc = c
  call update_hl(ich, jch, ns, ks, sb, cc)
end subroutine

subroutine ref_update_hw(ich, jch, v, w) bind(C, name="ref_update_hw")
  use iso_c_binding
  use spline_param
  implicit none
  integer(c_int), value :: ich, jch
  real(c_double) :: v(ns), w(ns)
  call update_hw(ich, jch, ns, v, w)
end subroutine

subroutine ref_update_hv(ich, ic, v, c) bind(C, name="ref_update_hv")
  use iso_c_binding
  use spline_param
  implicit none
  integer(c_int), value :: ich, ic
  real(c_double) :: v(ns)
  real(c_double), value :: c
  real(8) :: cc
  cc = c
  call update_hv(ich, ic, ns, v, cc)
end subroutine

subroutine ref_update_hb(ic, jc, c) bind(C, name="ref_update_hb")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ic, jc
  real(c_double), value :: c
  real(8) :: cc
  cc = c
  call update_hb(ic, jc, cc)
end subroutine

subroutine ref_update_cf(k, ich, jch, c) bind(C, name="ref_update_cf")
  use iso_c_binding
  implicit none
  integer(c_int), value :: k, ich, jch
  real(c_double), value :: c
  real(8) :: cc
  cc = c
  call update_cf(k, ich, jch, cc)
end subroutine

subroutine ref_get_matrix(hcco, hcbo, hbbo, acfo) bind(C, name="ref_get_matrix")
  use iso_c_binding
  use bsr_matrix
  implicit none
  real(c_double) :: hcco(*), hcbo(*), hbbo(*), acfo(*)
  hcco(1:size(hcc)) = reshape(hcc, [size(hcc)])
  if (kcp > 0) then
    hcbo(1:size(hcb)) = reshape(hcb, [size(hcb)])
    hbbo(1:size(hbb)) = hbb
  end if
  acfo(1:size(acf)) = reshape(acf, [size(acf)])
end subroutine

!----------------------------------------------------------------------
! One-electron operator with the core: Gen_Lval / INT_core of
! src/bsr/BSR_MAT3/l_core.f90.  db2 is handed in: the reference computes
! it in mdb2, an internal procedure of define_spline (which factorises sb
! with LAPACK), so it cannot be called on its own; the tests pin db2
! separately (exact integrals).
!----------------------------------------------------------------------
subroutine ref_bspd(d1o, d2o) bind(C, name="ref_bspd")
  use iso_c_binding
  use spline_param
  use spline_grid
  implicit none
  real(c_double) :: d1o(nv,ks,ks), d2o(nv,ks,ks)
  d1o = bspd(1:nv,:,:,1); d2o = bspd(1:nv,:,:,2)
end subroutine

subroutine ref_galerkin(rm1o, rm2o) bind(C, name="ref_galerkin")
  use iso_c_binding
  use spline_param
  use spline_galerkin
  implicit none
  real(c_double) :: rm1o(ns,ks), rm2o(ns,ks)
  rm1o = rm1; rm2o = rm2
end subroutine

subroutine ref_set_lcore(z_, kclosd_, nbf_, lbs_, iech_, db2in) bind(C, name="ref_set_lcore")
  use iso_c_binding
  use spline_param
  use spline_atomic
  use spline_galerkin
  use spline_orbitals
  use bsr_mat, only: nud
  implicit none
  real(c_double), value :: z_
  integer(c_int), value :: kclosd_, nbf_
  integer(c_int) :: lbs_(nbf_), iech_(nbf_)
  real(c_double) :: db2in(ns,ks)
  z = z_; kclosd = kclosd_; rel = .false.
  lbs(1:nbf_) = lbs_; iech(1:nbf_) = iech_; nbs(1:nbf_) = 0
  db2 = db2in
  nud = 0
end subroutine

subroutine ref_gen_lval(mlh_o, hlo, lio, lvo) bind(C, name="ref_gen_lval")
  use iso_c_binding
  use spline_param
  use spline_orbitals
  use L_core
  implicit none
  integer(c_int) :: mlh_o
  real(c_double) :: hlo(ns,ns,0:*), lio(nbf,nbf), lvo(ns,nbf)
  call Alloc_Lcore(0,0,0)
  call Gen_Lval
  mlh_o = mlh
  hlo(:,:,0:mlh) = hl_full(:,:,0:mlh)
  lio = L_int; lvo = L_vec
end subroutine

!----------------------------------------------------------------------
! Coefficient lists of module cmdata: Add_cdata into one block, and
! Merge_CDATA over sorted blocks (src/bsr/BSR_MAT3/add_cdata.f90,
! merge_cdata.f90).
!----------------------------------------------------------------------
subroutine ref_add_cdata(n, j1, j2, j3, j4, c, k1o, k2o, k3o, k4o, co, nout) bind(C, name="ref_add_cdata")
  use iso_c_binding
  use cmdata
  implicit none
  integer(c_int), value :: n
  integer(c_int) :: j1(n), j2(n), j3(n), j4(n), k1o(n), k2o(n), k3o(n), k4o(n), nout
  real(c_double) :: c(n), co(n)
  integer :: i, ip, jp
  if (allocated(CDATA)) deallocate(CDATA, K1, K2, K3, K4, IPT)
  allocate(CDATA(n + 1), K1(n + 1), K2(n + 1), K3(n + 1), K4(n + 1), IPT(n + 1))
  ip = 1; jp = 0
  do i = 1, n
    call Add_cdata(ip, jp, c(i), j1(i), j2(i), j3(i), j4(i))
  end do
  nout = jp
  k1o(1:jp) = K1(1:jp); k2o(1:jp) = K2(1:jp); k3o(1:jp) = K3(1:jp); k4o(1:jp) = K4(1:jp)
  co(1:jp) = CDATA(1:jp)
end subroutine

subroutine ref_merge_cdata(nn, ip, jp, n, k1i, k2i, k3i, k4i, c, eps, ipto, nc) bind(C, name="ref_merge_cdata")
  use iso_c_binding
  use cmdata
  implicit none
  integer(c_int), value :: nn, n
  integer(c_int) :: ip(nn), jp(nn), k1i(n), k2i(n), k3i(n), k4i(n), ipto(n), nc
  real(c_double) :: c(n)
  real(c_double), value :: eps
  integer :: ncc
  if (allocated(CDATA)) deallocate(CDATA, K1, K2, K3, K4, IPT)
  allocate(CDATA(n), K1(n), K2(n), K3(n), K4(n), IPT(n))
  K1 = k1i; K2 = k2i; K3 = k3i; K4 = k4i; CDATA = c
  call Merge_cdata(nn, ip, jp, ncc, eps)
  nc = ncc
  ipto(1:ncc) = IPT(1:ncc)
  c = CDATA
end subroutine
