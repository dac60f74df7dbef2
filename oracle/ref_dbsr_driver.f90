!======================================================================
!  ref_dbsr_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
!  C-callable entry points over the reference's own Dirac assembly
!  (src/dbsr/DBSR_MAT3_MPI: R_data, O_data, the UPDATE_* routines,
!  Target_h, Zero_cond, transform_matrix), which
!  ref.mk compiles together with the DBS library into
!  oracle/_ref/libdbs_ref.so.  This file holds no reference code: it fills
!  the module state that R_data reads (orbitals pq, ipbs, the <.|p> vectors
!  V_ch with their keys i_ch, j_ch, the channel and perturber counts, the
!  target pointers iptar, the list in module c_data), calls
!  alloc_dbsr_matrix, runs R_data(itype,kpol,0..3) exactly as Gen_matrix
!  does (gen_matrix.f90:11-14), and copies the accumulators out.
!  The grid comes from ref_dbs_init (ref_dbs_driver.f90).
!======================================================================

subroutine ref_dbsr_setup(nbf_, pqin, ipbs_in, nvch, ich_in, jch_in, vch_in, nch_, npert_, mk_, &
                          ntarg_, iptar_in, sizes) bind(C, name="ref_dbsr_setup")
  use iso_c_binding
  use dbsr_mat
  implicit none
  integer(c_int), value :: nbf_, nvch, nch_, npert_, mk_, ntarg_
  real(c_double) :: pqin(ns, 2, nbf_), vch_in(2*ns, *)
  integer(c_int) :: ipbs_in(nbf_), ich_in(*), jch_in(*), iptar_in(nch_), sizes(4)
  call alloc_DBS_orbitals_pq(0, ns)
  call alloc_DBS_orbitals_pq(nbf_, ns)
  nbf = nbf_
  pq(:, :, 1:nbf) = pqin
  ipbs(1:nbf) = ipbs_in
  if (allocated(i_ch)) deallocate(i_ch)
  if (allocated(j_ch)) deallocate(j_ch)
  if (allocated(V_ch)) deallocate(V_ch)
  nv_ch = nvch
  allocate(i_ch(max(1, nvch)), j_ch(max(1, nvch)), V_ch(2*ns, max(1, nvch)))
  if (nvch > 0) then
    i_ch(1:nvch) = ich_in(1:nvch); j_ch(1:nvch) = jch_in(1:nvch)
    V_ch(:, 1:nvch) = vch_in(1:2*ns, 1:nvch)
  end if
  nch = nch_; npert = npert_; mk = mk_; ntarg = ntarg_
  if (allocated(iptar)) deallocate(iptar)
  allocate(iptar(nch)); iptar = iptar_in
  ! work arrays are sized on first allocation only: drop them for a new ms
  if (allocated(x)) deallocate(x, xx)
  if (allocated(y)) deallocate(y, yy)
  nprocs = 1; myid = 0
  iicb = 0; iibb = 0   ! alloc_dbsr_matrix leaves them alone when npert = 0
  call alloc_dbsr_matrix
  sizes(1) = iicc; sizes(2) = iicb; sizes(3) = iibb; sizes(4) = ms
end subroutine

subroutine ref_dbsr_rdata(itype, kpol, n, k1in, k2in, k3in, k4in, cin) bind(C, name="ref_dbsr_rdata")
  use iso_c_binding
  use dbsr_mat
  use c_data
  implicit none
  integer(c_int), value :: itype, kpol, n
  integer(c_int) :: k1in(n), k2in(n), k3in(n), k4in(n)
  real(c_double) :: cin(n)
  integer :: i, b
  if (allocated(K1)) deallocate(K1, K2, K3, K4, IPT, CDATA)
  allocate(K1(n), K2(n), K3(n), K4(n), IPT(n), CDATA(n))
  K1 = k1in; K2 = k2in; K3 = k3in; K4 = k4in; CDATA = cin
  do i = 1, n
    IPT(i) = i
  end do
  ncdata = n
  do b = 0, 3
    call R_data(itype, kpol, b)
  end do
end subroutine

subroutine ref_dbsr_get(hch_o, hcp_o, hp_o, acf_o, htarg_o) bind(C, name="ref_dbsr_get")
  use iso_c_binding
  use dbsr_mat
  implicit none
  real(c_double) :: hch_o(ms, ms, *), hcp_o(ms, *), hp_o(*), acf_o(max(iicc, 1), 0:mk), htarg_o(*)
  integer :: nt
  if (iicc > 0) then
    hch_o(:, :, 1:iicc) = hch
    acf_o(1:iicc, :) = acf
  end if
  if (npert > 0) then
    if (iicb > 0) hcp_o(:, 1:iicb) = hcp
    if (iibb > 0) hp_o(1:iibb) = hp
  end if
  nt = ntarg * (ntarg + 1) / 2
  if (nt > 0) htarg_o(1:nt) = htarg(1:nt)
end subroutine

subroutine ref_dbsr_odata(itype, n, k1in, k2in, cin, eps, iitar_) bind(C, name="ref_dbsr_odata")
  ! the overlap pass of SUB1 (icase 0): O_data(itype) on one list
  use iso_c_binding
  use dbsr_mat
  use c_data
  implicit none
  integer(c_int), value :: itype, n, iitar_
  real(c_double), value :: eps
  integer(c_int) :: k1in(n), k2in(n)
  real(c_double) :: cin(n)
  integer :: i
  if (allocated(K1)) deallocate(K1, K2, K3, K4, IPT, CDATA)
  allocate(K1(n), K2(n), K3(n), K4(n), IPT(n), CDATA(n))
  K1 = k1in; K2 = k2in; K3 = 0; K4 = 0; CDATA = cin
  do i = 1, n
    IPT(i) = i
  end do
  ncdata = n
  eps_c = eps
  iitar = iitar_
  call O_data(itype)
end subroutine

subroutine ref_dbsr_diag_fppqq() bind(C, name="ref_dbsr_diag_fppqq")
  ! the B-spline overlaps of the channel diagonal blocks (sub1.f90:78)
  use dbsr_mat
  use DBS_gauss
  implicit none
  integer :: ich
  do ich = 1, nch
    call UPDATE_HX(ich, ich, fppqq, 1.d0)
  end do
end subroutine

subroutine ref_dbsr_get_otarg(otarg_o) bind(C, name="ref_dbsr_get_otarg")
  use iso_c_binding
  use dbsr_mat
  implicit none
  real(c_double) :: otarg_o(*)
  integer :: nt
  nt = ntarg * (ntarg + 1) / 2
  if (nt > 0) otarg_o(1:nt) = otarg(1:nt)
end subroutine

subroutine ref_dbsr_zero_cond(lch_in, kch_in, ilz, jlz, ibz, jbz, iprm_o, nhm_o) bind(C, name="ref_dbsr_zero_cond")
  ! Zero_cond for the channels of the current wave, channel l and kappa given
  use iso_c_binding
  use dbsr_mat
  use channel_jj, only: lch, kch
  implicit none
  integer(c_int) :: lch_in(*), kch_in(*), iprm_o(*), nhm_o
  integer(c_int), value :: ilz, jlz, ibz, jbz
  if (allocated(lch)) deallocate(lch)
  if (allocated(kch)) deallocate(kch)
  allocate(lch(nch), kch(nch))
  lch = lch_in(1:nch); kch = kch_in(1:nch)
  ilzero = ilz; jlzero = jlz; ibzero = ibz; jbzero = jbz
  mhm = nch * ms + npert
  call Zero_cond
  iprm_o(1:nch*ms) = iprm(1:nch*ms)
  nhm_o = nhm
end subroutine

subroutine ref_dbsr_transform(icase_, diag_in, ipsol_in) bind(C, name="ref_dbsr_transform")
  ! transform_matrix on the module matrices with the given channel solutions
  use iso_c_binding
  use dbsr_mat
  implicit none
  integer(c_int), value :: icase_
  real(c_double) :: diag_in(ms, ms, nch)
  integer(c_int) :: ipsol_in(nch)
  if (allocated(diag)) deallocate(diag)
  allocate(diag(ms, ms, nch)); diag = diag_in
  if (allocated(ipsol)) deallocate(ipsol)
  allocate(ipsol(0:nch)); ipsol(0) = 0; ipsol(1:nch) = ipsol_in
  icase = icase_
  call transform_matrix
end subroutine

subroutine ref_dbsr_get_overlaps(n, ov_o) bind(C, name="ref_dbsr_get_overlaps")
  ! the largest transformed overlaps transform_matrix records for icase 0,
  ! overlaps(nch+npert, nch+npert)
  use iso_c_binding
  use dbsr_mat
  implicit none
  integer(c_int), value :: n
  real(c_double) :: ov_o(n, n)
  ov_o = 0.d0
  if (allocated(overlaps)) ov_o(1:nch+npert, 1:nch+npert) = overlaps(1:nch+npert, 1:nch+npert)
end subroutine

subroutine ref_dbsr_record(path, n) bind(C, name="ref_dbsr_record")
  ! Record_matrix (record_matrix.f90) of the module matrices, written to a
  ! new file on the module's unit nui
  use iso_c_binding
  use dbsr_mat
  implicit none
  integer(c_int), value :: n
  character(kind=c_char) :: path(n)
  character(len=n) :: fn
  integer :: i
  do i = 1, n
    fn(i:i) = path(i)
  end do
  open(nui, file=fn, form='unformatted', status='replace')
  call Record_matrix
  close(nui)
end subroutine
