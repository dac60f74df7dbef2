!======================================================================
!  ref_lb_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
!  C-callable entry points over the reference's own routines of the
!  BSR_MAT3 list builder: Read_dets (mod_dets.f90) with the int_bnk
!  header readers of ZCONFLS, Decode_INT (incode_int.f90), Ifind_channel
!  (MOD_channel.f90), Jsym_int, Det_fact with Check_det, VDET and DET, and
!  Idef_itype (src/bsr/BSR_MAT3).  ref.mk compiles them with the libs_03
!  spline_orbitals module, the one that holds the IBORT and OBS arrays
!  check_det.f90 and vdet.f90 read, into oracle/_ref/liblb_ref.so.
!
!  This file holds no reference code.  It fills the module state those
!  routines read, and ref_lb_expand walks the int_bnk records over the
!  state pairs in the order of State_res (state_res.f90:45-150), calling
!  the routines above where State_res calls them and returning each
!  coefficient it would pass to Add_coef, with its structure and key.
!======================================================================

subroutine ref_lb_read(cap, c_o, ijt_o, int_o, idf_o, n_o, sz) bind(C, name="ref_lb_read")
  ! Read_dets on 'int_bnk.tst' in the current directory, then the
  ! coefficient records as State_res reads them (state_res.f90:37-41)
  use iso_c_binding
  use bsr_mat
  use dets
  implicit none
  integer(c_int), value :: cap
  real(c_double) :: c_o(cap)
  integer(c_int) :: ijt_o(cap), int_o(cap), idf_o(cap), n_o, sz(4)
  real(8) :: c
  integer :: ijt, ib, idf
  AF_bnk = 'int_bnk.tst'
  ndet = 0; kdet = 0; ndef = 0; kdef = 0
  call Read_dets
  n_o = 0
  do
    read(nub, end=2) c, ijt, ib, idf
    n_o = n_o + 1
    if (n_o <= cap) then
      c_o(n_o) = c; ijt_o(n_o) = ijt; int_o(n_o) = ib; idf_o(n_o) = idf
    end if
  end do
2 close(nub)
  sz(1) = ndet; sz(2) = kdet; sz(3) = ndef; sz(4) = kdef
end subroutine

subroutine ref_lb_get_dets(kpd_o, ipd_o, npd_o, kpf_o, ipf_o, npf_o) bind(C, name="ref_lb_get_dets")
  use iso_c_binding
  use dets
  implicit none
  integer(c_int) :: kpd_o(*), ipd_o(*), npd_o(*), kpf_o(*), ipf_o(*), npf_o(*)
  if (ndet > 0) then
    kpd_o(1:ndet) = KPD(1:ndet); ipd_o(1:ndet) = IPD(1:ndet); npd_o(1:kdet) = NPD(1:kdet)
  end if
  if (ndet > 0 .and. ndef > 0) then
    kpf_o(1:ndef) = KPF(1:ndef); ipf_o(1:ndef) = IPF(1:ndef); npf_o(1:kdef) = NPF(1:kdef)
  end if
end subroutine

subroutine ref_lb_set_dets(ndet_, kdet_, kpd_in, ipd_in, npd_in, ndef_, kdef_, kpf_in, ipf_in, npf_in) &
     bind(C, name="ref_lb_set_dets")
  use iso_c_binding
  use dets
  implicit none
  integer(c_int), value :: ndet_, kdet_, ndef_, kdef_
  integer(c_int) :: kpd_in(*), ipd_in(*), npd_in(*), kpf_in(*), ipf_in(*), npf_in(*)
  if (allocated(KPD)) deallocate(KPD, IPD, NPD, JPD)
  if (allocated(KPF)) deallocate(KPF, IPF, NPF, JPF)
  ndet = ndet_; kdet = kdet_; ndef = ndef_; kdef = kdef_
  allocate(KPD(max(ndet, 1)), IPD(max(ndet, 1)), NPD(max(kdet, 1)), JPD(max(ndet, 1)))
  allocate(KPF(max(ndef, 1)), IPF(max(ndef, 1)), NPF(max(kdef, 1)), JPF(max(ndef, 1)))
  if (ndet > 0) then
    KPD(1:ndet) = kpd_in(1:ndet); IPD(1:ndet) = ipd_in(1:ndet); NPD(1:kdet) = npd_in(1:kdet)
  end if
  if (ndef > 0) then
    KPF(1:ndef) = kpf_in(1:ndef); IPF(1:ndef) = ipf_in(1:ndef); NPF(1:kdef) = npf_in(1:kdef)
  end if
end subroutine

subroutine ref_lb_setup(nbf_, iech_in, lef_in, ibort_in, obs_in, nelc_, nch_, ipconf_in, npert_, ippert_in) &
     bind(C, name="ref_lb_setup")
  ! orbitals (spline_orbitals: iech, IBORT, OBS; orb_LS: LEF, IEF), the
  ! electron count (target) and the channel/perturber configuration
  ! pointers (channel: ipconf, ippert)
  use iso_c_binding
  use spline_param, only: ns
  use spline_orbitals, only: nbf, iech, IBORT, OBS
  use orb_LS, only: LEF, IEF
  use target, only: nelc
  use channel, only: nch, npert, ipconf, ippert
  implicit none
  integer(c_int), value :: nbf_, nelc_, nch_, npert_
  integer(c_int) :: iech_in(nbf_), lef_in(nbf_), ibort_in(nbf_, nbf_), ipconf_in(*), ippert_in(0:*)
  real(c_double) :: obs_in(nbf_, nbf_)
  ns = 1
  call allocate_bsorb(0)
  call allocate_bsorb(nbf_)
  nbf = nbf_
  iech(1:nbf_) = iech_in
  IBORT(1:nbf_, 1:nbf_) = ibort_in
  OBS(1:nbf_, 1:nbf_) = obs_in
  call Alloc_orb_LS(0)
  call Alloc_orb_LS(nbf_)
  LEF(1:nbf_) = lef_in; IEF(1:nbf_) = iech_in
  nelc = nelc_
  call Allocate_channel(0)
  call Allocate_channel(max(nch_, 1))
  nch = nch_
  ipconf(1:nch_) = ipconf_in(1:nch_)
  call Allocate_pert(0)
  call Allocate_pert(max(npert_, 1))
  npert = npert_
  ippert(0:npert_) = ippert_in(0:npert_)
end subroutine

subroutine ref_lb_decode(n, ints, out) bind(C, name="ref_lb_decode")
  ! Decode_INT of n packed integrals: out(6,n) = met, k, i1, i2, i3, i4
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  integer(c_int) :: ints(n), out(6, n)
  integer :: i
  do i = 1, n
    call Decode_INT(out(1, i), out(2, i), out(3, i), out(4, i), out(5, i), out(6, i), ints(i))
  end do
end subroutine

subroutine ref_lb_jsym(n, met, io) bind(C, name="ref_lb_jsym")
  ! Jsym_int on n index quadruples io(4,n) in place
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  integer(c_int) :: met(n), io(4, n)
  integer :: i, j1, j2, j3, j4
  do i = 1, n
    j1 = io(1, i); j2 = io(2, i); j3 = io(3, i); j4 = io(4, i)
    call Jsym_int(met(i), j1, j2, j3, j4)
    io(1, i) = j1; io(2, i) = j2; io(3, i) = j3; io(4, i) = j4
  end do
end subroutine

subroutine ref_lb_expand(icase_, mk_, eps_c_, eps_det_, eps_soo_, nrec, cbuf, ijtb, intb, idfb, ncfg, ip_state, &
                         no_state, wc, lorb, ip_orb, nsym, it1, it2, nstat, ip_stat, cap, o_itype, o_kpol, o_k, &
                         o_c, n_o) bind(C, name="ref_lb_expand")
  ! The State_res walk for one icase (1-6 and 11; the J-dependent cases
  ! 7-10 need the term tables and are not driven here)
  use iso_c_binding
  use bsr_mat, only: Eps_c, Eps_det, Eps_soo
  use param_LS, only: ibc
  use cmdata, only: ibi
  use new_defs
  use orb_LS, only: LEF
  use channel, only: nch
  implicit none
  integer(c_int), value :: icase_, mk_, nrec, ncfg, lorb, nsym, nstat, cap
  real(c_double), value :: eps_c_, eps_det_, eps_soo_
  real(c_double) :: cbuf(nrec), wc(ncfg), o_c(cap)
  integer(c_int) :: ijtb(nrec), intb(nrec), idfb(nrec), ip_state(ncfg), no_state(ncfg), ip_orb(lorb)
  integer(c_int) :: it1(nsym), it2(nsym), ip_stat(nstat), o_itype(cap), o_kpol(cap), o_k(4, cap), n_o
  integer, external :: Ifind_channel, Idef_itype
  integer :: ibuf, jcase, k, i1, i2, i3, i4, kpol, ijt, it, jt, is1, is2, js1, js2
  integer :: ik, jk, is, js, ip1, ip2, no1, no2, ich, jch, i, j, ic, jc, j1, j2, j3, j4
  integer :: io, jo, kk1, kk2, kk3, kk4, itype
  integer :: np1(256), np2(256)
  real(8) :: C, CC, CCC
  Eps_c = eps_c_; Eps_det = eps_det_; Eps_soo = eps_soo_
  call Allocate_ndets(-1)
  call Allocate_ndefs(-1)
  n_o = 0
  do ibuf = 1, nrec
    call Decode_INT(jcase, k, i1, i2, i3, i4, intb(ibuf))
    if (icase_ /= jcase) cycle
    kpol = k
    select case (icase_)
    case (4, 8, 9, 10); kpol = k - 1
    case (6, 7, 11); kpol = 0
    end select
    if (kpol > mk_) cycle
    ijt = ijtb(ibuf)
    it = ijt / ibc; jt = mod(ijt, ibc)
    is1 = it1(it); js1 = it1(jt)
    if (is1 == 0 .or. js1 == 0) cycle
    is2 = it2(it); js2 = it2(jt)
    do ik = is1, is2
      is = ip_stat(ik); ip1 = ip_state(is)
      no1 = no_state(is); np1(1:no1) = ip_orb(ip1+1:ip1+no1)
      ich = Ifind_channel(is)
      do jk = js1, js2
        js = ip_stat(jk); ip2 = ip_state(js)
        no2 = no_state(js); np2(1:no2) = ip_orb(ip2+1:ip2+no2)
        jch = Ifind_channel(js)
        if (it == jt .and. js > is) cycle
        select case (icase_)
        case (3, 4, 8, 9, 10); if (abs(wc(is) * wc(js)) < Eps_soo) cycle
        end select
        C = cbuf(ibuf)
        if (ich == jch .and. is /= js) C = C + C
        CC = C * wc(is) * wc(js)
        if (abs(CC) < Eps_c) cycle
        i = 0; if (ich > nch) i = ich - nch
        j = 0; if (jch > nch) j = jch - nch
        ic = max(i, j); jc = min(i, j)
        if (icase_ == 11) then
          j1 = 1; j2 = 1; j3 = 1; j4 = 1
        else
          j1 = np1(i1); j2 = np1(i2); j3 = np2(i3); j4 = np2(i4)
        end if
        if (icase_ == 6) kpol = LEF(j1)
        call Jsym_int(icase_, j1, j2, j3, j4)
        call Det_fact(idfb(ibuf), np1, np2)
        if (nndef == 0) cycle
        do i = 1, nndef
          CCC = CC * Adef(i); io = iof(i); jo = jof(i)
          itype = Idef_itype(j1, j2, j3, j4, io, jo, ic, jc, kk1, kk2, kk3, kk4, ibi)
          n_o = n_o + 1
          if (n_o <= cap) then
            o_itype(n_o) = itype; o_kpol(n_o) = kpol; o_c(n_o) = CCC
            o_k(1, n_o) = kk1; o_k(2, n_o) = kk2; o_k(3, n_o) = kk3; o_k(4, n_o) = kk4
          end if
        end do
      end do
    end do
  end do
end subroutine
