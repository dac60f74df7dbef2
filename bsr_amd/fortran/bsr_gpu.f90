!======================================================================
!  Module bsr_gpu -- ISO_C_BINDING interface to libbsr_gpu.so.
!
!  The entry points replace the reference's radial-core routines with the
!  same argument meaning (1-based indices, Fortran column-major arrays):
!    gpu_mrk_cell(k, rkb)          MRK_cell(k)        src/libs/BS/mrk_cell.f90:2
!    gpu_convol(k, a, d, icase)    convol(ns,ks,a,d,icase,'s','s')
!                                                     src/libs/BS/convol.f90:2
!    gpu_i_data(jtype, jpol, ...)  I_data(jtype,jpol) src/bsr/BSR_MAT3/I_data.f90:2
!    gpu_o_data(...)               O_data             src/bsr/BSR_MAT3/O_data.f90:2
!    gpu_update_hl(ich, jch, c)    UPDATE_HL(...,sb,c) src/bsr/BSR_MAT3/mod_intmat.f90:132
!    gpu_gen_lval(z, kclosd)       Gen_Lval + INT_core src/bsr/BSR_MAT3/l_core.f90:92
!    gpu_l_data(jtype, l, ...)     L_data(jtype,l)    src/bsr/BSR_MAT3/L_data.f90:2
!    bsrg_bs_orth(gpu_ctx, 0)      BS_ORTH            src/bsr/BSR_MAT3/bs_orth.f90:2
!    bsrg_cdata_merge(...)         Merge_CDATA        src/bsr/BSR_MAT3/merge_cdata.f90:2
!  and, for Dirac partial waves (DBSR_MAT3):
!    gpu_dbs_grid(ks,ns,ksp,ksq,t) DBS_grid / DBS_gauss module state
!    gpu_mrk_dirac(family, k, rkb) mrk_pppp/qqqq/pqpq/qpqp(k)  src/libs/DBS/mrk_pppp.f90:2
!    gpu_dirac_orbitals(...)       DBS_orbitals_pq state (pq, ipbs, V_ch)
!    gpu_dirac_r_data(...)         Gen_matrix icase 2 = R_data(itype,kpol,0..3)
!                                  over every list   src/dbsr/DBSR_MAT3_MPI/R_data.f90:2
!    gpu_dirac_get(...)            the accumulators hch, hcp, hp, acf, htarg
!    gpu_dirac_overlap(...)        UPDATE_HX(fppqq) + O_data  sub1.f90:77-79, O_data.f90:2
!    gpu_dirac_diag(...)           Zero_cond + Diag_channels  diag.f90:2, zero_cond.f90:2
!    bsrg_dirac_transform(ctx, w)  transform_matrix           transform_matrix.f90:2
!    gpu_dirac_record(...)         Record_matrix + SUB1's records  record_matrix.f90:2
!  the list builder (BSR_MAT3 State_res, src/bsr/BSR_MAT3/state_res.f90:2):
!    gpu_lb_open(path) .. gpu_lb_run(icase, mk) .. bsrg_lb_lists
!  and BSR_MAT4's interrupted runs: gpu_matrix_record, gpu_matrix_resume
!  (src/bsr/BSR_MAT4/state_res.f90:187-196, read_matrix.f90:2).
!  and the closing eigenvalue step of BSR_HD3 (diag_hd.f90:271-312, full
!  basis): bsrg_wave_eigen (assembled wave), bsrg_matfile_eigen (bsr_mat.nnn).
!  A non-zero status stops the program with the library's message, as the
!  reference's Stop does.
!======================================================================
module bsr_gpu
  use iso_c_binding
  implicit none
  type(c_ptr), save :: gpu_ctx = c_null_ptr

  interface
    function bsrg_last_error() bind(C, name="bsrg_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    integer(c_int) function bsrg_create(ctx, device) bind(C, name="bsrg_create")
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function
    integer(c_int) function bsrg_destroy(ctx) bind(C, name="bsrg_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function bsrg_set_grid(ctx, ks, ns, ml, me, h, t) bind(C, name="bsrg_set_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ks, ns, ml, me
      real(c_double), value :: h
      real(c_double) :: t(*)
    end function
    integer(c_int) function bsrg_mrk_cell(ctx, k, rkb) bind(C, name="bsrg_mrk_cell")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double) :: rkb(*)
    end function
    integer(c_int) function bsrg_convol(ctx, k, icase, d, a) bind(C, name="bsrg_convol")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k, icase
      real(c_double) :: d(*), a(*)
    end function
    integer(c_int) function bsrg_set_orbitals(ctx, nbf, pbs, iech, lbs, ibort, eps_ovl) &
        bind(C, name="bsrg_set_orbitals")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbf
      real(c_double) :: pbs(*)
      integer(c_int) :: iech(*), lbs(*), ibort(*)
      real(c_double), value :: eps_ovl
    end function
    integer(c_int) function bsrg_alloc_matrix(ctx, nch, ncp, mk) bind(C, name="bsrg_alloc_matrix")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nch, ncp, mk
    end function
    integer(c_int) function bsrg_update_hl(ctx, ich, jch, c) bind(C, name="bsrg_update_hl")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ich, jch
      real(c_double), value :: c
    end function
    integer(c_int) function bsrg_idata(ctx, icase, jtype, jpol, n, ipt, k1, k2, k3, k4, c) &
        bind(C, name="bsrg_idata")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: icase, jtype, jpol, n
      integer(c_int) :: ipt(*), k1(*), k2(*), k3(*), k4(*)
      real(c_double) :: c(*)
    end function
    integer(c_int) function bsrg_odata(ctx, n, ipt, k3, k4, c, eps_c) bind(C, name="bsrg_odata")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int) :: ipt(*), k3(*), k4(*)
      real(c_double) :: c(*)
      real(c_double), value :: eps_c
    end function
    integer(c_int) function bsrg_gen_lval(ctx, z, kclosd) bind(C, name="bsrg_gen_lval")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: z
      integer(c_int), value :: kclosd
    end function
    integer(c_int) function bsrg_ldata(ctx, jtype, l, n, ipt, k1, k2, k3, k4, c, iitar, eps_c) &
        bind(C, name="bsrg_ldata")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: jtype, l, n, iitar
      integer(c_int) :: ipt(*), k1(*), k2(*), k3(*), k4(*)
      real(c_double) :: c(*)
      real(c_double), value :: eps_c
    end function
    integer(c_int) function bsrg_cdata_merge(nn, ip, jp, n, k1, k2, k3, k4, c, eps_c, ipt, nc) &
        bind(C, name="bsrg_cdata_merge")
      import :: c_int, c_double
      integer(c_int), value :: nn, n
      integer(c_int) :: ip(*), jp(*), k1(*), k2(*), k3(*), k4(*), ipt(*), nc
      real(c_double) :: c(*)
      real(c_double), value :: eps_c
    end function
    integer(c_int) function bsrg_bs_orth(ctx, which) bind(C, name="bsrg_bs_orth")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
    end function
    integer(c_int) function bsrg_symmetrize_diag(ctx) bind(C, name="bsrg_symmetrize_diag")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function bsrg_scale(ctx, ec) bind(C, name="bsrg_scale")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ec
    end function
    integer(c_int) function bsrg_acf_finish(ctx, kcore) bind(C, name="bsrg_acf_finish")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: kcore
    end function
    integer(c_int) function bsrg_get_matrix(ctx, hcc, hcb, hbb, acf, htarg, otarg) &
        bind(C, name="bsrg_get_matrix")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: hcc(*), hcb(*), hbb(*), acf(*), htarg(*), otarg(*)
    end function
    integer(c_int) function bsrg_wave_assemble(ctx, nch, ncp, mk, nrec, jtype, kpol, k1, k2, k3, k4, c, &
        novl, ok3, ok4, oc, eps_c, ec, ntarg, etarg, iptar) bind(C, name="bsrg_wave_assemble")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nch, ncp, mk, nrec, novl, ntarg
      integer(c_int) :: jtype(*), kpol(*), k1(*), k2(*), k3(*), k4(*), ok3(*), ok4(*), iptar(*)
      real(c_double) :: c(*), oc(*), etarg(*)
      real(c_double), value :: eps_c, ec
    end function
    integer(c_int) function bsrg_lists_upload(ctx, nrec, jtype, kpol, k1, k2, k3, k4, c) &
        bind(C, name="bsrg_lists_upload")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrec
      integer(c_int) :: jtype(*), kpol(*), k1(*), k2(*), k3(*), k4(*)
      real(c_double) :: c(*)
    end function
    integer(c_int) function bsrg_wave_assemble_resident(ctx, nch, ncp, mk, novl, ok3, ok4, oc, eps_c, ec, &
        ntarg, etarg, iptar) bind(C, name="bsrg_wave_assemble_resident")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nch, ncp, mk, novl, ntarg
      integer(c_int) :: ok3(*), ok4(*), iptar(*)
      real(c_double) :: oc(*), etarg(*)
      real(c_double), value :: eps_c, ec
    end function
    integer(c_int) function bsrg_wave_write(ctx, path, tmax, ntarg, etarg, ec) bind(C, name="bsrg_wave_write")
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      real(c_double), value :: tmax, ec
      integer(c_int), value :: ntarg
      real(c_double) :: etarg(*)
    end function
    integer(c_int) function bsrg_wave_eigen(ctx, nev, e) bind(C, name="bsrg_wave_eigen")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nev
      real(c_double) :: e(*)
    end function
    integer(c_int) function bsrg_matfile_eigen(ctx, path, nev, e) bind(C, name="bsrg_matfile_eigen")
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      integer(c_int), value :: nev
      real(c_double) :: e(*)
    end function
    integer(c_int) function bsrg_dbs_set_grid(ctx, ks, ns, ksp, ksq, t) bind(C, name="bsrg_dbs_set_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ks, ns, ksp, ksq
      real(c_double) :: t(*)
    end function
    integer(c_int) function bsrg_dbs_mrk(ctx, family, k, rkb) bind(C, name="bsrg_dbs_mrk")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: family, k
      real(c_double) :: rkb(*)
    end function
    integer(c_int) function bsrg_dirac_set_orbitals(ctx, nbf, pq, ipbs, nv_ch, i_ch, j_ch, v_ch) &
        bind(C, name="bsrg_dirac_set_orbitals")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbf, nv_ch
      real(c_double) :: pq(*), v_ch(*)
      integer(c_int) :: ipbs(*), i_ch(*), j_ch(*)
    end function
    integer(c_int) function bsrg_dirac_assemble(ctx, nch, npert, mk, ntarg, iptar, nrec, itype, kpol, &
        k1, k2, k3, k4, c) bind(C, name="bsrg_dirac_assemble")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nch, npert, mk, ntarg, nrec
      integer(c_int) :: iptar(*), itype(*), kpol(*), k1(*), k2(*), k3(*), k4(*)
      real(c_double) :: c(*)
    end function
    integer(c_int) function bsrg_dirac_get(ctx, hch, hcp, hp, acf, htarg) bind(C, name="bsrg_dirac_get")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: hch(*), hcp(*), hp(*), acf(*), htarg(*)
    end function
    integer(c_int) function bsrg_dirac_overlap(ctx, nch, npert, ntarg, iptar, nrec, itype, k1, k2, cdata, &
        eps_c, iitar) bind(C, name="bsrg_dirac_overlap")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nch, npert, ntarg, nrec, iitar
      integer(c_int) :: iptar(*), itype(*), k1(*), k2(*)
      real(c_double) :: cdata(*)
      real(c_double), value :: eps_c
    end function
    integer(c_int) function bsrg_dirac_overlap_get(ctx, hch, hcp, hp, otarg) bind(C, name="bsrg_dirac_overlap_get")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: hch(*), hcp(*), hp(*), otarg(*)
    end function
    integer(c_int) function bsrg_dirac_update_hx(ctx, which, nmat, d, nrec, ich, jch, imat, c) &
        bind(C, name="bsrg_dirac_update_hx")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, nmat, nrec
      real(c_double) :: d(*), c(*)
      integer(c_int) :: ich(*), jch(*), imat(*)
    end function
    integer(c_int) function bsrg_dirac_energy_shift(ctx, ecore, etarg) bind(C, name="bsrg_dirac_energy_shift")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ecore
      real(c_double) :: etarg(*)
    end function
    integer(c_int) function bsrg_dirac_diag_channels(ctx, lch, kch, ilzero, jlzero, ibzero, jbzero, edmin, edmax, &
        egap, ipsol, nsol) bind(C, name="bsrg_dirac_diag_channels")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int) :: lch(*), kch(*), ipsol(*), nsol
      integer(c_int), value :: ilzero, jlzero, ibzero, jbzero
      real(c_double), value :: edmin, edmax, egap
    end function
    integer(c_int) function bsrg_dirac_solutions(ctx, diag, eval, iprm) bind(C, name="bsrg_dirac_solutions")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: diag(*), eval(*)
      integer(c_int) :: iprm(*)
    end function
    integer(c_int) function bsrg_dirac_transform(ctx, which) bind(C, name="bsrg_dirac_transform")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
    end function
    integer(c_int) function bsrg_dirac_overlaps(ctx, ov) bind(C, name="bsrg_dirac_overlaps")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: ov(*)
    end function
    integer(c_int) function bsrg_dirac_record(ctx, path, which, nelc_core) bind(C, name="bsrg_dirac_record")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      integer(c_int), value :: which, nelc_core
    end function
    integer(c_int) function bsrg_matrix_record(ctx, path, interrupt, intercase, max_nbuf) &
        bind(C, name="bsrg_matrix_record")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      integer(c_int), value :: interrupt, intercase, max_nbuf
    end function
    integer(c_int) function bsrg_matrix_resume(ctx, path, interrupt, intercase, max_nbuf) &
        bind(C, name="bsrg_matrix_resume")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: path(*)
      integer(c_int) :: interrupt, intercase, max_nbuf
    end function
    function bsrg_lb_create() bind(C, name="bsrg_lb_create") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    subroutine bsrg_lb_destroy(lb) bind(C, name="bsrg_lb_destroy")
      import :: c_ptr
      type(c_ptr), value :: lb
    end subroutine
    integer(c_int) function bsrg_lb_read_bank(lb, path, sizes5) bind(C, name="bsrg_lb_read_bank")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: lb
      character(kind=c_char) :: path(*)
      integer(c_int) :: sizes5(5)
    end function
    integer(c_int) function bsrg_lb_set_orbitals(lb, nbf, iech, lef, ibort, obs, nelc) &
        bind(C, name="bsrg_lb_set_orbitals")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: lb
      integer(c_int), value :: nbf, nelc
      integer(c_int) :: iech(*), lef(*), ibort(*)
      real(c_double) :: obs(*)
    end function
    integer(c_int) function bsrg_lb_set_channels(lb, nch, ipconf, npert, ippert) bind(C, name="bsrg_lb_set_channels")
      import :: c_ptr, c_int
      type(c_ptr), value :: lb
      integer(c_int), value :: nch, npert
      integer(c_int) :: ipconf(*), ippert(*)
    end function
    integer(c_int) function bsrg_lb_set_states(lb, ncfg, ip_state, no_state, wc, lorb, ip_orb, nsym, it_state1, &
        it_state2, nstat, ip_stat) bind(C, name="bsrg_lb_set_states")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: lb
      integer(c_int), value :: ncfg, lorb, nsym, nstat
      integer(c_int) :: ip_state(*), no_state(*), ip_orb(*), it_state1(*), it_state2(*), ip_stat(*)
      real(c_double) :: wc(*)
    end function
    integer(c_int) function bsrg_lb_set_params(lb, eps_c, eps_det, eps_soo, nb, mb, kb) &
        bind(C, name="bsrg_lb_set_params")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: lb
      real(c_double), value :: eps_c, eps_det, eps_soo
      integer(c_int), value :: nb, mb, kb
    end function
    integer(c_int) function bsrg_lb_run(lb, icase, mk, sizes3) bind(C, name="bsrg_lb_run")
      import :: c_ptr, c_int
      type(c_ptr), value :: lb
      integer(c_int), value :: icase, mk
      integer(c_int) :: sizes3(3)
    end function
    integer(c_int) function bsrg_lb_lists(lb, batch, jtype, jpol, k1, k2, k3, k4, c) bind(C, name="bsrg_lb_lists")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: lb
      integer(c_int) :: batch(*), jtype(*), jpol(*), k1(*), k2(*), k3(*), k4(*)
      real(c_double) :: c(*)
    end function
  end interface

contains

  subroutine gpu_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(256) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(bsrg_last_error(), msg, [256])
    text = ' '
    do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a,i6,2a)') 'bsr_gpu: status ', rc, ': ', trim(text)
    stop 'bsr_gpu failure'
  end subroutine

  subroutine gpu_init(device)
    integer, intent(in) :: device
    if (c_associated(gpu_ctx)) return
    call gpu_check(bsrg_create(gpu_ctx, int(device, c_int)))
  end subroutine

  subroutine gpu_finish()
    integer(c_int) :: rc
    if (.not. c_associated(gpu_ctx)) return
    rc = bsrg_destroy(gpu_ctx)
    gpu_ctx = c_null_ptr
  end subroutine

  subroutine gpu_grid(ks, ns, ml, me, h, t)
    integer, intent(in) :: ks, ns, ml, me
    real(8), intent(in) :: h, t(ns+ks)
    real(c_double) :: tt(ns+ks)
    tt = t
    call gpu_check(bsrg_set_grid(gpu_ctx, ks, ns, ml, me, h, tt))
  end subroutine

  subroutine gpu_mrk_cell(k, rkb)
    integer, intent(in) :: k
    real(8), intent(out) :: rkb(*)
    call gpu_check(bsrg_mrk_cell(gpu_ctx, k, rkb))
  end subroutine

  subroutine gpu_convol(k, a, d, icase)
    integer, intent(in) :: k, icase
    real(8) :: a(*), d(*)
    call gpu_check(bsrg_convol(gpu_ctx, k, icase, d, a))
  end subroutine

  subroutine gpu_orbitals(ns, nbf, pbs, iech, lbs, ibort, eps_ovl)
    integer, intent(in) :: ns, nbf, iech(nbf), lbs(nbf), ibort(nbf, nbf)
    real(8), intent(in) :: pbs(ns, nbf), eps_ovl
    real(c_double) :: p(ns, nbf)
    integer(c_int) :: ie(nbf), lb(nbf), ib(nbf, nbf)
    p = pbs; ie = iech; lb = lbs; ib = ibort
    call gpu_check(bsrg_set_orbitals(gpu_ctx, nbf, p, ie, lb, ib, eps_ovl))
  end subroutine

  subroutine gpu_allocate_matrix(nch, ncp, mk)
    integer, intent(in) :: nch, ncp, mk
    call gpu_check(bsrg_alloc_matrix(gpu_ctx, nch, ncp, mk))
  end subroutine

  subroutine gpu_update_hl(ich, jch, c)
    integer, intent(in) :: ich, jch
    real(8), intent(in) :: c
    call gpu_check(bsrg_update_hl(gpu_ctx, ich, jch, c))
  end subroutine

  subroutine gpu_i_data(jtype, jpol, ncdata, ipt, k1, k2, k3, k4, cdata)
    integer, intent(in) :: jtype, jpol, ncdata
    integer(c_int), intent(in) :: ipt(*), k1(*), k2(*), k3(*), k4(*)
    real(c_double), intent(in) :: cdata(*)
    call gpu_check(bsrg_idata(gpu_ctx, 5, jtype, jpol, ncdata, ipt, k1, k2, k3, k4, cdata))
  end subroutine

  subroutine gpu_o_data(ncdata, ipt, k3, k4, cdata, eps_c)
    integer, intent(in) :: ncdata
    integer(c_int), intent(in) :: ipt(*), k3(*), k4(*)
    real(c_double), intent(in) :: cdata(*)
    real(8), intent(in) :: eps_c
    call gpu_check(bsrg_odata(gpu_ctx, ncdata, ipt, k3, k4, cdata, eps_c))
  end subroutine

  ! Gen_Lval (l_core.f90:92-170): the one-electron operator with the core
  ! orbitals 1..kclosd, before the L-integral lists of a partial wave
  subroutine gpu_gen_lval(z, kclosd)
    real(8), intent(in) :: z
    integer, intent(in) :: kclosd
    call gpu_check(bsrg_gen_lval(gpu_ctx, z, kclosd))
  end subroutine

  ! L_data(jtype,l) (L_data.f90:2-149) for the records of module cmdata
  subroutine gpu_l_data(jtype, l, ncdata, ipt, k1, k2, k3, k4, cdata, iitar, eps_c)
    integer, intent(in) :: jtype, l, ncdata, iitar
    integer(c_int), intent(in) :: ipt(*), k1(*), k2(*), k3(*), k4(*)
    real(c_double), intent(in) :: cdata(*)
    real(8), intent(in) :: eps_c
    call gpu_check(bsrg_ldata(gpu_ctx, jtype, l, ncdata, ipt, k1, k2, k3, k4, cdata, iitar, eps_c))
  end subroutine

  ! ---- Dirac partial waves ----------------------------------------------
  subroutine gpu_dbs_grid(ks, ns, ksp, ksq, t)
    integer, intent(in) :: ks, ns, ksp, ksq
    real(8), intent(in) :: t(*)
    call gpu_check(bsrg_dbs_set_grid(gpu_ctx, ks, ns, ksp, ksq, t))
  end subroutine

  subroutine gpu_mrk_dirac(family, k, rkb)
    integer, intent(in) :: family, k   ! 0 pppp, 1 qqqq, 2 pqpq, 3 qpqp
    real(8), intent(out) :: rkb(*)
    call gpu_check(bsrg_dbs_mrk(gpu_ctx, family, k, rkb))
  end subroutine

  subroutine gpu_dirac_orbitals(nbf, pq, ipbs, nv_ch, i_ch, j_ch, v_ch)
    integer, intent(in) :: nbf, nv_ch
    real(8), intent(in) :: pq(*), v_ch(*)
    integer, intent(in) :: ipbs(*), i_ch(*), j_ch(*)
    call gpu_check(bsrg_dirac_set_orbitals(gpu_ctx, nbf, pq, ipbs, nv_ch, i_ch, j_ch, v_ch))
  end subroutine

  subroutine gpu_dirac_r_data(nch, npert, mk, ntarg, iptar, nrec, itype, kpol, k1, k2, k3, k4, c)
    integer, intent(in) :: nch, npert, mk, ntarg, nrec
    integer, intent(in) :: iptar(*), itype(*), kpol(*), k1(*), k2(*), k3(*), k4(*)
    real(8), intent(in) :: c(*)
    call gpu_check(bsrg_dirac_assemble(gpu_ctx, nch, npert, mk, ntarg, iptar, nrec, itype, kpol, &
                                       k1, k2, k3, k4, c))
  end subroutine

  subroutine gpu_dirac_get(hch, hcp, hp, acf, htarg)
    real(8), intent(out) :: hch(*), hcp(*), hp(*), acf(*), htarg(*)
    call gpu_check(bsrg_dirac_get(gpu_ctx, hch, hcp, hp, acf, htarg))
  end subroutine

  ! DBSR's overlap pass (sub1.f90:77-79, O_data.f90:2-90) into S
  subroutine gpu_dirac_overlap(nch, npert, ntarg, iptar, nrec, itype, k1, k2, cdata, eps_c, iitar)
    integer, intent(in) :: nch, npert, ntarg, nrec, iitar
    integer, intent(in) :: iptar(*), itype(*), k1(*), k2(*)
    real(8), intent(in) :: cdata(*), eps_c
    call gpu_check(bsrg_dirac_overlap(gpu_ctx, nch, npert, ntarg, iptar, nrec, itype, k1, k2, cdata, eps_c, iitar))
  end subroutine

  ! Zero_cond + Diag_channels (diag.f90:2-81) with the reference's defaults
  subroutine gpu_dirac_diag(lch, kch, ipsol, nsol)
    integer, intent(in) :: lch(*), kch(*)
    integer, intent(out) :: ipsol(*), nsol
    real(8), parameter :: c_au = 137.035999084d0
    call gpu_check(bsrg_dirac_diag_channels(gpu_ctx, lch, kch, -1, -1, 1, 1, -2.d0 * c_au * c_au, 1.d10, 1.d-3, &
                                            ipsol, nsol))
  end subroutine

  ! Record_matrix of the transformed S (which = 0, starts the file) or H
  ! (which = 1, appends H, mk and the asymptotic coefficients)
  subroutine gpu_dirac_record(path, which, nelc_core)
    character(*), intent(in) :: path
    integer, intent(in) :: which, nelc_core
    call gpu_check(bsrg_dirac_record(gpu_ctx, trim(path) // c_null_char, which, nelc_core))
  end subroutine

  ! ---- BSR_MAT4 interrupted runs -----------------------------------------
  subroutine gpu_matrix_record(path, interrupt, intercase, max_nbuf)
    character(*), intent(in) :: path
    integer, intent(in) :: interrupt, intercase, max_nbuf
    call gpu_check(bsrg_matrix_record(gpu_ctx, trim(path) // c_null_char, interrupt, intercase, max_nbuf))
  end subroutine

  subroutine gpu_matrix_resume(path, interrupt, intercase, max_nbuf)
    character(*), intent(in) :: path
    integer, intent(out) :: interrupt, intercase, max_nbuf
    integer(c_int) :: a, b, c
    call gpu_check(bsrg_matrix_resume(gpu_ctx, trim(path) // c_null_char, a, b, c))
    interrupt = a; intercase = b; max_nbuf = c
  end subroutine

  ! ---- list builder (State_res) -------------------------------------------
  ! a builder over the bank at path; sizes5 = ndet, kdet, ndef, kdef, nrec
  function gpu_lb_open(path, sizes5) result(lb)
    character(*), intent(in) :: path
    integer, intent(out) :: sizes5(5)
    type(c_ptr) :: lb
    integer(c_int) :: sz(5)
    lb = bsrg_lb_create()
    if (.not. c_associated(lb)) call gpu_check(-1_c_int)
    call gpu_check(bsrg_lb_read_bank(lb, trim(path) // c_null_char, sz))
    sizes5 = sz
  end function

  ! the record loop of State_res for icase; sizes3 = coefficients, list
  ! records, lists; then bsrg_lb_lists(lb, ...) with arrays of sizes3(2)
  subroutine gpu_lb_run(lb, icase, mk, sizes3)
    type(c_ptr), intent(in) :: lb
    integer, intent(in) :: icase, mk
    integer, intent(out) :: sizes3(3)
    integer(c_int) :: sz(3)
    call gpu_check(bsrg_lb_run(lb, icase, mk, sz))
    sizes3 = sz
  end subroutine

end module bsr_gpu
