!======================================================================
!  bsr_mat_gpu -- Fortran host driver for one partial wave.
!
!  Reads a wave description (stream file written by bsr_amd.jobfile),
!  runs the SUB1 sequence of src/bsr/BSR_MAT3/sub1.f90:70-219 through the
!  bsr_gpu module (overlaps, core shift, I_data for every list, target
!  energies, symmetrisation, ACF) and writes bsr_mat.nnn with Fortran
!  unformatted records in the layout of Record_matrix
!  (src/bsr/BSR_MAT3/record_matrix.f90:2-58).
!
!  usage: bsr_mat_gpu <wave file> <bsr_mat output>
!======================================================================
!  The lists reach the device through the drop-in I_data, L_data and
!  O_data of bsr_gpu_dropin.f90, which read them from module cmdata as the
!  reference's do; the two modules below hold the names those read
!  (mod_cmdata.f90:12-17, mod_param.f90:33,65).
module cmdata
  implicit none
  integer :: ncdata = 0
  real(8), allocatable :: CDATA(:)
  integer, allocatable :: K1(:), K2(:), K3(:), K4(:), IPT(:)
end module cmdata

module bsr_mat
  implicit none
  integer :: iitar = 0
  real(8) :: Eps_c = 1.0d-10
end module bsr_mat

program bsr_mat_gpu
  use bsr_gpu
  use cmdata, only: ncdata, CDATA, K1, K2, K3, K4, IPT
  use bsr_mat, only: Eps_c_m => Eps_c, iitar_m => iitar
  implicit none
  integer(c_int) :: ks, ns, ml, me, nbf, nch, ncp, mk, ntarg, nlists, novl, nllists, kclosd, iitar, orth
  real(c_double) :: h, ec, eps_c, eps_ovl, z
  real(c_double), allocatable :: t(:), pbs(:,:), etarg(:), oc(:)
  integer(c_int), allocatable :: iech(:), lbs(:), ibort(:,:), iptar(:), ok3(:), ok4(:)
  real(c_double), allocatable :: hcc(:,:,:), hcb(:,:,:), hbb(:), acf(:,:,:), htarg(:), otarg(:)
  character(512) :: fin, fout
  integer :: nu, nw, il, jtype, kpol, n, ich, i, nij
  integer(c_int) :: kcore

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=nu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(nu) ks, ns, ml, me, nbf, nch, ncp, mk, ntarg, nlists, novl
  read(nu) h, ec, eps_c, eps_ovl
  allocate(t(ns+ks), pbs(ns,nbf), iech(nbf), lbs(nbf), ibort(nbf,nbf), iptar(nch), etarg(ntarg))
  read(nu) t, pbs, iech, lbs, ibort, iptar, etarg
  allocate(ok3(novl), ok4(novl), oc(novl))
  read(nu) ok3, ok4, oc
  read(nu) nllists, kclosd, iitar, orth, z

  nij = nch*(nch+1)/2
  allocate(hcc(ns,ns,nij), hcb(ns,nch,max(ncp,1)), hbb(max(ncp*(ncp+1)/2,1)), &
           acf(nch,nch,0:mk), htarg(nij), otarg(nij))

  call gpu_init(0)
  call gpu_grid(ks, ns, ml, me, h, t)
  call gpu_orbitals(ns, nbf, pbs, iech, lbs, ibort, eps_ovl)
  call gpu_allocate_matrix(nch, ncp, mk)

  ! overlap matrix
  do ich = 1, nch
    call gpu_update_hl(ich, ich, 1.d0)
  end do
  Eps_c_m = eps_c
  iitar_m = iitar
  if (novl > 0) then
    ncdata = novl
    allocate(K3(novl), K4(novl), CDATA(novl), IPT(novl))
    K3 = ok3; K4 = ok4; CDATA = oc
    IPT = [(i, i = 1, novl)]
    call O_data
    deallocate(K3, K4, CDATA, IPT)
  end if
  call gpu_check(bsrg_symmetrize_diag(gpu_ctx))
  call gpu_check(bsrg_get_matrix(gpu_ctx, hcc, hcb, hbb, acf, htarg, otarg))
  open(newunit=nw, file=trim(fout), form='unformatted', status='replace')
  write(nw) ns, nch, ncp
  call record_matrix(nw)

  ! interaction matrix: EC*S, L-integrals (Gen_Lval, then icase 6), R^k
  call gpu_check(bsrg_scale(gpu_ctx, ec))
  if (nllists > 0) call gpu_gen_lval(z, kclosd)
  do il = 1, nllists
    read(nu) jtype, kpol, n
    ncdata = n
    allocate(K1(n), K2(n), K3(n), K4(n), CDATA(n), IPT(n))
    read(nu) K1, K2, K3, K4, CDATA
    IPT = [(i, i = 1, n)]
    call L_data(jtype, kpol)
    deallocate(K1, K2, K3, K4, CDATA, IPT)
  end do
  do il = 1, nlists
    read(nu) jtype, kpol, n
    ncdata = n
    allocate(K1(n), K2(n), K3(n), K4(n), CDATA(n), IPT(n))
    read(nu) K1, K2, K3, K4, CDATA
    IPT = [(i, i = 1, n)]
    call I_data(jtype, kpol)
    deallocate(K1, K2, K3, K4, CDATA, IPT)
  end do
  close(nu)
  do ich = 1, nch
    call gpu_update_hl(ich, ich, etarg(iptar(ich)) - ec)
  end do
  call gpu_check(bsrg_symmetrize_diag(gpu_ctx))
  if (orth /= 0) call gpu_check(bsrg_bs_orth(gpu_ctx, 0))
  ! core screening of ACF(i,i,0): 2(4l+2) per closed shell (sub1.f90:218-219)
  kcore = 0
  do i = 1, kclosd
    kcore = kcore + 2 * (4 * lbs(i) + 2)
  end do
  call gpu_check(bsrg_acf_finish(gpu_ctx, kcore))
  call gpu_check(bsrg_get_matrix(gpu_ctx, hcc, hcb, hbb, acf, htarg, otarg))
  call record_matrix(nw)
  write(nw) mk
  write(nw) acf
  write(nw) t(ns+1), ns
  write(nw) htarg
  write(nw) otarg
  write(nw) etarg
  write(nw) ec
  close(nw)
  call gpu_finish()

contains

  subroutine record_matrix(nu_out)
    integer, intent(in) :: nu_out
    integer :: ib, ich1, jch1, ip, jp, idiag
    idiag = 1
    do ib = 1, nch
      write(nu_out) hcc(:,:,ib*(ib+1)/2)
    end do
    do ich1 = 1, nch
      do jch1 = 1, ich1 - 1
        ib = ich1*(ich1-1)/2 + jch1
        if (sum(abs(hcc(:,:,ib))) == 0.d0) cycle
        write(nu_out) ich1, jch1
        write(nu_out) hcc(:,:,ib)
        idiag = 0
      end do
    end do
    if (ncp > 0) then
      do ich1 = 1, nch
        do ip = 1, ncp
          if (sum(abs(hcb(:,ich1,ip))) == 0.d0) cycle
          write(nu_out) ip + nch, ich1
          write(nu_out) hcb(:,ich1,ip)
          idiag = 0
        end do
      end do
      do ip = 1, ncp
        do jp = 1, ip
          ib = ip*(ip-1)/2 + jp
          if (hbb(ib) == 0.d0) cycle
          write(nu_out) ip + nch, jp + nch
          write(nu_out) hbb(ib)
          if (ip /= jp) idiag = 0
        end do
      end do
    end if
    write(nu_out) 0, idiag
  end subroutine

end program bsr_mat_gpu
