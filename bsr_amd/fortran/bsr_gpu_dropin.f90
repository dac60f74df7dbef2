!======================================================================
!  Drop-in bodies for BSR_MAT3's list consumers, with the reference's own
!  signatures.  Gen_matrix (src/bsr/BSR_MAT3/gen_matrix.f90:52-56) calls
!
!      Call I_data(jtype,jpol)   ! icase 1-5, 8-10
!      Call L_data(jtype,jpol)   ! icase 6
!      Call O_data               ! icase 11
!
!  and those read the merged list from module cmdata (ncdata, IPT,
!  K1..K4, CDATA; mod_cmdata.f90:12-17) and the cut Eps_c and the target
!  flag iitar from module bsr_mat (mod_param.f90:33,65).  This file
!  replaces I_data.f90, L_data.f90 and O_data.f90 in BSR_MAT3's build: the
!  same three subroutines hand the list to the device accumulators through
!  module bsr_gpu, so no call site changes.  It compiles against the
!  reference's own modules cmdata and bsr_mat.
!======================================================================

! I_data.f90:2 -- R^k lists of structure jtype and multipole jpol
Subroutine I_data(jtype, jpol)
  Use cmdata, only: ncdata, IPT, K1, K2, K3, K4, CDATA
  Use bsr_gpu, only: gpu_i_data
  Implicit none
  Integer, intent(in) :: jtype, jpol
  if (ncdata <= 0) Return
  Call gpu_i_data(jtype, jpol, ncdata, IPT, K1, K2, K3, K4, CDATA)
End Subroutine I_data

! L_data.f90:2 -- L-integral lists of structure jtype and orbital l
Subroutine L_data(jtype, l)
  Use cmdata, only: ncdata, IPT, K1, K2, K3, K4, CDATA
  Use bsr_mat, only: Eps_c, iitar
  Use bsr_gpu, only: gpu_l_data
  Implicit none
  Integer, intent(in) :: jtype, l
  if (ncdata <= 0) Return
  Call gpu_l_data(jtype, l, ncdata, IPT, K1, K2, K3, K4, CDATA, iitar, Eps_c)
End Subroutine L_data

! O_data.f90:2 -- overlap lists (-K3, -K4: the two sides of each overlap, O_data.f90:32)
Subroutine O_data
  Use cmdata, only: ncdata, IPT, K3, K4, CDATA
  Use bsr_mat, only: Eps_c, iitar
  Use bsr_gpu, only: gpu_o_data
  Implicit none
  ! the device overlap pass adds the target structure <i|j> to Target_h
  ! only; O_data.f90:66 also adds c*sb to off-diagonal blocks when iitar > 1
  if (iitar > 1) Stop ' O_data: iitar > 1 is not supported on the device '
  if (ncdata <= 0) Return
  Call gpu_o_data(ncdata, IPT, K3, K4, CDATA, Eps_c)
End Subroutine O_data
