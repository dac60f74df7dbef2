# Builds oracle/_ref/libbsr_ref.so from the reference's own Fortran sources
# (read in place under REF, never copied) plus oracle/ref_driver.f90, with
# amdflang.  Test infrastructure only.  LTO plus a version script that
# exports only the ref_* entry points lets the linker drop routines this
# driver never reaches (Target_new in mod_intmat.f90 calls LAPACK, which
# the image does not have); nothing is stubbed.
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
REF ?= /root/reference/src
FC := amdflang
OUT := $(HERE)_ref
FFLAGS := -O2 -fPIC -flto -module-dir $(OUT)/mod

SRCS := \
  $(REF)/libs/BS/spline_param.f90 $(REF)/libs/BS/spline_grid.f90 \
  $(REF)/libs/BS/spline_atomic.f90 $(REF)/libs/BS/spline_moments.f90 \
  $(REF)/libs/BS/spline_integrals.f90 $(REF)/libs/BS/spline_galerkin.f90 \
  $(REF)/libs/BS/spline_orbitals.f90 $(REF)/libs/ZCOM/read_par.f90 \
  $(REF)/libs/BS/define_grid.f90 $(REF)/libs/BS/gauss.f90 \
  $(REF)/libs/BS/vbsplvd.f90 $(REF)/libs/BS/mrm.f90 \
  $(REF)/libs/BS/moments.f90 $(REF)/libs/BS/rk_moments.f90 \
  $(REF)/libs/BS/mrk_cell.f90 $(REF)/libs/BS/density.f90 \
  $(REF)/libs/BS/convol.f90 $(REF)/libs/BS/sum_amb.f90 \
  $(REF)/libs/BS/bav.f90 $(REF)/libs/BS/bxv.f90 $(REF)/libs/BS/bvmv.f90 \
  $(REF)/libs/BS/quadr.f90 $(REF)/bsr/BSR_MAT3/mod_param.f90 \
  $(REF)/libs/ZCONFLS/MOD_target.f90 $(REF)/libs/ZCONFLS/MOD_channel.f90 \
  $(REF)/bsr/BSR_MAT3/mod_intmat.f90 \
  $(REF)/libs/BS/int_de.f90 $(REF)/libs/BS/azl.f90 $(REF)/libs/ZCOM/ipointer.f90 \
  $(REF)/libs/RADIAL/zcb.f90 $(REF)/bsr/BSR_MAT3/pri_int.f90 $(REF)/bsr/BSR_MAT3/l_core.f90 \
  $(REF)/bsr/BSR_MAT3/mod_cmdata.f90 $(REF)/bsr/BSR_MAT3/add_cdata.f90 $(REF)/bsr/BSR_MAT3/merge_cdata.f90

OBJS := $(addprefix $(OUT)/,$(notdir $(SRCS:.f90=.o)))

# Dirac B-spline routines and the Dirac assembly of DBSR_MAT3 (R_data,
# O_data, updates, Target_h, Zero_cond, transform_matrix, Record_matrix with the modules
# they read): a library of their own
# (routine names such as convol and density repeat those of the BS library)
DSRCS := \
  $(REF)/libs/ZCOM/MOD_zconst.f90 $(REF)/libs/ZCOM/read_par.f90 $(REF)/libs/ZCOM/rrtc.f90 \
  $(REF)/libs/ZCOM/check_file.f90 $(REF)/libs/ZCOM/gauss_points.f90 \
  $(REF)/libs/DBS/DBS_nuclear.f90 $(REF)/libs/DBS/DBS_grid.f90 \
  $(REF)/libs/DBS/DBS_gauss.f90 $(REF)/libs/DBS/DBS_moments.f90 \
  $(REF)/libs/DBS/DBS_integrals.f90 $(REF)/libs/DBS/DBS_debug.f90 \
  $(REF)/libs/DBS/zbsplvd.f90 $(REF)/libs/DBS/zinty.f90 \
  $(REF)/libs/DBS/mrk_pppp.f90 $(REF)/libs/DBS/mrk_qqqq.f90 \
  $(REF)/libs/DBS/mrk_pqpq.f90 $(REF)/libs/DBS/mrk_qpqp.f90 \
  $(REF)/libs/DBS/convol.f90 $(REF)/libs/DBS/density.f90 \
  $(REF)/libs/DBS/DBS_orbitals_pq.f90 $(REF)/libs/DBS/quadr_pq.f90 \
  $(REF)/libs/DBS/sum_amb.f90 $(REF)/libs/DBS/bav.f90 \
  $(REF)/libs/ZCONFJJ/mod_target_jj.f90 $(REF)/libs/ZCONFJJ/mod_orb_jj.f90 \
  $(REF)/libs/ZCONFJJ/mod_conf_jj.f90 $(REF)/libs/ZCONFJJ/mod_channel_jj.f90 \
  $(REF)/dbsr/DBSR_MAT3_MPI/mod_dbsr_mat.f90 $(REF)/dbsr/DBSR_MAT3_MPI/MOD_c_data.f90 \
  $(REF)/dbsr/DBSR_MAT3_MPI/updates.f90 $(REF)/dbsr/DBSR_MAT3_MPI/target_subs.f90 \
  $(REF)/dbsr/DBSR_MAT3_MPI/R_data.f90 $(REF)/dbsr/DBSR_MAT3_MPI/O_data.f90 \
  $(REF)/dbsr/DBSR_MAT3_MPI/zero_cond.f90 $(REF)/dbsr/DBSR_MAT3_MPI/transform_matrix.f90 \
  $(REF)/dbsr/DBSR_MAT3_MPI/record_matrix.f90
# The BSR_MAT3 list builder's routines (Read_dets with the int_bnk header
# readers, Decode_INT, Ifind_channel, Jsym_int, Det_fact, Check_det, VDET,
# DET, Idef_itype) with the modules they read: a library of their own,
# built with the libs_03 spline_orbitals module, whose IBORT and OBS arrays
# check_det.f90 and vdet.f90 use
LSRCS := \
  $(REF)/libs_03/BS/spline_param.f90 $(REF)/libs_03/BS/spline_orbitals.f90 \
  $(REF)/bsr/BSR_MAT3/mod_param.f90 $(REF)/libs/ZCONFLS/MOD_param_LS.f90 \
  $(REF)/libs/ZCONFLS/MOD_symc.f90 $(REF)/libs/ZCONFLS/MOD_symt.f90 $(REF)/libs/ZCONFLS/MOD_orb_LS.f90 \
  $(REF)/libs/ZCONFLS/MOD_target.f90 $(REF)/libs/ZCONFLS/MOD_channel.f90 $(REF)/libs/ZCONFLS/incode_int.f90 \
  $(REF)/libs/ZCOM/read_par.f90 $(REF)/libs/ZCOM/stop_mp.f90 $(REF)/libs/ZCOM/det.f90 \
  $(REF)/bsr/BSR_MAT3/mod_cmdata.f90 $(REF)/bsr/BSR_MAT3/mod_dets.f90 $(REF)/bsr/BSR_MAT3/mod_ndef.f90 \
  $(REF)/bsr/BSR_MAT3/mod_ndet.f90 $(REF)/bsr/BSR_MAT3/det_fact.f90 $(REF)/bsr/BSR_MAT3/check_det.f90 \
  $(REF)/bsr/BSR_MAT3/vdet.f90 $(REF)/bsr/BSR_MAT3/jsym_int.f90 $(REF)/bsr/BSR_MAT3/idef_type.f90
LOUT := $(OUT)/lb
LOBJS := $(addprefix $(LOUT)/,$(notdir $(LSRCS:.f90=.o)))

DOUT := $(OUT)/dbs
DOBJS := $(addprefix $(DOUT)/,$(notdir $(DSRCS:.f90=.o)))

all: $(OUT)/libbsr_ref.so $(OUT)/libdbs_ref.so $(OUT)/liblb_ref.so

$(OUT)/liblb_ref.so: $(LSRCS) $(HERE)ref_lb_driver.f90 $(HERE)ref.mk
	mkdir -p $(LOUT)/mod
	set -e; for f in $(LSRCS); do \
	  $(FC) -O2 -fPIC -flto -module-dir $(LOUT)/mod -c $$f -o $(LOUT)/$$(basename $$f .f90).o; done
	$(FC) -O2 -fPIC -flto -module-dir $(LOUT)/mod -c $(HERE)ref_lb_driver.f90 -o $(LOUT)/ref_lb_driver.o
	echo '{ global: ref_*; local: *; };' > $(LOUT)/exports.map
	$(FC) -flto -fuse-ld=lld -shared -o $@ $(LOBJS) $(LOUT)/ref_lb_driver.o \
	  -Wl,--version-script=$(LOUT)/exports.map

$(OUT)/libdbs_ref.so: $(DSRCS) $(HERE)ref_dbs_driver.f90 $(HERE)ref_dbsr_driver.f90 $(HERE)ref.mk
	mkdir -p $(DOUT)/mod
	set -e; for f in $(DSRCS); do \
	  $(FC) -O2 -fPIC -flto -module-dir $(DOUT)/mod -c $$f -o $(DOUT)/$$(basename $$f .f90).o; done
	$(FC) -O2 -fPIC -flto -module-dir $(DOUT)/mod -c $(HERE)ref_dbs_driver.f90 -o $(DOUT)/ref_dbs_driver.o
	$(FC) -O2 -fPIC -flto -module-dir $(DOUT)/mod -c $(HERE)ref_dbsr_driver.f90 -o $(DOUT)/ref_dbsr_driver.o
	echo '{ global: ref_*; local: *; };' > $(DOUT)/exports.map
	$(FC) -flto -fuse-ld=lld -shared -o $@ $(DOBJS) $(DOUT)/ref_dbs_driver.o $(DOUT)/ref_dbsr_driver.o \
	  -Wl,--version-script=$(DOUT)/exports.map

$(OUT)/libbsr_ref.so: $(SRCS) $(HERE)ref_driver.f90 $(HERE)ref.mk
	mkdir -p $(OUT)/mod
	set -e; for f in $(SRCS); do \
	  $(FC) $(FFLAGS) -c $$f -o $(OUT)/$$(basename $$f .f90).o; done
	$(FC) $(FFLAGS) -c $(HERE)ref_driver.f90 -o $(OUT)/ref_driver.o
	echo '{ global: ref_*; local: *; };' > $(OUT)/exports.map
	$(FC) -flto -fuse-ld=lld -shared -o $@ $(OBJS) $(OUT)/ref_driver.o \
	  -Wl,--version-script=$(OUT)/exports.map

.PHONY: all
