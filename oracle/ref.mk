# Builds oracle/_ref/libpgref.so from UNMODIFIED ParaGauss sources where they
# lie under the reference tree, plus oracle/ref_shim.f90 (C entry points).
# Only the standalone-compilable leaf modules are used: type_module, the
# autodiff instantiations ad2x2/ad2x3 (autodiff/module.mk:46-86 instantiates
# them by preprocessing autodiff2.f90 with ORD/NVAR/AUTODIFF; here the same
# values are passed as -D flags), pw_ldac_module, pbe_ggcxc_module,
# solhrules_module, solid_harmonics_module.  exchange.f90 is not buildable
# standalone (it uses vwnc -> cpksdervs_matrices -> occupied_levels_module ->
# the SCF/MPI state).  error_handler_ / assert_failed_ stay unresolved: they
# are reached only on invalid input, and the library is opened RTLD_LAZY.
#
#   make -f oracle/ref.mk REF=/root/reference
REF ?= /root/reference
OUT := oracle/_ref
FC := flang
FFLAGS := -cpp -I$(REF)/include -D_LINUX -O2 -fPIC -module-dir $(OUT)

OBJS := $(OUT)/type_module.o $(OUT)/ad2x2.o $(OUT)/ad2x3.o \
        $(OUT)/pw_ldac_module.o $(OUT)/pbe_ggcxc_module.o \
        $(OUT)/solhrules_module.o $(OUT)/solid_harmonics_module.o \
        $(OUT)/ref_shim.o

all: $(OUT)/libpgref.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/type_module.o: $(REF)/modules/type_module.f90 | $(OUT)
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/ad2x2.o: $(REF)/autodiff/autodiff2.f90 $(OUT)/type_module.o
	$(FC) $(FFLAGS) -DORD=2 -DNVAR=2 -DAUTODIFF=ad2x2 -c $< -o $@
$(OUT)/ad2x3.o: $(REF)/autodiff/autodiff2.f90 $(OUT)/type_module.o
	$(FC) $(FFLAGS) -DORD=2 -DNVAR=3 -DAUTODIFF=ad2x3 -c $< -o $@
$(OUT)/pw_ldac_module.o: $(REF)/modules/pw_ldac_module.f90 $(OUT)/ad2x2.o
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/pbe_ggcxc_module.o: $(REF)/modules/pbe_ggcxc_module.f90 $(OUT)/ad2x3.o $(OUT)/pw_ldac_module.o
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/solhrules_module.o: $(REF)/modules/solhrules_module.f90 $(OUT)/type_module.o
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/solid_harmonics_module.o: $(REF)/modules/solid_harmonics_module.f90 $(OUT)/solhrules_module.o
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/ref_shim.o: oracle/ref_shim.f90 $(OUT)/pbe_ggcxc_module.o $(OUT)/solid_harmonics_module.o
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/libpgref.so: $(OBJS)
	$(FC) -shared -fPIC -o $@ $(OBJS)

clean:
	rm -rf $(OUT)
.PHONY: all clean
