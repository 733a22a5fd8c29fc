# Builds the reference's own Fortran sources, where they lie, into oracle/_ref/
# (objects, module files and the ref_harness executable).  Test infrastructure:
# the outputs are the oracle's pin, never part of the product.
#   make -f oracle/ref.mk            (from the repo root)
# Needs amdflang (ROCm) and BLAS/LAPACK from MKL (/opt/conda/lib/libmkl_rt.so).
# Leaves out main_IMSRG, EOM_IMSRG, response and deuteron (the only ARPACK users).
REFDIR ?= /root/reference/im-SRG
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref
FC := amdflang
FFLAGS ?= -O2 -fopenmp
MKLDIR ?= /opt/conda/lib
LIBS := -L$(MKLDIR) -Wl,-rpath,$(MKLDIR) -lmkl_rt -lz

# module dependency order
MODS := gzipmod basic_IMSRG cross_coupled commutators isospin_operators \
        three_body_routines HF_mod TS_commutators tensor_products gaulag operators \
        operator_commutators IMSRG_MAGNUS EOM_scalar_commutators EOM_TS_commutators \
        EOM_dTZ_commutators brute_force_testing adams_ode IMSRG_ODE
OBJS := $(OUT)/anglib.o $(addprefix $(OUT)/,$(addsuffix .o,$(MODS))) $(OUT)/generators.o $(OUT)/interaction_IO.o

LIBDIR := $(abspath $(HERE)../eom-imsrg_full_amd/imsrg_mi355x/lib)

all: $(OUT)/ref_harness $(if $(wildcard $(LIBDIR)/libimsrg_mi355x.so),$(OUT)/drop_in_test $(OUT)/sub_harness)

$(OUT):
	@mkdir -p $(OUT)

$(OUT)/anglib.o: $(REFDIR)/anglib.f | $(OUT)
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)

# serialised: each module needs the .mod files of the ones before it
$(OUT)/gzipmod.o: $(REFDIR)/gzipmod.f90 | $(OUT)
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/basic_IMSRG.o: $(REFDIR)/basic_IMSRG.f90 $(OUT)/gzipmod.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/cross_coupled.o: $(REFDIR)/cross_coupled.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/commutators.o: $(REFDIR)/commutators.f90 $(OUT)/cross_coupled.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/isospin_operators.o: $(REFDIR)/isospin_operators.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/three_body_routines.o: $(REFDIR)/three_body_routines.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/HF_mod.o: $(REFDIR)/HF_mod.f90 $(OUT)/isospin_operators.o $(OUT)/three_body_routines.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/TS_commutators.o: $(REFDIR)/TS_commutators.f90 $(OUT)/cross_coupled.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/tensor_products.o: $(REFDIR)/tensor_products.f90 $(OUT)/cross_coupled.o $(OUT)/isospin_operators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/gaulag.o: $(REFDIR)/gaulag.f90 | $(OUT)
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/operators.o: $(REFDIR)/operators.f90 $(OUT)/gaulag.o $(OUT)/tensor_products.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/operator_commutators.o: $(REFDIR)/operator_commutators.f90 $(OUT)/cross_coupled.o $(OUT)/isospin_operators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/IMSRG_MAGNUS.o: $(REFDIR)/IMSRG_MAGNUS.f90 $(OUT)/HF_mod.o $(OUT)/TS_commutators.o $(OUT)/commutators.o $(OUT)/operators.o $(OUT)/operator_commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/EOM_scalar_commutators.o: $(REFDIR)/EOM_scalar_commutators.f90 $(OUT)/cross_coupled.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/EOM_TS_commutators.o: $(REFDIR)/EOM_TS_commutators.f90 $(OUT)/cross_coupled.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/EOM_dTZ_commutators.o: $(REFDIR)/EOM_dTZ_commutators.f90 $(OUT)/cross_coupled.o $(OUT)/isospin_operators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/brute_force_testing.o: $(REFDIR)/brute_force_testing.f90 $(OUT)/EOM_scalar_commutators.o $(OUT)/EOM_TS_commutators.o $(OUT)/EOM_dTZ_commutators.o $(OUT)/TS_commutators.o $(OUT)/operator_commutators.o $(OUT)/operators.o $(OUT)/commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/adams_ode.o: $(REFDIR)/adams_ode.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/IMSRG_ODE.o: $(REFDIR)/IMSRG_ODE.f90 $(OUT)/adams_ode.o $(OUT)/operators.o $(OUT)/commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/generators.o: $(REFDIR)/generators.f90 $(OUT)/commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)
$(OUT)/interaction_IO.o: $(REFDIR)/interaction_IO.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)

$(OUT)/harness_basis.o: $(HERE)harness_basis.f90 $(OUT)/basic_IMSRG.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)

$(OUT)/ref_harness: $(HERE)ref_harness.f90 $(OBJS) $(OUT)/harness_basis.o
	$(FC) $(FFLAGS) -o $@ $(HERE)ref_harness.f90 $(OBJS) $(OUT)/harness_basis.o -J$(OUT) $(LIBS)

# the product's Fortran drop-in, compiled against the reference modules, and
# its test program (reference routines vs drop-in on the same sq_op inputs)
$(OUT)/imsrg_gpu.o: $(HERE)../eom-imsrg_full_amd/fortran/imsrg_gpu.f90 $(OUT)/cross_coupled.o $(OUT)/three_body_routines.o
	$(FC) $(FFLAGS) -c -o $@ $< -J$(OUT)

$(OUT)/drop_in_test: $(HERE)../tests/fortran/drop_in_test.f90 $(OBJS) $(OUT)/harness_basis.o $(OUT)/imsrg_gpu.o $(LIBDIR)/libimsrg_mi355x.so
	$(FC) $(FFLAGS) -o $@ $(HERE)../tests/fortran/drop_in_test.f90 $(OBJS) $(OUT)/harness_basis.o \
	    $(OUT)/imsrg_gpu.o -J$(OUT) $(LIBS) -L$(LIBDIR) -limsrg_mi355x \
	    -Wl,-rpath,'$$ORIGIN/../../eom-imsrg_full_amd/imsrg_mi355x/lib'

# The link-time substitutes: the product's `module commutators`
# (eom-imsrg_full_amd/fortran/commutators_gpu.f90) and `module TS_commutators`
# (ts_commutators_gpu.f90) in place of the reference's, and the reference's
# callers of those modules -- IMSRG_MAGNUS, generators, IMSRG_ODE,
# brute_force_testing -- compiled unmodified against them into
# $(SUB); ref_harness.f90 linked the same way (with -DSUBSTITUTE, which adds
# the device generators to its generator choices) is sub_harness.
SUB := $(OUT)/sub
SUBFLAGS := -J$(SUB) -I$(SUB) -I$(OUT)
SUBOBJS := $(OUT)/anglib.o $(addprefix $(OUT)/,$(addsuffix .o,gzipmod basic_IMSRG cross_coupled \
        isospin_operators three_body_routines HF_mod tensor_products gaulag operators \
        operator_commutators EOM_scalar_commutators EOM_TS_commutators EOM_dTZ_commutators adams_ode \
        interaction_IO)) \
        $(addprefix $(SUB)/,commutators.o TS_commutators.o IMSRG_MAGNUS.o brute_force_testing.o IMSRG_ODE.o \
        generators.o)

$(SUB)/commutators.o: $(HERE)../eom-imsrg_full_amd/fortran/commutators_gpu.f90 $(OUT)/cross_coupled.o
	@mkdir -p $(SUB)
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)
$(SUB)/TS_commutators.o: $(HERE)../eom-imsrg_full_amd/fortran/ts_commutators_gpu.f90 $(OUT)/cross_coupled.o
	@mkdir -p $(SUB)
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)
$(SUB)/IMSRG_MAGNUS.o: $(REFDIR)/IMSRG_MAGNUS.f90 $(SUB)/commutators.o $(OUT)/HF_mod.o $(SUB)/TS_commutators.o $(OUT)/operators.o $(OUT)/operator_commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)
$(SUB)/brute_force_testing.o: $(REFDIR)/brute_force_testing.f90 $(SUB)/commutators.o $(OUT)/EOM_scalar_commutators.o $(OUT)/EOM_TS_commutators.o $(OUT)/EOM_dTZ_commutators.o $(SUB)/TS_commutators.o $(OUT)/operator_commutators.o $(OUT)/operators.o
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)
$(SUB)/IMSRG_ODE.o: $(REFDIR)/IMSRG_ODE.f90 $(SUB)/commutators.o $(OUT)/adams_ode.o $(OUT)/operators.o
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)
$(SUB)/generators.o: $(REFDIR)/generators.f90 $(SUB)/commutators.o
	$(FC) $(FFLAGS) -c -o $@ $< $(SUBFLAGS)

$(OUT)/sub_harness: $(HERE)ref_harness.f90 $(SUBOBJS) $(OUT)/harness_basis.o $(LIBDIR)/libimsrg_mi355x.so
	$(FC) $(FFLAGS) -cpp -DSUBSTITUTE -o $@ $(HERE)ref_harness.f90 $(SUBOBJS) $(OUT)/harness_basis.o \
	    $(SUBFLAGS) $(LIBS) -L$(LIBDIR) -limsrg_mi355x \
	    -Wl,-rpath,'$$ORIGIN/../../eom-imsrg_full_amd/imsrg_mi355x/lib'

clean:
	rm -rf $(OUT)

.PHONY: all clean
