# Compiles the reference's module files (type and variable declarations of
# the hot path's tables, no executable code) from where they lie under the
# reference checkout, with amdflang, into oracle/_ref/ (objects and .mod
# files).  Used only to compile ggi_tlm_amd/fortran/gpu_scn_init.F90 against
# the reference's own module layouts in a CPU test.  The sequential build's
# configuration: no MPI, no Ngspice.
#   make -f oracle/ref_modules.mk [REF=/root/reference/GGI_TLM/SRC]
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
REF ?= /root/reference/GGI_TLM/SRC
OUT := $(HERE)_ref
AMDFLANG ?= /opt/rocm/bin/amdflang
FFLAGS = -O0 -cpp -J $(OUT)

# dependency order
SRCS = CONSTANTS/constants.F90 POLYNOMIAL/polynomial_types.F90 FILTER/filter_types.F90 \
       TLM_MODULES/Geometry_Modules.F90 TLM_MODULES/TLM_Modules.F90 TLM_MODULES/Mesh_Modules.F90 \
       TLM_MODULES/PML_module.F90 TLM_MODULES/TLM_material_modules.F90 \
       TLM_MODULES/TLM_excitation_modules.F90 TLM_MODULES/TLM_output_modules.F90
OBJS = $(addprefix $(OUT)/,$(notdir $(SRCS:.F90=.o)))

all: $(OUT)/ref_modules.stamp

$(OUT)/ref_modules.stamp: $(addprefix $(REF)/,$(SRCS))
	mkdir -p $(OUT)
	for f in $(SRCS); do \
	  $(AMDFLANG) $(FFLAGS) -c $(REF)/$$f -o $(OUT)/$$(basename $${f%.F90}).o || exit 1; \
	done
	touch $@

clean:
	rm -rf $(OUT)
