# Links the reference's own Fortran driver (tea.f90 and the modules it uses,
# compiled from the sources where they lie under $(REF)) against
# libtealeaf_mi355x.so in place of the reference's OpenCL C++ objects: the
# drop-in test of the C-ABI (the link step a maintainer of the reference makes).
# Test infrastructure only; outputs go to oracle/_ref/ and nothing of the
# reference is copied into the repository.
#
#   make -f oracle/fortran_driver.mk        (after the library is built)
#
# Needs amdflang (ROCm) and MPICH's Fortran bindings (/opt/conda); the binary
# finds the library through an $ORIGIN-relative run path, and the system
# libstdc++ (which the HIP runtime needs) before MPICH's older one.
REF      ?= /root/reference
MPI      ?= /opt/conda
FORT     ?= /opt/rocm/bin/amdflang
CCOMP    ?= gcc
HERE     := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)/_ref
OBJ      := $(OUT)/fobj
LIBDIR   := $(HERE)/../tealeaf_opencl_amd/lib
FFLAGS   := -O2 -I$(MPI)/include -module-dir $(OBJ)

# module dependency order (USE statements of the reference sources)
UNITS := data definitions global_mpi report tea timer parse read_input initialise_chunk build_field \
         update_halo set_field start generate_chunk initialise field_summary calc_dt timestep \
         tea_leaf_common tea_leaf_cg tea_leaf_cheby tea_leaf_ppcg tea_leaf_jacobi tea_solve tea_leaf diffuse
OBJS  := $(patsubst %,$(OBJ)/%.o,$(UNITS)) $(OBJ)/timer_c.o

all: $(OUT)/tea_leaf_mi355x

$(OBJ):
	mkdir -p $(OBJ)

# one unit at a time, in order, so that each finds the modules it uses
$(OBJ)/units.stamp: $(patsubst %,$(REF)/%.f90,$(UNITS)) | $(OBJ)
	set -e; for u in $(UNITS); do $(FORT) $(FFLAGS) -c $(REF)/$$u.f90 -o $(OBJ)/$$u.o; done
	touch $@

$(OBJ)/timer_c.o: $(REF)/timer_c.c | $(OBJ)
	$(CCOMP) -O2 -c $< -o $@

$(OUT)/tea_leaf_mi355x: $(OBJ)/units.stamp $(OBJ)/timer_c.o $(LIBDIR)/libtealeaf_mi355x.so
	$(FORT) -O2 $(OBJS) -L$(LIBDIR) -ltealeaf_mi355x -L$(MPI)/lib -lmpifort -lmpi -lstdc++ \
	    -Wl,-rpath,/usr/lib/x86_64-linux-gnu:'$$ORIGIN/../../tealeaf_opencl_amd/lib':$(MPI)/lib -o $@

clean:
	rm -rf $(OBJ) $(OUT)/tea_leaf_mi355x

.PHONY: all clean
