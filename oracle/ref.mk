# Test infrastructure: builds the reference EQdyna 5.3.3 time step, from the
# reference's own Fortran sources where they lie, into oracle/_ref/.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Uses amdflang (ROCm LLVM Flang) and the MPICH that ships in /opt/conda.
# src/eqdyna3d.f90 (main program) and src/netcdf_io.f90 (netCDF, absent from
# this image) are not compiled; oracle/ref_main.f90 replaces them.
# No FMA contraction and no -march, so the build is the plain FP64 oracle.
# A second, timing-only build (oracle/_ref/eqd_ref_o3: -O3, the compiler's
# default contraction) is the CPU baseline the benchmark times; it is never
# used for parity.
REFSRC ?= /root/reference/src
OUT    := oracle/_ref
OBJ    := $(OUT)/obj
FC     := amdflang
MPIINC := /opt/conda/include
MPILIB := /opt/conda/lib
FFLAGS := -O2 -ffp-contract=off -I$(MPIINC) -J$(OBJ) -I$(OBJ)

REFUNITS := globalvar driver comdampv assembleGlobalKU calcElemKU calcB calcElemMass \
            qconstant hrglss faulting fric thermop meshgen mesh4num assembleGlobalMass \
            calcGlobalShapeFunc calcLocalShapeFunc library library_degeneration \
            library_output readInputFiles warning
REFOBJS  := $(addprefix $(OBJ)/,$(addsuffix .o,$(REFUNITS)))

OBJ3    := $(OUT)/obj_o3
FFLAGS3 := -O3 -I$(MPIINC) -J$(OBJ3) -I$(OBJ3)
REFOBJS3 := $(addprefix $(OBJ3)/,$(addsuffix .o,$(REFUNITS)))

all: $(OUT)/eqd_ref $(OUT)/eqd_ref_o3

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/globalvar.o: $(REFSRC)/globalvar.f90 | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/%.o: $(REFSRC)/%.f90 $(OBJ)/globalvar.o | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/ref_main.o: oracle/ref_main.f90 $(OBJ)/globalvar.o | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/eqd_ref: $(REFOBJS) $(OBJ)/ref_main.o
	$(FC) -o $@ $^ -L$(MPILIB) -lmpifort -lmpi -Wl,-rpath,$(MPILIB)

$(OBJ3):
	mkdir -p $(OBJ3)

$(OBJ3)/globalvar.o: $(REFSRC)/globalvar.f90 | $(OBJ3)
	$(FC) $(FFLAGS3) -c $< -o $@

$(OBJ3)/%.o: $(REFSRC)/%.f90 $(OBJ3)/globalvar.o | $(OBJ3)
	$(FC) $(FFLAGS3) -c $< -o $@

$(OBJ3)/ref_main.o: oracle/ref_main.f90 $(OBJ3)/globalvar.o | $(OBJ3)
	$(FC) $(FFLAGS3) -c $< -o $@

$(OUT)/eqd_ref_o3: $(REFOBJS3) $(OBJ3)/ref_main.o
	$(FC) -o $@ $^ -L$(MPILIB) -lmpifort -lmpi -Wl,-rpath,$(MPILIB)

clean:
	rm -rf $(OUT)

.PHONY: all clean
