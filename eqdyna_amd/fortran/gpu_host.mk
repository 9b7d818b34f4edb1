# Links the reference EQdyna host (its own Fortran sources, where they lie) with
# the GPU time step:  make -f eqdyna_amd/fortran/gpu_host.mk  (from the repo root)
#   -> eqdyna_amd/fortran/_build/eqdyna_gpu
# Left out: src/driver.f90 and the routines only the time loop calls (replaced
# by eqd_driver.f90 + libeqd.so), src/eqdyna3d.f90 and src/netcdf_io.f90
# (netCDF-Fortran is not in this image; eqd_host_main.f90 stands in, see its
# header).
REFSRC ?= /root/reference/src
OUT    := eqdyna_amd/fortran/_build
OBJ    := $(OUT)/obj
FC     := amdflang
MPIINC := /opt/conda/include
MPILIB := /opt/conda/lib
FFLAGS := -O2 -ffp-contract=off -I$(MPIINC) -J$(OBJ) -I$(OBJ)
HOSTUNITS := globalvar meshgen mesh4num assembleGlobalMass calcGlobalShapeFunc calcLocalShapeFunc \
             library library_degeneration library_output readInputFiles warning
HOSTOBJS  := $(addprefix $(OBJ)/,$(addsuffix .o,$(HOSTUNITS)))

all: $(OUT)/eqdyna_gpu

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/globalvar.o: $(REFSRC)/globalvar.f90 | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/%.o: $(REFSRC)/%.f90 $(OBJ)/globalvar.o | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/eqd_driver.o: eqdyna_amd/fortran/eqd_driver.f90 $(OBJ)/globalvar.o | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OBJ)/eqd_host_main.o: eqdyna_amd/fortran/eqd_host_main.f90 $(OBJ)/globalvar.o | $(OBJ)
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/eqdyna_gpu: $(HOSTOBJS) $(OBJ)/eqd_driver.o $(OBJ)/eqd_host_main.o eqdyna_amd/libeqd.so
	$(FC) -o $@ $(HOSTOBJS) $(OBJ)/eqd_driver.o $(OBJ)/eqd_host_main.o \
	    -L$(MPILIB) -lmpifort -lmpi -Wl,-rpath,$(MPILIB) \
	    -Leqdyna_amd -leqd -Wl,-rpath,'$$ORIGIN/../..'

clean:
	rm -rf $(OUT)

.PHONY: all clean
