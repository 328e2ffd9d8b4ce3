# Builds the reference's own MPI-free Fortran modules, read in place from
# the reference tree, plus oracle/ref_bridge.f90 into oracle/_ref/libref_lepic.so.
# TEST INFRASTRUCTURE ONLY.  Strict IEEE: no fast-math, no FP contraction.
# Single-threaded (no -fopenmp) so every loop runs in its source order.
REF ?= /root/reference
FLANG ?= /opt/rocm/llvm/bin/amdflang
FFLAGS = -O2 -fPIC -ffp-contract=off
OUT = _ref
M = $(REF)/modules
SRCS = $(M)/core/mod_constants.f90 $(M)/core/mod_RNG.f90 $(M)/particles/mod_particles.f90 \
       $(M)/particles/mod_particle_sorting.f90 $(M)/particles/mod_particleMover.f90 \
       $(M)/particles/mod_particleBC.f90 $(M)/fields/mod_chargeDeposition.f90 \
       $(M)/fields/mod_charge_weights.f90 $(M)/fields/mod_electricField.f90 \
       $(M)/config/mod_config.f90 $(M)/geometry/mod_domain.f90 $(M)/fields/mod_magneticField.f90 \
       $(M)/io/mod_output_2d.f90
OBJS = $(patsubst %.f90,$(OUT)/%.o,$(notdir $(SRCS)))

all: $(OUT)/libref_lepic.so

$(OUT)/libref_lepic.so: $(SRCS) ref_bridge.f90
	@mkdir -p $(OUT)
	cd $(OUT) && for f in $(SRCS); do $(FLANG) $(FFLAGS) -c $$f || exit 1; done
	cd $(OUT) && $(FLANG) $(FFLAGS) -c ../ref_bridge.f90
	cd $(OUT) && $(FLANG) -shared -o libref_lepic.so $(notdir $(OBJS)) ref_bridge.o

clean:
	rm -rf $(OUT)

.PHONY: all clean
