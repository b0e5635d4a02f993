# Partial build of the reference's own C sources, used only to pin the
# oracle.  Outputs go to oracle/_ref/ (git-ignored).  Sources are compiled
# where they lie under /root/reference; nothing is copied.
#
# Compiled: fp_shared_core.c FP_Coef.c CONSTANTS.c params.c tracer_fp_nsub.c DSA_MODELS.c
#   tracer_fp_dsa_reacc.c
# Not compilable here (GSL / MPI headers absent): Chang_Cooper.c HADRONIC.c
#   Synchrotron.c EMISSION.c tracer_fp_batch.c tracer_fp_setup.c tracer_fp_dsa.c
# The wrappers in ref_harness.c reach only the compiled functions; the link
# keeps only what they reach (--gc-sections) and fails on anything missing
# (-z defs).  One library per compile-time grid size.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
SIZES := 64 128 256 512
SRCS := $(REF)/src/fp_shared_core.c $(REF)/src/FP_Coef.c $(REF)/src/CONSTANTS.c \
        $(REF)/src/params.c $(REF)/src/tracer_fp_nsub.c $(REF)/src/DSA_MODELS.c \
        $(REF)/src/tracer_fp_dsa_reacc.c
CFLAGS_REF := -O2 -std=gnu11 -fPIC -fopenmp -ffp-contract=off -w \
              -ffunction-sections -fdata-sections -I$(REF)/include

LIBS := $(foreach n,$(SIZES),$(OUT)/libcroma_ref_n$(n).so)
LAYOUTS := $(foreach n,$(SIZES),$(OUT)/ref_layout_n$(n))

all: $(LIBS) $(LAYOUTS)

# struct layouts of the reference's GPU-backend headers (headers only)
$(OUT)/ref_layout_n%: $(HERE)ref_layout.c
	@mkdir -p $(OUT)
	gcc -std=gnu11 -w -Dnp=$* -Dnpe=$* -I$(REF)/include -o $@ $<

$(OUT)/libcroma_ref_n%.so: $(SRCS) $(HERE)ref_harness.c
	@mkdir -p $(OUT)
	gcc $(CFLAGS_REF) -Dnp=$* -Dnpe=$* -fvisibility=hidden -shared \
	    -Wl,--gc-sections -Wl,-z,defs -o $@ $(HERE)ref_harness.c $(SRCS) -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
