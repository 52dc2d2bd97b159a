# Builds the QLens reference (read-only sources under REF) into oracle/_ref/.
#
# This is the checker's "real reference": the default native path of QLens
# (no readline, MKL, FFTW, UMFPACK, MUMPS or FITS), compiled with g++ from the
# reference's own source files, plus oracle/ref_harness.cpp, a driver that sets
# up a QLens object from a command script, runs the pixel inversion and dumps the
# results as raw little-endian arrays.  Nothing from the reference is copied: the
# objects are compiled from the files where they lie, outputs go to oracle/_ref/.
#
#   make -f oracle/ref.mk -j8          (from the repository root)
#
# Flags follow the reference Makefile (Makefile:18-31) with OpenMP switched on;
# commands.cpp is compiled without -O exactly as the reference does.  modelparams.h
# uses std::vector without including <vector>; newer libstdc++ no longer pulls it
# in transitively, so the standard headers are force-included on the command line.

REF    ?= /root/reference
OUT    := oracle/_ref
OBJDIR := $(OUT)/obj
CXX    := g++
# -ffp-contract=off: x86-64 baseline has no FMA, stated explicitly so that a
# host compiler default can never change the reference arithmetic.
OPTS       := -Wno-write-strings -w -O3 -std=c++11 -fopenmp -ffp-contract=off
OPTS_NO_OPT:= -Wno-write-strings -w -std=c++11 -fopenmp -ffp-contract=off
FLAGS      := -DUSE_OPENMP -include vector -include string

SRCS := profile sbprofile egrad models commands params modelparams lenscalc \
        lens imgsrch pixelgrid cg mcmchdr errors brent sort gauss romberg spline \
        trirectangle GregsMathHdr hyp_2F1 cosmo simplex powell mcmceval qlens
OBJS := $(addprefix $(OBJDIR)/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/ref_harness

$(OBJDIR):
	mkdir -p $(OBJDIR)

# the reference CLI's main() is renamed so the harness can provide its own
$(OBJDIR)/qlens.o: $(REF)/qlens.cpp | $(OBJDIR)
	$(CXX) $(OPTS) $(FLAGS) -Dmain=qlens_cli_main -I$(REF) -c $< -o $@

$(OBJDIR)/commands.o: $(REF)/commands.cpp | $(OBJDIR)
	$(CXX) $(OPTS_NO_OPT) $(FLAGS) -I$(REF) -c $< -o $@

$(OBJDIR)/%.o: $(REF)/%.cpp | $(OBJDIR)
	$(CXX) $(OPTS) $(FLAGS) -I$(REF) -c $< -o $@

# the harness reads private grid state after an inversion (white-box checker only)
$(OBJDIR)/ref_harness.o: oracle/ref_harness.cpp | $(OBJDIR)
	$(CXX) $(OPTS) $(FLAGS) -fno-access-control -I$(REF) -c $< -o $@

$(OUT)/ref_harness: $(OBJS) $(OBJDIR)/ref_harness.o
	$(CXX) $(OPTS) $(FLAGS) -o $@ $^ -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
