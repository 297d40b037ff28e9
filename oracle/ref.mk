# Builds the reference's path from its own few source files, read in place
# under REF, into oracle/_ref/ (git-ignored).  Nothing from the reference is
# copied into this repository: the sources are compiled where they lie.
#
#   make -f oracle/ref.mk            (from the repo root)
#
# Products:
#   _ref/dym-mod-metro   the reference driver, flags as in its Makefile:1-3
#   _ref/llr, _ref/llr2d the LLR drivers (same flags)
#   _ref/dym-mod-metro-savecfg  the driver that also writes SU(3) configurations
#   _ref/ref_harness     ref_harness.cpp (ours) linked against the reference's
#                        own get_staples / get_rect_loops / update (main renamed)
#   _ref/ref_llr_harness ref_llr_harness.cpp (ours) linked against llr.cpp's own
#                        calibrate_rect_mult / seed_window / rm_solve (main renamed)
#   _ref/verify_group    the reference's table checker
REF ?= /root/reference
HERE := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)/_ref

# Reference Makefile:1-3
REF_CXXFLAGS = -O3 -g -pg -fopenmp -std=c++11
REF_CFLAGS   = -O3 -std=gnu99 -Wall -fopenmp
LDLIBS = -lm

SRC_C = $(REF)/lattice.c $(REF)/group.c

all: $(OUT)/dym-mod-metro $(OUT)/llr $(OUT)/llr2d $(OUT)/ref_harness $(OUT)/verify_group \
     $(OUT)/dym-mod-metro-savecfg $(OUT)/ref_llr_harness

$(OUT):
	mkdir -p $(OUT)

$(OUT)/lattice.o: $(REF)/lattice.c | $(OUT)
	gcc $(REF_CFLAGS) -w -I$(REF) -c $< -o $@
$(OUT)/group.o: $(REF)/group.c | $(OUT)
	gcc $(REF_CFLAGS) -w -I$(REF) -c $< -o $@
$(OUT)/timer.o: $(REF)/timer.cpp | $(OUT)
	g++ $(REF_CXXFLAGS) -w -I$(REF) -c $< -o $@

REF_OBJS = $(OUT)/lattice.o $(OUT)/group.o $(OUT)/timer.o

$(OUT)/dym-mod-metro: $(REF)/dym-mod-metro.cpp $(REF_OBJS)
	g++ $(REF_CXXFLAGS) -w -I$(REF) $< $(REF_OBJS) $(LDLIBS) -o $@
$(OUT)/llr: $(REF)/llr.cpp $(REF_OBJS)
	g++ $(REF_CXXFLAGS) -w -I$(REF) $< $(REF_OBJS) $(LDLIBS) -o $@
$(OUT)/llr2d: $(REF)/llr2d.cpp $(REF_OBJS)
	g++ $(REF_CXXFLAGS) -w -I$(REF) $< $(REF_OBJS) $(LDLIBS) -o $@
$(OUT)/dym-mod-metro-savecfg: $(REF)/dym-mod-metro-savecfg.cpp $(REF_OBJS)
	g++ $(REF_CXXFLAGS) -w -I$(REF) $< $(REF_OBJS) $(LDLIBS) -o $@
$(OUT)/verify_group: $(REF)/verify_group.c $(OUT)/group.o
	gcc $(REF_CFLAGS) -w -I$(REF) $< $(OUT)/group.o $(LDLIBS) -o $@

# The driver with main() renamed, so the harness can call its functions.
$(OUT)/metro_nomain.o: $(REF)/dym-mod-metro.cpp | $(OUT)
	g++ -O3 -fopenmp -std=c++11 -w -Dmain=ref_main -I$(REF) -c $< -o $@
$(OUT)/ref_harness: $(HERE)/ref_harness.cpp $(OUT)/metro_nomain.o $(REF_OBJS)
	g++ -O2 -fopenmp -std=c++11 -I$(REF) $< $(OUT)/metro_nomain.o $(REF_OBJS) $(LDLIBS) -pg -o $@

$(OUT)/llr_nomain.o: $(REF)/llr.cpp | $(OUT)
	g++ -O3 -fopenmp -std=c++11 -w -Dmain=llr_main -I$(REF) -c $< -o $@
$(OUT)/ref_llr_harness: $(HERE)/ref_llr_harness.cpp $(OUT)/llr_nomain.o $(REF_OBJS)
	g++ -O2 -fopenmp -std=c++11 -I$(REF) $< $(OUT)/llr_nomain.o $(REF_OBJS) $(LDLIBS) -pg -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
