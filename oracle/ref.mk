# Compiles the reference Fortran core (nls/nls.f90 of the reference tree,
# read in place, never copied) into _ref/:
#   _ref/libnls_ref32.so  -- as written: real(sp) single precision, BLAS sgbmv
#   _ref/libnls_ref64.so  -- the same source with its kind function mapped to
#                            8 bytes and sgbmv to dgbmv by preprocessor flags
#                            only, with default reals promoted to 8 bytes
#                            (-fdefault-real-8: the literal beta = 1.0 that
#                            rgbmv passes to gbmv must be 8 bytes), i.e. the
#                            reference algorithm in double precision
# BLAS comes from the MKL runtime that ships in this image (/opt/conda/lib).
# Used only by tests and the CPU-baseline leg of bench.py.
REFSRC  ?= /root/reference/nls/nls.f90
FLANG   ?= /opt/rocm/bin/amdflang
MKLDIR  ?= /opt/conda/lib
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
FFLAGS  := -O2 -fPIC -shared
LIBS    := -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)

all: $(OUT)/libnls_ref32.so $(OUT)/libnls_ref64.so

$(OUT)/libnls_ref32.so: $(REFSRC)
	mkdir -p $(OUT)/m32
	$(FLANG) $(FFLAGS) -module-dir $(OUT)/m32 -o $@ $(REFSRC) $(LIBS)

$(OUT)/libnls_ref64.so: $(REFSRC)
	mkdir -p $(OUT)/m64
	$(FLANG) $(FFLAGS) -cpp '-Dselected_real_kind(a,b)=8' -Dsgbmv=dgbmv -fdefault-real-8 \
	    -module-dir $(OUT)/m64 -o $@ $(REFSRC) $(LIBS)

.PHONY: all
