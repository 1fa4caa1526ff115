# Builds the reference engine from its own sources where they lie under
# /root/reference (read-only) into oracle/_ref/.  Test infrastructure only:
# used to pin the oracle restatement and as the CPU baseline in bench.py.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Nothing is copied from the reference; only compiled object code lands in
# oracle/_ref/ (git-ignored).  Two ISA variants are produced: the AVX-512
# build runs the reference's native W8 route, the AVX2 build its legacy W4
# route; callers pick the one the host CPU supports.

REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
CXXFLAGS_COMMON := -std=c++17 -O2 -fopenmp -fPIC -fvisibility=hidden -I$(REF)

CORE_SRC := $(REF)/core/math/general_math.cpp \
            $(REF)/core/shell/shell_math.cpp \
            $(REF)/core/shell/shell_contract.cpp \
            $(REF)/core/shell/cart2sph.cpp
ONEE_SRC := $(REF)/integrals/one_electron/onee_integrals.cpp \
            $(REF)/integrals/one_electron/onee_matrices.cpp
TWOE_SRC := $(REF)/integrals/two_electron/twoe_integrals.cpp \
            $(REF)/integrals/two_electron/fock/jk_direct.cpp

# the reference's existing one-electron C ABI (tlno_onee_*, with its named
# basis tables and .inp reader): fixture generation only, AVX-512 build
ONEE_CAPI_SRC := $(REF)/benchmarks/one_electron_engine/src/onee_capi.cpp \
                 $(REF)/benchmarks/one_electron_engine/src/basis_export.cpp \
                 $(REF)/benchmarks/common/molecule_io.cpp \
                 $(wildcard $(REF)/core/basis/*.cpp) \
                 $(wildcard $(REF)/core/molecule/*.cpp)

ISAS := avx512 avx2
march_avx512 := -march=x86-64-v4
march_avx2 := -march=x86-64-v3

API_LIBS := $(foreach i,$(ISAS),$(OUT)/libtlno_ref_$(i).so)
AUDIT_LIBS := $(foreach i,$(ISAS),$(OUT)/libtlno_ref_audit_$(i).so)

ONEE_CAPI_LIB := $(OUT)/libtlno_ref_onee_capi.so
# the engine with its stream-fingerprint validation switch on (fixture
# generation only, AVX-512 build)
FP_LIB := $(OUT)/libtlno_ref_fp_avx512.so

all: $(API_LIBS) $(AUDIT_LIBS) $(ONEE_CAPI_LIB) $(FP_LIB)

$(FP_LIB): oracle/ref_fp_shim.cpp $(CORE_SRC) | $(OUT)
	$(CXX) $(CXXFLAGS_COMMON) $(march_avx512) -shared -o $@ $^

$(ONEE_CAPI_LIB): $(CORE_SRC) $(ONEE_SRC) $(ONEE_CAPI_SRC) | $(OUT)
	$(CXX) -std=c++17 -O2 -fopenmp -fPIC -I$(REF) $(march_avx512) -shared -o $@ $^

$(OUT):
	mkdir -p $(OUT)

$(OUT)/libtlno_ref_%.so: oracle/ref_api_shim.cpp $(CORE_SRC) $(ONEE_SRC) $(TWOE_SRC) | $(OUT)
	$(CXX) $(CXXFLAGS_COMMON) $(march_$*) -shared -o $@ $^

$(OUT)/libtlno_ref_audit_%.so: oracle/ref_audit_shim.cpp $(CORE_SRC) | $(OUT)
	$(CXX) $(CXXFLAGS_COMMON) $(march_$*) -shared -o $@ $^

.PHONY: all
