# Build of the gfx950 solver library, the host-only problem generators, the
# C++ facade tests and the CPU oracle.  `make -B` rebuilds everything.
HIPCC    ?= /opt/rocm/bin/hipcc
CXX      ?= g++
ARCH     ?= gfx950
LIBDIR   := mcls_amd/lib
SRC      := mcls_amd/csrc
# -ffp-contract=off: no FMA contraction, so the domain tables and walk paths
# round exactly like the CPU restatement.  -packed-fp32-ops: no packed FP32.
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -ffp-contract=off \
            -Xclang -target-feature -Xclang -packed-fp32-ops \
            -Iinclude -I$(SRC) -Wall -Wno-unused-command-line-argument
CXXFLAGS := -O2 -std=c++17 -fPIC -Wall -Iinclude

GPU_SRCS := $(SRC)/kernels.hip $(SRC)/context.hip $(SRC)/mlp.hip
GPU_HDRS := $(SRC)/kernels.hpp $(SRC)/mlp.hpp include/mcls_gpu.h include/mcls_mlp.h

all: $(LIBDIR)/libmcls_gpu.so $(LIBDIR)/libmcls_problems.so tests/cpp/bin/test_facade tests/cpp/bin/test_mm \
     tests/cpp/bin/test_traits tests/cpp/bin/test_tpetra_adapter oracle

$(LIBDIR)/libmcls_gpu.so: $(GPU_SRCS) $(GPU_HDRS)
	mkdir -p $(LIBDIR)
	$(HIPCC) $(HIPFLAGS) -shared -o $@ $(GPU_SRCS)

$(LIBDIR)/libmcls_problems.so: $(SRC)/problems.cpp include/mcls_problems.h
	mkdir -p $(LIBDIR)
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRC)/problems.cpp -lpthread

tests/cpp/bin/test_facade: tests/cpp/test_facade.cpp mcls_amd/host/*.hpp include/mcls_gpu.h $(LIBDIR)/libmcls_gpu.so
	mkdir -p tests/cpp/bin
	$(CXX) $(CXXFLAGS) -Imcls_amd/host -o $@ tests/cpp/test_facade.cpp -L$(LIBDIR) -lmcls_gpu \
	    -L/opt/rocm/lib -Wl,-rpath,'$$ORIGIN/../../../$(LIBDIR)' -Wl,-rpath,/opt/rocm/lib

# the Tpetra binding against a double of the Tpetra API it calls (Trilinos is
# not installed here); host checks without a GPU, the managers with --gpu
tests/cpp/bin/test_tpetra_adapter: tests/cpp/test_tpetra_adapter.cpp tests/cpp/tpetra_double.hpp mcls_amd/host/*.hpp \
    include/mcls_gpu.h $(LIBDIR)/libmcls_gpu.so
	mkdir -p tests/cpp/bin
	$(CXX) $(CXXFLAGS) -Imcls_amd/host -o $@ tests/cpp/test_tpetra_adapter.cpp -L$(LIBDIR) -lmcls_gpu \
	    -L/opt/rocm/lib -Wl,-rpath,'$$ORIGIN/../../../$(LIBDIR)' -Wl,-rpath,/opt/rocm/lib

# MatrixMarket reader driver: host-only, no GPU library
tests/cpp/bin/test_mm: tests/cpp/test_mm.cpp mcls_amd/host/MCLS_MatrixMarket.hpp mcls_amd/host/MCLS_Core.hpp
	mkdir -p tests/cpp/bin
	$(CXX) $(CXXFLAGS) -Imcls_amd/host -o $@ tests/cpp/test_mm.cpp

# traits and linear-problem checks: host-only (headers only, no GPU library)
tests/cpp/bin/test_traits: tests/cpp/test_traits.cpp mcls_amd/host/*.hpp include/mcls_gpu.h
	mkdir -p tests/cpp/bin
	$(CXX) $(CXXFLAGS) -Imcls_amd/host -o $@ tests/cpp/test_traits.cpp

oracle:
	$(MAKE) -C oracle

clean:
	rm -rf $(LIBDIR) oracle/build

.PHONY: all oracle clean
