# Builds oracle/_ref/libref.so from the reference's own header-only functions,
# compiled in place under /root/reference (nothing is copied). Test
# infrastructure only; needs the reference tree, so it runs in the build
# container and the built library travels with the tree.
REFERENCE ?= /root/reference
CXX ?= g++
CXXFLAGS ?= -O2 -fPIC -std=c++17 -ffp-contract=off -w

all: _ref/libref.so

# The table hash and validation come from two of the reference's own source
# files, compiled where they lie.
REF_SOURCES = $(REFERENCE)/src/kokkos/em/tables/ProposalNativeTable.cpp $(REFERENCE)/src/kokkos/utility/Sha256.cpp

_ref/libref.so: ref_probe.cpp ../include/c8mi/c8mi_tables.h ../include/c8mi/c8mi_em.h
	@mkdir -p _ref
	$(CXX) $(CXXFLAGS) -include cstring -I$(REFERENCE) -I../include -shared ref_probe.cpp $(REF_SOURCES) -o $@

clean:
	rm -rf _ref

.PHONY: all clean
