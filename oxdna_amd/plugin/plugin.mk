# Builds the HIPDNA2 and HIPRNA2 interaction plugins (oxdna_amd/plugin/HIPDNA2.so,
# HIPRNA2.so) for the reference oxDNA program: compiled against the reference's headers where
# they lie (REF/src, read only) and linked to liboxh.so and the HIP runtime.
#
#   make -f oxdna_amd/plugin/plugin.mk          # needs oxdna_amd/liboxh.so
#
# The -D flags are those the reference's CMakeLists.txt:328 passes to its
# own sources (and oracle/ref.mk to the binary the tests load it into), so
# that the classes have the same layout on both sides.

REF      ?= /root/reference
SRC      := $(REF)/src
PKG      := oxdna_amd
ROCM     ?= /opt/rocm
CXX      ?= g++
CXXFLAGS := -std=c++14 -O2 -fPIC -DNOCUDA -DJSON_ENABLED -D_FORCE_INLINES -D__HIP_PLATFORM_AMD__ \
            -I$(SRC) -I$(SRC)/extern -Iinclude -I$(ROCM)/include -w

all: $(PKG)/plugin/HIPDNA2.so $(PKG)/plugin/HIPRNA2.so

$(PKG)/plugin/HIPDNA2.so: $(PKG)/plugin/HIPDNA2.cpp include/oxh.h $(PKG)/liboxh.so
	$(CXX) $(CXXFLAGS) -shared $< -o $@ -L$(PKG) -loxh -L$(ROCM)/lib -lamdhip64 \
		-Wl,-rpath,'$$ORIGIN/..' -Wl,-rpath,$(ROCM)/lib

$(PKG)/plugin/HIPRNA2.so: $(PKG)/plugin/HIPRNA2.cpp include/oxh.h $(PKG)/liboxh.so
	$(CXX) $(CXXFLAGS) -shared $< -o $@ -L$(PKG) -loxh -L$(ROCM)/lib -lamdhip64 \
		-Wl,-rpath,'$$ORIGIN/..' -Wl,-rpath,$(ROCM)/lib

clean:
	rm -f $(PKG)/plugin/HIPDNA2.so $(PKG)/plugin/HIPRNA2.so

.PHONY: all clean
