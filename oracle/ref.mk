# Builds the reference ODE 0.12 library straight from its own source files
# (g++ on each file; the reference's autotools build is not used) into
# oracle/_ref/.  Two precisions are produced:
#   oracle/_ref/libode_ref_double.so  (-DdDOUBLE, the parity reference)
#   oracle/_ref/libode_ref_single.so  (-DdSINGLE, the CPU timing baseline)
# Each .so also carries the scene driver (oracle/scene_driver.cpp), an
# ordinary ODE-API application, compiled against the reference headers.
#
# The reference tree ships ode/src/config.h (trimesh with OPCODE enabled), so
# OPCODE's sources are compiled too.  Nothing from the reference is copied
# into this repository; the outputs go to oracle/_ref/ only (git-ignored).
#
# usage: make -f oracle/ref.mk -j8        (from the repository root)

REF      ?= /root/reference/ode-0.12
OUT      := oracle/_ref
CXX      ?= g++
CC       ?= gcc
OPT      ?= -O2
INC      := -I$(REF)/include -I$(REF)/ode/src -I$(REF)/ode/src/joints \
            -I$(REF)/OPCODE -I$(REF)/OPCODE/Ice
COMMON   := $(OPT) -fPIC -w -D__ODE__ -DdNODEBUG $(INC)

ODE_CPP  := $(filter-out %/collision_trimesh_gimpact.cpp %/collision_libccd.cpp, \
              $(wildcard $(REF)/ode/src/*.cpp)) \
            $(wildcard $(REF)/ode/src/joints/*.cpp) \
            $(wildcard $(REF)/OPCODE/*.cpp) \
            $(wildcard $(REF)/OPCODE/Ice/*.cpp)
ODE_C    := $(wildcard $(REF)/ode/src/*.c)

objs = $(patsubst $(REF)/%.cpp,$(OUT)/obj_$(1)/%.o,$(ODE_CPP)) \
       $(patsubst $(REF)/%.c,$(OUT)/obj_$(1)/%.o,$(ODE_C)) \
       $(OUT)/obj_$(1)/scene_driver.o

OBJ_D := $(call objs,double)
OBJ_S := $(call objs,single)

all: $(OUT)/libode_ref_double.so $(OUT)/libode_ref_single.so $(OUT)/buildinfo.txt

# compiler and flags of the reference build, reported with the CPU baseline
$(OUT)/buildinfo.txt: oracle/ref.mk
	@mkdir -p $(OUT)
	@echo "$$($(CXX) --version | head -1); flags $(OPT) -fPIC -DdNODEBUG -DdSINGLE|-DdDOUBLE" > $@

$(OUT)/libode_ref_double.so: $(OBJ_D)
	$(CXX) -shared -Wl,-Bsymbolic -o $@ $^
$(OUT)/libode_ref_single.so: $(OBJ_S)
	$(CXX) -shared -Wl,-Bsymbolic -o $@ $^

$(OUT)/obj_double/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(COMMON) -DdDOUBLE -c $< -o $@
$(OUT)/obj_double/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(COMMON) -DdDOUBLE -c $< -o $@
$(OUT)/obj_single/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(COMMON) -DdSINGLE -c $< -o $@
$(OUT)/obj_single/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(COMMON) -DdSINGLE -c $< -o $@

$(OUT)/obj_double/scene_driver.o: oracle/scene_driver.cpp oracle/scene.h
	@mkdir -p $(dir $@)
	$(CXX) $(OPT) -fPIC -DdDOUBLE -DSCENE_DRIVER_REFERENCE -I$(REF)/include -Ioracle -c $< -o $@
$(OUT)/obj_single/scene_driver.o: oracle/scene_driver.cpp oracle/scene.h
	@mkdir -p $(dir $@)
	$(CXX) $(OPT) -fPIC -DdSINGLE -DSCENE_DRIVER_REFERENCE -I$(REF)/include -Ioracle -c $< -o $@

clean:
	rm -rf $(OUT)/obj_double $(OUT)/obj_single $(OUT)/*.so

.PHONY: all clean

objs-double: $(filter-out %/scene_driver.o,$(OBJ_D))
