# Builds the reference DPSNN (LIFCA target) from its own sources, where they
# lie under REFDIR, into oracle/_ref/.  Objects and the binary never leave
# oracle/_ref/.  Needs the MPICH wrapper shipped in the image (/opt/conda).
#
#   make -f oracle/ref.mk            -> oracle/_ref/LIFCA-DPSNN.out (uniform delays)
#
# The reference's getParameters.c uses strtok without <string.h>; the system
# header is force-included on the compile line (no stand-in is written).
REFDIR  ?= /root/reference
OUT     ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
MPICXX  ?= /opt/conda/bin/mpic++
export MPICH_CXX = g++ -include string.h

# Same dimensions and switches as the reference Makefile's CXXLIFCAFLAGS,
# plus DelayUniform (uniform integer delays, as the workload definitions ask).
FLAGS = -O3 -fno-omit-frame-pointer -fno-PIC -w \
	-D miniLocN=32000 -D makeMaxGlobH=256 -D makeStencil_Max=96 \
	-D makeActiveLTP -D makeActiveLTD -D miniM=7500 -D miniSimSpikes=6000 \
	-D makeMaxCMFract=1 -D LIFCAneuron -D makeInsertBarrierBeforeSendRec \
	-D DelayUniform -mcmodel=large -I$(REFDIR)

SRCS = DPSNN_mainUnderMPI.c DPSNN_synapse.c DPSNN_neuron_init.c DPSNN_neuron_sim.c \
	DPSNN_spike.c DPSNN_messagePassing.c DPSNN_localNet_init.c DPSNN_localNet_sim.c \
	DPSNN_synListSort.c localNetProcess.c DPSNN_getParameters.c dalEmulationUnderMPI.c \
	DPSNN_random.c DPSNN_stopWatch.c DPSNN_stat.c DPSNN_chrono.c DPSNN_memMeasure.c \
	DPSNN_LIFCAcorticalModuleGen.cpp DPSNN_LIFCAconnectome.cpp DPSNN_writeIniFiles.c \
	DPSNN_localNet_chrono.cpp DPSNN_messageMeasure.c erflib.c randdev.c

OBJS = $(addprefix $(OUT)/obj/,$(addsuffix .o,$(basename $(SRCS))))

all: $(OUT)/LIFCA-DPSNN.out

$(OUT)/obj/%.o: $(REFDIR)/%.c
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(FLAGS) -x c++ -c $< -o $@

$(OUT)/obj/%.o: $(REFDIR)/%.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(FLAGS) -c $< -o $@

$(OUT)/LIFCA-DPSNN.out: $(OBJS)
	$(MPICXX) $(FLAGS) $(OBJS) -o $@

.PHONY: all
