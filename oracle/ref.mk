# Builds the reference oxDNA CPU backend (the FP64 oracle binary) straight from
# the reference's own source files, with g++ and no cmake.  Outputs go only to
# oracle/_ref/ (git-ignored).  Test infrastructure only: nothing in the product
# links or runs what this produces.
#
#   make -f oracle/ref.mk -j8            # builds oracle/_ref/oxDNA + ForceDump.so
#
# Flags: -O2 without -ffast-math / -march=native, so the binary is strict IEEE
# FP64 and runs on any x86-64 host (the reference's default Release build adds
# -ffast-math -march=native, reference CMakeLists.txt:93-106).  The -D values
# below are the ones the reference's CMakeLists.txt:328 would pass.

REF      ?= /root/reference
SRC      := $(REF)/src
OUT      := oracle/_ref
CXX      ?= g++
CXXFLAGS := -std=c++14 -O2 -fPIC -DNOCUDA -DJSON_ENABLED -D_FORCE_INLINES \
            -DRELEASE='"oracle"' -DGIT_COMMIT='"none"' -DBUILD_TIME='"none"' \
            -I$(SRC)/extern -w

SRCS := \
	Observables/BaseObservable.cpp \
	Observables/ObservableFactory.cpp \
	Observables/ObservableOutput.cpp \
	Observables/Step.cpp \
	Observables/PotentialEnergy.cpp \
	Observables/StackingPropensity.cpp \
	Observables/CrossStackingPropensity.cpp \
	Observables/KineticEnergy.cpp \
	Observables/TotalEnergy.cpp \
	Observables/BackendInfo.cpp \
	Observables/HBEnergy.cpp \
	Observables/OrderParameterValues.cpp \
	Observables/HBList.cpp \
	Observables/StrandwiseBonds.cpp \
	Observables/PairEnergy.cpp \
	Observables/PairForce.cpp \
	Observables/ForceEnergy.cpp \
	Observables/Pressure.cpp \
	Observables/StressAutocorrelation.cpp \
	Observables/Density.cpp \
	Observables/Rdf.cpp \
	Observables/Distance.cpp \
	Observables/DensityProfile.cpp \
	Observables/ParticlePosition.cpp \
	Observables/CoaxVariables.cpp \
	Observables/Pitch.cpp \
	Observables/ExternalTorque.cpp \
	Observables/MeanVectorCosine.cpp \
	Observables/VectorAngle.cpp \
	Observables/Checkpoint.cpp \
	Observables/Contacts.cpp \
	Observables/Writhe.cpp \
	Observables/UnstackedList.cpp \
	Observables/PlectonemePosition.cpp \
	Observables/StretchedBonds.cpp \
	Observables/StructureFactor.cpp \
	Observables/FormFactor.cpp \
	Observables/TEPPlectonemePosition.cpp \
	Observables/AverageEnergy.cpp \
	Observables/ContactMap.cpp \
	Observables/AllVectors.cpp \
	Observables/ExternalForce.cpp \
	Observables/ParticleForceAndTorque.cpp \
	Observables/ParticleStress.cpp \
	Observables/Coordination.cpp \
	Observables/Configurations/Configuration.cpp \
	Observables/Configurations/BinaryConfiguration.cpp \
	Observables/Configurations/TEPtclOutput.cpp \
	Observables/Configurations/TEPxyzOutput.cpp \
	Backends/BackendFactory.cpp \
	Backends/SimBackend.cpp \
	Backends/AnalysisBackend.cpp \
	Backends/MDBackend.cpp \
	Backends/MCBackend.cpp \
	Backends/MD_CPUBackend.cpp \
	Backends/MC_CPUBackend.cpp \
	Backends/MC_CPUBackend2.cpp \
	Backends/VMMC_CPUBackend.cpp \
	Backends/MinBackend.cpp \
	Backends/FIREBackend.cpp \
	Backends/FFS_MD_CPUBackend.cpp \
	Backends/Thermostats/SRDThermostat.cpp \
	Backends/Thermostats/DPDThermostat.cpp \
	Backends/Thermostats/ThermostatFactory.cpp \
	Backends/Thermostats/BaseThermostat.cpp \
	Backends/Thermostats/NoThermostat.cpp \
	Backends/Thermostats/BrownianThermostat.cpp \
	Backends/Thermostats/RefreshThermostat.cpp \
	Backends/Thermostats/BussiThermostat.cpp \
	Backends/Thermostats/LangevinThermostat.cpp \
	Lists/BaseList.cpp \
	Lists/NoList.cpp \
	Lists/Cells.cpp \
	Lists/RodCells.cpp \
	Lists/VerletList.cpp \
	Lists/BinVerletList.cpp \
	Lists/ListFactory.cpp \
	Interactions/InteractionFactory.cpp \
	Interactions/InteractionUtils.cpp \
	Interactions/Mesh.cpp \
	Interactions/BaseInteraction.cpp \
	Interactions/LJInteraction.cpp \
	Interactions/DNAInteraction.cpp \
	Interactions/DNA2Interaction.cpp \
	Interactions/DNA2SDInteraction.cpp \
	Interactions/DNA3Interaction.cpp \
	Interactions/DNAInteraction_relax.cpp \
	Interactions/RNAInteraction.cpp \
	Interactions/RNAInteraction2.cpp \
	Interactions/RNAInteraction_relax.cpp \
	Interactions/TEPInteraction.cpp \
	Interactions/HSInteraction.cpp \
	Interactions/HardCylinderInteraction.cpp \
	Interactions/HardSpheroCylinderInteraction.cpp \
	Interactions/BoxInteraction.cpp \
	Interactions/PatchyInteraction.cpp \
	Interactions/KFInteraction.cpp \
	Interactions/DRHInteraction.cpp \
	Interactions/DRHInteraction_relax.cpp \
	Interactions/ANMInteraction.cpp \
	Interactions/DNANMInteraction.cpp \
	Interactions/RNANMInteraction.cpp \
	Backends/MCMoves/BaseMove.cpp \
	Backends/MCMoves/MoveFactory.cpp \
	Backends/MCMoves/MCTras.cpp \
	Backends/MCMoves/MCRot.cpp \
	Backends/MCMoves/Pivot.cpp \
	Backends/MCMoves/VMMC.cpp \
	Backends/MCMoves/VolumeMove.cpp \
	Backends/MCMoves/MoleculeVolumeMove.cpp \
	Backends/MCMoves/ShapeMove.cpp \
	Forces/ForceFactory.cpp \
	Forces/BaseForce.cpp \
	Forces/COMForce.cpp \
	Forces/ConstantRateForce.cpp \
	Forces/SawtoothForce.cpp \
	Forces/ConstantRateTorque.cpp \
	Forces/ConstantTrap.cpp \
	Forces/MovingTrap.cpp \
	Forces/LowdimMovingTrap.cpp \
	Forces/MutualTrap.cpp \
	Forces/RepulsiveEllipsoid.cpp \
	Forces/RepulsionPlane.cpp \
	Forces/RepulsionPlaneMoving.cpp \
	Forces/RepulsiveSphere.cpp \
	Forces/RepulsiveSphereSmooth.cpp \
	Forces/LJWall.cpp \
	Forces/HardWall.cpp \
	Forces/AlignmentField.cpp \
	Forces/GenericCentralForce.cpp \
	Forces/LJCone.cpp \
	Forces/YukawaSphere.cpp \
	Forces/Metadynamics/LTCoordination.cpp \
	Forces/Metadynamics/LTCOMTrap.cpp \
	Forces/Metadynamics/LTCOMAngleTrap.cpp \
	Forces/Metadynamics/LT2DCOMTrap.cpp \
	Forces/Metadynamics/LTAtanCOMTrap.cpp \
	Forces/Metadynamics/meta_utils.cpp \
	Forces/AttractionPlane.cpp \
	Forces/RepulsiveKeplerPoinsot.cpp \
	Forces/RepulsiveSphereMoving.cpp \
	Boxes/BaseBox.cpp \
	Boxes/CubicBox.cpp \
	Boxes/OrthogonalBox.cpp \
	Boxes/LeesEdwardsCubicBox.cpp \
	Boxes/BoxFactory.cpp \
	Particles/TEPParticle.cpp \
	Particles/RNANucleotide.cpp \
	Particles/BaseParticle.cpp \
	Particles/DNANucleotide.cpp \
	Particles/PatchyParticle.cpp \
	Particles/SpheroCylinder.cpp \
	Particles/Molecule.cpp \
	Particles/ANMParticle.cpp \
	Managers/SimManager.cpp \
	Utilities/OrderParameters.cpp \
	Utilities/Weights.cpp \
	Utilities/Histogram.cpp \
	Utilities/LR_vector.cpp \
	Utilities/LR_matrix.cpp \
	Utilities/Utils.cpp \
	Utilities/oxDNAException.cpp \
	Utilities/Logger.cpp \
	Utilities/parse_input/parse_input.cpp \
	Utilities/time_scales/time_scales.cpp \
	Utilities/SignalManager.cpp \
	Utilities/ConfigInfo.cpp \
	Utilities/FlattenedConfigInfo.cpp \
	Utilities/TopologyParser.cpp \
	PluginManagement/PluginManager.cpp \
	Utilities/Timings.cpp

OBJS := $(patsubst %.cpp,$(OUT)/obj/%.o,$(SRCS))

# the CPU baseline binary: the reference's Release flags (-O3 -ffast-math,
# CMakeLists.txt:93-106) with a portable x86-64-v3 target in place of
# -march=native, so the binary built here runs on the GPU box's host
RELFLAGS := -std=c++14 -O3 -ffast-math -march=x86-64-v3 -DNOCUDA -DJSON_ENABLED -D_FORCE_INLINES \
            -DRELEASE='"oracle-release"' -DGIT_COMMIT='"none"' -DBUILD_TIME='"none"' -I$(SRC)/extern -w
RELOBJS := $(patsubst %.cpp,$(OUT)/relobj/%.o,$(SRCS))

all: $(OUT)/oxDNA $(OUT)/ForceDump.so $(OUT)/oxDNA_release

$(OUT)/obj/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/tinyexpr.o: $(SRC)/extern/tinyexpr/tinyexpr.c
	@mkdir -p $(dir $@)
	gcc -O2 -fPIC -c $< -o $@

$(OUT)/obj/oxDNA_main.o: $(SRC)/oxDNA.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/oxDNA: $(OBJS) $(OUT)/obj/oxDNA_main.o $(OUT)/obj/tinyexpr.o
	$(CXX) -rdynamic -o $@ $^ -ldl

$(OUT)/relobj/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(RELFLAGS) -c $< -o $@

$(OUT)/relobj/oxDNA_main.o: $(SRC)/oxDNA.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(RELFLAGS) -c $< -o $@

$(OUT)/oxDNA_release: $(RELOBJS) $(OUT)/relobj/oxDNA_main.o $(OUT)/obj/tinyexpr.o
	$(CXX) -rdynamic -o $@ $^ -ldl

# full-precision per-particle force/torque dump, loaded by the reference's
# PluginManager (observable plugin ABI: extern "C" make_ForceDump)
$(OUT)/ForceDump.so: oracle/ForceDump.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -shared -I$(SRC) $< -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
