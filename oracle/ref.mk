# Builds the reference PhonoMC solver core from its own sources, in place under
# REF (read-only), with g++ and OpenMP only.  Outputs go to oracle/_ref/.
#
# Left out on purpose (they need HDF5, an external library, or are the CLI):
#   src/material/MaterialDataReader.cpp, src/material/ScatteringMatrixReader.cpp,
#   src/material/MaterialFactory.cpp, src/main.cpp
# The one reference to the HDF5 scattering-matrix reader (MonteCarloSolver.cpp,
# full-matrix branch, not a public TOML mode) stays an unresolved lazy symbol
# of the shared library and is never reached: materials are handed to the
# solver in memory by oracle/ref_harness.cpp.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
REF ?= /root/reference
OUT := oracle/_ref
OBJ := $(OUT)/obj

CORE_SRCS := \
  src/material/BzQuadrature.cpp src/material/MaterialData.cpp \
  src/material/ScatteringMatrix.cpp src/material/CallawayOperator.cpp \
  src/PhononMaterial.cpp src/material/Thermodynamics.cpp \
  src/material/PhononMaterialCollisionData.cpp \
  src/SimulationConfig.cpp src/config/ConfigParsing.cpp src/config/ConfigValidation.cpp \
  src/config/OutputDirectory.cpp src/SurfaceMesh.cpp src/geometry/DomainOutput.cpp \
  src/geometry/MeshReader.cpp src/geometry/SurfaceSampling.cpp src/SimulationDomain.cpp \
  src/MonteCarloSolver.cpp src/solver/ParticleStorage.cpp src/solver/CellParticleIndex.cpp \
  src/solver/RtaCollisionOperator.cpp src/solver/CarrierCollisionOperator.cpp \
  src/solver/NonlinearCallawayCollision.cpp src/solver/FullMatrixCollisionOperator.cpp \
  src/solver/PhononSourceOperatorFullMatrix.cpp src/solver/TemperatureGradientDrive.cpp \
  src/solver/Initialization.cpp src/solver/TransportOperator.cpp \
  src/solver/InterfaceInitialization.cpp src/solver/LayeredMaterialLayout.cpp \
  src/solver/BoundaryInteraction.cpp src/solver/InterfaceInteraction.cpp \
  src/solver/InterfaceModeSampler.cpp src/solver/AcousticInterfaceSampler.cpp \
  src/solver/ChannelCell.cpp src/solver/ReservoirInjector.cpp src/solver/ReservoirSampler.cpp \
  src/solver/ReservoirInitialization.cpp src/solver/RoughBoundaryInitialization.cpp \
  src/solver/RoughBoundarySampler.cpp src/solver/HeatSourceDriving.cpp \
  src/solver/PrescribedHeatSource.cpp src/solver/PhononSourceOperator.cpp \
  src/solver/Diagnostics.cpp src/solver/Observables.cpp src/solver/GridStatistics.cpp \
  src/solver/ModalResampler.cpp src/solver/ResultWriter.cpp src/solver/BoundaryLedger.cpp \
  src/solver/EnergyLedger.cpp

OBJS := $(patsubst %.cpp,$(OBJ)/%.o,$(CORE_SRCS))
CXXFLAGS := -std=c++17 -O3 -fPIC -fopenmp -DPHONOMC_USE_OPENMP=1 -include cstddef \
            -I$(REF)/include -I$(REF)/src

all: $(OUT)/libphonomc_ref.so

$(OBJ)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@

$(OBJ)/ref_harness.o: oracle/ref_harness.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@

$(OUT)/libphonomc_ref.so: $(OBJS) $(OBJ)/ref_harness.o
	g++ -shared -fopenmp -Wl,-z,lazy -o $@ $^

clean:
	rm -rf $(OUT)

.PHONY: all clean
