// IO_HDF5.h -- HDF5 output, restart input and the Xdmf wrapper of the .ini driver.
//
// Follows Save_HDF5::save, Load_HDF5::load (src/utils/io/IO_HDF5.h:76-525, 1527-2143) and
// writeXdmfForHdf5Wrapper (src/utils/io/IO_HDF5.cpp:16-378). The HDF5 calls live in
// libek_h5.so (ek_h5.cpp), built only where the HDF5 headers exist and opened on first use;
// when it is missing, asking for HDF5 output or an HDF5 restart is an error.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "ConfigMap.h"
#include "HydroParams.h"

namespace euler_kokkos {

// true when libek_h5.so could be opened (why: the loader's message otherwise)
bool hdf5_io_available(std::string* why = nullptr);

// <outputDir>/<outputPrefix>_<7-digit iStep>.h5 with the conservative variables (density,
// energy, momenta, and for MHD the face fields), interior cells or, with
// [output] ghostIncluded, the ghost-filled arrays. Usoa: interior SoA [var][k][j][i];
// Uref: the reference layout U(i,j[,k],var) with ghosts (read only with ghostIncluded).
// Returns the file name.
std::string save_HDF5(const std::vector<double>& Usoa, const std::vector<double>& Uref, const HydroParams& params,
                      const ConfigMap& configMap, const std::map<int, std::string>& variables_names, int iStep,
                      double totalTime);

// Restart input: the conservative variables of `filename` into Usoa (interior SoA), and
// its "time step" / "total time" attributes. [output] ghostIncluded says whether the file
// holds ghost cells; [run] restart_upscale reads a file of half the resolution
// (nx/2 x ny/2 [x nz/2]) and injects every coarse cell into its 2^dim fine cells, averaging
// the two coarse face fields for fine faces that fall between them.
void load_HDF5(const std::string& filename, const HydroParams& params, const ConfigMap& configMap,
               const std::map<int, std::string>& variables_names, std::vector<double>& Usoa, int& iStep,
               double& totalTime);

// <outputPrefix>.xmf (or <outputPrefix>_<7-digit step>.xmf for a single step) in the
// working directory: a temporal collection of the .h5 files 0..totalNumberOfSteps.
void writeXdmfForHdf5Wrapper(const HydroParams& params, const ConfigMap& configMap,
                             const std::map<int, std::string>& variables_names, int totalNumberOfSteps,
                             bool singleStep);

} // namespace euler_kokkos
