// IO_VTK.h -- VTK ImageData (.vti) writer with the reference's file layout.
//
// Follows save_VTK_2D/3D (src/utils/io/IO_VTK.cpp:25-210, 213-~410): file name
// <outputDir>/<outputPrefix>_<7-digit step>.vti, WholeExtent "0 nx 0 ny 0 nz", Origin and
// Spacing from the mesh, one cell array per variable named after m_variables_names, either
// ascii or raw-appended binary (UInt64 byte count before each array, x fastest). Only
// interior cells are written, which is exactly the device SoA layout.
#pragma once

#include <map>
#include <string>

#include "ConfigMap.h"
#include "HydroParams.h"

namespace euler_kokkos {

// U: SoA interior state [var][k][j][i] (FP64). Returns the file name written.
std::string save_VTK(const double* U, const HydroParams& params, const ConfigMap& configMap, int nbvar,
                     const std::map<int, std::string>& variables_names, int iStep,
                     const std::string& debug_name = "");

} // namespace euler_kokkos
