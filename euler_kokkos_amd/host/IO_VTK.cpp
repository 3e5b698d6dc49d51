// IO_VTK.cpp -- see IO_VTK.h.
#include "IO_VTK.h"

#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace euler_kokkos {

std::string save_VTK(const double* U, const HydroParams& p, const ConfigMap& c, int nbvar,
                     const std::map<int, std::string>& names, int iStep, const std::string& debug_name)
{
  const bool is3d = p.dimType == THREE_D;
  const int nz = is3d ? p.nz : 1;
  const std::string outputDir = c.getString("output", "outputDir", "./");
  const std::string outputPrefix = c.getString("output", "outputPrefix", "output");
  const bool ascii = c.getBool("output", "outputVtkAscii", false);
  std::ostringstream stepNum;
  stepNum.width(7);
  stepNum.fill('0');
  stepNum << iStep;
  const std::string filename = debug_name.empty()
                                 ? outputDir + "/" + outputPrefix + "_" + stepNum.str() + ".vti"
                                 : outputDir + "/" + outputPrefix + "_" + debug_name + "_" + stepNum.str() + ".vti";
  std::ofstream out(filename, std::ios::out | std::ios::binary);
  if (ascii)
    out << "<?xml version=\"1.0\"?>\n";
  out << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
  out << "  <ImageData WholeExtent=\"" << 0 << " " << p.nx << " " << 0 << " " << p.ny << " " << 0 << " "
      << (is3d ? p.nz : 0) << "\" " << "Origin=\"" << p.xmin << " " << p.ymin << " " << (is3d ? p.zmin : 0)
      << "\" " << "Spacing=\"" << p.dx << " " << p.dy << " " << (is3d ? p.dz : 0.0) << "\">\n";
  out << "  <Piece Extent=\"" << 0 << " " << p.nx << " " << 0 << " " << p.ny << " " << 0 << " "
      << (is3d ? p.nz : 0) << " " << "\">\n";
  out << "    <PointData>\n";
  out << "    </PointData>\n";
  const long long ncell = (long long)p.nx * p.ny * nz;
  if (ascii) {
    out << "    <CellData>\n";
    for (int v = 0; v < nbvar; v++) {
      out << "    <DataArray type=\"Float64\" Name=\"" << names.at(v) << "\" format=\"ascii\" >\n";
      for (long long cidx = 0; cidx < ncell; cidx++)
        out << U[v * ncell + cidx] << " ";
      out << "\n    </DataArray>\n";
    }
    out << "    </CellData>\n";
    out << "  </Piece>\n";
    out << "  </ImageData>\n";
    out << "</VTKFile>\n";
  } else {
    out << "    <CellData>" << std::endl;
    for (int v = 0; v < nbvar; v++)
      out << "     <DataArray type=\"Float64\" Name=\"" << names.at(v) << "\" format=\"appended\" offset=\""
          << v * ncell * (long long)sizeof(double) + v * (long long)sizeof(uint64_t) << "\" />" << std::endl;
    out << "    </CellData>" << std::endl;
    out << "  </Piece>" << std::endl;
    out << "  </ImageData>" << std::endl;
    out << "  <AppendedData encoding=\"raw\">" << std::endl;
    out << "_";
    const uint64_t nbOfWords = (uint64_t)ncell * sizeof(double);
    for (int v = 0; v < nbvar; v++) {
      out.write((const char*)&nbOfWords, sizeof(uint64_t));
      out.write((const char*)(U + v * ncell), (std::streamsize)(ncell * sizeof(double)));
    }
    out << "  </AppendedData>" << std::endl;
    out << "</VTKFile>" << std::endl;
  }
  return filename;
}

} // namespace euler_kokkos
