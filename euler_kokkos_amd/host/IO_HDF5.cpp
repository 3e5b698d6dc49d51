// IO_HDF5.cpp -- see IO_HDF5.h.
#include "IO_HDF5.h"

#include <dlfcn.h>

#include <fstream>
#include <iostream>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "ek_hip.h"

namespace euler_kokkos {

namespace {

using WriteFn = int (*)(const char*, int, const unsigned long long*, const unsigned long long*, int, int,
                        const char* const*, const double* const*, int, double, int, int, int, int, char*, int);
using ReadFn = int (*)(const char*, int, const unsigned long long*, int, const char* const*, double* const*, int*,
                       double*, char*, int);

struct H5Lib {
  WriteFn write = nullptr;
  ReadFn read = nullptr;
  std::string why;
};

// libek_h5.so next to the library holding this code
const H5Lib& h5lib()
{
  static H5Lib lib;
  static std::once_flag once;
  std::call_once(once, [] {
    Dl_info info;
    std::string dir = ".";
    if (dladdr((void*)&hdf5_io_available, &info) && info.dli_fname) {
      dir = info.dli_fname;
      const size_t s = dir.find_last_of('/');
      dir = s == std::string::npos ? "." : dir.substr(0, s);
    }
    const std::string path = dir + "/libek_h5.so";
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char* e = dlerror();
      lib.why = "HDF5 support is not available (" + path + ": " + (e ? e : "not found") + ")";
      return;
    }
    lib.write = (WriteFn)dlsym(h, "ek_h5_write");
    lib.read = (ReadFn)dlsym(h, "ek_h5_read");
    if (!lib.write || !lib.read)
      lib.why = path + " lacks ek_h5_write / ek_h5_read";
  });
  return lib;
}

const H5Lib& need_h5()
{
  const H5Lib& L = h5lib();
  if (!L.write || !L.read)
    throw std::runtime_error(L.why);
  return L;
}

// variables in the files, in the reference's write order: density, energy, momentum x, y
// (z in 3D hydro and in MHD), then the MHD face fields
std::vector<int> file_vars(const HydroParams& p)
{
  std::vector<int> v = { ID, IP, IU, IV };
  if (p.dimType == THREE_D || p.mhdEnabled)
    v.push_back(IW);
  if (p.mhdEnabled) {
    v.push_back(IA);
    v.push_back(IB);
    v.push_back(IC);
  }
  return v;
}

std::string step_name(const ConfigMap& c, int iStep, const char* ext)
{
  std::ostringstream n;
  n.width(7);
  n.fill('0');
  n << iStep;
  return c.getString("output", "outputPrefix", "output") + "_" + n.str() + ext;
}

} // namespace

bool hdf5_io_available(std::string* why)
{
  const H5Lib& L = h5lib();
  if (why)
    *why = L.why;
  return L.write && L.read;
}

std::string save_HDF5(const std::vector<double>& Usoa, const std::vector<double>& Uref, const HydroParams& params,
                      const ConfigMap& configMap, const std::map<int, std::string>& variables_names, int iStep,
                      double totalTime)
{
  const H5Lib& L = need_h5();
  const bool d3 = params.dimType == THREE_D;
  const bool ghostIncluded = configMap.getBool("output", "ghostIncluded", false);
  const int nx = params.nx, ny = params.ny, nz = d3 ? params.nz : 1;
  const int is = params.isize, js = params.jsize, ks = d3 ? params.ksize : 1;
  const int nv = params.nbvar;
  int level = configMap.getInteger("output", "outputHdf5CompressionLevel", 0);
  if (level < 0 || level > 9) {
    std::cerr << "Invalid value for compression level; must be an integer between 0 and 9 !!!" << std::endl;
    std::cerr << "compression level is then set to default value 0; i.e. no compression !!" << std::endl;
    level = 0;
  }
  const std::vector<int> vars = file_vars(params);
  std::vector<std::vector<double>> bufs;
  std::vector<const double*> data;
  std::vector<const char*> names;
  const long long ncell = (long long)nx * ny * nz;
  for (int v : vars) {
    names.push_back(variables_names.at(v).c_str());
    if (!ghostIncluded) {
      data.push_back(Usoa.data() + (long long)v * ncell); // [k][j][i] is the file's C order
    } else {
      // reference layout U(i,j[,k],var) -> (k, j, i) C order with the ghosts
      std::vector<double> b((size_t)is * js * ks);
      for (int k = 0; k < ks; k++)
        for (int j = 0; j < js; j++)
          for (int i = 0; i < is; i++)
            b[((size_t)k * js + j) * is + i] = Uref[(((size_t)i * js + j) * ks + k) * nv + v];
      bufs.push_back(std::move(b));
      data.push_back(bufs.back().data());
    }
  }
  unsigned long long dims[3], chunk[3];
  const int rank = d3 ? 3 : 2;
  if (d3) {
    dims[0] = ghostIncluded ? ks : nz;
    dims[1] = ghostIncluded ? js : ny;
    dims[2] = ghostIncluded ? is : nx;
    chunk[0] = nz;
    chunk[1] = ny;
    chunk[2] = nx;
  } else {
    dims[0] = ghostIncluded ? js : ny;
    dims[1] = ghostIncluded ? is : nx;
    chunk[0] = ny;
    chunk[1] = nx;
  }
  const std::string file = configMap.getString("output", "outputDir", "./") + "/" + step_name(configMap, iStep, ".h5");
  char err[512] = "";
  if (L.write(file.c_str(), rank, dims, chunk, level, (int)vars.size(), names.data(), data.data(), iStep, totalTime,
              params.nx, params.ny, params.nz, ghostIncluded ? 1 : 0, err, (int)sizeof(err)) != 0)
    throw std::runtime_error(std::string("HDF5 output: ") + err);
  return file;
}

void load_HDF5(const std::string& filename, const HydroParams& params, const ConfigMap& configMap,
               const std::map<int, std::string>& variables_names, std::vector<double>& Usoa, int& iStep,
               double& totalTime)
{
  const H5Lib& L = need_h5();
  const bool d3 = params.dimType == THREE_D;
  const bool ghostIncluded = configMap.getBool("output", "ghostIncluded", false);
  const bool half = configMap.getBool("run", "restart_upscale", false);
  const int gw = params.ghostWidth;
  const int nx = params.nx, ny = params.ny, nz = d3 ? params.nz : 1;
  // sizes in the file (Load_HDF5::load: nx_r / nx_rg)
  const int nxr = half ? nx / 2 : nx, nyr = half ? ny / 2 : ny, nzr = d3 ? (half ? nz / 2 : nz) : 1;
  const int nxg = nxr + 2 * gw, nyg = nyr + 2 * gw, nzg = d3 ? nzr + 2 * gw : 1;
  const int fx = ghostIncluded ? nxg : nxr, fy = ghostIncluded ? nyg : nyr, fz = d3 ? (ghostIncluded ? nzg : nzr) : 1;
  const int rank = d3 ? 3 : 2;
  unsigned long long dims[3];
  if (d3) {
    dims[0] = fz;
    dims[1] = fy;
    dims[2] = fx;
  } else {
    dims[0] = fy;
    dims[1] = fx;
  }
  const std::vector<int> vars = file_vars(params);
  std::vector<std::vector<double>> bufs(vars.size(), std::vector<double>((size_t)fx * fy * fz));
  std::vector<double*> out;
  std::vector<const char*> names;
  for (size_t n = 0; n < vars.size(); n++) {
    out.push_back(bufs[n].data());
    names.push_back(variables_names.at(vars[n]).c_str());
  }
  char err[512] = "";
  if (L.read(filename.c_str(), rank, dims, (int)vars.size(), names.data(), out.data(), &iStep, &totalTime, err,
             (int)sizeof(err)) != 0)
    throw std::runtime_error(std::string("HDF5 restart: ") + err);

  // file array -> the padded array of the file's resolution (ghost cells zero when absent),
  // then the interior of the run's resolution
  const long long ncell = (long long)nx * ny * nz;
  Usoa.assign((size_t)ncell * params.nbvar, 0.0);
  const int off = ghostIncluded ? 0 : gw; // position of the file's first value in the padded array
  for (size_t n = 0; n < vars.size(); n++) {
    const int v = vars[n];
    std::vector<double> pad((size_t)nxg * nyg * nzg, 0.0);
    for (int k = 0; k < fz; k++)
      for (int j = 0; j < fy; j++)
        for (int i = 0; i < fx; i++)
          pad[((size_t)(k + (d3 ? off : 0)) * nyg + (j + off)) * nxg + (i + off)] =
              bufs[n][((size_t)k * fy + j) * fx + i];
    auto P = [&](int i, int j, int k) { return pad[((size_t)k * nyg + j) * nxg + i]; };
    double* dst = Usoa.data() + (long long)v * ncell;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          // padded fine indices
          const int I = i + gw, J = j + gw, K = d3 ? k + gw : 0;
          double val;
          if (!half) {
            val = P(I, J, K);
          } else {
            // Load_HDF5::copy_buffer, halfResolution branch (IO_HDF5.h:1593-1650, 1684-1788)
            const int iL = (I + gw) / 2, jL = (J + gw) / 2, kL = d3 ? (K + gw) / 2 : 0;
            val = P(iL, jL, kL);
            if (v == IA && I + gw - 2 * iL != 0)
              val = (P(iL, jL, kL) + P(iL + 1, jL, kL)) / 2;
            else if (v == IB && J + gw - 2 * jL != 0)
              val = (P(iL, jL, kL) + P(iL, jL + 1, kL)) / 2;
            else if (d3 && v == IC && K + gw - 2 * kL != 0)
              val = (P(iL, jL, kL) + P(iL, jL, kL + 1)) / 2;
          }
          dst[((size_t)k * ny + j) * nx + i] = val;
        }
  }
}

void writeXdmfForHdf5Wrapper(const HydroParams& params, const ConfigMap& configMap,
                             const std::map<int, std::string>& variables_names, int totalNumberOfSteps,
                             bool singleStep)
{
  const bool d3 = params.dimType == THREE_D;
  const bool ghostIncluded = configMap.getBool("output", "ghostIncluded", false);
  int nxg = params.nx, nyg = params.ny, nzg = params.nz;
  if (ghostIncluded) {
    nxg += 2 * params.ghostWidth;
    nyg += 2 * params.ghostWidth;
    nzg += 2 * params.ghostWidth;
  }
  const std::string type = "Double";
  const std::string prefix = configMap.getString("output", "outputPrefix", "output");
  std::string xdmfFilename = prefix + ".xmf";
  if (singleStep)
    xdmfFilename = step_name(configMap, totalNumberOfSteps, ".xmf");
  std::ofstream x(xdmfFilename);
  x << "<?xml version=\"1.0\" ?>\n";
  x << "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n";
  x << "<Xdmf xmlns:xi=\"http://www.w3.org/2003/XInclude\" Version=\"2.2\">\n";
  x << "  <Domain>\n";
  x << "    <Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  const int start = singleStep ? totalNumberOfSteps : 0;
  const int stop = singleStep ? totalNumberOfSteps + 1 : totalNumberOfSteps;
  const std::string dims = d3 ? std::to_string(nzg) + " " + std::to_string(nyg) + " " + std::to_string(nxg)
                              : std::to_string(nyg) + " " + std::to_string(nxg);
  std::vector<int> vars = file_vars(params);
  for (int iStep = start; iStep <= stop; iStep++) {
    const std::string h5 = step_name(configMap, iStep, ".h5");
    const std::string base = h5.substr(0, h5.size() - 3);
    x << "    <Grid Name=\"" << base << "\" GridType=\"Uniform\">\n";
    x << "    <Time Value=\"" << iStep << "\" />\n";
    if (!d3)
      x << "      <Topology TopologyType=\"2DCoRectMesh\" NumberOfElements=\"" << dims << "\"/>\n";
    else
      x << "      <Topology TopologyType=\"3DCoRectMesh\" NumberOfElements=\"" << dims << "\"/>\n";
    x << "    <Geometry Type=\"" << (d3 ? "ORIGIN_DXDYDZ" : "ORIGIN_DXDY") << "\">\n";
    x << "    <DataStructure\n";
    x << "       Name=\"Origin\"\n";
    x << "       DataType=\"" << type << "\"\n";
    x << "       Dimensions=\"" << (d3 ? 3 : 2) << "\"\n";
    x << "       Format=\"XML\">\n";
    x << "       " << (d3 ? "0 0 0" : "0 0") << "\n";
    x << "    </DataStructure>\n";
    x << "    <DataStructure\n";
    x << "       Name=\"Spacing\"\n";
    x << "       DataType=\"" << type << "\"\n";
    x << "       Dimensions=\"" << (d3 ? 3 : 2) << "\"\n";
    x << "       Format=\"XML\">\n";
    x << "       " << (d3 ? "1 1 1" : "1 1") << "\n";
    x << "    </DataStructure>\n";
    x << "    </Geometry>\n";
    for (int v : vars) {
      const std::string& name = variables_names.at(v);
      x << "      <Attribute Center=\"Cell\" Name=\"" << name << "\">\n";
      x << "        <DataStructure\n";
      x << "           DataType=\"" << type << "\"\n";
      x << "           Dimensions=\"" << dims << "\"\n";
      x << "           Format=\"HDF\">\n";
      x << "           " << h5 << ":/" << name << "\n";
      x << "        </DataStructure>\n";
      x << "      </Attribute>\n";
    }
    x << "   </Grid>\n";
  }
  x << "   </Grid>\n";
  x << " </Domain>\n";
  x << "</Xdmf>\n";
}

} // namespace euler_kokkos
