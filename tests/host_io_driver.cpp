// Host-side driver for the sanitizer test (test_host_asan_cpu.py): the .ini parser, the
// parameters, the initial conditions and the VTK / HDF5 writers and the HDF5 reader of the
// host layer, with no device code, so that g++ can build all of it with AddressSanitizer
// and UndefinedBehaviorSanitizer.
//
// usage: host_io_driver file.ini   -> prints "ok <problem> <cells> vtk[ h5]"
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "ConfigMap.h"
#include "HydroParams.h"
#include "IO_HDF5.h"
#include "IO_VTK.h"
#include "Problems.h"

using namespace euler_kokkos;

int main(int argc, char** argv)
{
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s file.ini\n", argv[0]);
    return 2;
  }
  try {
    ConfigMap cfg = broadcast_parameters(argv[1]);
    HydroParams params(cfg);
    const std::map<int, std::string> names = { { ID, "rho" },    { IP, "energy" }, { IU, "rho_vx" },
                                               { IV, "rho_vy" }, { IW, "rho_vz" }, { IA, "bx" },
                                               { IB, "by" },     { IC, "bz" } };
    const std::string problem = cfg.getString("hydro", "problem", "unknown");
    std::vector<double> U;
    const bool known = params.mhdEnabled ? init_problem_mhd(problem, params, cfg, U)
                                         : init_problem(problem, params, cfg, U);
    if (!known) {
      std::fprintf(stderr, "no initial condition for %s\n", problem.c_str());
      return 3;
    }
    const long long ncell = (long long)params.nx * params.ny * (params.dimType == THREE_D ? params.nz : 1);
    if ((long long)U.size() != ncell * params.nbvar) {
      std::fprintf(stderr, "state size %zu, expected %lld\n", U.size(), ncell * params.nbvar);
      return 4;
    }
    save_VTK(U.data(), params, cfg, params.nbvar, names, 7);
    std::string extra;
    if (hdf5_io_available()) {
      const std::vector<double> Uref;  // ghostIncluded=false: interior cells only
      const std::string f = save_HDF5(U, Uref, params, cfg, names, 7, 0.25);
      std::vector<double> V(U.size(), -1.0);
      int it = -1;
      double t = -1.0;
      load_HDF5(f, params, cfg, names, V, it, t);
      if (it != 7 || t != 0.25) {
        std::fprintf(stderr, "restart attributes: step %d time %g\n", it, t);
        return 5;
      }
      // the file holds the conservative variables the reference writes (all of them here)
      for (size_t n = 0; n < U.size(); n++)
        if (V[n] != U[n]) {
          std::fprintf(stderr, "HDF5 round trip differs at %zu: %.17g vs %.17g\n", n, V[n], U[n]);
          return 6;
        }
      writeXdmfForHdf5Wrapper(params, cfg, names, 7, false);
      extra = " h5";
    }
    std::printf("ok %s %lld vtk%s\n", problem.c_str(), ncell, extra.c_str());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
