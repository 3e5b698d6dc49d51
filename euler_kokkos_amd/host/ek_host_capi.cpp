// ek_host_capi.cpp -- C entry points into the C++ host layer (declared in include/ek_host.h).
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "ConfigMap.h"
#include "HydroParams.h"
#include "SolverFactory.h"
#include "SolverHydroMuscl.h"
#include "SolverMHDMuscl.h"
#include "ek_host.h"

using namespace euler_kokkos;

namespace {
thread_local std::string g_err;

int copy_out(const std::string& s, char* out, int outlen)
{
  if (!out || outlen <= 0)
    return (int)s.size();
  const int n = (int)s.size() < outlen - 1 ? (int)s.size() : outlen - 1;
  std::memcpy(out, s.data(), (size_t)n);
  out[n] = '\0';
  return (int)s.size();
}
} // namespace

extern "C" int ek_ini_get(const char* text, int len, const char* section, const char* name,
                          int kind, char* out, int outlen)
{
  ConfigMap c(text, len);
  std::string v;
  switch (kind) {
    case 0: v = c.getString(section, name, "<unset>"); break;
    case 1: v = std::to_string(c.getInteger(section, name, -987654321)); break;
    case 2: {
      char buf[64];
      std::snprintf(buf, sizeof(buf), "%.17g", (double)c.getFloat(section, name, -1.5f));
      v = buf;
      break;
    }
    default: v = c.getBool(section, name, false) ? "true" : "false"; break;
  }
  return copy_out(v, out, outlen);
}

extern "C" int ek_ini_parse_error(const char* text, int len)
{
  return ConfigMap(text, len).ParseError();
}

extern "C" int ek_ini_params(const char* path, ek_params* out)
{
  try {
    ConfigMap c(path);
    if (c.ParseError() < 0) {
      g_err = std::string("cannot read ") + path;
      return EK_EINVAL;
    }
    HydroParams p(c);
    *out = p.to_ek();
    return EK_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return EK_EINVAL;
  }
}

extern "C" int ek_ini_run(const char* path, double* state, long long n, double* t_out, int* it_out)
{
  try {
    ConfigMap c(path);
    if (c.ParseError() < 0) {
      g_err = std::string("cannot read ") + path;
      return EK_EINVAL;
    }
    HydroParams p(c);
    const std::string name = c.getString("run", "solver_name", "Unknown");
    if (!SolverFactory::Instance().has(name)) {
      g_err = "unknown solver " + name;
      return EK_EINVAL;
    }
    SolverBase* s = SolverFactory::Instance().create(name, p, c);
    s->run();
    std::vector<double> U;
    if (auto* h2 = dynamic_cast<muscl::SolverHydroMuscl<2>*>(s))
      U = h2->state();
    else if (auto* h3 = dynamic_cast<muscl::SolverHydroMuscl<3>*>(s))
      U = h3->state();
    else if (auto* m2 = dynamic_cast<muscl::SolverMHDMuscl<2>*>(s))
      U = m2->state();
    else if (auto* m3 = dynamic_cast<muscl::SolverMHDMuscl<3>*>(s))
      U = m3->state();
    if ((long long)U.size() != n) {
      delete s;
      g_err = "state buffer size mismatch";
      return EK_EINVAL;
    }
    std::memcpy(state, U.data(), U.size() * sizeof(double));
    *t_out = s->m_t;
    *it_out = s->m_iteration;
    delete s;
    return EK_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return EK_EHIP;
  }
}

extern "C" const char* ek_host_last_error(void)
{
  return g_err.c_str();
}
