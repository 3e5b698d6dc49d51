/*
 * ek_host.h -- C entry points into the C++ host layer of libek_hip.so (the Solver /
 * ConfigMap side that keeps the reference's .ini interface). Used by Python callers and
 * tests; C++ callers use the classes in euler_kokkos_amd/host/ directly.
 */
#ifndef EK_HOST_H_
#define EK_HOST_H_

#include "ek_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Value of section.name in the .ini text `text` (len bytes) as ConfigMap reads it
 * (src/utils/config/ConfigMap.cpp, inih/INIReader.cpp): kind 0 getString (default
 * "<unset>"), 1 getInteger (default -987654321), 2 getFloat printed with %.17g (default
 * -1.5), 3 getBool (default false). Returns the full length of the value. */
int ek_ini_get(const char* text, int len, const char* section, const char* name, int kind, char* out,
               int outlen);

/* ConfigMap::ParseError of the .ini text: 0, or the line number of the first error. */
int ek_ini_parse_error(const char* text, int len);

/* ek_params of the .ini file at `path`, through HydroParams (src/shared/HydroParams.cpp). */
int ek_ini_params(const char* path, ek_params* out);

/* Builds the solver named in the .ini through SolverFactory (initial condition included),
 * runs the reference time loop to completion and copies the final state (SoA interior,
 * n doubles) to `state`, the final time to *t_out and iteration count to *it_out. */
int ek_ini_run(const char* path, double* state, long long n, double* t_out, int* it_out);

/* Message of the last failed ek_ini_* call on this thread. */
const char* ek_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* EK_HOST_H_ */
