/*
 * ek_oracle.h -- CPU restatement of euler_kokkos' MUSCL-Hancock step (TEST INFRASTRUCTURE).
 *
 * This library is the oracle that the HIP product path is checked against. It is
 * not part of the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. It restates the reference's impl-0 (default) algorithm
 * with the reference's data layout (Kokkos OpenMP LayoutRight: U(i,j[,k],var),
 * var fastest, i slowest, ghost cells included) and the reference's operation
 * order, compiled with -ffp-contract=off so that no multiply-add is fused.
 *
 * Every function cites the reference file:line it follows (paths relative to the
 * euler_kokkos source tree).
 */
#ifndef EK_ORACLE_H_
#define EK_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* variable indices: src/shared/enums.h:17-36 */
enum { EKO_ID = 0, EKO_IP = 1, EKO_IU = 2, EKO_IV = 3, EKO_IW = 4, EKO_IA = 5, EKO_IB = 6, EKO_IC = 7 };
/* riemann solver ids: src/shared/enums.h:142-149 */
enum { EKO_RIEMANN_APPROX = 0, EKO_RIEMANN_LLF = 1, EKO_RIEMANN_HLL = 2, EKO_RIEMANN_HLLC = 3, EKO_RIEMANN_HLLD = 4 };
/* boundary types: src/shared/enums.h:153-160 */
enum { EKO_BC_UNDEFINED = 0, EKO_BC_DIRICHLET = 1, EKO_BC_NEUMANN = 2, EKO_BC_PERIODIC = 3 };

/* Mirror of HydroParams + HydroSettings (src/shared/HydroParams.h:32-69,81-197). */
typedef struct {
  int dim;            /* 2 or 3 */
  int nx, ny, nz;
  int gw;             /* ghostWidth */
  int nbvar;
  int isize, jsize, ksize;
  int imin, imax, jmin, jmax, kmin, kmax;
  int bc[6];          /* xmin, xmax, ymin, ymax, zmin, zmax */
  int riemann;
  int mhd;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double dx, dy, dz;
  double gamma0, gamma6, cfl, slope_type, smallr, smallc, smallp, smallpp, cIso;
} eko_params;

/* Implode problem parameters (src/shared/problems/ImplodeParams.h). */
typedef struct {
  double rho_out, p_out, u_out, v_out, w_out;
  double rho_in, p_in, u_in, v_in, w_in;
  int shape;
} eko_implode;

/* Blast problem parameters (src/shared/problems/BlastParams.h:13-26), hydro part. */
typedef struct {
  double radius, center_x, center_y, center_z;
  double density_in, density_out, pressure_in, pressure_out;
} eko_blast;

/* derived sizes / constants: HydroParams::init (src/shared/HydroParams.cpp:326-359) */
void eko_params_init(eko_params* p);

/* ---- cell-level functions (exposed for known-answer tests) ---- */
void eko_compute_primitives(const eko_params* p, const double* u, double* c, double* q);
void eko_slope_scalar(const eko_params* p, double q, double qPlus, double qMinus, double* dq);
void eko_trace_along_dir(const eko_params* p, const double* q, const double* dqX, const double* dqY,
                         const double* dqZ, double dtdx, double dtdy, double dtdz, int faceId,
                         double* qface);
void eko_riemann_hydro(const eko_params* p, const double* qleft, const double* qright, double* flux);

/* ---- array-level kernels on the reference layout (ghosts included) ---- */
long eko_index(const eko_params* p, int i, int j, int k, int var);
void eko_init_implode(const eko_params* p, const eko_implode* ip, double* U);
void eko_init_blast(const eko_params* p, const eko_blast* bp, double* U);
void eko_make_boundaries(const eko_params* p, double* U);
double eko_compute_inv_dt(const eko_params* p, const double* U);
void eko_convert_to_primitives(const eko_params* p, const double* U, double* Q);
void eko_compute_and_store_fluxes(const eko_params* p, const double* Q, double* Fx, double* Fy,
                                  double* Fz, double dt);
void eko_update(const eko_params* p, double* U, const double* Fx, const double* Fy, const double* Fz);
void eko_godunov_unsplit_impl(const eko_params* p, double* Uin, double* Uout, double* Q, double* Fx,
                              double* Fy, double* Fz, double dt);

/* Full solver loop (SolverBase::finished / next_iteration / compute_dt and
 * SolverHydroMuscl::next_iteration_impl). U holds the state on entry and exit. Runs
 * until t >= tEnd - 1e-14 or it >= nStepmax. Returns 0 on success. */
int eko_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
            double* dt_last);

/* ---- 2D MHD (SolverMHDMuscl<2>, impl 0), in ek_oracle_mhd.c; nbvar = 8, gw = 3 ---- */
void eko_mhd_constoprim(const eko_params* p, const double* u, const double* mag, double* c, double* q);
void eko_mhd_speed_info(const eko_params* p, const double* q, double* out);
void eko_riemann_mhd(const eko_params* p, double* ql, double* qr, double* flux);
double eko_compute_emfz(const eko_params* p, double qEdge[4][8]);
void eko_mhd_trace_2d(const eko_params* p, const double qNb[3][3][8], const double bfNb[4][4][3],
                      double dtdx, double dtdy, double qm[2][8], double qp[2][8], double qEdge[4][8]);
void eko_mhd_init_orszag_tang(const eko_params* p, double* U);
void eko_mhd_convert_to_primitives(const eko_params* p, const double* U, double* Q);
double eko_mhd_inv_dt(const eko_params* p, const double* Q);
void eko_mhd_compute_trace(const eko_params* p, const double* U, const double* Q, double* const out[8],
                           double dtdx, double dtdy);
void eko_mhd_compute_fluxes(const eko_params* p, double* const tr[8], double* Fx, double* Fy);
void eko_mhd_compute_emf(const eko_params* p, double* const tr[8], double* Emf);
void eko_mhd_update(const eko_params* p, double* U, const double* Fx, const double* Fy, const double* Emf,
                    double dtdx, double dtdy);
/* MHD time loop; dt is computed inside each step from the primitives of the current state. */
int eko_mhd_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
                double* dt_last);

/* ---- 3D MHD (SolverMHDMuscl<3>, impl 0), in ek_oracle_mhd.c ---- */
void eko_mhd3d_convert_to_primitives(const eko_params* p, const double* U, double* Q);
double eko_mhd3d_inv_dt(const eko_params* p, const double* Q);
void eko_mhd3d_elec_field(const eko_params* p, const double* U, const double* Q, double* E);
void eko_mhd3d_mag_slopes(const eko_params* p, const double* U, double* dA, double* dB, double* dC);
void eko_mhd3d_trace(const eko_params* p, const double* q, double dq[3][8], const double bfNb[6],
                     const double dbf[12], const double E[3][2][2], double dtdx, double dtdy, double dtdz,
                     double qm[3][8], double qp[3][8], double qE[4][3][8]);
double eko_compute_emf_dir(const eko_params* p, double qEdge[4][8], int dir);
int eko_mhd_run_dt(const eko_params* p, double* U, int nsteps, double dt);
int eko_mhd3d_run_dt(const eko_params* p, double* U, int nsteps, double dt);
int eko_mhd3d_run(const eko_params* p, double* U, int nStepmax, double tEnd, double* t_io, int* it_io,
                  double* dt_last);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int eko_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* EK_ORACLE_H_ */
