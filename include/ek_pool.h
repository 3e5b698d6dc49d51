/*
 * ek_pool.h -- job pool: many independent grids advanced together on one GPU.
 *
 * No counterpart in the reference, which runs one grid per process; the pool serves the
 * batched and streaming uses of the same MUSCL-Hancock step. Each job is what one
 * reference run of a Hydro_Muscl_2D / Hydro_Muscl_3D / MHD_Muscl_3D .ini would compute: its own grid,
 * boundaries, gas constants, CFL, nStepmax and tEnd, and its own device-resident time
 * loop (SolverBase::compute_dt / finished / next_iteration, src/shared/SolverBase.cpp:124-195).
 * A job's result is the one its single-grid ek_run gives (same per-tile code).
 *
 * Jobs live in a fixed device arena (two ping-pong buffers); admission reserves the job's
 * state in both, retirement frees it. One ek_pool_step launch advances every unfinished
 * job by one step of its own dt, walking a work list of (job, tile) items with a
 * persistent grid; consecutive steps over an unchanged job set are replayed from a
 * captured HIP graph. Tile shapes are chosen by ek_pool_autotune. A 3D MHD pool also keeps
 * each job's ghost-padded work area (78 reals per padded cell) in a work buffer that takes
 * 84 % of arena_bytes (the two state buffers take the rest), and one pool step is the five
 * kernels of the 3D MHD step, each a persistent grid over all active jobs.
 *
 * All calls return 0 or an EK_E* code (include/ek_hip.h) with a message from
 * ek_pool_last_error. One pool per host thread.
 */
#ifndef EK_POOL_H_
#define EK_POOL_H_

#include "ek_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EK_EFULL 4 /* not enough free arena memory or job slots for this job */

typedef struct ek_pool ek_pool;

/* kind: the family every job of the pool belongs to -- dim, mhd, riemann, slope_type and
 * precision are taken from it (grid sizes, boundaries and constants may differ per job).
 * Supported: hydro 2D and 3D with HLLC, and 3D MHD with HLLD, HLL or LLF; slope_type 1 or 2;
 * FP64 or FP32.
 * arena_bytes: device bytes for job states (split over the two ping-pong buffers). */
int ek_pool_create(const ek_params* kind, long long arena_bytes, int max_jobs, ek_pool** out);
void ek_pool_destroy(ek_pool* pool);
const char* ek_pool_last_error(const ek_pool* pool);

/* Device bytes a job with these parameters needs in the arena (both buffers). */
long long ek_pool_job_bytes(const ek_pool* pool, const ek_params* job);
/* Arena bytes currently free (largest admissible job may be smaller: first-fit). */
long long ek_pool_free_bytes(const ek_pool* pool);
int ek_pool_active(const ek_pool* pool);

/* Admit a job: copies its initial state (layout as ek_upload) into the arena, sets t = 0,
 * it = 0 and reduces its first invDt. Returns EK_EFULL when it does not fit now. */
int ek_pool_admit(ek_pool* pool, const ek_params* job, const double* host_state, int layout, int nStepmax,
                  double tEnd, int* job_id);

/* ek_pool_admit without waiting for the upload (layout EK_LAYOUT_SOA): the copy of the
 * initial state (and its rounding in an FP32 pool) runs on the pool's upload stream while
 * ek_pool_step keeps advancing the other jobs, and the job joins the first ek_pool_step that
 * starts after the upload has completed (t = 0, it = 0 until then; an ek_pool_step with no
 * other job to advance waits for the earliest such upload). host_state must stay valid and
 * unchanged until the job is retired; page-locked memory (ek_host_alloc) lets the copy run
 * beside the steps. ek_pool_download / ek_pool_budgets / ek_pool_autotune wait for pending
 * uploads; ek_pool_retire of a job still uploading waits for its upload. (No reference
 * counterpart.) */
int ek_pool_admit_async(ek_pool* pool, const ek_params* job, const double* host_state, int nStepmax, double tEnd,
                        int* job_id);

/* Advance every unfinished job by nsteps steps (a job stops at its own tEnd / nStepmax).
 * Synchronizes and refreshes the jobs' t / it / finished. */
int ek_pool_step(ek_pool* pool, int nsteps);

/* Time, iteration and finished flag of a job after the last ek_pool_step. */
int ek_pool_status(const ek_pool* pool, int job_id, double* t, int* it, int* finished);

/* Copy a job's current state to the host (layout as ek_download). */
int ek_pool_download(ek_pool* pool, int job_id, double* host_state, int layout);

/* Compensated FP64 sums of the job's conserved variables (as ek_budgets). */
int ek_pool_budgets(ek_pool* pool, int job_id, double* sums);

/* Free a job's arena memory and slot. */
int ek_pool_retire(ek_pool* pool, int job_id);

/* Times `steps` pool steps with every candidate tile shape (2D) or z-chunk depth (3D) on
 * the current jobs, restores the jobs' state and time slots after each trial, and keeps the
 * fastest. *choice receives the chosen candidate index; ms[] (if not NULL, n_ms entries)
 * the measured time per step of each candidate. */
int ek_pool_autotune(ek_pool* pool, int steps, int* choice, double* ms, int n_ms);
int ek_pool_set_tile(ek_pool* pool, int choice);

/* Device time of the step launches of the last ek_pool_step (ms) and their count. */
int ek_pool_last_step_time(const ek_pool* pool, double* ms, int* launches);

/* Failure injection for tests of admission's rollback: with stage 1, 2 or 3, every following
 * ek_pool_admit fails (EK_EHIP) before its upload, after its upload and time-slot reset, or
 * after its device record is written; stage 0 turns it off. A failed admission leaves the
 * pool as it was (arena, job slots, work list). */
int ek_pool_test_fail_admit(ek_pool* pool, int stage);

/* Test hook for asynchronous admission: a job admitted by a following ek_pool_admit_async
 * joins no earlier than the `calls`-th ek_pool_step call after its admission while other jobs
 * are stepping (so that tests can make it join after an odd number of steps); 0 turns it off. */
int ek_pool_test_defer_join(ek_pool* pool, int calls);

#ifdef __cplusplus
}
#endif

#endif /* EK_POOL_H_ */
