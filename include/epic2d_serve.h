/* epic2d_serve.h — request loop over the batched ensemble (libepic2d_host).
 *
 * No reference counterpart: the reference runs one input.ini per process
 * (main.cpp:24-36).  Here a stream of input.ini-equivalent jobs arrives over
 * time; each job's particle and field state stays resident in HBM from
 * admission to completion; jobs are admitted first-come first-served
 * against a device-memory budget and a member cap, packed into one
 * epic_batch, and leave it when their step count is reached (continuous
 * batching: the batch changes between steps).  Reported: per-job
 * arrival/admission/completion times, p50/p99 time to completion and the
 * aggregate particle-steps/s.
 *
 * A job is a periodic plasma on an N x N-cell grid (N from grid_sizes),
 * ppc particles per cell per species, electron temperature te (eV), ion
 * temperature ti (eV); two-stream jobs split the electrons into two
 * counter-drifting halves at +-vd v_th,e; collisional jobs (never
 * two-stream: one electron species collides) have argon ions and add argon
 * neutrals with the cross sections of the caller's table (mcc_template).
 */
#ifndef EPIC2D_SERVE_H_
#define EPIC2D_SERVE_H_

#include <stdint.h>

#include "epic2d.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EPIC_SERVE_MAX_GRIDS 8

typedef struct {
  uint64_t seed;            /* draws arrivals and every job parameter */
  double rate_hz;           /* Poisson arrival rate */
  double duration_s;        /* arrivals in [0, duration_s) */
  int32_t max_jobs;         /* stop drawing after this many (0: no cap) */
  int32_t n_grids;
  int32_t grid_sizes[EPIC_SERVE_MAX_GRIDS]; /* cells per side, drawn uniformly */
  int32_t ppc_min, ppc_max; /* particles per cell per species, uniform */
  int32_t steps_min, steps_max;
  double te_min, te_max;    /* eV */
  double ti_min, ti_max;    /* eV */
  double vd_max;            /* two-stream drift, uniform in [0, vd_max] v_th,e */
  double two_stream_fraction;
  double collisional_fraction;
  double density;           /* m^-3 */
  double gas_density;       /* m^-3 argon for collisional jobs */
  double dt_coeff;          /* DT = dt_coeff / omega_pe */
  int32_t sort_interval;
  /* scheduling */
  double budget_bytes;      /* device memory for admitted jobs */
  int32_t max_members;
  int32_t use_graph;
  int32_t tune_steps;       /* > 0: tune kernel shapes on the first jobs before the clock starts */
  int32_t refine;           /* Poisson corrections per step (0: 2) */
  const epic_mcc_desc* mcc_template; /* tables for collisional jobs (species indices are set per job) */
  double progress_s;        /* > 0: a progress line on stderr at this interval */
  /* failure handling: every health_interval steps (0: 10) and at completion
   * each admitted job's device error counters and last Poisson residual are
   * read; a job with non-finite particles (EPIC_ERRC_NAN), dropped births
   * (CAPACITY), a fixed-point overflow, particles outside the grid, or a
   * relative residual that is not finite or above fail_relres (0: 1e-6) is
   * evicted and reported failed; the others go on undisturbed */
  int32_t health_interval;
  double fail_relres;
  int32_t fault_job;        /* fault injection for tests of that path: 1 + the index of a job whose
                               ions are loaded with a non-finite temperature (NaN velocities); 0: none */
} epic_serve_config;

typedef struct {
  int32_t index;
  double arrival_s, admit_s, done_s;  /* from the start of the stream */
  int32_t grid, ppc, steps, collisional, two_stream;
  double te, ti, vd;
  int64_t particles;                  /* at admission */
  int64_t bytes;                      /* admission estimate */
  double ke[3];                       /* digest at completion: kinetic energy of species 0..2 (x+y+z) */
  double phi_l2;                      /* digest: ||phi||_2 */
  int64_t collisions;                 /* digest: all MCC events */
  int32_t failed;                     /* 0: completed; else 1 + the EPIC_ERRC_* index that evicted it
                                         (EPIC_ERRC_NOCONV for a bad residual); done_s is then the eviction time */
  double relres;                      /* last Poisson ||b - A phi|| / ||b|| read for it */
  int64_t errors[EPIC_NERR];          /* its device error counters when last read */
} epic_serve_job;

typedef struct {
  int32_t jobs;
  int32_t jobs_done;
  double wall_s;              /* first arrival to last completion */
  double busy_s;              /* time with at least one job on the device */
  double latency_p50_s, latency_p99_s, latency_mean_s, latency_max_s;
  double wait_p50_s, wait_p99_s; /* arrival to admission */
  double particle_steps;      /* sum over jobs of particles x steps */
  double throughput;          /* particle_steps / wall_s */
  int32_t max_members;        /* largest batch */
  double peak_bytes;          /* largest admitted total */
  int64_t batch_steps, graph_captures;
  int32_t push_variant, node_block;
  int32_t jobs_failed;        /* evicted; latencies and particle_steps count completed jobs only */
} epic_serve_report;

/* The job stream of a configuration (no device work). */
int epic_serve_make_jobs(const epic_serve_config* cfg, epic_serve_job* out, int cap, int* n);
/* Descriptions of one job (norm scheme 1 of main.cpp:151-261), in the
 * given particle precision and solver; nsp is 2 or 3. */
int epic_serve_job_setup(const epic_serve_job* job, const epic_serve_config* cfg, int precision, int solver,
                         epic_domain_desc* dom, epic_species_desc sp[3], int* nsp);
/* Runs the stream in real time on the current device.  jobs_out (cap
 * entries, may be null) receives every job's record. */
int epic_serve_run(const epic_serve_config* cfg, epic_serve_report* rep, epic_serve_job* jobs_out, int cap);
const char* epic_serve_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* EPIC2D_SERVE_H_ */
