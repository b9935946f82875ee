/* nbg.h -- C ABI of libnbg_hip.so: the AHL21 integrator step with Jacobian
 * propagation and the transit-time search of NbodyGradient.jl, on gfx950.
 *
 * Conventions (shared by every entry point):
 *  - All arrays are host memory in the reference's Julia layout
 *    (column-major, 1-based indices become 0-based):
 *      x, v, xerror, verror : 3 x n        m : n        pair : n x n (uint8)
 *      jac_step, jac_error, jac_init : 7n x 7n        dqdt : 7n
 *      tt : n x ntt (TransitTiming) or 3 x n x ntt (TransitParameters.ttbv)
 *      dtdq0, dtdelements : n x ntt x 7 x n  (or 3 x n x ntt x 7 x n)
 *      count : n (int64)
 *  - Indices of bodies (ti) are 0-based.
 *  - The library copies in and out and keeps no host pointer after return.
 *  - Return value: 0 on success, negative on error (message in
 *    nbg_last_error).  Numerical non-convergence is not an error: the
 *    reference caps its iterations silently, and so does this library.
 *  - One call at a time per context (calls are serialised by a mutex).
 *  - Bodies per system: 2 <= n <= 12.  jac_step's mass rows must be unit rows
 *    and jac_error's mass rows zero (as State(ic) creates them; every step
 *    preserves this).
 */
#ifndef NBG_H
#define NBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nbg_ctx nbg_ctx;

/* Context on one GPU (HIP device index). */
int nbg_ctx_create(nbg_ctx** ctx, int device);
void nbg_ctx_destroy(nbg_ctx* ctx);
const char* nbg_last_error(const nbg_ctx* ctx);
/* Device name, compute-unit count, LDS bytes per workgroup serving n bodies. */
int nbg_device_info(nbg_ctx* ctx, char* name, int name_len, int* cu_count);
size_t nbg_lds_bytes_per_system(int n);

/* One AHL21 step, in place.
 * Replaces ahl21!(s, d, h) (src/integrator/ahl21/ahl21.jl:5-95) when grad != 0
 * and ahl21!(s, h) (src/integrator/ahl21/ahl21_no_grad.jl:6-18) when grad == 0;
 * the Julia scheme plug-in Integrator.scheme calls it (Integrator.jl:17-22,
 * Transits.jl:156-158).  With grad == 0, jac_step/jac_error/dqdt may be NULL. */
int nbg_ahl21_step(nbg_ctx* ctx, int32_t n, double h, int32_t grad, double* x, double* v, double* xerror,
                   double* verror, const double* m, const uint8_t* pair, double* jac_step, double* jac_error,
                   double* dqdt);

/* (intr::Integrator)(s, N; grad) -- Integrator.jl:211-234.  N may be
 * negative (integrates backward); *t advances with Kahan summation. */
int nbg_integrate_steps(nbg_ctx* ctx, int32_t n, double h, int64_t N, int32_t grad, double* t, double* x, double* v,
                        double* xerror, double* verror, const double* m, const uint8_t* pair, double* jac_step,
                        double* jac_error, double* dqdt);

/* (intr::Integrator)(s, time; grad) -- Integrator.jl:159-197: full steps of h
 * towards `time`, one partial step to land on it, *t = time. */
int nbg_integrate_to(nbg_ctx* ctx, int32_t n, double h, double time, int32_t grad, double* t, double* x, double* v,
                     double* xerror, double* verror, const double* m, const uint8_t* pair, double* jac_step,
                     double* jac_error, double* dqdt);

/* (intr::Integrator)(s, tt; grad) -- Transits.jl:140-180 with timing.jl:3-194:
 * round(tmax/h) steps from *t, transit search after every step (findtransit!
 * with dt/dq), calc_dtdelements! at the end when grad != 0.
 * ntbv = 1: TransitTiming (tt, dtdq0, dtdelements);
 * ntbv = 3: TransitParameters (ttbv, dtbvdq0, dtbvdelements).
 * The output arrays are filled in place (entries beyond count stay as given). */
int nbg_integrate_transits(nbg_ctx* ctx, int32_t n, double h, double tmax, int32_t ti, int32_t grad, int32_t ntbv,
                           double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                           const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt,
                           const double* jac_init, int64_t ntt, double* tt, double* dtdq0, double* dtdelements,
                           int64_t* count);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nbg_integrate_transits with the same arguments and the
 * same results, which calls no hipMalloc, hipFree or other allocating
 * routine.  Its device memory is one array of NBG_NOALLOC_SCRATCH_BYTES that
 * the library holds on every GPU it runs on until the process ends; a system
 * whose record (state, three 7n x 7n Jacobians, work area, tt, dtdq0 and
 * dtdelements: about 16 * ntbv * n * ntt * 7n + 24 * (7n)^2 bytes) does not
 * fit is rejected with an error before any device work.  Every call on one
 * GPU uses that GPU's one array: calls that overlap on the same GPU (through
 * different contexts, or other processes' threads sharing the library) race;
 * calls through one context are serialised. */
#define NBG_NOALLOC_SCRATCH_BYTES (32u << 20)
int nbg_integrate_transits_noalloc(nbg_ctx* ctx, int32_t n, double h, double tmax, int32_t ti, int32_t grad,
                                   int32_t ntbv, double* t, double* x, double* v, double* xerror, double* verror,
                                   const double* m, const uint8_t* pair, double* jac_step, double* jac_error,
                                   double* dqdt, const double* jac_init, int64_t ntt, double* tt, double* dtdq0,
                                   double* dtdelements, int64_t* count);

/* (intr::Integrator)(s, o::CartesianOutput) -- src/outputs/Outputs.jl:26-49:
 * nstep steps of h * check_step(*t, tmax) (tmax = intr.tmax sets only the
 * direction), recording the state before each step k = 0..nstep-1:
 *   out_t : nstep          out_x, out_v : 3 x n x nstep
 *   out_xerror, out_verror : 3 x n x nstep (both or neither; may be NULL)
 *   out_jac, out_jac_error : 7n x 7n x nstep (each may be NULL; need grad)
 * Afterwards the state holds the result of nstep steps and *t = t0 + h*nstep.
 * The reference always steps with the Jacobian; grad == 0 gives the same
 * x and v (its test_integrator.jl:184-185) without it.  Writing the records
 * to a JLD2 file (o.file) stays with the caller. */
int nbg_integrate_cartesian(nbg_ctx* ctx, int32_t n, double h, double tmax, int64_t nstep, int32_t grad,
                            double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                            const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt, double* out_t,
                            double* out_x, double* out_v, double* out_xerror, double* out_verror, double* out_jac,
                            double* out_jac_error);

/* (intr::Integrator)(s, ts::TransitSnapshot; grad) -- Transits.jl:186-211
 * with snapshot.jl:7-54 (the reference's constructor, Transits.jl:124, is
 * broken; its integrator method is what this follows): integrate from *t to
 * each of the nt times in turn as nbg_integrate_to does, and there record
 * v_sky and b^2 of every body relative to body 1 (vsky, bsky2: n x nt) and,
 * with grad, their derivatives as calc_dbvdq! forms them (dbvdq0:
 * 2 x n x nt x 7 x n; index 0 v_sky, 1 b^2) and the element gradients
 * calc_dbvdelements! forms from them with jac_init (dbvdelements, same
 * shape; zero without grad).  Body 1's entries are zero.  Afterwards the
 * state is at the last time, *t = times[nt-1]. */
int nbg_integrate_snapshot(nbg_ctx* ctx, int32_t n, double h, int64_t nt, const double* times, int32_t grad,
                           double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                           const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt,
                           const double* jac_init, double* vsky, double* bsky2, double* dbvdq0,
                           double* dbvdelements);

/* ---- Batched form: many independent systems in one persistent launch ----
 * Systems are packed back to back: system s uses n[s] bodies and ntt[s]
 * transit slots; x/v/m/pair/jac_init/tt/dtdq0/dtdelements/count hold the
 * per-system arrays above concatenated in system order (pair may be NULL =
 * all pairs Kepler, i.e. State's default).  Every system starts from
 * xerror = verror = 0, jac_step = I, jac_error = 0 at time t0[s]. */
typedef struct nbg_batch nbg_batch;

typedef struct {
  int64_t system_steps;    /* main steps + one per recorded transit (final gradient step) */
  int64_t transits;        /* transits recorded (count <= ntt) */
  int64_t newton_iters;    /* findtransit! Newton iterations */
  double kernel_ms;        /* device time of the last run (HIP events, all launches) */
  /* Per launch, indexed by its class c (the largest body count it serves:
   * n = 2..9 share class 9, every larger n has its own; index 1 is the
   * two-wave launch of nbg_set_coop): device time (HIP
   * events on the stream it ran on) and the work units it completed per body
   * count n, launch_units[c][n] (main steps + one per recorded transit when
   * grad != 0).  A launch runs its own class first, then takes systems of
   * smaller classes.  Entries of absent classes are 0. */
  double launch_ms[13];
  int64_t launch_units[13][13];
  int32_t launch_grid[13];  /* workgroups (one wave each) of that launch */
  int32_t paired_systems;   /* systems the queues paired two per wave (nbg_set_systems_per_wave) */
} nbg_batch_stats;

int nbg_batch_create(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                     const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                     const double* x, const double* v, const double* m, const uint8_t* pair,
                     const double* jac_init, nbg_batch** out);

/* nbg_batch_create with the initial conditions computed on the GPU: per system
 * the elements (n x 7, column-major rows m, P, t0, ecosw, esinw, I, Omega, as
 * ElementsIC holds them) and, optionally, the hierarchy matrix (n x n; NULL =
 * fully nested [n, 1, ..., 1]) become x, v and jac_init -- init_nbody
 * (src/ics/init_nbody.jl:13-162) with kepler_init (kepler_init.jl:66-210) and
 * ekepler (kepler.jl:1-41), which the reference runs on the host in State(ic)
 * (Integrator.jl:73-103).  Agreement with a host computation is to rounding
 * (the device math library's atan2, sin and cos), not bit for bit. */
int nbg_batch_create_elements(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                              const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                              const double* elements, const double* hierarchy, const uint8_t* pair,
                              nbg_batch** out);
/* The batch's initial x, v (3 x n per system) and jac_init (7n x 7n per
 * system), concatenated in system order; any pointer may be NULL. */
int nbg_batch_initial(nbg_batch* b, double* x, double* v, double* jac_init);

/* Autotuner for rounds (no counterpart in the reference): from the batch's
 * initial state, one round of `steps` steps (nbg_batch_round) under each
 * candidate launch configuration of nbg_batch_autotune_candidate; keeps the
 * fastest in b's context and resets the batch. */
int nbg_batch_autotune_round(nbg_batch* b, int64_t steps, int32_t* best, double* cand_ms, int32_t max_cand);
/* Autotuner (no counterpart in the reference): runs a copy of batch b with
 * every system's step count scaled by frac (0 < frac <= 1) under each
 * candidate launch configuration -- the workgroup form (two-wave workgroups
 * for the longest systems: off, automatic, or the 16/32/64 longest), the
 * one-wave workgroups per CU and the systems per wave (one, or two for the
 * paired body counts) -- and keeps the fastest in b's context (the values
 * nbg_set_coop, nbg_set_launch and nbg_set_systems_per_wave set).  cand_ms[c] receives candidate
 * c's device time for c < max_cand (may be NULL), *best the chosen index.
 * The copy needs as much device memory as b.  The batch itself is not run. */
int nbg_batch_autotune(nbg_batch* b, double frac, int32_t* best, double* cand_ms, int32_t max_cand);
/* Candidate `index` of the autotuner (its nbg_set_coop mode and maximum,
 * and its workgroups per CU, 0 = as LDS allows); returns the number of
 * candidates. */
int nbg_batch_autotune_candidate(int32_t index, int32_t* coop_mode, int32_t* coop_max, int32_t* groups_per_cu);
/* The systems per wave (1 or 2) of autotuner candidate `index`; -1 when
 * there is no such candidate. */
int32_t nbg_batch_autotune_systems_per_wave(int32_t index);
/* Restore every system to its initial state (device-side copy). */
int nbg_batch_reset(nbg_batch* b);
/* Integrate all systems to completion: one persistent launch per class of
 * body counts (2..9 together, then 10, 11, 12 alone), systems longest first
 * within a launch, launches overlapped on streams; the longest systems may
 * run on two-wave workgroups in a launch of their own (nbg_set_coop). */
int nbg_batch_run(nbg_batch* b, nbg_batch_stats* stats);
/* One round of a batch (the unit of the throughput benchmark; no
 * counterpart in the reference): every system advances at most `steps`
 * integrator steps of its own run -- Transits.jl:140-170, transit search and,
 * at a run's end, calc_dtdelements! included -- in one set of persistent
 * launches.  With restart != 0, systems whose run had ended before the round
 * first start it again from their initial records (state, Jacobians,
 * counters; earlier transit outputs are overwritten as the new run counts).
 * Stats are the work of this round alone: its steps plus one unit per
 * transit recorded, its transits and Newton iterations. */
int nbg_batch_round(nbg_batch* b, int64_t steps, int32_t restart, nbg_batch_stats* stats);
/* Copy results to host (any pointer may be NULL). */
int nbg_batch_download(nbg_batch* b, double* tt, double* dtdq0, double* dtdelements, int64_t* count, double* x,
                       double* v);
void nbg_batch_destroy(nbg_batch* b);

/* TransitSnapshot for many systems in one persistent launch (the batched
 * form of nbg_integrate_snapshot; no counterpart in the reference).  System
 * s has n[s] bodies, step h[s], starts at t0[s] from x, v, m, jac_init (its
 * slices of the concatenated arrays, as for nbg_batch_create; jac_step = I,
 * errors zero) and is integrated to its nt[s] >= 1 times (concatenated in
 * `times`).  Outputs concatenated in system order: vsky, bsky2 (n x nt per
 * system), dbvdq0, dbvdelements (2 x n x nt x 7 x n per system; may be NULL
 * without grad). */
int nbg_batch_snapshot(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* t0,
                       const int64_t* nt, const double* times, int32_t grad, const double* x, const double* v,
                       const double* m, const uint8_t* pair, const double* jac_init, double* vsky, double* bsky2,
                       double* dbvdq0, double* dbvdelements);

/* CartesianOutput for many systems in one persistent launch (the batched
 * form of nbg_integrate_cartesian; no counterpart in the reference).  System
 * s (n[s] bodies, x, v, m slices as for nbg_batch_create, jac_step = I,
 * errors zero, time t0[s]) takes nstep[s] >= 1 steps of h[s] in the direction
 * check_step(t0[s], tmax[s]) and records the state before each.  records:
 * 1 = xerror/verror, 2 = jac_step, 4 = jac_error (the latter two need grad).
 * Records are concatenated in system order, each as nbg_integrate_cartesian
 * lays them out (out_t: nstep; out_x ...: 3 x n x nstep; out_jac ...:
 * 7n x 7n x nstep). */
int nbg_batch_cartesian(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                        const double* t0, const int64_t* nstep, int32_t grad, int32_t records, const double* x,
                        const double* v, const double* m, const uint8_t* pair, double* out_t, double* out_x,
                        double* out_v, double* out_xerror, double* out_verror, double* out_jac, double* out_jac_error);

/* One-shot convenience: create, run, download, destroy. */
int nbg_batch_transits(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                       const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                       const double* x, const double* v, const double* m, const uint8_t* pair,
                       const double* jac_init, double* tt, double* dtdq0, double* dtdelements, int64_t* count,
                       nbg_batch_stats* stats);

/* ---- Serving: continuous batching of integrate-plus-gradient requests ----
 * No counterpart in the reference (its callers run one system per call); this
 * is the request loop of a timing fit's likelihood evaluations.  Each request
 * is one (intr)(s, tt; grad) run as in nbg_integrate_transits.  A server owns
 * a device arena of max_bytes; requests are admitted first come first served
 * while their HBM footprint (nbg_server_footprint) fits, and every call of
 * nbg_server_step advances all admitted requests in one set of persistent
 * launches, so requests that arrive while others run join at the next round.
 * Rounds are balanced in time: the requests of the largest body count among
 * the admitted ones advance up to chunk_steps steps, the others up to
 * chunk_steps times the ratio of the two body counts' per-step costs.
 * Results are bitwise those of the one-system call.
 * submit, finished, fetch and status may be called from other threads while
 * a step runs; steps are serialised with the context's other calls. */
typedef struct nbg_server nbg_server;
int nbg_server_create(nbg_ctx* ctx, int64_t max_bytes, int64_t chunk_steps, nbg_server** out);
void nbg_server_destroy(nbg_server* s);
/* Queue one request (0-based ti; x, v: 3 x n; m: n; pair: n x n or NULL;
 * jac_init: 7n x 7n, needed when grad != 0); *id identifies it. */
int nbg_server_submit(nbg_server* s, int32_t n, double h, double tmax, double t0, int32_t ti, int64_t ntt,
                      int32_t grad, int32_t ntbv, const double* x, const double* v, const double* m,
                      const uint8_t* pair, const double* jac_init, int64_t* id);
/* One round: admit, advance every admitted request, retire finished ones. */
int nbg_server_step(nbg_server* s, int32_t* n_finished, double* kernel_ms);
/* Ids of finished, not yet fetched requests (at most max_ids). */
int nbg_server_finished(nbg_server* s, int64_t* ids, int32_t max_ids, int32_t* n);
/* Copy a finished request's outputs (layouts as nbg_integrate_transits; any
 * pointer may be NULL) and forget it. */
int nbg_server_fetch(nbg_server* s, int64_t id, double* tt, double* dtdq0, double* dtdelements, int64_t* count,
                     double* x, double* v);
int nbg_server_status(nbg_server* s, int64_t* pending, int64_t* active, int64_t* finished, int64_t* bytes_free);
/* Work done by every round so far: system-steps (main steps plus one per
 * recorded transit when gradients are on), rounds and their device time. */
int nbg_server_work(nbg_server* s, int64_t* system_steps, int64_t* rounds, double* device_ms);
/* Device bytes one request of n bodies and ntt transit slots occupies. */
int64_t nbg_server_footprint(int32_t n, int32_t ntbv, int64_t ntt);

/* Tuning knobs for the persistent launches (0 = automatic). */
int nbg_set_launch(nbg_ctx* ctx, int32_t max_groups_per_cu, int32_t streams);
/* Upper bound on persistent workgroups per launch (0 = no bound). */
int nbg_set_grid_cap(nbg_ctx* ctx, int32_t grid_cap);
/* Cooperative workgroups of batch runs (no counterpart in the reference):
 * a system of at most 9 bodies with a transit search can run on a workgroup
 * of two waves, the second running its transit searches (findtransit!,
 * timing.jl:30-110) from snapshots while the first integrates; results are
 * bitwise those of the one-wave kernel.  mode 0: off; 1 (default): in
 * batch runs, the systems whose estimated time exceeds the batch's work
 * spread over every wave slot of the GPU (the ones that would set the
 * batch's end); 2: the longest ones regardless, in batch runs and serving
 * rounds.  At most max_systems per launch (default 128); n <= 10. */
int nbg_set_coop(nbg_ctx* ctx, int32_t mode, int32_t max_systems);
/* Systems per wave of batch runs and rounds (no counterpart in the
 * reference; the pair-to-lane mapping of the pair sweeps ahl21.jl:25-78).
 * k = 1 (default): one system per wave, a level's Kepler pairs one per lane.
 * k = 2: transit runs of equal body count run two per wave in lockstep
 * (every n <= 9; n = 10 too where the n = 12 launch takes it; two records
 * with their Jacobians in HBM and one shared work area fit the LDS of the
 * launch's one-system layouts, so occupancy is unchanged): the Kepler drifts
 * and pair Jacobians of both systems' levels share the lanes; the rest of
 * each step, and the transit searches, run for one system, then the other.
 * Results are bitwise those of k = 1.  Serving rounds keep one system per
 * wave. */
int nbg_set_systems_per_wave(nbg_ctx* ctx, int32_t k);
/* Form of the Jacobian updates (every launch of ctx): the chain product
 * J += jac_ij * J of the pair sweeps (ahl21.jl:31-42,60-72) and phisalpha!'s
 * jac_phi (ahl21.jl:607-698).  mode 0 (default): the reference's order,
 * results bitwise equal to the CPU restatement.  mode 1: factored -- a pair
 * (i, j) contributes +mj and -mi times one shared term to the rows of body i
 * and of body j, so that term is formed once per pair: kep*(a_i - a_j) in the
 * sweeps (36 instead of 144 fused multiply-adds per column and pair) and
 * jac_phi's column term; and the zero-addend compensated passes of
 * comp_sum_matrix! over rows an operation leaves unchanged, which only
 * re-normalise pairs their last update normalised, are left out.  It rounds
 * differently: jac_step, jac_error, dtdq0
 * and dtdelements agree with mode 0 to ~1e-11 relative (the tolerance asked
 * of Jacobian entries is 1e-10), dq/dt and the transit times, whose Newton
 * iteration reads dq/dt, within 1e-12 (measured equal); x and v, which read
 * neither, stay bitwise. */
int nbg_set_row_update(nbg_ctx* ctx, int32_t mode);
/* Diagnostic only (no counterpart in the reference): bound every wait between
 * the two waves of a cooperative workgroup to `polls` polls (0: the defaults,
 * ~1 s for a free search slot and ~2 min for the next search).  A wait that
 * runs out makes the launch return an error instead of hanging; tests use a
 * tiny bound to check that path. */
int nbg_set_coop_wait(nbg_ctx* ctx, int64_t polls);

/* Diagnostic only (no counterpart in the reference): a plain device copy of
 * `bytes` (a multiple of 16) through a grid-stride kernel, run `reps` times;
 * *ms = average duration.  It reads and writes `bytes` per run: the reference
 * reading for HBM byte counters and the box's achieved copy bandwidth. */
int nbg_copy_calibrate(nbg_ctx* ctx, int64_t bytes, int32_t reps, double* ms);

/* Diagnostic only (no counterpart in the reference): the Jacobian row update
 * of one Kepler pair at n = 9 (12 rows x 63 columns, J + jac_ij * J with
 * Kahan compensation, ahl21.jl:31-42) repeated `reps` times by each of `grid`
 * one-wave workgroups on synthetic data in LDS, with mode 0 on the VALU
 * (168 fused multiply-adds per lane, as the integrator does) or mode 1 on
 * the FP64 matrix cores (16 v_mfma_f64_16x16x4f64 per update); *ms = kernel
 * time, j_out (12 x 63 row-major, may be NULL) = workgroup 0's final rows.
 * The measurement behind the choice of unit for the chain products. */
int nbg_probe_row_update(nbg_ctx* ctx, int32_t mode, int32_t reps, int32_t grid, double* ms, double* j_out);

/* Diagnostic only (no counterpart in the reference; the reference calls
 * Julia's Base sin/cos/exp/sinh/cosh/cbrt): out[i] = f(x[i]) evaluated by the
 * device build of the portable libm (nbg_libm.h) for fn 0 sin, 1 cos, 2 exp,
 * 3 sinh, 4 cosh, 5 cbrt; x and out are host arrays of `count` doubles.  Lets
 * a test compare the device build bit for bit with the host build. */
int nbg_libm_eval_device(nbg_ctx* ctx, int32_t fn, int64_t count, const double* x, double* out);

/* Diagnostic only: evaluates the device pair formulas (nbg_pair.h) on the
 * host CPU so the test suite can compare them bit for bit with the CPU
 * restatement without a GPU.  Not used by any integration path.
 * out: delxv[6], jac_kepler[48] (6 x 8 column-major), jac_mass[6], iters. */
int nbg_selftest_pair_host(const double* x0, const double* v0, double k, double h, int32_t drift_first,
                           double* out);

#ifdef __cplusplus
}
#endif

#endif /* NBG_H */
