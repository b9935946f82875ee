// nbg_types.h -- per-system work descriptors shared by host and device code.
//
// A batch is a set of independent systems.  Every system owns slices of a few
// flat device buffers (offsets below, in elements); layouts follow the
// reference's Julia (column-major) arrays so host copies need no transposes.
//
//  state  (double): x[3n] v[3n] xerror[3n] verror[3n] m[n] dqdt[7n] gsave[n]
//                   t  (scalars at the end, STATE_T0_SLOT..)
//  jac    (double): jac_step[7n*7n] jac_error[7n*7n] jac_init[7n*7n]
//  ws     (double): two compressed 6n x 7n blocks (Jacobian copy during the
//                   final gradient step of a transit); for n >= JGLOBAL_MIN, and
//                   in every batch record, two more: the Jacobian and its error
//                   during the run when they are kept in HBM
//  tt     (double): ntbv x n x ntt   (TransitTiming.tt / TransitParameters.ttbv)
//  dq     (double): ntbv x n x ntt x 7 x n (dtdq0); del (double) same (dtdelements,
//                   at off_del)
//  ints   (int64) : count[n] istep ntransits_done newton_iters_total
//  pair   (uint8) : n x n
//  traj   (double): MODE_CARTESIAN only, the state before each step (traj_* below)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NBG_TD __host__ __device__ inline
#else
#define NBG_TD inline
#endif

namespace nbg {

enum SysMode : int32_t {
  MODE_STEPS = 0,     // (intr)(s, N) / (intr)(s, time): nsteps steps of h, then an optional step of hlast
  MODE_TRANSITS = 1,  // (intr)(s, tt): step loop with transit search
  MODE_CARTESIAN = 2, // (intr)(s, o::CartesianOutput): nsteps steps of h, the state recorded before each
  MODE_SNAPSHOT = 3,  // (intr)(s, ts::TransitSnapshot): integrate to each of nsteps times, b and v_sky there
};

struct SysDesc {
  int32_t n;
  int32_t mode;
  int32_t grad;
  int32_t ti;        // transited body, 0-based
  int32_t ntbv;      // 1 TransitTiming, 3 TransitParameters
  int32_t any_kick;  // some pair[i,j] (i<j) is true
  int32_t tag;       // reported in SysBufs.done_list when the system finishes
  int32_t traj;      // MODE_CARTESIAN: TRAJ_* flags of the optional records
  int32_t rowupd;    // Jacobian row update of the pair sweeps: 0 the reference's order, 1 factored (set per launch)
  int32_t pad_;
  int64_t nsteps;    // full steps in the run
  int64_t chunk;     // > 0: at most this many steps per visit (overrides the launch's chunk)
  int64_t ntt;
  double h;          // signed step used by the loop
  double hstep;      // intr.h (unsigned) for the first transit-time guess
  double hlast;      // MODE_STEPS: extra final step (0 = none)
  double t0;         // time at step 0 of the run (t = t0 + istep*h)
  int64_t off_state, off_jac, off_ws, off_tt, off_dq, off_del, off_int, off_pair, off_traj;
};

// The systems one launch may take: its own launch class first, then the
// classes of fewer bodies (their LDS layouts fit in its LDS), each a queue of
// descriptors ordered longest first with an atomic cursor.
constexpr int MAX_GROUPS = 8;
struct GroupList {
  int32_t ng;
  int32_t nsys[MAX_GROUPS];
  const SysDesc* descs[MAX_GROUPS];
  int* queue[MAX_GROUPS];
  // two systems per wave (nbg_run_duo): the queue hands out units of one or
  // two systems; unit u is descs[g][ustart[g][u] .. ustart[g][u + 1])
  int32_t nunits[MAX_GROUPS];
  const int* ustart[MAX_GROUPS];
  int32_t rowupd;  // SysDesc::rowupd of every system of the launch
};

struct SysBufs {
  double* state;
  double* jac;
  double* ws;
  double* tt;
  double* dq;
  double* del;
  int64_t* ints;
  uint8_t* pair;
  double* traj;        // trajectory records (null when no system records one)
  int* done_count;     // optional (null): tags of the systems that finished
  int32_t* done_list;  // during a launch are appended here
};

// Cooperative workgroups (two waves per system, batch launches only): the
// HBM slots that carry a transit search's Jacobian snapshot from the wave
// that integrates to the wave that searches.  Each workgroup owns `stride`
// doubles from jobs + blockIdx.x * stride: COOP_RQ slots of 2 * 6n * 7n.
constexpr int COOP_RQ = 4;
struct CoopBufs {
  double* jobs;
  int64_t stride;
  int64_t spin_slot;  // polls before a wait for a free ring slot gives up
  int64_t spin_job;   // polls before a wait for a job (or the end) gives up
};

// Initial conditions on the GPU (one system): inputs at in + off_el are the
// elements (n x 7, column-major: m, P, t0, ecosw, esinw, I, Omega) followed by
// the hierarchy matrix (n x n); x and v go to the state record at off_state,
// jac_init (7n x 7n, column-major) to the jac buffer at off_jinit.
struct IcDesc {
  int32_t n, pad;
  double t0;
  int64_t off_el, off_state, off_jinit;
};

// state slots
NBG_TD int64_t state_len(int n) { return 13 * (int64_t)n + 7 * n + n + 4; }
NBG_TD int64_t off_x(int) { return 0; }
NBG_TD int64_t off_v(int n) { return 3 * (int64_t)n; }
NBG_TD int64_t off_xe(int n) { return 6 * (int64_t)n; }
NBG_TD int64_t off_ve(int n) { return 9 * (int64_t)n; }
NBG_TD int64_t off_m(int n) { return 12 * (int64_t)n; }
NBG_TD int64_t off_dqdt(int n) { return 13 * (int64_t)n; }
NBG_TD int64_t off_gsave(int n) { return 20 * (int64_t)n; }
NBG_TD int64_t off_t(int n) { return 21 * (int64_t)n; }
NBG_TD int64_t jac_len(int n) { return 3 * 49 * (int64_t)n * n; }
// Systems of at least this many bodies keep their Jacobian (and its Kahan
// error) in an HBM work area at the end of their ws record instead of LDS.
constexpr int JGLOBAL_MIN = 7;
NBG_TD int64_t ws_len(int n) { return (n >= JGLOBAL_MIN ? 4 : 2) * 42 * (int64_t)n * n; }
// Batch records: room for the Jacobian in HBM at every n (two systems per
// wave keep it there; nbg_run_duo).
NBG_TD int64_t ws_len_batch(int n) { return 4 * 42 * (int64_t)n * n; }
NBG_TD int64_t int_len(int n) { return (int64_t)n + 3; }

// Trajectory record of a MODE_CARTESIAN run of ns steps: the state before
// step k (k = 0..ns-1) as arrays over k, each column-major with k last:
// t[ns], x[3n x ns], v[3n x ns], then, by flag, xerror and verror
// (TRAJ_ERR), jac_step[7n x 7n x ns] (TRAJ_JAC), jac_error (TRAJ_JACERR).
constexpr int TRAJ_ERR = 1, TRAJ_JAC = 2, TRAJ_JACERR = 4;
NBG_TD int64_t traj_off_x(int, int64_t ns) { return ns; }
NBG_TD int64_t traj_off_v(int n, int64_t ns) { return ns + 3 * (int64_t)n * ns; }
NBG_TD int64_t traj_off_xe(int n, int64_t ns) { return ns + 6 * (int64_t)n * ns; }
NBG_TD int64_t traj_off_ve(int n, int64_t ns) { return ns + 9 * (int64_t)n * ns; }
NBG_TD int64_t traj_off_jac(int n, int64_t ns, int f) { return ns + ((f & TRAJ_ERR) ? 12 : 6) * (int64_t)n * ns; }
NBG_TD int64_t traj_off_jacerr(int n, int64_t ns, int f) {
  return traj_off_jac(n, ns, f) + ((f & TRAJ_JAC) ? 49 * (int64_t)n * n * ns : 0);
}
NBG_TD int64_t traj_len(int n, int64_t ns, int f) {
  return traj_off_jacerr(n, ns, f) + ((f & TRAJ_JACERR) ? 49 * (int64_t)n * n * ns : 0);
}

// Record of a MODE_SNAPSHOT run over nt times (in the traj buffer): the
// times (input), then vsky and bsky2 [n x nt], dbvdq0 and dbvdelements
// [2 x n x nt x 7 x n], column-major.
NBG_TD int64_t snap_off_vsky(int, int64_t nt) { return nt; }
NBG_TD int64_t snap_off_bsky2(int n, int64_t nt) { return nt + (int64_t)n * nt; }
NBG_TD int64_t snap_off_dq(int n, int64_t nt) { return nt + 2 * (int64_t)n * nt; }
NBG_TD int64_t snap_off_del(int n, int64_t nt) { return snap_off_dq(n, nt) + 14 * (int64_t)n * n * nt; }
NBG_TD int64_t snap_len(int n, int64_t nt) { return snap_off_del(n, nt) + 14 * (int64_t)n * n * nt; }

}  // namespace nbg
