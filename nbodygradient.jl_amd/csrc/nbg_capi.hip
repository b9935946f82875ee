// nbg_capi.hip -- host side of libnbg_hip.so: the C ABI declared in
// include/nbg.h.  Packs systems into the device records described in
// nbg_types.h, launches one persistent kernel per body count (overlapped on
// streams), and copies results back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/nbg.h"
#include "nbg_pair.h"
#include "nbg_types.h"

extern "C" hipError_t nbg_launch_group(int cls, nbg::GroupList gl, nbg::SysBufs bufs, int64_t chunk, int grid,
                                       unsigned long long* units, hipStream_t stream);
extern "C" size_t nbg_lds_bytes(int nb);
extern "C" int nbg_class_of(int n);
extern "C" hipError_t nbg_launch_copy(const void* src, void* dst, size_t bytes, int cu_count, hipStream_t stream);
extern "C" hipError_t nbg_launch_coop(nbg::GroupList gl, nbg::SysBufs bufs, nbg::CoopBufs coop, int64_t chunk, int grid,
                                      unsigned long long* units, hipStream_t stream);
extern "C" int64_t nbg_coop_stride();
extern "C" hipError_t nbg_launch_rewind(const nbg::SysDesc* descs, int nsys, double* state, const double* state0,
                                        double* jac, const double* jac0, int64_t* ints, int cu_count,
                                        hipStream_t stream);
// Device memory of nbg_integrate_transits_noalloc: a file-scope array in this
// translation unit's code object, present on every GPU the library runs on
// until the process ends.  Host code never uses its name as an address (on
// the host it stands for a host-side shadow); it uses the device address that
// hipGetSymbolAddress returns for each device (scratch_dev below).
static __device__ __align__(256) unsigned char scratch[NBG_NOALLOC_SCRATCH_BYTES];

extern "C" hipError_t nbg_launch_probe(int mode, int reps, int grid, double* out, hipStream_t stream);
extern "C" hipError_t nbg_launch_libm(int fn, int64_t count, const double* x, double* out, hipStream_t stream);
// Initial conditions kernel (nbg_ic.hip)
extern "C" hipError_t nbg_launch_ic(const void* descs, int nsys, const double* in, double* state, double* jac,
                                    hipStream_t stream);
extern "C" int nbg_coop_max_bodies();
// Two systems per wave (nbg_duo.hip)
extern "C" hipError_t nbg_launch_group_duo(int cls, nbg::GroupList gl, nbg::SysBufs bufs, int64_t chunk, int grid,
                                           unsigned long long* units, hipStream_t stream);
extern "C" int nbg_duo_pairs(int cls, int n);

namespace {

constexpr int NB_MIN = 2, NB_MAX = 12;

template <class T> struct DevVec {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t count) {
    if (count <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t c = std::max<size_t>(count, 1);
    hipError_t e = hipMalloc(&p, c * sizeof(T));
    if (e == hipSuccess) cap = c;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Device storage of a set of systems.
struct Store {
  DevVec<double> state, state0, jac, jac0, ws, tt, dq, del;
  DevVec<int64_t> ints;
  DevVec<uint8_t> pair;
  DevVec<nbg::SysDesc> desc;
  DevVec<int> ustart;  // units of the two-per-wave queues (Layout::ustart)
  DevVec<int> queue;
  DevVec<unsigned long long> units;  // work units completed, [launch][n]
  DevVec<double> coop;               // Jacobian snapshot slots of the cooperative workgroups
  DevVec<double> traj;               // trajectory records (CartesianOutput runs)
  void release() {
    state.release(); state0.release(); jac.release(); jac0.release(); ws.release();
    tt.release(); dq.release(); del.release(); ints.release(); pair.release(); desc.release(); queue.release();
    units.release(); coop.release(); traj.release(); ustart.release();
  }
};

struct Layout {
  std::vector<nbg::SysDesc> desc;  // system order
  int64_t n_state = 0, n_jac = 0, n_ws = 0, n_tt = 0, n_dq = 0, n_int = 0, n_pair = 0, n_traj = 0;
  // launch groups: systems sorted by (launch class, cost desc); gn holds the
  // class (the largest body count it serves), gcost an estimate of its time
  std::vector<int> order;
  std::vector<int> gn, goff, gcnt;
  std::vector<double> gcost;
  std::vector<double> cost;  // optional per-system cost (default: from nsteps)
  int pairs = 0;             // queues ordered for two systems per wave (nbg_run_duo)
  int wave_slots = 0;        // one-wave workgroups the GPU holds at once (pairing keeps at least as many units)
  // units of the two-per-wave queues: for group g, ustart[uoff[g] ..] are the
  // offsets (from the group's start) of its units, then the group's count
  std::vector<int> ustart, uoff;
};

double check_step(double t0, double tmax) {  // Integrator.jl:249-259
  auto sg = [](double x) { return x < 0 ? -1.0 : (x > 0 ? 1.0 : x); };
  if (std::fabs(tmax) > std::fabs(t0)) return sg(tmax);
  if (sg(tmax) != sg(t0)) return sg(tmax);
  return -1.0 * sg(tmax);
}

long long jround(double x) { return (long long)std::nearbyint(x); }  // Julia round(Int64, x): half to even

void group_layout(Layout& L);

// Estimated device time of system i (microseconds): steps x per-step time
// (transit search included) measured on MI355X for n = 3, 5, 7, 9 and fitted
// by a quadratic.  It only orders work (longest first), shares the GPU
// between launches and picks the systems worth a second wave.
double step_cost(int nb) {
  const double n = nb;
  return std::max(10.0, 3.0 * n * n + 39.0 * n - 55.0);
}

double sys_cost(const Layout& L, int i) {
  const auto& d = L.desc[i];
  const double steps = L.cost.empty() ? (double)(d.nsteps + 1) : L.cost[i];
  return steps * step_cost(d.n);
}

void finalize_layout(Layout& L) {
  int64_t s = 0, j = 0, w = 0, t = 0, q = 0, in = 0, pr = 0, tr = 0;
  for (auto& d : L.desc) {
    const int n = d.n;
    d.off_state = s; s += nbg::state_len(n);
    d.off_jac = j; j += nbg::jac_len(n);
    d.off_ws = w; w += nbg::ws_len_batch(n);
    d.off_tt = t; t += (int64_t)d.ntbv * n * d.ntt;
    d.off_dq = q; d.off_del = q; q += (int64_t)d.ntbv * n * d.ntt * 7 * n;
    d.off_int = in; in += nbg::int_len(n);
    d.off_pair = pr; pr += (int64_t)n * n;
    d.off_traj = tr;
    tr += d.mode == nbg::MODE_CARTESIAN ? nbg::traj_len(n, d.nsteps, d.traj)
          : d.mode == nbg::MODE_SNAPSHOT ? nbg::snap_len(n, d.nsteps)
                                         : 0;
  }
  L.n_state = s; L.n_jac = j; L.n_ws = w; L.n_tt = t; L.n_dq = q; L.n_int = in; L.n_pair = pr; L.n_traj = tr;
  for (size_t k = 0; k < L.desc.size(); ++k) L.desc[k].tag = (int32_t)k;
  group_layout(L);
}

// Order systems by (launch class, cost desc) and form the launch groups.
void group_layout(Layout& L) {
  const int ns = (int)L.desc.size();
  L.order.resize(ns);
  std::iota(L.order.begin(), L.order.end(), 0);
  auto cost = [&](int i) { return sys_cost(L, i); };
  std::stable_sort(L.order.begin(), L.order.end(), [&](int a, int b) {
    const int ca = nbg_class_of(L.desc[a].n), cb = nbg_class_of(L.desc[b].n);
    if (ca != cb) return ca > cb;
    return cost(a) > cost(b);
  });
  L.gn.clear(); L.goff.clear(); L.gcnt.clear(); L.gcost.clear();
  for (int k = 0; k < ns; ++k) {
    const int c = nbg_class_of(L.desc[L.order[k]].n);
    if (L.gn.empty() || L.gn.back() != c) {
      L.gn.push_back(c);
      L.goff.push_back(k);
      L.gcnt.push_back(0);
      L.gcost.push_back(0.0);
    }
    L.gcnt.back() += 1;
    L.gcost.back() += cost(L.order[k]);
  }
  L.ustart.clear();
  L.uoff.clear();
  if (!L.pairs) return;
  // Two systems per wave: the queues hand out units of one or two systems.
  // Pairing trades latency for throughput -- a pair's two runs take about
  // 0.9 x the sum of their one-wave times, each run that much longer -- so
  // only systems of the same body count whose pair stays below the batch's
  // work spread over every wave slot are paired, the shortest first, and only
  // as many pairs as leave at least one unit per wave slot.  Units keep the
  // longest-first order of their first system.
  double total = 0.0;
  for (int i = 0; i < ns; ++i) total += cost(i);
  const double floor_cost = L.wave_slots > 0 ? total / L.wave_slots : total;
  int64_t budget = L.wave_slots > 0 ? std::max<int64_t>(0, (int64_t)ns - L.wave_slots) : ns / 2;
  std::vector<int> partner(ns, -1);  // by system index
  for (size_t g = 0; g < L.gn.size() && budget > 0; ++g) {
    const int off = L.goff[g], cnt = L.gcnt[g], cls = L.gn[g];
    std::vector<std::vector<int>> pos(NB_MAX + 1);  // systems by body count, longest first
    for (int k = 0; k < cnt; ++k) pos[L.desc[L.order[off + k]].n].push_back(L.order[off + k]);
    // the shortest systems first, across body counts
    std::vector<int> at(NB_MAX + 1);
    for (int n = NB_MIN; n <= NB_MAX; ++n) at[n] = (int)pos[n].size() - 1;
    while (budget > 0) {
      int bn = -1;
      double bc = 0.0;
      for (int n = NB_MIN; n <= NB_MAX; ++n) {
        if (at[n] < 1 || !nbg_duo_pairs(cls, n)) continue;
        const double c = cost(pos[n][at[n] - 1]);
        if (bn < 0 || c < bc) { bn = n; bc = c; }
      }
      if (bn < 0 || 1.8 * bc > floor_cost) break;
      const int a = pos[bn][at[bn]], b = pos[bn][at[bn] - 1];
      partner[a] = b;
      partner[b] = a;
      at[bn] -= 2;
      --budget;
    }
  }
  for (size_t g = 0; g < L.gn.size(); ++g) {
    const int off = L.goff[g], cnt = L.gcnt[g];
    std::vector<int> sub(L.order.begin() + off, L.order.begin() + off + cnt);
    std::vector<char> placed(ns, 0);
    L.uoff.push_back((int)L.ustart.size());
    int at = off;
    for (int i : sub) {
      if (placed[i]) continue;
      L.ustart.push_back(at - off);
      L.order[at++] = i;
      placed[i] = 1;
      if (partner[i] >= 0) {
        L.order[at++] = partner[i];
        placed[partner[i]] = 1;
        // the launch's estimated work: a pair takes ~0.8 x its two one-wave times
        L.gcost[g] -= 0.2 * (cost(i) + cost(partner[i]));
      }
    }
    L.ustart.push_back(cnt);
  }
}

bool mass_rows_ok(int n, const double* J, const double* JE) {
  const int s = 7 * n;
  for (int b = 0; b < n; ++b) {
    const int r = 7 * b + 6;
    for (int c = 0; c < s; ++c) {
      if (J && J[r + (size_t)s * c] != (c == r ? 1.0 : 0.0)) return false;
      if (JE && JE[r + (size_t)s * c] != 0.0) return false;
    }
  }
  return true;
}

}  // namespace

struct nbg_ctx {
  int device = 0;
  std::string err;
  std::mutex mu;
  int cu_count = 0;
  std::string name;
  int max_groups_per_cu = 0;
  int grid_cap = 0;
  int coop_mode = 1;   // 0 off, 1 systems longer than the batch's throughput bound, 2 the longest ones
  int coop_max = 128;  // at most this many cooperative systems per batch
  int systems_per_wave = 1;  // batches: 2 runs transit systems of equal body count two per wave (nbg_run_duo)
  int row_update = 0;        // Jacobian row update: 0 the reference's order (bitwise), 1 factored
  int64_t coop_polls = 0;  // > 0: bound on every inter-wave wait (diagnostic); 0: the defaults
  int nstreams = 4;
  std::vector<hipStream_t> streams;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t gev[2 * 16] = {};  // start/end of each body-count launch (timed on its own stream)
  Store single;  // reused by the single-system entry points
  unsigned char* scratch_dev = nullptr;  // device address of `scratch` on this context's device
};

struct nbg_batch {
  nbg_ctx* ctx = nullptr;
  Layout L;
  Store S;
  int64_t steps_main = 0;
  int64_t order_chunk = 0;  // the visit length the device queue order is for (0: whole runs)
  int order_pairs = 0;      // ... and whether it pairs systems (two per wave)
  int order_slots = 0;      // ... and for how many wave slots
};

#define NBG_HIP(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ctx->err = std::string(#expr) + " failed: " + hipGetErrorString(e_);            \
      return -2;                                                                      \
    }                                                                                 \
  } while (0)

#define NBG_ARG(cond, msg)         \
  do {                             \
    if (!(cond)) {                 \
      ctx->err = (msg);            \
      return -1;                   \
    }                              \
  } while (0)

namespace {

int grid_for(nbg_ctx* ctx, int nb, int nsys) {
  const size_t lds = nbg_lds_bytes(nb);
  int per_cu = (int)(163840 / std::max<size_t>(lds, 1));
  per_cu = std::max(1, std::min(per_cu, 8));
  if (ctx->max_groups_per_cu > 0) per_cu = std::min(per_cu, ctx->max_groups_per_cu);
  int g = ctx->cu_count * per_cu;
  if (ctx->grid_cap > 0) g = std::min(g, ctx->grid_cap);
  return std::max(1, std::min(g, nsys));
}

// Allocate device storage for a layout and upload descriptors.
int store_alloc(nbg_ctx* ctx, Store& S, const Layout& L, bool keep_initial) {
  NBG_HIP(S.state.ensure(L.n_state));
  NBG_HIP(S.jac.ensure(L.n_jac));
  NBG_HIP(S.ws.ensure(L.n_ws));
  NBG_HIP(S.tt.ensure(L.n_tt));
  NBG_HIP(S.dq.ensure(L.n_dq));
  NBG_HIP(S.del.ensure(L.n_dq));
  NBG_HIP(S.ints.ensure(L.n_int));
  NBG_HIP(S.pair.ensure(L.n_pair));
  if (L.n_traj > 0) NBG_HIP(S.traj.ensure(L.n_traj));
  NBG_HIP(S.desc.ensure(L.desc.size()));
  NBG_HIP(S.queue.ensure(2 * L.gn.size() + 1));
  NBG_HIP(S.units.ensure(13 * L.gn.size() + 14));
  if (keep_initial) {
    NBG_HIP(S.state0.ensure(L.n_state));
    NBG_HIP(S.jac0.ensure(L.n_jac));
  }
  std::vector<nbg::SysDesc> sorted(L.desc.size());
  for (size_t k = 0; k < L.order.size(); ++k) sorted[k] = L.desc[L.order[k]];
  NBG_HIP(hipMemcpy(S.desc.p, sorted.data(), sorted.size() * sizeof(nbg::SysDesc), hipMemcpyHostToDevice));
  return 0;
}

nbg::SysBufs bufs_of(Store& S) {
  nbg::SysBufs B;
  B.state = S.state.p; B.jac = S.jac.p; B.ws = S.ws.p; B.tt = S.tt.p; B.dq = S.dq.p; B.del = S.del.p;
  B.ints = S.ints.p; B.pair = S.pair.p;
  B.traj = S.traj.p;
  B.done_count = nullptr; B.done_list = nullptr;
  return B;
}

// Launch every group (one persistent kernel per body count) and wait.
int run_groups(nbg_ctx* ctx, Store& S, const Layout& L, int64_t chunk, double* kernel_ms,
               std::vector<float>* group_ms = nullptr, std::vector<unsigned long long>* units = nullptr,
               const nbg::SysBufs* bufs = nullptr, std::vector<int>* group_grid = nullptr) {
  NBG_HIP(hipSetDevice(ctx->device));
  const size_t G = L.gn.size();
  if (G > (size_t)nbg::MAX_GROUPS) {
    ctx->err = "too many launch classes";
    return -1;
  }
  NBG_HIP(hipMemset(S.queue.p, 0, sizeof(int) * (2 * G + 1)));
  NBG_HIP(hipMemset(S.units.p, 0, sizeof(unsigned long long) * (13 * G + 14)));
  const nbg::SysBufs B = bufs ? *bufs : bufs_of(S);
  const int ns = (int)ctx->streams.size();
  NBG_HIP(hipEventRecord(ctx->ev0, 0));
  for (int s = 0; s < ns; ++s) NBG_HIP(hipStreamWaitEvent(ctx->streams[s], ctx->ev0, 0));
  // One launch per launch class (groups are in decreasing class order).  A
  // launch first drains its own class, then takes systems of the smaller
  // classes, so no launch idles while smaller work remains.  Grids: with one
  // launch it takes the whole GPU; with several, the GPU is shared in
  // proportion to each class's estimated work (CU-time: the per-step cost
  // model times the CU fraction one workgroup holds, by LDS or by its wave
  // slot, one wave per SIMD at this kernel's register use), every class but
  // the smallest with a margin -- the largest the widest, since no launch can
  // take its systems -- so that every workgroup is resident from the start
  // and the long systems of the large-n classes start first.
  std::vector<int> grids(G);
  std::vector<double> frac(G), demand(G);
  double tot = 0.0;
  for (size_t g = 0; g < G; ++g) {
    frac[g] = std::max((double)nbg_lds_bytes(L.gn[g]) / 163840.0, 0.25);
    demand[g] = L.gcost[g] * frac[g];
    tot += demand[g];
  }
  double used = 0.0;
  for (size_t g = 0; g < G; ++g) {
    const int full = grid_for(ctx, L.gn[g], L.gcnt[g]);
    if (G == 1 || ctx->grid_cap > 0 || tot <= 0.0) {
      grids[g] = full;
      continue;
    }
    double cus = ctx->cu_count * (demand[g] / tot);
    if (g == 0) cus *= 1.35;
    else if (g + 1 < G) cus *= 1.1;
    else cus = std::max(ctx->cu_count - used, 0.25 * ctx->cu_count * (demand[g] / tot));
    grids[g] = std::max(1, std::min(full, (int)(cus / frac[g])));
    used += grids[g] * frac[g];
  }
  if (group_grid) *group_grid = grids;
  // Cooperative systems: the longest transit-search systems of the classes
  // the two-wave kernel serves (n <= nbg_coop_max_bodies()), from the front of
  // each cost-ordered group, get a workgroup of two waves (the second runs
  // their transit searches).  Auto mode, for whole runs (batches), takes the
  // systems whose estimated time exceeds the whole launch's work spread over
  // every wave slot of the GPU -- the ones that would otherwise set the end.
  // Serving rounds (chunked) use it only when forced (mode 2): measured on
  // the C4 trace, two-wave rounds raised p99 latency by 12 % (each round's
  // end waits for the searches still in flight).
  std::vector<int> kg(G, 0);
  int K = 0;
  if ((ctx->coop_mode == 2 || (ctx->coop_mode == 1 && chunk == 0)) && ctx->grid_cap == 0 &&
      ctx->max_groups_per_cu == 0 && ctx->streams.size() > 1) {
    double total = 0.0;
    for (size_t g = 0; g < G; ++g) total += L.gcost[g];
    const double floor_cost = total / (4.0 * ctx->cu_count);
    const int cap = std::min(ctx->coop_max, ctx->cu_count / 2);
    for (size_t g = 0; g < G; ++g) {
      if (L.gn[g] > nbg_coop_max_bodies()) continue;
      while (K < cap && kg[g] < L.gcnt[g]) {
        const int i = L.order[L.goff[g] + kg[g]];
        if (L.desc[i].mode != nbg::MODE_TRANSITS) break;
        if (ctx->coop_mode == 1 && sys_cost(L, i) <= floor_cost) break;
        ++kg[g];
        ++K;
      }
      if (L.pairs) {  // the cooperative systems must end on a unit boundary of the two-per-wave queue
        const int* us = L.ustart.data() + L.uoff[g];
        int u = 0;
        while (us[u] < kg[g]) ++u;
        if (us[u] != kg[g]) {
          --kg[g];
          --K;
        }
      }
    }
  }
  const bool timed_coop = K > 0 && group_ms != nullptr;
  if (K > 0) {
    NBG_HIP(S.coop.ensure((int64_t)K * nbg_coop_stride()));
    nbg::GroupList cg{};
    cg.rowupd = ctx->row_update;
    for (size_t g = 0; g < G; ++g) {
      if (kg[g] == 0) continue;
      cg.nsys[cg.ng] = kg[g];
      cg.descs[cg.ng] = S.desc.p + L.goff[g];
      cg.queue[cg.ng] = S.queue.p + G + 1 + g;
      cg.ng += 1;
      grids[g] = std::max(1, grids[g] - 2 * kg[g]);  // each holds two wave slots of that launch
    }
    if (group_grid) *group_grid = grids;
    // wait bounds in polls of ~500 cycles: ~1 s for a free ring slot (one
    // search frees it), ~2 min for the next job (a whole run may have none)
    const int64_t slot_polls = ctx->coop_polls > 0 ? ctx->coop_polls : (int64_t)1 << 22;
    const int64_t job_polls = ctx->coop_polls > 0 ? ctx->coop_polls : (int64_t)1 << 29;
    const nbg::CoopBufs cb{S.coop.p, nbg_coop_stride(), slot_polls, job_polls};
    hipStream_t st = ctx->streams[0];
    if (timed_coop) NBG_HIP(hipEventRecord(ctx->gev[2 * 15], st));
    NBG_HIP(nbg_launch_coop(cg, B, cb, chunk, K, S.units.p + 13 * G, st));
    if (timed_coop) NBG_HIP(hipEventRecord(ctx->gev[2 * 15 + 1], st));
    // a batch: let the two-wave workgroups (the largest LDS blocks) be placed
    // before the one-wave launches fill the CUs (a serving round is short and
    // leaves most of the GPU free)
    if (chunk == 0) std::this_thread::sleep_for(std::chrono::microseconds(2000));
  }
  // Launch order: the largest class first, alone for a moment, so that its
  // workgroups (the largest LDS blocks, the hardest to place) are resident
  // before the smaller classes fill the remaining LDS of every CU; then the
  // others, each on the stream with the least queued work (launches on one
  // stream run one after the other).
  std::vector<size_t> gord(G);
  std::iota(gord.begin(), gord.end(), 0);
  std::vector<double> load(ns, 0.0);
  for (size_t g : gord) {
    if (g == 1 && chunk == 0) std::this_thread::sleep_for(std::chrono::microseconds(2000));
    nbg::GroupList gl{};
    gl.rowupd = ctx->row_update;
    for (size_t h = g; h < G; ++h) {  // own class, then the smaller ones
      const int skip = kg[h];  // the cooperative systems
      gl.nsys[gl.ng] = L.gcnt[h] - skip;
      gl.descs[gl.ng] = S.desc.p + L.goff[h] + skip;
      gl.queue[gl.ng] = S.queue.p + h;
      if (L.pairs) {  // units from the first one after the cooperative systems; offsets from the group's start
        const int* us = L.ustart.data() + L.uoff[h];
        const int nu = (int)((h + 1 < G ? L.uoff[h + 1] : (int)L.ustart.size()) - L.uoff[h]) - 1;
        int u0 = 0;
        while (us[u0] < skip) ++u0;
        gl.descs[gl.ng] = S.desc.p + L.goff[h];
        gl.nunits[gl.ng] = nu - u0;
        gl.ustart[gl.ng] = S.ustart.p + L.uoff[h] + u0;
      }
      gl.ng += 1;
    }
    const int grid = grids[g];
    if (K > 0) load[0] += 1e30;  // the cooperative launch holds stream 0
    const int si = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[si] += L.gcost[g] / grid;
    hipStream_t st = ctx->streams[si];
    const bool timed = group_ms != nullptr;
    int64_t work = 0;  // systems this launch can take (none left when all went to cooperative workgroups)
    for (int e = 0; e < gl.ng; ++e) work += gl.nsys[e];
    if (timed) NBG_HIP(hipEventRecord(ctx->gev[2 * g], st));
    if (work > 0)
      NBG_HIP(L.pairs ? nbg_launch_group_duo(L.gn[g], gl, B, chunk, grid, S.units.p + 13 * g, st)
                      : nbg_launch_group(L.gn[g], gl, B, chunk, grid, S.units.p + 13 * g, st));
    if (timed) NBG_HIP(hipEventRecord(ctx->gev[2 * g + 1], st));
  }
  for (int s = 0; s < ns; ++s) {
    hipEvent_t e;
    NBG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NBG_HIP(hipEventRecord(e, ctx->streams[s]));
    NBG_HIP(hipStreamWaitEvent(0, e, 0));
    NBG_HIP(hipEventDestroy(e));
  }
  NBG_HIP(hipEventRecord(ctx->ev1, 0));
  NBG_HIP(hipEventSynchronize(ctx->ev1));
  NBG_HIP(hipGetLastError());
  if (kernel_ms) {
    float ms = 0;
    NBG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms = ms;
  }
  if (group_ms) {
    group_ms->assign(G, 0.0f);
    for (size_t g = 0; g < G; ++g) NBG_HIP(hipEventElapsedTime(&(*group_ms)[g], ctx->gev[2 * g], ctx->gev[2 * g + 1]));
  }
  if (K > 0) {
    unsigned long long werr = 0;
    NBG_HIP(hipMemcpy(&werr, S.units.p + 13 * G + 13, sizeof(werr), hipMemcpyDeviceToHost));
    if (werr) {
      ctx->err = "cooperative workgroup: a wait between its two waves timed out";
      return -2;
    }
  }
  if (units) {
    units->assign(13 * (G + 1), 0);
    NBG_HIP(hipMemcpy(units->data(), S.units.p, sizeof(unsigned long long) * 13 * (G + 1), hipMemcpyDeviceToHost));
  }
  if (group_ms) {
    group_ms->resize(G + 1, 0.0f);
    (*group_ms)[G] = 0.0f;
    if (timed_coop) NBG_HIP(hipEventElapsedTime(&(*group_ms)[G], ctx->gev[2 * 15], ctx->gev[2 * 15 + 1]));
  }
  if (group_grid) group_grid->push_back(K);
  return 0;
}

// Host-side staging of one system's record.
struct HostRec {
  std::vector<double> state, jac;
  std::vector<int64_t> ints;
  std::vector<uint8_t> pair;
};

void pack_system(const nbg::SysDesc& d, const double* x, const double* v, const double* xe, const double* ve,
                 const double* m, const uint8_t* pair, const double* J, const double* JE, const double* dqdt,
                 const double* jinit, double t, const double* gsave, const int64_t* count, HostRec& R) {
  const int n = d.n, s7 = 7 * n;
  R.state.assign(nbg::state_len(n), 0.0);
  for (int q = 0; q < 3 * n; ++q) {
    R.state[nbg::off_x(n) + q] = x[q];
    R.state[nbg::off_v(n) + q] = v[q];
    R.state[nbg::off_xe(n) + q] = xe ? xe[q] : 0.0;
    R.state[nbg::off_ve(n) + q] = ve ? ve[q] : 0.0;
  }
  for (int q = 0; q < n; ++q) {
    R.state[nbg::off_m(n) + q] = m[q];
    R.state[nbg::off_gsave(n) + q] = gsave ? gsave[q] : 0.0;
  }
  for (int q = 0; q < s7; ++q) R.state[nbg::off_dqdt(n) + q] = dqdt ? dqdt[q] : 0.0;
  R.state[nbg::off_t(n)] = t;
  const size_t NN = (size_t)s7 * s7;
  R.jac.assign(3 * NN, 0.0);
  for (size_t q = 0; q < NN; ++q) {
    R.jac[q] = J ? J[q] : ((q % (s7 + 1)) == 0 ? 1.0 : 0.0);
    R.jac[NN + q] = JE ? JE[q] : 0.0;
    R.jac[2 * NN + q] = jinit ? jinit[q] : ((q % (s7 + 1)) == 0 ? 1.0 : 0.0);
  }
  R.ints.assign(nbg::int_len(n), 0);
  if (count)
    for (int q = 0; q < n; ++q) R.ints[q] = count[q];
  R.pair.assign((size_t)n * n, 0);
  if (pair)
    for (int q = 0; q < n * n; ++q) R.pair[q] = pair[q] ? 1 : 0;
}

int any_kick(int n, const uint8_t* pair) {
  if (!pair) return 0;
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j)
      if (pair[i + n * j]) return 1;
  return 0;
}

// Host destinations of a MODE_CARTESIAN run's records (null: not recorded).
struct TrajOut {
  double *t, *x, *v, *xe, *ve, *jac, *jace;
  // MODE_SNAPSHOT: the times (input) and the b / v_sky outputs
  const double* times;
  double *vsky, *bsky2, *dbvdq0, *dbvdel;
};

// Run one system through the single-system store.
int run_single(nbg_ctx* ctx, nbg::SysDesc d, double* x, double* v, double* xe, double* ve, const double* m,
               const uint8_t* pair, double* J, double* JE, double* dqdt, const double* jinit, double* t,
               double* tt, double* dtdq0, double* dtdel, int64_t* count, const TrajOut* traj = nullptr) {
  Layout L;
  L.desc.push_back(d);
  finalize_layout(L);
  const nbg::SysDesc& D = L.desc[0];
  HostRec R;
  pack_system(D, x, v, xe, ve, m, pair, J, JE, dqdt, jinit, t ? *t : 0.0, nullptr, nullptr, R);
  Store& S = ctx->single;
  if (int rc = store_alloc(ctx, S, L, false)) return rc;
  NBG_HIP(hipMemcpy(S.state.p, R.state.data(), R.state.size() * sizeof(double), hipMemcpyHostToDevice));
  NBG_HIP(hipMemcpy(S.jac.p, R.jac.data(), R.jac.size() * sizeof(double), hipMemcpyHostToDevice));
  NBG_HIP(hipMemcpy(S.ints.p, R.ints.data(), R.ints.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  NBG_HIP(hipMemcpy(S.pair.p, R.pair.data(), R.pair.size(), hipMemcpyHostToDevice));
  const size_t ntt_len = (size_t)L.n_tt, ndq = (size_t)L.n_dq;
  if (D.mode == nbg::MODE_SNAPSHOT) {
    // the times in, every output zero (body 1's entries, and every
    // derivative without grad, stay so)
    NBG_HIP(hipMemset(S.traj.p + D.off_traj, 0, nbg::snap_len(D.n, D.nsteps) * sizeof(double)));
    NBG_HIP(hipMemcpy(S.traj.p + D.off_traj, traj->times, D.nsteps * sizeof(double), hipMemcpyHostToDevice));
  }
  if (D.mode == nbg::MODE_TRANSITS) {
    // outputs start from the caller's arrays (entries past count stay as given)
    NBG_HIP(hipMemcpy(S.tt.p, tt, ntt_len * sizeof(double), hipMemcpyHostToDevice));
    if (dtdq0) NBG_HIP(hipMemcpy(S.dq.p, dtdq0, ndq * sizeof(double), hipMemcpyHostToDevice));
    else NBG_HIP(hipMemset(S.dq.p, 0, ndq * sizeof(double)));
    if (dtdel) NBG_HIP(hipMemcpy(S.del.p, dtdel, ndq * sizeof(double), hipMemcpyHostToDevice));
    else NBG_HIP(hipMemset(S.del.p, 0, ndq * sizeof(double)));
  }
  if (int rc = run_groups(ctx, S, L, 0, nullptr)) return rc;
  NBG_HIP(hipMemcpy(R.state.data(), S.state.p, R.state.size() * sizeof(double), hipMemcpyDeviceToHost));
  NBG_HIP(hipMemcpy(R.ints.data(), S.ints.p, R.ints.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  const int n = D.n, s7 = 7 * n;
  for (int q = 0; q < 3 * n; ++q) {
    x[q] = R.state[nbg::off_x(n) + q];
    v[q] = R.state[nbg::off_v(n) + q];
    if (xe) xe[q] = R.state[nbg::off_xe(n) + q];
    if (ve) ve[q] = R.state[nbg::off_ve(n) + q];
  }
  if (D.grad) {
    const size_t NN = (size_t)s7 * s7;
    NBG_HIP(hipMemcpy(R.jac.data(), S.jac.p, 2 * NN * sizeof(double), hipMemcpyDeviceToHost));
    if (J) std::memcpy(J, R.jac.data(), NN * sizeof(double));
    if (JE) std::memcpy(JE, R.jac.data() + NN, NN * sizeof(double));
    if (dqdt)
      for (int q = 0; q < s7; ++q) dqdt[q] = R.state[nbg::off_dqdt(n) + q];
  }
  if (D.mode == nbg::MODE_SNAPSHOT) {
    const int64_t nt = D.nsteps;
    const double* base = S.traj.p + D.off_traj;
    NBG_HIP(hipMemcpy(traj->vsky, base + nbg::snap_off_vsky(n, nt), (size_t)n * nt * sizeof(double),
                      hipMemcpyDeviceToHost));
    NBG_HIP(hipMemcpy(traj->bsky2, base + nbg::snap_off_bsky2(n, nt), (size_t)n * nt * sizeof(double),
                      hipMemcpyDeviceToHost));
    const size_t nq = (size_t)14 * n * n * nt;
    if (traj->dbvdq0)
      NBG_HIP(hipMemcpy(traj->dbvdq0, base + nbg::snap_off_dq(n, nt), nq * sizeof(double), hipMemcpyDeviceToHost));
    if (traj->dbvdel)
      NBG_HIP(hipMemcpy(traj->dbvdel, base + nbg::snap_off_del(n, nt), nq * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (D.mode == nbg::MODE_CARTESIAN && traj) {
    const int64_t ns = D.nsteps;
    const int f = D.traj;
    auto get = [&](double* dst, int64_t off, int64_t len) -> int {
      if (dst) NBG_HIP(hipMemcpy(dst, S.traj.p + D.off_traj + off, (size_t)len * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
    };
    const int64_t n3 = 3 * (int64_t)n * ns, nj = 49 * (int64_t)n * n * ns;
    if (int rc = get(traj->t, 0, ns)) return rc;
    if (int rc = get(traj->x, nbg::traj_off_x(n, ns), n3)) return rc;
    if (int rc = get(traj->v, nbg::traj_off_v(n, ns), n3)) return rc;
    if (f & nbg::TRAJ_ERR) {
      if (int rc = get(traj->xe, nbg::traj_off_xe(n, ns), n3)) return rc;
      if (int rc = get(traj->ve, nbg::traj_off_ve(n, ns), n3)) return rc;
    }
    if (f & nbg::TRAJ_JAC)
      if (int rc = get(traj->jac, nbg::traj_off_jac(n, ns, f), nj)) return rc;
    if (f & nbg::TRAJ_JACERR)
      if (int rc = get(traj->jace, nbg::traj_off_jacerr(n, ns, f), nj)) return rc;
  }
  if (D.mode == nbg::MODE_TRANSITS) {
    if (t) *t = R.state[nbg::off_t(n)];
    for (int q = 0; q < n; ++q) count[q] = R.ints[q];
    NBG_HIP(hipMemcpy(tt, S.tt.p, ntt_len * sizeof(double), hipMemcpyDeviceToHost));
    if (dtdq0) NBG_HIP(hipMemcpy(dtdq0, S.dq.p, ndq * sizeof(double), hipMemcpyDeviceToHost));
    if (dtdel) NBG_HIP(hipMemcpy(dtdel, S.del.p, ndq * sizeof(double), hipMemcpyDeviceToHost));
  }
  return 0;
}

int check_common(nbg_ctx* ctx, int32_t n, double h, const double* x, const double* v, const double* m) {
  NBG_ARG(ctx != nullptr, "null context");
  NBG_ARG(n >= NB_MIN && n <= NB_MAX, "n must be in [2, 12] for the GPU path");
  NBG_ARG(x && v && m, "x, v and m are required");
  NBG_ARG(std::isfinite(h) && h != 0.0, "step h must be finite and non-zero");
  return 0;
}

}  // namespace

extern "C" {

int nbg_ctx_create(nbg_ctx** out, int device) {
  if (!out) return -1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return -3;
  auto* ctx = new nbg_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return -3; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return -3; }
  ctx->cu_count = prop.multiProcessorCount;
  ctx->name = prop.gcnArchName;
  ctx->streams.resize(ctx->nstreams);
  for (auto& s : ctx->streams)
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { delete ctx; return -3; }
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) { delete ctx; return -3; }
  for (auto& e : ctx->gev)
    if (hipEventCreate(&e) != hipSuccess) { delete ctx; return -3; }
  // the device address of the no-allocation scratch array on this device
  // (each GPU holds its own copy at its own address)
  {
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) { delete ctx; return -3; }
      it = addr.emplace(device, static_cast<unsigned char*>(p)).first;
    }
    ctx->scratch_dev = it->second;
  }
  *out = ctx;
  return 0;
}

void nbg_ctx_destroy(nbg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  ctx->single.release();
  for (auto& s : ctx->streams) (void)hipStreamDestroy(s);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (auto& e : ctx->gev)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
}

const char* nbg_last_error(const nbg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int nbg_device_info(nbg_ctx* ctx, char* name, int name_len, int* cu_count) {
  if (!ctx) return -1;
  if (name && name_len > 0) {
    std::snprintf(name, name_len, "%s", ctx->name.c_str());
  }
  if (cu_count) *cu_count = ctx->cu_count;
  return 0;
}

size_t nbg_lds_bytes_per_system(int n) { return (n >= 2 && n <= 12) ? nbg_lds_bytes(nbg_class_of(n)) : 0; }

int nbg_set_launch(nbg_ctx* ctx, int32_t max_groups_per_cu, int32_t streams) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->max_groups_per_cu = std::max(0, max_groups_per_cu);
  if (streams > 0 && streams != (int)ctx->streams.size()) {
    (void)hipDeviceSynchronize();
    for (auto& s : ctx->streams) (void)hipStreamDestroy(s);
    ctx->streams.assign(streams, nullptr);
    for (auto& s : ctx->streams) NBG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  return 0;
}

int nbg_set_coop(nbg_ctx* ctx, int32_t mode, int32_t max_systems) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(mode >= 0 && mode <= 2, "coop mode must be 0, 1 or 2");
  NBG_ARG(max_systems >= 0, "max_systems must be >= 0");
  ctx->coop_mode = mode;
  ctx->coop_max = max_systems;
  return 0;
}

int nbg_set_row_update(nbg_ctx* ctx, int32_t mode) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(mode == 0 || mode == 1, "row update mode must be 0 (reference order) or 1 (factored)");
  ctx->row_update = mode;
  return 0;
}

int nbg_set_systems_per_wave(nbg_ctx* ctx, int32_t k) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(k == 1 || k == 2, "systems per wave must be 1 or 2");
  ctx->systems_per_wave = k;
  return 0;
}

int nbg_set_coop_wait(nbg_ctx* ctx, int64_t polls) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(polls >= 0, "polls must be >= 0");
  ctx->coop_polls = polls;
  return 0;
}

int nbg_set_grid_cap(nbg_ctx* ctx, int32_t grid_cap) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->grid_cap = std::max(0, grid_cap);
  return 0;
}

int nbg_ahl21_step(nbg_ctx* ctx, int32_t n, double h, int32_t grad, double* x, double* v, double* xerror,
                   double* verror, const double* m, const uint8_t* pair, double* jac_step, double* jac_error,
                   double* dqdt) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(!grad || (jac_step && jac_error && dqdt), "grad step needs jac_step, jac_error and dqdt");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_STEPS; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1; d.any_kick = any_kick(n, pair);
  d.nsteps = 1; d.ntt = 0; d.h = h; d.hstep = h; d.hlast = 0.0; d.t0 = 0.0;
  return run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, nullptr, nullptr, nullptr,
                    nullptr, nullptr, nullptr);
}

int nbg_integrate_steps(nbg_ctx* ctx, int32_t n, double h, int64_t N, int32_t grad, double* t, double* x, double* v,
                        double* xerror, double* verror, const double* m, const uint8_t* pair, double* jac_step,
                        double* jac_error, double* dqdt) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t != nullptr, "t is required");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt), "grad integration needs jac_step, jac_error and dqdt");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  double hs = h;
  if (N < 0) { hs = -h; N = -N; }
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_STEPS; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1; d.any_kick = any_kick(n, pair);
  d.nsteps = N; d.ntt = 0; d.h = hs; d.hstep = h; d.hlast = 0.0; d.t0 = *t;
  if (N > 0) {
    int rc = run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  // s.t[1], s2 = comp_sum(s.t[1], s2, intr.h) per step (Integrator.jl:228)
  double tv = *t, s2 = 0.0;
  for (int64_t q = 0; q < N; ++q) {
    s2 += hs;
    const double tmp = tv + s2;
    s2 = (tv - tmp) + s2;
    tv = tmp;
  }
  *t = tv;
  return 0;
}

int nbg_integrate_to(nbg_ctx* ctx, int32_t n, double h, double time, int32_t grad, double* t, double* x, double* v,
                     double* xerror, double* verror, const double* m, const uint8_t* pair, double* jac_step,
                     double* jac_error, double* dqdt) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t != nullptr, "t is required");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt), "grad integration needs jac_step, jac_error and dqdt");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  const double t0 = *t;
  const long long nsteps = std::llabs(jround((time - t0) / h));
  const double hs = h * check_step(t0, time);
  const double tmax = t0 + (hs * (double)nsteps);
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_STEPS; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1; d.any_kick = any_kick(n, pair);
  d.nsteps = nsteps; d.ntt = 0; d.h = hs; d.hstep = h; d.hlast = (tmax != time) ? time - tmax : 0.0; d.t0 = t0;
  if (nsteps > 0 || d.hlast != 0.0) {
    int rc = run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  *t = time;
  return 0;
}

int nbg_integrate_transits(nbg_ctx* ctx, int32_t n, double h, double tmax, int32_t ti, int32_t grad, int32_t ntbv,
                           double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                           const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt,
                           const double* jac_init, int64_t ntt, double* tt, double* dtdq0, double* dtdelements,
                           int64_t* count) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t && tt && count, "t, tt and count are required");
  NBG_ARG(ti >= 0 && ti < n, "ti out of range");
  NBG_ARG(ntbv == 1 || ntbv == 3, "ntbv must be 1 (TransitTiming) or 3 (TransitParameters)");
  NBG_ARG(ntt >= 0, "ntt must be >= 0");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt && jac_init && dtdq0 && dtdelements),
          "grad run needs jac_step, jac_error, dqdt, jac_init, dtdq0 and dtdelements");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  const double t0 = *t;
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_TRANSITS; d.grad = grad ? 1 : 0; d.ti = ti; d.ntbv = ntbv;
  d.any_kick = any_kick(n, pair);
  d.nsteps = std::llabs(jround(tmax / h));
  d.ntt = ntt; d.h = h * check_step(t0, tmax + t0); d.hstep = h; d.hlast = 0.0; d.t0 = t0;
  for (int q = 0; q < n; ++q) count[q] = 0;
  return run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, jac_init, t, tt, dtdq0,
                    dtdelements, count);
}

int nbg_integrate_transits_noalloc(nbg_ctx* ctx, int32_t n, double h, double tmax, int32_t ti, int32_t grad,
                                   int32_t ntbv, double* t, double* x, double* v, double* xerror, double* verror,
                                   const double* m, const uint8_t* pair, double* jac_step, double* jac_error,
                                   double* dqdt, const double* jac_init, int64_t ntt, double* tt, double* dtdq0,
                                   double* dtdelements, int64_t* count) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t && tt && count, "t, tt and count are required");
  NBG_ARG(ti >= 0 && ti < n, "ti out of range");
  NBG_ARG(ntbv == 1 || ntbv == 3, "ntbv must be 1 (TransitTiming) or 3 (TransitParameters)");
  NBG_ARG(ntt >= 0, "ntt must be >= 0");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt && jac_init && dtdq0 && dtdelements),
          "grad run needs jac_step, jac_error, dqdt, jac_init, dtdq0 and dtdelements");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  NBG_ARG(ctx->scratch_dev != nullptr, "no scratch address for this device");
  // the system's record, piece by piece, in the scratch array (byte offsets
  // rounded up to 256)
  const int64_t s7 = 7 * (int64_t)n, NN = s7 * s7;
  const int64_t ntv = (int64_t)ntbv * n * ntt, ndq = ntv * s7;
  auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  int64_t o = 0;
  const int64_t o_desc = o; o += up((int64_t)sizeof(nbg::SysDesc));
  const int64_t o_queue = o; o += up(16);
  const int64_t o_units = o; o += up(14 * 8);
  const int64_t o_state = o; o += up(8 * nbg::state_len(n));
  const int64_t o_jac = o; o += up(8 * nbg::jac_len(n));
  const int64_t o_ws = o; o += up(8 * nbg::ws_len(n));
  const int64_t o_tt = o; o += up(8 * ntv);
  const int64_t o_dq = o; o += up(8 * ndq);
  const int64_t o_del = o; o += up(8 * ndq);
  const int64_t o_int = o; o += up(8 * nbg::int_len(n));
  const int64_t o_pair = o; o += up((int64_t)n * n);
  if (o > (int64_t)NBG_NOALLOC_SCRATCH_BYTES) {
    ctx->err = "nbg_integrate_transits_noalloc: this system needs " + std::to_string(o) +
               " bytes of device scratch, more than the " + std::to_string((long long)NBG_NOALLOC_SCRATCH_BYTES) +
               " the library holds (n, ntt or ntbv too large); use nbg_integrate_transits";
    return -1;
  }
  const double t0 = *t;
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_TRANSITS; d.grad = grad ? 1 : 0; d.ti = ti; d.ntbv = ntbv;
  d.any_kick = any_kick(n, pair);
  d.nsteps = std::llabs(jround(tmax / h));
  d.ntt = ntt; d.h = h * check_step(t0, tmax + t0); d.hstep = h; d.hlast = 0.0; d.t0 = t0;
  // every buffer pointer starts at its own piece, so every offset is zero
  d.off_state = d.off_jac = d.off_ws = d.off_tt = d.off_dq = d.off_del = d.off_int = d.off_pair = d.off_traj = 0;
  d.tag = 0;
  unsigned char* base = ctx->scratch_dev;
  auto dev = [&](int64_t off) { return base + off; };
  NBG_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->streams[0];
  const size_t n3 = 3 * (size_t)n;
  // inputs: descriptor, state (x, v, Kahan errors, m, dq/dt, gsave = 0, t),
  // Jacobians, counters (count starts at 0), pair mask, the output arrays'
  // current contents (entries beyond count stay as given)
  NBG_HIP(hipMemcpyAsync(dev(o_desc), &d, sizeof(d), hipMemcpyHostToDevice, st));
  NBG_HIP(hipMemsetAsync(dev(o_queue), 0, 16, st));
  NBG_HIP(hipMemsetAsync(dev(o_units), 0, 14 * 8, st));
  NBG_HIP(hipMemsetAsync(dev(o_state), 0, 8 * nbg::state_len(n), st));
  NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_x(n)), x, 8 * n3, hipMemcpyHostToDevice, st));
  NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_v(n)), v, 8 * n3, hipMemcpyHostToDevice, st));
  if (xerror) NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_xe(n)), xerror, 8 * n3, hipMemcpyHostToDevice, st));
  if (verror) NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_ve(n)), verror, 8 * n3, hipMemcpyHostToDevice, st));
  NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_m(n)), m, 8 * (size_t)n, hipMemcpyHostToDevice, st));
  if (dqdt) NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_dqdt(n)), dqdt, 8 * s7, hipMemcpyHostToDevice, st));
  NBG_HIP(hipMemcpyAsync(dev(o_state + 8 * nbg::off_t(n)), &t0, 8, hipMemcpyHostToDevice, st));
  if (grad) {
    NBG_HIP(hipMemcpyAsync(dev(o_jac), jac_step, 8 * NN, hipMemcpyHostToDevice, st));
    NBG_HIP(hipMemcpyAsync(dev(o_jac + 8 * NN), jac_error, 8 * NN, hipMemcpyHostToDevice, st));
    NBG_HIP(hipMemcpyAsync(dev(o_jac + 16 * NN), jac_init, 8 * NN, hipMemcpyHostToDevice, st));
  }
  NBG_HIP(hipMemsetAsync(dev(o_int), 0, 8 * nbg::int_len(n), st));
  if (pair) NBG_HIP(hipMemcpyAsync(dev(o_pair), pair, (size_t)n * n, hipMemcpyHostToDevice, st));
  else NBG_HIP(hipMemsetAsync(dev(o_pair), 0, (size_t)n * n, st));
  if (ntv) NBG_HIP(hipMemcpyAsync(dev(o_tt), tt, 8 * ntv, hipMemcpyHostToDevice, st));
  if (ndq) {
    if (dtdq0) NBG_HIP(hipMemcpyAsync(dev(o_dq), dtdq0, 8 * ndq, hipMemcpyHostToDevice, st));
    else NBG_HIP(hipMemsetAsync(dev(o_dq), 0, 8 * ndq, st));
    if (dtdelements) NBG_HIP(hipMemcpyAsync(dev(o_del), dtdelements, 8 * ndq, hipMemcpyHostToDevice, st));
    else NBG_HIP(hipMemsetAsync(dev(o_del), 0, 8 * ndq, st));
  }
  // one persistent workgroup of the system's launch class takes the system
  nbg::GroupList gl{};
  gl.rowupd = ctx->row_update;
  gl.ng = 1;
  gl.nsys[0] = 1;
  gl.descs[0] = reinterpret_cast<const nbg::SysDesc*>(dev(o_desc));
  gl.queue[0] = reinterpret_cast<int*>(dev(o_queue));
  nbg::SysBufs B;
  B.state = reinterpret_cast<double*>(dev(o_state));
  B.jac = reinterpret_cast<double*>(dev(o_jac));
  B.ws = reinterpret_cast<double*>(dev(o_ws));
  B.tt = reinterpret_cast<double*>(dev(o_tt));
  B.dq = reinterpret_cast<double*>(dev(o_dq));
  B.del = reinterpret_cast<double*>(dev(o_del));
  B.ints = reinterpret_cast<int64_t*>(dev(o_int));
  B.pair = dev(o_pair);
  B.traj = nullptr;
  B.done_count = nullptr;
  B.done_list = nullptr;
  NBG_HIP(nbg_launch_group(nbg_class_of(n), gl, B, 0, 1, reinterpret_cast<unsigned long long*>(dev(o_units)), st));
  // outputs, straight into the caller's arrays
  NBG_HIP(hipMemcpyAsync(x, dev(o_state + 8 * nbg::off_x(n)), 8 * n3, hipMemcpyDeviceToHost, st));
  NBG_HIP(hipMemcpyAsync(v, dev(o_state + 8 * nbg::off_v(n)), 8 * n3, hipMemcpyDeviceToHost, st));
  if (xerror) NBG_HIP(hipMemcpyAsync(xerror, dev(o_state + 8 * nbg::off_xe(n)), 8 * n3, hipMemcpyDeviceToHost, st));
  if (verror) NBG_HIP(hipMemcpyAsync(verror, dev(o_state + 8 * nbg::off_ve(n)), 8 * n3, hipMemcpyDeviceToHost, st));
  NBG_HIP(hipMemcpyAsync(t, dev(o_state + 8 * nbg::off_t(n)), 8, hipMemcpyDeviceToHost, st));
  NBG_HIP(hipMemcpyAsync(count, dev(o_int), 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  if (grad) {
    NBG_HIP(hipMemcpyAsync(jac_step, dev(o_jac), 8 * NN, hipMemcpyDeviceToHost, st));
    NBG_HIP(hipMemcpyAsync(jac_error, dev(o_jac + 8 * NN), 8 * NN, hipMemcpyDeviceToHost, st));
    NBG_HIP(hipMemcpyAsync(dqdt, dev(o_state + 8 * nbg::off_dqdt(n)), 8 * s7, hipMemcpyDeviceToHost, st));
  }
  if (ntv) NBG_HIP(hipMemcpyAsync(tt, dev(o_tt), 8 * ntv, hipMemcpyDeviceToHost, st));
  if (ndq && dtdq0) NBG_HIP(hipMemcpyAsync(dtdq0, dev(o_dq), 8 * ndq, hipMemcpyDeviceToHost, st));
  if (ndq && dtdelements) NBG_HIP(hipMemcpyAsync(dtdelements, dev(o_del), 8 * ndq, hipMemcpyDeviceToHost, st));
  NBG_HIP(hipStreamSynchronize(st));
  NBG_HIP(hipGetLastError());
  return 0;
}

int nbg_integrate_cartesian(nbg_ctx* ctx, int32_t n, double h, double tmax, int64_t nstep, int32_t grad,
                            double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                            const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt, double* out_t,
                            double* out_x, double* out_v, double* out_xerror, double* out_verror, double* out_jac,
                            double* out_jac_error) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t && out_t && out_x && out_v, "t, out_t, out_x and out_v are required");
  NBG_ARG(nstep >= 0, "nstep must be >= 0");
  NBG_ARG((out_xerror == nullptr) == (out_verror == nullptr), "out_xerror and out_verror go together");
  NBG_ARG(grad || (!out_jac && !out_jac_error), "Jacobian records need grad != 0");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt), "grad integration needs jac_step, jac_error and dqdt");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  // Outputs.jl:27-31: the direction from check_step(t0, intr.tmax), o.nstep steps
  const double t0 = *t;
  const double hs = h * check_step(t0, tmax);
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_CARTESIAN; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1;
  d.any_kick = any_kick(n, pair);
  d.traj = (out_xerror ? nbg::TRAJ_ERR : 0) | (out_jac ? nbg::TRAJ_JAC : 0) | (out_jac_error ? nbg::TRAJ_JACERR : 0);
  d.nsteps = nstep; d.ntt = 0; d.h = hs; d.hstep = h; d.hlast = 0.0; d.t0 = t0;
  if (nstep > 0) {
    TrajOut o{out_t, out_x, out_v, out_xerror, out_verror, out_jac, out_jac_error, nullptr, nullptr, nullptr, nullptr,
              nullptr};
    int rc = run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr, &o);
    if (rc) return rc;
  }
  *t = t0 + (hs * (double)nstep);  // s.t[1] = t0 + (h * i) after step i (Outputs.jl:42)
  return 0;
}

// ---------------------------------------------------------------- batches
int nbg_integrate_snapshot(nbg_ctx* ctx, int32_t n, double h, int64_t nt, const double* times, int32_t grad,
                           double* t, double* x, double* v, double* xerror, double* verror, const double* m,
                           const uint8_t* pair, double* jac_step, double* jac_error, double* dqdt,
                           const double* jac_init, double* vsky, double* bsky2, double* dbvdq0,
                           double* dbvdelements) {
  if (int rc = check_common(ctx, n, h, x, v, m)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(t && vsky && bsky2 && jac_init, "t, vsky, bsky2 and jac_init are required");
  NBG_ARG(nt >= 0 && (nt == 0 || times), "nt must be >= 0, with the times");
  NBG_ARG(!grad || (jac_step && jac_error && dqdt), "grad run needs jac_step, jac_error and dqdt");
  NBG_ARG(!grad || mass_rows_ok(n, jac_step, jac_error), "jac_step mass rows must be unit rows, jac_error's zero");
  for (int64_t j = 0; j < nt; ++j) NBG_ARG(std::isfinite(times[j]), "times must be finite");
  if (nt == 0) return 0;
  nbg::SysDesc d{};
  d.n = n; d.mode = nbg::MODE_SNAPSHOT; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1;
  d.any_kick = any_kick(n, pair);
  d.nsteps = nt; d.ntt = 0; d.h = h; d.hstep = h; d.hlast = 0.0; d.t0 = *t;
  TrajOut o{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, times, vsky, bsky2, dbvdq0, dbvdelements};
  int rc = run_single(ctx, d, x, v, xerror, verror, m, pair, jac_step, jac_error, dqdt, jac_init, t, nullptr, nullptr,
                      nullptr, nullptr, &o);  // t: the state's time at the start
  if (rc) return rc;
  *t = times[nt - 1];
  return 0;
}

// The batch's records, uploaded (state0 / jac0) but not yet copied to the
// working buffers (nbg_batch_reset does that).  With jac_init null, jac_init
// starts as the identity.
static int batch_build(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                       const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                       const double* x, const double* v, const double* m, const uint8_t* pair,
                       const double* jac_init, nbg_batch** out) {
  NBG_ARG(nsys > 0, "nsys must be positive");
  NBG_ARG(n && h && tmax && t0 && ntt && x && v && m, "missing per-system arrays");
  NBG_ARG(ntbv == 1 || ntbv == 3, "ntbv must be 1 or 3");
  auto* b = new nbg_batch();
  b->ctx = ctx;
  b->L.desc.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    if (n[s] < NB_MIN || n[s] > NB_MAX || !(std::isfinite(h[s]) && h[s] != 0.0) || ntt[s] < 0) {
      delete b;
      ctx->err = "system " + std::to_string(s) + ": n must be in [2,12], h finite non-zero, ntt >= 0";
      return -1;
    }
  }
  // per-system offsets into the packed host inputs
  std::vector<int64_t> ob(nsys + 1, 0), oj(nsys + 1, 0), op(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    ob[s + 1] = ob[s] + n[s];
    oj[s + 1] = oj[s] + 49LL * n[s] * n[s];
    op[s + 1] = op[s] + (int64_t)n[s] * n[s];
  }
  for (int s = 0; s < nsys; ++s) {
    nbg::SysDesc& d = b->L.desc[s];
    const int tis = ti ? ti[s] : 0;
    if (tis < 0 || tis >= n[s]) {
      delete b;
      ctx->err = "system " + std::to_string(s) + ": ti out of range";
      return -1;
    }
    d.n = n[s]; d.mode = nbg::MODE_TRANSITS; d.grad = grad ? 1 : 0; d.ti = tis; d.ntbv = ntbv;
    d.any_kick = pair ? any_kick(n[s], pair + op[s]) : 0;
    d.nsteps = std::llabs(jround(tmax[s] / h[s]));
    d.ntt = ntt[s]; d.h = h[s] * check_step(t0[s], tmax[s] + t0[s]); d.hstep = h[s]; d.hlast = 0.0; d.t0 = t0[s];
    b->steps_main += d.nsteps;
  }
  finalize_layout(b->L);
  if (int rc = store_alloc(ctx, b->S, b->L, true)) {
    b->S.release();
    delete b;
    return rc;
  }
  std::vector<double> hs((size_t)b->L.n_state), hj((size_t)b->L.n_jac);
  std::vector<int64_t> hi((size_t)b->L.n_int, 0);
  std::vector<uint8_t> hp((size_t)b->L.n_pair, 0);
  HostRec R;
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = b->L.desc[s];
    pack_system(d, x + 3 * ob[s], v + 3 * ob[s], nullptr, nullptr, m + ob[s], pair ? pair + op[s] : nullptr, nullptr,
                nullptr, nullptr, jac_init ? jac_init + oj[s] : nullptr, t0[s], nullptr, nullptr, R);
    std::copy(R.state.begin(), R.state.end(), hs.begin() + d.off_state);
    std::copy(R.jac.begin(), R.jac.end(), hj.begin() + d.off_jac);
    std::copy(R.ints.begin(), R.ints.end(), hi.begin() + d.off_int);
    std::copy(R.pair.begin(), R.pair.end(), hp.begin() + d.off_pair);
  }
  NBG_HIP(hipMemcpy(b->S.state0.p, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice));
  NBG_HIP(hipMemcpy(b->S.jac0.p, hj.data(), hj.size() * sizeof(double), hipMemcpyHostToDevice));
  NBG_HIP(hipMemcpy(b->S.pair.p, hp.data(), hp.size(), hipMemcpyHostToDevice));
  *out = b;
  return 0;
}

int nbg_batch_create(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                     const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                     const double* x, const double* v, const double* m, const uint8_t* pair,
                     const double* jac_init, nbg_batch** out) {
  NBG_ARG(ctx != nullptr && out != nullptr, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(!grad || jac_init, "grad batch needs jac_init");
  if (int rc = batch_build(ctx, nsys, n, h, tmax, t0, ti, ntt, grad, ntbv, x, v, m, pair, jac_init, out)) return rc;
  // ints start at zero; the reset also copies the initial state
  return nbg_batch_reset(*out);
}

// Fully nested hierarchy [n, 1, ..., 1] (setup_hierarchy.jl:9-145): row r
// pairs the inner r + 1 bodies (-1) with body r + 1 (+1); the last row is
// the centre of mass of all.
static void nested_hierarchy(int n, double* H) {
  for (int q = 0; q < n * n; ++q) H[q] = 0.0;
  for (int r = 0; r < n - 1; ++r) {
    for (int c = 0; c <= r; ++c) H[r + n * c] = -1.0;
    H[r + n * (r + 1)] = 1.0;
  }
  for (int c = 0; c < n; ++c) H[(n - 1) + n * c] = -1.0;
}

int nbg_batch_create_elements(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                              const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                              const double* elements, const double* hierarchy, const uint8_t* pair,
                              nbg_batch** out) {
  NBG_ARG(ctx != nullptr && out != nullptr, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(nsys > 0 && n && t0 && elements, "missing per-system arrays");
  int64_t nb = 0;
  for (int s = 0; s < nsys; ++s) {
    NBG_ARG(n[s] >= NB_MIN && n[s] <= NB_MAX, "n must be in [2, 12]");
    nb += n[s];
  }
  // masses from the elements' first column; x, v and jac_init come from the GPU
  std::vector<double> m(nb), zero(3 * nb, 0.0);
  std::vector<nbg::IcDesc> ic(nsys);
  std::vector<double> in;
  int64_t ob = 0, oe = 0, oh = 0;
  for (int s = 0; s < nsys; ++s) {
    const int k = n[s];
    for (int q = 0; q < k; ++q) m[ob + q] = elements[oe + q];
    for (int q = 0; q < 7 * k; ++q) {
      const double e = elements[oe + q];
      NBG_ARG(std::isfinite(e), "elements must be finite");
      in.push_back(e);
    }
    const size_t at = in.size();
    in.resize(at + (size_t)k * k);
    if (hierarchy) std::copy(hierarchy + oh, hierarchy + oh + (int64_t)k * k, in.begin() + at);
    else nested_hierarchy(k, in.data() + at);
    {
      // the orbits' element rows (init_nbody.jl:58-75) must exist
      const double* H = in.data() + at;
      int bb = 0;
      for (int i = 0; i < k - 1; ++i) {
        if (H[i] == 0.0) bb += 1;
        NBG_ARG(i + 1 + bb >= 0 && i + 1 + bb < k, "hierarchy does not match the body count");
        if (bb > 0) bb -= 2;
        else if (bb < 0) bb = 0;
      }
    }
    ic[s].n = k;
    ic[s].t0 = t0[s];
    ic[s].off_el = (int64_t)(at - 7 * (size_t)k);
    ob += k;
    oe += 7 * (int64_t)k;
    oh += (int64_t)k * k;
  }
  if (int rc = batch_build(ctx, nsys, n, h, tmax, t0, ti, ntt, grad, ntbv, zero.data(), zero.data(), m.data(), pair,
                           nullptr, out))
    return rc;
  nbg_batch* b = *out;
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = b->L.desc[s];
    ic[s].off_state = d.off_state;
    ic[s].off_jinit = d.off_jac + 2 * 49LL * d.n * d.n;
  }
  DevVec<double> din;
  DevVec<nbg::IcDesc> dds;
  auto fail = [&](hipError_t e) {
    ctx->err = std::string("initial conditions on the GPU failed: ") + hipGetErrorString(e);
    din.release();
    dds.release();
    nbg_batch_destroy(b);
    *out = nullptr;
    return -2;
  };
  hipError_t e = din.ensure(in.size());
  if (e == hipSuccess) e = dds.ensure(ic.size());
  if (e == hipSuccess) e = hipMemcpy(din.p, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dds.p, ic.data(), ic.size() * sizeof(nbg::IcDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = nbg_launch_ic(dds.p, nsys, din.p, b->S.state0.p, b->S.jac0.p, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(e);
  din.release();
  dds.release();
  return nbg_batch_reset(b);
}

int nbg_batch_reset(nbg_batch* b) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  NBG_HIP(hipSetDevice(ctx->device));
  NBG_HIP(hipMemcpy(b->S.state.p, b->S.state0.p, b->L.n_state * sizeof(double), hipMemcpyDeviceToDevice));
  NBG_HIP(hipMemcpy(b->S.jac.p, b->S.jac0.p, b->L.n_jac * sizeof(double), hipMemcpyDeviceToDevice));
  NBG_HIP(hipMemset(b->S.ints.p, 0, b->L.n_int * sizeof(int64_t)));
  NBG_HIP(hipMemset(b->S.tt.p, 0, std::max<int64_t>(b->L.n_tt, 1) * sizeof(double)));
  NBG_HIP(hipMemset(b->S.dq.p, 0, std::max<int64_t>(b->L.n_dq, 1) * sizeof(double)));
  NBG_HIP(hipMemset(b->S.del.p, 0, std::max<int64_t>(b->L.n_dq, 1) * sizeof(double)));
  NBG_HIP(hipDeviceSynchronize());
  return 0;
}

namespace {

// One-wave workgroups the GPU holds at once for a batch's largest launch
// class (as grid_for sizes its grid).
int wave_slots(nbg_ctx* ctx, const Layout& L) {
  int nmax = NB_MIN;
  for (const auto& d : L.desc) nmax = std::max(nmax, (int)d.n);
  int per_cu = (int)(163840 / std::max<size_t>(nbg_lds_bytes(nbg_class_of(nmax)), 1));
  per_cu = std::max(1, std::min(per_cu, 8));
  if (ctx->max_groups_per_cu > 0) per_cu = std::min(per_cu, ctx->max_groups_per_cu);
  return ctx->grid_cap > 0 ? std::min(ctx->grid_cap, ctx->cu_count * per_cu) : ctx->cu_count * per_cu;
}

// Upload a layout's queue order (descriptors, and the units of two-per-wave
// queues) to its store.
int upload_order(nbg_ctx* ctx, Store& S, const Layout& L) {
  std::vector<nbg::SysDesc> sorted(L.desc.size());
  for (size_t k = 0; k < L.order.size(); ++k) sorted[k] = L.desc[L.order[k]];
  NBG_HIP(hipMemcpy(S.desc.p, sorted.data(), sorted.size() * sizeof(nbg::SysDesc), hipMemcpyHostToDevice));
  if (!L.ustart.empty()) {
    NBG_HIP(S.ustart.ensure(L.ustart.size()));
    NBG_HIP(hipMemcpy(S.ustart.p, L.ustart.data(), L.ustart.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return 0;
}

// Order a batch's device queues for visits of `chunk` steps (0: whole runs):
// longest first by the steps one visit takes, min(chunk, nsteps + 1), times
// the per-step cost of the body count; the descriptors are uploaded again.
int batch_order(nbg_batch* b, int64_t chunk) {
  nbg_ctx* ctx = b->ctx;
  const int pairs = ctx->systems_per_wave == 2 ? 1 : 0;
  Layout& L = b->L;
  const int slots = pairs ? wave_slots(ctx, L) : 0;
  if (b->order_chunk == chunk && b->order_pairs == pairs && b->order_slots == slots) return 0;
  L.pairs = pairs;
  L.cost.clear();
  if (chunk > 0) {
    L.cost.resize(L.desc.size());
    for (size_t i = 0; i < L.desc.size(); ++i) L.cost[i] = (double)std::min<int64_t>(chunk, L.desc[i].nsteps + 1);
  }
  L.wave_slots = slots;
  group_layout(L);
  if (int rc = upload_order(ctx, b->S, L)) return rc;
  b->order_chunk = chunk;
  b->order_pairs = pairs;
  b->order_slots = slots;
  return 0;
}

// Per-launch figures of one run_groups call (launch classes, then the
// cooperative launch as class 1).  With per_visit, the totals are the work
// of this call's visits (the launches' unit counters); otherwise they are
// the batch's run totals from its counters.
int fill_stats(nbg_batch* b, double ms, const std::vector<float>& gms, const std::vector<unsigned long long>& gun,
               const std::vector<int>& ggrid, bool per_visit, nbg_batch_stats* stats) {
  nbg_ctx* ctx = b->ctx;
  std::memset(stats, 0, sizeof(*stats));
  const size_t G = b->L.gn.size();
  const bool coop = ggrid.size() > G && ggrid[G] > 0;
  if (per_visit) {
    for (size_t g = 0; g <= G; ++g) {
      if (g == G && !coop) break;
      stats->transits += (int64_t)gun[13 * g + 0];
      stats->newton_iters += (int64_t)gun[13 * g + 1];
      for (int n = 2; n < 13; ++n) stats->system_steps += (int64_t)gun[13 * g + n];
    }
  } else {
    std::vector<int64_t> hi((size_t)b->L.n_int);
    NBG_HIP(hipMemcpy(hi.data(), b->S.ints.p, hi.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t tr = 0, it = 0;
    for (const auto& d : b->L.desc) {
      tr += hi[d.off_int + d.n + 1];
      it += hi[d.off_int + d.n + 2];
    }
    stats->transits = tr;
    stats->newton_iters = it;
    stats->system_steps = b->steps_main + (b->L.desc.empty() || !b->L.desc[0].grad ? 0 : tr);
  }
  stats->kernel_ms = ms;
  if (b->L.pairs)
    for (size_t g = 0; g < G; ++g) {
      const int* us = b->L.ustart.data() + b->L.uoff[g];
      for (int u = 0; us[u] < b->L.gcnt[g]; ++u)
        if (us[u + 1] - us[u] == 2) stats->paired_systems += 2;
    }
  for (size_t g = 0; g < G; ++g) {
    const int c = b->L.gn[g];
    stats->launch_ms[c] = gms[g];
    stats->launch_grid[c] = ggrid[g];
    for (int n = 2; n < 13; ++n) stats->launch_units[c][n] = (int64_t)gun[13 * g + n];
  }
  if (coop) {
    stats->launch_ms[1] = gms[G];
    stats->launch_grid[1] = ggrid[G];
    for (int n = 2; n < 13; ++n) stats->launch_units[1][n] = (int64_t)gun[13 * G + n];
  }
  return 0;
}

// One round of a batch: with restart, systems that have finished start
// again from their initial records; then every system advances up to
// `steps` steps (0: to its end).
int batch_round(nbg_batch* b, int64_t steps, bool restart, nbg_batch_stats* stats) {
  nbg_ctx* ctx = b->ctx;
  NBG_HIP(hipSetDevice(ctx->device));
  if (int rc = batch_order(b, steps)) return rc;
  if (restart)
    NBG_HIP(nbg_launch_rewind(b->S.desc.p, (int)b->L.desc.size(), b->S.state.p, b->S.state0.p, b->S.jac.p,
                              b->S.jac0.p, b->S.ints.p, ctx->cu_count, 0));
  double ms = 0.0;
  std::vector<float> gms;
  std::vector<unsigned long long> gun;
  std::vector<int> ggrid;
  if (int rc = run_groups(ctx, b->S, b->L, steps, &ms, &gms, &gun, nullptr, &ggrid)) return rc;
  nbg_batch_stats tmp;
  return fill_stats(b, ms, gms, gun, ggrid, steps > 0, stats ? stats : &tmp);
}

}  // namespace

int nbg_batch_run(nbg_batch* b, nbg_batch_stats* stats) {
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  return batch_round(b, 0, false, stats);
}

int nbg_batch_round(nbg_batch* b, int64_t steps, int32_t restart, nbg_batch_stats* stats) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(steps > 0, "steps must be positive");
  return batch_round(b, steps, restart != 0, stats);
}

// Autotuner: each candidate launch configuration runs a copy of the batch
// with every step count scaled by `frac`; the fastest one is kept in the
// context.  The copy has the batch's systems, body counts and transit slots,
// so its launch classes, LPT order and two-wave selection scale with it.
namespace {
struct TuneCand {
  int coop_mode, coop_max, groups_per_cu, systems_per_wave;
};
const TuneCand kTune[] = {
    {1, 128, 0, 1},  // two-wave workgroups for the systems beyond the batch's throughput bound
    {0, 0, 0, 1},    // one wave per system throughout
    {2, 16, 0, 1},   // two-wave workgroups for the 16, 32, 64 longest systems
    {2, 32, 0, 1},
    {2, 64, 0, 1},
    {1, 128, 3, 1},  // three one-wave workgroups per CU instead of four
    {1, 128, 0, 2},  // two systems per wave, with and without two-wave workgroups
    {0, 0, 0, 2},
};

struct TuneKeep {
  int coop_mode, coop_max, groups_per_cu, systems_per_wave;
  explicit TuneKeep(const nbg_ctx* c)
      : coop_mode(c->coop_mode), coop_max(c->coop_max), groups_per_cu(c->max_groups_per_cu),
        systems_per_wave(c->systems_per_wave) {}
  void restore(nbg_ctx* c) const {
    c->coop_mode = coop_mode; c->coop_max = coop_max; c->max_groups_per_cu = groups_per_cu;
    c->systems_per_wave = systems_per_wave;
  }
};
void tune_apply(nbg_ctx* c, const TuneCand& t) {
  c->coop_mode = t.coop_mode; c->coop_max = t.coop_max; c->max_groups_per_cu = t.groups_per_cu;
  c->systems_per_wave = t.systems_per_wave;
}
}  // namespace

int nbg_batch_autotune(nbg_batch* b, double frac, int32_t* best, double* cand_ms, int32_t max_cand) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(frac > 0.0 && frac <= 1.0, "frac must be in (0, 1]");
  Layout Ls = b->L;  // offsets depend on n and ntt only: the copy keeps them
  for (auto& d : Ls.desc) d.nsteps = std::max<int64_t>(1, (int64_t)std::llround((double)d.nsteps * frac));
  Ls.cost.clear();
  group_layout(Ls);
  Store T;
  auto done = [&](int rc) {
    T.release();
    return rc;
  };
  if (int rc = store_alloc(ctx, T, Ls, false)) return done(rc);
  const int nc = (int)(sizeof(kTune) / sizeof(kTune[0]));
  const TuneKeep keep(ctx);
  int bi = 0;
  double bms = 0.0;
  for (int c = 0; c < nc; ++c) {
    hipError_t e = hipSuccess;
    if (Ls.pairs != (kTune[c].systems_per_wave == 2 ? 1 : 0)) {  // queue order for this candidate
      Ls.pairs = kTune[c].systems_per_wave == 2 ? 1 : 0;
      Ls.wave_slots = wave_slots(ctx, Ls);
      group_layout(Ls);
      if (int rc = upload_order(ctx, T, Ls)) {
        keep.restore(ctx);
        return done(rc);
      }
    }
    if (e == hipSuccess) e = hipMemcpy(T.state.p, b->S.state0.p, Ls.n_state * sizeof(double), hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(T.jac.p, b->S.jac0.p, Ls.n_jac * sizeof(double), hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(T.pair.p, b->S.pair.p, std::max<int64_t>(Ls.n_pair, 1), hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemset(T.ints.p, 0, Ls.n_int * sizeof(int64_t));
    if (e != hipSuccess) {
      ctx->err = std::string("autotune: ") + hipGetErrorString(e);
      keep.restore(ctx);
      return done(-2);
    }
    tune_apply(ctx, kTune[c]);
    double ms = 0.0;
    const int rc = run_groups(ctx, T, Ls, 0, &ms);
    if (rc) {
      keep.restore(ctx);
      return done(rc);
    }
    if (cand_ms && c < max_cand) cand_ms[c] = ms;
    if (c == 0 || ms < bms) { bms = ms; bi = c; }
  }
  tune_apply(ctx, kTune[bi]);
  if (best) *best = bi;
  return done(0);
}

int nbg_batch_autotune_round(nbg_batch* b, int64_t steps, int32_t* best, double* cand_ms, int32_t max_cand) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(steps > 0, "steps must be positive");
  const int nc = (int)(sizeof(kTune) / sizeof(kTune[0]));
  const TuneKeep keep(ctx);
  int bi = 0;
  double bms = 0.0;
  for (int c = 0; c < nc; ++c) {
    int rc = nbg_batch_reset(b);
    tune_apply(ctx, kTune[c]);
    nbg_batch_stats st;
    if (rc == 0) rc = batch_round(b, steps, false, &st);
    if (rc) {
      keep.restore(ctx);
      return rc;
    }
    if (cand_ms && c < max_cand) cand_ms[c] = st.kernel_ms;
    if (c == 0 || st.kernel_ms < bms) { bms = st.kernel_ms; bi = c; }
  }
  tune_apply(ctx, kTune[bi]);
  if (best) *best = bi;
  return nbg_batch_reset(b);
}

int nbg_batch_autotune_candidate(int32_t index, int32_t* coop_mode, int32_t* coop_max, int32_t* groups_per_cu) {
  const int nc = (int)(sizeof(kTune) / sizeof(kTune[0]));
  if (index < 0 || index >= nc) return nc;
  if (coop_mode) *coop_mode = kTune[index].coop_mode;
  if (coop_max) *coop_max = kTune[index].coop_max;
  if (groups_per_cu) *groups_per_cu = kTune[index].groups_per_cu;
  return nc;
}

int32_t nbg_batch_autotune_systems_per_wave(int32_t index) {
  const int nc = (int)(sizeof(kTune) / sizeof(kTune[0]));
  return (index < 0 || index >= nc) ? -1 : kTune[index].systems_per_wave;
}

int nbg_batch_download(nbg_batch* b, double* tt, double* dtdq0, double* dtdelements, int64_t* count, double* x,
                       double* v) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const Layout& L = b->L;
  // device layout == packed host layout in system order for tt / dq / del
  if (tt) NBG_HIP(hipMemcpy(tt, b->S.tt.p, L.n_tt * sizeof(double), hipMemcpyDeviceToHost));
  if (dtdq0) NBG_HIP(hipMemcpy(dtdq0, b->S.dq.p, L.n_dq * sizeof(double), hipMemcpyDeviceToHost));
  if (dtdelements) NBG_HIP(hipMemcpy(dtdelements, b->S.del.p, L.n_dq * sizeof(double), hipMemcpyDeviceToHost));
  if (count || x || v) {
    std::vector<int64_t> hi((size_t)L.n_int);
    std::vector<double> hs((size_t)L.n_state);
    NBG_HIP(hipMemcpy(hi.data(), b->S.ints.p, hi.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    NBG_HIP(hipMemcpy(hs.data(), b->S.state.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
    int64_t ob = 0;
    for (const auto& d : L.desc) {
      for (int q = 0; q < d.n; ++q) {
        if (count) count[ob + q] = hi[d.off_int + q];
      }
      for (int q = 0; q < 3 * d.n; ++q) {
        if (x) x[3 * ob + q] = hs[d.off_state + nbg::off_x(d.n) + q];
        if (v) v[3 * ob + q] = hs[d.off_state + nbg::off_v(d.n) + q];
      }
      ob += d.n;
    }
  }
  return 0;
}

int nbg_batch_initial(nbg_batch* b, double* x, double* v, double* jac_init) {
  if (!b) return -1;
  nbg_ctx* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const Layout& L = b->L;
  std::vector<double> hs((size_t)L.n_state);
  NBG_HIP(hipMemcpy(hs.data(), b->S.state0.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
  int64_t ob = 0, oj = 0;
  for (const auto& d : L.desc) {
    for (int q = 0; q < 3 * d.n; ++q) {
      if (x) x[3 * ob + q] = hs[d.off_state + nbg::off_x(d.n) + q];
      if (v) v[3 * ob + q] = hs[d.off_state + nbg::off_v(d.n) + q];
    }
    const int64_t nn = 49LL * d.n * d.n;
    if (jac_init)
      NBG_HIP(hipMemcpy(jac_init + oj, b->S.jac0.p + d.off_jac + 2 * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    ob += d.n;
    oj += nn;
  }
  return 0;
}

void nbg_batch_destroy(nbg_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipDeviceSynchronize();
  b->S.release();
  delete b;
}

int nbg_batch_transits(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                       const double* t0, const int32_t* ti, const int64_t* ntt, int32_t grad, int32_t ntbv,
                       const double* x, const double* v, const double* m, const uint8_t* pair,
                       const double* jac_init, double* tt, double* dtdq0, double* dtdelements, int64_t* count,
                       nbg_batch_stats* stats) {
  nbg_batch* b = nullptr;
  int rc = nbg_batch_create(ctx, nsys, n, h, tmax, t0, ti, ntt, grad, ntbv, x, v, m, pair, jac_init, &b);
  if (rc) return rc;
  rc = nbg_batch_run(b, stats);
  if (!rc) rc = nbg_batch_download(b, tt, dtdq0, dtdelements, count, nullptr, nullptr);
  nbg_batch_destroy(b);
  return rc;
}

int nbg_batch_snapshot(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* t0,
                       const int64_t* nt, const double* times, int32_t grad, const double* x, const double* v,
                       const double* m, const uint8_t* pair, const double* jac_init, double* vsky, double* bsky2,
                       double* dbvdq0, double* dbvdelements) {
  NBG_ARG(ctx != nullptr, "null context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(nsys > 0 && n && h && t0 && nt && times && x && v && m && jac_init && vsky && bsky2,
          "missing per-system arrays");
  NBG_ARG(!grad || (dbvdq0 && dbvdelements), "grad run needs dbvdq0 and dbvdelements");
  Layout L;
  L.desc.resize(nsys);
  std::vector<int64_t> ob(nsys + 1, 0), oj(nsys + 1, 0), op(nsys + 1, 0), ot(nsys + 1, 0), oo(nsys + 1, 0),
      oq(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    NBG_ARG(n[s] >= NB_MIN && n[s] <= NB_MAX, "n must be in [2, 12]");
    NBG_ARG(std::isfinite(h[s]) && h[s] != 0.0 && nt[s] >= 1, "h finite and non-zero, nt >= 1");
    ob[s + 1] = ob[s] + n[s];
    oj[s + 1] = oj[s] + 49LL * n[s] * n[s];
    op[s + 1] = op[s] + (int64_t)n[s] * n[s];
    ot[s + 1] = ot[s] + nt[s];
    oo[s + 1] = oo[s] + (int64_t)n[s] * nt[s];
    oq[s + 1] = oq[s] + 14LL * n[s] * n[s] * nt[s];
  }
  for (int64_t q = 0; q < ot[nsys]; ++q) NBG_ARG(std::isfinite(times[q]), "times must be finite");
  for (int s = 0; s < nsys; ++s) {
    nbg::SysDesc& d = L.desc[s];
    d.n = n[s]; d.mode = nbg::MODE_SNAPSHOT; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1;
    d.any_kick = pair ? any_kick(n[s], pair + op[s]) : 0;
    d.nsteps = nt[s]; d.ntt = 0; d.h = h[s]; d.hstep = h[s]; d.hlast = 0.0; d.t0 = t0[s];
  }
  finalize_layout(L);
  // order the queue by the integration each system has ahead (its time span
  // in steps), not by its number of snapshot times
  L.cost.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    double span = 0.0, tc = t0[s];
    for (int64_t j = 0; j < nt[s]; ++j) {
      span += std::fabs(times[ot[s] + j] - tc);
      tc = times[ot[s] + j];
    }
    L.cost[s] = span / std::fabs(h[s]) + 1.0;
  }
  group_layout(L);
  Store S;
  auto done = [&](int rc) {
    S.release();
    return rc;
  };
  if (int rc = store_alloc(ctx, S, L, false)) return done(rc);
  std::vector<double> hs((size_t)L.n_state), hj((size_t)L.n_jac), ht((size_t)L.n_traj, 0.0);
  std::vector<int64_t> hi((size_t)L.n_int, 0);
  std::vector<uint8_t> hp((size_t)L.n_pair, 0);
  HostRec R;
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = L.desc[s];
    pack_system(d, x + 3 * ob[s], v + 3 * ob[s], nullptr, nullptr, m + ob[s], pair ? pair + op[s] : nullptr, nullptr,
                nullptr, nullptr, jac_init + oj[s], t0[s], nullptr, nullptr, R);
    std::copy(R.state.begin(), R.state.end(), hs.begin() + d.off_state);
    std::copy(R.jac.begin(), R.jac.end(), hj.begin() + d.off_jac);
    std::copy(R.pair.begin(), R.pair.end(), hp.begin() + d.off_pair);
    std::copy(times + ot[s], times + ot[s + 1], ht.begin() + d.off_traj);
  }
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    NBG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  if (int rc = up(S.state.p, hs.data(), hs.size() * sizeof(double))) return done(rc);
  if (int rc = up(S.jac.p, hj.data(), hj.size() * sizeof(double))) return done(rc);
  if (int rc = up(S.ints.p, hi.data(), hi.size() * sizeof(int64_t))) return done(rc);
  if (int rc = up(S.pair.p, hp.data(), std::max<size_t>(hp.size(), 1))) return done(rc);
  if (int rc = up(S.traj.p, ht.data(), ht.size() * sizeof(double))) return done(rc);
  if (int rc = run_groups(ctx, S, L, 0, nullptr)) return done(rc);
  hipError_t e = hipMemcpy(ht.data(), S.traj.p, ht.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    ctx->err = std::string("snapshot download: ") + hipGetErrorString(e);
    return done(-2);
  }
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = L.desc[s];
    const int k = d.n;
    const int64_t ns = d.nsteps;
    const double* base = ht.data() + d.off_traj;
    std::copy(base + nbg::snap_off_vsky(k, ns), base + nbg::snap_off_vsky(k, ns) + k * ns, vsky + oo[s]);
    std::copy(base + nbg::snap_off_bsky2(k, ns), base + nbg::snap_off_bsky2(k, ns) + k * ns, bsky2 + oo[s]);
    if (dbvdq0) std::copy(base + nbg::snap_off_dq(k, ns), base + nbg::snap_off_dq(k, ns) + (oq[s + 1] - oq[s]),
                          dbvdq0 + oq[s]);
    if (dbvdelements)
      std::copy(base + nbg::snap_off_del(k, ns), base + nbg::snap_off_del(k, ns) + (oq[s + 1] - oq[s]),
                dbvdelements + oq[s]);
  }
  return done(0);
}

int nbg_batch_cartesian(nbg_ctx* ctx, int32_t nsys, const int32_t* n, const double* h, const double* tmax,
                        const double* t0, const int64_t* nstep, int32_t grad, int32_t records, const double* x,
                        const double* v, const double* m, const uint8_t* pair, double* out_t, double* out_x,
                        double* out_v, double* out_xerror, double* out_verror, double* out_jac, double* out_jac_error) {
  NBG_ARG(ctx != nullptr, "null context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(nsys > 0 && n && h && tmax && t0 && nstep && x && v && m && out_t && out_x && out_v,
          "missing per-system arrays");
  NBG_ARG((records & ~(nbg::TRAJ_ERR | nbg::TRAJ_JAC | nbg::TRAJ_JACERR)) == 0, "unknown record flags");
  NBG_ARG(!(records & nbg::TRAJ_ERR) || (out_xerror && out_verror), "error records need out_xerror and out_verror");
  NBG_ARG(!(records & nbg::TRAJ_JAC) || out_jac, "Jacobian records need out_jac");
  NBG_ARG(!(records & nbg::TRAJ_JACERR) || out_jac_error, "Jacobian error records need out_jac_error");
  NBG_ARG(grad || !(records & (nbg::TRAJ_JAC | nbg::TRAJ_JACERR)), "Jacobian records need grad != 0");
  Layout L;
  L.desc.resize(nsys);
  std::vector<int64_t> ob(nsys + 1, 0), op(nsys + 1, 0), o1(nsys + 1, 0), o3(nsys + 1, 0), oj(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    NBG_ARG(n[s] >= NB_MIN && n[s] <= NB_MAX, "n must be in [2, 12]");
    NBG_ARG(std::isfinite(h[s]) && h[s] != 0.0 && nstep[s] >= 1, "h finite and non-zero, nstep >= 1");
    ob[s + 1] = ob[s] + n[s];
    op[s + 1] = op[s] + (int64_t)n[s] * n[s];
    o1[s + 1] = o1[s] + nstep[s];
    o3[s + 1] = o3[s] + 3LL * n[s] * nstep[s];
    oj[s + 1] = oj[s] + 49LL * n[s] * n[s] * nstep[s];
  }
  for (int s = 0; s < nsys; ++s) {
    nbg::SysDesc& d = L.desc[s];
    d.n = n[s]; d.mode = nbg::MODE_CARTESIAN; d.grad = grad ? 1 : 0; d.ti = 0; d.ntbv = 1;
    d.any_kick = pair ? any_kick(n[s], pair + op[s]) : 0;
    d.traj = records;
    d.nsteps = nstep[s]; d.ntt = 0; d.h = h[s] * check_step(t0[s], tmax[s]); d.hstep = h[s]; d.hlast = 0.0;
    d.t0 = t0[s];
  }
  finalize_layout(L);
  Store S;
  auto done = [&](int rc) {
    S.release();
    return rc;
  };
  if (int rc = store_alloc(ctx, S, L, false)) return done(rc);
  std::vector<double> hs((size_t)L.n_state), hj((size_t)L.n_jac);
  std::vector<int64_t> hi((size_t)L.n_int, 0);
  std::vector<uint8_t> hp((size_t)L.n_pair, 0);
  HostRec R;
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = L.desc[s];
    pack_system(d, x + 3 * ob[s], v + 3 * ob[s], nullptr, nullptr, m + ob[s], pair ? pair + op[s] : nullptr, nullptr,
                nullptr, nullptr, nullptr, t0[s], nullptr, nullptr, R);
    std::copy(R.state.begin(), R.state.end(), hs.begin() + d.off_state);
    std::copy(R.jac.begin(), R.jac.end(), hj.begin() + d.off_jac);
    std::copy(R.pair.begin(), R.pair.end(), hp.begin() + d.off_pair);
  }
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    NBG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  if (int rc = up(S.state.p, hs.data(), hs.size() * sizeof(double))) return done(rc);
  if (int rc = up(S.jac.p, hj.data(), hj.size() * sizeof(double))) return done(rc);
  if (int rc = up(S.ints.p, hi.data(), hi.size() * sizeof(int64_t))) return done(rc);
  if (int rc = up(S.pair.p, hp.data(), std::max<size_t>(hp.size(), 1))) return done(rc);
  if (int rc = run_groups(ctx, S, L, 0, nullptr)) return done(rc);
  for (int s = 0; s < nsys; ++s) {
    const nbg::SysDesc& d = L.desc[s];
    const int k = d.n;
    const int64_t ns = d.nsteps;
    const double* base = S.traj.p + d.off_traj;
    auto get = [&](double* dst, int64_t off, int64_t len) -> int {
      NBG_HIP(hipMemcpy(dst, base + off, (size_t)len * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
    };
    const int64_t n3 = 3LL * k * ns, nj = 49LL * k * k * ns;
    int rc = get(out_t + o1[s], 0, ns);
    if (!rc) rc = get(out_x + o3[s], nbg::traj_off_x(k, ns), n3);
    if (!rc) rc = get(out_v + o3[s], nbg::traj_off_v(k, ns), n3);
    if (!rc && (records & nbg::TRAJ_ERR)) rc = get(out_xerror + o3[s], nbg::traj_off_xe(k, ns), n3);
    if (!rc && (records & nbg::TRAJ_ERR)) rc = get(out_verror + o3[s], nbg::traj_off_ve(k, ns), n3);
    if (!rc && (records & nbg::TRAJ_JAC)) rc = get(out_jac + oj[s], nbg::traj_off_jac(k, ns, records), nj);
    if (!rc && (records & nbg::TRAJ_JACERR)) rc = get(out_jac_error + oj[s], nbg::traj_off_jacerr(k, ns, records), nj);
    if (rc) return done(rc);
  }
  return done(0);
}

// ---------------------------------------------------------------------------
// Serving: requests (one system each) admitted into a device arena as their
// HBM footprint allows, advanced `chunk` steps per round by the persistent
// class kernels together with every other active request, retired when done.
// ---------------------------------------------------------------------------
namespace {

// First-fit allocator of 256-byte-aligned ranges of a device buffer.
struct Arena {
  int64_t cap = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
  static int64_t round(int64_t len) { return (len + 255) & ~(int64_t)255; }
  void init(int64_t c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int64_t alloc(int64_t len) {
    len = round(len);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      const int64_t off = it->first, rem = it->second - len;
      free_.erase(it);
      if (rem > 0) free_[off + len] = rem;
      return off;
    }
    return -1;
  }
  void release(int64_t off, int64_t len) {
    len = round(len);
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t free_bytes() const {
    int64_t f = 0;
    for (const auto& kv : free_) f += kv.second;
    return f;
  }
};

// Byte offsets of a request's record inside its arena block.
struct ReqLayout {
  int64_t state, jac, ws, tt, dq, del, ints, pair, bytes;
};

ReqLayout req_layout(int n, int ntbv, int64_t ntt) {
  ReqLayout R;
  int64_t o = 0;
  R.state = o; o += 8 * nbg::state_len(n);
  R.jac = o; o += 8 * nbg::jac_len(n);
  R.ws = o; o += 8 * nbg::ws_len(n);
  R.tt = o; o += 8 * (int64_t)ntbv * n * ntt;
  R.dq = o; o += 8 * (int64_t)ntbv * n * ntt * 7 * n;
  R.del = o; o += 8 * (int64_t)ntbv * n * ntt * 7 * n;
  R.ints = o; o += 8 * nbg::int_len(n);
  R.pair = o; o += ((int64_t)n * n + 7) / 8 * 8;
  R.bytes = o;
  return R;
}

struct ServeReq {
  int64_t id = 0;
  nbg::SysDesc d{};
  ReqLayout lay{};
  int64_t off = -1;       // arena block (bytes)
  int64_t steps_run = 0;  // steps advanced so far (chunk per round)
  HostRec init;           // initial record, until admitted
  std::vector<double> tt, dq, del, x, v;
  std::vector<int64_t> count;
};

}  // namespace

struct nbg_server {
  nbg_ctx* ctx = nullptr;
  int64_t chunk = 0;
  DevVec<unsigned char> arena_mem;
  Arena arena;
  DevVec<int> done;  // [0]: count, [1..]: tags
  Store S;           // descriptors, queues and unit counters of a round
  std::mutex mu;     // pending, finished, next_id
  std::deque<ServeReq*> pending;
  std::vector<ServeReq*> active;
  std::map<int64_t, ServeReq*> finished;
  int64_t next_id = 1;
  int64_t rounds = 0;
  int64_t units = 0;       // system-steps (main steps + one per recorded transit) over every round
  double device_ms = 0.0;  // device time over every round
};

int nbg_server_create(nbg_ctx* ctx, int64_t max_bytes, int64_t chunk_steps, nbg_server** out) {
  NBG_ARG(ctx && out && max_bytes >= 4096 && chunk_steps > 0, "nbg_server_create: bad arguments");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_HIP(hipSetDevice(ctx->device));
  auto* sv = new nbg_server();
  sv->ctx = ctx;
  sv->chunk = chunk_steps;
  const int64_t cap = Arena::round(max_bytes);
  if (sv->arena_mem.ensure((size_t)cap) != hipSuccess) {
    delete sv;
    ctx->err = "nbg_server_create: cannot allocate the device arena";
    return -2;
  }
  sv->arena.init(cap);
  *out = sv;
  return 0;
}

void nbg_server_destroy(nbg_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->ctx->mu);
    (void)hipSetDevice(sv->ctx->device);
    sv->arena_mem.release();
    sv->done.release();
    sv->S.release();
  }
  for (auto* r : sv->pending) delete r;
  for (auto* r : sv->active) delete r;
  for (auto& kv : sv->finished) delete kv.second;
  delete sv;
}

int nbg_server_submit(nbg_server* sv, int32_t n, double h, double tmax, double t0, int32_t ti, int64_t ntt,
                      int32_t grad, int32_t ntbv, const double* x, const double* v, const double* m,
                      const uint8_t* pair, const double* jac_init, int64_t* id) {
  if (!sv) return -1;
  nbg_ctx* ctx = sv->ctx;
  NBG_ARG(n >= NB_MIN && n <= NB_MAX, "n must be in [2, 12]");
  NBG_ARG(std::isfinite(h) && h != 0.0 && std::isfinite(tmax) && std::isfinite(t0) && ntt >= 0,
          "h finite non-zero, tmax and t0 finite, ntt >= 0");
  NBG_ARG(ti >= 0 && ti < n, "ti out of range");
  NBG_ARG(ntbv == 1 || ntbv == 3, "ntbv must be 1 or 3");
  NBG_ARG(x && v && m && id && (!grad || jac_init), "missing arrays");
  auto* r = new ServeReq();
  nbg::SysDesc& d = r->d;
  d.n = n; d.mode = nbg::MODE_TRANSITS; d.grad = grad ? 1 : 0; d.ti = ti; d.ntbv = ntbv;
  d.any_kick = pair ? any_kick(n, pair) : 0;
  d.nsteps = std::llabs(jround(tmax / h));
  d.ntt = ntt; d.h = h * check_step(t0, tmax + t0); d.hstep = h; d.hlast = 0.0; d.t0 = t0;
  r->lay = req_layout(n, ntbv, ntt);
  pack_system(d, x, v, nullptr, nullptr, m, pair, nullptr, nullptr, nullptr, jac_init, t0, nullptr, nullptr, r->init);
  if (Arena::round(r->lay.bytes) > sv->arena.cap) {
    delete r;
    ctx->err = "request footprint exceeds the server's arena";
    return -1;
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  r->id = sv->next_id++;
  sv->pending.push_back(r);
  *id = r->id;
  return 0;
}

int nbg_server_step(nbg_server* sv, int32_t* n_finished, double* kernel_ms) {
  if (!sv) return -1;
  nbg_ctx* ctx = sv->ctx;
  std::lock_guard<std::mutex> glk(ctx->mu);
  NBG_HIP(hipSetDevice(ctx->device));
  unsigned char* base = sv->arena_mem.p;
  // admission, first come first served, while the arena has room
  std::vector<ServeReq*> admitted;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    while (!sv->pending.empty()) {
      ServeReq* r = sv->pending.front();
      const int64_t off = sv->arena.alloc(r->lay.bytes);
      if (off < 0) break;
      sv->pending.pop_front();
      r->off = off;
      admitted.push_back(r);
    }
  }
  for (ServeReq* r : admitted) {
    nbg::SysDesc& d = r->d;
    const ReqLayout& Ly = r->lay;
    d.off_state = (r->off + Ly.state) / 8;
    d.off_jac = (r->off + Ly.jac) / 8;
    d.off_ws = (r->off + Ly.ws) / 8;
    d.off_tt = (r->off + Ly.tt) / 8;
    d.off_dq = (r->off + Ly.dq) / 8;
    d.off_del = (r->off + Ly.del) / 8;
    d.off_int = (r->off + Ly.ints) / 8;
    d.off_pair = r->off + Ly.pair;
    NBG_HIP(hipMemcpy(base + r->off + Ly.state, r->init.state.data(), 8 * r->init.state.size(), hipMemcpyHostToDevice));
    NBG_HIP(hipMemcpy(base + r->off + Ly.jac, r->init.jac.data(), 8 * r->init.jac.size(), hipMemcpyHostToDevice));
    NBG_HIP(hipMemset(base + r->off + Ly.tt, 0, (size_t)(Ly.pair - Ly.tt)));  // tt, dq, del, ints
    NBG_HIP(hipMemcpy(base + r->off + Ly.pair, r->init.pair.data(), r->init.pair.size(), hipMemcpyHostToDevice));
    r->init = HostRec();
    sv->active.push_back(r);
  }
  if (n_finished) *n_finished = 0;
  if (kernel_ms) *kernel_ms = 0.0;
  if (sv->active.empty()) return 0;
  // one round over every active request.  Rounds are balanced in time
  // rather than in steps: the costliest body count among the active
  // requests advances `chunk` steps, and a request whose steps cost less
  // (per the per-step cost model) proportionally more, so that no workgroup
  // idles through most of a round waiting for the costliest ones.
  Layout L;
  L.desc.resize(sv->active.size());
  L.cost.resize(sv->active.size());
  int nmax = 2;
  for (ServeReq* r : sv->active) nmax = std::max(nmax, r->d.n);
  const double cmax = step_cost(nmax);
  for (size_t k = 0; k < sv->active.size(); ++k) {
    ServeReq* r = sv->active[k];
    L.desc[k] = r->d;
    L.desc[k].tag = (int32_t)k;
    L.desc[k].chunk = std::max<int64_t>(1, std::llround((double)sv->chunk * cmax / step_cost(r->d.n)));
    L.cost[k] = (double)std::max<int64_t>(1, std::min<int64_t>(L.desc[k].chunk, r->d.nsteps + 1 - r->steps_run));
  }
  // one request per wave: pairing requests (nbg_set_systems_per_wave) was
  // measured to add 1 % to the closed-loop capacity, for longer requests
  group_layout(L);
  const size_t na = sv->active.size();
  NBG_HIP(sv->S.desc.ensure(na));
  NBG_HIP(sv->S.queue.ensure(2 * L.gn.size() + 1));
  NBG_HIP(sv->S.units.ensure(13 * L.gn.size() + 14));
  NBG_HIP(sv->done.ensure(na + 1));
  if (int rc = upload_order(ctx, sv->S, L)) return rc;
  NBG_HIP(hipMemset(sv->done.p, 0, sizeof(int)));
  nbg::SysBufs B;
  B.state = B.jac = B.ws = B.tt = B.dq = B.del = reinterpret_cast<double*>(base);
  B.ints = reinterpret_cast<int64_t*>(base);
  B.pair = base;
  B.traj = nullptr;
  B.done_count = sv->done.p;
  B.done_list = sv->done.p + 1;
  double ms = 0.0;
  std::vector<unsigned long long> gun;
  if (int rc = run_groups(ctx, sv->S, L, sv->chunk, &ms, nullptr, &gun, &B)) return rc;
  sv->rounds += 1;
  for (size_t g = 0; g < gun.size() / 13; ++g)
    for (int n = 2; n < 13; ++n) sv->units += (int64_t)gun[13 * g + n];
  sv->device_ms += ms;
  if (kernel_ms) *kernel_ms = ms;
  for (size_t k = 0; k < na; ++k) {
    ServeReq* r = sv->active[k];
    r->steps_run = std::min<int64_t>(r->d.nsteps + 1, r->steps_run + L.desc[k].chunk);
  }
  int nd = 0;
  NBG_HIP(hipMemcpy(&nd, sv->done.p, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> tags((size_t)std::max(nd, 0));
  if (nd > 0) NBG_HIP(hipMemcpy(tags.data(), sv->done.p + 1, sizeof(int) * nd, hipMemcpyDeviceToHost));
  std::vector<char> gone(na, 0);
  for (int tg : tags) {
    if (tg < 0 || (size_t)tg >= na || gone[tg]) continue;
    ServeReq* r = sv->active[tg];
    const nbg::SysDesc& d = r->d;
    const ReqLayout& Ly = r->lay;
    const int n = d.n;
    const size_t ntv = (size_t)d.ntbv * n * d.ntt, ndq = ntv * 7 * n;
    r->tt.resize(ntv);
    r->dq.resize(d.grad ? ndq : 0);
    r->del.resize(d.grad ? ndq : 0);
    r->count.resize(n);
    std::vector<double> st((size_t)nbg::state_len(n));
    if (ntv) NBG_HIP(hipMemcpy(r->tt.data(), base + r->off + Ly.tt, 8 * ntv, hipMemcpyDeviceToHost));
    if (d.grad && ndq) {
      NBG_HIP(hipMemcpy(r->dq.data(), base + r->off + Ly.dq, 8 * ndq, hipMemcpyDeviceToHost));
      NBG_HIP(hipMemcpy(r->del.data(), base + r->off + Ly.del, 8 * ndq, hipMemcpyDeviceToHost));
    }
    NBG_HIP(hipMemcpy(r->count.data(), base + r->off + Ly.ints, 8 * (size_t)n, hipMemcpyDeviceToHost));
    NBG_HIP(hipMemcpy(st.data(), base + r->off + Ly.state, 8 * st.size(), hipMemcpyDeviceToHost));
    r->x.assign(st.begin() + nbg::off_x(n), st.begin() + nbg::off_x(n) + 3 * n);
    r->v.assign(st.begin() + nbg::off_v(n), st.begin() + nbg::off_v(n) + 3 * n);
    gone[tg] = 1;
  }
  std::vector<ServeReq*> keep;
  int fin = 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (size_t k = 0; k < na; ++k) {
      ServeReq* r = sv->active[k];
      if (gone[k]) {
        sv->arena.release(r->off, r->lay.bytes);
        sv->finished[r->id] = r;
        ++fin;
      } else {
        keep.push_back(r);
      }
    }
  }
  sv->active.swap(keep);
  if (n_finished) *n_finished = fin;
  return 0;
}

int nbg_server_finished(nbg_server* sv, int64_t* ids, int32_t max_ids, int32_t* n) {
  if (!sv || !n || (max_ids > 0 && !ids)) return -1;
  std::lock_guard<std::mutex> lk(sv->mu);
  int32_t k = 0;
  for (const auto& kv : sv->finished) {
    if (k >= max_ids) break;
    ids[k++] = kv.first;
  }
  *n = k;
  return 0;
}

int nbg_server_fetch(nbg_server* sv, int64_t id, double* tt, double* dtdq0, double* dtdelements, int64_t* count,
                     double* x, double* v) {
  if (!sv) return -1;
  nbg_ctx* ctx = sv->ctx;
  ServeReq* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->finished.find(id);
    if (it != sv->finished.end()) {
      r = it->second;
      sv->finished.erase(it);
    }
  }
  NBG_ARG(r != nullptr, "nbg_server_fetch: no finished request with this id");
  if (tt) std::copy(r->tt.begin(), r->tt.end(), tt);
  if (dtdq0) std::copy(r->dq.begin(), r->dq.end(), dtdq0);
  if (dtdelements) std::copy(r->del.begin(), r->del.end(), dtdelements);
  if (count) std::copy(r->count.begin(), r->count.end(), count);
  if (x) std::copy(r->x.begin(), r->x.end(), x);
  if (v) std::copy(r->v.begin(), r->v.end(), v);
  delete r;
  return 0;
}

int nbg_server_status(nbg_server* sv, int64_t* pending, int64_t* active, int64_t* finished, int64_t* bytes_free) {
  if (!sv) return -1;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (pending) *pending = (int64_t)sv->pending.size();
  if (active) *active = (int64_t)sv->active.size();
  if (finished) *finished = (int64_t)sv->finished.size();
  if (bytes_free) *bytes_free = sv->arena.free_bytes();
  return 0;
}

int nbg_server_work(nbg_server* sv, int64_t* system_steps, int64_t* rounds, double* device_ms) {
  if (!sv) return -1;
  std::lock_guard<std::mutex> lk(sv->ctx->mu);  // steps update these under the context's lock
  if (system_steps) *system_steps = sv->units;
  if (rounds) *rounds = sv->rounds;
  if (device_ms) *device_ms = sv->device_ms;
  return 0;
}

int64_t nbg_server_footprint(int32_t n, int32_t ntbv, int64_t ntt) { return Arena::round(req_layout(n, ntbv, ntt).bytes); }

int nbg_copy_calibrate(nbg_ctx* ctx, int64_t bytes, int32_t reps, double* ms) {
  NBG_ARG(ctx && bytes > 0 && bytes % 16 == 0 && reps > 0 && ms, "nbg_copy_calibrate: bad arguments");
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_HIP(hipSetDevice(ctx->device));
  void *src = nullptr, *dst = nullptr;
  NBG_HIP(hipMalloc(&src, (size_t)bytes));
  if (hipMalloc(&dst, (size_t)bytes) != hipSuccess) {
    (void)hipFree(src);
    ctx->err = "nbg_copy_calibrate: out of device memory";
    return -2;
  }
  hipError_t e = hipMemset(src, 0, (size_t)bytes);
  if (e == hipSuccess) e = nbg_launch_copy(src, dst, (size_t)bytes, ctx->cu_count, 0);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(ctx->ev0, 0);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = nbg_launch_copy(src, dst, (size_t)bytes, ctx->cu_count, 0);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) {
    ctx->err = std::string("nbg_copy_calibrate: ") + hipGetErrorString(e);
    return -2;
  }
  *ms = t / reps;
  return 0;
}

int nbg_probe_row_update(nbg_ctx* ctx, int32_t mode, int32_t reps, int32_t grid, double* ms, double* j_out) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(mode == 0 || mode == 1, "mode must be 0 (VALU) or 1 (MFMA)");
  NBG_ARG(reps > 0 && grid > 0 && ms, "reps and grid must be positive, ms non-null");
  NBG_HIP(hipSetDevice(ctx->device));
  DevVec<double> out;
  NBG_HIP(out.ensure(12 * 63));
  hipStream_t st = ctx->streams[0];
  NBG_HIP(hipEventRecord(ctx->ev0, st));
  hipError_t e = nbg_launch_probe(mode, reps, grid, out.p, st);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, st);
  if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  if (e == hipSuccess && j_out) e = hipMemcpy(j_out, out.p, 12 * 63 * sizeof(double), hipMemcpyDeviceToHost);
  out.release();
  if (e != hipSuccess) {
    ctx->err = std::string("row-update probe: ") + hipGetErrorString(e);
    return -2;
  }
  *ms = t;
  return 0;
}

int nbg_libm_eval_device(nbg_ctx* ctx, int32_t fn, int64_t count, const double* x, double* out) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  NBG_ARG(fn >= 0 && fn <= 5, "fn must be 0 (sin) .. 5 (cbrt)");
  NBG_ARG(count >= 0 && (count == 0 || (x && out)), "count must be >= 0 and x, out non-null");
  if (count == 0) return 0;
  NBG_HIP(hipSetDevice(ctx->device));
  DevVec<double> dx, dout;
  hipStream_t st = ctx->streams[0];
  hipError_t e = dx.ensure(count);
  if (e == hipSuccess) e = dout.ensure(count);
  if (e == hipSuccess) e = hipMemcpyAsync(dx.p, x, count * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = nbg_launch_libm(fn, count, dx.p, dout.p, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout.p, count * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dx.release();
  dout.release();
  if (e != hipSuccess) {
    ctx->err = std::string("nbg_libm_eval_device: ") + hipGetErrorString(e);
    return -2;
  }
  return 0;
}

int nbg_selftest_pair_host(const double* x0, const double* v0, double k, double h, int32_t drift_first,
                           double* out) {
  if (!x0 || !v0 || !out) return -1;
  double delxv[6];
  const nbg::KParams P = nbg::jac_delxv_gamma(x0, v0, delxv, k, h, drift_first != 0);
  double kep[48], jm[6];
  for (int q = 0; q < 48; ++q) kep[q] = 0.0;
  for (int q = 0; q < 6; ++q) jm[q] = 0.0;
  nbg::compute_jacobian_gamma(P, x0, v0, kep, jm, drift_first != 0);
  for (int q = 0; q < 6; ++q) out[q] = delxv[q];
  for (int q = 0; q < 48; ++q) out[6 + q] = kep[q];
  for (int q = 0; q < 6; ++q) out[54 + q] = jm[q];
  out[60] = (double)P.iters;
  return 0;
}

}  // extern "C"
