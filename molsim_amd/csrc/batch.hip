// Batched ensemble: many resident systems (contexts) evaluated together.
//
// UTotal: every stage of the pass (cell binning, scan, scatter, finalize, pair loop, spread,
// the five FFT stages, gather, reduction) is ONE segmented launch covering the workgroups of
// all systems (one launch per template variant: precision, type-pair count, pair tile,
// spline order, mesh size), so a batch step costs about as many launches as one system's
// pass, independent of the number of systems.  The stage sequence is captured once into a
// hipGraph: the pair stage runs on a second stream beside the mesh stages, the gather waits
// for it, and a last kernel collects every system's energies into one buffer.
//
// DUTotal: the trials of all systems run as one segmented launch per spline order, with one
// upload of the trial inputs and one download of the results.
#include <cstdlib>
#include <vector>

#include "common.h"

using namespace msg;

namespace {

__global__ void k_collect(const double* const* __restrict__ outp, double* __restrict__ outs) {
  outs[64 * blockIdx.x + threadIdx.x] = outp[blockIdx.x][threadIdx.x];
}

constexpr int kDeStride = kMaxTP + 4;

}  // namespace

namespace msg {

// Conditions for a system to join a batched pass.
int check_batchable(Ctx* c) {
  if (!c->have_table || !c->have_particles || !c->have_positions ||
      (c->P.ewald_mode != MSG_EWALD_NONE && !c->have_ewald))
    return set_error(MSG_ERR_STATE, "batched pass: a system is not fully set up");
  if (c->P.list_mode == MSG_LIST_VERLET)
    return set_error(MSG_ERR_ARG, "batched pass: Verlet-list systems run unbatched");
  if (c->P.ewald_mode == MSG_EWALD_STD)
    return set_error(MSG_ERR_ARG, "batched pass: standard-Ewald systems run unbatched");
  if (c->prof) return set_error(MSG_ERR_STATE, "batched pass: profiling is on");
  return MSG_OK;
}

int build_pass(const std::vector<Ctx*>& cs, SegOwner& own, Pass& p) {
  p = Pass();
  p.cs = cs;
  for (Ctx* c : cs) {
    if (int e = check_batchable(c)) return e;
    p.any_spme |= c->P.ewald_mode == MSG_EWALD_SPM;
  }
  if (cs.empty()) return MSG_OK;
  if (int e = plan_bin(cs, own, p.bin)) return e;
  if (int e = plan_sort(cs, own, p.scan_c, p.scan_b, p.scatter_c, p.scatter_b, p.finalize)) return e;
  if (int e = plan_pair(cs, own, p.pair)) return e;
  if (int e = plan_pair_sum(cs, own, p.pair_sum)) return e;
  if (int e = plan_spme(cs, own, p.spread, p.fft, p.gather)) return e;
  if (int e = plan_reduce(cs, own, p.reduce)) return e;
  std::vector<const double*> outp;
  for (Ctx* c : cs) outp.push_back(c->d_out);
  if (int e = own.upload(outp, (const double* const**)&p.d_outp)) return e;
  const double* d = nullptr;
  if (int e = own.upload(std::vector<double>(64 * cs.size(), 0.0), &d)) return e;
  p.d_outs = (double*)d;
  return MSG_OK;
}

static int run_stage(const SegStage& st, hipStream_t s) {
  for (const SegGroup& g : st) g.launch(g, s);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// The batched UTotal sequence on stream s (capturable).  With SPME the work forks after the
// binning pass: the cell list, pair kernel and pair-row sum run on `side` beside the brick
// scan/scatter and the mesh stages; the gather waits for them.
// With MOLSIM_GPU_SERIAL_PASS=1 in the environment every stage runs on one stream, so a
// kernel profiler sees each stage's own duration (a measurement aid; results are identical).
static bool serial_pass() {
  static const bool v = [] {
    const char* e = std::getenv("MOLSIM_GPU_SERIAL_PASS");
    return e && e[0] == '1';
  }();
  return v;
}

int launch_pass(const Pass& p, hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join) {
  if (p.cs.empty()) return MSG_OK;
  if (int e = run_stage(p.bin, s)) return e;
  const bool branch = p.any_spme && !serial_pass();
  auto mesh = [&]() -> int {
    if (int e = run_stage(p.scan_b, s)) return e;
    if (int e = run_stage(p.scatter_b, s)) return e;
    if (!p.any_spme) return MSG_OK;
    if (int e = run_stage(p.spread, s)) return e;
    for (int k = 0; k < 5; ++k)
      if (int e = run_stage(p.fft[k], s)) return e;
    return MSG_OK;
  };
  auto cells_pair = [&](hipStream_t cs) -> int {
    if (int e = run_stage(p.scan_c, cs)) return e;
    if (int e = run_stage(p.scatter_c, cs)) return e;
    if (int e = run_stage(p.finalize, cs)) return e;
    if (int e = run_stage(p.pair, cs)) return e;
    return run_stage(p.pair_sum, cs);
  };
  if (branch) {
    // the mesh branch is the longer chain: recorded first, it is dispatched first
    MSG_HIP(hipEventRecord(fork, s));
    if (int e = mesh()) return e;
    MSG_HIP(hipStreamWaitEvent(side, fork, 0));
    if (int e = cells_pair(side)) return e;
    MSG_HIP(hipEventRecord(join, side));
    MSG_HIP(hipStreamWaitEvent(s, join, 0));
  } else {
    if (int e = cells_pair(s)) return e;
    if (int e = mesh()) return e;
  }
  if (p.any_spme)
    if (int e = run_stage(p.gather, s)) return e;
  if (int e = run_stage(p.reduce, s)) return e;
  hipLaunchKernelGGL(k_collect, dim3((unsigned)p.cs.size()), dim3(64), 0, s, p.d_outp, p.d_outs);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// Orders each moved system's own stream work (commits, new positions) before stream s.
static int order_after_contexts(const std::vector<Ctx*>& cs, hipStream_t s) {
  for (Ctx* c : cs)
    if (!c->fresh) {
      if (!c->ev_batch) MSG_HIP(hipEventCreateWithFlags(&c->ev_batch, hipEventDisableTiming));
      MSG_HIP(hipEventRecord(c->ev_batch, c->stream));
      MSG_HIP(hipStreamWaitEvent(s, c->ev_batch, 0));
    }
  return MSG_OK;
}

// Energies of system s from a collected 64-double block; error flags raise.
static int read_out(const double* o, const Ctx* c, int s, double* u_tot, double* u_rec, double* virial) {
  const int ntp = c->nptpt;
  if (o[ntp + 7] != 0.0)
    return set_error(MSG_ERR_R2UMIN, "system " + std::to_string(s) + " raised error flags " +
                                         std::to_string((int)o[ntp + 7]));
  if (u_rec) u_rec[s] = o[ntp + 1];
  if (u_tot) u_tot[s] = o[ntp + 2];
  if (virial) virial[s] = o[ntp + 3];
  return MSG_OK;
}

// Trial inputs of the systems cs (ntrial[s], ip 1-based, rotm; concatenated in cs order)
// packed group by group; each group's trial-start table and the inputs go through `own`.
struct DeLaunch {
  std::vector<const int*> d_tstart;
  std::vector<long> gbase;           // first packed trial of each group
  std::vector<long> pos;             // packed position of each system's first trial
  const int* d_ip = nullptr;
  const double* d_r = nullptr;
  long ntot = 0;
};

static long count_trials(const int* ntrial, int n) {
  long t = 0;
  for (int s = 0; s < n; ++s) t += std::max(ntrial[s], 0);
  return t;
}

// Arena bytes pack_trials needs.
static size_t trial_bytes(long ntot, int n, size_t ngroups) {
  return (size_t)ntot * 28 + (size_t)(n + 2) * 4 * ngroups + 256 * (ngroups + 3);
}

static int pack_trials(const std::vector<DeGroup>& de, const std::vector<Ctx*>& cs, const int* ntrial,
                       const int* ip, const double* rotm, SegOwner& own, DeLaunch& L) {
  const int n = (int)cs.size();
  std::vector<long> off(n + 1, 0);
  for (int s = 0; s < n; ++s) {
    if (ntrial[s] < 0) return set_error(MSG_ERR_ARG, "DUTotal batch: negative trial count");
    off[s + 1] = off[s] + ntrial[s];
  }
  L.ntot = off[n];
  if (L.ntot > (1L << 30)) return set_error(MSG_ERR_ARG, "DUTotal batch: too many trials");
  for (int s = 0; s < n; ++s)
    for (long t = off[s]; t < off[s + 1]; ++t)
      if (ip[t] < 1 || ip[t] > cs[s]->np)
        return set_error(MSG_ERR_ARG, "DUTotal batch: particle index out of range in system " + std::to_string(s));
  std::vector<int> tip(L.ntot);
  std::vector<double> tr(3 * L.ntot);
  L.pos.assign(n, 0);
  L.gbase.assign(de.size() + 1, 0);
  L.d_tstart.assign(de.size(), nullptr);
  long k = 0;
  for (size_t gi = 0; gi < de.size(); ++gi) {
    L.gbase[gi] = k;
    std::vector<int> ts(de[gi].sys.size() + 1, 0);
    for (size_t m = 0; m < de[gi].sys.size(); ++m) {
      const int s = de[gi].sys[m];
      ts[m] = (int)(k - L.gbase[gi]);
      L.pos[s] = k;
      for (long t = off[s]; t < off[s + 1]; ++t, ++k) {
        tip[k] = ip[t] - 1;
        for (int d = 0; d < 3; ++d) tr[3 * k + d] = rotm[3 * t + d];
      }
    }
    ts[de[gi].sys.size()] = (int)(k - L.gbase[gi]);
    if (int e = own.upload(ts, &L.d_tstart[gi])) return e;
  }
  L.gbase[de.size()] = k;
  if (int e = own.upload(tip, &L.d_ip)) return e;
  return own.upload(tr, &L.d_r);
}

static int launch_trials(const std::vector<DeGroup>& de, const DeLaunch& L, double* d_out, hipStream_t s) {
  for (size_t gi = 0; gi < de.size(); ++gi) {
    const long g0 = L.gbase[gi], gn = L.gbase[gi + 1] - g0;
    if (int e = launch_de_seg(de[gi], L.d_tstart[gi], L.d_ip + g0, L.d_r + 3 * g0, d_out + g0 * kDeStride,
                              (int)gn, s))
      return e;
  }
  return MSG_OK;
}

static void unpack_trials(const std::vector<Ctx*>& cs, const int* ntrial, const DeLaunch& L,
                          const double* h_de, int* ovl, double* du_tot) {
  long t = 0;
  for (size_t s = 0; s < cs.size(); ++s) {
    const int ntp = cs[s]->nptpt;
    for (int k = 0; k < ntrial[s]; ++k, ++t) {
      const double* o = h_de + (L.pos[s] + k) * kDeStride;
      ovl[t] = o[ntp + 3] != 0.0;
      du_tot[t] = o[ntp + 2];
    }
  }
}

}  // namespace msg

struct msg_batch {
  std::vector<msg_ctx*> ctx;
  std::vector<Ctx*> cs;
  std::vector<long long> box_gen;  // each context's box generation when the batch was planned
  hipStream_t origin = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  SegOwner own;                 // the pass plan and trial groups (built once)
  SegOwner trial_own;           // per-call trial inputs (arena)
  Pass pass;
  double* h_outs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<DeGroup> de;
  long de_cap = 0;
  double* d_de = nullptr;
  std::vector<double> h_de;
};

extern "C" {

void msg_batch_destroy(msg_batch* b);

int msg_batch_create(msg_ctx** ctxs, int n, msg_batch** out) {
  if (!ctxs || n < 1 || !out) return set_error(MSG_ERR_ARG, "msg_batch_create: bad arguments");
  *out = nullptr;
  msg_batch* b = new msg_batch();
  auto fail = [&](int code) {
    msg_batch_destroy(b);
    return code;
  };
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(set_error(MSG_ERR_ARG, "msg_batch_create: null context"));
    Ctx* c = &ctxs[i]->c;
    if (int e = check_batchable(c)) return fail(e);
    b->ctx.push_back(ctxs[i]);
    b->cs.push_back(c);
    b->box_gen.push_back(c->box_gen);
  }
  // the contexts' own pending work must be done before the batch touches their buffers
  for (Ctx* c : b->cs)
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return fail(set_error(MSG_ERR_HIP, "msg_batch_create: context stream failed"));
  if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(MSG_ERR_HIP, "msg_batch_create: stream/event creation failed"));
  int e;
  if ((e = build_pass(b->cs, b->own, b->pass))) return fail(e);
  if ((e = plan_de(b->cs, b->own, b->de))) return fail(e);
  if (hipHostMalloc((void**)&b->h_outs, 64 * n * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(set_error(MSG_ERR_NOMEM, "msg_batch_create: hipHostMalloc failed"));
  if (hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(set_error(MSG_ERR_HIP, "msg_batch_create: capture failed"));
  const int err = launch_pass(b->pass, b->origin, b->side, b->fork, b->join);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->origin, &g);
  if (err) {
    if (g) hipGraphDestroy(g);
    return fail(err);
  }
  if (ce != hipSuccess) return fail(set_error(MSG_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(ce)));
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess)
    return fail(set_error(MSG_ERR_HIP, "msg_batch_create: graph instantiation failed"));
  *out = b;
  return MSG_OK;
}

void msg_batch_destroy(msg_batch* b) {
  if (!b) return;
  if (b->origin) hipStreamSynchronize(b->origin);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->d_de) hipFree(b->d_de);
  if (b->h_outs) hipHostFree(b->h_outs);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->join) hipEventDestroy(b->join);
  if (b->side) hipStreamDestroy(b->side);
  if (b->origin) hipStreamDestroy(b->origin);
  delete b;
}

// The plan holds every system's box: a box changed since (msg_set_box) makes it stale.
static int batch_current(const msg_batch* b, const char* who) {
  for (size_t s = 0; s < b->cs.size(); ++s)
    if (b->cs[s]->box_gen != b->box_gen[s])
      return set_error(MSG_ERR_STATE, std::string(who) + ": a system's box changed after msg_batch_create");
  return MSG_OK;
}

int msg_batch_utotal(msg_batch* b) {
  if (!b) return set_error(MSG_ERR_ARG, "msg_batch_utotal: null batch");
  if (int e = batch_current(b, "msg_batch_utotal")) return e;
  if (int e = order_after_contexts(b->cs, b->origin)) return e;
  MSG_HIP(hipGraphLaunch(b->exec, b->origin));
  for (Ctx* c : b->cs) {
    c->fresh = c->cells_valid = true;
    c->cells_maintained = false;  // the batched pass sorted the cells afresh
  }
  return MSG_OK;
}

int msg_batch_sync(msg_batch* b) {
  if (!b) return set_error(MSG_ERR_ARG, "msg_batch_sync: null batch");
  MSG_HIP(hipStreamSynchronize(b->origin));
  return MSG_OK;
}

int msg_batch_results(msg_batch* b, double* u_tot, double* u_rec, double* virial) {
  if (!b) return set_error(MSG_ERR_ARG, "msg_batch_results: null batch");
  const int n = (int)b->cs.size();
  MSG_HIP(hipMemcpyAsync(b->h_outs, b->pass.d_outs, 64 * n * sizeof(double), hipMemcpyDeviceToHost, b->origin));
  MSG_HIP(hipStreamSynchronize(b->origin));
  for (int s = 0; s < n; ++s)
    if (int e = read_out(b->h_outs + 64 * s, b->cs[s], s, u_tot, u_rec, virial)) return e;
  return MSG_OK;
}

int msg_batch_dutotal(msg_batch* b, const int* ntrial, const int* ip, const double* rotm,
                      int* ovl, double* du_tot) {
  if (!b || !ntrial || !ip || !rotm || !ovl || !du_tot)
    return set_error(MSG_ERR_ARG, "msg_batch_dutotal: bad arguments");
  if (int e = batch_current(b, "msg_batch_dutotal")) return e;
  // the trials read phi / S / cells of the current positions
  bool stale = false;
  for (Ctx* c : b->cs) stale |= !c->fresh;
  if (stale)
    if (int e = msg_batch_utotal(b)) return e;
  MSG_HIP(hipStreamSynchronize(b->origin));  // the arena and outputs of the last call are free
  const int n = (int)b->cs.size();
  if (int e = b->trial_own.reserve(trial_bytes(count_trials(ntrial, n), n, b->de.size()))) return e;
  b->trial_own.reset();
  DeLaunch L;
  if (int e = pack_trials(b->de, b->cs, ntrial, ip, rotm, b->trial_own, L)) return e;
  if (L.ntot == 0) return MSG_OK;
  if (L.ntot > b->de_cap) {
    if (b->d_de) hipFree(b->d_de);
    b->d_de = nullptr;
    b->de_cap = 0;
    const long cap = std::max(L.ntot, 4096L);
    MSG_HIP(hipMalloc((void**)&b->d_de, cap * kDeStride * sizeof(double)));
    b->de_cap = cap;
  }
  hipStream_t st = b->origin;
  if (int e = b->trial_own.flush(st)) return e;
  if (int e = launch_trials(b->de, L, b->d_de, st)) return e;
  b->h_de.resize(L.ntot * kDeStride);
  MSG_HIP(hipMemcpyAsync(b->h_de.data(), b->d_de, L.ntot * kDeStride * sizeof(double), hipMemcpyDeviceToHost, st));
  MSG_HIP(hipStreamSynchronize(st));
  unpack_trials(b->cs, ntrial, L, b->h_de.data(), ovl, du_tot);
  return MSG_OK;
}

int msg_batch_mc_pass(msg_batch* b, int ntrial, int isamp, const double* dtran, double beta, int* iseed,
                      int* rng_state, double* am, double* du_sum, int* naccept, int* events) {
  if (!b || ntrial < 0 || !dtran || !iseed || !rng_state || !am)
    return set_error(MSG_ERR_ARG, "msg_batch_mc_pass: bad arguments");
  if (int e = batch_current(b, "msg_batch_mc_pass")) return e;
  bool stale = false;
  for (Ctx* c : b->cs) stale |= !c->fresh;
  if (stale)
    if (int e = msg_batch_utotal(b)) return e;
  return mc_chain_batch(b->cs, b->origin, [b]() { return msg_batch_utotal(b); }, ntrial, isamp, dtran, beta,
                        iseed, rng_state, am, du_sum, naccept, events);
}

// ---- execution units (continuous batching) --------------------------------------------------
// A unit executes one batch of requests at a time on its own streams: a UTotal pass over
// the systems that asked for energy + forces (and over trial systems whose state is stale),
// then the MC trials of the trial systems, then one download of energies, forces and trial
// results.  Several units in flight overlap on the device; a system may be in one in-flight
// unit at a time (the caller's scheduler guarantees it).
}  // extern "C"

struct msg_unit {
  hipStream_t s = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
  SegOwner own;
  Pass pass;
  std::vector<Ctx*> ut, dec;
  std::vector<int> ntrial;
  std::vector<DeGroup> de;
  DeLaunch L;
  bool launched = false, want_forces = false;
  double* d_de = nullptr;
  long de_cap = 0;
  double* h_outs = nullptr;
  long outs_cap = 0;
  double* h_de = nullptr;
  long hde_cap = 0;
  double* h_f = nullptr;
  long hf_cap = 0;
  std::vector<long> foff;
};

namespace {

int grow_pinned(double** p, long* cap, long need) {
  if (need <= *cap) return MSG_OK;
  const long n = std::max(need, 2 * *cap + 1024);
  if (*p) hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  MSG_HIP(hipHostMalloc((void**)p, n * sizeof(double), hipHostMallocDefault));
  *cap = n;
  return MSG_OK;
}

// Plans the unit's pass and trials into its arena, doubling the arena until they fit.
int plan_unit(msg_unit* u, const std::vector<Ctx*>& pcs, const int* ip, const double* rotm) {
  for (int attempt = 0; attempt < 8; ++attempt) {
    u->own.reset();
    u->de.clear();
    int e = build_pass(pcs, u->own, u->pass);
    if (!e && !u->dec.empty()) e = plan_de(u->dec, u->own, u->de);
    if (!e && !u->dec.empty()) e = pack_trials(u->de, u->dec, u->ntrial.data(), ip, rotm, u->own, u->L);
    if (e != MSG_ERR_NOMEM) return e;
    if (int e2 = u->own.reserve(2 * u->own.cap)) return e2;
  }
  return set_error(MSG_ERR_NOMEM, "unit plan does not fit its arena");
}

}  // namespace

extern "C" {

void msg_unit_destroy(msg_unit* u) {
  if (!u) return;
  if (u->s) hipStreamSynchronize(u->s);
  if (u->d_de) hipFree(u->d_de);
  if (u->h_outs) hipHostFree(u->h_outs);
  if (u->h_de) hipHostFree(u->h_de);
  if (u->h_f) hipHostFree(u->h_f);
  if (u->fork) hipEventDestroy(u->fork);
  if (u->join) hipEventDestroy(u->join);
  if (u->done) hipEventDestroy(u->done);
  if (u->side) hipStreamDestroy(u->side);
  if (u->s) hipStreamDestroy(u->s);
  delete u;
}

int msg_unit_create(msg_unit** out) {
  if (!out) return set_error(MSG_ERR_ARG, "msg_unit_create: null output");
  msg_unit* u = new msg_unit();
  *out = nullptr;
  if (hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&u->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&u->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&u->join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&u->done, hipEventDisableTiming) != hipSuccess) {
    msg_unit_destroy(u);
    return set_error(MSG_ERR_HIP, "msg_unit_create: stream/event creation failed");
  }
  if (int e = u->own.init_arena(16 << 20)) {
    msg_unit_destroy(u);
    return e;
  }
  *out = u;
  return MSG_OK;
}

int msg_unit_query(msg_unit* u) {
  if (!u) return set_error(MSG_ERR_ARG, "msg_unit_query: null unit");
  if (!u->launched) return 1;
  const hipError_t e = hipEventQuery(u->done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return set_error(MSG_ERR_HIP, std::string("msg_unit_query: ") + hipGetErrorString(e));
}

// Launches (asynchronously) a UTotal pass over ut[0..n_u) (forces downloaded when
// want_forces) and ntrial[s] trials of each dec[s] (ip 1-based and rotm concatenated over
// dec).  A system may appear in both lists (its pass runs first), but not twice in one.
int msg_unit_launch(msg_unit* u, msg_ctx** ut, int n_u, int want_forces, msg_ctx** dec, int n_d,
                    const int* ntrial, const int* ip, const double* rotm) {
  if (!u || n_u < 0 || n_d < 0 || (n_u && !ut) || (n_d && (!dec || !ntrial || !ip || !rotm)))
    return set_error(MSG_ERR_ARG, "msg_unit_launch: bad arguments");
  if (u->launched) {
    const hipError_t q = hipEventQuery(u->done);
    if (q == hipErrorNotReady) return set_error(MSG_ERR_STATE, "msg_unit_launch: unit still running");
    if (q != hipSuccess) return set_error(MSG_ERR_HIP, std::string("msg_unit_launch: ") + hipGetErrorString(q));
  }
  u->launched = false;
  u->ut.clear();
  u->dec.clear();
  u->ntrial.assign(ntrial, ntrial + n_d);
  std::vector<Ctx*> pcs;
  for (int i = 0; i < n_u; ++i) {
    if (!ut[i]) return set_error(MSG_ERR_ARG, "msg_unit_launch: null context");
    u->ut.push_back(&ut[i]->c);
  }
  for (int i = 0; i < n_d; ++i) {
    if (!dec[i]) return set_error(MSG_ERR_ARG, "msg_unit_launch: null context");
    u->dec.push_back(&dec[i]->c);
  }
  for (size_t i = 0; i < u->ut.size(); ++i)
    for (size_t j = 0; j < i; ++j)
      if (u->ut[i] == u->ut[j]) return set_error(MSG_ERR_ARG, "msg_unit_launch: system listed twice");
  for (size_t i = 0; i < u->dec.size(); ++i)
    for (size_t j = 0; j < i; ++j)
      if (u->dec[i] == u->dec[j]) return set_error(MSG_ERR_ARG, "msg_unit_launch: system listed twice");
  pcs = u->ut;
  for (Ctx* c : u->dec)
    if (!c->fresh && std::find(pcs.begin(), pcs.end(), c) == pcs.end()) pcs.push_back(c);
  for (Ctx* c : u->dec)
    if (int e = check_batchable(c)) return e;
  if (int e = plan_unit(u, pcs, ip, rotm)) return e;
  // buffers: trial outputs on the device, pinned host copies of everything downloaded
  const long ntot = u->dec.empty() ? 0 : u->L.ntot;
  if (ntot > u->de_cap) {
    if (u->d_de) hipFree(u->d_de);
    u->d_de = nullptr;
    u->de_cap = 0;
    const long cap = std::max(ntot, 4096L);
    MSG_HIP(hipMalloc((void**)&u->d_de, cap * kDeStride * sizeof(double)));
    u->de_cap = cap;
  }
  u->want_forces = want_forces != 0;
  u->foff.assign(u->ut.size() + 1, 0);
  for (size_t i = 0; i < u->ut.size(); ++i) u->foff[i + 1] = u->foff[i] + 3L * u->ut[i]->np;
  if (int e = grow_pinned(&u->h_outs, &u->outs_cap, 64L * (long)pcs.size())) return e;
  if (int e = grow_pinned(&u->h_de, &u->hde_cap, ntot * kDeStride)) return e;
  if (u->want_forces)
    if (int e = grow_pinned(&u->h_f, &u->hf_cap, u->foff.back())) return e;
  // launch
  std::vector<Ctx*> all = pcs;
  for (Ctx* c : u->dec) all.push_back(c);
  if (int e = order_after_contexts(all, u->s)) return e;
  if (int e = u->own.flush(u->s)) return e;
  if (int e = launch_pass(u->pass, u->s, u->side, u->fork, u->join)) return e;
  for (Ctx* c : pcs) {
    c->fresh = c->cells_valid = true;
    c->cells_maintained = false;
  }
  if (ntot)
    if (int e = launch_trials(u->de, u->L, u->d_de, u->s)) return e;
  if (!pcs.empty())
    MSG_HIP(hipMemcpyAsync(u->h_outs, u->pass.d_outs, 64 * pcs.size() * sizeof(double), hipMemcpyDeviceToHost, u->s));
  if (ntot)
    MSG_HIP(hipMemcpyAsync(u->h_de, u->d_de, ntot * kDeStride * sizeof(double), hipMemcpyDeviceToHost, u->s));
  if (u->want_forces)
    for (size_t i = 0; i < u->ut.size(); ++i)
      MSG_HIP(hipMemcpyAsync(u->h_f + u->foff[i], u->ut[i]->d_force, 3 * u->ut[i]->np * sizeof(double),
                             hipMemcpyDeviceToHost, u->s));
  MSG_HIP(hipEventRecord(u->done, u->s));
  u->launched = true;
  return MSG_OK;
}

// Waits for the unit and copies its results: u_tot, u_rec, virial of the n_u UTotal
// systems, their forces (concatenated force(3,np) per system; needs want_forces), and the
// overlap flags and energy changes of all trials (in launch order).  Any may be null.
int msg_unit_results(msg_unit* u, double* u_tot, double* u_rec, double* virial, double* force,
                     int* ovl, double* du_tot) {
  if (!u || !u->launched) return set_error(MSG_ERR_STATE, "msg_unit_results: nothing launched");
  MSG_HIP(hipEventSynchronize(u->done));
  for (size_t s = 0; s < u->pass.cs.size(); ++s) {
    const bool want = s < u->ut.size();
    if (int e = read_out(u->h_outs + 64 * s, u->pass.cs[s], (int)s, want ? u_tot : nullptr,
                         want ? u_rec : nullptr, want ? virial : nullptr))
      return e;
  }
  if (force && u->want_forces) std::copy(u->h_f, u->h_f + u->foff.back(), force);
  if (!u->dec.empty() && ovl && du_tot) unpack_trials(u->dec, u->ntrial.data(), u->L, u->h_de, ovl, du_tot);
  return MSG_OK;
}

// Completion query of a context's stream (1 = idle), for request loops that poll.
int msg_query(msg_ctx* h) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_query: null context");
  const hipError_t e = hipStreamQuery(h->c.stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return set_error(MSG_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

// Completion query of a batch's stream (1 = idle).
int msg_batch_query(msg_batch* b) {
  if (!b) return set_error(MSG_ERR_ARG, "msg_batch_query: null batch");
  const hipError_t e = hipStreamQuery(b->origin);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return set_error(MSG_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

}  // extern "C"
