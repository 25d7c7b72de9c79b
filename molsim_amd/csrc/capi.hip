// C-ABI of the MI355X non-bonded path: context life cycle, state transfer, the UTotal launch
// sequence (captured once into a hipGraph for the mixed-precision path), the final FP64
// reduction, error flags and per-kernel event timing.
#include <cstring>
#include <mutex>

#include "common.h"
#include "reduce.h"

namespace msg {

static thread_local std::string g_err = "";

int set_error(int code, const std::string& text) {
  g_err = text;
  return code;
}

int dalloc(Ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return set_error(MSG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->bytes += (long long)bytes;
  return MSG_OK;
}

static const char* kKernelNames[MSG_K_COUNT] = {
    "bin", "scan", "pair", "spread", "fft_x", "fft_y", "fft_zconv", "fft_yi", "fft_xi",
    "gather", "reduce", "ewald_std", "dutotal", "vlist", "scatter", "finalize"};

void prof_begin(Ctx* c, int kid) {
  if (!c->prof) return;
  if (c->ev_used + 2 > (int)(sizeof(c->ev) / sizeof(c->ev[0]))) {
    // pool exhausted: drain it
    hipStreamSynchronize(c->stream);
    for (int k = 0; k + 1 < c->ev_used; k += 2) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]);
      c->prof_ms[c->ev_kernel[k / 2]] += ms;
      c->prof_n[c->ev_kernel[k / 2]] += 1;
    }
    c->ev_used = 0;
    c->ev_kernel.clear();
  }
  hipEventRecord(c->ev[c->ev_used], c->stream);
  c->ev_kernel.push_back(kid);
}

void prof_end(Ctx* c, int kid) {
  (void)kid;
  if (!c->prof) return;
  hipEventRecord(c->ev[c->ev_used + 1], c->stream);
  c->ev_used += 2;
}

static void prof_drain(Ctx* c) {
  if (!c->prof || c->ev_used == 0) return;
  hipStreamSynchronize(c->stream);
  for (int k = 0; k + 1 < c->ev_used; k += 2) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]);
    c->prof_ms[c->ev_kernel[k / 2]] += ms;
    c->prof_n[c->ev_kernel[k / 2]] += 1;
  }
  c->ev_used = 0;
  c->ev_kernel.clear();
}

namespace {

__global__ void k_zero(int* a, int na, int* b, int nb, int* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0;
  if (i < nb) b[i] = 0;
  if (i == 0 && flags) flags[0] = 0;
}

struct PairSumArgs {
  const double* rows;
  int nrow, np1;
  double* out;
};

// The pair rows summed into one row.  With SPME this runs on the pair branch, beside the
// mesh kernels, so the final reduction reads one row (the same totals it would form from
// the rows itself).
__device__ __forceinline__ void pair_sum_body(const PairSumArgs& a) {
  __shared__ double red[1024];
  pair_rows_total(a.rows, a.nrow, a.np1, red);
  if (threadIdx.x < a.np1) a.out[threadIdx.x] = red[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_pair_sum(PairSumArgs a) { pair_sum_body(a); }

__global__ __launch_bounds__(1024) void k_pair_sum_seg(const PairSumArgs* __restrict__ args,
                                                       const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  pair_sum_body(args[s.x]);
}

void seg_pair_sum(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_sum_seg, dim3(g.nblk), dim3(1024), 0, s, (const PairSumArgs*)g.d_args, g.d_map);
}

__global__ __launch_bounds__(1024) void k_reduce(ReduceArgs a) { reduce_body<1024>(a); }

__global__ __launch_bounds__(1024) void k_reduce_seg(const ReduceArgs* __restrict__ args,
                                                     const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  reduce_body<1024>(args[s.x]);
}

void seg_reduce(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_seg, dim3(g.nblk), dim3(1024), 0, s, (const ReduceArgs*)g.d_args, g.d_map);
}

}  // namespace

int launch_zero(Ctx* c) {
  const int na = c->grid.ncell;
  const int nb = c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.nkey : 0;
  const int n = std::max(std::max(na, nb), 1);
  hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d_cell_count, na,
                     c->d_brick_count, nb, c->d_flags);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// The cell counters a list rebuild's binning filled (the brick counters, which that binning
// leaves alone, and the flags, which a following list build may raise, are kept).
int launch_zero_counts(Ctx* c) {
  const int na = c->grid.ncell;
  hipLaunchKernelGGL(k_zero, dim3((na + 255) / 256), dim3(256), 0, c->stream, c->d_cell_count, na,
                     c->d_brick_count, 0, (int*)nullptr);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

ReduceArgs reduce_args(Ctx* c, bool presum) {
  ReduceArgs a;
  a.pair = c->d_part + (presum ? c->part_off_pairsum : c->part_off_pair);
  a.npair = presum ? 1 : c->npart_pair;
  a.ntp = c->nptpt;
  a.spme = c->d_part + c->part_off_spme;
  a.nspme = c->nzblocks;
  a.ewstd = c->d_part + c->part_off_ewstd;
  a.newstd = c->npart_ewstd;
  a.M = c->d_M;
  a.ewald = c->P.ewald_mode;
  a.lsurf = c->P.lsurf;
  a.self_raw = -c->P.alpha / std::sqrt(M_PI) * c->sumq2;
  const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
  a.surf_fac = 2.0 * M_PI / (3.0 * vol);
  a.epsi4pi = c->P.epsi4pi;
  a.out = c->d_out;
  a.flags = c->d_flags;
  a.cell_count = c->d_cell_count;
  a.ncell = c->grid.ncell;
  a.brick_count = c->d_brick_count;
  a.nbrick = c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.nkey : 0;
  return a;
}

static PairSumArgs pair_sum_args(Ctx* c) {
  PairSumArgs a;
  a.rows = c->d_part + c->part_off_pair;
  a.nrow = c->npart_pair;
  a.np1 = c->nptpt + 1;
  a.out = c->d_part + c->part_off_pairsum;
  return a;
}

int plan_pair_sum(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out) {
  SegBuilder<PairSumArgs> b;
  for (Ctx* c : cs) b.add(0, seg_pair_sum, pair_sum_args(c), 1);
  return b.finish(own, out);
}

int launch_pair_sum(Ctx* c) {
  hipLaunchKernelGGL(k_pair_sum, dim3(1), dim3(1024), 0, c->stream, pair_sum_args(c));
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int plan_reduce(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out) {
  SegBuilder<ReduceArgs> b;
  for (Ctx* c : cs) b.add(0, seg_reduce, reduce_args(c, true), 1);  // batches pre-sum the pair rows
  return b.finish(own, out);
}

int launch_reduce(Ctx* c, bool presum) {
  const ReduceArgs a = reduce_args(c, presum);
  prof_begin(c, MSG_K_REDUCE);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, a);
  prof_end(c, MSG_K_REDUCE);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// The UTotal launch sequence (capturable; with a Verlet list the host waits once, for the
// skin test, which is why that mode launches eagerly).  Counters and
// flags are zero on entry (cleared at allocation and by the previous pass's reduction).
// With SPME the work forks after the binning pass: the cell list (scan, scatter, ranking,
// Verlet rebuild), the pair kernel and the pair-row sum run on a second stream beside the
// brick scan/scatter and the mesh kernels; the gather, which adds to the pair forces, waits
// for them.  One fork only: each cross-stream dependency costs tens of microseconds of
// graph execution time.  With per-kernel timing on, everything runs on one stream so each kernel is
// timed alone.
int utotal_sequence(Ctx* c) {
  int e;
  bool rebuild = false;
  const bool verlet = c->P.list_mode == MSG_LIST_VERLET;
  const bool branch = c->P.ewald_mode == MSG_EWALD_SPM && !c->prof;
  // Verlet list with SPME: the binning without cells and the brick sort do not depend on the
  // skin test, so they are queued before the host waits for its result (and run before the
  // fork, where the pair kernel cannot crowd out the brick scatter's atomics); a rebuild
  // bins the cells on the pair branch.
  // Within msg_md_run the rebuilds are decided at the end of each step (NList after VelVer);
  // a pass then builds a list only when none is valid.
  const bool early = verlet && branch;
  if (early) {
    bool pending = false;
    if (c->md_managed) rebuild = !c->vl_valid;
    else if ((e = verlet_check_begin(c, pending))) return e;
    if ((e = launch_bin(c, false)) || (e = launch_sort(c, 2))) return e;
    if (!c->md_managed) {
      if (pending) {
        if ((e = verlet_check_end(c, rebuild))) return e;
      } else {
        rebuild = true;
      }
    }
  } else if (verlet) {
    if (c->md_managed) rebuild = !c->vl_valid;
    else if ((e = verlet_check(c, rebuild))) return e;
  }
  // with a Verlet list the cell order is only needed to rebuild the list (later users such
  // as DUTotal complete it on demand); a cell order that accepted moves have patched
  // (UpdateCellIp) is kept, as the reference keeps its linked lists between moves
  const bool maintained = c->cells_maintained && c->cells_valid && !verlet;
  const bool cells_needed = (!verlet || rebuild) && !maintained;
  // the wave-per-cell-chunk pair kernel needs the chunk table of a fresh sort: a patched
  // order takes the neighbour-parallel kernel for this pass
  const int lph_saved = c->pair_lph, npair_saved = c->npart_pair;
  if (maintained && c->pair_lph == 0) {
    c->pair_lph = 8;
    c->npart_pair = pair_blocks(c);
  }
  struct Restore {
    Ctx* c;
    int lph, np;
    ~Restore() {
      c->pair_lph = lph;
      c->npart_pair = np;
    }
  } restore{c, lph_saved, npair_saved};
  if (!early && (e = launch_bin(c, cells_needed))) return e;
  auto cell_work = [&]() -> int {
    if (early && rebuild)
      if (int e2 = launch_bin(c, true, false)) return e2;
    // without a rebuild the Verlet pair kernel reads its slot records, refreshed here from
    // the current positions, and the cell slots are not sorted at all (complete_cells does
    // it for later users)
    if (verlet && !rebuild)
      if (int e2 = launch_vl_records(c)) return e2;
    const int which = (cells_needed ? 1 : 0) | (branch ? 0 : 2);
    if (int e2 = launch_sort(c, which)) return e2;
    if (cells_needed)
      if (int e2 = launch_finalize(c)) return e2;
    if (rebuild) {
      if (int e2 = launch_vlist_build(c)) return e2;
      c->vl_valid = true;
    }
    return MSG_OK;
  };
  c->cells_valid = cells_needed || maintained;
  if (cells_needed) c->cells_maintained = false;
  if (branch) {
    MSG_HIP(hipEventRecord(c->ev_fork, c->stream));
    // the mesh branch is the longer chain: recorded first, it is dispatched first
    if (!early && (e = launch_sort(c, 2))) return e;
    if ((e = launch_spme_mesh(c))) return e;
    MSG_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    hipStream_t main = c->stream;
    c->stream = c->side;
    e = cell_work();
    if (!e) e = launch_pair(c, false);
    if (!e) e = launch_pair_sum(c);
    c->stream = main;
    if (e) return e;
    MSG_HIP(hipEventRecord(c->ev_join, c->side));
    MSG_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    // the final reduction rides as one extra workgroup of the gather launch
    return launch_spme_gather(c, true);
  } else {
    if ((e = cell_work())) return e;
    if (c->P.ewald_mode == MSG_EWALD_SPM) {
      // per-kernel timing: one stream, so no kernel's time includes another's
      if ((e = launch_pair(c, false))) return e;
      if ((e = launch_spme_mesh(c))) return e;
      if ((e = launch_spme_gather(c))) return e;
    } else {
      if ((e = launch_pair(c, c->P.ewald_mode == MSG_EWALD_NONE))) return e;
      if (c->P.ewald_mode == MSG_EWALD_STD)
        if ((e = launch_ewstd(c))) return e;
    }
  }
  return launch_reduce(c, false);
}

static int flags_to_error(int f);

// The cell order of the current positions for users that need it after a Verlet-list pass
// that left it out: the binning runs again on the same positions (same cells, same counts),
// the cells are sorted and ranked, and the counters the binning filled are cleared.
int complete_cells(Ctx* c) {
  if (c->cells_valid) return MSG_OK;
  int e;
  if ((e = launch_bin(c)) || (e = launch_sort(c, 1)) || (e = launch_finalize(c)) || (e = launch_zero(c)))
    return e;
  c->cells_valid = true;
  return MSG_OK;
}

int utotal_refresh(Ctx* c) {
  if (int e = utotal_sequence(c)) return e;
  MSG_HIP(hipMemcpyAsync(c->h_out, c->d_out, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  if (int e = flags_to_error((int)c->h_out[c->nptpt + 7])) return e;
  c->fresh = true;
  return MSG_OK;
}

int check_ready_state(Ctx* c) {
  if (!c->have_table) return set_error(MSG_ERR_STATE, "msg_set_table has not been called");
  if (!c->have_particles) return set_error(MSG_ERR_STATE, "msg_set_particles has not been called");
  if (!c->have_positions) return set_error(MSG_ERR_STATE, "positions have not been set");
  if (c->P.ewald_mode != MSG_EWALD_NONE && !c->have_ewald)
    return set_error(MSG_ERR_STATE, "msg_ewald_setup has not been called");
  return MSG_OK;
}

static int flags_to_error(int f) {
  if (f & FLAG_R2UMIN) return set_error(MSG_ERR_R2UMIN, "pair below the table: r2 < r2umin");
  if (f & FLAG_CELL) return set_error(MSG_ERR_CELL, "position outside the box");
  if (f & FLAG_NLIST) return set_error(MSG_ERR_NLIST, "neighbour list overflow");
  if (f & FLAG_IBUF) return set_error(MSG_ERR_IBUF, "table scan past its end");
  return MSG_OK;
}

}  // namespace msg

using namespace msg;

extern "C" {

const char* msg_last_error(void) { return g_err.c_str(); }

const char* msg_kernel_name(int k) {
  return (k >= 0 && k < MSG_K_COUNT) ? kKernelNames[k] : "?";
}

int msg_create(const msg_params* p, msg_ctx** out) {
  if (!p || !out) return set_error(MSG_ERR_ARG, "msg_create: null argument");
  *out = nullptr;
  if (p->np < 1 || p->npt < 1 || p->npt > MSG_MAX_TYPES)
    return set_error(MSG_ERR_ARG, "msg_create: np >= 1 and 1 <= npt <= 8 required");
  for (int d = 0; d < 3; ++d)
    if (!(p->boxlen[d] > 0)) return set_error(MSG_ERR_ARG, "msg_create: box lengths must be > 0");
  if (!(p->rcut > 0)) return set_error(MSG_ERR_CELL, "msg_create: rcut must be > 0");
  if (p->list_mode < 0 || p->list_mode > 2 || p->ewald_mode < 0 || p->ewald_mode > 2)
    return set_error(MSG_ERR_ARG, "msg_create: bad list/ewald mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(MSG_ERR_HIP, "msg_create: no HIP device");
  msg_ctx* h = new msg_ctx();
  Ctx* c = &h->c;
  c->P = *p;
  if (c->P.force_scale == 0.0) c->P.force_scale = 1.0;
  c->np = p->np;
  c->npt = p->npt;
  c->nptpt = p->npt * (p->npt + 1) / 2;
  c->fp64 = p->precision == MSG_PREC_FP64;
  hipGetDevice(&c->device);
  int e = MSG_OK;
  auto fail = [&](int code) {
    msg_destroy(h);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_skin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_cells, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(MSG_ERR_HIP, "stream/event creation failed"));
  const int np = c->np;
  if ((e = dalloc(c, (void**)&c->d_ro, 3 * np * sizeof(double)))) return fail(e);
  if ((e = dalloc(c, (void**)&c->d_az, np * sizeof(double)))) return fail(e);
  if ((e = dalloc(c, (void**)&c->d_ipt, np * sizeof(int)))) return fail(e);
  if ((e = dalloc(c, (void**)&c->d_force, 3 * np * sizeof(double)))) return fail(e);
  if ((e = lists_setup(c))) return fail(e);
  if (p->list_mode == MSG_LIST_VERLET && (e = verlet_setup(c))) return fail(e);
  // partial-sum layout
  c->pair_lph = default_pair_tile(c);
  c->npart_pair = pair_blocks(c);
  c->npart_pair_cap = std::max(c->npart_pair, std::max((np + 3) / 4, 2048 + 1));
  c->npart_dip = (np + 255) / 256;
  c->part_off_pair = 0;
  c->part_off_dip = c->part_off_pair + c->npart_pair_cap * (kMaxTP + 1);
  c->part_off_spme = c->part_off_dip + 3 * c->npart_dip;
  const int max_spme_blocks = 2 * 2 * 256 * 130;  // upper bound for nzblocks * 2 (K <= 256)
  c->part_off_ewstd = c->part_off_spme + max_spme_blocks;
  c->part_off_pairsum = c->part_off_ewstd + 2 * 4096;
  const int nparts = c->part_off_pairsum + kMaxTP + 1;
  if ((e = dalloc(c, (void**)&c->d_part, nparts * sizeof(double)))) return fail(e);
  if ((e = dalloc(c, (void**)&c->d_out, 64 * sizeof(double)))) return fail(e);
  if ((e = dalloc(c, (void**)&c->d_flags, 16 * sizeof(int)))) return fail(e);
  // word 0: the pass's flags, 1: sticky over a multi-step call, 2: the list export's own
  if (hipMemset(c->d_flags, 0, 16 * sizeof(int)) != hipSuccess) return fail(set_error(MSG_ERR_HIP, "hipMemset failed"));
  if (hipHostMalloc((void**)&c->h_out, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(set_error(MSG_ERR_NOMEM, "hipHostMalloc failed"));
  if (hipHostMalloc((void**)&c->h_flags, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess)
    return fail(set_error(MSG_ERR_NOMEM, "hipHostMalloc failed"));
  if ((e = launch_zero(c))) return fail(e);
  if ((e = noalloc_setup(c->device))) return fail(e);  // the scratch of msg_dutotal_batch_noalloc
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_error(MSG_ERR_HIP, "sync failed"));
  *out = h;
  return MSG_OK;
}

void msg_destroy(msg_ctx* h) {
  if (!h) return;
  Ctx* c = &h->c;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  void* ptrs[] = {c->d_ro, c->d_az, c->d_ipt, c->d_tcoef, c->d_tbins, c->d_cellid, c->d_cell_chunk,
                  c->d_cell_count, c->d_cell_start, c->d_cell_fill, c->d_cell_tmp, c->d_cell_ip,
                  c->d_slot_of, c->d_cpos, c->d_crel, c->d_ctype, c->d_slot_cell, c->d_brick_count,
                  c->d_brick_start, c->d_brick_fill, c->d_brick_ip, c->d_mm,
                  c->d_frac, c->d_Q, c->d_C, c->d_G, c->d_tw, c->d_kfac, c->d_kvec, c->d_sk,
                  c->d_force, c->d_part, c->d_out, c->d_flags, c->d_de, c->d_de_ip, c->d_de_r,
                  c->d_tcoef64, c->d_gk, c->d_item_cell, c->d_bq, c->d_bmm, c->d_M,
                  c->d_vl_home, c->d_vl_n, c->d_vl_nbr, c->d_vl_pos, c->d_vl_rel, c->d_vl_slot, c->d_ro_ref, c->d_vel,
                  c->d_acc, c->d_minv, c->d_ke_part, c->d_rank, c->d_laz, c->d_drosum, c->d_r2, c->d_bm,
                  c->d_pm, c->d_skin};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (c->h_out) hipHostFree(c->h_out);
  if (c->h_flags) hipHostFree(c->h_flags);
  if (c->h_skin) hipHostFree(c->h_skin);
  for (auto& ev : c->ev)
    if (ev) hipEventDestroy(ev);
  if (c->side) hipStreamDestroy(c->side);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->ev_skin) hipEventDestroy(c->ev_skin);
  if (c->ev_cells) hipEventDestroy(c->ev_cells);
  if (c->ev_batch) hipEventDestroy(c->ev_batch);
  if (c->stream) hipStreamDestroy(c->stream);
  delete h;
}

int msg_set_table(msg_ctx* h, const double* ubuf, int nbuf, const int* iubuflow, const int* nugrid,
                  const double* r2umin, const double* r2atat) {
  if (!h || !ubuf || nbuf < 12 || !iubuflow || !nugrid || !r2umin || !r2atat)
    return set_error(MSG_ERR_ARG, "msg_set_table: bad arguments");
  Ctx* c = &h->c;
  const int ntp = c->nptpt;
  c->h_ubuf.assign(ubuf, ubuf + nbuf);
  c->h_iubuflow.assign(iubuflow, iubuflow + ntp);
  c->h_nugrid.assign(nugrid, nugrid + ntp);
  c->h_r2umin.assign(r2umin, r2umin + ntp);
  c->h_r2atat.assign(r2atat, r2atat + ntp);
  if (int e = pack_table(c)) return e;
  c->have_table = true;
  c->graph_ok = false;
  return MSG_OK;
}

int msg_set_particles(msg_ctx* h, const double* az, const int* iptpn) {
  if (!h || !az || !iptpn) return set_error(MSG_ERR_ARG, "msg_set_particles: null argument");
  Ctx* c = &h->c;
  std::vector<int> t(c->np);
  double s = 0.0;
  for (int i = 0; i < c->np; ++i) {
    t[i] = iptpn[i] - 1;
    if (t[i] < 0 || t[i] >= c->npt) return set_error(MSG_ERR_ARG, "msg_set_particles: type out of range");
    s += az[i] * az[i];
  }
  c->sumq2 = s;
  MSG_HIP(hipMemcpy(c->d_az, az, c->np * sizeof(double), hipMemcpyHostToDevice));
  MSG_HIP(hipMemcpy(c->d_ipt, t.data(), c->np * sizeof(int), hipMemcpyHostToDevice));
  c->h_ipt = t;
  c->h_az.assign(az, az + c->np);
  spme_charge_scale(c);  // the spread's fixed-point unit (captured with the pass)
  c->have_particles = true;
  c->graph_ok = false;   // the captured reduction holds the self term
  c->fresh = false;
  c->vl_valid = false;   // the Verlet list's records hold the types
  return MSG_OK;
}

int msg_set_weak_charges(msg_ctx* h, const int* laz, const double* zat) {
  if (!h || !laz || !zat) return set_error(MSG_ERR_ARG, "msg_set_weak_charges: null argument");
  Ctx* c = &h->c;
  if (!c->have_particles) return set_error(MSG_ERR_STATE, "msg_set_particles has not been called");
  if (c->P.ewald_mode == MSG_EWALD_SPM)  // Src/denergy.F90:1565-1570
    return set_error(MSG_ERR_ARG, "weak charges need std Ewald or no Ewald (SPME is not supported with them)");
  if (!c->d_laz)
    if (int e = dalloc(c, (void**)&c->d_laz, c->np * sizeof(int))) return e;
  c->h_zat.assign(zat, zat + c->npt);
  c->h_laz.resize(c->np);
  double s = 0.0;
  for (int i = 0; i < c->np; ++i) {
    c->h_laz[i] = laz[i] != 0;
    c->h_az[i] = c->h_laz[i] ? c->h_zat[c->h_ipt[i]] : 0.0;  // as IOCnf sets az, Src/molsim.F90:895-898
    s += c->h_az[i] * c->h_az[i];
  }
  c->sumq2 = s;
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipMemcpy(c->d_laz, c->h_laz.data(), c->np * sizeof(int), hipMemcpyHostToDevice));
  MSG_HIP(hipMemcpy(c->d_az, c->h_az.data(), c->np * sizeof(double), hipMemcpyHostToDevice));
  c->graph_ok = false;  // the captured pass must see the charge states
  c->fresh = false;
  c->vl_valid = false;  // the Verlet list's records hold the charge states
  c->cells_maintained = false;  // the cell records' tags are set by a fresh sort
  return MSG_OK;
}

int msg_commit_charge(msg_ctx* h, int nptm, const int* ipnptm) {
  if (!h || !ipnptm || nptm < 1 || nptm > 2) return set_error(MSG_ERR_ARG, "msg_commit_charge: one or two particles");
  Ctx* c = &h->c;
  if (c->h_laz.empty()) return set_error(MSG_ERR_STATE, "msg_commit_charge: weak charges are not set");
  for (int m = 0; m < nptm; ++m)
    if (ipnptm[m] < 1 || ipnptm[m] > c->np) return set_error(MSG_ERR_ARG, "msg_commit_charge: ip out of range");
  if (nptm == 2 && ipnptm[0] == ipnptm[1]) return set_error(MSG_ERR_ARG, "msg_commit_charge: the same particle twice");
  MSG_HIP(hipStreamSynchronize(c->stream));
  for (int m = 0; m < nptm; ++m) {  // laz and az of the flipped particles, Src/mc.F90:3781-3785
    const int i = ipnptm[m] - 1;
    c->sumq2 -= c->h_az[i] * c->h_az[i];
    c->h_laz[i] = !c->h_laz[i];
    c->h_az[i] = c->h_laz[i] ? c->h_zat[c->h_ipt[i]] : 0.0;
    c->sumq2 += c->h_az[i] * c->h_az[i];
    MSG_HIP(hipMemcpy(c->d_laz + i, &c->h_laz[i], sizeof(int), hipMemcpyHostToDevice));
    MSG_HIP(hipMemcpy(c->d_az + i, &c->h_az[i], sizeof(double), hipMemcpyHostToDevice));
    if (int e = set_charge_tag(c, i)) return e;  // a patched cell order keeps its records
  }
  c->graph_ok = false;  // self term of the captured reduction
  c->fresh = false;     // structure factors and the cell records' charge tags are rebuilt
  c->vl_valid = false;  // so are the Verlet list's (rebuilt at the next pass)
  return MSG_OK;
}

int msg_set_positions(msg_ctx* h, const double* ro) {
  if (!h || !ro) return set_error(MSG_ERR_ARG, "msg_set_positions: null argument");
  Ctx* c = &h->c;
  MSG_HIP(hipMemcpyAsync(c->d_ro, ro, 3 * c->np * sizeof(double), hipMemcpyHostToDevice, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->have_positions = true;
  c->fresh = false;
  c->cells_maintained = false;  // new positions: the next pass sorts the cells (SetCellList)
  return MSG_OK;
}

int msg_set_positions_device(msg_ctx* h, const double* d_ro) {
  if (!h || !d_ro) return set_error(MSG_ERR_ARG, "msg_set_positions_device: null argument");
  Ctx* c = &h->c;
  MSG_HIP(hipMemcpyAsync(c->d_ro, d_ro, 3 * c->np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  c->have_positions = true;
  c->fresh = false;
  c->cells_maintained = false;
  return MSG_OK;
}

int msg_ewald_setup(msg_ctx* h) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_ewald_setup: null context");
  Ctx* c = &h->c;
  int e = MSG_OK;
  if (c->P.ewald_mode == MSG_EWALD_SPM) e = spme_setup(c);
  else if (c->P.ewald_mode == MSG_EWALD_STD) e = ewstd_setup(c);
  if (e) return e;
  c->have_ewald = true;
  c->have_gk = false;
  c->fresh = false;
  c->graph_ok = false;
  return MSG_OK;
}

// Box change (VolChange and the other callers of SetBoxParam + EwaldSetup + InitCellList,
// Src/mc.F90:3423-3436, md.F90:699-700): the cell grid and its buffers, the Verlet list
// width, the SPME influence function or the std-Ewald k-vectors follow the new box; the
// caller then sets the (scaled) positions.  Everything derived from the old box is dropped.
int msg_set_box(msg_ctx* h, const double* boxlen) {
  if (!h || !boxlen) return set_error(MSG_ERR_ARG, "msg_set_box: null argument");
  Ctx* c = &h->c;
  for (int d = 0; d < 3; ++d)
    if (!(boxlen[d] > 0.0)) return set_error(MSG_ERR_ARG, "msg_set_box: box lengths must be > 0");
  MSG_HIP(hipStreamSynchronize(c->stream));
  // Everything derived from the old box is dropped before any buffer is replaced, so a failed
  // call leaves no graph or batch that still points at freed memory; on failure the context
  // keeps the old box lengths and is usable again after a successful msg_set_box.
  c->fresh = false;
  c->cells_valid = false;
  c->cells_maintained = false;
  c->vl_valid = false;
  c->have_gk = false;
  c->graph_ok = false;
  if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
  c->box_gen++;
  double old[3];
  for (int d = 0; d < 3; ++d) {
    old[d] = c->P.boxlen[d];
    c->P.boxlen[d] = boxlen[d];
  }
  int e = lists_setup(c);
  if (!e && c->P.list_mode == MSG_LIST_VERLET) e = verlet_setup(c);
  if (!e && c->have_ewald) e = msg_ewald_setup(h);
  if (!e) {
    c->npart_pair = pair_blocks(c);
    if (c->npart_pair > c->npart_pair_cap) e = set_error(MSG_ERR_ARG, "msg_set_box: pair partials exceed reserve");
  }
  if (e) {
    // back to the old box: a failed allocation left the old buffers in place and they hold
    // the old grid, so this set-up allocates nothing; the Ewald arrays may be half rewritten,
    // so the Ewald set-up must be done again (msg_ewald_setup or another msg_set_box)
    for (int d = 0; d < 3; ++d) c->P.boxlen[d] = old[d];
    lists_setup(c);
    if (c->P.list_mode == MSG_LIST_VERLET) verlet_setup(c);
    c->npart_pair = pair_blocks(c);
    c->have_ewald = false;
    return e;
  }
  if ((e = launch_zero(c))) return e;  // the counters of the new grid
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->have_gk = false;
  return MSG_OK;
}

int msg_nlist(msg_ctx* h) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_nlist: null context");
  Ctx* c = &h->c;
  if (!c->have_positions) return set_error(MSG_ERR_STATE, "positions have not been set");
  int e;
  if ((e = launch_bin(c)) || (e = launch_sort(c, 3)) || (e = launch_finalize(c))) return e;
  c->cells_valid = true;
  c->cells_maintained = false;  // SetCellList: the fresh order
  if (c->P.list_mode == MSG_LIST_VERLET) {
    if ((e = launch_vlist_build(c))) return e;
    c->vl_valid = true;
  }
  MSG_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if ((e = launch_zero(c))) return e;  // counters and flags back to zero for the next pass
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->fresh = false;
  return flags_to_error(c->h_flags[0]);
}

int msg_utotal_async(msg_ctx* h, double* d_force) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_utotal_async: null context");
  Ctx* c = &h->c;
  if (int e = check_ready_state(c)) return e;
  // the Verlet-list skin test is a host decision, so that mode launches eagerly
  const bool use_graph = !c->fp64 && !c->prof && c->P.list_mode != MSG_LIST_VERLET && !c->cells_maintained;
  if (use_graph) {
    if (!c->graph_ok) {
      if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
      if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
      MSG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int e = utotal_sequence(c);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (e) { if (g) hipGraphDestroy(g); return e; }
      if (ce != hipSuccess) return set_error(MSG_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
      c->graph = g;
      MSG_HIP(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
      c->graph_ok = true;
    }
    MSG_HIP(hipGraphLaunch(c->gexec, c->stream));
  } else {
    if (int e = utotal_sequence(c)) return e;
  }
  c->fresh = true;
  if (d_force && d_force != c->d_force)
    MSG_HIP(hipMemcpyAsync(d_force, c->d_force, 3 * c->np * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return MSG_OK;
}

int msg_fetch(msg_ctx* h, double* u_twob, double* u_rec, double* u_tot, double* virial) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_fetch: null context");
  Ctx* c = &h->c;
  MSG_HIP(hipMemcpyAsync(c->h_out, c->d_out, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  prof_drain(c);
  if (int e = flags_to_error((int)c->h_out[c->nptpt + 7])) return e;
  const int ntp = c->nptpt;
  if (u_twob)
    for (int k = 0; k <= ntp; ++k) u_twob[k] = c->h_out[k];
  if (u_rec) *u_rec = c->h_out[ntp + 1];
  if (u_tot) *u_tot = c->h_out[ntp + 2];
  if (virial) *virial = c->h_out[ntp + 3];
  return MSG_OK;
}

int msg_utotal(msg_ctx* h, double* u_twob, double* u_rec, double* u_tot, double* force,
               double* virial) {
  if (int e = msg_utotal_async(h, nullptr)) return e;
  Ctx* c = &h->c;
  if (force)
    MSG_HIP(hipMemcpyAsync(force, c->d_force, 3 * c->np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return msg_fetch(h, u_twob, u_rec, u_tot, virial);
}

int msg_dutotal(msg_ctx* h, int nptm, const int* ipnptm, const double* rotm, int* lhsoverlap,
                double* du_twob, double* du_rec, double* du_tot) {
  if (!h || nptm != 1 || !ipnptm || !rotm || !lhsoverlap)
    return set_error(MSG_ERR_ARG, "msg_dutotal: single-particle moves (nptm = 1) only");
  Ctx* c = &h->c;
  if (int e = check_ready_state(c)) return e;
  double dt, dr;
  std::vector<double> twob(c->nptpt + 1);
  int ov = 0;
  if (int e = dutotal_batch(c, 1, ipnptm, rotm, &ov, &dt, &dr, twob.data())) return e;
  *lhsoverlap = ov;
  if (du_twob)
    for (int k = 0; k <= c->nptpt; ++k) du_twob[k] = twob[k];
  if (du_rec) *du_rec = dr;
  if (du_tot) *du_tot = dt;
  return MSG_OK;
}

int msg_dutotal_charge(msg_ctx* h, int nptm, const int* ipnptm, int* lhsoverlap, double* du_twob,
                       double* du_rec, double* du_tot) {
  if (!h || !ipnptm || !lhsoverlap) return set_error(MSG_ERR_ARG, "msg_dutotal_charge: null argument");
  Ctx* c = &h->c;
  if (int e = check_ready_state(c)) return e;
  std::vector<double> twob(c->nptpt + 1);
  double dt = 0.0, dr = 0.0;
  int ov = 0;
  if (int e = dutotal_charge(c, nptm, ipnptm, &ov, twob.data(), &dr, &dt)) return e;
  *lhsoverlap = ov;
  if (du_twob)
    for (int k = 0; k <= c->nptpt; ++k) du_twob[k] = twob[k];
  if (du_rec) *du_rec = dr;
  if (du_tot) *du_tot = dt;
  return MSG_OK;
}

int msg_dutotal_batch(msg_ctx* h, int ntrial, const int* ip, const double* rotm, int* lhsoverlap,
                      double* du_tot, double* du_rec) {
  if (!h || ntrial < 1 || !ip || !rotm || !lhsoverlap || !du_tot)
    return set_error(MSG_ERR_ARG, "msg_dutotal_batch: bad arguments");
  Ctx* c = &h->c;
  if (int e = check_ready_state(c)) return e;
  return dutotal_batch(c, ntrial, ip, rotm, lhsoverlap, du_tot, du_rec, nullptr);
}

int msg_commit(msg_ctx* h, int nptm, const int* ipnptm, const double* rotm) {
  if (!h || nptm != 1 || !ipnptm || !rotm)
    return set_error(MSG_ERR_ARG, "msg_commit: single-particle moves (nptm = 1) only");
  Ctx* c = &h->c;
  if (ipnptm[0] < 1 || ipnptm[0] > c->np) return set_error(MSG_ERR_ARG, "msg_commit: ip out of range");
  return commit_move(c, ipnptm[0] - 1, rotm);
}

int msg_get_cell_list(msg_ctx* h, int* nc, int* cell_start, int* cell_ip) {
  if (!h || !nc) return set_error(MSG_ERR_ARG, "msg_get_cell_list: null argument");
  Ctx* c = &h->c;
  for (int d = 0; d < 3; ++d) nc[d] = c->grid.nc[d];
  if (int e = complete_cells(c)) return e;
  MSG_HIP(hipStreamSynchronize(c->stream));
  if (cell_start)
    MSG_HIP(hipMemcpy(cell_start, c->d_cell_start, (c->grid.ncell + 1) * sizeof(int), hipMemcpyDeviceToHost));
  if (cell_ip) {
    MSG_HIP(hipMemcpy(cell_ip, c->d_cell_ip, c->np * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->np; ++i) cell_ip[i] += 1;
  }
  return MSG_OK;
}

void* msg_stream(msg_ctx* h) { return h ? (void*)h->c.stream : nullptr; }

int msg_set_tile(msg_ctx* h, int pair_lanes) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_set_tile: null context");
  if (pair_lanes != 0 && pair_lanes != 4 && pair_lanes != 8 && pair_lanes != 16 && pair_lanes != 32 &&
      pair_lanes != 64 && pair_lanes != 108 && pair_lanes != 116)
    return set_error(MSG_ERR_ARG, "msg_set_tile: pair_lanes must be 0, 4, 8, 16, 32, 64, 108 or 116");
  Ctx* c = &h->c;
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->pair_lph = pair_lanes;
  c->npart_pair = pair_blocks(c);
  if (c->npart_pair > c->npart_pair_cap) return set_error(MSG_ERR_ARG, "msg_set_tile: tile exceeds reserve");
  c->graph_ok = false;
  return MSG_OK;
}

int msg_get_tile(msg_ctx* h) { return h ? h->c.pair_lph : MSG_ERR_ARG; }

int msg_get_forces(msg_ctx* h, double* force) {
  if (!h || !force) return set_error(MSG_ERR_ARG, "msg_get_forces: bad arguments");
  Ctx* c = &h->c;
  MSG_HIP(hipMemcpyAsync(force, c->d_force, 3 * c->np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

int msg_profile(msg_ctx* h, int enable) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_profile: null context");
  Ctx* c = &h->c;
  prof_drain(c);
  if (enable && !c->ev[0])  // timing events are created on first use
    for (auto& ev : c->ev)
      if (hipEventCreate(&ev) != hipSuccess) return set_error(MSG_ERR_HIP, "msg_profile: event creation failed");
  c->prof = enable != 0;
  for (int k = 0; k < MSG_K_COUNT; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return MSG_OK;
}

int msg_profile_read(msg_ctx* h, int kid, double* total_ms, int* launches) {
  if (!h || kid < 0 || kid >= MSG_K_COUNT) return set_error(MSG_ERR_ARG, "msg_profile_read: bad arguments");
  Ctx* c = &h->c;
  prof_drain(c);
  if (total_ms) *total_ms = c->prof_ms[kid];
  if (launches) *launches = c->prof_n[kid];
  return MSG_OK;
}

long long msg_device_bytes(msg_ctx* h) { return h ? h->c.bytes : 0; }

}  // extern "C"
