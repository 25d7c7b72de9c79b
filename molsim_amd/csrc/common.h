// Internal declarations of the MI355X non-bonded path (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/molsim_gpu.h"

namespace msg {

// ------------------------------------------------------------------------------------------
// Error handling: every failure ends in set_error(code, text) and a negative return value.
int set_error(int code, const std::string& text);

#define MSG_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::msg::set_error(MSG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device-side error flags (OR-ed into ctx->d_flags[0]; each pass's reduction hands them to the
// output block, clears word 0 and ORs them into the sticky word d_flags[1], which only the
// multi-step callers clear).
enum : int { FLAG_R2UMIN = 1, FLAG_IBUF = 2, FLAG_CELL = 4, FLAG_NLIST = 8 };

constexpr int kMaxT = MSG_MAX_TYPES;
constexpr int kMaxTP = kMaxT * (kMaxT + 1) / 2;
constexpr int kTabBins = 256;        // r^2 bins of the interval index per type pair
// Weak charges: the type field of an uncharged particle's cell-ordered record (cpos.w) carries
// this offset, so the pair kernels see the charge state in the record they load anyway.
constexpr int kUncharged = 8;
constexpr int kMaxTabIntervals = 512;   // records below rcut staged in LDS (<= 48 KB in FP64)
constexpr int kLdsTypes = 4;            // particle types whose table the LDS-staging kernels take
constexpr int kMaxTabIntervalsMany = 4096;  // records of a table of more types (read from memory)
constexpr int kMaxStencil = 27;
constexpr int kPairHB = 4;           // home particles per pair-kernel work item
// SPME bricks: each mesh point belongs to one brick; spread and gather work per brick.  Meshes
// of a power-of-two K in 16..256 with orders 3..8 take bricks of exactly bx x 8 x 8 points,
// bx = 32 for K >= 32 (fewer candidates per brick point: the stencil halo is shared by more
// points) and 16 for K = 16 (brick_x); other meshes and orders (the general path) split each
// dimension into nb bricks of floor(K / nb) or ceil(K / nb) points, nb = max(1, floor(K / kBG)),
// kBG >= order - 1, so a particle's stencil reaches only its own brick and the next one in
// every dimension.
constexpr int kBY = 8, kBZ = 8;
constexpr int brick_x(int K) { return K >= 32 ? 32 : 16; }
// The fused sort (every scatter workgroup scans all counts itself) is used while the counts
// number at most kLsMax and all workgroups' scans read at most kLsWork counts together.
constexpr int kLsMax = 8192;
constexpr long long kLsWork = 1ll << 22;
// Brick classes (brick_key) multiply the brick sort's counters by 8.  They pay where the counters
// stay few (at most 2048: the small meshes of batched ensembles, whose pass got 7 % shorter);
// on the 32 768-particle 128^3 pass they lengthened the fused sort's per-workgroup scans more
// than they shortened the spread, and at 1 M particles the longer scan and binning ate the
// spread's and the gather's gain.
inline bool brick_classes(int nbrick) { return 8ll * nbrick <= 2048; }
constexpr int kMinOrder = 2, kMaxOrder = 11;  // SPME orders (Src/potential.F90:359-360,3268)
constexpr int kMaxRadices = 16;               // FFT stages of the general path

inline int tp_index(int npt, int i, int j) {
  if (i > j) { int t = i; i = j; j = t; }
  return i * npt - i * (i - 1) / 2 + (j - i);
}

// ------------------------------------------------------------------------------------------
// Packed tabulated potential for the GPU (built from MOLSIM's ubuf on the host).
// Intervals of all type pairs are concatenated in the reference's scan order
// (decreasing zlow); per type pair: [first, last] interval range and a bin table over
// r^2 in [zlo, rcut^2] that gives the first interval to scan from.
struct GpuTableMeta {
  int first[kMaxTP], last[kMaxTP];
  double zlo[kMaxTP], invw[kMaxTP], r2atat[kMaxTP], r2umin[kMaxTP];
  int tpidx[kMaxT * kMaxT];
  int nint, ntp;
};

// Kernel argument blocks -------------------------------------------------------------------
struct Box {
  double L[3], half[3], inv[3];
};

struct CellGrid {
  int nc[3];
  int ncell;
  double csi[3];  // inverse cell size
  int nst;
  int st[3 * kMaxStencil];
};

struct MeshGrid {
  int K, H, order;
  int nb[3];      // bricks per dimension
  int nbrick;
  int bx;         // brick extent in x on the power-of-two path
  int ncls;       // classes per brick (8 on the power-of-two path, else 1), see brick_key
  int nkey;       // nbrick * ncls: counters of the brick sort
  double dkdr[3]; // K / L
  int gen;        // general path: any even K, orders 2..11 (runtime-K FFT, bricks of any extent)
  int qshift;     // FP32 spread: contributions are integers in units of 2^-qshift (30 - e, 2^e >= max |q|)
};

// Brick of mesh point m (0 <= m < K) along a dimension split into nb bricks whose first points
// are floor(b K / nb): the largest b with floor(b K / nb) <= m.  With K a multiple of nb this
// is m / (K / nb).
__host__ __device__ inline int brick_of(int m, int nb, int K) { return ((m + 1) * nb - 1) / K; }
__host__ __device__ inline int brick_index(const MeshGrid& g, int mx, int my, int mz) {
  return brick_of(mx, g.nb[0], g.K) + g.nb[0] * (brick_of(my, g.nb[1], g.K) + g.nb[1] * brick_of(mz, g.nb[2], g.K));
}
// Sort key of a particle with base point (mx, my, mz): its brick, and on the power-of-two path
// (bricks of bx x kBY x kBZ points) its class within the brick, bit d set when the base point
// lies within order - 1 points of the brick's low side in dimension d.  Only such particles
// have stencil points in the previous brick along d, so the spread of a brick reads, from each
// of its next bricks, just the classes that reach back into it.
__host__ __device__ inline int brick_key(const MeshGrid& g, int mx, int my, int mz) {
  const int b = brick_index(g, mx, my, mz);
  if (g.ncls == 1) return b;
  const int r = g.order - 1;
  const int cls = ((mx & (g.bx - 1)) < r ? 1 : 0) | ((my & (kBY - 1)) < r ? 2 : 0) | ((mz & (kBZ - 1)) < r ? 4 : 0);
  return b * 8 + cls;
}

enum KernelId : int;

// ------------------------------------------------------------------------------------------
struct Ctx {
  msg_params P{};
  int np = 0, npt = 0, nptpt = 0;
  bool fp64 = false;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;        // second branch of the UTotal graph (pair kernel)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_skin = nullptr;  // the skin test's (r2max1, r2max2) copy to the host has landed
  hipEvent_t ev_cells = nullptr; // a list rebuild's cell order is complete (its list build may follow on the side stream)
  hipEvent_t ev_batch = nullptr;     // orders the context's own work before a batched pass
  int device = 0;
  long long bytes = 0;

  // particles
  double* d_ro = nullptr;    // 3*np, caller order
  double* d_az = nullptr;    // np
  int* d_ipt = nullptr;      // np, 0-based
  bool have_particles = false, have_positions = false;
  double sumq2 = 0.0;

  // table
  bool have_table = false;
  GpuTableMeta tmeta{};
  void* d_tcoef = nullptr;   // nint * 12 reals (float or double)
  int* d_tbins = nullptr;    // ntp * kTabBins
  std::vector<double> h_ubuf;
  std::vector<int> h_iubuflow, h_nugrid;
  std::vector<double> h_r2umin, h_r2atat;

  // cell list
  CellGrid grid{};
  int* d_cellid = nullptr;       // np (per particle, caller order)
  int* d_cell_count = nullptr;   // ncell
  int* d_cell_start = nullptr;   // ncell+1
  int* d_cell_fill = nullptr;    // ncell
  int* d_cell_chunk = nullptr;   // ncell+1: offsets of the pair kernel's home chunks
  int* d_item_cell = nullptr;    // pair-kernel work item -> cell
  double* d_M = nullptr;         // total dipole sum q r of the current pass (3)
  int* d_cell_tmp = nullptr;     // np (unsorted members)
  int* d_rank = nullptr;         // 2*np: place of each particle in its cell, then in its brick
  // weak (titrating) charges: charge state per particle (1 charged), null when off; host
  // copies of the states, charges, full charges per type and types
  int* d_laz = nullptr;
  std::vector<int> h_laz, h_ipt;
  std::vector<double> h_az, h_zat;
  int* d_cell_ip = nullptr;      // np (sorted: cell-major, descending ip in a cell)
  int* d_slot_of = nullptr;      // np: slot of particle ip in the cell order
  void* d_cpos = nullptr;        // np double4 in cell order: x, y, z, particle type
  void* d_crel = nullptr;        // mixed path: np float4 in cell order, the position relative to
                                 // its cell's lower corner and the type (tag) as cpos
  int* d_ctype = nullptr;        // np
  int* d_slot_cell = nullptr;    // np

  // SPME
  bool have_ewald = false;
  MeshGrid mesh{};
  int* d_brick_count = nullptr, *d_brick_start = nullptr, *d_brick_fill = nullptr;
  int* d_brick_ip = nullptr;
  int* d_mm = nullptr;           // np, packed meshmax (caller order)
  void* d_bq = nullptr;          // np: (frac x, y, z, q) in brick order (real4)
  int* d_bmm = nullptr;          // np: packed meshmax in brick order
  void* d_frac = nullptr;        // np * 3 reals, caller order
  void* d_Q = nullptr;           // K^3 reals (charge mesh, then phi)
  void* d_C = nullptr;           // H*K*K complex
  void* d_G = nullptr;           // H^3 reals (meshfac)
  void* d_tw = nullptr;          // K complex twiddles exp(-2 pi i m / K)
  int nzblocks = 0;              // partial sums of the z/conv pass
  int nrad = 0;                  // general path: FFT radices, product K
  int rad[kMaxRadices] = {};

  // std Ewald
  int nkvec = 0;
  double* d_kfac = nullptr;      // nkvec
  int* d_kvec = nullptr;         // nkvec * 3
  double* d_sk = nullptr;        // nkvec * 4 * 2 (structure factors, 4 sign octants)

  // outputs and partials
  double* d_force = nullptr;     // 3*np (caller order)
  double* d_part = nullptr;      // partial sums (pair, spme, dipole)
  int npart_pair = 0, npart_dip = 0, part_off_pair = 0, part_off_spme = 0, part_off_dip = 0;
  int part_off_ewstd = 0, npart_ewstd = 0;
  int part_off_pairsum = 0;  // the pair rows summed into one row (SPME passes)
  double* d_out = nullptr;       // [u_twob(0..nptpt), u_rec, u_tot, virial, dip x,y,z]
  int* d_flags = nullptr;        // error flags
  double* h_out = nullptr;       // pinned
  int* h_flags = nullptr;        // pinned

  // graph of the utotal launch sequence
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  double* graph_force_dst = nullptr;

  // profiling
  bool prof = false;
  hipEvent_t ev[2 * MSG_K_COUNT * 64] = {};  // per-kernel timing events (created on first use)
  int ev_used = 0;
  std::vector<int> ev_kernel;
  double prof_ms[MSG_K_COUNT] = {0};
  int prof_n[MSG_K_COUNT] = {0};

  // Verlet list (MSG_LIST_VERLET): full list at rcut + drnlist built from the cell list;
  // home order = cell order at build ("slots"); neighbours are slots, column-major [maxn][np]
  int vl_maxn = 0;
  std::vector<int> h_de_ip;  // staged trial indices (host)
  int pair_lph = 32;        // pair tile: lanes per home particle (8..64), 0 = wave-item kernel,
                            // 100 + lanes (108, 116) = cell-tile kernel (mixed path, cell lists)
  int npart_pair_cap = 0;   // pair partial slots reserved (largest tile)
  int* d_vl_home = nullptr;
  int* d_vl_n = nullptr;
  int* d_vl_nbr = nullptr;
  double* d_vl_pos = nullptr;    // double4 per slot: position, type (+ kUncharged when uncharged)
  int* d_vl_rel = nullptr;       // mixed path, int4 per slot: position relative to the corner of its
                                 // cell at the build in units of cellsize / 2^28; w = that cell
                                 // (9 bits per dimension) | type tag << 27
  int* d_vl_slot = nullptr;      // slot of each particle
  double* d_ro_ref = nullptr;    // positions at the last build
  // NList skin test (Src/nlist.F90:597-621): per-particle squared displacement since the build
  // (d_r2, index order), per-256 block maxima and their exclusive prefix maxima, and the
  // result (r2max1, r2max2) as the bits of two non-negative doubles (pinned copy h_skin)
  double* d_r2 = nullptr;
  double* d_bm = nullptr;
  double* d_pm = nullptr;
  unsigned long long* d_skin = nullptr;
  unsigned long long* h_skin = nullptr;
  double* d_drosum = nullptr;    // 3*np: MD displacements summed since the build (drosum)
  bool drosum_reset = false;     // a list was built: the next drift starts the sums again
  int inlist = 0;                // 0: rebuild by the skin test; > 0: every inlist-th MD step
  int vl_width_user = 0;         // list width the caller set (MOLSIM's maxnneigh); 0: from the density
  long long istep2 = 0;          // MD steps of the current macrostep (NList's istep2)
  bool md_managed = false;       // msg_md_run decides rebuilds itself (at the end of a step)
  bool vl_valid = false;
  int ncell_cap = 0;             // cells the cell buffers hold
  int vl_cap = 0;                // list width vl_nbr holds
  long long box_gen = 0;         // bumped by msg_set_box (a batch planned before is stale)
  long long vl_builds = 0;

  // molecular dynamics (velocity Verlet)
  double* d_vel = nullptr;       // 3*np
  double* d_acc = nullptr;       // 3*np
  double* d_minv = nullptr;      // per type 1/mass
  double* d_ke_part = nullptr;
  bool md_ready = false;

  // dE: FP64 copy of the packed table, real-space influence kernel g (H^3, folded by
  // symmetry), per-trial inputs/outputs, and whether phi / S / cells / M match the positions
  double* d_tcoef64 = nullptr;
  double* d_gk = nullptr;
  bool have_gk = false;
  double* d_de = nullptr;        // per trial: [du_twob(0..ntp), du_rec, du_tot, overlap]
  int de_cap = 0;
  int* d_de_ip = nullptr;
  double* d_de_r = nullptr;
  bool fresh = false;            // a UTotal has run on the current positions
  bool cells_valid = false;      // the cell order (cell_ip, cpos, ...) matches the last pass
  bool cells_maintained = false; // accepted moves have patched the cell order (UpdateCellIp):
                                 // passes keep it instead of sorting the cells afresh
  int mc_next = 0;               // next particle of a sequential MC pass (isamp = 0)
};

// allocation helper (tracked in ctx->bytes)
int dalloc(Ctx* c, void** p, size_t bytes);

// Host-side table packing (table.cpp)
int pack_table(Ctx* c);

// Launch sequences (each .hip file)
int lists_setup(Ctx* c);                           // cell grid + buffers
int launch_bin(Ctx* c, bool cells = true, bool mesh = true);  // K1 cell id + mesh bin + counts + dipole
int launch_sort(Ctx* c, int which);                // scan + scatter: 1 cells, 2 bricks + dipole, 3 both
int launch_finalize(Ctx* c);                       // cell order (after launch_sort)
int launch_pair(Ctx* c, bool last);                // K4
int pair_blocks(Ctx* c);                           // workgroups of the pair kernel
bool pair_nbpar(const Ctx* c);                     // neighbour-parallel kernel in use
bool pair_tiled(const Ctx* c);                     // cell-tile kernel in use
int pair_tile_tx(const Ctx* c);                    // its cells per tile along x
bool pair_many(const Ctx* c);                      // more types than the LDS-table kernels take
int default_pair_tile(const Ctx* c);               // lanes per home for a new context
int spme_setup(Ctx* c);                            // K-space arrays
bool vl_rel_ok(const Ctx* c);                      // the Verlet pair loop takes the fixed-point records
constexpr double kVlUnits = 268435456.0;           // 2^28 fixed-point units per cell of the Verlet grid
void spme_charge_scale(Ctx* c);                    // MeshGrid::qshift from the charges
int launch_spme_mesh(Ctx* c);                      // spread, FFTs + influence function
int launch_spme_gather(Ctx* c, bool with_reduce = false);                    // force gather (after the pair kernel)
int ewstd_setup(Ctx* c);
int launch_ewstd(Ctx* c);
int launch_reduce(Ctx* c, bool presum);           // final energies (presum: pair rows pre-summed)
int launch_pair_sum(Ctx* c);                       // pair rows summed into one row
int launch_zero(Ctx* c);                           // counters and flags
int complete_cells(Ctx* c);                        // cell order of the current positions, if a pass left it out
int dutotal_batch(Ctx* c, int ntrial, const int* ip, const double* rotm, int* ovl,
                  double* du_tot, double* du_rec, double* du_twob);
int commit_move(Ctx* c, int ip, const double* rnew);   // MCUpdate: S(k), dipole, position, UpdateCellIp
int set_charge_tag(Ctx* c, int ip);                  // cell record tag after a charge-state flip
bool chain_eligible(const Ctx* c);                 // device-resident MC chain applies
int spme_trial_setup(Ctx* c);                      // SPME trials' real-space influence kernel g
int mc_chain_batch(const std::vector<Ctx*>& cs, hipStream_t s, const std::function<int()>& refresh, int ntrial,
                   int isamp, const double* dtran, double beta, int* iseed, int* rng_state, double* am,
                   double* du_sum, int* naccept, int* events);  // one chain per system, side by side
int mc_chain(Ctx* c, int ntrial, int isamp, const double* dtran, double beta, int* iseed,
             int* rng_state, double* am, double* du_sum, int* naccept, int* events);
int dutotal_launch(Ctx* c, int ntrial, const int* ip, const double* rotm);
int dutotal_collect(Ctx* c, int ntrial, int* ovl, double* du_tot, double* du_rec, double* du_twob);
int dutotal_charge(Ctx* c, int nptm, const int* ip1, int* ovl, double* du_twob, double* du_rec,
                   double* du_tot);
int utotal_refresh(Ctx* c);                        // UTotal sequence + sync (for dE state)
int check_ready_state(Ctx* c);                     // table, particles, positions, Ewald set up
int noalloc_setup(int device);                     // scratch address of msg_dutotal_batch_noalloc
int utotal_sequence(Ctx* c);
int verlet_setup(Ctx* c);
int verlet_check(Ctx* c, bool& rebuild);            // skin test of the current positions (host sync)
int launch_vl_records(Ctx* c);                     // Verlet slot records of the current positions
int verlet_check_begin(Ctx* c, bool& pending);       // its first half: queue the test and its copy
int verlet_check_end(Ctx* c, bool& rebuild);         // its second half: wait and decide
int launch_vlist_build(Ctx* c);
int launch_skin(Ctx* c);                           // r2max1 / r2max2 of d_r2 and d_bm, copied to h_skin
bool skin_decide(const Ctx* c);                    // sqrt(r2max1) + sqrt(r2max2) > drnlist (after ev_skin)
int launch_zero_counts(Ctx* c);                    // cell and brick counters (the flags kept)

// profiling wrappers around a launch
void prof_begin(Ctx* c, int kid);
void prof_end(Ctx* c, int kid);

// ------------------------------------------------------------------------------------------
// Segmented launches: one grid holds the workgroups of many systems.  map[b] = (system,
// workgroup within that system) and args[system] is that system's argument block, so a
// kernel body written for one system (argument block + workgroup index) serves both the
// single-system launch and the batched one.
struct SegGroup;
typedef void (*SegLaunchFn)(const SegGroup&, hipStream_t);
struct SegGroup {
  int nblk = 0;                  // workgroups of all systems
  size_t lds = 0;                // dynamic LDS bytes (the largest of the systems)
  const void* d_args = nullptr;  // per-system argument blocks
  const int2* d_map = nullptr;   // workgroup -> (system, local workgroup)
  SegLaunchFn launch = nullptr;
};

// Device memory of launch plans.  Default: each array is its own allocation, copied at
// once (plans built once).  Arena mode (plans built per launch): arrays are packed into a
// pinned host staging buffer and land in a device arena by one asynchronous copy (flush);
// the caller resets the arena only after the launches that read it have completed.
struct SegOwner {
  std::vector<void*> ptrs;
  char* arena = nullptr;   // device
  char* stage = nullptr;   // pinned host
  size_t cap = 0, used = 0;
  ~SegOwner() {
    for (void* p : ptrs) hipFree(p);
    if (arena) hipFree(arena);
    if (stage) hipHostFree(stage);
  }
  int init_arena(size_t bytes) {
    MSG_HIP(hipMalloc((void**)&arena, bytes));
    MSG_HIP(hipHostMalloc((void**)&stage, bytes, hipHostMallocDefault));
    cap = bytes;
    used = 0;
    return MSG_OK;
  }
  void reset() { used = 0; }
  // Grows the arena to at least `bytes` (only while nothing in flight reads it).
  int reserve(size_t bytes) {
    if (arena && bytes <= cap) return MSG_OK;
    if (arena) hipFree(arena);
    if (stage) hipHostFree(stage);
    arena = nullptr;
    stage = nullptr;
    return init_arena(bytes);
  }
  template <class T>
  int upload(const std::vector<T>& v, const T** out) {
    const size_t b = v.size() * sizeof(T);
    if (arena) {
      const size_t nb = (b + 255) & ~size_t(255);
      if (used + nb > cap) return set_error(MSG_ERR_NOMEM, "launch plan exceeds the unit's argument arena");
      std::copy((const char*)v.data(), (const char*)v.data() + b, stage + used);
      *out = (const T*)(arena + used);
      used += nb;
      return MSG_OK;
    }
    void* p = nullptr;
    MSG_HIP(hipMalloc(&p, std::max<size_t>(1, b)));
    ptrs.push_back(p);
    MSG_HIP(hipMemcpy(p, v.data(), b, hipMemcpyHostToDevice));
    *out = (const T*)p;
    return MSG_OK;
  }
  int flush(hipStream_t s) {
    if (arena && used) MSG_HIP(hipMemcpyAsync(arena, stage, used, hipMemcpyHostToDevice, s));
    return MSG_OK;
  }
};

// XCD-aware placement: the dispatcher hands workgroup b of a launch to XCD b % kXcd, and
// each XCD has its own L2.  A segmented launch therefore gives every system one XCD (the
// least loaded, systems taken largest first) and interleaves the per-XCD workgroup lists,
// so all of a system's workgroups share one L2 (the pair kernel re-reads each particle
// ~100 times).  Slots an XCD has no work for are map entries with system -1.
constexpr int kXcd = 8;

// Workgroup b of a single-system launch of n workgroups runs on XCD b mod 8 (round-robin
// dispatch); renumbered so that XCD x takes the x-th contiguous run of the logical workgroups,
// consecutive work items (spatially close in cell or brick order) share an L2 and the data
// they read around them is fetched into one L2 instead of eight.
__device__ inline int xcd_logical(int b, int n) {
  const int x = b % kXcd, k = b / kXcd, q = n / kXcd, r = n % kXcd;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Collects the argument blocks of one stage per template variant, then uploads them.
template <class A>
struct SegBuilder {
  struct G {
    std::vector<A> args;
    std::vector<int> nblk;
    size_t lds = 0;
    SegLaunchFn fn = nullptr;
  };
  std::map<int, G> groups;
  void add(int variant, SegLaunchFn fn, const A& a, int nblk, size_t lds = 0) {
    if (nblk <= 0) return;
    G& g = groups[variant];
    g.fn = fn;
    g.lds = std::max(g.lds, lds);
    g.args.push_back(a);
    g.nblk.push_back(nblk);
  }
  static std::vector<int2> place(const std::vector<int>& nblk) {
    const int n = (int)nblk.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nblk[x] > nblk[y]; });
    std::vector<std::vector<int2>> lists(kXcd);
    std::vector<long> load(kXcd, 0);
    for (int sys : order) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      for (int b = 0; b < nblk[sys]; ++b) lists[x].push_back(make_int2(sys, b));
      load[x] += nblk[sys];
    }
    size_t len = 0;
    for (auto& l : lists) len = std::max(len, l.size());
    std::vector<int2> map(len * kXcd, make_int2(-1, 0));
    for (int x = 0; x < kXcd; ++x)
      for (size_t k = 0; k < lists[x].size(); ++k) map[k * kXcd + x] = lists[x][k];
    return map;
  }
  int finish(SegOwner& own, std::vector<SegGroup>& out) {
    for (auto& kv : groups) {
      const std::vector<int2> map = place(kv.second.nblk);
      SegGroup sg;
      sg.nblk = (int)map.size();
      sg.lds = kv.second.lds;
      sg.launch = kv.second.fn;
      const A* da = nullptr;
      if (int e = own.upload(kv.second.args, &da)) return e;
      if (int e = own.upload(map, &sg.d_map)) return e;
      sg.d_args = da;
      out.push_back(sg);
    }
    return MSG_OK;
  }
};

// One stage of the batched UTotal sequence over many systems (each .hip file).
typedef std::vector<SegGroup> SegStage;
int plan_bin(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out);
int plan_sort(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& scan_c, SegStage& scan_b,
              SegStage& scatter_c, SegStage& scatter_b, SegStage& finalize);
int plan_pair(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out);
int plan_spme(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& spread, SegStage* fft,
              SegStage& gather);   // fft[5]: x, y, zconv, yi, xi
int plan_reduce(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out);
int plan_pair_sum(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out);

// A batched UTotal pass over many systems: its stages and the collection of every
// system's energies (64 doubles each) into d_outs.
struct Pass {
  std::vector<Ctx*> cs;
  SegStage bin, scan_c, scan_b, scatter_c, scatter_b, finalize, pair, pair_sum, spread, fft[5], gather, reduce;
  bool any_spme = false;
  const double* const* d_outp = nullptr;
  double* d_outs = nullptr;
};
int build_pass(const std::vector<Ctx*>& cs, SegOwner& own, Pass& p);
int launch_pass(const Pass& p, hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join);
int check_batchable(Ctx* c);

// Batched MC trials: systems grouped by spline order; d_args holds one block per member.
struct DeGroup {
  int order = 0;
  std::vector<int> sys;          // batch indices of the members
  const void* d_args = nullptr;
};
int plan_de(const std::vector<Ctx*>& cs, SegOwner& own, std::vector<DeGroup>& out);
int launch_de_seg(const DeGroup& g, const int* d_tstart, const int* d_ip, const double* d_r,
                  double* d_out, int ntot, hipStream_t s);

}  // namespace msg

struct msg_ctx {
  msg::Ctx c;
};

// ------------------------------------------------------------------------------------------
// Device helpers shared by the kernel files.
namespace msgd {

// Wave-local LDS hand-off (all lanes of one wavefront, no workgroup barrier).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cardinal B-spline values (and derivatives) of order n at fractional offset x in [0,1):
// vals[i] = M_n(x+i), ders[i] = M_n'(x+i) (Schoenberg recursion, Src/mollib.F90:1629-1683).
template <typename real, int N>
__device__ inline void bspline(real x, real* vals, real* ders) {
  real last[N];
  vals[0] = x;
  vals[1] = real(1) - x;
#pragma unroll
  for (int i = 2; i < N; ++i) vals[i] = real(0);
#pragma unroll
  for (int i = 0; i < N; ++i) last[i] = real(0);
#pragma unroll
  for (int m = 3; m <= N; ++m) {
#pragma unroll
    for (int i = 0; i < N; ++i) last[i] = vals[i];
    const real inv = real(1) / real(m - 1);
    vals[0] = x * last[0] * inv;
#pragma unroll
    for (int i = 1; i <= m - 2; ++i) {
      const real a = real(i) + x;
      vals[i] = (a * last[i] + (real(m) - a) * last[i - 1]) * inv;
    }
    vals[m - 1] = (real(1) - x) * last[m - 2] * inv;
  }
  if (ders) {
    ders[0] = last[0];
#pragma unroll
    for (int i = 1; i <= N - 2; ++i) ders[i] = last[i] - last[i - 1];
    ders[N - 1] = -last[N - 2];
  }
}

// The value of lane b ^ D, b = the lane's index (only its bit D matters): DPP within a row of
// 16 lanes (quad permutes for 1 and 2, row shifts for 4, a row rotation for 8), a swizzle for
// 16, a lane permute for 32.  row_shl:n reads lane + n, row_shr:n lane - n.  Every lane of
// the wavefront must be active.
template <int D>
__device__ __forceinline__ int xor_lanes(int v, int b) {
  if constexpr (D == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  if constexpr (D == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  if constexpr (D == 4) {
    const int up = __builtin_amdgcn_mov_dpp(v, 0x104, 0xF, 0xF, false);    // row_shl:4
    const int down = __builtin_amdgcn_mov_dpp(v, 0x114, 0xF, 0xF, false);  // row_shr:4
    return (b & 4) ? down : up;
  }
  if constexpr (D == 8) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  if constexpr (D == 16) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (16 << 10));  // bit mask mode: lane ^ 16
  static_assert(D > 0 && D < 64 && (D & (D - 1)) == 0, "xor_lanes: D is a power of two below 64");
  return __shfl_xor(v, 32, 64);
}
template <int D>
__device__ __forceinline__ float xor_lanes(float v, int b) { return __int_as_float(xor_lanes<D>(__float_as_int(v), b)); }
template <int D>
__device__ __forceinline__ double xor_lanes(double v, int b) {
  const long long u = __double_as_longlong(v);
  const int lo = xor_lanes<D>((int)(unsigned)u, b), hi = xor_lanes<D>((int)(u >> 32), b);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over aligned groups of L lanes (a power of two, at most 64), valid in every lane of the
// group: the butterfly over partners at distance L/2, ..., 2, 1 (the order of a shuffle
// butterfly), through xor_lanes.  b = the lane's index.
template <int L, typename T>
__device__ __forceinline__ T xor_sum(T v, int b) {
  if constexpr (L >= 64) v += xor_lanes<32>(v, b);
  if constexpr (L >= 32) v += xor_lanes<16>(v, b);
  if constexpr (L >= 16) v += xor_lanes<8>(v, b);
  if constexpr (L >= 8) v += xor_lanes<4>(v, b);
  if constexpr (L >= 4) v += xor_lanes<2>(v, b);
  if constexpr (L >= 2) v += xor_lanes<1>(v, b);
  return v;
}

// Inclusive prefix sum over a wavefront's 64 lanes by DPP (no LDS round trip): row shifts by
// 1, 2, 4, 8 scan each row of 16 lanes (a lane whose source lies in another row adds 0), then
// lane 15 of rows 0 and 2 is added to rows 1 and 3 (row_bcast:15) and lane 31 to rows 2 and 3
// (row_bcast:31).  Every lane of the wavefront must be active.
__device__ __forceinline__ int wave_incl_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1 and 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2 and 3
  return x;
}

// One DPP move of a double (both halves, the same lane pattern).
template <int CTL>
__device__ inline double dpp_f64(double x) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over a wavefront, valid in every lane, in a fixed order: within each row of 16 lanes by
// DPP rotations (8, 4, 2, 1), then the four row totals of lanes 0, 16, 32, 48 in that order.
// No LDS round trip and no wait between the steps.
__device__ inline double wave_sum_rows(double v) {
  v += dpp_f64<0x128>(v);  // row_ror:8
  v += dpp_f64<0x124>(v);  // row_ror:4
  v += dpp_f64<0x122>(v);  // row_ror:2
  v += dpp_f64<0x121>(v);  // row_ror:1
  const long long u = __double_as_longlong(v);
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, 16 * k);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 16 * k);
    r[k] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
  return ((r[0] + r[1]) + r[2]) + r[3];
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// FP64 sums take the DPP form.
__device__ inline double wave_sum(double v) { return wave_sum_rows(v); }

// Sums of two FP64 values over a workgroup of blockDim.x (a multiple of 64, at most 1024)
// threads, valid in thread 0: the same order as block_sum (wavefront sums, then the
// wavefronts in order) with one barrier for both.  sh holds 2 * 16 values; it must not be
// reused before a later barrier.
__device__ inline void block_sum2(double& a, double& b, double* sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w] = a;
    sh[16 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < nw; ++i) {
      sa += sh[i];
      sb += sh[16 + i];
    }
    a = sa;
    b = sb;
  }
}

// Sum over a workgroup of blockDim.x (multiple of 64) threads; result valid in thread 0.
template <typename T>
__device__ inline T block_sum(T v, T* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

}  // namespace msgd
