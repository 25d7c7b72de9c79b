// K1: counting-sort cell list (and the SPME brick binning that shares its passes).
//
// MOLSIM keeps a doubly linked list per cell, filled by head insertion in particle order
// (SetCellList / AddIpToCell, Src/celllist.F90:307-327,384-410), so a traversal visits a
// cell's particles in DESCENDING index order.  Here the cell list is a counting sort:
//   k_bin       cell id (FP64 floor((ro + L/2) / cellsize), pcellro Src/celllist.F90:295-305)
//               and per-cell counts; the same pass bins particles by the SPME mesh brick of
//               their B-spline base point and forms per-block partial dipoles (UEwaldSurf);
//   k_scan      exclusive scan of the counts (one workgroup per array);
//   k_scatter   each particle takes a slot of its cell and of its brick, and writes the
//               brick-ordered spline data directly (no result depends on the order within
//               a brick);
//   k_finalize  one wavefront per cell ranks its members by index (descending, so the cell
//               order equals the reference's traversal order bit for bit) and writes the
//               cell-ordered position/type arrays.  It runs beside the mesh branch.
#include <type_traits>

#include "common.h"

namespace msg {

namespace {

struct BinArgs {
  int np;
  const double* ro;
  const double* az;
  Box box;
  CellGrid g;
  int use_cells;    // cell ids and counts (0: only the box check, for passes with no cell sort)
  int use_mesh;
  MeshGrid m;
  int* cellid;
  int* cell_count;
  int* mm;
  void* frac;       // real[3*np]
  int* brick_count;
  double* dip_part;
  int* flags;
  int* rank_cell;   // each particle's place in its cell / brick as the counters hand them
  int* rank_brick;  // out, for a role whose scatter takes slots from them (null where the
                    // scatter keeps its own counters: large grids, batched passes)
};

// Counter increments of a wavefront aggregated over runs of consecutive lanes with the same
// key (particles of a chain are consecutive in index and space, so they fill the same cell or
// brick in runs): the first lane of a run adds the run's length, and every lane's place is
// the value its run's first lane got plus its position in the run.  A wavefront without runs
// takes one increment per lane.  A lane with key < 0 adds nothing.  Every lane of the
// wavefront must take part.  Returns the lane's place when want is set (uniform), else 0.
__device__ __forceinline__ int run_add(int* cnt, int key, int lane, bool want) {
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long H = __ballot(head);
  if (H == ~0ull) {  // no runs in this wavefront
    if (key < 0) return 0;
    if (want) return atomicAdd(cnt + key, 1);
    atomicAdd(cnt + key, 1);
    return 0;
  }
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0 .. lane
  const int hl = 63 - __clzll(H & upto);                  // this run's first lane
  const unsigned long long above = H & ~upto;
  const int next = above ? __ffsll(above) - 1 : 64;       // the next run's first lane
  int base = 0;
  if (head && key >= 0) {
    if (want)
      base = atomicAdd(cnt + key, next - lane);
    else
      atomicAdd(cnt + key, next - lane);
  }
  if (!want) return 0;
  return __shfl(base, hl, 64) + (lane - hl);
}

template <typename real>
__device__ __forceinline__ void bin_body(const BinArgs& a, int bid) {
  __shared__ double sh[4][3];
  const int i = bid * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  double dx = 0.0, dy = 0.0, dz = 0.0;
  int ckey = -1, bkey = -1;  // the particle's cell and brick (-1: no particle)
  if (i < a.np) {
    const double r[3] = {a.ro[3 * i], a.ro[3 * i + 1], a.ro[3 * i + 2]};
    const double q = a.az[i];
    int ci[3];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int k = (int)floor((r[d] + a.box.half[d]) * a.g.csi[d]);
      if (k < 0 || k > a.g.nc[d]) bad = true;
      ci[d] = min(max(k, 0), a.g.nc[d] - 1);
    }
    if (bad) atomicOr(a.flags, FLAG_CELL);
    if (a.use_cells) {
      ckey = ci[0] + a.g.nc[0] * (ci[1] + a.g.nc[1] * ci[2]);
      a.cellid[i] = ckey;
    }
    if (a.use_mesh) {
      real* frac = (real*)a.frac;
      int mi[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double p = a.m.dkdr[d] * r[d];  // UEwaldRecSPM: dkdr * r, Src/energy.F90:1230
        const double v = floor(p);
        frac[3 * i + d] = (real)(p - v);
        int k = (int)v % a.m.K;
        mi[d] = k < 0 ? k + a.m.K : k;
      }
      a.mm[i] = mi[0] | (mi[1] << 10) | (mi[2] << 20);
      bkey = brick_key(a.m, mi[0], mi[1], mi[2]);
    }
    dx = q * r[0];
    dy = q * r[1];
    dz = q * r[2];
  }
  // the cell and brick counters (uniform conditions: the whole wavefront takes part)
  if (a.use_cells) {
    const int rk = run_add(a.cell_count, ckey, lane, a.rank_cell != nullptr);
    if (a.rank_cell && i < a.np) a.rank_cell[i] = rk;
  }
  if (a.use_mesh) {
    const int rk = run_add(a.brick_count, bkey, lane, a.rank_brick != nullptr);
    if (a.rank_brick && i < a.np) a.rank_brick[i] = rk;
  }
  if (!a.dip_part) return;  // the same for the whole grid
  // the workgroup's partial dipole: wavefront sums (DPP), then the four wavefronts in order
  dx = msgd::wave_sum_rows(dx);
  dy = msgd::wave_sum_rows(dy);
  dz = msgd::wave_sum_rows(dz);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w][0] = dx;
    sh[w][1] = dy;
    sh[w][2] = dz;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    a.dip_part[3 * bid + threadIdx.x] = ((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) + sh[3][threadIdx.x];
}

template <typename real>
__global__ __launch_bounds__(256) void k_bin(BinArgs a) { bin_body<real>(a, blockIdx.x); }

template <typename real>
__global__ __launch_bounds__(256) void k_bin_seg(const BinArgs* __restrict__ args,
                                                 const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  bin_body<real>(args[s.x], s.y);
}

// Exclusive scans of cnt[0..n) and of ceil(cnt / DIV) (out2, may be null) by one workgroup of
// NT threads, in tiles of NT * V values.  Thread t takes the values t, t + NT, ... of a tile
// (coalesced loads and stores), the wavefronts scan the V rows by DPP, one barrier
// exchanges the wavefront totals, and the running totals carry into the next tile.
// out1[n] / out2[n] get the totals.  out1 / out2 may point to LDS or global memory.
template <int NT, int V, int DIV>
__device__ void block_scan_tiled(int n, const int* __restrict__ cnt, int* out1, int* out2,
                                 int (*wsum)[NT / 64]) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int c1 = 0, c2 = 0;
  for (int base = 0; base < n; base += NT * V) {
    int v[V], i1[V], i2[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int idx = base + u * NT + threadIdx.x;
      v[u] = idx < n ? cnt[idx] : 0;
      i1[u] = v[u];
      i2[u] = (v[u] + DIV - 1) / DIV;
    }
#pragma unroll
    for (int u = 0; u < V; ++u) {
      i1[u] = msgd::wave_incl_scan(i1[u]);
      i2[u] = msgd::wave_incl_scan(i2[u]);
    }
    if (lane == 63)
#pragma unroll
      for (int u = 0; u < V; ++u) {
        wsum[u][w] = i1[u];
        wsum[V + u][w] = i2[u];
      }
    __syncthreads();
    int b1 = c1, b2 = c2;
#pragma unroll
    for (int u = 0; u < V; ++u) {
      int t1 = 0, t2 = 0, p1 = 0, p2 = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const int x = wsum[u][k], y = wsum[V + u][k];
        t1 += x;
        t2 += y;
        if (k < w) {
          p1 += x;
          p2 += y;
        }
      }
      const int idx = base + u * NT + threadIdx.x;
      if (idx < n) {
        out1[idx] = b1 + p1 + i1[u] - v[u];
        if (out2) out2[idx] = b2 + p2 + i2[u] - (v[u] + DIV - 1) / DIV;
      }
      b1 += t1;
      b2 += t2;
    }
    c1 = b1;
    c2 = b2;
    __syncthreads();  // the next tile rewrites wsum
  }
  if (threadIdx.x == 0) {
    out1[n] = c1;
    if (out2) out2[n] = c2;
  }
  __syncthreads();
}

constexpr int kScanV = 8;

// Total dipole M = sum of k_bin's per-block partials, in the order of a 1024-thread
// reduction whatever the workgroup size NT (virtual thread v = t + NT*k sums partials v,
// v + 1024, ...; the 16 virtual wavefront totals are added in order), so the single-system
// and the batched passes give the same bits.
template <int NT>
__device__ void dipole_sum(const double* __restrict__ part, int n, double* M, double (*dsum)[3]) {
  constexpr int R = 1024 / NT, NW = NT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x + NT * k; b < n; b += 1024)
#pragma unroll
      for (int d = 0; d < 3; ++d) m[d] += part[3 * b + d];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t = msgd::wave_sum(m[d]);
      if (lane == 0) dsum[w + NW * k][d] = t;
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < 16; ++v) t += dsum[v][threadIdx.x];
    M[threadIdx.x] = t;
  }
}

struct ScanArgs {
  int bid0;               // role of workgroup 0 of this launch (0 cells, 1 bricks, 2 dipole)
  int n0;                 // cells
  const int* cnt0;
  int* start0;
  int* fill0;
  int* chunk0;
  int n1;                 // brick keys (0 without SPME)
  const int* cnt1;
  int* start1;
  int* fill1;
  const double* dip_part;
  int ndip;
  double* M;
};

// Role 0: cell offsets and the per-cell home-chunk offsets of the pair kernel; role 1:
// brick offsets.  Both clear their fill counters.  Role 2: total dipole.  A launch covers
// roles bid0, bid0 + 1, ... (with SPME the cell role runs on the pair branch and the brick
// and dipole roles on the mesh branch).
__device__ __forceinline__ void scan_body(const ScanArgs& a, int bid) {
  __shared__ int wsum[2 * kScanV][16];
  __shared__ double dsum[16][3];
  bid += a.bid0;
  if (bid == 2) {  // total dipole sum q r (surface term), from k_bin's partials
    dipole_sum<1024>(a.dip_part, a.ndip, a.M, dsum);
    return;
  }
  const int n = bid == 0 ? a.n0 : a.n1;
  const int* cnt = bid == 0 ? a.cnt0 : a.cnt1;
  int* fill = bid == 0 ? a.fill0 : a.fill1;
  if (n <= 0) return;
  block_scan_tiled<1024, kScanV, kPairHB>(n, cnt, bid == 0 ? a.start0 : a.start1, bid == 0 ? a.chunk0 : nullptr,
                                          wsum);
  for (int i = threadIdx.x; i < n; i += 1024) fill[i] = 0;
}

__global__ __launch_bounds__(1024) void k_scan(ScanArgs a) { scan_body(a, blockIdx.x); }

__global__ __launch_bounds__(1024) void k_scan_seg(const ScanArgs* __restrict__ args,
                                                   const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  scan_body(args[s.x], s.y);
}

struct ScatterArgs {
  int np;
  int parts;      // 1: cell slots, 2: brick slots and brick-ordered spline data
  const int* cellid;
  const int* cell_start;
  int* cell_fill;
  int* cell_tmp;
  int use_mesh;
  const int* mm;
  MeshGrid m;
  const int* brick_start;
  int* brick_fill;
  int* brick_ip;
  const void* frac;
  const double* az;
  void* bq;       // brick order: (frac x, frac y, frac z, q) reals
  int* bmm;       // brick order: packed meshmax
};

// Each particle takes a slot of its cell (ranked later) and of its mesh brick.  The brick
// order within a brick is left as the atomics hand it out: spreading sums exact 64-bit
// fixed-point contributions and the gather works per particle, so no result depends on it,
// and the mesh branch does not wait for a ranking pass.
template <typename real>
__device__ __forceinline__ void scatter_body(const ScatterArgs& a, int bid) {
  // slots from the fill counters, increments aggregated over runs (the whole wavefront takes
  // part; lanes past the end add nothing)
  const int i = bid * 256 + threadIdx.x, lane = threadIdx.x & 63;
  const bool ok = i < a.np;
  if (a.parts & 1) {
    const int c = ok ? a.cellid[i] : -1;
    const int rk = run_add(a.cell_fill, c, lane, true);
    if (ok) a.cell_tmp[a.cell_start[c] + rk] = i;
  }
  if ((a.parts & 2) && a.use_mesh) {
    const int p = ok ? a.mm[i] : 0;
    const int b = ok ? brick_key(a.m, p & 1023, (p >> 10) & 1023, p >> 20) : -1;
    const int rk = run_add(a.brick_fill, b, lane, true);
    if (!ok) return;
    const int dst = a.brick_start[b] + rk;
    typedef typename std::conditional<sizeof(real) == 4, float4, double4>::type R4;
    const real* fr = (const real*)a.frac;
    R4 v;
    v.x = fr[3 * i];
    v.y = fr[3 * i + 1];
    v.z = fr[3 * i + 2];
    v.w = (real)a.az[i];
    ((R4*)a.bq)[dst] = v;
    a.bmm[dst] = p;
    a.brick_ip[dst] = i;
  }
}

template <typename real>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a) { scatter_body<real>(a, blockIdx.x); }

template <typename real>
__global__ __launch_bounds__(256) void k_scatter_seg(const ScatterArgs* __restrict__ args,
                                                     const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  scatter_body<real>(args[s.x], s.y);
}

// Scatter with the offsets scanned by the workgroup itself, for count arrays that fit in
// LDS: every workgroup scans the counts k_bin left (cells or bricks) into LDS; workgroup 0
// also publishes them (with the pair kernel's home-chunk offsets for cells, the total dipole
// for bricks); a particle's slot is its group's offset plus the rank k_bin's counter gave it.
// This replaces a separate scan launch and the scatter's own counters.
constexpr int kLsV = 8;

struct ScatterLsArgs {
  ScatterArgs s;
  int role;            // 1: cells, 2: bricks
  int n;               // ncell or nkey (<= kLsMax)
  const int* cnt;
  int* start;          // offsets, n + 1 (written by workgroup 0)
  int* chunk;          // cells: home-chunk offsets, n + 1 (written by workgroup 0)
  const int* rank;     // per particle (cell ranks, or brick ranks)
  const double* dip_part;
  int ndip;
  double* M;           // total dipole, formed by workgroup 0 when dipole is set
  int dipole;
};

template <typename real>
__global__ __launch_bounds__(256) void k_scatter_ls(ScatterLsArgs a) {
  __shared__ int s_start[kLsMax + 1];
  __shared__ int wsum[2 * kLsV][4];
  __shared__ double dsum[16][3];
  const bool first = blockIdx.x == 0;
  block_scan_tiled<256, kLsV, kPairHB>(a.n, a.cnt, s_start, (first && a.role == 1) ? a.chunk : nullptr, wsum);
  if (first) {
    for (int k = threadIdx.x; k <= a.n; k += 256) a.start[k] = s_start[k];
    if (a.dipole) dipole_sum<256>(a.dip_part, a.ndip, a.M, dsum);
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.s.np) return;
  const int rk = a.rank[i];
  if (a.role == 1) {
    a.s.cell_tmp[s_start[a.s.cellid[i]] + rk] = i;
  } else {
    const int p = a.s.mm[i];
    const int b = brick_key(a.s.m, p & 1023, (p >> 10) & 1023, p >> 20);
    const int dst = s_start[b] + rk;
    typedef typename std::conditional<sizeof(real) == 4, float4, double4>::type R4;
    const real* fr = (const real*)a.s.frac;
    R4 v;
    v.x = fr[3 * i];
    v.y = fr[3 * i + 1];
    v.z = fr[3 * i + 2];
    v.w = (real)a.s.az[i];
    ((R4*)a.s.bq)[dst] = v;
    a.s.bmm[dst] = p;
    a.s.brick_ip[dst] = i;
  }
}

struct FinalizeArgs {
  int ncell, np;
  const int* cell_start;
  const int* cell_tmp;
  const int* cell_chunk;
  const double* ro;
  const int* ipt;
  int* cell_ip;
  int* slot_of;
  double4* cpos;
  int* ctype;
  int* slot_cell;
  int* item_cell;
  const int* laz;  // weak charges (null when off): an uncharged particle's record type gets + 8
  float4* crel;    // mixed path: cell-relative FP32 records (null on the FP64 path)
  CellGrid g;
  double half[3];  // L / 2
};

// One wavefront per cell: rank members by particle index (descending, the reference's
// linked-list order) and write the cell-ordered arrays; cells also list their pair-kernel
// work items.  The members are compared 64 at a time (one coalesced load, then each member
// read from its lane into a scalar register: no LDS round trip);
// a cell that holds every particle (one cell for the whole box) has the closed-form rank
// np - 1 - ip.
__device__ __forceinline__ void finalize_body(const FinalizeArgs& a, int bid) {
  const int wv = bid * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wv < a.ncell) {
    const int s0 = a.cell_start[wv], n = a.cell_start[wv + 1] - s0;
    for (int k = lane; k < ((n + 63) & ~63); k += 64) {
      const int ip = k < n ? a.cell_tmp[s0 + k] : -1;
      int rank = 0;
      if (n == a.np) {
        rank = n - 1 - ip;
      } else {
        for (int j0 = 0; j0 < n; j0 += 64) {
          const int jv = j0 + lane < n ? a.cell_tmp[s0 + j0 + lane] : -1;
          const int m = min(64, n - j0);
          for (int j = 0; j < m; ++j) rank += __builtin_amdgcn_readlane(jv, j) > ip;
        }
      }
      if (k >= n) continue;
      const int dst = s0 + rank;
      a.cell_ip[dst] = ip;
      a.slot_of[ip] = dst;
      const int tw = a.ipt[ip] + ((a.laz && !a.laz[ip]) ? kUncharged : 0);
      a.cpos[dst] = make_double4(a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2], (double)tw);
      if (a.crel) {
        const int cc[3] = {wv % a.g.nc[0], (wv / a.g.nc[0]) % a.g.nc[1], wv / (a.g.nc[0] * a.g.nc[1])};
        float r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = (float)(a.ro[3 * ip + d] + a.half[d] - cc[d] / a.g.csi[d]);
        a.crel[dst] = make_float4(r[0], r[1], r[2], (float)tw);
      }
      a.ctype[dst] = a.ipt[ip];
      a.slot_cell[dst] = wv;
    }
    const int i0 = a.cell_chunk[wv], ni = a.cell_chunk[wv + 1] - i0;
    for (int k = lane; k < ni; k += 64) a.item_cell[i0 + k] = wv;
  }
}

__global__ __launch_bounds__(256) void k_finalize(FinalizeArgs a) { finalize_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_finalize_seg(const FinalizeArgs* __restrict__ args,
                                                      const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  finalize_body(args[s.x], s.y);
}

}  // namespace

// Cell grid of InitCellList (Src/celllist.F90:108-152): ncell = max(1, floor(L / rcell)),
// collapsed to one cell where 2*ceil(rcut/cellsize)+1 > ncell.  With cells >= rcut the
// stencil is every offset in {-1,0,1} of the non-collapsed dimensions, in the reference's
// order (heap sort by |dx|+|dy|+|dz|; the order only matters for summation order).
int lists_setup(Ctx* c) {
  const msg_params& P = c->P;
  CellGrid& g = c->grid;
  const double rcell = P.rcut + P.drnlist;
  if (!(rcell > 0.0)) return set_error(MSG_ERR_CELL, "rcell <= 0");
  int range[3];
  double cs[3];
  for (int d = 0; d < 3; ++d) {
    if (P.list_mode == MSG_LIST_ALL) {
      g.nc[d] = 1;
      cs[d] = P.boxlen[d];
      range[d] = 0;
      continue;
    }
    g.nc[d] = std::max(1, (int)std::floor(P.boxlen[d] * (1.0 / rcell)));
    cs[d] = P.boxlen[d] / g.nc[d];
    range[d] = (int)std::ceil(P.rcut * (1.0 / cs[d]));
    if (2 * range[d] + 1 > g.nc[d]) {
      g.nc[d] = 1;
      range[d] = 0;
      cs[d] = P.boxlen[d];
    }
    if (range[d] > 1) return set_error(MSG_ERR_CELL, "cell smaller than rcut");
  }
  for (int d = 0; d < 3; ++d) g.csi[d] = 1.0 / cs[d];
  g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
  // stencil: all offsets with range 0/1, stable-ordered by Manhattan distance
  std::vector<std::array<int, 3>> st;
  for (int man = 0; man <= 3; ++man)
    for (int ix = -range[0]; ix <= range[0]; ++ix)
      for (int iy = -range[1]; iy <= range[1]; ++iy)
        for (int iz = -range[2]; iz <= range[2]; ++iz)
          if (std::abs(ix) + std::abs(iy) + std::abs(iz) == man) st.push_back({ix, iy, iz});
  g.nst = (int)st.size();
  for (int s = 0; s < g.nst; ++s)
    for (int d = 0; d < 3; ++d) g.st[3 * s + d] = st[s][d];

  const int np = c->np;
  auto A = [&](void** p, size_t b) { return dalloc(c, p, b); };
  // per-cell buffers: (re)allocated when the grid outgrows them (a new box); the counters
  // of a new grid are cleared by the caller
  if (g.ncell > c->ncell_cap) {
    // the new buffers first: if one allocation fails, the old ones are still in place
    void** cellbuf[5] = {(void**)&c->d_cell_count, (void**)&c->d_cell_start, (void**)&c->d_cell_fill,
                         (void**)&c->d_cell_chunk, (void**)&c->d_item_cell};
    const size_t sz[5] = {g.ncell * sizeof(int), (g.ncell + 1) * sizeof(int), g.ncell * sizeof(int),
                          (g.ncell + 1) * sizeof(int), (g.ncell + np / kPairHB + 2) * sizeof(int)};
    void* fresh[5] = {};
    for (int k = 0; k < 5; ++k)
      if (int e = A(&fresh[k], sz[k])) {
        for (int q = 0; q < k; ++q) {
          hipFree(fresh[q]);
          c->bytes -= (long long)sz[q];
        }
        return e;
      }
    for (int k = 0; k < 5; ++k) {
      if (*cellbuf[k]) hipFree(*cellbuf[k]);
      *cellbuf[k] = fresh[k];
    }
    if (c->ncell_cap > 0)
      c->bytes -= (long long)(5 * c->ncell_cap + np / kPairHB + 4) * (long long)sizeof(int);
    c->ncell_cap = g.ncell;
  }
  if (c->d_cellid) return MSG_OK;  // the per-particle buffers exist already
  if (int e = A((void**)&c->d_cellid, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_M, 4 * sizeof(double))) return e;
  if (int e = A((void**)&c->d_cell_tmp, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_ip, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_slot_of, np * sizeof(int))) return e;
  if (int e = A(&c->d_cpos, np * sizeof(double4))) return e;
  if (!c->fp64)
    if (int e = A(&c->d_crel, np * sizeof(float4))) return e;
  if (int e = A((void**)&c->d_ctype, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_slot_cell, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_rank, 2 * (size_t)np * sizeof(int))) return e;
  return MSG_OK;
}

static Box make_box(const msg_params& P) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = P.boxlen[d];
    b.half[d] = 0.5 * P.boxlen[d];
    b.inv[d] = 1.0 / P.boxlen[d];
  }
  return b;
}

// Fused scatter (offsets scanned per workgroup, slots from k_bin's ranks) for a role whose
// counts fit in LDS (kLsMax) and whose per-workgroup scans (every workgroup reads all counts)
// stay below kLsWork count reads: 1 cells, 2 bricks.
static bool sort_fused(const Ctx* c, int role) {
  const long long n = role == 1 ? c->grid.ncell : c->mesh.nkey;
  return n <= kLsMax && n * ((c->np + 255) / 256) <= kLsWork;
}

static BinArgs bin_args(Ctx* c) {
  BinArgs a;
  a.np = c->np;
  a.ro = c->d_ro;
  a.az = c->d_az;
  a.box = make_box(c->P);
  a.g = c->grid;
  a.use_cells = 1;
  a.use_mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  a.m = c->mesh;
  a.cellid = c->d_cellid;
  a.cell_count = c->d_cell_count;
  a.mm = c->d_mm;
  a.frac = c->d_frac;
  a.brick_count = c->d_brick_count;
  a.dip_part = c->d_part + c->part_off_dip;
  a.flags = c->d_flags;
  a.rank_cell = sort_fused(c, 1) ? c->d_rank : nullptr;
  a.rank_brick = (a.use_mesh && sort_fused(c, 2)) ? c->d_rank + c->np : nullptr;
  return a;
}

static ScanArgs scan_args(Ctx* c, int bid0) {
  const bool mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  ScanArgs a;
  a.bid0 = bid0;
  a.n0 = c->grid.ncell;
  a.cnt0 = c->d_cell_count;
  a.start0 = c->d_cell_start;
  a.fill0 = c->d_cell_fill;
  a.chunk0 = c->d_cell_chunk;
  a.n1 = mesh ? c->mesh.nkey : 0;
  a.cnt1 = c->d_brick_count;
  a.start1 = c->d_brick_start;
  a.fill1 = c->d_brick_fill;
  a.dip_part = c->d_part + c->part_off_dip;
  a.ndip = c->npart_dip;
  a.M = c->d_M;
  return a;
}

static ScatterArgs scatter_args(Ctx* c, int parts) {
  ScatterArgs a;
  a.np = c->np;
  a.parts = parts;
  a.cellid = c->d_cellid;
  a.cell_start = c->d_cell_start;
  a.cell_fill = c->d_cell_fill;
  a.cell_tmp = c->d_cell_tmp;
  a.use_mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  a.mm = c->d_mm;
  a.m = c->mesh;
  a.brick_start = c->d_brick_start;
  a.brick_fill = c->d_brick_fill;
  a.brick_ip = c->d_brick_ip;
  a.frac = c->d_frac;
  a.az = c->d_az;
  a.bq = c->d_bq;
  a.bmm = c->d_bmm;
  return a;
}

static FinalizeArgs finalize_args(Ctx* c) {
  FinalizeArgs f;
  f.ncell = c->grid.ncell;
  f.np = c->np;
  f.cell_start = c->d_cell_start;
  f.cell_tmp = c->d_cell_tmp;
  f.cell_chunk = c->d_cell_chunk;
  f.ro = c->d_ro;
  f.ipt = c->d_ipt;
  f.cell_ip = c->d_cell_ip;
  f.slot_of = c->d_slot_of;
  f.cpos = (double4*)c->d_cpos;
  f.crel = (float4*)c->d_crel;
  f.g = c->grid;
  for (int d = 0; d < 3; ++d) f.half[d] = 0.5 * c->P.boxlen[d];
  f.ctype = c->d_ctype;
  f.slot_cell = c->d_slot_cell;
  f.item_cell = c->d_item_cell;
  f.laz = c->d_laz;
  return f;
}

static int finalize_blocks(const FinalizeArgs& f) { return (f.ncell + 3) / 4; }

int launch_bin(Ctx* c, bool cells, bool mesh) {
  const int nb = (c->np + 255) / 256;
  BinArgs a = bin_args(c);
  if (!cells) {
    a.use_cells = 0;
    a.rank_cell = nullptr;
  }
  if (!mesh) {  // cell ids and counts alone, for a pass whose bricks were binned before
    a.use_mesh = 0;
    a.rank_brick = nullptr;
    a.dip_part = nullptr;
  }
  prof_begin(c, MSG_K_BIN);
  if (c->fp64)
    hipLaunchKernelGGL(k_bin<double>, dim3(nb), dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_bin<float>, dim3(nb), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_BIN);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// Scan and scatter.  which = 1: cell slots (the cell order is completed by launch_finalize);
// 2: brick slots, brick-ordered spline data and the total dipole (the mesh branch may start
// after these); 3: both.
static void launch_scatter_ls(Ctx* c, int role, bool dipole) {
  ScatterLsArgs a;
  a.s = scatter_args(c, role);
  a.role = role;
  a.n = role == 1 ? c->grid.ncell : c->mesh.nkey;
  a.cnt = role == 1 ? c->d_cell_count : c->d_brick_count;
  a.start = role == 1 ? c->d_cell_start : c->d_brick_start;
  a.chunk = c->d_cell_chunk;
  a.rank = c->d_rank + (role == 1 ? 0 : c->np);  // as bin_args hands them to k_bin
  a.dip_part = c->d_part + c->part_off_dip;
  a.ndip = c->npart_dip;
  a.M = c->d_M;
  a.dipole = dipole;
  const int nb = (c->np + 255) / 256;
  if (c->fp64)
    hipLaunchKernelGGL(k_scatter_ls<double>, dim3(nb), dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_scatter_ls<float>, dim3(nb), dim3(256), 0, c->stream, a);
}

int launch_sort(Ctx* c, int which) {
  const bool mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  // per role: the fused scatter where the counts fit in LDS, else scan + scatter
  const bool fc = (which & 1) && sort_fused(c, 1);
  const bool fb = (which & 2) && mesh && sort_fused(c, 2);
  const bool oc = (which & 1) && !fc;
  const bool ob = (which & 2) && mesh && !fb;
  // the total dipole (which & 2) goes with the brick role; without a mesh, with the cell role
  const bool dip_fc = (which & 2) && !mesh && fc;
  const bool dip_old = (which & 2) && !fb && !dip_fc;
  if (fc || fb) {
    prof_begin(c, MSG_K_SCATTER);
    if (fc) launch_scatter_ls(c, 1, dip_fc);
    if (fb) launch_scatter_ls(c, 2, true);
    prof_end(c, MSG_K_SCATTER);
  }
  if (oc || ob || dip_old) {
    // scan roles: 0 cells, 1 bricks, 2 dipole; one launch over the contiguous range needed
    const int r0 = oc ? 0 : (ob ? 1 : 2), r1 = dip_old ? 2 : (ob ? 1 : 0);
    prof_begin(c, MSG_K_SCAN);
    hipLaunchKernelGGL(k_scan, dim3(r1 - r0 + 1), dim3(1024), 0, c->stream, scan_args(c, r0));
    prof_end(c, MSG_K_SCAN);
    const int parts = (oc ? 1 : 0) | (ob ? 2 : 0);
    if (parts) {
      const int np = c->np;
      prof_begin(c, MSG_K_SCATTER);
      if (c->fp64)
        hipLaunchKernelGGL(k_scatter<double>, dim3((np + 255) / 256), dim3(256), 0, c->stream, scatter_args(c, parts));
      else
        hipLaunchKernelGGL(k_scatter<float>, dim3((np + 255) / 256), dim3(256), 0, c->stream, scatter_args(c, parts));
      prof_end(c, MSG_K_SCATTER);
    }
  }
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int launch_finalize(Ctx* c) {
  const FinalizeArgs f = finalize_args(c);
  prof_begin(c, MSG_K_FINALIZE);
  hipLaunchKernelGGL(k_finalize, dim3(finalize_blocks(f)), dim3(256), 0, c->stream, f);
  prof_end(c, MSG_K_FINALIZE);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// ---- batched stages ------------------------------------------------------------------------
template <typename real>
static void seg_bin(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_seg<real>, dim3(g.nblk), dim3(256), 0, s, (const BinArgs*)g.d_args, g.d_map);
}
static void seg_scan(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_seg, dim3(g.nblk), dim3(1024), 0, s, (const ScanArgs*)g.d_args, g.d_map);
}
template <typename real>
static void seg_scatter(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter_seg<real>, dim3(g.nblk), dim3(256), 0, s, (const ScatterArgs*)g.d_args, g.d_map);
}
static void seg_finalize(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_seg, dim3(g.nblk), dim3(256), 0, s, (const FinalizeArgs*)g.d_args, g.d_map);
}

int plan_bin(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out) {
  SegBuilder<BinArgs> b;
  for (Ctx* c : cs) {
    BinArgs a = bin_args(c);
    a.rank_cell = a.rank_brick = nullptr;  // the batched scatter counts slots itself
    b.add(c->fp64, c->fp64 ? seg_bin<double> : seg_bin<float>, a, (c->np + 255) / 256);
  }
  return b.finish(own, out);
}

// Batched list stages: cell scan and scatter (the pair branch), brick/dipole scan and brick
// scatter (the mesh branch), cell ranking.
int plan_sort(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& scan_c, SegStage& scan_b,
              SegStage& scatter_c, SegStage& scatter_b, SegStage& finalize) {
  SegBuilder<ScanArgs> bsc, bsb;
  SegBuilder<ScatterArgs> bcc, bcb;
  SegBuilder<FinalizeArgs> bf;
  for (Ctx* c : cs) {
    const int nb = (c->np + 255) / 256;
    bsc.add(0, seg_scan, scan_args(c, 0), 1);
    bsb.add(0, seg_scan, scan_args(c, 1), 2);
    bcc.add(c->fp64, c->fp64 ? seg_scatter<double> : seg_scatter<float>, scatter_args(c, 1), nb);
    if (c->P.ewald_mode == MSG_EWALD_SPM)
      bcb.add(c->fp64, c->fp64 ? seg_scatter<double> : seg_scatter<float>, scatter_args(c, 2), nb);
    const FinalizeArgs f = finalize_args(c);
    bf.add(0, seg_finalize, f, finalize_blocks(f));
  }
  if (int e = bsc.finish(own, scan_c)) return e;
  if (int e = bsb.finish(own, scan_b)) return e;
  if (int e = bcc.finish(own, scatter_c)) return e;
  if (int e = bcb.finish(own, scatter_b)) return e;
  return bf.finish(own, finalize);
}

}  // namespace msg
