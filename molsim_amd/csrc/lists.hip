// K1: counting-sort cell list (and the SPME brick binning that shares its passes).
//
// MOLSIM keeps a doubly linked list per cell, filled by head insertion in particle order
// (SetCellList / AddIpToCell, Src/celllist.F90:307-327,384-410), so a traversal visits a
// cell's particles in DESCENDING index order.  Here the cell list is a counting sort:
//   k_bin       cell id (FP64 floor((ro + L/2) / cellsize), pcellro Src/celllist.F90:295-305)
//               and per-cell counts; the same pass bins particles by the SPME mesh brick of
//               their B-spline base point and forms per-block partial dipoles (UEwaldSurf);
//   k_scan      exclusive scan of the counts (one workgroup per array);
//   k_scatter   each particle takes a slot of its cell / brick;
//   k_finalize  one wavefront per cell ranks its members by index (descending, so the cell
//               order equals the reference's traversal order bit for bit) and writes the
//               cell-ordered position/charge/type arrays; bricks are ranked ascending.
#include <type_traits>

#include "common.h"

namespace msg {

namespace {

template <typename real>
__global__ __launch_bounds__(256) void k_bin(int np, const double* __restrict__ ro,
                                             const double* __restrict__ az, Box box, CellGrid g,
                                             int use_mesh, MeshGrid m, int* __restrict__ cellid,
                                             int* __restrict__ cell_count, int* __restrict__ mm,
                                             real* __restrict__ frac, int* __restrict__ brick_count,
                                             double* __restrict__ dip_part, int* __restrict__ flags) {
  __shared__ double sh[8];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double dx = 0.0, dy = 0.0, dz = 0.0;
  if (i < np) {
    const double r[3] = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
    const double q = az[i];
    int ci[3];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int k = (int)floor((r[d] + box.half[d]) * g.csi[d]);
      if (k < 0 || k > g.nc[d]) bad = true;
      ci[d] = min(max(k, 0), g.nc[d] - 1);
    }
    if (bad) atomicOr(flags, FLAG_CELL);
    const int c = ci[0] + g.nc[0] * (ci[1] + g.nc[1] * ci[2]);
    cellid[i] = c;
    atomicAdd(&cell_count[c], 1);
    if (use_mesh) {
      int mi[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double p = m.dkdr[d] * r[d];  // UEwaldRecSPM: dkdr * r, Src/energy.F90:1230
        const double v = floor(p);
        frac[3 * i + d] = (real)(p - v);
        int k = (int)v % m.K;
        mi[d] = k < 0 ? k + m.K : k;
      }
      mm[i] = mi[0] | (mi[1] << 10) | (mi[2] << 20);
      const int b = mi[0] / kBX + m.nb[0] * (mi[1] / kBY + m.nb[1] * (mi[2] / kBZ));
      atomicAdd(&brick_count[b], 1);
    }
    dx = q * r[0];
    dy = q * r[1];
    dz = q * r[2];
  }
  dx = msgd::block_sum(dx, sh);
  dy = msgd::block_sum(dy, sh);
  dz = msgd::block_sum(dz, sh);
  if (threadIdx.x == 0) {
    dip_part[3 * blockIdx.x + 0] = dx;
    dip_part[3 * blockIdx.x + 1] = dy;
    dip_part[3 * blockIdx.x + 2] = dz;
  }
}

// Exclusive scan over n values of one 1024-thread workgroup; out[n] receives the total.
// div > 1 scans ceil(cnt / div) instead of cnt.  Each thread owns a contiguous run of
// values, loaded together, so the scan costs two memory round trips, not one per 1024.
__device__ void block_scan(int n, const int* __restrict__ cnt, int div, int* out, int* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int per = (n + 1023) / 1024;
  const int b0 = threadIdx.x * per, b1 = min(n, b0 + per);
  int sum = 0;
  for (int i0 = b0; i0 < b1; i0 += 8) {  // eight independent loads in flight
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = i0 + u < b1 ? cnt[i0 + u] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) sum += div > 1 ? (v[u] + div - 1) / div : v[u];
  }
  int incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int before = 0, total = 0;
  for (int k = 0; k < 16; ++k) {
    if (k < w) before += wsum[k];
    total += wsum[k];
  }
  int run = before + incl - sum;
  for (int i0 = b0; i0 < b1; i0 += 8) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = i0 + u < b1 ? cnt[i0 + u] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i0 + u < b1) out[i0 + u] = run;
      run += div > 1 ? (v[u] + div - 1) / div : v[u];
    }
  }
  if (threadIdx.x == 0) out[n] = total;
  __syncthreads();
}

// Workgroup 0: cell offsets and the per-cell home-chunk offsets of the pair kernel;
// workgroup 1: brick offsets.  Both clear their fill counters.
__global__ __launch_bounds__(1024) void k_scan(int n0, const int* __restrict__ cnt0, int* start0,
                                              int* fill0, int hb, int* chunk0, int n1,
                                              const int* __restrict__ cnt1, int* start1,
                                              int* fill1, const double* __restrict__ dip_part,
                                              int ndip, double* __restrict__ M) {
  __shared__ int wsum[16];
  __shared__ double dsum[16][3];
  if (blockIdx.x == 2) {  // total dipole sum q r (surface term), from k_bin's partials
    double m[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < ndip; b += 1024)
#pragma unroll
      for (int d = 0; d < 3; ++d) m[d] += dip_part[3 * b + d];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t = msgd::wave_sum(m[d]);
      if ((threadIdx.x & 63) == 0) dsum[threadIdx.x >> 6][d] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < 16; ++w) t += dsum[w][threadIdx.x];
      M[threadIdx.x] = t;
    }
    return;
  }
  const int n = blockIdx.x == 0 ? n0 : n1;
  const int* cnt = blockIdx.x == 0 ? cnt0 : cnt1;
  int* fill = blockIdx.x == 0 ? fill0 : fill1;
  if (n <= 0) return;
  block_scan(n, cnt, 1, blockIdx.x == 0 ? start0 : start1, wsum);
  if (blockIdx.x == 0) block_scan(n, cnt, hb, chunk0, wsum);
  for (int i = threadIdx.x; i < n; i += 1024) fill[i] = 0;
}

__global__ __launch_bounds__(256) void k_scatter(int np, const int* __restrict__ cellid,
                                                 const int* __restrict__ cell_start, int* cell_fill,
                                                 int* __restrict__ cell_tmp, int use_mesh,
                                                 const int* __restrict__ mm, MeshGrid m,
                                                 const int* __restrict__ brick_start,
                                                 int* brick_fill, int* __restrict__ brick_tmp) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  const int c = cellid[i];
  cell_tmp[cell_start[c] + atomicAdd(&cell_fill[c], 1)] = i;
  if (use_mesh) {
    const int p = mm[i];
    const int b = (p & 1023) / kBX + m.nb[0] * (((p >> 10) & 1023) / kBY + m.nb[1] * ((p >> 20) / kBZ));
    brick_tmp[brick_start[b] + atomicAdd(&brick_fill[b], 1)] = i;
  }
}

struct FinalizeArgs {
  int ncell, nbrick;
  const int* cell_start;
  const int* cell_tmp;
  const int* cell_chunk;
  const double* ro;
  const double* az;
  const int* ipt;
  int* cell_ip;
  int* slot_of;
  double4* cpos;
  int* ctype;
  int* slot_cell;
  int* item_cell;
  const int* brick_start;
  const int* brick_tmp;
  int* brick_ip;
  const int* mm;
  const void* frac;
  void* bq;       // brick order: (frac x, frac y, frac z, q) reals
  int* bmm;       // brick order: packed meshmax
};

// One wavefront per cell (then per brick): rank members by particle index and write the
// ordered arrays; cells also list their pair-kernel work items.
template <typename real>
__global__ __launch_bounds__(256) void k_finalize(FinalizeArgs a) {
  const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wv < a.ncell) {
    const int s0 = a.cell_start[wv], n = a.cell_start[wv + 1] - s0;
    for (int k = lane; k < ((n + 63) & ~63); k += 64) {
      const int ip = k < n ? a.cell_tmp[s0 + k] : -1;
      int rank = 0;
      if (n <= 64) {  // members in registers: rank by broadcasting each one
        for (int j = 0; j < n; ++j) rank += __shfl(ip, j, 64) > ip;
      } else {
        for (int j = 0; j < n; ++j) rank += a.cell_tmp[s0 + j] > ip;
      }
      if (k >= n) continue;
      const int dst = s0 + rank;
      a.cell_ip[dst] = ip;
      a.slot_of[ip] = dst;
      a.cpos[dst] = make_double4(a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2], a.az[ip]);
      a.ctype[dst] = a.ipt[ip];
      a.slot_cell[dst] = wv;
    }
    const int i0 = a.cell_chunk[wv], ni = a.cell_chunk[wv + 1] - i0;
    for (int k = lane; k < ni; k += 64) a.item_cell[i0 + k] = wv;
  } else if (wv - a.ncell < a.nbrick) {
    const int b = wv - a.ncell;
    const int s0 = a.brick_start[b], n = a.brick_start[b + 1] - s0;
    const real* fr = (const real*)a.frac;
    typedef typename std::conditional<sizeof(real) == 4, float4, double4>::type R4;
    R4* bq = (R4*)a.bq;
    for (int k = lane; k < ((n + 63) & ~63); k += 64) {
      const int ip = k < n ? a.brick_tmp[s0 + k] : 0x7fffffff;
      int rank = 0;
      if (n <= 64) {
        for (int j = 0; j < n; ++j) rank += __shfl(ip, j, 64) < ip;
      } else {
        for (int j = 0; j < n; ++j) rank += a.brick_tmp[s0 + j] < ip;
      }
      if (k >= n) continue;
      const int dst = s0 + rank;
      a.brick_ip[dst] = ip;
      R4 v;
      v.x = fr[3 * ip];
      v.y = fr[3 * ip + 1];
      v.z = fr[3 * ip + 2];
      v.w = (real)a.az[ip];
      bq[dst] = v;
      a.bmm[dst] = a.mm[ip];
    }
  }
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
  if (int e = A((void**)&c->d_cellid, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_count, g.ncell * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_start, (g.ncell + 1) * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_fill, g.ncell * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_chunk, (g.ncell + 1) * sizeof(int))) return e;
  if (int e = A((void**)&c->d_M, 4 * sizeof(double))) return e;
  if (int e = A((void**)&c->d_item_cell, (g.ncell + np / kPairHB + 2) * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_tmp, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_cell_ip, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_slot_of, np * sizeof(int))) return e;
  if (int e = A(&c->d_cpos, np * sizeof(double4))) return e;
  if (int e = A((void**)&c->d_ctype, np * sizeof(int))) return e;
  if (int e = A((void**)&c->d_slot_cell, np * sizeof(int))) return e;
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

int launch_bin(Ctx* c) {
  const int np = c->np, nb = (np + 255) / 256;
  const bool mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  prof_begin(c, MSG_K_BIN);
  if (c->fp64)
    hipLaunchKernelGGL(k_bin<double>, dim3(nb), dim3(256), 0, c->stream, np, c->d_ro, c->d_az,
                       make_box(c->P), c->grid, (int)mesh, c->mesh, c->d_cellid, c->d_cell_count,
                       c->d_mm, (double*)c->d_frac, c->d_brick_count,
                       c->d_part + c->part_off_dip, c->d_flags);
  else
    hipLaunchKernelGGL(k_bin<float>, dim3(nb), dim3(256), 0, c->stream, np, c->d_ro, c->d_az,
                       make_box(c->P), c->grid, (int)mesh, c->mesh, c->d_cellid, c->d_cell_count,
                       c->d_mm, (float*)c->d_frac, c->d_brick_count,
                       c->d_part + c->part_off_dip, c->d_flags);
  prof_end(c, MSG_K_BIN);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int launch_sort(Ctx* c) {
  const int np = c->np;
  const bool mesh = c->P.ewald_mode == MSG_EWALD_SPM;
  const int nbrick = mesh ? c->mesh.nbrick : 0;
  prof_begin(c, MSG_K_CELLSORT);
  hipLaunchKernelGGL(k_scan, dim3(3), dim3(1024), 0, c->stream, c->grid.ncell, c->d_cell_count,
                     c->d_cell_start, c->d_cell_fill, kPairHB, c->d_cell_chunk, nbrick,
                     c->d_brick_count, c->d_brick_start, c->d_brick_fill,
                     c->d_part + c->part_off_dip, c->npart_dip, c->d_M);
  hipLaunchKernelGGL(k_scatter, dim3((np + 255) / 256), dim3(256), 0, c->stream, np, c->d_cellid,
                     c->d_cell_start, c->d_cell_fill, c->d_cell_tmp, (int)mesh, c->d_mm, c->mesh,
                     c->d_brick_start, c->d_brick_fill, c->d_brick_tmp);
  const int waves = c->grid.ncell + nbrick;
  FinalizeArgs f;
  f.ncell = c->grid.ncell;
  f.nbrick = nbrick;
  f.cell_start = c->d_cell_start;
  f.cell_tmp = c->d_cell_tmp;
  f.cell_chunk = c->d_cell_chunk;
  f.ro = c->d_ro;
  f.az = c->d_az;
  f.ipt = c->d_ipt;
  f.cell_ip = c->d_cell_ip;
  f.slot_of = c->d_slot_of;
  f.cpos = (double4*)c->d_cpos;
  f.ctype = c->d_ctype;
  f.slot_cell = c->d_slot_cell;
  f.item_cell = c->d_item_cell;
  f.brick_start = c->d_brick_start;
  f.brick_tmp = c->d_brick_tmp;
  f.brick_ip = c->d_brick_ip;
  f.mm = c->d_mm;
  f.frac = c->d_frac;
  f.bq = c->d_bq;
  f.bmm = c->d_bmm;
  if (c->fp64)
    hipLaunchKernelGGL(k_finalize<double>, dim3((waves + 3) / 4), dim3(256), 0, c->stream, f);
  else
    hipLaunchKernelGGL(k_finalize<float>, dim3((waves + 3) / 4), dim3(256), 0, c->stream, f);
  prof_end(c, MSG_K_CELLSORT);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg
