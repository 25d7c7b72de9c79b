// K4: real-space pair loop over MOLSIM's tabulated potential (UTwoBodyACellList,
// Src/energy.F90:669-777; UTwoBodyA, :445-544, for the all-pairs mode).
//
// Work items are (cell, chunk of 4 home particles); a wavefront takes an item, spreads the
// candidates of the 27 stencil cells over its lanes and wave-reduces the home forces.  Every
// pair is evaluated from both of its particles (no atomics, deterministic forces; energies
// and the virial are halved in the final reduction).  The packed table (records of 12 coefficients in the
// reference's scan order + an r^2 bin index per type pair) is staged in LDS.
// Per pair, as in the reference:
//   PBCr2 single shift (FP64 difference, Src/molaux.F90:149-184) -> r2 in the kernel's real;
//   r2 >= rcut2 (cell list) / r2 > rcut2 (all pairs): skipped;
//   r2 < r2atat: usum = fsum = 1e10 (hard-core emulation);
//   r2 < r2umin: error flag (the reference stops);
//   else scan records from the bin's start while r2 < zlow, then Horner in d = r2 - zlow.
#include "common.h"

namespace msg {

namespace {

struct PairArgs {
  int np;
  const double4* cpos;
  const float4* crel;   // mixed path: FP32 cell-relative records (the neighbour-parallel kernel reads these)
  int xcd;              // single-system launch: workgroups renumbered so each XCD takes a contiguous run
  int tile_tx;          // cell-tile kernel: cells per tile along x
  float csz[3];         // cell edge lengths (FP32 of 1 / csi)
  float ncinv[2];       // 1 / nc[0], 1 / (nc[0] * nc[1]) in FP32 (cell coordinates of a cell index)
  const int* ctype;
  const int* cell_ip;
  const int* cell_start;
  const int* cell_chunk;
  const int* item_cell;
  const int* slot_cell;
  CellGrid g;
  Box box;
  double rcut2;
  int cut_inclusive;
  const void* tcoef;
  const int* tbins;
  GpuTableMeta tm;
  double* force;      // 3*np caller order
  double* part;       // per block: ntp energies + virial
  int* flags;
  double scale;       // applied to the force when this kernel is the last writer
  int nblk;           // workgroups of this system (stride of the work-item loop)
};

// Copy of nbytes (a multiple of 16, both ends 16-byte aligned) from global memory to LDS by
// the 256 threads of a workgroup: 16-byte loads, eight per thread issued before any store,
// so the copy costs one memory round trip instead of one per element.  Every element of
// the register array is assigned on both sides of the bounds test (a conditionally filled
// array is demoted to scratch memory, with one wait per load).
__device__ inline void lds_copy16(void* dst, const void* src, int nbytes) {
  const int4* s = (const int4*)src;
  int4* d = (int4*)dst;
  const int n = nbytes >> 4;
  constexpr int B = 8;
  for (int base = threadIdx.x; base < n; base += 256 * B) {
    int4 v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = base + u * 256 < n ? s[base + u * 256] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (base + u * 256 < n) d[base + u * 256] = v[u];
  }
}

using msgd::wave_sync;

constexpr int kWaves = 4;  // wavefronts per workgroup; each works through its own items

// Work item = (cell, chunk of up to kPairHB home particles).  The wavefront spreads the
// candidates of the home cell's stencil over its 64 lanes; every lane evaluates its
// candidate against all homes of the chunk; home forces are wave-reduced at the end.
// Energies (per type pair) and the virial accumulate per lane over all items.
template <typename real, int NTP>
__device__ __forceinline__ void pair_items_body(const PairArgs& a, int bid) {
  extern __shared__ __align__(16) double s_dyn[];  // table records, then the bin index (sized at launch)
  __shared__ int s_nbs[kWaves][kMaxStencil + 1], s_pref[kWaves][kMaxStencil + 1];
  __shared__ real s_zlo[kMaxTP], s_invw[kMaxTP], s_r2at[kMaxTP], s_r2um[kMaxTP];
  __shared__ int s_last[kMaxTP], s_tp[kMaxT * kMaxT];
  __shared__ double s_red[kWaves][kMaxTP + 1];
  const int nint = a.tm.nint;
  real* s_coef = (real*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  {
    const real* src = (const real*)a.tcoef;
    lds_copy16(s_coef, src, nint * 12 * (int)sizeof(real));
    lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
    if (threadIdx.x < NTP) {
      const int tp = threadIdx.x;
      s_zlo[tp] = (real)a.tm.zlo[tp];
      s_invw[tp] = (real)a.tm.invw[tp];
      s_r2at[tp] = (real)a.tm.r2atat[tp];
      s_r2um[tp] = (real)a.tm.r2umin[tp];
      s_last[tp] = a.tm.last[tp];
    }
    if (threadIdx.x < kMaxT * kMaxT) s_tp[threadIdx.x] = a.tm.tpidx[threadIdx.x];
  }
  __syncthreads();

  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nitems = a.cell_chunk[a.g.ncell];
  const real rc2 = (real)a.rcut2;
  double et[NTP];
#pragma unroll
  for (int k = 0; k < NTP; ++k) et[k] = 0.0;
  double vir = 0.0;
  int err = 0;
  for (int item = bid * kWaves + wv; item < nitems; item += a.nblk * kWaves) {
    const int c = a.item_cell[item];
    const int h0 = a.cell_start[c] + (item - a.cell_chunk[c]) * kPairHB;
    const int nh = min(kPairHB, a.cell_start[c + 1] - h0);
    // stencil cells: lane k < nst records start and count, then a wave prefix sum
    const int cx = c % a.g.nc[0], cy = (c / a.g.nc[0]) % a.g.nc[1], cz = c / (a.g.nc[0] * a.g.nc[1]);
    int cnt = 0, st0 = 0;
    if (lane < a.g.nst) {
      int x = cx + a.g.st[3 * lane], y = cy + a.g.st[3 * lane + 1], z = cz + a.g.st[3 * lane + 2];
      x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
      y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
      z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
      const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
      st0 = a.cell_start[nb];
      cnt = a.cell_start[nb + 1] - st0;
    }
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    const int ncand = __shfl(incl, a.g.nst - 1, 64);
    if (lane < a.g.nst) {
      s_nbs[wv][lane] = st0;
      s_pref[wv][lane] = incl - cnt;
    }
    if (lane == 0) s_pref[wv][a.g.nst] = ncand;
    wave_sync();
    double4 hp[kPairHB];
    int ht[kPairHB];
#pragma unroll
    for (int h = 0; h < kPairHB; ++h) {
      const int s = h0 + min(h, nh - 1);
      hp[h] = a.cpos[s];
      ht[h] = (int)a.cpos[s].w;  // + kUncharged for an uncharged home (weak charges)
    }
    double f[kPairHB][3];
#pragma unroll
    for (int h = 0; h < kPairHB; ++h) f[h][0] = f[h][1] = f[h][2] = 0.0;
    for (int idx = lane; idx < ncand; idx += 64) {
      int k = 0;
      while (s_pref[wv][k + 1] <= idx) ++k;
      const int j = s_nbs[wv][k] + (idx - s_pref[wv][k]);
      const double4 pj = a.cpos[j];
      const int tj = (int)pj.w;
      if (tj >= kUncharged) continue;  // weak charges: uncharged partner
#pragma unroll
      for (int h = 0; h < kPairHB; ++h) {
        if (h >= nh || j == h0 + h || ht[h] >= kUncharged) continue;
        double d[3] = {hp[h].x - pj.x, hp[h].y - pj.y, hp[h].z - pj.z};
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (fabs(d[q]) > a.box.half[q]) d[q] -= copysign(a.box.L[q], d[q]);
        const real dx = (real)d[0], dy = (real)d[1], dz = (real)d[2];
        const real r2 = dx * dx + dy * dy + dz * dz;
        if (a.cut_inclusive ? (r2 > rc2) : !(r2 < rc2)) continue;
        const int tp = s_tp[ht[h] * kMaxT + tj];
        real us, fs;
        if (r2 < s_r2at[tp]) {
          us = real(1e10);
          fs = real(1e10);
        } else if (r2 < s_r2um[tp]) {
          err |= FLAG_R2UMIN;
          continue;
        } else {
          int b = (int)((r2 - s_zlo[tp]) * s_invw[tp]);
          b = min(max(b, 0), kTabBins - 1);
          int ti = s_bins[tp * kTabBins + b];
          const int last = s_last[tp];
          while (ti < last && r2 < s_coef[12 * ti]) ++ti;
          const real* t = s_coef + 12 * ti;
          const real dd = r2 - t[0];
          us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
          fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
        }
#pragma unroll
        for (int k2 = 0; k2 < NTP; ++k2) et[k2] += (k2 == tp) ? (double)us : 0.0;
        f[h][0] += (double)(fs * dx);
        f[h][1] += (double)(fs * dy);
        f[h][2] += (double)(fs * dz);
        vir -= (double)(fs * r2);
      }
    }
#pragma unroll
    for (int h = 0; h < kPairHB; ++h) {
      const double fx = msgd::wave_sum(f[h][0]), fy = msgd::wave_sum(f[h][1]), fz = msgd::wave_sum(f[h][2]);
      if (lane == 0 && h < nh) {
        const int ip = a.cell_ip[h0 + h];
        a.force[3 * ip + 0] = fx * a.scale;
        a.force[3 * ip + 1] = fy * a.scale;
        a.force[3 * ip + 2] = fz * a.scale;
      }
    }
    wave_sync();  // s_nbs / s_pref are rewritten by the next item
  }
  if (err) atomicOr(a.flags, err);
  // per-workgroup sums per type pair and the virial, fixed order
#pragma unroll
  for (int k = 0; k < NTP; ++k) {
    const double v = msgd::wave_sum(et[k]);
    if (lane == 0) s_red[wv][k] = v;
  }
  {
    const double v = msgd::wave_sum(vir);
    if (lane == 0) s_red[wv][NTP] = v;
  }
  __syncthreads();
  if (threadIdx.x <= NTP) {
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += s_red[w][threadIdx.x];
    a.part[bid * (a.tm.ntp + 1) + (threadIdx.x == NTP ? a.tm.ntp : threadIdx.x)] = s;
  }
}

template <typename real, int NTP>
__global__ __launch_bounds__(256) void k_pair(PairArgs a) { pair_items_body<real, NTP>(a, blockIdx.x); }

template <typename real, int NTP>
__global__ __launch_bounds__(256) void k_pair_seg(const PairArgs* __restrict__ args,
                                                  const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  pair_items_body<real, NTP>(args[s.x], s.y);
}

// Pair evaluation shared by the kernels: table lookup and accumulation for home h (type ht)
// and candidate pj (type tj).
template <typename real, int NTP>
struct PairAcc {
  // energies of one home particle's pairs by partner type (the home's type fixes the type
  // pair); to_tp() folds them into type-pair slots at the end
  static constexpr int NT = NTP == 1 ? 1 : NTP == 3 ? 2 : NTP == 6 ? 3 : 4;
  double et[NT];
  double vir, fx, fy, fz;
  int err;
  __device__ void zero() {
#pragma unroll
    for (int q = 0; q < NT; ++q) et[q] = 0.0;
    vir = fx = fy = fz = 0.0;
    err = 0;
  }
  __device__ double to_tp(int q, int ht, const int* s_tp) const {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) v += (s_tp[ht * kMaxT + t] == q) ? et[t] : 0.0;
    return v;
  }
};

template <typename real, int NTP>
__device__ inline void pair_eval(const double4& hp, int ht, const double4& pj, int tj, const Box& box,
                                 real rc2, bool inclusive, const real* s_coef, const int* s_bins,
                                 const real* s_zlo, const real* s_invw, const real* s_r2at,
                                 const real* s_r2um, const int* s_last, const int* s_tp,
                                 PairAcc<real, NTP>& acc) {
  double d[3] = {hp.x - pj.x, hp.y - pj.y, hp.z - pj.z};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (fabs(d[q]) > box.half[q]) d[q] -= copysign(box.L[q], d[q]);
  const real dx = (real)d[0], dy = (real)d[1], dz = (real)d[2];
  const real r2 = dx * dx + dy * dy + dz * dz;
  if (inclusive ? (r2 > rc2) : !(r2 < rc2)) return;
  if (tj >= kUncharged) return;  // weak charges: no pair energy with an uncharged partner (UWeakChargeA)
  const int tp = s_tp[ht * kMaxT + tj];
  real us, fs;
  if (r2 < s_r2at[tp]) {
    us = real(1e10);
    fs = real(1e10);
  } else if (r2 < s_r2um[tp]) {
    acc.err |= FLAG_R2UMIN;
    return;
  } else {
    int b = (int)((r2 - s_zlo[tp]) * s_invw[tp]);
    b = min(max(b, 0), kTabBins - 1);
    int ti = s_bins[tp * kTabBins + b];
    const int last = s_last[tp];
    while (ti < last && r2 < s_coef[12 * ti]) ++ti;
    const real* t = s_coef + 12 * ti;
    const real dd = r2 - t[0];
    us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
    fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
  }
#pragma unroll
  for (int t = 0; t < PairAcc<real, NTP>::NT; ++t) acc.et[t] += (t == tj) ? (double)us : 0.0;
  acc.fx += (double)(fs * dx);
  acc.fy += (double)(fs * dy);
  acc.fz += (double)(fs * dz);
  acc.vir -= (double)(fs * r2);
}

// Table metadata of one (home type, partner type) combination, read with one LDS access:
// the bin range and overlap/table limits of its type pair, the type pair and its last
// interval.
template <typename real>
struct alignas(4 * sizeof(real)) PairMeta {
  real zlo, invw, r2at, r2um;
};

// pair_eval with the metadata of the combination in one record (s_pm / s_pi indexed by
// ht * kMaxT + tj) and the interval's 12 coefficients read whole before the interval
// search checks them (the search advances for a few per cent of pairs): three dependent LDS
// accesses per pair instead of six.  Same arithmetic as pair_eval.
template <typename real, int NTP>
__device__ inline void pair_eval_m(const double4& hp, int ht, const double4& pj, int tj, const Box& box,
                                   real rc2, bool inclusive, const real* s_coef, const int* s_bins,
                                   const PairMeta<real>* s_pm, const int2* s_pi, PairAcc<real, NTP>& acc) {
  typedef typename std::conditional<sizeof(real) == 4, float4, double2>::type V16;
  constexpr int NV = 12 * (int)sizeof(real) / 16;
  double d[3] = {hp.x - pj.x, hp.y - pj.y, hp.z - pj.z};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (fabs(d[q]) > box.half[q]) d[q] -= copysign(box.L[q], d[q]);
  const real dx = (real)d[0], dy = (real)d[1], dz = (real)d[2];
  const real r2 = dx * dx + dy * dy + dz * dz;
  if (inclusive ? (r2 > rc2) : !(r2 < rc2)) return;
  if (tj >= kUncharged) return;  // weak charges: no pair energy with an uncharged partner (UWeakChargeA)
  const int mi = ht * kMaxT + tj;
  const PairMeta<real> pm = s_pm[mi];
  const int2 pi = s_pi[mi];
  real us, fs;
  if (r2 < pm.r2at) {
    us = real(1e10);
    fs = real(1e10);
  } else if (r2 < pm.r2um) {
    acc.err |= FLAG_R2UMIN;
    return;
  } else {
    int b = (int)((r2 - pm.zlo) * pm.invw);
    b = min(max(b, 0), kTabBins - 1);
    int ti = s_bins[pi.x * kTabBins + b];
    alignas(16) real t[12];
    {
      const V16* rp = (const V16*)(s_coef + 12 * ti);
#pragma unroll
      for (int k = 0; k < NV; ++k) *(V16*)(t + k * 16 / (int)sizeof(real)) = rp[k];
    }
    while (ti < pi.y && r2 < t[0]) {
      ++ti;
      const V16* rp = (const V16*)(s_coef + 12 * ti);
#pragma unroll
      for (int k = 0; k < NV; ++k) *(V16*)(t + k * 16 / (int)sizeof(real)) = rp[k];
    }
    const real dd = r2 - t[0];
    us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
    fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
  }
#pragma unroll
  for (int t = 0; t < PairAcc<real, NTP>::NT; ++t) acc.et[t] += (t == tj) ? (double)us : 0.0;
  acc.fx += (double)(fs * dx);
  acc.fy += (double)(fs * dy);
  acc.fz += (double)(fs * dz);
  acc.vir -= (double)(fs * r2);
}

// Fills the combination records (threads 0 .. kMaxT^2 - 1) and the type-pair index.
template <typename real>
__device__ inline void fill_pair_meta(const GpuTableMeta& tm, PairMeta<real>* s_pm, int2* s_pi, int* s_tp) {
  if (threadIdx.x < kMaxT * kMaxT) {
    const int tp = tm.tpidx[threadIdx.x];
    s_tp[threadIdx.x] = tp;
    PairMeta<real> m;
    m.zlo = (real)tm.zlo[tp];
    m.invw = (real)tm.invw[tp];
    m.r2at = (real)tm.r2atat[tp];
    m.r2um = (real)tm.r2umin[tp];
    s_pm[threadIdx.x] = m;
    s_pi[threadIdx.x] = make_int2(tp, tm.last[tp]);
  }
}

// Neighbour-parallel variant for the full 27-cell stencil: LPH lanes per home particle (the
// tile shape, 8..64, chosen per system size by the autotuner; 256 / LPH homes per workgroup).
// The candidates of the 27 stencil cells are split into LPH contiguous, equal shares (so dense
// cells do not serialise one lane); each lane walks its share two candidates at a time with
// both loads issued first.  The group reduces the home force with shuffles.
template <typename real, int NTP, int LPH>
__device__ __forceinline__ void pair_nb_body(const PairArgs& a, int bid) {
  constexpr int HPB = 256 / LPH;                          // homes per workgroup
  constexpr int MPL = (kMaxStencil + LPH - 1) / LPH;      // stencil cells per lane
  extern __shared__ __align__(16) double s_dyn[];
  __shared__ PairMeta<real> s_pm[kMaxT * kMaxT];
  __shared__ int2 s_pi[kMaxT * kMaxT];
  __shared__ int s_tp[kMaxT * kMaxT];
  __shared__ double s_red[4][kMaxTP + 1];
  __shared__ int s_cst[HPB][kMaxStencil + 1], s_cpf[HPB][kMaxStencil + 1];
  const int nint = a.tm.nint;
  real* s_coef = (real*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  {
    const real* src = (const real*)a.tcoef;
    lds_copy16(s_coef, src, nint * 12 * (int)sizeof(real));
    lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
    fill_pair_meta<real>(a.tm, s_pm, s_pi, s_tp);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = threadIdx.x / LPH;
  const int s = bid * HPB + grp;
  const int k = threadIdx.x & (LPH - 1);
  const bool home = s < a.np;
  const int hs = home ? s : a.np - 1;
  const double4 hp = a.cpos[hs];
  const int hw = (int)hp.w;
  const int ht = hw & (kUncharged - 1);
  const bool hcharged = hw < kUncharged;  // an uncharged home has no pair energy (weak charges)
  const int c = a.slot_cell[hs];
  const int cx = c % a.g.nc[0], cy = (c / a.g.nc[0]) % a.g.nc[1], cz = c / (a.g.nc[0] * a.g.nc[1]);
  // home position inside its cell, for skipping stencil cells whose nearest point is
  // beyond the cutoff (every candidate in them would fail the cutoff test anyway)
  const int cc3[3] = {cx, cy, cz};
  const double hh[3] = {hp.x, hp.y, hp.z};
  double hrel[3], csz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    csz[d] = a.box.L[d] / a.g.nc[d];
    hrel[d] = (hh[d] + a.box.half[d]) - cc3[d] * csz[d];
  }
  const double cull2 = a.rcut2 * (1.0 + 1e-5);
  int st0[MPL], cnt[MPL], tot = 0;
#pragma unroll
  for (int m = 0; m < MPL; ++m) {
    const int sc = k * MPL + m;
    cnt[m] = 0;
    st0[m] = 0;
    double bd2 = 0.0;
    if (sc < a.g.nst) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int o = a.g.st[3 * sc + d];
        const double dist = o < 0 ? hrel[d] : (o > 0 ? csz[d] - hrel[d] : 0.0);
        bd2 += dist > 0.0 ? dist * dist : 0.0;
      }
    }
    if (sc < a.g.nst && bd2 <= cull2) {
      int x = cx + a.g.st[3 * sc], y = cy + a.g.st[3 * sc + 1], z = cz + a.g.st[3 * sc + 2];
      x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
      y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
      z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
      const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
      st0[m] = a.cell_start[nb];
      cnt[m] = a.cell_start[nb + 1] - st0[m];
    }
    tot += cnt[m];
  }
  int incl = tot;
#pragma unroll
  for (int off = 1; off < LPH; off <<= 1) {
    const int t = __shfl_up(incl, off, LPH);
    if (k >= off) incl += t;
  }
  const int ncand = __shfl(incl, LPH - 1, LPH);
  {
    int pre = incl - tot;
#pragma unroll
    for (int m = 0; m < MPL; ++m) {
      const int sc = k * MPL + m;
      if (sc < a.g.nst) {
        s_cst[grp][sc] = st0[m];
        s_cpf[grp][sc] = pre;
      }
      pre += cnt[m];
    }
  }
  if (k == 0) s_cpf[grp][a.g.nst] = ncand;
  __syncthreads();  // table in LDS and the per-group cell ranges

  const real rc2 = (real)a.rcut2;
  PairAcc<real, NTP> acc;
  acc.zero();
  const int share = (ncand + LPH - 1) / LPH;
  int i = k * share;
  const int i1 = (home && hcharged) ? min(ncand, i + share) : 0;
  if (i < i1) {
    int cc = 0;
    while (s_cpf[grp][cc + 1] <= i) ++cc;
    int j = s_cst[grp][cc] + (i - s_cpf[grp][cc]);
    int cend = s_cst[grp][cc] + (s_cpf[grp][cc + 1] - s_cpf[grp][cc]);
    constexpr int U = 4;  // candidates in flight per lane
    while (i < i1) {
      int jv[U];
      bool hv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        hv[u] = i < i1;
        jv[u] = hv[u] ? j : s;
        if (hv[u]) {
          ++i;
          ++j;
          if (j == cend && i < i1) {
            do { ++cc; } while (s_cpf[grp][cc + 1] <= i);
            j = s_cst[grp][cc];
            cend = j + (s_cpf[grp][cc + 1] - s_cpf[grp][cc]);
          }
        }
      }
      double4 pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) pv[u] = a.cpos[jv[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (hv[u] && jv[u] != s)
          pair_eval_m<real, NTP>(hp, ht, pv[u], (int)pv[u].w, a.box, rc2, a.cut_inclusive, s_coef, s_bins,
                                 s_pm, s_pi, acc);
    }
  }
  acc.fx = msgd::xor_sum<LPH>(acc.fx, threadIdx.x);
  acc.fy = msgd::xor_sum<LPH>(acc.fy, threadIdx.x);
  acc.fz = msgd::xor_sum<LPH>(acc.fz, threadIdx.x);
  if (home && k == 0) {
    const int ip = a.cell_ip[s];
    a.force[3 * ip + 0] = acc.fx * a.scale;
    a.force[3 * ip + 1] = acc.fy * a.scale;
    a.force[3 * ip + 2] = acc.fz * a.scale;
  }
  if (acc.err) atomicOr(a.flags, acc.err);
#pragma unroll
  for (int q = 0; q < NTP; ++q) {
    const double v = msgd::wave_sum(acc.to_tp(q, ht, s_tp));
    if (lane == 0) s_red[wv][q] = v;
  }
  {
    const double v = msgd::wave_sum(acc.vir);
    if (lane == 0) s_red[wv][NTP] = v;
  }
  __syncthreads();
  if (threadIdx.x <= NTP) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][threadIdx.x];
    a.part[bid * (a.tm.ntp + 1) + (threadIdx.x == NTP ? a.tm.ntp : threadIdx.x)] = t;
  }
}


constexpr int kNbRound = 4;    // candidates per lane and test round (loads in flight)
constexpr int kNbEval = 4;     // in-cut pairs per lane evaluated together (loads in flight)
constexpr int kNbSlotBits = 27;  // a code (unsigned) holds a slot below 2^27 and a stencil index above

// Quotient of 0 <= x < 2^24 by d >= 1 through the FP32 reciprocal dinv ~ 1 / d: the product
// is off by at most one, and one correction step makes it exact.
__device__ __forceinline__ int div_small(int x, int d, float dinv) {
  int q = (int)((float)x * dinv);
  const int r = x - q * d;
  q += r < 0 ? -1 : (r >= d ? 1 : 0);
  return q;
}

// Inclusive prefix sum and total of x over aligned segments of L = 8 or 16 lanes (inside a row
// of 16 lanes), with DPP row shifts: no LDS round trip.  k = lane index within the segment.
template <int L, int OFF>
__device__ __forceinline__ void seg_scan_step(int& inc, int& suf, int k) {
  // row_shr:OFF / row_shl:OFF; lanes shifted in from outside the row read 0, lanes from the
  // other segment of the row are masked
  const int lo = __builtin_amdgcn_update_dpp(0, inc, 0x110 + OFF, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, suf, 0x100 + OFF, 0xF, 0xF, true);
  inc += k >= OFF ? lo : 0;
  suf += k + OFF < L ? hi : 0;
}

template <int L>
__device__ __forceinline__ int seg_scan(int x, int k, int& total) {
  static_assert(L == 8 || L == 16, "segments inside a DPP row");
  int inc = x, suf = x;
  seg_scan_step<L, 1>(inc, suf, k);
  seg_scan_step<L, 2>(inc, suf, k);
  seg_scan_step<L, 4>(inc, suf, k);
  if constexpr (L == 16) seg_scan_step<L, 8>(inc, suf, k);
  total = inc + suf - x;
  return inc;
}

// Displacement of a home (position hr relative to its cell) and a candidate (record p relative
// to its own cell) whose cell is the home's cell + stencil offset o: (hr - o * cellsize) - p,
// with PBCr2's single shift in FP32 in a dimension collapsed to one cell.  Returns r^2.
__device__ __forceinline__ float rel_disp(const float* hr, float4 sh, float4 p, bool col, const float* Lf,
                                          const float* Hf, float& dx, float& dy, float& dz) {
  dx = (hr[0] - sh.x) - p.x;
  dy = (hr[1] - sh.y) - p.y;
  dz = (hr[2] - sh.z) - p.z;
  if (col) {
    if (fabsf(dx) > Hf[0]) dx -= copysignf(Lf[0], dx);
    if (fabsf(dy) > Hf[1]) dy -= copysignf(Lf[1], dy);
    if (fabsf(dz) > Hf[2]) dz -= copysignf(Lf[2], dz);
  }
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Mixed-path neighbour-parallel pair kernel over FP32 cell-relative records: a home's
// displacement to a candidate of stencil cell s is (rel_home - s * cellsize) - rel_cand, which
// is the minimum image for every pair within the cut-off when a dimension has at least three
// cells (cells >= rcut); a dimension collapsed to one cell takes PBCr2's single shift in FP32.
// The candidate records are 16 bytes (position and type); no FP64 arithmetic is spent per
// candidate.
//
// LPH lanes take one home.  Its candidates (the particles of the stencil cells not beyond the
// cut-off from the home) are written as codes (slot | stencil index << 27) into an LDS list
// of CAP entries per home, a chunk at a time.  The lanes then test the list interleaved
// (entry e by lane e mod LPH: balanced, no search for cell boundaries) and compact the codes of
// the pairs inside the cut-off to the front of the list, in list order; the table lookup and
// accumulation then run over the compacted codes, again interleaved, so they cost once per
// in-cut pair instead of once per candidate slot of the wavefront.  Every step is a fixed
// function of the cell list, so the sums are reproducible.
template <int NTP, int LPH>
__device__ __forceinline__ void pair_nb_rel_body(const PairArgs& a, int bid) {
  constexpr int HPB = 256 / LPH;
  constexpr int MPL = (kMaxStencil + LPH - 1) / LPH;
  constexpr int KU = kNbRound;
  constexpr int CAP = LPH * KU > 128 ? LPH * KU : 128;  // list entries per home (a chunk)
  constexpr int PB = kNbEval;
  constexpr unsigned kMask = (1u << kNbSlotBits) - 1;
  constexpr int NT = PairAcc<float, NTP>::NT;
  extern __shared__ __align__(16) double s_dyn[];
  __shared__ PairMeta<float> s_pm[kMaxT * kMaxT];
  __shared__ int2 s_pi[kMaxT * kMaxT];
  __shared__ int s_tp[kMaxT * kMaxT];
  __shared__ double s_red[4][kMaxTP + 1];
  __shared__ float4 s_sh[kMaxStencil];
  __shared__ int s_so[kMaxStencil];  // stencil offsets + 1 packed as ox | oy << 2 | oz << 4
  __shared__ unsigned s_list[HPB][CAP];
  const int nint = a.tm.nint;
  float* s_coef = (float*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  const float csz[3] = {a.csz[0], a.csz[1], a.csz[2]};
  {
    lds_copy16(s_coef, a.tcoef, nint * 12 * (int)sizeof(float));
    lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
    fill_pair_meta<float>(a.tm, s_pm, s_pi, s_tp);
    if (threadIdx.x < a.g.nst) {
      const int* o = a.g.st + 3 * threadIdx.x;
      s_sh[threadIdx.x] = make_float4(o[0] * csz[0], o[1] * csz[1], o[2] * csz[2], 0.f);
      s_so[threadIdx.x] = (o[0] + 1) | (o[1] + 1) << 2 | (o[2] + 1) << 4;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = threadIdx.x / LPH;
  const int s = bid * HPB + grp;
  const int k = threadIdx.x & (LPH - 1);
  const bool home = s < a.np;
  const int hs = home ? s : a.np - 1;
  const float4 hp = a.crel[hs];
  const int c = a.slot_cell[hs];
  const int ip = a.cell_ip[hs];  // the home's particle index, for the force store at the end
  __syncthreads();  // table, metadata and stencil in LDS
  const int hw = (int)hp.w;
  const int ht = hw & (kUncharged - 1);
  const bool hcharged = hw < kUncharged;
  const float hr[3] = {hp.x, hp.y, hp.z};
  // the home's cell coordinates, and per dimension and offset -1 / +1 the wrapped neighbour
  // coordinate and the squared distance from the home to that side of its cell
  int cc3[3];
  cc3[2] = div_small(c, a.g.nc[0] * a.g.nc[1], a.ncinv[1]);
  const int cxy = c - cc3[2] * a.g.nc[0] * a.g.nc[1];
  cc3[1] = div_small(cxy, a.g.nc[0], a.ncinv[0]);
  cc3[0] = cxy - cc3[1] * a.g.nc[0];
  int cm[3], cp[3];
  float dm[3], dp[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cm[d] = cc3[d] == 0 ? a.g.nc[d] - 1 : cc3[d] - 1;
    cp[d] = cc3[d] == a.g.nc[d] - 1 ? 0 : cc3[d] + 1;
    dm[d] = hr[d] * hr[d];
    const float u = csz[d] - hr[d];
    dp[d] = u > 0.f ? u * u : 0.f;
  }
  // stencil cells whose nearest point is beyond the cut-off (with a margin) are skipped
  const float cull2 = (float)(a.rcut2 * (1.0 + 1e-4)) + 1e-3f;
  int st0[MPL], cnt[MPL], tot = 0;
#pragma unroll
  for (int m = 0; m < MPL; ++m) {
    const int sc = k * MPL + m;
    cnt[m] = 0;
    st0[m] = 0;
    if (home && hcharged && sc < a.g.nst) {
      const int so = s_so[sc];
      int xyz[3];
      float bd2 = 0.f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int o = (so >> (2 * d)) & 3;
        xyz[d] = o == 0 ? cm[d] : (o == 2 ? cp[d] : cc3[d]);
        bd2 += o == 0 ? dm[d] : (o == 2 ? dp[d] : 0.f);
      }
      if (bd2 <= cull2) {
        const int nb = xyz[0] + a.g.nc[0] * (xyz[1] + a.g.nc[1] * xyz[2]);
        st0[m] = a.cell_start[nb];
        cnt[m] = a.cell_start[nb + 1] - st0[m];
      }
    }
    tot += cnt[m];
  }
  int ncand, incl;
  if constexpr (LPH == 8 || LPH == 16) {
    incl = seg_scan<LPH>(tot, k, ncand);
  } else {
    incl = tot;
#pragma unroll
    for (int off = 1; off < LPH; off <<= 1) {
      const int t = __shfl_up(incl, off, LPH);
      if (k >= off) incl += t;
    }
    ncand = __shfl(incl, LPH - 1, LPH);
  }
  const int pre = incl - tot;  // this lane's cells start at this candidate index

  const bool col = a.g.nc[0] == 1 || a.g.nc[1] == 1 || a.g.nc[2] == 1;
  const float Lf[3] = {(float)a.box.L[0], (float)a.box.L[1], (float)a.box.L[2]};
  const float Hf[3] = {(float)a.box.half[0], (float)a.box.half[1], (float)a.box.half[2]};
  const float rc2 = (float)a.rcut2;
  unsigned* list = s_list[grp];
  double et[NT], vir = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
  for (int q = 0; q < NT; ++q) et[q] = 0.0;
  int err = 0;
  for (int base = 0; base < ncand; base += CAP) {  // the same chunks for the home's lanes
    const int nc = min(CAP, ncand - base);
    {
      int p0 = pre;
#pragma unroll
      for (int m = 0; m < MPL; ++m) {
        const int lo = max(p0, base), hi = min(p0 + cnt[m], base + CAP);
        const unsigned code0 = (unsigned)(st0[m] - p0) + ((unsigned)(k * MPL + m) << kNbSlotBits);  // + t
        for (int t = lo; t < hi; ++t) list[t - base] = code0 + t;
        p0 += cnt[m];
      }
    }
    wave_sync();
    int nq = 0;  // codes of in-cut pairs at the front of the list
    for (int r0 = 0; r0 < nc; r0 += LPH * KU) {
      unsigned code[KU];
      float4 sh[KU], pv[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) code[u] = list[min(r0 + u * LPH + k, nc - 1)];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        sh[u] = s_sh[code[u] >> kNbSlotBits];
        pv[u] = a.crel[code[u] & kMask];
      }
      // in-cut candidates of this round (the home itself and uncharged partners are dropped
      // when the queue is evaluated)
      unsigned hm = 0;
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        float dx, dy, dz;
        const float r2 = rel_disp(hr, sh[u], pv[u], col, Lf, Hf, dx, dy, dz);
        if (r0 + u * LPH + k < nc && r2 < rc2) hm |= 1u << u;
      }
      const int nhit = __popc(hm);
      int ng, hi;
      if constexpr (LPH == 8 || LPH == 16) {
        hi = seg_scan<LPH>(nhit, k, ng);
      } else {
        hi = nhit;
#pragma unroll
        for (int off = 1; off < LPH; off <<= 1) {
          const int t = __shfl_up(hi, off, LPH);
          if (k >= off) hi += t;
        }
        ng = __shfl(hi, LPH - 1, LPH);
      }
      // every entry of this round was read above: the compacted codes go below its end
      int pos = nq + hi - nhit;
#pragma unroll
      for (int u = 0; u < KU; ++u)
        if (hm & (1u << u)) list[pos++] = code[u];
      nq = min(nq + ng, nc);
    }
    wave_sync();
    for (int e0 = 0; e0 < nq; e0 += LPH * PB) {
      unsigned cu[PB];
      float4 pj[PB], shv[PB];
#pragma unroll
      for (int w = 0; w < PB; ++w) cu[w] = list[min(e0 + w * LPH + k, nq - 1)];
#pragma unroll
      for (int w = 0; w < PB; ++w) {
        pj[w] = a.crel[cu[w] & kMask];
        shv[w] = s_sh[cu[w] >> kNbSlotBits];
      }
      // the batch's sums in FP32 (at most PB terms), added to the FP64 sums once per batch
      float ef[NT], vf = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int q = 0; q < NT; ++q) ef[q] = 0.f;
#pragma unroll
      for (int w = 0; w < PB; ++w) {
        const int tj = (int)pj[w].w;
        // tj >= kUncharged: weak charges, no pair energy with an uncharged partner
        if (e0 + w * LPH + k >= nq || (int)(cu[w] & kMask) == s || tj >= kUncharged) continue;
        float dx, dy, dz;
        const float r2 = rel_disp(hr, shv[w], pj[w], col, Lf, Hf, dx, dy, dz);
        const int mi = ht * kMaxT + tj;
        const PairMeta<float> pm = s_pm[mi];
        float us, fs;
        if (r2 < pm.r2at) {
          us = 1e10f;
          fs = 1e10f;
        } else if (r2 < pm.r2um) {
          err |= FLAG_R2UMIN;
          continue;
        } else {
          const int2 pi = s_pi[mi];
          int b = (int)((r2 - pm.zlo) * pm.invw);
          b = min(max(b, 0), kTabBins - 1);
          int ti = s_bins[pi.x * kTabBins + b];
          alignas(16) float t[12];
          {
            const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
            for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
          }
          while (ti < pi.y && r2 < t[0]) {
            ++ti;
            const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
            for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
          }
          const float dd = r2 - t[0];
          us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
          fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
        }
#pragma unroll
        for (int q = 0; q < NT; ++q) ef[q] += q == tj ? us : 0.f;
        gx += fs * dx;
        gy += fs * dy;
        gz += fs * dz;
        vf -= fs * r2;
      }
#pragma unroll
      for (int q = 0; q < NT; ++q) et[q] += (double)ef[q];
      fx += (double)gx;
      fy += (double)gy;
      fz += (double)gz;
      vir += (double)vf;
    }
    wave_sync();  // the next chunk overwrites the list
  }
  fx = msgd::xor_sum<LPH>(fx, threadIdx.x);
  fy = msgd::xor_sum<LPH>(fy, threadIdx.x);
  fz = msgd::xor_sum<LPH>(fz, threadIdx.x);
  if (home && k == 0) {
    a.force[3 * ip + 0] = fx * a.scale;
    a.force[3 * ip + 1] = fy * a.scale;
    a.force[3 * ip + 2] = fz * a.scale;
  }
  if (err) atomicOr(a.flags, err);
#pragma unroll
  for (int q = 0; q < NTP; ++q) {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) v += (s_tp[ht * kMaxT + t] == q) ? et[t] : 0.0;
    v = msgd::wave_sum(v);
    if (lane == 0) s_red[wv][q] = v;
  }
  {
    const double v = msgd::wave_sum(vir);
    if (lane == 0) s_red[wv][NTP] = v;
  }
  __syncthreads();
  if (threadIdx.x <= NTP) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][threadIdx.x];
    a.part[bid * (a.tm.ntp + 1) + (threadIdx.x == NTP ? a.tm.ntp : threadIdx.x)] = t;
  }
}

// Cell-tile pair kernel (mixed path): a workgroup takes a tile of up to tile_tx consecutive
// cells along x of one (y, z) cell row.  The stencil cells of all its homes are the cells of
// the 3 x 3 neighbouring rows over the tile's x range widened by one cell on each side; in
// cell order each such row segment is a run of consecutive cells, so the workgroup stages
// their records into LDS with coalesced loads, converted to coordinates relative to the tile's
// first cell (the cell-relative FP32 position plus the row's cell offset), and every candidate
// is then read from LDS.  The displacement of a home and a candidate is the difference of
// their tile coordinates: the minimum image for every pair within the cut-off when each
// dimension has at least three cells; a dimension collapsed to one cell takes PBCr2's single
// shift in FP32.  LPH lanes per home walk its 27 stencil cells (cells beyond the cut-off from
// the home skipped) as contiguous LDS ranges.  A tile whose neighbourhood holds more records
// than the LDS area (very dense cells) reads its candidates from global memory instead.
constexpr int kTileRec = 2048;   // staged records (32 KB)
constexpr int kTileCells = 128;  // cells of a tile's neighbourhood (9 rows of tile_tx + 2 cells)

template <int NTP, int LPH>
__device__ __forceinline__ void pair_tile_body(const PairArgs& a, int bid) {
  constexpr int HPB = 256 / LPH;
  constexpr int MPL = (kMaxStencil + LPH - 1) / LPH;
  extern __shared__ __align__(16) double s_dyn[];
  __shared__ PairMeta<float> s_pm[kMaxT * kMaxT];
  __shared__ int2 s_pi[kMaxT * kMaxT];
  __shared__ int s_tp[kMaxT * kMaxT];
  __shared__ double s_red[4][kMaxTP + 1];
  __shared__ int s_cst[HPB][kMaxStencil + 1], s_cpf[HPB][kMaxStencil + 1];
  __shared__ int s_cell[kTileCells], s_coff[kTileCells + 1], s_cbeg[kTileCells];
  __shared__ float4 s_rec[kTileRec];
  static_assert(HPB * kMaxStencil <= kTileRec, "stencil offsets share the staging area");
  float4 (*s_csh)[kMaxStencil] = (float4 (*)[kMaxStencil])s_rec;  // global mode: the stencil cells' offsets
  __shared__ int s_wsum[4];
  const int nint = a.tm.nint;
  float* s_coef = (float*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  {
    lds_copy16(s_coef, a.tcoef, nint * 12 * (int)sizeof(float));
    lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
    fill_pair_meta<float>(a.tm, s_pm, s_pi, s_tp);
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nx = a.g.nc[0], ny = a.g.nc[1], nz = a.g.nc[2];
  const int ntx = (nx + a.tile_tx - 1) / a.tile_tx;
  const int tx_i = bid % ntx, cy = (bid / ntx) % ny, cz = bid / (ntx * ny);
  const int x0 = tx_i * a.tile_tx, tx = min(a.tile_tx, nx - x0);
  const int rx = nx > 1 ? 1 : 0, ry = ny > 1 ? 1 : 0, rz = nz > 1 ? 1 : 0;
  const int wseg = tx + 2 * rx;                 // cells per row segment
  const int nrow = (2 * ry + 1) * (2 * rz + 1);
  const int ncells = nrow * wseg;               // <= kTileCells (tile_tx <= 12)
  float csz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) csz[d] = (float)(1.0 / a.g.csi[d]);
  // the neighbourhood's cells: row r = (dy, dz), segment position q (unwrapped x = x0 - rx + q)
  int cnt = 0;
  if (tid < ncells) {
    const int r = tid / wseg, q = tid - r * wseg;
    const int dy = r % (2 * ry + 1) - ry, dz = r / (2 * ry + 1) - rz;
    int x = x0 - rx + q, y = cy + dy, z = cz + dz;
    x += x < 0 ? nx : (x >= nx ? -nx : 0);
    y += y < 0 ? ny : (y >= ny ? -ny : 0);
    z += z < 0 ? nz : (z >= nz ? -nz : 0);
    const int c = x + nx * (y + ny * z);
    s_cell[tid] = c;
    s_cbeg[tid] = a.cell_start[c];
    cnt = a.cell_start[c + 1] - s_cbeg[tid];
  }
  // offsets of the cells' records in the staged area (ncells <= 128: two wavefronts)
  {
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    const int base = wv == 1 ? s_wsum[0] : 0;
    if (tid < ncells) s_coff[tid + 1] = base + incl;
    if (tid == 0) s_coff[0] = 0;
  }
  __syncthreads();
  const int total = s_coff[ncells];
  const bool staged = total <= kTileRec;
  if (staged) {
    // cooperative copy: record k belongs to the cell whose offsets bracket it
    for (int k = tid; k < total; k += 256) {
      int lo = 0, hi = ncells - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_coff[mid] <= k) lo = mid;
        else hi = mid - 1;
      }
      const int r = lo / wseg, q = lo - r * wseg;
      const int dy = r % (2 * ry + 1) - ry, dz = r / (2 * ry + 1) - rz;
      const float4 v = a.crel[s_cbeg[lo] + (k - s_coff[lo])];
      s_rec[k] = make_float4(v.x + (q - rx) * csz[0], v.y + dy * csz[1], v.z + dz * csz[2], v.w);
    }
  }
  __syncthreads();
  // homes: the tile's cells, row (0, 0) at segment positions rx .. rx + tx - 1
  const int r0 = ry + (2 * ry + 1) * rz;
  const int hc0 = r0 * wseg + rx;
  const int hslot0 = s_cbeg[hc0];
  const int nh = s_coff[hc0 + tx] - s_coff[hc0];
  const bool col = nx == 1 || ny == 1 || nz == 1;
  const float Lf[3] = {(float)a.box.L[0], (float)a.box.L[1], (float)a.box.L[2]};
  const float Hf[3] = {(float)a.box.half[0], (float)a.box.half[1], (float)a.box.half[2]};
  const float rc2 = (float)a.rcut2;
  const float cull2 = (float)(a.rcut2 * (1.0 + 1e-4)) + 1e-3f;
  const int grp = tid / LPH, k = tid & (LPH - 1);
  constexpr int NT = PairAcc<float, NTP>::NT;  // partner types
  double et[NT], etp[NTP];                      // a home's energies by partner type; this lane's by type pair
#pragma unroll
  for (int q = 0; q < NT; ++q) et[q] = 0.0;
#pragma unroll
  for (int q = 0; q < NTP; ++q) etp[q] = 0.0;
  double vacc = 0.0;
  int errs = 0;
  for (int h0 = 0; h0 < nh; h0 += HPB) {
    const int h = h0 + grp;
    const bool home = h < nh;
    const int hh = home ? h : nh - 1;
    // the home's cell: segment position hq (row 0), its record and tile coordinates
    int hq = 0;
    while (hq + 1 < tx && s_coff[hc0 + hq + 1] - s_coff[hc0] <= hh) ++hq;
    const int hidx = s_coff[hc0] + hh;            // index in the staged area
    float4 hp;
    if (staged) {
      hp = s_rec[hidx];
    } else {
      hp = a.crel[hslot0 + hh];                   // relative to its own cell
    }
    const int hw = (int)hp.w;
    const int ht = hw & (kUncharged - 1);
    const bool hcharged = hw < kUncharged;
    // the home's cell box in the same coordinates (for the stencil-cell distance bound)
    const float bx0 = staged ? hq * csz[0] : 0.f;
    const float hr[3] = {hp.x - bx0, hp.y, hp.z};  // position within the home cell
    int st0[MPL], cn[MPL], tot = 0;
#pragma unroll
    for (int m = 0; m < MPL; ++m) {
      const int sc = k * MPL + m;
      st0[m] = 0;
      cn[m] = 0;
      if (sc < a.g.nst) {
        const int ox = a.g.st[3 * sc], oy = a.g.st[3 * sc + 1], oz = a.g.st[3 * sc + 2];
        const int o3[3] = {ox, oy, oz};
        float bd2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float dist = o3[d] < 0 ? hr[d] : (o3[d] > 0 ? csz[d] - hr[d] : 0.f);
          bd2 += dist > 0.f ? dist * dist : 0.f;
        }
        if (bd2 <= cull2) {
          const int cidx = ((oy + ry) + (2 * ry + 1) * (oz + rz)) * wseg + (hq + rx + ox);
          cn[m] = s_coff[cidx + 1] - s_coff[cidx];
          if (staged) {
            st0[m] = s_coff[cidx];
          } else {
            st0[m] = s_cbeg[cidx];
            s_csh[grp][sc] = make_float4(ox * csz[0], oy * csz[1], oz * csz[2], 0.f);
          }
        }
      }
      tot += cn[m];
    }
    int incl = tot;
#pragma unroll
    for (int off = 1; off < LPH; off <<= 1) {
      const int t = __shfl_up(incl, off, LPH);
      if (k >= off) incl += t;
    }
    const int ncand = __shfl(incl, LPH - 1, LPH);
    {
      int pre = incl - tot;
#pragma unroll
      for (int m = 0; m < MPL; ++m) {
        const int sc = k * MPL + m;
        if (sc < a.g.nst) {
          s_cst[grp][sc] = st0[m];
          s_cpf[grp][sc] = pre;
        }
        pre += cn[m];
      }
    }
    if (k == 0) s_cpf[grp][a.g.nst] = ncand;
    __syncthreads();
    const int self = staged ? hidx : hslot0 + hh;
    const int share = (ncand + LPH - 1) / LPH;
    int i = k * share;
    const int i1 = (home && hcharged) ? min(ncand, i + share) : 0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < i1) {
      int cc = 0;
      while (s_cpf[grp][cc + 1] <= i) ++cc;
      int j = s_cst[grp][cc] + (i - s_cpf[grp][cc]);
      int cend = s_cst[grp][cc] + (s_cpf[grp][cc + 1] - s_cpf[grp][cc]);
      float4 sh = staged ? make_float4(0.f, 0.f, 0.f, 0.f) : s_csh[grp][cc];
      constexpr int U = 4;
      while (i < i1) {
        int jv[U];
        bool hv[U];
        float4 shv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          hv[u] = i < i1;
          jv[u] = hv[u] ? j : self;
          shv[u] = sh;
          if (hv[u]) {
            ++i;
            ++j;
            if (j == cend && i < i1) {
              do { ++cc; } while (s_cpf[grp][cc + 1] <= i);
              j = s_cst[grp][cc];
              cend = j + (s_cpf[grp][cc + 1] - s_cpf[grp][cc]);
              if (!staged) sh = s_csh[grp][cc];
            }
          }
        }
        float4 pv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pv[u] = staged ? s_rec[jv[u]] : a.crel[jv[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!hv[u] || jv[u] == self) continue;
          float dx = hp.x - (pv[u].x + shv[u].x), dy = hp.y - (pv[u].y + shv[u].y),
                dz = hp.z - (pv[u].z + shv[u].z);
          if (col) {
            if (fabsf(dx) > Hf[0]) dx -= copysignf(Lf[0], dx);
            if (fabsf(dy) > Hf[1]) dy -= copysignf(Lf[1], dy);
            if (fabsf(dz) > Hf[2]) dz -= copysignf(Lf[2], dz);
          }
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < rc2)) continue;
          const int tj = (int)pv[u].w;
          if (tj >= kUncharged) continue;  // weak charges: uncharged partner
          const int mi = ht * kMaxT + tj;
          const PairMeta<float> pm = s_pm[mi];
          const int2 pi = s_pi[mi];
          float us, fs;
          if (r2 < pm.r2at) {
            us = 1e10f;
            fs = 1e10f;
          } else if (r2 < pm.r2um) {
            errs |= FLAG_R2UMIN;
            continue;
          } else {
            int b = (int)((r2 - pm.zlo) * pm.invw);
            b = min(max(b, 0), kTabBins - 1);
            int ti = s_bins[pi.x * kTabBins + b];
            alignas(16) float t[12];
            {
              const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
              for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
            }
            while (ti < pi.y && r2 < t[0]) {
              ++ti;
              const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
              for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
            }
            const float dd = r2 - t[0];
            us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
            fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
          }
#pragma unroll
          for (int q = 0; q < NT; ++q) et[q] += (q == tj) ? (double)us : 0.0;
          fx += (double)(fs * dx);
          fy += (double)(fs * dy);
          fz += (double)(fs * dz);
          vacc -= (double)(fs * r2);
        }
      }
    }
#pragma unroll
    for (int off = LPH / 2; off > 0; off >>= 1) {
      fx += __shfl_xor(fx, off, 64);
      fy += __shfl_xor(fy, off, 64);
      fz += __shfl_xor(fz, off, 64);
    }
    if (home && k == 0) {
      const int ip = a.cell_ip[hslot0 + h];
      a.force[3 * ip + 0] = fx * a.scale;
      a.force[3 * ip + 1] = fy * a.scale;
      a.force[3 * ip + 2] = fz * a.scale;
    }
    // this home's energies by partner type into type-pair slots (the home type fixes the pair)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double v = et[t];
      et[t] = 0.0;
#pragma unroll
      for (int q = 0; q < NTP; ++q) etp[q] += (s_tp[ht * kMaxT + t] == q) ? v : 0.0;
    }
    __syncthreads();  // s_cst / s_cpf / s_csh are rewritten for the next homes
  }
  if (errs) atomicOr(a.flags, errs);
#pragma unroll
  for (int q = 0; q < NTP; ++q) {
    const double v = msgd::wave_sum(etp[q]);
    if (lane == 0) s_red[wv][q] = v;
  }
  {
    const double v = msgd::wave_sum(vacc);
    if (lane == 0) s_red[wv][NTP] = v;
  }
  __syncthreads();
  if (tid <= NTP) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][tid];
    a.part[bid * (a.tm.ntp + 1) + (tid == NTP ? a.tm.ntp : tid)] = t;
  }
}

template <int NTP, int LPH>
__global__ __launch_bounds__(256) void k_pair_tile(PairArgs a) {
  pair_tile_body<NTP, LPH>(a, a.xcd ? xcd_logical(blockIdx.x, gridDim.x) : blockIdx.x);
}

template <int NTP, int LPH>
__global__ __launch_bounds__(256) void k_pair_tile_seg(const PairArgs* __restrict__ args,
                                                       const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  pair_tile_body<NTP, LPH>(args[s.x], s.y);
}

template <typename real, int NTP, int LPH>
__global__ __launch_bounds__(256) void k_pair_nb(PairArgs a) {
  const int bid = a.xcd ? xcd_logical(blockIdx.x, gridDim.x) : blockIdx.x;
  if constexpr (sizeof(real) == 4)
    pair_nb_rel_body<NTP, LPH>(a, bid);
  else
    pair_nb_body<real, NTP, LPH>(a, bid);
}

template <typename real, int NTP, int LPH>
__global__ __launch_bounds__(256) void k_pair_nb_seg(const PairArgs* __restrict__ args,
                                                     const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  if constexpr (sizeof(real) == 4)
    pair_nb_rel_body<NTP, LPH>(args[s.x], s.y);
  else
    pair_nb_body<real, NTP, LPH>(args[s.x], s.y);
}

// Verlet list build (SetVListMC semantics, Src/nlist.F90:820-884, as a full list that also
// serves MD).  Homes are taken in cell order (their "slots"); kVlLph lanes per home walk the
// candidates of its stencil cells in cell order (cells whose nearest point lies beyond rcut +
// skin are skipped), kVlLph candidates per round; a ballot over the group ranks the partners
// with r2 < (rcut + skin)^2 (PBCr2's single shift and sum of squares in FP64, not contracted,
// so the pair test is the reference's bit for bit), which are appended as slots in that order,
// so the list is the one a single thread per home walking the cells would write.  The
// slot-ordered positions and types the pair loop reads are copied from the cell records; the
// slot-record kernel keeps their positions current until the next build.  Also snapshots the
// positions (from the cell records) the displacement test of a UTotal caller measures from.
struct VlistArgs {
  int np, maxn;
  const double4* cpos;
  const int* cell_ip;
  const int* cell_start;
  const int* slot_cell;
  CellGrid g;
  Box box;
  double rl2;          // (rcut + skin)^2
  double* ro_ref;
  int* vl_home;
  int* vl_n;
  int* vl_nbr;
  double4* vl_pos;
  int4* vl_rel;        // mixed path: fixed-point records relative to the cell (null otherwise)
  int* vl_slot;
  int* flags;
  int stage;           // lists staged in LDS (rows of maxn + 1) and written out row by row
};

constexpr int kVlLph = 16;
constexpr int kVlStageBytes = 64 * 1024;  // LDS for staged lists: maxn up to 1023

__global__ __launch_bounds__(256) void k_vlist_build(VlistArgs a) {
  constexpr int LPH = kVlLph, HPB = 256 / LPH, MPL = (kMaxStencil + LPH - 1) / LPH;
  __shared__ int s_cst[HPB][kMaxStencil + 1], s_cpf[HPB][kMaxStencil + 1];
  __shared__ int s_n[HPB];
  extern __shared__ int s_list[];  // [HPB][maxn + 1] with a.stage
  const int ls = a.maxn + 1;
  const int tid = threadIdx.x, lane = tid & 63, grp = tid / LPH, k = tid & (LPH - 1);
  const int s = blockIdx.x * HPB + grp;
  const bool home = s < a.np;
  const int hs = home ? s : a.np - 1;
  const double4 hp = a.cpos[hs];
  if (home && k == 0) {
    const int ip = a.cell_ip[s];
    a.vl_home[s] = ip;
    a.vl_pos[s] = hp;
    a.vl_slot[ip] = s;
    a.ro_ref[3 * ip] = hp.x;
    a.ro_ref[3 * ip + 1] = hp.y;
    a.ro_ref[3 * ip + 2] = hp.z;
  }
  const int c = a.slot_cell[hs];
  const int cx = c % a.g.nc[0], cy = (c / a.g.nc[0]) % a.g.nc[1], cz = c / (a.g.nc[0] * a.g.nc[1]);
  // home position inside its cell, for the distance bound of each stencil cell (the bound
  // carries a margin, so no partner is lost)
  const int cc3[3] = {cx, cy, cz};
  const double hh[3] = {hp.x, hp.y, hp.z};
  double hrel[3], csz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    csz[d] = a.box.L[d] / a.g.nc[d];
    hrel[d] = (hh[d] + a.box.half[d]) - cc3[d] * csz[d];
  }
  if (home && k == 0 && a.vl_rel)  // the cell of the build and the type tag in w
    a.vl_rel[s] = make_int4(__double2int_rn(hrel[0] * (kVlUnits / csz[0])), __double2int_rn(hrel[1] * (kVlUnits / csz[1])),
                            __double2int_rn(hrel[2] * (kVlUnits / csz[2])), cx | (cy << 9) | (cz << 18) | ((int)hp.w << 27));
  const double cull2 = a.rl2 * (1.0 + 1e-5);
  int st0[MPL], cnt[MPL], tot = 0;
#pragma unroll
  for (int m = 0; m < MPL; ++m) {
    const int sc = k * MPL + m;
    st0[m] = 0;
    cnt[m] = 0;
    if (sc < a.g.nst) {
      double bd2 = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int o = a.g.st[3 * sc + d];
        const double dist = o < 0 ? hrel[d] : (o > 0 ? csz[d] - hrel[d] : 0.0);
        bd2 += dist > 0.0 ? dist * dist : 0.0;
      }
      if (bd2 <= cull2) {
        int x = cx + a.g.st[3 * sc], y = cy + a.g.st[3 * sc + 1], z = cz + a.g.st[3 * sc + 2];
        x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
        y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
        z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
        const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
        st0[m] = a.cell_start[nb];
        cnt[m] = a.cell_start[nb + 1] - st0[m];
      }
    }
    tot += cnt[m];
  }
  int incl = tot;
#pragma unroll
  for (int off = 1; off < LPH; off <<= 1) {
    const int t = __shfl_up(incl, off, LPH);
    if (k >= off) incl += t;
  }
  const int ncand = __shfl(incl, LPH - 1, LPH);
  {
    int pre = incl - tot;
#pragma unroll
    for (int m = 0; m < MPL; ++m) {
      const int sc = k * MPL + m;
      if (sc < a.g.nst) {
        s_cst[grp][sc] = st0[m];
        s_cpf[grp][sc] = pre;
      }
      pre += cnt[m];
    }
  }
  if (k == 0) s_cpf[grp][a.g.nst] = ncand;
  __syncthreads();
  // rounds: the most any group of the wavefront needs, so every lane reaches each ballot
  int rounds = (ncand + LPH - 1) / LPH;
#pragma unroll
  for (int off = LPH; off < 64; off <<= 1) rounds = max(rounds, __shfl_xor(rounds, off, 64));
  const int gsh = lane & ~(LPH - 1);  // the group's first lane in the wavefront
  // four rounds per iteration: their candidate loads are issued before the first test
  constexpr int U = 4;
  int n = 0, cc = 0;
  for (int r = 0; r < rounds; r += U) {
    int jv[U];
    bool val[U];
    double4 pv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = (r + u) * LPH + k;
      val[u] = home && i < ncand;
      jv[u] = s;
      if (val[u]) {
        while (s_cpf[grp][cc + 1] <= i) ++cc;
        jv[u] = s_cst[grp][cc] + (i - s_cpf[grp][cc]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pv[u] = a.cpos[val[u] ? jv[u] : hs];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bool hit = false;
      if (val[u] && jv[u] != s) {
#pragma clang fp contract(off)
        double d[3] = {hp.x - pv[u].x, hp.y - pv[u].y, hp.z - pv[u].z};
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (fabs(d[q]) > a.box.half[q]) d[q] -= copysign(a.box.L[q], d[q]);
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        hit = r2 < a.rl2;
      }
      const unsigned gm = (unsigned)(__ballot(hit) >> gsh) & ((1u << LPH) - 1u);
      if (hit) {
        const int m = n + __popc(gm & ((1u << k) - 1u));
        if (m < a.maxn) {
          if (a.stage)
            s_list[grp * ls + m] = jv[u];
          else
            a.vl_nbr[(size_t)m * a.np + s] = jv[u];
        }
      }
      n += __popc(gm);
    }
  }
  if (k == 0) {
    if (home && n > a.maxn) atomicOr(a.flags, FLAG_NLIST);
    n = home ? min(n, a.maxn) : 0;
    s_n[grp] = n;
    if (home) a.vl_n[s] = n;
  }
  if (a.stage) {
    // the workgroup's lists, row by row: the HPB homes of a row are contiguous in vl_nbr
    __syncthreads();
    int nmax = 0;
#pragma unroll
    for (int h = 0; h < HPB; ++h) nmax = max(nmax, s_n[h]);
    const int s0 = blockIdx.x * HPB;
    for (int idx = tid; idx < nmax * HPB; idx += 256) {
      const int m = idx / HPB, h = idx - m * HPB;
      if (m < s_n[h]) a.vl_nbr[(size_t)m * a.np + s0 + h] = s_list[h * ls + m];
    }
  }
}

// Pair loop over the Verlet list (UTwoBodyA over jpnlist, Src/energy.F90:445-544): one thread
// per home slot, partner records (position, tagged type) read from the slot-ordered array,
// so neighbours of nearby homes share cache lines; two candidates per iteration with both
// loads issued first.
struct VlPairArgs {
  PairArgs p;
  int maxn;
  const int* vl_home;
  const int* vl_n;
  const int* vl_nbr;
  const double4* vl_pos;
};

template <typename real, int NTP>
__global__ __launch_bounds__(256) void k_pair_vl(VlPairArgs v) {
  const PairArgs& a = v.p;
  extern __shared__ __align__(16) double s_dyn[];
  __shared__ PairMeta<real> s_pm[kMaxT * kMaxT];
  __shared__ int2 s_pi[kMaxT * kMaxT];
  __shared__ int s_tp[kMaxT * kMaxT];
  __shared__ double s_red[4][kMaxTP + 1];
  const int nint = a.tm.nint;
  real* s_coef = (real*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  {
    const real* src = (const real*)a.tcoef;
    lds_copy16(s_coef, src, nint * 12 * (int)sizeof(real));
    lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
    fill_pair_meta<real>(a.tm, s_pm, s_pi, s_tp);
  }
  __syncthreads();
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const real rc2 = (real)a.rcut2;
  PairAcc<real, NTP> acc;
  acc.zero();
  int ht = 0;
  if (s < a.np) {
    const int ip = v.vl_home[s];
    const double4 hp = v.vl_pos[s];
    const int hw = (int)hp.w;
    ht = hw & (kUncharged - 1);
    const int n = hw >= kUncharged ? 0 : v.vl_n[s];  // an uncharged home has no pair energy
    // the next pair of neighbour slots is loaded one iteration ahead
    int jA = n > 0 ? v.vl_nbr[s] : 0;
    int jB = n > 1 ? v.vl_nbr[(size_t)a.np + s] : jA;
    for (int k = 0; k < n; k += 2) {
      const bool haveB = k + 1 < n;
      const double4 pA = v.vl_pos[jA];
      const double4 pB = v.vl_pos[jB];
      const int nA = k + 2 < n ? v.vl_nbr[(size_t)(k + 2) * a.np + s] : jA;
      const int nB = k + 3 < n ? v.vl_nbr[(size_t)(k + 3) * a.np + s] : nA;
      pair_eval_m<real, NTP>(hp, ht, pA, (int)pA.w, a.box, rc2, true, s_coef, s_bins, s_pm, s_pi, acc);
      if (haveB)
        pair_eval_m<real, NTP>(hp, ht, pB, (int)pB.w, a.box, rc2, true, s_coef, s_bins, s_pm, s_pi, acc);
      jA = nA;
      jB = nB;
    }
    a.force[3 * ip + 0] = acc.fx * a.scale;
    a.force[3 * ip + 1] = acc.fy * a.scale;
    a.force[3 * ip + 2] = acc.fz * a.scale;
  }
  if (acc.err) atomicOr(a.flags, acc.err);
#pragma unroll
  for (int q = 0; q < NTP; ++q) {
    const double t = msgd::wave_sum(acc.to_tp(q, ht, s_tp));
    if (lane == 0) s_red[wv][q] = t;
  }
  {
    const double t = msgd::wave_sum(acc.vir);
    if (lane == 0) s_red[wv][NTP] = t;
  }
  __syncthreads();
  if (threadIdx.x <= NTP) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][threadIdx.x];
    a.part[blockIdx.x * (a.tm.ntp + 1) + (threadIdx.x == NTP ? a.tm.ntp : threadIdx.x)] = t;
  }
}

// Mixed-path pair loop over the Verlet list with fixed-point records relative to the corner of
// each slot's cell at the build (16 bytes per partner instead of 32; units of cellsize / 2^28,
// 7.5e-8 A for 20 A cells): the displacement of home i and partner j is
// (rel_i - rel_j) + (c_i - c_j) 2^28 units, exact in 32-bit integers, with the cell difference
// taken to -1..1 per dimension (the partners of a home sat in its cell or a neighbouring one at
// the build, and every dimension has at least three cells), which is the minimum image; it is
// converted to FP32 with one rounding, as the FP64 difference is in the other kernels.
// Positions are refreshed each pass relative to the same cells.  Four partners per round with
// their loads issued first; the round's sums in FP32 (at most four terms), added to the FP64
// sums once per round.
struct VlRelArgs {
  PairArgs p;
  const int* vl_home;
  const int* vl_n;
  const int* vl_nbr;
  const int4* vl_rel;
  float unit[3];  // cellsize / 2^28
  int nc[3];
};

template <int NTP>
__global__ __launch_bounds__(256) void k_pair_vl_rel(VlRelArgs v) {
  constexpr int NT = PairAcc<float, NTP>::NT, PB = 4;
  const PairArgs& a = v.p;
  extern __shared__ __align__(16) double s_dyn[];
  __shared__ PairMeta<float> s_pm[kMaxT * kMaxT];
  __shared__ int2 s_pi[kMaxT * kMaxT];
  __shared__ int s_tp[kMaxT * kMaxT];
  __shared__ double s_red[4][kMaxTP + 1];
  const int nint = a.tm.nint;
  float* s_coef = (float*)s_dyn;
  int* s_bins = (int*)(s_coef + nint * 12);
  lds_copy16(s_coef, (const float*)a.tcoef, nint * 12 * (int)sizeof(float));
  lds_copy16(s_bins, a.tbins, NTP * kTabBins * (int)sizeof(int));
  fill_pair_meta<float>(a.tm, s_pm, s_pi, s_tp);
  __syncthreads();
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const float rc2 = (float)a.rcut2;
  double et[NT], vir = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
  for (int q = 0; q < NT; ++q) et[q] = 0.0;
  int err = 0, ht = 0;
  if (s < a.np) {
    const int ip = v.vl_home[s];
    const int4 hr = v.vl_rel[s];
    const int hw = hr.w;
    const int htag = hw >> 27;
    ht = htag & (kUncharged - 1);
    const int n = htag >= kUncharged ? 0 : v.vl_n[s];  // an uncharged home has no pair energy
    const int hx[3] = {hr.x, hr.y, hr.z};
    for (int k0 = 0; k0 < n; k0 += PB) {
      int j[PB];
      int4 p[PB];
#pragma unroll
      for (int u = 0; u < PB; ++u) j[u] = k0 + u < n ? v.vl_nbr[(size_t)(k0 + u) * a.np + s] : s;
#pragma unroll
      for (int u = 0; u < PB; ++u) p[u] = v.vl_rel[j[u]];
      float ef[NT], vf = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int q = 0; q < NT; ++q) ef[q] = 0.f;
#pragma unroll
      for (int u = 0; u < PB; ++u) {
        const int pw = p[u].w;
        const int tj = pw >> 27;
        // tj >= kUncharged: weak charges, no pair energy with an uncharged partner
        if (k0 + u >= n || tj >= kUncharged) continue;
        const int pr[3] = {p[u].x, p[u].y, p[u].z};
        float d[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          int dc = ((hw >> (9 * q)) & 511) - ((pw >> (9 * q)) & 511);
          dc -= dc > 1 ? v.nc[q] : 0;
          dc += dc < -1 ? v.nc[q] : 0;
          d[q] = (float)((hx[q] - pr[q]) + (dc << 28)) * v.unit[q];
        }
        const float r2 = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
        if (r2 > rc2) continue;  // UTwoBodyA's r2 > rcut2 test
        const int mi = ht * kMaxT + tj;
        const PairMeta<float> pm = s_pm[mi];
        float us, fs;
        if (r2 < pm.r2at) {
          us = 1e10f;
          fs = 1e10f;
        } else if (r2 < pm.r2um) {
          err |= FLAG_R2UMIN;
          continue;
        } else {
          const int2 pi = s_pi[mi];
          int b = (int)((r2 - pm.zlo) * pm.invw);
          b = min(max(b, 0), kTabBins - 1);
          int ti = s_bins[pi.x * kTabBins + b];
          alignas(16) float t[12];
          {
            const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
            for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
          }
          while (ti < pi.y && r2 < t[0]) {
            ++ti;
            const float4* rp = (const float4*)(s_coef + 12 * ti);
#pragma unroll
            for (int q = 0; q < 3; ++q) *(float4*)(t + 4 * q) = rp[q];
          }
          const float dd = r2 - t[0];
          us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
          fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
        }
#pragma unroll
        for (int q = 0; q < NT; ++q) ef[q] += q == tj ? us : 0.f;
        gx += fs * d[0];
        gy += fs * d[1];
        gz += fs * d[2];
        vf -= fs * r2;
      }
#pragma unroll
      for (int q = 0; q < NT; ++q) et[q] += (double)ef[q];
      fx += (double)gx;
      fy += (double)gy;
      fz += (double)gz;
      vir += (double)vf;
    }
    a.force[3 * ip + 0] = fx * a.scale;
    a.force[3 * ip + 1] = fy * a.scale;
    a.force[3 * ip + 2] = fz * a.scale;
  }
  if (err) atomicOr(a.flags, err);
#pragma unroll
  for (int q = 0; q < NTP; ++q) {
    double val = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) val += (s_tp[ht * kMaxT + t] == q) ? et[t] : 0.0;
    const double tot = msgd::wave_sum(val);
    if (lane == 0) s_red[wv][q] = tot;
  }
  {
    const double tot = msgd::wave_sum(vir);
    if (lane == 0) s_red[wv][NTP] = tot;
  }
  __syncthreads();
  if (threadIdx.x <= NTP) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][threadIdx.x];
    a.part[blockIdx.x * (a.tm.ntp + 1) + (threadIdx.x == NTP ? a.tm.ntp : threadIdx.x)] = t;
  }
}

// The list in MOLSIM's layout (Src/nlist.F90:724-884): jpnlist(maxnneigh, np), column ip
// holding the 1-based partners of particle ip in ascending order, nneighpn(ip) of them, as
// SetVListMD (kind 0: a pair (ip, jp) is kept by ip when ip < jp and ip + jp is odd, or ip > jp
// and ip + jp is even, so every pair appears once) or SetVListMC (kind 1: the full list) writes
// them.  One wavefront per home: its partners (slots) become particle indices, the kind's rule
// filters them, a ballot compacts them into LDS, and each is placed by its rank (the number of
// smaller partners; indices are distinct).
struct VlExportArgs {
  int np, kind, ldo;     // ldo: the caller's maxnneigh
  int width, wpb;        // LDS row width (the list width) and wavefronts per workgroup
  const int* vl_home;
  const int* vl_n;
  const int* vl_nbr;
  int* nneigh;
  int* jpn;
  int* flags;
};

__global__ __launch_bounds__(256) void k_vlist_export(VlExportArgs a) {
  extern __shared__ int s_rows[];  // [wavefronts][row width]
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * a.wpb + wv;
  if (s >= a.np) return;  // whole wavefronts leave; no workgroup barrier follows
  int* row = s_rows + wv * a.width;
  const int ip = a.vl_home[s], n = a.vl_n[s];
  int cnt = 0;
  for (int k0 = 0; k0 < n; k0 += 64) {
    const int k = k0 + lane;
    int j = -1;
    bool keep = false;
    if (k < n) {
      j = a.vl_home[a.vl_nbr[(size_t)k * a.np + s]];
      const int par = (ip + j) & 1;
      keep = a.kind == 1 || (ip < j ? par != 0 : par == 0);
    }
    const unsigned long long bal = __ballot(keep);
    if (keep) row[cnt + __popcll(bal & ((1ull << lane) - 1ull))] = j;
    cnt += __popcll(bal);
  }
  wave_sync();
  for (int k = lane; k < cnt; k += 64) {
    const int x = row[k];
    int r = 0;
    for (int m = 0; m < cnt; ++m) r += row[m] < x;
    if (r < a.ldo) a.jpn[(size_t)ip * a.ldo + r] = x + 1;
  }
  if (lane == 0) {
    a.nneigh[ip] = min(cnt, a.ldo);
    if (cnt > a.ldo) atomicOr(a.flags, FLAG_NLIST);
  }
}

}  // namespace

// Work items: sum over cells of ceil(count / kPairHB) <= ncell + np / kPairHB; a grid of
// up to 8 workgroups per CU walks them with a stride.
// The neighbour-parallel kernel serves every cell-list and all-pairs system (any stencil:
// 27 cells, or fewer where dimensions collapse to one cell) unless tile 0 is chosen.
namespace {

// Pair kernel of systems with more than kLdsTypes particle types (up to MSG_MAX_TYPES, so up
// to 36 type pairs): the table's records and bins are read from global memory (they no longer
// fit the LDS of a workgroup) and every pair is evaluated in FP64, one thread per home, over
// the cell list's stencil or the home's Verlet-list row.  The pair rules are UTwoBodyACellList
// / UTwoBodyA's (PBCr2 single shift, r2 < rcut2 or, for all pairs and Verlet lists,
// r2 <= rcut2, hard-core emulation below r2atat, an error below r2umin).  Per-type-pair
// energies and the virial go to the block's partial row as in the other pair kernels.
struct ManyArgs {
  PairArgs p;
  const double* coef;   // FP64 records (12 per interval), first/last per type pair in p.tm
  int verlet;
  const int* vl_home;
  const int* vl_n;
  const int* vl_nbr;
  const double4* vl_pos;
};

__device__ __forceinline__ void many_eval(const ManyArgs& g, const double4& hp, int ht, const double4& pj,
                                          double* et, double& vir, double* f, int& err) {
  const PairArgs& a = g.p;
  const int tj = (int)pj.w;
  if (tj >= kUncharged) return;  // weak charges: uncharged partner
  double d[3] = {hp.x - pj.x, hp.y - pj.y, hp.z - pj.z};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (fabs(d[q]) > a.box.half[q]) d[q] -= copysign(a.box.L[q], d[q]);
  const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (a.cut_inclusive ? (r2 > a.rcut2) : !(r2 < a.rcut2)) return;
  const int tp = a.tm.tpidx[ht * kMaxT + tj];
  double us, fs;
  if (r2 < a.tm.r2atat[tp]) {
    us = 1e10;
    fs = 1e10;
  } else if (r2 < a.tm.r2umin[tp]) {
    err |= FLAG_R2UMIN;
    return;
  } else {
    int b = (int)((r2 - a.tm.zlo[tp]) * a.tm.invw[tp]);
    b = min(max(b, 0), kTabBins - 1);
    int ti = a.tbins[tp * kTabBins + b];
    const int last = a.tm.last[tp];
    while (ti < last && r2 < g.coef[12 * ti]) ++ti;
    const double* t = g.coef + 12 * ti;
    const double dd = r2 - t[0];
    us = t[1] + dd * (t[2] + dd * (t[3] + dd * (t[4] + dd * (t[5] + dd * t[6]))));
    fs = t[7] + dd * (t[8] + dd * (t[9] + dd * (t[10] + dd * t[11])));
  }
#pragma unroll
  for (int q = 0; q < kMaxT; ++q) et[q] += q == tj ? us : 0.0;
  f[0] += fs * d[0];
  f[1] += fs * d[1];
  f[2] += fs * d[2];
  vir -= fs * r2;
}

__global__ __launch_bounds__(256) void k_pair_many(ManyArgs g) {
  const PairArgs& a = g.p;
  __shared__ double s_red[4][kMaxTP + 1];
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double et[kMaxT], vir = 0.0, f[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < kMaxT; ++q) et[q] = 0.0;
  int err = 0, ht = 0;
  if (s < a.np) {
    const double4 hp = g.verlet ? g.vl_pos[s] : a.cpos[s];
    const int hw = (int)hp.w;
    ht = hw & (kUncharged - 1);
    if (hw < kUncharged) {  // an uncharged home has no pair energy
      if (g.verlet) {
        const int n = g.vl_n[s];
        for (int k = 0; k < n; ++k) many_eval(g, hp, ht, g.vl_pos[g.vl_nbr[(size_t)k * a.np + s]], et, vir, f, err);
      } else {
        const int c = a.slot_cell[s];
        const int cx = c % a.g.nc[0], cy = (c / a.g.nc[0]) % a.g.nc[1], cz = c / (a.g.nc[0] * a.g.nc[1]);
        for (int k = 0; k < a.g.nst; ++k) {
          int x = cx + a.g.st[3 * k], y = cy + a.g.st[3 * k + 1], z = cz + a.g.st[3 * k + 2];
          x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
          y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
          z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
          const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
          const int j1 = a.cell_start[nb + 1];
          for (int j = a.cell_start[nb]; j < j1; ++j)
            if (j != s) many_eval(g, hp, ht, a.cpos[j], et, vir, f, err);
        }
      }
    }
    const int ip = g.verlet ? g.vl_home[s] : a.cell_ip[s];
    a.force[3 * ip + 0] = f[0] * a.scale;
    a.force[3 * ip + 1] = f[1] * a.scale;
    a.force[3 * ip + 2] = f[2] * a.scale;
  }
  if (err) atomicOr(a.flags, err);
  const int ntp = a.tm.ntp;
  for (int q = 0; q < ntp; ++q) {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxT; ++t) v += a.tm.tpidx[ht * kMaxT + t] == q ? et[t] : 0.0;
    v = msgd::wave_sum(v);
    if (lane == 0) s_red[wv][q] = v;
  }
  vir = msgd::wave_sum(vir);
  if (lane == 0) s_red[wv][ntp] = vir;
  __syncthreads();
  if (threadIdx.x <= ntp) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += s_red[w][threadIdx.x];
    a.part[blockIdx.x * (ntp + 1) + threadIdx.x] = t;
  }
}

}  // namespace

// Systems with more particle types than the LDS-table kernels take.
bool pair_many(const Ctx* c) { return c->npt > kLdsTypes; }

bool pair_nbpar(const Ctx* c) {
  // the mixed-path kernel codes a slot in 27 bits: larger systems take the work-item kernel
  return c->P.list_mode != MSG_LIST_VERLET && c->pair_lph > 0 && !pair_tiled(c) && !pair_many(c) &&
         (c->fp64 || c->np < (1 << kNbSlotBits));
}

// Cell-tile kernel (tile codes 100 + lanes per home): mixed path, cell lists.
bool pair_tiled(const Ctx* c) {
  return c->P.list_mode == MSG_LIST_CELL && !c->fp64 && c->pair_lph >= 100 && !pair_many(c);
}

// Lanes per home of the neighbour-parallel kernel (a cell-tile code on the FP64 path runs it
// with the same lanes).
static int nbpar_lanes(const Ctx* c) { return c->pair_lph >= 100 ? c->pair_lph - 100 : c->pair_lph; }

// Cells per tile along x: about 16 homes per tile at the mean density, at most 12 cells (the
// neighbourhood's 9 rows of tile + 2 cells fit the kernel's cell table), and few enough tiles
// for the pair partial rows reserved.
int pair_tile_tx(const Ctx* c) {
  const int nx = c->grid.nc[0];
  const double ppc = (double)c->np / std::max(1, c->grid.ncell);
  int tx = (int)std::lround(16.0 / std::max(ppc, 1e-9));
  tx = std::max(1, std::min(std::min(tx, 12), nx));
  const long rows = (long)c->grid.nc[1] * c->grid.nc[2];
  while (tx < std::min(12, nx) && rows * ((nx + tx - 1) / tx) > c->npart_pair_cap) ++tx;
  return tx;
}

// Default tile: about 16 candidates per lane, i.e. lanes per home = expected candidates
// (particles in the stencil cells) / 16, rounded up to a power of two in 8..64.  Measured
// on the 32 768-particle polyelectrolyte (~110 candidates per home), 8 lanes give the
// shortest UTotal step; all-pairs systems of ~1000 particles take 64.  The autotuner
// refines the tile per system size.
int default_pair_tile(const Ctx* c) {
  const double cand = (double)c->np * c->grid.nst / std::max(1, c->grid.ncell);
  int lph = 8;
  while (lph < 64 && lph * 16 < cand) lph *= 2;
  return lph;
}

int pair_blocks(Ctx* c) {
  if (c->P.list_mode == MSG_LIST_VERLET || pair_many(c)) return (c->np + 255) / 256;
  if (pair_tiled(c)) {
    const int tx = pair_tile_tx(c);
    return c->grid.nc[1] * c->grid.nc[2] * ((c->grid.nc[0] + tx - 1) / tx);
  }
  if (pair_nbpar(c) || c->pair_lph >= 100) {
    const int l = nbpar_lanes(c);
    return (c->np + 256 / l - 1) / (256 / l);
  }
  const long items = (long)c->grid.ncell + c->np / kPairHB + 1;
  return (int)std::max(1L, std::min((items + kWaves - 1) / kWaves, 2048L));
}

static PairArgs pair_args(Ctx* c, bool last) {
  PairArgs a;
  a.np = c->np;
  a.cpos = (const double4*)c->d_cpos;
  a.crel = (const float4*)c->d_crel;
  a.xcd = 1;
  a.ctype = c->d_ctype;
  a.cell_ip = c->d_cell_ip;
  a.cell_start = c->d_cell_start;
  a.slot_cell = c->d_slot_cell;
  a.cell_chunk = c->d_cell_chunk;
  a.item_cell = c->d_item_cell;
  a.g = c->grid;
  for (int d = 0; d < 3; ++d) {
    a.box.L[d] = c->P.boxlen[d];
    a.box.half[d] = 0.5 * c->P.boxlen[d];
    a.box.inv[d] = 1.0 / c->P.boxlen[d];
  }
  a.rcut2 = c->P.rcut * c->P.rcut;
  a.cut_inclusive = c->P.list_mode == MSG_LIST_ALL;
  a.tcoef = c->d_tcoef;
  a.tbins = c->d_tbins;
  a.tm = c->tmeta;
  a.force = c->d_force;
  a.part = c->d_part + c->part_off_pair;
  a.flags = c->d_flags;
  a.scale = last ? c->P.force_scale : 1.0;
  a.nblk = c->npart_pair;
  a.tile_tx = pair_tiled(c) ? pair_tile_tx(c) : 0;
  for (int d = 0; d < 3; ++d) a.csz[d] = (float)(1.0 / c->grid.csi[d]);
  a.ncinv[0] = 1.0f / (float)c->grid.nc[0];
  a.ncinv[1] = 1.0f / (float)(c->grid.nc[0] * c->grid.nc[1]);
  return a;
}

static size_t pair_lds(const Ctx* c) {
  return (size_t)c->tmeta.nint * 12 * (c->fp64 ? 8 : 4) + 8 + (size_t)c->tmeta.ntp * kTabBins * sizeof(int);
}

// The mixed path keeps the fixed-point Verlet records when every dimension of the list's cell
// grid has 3 to 512 cells (the cell difference of two listed particles is then -1..1 after
// wrapping, and a cell coordinate fits 9 bits) and the table is staged in LDS; its pair loop
// takes them under the skin rule (inlist = 0), which bounds every particle's drift since the
// build to the skin, so the records stay far inside 32 bits.
bool vl_rel_ok(const Ctx* c) {
  if (c->fp64 || !c->d_vl_rel || pair_many(c)) return false;
  for (int d = 0; d < 3; ++d)
    if (c->grid.nc[d] < 3 || c->grid.nc[d] > 512) return false;
  return true;
}

int launch_pair(Ctx* c, bool last) {
  const PairArgs a = pair_args(c, last);
  const int nb = c->npart_pair;
  const size_t lds = pair_lds(c);
  prof_begin(c, MSG_K_PAIR);
  if (pair_many(c)) {
    ManyArgs g;
    g.p = a;
    g.coef = c->d_tcoef64;
    g.verlet = c->P.list_mode == MSG_LIST_VERLET;
    if (g.verlet) g.p.cut_inclusive = 1;  // UTwoBodyA's r2 > rcut2 test
    g.vl_home = c->d_vl_home;
    g.vl_n = c->d_vl_n;
    g.vl_nbr = c->d_vl_nbr;
    g.vl_pos = (const double4*)c->d_vl_pos;
    hipLaunchKernelGGL(k_pair_many, dim3(nb), dim3(256), 0, c->stream, g);
    prof_end(c, MSG_K_PAIR);
    MSG_HIP(hipGetLastError());
    return MSG_OK;
  }
  const bool nbpar = pair_nbpar(c);
  if (c->P.list_mode == MSG_LIST_VERLET && vl_rel_ok(c) && c->inlist == 0) {
    VlRelArgs v;
    v.p = a;
    v.p.cut_inclusive = 1;
    v.vl_home = c->d_vl_home;
    v.vl_n = c->d_vl_n;
    v.vl_nbr = c->d_vl_nbr;
    v.vl_rel = (const int4*)c->d_vl_rel;
    for (int d = 0; d < 3; ++d) {
      v.unit[d] = (float)(c->P.boxlen[d] / c->grid.nc[d] / kVlUnits);
      v.nc[d] = c->grid.nc[d];
    }
    switch (c->nptpt) {
      case 1: hipLaunchKernelGGL((k_pair_vl_rel<1>), dim3(nb), dim3(256), lds, c->stream, v); break;
      case 3: hipLaunchKernelGGL((k_pair_vl_rel<3>), dim3(nb), dim3(256), lds, c->stream, v); break;
      case 6: hipLaunchKernelGGL((k_pair_vl_rel<6>), dim3(nb), dim3(256), lds, c->stream, v); break;
      default: hipLaunchKernelGGL((k_pair_vl_rel<10>), dim3(nb), dim3(256), lds, c->stream, v); break;
    }
    prof_end(c, MSG_K_PAIR);
    MSG_HIP(hipGetLastError());
    return MSG_OK;
  }
  if (c->P.list_mode == MSG_LIST_VERLET) {
    VlPairArgs v;
    v.p = a;
    v.p.cut_inclusive = 1;  // UTwoBodyA's r2 > rcut2 test
    v.maxn = c->vl_maxn;
    v.vl_home = c->d_vl_home;
    v.vl_n = c->d_vl_n;
    v.vl_nbr = c->d_vl_nbr;
    v.vl_pos = (const double4*)c->d_vl_pos;
    switch (c->nptpt) {
#define MSG_VL(NTP)                                                                            \
  if (c->fp64) hipLaunchKernelGGL((k_pair_vl<double, NTP>), dim3(nb), dim3(256), lds, c->stream, v); \
  else hipLaunchKernelGGL((k_pair_vl<float, NTP>), dim3(nb), dim3(256), lds, c->stream, v);
      case 1: MSG_VL(1) break;
      case 3: MSG_VL(3) break;
      case 6: MSG_VL(6) break;
      default: MSG_VL(10) break;
#undef MSG_VL
    }
    prof_end(c, MSG_K_PAIR);
    MSG_HIP(hipGetLastError());
    return MSG_OK;
  }
  const int lph = nbpar_lanes(c);
  if (pair_tiled(c)) {
#define MSG_TILE(NTP)                                                                             \
  if (lph == 8) hipLaunchKernelGGL((k_pair_tile<NTP, 8>), dim3(nb), dim3(256), lds, c->stream, a); \
  else hipLaunchKernelGGL((k_pair_tile<NTP, 16>), dim3(nb), dim3(256), lds, c->stream, a);
    switch (c->nptpt) {
      case 1: MSG_TILE(1) break;
      case 3: MSG_TILE(3) break;
      case 6: MSG_TILE(6) break;
      default: MSG_TILE(10) break;
    }
#undef MSG_TILE
    prof_end(c, MSG_K_PAIR);
    MSG_HIP(hipGetLastError());
    return MSG_OK;
  }
  const bool nbp = nbpar || c->pair_lph >= 100;
#define MSG_PAIR_LAUNCH(R, NTP)                                                              \
  do {                                                                                       \
    if (nbp && lph == 4)                                                                     \
      hipLaunchKernelGGL((k_pair_nb<R, NTP, 4>), dim3(nb), dim3(256), lds, c->stream, a);   \
    else if (nbp && lph == 8)                                                                \
      hipLaunchKernelGGL((k_pair_nb<R, NTP, 8>), dim3(nb), dim3(256), lds, c->stream, a);   \
    else if (nbp && lph == 16)                                                               \
      hipLaunchKernelGGL((k_pair_nb<R, NTP, 16>), dim3(nb), dim3(256), lds, c->stream, a);  \
    else if (nbp && lph == 32)                                                               \
      hipLaunchKernelGGL((k_pair_nb<R, NTP, 32>), dim3(nb), dim3(256), lds, c->stream, a);  \
    else if (nbp)                                                                            \
      hipLaunchKernelGGL((k_pair_nb<R, NTP, 64>), dim3(nb), dim3(256), lds, c->stream, a);  \
    else                                                                                     \
      hipLaunchKernelGGL((k_pair<R, NTP>), dim3(nb), dim3(256), lds, c->stream, a);         \
  } while (0)
  switch (c->nptpt) {
    case 1: if (c->fp64) MSG_PAIR_LAUNCH(double, 1); else MSG_PAIR_LAUNCH(float, 1); break;
    case 3: if (c->fp64) MSG_PAIR_LAUNCH(double, 3); else MSG_PAIR_LAUNCH(float, 3); break;
    case 6: if (c->fp64) MSG_PAIR_LAUNCH(double, 6); else MSG_PAIR_LAUNCH(float, 6); break;
    default: if (c->fp64) MSG_PAIR_LAUNCH(double, 10); else MSG_PAIR_LAUNCH(float, 10); break;
  }
#undef MSG_PAIR_LAUNCH
  prof_end(c, MSG_K_PAIR);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg

namespace msg {

int launch_vlist_build(Ctx* c) {
  VlistArgs v;
  v.np = c->np;
  v.maxn = c->vl_maxn;
  v.cpos = (const double4*)c->d_cpos;
  v.cell_ip = c->d_cell_ip;
  v.cell_start = c->d_cell_start;
  v.slot_cell = c->d_slot_cell;
  v.g = c->grid;
  for (int d = 0; d < 3; ++d) {
    v.box.L[d] = c->P.boxlen[d];
    v.box.half[d] = 0.5 * c->P.boxlen[d];
    v.box.inv[d] = 1.0 / c->P.boxlen[d];
  }
  const double rl = c->P.rcut + c->P.drnlist;
  v.rl2 = rl * rl;
  v.ro_ref = c->d_ro_ref;
  v.vl_home = c->d_vl_home;
  v.vl_n = c->d_vl_n;
  v.vl_nbr = c->d_vl_nbr;
  v.vl_pos = (double4*)c->d_vl_pos;
  v.vl_rel = vl_rel_ok(c) ? (int4*)c->d_vl_rel : nullptr;
  v.vl_slot = c->d_vl_slot;
  v.flags = c->d_flags;
  prof_begin(c, MSG_K_VLIST);
  const size_t lds = (size_t)(256 / kVlLph) * (c->vl_maxn + 1) * sizeof(int);
  v.stage = lds <= (size_t)kVlStageBytes;
  hipLaunchKernelGGL(k_vlist_build, dim3((c->np + 256 / kVlLph - 1) / (256 / kVlLph)), dim3(256),
                     v.stage ? lds : 0, c->stream, v);
  prof_end(c, MSG_K_VLIST);
  MSG_HIP(hipGetLastError());
  c->vl_builds++;
  c->drosum_reset = true;  // the displacement sums start again with the new list
  return MSG_OK;
}

}  // namespace msg

namespace msg {

namespace {

template <typename real, int NTP, int LPH>
void seg_pair_nb(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_pair_nb_seg<real, NTP, LPH>), dim3(g.nblk), dim3(256), g.lds, s,
                     (const PairArgs*)g.d_args, g.d_map);
}

template <typename real, int NTP>
void seg_pair_items(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_pair_seg<real, NTP>), dim3(g.nblk), dim3(256), g.lds, s,
                     (const PairArgs*)g.d_args, g.d_map);
}

template <typename real, int NTP>
SegLaunchFn pair_fn(int lph) {
  switch (lph) {
    case 4: return seg_pair_nb<real, NTP, 4>;
    case 8: return seg_pair_nb<real, NTP, 8>;
    case 16: return seg_pair_nb<real, NTP, 16>;
    case 32: return seg_pair_nb<real, NTP, 32>;
    case 64: return seg_pair_nb<real, NTP, 64>;
    default: return seg_pair_items<real, NTP>;
  }
}

template <int NTP, int LPH>
void seg_pair_tile(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_pair_tile_seg<NTP, LPH>), dim3(g.nblk), dim3(256), g.lds, s,
                     (const PairArgs*)g.d_args, g.d_map);
}

SegLaunchFn pair_tile_fn(int ntp, int lph) {
  switch (ntp) {
    case 1: return lph == 8 ? seg_pair_tile<1, 8> : seg_pair_tile<1, 16>;
    case 3: return lph == 8 ? seg_pair_tile<3, 8> : seg_pair_tile<3, 16>;
    case 6: return lph == 8 ? seg_pair_tile<6, 8> : seg_pair_tile<6, 16>;
    default: return lph == 8 ? seg_pair_tile<10, 8> : seg_pair_tile<10, 16>;
  }
}

template <typename real>
SegLaunchFn pair_fn_ntp(int ntp, int lph) {
  switch (ntp) {
    case 1: return pair_fn<real, 1>(lph);
    case 3: return pair_fn<real, 3>(lph);
    case 6: return pair_fn<real, 6>(lph);
    default: return pair_fn<real, 10>(lph);
  }
}

}  // namespace

// Pair stage of a batch: one launch per (precision, type-pair count, tile).  The forces are
// not scaled here (the gather, or the reduction-free path, is the last writer).
int plan_pair(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& out) {
  SegBuilder<PairArgs> b;
  for (Ctx* c : cs) {
    if (c->P.list_mode == MSG_LIST_VERLET) return set_error(MSG_ERR_ARG, "batch: Verlet-list systems run unbatched");
    if (pair_many(c)) return set_error(MSG_ERR_ARG, "batch: systems of more than 4 particle types run unbatched");
    const int ntp = c->nptpt <= 1 ? 1 : c->nptpt <= 3 ? 3 : c->nptpt <= 6 ? 6 : 10;
    if (pair_tiled(c)) {
      const int key = 200000 + ntp * 1000 + c->pair_lph;
      b.add(key, pair_tile_fn(ntp, nbpar_lanes(c)), pair_args(c, c->P.ewald_mode == MSG_EWALD_NONE), c->npart_pair,
            pair_lds(c));
      continue;
    }
    const int lph = (pair_nbpar(c) || c->pair_lph >= 100) ? nbpar_lanes(c) : 0;
    const int key = (c->fp64 ? 100000 : 0) + ntp * 100 + lph;
    SegLaunchFn fn = c->fp64 ? pair_fn_ntp<double>(ntp, lph) : pair_fn_ntp<float>(ntp, lph);
    b.add(key, fn, pair_args(c, c->P.ewald_mode == MSG_EWALD_NONE), c->npart_pair, pair_lds(c));
  }
  return b.finish(own, out);
}

}  // namespace msg

using namespace msg;

extern "C" {

// The current Verlet list in MOLSIM's layout (kind 0: SetVListMD, 1: SetVListMC); built for
// the current positions first if none is valid (as msg_nlist).
int msg_get_verlet_list(msg_ctx* h, int kind, int maxnneigh, int* nneighpn, int* jpnlist) {
  if (!h || (kind != 0 && kind != 1) || maxnneigh < 1 || !nneighpn || !jpnlist)
    return set_error(MSG_ERR_ARG, "msg_get_verlet_list: bad arguments");
  Ctx* c = &h->c;
  if (c->P.list_mode != MSG_LIST_VERLET) return set_error(MSG_ERR_STATE, "msg_get_verlet_list: not in Verlet-list mode");
  if (!c->vl_valid)
    if (int e = msg_nlist(h)) return e;
  const int width = c->vl_maxn;
  const int wpb = width <= 4096 ? 4 : 1;
  if ((size_t)wpb * width * sizeof(int) > 65536) return set_error(MSG_ERR_ARG, "msg_get_verlet_list: list width above 16384");
  const int np = c->np;
  int *d_nn = nullptr, *d_jl = nullptr;
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipMalloc(&d_nn, np * sizeof(int)));
  if (hipMalloc(&d_jl, (size_t)np * maxnneigh * sizeof(int)) != hipSuccess) {
    hipFree(d_nn);
    return set_error(MSG_ERR_NOMEM, "msg_get_verlet_list: no device memory for the copy");
  }
  VlExportArgs a;
  a.np = np;
  a.kind = kind;
  a.ldo = maxnneigh;
  a.width = width;
  a.wpb = wpb;
  a.vl_home = c->d_vl_home;
  a.vl_n = c->d_vl_n;
  a.vl_nbr = c->d_vl_nbr;
  a.nneigh = d_nn;
  a.jpn = d_jl;
  a.flags = c->d_flags + 2;  // a word of its own: an overflow of the caller's width is this call's error
  int e = MSG_OK;
  hipLaunchKernelGGL(k_vlist_export, dim3((np + wpb - 1) / wpb), dim3(64 * wpb), (size_t)wpb * width * sizeof(int),
                     c->stream, a);
  hipError_t he = hipGetLastError();
  int ovf = 0;
  if (he == hipSuccess) he = hipMemcpyAsync(nneighpn, d_nn, np * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(jpnlist, d_jl, (size_t)np * maxnneigh * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(&ovf, c->d_flags + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipMemsetAsync(c->d_flags + 2, 0, sizeof(int), c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  hipFree(d_nn);
  hipFree(d_jl);
  if (he != hipSuccess) return set_error(MSG_ERR_HIP, std::string("msg_get_verlet_list: ") + hipGetErrorString(he));
  if (ovf) e = set_error(MSG_ERR_NLIST, "msg_get_verlet_list: a list is longer than maxnneigh");
  return e;
}

}  // extern "C"
