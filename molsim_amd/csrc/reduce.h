// Final FP64 reduction of a UTotal pass (included by capi.hip and spme.hip, which runs it
// as one extra workgroup of the gather launch).
#pragma once

#include "common.h"

namespace msg {

struct ReduceArgs {
  const double* pair;  int npair; int ntp;   // npair rows of ntp energies + virial (1: pre-summed)
  const double* spme;  int nspme;
  const double* ewstd; int newstd;
  const double* M;     // total dipole sum q r, formed by k_scan
  int ewald;           // 0 none, 1 std, 2 spm
  int lsurf;
  double self_raw;     // -alpha/sqrt(pi) sum q^2
  double surf_fac;     // 2 pi / (3 V)
  double epsi4pi;
  double* out;
  int* flags;
  int* cell_count;
  int ncell;
  int* brick_count;
  int nbrick;          // brick sort counters to clear (MeshGrid::nkey)
};

// Fixed-order column totals of the pair kernel's rows (np1 columns: energies by type pair +
// virial, one row per workgroup) by one 1024-thread workgroup.  Thread t owns column
// t mod np1 and adds rows t div np1, t div np1 + R, ... (R = 1024 div np1 threads per column,
// four loads in flight); the R partial sums of a column are then added in a fixed binary tree
// in LDS.  The totals end in red[0 .. np1).  With one row the totals are the row itself.
__device__ inline void pair_rows_total(const double* __restrict__ rows, int nrow, int np1,
                                                double* red) {
  const int R = 1024 / np1;
  const int c = threadIdx.x % np1, r0 = threadIdx.x / np1;
  double s = 0.0;
  if (r0 < R) {
    for (int r = r0; r < nrow; r += 4 * R) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = r + u * R < nrow ? rows[(size_t)(r + u * R) * np1 + c] : 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u) s += x[u];
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  int p2 = 1;
  while (p2 < R) p2 <<= 1;
  for (int st = p2 >> 1; st > 0; st >>= 1) {
    if (r0 < st && r0 + st < R) red[threadIdx.x] += red[threadIdx.x + st * np1];
    __syncthreads();
  }
}

// Final FP64 reduction of all partial sums in a fixed order (deterministic).  It is the last
// kernel of a UTotal pass, so it also hands the error flags to the output block and clears
// the cell/brick counters and flags for the next pass.  The k-space partials, dipole, flags
// and a pre-summed pair row are loaded together at the start.
//
// NT threads (1024, or 256 when the reduction rides as one extra workgroup of the gather
// launch): the k-space partials are added in the order of a 1024-thread workgroup whatever
// NT is (virtual thread v = t + NT*k takes partials v, v + 1024, ...), so both forms give
// the same bits.  With NT < 1024 the pair rows must be pre-summed (npair == 1).
template <int NT>
__device__ inline void reduce_body(const ReduceArgs& a) {
  constexpr int R = 1024 / NT, NW = NT / 64;
  __shared__ double red[NT == 1024 ? 1024 : 1];
  __shared__ double wred[16][2];
  const int tid = threadIdx.x;
  const double* ew = a.ewald == 2 ? a.spme : a.ewstd;
  const int nev = a.ewald == 2 ? a.nspme : (a.ewald == 1 ? a.newstd : 0);
  const int np1 = a.ntp + 1;
  double m[3] = {0.0, 0.0, 0.0};
  int fl = 0;
  if (tid == 0) {
    m[0] = a.M[0];
    m[1] = a.M[1];
    m[2] = a.M[2];
    fl = a.flags[0];
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int v = tid + NT * k;
    double e0 = 0.0, e1 = 0.0;
    if (v < nev) {
      e0 = ew[2 * v];
      e1 = ew[2 * v + 1];
    }
    for (int b = v + 1024; b < nev; b += 1024) {
      e0 += ew[2 * b];
      e1 += ew[2 * b + 1];
    }
    e0 = msgd::wave_sum(e0);
    e1 = msgd::wave_sum(e1);
    if ((tid & 63) == 0) {
      wred[(tid >> 6) + NW * k][0] = e0;
      wred[(tid >> 6) + NW * k][1] = e1;
    }
  }
  // the pair row totals: one row already, or summed here (into LDS)
  const double* pt = a.pair;
  if constexpr (NT == 1024) {
    if (a.npair != 1) {
      pair_rows_total(a.pair, a.npair, np1, red);
      pt = red;
    }
  }
  __syncthreads();
  if (tid == 0) {
    double es = 0.0, vs = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      es += wred[w][0];
      vs += wred[w][1];
    }
    double* o = a.out;
    double tw = 0.0;
    for (int tp = 0; tp < a.ntp; ++tp) {
      o[1 + tp] = 0.5 * pt[tp];  // every pair was evaluated from both of its particles
      tw += 0.5 * pt[tp];
    }
    o[0] = tw;
    double urec = 0.0, virrec = 0.0;
    if (a.ewald) {
      // SPME: E = sum w G |F|^2, virrec -= sum w G (1 + 2 m^2 expf) |F|^2
      // std:  E = sum kfac |S|^2, virrec -= sum (1 - k^2/(2 alpha^2)) kfac |S|^2
      urec = es + a.self_raw;
      virrec = -vs;
      if (a.lsurf) {
        const double term = a.surf_fac * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        urec += term;
        virrec -= term;
      }
    }
    o[a.ntp + 1] = a.epsi4pi * urec;
    o[a.ntp + 2] = tw + a.epsi4pi * urec;
    o[a.ntp + 3] = 0.5 * pt[a.ntp] + a.epsi4pi * virrec;
    o[a.ntp + 4] = m[0];
    o[a.ntp + 5] = m[1];
    o[a.ntp + 6] = m[2];
    o[a.ntp + 7] = (double)fl;
    a.flags[0] = 0;
    a.flags[1] |= fl;  // sticky over the passes of a multi-step call
  }
  for (int i = tid; i < a.ncell; i += NT) a.cell_count[i] = 0;
  for (int i = tid; i < a.nbrick; i += NT) a.brick_count[i] = 0;
}

// The reduction's arguments for a context (capi.hip); presum: the pair rows were summed into
// one row by k_pair_sum.
ReduceArgs reduce_args(Ctx* c, bool presum);

}  // namespace msg
