// UpdateCellIp on the device (Src/celllist.F90:329-382): after an accepted move the particle
// leaves its cell's run of slots and becomes the first slot (the head, AddIpToCell :307-327) of
// its new cell; the other particles keep their order.  The cell list is a counting-sort layout
// (cell_start offsets, one run of slots per cell), so the particles between the old slot and
// the new cell's head move by one slot and the offsets of the cells in between by one.  One
// workgroup of NT threads does it, in chunks of NT slots (all of a chunk's slots are read before
// any is written, and chunks go in the direction that never overwrites an unread slot).
#pragma once

#include "common.h"

namespace msg {

struct CellMoveArgs {
  CellGrid g;
  Box box;
  int* cell_start;
  int* cell_ip;
  double4* cpos;
  int* ctype;
  int* slot_of;
  int* slot_cell;
  float4* crel;   // mixed path: cell-relative FP32 records (null on the FP64 path)
};

// A position's FP32 record relative to the lower corner of cell c (as the cell sort writes it).
__device__ inline float4 cell_rel(const CellGrid& g, const Box& box, int c, const double* r, double w) {
  const int cc[3] = {c % g.nc[0], (c / g.nc[0]) % g.nc[1], c / (g.nc[0] * g.nc[1])};
  float v[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] = (float)(r[d] + box.half[d] - cc[d] / g.csi[d]);
  return make_float4(v[0], v[1], v[2], (float)w);
}

// Cell of a position as pcellro (Src/celllist.F90:295-305), clamped as the binning clamps.
__device__ inline int cell_of_pos(const CellGrid& g, const Box& box, const double* r) {
  int ci[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int k = (int)floor((r[d] + box.half[d]) * g.csi[d]);
    ci[d] = min(max(k, 0), g.nc[d] - 1);
  }
  return ci[0] + g.nc[0] * (ci[1] + g.nc[1] * ci[2]);
}

// Particle ip moved to rn: its cell-ordered record follows, and if its cell changed the slots
// shift as described above.  Every thread of the workgroup calls it; returns 1 when the cell
// changed.  Global writes are made visible to the workgroup by the barriers.
template <int NT>
__device__ int cell_move(const CellMoveArgs& a, int ip, const double* rn) {
  __shared__ int s_old, s_A, s_B;
  if (threadIdx.x == 0) {
    s_old = a.slot_of[ip];
    s_A = a.slot_cell[s_old];
    s_B = cell_of_pos(a.g, a.box, rn);
  }
  __syncthreads();
  const int so = s_old, A = s_A, B = s_B;
  const double4 rec = a.cpos[so];
  const int tp = a.ctype[so];
  const double4 moved = make_double4(rn[0], rn[1], rn[2], rec.w);
  if (A == B) {
    __syncthreads();
    if (threadIdx.x == 0) {
      a.cpos[so] = moved;
      if (a.crel) a.crel[so] = cell_rel(a.g, a.box, B, rn, rec.w);
    }
    __syncthreads();
    return 0;
  }
  int dst;
  if (A < B) {
    dst = a.cell_start[B] - 1;  // B's first slot once the run (so, dst] has moved down
    __syncthreads();
    for (int s0 = so + 1; s0 <= dst; s0 += NT) {
      const int s = s0 + (int)threadIdx.x;
      const bool v = s <= dst;
      int ci = 0, ct = 0, sc = 0;
      double4 p = make_double4(0, 0, 0, 0);
      float4 pr = make_float4(0, 0, 0, 0);
      if (v) {
        ci = a.cell_ip[s];
        ct = a.ctype[s];
        sc = a.slot_cell[s];
        p = a.cpos[s];
        if (a.crel) pr = a.crel[s];
      }
      __syncthreads();
      if (v) {
        a.cell_ip[s - 1] = ci;
        a.ctype[s - 1] = ct;
        a.slot_cell[s - 1] = sc;
        a.cpos[s - 1] = p;
        if (a.crel) a.crel[s - 1] = pr;
        a.slot_of[ci] = s - 1;
      }
      __syncthreads();
    }
    for (int c = A + 1 + (int)threadIdx.x; c <= B; c += NT) a.cell_start[c] -= 1;
  } else {
    dst = a.cell_start[B];      // B's first slot; the run [dst, so) moves up
    __syncthreads();
    for (int s1 = so - 1; s1 >= dst; s1 -= NT) {
      const int s = s1 - (int)threadIdx.x;
      const bool v = s >= dst;
      int ci = 0, ct = 0, sc = 0;
      double4 p = make_double4(0, 0, 0, 0);
      float4 pr = make_float4(0, 0, 0, 0);
      if (v) {
        ci = a.cell_ip[s];
        ct = a.ctype[s];
        sc = a.slot_cell[s];
        p = a.cpos[s];
        if (a.crel) pr = a.crel[s];
      }
      __syncthreads();
      if (v) {
        a.cell_ip[s + 1] = ci;
        a.ctype[s + 1] = ct;
        a.slot_cell[s + 1] = sc;
        a.cpos[s + 1] = p;
        if (a.crel) a.crel[s + 1] = pr;
        a.slot_of[ci] = s + 1;
      }
      __syncthreads();
    }
    for (int c = B + 1 + (int)threadIdx.x; c <= A; c += NT) a.cell_start[c] += 1;
  }
  if (threadIdx.x == 0) {
    a.cell_ip[dst] = ip;
    a.ctype[dst] = tp;
    a.slot_cell[dst] = B;
    a.cpos[dst] = moved;
    if (a.crel) a.crel[dst] = cell_rel(a.g, a.box, B, rn, rec.w);
    a.slot_of[ip] = dst;
  }
  __syncthreads();
  return 1;
}

}  // namespace msg
