// K11-K14: Monte Carlo energy differences of single-particle trial moves (DUTotal,
// Src/denergy.F90:98-279) and the commit of an accepted move (MCUpdate,
// Src/mc.F90:5203-5235), all in FP64.
//
// One workgroup per trial; any number of independent trials against the same state run in
// one launch.  Per trial:
//   real space   (UTwoBodyANew/OldCellList, Src/denergy.F90:728-928): the stencil cells of the
//                trial position's cell and of the particle's cell; hard-core or below-table
//                pairs at the new position flag an overlap; FP64 table records.
//   SPME         the reference spreads the moved charge into a copy of the mesh and runs a
//                full r2c per trial (DUTwoBodyEwaldRecSPM, Src/denergy.F90:1413-1495).  Here
//                the same energy difference is evaluated exactly and locally:
//                  dE = sum_m dQ(m) phi(m) + 1/2 sum_{m,m'} dQ(m) g(m - m') dQ(m')
//                where phi = g * Q is the convolved mesh UTotal leaves behind and g the real
//                -space influence kernel (inverse transform of meshfac).  dQ has 2 p^3 points,
//                so a trial costs O(p^6) instead of O(K^3 log K).
//   std Ewald    dS(k) = q (e^{ik.r_new} - e^{ik.r_old}) on the resident structure factors,
//                dE = sum kfac (|S + dS|^2 - |S|^2) (DUTwoBodyEwaldRecStd, :1275-1335).
//   surface      (|M + q dr|^2 - |M|^2) 2 pi / (3 V)  (DUTwoBodyEwaldSurf, :1504-1539).
// The state (cell list, phi or S, dipole M) must match the current positions: after a
// commit or new positions it is rebuilt by a UTotal pass before the next trials.
#include <mutex>

#include "common.h"
#include "cellmove.h"

namespace msg {

namespace {

struct DeArgs {
  int np, ntp;
  const double* ro;
  const double* az;
  const int* ipt;
  // cell list
  CellGrid g;
  Box box;
  const int* cell_start;
  const int* cell_ip;
  const double4* cpos;
  const int* ctype;
  const int* slot_of;
  const int* slot_cell;
  double rcut2;
  int cut_inclusive;
  // table (FP64)
  const double* tcoef;
  const int* tbins;
  GpuTableMeta tm;
  // Ewald
  int ewald;
  int lsurf;
  double surf_fac;
  const double* Mdip;     // 3 doubles (d_out + ntp + 4)
  // SPME
  MeshGrid m;
  int order;
  const void* phi;        // K^3 reals (float or double)
  int phi_fp64;
  const double* gk;       // H^3 folded real-space kernel
  // std
  int nk;
  const int* kvec;
  const double* kfac;
  const double* sk;
  double tpb[3];
  double epsi4pi;
  // trials
  const int* tip;
  int tip_base;           // 0, or 1 when tip holds the caller's 1-based indices
  const double* tr;
  double* out;            // per trial: ntp + 4 doubles
  const int* laz;         // weak charges (null when off)
};

__device__ inline double table_u(const DeArgs& a, int tp, double r2) {
  int b = (int)((r2 - a.tm.zlo[tp]) * a.tm.invw[tp]);
  b = min(max(b, 0), kTabBins - 1);
  int idx = a.tbins[tp * kTabBins + b];
  const int last = a.tm.last[tp];
  while (idx < last && r2 < a.tcoef[12 * idx]) ++idx;
  const double* t = a.tcoef + 12 * idx;
  const double d = r2 - t[0];
  return t[1] + d * (t[2] + d * (t[3] + d * (t[4] + d * (t[5] + d * t[6]))));
}

template <int P>
__device__ inline void spline64(double pos, int K, int& mm, double* th) {
  const double v = floor(pos);
  msgd::bspline<double, P>(pos - v, th, (double*)nullptr);
  int k = (int)v % K;
  mm = k < 0 ? k + K : k;
}

__device__ inline int fold(int d, int K) {
  d %= K;
  if (d < 0) d += K;
  return d <= K / 2 ? d : K - d;
}

// One trial: particle ip (0-based) to position trn[0..2]; results to o[0 .. ntp+3].
template <int P>
__device__ __forceinline__ void de_body(const DeArgs& a, int ip, const double* trn, double* o) {
  constexpr int W = 2 * P - 1;
  // windows of g staged in LDS up to order 8; above, they are read from g itself
  constexpr bool kStage = P <= 8;
  constexpr int W3 = kStage ? W * W * W : 1;
  __shared__ double sh[8];
  __shared__ double thn[3][P], tho[3][P];
  __shared__ int mmn[3], mmo[3];
  __shared__ double w0[W3], w1[W3];
  __shared__ int s_ovl;
  const double rn[3] = {trn[0], trn[1], trn[2]};
  const double ro[3] = {a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2]};
  const double q = a.az[ip];
  const int ti = a.ipt[ip];
  if (threadIdx.x == 0) s_ovl = 0;
  __syncthreads();

  // ---------------- real space: new minus old ----------------
  double e[kMaxT];
#pragma unroll
  for (int k = 0; k < kMaxT; ++k) e[k] = 0.0;
  int ovl = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const double* r = pass == 0 ? rn : ro;
    int cc;
    if (pass == 0) {
      int ci[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int k = (int)floor((r[d] + a.box.half[d]) * a.g.csi[d]);
        ci[d] = min(max(k, 0), a.g.nc[d] - 1);
      }
      cc = ci[0] + a.g.nc[0] * (ci[1] + a.g.nc[1] * ci[2]);
    } else {
      cc = a.slot_cell[a.slot_of[ip]];
    }
    const int cx = cc % a.g.nc[0], cy = (cc / a.g.nc[0]) % a.g.nc[1], cz = cc / (a.g.nc[0] * a.g.nc[1]);
    const double sgn = pass == 0 ? 1.0 : -1.0;
    for (int k = 0; k < a.g.nst; ++k) {
      int x = cx + a.g.st[3 * k], y = cy + a.g.st[3 * k + 1], z = cz + a.g.st[3 * k + 2];
      x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
      y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
      z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
      const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
      const int j1 = a.cell_start[nb + 1];
      for (int j = a.cell_start[nb] + threadIdx.x; j < j1; j += 256) {
        if (a.cell_ip[j] == ip) continue;
        const double4 pj = a.cpos[j];
        double d[3] = {r[0] - pj.x, r[1] - pj.y, r[2] - pj.z};
#pragma unroll
        for (int qd = 0; qd < 3; ++qd)
          if (fabs(d[qd]) > a.box.half[qd]) d[qd] -= copysign(a.box.L[qd], d[qd]);
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (a.cut_inclusive ? (r2 > a.rcut2) : !(r2 < a.rcut2)) continue;
        const int tj = a.ctype[j];
        const int tp = a.tm.tpidx[ti * kMaxT + tj];
        if (pass == 0 && (r2 < a.tm.r2atat[tp] || r2 < a.tm.r2umin[tp])) {
          ovl = 1;
          continue;
        }
        // weak charges: UWeakChargeANew/AOld skip an uncharged pair after the overlap checks
        // (Src/denergy.F90:3955-3959, 4066-4068); the partner's state is in its record
        if (a.laz && (!a.laz[ip] || (int)pj.w >= kUncharged)) continue;
        const double u = table_u(a, tp, r2);
#pragma unroll
        for (int k2 = 0; k2 < kMaxT; ++k2) e[k2] += (k2 == tj) ? sgn * u : 0.0;
      }
    }
  }
  if (ovl) atomicOr(&s_ovl, 1);

  // ---------------- reciprocal space ----------------
  double drec = 0.0;
  if (a.ewald == 2) {
    const int K = a.m.K;
    if (threadIdx.x < 3) {
      const int d = threadIdx.x;
      spline64<P>(a.m.dkdr[d] * rn[d], K, mmn[d], thn[d]);
      spline64<P>(a.m.dkdr[d] * ro[d], K, mmo[d], tho[d]);
    }
    __syncthreads();
    // windows of g around offset 0 and around mm_new - mm_old
    const int H = a.m.H;
    if constexpr (kStage) {
      for (int k = threadIdx.x; k < W3; k += 256) {
        const int ox = k % W - (P - 1), oy = (k / W) % W - (P - 1), oz = k / (W * W) - (P - 1);
        w0[k] = a.gk[fold(ox, K) + H * (fold(oy, K) + H * fold(oz, K))];
        w1[k] = a.gk[fold(mmn[0] - mmo[0] + ox, K) +
                     H * (fold(mmn[1] - mmo[1] + oy, K) + H * fold(mmn[2] - mmo[2] + oz, K))];
      }
    }
    __syncthreads();
    // term 1: sum dQ phi over the two stencils; term 2: pair sums with g
    double s1 = 0.0, s2 = 0.0;
    constexpr int P3 = P * P * P;
    for (int a_ = threadIdx.x; a_ < P3; a_ += 256) {
      const int ix = a_ % P, iy = (a_ / P) % P, iz = a_ / (P * P);
      const double tn = thn[0][ix] * thn[1][iy] * thn[2][iz];
      const double to = tho[0][ix] * tho[1][iy] * tho[2][iz];
      int xn = mmn[0] - ix, yn = mmn[1] - iy, zn = mmn[2] - iz;
      int xo = mmo[0] - ix, yo = mmo[1] - iy, zo = mmo[2] - iz;
      xn += xn < 0 ? K : 0; yn += yn < 0 ? K : 0; zn += zn < 0 ? K : 0;
      xo += xo < 0 ? K : 0; yo += yo < 0 ? K : 0; zo += zo < 0 ? K : 0;
      const size_t in = xn + (size_t)K * (yn + (size_t)K * zn), io = xo + (size_t)K * (yo + (size_t)K * zo);
      const double pn = a.phi_fp64 ? ((const double*)a.phi)[in] : (double)((const float*)a.phi)[in];
      const double po = a.phi_fp64 ? ((const double*)a.phi)[io] : (double)((const float*)a.phi)[io];
      s1 += tn * pn - to * po;
      // point a_ of each stencil against every point b of both stencils:
      //   offset (new a - new b) = (old a - old b) = (ib - ia) in stencil units (window w0)
      //   offset (new a - old b) = mmn - mmo + (ib - ia)                       (window w1)
      double nn = 0.0, oo = 0.0, no = 0.0;
      for (int b = 0; b < P3; ++b) {
        const int jx = b % P, jy = (b / P) % P, jz = b / (P * P);
        const double tbn = thn[0][jx] * thn[1][jy] * thn[2][jz];
        const double tbo = tho[0][jx] * tho[1][jy] * tho[2][jz];
        double g0, g1;
        if constexpr (kStage) {
          const int k = (jx - ix + P - 1) + W * ((jy - iy + P - 1) + W * (jz - iz + P - 1));
          g0 = w0[k];
          g1 = w1[k];
        } else {
          g0 = a.gk[fold(jx - ix, K) + H * (fold(jy - iy, K) + H * fold(jz - iz, K))];
          g1 = a.gk[fold(mmn[0] - mmo[0] + jx - ix, K) +
                    H * (fold(mmn[1] - mmo[1] + jy - iy, K) + H * fold(mmn[2] - mmo[2] + jz - iz, K))];
        }
        nn += tbn * g0;
        oo += tbo * g0;
        no += tbo * g1;
      }
      s2 += tn * nn + to * oo - 2.0 * tn * no;
    }
    drec = q * s1 + 0.5 * q * q * s2;
  } else if (a.ewald == 1) {
    for (int kn = threadIdx.x; kn < a.nk; kn += 256) {
      const double kx = a.kvec[3 * kn] * a.tpb[0], ky = a.kvec[3 * kn + 1] * a.tpb[1],
                   kz = a.kvec[3 * kn + 2] * a.tpb[2];
      double sxn, cxn, syn, cyn, szn, czn, sxo, cxo, syo, cyo, szo, czo;
      sincos(kx * rn[0], &sxn, &cxn);
      sincos(ky * rn[1], &syn, &cyn);
      sincos(kz * rn[2], &szn, &czn);
      sincos(kx * ro[0], &sxo, &cxo);
      sincos(ky * ro[1], &syo, &cyo);
      sincos(kz * ro[2], &szo, &czo);
      double tn = 0.0, to = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double gxn = (c & 2) ? sxn : -sxn, gyn = (c & 1) ? syn : -syn;
        const double gxo = (c & 2) ? sxo : -sxo, gyo = (c & 1) ? syo : -syo;
        const double r1n = cxn * cyn - gxn * gyn, i1n = cxn * gyn + gxn * cyn;
        const double r1o = cxo * cyo - gxo * gyo, i1o = cxo * gyo + gxo * cyo;
        const double dr = q * ((r1n * czn - i1n * szn) - (r1o * czo - i1o * szo));
        const double di = q * ((r1n * szn + i1n * czn) - (r1o * szo + i1o * czo));
        const double Sr = a.sk[8 * kn + 2 * c], Si = a.sk[8 * kn + 2 * c + 1];
        tn += (Sr + dr) * (Sr + dr) + (Si + di) * (Si + di);
        to += Sr * Sr + Si * Si;
      }
      drec += a.kfac[kn] * (tn - to);
    }
  }
  drec = msgd::block_sum(drec, sh);
  double tw = 0.0;  // thread 0: the real-space sum over type pairs
  for (int tp = 0; tp < a.ntp; ++tp) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxT; ++k) v += (a.tm.tpidx[ti * kMaxT + k] == tp) ? e[k] : 0.0;
    v = msgd::block_sum(v, sh);
    if (threadIdx.x == 0) {
      o[1 + tp] = v;
      tw += v;
    }
  }
  if (threadIdx.x == 0) {
    o[0] = tw;
    if (a.ewald && a.lsurf) {
      const double M[3] = {a.Mdip[0], a.Mdip[1], a.Mdip[2]};
      double m2o = 0.0, m2n = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double mn = M[d] + q * (rn[d] - ro[d]);
        m2o += M[d] * M[d];
        m2n += mn * mn;
      }
      drec += a.surf_fac * (m2n - m2o);
    }
    const double du_rec = a.ewald ? a.epsi4pi * drec : 0.0;
    o[a.ntp + 1] = du_rec;
    o[a.ntp + 2] = tw + du_rec;
    o[a.ntp + 3] = (double)s_ovl;
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_dutotal(DeArgs a) {
  const int t = blockIdx.x;
  de_body<P>(a, a.tip[t] - a.tip_base, a.tr + 3 * t, a.out + (size_t)t * (a.ntp + 4));
}

// Trials of many systems in one grid: workgroup b is trial b of the concatenation, system
// s = #{k >= 1 : tstart[k] <= b}; outputs are strided by kMaxTP + 4.
template <int P>
__global__ __launch_bounds__(256) void k_dutotal_seg(const DeArgs* __restrict__ args,
                                                     const int* __restrict__ tstart, int nsys,
                                                     const int* __restrict__ tip,
                                                     const double* __restrict__ tr,
                                                     double* __restrict__ out) {
  const int b = blockIdx.x;
  int cnt = 0;
  for (int k = 1 + threadIdx.x; k < nsys; k += 256) cnt += tstart[k] <= b;
  __shared__ int s_sys;
  if (threadIdx.x == 0) s_sys = 0;
  __syncthreads();
  if (cnt) atomicAdd(&s_sys, cnt);
  __syncthreads();
  de_body<P>(args[s_sys], tip[b], tr + 3 * b, out + (size_t)b * (kMaxTP + 4));
}

// ChargeChange trial with weak charges (Src/mc.F90:3683-3800; DUTotal's weak-charge branch,
// Src/denergy.F90:154-164): the charge states of nptm (1 or 2) particles flip, positions stay.
// Real space: each flipped particle against its cell-list partners with the new states minus
// the old (UWeakChargeANew/AOld, :3911-4145; the pair of the two flipped particles once,
// lptmdutwob), overlap checks before the charge test.  Reciprocal space from the resident
// structure factors changed by (aztm - az) e^{ik.r} (DUWeakChargeEwaldRecStd, :1586-1650),
// the self term (:1654-1667) and the surface term (:1671-1695).  One workgroup, all FP64.
struct DeChargeArgs {
  DeArgs a;
  int nptm;
  int ip[2];
  int laznew[2];
  double aznew[2];
  double alpha_spi;  // alpha / sqrt(pi)
};

__global__ __launch_bounds__(256) void k_dutotal_charge(DeChargeArgs d) {
  const DeArgs& a = d.a;
  __shared__ double sh[8];
  __shared__ int s_ovl;
  if (threadIdx.x == 0) s_ovl = 0;
  __syncthreads();
  double e[kMaxTP];
#pragma unroll
  for (int k = 0; k < kMaxTP; ++k) e[k] = 0.0;
  int ovl = 0;
  for (int m = 0; m < d.nptm; ++m) {
    const int ip = d.ip[m];
    const int ti = a.ipt[ip];
    const int li_old = a.laz[ip], li_new = d.laznew[m];
    const double ri[3] = {a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2]};
    const int cc = a.slot_cell[a.slot_of[ip]];
    const int cx = cc % a.g.nc[0], cy = (cc / a.g.nc[0]) % a.g.nc[1], cz = cc / (a.g.nc[0] * a.g.nc[1]);
    for (int k = 0; k < a.g.nst; ++k) {
      int x = cx + a.g.st[3 * k], y = cy + a.g.st[3 * k + 1], z = cz + a.g.st[3 * k + 2];
      x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
      y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
      z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
      const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
      const int j1 = a.cell_start[nb + 1];
      for (int j = a.cell_start[nb] + threadIdx.x; j < j1; j += 256) {
        const int jp = a.cell_ip[j];
        if (jp == ip) continue;
        int mj = -1;
        for (int q = 0; q < d.nptm; ++q) mj = d.ip[q] == jp ? q : mj;
        if (mj >= 0 && jp < ip) continue;  // the flipped pair once
        const double4 pj = a.cpos[j];
        double dd[3] = {ri[0] - pj.x, ri[1] - pj.y, ri[2] - pj.z};
#pragma unroll
        for (int qd = 0; qd < 3; ++qd)
          if (fabs(dd[qd]) > a.box.half[qd]) dd[qd] -= copysign(a.box.L[qd], dd[qd]);
        const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        if (a.cut_inclusive ? (r2 > a.rcut2) : !(r2 < a.rcut2)) continue;
        const int tj = a.ctype[j];
        const int tp = a.tm.tpidx[ti * kMaxT + tj];
        if (r2 < a.tm.r2atat[tp] || r2 < a.tm.r2umin[tp]) {
          ovl = 1;
          continue;
        }
        const int lj_old = (int)pj.w < kUncharged;
        const int lj_new = mj >= 0 ? d.laznew[mj] : lj_old;
        const int on_new = li_new && lj_new, on_old = li_old && lj_old;
        if (on_new == on_old) continue;
        const double u = table_u(a, tp, r2);
        const double du = on_new ? u : -u;
#pragma unroll
        for (int k2 = 0; k2 < kMaxTP; ++k2) e[k2] += (k2 == tp) ? du : 0.0;
      }
    }
  }
  if (ovl) atomicOr(&s_ovl, 1);
  double dq[2] = {0.0, 0.0}, rm[2][3];
  for (int m = 0; m < d.nptm; ++m) {
    dq[m] = d.aznew[m] - a.az[d.ip[m]];
    for (int k = 0; k < 3; ++k) rm[m][k] = a.ro[3 * d.ip[m] + k];
  }
  double drec = 0.0;
  if (a.ewald == 1) {
    for (int kn = threadIdx.x; kn < a.nk; kn += 256) {
      const double kx = a.kvec[3 * kn] * a.tpb[0], ky = a.kvec[3 * kn + 1] * a.tpb[1],
                   kz = a.kvec[3 * kn + 2] * a.tpb[2];
      double dr[4] = {0, 0, 0, 0}, di[4] = {0, 0, 0, 0};
      for (int m = 0; m < d.nptm; ++m) {
        double sx, cx, sy, cy, sz, cz;
        sincos(kx * rm[m][0], &sx, &cx);
        sincos(ky * rm[m][1], &sy, &cy);
        sincos(kz * rm[m][2], &sz, &cz);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double gx = (c & 2) ? sx : -sx, gy = (c & 1) ? sy : -sy;
          const double r1 = cx * cy - gx * gy, i1 = cx * gy + gx * cy;
          dr[c] += dq[m] * (r1 * cz - i1 * sz);
          di[c] += dq[m] * (r1 * sz + i1 * cz);
        }
      }
      double tn = 0.0, to = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double Sr = a.sk[8 * kn + 2 * c], Si = a.sk[8 * kn + 2 * c + 1];
        tn += (Sr + dr[c]) * (Sr + dr[c]) + (Si + di[c]) * (Si + di[c]);
        to += Sr * Sr + Si * Si;
      }
      drec += a.kfac[kn] * (tn - to);
    }
  }
  drec = msgd::block_sum(drec, sh);
  double tw = 0.0;  // thread 0: the real-space sum over type pairs
#pragma unroll
  for (int tp = 0; tp < kMaxTP; ++tp) {
    if (tp >= a.ntp) break;
    const double v = msgd::block_sum(e[tp], sh);
    if (threadIdx.x == 0) {
      a.out[1 + tp] = v;
      tw += v;
    }
  }
  if (threadIdx.x == 0) {
    double* o = a.out;
    o[0] = tw;
    if (a.ewald) {
      for (int m = 0; m < d.nptm; ++m) {
        const double q0 = a.az[d.ip[m]];
        drec += d.alpha_spi * (q0 * q0 - d.aznew[m] * d.aznew[m]);
      }
      if (a.lsurf) {
        const double M[3] = {a.Mdip[0], a.Mdip[1], a.Mdip[2]};
        double m2o = 0.0, m2n = 0.0;
        for (int k = 0; k < 3; ++k) {
          double mn = M[k];
          for (int m = 0; m < d.nptm; ++m) mn += dq[m] * rm[m][k];
          m2o += M[k] * M[k];
          m2n += mn * mn;
        }
        drec += a.surf_fac * (m2n - m2o);
      }
    }
    const double du_rec = a.ewald ? a.epsi4pi * drec : 0.0;
    o[a.ntp + 1] = du_rec;
    o[a.ntp + 2] = tw + du_rec;
    o[a.ntp + 3] = (double)s_ovl;
  }
}

// g(x) = sum_k meshfac(|k|) cos(2 pi k.x / K) over the full spectrum, folded to [0, K/2]^3
// by symmetry: three cosine stages, each out[.., x, ..] = sum_k cw[x][k] in[.., k, ..].
__global__ void k_cos_stage(const double* __restrict__ in, double* __restrict__ out,
                            const double* __restrict__ cw, int H, int axis) {
  const int n = H * H * H;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  int c[3] = {o % H, (o / H) % H, o / (H * H)};
  const int x = c[axis];
  double s = 0.0;
  for (int k = 0; k < H; ++k) {
    c[axis] = k;
    s += cw[x * H + k] * in[c[0] + H * (c[1] + H * c[2])];
  }
  out[o] = s;
}

template <int P>
void launch_de(Ctx* c, const DeArgs& a, int ntrial) {
  hipLaunchKernelGGL(k_dutotal<P>, dim3(ntrial), dim3(256), 0, c->stream, a);
}

int make_gk(Ctx* c) {
  const int K = c->mesh.K, H = c->mesh.H;
  const size_t n = (size_t)H * H * H;
  if (!c->d_gk)
    if (int e = dalloc(c, (void**)&c->d_gk, n * sizeof(double))) return e;
  // meshfac in FP64 (the device copy may be FP32)
  const msg_params& P = c->P;
  std::vector<double> spl(P.order, 0.0), last;
  spl[1] = 1.0;
  for (int m = 3; m <= P.order; ++m) {
    last = spl;
    spl[0] = 0.0;
    for (int i = 1; i <= m - 2; ++i) spl[i] = (i * last[i] + (m - i) * last[i - 1]) / (m - 1);
    spl[m - 1] = last[m - 2] / (m - 1);
  }
  std::vector<double> b2(K);
  for (int mm = 0; mm < K; ++mm) {
    double re = 0, im = 0;
    for (int k = 0; k <= P.order - 2; ++k) {
      re += spl[k + 1] * std::cos(2.0 * M_PI * mm * k / K);
      im += spl[k + 1] * std::sin(2.0 * M_PI * mm * k / K);
    }
    b2[mm] = ((P.order % 2 == 1) && (2 * mm == K)) ? 0.0 : 1.0 / (re * re + im * im);
  }
  const double vol = P.boxlen[0] * P.boxlen[1] * P.boxlen[2];
  std::vector<double> G(n), cw((size_t)H * H);
  for (int nz = 0; nz < H; ++nz)
    for (int ny = 0; ny < H; ++ny)
      for (int nx = 0; nx < H; ++nx) {
        const size_t o = nx + (size_t)H * (ny + (size_t)H * nz);
        if (!nx && !ny && !nz) { G[o] = 0.0; continue; }
        const double m2 = std::pow(nx * 2.0 * M_PI / P.boxlen[0], 2) +
                          std::pow(ny * 2.0 * M_PI / P.boxlen[1], 2) +
                          std::pow(nz * 2.0 * M_PI / P.boxlen[2], 2);
        G[o] = 4.0 * M_PI / vol * std::exp(-m2 / (4.0 * P.alpha * P.alpha)) / m2 * b2[nx] * b2[ny] * b2[nz];
      }
  for (int x = 0; x < H; ++x)
    for (int k = 0; k < H; ++k)
      cw[(size_t)x * H + k] = ((k == 0 || k == K / 2) ? 1.0 : 2.0) * std::cos(2.0 * M_PI * k * x / K);
  double *d_a = nullptr, *d_b = nullptr, *d_cw = nullptr;
  MSG_HIP(hipMalloc(&d_a, n * sizeof(double)));
  MSG_HIP(hipMalloc(&d_b, n * sizeof(double)));
  MSG_HIP(hipMalloc(&d_cw, cw.size() * sizeof(double)));
  MSG_HIP(hipMemcpy(d_a, G.data(), n * sizeof(double), hipMemcpyHostToDevice));
  MSG_HIP(hipMemcpy(d_cw, cw.data(), cw.size() * sizeof(double), hipMemcpyHostToDevice));
  const int nb = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_cos_stage, dim3(nb), dim3(256), 0, c->stream, d_a, d_b, d_cw, H, 0);
  hipLaunchKernelGGL(k_cos_stage, dim3(nb), dim3(256), 0, c->stream, d_b, d_a, d_cw, H, 1);
  hipLaunchKernelGGL(k_cos_stage, dim3(nb), dim3(256), 0, c->stream, d_a, c->d_gk, d_cw, H, 2);
  MSG_HIP(hipStreamSynchronize(c->stream));
  hipFree(d_a);
  hipFree(d_b);
  hipFree(d_cw);
  c->have_gk = true;
  return MSG_OK;
}

}  // namespace

int spme_trial_setup(Ctx* c) {
  if (c->P.ewald_mode != MSG_EWALD_SPM || c->have_gk) return MSG_OK;
  return make_gk(c);
}

static DeArgs de_args(Ctx* c);

int dutotal_launch(Ctx* c, int ntrial, const int* ip, const double* rotm) {
  for (int t = 0; t < ntrial; ++t)
    if (ip[t] < 1 || ip[t] > c->np) return set_error(MSG_ERR_ARG, "DUTotal: particle index out of range");
  if (c->P.ewald_mode == MSG_EWALD_SPM && !c->have_gk)
    if (int e = make_gk(c)) return e;
  if (!c->fresh)
    if (int e = utotal_refresh(c)) return e;
  if (int e = complete_cells(c)) return e;  // a Verlet-list pass may have left the cell order out
  if (ntrial > c->de_cap) {
    MSG_HIP(hipStreamSynchronize(c->stream));
    if (c->d_de) { hipFree(c->d_de); hipFree(c->d_de_ip); hipFree(c->d_de_r); }
    c->d_de = nullptr;
    const int cap = std::max(ntrial, 256);
    if (int e = dalloc(c, (void**)&c->d_de, (size_t)cap * (kMaxTP + 4) * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_de_ip, cap * sizeof(int))) return e;
    if (int e = dalloc(c, (void**)&c->d_de_r, (size_t)cap * 3 * sizeof(double))) return e;
    c->de_cap = cap;
  }
  // 0-based indices staged in the context: the copy may still be pending when this returns
  std::vector<int>& ip0 = c->h_de_ip;
  ip0.resize(ntrial);
  for (int t = 0; t < ntrial; ++t) ip0[t] = ip[t] - 1;
  MSG_HIP(hipMemcpyAsync(c->d_de_ip, ip0.data(), ntrial * sizeof(int), hipMemcpyHostToDevice, c->stream));
  MSG_HIP(hipMemcpyAsync(c->d_de_r, rotm, ntrial * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  DeArgs a = de_args(c);
  a.tip = c->d_de_ip;
  a.tr = c->d_de_r;
  a.out = c->d_de;
  prof_begin(c, MSG_K_DE);
  switch (c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.order : 4) {
    case 2: launch_de<2>(c, a, ntrial); break;
    case 3: launch_de<3>(c, a, ntrial); break;
    case 4: launch_de<4>(c, a, ntrial); break;
    case 5: launch_de<5>(c, a, ntrial); break;
    case 6: launch_de<6>(c, a, ntrial); break;
    case 7: launch_de<7>(c, a, ntrial); break;
    case 8: launch_de<8>(c, a, ntrial); break;
    case 9: launch_de<9>(c, a, ntrial); break;
    case 10: launch_de<10>(c, a, ntrial); break;
    case 11: launch_de<11>(c, a, ntrial); break;
    default: return set_error(MSG_ERR_ARG, "DUTotal: SPME order must be 2..11");
  }
  prof_end(c, MSG_K_DE);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

static DeArgs de_args(Ctx* c) {
  const int ntp = c->nptpt;
  DeArgs a;
  a.np = c->np;
  a.ntp = ntp;
  a.ro = c->d_ro;
  a.az = c->d_az;
  a.ipt = c->d_ipt;
  a.g = c->grid;
  for (int d = 0; d < 3; ++d) {
    a.box.L[d] = c->P.boxlen[d];
    a.box.half[d] = 0.5 * c->P.boxlen[d];
    a.box.inv[d] = 1.0 / c->P.boxlen[d];
  }
  a.cell_start = c->d_cell_start;
  a.cell_ip = c->d_cell_ip;
  a.cpos = (const double4*)c->d_cpos;
  a.ctype = c->d_ctype;
  a.slot_of = c->d_slot_of;
  a.slot_cell = c->d_slot_cell;
  a.rcut2 = c->P.rcut * c->P.rcut;
  a.cut_inclusive = c->P.list_mode != MSG_LIST_CELL;
  a.tcoef = c->d_tcoef64;
  a.tbins = c->d_tbins;
  a.tm = c->tmeta;
  a.ewald = c->P.ewald_mode;
  a.lsurf = c->P.lsurf;
  const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
  a.surf_fac = 2.0 * M_PI / (3.0 * vol);
  a.Mdip = c->d_out + ntp + 4;
  a.m = c->mesh;
  a.order = c->mesh.order;
  a.phi = c->d_Q;
  a.phi_fp64 = c->fp64;
  a.gk = c->d_gk;
  a.nk = c->nkvec;
  a.kvec = c->d_kvec;
  a.kfac = c->d_kfac;
  a.sk = c->d_sk;
  for (int d = 0; d < 3; ++d) a.tpb[d] = 2.0 * M_PI / c->P.boxlen[d];
  a.epsi4pi = c->P.epsi4pi;
  a.tip = nullptr;
  a.tip_base = 0;
  a.tr = nullptr;
  a.out = nullptr;
  a.laz = c->d_laz;
  return a;
}

// Batched trials: systems grouped by B-spline order (the kernel's template), one argument
// block per system uploaded once.
int plan_de(const std::vector<Ctx*>& cs, SegOwner& own, std::vector<DeGroup>& out) {
  std::map<int, std::vector<int>> by_order;
  for (int s = 0; s < (int)cs.size(); ++s) {
    Ctx* c = cs[s];
    if (c->P.ewald_mode == MSG_EWALD_SPM && !c->have_gk)
      if (int e = make_gk(c)) return e;
    by_order[c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.order : 4].push_back(s);
  }
  for (auto& kv : by_order) {
    if (kv.first < kMinOrder || kv.first > kMaxOrder) return set_error(MSG_ERR_ARG, "DUTotal: SPME order must be 2..11");
    std::vector<DeArgs> args;
    for (int s : kv.second) args.push_back(de_args(cs[s]));
    DeGroup g;
    g.order = kv.first;
    g.sys = kv.second;
    const DeArgs* d = nullptr;
    if (int e = own.upload(args, &d)) return e;
    g.d_args = d;
    out.push_back(g);
  }
  return MSG_OK;
}

int launch_de_seg(const DeGroup& g, const int* d_tstart, const int* d_ip, const double* d_r,
                  double* d_out, int ntot, hipStream_t s) {
  if (ntot <= 0) return MSG_OK;
  const DeArgs* a = (const DeArgs*)g.d_args;
  const int n = (int)g.sys.size();
  switch (g.order) {
#define MSG_DE_SEG(P) \
  case P: hipLaunchKernelGGL(k_dutotal_seg<P>, dim3(ntot), dim3(256), 0, s, a, d_tstart, n, d_ip, d_r, d_out); break;
    MSG_DE_SEG(2) MSG_DE_SEG(3) MSG_DE_SEG(4) MSG_DE_SEG(5) MSG_DE_SEG(6) MSG_DE_SEG(7) MSG_DE_SEG(8)
    MSG_DE_SEG(9) MSG_DE_SEG(10) MSG_DE_SEG(11)
#undef MSG_DE_SEG
    default: return set_error(MSG_ERR_ARG, "DUTotal: SPME order must be 2..11");
  }
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int dutotal_collect(Ctx* c, int ntrial, int* ovl, double* du_tot, double* du_rec, double* du_twob) {
  const int ntp = c->nptpt;
  std::vector<double> res((size_t)ntrial * (ntp + 4));
  MSG_HIP(hipMemcpyAsync(res.data(), c->d_de, res.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  for (int t = 0; t < ntrial; ++t) {
    const double* o = res.data() + (size_t)t * (ntp + 4);
    ovl[t] = o[ntp + 3] != 0.0;
    du_tot[t] = o[ntp + 2];
    if (du_rec) du_rec[t] = o[ntp + 1];
    if (du_twob)
      for (int k = 0; k <= ntp; ++k) du_twob[(size_t)t * (ntp + 1) + k] = o[k];
  }
  return MSG_OK;
}

int dutotal_batch(Ctx* c, int ntrial, const int* ip, const double* rotm, int* ovl, double* du_tot,
                  double* du_rec, double* du_twob) {
  if (int e = dutotal_launch(c, ntrial, ip, rotm)) return e;
  return dutotal_collect(c, ntrial, ovl, du_tot, du_rec, du_twob);
}

// ChargeChange trial: the charge states of nptm particles ip1[] (1-based) flip.
int dutotal_charge(Ctx* c, int nptm, const int* ip1, int* ovl, double* du_twob, double* du_rec,
                   double* du_tot) {
  if (c->h_laz.empty()) return set_error(MSG_ERR_STATE, "ChargeChange: weak charges are not set");
  if (nptm < 1 || nptm > 2) return set_error(MSG_ERR_ARG, "ChargeChange: one or two particles");
  for (int m = 0; m < nptm; ++m)
    if (ip1[m] < 1 || ip1[m] > c->np) return set_error(MSG_ERR_ARG, "ChargeChange: particle index out of range");
  if (nptm == 2 && ip1[0] == ip1[1]) return set_error(MSG_ERR_ARG, "ChargeChange: the same particle twice");
  if (!c->fresh)
    if (int e = utotal_refresh(c)) return e;
  if (int e = complete_cells(c)) return e;
  if (c->de_cap < 1) {
    MSG_HIP(hipStreamSynchronize(c->stream));
    if (int e = dalloc(c, (void**)&c->d_de, (size_t)256 * (kMaxTP + 4) * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_de_ip, 256 * sizeof(int))) return e;
    if (int e = dalloc(c, (void**)&c->d_de_r, (size_t)256 * 3 * sizeof(double))) return e;
    c->de_cap = 256;
  }
  DeChargeArgs d;
  d.a = de_args(c);
  d.a.out = c->d_de;
  d.nptm = nptm;
  for (int m = 0; m < 2; ++m) {
    const int ip = ip1[m < nptm ? m : 0] - 1;
    d.ip[m] = ip;
    d.laznew[m] = !c->h_laz[ip];
    d.aznew[m] = d.laznew[m] ? c->h_zat[c->h_ipt[ip]] : 0.0;  // aztm, Src/mc.F90:3740-3744
  }
  d.alpha_spi = c->P.alpha / std::sqrt(M_PI);
  prof_begin(c, MSG_K_DE);
  hipLaunchKernelGGL(k_dutotal_charge, dim3(1), dim3(256), 0, c->stream, d);
  prof_end(c, MSG_K_DE);
  MSG_HIP(hipGetLastError());
  return dutotal_collect(c, 1, ovl, du_tot, du_rec, du_twob);
}

}  // namespace msg

namespace msg {

namespace {

// MCUpdate for one accepted single-particle move (Src/mc.F90:5203-5235): EwaldUpdateArray for
// standard Ewald (S(k) += q (e^{ik.r_new} - e^{ik.r_old}), Src/denergy.F90:1776-1819), the dipole
// sum, ro <- rotm and UpdateCellIp (the cell list patched in place, head insertion), all in one
// workgroup; no UTotal pass follows for standard Ewald or no Ewald.  With SPME the convolved
// mesh is refreshed by the next pass (which keeps the patched cell order).

struct CommitArgs {
  int ip;
  double rn[3];
  double* ro;
  const double* az;
  int nk;                 // standard Ewald: k-vectors to update (0: none)
  const int* kvec;
  double tpb[3];
  double* sk;
  double* Mout;
  double* Mtot;
  int cells;              // the cell list is valid: patch it
  CellMoveArgs cm;
};

__global__ __launch_bounds__(1024) void k_commit(CommitArgs a) {
  const int ip = a.ip;
  const double q = a.az[ip];
  const double ro_old[3] = {a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2]};
  const double* rn = a.rn;
  for (int kn = threadIdx.x; kn < a.nk; kn += 1024) {
    const double kx = a.kvec[3 * kn] * a.tpb[0], ky = a.kvec[3 * kn + 1] * a.tpb[1], kz = a.kvec[3 * kn + 2] * a.tpb[2];
    double sxn, cxn, syn, cyn, szn, czn, sxo, cxo, syo, cyo, szo, czo;
    sincos(kx * rn[0], &sxn, &cxn);
    sincos(ky * rn[1], &syn, &cyn);
    sincos(kz * rn[2], &szn, &czn);
    sincos(kx * ro_old[0], &sxo, &cxo);
    sincos(ky * ro_old[1], &syo, &cyo);
    sincos(kz * ro_old[2], &szo, &czo);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double gxn = (c & 2) ? sxn : -sxn, gyn = (c & 1) ? syn : -syn;
      const double gxo = (c & 2) ? sxo : -sxo, gyo = (c & 1) ? syo : -syo;
      const double r1n = cxn * cyn - gxn * gyn, i1n = cxn * gyn + gxn * cyn;
      const double r1o = cxo * cyo - gxo * gyo, i1o = cxo * gyo + gxo * cyo;
      a.sk[8 * kn + 2 * c] += q * ((r1n * czn - i1n * szn) - (r1o * czo - i1o * szo));
      a.sk[8 * kn + 2 * c + 1] += q * ((r1n * szn + i1n * czn) - (r1o * szo + i1o * czo));
    }
  }
  if (a.cells) cell_move<1024>(a.cm, ip, rn);
  __syncthreads();  // every thread has read the old position
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) {
      a.Mout[d] += q * (rn[d] - ro_old[d]);
      a.Mtot[d] = a.Mout[d];
      a.ro[3 * ip + d] = rn[d];
    }
  }
}

// The cell-ordered record of a particle whose charge state flipped: its type tag.
__global__ void k_set_tag(int ip, const int* slot_of, double4* cpos, float4* crel, double tag) {
  if (threadIdx.x == 0) {
    cpos[slot_of[ip]].w = tag;
    if (crel) crel[slot_of[ip]].w = (float)tag;
  }
}

}  // namespace

CellMoveArgs cell_move_args(Ctx* c) {
  CellMoveArgs m;
  m.g = c->grid;
  for (int d = 0; d < 3; ++d) {
    m.box.L[d] = c->P.boxlen[d];
    m.box.half[d] = 0.5 * c->P.boxlen[d];
    m.box.inv[d] = 1.0 / c->P.boxlen[d];
  }
  m.cell_start = c->d_cell_start;
  m.cell_ip = c->d_cell_ip;
  m.cpos = (double4*)c->d_cpos;
  m.ctype = c->d_ctype;
  m.slot_of = c->d_slot_of;
  m.slot_cell = c->d_slot_cell;
  m.crel = (float4*)c->d_crel;
  return m;
}

int commit_move(Ctx* c, int ip, const double* rnew) {
  CommitArgs a;
  a.ip = ip;
  for (int d = 0; d < 3; ++d) {
    a.rn[d] = rnew[d];
    a.tpb[d] = 2.0 * M_PI / c->P.boxlen[d];
  }
  a.ro = c->d_ro;
  a.az = c->d_az;
  // the structure factors are updated when they hold the current state (else a pass rebuilds them)
  a.nk = (c->P.ewald_mode == MSG_EWALD_STD && c->fresh) ? c->nkvec : 0;
  a.kvec = c->d_kvec;
  a.sk = c->d_sk;
  a.Mout = c->d_out + c->nptpt + 4;
  a.Mtot = c->d_M;
  a.cells = c->cells_valid ? 1 : 0;
  a.cm = cell_move_args(c);
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(1024), 0, c->stream, a);
  MSG_HIP(hipGetLastError());
  if (c->P.ewald_mode == MSG_EWALD_SPM) c->fresh = false;  // the convolved mesh is the old one
  if (c->cells_valid) c->cells_maintained = true;
  return MSG_OK;
}

int set_charge_tag(Ctx* c, int ip) {
  if (!c->cells_valid) return MSG_OK;  // the next pass ranks the cells afresh
  const double tag = (double)(c->h_ipt[ip] + (c->h_laz[ip] ? 0 : kUncharged));
  hipLaunchKernelGGL(k_set_tag, dim3(1), dim3(64), 0, c->stream, ip, c->d_slot_of, (double4*)c->d_cpos,
                     (float4*)c->d_crel, tag);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg

// ------------------------------------------------------------------------------------------
// DUTotal for callers that cannot allocate (graph capture, tight serving loops): the trial
// inputs and the per-trial records live in `scratch`, a file-scope device array of a fixed
// size that each GPU the library runs on holds until the process ends.  Its device address is
// looked up once per device at context creation (hipGetSymbolAddress, with that device
// current); host code uses only that address, never the array's name.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU (from several threads or contexts)
// race.  A call writes every slot it reads back, so nothing needs clearing between calls.
namespace msg {

namespace {

constexpr int kNoAllocTrials = 16384;             // trials per call at most
constexpr size_t kRecD = kMaxTP + 4;              // doubles per trial record
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kOffRec = 0;
constexpr size_t kOffIp = kOffRec + align256(kNoAllocTrials * kRecD * sizeof(double));
constexpr size_t kOffR = kOffIp + align256(kNoAllocTrials * sizeof(int));
constexpr size_t kOffOvl = kOffR + align256(3 * (size_t)kNoAllocTrials * sizeof(double));
constexpr size_t kOffTot = kOffOvl + align256(kNoAllocTrials * sizeof(int));
constexpr size_t kOffDrec = kOffTot + align256(kNoAllocTrials * sizeof(double));
constexpr size_t kScratchBytes = kOffDrec + align256(kNoAllocTrials * sizeof(double));  // 6.0 MB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The records of the trials unpacked into the caller's output layout (overlap flags,
// du_tot, du_rec), so each comes back to the host with one copy.
__global__ __launch_bounds__(256) void k_de_unpack(const double* __restrict__ rec, int stride, int ntp,
                                                   int ntrial, int* ovl, double* tot, double* drec) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntrial) return;
  const double* o = rec + (size_t)t * stride;
  ovl[t] = o[ntp + 3] != 0.0;
  tot[t] = o[ntp + 2];
  drec[t] = o[ntp + 1];
}

}  // namespace

// Once per device, with that device current (context creation): the address of this
// device's copy of the scratch array.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return set_error(MSG_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return MSG_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(MSG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = (unsigned char*)ptr;
  return MSG_OK;
}

}  // namespace msg

using namespace msg;

extern "C" {

int msg_dutotal_batch_noalloc(msg_ctx* h, int ntrial, const int* ip, const double* rotm, int* lhsoverlap,
                              double* du_tot, double* du_rec) {
  if (!h || !ip || !rotm || !lhsoverlap || !du_tot)
    return set_error(MSG_ERR_ARG, "msg_dutotal_batch_noalloc: null argument");
  if (ntrial < 1 || ntrial > kNoAllocTrials)
    return set_error(MSG_ERR_ARG, "msg_dutotal_batch_noalloc: ntrial must be 1..16384 (the scratch holds 16384 trials)");
  Ctx* c = &h->c;
  if (int e = check_ready_state(c)) return e;
  for (int t = 0; t < ntrial; ++t)
    if (ip[t] < 1 || ip[t] > c->np) return set_error(MSG_ERR_ARG, "msg_dutotal_batch_noalloc: particle index out of range");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = g_scratch[c->device];
  }
  if (!base) return set_error(MSG_ERR_STATE, "msg_dutotal_batch_noalloc: no scratch address for this device");
  if (c->P.ewald_mode == MSG_EWALD_SPM && !c->have_gk)
    return set_error(MSG_ERR_STATE,
                     "msg_dutotal_batch_noalloc: the SPME trial tables are built by the first msg_dutotal / "
                     "msg_dutotal_batch call (it allocates); call one of them once first");
  if (!c->fresh)
    if (int e = utotal_refresh(c)) return e;
  if (int e = complete_cells(c)) return e;
  double* rec = (double*)(base + kOffRec);
  int* d_ip = (int*)(base + kOffIp);
  double* d_r = (double*)(base + kOffR);
  int* d_ovl = (int*)(base + kOffOvl);
  double* d_tot = (double*)(base + kOffTot);
  double* d_drec = (double*)(base + kOffDrec);
  MSG_HIP(hipMemcpyAsync(d_ip, ip, ntrial * sizeof(int), hipMemcpyHostToDevice, c->stream));
  MSG_HIP(hipMemcpyAsync(d_r, rotm, 3 * (size_t)ntrial * sizeof(double), hipMemcpyHostToDevice, c->stream));
  DeArgs a = de_args(c);
  a.tip = d_ip;
  a.tip_base = 1;
  a.tr = d_r;
  a.out = rec;
  // (no per-kernel timing here: its event bookkeeping grows a host vector)
  switch (c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.order : 4) {
    case 2: launch_de<2>(c, a, ntrial); break;
    case 3: launch_de<3>(c, a, ntrial); break;
    case 4: launch_de<4>(c, a, ntrial); break;
    case 5: launch_de<5>(c, a, ntrial); break;
    case 6: launch_de<6>(c, a, ntrial); break;
    case 7: launch_de<7>(c, a, ntrial); break;
    case 8: launch_de<8>(c, a, ntrial); break;
    case 9: launch_de<9>(c, a, ntrial); break;
    case 10: launch_de<10>(c, a, ntrial); break;
    case 11: launch_de<11>(c, a, ntrial); break;
    default: return set_error(MSG_ERR_ARG, "DUTotal: SPME order must be 2..11");
  }
  hipLaunchKernelGGL(k_de_unpack, dim3((ntrial + 255) / 256), dim3(256), 0, c->stream, rec, c->nptpt + 4,
                     c->nptpt, ntrial, d_ovl, d_tot, d_drec);
  MSG_HIP(hipGetLastError());
  MSG_HIP(hipMemcpyAsync(lhsoverlap, d_ovl, ntrial * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipMemcpyAsync(du_tot, d_tot, ntrial * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (du_rec) MSG_HIP(hipMemcpyAsync(du_rec, d_drec, ntrial * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  return MSG_OK;
}

}  // extern "C"
