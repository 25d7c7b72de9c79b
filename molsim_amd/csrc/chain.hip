// Device-resident Monte Carlo chain: MCPass / SPartMove / GetRandomTrialPos / Metropolis /
// MCUpdate (Src/mc.F90:1470-1750,4527-4573,5135-5235) for single-particle moves, the
// trials run one after another inside ONE workgroup of 1024 threads, with MOLSIM's random
// number generator (Random, Src/mollib.F90:2689-2713) on the device, so a pass of any
// length costs one launch and no host round trip per trial.
//
// Per trial (all FP64):
//   thread 0   particle (sequential or random), trial position (Random - 1/2) * dtran per
//              axis, periodic wrap;
//   all        real-space energy difference over the cell list (DUTwoBody semantics: a
//              hard-core or below-table pair at the new position is an overlap), standard-
//              Ewald structure-factor change dS(k) = q (e^{ik.r_new} - e^{ik.r_old}) with
//              dE = sum kfac (|S + dS|^2 - |S|^2), surface term;
//   thread 0   Metropolis (expmax = 87 guards, weight 1);
//   accepted   S += dS (kept in registers from the energy step), the dipole sum, the position,
//              and UpdateCellIp (Src/celllist.F90:366-382) on the cell list in place: a move
//              into another cell makes the particle its new cell's head and shifts the slots in
//              between, so the chain runs on without leaving the launch.
// SPME (DUTwoBodyEwaldRecSPM, Src/denergy.F90:1413-1495, evaluated locally): with phi = g * Q
// the convolved mesh of the last UTotal pass and g the real-space influence kernel,
//   dE = sum_m dQ(m) phi'(m) + 1/2 sum_{m,m'} dQ(m) g(m - m') dQ(m'),
//   phi' = phi + sum_k g * dQ_k over the moves k accepted since phi was computed.
// Each double sum over two stencils u, v (products of per-axis B-spline weights) collapses to a
// sum over the (2p - 1)^3 mesh offsets d of g(mm_u - mm_v + d) times the product of the three
// per-axis cross-correlations sum_a u(a) v(a + d), so the self term costs 3 (2p - 1)^3 and
// each pending move 4 (2p - 1)^3 multiply-adds.  After kChainLog accepted moves (batched
// chains: up to kChainLogMax at orders up to 6) the launch ends; the host runs a UTotal pass
// (phi of the current positions, the patched cell order kept) and launches the chain again
// where it stopped.
#include <cmath>
#include <functional>

#include "cellmove.h"
#include "common.h"

namespace msg {

namespace {

constexpr int kChainThreads = 1024;
constexpr int kChainKvec = 4;   // k-vectors per thread at most: nk <= 4096
constexpr int kChainNcut = 31;  // largest k-vector index per axis
constexpr int kChainLog = 16;   // SPME: accepted moves corrected for before phi is recomputed
constexpr int kChainLogMax = 32; // the log's capacity at orders up to 6 (batched chains use it)

__host__ __device__ constexpr int chain_log_cap(int order) { return order <= 6 ? kChainLogMax : kChainLog; }

struct ChainArgs {
  int np, ntrial, isamp, ewald, lsurf;
  double* ro;
  const double* az;
  const int* ipt;
  double4* cpos;
  int* cell_ip;
  int* cell_start;
  int* slot_of;
  int* slot_cell;
  int* ctype;
  float4* crel;
  CellGrid g;
  Box box;
  double rcut2;
  int cut_inclusive;
  const double* tcoef;  // FP64 table records
  const int* tbins;
  GpuTableMeta tm;
  int nk, ncut;
  const int* kvec;
  const double* kfac;
  double* sk;
  double tpb[3];
  // SPME
  MeshGrid m;
  const void* phi;      // K^3 reals (float or double): g * Q of the last UTotal pass
  int phi_fp64;
  const double* gk;     // H^3 folded real-space influence kernel
  double* Mout;         // dipole sum kept with the energies
  double* Mtot;         // dipole sum read by the force kernels
  double surf_fac;
  double epsi4pi;
  const double* dtran;  // per particle type
  double beta;
  int* rng;             // iseed, xorshift word, Park-Miller word
  double* am;
  int* next;            // sequential sampling cursor
  int* events;          // per trial: -1 overlap, 0 rejected, 1 accepted
  double* stat;         // [0] sum of accepted dE, [1] accepted, [2] trials done, [3] moves into another cell
  int maxlog;           // SPME: accepted moves per launch (at most the log's capacity)
};

// MOLSIM's Random: Marsaglia xorshift combined with Park-Miller by Schrage's method.
__device__ double chain_random(int* st, double* am) {
  const int a = 16807, m = 2147483647, q = 127773, r = 2836;
  if (st[0] <= 0 || st[2] < 0) {
    const int s = st[0] < 0 ? -st[0] : st[0];
    // the single-precision real below 1 divided by m in single precision, where m rounds
    // to 2^31: (1 - 2^-24) 2^-31, exact
    *am = (double)__int_as_float(0x3f7fffff) * 0x1p-31;
    st[2] = (888889999 ^ s) | 1;
    st[1] = 777755555 ^ s;
    st[0] = s + 1;
  }
  unsigned int x = (unsigned int)st[1];
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  st[1] = (int)x;
  const int k = st[2] / q;
  st[2] = a * (st[2] - k * q) - r * k;
  if (st[2] < 0) st[2] += m;
  return *am * (double)((m & (st[1] ^ st[2])) | 1);
}

__device__ inline double chain_table_u(const ChainArgs& a, int tp, double r2) {
  int b = (int)((r2 - a.tm.zlo[tp]) * a.tm.invw[tp]);
  b = min(max(b, 0), kTabBins - 1);
  int idx = a.tbins[tp * kTabBins + b];
  const int last = a.tm.last[tp];
  while (idx < last && r2 < a.tcoef[12 * idx]) ++idx;
  const double* t = a.tcoef + 12 * idx;
  const double d = r2 - t[0];
  return t[1] + d * (t[2] + d * (t[3] + d * (t[4] + d * (t[5] + d * t[6]))));
}

__device__ inline int chain_cell(const ChainArgs& a, const double* r) {
  int ci[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int k = (int)floor((r[d] + a.box.half[d]) * a.g.csi[d]);
    ci[d] = min(max(k, 0), a.g.nc[d] - 1);
  }
  return ci[0] + a.g.nc[0] * (ci[1] + a.g.nc[1] * ci[2]);
}

// |d| folded into one period of K (g is even along every axis)
__device__ inline int chain_fold(int d, int K) {
  d %= K;
  if (d < 0) d += K;
  return d <= K / 2 ? d : K - d;
}

// sum_a u(a) v(a + d) of two stencils' weights along one axis, d = dd - (P - 1)
template <int P>
__device__ inline double chain_xcorr(const double* u, const double* v, int dd) {
  const int d = dd - (P - 1);
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
    const int b = a + d;
    if (b >= 0 && b < P) s += u[a] * v[b];
  }
  return s;
}

// Sum over the workgroup; the result is valid in every thread.
__device__ inline double chain_sum(double v, double* sh) {
  v = msgd::wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kChainThreads / 64; ++w) s += sh[w];
  return s;
}

// KV: k-vectors per thread (the smallest of 1, 2, 4 that covers nk, so the structure-factor
// changes kept for the commit fit in registers).  P: SPME order, 0 for standard or no Ewald.
template <int KV, int P>
__device__ __forceinline__ void chain_body(const ChainArgs& a) {
  constexpr int PP = P > 0 ? P : 1;
  constexpr int W = 2 * PP - 1;
  constexpr int NL = P > 0 ? chain_log_cap(P) : 1;
  constexpr int NT = 3 + 4 * NL;  // terms: the trial's self terms, 4 per pending move
  __shared__ double sh[kChainThreads / 64];
  // e^{i n (2 pi / L_d) r_d} for n = 0..ncut, per axis d, at the trial (0) and old (1) position
  __shared__ double2 s_eik[P > 0 ? 1 : 2][3][P > 0 ? 1 : kChainNcut + 1];
  // SPME: the trial's per-axis weights and base points ([0] new, [1] old position), those of
  // the moves accepted in this launch, and per term its cross-correlations, offset and factor
  __shared__ double s_th[2][3][PP];
  __shared__ int s_mm[2][3];
  __shared__ double s_lth[NL][2][3][PP];
  __shared__ int s_lmm[NL][2][3];
  __shared__ double s_lq[NL];
  __shared__ double s_cv[NT][3][W];
  __shared__ int s_cd[NT][3];
  __shared__ double s_cf[NT];
  int npend = 0;
  __shared__ int s_ip, s_acc, s_ovl;
  __shared__ double s_rn[3];
  __shared__ int s_rng[3];
  __shared__ double s_am;
  __shared__ int s_next;
  const int tid = threadIdx.x;
  if (tid == 0) {
    for (int k = 0; k < 3; ++k) s_rng[k] = a.rng[k];
    s_am = *a.am;
    s_next = *a.next;
  }
  CellMoveArgs cm;
  cm.g = a.g;
  cm.box = a.box;
  cm.cell_start = a.cell_start;
  cm.cell_ip = a.cell_ip;
  cm.cpos = a.cpos;
  cm.ctype = a.ctype;
  cm.slot_of = a.slot_of;
  cm.slot_cell = a.slot_cell;
  cm.crel = a.crel;
  double dsum = 0.0;
  int nacc = 0, done = 0, ncross = 0;
  const double expmax = 87.0;  // Src/mol.F90:62
  for (int t = 0; t < a.ntrial; ++t) {
    if (tid == 0) {
      int ip;
      if (a.isamp == 0) {
        ip = s_next;
        s_next = (s_next + 1) % a.np;
      } else {
        ip = (int)(a.np * chain_random(s_rng, &s_am));
        ip = min(max(ip, 0), a.np - 1);
      }
      const double dt = a.dtran[a.ipt[ip]];
      for (int d = 0; d < 3; ++d) s_rn[d] = a.ro[3 * ip + d] + (chain_random(s_rng, &s_am) - 0.5) * dt;
      for (int d = 0; d < 3; ++d) {
        const double L = a.box.L[d];
        if (fabs(s_rn[d]) > 0.5 * L) s_rn[d] -= L * round(s_rn[d] / L);
      }
      s_ip = ip;
      s_ovl = 0;
    }
    __syncthreads();
    const int ip = s_ip;
    const double rn[3] = {s_rn[0], s_rn[1], s_rn[2]};
    const double ro[3] = {a.ro[3 * ip], a.ro[3 * ip + 1], a.ro[3 * ip + 2]};
    const double q = a.az[ip];
    const int ti = a.ipt[ip];
    if constexpr (P > 0) {
      if (tid < 6) {  // B-spline weights and base point per (position, axis), in FP64
        const int w = tid / 3, d = tid % 3;
        const double x = a.m.dkdr[d] * (w == 0 ? rn[d] : ro[d]);
        const double v = floor(x);
        double th[PP];
        msgd::bspline<double, PP>(x - v, th, (double*)nullptr);
#pragma unroll
        for (int k = 0; k < PP; ++k) s_th[w][d][k] = th[k];
        const int k = (int)v % a.m.K;
        s_mm[w][d] = k < 0 ? k + a.m.K : k;
      }
    } else if (a.ewald == 1 && tid < 6 * (a.ncut + 1)) {
      // one sincos per (position, axis, n); the k-vector loop below reads them
      const int n = tid % (a.ncut + 1), d = (tid / (a.ncut + 1)) % 3, w = tid / (3 * (a.ncut + 1));
      const double k = n * a.tpb[d];
      double sv, cv;
      sincos(k * (w == 0 ? rn[d] : ro[d]), &sv, &cv);
      s_eik[w][d][n] = make_double2(cv, sv);
    }
    // real space: new minus old
    double e = 0.0;
    int ovl = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const double r[3] = {pass == 0 ? rn[0] : ro[0], pass == 0 ? rn[1] : ro[1], pass == 0 ? rn[2] : ro[2]};
      const int cc = pass == 0 ? chain_cell(a, rn) : a.slot_cell[a.slot_of[ip]];
      const int cx = cc % a.g.nc[0], cy = (cc / a.g.nc[0]) % a.g.nc[1], cz = cc / (a.g.nc[0] * a.g.nc[1]);
      const double sgn = pass == 0 ? 1.0 : -1.0;
      for (int k = 0; k < a.g.nst; ++k) {
        int x = cx + a.g.st[3 * k], y = cy + a.g.st[3 * k + 1], z = cz + a.g.st[3 * k + 2];
        x += x < 0 ? a.g.nc[0] : (x >= a.g.nc[0] ? -a.g.nc[0] : 0);
        y += y < 0 ? a.g.nc[1] : (y >= a.g.nc[1] ? -a.g.nc[1] : 0);
        z += z < 0 ? a.g.nc[2] : (z >= a.g.nc[2] ? -a.g.nc[2] : 0);
        const int nb = x + a.g.nc[0] * (y + a.g.nc[1] * z);
        const int j1 = a.cell_start[nb + 1];
        for (int j = a.cell_start[nb] + tid; j < j1; j += kChainThreads) {
          if (a.cell_ip[j] == ip) continue;
          const double4 pj = a.cpos[j];
          double d[3] = {r[0] - pj.x, r[1] - pj.y, r[2] - pj.z};
#pragma unroll
          for (int qd = 0; qd < 3; ++qd)
            if (fabs(d[qd]) > a.box.half[qd]) d[qd] -= copysign(a.box.L[qd], d[qd]);
          const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (a.cut_inclusive ? (r2 > a.rcut2) : !(r2 < a.rcut2)) continue;
          const int tp = a.tm.tpidx[ti * kMaxT + (int)pj.w];
          if (pass == 0 && (r2 < a.tm.r2atat[tp] || r2 < a.tm.r2umin[tp])) {
            ovl = 1;
            continue;
          }
          e += sgn * chain_table_u(a, tp, r2);
        }
      }
    }
    if (ovl) atomicOr(&s_ovl, 1);
    // standard Ewald: dS per k-vector kept for the commit
    __syncthreads();  // s_eik, s_th
    double dS[KV][8];
    double drec = 0.0;
    if constexpr (P > 0) {
      const int nt = 3 + 4 * npend;
      for (int w = tid; w < nt * 3 * W; w += kChainThreads) {
        const int dd = w % W, d = (w / W) % 3, tm = w / (3 * W);
        const double* u;
        const double* v;
        int off;
        double cf;
        if (tm < 3) {  // new-new, old-old, new-old
          u = s_th[tm == 1][d];
          v = s_th[tm >= 1][d];
          off = s_mm[tm == 1][d] - s_mm[tm >= 1][d];
          cf = tm < 2 ? 0.5 * q * q : -q * q;
        } else {       // trial new/old against pending move k's new/old
          const int k = (tm - 3) >> 2, j = (tm - 3) & 3;
          u = s_th[j >> 1][d];
          v = s_lth[k][j & 1][d];
          off = s_mm[j >> 1][d] - s_lmm[k][j & 1][d];
          cf = ((j == 0 || j == 3) ? q : -q) * s_lq[k];
        }
        s_cv[tm][d][dd] = chain_xcorr<PP>(u, v, dd);
        if (dd == 0) s_cd[tm][d] = off;
        if (dd == 0 && d == 0) s_cf[tm] = cf;
      }
      __syncthreads();
      const int K = a.m.K, H = a.m.H;
      for (int r = tid; r < nt * W * W; r += kChainThreads) {
        const int dy = r % W, dz = (r / W) % W, tm = r / (W * W);
        const int oy = chain_fold(s_cd[tm][1] + dy - (PP - 1), K), oz = chain_fold(s_cd[tm][2] + dz - (PP - 1), K);
        const double* grow = a.gk + (size_t)H * (oy + (size_t)H * oz);
        double sx = 0.0;
#pragma unroll
        for (int dx = 0; dx < W; ++dx) sx += s_cv[tm][0][dx] * grow[chain_fold(s_cd[tm][0] + dx - (PP - 1), K)];
        drec += s_cf[tm] * s_cv[tm][1][dy] * s_cv[tm][2][dz] * sx;
      }
      // sum dQ phi over the trial's two stencils
      for (int i3 = tid; i3 < PP * PP * PP; i3 += kChainThreads) {
        const int ix = i3 % PP, iy = (i3 / PP) % PP, iz = i3 / (PP * PP);
        double sp = 0.0;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          int x = s_mm[w][0] - ix, y = s_mm[w][1] - iy, z = s_mm[w][2] - iz;
          x += x < 0 ? K : 0;
          y += y < 0 ? K : 0;
          z += z < 0 ? K : 0;
          const size_t mi = x + (size_t)K * (y + (size_t)K * z);
          const double ph = a.phi_fp64 ? ((const double*)a.phi)[mi] : (double)((const float*)a.phi)[mi];
          const double wt = s_th[w][0][ix] * s_th[w][1][iy] * s_th[w][2][iz];
          sp += w == 0 ? wt * ph : -wt * ph;
        }
        drec += q * sp;
      }
    } else if (a.ewald == 1) {
#pragma unroll
      for (int u = 0; u < KV; ++u) {
        const int kn = tid + u * kChainThreads;
        if (kn >= a.nk) continue;
        const int nx = a.kvec[3 * kn], ny = a.kvec[3 * kn + 1], nz = a.kvec[3 * kn + 2];
        const double2 ex = s_eik[0][0][nx], ey = s_eik[0][1][ny], ez = s_eik[0][2][nz];
        const double2 fx = s_eik[1][0][nx], fy = s_eik[1][1][ny], fz = s_eik[1][2][nz];
        const double cxn = ex.x, sxn = ex.y, cyn = ey.x, syn = ey.y, czn = ez.x, szn = ez.y;
        const double cxo = fx.x, sxo = fx.y, cyo = fy.x, syo = fy.y, czo = fz.x, szo = fz.y;
        double tn = 0.0, to = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double gxn = (c & 2) ? sxn : -sxn, gyn = (c & 1) ? syn : -syn;
          const double gxo = (c & 2) ? sxo : -sxo, gyo = (c & 1) ? syo : -syo;
          const double r1n = cxn * cyn - gxn * gyn, i1n = cxn * gyn + gxn * cyn;
          const double r1o = cxo * cyo - gxo * gyo, i1o = cxo * gyo + gxo * cyo;
          const double dr = q * ((r1n * czn - i1n * szn) - (r1o * czo - i1o * szo));
          const double di = q * ((r1n * szn + i1n * czn) - (r1o * szo + i1o * czo));
          dS[u][2 * c] = dr;
          dS[u][2 * c + 1] = di;
          const double Sr = a.sk[8 * kn + 2 * c], Si = a.sk[8 * kn + 2 * c + 1];
          tn += (Sr + dr) * (Sr + dr) + (Si + di) * (Si + di);
          to += Sr * Sr + Si * Si;
        }
        drec += a.kfac[kn] * (tn - to);
      }
    }
    const double etot = chain_sum(e, sh);
    const double rtot = chain_sum(drec, sh);
    __syncthreads();
    if (tid == 0) {
      double du_rec = 0.0;
      if (a.ewald) {
        double sr = rtot;
        if (a.lsurf) {
          double m2o = 0.0, m2n = 0.0;
          for (int d = 0; d < 3; ++d) {
            const double mn = a.Mout[d] + q * (rn[d] - ro[d]);
            m2o += a.Mout[d] * a.Mout[d];
            m2n += mn * mn;
          }
          sr += a.surf_fac * (m2n - m2o);
        }
        du_rec = a.epsi4pi * sr;
      }
      const double du = etot + du_rec;
      bool acc;
      if (s_ovl) {
        acc = false;
      } else {
        const double dured = a.beta * du;
        if (dured > expmax) acc = false;
        else if (dured < -expmax) acc = true;
        else {
          const double fac = exp(-dured);
          acc = fac > 1.0 ? true : fac > chain_random(s_rng, &s_am);
        }
      }
      a.events[t] = s_ovl ? -1 : (acc ? 1 : 0);
      s_acc = acc;
      if (acc) {
        dsum += du;
        ++nacc;
      }
    }
    __syncthreads();
    ++done;
    if (s_acc) {  // MCUpdate in place
      if constexpr (P > 0) {  // the move joins the pending corrections of phi
        if (tid < 6 * PP) {
          const int w = tid / (3 * PP), d = (tid / PP) % 3, k = tid % PP;
          s_lth[npend][w][d][k] = s_th[w][d][k];
        }
        if (tid < 6) s_lmm[npend][tid / 3][tid % 3] = s_mm[tid / 3][tid % 3];
        if (tid == 0) s_lq[npend] = q;
        ++npend;
      } else if (a.ewald == 1) {
#pragma unroll
        for (int u = 0; u < KV; ++u) {
          const int kn = tid + u * kChainThreads;
          if (kn >= a.nk) continue;
#pragma unroll
          for (int c = 0; c < 8; ++c) a.sk[8 * kn + c] += dS[u][c];
        }
      }
      ncross += cell_move<kChainThreads>(cm, ip, rn);
      if (tid == 0) {
        for (int d = 0; d < 3; ++d) {
          a.Mout[d] += q * (rn[d] - ro[d]);
          a.Mtot[d] = a.Mout[d];
          a.ro[3 * ip + d] = rn[d];
        }
      }
      __threadfence_block();
    }
    __syncthreads();
    if (P > 0 && npend >= a.maxlog) break;  // phi is recomputed before the chain goes on
  }
  if (tid == 0) {
    for (int k = 0; k < 3; ++k) a.rng[k] = s_rng[k];
    *a.am = s_am;
    *a.next = s_next;
    a.stat[0] = dsum;
    a.stat[1] = nacc;
    a.stat[2] = done;
    a.stat[3] = ncross;
  }
}

template <int KV, int P>
__global__ __launch_bounds__(kChainThreads) void k_mc_chain(ChainArgs a) {
  chain_body<KV, P>(a);
}

// One chain per system: workgroup b runs the chain of args[b] (an ensemble's passes side by
// side, one workgroup each).
template <int KV, int P>
__global__ __launch_bounds__(kChainThreads) void k_mc_chain_seg(const ChainArgs* __restrict__ args) {
  chain_body<KV, P>(args[blockIdx.x]);
}

}  // namespace

bool chain_eligible(const Ctx* c) {
  // systems with weak charges run their trials from the host (msg_dutotal honours the states)
  if (c->P.list_mode == MSG_LIST_VERLET || !c->h_laz.empty()) return false;
  if (c->P.ewald_mode == MSG_EWALD_SPM) return c->mesh.order >= 2 && c->mesh.order <= 11;
  return c->P.ewald_mode != MSG_EWALD_STD || (c->nkvec <= kChainThreads * kChainKvec && c->P.ncut <= kChainNcut);
}

namespace {

// Per-chain device state: ints {iseed, xorshift word, Park-Miller word, sequential cursor};
// doubles {am, stat[4] (sum of accepted dE, accepted, trials done, moves into another cell),
// 3 unused, dtran per type}.
constexpr int kChainInts = 4;
constexpr int kChainDoubles = 8 + kMaxT;

// The chain's view of a context; the state pointers and the trial count are set per launch.
ChainArgs chain_args(Ctx* c, int isamp, double beta) {
  ChainArgs a;
  a.np = c->np;
  a.ntrial = 0;
  a.isamp = isamp;
  a.ewald = c->P.ewald_mode;
  a.lsurf = c->P.lsurf;
  a.ro = c->d_ro;
  a.az = c->d_az;
  a.ipt = c->d_ipt;
  a.cpos = (double4*)c->d_cpos;
  a.cell_ip = c->d_cell_ip;
  a.cell_start = c->d_cell_start;
  a.slot_of = c->d_slot_of;
  a.slot_cell = c->d_slot_cell;
  a.ctype = c->d_ctype;
  a.crel = (float4*)c->d_crel;
  a.g = c->grid;
  for (int d = 0; d < 3; ++d) {
    a.box.L[d] = c->P.boxlen[d];
    a.box.half[d] = 0.5 * c->P.boxlen[d];
    a.box.inv[d] = 1.0 / c->P.boxlen[d];
    a.tpb[d] = 2.0 * M_PI / c->P.boxlen[d];
  }
  a.rcut2 = c->P.rcut * c->P.rcut;
  a.cut_inclusive = c->P.list_mode != MSG_LIST_CELL;
  a.tcoef = c->d_tcoef64;
  a.tbins = c->d_tbins;
  a.tm = c->tmeta;
  a.nk = c->P.ewald_mode == MSG_EWALD_STD ? c->nkvec : 0;
  a.ncut = c->P.ewald_mode == MSG_EWALD_STD ? c->P.ncut : 0;
  a.kvec = c->d_kvec;
  a.kfac = c->d_kfac;
  a.sk = c->d_sk;
  a.m = c->mesh;
  a.phi = c->d_Q;
  a.phi_fp64 = c->fp64;
  a.gk = c->d_gk;
  a.Mout = c->d_out + c->nptpt + 4;
  a.Mtot = c->d_M;
  const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
  a.surf_fac = 2.0 * M_PI / (3.0 * vol);
  a.epsi4pi = c->P.epsi4pi;
  a.beta = beta;
  a.maxlog = kChainLog;
  return a;
}

void chain_state(ChainArgs& a, int* d_i, double* d_d, int* d_events) {
  a.rng = d_i;
  a.next = d_i + 3;
  a.am = d_d;
  a.stat = d_d + 1;
  a.dtran = d_d + 8;
  a.events = d_events;
}

// order: the SPME order, 0 without SPME (kv then picks the k-vectors per thread)
template <int KV, int P>
void launch_one(hipStream_t s, const ChainArgs* a, const ChainArgs* d_args, int n) {
  if (d_args)
    hipLaunchKernelGGL((k_mc_chain_seg<KV, P>), dim3(n), dim3(kChainThreads), 0, s, d_args);
  else
    hipLaunchKernelGGL((k_mc_chain<KV, P>), dim3(1), dim3(kChainThreads), 0, s, *a);
}

void launch_chain(hipStream_t s, int order, int nk, const ChainArgs* a, const ChainArgs* d_args, int n) {
  switch (order) {
    case 0:
      if (nk <= kChainThreads) launch_one<1, 0>(s, a, d_args, n);
      else if (nk <= 2 * kChainThreads) launch_one<2, 0>(s, a, d_args, n);
      else launch_one<4, 0>(s, a, d_args, n);
      break;
    case 2: launch_one<1, 2>(s, a, d_args, n); break;
    case 3: launch_one<1, 3>(s, a, d_args, n); break;
    case 4: launch_one<1, 4>(s, a, d_args, n); break;
    case 5: launch_one<1, 5>(s, a, d_args, n); break;
    case 6: launch_one<1, 6>(s, a, d_args, n); break;
    case 7: launch_one<1, 7>(s, a, d_args, n); break;
    case 8: launch_one<1, 8>(s, a, d_args, n); break;
    case 9: launch_one<1, 9>(s, a, d_args, n); break;
    case 10: launch_one<1, 10>(s, a, d_args, n); break;
    default: launch_one<1, 11>(s, a, d_args, n); break;
  }
}

int chain_order(const Ctx* c) { return c->P.ewald_mode == MSG_EWALD_SPM ? c->mesh.order : 0; }

struct DeviceBuffers {
  std::vector<void*> p;
  ~DeviceBuffers() {
    for (void* q : p) hipFree(q);
  }
  int alloc(void** out, size_t bytes) {
    *out = nullptr;
    if (hipMalloc(out, std::max(bytes, (size_t)16)) != hipSuccess)
      return set_error(MSG_ERR_NOMEM, "mc_chain: device allocation failed");
    p.push_back(*out);
    return MSG_OK;
  }
};

}  // namespace

// Runs ntrial trials of the chain: one launch, or with SPME one launch per kChainLog accepted
// moves with a UTotal pass (phi of the current positions) between them.
int mc_chain(Ctx* c, int ntrial, int isamp, const double* dtran, double beta, int* iseed, int* rng_state,
             double* am, double* du_sum, int* naccept, int* events) {
  DeviceBuffers buf;
  int* d_i = nullptr;
  double* d_d = nullptr;
  int* d_ev = nullptr;
  if (int e = buf.alloc((void**)&d_i, kChainInts * sizeof(int))) return e;
  if (int e = buf.alloc((void**)&d_d, kChainDoubles * sizeof(double))) return e;
  if (int e = buf.alloc((void**)&d_ev, (size_t)ntrial * sizeof(int))) return e;
  int hi[kChainInts] = {*iseed, rng_state[0], rng_state[1], c->mc_next};
  double hd[kChainDoubles] = {};
  hd[0] = *am;
  for (int t = 0; t < c->npt; ++t) hd[8 + t] = dtran[t];
  MSG_HIP(hipMemcpyAsync(d_i, hi, sizeof(hi), hipMemcpyHostToDevice, c->stream));
  MSG_HIP(hipMemcpyAsync(d_d, hd, sizeof(hd), hipMemcpyHostToDevice, c->stream));
  const bool spm = c->P.ewald_mode == MSG_EWALD_SPM;
  if (int e = spme_trial_setup(c)) return e;
  ChainArgs a = chain_args(c, isamp, beta);
  chain_state(a, d_i, d_d, d_ev);
  double dsum = 0.0;
  int nacc = 0, done = 0;
  do {
    if (!c->fresh)
      if (int e = utotal_refresh(c)) return e;
    a.ntrial = ntrial - done;
    a.events = d_ev + done;
    prof_begin(c, MSG_K_DE);
    launch_chain(c->stream, chain_order(c), a.nk, &a, nullptr, 1);
    prof_end(c, MSG_K_DE);
    MSG_HIP(hipGetLastError());
    double st[4];
    MSG_HIP(hipMemcpyAsync(st, d_d + 1, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    MSG_HIP(hipStreamSynchronize(c->stream));
    dsum += st[0];
    nacc += (int)st[1];
    done += (int)st[2];
    if (st[3] > 0.0) c->cells_maintained = true;  // the cell order carries the chain's moves
    if (spm && st[1] > 0.0) c->fresh = false;     // phi is that of the positions before the moves
    if (st[2] < 1.0 && done < ntrial) return set_error(MSG_ERR_HIP, "mc_chain: a launch made no trial");
  } while (done < ntrial);
  MSG_HIP(hipMemcpyAsync(hi, d_i, sizeof(hi), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipMemcpyAsync(hd, d_d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (events && ntrial > 0)
    MSG_HIP(hipMemcpyAsync(events, d_ev, ntrial * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  *iseed = hi[0];
  rng_state[0] = hi[1];
  rng_state[1] = hi[2];
  *am = hd[0];
  c->mc_next = hi[3];
  if (du_sum) *du_sum = dsum;
  if (naccept) *naccept = nacc;
  return MSG_OK;
}

// The chains of many systems side by side: one workgroup per system and one launch per SPME
// order present; refresh() runs a UTotal pass of every system (phi of the current positions)
// after a launch in which an SPME chain filled its log of accepted moves.
int mc_chain_batch(const std::vector<Ctx*>& cs, hipStream_t s, const std::function<int()>& refresh, int ntrial,
                   int isamp, const double* dtran, double beta, int* iseed, int* rng_state, double* am,
                   double* du_sum, int* naccept, int* events) {
  const int n = (int)cs.size();
  for (Ctx* c : cs)
    if (!chain_eligible(c) || c->P.ewald_mode == MSG_EWALD_STD)
      return set_error(MSG_ERR_ARG, "msg_batch_mc_pass: every system must be SPME or Ewald-free, without weak charges");
  DeviceBuffers buf;
  int* d_i = nullptr;
  double* d_d = nullptr;
  int* d_ev = nullptr;
  ChainArgs* d_args = nullptr;
  if (int e = buf.alloc((void**)&d_i, (size_t)n * kChainInts * sizeof(int))) return e;
  if (int e = buf.alloc((void**)&d_d, (size_t)n * kChainDoubles * sizeof(double))) return e;
  if (int e = buf.alloc((void**)&d_ev, (size_t)n * ntrial * sizeof(int))) return e;
  if (int e = buf.alloc((void**)&d_args, (size_t)n * sizeof(ChainArgs))) return e;
  std::vector<int> hi((size_t)n * kChainInts);
  std::vector<double> hd((size_t)n * kChainDoubles, 0.0);
  for (int k = 0; k < n; ++k) {
    hi[kChainInts * k] = iseed[k];
    hi[kChainInts * k + 1] = rng_state[2 * k];
    hi[kChainInts * k + 2] = rng_state[2 * k + 1];
    hi[kChainInts * k + 3] = cs[k]->mc_next;
    hd[(size_t)kChainDoubles * k] = am[k];
    for (int t = 0; t < cs[k]->npt; ++t) hd[(size_t)kChainDoubles * k + 8 + t] = dtran[(size_t)MSG_MAX_TYPES * k + t];
    if (int e = spme_trial_setup(cs[k])) return e;
  }
  MSG_HIP(hipMemcpyAsync(d_i, hi.data(), hi.size() * sizeof(int), hipMemcpyHostToDevice, s));
  MSG_HIP(hipMemcpyAsync(d_d, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice, s));
  // systems grouped by SPME order (one kernel instance each), contiguous in the argument array
  std::vector<int> idx(n);
  for (int k = 0; k < n; ++k) idx[k] = k;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return chain_order(cs[x]) < chain_order(cs[y]); });
  std::vector<ChainArgs> args(n);
  for (int r = 0; r < n; ++r) {
    const int k = idx[r];
    args[r] = chain_args(cs[k], isamp, beta);
    args[r].maxlog = chain_log_cap(chain_order(cs[k]));  // fewer batched refreshes
    chain_state(args[r], d_i + kChainInts * k, d_d + (size_t)kChainDoubles * k, d_ev + (size_t)ntrial * k);
  }
  std::vector<int> done(n, 0);
  std::vector<double> dsum(n, 0.0), st((size_t)4 * n);
  std::vector<long long> nacc(n, 0);
  bool any_spm_accept = false;
  while (true) {
    bool left = false;
    for (int k = 0; k < n; ++k) left |= done[k] < ntrial;
    if (!left) break;
    for (int r = 0; r < n; ++r) {
      const int k = idx[r];
      args[r].ntrial = ntrial - done[k];
      args[r].events = d_ev + (size_t)ntrial * k + done[k];
    }
    MSG_HIP(hipMemcpyAsync(d_args, args.data(), n * sizeof(ChainArgs), hipMemcpyHostToDevice, s));
    for (int r0 = 0; r0 < n;) {
      int r1 = r0;
      while (r1 < n && chain_order(cs[idx[r1]]) == chain_order(cs[idx[r0]])) ++r1;
      launch_chain(s, chain_order(cs[idx[r0]]), 0, nullptr, d_args + r0, r1 - r0);
      MSG_HIP(hipGetLastError());
      r0 = r1;
    }
    for (int k = 0; k < n; ++k)
      MSG_HIP(hipMemcpyAsync(st.data() + 4 * k, d_d + (size_t)kChainDoubles * k + 1, 4 * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    MSG_HIP(hipStreamSynchronize(s));
    any_spm_accept = false;
    for (int k = 0; k < n; ++k) {
      dsum[k] += st[4 * k];
      nacc[k] += (long long)st[4 * k + 1];
      done[k] += (int)st[4 * k + 2];
      if (st[4 * k + 3] > 0.0) cs[k]->cells_maintained = true;
      if (cs[k]->P.ewald_mode == MSG_EWALD_SPM && st[4 * k + 1] > 0.0) {
        cs[k]->fresh = false;
        any_spm_accept = true;
      }
    }
    left = false;
    for (int k = 0; k < n; ++k) left |= done[k] < ntrial;
    if (left && any_spm_accept)
      if (int e = refresh()) return e;
  }
  MSG_HIP(hipMemcpyAsync(hi.data(), d_i, hi.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  MSG_HIP(hipMemcpyAsync(hd.data(), d_d, hd.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (events && ntrial > 0)
    MSG_HIP(hipMemcpyAsync(events, d_ev, (size_t)n * ntrial * sizeof(int), hipMemcpyDeviceToHost, s));
  MSG_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < n; ++k) {
    iseed[k] = hi[kChainInts * k];
    rng_state[2 * k] = hi[kChainInts * k + 1];
    rng_state[2 * k + 1] = hi[kChainInts * k + 2];
    cs[k]->mc_next = hi[kChainInts * k + 3];
    am[k] = hd[(size_t)kChainDoubles * k];
    if (du_sum) du_sum[k] = dsum[k];
    if (naccept) naccept[k] = (int)nacc[k];
  }
  return MSG_OK;
}

}  // namespace msg
