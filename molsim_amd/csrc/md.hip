// Verlet-list skin test (NList, Src/nlist.F90:553-669) and the velocity-Verlet caller of
// UTotal (VelVer, Src/md.F90:252-410), on the device.
//
//   k_drift  drostep = dt v + (dt^2/2) a, ro += drostep, PBC with ANINT (Src/md.F90:289-298,
//            Src/molaux.F90:104-140); drosum += drostep and |drosum|^2 per particle with the
//            maxima of 256-particle blocks, for the skin test.
//   k_disp   |ro - ro_ref|^2 (minimum image) per particle and the block maxima: the skin test of
//            a UTotal caller that moves particles itself (displacement since the build).
//   k_skin_scan, k_skin
//            the skin test's r2max1 and r2max2 over the particles in index order
//            (Src/nlist.F90:600-612): r2max1 is the maximum; r2max2 is the largest value that
//            was not a new maximum when the reference's loop reached it (a new maximum replaces
//            r2max1 without passing the old one on).  One workgroup scans the block maxima into
//            exclusive prefix maxima; then every block finds, from its prefix and a prefix-max
//            scan of its own values, the values that are no new maximum, and adds their maximum
//            to r2max2 with one atomic.  The host rebuilds when sqrt(r2max1) + sqrt(r2max2) >
//            drnlist.
//   k_vl_records  the list's slot-ordered records of the current positions (each pass
//            without a rebuild; the build writes them itself).
//   k_kick   a = F / m, v += (dt/2) (a + a_old) (Src/md.F90:348-358), kinetic-energy partials.
//   k_mdlog  per-step record of U_tot and the kinetic energy.
#include <cstring>

#include "common.h"

namespace msg {

namespace {

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// The maximum of a 256-thread workgroup's values into bm[blockIdx.x].
__device__ inline void block_max_out(double v, double* bm) {
  __shared__ double sh[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) bm[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// The Verlet list's slot-ordered records (position, type + kUncharged when uncharged) of the
// current positions: one thread per slot gathers its particle's (a gather with coalesced
// stores, where a scatter from the drift would read each record's line before writing it).
// With vl_rel (mixed path) also the fixed-point record relative to the corner of the slot's
// cell at the build (the cell is kept in the record): (ro + L/2) - cell * cellsize in FP64,
// taken back by L when the particle crossed the periodic boundary since the build, in units of
// cellsize / 2^28 (rounded to nearest).
struct VlRecArgs {
  int np;
  const int* vl_home;
  const double* ro;
  const int* ipt;
  const int* laz;
  double4* vl_pos;   // null when the pass's pair loop reads only the fixed-point records
  int4* vl_rel;
  Box box;
  double csz[3];
};

__global__ __launch_bounds__(256) void k_vl_records(VlRecArgs a) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= a.np) return;
  const int i = a.vl_home[s];
  const int tag = a.ipt[i] + ((a.laz && !a.laz[i]) ? kUncharged : 0);
  const double r[3] = {a.ro[3 * i], a.ro[3 * i + 1], a.ro[3 * i + 2]};
  if (a.vl_pos) a.vl_pos[s] = make_double4(r[0], r[1], r[2], (double)tag);
  if (a.vl_rel) {
    const int cw = a.vl_rel[s].w & ((1 << 27) - 1);
    int q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double x = (r[d] + a.box.half[d]) - ((cw >> (9 * d)) & 511) * a.csz[d];
      if (fabs(x) > a.box.half[d]) x -= copysign(a.box.L[d], x);
      q[d] = __double2int_rn(x * (kVlUnits / a.csz[d]));
    }
    a.vl_rel[s] = make_int4(q[0], q[1], q[2], cw | (tag << 27));
  }
}

__global__ __launch_bounds__(256) void k_disp(int np, const double* __restrict__ ro,
                                              const double* __restrict__ ref, Box box,
                                              double* __restrict__ r2, double* __restrict__ bm) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < np) {
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double x = ro[3 * i + k] - ref[3 * i + k];
      if (fabs(x) > box.half[k]) x -= copysign(box.L[k], x);
      d[k] = x;
    }
    s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    r2[i] = s;
  }
  block_max_out(s, bm);
}

// Exclusive prefix maxima pm[b] = max(bm[0..b)) (0 for b = 0) by one 1024-thread workgroup;
// skin[0] = bits of the maximum over all blocks (r2max1), skin[1] = 0 (r2max2 is added by
// k_skin).
__global__ __launch_bounds__(1024) void k_skin_scan(int nb, const double* __restrict__ bm,
                                                    double* __restrict__ pm,
                                                    unsigned long long* skin) {
  __shared__ double wtot[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (nb + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nb, b0 + per);
  // up to eight of the thread's block maxima kept in registers (their loads in flight
  // together), the rest read in the loops
  constexpr int R = 8;
  double keep[R];
#pragma unroll
  for (int u = 0; u < R; ++u) keep[u] = b0 + u < b1 ? bm[b0 + u] : 0.0;
  double own = 0.0;
#pragma unroll
  for (int u = 0; u < R; ++u) own = fmax(own, keep[u]);
  for (int b = b0 + R; b < b1; ++b) own = fmax(own, bm[b]);
  // inclusive scan of the threads' maxima: wavefront, then the wavefront totals
  double inc = own;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(inc, off, 64);
    if (lane >= off) inc = fmax(inc, t);
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  double before = 0.0;
  for (int k = 0; k < w; ++k) before = fmax(before, wtot[k]);
  double run = fmax(before, __shfl_up(inc, 1, 64));
  if (lane == 0) run = before;
#pragma unroll
  for (int u = 0; u < R; ++u)
    if (b0 + u < b1) {
      pm[b0 + u] = run;
      run = fmax(run, keep[u]);
    }
  for (int b = b0 + R; b < b1; ++b) {
    pm[b] = run;
    run = fmax(run, bm[b]);
  }
  if (tid == 0) {
    double all = 0.0;
    for (int k = 0; k < 16; ++k) all = fmax(all, wtot[k]);
    skin[0] = (unsigned long long)__double_as_longlong(all);
    skin[1] = 0ull;
  }
}

// r2max2: per block, the values r2[i] that are not above the maximum of r2[0..i) (a running
// maximum starting at 0); their maximum is added to skin[1] (non-negative doubles order as
// their bit patterns).
__global__ __launch_bounds__(256) void k_skin(int np, const double* __restrict__ r2,
                                              const double* __restrict__ pm,
                                              unsigned long long* skin) {
  __shared__ double wtot[4];
  __shared__ double cand[4];
  const int i = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double v = i < np ? r2[i] : 0.0;
  double inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(inc, off, 64);
    if (lane >= off) inc = fmax(inc, t);
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  double before = pm[blockIdx.x];
  for (int k = 0; k < w; ++k) before = fmax(before, wtot[k]);
  const double up = __shfl_up(inc, 1, 64);
  const double excl = lane == 0 ? before : fmax(before, up);
  // the reference's loop: v > r2max1 makes v the new r2max1; otherwise v may raise r2max2
  double c = (i < np && !(v > excl)) ? v : 0.0;
  c = wave_max(c);
  if (lane == 0) cand[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmax(fmax(cand[0], cand[1]), fmax(cand[2], cand[3]));
    // the atomic only when the value may raise the maximum so far (a stale read only lets
    // an unneeded atomic through): thousands of workgroups on one word otherwise queue up
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    if (m > 0.0 && bits > __atomic_load_n(skin + 1, __ATOMIC_RELAXED)) atomicMax(skin + 1, bits);
  }
}

struct MdArgs {
  int np;
  double dt, fac;       // tstep and Half * tstep**2
  Box box;
  double* ro;
  double* vel;
  double* acc;
  const double* force;
  const double* minv;   // per type
  const int* ipt;
  double* drosum;       // displacements summed since the list build (null: no skin test)
  int reset;            // a list was built since the last drift: the sums start again
  double* r2;           // |drosum|^2 per particle and block maxima for the skin test
  double* bm;
  double* ke_part;
};

__global__ __launch_bounds__(256) void k_drift(MdArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < a.np) {
    double ds[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ds[d] = a.dt * a.vel[3 * i + d] + a.fac * a.acc[3 * i + d];
      double x = a.ro[3 * i + d] + ds[d];
      // PBC: if (abs(x) > L/2) x = x - L * ANINT(x / L)
      if (fabs(x) > a.box.half[d]) x -= a.box.L[d] * round(x * a.box.inv[d]);
      a.ro[3 * i + d] = x;
    }
    if (a.drosum) {
      double q[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        q[d] = (a.reset ? 0.0 : a.drosum[3 * i + d]) + ds[d];
        a.drosum[3 * i + d] = q[d];
      }
      s = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
      a.r2[i] = s;
    }
  }
  if (a.drosum) block_max_out(s, a.bm);  // a.drosum is the same for the whole grid
}

__global__ __launch_bounds__(256) void k_kick(MdArgs a, int set_only) {
#pragma clang fp contract(off)
  __shared__ double sh[8];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double ke = 0.0;
  if (i < a.np) {
    const double mi = a.minv[a.ipt[i]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double an = mi * a.force[3 * i + d];
      double v = a.vel[3 * i + d];
      if (!set_only) v += (0.5 * a.dt) * (an + a.acc[3 * i + d]);
      a.vel[3 * i + d] = v;
      a.acc[3 * i + d] = an;
      ke += 0.5 * v * v / mi;
    }
  }
  ke = msgd::block_sum(ke, sh);
  if (threadIdx.x == 0) a.ke_part[blockIdx.x] = ke;
}

__global__ __launch_bounds__(256) void k_mdlog(const double* __restrict__ ke_part, int nb,
                                               const double* __restrict__ out, int ntp,
                                               double* __restrict__ log) {
  __shared__ double sh[8];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += ke_part[b];
  s = msgd::block_sum(s, sh);
  if (threadIdx.x == 0) {
    log[0] = out[ntp + 2];
    log[1] = s;
  }
}

__global__ void k_clear_sticky(int* flags) {
  if (threadIdx.x == 0) flags[1] = 0;
}

Box make_box(const msg_params& P) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = P.boxlen[d];
    b.half[d] = 0.5 * P.boxlen[d];
    b.inv[d] = 1.0 / P.boxlen[d];
  }
  return b;
}

MdArgs md_args(Ctx* c, double dt, bool skin) {
  MdArgs a;
  a.np = c->np;
  a.dt = dt;
  a.fac = 0.5 * (dt * dt);
  a.box = make_box(c->P);
  a.ro = c->d_ro;
  a.vel = c->d_vel;
  a.acc = c->d_acc;
  a.force = c->d_force;
  a.minv = c->d_minv;
  a.ipt = c->d_ipt;
  a.drosum = skin ? c->d_drosum : nullptr;
  a.reset = c->drosum_reset ? 1 : 0;
  a.r2 = c->d_r2;
  a.bm = c->d_bm;
  a.ke_part = c->d_ke_part;
  return a;
}

// Cell order and Verlet list of the current positions at the end of an MD step (NListSub of
// NList(iSimulationStep)).  The binning, sorting and ranking read the positions, so they run on
// the main stream before the next drift; the list build reads only the cell-ordered records and
// runs on the side stream, beside the next step's drift and mesh work, ahead of the pair kernel
// queued there.
int md_rebuild(Ctx* c) {
  int e;
  if ((e = launch_bin(c, true, false)) || (e = launch_sort(c, 1)) || (e = launch_finalize(c)) ||
      (e = launch_zero_counts(c)))
    return e;
  MSG_HIP(hipEventRecord(c->ev_cells, c->stream));
  MSG_HIP(hipStreamWaitEvent(c->side, c->ev_cells, 0));
  hipStream_t main = c->stream;
  c->stream = c->side;
  e = launch_vlist_build(c);
  c->stream = main;
  if (e) return e;
  c->vl_valid = true;
  c->cells_valid = true;
  return MSG_OK;
}

}  // namespace

int verlet_setup(Ctx* c) {
  const msg_params& P = c->P;
  if (!(P.drnlist >= 0.0)) return set_error(MSG_ERR_ARG, "Verlet list: drnlist must be >= 0");
  const double vol = P.boxlen[0] * P.boxlen[1] * P.boxlen[2];
  const double rl = P.rcut + P.drnlist;
  const double expect = c->np / vol * 4.0 / 3.0 * M_PI * rl * rl * rl;
  int maxn = (int)std::ceil(2.0 * expect + 48.0);
  maxn = std::min(((maxn + 31) / 32) * 32, std::max(c->np - 1, 32));
  if (c->vl_width_user > 0) maxn = c->vl_width_user;
  const int np = c->np;
  if (maxn > c->vl_cap) {  // the list width grows with the density (a smaller box)
    int* nbr = nullptr;
    if (int e = dalloc(c, (void**)&nbr, (size_t)maxn * np * sizeof(int))) return e;
    if (c->d_vl_nbr) {
      hipFree(c->d_vl_nbr);
      c->bytes -= (long long)c->vl_cap * np * (long long)sizeof(int);
    }
    c->d_vl_nbr = nbr;
    c->vl_cap = maxn;
  }
  c->vl_maxn = maxn;
  if (c->d_vl_home) return MSG_OK;  // the per-particle buffers exist already
  const int nb = (np + 255) / 256;
  if (int e = dalloc(c, (void**)&c->d_vl_home, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_n, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_pos, 4 * (size_t)np * sizeof(double))) return e;
  if (!c->fp64)
    if (int e = dalloc(c, (void**)&c->d_vl_rel, 4 * (size_t)np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_slot, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_ro_ref, 3 * (size_t)np * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_drosum, 3 * (size_t)np * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_r2, (size_t)np * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_bm, (size_t)nb * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_pm, (size_t)nb * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_skin, 2 * sizeof(unsigned long long))) return e;
  if (hipHostMalloc((void**)&c->h_skin, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
    return set_error(MSG_ERR_NOMEM, "hipHostMalloc failed");
  MSG_HIP(hipMemset(c->d_drosum, 0, 3 * (size_t)np * sizeof(double)));
  return MSG_OK;
}

int launch_vl_records(Ctx* c) {
  VlRecArgs a;
  a.np = c->np;
  a.vl_home = c->d_vl_home;
  a.ro = c->d_ro;
  a.ipt = c->d_ipt;
  a.laz = c->d_laz;
  a.vl_rel = vl_rel_ok(c) ? (int4*)c->d_vl_rel : nullptr;
  a.vl_pos = (a.vl_rel && c->inlist == 0) ? nullptr : (double4*)c->d_vl_pos;  // as launch_pair chooses
  a.box = make_box(c->P);
  for (int d = 0; d < 3; ++d) a.csz[d] = c->P.boxlen[d] / c->grid.nc[d];
  hipLaunchKernelGGL(k_vl_records, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, a);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// The skin test over d_r2 / d_bm (filled by k_drift or k_disp): r2max1 and r2max2 into d_skin
// and their copy to the pinned h_skin, marked by ev_skin.
int launch_skin(Ctx* c) {
  const int nb = (c->np + 255) / 256;
  hipLaunchKernelGGL(k_skin_scan, dim3(1), dim3(1024), 0, c->stream, nb, c->d_bm, c->d_pm, c->d_skin);
  hipLaunchKernelGGL(k_skin, dim3(nb), dim3(256), 0, c->stream, c->np, c->d_r2, c->d_pm, c->d_skin);
  MSG_HIP(hipGetLastError());
  MSG_HIP(hipMemcpyAsync(c->h_skin, c->d_skin, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipEventRecord(c->ev_skin, c->stream));
  return MSG_OK;
}

bool skin_decide(const Ctx* c) {
  double m1, m2;
  std::memcpy(&m1, &c->h_skin[0], sizeof(double));
  std::memcpy(&m2, &c->h_skin[1], sizeof(double));
  return std::sqrt(m1) + std::sqrt(m2) > c->P.drnlist;
}

// The skin test of a UTotal caller in two halves, so that work that does not depend on its
// outcome can be queued in between: the first queues the displacement kernels and the copy of
// (r2max1, r2max2) (pending = false when no list exists yet: a build is due anyway); the second
// waits for the copy alone and decides.  The displacement of a particle is its minimum-image
// distance from where it was at the build.
int verlet_check_begin(Ctx* c, bool& pending) {
  pending = false;
  if (!c->vl_valid) return MSG_OK;
  const int nb = (c->np + 255) / 256;
  hipLaunchKernelGGL(k_disp, dim3(nb), dim3(256), 0, c->stream, c->np, c->d_ro, c->d_ro_ref, make_box(c->P),
                     c->d_r2, c->d_bm);
  MSG_HIP(hipGetLastError());
  if (int e = launch_skin(c)) return e;
  pending = true;
  return MSG_OK;
}

int verlet_check(Ctx* c, bool& rebuild) {
  bool pending;
  if (int e = verlet_check_begin(c, pending)) return e;
  if (!pending) {
    rebuild = true;
    return MSG_OK;
  }
  return verlet_check_end(c, rebuild);
}

int verlet_check_end(Ctx* c, bool& rebuild) {
  MSG_HIP(hipEventSynchronize(c->ev_skin));
  rebuild = skin_decide(c);
  return MSG_OK;
}

}  // namespace msg

using namespace msg;

extern "C" {

int msg_md_init(msg_ctx* h, const double* mass, const double* vel) {
  if (!h || !mass) return set_error(MSG_ERR_ARG, "msg_md_init: null argument");
  Ctx* c = &h->c;
  std::vector<double> minv(c->npt);
  for (int t = 0; t < c->npt; ++t) {
    if (!(mass[t] > 0.0)) return set_error(MSG_ERR_ARG, "msg_md_init: masses must be > 0");
    minv[t] = 1.0 / mass[t];
  }
  const int np = c->np;
  if (!c->d_vel) {
    if (int e = dalloc(c, (void**)&c->d_vel, 3 * (size_t)np * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_acc, 3 * (size_t)np * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_minv, kMaxT * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_ke_part, ((np + 255) / 256) * sizeof(double))) return e;
  }
  MSG_HIP(hipMemcpy(c->d_minv, minv.data(), c->npt * sizeof(double), hipMemcpyHostToDevice));
  if (vel)
    MSG_HIP(hipMemcpy(c->d_vel, vel, 3 * (size_t)np * sizeof(double), hipMemcpyHostToDevice));
  else
    MSG_HIP(hipMemset(c->d_vel, 0, 3 * (size_t)np * sizeof(double)));
  // initial accelerations from the forces of the current positions (the list, in Verlet mode,
  // is built for them: NList(iWriteInput))
  if (int e = utotal_refresh(c)) return e;
  MdArgs a = md_args(c, 0.0, false);
  hipLaunchKernelGGL(k_kick, dim3((np + 255) / 256), dim3(256), 0, c->stream, a, 1);
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->md_ready = true;
  return MSG_OK;
}

// nsteps velocity-Verlet steps, each followed by NList(iSimulationStep) in Verlet mode: with
// inlist = 0 the displacements summed since the last build decide by the skin test, with
// inlist > 0 the list is rebuilt every inlist-th step of the macrostep; a rebuild takes the
// positions of the step just finished, as the reference's does.  The skin test's result is
// read back while the step's UTotal runs, so the host never holds the device up.  Error flags
// of every step are kept (sticky word) and reported at the end.
int msg_md_run(msg_ctx* h, int nsteps, double dt, double* log) {
  if (!h || nsteps < 0) return set_error(MSG_ERR_ARG, "msg_md_run: bad arguments");
  Ctx* c = &h->c;
  if (!c->md_ready) return set_error(MSG_ERR_STATE, "msg_md_init has not been called");
  const bool verlet = c->P.list_mode == MSG_LIST_VERLET;
  double* d_log = nullptr;
  if (log && nsteps > 0) MSG_HIP(hipMalloc(&d_log, 2 * (size_t)nsteps * sizeof(double)));
  const int nb = (c->np + 255) / 256;
  hipLaunchKernelGGL(k_clear_sticky, dim3(1), dim3(64), 0, c->stream, c->d_flags);
  int e = MSG_OK;
  c->md_managed = true;
  for (int s = 0; s < nsteps && !e; ++s) {
    const bool skin = verlet && c->inlist == 0 && c->vl_valid;
    MdArgs a = md_args(c, dt, skin);
    hipLaunchKernelGGL(k_drift, dim3(nb), dim3(256), 0, c->stream, a);
    if (skin) c->drosum_reset = false;
    c->fresh = false;
    c->cells_maintained = false;  // every particle moved: the cells are sorted afresh
    if (skin && (e = launch_skin(c))) break;
    if ((e = utotal_sequence(c))) break;
    a = md_args(c, dt, false);
    hipLaunchKernelGGL(k_kick, dim3(nb), dim3(256), 0, c->stream, a, 0);
    if (d_log)
      hipLaunchKernelGGL(k_mdlog, dim3(1), dim3(256), 0, c->stream, c->d_ke_part, nb, c->d_out,
                         c->nptpt, d_log + 2 * (size_t)s);
    if (verlet) {
      c->istep2++;
      bool rebuild;
      if (c->inlist > 0) {
        rebuild = c->istep2 % c->inlist == 0;
      } else if (skin) {
        if (hipEventSynchronize(c->ev_skin) != hipSuccess) { e = set_error(MSG_ERR_HIP, "skin test copy failed"); break; }
        rebuild = skin_decide(c);
      } else {
        rebuild = false;  // the pass just built the list for these positions
      }
      if (rebuild && (e = md_rebuild(c))) break;
    }
  }
  c->md_managed = false;
  if (e) {
    hipStreamSynchronize(c->stream);
    if (d_log) hipFree(d_log);
    return e;
  }
  c->fresh = true;
  MSG_HIP(hipMemcpyAsync(c->h_out, c->d_out, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (d_log) MSG_HIP(hipMemcpyAsync(log, d_log, 2 * (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  if (d_log) hipFree(d_log);
  // flags of every step's pass (sticky word) and of a rebuild after the last step (word 0,
  // kept for the next pass to report as well)
  const int f = c->h_flags[0] | c->h_flags[1];
  if (f & FLAG_NLIST) c->vl_valid = false;  // a truncated list is not used again
  if (f) {
    const int code = f & FLAG_NLIST ? set_error(MSG_ERR_NLIST, "Verlet list overflow during the MD run")
                     : f & FLAG_R2UMIN ? set_error(MSG_ERR_R2UMIN, "pair below the table: r2 < r2umin")
                     : f & FLAG_CELL   ? set_error(MSG_ERR_CELL, "position outside the box")
                                       : set_error(MSG_ERR_IBUF, "table scan past its end");
    return code;
  }
  return MSG_OK;
}

int msg_nlist_interval(msg_ctx* h, int inlist) {
  if (!h || inlist < 0) return set_error(MSG_ERR_ARG, "msg_nlist_interval: inlist must be >= 0");  // Src/nlist.F90:617-618
  // the mixed pair loop's fixed-point records assume the skin rule's bounded drift since the
  // build: back from a fixed interval, the next pass builds the list afresh
  if (h->c.inlist > 0 && inlist == 0) h->c.vl_valid = false;
  h->c.inlist = inlist;
  return MSG_OK;
}

int msg_set_nlist_width(msg_ctx* h, int maxnneigh) {
  if (!h || maxnneigh < 0) return set_error(MSG_ERR_ARG, "msg_set_nlist_width: maxnneigh must be >= 0");
  Ctx* c = &h->c;
  if (c->P.list_mode != MSG_LIST_VERLET) return set_error(MSG_ERR_STATE, "msg_set_nlist_width: not in Verlet-list mode");
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipStreamSynchronize(c->side));
  c->vl_width_user = maxnneigh;
  c->vl_valid = false;  // the list is built again at the new width
  return verlet_setup(c);
}

int msg_nlist_macrostep(msg_ctx* h) {
  if (!h) return set_error(MSG_ERR_ARG, "msg_nlist_macrostep: null context");
  Ctx* c = &h->c;
  c->istep2 = 0;
  c->drosum_reset = true;  // drosum = 0 without a rebuild, as NList(iBeforeMacrostep)
  return MSG_OK;
}

int msg_get_positions(msg_ctx* h, double* ro) {
  if (!h || !ro) return set_error(MSG_ERR_ARG, "msg_get_positions: null argument");
  Ctx* c = &h->c;
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipMemcpy(ro, c->d_ro, 3 * (size_t)c->np * sizeof(double), hipMemcpyDeviceToHost));
  return MSG_OK;
}

int msg_get_velocities(msg_ctx* h, double* vel) {
  if (!h || !vel) return set_error(MSG_ERR_ARG, "msg_get_velocities: null argument");
  Ctx* c = &h->c;
  if (!c->md_ready) return set_error(MSG_ERR_STATE, "msg_md_init has not been called");
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipMemcpy(vel, c->d_vel, 3 * (size_t)c->np * sizeof(double), hipMemcpyDeviceToHost));
  return MSG_OK;
}

long long msg_nlist_builds(msg_ctx* h) { return h ? h->c.vl_builds : 0; }

}  // extern "C"
