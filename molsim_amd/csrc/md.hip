// Verlet-list displacement test (NList skin rule, Src/nlist.F90:597-621) and the
// velocity-Verlet caller of UTotal (VelVer, Src/md.F90:252-410), on the device.
//
//   k_disp   max |ro - ro_ref|^2 over all particles since the last list build (minimum image).
//            The reference rebuilds when sqrt(r2max1) + sqrt(r2max2) > drnlist over its two
//            largest accumulated displacements; here the list is rebuilt when
//            2 * max |dr| > drnlist, which is never later than the reference's rule.
//   k_drift  ro += dt v + dt^2/2 a, wrapped into the box (PBC, Src/md.F90:289-298), and the
//            displacement maximum of the moved positions.
//   k_vl_records  the list's slot-ordered records of the current positions (each pass
//            without a rebuild; the build writes them itself).
//   k_kick   a = F / m, v += dt/2 (a_old + a) (Src/md.F90:348-358), kinetic-energy partials.
//   k_mdlog  per-step record of U_tot and the kinetic energy.
#include <cstring>

#include "common.h"

namespace msg {

namespace {

__device__ inline void atomic_max_pos(unsigned int* p, float v) {
  atomicMax(p, __float_as_uint(v));  // non-negative floats order as their bit patterns
}

__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// The workgroup's maximum (256 lanes) into its slot: one atomic per workgroup.
__device__ inline void dmax_add(unsigned int* dmax, float d2) {
  __shared__ float sh[4];
  d2 = wave_max(d2);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0)
    atomic_max_pos(dmax + (blockIdx.x % kDmaxSlots) * kDmaxStride,
                   fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3])));
}

// The Verlet list's slot-ordered records (position, type + kUncharged when uncharged) of the
// current positions: one thread per slot gathers its particle's (a gather with coalesced
// stores, where a scatter from the drift would read each record's line before writing it).
__global__ __launch_bounds__(256) void k_vl_records(int np, const int* __restrict__ vl_home,
                                                    const double* __restrict__ ro,
                                                    const int* __restrict__ ipt,
                                                    const int* __restrict__ laz, double4* vl_pos) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= np) return;
  const int i = vl_home[s];
  vl_pos[s] = make_double4(ro[3 * i], ro[3 * i + 1], ro[3 * i + 2],
                           (double)(ipt[i] + ((laz && !laz[i]) ? kUncharged : 0)));
}

__global__ __launch_bounds__(256) void k_disp(int np, const double* __restrict__ ro,
                                              const double* __restrict__ ref, Box box,
                                              unsigned int* dmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  float d2 = 0.f;
  if (i < np) {
    double s = 0.0;
    const double r[3] = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double x = r[d] - ref[3 * i + d];
      x -= box.L[d] * rint(x * box.inv[d]);
      s += x * x;
    }
    d2 = (float)s;
  }
  dmax_add(dmax, d2);
}

struct MdArgs {
  int np;
  double dt;
  Box box;
  double* ro;
  double* vel;
  double* acc;
  const double* force;
  const double* minv;   // per type
  const int* ipt;
  const double* ref;    // Verlet-list reference positions (or null)
  unsigned int* dmax;
  double* ke_part;
};

__global__ __launch_bounds__(256) void k_drift(MdArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  float d2 = 0.f;
  if (i < a.np) {
    double s = 0.0;
    double xs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double x = a.ro[3 * i + d] + a.dt * a.vel[3 * i + d] + 0.5 * a.dt * a.dt * a.acc[3 * i + d];
      x -= a.box.L[d] * rint(x * a.box.inv[d]);  // PBC: back into [-L/2, L/2]
      if (x >= a.box.half[d]) x -= a.box.L[d];
      a.ro[3 * i + d] = x;
      xs[d] = x;
      if (a.ref) {
        double y = x - a.ref[3 * i + d];
        y -= a.box.L[d] * rint(y * a.box.inv[d]);
        s += y * y;
      }
    }
    d2 = (float)s;
  }
  if (a.ref) dmax_add(a.dmax, d2);  // a.ref is the same for the whole grid
}

__global__ __launch_bounds__(256) void k_kick(MdArgs a, int set_only) {
  __shared__ double sh[8];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double ke = 0.0;
  if (i < a.np) {
    const double mi = a.minv[a.ipt[i]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double an = a.force[3 * i + d] * mi;
      double v = a.vel[3 * i + d];
      if (!set_only) v += 0.5 * a.dt * (a.acc[3 * i + d] + an);
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

MdArgs md_args(Ctx* c, double dt) {
  MdArgs a;
  a.np = c->np;
  a.dt = dt;
  for (int d = 0; d < 3; ++d) {
    a.box.L[d] = c->P.boxlen[d];
    a.box.half[d] = 0.5 * c->P.boxlen[d];
    a.box.inv[d] = 1.0 / c->P.boxlen[d];
  }
  a.ro = c->d_ro;
  a.vel = c->d_vel;
  a.acc = c->d_acc;
  a.force = c->d_force;
  a.minv = c->d_minv;
  a.ipt = c->d_ipt;
  a.ref = c->P.list_mode == MSG_LIST_VERLET && c->vl_valid ? c->d_ro_ref : nullptr;
  a.dmax = c->d_dmax;
  a.ke_part = c->d_ke_part;
  return a;
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
  c->vl_maxn = maxn;
  const int np = c->np;
  if (maxn > c->vl_cap) {  // the list width grows with the density (a smaller box)
    if (c->d_vl_nbr) {
      hipFree(c->d_vl_nbr);
      c->d_vl_nbr = nullptr;
      c->bytes -= (long long)c->vl_cap * np * (long long)sizeof(int);
    }
    if (int e = dalloc(c, (void**)&c->d_vl_nbr, (size_t)maxn * np * sizeof(int))) return e;
    c->vl_cap = maxn;
  }
  if (c->d_vl_home) return MSG_OK;  // the per-particle buffers exist already
  if (int e = dalloc(c, (void**)&c->d_vl_home, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_n, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_pos, 4 * (size_t)np * sizeof(double))) return e;
  if (int e = dalloc(c, (void**)&c->d_vl_slot, np * sizeof(int))) return e;
  if (int e = dalloc(c, (void**)&c->d_ro_ref, 3 * (size_t)np * sizeof(double))) return e;
  const size_t dbytes = (size_t)kDmaxSlots * kDmaxStride * sizeof(unsigned int);
  if (int e = dalloc(c, (void**)&c->d_dmax, dbytes)) return e;
  if (hipHostMalloc((void**)&c->h_dmax, dbytes, hipHostMallocDefault) != hipSuccess)
    return set_error(MSG_ERR_NOMEM, "hipHostMalloc failed");
  MSG_HIP(hipMemset(c->d_dmax, 0, dbytes));
  return MSG_OK;
}

int launch_vl_records(Ctx* c) {
  hipLaunchKernelGGL(k_vl_records, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, c->np, c->d_vl_home,
                     c->d_ro, c->d_ipt, c->d_laz, (double4*)c->d_vl_pos);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int launch_disp(Ctx* c) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = c->P.boxlen[d];
    b.half[d] = 0.5 * c->P.boxlen[d];
    b.inv[d] = 1.0 / c->P.boxlen[d];
  }
  hipLaunchKernelGGL(k_disp, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, c->np, c->d_ro,
                     c->d_ro_ref, b, c->d_dmax);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

// The skin test in two halves, so that work that does not depend on its outcome can be queued
// in between: the first queues the copy of the displacement maximum to the host (pending =
// false when no list exists yet: a build is due anyway); the second waits for the copy alone
// and decides.
int verlet_check_begin(Ctx* c, bool& pending) {
  pending = false;
  if (!c->vl_valid) return MSG_OK;
  if (!c->disp_done)
    if (int e = launch_disp(c)) return e;
  c->disp_done = false;
  MSG_HIP(hipMemcpyAsync(c->h_dmax, c->d_dmax, (size_t)kDmaxSlots * kDmaxStride * sizeof(unsigned int),
                         hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipEventRecord(c->ev_dmax, c->stream));
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
  MSG_HIP(hipEventSynchronize(c->ev_dmax));
  unsigned int bits = 0;
  for (int k = 0; k < kDmaxSlots; ++k) bits = std::max(bits, c->h_dmax[k * kDmaxStride]);
  float d2;
  std::memcpy(&d2, &bits, sizeof(float));
  rebuild = 2.0 * std::sqrt((double)d2) > c->P.drnlist;
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
  // initial accelerations from the forces of the current positions
  if (int e = utotal_refresh(c)) return e;
  MdArgs a = md_args(c, 0.0);
  hipLaunchKernelGGL(k_kick, dim3((np + 255) / 256), dim3(256), 0, c->stream, a, 1);
  MSG_HIP(hipStreamSynchronize(c->stream));
  c->md_ready = true;
  return MSG_OK;
}

int msg_md_run(msg_ctx* h, int nsteps, double dt, double* log) {
  if (!h || nsteps < 0) return set_error(MSG_ERR_ARG, "msg_md_run: bad arguments");
  Ctx* c = &h->c;
  if (!c->md_ready) return set_error(MSG_ERR_STATE, "msg_md_init has not been called");
  double* d_log = nullptr;
  if (log && nsteps > 0) MSG_HIP(hipMalloc(&d_log, 2 * (size_t)nsteps * sizeof(double)));
  const int nb = (c->np + 255) / 256;
  for (int s = 0; s < nsteps; ++s) {
    MdArgs a = md_args(c, dt);
    hipLaunchKernelGGL(k_drift, dim3(nb), dim3(256), 0, c->stream, a);
    c->disp_done = a.ref != nullptr;
    c->fresh = false;
    int e = utotal_sequence(c);
    if (e) { if (d_log) hipFree(d_log); return e; }
    a = md_args(c, dt);
    hipLaunchKernelGGL(k_kick, dim3(nb), dim3(256), 0, c->stream, a, 0);
    if (d_log)
      hipLaunchKernelGGL(k_mdlog, dim3(1), dim3(256), 0, c->stream, c->d_ke_part, nb, c->d_out,
                         c->nptpt, d_log + 2 * (size_t)s);
  }
  c->fresh = true;
  MSG_HIP(hipMemcpyAsync(c->h_out, c->d_out, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (d_log) MSG_HIP(hipMemcpyAsync(log, d_log, 2 * (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MSG_HIP(hipStreamSynchronize(c->stream));
  if (d_log) hipFree(d_log);
  const int f = (int)c->h_out[c->nptpt + 7];
  if (f & FLAG_NLIST) return set_error(MSG_ERR_NLIST, "Verlet list overflow");
  if (f & FLAG_R2UMIN) return set_error(MSG_ERR_R2UMIN, "pair below the table: r2 < r2umin");
  if (f & FLAG_CELL) return set_error(MSG_ERR_CELL, "position outside the box");
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
