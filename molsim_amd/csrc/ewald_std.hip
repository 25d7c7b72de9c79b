// K9/K10: standard Ewald reciprocal space in FP64 (UEwaldRecStd, Src/energy.F90:964-1102,
// with the k-factors of EwaldSetup, :6736-6760, and the surface term of UEwaldSurf,
// :1349-1371).
//
// k-vectors are MOLSIM's octant (nx, ny, nz >= 0, n^2 <= ncut^2 for 'sphere' or 3 ncut^2 for
// 'cube', origin excluded) with the four sign combinations (-,-,+), (-,+,+), (+,-,+),
// (+,+,+) of (kx, ky) per octant vector, and kfac carrying the 1/2 weights of zero indices
// and the factor 2 of nz > 0.
//   k_ewstd_sk     one workgroup per octant vector: the 4 structure factors
//                  S = sum_a q_a exp(i k.r_a) (direct sincos in FP64), kfac sum |S|^2 and the
//                  virial weight; S stays resident for MC energy differences.
//   k_ewstd_force  one wavefront per particle: F_a += 2 EpsiFourPi kfac k Im(e_a conj(S)) over
//                  all vectors (the reference's facmm/facmp/facpm/facpp form), plus the
//                  surface force; adds to the pair force and applies the force unit scale.
#include "common.h"

namespace msg {

namespace {

struct EwArgs {
  int np, nk, ncut;
  const double* ro;
  const double* az;
  const int* kvec;     // nk * 3
  const double* kfac;  // nk
  double tpb[3];       // 2 pi / L
  double inv2a2;       // 1 / (2 alpha^2)
  double* sk;          // nk * 8: 4 complex structure factors
  double* part;        // nk * 2: kfac sum|S|^2, (1 - k^2/(2 alpha^2)) * that
};

__global__ __launch_bounds__(256) void k_ewstd_sk(EwArgs a) {
  __shared__ double sh[8];
  const int kn = blockIdx.x;
  const double kx = a.kvec[3 * kn] * a.tpb[0], ky = a.kvec[3 * kn + 1] * a.tpb[1],
               kz = a.kvec[3 * kn + 2] * a.tpb[2];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < a.np; i += 256) {
    const double q = a.az[i];
    const double x = a.ro[3 * i], y = a.ro[3 * i + 1], z = a.ro[3 * i + 2];
    double sx, cx, sy, cy, sz, cz;
    sincos(kx * x, &sx, &cx);
    sincos(ky * y, &sy, &cy);
    sincos(kz * z, &sz, &cz);
    // exp(i(sgx kx x + sgy ky y + kz z)) for (sgx, sgy) = (-,-), (-,+), (+,-), (+,+)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double gx = (c & 2) ? sx : -sx, gy = (c & 1) ? sy : -sy;
      // (cx + i gx)(cy + i gy)(cz + i sz)
      const double re1 = cx * cy - gx * gy, im1 = cx * gy + gx * cy;
      s[2 * c] += q * (re1 * cz - im1 * sz);
      s[2 * c + 1] += q * (re1 * sz + im1 * cz);
    }
  }
  double tot[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) tot[k] = msgd::block_sum(s[k], sh);
  if (threadIdx.x == 0) {
    double s2 = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a.sk[8 * kn + k] = tot[k];
      s2 += tot[k] * tot[k];
    }
    const double term = a.kfac[kn] * s2;
    const double k2 = kx * kx + ky * ky + kz * kz;
    a.part[2 * kn] = term;
    a.part[2 * kn + 1] = (1.0 - k2 * a.inv2a2) * term;
  }
}

struct EwForceArgs {
  EwArgs e;
  double epsi4pi;
  double surf2fac;
  const double* Mtot;
  double* force;
  double scale;
};

constexpr int kEwNmax = 32;  // phase-table entries per axis: ncut <= 31

// One wavefront per particle; its lanes take the k-vectors lane, lane + 64, ... and the
// force components are wave-summed.  The phase factors e^{i n (2 pi / L_d) x_d},
// n = 0..ncut, are computed once per particle (one sincos per lane) into an LDS table from
// which every k-vector reads its three factors.
__global__ __launch_bounds__(256) void k_ewstd_force(EwForceArgs f) {
  const EwArgs& a = f.e;
  __shared__ double2 s_e[4][3][kEwNmax];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + w;
  if (i >= a.np) return;  // the whole wavefront leaves; no workgroup barrier follows
  const double q = a.az[i];
  const double x = a.ro[3 * i], y = a.ro[3 * i + 1], z = a.ro[3 * i + 2];
  const int nc1 = a.ncut + 1;
  for (int t = lane; t < 3 * nc1; t += 64) {
    const int d = t / nc1, n = t - d * nc1;
    const double k = n * a.tpb[d];
    double sv, cv;
    sincos(k * (d == 0 ? x : (d == 1 ? y : z)), &sv, &cv);
    s_e[w][d][n] = make_double2(cv, sv);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int kn = lane; kn < a.nk; kn += 64) {
    const int nx = a.kvec[3 * kn], ny = a.kvec[3 * kn + 1], nz = a.kvec[3 * kn + 2];
    const double kx = nx * a.tpb[0], ky = ny * a.tpb[1], kz = nz * a.tpb[2];
    const double2 ex = s_e[w][0][nx], ey = s_e[w][1][ny], ez = s_e[w][2][nz];
    const double cx = ex.x, sx = ex.y, cy = ey.x, sy = ey.y, cz = ez.x, sz = ez.y;
    double fc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double gx = (c & 2) ? sx : -sx, gy = (c & 1) ? sy : -sy;
      const double re1 = cx * cy - gx * gy, im1 = cx * gy + gx * cy;
      const double er = q * (re1 * cz - im1 * sz), ei = q * (re1 * sz + im1 * cz);
      const double Sr = a.sk[8 * kn + 2 * c], Si = a.sk[8 * kn + 2 * c + 1];
      fc[c] = -er * Si + ei * Sr;
    }
    const double kf2 = f.epsi4pi * 2.0 * a.kfac[kn];
    fx += kf2 * kx * (-fc[0] - fc[1] + fc[2] + fc[3]);
    fy += kf2 * ky * (-fc[0] + fc[1] - fc[2] + fc[3]);
    fz += kf2 * kz * (fc[0] + fc[1] + fc[2] + fc[3]);
  }
  fx = msgd::wave_sum(fx);
  fy = msgd::wave_sum(fy);
  fz = msgd::wave_sum(fz);
  if (lane < 3) {
    const double fr = lane == 0 ? fx : (lane == 1 ? fy : fz);
    const double M = f.Mtot[lane];
    const double fs = -f.epsi4pi * f.surf2fac * q * M;
    f.force[3 * i + lane] = (f.force[3 * i + lane] + fr + fs) * f.scale;
  }
}

EwArgs make_args(Ctx* c) {
  EwArgs a;
  a.np = c->np;
  a.nk = c->nkvec;
  a.ncut = c->P.ncut;
  a.ro = c->d_ro;
  a.az = c->d_az;
  a.kvec = c->d_kvec;
  a.kfac = c->d_kfac;
  for (int d = 0; d < 3; ++d) a.tpb[d] = 2.0 * M_PI / c->P.boxlen[d];
  a.inv2a2 = 1.0 / (2.0 * c->P.alpha * c->P.alpha);
  a.sk = c->d_sk;
  a.part = c->d_part + c->part_off_ewstd;
  return a;
}

}  // namespace

int ewstd_setup(Ctx* c) {
  const msg_params& P = c->P;
  if (P.ncut < 1) return set_error(MSG_ERR_ARG, "std Ewald: ncut must be >= 1");
  if (P.ncut >= kEwNmax) return set_error(MSG_ERR_ARG, "std Ewald: ncut must be < 32");
  const int ncut = P.ncut, ncut2 = P.ncutregion == 1 ? 3 * ncut * ncut : ncut * ncut;
  const double vol = P.boxlen[0] * P.boxlen[1] * P.boxlen[2];
  const double fac0 = 2.0 * M_PI / vol;
  std::vector<int> kv;
  std::vector<double> kf;
  for (int nz = 0; nz <= ncut; ++nz) {
    const double fz = nz == 0 ? 1.0 : 2.0;
    for (int ny = 0; ny <= ncut; ++ny) {
      const double fy = ny == 0 ? 0.5 : 1.0;
      for (int nx = 0; nx <= ncut; ++nx) {
        if (nx * nx + ny * ny + nz * nz > ncut2) continue;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        const double fx = nx == 0 ? 0.5 : 1.0;
        const double k2 = std::pow(nx * 2.0 * M_PI / P.boxlen[0], 2) +
                          std::pow(ny * 2.0 * M_PI / P.boxlen[1], 2) +
                          std::pow(nz * 2.0 * M_PI / P.boxlen[2], 2);
        kv.push_back(nx);
        kv.push_back(ny);
        kv.push_back(nz);
        kf.push_back(fac0 * fz * fy * fx * std::exp(-k2 / (4.0 * P.alpha * P.alpha)) / k2);
      }
    }
  }
  const int nk = (int)kf.size();
  if (nk > 4096) return set_error(MSG_ERR_ARG, "std Ewald: more than 4096 octant k-vectors");
  if (!c->d_kfac) {
    if (int e = dalloc(c, (void**)&c->d_kfac, 4096 * sizeof(double))) return e;
    if (int e = dalloc(c, (void**)&c->d_kvec, 4096 * 3 * sizeof(int))) return e;
    if (int e = dalloc(c, (void**)&c->d_sk, 4096 * 8 * sizeof(double))) return e;
  }
  MSG_HIP(hipMemcpy(c->d_kfac, kf.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  MSG_HIP(hipMemcpy(c->d_kvec, kv.data(), nk * 3 * sizeof(int), hipMemcpyHostToDevice));
  c->nkvec = nk;
  c->npart_ewstd = nk;
  return MSG_OK;
}

int launch_ewstd(Ctx* c) {
  EwForceArgs f;
  f.e = make_args(c);
  f.epsi4pi = c->P.epsi4pi;
  const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
  f.surf2fac = c->P.lsurf ? 2.0 * (2.0 * M_PI / (3.0 * vol)) : 0.0;
  f.Mtot = c->d_M;
  f.force = c->d_force;
  f.scale = c->P.force_scale;
  prof_begin(c, MSG_K_EWSTD);
  hipLaunchKernelGGL(k_ewstd_sk, dim3(c->nkvec), dim3(256), 0, c->stream, f.e);
  hipLaunchKernelGGL(k_ewstd_force, dim3((c->np + 3) / 4), dim3(256), 0, c->stream, f);
  prof_end(c, MSG_K_EWSTD);
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg
