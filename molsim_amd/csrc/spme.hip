// K5-K8: smooth particle-mesh Ewald reciprocal space (UEwaldRecSPM + SPMFFTRec,
// Src/energy.F90:1207-1297,6525-6587; influence function of EwaldSetup, :6969-7018).
//
//   k_spread    one workgroup per mesh brick (16x8x8 points).  Particles were binned by the
//               brick of their B-spline base point (meshmax); a brick's contributors are the
//               particles of its own and the next brick in each dimension.  Each wavefront
//               accumulates whole particles (one lane per (iz,iy) stencil row) into its own
//               LDS copy of the brick, sequentially, so there are no atomics at all and the
//               result is deterministic; the copies are summed in fixed order and every mesh
//               point is written exactly once.
//   k_fft_x     r2c along x: two real rows packed as one complex line, Stockham radix-4/2
//               in LDS, split into the two half spectra.
//   k_fft_y     complex FFT along y on tiles of T consecutive kx (forward or inverse).
//   k_fft_zconv forward FFT along z, influence-function multiply with the FP64 energy and
//               virial reduction (the reference's CalcGIF loop), inverse FFT along z.
//   k_fft_xi    c2r along x (half spectra of two rows -> one complex line -> two real rows).
//   k_gather    per particle (brick order): field from the convolved mesh, force
//               F -= q * (sum dtheta theta theta phi) * dkdr * EpsiFourPi plus the surface
//               term -EpsiFourPi * 2 fac * q * M; adds to the pair force and applies the
//               force unit scale.
// FFT conventions are FFTW's: forward e^{-i}, backward e^{+i}, no normalisation.
#include "common.h"

namespace msg {

namespace {

template <typename real> struct C2T;
template <> struct C2T<float> { typedef float2 t; };
template <> struct C2T<double> { typedef double2 t; };

template <typename C2>
__device__ inline C2 cmul(C2 a, C2 b) {
  C2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C2>
__device__ inline C2 cadd(C2 a, C2 b) { C2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C2>
__device__ inline C2 csub(C2 a, C2 b) { C2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C2>
__device__ inline C2 cconj(C2 a) { a.y = -a.y; return a; }
template <typename C2>
__device__ inline C2 cmk(decltype(C2().x) x, decltype(C2().x) y) { C2 r; r.x = x; r.y = y; return r; }

// Elements of one workgroup's LDS line buffer and its thread count.
template <typename real> struct FftCfg;
template <> struct FftCfg<float> { static constexpr int E = 2048, NT = 256; };
template <> struct FftCfg<double> { static constexpr int E = 1024, NT = 256; };

// Stockham autosort FFT of T lines of length K held line-major in LDS buffer A (B is
// scratch).  Radix-4 stages, then one radix-2 stage when log2 K is odd.  Returns the
// buffer holding the result.  All NT threads of the workgroup must call it.
template <typename C2, int K, int T>
__device__ C2* lds_fft(C2* A, C2* B, const C2* __restrict__ tw, bool inverse) {
  constexpr int NT = 256;
  constexpr int TPL = NT / T;  // threads per line
  const int line = threadIdx.x / TPL, tl = threadIdx.x % TPL;
  C2* src = A;
  C2* dst = B;
  int Ns = 1;
  for (; Ns * 4 <= K; Ns *= 4) {
    const int step = K / (Ns * 4);
    for (int j = tl; j < K / 4; j += TPL) {
      const int jr = j % Ns;
      const C2* s = src + line * K;
      C2 v0 = s[j], v1 = s[j + K / 4], v2 = s[j + K / 2], v3 = s[j + 3 * K / 4];
      C2 w1 = tw[jr * step], w2 = tw[2 * jr * step], w3 = tw[3 * jr * step];
      if (inverse) { w1 = cconj(w1); w2 = cconj(w2); w3 = cconj(w3); }
      v1 = cmul(v1, w1);
      v2 = cmul(v2, w2);
      v3 = cmul(v3, w3);
      const C2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
      const C2 d13 = csub(v1, v3);
      // forward: a3 = -i (v1 - v3); inverse: +i (v1 - v3)
      const C2 a3 = inverse ? cmk<C2>(-d13.y, d13.x) : cmk<C2>(d13.y, -d13.x);
      C2* d = dst + line * K + (j / Ns) * Ns * 4 + jr;
      d[0] = cadd(a0, a2);
      d[Ns] = cadd(a1, a3);
      d[2 * Ns] = csub(a0, a2);
      d[3 * Ns] = csub(a1, a3);
    }
    __syncthreads();
    C2* t = src; src = dst; dst = t;
  }
  if (Ns < K) {
    const int step = K / (Ns * 2);
    for (int j = tl; j < K / 2; j += TPL) {
      const int jr = j % Ns;
      const C2* s = src + line * K;
      C2 w = tw[jr * step];
      if (inverse) w = cconj(w);
      const C2 v0 = s[j], v1 = cmul(s[j + K / 2], w);
      C2* d = dst + line * K + (j / Ns) * Ns * 2 + jr;
      d[0] = cadd(v0, v1);
      d[Ns] = csub(v0, v1);
    }
    __syncthreads();
    C2* t = src; src = dst; dst = t;
  }
  return src;
}

// r2c along x.  Workgroup: T line pairs (2T real rows); row = y + K*z.
template <typename real, int K>
__global__ __launch_bounds__(256) void k_fft_x(const real* __restrict__ Q,
                                               typename C2T<real>::t* __restrict__ Cm,
                                               const typename C2T<real>::t* __restrict__ tw) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E, T = E / K, TPL = 256 / T, H = K / 2 + 1;
  __shared__ C2 A[E], B[E];
  const int line = threadIdx.x / TPL, tl = threadIdx.x % TPL;
  const size_t r0 = 2 * ((size_t)blockIdx.x * T + line), r1 = r0 + 1;
  for (int k = tl; k < K; k += TPL) A[line * K + k] = cmk<C2>(Q[r0 * K + k], Q[r1 * K + k]);
  __syncthreads();
  const C2* Z = lds_fft<C2, K, T>(A, B, tw, false) + line * K;
  const real h = real(0.5);
  for (int k = tl; k < H; k += TPL) {
    const C2 z = Z[k], zc = cconj(Z[(K - k) & (K - 1)]);
    const C2 s = cadd(z, zc), d = csub(z, zc);
    Cm[r0 * H + k] = cmk<C2>(h * s.x, h * s.y);
    Cm[r1 * H + k] = cmk<C2>(h * d.y, -h * d.x);  // (z - zc) / (2i)
  }
}

// c2r along x (the imaginary parts of the kx = 0 and K/2 terms are ignored, as in c2r).
template <typename real, int K>
__global__ __launch_bounds__(256) void k_fft_xi(const typename C2T<real>::t* __restrict__ Cm,
                                                real* __restrict__ Q,
                                                const typename C2T<real>::t* __restrict__ tw) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E, T = E / K, TPL = 256 / T, H = K / 2 + 1;
  __shared__ C2 A[E], B[E];
  const int line = threadIdx.x / TPL, tl = threadIdx.x % TPL;
  const size_t r0 = 2 * ((size_t)blockIdx.x * T + line), r1 = r0 + 1;
  for (int k = tl; k < H; k += TPL) {
    C2 a = Cm[r0 * H + k], b = Cm[r1 * H + k];
    if (k == 0 || k == K / 2) { a.y = 0; b.y = 0; }
    A[line * K + k] = cmk<C2>(a.x - b.y, a.y + b.x);                 // a + i b
    if (k > 0 && k < K / 2) A[line * K + K - k] = cmk<C2>(a.x + b.y, b.x - a.y);  // conj(a) + i conj(b)
  }
  __syncthreads();
  const C2* Z = lds_fft<C2, K, T>(A, B, tw, true) + line * K;
  for (int k = tl; k < K; k += TPL) {
    const C2 z = Z[k];
    Q[r0 * K + k] = z.x;
    Q[r1 * K + k] = z.y;
  }
}

// Complex FFT along y (stride H) for tiles of T consecutive kx at one z.
template <typename real, int K>
__global__ __launch_bounds__(256) void k_fft_y(typename C2T<real>::t* __restrict__ Cm,
                                               const typename C2T<real>::t* __restrict__ tw,
                                               int inverse) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E, T = E / K, H = K / 2 + 1, NTX = (H + T - 1) / T;
  __shared__ C2 A[E], B[E];
  const int tx = blockIdx.x % NTX, z = blockIdx.x / NTX, kx0 = tx * T;
  for (int e = threadIdx.x; e < E; e += 256) {
    const int kl = e % T, y = e / T, kx = kx0 + kl;
    A[kl * K + y] = kx < H ? Cm[kx + (size_t)H * (y + (size_t)K * z)] : cmk<C2>(0, 0);
  }
  __syncthreads();
  const C2* R = lds_fft<C2, K, T>(A, B, tw, inverse != 0);
  for (int e = threadIdx.x; e < E; e += 256) {
    const int kl = e % T, y = e / T, kx = kx0 + kl;
    if (kx < H) Cm[kx + (size_t)H * (y + (size_t)K * z)] = R[kl * K + y];
  }
}

struct ConvArgs {
  double tpl2[3];   // (2 pi / L)^2
  double expf;      // -1 / (4 alpha^2)
};

// Forward FFT along z, influence-function multiply + FP64 energy/virial, inverse FFT along z.
template <typename real, int K>
__global__ __launch_bounds__(256) void k_fft_zconv(typename C2T<real>::t* __restrict__ Cm,
                                                   const typename C2T<real>::t* __restrict__ tw,
                                                   const real* __restrict__ G, ConvArgs ca,
                                                   double* __restrict__ part) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E, T = E / K, H = K / 2 + 1, NTX = (H + T - 1) / T;
  __shared__ C2 A[E], B[E];
  __shared__ double red[8];
  const int tx = blockIdx.x % NTX, y = blockIdx.x / NTX, kx0 = tx * T;
  const size_t plane = (size_t)H * K;
  for (int e = threadIdx.x; e < E; e += 256) {
    const int kl = e % T, z = e / T, kx = kx0 + kl;
    A[kl * K + z] = kx < H ? Cm[kx + (size_t)H * y + plane * z] : cmk<C2>(0, 0);
  }
  __syncthreads();
  C2* R = lds_fft<C2, K, T>(A, B, tw, false);
  const int ky = y <= K / 2 ? y : K - y;
  double esum = 0.0, vsum = 0.0;
  for (int e = threadIdx.x; e < E; e += 256) {
    const int kl = e / K, kz = e % K, kx = kx0 + kl;
    if (kx >= H) continue;
    const int kzz = kz <= K / 2 ? kz : K - kz;
    const real g = G[kx + H * (ky + H * kzz)];
    C2 f = R[e];
    const double f2 = (double)f.x * (double)f.x + (double)f.y * (double)f.y;
    const double w = (kx == 0 || kx == K / 2) ? 0.5 : 1.0;
    const double m2 = ca.tpl2[0] * kx * kx + ca.tpl2[1] * ky * ky + ca.tpl2[2] * kzz * kzz;
    const double eg = w * (double)g * f2;
    esum += eg;
    vsum += eg * (1.0 + 2.0 * m2 * ca.expf);
    f.x *= g;
    f.y *= g;
    R[e] = f;
  }
  __syncthreads();
  C2* other = (R == A) ? B : A;
  const C2* S = lds_fft<C2, K, T>(R, other, tw, true);
  for (int e = threadIdx.x; e < E; e += 256) {
    const int kl = e % T, z = e / T, kx = kx0 + kl;
    if (kx < H) Cm[kx + (size_t)H * y + plane * z] = S[kl * K + z];
  }
  esum = msgd::block_sum(esum, red);
  vsum = msgd::block_sum(vsum, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = esum;
    part[2 * blockIdx.x + 1] = vsum;
  }
}

struct SpreadArgs {
  MeshGrid m;
  const int* brick_start;
  const int* brick_ip;
  const int* mm;
  const void* frac;
  const double* az;
  void* Q;
};

// One workgroup per brick; W wavefronts each own an LDS copy of the brick.
template <typename real, int P>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a) {
  constexpr int W = 4, NBP = kBX * kBY * kBZ, CAP = sizeof(real) == 4 ? 256 : 128;
  constexpr int NACC = sizeof(real) == 4 ? W : 2;  // FP64: two copies (waves pair up, see below)
  __shared__ real acc[NACC][NBP];
  __shared__ real th[CAP][3 * P];
  __shared__ int loc[CAP];
  __shared__ int r_beg[8], r_pref[9];
  const int K = a.m.K;
  const int b = blockIdx.x;
  const int bx = b % a.m.nb[0], by = (b / a.m.nb[0]) % a.m.nb[1], bz = b / (a.m.nb[0] * a.m.nb[1]);
  const int lo[3] = {bx * kBX, by * kBY, bz * kBZ};
  if (threadIdx.x == 0) {
    const int ox = a.m.nb[0] > 1 ? 2 : 1, oy = a.m.nb[1] > 1 ? 2 : 1, oz = a.m.nb[2] > 1 ? 2 : 1;
    int n = 0, pref = 0;
    for (int iz = 0; iz < oz; ++iz)
      for (int iy = 0; iy < oy; ++iy)
        for (int ix = 0; ix < ox; ++ix) {
          const int mb = (bx + ix) % a.m.nb[0] +
                         a.m.nb[0] * ((by + iy) % a.m.nb[1] + a.m.nb[1] * ((bz + iz) % a.m.nb[2]));
          r_beg[n] = a.brick_start[mb];
          r_pref[n] = pref;
          pref += a.brick_start[mb + 1] - a.brick_start[mb];
          ++n;
        }
    for (; n < 9; ++n) r_pref[n] = pref;
  }
  for (int k = threadIdx.x; k < NACC * NBP; k += 256) (&acc[0][0])[k] = real(0);
  __syncthreads();
  const int total = r_pref[8];
  const real* frac = (const real*)a.frac;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // FP64: waves 0/1 own copies 0/1 and waves 2/3 take turns with them (see phase 2)
  for (int base = 0; base < total; base += CAP) {
    const int cnt = min(CAP, total - base);
    if ((int)threadIdx.x < cnt) {
      const int t = base + threadIdx.x;
      int r = 0;
      while (r < 7 && t >= r_pref[r + 1]) ++r;
      const int ip = a.brick_ip[r_beg[r] + (t - r_pref[r])];
      const int packed = a.mm[ip];
      const int mmv[3] = {packed & 1023, (packed >> 10) & 1023, packed >> 20};
      real* T = th[threadIdx.x];
#pragma unroll
      for (int d = 0; d < 3; ++d) msgd::bspline<real, P>(frac[3 * ip + d], T + d * P, (real*)nullptr);
      const real q = (real)a.az[ip];
#pragma unroll
      for (int i = 0; i < P; ++i) T[2 * P + i] *= q;
      int l[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int v = mmv[d] - lo[d];
        l[d] = v < 0 ? v + K : v;
      }
      loc[threadIdx.x] = l[0] | (l[1] << 10) | (l[2] << 20);
    }
    __syncthreads();
    // phase 2: each wave takes every W-th particle in order; lane = (iz, iy) row
    const int iz = lane / P, iy = lane % P;
    for (int pass = 0; pass < (NACC == W ? 1 : 2); ++pass) {
      const bool mine = NACC == W ? true : ((wv >> 1) == pass);
      if (mine && lane < P * P) {
        real* A = acc[NACC == W ? wv : (wv & 1)];
        for (int p = wv; p < cnt; p += W) {
          const int lp = loc[p];
          int ly = ((lp >> 10) & 1023) - iy, lz = (lp >> 20) - iz;
          ly += ly < 0 ? K : 0;
          lz += lz < 0 ? K : 0;
          if (ly >= kBY || lz >= kBZ) continue;
          const real wyz = th[p][P + iy] * th[p][2 * P + iz];
          const int lx0 = lp & 1023;
          real* row = A + kBX * (ly + kBY * lz);
#pragma unroll
          for (int ix = 0; ix < P; ++ix) {
            int lx = lx0 - ix;
            lx += lx < 0 ? K : 0;
            if (lx < kBX) row[lx] += th[p][ix] * wyz;
          }
        }
      }
      if (NACC != W) __syncthreads();
    }
    __syncthreads();
  }
  real* Q = (real*)a.Q;
  for (int k = threadIdx.x; k < NBP; k += 256) {
    real s = acc[0][k];
#pragma unroll
    for (int w = 1; w < NACC; ++w) s += acc[w][k];
    const int lx = k % kBX, ly = (k / kBX) % kBY, lz = k / (kBX * kBY);
    Q[(lo[0] + lx) + (size_t)K * ((lo[1] + ly) + (size_t)K * (lo[2] + lz))] = s;
  }
}

struct GatherArgs {
  MeshGrid m;
  int np;
  const int* brick_ip;
  const int* mm;
  const void* frac;
  const double* az;
  const void* phi;
  double epsi4pi;
  double surf2fac;          // 2 * (2 pi / (3 V)) when lsurf, else 0
  const double* dip_part;   // per-block partial dipoles (sum q r)
  int ndip;
  double* force;            // 3*np caller order (pair force already there)
  double scale;
};

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
  __shared__ double M[3];
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int k = 0; k < a.ndip; ++k) s += a.dip_part[3 * k + threadIdx.x];
    M[threadIdx.x] = s;
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.np) return;
  const int K = a.m.K;
  const int ip = a.brick_ip[i];
  const int packed = a.mm[ip];
  const int mmv[3] = {packed & 1023, (packed >> 10) & 1023, packed >> 20};
  const real* frac = (const real*)a.frac;
  real th[3][P], dth[3][P];
#pragma unroll
  for (int d = 0; d < 3; ++d) msgd::bspline<real, P>(frac[3 * ip + d], th[d], dth[d]);
  const real* phi = (const real*)a.phi;
  int xs[P];
#pragma unroll
  for (int ix = 0; ix < P; ++ix) {
    int x = mmv[0] - ix;
    xs[ix] = x < 0 ? x + K : x;
  }
  real e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
  for (int iz = 0; iz < P; ++iz) {
    int z = mmv[2] - iz;
    z += z < 0 ? K : 0;
#pragma unroll
    for (int iy = 0; iy < P; ++iy) {
      int y = mmv[1] - iy;
      y += y < 0 ? K : 0;
      const real* row = phi + (size_t)K * (y + (size_t)K * z);
      real sx = 0, dsx = 0;
#pragma unroll
      for (int ix = 0; ix < P; ++ix) {
        const real v = row[xs[ix]];
        sx += th[0][ix] * v;
        dsx += dth[0][ix] * v;
      }
      e0 += dsx * th[1][iy] * th[2][iz];
      e1 += sx * dth[1][iy] * th[2][iz];
      e2 += sx * th[1][iy] * dth[2][iz];
    }
  }
  const double q = a.az[ip];
  const double es[3] = {(double)e0, (double)e1, (double)e2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double f = -q * es[d] * a.m.dkdr[d] * a.epsi4pi - a.epsi4pi * a.surf2fac * q * M[d];
    a.force[3 * ip + d] = (a.force[3 * ip + d] + f) * a.scale;
  }
}

// ------------------------------------------------------------------------------------------
template <typename real, int K>
void launch_fft_all(Ctx* c) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E, T = E / K, H = K / 2 + 1, NTX = (H + T - 1) / T;
  C2* Cm = (C2*)c->d_C;
  const C2* tw = (const C2*)c->d_tw;
  real* Q = (real*)c->d_Q;
  const int nx = K * K / 2 / T;
  prof_begin(c, MSG_K_FFT_X);
  hipLaunchKernelGGL((k_fft_x<real, K>), dim3(nx), dim3(256), 0, c->stream, Q, Cm, tw);
  prof_end(c, MSG_K_FFT_X);
  prof_begin(c, MSG_K_FFT_Y);
  hipLaunchKernelGGL((k_fft_y<real, K>), dim3(NTX * K), dim3(256), 0, c->stream, Cm, tw, 0);
  prof_end(c, MSG_K_FFT_Y);
  ConvArgs ca;
  for (int d = 0; d < 3; ++d) ca.tpl2[d] = std::pow(2.0 * M_PI / c->P.boxlen[d], 2);
  ca.expf = -1.0 / (4.0 * c->P.alpha * c->P.alpha);
  prof_begin(c, MSG_K_FFT_ZCONV);
  hipLaunchKernelGGL((k_fft_zconv<real, K>), dim3(NTX * K), dim3(256), 0, c->stream, Cm, tw,
                     (const real*)c->d_G, ca, c->d_part + c->part_off_spme);
  prof_end(c, MSG_K_FFT_ZCONV);
  prof_begin(c, MSG_K_FFT_YI);
  hipLaunchKernelGGL((k_fft_y<real, K>), dim3(NTX * K), dim3(256), 0, c->stream, Cm, tw, 1);
  prof_end(c, MSG_K_FFT_YI);
  prof_begin(c, MSG_K_FFT_XI);
  hipLaunchKernelGGL((k_fft_xi<real, K>), dim3(nx), dim3(256), 0, c->stream, (const C2*)Cm, Q, tw);
  prof_end(c, MSG_K_FFT_XI);
}

template <typename real>
int launch_fft_dispatch(Ctx* c) {
  switch (c->mesh.K) {
    case 16: launch_fft_all<real, 16>(c); break;
    case 32: launch_fft_all<real, 32>(c); break;
    case 64: launch_fft_all<real, 64>(c); break;
    case 128: launch_fft_all<real, 128>(c); break;
    case 256: launch_fft_all<real, 256>(c); break;
    default: return set_error(MSG_ERR_MESH, "nmesh must be 16, 32, 64, 128 or 256");
  }
  return MSG_OK;
}

template <typename real, int P>
void launch_spread_gather(Ctx* c, bool spread) {
  if (spread) {
    SpreadArgs a{c->mesh, c->d_brick_start, c->d_brick_ip, c->d_mm, c->d_frac, c->d_az, c->d_Q};
    prof_begin(c, MSG_K_SPREAD);
    hipLaunchKernelGGL((k_spread<real, P>), dim3(c->mesh.nbrick), dim3(256), 0, c->stream, a);
    prof_end(c, MSG_K_SPREAD);
  } else {
    GatherArgs g;
    g.m = c->mesh;
    g.np = c->np;
    g.brick_ip = c->d_brick_ip;
    g.mm = c->d_mm;
    g.frac = c->d_frac;
    g.az = c->d_az;
    g.phi = c->d_Q;
    g.epsi4pi = c->P.epsi4pi;
    const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
    g.surf2fac = c->P.lsurf ? 2.0 * (2.0 * M_PI / (3.0 * vol)) : 0.0;
    g.dip_part = c->d_part + c->part_off_dip;
    g.ndip = c->npart_dip;
    g.force = c->d_force;
    g.scale = c->P.force_scale;
    prof_begin(c, MSG_K_GATHER);
    hipLaunchKernelGGL((k_gather<real, P>), dim3((c->np + 255) / 256), dim3(256), 0, c->stream, g);
    prof_end(c, MSG_K_GATHER);
  }
}

template <typename real>
int spread_gather_dispatch(Ctx* c, bool spread) {
  switch (c->mesh.order) {
    case 3: launch_spread_gather<real, 3>(c, spread); break;
    case 4: launch_spread_gather<real, 4>(c, spread); break;
    case 5: launch_spread_gather<real, 5>(c, spread); break;
    case 6: launch_spread_gather<real, 6>(c, spread); break;
    case 7: launch_spread_gather<real, 7>(c, spread); break;
    case 8: launch_spread_gather<real, 8>(c, spread); break;
    default: return set_error(MSG_ERR_ARG, "SPME order must be 3..8");
  }
  return MSG_OK;
}

}  // namespace

// EwaldSetup, SPME branch (Src/energy.F90:6969-7018): B-spline moduli b^2(m) and the
// influence function meshfac over [0..K/2]^3, computed in FP64 on the host.
int spme_setup(Ctx* c) {
  const msg_params& P = c->P;
  const int K = P.nmesh, p = P.order;
  if (K % 2) return set_error(MSG_ERR_ODD_MESH, "nmesh must be even");
  if (K < 16 || K > 256 || (K & (K - 1))) return set_error(MSG_ERR_MESH, "nmesh must be 16, 32, 64, 128 or 256");
  if (p < 3 || p > 8) return set_error(MSG_ERR_ARG, "SPME order must be 3..8");
  MeshGrid& m = c->mesh;
  m.K = K;
  m.H = K / 2 + 1;
  m.order = p;
  m.nb[0] = K / kBX;
  m.nb[1] = K / kBY;
  m.nb[2] = K / kBZ;
  m.nbrick = m.nb[0] * m.nb[1] * m.nb[2];
  for (int d = 0; d < 3; ++d) m.dkdr[d] = K / P.boxlen[d];
  // spline at pos = 0: M_p(1..p-1) are the only non-zero integer-point values
  std::vector<double> spl(p, 0.0);
  {
    std::vector<double> last(p, 0.0);
    spl[0] = 0.0;
    spl[1] = 1.0;
    for (int n = 3; n <= p; ++n) {
      last = spl;
      spl[0] = 0.0;
      for (int i = 1; i <= n - 2; ++i) spl[i] = (i * last[i] + (n - i) * last[i - 1]) / (n - 1);
      spl[n - 1] = last[n - 2] / (n - 1);
    }
  }
  std::vector<double> b2(K);
  for (int mm = 0; mm < K; ++mm) {
    if ((p % 2 == 1) && (2 * mm == K)) { b2[mm] = 0.0; continue; }
    double re = 0, im = 0;
    for (int k = 0; k <= p - 2; ++k) {
      const double ang = 2.0 * M_PI * mm * k / K;
      re += spl[k + 1] * std::cos(ang);
      im += spl[k + 1] * std::sin(ang);
    }
    b2[mm] = 1.0 / (re * re + im * im);
  }
  const int H = m.H;
  const double vol = P.boxlen[0] * P.boxlen[1] * P.boxlen[2];
  const double fac0 = 4.0 * M_PI / vol, expf = -1.0 / (4.0 * P.alpha * P.alpha);
  std::vector<double> G((size_t)H * H * H);
  for (int nz = 0; nz < H; ++nz)
    for (int ny = 0; ny < H; ++ny)
      for (int nx = 0; nx < H; ++nx) {
        const size_t o = nx + (size_t)H * (ny + (size_t)H * nz);
        if (nx == 0 && ny == 0 && nz == 0) { G[o] = 0.0; continue; }
        const double m2 = std::pow(nx * 2.0 * M_PI / P.boxlen[0], 2) +
                          std::pow(ny * 2.0 * M_PI / P.boxlen[1], 2) +
                          std::pow(nz * 2.0 * M_PI / P.boxlen[2], 2);
        G[o] = fac0 * std::exp(m2 * expf) / m2 * b2[nx] * b2[ny] * b2[nz];
      }
  const size_t rs = c->fp64 ? 8 : 4;
  const int np = c->np;
  auto A = [&](void** ptr, size_t b) { return dalloc(c, ptr, b); };
  if (!c->d_Q) {
    if (int e = A(&c->d_Q, (size_t)K * K * K * rs)) return e;
    if (int e = A(&c->d_C, (size_t)H * K * K * 2 * rs)) return e;
    if (int e = A(&c->d_G, (size_t)H * H * H * rs)) return e;
    if (int e = A(&c->d_tw, (size_t)K * 2 * rs)) return e;
    if (int e = A((void**)&c->d_brick_count, m.nbrick * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_start, (m.nbrick + 1) * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_fill, m.nbrick * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_tmp, np * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_ip, np * sizeof(int))) return e;
    if (int e = A((void**)&c->d_mm, np * sizeof(int))) return e;
    if (int e = A(&c->d_frac, (size_t)np * 3 * rs)) return e;
  }
  std::vector<double> tw(2 * K);
  for (int i = 0; i < K; ++i) {
    tw[2 * i] = std::cos(2.0 * M_PI * i / K);
    tw[2 * i + 1] = -std::sin(2.0 * M_PI * i / K);
  }
  if (c->fp64) {
    MSG_HIP(hipMemcpy(c->d_G, G.data(), G.size() * 8, hipMemcpyHostToDevice));
    MSG_HIP(hipMemcpy(c->d_tw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> Gf(G.begin(), G.end()), twf(tw.begin(), tw.end());
    MSG_HIP(hipMemcpy(c->d_G, Gf.data(), Gf.size() * 4, hipMemcpyHostToDevice));
    MSG_HIP(hipMemcpy(c->d_tw, twf.data(), twf.size() * 4, hipMemcpyHostToDevice));
  }
  const int T = (c->fp64 ? FftCfg<double>::E : FftCfg<float>::E) / K;
  c->nzblocks = ((H + T - 1) / T) * K;
  return MSG_OK;
}

int launch_spme(Ctx* c) {
  int e;
  if (c->fp64) {
    if ((e = spread_gather_dispatch<double>(c, true))) return e;
    if ((e = launch_fft_dispatch<double>(c))) return e;
    if ((e = spread_gather_dispatch<double>(c, false))) return e;
  } else {
    if ((e = spread_gather_dispatch<float>(c, true))) return e;
    if ((e = launch_fft_dispatch<float>(c))) return e;
    if ((e = spread_gather_dispatch<float>(c, false))) return e;
  }
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg
