// K5-K8: smooth particle-mesh Ewald reciprocal space (UEwaldRecSPM + SPMFFTRec,
// Src/energy.F90:1207-1297,6525-6587; influence function of EwaldSetup, :6969-7018).
//
//   k_spread    one workgroup per mesh brick (32x8x8 points, 16x8x8 for K = 16).  Particles
//               were binned by the brick of their B-spline base point (meshmax); a brick's
//               contributors are the particles of its own and the next brick in each
//               dimension.  Contributions are accumulated in LDS as fixed-point integers
//               (32-bit with a per-brick scale in FP32, 64-bit in FP64: exact,
//               order-independent, hence deterministic; no global atomics) and every mesh
//               point is written once.
//   k_fft_x     r2c along x: two real rows packed as one complex line, a line transform held
//               in registers (below), split into the two half spectra.
//   k_fft_y     complex FFT along y on tiles of T lines (forward or inverse).
// Half-complex layout: kx = 0..K/2-1 as Cm[kx + (K/2)(y + K z)] (rows of K/2 values, so a
// tile of consecutive kx is an aligned segment), and the Nyquist column kx = K/2 as its own
// K x K plane Nq[y + K z] behind it.  The y and z passes index lines by L = kx + (K/2) * (z or
// y) for the K/2 regular columns, followed by the K Nyquist lines, so no workgroup computes
// padding lines.
//   k_fft_zconv forward FFT along z, influence-function multiply with the FP64 energy and
//               virial reduction (the reference's CalcGIF loop), inverse FFT along z.
//   k_fft_xi    c2r along x (half spectra of two rows -> one complex line -> two real rows).
//   k_gather    per particle (brick order): field from the convolved mesh, force
//               F -= q * (sum dtheta theta theta phi) * dkdr * EpsiFourPi plus the surface
//               term -EpsiFourPi * 2 fac * q * M; adds to the pair force and applies the
//               force unit scale.
// FFT conventions are FFTW's: forward e^{-i}, backward e^{+i}, no normalisation.
#include "common.h"
#include "reduce.h"

namespace msg {

namespace {

template <typename real> struct C2T;
template <> struct C2T<float> { typedef float2 t; };
template <> struct C2T<double> { typedef double2 t; };

// Complex product with its fused multiply-adds written out, so every kernel that inlines it
// (separate and fused FFT passes, single and batched launches) rounds it the same way.
template <typename C2>
__device__ inline C2 cmul(C2 a, C2 b) {
  C2 r;
  r.x = fma(a.x, b.x, -(a.y * b.y));
  r.y = fma(a.x, b.y, a.y * b.x);
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

// ------------------------------------------------------------------------------------------
// Power-of-two meshes: register-resident line transforms.  A line of K = 8 M points is held by
// M consecutive lanes, 8 points per lane: lane b holds x[M n1 + b], n1 = 0..7.  The transform
// is the K = 8 x M Cooley-Tukey split: a radix-8 DFT over n1 in registers (three radix-2
// decimation-in-frequency stages), the twiddles W_K^(b k1), then the M-point DFT over the lanes
// by log2 M radix-2 decimation-in-frequency stages between partner lanes (shuffles, no LDS).
// Afterwards register r of lane b holds X[br3(r) + 8 brM(b)] (br: bit reversal).
// W = e^(-2 pi i / K) forward, its conjugate for the inverse; unnormalised, as FFTW's.
constexpr int kLineElems = 2048;  // points per 256-thread workgroup of a separate pass
template <int K>
struct LineCfg {
  static constexpr int M = K / 8;  // lanes per line
  static constexpr int LB = M == 2 ? 1 : M == 4 ? 2 : M == 8 ? 3 : M == 16 ? 4 : 5;
  static constexpr int T = kLineElems / K;  // lines per 256-thread workgroup
  // LDS line stride: a point k of a line sits at lpos(k) = k + k / 8, so the eight points a lane
  // holds after the transform (stride 8) and the lanes' points before it fall into distinct
  // banks; the odd extra point staggers the lines
  static constexpr int LSX = K + K / 8 + 1;
  static_assert(K >= 16 && K <= 256 && (K & (K - 1)) == 0, "power-of-two mesh 16..256");
};

__device__ __forceinline__ int lpos(int k) { return k + (k >> 3); }

__device__ __forceinline__ int br3(int r) { return ((r & 1) << 2) | (r & 2) | ((r >> 2) & 1); }

// The point of the line held by register r of lane b after line_fft.
template <int K>
__device__ __forceinline__ int line_pos(int r, int b) {
  return br3(r) + 8 * (int)(__brev((unsigned)b) >> (32 - LineCfg<K>::LB));
}

// a * W8^j (j = 1, 2, 3), W8 = e^(-2 pi i / 8); conjugated for the inverse.
template <bool INV, typename C2>
__device__ __forceinline__ C2 mul_w8(C2 a, int j) {
  typedef decltype(a.x) R;
  const R s = R(0.70710678118654752440);
  if (j == 1) return INV ? cmk<C2>(s * (a.x - a.y), s * (a.x + a.y)) : cmk<C2>(s * (a.x + a.y), s * (a.y - a.x));
  if (j == 2) return INV ? cmk<C2>(-a.y, a.x) : cmk<C2>(a.y, -a.x);
  return INV ? cmk<C2>(-(s * (a.x + a.y)), s * (a.x - a.y)) : cmk<C2>(s * (a.y - a.x), -(s * (a.x + a.y)));
}

template <bool INV, typename C2>
__device__ __forceinline__ C2 twiddle(const C2* s_tw, int j) {
  const C2 w = s_tw[j];
  return INV ? cconj(w) : w;
}

using msgd::xor_lanes;

// One radix-2 stage between partner lanes at distance D (see line_fft).
template <int K, int D, bool INV, typename C2>
__device__ __forceinline__ void lane_stage(C2 (&v)[8], int b, const C2* s_tw) {
  const bool up = (b & D) != 0;
  const C2 w = twiddle<INV>(s_tw, (b & (D - 1)) * (K / (2 * D)));
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    C2 p;
    p.x = xor_lanes<D>(v[r].x, b);
    p.y = xor_lanes<D>(v[r].y, b);
    v[r] = up ? cmul(csub(p, v[r]), w) : cadd(v[r], p);
  }
}

// The transform of the line whose points lane b of the line's M lanes holds (see above); s_tw
// holds W_K^j, j < K.  All M lanes of the line take part.
template <int K, bool INV, typename C2>
__device__ __forceinline__ void line_fft(C2 (&v)[8], int b, const C2* s_tw) {
  constexpr int M = LineCfg<K>::M;
#pragma unroll
  for (int j = 0; j < 4; ++j) {  // span 4
    const C2 x = v[j], y = v[j + 4];
    v[j] = cadd(x, y);
    v[j + 4] = j == 0 ? csub(x, y) : mul_w8<INV>(csub(x, y), j);
  }
#pragma unroll
  for (int g = 0; g < 8; g += 4) {  // span 2
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const C2 x = v[g + j], y = v[g + j + 2];
      v[g + j] = cadd(x, y);
      v[g + j + 2] = j == 0 ? csub(x, y) : mul_w8<INV>(csub(x, y), 2);
    }
  }
#pragma unroll
  for (int g = 0; g < 8; g += 2) {  // span 1
    const C2 x = v[g], y = v[g + 1];
    v[g] = cadd(x, y);
    v[g + 1] = csub(x, y);
  }
  // register r holds k1 = br3(r): twiddle W_K^(b k1)
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], twiddle<INV>(s_tw, b * br3(r)));
  // the M-point DFT over the lanes: partners at distance d, twiddle W_2d^(b mod d) on the
  // difference
  if constexpr (M >= 32) lane_stage<K, 16, INV>(v, b, s_tw);
  if constexpr (M >= 16) lane_stage<K, 8, INV>(v, b, s_tw);
  if constexpr (M >= 8) lane_stage<K, 4, INV>(v, b, s_tw);
  if constexpr (M >= 4) lane_stage<K, 2, INV>(v, b, s_tw);
  lane_stage<K, 1, INV>(v, b, s_tw);
}

struct ConvArgs {
  double tpl2[3];   // (2 pi / L)^2
  double expf;      // -1 / (4 alpha^2)
};

// Arguments of the FFT kernels of one system's mesh.
struct FftArgs {
  void* Q;            // real mesh K^3
  void* Cm;           // half-complex mesh H*K*K
  const void* tw;     // K twiddles
  const void* G;      // influence function H^3
  ConvArgs ca;
  double* part;       // per zconv workgroup: energy, virial
  // general path (runtime K): mesh size, the Stockham radices (product K)
  int K, nrad;
  int rad[kMaxRadices];
};

// Twiddles W_K^j into LDS (the caller's barrier publishes them).
template <typename C2, int K, int NT>
__device__ __forceinline__ void stage_twiddles(C2* s_tw, const void* tw) {
  const C2* __restrict__ t = (const C2*)tw;
  for (int k = threadIdx.x; k < K; k += NT) s_tw[k] = t[k];
}

// The r2c split of a packed line Z (rows y0 = 2p real part, y0 + 1 imaginary part; points at
// lpos) at kx:
// the spectrum of row y0 (odd = false) or y0 + 1 (odd = true).
template <typename C2, int K>
__device__ __forceinline__ C2 r2c_split(const C2* Z, int kx, bool odd) {
  typedef decltype(C2().x) R;
  const R h = R(0.5);
  const C2 zz = Z[lpos(kx)], zc = cconj(Z[lpos((K - kx) & (K - 1))]);
  if (!odd) {
    const C2 sm = cadd(zz, zc);
    return cmk<C2>(h * sm.x, h * sm.y);
  }
  const C2 d = csub(zz, zc);
  return cmk<C2>(h * d.y, -h * d.x);  // (z - zc) / (2i)
}

// The c2r merge of the half spectra a (row 2p) and b (row 2p + 1) into point k of the packed
// line: a + i b for k <= K/2 (imaginary parts of the kx = 0 and K/2 terms ignored), conj(a) +
// i conj(b) taken from kx = K - k above.  spec(kx, odd) returns the half spectra.
template <typename C2, int K, typename F>
__device__ __forceinline__ C2 c2r_merge(int k, F spec) {
  if (k <= K / 2) {
    C2 a = spec(k, false), b = spec(k, true);
    if (k == 0 || k == K / 2) {
      a.y = 0;
      b.y = 0;
    }
    return cmk<C2>(a.x - b.y, a.y + b.x);
  }
  const C2 a = spec(K - k, false), b = spec(K - k, true);
  return cmk<C2>(a.x + b.y, b.x - a.y);
}

// r2c along x.  Workgroup: T packed lines (rows 2p, 2p + 1 as one complex line; row = y + K z),
// M lanes each.
template <typename real, int K>
__device__ __forceinline__ void fft_x_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int M = LineCfg<K>::M, T = LineCfg<K>::T, H = K / 2 + 1, LSX = LineCfg<K>::LSX;
  const real* __restrict__ Q = (const real*)fa.Q;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  __shared__ C2 A[T * LSX];
  __shared__ C2 s_tw[K];
  stage_twiddles<C2, K, 256>(s_tw, fa.tw);
  const int line = threadIdx.x / M, b = threadIdx.x % M;
  const size_t r0 = 2 * ((size_t)bid * T + line), r1 = r0 + 1;
  C2 v[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) v[n] = cmk<C2>(Q[r0 * K + M * n + b], Q[r1 * K + M * n + b]);
  __syncthreads();
  line_fft<K, false>(v, b, s_tw);
  C2* Z = A + line * LSX;
#pragma unroll
  for (int r = 0; r < 8; ++r) Z[lpos(line_pos<K>(r, b))] = v[r];
  __syncthreads();
  C2* __restrict__ Nq = Cm + (size_t)(K / 2) * K * K;
  for (int k = b; k < H; k += M) {
    const C2 c0 = r2c_split<C2, K>(Z, k, false), c1 = r2c_split<C2, K>(Z, k, true);
    if (k < K / 2) {
      Cm[r0 * (K / 2) + k] = c0;
      Cm[r1 * (K / 2) + k] = c1;
    } else {
      Nq[r0] = c0;
      Nq[r1] = c1;
    }
  }
}

// c2r along x.
template <typename real, int K>
__device__ __forceinline__ void fft_xi_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int M = LineCfg<K>::M, T = LineCfg<K>::T, H = K / 2 + 1, LSX = LineCfg<K>::LSX;
  static_assert(LSX >= 2 * H, "the half spectra fit a line");
  const C2* __restrict__ Cm = (const C2*)fa.Cm;
  real* __restrict__ Q = (real*)fa.Q;
  // first the two rows' half spectra of each line (2H points), then the transformed lines
  __shared__ C2 A[T * LSX];
  C2* S = A;
  __shared__ C2 s_tw[K];
  stage_twiddles<C2, K, 256>(s_tw, fa.tw);
  const int line = threadIdx.x / M, b = threadIdx.x % M;
  const size_t row0 = 2 * (size_t)bid * T;  // the workgroup's 2T rows are consecutive
  const C2* __restrict__ Nq = Cm + (size_t)(K / 2) * K * K;
  for (int e = threadIdx.x; e < 2 * T * (K / 2); e += 256) {
    const int rr = e / (K / 2), k = e % (K / 2);
    S[(rr >> 1) * LSX + (rr & 1) * H + k] = Cm[(row0 + rr) * (K / 2) + k];
  }
  for (int rr = threadIdx.x; rr < 2 * T; rr += 256) S[(rr >> 1) * LSX + (rr & 1) * H + K / 2] = Nq[row0 + rr];
  __syncthreads();
  const C2* Sl = S + line * LSX;
  C2 v[8];
#pragma unroll
  for (int n = 0; n < 8; ++n)
    v[n] = c2r_merge<C2, K>(M * n + b, [&](int kx, bool odd) { return Sl[(odd ? H : 0) + kx]; });
  line_fft<K, true>(v, b, s_tw);
  __syncthreads();  // every half spectrum read before the lines overwrite them
  C2* Z = A + line * LSX;
#pragma unroll
  for (int r = 0; r < 8; ++r) Z[lpos(line_pos<K>(r, b))] = v[r];
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * T * K; e += 256) {  // the 2T rows, coalesced
    const int rr = e / K, k = e % K;
    const C2 z = A[(rr >> 1) * LSX + lpos(k)];
    Q[(row0 + rr) * K + k] = (rr & 1) ? z.y : z.x;
  }
}

// Address of element w (y for the y pass, z for the z pass) of line L of the half-complex
// mesh: regular lines L < K^2/2 are (kx, v) = (L mod K/2, L div K/2) with v the other
// coordinate; lines beyond are the Nyquist column at v = L - K^2/2.
template <int K, bool YPASS>
__device__ __forceinline__ size_t hc_index(int L, int w) {
  constexpr int NH = K / 2, NREG = NH * K;
  if (L < NREG) {
    const int kx = L % NH, v = L / NH;
    return YPASS ? kx + (size_t)NH * (w + (size_t)K * v) : kx + (size_t)NH * (v + (size_t)K * w);
  }
  const int v = L - NREG;
  return (size_t)NH * K * K + (YPASS ? w + (size_t)K * v : v + (size_t)K * w);
}

// Lines of the y and z passes and the workgroups that cover them (T lines each).
template <int K, int T>
struct HcLines {
  static constexpr int N = (K / 2 + 1) * K;
  static constexpr int NB = (N + T - 1) / T;
};

// T lines from the half-complex mesh into LDS (line-major, padded rows of LS), and back.
template <typename C2, int K, bool YPASS>
__device__ __forceinline__ void hc_load(C2* A, const C2* __restrict__ Cm, int L0) {
  constexpr int T = LineCfg<K>::T, LSX = LineCfg<K>::LSX, NL = HcLines<K, T>::N;
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, w = e / T, L = L0 + kl;
    A[kl * LSX + lpos(w)] = L < NL ? Cm[hc_index<K, YPASS>(L, w)] : cmk<C2>(0, 0);
  }
}
template <typename C2, int K, bool YPASS>
__device__ __forceinline__ void hc_store(C2* __restrict__ Cm, const C2* A, int L0) {
  constexpr int T = LineCfg<K>::T, LSX = LineCfg<K>::LSX, NL = HcLines<K, T>::N;
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, w = e / T, L = L0 + kl;
    if (L < NL) Cm[hc_index<K, YPASS>(L, w)] = A[kl * LSX + lpos(w)];
  }
}

// Complex FFT along y for T lines (consecutive kx at one z, or Nyquist lines).
template <typename real, int K>
__device__ __forceinline__ void fft_y_body(const FftArgs& fa, int bid, int inverse) {
  typedef typename C2T<real>::t C2;
  constexpr int M = LineCfg<K>::M, T = LineCfg<K>::T, LSX = LineCfg<K>::LSX;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  __shared__ C2 A[T * LSX];
  __shared__ C2 s_tw[K];
  stage_twiddles<C2, K, 256>(s_tw, fa.tw);
  const int L0 = bid * T;
  hc_load<C2, K, true>(A, Cm, L0);
  __syncthreads();
  const int line = threadIdx.x / M, b = threadIdx.x % M;
  C2* Al = A + line * LSX;
  C2 v[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) v[n] = Al[lpos(M * n + b)];
  if (inverse) line_fft<K, true>(v, b, s_tw);
  else line_fft<K, false>(v, b, s_tw);
  __syncthreads();  // every line read before any is overwritten
#pragma unroll
  for (int r = 0; r < 8; ++r) Al[lpos(line_pos<K>(r, b))] = v[r];
  __syncthreads();
  hc_store<C2, K, true>(Cm, A, L0);
}

// Forward FFT along z, influence-function multiply + FP64 energy/virial, inverse FFT along z.
template <typename real, int K>
__device__ __forceinline__ void fft_zconv_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int M = LineCfg<K>::M, T = LineCfg<K>::T, H = K / 2 + 1, NL = HcLines<K, T>::N;
  constexpr int NH = K / 2, NREG = NH * K, LSX = LineCfg<K>::LSX;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  const real* __restrict__ G = (const real*)fa.G;
  const ConvArgs& ca = fa.ca;
  double* __restrict__ part = fa.part;
  __shared__ C2 A[T * LSX];
  __shared__ C2 s_tw[K];
  __shared__ double red[32];
  stage_twiddles<C2, K, 256>(s_tw, fa.tw);
  const int L0 = bid * T;
  hc_load<C2, K, false>(A, Cm, L0);
  const int line = threadIdx.x / M, b = threadIdx.x % M, L = L0 + line;
  const bool valid = L < NL;
  const int kx = L < NREG ? L % NH : NH, y = L < NREG ? L / NH : L - NREG;
  const int ky = y <= K / 2 ? y : K - y;
  // the influence function at the points this lane will hold, loaded beside the lines
  real g[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int kz = line_pos<K>(r, b), kzz = kz <= K / 2 ? kz : K - kz;
    g[r] = valid ? G[kx + H * (ky + H * kzz)] : real(0);
  }
  __syncthreads();
  C2* Al = A + line * LSX;
  C2 v[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) v[n] = Al[lpos(M * n + b)];
  line_fft<K, false>(v, b, s_tw);
  double esum = 0.0, vsum = 0.0;
  if (valid) {
    // no contraction of the FP64 energy/virial terms: the value must not depend on how
    // the compiler fuses them in the single-system and the batched instantiation
#pragma clang fp contract(off)
    const double w = (kx == 0 || kx == K / 2) ? 0.5 : 1.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int kz = line_pos<K>(r, b), kzz = kz <= K / 2 ? kz : K - kz;
      C2 f = v[r];
      const double f2 = (double)f.x * (double)f.x + (double)f.y * (double)f.y;
      const double m2 = ca.tpl2[0] * kx * kx + ca.tpl2[1] * ky * ky + ca.tpl2[2] * kzz * kzz;
      const double eg = w * (double)g[r] * f2;
      esum += eg;
      vsum += eg * (1.0 + 2.0 * m2 * ca.expf);
      f.x *= g[r];
      f.y *= g[r];
      v[r] = f;
    }
  }
  __syncthreads();  // every line read before any is overwritten
#pragma unroll
  for (int r = 0; r < 8; ++r) Al[lpos(line_pos<K>(r, b))] = v[r];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < 8; ++n) v[n] = Al[lpos(M * n + b)];
  line_fft<K, true>(v, b, s_tw);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) Al[lpos(line_pos<K>(r, b))] = v[r];
  __syncthreads();
  hc_store<C2, K, false>(Cm, A, L0);
  msgd::block_sum2(esum, vsum, red);
  if (threadIdx.x == 0) {
    part[2 * bid] = esum;
    part[2 * bid + 1] = vsum;
  }
}

// Entry points: one system (argument block by value) or a batch (segmented).
#define MSG_FFT_ENTRY(name, body, ...)                                                      \
  template <typename real, int K>                                                           \
  __global__ __launch_bounds__(256) void name(FftArgs a) { body<real, K>(a, blockIdx.x __VA_ARGS__); } \
  template <typename real, int K>                                                           \
  __global__ __launch_bounds__(256) void name##_seg(const FftArgs* __restrict__ args,       \
                                                    const int2* __restrict__ map) {         \
    const int2 s = map[blockIdx.x];                                                         \
    if (s.x < 0) return;                                                                    \
    body<real, K>(args[s.x], s.y __VA_ARGS__);                                              \
  }
MSG_FFT_ENTRY(k_fft_x, fft_x_body)
MSG_FFT_ENTRY(k_fft_xi, fft_xi_body)
MSG_FFT_ENTRY(k_fft_y, fft_y_body, , 0)
MSG_FFT_ENTRY(k_fft_yi, fft_y_body, , 1)
MSG_FFT_ENTRY(k_fft_zconv, fft_zconv_body)
#undef MSG_FFT_ENTRY

// Fused plane passes (FP32, K <= 128, one system): a workgroup of 1024 threads per z-plane
// holds the plane in LDS and does the x r2c and then the y FFT (k_fft_xy), or the inverse y
// FFT and then the x c2r (k_fft_yxi), so the mesh crosses memory once per pair of passes and
// two launches of the chain are saved.  Per line the arithmetic is that of k_fft_x / k_fft_y
// / k_fft_yi / k_fft_xi (the same line transform, split and merge), so the results are the
// same bits as the separate passes (which the batched passes keep).
constexpr int kPlaneNT = 1024;
template <int K>
struct PlaneCfg {
  static constexpr int H = K / 2 + 1, LS = LineCfg<K>::LSX, NH = K / 2;
  static constexpr int NB = H * LS;  // complex elements per buffer (NH x lines or H y lines)
  static constexpr size_t lds = (2 * (size_t)NB + K) * sizeof(float2);
};

template <int K>
__global__ __launch_bounds__(kPlaneNT) void k_fft_xy(FftArgs fa) {
  typedef float2 C2;
  constexpr int H = PlaneCfg<K>::H, LS = PlaneCfg<K>::LS, NB = PlaneCfg<K>::NB, NH = PlaneCfg<K>::NH;
  constexpr int M = LineCfg<K>::M, LPR = kPlaneNT / M;  // lines per round of the workgroup
  extern __shared__ __align__(16) double s_dyn[];
  C2* A = (C2*)s_dyn;
  C2* B = A + NB;
  C2* s_tw = B + NB;
  const float* __restrict__ Q = (const float*)fa.Q;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  stage_twiddles<C2, K, kPlaneNT>(s_tw, fa.tw);
  const size_t z = blockIdx.x;
  const float* __restrict__ Qz = Q + z * K * K;
  const int b = threadIdx.x % M;
  // x: packed line p = rows y = 2p and 2p + 1
  {
    const int p = threadIdx.x / M;
    C2 v[8];
    if (p < NH) {
#pragma unroll
      for (int n = 0; n < 8; ++n) v[n] = cmk<C2>(Qz[(2 * p) * K + M * n + b], Qz[(2 * p + 1) * K + M * n + b]);
    }
    __syncthreads();  // twiddles
    if (p < NH) {
      line_fft<K, false>(v, b, s_tw);
#pragma unroll
      for (int r = 0; r < 8; ++r) A[p * LS + lpos(line_pos<K>(r, b))] = v[r];
    }
  }
  __syncthreads();
  // y: line kx = 0 .. K/2 over y, its points split from the packed x lines
  for (int kx = threadIdx.x / M; kx < H; kx += LPR) {
    C2 v[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int y = M * n + b;
      v[n] = r2c_split<C2, K>(A + (y >> 1) * LS, kx, (y & 1) != 0);
    }
    line_fft<K, false>(v, b, s_tw);
#pragma unroll
    for (int r = 0; r < 8; ++r) B[kx * LS + lpos(line_pos<K>(r, b))] = v[r];
  }
  __syncthreads();
  C2* __restrict__ Nq = Cm + (size_t)NH * K * K;
  for (int e = threadIdx.x; e < NH * K; e += kPlaneNT) {
    const int y = e / NH, kx = e % NH;
    Cm[kx + (size_t)NH * (y + K * z)] = B[kx * LS + lpos(y)];
  }
  for (int y = threadIdx.x; y < K; y += kPlaneNT) Nq[y + K * z] = B[NH * LS + lpos(y)];
}

template <int K>
__global__ __launch_bounds__(kPlaneNT) void k_fft_yxi(FftArgs fa) {
  typedef float2 C2;
  constexpr int H = PlaneCfg<K>::H, LS = PlaneCfg<K>::LS, NB = PlaneCfg<K>::NB, NH = PlaneCfg<K>::NH;
  constexpr int M = LineCfg<K>::M, LPR = kPlaneNT / M;
  extern __shared__ __align__(16) double s_dyn[];
  C2* A = (C2*)s_dyn;
  C2* B = A + NB;
  C2* s_tw = B + NB;
  const C2* __restrict__ Cm = (const C2*)fa.Cm;
  float* __restrict__ Q = (float*)fa.Q;
  stage_twiddles<C2, K, kPlaneNT>(s_tw, fa.tw);
  const size_t z = blockIdx.x;
  const C2* __restrict__ Nq = Cm + (size_t)NH * K * K;
  for (int e = threadIdx.x; e < NH * K; e += kPlaneNT) {
    const int y = e / NH, kx = e % NH;
    A[kx * LS + lpos(y)] = Cm[kx + (size_t)NH * (y + K * z)];
  }
  for (int y = threadIdx.x; y < K; y += kPlaneNT) A[NH * LS + lpos(y)] = Nq[y + K * z];
  __syncthreads();
  const int b = threadIdx.x % M;
  // inverse y: lines kx = 0 .. K/2
  for (int kx = threadIdx.x / M; kx < H; kx += LPR) {
    C2 v[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) v[n] = A[kx * LS + lpos(M * n + b)];
    line_fft<K, true>(v, b, s_tw);
#pragma unroll
    for (int r = 0; r < 8; ++r) B[kx * LS + lpos(line_pos<K>(r, b))] = v[r];
  }
  __syncthreads();
  // inverse x: packed line p merged from the half spectra of rows 2p and 2p + 1
  {
    const int p = threadIdx.x / M;
    if (p < NH) {
      C2 v[8];
#pragma unroll
      for (int n = 0; n < 8; ++n)
        v[n] = c2r_merge<C2, K>(M * n + b, [&](int kx, bool odd) { return B[kx * LS + lpos(2 * p + (odd ? 1 : 0))]; });
      line_fft<K, true>(v, b, s_tw);
#pragma unroll
      for (int r = 0; r < 8; ++r) A[p * LS + lpos(line_pos<K>(r, b))] = v[r];
    }
  }
  __syncthreads();
  float* __restrict__ Qz = Q + z * K * K;
  for (int e = threadIdx.x; e < K * K; e += kPlaneNT) {  // rows y, coalesced
    const int y = e / K, x = e % K;
    const C2 v = A[(y >> 1) * LS + lpos(x)];
    Qz[y * K + x] = (y & 1) ? v.y : v.x;
  }
}

// The fused plane passes serve the single-system mixed path at K <= 128 (the plane and its
// scratch copy fit the LDS of one workgroup).
template <typename real, int K>
constexpr bool plane_fft() {
  return sizeof(real) == 4 && K <= 128;
}

// ------------------------------------------------------------------------------------------
// General path: any even K (the reference's EwaldSetup takes any even nmesh, e.g. its default
// 48 and the 40 of its test decks).  The same five passes with K known only at run time: a
// Stockham autosort FFT over K's factors (4, then 2, 3, 5, 7, ... in any order) where each
// output of a stage is formed by one thread as
//   y[jh Ns r + q Ns + jr] = sum_{m<r} x[jh Ns + jr + m K/r] w^{m (jr + q Ns) K / (Ns r)}
// (w = e^{-2 pi i / K}, conjugated for the inverse; Ns = product of the earlier radices), the
// radix-r generalisation of the radix-4 butterfly above.  Lines are held in LDS, T per
// workgroup (as many as fit).
template <typename C2>
__device__ C2* lds_fft_g(C2* A, C2* B, const C2* __restrict__ tw, int K, int T, int LS, int nrad,
                         const int* rad, bool inverse) {
  C2* src = A;
  C2* dst = B;
  int Ns = 1;
  for (int st = 0; st < nrad; ++st) {
    const int r = rad[st], Kr = K / r, step = K / (Ns * r);
    for (int idx = threadIdx.x; idx < T * K; idx += 256) {
      const int line = idx / K, o = idx - line * K;
      const int jr = o % Ns, q = (o / Ns) % r, jh = o / (Ns * r);
      const C2* x = src + line * LS + jh * Ns + jr;
      const int e = (jr + q * Ns) * step;  // < K
      C2 acc = x[0];
      int ex = 0;
      for (int m = 1; m < r; ++m) {
        ex += e;
        if (ex >= K) ex -= K;
        C2 w = tw[ex];
        if (inverse) w = cconj(w);
        acc = cadd(acc, cmul(x[m * Kr], w));
      }
      dst[line * LS + o] = acc;
    }
    __syncthreads();
    C2* t = src; src = dst; dst = t;
    Ns *= r;
  }
  return src;
}

template <typename real>
__device__ __forceinline__ size_t hc_index_g(int K, int L, int w, bool ypass) {
  const int NH = K / 2, NREG = NH * K;
  if (L < NREG) {
    const int kx = L % NH, v = L / NH;
    return ypass ? kx + (size_t)NH * (w + (size_t)K * v) : kx + (size_t)NH * (v + (size_t)K * w);
  }
  const int v = L - NREG;
  return (size_t)NH * K * K + (ypass ? w + (size_t)K * v : v + (size_t)K * w);
}

// Lines per workgroup of the general passes: x passes hold T line pairs of K, the y/z passes
// T lines of K + 1 (padded) elements.
template <typename real>
__host__ __device__ inline int gen_lines_x(int K) { return FftCfg<real>::E / K; }
template <typename real>
__host__ __device__ inline int gen_lines_yz(int K) { return FftCfg<real>::E / (K + 1); }

template <typename real>
__device__ __forceinline__ void gfft_x_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E;
  __shared__ C2 A[E], B[E], s_tw[E];
  const int K = fa.K, H = K / 2 + 1, T = gen_lines_x<real>(K), NP = K * K / 2;
  const real* __restrict__ Q = (const real*)fa.Q;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  for (int k = threadIdx.x; k < K; k += 256) s_tw[k] = ((const C2*)fa.tw)[k];
  for (int idx = threadIdx.x; idx < T * K; idx += 256) {
    const int line = idx / K, k = idx - line * K, p = bid * T + line;
    A[idx] = p < NP ? cmk<C2>(Q[(size_t)(2 * p) * K + k], Q[(size_t)(2 * p + 1) * K + k]) : cmk<C2>(0, 0);
  }
  __syncthreads();
  const C2* Z = lds_fft_g<C2>(A, B, s_tw, K, T, K, fa.nrad, fa.rad, false);
  const real h = real(0.5);
  C2* __restrict__ Nq = Cm + (size_t)(K / 2) * K * K;
  for (int idx = threadIdx.x; idx < T * H; idx += 256) {
    const int line = idx / H, k = idx - line * H, p = bid * T + line;
    if (p >= NP) continue;
    const size_t r0 = 2 * (size_t)p, r1 = r0 + 1;
    const C2 z = Z[line * K + k], zc = cconj(Z[line * K + (k == 0 ? 0 : K - k)]);
    const C2 sm = cadd(z, zc), df = csub(z, zc);
    const C2 c0 = cmk<C2>(h * sm.x, h * sm.y), c1 = cmk<C2>(h * df.y, -h * df.x);
    if (k < K / 2) {
      Cm[r0 * (K / 2) + k] = c0;
      Cm[r1 * (K / 2) + k] = c1;
    } else {
      Nq[r0] = c0;
      Nq[r1] = c1;
    }
  }
}

template <typename real>
__device__ __forceinline__ void gfft_xi_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E;
  __shared__ C2 A[E], B[E], s_tw[E];
  const int K = fa.K, H = K / 2 + 1, T = gen_lines_x<real>(K), NP = K * K / 2;
  const C2* __restrict__ Cm = (const C2*)fa.Cm;
  real* __restrict__ Q = (real*)fa.Q;
  for (int k = threadIdx.x; k < K; k += 256) s_tw[k] = ((const C2*)fa.tw)[k];
  const C2* __restrict__ Nq = Cm + (size_t)(K / 2) * K * K;
  for (int idx = threadIdx.x; idx < T * H; idx += 256) {
    const int line = idx / H, k = idx - line * H, p = bid * T + line;
    C2 a = cmk<C2>(0, 0), b = cmk<C2>(0, 0);
    if (p < NP) {
      const size_t r0 = 2 * (size_t)p, r1 = r0 + 1;
      a = k < K / 2 ? Cm[r0 * (K / 2) + k] : Nq[r0];
      b = k < K / 2 ? Cm[r1 * (K / 2) + k] : Nq[r1];
    }
    if (k == 0 || k == K / 2) { a.y = 0; b.y = 0; }
    A[line * K + k] = cmk<C2>(a.x - b.y, a.y + b.x);                                   // a + i b
    if (k > 0 && k < K / 2) A[line * K + K - k] = cmk<C2>(a.x + b.y, b.x - a.y);     // conj(a) + i conj(b)
  }
  __syncthreads();
  const C2* Z = lds_fft_g<C2>(A, B, s_tw, K, T, K, fa.nrad, fa.rad, true);
  for (int idx = threadIdx.x; idx < T * K; idx += 256) {
    const int line = idx / K, k = idx - line * K, p = bid * T + line;
    if (p >= NP) continue;
    const C2 z = Z[idx];
    Q[(size_t)(2 * p) * K + k] = z.x;
    Q[(size_t)(2 * p + 1) * K + k] = z.y;
  }
}

template <typename real>
__device__ __forceinline__ void gfft_y_body(const FftArgs& fa, int bid, int inverse) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E;
  __shared__ C2 A[E], B[E], s_tw[E];
  const int K = fa.K, T = gen_lines_yz<real>(K), LS = K + 1, NL = (K / 2 + 1) * K;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  for (int k = threadIdx.x; k < K; k += 256) s_tw[k] = ((const C2*)fa.tw)[k];
  const int L0 = bid * T;
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, y = e / T, L = L0 + kl;
    A[kl * LS + y] = L < NL ? Cm[hc_index_g<real>(K, L, y, true)] : cmk<C2>(0, 0);
  }
  __syncthreads();
  const C2* R = lds_fft_g<C2>(A, B, s_tw, K, T, LS, fa.nrad, fa.rad, inverse != 0);
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, y = e / T, L = L0 + kl;
    if (L < NL) Cm[hc_index_g<real>(K, L, y, true)] = R[kl * LS + y];
  }
}

template <typename real>
__device__ __forceinline__ void gfft_zconv_body(const FftArgs& fa, int bid) {
  typedef typename C2T<real>::t C2;
  constexpr int E = FftCfg<real>::E;
  __shared__ C2 A[E], B[E], s_tw[E];
  __shared__ double red[32];
  const int K = fa.K, H = K / 2 + 1, T = gen_lines_yz<real>(K), LS = K + 1, NL = H * K;
  const int NH = K / 2, NREG = NH * K;
  C2* __restrict__ Cm = (C2*)fa.Cm;
  const real* __restrict__ G = (const real*)fa.G;
  const ConvArgs& ca = fa.ca;
  for (int k = threadIdx.x; k < K; k += 256) s_tw[k] = ((const C2*)fa.tw)[k];
  const int L0 = bid * T;
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, z = e / T, L = L0 + kl;
    A[kl * LS + z] = L < NL ? Cm[hc_index_g<real>(K, L, z, false)] : cmk<C2>(0, 0);
  }
  __syncthreads();
  C2* R = lds_fft_g<C2>(A, B, s_tw, K, T, LS, fa.nrad, fa.rad, false);
  double esum = 0.0, vsum = 0.0;
  for (int e = threadIdx.x; e < T * K; e += 256) {
#pragma clang fp contract(off)
    const int kl = e / K, kz = e % K, L = L0 + kl;
    if (L >= NL) continue;
    const int kx = L < NREG ? L % NH : NH, y = L < NREG ? L / NH : L - NREG;
    const int ky = y <= K / 2 ? y : K - y;
    const int kzz = kz <= K / 2 ? kz : K - kz;
    const real g = G[kx + H * (ky + H * kzz)];
    C2 f = R[kl * LS + kz];
    const double f2 = (double)f.x * (double)f.x + (double)f.y * (double)f.y;
    const double w = (kx == 0 || kx == K / 2) ? 0.5 : 1.0;
    const double m2 = ca.tpl2[0] * kx * kx + ca.tpl2[1] * ky * ky + ca.tpl2[2] * kzz * kzz;
    const double eg = w * (double)g * f2;
    esum += eg;
    vsum += eg * (1.0 + 2.0 * m2 * ca.expf);
    f.x *= g;
    f.y *= g;
    R[kl * LS + kz] = f;
  }
  __syncthreads();
  C2* other = (R == A) ? B : A;
  const C2* S = lds_fft_g<C2>(R, other, s_tw, K, T, LS, fa.nrad, fa.rad, true);
  for (int e = threadIdx.x; e < T * K; e += 256) {
    const int kl = e % T, z = e / T, L = L0 + kl;
    if (L < NL) Cm[hc_index_g<real>(K, L, z, false)] = S[kl * LS + z];
  }
  msgd::block_sum2(esum, vsum, red);
  if (threadIdx.x == 0) {
    fa.part[2 * bid] = esum;
    fa.part[2 * bid + 1] = vsum;
  }
}

#define MSG_GFFT_ENTRY(name, body, ...)                                                     \
  template <typename real>                                                                  \
  __global__ __launch_bounds__(256) void name(FftArgs a) { body<real>(a, blockIdx.x __VA_ARGS__); } \
  template <typename real>                                                                  \
  __global__ __launch_bounds__(256) void name##_seg(const FftArgs* __restrict__ args,       \
                                                    const int2* __restrict__ map) {         \
    const int2 s = map[blockIdx.x];                                                         \
    if (s.x < 0) return;                                                                    \
    body<real>(args[s.x], s.y __VA_ARGS__);                                                 \
  }
MSG_GFFT_ENTRY(k_gfft_x, gfft_x_body)
MSG_GFFT_ENTRY(k_gfft_xi, gfft_xi_body)
MSG_GFFT_ENTRY(k_gfft_y, gfft_y_body, , 0)
MSG_GFFT_ENTRY(k_gfft_yi, gfft_y_body, , 1)
MSG_GFFT_ENTRY(k_gfft_zconv, gfft_zconv_body)
#undef MSG_GFFT_ENTRY

struct SpreadArgs {
  MeshGrid m;
  const int* brick_start;
  const void* bq;     // brick order (frac x, y, z, q)
  const int* bmm;     // brick order packed meshmax
  void* Q;
};

template <typename real> struct R4T;
template <> struct R4T<float> { typedef float4 t; };
template <> struct R4T<double> { typedef double4 t; };

// Fixed-point accumulation of the brick: contributions are added to integer LDS accumulators,
// so the sum is exact and independent of the order the threads add in (deterministic without
// ordering the threads).
//   FP64 path: 64-bit accumulators, scale 2^48 (|Q| < 32768 per mesh point, resolution
//   3.6e-15, below FP64 rounding of the mesh values), rounded to nearest.
//   FP32 path: contributions are rounded half up to integers in units of 2^-s.  The charge
//   weights carry the scale 2^s (exact: a power of two).  The general-mesh spread uses 64-bit
//   accumulators with s = qshift = 30 - e, 2^e >= the largest |charge|: every contribution (a
//   product of three B-spline weights <= 1 and a charge) is below 2^30 in magnitude.  The
//   power-of-two spread uses 32-bit accumulators with a scale per brick (brick_shift).
constexpr double kQScale = 281474976710656.0;  // 2^48

// floor(x) as a 32-bit integer in one instruction (V_CVT_FLR_I32_F32), |x| < 2^31.
__device__ __forceinline__ int cvt_flr_i32(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
// The contribution a * b (scaled) as a 64-bit accumulator increment: FP32, a * b + 0.5 in one
// rounding, then floor (round half up); FP64, a * b * 2^48 rounded to nearest.
__device__ __forceinline__ unsigned long long to_fixed(float a, float b) {
  return (unsigned long long)(long long)cvt_flr_i32(__builtin_fmaf(a, b, 0.5f));
}
__device__ __forceinline__ unsigned long long to_fixed(double a, double b) {
  return (unsigned long long)__double2ll_rn(a * b * kQScale);
}
// Largest product of three order-P B-spline weights, (max_x M_P(x))^3, rounded up.
__host__ __device__ constexpr float wmax3(int P) {
  return P <= 2 ? 1.0f : P == 3 ? 0.4220f : P == 4 ? 0.2964f : P == 5 ? 0.2150f : P == 6 ? 0.1665f
       : P == 7 ? 0.1335f : 0.1102f;
}
// ceil(log2(x)) for x > 0 (the power of two above or at x); 0 for x = 0.
__device__ __forceinline__ int ceil_log2f(float x) {
  int e;
  const float m = frexpf(x, &e);
  return m == 0.5f ? e - 1 : e;
}
// Scale exponent s of a brick's 32-bit accumulators.  A mesh point receives at most one
// contribution from each of the brick's ncand candidates, each at most 2^e * wmax3(P) in
// magnitude, so 2^s * ncand * 2^e * wmax3(P) <= 2^30 bounds every contribution and every sum
// (with the half units of rounding) below 2^31: the integer sums are exact.
template <int P>
__device__ __forceinline__ int brick_shift(const MeshGrid& m, int ncand) {
  return m.qshift - ceil_log2f((float)ncand * wmax3(P));
}
__device__ __forceinline__ void spread_add(unsigned* p, float a, float b) {
  atomicAdd(p, (unsigned)cvt_flr_i32(__builtin_fmaf(a, b, 0.5f)));
}
__device__ __forceinline__ void spread_add(unsigned long long* p, double a, double b) {
  atomicAdd(p, to_fixed(a, b));
}
__device__ __forceinline__ double acc_value(unsigned v) { return (double)(int)v; }
__device__ __forceinline__ double acc_value(unsigned long long v) { return (double)(long long)v; }

// Weight scale folded into the charge, and the factor back to mesh values.
template <typename real>
__device__ __forceinline__ real charge_scale(const MeshGrid& m) {
  if constexpr (sizeof(real) == 4) return ldexpf(1.0f, m.qshift);
  else return (real)1.0;
}
template <typename real>
__device__ __forceinline__ double fixed_unit(const MeshGrid& m) {
  if constexpr (sizeof(real) == 4) return ldexp(1.0, -m.qshift);
  else return 1.0 / kQScale;
}

// Stencil range of one dimension inside a brick of extent B starting at the particle's
// local base point l (the stencil covers l, l-1, ..., l-P+1): with one brick across the
// mesh every point is inside (indices wrap); otherwise only i with 0 <= l - i < B.
template <int P>
__device__ inline void brick_range(int l, int B, bool whole, int& i0, int& n) {
  if (whole) {
    i0 = 0;
    n = P;
  } else {
    i0 = max(0, l - B + 1);
    n = max(0, min(P - 1, l) - i0 + 1);
  }
}

// Exclusive prefix over the wavefront's lanes of v (0 <= v < 16) and the wavefront total,
// from four ballots of its bits.
__device__ __forceinline__ int wave_prefix4(int v, int lane, int& total) {
  const unsigned long long lt = (1ull << lane) - 1ull;
  int pre = 0, tot = 0;
#pragma unroll
  for (int bit = 0; bit < 4; ++bit) {
    const unsigned long long m = __ballot((v >> bit) & 1);
    pre += __popcll(m & lt) << bit;
    tot += __popcll(m) << bit;
  }
  total = tot;
  return pre;
}

// One workgroup per brick.  The candidates are the particles of the brick and of the next
// brick in each dimension (their stencils reach back into this one).  The four wavefronts
// work independently (no workgroup barrier between the zeroed accumulator and the final
// write): each takes 64 candidates at a time, the next 64 loaded while it works on these.
// (1) Every lane clips its candidate's stencil to the brick from the packed base point; the
// candidates with points inside are ranked by ballot.  (2) A contributor writes its record to
// the wavefront's LDS area (x weights with those of points outside the brick set to zero, y
// weights, z weights times the charge, local base point, clipped y extent) and its tasks,
// one per clipped z plane, at offsets from a ballot prefix sum; a task code carries the
// contributor, the z index and the first clipped y offset, so a task's reads (record, y
// weights, z weight) are issued together.  (3) The lanes take the tasks and add the clipped
// y rows of P x points into the brick with integer LDS atomics (32-bit in FP32).  The x points outside
// the brick add zero at addresses that stay inside the padded array.  Finally every mesh
// point of the brick is written once.
template <typename real, int P, int BX, bool CLS>
__device__ __forceinline__ void spread_body(const SpreadArgs& a, int bid) {
  typedef typename R4T<real>::t R4;
  // wavefronts, candidates per wavefront round, contributors per record pass (a round with
  // more contributors takes two passes: half the record area, more workgroups per CU)
  constexpr int NW = 4, WC = 64, HC = 32;
  constexpr int NBP = BX * kBY * kBZ;
  // rows padded to BX + 1 accumulators: the lanes adding one particle's rows (same x,
  // consecutive y/z) then fall into different LDS banks instead of one.  A row's x points
  // lie at lx0 - ix, lx0 < 2 BX (own or next brick), ix < P: the array is padded by P - 1
  // in front and 2 BX behind, so the zero-weight points of any row stay inside it.
  constexpr int ROW = BX + 1, NACC = ROW * kBY * kBZ;
  constexpr int PADL = P - 1;
  // contributor record: wx[P] | wy[P] | wz[P] | base point | ny, padded to 16 bytes
  constexpr int RW = sizeof(real) == 4 ? ((3 * P + 2 + 3) & ~3) : ((3 * P + 2 + 1) & ~1);
  typedef typename std::conditional<sizeof(real) == 4, unsigned, unsigned long long>::type acc_t;
  __shared__ acc_t acc_s[PADL + NACC + 2 * BX];
  acc_t* const acc = acc_s + PADL;
  __shared__ __align__(16) real rec_s[NW][HC][RW];
  __shared__ unsigned short tsk_s[NW][HC * P];  // tasks: contributor | z index << 8 | y0 << 11
  const int K = a.m.K;
  const int b = bid;
  const int bx = b % a.m.nb[0], by = (b / a.m.nb[0]) % a.m.nb[1], bz = b / (a.m.nb[0] * a.m.nb[1]);
  const int lo[3] = {bx * BX, by * kBY, bz * kBZ};
  const bool whole[3] = {a.m.nb[0] == 1, a.m.nb[1] == 1, a.m.nb[2] == 1};
  // source ranges: the brick's own particles and its next bricks' in each dimension.  With
  // classes (CLS, brick_key) a next brick contributes only the classes whose stencils reach back
  // into this brick: 14 ranges of consecutive classes, range t = (next-brick offset bits
  // x | y << 1 | z << 2, first class, last class), packed three bits per range; without, the 8
  // bricks whole.  Lanes 0..NR-1 of every wavefront load them; every wavefront then holds them
  // in scalar registers (no LDS round trip, no barrier), so a candidate's place is found
  // without LDS reads.
  constexpr int NR = CLS ? 14 : 8;
  constexpr unsigned long long kOff = 0ull | 1ull << 3 | 1ull << 6 | 1ull << 9 | 1ull << 12 | 2ull << 15 |
                                      2ull << 18 | 4ull << 21 | 3ull << 24 | 3ull << 27 | 5ull << 30 |
                                      5ull << 33 | 6ull << 36 | 7ull << 39;
  constexpr unsigned long long kC0 = 0ull | 1ull << 3 | 3ull << 6 | 5ull << 9 | 7ull << 12 | 2ull << 15 |
                                     6ull << 18 | 4ull << 21 | 3ull << 24 | 7ull << 27 | 5ull << 30 |
                                     7ull << 33 | 6ull << 36 | 7ull << 39;
  constexpr unsigned long long kC1 = 7ull | 1ull << 3 | 3ull << 6 | 5ull << 9 | 7ull << 12 | 3ull << 15 |
                                     7ull << 18 | 7ull << 21 | 3ull << 24 | 7ull << 27 | 5ull << 30 |
                                     7ull << 33 | 7ull << 36 | 7ull << 39;  // last class of the range
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int beg = 0, len = 0;
  if constexpr (CLS) {
    if (lane < NR) {
      const int sh3 = 3 * lane;
      const int off = (int)((kOff >> sh3) & 7), c0 = (int)((kC0 >> sh3) & 7), c1 = (int)((kC1 >> sh3) & 7) + 1;
      const int ix = off & 1, iy = (off >> 1) & 1, iz = off >> 2;
      if (!((ix && whole[0]) || (iy && whole[1]) || (iz && whole[2]))) {
        const int mb = (bx + ix) % a.m.nb[0] +
                       a.m.nb[0] * ((by + iy) % a.m.nb[1] + a.m.nb[1] * ((bz + iz) % a.m.nb[2]));
        beg = a.brick_start[8 * mb + c0];
        len = a.brick_start[8 * mb + c1] - beg;
      }
    }
  } else if (lane < 8) {  // without classes: the (up to) 8 bricks whole
    const int ox = whole[0] ? 1 : 2, oy = whole[1] ? 1 : 2, oz = whole[2] ? 1 : 2;
    if (lane < ox * oy * oz) {
      const int ix = lane % ox, iy = (lane / ox) % oy, iz = lane / (ox * oy);
      const int mb = (bx + ix) % a.m.nb[0] +
                     a.m.nb[0] * ((by + iy) % a.m.nb[1] + a.m.nb[1] * ((bz + iz) % a.m.nb[2]));
      beg = a.brick_start[mb];
      len = a.brick_start[mb + 1] - beg;
    }
  }
  for (int k = threadIdx.x; k < PADL + NACC + 2 * BX; k += 256) acc_s[k] = 0;
  int pref[NR + 1], dlt[NR];
  {
    int run = 0;
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      pref[t] = run;
      dlt[t] = __builtin_amdgcn_readlane(beg, t) - run;
      run += __builtin_amdgcn_readlane(len, t);
    }
    pref[NR] = run;
  }
  const int total = pref[NR];
  __syncthreads();  // the accumulator is zeroed
  const R4* bq = (const R4*)a.bq;
  // FP32: the brick's scale from its candidate count; FP64: 2^48 in the conversion
  const int sh = sizeof(real) == 4 ? brick_shift<P>(a.m, total) : 0;
  const real qs = sizeof(real) == 4 ? (real)ldexpf(1.0f, sh) : (real)1;
  real (*rec)[RW] = rec_s[w];
  unsigned short* tsk = tsk_s[w];
  int t = w * WC + lane;
  int packed = 0;
  R4 v{};
  if (t < total) {
    int d = dlt[0];
#pragma unroll
    for (int r = 1; r < NR; ++r) d = t >= pref[r] ? dlt[r] : d;
    packed = a.bmm[t + d];
    v = bq[t + d];
  }
  for (; t - lane < total; t += NW * WC) {
    // the next round's candidate, loaded while this one is worked on
    const int tn = t + NW * WC;
    int packed_n = 0;
    R4 v_n{};
    if (tn < total) {
      int d = dlt[0];
#pragma unroll
      for (int r = 1; r < NR; ++r) d = tn >= pref[r] ? dlt[r] : d;
      packed_n = a.bmm[tn + d];
      v_n = bq[tn + d];
    }
    bool hit = false;
    int lp = 0, rr = 0;
    if (t < total) {
      int l[3] = {(packed & 1023) - lo[0], ((packed >> 10) & 1023) - lo[1], (packed >> 20) - lo[2]};
#pragma unroll
      for (int e = 0; e < 3; ++e) l[e] += l[e] < 0 ? K : 0;
      int x0, nx, y0, ny, z0, nz;
      brick_range<P>(l[0], BX, whole[0], x0, nx);
      brick_range<P>(l[1], kBY, whole[1], y0, ny);
      brick_range<P>(l[2], kBZ, whole[2], z0, nz);
      hit = nx > 0 && ny > 0 && nz > 0;
      lp = l[0] | (l[1] << 10) | (l[2] << 20);
      rr = x0 | (nx << 4) | (y0 << 8) | (ny << 12) | (z0 << 16) | (nz << 20);
    }
    const unsigned long long bal = __ballot(hit);
    const int nc = __popcll(bal);  // wavefront-uniform
    const int kk = __popcll(bal & ((1ull << lane) - 1ull));
    for (int kb = 0; kb < nc; kb += HC) {
      const bool mine = hit && kk >= kb && kk < kb + HC;
      int ntask;
      const int t0 = wave_prefix4(mine ? (rr >> 20) & 15 : 0, lane, ntask);
      if (mine) {
        const int k = kk - kb;
        const int x0 = rr & 15, nx = (rr >> 4) & 15, y0 = (rr >> 8) & 15, ny = (rr >> 12) & 15;
        const int z0 = (rr >> 16) & 15, nz = (rr >> 20) & 15;
        real wx[P], wy[P], wz[P];  // each contributor is in exactly one pass
        msgd::bspline<real, P>(v.x, wx, (real*)nullptr);
        msgd::bspline<real, P>(v.y, wy, (real*)nullptr);
        msgd::bspline<real, P>(v.z, wz, (real*)nullptr);
        const real qw = v.w * qs;
        real* R = rec[k];
#pragma unroll
        for (int i = 0; i < P; ++i) {
          R[i] = (i >= x0 && i < x0 + nx) ? wx[i] : (real)0;
          R[P + i] = wy[i];
          R[2 * P + i] = wz[i] * qw;
        }
        ((int*)(R + 3 * P))[0] = lp;
        ((int*)(R + 3 * P))[sizeof(real) / sizeof(int)] = ny;
        for (int j = 0; j < nz; ++j) tsk[t0 + j] = (unsigned short)(k | ((z0 + j) << 8) | (y0 << 11));
      }
      msgd::wave_sync();
      for (int task = lane; task < ntask; task += 64) {
        const int code = tsk[task];
        const int k = code & 255, iz = (code >> 8) & 7, y0 = code >> 11;
        const real* R = rec[k];
        // every read of the task depends only on its code
        real tx[P], ty[P];
#pragma unroll
        for (int i = 0; i < P; ++i) {
          tx[i] = R[i];
          ty[i] = R[P + y0 + i];  // P + y0 + i < 3 P: inside the record; beyond ny unused
        }
        const real wz = R[2 * P + iz];
        const int cl = ((const int*)(R + 3 * P))[0];
        const int ny = ((const int*)(R + 3 * P))[sizeof(real) / sizeof(int)];
        const int lx0 = cl & 1023, ly0 = ((cl >> 10) & 1023) - y0;
        int lz = (cl >> 20) - iz;
        lz += lz < 0 ? K : 0;  // wraps only where one brick spans the mesh
#pragma unroll
        for (int ys = 0; ys < P; ++ys) {
          if (ys >= ny) break;
          int ly = ly0 - ys;
          ly += ly < 0 ? K : 0;
          const real wyz = ty[ys] * wz;
          acc_t* rowp = acc + ROW * (ly + kBY * lz);
          if (whole[0]) {
#pragma unroll
            for (int ix = 0; ix < P; ++ix) {
              int lx = lx0 - ix;
              lx += lx < 0 ? K : 0;
              spread_add(rowp + lx, tx[ix], wyz);
            }
          } else {
            // x points lx0 - ix as constant offsets from one base
            acc_t* rp = rowp + lx0 - (P - 1);
#pragma unroll
            for (int ix = 0; ix < P; ++ix) spread_add(rp + (P - 1 - ix), tx[ix], wyz);
          }
        }
      }
      msgd::wave_sync();  // the records and tasks are rewritten by the next pass
    }
    packed = packed_n;
    v = v_n;
  }
  __syncthreads();
  real* Q = (real*)a.Q;
  const double unit = sizeof(real) == 4 ? ldexp(1.0, -sh) : 1.0 / kQScale;
  for (int k = threadIdx.x; k < NBP; k += 256) {
    const int lx = k % BX, ly = (k / BX) % kBY, lz = k / (BX * kBY);
    const double v = acc_value(acc[ROW * (ly + kBY * lz) + lx]) * unit;
    Q[(lo[0] + lx) + (size_t)K * ((lo[1] + ly) + (size_t)K * (lo[2] + lz))] = (real)v;
  }
}

// Registers capped for six workgroups per CU (the LDS use allows seven).
template <typename real, int P, int BX, bool CLS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_spread(SpreadArgs a) {
  spread_body<real, P, BX, CLS>(a, blockIdx.x);
}

template <typename real, int P, int BX, bool CLS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_spread_seg(const SpreadArgs* __restrict__ args,
                                                    const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  spread_body<real, P, BX, CLS>(args[s.x], s.y);
}

// General-path spread: the same algorithm over bricks of any extent.  Brick b of a dimension
// split into nb bricks spans [floor(b K / nb), floor((b + 1) K / nb)), at least order - 1
// points, so again only a brick's own and next bricks hold its contributors.  The fixed-point
// accumulator is sized at launch for the largest brick (dynamic LDS); fewer candidates are
// staged per round for high orders.
template <typename real, int P>
__device__ __forceinline__ void spread_body_g(const SpreadArgs& a, int bid) {
  typedef typename R4T<real>::t R4;
  constexpr int CAP = P <= 8 ? 128 : 64;
  extern __shared__ unsigned long long acc_g[];
  __shared__ real th[CAP][3 * P];
  __shared__ int loc[CAP];
  __shared__ int rng[CAP];
  __shared__ unsigned short tsk[CAP * P];  // tasks: contributor | z plane << 8
  __shared__ int wcnt[4];
  __shared__ int ntsk;
  __shared__ int r_beg[8], r_pref[9], r_len[8];
  const int K = a.m.K;
  const int nbx = a.m.nb[0], nby = a.m.nb[1], nbz = a.m.nb[2];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  const int bb[3] = {bx, by, bz}, nbd[3] = {nbx, nby, nbz};
  int lo[3], ext[3];
  bool whole[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = bb[d] * K / nbd[d];
    ext[d] = (bb[d] + 1) * K / nbd[d] - lo[d];
    whole[d] = nbd[d] == 1;
  }
  const int ROW = ext[0] + 1, NACC = ROW * ext[1] * ext[2];
  if (threadIdx.x < 8) {
    const int ox = whole[0] ? 1 : 2, oy = whole[1] ? 1 : 2, oz = whole[2] ? 1 : 2;
    const int t = threadIdx.x;
    int beg = 0, len = 0;
    if (t < ox * oy * oz) {
      const int ix = t % ox, iy = (t / ox) % oy, iz = t / (ox * oy);
      const int mb = (bx + ix) % nbx + nbx * ((by + iy) % nby + nby * ((bz + iz) % nbz));
      beg = a.brick_start[mb];
      len = a.brick_start[mb + 1] - beg;
    }
    r_beg[t] = beg;
    r_len[t] = len;
    if (t == 0) ntsk = 0;
  }
  for (int k = threadIdx.x; k < NACC; k += 256) acc_g[k] = 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    int pref = 0;
    for (int t = 0; t < 8; ++t) {
      r_pref[t] = pref;
      pref += r_len[t];
    }
    r_pref[8] = pref;
  }
  __syncthreads();
  const int total = r_pref[8];
  const R4* bq = (const R4*)a.bq;
  const real qs = charge_scale<real>(a.m);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int base = 0; base < total; base += CAP) {
    const int t = base + threadIdx.x;
    bool hit = false;
    int lp = 0, rr = 0;
    R4 v;
    if (threadIdx.x < CAP && t < total) {
      int r = 0;
      while (r < 7 && t >= r_pref[r + 1]) ++r;
      const int idx = r_beg[r] + (t - r_pref[r]);
      const int packed = a.bmm[idx];
      v = bq[idx];
      int l[3] = {(packed & 1023) - lo[0], ((packed >> 10) & 1023) - lo[1], (packed >> 20) - lo[2]};
#pragma unroll
      for (int d = 0; d < 3; ++d) l[d] += l[d] < 0 ? K : 0;
      int x0, nx, y0, ny, z0, nz;
      brick_range<P>(l[0], ext[0], whole[0], x0, nx);
      brick_range<P>(l[1], ext[1], whole[1], y0, ny);
      brick_range<P>(l[2], ext[2], whole[2], z0, nz);
      hit = nx > 0 && ny > 0 && nz > 0;
      lp = l[0] | (l[1] << 10) | (l[2] << 20);
      rr = x0 | (nx << 4) | (y0 << 8) | (ny << 12) | (z0 << 16) | (nz << 20);
    }
    const unsigned long long bal = __ballot(hit);
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int off = 0;
    for (int k = 0; k < w; ++k) off += wcnt[k];
    if (hit) {
      const int k = off + __popcll(bal & ((1ull << lane) - 1ull));
      loc[k] = lp;
      rng[k] = rr;
      real* T = th[k];
      msgd::bspline<real, P>(v.x, T, (real*)nullptr);
      msgd::bspline<real, P>(v.y, T + P, (real*)nullptr);
      msgd::bspline<real, P>(v.z, T + 2 * P, (real*)nullptr);
      const real qw = v.w * qs;
#pragma unroll
      for (int i = 0; i < P; ++i) T[2 * P + i] *= qw;
      const int nz = (rr >> 20) & 15;
      const int t0 = atomicAdd(&ntsk, nz);  // task order varies; integer sums do not
      for (int j = 0; j < nz; ++j) tsk[t0 + j] = (unsigned short)(k | (j << 8));
    }
    __syncthreads();
    const int ntask = ntsk;
    for (int task = threadIdx.x; task < ntask; task += 256) {
      const int code = tsk[task];
      const int k = code & 255, zs = code >> 8;
      const int cr = rng[k];
      const int x0 = cr & 15, nx = (cr >> 4) & 15, y0 = (cr >> 8) & 15, ny = (cr >> 12) & 15;
      const int iz = ((cr >> 16) & 15) + zs;
      const int cl = loc[k];
      const int lx0 = cl & 1023, ly0 = (cl >> 10) & 1023;
      int lz = (cl >> 20) - iz;
      lz += lz < 0 ? K : 0;
      const real* T = th[k];
      const real wz = T[2 * P + iz];
      for (int ys = 0; ys < ny; ++ys) {
        const int iy = y0 + ys;
        int ly = ly0 - iy;
        ly += ly < 0 ? K : 0;
        const real wyz = T[P + iy] * wz;
        unsigned long long* rowp = acc_g + ROW * (ly + ext[1] * lz);
        for (int ix = x0; ix < x0 + nx; ++ix) {
          int lx = lx0 - ix;
          lx += lx < 0 ? K : 0;
          atomicAdd(rowp + lx, to_fixed(T[ix], wyz));
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) ntsk = 0;  // read by every thread before the barrier above
  }
  real* Q = (real*)a.Q;
  const int npts = ext[0] * ext[1] * ext[2];
  const double unit = fixed_unit<real>(a.m);
  for (int k = threadIdx.x; k < npts; k += 256) {
    const int lx = k % ext[0], ly = (k / ext[0]) % ext[1], lz = k / (ext[0] * ext[1]);
    const double v = (double)(long long)acc_g[ROW * (ly + ext[1] * lz) + lx] * unit;
    Q[(lo[0] + lx) + (size_t)K * ((lo[1] + ly) + (size_t)K * (lo[2] + lz))] = (real)v;
  }
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_spread_g(SpreadArgs a) { spread_body_g<real, P>(a, blockIdx.x); }

template <typename real, int P>
__global__ __launch_bounds__(256) void k_spread_g_seg(const SpreadArgs* __restrict__ args,
                                                      const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  spread_body_g<real, P>(args[s.x], s.y);
}

// Dynamic LDS of the general spread: the accumulator of the largest brick.
inline size_t spread_g_lds(const MeshGrid& m) {
  int e[3];
  for (int d = 0; d < 3; ++d) e[d] = (m.K + m.nb[d] - 1) / m.nb[d];
  return (size_t)(e[0] + 1) * e[1] * e[2] * sizeof(unsigned long long);
}
struct GatherArgs {
  MeshGrid m;
  int np;
  const int* brick_ip;
  const void* bq;
  const int* bmm;
  const void* phi;
  double epsi4pi;
  double surf2fac;          // 2 * (2 pi / (3 V)) when lsurf, else 0
  const double* Mtot;       // total dipole sum q r
  double* force;            // 3*np caller order (pair force already there)
  double scale;
};

// LPP lanes per particle (brick order; 8, or 16 for orders above 8); lane ix < P takes x
// offset ix of the stencil and walks its P x P (y, z) rows, so the lanes of a particle read
// neighbouring words of one row (one cache line per particle and load instead of one per
// z-plane).
template <typename real, int P, int LPP = 8>
__device__ __forceinline__ void gather_body(const GatherArgs& a, int bid) {
  typedef typename R4T<real>::t R4;
  const double M[3] = {a.Mtot[0], a.Mtot[1], a.Mtot[2]};
  const int gid = bid * 256 + threadIdx.x;
  const int i = gid / LPP, ix = gid % LPP;
  const bool valid = i < a.np;
  const int K = a.m.K;
  real e0 = 0, e1 = 0, e2 = 0;
  R4 v;
  v.x = v.y = v.z = v.w = real(0);
  // the particle's index and pair force, loaded first so the loads overlap the field sums
  int ip = 0;
  double pf[3] = {0.0, 0.0, 0.0};
  if (valid && ix == 0) {
    ip = a.brick_ip[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) pf[d] = a.force[3 * ip + d];
  }
  if (valid) {
    v = ((const R4*)a.bq)[i];
    const int packed = a.bmm[i];
    if (ix < P) {
      real tx[P], dtx[P], ty[P], dty[P], tz[P], dtz[P];
      msgd::bspline<real, P>(v.x, tx, dtx);
      msgd::bspline<real, P>(v.y, ty, dty);
      msgd::bspline<real, P>(v.z, tz, dtz);
      const int mx = packed & 1023, my = (packed >> 10) & 1023, mz = packed >> 20;
      int x = mx - ix;
      x += x < 0 ? K : 0;
      real thx = real(0), dthx = real(0);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        thx = (k == ix) ? tx[k] : thx;
        dthx = (k == ix) ? dtx[k] : dthx;
      }
      int ys[P];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int y = my - k;
        ys[k] = K * (y < 0 ? y + K : y);
      }
      const real* phi = (const real*)a.phi + x;
#pragma unroll
      for (int jz = 0; jz < P; ++jz) {
        int z = mz - jz;
        z += z < 0 ? K : 0;
        const real* plane = phi + (size_t)K * K * z;
        real sy = 0, dsy = 0;
#pragma unroll
        for (int jy = 0; jy < P; ++jy) {
          const real w = plane[ys[jy]];
          sy += ty[jy] * w;
          dsy += dty[jy] * w;
        }
        e0 += dthx * tz[jz] * sy;
        e1 += thx * tz[jz] * dsy;
        e2 += thx * dtz[jz] * sy;
      }
    }
  }
  e0 = msgd::xor_sum<LPP>(e0, threadIdx.x);
  e1 = msgd::xor_sum<LPP>(e1, threadIdx.x);
  e2 = msgd::xor_sum<LPP>(e2, threadIdx.x);
  if (valid && ix == 0) {
    const double q = (double)v.w;
    const double es[3] = {(double)e0, (double)e1, (double)e2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double f = -q * es[d] * a.m.dkdr[d] * a.epsi4pi - a.epsi4pi * a.surf2fac * q * M[d];
      a.force[3 * ip + d] = (pf[d] + f) * a.scale;
    }
  }
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
  gather_body<real, P>(a, xcd_logical(blockIdx.x, gridDim.x));
}

template <typename real, int P>
constexpr int gather_lanes() { return P <= 8 ? 8 : 16; }

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_g(GatherArgs a) {
  gather_body<real, P, gather_lanes<real, P>()>(a, xcd_logical(blockIdx.x, gridDim.x));
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_g_seg(const GatherArgs* __restrict__ args,
                                                      const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  gather_body<real, P, gather_lanes<real, P>()>(args[s.x], s.y);
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_reduce_g(GatherArgs a, ReduceArgs r) {
  if (blockIdx.x == gridDim.x - 1)
    reduce_body<256>(r);
  else
    gather_body<real, P, gather_lanes<real, P>()>(a, xcd_logical(blockIdx.x, gridDim.x - 1));
}

// The gather with the pass's final reduction as one extra (last) workgroup.  The reduction
// reads only what earlier kernels of the pass wrote (k-space partials, the pre-summed pair
// row, the dipole, the flags), so it needs no ordering against the gather's workgroups, and
// the pass saves one dependent launch.
template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_reduce(GatherArgs a, ReduceArgs r) {
  if (blockIdx.x == gridDim.x - 1)
    reduce_body<256>(r);
  else
    gather_body<real, P>(a, xcd_logical(blockIdx.x, gridDim.x - 1));
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_seg(const GatherArgs* __restrict__ args,
                                                    const int2* __restrict__ map) {
  const int2 s = map[blockIdx.x];
  if (s.x < 0) return;  // no work for this XCD slot
  gather_body<real, P>(args[s.x], s.y);
}

// ------------------------------------------------------------------------------------------
FftArgs fft_args(Ctx* c) {
  FftArgs a;
  a.Q = c->d_Q;
  a.Cm = c->d_C;
  a.tw = c->d_tw;
  a.G = c->d_G;
  for (int d = 0; d < 3; ++d) a.ca.tpl2[d] = std::pow(2.0 * M_PI / c->P.boxlen[d], 2);
  a.ca.expf = -1.0 / (4.0 * c->P.alpha * c->P.alpha);
  a.part = c->d_part + c->part_off_spme;
  a.K = c->mesh.K;
  a.nrad = c->nrad;
  for (int k = 0; k < kMaxRadices; ++k) a.rad[k] = c->rad[k];
  return a;
}

// The five general passes (runtime K): x, y, zconv, yi, xi.
template <typename real>
void gfft_blocks(int K, int* nb) {
  const int tx = gen_lines_x<real>(K), tyz = gen_lines_yz<real>(K);
  nb[0] = nb[4] = (K * K / 2 + tx - 1) / tx;
  nb[1] = nb[2] = nb[3] = ((K / 2 + 1) * K + tyz - 1) / tyz;
}

template <typename real>
void launch_gfft_all(Ctx* c) {
  int nb[5];
  gfft_blocks<real>(c->mesh.K, nb);
  const FftArgs a = fft_args(c);
  prof_begin(c, MSG_K_FFT_X);
  hipLaunchKernelGGL((k_gfft_x<real>), dim3(nb[0]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_X);
  prof_begin(c, MSG_K_FFT_Y);
  hipLaunchKernelGGL((k_gfft_y<real>), dim3(nb[1]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_Y);
  prof_begin(c, MSG_K_FFT_ZCONV);
  hipLaunchKernelGGL((k_gfft_zconv<real>), dim3(nb[2]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_ZCONV);
  prof_begin(c, MSG_K_FFT_YI);
  hipLaunchKernelGGL((k_gfft_yi<real>), dim3(nb[3]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_YI);
  prof_begin(c, MSG_K_FFT_XI);
  hipLaunchKernelGGL((k_gfft_xi<real>), dim3(nb[4]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_XI);
}

// Workgroups of the five FFT stages for mesh K: x, y, zconv, yi, xi.
template <typename real, int K>
void fft_blocks(int* nb) {
  constexpr int T = LineCfg<K>::T;
  nb[0] = nb[4] = K * K / 2 / T;
  nb[1] = nb[2] = nb[3] = HcLines<K, T>::NB;
}

template <typename real, int K>
void launch_fft_all(Ctx* c) {
  int nb[5];
  fft_blocks<real, K>(nb);
  const FftArgs a = fft_args(c);
  if constexpr (plane_fft<real, K>()) {
    prof_begin(c, MSG_K_FFT_X);
    hipLaunchKernelGGL((k_fft_xy<K>), dim3(K), dim3(kPlaneNT), PlaneCfg<K>::lds, c->stream, a);
    prof_end(c, MSG_K_FFT_X);
    prof_begin(c, MSG_K_FFT_ZCONV);
    hipLaunchKernelGGL((k_fft_zconv<real, K>), dim3(nb[2]), dim3(256), 0, c->stream, a);
    prof_end(c, MSG_K_FFT_ZCONV);
    prof_begin(c, MSG_K_FFT_XI);
    hipLaunchKernelGGL((k_fft_yxi<K>), dim3(K), dim3(kPlaneNT), PlaneCfg<K>::lds, c->stream, a);
    prof_end(c, MSG_K_FFT_XI);
    return;
  }
  prof_begin(c, MSG_K_FFT_X);
  hipLaunchKernelGGL((k_fft_x<real, K>), dim3(nb[0]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_X);
  prof_begin(c, MSG_K_FFT_Y);
  hipLaunchKernelGGL((k_fft_y<real, K>), dim3(nb[1]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_Y);
  prof_begin(c, MSG_K_FFT_ZCONV);
  hipLaunchKernelGGL((k_fft_zconv<real, K>), dim3(nb[2]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_ZCONV);
  prof_begin(c, MSG_K_FFT_YI);
  hipLaunchKernelGGL((k_fft_yi<real, K>), dim3(nb[3]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_YI);
  prof_begin(c, MSG_K_FFT_XI);
  hipLaunchKernelGGL((k_fft_xi<real, K>), dim3(nb[4]), dim3(256), 0, c->stream, a);
  prof_end(c, MSG_K_FFT_XI);
}

template <typename real>
int launch_fft_dispatch(Ctx* c) {
  if (c->mesh.gen) {
    launch_gfft_all<real>(c);
    return MSG_OK;
  }
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

SpreadArgs spread_args(Ctx* c) { return SpreadArgs{c->mesh, c->d_brick_start, c->d_bq, c->d_bmm, c->d_Q}; }

GatherArgs gather_args(Ctx* c) {
  GatherArgs g;
  g.m = c->mesh;
  g.np = c->np;
  g.brick_ip = c->d_brick_ip;
  g.bq = c->d_bq;
  g.bmm = c->d_bmm;
  g.phi = c->d_Q;
  g.epsi4pi = c->P.epsi4pi;
  const double vol = c->P.boxlen[0] * c->P.boxlen[1] * c->P.boxlen[2];
  g.surf2fac = c->P.lsurf ? 2.0 * (2.0 * M_PI / (3.0 * vol)) : 0.0;
  g.Mtot = c->d_M;
  g.force = c->d_force;
  g.scale = c->P.force_scale;
  return g;
}

template <typename real, int P>
void launch_spread_gather(Ctx* c, bool spread, bool with_reduce) {
  if (spread) {
    const SpreadArgs a = spread_args(c);
    prof_begin(c, MSG_K_SPREAD);
    // K = 16 (16-point bricks) always has classes (at most 4 bricks)
    if (c->mesh.bx == 16)
      hipLaunchKernelGGL((k_spread<real, P, 16, true>), dim3(c->mesh.nbrick), dim3(256), 0, c->stream, a);
    else if (c->mesh.ncls == 8)
      hipLaunchKernelGGL((k_spread<real, P, 32, true>), dim3(c->mesh.nbrick), dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_spread<real, P, 32, false>), dim3(c->mesh.nbrick), dim3(256), 0, c->stream, a);
    prof_end(c, MSG_K_SPREAD);
  } else {
    const GatherArgs g = gather_args(c);
    const int nb = (8 * c->np + 255) / 256;
    prof_begin(c, MSG_K_GATHER);
    if (with_reduce)
      hipLaunchKernelGGL((k_gather_reduce<real, P>), dim3(nb + 1), dim3(256), 0, c->stream, g,
                         reduce_args(c, true));
    else
      hipLaunchKernelGGL((k_gather<real, P>), dim3(nb), dim3(256), 0, c->stream, g);
    prof_end(c, MSG_K_GATHER);
  }
}

template <typename real, int P>
void launch_spread_gather_g(Ctx* c, bool spread, bool with_reduce) {
  if (spread) {
    const SpreadArgs a = spread_args(c);
    prof_begin(c, MSG_K_SPREAD);
    hipLaunchKernelGGL((k_spread_g<real, P>), dim3(c->mesh.nbrick), dim3(256), spread_g_lds(c->mesh), c->stream, a);
    prof_end(c, MSG_K_SPREAD);
  } else {
    const GatherArgs g = gather_args(c);
    const int nb = (gather_lanes<real, P>() * c->np + 255) / 256;
    prof_begin(c, MSG_K_GATHER);
    if (with_reduce)
      hipLaunchKernelGGL((k_gather_reduce_g<real, P>), dim3(nb + 1), dim3(256), 0, c->stream, g,
                         reduce_args(c, true));
    else
      hipLaunchKernelGGL((k_gather_g<real, P>), dim3(nb), dim3(256), 0, c->stream, g);
    prof_end(c, MSG_K_GATHER);
  }
}

template <typename real>
int spread_gather_dispatch(Ctx* c, bool spread, bool with_reduce = false) {
  if (c->mesh.gen) {
    switch (c->mesh.order) {
#define MSG_SG_G(P) case P: launch_spread_gather_g<real, P>(c, spread, with_reduce); break;
      MSG_SG_G(2) MSG_SG_G(3) MSG_SG_G(4) MSG_SG_G(5) MSG_SG_G(6) MSG_SG_G(7) MSG_SG_G(8) MSG_SG_G(9)
      MSG_SG_G(10) MSG_SG_G(11)
#undef MSG_SG_G
      default: return set_error(MSG_ERR_ARG, "SPME order must be 2..11");
    }
    return MSG_OK;
  }
  switch (c->mesh.order) {
    case 3: launch_spread_gather<real, 3>(c, spread, with_reduce); break;
    case 4: launch_spread_gather<real, 4>(c, spread, with_reduce); break;
    case 5: launch_spread_gather<real, 5>(c, spread, with_reduce); break;
    case 6: launch_spread_gather<real, 6>(c, spread, with_reduce); break;
    case 7: launch_spread_gather<real, 7>(c, spread, with_reduce); break;
    case 8: launch_spread_gather<real, 8>(c, spread, with_reduce); break;
    default: return set_error(MSG_ERR_ARG, "SPME order must be 3..8");
  }
  return MSG_OK;
}

}  // namespace

// EwaldSetup, SPME branch (Src/energy.F90:6969-7018): B-spline moduli b^2(m) and the
// influence function meshfac over [0..K/2]^3, computed in FP64 on the host.
// 30 - e for the smallest e >= 0 with 2^e >= max |az|: a charge times three B-spline weights
// (each <= 1) scaled by 2^qshift is below 2^30 in magnitude.
void spme_charge_scale(Ctx* c) {
  double qmax = 0.0;
  for (double q : c->h_az) qmax = std::max(qmax, std::fabs(q));
  int e = 0;
  while (e < 60 && std::ldexp(1.0, e) < qmax) ++e;
  c->mesh.qshift = 30 - e;
}

int spme_setup(Ctx* c) {
  const msg_params& P = c->P;
  const int K = P.nmesh, p = P.order;
  if (K % 2) return set_error(MSG_ERR_ODD_MESH, "nmesh must be even");
  if (p < kMinOrder || p > kMaxOrder) return set_error(MSG_ERR_ARG, "SPME order must be 2..11");
  if (K < p || K < 4 || K > 1024)
    return set_error(MSG_ERR_MESH, "nmesh must be at least the order (and 4) and at most 1024");
  MeshGrid& m = c->mesh;
  m.K = K;
  m.H = K / 2 + 1;
  m.order = p;
  const bool pow2 = K >= 16 && K <= 256 && (K & (K - 1)) == 0;
  m.gen = !(pow2 && p >= 3 && p <= 8);
  if (!m.gen) {
    m.nb[0] = K / brick_x(K);
    m.nb[1] = K / kBY;
    m.nb[2] = K / kBZ;
    if (!c->fp64 && K <= 128) {
      // the fused plane passes take more dynamic LDS than the default limit
      hipError_t e = hipSuccess;
      switch (K) {
#define MSG_PLANE_ATTR(KK)                                                                          \
  case KK:                                                                                        \
    e = hipFuncSetAttribute((const void*)k_fft_xy<KK>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)PlaneCfg<KK>::lds);                                                \
    if (e == hipSuccess)                                                                          \
      e = hipFuncSetAttribute((const void*)k_fft_yxi<KK>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                              (int)PlaneCfg<KK>::lds);                                              \
    break;
        MSG_PLANE_ATTR(16)
        MSG_PLANE_ATTR(32)
        MSG_PLANE_ATTR(64)
        MSG_PLANE_ATTR(128)
#undef MSG_PLANE_ATTR
      }
      if (e != hipSuccess) return set_error(MSG_ERR_HIP, "hipFuncSetAttribute (plane FFT LDS)");
    }
  } else {
    // bricks of at least max(8, p - 1) points, so a stencil reaches only the next brick
    const int bmin = std::max(8, p - 1);
    for (int d = 0; d < 3; ++d) m.nb[d] = std::max(1, K / bmin);
    const size_t lds = spread_g_lds(m) + (p <= 8 ? 128 : 64) * (3 * p * (c->fp64 ? 8 : 4) + 8 + 2 * p) + 256;
    if (lds > 65536) return set_error(MSG_ERR_MESH, "nmesh too small for this SPME order (one brick spans the mesh)");
    if (K > (c->fp64 ? FftCfg<double>::E : FftCfg<float>::E) - 1)
      return set_error(MSG_ERR_MESH, "nmesh too large for the general FFT");
    // Stockham radices: fours, a two, then the odd prime factors
    int rest = K, n = 0;
    while (rest % 4 == 0 && n < kMaxRadices) { c->rad[n++] = 4; rest /= 4; }
    while (rest % 2 == 0 && n < kMaxRadices) { c->rad[n++] = 2; rest /= 2; }
    for (int f = 3; rest > 1 && n < kMaxRadices; f += 2)
      while (rest % f == 0 && n < kMaxRadices) { c->rad[n++] = f; rest /= f; }
    if (rest != 1) return set_error(MSG_ERR_MESH, "nmesh has too many factors for the general FFT");
    c->nrad = n;
  }
  m.nbrick = m.nb[0] * m.nb[1] * m.nb[2];
  m.bx = m.gen ? 0 : brick_x(K);
  m.ncls = !m.gen && brick_classes(m.nbrick) ? 8 : 1;
  m.nkey = m.nbrick * m.ncls;
  for (int d = 0; d < 3; ++d) m.dkdr[d] = K / P.boxlen[d];
  spme_charge_scale(c);
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
    if (int e = A((void**)&c->d_brick_count, m.nkey * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_start, (m.nkey + 1) * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_fill, m.nkey * sizeof(int))) return e;
    if (int e = A((void**)&c->d_brick_ip, np * sizeof(int))) return e;
    if (int e = A((void**)&c->d_mm, np * sizeof(int))) return e;
    if (int e = A(&c->d_bq, (size_t)np * 4 * rs)) return e;
    if (int e = A((void**)&c->d_bmm, np * sizeof(int))) return e;
    if (int e = A(&c->d_frac, (size_t)np * 3 * rs)) return e;
  }
  MSG_HIP(hipMemset(c->d_brick_count, 0, m.nkey * sizeof(int)));
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
  if (m.gen) {
    const int T = c->fp64 ? gen_lines_yz<double>(K) : gen_lines_yz<float>(K);
    c->nzblocks = (H * K + T - 1) / T;  // workgroups of the general z pass
  } else {
    const int T = kLineElems / K;
    c->nzblocks = (H * K + T - 1) / T;  // workgroups of the z pass (HcLines<K, T>::NB)
  }
  return MSG_OK;
}

int launch_spme_mesh(Ctx* c) {
  int e;
  if (c->fp64) {
    if ((e = spread_gather_dispatch<double>(c, true))) return e;
    if ((e = launch_fft_dispatch<double>(c))) return e;
  } else {
    if ((e = spread_gather_dispatch<float>(c, true))) return e;
    if ((e = launch_fft_dispatch<float>(c))) return e;
  }
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

int launch_spme_gather(Ctx* c, bool with_reduce) {
  int e = c->fp64 ? spread_gather_dispatch<double>(c, false, with_reduce)
                  : spread_gather_dispatch<float>(c, false, with_reduce);
  if (e) return e;
  MSG_HIP(hipGetLastError());
  return MSG_OK;
}

}  // namespace msg

namespace msg {

namespace {

template <typename real, int K>
void seg_fft(const SegGroup& g, hipStream_t s, int stage) {
  const FftArgs* a = (const FftArgs*)g.d_args;
  switch (stage) {
    case 0: hipLaunchKernelGGL((k_fft_x_seg<real, K>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 1: hipLaunchKernelGGL((k_fft_y_seg<real, K>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 2: hipLaunchKernelGGL((k_fft_zconv_seg<real, K>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 3: hipLaunchKernelGGL((k_fft_yi_seg<real, K>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    default: hipLaunchKernelGGL((k_fft_xi_seg<real, K>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
  }
}
template <typename real, int K, int STAGE>
void seg_fft_stage(const SegGroup& g, hipStream_t s) { seg_fft<real, K>(g, s, STAGE); }

template <typename real, int K>
void add_fft(Ctx* c, SegBuilder<FftArgs>* b, int key) {
  int nb[5];
  fft_blocks<real, K>(nb);
  const FftArgs a = fft_args(c);
  b[0].add(key, seg_fft_stage<real, K, 0>, a, nb[0]);
  b[1].add(key, seg_fft_stage<real, K, 1>, a, nb[1]);
  b[2].add(key, seg_fft_stage<real, K, 2>, a, nb[2]);
  b[3].add(key, seg_fft_stage<real, K, 3>, a, nb[3]);
  b[4].add(key, seg_fft_stage<real, K, 4>, a, nb[4]);
}

template <typename real, int STAGE>
void seg_gfft_stage(const SegGroup& g, hipStream_t s) {
  const FftArgs* a = (const FftArgs*)g.d_args;
  switch (STAGE) {
    case 0: hipLaunchKernelGGL((k_gfft_x_seg<real>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 1: hipLaunchKernelGGL((k_gfft_y_seg<real>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 2: hipLaunchKernelGGL((k_gfft_zconv_seg<real>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    case 3: hipLaunchKernelGGL((k_gfft_yi_seg<real>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
    default: hipLaunchKernelGGL((k_gfft_xi_seg<real>), dim3(g.nblk), dim3(256), 0, s, a, g.d_map); break;
  }
}

template <typename real>
int add_fft_k(Ctx* c, SegBuilder<FftArgs>* b) {
  if (c->mesh.gen) {  // one launch per stage covers every general-path mesh of this precision
    int nb[5];
    gfft_blocks<real>(c->mesh.K, nb);
    const FftArgs a = fft_args(c);
    const int key = sizeof(real) == 8 ? 2001 : 2000;
    b[0].add(key, seg_gfft_stage<real, 0>, a, nb[0]);
    b[1].add(key, seg_gfft_stage<real, 1>, a, nb[1]);
    b[2].add(key, seg_gfft_stage<real, 2>, a, nb[2]);
    b[3].add(key, seg_gfft_stage<real, 3>, a, nb[3]);
    b[4].add(key, seg_gfft_stage<real, 4>, a, nb[4]);
    return MSG_OK;
  }
  const int key = (sizeof(real) == 8 ? 1000 : 0) + c->mesh.K;
  switch (c->mesh.K) {
    case 16: add_fft<real, 16>(c, b, key); break;
    case 32: add_fft<real, 32>(c, b, key); break;
    case 64: add_fft<real, 64>(c, b, key); break;
    case 128: add_fft<real, 128>(c, b, key); break;
    case 256: add_fft<real, 256>(c, b, key); break;
    default: return set_error(MSG_ERR_MESH, "nmesh must be 16, 32, 64, 128 or 256");
  }
  return MSG_OK;
}

template <typename real, int P, int BX, bool CLS>
void seg_spread(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_spread_seg<real, P, BX, CLS>), dim3(g.nblk), dim3(256), 0, s, (const SpreadArgs*)g.d_args,
                     g.d_map);
}
template <typename real, int P>
void seg_gather(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_gather_seg<real, P>), dim3(g.nblk), dim3(256), 0, s, (const GatherArgs*)g.d_args, g.d_map);
}

template <typename real, int P>
void seg_spread_g(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_spread_g_seg<real, P>), dim3(g.nblk), dim3(256), g.lds, s, (const SpreadArgs*)g.d_args,
                     g.d_map);
}
template <typename real, int P>
void seg_gather_g(const SegGroup& g, hipStream_t s) {
  hipLaunchKernelGGL((k_gather_g_seg<real, P>), dim3(g.nblk), dim3(256), 0, s, (const GatherArgs*)g.d_args, g.d_map);
}

template <typename real>
int add_spread_gather_g(Ctx* c, SegBuilder<SpreadArgs>& bs, SegBuilder<GatherArgs>& bg) {
  const int key = (sizeof(real) == 8 ? 300 : 200) + c->mesh.order;
  SegLaunchFn fs, fg;
  int lanes = 8;
  switch (c->mesh.order) {
#define MSG_SGG(P)               \
  case P:                        \
    fs = seg_spread_g<real, P>;  \
    fg = seg_gather_g<real, P>;  \
    lanes = gather_lanes<real, P>(); \
    break;
    MSG_SGG(2) MSG_SGG(3) MSG_SGG(4) MSG_SGG(5) MSG_SGG(6) MSG_SGG(7) MSG_SGG(8) MSG_SGG(9) MSG_SGG(10) MSG_SGG(11)
#undef MSG_SGG
    default: return set_error(MSG_ERR_ARG, "SPME order must be 2..11");
  }
  bs.add(key, fs, spread_args(c), c->mesh.nbrick, spread_g_lds(c->mesh));
  bg.add(key, fg, gather_args(c), (lanes * c->np + 255) / 256);
  return MSG_OK;
}

template <typename real>
int add_spread_gather(Ctx* c, SegBuilder<SpreadArgs>& bs, SegBuilder<GatherArgs>& bg) {
  if (c->mesh.gen) return add_spread_gather_g<real>(c, bs, bg);
  const bool wide = c->mesh.bx == 32, cls = c->mesh.ncls == 8;
  // the spread depends on the brick shape and classes, the gather does not
  const int key = (sizeof(real) == 8 ? 100 : 0) + c->mesh.order;
  const int skey = key + (wide ? 1000 : 0) + (cls ? 2000 : 0);
  SegLaunchFn fs, fg;
#define MSG_SG(P)                                                                             \
  fs = !wide ? seg_spread<real, P, 16, true>                                                 \
             : (cls ? seg_spread<real, P, 32, true> : seg_spread<real, P, 32, false>); \
  fg = seg_gather<real, P>;
  switch (c->mesh.order) {
    case 3: MSG_SG(3) break;
    case 4: MSG_SG(4) break;
    case 5: MSG_SG(5) break;
    case 6: MSG_SG(6) break;
    case 7: MSG_SG(7) break;
    case 8: MSG_SG(8) break;
    default: return set_error(MSG_ERR_ARG, "SPME order must be 3..8");
  }
#undef MSG_SG
  bs.add(skey, fs, spread_args(c), c->mesh.nbrick);
  bg.add(key, fg, gather_args(c), (8 * c->np + 255) / 256);
  return MSG_OK;
}

}  // namespace

// Mesh stages of a batch: one launch per (precision, order) for spread and gather and per
// (precision, K) for each FFT stage.  Systems without SPME are skipped.
int plan_spme(const std::vector<Ctx*>& cs, SegOwner& own, SegStage& spread, SegStage* fft,
              SegStage& gather) {
  SegBuilder<SpreadArgs> bs;
  SegBuilder<GatherArgs> bg;
  SegBuilder<FftArgs> bf[5];
  for (Ctx* c : cs) {
    if (c->P.ewald_mode != MSG_EWALD_SPM) continue;
    int e = c->fp64 ? add_spread_gather<double>(c, bs, bg) : add_spread_gather<float>(c, bs, bg);
    if (e) return e;
    e = c->fp64 ? add_fft_k<double>(c, bf) : add_fft_k<float>(c, bf);
    if (e) return e;
  }
  if (int e = bs.finish(own, spread)) return e;
  for (int k = 0; k < 5; ++k)
    if (int e = bf[k].finish(own, fft[k])) return e;
  return bg.finish(own, gather);
}

}  // namespace msg
