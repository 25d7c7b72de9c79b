// ORACLE -- test infrastructure only, never part of the product path.
//
// FP64 CPU restatement of MOLSIM's non-bonded energy/force path, written from the
// semantics of the reference (no reference source is copied).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and
// only as the checker.
//
// Functions and the reference routine each one restates:
//   orc_bspline        Src/mollib.F90:1629-1683   CardinalBSpline
//   orc_random         Src/mollib.F90:2689-2713   Random (Marsaglia xorshift + Park-Miller)
//   orc_heapsort_index Src/mollib.F90:1749-1796   HeapSortIndex
//   orc_pot_eval       Src/potential.F90:1266-1425,2938-2960  PotDefault, Pot_1_6_12, CalcEwaldUfac
//   orc_build_table    Src/potential.F90:868-1258 PotTwoBodyTab1/2/3, SetUBuffer, CheckUBuffer
//   orc_cell_init      Src/celllist.F90:76-265    InitCellList (grid + stencil order)
//   orc_cell_build     Src/celllist.F90:295-327,384-410  pcellro, AddIpToCell, SetCellList
//   orc_utwob_cell     Src/energy.F90:669-777     UTwoBodyACellList
//   orc_utwob_all      Src/energy.F90:445-544     UTwoBodyA over all pairs (j > i)
//   orc_ewald_std_*    Src/energy.F90:964-1102,1401-1426,6736-6760  UEwaldRecStd, EwaldSetArray, kfac
//   orc_spme_*         Src/energy.F90:1207-1297,6525-6587,6855-7019 UEwaldRecSPM, SPMFFTRec, EwaldSetup(spm)
//   orc_ewald_self/surf Src/energy.F90:1301-1371 UEwaldSelf (atomic part), UEwaldSurf
//   orc_utotal         Src/energy.F90:280-434     UTotal, charge/monoatomic branch
//   orc_dutotal        Src/denergy.F90:98-279,728-928,1240-1543  DUTotal single-particle move
//   orc_dutotal_charge Src/mc.F90:3683-3800, Src/denergy.F90:154-164,1552-1699,3911-4145
//                      ChargeChange trial with weak charges (DUWeakChargeEwald, UWeakChargeANew/AOld)
//   orc_fft3_*         FFTW r2c/c2r conventions (unnormalised, forward e^{-i}), own radix-2 code
//   orc_vlist          Src/nlist.F90:724-759,820-884  SetVListMD (parity half list), SetVListMC
//   orc_vlist_cells    the same rows from a cell list (large systems)
//   orc_nlist_skin     Src/nlist.F90:597-621      NList skin test (inlist = 0), r2max1/r2max2 order
//   orc_celllist_set   Src/celllist.F90:307-327,384-410  SetCellList with ipnext/ipprev
//   orc_celllist_update Src/celllist.F90:329-382  UpdateCellIp = RmIpFromCell + AddIpToCell
//   orc_trials_*       Src/denergy.F90:1275-1335,1413-1539,1708-1731  DUTotal with resident
//                      structure factors / charge mesh / dipole (incremental trial energies)
//
// Conventions: positions ro[3*i+d] (Fortran ro(3,np)), particle types 0-based, type pair
// index tp(i,j) = MOLSIM iptpt(i+1,j+1)-1, u_twob[0] = sum, u_twob[1+tp] per pair.
// Meshes are x-fastest: Q[x + K*(y + K*z)], FQ[kx + (K/2+1)*(ky + K*kz)].

#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> cplx;

static const double kPi = 3.14159265358979323846;

extern "C" {

// ----------------------------------------------------------------------------------
// Errors (the reference calls Stop(); the oracle returns these codes)
enum {
  ORC_OK = 0,
  ORC_ERR_ARG = -1,
  ORC_ERR_R2UMIN = -2,     // r2 < r2umin: below the table (energy.F90:710-711)
  ORC_ERR_IBUF = -3,       // ibuf > nbuf (energy.F90:717)
  ORC_ERR_TABLE = -4,      // table did not converge (potential.F90:1117,1127)
  ORC_ERR_TABLE_MEM = -5,  // table needs more space than given
  ORC_ERR_ODD_MESH = -6,   // nmesh must be even (energy.F90:6891)
};

// ----------------------------------------------------------------------------------
// CardinalBSpline, Src/mollib.F90:1629-1683 (Schoenberg recursion).
// vals[i] = M_n(x+i), ders[i] = dM_n/dx(x+i), x = pos - floor(pos); *kmax = floor(pos).
void orc_bspline(int n, double pos, int* kmax, double* vals, double* ders) {
  int v = (int)std::floor(pos);
  double x = pos - v;
  *kmax = v;
  std::vector<double> last(n, 0.0);
  vals[0] = x;
  if (n > 1) {
    vals[1] = 1.0 - x;
    for (int m = 3; m <= n; ++m) {
      for (int i = 0; i < n; ++i) last[i] = vals[i];
      vals[0] = x * last[0] / (m - 1);
      for (int i = 1; i <= m - 2; ++i) {
        double a = i + x;
        vals[i] = (a * last[i] + (m - a) * last[i - 1]) / (m - 1);
      }
      vals[m - 1] = (1.0 - x) * last[m - 2] / (m - 1);
    }
    if (ders) {  // for n == 2 'last' is still all zero, as in the reference
      ders[0] = last[0];
      for (int i = 1; i <= n - 2; ++i) ders[i] = last[i] - last[i - 1];
      ders[n - 1] = -last[n - 2];
    }
  } else if (ders) {
    ders[0] = 0.0;
  }
}

// ----------------------------------------------------------------------------------
// Random, Src/mollib.F90:2689-2713.  state = {ix, iy}; *am is the scale; *idum the seed.
// Threads the oracle's loops use: 1, or OpenMP's count in liboracle_omp.so.
int orc_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double orc_random(int32_t* idum, int32_t* state, double* am) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  int32_t& ix = state[0];
  int32_t& iy = state[1];
  if (*idum <= 0 || iy < 0) {
    // am = nearest(1.0,-1.0)/im is evaluated in default (4-byte) real: 1 - 2^-24 divided by
    // im converted to real(4), which rounds to 2^31; the quotient (1 - 2^-24) 2^-31 is exact
    *am = (double)(1.0f - 5.9604644775390625e-08f) * 0x1p-31;
    int32_t a = *idum < 0 ? -*idum : *idum;
    iy = ((int32_t)888889999 ^ a) | 1;
    ix = (int32_t)777755555 ^ a;
    *idum = a + 1;
  }
  uint32_t ux = (uint32_t)ix;
  ux ^= ux << 13;
  ux ^= ux >> 17;  // ishft with negative count is a logical shift
  ux ^= ux << 5;
  ix = (int32_t)ux;
  int32_t k = iy / iq;
  iy = ia * (iy - k * iq) - ir * k;
  if (iy < 0) iy += im;
  return *am * (double)((im & (ix ^ iy)) | 1);
}

// ----------------------------------------------------------------------------------
// HeapSortIndex, Src/mollib.F90:1749-1796 (Numerical Recipes indexx).  index is 1-based.
void orc_heapsort_index(int n, const double* vec1, int* index1) {
  // 1-based views
  const double* vec = vec1 - 1;
  int* idx = index1 - 1;
  for (int i = 1; i <= n; ++i) idx[i] = i;
  if (n <= 1) return;
  int l = n / 2 + 1, ir = n, indext, i, j;
  double vect;
  for (;;) {
    if (l > 1) {
      --l;
      indext = idx[l];
      vect = vec[indext];
    } else {
      indext = idx[ir];
      vect = vec[indext];
      idx[ir] = idx[1];
      --ir;
      if (ir == 1) {
        idx[1] = indext;
        return;
      }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir && vec[idx[j]] < vec[idx[j + 1]]) ++j;
      if (vect < vec[idx[j]]) {
        idx[i] = idx[j];
        i = j;
        j += j;
      } else {
        j = ir + 1;
      }
    }
    idx[i] = indext;
  }
}

// ----------------------------------------------------------------------------------
// Two-body potential.  form 0 = MOLSIM 'default' with only the 1/r term (ucoff(1) =
// EpsiFourPi*zi*zj, potential.F90:1266-1341); form 1 = '(1,6,12)' (potential.F90:1349-1425).
struct OrcPot {
  int npt, form, lewald, pad;
  double alpha, rcut, epsi4pi;
  double utoltab, ftoltab, umaxtab, fmaxtab, r2uminin;
  double z[8], sig[8], eps[8], rad[8];
};

static inline int tp_index(int npt, int i, int j) {
  if (i > j) std::swap(i, j);
  return i * npt - i * (i - 1) / 2 + (j - i);
}

// CalcEwaldUfac, Src/potential.F90:2938-2960 (erfc via libm; the reference uses its own
// Chebyshev erf accurate to ~1e-16, far below the table tolerance).
static void ewald_ufac(double r, double alpha, double& c0, double& c1, double& c2) {
  double ra = alpha * r;
  double ex = 2.0 / std::sqrt(kPi) * std::exp(-ra * ra);
  double ec = std::erfc(ra);
  c0 = ec;
  c1 = ec + ra * ex;
  c2 = ec + ra * (1.0 + ra * ra) * ex;
}

void orc_pot_eval(const OrcPot* P, int it, int jt, double r, double* u0, double* u1, double* u2) {
  double qq = P->z[it] * P->z[jt] * P->epsi4pi;
  double c0 = 1, c1 = 1, c2 = 1;
  if (P->lewald) ewald_ufac(r, P->alpha, c0, c1, c2);
  double ri = 1.0 / r;
  if (P->form == 0) {
    double w0 = c0 * (qq * ri);
    double w1 = c1 * (-1.0) * (qq * ri);
    double w2 = c2 * (-1.0) * (-2.0) * (qq * ri);
    *u0 = w0;
    *u1 = w1 * ri;
    *u2 = w2 * ri * ri;
  } else {
    double sig = 0.5 * (P->sig[it] + P->sig[jt]);
    double eps = std::sqrt(P->eps[it] * P->eps[jt]);
    double f = std::pow(sig * ri, 6);
    double s0 = 4.0 * eps * (f * f - f);
    double s1 = -24.0 * eps * (2.0 * f * f - f) * ri;
    double s2 = 24.0 * eps * (26.0 * f * f - 7.0 * f) * ri * ri;
    *u0 = qq * c0 * ri + s0;
    *u1 = -qq * c1 * ri * ri + s1;
    *u2 = 2.0 * qq * c2 * ri * ri * ri + s2;
  }
}

// SetUBuffer, Src/potential.F90:1139-1216: quintic in z = r^2 matching w, w', w'' at both ends.
static void set_ubuf(double zlow, double u0l, double u1l, double u2l, double zupp, double u0u,
                     double u1u, double u2u, double* t) {
  if (u0l == 0.0 && u0u == 0.0) {
    t[0] = zlow;
    for (int k = 1; k < 12; ++k) t[k] = 0.0;
    return;
  }
  double rl = std::sqrt(zlow), ru = std::sqrt(zupp);
  double d1 = zupp - zlow, d2 = d1 * d1, d3 = d2 * d1;
  double w0l = u0l, w0u = u0u;
  double w1l = 0.5 * u1l / rl, w1u = 0.5 * u1u / ru;
  double w2l = 0.25 * (u2l - u1l / rl) / zlow, w2u = 0.25 * (u2u - u1u / ru) / zupp;
  double a = 6.0 * (w0u - w0l - w1l * d1 - 0.5 * w2l * d2) / d3;
  double b = 2.0 * (w1u - w1l - w2l * d1) / d2;
  double c = (w2u - w2l) / d1;
  t[0] = zlow;
  t[1] = w0l;
  t[2] = w1l;
  t[3] = 0.5 * w2l;
  t[4] = (10.0 * a - 12.0 * b + 3.0 * c) / 6.0;
  t[5] = (-15.0 * a + 21.0 * b - 6.0 * c) / (6.0 * d1);
  t[6] = (2.0 * a - 3.0 * b + c) / (2.0 * d2);
  t[7] = -2.0 * t[2];
  t[8] = -4.0 * t[3];
  t[9] = -6.0 * t[4];
  t[10] = -8.0 * t[5];
  t[11] = -10.0 * t[6];
}

// CheckUBuffer, Src/potential.F90:1224-1258: 11 equispaced r checks of U and F.
static void check_ubuf(const OrcPot* P, int it, int jt, double rl, double ru, const double* t,
                       bool& ok, bool& repul) {
  ok = false;
  repul = false;
  const int nchk = 11;
  double dr = (ru - rl) / (nchk - 1);
  for (int k = 0; k < nchk; ++k) {
    double r1 = rl + dr * k;
    double u0, u1, u2;
    orc_pot_eval(P, it, jt, r1, &u0, &u1, &u2);
    double dz = r1 * r1 - rl * rl;
    double us = t[1] + dz * (t[2] + dz * (t[3] + dz * (t[4] + dz * (t[5] + dz * t[6]))));
    double fs = t[7] + dz * (t[8] + dz * (t[9] + dz * (t[10] + dz * t[11])));
    fs = r1 * fs;
    if (std::fabs(us - u0) > P->utoltab) return;
    if (std::fabs(fs + u1) > P->ftoltab) return;
    if (std::fabs(us) > P->umaxtab) repul = true;
    if (std::fabs(fs) > P->fmaxtab) repul = true;
  }
  ok = true;
}

// PotTwoBodyTab1/2/3, Src/potential.F90:868-1131 (Andrea-Swope-Andersen intervals from
// rumax = 1.1*rcut inwards; dr shrinks x0.9 until the interval passes CheckUBuffer).
// Outputs per type pair tp: iubuflow[tp] (0-based offset into ubuf), nugrid[tp], r2umin[tp],
// r2atat[tp] = (rad_i+rad_j)^2.  Returns the number of doubles used, or an error code.
int orc_build_table(const OrcPot* P, double* ubuf, int maxbuf, int* iubuflow, int* nugrid,
                    double* r2umin, double* r2atat) {
  const int mnugrid = 1200, ndr = 100;
  const double drfac = 0.9;
  int ibuf = 0;
  for (int it = 0; it < P->npt; ++it) {
    for (int jt = it; jt < P->npt; ++jt) {
      int tp = tp_index(P->npt, it, jt);
      double rumax = std::sqrt((1.1 * P->rcut) * (1.1 * P->rcut));
      double rumin = std::sqrt(P->r2uminin);
      r2atat[tp] = (P->rad[it] + P->rad[jt]) * (P->rad[it] + P->rad[jt]);
      iubuflow[tp] = ibuf;
      double rupp = rumax;
      double dr = 0.5 * (rumax - rumin);
      int igrid;
      for (igrid = 1; igrid <= mnugrid; ++igrid) {
        double zupp = rupp * rupp;
        double rlow = 0, t[12];
        bool ok = false, repul = false;
        int idr;
        for (idr = 1; idr <= ndr; ++idr) {
          rlow = std::max(rumin, rupp - dr);
          double zlow = rlow * rlow;
          double a0, a1, a2, b0, b1, b2;
          orc_pot_eval(P, it, jt, rlow, &a0, &a1, &a2);
          orc_pot_eval(P, it, jt, rupp, &b0, &b1, &b2);
          set_ubuf(zlow, a0, a1, a2, zupp, b0, b1, b2, t);
          check_ubuf(P, it, jt, rlow, rupp, t, ok, repul);
          if (ok) {
            rupp = rlow;
            break;
          }
          dr *= drfac;
        }
        if (!ok) return ORC_ERR_TABLE;
        if (ibuf + 12 > maxbuf) return ORC_ERR_TABLE_MEM;
        std::memcpy(ubuf + ibuf, t, sizeof(t));
        ibuf += 12;
        if (repul) rumin = rlow;
        if (rlow <= rumin) break;
        if (repul) break;
      }
      if (igrid > mnugrid) return ORC_ERR_TABLE;
      nugrid[tp] = igrid;
      r2umin[tp] = rumin * rumin;
    }
  }
  return ibuf;
}

// Table lookup as in UTwoBodyACellList (energy.F90:707-728).  Returns 0, or an error code.
static inline int table_uf(const double* ubuf, int nbuf, int low, double r2, double r2at,
                           double r2min, double& us, double& fs) {
  if (r2 < r2at) {
    us = 1e10;
    fs = 1e10;
    return 0;
  }
  if (r2 < r2min) return ORC_ERR_R2UMIN;
  int ib = low;
  while (r2 < ubuf[ib]) {
    ib += 12;
    if (ib >= nbuf) return ORC_ERR_IBUF;
  }
  double d = r2 - ubuf[ib];
  const double* t = ubuf + ib;
  us = t[1] + d * (t[2] + d * (t[3] + d * (t[4] + d * (t[5] + d * t[6]))));
  fs = t[7] + d * (t[8] + d * (t[9] + d * (t[10] + d * t[11])));
  return 0;
}

// PBCr2 for a box-like cell, Src/molaux.F90:149-184: one shift by the box length.
static inline void pbc_r2(const double* box, double* d, double& r2) {
  for (int k = 0; k < 3; ++k) {
    if (std::fabs(d[k]) > 0.5 * box[k]) d[k] -= std::copysign(box[k], d[k]);
  }
  r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// ----------------------------------------------------------------------------------
// InitCellList, Src/celllist.F90:76-265.  nc[3] cells; stencil[3*nst] offsets in the
// reference's order (heap sort by |dx|+|dy|+|dz|).  Returns nst.
int orc_cell_init(const double* box, double rcell, double rcut, int* nc, double* csize,
                  int* stencil, int maxst) {
  int range[3];
  for (int k = 0; k < 3; ++k) {
    nc[k] = std::max(1, (int)std::floor(box[k] * (1.0 / rcell)));
    csize[k] = box[k] / nc[k];
    range[k] = (int)std::ceil(rcut * (1.0 / csize[k]));
  }
  for (int k = 0; k < 3; ++k) {
    if (2 * range[k] + 1 > nc[k]) {
      nc[k] = 1;
      range[k] = 0;
      csize[k] = box[k];
    }
  }
  std::vector<int> dirs;
  for (int ix = -range[0]; ix <= range[0]; ++ix)
    for (int iy = -range[1]; iy <= range[1]; ++iy)
      for (int iz = -range[2]; iz <= range[2]; ++iz) {
        int o[3] = {ix, iy, iz};
        double s = 0;
        for (int k = 0; k < 3; ++k) {
          double dd = std::max(0, std::abs(o[k]) - 1) * csize[k];
          s += dd * dd;
        }
        if (s > rcut * rcut) continue;
        dirs.push_back(ix);
        dirs.push_back(iy);
        dirs.push_back(iz);
      }
  int n = (int)dirs.size() / 3;
  if (n > maxst) return ORC_ERR_ARG;
  std::vector<double> key(n);
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i)
    key[i] = std::abs(dirs[3 * i]) + std::abs(dirs[3 * i + 1]) + std::abs(dirs[3 * i + 2]);
  orc_heapsort_index(n, key.data(), idx.data());
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) stencil[3 * i + k] = dirs[3 * (idx[i] - 1) + k];
  return n;
}

static inline int cell_coord(double x, double box, double csi) {
  return (int)std::floor((x + 0.5 * box) * csi);
}

// pcellro + AddIpToCell + SetCellList, Src/celllist.F90:295-327,384-410.  Cell id is
// x-fastest (idCellPos).  head[c] = first particle (-1 none), next[i], npart[c], cellof[i].
// Coordinates at +L/2 exactly would index past the grid in the reference; here they are
// clamped to the last cell.
int orc_cell_build(int np, const double* ro, const double* box, const int* nc, const double* csize,
                   int* head, int* next, int* npart, int* cellof) {
  int ncell = nc[0] * nc[1] * nc[2];
  for (int c = 0; c < ncell; ++c) {
    head[c] = -1;
    npart[c] = 0;
  }
  for (int i = 0; i < np; ++i) {
    int ci[3];
    for (int k = 0; k < 3; ++k) {
      ci[k] = cell_coord(ro[3 * i + k], box[k], 1.0 / csize[k]);
      if (ci[k] < 0 || ci[k] > nc[k]) return ORC_ERR_ARG;
      if (ci[k] == nc[k]) ci[k] = nc[k] - 1;
    }
    int c = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
    next[i] = head[c];
    head[c] = i;
    npart[c]++;
    cellof[i] = c;
  }
  return 0;
}

// Neighbour cells of cell c in stencil order (periodic wrap where the box is periodic).
static int cell_neigh(int c, const int* nc, const int* st, int nst, int* out) {
  int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
  int n = 0;
  for (int s = 0; s < nst; ++s) {
    int x = cx + st[3 * s], y = cy + st[3 * s + 1], z = cz + st[3 * s + 2];
    x = ((x % nc[0]) + nc[0]) % nc[0];
    y = ((y % nc[1]) + nc[1]) % nc[1];
    z = ((z % nc[2]) + nc[2]) % nc[2];
    out[n++] = x + nc[0] * (y + nc[1] * z);
  }
  return n;
}

struct OrcTable {
  const double* ubuf;
  int nbuf;
  const int* iubuflow;
  const double* r2umin;
  const double* r2atat;
};

// UTwoBodyACellList, Src/energy.F90:669-777.  u_twob[1+tp] accumulates, u_twob[0] = sum.
// force[3*np] and *virial are ACCUMULATED into (the caller zeroes them, as UTotal does).
// With weak charges (laz non-null) a pair with an uncharged particle is skipped before
// anything else, as UWeakChargeA does (Src/energy.F90:7197-7201).
int orc_utwob_cell(int np, int npt, const double* ro, const int* ipt, const double* box,
                   double rcut, const double* ubuf, int nbuf, const int* iubuflow,
                   const double* r2umin, const double* r2atat, const int* nc, const double* csize,
                   const int* stencil, int nst, double* u_twob, double* force, double* virial,
                   const int* laz) {
  int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> head(ncell), next(np), npart(ncell), cellof(np);
  int rc = orc_cell_build(np, ro, box, nc, csize, head.data(), next.data(), npart.data(),
                          cellof.data());
  if (rc) return rc;
  int nptpt = npt * (npt + 1) / 2;
  for (int k = 0; k <= nptpt; ++k) u_twob[k] = 0;
  double vir = 0, rcut2 = rcut * rcut;
  // the pairs (i, j > i) of home i, added into f / ut / v
  auto home = [&](int i, int* nb, double* f, double* ut, double& v) -> int {
    int nn = cell_neigh(cellof[i], nc, stencil, nst, nb);
    for (int s = 0; s < nn; ++s) {
      for (int j = head[nb[s]]; j >= 0; j = next[j]) {
        if (j <= i) continue;
        if (laz && (!laz[i] || !laz[j])) continue;
        int tp = tp_index(npt, ipt[i], ipt[j]);
        double d[3] = {ro[3 * i] - ro[3 * j], ro[3 * i + 1] - ro[3 * j + 1],
                       ro[3 * i + 2] - ro[3 * j + 2]};
        double r2;
        pbc_r2(box, d, r2);
        if (!(r2 < rcut2)) continue;
        double us, fs;
        int e = table_uf(ubuf, nbuf, iubuflow[tp], r2, r2atat[tp], r2umin[tp], us, fs);
        if (e) return e;
        ut[1 + tp] += us;
        for (int k = 0; k < 3; ++k) {
          f[3 * i + k] += fs * d[k];
          f[3 * j + k] -= fs * d[k];
        }
        v -= fs * r2;
      }
    }
    return 0;
  };
#ifdef _OPENMP
  // liboracle_omp.so (CPU timing on all host cores only): homes shared out over the threads,
  // each adding into arrays of its own, which are summed at the end
  int err = 0;
#pragma omp parallel
  {
    std::vector<int> nbt(nst);
    std::vector<double> ft(3 * (size_t)np, 0.0), utt(nptpt + 1, 0.0);
    double vt = 0;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < np; ++i) {
      int e = home(i, nbt.data(), ft.data(), utt.data(), vt);
      if (e) {
#pragma omp atomic write
        err = e;
      }
    }
#pragma omp critical
    {
      for (size_t k = 0; k < ft.size(); ++k) force[k] += ft[k];
      for (int k = 1; k <= nptpt; ++k) u_twob[k] += utt[k];
      vir += vt;
    }
  }
  if (err) return err;
#else
  std::vector<int> nb(nst);
  for (int i = 0; i < np; ++i)
    if (int e = home(i, nb.data(), force, u_twob, vir)) return e;
#endif
  double s = 0;
  for (int k = 1; k <= nptpt; ++k) s += u_twob[k];
  u_twob[0] = s;
  *virial += vir;
  return 0;
}

// UTwoBodyA over all pairs j > i (no list), Src/energy.F90:445-544 (r2 > rcut2 skipped;
// hard-core pairs get usum = fsum = 1e10 as in the cell-list routine).
int orc_utwob_all(int np, int npt, const double* ro, const int* ipt, const double* box, double rcut,
                  const double* ubuf, int nbuf, const int* iubuflow, const double* r2umin,
                  const double* r2atat, double* u_twob, double* force, double* virial,
                  const int* laz) {
  int nptpt = npt * (npt + 1) / 2;
  for (int k = 0; k <= nptpt; ++k) u_twob[k] = 0;
  double vir = 0, rcut2 = rcut * rcut;
  for (int i = 0; i < np; ++i)
    for (int j = i + 1; j < np; ++j) {
      if (laz && (!laz[i] || !laz[j])) continue;
      int tp = tp_index(npt, ipt[i], ipt[j]);
      double d[3] = {ro[3 * i] - ro[3 * j], ro[3 * i + 1] - ro[3 * j + 1],
                     ro[3 * i + 2] - ro[3 * j + 2]};
      double r2;
      pbc_r2(box, d, r2);
      if (r2 > rcut2) continue;
      double us, fs;
      int e = table_uf(ubuf, nbuf, iubuflow[tp], r2, r2atat[tp], r2umin[tp], us, fs);
      if (e) return e;
      u_twob[1 + tp] += us;
      for (int k = 0; k < 3; ++k) {
        force[3 * i + k] += fs * d[k];
        force[3 * j + k] -= fs * d[k];
      }
      vir -= fs * r2;
    }
  double s = 0;
  for (int k = 1; k <= nptpt; ++k) s += u_twob[k];
  u_twob[0] = s;
  *virial += vir;
  return 0;
}

// ----------------------------------------------------------------------------------
// Standard Ewald.  kfac over one octant (nx,ny,nz >= 0), Src/energy.F90:6736-6760.
// region 0 = sphere (n^2 <= ncut^2), 1 = cube (n^2 <= 3 ncut^2).  Returns nkvec.
int orc_ewald_std_kfac(const double* box, double alpha, int ncut, int region, double* kfac,
                       int maxk) {
  int ncut2 = region == 0 ? ncut * ncut : 3 * ncut * ncut;
  double vol = box[0] * box[1] * box[2];
  double fac0 = 2 * kPi / vol;
  int nk = 0;
  for (int nz = 0; nz <= ncut; ++nz) {
    double fz = nz == 0 ? 1.0 : 2.0;
    for (int ny = 0; ny <= ncut; ++ny) {
      double fy = ny == 0 ? 0.5 : 1.0;
      for (int nx = 0; nx <= ncut; ++nx) {
        if (nx * nx + ny * ny + nz * nz > ncut2) continue;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double fx = nx == 0 ? 0.5 : 1.0;
        double kx = nx * 2 * kPi / box[0], ky = ny * 2 * kPi / box[1], kz = nz * 2 * kPi / box[2];
        double k2 = kx * kx + ky * ky + kz * kz;
        if (nk >= maxk) return ORC_ERR_ARG;
        kfac[nk++] = fac0 * fz * fy * fx * std::exp(-k2 / (4 * alpha * alpha)) / k2;
      }
    }
  }
  return nk;
}

// UEwaldRecStd (serial branch), Src/energy.F90:1050-1098 with EwaldSetArray :1401-1426.
// Returns urec WITHOUT EpsiFourPi (as u%rec before UEwald's scaling); forces include
// EpsiFourPi (accumulated); *virrec accumulates without EpsiFourPi.  sumeikr (optional,
// nkvec*4 complex as interleaved doubles) receives the structure factors.
int orc_ewald_std_rec(int na, const double* r, const double* az, const double* box, double alpha,
                      int ncut, int region, double epsi4pi, const double* kfac, double* urec,
                      double* force, double* virrec, double* sumeikr_out) {
  int ncut2 = region == 0 ? ncut * ncut : 3 * ncut * ncut;
  double tpb[3] = {2 * kPi / box[0], 2 * kPi / box[1], 2 * kPi / box[2]};
  std::vector<cplx> ex(na * (ncut + 1)), ey(na * (ncut + 1)), ez(na * (ncut + 1));
  for (int a = 0; a < na; ++a) {
    ex[a * (ncut + 1)] = ey[a * (ncut + 1)] = ez[a * (ncut + 1)] = cplx(1, 0);
    if (ncut >= 1) {
      ex[a * (ncut + 1) + 1] = cplx(std::cos(tpb[0] * r[3 * a]), std::sin(tpb[0] * r[3 * a]));
      ey[a * (ncut + 1) + 1] =
          cplx(std::cos(tpb[1] * r[3 * a + 1]), std::sin(tpb[1] * r[3 * a + 1]));
      ez[a * (ncut + 1) + 1] =
          cplx(std::cos(tpb[2] * r[3 * a + 2]), std::sin(tpb[2] * r[3 * a + 2]));
    }
    for (int c = 2; c <= ncut; ++c) {
      ex[a * (ncut + 1) + c] = ex[a * (ncut + 1) + c - 1] * ex[a * (ncut + 1) + 1];
      ey[a * (ncut + 1) + c] = ey[a * (ncut + 1) + c - 1] * ey[a * (ncut + 1) + 1];
      ez[a * (ncut + 1) + c] = ez[a * (ncut + 1) + c - 1] * ez[a * (ncut + 1) + 1];
    }
  }
  std::vector<cplx> yzm(na), yzp(na), er(4 * na);
  double u = 0, vir = 0;
  int kn = 0;
  for (int nz = 0; nz <= ncut; ++nz)
    for (int ny = 0; ny <= ncut; ++ny) {
      if (ny * ny + nz * nz > ncut2) continue;
      for (int a = 0; a < na; ++a) {
        yzm[a] = std::conj(ey[a * (ncut + 1) + ny]) * ez[a * (ncut + 1) + nz];
        yzp[a] = ey[a * (ncut + 1) + ny] * ez[a * (ncut + 1) + nz];
      }
      for (int nx = 0; nx <= ncut; ++nx) {
        if (nx * nx + ny * ny + nz * nz > ncut2) continue;
        if (nx == 0 && ny == 0 && nz == 0) continue;
        cplx s[4] = {0, 0, 0, 0};
        for (int a = 0; a < na; ++a) {
          cplx exa = ex[a * (ncut + 1) + nx];
          er[4 * a + 0] = az[a] * std::conj(exa) * yzm[a];
          er[4 * a + 1] = az[a] * std::conj(exa) * yzp[a];
          er[4 * a + 2] = az[a] * exa * yzm[a];
          er[4 * a + 3] = az[a] * exa * yzp[a];
          for (int q = 0; q < 4; ++q) s[q] += er[4 * a + q];
        }
        if (sumeikr_out)
          for (int q = 0; q < 4; ++q) {
            sumeikr_out[2 * (4 * kn + q)] = s[q].real();
            sumeikr_out[2 * (4 * kn + q) + 1] = s[q].imag();
          }
        double sq = 0;
        for (int q = 0; q < 4; ++q) sq += std::norm(s[q]);
        double term = kfac[kn] * sq;
        u += term;
        double kf2 = epsi4pi * 2 * kfac[kn];
        double kx = kf2 * nx * tpb[0], ky = kf2 * ny * tpb[1], kz = kf2 * nz * tpb[2];
        for (int a = 0; a < na; ++a) {
          double f[4];
          for (int q = 0; q < 4; ++q)
            f[q] = -er[4 * a + q].real() * s[q].imag() + er[4 * a + q].imag() * s[q].real();
          force[3 * a + 0] += kx * (-f[0] - f[1] + f[2] + f[3]);
          force[3 * a + 1] += ky * (-f[0] + f[1] - f[2] + f[3]);
          force[3 * a + 2] += kz * (+f[0] + f[1] + f[2] + f[3]);
        }
        double k2 = std::pow(nx * tpb[0], 2) + std::pow(ny * tpb[1], 2) + std::pow(nz * tpb[2], 2);
        vir -= (1.0 - k2 / (2 * alpha * alpha)) * term;
        ++kn;
      }
    }
  *urec = u;
  *virrec += vir;
  return kn;
}

// ----------------------------------------------------------------------------------
// FFT (own code).  Radix-2 for powers of two; otherwise a recursive mixed-radix
// decimation in time (Cooley-Tukey over the prime factors of n, a plain DFT for a prime
// length).  sign -1 = forward.
//
// out[k] = sum_j in[j * is] W^{jk}, W = e^{sign 2 pi i / n}; tw holds e^{sign 2 pi i e / N}
// for e < N and W^e = tw[(e mod n) * (N / n)].
static void fft_mixed(const cplx* in, int is, cplx* out, int n, int N, const std::vector<cplx>& tw) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int step = N / n;
  int p = 2;
  while (n % p) ++p;
  if (p == n) {
    for (int k = 0; k < n; ++k) {
      cplx s = 0;
      for (int j = 0; j < n; ++j) s += in[(size_t)j * is] * tw[(size_t)((long)j * k % n) * step];
      out[k] = s;
    }
    return;
  }
  const int m = n / p;
  // sub-sequence r (in[j p + r], j < m) transformed into out[r m .. r m + m)
  for (int r = 0; r < p; ++r) fft_mixed(in + (size_t)r * is, is * p, out + (size_t)r * m, m, N, tw);
  std::vector<cplx> comb(n);
  for (int k = 0; k < m; ++k)
    for (int s = 0; s < p; ++s) {
      const int kk = k + m * s;
      cplx v = 0;
      for (int r = 0; r < p; ++r) v += tw[(size_t)((long)r * kk % n) * step] * out[(size_t)r * m + k];
      comb[kk] = v;
    }
  for (int k = 0; k < n; ++k) out[k] = comb[k];
}

static void fft1(cplx* a, int n, int stride, int sign, std::vector<cplx>& buf) {
  buf.resize(n);
  for (int i = 0; i < n; ++i) buf[i] = a[(size_t)i * stride];
  if ((n & (n - 1)) == 0) {
    // bit reversal
    for (int i = 1, j = 0; i < n; ++i) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(buf[i], buf[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      for (int k = 0; k < len / 2; ++k) {
        double ang = sign * 2 * kPi * k / len;
        cplx w(std::cos(ang), std::sin(ang));
        for (int i = 0; i < n; i += len) {
          cplx u = buf[i + k], v = buf[i + k + len / 2] * w;
          buf[i + k] = u + v;
          buf[i + k + len / 2] = u - v;
        }
      }
    }
    for (int i = 0; i < n; ++i) a[(size_t)i * stride] = buf[i];
  } else {
    std::vector<cplx> tw(n), out(n);
    for (int e = 0; e < n; ++e) {
      const double ang = sign * 2 * kPi * (double)e / n;
      tw[e] = cplx(std::cos(ang), std::sin(ang));
    }
    fft_mixed(buf.data(), 1, out.data(), n, n, tw);
    for (int k = 0; k < n; ++k) a[(size_t)k * stride] = out[k];
  }
}

// (the line loops are shared out over the threads in liboracle_omp.so; each line is
// transformed the same way either way)
static void fft3_full(cplx* a, int K, int sign) {
  size_t K2 = (size_t)K * K;
#pragma omp parallel
  {
    std::vector<cplx> buf;
#pragma omp for collapse(2)
    for (int z = 0; z < K; ++z)
      for (int y = 0; y < K; ++y) fft1(a + (size_t)K * (y + (size_t)K * z), K, 1, sign, buf);
#pragma omp for collapse(2)
    for (int z = 0; z < K; ++z)
      for (int x = 0; x < K; ++x) fft1(a + x + K2 * z, K, K, sign, buf);
#pragma omp for collapse(2)
    for (int y = 0; y < K; ++y)
      for (int x = 0; x < K; ++x) fft1(a + x + (size_t)K * y, K, (int)K2, sign, buf);
  }
}

// fftw_plan_dft_r2c_3d(K,K,K) semantics on an x-fastest array: out[(K/2+1)*K*K] complex.
void orc_fft3_r2c(int K, const double* in, double* out_c) {
  size_t n = (size_t)K * K * K;
  std::vector<cplx> a(n);
  for (size_t i = 0; i < n; ++i) a[i] = cplx(in[i], 0);
  fft3_full(a.data(), K, -1);
  int H = K / 2 + 1;
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < K; ++y)
      for (int x = 0; x < H; ++x) {
        cplx v = a[x + (size_t)K * (y + (size_t)K * z)];
        size_t o = x + (size_t)H * (y + (size_t)K * z);
        out_c[2 * o] = v.real();
        out_c[2 * o + 1] = v.imag();
      }
}

// fftw_plan_dft_c2r_3d semantics (unnormalised, uses the half spectrum only).
void orc_fft3_c2r(int K, const double* in_c, double* out) {
  size_t n = (size_t)K * K * K;
  int H = K / 2 + 1;
  std::vector<cplx> a(n);
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < K; ++y)
      for (int x = 0; x < K; ++x) {
        cplx v;
        if (x < H) {
          size_t o = x + (size_t)H * (y + (size_t)K * z);
          v = cplx(in_c[2 * o], in_c[2 * o + 1]);
        } else {
          int xm = K - x, ym = (K - y) % K, zm = (K - z) % K;
          size_t o = xm + (size_t)H * (ym + (size_t)K * zm);
          v = std::conj(cplx(in_c[2 * o], in_c[2 * o + 1]));
        }
        a[x + (size_t)K * (y + (size_t)K * z)] = v;
      }
  // c2r treats the x=0 and x=K/2 planes as Hermitian: use only their real-consistent part
  fft3_full(a.data(), K, +1);
  for (size_t i = 0; i < n; ++i) out[i] = a[i].real();
}

// ----------------------------------------------------------------------------------
// EwaldSetup SPME branch, Src/energy.F90:6969-7018.  Arrays (K/2+1)^3, index
// nx + H*(ny + H*nz).  b2 from the order-p spline at pos = 0.
int orc_spme_setup(const double* box, double alpha, int K, int order, double* meshfac,
                   double* energyfac, double* virfac) {
  if (K % 2) return ORC_ERR_ODD_MESH;
  int H = K / 2 + 1;
  std::vector<double> spl(order), b2(3 * K);
  int kmax;
  orc_bspline(order, 0.0, &kmax, spl.data(), nullptr);
  for (int d = 0; d < 3; ++d)
    for (int m = 0; m < K; ++m) {
      if ((order % 2 == 1) && (2 * m == K)) {
        b2[d * K + m] = 0.0;  // left undefined by the reference for odd order; zero here
        continue;
      }
      cplx a = 0;
      for (int k = 0; k <= order - 2; ++k) {
        double ang = 2 * kPi * m * k / K;
        a += spl[k + 1] * cplx(std::cos(ang), std::sin(ang));
      }
      b2[d * K + m] = 1.0 / std::norm(a);
    }
  double vol = box[0] * box[1] * box[2];
  double fac0 = 4 * kPi / vol, expf = -1.0 / (4 * alpha * alpha);
  for (int nz = 0; nz < H; ++nz)
    for (int ny = 0; ny < H; ++ny)
      for (int nx = 0; nx < H; ++nx) {
        size_t o = nx + (size_t)H * (ny + (size_t)H * nz);
        if (nx == 0 && ny == 0 && nz == 0) {
          meshfac[o] = energyfac[o] = virfac[o] = 0;
          continue;
        }
        double m2 = std::pow(nx * 2 * kPi / box[0], 2) + std::pow(ny * 2 * kPi / box[1], 2) +
                    std::pow(nz * 2 * kPi / box[2], 2);
        double c = std::exp(m2 * expf) / m2;
        double mf = fac0 * c * b2[nx] * b2[K + ny] * b2[2 * K + nz];
        meshfac[o] = mf;
        if (nx == (K - nx) % K) mf /= 2;
        if (ny == (K - ny) % K) mf /= 2;
        if (nz == (K - nz) % K) mf /= 2;
        energyfac[o] = mf;
        virfac[o] = (1.0 + 2 * m2 * expf) * mf;
      }
  return 0;
}

// Spreading of UEwaldRecSPM, Src/energy.F90:1226-1245 (Q zeroed first).
void orc_spme_spread(int na, const double* r, const double* az, const double* box, int K,
                     int order, double* Q) {
  size_t n = (size_t)K * K * K;
  for (size_t i = 0; i < n; ++i) Q[i] = 0;
#ifdef _OPENMP
  // liboracle_omp.so (CPU timing on all host cores only): particles shared out over the
  // threads, each spreading into a mesh of its own; the meshes are summed at the end
  int nth = omp_get_max_threads();
  std::vector<std::vector<double>> Qt(nth);
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    Qt[t].assign(n, 0.0);
    double* Qp = Qt[t].data();
    std::vector<double> th(3 * order);
    int mm[3];
#pragma omp for
    for (int a = 0; a < na; ++a) {
      for (int d = 0; d < 3; ++d) {
        int km;
        orc_bspline(order, K / box[d] * r[3 * a + d], &km, &th[d * order], nullptr);
        mm[d] = ((km + K) % K + K) % K;
      }
      for (int iz = 0; iz < order; ++iz) {
        int z = (K + mm[2] - iz) % K;
        double qz = az[a] * th[2 * order + iz];
        for (int iy = 0; iy < order; ++iy) {
          int y = (K + mm[1] - iy) % K;
          double qyz = qz * th[order + iy];
          for (int ix = 0; ix < order; ++ix) {
            int x = (K + mm[0] - ix) % K;
            Qp[x + (size_t)K * (y + (size_t)K * z)] += qyz * th[ix];
          }
        }
      }
    }
#pragma omp for
    for (long long i = 0; i < (long long)n; ++i)
      for (int u = 0; u < nth; ++u) Q[i] += Qt[u][i];
  }
#else
  std::vector<double> th(3 * order);
  int mm[3];
  for (int a = 0; a < na; ++a) {
    for (int d = 0; d < 3; ++d) {
      int km;
      orc_bspline(order, K / box[d] * r[3 * a + d], &km, &th[d * order], nullptr);
      mm[d] = ((km + K) % K + K) % K;
    }
    for (int iz = 0; iz < order; ++iz) {
      int z = (K + mm[2] - iz) % K;
      double qz = az[a] * th[2 * order + iz];
      for (int iy = 0; iy < order; ++iy) {
        int y = (K + mm[1] - iy) % K;
        double qyz = qz * th[order + iy];
        for (int ix = 0; ix < order; ++ix) {
          int x = (K + mm[0] - ix) % K;
          Q[x + (size_t)K * (y + (size_t)K * z)] += qyz * th[ix];
        }
      }
    }
  }
#endif
}

// SPMFFTRec, Src/energy.F90:6525-6587: r2c(Q); E = sum energyfac |F|^2; W -= sum virfac
// |F|^2; F *= meshfac; phi = c2r(F).  Sums over the half spectrum with the weights the
// reference's 4-point mirror loop gives each stored element.  Q is overwritten by phi.
void orc_spme_conv(int K, const double* meshfac, const double* energyfac, const double* virfac,
                   double* Q, double* u, double* vir) {
  int H = K / 2 + 1;
  size_t nc = (size_t)H * K * K;
  std::vector<double> F(2 * nc);
  orc_fft3_r2c(K, Q, F.data());
  double e = 0, w = 0;
  for (int nz = 0; nz <= K / 2; ++nz) {
    int nza = (K - nz) % K;
    for (int ny = 0; ny <= K / 2; ++ny) {
      int nya = (K - ny) % K;
      for (int nx = 0; nx <= K / 2; ++nx) {
        size_t g = nx + (size_t)H * (ny + (size_t)H * nz);
        size_t idx[4] = {nx + (size_t)H * (ny + (size_t)K * nz),
                         nx + (size_t)H * (ny + (size_t)K * nza),
                         nx + (size_t)H * (nya + (size_t)K * nz),
                         nx + (size_t)H * (nya + (size_t)K * nza)};
        cplx v[4];
        double qs = 0;
        for (int q = 0; q < 4; ++q) {
          v[q] = cplx(F[2 * idx[q]], F[2 * idx[q] + 1]);
          qs += std::norm(v[q]);
        }
        e += qs * energyfac[g];
        w -= qs * virfac[g];
        for (int q = 0; q < 4; ++q) {
          cplx m = v[q] * meshfac[g];
          F[2 * idx[q]] = m.real();
          F[2 * idx[q] + 1] = m.imag();
        }
      }
    }
  }
  *u = e;
  *vir += w;
  orc_fft3_c2r(K, F.data(), Q);
}

// Energy of a mesh only (DUTwoBodyEwaldRecSPM's CalcGIF loop, Src/denergy.F90:1476-1489).
double orc_spme_energy(int K, const double* energyfac, const double* Q) {
  int H = K / 2 + 1;
  size_t nc = (size_t)H * K * K;
  std::vector<double> F(2 * nc);
  orc_fft3_r2c(K, Q, F.data());
  double e = 0;
  for (int nz = 0; nz <= K / 2; ++nz) {
    int nza = (K - nz) % K;
    for (int ny = 0; ny <= K / 2; ++ny) {
      int nya = (K - ny) % K;
      for (int nx = 0; nx <= K / 2; ++nx) {
        size_t g = nx + (size_t)H * (ny + (size_t)H * nz);
        size_t idx[4] = {nx + (size_t)H * (ny + (size_t)K * nz),
                         nx + (size_t)H * (ny + (size_t)K * nza),
                         nx + (size_t)H * (nya + (size_t)K * nz),
                         nx + (size_t)H * (nya + (size_t)K * nza)};
        double qs = 0;
        for (int q = 0; q < 4; ++q) qs += F[2 * idx[q]] * F[2 * idx[q]] + F[2 * idx[q] + 1] * F[2 * idx[q] + 1];
        e += qs * energyfac[g];
      }
    }
  }
  return e;
}

// Force gather of UEwaldRecSPM, Src/energy.F90:1258-1286 (phi = convolved mesh).
void orc_spme_gather(int na, const double* r, const double* az, const double* box, int K,
                     int order, double epsi4pi, const double* phi, double* force) {
#pragma omp parallel
  {
  std::vector<double> th(3 * order), dth(3 * order);
  int mm[3];
#pragma omp for
  for (int a = 0; a < na; ++a) {
    for (int d = 0; d < 3; ++d) {
      int km;
      orc_bspline(order, K / box[d] * r[3 * a + d], &km, &th[d * order], &dth[d * order]);
      mm[d] = ((km + K) % K + K) % K;
    }
    double es[3] = {0, 0, 0};
    for (int iz = 0; iz < order; ++iz) {
      int z = (K + mm[2] - iz) % K;
      for (int iy = 0; iy < order; ++iy) {
        int y = (K + mm[1] - iy) % K;
        for (int ix = 0; ix < order; ++ix) {
          int x = (K + mm[0] - ix) % K;
          double q = phi[x + (size_t)K * (y + (size_t)K * z)];
          es[0] += dth[ix] * th[order + iy] * th[2 * order + iz] * q;
          es[1] += th[ix] * dth[order + iy] * th[2 * order + iz] * q;
          es[2] += th[ix] * th[order + iy] * dth[2 * order + iz] * q;
        }
      }
    }
    for (int d = 0; d < 3; ++d) force[3 * a + d] -= az[a] * es[d] * (K / box[d]) * epsi4pi;
  }
  }
}

// ----------------------------------------------------------------------------------
// UEwaldSelf (atomic part) and UEwaldSurf (3D), Src/energy.F90:1301-1371.  Energies are
// returned without EpsiFourPi; surface forces include it and are accumulated.
double orc_ewald_self(int na, const double* az, double alpha) {
  double s = 0;
  for (int a = 0; a < na; ++a) s += az[a] * az[a];
  return -alpha / std::sqrt(kPi) * s;
}

double orc_ewald_surf(int na, const double* r, const double* az, const double* box,
                      double epsi4pi, double* force, double* virrec) {
  double vol = box[0] * box[1] * box[2];
  double fac = 2 * kPi / (3 * vol);
  double m[3] = {0, 0, 0};
  for (int a = 0; a < na; ++a)
    for (int d = 0; d < 3; ++d) m[d] += az[a] * r[3 * a + d];
  double term = fac * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  if (force)
    for (int a = 0; a < na; ++a)
      for (int d = 0; d < 3; ++d) force[3 * a + d] -= epsi4pi * (2 * fac) * az[a] * m[d];
  if (virrec) *virrec -= term;
  return term;
}

// ----------------------------------------------------------------------------------
// Full system description for utotal / dutotal.
struct OrcSys {
  int np, npt;
  const double* ro;
  const int* ipt;
  const double* az;
  double box[3];
  double rcut, drnlist, epsi4pi;
  // table
  const double* ubuf;
  int nbuf;
  const int* iubuflow;
  const double* r2umin;
  const double* r2atat;
  // list: 0 = all pairs (vlist-equivalent), 1 = cell list
  int list;
  // Ewald: 0 none, 1 std, 2 spm
  int ewald;
  int surf;
  double alpha;
  int ncut, region;  // std
  int K, order;      // spm
  // weak (titrating) charges: laz[i] = 1 if particle i carries its charge zat[type]; null
  // when every particle is charged (lweakcharge off)
  const int* laz;
  const double* zat;
};

// UTotal charge/monoatomic branch, Src/energy.F90:280-434 (forces left unscaled, i.e. in
// energy/length units; UTotal's unit scaling is a caller-side multiply).
// Outputs: u_twob[nptpt+1], *u_rec (with self/surf, x EpsiFourPi), *u_tot, force[3np], *virial.
int orc_utotal(const OrcSys* S, double* u_twob, double* u_rec, double* u_tot, double* force,
               double* virial) {
  int np = S->np;
  for (int i = 0; i < 3 * np; ++i) force[i] = 0;
  *virial = 0;
  int rc;
  if (S->list == 1) {
    int nc[3];
    double cs[3];
    std::vector<int> st(3 * 1000);
    int nst = orc_cell_init(S->box, S->rcut + S->drnlist, S->rcut, nc, cs, st.data(), 1000);
    if (nst < 0) return nst;
    rc = orc_utwob_cell(np, S->npt, S->ro, S->ipt, S->box, S->rcut, S->ubuf, S->nbuf, S->iubuflow,
                        S->r2umin, S->r2atat, nc, cs, st.data(), nst, u_twob, force, virial, S->laz);
  } else {
    rc = orc_utwob_all(np, S->npt, S->ro, S->ipt, S->box, S->rcut, S->ubuf, S->nbuf, S->iubuflow,
                       S->r2umin, S->r2atat, u_twob, force, virial, S->laz);
  }
  if (rc) return rc;
  double urec = 0, virrec = 0;
  if (S->ewald == 1) {
    int H = S->ncut + 1;
    std::vector<double> kfac(4 * H * H * H);
    int nk = orc_ewald_std_kfac(S->box, S->alpha, S->ncut, S->region, kfac.data(), (int)kfac.size());
    if (nk < 0) return nk;
    orc_ewald_std_rec(np, S->ro, S->az, S->box, S->alpha, S->ncut, S->region, S->epsi4pi,
                      kfac.data(), &urec, force, &virrec, nullptr);
  } else if (S->ewald == 2) {
    int K = S->K, H = K / 2 + 1;
    size_t g = (size_t)H * H * H;
    std::vector<double> mf(g), ef(g), vf(g), Q((size_t)K * K * K);
    rc = orc_spme_setup(S->box, S->alpha, K, S->order, mf.data(), ef.data(), vf.data());
    if (rc) return rc;
    orc_spme_spread(np, S->ro, S->az, S->box, K, S->order, Q.data());
    orc_spme_conv(K, mf.data(), ef.data(), vf.data(), Q.data(), &urec, &virrec);
    orc_spme_gather(np, S->ro, S->az, S->box, K, S->order, S->epsi4pi, Q.data(), force);
  }
  if (S->ewald) {
    urec += orc_ewald_self(np, S->az, S->alpha);
    if (S->surf) urec += orc_ewald_surf(np, S->ro, S->az, S->box, S->epsi4pi, force, &virrec);
    urec *= S->epsi4pi;
    *virial += S->epsi4pi * virrec;
  }
  *u_rec = urec;
  *u_tot = u_twob[0] + urec;
  return 0;
}

// DUTotal for a single moved particle ip to rnew[3], charge/monoatomic branch:
// Src/denergy.F90:98-279 (DUTwoBody with the cell-list or all-pairs New/Old routines
// :728-928 / :289-495) and DUTwoBodyEwald :1240-1543 (std: sumeikr update; spm: the
// reference's full r2c of the modified mesh).  *overlap = 1 on a hard-core overlap
// (then only the overlap flag is meaningful, as in the reference).
// Cell list of a configuration for the real-space part of DUTotal (InitCellList +
// SetCellList with the list's cut-off rcut + drnlist).
struct CellIndex {
  int nc[3];
  double cs[3];
  std::vector<int> st;
  int nst = 0;
  std::vector<int> head, next, npart, cellof;
};

static int cell_index_build(const OrcSys* S, CellIndex& C) {
  C.st.assign(3 * 1000, 0);
  C.nst = orc_cell_init(S->box, S->rcut + S->drnlist, S->rcut, C.nc, C.cs, C.st.data(), 1000);
  if (C.nst < 0) return C.nst;
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  C.head.assign(ncell, 0);
  C.next.assign(S->np, 0);
  C.npart.assign(ncell, 0);
  C.cellof.assign(S->np, 0);
  return orc_cell_build(S->np, S->ro, S->box, C.nc, C.cs, C.head.data(), C.next.data(), C.npart.data(),
                        C.cellof.data());
}

// Real-space part of DUTotal (UTwoBodyANew / UTwoBodyAOld, Src/denergy.F90:728-928, or their
// all-pairs forms): per type pair, the new position's and the old position's pair energies
// with every other particle.  Returns 0, 1 when the new position overlaps, or an error.
static int dut_real(const OrcSys* S, const CellIndex* C, int ip, const double* rnew, double* unew,
                    double* uold) {
  const int np = S->np, npt = S->npt;
  const double rcut2 = S->rcut * S->rcut;
  if (C) {
    // new: cell of the trial position; old: cell of ip
    int ci[3];
    for (int k = 0; k < 3; ++k) {
      ci[k] = cell_coord(rnew[k], S->box[k], 1.0 / C->cs[k]);
      if (ci[k] == C->nc[k]) ci[k] = C->nc[k] - 1;
      if (ci[k] < 0 || ci[k] >= C->nc[k]) return ORC_ERR_ARG;
    }
    const int cnew = ci[0] + C->nc[0] * (ci[1] + C->nc[1] * ci[2]);
    std::vector<int> nb(C->nst);
    for (int pass = 0; pass < 2; ++pass) {
      const int c = pass == 0 ? cnew : C->cellof[ip];
      const double* ri = pass == 0 ? rnew : S->ro + 3 * ip;
      const int nn = cell_neigh(c, C->nc, C->st.data(), C->nst, nb.data());
      for (int sidx = 0; sidx < nn; ++sidx)
        for (int j = C->head[nb[sidx]]; j >= 0; j = C->next[j]) {
          if (j == ip) continue;
          int tp = tp_index(npt, S->ipt[ip], S->ipt[j]);
          double d[3] = {ri[0] - S->ro[3 * j], ri[1] - S->ro[3 * j + 1], ri[2] - S->ro[3 * j + 2]};
          double r2;
          pbc_r2(S->box, d, r2);
          if (!(r2 < rcut2)) continue;
          if (pass == 0 && (r2 < S->r2atat[tp] || r2 < S->r2umin[tp])) return 1;  // overlap
          // weak charges: UWeakChargeANew/AOld skip uncharged pairs after the overlap checks
          // (Src/denergy.F90:3955-3959, 4066-4068)
          if (S->laz && (!S->laz[ip] || !S->laz[j])) continue;
          double us, fs;
          int e = table_uf(S->ubuf, S->nbuf, S->iubuflow[tp], r2, 0.0, S->r2umin[tp], us, fs);
          if (e) return e;
          (pass == 0 ? unew : uold)[1 + tp] += us;
        }
    }
  } else {
    for (int pass = 0; pass < 2; ++pass) {
      const double* ri = pass == 0 ? rnew : S->ro + 3 * ip;
      for (int j = 0; j < np; ++j) {
        if (j == ip) continue;
        int tp = tp_index(npt, S->ipt[ip], S->ipt[j]);
        double d[3] = {ri[0] - S->ro[3 * j], ri[1] - S->ro[3 * j + 1], ri[2] - S->ro[3 * j + 2]};
        double r2;
        pbc_r2(S->box, d, r2);
        if (r2 > rcut2) continue;
        if (pass == 0 && (r2 < S->r2atat[tp] || r2 < S->r2umin[tp])) return 1;
        if (S->laz && (!S->laz[ip] || !S->laz[j])) continue;
        double us, fs;
        int e = table_uf(S->ubuf, S->nbuf, S->iubuflow[tp], r2, 0.0, S->r2umin[tp], us, fs);
        if (e) return e;
        (pass == 0 ? unew : uold)[1 + tp] += us;
      }
    }
  }
  return 0;
}

int orc_dutotal(const OrcSys* S, int ip, const double* rnew, int* overlap, double* du_twob,
                double* du_rec, double* du_tot) {
  int np = S->np, npt = S->npt, nptpt = npt * (npt + 1) / 2;
  std::vector<double> unew(nptpt + 1, 0.0), uold(nptpt + 1, 0.0);
  *overlap = 1;
  *du_rec = 0;
  *du_tot = 0;
  for (int k = 0; k <= nptpt; ++k) du_twob[k] = 0;
  {
    CellIndex C;
    if (S->list == 1) {
      int rc = cell_index_build(S, C);
      if (rc) return rc;
    }
    int r = dut_real(S, S->list == 1 ? &C : nullptr, ip, rnew, unew.data(), uold.data());
    if (r == 1) return 0;  // overlap
    if (r) return r;
  }
  *overlap = 0;
  double s = 0;
  for (int k = 1; k <= nptpt; ++k) {
    du_twob[k] = unew[k] - uold[k];
    s += du_twob[k];
  }
  du_twob[0] = s;
  double drec = 0;
  if (S->ewald == 1) {
    int H = S->ncut + 1;
    std::vector<double> kfac(4 * H * H * H);
    int nk = orc_ewald_std_kfac(S->box, S->alpha, S->ncut, S->region, kfac.data(), (int)kfac.size());
    if (nk < 0) return nk;
    std::vector<double> s0(8 * nk), s1(8 * nk), f(3 * np, 0.0);
    std::vector<double> rtm(S->ro, S->ro + 3 * np);
    for (int k = 0; k < 3; ++k) rtm[3 * ip + k] = rnew[k];
    double u0, u1, v = 0;
    orc_ewald_std_rec(np, S->ro, S->az, S->box, S->alpha, S->ncut, S->region, S->epsi4pi,
                      kfac.data(), &u0, f.data(), &v, s0.data());
    orc_ewald_std_rec(np, rtm.data(), S->az, S->box, S->alpha, S->ncut, S->region, S->epsi4pi,
                      kfac.data(), &u1, f.data(), &v, s1.data());
    // termnew - termold per k as in DUTwoBodyEwaldRecStd (denergy.F90:1318-1325)
    for (int kn = 0; kn < nk; ++kn) {
      double tn = 0, to = 0;
      for (int q = 0; q < 4; ++q) {
        tn += s1[2 * (4 * kn + q)] * s1[2 * (4 * kn + q)] +
              s1[2 * (4 * kn + q) + 1] * s1[2 * (4 * kn + q) + 1];
        to += s0[2 * (4 * kn + q)] * s0[2 * (4 * kn + q)] +
              s0[2 * (4 * kn + q) + 1] * s0[2 * (4 * kn + q) + 1];
      }
      drec += kfac[kn] * (tn - to);
    }
  } else if (S->ewald == 2) {
    int K = S->K, H = K / 2 + 1;
    size_t g = (size_t)H * H * H, n3 = (size_t)K * K * K;
    std::vector<double> mf(g), ef(g), vf(g), Q(n3), Qtm(n3);
    int rc = orc_spme_setup(S->box, S->alpha, K, S->order, mf.data(), ef.data(), vf.data());
    if (rc) return rc;
    orc_spme_spread(np, S->ro, S->az, S->box, K, S->order, Q.data());
    double uold_rec = orc_spme_energy(K, ef.data(), Q.data());
    // QMeshTM = QMesh - old stencil + new stencil (denergy.F90:1435-1466)
    std::vector<double> dq(n3);
    double one_r[3] = {S->ro[3 * ip], S->ro[3 * ip + 1], S->ro[3 * ip + 2]};
    double qi = S->az[ip], mq = -qi;
    orc_spme_spread(1, one_r, &mq, S->box, K, S->order, dq.data());
    for (size_t i = 0; i < n3; ++i) Qtm[i] = Q[i] + dq[i];
    orc_spme_spread(1, rnew, &qi, S->box, K, S->order, dq.data());
    for (size_t i = 0; i < n3; ++i) Qtm[i] += dq[i];
    double unew_rec = orc_spme_energy(K, ef.data(), Qtm.data());
    drec = unew_rec - uold_rec;
  }
  if (S->ewald && S->surf) {
    // DUTwoBodyEwaldSurf, Src/denergy.F90:1504-1539
    double vol = S->box[0] * S->box[1] * S->box[2];
    double fac = 2 * kPi / (3 * vol);
    double m[3] = {0, 0, 0};
    for (int a = 0; a < np; ++a)
      for (int d = 0; d < 3; ++d) m[d] += S->az[a] * S->ro[3 * a + d];
    double mt[3];
    for (int d = 0; d < 3; ++d) mt[d] = m[d] + S->az[ip] * (rnew[d] - S->ro[3 * ip + d]);
    drec += fac * ((mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2]) -
                   (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
  }
  *du_rec = S->epsi4pi * drec;
  *du_tot = du_twob[0] + *du_rec;
  return 0;
}

// ChargeChange trial (Src/mc.F90:3683-3800) of a system with weak charges: the charge states of
// the nptm (1 or 2) particles ips[] flip (laztm = .not.laz, aztm = zat[type] or 0), positions
// stay.  DUTotal's weak-charge branch (Src/denergy.F90:154-164): real space over the flipped
// particles' partners with the new and the old states (UWeakChargeANew/AOld, :3911-4145; the
// pair of the two flipped particles counted once, lptmdutwob), reciprocal space from the
// changed structure factors (DUWeakChargeEwaldRecStd, :1586-1650), the self term (:1654-1667)
// and the surface term (:1671-1695).  Std Ewald or none: the reference stops with SPME.
int orc_dutotal_charge(const OrcSys* S, int nptm, const int* ips, int* overlap, double* du_twob,
                       double* du_rec, double* du_tot) {
  if (!S->laz || !S->zat || nptm < 1 || nptm > 2 || S->ewald == 2) return ORC_ERR_ARG;
  const int np = S->np, npt = S->npt, nptpt = npt * (npt + 1) / 2;
  const bool inclusive = S->list != 1;  // all pairs: UTwoBodyA's r2 <= rcut2
  const double rcut2 = S->rcut * S->rcut;
  std::vector<int> lazn(S->laz, S->laz + np);
  std::vector<double> azn(S->az, S->az + np);
  std::vector<char> moved(np, 0);
  for (int m = 0; m < nptm; ++m) {
    const int i = ips[m];
    if (i < 0 || i >= np) return ORC_ERR_ARG;
    moved[i] = 1;
    lazn[i] = !S->laz[i];
    azn[i] = lazn[i] ? S->zat[S->ipt[i]] : 0.0;
  }
  std::vector<double> unew(nptpt + 1, 0.0), uold(nptpt + 1, 0.0);
  *overlap = 1;
  *du_rec = 0;
  *du_tot = 0;
  for (int k = 0; k <= nptpt; ++k) du_twob[k] = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int* lz = pass == 0 ? lazn.data() : S->laz;
    for (int m = 0; m < nptm; ++m) {
      const int i = ips[m];
      for (int j = 0; j < np; ++j) {
        if (j == i || (moved[j] && j < i)) continue;
        const int tp = tp_index(npt, S->ipt[i], S->ipt[j]);
        double d[3] = {S->ro[3 * i] - S->ro[3 * j], S->ro[3 * i + 1] - S->ro[3 * j + 1],
                       S->ro[3 * i + 2] - S->ro[3 * j + 2]};
        double r2;
        pbc_r2(S->box, d, r2);
        if (inclusive ? r2 > rcut2 : !(r2 < rcut2)) continue;
        if (r2 < S->r2atat[tp] || r2 < S->r2umin[tp]) return 0;  // overlap
        if (!lz[i] || !lz[j]) continue;
        double us, fs;
        const int e = table_uf(S->ubuf, S->nbuf, S->iubuflow[tp], r2, 0.0, S->r2umin[tp], us, fs);
        if (e) return e;
        (pass == 0 ? unew : uold)[1 + tp] += us;
      }
    }
  }
  *overlap = 0;
  double s = 0;
  for (int k = 1; k <= nptpt; ++k) {
    du_twob[k] = unew[k] - uold[k];
    s += du_twob[k];
  }
  du_twob[0] = s;
  double drec = 0;
  if (S->ewald == 1) {
    const int H = S->ncut + 1;
    std::vector<double> kfac(4 * H * H * H);
    const int nk = orc_ewald_std_kfac(S->box, S->alpha, S->ncut, S->region, kfac.data(), (int)kfac.size());
    if (nk < 0) return nk;
    std::vector<double> s0(8 * nk), s1(8 * nk), f(3 * np, 0.0);
    double u0, u1, v = 0;
    orc_ewald_std_rec(np, S->ro, S->az, S->box, S->alpha, S->ncut, S->region, S->epsi4pi, kfac.data(),
                      &u0, f.data(), &v, s0.data());
    orc_ewald_std_rec(np, S->ro, azn.data(), S->box, S->alpha, S->ncut, S->region, S->epsi4pi,
                      kfac.data(), &u1, f.data(), &v, s1.data());
    for (int kn = 0; kn < nk; ++kn) {
      double tn = 0, to = 0;
      for (int q = 0; q < 4; ++q) {
        tn += s1[2 * (4 * kn + q)] * s1[2 * (4 * kn + q)] + s1[2 * (4 * kn + q) + 1] * s1[2 * (4 * kn + q) + 1];
        to += s0[2 * (4 * kn + q)] * s0[2 * (4 * kn + q)] + s0[2 * (4 * kn + q) + 1] * s0[2 * (4 * kn + q) + 1];
      }
      drec += kfac[kn] * (tn - to);
    }
  }
  if (S->ewald) {
    const double fac = S->alpha / std::sqrt(kPi);
    for (int m = 0; m < nptm; ++m) {
      const int i = ips[m];
      drec += fac * (S->az[i] * S->az[i] - azn[i] * azn[i]);
    }
    if (S->surf) {
      const double vol = S->box[0] * S->box[1] * S->box[2], sf = 2 * kPi / (3 * vol);
      double m0[3] = {0, 0, 0}, m1[3] = {0, 0, 0};
      for (int a = 0; a < np; ++a)
        for (int dd = 0; dd < 3; ++dd) {
          m0[dd] += S->az[a] * S->ro[3 * a + dd];
          m1[dd] += azn[a] * S->ro[3 * a + dd];
        }
      drec += sf * ((m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]) - (m0[0] * m0[0] + m0[1] * m0[1] + m0[2] * m0[2]));
    }
  }
  *du_rec = S->epsi4pi * drec;
  *du_tot = du_twob[0] + *du_rec;
  return 0;
}

// ----------------------------------------------------------------------------------
// Verlet lists at rcut + drnlist, serial branch of SetVList (Src/nlist.F90:724-759 MD,
// :820-884 MC; rs2cut = (rcut + drnlist)^2 and r2 < rs2cut with PBCr2).
//   kind 0, SetVListMD: row ip holds partner jp only by the parity rule of :742-745 (1-based
//          ip + jp odd: the lower index keeps the pair; even: the higher one), so every pair
//          is listed once; partners in ascending order (the jp loop runs 1..np).
//   kind 1, SetVListMC: the full symmetric list; the ip < jp loop appends jp to row ip and
//          ip to row jp, which also leaves every row in ascending order.
// jpnlist[ip * maxn + k] holds 0-based indices, nneighpn[ip] the row lengths.  Returns 0, or
// ORC_ERR_NLIST when a row would exceed maxn (the reference stops, :752, :875-877).  O(np^2)
// pair tests, as the reference.
enum { ORC_ERR_NLIST = -11 };

static inline bool vlist_md_keeps(int ip, int jp) {
  // (ip+1) + (jp+1) has the parity of ip + jp
  const int itemp = (ip + jp) & 1;
  if (ip < jp && itemp == 0) return false;
  if (ip > jp && itemp != 0) return false;
  return ip != jp;
}

int orc_vlist(int kind, int np, const double* ro, const double* box, double rcut, double drnlist,
              int maxn, int* nneighpn, int* jpnlist) {
  const double rs2cut = (rcut + drnlist) * (rcut + drnlist);
  for (int i = 0; i < np; ++i) nneighpn[i] = 0;
  if (kind == 0) {
    for (int ip = 0; ip < np; ++ip)
      for (int jp = 0; jp < np; ++jp) {
        if (!vlist_md_keeps(ip, jp)) continue;
        double d[3] = {ro[3 * ip] - ro[3 * jp], ro[3 * ip + 1] - ro[3 * jp + 1], ro[3 * ip + 2] - ro[3 * jp + 2]};
        double r2;
        pbc_r2(box, d, r2);
        if (r2 < rs2cut) {
          if (nneighpn[ip] >= maxn) return ORC_ERR_NLIST;
          jpnlist[(size_t)ip * maxn + nneighpn[ip]++] = jp;
        }
      }
    return 0;
  }
  for (int ip = 0; ip < np; ++ip)
    for (int jp = ip + 1; jp < np; ++jp) {
      double d[3] = {ro[3 * ip] - ro[3 * jp], ro[3 * ip + 1] - ro[3 * jp + 1], ro[3 * ip + 2] - ro[3 * jp + 2]};
      double r2;
      pbc_r2(box, d, r2);
      if (r2 < rs2cut) {
        if (nneighpn[ip] >= maxn || nneighpn[jp] >= maxn) return ORC_ERR_NLIST;
        jpnlist[(size_t)ip * maxn + nneighpn[ip]++] = jp;
        jpnlist[(size_t)jp * maxn + nneighpn[jp]++] = ip;
      }
    }
  return 0;
}

// The same lists from a cell list (InitCellList at rcell = rcut + drnlist, Src/celllist.F90:
// 76-265): the candidates of each particle are the particles of its stencil cells, the same
// pair test and parity rule are applied, and each row is sorted ascending -- the rows
// orc_vlist writes, in O(np) instead of O(np^2) (this is the large-system form, for lists of
// a million particles; its rows equal orc_vlist's bit for bit, which the tests check).
int orc_vlist_cells(int kind, int np, const double* ro, const double* box, double rcut, double drnlist,
                    int maxn, int* nneighpn, int* jpnlist) {
  const double rl = rcut + drnlist, rs2cut = rl * rl;
  int nc[3];
  double cs[3];
  std::vector<int> st(3 * 4096);
  const int nst = orc_cell_init(box, rl, rl, nc, cs, st.data(), 4096);
  if (nst < 0) return nst;
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> head(ncell), nxt(np), npart(ncell), cellof(np);
  if (int e = orc_cell_build(np, ro, box, nc, cs, head.data(), nxt.data(), npart.data(), cellof.data())) return e;
  // the cells' members copied out contiguously (positions with them), so a candidate loop
  // reads consecutive memory instead of following the linked lists through the whole array
  std::vector<int> cstart(ncell + 1, 0), cidx(np);
  std::vector<double> cpos(3 * (size_t)np);
  for (int c = 0; c < ncell; ++c) cstart[c + 1] = cstart[c] + npart[c];
  for (int c = 0; c < ncell; ++c) {
    int k = cstart[c];
    for (int j = head[c]; j >= 0; j = nxt[j], ++k) {
      cidx[k] = j;
      for (int d = 0; d < 3; ++d) cpos[3 * (size_t)k + d] = ro[3 * j + d];
    }
  }
  int err = 0;
#pragma omp parallel
  {
    std::vector<int> nb(nst), row;
#pragma omp for schedule(dynamic, 256)
    for (int ip = 0; ip < np; ++ip) {
      row.clear();
      const int nn = cell_neigh(cellof[ip], nc, st.data(), nst, nb.data());
      for (int s = 0; s < nn; ++s)
        for (int k = cstart[nb[s]]; k < cstart[nb[s] + 1]; ++k) {
          const int jp = cidx[k];
          if (kind == 0 ? !vlist_md_keeps(ip, jp) : jp == ip) continue;
          const double* pj = &cpos[3 * (size_t)k];
          double d[3] = {ro[3 * ip] - pj[0], ro[3 * ip + 1] - pj[1], ro[3 * ip + 2] - pj[2]};
          double r2;
          pbc_r2(box, d, r2);
          if (r2 < rs2cut) row.push_back(jp);
        }
      std::sort(row.begin(), row.end());
      if ((int)row.size() > maxn) {
#pragma omp atomic write
        err = ORC_ERR_NLIST;
        nneighpn[ip] = maxn;
      } else {
        nneighpn[ip] = (int)row.size();
      }
      for (int k = 0; k < nneighpn[ip]; ++k) jpnlist[(size_t)ip * maxn + k] = row[k];
    }
  }
  return err;
}

// NList skin test with an adjusted update interval (inlist = 0), Src/nlist.F90:600-612:
// over the particles in index order, r2 = sum(drosum(1:3,ip)**2); a value above r2max1
// replaces r2max1 WITHOUT passing the old maximum on to r2max2, otherwise a value above
// r2max2 replaces r2max2 (so r2max2 is not always the second largest value).  drosum holds
// each particle's displacement summed since the last build.  Returns 1 when
// sqrt(r2max1) + sqrt(r2max2) > drnlist (the list is to be rebuilt), else 0.
int orc_nlist_skin(int np, const double* drosum, double drnlist, double* r2max1, double* r2max2) {
  double m1 = 0.0, m2 = 0.0;
  for (int ip = 0; ip < np; ++ip) {
    const double* d = drosum + 3 * ip;
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 > m1)
      m1 = r2;
    else if (r2 > m2)
      m2 = r2;
  }
  *r2max1 = m1;
  *r2max2 = m2;
  return std::sqrt(m1) + std::sqrt(m2) > drnlist ? 1 : 0;
}

// SetCellList with MOLSIM's doubly linked lists (AddIpToCell by head insertion in particle
// order, Src/celllist.F90:307-327,384-410).  head[c] (-1: empty, the reference's 0), npart[c],
// next[i] / prev[i] (-1: none), cellof[i].  Cell of a particle as pcellro (:295-305).
int orc_celllist_set(int np, const double* ro, const double* box, const int* nc, const double* csize,
                     int* head, int* npart, int* next, int* prev, int* cellof) {
  const int ncell = nc[0] * nc[1] * nc[2];
  if (int e = orc_cell_build(np, ro, box, nc, csize, head, next, npart, cellof)) return e;
  for (int i = 0; i < np; ++i) prev[i] = -1;
  for (int c = 0; c < ncell; ++c)
    for (int i = head[c]; i >= 0; i = next[i])
      if (next[i] >= 0) prev[next[i]] = i;
  return 0;
}

// UpdateCellIp(ip), Src/celllist.F90:329-382: when the cell of ro[ip] differs from the cell ip
// is linked in, RmIpFromCell unlinks it (its four cases: alone, head, tail, inside) and
// AddIpToCell makes it the head of its new cell.  Returns 1 when the particle moved to another
// cell, 0 when it stayed, or an error code.
int orc_celllist_update(int ip, const double* ro, const double* box, const int* nc, const double* csize,
                        int* head, int* npart, int* next, int* prev, int* cellof) {
  int ci[3];
  for (int k = 0; k < 3; ++k) {
    ci[k] = cell_coord(ro[3 * ip + k], box[k], 1.0 / csize[k]);
    if (ci[k] < 0 || ci[k] > nc[k]) return ORC_ERR_ARG;
    if (ci[k] == nc[k]) ci[k] = nc[k] - 1;
  }
  const int cnew = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]), cold = cellof[ip];
  if (cnew == cold) return 0;
  // RmIpFromCell
  npart[cold]--;
  const int nx = next[ip], pv = prev[ip];
  if (nx < 0) {
    if (head[cold] == ip) head[cold] = -1;
    else next[pv] = -1;
  } else {
    if (head[cold] == ip) {
      head[cold] = nx;
      prev[nx] = -1;
    } else {
      prev[nx] = pv;
      next[pv] = nx;
    }
  }
  // AddIpToCell
  npart[cnew]++;
  const int jp = head[cnew];
  head[cnew] = ip;
  next[ip] = jp;
  if (jp >= 0) prev[jp] = ip;
  prev[ip] = -1;
  cellof[ip] = cnew;
  return 1;
}

// ----------------------------------------------------------------------------------
// Trial energies the way MOLSIM's MC loop gets them (the CPU baseline of batched trials):
// the configuration's cell list, structure factors S(k) (EwaldSetArray, Src/energy.F90:
// 1401-1426), charge mesh and dipole stay resident, and each trial costs
//   real space   the old and new positions' partners (UTwoBodyANew/AOld, as dut_real);
//   std Ewald    O(nkvec): S_tm(k) = S(k) + q (e^{ik.rnew} - e^{ik.rold}) per octant sign
//                combination and dE = sum kfac (|S_tm|^2 - |S|^2) (DUTwoBodyEwaldRecStd with
//                EwaldSetArrayTM, Src/denergy.F90:1275-1335,1708-1731);
//   SPME         QMeshTM = QMesh - old stencil + new stencil and one FFT energy
//                (DUTwoBodyEwaldRecSPM, Src/denergy.F90:1413-1495);
//   surface      the resident dipole (DUTwoBodyEwaldSurf, Src/denergy.F90:1504-1539).
// Trials are independent (evaluated against the same configuration).
struct OrcTrials {
  OrcSys S;
  CellIndex C;
  std::vector<double> kfac;
  int nk = 0;
  std::vector<cplx> sk;  // 4 per k vector, in kfac order
  std::vector<double> ef, Q;
  double m[3] = {0, 0, 0};
};

// e^{i 2pi n x / L} for n = 0..ncut, by the recurrence orc_ewald_std_rec uses.
static void eikr_line(double x, double tpb, int ncut, cplx* e) {
  e[0] = cplx(1, 0);
  if (ncut >= 1) e[1] = cplx(std::cos(tpb * x), std::sin(tpb * x));
  for (int c = 2; c <= ncut; ++c) e[c] = e[c - 1] * e[1];
}

void* orc_trials_open(const OrcSys* S) {
  OrcTrials* T = new OrcTrials;
  T->S = *S;
  if (S->list == 1 && cell_index_build(S, T->C)) {
    delete T;
    return nullptr;
  }
  if (S->ewald == 1) {
    const int H = S->ncut + 1;
    T->kfac.assign(4 * H * H * H, 0.0);
    T->nk = orc_ewald_std_kfac(S->box, S->alpha, S->ncut, S->region, T->kfac.data(), (int)T->kfac.size());
    if (T->nk < 0) {
      delete T;
      return nullptr;
    }
    std::vector<double> f(3 * S->np, 0.0), sr(8 * T->nk);
    double u, v = 0;
    orc_ewald_std_rec(S->np, S->ro, S->az, S->box, S->alpha, S->ncut, S->region, S->epsi4pi,
                      T->kfac.data(), &u, f.data(), &v, sr.data());
    T->sk.resize(4 * T->nk);
    for (int q = 0; q < 4 * T->nk; ++q) T->sk[q] = cplx(sr[2 * q], sr[2 * q + 1]);
  } else if (S->ewald == 2) {
    const int K = S->K, H = K / 2 + 1;
    const size_t g = (size_t)H * H * H;
    std::vector<double> mf(g), vf(g);
    T->ef.assign(g, 0.0);
    T->Q.assign((size_t)K * K * K, 0.0);
    if (orc_spme_setup(S->box, S->alpha, K, S->order, mf.data(), T->ef.data(), vf.data())) {
      delete T;
      return nullptr;
    }
    orc_spme_spread(S->np, S->ro, S->az, S->box, K, S->order, T->Q.data());
  }
  for (int a = 0; a < S->np; ++a)
    for (int d = 0; d < 3; ++d) T->m[d] += S->az[a] * S->ro[3 * a + d];
  return T;
}

void orc_trials_close(void* h) { delete (OrcTrials*)h; }

int orc_trials_dutotal(void* h, int ip, const double* rnew, int* overlap, double* du_twob, double* du_rec,
                       double* du_tot) {
  OrcTrials* T = (OrcTrials*)h;
  const OrcSys* S = &T->S;
  const int np = S->np, npt = S->npt, nptpt = npt * (npt + 1) / 2;
  if (ip < 0 || ip >= np) return ORC_ERR_ARG;
  std::vector<double> unew(nptpt + 1, 0.0), uold(nptpt + 1, 0.0);
  *overlap = 1;
  *du_rec = 0;
  *du_tot = 0;
  for (int k = 0; k <= nptpt; ++k) du_twob[k] = 0;
  const int r = dut_real(S, S->list == 1 ? &T->C : nullptr, ip, rnew, unew.data(), uold.data());
  if (r == 1) return 0;
  if (r) return r;
  *overlap = 0;
  double s = 0;
  for (int k = 1; k <= nptpt; ++k) {
    du_twob[k] = unew[k] - uold[k];
    s += du_twob[k];
  }
  du_twob[0] = s;
  double drec = 0;
  const double qi = S->az[ip];
  const double* ro = S->ro + 3 * ip;
  if (S->ewald == 1) {
    const int nc = S->ncut, ncut2 = S->region == 0 ? nc * nc : 3 * nc * nc;
    std::vector<cplx> eo(3 * (nc + 1)), en(3 * (nc + 1));
    for (int d = 0; d < 3; ++d) {
      const double tpb = 2 * kPi / S->box[d];
      eikr_line(ro[d], tpb, nc, eo.data() + d * (nc + 1));
      eikr_line(rnew[d], tpb, nc, en.data() + d * (nc + 1));
    }
    const cplx* ox = eo.data();
    const cplx* oy = ox + nc + 1;
    const cplx* oz = oy + nc + 1;
    const cplx* nx_ = en.data();
    const cplx* ny_ = nx_ + nc + 1;
    const cplx* nz_ = ny_ + nc + 1;
    int kn = 0;
    for (int nz = 0; nz <= nc; ++nz)
      for (int ny = 0; ny <= nc; ++ny) {
        if (ny * ny + nz * nz > ncut2) continue;
        const cplx om = std::conj(oy[ny]) * oz[nz], op = oy[ny] * oz[nz];
        const cplx nm = std::conj(ny_[ny]) * nz_[nz], npp = ny_[ny] * nz_[nz];
        for (int nx = 0; nx <= nc; ++nx) {
          if (nx * nx + ny * ny + nz * nz > ncut2) continue;
          if (nx == 0 && ny == 0 && nz == 0) continue;
          const cplx e_old[4] = {std::conj(ox[nx]) * om, std::conj(ox[nx]) * op, ox[nx] * om, ox[nx] * op};
          const cplx e_new[4] = {std::conj(nx_[nx]) * nm, std::conj(nx_[nx]) * npp, nx_[nx] * nm, nx_[nx] * npp};
          double term = 0;
          for (int q = 0; q < 4; ++q) {
            const cplx s0 = T->sk[4 * kn + q];
            const cplx s1 = s0 + qi * (e_new[q] - e_old[q]);
            term += std::norm(s1) - std::norm(s0);
          }
          drec += T->kfac[kn] * term;
          ++kn;
        }
      }
  } else if (S->ewald == 2) {
    const int K = S->K;
    const size_t n3 = (size_t)K * K * K;
    std::vector<double> Qtm(T->Q), dq(n3);
    const double uold_rec = orc_spme_energy(K, T->ef.data(), T->Q.data());
    const double one_r[3] = {ro[0], ro[1], ro[2]};
    const double mq = -qi;
    orc_spme_spread(1, one_r, &mq, S->box, K, S->order, dq.data());
    for (size_t i = 0; i < n3; ++i) Qtm[i] += dq[i];
    orc_spme_spread(1, rnew, &qi, S->box, K, S->order, dq.data());
    for (size_t i = 0; i < n3; ++i) Qtm[i] += dq[i];
    drec = orc_spme_energy(K, T->ef.data(), Qtm.data()) - uold_rec;
  }
  if (S->ewald && S->surf) {
    const double vol = S->box[0] * S->box[1] * S->box[2];
    const double fac = 2 * kPi / (3 * vol);
    double mt[3];
    for (int d = 0; d < 3; ++d) mt[d] = T->m[d] + qi * (rnew[d] - ro[d]);
    drec += fac * ((mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2]) -
                   (T->m[0] * T->m[0] + T->m[1] * T->m[1] + T->m[2] * T->m[2]));
  }
  *du_rec = S->epsi4pi * drec;
  *du_tot = du_twob[0] + *du_rec;
  return 0;
}

}  // extern "C"
