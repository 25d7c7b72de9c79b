// Host-side potential tables: construction (msg_build_table, restating PotTwoBodyTab1/2/3,
// SetUBuffer and CheckUBuffer of Src/potential.F90:868-1258) and packing of MOLSIM's table
// state into the GPU layout (pack_table).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// u(r), u'(r), u''(r) of one type pair.  Form 0 is MOLSIM's 'default' potential with the
// 1/r term only (Src/potential.F90:1266-1341), form 1 is '(1,6,12)' (:1349-1425); with
// lewald the 1/r term is damped by erfc(alpha r) and its derivatives (CalcEwaldUfac,
// :2938-2960).
class PairPotential {
 public:
  PairPotential(const msg_pot& p, int it, int jt) : p_(p) {
    qq_ = p.z[it] * p.z[jt] * p.epsi4pi;
    sig_ = 0.5 * (p.sig[it] + p.sig[jt]);
    eps_ = std::sqrt(p.eps[it] * p.eps[jt]);
  }
  void operator()(double r, double& u0, double& u1, double& u2) const {
    double c0 = 1.0, c1 = 1.0, c2 = 1.0;
    if (p_.lewald) {
      const double ra = p_.alpha * r;
      const double g = 2.0 / std::sqrt(kPi) * std::exp(-ra * ra);
      const double ec = std::erfc(ra);
      c0 = ec;
      c1 = ec + ra * g;
      c2 = ec + ra * (1.0 + ra * ra) * g;
    }
    const double ri = 1.0 / r;
    if (p_.form == 0) {
      u0 = c0 * (qq_ * ri);
      u1 = -c1 * (qq_ * ri) * ri;
      u2 = 2.0 * c2 * (qq_ * ri) * ri * ri;
      return;
    }
    const double s6 = std::pow(sig_ * ri, 6);
    u0 = qq_ * c0 * ri + 4.0 * eps_ * (s6 * s6 - s6);
    u1 = -qq_ * c1 * ri * ri - 24.0 * eps_ * (2.0 * s6 * s6 - s6) * ri;
    u2 = 2.0 * qq_ * c2 * ri * ri * ri + 24.0 * eps_ * (26.0 * s6 * s6 - 7.0 * s6) * ri * ri;
  }

 private:
  const msg_pot& p_;
  double qq_, sig_, eps_;
};

// One table record: quintic U and quartic F/r in d = r^2 - zlow that match w, dw/dz and
// d2w/dz2 (w(z) = u(sqrt z)) at both interval ends.
struct Record {
  double c[12];
  static Record fit(double zl, const double ul[3], double zu, const double uu[3]) {
    Record t{};
    t.c[0] = zl;
    if (ul[0] == 0.0 && uu[0] == 0.0) return t;
    const double rl = std::sqrt(zl), ru = std::sqrt(zu);
    const double h = zu - zl, h2 = h * h, h3 = h2 * h;
    const double w0l = ul[0], w1l = 0.5 * ul[1] / rl, w2l = 0.25 * (ul[2] - ul[1] / rl) / zl;
    const double w0u = uu[0], w1u = 0.5 * uu[1] / ru, w2u = 0.25 * (uu[2] - uu[1] / ru) / zu;
    const double a = 6.0 * (w0u - w0l - w1l * h - 0.5 * w2l * h2) / h3;
    const double b = 2.0 * (w1u - w1l - w2l * h) / h2;
    const double c = (w2u - w2l) / h;
    t.c[1] = w0l;
    t.c[2] = w1l;
    t.c[3] = 0.5 * w2l;
    t.c[4] = (10.0 * a - 12.0 * b + 3.0 * c) / 6.0;
    t.c[5] = (-15.0 * a + 21.0 * b - 6.0 * c) / (6.0 * h);
    t.c[6] = (2.0 * a - 3.0 * b + c) / (2.0 * h2);
    for (int k = 0; k < 5; ++k) t.c[7 + k] = -2.0 * (k + 1) * t.c[2 + k];
    return t;
  }
  double u(double dz) const {
    return c[1] + dz * (c[2] + dz * (c[3] + dz * (c[4] + dz * (c[5] + dz * c[6]))));
  }
  double f(double dz) const {
    return c[7] + dz * (c[8] + dz * (c[9] + dz * (c[10] + dz * c[11])));
  }
};

// Accuracy check of one candidate interval at 11 equispaced r values: |dU| <= utoltab and
// |dF| <= ftoltab; 'repulsive' when |U| or |F| exceeds umaxtab / fmaxtab.
bool interval_ok(const PairPotential& pot, const msg_pot& p, double rl, double ru, const Record& t,
                 bool& repulsive) {
  repulsive = false;
  const double step = (ru - rl) / 10.0;
  for (int k = 0; k <= 10; ++k) {
    const double r = rl + step * k;
    double u0, u1, u2;
    pot(r, u0, u1, u2);
    const double dz = r * r - rl * rl;
    const double us = t.u(dz), fs = r * t.f(dz);
    if (std::fabs(us - u0) > p.utoltab || std::fabs(fs + u1) > p.ftoltab) return false;
    if (std::fabs(us) > p.umaxtab || std::fabs(fs) > p.fmaxtab) repulsive = true;
  }
  return true;
}

}  // namespace

extern "C" int msg_build_table(const msg_pot* p, double* ubuf, int maxbuf, int* iubuflow,
                               int* nugrid, double* r2umin, double* r2atat) {
  using msg::set_error;
  if (!p || !ubuf || !iubuflow || !nugrid || !r2umin || !r2atat || p->npt < 1 || p->npt > 8)
    return set_error(MSG_ERR_ARG, "msg_build_table: bad arguments");
  const int max_intervals = 1200, max_shrink = 100;
  int used = 0;
  for (int it = 0; it < p->npt; ++it) {
    for (int jt = it; jt < p->npt; ++jt) {
      const int tp = msg::tp_index(p->npt, it, jt);
      const PairPotential pot(*p, it, jt);
      const double rumax = std::sqrt((1.1 * p->rcut) * (1.1 * p->rcut));  // rumax = sqrt(r2umax)
      double rumin = std::sqrt(p->r2uminin);
      r2atat[tp] = (p->rad[it] + p->rad[jt]) * (p->rad[it] + p->rad[jt]);
      iubuflow[tp] = used + 1;
      double rupp = rumax, dr = 0.5 * (rumax - rumin);
      int n = 0;
      bool done = false;
      while (!done) {
        if (n == max_intervals) return set_error(MSG_ERR_TABLE, "msg_build_table: too many intervals");
        double uu[3];
        pot(rupp, uu[0], uu[1], uu[2]);
        Record rec;
        double rlow = rupp;
        bool ok = false, repulsive = false;
        for (int s = 0; s < max_shrink && !ok; ++s) {
          rlow = std::max(rumin, rupp - dr);
          double ul[3];
          pot(rlow, ul[0], ul[1], ul[2]);
          rec = Record::fit(rlow * rlow, ul, rupp * rupp, uu);
          ok = interval_ok(pot, *p, rlow, rupp, rec, repulsive);
          if (!ok) dr *= 0.9;
        }
        if (!ok) return set_error(MSG_ERR_TABLE, "msg_build_table: interval did not converge");
        if (used + 12 > maxbuf) return set_error(MSG_ERR_TABLE_MEM, "msg_build_table: ubuf too small");
        std::memcpy(ubuf + used, rec.c, sizeof(rec.c));
        used += 12;
        ++n;
        rupp = rlow;
        if (repulsive) rumin = rlow;
        done = rlow <= rumin || repulsive;
      }
      nugrid[tp] = n;
      r2umin[tp] = rumin * rumin;
    }
  }
  return used;
}

namespace msg {

// MOLSIM's table (ubuf with 1-based iubuflow) -> GPU layout.  Only intervals that a pair at
// max(r2atat, r2umin) <= r2 < rcut^2 can select are kept; the scan order is preserved.
int pack_table(Ctx* c) {
  const int ntp = c->nptpt;
  const double rc2 = c->P.rcut * c->P.rcut;
  GpuTableMeta& m = c->tmeta;
  std::vector<double> coef;
  for (int tp = 0; tp < ntp; ++tp) {
    const int lo = c->h_iubuflow[tp] - 1, n = c->h_nugrid[tp];
    if (lo < 0 || lo + 12 * n > (int)c->h_ubuf.size() || n < 1)
      return set_error(MSG_ERR_ARG, "msg_set_table: iubuflow/nugrid outside ubuf");
    const double zmin = std::max(c->h_r2atat[tp], c->h_r2umin[tp]);
    // first interval a pair with r2 < rc2 can select: the first with zlow < rc2
    int a = 0;
    while (a < n - 1 && c->h_ubuf[lo + 12 * a] >= rc2) ++a;
    // last interval needed: the first with zlow <= zmin
    int b = a;
    while (b < n - 1 && c->h_ubuf[lo + 12 * b] > zmin) ++b;
    m.first[tp] = (int)coef.size() / 12;
    for (int k = a; k <= b; ++k)
      coef.insert(coef.end(), c->h_ubuf.begin() + lo + 12 * k, c->h_ubuf.begin() + lo + 12 * k + 12);
    m.last[tp] = (int)coef.size() / 12 - 1;
    const double zl = std::min(zmin, rc2), zh = rc2;
    m.zlo[tp] = zl;
    m.invw[tp] = zh > zl ? kTabBins / (zh - zl) : 0.0;
    m.r2atat[tp] = c->h_r2atat[tp];
    m.r2umin[tp] = c->h_r2umin[tp];
  }
  m.nint = (int)coef.size() / 12;
  m.ntp = ntp;
  const int cap = c->npt > kLdsTypes ? kMaxTabIntervalsMany : kMaxTabIntervals;
  if (m.nint > cap)
    return set_error(MSG_ERR_TABLE_MEM, c->npt > kLdsTypes ? "msg_set_table: more than 4096 intervals below rcut"
                                                           : "msg_set_table: more than 512 intervals below rcut");
  for (int i = 0; i < kMaxT; ++i)
    for (int j = 0; j < kMaxT; ++j)
      m.tpidx[i * kMaxT + j] = (i < c->npt && j < c->npt) ? tp_index(c->npt, i, j) : 0;
  // bins: the first interval (scan order) with zlow <= the upper edge of the next bin
  std::vector<int> bins((size_t)ntp * kTabBins);
  for (int tp = 0; tp < ntp; ++tp) {
    const double w = m.invw[tp] > 0 ? 1.0 / m.invw[tp] : 0.0;
    for (int b = 0; b < kTabBins; ++b) {
      const double edge = m.zlo[tp] + (b + 2) * w;  // one bin of margin for FP32 rounding of the bin index
      int k = m.first[tp];
      while (k < m.last[tp] && coef[12 * k] > edge) ++k;
      bins[(size_t)tp * kTabBins + b] = k;
    }
  }
  const size_t rs = c->fp64 ? sizeof(double) : sizeof(float);
  if (!c->d_tcoef) {
    const size_t cap_b = (size_t)cap * 12 * sizeof(double);
    if (int e = dalloc(c, &c->d_tcoef, cap_b)) return e;
    if (int e = dalloc(c, (void**)&c->d_tcoef64, cap_b)) return e;
    if (int e = dalloc(c, (void**)&c->d_tbins, kMaxTP * kTabBins * sizeof(int))) return e;
  }
  if (c->fp64) {
    MSG_HIP(hipMemcpy(c->d_tcoef, coef.data(), coef.size() * rs, hipMemcpyHostToDevice));
  } else {
    std::vector<float> cf(coef.begin(), coef.end());
    MSG_HIP(hipMemcpy(c->d_tcoef, cf.data(), cf.size() * rs, hipMemcpyHostToDevice));
  }
  MSG_HIP(hipMemcpy(c->d_tcoef64, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
  MSG_HIP(hipMemcpy(c->d_tbins, bins.data(), bins.size() * sizeof(int), hipMemcpyHostToDevice));
  return MSG_OK;
}

}  // namespace msg
