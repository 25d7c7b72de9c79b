// Single-particle Monte Carlo caller of DUTotal on the host: MCPass / SPartMove /
// GetRandomTrialPos / Metropolis / MCUpdate (Src/mc.F90:1470-1750,4527-4573,5135-5235) with
// MOLSIM's random number generator (Random, Src/mollib.F90:2689-2713): a Marsaglia xorshift
// combined with a Park-Miller sequence computed by Schrage's method.
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.h"

namespace {

// Generator state as the reference keeps it: the seed (iseed, positive after the first call),
// the two 32-bit sequences and the output scale.
struct Generator {
  int32_t* seed;
  int32_t* st;  // [0] xorshift word, [1] Park-Miller word
  double* scale;

  double next() {
    const int32_t a = 16807, m = 2147483647, q = 127773, r = 2836;
    if (*seed <= 0 || st[1] < 0) {
      const int32_t s = *seed < 0 ? -*seed : *seed;
      // the default-kind (single precision) real just below 1 divided by m in single
      // precision, where m rounds to 2^31: (1 - 2^-24) 2^-31, exact
      *scale = (double)std::nextafter(1.0f, 0.0f) * 0x1p-31;
      st[1] = (888889999 ^ s) | 1;
      st[0] = 777755555 ^ s;
      *seed = s + 1;
    }
    uint32_t x = (uint32_t)st[0];
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    st[0] = (int32_t)x;
    const int32_t k = st[1] / q;
    st[1] = a * (st[1] - k * q) - r * k;
    if (st[1] < 0) st[1] += m;
    return *scale * (double)((m & (st[0] ^ st[1])) | 1);
  }
};

}  // namespace

using namespace msg;

extern "C" {

double msg_random(int* iseed, int* state, double* am) {
  Generator g{iseed, state, am};
  return g.next();
}

int msg_mc_pass(msg_ctx* h, int ntrial, int isamp, const double* dtran, double beta, int* iseed,
                int* rng_state, double* am, double* du_sum, int* naccept, int* events) {
  if (!h || ntrial < 0 || !dtran || !iseed || !rng_state || !am)
    return set_error(MSG_ERR_ARG, "msg_mc_pass: bad arguments");
  Ctx* c = &h->c;
  if (chain_eligible(c) && !c->prof)  // the whole pass on the device
    return mc_chain(c, ntrial, isamp, dtran, beta, iseed, rng_state, am, du_sum, naccept, events);
  const double expmax = 87.0;  // Src/mol.F90:62
  std::vector<double> ro(3 * (size_t)c->np);
  MSG_HIP(hipStreamSynchronize(c->stream));
  MSG_HIP(hipMemcpy(ro.data(), c->d_ro, ro.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int> ipt(c->np);
  MSG_HIP(hipMemcpy(ipt.data(), c->d_ipt, c->np * sizeof(int), hipMemcpyDeviceToHost));
  Generator rnd{iseed, rng_state, am};
  double dsum = 0.0;
  int nacc = 0;
  for (int t = 0; t < ntrial; ++t) {
    // MCPass: the particle to move
    int ip;
    if (isamp == 0) {
      ip = c->mc_next;
      c->mc_next = (c->mc_next + 1) % c->np;
    } else {
      ip = (int)(c->np * rnd.next());
      ip = std::min(std::max(ip, 0), c->np - 1);
    }
    // GetRandomTrialPos (box region) + CheckPartBCTM (PBC)
    const double dt = dtran[ipt[ip]];
    double rn[3];
    for (int d = 0; d < 3; ++d) rn[d] = ro[3 * ip + d] + (rnd.next() - 0.5) * dt;
    for (int d = 0; d < 3; ++d) {
      const double L = c->P.boxlen[d];
      if (std::fabs(rn[d]) > 0.5 * L) rn[d] -= L * std::round(rn[d] / L);
    }
    int ovl = 0;
    double du = 0.0, drec = 0.0;
    const int ip1 = ip + 1;
    if (int e = dutotal_batch(c, 1, &ip1, rn, &ovl, &du, &drec, nullptr)) return e;
    // Metropolis (weight 1)
    bool accept;
    if (ovl) {
      accept = false;
    } else {
      const double dured = beta * du;
      if (dured > expmax) accept = false;
      else if (dured < -expmax) accept = true;
      else {
        const double fac = std::exp(-dured);
        accept = fac > 1.0 ? true : fac > rnd.next();
      }
    }
    if (events) events[t] = ovl ? -1 : (accept ? 1 : 0);
    if (accept) {
      if (int e = commit_move(c, ip, rn)) return e;
      for (int d = 0; d < 3; ++d) ro[3 * ip + d] = rn[d];
      dsum += du;
      ++nacc;
    }
  }
  if (du_sum) *du_sum = dsum;
  if (naccept) *naccept = nacc;
  return MSG_OK;
}

}  // extern "C"
