// rng.cpp -- seeded, counter-based generator of test matrices.
//
// The reference tests draw their matrices with DLARNV/DLARUV, a sequential
// 48-bit congruential generator (SRC/dlarnv.f:94-172, SRC/dlaruv.f).  Large
// GPU inputs need something that can be produced in parallel and reproduced
// bit-for-bit by an independent implementation (the test suite's numpy copy),
// so this generator is its own design:
//   r(key, idx, attempt) = mix64(key ^ mix64(idx*C + attempt)),  key = mix64(seed)
// with mix64 the splitmix64 finaliser, u = (r >> 11) * 2^-53 in [0,1).
//   idist 1: u;   idist 2: 2u - 1 in [-1,1);
//   idist 3: Marsaglia's polar method on attempts (2a, 2a+1), first accepted
//            pair, z = u*sqrt(-2 ln s / s) with ln computed by an
//            arithmetic-only series (frexp split + atanh series), so no libm
//            transcendental enters and every IEEE-754 host gives the same bits.
// Compiled with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/lapack_amd.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t draw(uint64_t key, uint64_t idx, uint64_t attempt) {
  return mix64(key ^ mix64(idx * 0xD1B54A32D192ED03ull + attempt));
}
inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1p-53; }

// ln(s) for s in (0, 1]: s = m*2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(t), t = (m-1)/(m+1), |t| < 0.1716, series to t^21.
double det_log(double s) {
  int e = 0;
  double m = std::frexp(s, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double t = (m - 1.0) / (m + 1.0);
  const double t2 = t * t;
  double p = 1.0 / 21.0;
  p = p * t2 + 1.0 / 19.0;
  p = p * t2 + 1.0 / 17.0;
  p = p * t2 + 1.0 / 15.0;
  p = p * t2 + 1.0 / 13.0;
  p = p * t2 + 1.0 / 11.0;
  p = p * t2 + 1.0 / 9.0;
  p = p * t2 + 1.0 / 7.0;
  p = p * t2 + 1.0 / 5.0;
  p = p * t2 + 1.0 / 3.0;
  p = p * t2 + 1.0;
  const double lnm = 2.0 * t * p;
  const double ln2_hi = 0x1.62e42fefa3800p-1, ln2_lo = 0x1.ef35793c76730p-45;
  const double de = (double)e;
  return de * ln2_hi + (lnm + de * ln2_lo);
}

double one(int idist, uint64_t key, uint64_t idx) {
  if (idist == 1) return u01(draw(key, idx, 0));
  if (idist == 2) return 2.0 * u01(draw(key, idx, 0)) - 1.0;
  for (uint64_t a = 0;; a += 2) {
    const double u = 2.0 * u01(draw(key, idx, a)) - 1.0;
    const double v = 2.0 * u01(draw(key, idx, a + 1)) - 1.0;
    const double uu = u * u;
    const double vv = v * v;
    const double s = uu + vv;
    if (s > 0.0 && s < 1.0) {
      const double l = det_log(s);
      const double q = (-2.0 * l) / s;
      return u * std::sqrt(q);
    }
  }
}

}  // namespace

extern "C" void lapack_amd_dlarnv(int idist, unsigned long long seed, unsigned long long offset, long long n,
                                  double* x) {
  if (n <= 0 || idist < 1 || idist > 3) return;
  const uint64_t key = mix64((uint64_t)seed);
  auto work = [&](long long b, long long e) {
    for (long long k = b; k < e; ++k) x[k] = one(idist, key, (uint64_t)offset + (uint64_t)k);
  };
  unsigned nt = std::thread::hardware_concurrency();
  if (nt > 16) nt = 16;
  if (nt < 1) nt = 1;
  if (n < (1 << 18) || nt == 1) {
    work(0, n);
    return;
  }
  std::vector<std::thread> th;
  const long long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const long long b = (long long)t * chunk, e = b + chunk < n ? b + chunk : n;
    if (b < e) th.emplace_back(work, b, e);
  }
  for (auto& t : th) t.join();
}
