// Seeded synthetic systems: Philox-4x64-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) with Box-Muller
// normals.  Every element is a pure function of (seed, indices), so the same
// system is produced whatever the thread count.
//
//   Gamma[G,p,q] = Gamma[G,q,p] = z * exp(-G/100) / sqrt(NG), p <= q,
//     z = sqrt(-2 ln(1 - u1)) cos(2 pi u2), (u1, u2) from the first two words
//     of philox(ctr = (G, p, q, 0), key = (seed, 0));
//   eps_i = -2.0 + 1.5 u, u from philox((i, 0, 0, 1), key);
//   eps_a =  0.2 + 2.8 u, u from philox((a, 0, 0, 2), key);
//   u = (word >> 11) * 2^-53.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/sisi_engine.h"

namespace {

inline void mulhilo(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
  const unsigned __int128 p = static_cast<unsigned __int128>(a) * b;
  lo = static_cast<uint64_t>(p);
  hi = static_cast<uint64_t>(p >> 64);
}

void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t lo0, hi0, lo1, hi1;
    mulhilo(M0, c[0], lo0, hi0);
    mulhilo(M1, c[2], lo1, hi1);
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

inline double unit(uint64_t w) { return double(w >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" int sisi_generate_system_impl(uint64_t seed, int No, int Nv, int NG, double *gamma,
                                         double *eps_i, double *eps_a) {
  if (No <= 0 || Nv <= 0 || NG <= 0) return SISI_ERR_INVALID;
  const int64_t Np = int64_t(No) + Nv;
  const double two_pi = 6.283185307179586476925286766559;
  if (gamma) {
    // rows G are independent: split them over host threads
    auto rows = [&](int64_t g0, int64_t g1) {
      for (int64_t G = g0; G < g1; ++G) {
        const double scale = std::exp(-double(G) / 100.0) / std::sqrt(double(NG));
        for (int64_t q = 0; q < Np; ++q)
          for (int64_t p = 0; p <= q; ++p) {
            uint64_t c[4] = {uint64_t(G), uint64_t(p), uint64_t(q), 0};
            philox4x64_10(c, seed, 0);
            const double u1 = unit(c[0]), u2 = unit(c[1]);
            const double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
            const double v = z * scale;
            gamma[G + NG * (p + Np * q)] = v;
            gamma[G + NG * (q + Np * p)] = v;
          }
      }
    };
    const int64_t nt = std::max<int64_t>(
        1, std::min<int64_t>(NG, std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int64_t t = 0; t < nt; ++t) pool.emplace_back(rows, NG * t / nt, NG * (t + 1) / nt);
    for (auto &th : pool) th.join();
  }
  if (eps_i)
    for (int i = 0; i < No; ++i) {
      uint64_t c[4] = {uint64_t(i), 0, 0, 1};
      philox4x64_10(c, seed, 0);
      eps_i[i] = -2.0 + 1.5 * unit(c[0]);
    }
  if (eps_a)
    for (int a = 0; a < Nv; ++a) {
      uint64_t c[4] = {uint64_t(a), 0, 0, 2};
      philox4x64_10(c, seed, 0);
      eps_a[a] = 0.2 + 2.8 * unit(c[0]);
    }
  return SISI_OK;
}
