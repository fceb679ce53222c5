// ref_rng.cpp -- replays the reference's own input vector (test infrastructure).
// dawgsMain.cpp:326-330 fills q with std::mt19937(rank) and
// std::uniform_real_distribution<>(-0.25, 0.25); under libstdc++ this
// regenerates the same doubles bit for bit.
#include <cstdint>
#include <random>

#include "ogs_oracle.h"

extern "C" void oracle_reference_q(uint32_t seed, dlong n, double *q) {
  std::mt19937 RNG = std::mt19937(seed);
  std::uniform_real_distribution<> distrib(-0.25, 0.25);
  for (dlong i = 0; i < n; i++) q[i] = distrib(RNG);
}
