// Host-side QMC and RNG tables shared by the render and photon passes.
#pragma once

#include <stdint.h>

#include <vector>

// Values of Halton(base)::getNext() for the first n calls (table[k] = call k+1).
// The reference generator is incremental with a 1e-10 carry tolerance
// (mcqmc.h:35-49); near powers of the base that tolerance makes it differ from
// the exact radical inverse by up to ~100 ulp, so the device reads these
// tables instead of recomputing radical inverses.
void halton_incremental(int base, uint64_t n, std::vector<float>& out);

// nextPrime (mcqmc.h:84-97)
int next_prime(int last);

// glibc TYPE_3 rand() outputs after srand(seed): the generator behind
// blockSpliter_t's shuffle (renderblock.cc:82).
void glibc_rand(uint32_t seed, int64_t n, std::vector<int32_t>& out);

// 16807^(2^k) mod (2^31-1), k = 0..63: jump-ahead table for ourRandom()
// (vector3d.h:276-297).
void lcg_pow_table(uint64_t out[64]);
