// CPU ORACLE -- test infrastructure only, never part of the product.
//
// A plain C++17 restatement, in double precision, of the reference engine's CPU algorithms. Only the tests, the
// smoke check and the CPU-baseline leg of the benchmark load it (as liboracle.so through oracle.h); the HIP
// library never links it.
//
//   Philox4x32-10, to_unit, Box-Muller (CPU form) .. packages/native/cuda/include/optlab_cuda/philox.hpp:22-83
//   NormalStream and the aux stream tag ........... paths.hpp:207-233
//   table_value, record, simulate_pair ............ paths.hpp:196-361
//   cpu_paths (blocks of 512 pairs, threads) ...... src/paths_cpu.cpp:14-52
//   QE-M constants, qe_step, path units ........... schemes.hpp:159-311
//   estimate_from_sums ............................ src/cpu_mc.cpp:14-18
//   particle leverage (host algorithm + CPU bins) . src/leverage.cpp:17-194
//   geometric Asian closed form ................... src/closed_forms.cpp:21-33
//
// Where the reference leaves a choice open the oracle fixes it and the HIP engine follows (each is pinned in the
// tests): the fixed products read Philox stream 0 of pair p (the reference GPU used cuRAND, whose mapping is not
// reproducible); partial sums are combined in a fixed block order, so results do not depend on the thread count.
#pragma once

#include <cstdint>

namespace oracle {

struct Words {
    std::uint32_t x, y, z, w;
};

// philox.hpp:22-43: ten rounds of two 32x32->64 multiplies, key bumped by the Weyl constants after each round.
inline Words philox(Words c, std::uint32_t k0, std::uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        const std::uint64_t p0 = static_cast<std::uint64_t>(0xD2511F53u) * c.x;
        const std::uint64_t p1 = static_cast<std::uint64_t>(0xCD9E8D57u) * c.z;
        const std::uint32_t hi0 = static_cast<std::uint32_t>(p0 >> 32), lo0 = static_cast<std::uint32_t>(p0);
        const std::uint32_t hi1 = static_cast<std::uint32_t>(p1 >> 32), lo1 = static_cast<std::uint32_t>(p1);
        c = Words{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// philox.hpp:45-49: (i + 1/2) 2^-32, in (0, 1).
inline double unit(std::uint32_t i) { return static_cast<double>(i) * 2.3283064365386963e-10 + 1.1641532182693481e-10; }

// philox.hpp:68-72, the CPU form: radius sqrt(-2 ln u), angle 2 pi v with 2 pi rounded to double.
void gauss_pair(std::uint32_t a, std::uint32_t b, double& z0, double& z1);

constexpr std::uint32_t kAuxTag = 0x80000000u;  // paths.hpp:209
// The stochastic mixing weight's driver (no counterpart in the reference engine, which refuses that mixer): a
// third stream of the pair, so that the spot normals are those of the constant matrix whatever the weight does.
constexpr std::uint32_t kMixTag = 0x40000000u;

// paths.hpp:213-233: the normals of one pair on one stream, four per Philox block, none wasted.
struct Draws {
    std::uint64_t pair;
    std::uint32_t k0, k1;
    std::uint32_t tag = 0;
    std::uint64_t block = 0;
    double buf[4] = {0, 0, 0, 0};
    int used = 4;
    double next();
};

}  // namespace oracle
