// Device building blocks shared by the kernels of liboptlab_mi: Philox4x32-10, Box-Muller, the per-pair normal
// streams, math in the path precision R (float or double), and wave64 reductions.
//
// Draws follow packages/native/cuda/include/optlab_cuda/philox.hpp:22-83 and paths.hpp:207-233: the integer stream
// is bit-exact to the reference; the double-precision Box-Muller takes the reference CPU form cos/sin(2 pi v), the
// float one the sine and cosine of 2 pi v that the reference GPU took with sincospi(2 v).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace optlab_mi {

struct U4 {
    std::uint32_t x, y, z, w;
};

// Each round's two 32 x 32 -> 64 products are formed as 64-bit products, which the compiler maps to one
// v_mad_u64_u32 each instead of a v_mul_hi_u32 / v_mul_lo_u32 pair.
__device__ __forceinline__ U4 philox10(U4 c, std::uint32_t k0, std::uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const std::uint64_t p0 = static_cast<std::uint64_t>(c.x) * 0xD2511F53u;
        const std::uint64_t p1 = static_cast<std::uint64_t>(c.z) * 0xCD9E8D57u;
        const std::uint32_t hi0 = static_cast<std::uint32_t>(p0 >> 32), lo0 = static_cast<std::uint32_t>(p0);
        const std::uint32_t hi1 = static_cast<std::uint32_t>(p1 >> 32), lo1 = static_cast<std::uint32_t>(p1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// ---- kernel-constant reads ------------------------------------------------------------------------------------
// Element i of an array that no kernel writes while the reading kernel runs (request data, tables built by an
// earlier launch), read through the constant address space. With a wave-uniform index the compiler issues a scalar
// load whose wait counts scalar and LDS traffic only; the same read through a generic pointer is a vector load,
// because the compiler cannot prove the kernel's own stores leave it alone, and waiting on a vector load also waits
// for every store the lane has pending. With a divergent index it is an ordinary vector load.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, long long i) {
    typedef const __attribute__((address_space(4))) T* ConstPtr;
    return ((ConstPtr)p)[i];
}

// ---- math in the path precision ---------------------------------------------------------------------------------
__device__ __forceinline__ float r_exp(float x) { return expf(x); }
__device__ __forceinline__ double r_exp(double x) { return exp(x); }
__device__ __forceinline__ float r_log(float x) { return logf(x); }
__device__ __forceinline__ double r_log(double x) { return log(x); }
// FP32 square root: the hardware v_sqrt_f32 (1 ulp; denormals flushed), as -prec-sqrt=false gave the reference's
// GPU build. Named explicitly: sqrtf carries its accuracy as metadata, which the optimizer drops where it merges
// the calls of several branches, and without it the compiler expands the correctly rounded sequence (an rsq and
// Newton steps, about 15 instructions).
__device__ __forceinline__ float r_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float r_erfc(float x) { return erfcf(x); }

// FP32 reciprocal, natural logarithm and exponential by the hardware instructions (v_rcp_f32, v_log_f32 = log2,
// v_exp_f32 = exp2; about 1 ulp each, 2 with the scaling by ln 2), the accuracy the reference's FP32 GPU build asked
// for (-prec-div=false). The library forms add range handling that these call sites never need: denominators far
// below 2^96 in magnitude, logarithms of positive finite numbers, exponents within the normal range (smaller results
// flush to zero, as FTZ does anyway). logf expands to 7 instructions and a division to 5; these take 2 and 2.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
__device__ __forceinline__ float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ __forceinline__ double r_erfc(double x) { return erfc(x); }
template <typename R>
__device__ __forceinline__ R r_max(R a, R b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R r_min(R a, R b) { return a < b ? a : b; }

// (i + 1/2) 2^-32 in (0, 1), in the precision R (philox.hpp:45-49).
template <typename R>
__device__ __forceinline__ R unit01(std::uint32_t i) {
    return static_cast<R>(i) * static_cast<R>(2.3283064365386963e-10) + static_cast<R>(1.1641532182693481e-10);
}

// The FP32 pair takes the angle's sine and cosine from the hardware (v_sin_f32 / v_cos_f32 take their argument in
// revolutions, so they give sin and cos of 2 pi v, the reference GPU's sincospi(2 v), directly). Measured on gfx950
// over 2^20 draws: 1.2e-7 largest and 2.5e-8 median absolute error (sincospif: 5e-8 and 1e-8), for 2 instructions
// instead of about 40.
__device__ __forceinline__ void gauss2(std::uint32_t a, std::uint32_t b, float& z0, float& z1) {
    const float u = unit01<float>(a), v = unit01<float>(b);
    const float radius = r_sqrt(-2.0f * hw_log(u));
    z0 = radius * __builtin_amdgcn_cosf(v);
    z1 = radius * __builtin_amdgcn_sinf(v);
}

__device__ __forceinline__ void gauss2(std::uint32_t a, std::uint32_t b, double& z0, double& z1) {
    const double u = unit01<double>(a), v = unit01<double>(b);
    const double radius = sqrt(-2.0 * log(u));
    const double angle = 6.283185307179586 * v;
    z0 = radius * cos(angle);
    z1 = radius * sin(angle);
}

constexpr std::uint32_t kAuxTag = 0x80000000u;
// The stochastic mixing weight's own driver: a third stream of the pair, so that the spot normals of a basket are
// those of the constant matrix whatever the weight does.
constexpr std::uint32_t kMixTag = 0x40000000u;

// Normals of one pair on one stream, four per Philox block (paths.hpp:213-233). The consumption pattern is the
// same on every lane of a launch, so `used` and `block` stay wave-uniform.
template <typename R>
struct NormalStream {
    std::uint32_t p0, p1, k0, k1, tag;
    std::uint32_t block;
    int used;
    R buf[4];

    __device__ __forceinline__ void init(std::uint64_t pair, std::uint32_t key0, std::uint32_t key1, std::uint32_t t) {
        p0 = static_cast<std::uint32_t>(pair);
        p1 = static_cast<std::uint32_t>(pair >> 32);
        k0 = key0;
        k1 = key1;
        tag = t;
        block = 0;
        used = 4;
    }
    __device__ __forceinline__ R next() {
        if (used == 4) {
            const U4 r = philox10(U4{p0, p1, block, tag}, k0, k1);
            gauss2(r.x, r.y, buf[0], buf[1]);
            gauss2(r.z, r.w, buf[2], buf[3]);
            ++block;
            used = 0;
        }
        const int u = used++;
        return u == 0 ? buf[0] : u == 1 ? buf[1] : u == 2 ? buf[2] : buf[3];
    }
};

// Normals of one pair on one stream drawn by the pair's two lanes together, for kernels that carry one leg a lane
// (the legs of pair p on lanes 2p and 2p + 1 of a wave; both lanes must be active whenever one draws). Each refill
// evaluates Philox and Box-Muller once a lane, for two consecutive blocks -- block b on the even lane, b + 1 on the odd
// one -- and a quad permutation hands each lane the other's four normals: eight normals a refill for the work of
// four, the same values as NormalStream block by block, since each block's arithmetic is the same wherever it runs.
template <typename R>
struct PairNormalStream {
    std::uint32_t p0, p1, k0, k1, tag;
    std::uint32_t block;
    int used;
    R buf[8];

    __device__ __forceinline__ void init(std::uint64_t pair, std::uint32_t key0, std::uint32_t key1, std::uint32_t t) {
        p0 = static_cast<std::uint32_t>(pair);
        p1 = static_cast<std::uint32_t>(pair >> 32);
        k0 = key0;
        k1 = key1;
        tag = t;
        block = 0;
        used = 8;
    }
    // value v of lane 2q + L of each quad pair, on both of its lanes (DPP quad_perm [L, L, 2 + L, 2 + L])
    template <int L>
    __device__ __forceinline__ static float from_lane(float v) {
        constexpr int kPerm = L == 0 ? 0xA0 : 0xF5;
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kPerm, 0xF, 0xF,
                                                                      false));
    }
    template <int L>
    __device__ __forceinline__ static double from_lane(double v) {
        constexpr int kPerm = L == 0 ? 0xA0 : 0xF5;
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), kPerm, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), kPerm, 0xF, 0xF, false);
        return __builtin_bit_cast(double, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                              static_cast<unsigned>(lo));
    }
    __device__ __forceinline__ R next() {
        if (used == 8) {
            const std::uint32_t odd = threadIdx.x & 1u;
            // the keys made opaque here, so that the compiler forms the round keys at each refill (scalar adds)
            // instead of holding all twenty in scalar registers across the kernel
            std::uint32_t r0 = k0, r1 = k1;
            asm volatile("" : "+s"(r0), "+s"(r1));
            const U4 r = philox10(U4{p0, p1, block + odd, tag}, r0, r1);
            R own[4];
            gauss2(r.x, r.y, own[0], own[1]);
            gauss2(r.z, r.w, own[2], own[3]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                buf[q] = from_lane<0>(own[q]);
                buf[4 + q] = from_lane<1>(own[q]);
            }
            block += 2;
            used = 0;
        }
        const int u = used++;
        R v = buf[0];
#pragma unroll
        for (int q = 1; q < 8; ++q)
            if (u == q) v = buf[q];
        return v;
    }
};

// ---- wave64 reductions -----------------------------------------------------------------------------------------

__device__ __forceinline__ double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }

// Reduce-scatter of 16 doubles over the 64 lanes of a wave: on return, v[0] of lane l holds the wave total of
// value (l & 15) (every group of 16 lanes holds the same totals). Two all-reduce levels (xor 32, 16) then four
// halving levels (xor 8, 4, 2, 1): 47 shuffles and adds for 16 values instead of 96 for one all-reduce each.
__device__ __forceinline__ void reduce_scatter16(double v[16], int lane) {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] += shfl_xor_d(v[i], 32);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] += shfl_xor_d(v[i], 16);
#pragma unroll
    for (int half = 8; half >= 1; half >>= 1) {
        const bool upper = (lane & half) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const double keep = upper ? v[i + half] : v[i];
            const double give = upper ? v[i] : v[i + half];
            v[i] = keep + shfl_xor_d(give, half);
        }
    }
}

// Neumaier-compensated FP64 accumulator: s + c is the sum.
struct Comp {
    double s, c;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

}  // namespace optlab_mi
