// Fixed products on one underlying -- Black-Scholes exact log steps or Heston by Andersen's QE-M -- and strips of
// vanillas, with every payoff summed on the device in compensated FP64.
// Semantics: packages/native/cuda/include/optlab_cuda/schemes.hpp:159-311 (make_consts, qe_step, path units),
// src/cpu_mc.cpp:14-18 (estimates on pair averages). Pair p reads the normals of Philox stream 0 of pair p
// (paths.hpp:213-233), two per Heston step (Zv then Zx) and one per Black-Scholes step.
//
// MI355X layout: one antithetic pair per lane per iteration of a grid-stride loop, path state in VGPRs (FP32 or
// FP64). At each payoff event the wave's pair averages (widened to FP64) and their squares are reduced across the
// 64 lanes by reduce-scatters of 16 values, so that lane l ends up owning output slot l and keeps it in a
// Neumaier-compensated FP64 register pair. Each wave writes its 64 slots once at the end; a second kernel adds the
// waves in a fixed order. No atomics: results are deterministic for a given launch geometry.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

// Device memory of the allocation-free strip (optlab_strip_noalloc): file scope, one copy on every device the
// library runs on, held until the process ends. Laid out per call as [slot table][totals][chunk sums][per-block
// rows]. Host code never passes this name to a kernel: it is a host-side stand-in, and the device address is what
// hipGetSymbolAddress returns for the current device (noalloc_base below).
constexpr size_t kStripNoAllocBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kStripNoAllocBytes];

namespace optlab_mi {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxStrikes = 31;
constexpr int kMaxEvents = 16;
#ifndef OPTLAB_STRIP_WAVES
#define OPTLAB_STRIP_WAVES 1
#endif
constexpr int kStripWavesF32 = OPTLAB_STRIP_WAVES;

template <typename R>
struct Consts {  // schemes.hpp:159-207, computed in double on the host and cast once
    R S0, D, bs_drift, bs_vol, ln_h, K[3];
    int leg[3], n_steps;
    R inv_n;
    R v0, theta, mu_dt, E, c1, c2, K1, K2, K3, K4, A, K0p, psi_c;
    // theta (1 - E), 2 A and K1 + K3 / 2 for the FP32 step: the scalar unit has no float arithmetic, so a float
    // expression of kernel arguments formed in the step would cost a vector instruction every step
    R th1E, twoA, k1h;
    // the derivatives in sigma of the Black-Scholes step's drift and volatility (-sigma dt, sqrt dt): the tangent of
    // the path along sigma in the strip's sensitivities
    R g_drift, g_vol;
};

struct HostConsts {
    double S0, D, bs_drift, bs_vol, ln_h, K[3];
    int leg[3], n_steps;
    double inv_n, v0, theta, mu_dt, E, c1, c2, K1, K2, K3, K4, A, K0p, psi_c;
    double g_drift, g_vol;
    template <typename R>
    Consts<R> as() const {
        Consts<R> c{};
        c.S0 = R(S0), c.D = R(D), c.bs_drift = R(bs_drift), c.bs_vol = R(bs_vol), c.ln_h = R(ln_h);
        for (int k = 0; k < 3; ++k) c.K[k] = R(K[k]), c.leg[k] = leg[k];
        c.n_steps = n_steps, c.inv_n = R(inv_n), c.v0 = R(v0), c.theta = R(theta), c.mu_dt = R(mu_dt), c.E = R(E);
        c.c1 = R(c1), c.c2 = R(c2), c.K1 = R(K1), c.K2 = R(K2), c.K3 = R(K3), c.K4 = R(K4), c.A = R(A);
        c.K0p = R(K0p), c.psi_c = R(psi_c);
        c.th1E = R(theta * (1 - E)), c.twoA = R(2 * A), c.k1h = R(K1 + 0.5 * K3);
        c.g_drift = R(g_drift), c.g_vol = R(g_vol);
        return c;
    }
};

HostConsts make_consts(int n_steps, double S0, double r, double q, double T, double sigma, double v0, double kappa,
                       double theta, double xi, double rho, double H, const int* legs, const double* strikes,
                       double psi_c) {
    const double dt = T / n_steps;
    HostConsts c{};
    c.S0 = S0;
    c.D = std::exp(-r * T);
    c.bs_drift = (r - q - 0.5 * sigma * sigma) * dt;
    c.bs_vol = sigma * std::sqrt(dt);
    c.g_drift = -sigma * dt;
    c.g_vol = std::sqrt(dt);
    c.ln_h = H > 0 ? std::log(H / S0) : 0.0;
    for (int k = 0; k < 3; ++k) c.K[k] = strikes ? strikes[k] : 0.0, c.leg[k] = legs ? legs[k] : 0;
    c.n_steps = n_steps;
    c.inv_n = 1.0 / n_steps;
    const double E = std::exp(-kappa * dt), xi2 = xi * xi;
    const double k12 = 0.5 * dt * (kappa * rho / xi - 0.5);
    c.v0 = v0, c.theta = theta, c.mu_dt = (r - q) * dt, c.E = E;
    c.c1 = xi2 * E * (1 - E) / kappa;
    c.c2 = theta * xi2 * (1 - E) * (1 - E) / (2 * kappa);
    c.K1 = k12 - rho / xi;
    c.K2 = k12 + rho / xi;
    c.K3 = 0.5 * dt * (1 - rho * rho);
    c.K4 = c.K3;
    c.A = (k12 + rho / xi) + 0.25 * dt * (1 - rho * rho);
    c.K0p = -rho * kappa * theta * dt / xi;
    c.psi_c = psi_c;
    return c;
}

template <typename R>
__device__ __forceinline__ R leg_payoff(int leg, R S, R K) {  // schemes.hpp:209-217
    return leg == 1 ? r_max(S - K, R(0)) : leg == 2 ? r_max(K - S, R(0)) : leg == 3 ? S - K : R(0);
}

// schemes.hpp:219-252: one QE-M step of (y, v); the martingale correction's log is taken once after the branch.
template <typename R>
__device__ __forceinline__ void qe_step(const Consts<R>& c, R& y, R& v, R zv, R zx) {
    const R m = c.theta + (v - c.theta) * c.E;
    const R psi = (v * c.c1 + c.c2) / (m * m);
    R vn, lin, w, arg;
    bool finite;
    if (psi <= c.psi_c) {
        const R inv = R(2) / psi;
        const R b2 = inv - R(1) + r_sqrt(inv) * r_sqrt(r_max(inv - R(1), R(0)));
        const R a = m / (R(1) + b2), b = r_sqrt(b2);
        vn = a * (b + zv) * (b + zv);
        arg = R(1) - R(2) * c.A * a;
        finite = c.A < R(1) / (R(2) * a);
        lin = -c.A * b2 * a / arg;
        w = R(0.5);
    } else {
        const R p = (psi - R(1)) / (psi + R(1)), beta = (R(1) - p) / m;
        // 1 - U = Phi(-Zv) = erfc(Zv / sqrt 2) / 2, exact in both tails
        const R one_minus_u = R(0.5) * r_erfc(zv * R(0.70710678118654752));
        vn = one_minus_u >= R(1) - p ? R(0) : r_log((R(1) - p) / one_minus_u) / beta;
        arg = p + beta * (R(1) - p) / (beta - c.A);
        finite = c.A < beta;
        lin = R(0);
        w = R(-1);
    }
    const R K0 = finite ? lin + w * r_log(arg) - (c.K1 + R(0.5) * c.K3) * v : c.K0p;
    y += c.mu_dt + K0 + c.K1 * v + c.K2 * vn + r_sqrt(r_max(c.K3 * v + c.K4 * vn, R(0))) * zx;
    v = vn;
}

// erfc by the Chebyshev fit of Numerical Recipes in C (2nd ed., section 6.2, "erfcc"): fractional error below
// 1.2e-7 on the whole line, one reciprocal and one exponential.
__device__ __forceinline__ float erfc_fast(float x) {
    const float z = fabsf(x), t = hw_rcp(1.0f + 0.5f * z);
    const float poly = -1.26551223f + t * (1.00002368f + t * (0.37409196f + t * (0.09678418f + t * (-0.18628806f +
                       t * (0.27886807f + t * (-1.13520398f + t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
    const float r = t * hw_exp(-z * z + poly);
    return x >= 0.0f ? r : 2.0f - r;
}

// The FP32 step: the scheme of qe_step with the divisions shared between the quantities that need them. With
// s2 = v c1 + c2 and m2 = m^2 (psi = s2 / m2): the branch tests s2 <= psi_c m2; the quadratic branch takes
// 2 / psi = 2 m2 / s2 and sqrt(inv) sqrt(inv - 1) as one root; A < 1 / (2 a) is arg = 1 - 2 A a > 0; the
// exponential branch takes p, 1 - p and beta from the one reciprocal of s2 + m2. Four divisions and three roots
// where the literal transcription has seven and four; the FP64 kernels keep the reference's order of operations.
//
// Both legs of a pair take one step together (legs 0 and 1 take normals (zv, zx) and (-zv, -zx)). On the strip's
// Heston workload about one leg-step in six takes the exponential branch (psi > psi_c), but nearly every wave has
// lanes on both sides, so a per-leg step ran both branches for each leg. Here the quadratic branch runs per leg,
// the exponential branch once, on the leg of each lane that needs it (leg 0 when both do), and a second time only
// for lanes whose two legs both need it (about one wave in a hundred has such a lane): each leg gets the same
// arithmetic, and a pair-step issues one exponential branch less.
struct QeBranch {
    float vn, lin, w, arg;
    bool finite;
};

__device__ __forceinline__ QeBranch qe_quadratic(const Consts<float>& c, float m, float s2, float m2, float zv) {
    QeBranch o;
    const float inv = 2.0f * m2 * hw_rcp(s2);
    const float b2 = inv - 1.0f + r_sqrt(inv * fmaxf(inv - 1.0f, 0.0f));
    const float a = m * hw_rcp(1.0f + b2), b = r_sqrt(b2);
    o.vn = a * (b + zv) * (b + zv);
    o.arg = 1.0f - c.twoA * a;
    o.finite = o.arg > 0.0f;
    o.lin = -c.A * b2 * a * hw_rcp(o.arg);
    o.w = 0.5f;
    return o;
}

__device__ __forceinline__ QeBranch qe_exponential(const Consts<float>& c, float m, float s2, float m2, float zv) {
    QeBranch o;
    const float den = hw_rcp(s2 + m2);
    const float p = (s2 - m2) * den, q = 2.0f * m2 * den, beta = 2.0f * m * den;
    const float one_minus_u = 0.5f * erfc_fast(zv * 0.70710678f);
    o.vn = one_minus_u >= q ? 0.0f : hw_log(q * hw_rcp(one_minus_u)) * (0.5f * (s2 + m2) * hw_rcp(m));
    o.arg = p + beta * q * hw_rcp(beta - c.A);
    o.finite = c.A < beta;
    o.lin = 0.0f;
    o.w = -1.0f;
    return o;
}

__device__ __forceinline__ void qe_tail(const Consts<float>& c, float& y, float& v, const QeBranch& o, float zx) {
    const float K0 = o.finite ? o.lin + o.w * hw_log(o.arg) - c.k1h * v : c.K0p;
    y += c.mu_dt + K0 + c.K1 * v + c.K2 * o.vn + r_sqrt(fmaxf(c.K3 * v + c.K4 * o.vn, 0.0f)) * zx;
    v = o.vn;
}

__device__ __forceinline__ void qe_pair(const Consts<float>& c, float (&y)[2], float (&v)[2], float zv, float zx) {
    float m[2], s2[2], m2[2];
    bool ex[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        m[a] = v[a] * c.E + c.th1E;  // theta + (v - theta) E
        s2[a] = v[a] * c.c1 + c.c2;
        m2[a] = m[a] * m[a];
        ex[a] = !(s2[a] <= c.psi_c * m2[a]);
    }
    // The quadratic branch for both legs in every lane (a wave almost always has a lane that needs it, so a
    // condition would not skip it; without one there are no defaults to set and no exec mask to save): a lane past
    // psi_c overwrites it with the exponential branch below, and selects drop whatever the quadratic made there.
    QeBranch o[2] = {qe_quadratic(c, m[0], s2[0], m2[0], zv), qe_quadratic(c, m[1], s2[1], m2[1], -zv)};
    if (ex[0] || ex[1]) {
        const bool first = ex[0];  // the leg this pass serves: 0 when leg 0 needs it, else 1
        const QeBranch r = qe_exponential(c, first ? m[0] : m[1], first ? s2[0] : s2[1], first ? m2[0] : m2[1],
                                          first ? zv : -zv);
        if (first)
            o[0] = r;
        else
            o[1] = r;
    }
    if (ex[0] && ex[1]) o[1] = qe_exponential(c, m[1], s2[1], m2[1], -zv);
    qe_tail(c, y[0], v[0], o[0], zx);
    qe_tail(c, y[1], v[1], o[1], -zx);
}

// The pair step from code templated on the precision: the FP32 one above; FP64 keeps the reference's per-leg
// order of operations (qe_step<double>), so this overload is never called.
__device__ __forceinline__ void qe_pair_any(const Consts<float>& c, float (&y)[2], float (&v)[2], float zv, float zx) {
    qe_pair(c, y, v, zv, zx);
}
__device__ __forceinline__ void qe_pair_any(const Consts<double>&, double (&)[2], double (&)[2], double, double) {}

// schemes.hpp:254-311: European / Asian / down-and-out (+ up-and-out) pair averages; out[4] = out[0] out[1].
template <typename R, int MODEL, int PRODUCT>
__device__ __forceinline__ void fixed_unit(const Consts<R>& c, NormalStream<R>& src, R out[5]) {
    R o[4] = {R(0), R(0), R(0), R(0)};
    if (MODEL == 0 && PRODUCT == 0) {
        const R z = src.next();
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const R S = c.S0 * r_exp(c.bs_drift + c.bs_vol * (a == 0 ? z : -z));
#pragma unroll
            for (int k = 0; k < 3; ++k) o[k] += c.D * leg_payoff(c.leg[k], S, c.K[k]);
            o[3] += S;
        }
    } else {
        R y[2] = {R(0), R(0)}, v[2] = {c.v0, c.v0}, sum_s[2] = {R(0), R(0)}, sum_y[2] = {R(0), R(0)};
        bool alive[2] = {true, true};
        auto one_step = [&]() __attribute__((always_inline)) {
            R zv = R(0), zx = R(0), z = R(0);
            if (MODEL == 1) {
                zv = src.next();
                zx = src.next();
            } else {
                z = src.next();
            }
            if (MODEL == 1 && std::is_same<R, float>::value) qe_pair_any(c, y, v, zv, zx);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const R sg = a == 0 ? R(1) : R(-1);
                if (MODEL == 1) {
                    if (!std::is_same<R, float>::value) qe_step(c, y[a], v[a], sg * zv, sg * zx);
                } else {
                    y[a] += c.bs_drift + c.bs_vol * sg * z;
                }
                if (PRODUCT == 1) {
                    sum_s[a] += r_exp(y[a]);
                    sum_y[a] += y[a];
                }
                if (PRODUCT == 2) alive[a] = alive[a] && !(y[a] <= c.ln_h);
                if (PRODUCT == 3) alive[a] = alive[a] && !(y[a] >= c.ln_h);
            }
        };
        // Black-Scholes: the four steps a Philox block serves run together, so the stream's buffer position is a
        // compile-time constant at every draw (measured: the down-and-out 1.12 -> 0.96 ms, the same price; two Heston
        // steps a pass moved its times by -3.5 % and +0.7 % and, contracted differently, its FP32 bits)
        int i = 0;
        if constexpr (MODEL == 0) {
            for (; i + 4 <= c.n_steps; i += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) one_step();
            }
        }
        for (; i < c.n_steps; ++i) one_step();
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const R S = c.S0 * r_exp(y[a]);
            if (PRODUCT == 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k) o[k] += c.D * leg_payoff(c.leg[k], S, c.K[k]);
            } else if (PRODUCT == 1) {
                o[0] += c.D * leg_payoff(c.leg[0], c.S0 * sum_s[a] * c.inv_n, c.K[0]);
                o[1] += c.D * leg_payoff(c.leg[0], c.S0 * r_exp(sum_y[a] * c.inv_n), c.K[0]);
            } else {
                const R pay = c.D * leg_payoff(c.leg[0], S, c.K[0]);
                o[0] += alive[a] ? pay : R(0);
                o[1] += pay;
            }
            o[3] += S;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = R(0.5) * o[k];
    out[4] = out[0] * out[1];
}

// Slots of a fixed product: 0..4 the sums of out[0..4], 5..9 the sums of their squares (in FP64).
template <typename R, int MODEL, int PRODUCT>
__global__ void __launch_bounds__(kBlock) fixed_kernel(Consts<R> c, long long n_pairs, std::uint32_t k0,
                                                        std::uint32_t k1, double* partials) {
    const int lane = threadIdx.x & 63;
    // each lane's own pairs in FP64 (a few to a few dozen terms a lane), reduced over the wave once at the end: a
    // wave reduction per pair cost more than a one-step product itself
    double sum[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) sum[k] = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long base = static_cast<long long>(blockIdx.x) * blockDim.x; base < n_pairs; base += stride) {
        const long long p = base + threadIdx.x;
        const bool active = p < n_pairs;
        NormalStream<R> src;
        src.init(static_cast<std::uint64_t>(active ? p : 0), k0, k1, 0u);
        R out[5];
        fixed_unit<R, MODEL, PRODUCT>(c, src, out);
        if (active) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double yk = static_cast<double>(out[k]);
                sum[k] += yk;
                sum[5 + k] += yk * yk;
            }
        }
    }
    double vals[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) vals[k] = k < 10 ? sum[k] : 0.0;
    reduce_scatter16(vals, lane);
    const int v = lane & 15;
    Comp acc{lane < 16 && v < 10 ? vals[0] : 0.0, 0.0};
    // the block's waves in a fixed order: one row of 16 partial sums per block
    __shared__ double part_s[kBlock / 64][16], part_c[kBlock / 64][16];
    const int wib = threadIdx.x >> 6;
    if (lane < 16) part_s[wib][lane] = acc.s, part_c[wib][lane] = acc.c;
    __syncthreads();
    if (threadIdx.x < 16) {
        Comp a{part_s[0][lane], part_c[0][lane]};
        for (int w = 1; w < kBlock / 64; ++w) {
            a.add(part_s[w][lane]);
            a.c += part_c[w][lane];
        }
        partials[static_cast<long long>(blockIdx.x) * 16 + lane] = a.value();
    }
}

// The strip: events at mat_step[e] and 32 value slots per event. Slot k < n_strikes is the discounted
// pair-average payoff of strike k, slot n_strikes the pair average of S (a call struck at 0, undiscounted), the
// rest zero. Each slot is D max(sign (S - K), 0) averaged over the two legs, with (K, D, sign) from a per-event
// table, so the unrolled payoff code carries no per-slot conditions. Output slots of event e: 0..31 the sums of the
// values, 32..63 the sums of their squares.
template <typename R>
struct StripSlot {
    R K, D, sign;
};

// Bytes of the largest (FP64) device table of a strip: (K, D, sign) per event and slot, and the event steps.
constexpr size_t kStripTableBytes = kMaxEvents * 32 * 3 * sizeof(double) + kMaxEvents * sizeof(int);

template <typename R>
struct StripArgs {
    Consts<R> c;
    int model, put, n_strikes, n_mats;
    const StripSlot<R>* slots;  // n_mats x 32
    const int* mat_step;        // n_mats
};

template <typename R>
__device__ __forceinline__ R strip_value(const StripSlot<R>& t, R S0, R S1) {
    const R o = t.D * r_max(t.sign * (S0 - t.K), R(0)) + t.D * r_max(t.sign * (S1 - t.K), R(0));
    return R(0.5) * o;
}

// One path step of both legs from normals z (z[0], z[1] for Heston: Zv then Zx; z[0] for Black-Scholes).
template <typename R, int MODEL>
__device__ __forceinline__ void strip_step(const Consts<R>& c, R& y0, R& y1, R& v0, R& v1, R za, R zb) {
    if (MODEL == 1 && std::is_same<R, float>::value) {
        R y[2] = {y0, y1}, v[2] = {v0, v1};
        qe_pair_any(c, y, v, za, zb);
        y0 = y[0], y1 = y[1], v0 = v[0], v1 = v[1];
    } else if (MODEL == 1) {
        qe_step(c, y0, v0, za, zb);
        qe_step(c, y1, v1, -za, -zb);
    } else {
        y0 += c.bs_drift + c.bs_vol * za;
        y1 += c.bs_drift + c.bs_vol * -za;
    }
}

// Both legs of pair p through every step. The normals are taken four per Philox block exactly as the pair's
// stream 0 hands them out (Heston: two steps per block, Black-Scholes: four), straight from registers. At an event
// the two spots are parked in the wave's LDS slice; the payoffs and their reductions run after the path, so the
// path loop carries only the path state.
template <typename R, int MODEL>
__device__ __forceinline__ void strip_path(const StripArgs<R>& s, std::uint64_t pair, std::uint32_t k0,
                                           std::uint32_t k1, R* spot0, R* spot1, int lane) {
    const Consts<R>& c = s.c;
    R y0 = R(0), y1 = R(0), v0 = c.v0, v1 = c.v0;
    int e = 0, next_event = s.mat_step[0];
    const int n = c.n_steps;
    constexpr int per_block = MODEL == 1 ? 2 : 4;
    const std::uint32_t p0 = static_cast<std::uint32_t>(pair), p1 = static_cast<std::uint32_t>(pair >> 32);
    for (int i0 = 0; i0 < n; i0 += per_block) {
        const U4 r = philox10(U4{p0, p1, static_cast<std::uint32_t>(i0 / per_block), 0u}, k0, k1);
        R z[4];
        gauss2(r.x, r.y, z[0], z[1]);
        gauss2(r.z, r.w, z[2], z[3]);
#pragma unroll
        for (int j = 0; j < per_block; ++j) {
            const int i = i0 + j + 1;
            if (i > n) break;
            if (MODEL == 1)
                strip_step<R, 1>(c, y0, y1, v0, v1, z[2 * j], z[2 * j + 1]);
            else
                strip_step<R, 0>(c, y0, y1, v0, v1, z[j], R(0));
            if (i == next_event) {
                spot0[e * 64 + lane] = c.S0 * r_exp(y0);
                spot1[e * 64 + lane] = c.S0 * r_exp(y1);
                ++e;
                next_event = e < s.n_mats ? s.mat_step[e] : -1;
            }
        }
    }
}

template <typename R, int MODEL>
__device__ __forceinline__ void strip_body(const StripArgs<R>& s, long long n_pairs, std::uint32_t k0, std::uint32_t k1,
                                           double* partials) {
    // LDS per wave: [n_mats][64] compensated sums (s, c), then [n_mats][64] spots of each leg
    extern __shared__ double strip_lds[];
    __shared__ StripSlot<R> table[kMaxEvents * 32];  // read with broadcast loads at the payoffs
    for (int t = threadIdx.x; t < s.n_mats * 32; t += blockDim.x) table[t] = s.slots[t];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int n_values = s.n_strikes + 1, n_mats = s.n_mats;
    double* acc_s = strip_lds + (threadIdx.x >> 6) * (2 * 64 * n_mats + (2 * 64 * n_mats * sizeof(R) + 7) / 8);
    double* acc_c = acc_s + 64 * n_mats;
    R* spot0 = reinterpret_cast<R*>(acc_c + 64 * n_mats);
    R* spot1 = spot0 + 64 * n_mats;
    for (int e = 0; e < n_mats; ++e) acc_s[e * 64 + lane] = 0.0, acc_c[e * 64 + lane] = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long base = static_cast<long long>(blockIdx.x) * blockDim.x; base < n_pairs; base += stride) {
        const long long p = base + threadIdx.x;
        const bool active = p < n_pairs;
        strip_path<R, MODEL>(s, static_cast<std::uint64_t>(active ? p : 0), k0, k1, spot0, spot1, lane);
        for (int e = 0; e < n_mats; ++e) {
            const R S0 = spot0[e * 64 + lane], S1 = spot1[e * 64 + lane];
#pragma unroll
            for (int chunk = 0; chunk < 4; ++chunk) {
                const int first = (chunk & 1) * 16;
                if (first >= n_values) continue;
                double vals[16];
#pragma unroll
                for (int i2 = 0; i2 < 16; ++i2) {
                    const double yk = active ? static_cast<double>(strip_value(table[e * 32 + first + i2], S0, S1)) : 0.0;
                    vals[i2] = chunk < 2 ? yk : yk * yk;
                }
                reduce_scatter16(vals, lane);
                if ((lane >> 4) == chunk) {
                    Comp a{acc_s[e * 64 + lane], acc_c[e * 64 + lane]};
                    a.add(vals[0]);
                    acc_s[e * 64 + lane] = a.s;
                    acc_c[e * 64 + lane] = a.c;
                }
            }
        }
    }
    // the block's waves in a fixed order: one row of n_mats x 64 partial sums per block
    __syncthreads();
    if (threadIdx.x < 64) {
        const int per_wave = 2 * 64 * n_mats + (2 * 64 * n_mats * static_cast<int>(sizeof(R)) + 7) / 8;
        for (int e = 0; e < n_mats; ++e) {
            Comp a{acc_s[e * 64 + lane], acc_c[e * 64 + lane]};
            for (int w = 1; w < kBlock / 64; ++w) {
                a.add(strip_lds[w * per_wave + e * 64 + lane]);
                a.c += strip_lds[w * per_wave + 64 * n_mats + e * 64 + lane];
            }
            partials[(static_cast<long long>(blockIdx.x) * n_mats + e) * 64 + lane] = a.value();
        }
    }
    (void)wave;
}

// ---- pathwise sensitivities of the strip ------------------------------------------------------------------------
// dPrice/dS0 and the derivative in the model's variance parameter (Heston: v0; Black-Scholes: sigma) of every price
// of the strip, by forward (tangent) differentiation of each path: S_t = S0 exp(y_t) gives dS_t/dS0 = S_t / S0 (the
// path in y does not depend on S0), and the tangent (dy, dv) of the state along v0 (sigma) is carried through every
// step beside the state. A payoff D max(s (S - K), 0) differentiates to D 1{s (S - K) > 0} s dS. Many outputs and
// two inputs make the tangent the cheap direction here: two more registers a leg and no stored path.

// The special functions of the tangent step: in FP32 the price kernels' (hardware log, the Chebyshev erfc), in
// FP64 the library's, as qe_step takes them.
__device__ __forceinline__ float tan_log(float x) { return hw_log(x); }
__device__ __forceinline__ double tan_log(double x) { return r_log(x); }
__device__ __forceinline__ float tan_erfc(float x) { return erfc_fast(x); }
__device__ __forceinline__ double tan_erfc(double x) { return r_erfc(x); }

// A QE-M step of (y, v) in qe_step's operations, with the tangent (ty, tv) of (y, v) along v0 (the step's
// constants do not depend on v0). The exponential branch's v' = 0 region (1 - U >= 1 - p) has a zero tangent. The
// switch between the branches at psi_c and the martingale correction's switch to K0p are steps of the scheme that a
// pathwise derivative does not see: on paths that cross them under a bump the tangent and a finite difference part.
template <typename R>
struct TanBranch {
    R vn, lin, w, arg, dvn, dlin, darg;
    bool finite;
};

// psi = s2 / m^2 with s2 = v c1 + c2, m = theta + (v - theta) E; their tangents along v0 (dv = tv)
template <typename R>
struct TanMoments {
    R m, psi, dm, dpsi;
};

template <typename R>
__device__ __forceinline__ TanMoments<R> tan_moments(const Consts<R>& c, R v, R tv) {
    TanMoments<R> o;
    o.m = c.theta + (v - c.theta) * c.E;
    o.psi = (v * c.c1 + c.c2) / (o.m * o.m);
    o.dm = c.E * tv;
    o.dpsi = (c.c1 * tv - R(2) * o.psi * o.m * o.dm) / (o.m * o.m);
    return o;
}

template <typename R>
__device__ __forceinline__ TanBranch<R> tan_quadratic(const Consts<R>& c, const TanMoments<R>& q, R zv) {
    TanBranch<R> o;
    const R inv = R(2) / q.psi;
    const R root = r_sqrt(inv) * r_sqrt(r_max(inv - R(1), R(0)));
    const R b2 = inv - R(1) + root;
    const R a = q.m / (R(1) + b2), b = r_sqrt(b2);
    o.vn = a * (b + zv) * (b + zv);
    o.arg = R(1) - R(2) * c.A * a;
    o.finite = c.A < R(1) / (R(2) * a);
    o.lin = -c.A * b2 * a / o.arg;
    o.w = R(0.5);
    const R dinv = -inv * q.dpsi / q.psi;
    // root = sqrt(inv^2 - inv): d root = (2 inv - 1) d inv / (2 root)
    const R db2 = dinv + (root > R(0) ? (R(2) * inv - R(1)) * dinv / (R(2) * root) : R(0));
    const R da = (q.dm - a * db2) / (R(1) + b2), db = db2 / (R(2) * b);
    o.dvn = da * (b + zv) * (b + zv) + R(2) * a * (b + zv) * db;
    o.darg = R(-2) * c.A * da;
    o.dlin = (-c.A * (db2 * a + b2 * da) - o.lin * o.darg) / o.arg;
    return o;
}

template <typename R>
__device__ __forceinline__ TanBranch<R> tan_exponential(const Consts<R>& c, const TanMoments<R>& q, R zv) {
    TanBranch<R> o;
    const R p = (q.psi - R(1)) / (q.psi + R(1)), beta = (R(1) - p) / q.m;
    const R one_minus_u = R(0.5) * tan_erfc(zv * R(0.70710678118654752));
    const bool zero = one_minus_u >= R(1) - p;
    o.vn = zero ? R(0) : tan_log((R(1) - p) / one_minus_u) / beta;
    o.arg = p + beta * (R(1) - p) / (beta - c.A);
    o.finite = c.A < beta;
    o.lin = R(0);
    o.w = R(-1);
    const R dp = R(2) * q.dpsi / ((q.psi + R(1)) * (q.psi + R(1)));
    const R dbeta = (-dp - beta * q.dm) / q.m;
    o.dvn = zero ? R(0) : (-dp / (R(1) - p) - o.vn * dbeta) / beta;
    const R gap = beta - c.A;
    o.darg = dp + (dbeta * (R(1) - p) - beta * dp) / gap - beta * (R(1) - p) * dbeta / (gap * gap);
    o.dlin = R(0);
    return o;
}

template <typename R>
__device__ __forceinline__ void tan_tail(const Consts<R>& c, R& y, R& v, R& ty, R& tv, const TanBranch<R>& o, R zx) {
    const R K0 = o.finite ? o.lin + o.w * tan_log(o.arg) - (c.K1 + R(0.5) * c.K3) * v : c.K0p;
    const R dK0 = o.finite ? o.dlin + o.w * o.darg / o.arg - (c.K1 + R(0.5) * c.K3) * tv : R(0);
    const R s = c.K3 * v + c.K4 * o.vn;
    const R sd = r_sqrt(r_max(s, R(0)));
    y += c.mu_dt + K0 + c.K1 * v + c.K2 * o.vn + sd * zx;
    ty += dK0 + c.K1 * tv + c.K2 * o.dvn + (s > R(0) ? (c.K3 * tv + c.K4 * o.dvn) * zx / (R(2) * sd) : R(0));
    v = o.vn;
    tv = o.dvn;
}

// Both legs of a pair (normals (zv, zx) and (-zv, -zx)) as qe_pair takes them: the quadratic branch for both legs
// in every lane, the exponential branch once, on the leg of each lane that needs it (leg 0 when both do), and again
// only for lanes whose two legs both need it. Each leg gets the numbers of the branch its own psi selects.
template <typename R>
__device__ __forceinline__ void qe_pair_tangent(const Consts<R>& c, R (&y)[2], R (&v)[2], R (&ty)[2], R (&tv)[2],
                                                R zv, R zx) {
    const TanMoments<R> q[2] = {tan_moments(c, v[0], tv[0]), tan_moments(c, v[1], tv[1])};
    const bool ex[2] = {!(q[0].psi <= c.psi_c), !(q[1].psi <= c.psi_c)};
    TanBranch<R> o[2] = {tan_quadratic(c, q[0], zv), tan_quadratic(c, q[1], -zv)};
    if (ex[0] || ex[1]) {
        const bool first = ex[0];
        const TanBranch<R> r = tan_exponential(c, first ? q[0] : q[1], first ? zv : -zv);
        if (first)
            o[0] = r;
        else
            o[1] = r;
    }
    if (ex[0] && ex[1]) o[1] = tan_exponential(c, q[1], -zv);
    tan_tail(c, y[0], v[0], ty[0], tv[0], o[0], zx);
    tan_tail(c, y[1], v[1], ty[1], tv[1], o[1], -zx);
}

// The sums of one maturity's sensitivities (g 0: delta, 1: the variance parameter) over the wave's pairs, added to
// the wave's accumulators acc + g * stride + lane in the layout of the price's (lanes 0-31: value sums of slots
// 0-31, lanes 32-63: sums of squares). Called by the whole wave at once (the maturity steps are the same on every
// lane); inactive lanes add zeros.
template <typename R>
__device__ __forceinline__ void greek_event_sums(const StripSlot<R>* row, R S0, R S1, R dS0, R dS1, R inv_S0,
                                                 int n_values, bool active, double* acc_s, double* acc_c, int stride,
                                                 int lane) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        const R t0 = g == 0 ? S0 * inv_S0 : dS0, t1 = g == 0 ? S1 * inv_S0 : dS1;
#pragma unroll
        for (int chunk = 0; chunk < 4; ++chunk) {
            const int first = (chunk & 1) * 16;
            if (first >= n_values) continue;
            double vals[16];
#pragma unroll
            for (int i2 = 0; i2 < 16; ++i2) {
                const StripSlot<R>& t = row[first + i2];
                const R o = (t.sign * (S0 - t.K) > R(0) ? t.D * t.sign * t0 : R(0)) +
                            (t.sign * (S1 - t.K) > R(0) ? t.D * t.sign * t1 : R(0));
                const double yk = active ? static_cast<double>(R(0.5) * o) : 0.0;
                vals[i2] = chunk < 2 ? yk : yk * yk;
            }
            reduce_scatter16(vals, lane);
            if ((lane >> 4) == chunk) {
                Comp a{acc_s[g * stride + lane], acc_c[g * stride + lane]};
                a.add(vals[0]);
                acc_s[g * stride + lane] = a.s;
                acc_c[g * stride + lane] = a.c;
            }
        }
    }
}

// The strip's sensitivities: the strip's paths (one pair a lane, stream 0 of the pair) with the tangent carried
// along; at each maturity the wave adds the pairs' sensitivities to its LDS accumulators. One row of 2 n_mats x 64
// partial sums a block: [greek][maturity][slot sums, then sums of squares].
template <typename R, int MODEL>
__device__ __forceinline__ void strip_greeks_body(const StripArgs<R>& s, long long n_pairs, std::uint32_t k0,
                                                  std::uint32_t k1, double* partials) {
    extern __shared__ double strip_lds[];
    __shared__ StripSlot<R> table[kMaxEvents * 32];
    for (int t = threadIdx.x; t < s.n_mats * 32; t += blockDim.x) table[t] = s.slots[t];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int n_values = s.n_strikes + 1, n_mats = s.n_mats, stride = 64 * n_mats;
    const int per_wave = 4 * stride;
    double* acc_s = strip_lds + (threadIdx.x >> 6) * per_wave;  // [greek][maturity][64], then the compensations
    double* acc_c = acc_s + 2 * stride;
    for (int i = lane; i < 2 * stride; i += 64) acc_s[i] = 0.0, acc_c[i] = 0.0;
    const Consts<R>& c = s.c;
    const R inv_S0 = R(1) / c.S0;
    const int n = c.n_steps;
    constexpr int per_block = MODEL == 1 ? 2 : 4;  // steps a Philox block serves
    const long long grid_stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long base = static_cast<long long>(blockIdx.x) * blockDim.x; base < n_pairs; base += grid_stride) {
        const long long p = base + threadIdx.x;
        const bool active = p < n_pairs;
        const std::uint64_t pair = static_cast<std::uint64_t>(active ? p : 0);
        const std::uint32_t p0 = static_cast<std::uint32_t>(pair), p1 = static_cast<std::uint32_t>(pair >> 32);
        R y0 = R(0), y1 = R(0), v0 = c.v0, v1 = c.v0;
        R ty0 = R(0), ty1 = R(0), tv0 = R(1), tv1 = R(1);  // d/dv0 (Heston) or d/dsigma (Black-Scholes)
        int e = 0, next_event = s.mat_step[0];
        R z[4] = {R(0), R(0), R(0), R(0)};
        // FP32: a Philox block's steps a pass, unrolled, so each step's place in its block (and its normals) is a
        // compile-time constant; FP64 keeps one step a pass (unrolled it measured 13.5 -> 17.3 ms)
        constexpr int kUnroll = std::is_same<R, float>::value ? per_block : 1;
        for (int i0 = 0; i0 < n; i0 += kUnroll) {
#pragma unroll
        for (int jj = 0; jj < kUnroll; ++jj) {
            const int i = i0 + jj + 1;
            if (i > n) break;
            const int j = (i - 1) & (per_block - 1);
            if (j == 0) {
                const U4 r = philox10(U4{p0, p1, static_cast<std::uint32_t>((i - 1) / per_block), 0u}, k0, k1);
                gauss2(r.x, r.y, z[0], z[1]);
                gauss2(r.z, r.w, z[2], z[3]);
            }
            if (MODEL == 1) {
                const R zv = j == 0 ? z[0] : z[2], zx = j == 0 ? z[1] : z[3];
                R y[2] = {y0, y1}, v[2] = {v0, v1}, ty[2] = {ty0, ty1}, tv[2] = {tv0, tv1};
                qe_pair_tangent(c, y, v, ty, tv, zv, zx);
                y0 = y[0], y1 = y[1], v0 = v[0], v1 = v[1], ty0 = ty[0], ty1 = ty[1], tv0 = tv[0], tv1 = tv[1];
            } else {
                const R zb = j == 0 ? z[0] : j == 1 ? z[1] : j == 2 ? z[2] : z[3];
                y0 += c.bs_drift + c.bs_vol * zb;
                y1 += c.bs_drift + c.bs_vol * -zb;
                ty0 += c.g_drift + c.g_vol * zb;
                ty1 += c.g_drift + c.g_vol * -zb;
            }
            if (i == next_event) {
                const R S0 = c.S0 * r_exp(y0), S1 = c.S0 * r_exp(y1);
                greek_event_sums(table + e * 32, S0, S1, S0 * ty0, S1 * ty1, inv_S0, n_values, active,
                                 acc_s + e * 64, acc_c + e * 64, stride, lane);
                ++e;
                next_event = e < n_mats ? s.mat_step[e] : -1;
            }
        }
        }
    }
    // the block's waves in a fixed order
    __syncthreads();
    if (threadIdx.x < 64) {
        for (int r = 0; r < 2 * n_mats; ++r) {
            Comp a{acc_s[r * 64 + lane], acc_c[r * 64 + lane]};
            for (int w = 1; w < kBlock / 64; ++w) {
                a.add(strip_lds[w * per_wave + r * 64 + lane]);
                a.c += strip_lds[w * per_wave + 2 * stride + r * 64 + lane];
            }
            partials[(static_cast<long long>(blockIdx.x) * 2 * n_mats + r) * 64 + lane] = a.value();
        }
    }
}

template <typename R, int MODEL>
__global__ void __launch_bounds__(kBlock) strip_greeks_kernel(StripArgs<R> s, long long n_pairs, std::uint32_t k0,
                                                              std::uint32_t k1, double* partials) {
    strip_greeks_body<R, MODEL>(s, n_pairs, k0, k1, partials);
}

// Separate FP32 and FP64 kernels, so an occupancy hint can be given to one without the other (OPTLAB_STRIP_WAVES:
// minimum waves per SIMD asked of the FP32 kernel; 1 = the compiler's choice, measured fastest).
template <int MODEL>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kStripWavesF32, 8)))
strip_kernel_f32(StripArgs<float> s, long long n_pairs, std::uint32_t k0, std::uint32_t k1, double* partials) {
    strip_body<float, MODEL>(s, n_pairs, k0, k1, partials);
}

template <int MODEL>
__global__ void __launch_bounds__(kBlock) strip_kernel_f64(StripArgs<double> s, long long n_pairs, std::uint32_t k0,
                                                            std::uint32_t k1, double* partials) {
    strip_body<double, MODEL>(s, n_pairs, k0, k1, partials);
}

// The per-block rows (n_rows x n_slots) reduced to totals in two fixed-order stages, Neumaier-compensated:
// stage 1, one workgroup per (group of 64 slots, chunk of rows): wave w adds rows w, w + 16, ... of the chunk, then
// the 16 waves are combined in order; stage 2, one thread per slot combines the chunks in order. Chunks are 64 rows
// or more, at most kMaxChunks of them. A result depends only on the rows (the strip writes 2048 rows of 512 slots:
// one chain of 512 rows a wave took 163 us, the chunks take a few).
constexpr int kMinChunkRows = 64, kMaxChunks = 32;
constexpr int kChunkThreads = 1024;

__global__ void __launch_bounds__(kChunkThreads) reduce_chunks_kernel(const double* partials, long long n_rows,
                                                                       int n_slots, long long chunk_rows,
                                                                       double* chunks) {
    constexpr int W = kChunkThreads / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const long long r0 = static_cast<long long>(blockIdx.y) * chunk_rows;
    const long long r1 = r0 + chunk_rows < n_rows ? r0 + chunk_rows : n_rows;
    Comp acc{0.0, 0.0};
    if (slot < n_slots)
        for (long long r = r0 + w; r < r1; r += W) acc.add(partials[r * n_slots + slot]);
    __shared__ double hi[W][64], lo[W][64];
    hi[w][lane] = acc.s;
    lo[w][lane] = acc.c;
    __syncthreads();
    if (w == 0 && slot < n_slots) {
        Comp a{hi[0][lane], lo[0][lane]};
        for (int k = 1; k < W; ++k) {
            a.add(hi[k][lane]);
            a.c += lo[k][lane];
        }
        const long long at = (static_cast<long long>(blockIdx.y) * n_slots + slot) * 2;
        chunks[at] = a.s;
        chunks[at + 1] = a.c;
    }
}

__global__ void __launch_bounds__(256) reduce_totals_kernel(const double* chunks, int n_chunks, int n_slots,
                                                            double* totals) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n_slots) return;
    Comp a{chunks[2 * slot], chunks[2 * slot + 1]};
#pragma unroll 8
    for (int c = 1; c < n_chunks; ++c) {
        const long long at = (static_cast<long long>(c) * n_slots + slot) * 2;
        a.add(chunks[at]);
        a.c += chunks[at + 1];
    }
    totals[slot] = a.value();
}

// chunks: device memory for the chunk sums (kMaxChunks x n_slots x 2 doubles at most), or null for the cached
// scratch slot.
void reduce_rows(const double* partials, long long n_rows, int n_slots, double* totals, hipStream_t st,
                 double* chunks = nullptr) {
    const long long chunk_rows = std::max<long long>(kMinChunkRows, (n_rows + kMaxChunks - 1) / kMaxChunks);
    const int n_chunks = static_cast<int>((n_rows + chunk_rows - 1) / chunk_rows);
    if (!chunks)
        chunks = static_cast<double*>(scratch(kSlotChunks, static_cast<size_t>(n_chunks) * n_slots * 2 * sizeof(double)));
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3((n_slots + 63) / 64, n_chunks), dim3(kChunkThreads), 0, st, partials,
                       n_rows, n_slots, chunk_rows, chunks);
    hipLaunchKernelGGL(reduce_totals_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, st, chunks, n_chunks, n_slots,
                       totals);
    hip_check(hipGetLastError(), "reduce kernel launch");
}

void estimate(double sum, double sumsq, long long n_pairs, double& price, double& se) {  // cpu_mc.cpp:14-18
    const double n = static_cast<double>(n_pairs), mean = sum / n;
    const double var = sumsq / n - mean * mean;
    price = mean;
    se = n_pairs > 1 ? std::sqrt(std::fmax(var, 0.0) / (n - 1.0)) : 0.0;
}

double norm_cdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }

// Geometric Asian on n equally spaced fixings (the arithmetic Asian's control variate): ln G is Gaussian with mean
// ln S + (r - q - sigma^2 / 2) mean(t_i) and variance sigma^2 mean_ij min(t_i, t_j).
double geometric_asian(double S, double K, double T, double r, double q, double sigma, int n, bool call) {
    double mean_t = 0.0, mean_min = 0.0;
    for (int i = 1; i <= n; ++i) {
        mean_t += i * T / n;
        for (int j = 1; j <= n; ++j) mean_min += (i < j ? i : j) * T / n;
    }
    mean_t /= n;
    mean_min /= double(n) * n;
    const double mu = std::log(S) + (r - q - 0.5 * sigma * sigma) * mean_t, sd = sigma * std::sqrt(mean_min);
    const double d2 = (mu - std::log(K)) / sd, d1 = d2 + sd, eg = std::exp(mu + 0.5 * sd * sd), D = std::exp(-r * T);
    return call ? D * (eg * norm_cdf(d1) - K * norm_cdf(d2)) : D * (K * norm_cdf(-d2) - eg * norm_cdf(-d1));
}

// One pair per lane up to eight rounds of resident workgroups (as many blocks per CU as the kernel's registers and
// LDS allow); beyond that the grid-stride loop gives lanes more pairs. Several rounds let the dispatcher balance
// the uneven wave times of QE-M's branches (measured: one resident round with two pairs per lane is 4 % slower).
template <typename K>
unsigned resident_grid(K kernel, size_t lds, long long n_pairs) {
    // the occupancy of (kernel, LDS bytes) on a device does not change: asked once
    static std::map<std::tuple<const void*, size_t, int>, int> known;
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), lds, device);
    auto it = known.find(key);
    if (it == known.end()) {
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, lds), "occupancy");
        it = known.emplace(key, per_cu).first;
    }
    const int per_cu = it->second;
    const long long want = (n_pairs + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(device_cu_count()) * std::max(per_cu, 1) * 8;
    return static_cast<unsigned>(std::max<long long>(1, std::min(want, cap)));
}

// Pinned host memory for a call's totals (per device, grown on demand): the totals come back by an asynchronous
// copy on the call's stream, so the one wait at the end of the call covers the kernels and the copy.
double* pinned_totals(size_t n) {
    static std::vector<std::pair<double*, size_t>> bufs;  // per device
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    if (static_cast<int>(bufs.size()) <= device) bufs.resize(device + 1, {nullptr, 0});
    auto& b = bufs[device];
    if (b.second < n) {
        if (b.first) hip_check(hipHostFree(b.first), "free pinned totals");
        b.first = nullptr;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&b.first), n * sizeof(double), hipHostMallocDefault),
                  "pinned totals");
        b.second = n;
    }
    return b.first;
}

// HIP events around the main kernel of a call (mark() after it); stop() waits for the stream and returns the
// main kernel's time in ms. The two events are created once and reused (calls on a device are serialised).
struct Timer {
    hipStream_t st;
    hipEvent_t e0, e1;
    bool marked = false;
    explicit Timer(hipStream_t s) : st(s) {
        static hipEvent_t events[2] = {nullptr, nullptr};
        if (!events[0]) {
            hip_check(hipEventCreate(&events[0]), "event");
            hip_check(hipEventCreate(&events[1]), "event");
        }
        e0 = events[0];
        e1 = events[1];
        hip_check(hipEventRecord(e0, st), "event record");
    }
    void mark() {
        hip_check(hipEventRecord(e1, st), "event record");
        marked = true;
    }
    double stop() {
        if (!marked) hip_check(hipEventRecord(e1, st), "event record");
        hip_check(hipStreamSynchronize(st), "kernel");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        return ms;
    }
};

// Calls f with the fixed-product kernel of (R, model, product).
template <typename R, typename F>
void with_fixed_kernel(int model, int product, F f) {
    if (model == 0) {
        if (product == 0) f(fixed_kernel<R, 0, 0>);
        else if (product == 1) f(fixed_kernel<R, 0, 1>);
        else if (product == 2) f(fixed_kernel<R, 0, 2>);
        else f(fixed_kernel<R, 0, 3>);
    } else {
        if (product == 0) f(fixed_kernel<R, 1, 0>);
        else if (product == 1) f(fixed_kernel<R, 1, 1>);
        else if (product == 2) f(fixed_kernel<R, 1, 2>);
        else f(fixed_kernel<R, 1, 3>);
    }
}

// Calls f with the strip kernel of (R, model).
template <typename R, typename F>
void with_strip_kernel(int model, F f) {
    if constexpr (sizeof(R) == 4) {
        if (model == 1) f(strip_kernel_f32<1>);
        else f(strip_kernel_f32<0>);
    } else {
        if (model == 1) f(strip_kernel_f64<1>);
        else f(strip_kernel_f64<0>);
    }
}

template <typename R>
size_t strip_lds_bytes(int n_mats) {
    const size_t per_wave = 2 * 64 * n_mats * sizeof(double) + ((2 * 64 * n_mats * sizeof(R) + 7) / 8) * 8;
    return static_cast<size_t>(kBlock / 64) * per_wave;
}

// One fixed-product pricing: the resident grid, the partial sums (one row per block), the main kernel between
// the timer's events, then the fixed-order reduction. Returns the main kernel's time in ms.
template <typename R>
double price_on_device(const HostConsts& hc, int model, int product, long long n_pairs, std::uint64_t seed,
                       double* totals, double* host_totals) {
    const Consts<R> c = hc.as<R>();
    const std::uint32_t k0 = static_cast<std::uint32_t>(seed), k1 = static_cast<std::uint32_t>(seed >> 32);
    unsigned grid = 1;
    with_fixed_kernel<R>(model, product, [&](auto kernel) { grid = resident_grid(kernel, 0, n_pairs); });
    double* partials = static_cast<double*>(scratch(kSlotPartials, static_cast<size_t>(grid) * 16 * sizeof(double)));
    Timer timer(nullptr);
    with_fixed_kernel<R>(model, product, [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, nullptr, c, n_pairs, k0, k1, partials);
    });
    hip_check(hipGetLastError(), "fixed kernel launch");
    timer.mark();
    reduce_rows(partials, static_cast<long long>(grid), 16, totals, nullptr);
    hip_check(hipMemcpyAsync(host_totals, totals, 16 * sizeof(double), hipMemcpyDeviceToHost, nullptr), "copy totals");
    return timer.stop();
}

template <typename R>
double strip_on_device(const StripArgs<R>& s, long long n_pairs, std::uint64_t seed, double* totals,
                       double* host_totals) {
    const std::uint32_t k0 = static_cast<std::uint32_t>(seed), k1 = static_cast<std::uint32_t>(seed >> 32);
    const size_t lds = strip_lds_bytes<R>(s.n_mats);
    const int n_slots = s.n_mats * 64;
    int dev = 0, lds_max = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attribute");
    if (lds > static_cast<size_t>(lds_max))
        invalid("too many maturities for one workgroup's LDS in this precision (" + std::to_string(lds) + " > " +
                std::to_string(lds_max) + " bytes)");
    unsigned grid = 1;
    with_strip_kernel<R>(s.model, [&](auto kernel) { grid = resident_grid(kernel, lds, n_pairs); });
    double* partials =
        static_cast<double*>(scratch(kSlotPartials, static_cast<size_t>(grid) * n_slots * sizeof(double)));
    Timer timer(nullptr);
    with_strip_kernel<R>(s.model, [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), lds, nullptr, s, n_pairs, k0, k1, partials);
    });
    hip_check(hipGetLastError(), "strip kernel launch");
    timer.mark();
    reduce_rows(partials, static_cast<long long>(grid), n_slots, totals, nullptr);
    hip_check(hipMemcpyAsync(host_totals, totals, static_cast<size_t>(n_slots) * sizeof(double), hipMemcpyDeviceToHost,
                             nullptr),
              "copy totals");
    return timer.stop();
}

// The sensitivities' pass: as strip_on_device with the tangent kernel, whose rows hold 2 n_mats x 64 slots.
template <typename R>
double strip_greeks_on_device(const StripArgs<R>& s, long long n_pairs, std::uint64_t seed, double* totals,
                              double* host_totals) {
    const std::uint32_t k0 = static_cast<std::uint32_t>(seed), k1 = static_cast<std::uint32_t>(seed >> 32);
    const size_t lds = static_cast<size_t>(kBlock / 64) * 4 * 64 * s.n_mats * sizeof(double);
    const int n_slots = 2 * s.n_mats * 64;
    int dev = 0, lds_max = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attribute");
    if (lds > static_cast<size_t>(lds_max))
        invalid("too many maturities for one workgroup's LDS (" + std::to_string(lds) + " > " +
                std::to_string(lds_max) + " bytes)");
    auto run = [&](auto kernel) {
        const unsigned grid = resident_grid(kernel, lds, n_pairs);
        double* partials =
            static_cast<double*>(scratch(kSlotPartials, static_cast<size_t>(grid) * n_slots * sizeof(double)));
        Timer timer(nullptr);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), lds, nullptr, s, n_pairs, k0, k1, partials);
        hip_check(hipGetLastError(), "strip sensitivities kernel launch");
        timer.mark();
        reduce_rows(partials, static_cast<long long>(grid), n_slots, totals, nullptr);
        hip_check(hipMemcpyAsync(host_totals, totals, static_cast<size_t>(n_slots) * sizeof(double),
                                 hipMemcpyDeviceToHost, nullptr),
                  "copy totals");
        return timer.stop();
    };
    return s.model == 1 ? run(strip_greeks_kernel<R, 1>) : run(strip_greeks_kernel<R, 0>);
}

// The bytes of the last strip table copied to each device's strip buffer, whatever its precision.
std::vector<std::vector<unsigned char>> g_strip_sent;

// The strip's constants on the host, and its (event, slot) table and event steps copied to device memory.
template <typename R>
StripArgs<R> strip_args(const optlab_strip_request& rq) {
    StripArgs<R> s{};
    s.c = make_consts(rq.n_steps, rq.S0, rq.r, rq.q, rq.T, rq.sigma, rq.v0, rq.kappa, rq.theta, rq.xi, rq.rho, 0.0,
                      nullptr, nullptr, rq.psi_c)
              .template as<R>();
    s.model = rq.model, s.put = rq.put, s.n_strikes = rq.n_strikes, s.n_mats = rq.n_mats;
    struct Block {
        StripSlot<R> slots[kMaxEvents * 32];
        int mat_step[kMaxEvents];
    } host{};
    for (int e = 0; e < rq.n_mats; ++e) {
        host.mat_step[e] = rq.mat_step[e];
        const R disc = static_cast<R>(std::exp(-rq.r * (rq.T * rq.mat_step[e] / rq.n_steps)));
        for (int k = 0; k < 32; ++k) {
            StripSlot<R>& t = host.slots[e * 32 + k];
            if (k < rq.n_strikes)
                t = StripSlot<R>{static_cast<R>(rq.strikes[k]), disc, rq.put ? R(-1) : R(1)};
            else if (k == rq.n_strikes)
                t = StripSlot<R>{R(0), R(1), R(1)};  // max(S - 0, 0) = S: the forward, undiscounted
            else
                t = StripSlot<R>{R(0), R(0), R(1)};
        }
    }
    // The strip's own buffer, sized for the FP64 table so it is never reallocated; the table is uploaded only when
    // it differs from the last one sent to this device (repeated calls on one strip skip the copy). One record of the
    // last upload for both precisions, since they share the buffer: a float table after a double one of the same
    // strip differs in size, so it is sent again.
    std::vector<std::vector<unsigned char>>& last = g_strip_sent;  // per device
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    if (static_cast<int>(last.size()) <= device) last.resize(device + 1);
    Block* dev = static_cast<Block*>(scratch(kSlotStrip, kStripTableBytes));
    const unsigned char* bytes = reinterpret_cast<const unsigned char*>(&host);
    std::vector<unsigned char>& seen = last[device];
    if (seen.size() != sizeof(Block) || std::memcmp(seen.data(), bytes, sizeof(Block)) != 0) {
        hip_check(hipMemcpy(dev, &host, sizeof(Block), hipMemcpyHostToDevice), "copy strip");
        seen.assign(bytes, bytes + sizeof(Block));
    }
    s.slots = dev->slots;
    s.mat_step = dev->mat_step;
    return s;
}

// ---- the strip without allocation ------------------------------------------------------------------------------
// What optlab_strip_noalloc needs beyond the strip's own kernels: the slot table built on the device from values
// passed by value (no host-to-device copy), the estimates formed on the device into the caller's arrays, and a
// per-device record of the file-scope array's address and of the strip kernels' occupancy (filled on the first call
// on a device; no allocation after that, and none by the queries themselves).

struct StripTableValues {  // 460 bytes: within a kernel's argument space
    double K[kMaxStrikes];
    double D[kMaxEvents];   // the discount to each maturity, exp(-r t_j), formed on the host as strip_args forms it
    int mat_step[kMaxEvents];
    int n_strikes, n_mats, put;
};

template <typename R>
struct StripBlock {  // the layout strip_args copies to the device
    StripSlot<R> slots[kMaxEvents * 32];
    int mat_step[kMaxEvents];
};
static_assert(sizeof(StripBlock<double>) <= kStripTableBytes, "strip table layout");

template <typename R>
__global__ void __launch_bounds__(512) strip_table_kernel(StripTableValues v, StripBlock<R>* out) {
    const int t = threadIdx.x;  // one thread per (maturity, slot): 16 x 32
    const int e = t >> 5, k = t & 31;
    if (e >= v.n_mats) return;
    StripSlot<R> slot;
    if (k < v.n_strikes)
        slot = StripSlot<R>{static_cast<R>(v.K[k]), static_cast<R>(v.D[e]), v.put ? R(-1) : R(1)};
    else if (k == v.n_strikes)
        slot = StripSlot<R>{R(0), R(1), R(1)};  // the forward, undiscounted
    else
        slot = StripSlot<R>{R(0), R(0), R(1)};
    out->slots[e * 32 + k] = slot;
    if (k == 0) out->mat_step[e] = v.mat_step[e];
}

// price and SE of every (maturity, column) from the totals (cpu_mc.cpp:14-18), in the host estimate's operation
// order and roundings (no contraction into fused multiply-adds)
__global__ void __launch_bounds__(256) strip_estimate_kernel(const double* totals, int n_mats, int width,
                                                             long long n_pairs, double* price, double* se) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_mats * width) return;
    const int e = i / width, k = i - e * width;
    const double sum = totals[e * 64 + k], sumsq = totals[e * 64 + 32 + k];
    const double n = static_cast<double>(n_pairs), mean = sum / n;
    const double var = sumsq / n - mean * mean;
    price[i] = mean;
    se[i] = n_pairs > 1 ? sqrt(fmax(var, 0.0) / (n - 1.0)) : 0.0;
}

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
    unsigned char* base = nullptr;       // the device address of `scratch` on this device
    int per_cu[2][2][kMaxEvents] = {};   // [precision][model][n_mats - 1]: resident strip workgroups a CU (0: unknown)
};
NoAllocDevice g_noalloc[kNoAllocDevices];

NoAllocDevice& noalloc_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocDevices) invalid("device index beyond the allocation-free strip's table");
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.base) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(OPTLAB_EDEVICE, "hipGetSymbolAddress(scratch) gave a null address");
        d.base = static_cast<unsigned char*>(p);
    }
    return d;
}

constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }

template <typename R>
void strip_noalloc_on_device(NoAllocDevice& d, const StripArgs<R>& s_in, const StripTableValues& v, long long n_pairs,
                             std::uint64_t seed, double* price, double* se, hipStream_t st) {
    const std::uint32_t k0 = static_cast<std::uint32_t>(seed), k1 = static_cast<std::uint32_t>(seed >> 32);
    const size_t lds = strip_lds_bytes<R>(s_in.n_mats);
    const int n_slots = s_in.n_mats * 64;
    int dev = 0, lds_max = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attribute");
    if (lds > static_cast<size_t>(lds_max))
        invalid("too many maturities for one workgroup's LDS in this precision (" + std::to_string(lds) + " > " +
                std::to_string(lds_max) + " bytes)");
    int& per_cu = d.per_cu[sizeof(R) == 4][s_in.model][s_in.n_mats - 1];
    if (!per_cu)
        with_strip_kernel<R>(s_in.model, [&](auto kernel) {
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, lds), "occupancy");
        });
    // the array's layout for this call; the per-block rows take what is left, which caps the grid
    const size_t off_totals = align256(kStripTableBytes);
    const size_t off_chunks = off_totals + align256(static_cast<size_t>(n_slots) * sizeof(double));
    const size_t off_rows = off_chunks + align256(static_cast<size_t>(kMaxChunks) * n_slots * 2 * sizeof(double));
    if (off_rows >= kStripNoAllocBytes) invalid("the strip does not fit the allocation-free strip's memory");
    const long long max_rows = static_cast<long long>((kStripNoAllocBytes - off_rows) / (n_slots * sizeof(double)));
    const long long want = (n_pairs + kBlock - 1) / kBlock;
    const long long resident = static_cast<long long>(device_cu_count()) * std::max(per_cu, 1) * 8;
    const long long grid = std::max<long long>(1, std::min({want, resident, max_rows}));
    StripBlock<R>* table = reinterpret_cast<StripBlock<R>*>(d.base);
    double* totals = reinterpret_cast<double*>(d.base + off_totals);
    double* chunks = reinterpret_cast<double*>(d.base + off_chunks);
    double* partials = reinterpret_cast<double*>(d.base + off_rows);
    hipLaunchKernelGGL(strip_table_kernel<R>, dim3(1), dim3(kMaxEvents * 32), 0, st, v, table);
    StripArgs<R> s = s_in;
    s.slots = table->slots;
    s.mat_step = table->mat_step;
    with_strip_kernel<R>(s.model, [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(kBlock), lds, st, s, n_pairs, k0, k1,
                           partials);
    });
    hip_check(hipGetLastError(), "strip kernel launch");
    reduce_rows(partials, grid, n_slots, totals, st, chunks);
    const int width = s.n_strikes + 1, n_out = s.n_mats * width;
    hipLaunchKernelGGL(strip_estimate_kernel, dim3((n_out + 255) / 256), dim3(256), 0, st, totals, s.n_mats, width,
                       n_pairs, price, se);
    hip_check(hipGetLastError(), "strip estimate launch");
}

}  // namespace

void check_price_request(const optlab_request& c) {  // requests.cpp:140-166, with product 3 added
    if (c.model < 0 || c.model > 1) invalid("model must be 0 (Black-Scholes) or 1 (Heston)");
    if (c.product < 0 || c.product > 3) invalid("product must be 0, 1, 2 or 3");
    if (c.n_steps < 1) invalid("n_steps must be >= 1");
    if (c.n_paths < 4 || c.n_paths % 2) invalid("n_paths must be even and >= 4");
    if (!(c.S0 > 0 && c.T > 0)) invalid("S0 and T must be positive");
    if (c.model == 0 && !(c.sigma > 0)) invalid("sigma must be positive");
    if (c.model == 1 && !(c.v0 >= 0 && c.kappa > 0 && c.theta > 0 && c.xi > 0 && std::fabs(c.rho) <= 1))
        invalid("Heston parameters out of range");
    if (c.product == 2 && !(c.H > 0 && c.H < c.S0)) invalid("barrier: need 0 < H < S0");
    if (c.product == 3 && !(c.H > c.S0)) invalid("up-and-out barrier: need H > S0");
    for (int k = 0; k < 3; ++k)
        if (c.legs[k] < 0 || c.legs[k] > 3) invalid("legs must be 0 to 3");
}

void run_price(const optlab_request& c, int prec, optlab_result& out) {
    const auto t_start = std::chrono::steady_clock::now();
    const HostConsts hc = make_consts(c.n_steps, c.S0, c.r, c.q, c.T, c.sigma, c.v0, c.kappa, c.theta, c.xi, c.rho,
                                      c.H, c.legs, c.strikes, 1.5);
    const long long n_pairs = c.n_paths / 2;
    double* totals = static_cast<double*>(scratch(kSlotTotals, 16 * sizeof(double)));
    double* host = pinned_totals(16);
    const double kernel_ms = prec == 0 ? price_on_device<double>(hc, c.model, c.product, n_pairs, c.seed, totals, host)
                                       : price_on_device<float>(hc, c.model, c.product, n_pairs, c.seed, totals, host);
    double t[16];
    std::memcpy(t, host, sizeof t);
    std::memset(&out, 0, sizeof out);
    for (int k = 0; k < 4; ++k) estimate(t[k], t[5 + k], n_pairs, out.price[k], out.se[k]);
    if (c.product == 1) {
        out.price[2] = out.se[2] = 0.0;
        if (c.model == 0 && (c.legs[0] == 1 || c.legs[0] == 2)) {
            const double n = static_cast<double>(n_pairs);
            const double ma = t[0] / n, mg = t[1] / n;
            const double va = t[5] / n - ma * ma, vg = t[6] / n - mg * mg;
            const double cov = t[4] / n - ma * mg;
            const double beta = vg > 0 ? cov / vg : 0.0;
            const double exact = geometric_asian(c.S0, c.strikes[0], c.T, c.r, c.q, c.sigma, c.n_steps, c.legs[0] == 1);
            out.price[2] = ma - beta * (mg - exact);
            const double vr = va - 2.0 * beta * cov + beta * beta * vg;
            out.se[2] = std::sqrt(std::fmax(vr, 0.0) / (n - 1.0));
        }
    }
    out.n_paths = c.n_paths;
    out.kernel_ms = kernel_ms;
    out.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
}

void check_strip_request(const optlab_strip_request& c) {
    if (c.model < 0 || c.model > 1) invalid("model must be 0 (Black-Scholes) or 1 (Heston)");
    if (c.n_steps < 1) invalid("n_steps must be >= 1");
    if (c.n_paths < 4 || c.n_paths % 2) invalid("n_paths must be even and >= 4");
    if (!(c.S0 > 0 && c.T > 0)) invalid("S0 and T must be positive");
    if (c.model == 0 && !(c.sigma > 0)) invalid("sigma must be positive");
    if (c.model == 1 && !(c.v0 >= 0 && c.kappa > 0 && c.theta > 0 && c.xi > 0 && std::fabs(c.rho) <= 1))
        invalid("Heston parameters out of range");
    if (c.n_strikes < 1 || c.n_strikes > kMaxStrikes || !c.strikes) invalid("n_strikes must be 1 to 31");
    if (c.n_mats < 1 || c.n_mats > kMaxEvents || !c.mat_step) invalid("n_mats must be 1 to 16");
    for (int j = 0; j < c.n_mats; ++j)
        if (c.mat_step[j] < 1 || c.mat_step[j] > c.n_steps || (j && c.mat_step[j] <= c.mat_step[j - 1]))
            invalid("mat_step must increase within [1, n_steps]");
    if (!(c.psi_c >= 1.0 && c.psi_c <= 2.0)) invalid("psi_c must be in [1, 2]");
    if (!c.price || !c.se) invalid("null output");
}

void run_strip(optlab_strip_request& rq, int prec) {
    const auto t_start = std::chrono::steady_clock::now();
    const long long n_pairs = rq.n_paths / 2;
    const int n_slots = rq.n_mats * 64;
    double* totals = static_cast<double*>(scratch(kSlotTotals, static_cast<size_t>(n_slots) * sizeof(double)));
    double* host = pinned_totals(static_cast<size_t>(n_slots));
    rq.kernel_ms = prec == 0 ? strip_on_device(strip_args<double>(rq), n_pairs, rq.seed, totals, host)
                             : strip_on_device(strip_args<float>(rq), n_pairs, rq.seed, totals, host);
    const std::vector<double> t(host, host + n_slots);
    const int width = rq.n_strikes + 1;
    for (int e = 0; e < rq.n_mats; ++e)
        for (int k = 0; k < width; ++k)
            estimate(t[e * 64 + k], t[e * 64 + 32 + k], n_pairs, rq.price[e * width + k], rq.se[e * width + k]);
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
}

void run_strip_greeks(optlab_strip_request& rq, int prec, double* delta, double* delta_se, double* vega,
                      double* vega_se) {
    const auto t_start = std::chrono::steady_clock::now();
    const long long n_pairs = rq.n_paths / 2;
    const int n_slots = 2 * rq.n_mats * 64;
    double* totals = static_cast<double*>(scratch(kSlotTotals, static_cast<size_t>(n_slots) * sizeof(double)));
    double* host = pinned_totals(static_cast<size_t>(n_slots));
    rq.kernel_ms = prec == 0 ? strip_greeks_on_device(strip_args<double>(rq), n_pairs, rq.seed, totals, host)
                             : strip_greeks_on_device(strip_args<float>(rq), n_pairs, rq.seed, totals, host);
    const std::vector<double> t(host, host + n_slots);
    const int width = rq.n_strikes + 1, n_mats = rq.n_mats;
    for (int g = 0; g < 2; ++g) {
        double* value = g == 0 ? delta : vega;
        double* err = g == 0 ? delta_se : vega_se;
        for (int e = 0; e < n_mats; ++e) {
            const int row = (g * n_mats + e) * 64;
            for (int k = 0; k < width; ++k)
                estimate(t[row + k], t[row + 32 + k], n_pairs, value[e * width + k], err[e * width + k]);
        }
    }
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
}

void run_strip_noalloc(const optlab_strip_request& rq, int prec, double* price_dev, double* se_dev, hipStream_t st) {
    NoAllocDevice& d = noalloc_device();
    StripTableValues v{};
    for (int k = 0; k < rq.n_strikes; ++k) v.K[k] = rq.strikes[k];
    for (int e = 0; e < rq.n_mats; ++e) {
        v.mat_step[e] = rq.mat_step[e];
        v.D[e] = std::exp(-rq.r * (rq.T * rq.mat_step[e] / rq.n_steps));
    }
    v.n_strikes = rq.n_strikes, v.n_mats = rq.n_mats, v.put = rq.put;
    const HostConsts hc = make_consts(rq.n_steps, rq.S0, rq.r, rq.q, rq.T, rq.sigma, rq.v0, rq.kappa, rq.theta,
                                      rq.xi, rq.rho, 0.0, nullptr, nullptr, rq.psi_c);
    const long long n_pairs = rq.n_paths / 2;
    if (prec == 0) {
        StripArgs<double> s{};
        s.c = hc.as<double>();
        s.model = rq.model, s.put = rq.put, s.n_strikes = rq.n_strikes, s.n_mats = rq.n_mats;
        strip_noalloc_on_device(d, s, v, n_pairs, rq.seed, price_dev, se_dev, st);
    } else {
        StripArgs<float> s{};
        s.c = hc.as<float>();
        s.model = rq.model, s.put = rq.put, s.n_strikes = rq.n_strikes, s.n_mats = rq.n_mats;
        strip_noalloc_on_device(d, s, v, n_pairs, rq.seed, price_dev, se_dev, st);
    }
}

}  // namespace optlab_mi
