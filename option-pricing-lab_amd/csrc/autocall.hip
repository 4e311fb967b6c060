// Worst-of autocall (Phoenix) on 1 to 5 names under Dupire local vol: price by Monte Carlo and Greeks by a pathwise
// adjoint (deltas per name, vegas per node of the risk grid per surface).
// Product: packages/structured/src/optlab/structured/compiler.py:114-161, 270-295 (coupon with memory, autocall,
// knock-in put at maturity on the worst performance, a daily knock-in on the lowest close), every comparison
// smoothed as in packages/structured/src/optlab/structured/script/evaluator.py:93-108. Paths: the local-vol step of
// packages/native/cuda/include/optlab_cuda/paths.hpp:274-361 (kind 1: table variance, n_names normals per step
// on stream 0 of the pair, z = +-root . zs). Bump Greeks it replaces: structured/greeks.py:164-313 and
// packages/risk/src/optlab/risk/greeks.py:199-244.
//
// MI355X design:
//   forward   one leg per lane (lanes 2p, 2p + 1 are the two legs of pair p, so a wave reduces pair averages with
//             one shuffle), path state in VGPRs, the step's table rows read through the L1/L2 (every lane of the
//             launch reads the same rows at the same step); with Greeks on, the state every SEG steps goes to a
//             checkpoint array (structure of arrays, coalesced) and the payoff's adjoint seeds are stored per leg.
//   backward  persistent workgroups; segments of SEG steps from the last to the first; within a segment each
//             workgroup recomputes its legs' SEG steps from the checkpoint (the two lanes of a pair sharing each
//             Philox and Box-Muller evaluation: PairNormalStream), walks them backwards and adds the
//             adjoint of every table entry it read into an LDS copy of the segment's rows as 64-bit fixed point
//             (integer adds: the result does not depend on the order, so the Greeks are bit-reproducible); the
//             segment's rows are then added into the table of the workgroup's batch (16 batches of consecutive
//             workgroups) with 64-bit integer atomics in HBM, exact like the LDS adds. The batch tables' adjoints
//             are contracted with dLV/dH on the device (surface.hip).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "device.hpp"
#include "engine.hpp"
#include "localvol.hpp"
#include "surface.hpp"

namespace optlab_mi {

namespace {

constexpr int kFwdBlock = 256;
// Backward workgroup: 1024 lanes in FP32 (one workgroup a CU, so few private tables), 256 in FP64 (its state
// needs twice the registers).
template <typename R>
constexpr int bwd_block() {
    return sizeof(R) == 4 ? 1024 : 256;
}
// SEG, the steps between checkpoints, is a multiple of 4, so a segment's SEG n normals fill whole Philox blocks and
// the backward pass can start a segment's stream at its block. Candidates: 4; 8 up to 3 names; 16 for one name
// (the segment's LDS rows, SEG x n x n_x x 8 B, must fit in a workgroup's 160 KiB).
constexpr int kBatches = 16;
constexpr int kMaxNames = 5;
constexpr int kMaxObs = 64;
constexpr double kFixed = 4294967296.0;  // 2^32: the table adjoint in units of 2^-32

template <typename R>
struct AcArgs {
    int n, n_steps, n_obs, n_x;
    long long n_legs, first_pair;
    std::uint32_t key0, key1;
    R dt;
    const R* root;                 // n x n
    const R* table[kMaxNames];     // per name: n_steps x n_x
    const R* lo[kMaxNames];        // per name: n_steps, each row's first abscissa in the path precision
    const R* inv[kMaxNames];       // per name: n_steps, each row's 1 / dx in the path precision
    const R* lf;                   // n x (n_steps + 1): ln(F_a(s dt) / fixing_a)
    const int* obs_step;           // n_obs
    const int* obs_of;             // n_steps + 1: the observation at that step, or -1
    const R* disc;                 // n_obs
    Terms<R> pr;
    R spot[kMaxNames];
    int greeks;
    R* ck;                // ((c n + a) n_legs + leg)
    R* perf;              // n_obs x n_legs: perf_j, then dV / dx of the worst name at date j
    unsigned char* worst; // n_obs x n_legs
    R* lseed;             // n_legs: dV / dx of the lowest close's name at that close
    int* lat;             // n_legs: 8 step + name of the lowest close
    double* partials;     // forward workgroups x 16
    // Variance state (the kernels' VAR form): name a's spot variance is its table entry times max(v_a, 0) with v_a a
    // Heston variance stepped by Euler with full truncation (SLV: the table holds the squared leverage; Heston: a
    // table of ones; constant variance: a table of sigma^2 with v_a frozen at 1). A local-vol name has v_a frozen at
    // 1 (kappa = xi = 0, v0 = 1), so its steps are the local-vol steps exactly.
    R hv0[kMaxNames], hkappa[kMaxNames], htheta[kMaxNames], hxi[kMaxNames], hrho[kMaxNames], hrhoc[kMaxNames];
    int var_aux;                  // variance normals from the pair's aux stream (any SLV name), else from stream 0
    int same;                     // variance normals correlated by root, as the spot normals
    const R* ratio[kMaxNames];    // n_steps x n_x: d table / d local variance (SLV: L^2 / LV), or null: 1
    int grad_mask;                // bit a: name a's table moves with its surface (local vol and SLV)
    R* ckv;                       // checkpoints of v, laid out as ck
};

// The normals of one step: n spot normals from stream 0, then (VAR) n variance normals from the aux stream or
// stream 0 -- the order of the path engine's steps -- correlated into z and w for the leg of sign `sign`.
template <typename R, int N, bool VAR, typename Stream>
__device__ __forceinline__ void step_normals(const AcArgs<R>& s, Stream& rng, Stream& aux, R sign,
                                             R (&z)[N], R (&w)[N]) {
    R zs[N], zr[N];
#pragma unroll
    for (int a = 0; a < N; ++a) zs[a] = rng.next();
    if (VAR) {
#pragma unroll
        for (int a = 0; a < N; ++a) zr[a] = s.var_aux ? aux.next() : rng.next();
    }
    correlate<R, N>(s.root, sign, zs, z);
    if (VAR) {
        if (s.same) {
            correlate<R, N>(s.root, sign, zr, w);
        } else {
#pragma unroll
            for (int a = 0; a < N; ++a) w[a] = sign * zr[a];
        }
    }
}

// The spot variance of name a at this step from its interpolated table entry t (and v: max(v, 0)), and v's Euler step
// (paths.hip: the variance is read before v moves).
template <typename R, bool VAR>
__device__ __forceinline__ R step_variance(const AcArgs<R>& s, int a, R t, R& v, R za, R wa, R& vp) {
    if (!VAR) {
        vp = R(1);
        return t;
    }
    vp = r_max(v, R(0));
    const R var = t * vp;
    const R zv = s.hrho[a] * za + s.hrhoc[a] * wa;
    v += s.hkappa[a] * (s.htheta[a] - vp) * s.dt + s.hxi[a] * r_sqrt(vp * s.dt) * zv;
    return var;
}

template <typename R, int N, int SEG, bool VAR>
__global__ void __launch_bounds__(kFwdBlock) autocall_forward(AcArgs<R> s) {
    __shared__ double wave_part[kFwdBlock / 64][16];
    const long long i = static_cast<long long>(blockIdx.x) * kFwdBlock + threadIdx.x;
    const bool active = i < s.n_legs;
    const long long leg = active ? i : 0;
    const R sign = (i & 1) ? R(-1) : R(1);
    const int stride = s.n_steps + 1;
    NormalStream<R> rng, aux;
    rng.init(static_cast<std::uint64_t>(s.first_pair + (leg >> 1)), s.key0, s.key1, 0u);
    if (VAR) aux.init(static_cast<std::uint64_t>(s.first_pair + (leg >> 1)), s.key0, s.key1, kAuxTag);

    R x[N], vs[N];
#pragma unroll
    for (int a = 0; a < N; ++a) x[a] = R(0), vs[a] = VAR ? s.hv0[a] : R(1);
    R ymin = ldc(s.lf, 0);
    int at = 0;
#pragma unroll
    for (int a = 1; a < N; ++a)
        if (ldc(s.lf, a * stride) < ymin) ymin = ldc(s.lf, a * stride), at = a;
    Note<R> note{R(1), R(0), R(0)};
    int j = 0, next = ldc(s.obs_step, 0);

    auto one_step = [&](int step) __attribute__((always_inline)) {
        if (s.greeks && step % SEG == 0 && active) {
            const long long c = step / SEG;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                s.ck[(c * N + a) * s.n_legs + i] = x[a];
                if (VAR) s.ckv[(c * N + a) * s.n_legs + i] = vs[a];
            }
        }
        R z[N], wz[N];
        step_normals<R, N, VAR>(s, rng, aux, sign, z, wz);
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const R* y = s.table[a] + static_cast<long long>(step) * s.n_x;
            const Lerp<R> l =
                lerp_at(y, ldc(s.lo[a], step), ldc(s.inv[a], step), s.n_x, x[a]);
            R vp;
            const R var = step_variance<R, VAR>(s, a, l.var, vs[a], z[a], VAR ? wz[a] : R(0), vp);
            x[a] += R(-0.5) * var * s.dt + r_sqrt(var * s.dt) * z[a];
        }
        const int t1 = step + 1;
        R yw = ldc(s.lf, t1) + x[0];
        int w = 0;
#pragma unroll
        for (int a = 1; a < N; ++a) {
            const R ya = ldc(s.lf, a * stride + t1) + x[a];
            if (ya < yw) yw = ya, w = a;
        }
        if (yw < ymin) ymin = yw, at = t1 * 8 + w;
        if (t1 == next) {
            const R p = r_exp(yw);
            const bool last = j == s.n_obs - 1;
            if (s.greeks && active) {
                s.perf[static_cast<long long>(j) * s.n_legs + i] = p;
                s.worst[static_cast<long long>(j) * s.n_legs + i] = static_cast<unsigned char>(w);
            }
            observe(s.pr, last, p, r_exp(ymin), ldc(s.disc, j), note);
            ++j;
            next = j < s.n_obs ? ldc(s.obs_step, j) : -1;
        }
    };
    // Without a variance state a step draws exactly N normals from stream 0, so four steps draw N whole Philox
    // blocks: unrolled by four, the stream's buffer position is a compile-time constant at every draw.
    int step = 0;
    if constexpr (!VAR) {
        for (; step + 4 <= s.n_steps; step += 4) {
            one_step(step);
            one_step(step + 1);
            one_step(step + 2);
            one_step(step + 3);
        }
    }
    for (; step < s.n_steps; ++step) one_step(step);

    // Reverse sweep of the script for dV / dperf_j and dV / dlow (the state before date j recomputed from the
    // stored performances: n_obs^2 / 2 cheap steps instead of n_obs saved states).
    R dsum[N];
#pragma unroll
    for (int a = 0; a < N; ++a) dsum[a] = R(0);
    if (s.greeks && active) {
        const Terms<R>& t = s.pr;
        const R low = r_exp(ymin);
        R abar = R(0), mbar = R(0), lbar = R(0);
        for (int jj = s.n_obs - 1; jj >= 0; --jj) {
            R alive = R(1), missed = R(0);
            for (int q = 0; q < jj; ++q) {
                const R pq = s.perf[static_cast<long long>(q) * s.n_legs + i];
                const R c = ramp_ge(pq, t.coupon_barrier, t.smoothing).v;
                if (t.memory) missed = (R(1) - c) * (missed + R(1));
                alive *= R(1) - ramp_ge(pq, t.autocall, t.smoothing).v;
            }
            const long long o = static_cast<long long>(jj) * s.n_legs + i;
            const R p = s.perf[o], D = ldc(s.disc, jj);
            const Ramp<R> c = ramp_ge(p, t.coupon_barrier, t.smoothing);
            R pbar = R(0);
            if (jj == s.n_obs - 1) {
                const R h = r_max(t.put_strike - p, R(0)), dh = t.put_strike - p > R(0) ? R(-1) : R(0);
                const Ramp<R> ki = t.ki_barrier > R(0) ? ramp_lt(low, t.ki_barrier, t.smoothing) : Ramp<R>{R(0), R(0)};
                abar += D * (R(1) - ki.v * h);
                pbar += alive * D * (-ki.v * dh);
                lbar += -alive * D * h * ki.d;
            } else {
                const Ramp<R> u = ramp_ge(p, t.autocall, t.smoothing);
                const R ubar = alive * (D - abar);
                abar = abar * (R(1) - u.v) + u.v * D;
                pbar += ubar * u.d;
            }
            R cbar = R(0);
            if (t.memory) {
                cbar -= mbar * (missed + R(1));
                mbar *= R(1) - c.v;
            }
            const R f = t.memory ? R(1) + missed : R(1);
            abar += c.v * t.coupon * f * D;
            cbar += alive * t.coupon * f * D;
            if (t.memory) mbar += alive * c.v * t.coupon * D;
            pbar += cbar * c.d;
            const R seed = pbar * p;
            s.perf[o] = seed;
            const int w = s.worst[o];
#pragma unroll
            for (int a = 0; a < N; ++a)
                if (a == w) dsum[a] += seed;
        }
        const R lseed = lbar * low;
        s.lseed[i] = lseed;
        s.lat[i] = at;
#pragma unroll
        for (int a = 0; a < N; ++a)
            if (a == (at & 7)) dsum[a] += lseed;
    }

    // pair averages (lanes 2p and 2p + 1), then sums and sums of squares over the wave and the workgroup
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = 0.0;
    {
        const double mine = active ? static_cast<double>(note.value) : 0.0;
        const double pair = 0.5 * (mine + shfl_xor_d(mine, 1));
        const bool even = (threadIdx.x & 1) == 0;
        v[0] = even ? pair : 0.0;
        v[1] = even ? pair * pair : 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const double d = active ? static_cast<double>(dsum[a] / s.spot[a]) : 0.0;
            const double pd = 0.5 * (d + shfl_xor_d(d, 1));
            v[2 + 2 * a] = even ? pd : 0.0;
            v[3 + 2 * a] = even ? pd * pd : 0.0;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    reduce_scatter16(v, lane);
    if (lane < 16) wave_part[wave][lane] = v[0];
    __syncthreads();
    if (threadIdx.x < 16) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kFwdBlock / 64; ++w) t += wave_part[w][threadIdx.x];
        s.partials[static_cast<long long>(blockIdx.x) * 16 + threadIdx.x] = t;
    }
}

// batches[b * 16 + k] = sum over the forward workgroups of batch b (in order) of value k.
__global__ void batch_sums_kernel(const double* partials, long long n_blocks, double* batches) {
    const int b = threadIdx.x / 16, k = threadIdx.x % 16;
    const long long b0 = b * n_blocks / kBatches, b1 = (b + 1) * n_blocks / kBatches;
    double t = 0.0;
    long long blk = b0;
    // eight loads in flight a pass, added in block order
    for (; blk + 8 <= b1; blk += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partials[(blk + u) * 16 + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) t += v[u];
    }
    for (; blk < b1; ++blk) t += partials[blk * 16 + k];
    batches[threadIdx.x] = t;
}

template <typename R>
__device__ __forceinline__ long long to_fixed(R v) {
    return static_cast<long long>(rint(v * static_cast<R>(kFixed)));
}

// FP32: the same integer, rint(v 2^32), from two 32-bit conversions (gfx950 has no float-to-64-bit conversion, and
// its generic expansion is 13 instructions): q = hi 2^16 + lo with hi = floor(q / 2^16) and 0 <= lo < 2^16, both exact
// in float, so the 64-bit integer's low word is hi << 16 | lo and its high word hi >> 16 (arithmetic). Valid while
// |v| < 2^15, far beyond any adjoint weight here.
template <>
__device__ __forceinline__ long long to_fixed<float>(float v) {
    const float q = rintf(v * static_cast<float>(kFixed));
    const float hq = floorf(q * (1.0f / 65536.0f));
    const int hi = static_cast<int>(hq);
    const unsigned lo = static_cast<unsigned>(fmaf(hq, -65536.0f, q));
    const unsigned low = (static_cast<unsigned>(hi) << 16) | lo;
    const unsigned high = static_cast<unsigned>(hi >> 16);
    return static_cast<long long>((static_cast<unsigned long long>(high) << 32) | low);
}

// 1 / x for the adjoint's weights: the hardware reciprocal in FP32 (1 ulp; an IEEE division is ten instructions in
// a kernel that is bound by its instruction issue), the division in FP64.
__device__ __forceinline__ float adj_rcp(float x) { return hw_rcp(x); }
__device__ __forceinline__ double adj_rcp(double x) { return 1.0 / x; }

template <typename R, int N, int SEG, bool VAR>
__global__ void __launch_bounds__(bwd_block<R>()) autocall_backward(AcArgs<R> s, long long n_chunks, R* xbar,
                                                               unsigned long long* batch, int xl_legs) {
    constexpr int kBwdBlock = bwd_block<R>();
    extern __shared__ unsigned char lds_raw[];
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds_raw);  // SEG x N x n_x
    const long long G = gridDim.x, blk = blockIdx.x;
    const long long c0 = blk * n_chunks / G, c1 = (blk + 1) * n_chunks / G;
    const int n_seg = (s.n_steps + SEG - 1) / SEG;
    const int row_len = N * s.n_x;
    // x-bar of the workgroup's legs from one segment to the one before: in LDS after the rows when xl_legs > 0
    // ([name][xl_legs], a lane's own entries only), else in device memory
    R* xl = reinterpret_cast<R*>(acc + SEG * row_len);
    // The rows start at zero; the flush at the end of each segment clears what it reads, so they start the next
    // segment at zero too.
    for (int e = threadIdx.x; e < SEG * row_len; e += kBwdBlock) acc[e] = 0ull;
    // The flush walks the entries e = threadIdx.x + m kBwdBlock as (row, column) = (e / n_x, e % n_x), advanced by
    // (kBwdBlock / n_x, kBwdBlock % n_x) each time instead of dividing.
    const int adv_rows = kBwdBlock / s.n_x, adv_cols = kBwdBlock - adv_rows * s.n_x;
    const int first_row = static_cast<int>(threadIdx.x) / s.n_x;
    const int first_col = static_cast<int>(threadIdx.x) - first_row * s.n_x;
    long long b = blk * kBatches / G;
    while (b > 0 && b * G / kBatches > blk) --b;
    while (b + 1 < kBatches && (b + 1) * G / kBatches <= blk) ++b;
    unsigned long long* const tb = batch + b * static_cast<long long>(N) * s.n_steps * s.n_x;
    __syncthreads();
    for (int seg = n_seg - 1; seg >= 0; --seg) {
        const int s0 = seg * SEG, len = min(SEG, s.n_steps - s0);
        for (long long chunk = c0; chunk < c1; ++chunk) {
            const long long i = chunk * kBwdBlock + threadIdx.x;
            const bool active = i < s.n_legs;
            const long long leg = active ? i : 0;
            const int local = static_cast<int>(chunk - c0) * kBwdBlock + static_cast<int>(threadIdx.x);
            const R sign = (i & 1) ? R(-1) : R(1);
            R x[N], xb[N], v[N];
#pragma unroll
            for (int a = 0; a < N; ++a) {
                x[a] = s.ck[(static_cast<long long>(seg) * N + a) * s.n_legs + leg];
                v[a] = VAR ? s.ckv[(static_cast<long long>(seg) * N + a) * s.n_legs + leg] : R(1);
                xb[a] = seg == n_seg - 1 ? R(0) : (xl_legs ? xl[a * xl_legs + local] : xbar[a * s.n_legs + leg]);
            }
            // the pair's two legs are lanes 2p and 2p + 1 (chunks start at multiples of the workgroup size), and
            // every lane draws, active or not (an inactive pair draws pair 0's normals on both of its lanes): the
            // two lanes of a pair share each refill's Philox and Box-Muller work
            PairNormalStream<R> rng, aux;
            rng.init(static_cast<std::uint64_t>(s.first_pair + (leg >> 1)), s.key0, s.key1, 0u);
            // a segment's normals start at a Philox block: SEG n of stream 0 a segment (2 SEG n when the variance
            // normals come from stream 0 too), SEG n of the aux stream
            const int per_seg = (VAR && !s.var_aux ? 2 : 1) * (SEG * N / 4);
            rng.block = static_cast<std::uint32_t>(seg * per_seg);
            if (VAR) {
                aux.init(static_cast<std::uint64_t>(s.first_pair + (leg >> 1)), s.key0, s.key1, kAuxTag);
                aux.block = static_cast<std::uint32_t>(seg * (SEG * N / 4));
            }
            // What the walk back needs of each (step, name). Up to 12 (step, name) pairs a segment it is kept from
            // the recompute: the lerp's grid coordinate (its entry and weight), its slope and dx'/dt =
            // v+ dt (-1/2 + z / (2 sqrt(var dt))) for the table entry t (var = t v+; v+ = 1 without a variance
            // state), which spares the walk back its table reads. Beyond that the three values would not fit the
            // 128 VGPRs of a 1024-lane workgroup: the state x, the normal z (and v+) are kept instead and the walk
            // back reads the table again.
            constexpr bool kKeep = SEG * N <= 12;
            constexpr int KS = kKeep ? SEG : 1, XS = kKeep ? 1 : SEG, VS = (!kKeep && VAR) ? SEG : 1;
            R lu[KS][N], ls[KS][N], lw[KS][N], xs[XS][N], zz[XS][N], vv[VS][N];
#pragma unroll
            for (int k = 0; k < SEG; ++k) {
                if (k >= len) break;
                const int step = s0 + k;
                R z[N], w[N];
                step_normals<R, N, VAR>(s, rng, aux, sign, z, w);
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    const R* y = s.table[a] + static_cast<long long>(step) * s.n_x;
                    const Lerp<R> l = lerp_at(y, ldc(s.lo[a], step), ldc(s.inv[a], step),
                                              s.n_x, x[a]);
                    if constexpr (!kKeep) {
                        xs[k][a] = x[a];
                        zz[k][a] = z[a];
                    }
                    R vp;
                    const R var = step_variance<R, VAR>(s, a, l.var, v[a], z[a], VAR ? w[a] : R(0), vp);
                    const R sd = r_sqrt(var * s.dt);
                    if constexpr (kKeep) {
                        lu[k][a] = l.u;
                        ls[k][a] = l.slope;
                        // a step with no variance (v+ = 0) does not move with its table entry: weight 0, not 0 inf
                        lw[k][a] = sd > R(0) ? vp * (s.dt * (R(-0.5) + R(0.5) * z[a] * adj_rcp(sd))) : R(0);
                    } else if constexpr (VAR) {
                        vv[k][a] = vp;
                    }
                    x[a] += R(-0.5) * var * s.dt + sd * z[a];
                }
            }
            const int lat = s.lat[leg];
            const R lseed = s.lseed[leg];
#pragma unroll
            for (int k = SEG - 1; k >= 0; --k) {
                if (k >= len) continue;
                const int t1 = s0 + k + 1;
                const int j = ldc(s.obs_of, t1);
                if (j >= 0) {
                    const long long o = static_cast<long long>(j) * s.n_legs + leg;
                    const int w = s.worst[o];
                    const R g = s.perf[o];
#pragma unroll
                    for (int a = 0; a < N; ++a)
                        if (a == w) xb[a] += g;
                }
                if ((lat >> 3) == t1) {
#pragma unroll
                    for (int a = 0; a < N; ++a)
                        if (a == (lat & 7)) xb[a] += lseed;
                }
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    R u, slope, wgt;
                    if constexpr (kKeep) {
                        u = lu[k][a], slope = ls[k][a], wgt = lw[k][a];
                    } else {
                        const int step = s0 + k;
                        const R* y = s.table[a] + static_cast<long long>(step) * s.n_x;
                        const Lerp<R> l = lerp_at(y, ldc(s.lo[a], step),
                                                  ldc(s.inv[a], step), s.n_x, xs[k][a]);
                        const R vp = VAR ? vv[VAR ? k : 0][a] : R(1);
                        const R sd = r_sqrt(l.var * vp * s.dt);
                        u = l.u, slope = l.slope;
                        wgt = sd > R(0) ? vp * (s.dt * (R(-0.5) + R(0.5) * zz[k][a] * adj_rcp(sd))) : R(0);
                    }
                    const R vbar = xb[a] * wgt;  // d V / d (the table entry interpolated at x)
                    const int i0 = static_cast<int>(u);
                    const R f = u - static_cast<R>(i0);
                    if (active && (!VAR || ((s.grad_mask >> a) & 1))) {
                        // a table of the squared leverage moves with the local variance as L^2 / LV of each entry
                        R r0 = R(1), r1 = R(1);
                        if (VAR && s.ratio[a]) {
                            const R* rr = s.ratio[a] + static_cast<long long>(s0 + k) * s.n_x + i0;
                            r0 = rr[0];
                            r1 = rr[1];  // i0 + 1 <= n_x - 1: the lerp keeps u below n_x - 1
                        }
                        unsigned long long* r = acc + (k * N + a) * s.n_x + i0;
                        // the two entries' shares: with no ratio the left one is the whole weight's integer less
                        // the right one's (one conversion fewer; the sum of the pair is then exactly the whole's)
                        long long q0, q1;
                        if (VAR && s.ratio[a]) {
                            q0 = to_fixed(vbar * (R(1) - f) * r0);
                            q1 = to_fixed(vbar * f * r1);
                        } else {
                            q1 = to_fixed(vbar * f);
                            q0 = to_fixed(vbar) - q1;
                        }
                        if (q0) atomicAdd(r, static_cast<unsigned long long>(q0));
                        if (q1) atomicAdd(r + 1, static_cast<unsigned long long>(q1));
                    }
                    xb[a] += vbar * slope;
                }
            }
            if (seg > 0 && active) {
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    if (xl_legs)
                        xl[a * xl_legs + local] = xb[a];
                    else
                        xbar[a * s.n_legs + i] = xb[a];
                }
            }
        }
        __syncthreads();
        // the segment's rows into this workgroup's batch table (batch b = the workgroups [b G / 16, (b + 1) G / 16)):
        // integer atomics, so the batch's table does not depend on the order of its workgroups' additions
        {
            int row = first_row, col = first_col;
            for (int e = threadIdx.x; e < len * row_len; e += kBwdBlock) {
                const unsigned long long v = acc[e];
                if (v) {
                    acc[e] = 0ull;
                    const int k = row / N, a = row - k * N;
                    __hip_atomic_fetch_add(tb + (static_cast<long long>(a) * s.n_steps + s0 + k) * s.n_x + col, v,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                row += adv_rows;
                col += adv_cols;
                if (col >= s.n_x) {
                    col -= s.n_x;
                    ++row;
                }
            }
        }
        __syncthreads();
    }
}

// An SLV name's table from its squared leverage rows (lev2) and the surface's local variance: ratio = L^2 / LV of
// every entry (the adjoint's d table / d LV), and the table itself -- lev2 at the unbumped surface, ratio times the
// bumped local variance under a bump (the conditional expectation E[v | x] = LV / L^2 held).
template <typename R>
__global__ void slv_tables_kernel(const double* lev2, const double* lv_base, const R* lv_moved, R* ratio, R* table,
                                  long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = lev2[i] / lv_base[i];
    if (ratio) ratio[i] = static_cast<R>(r);
    table[i] = lv_moved ? static_cast<R>(r * static_cast<double>(lv_moved[i])) : static_cast<R>(lev2[i]);
}

// A name's row ends (first abscissa, 1 / dx) in the path precision: converted once a call, so that the step loops
// read them in the precision they use (the same conversion the loops made at every read).
template <typename R>
__global__ void row_ends_kernel(const double* lo, const double* inv, int n, R* out) {
    const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = static_cast<R>(lo[i]);
        out[n + i] = static_cast<R>(inv[i]);
    }
}

template <typename R>
__global__ void fill_table_kernel(R* table, R value, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) table[i] = value;
}

// ---- host -------------------------------------------------------------------------------------------------------

struct Arena {
    std::vector<unsigned char> host;
    size_t put(const void* src, size_t bytes) {
        const size_t at = (host.size() + 255) & ~size_t(255);
        host.resize(at + bytes);
        if (bytes) std::memcpy(host.data() + at, src, bytes);
        return at;
    }
    template <typename R>
    size_t reals(const std::vector<double>& v) {
        std::vector<R> tmp(v.begin(), v.end());
        return put(tmp.data(), tmp.size() * sizeof(R));
    }
};

struct Events {
    hipEvent_t e[4];
    Events() {
        for (auto& x : e) hip_check(hipEventCreate(&x), "hipEventCreate");
    }
    ~Events() {
        for (auto& x : e) (void)hipEventDestroy(x);
    }
    float ms(int a, int b) const {
        float t = 0.f;
        hip_check(hipEventElapsedTime(&t, e[a], e[b]), "hipEventElapsedTime");
        return t;
    }
};

void mean_se(double sum, double sum2, double n, double& mean, double& se) {  // cpu_mc.cpp:14-18
    mean = sum / n;
    const double var = n > 1 ? std::max(sum2 - sum * sum / n, 0.0) / (n - 1) : 0.0;
    se = std::sqrt(var / n);
}

int model_kind(const optlab_autocall_models* m, int a) { return m ? m->kind[a] : 1; }

template <typename R, int N, int SEG, bool VAR>
void run_typed(optlab_autocall_request& rq, const optlab_autocall_models* models, bool adjoint) {
    static_assert(SEG % 4 == 0, "a segment's normals must fill whole Philox blocks");
    const auto wall0 = std::chrono::steady_clock::now();
    const int n = rq.n_names, ns = rq.n_steps;
    const long long n_legs = 2 * rq.n_pairs;
    const TentGrid grid = make_tent_grid();
    const int n_nodes = grid.n_strikes * grid.n_tenors;
    constexpr int kBwdBlock = bwd_block<R>();
    hipStream_t stream = nullptr;

    // tables: cached unmoved ones, or bumped ones built here
    const SurfaceRows* rows[kMaxNames];
    const R* tables[kMaxNames];
    int n_x = 0;
    int n_bumped = 0;
    for (int a = 0; a < n; ++a) {
        rows[a] = &surface_rows(rq.surface[a]);
        n_x = static_cast<int>(rows[a]->n_x);
        bool moved = false;
        if (rq.nodes)
            for (int k = 0; k < n_nodes; ++k) moved |= rq.nodes[a * n_nodes + k] != 0.0;
        n_bumped += moved;
    }
    const size_t table_bytes = static_cast<size_t>(ns) * n_x * sizeof(R);
    const size_t node_bytes = static_cast<size_t>(n_nodes) * sizeof(double);
    unsigned char* bump_mem = n_bumped ? static_cast<unsigned char*>(scratch(
                                             kSlotTables, n_bumped * (((table_bytes + 255) & ~size_t(255)) + node_bytes)))
                                       : nullptr;
    // the other kinds' tables: constants (kinds 0 and 2), and SLV's from its leverage rows with their ratios
    const long long entries_a = static_cast<long long>(ns) * n_x;
    const size_t real_b = (static_cast<size_t>(entries_a) * sizeof(R) + 255) & ~size_t(255);
    const size_t dbl_b = (static_cast<size_t>(entries_a) * sizeof(double) + 255) & ~size_t(255);
    int n_other = 0, n_slv = 0;
    for (int a = 0; a < n; ++a) {
        const int k = model_kind(models, a);
        n_other += k != 1;
        n_slv += k == 3;
    }
    unsigned char* model_mem =
        n_other ? static_cast<unsigned char*>(scratch(kSlotModels, n_other * real_b + n_slv * (real_b + dbl_b)))
                : nullptr;
    const R* ratios[kMaxNames] = {};
    const unsigned blocks_a = static_cast<unsigned>((entries_a + 255) / 256);
    for (int a = 0, b = 0, m = 0, o = 0; a < n; ++a) {
        bool moved = false;
        if (rq.nodes)
            for (int k = 0; k < n_nodes; ++k) moved |= rq.nodes[a * n_nodes + k] != 0.0;
        const int kind = model_kind(models, a);
        if (kind == 0 || kind == 2) {
            R* t = reinterpret_cast<R*>(model_mem + o++ * real_b);
            const R value = kind == 0 ? static_cast<R>(models->cvar[a]) : R(1);
            hipLaunchKernelGGL(fill_table_kernel<R>, dim3(blocks_a), dim3(256), 0, stream, t, value, entries_a);
            hip_check(hipGetLastError(), "table fill launch");
            tables[a] = t;
            continue;
        }
        if (kind == 3) {
            R* t = reinterpret_cast<R*>(model_mem + o++ * real_b);
            unsigned char* at = model_mem + n_other * real_b + m++ * (real_b + dbl_b);
            R* ratio = reinterpret_cast<R*>(at);
            double* lev2 = reinterpret_cast<double*>(at + real_b);
            hip_check(hipMemcpyAsync(lev2, models->lev2 + a * entries_a, entries_a * sizeof(double),
                                     hipMemcpyHostToDevice, stream),
                      "copy leverage rows");
            const double* lv_base = static_cast<const double*>(surface_base_table(rq.surface[a], 0, stream));
            const R* lv_moved = nullptr;
            if (moved) {
                unsigned char* bt = bump_mem + b++ * (((table_bytes + 255) & ~size_t(255)) + node_bytes);
                double* d_nodes = reinterpret_cast<double*>(bt + ((table_bytes + 255) & ~size_t(255)));
                hip_check(hipMemcpyAsync(d_nodes, rq.nodes + a * n_nodes, node_bytes, hipMemcpyHostToDevice, stream),
                          "copy nodes");
                launch_bumped_table(*rows[a], grid, d_nodes, SurfaceMove{0.0, 0.0},
                                    std::is_same<R, double>::value ? 0 : 1, bt, stream);
                lv_moved = reinterpret_cast<const R*>(bt);
            }
            hipLaunchKernelGGL(slv_tables_kernel<R>, dim3(blocks_a), dim3(256), 0, stream, lev2, lv_base, lv_moved,
                               adjoint ? ratio : nullptr, t, entries_a);
            hip_check(hipGetLastError(), "SLV table launch");
            tables[a] = t;
            ratios[a] = ratio;
            continue;
        }
        if (!moved) {
            tables[a] = static_cast<const R*>(surface_base_table(rq.surface[a], std::is_same<R, double>::value ? 0 : 1,
                                                                 stream));
            continue;
        }
        unsigned char* at = bump_mem + b++ * (((table_bytes + 255) & ~size_t(255)) + node_bytes);
        double* d_nodes = reinterpret_cast<double*>(at + ((table_bytes + 255) & ~size_t(255)));
        hip_check(hipMemcpyAsync(d_nodes, rq.nodes + a * n_nodes, node_bytes, hipMemcpyHostToDevice, stream),
                  "copy nodes");
        launch_bumped_table(*rows[a], grid, d_nodes, SurfaceMove{0.0, 0.0}, std::is_same<R, double>::value ? 0 : 1,
                            at, stream);
        tables[a] = reinterpret_cast<const R*>(at);
    }

    // the request's small arrays
    std::vector<double> root(rq.root, rq.root + n * n), lf(static_cast<size_t>(n) * (ns + 1)), disc(rq.disc,
                                                                                                  rq.disc + rq.n_obs);
    double spot[kMaxNames];
    for (int a = 0; a < n; ++a) {
        const double factor = rq.spot ? rq.spot[a] : 1.0;
        spot[a] = rq.fwd[a * (ns + 1)] * factor;
        for (int t = 0; t <= ns; ++t) lf[a * (ns + 1) + t] = std::log(rq.fwd[a * (ns + 1) + t] * factor / rq.fixing[a]);
    }
    std::vector<int> obs_of(ns + 1, -1);
    for (int j = 0; j < rq.n_obs; ++j) obs_of[rq.obs_step[j]] = j;
    Arena ar;
    const size_t o_root = ar.reals<R>(root), o_lf = ar.reals<R>(lf), o_disc = ar.reals<R>(disc);
    const size_t o_obs = ar.put(rq.obs_step, rq.n_obs * sizeof(int)), o_of = ar.put(obs_of.data(), obs_of.size() * 4);
    const std::vector<R> ends_room(static_cast<size_t>(2) * n * ns, R(0));  // filled on the device below
    const size_t o_ends = ar.put(ends_room.data(), ends_room.size() * sizeof(R));
    unsigned char* dev = static_cast<unsigned char*>(scratch(kSlotSpec, std::max<size_t>(ar.host.size(), 256)));
    hip_check(hipMemcpyAsync(dev, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice, stream), "copy spec");

    AcArgs<R> s{};
    s.n = n, s.n_steps = ns, s.n_obs = rq.n_obs, s.n_x = n_x;
    s.n_legs = n_legs, s.first_pair = rq.first_pair;
    s.key0 = static_cast<std::uint32_t>(rq.seed), s.key1 = static_cast<std::uint32_t>(rq.seed >> 32);
    s.dt = static_cast<R>(rq.dt);
    s.root = reinterpret_cast<const R*>(dev + o_root);
    for (int a = 0; a < n; ++a) {
        R* ends = reinterpret_cast<R*>(dev + o_ends) + static_cast<size_t>(2) * a * ns;
        hipLaunchKernelGGL(row_ends_kernel<R>, dim3((ns + 255) / 256), dim3(256), 0, stream, rows[a]->lo,
                           rows[a]->inv_dx, ns, ends);
        hip_check(hipGetLastError(), "row ends launch");
        s.table[a] = tables[a];
        s.lo[a] = ends;
        s.inv[a] = ends + ns;
        s.spot[a] = static_cast<R>(spot[a]);
    }
    s.lf = reinterpret_cast<const R*>(dev + o_lf);
    s.obs_step = reinterpret_cast<const int*>(dev + o_obs);
    s.obs_of = reinterpret_cast<const int*>(dev + o_of);
    s.disc = reinterpret_cast<const R*>(dev + o_disc);
    s.pr = Terms<R>{static_cast<R>(rq.autocall), static_cast<R>(rq.coupon_barrier), static_cast<R>(rq.coupon),
                    static_cast<R>(rq.ki_barrier), static_cast<R>(rq.put_strike), static_cast<R>(rq.smoothing),
                    rq.memory};
    s.greeks = adjoint ? 1 : 0;
    s.grad_mask = 0;
    s.var_aux = 0;
    s.same = models ? models->same : 0;
    for (int a = 0; a < n; ++a) {
        const int kind = model_kind(models, a);
        if (kind == 1 || kind == 3) s.grad_mask |= 1 << a;
        s.var_aux |= kind == 3;
        s.ratio[a] = ratios[a];
        if (kind == 2 || kind == 3) {
            const double* h = models->heston + 5 * a;
            s.hv0[a] = static_cast<R>(h[0]), s.hkappa[a] = static_cast<R>(h[1]), s.htheta[a] = static_cast<R>(h[2]);
            s.hxi[a] = static_cast<R>(h[3]), s.hrho[a] = static_cast<R>(h[4]);
            s.hrhoc[a] = static_cast<R>(std::sqrt(1.0 - h[4] * h[4]));
        } else {  // a variance frozen at 1: the name's own table is its spot variance
            s.hv0[a] = R(1), s.hkappa[a] = R(0), s.htheta[a] = R(0), s.hxi[a] = R(0), s.hrho[a] = R(0);
            s.hrhoc[a] = R(1);
        }
    }

    const long long fwd_blocks = (n_legs + kFwdBlock - 1) / kFwdBlock;
    const int n_seg = (ns + SEG - 1) / SEG;
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) & ~(align - 1); };
    const size_t part_bytes = up(static_cast<size_t>(fwd_blocks) * 16 * sizeof(double));
    double* partials = static_cast<double*>(scratch(kSlotPartials, part_bytes + 256 * sizeof(double)));
    double* batches = partials + part_bytes / sizeof(double);
    s.partials = partials;
    R* xbar = nullptr;
    if (adjoint) {
        const size_t ck_b = up(static_cast<size_t>(n_seg) * n * n_legs * sizeof(R));
        s.ck = static_cast<R*>(scratch(kSlotCheckpoints, (VAR ? 2 : 1) * ck_b));
        s.ckv = VAR ? reinterpret_cast<R*>(reinterpret_cast<unsigned char*>(s.ck) + ck_b) : nullptr;
        const size_t perf_b = up(static_cast<size_t>(rq.n_obs) * n_legs * sizeof(R));
        const size_t worst_b = up(static_cast<size_t>(rq.n_obs) * n_legs);
        const size_t leg_b = up(static_cast<size_t>(n_legs) * sizeof(R));
        const size_t lat_b = up(static_cast<size_t>(n_legs) * sizeof(int));
        const size_t xbar_b = up(static_cast<size_t>(n) * n_legs * sizeof(R));
        unsigned char* m = static_cast<unsigned char*>(scratch(kSlotSeeds, perf_b + worst_b + leg_b + lat_b + xbar_b));
        s.perf = reinterpret_cast<R*>(m);
        s.worst = m + perf_b;
        s.lseed = reinterpret_cast<R*>(m + perf_b + worst_b);
        s.lat = reinterpret_cast<int*>(m + perf_b + worst_b + leg_b);
        xbar = reinterpret_cast<R*>(m + perf_b + worst_b + leg_b + lat_b);
    }

    Events ev;
    hip_check(hipEventRecord(ev.e[0], stream), "event");
    hipLaunchKernelGGL((autocall_forward<R, N, SEG, VAR>), dim3(static_cast<unsigned>(fwd_blocks)), dim3(kFwdBlock), 0, stream,
                       s);
    hip_check(hipGetLastError(), "autocall forward launch");
    hipLaunchKernelGGL(batch_sums_kernel, dim3(1), dim3(kBatches * 16), 0, stream, partials, fwd_blocks, batches);
    hip_check(hipGetLastError(), "batch sums launch");
    hip_check(hipEventRecord(ev.e[1], stream), "event");

    std::vector<double> node_out;
    long long bwd_blocks = 0, n_chunks = 0;
    int xl_legs = 0;
    if (adjoint) {
        size_t lds = static_cast<size_t>(SEG) * N * n_x * sizeof(unsigned long long);
        int max_lds = 0;
        int dev_id = 0;
        hip_check(hipGetDevice(&dev_id), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id), "attribute");
        if (lds > static_cast<size_t>(max_lds)) invalid("n_x too large for the adjoint's LDS rows");
        const void* bwd_fn = reinterpret_cast<const void*>(autocall_backward<R, N, SEG, VAR>);
        auto resident = [&](size_t bytes) {
            hip_check(hipFuncSetAttribute(bwd_fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)),
                      "LDS attribute");
            int per = 0;
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, autocall_backward<R, N, SEG, VAR>, kBwdBlock,
                                                                   bytes),
                      "occupancy");
            return per;
        };
        const int per_cu = resident(lds);
        n_chunks = (n_legs + kBwdBlock - 1) / kBwdBlock;
        bwd_blocks = std::min<long long>(n_chunks, static_cast<long long>(std::max(per_cu, 1)) * device_cu_count());
        bwd_blocks = std::max<long long>(bwd_blocks, 1);
        // x-bar in LDS beside the rows where it fits without costing a resident workgroup: no HBM round trip of
        // n_names x n_legs values a segment
        const long long xl = (n_chunks + bwd_blocks - 1) / bwd_blocks * kBwdBlock;
        const size_t xl_bytes = static_cast<size_t>(N) * xl * sizeof(R);
        if (lds + xl_bytes <= static_cast<size_t>(max_lds) && resident(lds + xl_bytes) >= per_cu) {
            lds += xl_bytes;
            xl_legs = static_cast<int>(xl);
        } else {
            resident(lds);
        }
        const long long entries = static_cast<long long>(n) * ns * n_x;
        // the 16 batch tables, one after the other ([batch][name][step][x], 64-bit fixed point), filled by the
        // backward kernel's atomics: zeroed first
        long long* batch = static_cast<long long*>(
            scratch(kSlotBatches, static_cast<size_t>(kBatches) * entries * sizeof(long long) +
                                      static_cast<size_t>(n) * kBatches * n_nodes * sizeof(double)));
        double* vega_dev = reinterpret_cast<double*>(batch + kBatches * entries);
        hip_check(hipMemsetAsync(batch, 0, static_cast<size_t>(kBatches) * entries * sizeof(long long), stream),
                  "clear batch tables");
        hipLaunchKernelGGL((autocall_backward<R, N, SEG, VAR>), dim3(static_cast<unsigned>(bwd_blocks)), dim3(kBwdBlock), lds,
                           stream, s, n_chunks, xbar, reinterpret_cast<unsigned long long*>(batch), xl_legs);
        hip_check(hipGetLastError(), "autocall backward launch");
        SurfaceSet set{};
        DerivSet dset{};
        set.n = n;
        for (int a = 0; a < n; ++a) {
            set.s[a] = *rows[a];
            dset.d[a] = surface_node_derivatives(rq.surface[a], stream);
        }
        launch_node_contraction(set, dset, grid, batch, kBatches, 1.0 / kFixed, vega_dev, stream, true);
        node_out.resize(static_cast<size_t>(n) * kBatches * n_nodes);
        hip_check(hipEventRecord(ev.e[2], stream), "event");
        hip_check(hipMemcpyAsync(node_out.data(), vega_dev, node_out.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 stream),
                  "copy vegas");
    } else {
        hip_check(hipEventRecord(ev.e[2], stream), "event");
    }
    std::vector<double> bsum(kBatches * 16);
    hip_check(hipMemcpyAsync(bsum.data(), batches, bsum.size() * sizeof(double), hipMemcpyDeviceToHost, stream),
              "copy sums");
    hip_check(hipStreamSynchronize(stream), "synchronize");

    // price and deltas from pair averages; batch prices over the forward workgroups of each batch
    double tot[16] = {0};
    for (int b = 0; b < kBatches; ++b)
        for (int k = 0; k < 16; ++k) tot[k] += bsum[b * 16 + k];
    const double pairs = static_cast<double>(rq.n_pairs);
    mean_se(tot[0], tot[1], pairs, rq.price, rq.se);
    for (int b = 0; b < kBatches; ++b) {
        const long long l0 = std::min(n_legs, (b * fwd_blocks / kBatches) * kFwdBlock);
        const long long l1 = std::min(n_legs, ((b + 1) * fwd_blocks / kBatches) * kFwdBlock);
        rq.batch[b] = l1 > l0 ? bsum[b * 16] / (0.5 * static_cast<double>(l1 - l0)) : 0.0;
    }
    for (int a = 0; a < kMaxNames; ++a) rq.delta[a] = rq.delta_se[a] = 0.0;
    if (adjoint) {
        for (int a = 0; a < n; ++a) mean_se(tot[2 + 2 * a], tot[3 + 2 * a], pairs, rq.delta[a], rq.delta_se[a]);
        // vegas: sum over legs / n_legs, per vol point; SE by batch means weighted by each batch's legs
        double legs_b[kBatches];
        for (int b = 0; b < kBatches; ++b) {
            const long long w0 = b * bwd_blocks / kBatches, w1 = (b + 1) * bwd_blocks / kBatches;
            const long long l0 = std::min(n_legs, (w0 * n_chunks / bwd_blocks) * kBwdBlock);
            const long long l1 = std::min(n_legs, (w1 * n_chunks / bwd_blocks) * kBwdBlock);
            legs_b[b] = static_cast<double>(l1 - l0);
        }
        int used = 0;
        for (int b = 0; b < kBatches; ++b) used += legs_b[b] > 0;
        for (int a = 0; a < n; ++a)
            for (int k = 0; k < n_nodes; ++k) {
                double total = 0.0;
                for (int b = 0; b < kBatches; ++b) total += node_out[(static_cast<size_t>(a) * kBatches + b) * n_nodes + k];
                const double mean = 0.01 * total / static_cast<double>(n_legs);
                double var = 0.0;
                for (int b = 0; b < kBatches; ++b) {
                    if (legs_b[b] <= 0) continue;
                    const double vb = 0.01 * node_out[(static_cast<size_t>(a) * kBatches + b) * n_nodes + k] / legs_b[b];
                    const double wgt = legs_b[b] / static_cast<double>(n_legs);
                    var += wgt * wgt * (vb - mean) * (vb - mean);
                }
                if (rq.vega) rq.vega[a * n_nodes + k] = mean;
                if (rq.vega_se) rq.vega_se[a * n_nodes + k] = used > 1 ? std::sqrt(var * used / (used - 1.0)) : 0.0;
            }
    }
    rq.forward_ms = ev.ms(0, 1);
    rq.backward_ms = adjoint ? ev.ms(1, 2) : 0.0;
    rq.checkpoint_used = adjoint ? SEG : 0;
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

// The checkpoint intervals compiled for N names (a segment's rows, SEG x N x 801 x 8 B at the risk grid's n_x,
// within 160 KiB of LDS).
template <int N>
constexpr bool seg_compiled(int seg) {
    return seg == 4 || (seg == 8 && N <= 3) || (seg == 16 && N == 1);
}

template <typename R, int N, bool VAR>
void run_seg(optlab_autocall_request& rq, const optlab_autocall_models* m, bool adjoint, int seg) {
    if (seg == 8 && N <= 3) return run_typed<R, N, (N <= 3 ? 8 : 4), VAR>(rq, m, adjoint);
    if (seg == 16 && N == 1) return run_typed<R, N, (N == 1 ? 16 : 4), VAR>(rq, m, adjoint);
    return run_typed<R, N, 4, VAR>(rq, m, adjoint);
}

// Autotuned checkpoint interval per shape (precision, names, steps, legs, n_x): more steps a segment means fewer
// checkpoints to write and read and fewer segment passes, but more state held in registers while a segment is walked
// back and more LDS a workgroup. The first adjoint of a shape with checkpoint_every = 0 runs once untimed (buffers
// sized, tables built and cached), then every candidate three times, and keeps the fastest by its best device time
// (the same request: the outputs are overwritten by the call that follows).
struct SegKey {
    int prec, n, n_steps, n_x;
    long long n_legs;
    bool var;
    bool operator<(const SegKey& o) const {
        if (var != o.var) return var < o.var;
        if (prec != o.prec) return prec < o.prec;
        if (n != o.n) return n < o.n;
        if (n_steps != o.n_steps) return n_steps < o.n_steps;
        if (n_x != o.n_x) return n_x < o.n_x;
        return n_legs < o.n_legs;
    }
};
std::map<SegKey, int> g_seg;

template <typename R, int N, bool VAR>
int pick_seg(optlab_autocall_request& rq, const optlab_autocall_models* m) {
    const int n_x = static_cast<int>(surface_rows(rq.surface[0]).n_x);
    const SegKey key{sizeof(R) == 4 ? 1 : 0, N, rq.n_steps, n_x, 2 * rq.n_pairs, VAR};
    const auto it = g_seg.find(key);
    if (it != g_seg.end()) return it->second;
    int dev_id = 0, max_lds = 0;
    hip_check(hipGetDevice(&dev_id), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id), "attribute");
    int best = 4;
    double best_ms = 0.0;
    {
        optlab_autocall_request t = rq;
        run_seg<R, N, VAR>(t, m, true, 4);
    }
    for (const int seg : {4, 8, 16}) {
        if (!seg_compiled<N>(seg) || static_cast<size_t>(seg) * N * n_x * 8 > static_cast<size_t>(max_lds)) continue;
        double ms = 0.0;
        for (int rep = 0; rep < 3; ++rep) {
            optlab_autocall_request t = rq;
            run_seg<R, N, VAR>(t, m, true, seg);
            const double m = t.forward_ms + t.backward_ms;
            ms = rep == 0 ? m : std::min(ms, m);
        }
        if (seg == 4 || ms < best_ms) best = seg, best_ms = ms;
    }
    g_seg[key] = best;
    return best;
}

template <typename R, int N, bool VAR>
void run_names(optlab_autocall_request& rq, const optlab_autocall_models* m, bool adjoint) {
    if (!adjoint) return run_typed<R, N, 4, VAR>(rq, m, false);  // a price takes no checkpoints
    int seg = rq.checkpoint_every;
    if (seg == 0) seg = pick_seg<R, N, VAR>(rq, m);
    if (!seg_compiled<N>(seg)) invalid("checkpoint_every must be 0, 4, 8 (up to 3 names) or 16 (one name)");
    run_seg<R, N, VAR>(rq, m, true, seg);
}

template <typename R, bool VAR>
void run_prec(optlab_autocall_request& rq, const optlab_autocall_models* m, bool adjoint) {
    switch (rq.n_names) {
        case 1: return run_names<R, 1, VAR>(rq, m, adjoint);
        case 2: return run_names<R, 2, VAR>(rq, m, adjoint);
        case 3: return run_names<R, 3, VAR>(rq, m, adjoint);
        case 4: return run_names<R, 4, VAR>(rq, m, adjoint);
        default: return run_names<R, 5, VAR>(rq, m, adjoint);
    }
}

}  // namespace

void check_autocall_request(const optlab_autocall_request& c, bool adjoint) {
    if (c.n_names < 1 || c.n_names > kMaxNames) invalid("n_names must be 1 to 5");
    if (c.n_steps < 1 || !(c.dt > 0)) invalid("n_steps must be >= 1 and dt > 0");
    if (!c.surface || !c.root || !c.fwd || !c.fixing || !c.obs_step || !c.disc) invalid("null array");
    if (c.n_obs < 1 || c.n_obs > kMaxObs) invalid("n_obs must be 1 to 64");
    int prev = 0;
    for (int j = 0; j < c.n_obs; ++j) {
        if (c.obs_step[j] <= prev || c.obs_step[j] > c.n_steps) invalid("obs_step must increase within [1, n_steps]");
        prev = c.obs_step[j];
    }
    if (prev != c.n_steps) invalid("the last observation must be at n_steps");
    for (int a = 0; a < c.n_names; ++a) {
        if (!(c.fixing[a] > 0)) invalid("fixings must be > 0");
        for (int t = 0; t <= c.n_steps; ++t)
            if (!(c.fwd[a * (c.n_steps + 1) + t] > 0)) invalid("forwards must be > 0");
        if (c.spot && !(c.spot[a] > 0)) invalid("spot factors must be > 0");
    }
    if (!(c.smoothing >= 0)) invalid("smoothing must be >= 0");
    if (c.n_pairs < 1 || c.first_pair < 0) invalid("n_pairs must be >= 1 and first_pair >= 0");
    if (c.n_pairs > (1ll << 40)) invalid("n_pairs too large");
    if (adjoint && c.nodes) invalid("adjoint Greeks are taken at the unbumped surfaces: nodes must be NULL");
    const int k = c.checkpoint_every;
    if (adjoint && !(k == 0 || k == 4 || (k == 8 && c.n_names <= 3) || (k == 16 && c.n_names == 1)))
        invalid("checkpoint_every must be 0, 4, 8 (up to 3 names) or 16 (one name)");
    int n_x = -1;
    for (int a = 0; a < c.n_names; ++a) {
        const SurfaceRows& r = surface_rows(c.surface[a]);
        if (r.n_rows != c.n_steps) invalid("each surface needs one row per step");
        if (n_x >= 0 && r.n_x != n_x) invalid("the surfaces must share n_x");
        n_x = r.n_x;
    }
}

void check_autocall_models(const optlab_autocall_request& c, const optlab_autocall_models& m) {
    if (!m.kind) invalid("null array");
    bool slv = false;
    for (int a = 0; a < c.n_names; ++a) {
        const int k = m.kind[a];
        if (k < 0 || k > 3) invalid("kind must be 0 (constant), 1 (local vol), 2 (Heston) or 3 (SLV)");
        if (k == 0 && !(m.cvar && m.cvar[a] > 0)) invalid("constant variance must be > 0");
        if (k == 2 || k == 3) {
            if (!m.heston) invalid("null array");
            const double* h = m.heston + 5 * a;
            if (!(h[0] >= 0 && h[1] >= 0 && h[2] >= 0 && h[3] >= 0 && std::fabs(h[4]) <= 1))
                invalid("Heston parameters out of range");
        }
        slv |= k == 3;
    }
    if (slv && !m.lev2) invalid("null array");
}

void run_autocall(optlab_autocall_request& rq, int prec, bool adjoint, const optlab_autocall_models* models) {
    bool var = false;
    if (models)
        for (int a = 0; a < rq.n_names; ++a) var |= models->kind[a] == 2 || models->kind[a] == 3;
    if (prec == 0) {
        if (var) run_prec<double, true>(rq, models, adjoint);
        else run_prec<double, false>(rq, models, adjoint);
    } else {
        if (var) run_prec<float, true>(rq, models, adjoint);
        else run_prec<float, false>(rq, models, adjoint);
    }
}

}  // namespace optlab_mi
