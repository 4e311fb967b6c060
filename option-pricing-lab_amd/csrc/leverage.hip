// SLV leverage by the particle method (Guyon and Henry-Labordere), entirely on the device.
// Semantics: packages/native/cuda/src/leverage.cpp:17-186 (the host algorithm and the CPU particles) and
// include/optlab_cuda/leverage.hpp:13-53.
//
// Per step i, five kernels on one stream, no host round trip (the reference copies the bins to the host and the
// leverage row back, four synchronous copies per step):
//   bin   -- every block first reduces the previous step's per-block ranges (min x, max x, max |v|) in a fixed
//            order, then bins its particles in LDS: a fine histogram (u32) for the quantiles and, on the regression
//            bins, count and four fixed-point sums (v, d, d^2, d v with d = (x - centre) / width in [-1/2, 1/2]),
//            added with 64-bit integer LDS atomics, so the totals do not depend on the order of the additions;
//   reduce-- adds the per-block bins in a fixed order;
//   prep  -- one workgroup: quantiles by a block scan of the fine histogram, the 81-point grid, the bandwidth
//            and each regression bin's moments;
//   fit   -- one wave per grid point, across the chip: the Gaussian Nadaraya-Watson (or local-linear) regression
//            of E[v | x] on the bin centres in FP64 (fixed-order wave reductions), the FP64 division
//            L = clip(sqrt(sigma_loc^2 / E[v | x])), and the leverage row for the step kernel (the regression is
//            most of a step's FP64 work: on one CU it was the longest kernel of the step);
//   step  -- one pair per lane, the leverage row staged in LDS, one Euler step of both legs (Philox counter
//            (pair, i + 1, 0), leverage.cpp:197-208), block min / max / max |v| for the next bin.
// Particles are FP32 (backend 1) or FP64 (backend 0); everything after the bins is FP64.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

constexpr int kBlock = 256;
constexpr int kRowThreads = 1024;
constexpr int kFitThreads = 256;  // four grid points a workgroup
constexpr int kMaxFine = 8192;
constexpr int kMaxBins = 512;
constexpr int kMaxX = 1024;

struct Shared {  // device-resident state of the calibration between kernels
    double lo, hi, vmax;  // range of x and max |v| of the particles before this step's binning
    double a, b;          // ends of the current leverage row's grid
    int shift_d, shift_v;
    double bw, fallback;  // this step's kernel bandwidth and E[v] (the regression's fallback)
};

struct LevArgs {
    long long n_pairs;
    int n_steps, n_x, n_bins, n_fine, n_lv, local_linear, positive_part;
    std::uint32_t k0, k1;
    double v0, kappa, theta, xi, rho, rc;
    double atm_vol, bandwidth_scale, l_min, l_max, q_lo, q_hi;
    int bw_rule;     // 0: the reference's bandwidth; 1: bw_coef std(x) N^-0.2 over the particles
    double bw_coef;
    double n_pow;    // N^-0.2, N the number of particles (formed on the host, as leverage.cpp:63 does)
    const double* dt;
    const double* t0;
    const double* lv_lo;
    const double* lv_inv_dx;
    const double* lv_rows;
    double* x_grids;  // device, n_steps x n_x
    double* values;   // device, n_steps x n_x
    Shared* shared;
    double* range_part;            // [step blocks][3]
    int n_step_blocks;
    unsigned* fine_part;           // [bin blocks][n_fine]
    long long* sum_part;           // [bin blocks][5 n_bins]
    int n_bin_blocks;
    unsigned long long* fine_tot;  // [n_fine]
    long long* sum_tot;            // [5 n_bins]
    double* x_eval;                // [n_x]: this step's grid
    double* bins;                  // [6][n_bins]: centre, count, sum v, sum x, sum x^2, sum x v of each bin
};

template <typename R>
struct Particles {
    R* x;
    R* v;
    R* lev;  // the current leverage row, n_x values
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

constexpr int kBinThreads = 1024;  // four waves a SIMD: the LDS atomics' and loads' latencies overlap

// Reduces the step kernel's per-block (min x, max x, max |v|): minima and maxima, the same in any order (wave
// butterflies, then the waves' results in LDS).
template <int T>
__device__ void reduce_range(const LevArgs& s, double& lo, double& hi, double& vmax) {
    __shared__ double red[3][T / 64];
    double mn = INFINITY, mx = -INFINITY, mv = 0.0;
    for (int b = threadIdx.x; b < s.n_step_blocks; b += blockDim.x) {
        mn = fmin(mn, s.range_part[3 * b]);
        mx = fmax(mx, s.range_part[3 * b + 1]);
        mv = fmax(mv, s.range_part[3 * b + 2]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, m, 64));
        mx = fmax(mx, __shfl_xor(mx, m, 64));
        mv = fmax(mv, __shfl_xor(mv, m, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[0][wave] = mn, red[1][wave] = mx, red[2][wave] = mv;
    __syncthreads();
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        mn = fmin(mn, red[0][w]);
        mx = fmax(mx, red[1][w]);
        mv = fmax(mv, red[2][w]);
    }
    lo = mn, hi = mx, vmax = mv;
    __syncthreads();
}

__device__ __forceinline__ int shift_bits(long long n_particles) {
    int b = 0;
    while ((1LL << b) <= n_particles) ++b;  // n < 2^b
    return b;
}

template <typename R>
__global__ void __launch_bounds__(kBinThreads) bin_kernel(LevArgs s, Particles<R> P) {
    __shared__ unsigned fine[kMaxFine];
    __shared__ long long sums[5 * kMaxBins];
    double lo, hi, vmax;
    reduce_range<kBinThreads>(s, lo, hi, vmax);
    const int nf = s.n_fine, nb = s.n_bins;
    const long long n = 2 * s.n_pairs;
    // fixed-point scales: n |d| < 2^62 with |d| <= 1/2 + eps, and n max|v| < 2^61
    const int bn = shift_bits(n);
    int ev = 0;
    if (vmax > 0) frexp(vmax, &ev);  // vmax < 2^ev
    const int shift_d = 61 - bn;
    const int shift_v = min(52, 61 - bn - max(ev, -8));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s.shared->lo = lo, s.shared->hi = hi, s.shared->vmax = vmax;
        s.shared->shift_d = shift_d, s.shared->shift_v = shift_v;
    }
    for (int k = threadIdx.x; k < nf; k += blockDim.x) fine[k] = 0u;
    for (int k = threadIdx.x; k < 5 * nb; k += blockDim.x) sums[k] = 0;
    __syncthreads();
    const double sd = ldexp(1.0, shift_d), sv = ldexp(1.0, shift_v);
    // bin indices in the particle precision (the reference GPU bins in float, leverage_gpu.cu:43-50)
    const R rlo = static_cast<R>(lo), rhi = static_cast<R>(hi);
    const R wf = (rhi - rlo) / static_cast<R>(nf);
    const R we = (rhi + static_cast<R>(1e-12) - rlo) / static_cast<R>(nb);
    const double wed = (hi + 1e-12 - lo) / nb;
    for (long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
         j += static_cast<long long>(gridDim.x) * blockDim.x) {
        const R xr = P.x[j];
        const R vr = s.positive_part ? r_max(P.v[j], R(0)) : P.v[j];
        int f = wf > R(0) ? static_cast<int>((xr - rlo) / wf) : 0;
        f = min(max(f, 0), nf - 1);
        int k = static_cast<int>((xr - rlo) / we);
        k = min(max(k, 0), nb - 1);
        atomicAdd(&fine[f], 1u);
        const double xd = static_cast<double>(xr), vd = static_cast<double>(vr);
        const double centre = lo + (k + 0.5) * wed;
        const double d = (xd - centre) / wed;
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[k]), 1ull);
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[nb + k]),
                  static_cast<unsigned long long>(__double2ll_rn(vd * sv)));
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * nb + k]),
                  static_cast<unsigned long long>(__double2ll_rn(d * sd)));
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[3 * nb + k]),
                  static_cast<unsigned long long>(__double2ll_rn(d * d * sd)));
        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[4 * nb + k]),
                  static_cast<unsigned long long>(__double2ll_rn(d * vd * sv)));
    }
    __syncthreads();
    unsigned* fp = s.fine_part + static_cast<long long>(blockIdx.x) * nf;
    long long* sp = s.sum_part + static_cast<long long>(blockIdx.x) * 5 * nb;
    for (int k = threadIdx.x; k < nf; k += blockDim.x) fp[k] = fine[k];
    for (int k = threadIdx.x; k < 5 * nb; k += blockDim.x) sp[k] = sums[k];
}

// Totals of the workgroups' bins: a workgroup takes 64 slots, its 16 waves each add every sixteenth workgroup's row
// (a short chain of loads a wave; integer sums give the same bits in any order), then wave 0 adds the partials.
constexpr int kReduceThreads = 1024;

__global__ void __launch_bounds__(kReduceThreads) reduce_bins_kernel(LevArgs s) {
    __shared__ unsigned long long part[kReduceThreads / 64][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = kReduceThreads / 64;
    const int slot = blockIdx.x * 64 + lane;
    const int nf = s.n_fine, ns = 5 * s.n_bins;
    unsigned long long t = 0;
    if (slot < nf) {
#pragma unroll 8
        for (int b = wave; b < s.n_bin_blocks; b += n_waves) t += s.fine_part[static_cast<long long>(b) * nf + slot];
    } else if (slot < nf + ns) {
        const int k = slot - nf;
#pragma unroll 8
        for (int b = wave; b < s.n_bin_blocks; b += n_waves)
            t += static_cast<unsigned long long>(s.sum_part[static_cast<long long>(b) * ns + k]);
    }
    part[wave][lane] = t;
    __syncthreads();
    if (wave == 0 && slot < nf + ns) {
        unsigned long long v = 0;
        for (int w = 0; w < n_waves; ++w) v += part[w][lane];
        if (slot < nf)
            s.fine_tot[slot] = v;
        else
            s.sum_tot[slot - nf] = static_cast<long long>(v);
    }
}

__device__ __forceinline__ double interp_uniform(double x, double a, double b, const double* y, int n) {
    double u = (x - a) / (b - a) * (n - 1);  // leverage.cpp:17-24
    if (!(u > 0)) return y[0];
    if (u >= n - 1) return y[n - 1];
    const int i = static_cast<int>(u);
    const double w = u - i;
    return y[i] * (1 - w) + y[i + 1] * w;
}

// leverage.cpp:26-38, 56-66: the grid of step i from the quantiles of the fine histogram (or the flat start at
// step 0), the bandwidth, E[v] and each regression bin's moments, for the fit kernel.
template <typename R>
__global__ void __launch_bounds__(kRowThreads) prep_kernel(LevArgs s, int i) {
    constexpr int kPer = kMaxFine / kRowThreads;  // fine bins a thread owns, at most
    __shared__ double scan[kRowThreads];
    __shared__ double qval[2];
    __shared__ int first_hit[2];
    const int nx = s.n_x, nb = s.n_bins, nf = s.n_fine;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n_waves = blockDim.x >> 6;
    const double n_particles = 2.0 * static_cast<double>(s.n_pairs);
    const double h = s.dt[i];
    if (i == 0) {
        for (int g = tid; g < nx; g += blockDim.x) s.x_eval[g] = (-6.0 + 12.0 * g / (nx - 1)) * s.atm_vol * sqrt(fmax(h, 1e-4));
        if (tid == 0) {
            s.shared->a = (-6.0) * s.atm_vol * sqrt(fmax(h, 1e-4));
            s.shared->b = (-6.0 + 12.0 * (nx - 1) / (nx - 1)) * s.atm_vol * sqrt(fmax(h, 1e-4));
        }
        return;
    }
    const double lo = s.shared->lo, hi = s.shared->hi, t0 = s.t0[i];
    // leverage.cpp:26-38 by a block scan: each thread owns a contiguous run of fine bins; the cumulative counts are
    // integers below 2^53, so every grouping of the additions gives the reference's sequential values.
    const int per = (nf + blockDim.x - 1) / blockDim.x;  // <= kPer
    const int k0 = min(nf, tid * per), k1 = min(nf, k0 + per);
    double fc[kPer];  // this thread's fine counts, read once
#pragma unroll
    for (int m = 0; m < kPer; ++m) fc[m] = k0 + m < k1 ? static_cast<double>(s.fine_tot[k0 + m]) : 0.0;
    double run = 0.0;
#pragma unroll
    for (int m = 0; m < kPer; ++m)
        if (k0 + m < k1) run += fc[m];
    // inclusive scan of the runs: in the wave by shuffles, then the waves' totals in order
    double incl = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) scan[wave] = incl;
    if (tid < 2) first_hit[tid] = INT_MAX;
    __syncthreads();
    double before = 0.0, total = 0.0;
    for (int w = 0; w < n_waves; ++w) {
        if (w < wave) before += scan[w];
        total += scan[w];
    }
    const double start = before + incl - run;  // the counts before this thread's first fine bin
    __syncthreads();
    const double target[2] = {s.q_lo * total, s.q_hi * total};
    for (int q = 0; q < 2; ++q) {
        double cum = start;
        bool hit = false;
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            if (hit || k0 + m >= k1) continue;
            const double c = fc[m];
            if (cum + c >= target[q] && c > 0) {
                atomicMin(&first_hit[q], k0 + m);
                hit = true;
            } else {
                cum += c;
            }
        }
    }
    __syncthreads();
    const double width = (hi - lo) / static_cast<double>(nf);
    for (int q = 0; q < 2; ++q) {
        const int k = first_hit[q];
        if (k == INT_MAX) {
            if (tid == 0) qval[q] = hi;
        } else if (k >= k0 && k < k1) {
            double cum = start, c = 0.0;
#pragma unroll
            for (int m = 0; m < kPer; ++m) {
                if (k0 + m < k) cum += fc[m];
                if (k0 + m == k) c = fc[m];
            }
            qval[q] = lo + (static_cast<double>(k) + (target[q] - cum) / c) * width;
        }
    }
    __syncthreads();
    const double a = qval[0], b = qval[1];
    for (int g = tid; g < nx; g += blockDim.x) s.x_eval[g] = a + (b - a) * g / (nx - 1);
    if (tid == 0) s.shared->a = a, s.shared->b = a + (b - a) * (nx - 1) / (nx - 1);
    // each regression bin's moments from the fixed-point totals:
    // sum x = n c + w S_d, sum x^2 = n c^2 + 2 c w S_d + w^2 S_dd, sum x v = c S_v + w S_dv
    const double we = (hi + 1e-12 - lo) / nb;
    const double inv_d = ldexp(1.0, -s.shared->shift_d), inv_v = ldexp(1.0, -s.shared->shift_v);
    const bool moments = s.local_linear || s.bw_rule == 1;
    double* B = s.bins;
    double sx = 0.0, sxx = 0.0, vs = 0.0;
    for (int k = tid; k < nb; k += blockDim.x) {
        const double centre = lo + (k + 0.5) * we, cnt = static_cast<double>(s.sum_tot[k]);
        const double sv = static_cast<double>(s.sum_tot[nb + k]) * inv_v;
        B[k] = centre, B[nb + k] = cnt, B[2 * nb + k] = sv;
        if (moments) {
            const double sd = static_cast<double>(s.sum_tot[2 * nb + k]) * inv_d;
            const double sdd = static_cast<double>(s.sum_tot[3 * nb + k]) * inv_d;
            const double sdv = static_cast<double>(s.sum_tot[4 * nb + k]) * inv_v;
            const double bx = cnt * centre + we * sd;
            const double bxx = cnt * centre * centre + 2.0 * centre * we * sd + we * we * sdd;
            B[3 * nb + k] = bx, B[4 * nb + k] = bxx, B[5 * nb + k] = centre * sv + we * sdv;
            sx += bx;
            sxx += bxx;
        }
    }
    // E[v] (the fallback of an empty neighbourhood): wave 0, lanes over the bins in the fit's order
    if (wave == 0) {
        for (int k = lane; k < nb; k += 64) vs += static_cast<double>(s.sum_tot[nb + k]) * inv_v;
        vs = wave_sum(vs);
    }
    double bw = s.bandwidth_scale * s.atm_vol * sqrt(fmax(t0, 0.15)) * s.n_pow;
    if (s.bw_rule == 1) {  // std(x) over the particles from the bins
        sx = wave_sum(sx);
        sxx = wave_sum(sxx);
        if (lane == 0) scan[wave] = sx, scan[n_waves + wave] = sxx;
        __syncthreads();
        double tx = 0.0, txx = 0.0;
        for (int w = 0; w < n_waves; ++w) tx += scan[w], txx += scan[n_waves + w];
        const double mean = tx / n_particles, var = fmax(txx / n_particles - mean * mean, 0.0);
        bw = s.bw_coef * sqrt(var) * s.n_pow;
    }
    if (tid == 0) s.shared->bw = bw, s.shared->fallback = vs / n_particles;
}

// leverage.cpp:67-92, 141-148: one wave per grid point g of step i: the regression of E[v | x] at x_g over the
// bins (lane l takes bins l, l + 64, ..., then a fixed butterfly), and the leverage there.
template <typename R>
__global__ void __launch_bounds__(kFitThreads) fit_kernel(LevArgs s, Particles<R> P, int i) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (kFitThreads / 64) + (threadIdx.x >> 6);
    const int nx = s.n_x, nb = s.n_bins;
    if (g >= nx) return;
    const double x0 = s.x_eval[g];
    double c = s.v0;
    if (i > 0) {
        const double bw = s.shared->bw;
        const double* B = s.bins;
        double S0 = 0, T0 = 0, Ax = 0, Axx = 0, Axv = 0;
        for (int k = lane; k < nb; k += 64) {
            const double d = (x0 - B[k]) / bw, w = exp(-0.5 * d * d);
            S0 += w * B[nb + k];
            T0 += w * B[2 * nb + k];
            if (s.local_linear) {
                Ax += w * B[3 * nb + k];
                Axx += w * B[4 * nb + k];
                Axv += w * B[5 * nb + k];
            }
        }
        S0 = wave_sum(S0);
        T0 = wave_sum(T0);
        if (s.local_linear) {
            Ax = wave_sum(Ax);
            Axx = wave_sum(Axx);
            Axv = wave_sum(Axv);
        }
        const double nw = S0 > 1e-8 ? T0 / fmax(S0, 1e-300) : s.shared->fallback;
        c = nw;
        if (s.local_linear) {
            const double S1 = Ax - x0 * S0, S2 = Axx - 2 * x0 * Ax + x0 * x0 * S0, T1 = Axv - x0 * T0;
            const double det = S0 * S2 - S1 * S1;
            const bool ok = S0 >= 50 && det > 1e-12 * S0 * fabs(S2);
            c = fmax(ok ? (S2 * T0 - S1 * T1) / det : nw, 1e-8);
        }
    }
    if (lane != 0) return;
    const double* row = s.lv_rows + static_cast<long long>(i) * s.n_lv;
    const double lv_lo = s.lv_lo[i], lv_hi = lv_lo + (s.n_lv - 1) / s.lv_inv_dx[i];
    const double loc = interp_uniform(x0, lv_lo, lv_hi, row, s.n_lv);
    const double L = fmin(fmax(sqrt(loc / fmax(c, 1e-8)), s.l_min), s.l_max);
    s.x_grids[static_cast<long long>(i) * nx + g] = x0;
    s.values[static_cast<long long>(i) * nx + g] = L;
    P.lev[g] = static_cast<R>(L);
}

// leverage.cpp:189-214: one Euler step of both legs of every pair on the leverage row.
template <typename R>
__global__ void __launch_bounds__(kBlock) step_kernel(LevArgs s, Particles<R> P, int i) {
    __shared__ R lev[kMaxX];
    __shared__ double red[3][kBlock / 64];
    const int nx = s.n_x;
    for (int g = threadIdx.x; g < nx; g += blockDim.x) lev[g] = P.lev[g];
    __syncthreads();
    const R a = static_cast<R>(s.shared->a), b = static_cast<R>(s.shared->b);
    const R h = static_cast<R>(s.dt[i]);
    const R rho = static_cast<R>(s.rho), rc = static_cast<R>(s.rc), kappa = static_cast<R>(s.kappa);
    const R theta = static_cast<R>(s.theta), xi = static_cast<R>(s.xi);
    const R span = b - a, top = static_cast<R>(nx - 1);
    double mn = INFINITY, mx = -INFINITY, mv = 0.0;
    const long long np = s.n_pairs;
    for (long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
         p += static_cast<long long>(gridDim.x) * blockDim.x) {
        const U4 r4 = philox10(U4{static_cast<std::uint32_t>(p), static_cast<std::uint32_t>(static_cast<std::uint64_t>(p) >> 32),
                                  static_cast<std::uint32_t>(i + 1), 0u},
                               s.k0, s.k1);
        R z0, z1;
        gauss2(r4.x, r4.y, z0, z1);
#pragma unroll
        for (int leg = 0; leg < 2; ++leg) {
            const long long j = leg ? np + p : p;
            const R sg = leg ? R(-1) : R(1), zs = sg * z0, zv = sg * z1;
            R x = P.x[j], v = P.v[j];
            const R u = (x - a) / span * top;
            R L;
            if (!(u > R(0))) {
                L = lev[0];
            } else if (u >= top) {
                L = lev[nx - 1];
            } else {
                const int k = static_cast<int>(u);
                const R w = u - static_cast<R>(k);
                L = lev[k] * (R(1) - w) + lev[k + 1] * w;
            }
            const R vp = r_max(v, R(0)), var = L * L * vp;
            const R zx = rho * zv + rc * zs;
            v += kappa * (theta - vp) * h + xi * r_sqrt(vp * h) * zv;
            x += R(-0.5) * var * h + r_sqrt(var * h) * zx;
            P.x[j] = x;
            P.v[j] = v;
            mn = fmin(mn, static_cast<double>(x));
            mx = fmax(mx, static_cast<double>(x));
            mv = fmax(mv, fabs(static_cast<double>(v)));
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, m, 64));
        mx = fmax(mx, __shfl_xor(mx, m, 64));
        mv = fmax(mv, __shfl_xor(mv, m, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[0][wave] = mn, red[1][wave] = mx, red[2][wave] = mv;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            mn = fmin(mn, red[0][w]);
            mx = fmax(mx, red[1][w]);
            mv = fmax(mv, red[2][w]);
        }
        mn = fmin(mn, red[0][0]), mx = fmax(mx, red[1][0]), mv = fmax(mv, red[2][0]);
        s.range_part[3 * blockIdx.x] = mn;
        s.range_part[3 * blockIdx.x + 1] = mx;
        s.range_part[3 * blockIdx.x + 2] = mv;
    }
}

template <typename R>
__global__ void init_kernel(Particles<R> P, long long n, R v0) {
    for (long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
         j += static_cast<long long>(gridDim.x) * blockDim.x) {
        P.x[j] = R(0);
        P.v[j] = v0;
    }
}

template <typename R>
void run_typed(optlab_leverage_request& rq, int bw_rule, double bw_coef) {
    const auto t_start = std::chrono::steady_clock::now();
    const int cus = device_cu_count();
    const long long np = rq.n_pairs, n = 2 * np;
    const int nx = rq.n_x, nb = rq.n_bins, nf = rq.n_fine;
    LevArgs s{};
    s.n_pairs = np, s.n_steps = rq.n_steps, s.n_x = nx, s.n_bins = nb, s.n_fine = nf, s.n_lv = rq.n_lv;
    s.local_linear = rq.local_linear, s.positive_part = rq.positive_part;
    s.k0 = static_cast<std::uint32_t>(rq.seed), s.k1 = static_cast<std::uint32_t>(rq.seed >> 32);
    s.v0 = rq.v0, s.kappa = rq.kappa, s.theta = rq.theta, s.xi = rq.xi, s.rho = rq.rho;
    s.rc = std::sqrt(1 - rq.rho * rq.rho);
    s.atm_vol = rq.atm_vol, s.bandwidth_scale = rq.bandwidth_scale, s.l_min = rq.l_min, s.l_max = rq.l_max;
    s.q_lo = rq.q_lo, s.q_hi = rq.q_hi;
    s.bw_rule = bw_rule, s.bw_coef = bw_coef;
    s.n_pow = std::pow(static_cast<double>(n), -0.2);
    s.n_step_blocks = static_cast<int>(std::max<long long>(1, std::min<long long>((np + kBlock - 1) / kBlock, 4LL * cus)));
    s.n_bin_blocks = static_cast<int>(std::max<long long>(1, std::min<long long>((n + kBinThreads - 1) / kBinThreads, cus)));
    const size_t rows = static_cast<size_t>(rq.n_steps);
    // one device block of memory for everything, carved by offsets
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t o_dt = carve(rows * 8), o_t0 = carve(rows * 8), o_lo = carve(rows * 8), o_inv = carve(rows * 8);
    const size_t o_lv = carve(rows * rq.n_lv * 8), o_xg = carve(rows * nx * 8), o_val = carve(rows * nx * 8);
    const size_t o_sh = carve(sizeof(Shared)), o_rp = carve(static_cast<size_t>(s.n_step_blocks) * 3 * 8);
    const size_t o_fp = carve(static_cast<size_t>(s.n_bin_blocks) * nf * 4);
    const size_t o_spart = carve(static_cast<size_t>(s.n_bin_blocks) * 5 * nb * 8);
    const size_t o_ft = carve(static_cast<size_t>(nf) * 8), o_st = carve(static_cast<size_t>(5 * nb) * 8);
    const size_t o_x = carve(static_cast<size_t>(n) * sizeof(R)), o_v = carve(static_cast<size_t>(n) * sizeof(R));
    const size_t o_lev = carve(static_cast<size_t>(nx) * sizeof(R));
    const size_t o_xe = carve(static_cast<size_t>(nx) * 8), o_bins = carve(static_cast<size_t>(6 * nb) * 8);
    unsigned char* base = static_cast<unsigned char*>(scratch(kSlotParticles, off));
    hip_check(hipMemcpy(base + o_dt, rq.dt, rows * 8, hipMemcpyHostToDevice), "copy dt");
    hip_check(hipMemcpy(base + o_t0, rq.t0, rows * 8, hipMemcpyHostToDevice), "copy t0");
    hip_check(hipMemcpy(base + o_lo, rq.lv_lo, rows * 8, hipMemcpyHostToDevice), "copy lv_lo");
    hip_check(hipMemcpy(base + o_inv, rq.lv_inv_dx, rows * 8, hipMemcpyHostToDevice), "copy lv_inv_dx");
    hip_check(hipMemcpy(base + o_lv, rq.lv_rows, rows * rq.n_lv * 8, hipMemcpyHostToDevice), "copy lv_rows");
    s.dt = reinterpret_cast<const double*>(base + o_dt), s.t0 = reinterpret_cast<const double*>(base + o_t0);
    s.lv_lo = reinterpret_cast<const double*>(base + o_lo), s.lv_inv_dx = reinterpret_cast<const double*>(base + o_inv);
    s.lv_rows = reinterpret_cast<const double*>(base + o_lv);
    s.x_grids = reinterpret_cast<double*>(base + o_xg), s.values = reinterpret_cast<double*>(base + o_val);
    s.shared = reinterpret_cast<Shared*>(base + o_sh), s.range_part = reinterpret_cast<double*>(base + o_rp);
    s.fine_part = reinterpret_cast<unsigned*>(base + o_fp), s.sum_part = reinterpret_cast<long long*>(base + o_spart);
    s.fine_tot = reinterpret_cast<unsigned long long*>(base + o_ft), s.sum_tot = reinterpret_cast<long long*>(base + o_st);
    s.x_eval = reinterpret_cast<double*>(base + o_xe), s.bins = reinterpret_cast<double*>(base + o_bins);
    Particles<R> P{reinterpret_cast<R*>(base + o_x), reinterpret_cast<R*>(base + o_v), reinterpret_cast<R*>(base + o_lev)};

    // The 4 kernels of every step, 4 n_steps launches that depend each on the last, are captured in one HIP graph
    // and launched at once: the gaps between dependent launches cost more than the kernels at 2^20 particles.
    hipStream_t st = engine_stream();
    const unsigned init_grid = static_cast<unsigned>(std::min<long long>((n + kBlock - 1) / kBlock, 4LL * cus));
    hipLaunchKernelGGL(init_kernel<R>, dim3(init_grid), dim3(kBlock), 0, st, P, n, static_cast<R>(rq.v0));
    hip_check(hipGetLastError(), "init kernel launch");
    const unsigned red_grid = static_cast<unsigned>((nf + 5 * nb + 63) / 64);  // 64 slots a workgroup
    const unsigned fit_grid = static_cast<unsigned>((nx + kFitThreads / 64 - 1) / (kFitThreads / 64));
    hip_check(hipStreamSynchronize(st), "synchronize");
    hipGraph_t graph;
    hipGraphExec_t exec;
    hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int i = 0; i < rq.n_steps; ++i) {
        if (i > 0) {
            hipLaunchKernelGGL(bin_kernel<R>, dim3(s.n_bin_blocks), dim3(kBinThreads), 0, st, s, P);
            hipLaunchKernelGGL(reduce_bins_kernel, dim3(red_grid), dim3(kReduceThreads), 0, st, s);
        }
        hipLaunchKernelGGL(prep_kernel<R>, dim3(1), dim3(kRowThreads), 0, st, s, i);
        hipLaunchKernelGGL(fit_kernel<R>, dim3(fit_grid), dim3(kFitThreads), 0, st, s, P, i);
        hipLaunchKernelGGL(step_kernel<R>, dim3(s.n_step_blocks), dim3(kBlock), 0, st, s, P, i);
    }
    hip_check(hipStreamEndCapture(st, &graph), "end capture");
    hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    hip_check(hipGraphLaunch(exec, st), "graph launch");
    hip_check(hipStreamSynchronize(st), "synchronize");
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    hip_check(hipMemcpy(rq.x_grids, s.x_grids, rows * nx * 8, hipMemcpyDeviceToHost), "copy x_grids");
    hip_check(hipMemcpy(rq.values, s.values, rows * nx * 8, hipMemcpyDeviceToHost), "copy values");
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace

void check_leverage_request(const optlab_leverage_request& rq) {  // requests.cpp:282-288
    if (!rq.dt || !rq.t0 || !rq.lv_lo || !rq.lv_inv_dx || !rq.lv_rows || !rq.x_grids || !rq.values)
        invalid("null pointer");
    if (rq.n_steps < 1 || rq.n_pairs < 2 || rq.n_x < 2 || rq.n_bins < 2 || rq.n_fine < 16 || rq.n_lv < 2)
        invalid("n_steps, n_pairs, n_x, n_bins, n_fine or n_lv too small");
    if (!(rq.kappa > 0 && rq.theta > 0 && rq.xi >= 0 && std::fabs(rq.rho) <= 1 && rq.v0 >= 0))
        invalid("Heston parameters out of range");
    if (rq.n_bins > kMaxBins) invalid("n_bins must be <= 512 on the GPU");
    if (rq.n_fine > kMaxFine) invalid("n_fine must be <= 8192 on the GPU");
    if (rq.n_x > kMaxX) invalid("n_x must be <= 1024 on the GPU");
    if (2 * rq.n_pairs >= (1LL << 40)) invalid("too many particles");
}

void run_leverage(optlab_leverage_request& rq, const optlab_leverage_options* opt, int prec) {
    if (opt && (opt->bandwidth_rule < 0 || opt->bandwidth_rule > 1)) invalid("bandwidth_rule must be 0 or 1");
    if (opt && opt->bandwidth_rule == 1 && !(opt->bandwidth_coef > 0)) invalid("bandwidth_coef must be positive");
    const int rule = opt ? opt->bandwidth_rule : 0;
    const double coef = opt ? opt->bandwidth_coef : 0.0;
    if (prec == 0)
        run_typed<double>(rq, rule, coef);
    else
        run_typed<float>(rq, rule, coef);
}

}  // namespace optlab_mi
