// SLV leverage by the particle method (Guyon and Henry-Labordere), entirely on the device.
// Semantics: packages/native/cuda/src/leverage.cpp:17-186 (the host algorithm and the CPU particles) and
// include/optlab_cuda/leverage.hpp:13-53.
//
// Two forms, the same bits: by default one persistent launch for the whole calibration (below, "the calibration as
// one persistent launch"); with OPTLAB_SLV_WORKGROUPS=0, or where the persistent launch cannot be resident, per step
// i five kernels on one stream, no host round trip (the reference copies the bins to the host and the leverage row
// back, four synchronous copies per step):
//   bin   -- every block first reduces the previous step's per-block ranges (min x, max x, max |v|) in a fixed
//            order, then bins its particles in LDS: a fine histogram (u32) for the quantiles and, on the regression
//            bins, count and four fixed-point sums (v, d, d^2, d v with d = (x - centre) / width in [-1/2, 1/2]),
//            added with 64-bit integer LDS atomics, so the totals do not depend on the order of the additions;
//   reduce-- adds the per-block bins in a fixed order;
//   prep  -- one workgroup: quantiles by a block scan of the fine histogram, the 81-point grid, the bandwidth
//            and each regression bin's moments;
//   fit   -- one lane per grid point: the Gaussian Nadaraya-Watson (or local-linear) regression of E[v | x] on
//            the bin centres in FP64, in the reference's order over the bins within 12 bandwidths of the point,
//            the FP64 division L = clip(sqrt(sigma_loc^2 / E[v | x])), and the leverage row for the step kernel;
//   step  -- one pair per lane, the leverage row staged in LDS, one Euler step of both legs (Philox counter
//            (pair, i + 1, 0), leverage.cpp:197-208), block min / max / max |v| for the next bin.
// Particles are FP32 (backend 1) or FP64 (backend 0); everything after the bins is FP64.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

constexpr int kBlock = 256;
constexpr int kRowThreads = 1024;
constexpr int kFitThreads = 64;  // a grid point a lane
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
    // the local variance from slice records instead of lv_rows (null: the rows): per step the last grid point and
    // a 24-double record (c0, c1, dc0, dc1, p0[5], p1[5], dp0[5], dp1[5]), and the Dupire clip
    const double* lv_hi;
    const double* lv_rec;
    double lv_floor, lv_cap;
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
#pragma clang fp contract(off)
    double u = (x - a) / (b - a) * (n - 1);  // leverage.cpp:17-24
    if (!(u > 0)) return y[0];
    if (u >= n - 1) return y[n - 1];
    const int i = static_cast<int>(u);
    const double w = u - i;
    return y[i] * (1 - w) + y[i + 1] * w;
}

// The Dupire local variance at log forward moneyness k of a row given as raw SVI slices: w = c0 S0 + c1 S1 with
// S = a + b (rho x + sqrt(x^2 + s2)), x = k - m, its k- and T-derivatives, Gatheral's g, clip(w_T / max(g, 1e-6),
// floor, cap). The operations and their order are the host table builder's (optlab/core/slices.py), with no
// contraction into fused multiply-adds: the value of every table entry is the host's, bit for bit.
__device__ double slice_dupire(const double* rec, double k, double floor, double cap) {
#pragma clang fp contract(off)
    double w = 0.0, wk = 0.0, wkk = 0.0, wT = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double c = rec[j], dc = rec[2 + j];
        const double a = rec[4 + 5 * j], b = rec[5 + 5 * j], rho = rec[6 + 5 * j], m = rec[7 + 5 * j];
        const double s2 = rec[8 + 5 * j];
        const double da = rec[14 + 5 * j], db = rec[15 + 5 * j], drho = rec[16 + 5 * j], dm = rec[17 + 5 * j];
        const double ds2 = rec[18 + 5 * j];
        const double x = k - m;
        const double root = sqrt(x * x + s2);
        const double lin = rho * x + root;
        const double s_k = b * (rho + x / root);
        const double s_kk = b * s2 / (root * root * root);
        const double s_t = da + db * lin + b * drho * x - s_k * dm + 0.5 * b * ds2 / root;
        const double S = a + b * lin;
        const double t0 = c * S, t1 = c * s_k, t2 = c * s_kk, t3 = dc * S + c * s_t;
        if (j == 0) {
            w = t0, wk = t1, wkk = t2, wT = t3;
        } else {
            w = w + t0, wk = wk + t1, wkk = wkk + t2, wT = wT + t3;
        }
    }
    const double u = 1.0 - 0.5 * k * wk / w;
    const double g = u * u - 0.25 * wk * wk * (1.0 / w + 0.25) + 0.5 * wkk;
    return fmin(fmax(wT / fmax(g, 1e-6), floor), cap);
}

// Entry j of a row of n points on [lo, hi] from its slice record: k = j (hi - lo) / (n - 1) + lo (the last point hi),
// rounded as the host's grid is (a product, then a sum; no fused multiply-add).
__device__ double slice_entry(const double* rec, int j, int n, double lo, double hi, double floor, double cap) {
#pragma clang fp contract(off)
    const double step = (hi - lo) / (n - 1);
    const double k = j == n - 1 ? hi : static_cast<double>(j) * step + lo;
    return slice_dupire(rec, k, floor, cap);
}

// sigma_loc^2 of step i at x0 (leverage.cpp:141-148): the row's table read by interp_uniform, or -- from slice
// records -- the same two table entries evaluated where they are needed.
__device__ double local_var_at(const LevArgs& s, int i, double x0) {
#pragma clang fp contract(off)
    const int n = s.n_lv;
    const double lo = s.lv_lo[i], end = lo + (n - 1) / s.lv_inv_dx[i];
    if (!s.lv_rec) return interp_uniform(x0, lo, end, s.lv_rows + static_cast<long long>(i) * n, n);
    const double* rec = s.lv_rec + static_cast<long long>(i) * 24;
    const double hi = s.lv_hi[i];
    const double u = (x0 - lo) / (end - lo) * (n - 1);
    if (!(u > 0)) return slice_entry(rec, 0, n, lo, hi, s.lv_floor, s.lv_cap);
    if (u >= n - 1) return slice_entry(rec, n - 1, n, lo, hi, s.lv_floor, s.lv_cap);
    const int j = static_cast<int>(u);
    const double w = u - j;
    const double y0 = slice_entry(rec, j, n, lo, hi, s.lv_floor, s.lv_cap);
    const double y1 = slice_entry(rec, j + 1, n, lo, hi, s.lv_floor, s.lv_cap);
    return y0 * (1 - w) + y1 * w;
}

// local_var_at in two parts, so that different lanes evaluate the two table entries of a point: part e (0 or 1) is
// the entry at or below x0 (e = 0) or above it (e = 1, used only between the ends); local_var_join combines them
// as local_var_at does.
__device__ double local_var_part(const LevArgs& s, int i, double x0, int e) {
#pragma clang fp contract(off)
    const int n = s.n_lv;
    const double lo = s.lv_lo[i], end = lo + (n - 1) / s.lv_inv_dx[i];
    const double u = (x0 - lo) / (end - lo) * (n - 1);
    int j;
    if (!(u > 0))
        j = 0;
    else if (u >= n - 1)
        j = n - 1;
    else
        j = static_cast<int>(u) + e;
    if (!s.lv_rec) return s.lv_rows[static_cast<long long>(i) * n + j];
    return slice_entry(s.lv_rec + static_cast<long long>(i) * 24, j, n, lo, s.lv_hi[i], s.lv_floor, s.lv_cap);
}

__device__ double local_var_join(const LevArgs& s, int i, double x0, double y0, double y1) {
#pragma clang fp contract(off)
    const int n = s.n_lv;
    const double lo = s.lv_lo[i], end = lo + (n - 1) / s.lv_inv_dx[i];
    const double u = (x0 - lo) / (end - lo) * (n - 1);
    if (!(u > 0) || u >= n - 1) return y0;
    const int j = static_cast<int>(u);
    const double w = u - j;
    return y0 * (1 - w) + y1 * w;
}

// E[v | x0] of step i > 0 from the bins (leverage.cpp:67-84): the Gaussian Nadaraya-Watson ratio, or the local-linear
// estimate, over the bins in increasing order as the reference adds them, skipping the bins more than 12 bandwidths
// from x0 (one bin of margin for the rounding of the window's ends): their weights exp(-d^2 / 2) < exp(-72) carry,
// over at most 2^22 particles, a total below 3e-25, under 1e-16 of any S0 the estimate uses (S0 > 1e-8, else the
// fallback; local-linear: S0 >= 50), so the result is the full sum's to the last digits. d = (x0 - c) / bw as the
// reference forms it. B: centre, count, sum v, sum x, sum x^2, sum x v of each bin (6 x nb); lo, hi the bins' range.
__device__ double fit_point(const LevArgs& s, const double* B, double x0, double bw, double fallback, double lo,
                            double hi) {
#pragma clang fp contract(off)
    const int nb = s.n_bins;
    int ka = 0, kb = nb - 1;
    const double reach = 12.0 * bw, we = (hi + 1e-12 - lo) / nb;
    if (bw > 0 && reach < INFINITY && we > 0) {
        const double fa = floor((x0 - reach - lo) / we - 0.5) - 1.0, fb = ceil((x0 + reach - lo) / we - 0.5) + 1.0;
        ka = fa <= 0.0 ? 0 : (fa >= nb ? nb : static_cast<int>(fa));
        kb = fb >= nb - 1 ? nb - 1 : (fb < 0.0 ? -1 : static_cast<int>(fb));
    }
    double S0 = 0, T0 = 0, Ax = 0, Axx = 0, Axv = 0;
    if (s.local_linear) {
        for (int k = ka; k <= kb; ++k) {
            const double d = (x0 - B[k]) / bw, w = exp(-0.5 * d * d);
            S0 += w * B[nb + k], T0 += w * B[2 * nb + k];
            Ax += w * B[3 * nb + k], Axx += w * B[4 * nb + k], Axv += w * B[5 * nb + k];
        }
    } else {
#pragma unroll 4
        for (int k = ka; k <= kb; ++k) {
            const double d = (x0 - B[k]) / bw, w = exp(-0.5 * d * d);
            S0 += w * B[nb + k], T0 += w * B[2 * nb + k];
        }
    }
    const double nw = S0 > 1e-8 ? T0 / fmax(S0, 1e-300) : fallback;
    if (!s.local_linear) return nw;
    const double S1 = Ax - x0 * S0, S2 = Axx - 2 * x0 * Ax + x0 * x0 * S0, T1 = Axv - x0 * T0;
    const double det = S0 * S2 - S1 * S1;
    const bool ok = S0 >= 50 && det > 1e-12 * S0 * fabs(S2);
    return fmax(ok ? (S2 * T0 - S1 * T1) / det : nw, 1e-8);
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
    for (int q = 0; q < 2; ++q) {  // each thread's first bin reaching the target, the wave's least, one LDS atomic
        double cum = start;
        int cand = INT_MAX;
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            if (cand != INT_MAX || k0 + m >= k1) continue;
            const double c = fc[m];
            if (cum + c >= target[q] && c > 0)
                cand = k0 + m;
            else
                cum += c;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) cand = min(cand, __shfl_xor(cand, off, 64));
        if ((threadIdx.x & 63) == 0 && cand != INT_MAX) atomicMin(&first_hit[q], cand);
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

// leverage.cpp:67-92, 141-148: one thread per grid point g of step i: E[v | x_g] from the bins (fit_point) and the
// leverage there.
template <typename R>
__global__ void __launch_bounds__(kFitThreads) fit_kernel(LevArgs s, Particles<R> P, int i) {
    const int g = blockIdx.x * kFitThreads + threadIdx.x;
    if (g >= s.n_x) return;
    const double x0 = s.x_eval[g];
    const double c = i > 0 ? fit_point(s, s.bins, x0, s.shared->bw, s.shared->fallback, s.shared->lo, s.shared->hi)
                           : s.v0;
    const double loc = local_var_at(s, i, x0);
    const double L = fmin(fmax(sqrt(loc / fmax(c, 1e-8)), s.l_min), s.l_max);
    s.x_grids[static_cast<long long>(i) * s.n_x + g] = x0;
    s.values[static_cast<long long>(i) * s.n_x + g] = L;
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

// ---- the calibration as one persistent launch ------------------------------------------------------------------
// Every workgroup keeps its particles in registers for the whole calibration (PPL pairs a lane: pair p = global
// lane + k x lanes), and a step is three phases between two grid-wide waits:
//   A. from the totals of step i's bins: the quantiles, the grid, the bins' moments and the bandwidth, then the
//      regression at every grid point and the leverage row in LDS. Every workgroup computes the same row, in the
//      five-kernel form's arithmetic (fit_point, a lane a grid point): a few hundred FP64 operations a lane,
//      cheaper than a third wait to broadcast it;
//   B. one Euler step of the workgroup's particles (leverage.cpp:189-214), the workgroup's range published;
//   C. after the wait for every range: the particles binned in LDS and the workgroup's nonzero bins added into the
//      step's totals with integer atomics (order-independent, so the totals are the five-kernel form's bits).
// Exchanged data is written and read with agent-scope relaxed atomics (write-through stores and loads that bypass
// the XCD's L2, coherent across XCDs) and every wave waits for its own memory operations before a wait: no L2
// write-back or invalidate a step. A wait is a two-level counter (eight group counters, whose last arrival adds to
// one top counter) spun on with a wall-clock bound: past it the launch sets an abort flag and every workgroup
// leaves, so a launch always ends. Totals are double-buffered by step parity and zeroed a step ahead.

constexpr int kPersistThreads = 512;
constexpr int kPersistMaxX = 256;
constexpr int kSyncStride = 32;  // unsigned ints between two counters: a 128-byte line each
constexpr int kSyncGroups = 8;

struct PersistState {
    double* ranges;        // [2][gridDim.x][4]: min x, max x, max |v| of each workgroup's particles
    unsigned* fine_tot;    // [2][n_fine]
    long long* sum_tot;    // [2][5 n_bins]
    unsigned* sync;        // [kSyncGroups group counters, top counter, abort flag] x kSyncStride
    unsigned long long spin_limit;  // wall-clock ticks a wait may spin
};

__device__ __forceinline__ unsigned load_agent(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait number b of the launch (0, 1, 2, ...): true when every workgroup has arrived, false when the launch is
// abandoned (a wait spun past its bound here or elsewhere).
__device__ bool grid_wait(const PersistState& ps, unsigned b) {
    __shared__ int go;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores and atomics have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, g = blockIdx.x % kSyncGroups;
        const unsigned in_group = (G - g + kSyncGroups - 1) / kSyncGroups;
        const unsigned groups = G < kSyncGroups ? G : kSyncGroups;
        unsigned* top = ps.sync + kSyncGroups * kSyncStride;
        unsigned* abort_flag = top + kSyncStride;
        const unsigned old = __hip_atomic_fetch_add(ps.sync + g * kSyncStride, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == in_group * (b + 1)) __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = groups * (b + 1);
        const unsigned long long t0 = wall_clock64();
        int ok = 1;
        while (load_agent(top) < target) {
            if (load_agent(abort_flag)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > ps.spin_limit) {
                __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        go = ok;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(go) != 0;
}

template <typename R, int PPL>
__global__ void __launch_bounds__(kPersistThreads) slv_persistent_kernel(LevArgs s, PersistState ps) {
    __shared__ alignas(16) unsigned fine[kMaxFine];  // phase C: the LDS histogram; phase A: the bins' moments (6 n_bins doubles)
    __shared__ long long sums[5 * kMaxBins];  // phase C: the regression bins' fixed-point sums
    __shared__ double xev[kPersistMaxX];
    __shared__ R lev[kPersistMaxX];
    __shared__ double red[3][kPersistThreads / 64];   // the workgroup's range
    __shared__ double red2[3][kPersistThreads / 64];  // the launch's range
    __shared__ double scan[4 * (kPersistThreads / 64)];
    __shared__ double lvv[2][kPersistMaxX];  // the two table entries around each grid point (or its one entry)
    __shared__ double qval[2];
    __shared__ int first_hit[2];
    __shared__ double sh_bw, sh_fallback;
    __shared__ double fitv[kPersistMaxX];  // E[v | x] at the grid points
    constexpr int kPer = kMaxFine / kPersistThreads;
    constexpr int n_waves = kPersistThreads / 64;
    double* B = reinterpret_cast<double*>(fine);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nx = s.n_x, nb = s.n_bins, nf = s.n_fine;
    const long long np = s.n_pairs, lanes = static_cast<long long>(gridDim.x) * kPersistThreads;
    const long long gid = static_cast<long long>(blockIdx.x) * kPersistThreads + tid;
    const double n_particles = 2.0 * static_cast<double>(np);
    const int bn = shift_bits(2 * np);
    // bin sums published: count and sum of v always; the offsets' sums where the moments are read (local-linear:
    // d, d^2, d v; the std rule: d, d^2)
    const int n_sums = s.local_linear ? 5 * nb : (s.bw_rule == 1 ? 4 * nb : 2 * nb);
    const R rho = static_cast<R>(s.rho), rc = static_cast<R>(s.rc), kappa = static_cast<R>(s.kappa);
    const R theta = static_cast<R>(s.theta), xi = static_cast<R>(s.xi);

    R xs[PPL][2], vs[PPL][2];
#pragma unroll
    for (int k = 0; k < PPL; ++k)
#pragma unroll
        for (int leg = 0; leg < 2; ++leg) xs[k][leg] = R(0), vs[k][leg] = static_cast<R>(s.v0);

    double lo = 0.0, hi = 0.0;  // the range of the particles at the start of the step (its bins' range)
    R zn[PPL][2];
    auto draw = [&](int step) {
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const long long p = gid + k * lanes;
            zn[k][0] = zn[k][1] = R(0);
            if (p >= np) continue;
            const U4 r4 = philox10(U4{static_cast<std::uint32_t>(p), static_cast<std::uint32_t>(static_cast<std::uint64_t>(p) >> 32),
                                      static_cast<std::uint32_t>(step + 1), 0u},
                                   s.k0, s.k1);
            gauss2(r4.x, r4.y, zn[k][0], zn[k][1]);
        }
    };
    draw(0);
    int shift_d = 0, shift_v = 0;
    unsigned wait_no = 0;
    for (int i = 0; i < s.n_steps; ++i) {
        const int par = i & 1, nxt = par ^ 1;
        {  // the next step's totals, last read at step i - 1 (before the waits since)
            unsigned* f = ps.fine_tot + static_cast<long long>(nxt) * nf;
            long long* t = ps.sum_tot + static_cast<long long>(nxt) * 5 * nb;
            for (long long e = gid; e < nf; e += lanes) __hip_atomic_store(f + e, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (long long e = gid; e < 5 * nb; e += lanes)
                __hip_atomic_store(t + e, 0ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // ---- A: the grid and the leverage row of step i (leverage.cpp:42-97). Latency-bound: the totals are read
        // in one batch of loads, the table entries of every grid point are evaluated by different lanes, and the
        // regression of a grid point visits only the bins within 12 bandwidths of it (fit_point: what lies beyond
        // is below 1e-16 of the sums the estimate uses).
        const double h = s.dt[i];
        double ga, gb;
        if (i == 0) {
            for (int g = tid; g < nx; g += kPersistThreads) xev[g] = (-6.0 + 12.0 * g / (nx - 1)) * s.atm_vol * sqrt(fmax(h, 1e-4));
            ga = (-6.0) * s.atm_vol * sqrt(fmax(h, 1e-4));
            gb = (-6.0 + 12.0 * (nx - 1) / (nx - 1)) * s.atm_vol * sqrt(fmax(h, 1e-4));
        } else {
            const unsigned* ftot = ps.fine_tot + static_cast<long long>(par) * nf;
            const long long* stot = ps.sum_tot + static_cast<long long>(par) * 5 * nb;
            const int per = (nf + kPersistThreads - 1) / kPersistThreads;
            const int k0 = min(nf, tid * per), k1 = min(nf, k0 + per);
            unsigned fc[kPer];  // counts below 2^32 (the persistent launch holds at most 2^22 particles)
            long long sl[(5 * kMaxBins + kPersistThreads - 1) / kPersistThreads];
#pragma unroll
            for (int m = 0; m < kPer; ++m) fc[m] = k0 + m < k1 ? load_agent(ftot + k0 + m) : 0u;
#pragma unroll
            for (int m = 0; m < (5 * kMaxBins + kPersistThreads - 1) / kPersistThreads; ++m) {
                const int e = tid + m * kPersistThreads;
                sl[m] = e < n_sums ? __hip_atomic_load(stot + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            }
#pragma unroll
            for (int m = 0; m < (5 * kMaxBins + kPersistThreads - 1) / kPersistThreads; ++m) {
                const int e = tid + m * kPersistThreads;
                if (e < 5 * nb) sums[e] = sl[m];
            }
            // the cumulative counts (integers: the same values in any grouping)
            unsigned run = 0;
#pragma unroll
            for (int m = 0; m < kPer; ++m) run += fc[m];
            unsigned incl = run;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const unsigned y = __shfl_up(incl, off, 64);
                if (lane >= off) incl += y;
            }
            if (lane == 63) scan[wave] = static_cast<double>(incl);
            if (tid < 2) first_hit[tid] = INT_MAX;
            __syncthreads();
            double before = 0.0, total = 0.0;
            for (int w = 0; w < n_waves; ++w) {
                if (w < wave) before += scan[w];
                total += scan[w];
            }
            const double start = before + static_cast<double>(incl - run);
            const double target[2] = {s.q_lo * total, s.q_hi * total};
            for (int q = 0; q < 2; ++q) {  // each thread's first bin reaching the target, the wave's least, one LDS atomic
                double cum = start;
                int cand = INT_MAX;
#pragma unroll
                for (int m = 0; m < kPer; ++m) {
                    if (cand != INT_MAX || k0 + m >= k1) continue;
                    const double c = static_cast<double>(fc[m]);
                    if (cum + c >= target[q] && c > 0)
                        cand = k0 + m;
                    else
                        cum += c;
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) cand = min(cand, __shfl_xor(cand, off, 64));
                if (lane == 0 && cand != INT_MAX) atomicMin(&first_hit[q], cand);
            }
            // the bins' moments from the fixed-point totals (as prep_kernel), into B
            const double we = (hi + 1e-12 - lo) / nb;
            const double inv_d = ldexp(1.0, -shift_d), inv_v = ldexp(1.0, -shift_v);
            const bool moments = s.local_linear || s.bw_rule == 1;
            double sx = 0.0, sxx = 0.0;
            for (int k = tid; k < nb; k += kPersistThreads) {
                const double centre = lo + (k + 0.5) * we, cnt = static_cast<double>(sums[k]);
                const double sv = static_cast<double>(sums[nb + k]) * inv_v;
                B[k] = centre, B[nb + k] = cnt, B[2 * nb + k] = sv;
                if (moments) {
                    const double sd = static_cast<double>(sums[2 * nb + k]) * inv_d;
                    const double sdd = static_cast<double>(sums[3 * nb + k]) * inv_d;
                    const double sdv = static_cast<double>(sums[4 * nb + k]) * inv_v;
                    const double bx = cnt * centre + we * sd;
                    const double bxx = cnt * centre * centre + 2.0 * centre * we * sd + we * we * sdd;
                    B[3 * nb + k] = bx, B[4 * nb + k] = bxx, B[5 * nb + k] = centre * sv + we * sdv;
                    sx += bx;
                    sxx += bxx;
                }
            }
            if (s.bw_rule == 1) {
                sx = wave_sum(sx);
                sxx = wave_sum(sxx);
                if (lane == 0) scan[2 * n_waves + wave] = sx, scan[3 * n_waves + wave] = sxx;
            }
            __syncthreads();  // first_hit, B, the std's partial sums
            const double width = (hi - lo) / static_cast<double>(nf);
            for (int q = 0; q < 2; ++q) {
                const int k = first_hit[q];
                if (k == INT_MAX) {
                    if (tid == 0) qval[q] = hi;
                } else if (k >= k0 && k < k1) {
                    double cum = start, c = 0.0;
#pragma unroll
                    for (int m = 0; m < kPer; ++m) {
                        if (k0 + m < k) cum += static_cast<double>(fc[m]);
                        if (k0 + m == k) c = static_cast<double>(fc[m]);
                    }
                    qval[q] = lo + (static_cast<double>(k) + (target[q] - cum) / c) * width;
                }
            }
            if (wave == 0) {  // E[v]: the fallback of an empty neighbourhood (lanes over the bins, as the fit)
                double vsum = 0.0;
                for (int k = lane; k < nb; k += 64) vsum += B[2 * nb + k];
                vsum = wave_sum(vsum);
                double bw = s.bandwidth_scale * s.atm_vol * sqrt(fmax(s.t0[i], 0.15)) * s.n_pow;
                if (s.bw_rule == 1) {
                    double tx = 0.0, txx = 0.0;
                    for (int w = 0; w < n_waves; ++w) tx += scan[2 * n_waves + w], txx += scan[3 * n_waves + w];
                    const double mean = tx / n_particles, var = fmax(txx / n_particles - mean * mean, 0.0);
                    bw = s.bw_coef * sqrt(var) * s.n_pow;
                }
                if (lane == 0) sh_bw = bw, sh_fallback = vsum / n_particles;
            }
            __syncthreads();
            const double qa = qval[0], qb = qval[1];
            for (int g = tid; g < nx; g += kPersistThreads) xev[g] = qa + (qb - qa) * g / (nx - 1);
            ga = qa, gb = qa + (qb - qa) * (nx - 1) / (nx - 1);
        }
        __syncthreads();  // xev
        // at once, on different waves: the table entries around every grid point (two lanes a point, the first half
        // of the workgroup) and the regression at every grid point (as fit_kernel: a lane a point, the points spread
        // over the second half's waves); then their quotient
        constexpr int half = kPersistThreads / 2;
        if (tid < half) {
            for (int t = tid; t < 2 * nx; t += half) {
                const int g = t >> 1, e = t & 1;
                lvv[e][g] = local_var_part(s, i, xev[g], e);
            }
        } else {
            const int t = tid - half, g = (t & 63) * (half / 64) + (t >> 6);
            for (int gg = g; gg < nx; gg += half) fitv[gg] = i > 0 ? fit_point(s, B, xev[gg], sh_bw, sh_fallback, lo, hi) : s.v0;
        }
        __syncthreads();
        for (int g = tid; g < nx; g += kPersistThreads) {
            const double x0 = xev[g];
            const double c = fitv[g];
            const double loc = local_var_join(s, i, x0, lvv[0][g], lvv[1][g]);
            const double L = fmin(fmax(sqrt(loc / fmax(c, 1e-8)), s.l_min), s.l_max);
            lev[g] = static_cast<R>(L);
            if (blockIdx.x == 0) {
                s.x_grids[static_cast<long long>(i) * nx + g] = x0;
                s.values[static_cast<long long>(i) * nx + g] = L;
            }
        }
        __syncthreads();
        // the LDS bins of phase C cleared now (fine doubles as the bins' moments, read for the last time above); the
        // waits below order these stores before the binning
        if (i + 1 < s.n_steps) {
            for (int k = tid; k < nf; k += kPersistThreads) fine[k] = 0u;
            for (int k = tid; k < 5 * nb; k += kPersistThreads) sums[k] = 0;
        }
        // ---- B: one Euler step of both legs of the workgroup's pairs (as step_kernel)
        const R a = static_cast<R>(ga), b = static_cast<R>(gb);
        const R hr = static_cast<R>(h);
        const R span = b - a, top = static_cast<R>(nx - 1);
        double mn = INFINITY, mx = -INFINITY, mv = 0.0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const long long p = gid + k * lanes;
            if (p >= np) continue;
            const R z0 = zn[k][0], z1 = zn[k][1];
#pragma unroll
            for (int leg = 0; leg < 2; ++leg) {
                const R sg = leg ? R(-1) : R(1), zs = sg * z0, zv = sg * z1;
                R x = xs[k][leg], v = vs[k][leg];
                const R u = (x - a) / span * top;
                R L;
                if (!(u > R(0))) {
                    L = lev[0];
                } else if (u >= top) {
                    L = lev[nx - 1];
                } else {
                    const int q = static_cast<int>(u);
                    const R w = u - static_cast<R>(q);
                    L = lev[q] * (R(1) - w) + lev[q + 1] * w;
                }
                const R vp = r_max(v, R(0)), var = L * L * vp;
                const R zx = rho * zv + rc * zs;
                v += kappa * (theta - vp) * hr + xi * r_sqrt(vp * hr) * zv;
                x += R(-0.5) * var * hr + r_sqrt(var * hr) * zx;
                xs[k][leg] = x;
                vs[k][leg] = v;
                mn = fmin(mn, static_cast<double>(x));
                mx = fmax(mx, static_cast<double>(x));
                mv = fmax(mv, fabs(static_cast<double>(v)));
            }
        }
        if (i + 1 == s.n_steps) break;
        // the workgroup's range, then every workgroup's
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, m, 64));
            mx = fmax(mx, __shfl_xor(mx, m, 64));
            mv = fmax(mv, __shfl_xor(mv, m, 64));
        }
        if (lane == 0) red[0][wave] = mn, red[1][wave] = mx, red[2][wave] = mv;
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < n_waves; ++w) {
                mn = fmin(mn, red[0][w]);
                mx = fmax(mx, red[1][w]);
                mv = fmax(mv, red[2][w]);
            }
            double* r = ps.ranges + (static_cast<long long>(nxt) * gridDim.x + blockIdx.x) * 4;
            __hip_atomic_store(r, mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(r + 1, mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(r + 2, mv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!grid_wait(ps, wait_no++)) return;
        mn = INFINITY, mx = -INFINITY, mv = 0.0;
        for (int w = tid; w < static_cast<int>(gridDim.x); w += kPersistThreads) {
            const double* r = ps.ranges + (static_cast<long long>(nxt) * gridDim.x + w) * 4;
            mn = fmin(mn, __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            mx = fmax(mx, __hip_atomic_load(r + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            mv = fmax(mv, __hip_atomic_load(r + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, m, 64));
            mx = fmax(mx, __shfl_xor(mx, m, 64));
            mv = fmax(mv, __shfl_xor(mv, m, 64));
        }
        if (lane == 0) red2[0][wave] = mn, red2[1][wave] = mx, red2[2][wave] = mv;
        __syncthreads();
        for (int w = 0; w < n_waves; ++w) {
            mn = fmin(mn, red2[0][w]);
            mx = fmax(mx, red2[1][w]);
            mv = fmax(mv, red2[2][w]);
        }
        lo = mn, hi = mx;
        // ---- C: bin the particles of step i + 1 (as bin_kernel) and add the workgroup's bins to the totals
        int ev = 0;
        if (mv > 0) frexp(mv, &ev);
        shift_d = 61 - bn;
        shift_v = min(52, 61 - bn - max(ev, -8));
        const double sd = ldexp(1.0, shift_d), sv = ldexp(1.0, shift_v);
        const R rlo = static_cast<R>(lo), rhi = static_cast<R>(hi);
        const R wf = (rhi - rlo) / static_cast<R>(nf);
        const R we = (rhi + static_cast<R>(1e-12) - rlo) / static_cast<R>(nb);
        const double wed = (hi + 1e-12 - lo) / nb;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (gid + k * lanes >= np) continue;
#pragma unroll
            for (int leg = 0; leg < 2; ++leg) {
                const R xr = xs[k][leg];
                const R vr = s.positive_part ? r_max(vs[k][leg], R(0)) : vs[k][leg];
                int f = wf > R(0) ? static_cast<int>((xr - rlo) / wf) : 0;
                f = min(max(f, 0), nf - 1);
                int q = static_cast<int>((xr - rlo) / we);
                q = min(max(q, 0), nb - 1);
                atomicAdd(&fine[f], 1u);
                const double xd = static_cast<double>(xr), vd = static_cast<double>(vr);
                atomicAdd(reinterpret_cast<unsigned long long*>(&sums[q]), 1ull);
                atomicAdd(reinterpret_cast<unsigned long long*>(&sums[nb + q]),
                          static_cast<unsigned long long>(__double2ll_rn(vd * sv)));
                if (n_sums > 2 * nb) {  // the offsets' sums only where the moments are read
                    const double centre = lo + (q + 0.5) * wed;
                    const double d = (xd - centre) / wed;
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sums[2 * nb + q]),
                              static_cast<unsigned long long>(__double2ll_rn(d * sd)));
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sums[3 * nb + q]),
                              static_cast<unsigned long long>(__double2ll_rn(d * d * sd)));
                    if (n_sums > 4 * nb)
                        atomicAdd(reinterpret_cast<unsigned long long*>(&sums[4 * nb + q]),
                                  static_cast<unsigned long long>(__double2ll_rn(d * vd * sv)));
                }
            }
        }
        __syncthreads();
        {
            unsigned* f = ps.fine_tot + static_cast<long long>(nxt) * nf;
            unsigned long long* t = reinterpret_cast<unsigned long long*>(ps.sum_tot + static_cast<long long>(nxt) * 5 * nb);
            if ((nf & 1) == 0) {  // two bins a 64-bit atomic (a bin's total stays below 2^32: no carry crosses)
                unsigned long long* f2 = reinterpret_cast<unsigned long long*>(f);
                for (int q = tid; q < nf / 2; q += kPersistThreads) {
                    const unsigned long long v = fine[2 * q] | (static_cast<unsigned long long>(fine[2 * q + 1]) << 32);
                    if (v) __hip_atomic_fetch_add(f2 + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                for (int k = tid; k < nf; k += kPersistThreads)
                    if (fine[k]) __hip_atomic_fetch_add(f + k, fine[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            for (int k = tid; k < n_sums; k += kPersistThreads)
                if (sums[k])
                    __hip_atomic_fetch_add(t + k, static_cast<unsigned long long>(sums[k]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        }
        draw(i + 1);
        if (!grid_wait(ps, wait_no++)) return;
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

// Calls f with the persistent kernel of (R, pairs per lane).
template <typename R, typename F>
bool with_persistent_kernel(int ppl, F f) {
    switch (ppl) {
        case 1: f(slv_persistent_kernel<R, 1>); return true;
        case 2: f(slv_persistent_kernel<R, 2>); return true;
        case 4: f(slv_persistent_kernel<R, 4>); return true;
        case 8:
            if constexpr (sizeof(R) == 4) {
                f(slv_persistent_kernel<R, 8>);
                return true;
            }
            return false;
        default: return false;
    }
}

// The persistent launch's shape: workgroups of 512 lanes (the kernel holds up to 256 VGPRs a lane: its FP64
// constants stay live across the step loop) and pairs a lane, a power of two up to 8 (FP32) or 4 (FP64), beyond
// which the registers spill. The launch is latency-bound: a wait costs more with more workgroups (2.2 us at 64,
// 2.8 us at 256 on MI355X) and so does the integer atomics' publication of the bins, while fewer workgroups step
// and bin more pairs a lane; at 2^19 pairs the calibration measured 45.6 ms at 256 workgroups, 47.2 at 64 to 128.
// OPTLAB_SLV_WORKGROUPS sets the preferred count (0: the five-kernel form).
struct PersistShape {
    int groups = 0, ppl = 0;
};

template <typename R>
PersistShape persistent_shape(long long n_pairs, int n_x) {
    PersistShape sh;
    if (n_x > kPersistMaxX) return sh;
    int want = 256;
    if (const char* e = std::getenv("OPTLAB_SLV_WORKGROUPS")) want = std::atoi(e);
    if (want <= 0) return sh;  // 0: the five-kernel form
    const long long need = (n_pairs + kPersistThreads - 1) / kPersistThreads;  // workgroups at one pair a lane
    long long groups = std::min<long long>(want, need);
    long long ppl = 1;
    while (groups * kPersistThreads * ppl < n_pairs) ppl *= 2;
    const int max_ppl = sizeof(R) == 4 ? 8 : 4;
    if (ppl > max_ppl) {  // more workgroups instead
        ppl = max_ppl;
        groups = (n_pairs + kPersistThreads * ppl - 1) / (kPersistThreads * ppl);
    }
    int per_cu = 0;
    with_persistent_kernel<R>(static_cast<int>(ppl), [&](auto kernel) {
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kPersistThreads, 0), "occupancy");
    });
    if (per_cu < 1 || groups > static_cast<long long>(per_cu) * device_cu_count()) return sh;  // not co-resident
    sh.groups = static_cast<int>(groups), sh.ppl = static_cast<int>(ppl);
    return sh;
}

template <typename R>
void run_typed(optlab_leverage_request& rq, int bw_rule, double bw_coef, const optlab_lv_slices* lvs) {
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
    const size_t o_lv = carve(lvs ? 8 : rows * rq.n_lv * 8), o_xg = carve(rows * nx * 8), o_val = carve(rows * nx * 8);
    const size_t o_hi = carve(lvs ? rows * 8 : 8), o_rec = carve(lvs ? rows * 24 * 8 : 8);
    const size_t o_sh = carve(sizeof(Shared)), o_rp = carve(static_cast<size_t>(s.n_step_blocks) * 3 * 8);
    const size_t o_fp = carve(static_cast<size_t>(s.n_bin_blocks) * nf * 4);
    const size_t o_spart = carve(static_cast<size_t>(s.n_bin_blocks) * 5 * nb * 8);
    const size_t o_ft = carve(static_cast<size_t>(nf) * 8), o_st = carve(static_cast<size_t>(5 * nb) * 8);
    const size_t o_x = carve(static_cast<size_t>(n) * sizeof(R)), o_v = carve(static_cast<size_t>(n) * sizeof(R));
    const size_t o_lev = carve(static_cast<size_t>(nx) * sizeof(R));
    const size_t o_xe = carve(static_cast<size_t>(nx) * 8), o_bins = carve(static_cast<size_t>(6 * nb) * 8);
    const PersistShape shape = persistent_shape<R>(np, nx);
    const int G = std::max(shape.groups, 1);
    const size_t o_pr = carve(static_cast<size_t>(2) * G * 4 * 8), o_pf = carve(static_cast<size_t>(2) * nf * 4);
    const size_t o_ps = carve(static_cast<size_t>(2) * 5 * nb * 8);
    const size_t o_sync = carve(static_cast<size_t>(kSyncGroups + 2) * kSyncStride * 4);
    unsigned char* base = static_cast<unsigned char*>(scratch(kSlotParticles, off));
    hip_check(hipMemcpy(base + o_dt, rq.dt, rows * 8, hipMemcpyHostToDevice), "copy dt");
    hip_check(hipMemcpy(base + o_t0, rq.t0, rows * 8, hipMemcpyHostToDevice), "copy t0");
    hip_check(hipMemcpy(base + o_lo, rq.lv_lo, rows * 8, hipMemcpyHostToDevice), "copy lv_lo");
    hip_check(hipMemcpy(base + o_inv, rq.lv_inv_dx, rows * 8, hipMemcpyHostToDevice), "copy lv_inv_dx");
    if (lvs) {
        hip_check(hipMemcpy(base + o_hi, lvs->hi, rows * 8, hipMemcpyHostToDevice), "copy lv hi");
        hip_check(hipMemcpy(base + o_rec, lvs->records, rows * 24 * 8, hipMemcpyHostToDevice), "copy lv records");
        s.lv_hi = reinterpret_cast<const double*>(base + o_hi);
        s.lv_rec = reinterpret_cast<const double*>(base + o_rec);
        s.lv_floor = lvs->floor, s.lv_cap = lvs->cap;
    } else {
        hip_check(hipMemcpy(base + o_lv, rq.lv_rows, rows * rq.n_lv * 8, hipMemcpyHostToDevice), "copy lv_rows");
    }
    s.dt = reinterpret_cast<const double*>(base + o_dt), s.t0 = reinterpret_cast<const double*>(base + o_t0);
    s.lv_lo = reinterpret_cast<const double*>(base + o_lo), s.lv_inv_dx = reinterpret_cast<const double*>(base + o_inv);
    s.lv_rows = reinterpret_cast<const double*>(base + o_lv);
    s.x_grids = reinterpret_cast<double*>(base + o_xg), s.values = reinterpret_cast<double*>(base + o_val);
    s.shared = reinterpret_cast<Shared*>(base + o_sh), s.range_part = reinterpret_cast<double*>(base + o_rp);
    s.fine_part = reinterpret_cast<unsigned*>(base + o_fp), s.sum_part = reinterpret_cast<long long*>(base + o_spart);
    s.fine_tot = reinterpret_cast<unsigned long long*>(base + o_ft), s.sum_tot = reinterpret_cast<long long*>(base + o_st);
    s.x_eval = reinterpret_cast<double*>(base + o_xe), s.bins = reinterpret_cast<double*>(base + o_bins);
    Particles<R> P{reinterpret_cast<R*>(base + o_x), reinterpret_cast<R*>(base + o_v), reinterpret_cast<R*>(base + o_lev)};

    hipStream_t st = engine_stream();
    if (shape.groups > 0) {  // one persistent launch
        PersistState ps{};
        ps.ranges = reinterpret_cast<double*>(base + o_pr);
        ps.fine_tot = reinterpret_cast<unsigned*>(base + o_pf);
        ps.sum_tot = reinterpret_cast<long long*>(base + o_ps);
        ps.sync = reinterpret_cast<unsigned*>(base + o_sync);
        int dev = 0, rate_khz = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
        ps.spin_limit = 2000ull * static_cast<unsigned long long>(std::max(rate_khz, 1));  // 2 s a wait
        const size_t sync_bytes = static_cast<size_t>(kSyncGroups + 2) * kSyncStride * 4;
        hip_check(hipMemsetAsync(ps.sync, 0, sync_bytes, st), "clear waits");
        hip_check(hipMemsetAsync(ps.fine_tot, 0, static_cast<size_t>(2) * nf * 4, st), "clear totals");
        hip_check(hipMemsetAsync(ps.sum_tot, 0, static_cast<size_t>(2) * 5 * nb * 8, st), "clear totals");
        void* args[] = {&s, &ps};
        with_persistent_kernel<R>(shape.ppl, [&](auto kernel) {
            hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kernel), dim3(shape.groups),
                                                 dim3(kPersistThreads), args, 0, st),
                      "persistent SLV launch");
        });
        unsigned flags[kSyncGroups + 2];
        hip_check(hipStreamSynchronize(st), "persistent SLV kernel");
        for (int k = 0; k < kSyncGroups + 2; ++k)
            hip_check(hipMemcpy(&flags[k], ps.sync + k * kSyncStride, 4, hipMemcpyDeviceToHost), "copy waits");
        if (flags[kSyncGroups + 1])
            throw Error(OPTLAB_EDEVICE, "persistent SLV launch abandoned: a grid-wide wait passed its time bound");
        hip_check(hipMemcpy(rq.x_grids, s.x_grids, rows * nx * 8, hipMemcpyDeviceToHost), "copy x_grids");
        hip_check(hipMemcpy(rq.values, s.values, rows * nx * 8, hipMemcpyDeviceToHost), "copy values");
        rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        return;
    }
    // The five-kernel form (more particles than 16 pairs a lane on every CU, or more than 256 grid points): the
    // kernels of every step, launches that depend each on the last, are captured in one HIP graph and launched at
    // once: the gaps between dependent launches cost more than the kernels at 2^20 particles.
    const unsigned init_grid = static_cast<unsigned>(std::min<long long>((n + kBlock - 1) / kBlock, 4LL * cus));
    hipLaunchKernelGGL(init_kernel<R>, dim3(init_grid), dim3(kBlock), 0, st, P, n, static_cast<R>(rq.v0));
    hip_check(hipGetLastError(), "init kernel launch");
    const unsigned red_grid = static_cast<unsigned>((nf + 5 * nb + 63) / 64);  // 64 slots a workgroup
    const unsigned fit_grid = static_cast<unsigned>((nx + kFitThreads - 1) / kFitThreads);
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

void check_leverage_request(const optlab_leverage_request& rq, bool need_rows) {  // requests.cpp:282-288
    if (!rq.dt || !rq.t0 || !rq.lv_lo || !rq.lv_inv_dx || (need_rows && !rq.lv_rows) || !rq.x_grids || !rq.values)
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

void run_leverage(optlab_leverage_request& rq, const optlab_leverage_options* opt, int prec,
                  const optlab_lv_slices* lvs) {
    if (opt && (opt->bandwidth_rule < 0 || opt->bandwidth_rule > 1)) invalid("bandwidth_rule must be 0 or 1");
    if (opt && opt->bandwidth_rule == 1 && !(opt->bandwidth_coef > 0)) invalid("bandwidth_coef must be positive");
    const int rule = opt ? opt->bandwidth_rule : 0;
    const double coef = opt ? opt->bandwidth_coef : 0.0;
    if (lvs) {
        if (lvs->n_rows != rq.n_steps) invalid("the slices need one row per step");
        if (!lvs->hi || !lvs->records) invalid("null slices");
        if (!(lvs->floor > 0 && lvs->cap >= lvs->floor)) invalid("the local variance clip needs 0 < floor <= cap");
    }
    if (prec == 0)
        run_typed<double>(rq, rule, coef, lvs);
    else
        run_typed<float>(rq, rule, coef, lvs);
}

}  // namespace optlab_mi
