// Local-variance tables of implied-vol surfaces bumped by the tents of the risk grid (and by a parallel shift and a
// skew rotation), built on the device in FP64, and the exact derivative of every table entry with respect to the
// grid's nodes contracted with a table adjoint.
// Semantics: packages/risk/src/optlab/risk/grid.py:32-166 (strike tents: Gaussian-smoothed hats in log forward
// moneyness; tenor tents: a smootherstep partition of unity), packages/risk/src/optlab/risk/surfaces.py:30-165
// (sigma' = sigma + a, a = parallel + skew + sum_ij H_ij phi_i(k, T) psi_j(T); w = T sigma'^2 with exact w_k, w_kk,
// w_T; Dupire on the base surface's table grid) and packages/structured/src/optlab/structured/surfaces.py:10-50
// (the skew rotation).
//
// Layout: the base surface arrives as rows (one per simulation step) of w, w_k, w_kk, w_T on the engine's table
// grid; the table kernel runs one thread per entry; the contraction runs one workgroup per row, its threads split
// over (active node, x phase), each accumulating every batch of the adjoint.
#include <algorithm>
#include <cmath>
#include <vector>

#include "device.hpp"
#include "engine.hpp"
#include "surface.hpp"

namespace optlab_mi {

namespace {

constexpr double kSkewWidth = 0.3, kSkewFloor = 1.0 / 12;
constexpr int kMaxBatches = 16;

struct Width {  // l(k, T) = c(k) sqrt(max(T, floor)) and its derivatives (grid.py:100-112)
    double l, lk, lkk, lT;
};

__device__ Width smoothing_width(const TentGrid& g, double k, double T) {
    const double root = sqrt(fmax(T, g.t_floor));
    const double span = g.wing - g.atm, a2 = g.scale * g.scale;
    const double e = exp(-k * k / a2);
    const double c = g.wing - span * e;
    const double ck = span * e * 2.0 * k / a2;
    const double ckk = span * e * (2.0 / a2 - 4.0 * k * k / (a2 * a2));
    return Width{c * root, ck * root, ckk * root, T < g.t_floor ? 0.0 : 0.5 * c / root};
}

struct Tent {
    double f, fk, fkk, fT;
};

// Strike tent i at k: const + sum_m c_m R(k - a_m), R = u Phi(z) + l n(z), z = u / l, R_k = Phi + n l_k,
// R_kk = n ((1 - z l_k)^2 / l + l_kk), R_T = n l_T (grid.py:115-140); the hat's constants read into registers once.
struct Hat {
    double c0;
    int n;
    double a[3], c[3];
};

__device__ __forceinline__ Hat hat_of(const TentGrid& g, int i) {
    Hat h{g.hat_const[i], g.hat_n[i], {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int m = 0; m < 3; ++m)
        if (m < h.n) h.a[m] = g.hat_a[i][m], h.c[m] = g.hat_c[i][m];
    return h;
}

__device__ __forceinline__ Tent hat_tent(const Hat& h, double k, const Width& w) {
    Tent t{h.c0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        if (m >= h.n) break;
        const double u = k - h.a[m], c = h.c[m];
        const double z = u / w.l;
        const double pdf = exp(-0.5 * z * z) * 0.3989422804014327, cdf = normcdf(z);
        const double q = 1.0 - z * w.lk;
        t.f += c * (u * cdf + w.l * pdf);
        t.fk += c * (cdf + pdf * w.lk);
        t.fkk += c * (pdf * (q * q / w.l + w.lkk));
        t.fT += c * (pdf * w.lT);
    }
    return t;
}

// The (at most two) tenor tents that are not zero at T, j0 and j0 + 1, and their T-derivatives (grid.py:143-162).
struct Tenors {
    int j0;
    double psi[2], dpsi[2];
};

__device__ Tenors tenor_tents(const TentGrid& g, double T) {
    Tenors r{0, {0.0, 0.0}, {0.0, 0.0}};
    const int n = g.n_tenors;
    if (T <= g.tenors[0]) {
        r.psi[0] = 1.0;
        return r;
    }
    if (T >= g.tenors[n - 1]) {
        r.j0 = n - 1, r.psi[0] = 1.0;
        return r;
    }
    int i = 0;
    while (i + 1 < n && g.tenors[i + 1] <= T) ++i;
    r.j0 = i;
    if (T == g.tenors[i]) {
        r.psi[0] = 1.0;
        return r;
    }
    const double a = g.tenors[i], b = g.tenors[i + 1], u = (T - a) / (b - a);
    const double s = u * u * u * (10.0 - 15.0 * u + 6.0 * u * u), ds = 30.0 * u * u * (1.0 - u) * (1.0 - u);
    r.psi[0] = 1.0 - s, r.psi[1] = s;
    r.dpsi[0] = -ds / (b - a), r.dpsi[1] = ds / (b - a);
    return r;
}

struct Point {
    double k, T, w, wk, wkk, wT;
};

__device__ Point base_point(const SurfaceRows& s, long long row, int x) {
    const long long at = row * s.n_x + x;
    return Point{s.lo[row] + x / s.inv_dx[row], s.T[row], s.w[at], s.wk[at], s.wkk[at], s.wT[at]};
}

struct Sigma {  // sigma and its derivatives from w (surfaces.py:37-43)
    double s, sk, skk, sT;
};

__device__ Sigma sigma_of(const Point& p) {
    const double s = sqrt(fmax(p.w, 1e-300) / p.T);
    const double sk = p.wk / (2.0 * p.T * s);
    const double skk = (p.wkk - 2.0 * p.T * sk * sk) / (2.0 * p.T * s);
    const double sT = (p.wT - s * s) / (2.0 * p.T * s);
    return Sigma{s, sk, skk, sT};
}

// Durrleman's g (core/svi.py:31-35); the local variance is clip(w_T / max(g, 1e-6), floor, cap).
__device__ double g_function(double k, double w, double wk, double wkk) {
    const double A = 1.0 - k * wk / (2.0 * w);
    return A * A - wk * wk / 4.0 * (1.0 / w + 0.25) + wkk / 2.0;
}

// The skew rotation and its derivatives in k, k twice and T (structured/surfaces.py:29-50).
__device__ Tent skew_bump(double skew, double k, double T) {
    if (skew == 0.0) return Tent{0.0, 0.0, 0.0, 0.0};
    const double norm = 1.0 / (tanh(-log(0.9) / kSkewWidth) + tanh(log(1.1) / kSkewWidth));
    const double Tf = fmax(T, kSkewFloor);
    const double scale = sqrt(1.0 / Tf), dscale = T < kSkewFloor ? 0.0 : -0.5 * scale / Tf;
    const double a = -skew * norm, th = tanh(k / kSkewWidth), sech2 = 1.0 - th * th;
    return Tent{a * scale * th, a * scale * sech2 / kSkewWidth,
                -2.0 * a * scale * th * sech2 / (kSkewWidth * kSkewWidth), a * dscale * th};
}

template <typename R>
__global__ void __launch_bounds__(256) bumped_table_kernel(SurfaceRows s, TentGrid g, const double* H, SurfaceMove mv,
                                                           R* out, double floor, double cap) {
    const long long n = s.n_rows * static_cast<long long>(s.n_x);
    for (long long at = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; at < n;
         at += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long row = at / s.n_x;
        const int x = static_cast<int>(at - row * s.n_x);
        const Point p = base_point(s, row, x);
        const Sigma q = sigma_of(p);
        const Tent sk = skew_bump(mv.skew, p.k, p.T);
        double a = mv.parallel + sk.f, ak = sk.fk, akk = sk.fkk, aT = sk.fT;
        if (H) {
            const Width wd = smoothing_width(g, p.k, p.T);
            const Tenors tn = tenor_tents(g, p.T);
            for (int i = 0; i < g.n_strikes; ++i) {
                double h = 0.0, hT = 0.0;
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = tn.j0 + jj;
                    if (j >= g.n_tenors) break;
                    h += H[i * g.n_tenors + j] * tn.psi[jj];
                    hT += H[i * g.n_tenors + j] * tn.dpsi[jj];
                }
                if (h == 0.0 && hT == 0.0) continue;
                const Tent t = hat_tent(hat_of(g, i), p.k, wd);
                a += t.f * h, ak += t.fk * h, akk += t.fkk * h, aT += t.f * hT + t.fT * h;
            }
        }
        const double s2 = q.s + a, sk2 = q.sk + ak, skk2 = q.skk + akk, sT2 = q.sT + aT;
        const double w2 = p.T * s2 * s2, wk2 = 2.0 * p.T * s2 * sk2;
        const double wkk2 = 2.0 * p.T * (sk2 * sk2 + s2 * skk2), wT2 = s2 * s2 + 2.0 * p.T * s2 * sT2;
        const double gg = g_function(p.k, w2, wk2, wkk2);
        out[at] = static_cast<R>(fmin(fmax(wT2 / fmax(gg, 1e-6), floor), cap));
    }
}

// One workgroup per (row, surface), one wave per strike tent; lane l takes the x with x % 64 == l. At each x the
// wave evaluates its strike's tent once and the entry's derivative for the two tenor tents that are not zero at the
// row's maturity, then multiplies it into every batch of the adjoint (the batches of one entry are contiguous).
// The wave sums its lanes with shuffles in a fixed order.
// D[(row * n_x + x) * n_active + 2 i + jj] = dLV[row][x] / dH at H = 0 for strike tent i and tenor tent j0(row) + jj
// (the two tenor tents that are not zero at the row's maturity); zero where the entry is clipped. Computed once per
// surface and kept with it: the base surface does not move between adjoint calls.
__global__ void __launch_bounds__(256) node_derivative_kernel(SurfaceRows s, TentGrid g, double* D, double floor,
                                                              double cap) {
    const long long n = s.n_rows * static_cast<long long>(s.n_x);
    const int n_active = 2 * g.n_strikes;
    for (long long at = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; at < n;
         at += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long row = at / s.n_x;
        const int x = static_cast<int>(at - row * s.n_x);
        double* out = D + at * n_active;
        const Point p = base_point(s, row, x);
        const double gg = g_function(p.k, p.w, p.wk, p.wkk);
        const double lv_raw = p.wT / fmax(gg, 1e-6);
        if (lv_raw <= floor || lv_raw >= cap) {  // clipped: the entry does not move
            for (int na = 0; na < n_active; ++na) out[na] = 0.0;
            continue;
        }
        const Sigma q = sigma_of(p);
        const Width wd = smoothing_width(g, p.k, p.T);
        const Tenors tn = tenor_tents(g, p.T);
        const double A = 1.0 - p.k * p.wk / (2.0 * p.w);
        const double g_w = A * p.k * p.wk / (p.w * p.w) + p.wk * p.wk / (4.0 * p.w * p.w);
        const double g_wk = -A * p.k / p.w - 0.5 * p.wk * (1.0 / p.w + 0.25);
        for (int i = 0; i < g.n_strikes; ++i) {
            const Tent t = hat_tent(hat_of(g, i), p.k, wd);
            for (int jj = 0; jj < 2; ++jj) {
                const double psi = tn.psi[jj], dpsi = tn.dpsi[jj];
                const double da = t.f * psi, dak = t.fk * psi, dakk = t.fkk * psi, daT = t.fT * psi + t.f * dpsi;
                const double dw = 2.0 * p.T * q.s * da;
                const double dwk = 2.0 * p.T * (q.sk * da + q.s * dak);
                const double dwkk = 2.0 * p.T * (2.0 * q.sk * dak + q.skk * da + q.s * dakk);
                const double dwT = 2.0 * q.s * da + 2.0 * p.T * (q.sT * da + q.s * daT);
                out[2 * i + jj] =
                    gg > 1e-6 ? (dwT - lv_raw * (g_w * dw + g_wk * dwk + 0.5 * dwkk)) / gg : dwT / 1e-6;
            }
        }
    }
}

// One workgroup per (row, surface): for that row, the (batch x node) products sum over x of the batch adjoint of
// each entry times its node derivative -- a 16 x n_x by n_x x 20 matrix product -- on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: 16 batches x 16 nodes, 4 entries a step; two node tiles cover 20 nodes). Each of the
// workgroup's four waves takes a quarter of the row's entries, reads its adjoints and derivatives straight from
// memory (every byte of the 520 MB once), and the four partial tiles are added in LDS in wave order: a fixed order,
// so the result does not depend on scheduling. (A thread per (batch, node) summing in order took 0.65 ms a call
// reading from memory and 0.385 ms staged through LDS.)
// Lane layout of the 16 x 16 x 4 FP64 product (checked on the device): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 k + l / 16][l % 16] in its k-th accumulator.
constexpr int kContractWaves = 4;
typedef double v4d __attribute__((ext_vector_type(4)));

// The adjoint of entry (a, row, x) of batch b is tbar[((a n_rows + row) n_x + x) x_stride + b batch_stride]: batches
// interleaved (x_stride = n_batches, batch_stride = 1) or one table after the other (x_stride = 1, batch_stride =
// the entries of a table).
constexpr int kContractUnroll = 8;

template <typename Adj>
__global__ void __launch_bounds__(64 * kContractWaves) node_contraction_kernel(SurfaceSet set, DerivSet dset,
                                                                             int n_active, const Adj* tbar,
                                                                             int n_batches, long long x_stride,
                                                                             long long batch_stride, double unit,
                                                                             double* out) {
    __shared__ double part[kContractWaves][2][4][64];
    const long long row = blockIdx.x;
    const int a = blockIdx.y;
    const int n_x = set.s[a].n_x;
    const long long n_rows = set.s[a].n_rows;
    const long long base = (static_cast<long long>(a) * n_rows + row) * n_x;  // the row's first entry
    const double* drow = dset.d[a] + row * n_x * n_active;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kk = lane >> 4;  // A row (batch) / B column (node), and the entry within the step
    const bool batch_ok = m < n_batches;
    const Adj* tb = tbar + (batch_ok ? m : 0) * batch_stride;
    const Adj* tb_row = tb + base * x_stride;
    const int m0 = min(m, n_active - 1), m1 = min(m + 16, n_active - 1);
    // this wave's entries: steps of 4 split evenly over the waves
    const int steps = (n_x + 3) / 4, per = (steps + kContractWaves - 1) / kContractWaves;
    const int x_lo = 4 * per * wave, x_hi = min(n_x, 4 * per * (wave + 1));
    v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
    // kContractUnroll steps of 4 entries a pass: every load of the pass is issued before its first product, so a wave
    // has that many loads in flight instead of one; the products still accumulate step after step, in entry order
    for (int x0 = x_lo; x0 < x_hi; x0 += 4 * kContractUnroll) {
        // the loads read clamped, always valid addresses with no branch around them (a branch would make the
        // compiler wait for each load inside it); the operands outside the row or the tiles are zeroed afterwards
        Adj raw[kContractUnroll];
        double d0[kContractUnroll], d1[kContractUnroll];
#pragma unroll
        for (int u = 0; u < kContractUnroll; ++u) {
            const int x = min(x0 + 4 * u + kk, x_hi - 1);
            raw[u] = tb_row[static_cast<long long>(x) * x_stride];
            const double* d = drow + static_cast<long long>(x) * n_active;
            d0[u] = d[m0];
            d1[u] = d[m1];
        }
#pragma unroll
        for (int u = 0; u < kContractUnroll; ++u) {
            const bool in = x0 + 4 * u + kk < x_hi;
            const double av = (in && batch_ok) ? static_cast<double>(raw[u]) : 0.0;
            const double b0 = (in && m < n_active) ? d0[u] : 0.0;
            const double b1 = (in && m + 16 < n_active) ? d1[u] : 0.0;
            c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b0, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1, c1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) part[wave][0][k][lane] = c0[k], part[wave][1][k][lane] = c1[k];
    __syncthreads();
    // thread t < 2 x 4 x 64 adds the waves' values of one (tile, k, lane) slot in wave order and writes it
    for (int t = threadIdx.x; t < 2 * 4 * 64; t += blockDim.x) {
        const int tile = t >> 8, k = (t >> 6) & 3, l = t & 63;
        double v = part[0][tile][k][l];
#pragma unroll
        for (int w = 1; w < kContractWaves; ++w) v += part[w][tile][k][l];
        const int b = 4 * k + (l >> 4), na = (l & 15) + 16 * tile;
        if (b < n_batches && na < n_active)
            out[((static_cast<long long>(a) * n_batches + b) * n_rows + row) * n_active + na] = v * unit;
    }
}

// vega[(a * n_batches + b) * n_nodes + node] = the sum over rows of each row's part of node (i, j): one wave per
// (node, batch, surface), lanes striding the rows, then a fixed-order shuffle sum.
__global__ void __launch_bounds__(64) node_sum_kernel(SurfaceSet set, TentGrid g, const double* per_row,
                                                      int n_batches, double* vega) {
    const int node = blockIdx.x, b = blockIdx.y, a = blockIdx.z, lane = threadIdx.x;
    const int n_nodes = g.n_strikes * g.n_tenors, n_active = 2 * g.n_strikes;
    const int i = node / g.n_tenors, j = node % g.n_tenors;
    const SurfaceRows s = set.s[a];
    const double* pr = per_row + (static_cast<long long>(a) * n_batches + b) * s.n_rows * n_active;
    double v = 0.0;
    for (long long row = lane; row < s.n_rows; row += 64) {
        const int jj = j - tenor_tents(g, s.T[row]).j0;
        if (jj == 0 || jj == 1) v += pr[row * n_active + 2 * i + jj];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_d(v, m);
    if (lane == 0) vega[(static_cast<long long>(a) * n_batches + b) * n_nodes + node] = v;
}

}  // namespace

TentGrid make_tent_grid() {
    // grid.py:25-26 and 73-77: strikes in forward moneyness, tenors in years, the smoothing constants
    static const double strikes[] = {0.6, 0.7, 0.8, 0.9, 0.95, 1.0, 1.05, 1.1, 1.2, 1.4};
    static const double tenors[] = {1.0 / 12, 0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 5.0};
    TentGrid g{};
    g.n_strikes = 10, g.n_tenors = 8;
    for (int i = 0; i < g.n_strikes; ++i) g.log_strikes[i] = std::log(strikes[i]);
    for (int j = 0; j < g.n_tenors; ++j) g.tenors[j] = tenors[j];
    g.atm = 0.04, g.wing = 0.12, g.scale = 0.15, g.t_floor = 1.0 / 52;
    // each hat as const + sum of ramps (x - a)_+ (grid.py:80-97)
    const double* x = g.log_strikes;
    const int n = g.n_strikes;
    for (int i = 0; i < n; ++i) {
        if (i == 0) {
            const double d = x[1] - x[0];
            g.hat_const[i] = 1.0, g.hat_n[i] = 2;
            g.hat_a[i][0] = x[0], g.hat_c[i][0] = -1 / d, g.hat_a[i][1] = x[1], g.hat_c[i][1] = 1 / d;
        } else if (i == n - 1) {
            const double d = x[n - 1] - x[n - 2];
            g.hat_const[i] = 0.0, g.hat_n[i] = 2;
            g.hat_a[i][0] = x[n - 2], g.hat_c[i][0] = 1 / d, g.hat_a[i][1] = x[n - 1], g.hat_c[i][1] = -1 / d;
        } else {
            const double dl = x[i] - x[i - 1], dr = x[i + 1] - x[i];
            g.hat_const[i] = 0.0, g.hat_n[i] = 3;
            g.hat_a[i][0] = x[i - 1], g.hat_c[i][0] = 1 / dl;
            g.hat_a[i][1] = x[i], g.hat_c[i][1] = -1 / dl - 1 / dr;
            g.hat_a[i][2] = x[i + 1], g.hat_c[i][2] = 1 / dr;
        }
    }
    return g;
}

void launch_bumped_table(const SurfaceRows& s, const TentGrid& g, const double* H, SurfaceMove move, int prec,
                         void* out, hipStream_t stream) {
    const long long n = s.n_rows * static_cast<long long>(s.n_x);
    const unsigned grid = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 65536));
    if (prec == 0)
        hipLaunchKernelGGL(bumped_table_kernel<double>, dim3(grid), dim3(256), 0, stream, s, g, H, move,
                           static_cast<double*>(out), 1e-4, 4.0);
    else
        hipLaunchKernelGGL(bumped_table_kernel<float>, dim3(grid), dim3(256), 0, stream, s, g, H, move,
                           static_cast<float*>(out), 1e-4, 4.0);
    hip_check(hipGetLastError(), "bumped table kernel launch");
}

template <typename A>
void contract(const SurfaceSet& set, const DerivSet& dset, const TentGrid& g, const A* tbar, int n_batches,
              double unit, double* out, hipStream_t stream, bool batch_major = false) {
    if (n_batches < 1 || n_batches > kMaxBatches) invalid("n_batches must be 1 to 16");
    if (set.n < 1 || set.n > kMaxSurfaces) invalid("1 to 5 surfaces");
    for (int a = 1; a < set.n; ++a)
        if (set.s[a].n_rows != set.s[0].n_rows || set.s[a].n_x != set.s[0].n_x) invalid("surfaces of one shape");
    const int n_active = 2 * g.n_strikes;
    const long long n_rows = set.s[0].n_rows;
    double* per_row = static_cast<double*>(
        scratch(kSlotNodes, static_cast<size_t>(set.n) * n_batches * n_rows * n_active * sizeof(double)));
    if (n_active > 32) invalid("at most 32 active nodes a row (two 16-node tiles)");
    const int threads = 64 * kContractWaves;
    const long long table_entries = static_cast<long long>(set.n) * n_rows * set.s[0].n_x;
    const long long x_stride = batch_major ? 1 : n_batches, batch_stride = batch_major ? table_entries : 1;
    hipLaunchKernelGGL(node_contraction_kernel<A>, dim3(static_cast<unsigned>(n_rows), set.n), dim3(threads), 0,
                       stream, set, dset, n_active, tbar, n_batches, x_stride, batch_stride, unit, per_row);
    hip_check(hipGetLastError(), "node contraction kernel launch");
    hipLaunchKernelGGL(node_sum_kernel, dim3(g.n_strikes * g.n_tenors, n_batches, set.n), dim3(64), 0, stream, set, g,
                       per_row, n_batches, out);
    hip_check(hipGetLastError(), "node sum kernel launch");
}

void launch_node_contraction(const SurfaceSet& set, const DerivSet& dset, const TentGrid& g, const long long* tbar,
                             int n_batches, double unit, double* out, hipStream_t stream, bool batch_major) {
    contract(set, dset, g, tbar, n_batches, unit, out, stream, batch_major);
}

// ---- the registry of uploaded surfaces ----------------------------------------------------------------------------

namespace {

struct Held {
    int device = -1;
    SurfaceRows rows{};
    void* mem = nullptr;
    void* table[2] = {nullptr, nullptr};  // the unmoved local-variance table in double and in float, built once
    double* deriv = nullptr;              // its node derivatives, built once
};
std::vector<Held> g_surfaces;

int current_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    return dev;
}

}  // namespace

int surface_upload(const optlab_surface_rows& r) {
    if (r.n_rows < 1 || r.n_x < 2) invalid("a surface needs n_rows >= 1 and n_x >= 2");
    if (!r.T || !r.lo || !r.inv_dx || !r.w || !r.w_k || !r.w_kk || !r.w_T) invalid("null surface array");
    for (int i = 0; i < r.n_rows; ++i)
        if (!(r.T[i] > 0) || !(r.inv_dx[i] > 0)) invalid("surface rows need T > 0 and inv_dx > 0");
    const size_t nr = static_cast<size_t>(r.n_rows), ne = nr * r.n_x;
    const size_t bytes = (3 * nr + 4 * ne) * sizeof(double);
    Held h;
    h.device = current_device();
    hip_check(hipMalloc(&h.mem, bytes), "hipMalloc surface");
    double* d = static_cast<double*>(h.mem);
    const double* src[7] = {r.T, r.lo, r.inv_dx, r.w, r.w_k, r.w_kk, r.w_T};
    const size_t len[7] = {nr, nr, nr, ne, ne, ne, ne};
    const double* at[7];
    for (int i = 0; i < 7; ++i) {
        hip_check(hipMemcpy(d, src[i], len[i] * sizeof(double), hipMemcpyHostToDevice), "copy surface");
        at[i] = d;
        d += len[i];
    }
    h.rows = SurfaceRows{r.n_rows, r.n_x, at[0], at[1], at[2], at[3], at[4], at[5], at[6]};
    for (size_t i = 0; i < g_surfaces.size(); ++i)
        if (!g_surfaces[i].mem) {
            g_surfaces[i] = h;
            return static_cast<int>(i) + 1;
        }
    g_surfaces.push_back(h);
    return static_cast<int>(g_surfaces.size());
}

static Held& held(int handle) {
    if (handle < 1 || handle > static_cast<int>(g_surfaces.size()) || !g_surfaces[handle - 1].mem)
        invalid("no such surface handle");
    Held& h = g_surfaces[handle - 1];
    if (h.device != current_device()) invalid("the surface was uploaded to another device");
    return h;
}

void surface_free(int handle) {
    Held& h = held(handle);
    hip_check(hipDeviceSynchronize(), "synchronize before free");
    for (void*& t : h.table)
        if (t) hip_check(hipFree(t), "hipFree table"), t = nullptr;
    if (h.deriv) hip_check(hipFree(h.deriv), "hipFree derivatives"), h.deriv = nullptr;
    hip_check(hipFree(h.mem), "hipFree surface");
    h = Held{};
}

const SurfaceRows& surface_rows(int handle) { return held(handle).rows; }

const void* surface_base_table(int handle, int prec, hipStream_t stream) {
    Held& h = held(handle);
    if (!h.table[prec]) {
        const size_t n = static_cast<size_t>(h.rows.n_rows) * h.rows.n_x;
        hip_check(hipMalloc(&h.table[prec], n * (prec == 0 ? sizeof(double) : sizeof(float))), "hipMalloc table");
        launch_bumped_table(h.rows, make_tent_grid(), nullptr, SurfaceMove{0.0, 0.0}, prec, h.table[prec], stream);
    }
    return h.table[prec];
}

const double* surface_node_derivatives(int handle, hipStream_t stream) {
    Held& h = held(handle);
    if (!h.deriv) {
        const TentGrid g = make_tent_grid();
        const long long n = static_cast<long long>(h.rows.n_rows) * h.rows.n_x;
        hip_check(hipMalloc(&h.deriv, static_cast<size_t>(n) * 2 * g.n_strikes * sizeof(double)), "hipMalloc deriv");
        const unsigned grid = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 65536));
        hipLaunchKernelGGL(node_derivative_kernel, dim3(grid), dim3(256), 0, stream, h.rows, g, h.deriv, 1e-4, 4.0);
        hip_check(hipGetLastError(), "node derivative kernel launch");
    }
    return h.deriv;
}

void surface_table(int handle, const double* nodes, double parallel, double skew, double* out_host) {
    const SurfaceRows& s = held(handle).rows;
    const size_t n = static_cast<size_t>(s.n_rows) * s.n_x;
    const TentGrid g = make_tent_grid();
    const size_t n_nodes = static_cast<size_t>(g.n_strikes) * g.n_tenors;
    double* d_nodes = nullptr;
    unsigned char* buf = static_cast<unsigned char*>(scratch(kSlotTables, n * sizeof(double) + n_nodes * sizeof(double)));
    double* d_out = reinterpret_cast<double*>(buf);
    if (nodes) {
        d_nodes = d_out + n;
        hip_check(hipMemcpy(d_nodes, nodes, n_nodes * sizeof(double), hipMemcpyHostToDevice), "copy nodes");
    }
    launch_bumped_table(s, g, d_nodes, SurfaceMove{parallel, skew}, 0, d_out, nullptr);
    hip_check(hipMemcpy(out_host, d_out, n * sizeof(double), hipMemcpyDeviceToHost), "copy table");
}

void surface_node_grad(int handle, const double* tbar_host, double* out_host) {
    const SurfaceRows& s = held(handle).rows;
    const size_t n = static_cast<size_t>(s.n_rows) * s.n_x;
    const TentGrid g = make_tent_grid();
    const size_t n_nodes = static_cast<size_t>(g.n_strikes) * g.n_tenors;
    double* buf = static_cast<double*>(scratch(kSlotTables, (n + n_nodes) * sizeof(double)));
    hip_check(hipMemcpy(buf, tbar_host, n * sizeof(double), hipMemcpyHostToDevice), "copy tbar");
    SurfaceSet set{};
    set.s[0] = s;
    set.n = 1;
    DerivSet dset{};
    dset.d[0] = surface_node_derivatives(handle, nullptr);
    contract(set, dset, g, buf, 1, 1.0, buf + n, nullptr);
    hip_check(hipMemcpy(out_host, buf + n, n_nodes * sizeof(double), hipMemcpyDeviceToHost), "copy node gradient");
}

}  // namespace optlab_mi
