// American options by Longstaff-Schwartz on the engine's paths.
// Algorithm: packages/core/src/optlab/core/products.py:105-169 (american_lsm, _basis) and :30-35 (_result): the
// exercise rule is estimated on one set of paths by a backward sweep of least-squares regressions of the discounted
// continuation value of the in-the-money paths on (1, s, s^2, s^3) [+ (v, v^2, s v)], s = S / K, and applied to an
// independent set.
//
// MI355X design: the path engine writes the spots (and the variance factor) at the exercise dates into HBM,
// structure of arrays (a date's paths contiguous). Each date of the backward sweep is one launch on one stream,
// with no host round trip: every workgroup first solves the later date from the per-workgroup rows of sums the
// previous launch wrote (the same fixed-order sums and FP64 Cholesky factorisation in each, so the same bits; the
// first workgroup publishes them), then a grid-stride pass brings every path's discounted value to the later date
// (its payoff at maturity, or that date's exercise decision) and forms the normal equations of the in-the-money
// paths of this date (X'X, X'y and the count: 15 or 36 FP64 sums) as per-workgroup rows, in alternate halves of
// one buffer. One workgroup solves the first date at the end. The rule is then applied to the second set of paths
// inside the path kernel itself, at each date as the paths reach it (ExerciseRule), which writes only its
// workgroups' sums of the pair averages: the second set's spots never reach HBM. Every sum has a fixed order, so a
// result does not depend on anything but the request.
//
// The kernel boundary publishes the rows to the next launch: a device-scope fence in every workgroup (to let the
// last one solve) writes back the L2 of its XCD each time, which cost more than the launch it saves; a separate
// one-workgroup solve launch a date cost 6-9 us more than redoing the solve in every workgroup of the next one.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

constexpr int kLsmBlock = 256;
constexpr int kMaxBasis = 7;
constexpr int kCoefStride = 7;  // coefficients of a date in the request's coefs array

template <int NB>
constexpr int n_sums() {
    return NB * (NB + 1) / 2 + NB + 1;  // X'X (lower triangle), X'y, the count
}

struct LsmArgs {
    const double* S;  // n_obs x rows: the spots at the dates
    const double* V;  // n_obs x rows: the variance factor (basis with v), or null
    double* value;    // rows: the discounted value of each path (pass 1)
    long long rows;   // 2 n_pairs
    int n_obs, min_itm;
    double inv_strike, phi;  // phi: +1 call, -1 put
    const double* disc;      // n_obs
    double* coefs;           // n_obs x kCoefStride
    int* valid;              // n_obs
    double* partial;         // 2 x workgroups x n_sums: the rows of date j in half j & 1
};

// products.py:105-109, with s^3 formed as s s s.
template <int NB>
__device__ __forceinline__ void basis(double s, double v, double (&X)[NB]) {
    X[0] = 1.0;
    X[1] = s;
    X[2] = s * s;
    X[3] = s * s * s;
    if constexpr (NB == 7) {
        X[4] = v;
        X[5] = v * v;
        X[6] = s * v;
    }
}

// Sum over a workgroup of K values per lane; thread k < K of the workgroup returns total k (others: 0). In a wave
// the K values (padded to P, a power of two) are reduce-scattered by halving: at the level of xor mask h a lane
// keeps the half of its values whose index has bit h equal to its lane's and adds its partner's copy of them, so
// after log2 P levels lane l holds value l & (P - 1) summed over its group of P lanes; the remaining levels add the
// groups. P - 1 + log2(64 / P) shuffles instead of 6 K. Then wave w's lane k < K writes slot (w, k), and thread k
// adds the waves in order. The order of every addition is fixed.
template <int K>
__device__ __forceinline__ double workgroup_sums(double (&acc)[K], double* lds /* [waves][K] */) {
    static_assert(K >= 1 && K <= 64, "at most 64 sums");
    constexpr int P = K <= 1 ? 1 : K <= 2 ? 2 : K <= 4 ? 4 : K <= 8 ? 8 : K <= 16 ? 16 : K <= 32 ? 32 : 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
    double v[P];
#pragma unroll
    for (int i = 0; i < P; ++i) v[i] = i < K ? acc[i] : 0.0;
#pragma unroll
    for (int half = P / 2; half >= 1; half >>= 1) {
        const bool upper = (lane & half) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const double keep = upper ? v[i + half] : v[i];
            const double give = upper ? v[i] : v[i + half];
            v[i] = keep + shfl_xor_d(give, half);
        }
    }
    double t = v[0];
#pragma unroll
    for (int m = P; m < 64; m <<= 1) t += shfl_xor_d(t, m);
    if (lane < K) lds[wave * K + lane] = t;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < K)
        for (int w = 0; w < n_waves; ++w) r += lds[w * K + threadIdx.x];
    return r;
}

// FP64 Cholesky solve of X'X beta = X'y from the sums (lower triangle of X'X, X'y, the count), by one lane; one
// division a column (the reciprocal of the diagonal, then products). False when the date has min_itm paths or fewer
// in the money, or X'X is not numerically positive definite.
template <int NB>
__device__ bool cholesky_solve(const double* tot, int min_itm, double (&beta)[NB]) {
    double L[NB][NB], b[NB], inv[NB];
    int m = 0;
#pragma unroll
    for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) L[r][c] = tot[m++];
#pragma unroll
    for (int r = 0; r < NB; ++r) b[r] = tot[m++];
    if (!(tot[m] > static_cast<double>(min_itm))) return false;
#pragma unroll
    for (int c = 0; c < NB; ++c) {  // A = L L' in place (lower triangle)
        double d = L[c][c];
#pragma unroll
        for (int k = 0; k < c; ++k) d -= L[c][k] * L[c][k];
        if (!(d > 0.0)) return false;
        L[c][c] = sqrt(d);
        inv[c] = 1.0 / L[c][c];
#pragma unroll
        for (int r = c + 1; r < NB; ++r) {
            double e = L[r][c];
#pragma unroll
            for (int k = 0; k < c; ++k) e -= L[r][k] * L[c][k];
            L[r][c] = e * inv[c];
        }
    }
    double y[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        double e = b[r];
#pragma unroll
        for (int k = 0; k < r; ++k) e -= L[r][k] * y[k];
        y[r] = e * inv[r];
    }
#pragma unroll
    for (int r = NB - 1; r >= 0; --r) {
        double e = y[r];
#pragma unroll
        for (int k = r + 1; k < NB; ++k) e -= L[k][r] * beta[k];
        beta[r] = e * inv[r];
    }
    return true;
}

// rows a lane loads before it uses any: at 2^20 rows and two workgroups a CU, a lane's whole share in one round
// of loads (four rows at a time took 1 us more a date with the variance basis, the same without)
constexpr int kLsmUnroll = 8;

// The solve of date d by one workgroup: the n_rows rows of its half of `partial` added in a fixed order (lane l
// takes rows l, l + 256, ..., then a fixed tree), then lane 0's Cholesky; beta and ok land in LDS for the whole
// workgroup. Every workgroup that calls it for a date computes the same bits.
template <int NB>
__device__ void solve_date(const LsmArgs& a, int d, int n_rows, double* lds, double* tot, double* beta_s, int* ok_s) {
    constexpr int K = n_sums<NB>();
    const double* rows = a.partial + static_cast<long long>(d & 1) * n_rows * K;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int r = threadIdx.x; r < n_rows; r += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += rows[static_cast<long long>(r) * K + k];
    const double total = workgroup_sums<K>(acc, lds);
    if (threadIdx.x < K) tot[threadIdx.x] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        double beta[NB];
        const bool ok = cholesky_solve<NB>(tot, a.min_itm, beta);
#pragma unroll
        for (int k = 0; k < NB; ++k) beta_s[k] = ok ? beta[k] : 0.0;
        *ok_s = ok ? 1 : 0;
    }
    __syncthreads();
}

// The coefficients of date d (NaN when not regressed) and its flag, for the apply pass and the caller.
template <int NB>
__device__ void publish_date(const LsmArgs& a, int d, const double* beta_s, int ok) {
    for (int k = threadIdx.x; k < kCoefStride; k += blockDim.x)
        a.coefs[d * kCoefStride + k] = ok ? (k < NB ? beta_s[k] : 0.0) : NAN;
    if (threadIdx.x == 0) a.valid[d] = ok;
}

// Date j of the backward sweep (products.py:151-158), one launch. Unless j + 1 is the last date, the workgroup first
// solves date j + 1 from the rows the previous launch wrote (every workgroup the same bits; workgroup 0 publishes
// them). Each path's discounted value is then brought to date j + 1: the payoff at maturity when j + 1 is the last
// date, else date j + 1's exercise decision (an in-the-money path exercises when its exercise value beats the
// fitted continuation). Then the in-the-money paths of date j add their terms of the normal equations to a
// per-workgroup row, in the half of `partial` of date j.
template <int NB>
__global__ void __launch_bounds__(kLsmBlock) lsm_date_kernel(LsmArgs a, int j) {
    constexpr int K = n_sums<NB>();
    constexpr int U = kLsmUnroll;
    __shared__ double lds[(kLsmBlock / 64) * K];
    __shared__ double tot[K], beta_s[NB];
    __shared__ int ok_s;
    const int nx = j + 1;
    const bool at_maturity = nx == a.n_obs - 1;
    bool rule = false;
    double bn[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) bn[k] = 0.0;
    if (!at_maturity) {
        solve_date<NB>(a, nx, gridDim.x, lds, tot, beta_s, &ok_s);
        rule = ok_s != 0;
#pragma unroll
        for (int k = 0; k < NB; ++k) bn[k] = beta_s[k];
        if (blockIdx.x == 0) publish_date<NB>(a, nx, beta_s, ok_s);
    }
    const double* Sn = a.S + static_cast<long long>(nx) * a.rows;
    const double* Vn = a.V ? a.V + static_cast<long long>(nx) * a.rows : nullptr;
    const double* Sj = a.S + static_cast<long long>(j) * a.rows;
    const double* Vj = a.V ? a.V + static_cast<long long>(j) * a.rows : nullptr;
    const double dn = ldc(a.disc, nx), dj = ldc(a.disc, j);
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    const long long tile = static_cast<long long>(kLsmBlock) * U;
    for (long long base = blockIdx.x * tile; base < a.rows; base += gridDim.x * tile) {
        double sn[U], sj[U], wn[U], wj[U], val[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = base + u * kLsmBlock + threadIdx.x;
            const long long k = i < a.rows ? i : 0;
            sn[u] = Sn[k];
            sj[u] = Sj[k];
            wn[u] = Vn ? Vn[k] : 0.0;
            wj[u] = Vj ? Vj[k] : 0.0;
            val[u] = at_maturity ? 0.0 : a.value[k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = base + u * kLsmBlock + threadIdx.x;
            if (i >= a.rows) continue;
            const double s1 = sn[u] * a.inv_strike;
            const double exn = fmax(a.phi * (s1 - 1.0), 0.0);
            double v = val[u];
            if (at_maturity) {
                v = exn * dn;
                a.value[i] = v;
            } else if (rule && exn > 0.0) {
                double X[NB];
                basis<NB>(s1, wn[u], X);
                double cont = 0.0;
#pragma unroll
                for (int k = 0; k < NB; ++k) cont += X[k] * bn[k];
                if (exn > cont) {
                    v = exn * dn;
                    a.value[i] = v;
                }
            }
            const double s = sj[u] * a.inv_strike;
            if (a.phi * (s - 1.0) > 0.0) {
                double X[NB];
                basis<NB>(s, wj[u], X);
                const double y = v / dj;
                int m = 0;
#pragma unroll
                for (int r = 0; r < NB; ++r)
#pragma unroll
                    for (int c = 0; c <= r; ++c) acc[m++] += X[r] * X[c];
#pragma unroll
                for (int r = 0; r < NB; ++r) acc[m++] += X[r] * y;
                acc[m] += 1.0;
            }
        }
    }
    const double t = workgroup_sums<K>(acc, lds);
    if (threadIdx.x < K)
        a.partial[(static_cast<long long>(j & 1) * gridDim.x + blockIdx.x) * K + threadIdx.x] = t;
}

// The first date's solve (one workgroup), after the last date kernel wrote its rows.
template <int NB>
__global__ void __launch_bounds__(kLsmBlock) lsm_solve_kernel(LsmArgs a, int j, int n_rows) {
    constexpr int K = n_sums<NB>();
    __shared__ double lds[(kLsmBlock / 64) * K];
    __shared__ double tot[K], beta_s[NB];
    __shared__ int ok_s;
    solve_date<NB>(a, j, n_rows, lds, tot, beta_s, &ok_s);
    publish_date<NB>(a, j, beta_s, ok_s);
}

struct Events {
    hipEvent_t e[3];
    Events() {
        for (auto& x : e) hip_check(hipEventCreate(&x), "event");
    }
    ~Events() {
        for (auto& x : e) (void)hipEventDestroy(x);
    }
    double ms(int a, int b) const {
        float t = 0.f;
        hip_check(hipEventElapsedTime(&t, e[a], e[b]), "elapsed");
        return t;
    }
};

template <int NB>
void sweep(optlab_lsm_request& rq, int prec) {
    const auto wall0 = std::chrono::steady_clock::now();
    hipStream_t st = engine_stream();
    const int n_obs = rq.paths.n_obs;
    const long long n_pairs = rq.paths.n_pairs, rows = 2 * n_pairs;
    constexpr int K = n_sums<NB>();
    // the sweep: two workgroups a CU, kLsmUnroll rows a lane at a time (the solve adds two rows a lane; four
    // workgroups a CU measured slower: the reduction is per workgroup)
    const long long tile = static_cast<long long>(kLsmBlock) * kLsmUnroll;
    const int blocks = static_cast<int>(
        std::max<long long>(1, std::min<long long>((rows + tile - 1) / tile, 2LL * device_cu_count())));
    const long long apply_blocks = exercise_workgroups(n_pairs);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t s_bytes = up(static_cast<size_t>(n_obs) * rows * sizeof(double));
    const size_t v_bytes = NB == 7 ? s_bytes : 0;
    const size_t val_bytes = up(static_cast<size_t>(rows) * sizeof(double));
    const size_t part_bytes = up(2 * static_cast<size_t>(blocks) * K * sizeof(double));
    const size_t coef_bytes = up(static_cast<size_t>(n_obs) * kCoefStride * sizeof(double));
    const size_t valid_bytes = up(static_cast<size_t>(n_obs) * sizeof(int));
    const size_t disc_bytes = up(static_cast<size_t>(n_obs) * sizeof(double));
    const size_t sums_bytes = up(static_cast<size_t>(apply_blocks) * 2 * sizeof(double));
    unsigned char* m = static_cast<unsigned char*>(scratch(
        kSlotLsm, s_bytes + v_bytes + val_bytes + part_bytes + coef_bytes + valid_bytes + disc_bytes + sums_bytes));
    LsmArgs a{};
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        unsigned char* p = m + off;
        off += bytes;
        return p;
    };
    double* S = reinterpret_cast<double*>(carve(s_bytes));
    double* V = NB == 7 ? reinterpret_cast<double*>(carve(v_bytes)) : nullptr;
    a.S = S, a.V = V;
    a.value = reinterpret_cast<double*>(carve(val_bytes));
    a.partial = reinterpret_cast<double*>(carve(part_bytes));
    a.coefs = reinterpret_cast<double*>(carve(coef_bytes));
    a.valid = reinterpret_cast<int*>(carve(valid_bytes));
    double* disc = reinterpret_cast<double*>(carve(disc_bytes));
    double* sums = reinterpret_cast<double*>(carve(sums_bytes));
    a.disc = disc;
    a.rows = rows, a.n_obs = n_obs, a.min_itm = rq.min_itm;
    a.inv_strike = 1.0 / rq.strike, a.phi = rq.call ? 1.0 : -1.0;
    hip_check(hipMemcpyAsync(disc, rq.disc, n_obs * sizeof(double), hipMemcpyHostToDevice, st), "copy disc");

    Events ev;
    // pass 1: paths, then the backward sweep of regressions
    hip_check(hipEventRecord(ev.e[0], st), "event");
    double paths_ms = launch_paths(rq.paths, prec, S, st, V);
    hip_check(hipEventRecord(ev.e[1], st), "event");
    // every date's rows are solved by the next launch (date 0's by the solve kernel) before they are overwritten
    for (int j = n_obs - 2; j >= 0; --j)
        hipLaunchKernelGGL((lsm_date_kernel<NB>), dim3(blocks), dim3(kLsmBlock), 0, st, a, j);
    hipLaunchKernelGGL((lsm_solve_kernel<NB>), dim3(1), dim3(kLsmBlock), 0, st, a, 0, blocks);
    hip_check(hipGetLastError(), "lsm sweep launch");
    hip_check(hipEventRecord(ev.e[2], st), "event");
    // pass 2: the next n_pairs pairs of the same seed, the rule applied as the paths reach each date
    optlab_paths_request p2 = rq.paths;
    p2.first_pair = rq.paths.first_pair + n_pairs;
    const ExerciseRule rule{a.coefs, a.valid, disc, rq.strike, a.phi, sums};
    const double apply_ms = launch_paths_exercise(p2, prec, rule, st);
    std::vector<double> h(static_cast<size_t>(apply_blocks) * 2);
    hip_check(hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, st), "copy sums");
    std::vector<double> coefs(static_cast<size_t>(n_obs) * kCoefStride);
    hip_check(hipMemcpyAsync(coefs.data(), a.coefs, coefs.size() * sizeof(double), hipMemcpyDeviceToHost, st),
              "copy coefficients");
    hip_check(hipStreamSynchronize(st), "synchronize");

    double s1 = 0.0, s2 = 0.0;
    for (long long b = 0; b < apply_blocks; ++b) s1 += h[2 * b], s2 += h[2 * b + 1];
    const double n = static_cast<double>(n_pairs);
    rq.price = s1 / n;
    rq.se = n > 1 ? std::sqrt(std::max((s2 - s1 * s1 / n) / (n - 1.0), 0.0) / n) : 0.0;
    const double immediate = rq.strike * std::max((rq.call ? 1.0 : -1.0) * (rq.spot / rq.strike - 1.0), 0.0);
    if (immediate > rq.price) rq.price = immediate, rq.se = 0.0;  // products.py:165-168
    for (int k = 0; k < kCoefStride; ++k) coefs[static_cast<size_t>(n_obs - 1) * kCoefStride + k] = NAN;  // maturity
    if (rq.coefs) std::memcpy(rq.coefs, coefs.data(), coefs.size() * sizeof(double));
    rq.paths_ms = paths_ms;
    rq.regression_ms = ev.ms(1, 2);
    rq.apply_ms = apply_ms;
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

}  // namespace

void check_lsm_request(const optlab_lsm_request& rq) {
    const optlab_paths_request& p = rq.paths;
    check_paths_request(p);
    if (p.n_names != 1) invalid("an American option is on one name");
    if (p.n_obs < 2) invalid("n_obs (exercise dates) must be >= 2");
    if (p.n_windows != 0 || p.n_keep != p.n_obs || p.n_cols != p.n_obs) invalid("keep every exercise date, no windows");
    for (int j = 0; j < p.n_obs; ++j)
        if (p.col_of[j] != j) invalid("keep every exercise date, no windows");
    if (p.n_pairs < 1) invalid("n_pairs must be >= 1");
    if (!(rq.strike > 0)) invalid("strike must be > 0");
    if (!rq.disc) invalid("null array");
    for (int j = 0; j < p.n_obs; ++j)
        if (!(rq.disc[j] > 0)) invalid("discount factors must be > 0");
    if (rq.min_itm < 0) invalid("min_itm must be >= 0");
    if (rq.basis_var && !(p.kind[0] == 2 || p.kind[0] == 3)) invalid("basis_var needs a Heston or SLV name");
}

void run_lsm(optlab_lsm_request& rq, int prec) {
    if (rq.basis_var)
        sweep<kMaxBasis>(rq, prec);
    else
        sweep<4>(rq, prec);
}

}  // namespace optlab_mi
