// The issuer's call regression: a least-squares fit of y on polynomial and hinge terms of a state, in FP64 normal
// equations on the device.
// Algorithm: packages/structured/src/optlab/structured/exercise.py:191-238 (fit: at each call date, the note's
// continuation regressed on the terms of the state over the paths the call reaches) and :150-171 (Rule.features,
// Rule.continuation). The reference solves with NumPy's lstsq; here the terms are formed in-kernel from the state,
// X'X and X'y are summed in FP64 in a fixed order and the system is solved by a Cholesky factorisation.
//
// MI355X design: one pass over the rows, 256-row tiles of terms staged in LDS by the workgroup, each lane owning
// one or two of the (at most 325 + 25) entries of X'X and X'y and adding the tile's rows to them in order: one row
// of sums per workgroup. One workgroup then adds the rows in order and factorises, the entries of each column of L
// spread over its lanes (each dot product in the serial order). The fitted continuation is evaluated on a second
// set of states in one more pass, its terms formed on the fly.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

constexpr int kRegThreads = 256;
constexpr int kMaxTerms = 25;  // exercise.py:28 (MAX_TERMS)
constexpr int kMaxVars = 8;
// A pivot of the factorisation below this fraction of its diagonal means the term lies numerically in the span of
// the terms before it (a repeated term, a hinge outside the states): the normal equations are then meaningless, and
// the call reports ok = 0 instead of a solution.
constexpr double kPivotTol = 1e-13;

struct RegSpec {  // device copy of the terms
    double mean[kMaxVars], scale[kMaxVars];
    int powers[kMaxTerms * kMaxVars];  // [term][variable], stride kMaxVars
    int knot_var[kMaxTerms];
    double knot[kMaxTerms];
};

struct RegArgs {
    const RegSpec* spec;
    const double* state;  // n x n_vars
    const double* y;      // n
    long long n;
    int n_vars, n_powers, n_knots;
    double* partial;  // workgroups x n_sums
};

__device__ __forceinline__ double ipow(double z, int e) {  // exponents 0..3, the cube as z z z
    return e == 0 ? 1.0 : e == 1 ? z : e == 2 ? z * z : z * z * z;
}

// z = (x - mean) / scale of one row (exercise.py:165).
__device__ __forceinline__ void normalise(const RegSpec* sp, const double* x, int nv, double (&z)[kMaxVars]) {
#pragma unroll
    for (int i = 0; i < kMaxVars; ++i) z[i] = i < nv ? (x[i] - ldc(sp->mean, i)) / ldc(sp->scale, i) : 0.0;
}

// Term t of a row (exercise.py:166-168): prod_i z_i^p_i for t < n_powers (the product in variable order), else
// the hinge max(z_v - k, 0).
__device__ __forceinline__ double term(const RegSpec* sp, const double (&z)[kMaxVars], int nv, int n_powers, int t) {
    if (t < n_powers) {
        double f = 1.0;
#pragma unroll
        for (int i = 0; i < kMaxVars; ++i)
            if (i < nv) f *= ipow(z[i], ldc(sp->powers, t * kMaxVars + i));
        return f;
    }
    const int h = t - n_powers;
    const int v = ldc(sp->knot_var, h);
    double zv = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxVars; ++i)
        if (i == v) zv = z[i];
    return fmax(zv - ldc(sp->knot, h), 0.0);
}

__device__ __forceinline__ int n_sums(int p) { return p * (p + 1) / 2 + p; }  // X'X (lower triangle), X'y

// Entry e of the sums -> (r, c): X'X[r][c] for e < p (p + 1) / 2 (rows of the lower triangle in order), else
// X'y[r] (c = p).
__device__ __forceinline__ void entry(int e, int p, int& r, int& c) {
    const int tri = p * (p + 1) / 2;
    if (e >= tri) {
        r = e - tri, c = p;
        return;
    }
    r = 0;
    while ((r + 1) * (r + 2) / 2 <= e) ++r;
    c = e - r * (r + 1) / 2;
}

constexpr int kEntriesPerLane = (kMaxTerms * (kMaxTerms + 1) / 2 + kMaxTerms + kRegThreads - 1) / kRegThreads;

__global__ void __launch_bounds__(kRegThreads) regress_sums_kernel(RegArgs a) {
    extern __shared__ double lds[];  // terms [256][p], then y [256]
    const int p = a.n_powers + a.n_knots, K = n_sums(p), tid = threadIdx.x;
    double* feat = lds;
    double* yv = lds + kRegThreads * p;
    int er[kEntriesPerLane], ec[kEntriesPerLane];
    double acc[kEntriesPerLane];
#pragma unroll
    for (int m = 0; m < kEntriesPerLane; ++m) {
        const int e = tid + m * kRegThreads;
        if (e < K) entry(e, p, er[m], ec[m]);
        else er[m] = ec[m] = 0;
        acc[m] = 0.0;
    }
    const long long n_tiles = (a.n + kRegThreads - 1) / kRegThreads;
    for (long long tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const long long row = tile * kRegThreads + tid;
        if (row < a.n) {
            double x[kMaxVars], z[kMaxVars];
#pragma unroll
            for (int i = 0; i < kMaxVars; ++i) x[i] = i < a.n_vars ? a.state[row * a.n_vars + i] : 0.0;
            normalise(a.spec, x, a.n_vars, z);
            for (int t = 0; t < p; ++t) feat[tid * p + t] = term(a.spec, z, a.n_vars, a.n_powers, t);
            yv[tid] = a.y[row];
        } else {  // rows past the end add exact zeros
            for (int t = 0; t < p; ++t) feat[tid * p + t] = 0.0;
            yv[tid] = 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kEntriesPerLane; ++m) {
            if (tid + m * kRegThreads >= K) continue;
            const int r = er[m], c = ec[m];
            double s = acc[m];
            if (c < p) {
                for (int i = 0; i < kRegThreads; ++i) s += feat[i * p + r] * feat[i * p + c];
            } else {
                for (int i = 0; i < kRegThreads; ++i) s += feat[i * p + r] * yv[i];
            }
            acc[m] = s;
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < kEntriesPerLane; ++m) {
        const int e = tid + m * kRegThreads;
        if (e < K) a.partial[static_cast<long long>(blockIdx.x) * K + e] = acc[m];
    }
}

// One workgroup: the rows added in order (lane e owns entry e), the Cholesky factorisation of X'X column by column
// (the diagonal by lane 0, the column's other entries by one lane each, every dot product in the serial order, one
// reciprocal a column), the two triangular solves by lane 0.
__global__ void __launch_bounds__(kRegThreads) regress_solve_kernel(RegArgs a, int n_rows, double* coef, int* ok_out) {
    __shared__ double tot[kMaxTerms * (kMaxTerms + 1) / 2 + kMaxTerms];
    __shared__ double L[kMaxTerms][kMaxTerms];
    __shared__ double inv[kMaxTerms], w[kMaxTerms];
    __shared__ int ok;
    const int p = a.n_powers + a.n_knots, K = n_sums(p), tid = threadIdx.x;
    for (int e = tid; e < K; e += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < n_rows; ++b) s += a.partial[static_cast<long long>(b) * K + e];
        tot[e] = s;
    }
    if (tid == 0) ok = 1;
    __syncthreads();
    for (int e = tid; e < p * (p + 1) / 2; e += blockDim.x) {
        int r, c;
        entry(e, p, r, c);
        L[r][c] = tot[e];
    }
    __syncthreads();
    for (int c = 0; c < p; ++c) {
        if (tid == 0) {
            const double diag = L[c][c];
            double d = diag;
            for (int k = 0; k < c; ++k) d -= L[c][k] * L[c][k];
            if (!(d > kPivotTol * diag)) {
                ok = 0;
            } else {
                L[c][c] = sqrt(d);
                inv[c] = 1.0 / L[c][c];
            }
        }
        __syncthreads();
        if (!ok) break;
        const int r = c + 1 + tid;
        if (r < p) {
            double e = L[r][c];
            for (int k = 0; k < c; ++k) e -= L[r][k] * L[c][k];
            L[r][c] = e * inv[c];
        }
        __syncthreads();
    }
    if (tid != 0) return;
    *ok_out = ok;
    if (!ok) return;
    const int tri = p * (p + 1) / 2;
    for (int r = 0; r < p; ++r) {  // L w = X'y
        double e = tot[tri + r];
        for (int k = 0; k < r; ++k) e -= L[r][k] * w[k];
        w[r] = e * inv[r];
    }
    for (int r = p - 1; r >= 0; --r) {  // L' beta = w
        double e = w[r];
        for (int k = r + 1; k < p; ++k) e -= L[k][r] * coef[k];
        coef[r] = e * inv[r];
    }
}

// exercise.py:170-171: the fitted continuation on each state, terms formed on the fly and added in term order
// (NaN everywhere when the solve failed).
__global__ void __launch_bounds__(kRegThreads) regress_eval_kernel(RegArgs a, const double* coef, const int* ok,
                                                                   double* out) {
    const int p = a.n_powers + a.n_knots;
    const bool solved = ldc(ok, 0) != 0;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long row = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; row < a.n; row += stride) {
        double x[kMaxVars], z[kMaxVars];
#pragma unroll
        for (int i = 0; i < kMaxVars; ++i) x[i] = i < a.n_vars ? a.state[row * a.n_vars + i] : 0.0;
        normalise(a.spec, x, a.n_vars, z);
        double s = 0.0;
        for (int t = 0; t < p; ++t) s += term(a.spec, z, a.n_vars, a.n_powers, t) * ldc(coef, t);
        out[row] = solved ? s : NAN;
    }
}

}  // namespace

void check_regress_request(const optlab_regress_request& rq) {
    if (rq.n_rows < 1) invalid("n_rows must be >= 1");
    if (rq.n_vars < 1 || rq.n_vars > kMaxVars) invalid("n_vars must be 1 to 8");
    if (!rq.state || !rq.y || !rq.mean || !rq.scale || !rq.coef) invalid("null array");
    if (rq.n_powers < 0 || rq.n_knots < 0 || rq.n_powers + rq.n_knots < 1 || rq.n_powers + rq.n_knots > kMaxTerms)
        invalid("n_powers + n_knots must be 1 to 25");
    if (rq.n_powers > 0 && !rq.powers) invalid("null array");
    if (rq.n_knots > 0 && (!rq.knot_var || !rq.knot)) invalid("null array");
    for (int i = 0; i < rq.n_vars; ++i)
        if (!(rq.scale[i] != 0.0) || !std::isfinite(rq.scale[i]) || !std::isfinite(rq.mean[i]))
            invalid("scale must be finite and nonzero, mean finite");
    for (int i = 0; i < rq.n_powers * rq.n_vars; ++i)
        if (rq.powers[i] < 0 || rq.powers[i] > 3) invalid("powers must be 0 to 3");
    for (int h = 0; h < rq.n_knots; ++h)
        if (rq.knot_var[h] < 0 || rq.knot_var[h] >= rq.n_vars || !std::isfinite(rq.knot[h]))
            invalid("knot_var must name a variable and knots be finite");
    if (rq.n_eval < 0 || (rq.n_eval > 0 && (!rq.eval_state || !rq.eval_out))) invalid("bad evaluation arrays");
}

void run_regress(optlab_regress_request& rq) {
    const auto wall0 = std::chrono::steady_clock::now();
    hipStream_t st = engine_stream();
    const int p = rq.n_powers + rq.n_knots, K = p * (p + 1) / 2 + p, nv = rq.n_vars;
    const long long n = rq.n_rows, n_tiles = (n + kRegThreads - 1) / kRegThreads;
    const int blocks = static_cast<int>(std::max<long long>(1, std::min<long long>(n_tiles, 2LL * device_cu_count())));
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t spec_b = up(sizeof(RegSpec)), state_b = up(static_cast<size_t>(n) * nv * 8), y_b = up(static_cast<size_t>(n) * 8);
    const size_t part_b = up(static_cast<size_t>(blocks) * K * 8), coef_b = up(kMaxTerms * 8), ok_b = up(sizeof(int));
    const size_t ev_state_b = up(static_cast<size_t>(rq.n_eval) * nv * 8), ev_out_b = up(static_cast<size_t>(rq.n_eval) * 8);
    unsigned char* m = static_cast<unsigned char*>(
        scratch(kSlotRegress, spec_b + state_b + y_b + part_b + coef_b + ok_b + ev_state_b + ev_out_b));
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        unsigned char* q = m + off;
        off += bytes;
        return q;
    };
    RegSpec* d_spec = reinterpret_cast<RegSpec*>(carve(spec_b));
    double* d_state = reinterpret_cast<double*>(carve(state_b));
    double* d_y = reinterpret_cast<double*>(carve(y_b));
    double* d_part = reinterpret_cast<double*>(carve(part_b));
    double* d_coef = reinterpret_cast<double*>(carve(coef_b));
    int* d_ok = reinterpret_cast<int*>(carve(ok_b));
    double* d_ev_state = reinterpret_cast<double*>(carve(ev_state_b));
    double* d_ev_out = reinterpret_cast<double*>(carve(ev_out_b));

    RegSpec spec{};
    for (int i = 0; i < nv; ++i) spec.mean[i] = rq.mean[i], spec.scale[i] = rq.scale[i];
    for (int t = 0; t < rq.n_powers; ++t)
        for (int i = 0; i < nv; ++i) spec.powers[t * kMaxVars + i] = rq.powers[t * nv + i];
    for (int h = 0; h < rq.n_knots; ++h) spec.knot_var[h] = rq.knot_var[h], spec.knot[h] = rq.knot[h];
    hip_check(hipMemcpyAsync(d_spec, &spec, sizeof spec, hipMemcpyHostToDevice, st), "copy terms");
    hip_check(hipMemcpyAsync(d_state, rq.state, static_cast<size_t>(n) * nv * 8, hipMemcpyHostToDevice, st), "copy state");
    hip_check(hipMemcpyAsync(d_y, rq.y, static_cast<size_t>(n) * 8, hipMemcpyHostToDevice, st), "copy y");
    if (rq.n_eval > 0)
        hip_check(hipMemcpyAsync(d_ev_state, rq.eval_state, static_cast<size_t>(rq.n_eval) * nv * 8,
                                 hipMemcpyHostToDevice, st),
                  "copy evaluation states");

    hipEvent_t ev[3];
    for (auto& e : ev) hip_check(hipEventCreate(&e), "event");
    RegArgs a{d_spec, d_state, d_y, n, nv, rq.n_powers, rq.n_knots, d_part};
    hip_check(hipEventRecord(ev[0], st), "event");
    const size_t lds = static_cast<size_t>(kRegThreads) * (p + 1) * sizeof(double);
    hipLaunchKernelGGL(regress_sums_kernel, dim3(blocks), dim3(kRegThreads), lds, st, a);
    hipLaunchKernelGGL(regress_solve_kernel, dim3(1), dim3(kRegThreads), 0, st, a, blocks, d_coef, d_ok);
    hip_check(hipGetLastError(), "regression launch");
    hip_check(hipEventRecord(ev[1], st), "event");
    if (rq.n_eval > 0) {
        RegArgs e = a;
        e.state = d_ev_state, e.n = rq.n_eval;
        const int eb = static_cast<int>(std::max<long long>(
            1, std::min<long long>((rq.n_eval + kRegThreads - 1) / kRegThreads, 4LL * device_cu_count())));
        hipLaunchKernelGGL(regress_eval_kernel, dim3(eb), dim3(kRegThreads), 0, st, e, d_coef, d_ok, d_ev_out);
        hip_check(hipGetLastError(), "evaluation launch");
    }
    hip_check(hipEventRecord(ev[2], st), "event");
    int ok = 0;
    std::vector<double> coef(p);
    hip_check(hipMemcpyAsync(&ok, d_ok, sizeof ok, hipMemcpyDeviceToHost, st), "copy flag");
    hip_check(hipMemcpyAsync(coef.data(), d_coef, p * 8, hipMemcpyDeviceToHost, st), "copy coefficients");
    if (rq.n_eval > 0)
        hip_check(hipMemcpyAsync(rq.eval_out, d_ev_out, static_cast<size_t>(rq.n_eval) * 8, hipMemcpyDeviceToHost, st),
                  "copy continuation");
    hip_check(hipStreamSynchronize(st), "synchronize");
    float f = 0.f, g = 0.f;
    hip_check(hipEventElapsedTime(&f, ev[0], ev[1]), "elapsed");
    hip_check(hipEventElapsedTime(&g, ev[1], ev[2]), "elapsed");
    for (auto& e : ev) (void)hipEventDestroy(e);
    rq.ok = ok;
    if (ok) std::memcpy(rq.coef, coef.data(), p * sizeof(double));
    rq.fit_ms = f, rq.eval_ms = g;
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

}  // namespace optlab_mi
