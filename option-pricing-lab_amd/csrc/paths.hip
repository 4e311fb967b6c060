// Path engine kernel: every model of optlab (constant variance, Dupire local variance table, Heston, SLV), 1 to 8
// correlated names, optional local correlation, observations and running window extremes written as doubles.
// Semantics: packages/native/cuda/include/optlab_cuda/paths.hpp:196-361 (simulate_pair, record, table_value).
//
// MI355X layout: one antithetic pair per lane, state (x, v, window extremes, the normal buffer) in VGPRs with every
// per-name loop unrolled to the compile-time bound N so nothing spills; the table rows of the current step are read
// where they lie (a step's rows are a few KB, every lane of the launch reads the same ones: they stay in L2 and the
// lanes' gathers hit L1/L2; staging them in LDS with a barrier a step was measured 1.2-2.7x slower on every model);
// output columns are structure-of-arrays, so a wave writes 64 consecutive doubles per column.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

// A close from its log-forward-moneyness: FP32 by the hardware exp2 (exp(x) = 2^(x log2 e), about 2 ulp; the library
// expf adds a range reduction and clamps that a close never needs, 11 instructions against 2); FP64 by exp.
__device__ __forceinline__ float close_exp(float x) { return hw_exp(x); }
__device__ __forceinline__ double close_exp(double x) { return exp(x); }

constexpr int kBlock = 256;
constexpr int kMaxNames = 8;

// An exercise rule in device memory (one pointer in the kernel's arguments: the fields are loaded where the rule is
// applied, so they hold no scalar registers over the step loop).
struct ExerciseDev {
    const double* coefs;  // n_obs x 7: the continuation value's coefficients of each date
    const int* valid;     // n_obs: 1 where the date has a rule
    const double* disc;   // n_obs: D(t_j)
    double strike, inv_strike, phi;
    double* sums;         // 2 a workgroup: the sum of the pairs' average pay and of its square
};
typedef const __attribute__((address_space(4))) ExerciseDev* ExerciseConst;

// One observation of the schedule, loaded with one scalar load: the step after which it falls, its output column
// (-1: not kept) and its window flags (bit w: inside window w; bit 2 + w: its first observation; bit 4 + w: its
// last), formed on the host from the windows' index ranges. (The step, the column and the windows' bounds read and
// compared one by one cost the fixing-heavy pass 4-6 % and the phoenix script's daily window 8 %.)
struct ObsRec {
    int step, col, flags, pad;
};

template <typename R>
struct PathArgs {
    int n_names, n_steps, n_obs, n_x;
    long long n_pairs, first_pair, rows;
    std::uint32_t key0, key1;
    const R* dt;
    const R* t0;
    const ObsRec* orec;  // n_obs
    const R* root;
    const int* kind;
    const R* cvar;
    const R* hes;
    int same;
    const int* step_row;
    const R* lo;
    const R* inv_dx;
    const R* table;
    const R* fwd;
    int mixer;
    R mix[8];
    const R* low_root;
    int n_cols, n_keep;
    double* out;
    double* out_var;     // or null: the variance factor of name 0 at the kept columns, as out's first name
    int n_tab;           // names reading a table (the form's choice)
    int has_var, has_slv;
    const ExerciseDev* ex;  // kFormExercise: the rule applied at the observations of one name, in place of writing
};

// paths.hpp:196-205 on a row held at y.
template <typename R>
__device__ __forceinline__ R lerp_row(const R* y, R lo, R inv_dx, int n_x, R x) {
    R u = (x - lo) * inv_dx;
    const R top = static_cast<R>(n_x - 1) - static_cast<R>(1e-4);
    u = r_min(r_max(u, R(0)), top);
    const int i = static_cast<int>(u);
    const R w = u - static_cast<R>(i);
    return y[i] * (R(1) - w) + y[i + 1] * w;
}

template <typename R, int N>
struct Extremes {
    R lo[2][N][2], hi[2][N][2];
};

// paths.hpp:241-272, for the two legs of a lane's pair (output rows row[0], row[1]).
template <typename R, int N>
__device__ __forceinline__ void record(const PathArgs<R>& s, int n, const long long (&row)[2], bool active, int flags, int c,
                                       const R (&fw)[N], const R (&x)[2][N], const R (&v)[2][N], Extremes<R, N>& e) {
    if (s.out_var && c >= 0 && active) {
#pragma unroll
        for (int l = 0; l < 2; ++l) s.out_var[c * s.rows + row[l]] = static_cast<double>(v[l][0]);
    }
#pragma unroll
    for (int a = 0; a < N; ++a) {
        if (a >= n) break;
        const R f = fw[a];
        R S[2];
#pragma unroll
        for (int l = 0; l < 2; ++l) S[l] = f * close_exp(x[l][a]);
        double* base = s.out + static_cast<long long>(a) * s.n_cols * s.rows;
        if (c >= 0 && active) {
#pragma unroll
            for (int l = 0; l < 2; ++l) base[c * s.rows + row[l]] = static_cast<double>(S[l]);
        }
        if (!flags) continue;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            if (!(flags & (1 << w))) continue;
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                if (flags & (4 << w)) {
                    e.lo[l][a][w] = e.hi[l][a][w] = S[l];
                } else {
                    e.lo[l][a][w] = r_min(e.lo[l][a][w], S[l]);
                    e.hi[l][a][w] = r_max(e.hi[l][a][w], S[l]);
                }
            }
            if ((flags & (16 << w)) && active) {
                const long long cl = s.n_keep + 2 * w, ch = cl + 1;
#pragma unroll
                for (int l = 0; l < 2; ++l) {
                    base[cl * s.rows + row[l]] = static_cast<double>(e.lo[l][a][w]);
                    base[ch * s.rows + row[l]] = static_cast<double>(e.hi[l][a][w]);
                }
            }
        }
    }
}

// paths.hpp:274-361: lane p carries both legs of pair p (output rows p and n_pairs + p).
// The kernel's form, by what the request needs:
// kFormCV: every name has a constant variance and there is no mixer (Black-Scholes names and baskets), so the
// model dispatch, the variance factor and the table reads compile away: the same arithmetic on fewer scalar
// instructions a step;
// kFormGeneric: tables, Heston and SLV names, no mixer;
// kFormMixed: the generic form with a mixer (local correlation, or the stochastic weight with its state and its
// driver's stream), kept apart so that the other forms carry none of its registers;
// kFormExercise, added to kFormCV or kFormGeneric (one name): nothing is written at the observations; an exercise
// rule is applied there instead and the workgroup's sums of the pay are written.
// kFormTable: every name reads a local-variance table and none has a variance state or a mixer (local-vol names and
// baskets): the model dispatch and the variance factor compile away.
constexpr int kFormCV = 0, kFormGeneric = 1, kFormMixed = 2, kFormTable = 3, kFormExercise = 4;

// The exercise state of a lane's two legs: the exercise value (in units of the strike) and its date.
struct Exercise {
    double cash[2];
    int tau[2];
    bool done[2];
};

// products.py:137-150 at observation j, for both legs: a leg exercises at the first date whose rule says so (the
// reference's backward loop keeps the earliest), else at maturity, where it takes its payoff. The spot is formed
// as record forms it (f exp(x) in the path precision, then widened), so the rule sees the values its regressions
// were fitted on; the basis is (1, s, s^2, s^3) in s = S / K, then (v, v^2, s v), with s^3 formed as s s s.
template <typename R, int N>
__device__ __forceinline__ void exercise(const PathArgs<R>& s, int j, const R (&x)[2][N], const R (&v)[2][N],
                                         Exercise& e) {
    const ExerciseConst r = (ExerciseConst)s.ex;
    const int last = s.n_obs - 1;
    const R f = ldc(s.fwd, j);
    const bool rule = j < last && ldc(r->valid, j);
    // every date has its row (unread where there is no rule): loaded unconditionally, the row's seven loads are
    // issued together and waited for once, not one branch and one wait each
    double beta[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) beta[k] = ldc(r->coefs, j * 7 + k);
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (e.done[l]) continue;
        const double sx = static_cast<double>(f * close_exp(x[l][0])) * r->inv_strike;
        const double ex = fmax(r->phi * (sx - 1.0), 0.0);
        if (j == last) {
            e.cash[l] = ex;
            continue;
        }
        if (!rule || !(ex > 0.0)) continue;
        const double w = static_cast<double>(v[l][0]);
        const double X[7] = {1.0, sx, sx * sx, sx * sx * sx, w, w * w, sx * w};
        // the coefficients of the variance terms are 0 with the basis without v: each adds an exact 0
        double cont = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k) cont += X[k] * beta[k];
        if (ex > cont) {
            e.cash[l] = ex;
            e.tau[l] = j;
            e.done[l] = true;
        }
    }
}

// The workgroup's sums of a and b (every lane one value each), in a fixed order: a butterfly over each wave, then
// the waves in order; written to out[2 blockIdx.x], out[2 blockIdx.x + 1].
__device__ __forceinline__ void workgroup_pair_sums(double a, double b, double* out) {
    __shared__ double part[kBlock / 64][2];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        a += __shfl_xor(a, m);
        b += __shfl_xor(b, m);
    }
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64][0] = a, part[threadIdx.x / 64][1] = b;
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += part[w][threadIdx.x];
        out[2 * blockIdx.x + threadIdx.x] = t;
    }
}

template <typename R, int N, int FORM>
__global__ void __launch_bounds__(kBlock) paths_kernel(PathArgs<R> s) {
    constexpr bool EXER = (FORM & kFormExercise) != 0;
    constexpr bool CV = (FORM & 3) == kFormCV;
    constexpr bool TAB = (FORM & 3) == kFormTable;
    static_assert(!EXER || (N == 1 && (FORM & 3) != kFormMixed), "an exercise rule is on one name, no mixer");
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = p < s.n_pairs;
    // the dispatch launches N = 1, 2, 3, 4 for exactly that many names (N = 8 for 5 to 8): a constant count folds
    // every per-name bound away
    const int n = N <= 4 ? N : s.n_names;
    const long long row[2] = {p, s.n_pairs + p};
    const std::uint64_t pair = static_cast<std::uint64_t>(s.first_pair + (active ? p : 0));

    R x[2][N], v[2][N], cvar[N];
    int kind[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
        kind[a] = TAB ? 1 : (a < n ? ldc(s.kind, a) : 0);
        cvar[a] = a < n ? ldc(s.cvar, a) : R(0);
        const bool hv = kind[a] == 2 || kind[a] == 3;
        x[0][a] = x[1][a] = R(0);
        v[0][a] = v[1][a] = (a < n && hv) ? ldc(s.hes, a * 5) : R(0);
    }
    const bool has_var = !CV && !TAB && s.has_var;
    constexpr bool mixer = FORM == kFormMixed;
    Extremes<R, N> ext;
    Exercise exs{{0.0, 0.0}, {s.n_obs - 1, s.n_obs - 1}, {false, false}};
    // The next observation's step, output column and forwards are loaded one observation ahead, and the next
    // step's dt one step ahead: a scalar load's wait then falls after a step of vector work instead of at its
    // start (every step waited for its constants when each was loaded where it was used).
    int next_step = -1, next_col = -1, next_flags = 0;
    R next_fw[N];
    auto load_obs = [&](int jj) {
        const bool in = jj < s.n_obs;
        typedef int v4i __attribute__((ext_vector_type(4)));
        const v4i o = in ? ldc(reinterpret_cast<const v4i*>(s.orec), jj) : v4i{-1, -1, 0, 0};
        next_step = o.x;
        if constexpr (!EXER) {
            next_col = o.y, next_flags = o.z;
#pragma unroll
            for (int a = 0; a < N; ++a) next_fw[a] = (in && a < n) ? ldc(s.fwd, a * s.n_obs + jj) : R(0);
        }
    };
    auto observe = [&](int jj) {
        if constexpr (EXER) {
            exercise<R, N>(s, jj, x, v, exs);
        } else {
            R fw[N];
#pragma unroll
            for (int a = 0; a < N; ++a) fw[a] = next_fw[a];
            record<R, N>(s, n, row, active, next_flags, next_col, fw, x, v, ext);
        }
    };
    int j = 0;
    load_obs(0);
    while (next_step == 0) {
        observe(j++);
        load_obs(j);
    }

    NormalStream<R> rng, aux, drv;
    rng.init(pair, s.key0, s.key1, 0u);
    aux.init(pair, s.key0, s.key1, kAuxTag);
    if (mixer) drv.init(pair, s.key0, s.key1, kMixTag);
    const bool jacobi = mixer && s.mixer == 2;  // the stochastic weight: per-leg state, its own driver
    R wst[2] = {s.mix[0], s.mix[0]};

    R dt_next = s.n_steps > 0 ? ldc(s.dt, 0) : R(0);
    auto one_step = [&](int i) __attribute__((always_inline)) {
        const R dt = dt_next, t0 = mixer ? ldc(s.t0, i - 1) : R(0);
        if (i < s.n_steps) dt_next = ldc(s.dt, i);
        R zs[N], zr[N], fresh = R(0);
#pragma unroll
        for (int a = 0; a < N; ++a)
            if (a < n) zs[a] = rng.next();
        if (has_var) {
#pragma unroll
            for (int a = 0; a < N; ++a)
                if (a < n) zr[a] = s.has_slv ? aux.next() : rng.next();
        }
        R drive = R(0);
        if (mixer) fresh = rng.next();
        if (jacobi) drive = drv.next();
        // root . zs (and root . zr for Heston/SLV names with the spot normals' correlation) do not depend on the
        // leg: the - leg takes their negatives
        R zc[N], wc[N];
        if (!mixer) {
#pragma unroll
            for (int a = 0; a < N; ++a) {
                if (a >= n) break;
                R acc = R(0);
#pragma unroll
                for (int b = 0; b < N; ++b)
                    if (b < n) acc += ldc(s.root, a * n + b) * zs[b];
                zc[a] = acc;
            }
        }
        if (has_var && s.same) {
#pragma unroll
            for (int a = 0; a < N; ++a) {
                if (a >= n) break;
                R acc = R(0);
#pragma unroll
                for (int b = 0; b < N; ++b)
                    if (b < n) acc += ldc(s.root, a * n + b) * zr[b];
                wc[a] = acc;
            }
        }
#pragma unroll
        for (int leg = 0; leg < 2; ++leg) {
            const R sign = leg ? R(-1) : R(1);
            R z[N], w[N];
            if (mixer) {
                R wt;
                if (jacobi) {
                    wt = wst[leg];
                } else {
                    R m = R(0);
#pragma unroll
                    for (int a = 0; a < N; ++a)
                        if (a < n) m += r_exp(x[leg][a]);
                    const R y = r_log(m / static_cast<R>(n));
                    const R u = y / (s.mix[4] * r_sqrt(r_max(t0, R(1.0 / 12))));
                    wt = s.mix[0] + s.mix[1] + s.mix[2] * r_max(-u, R(0)) - s.mix[3] * r_max(u, R(0));
                    wt = r_min(r_max(wt, R(0)), R(1));
                }
                R zlast = zs[0];
#pragma unroll
                for (int a = 0; a < N; ++a)
                    if (a == n - 1) zlast = zs[a];
                const R market = s.mix[5] * sign * zlast, fr = sign * fresh;
                const R common = s.mix[6] * market + s.mix[7] * fr;
                R other[N];
#pragma unroll
                for (int a = 0; a < N; ++a) other[a] = (a == n - 1) ? s.mix[7] * market - s.mix[6] * fr : sign * zs[a];
                const R sw = r_sqrt(wt), sl = r_sqrt(R(1) - wt);
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    if (a >= n) break;
                    R acc = R(0);
#pragma unroll
                    for (int b = 0; b < N; ++b)
                        if (b < n) acc += ldc(s.low_root, a * n + b) * other[b];
                    z[a] = sw * common + sl * acc;
                }
                if (jacobi) {  // localcorr.py:96-108: dw = kappa (theta - w) dt + nu sqrt(w (1 - w)) dB, clipped
                    const R rho = s.mix[3];
                    const R zb = rho * market + r_sqrt(R(1) - rho * rho) * (sign * drive);
                    const R w = wt + s.mix[4] * (s.mix[1] - wt) * dt + s.mix[2] * r_sqrt(r_max(wt * (R(1) - wt), R(0)) * dt) * zb;
                    wst[leg] = r_min(r_max(w, R(0)), R(1));
                }
            } else {
#pragma unroll
                for (int a = 0; a < N; ++a)
                    if (a < n) z[a] = sign * zc[a];
            }
            if (has_var) {
#pragma unroll
                for (int a = 0; a < N; ++a) {
                    if (a >= n) break;
                    if (s.same) {
                        w[a] = sign * wc[a];
                    } else {
                        w[a] = sign * zr[a];
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < N; ++a) {
                if (a >= n) break;
                const int k = CV ? 0 : (TAB ? 1 : kind[a]);
                R var;
                const R vp = r_max(v[leg][a], R(0));
                if (k == 0) {
                    var = cvar[a];
                } else if (k == 2) {
                    var = vp;
                } else {
                    const int row = ldc(s.step_row, a * s.n_steps + (i - 1));
                    const R* y = s.table + static_cast<long long>(row) * s.n_x;
                    const R t = lerp_row(y, ldc(s.lo, row), ldc(s.inv_dx, row), s.n_x, x[leg][a]);
                    var = k == 1 ? t : t * vp;
                }
                if (k == 2 || k == 3) {
                    const R kappa = ldc(s.hes, a * 5 + 1), theta = ldc(s.hes, a * 5 + 2);
                    const R xi = ldc(s.hes, a * 5 + 3), rho = ldc(s.hes, a * 5 + 4);
                    const R zv = rho * z[a] + r_sqrt(R(1) - rho * rho) * w[a];
                    v[leg][a] += kappa * (theta - vp) * dt + xi * r_sqrt(vp * dt) * zv;
                }
                x[leg][a] += R(-0.5) * var * dt + r_sqrt(var * dt) * z[a];
            }
        }
        while (next_step == i) {
            observe(j++);
            load_obs(j);
        }
    };
    // Constant-variance and table forms of up to four names draw exactly N normals a step from stream 0, so four
    // steps draw N whole Philox blocks: unrolled by four, the stream's buffer position is a compile-time constant
    // at every draw (no select among the buffered normals, no refill test).
    constexpr bool kWholeBlocks = (CV || TAB) && N <= 4;
    int i = 1;
    if constexpr (kWholeBlocks) {
        for (; i + 3 <= s.n_steps; i += 4) {
            one_step(i);
            one_step(i + 1);
            one_step(i + 2);
            one_step(i + 3);
        }
    } else if constexpr (!mixer && N <= 2) {
        // the generic form of one or two names whose variance normals come from stream 0 (no SLV name) draws 2 N
        // normals a step from it: two steps are N whole blocks, and under this test the draws' source is known
        if (has_var && !s.has_slv) {
            for (; i + 1 <= s.n_steps; i += 2) {
                one_step(i);
                one_step(i + 1);
            }
        }
    }
    for (; i <= s.n_steps; ++i) one_step(i);
    if constexpr (EXER) {  // products.py:150 and :30-35: K cash D(t_tau) of each leg, the pair's average
        const ExerciseConst r = (ExerciseConst)s.ex;
        double d = 0.0;
        if (active) {
            double pay = 0.0;
#pragma unroll
            for (int l = 0; l < 2; ++l) pay += r->strike * exs.cash[l] * ldc(r->disc, exs.tau[l]);
            d = 0.5 * pay;
        }
        workgroup_pair_sums(d, d * d, r->sums);
    }
}

// ---- device copy of a request --------------------------------------------------------------------------------

struct Arena {
    std::vector<unsigned char> host;
    size_t put(const void* src, size_t bytes) {
        const size_t at = (host.size() + 255) & ~size_t(255);
        host.resize(at + bytes);
        if (bytes) std::memcpy(host.data() + at, src, bytes);
        return at;
    }
    template <typename R>
    size_t reals(const double* src, long long n) {
        std::vector<R> tmp(src ? static_cast<size_t>(n) : 0);
        for (size_t i = 0; i < tmp.size(); ++i) tmp[i] = static_cast<R>(src[i]);
        return put(tmp.data(), tmp.size() * sizeof(R));
    }
    size_t ints(const int* src, long long n) { return put(src, src ? static_cast<size_t>(n) * sizeof(int) : 0); }
};

template <typename R, int N, int FORM>
void launch_form(const PathArgs<R>& s, hipStream_t stream) {
    const unsigned grid = static_cast<unsigned>((s.n_pairs + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((paths_kernel<R, N, FORM>), dim3(grid), dim3(kBlock), 0, stream, s);
}

template <typename R, int N>
void launch_n(const PathArgs<R>& s, hipStream_t stream) {
    if constexpr (N == 1) {
        if (s.ex) {
            if (s.n_tab == 0 && !s.has_var)
                launch_form<R, 1, kFormCV | kFormExercise>(s, stream);
            else
                launch_form<R, 1, kFormGeneric | kFormExercise>(s, stream);
            return;
        }
    }
    if (s.mixer)
        launch_form<R, N, kFormMixed>(s, stream);
    else if (s.n_tab == 0 && !s.has_var)
        launch_form<R, N, kFormCV>(s, stream);
    else if (s.n_tab == s.n_names && !s.has_var)
        launch_form<R, N, kFormTable>(s, stream);
    else
        launch_form<R, N, kFormGeneric>(s, stream);
}

template <typename R>
double launch_typed(const optlab_paths_request& rq, double* d_out, double* d_var, hipStream_t stream,
                    const ExerciseRule* rule = nullptr) {
    const int n = rq.n_names;
    const long long nsn = static_cast<long long>(n) * rq.n_steps;
    Arena ar;
    const size_t o_dt = ar.reals<R>(rq.dt, rq.n_steps), o_t0 = ar.reals<R>(rq.t0, rq.n_steps);
    std::vector<ObsRec> orec(static_cast<size_t>(std::max(rq.n_obs, 1)));
    for (int j = 0; j < rq.n_obs; ++j) {
        int flags = 0;
        for (int w = 0; w < rq.n_windows && w < 2; ++w) {
            if (j < rq.win_lo[w] || j > rq.win_hi[w]) continue;
            flags |= 1 << w;
            if (j == rq.win_lo[w]) flags |= 4 << w;
            if (j == rq.win_hi[w]) flags |= 16 << w;
        }
        orec[j] = ObsRec{rq.obs_step[j], rq.col_of ? rq.col_of[j] : -1, flags, 0};
    }
    const size_t o_obs = ar.put(orec.data(), orec.size() * sizeof(ObsRec));
    const size_t o_root = ar.reals<R>(rq.root, n * n), o_kind = ar.ints(rq.kind, n);
    const size_t o_cvar = ar.reals<R>(rq.cvar, n), o_hes = ar.reals<R>(rq.heston, 5 * n);
    bool any_table = false;
    for (int a = 0; a < n; ++a) any_table |= rq.kind[a] == 1 || rq.kind[a] == 3;
    const size_t o_row = ar.ints(any_table ? rq.step_row : nullptr, nsn);
    const size_t o_lo = ar.reals<R>(any_table ? rq.row_lo : nullptr, rq.n_rows);
    const size_t o_inv = ar.reals<R>(any_table ? rq.row_inv_dx : nullptr, rq.n_rows);
    const size_t o_tab = ar.reals<R>(any_table ? rq.table : nullptr, static_cast<long long>(rq.n_rows) * rq.n_x);
    const size_t o_fwd = ar.reals<R>(rq.fwd, static_cast<long long>(n) * rq.n_obs);
    const size_t o_low = ar.reals<R>(rq.mixer ? rq.low_root : nullptr, n * n);
    size_t o_ex = 0;
    if (rule) {
        if (n != 1 || rq.mixer) invalid("an exercise rule is on one name, no mixer");
        const ExerciseDev e{rule->coefs, rule->valid, rule->disc, rule->strike, 1.0 / rule->strike, rule->phi,
                            rule->sums};
        o_ex = ar.put(&e, sizeof(e));
    }
    ar.put(nullptr, 0);
    unsigned char* dev = static_cast<unsigned char*>(scratch(kSlotSpec, std::max<size_t>(ar.host.size(), 256)));
    hip_check(hipMemcpyAsync(dev, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice, stream), "copy spec");

    PathArgs<R> s{};
    s.n_names = n, s.n_steps = rq.n_steps, s.n_obs = rq.n_obs, s.n_x = rq.n_x;
    s.n_pairs = rq.n_pairs, s.first_pair = rq.first_pair, s.rows = 2 * rq.n_pairs;
    s.key0 = static_cast<std::uint32_t>(rq.seed), s.key1 = static_cast<std::uint32_t>(rq.seed >> 32);
    auto at = [&](size_t o) { return dev + o; };
    s.dt = reinterpret_cast<const R*>(at(o_dt)), s.t0 = reinterpret_cast<const R*>(at(o_t0));
    s.orec = reinterpret_cast<const ObsRec*>(at(o_obs));
    s.root = reinterpret_cast<const R*>(at(o_root)), s.kind = reinterpret_cast<const int*>(at(o_kind));
    s.cvar = reinterpret_cast<const R*>(at(o_cvar)), s.hes = reinterpret_cast<const R*>(at(o_hes));
    s.same = rq.same;
    s.step_row = reinterpret_cast<const int*>(at(o_row));
    s.lo = reinterpret_cast<const R*>(at(o_lo)), s.inv_dx = reinterpret_cast<const R*>(at(o_inv));
    s.table = reinterpret_cast<const R*>(at(o_tab)), s.fwd = reinterpret_cast<const R*>(at(o_fwd));
    s.mixer = rq.mixer;
    for (int k = 0; k < 8; ++k) s.mix[k] = static_cast<R>(rq.mix[k]);
    s.low_root = reinterpret_cast<const R*>(at(o_low));
    s.n_cols = rq.n_cols, s.n_keep = rq.n_keep;
    s.out = d_out;
    s.out_var = d_var;
    if (rule) {
        s.out = s.out_var = nullptr;
        s.ex = reinterpret_cast<const ExerciseDev*>(at(o_ex));
    }
    s.n_tab = 0;
    s.has_var = s.has_slv = 0;
    for (int a = 0; a < n; ++a) {
        const int k = rq.kind[a];
        s.n_tab += k == 1 || k == 3;
        s.has_var |= (k == 2 || k == 3);
        s.has_slv |= k == 3;
    }

    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream), "event record");
    if (n <= 1) launch_n<R, 1>(s, stream);
    else if (n <= 2) launch_n<R, 2>(s, stream);
    else if (n == 3) launch_n<R, 3>(s, stream);
    else if (n <= 4) launch_n<R, 4>(s, stream);
    else launch_n<R, kMaxNames>(s, stream);
    hip_check(hipGetLastError(), "paths kernel launch");
    hip_check(hipEventRecord(e1, stream), "event record");
    hip_check(hipEventSynchronize(e1), "paths kernel");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

__global__ void philox_kernel(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                              unsigned* out) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const U4 w = philox10(U4{c0 + static_cast<unsigned>(i), c1, c2, c3}, k0, k1);
    out[4 * i] = w.x, out[4 * i + 1] = w.y, out[4 * i + 2] = w.z, out[4 * i + 3] = w.w;
}

template <typename R>
__global__ void normals_kernel(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned k0,
                               unsigned k1, double* out) {
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    NormalStream<R> d;
    d.init(static_cast<std::uint64_t>(first_pair + p), k0, k1, tag);
    for (int k = 0; k < n_normals; ++k) out[p * n_normals + k] = static_cast<double>(d.next());
}

}  // namespace

double launch_paths(const optlab_paths_request& rq, int prec, double* d_out, hipStream_t stream, double* d_var) {
    return prec == 0 ? launch_typed<double>(rq, d_out, d_var, stream) : launch_typed<float>(rq, d_out, d_var, stream);
}

long long exercise_workgroups(long long n_pairs) { return (n_pairs + kBlock - 1) / kBlock; }

double launch_paths_exercise(const optlab_paths_request& rq, int prec, const ExerciseRule& rule, hipStream_t stream) {
    return prec == 0 ? launch_typed<double>(rq, nullptr, nullptr, stream, &rule)
                     : launch_typed<float>(rq, nullptr, nullptr, stream, &rule);
}

void launch_philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                   unsigned* out, hipStream_t stream) {
    if (n <= 0) return;
    hipLaunchKernelGGL(philox_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, c0, c1, c2,
                       c3, k0, k1, n, out);
    hip_check(hipGetLastError(), "philox kernel launch");
}

void launch_normals(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                    int prec, double* out, hipStream_t stream) {
    if (n_pairs <= 0 || n_normals <= 0) return;
    const dim3 grid(static_cast<unsigned>((n_pairs + 255) / 256));
    const unsigned k0 = static_cast<unsigned>(seed), k1 = static_cast<unsigned>(seed >> 32);
    if (prec == 0)
        hipLaunchKernelGGL(normals_kernel<double>, grid, dim3(256), 0, stream, first_pair, n_pairs, n_normals, tag, k0,
                           k1, out);
    else
        hipLaunchKernelGGL(normals_kernel<float>, grid, dim3(256), 0, stream, first_pair, n_pairs, n_normals, tag, k0,
                           k1, out);
    hip_check(hipGetLastError(), "normals kernel launch");
}

}  // namespace optlab_mi
