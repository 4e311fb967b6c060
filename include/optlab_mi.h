/* optlab_mi.h -- C ABI of the MI355X Monte Carlo path engine (liboptlab_mi.so).
 *
 * The entry points replace the checked requests of the reference engine, `optlab::api` in
 * packages/native/cuda/include/optlab_cuda/requests.hpp:95-110, which its nanobind module binds
 * (packages/native/bindings/ext.cpp:57-178). The request structs keep the reference layouts field for
 * field (requests.hpp:12-93) so a binding written for them fills these unchanged; the additions are new
 * structs (strip, SLV particle run, adjoint, book), never new fields in the old ones.
 *
 * Conventions, all entry points:
 *   - every function returns OPTLAB_OK (0) or an error code; optlab_last_error() gives the message of the
 *     last failure on the calling thread (the reference throws std::invalid_argument / std::runtime_error,
 *     requests.cpp:140-219, which ext.cpp:43-53 turns into EngineError);
 *   - backend 0 is double precision: the numbers of the reference CPU engine (the same Philox draws and the
 *     same operations in FP64), computed on the GPU; backend 1 is the float engine of the reference GPU
 *     (FP32 path state, FP64 sums). Any other backend is OPTLAB_EINVAL;
 *   - n_threads is accepted for ABI compatibility and ignored: all work runs on the current HIP device;
 *   - buffers named "host" are caller-owned host memory, written in place; "_dev" entry points take
 *     caller-owned device pointers and an optional hipStream_t passed as void* (NULL: the null stream);
 *   - calls on one device are not thread-safe (requests.hpp:103).
 */
#ifndef OPTLAB_MI_H
#define OPTLAB_MI_H

#ifdef __cplusplus
extern "C" {
#endif

#define OPTLAB_MI_ABI_VERSION 1

#if defined(__GNUC__) || defined(__clang__)
#define OPTLAB_API __attribute__((visibility("default")))
#else
#define OPTLAB_API
#endif

enum {
    OPTLAB_OK = 0,
    OPTLAB_EINVAL = 1,   /* bad request: std::invalid_argument in the reference */
    OPTLAB_EDEVICE = 2,  /* no device, or a HIP error: std::runtime_error in the reference */
    OPTLAB_ENOMEM = 3,   /* device allocation failed */
    OPTLAB_EINTERNAL = 4
};

/* ---- fixed products: requests.hpp:12-29, unchanged layout -------------------------------------------------- */
/* model: 0 Black-Scholes (exact log step), 1 Heston (Andersen QE-M). product: 0 European, 1 Asian, 2 down-and-out
   barrier (0 < H < S0), 3 up-and-out barrier (H > S0; an addition: the reference refuses product 3).
   legs: 0 none, 1 call, 2 put, 3 forward. Outputs (params.hpp:27-32): European out[k] = leg k; Asian out[0]
   arithmetic, out[1] geometric, out[2] the arithmetic with the geometric control variate (an addition: the
   reference leaves out[2] at 0; Black-Scholes only, else 0); barrier out[0] knocked-out leg 0, out[1] same leg
   without barrier; out[3] = E[S_T] undiscounted. Pair p reads the normals of Philox stream 0 of pair p. */
typedef struct {
    int model, product;
    int legs[3];
    double strikes[3];
    double H;
    int n_steps;
    double S0, r, q, T;
    double sigma;
    double v0, kappa, theta, xi, rho;
    long long n_paths;
    unsigned long long seed;
} optlab_request;

typedef struct {
    double price[4], se[4];
    long long n_paths;
    double kernel_ms, total_ms;
} optlab_result;

/* ---- path engine: requests.hpp:32-63, unchanged layout ----------------------------------------------------- */
/* Every model of optlab (0 constant variance, 1 tabulated local variance, 2 Heston, 3 SLV = table of squared
   leverage times v+), 1 to 8 correlated names, optional local correlation. Arrays in double, row major; out has
   n_names * n_cols columns of 2 n_pairs rows, column (a, c) at out + (a * n_cols + c) * 2 n_pairs. */
typedef struct {
    int n_names, n_steps, n_obs, n_x, n_rows;
    long long n_pairs, first_pair;
    unsigned long long seed;
    const double* dt;
    const double* t0;
    const int* obs_step;
    const double* root;
    const int* kind;
    const double* cvar;
    const double* heston;
    int same;
    const int* step_row;
    const double* row_lo;
    const double* row_inv_dx;
    const double* table;
    const double* fwd;
    int mixer;
    double mix[8];
    const double* low_root;
    int n_cols, n_keep;
    const int* col_of;
    int n_windows;
    const int* win_lo;
    const int* win_hi;
    double* out;
    double kernel_ms, total_ms;
} optlab_paths_request;

/* requests.hpp:66-71, unchanged layout (sm_count = compute units, major/minor = gfx major/minor). */
typedef struct {
    char name[256];
    int major, minor, sm_count, clock_mhz;
    long long total_mem, free_mem;
    int integrated, kernel_timeout;
} optlab_device;

/* ---- SLV leverage by particles: requests.hpp:74-93, unchanged layout -------------------------------------- */
typedef struct {
    int n_steps;
    const double* dt;
    const double* t0;
    long long n_pairs;
    unsigned long long seed;
    double v0, kappa, theta, xi, rho;
    int n_x, n_bins, n_fine;
    double atm_vol, bandwidth_scale, l_min, l_max, q_lo, q_hi;
    int local_linear, positive_part;
    int n_lv;
    const double* lv_lo;
    const double* lv_inv_dx;
    const double* lv_rows;
    double* x_grids;
    double* values;
    double total_ms;
} optlab_leverage_request;

/* ---- addition: options of the particle method beyond the reference request ------------------------------- */
/* bandwidth_rule 0: the reference's bandwidth_scale * atm_vol * sqrt(max(t, 0.15)) * N^-0.2 (leverage.cpp:118);
   1: bandwidth_coef * std(ln S) * N^-0.2, the standard deviation taken over the particles at the step. */
typedef struct {
    int bandwidth_rule;
    double bandwidth_coef;
} optlab_leverage_options;

/* ---- addition: the particle method's local variance as surface slices ------------------------------------- */
/* Instead of the rows lv_rows of the request (the Dupire local variance tabulated at each step start, host data the
   reference builds in NumPy, native/__init__.py:129-144), the surface at each step's maturity given as at most two
   raw SVI slices w = c0 S(k; p0) + c1 S(k; p1), S(k; a, b, rho, m, s2) = a + b (rho (k - m) + sqrt((k - m)^2 + s2)),
   with their derivatives in T; the engine evaluates the table entries it reads (the Dupire formula with Gatheral's g,
   clipped to [floor, cap]) on the device, at k = j (hi - lo) / (n_lv - 1) + lo with lo = lv_lo[i] and the last point
   hi[i], the entries of the table the rows would hold. Row i of records: c0, c1, dc0/dT, dc1/dT, p0, p1, dp0/dT,
   dp1/dT (5 each: a, b, rho, m, s2). */
typedef struct {
    int n_rows;              /* = n_steps */
    const double* hi;        /* host, n_rows */
    const double* records;   /* host, n_rows x 24 */
    double floor, cap;       /* the clip of the local variance (the reference: 1e-4, 4) */
} optlab_lv_slices;

/* ---- addition: a strip of vanillas on one simulation (Heston QE-M or Black-Scholes) ------------------------ */
/* Calls (or puts) at n_strikes strikes and n_mats maturities t_j = T * mat_step[j] / n_steps, all on the same
   paths; price[j * (n_strikes + 1) + k] is strike k at maturity j, discounted by exp(-r t_j), and column
   n_strikes holds E[S_{t_j}] undiscounted (the martingale check). Sums of pair averages in compensated FP64. */
typedef struct {
    int model;              /* 0 Black-Scholes, 1 Heston QE-M */
    int put;                /* 0 calls, 1 puts */
    int n_steps;            /* uniform steps of T / n_steps */
    double S0, r, q, T;
    double sigma;
    double v0, kappa, theta, xi, rho;
    double psi_c;           /* QE switching level (Andersen: 1.5) */
    int n_strikes;          /* 1 to 31 */
    const double* strikes;  /* host, n_strikes */
    int n_mats;             /* 1 to 16 */
    const int* mat_step;    /* host, n_mats, increasing in [1, n_steps] */
    long long n_paths;      /* even, >= 4 */
    unsigned long long seed;
    double* price;          /* host, n_mats x (n_strikes + 1) */
    double* se;             /* host, n_mats x (n_strikes + 1) */
    double kernel_ms, total_ms;
} optlab_strip_request;

/* ---- addition: implied-vol surfaces held on the device ----------------------------------------------------- */
/* One surface on the engine's table grid: row r (one per simulation step) at maturity T[r] (>= 1/365), on the log
   forward moneyness k = lo[r] + x / inv_dx[r], x < n_x, with the total variance w and its derivatives w_k, w_kk,
   w_T there: the rows packages/risk/src/optlab/risk/surfaces.py:180-205 keeps for a surface. Uploaded once; the
   local-variance tables of the surface moved by node tents of the risk grid, a parallel shift or a skew tilt are
   then built on the device (surfaces.py:95-170) in double. */
typedef struct {
    int n_rows, n_x;
    const double* T;       /* host, n_rows */
    const double* lo;      /* host, n_rows */
    const double* inv_dx;  /* host, n_rows */
    const double* w;       /* host, n_rows x n_x */
    const double* w_k;
    const double* w_kk;
    const double* w_T;
} optlab_surface_rows;

/* The risk grid: 10 strikes in forward moneyness x 8 tenors (risk/grid.py:25-26). */
#define OPTLAB_GRID_STRIKES 10
#define OPTLAB_GRID_TENORS 8
#define OPTLAB_GRID_NODES (OPTLAB_GRID_STRIKES * OPTLAB_GRID_TENORS)

/* ---- addition: worst-of autocall (Phoenix) under Dupire local vol, with adjoint Greeks ------------------- */
/* The product of packages/structured/src/optlab/structured/compiler.py:114-161, 270-281 with a daily knock-in:
   at observation j the performance is perf_j = min_a S_a(t_j) / fixing_a; a coupon `coupon` (times 1 + missed
   with memory) is paid when perf_j >= coupon_barrier; before the last date the note redeems at par and ends when
   perf_j >= autocall; at the last date it pays 1 - ki max(put_strike - perf, 0), ki = [min over every close of
   min_a S_a / fixing_a < ki_barrier]. Every comparison is the ramp clip(d / (smoothing |b|) + 1/2, 0, 1) of
   packages/structured/src/optlab/structured/script/evaluator.py:93-108 (smoothing 0: hard). Paths: the
   engine's local-vol step (paths.hpp:274-361, kind 1) with n_names normals per step on Philox stream 0 of
   each antithetic pair, one table row per step. Greeks (optlab_adjoint): delta per name at sticky moneyness
   (the spot scales the paths) and vega per node of the risk grid per name, per vol point, at the unbumped
   surfaces, by a pathwise adjoint. Standard errors: price and deltas from pair averages, vegas by batch means
   over 16 batches. */
typedef struct {
    int n_names;             /* 1 to 5 */
    int n_steps;             /* uniform steps of dt; row s of every surface is read at step s */
    double dt;
    const int* surface;      /* host, n_names surface handles, each with n_steps rows of one n_x */
    const double* root;      /* host, n_names x n_names: correlated normals z = root . independent normals */
    const double* fwd;       /* host, n_names x (n_steps + 1): forward F_a(s dt); s = 0 is the spot */
    const double* fixing;    /* host, n_names: the levels performances are read against */
    int n_obs;               /* 1 to 64 */
    const int* obs_step;     /* host, n_obs increasing in [1, n_steps]; the last is n_steps (maturity) */
    const double* disc;      /* host, n_obs discount factors of the payments */
    double autocall, coupon_barrier, coupon;
    int memory;
    double ki_barrier;       /* 0: no knock-in (the note redeems at par) */
    double put_strike;
    double smoothing;
    long long n_pairs, first_pair;
    unsigned long long seed;
    const double* nodes;     /* host or NULL: n_names x OPTLAB_GRID_NODES node bumps (vol units); price only */
    const double* spot;      /* host or NULL: n_names factors on fwd (a spot bump at sticky moneyness) */
    double price, se;
    double batch[16];        /* the price of each of 16 batches of consecutive pairs (CRN differences) */
    double delta[5], delta_se[5];
    double* vega;            /* host or NULL: n_names x OPTLAB_GRID_NODES, per vol point (optlab_adjoint) */
    double* vega_se;         /* host or NULL: the same shape */
    double forward_ms, backward_ms, total_ms;
    int checkpoint_every;    /* optlab_adjoint: steps between checkpoints of the forward state, 4, 8 (up to 3
                                names) or 16 (one name), where the segment's rows fit in LDS; 0: the autotuner
                                times the candidates on the first adjoint of a shape and keeps the fastest */
    int checkpoint_used;     /* out: the interval the adjoint ran with (0 for a price) */
} optlab_autocall_request;

/* ---- addition: the names' models of a worst-of autocall ----------------------------------------------------- */
/* Per name: kind 0 constant variance cvar, 1 the surface's local variance (what optlab_autocall assumes for every
   name), 2 Heston, 3 SLV: spot variance lev2(x, t) max(v, 0) with v a Heston variance (Euler, full truncation) --
   the steps of paths.hpp:196-361 for these kinds, with the variance normals (n_names a step) drawn from the pair's
   aux stream when any name is SLV, else from stream 0 after the step's spot normals, and correlated by root when
   `same`. lev2 holds each SLV name's squared leverage on its surface's table grid (row s read at step s, entry j
   at x = lo[s] + j / inv_dx[s]). Greeks: deltas for every kind (sticky moneyness); node vegas for kinds 1 and 3
   (0 for the others), an SLV name's with its conditional expectation E[v | x] = LV / L^2 held at every entry, so
   that a surface bump moves lev2 to lev2 LV_bumped / LV (the move a bumped price, nodes != NULL, makes too). */
typedef struct {
    const int* kind;         /* host, n_names */
    const double* cvar;      /* host or NULL: n_names, kind 0's variance */
    const double* heston;    /* host or NULL: n_names x 5 (v0, kappa, theta, xi, rho), kinds 2 and 3 */
    const double* lev2;      /* host or NULL: n_names x n_steps x n_x, read for kind 3 names */
    int same;
} optlab_autocall_models;

/* ---- addition: a payoff script fused onto the device ----------------------------------------------------------- */
/* What optlab_script returns: the price (the mean over pairs of the pair-averaged discounted payoff) and its standard
   error, the path kernel's and the payoff kernel's times, the hipRTC compile time of the script (0 when the process
   had compiled it already on this device) and the call's wall time, in ms. */
typedef struct {
    double price, se;
    double paths_ms, payoff_ms, compile_ms, total_ms;
} optlab_script_result;

/* ---- addition: a book of trades revalued under scenarios, in one persistent launch --------------------------- */
/* Trades on the names of one book (each name: a surface handle on the book's grid of uniform steps, and its
   forwards); paths by the local-vol step of paths.hpp:274-361 on stream 0 of each trade's own seed, so every
   scenario of a trade sees the same draws. Scenario s multiplies name a's spot by 1 + spot_shock[s][a] (the paths
   scale: sticky moneyness) and moves every implied vol by vol_shift[s] plus the skew tilt skew[s]
   (structured/surfaces.py:18-50); the local-variance tables of every (scenario, name) are built on the device
   (risk/surfaces.py:141-165). Every scenario is simulated in full: the revaluation of
   packages/risk/src/optlab/risk/engine.py:600-645, the products of structured/engine.py:199-330. Strikes, barriers
   and fixings are in units of each name's base spot fwd[a][0]. */
typedef struct {
    int product;             /* 0 European call, 1 arithmetic Asian call, 2 up-and-out call on every close,
                                3 worst-of autocall (the phoenix of optlab_autocall_request, hard comparisons) */
    int n_names;             /* 1 for products 0-2; 1 to 5 for product 3 */
    int names[5];            /* indices into the book's names */
    double root[25];         /* n_names x n_names correlation root of the trade's names (row-major) */
    int n_steps;             /* maturity, in steps of the book's grid (1 to the book's n_steps) */
    int obs_every;           /* Asian fixings / autocall dates every obs_every steps; n_steps % obs_every == 0 */
    double strike;           /* products 0-2: strike; 3: put strike (performance) */
    double barrier;          /* 2: the up-and-out level; 3: the knock-in level (0: none) */
    double coupon, coupon_barrier, autocall;  /* product 3 */
    int memory;              /* product 3 */
    double quantity;         /* units held: the weight of the trade in the book's P&L */
    long long n_pairs;
    unsigned long long seed;
} optlab_trade;

typedef struct {
    int n_names, n_steps;    /* the book's grid: n_steps uniform steps of dt from the valuation date */
    double dt;
    const int* surface;      /* host, n_names surface handles with n_steps rows each, one n_x */
    const double* fwd;       /* host, n_names x (n_steps + 1) */
    const double* disc;      /* host, n_steps + 1 discount factors to each step's date */
    int n_trades;            /* >= 1 */
    const optlab_trade* trades;
    int n_scen;              /* >= 1; scenario 0 is the base when its moves are zero */
    const double* spot_shock;/* host, n_scen x n_names */
    const double* vol_shift; /* host, n_scen */
    const double* skew;      /* host, n_scen */
    int mode;                /* 0: one persistent launch drains the queue of (trade, scenario, chunk) items,
                                captured with its reset and reduction in a HIP graph; 1: one launch per trade
                                (the loop); 2: mode 0 launched directly, without the graph */
    int block;               /* threads per workgroup: 128, 256 or 512; 0: the autotuner chooses */
    int pairs_per_lane;      /* 1, 2, 4 or 8: an item is one wave's 64 x pairs_per_lane pairs; 0: the autotuner */
    double* value;           /* host, n_trades x n_scen prices */
    double* se;              /* host, n_trades x n_scen standard errors (pair averages) */
    double* finish_ms;       /* host or NULL, n_trades: mode 0, when the trade's last item finished after the
                                launch began (device clock); mode 1, the trade's own launch time */
    double tables_ms, kernel_ms, total_ms;
    long long name_steps;    /* simulated: sum over items of legs x steps x names */
    int block_used, pairs_per_lane_used;
} optlab_book_request;

/* ---- addition: an American option by Longstaff-Schwartz on the engine's paths ------------------------------- */
/* packages/core/src/optlab/core/products.py:105-169 (american_lsm): one name; at each exercise date, walking back
   from the last, the discounted continuation value of the in-the-money paths is regressed on the basis
   (1, s, s^2, s^3) in s = S / K, plus (v, v^2, s v) of the variance factor when basis_var; a path exercises where
   its exercise value beats the fitted continuation. The rule is estimated on the pairs of `paths` (pass 1) and
   applied to the next n_pairs pairs of the same seed (pass 2: first_pair + n_pairs on), so the price is a lower
   bound up to Monte Carlo error. The reference solves each regression with NumPy's lstsq; here the normal
   equations are formed and solved on the device in FP64 (sums in a fixed order, a Cholesky factorisation). A date
   with min_itm or fewer paths in the money, or whose normal matrix is not positive definite, exercises nothing.
   The price is the pair-average mean of K cash D(t_tau) (products.py:30-35), or the immediate exercise value when
   that is higher (then se = 0). */
typedef struct {
    optlab_paths_request paths; /* one name, kept columns = the exercise dates (n_keep = n_obs = n_exercise,
                                   no windows); out unused */
    double strike;
    int call;                   /* 1 call, 0 put */
    const double* disc;         /* host, n_obs discount factors D(t_j) of the exercise dates */
    double spot;                /* today's spot (the immediate exercise value) */
    int basis_var;              /* 1: add v, v^2, s v (a Heston or SLV name) */
    int min_itm;                /* regress only with more than this many paths in the money (50) */
    double price, se;
    double* coefs;              /* host or NULL: n_obs x 7 regression coefficients, NaN at dates that exercise nothing */
    double paths_ms;            /* pass 1's path kernel (device events) */
    double regression_ms;       /* the backward sweep: one launch a date (device events) */
    double apply_ms;            /* pass 2: its path kernel, which applies the rule as the paths reach each date and
                                   writes no spots (device events) */
    double total_ms;            /* the whole call, host wall clock */
} optlab_lsm_request;

/* ---- addition: the issuer's call regression on a pilot's states --------------------------------------------- */
/* packages/structured/src/optlab/structured/exercise.py:191-238 (fit) and :150-171 (Rule): the least-squares fit of
   y on the terms of a state x (n_vars variables), with z = (x - mean) / scale: polynomial terms prod_i z_i^p_i
   (powers, exponents 0 to 3, the constant included) then hinges max(z_v - k, 0) (knots), at most 25 terms. The
   reference solves with NumPy's lstsq; here the normal equations are formed in FP64 on the device (fixed-order
   sums) and solved by a Cholesky factorisation. ok = 0 when the normal matrix is not numerically positive definite,
   a pivot below 1e-13 of its diagonal (a repeated term, a hinge no state reaches): coef is then left as it was and
   the caller decides (the reference's lstsq would return a minimum-norm solution). With n_eval > 0 the fitted continuation is evaluated on a
   second set of states (exercise.py:170-171: Rule.continuation), in the same terms. Both backends compute in
   FP64. */
typedef struct {
    long long n_rows;           /* rows of the fit, >= 1 */
    int n_vars;                 /* 1 to 8 */
    const double* state;        /* host, n_rows x n_vars, row major */
    const double* y;            /* host, n_rows */
    const double* mean;         /* host, n_vars */
    const double* scale;        /* host, n_vars, nonzero */
    int n_powers;               /* polynomial terms */
    const int* powers;          /* host, n_powers x n_vars exponents in 0..3 */
    int n_knots;                /* hinge terms; n_powers + n_knots in 1..25 */
    const int* knot_var;        /* host, n_knots variable indices */
    const double* knot;         /* host, n_knots levels, in z */
    double* coef;               /* host, n_powers + n_knots (out) */
    int ok;                     /* out: 1 when solved */
    long long n_eval;           /* states to evaluate the continuation on (0: none) */
    const double* eval_state;   /* host, n_eval x n_vars */
    double* eval_out;           /* host, n_eval (out) */
    double fit_ms, eval_ms, total_ms;
} optlab_regress_request;

/* ---- entry points ------------------------------------------------------------------------------------------- */

/* requests.hpp:97-101 */
OPTLAB_API int optlab_gpu_count(void);
OPTLAB_API int optlab_has_hip(void);                 /* has_cuda(): built with the HIP backend (always 1 for this library) */
OPTLAB_API int optlab_cpu_threads(void);
OPTLAB_API int optlab_gpu_name(char* buf, int len);  /* device 0 as "name (gfxNNN, N CU)", "" without a device */
OPTLAB_API int optlab_device_info(int device, optlab_device* out);
OPTLAB_API const char* optlab_last_error(void);
OPTLAB_API int optlab_abi_version(void);

/* requests.hpp:104 (api::price). */
OPTLAB_API int optlab_price(const optlab_request* rq, int backend, int n_threads, optlab_result* out);

/* requests.hpp:106 (api::paths): host arrays in, host out. */
OPTLAB_API int optlab_paths(optlab_paths_request* rq, int backend, int n_threads);

/* Addition: api::paths with the output left on the device (d_out: n_names * n_cols * 2 n_pairs doubles); rq->out
   is not read. The specification arrays are still host memory. */
OPTLAB_API int optlab_paths_dev(optlab_paths_request* rq, int backend, double* d_out, void* stream);

/* requests.hpp:108 (api::leverage). */
OPTLAB_API int optlab_leverage(optlab_leverage_request* rq, int backend, int n_threads);

/* Addition: api::leverage with the options above (opt NULL: the reference's rule). */
OPTLAB_API int optlab_leverage_ex(optlab_leverage_request* rq, const optlab_leverage_options* opt, int backend);

/* Addition: optlab_leverage_ex with the local variance from surface slices (lv_rows may be NULL): the calibration
   the reference runs through leverage() after building the rows on the host (native/__init__.py:460-495), with the
   rows' host build and upload replaced by one record a step. opt may be NULL. */
OPTLAB_API int optlab_leverage_slices(optlab_leverage_request* rq, const optlab_leverage_options* opt,
                                      const optlab_lv_slices* lv, int backend);

/* Addition: the strip of vanillas (no counterpart in the reference's fixed products, which price three legs at
   one maturity: params.hpp:27-40). */
OPTLAB_API int optlab_strip(optlab_strip_request* rq, int backend);

/* Addition: pathwise sensitivities of every price of the strip, on the strip's paths (no counterpart in the
   reference, whose Greeks are central bumps on common random numbers: structured/greeks.py:164-313, a spot bump
   scaling the paths). delta[j * (n_strikes + 1) + k] is d price / d S0 and vega[...] the derivative in the variance
   parameter: v0 under Heston QE-M (through every step of the scheme), sigma under Black-Scholes; column n_strikes
   holds those of E[S_{t_j}]. Each with its standard error (over pairs). rq->price and rq->se are not written;
   kernel_ms and total_ms are. */
OPTLAB_API int optlab_strip_greeks(optlab_strip_request* rq, int backend, double* delta, double* delta_se,
                                   double* vega, double* vega_se);

/* Addition with no counterpart in the reference: optlab_strip for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc, hipFree or
   other allocating routine: its device memory is one 16 MiB array per device, fixed at load time and held until the
   process ends (the slot table, the totals, the chunk sums and the per-workgroup rows; the rows take what is left,
   so a strip with many maturities runs fewer workgroups, and a request whose fixed parts do not fit is rejected).
   Asynchronous on `stream` (a hipStream_t, NULL: the null stream): it neither waits nor copies to the host. The
   estimates go to price_dev and se_dev, device arrays of n_mats x (n_strikes + 1) doubles laid out as rq->price;
   rq->price, rq->se and the timings are not used. Every call on a device shares that device's array and leaves its
   contents behind: calls that overlap on one device (two streams, two threads) race. */
OPTLAB_API int optlab_strip_noalloc(const optlab_strip_request* rq, int backend, double* price_dev, double* se_dev,
                                    void* stream);

/* Addition: upload a surface's rows to the current device; *handle names it in later calls (the reference keeps
   the same rows per surface in a host cache, risk/surfaces.py:180-205). */
OPTLAB_API int optlab_surface_upload(const optlab_surface_rows* rows, int* handle);
OPTLAB_API int optlab_surface_free(int handle);

/* Addition: the local-variance table of surface `handle` moved by node bumps (OPTLAB_GRID_NODES, strike-major,
   vol units, or NULL), a parallel shift and a skew tilt (surfaces.py:141-165): out host, n_rows x n_x, double. */
OPTLAB_API int optlab_surface_table(int handle, const double* nodes, double parallel, double skew, double* out);

/* Addition: sum over entries of tbar[row][x] dLV[row][x] / dH_node at the unbumped surface, for every node of the
   risk grid: tbar host n_rows x n_x, out host OPTLAB_GRID_NODES. The contraction the adjoint Greeks end with. */
OPTLAB_API int optlab_surface_node_grad(int handle, const double* tbar, double* out);

/* Addition: the worst-of autocall above, price only (the bump-and-revalue comparator calls this; structured/
   greeks.py:164-313 revalues through the script evaluator). */
OPTLAB_API int optlab_autocall(optlab_autocall_request* rq, int backend);

/* Addition: price, deltas and node vegas by the pathwise adjoint (nodes must be NULL). */
OPTLAB_API int optlab_adjoint(optlab_autocall_request* rq, int backend);

/* Addition: optlab_autocall (adjoint 0) or optlab_adjoint (adjoint 1) with each name's model; models NULL: every
   name local vol. */
OPTLAB_API int optlab_autocall_ex(optlab_autocall_request* rq, const optlab_autocall_models* models, int adjoint,
                                  int backend);

/* Addition: a product in the payoff language priced on the engine's paths without bringing them to the host (the
   reference evaluates the script in NumPy on paths copied back: structured/script/evaluator.py, structured/
   engine.py:199-330). `source` is the HIP source the host lowers the script to (optlab.native.script), compiled at
   run time by hipRTC for the current device and cached per process and device by its text; it defines
   extern "C" __global__ void script_payoff(const double* out, long long n_pairs, double* rows), run with 256 lanes
   a workgroup over the paths of rq (rq->out is not used: the paths stay in device memory). The payoff is evaluated
   in double whatever the backend of the paths. */
OPTLAB_API int optlab_script(optlab_paths_request* rq, int backend, const char* source, optlab_script_result* out);

/* Addition: the book above (backend 0 double paths, 1 float paths; sums in double either way). */
OPTLAB_API int optlab_book(optlab_book_request* rq, int backend);

/* Addition: the American option above (backend 0 double paths, 1 float paths; the regressions in double either way).
   Replaces core/products.py:112-169 (NumPy paths, lstsq per date). */
OPTLAB_API int optlab_lsm(optlab_lsm_request* rq, int backend);

/* Addition: the issuer's call regression in FP64 normal equations (exercise.py:191-238); the reference calls
   np.linalg.lstsq on the host. */
OPTLAB_API int optlab_regress(optlab_regress_request* rq, int backend);

/* Addition, for tests of the integer stream: the Philox4x32-10 words of counters (c0 + i, c1, c2, c3), i < n,
   key (k0, k1), computed by the device function the kernels use; out_dev holds 4 n uint32 (device memory). */
OPTLAB_API int optlab_philox_dev(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                      unsigned* out_dev, void* stream);

/* Addition, for tests: the normals NormalStream (paths.hpp:213-233) gives pair p, stream tag, in the given
   precision (backend 0 double, 1 float widened to double): out_dev holds n_pairs * n_normals doubles, pair-major. */
OPTLAB_API int optlab_normals_dev(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                       int backend, double* out_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPTLAB_MI_H */
