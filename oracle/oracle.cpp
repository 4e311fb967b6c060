// CPU ORACLE -- test infrastructure only (see oracle.hpp for the map to the reference files). Double precision,
// expressions in the reference's order of operations (built with -ffp-contract=off, like the reference build in
// build_ref.sh), so the restatement and the compiled reference agree bit for bit on the same machine.
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"

namespace oracle {

void gauss_pair(std::uint32_t a, std::uint32_t b, double& z0, double& z1) {
    const double u = unit(a), v = unit(b);
    const double radius = std::sqrt(-2.0 * std::log(u));
    const double angle = 6.283185307179586 * v;
    z0 = radius * std::cos(angle);
    z1 = radius * std::sin(angle);
}

double Draws::next() {
    if (used == 4) {
        const Words r = philox(Words{static_cast<std::uint32_t>(pair), static_cast<std::uint32_t>(pair >> 32),
                                     static_cast<std::uint32_t>(block), static_cast<std::uint32_t>(block >> 32) | tag},
                               k0, k1);
        gauss_pair(r.x, r.y, buf[0], buf[1]);
        gauss_pair(r.z, r.w, buf[2], buf[3]);
        ++block;
        used = 0;
    }
    return buf[used++];
}

namespace {

thread_local std::string g_error;

struct Failure {
    int code;
    std::string what;
};

[[noreturn]] void bad(const std::string& what) { throw Failure{OPTLAB_EINVAL, what}; }

inline double dmax(double a, double b) { return a > b ? a : b; }
inline double dmin(double a, double b) { return a < b ? a : b; }

int threads_for(int n_threads, long long n_blocks) {
    int hw = static_cast<int>(std::thread::hardware_concurrency());
    int t = n_threads > 0 ? n_threads : (hw > 0 ? hw : 1);
    return static_cast<int>(std::max<long long>(1, std::min<long long>(t, n_blocks)));
}

// Runs body(block) for block in [0, n_blocks) on t threads pulling from an atomic counter (paths_cpu.cpp:30-52).
template <typename F>
void for_blocks(long long n_blocks, int n_threads, F body) {
    const int t = threads_for(n_threads, n_blocks);
    std::atomic<long long> next{0};
    auto work = [&]() {
        for (long long k; (k = next.fetch_add(1)) < n_blocks;) body(k);
    };
    if (t == 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int i = 0; i < t; ++i) pool.emplace_back(work);
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------------------------------------ path engine

constexpr int kNames = 8, kWindows = 2;

struct Spec {
    const optlab_paths_request* rq;
    long long rows;
    std::uint32_t k0, k1;
    double* var = nullptr;  // or the variance factor of name 0 at the kept columns (n_keep x rows)
};

// paths.hpp:196-205: uniform-grid lerp, u clamped to [0, n_x - 1 - 1e-4].
double table_at(const optlab_paths_request& r, int row, double x) {
    double u = (x - r.row_lo[row]) * r.row_inv_dx[row];
    const double top = static_cast<double>(r.n_x - 1) - 1e-4;
    u = dmin(dmax(u, 0.0), top);
    const int i = static_cast<int>(u);
    const double w = u - static_cast<double>(i);
    const double* y = r.table + static_cast<long long>(row) * r.n_x;
    return y[i] * (1.0 - w) + y[i + 1] * w;
}

struct Extremes {
    double lo[2][kNames][kWindows], hi[2][kNames][kWindows];
};

// paths.hpp:241-272: spots of both legs at observation j, window extremes at the window's last observation.
void observe(const Spec& s, long long p, int j, const double* x0, const double* x1, Extremes& e) {
    const optlab_paths_request& r = *s.rq;
    const long long row[2] = {p, r.n_pairs + p};
    const int c = r.col_of[j];
    for (int a = 0; a < r.n_names; ++a) {
        const double f = r.fwd[a * r.n_obs + j];
        const double S[2] = {f * std::exp(x0[a]), f * std::exp(x1[a])};
        double* base = r.out + static_cast<long long>(a) * r.n_cols * s.rows;
        if (c >= 0) {
            base[c * s.rows + row[0]] = S[0];
            base[c * s.rows + row[1]] = S[1];
        }
        for (int w = 0; w < r.n_windows; ++w) {
            if (j < r.win_lo[w] || j > r.win_hi[w]) continue;
            for (int leg = 0; leg < 2; ++leg) {
                if (j == r.win_lo[w]) {
                    e.lo[leg][a][w] = e.hi[leg][a][w] = S[leg];
                } else {
                    e.lo[leg][a][w] = dmin(e.lo[leg][a][w], S[leg]);
                    e.hi[leg][a][w] = dmax(e.hi[leg][a][w], S[leg]);
                }
            }
            if (j == r.win_hi[w]) {
                const long long cl = r.n_keep + 2 * w, ch = cl + 1;
                for (int leg = 0; leg < 2; ++leg) {
                    base[cl * s.rows + row[leg]] = e.lo[leg][a][w];
                    base[ch * s.rows + row[leg]] = e.hi[leg][a][w];
                }
            }
        }
    }
}

// paths.hpp:274-361: both legs of local pair p through every step.
void pair_paths(const Spec& s, long long p) {
    const optlab_paths_request& r = *s.rq;
    const int n = r.n_names;
    const std::uint64_t pair = static_cast<std::uint64_t>(r.first_pair + p);
    double x[2][kNames], v[2][kNames];
    bool has_var = false, has_slv = false;
    for (int a = 0; a < n; ++a) {
        const int k = r.kind[a];
        const bool hv = k == 2 || k == 3;
        has_var |= hv;
        has_slv |= k == 3;
        x[0][a] = x[1][a] = 0.0;
        v[0][a] = v[1][a] = hv ? r.heston[a * 5] : 0.0;
    }
    Extremes ext;
    auto obs = [&](int jj) {
        const int c = r.col_of[jj];
        if (s.var && c >= 0) s.var[c * s.rows + p] = v[0][0], s.var[c * s.rows + r.n_pairs + p] = v[1][0];
        observe(s, p, jj, x[0], x[1], ext);
    };
    int j = 0;
    while (j < r.n_obs && r.obs_step[j] == 0) obs(j++);
    double zs[kNames], zr[kNames], fresh = 0.0, drive = 0.0;
    Draws main{pair, s.k0, s.k1};
    Draws aux{pair, s.k0, s.k1, kAuxTag};
    Draws drv{pair, s.k0, s.k1, kMixTag};
    double wst[2] = {r.mix[0], r.mix[0]};  // the stochastic weight of each leg (mixer 2)
    for (int i = 1; i <= r.n_steps; ++i) {
        const double dt = r.dt[i - 1], t0 = r.t0[i - 1];
        for (int a = 0; a < n; ++a) zs[a] = main.next();
        if (has_var)
            for (int a = 0; a < n; ++a) zr[a] = has_slv ? aux.next() : main.next();
        if (r.mixer) fresh = main.next();
        if (r.mixer == 2) drive = drv.next();
        for (int leg = 0; leg < 2; ++leg) {
            const double sign = leg ? -1.0 : 1.0;
            double z[kNames], w[kNames];
            if (r.mixer) {
                double wt;
                if (r.mixer == 2) {  // localcorr.py:87-105: the weight is the state
                    wt = wst[leg];
                } else {
                    double m = 0.0;
                    for (int a = 0; a < n; ++a) m += std::exp(x[leg][a]);
                    const double y = std::log(m / static_cast<double>(n));
                    const double u = y / (r.mix[4] * std::sqrt(dmax(t0, 1.0 / 12)));
                    wt = r.mix[0] + r.mix[1] + r.mix[2] * dmax(-u, 0.0) - r.mix[3] * dmax(u, 0.0);
                    wt = dmin(dmax(wt, 0.0), 1.0);
                }
                const double market = r.mix[5] * sign * zs[n - 1], fr = sign * fresh;
                const double common = r.mix[6] * market + r.mix[7] * fr;
                double other[kNames];
                for (int a = 0; a < n; ++a) other[a] = sign * zs[a];
                other[n - 1] = r.mix[7] * market - r.mix[6] * fr;
                const double sw = std::sqrt(wt), sl = std::sqrt(1.0 - wt);
                for (int a = 0; a < n; ++a) {
                    double acc = 0.0;
                    for (int b = 0; b < n; ++b) acc += r.low_root[a * n + b] * other[b];
                    z[a] = sw * common + sl * acc;
                }
                if (r.mixer == 2) {
                    // localcorr.py:99-108 (advance), called after the step's normals are mixed
                    // (multiasset.py:122-124): z_B = rho market + sqrt(1 - rho^2) e, e the driver's (antithetic)
                    // normal; w += kappa (theta - w) dt + nu sqrt(max(w (1 - w), 0) dt) z_B, clipped to [0, 1].
                    const double rho = r.mix[3];
                    const double zb = rho * market + std::sqrt(1.0 - rho * rho) * (sign * drive);
                    const double w =
                        wt + r.mix[4] * (r.mix[1] - wt) * dt + r.mix[2] * std::sqrt(dmax(wt * (1.0 - wt), 0.0) * dt) * zb;
                    wst[leg] = dmin(dmax(w, 0.0), 1.0);
                }
            } else {
                for (int a = 0; a < n; ++a) {
                    double acc = 0.0;
                    for (int b = 0; b < n; ++b) acc += r.root[a * n + b] * zs[b];
                    z[a] = sign * acc;
                }
            }
            if (has_var) {
                for (int a = 0; a < n; ++a) {
                    if (r.same) {
                        double acc = 0.0;
                        for (int b = 0; b < n; ++b) acc += r.root[a * n + b] * zr[b];
                        w[a] = sign * acc;
                    } else {
                        w[a] = sign * zr[a];
                    }
                }
            }
            for (int a = 0; a < n; ++a) {
                const int k = r.kind[a];
                double var;
                const double vp = dmax(v[leg][a], 0.0);
                if (k == 0) {
                    var = r.cvar[a];
                } else if (k == 2) {
                    var = vp;
                } else {
                    const double t = table_at(r, r.step_row[a * r.n_steps + (i - 1)], x[leg][a]);
                    var = k == 1 ? t : t * vp;
                }
                if (k == 2 || k == 3) {
                    const double* h = r.heston + a * 5;
                    const double zv = h[4] * z[a] + std::sqrt(1.0 - h[4] * h[4]) * w[a];
                    v[leg][a] += h[1] * (h[2] - vp) * dt + h[3] * std::sqrt(vp * dt) * zv;
                }
                x[leg][a] += -0.5 * var * dt + std::sqrt(var * dt) * z[a];
            }
        }
        while (j < r.n_obs && r.obs_step[j] == i) obs(j++);
    }
}

// requests.cpp:168-219: the checks of a path request.
void check_paths(const optlab_paths_request& c) {
    if (c.n_names < 1 || c.n_names > kNames) bad("n_names must be 1 to 8");
    if (c.n_pairs < 1 || c.first_pair < 0) bad("n_pairs must be >= 1 and first_pair >= 0");
    if (c.n_steps < 0 || c.n_obs < 0) bad("n_steps and n_obs must be >= 0");
    if (!c.out && c.n_obs > 0) bad("null output");
    if (c.n_steps > 0 && (!c.dt || !c.t0)) bad("null time grid");
    if (!c.root || !c.kind || !c.cvar || !c.heston || !c.fwd || (c.n_obs && !c.obs_step)) bad("null array");
    bool table = false;
    for (int a = 0; a < c.n_names; ++a) {
        const int k = c.kind[a];
        if (k < 0 || k > 3) bad("kind must be 0 to 3");
        table |= k == 1 || k == 3;
        if ((k == 2 || k == 3) && !(c.heston[a * 5 + 1] > 0 && c.heston[a * 5 + 2] > 0 && c.heston[a * 5 + 3] >= 0 &&
                                   std::fabs(c.heston[a * 5 + 4]) <= 1))
            bad("Heston parameters out of range");
    }
    if (table) {
        if (c.n_x < 2 || c.n_rows < 1 || !c.step_row || !c.row_lo || !c.row_inv_dx || !c.table)
            bad("a tabulated variance needs n_x >= 2, rows and step_row");
        for (long long i = 0; i < static_cast<long long>(c.n_names) * c.n_steps; ++i) {
            const int k = c.kind[i / (c.n_steps ? c.n_steps : 1)];
            if ((k == 1 || k == 3) && (c.step_row[i] < 0 || c.step_row[i] >= c.n_rows)) bad("step_row out of range");
        }
    }
    int prev = 0;
    for (int j = 0; j < c.n_obs; ++j) {
        if (c.obs_step[j] < prev || c.obs_step[j] > c.n_steps) bad("obs_step must be nondecreasing in [0, n_steps]");
        prev = c.obs_step[j];
    }
    if (c.mixer < 0 || c.mixer > 2 || (c.mixer == 1 && (!c.low_root || !(c.mix[4] > 0)))) bad("bad local correlation");
    if (c.mixer == 2 && (!c.low_root || !(c.mix[0] >= 0 && c.mix[0] <= 1) || !(c.mix[1] >= 0 && c.mix[1] <= 1) ||
                         !(c.mix[2] >= 0) || !(std::fabs(c.mix[3]) <= 1) || !(c.mix[4] >= 0)))
        bad("bad stochastic mix");
    if (c.n_windows < 0 || c.n_windows > kWindows || (c.n_windows && (!c.win_lo || !c.win_hi)))
        bad("n_windows must be 0 to 2, with win_lo and win_hi");
    if (c.n_obs && !c.col_of) bad("null col_of");
    if (c.n_keep < 0 || c.n_cols != c.n_keep + 2 * c.n_windows) bad("n_cols must be n_keep + 2 n_windows");
    for (int j = 0; j < c.n_obs; ++j)
        if (c.col_of[j] < -1 || c.col_of[j] >= c.n_keep) bad("col_of out of range");
    for (int w = 0; w < c.n_windows; ++w)
        if (c.win_lo[w] < 0 || c.win_hi[w] < c.win_lo[w] || c.win_hi[w] >= c.n_obs) bad("bad window");
}

// -------------------------------------------------------------------------------------------- fixed products

struct Consts {  // schemes.hpp:159-207 in double
    double S0, D, bs_drift, bs_vol, ln_h, K[3];
    int leg[3], n_steps;
    double inv_n;
    double v0, theta, mu_dt, E, c1, c2, K1, K2, K3, K4, A, K0p, psi_c;
};

Consts make_consts(int n_steps, double S0, double r, double q, double T, double sigma, double v0, double kappa,
                   double theta, double xi, double rho, double H, const int* legs, const double* strikes) {
    const double dt = T / n_steps;
    Consts c{};
    c.S0 = S0;
    c.D = std::exp(-r * T);
    c.bs_drift = (r - q - 0.5 * sigma * sigma) * dt;
    c.bs_vol = sigma * std::sqrt(dt);
    c.ln_h = H > 0 ? std::log(H / S0) : 0.0;
    for (int k = 0; k < 3; ++k) {
        c.K[k] = strikes ? strikes[k] : 0.0;
        c.leg[k] = legs ? legs[k] : 0;
    }
    c.n_steps = n_steps;
    c.inv_n = 1.0 / n_steps;
    const double E = std::exp(-kappa * dt), xi2 = xi * xi;
    const double k12 = 0.5 * dt * (kappa * rho / xi - 0.5);
    c.v0 = v0;
    c.theta = theta;
    c.mu_dt = (r - q) * dt;
    c.E = E;
    c.c1 = xi2 * E * (1 - E) / kappa;
    c.c2 = theta * xi2 * (1 - E) * (1 - E) / (2 * kappa);
    c.K1 = k12 - rho / xi;
    c.K2 = k12 + rho / xi;
    c.K3 = 0.5 * dt * (1 - rho * rho);
    c.K4 = c.K3;
    c.A = (k12 + rho / xi) + 0.25 * dt * (1 - rho * rho);
    c.K0p = -rho * kappa * theta * dt / xi;
    c.psi_c = 1.5;
    return c;
}

double leg_payoff(int leg, double S, double K) {  // schemes.hpp:209-217
    switch (leg) {
        case 1: return dmax(S - K, 0.0);
        case 2: return dmax(K - S, 0.0);
        case 3: return S - K;
        default: return 0.0;
    }
}

double ncdf(double x) { return 0.5 * std::erfc(-x * 0.70710678118654752); }

// schemes.hpp:219-252: Andersen QE-M, log of the martingale correction taken once after the branch.
void qe_step(const Consts& c, double& y, double& v, double zv, double zx) {
    const double m = c.theta + (v - c.theta) * c.E;
    const double psi = (v * c.c1 + c.c2) / (m * m);
    double vn, lin, w, arg;
    bool finite;
    if (psi <= c.psi_c) {
        const double inv = 2.0 / psi;
        const double b2 = inv - 1.0 + std::sqrt(inv) * std::sqrt(dmax(inv - 1.0, 0.0));
        const double a = m / (1.0 + b2), b = std::sqrt(b2);
        vn = a * (b + zv) * (b + zv);
        arg = 1.0 - 2.0 * c.A * a;
        finite = c.A < 1.0 / (2.0 * a);
        lin = -c.A * b2 * a / arg;
        w = 0.5;
    } else {
        const double p = (psi - 1.0) / (psi + 1.0), beta = (1.0 - p) / m;
        const double one_minus_u = ncdf(-zv);
        vn = one_minus_u >= 1.0 - p ? 0.0 : std::log((1.0 - p) / one_minus_u) / beta;
        arg = p + beta * (1.0 - p) / (beta - c.A);
        finite = c.A < beta;
        lin = 0.0;
        w = -1.0;
    }
    const double K0 = finite ? lin + w * std::log(arg) - (c.K1 + 0.5 * c.K3) * v : c.K0p;
    y += c.mu_dt + K0 + c.K1 * v + c.K2 * vn + std::sqrt(dmax(c.K3 * v + c.K4 * vn, 0.0)) * zx;
    v = vn;
}

constexpr int kFixedOut = 5;  // the four outputs of params.hpp:27-32 and the arithmetic x geometric product

// schemes.hpp:254-311: the pair averages of one unit. model 0/1, product 0 European, 1 Asian, 2 down-and-out,
// 3 up-and-out; out[4] = out[0] * out[1] for the Asian control variate.
void fixed_unit(const Consts& c, int model, int product, Draws& src, double out[kFixedOut]) {
    double o[4] = {0, 0, 0, 0};
    if (model == 0 && product == 0) {
        const double z = src.next();
        for (int a = 0; a < 2; ++a) {
            const double S = c.S0 * std::exp(c.bs_drift + c.bs_vol * (a == 0 ? z : -z));
            for (int k = 0; k < 3; ++k) o[k] += c.D * leg_payoff(c.leg[k], S, c.K[k]);
            o[3] += S;
        }
    } else {
        double y[2] = {0, 0}, v[2] = {c.v0, c.v0}, sum_s[2] = {0, 0}, sum_y[2] = {0, 0};
        bool alive[2] = {true, true};
        for (int i = 0; i < c.n_steps; ++i) {
            double zv = 0, zx = 0, z = 0;
            if (model == 1) {
                zv = src.next();
                zx = src.next();
            } else {
                z = src.next();
            }
            for (int a = 0; a < 2; ++a) {
                const double sg = a == 0 ? 1.0 : -1.0;
                if (model == 1)
                    qe_step(c, y[a], v[a], sg * zv, sg * zx);
                else
                    y[a] += c.bs_drift + c.bs_vol * sg * z;
                if (product == 1) {
                    sum_s[a] += std::exp(y[a]);
                    sum_y[a] += y[a];
                }
                if (product == 2) alive[a] = alive[a] && !(y[a] <= c.ln_h);
                if (product == 3) alive[a] = alive[a] && !(y[a] >= c.ln_h);
            }
        }
        for (int a = 0; a < 2; ++a) {
            const double S = c.S0 * std::exp(y[a]);
            if (product == 0) {
                for (int k = 0; k < 3; ++k) o[k] += c.D * leg_payoff(c.leg[k], S, c.K[k]);
            } else if (product == 1) {
                o[0] += c.D * leg_payoff(c.leg[0], c.S0 * sum_s[a] * c.inv_n, c.K[0]);
                o[1] += c.D * leg_payoff(c.leg[0], c.S0 * std::exp(sum_y[a] * c.inv_n), c.K[0]);
            } else {
                const double pay = c.D * leg_payoff(c.leg[0], S, c.K[0]);
                o[0] += alive[a] ? pay : 0.0;
                o[1] += pay;
            }
            o[3] += S;
        }
    }
    for (int k = 0; k < 4; ++k) out[k] = 0.5 * o[k];
    out[4] = out[0] * out[1];
}

struct Moments {
    double s = 0.0, s2 = 0.0;
};

void estimate(double sum, double sumsq, long long n_pairs, double& price, double& se) {  // cpu_mc.cpp:14-18
    const double n = static_cast<double>(n_pairs), mean = sum / n;
    const double var = sumsq / n - mean * mean;
    price = mean;
    se = n_pairs > 1 ? std::sqrt(std::fmax(var, 0.0) / (n - 1.0)) : 0.0;
}

double norm_cdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }

}  // namespace

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

namespace {

void check_request(const optlab_request& c) {  // requests.cpp:140-166, with product 3 (up-and-out) added
    if (c.model < 0 || c.model > 1) bad("model must be 0 (Black-Scholes) or 1 (Heston)");
    if (c.product < 0 || c.product > 3) bad("product must be 0, 1, 2 or 3");
    if (c.n_steps < 1) bad("n_steps must be >= 1");
    if (c.n_paths < 4 || c.n_paths % 2) bad("n_paths must be even and >= 4");
    if (!(c.S0 > 0 && c.T > 0)) bad("S0 and T must be positive");
    if (c.model == 0 && !(c.sigma > 0)) bad("sigma must be positive");
    if (c.model == 1 && !(c.v0 >= 0 && c.kappa > 0 && c.theta > 0 && c.xi > 0 && std::fabs(c.rho) <= 1))
        bad("Heston parameters out of range");
    if (c.product == 2 && !(c.H > 0 && c.H < c.S0)) bad("barrier: need 0 < H < S0");
    if (c.product == 3 && !(c.H > c.S0)) bad("up-and-out barrier: need H > S0");
    for (int k = 0; k < 3; ++k)
        if (c.legs[k] < 0 || c.legs[k] > 3) bad("legs must be 0 to 3");
}

// ------------------------------------------------------------------------------------------------- leverage

double interp_uniform(double x, double a, double b, const double* y, int n) {  // leverage.cpp:17-24
    double u = (x - a) / (b - a) * (n - 1);
    if (!(u > 0)) return y[0];
    if (u >= n - 1) return y[n - 1];
    const int i = static_cast<int>(u);
    const double w = u - i;
    return y[i] * (1 - w) + y[i + 1] * w;
}

struct Bins {
    std::vector<long long> fine;
    std::vector<double> acc;  // 5 x n_bins: count, v, x, x^2, x v
    double v_sum = 0.0;
};

double quantile(const Bins& b, double lo, double hi, double q) {  // leverage.cpp:26-38
    long long total = 0;
    for (long long c : b.fine) total += c;
    const double target = q * static_cast<double>(total);
    const double width = (hi - lo) / static_cast<double>(b.fine.size());
    double cum = 0;
    for (size_t k = 0; k < b.fine.size(); ++k) {
        const double next = cum + static_cast<double>(b.fine[k]);
        if (next >= target && b.fine[k] > 0) return lo + (static_cast<double>(k) + (target - cum) / b.fine[k]) * width;
        cum = next;
    }
    return hi;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

void oracle_philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                   unsigned* out) {
    for (long long i = 0; i < n; ++i) {
        const Words w = philox(Words{c0 + static_cast<unsigned>(i), c1, c2, c3}, k0, k1);
        out[4 * i] = w.x, out[4 * i + 1] = w.y, out[4 * i + 2] = w.z, out[4 * i + 3] = w.w;
    }
}

void oracle_normals(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                    double* out) {
    for (long long p = 0; p < n_pairs; ++p) {
        Draws d{static_cast<std::uint64_t>(first_pair + p), static_cast<std::uint32_t>(seed),
                static_cast<std::uint32_t>(seed >> 32), tag};
        for (int k = 0; k < n_normals; ++k) out[p * n_normals + k] = d.next();
    }
}

int oracle_paths(optlab_paths_request* rq, int n_threads) { return oracle_paths_var(rq, n_threads, nullptr); }

int oracle_paths_var(optlab_paths_request* rq, int n_threads, double* var) {
    try {
        if (!rq) bad("null request");
        check_paths(*rq);
        Spec s{rq, 2 * rq->n_pairs, static_cast<std::uint32_t>(rq->seed), static_cast<std::uint32_t>(rq->seed >> 32),
               var};
        const long long block = 512, n_blocks = (rq->n_pairs + block - 1) / block;
        for_blocks(n_blocks, n_threads, [&](long long k) {
            const long long a = k * block, b = std::min(rq->n_pairs, a + block);
            for (long long p = a; p < b; ++p) pair_paths(s, p);
        });
        rq->kernel_ms = rq->total_ms = 0.0;
        return OPTLAB_OK;
    } catch (const Failure& f) {
        g_error = f.what;
        return f.code;
    }
}

int oracle_price(const optlab_request* rq, int n_threads, optlab_result* out) {
    try {
        if (!rq || !out) bad("null request");
        const optlab_request& c = *rq;
        check_request(c);
        const Consts k = make_consts(c.n_steps, c.S0, c.r, c.q, c.T, c.sigma, c.v0, c.kappa, c.theta, c.xi, c.rho,
                                     c.H, c.legs, c.strikes);
        const long long n_pairs = c.n_paths / 2, block = 512, n_blocks = (n_pairs + block - 1) / block;
        std::vector<Moments> part(static_cast<size_t>(n_blocks) * kFixedOut);
        const std::uint32_t k0 = static_cast<std::uint32_t>(c.seed), k1 = static_cast<std::uint32_t>(c.seed >> 32);
        for_blocks(n_blocks, n_threads, [&](long long b) {
            Moments m[kFixedOut];
            double y[kFixedOut];
            for (long long p = b * block; p < std::min(n_pairs, (b + 1) * block); ++p) {
                Draws src{static_cast<std::uint64_t>(p), k0, k1};
                fixed_unit(k, c.model, c.product, src, y);
                for (int o = 0; o < kFixedOut; ++o) m[o].s += y[o], m[o].s2 += y[o] * y[o];
            }
            for (int o = 0; o < kFixedOut; ++o) part[b * kFixedOut + o] = m[o];
        });
        Moments tot[kFixedOut];
        for (long long b = 0; b < n_blocks; ++b)
            for (int o = 0; o < kFixedOut; ++o) tot[o].s += part[b * kFixedOut + o].s, tot[o].s2 += part[b * kFixedOut + o].s2;
        std::memset(out, 0, sizeof *out);
        for (int o = 0; o < 4; ++o) estimate(tot[o].s, tot[o].s2, n_pairs, out->price[o], out->se[o]);
        if (c.product == 1) {
            out->price[2] = out->se[2] = 0.0;
            if (c.model == 0 && (c.legs[0] == 1 || c.legs[0] == 2)) {
                // arithmetic with the geometric control variate: beta from the pair averages' covariance
                const double n = static_cast<double>(n_pairs);
                const double ma = tot[0].s / n, mg = tot[1].s / n;
                const double va = tot[0].s2 / n - ma * ma, vg = tot[1].s2 / n - mg * mg;
                const double cov = tot[4].s / n - ma * mg;
                const double beta = vg > 0 ? cov / vg : 0.0;
                const double exact = geometric_asian(c.S0, c.strikes[0], c.T, c.r, c.q, c.sigma, c.n_steps, c.legs[0] == 1);
                out->price[2] = ma - beta * (mg - exact);
                const double vr = va - 2.0 * beta * cov + beta * beta * vg;
                out->se[2] = std::sqrt(std::fmax(vr, 0.0) / (n - 1.0));
            }
        }
        out->n_paths = c.n_paths;
        return OPTLAB_OK;
    } catch (const Failure& f) {
        g_error = f.what;
        return f.code;
    }
}

int oracle_strip(optlab_strip_request* rq, int n_threads) {
    try {
        if (!rq) bad("null request");
        const optlab_strip_request& c = *rq;
        if (c.model < 0 || c.model > 1) bad("model must be 0 (Black-Scholes) or 1 (Heston)");
        if (c.n_steps < 1) bad("n_steps must be >= 1");
        if (c.n_paths < 4 || c.n_paths % 2) bad("n_paths must be even and >= 4");
        if (!(c.S0 > 0 && c.T > 0)) bad("S0 and T must be positive");
        if (c.model == 0 && !(c.sigma > 0)) bad("sigma must be positive");
        if (c.model == 1 && !(c.v0 >= 0 && c.kappa > 0 && c.theta > 0 && c.xi > 0 && std::fabs(c.rho) <= 1))
            bad("Heston parameters out of range");
        if (c.n_strikes < 1 || c.n_strikes > 31 || !c.strikes) bad("n_strikes must be 1 to 31");
        if (c.n_mats < 1 || c.n_mats > 64 || !c.mat_step) bad("n_mats must be 1 to 64");
        for (int j = 0; j < c.n_mats; ++j)
            if (c.mat_step[j] < 1 || c.mat_step[j] > c.n_steps || (j && c.mat_step[j] <= c.mat_step[j - 1]))
                bad("mat_step must increase within [1, n_steps]");
        if (!(c.psi_c >= 1.0 && c.psi_c <= 2.0)) bad("psi_c must be in [1, 2]");
        if (!c.price || !c.se) bad("null output");
        Consts k = make_consts(c.n_steps, c.S0, c.r, c.q, c.T, c.sigma, c.v0, c.kappa, c.theta, c.xi, c.rho, 0.0,
                               nullptr, nullptr);
        k.psi_c = c.psi_c;
        const int nk = c.n_strikes, nm = c.n_mats, width = nk + 1;
        std::vector<double> disc(nm);
        for (int j = 0; j < nm; ++j) disc[j] = std::exp(-c.r * (c.T * c.mat_step[j] / c.n_steps));
        const long long n_pairs = c.n_paths / 2, block = 512, n_blocks = (n_pairs + block - 1) / block;
        const size_t n_out = static_cast<size_t>(nm) * width;
        std::vector<Moments> part(static_cast<size_t>(n_blocks) * n_out);
        const std::uint32_t k0 = static_cast<std::uint32_t>(c.seed), k1 = static_cast<std::uint32_t>(c.seed >> 32);
        for_blocks(n_blocks, n_threads, [&](long long b) {
            Moments* m = &part[static_cast<size_t>(b) * n_out];
            for (long long p = b * block; p < std::min(n_pairs, (b + 1) * block); ++p) {
                Draws src{static_cast<std::uint64_t>(p), k0, k1};
                double y[2] = {0, 0}, v[2] = {k.v0, k.v0};
                int j = 0;
                for (int i = 1; i <= c.n_steps; ++i) {
                    double zv = 0, zx = 0, z = 0;
                    if (c.model == 1) {
                        zv = src.next();
                        zx = src.next();
                    } else {
                        z = src.next();
                    }
                    for (int a = 0; a < 2; ++a) {
                        const double sg = a == 0 ? 1.0 : -1.0;
                        if (c.model == 1)
                            qe_step(k, y[a], v[a], sg * zv, sg * zx);
                        else
                            y[a] += k.bs_drift + k.bs_vol * sg * z;
                    }
                    if (j < nm && c.mat_step[j] == i) {
                        const double S[2] = {k.S0 * std::exp(y[0]), k.S0 * std::exp(y[1])};
                        for (int q = 0; q < nk; ++q) {
                            const double K = c.strikes[q];
                            double o = 0.0;
                            for (int a = 0; a < 2; ++a)
                                o += disc[j] * (c.put ? dmax(K - S[a], 0.0) : dmax(S[a] - K, 0.0));
                            const double pa = 0.5 * o;
                            m[j * width + q].s += pa, m[j * width + q].s2 += pa * pa;
                        }
                        const double fa = 0.5 * (S[0] + S[1]);
                        m[j * width + nk].s += fa, m[j * width + nk].s2 += fa * fa;
                        ++j;
                    }
                }
            }
        });
        for (size_t o = 0; o < n_out; ++o) {
            Moments t;
            for (long long b = 0; b < n_blocks; ++b) t.s += part[b * n_out + o].s, t.s2 += part[b * n_out + o].s2;
            estimate(t.s, t.s2, n_pairs, rq->price[o], rq->se[o]);
        }
        rq->kernel_ms = rq->total_ms = 0.0;
        return OPTLAB_OK;
    } catch (const Failure& f) {
        g_error = f.what;
        return f.code;
    }
}

int oracle_leverage(optlab_leverage_request* rq, int n_threads, int n_parts) {
    return oracle_leverage_ex(rq, nullptr, n_threads, n_parts);
}

int oracle_leverage_ex(optlab_leverage_request* rq, const optlab_leverage_options* opt, int n_threads, int n_parts) {
    return oracle_leverage_trace(rq, opt, n_threads, n_parts, nullptr, nullptr, -1, nullptr);
}

int oracle_leverage_trace(optlab_leverage_request* rq, const optlab_leverage_options* opt, int n_threads,
                          int n_parts, double* bw_out, double* std_out, int x_step, double* x_out) {
    try {
        const int rule = opt ? opt->bandwidth_rule : 0;
        if (rule < 0 || rule > 1) bad("bandwidth_rule must be 0 or 1");
        if (rule == 1 && !(opt->bandwidth_coef > 0)) bad("bandwidth_coef must be positive");
        if (!rq) bad("null request");
        const optlab_leverage_request& s = *rq;
        if (!s.dt || !s.t0 || !s.lv_lo || !s.lv_inv_dx || !s.lv_rows || !s.x_grids || !s.values) bad("null pointer");
        if (s.n_steps < 1 || s.n_pairs < 2 || s.n_x < 2 || s.n_bins < 2 || s.n_fine < 16 || s.n_lv < 2)
            bad("n_steps, n_pairs, n_x, n_bins, n_fine or n_lv too small");
        if (!(s.kappa > 0 && s.theta > 0 && s.xi >= 0 && std::fabs(s.rho) <= 1 && s.v0 >= 0))
            bad("Heston parameters out of range");
        const int hw = static_cast<int>(std::thread::hardware_concurrency());
        const int threads = n_threads > 0 ? n_threads : (hw > 0 ? hw : 1);
        const int parts = n_parts > 0 ? n_parts : threads;
        const long long np = s.n_pairs;
        const size_t n = static_cast<size_t>(2 * np);
        std::vector<double> X(n, 0.0), V(n, s.v0);
        const int nx = s.n_x, nb = s.n_bins;
        const double n_particles = 2.0 * static_cast<double>(np);
        std::vector<double> x_eval(nx), cond(nx), lev(nx), centres(nb);
        Bins bins;
        bins.fine.assign(s.n_fine, 0);
        bins.acc.assign(5 * static_cast<size_t>(nb), 0.0);
        const std::uint32_t k0 = static_cast<std::uint32_t>(s.seed), k1 = static_cast<std::uint32_t>(s.seed >> 32);
        const double rc = std::sqrt(1 - s.rho * s.rho);
        double lo = 0, hi = 0;
        for (int i = 0; i < s.n_steps; ++i) {
            const double t0 = s.t0[i], h = s.dt[i];
            if (x_out && i == x_step) std::copy(X.begin(), X.end(), x_out);
            if (bw_out) bw_out[i] = 0.0;
            if (std_out) std_out[i] = 0.0;
            if (i == 0) {
                for (int g = 0; g < nx; ++g)
                    x_eval[g] = (-6.0 + 12.0 * g / (nx - 1)) * s.atm_vol * std::sqrt(std::max(h, 1e-4)), cond[g] = s.v0;
            } else {
                // leverage.cpp:164-187: contiguous parts summed in order, then added part by part
                const int nf = s.n_fine;
                const double wf = (hi - lo) / nf, we = (hi + 1e-12 - lo) / nb;
                std::vector<Bins> part(parts);
                const size_t per = (n + parts - 1) / parts;
                for_blocks(parts, threads, [&](long long t) {
                    Bins& pb = part[t];
                    pb.fine.assign(nf, 0);
                    pb.acc.assign(5 * static_cast<size_t>(nb), 0.0);
                    const size_t a = t * per, b = std::min(n, a + per);
                    for (size_t j = a; j < b; ++j) {
                        const double x = X[j], v = s.positive_part ? std::max(V[j], 0.0) : V[j];
                        int f = wf > 0 ? static_cast<int>((x - lo) / wf) : 0;
                        pb.fine[std::min(std::max(f, 0), nf - 1)] += 1;
                        int k = static_cast<int>((x - lo) / we);
                        k = std::min(std::max(k, 0), nb - 1);
                        pb.acc[k] += 1, pb.acc[nb + k] += v, pb.acc[2 * nb + k] += x, pb.acc[3 * nb + k] += x * x;
                        pb.acc[4 * nb + k] += x * v;
                        pb.v_sum += v;
                    }
                });
                std::fill(bins.fine.begin(), bins.fine.end(), 0);
                std::fill(bins.acc.begin(), bins.acc.end(), 0.0);
                bins.v_sum = 0.0;
                for (const auto& pb : part) {
                    if (pb.fine.empty()) continue;
                    for (int k = 0; k < nf; ++k) bins.fine[k] += pb.fine[k];
                    for (size_t k = 0; k < bins.acc.size(); ++k) bins.acc[k] += pb.acc[k];
                    bins.v_sum += pb.v_sum;
                }
                // leverage.cpp:116-139: quantile grid, kernel regression on the bin centres
                const double a = quantile(bins, lo, hi, s.q_lo), b = quantile(bins, lo, hi, s.q_hi);
                for (int g = 0; g < nx; ++g) x_eval[g] = a + (b - a) * g / (nx - 1);
                double bw = s.bandwidth_scale * s.atm_vol * std::sqrt(std::max(t0, 0.15)) * std::pow(n_particles, -0.2);
                if (rule == 1) {  // coef * std(x) * N^-0.2 over the particles, from the bins' sums of x and x^2
                    double sx = 0.0, sxx = 0.0;
                    for (int k = 0; k < nb; ++k) sx += bins.acc[2 * nb + k], sxx += bins.acc[3 * nb + k];
                    const double mean = sx / n_particles, var = std::max(sxx / n_particles - mean * mean, 0.0);
                    bw = opt->bandwidth_coef * std::sqrt(var) * std::pow(n_particles, -0.2);
                    if (std_out) std_out[i] = std::sqrt(var);
                }
                if (bw_out) bw_out[i] = bw;
                const double width = (hi + 1e-12 - lo) / nb;
                for (int k = 0; k < nb; ++k) centres[k] = lo + (k + 0.5) * width;
                const double fallback = bins.v_sum / n_particles;
                const double *cnt = bins.acc.data(), *sv = cnt + nb, *sx = sv + nb, *sxx = sx + nb, *sxv = sxx + nb;
                for (int g = 0; g < nx; ++g) {
                    double S0 = 0, T0 = 0, Ax = 0, Axx = 0, Axv = 0;
                    const double x0 = x_eval[g];
                    for (int k = 0; k < nb; ++k) {
                        const double d = (x0 - centres[k]) / bw, w = std::exp(-0.5 * d * d);
                        S0 += w * cnt[k], T0 += w * sv[k];
                        if (s.local_linear) Ax += w * sx[k], Axx += w * sxx[k], Axv += w * sxv[k];
                    }
                    const double nw = S0 > 1e-8 ? T0 / std::max(S0, 1e-300) : fallback;
                    cond[g] = nw;
                    if (s.local_linear) {
                        const double S1 = Ax - x0 * S0, S2 = Axx - 2 * x0 * Ax + x0 * x0 * S0, T1 = Axv - x0 * T0;
                        const double det = S0 * S2 - S1 * S1;
                        const bool ok = S0 >= 50 && det > 1e-12 * S0 * std::fabs(S2);
                        cond[g] = std::max(ok ? (S2 * T0 - S1 * T1) / det : nw, 1e-8);
                    }
                }
            }
            // leverage.cpp:141-148: L = clip(sqrt(sigma_loc^2 / E[v | x]), l_min, l_max)
            const double* row = s.lv_rows + static_cast<long long>(i) * s.n_lv;
            const double lv_lo = s.lv_lo[i], lv_hi = lv_lo + (s.n_lv - 1) / s.lv_inv_dx[i];
            for (int g = 0; g < nx; ++g) {
                const double loc = interp_uniform(x_eval[g], lv_lo, lv_hi, row, s.n_lv);
                lev[g] = std::min(std::max(std::sqrt(loc / std::max(cond[g], 1e-8)), s.l_min), s.l_max);
                s.x_grids[static_cast<long long>(i) * nx + g] = x_eval[g];
                s.values[static_cast<long long>(i) * nx + g] = lev[g];
            }
            // leverage.cpp:189-214: one Euler step of every particle on the leverage row
            const double ga = x_eval[0], gb = x_eval[nx - 1];
            const long long per = (np + threads - 1) / threads;
            std::vector<double> los(threads, std::numeric_limits<double>::infinity());
            std::vector<double> his(threads, -std::numeric_limits<double>::infinity());
            for_blocks(threads, threads, [&](long long t) {
                double mn = los[t], mx = his[t];
                const long long pa = t * per, pb = std::min(np, pa + per);
                for (long long p = pa; p < pb; ++p) {
                    const Words r4 = philox(Words{static_cast<std::uint32_t>(p), static_cast<std::uint32_t>(
                                                      static_cast<std::uint64_t>(p) >> 32),
                                                  static_cast<std::uint32_t>(i + 1), 0u},
                                            k0, k1);
                    double z[4];
                    gauss_pair(r4.x, r4.y, z[0], z[1]);
                    gauss_pair(r4.z, r4.w, z[2], z[3]);
                    for (int leg = 0; leg < 2; ++leg) {
                        const size_t j = static_cast<size_t>(leg ? np + p : p);
                        const double sg = leg ? -1.0 : 1.0, z1 = sg * z[0], z2 = sg * z[1];
                        const double L = interp_uniform(X[j], ga, gb, lev.data(), nx);
                        const double vp = std::max(V[j], 0.0), var = L * L * vp;
                        const double zx = s.rho * z2 + rc * z1;
                        V[j] += s.kappa * (s.theta - vp) * h + s.xi * std::sqrt(vp * h) * z2;
                        X[j] += -0.5 * var * h + std::sqrt(var * h) * zx;
                        mn = std::min(mn, X[j]), mx = std::max(mx, X[j]);
                    }
                }
                los[t] = mn, his[t] = mx;
            });
            lo = *std::min_element(los.begin(), los.end());
            hi = *std::max_element(his.begin(), his.end());
        }
        rq->total_ms = 0.0;
        return OPTLAB_OK;
    } catch (const Failure& f) {
        g_error = f.what;
        return f.code;
    }
}

double oracle_geometric_asian(double S, double K, double T, double r, double q, double sigma, int n, int call) {
    return geometric_asian(S, K, T, r, q, sigma, n, call != 0);
}

}  // extern "C"
