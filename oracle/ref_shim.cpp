// CPU ORACLE -- test infrastructure only. C linkage over the reference's own CPU engine, compiled from its sources
// where they lie (build_ref.sh) into oracle/_ref/libref.so, so the tests can check the oracle's restatement
// against the compiled reference in this container. Nothing here ships, and the GPU runs never need it.
//
// The struct layouts are those of packages/native/cuda/include/optlab_cuda/requests.hpp:12-93.
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

#include "optlab_cuda/closed_forms.hpp"
#include "optlab_cuda/paths.hpp"
#include "optlab_cuda/philox.hpp"
#include "optlab_cuda/requests.hpp"
#include "optlab_cuda/schemes.hpp"

namespace {

thread_local std::string g_error;

// The normals of schemes.hpp:266-267 (Src::next4) taken from the path engine's stream 0 of the pair
// (paths.hpp:213-233): the reference's QE-M scheme on the draws the oracle uses for its fixed products.
struct PhiloxSource {
    optlab::NormalStream<double> stream;
    void next4(double z[4]) {
        for (int k = 0; k < 4; ++k) z[k] = stream.next();
    }
};

template <optlab::Model M, optlab::Product P>
void sums(const optlab::PathConsts<double>& c, std::uint64_t seed, long long n_pairs, double* s, double* s2) {
    for (long long p = 0; p < n_pairs; ++p) {
        PhiloxSource src{optlab::NormalStream<double>{static_cast<std::uint64_t>(p), static_cast<std::uint32_t>(seed),
                                                      static_cast<std::uint32_t>(seed >> 32)}};
        double out[optlab::kOutputs];
        if constexpr (M == optlab::Model::BlackScholes && P == optlab::Product::European)
            optlab::bs_european_unit(c, src.stream.next(), out);
        else
            optlab::path_unit<double, M, P>(c, src, out);
        for (int k = 0; k < optlab::kOutputs; ++k) s[k] += out[k], s2[k] += out[k] * out[k];
    }
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

int ref_paths(optlab_paths_request* rq, int n_threads) {
    try {
        optlab::api::paths(*rq, 0, n_threads);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

int ref_leverage(optlab_leverage_request* rq, int n_threads) {
    try {
        optlab::api::leverage(*rq, 0, n_threads);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

void ref_philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                unsigned* out) {
    for (long long i = 0; i < n; ++i) {
        const optlab::U4 w = optlab::philox4x32_10({c0 + static_cast<unsigned>(i), c1, c2, c3}, k0, k1);
        out[4 * i] = w.x, out[4 * i + 1] = w.y, out[4 * i + 2] = w.z, out[4 * i + 3] = w.w;
    }
}

void ref_normals(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                 double* out) {
    for (long long p = 0; p < n_pairs; ++p) {
        optlab::NormalStream<double> d{static_cast<std::uint64_t>(first_pair + p), static_cast<std::uint32_t>(seed),
                                       static_cast<std::uint32_t>(seed >> 32), tag};
        for (int k = 0; k < n_normals; ++k) out[p * n_normals + k] = d.next();
    }
}

// The reference's fixed products (schemes.hpp path units, make_consts) on Philox stream 0 of each pair, single
// thread, in pair order: sums and sums of squares of the four pair-average outputs.
int ref_fixed_sums(const optlab_request* c, double* s, double* s2) {
    try {
        optlab::McRequest rq;
        rq.model = static_cast<optlab::Model>(c->model);
        rq.product.product = static_cast<optlab::Product>(c->product);
        for (int k = 0; k < 3; ++k) {
            rq.product.legs[k] = static_cast<optlab::Leg>(c->legs[k]);
            rq.product.strikes[k] = c->strikes[k];
        }
        rq.product.H = c->H;
        rq.product.n_steps = c->n_steps;
        rq.market = {c->S0, c->r, c->q, c->T};
        rq.bs.sigma = c->sigma;
        rq.heston = {c->v0, c->kappa, c->theta, c->xi, c->rho};
        const optlab::PathConsts<double> k = optlab::make_consts<double>(rq);
        const long long n_pairs = c->n_paths / 2;
        for (int o = 0; o < optlab::kOutputs; ++o) s[o] = s2[o] = 0.0;
        using optlab::Model;
        using optlab::Product;
        const bool bs = c->model == 0;
        switch (c->product) {
            case 0: bs ? sums<Model::BlackScholes, Product::European>(k, c->seed, n_pairs, s, s2)
                       : sums<Model::Heston, Product::European>(k, c->seed, n_pairs, s, s2); break;
            case 1: bs ? sums<Model::BlackScholes, Product::Asian>(k, c->seed, n_pairs, s, s2)
                       : sums<Model::Heston, Product::Asian>(k, c->seed, n_pairs, s, s2); break;
            case 2: bs ? sums<Model::BlackScholes, Product::Barrier>(k, c->seed, n_pairs, s, s2)
                       : sums<Model::Heston, Product::Barrier>(k, c->seed, n_pairs, s, s2); break;
            default: g_error = "product must be 0, 1 or 2"; return 1;
        }
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

double ref_geometric_asian(double S, double K, double T, double r, double q, double sigma, int n, int call) {
    return optlab::geometric_asian_price(S, K, T, r, q, sigma, n, call != 0);
}

double ref_bs_price(double S, double K, double T, double r, double q, double sigma, int call) {
    return optlab::bs_price(S, K, T, r, q, sigma, call != 0);
}

double ref_down_out_call_bgk(double S, double K, double H, double T, double r, double q, double sigma, int n) {
    return optlab::down_out_call_bgk(S, K, H, T, r, q, sigma, n);
}

}  // extern "C"
