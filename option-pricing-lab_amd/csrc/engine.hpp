// Host-side declarations shared by the translation units of liboptlab_mi (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/optlab_mi.h"

namespace optlab_mi {

// Errors cross the C boundary as status codes; inside the library they are exceptions.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};
[[noreturn]] inline void invalid(const std::string& what) { throw Error(OPTLAB_EINVAL, what); }
void hip_check(hipError_t e, const char* what);

// Grow-only device buffers reused across calls (calls on one device are serialised by contract).
void* scratch(int slot, size_t bytes);
enum Slot {
    kSlotSpec = 0,
    kSlotOut,
    kSlotPartials,
    kSlotTotals,
    kSlotParticles,
    kSlotStrip,
    kSlotAux,
    kSlotNodes,        // per-row node contractions
    kSlotTables,       // bumped local-variance tables of one call
    kSlotCheckpoints,  // adjoint: path state every few steps
    kSlotSeeds,        // adjoint: payoff seeds and adjoint state
    kSlotAdjoint,      // adjoint: per-workgroup table adjoints
    kSlotBatches,      // adjoint: per-batch table adjoints
    kSlotBook,         // book: the tables of every (scenario, name) and the book's small arrays
    kSlotBookQueue,    // book: the work queue, trades, per-item sums and clocks
    kSlotLsm,          // Longstaff-Schwartz: spots and variances at the dates, values, partial sums, coefficients
    kSlotChunks,       // fixed products and the strip: compensated sums of chunks of the per-block rows
    kSlotRegress,      // the issuer's call regression: terms, states, sums, coefficients
    kSlotModels,       // autocall: the tables of names that are not local vol, SLV leverage rows and ratios
    kSlotScript,       // script payoffs: the path engine's output the payoff kernel reads
    kSlotScriptRows,   // script payoffs: each workgroup's (sum, sum of squares)
    kSlotCount
};

int device_cu_count();

// A stream of the library's own (created once, default priority), for work captured in HIP graphs: the null stream
// cannot be captured.
hipStream_t engine_stream();

// ---- path engine (paths.hip) ----------------------------------------------------------------------------------
// Launch the path kernel for every pair of `rq` in precision prec (0 double, 1 float), output to d_out (device,
// n_names x n_cols x 2 n_pairs doubles) on `stream`; d_var, when not null, receives the variance factor of name 0 at
// the kept columns (n_keep x 2 n_pairs doubles). Returns the kernel time in ms (HIP events).
double launch_paths(const optlab_paths_request& rq, int prec, double* d_out, hipStream_t stream,
                    double* d_var = nullptr);
void check_paths_request(const optlab_paths_request& rq);

// ---- script payoffs (script.hip) ----------------------------------------------------------------------------------
// The paths of rq into device memory, then the payoff kernel compiled from `source` by hipRTC over them.
void run_script(optlab_paths_request& rq, int prec, const char* source, optlab_script_result& res);

// An exercise rule for the path kernel of one name (the pricing pass of Longstaff-Schwartz): at observation j a leg
// exercises when its exercise value max(phi (S / K - 1), 0) is positive and beats coefs[j] . basis(S / K, v), the
// first such date or maturity; nothing is written but the sums. Device pointers.
struct ExerciseRule {
    const double* coefs;  // n_obs x 7 over (1, s, s^2, s^3, v, v^2, s v), 0 for a term not in the basis
    const int* valid;     // n_obs: 1 where date j has a rule (never the last)
    const double* disc;   // n_obs
    double strike, phi;   // phi: +1 call, -1 put
    double* sums;         // 2 x exercise_workgroups(n_pairs): the sums of the pairs' average pay and of its square
};
long long exercise_workgroups(long long n_pairs);
// The path kernel with the rule applied at the observations (no output columns); the kernel time in ms.
double launch_paths_exercise(const optlab_paths_request& rq, int prec, const ExerciseRule& rule, hipStream_t stream);

// ---- fixed products and strips (fixed.hip) ---------------------------------------------------------------------
void check_price_request(const optlab_request& rq);
void run_price(const optlab_request& rq, int prec, optlab_result& out);
void check_strip_request(const optlab_strip_request& rq);
void run_strip(optlab_strip_request& rq, int prec);
// The strip's pathwise sensitivities: delta and d/dv0 (Heston) or d/dsigma (Black-Scholes), n_mats x (n_strikes + 1)
// each with its standard error.
void run_strip_greeks(optlab_strip_request& rq, int prec, double* delta, double* delta_se, double* vega,
                      double* vega_se);
// The strip in the file-scope device array of fixed.hip (no allocation), asynchronous on `st`: the estimates go to
// price_dev and se_dev (device, n_mats x (n_strikes + 1) doubles each).
void run_strip_noalloc(const optlab_strip_request& rq, int prec, double* price_dev, double* se_dev, hipStream_t st);

// ---- SLV particles (leverage.hip) -----------------------------------------------------------------------------
void check_leverage_request(const optlab_leverage_request& rq, bool need_rows = true);
// lvs: the local variance from slice records instead of rq.lv_rows (null: the rows)
void run_leverage(optlab_leverage_request& rq, const optlab_leverage_options* opt, int prec,
                  const optlab_lv_slices* lvs = nullptr);

// ---- surfaces on the device (surface.hip) and the worst-of autocall (autocall.hip) -----------------------------
int surface_upload(const optlab_surface_rows& rows);
void surface_free(int handle);
void surface_table(int handle, const double* nodes, double parallel, double skew, double* out_host);
void surface_node_grad(int handle, const double* tbar_host, double* out_host);
void check_autocall_request(const optlab_autocall_request& rq, bool adjoint);
void check_autocall_models(const optlab_autocall_request& rq, const optlab_autocall_models& m);
// models: NULL for every name local vol
void run_autocall(optlab_autocall_request& rq, int prec, bool adjoint, const optlab_autocall_models* models = nullptr);
void check_book_request(const optlab_book_request& rq);
void run_book(optlab_book_request& rq, int prec);

// ---- American options by Longstaff-Schwartz (lsm.hip) ----------------------------------------------------------
void check_lsm_request(const optlab_lsm_request& rq);
void run_lsm(optlab_lsm_request& rq, int prec);
void check_regress_request(const optlab_regress_request& rq);
void run_regress(optlab_regress_request& rq);

// ---- test entry points (paths.hip) -----------------------------------------------------------------------------
void launch_philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                   unsigned* out, hipStream_t stream);
void launch_normals(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                    int prec, double* out, hipStream_t stream);

}  // namespace optlab_mi
