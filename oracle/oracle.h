/* CPU ORACLE -- test infrastructure only. C entry points of liboracle.so, loaded by the tests, the smoke check
   and the CPU-baseline leg of the benchmark to check the HIP engine. Requests are the structs of
   include/optlab_mi.h; every function returns 0 or an error code (message: oracle_last_error()). */
#ifndef OPTLAB_ORACLE_H
#define OPTLAB_ORACLE_H

#include "../include/optlab_mi.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* oracle_last_error(void);

/* philox.hpp:33-43 on counters (c0 + i, c1, c2, c3), i < n: out holds 4 n words. */
void oracle_philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                   unsigned* out);

/* paths.hpp:213-233: n_normals of each pair first_pair .. first_pair + n_pairs - 1 on a stream, pair-major. */
void oracle_normals(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                    double* out);

/* requests.cpp:264-280 with backend 0 (paths_cpu.cpp): every pair in double on n_threads threads (0: all). */
int oracle_paths(optlab_paths_request* rq, int n_threads);
/* The same, also writing the variance factor v of name 0 (before its floor at 0) at the kept columns into var
   (n_keep x 2 n_pairs doubles), as the device's Longstaff-Schwartz pass keeps it for the basis with v. */
int oracle_paths_var(optlab_paths_request* rq, int n_threads, double* var);

/* The fixed products on Philox stream 0 of each pair, in double (schemes.hpp + cpu_mc.cpp:14-18). */
int oracle_price(const optlab_request* rq, int n_threads, optlab_result* out);

/* The strip of vanillas on QE-M or exact Black-Scholes paths, in double. */
int oracle_strip(optlab_strip_request* rq, int n_threads);

/* leverage.cpp:42-97 with the CPU particles of leverage.cpp:101-186. n_parts fixes the partition of the bin sums
   (the reference sums one contiguous part per thread and adds the parts in order; 0: one part per thread). */
int oracle_leverage(optlab_leverage_request* rq, int n_threads, int n_parts);

/* The same with the options of include/optlab_mi.h (opt NULL: the reference's bandwidth). */
int oracle_leverage_ex(optlab_leverage_request* rq, const optlab_leverage_options* opt, int n_threads, int n_parts);

/* oracle_leverage_ex, also reporting each step's kernel bandwidth (bw_out, n_steps; 0 at step 0, which has none)
   and, under bandwidth rule 1, the standard deviation of the particles' log-spot offsets that set it (std_out,
   n_steps; 0 under rule 0), and the 2 n_pairs log-spot offsets x the particles hold when step x_step starts
   (x_out; rows p and n_pairs + p are pair p). Any output may be NULL. */
int oracle_leverage_trace(optlab_leverage_request* rq, const optlab_leverage_options* opt, int n_threads,
                          int n_parts, double* bw_out, double* std_out, int x_step, double* x_out);

/* closed_forms.cpp:21-33: geometric Asian on n equally spaced fixings. */
double oracle_geometric_asian(double S, double K, double T, double r, double q, double sigma, int n, int call);

#ifdef __cplusplus
}
#endif

#endif
