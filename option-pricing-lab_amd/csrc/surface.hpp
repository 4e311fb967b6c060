// Implied-vol surfaces on the device: the base surface's rows (w and its derivatives on the engine's table grid),
// the risk grid's tents, and the launches that turn them into (bumped) local-variance tables and node derivatives.
#pragma once

#include <hip/hip_runtime.h>

namespace optlab_mi {

constexpr int kTentMaxStrikes = 16;
constexpr int kTentMaxTenors = 16;

// packages/risk/src/optlab/risk/grid.py:25-26, 73-97: strike nodes in log forward moneyness, tenor nodes in years,
// each strike hat as hat_const + sum_m hat_c[m] (k - hat_a[m])_+, and the Gaussian smoothing width constants.
struct TentGrid {
    int n_strikes, n_tenors;
    double log_strikes[kTentMaxStrikes];
    double tenors[kTentMaxTenors];
    double hat_const[kTentMaxStrikes];
    int hat_n[kTentMaxStrikes];
    double hat_a[kTentMaxStrikes][3], hat_c[kTentMaxStrikes][3];
    double atm, wing, scale, t_floor;
};

// Device pointers to a surface's rows: row r is read at maturity T[r] on k = lo[r] + x / inv_dx[r], x < n_x.
struct SurfaceRows {
    long long n_rows;
    int n_x;
    const double* T;
    const double* lo;
    const double* inv_dx;
    const double* w;
    const double* wk;
    const double* wkk;
    const double* wT;
};

// Moves of the whole surface (packages/structured/src/optlab/structured/surfaces.py:18-50): a parallel shift of
// the implied vols and the skew rotation -skew c sqrt(1 / max(T, 1/12)) tanh(k / 0.3).
struct SurfaceMove {
    double parallel, skew;
};

TentGrid make_tent_grid();

// The local-variance table (n_rows x n_x, R = float or double by prec 1 / 0) of the surface moved by the node bumps
// H (n_strikes x n_tenors in vol units, or null) and by `move`.
void launch_bumped_table(const SurfaceRows& s, const TentGrid& g, const double* H, SurfaceMove move, int prec,
                         void* out, hipStream_t stream);

constexpr int kMaxSurfaces = 5;

struct SurfaceSet {  // surfaces of one shape (n_rows, n_x), contracted in one launch
    SurfaceRows s[kMaxSurfaces];
    int n;
};

struct DerivSet {  // each surface's node derivatives (surface_node_derivatives)
    const double* d[kMaxSurfaces];
};

// out[(a * n_batches + b) * n_nodes + node] = sum over (row, x) of tbar[a][row][x][b] * unit * dLV_a[row][x] /
// dH_node at H = 0: the adjoint of surface a's table (n_rows x n_x entries, fixed-point integers of weight `unit`)
// contracted with the table's node derivatives. The batches of an entry are contiguous, or with batch_major the
// batches' tables follow one another (tbar[b][a][row][x]).
void launch_node_contraction(const SurfaceSet& set, const DerivSet& dset, const TentGrid& g, const long long* tbar,
                             int n_batches, double unit, double* out, hipStream_t stream,
                             bool batch_major = false);

// The rows of an uploaded surface (handle from surface_upload) and its unmoved local-variance table in precision
// prec (0 double, 1 float), built on first use and kept with the surface.
const SurfaceRows& surface_rows(int handle);
const void* surface_base_table(int handle, int prec, hipStream_t stream);
// dLV / dH for the two active tenor tents of each row and every strike tent, (n_rows x n_x) x 2 n_strikes doubles,
// built on first use and kept with the surface.
const double* surface_node_derivatives(int handle, hipStream_t stream);

}  // namespace optlab_mi
