// A book of trades revalued under scenarios in one persistent launch.
// What it replaces: the risk engine's full revaluation (packages/risk/src/optlab/risk/engine.py:600-645: every
// scenario re-simulated on the trade's seed) driving the structured engine's batch pipeline
// (packages/structured/src/optlab/structured/engine.py:199-330), one trade and one scenario at a time.
// Products: European call, arithmetic Asian call, up-and-out call on every close (structured/engine.py products on
// the path engine's spots) and the worst-of autocall of autocall.hip with hard comparisons. Paths: the local-vol
// step of packages/native/cuda/include/optlab_cuda/paths.hpp:274-361 on stream 0 of the trade's seed.
//
// MI355X design: the queue holds (trade, scenario, chunk of pairs) items ordered longest first; a persistent grid
// of exactly the resident workgroups walks it with a stride of the grid, so trades of 1 to 5 names, 63 to 1260
// steps and 2^14 to 2^17 paths share the machine (longest first keeps the tail short). An item is blockDim x
// pairs_per_lane pairs of one (trade, scenario), taken by one workgroup: one pair per lane (both legs in
// registers), pairs_per_lane pairs in turn, the table rows of each step staged in LDS once for all the workgroup's
// waves (double-buffered, one barrier a step) instead of gathered per lane from HBM. Each item's pair sums go to its
// own slot and a reduction adds the slots of a (trade, scenario) in chunk order, so results do not depend on which
// workgroup ran what. Trades are split by name count, each class run by a persistent kernel compiled for exactly
// that many names; the clock stamp, the class kernels and the reduction are captured in one HIP graph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "device.hpp"
#include "engine.hpp"
#include "localvol.hpp"
#include "surface.hpp"

namespace optlab_mi {

namespace {

constexpr int kNames = 5;       // names of one trade
constexpr int kBookNames = 64;  // names of one book
constexpr int kMaxBlock = 512;

template <typename R>
struct TradeDev {
    int product, n, n_steps, obs_every, memory, chunks;
    int names[kNames];
    R root[kNames * kNames];
    R strike, barrier, coupon, coupon_barrier, autocall, put_strike;
    R lfix[kNames];  // ln of each name's fixing (its base spot)
    long long n_pairs, slot0;
    std::uint32_t key0, key1;
};

struct Item {
    int trade, scen, chunk, pad;
    long long slot;
};

// The fields of an item and its trade that steer control flow, read as scalars (readfirstlane): every loop around a
// barrier then has an exit the compiler knows to be the same in all lanes.
struct Shape {
    int product, n, n_steps, obs_every, memory, scen;
    int names[kNames];
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// What a step reads of each name besides its table row, in one scalar load: (lo, 1 / dx of the row's grid, ln F at
// the end of the step, 0), at name * n_steps + step.
template <typename R>
struct StepRec {
    typedef R type __attribute__((ext_vector_type(4)));
};

template <typename R>
struct BookArgs {
    int n_book, n_steps, n_x, n_scen, n_items, pairs_per_lane;
    R dt;
    const R* tables;  // ((scen * n_book + name) * n_steps + step) * n_x
    const typename StepRec<R>::type* steps;  // name * n_steps + step
    const R* lf;      // name * (n_steps + 1) + step: ln F
    const R* lsh;     // scen * n_book + name: ln(1 + shock)
    const R* disc;    // n_steps + 1
    const TradeDev<R>* trades;
    const Item* items;
    double* partial;              // slot x 2
    unsigned long long* finish;   // slot
    unsigned long long* start;
};

// The rows of step `step` of the NM names into LDS (NM x n_x), by the workgroup; base[k] is name k's first row
// under the trade's scenario.
template <typename R, int NM>
__device__ __forceinline__ void stage_rows(const BookArgs<R>& a, const R* const (&base)[NM], int step, R* dst) {
    const int off = step * a.n_x;
#pragma unroll
    for (int k = 0; k < NM; ++k) {  // unrolled: the names stay in registers (a runtime index would put them on
                                    // the stack)
        const R* src = base[k] + off;
        for (int c = threadIdx.x; c < a.n_x; c += blockDim.x) dst[k * a.n_x + c] = src[c];
    }
}

// Value of pair p of trade t under scenario s: the average of its two legs. Every lane of the workgroup calls it
// together (lanes past the trade's pairs compute a pair they then drop): the step's table rows are staged in LDS,
// double-buffered, with one barrier a step.
template <typename R, int NM>
__device__ double pair_value(const BookArgs<R>& a, const TradeDev<R>& t, const Shape& sh, long long p, R* rows) {
    NormalStream<R> rng;
    rng.init(static_cast<std::uint64_t>(p), ldc(&t.key0, 0), ldc(&t.key1, 0), 0u);
    constexpr int n = NM;  // the kernel of a trade's class is compiled for its name count
    const int scen = sh.scen;
    const int stride = a.n_steps + 1;
    R x[2][NM];
    R lsh[NM];
    int rec[NM];  // the name's first step record; its ln F row starts at name * (n_steps + 1)
    int lfr[NM];
    const R* base[NM];  // the name's first table row under this scenario
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        x[0][k] = x[1][k] = R(0);
        const int nm = k < n ? sh.names[k] : sh.names[0];
        lsh[k] = ldc(a.lsh, scen * a.n_book + nm);
        rec[k] = nm * a.n_steps;
        lfr[k] = nm * stride;
        base[k] = a.tables + (static_cast<long long>(scen) * a.n_book + nm) * a.n_steps * a.n_x;
    }
    // product state, per leg
    R acc[2] = {R(0), R(0)};        // Asian: sum of fixings; barrier: running max of ln S
    R ymin[2] = {R(0), R(0)};       // autocall: lowest ln performance over every close
    Note<R> note[2] = {{R(1), R(0), R(0)}, {R(1), R(0), R(0)}};
    const Terms<R> terms{ldc(&t.autocall, 0), ldc(&t.coupon_barrier, 0), ldc(&t.coupon, 0), ldc(&t.barrier, 0),
                         ldc(&t.put_strike, 0), R(0), sh.memory};
    const int product = sh.product, n_steps = sh.n_steps, obs_every = sh.obs_every;
    if (product == 2) acc[0] = acc[1] = R(-1e30);
    if (product == 3) {
        R y0 = ldc(a.lf, lfr[0]) + lsh[0] - ldc(t.lfix, 0);
#pragma unroll
        for (int k = 1; k < NM; ++k)
            if (k < n) y0 = r_min(y0, ldc(a.lf, lfr[k]) + lsh[k] - ldc(t.lfix, k));
        ymin[0] = ymin[1] = y0;
    }
    int obs_j = 0;
    const int n_obs = product == 3 ? n_steps / obs_every : 0;
    const int stage_len = n * a.n_x;
    __syncthreads();  // the previous pass has read its last rows
    stage_rows<R, NM>(a, base, 0, rows);
    int to_fix = obs_every;  // steps to the next fixing (a countdown: a runtime modulo costs a division a step)
    auto one_step = [&](int step) __attribute__((always_inline)) {
        const R* cur = rows + (step & 1) * stage_len;
        __syncthreads();  // this step's rows are in; everyone is done with the last step's buffer
        if (step + 1 < n_steps) stage_rows<R, NM>(a, base, step + 1, rows + ((step + 1) & 1) * stage_len);
        R zs[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k)
            if (k < n) zs[k] = rng.next();
        R zc[NM];  // correlated normals of the + leg (the - leg takes their negatives)
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            if (k >= n) break;
            R acc_z = R(0);
#pragma unroll
            for (int b = 0; b < NM; ++b)
                if (b < n) acc_z += ldc(t.root, k * n + b) * zs[b];
            zc[k] = acc_z;
        }
        R var[NM];  // the table reads do not depend on the leg's sign until x differs: read per leg
        typename StepRec<R>::type g[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k)
            if (k < n) g[k] = ldc(a.steps, rec[k] + step);
        const int t1 = step + 1;
        const bool fix = --to_fix == 0;  // t1 % obs_every == 0
        if (fix) to_fix = obs_every;
#pragma unroll
        for (int leg = 0; leg < 2; ++leg) {
            const R sign = leg ? R(-1) : R(1);
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                if (k >= n) break;
                const R z = sign * zc[k];
                const Lerp<R> l = lerp_at(cur + k * a.n_x, g[k].x, g[k].y, a.n_x, x[leg][k]);
                var[k] = l.var;
                x[leg][k] += R(-0.5) * var[k] * a.dt + r_sqrt(var[k] * a.dt) * z;
            }
            const R y = g[0].z + lsh[0] + x[leg][0];  // ln S of the first name
            if (product == 1) {
                if (fix) acc[leg] += r_exp(y);
            } else if (product == 2) {
                acc[leg] = r_max(acc[leg], y);
            } else if (product == 3) {
                R yw = y - ldc(t.lfix, 0);
#pragma unroll
                for (int k = 1; k < NM; ++k)
                    if (k < n) yw = r_min(yw, g[k].z + lsh[k] + x[leg][k] - ldc(t.lfix, k));
                ymin[leg] = r_min(ymin[leg], yw);
                if (fix) observe(terms, obs_j == n_obs - 1, r_exp(yw), r_exp(ymin[leg]), ldc(a.disc, t1), note[leg]);
            }
        }
        if (product == 3 && fix) ++obs_j;
    };
    // a step draws exactly NM normals, so four steps draw NM whole Philox blocks: unrolled by four, the stream's
    // buffer position is a compile-time constant at every draw
    int step = 0;
    for (; step + 4 <= n_steps; step += 4) {
        one_step(step);
        one_step(step + 1);
        one_step(step + 2);
        one_step(step + 3);
    }
    for (; step < n_steps; ++step) one_step(step);
    const R D = ldc(a.disc, n_steps);
    const R strike = ldc(&t.strike, 0);
    double v = 0.0;
#pragma unroll
    for (int leg = 0; leg < 2; ++leg) {
        const R ST = r_exp(ldc(a.lf, lfr[0] + n_steps) + lsh[0] + x[leg][0]);
        R pay;
        if (product == 0) {
            pay = D * r_max(ST - strike, R(0));
        } else if (product == 1) {
            const R avg = acc[leg] / static_cast<R>(n_steps / obs_every);
            pay = D * r_max(avg - strike, R(0));
        } else if (product == 2) {
            pay = acc[leg] < r_log(terms.ki_barrier) ? D * r_max(ST - strike, R(0)) : R(0);
        } else {
            pay = note[leg].value;
        }
        v += static_cast<double>(pay);
    }
    return 0.5 * v;
}

// Waves a SIMD asked of the compiler for the kernel of NM names: 6 for FP32 (80 VGPRs, three 512-lane workgroups a
// CU), which the 5-name kernel (88 VGPRs, two workgroups) then meets; 1 leaves the choice to the compiler.
template <typename R, int NM>
constexpr int book_waves() {
    return sizeof(R) == 4 ? 6 : 1;
}

// Persistent, one launch per class of trades by name count (NM: the trades' exact name count): a grid of the
// resident workgroups walks its class's queue (ordered longest first) with a stride of the grid, item it on
// workgroup it % grid. An item is blockDim x pairs_per_lane pairs of one (trade, scenario): every wave of the
// workgroup reads the same staged rows. Every loop bound is the same in all lanes (kernel arguments and the item
// index), so each barrier is reached by the whole workgroup. Each item's pair sums go to its own slot.
template <typename R, int NM>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(book_waves<R, NM>(), 8)))
book_kernel(BookArgs<R> a) {
    extern __shared__ unsigned char lds_raw[];
    R* rows = reinterpret_cast<R*>(lds_raw);  // 2 x NM x n_x
    __shared__ double part[kMaxBlock / 64][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
    for (long long it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const Item* item = a.items + it;
        const TradeDev<R>& t = a.trades[uniform(ldc(&item->trade, 0))];
        Shape sh;
        sh.product = uniform(ldc(&t.product, 0)), sh.n = uniform(ldc(&t.n, 0)), sh.n_steps = uniform(ldc(&t.n_steps, 0));
        sh.obs_every = uniform(ldc(&t.obs_every, 0)), sh.memory = uniform(ldc(&t.memory, 0));
        sh.scen = uniform(ldc(&item->scen, 0));
#pragma unroll
        for (int k = 0; k < kNames; ++k) sh.names[k] = uniform(ldc(t.names, k));
        const int chunk = uniform(ldc(&item->chunk, 0));
        const long long slot = ldc(&item->slot, 0);
        const long long n_pairs = ldc(&t.n_pairs, 0);
        double s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < a.pairs_per_lane; ++k) {
            const long long p = (static_cast<long long>(chunk) * a.pairs_per_lane + k) * blockDim.x + threadIdx.x;
            const bool live = p < n_pairs;
            const double v = pair_value<R, NM>(a, t, sh, live ? p : 0, rows);
            if (live) {
                s1 += v;
                s2 += v * v;
            }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            s1 += shfl_xor_d(s1, m);
            s2 += shfl_xor_d(s2, m);
        }
        if (lane == 0) part[wave][0] = s1, part[wave][1] = s2;
        __syncthreads();
        if (threadIdx.x < 2) {
            double v = 0.0;
            for (int w = 0; w < n_waves; ++w) v += part[w][threadIdx.x];
            a.partial[slot * 2 + threadIdx.x] = v;
            if (threadIdx.x == 0) a.finish[slot] = wall_clock64();
        }
        __syncthreads();  // part is read before the next item writes it
    }
}

__global__ void stamp_kernel(unsigned long long* start) { *start = wall_clock64(); }

// sums[(t * n_scen + s) * 2 + k] = the slots of (t, s) added in chunk order.
template <typename R>
__global__ void book_reduce_kernel(const TradeDev<R>* trades, int n_trades, int n_scen, const double* partial,
                                   double* sums) {
    const long long ts = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ts >= static_cast<long long>(n_trades) * n_scen) return;
    const int ti = static_cast<int>(ts / n_scen), s = static_cast<int>(ts % n_scen);
    const TradeDev<R>& t = trades[ti];
    double v1 = 0.0, v2 = 0.0;
    for (int c = 0; c < t.chunks; ++c) {
        const long long slot = t.slot0 + static_cast<long long>(s) * t.chunks + c;
        v1 += partial[slot * 2];
        v2 += partial[slot * 2 + 1];
    }
    sums[ts * 2] = v1;
    sums[ts * 2 + 1] = v2;
}

// ---- host -------------------------------------------------------------------------------------------------------


template <typename R>
struct Prepared {
    std::vector<TradeDev<R>> trades;
    long long n_slots = 0;
};

template <typename R>
Prepared<R> prepare(const optlab_book_request& rq, const std::vector<double>& spot, int block, int ppl) {
    Prepared<R> pr;
    const long long per_item = static_cast<long long>(block) * ppl;  // an item is one workgroup's
    for (int i = 0; i < rq.n_trades; ++i) {
        const optlab_trade& c = rq.trades[i];
        TradeDev<R> t{};
        t.product = c.product, t.n = c.n_names, t.n_steps = c.n_steps, t.obs_every = c.obs_every;
        t.memory = c.memory;
        for (int k = 0; k < kNames; ++k) t.names[k] = k < c.n_names ? c.names[k] : 0;
        for (int k = 0; k < kNames * kNames; ++k) t.root[k] = static_cast<R>(c.root[k]);
        const double s0 = spot[c.names[0]];
        const bool vanilla = c.product != 3;
        t.strike = static_cast<R>(vanilla ? c.strike * s0 : 0.0);
        t.barrier = static_cast<R>(c.product == 2 ? c.barrier * s0 : c.barrier);
        t.put_strike = static_cast<R>(vanilla ? 0.0 : c.strike);
        t.coupon = static_cast<R>(c.coupon), t.coupon_barrier = static_cast<R>(c.coupon_barrier);
        t.autocall = static_cast<R>(c.autocall);
        for (int k = 0; k < kNames; ++k) t.lfix[k] = static_cast<R>(k < c.n_names ? std::log(spot[c.names[k]]) : 0.0);
        t.n_pairs = c.n_pairs;
        t.chunks = static_cast<int>((c.n_pairs + per_item - 1) / per_item);
        t.slot0 = pr.n_slots;
        pr.n_slots += static_cast<long long>(rq.n_scen) * t.chunks;
        t.key0 = static_cast<std::uint32_t>(c.seed), t.key1 = static_cast<std::uint32_t>(c.seed >> 32);
        pr.trades.push_back(t);
    }
    return pr;
}

// The queue: every (trade, scenario, chunk) of the chosen trades, longest first (steps x names; ties by index).
template <typename R>
std::vector<Item> queue_of(const Prepared<R>& pr, int n_scen, const std::vector<int>& which) {
    std::vector<Item> items;
    for (int ti : which) {
        const TradeDev<R>& t = pr.trades[ti];
        for (int s = 0; s < n_scen; ++s)
            for (int c = 0; c < t.chunks; ++c)
                items.push_back(Item{ti, s, c, 0, t.slot0 + static_cast<long long>(s) * t.chunks + c});
    }
    std::stable_sort(items.begin(), items.end(), [&](const Item& x, const Item& y) {
        const TradeDev<R>&a = pr.trades[x.trade], &b = pr.trades[y.trade];
        return static_cast<long long>(a.n_steps) * a.n > static_cast<long long>(b.n_steps) * b.n;
    });
    return items;
}

template <typename R>
struct Device {
    BookArgs<R> args{};
    TradeDev<R>* trades = nullptr;
    double* sums = nullptr;
};

template <typename R, int NM>
int resident_blocks(int block, size_t lds) {
    int per_cu = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, book_kernel<R, NM>, block, lds), "occupancy");
    return std::max(per_cu, 1) * device_cu_count();
}

constexpr int kClasses = kNames;  // class c: the trades on c + 1 names

template <typename R>
int class_of(const TradeDev<R>& t) {
    return t.n - 1;
}

template <typename R>
int class_resident(int c, int block, size_t lds) {
    switch (c) {
        case 0: return resident_blocks<R, 1>(block, lds);
        case 1: return resident_blocks<R, 2>(block, lds);
        case 2: return resident_blocks<R, 3>(block, lds);
        case 3: return resident_blocks<R, 4>(block, lds);
        default: return resident_blocks<R, 5>(block, lds);
    }
}

template <typename R>
void launch_class(int c, const BookArgs<R>& ac, int grid, int block, size_t lds, hipStream_t st) {
    switch (c) {
        case 0: hipLaunchKernelGGL((book_kernel<R, 1>), dim3(grid), dim3(block), lds, st, ac); break;
        case 1: hipLaunchKernelGGL((book_kernel<R, 2>), dim3(grid), dim3(block), lds, st, ac); break;
        case 2: hipLaunchKernelGGL((book_kernel<R, 3>), dim3(grid), dim3(block), lds, st, ac); break;
        case 3: hipLaunchKernelGGL((book_kernel<R, 4>), dim3(grid), dim3(block), lds, st, ac); break;
        default: hipLaunchKernelGGL((book_kernel<R, 5>), dim3(grid), dim3(block), lds, st, ac); break;
    }
}

// One launch of the queue `items` with its clock stamp and its reduction (captured in a HIP graph and the graph
// launched, when `graph`), synchronised. The items are split by the name count of their trade into up to five
// persistent launches, each of the kernel compiled for exactly that many names (no per-name conditions, the
// correlation root at fixed offsets), each over its own longest-first queue.
template <typename R>
float launch_queue(BookArgs<R> args, const std::vector<Item>& items, const std::vector<TradeDev<R>>& trades,
                   Item* d_items, int block, int n_trades, int n_scen, const TradeDev<R>* d_trades, double* d_sums,
                   bool graph) {
    hipStream_t st = engine_stream();
    std::vector<Item> ordered;
    ordered.reserve(items.size());
    int count[kClasses] = {}, first[kClasses] = {};
    for (int c = 0; c < kClasses; ++c) {
        first[c] = static_cast<int>(ordered.size());
        for (const Item& it : items)
            if (class_of(trades[it.trade]) == c) ordered.push_back(it), ++count[c];
    }
    hip_check(hipMemcpyAsync(d_items, ordered.data(), ordered.size() * sizeof(Item), hipMemcpyHostToDevice, st),
              "copy queue");
    int grid[kClasses];
    size_t lds[kClasses];
    for (int c = 0; c < kClasses; ++c) {
        lds[c] = 2 * static_cast<size_t>(c + 1) * args.n_x * sizeof(R);
        grid[c] = std::max(1, std::min(class_resident<R>(c, block, lds[c]), count[c]));
    }
    const unsigned red = static_cast<unsigned>((static_cast<long long>(n_trades) * n_scen + 255) / 256);
    auto body = [&]() {
        hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(1), 0, st, args.start);
        for (int c = 0; c < kClasses; ++c) {
            if (!count[c]) continue;
            BookArgs<R> ac = args;
            ac.items = d_items + first[c];
            ac.n_items = count[c];
            launch_class<R>(c, ac, grid[c], block, lds[c], st);
        }
        hipLaunchKernelGGL(book_reduce_kernel<R>, dim3(red), dim3(256), 0, st, d_trades, n_trades, n_scen,
                           args.partial, d_sums);
    };
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    if (graph) {
        hipGraph_t g;
        hipGraphExec_t ge;
        hip_check(hipStreamSynchronize(st), "synchronize");
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
        body();
        hip_check(hipStreamEndCapture(st, &g), "end capture");
        hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "instantiate");
        hip_check(hipEventRecord(e0, st), "event");
        hip_check(hipGraphLaunch(ge, st), "graph launch");
        hip_check(hipEventRecord(e1, st), "event");
        hip_check(hipStreamSynchronize(st), "synchronize");
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
    } else {
        hip_check(hipEventRecord(e0, st), "event");
        body();
        hip_check(hipGetLastError(), "book launch");
        hip_check(hipEventRecord(e1, st), "event");
        hip_check(hipStreamSynchronize(st), "synchronize");
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

std::map<std::tuple<int, int, long long, int>, std::pair<int, int>> g_tuned;  // (trades, scen, pairs, prec) -> choice

template <typename R>
void run_typed(optlab_book_request& rq) {
    const auto wall0 = std::chrono::steady_clock::now();
    const int nb = rq.n_names, ns = rq.n_steps, nsc = rq.n_scen;
    hipStream_t st = engine_stream();
    const int prec = std::is_same<R, double>::value ? 0 : 1;
    const TentGrid grid = make_tent_grid();

    // the tables of every (scenario, name), and the rows' grids
    const SurfaceRows* rows[kBookNames];
    for (int a = 0; a < nb; ++a) rows[a] = &surface_rows(rq.surface[a]);
    const int n_x = rows[0]->n_x;
    const size_t table = static_cast<size_t>(ns) * n_x;
    // step records (4 values each, 64-byte aligned after the tables), then ln F, ln(1 + shock), discount factors
    const size_t tables_n = (static_cast<size_t>(nsc) * nb * table + 63) & ~size_t(63);
    const size_t small = static_cast<size_t>(nb) * ns * 4 + static_cast<size_t>(nb) * (ns + 1) +
                         static_cast<size_t>(nsc) * nb + (ns + 1);
    R* mem = static_cast<R*>(scratch(kSlotBook, (tables_n + small) * sizeof(R)));
    R* tables = mem;
    R* steps = tables + tables_n;
    R* lf = steps + static_cast<size_t>(nb) * ns * 4;
    R* lsh = lf + static_cast<size_t>(nb) * (ns + 1);
    R* disc = lsh + static_cast<size_t>(nsc) * nb;
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    hip_check(hipEventRecord(t0, st), "event");
    for (int s = 0; s < nsc; ++s)
        for (int a = 0; a < nb; ++a)
            launch_bumped_table(*rows[a], grid, nullptr, SurfaceMove{rq.vol_shift[s], rq.skew[s]}, prec,
                                tables + (static_cast<size_t>(s) * nb + a) * table, st);
    hip_check(hipEventRecord(t1, st), "event");
    std::vector<double> spot(nb);
    {
        std::vector<double> hlo(static_cast<size_t>(nb) * ns), hinv(hlo.size());
        for (int a = 0; a < nb; ++a) {
            hip_check(hipMemcpy(hlo.data() + static_cast<size_t>(a) * ns, rows[a]->lo, ns * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "copy rows");
            hip_check(hipMemcpy(hinv.data() + static_cast<size_t>(a) * ns, rows[a]->inv_dx, ns * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "copy rows");
            spot[a] = rq.fwd[static_cast<size_t>(a) * (ns + 1)];
        }
        std::vector<R> host(small);
        size_t k = 0;
        for (int a = 0; a < nb; ++a)
            for (int i = 0; i < ns; ++i) {
                const size_t r = static_cast<size_t>(a) * ns + i;
                host[k++] = static_cast<R>(hlo[r]);
                host[k++] = static_cast<R>(hinv[r]);
                host[k++] = static_cast<R>(std::log(rq.fwd[static_cast<size_t>(a) * (ns + 1) + i + 1]));
                host[k++] = R(0);
            }
        for (size_t i = 0; i < static_cast<size_t>(nb) * (ns + 1); ++i) host[k++] = static_cast<R>(std::log(rq.fwd[i]));
        for (size_t i = 0; i < static_cast<size_t>(nsc) * nb; ++i) host[k++] = static_cast<R>(std::log1p(rq.spot_shock[i]));
        for (int i = 0; i <= ns; ++i) host[k++] = static_cast<R>(rq.disc[i]);
        hip_check(hipMemcpy(steps, host.data(), small * sizeof(R), hipMemcpyHostToDevice), "copy book");  // `host` ends here
    }

    BookArgs<R> args{};
    args.n_book = nb, args.n_steps = ns, args.n_x = n_x, args.n_scen = nsc;
    args.dt = static_cast<R>(rq.dt);
    args.tables = tables, args.lf = lf, args.lsh = lsh, args.disc = disc;
    args.steps = reinterpret_cast<const typename StepRec<R>::type*>(steps);

    // workgroup size and pairs per lane: the request's, or the fastest of a few on a sample of the book
    long long total_pairs = 0;
    for (int i = 0; i < rq.n_trades; ++i) total_pairs += rq.trades[i].n_pairs;
    int block = rq.block, ppl = rq.pairs_per_lane;
    const auto key = std::make_tuple(rq.n_trades, nsc, total_pairs, prec);

    auto run = [&](int blk, int pp, const std::vector<int>& which, int n_scen_run, bool graph, float* ms_out,
                   Prepared<R>* keep) {
        Prepared<R> pr = prepare<R>(rq, spot, blk, pp);
        std::vector<Item> items = queue_of(pr, n_scen_run, which);
        const size_t slots = static_cast<size_t>(pr.n_slots);
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };  // every piece starts 256-byte aligned
        const size_t bytes = up(items.size() * sizeof(Item)) + up(pr.trades.size() * sizeof(TradeDev<R>)) +
                             up(slots * 2 * sizeof(double)) + up(slots * sizeof(unsigned long long)) +
                             up(static_cast<size_t>(rq.n_trades) * nsc * 2 * sizeof(double)) +
                             up(sizeof(unsigned long long));
        unsigned char* m = static_cast<unsigned char*>(scratch(kSlotBookQueue, bytes));
        auto carve = [&](size_t bytes) {
            unsigned char* p = m;
            m += (bytes + 255) & ~size_t(255);
            return p;
        };
        Item* d_items = reinterpret_cast<Item*>(carve(items.size() * sizeof(Item)));
        TradeDev<R>* d_trades = reinterpret_cast<TradeDev<R>*>(carve(pr.trades.size() * sizeof(TradeDev<R>)));
        double* d_partial = reinterpret_cast<double*>(carve(slots * 2 * sizeof(double)));
        unsigned long long* d_finish = reinterpret_cast<unsigned long long*>(carve(slots * sizeof(unsigned long long)));
        double* d_sums = reinterpret_cast<double*>(carve(static_cast<size_t>(rq.n_trades) * nsc * 2 * sizeof(double)));
        unsigned long long* d_start = reinterpret_cast<unsigned long long*>(carve(sizeof(unsigned long long)));
        hip_check(hipMemcpyAsync(d_trades, pr.trades.data(), pr.trades.size() * sizeof(TradeDev<R>),
                                 hipMemcpyHostToDevice, st),
                  "copy trades");
        BookArgs<R> ar = args;
        ar.trades = d_trades, ar.pairs_per_lane = pp, ar.partial = d_partial;
        ar.finish = d_finish, ar.start = d_start;
        *ms_out = launch_queue(ar, items, pr.trades, d_items, blk, rq.n_trades, n_scen_run, d_trades, d_sums, graph);
        if (keep) *keep = pr;
        return std::make_tuple(d_sums, d_finish, d_start);
    };

    std::vector<int> all(rq.n_trades);
    for (int i = 0; i < rq.n_trades; ++i) all[i] = i;
    if (block == 0 || ppl == 0) {
        auto hit = g_tuned.find(key);
        if (hit != g_tuned.end()) {
            block = hit->second.first, ppl = hit->second.second;
        } else {
            // the sample: every trade under one scenario (the queue's shape without its length)
            const int cand[][2] = {{256, 1}, {256, 2}, {512, 1}, {128, 2}, {256, 4}, {512, 2}};
            float best = 1e30f;
            for (const auto& c : cand) {
                float ms = 0.f;
                run(c[0], c[1], all, 1, false, &ms, nullptr);  // first use of a shape pays its warm-up
                run(c[0], c[1], all, 1, false, &ms, nullptr);
                if (ms < best) best = ms, block = c[0], ppl = c[1];
            }
            g_tuned[key] = {block, ppl};
        }
    }
    if (block != 128 && block != 256 && block != 512) invalid("block must be 128, 256 or 512");
    if (ppl != 1 && ppl != 2 && ppl != 4 && ppl != 8) invalid("pairs_per_lane must be 1, 2, 4 or 8");
    rq.block_used = block, rq.pairs_per_lane_used = ppl;

    std::vector<double> sums(static_cast<size_t>(rq.n_trades) * nsc * 2);
    Prepared<R> pr;
    double kernel_ms = 0.0;
    const double clock_khz = [] {
        int dev = 0, khz = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
        return static_cast<double>(khz > 0 ? khz : 100000);
    }();
    if (rq.mode != 1) {
        float ms = 0.f;
        auto [d_sums, d_finish, d_start] = run(block, ppl, all, nsc, rq.mode == 0, &ms, &pr);
        kernel_ms = ms;
        hip_check(hipMemcpy(sums.data(), d_sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost), "copy sums");
        if (rq.finish_ms) {
            std::vector<unsigned long long> fin(static_cast<size_t>(pr.n_slots));
            unsigned long long start = 0;
            hip_check(hipMemcpy(fin.data(), d_finish, fin.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost),
                      "copy clocks");
            hip_check(hipMemcpy(&start, d_start, sizeof start, hipMemcpyDeviceToHost), "copy clock");
            for (int i = 0; i < rq.n_trades; ++i) {
                const TradeDev<R>& t = pr.trades[i];
                unsigned long long last = start;
                for (long long k = 0; k < static_cast<long long>(nsc) * t.chunks; ++k)
                    last = std::max(last, fin[t.slot0 + k]);
                rq.finish_ms[i] = static_cast<double>(last - start) / clock_khz;
            }
        }
    } else {
        for (int i = 0; i < rq.n_trades; ++i) {
            float ms = 0.f;
            auto [d_sums, d_finish, d_start] = run(block, ppl, std::vector<int>{i}, nsc, false, &ms, &pr);
            (void)d_finish, (void)d_start;
            kernel_ms += ms;
            if (rq.finish_ms) rq.finish_ms[i] = ms;
            hip_check(hipMemcpy(sums.data() + static_cast<size_t>(i) * nsc * 2, d_sums + static_cast<size_t>(i) * nsc * 2,
                                static_cast<size_t>(nsc) * 2 * sizeof(double), hipMemcpyDeviceToHost),
                      "copy sums");
        }
    }
    long long name_steps = 0;
    for (int i = 0; i < rq.n_trades; ++i) {
        const optlab_trade& c = rq.trades[i];
        const double n = static_cast<double>(c.n_pairs);
        name_steps += 2 * c.n_pairs * c.n_steps * static_cast<long long>(c.n_names) * nsc;
        for (int s = 0; s < nsc; ++s) {
            const double s1 = sums[(static_cast<size_t>(i) * nsc + s) * 2], s2 = sums[(static_cast<size_t>(i) * nsc + s) * 2 + 1];
            rq.value[static_cast<size_t>(i) * nsc + s] = s1 / n;
            const double var = n > 1 ? std::max(s2 - s1 * s1 / n, 0.0) / (n - 1) : 0.0;
            rq.se[static_cast<size_t>(i) * nsc + s] = std::sqrt(var / n);
        }
    }
    float tables_ms = 0.f;
    hip_check(hipEventElapsedTime(&tables_ms, t0, t1), "elapsed");
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    rq.tables_ms = tables_ms;
    rq.kernel_ms = kernel_ms;
    rq.name_steps = name_steps;
    rq.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

}  // namespace

void check_book_request(const optlab_book_request& c) {
    if (c.n_names < 1 || c.n_names > kBookNames) invalid("n_names must be 1 to 64");
    if (c.n_steps < 1 || !(c.dt > 0)) invalid("n_steps must be >= 1 and dt > 0");
    if (!c.surface || !c.fwd || !c.disc || !c.trades || !c.spot_shock || !c.vol_shift || !c.skew || !c.value || !c.se)
        invalid("null array");
    if (c.n_trades < 1 || c.n_scen < 1) invalid("n_trades and n_scen must be >= 1");
    if (c.mode < 0 || c.mode > 2) invalid("mode must be 0 (one graph launch), 1 (a launch per trade) or 2 (one launch)");
    int n_x = -1;
    for (int a = 0; a < c.n_names; ++a) {
        const SurfaceRows& r = surface_rows(c.surface[a]);
        if (r.n_rows != c.n_steps) invalid("each surface needs one row per step of the book");
        if (n_x >= 0 && r.n_x != n_x) invalid("the surfaces must share n_x");
        n_x = r.n_x;
        for (int t = 0; t <= c.n_steps; ++t)
            if (!(c.fwd[static_cast<size_t>(a) * (c.n_steps + 1) + t] > 0)) invalid("forwards must be > 0");
    }
    for (int s = 0; s < c.n_scen; ++s)
        for (int a = 0; a < c.n_names; ++a)
            if (!(c.spot_shock[static_cast<size_t>(s) * c.n_names + a] > -1)) invalid("spot shocks must be > -1");
    for (int i = 0; i < c.n_trades; ++i) {
        const optlab_trade& t = c.trades[i];
        if (t.product < 0 || t.product > 3) invalid("product must be 0 to 3");
        if (t.n_names < 1 || t.n_names > kNames || (t.product != 3 && t.n_names != 1))
            invalid("a vanilla trade has one name, an autocall 1 to 5");
        for (int k = 0; k < t.n_names; ++k)
            if (t.names[k] < 0 || t.names[k] >= c.n_names) invalid("trade name out of range");
        if (t.n_steps < 1 || t.n_steps > c.n_steps) invalid("trade n_steps must be 1 to the book's n_steps");
        if (t.obs_every < 1 || t.n_steps % t.obs_every) invalid("obs_every must divide n_steps");
        if (t.n_pairs < 1 || t.n_pairs > (1ll << 30)) invalid("n_pairs must be 1 to 2^30");
        if (t.product == 2 && !(t.barrier > 0)) invalid("an up-and-out barrier must be > 0");
    }
}

void run_book(optlab_book_request& rq, int prec) {
    if (prec == 0)
        run_typed<double>(rq);
    else
        run_typed<float>(rq);
}

}  // namespace optlab_mi
