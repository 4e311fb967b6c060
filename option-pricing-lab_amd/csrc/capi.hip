// The C ABI of include/optlab_mi.h: checks (the reference's requests.cpp:140-219 messages), HIP device handling,
// the per-device scratch buffers, and the translation of exceptions into status codes plus a thread-local message.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace optlab_mi {

namespace {

thread_local std::string g_error;

struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::vector<std::vector<Buffer>> g_scratch;  // [device][slot]

int current_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    return dev;
}

int device_count_nothrow() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

void need_device() {
    if (device_count_nothrow() == 0) throw Error(OPTLAB_EDEVICE, "no HIP device");
}

template <typename F>
int guarded(F body) {
    try {
        body();
        return OPTLAB_OK;
    } catch (const Error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return OPTLAB_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return OPTLAB_EINTERNAL;
    }
}

int precision_of(int backend) {
    if (backend != 0 && backend != 1) invalid("backend must be 0 (double) or 1 (float)");
    return backend;
}

}  // namespace

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof buf, "HIP error in %s: %s", what, hipGetErrorString(e));
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? OPTLAB_ENOMEM : OPTLAB_EDEVICE, buf);
    }
}

void* scratch(int slot, size_t bytes) {
    const int dev = current_device();
    if (static_cast<int>(g_scratch.size()) <= dev) g_scratch.resize(dev + 1);
    auto& slots = g_scratch[dev];
    if (slots.empty()) slots.resize(kSlotCount);
    Buffer& b = slots[slot];
    if (b.bytes < bytes) {
        const size_t want = std::max(bytes, b.bytes + b.bytes / 2);  // grow by half at least: fewer reallocations
        if (b.ptr) hip_check(hipFree(b.ptr), "hipFree");
        b.ptr = nullptr;
        b.bytes = 0;
        hip_check(hipMalloc(&b.ptr, want), "hipMalloc");
        b.bytes = want;
    }
    return b.ptr;
}

hipStream_t engine_stream() {
    static hipStream_t s = [] {
        hipStream_t x = nullptr;
        hip_check(hipStreamCreate(&x), "hipStreamCreate");
        return x;
    }();
    return s;
}

int device_cu_count() {
    static int cached = 0;
    if (!cached) {
        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, current_device()), "attribute");
        cached = cus > 0 ? cus : 1;
    }
    return cached;
}

void check_paths_request(const optlab_paths_request& c) {  // requests.cpp:168-207
    if (c.n_names < 1 || c.n_names > 8) invalid("n_names must be 1 to 8");
    if (c.n_pairs < 1 || c.first_pair < 0) invalid("n_pairs must be >= 1 and first_pair >= 0");
    if (c.n_steps < 0 || c.n_obs < 0) invalid("n_steps and n_obs must be >= 0");
    if (c.n_steps > 0 && (!c.dt || !c.t0)) invalid("null time grid");
    if (!c.root || !c.kind || !c.cvar || !c.heston || !c.fwd || (c.n_obs && !c.obs_step)) invalid("null array");
    bool table = false;
    for (int a = 0; a < c.n_names; ++a) {
        const int k = c.kind[a];
        if (k < 0 || k > 3) invalid("kind must be 0 to 3");
        table |= k == 1 || k == 3;
        if ((k == 2 || k == 3) && !(c.heston[a * 5 + 1] > 0 && c.heston[a * 5 + 2] > 0 && c.heston[a * 5 + 3] >= 0 &&
                                   std::fabs(c.heston[a * 5 + 4]) <= 1))
            invalid("Heston parameters out of range");
    }
    if (table) {
        if (c.n_x < 2 || c.n_rows < 1 || !c.step_row || !c.row_lo || !c.row_inv_dx || !c.table)
            invalid("a tabulated variance needs n_x >= 2, rows and step_row");
        for (long long i = 0; i < static_cast<long long>(c.n_names) * c.n_steps; ++i) {
            const int k = c.kind[i / (c.n_steps ? c.n_steps : 1)];
            if ((k == 1 || k == 3) && (c.step_row[i] < 0 || c.step_row[i] >= c.n_rows))
                invalid("step_row out of range");
        }
    }
    int prev = 0;
    for (int j = 0; j < c.n_obs; ++j) {
        if (c.obs_step[j] < prev || c.obs_step[j] > c.n_steps) invalid("obs_step must be nondecreasing in [0, n_steps]");
        prev = c.obs_step[j];
    }
    if (c.mixer < 0 || c.mixer > 2 || (c.mixer == 1 && (!c.low_root || !(c.mix[4] > 0))))
        invalid("bad local correlation");
    // mixer 2 (an addition: the reference engine refuses the stochastic weight): w0, theta in [0, 1], nu >= 0,
    // |rho| <= 1, kappa >= 0
    if (c.mixer == 2 && (!c.low_root || !(c.mix[0] >= 0 && c.mix[0] <= 1) || !(c.mix[1] >= 0 && c.mix[1] <= 1) ||
                         !(c.mix[2] >= 0) || !(std::fabs(c.mix[3]) <= 1) || !(c.mix[4] >= 0)))
        invalid("bad stochastic mix");
    if (c.n_windows < 0 || c.n_windows > 2 || (c.n_windows && (!c.win_lo || !c.win_hi)))
        invalid("n_windows must be 0 to 2, with win_lo and win_hi");
    if (c.n_obs && !c.col_of) invalid("null col_of");
    if (c.n_keep < 0 || c.n_cols != c.n_keep + 2 * c.n_windows) invalid("n_cols must be n_keep + 2 n_windows");
    for (int j = 0; j < c.n_obs; ++j)
        if (c.col_of[j] < -1 || c.col_of[j] >= c.n_keep) invalid("col_of out of range");
    for (int w = 0; w < c.n_windows; ++w)
        if (c.win_lo[w] < 0 || c.win_hi[w] < c.win_lo[w] || c.win_hi[w] >= c.n_obs) invalid("bad window");
}

}  // namespace optlab_mi

using namespace optlab_mi;

extern "C" {

int optlab_abi_version(void) { return OPTLAB_MI_ABI_VERSION; }

const char* optlab_last_error(void) { return g_error.c_str(); }

int optlab_gpu_count(void) { return device_count_nothrow(); }

int optlab_has_hip(void) { return 1; }

int optlab_cpu_threads(void) { return static_cast<int>(std::thread::hardware_concurrency()); }

int optlab_gpu_name(char* buf, int len) {
    if (!buf || len <= 0) return OPTLAB_EINVAL;
    buf[0] = 0;
    if (device_count_nothrow() == 0) return OPTLAB_OK;
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) {
        (void)hipGetLastError();
        return OPTLAB_OK;
    }
    std::snprintf(buf, len, "%s (%s, %d CU)", p.name, p.gcnArchName, p.multiProcessorCount);
    return OPTLAB_OK;
}

int optlab_device_info(int device, optlab_device* out) {
    return guarded([&] {
        if (!out) invalid("null output");
        std::memset(out, 0, sizeof *out);
        if (device < 0 || device >= device_count_nothrow()) throw Error(OPTLAB_EDEVICE, "no such HIP device");
        hipDeviceProp_t p{};
        hip_check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
        std::snprintf(out->name, sizeof out->name, "%s", p.name);
        out->major = p.major, out->minor = p.minor, out->sm_count = p.multiProcessorCount;
        out->clock_mhz = p.clockRate / 1000;
        out->total_mem = static_cast<long long>(p.totalGlobalMem);
        out->integrated = p.integrated;
        out->kernel_timeout = p.kernelExecTimeoutEnabled;
        int prev = 0;
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        size_t fr = 0, tot = 0;
        if (hipSetDevice(device) == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess)
            out->free_mem = static_cast<long long>(fr);
        (void)hipSetDevice(prev);
        (void)hipGetLastError();
    });
}

int optlab_price(const optlab_request* rq, int backend, int n_threads, optlab_result* out) {
    (void)n_threads;
    return guarded([&] {
        if (!rq || !out) invalid("null request");
        check_price_request(*rq);
        const int prec = precision_of(backend);
        need_device();
        run_price(*rq, prec, *out);
    });
}

int optlab_paths(optlab_paths_request* rq, int backend, int n_threads) {
    (void)n_threads;
    return guarded([&] {
        if (!rq) invalid("null request");
        check_paths_request(*rq);
        if (!rq->out && rq->n_obs > 0) invalid("null output");
        const int prec = precision_of(backend);
        need_device();
        const auto t0 = std::chrono::steady_clock::now();
        const size_t n_out = static_cast<size_t>(rq->n_names) * rq->n_cols * 2 * rq->n_pairs;
        double* d_out = static_cast<double*>(scratch(kSlotOut, std::max<size_t>(n_out, 1) * sizeof(double)));
        rq->kernel_ms = launch_paths(*rq, prec, d_out, nullptr);
        if (n_out) hip_check(hipMemcpy(rq->out, d_out, n_out * sizeof(double), hipMemcpyDeviceToHost), "copy paths");
        rq->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int optlab_paths_dev(optlab_paths_request* rq, int backend, double* d_out, void* stream) {
    return guarded([&] {
        if (!rq) invalid("null request");
        check_paths_request(*rq);
        const size_t n_out = static_cast<size_t>(rq->n_names) * rq->n_cols * 2 * rq->n_pairs;
        if (!d_out && n_out) invalid("null output");
        const int prec = precision_of(backend);
        need_device();
        const auto t0 = std::chrono::steady_clock::now();
        rq->kernel_ms = launch_paths(*rq, prec, d_out, static_cast<hipStream_t>(stream));
        rq->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int optlab_script(optlab_paths_request* rq, int backend, const char* source, optlab_script_result* out) {
    return guarded([&] {
        if (!rq || !source || !out) invalid("null argument");
        check_paths_request(*rq);
        if (rq->n_pairs < 2) invalid("n_pairs must be >= 2");
        const int prec = precision_of(backend);
        need_device();
        std::memset(out, 0, sizeof *out);
        run_script(*rq, prec, source, *out);
    });
}

int optlab_leverage(optlab_leverage_request* rq, int backend, int n_threads) {
    (void)n_threads;
    return guarded([&] {
        if (!rq) invalid("null request");
        check_leverage_request(*rq);
        const int prec = precision_of(backend);
        need_device();
        run_leverage(*rq, nullptr, prec);
    });
}

int optlab_leverage_ex(optlab_leverage_request* rq, const optlab_leverage_options* opt, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        check_leverage_request(*rq);
        const int prec = precision_of(backend);
        need_device();
        run_leverage(*rq, opt, prec);
    });
}

int optlab_leverage_slices(optlab_leverage_request* rq, const optlab_leverage_options* opt,
                           const optlab_lv_slices* lv, int backend) {
    return guarded([&] {
        if (!rq || !lv) invalid("null request");
        check_leverage_request(*rq, false);
        const int prec = precision_of(backend);
        need_device();
        run_leverage(*rq, opt, prec, lv);
    });
}

int optlab_strip(optlab_strip_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        check_strip_request(*rq);
        const int prec = precision_of(backend);
        need_device();
        run_strip(*rq, prec);
    });
}

int optlab_strip_greeks(optlab_strip_request* rq, int backend, double* delta, double* delta_se, double* vega,
                        double* vega_se) {
    return guarded([&] {
        if (!rq) invalid("null request");
        if (!delta || !delta_se || !vega || !vega_se) invalid("null output");
        optlab_strip_request c = *rq;
        c.price = delta;  // the checks ask for the price's outputs, which this call does not write
        c.se = delta_se;
        check_strip_request(c);
        const int prec = precision_of(backend);
        need_device();
        run_strip_greeks(*rq, prec, delta, delta_se, vega, vega_se);
    });
}

int optlab_strip_noalloc(const optlab_strip_request* rq, int backend, double* price_dev, double* se_dev,
                         void* stream) {
    return guarded([&] {
        if (!rq) invalid("null request");
        if (!price_dev || !se_dev) invalid("null device output");
        optlab_strip_request c = *rq;
        c.price = price_dev;  // the checks ask for outputs; these are device arrays here
        c.se = se_dev;
        check_strip_request(c);
        const int prec = precision_of(backend);
        need_device();
        run_strip_noalloc(c, prec, price_dev, se_dev, static_cast<hipStream_t>(stream));
    });
}

int optlab_surface_upload(const optlab_surface_rows* rows, int* handle) {
    return guarded([&] {
        if (!rows || !handle) invalid("null argument");
        need_device();
        *handle = surface_upload(*rows);
    });
}

int optlab_surface_free(int handle) {
    return guarded([&] {
        need_device();
        surface_free(handle);
    });
}

int optlab_surface_table(int handle, const double* nodes, double parallel, double skew, double* out) {
    return guarded([&] {
        if (!out) invalid("null output");
        need_device();
        surface_table(handle, nodes, parallel, skew, out);
    });
}

int optlab_surface_node_grad(int handle, const double* tbar, double* out) {
    return guarded([&] {
        if (!tbar || !out) invalid("null argument");
        need_device();
        surface_node_grad(handle, tbar, out);
    });
}

int optlab_autocall(optlab_autocall_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        const int prec = precision_of(backend);
        need_device();
        check_autocall_request(*rq, false);
        run_autocall(*rq, prec, false);
    });
}

int optlab_adjoint(optlab_autocall_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        const int prec = precision_of(backend);
        need_device();
        check_autocall_request(*rq, true);
        run_autocall(*rq, prec, true);
    });
}

int optlab_autocall_ex(optlab_autocall_request* rq, const optlab_autocall_models* models, int adjoint, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        const int prec = precision_of(backend);
        need_device();
        check_autocall_request(*rq, adjoint != 0);
        if (models) check_autocall_models(*rq, *models);
        run_autocall(*rq, prec, adjoint != 0, models);
    });
}

int optlab_book(optlab_book_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        const int prec = precision_of(backend);
        need_device();
        check_book_request(*rq);
        run_book(*rq, prec);
    });
}

int optlab_lsm(optlab_lsm_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        const int prec = precision_of(backend);
        need_device();
        check_lsm_request(*rq);
        run_lsm(*rq, prec);
    });
}

int optlab_regress(optlab_regress_request* rq, int backend) {
    return guarded([&] {
        if (!rq) invalid("null request");
        (void)precision_of(backend);  // either backend: the regression is FP64
        need_device();
        check_regress_request(*rq);
        run_regress(*rq);
    });
}

int optlab_philox_dev(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, long long n,
                      unsigned* out_dev, void* stream) {
    return guarded([&] {
        if (n < 0 || (n > 0 && !out_dev)) invalid("bad output");
        need_device();
        launch_philox(c0, c1, c2, c3, k0, k1, n, out_dev, static_cast<hipStream_t>(stream));
    });
}

int optlab_normals_dev(long long first_pair, long long n_pairs, int n_normals, unsigned tag, unsigned long long seed,
                       int backend, double* out_dev, void* stream) {
    return guarded([&] {
        if (first_pair < 0 || n_pairs < 0 || n_normals < 0 || ((n_pairs > 0 && n_normals > 0) && !out_dev))
            invalid("bad request");
        const int prec = precision_of(backend);
        need_device();
        launch_normals(first_pair, n_pairs, n_normals, tag, seed, prec, out_dev, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
