// Fused script payoffs: a product written in the payoff language (packages/structured/src/optlab/structured/script/
// evaluator.py evaluates it in NumPy on paths brought to the host) is lowered by the host to HIP source and compiled
// here at run time by hipRTC for the current device. The generated kernel reads the path engine's output where the
// path kernel left it in HBM (names x columns x legs, structure of arrays: a wave reads 64 consecutive legs of a
// column) and reduces each workgroup's pair averages to one (sum, sum of squares) row; only those rows come back.
//
// The contract of the generated source: one kernel
//     extern "C" __global__ void script_payoff(const double* out, long long n_pairs, double* rows)
// launched with 256 lanes a workgroup and ceil(n_pairs / 256) workgroups; lane p values pair p (legs p and
// n_pairs + p) and workgroup b writes rows[2 b] (sum of its pairs' averages) and rows[2 b + 1] (sum of squares).
#include <hip/hiprtc.h>

#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace optlab_mi {

namespace {

constexpr int kScriptBlock = 256;

struct ScriptKernel {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    double compile_ms = 0.0;
};

std::string device_arch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop{};
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    return prop.gcnArchName;
}

// Compiled kernels by (device, source): a script is compiled once a process and device.
ScriptKernel& script_kernel(const std::string& source, bool& fresh) {
    static std::map<std::pair<int, std::string>, ScriptKernel> compiled;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    const auto key = std::make_pair(dev, source);
    auto it = compiled.find(key);
    fresh = it == compiled.end();
    if (!fresh) return it->second;
    const auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, source.c_str(), "script_payoff.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        throw Error(OPTLAB_EDEVICE, "hiprtcCreateProgram failed");
    const std::string arch = "--offload-arch=" + device_arch();
    // the library's own flags for device code: no packed FP32 forms
    const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, static_cast<int>(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        invalid("the script payoff does not compile: " + log);
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    std::vector<char> code(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    ScriptKernel k;
    hip_check(hipModuleLoadData(&k.module, code.data()), "hipModuleLoadData (script payoff)");
    hip_check(hipModuleGetFunction(&k.fn, k.module, "script_payoff"), "hipModuleGetFunction (script_payoff)");
    k.compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return compiled.emplace(key, k).first->second;
}

}  // namespace

void run_script(optlab_paths_request& rq, int prec, const char* source, optlab_script_result& res) {
    const auto wall0 = std::chrono::steady_clock::now();
    const long long n_pairs = rq.n_pairs;
    const size_t n_out = static_cast<size_t>(rq.n_names) * rq.n_cols * 2 * n_pairs;
    double* paths = static_cast<double*>(scratch(kSlotScript, n_out * sizeof(double)));
    bool fresh = false;
    ScriptKernel& k = script_kernel(source, fresh);
    const long long grid = (n_pairs + kScriptBlock - 1) / kScriptBlock;
    if (grid > 0x7fffffffll) invalid("too many pairs for one payoff launch");
    double* rows = static_cast<double*>(scratch(kSlotScriptRows, static_cast<size_t>(grid) * 2 * sizeof(double)));
    hipStream_t st = nullptr;
    res.paths_ms = launch_paths(rq, prec, paths, st);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, st), "event record");
    long long np = n_pairs;
    void* args[] = {&paths, &np, &rows};
    hip_check(hipModuleLaunchKernel(k.fn, static_cast<unsigned>(grid), 1, 1, kScriptBlock, 1, 1, 0, st, args, nullptr),
              "script payoff launch");
    hip_check(hipEventRecord(e1, st), "event record");
    std::vector<double> host(static_cast<size_t>(grid) * 2);
    hip_check(hipMemcpyAsync(host.data(), rows, host.size() * sizeof(double), hipMemcpyDeviceToHost, st), "copy rows");
    hip_check(hipStreamSynchronize(st), "script payoff");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the rows in their order, compensated (the sum does not depend on the launch's scheduling)
    double s = 0.0, sc = 0.0, q = 0.0, qc = 0.0;
    auto add = [](double& sum, double& c, double x) {
        const double t = sum + x;
        c += std::fabs(sum) >= std::fabs(x) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    };
    for (long long b = 0; b < grid; ++b) add(s, sc, host[2 * b]), add(q, qc, host[2 * b + 1]);
    const double n = static_cast<double>(n_pairs), mean = (s + sc) / n;
    const double var = (q + qc) / n - mean * mean;  // cpu_mc.cpp:14-18
    res.price = mean;
    res.se = n_pairs > 1 ? std::sqrt(std::fmax(var, 0.0) / (n - 1.0)) : 0.0;
    res.payoff_ms = ms;
    res.compile_ms = fresh ? k.compile_ms : 0.0;
    res.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

}  // namespace optlab_mi
