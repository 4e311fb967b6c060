"""Benchmark: Heston QE-M path-steps/s on one MI355X, FP32 paths against the FP64 mode of the same kernel.

Workload (one "step" = one pass of the hot path over one batch): a strip of 21 European calls (K = 70..130 step 3)
at 8 maturities (0.25..2 y) on 2^20 Heston paths (Andersen QE-M, psi_c 1.5; S0 100, v0 .04, kappa 1.5, theta .04,
xi .6, rho -.7, r .03, q 0) of 504 uniform steps to T = 2 y, FP32 path state with compensated FP64 payoff sums.
Inputs are parameters (nothing to stage in HBM); the outputs are 9 x 22 estimates.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--fp64-steps 3] [--no-cpu] [--adjoint-paths 524288]

Prints one JSON line: value = FP32 path-steps/s of the whole call (wall clock around the timed steps), the FP64
figure and the FP32/FP64 ratio beside it, the roofline of the strip kernel (its average duration from HIP events
recorded around it on its stream, by the library) and the CPU oracle timed on a bounded sample on this host; then the
second half of the metric, the worst-of autocall's adjoint Greeks against bump-and-revalue at matched standard
error ("adjoint")."""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "option-pricing-lab_amd"))

METRIC = "path-steps/s Heston QE FP32 vs FP64; AAD-vs-bump Greeks speedup at ≤4 SE"
N_PATHS = 1 << 20
N_STEPS = 504
T = 2.0
STRIKES = [70.0 + 3.0 * k for k in range(21)]
MATS = [0.25 * (j + 1) for j in range(8)]
# Algorithmic VALU lane-slots per path-step of FP32 QE-M (one leg, one step): 10 special functions (div, sqrt, log,
# erfc) at 4 slots, 28 other FP32 ops, and the amortised normals -- a quarter of a Philox4x32-10 block (40 32-bit
# multiplies at 4 slots + 60 integer ops) and half a Box-Muller (3 specials + 10 ops). 40 + 28 + 55 + 11 = 134.
SLOTS_PER_PATH_STEP = 134
# MI355X: 256 CUs x 64 lanes x 2.4 GHz = 3.93e13 unpacked VALU lane-slots/s, i.e. 78.6 TFLOP/s of FMA (the FP64
# vector peak and the unpacked FP32 peak; packed FP32 is not used).
LANE_SLOTS_PER_S = 256 * 64 * 2.4e9
PEAK_TFLOPS = 78.6
# HBM bytes a launch of the FP32 strip kernel moves, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on this
# workload (480.2 and 8192 KiB a dispatch), scaled by what the counters read for a 256 MiB device copy (x 2.0 and
# x 1.0): the constants and strike table read, one row of 8 maturities x 64 sums written per workgroup.
STRIP_TRAFFIC_BYTES = int((480.2 * 2.0 + 8192.0) * 1024)


# ---- the SLV particle workload (SSVI surface, Heston factor, uniform daily grid) --------------------------------

SLV_PILLARS = [0.05, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0]
# Algorithmic VALU lane-slots per particle-step of the SLV calibration (one particle, one step): half a Philox4x32-10
# block (20 32-bit multiplies at 4 slots + 30 integer ops = 110), half of a pair's Box-Muller (2 specials at 4 + 3
# = 11), the Euler step (the leverage lerp's division and 2 square roots at 4 slots + 26 FP32 ops = 38) and the
# binning (2 divisions at 4 + 12 FP32/FP64 ops = 20): 179. The per-step regression (81 points x 256 bins) adds
# under 2 % at 2^20 particles and is left out.
SLV_SLOTS_PER_PARTICLE_STEP = 179
# Algorithmic HBM bytes a step with the particles in registers: the bin totals (8192 fine counts x 4 B + 4 x 256
# fixed-point sums x 8 B) written once and read once, and the leverage row (81 x 2 doubles) written.
SLV_BYTES_PER_STEP = 2 * (8192 * 4 + 4 * 256 * 8) + 81 * 16
SLV_STRIKES = [70.0 + 3.0 * k for k in range(21)]
SLV_MATS = [0.375 * (j + 1) for j in range(8)]  # 8 maturities to 3 y


def slv_inputs():
    """The synthetic SLV case: power-law SSVI (theta_t = 0.04 t, rho -0.6, eta 1.2, gamma 0.5) on the pillars above,
    Heston factor (v0 .04, kappa 2, theta .04, xi .5, rho -.7), flat market S0 100, r .03, q 0."""
    import numpy as np

    from optlab.core.market import Market
    from optlab.core.models import HestonParams
    from optlab.core.ssvi import SSVI

    pillars = np.array(SLV_PILLARS)
    surface = SSVI.power_law(-0.6, 1.2, 0.5, pillars, 0.04 * pillars)
    return Market.flat(100.0, 0.03, 0.0), surface, HestonParams(0.04, 2.0, 0.04, 0.5, -0.7)


def slv_calibrate(n_particles, T, backend="gpu", n_bins=256, rng=17):
    """The leverage by particles on a uniform 365/y grid, 256 bins, bandwidth 0.25 std(ln S) N^-1/5."""
    from optlab import native

    market, surface, heston = slv_inputs()
    return native.calibrate_leverage(heston, surface, market, T, 1.0, n_particles, dt_max=1 / 365, dt_min=None,
                                     rng=rng, backend=backend, n_bins=n_bins, bandwidth=("std", 0.25))


def slv_reprice(model, n_paths, maturities, strikes, seed=23, backend="gpu"):
    """Calls on SLV paths (uniform daily steps) against the surface: price gaps, SEs and implied-vol gaps in bp."""
    import numpy as np

    from optlab import native
    from optlab.core.blackscholes import black_price, implied_vol

    market, surface, _ = slv_inputs()
    S = native.simulate(model, market, maturities, n_paths, seed=seed, dt_max=1 / 365, dt_min=None, backend=backend)
    K = np.asarray(strikes)
    h = S.shape[0] // 2
    out = {"price_gap": [], "se": [], "vol_bp": []}
    for j, T in enumerate(maturities):
        F, D = float(market.F(T)), float(market.D(T))
        pay = D * np.maximum(S[:, j][:, None] - K, 0.0)
        pay = 0.5 * (pay[:h] + pay[h:])
        p, se = pay.mean(axis=0), pay.std(axis=0, ddof=1) / np.sqrt(h)
        target_vol = surface.iv(np.log(K / F), T)
        target = black_price(F, K, T, target_vol, D)
        vol = np.array([implied_vol(pi, F, k, T, D) for pi, k in zip(p, K)])
        out["price_gap"].append(p - target)
        out["se"].append(se)
        out["vol_bp"].append((vol - target_vol) * 1e4)
    return {k: np.array(v) for k, v in out.items()}


def _request(backend_model, n_paths):
    from optlab.core.market import Market
    from optlab.core.models import HestonModel, HestonParams
    from optlab.native import gpu

    market = Market.flat(100.0, 0.03, 0.0)
    model = HestonModel(HestonParams(0.04, 1.5, 0.04, 0.6, -0.7))
    steps = [round(m / T * N_STEPS) for m in MATS]
    return gpu.strip_request(model, market, STRIKES, steps, N_STEPS, T, n_paths, 42)


def strip_vol_bp(p32, p64) -> float:
    """The largest |implied vol(FP32 price) - implied vol(FP64 price)| over the strip's 21 x 8 calls, in basis
    points (Black on the forward 100 e^{rT}, discount e^{-rT}); calls whose FP64 price has no implied vol are left
    out."""
    import numpy as np

    from optlab.core.blackscholes import implied_vol

    worst = 0.0
    for j, t in enumerate(MATS):
        F, D = 100.0 * math.exp(0.03 * t), math.exp(-0.03 * t)
        for q, K in enumerate(STRIKES):
            a, b = implied_vol(p32[j, q], F, K, t, D), implied_vol(p64[j, q], F, K, t, D)
            if np.isfinite(a) and np.isfinite(b):
                worst = max(worst, abs(a - b) * 1e4)
    return worst


def _time_gpu(lib, backend: int, steps: int, warmup: int):
    import torch

    rq, out = _request(backend, N_PATHS)
    for _ in range(warmup):
        assert lib.optlab_strip(C.byref(rq), backend) == 0, lib.optlab_last_error()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        assert lib.optlab_strip(C.byref(rq), backend) == 0, lib.optlab_last_error()
        kernel_ms.append(rq.kernel_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps * 1e3, sum(kernel_ms) / len(kernel_ms), out["price"].copy(), out["se"].copy()


def host_threads() -> dict:
    """The host threads this process may use: the CPUs of its affinity mask (what nproc reports), capped by a cgroup
    CPU quota where one is set; the CPU model beside it."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = min(affinity, max(1, math.ceil(quota))) if quota else affinity
    model = "unknown"
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": threads, "nproc": affinity, "cgroup_cpu_quota": quota, "cpu_model": model}


def _time_cpu(threads: int, warmup_s: float = 2.0, min_rep_s: float = 0.3, reps: int = 5):
    """The oracle (FP64, the reference CPU engine's arithmetic) on a bounded sample, the same strip on 2^18 paths, in
    the protocol of the reference's benchmark (packages/native/cuda/bench/bench.cpp:117-123,181-204): warm-up for
    at least 2 s, then 5 repetitions of at least 300 ms each; (median, min, max) path-steps/s."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import oracle

    n = 1 << 18
    rq, out = _request(0, n)  # `out` owns the arrays rq points to
    lib = oracle()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < warmup_s:  # page faults, thread start, clocks
        assert lib.oracle_strip(C.byref(rq), threads) == 0
    rates = []
    for _ in range(reps):
        calls, t0 = 0, time.perf_counter()
        while True:
            assert lib.oracle_strip(C.byref(rq), threads) == 0
            calls += 1
            dt = time.perf_counter() - t0
            if dt >= min_rep_s:
                break
        rates.append(calls * n * N_STEPS / dt)
    rates.sort()
    return rates[len(rates) // 2], rates[0], rates[-1], n


# ---- the worst-of autocall adjoint workload ----------------------------------------------------------------------

AC_SCALES, AC_RHOS = (0.03, 0.04, 0.05), (-0.5, -0.6, -0.7)


def autocall_note():
    """Worst-of Phoenix on 3 names under Dupire local vol from SSVI surfaces (theta_t = s t, s .03/.04/.05, rho
    -.5/-.6/-.7, eta 1.2, gamma .5, pillars 0.05..3 y), correlation 0.6, 3 y quarterly, 756 daily steps, autocall
    100 %, coupon 8 % p.a. at a 70 % barrier with memory, knock-in put at 60 % on every close, ramps of 0.5 %;
    flat markets S0 100, r .03, q 0."""
    import numpy as np

    from optlab.core.market import Market
    from optlab.core.ssvi import SSVI
    from optlab.native.adjoint import WorstOfAutocall

    pillars = np.array(SLV_PILLARS)
    surfaces = [SSVI.power_law(r, 1.2, 0.5, pillars, s * pillars) for s, r in zip(AC_SCALES, AC_RHOS)]
    corr = np.full((3, 3), 0.6) + 0.4 * np.eye(3)
    return WorstOfAutocall(surfaces, [Market.flat(100.0, 0.03, 0.0)] * 3, corr)


def bench_adjoint(n_paths: int, seed: int = 42, reps: int = 3, bumps: bool = True) -> dict:
    """Adjoint Greeks (3 deltas + 240 node vegas) against bump-and-revalue of all 486 bumps on common random numbers,
    both in FP32 on the same engine; the speedup at matched standard error is, per node whose vega is beyond 3 SE,
    t_bump / t_adjoint times (SE_bump / SE_adjoint)^2 (the paths the bumps need for the adjoint's SE): its median,
    10th percentile and worst node."""
    import numpy as np

    note = autocall_note()
    chosen = note.adjoint(n_paths, seed).seconds["checkpoint_every"]  # warm-up; the autotuner picks the interval
    note.price(n_paths, seed)
    by_interval = {}
    for k in (4, 8):  # the candidates for 3 names, each timed once more (forward + backward stage, device events)
        t = note.adjoint(n_paths, seed, checkpoint_every=k).seconds
        by_interval[str(k)] = t["forward_ms"] + t["backward_ms"]
    walls, fwd, bwd = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        g = note.adjoint(n_paths, seed)
        walls.append(time.perf_counter() - t0)
        fwd.append(g.seconds["forward_ms"])
        bwd.append(g.seconds["backward_ms"])
    t_adj = sorted(walls)[reps // 2]
    if not bumps:
        note.close()
        return {"adjoint_ms": t_adj * 1e3, "adjoint_forward_kernel_ms": float(np.median(fwd)),
                "adjoint_backward_ms": float(np.median(bwd)), "checkpoint_every": chosen,
                "device_ms_by_checkpoint_interval": by_interval, "price": g.price, "se": g.se}
    t0 = time.perf_counter()
    b = note.bump(n_paths, seed)
    t_bump = time.perf_counter() - t0
    live = np.abs(g.vega) > 3 * g.vega_se
    node_ratio = (b.vega_se[live] / g.vega_se[live]) ** 2 if live.any() else np.array([np.nan])
    ratio = float(np.median(node_ratio))
    moving = np.hypot(g.vega_se, b.vega_se) > 0  # nodes beyond the last date move nothing: both vegas are 0
    z = np.abs(g.vega - b.vega)[moving] / np.hypot(g.vega_se, b.vega_se)[moving]
    zd = np.abs(g.delta - b.delta) / np.hypot(g.delta_se, b.delta_se)
    name_steps = n_paths * note.n_steps * note.n_names
    note.close()
    return {
        "workload": "worst-of Phoenix, 3 names, local vol from 3 SSVI surfaces, corr 0.6, 3 y quarterly, 756 daily "
                    f"steps, {n_paths} paths, FP32 paths; 3 deltas + 3 x 80 node vegas",
        "price": g.price,
        "se": g.se,
        "adjoint_ms": t_adj * 1e3,
        "adjoint_forward_kernel_ms": float(np.median(fwd)),
        "adjoint_backward_ms": float(np.median(bwd)),
        "checkpoint_every": chosen,
        "device_ms_by_checkpoint_interval": by_interval,
        "bump_ms": t_bump * 1e3,
        "revaluations": b.seconds["revaluations"],
        "speedup_time": t_bump / t_adj,
        "se_ratio_sq_median": ratio,
        "speedup_matched_se": t_bump / t_adj * ratio,
        "speedup_matched_se_worst_node": t_bump / t_adj * float(np.min(node_ratio)),
        "speedup_matched_se_p10_node": t_bump / t_adj * float(np.percentile(node_ratio, 10)),
        "live_nodes": int(live.sum()),
        "live_nodes_at_10x_matched_se": float(np.mean(t_bump / t_adj * node_ratio >= 10)),
        "moving_nodes": int(moving.sum()),
        "moving_nodes_within_4se": float(np.mean(z < 4)),
        "max_z_vega": float(z.max()),
        "max_z_delta": float(zd.max()),
        "delta": g.delta.tolist(),
        "delta_se": g.delta_se.tolist(),
        "forward_name_steps_per_s": name_steps / (float(np.median(fwd)) * 1e-3),
    }


def bench_adjoint_models(n_paths: int = 1 << 19, seed: int = 42, reps: int = 3) -> dict:
    """The autocall's adjoint with other models a name: the note of `autocall_note` with name 0 under SLV (its
    leverage calibrated by the particle method on name 0's surface, 2^18 particles), name 1 under local vol and name
    2 under Heston; device time of the forward and backward stages, against the local-vol note's."""
    import numpy as np

    from optlab import native
    from optlab.core.market import Market
    from optlab.core.models import HestonModel, HestonParams
    from optlab.native.adjoint import WorstOfAutocall

    base = autocall_note()
    surfaces = [s.surface for s in base.surfaces]
    market = Market.flat(100.0, 0.03, 0.0)
    heston = HestonParams(0.04, 2.0, 0.04, 0.5, -0.7)
    slv = native.calibrate_leverage(heston, surfaces[0], market, 3.0, 1.0, 1 << 18, rng=17, backend="gpu",
                                    n_bins=256, bandwidth=("std", 0.25))
    models = [slv, None, HestonModel(HestonParams(0.05, 2.0, 0.05, 0.5, -0.6))]
    note = WorstOfAutocall(surfaces, [market] * 3, base.root @ base.root.T, models=models, same=True)
    out = {}
    for name, nt in (("slv_lv_heston", note), ("local_vol", base)):
        nt.adjoint(n_paths, seed)  # warm-up, the autotuner's choice
        dev, walls = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            g = nt.adjoint(n_paths, seed)
            walls.append(time.perf_counter() - t0)
            dev.append(g.seconds["forward_ms"] + g.seconds["backward_ms"])
        out[name] = {"device_ms": float(np.median(dev)), "wall_ms": float(np.median(walls)) * 1e3,
                     "checkpoint_every": g.seconds["checkpoint_every"], "price": g.price,
                     "live_vegas": int((np.abs(g.vega) > 3 * g.vega_se).sum())}
        nt.close()
    out["workload"] = ("the worst-of phoenix of the adjoint block with name 0 under SLV (leverage calibrated on its "
                       f"surface), name 1 local vol, name 2 Heston; {n_paths} FP32 paths; 3 deltas + node vegas of "
                       "the SLV and local-vol names")
    return out


def bench_strip_greeks(reps: int = 5) -> dict:
    """The strip's sensitivities (delta and d/dv0 of all 21 x 8 calls, optlab_strip_greeks) on the benchmark
    strip's paths, against the strip itself and against central bumps of S0 and v0 on common random numbers (four
    more strips)."""
    import numpy as np
    import torch

    from optlab.native import _abi

    lib = _abi.load()
    rq, out = _request(1, N_PATHS)
    shape = out["price"].shape
    d, dse, g, gse = (np.empty(shape) for _ in range(4))

    def greeks(backend):
        assert lib.optlab_strip_greeks(C.byref(rq), backend, _abi.ptr(d), _abi.ptr(dse), _abi.ptr(g),
                                       _abi.ptr(gse)) == 0, lib.optlab_last_error()
        return rq.kernel_ms

    def strip():
        assert lib.optlab_strip(C.byref(rq), 1) == 0, lib.optlab_last_error()
        return rq.kernel_ms

    greeks(1), strip()
    torch.cuda.synchronize()
    t_g, k_g, t_s = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        k_g.append(greeks(1))
        t_g.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        strip()
        t_s.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    greeks(0)
    t64 = time.perf_counter() - t0
    wall, price = float(np.median(t_g)) * 1e3, float(np.median(t_s)) * 1e3
    return {"workload": "delta and d/dv0 of the strip's 21 x 8 calls on its 2^20 FP32 paths x 504 steps (a tangent "
                        "through every QE-M step)",
            "greeks_ms": wall, "greeks_kernel_ms": float(np.median(k_g)), "greeks_fp64_ms": t64 * 1e3,
            "strip_ms": price, "greeks_over_strip": wall / price,
            "speedup_vs_central_bumps": 4 * price / wall,
            "path_steps_per_s": N_PATHS * N_STEPS / (wall * 1e-3),
            "atm_1y": {"delta": float(d[3, 10]), "delta_se": float(dse[3, 10]), "dv0": float(g[3, 10]),
                       "dv0_se": float(gse[3, 10])}}


def bench_script(n_paths: int = 1 << 20, reps: int = 5, cpu: bool = True) -> dict:
    """The phoenix of the preset list written in the payoff language (3 local-vol names, 3 y quarterly, knock-in on
    756 daily closes), priced by optlab_script: the paths into HBM, the payoff compiled by hipRTC and run over them,
    pair sums back. The CPU baseline is the reference's way, the NumPy evaluator (restated in oracle/script_oracle.py)
    on the same paths copied to the host, on a bounded sample."""
    import numpy as np

    from optlab.core.market import Market
    from optlab.core.models import LocalVolModel
    from optlab.native.script import ScriptPricer
    from optlab.structured import script as sc
    from optlab.structured.multiasset import Basket

    base = autocall_note()
    names = ("A", "B", "C")
    market = Market.flat(100.0, 0.03, 0.0)
    basket = Basket(names, {n: LocalVolModel(s.surface) for n, s in zip(names, base.surfaces)},
                    {n: market for n in names}, base.root @ base.root.T)
    base.close()
    pricer = ScriptPricer(sc.phoenix_script(names, 3.0, 0.25, daily_ki=756), basket, smoothing=0.005)
    first = pricer.price(n_paths, 7, "fp32")
    runs = [pricer.price(n_paths, 7, "fp32") for _ in range(reps)]
    med = {k: float(np.median([r[k] for r in runs])) for k in ("paths_ms", "payoff_ms", "total_ms")}
    out = {"workload": f"phoenix script, 3 local-vol names, 3 y quarterly, daily knock-in window; {n_paths} FP32 "
                       "paths, payoff in FP64 on the device",
           "price": runs[0]["price"], "se": runs[0]["se"], "compile_ms_first_call": first["compile_ms"], **med,
           "name_steps_per_s": n_paths * 756 * 3 / (med["total_ms"] * 1e-3)}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import script_oracle

        sample = 1 << 16
        t0 = time.perf_counter()
        spots, wins = pricer.paths(sample, 7, "fp32")
        t_paths = time.perf_counter() - t0
        t0 = time.perf_counter()
        values = script_oracle.leg_values(pricer.program, None, 0.005, spots, wins, lambda t: float(market.D(t)))
        t_eval = time.perf_counter() - t0
        price, _ = script_oracle.pair_mean(values)
        out["cpu_baseline"] = {"value": t_eval * 1e3 * n_paths / sample, "unit": "ms", "cores": 1, "kind": "port",
                               "sample": f"the evaluator in NumPy on {sample} of the same paths brought to the host "
                                         f"({t_paths * 1e3:.1f} ms to simulate and copy them), scaled to {n_paths}",
                               "price_on_sample": price}
    return out


def bench_book(path_scale: int = 1, n_scen: int = 250) -> dict:
    """The batched book: 512 trades (seed 7) x 250 scenarios plus the base, 2^14-2^18 paths by product (divided by
    `path_scale`); one persistent launch captured in a HIP graph against the per-trade loop, same numbers. The first
    batched call runs the autotuner (six candidates on one scenario) before its timed launch; `batched_ms` is the
    launch's device time."""
    import numpy as np

    from optlab.risk.book import pnl, revalue, synthetic_book, synthetic_scenarios, var_es

    small = synthetic_book(path_scale=256)  # warm-up: the kernels loaded, the surfaces' tables built once
    revalue(small, synthetic_scenarios(len(small.surfaces), 4))
    book = synthetic_book(path_scale=path_scale)
    scen = synthetic_scenarios(len(book.surfaces), n_scen)
    book.device_surfaces()
    t0 = time.perf_counter()
    rb = revalue(book, scen, mode="batched")
    wall_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    rl = revalue(book, scen, mode="loop", block=rb.block, pairs_per_lane=rb.pairs_per_lane)
    wall_l = time.perf_counter() - t0
    same = bool(np.array_equal(rb.value, rl.value))
    risk = var_es(pnl(book, rb))
    return {
        "workload": f"512 trades (40 % European, 20 % Asian, 20 % daily up-and-out, 20 % worst-of autocall on 1-5 "
                    f"names; 0.25-5 y; 2^14-2^18 paths by product{'' if path_scale == 1 else f' / {path_scale}'}) x "
                    f"{n_scen} scenarios + base, local vol from 5 SSVI surfaces, FP32 paths",
        "total_paths_per_scenario": int(sum(t.n_paths for t in book.trades)),
        "name_steps": rb.name_steps,
        "batched_ms": rb.kernel_ms,
        "batched_wall_ms": wall_b * 1e3,
        "batched_name_steps_per_s": rb.name_steps / (rb.kernel_ms * 1e-3),
        "loop_ms": rl.kernel_ms,
        "loop_wall_ms": wall_l * 1e3,
        "loop_name_steps_per_s": rl.name_steps / (rl.kernel_ms * 1e-3),
        "batched_over_loop": rl.kernel_ms / rb.kernel_ms,
        "tables_ms": rb.tables_ms,
        "block": rb.block,
        "pairs_per_lane": rb.pairs_per_lane,
        "latency_ms_batched": {"p50": float(np.percentile(rb.finish_ms, 50)), "p99": float(np.percentile(rb.finish_ms, 99))},
        "latency_ms_loop": {"p50": float(np.percentile(rl.finish_ms, 50)), "p99": float(np.percentile(rl.finish_ms, 99))},
        "var99": risk["var"],
        "es975": risk["es"],
        "batched_equals_loop": same,
    }


def bench_slv(n_particles: int = 1 << 20, T: float = 3.0, n_paths: int = 1 << 20) -> dict:
    """The SLV leverage by particles (2^20 particles, 365 steps a year to 3 y, 256 bins, bandwidth 0.25 std(ln S)
    N^-1/5, FP64 leverage division) and the repricing of the 21 x 8 grid on fresh FP32 paths."""
    import numpy as np

    slv_calibrate(1 << 14, 0.25)  # warm-up
    t0 = time.perf_counter()
    model = slv_calibrate(n_particles, T)
    t_cal = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = slv_reprice(model, n_paths, SLV_MATS, SLV_STRIKES)
    t_rep = time.perf_counter() - t0
    n_steps = int(round(T * 365))
    inner = slice(3, 18)  # K 79..121
    return {
        "workload": f"SLV particles: SSVI theta .04 t, rho -.6, eta 1.2, gamma .5; Heston .04/2/.04/.5/-.7; {n_particles} "
                    f"particles x {n_steps} steps, 256 bins, bandwidth 0.25 std(ln S) N^-1/5; reprice 21 x 8 calls on "
                    f"{n_paths} FP32 paths",
        "calibration_s": t_cal,
        "particle_steps_per_s": n_particles * n_steps / t_cal,
        "roofline": {
            "bound": "latency (two grid-wide waits and two exchanges of the bins a step)",
            "valu_lane_slots_per_particle_step": SLV_SLOTS_PER_PARTICLE_STEP,
            "valu_frac": SLV_SLOTS_PER_PARTICLE_STEP * n_particles * n_steps / t_cal / LANE_SLOTS_PER_S,
            "hbm_bytes_per_step": SLV_BYTES_PER_STEP,
            "hbm_GB_per_s": SLV_BYTES_PER_STEP * n_steps / t_cal / 1e9,
            "hbm_frac": SLV_BYTES_PER_STEP * n_steps / t_cal / 8e12,
            "us_per_step": t_cal / n_steps * 1e6,
        },
        "reprice_s": t_rep,
        "max_abs_vol_bp_k79_121": float(np.nanmax(np.abs(r["vol_bp"][:, inner]))),
        "max_abs_vol_bp_all": float(np.nanmax(np.abs(r["vol_bp"]))),
        "max_price_gap_in_se": float(np.nanmax(np.abs(r["price_gap"]) / r["se"])),
        "vol_bp_by_maturity_k79_121": [float(np.nanmax(np.abs(v[inner]))) for v in r["vol_bp"]],
    }


# The reference's published GPU kernel times (packages/native/cuda/README.md:26-40, RTX 2070 Super, float paths)
# on its own benchmark requests (bench/bench.cpp:67-77, params.hpp:10-49: S0 100, r .04, q .01, T 1, sigma .2, the
# SPX Heston parameters, calls at K 100, barrier H 80, seed 20261029), and its SLV leverage time
# (ACCELERATION.md:36: 1 y, 200 000 particles, GPU).
REFERENCE_KERNEL_MS = {
    "bs_european_1e8x1": 0.98,
    "bs_down_and_out_1e7x252": 8.52,
    "heston_european_1e7x252": 92.2,
    "heston_down_and_out_1e7x252": 91.2,
}
REFERENCE_SLV_1Y_200K_S = 0.16


def bench_reference_workloads(reps: int = 5) -> dict:
    """The reference's own benchmark workloads, FP32 paths as its GPU engine: the kernel time by the library's HIP
    events (median of reps after a warm-up) against the published one, and the SLV leverage of 1 y on 200 000
    particles (the reference's 400 bins and bandwidth rule; on this benchmark's SSVI case, the reference's surface
    for that row not being among its inputs) against the published wall time."""
    import numpy as np

    from optlab import native
    from optlab.core.market import Market
    from optlab.core.models import BlackScholesModel, HestonModel, HestonParams
    from optlab.native import gpu

    market = Market.flat(100.0, 0.04, 0.01)
    bs = BlackScholesModel(0.20)
    hes = HestonModel(HestonParams(0.022846642482989885, 3.0892635624407627, 0.05372458999932075, 1.307745170926575,
                                   -0.6879606734541107))
    cases = {  # name: (model, product, n_steps, n_paths)
        "bs_european_1e8x1": (bs, 0, 1, 10**8),
        "bs_down_and_out_1e7x252": (bs, 2, 252, 10**7),
        "heston_european_1e7x252": (hes, 0, 252, 10**7),
        "heston_down_and_out_1e7x252": (hes, 2, 252, 10**7),
    }
    out = {}
    for name, (model, product, n_steps, n_paths) in cases.items():
        rq = gpu.request(model, market, 1.0, "call", product, n_steps, n_paths, 20261029, 100.0, 80.0)
        gpu.price(rq, "gpu")
        runs = [gpu.price(rq, "gpu") for _ in range(reps)]
        k = float(np.median([r.kernel_ms for r in runs]))
        ref = REFERENCE_KERNEL_MS[name]
        out[name] = {"kernel_ms": k, "total_ms": float(np.median([r.total_ms for r in runs])),
                     "path_steps_per_s": n_paths * n_steps / (k * 1e-3), "reference_kernel_ms": ref,
                     "speedup_vs_reference_gpu": ref / k, "price": runs[0].price[0], "se": runs[0].se[0]}
    market_s, surface, heston = slv_inputs()
    native.calibrate_leverage(heston, surface, market_s, 0.25, 1.0, 1 << 14, rng=17, backend="gpu")  # warm-up
    t0 = time.perf_counter()
    native.calibrate_leverage(heston, surface, market_s, 1.0, 1.0, 200_000, rng=17, backend="gpu")
    t = time.perf_counter() - t0
    out["slv_leverage_1y_200k"] = {"calibration_s": t, "reference_s": REFERENCE_SLV_1Y_200K_S,
                                   "speedup_vs_reference_gpu": REFERENCE_SLV_1Y_200K_S / t}
    out["hardware_of_the_reference"] = "RTX 2070 Super (TU104, 40 SMs, FP32 paths)"
    return out


def bench_fixings(n_paths: int = 1 << 18, reps: int = 5) -> dict:
    """The fixing-heavy pass of the path engine: Black-Scholes (S0 100, r .03, q .01, sigma .2), 252 daily steps,
    every close kept, 2^18 paths, output left in HBM (optlab_paths_dev): 8 B written per path, close and name, so the
    roofline is HBM write bandwidth."""
    import numpy as np
    import torch

    from optlab import native
    from optlab.core.market import Market
    from optlab.core.models import BlackScholesModel

    market = Market.flat(100.0, 0.03, 0.01)
    times = [k / 252 for k in range(1, 253)]
    p = native.plan(BlackScholesModel(0.2), market, times)
    out = {}
    for backend in ("fp32", "fp64"):
        buf, _ = native.run_device(p, n_paths // 2, 42, backend=backend)  # warm-up, and the output buffer
        ms = []
        for _ in range(reps):
            _, t = native.run_device(p, n_paths // 2, 42, backend=backend, out=buf)
            ms.append(t["kernel_ms"])
        torch.cuda.synchronize()
        k = float(np.median(ms))
        bytes_out = buf.numel() * 8
        out[backend] = {"kernel_ms": k, "written_GB": bytes_out / 1e9, "write_GB_per_s": bytes_out / (k * 1e-3) / 1e9,
                        "frac_of_8TBps": bytes_out / (k * 1e-3) / 8e12,
                        "path_steps_per_s": n_paths * len(times) / (k * 1e-3)}
        del buf
    out["workload"] = (f"Black-Scholes daily closes: {n_paths} paths x 252 steps, every close kept (1 x 252 x "
                       f"{n_paths} doubles in HBM)")
    return out


def bench_lsm(n_paths: int = 1 << 20, n_ex: int = 50, reps: int = 5, cpu: bool = True) -> dict:
    """American puts by Longstaff-Schwartz on the device: the first row of Longstaff and Schwartz (2001), table 1
    (S0 36, K 40, r .06, sigma .2, T 1, 50 exercise dates; Black-Scholes paths of 300 steps, 6 per date), n_paths
    paths for the rule and n_paths more for the estimate, in FP32 and FP64; and a Heston put (S0 100, K 110,
    the basis with the variance factor). Medians of reps calls after a warm-up: the whole call, and the library's
    HIP events around the paths, the backward sweep of regressions and the pass that applies the rule. With cpu,
    the oracle's NumPy regression (normal equations, the reference's algorithm) on the same FP64 paths, on a sample
    of 2^17 paths, as the CPU baseline of the sweep."""
    import numpy as np

    from optlab import native
    from optlab.core.market import Market
    from optlab.core.models import BlackScholesModel, HestonModel, HestonParams

    cases = {"bs_put_ls_table1": (BlackScholesModel(0.2), Market.flat(36.0, 0.06, 0.0), 40.0),
             "heston_put_var_basis": (HestonModel(HestonParams(0.04, 1.5, 0.04, 0.6, -0.7)),
                                      Market.flat(100.0, 0.03, 0.0), 110.0)}
    out = {}
    for name, (model, market, strike) in cases.items():
        for backend in ("fp32", "fp64"):
            native.american_lsm(model, market, strike, 1.0, "put", n_ex, n_paths, rng=1, backend=backend)
            runs = [native.american_lsm(model, market, strike, 1.0, "put", n_ex, n_paths, rng=1, backend=backend)
                    for _ in range(reps)]
            med = {k: float(np.median([r.timing[k] for r in runs]))
                   for k in ("total_ms", "paths_ms", "regression_ms", "apply_ms")}
            med.update(price=runs[0].price, stderr=runs[0].stderr,
                       path_dates_per_s=2 * n_paths * n_ex / (med["total_ms"] * 1e-3))
            out[f"{name}_{backend}"] = med
    out["workload"] = (f"American put, {n_ex} exercise dates, {n_paths} paths for the rule + {n_paths} for the "
                       "estimate, FP64 normal equations on the device")
    out["issuer_call_regression"] = bench_regress(cpu=cpu)
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import lsm_oracle

        model, market, strike = cases["bs_put_ls_table1"]
        times = np.linspace(1.0 / n_ex, 1.0, n_ex)
        p = native.plan(model, market, times)
        n_cpu = 1 << 17
        S1 = native.run(p, n_cpu // 2, native.seed64(1), backend="fp64")[0][0]
        disc = market.D(times)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            lsm_oracle.lsm_rule(S1, None, disc, strike)
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        gpu = out["bs_put_ls_table1_fp64"]["regression_ms"]
        out["cpu_baseline"] = {"value": n_cpu * (n_ex - 1) / (ms * 1e-3), "unit": "path-dates/s of the sweep",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle/lsm_oracle.py lsm_rule (NumPy) on {n_cpu} FP64 paths x {n_ex} dates"}
        out["sweep_gpu_path_dates_per_s"] = n_paths * (n_ex - 1) / (gpu * 1e-3)
    return out


def bench_regress(n: int = 100_000, n_eval: int = 1 << 20, reps: int = 5, cpu: bool = True) -> dict:
    """The issuer's call regression of a worst-of note (exercise.py:191-238): 3 names' spots at a call date on n
    pilot paths (lognormal, correlation 0.6), the terms exercise.py's rules give them (the quadratic and 5 hinges a
    name: 25), FP64 normal equations on the device, and the fitted continuation on n_eval states. Medians of reps
    calls after a warm-up; with cpu, the reference's np.linalg.lstsq on the oracle's terms of the same rows (and
    the terms of the n_eval states times the coefficients) timed beside it."""
    import numpy as np

    from optlab import native

    from optlab.structured.exercise import terms

    rng = np.random.default_rng(7)
    c = np.full((3, 3), 0.6) + 0.4 * np.eye(3)
    s = np.exp(0.25 * rng.standard_normal((n + n_eval, 3)) @ np.linalg.cholesky(c).T - 0.03)
    x, ev = s[:n], s[n:]
    w = x.min(axis=1)
    y = 0.9 + 0.05 * np.minimum(w, 1.2) - 0.8 * np.maximum(0.6 - w, 0.0) + 0.02 * rng.standard_normal(n)
    mean, scale = x.mean(axis=0), x.std(axis=0)
    pw, kn = terms((x - mean) / scale)
    native.regress(x, y, mean, scale, pw, kn, eval_state=ev)
    runs = [native.regress(x, y, mean, scale, pw, kn, eval_state=ev) for _ in range(reps)]
    out = {k: float(np.median([r.timing[k] for r in runs])) for k in ("fit_ms", "eval_ms", "total_ms")}
    out.update(terms=len(pw) + len(kn), rows=n, eval_rows=n_eval)
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import regress_oracle as ro

        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            beta = ro.fit(ro.features(x, mean, scale, pw, kn), y, "lstsq")
            ro.features(ev, mean, scale, pw, kn) @ beta
            ts.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": float(np.median(ts)) * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "oracle/regress_oracle.py terms + np.linalg.lstsq (the reference's solve) "
                                         "on the same rows, and the continuation on the same states"}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-paths", type=int, default=1 << 19,
                    help="paths of the worst-of autocall adjoint-vs-bump comparison (0: skip it)")
    ap.add_argument("--book-scale", type=int, default=1,
                    help="divide every trade's paths of the batched book by this (0: skip the book)")
    ap.add_argument("--slv-particles", type=int, default=1 << 20, help="SLV particles (0: skip the SLV block)")
    ap.add_argument("--only", choices=("adjoint", "adjoint_models", "book", "slv", "fixings", "lsm", "reference",
                                       "strip_greeks", "script"), default=None,
                    help="run only this block and print its JSON (for profiling runs; the benchmark line needs all)")
    ap.add_argument("--no-bumps", action="store_true", help="with --only adjoint: skip the bump-and-revalue comparator")
    ap.add_argument("--no-blocks", action="store_true",
                    help="the strip only (the line's value, roofline and CPU baseline), none of the other workloads' "
                         "blocks: for profiling runs of the strip kernel")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed region, one device-to-device copy of this many MiB: the known-size kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path (the reference has none either); "
              "the per-GPU throughput is the figure, at --gpus 1", file=sys.stderr)
        return 2

    import torch

    from optlab.native import _abi

    lib = _abi.load()
    if lib.optlab_gpu_count() < 1 or not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 1
    torch.cuda.init()
    if args.only:
        block = {"adjoint": lambda: bench_adjoint(args.adjoint_paths, bumps=not args.no_bumps),
                 "book": lambda: bench_book(args.book_scale),
                 "slv": lambda: bench_slv(args.slv_particles),
                 "fixings": lambda: bench_fixings(),
                 "lsm": lambda: bench_lsm(cpu=not args.no_cpu),
                 "reference": lambda: bench_reference_workloads(),
                 "adjoint_models": lambda: bench_adjoint_models(),
                 "strip_greeks": lambda: bench_strip_greeks(),
                 "script": lambda: bench_script(cpu=not args.no_cpu)}[args.only]()
        print(json.dumps({args.only: block}), flush=True)
        return 0
    ms32, kern32, p32, se32 = _time_gpu(lib, 1, args.steps, args.warmup)
    ms64, kern64, p64, se64 = _time_gpu(lib, 0, max(1, args.fp64_steps), 1)
    work = N_PATHS * N_STEPS
    v32, v64 = work / (ms32 * 1e-3), work / (ms64 * 1e-3)
    k = len(STRIKES)
    gap = (p32[:, :k] - p64[:, :k]) / se64[:, :k]
    achieved_slots = SLOTS_PER_PATH_STEP * work / (kern32 * 1e-3)
    achieved = 2 * achieved_slots / 1e12  # one lane-slot = one FMA-equivalent = 2 flops
    line = {
        "metric": METRIC,
        "value": v32,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Heston QE-M strip: 21 calls x 8 maturities on 2^20 paths x 504 steps (T 2y), FP32 paths, "
                        "compensated FP64 sums; FP64 mode of the same kernel timed beside it",
            "paths": N_PATHS,
            "time_steps": N_STEPS,
            "strikes": k,
            "maturities": len(MATS),
            "psi_c": 1.5,
        },
        "fp64": {"value": v64, "ms_per_step": ms64, "kernel_ms": kern64, "steps": max(1, args.fp64_steps)},
        "fp32_over_fp64": v32 / v64,
        "parity": {
            "max_abs_gap_fp32_vs_fp64": float(abs(p32[:, :k] - p64[:, :k]).max()),
            "max_gap_in_se": float(abs(gap).max()),
            "max_abs_vol_bp_fp32_vs_fp64": strip_vol_bp(p32, p64),
            "forward_rel_err_fp32": float(max(abs(p32[j, k] / (100 * math.exp(0.03 * MATS[j])) - 1) for j in range(8))),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS,
            "traffic": STRIP_TRAFFIC_BYTES,
            "kernel_ms": kern32,
            "slots_per_path_step": SLOTS_PER_PATH_STEP,
            "roofline_path_steps_per_s": LANE_SLOTS_PER_S / SLOTS_PER_PATH_STEP,
        },
    }
    if not args.no_cpu:
        host = host_threads()
        threads = host["threads"]
        cpu_rate, cpu_min, cpu_max, n_cpu = _time_cpu(threads)
        line["cpu_baseline"] = {
            "value": cpu_rate,
            "unit": "path-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"the same QE-M strip in FP64 on {n_cpu} paths x {N_STEPS} steps, oracle/oracle.cpp on every "
                      f"host thread this process may use ({threads}); warm-up 2 s, median of 5 repetitions of at "
                      "least 300 ms",
            "min": cpu_min,
            "max": cpu_max,
            "nproc": host["nproc"],
            "cgroup_cpu_quota": host["cgroup_cpu_quota"],
            "cpu_model": host["cpu_model"],
        }
        line["gpu_over_cpu"] = v32 / cpu_rate
    if args.no_blocks:
        print(json.dumps(line), flush=True)
        return 0
    if args.adjoint_paths:
        line["adjoint"] = bench_adjoint(args.adjoint_paths)
        print("bench: adjoint block done", file=sys.stderr, flush=True)
    if args.book_scale:
        line["book"] = bench_book(args.book_scale)
        print("bench: book block done", file=sys.stderr, flush=True)
    if args.slv_particles:
        line["slv"] = bench_slv(args.slv_particles)
        print("bench: slv block done", file=sys.stderr, flush=True)
    line["fixings"] = bench_fixings()
    print("bench: fixings block done", file=sys.stderr, flush=True)
    line["lsm"] = bench_lsm(cpu=not args.no_cpu)
    print("bench: lsm block done", file=sys.stderr, flush=True)
    line["reference_workloads"] = bench_reference_workloads()
    print("bench: reference workloads done", file=sys.stderr, flush=True)
    line["strip_greeks"] = bench_strip_greeks()
    print("bench: strip sensitivities done", file=sys.stderr, flush=True)
    line["script"] = bench_script(cpu=not args.no_cpu)
    print("bench: script payoff done", file=sys.stderr, flush=True)
    if args.adjoint_paths:
        line["adjoint_models"] = bench_adjoint_models(args.adjoint_paths)
        print("bench: adjoint with models done", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
