// Device pieces shared by the kernels that step local-vol paths and evaluate the structured payoffs on them (the
// worst-of autocall with its adjoint, the batched book): the table lerp of
// packages/native/cuda/include/optlab_cuda/paths.hpp:196-205 with its slope, the correlation of the names'
// normals, the smoothed comparisons of packages/structured/src/optlab/structured/script/evaluator.py:93-108 and one
// observation date of the phoenix script (packages/structured/src/optlab/structured/compiler.py:114-161, 270-295).
#pragma once

#include "device.hpp"

namespace optlab_mi {

template <typename R>
struct Terms {  // compiler.py:75-111 parameters, in the path precision
    R autocall, coupon_barrier, coupon, ki_barrier, put_strike, smoothing;
    int memory;
};

// Smoothed comparisons (evaluator.py:93-108): the value and its derivative in the left operand.
template <typename R>
struct Ramp {
    R v, d;
};

template <typename R>
__device__ __forceinline__ Ramp<R> ramp_ge(R a, R b, R eps) {  // a >= b
    const R scale = eps * (b < R(0) ? -b : b);
    if (!(scale > R(0))) return Ramp<R>{a >= b ? R(1) : R(0), R(0)};
    const R r = (a - b) / scale + R(0.5);
    return Ramp<R>{r_min(r_max(r, R(0)), R(1)), (r > R(0) && r < R(1)) ? R(1) / scale : R(0)};
}

template <typename R>
__device__ __forceinline__ Ramp<R> ramp_lt(R a, R b, R eps) {  // a < b
    const R scale = eps * (b < R(0) ? -b : b);
    if (!(scale > R(0))) return Ramp<R>{a < b ? R(1) : R(0), R(0)};
    const R r = (b - a) / scale + R(0.5);
    return Ramp<R>{r_min(r_max(r, R(0)), R(1)), (r > R(0) && r < R(1)) ? R(-1) / scale : R(0)};
}

template <typename R>
struct Note {  // the script's state: alive (1 until called), missed coupons, value paid so far
    R alive, missed, value;
};

// One observation date of the script (compiler.py:114-127 coupon, 270-281 call, 141-161 redemption).
template <typename R>
__device__ __forceinline__ void observe(const Terms<R>& t, bool last, R p, R low, R D, Note<R>& s) {
    const R c = ramp_ge(p, t.coupon_barrier, t.smoothing).v;
    const R f = t.memory ? R(1) + s.missed : R(1);
    s.value += s.alive * c * t.coupon * f * D;
    if (t.memory) s.missed = (R(1) - c) * (s.missed + R(1));
    if (!last) {
        const R u = ramp_ge(p, t.autocall, t.smoothing).v;
        s.value += s.alive * u * D;
        s.alive *= R(1) - u;
    } else {
        const R ki = t.ki_barrier > R(0) ? ramp_lt(low, t.ki_barrier, t.smoothing).v : R(0);
        s.value += s.alive * D * (R(1) - ki * r_max(t.put_strike - p, R(0)));
    }
}

// The lerp of paths.hpp:196-205 with its slope in x (zero where u is clamped) and the weight of entry i0 + 1; u is
// the clamped grid coordinate (i0 = int(u), f = u - i0).
template <typename R>
struct Lerp {
    int i0;
    R f, var, slope, u;
};

template <typename R>
__device__ __forceinline__ Lerp<R> lerp_at(const R* y, R lo, R inv, int n_x, R x) {
    const R u0 = (x - lo) * inv;
    const R top = static_cast<R>(n_x - 1) - static_cast<R>(1e-4);
    const R u = r_min(r_max(u0, R(0)), top);
    const int i = static_cast<int>(u);
    const R f = u - static_cast<R>(i);
    const R y0 = y[i], y1 = y[i + 1];
    const bool inside = u0 > R(0) && u0 < top;
    return Lerp<R>{i, f, y0 * (R(1) - f) + y1 * f, inside ? (y1 - y0) * inv : R(0), u};
}

template <typename R, int N>
__device__ __forceinline__ void correlate(const R* root, R sign, const R (&zs)[N], R (&z)[N]) {
#pragma unroll
    for (int a = 0; a < N; ++a) {
        R acc = R(0);
#pragma unroll
        for (int b = 0; b < N; ++b) acc += ldc(root, a * N + b) * zs[b];
        z[a] = sign * acc;
    }
}

}  // namespace optlab_mi
