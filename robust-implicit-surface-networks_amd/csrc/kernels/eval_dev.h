// FP64 implicit-function evaluation at grid vertices (the reference calls the
// external implicit_functions::load_functions on the host,
// app/implicit_arrangement.cpp:59).  Formulas: plane (x-p).n; sphere
// r^2-|x-c|^2 or r-|x-c| (positive inside); cylinder r - |(x-c) - ((x-c).a)a|;
// torus r - sqrt((|x_perp|-R)^2 + h^2); zero.  Axes arrive unit length.
// Evaluated left to right without contraction; square roots are correctly
// rounded so the device values equal the host's bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "exact_dev.h"

namespace rin {

struct FuncDesc {
    int32_t type;  // 0 plane, 1 sphere, 2 cylinder, 3 torus, 4 zero
    int32_t pad;
    double p[12];
};

namespace dev {

// sign of m - (y + h)^2 computed exactly (h a power of two), as an expansion
__device__ inline int mid_cmp(double m, double y, double h) {
    double e[6];
    int n = 1;
    e[0] = m;
    double p, q;
    two_prod(y, y, p, q);
    n = grow(e, n, -p, e);
    n = grow(e, n, -q, e);
    n = grow(e, n, -(2.0 * h) * y, e);
    n = grow(e, n, -(h * h), e);
    return exp_sign(e, n);
}

// correctly rounded square root: any close approximation, then exact
// midpoint tests move it to the nearest double
__device__ inline double sqrt_rn(double x) {
    if (!(x > 0.0) || x == __builtin_huge_val()) return (x == 0.0 || x == __builtin_huge_val()) ? x : __builtin_nan("");
    int e;
    double m = frexp(x, &e);  // [0.5,1)
    m *= 2.0;
    e -= 1;                   // [1,2)
    if (e & 1) {
        m *= 2.0;
        e -= 1;               // [1,4), even exponent
    }
    double y = __builtin_sqrt(m);
    // Fast acceptance.  For the correctly rounded root the residual m - y*y
    // is a double (Boldo and Daumas, "Representable correcting terms for
    // possibly underflowing floating point operations", 2003), so one fused
    // multiply-add gives it exactly; y in [1, 2] is the correctly rounded root
    // when the residual lies within 2 h y of zero (h = 2^-53, half an ulp), and
    // a root off by an ulp has a residual beyond (1 - 2^-52) of that band,
    // exact or rounded, so a test with a margin of 2^-40 accepts only the
    // correctly rounded root (and nearly every one: a lane that fails sends
    // its whole wavefront through the slow path).  Roots within 2^-40 of a
    // midpoint go on to the exact tests below.
    if (__builtin_fabs(__builtin_fma(-y, y, m)) <= (1.0 - 0x1p-40) * 0x1p-52 * y) return ldexp(y, e / 2);
    const double u = 0x1p-52;
    for (int it = 0; it < 3; ++it) {
        if (y < 2.0 && mid_cmp(m, y, 0.5 * u) > 0) {
            y += u;
            continue;
        }
        const double ud = (y > 1.0) ? u : 0.5 * u;
        if (y > 1.0 && mid_cmp(m, y, -0.5 * ud) < 0) {
            y -= ud;
            continue;
        }
        break;
    }
    return ldexp(y, e / 2);
}

__device__ inline double eval_func(const FuncDesc& f, double x0, double x1, double x2) {
    const double* p = f.p;
    switch (f.type) {
        case 0:
            return (x0 - p[0]) * p[3] + (x1 - p[1]) * p[4] + (x2 - p[2]) * p[5];
        case 1: {
            double d0 = x0 - p[0], d1 = x1 - p[1], d2 = x2 - p[2];
            double q = d0 * d0 + d1 * d1 + d2 * d2;
            return p[4] != 0.0 ? p[3] * p[3] - q : p[3] - sqrt_rn(q);
        }
        case 2: {
            double d0 = x0 - p[0], d1 = x1 - p[1], d2 = x2 - p[2];
            double h = d0 * p[3] + d1 * p[4] + d2 * p[5];
            double e0 = d0 - h * p[3], e1 = d1 - h * p[4], e2 = d2 - h * p[5];
            double v = p[6] - sqrt_rn(e0 * e0 + e1 * e1 + e2 * e2);
            return p[7] != 0.0 ? -v : v;
        }
        case 3: {
            double d0 = x0 - p[0], d1 = x1 - p[1], d2 = x2 - p[2];
            double h = d0 * p[3] + d1 * p[4] + d2 * p[5];
            double e0 = d0 - h * p[3], e1 = d1 - h * p[4], e2 = d2 - h * p[5];
            double u = sqrt_rn(e0 * e0 + e1 * e1 + e2 * e2) - p[6];
            double v = p[7] - sqrt_rn(u * u + h * h);
            return p[8] != 0.0 ? -v : v;
        }
        default:
            return 0.0;
    }
}

}  // namespace dev
}  // namespace rin
