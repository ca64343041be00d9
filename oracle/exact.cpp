// ORACLE — test infrastructure only. Nothing in the product links this code.
// Exact determinant signs via an arbitrary-precision binary float; see exact.h.
#include "exact.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace oracle {

uint64_t g_exact_fallbacks = 0;

static void trim(std::vector<uint32_t>& m) {
    while (!m.empty() && m.back() == 0) m.pop_back();
}

BigF BigF::from_double(double d) {
    BigF r;
    if (d == 0.0) return r;
    if (!std::isfinite(d)) throw std::runtime_error("oracle: non-finite value in predicate");
    int e;
    double m = std::frexp(std::fabs(d), &e);     // d = m * 2^e, m in [0.5,1)
    uint64_t mi = (uint64_t)std::ldexp(m, 64 - 11);  // 53 significant bits
    r.sign = d < 0 ? -1 : 1;
    r.exp = (long)e - 53;
    r.mag.push_back((uint32_t)(mi & 0xffffffffu));
    r.mag.push_back((uint32_t)(mi >> 32));
    trim(r.mag);
    return r;
}

BigF big_neg(const BigF& a) {
    BigF r = a;
    r.sign = -r.sign;
    return r;
}

BigF big_mul(const BigF& a, const BigF& b) {
    BigF r;
    if (a.sign == 0 || b.sign == 0) return r;
    r.sign = a.sign * b.sign;
    r.exp = a.exp + b.exp;
    r.mag.assign(a.mag.size() + b.mag.size(), 0);
    for (size_t i = 0; i < a.mag.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.mag.size(); ++j) {
            uint64_t cur = (uint64_t)a.mag[i] * b.mag[j] + r.mag[i + j] + carry;
            r.mag[i + j] = (uint32_t)cur;
            carry = cur >> 32;
        }
        size_t k = i + b.mag.size();
        while (carry) {
            uint64_t cur = (uint64_t)r.mag[k] + carry;
            r.mag[k] = (uint32_t)cur;
            carry = cur >> 32;
            ++k;
        }
    }
    trim(r.mag);
    return r;
}

// magnitude shifted left by s bits
static std::vector<uint32_t> shl(const std::vector<uint32_t>& m, long s) {
    long limbs = s / 32, bits = s % 32;
    std::vector<uint32_t> r(m.size() + limbs + 1, 0);
    for (size_t i = 0; i < m.size(); ++i) {
        uint64_t v = (uint64_t)m[i] << bits;
        r[i + limbs] |= (uint32_t)v;
        r[i + limbs + 1] |= (uint32_t)(v >> 32);
    }
    trim(r);
    return r;
}

static int cmp_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;) {
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}

static std::vector<uint32_t> add_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        uint64_t cur = carry;
        if (i < a.size()) cur += a[i];
        if (i < b.size()) cur += b[i];
        r[i] = (uint32_t)cur;
        carry = cur >> 32;
    }
    trim(r);
    return r;
}

// a - b with |a| >= |b|
static std::vector<uint32_t> sub_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    std::vector<uint32_t> r(a.size(), 0);
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        int64_t cur = (int64_t)a[i] - borrow - (i < b.size() ? (int64_t)b[i] : 0);
        if (cur < 0) {
            cur += ((int64_t)1 << 32);
            borrow = 1;
        } else {
            borrow = 0;
        }
        r[i] = (uint32_t)cur;
    }
    trim(r);
    return r;
}

BigF big_add(const BigF& a, const BigF& b) {
    if (a.sign == 0) return b;
    if (b.sign == 0) return a;
    long e = std::min(a.exp, b.exp);
    std::vector<uint32_t> am = a.exp > e ? shl(a.mag, a.exp - e) : a.mag;
    std::vector<uint32_t> bm = b.exp > e ? shl(b.mag, b.exp - e) : b.mag;
    BigF r;
    r.exp = e;
    if (a.sign == b.sign) {
        r.sign = a.sign;
        r.mag = add_mag(am, bm);
    } else {
        int c = cmp_mag(am, bm);
        if (c == 0) return BigF();
        if (c > 0) {
            r.sign = a.sign;
            r.mag = sub_mag(am, bm);
        } else {
            r.sign = b.sign;
            r.mag = sub_mag(bm, am);
        }
    }
    // drop whole trailing zero limbs to keep sizes small
    size_t z = 0;
    while (z < r.mag.size() && r.mag[z] == 0) ++z;
    if (z) {
        r.mag.erase(r.mag.begin(), r.mag.begin() + z);
        r.exp += 32 * (long)z;
    }
    return r;
}

// Laplace expansion along the first row over the column subset `cols` (bitmask),
// rows r..n-1.
static BigF det_big_rec(int n, const BigF* M, int r, unsigned cols) {
    if (r == n - 1) {
        for (int c = 0; c < n; ++c)
            if (cols & (1u << c)) return M[r * n + c];
        return BigF();
    }
    BigF acc;
    int k = 0;
    for (int c = 0; c < n; ++c) {
        if (!(cols & (1u << c))) continue;
        const BigF& a = M[r * n + c];
        if (a.sign != 0) {
            BigF minor = det_big_rec(n, M, r + 1, cols & ~(1u << c));
            BigF t = big_mul(a, minor);
            if (k & 1) t = big_neg(t);
            acc = big_add(acc, t);
        }
        ++k;
    }
    return acc;
}

int det_sign_exact(int n, const double* M) {
    std::vector<BigF> B(n * n);
    for (int i = 0; i < n * n; ++i) B[i] = BigF::from_double(M[i]);
    BigF d = det_big_rec(n, B.data(), 0, (1u << n) - 1);
    return d.sign;
}

// double evaluation of the same expansion plus the permanent of |M|
static void det_fp_rec(int n, const double* M, int r, unsigned cols, double& det, double& perm) {
    if (r == n - 1) {
        for (int c = 0; c < n; ++c)
            if (cols & (1u << c)) {
                det = M[r * n + c];
                perm = std::fabs(det);
                return;
            }
    }
    double acc = 0, pacc = 0;
    int k = 0;
    for (int c = 0; c < n; ++c) {
        if (!(cols & (1u << c))) continue;
        double a = M[r * n + c];
        if (a != 0) {
            double d, p;
            det_fp_rec(n, M, r + 1, cols & ~(1u << c), d, p);
            double t = a * d;
            acc += (k & 1) ? -t : t;
            pacc += std::fabs(a) * p;
        }
        ++k;
    }
    det = acc;
    perm = pacc;
}

int det_sign(int n, const double* M) {
    double d, p;
    det_fp_rec(n, M, 0, (1u << n) - 1, d, p);
    // Each of the n! products carries at most n-1 roundings and each partial
    // sum at most n more per level; 4*n*n*eps*perm bounds the total error with a
    // wide margin.  Tiny permanents are sent to the exact path because the
    // bound ignores underflow.
    const double eps = std::ldexp(1.0, -53);
    if (p > 1e-250 && p < 1e250) {
        double bound = 4.0 * n * n * eps * p;
        if (d > bound) return 1;
        if (d < -bound) return -1;
    }
    ++g_exact_fallbacks;
    return det_sign_exact(n, M);
}

}  // namespace oracle
