// ORACLE — test infrastructure only. Nothing in the product links this code.
//
// Exact sign of small determinants whose entries are IEEE doubles.
//
// The reference's per-tet cut library (simplicial_arrangement, an external
// dependency fetched at configure time, absent here) decides every vertex /
// plane classification with exact predicates.  Its call sites are
// src/implicit_arrangement.cpp:265-283 and src/material_interface.cpp:309-327.
// The oracle restates those predicates as signs of determinants and evaluates
// them exactly with an arbitrary-precision binary float (BigF): every double is
// an integer times a power of two, so products and sums of doubles are exact
// in BigF.  A plain double evaluation with a permanent-based error bound is
// tried first; BigF decides whenever that bound does not separate the value
// from zero.
#pragma once
#include <cstdint>
#include <vector>

namespace oracle {

// value = sign * mag * 2^exp, mag little-endian base 2^32, no leading zero limbs.
struct BigF {
    int sign = 0;       // -1, 0, +1
    long exp = 0;
    std::vector<uint32_t> mag;
    static BigF from_double(double d);
};

BigF big_mul(const BigF& a, const BigF& b);
BigF big_add(const BigF& a, const BigF& b);
BigF big_neg(const BigF& a);

// Exact sign of det(M), M is n x n row-major, n <= 6.
int det_sign_exact(int n, const double* M);
// Same, trying a floating-point filter first.
int det_sign(int n, const double* M);

// Counter of how many determinant signs needed the exact path (diagnostics).
extern uint64_t g_exact_fallbacks;

}  // namespace oracle
