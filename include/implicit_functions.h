// load_functions of the external implicit_functions library, with the
// signature its callers use: app/implicit_arrangement.cpp:59,
// app/material_interface.cpp:62, app/csg.cpp:184 and
// tests/test_implicit_networks.cpp:62.
//
// The function file is the reference's JSON list of typed functions: plane
// (point, normal), sphere (center, radius, optional squared), cylinder
// (axis_point, axis_vector, radius, optional is_flipped), torus (center,
// axis_vector, major_radius, minor_radius, optional is_flipped) and zero.
// Every function is evaluated at every point on the GPU (rin_eval_points of
// include/rin.h).  funcVals is resized to (points x functions) and must be
// stored row-major, like the reference's
// Eigen::Matrix<double, Dynamic, Dynamic, RowMajor>.  Returns false with a
// message on stdout when the file cannot be read or names another type.
#pragma once
#include <algorithm>
#include <array>
#include <cstddef>
#include <string>
#include <type_traits>
#include <vector>

namespace rin {

template <class M, class = void>
struct matrix_is_row_major : std::true_type {};
template <class M>
struct matrix_is_row_major<M, std::void_t<decltype(M::IsRowMajor)>> : std::bool_constant<bool(M::IsRowMajor)> {};

// reads the file and evaluates on the GPU: values row-major (points x functions)
bool load_function_values(const std::string& filename, const std::vector<std::array<double, 3>>& pts,
                          std::vector<double>& values, size_t& n_funcs);

}  // namespace rin

template <class Matrix>
bool load_functions(const std::string& filename, const std::vector<std::array<double, 3>>& pts, Matrix& funcVals) {
    static_assert(rin::matrix_is_row_major<Matrix>::value, "funcVals must be stored row-major (points x functions)");
    std::vector<double> v;
    size_t F = 0;
    if (!rin::load_function_values(filename, pts, v, F)) return false;
    funcVals.resize(pts.size(), F);
    std::copy(v.begin(), v.end(), funcVals.data());
    return true;
}
