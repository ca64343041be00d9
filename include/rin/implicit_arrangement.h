// Drop-in C++ entry points with the reference's signatures:
//   implicit_arrangement(...)  src/implicit_arrangement.h:39-62
//   material_interface(...)    src/material_interface.h:38-61
// Callers (app/implicit_arrangement.cpp:84, app/material_interface.cpp:87,
// src/csg.cpp:38, tests/test_implicit_networks.cpp) compile unchanged: the
// function-value matrix parameter accepts the reference's
// Eigen::Matrix<double, Dynamic, Dynamic, RowMajor> (or any row-major type
// with data(), rows() and cols()).
//
// The per-tet path (signs, filter, per-tet cut, iso-mesh extraction, vertex
// coordinates) runs on the GPU through the C ABI of include/rin.h; the graph
// stages (edges, patches, chains, shells, arrangement cells) run on the host.
// Outputs are appended, as in the reference.  Both functions return false and
// print a message where the reference does, and also when the device fails.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "rin/mesh.h"

namespace rin {

template <class M, class = void>
struct storage_is_row_major : std::true_type {};
template <class M>
struct storage_is_row_major<M, std::void_t<decltype(M::IsRowMajor)>> : std::bool_constant<bool(M::IsRowMajor)> {};

// Read-only view of a row-major (vertices x functions) matrix of doubles.
struct RowMajorValues {
    const double* data = nullptr;
    size_t rows = 0, cols = 0;
    RowMajorValues(const double* d, size_t r, size_t c) : data(d), rows(r), cols(c) {}
    template <class M>
    RowMajorValues(const M& m) : data(m.data()), rows((size_t)m.rows()), cols((size_t)m.cols()) {
        static_assert(storage_is_row_major<M>::value, "funcVals must be stored row-major (vertices x functions)");
    }
};

// Device used by the drop-in entry points on the calling thread (default 0).
void set_device(int device);

// The -R verdict of src/implicit_arrangement.cpp:329-343 (and
// src/material_interface.cpp:374-386), which the reference computes inline:
// type 2 when the cut failed in the normal order, type 3 when it failed in the
// reverse order, type 1 when a tet's (vertex, face, cell) counts differ
// between the orders, success otherwise.  Prints the reference's message and
// returns whether the check passed.
bool robust_verdict(bool forward_ok, bool reverse_ok, const std::vector<uint64_t>& forward_counts,
                    const std::vector<uint64_t>& reverse_counts);

}  // namespace rin

bool implicit_arrangement(bool robust_test, bool use_lookup, bool use_secondary_lookup, bool use_topo_ray_shooting,
                          //
                          const std::vector<std::array<double, 3>>& pts,
                          const std::vector<std::array<size_t, 4>>& tets, const rin::RowMajorValues& funcVals,
                          //
                          std::vector<std::array<double, 3>>& iso_pts, std::vector<PolygonFace>& iso_faces,
                          std::vector<std::vector<size_t>>& patches, std::vector<size_t>& patch_function_label,
                          std::vector<Edge>& iso_edges, std::vector<std::vector<size_t>>& chains,
                          std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                          std::vector<std::vector<size_t>>& shells,
                          std::vector<std::vector<size_t>>& arrangement_cells,
                          std::vector<std::vector<bool>>& cell_function_label, std::vector<std::string>& timing_labels,
                          std::vector<double>& timings, std::vector<std::string>& stats_labels,
                          std::vector<size_t>& stats);

bool material_interface(bool robust_test, bool use_lookup, bool use_secondary_lookup, bool use_topo_ray_shooting,
                        //
                        const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<size_t, 4>>& tets,
                        const rin::RowMajorValues& funcVals,
                        //
                        std::vector<std::array<double, 3>>& MI_pts, std::vector<PolygonFace>& MI_faces,
                        std::vector<std::vector<size_t>>& patches,
                        std::vector<std::pair<size_t, size_t>>& patch_function_label, std::vector<Edge>& MI_edges,
                        std::vector<std::vector<size_t>>& chains,
                        std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                        std::vector<std::vector<size_t>>& shells, std::vector<std::vector<size_t>>& material_cells,
                        std::vector<size_t>& cell_function_label, std::vector<std::string>& timing_labels,
                        std::vector<double>& timings, std::vector<std::string>& stats_labels,
                        std::vector<size_t>& stats);
