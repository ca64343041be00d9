// Output record types of the drop-in C++ entry points.
//
// Field names, meanings and element types follow the reference's src/mesh.h:11-71
// (Mesh_None, PolygonFace, IsoVert, MI_Vert, Edge) so that code written against
// the reference compiles unchanged against this build.
#pragma once
#include <array>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

static constexpr size_t Mesh_None = std::numeric_limits<size_t>::max();

// A polygon of the output network (src/mesh.h:15-25).
struct PolygonFace {
    std::vector<size_t> vert_indices;                        // cycle of global vertex ids
    std::vector<std::pair<size_t, size_t>> tet_face_indices; // (tet, local face) occurrences
    std::pair<size_t, size_t> func_index;                    // IA: .first; MI: (pos, neg) material
};

// A vertex of the implicit arrangement (src/mesh.h:28-41).
struct IsoVert {
    size_t tet_index;                  // tet of the first occurrence
    size_t tet_vert_index;             // local vertex index there
    size_t simplex_size;               // 1 point, 2 edge, 3 triangle, 4 tet
    std::array<size_t, 4> simplex_vert_indices;
    std::array<size_t, 3> func_indices = {Mesh_None, Mesh_None, Mesh_None};
};

// A vertex of the material interface (src/mesh.h:45-57).
struct MI_Vert {
    size_t tet_index;
    size_t tet_vert_index;
    size_t simplex_size;
    std::array<size_t, 4> simplex_vert_indices;
    std::array<size_t, 4> material_indices;
};

// An edge of the output network (src/mesh.h:62-71).
struct Edge {
    size_t v1;
    size_t v2;
    std::vector<std::pair<size_t, size_t>> face_edge_indices;  // (face, edge-in-face)
};
