// Decoded per-tet cell complex: the fields of simplicial_arrangement's
// Arrangement<3> / MaterialInterface<3> that the reference's host graph code
// reads (src/extract_mesh.cpp:54-260, src/pair_faces.cpp:138-718,
// src/topo_ray_shooting.cpp:529-1129).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "rin/cut_record.h"
#include "rin/mesh.h"

namespace rin {

struct CutFace {
    std::vector<size_t> vertices;
    // IA
    size_t supporting_plane = Mesh_None;
    size_t positive_cell = Mesh_None;
    size_t negative_cell = Mesh_None;
    // MI
    size_t positive_material_label = Mesh_None;
    size_t negative_material_label = Mesh_None;
};
struct CutCell {
    std::vector<size_t> faces;
    size_t material_label = Mesh_None;
};
struct CutView {
    bool mi = false;
    size_t nlabels = 0;
    std::vector<std::array<size_t, 4>> vertices;  // IA uses 3 entries
    std::vector<CutFace> faces;
    std::vector<CutCell> cells;
    // coincidence groups (empty when every plane/material is unique)
    std::vector<std::vector<size_t>> unique_planes;
    std::vector<size_t> unique_plane_indices;
    std::vector<bool> unique_plane_orientations;
};

// Decode one record; throws std::runtime_error on a malformed record.
CutView decode_cut(const uint16_t* rec);

// Source of per-tet complexes for the host graph stages.
struct CutSource {
    virtual ~CutSource() = default;
    virtual const CutView& get(size_t tet) = 0;
    // Hint: these tets are about to be read (lets a device-backed source
    // materialise them in one batch).
    virtual void prefetch(const std::vector<size_t>& tets) { (void)tets; }
};

}  // namespace rin
