// Host graph stages that follow the per-tet hot path: iso-edges, patches,
// chains, face order around chains, shells/components, topological ray
// shooting, arrangement cells and sign propagation.  They stay on the CPU and
// read the device-produced arrays plus the per-tet complexes of the few tets
// they need (through CutSource).
#pragma once
#include <array>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "cut_view.h"
#include "rin/mesh.h"

namespace rin {

using HalfPair = std::pair<std::pair<size_t, int>, std::pair<size_t, int>>;

// src/mesh_connectivity.cpp:10-56
void compute_mesh_edges(const std::vector<PolygonFace>& faces, std::vector<std::vector<size_t>>& edges_of_face,
                        std::vector<Edge>& edges);
// src/mesh_connectivity.cpp:58-94 (IA) and :156-192 (MI)
void compute_patches(const std::vector<std::vector<size_t>>& edges_of_face, const std::vector<Edge>& edges,
                     const std::vector<PolygonFace>& faces, std::vector<std::vector<size_t>>& patches,
                     std::vector<std::pair<size_t, size_t>>& labels);
// src/mesh_connectivity.cpp:196-241
void compute_chains(const std::vector<Edge>& edges, const std::vector<std::vector<size_t>>& nm_edges_of_vert,
                    std::vector<std::vector<size_t>>& chains);
// src/mesh_connectivity.cpp:243-344
void compute_shells_and_components(size_t num_patch, const std::vector<std::vector<HalfPair>>& half_patch_pairs,
                                   std::vector<std::vector<size_t>>& shells, std::vector<size_t>& shell_of_half_patch,
                                   std::vector<std::vector<size_t>>& components,
                                   std::vector<size_t>& component_of_patch);
// src/mesh_connectivity.cpp:347-405
void compute_arrangement_cells(size_t num_shell, const std::vector<std::pair<size_t, size_t>>& links,
                               std::vector<std::vector<size_t>>& cells);
// src/mesh_connectivity.cpp:407-481
std::vector<std::vector<bool>> sign_propagation(const std::vector<std::vector<size_t>>& cells,
                                                const std::vector<size_t>& shell_of_half_patch,
                                                const std::vector<std::vector<size_t>>& shells,
                                                const std::vector<size_t>& patch_label, size_t n_func,
                                                const std::vector<bool>& sample_label);
// src/mesh_connectivity.cpp:483-518
std::vector<size_t> sign_propagation_MI(const std::vector<std::vector<size_t>>& cells,
                                        const std::vector<size_t>& shell_of_half_patch,
                                        const std::vector<std::vector<size_t>>& shells,
                                        const std::vector<std::pair<size_t, size_t>>& patch_label, size_t n_func,
                                        const std::vector<double>& sample_values);

// Everything the graph stages read.
struct GraphIn {
    bool mi = false;
    const std::vector<std::array<double, 3>>* pts = nullptr;
    const std::vector<std::array<size_t, 4>>* tets = nullptr;
    // iso/MI vertices in a common form (IsoVert fields; MI uses all four
    // function slots through `fn4`)
    std::vector<IsoVert>* iso_verts = nullptr;
    std::vector<MI_Vert>* mi_verts = nullptr;
    const std::vector<PolygonFace>* faces = nullptr;
    const std::vector<size_t>* func_in_tet = nullptr;
    const std::vector<size_t>* start_index = nullptr;
    const std::vector<uint8_t>* degenerate_vertex = nullptr;  // per grid vertex
    bool found_degenerate = false;
    size_t n_func = 0;
    std::vector<bool> sample_sign;       // IA: sign of each function at vertex 0
    std::vector<double> sample_value;    // MI: value of each function at vertex 0
    CutSource* cuts = nullptr;
    // when set, the iso-edges come from here (e.g. computed on the device)
    // instead of compute_mesh_edges over the faces
    std::function<void(std::vector<std::vector<size_t>>&, std::vector<Edge>&)> edges_from;
    // when set, the patches come from here instead of compute_patches
    std::function<void(std::vector<std::vector<size_t>>&)> patches_from;
    // when set, the chains come from here instead of compute_chains
    std::function<void(std::vector<std::vector<size_t>>&)> chains_from;
};

struct GraphOut {
    std::vector<Edge> edges;
    std::vector<std::vector<size_t>> patches;
    std::vector<size_t> patch_label;                       // IA
    std::vector<std::pair<size_t, size_t>> patch_label_mi; // MI
    std::vector<std::vector<size_t>> chains;
    std::vector<std::vector<size_t>> nm_edges_of_vert;
    std::vector<std::vector<size_t>> shells;
    std::vector<std::vector<size_t>> cells;
    std::vector<std::vector<bool>> cell_label;             // IA
    std::vector<size_t> cell_label_mi;                     // MI
    size_t num_components = 0;
    std::vector<std::pair<std::string, double>> timings;
    std::vector<std::pair<std::string, size_t>> stats;
    std::string error;
};

// Runs the stages in the order of src/implicit_arrangement.cpp:404-645 /
// src/material_interface.cpp:449-694.  Returns false (with out.error) where
// the reference returns false.
bool run_graph(const GraphIn& in, GraphOut& out);

}  // namespace rin
