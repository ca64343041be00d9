// The reference's input/output routines (src/io.h:12-98), with the same
// names, signatures and file formats, so its callers (app/*.cpp,
// tests/test_implicit_networks.cpp) compile unchanged against this header and
// librin_io.so in place of src/io.h / src/io.cpp:
//   parse_config_file   src/io.cpp:19-63   CLI configuration JSON
//   load_tet_mesh       src/io.cpp:65-93   [points, tets] JSON
//   generate_tet_mesh   src/io.cpp:95-152  5-tet regular grid
//   save_result         src/io.cpp:154-233 mesh.json of an arrangement
//   save_result_MI      src/io.cpp:235-312 mesh.json of a material interface
//   save_timings        src/io.cpp:556-572 timings.json
//   save_statistics     src/io.cpp:574-590 stats.json
// The JSON files are written as the reference writes them: one line, keys in
// ascending order, numbers in shortest round-trip form, empty lists as null.
#pragma once
#include <array>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "rin/mesh.h"

struct Config {
    // Input
    std::string tet_mesh_file;
    std::string func_file;
    std::string output_dir;
    // Lookup setting.
    bool use_lookup = true;
    bool use_secondary_lookup = true;
    bool use_topo_ray_shooting = true;
    // Parameters for tet grid generation (only used if tet_mesh_file is empty).
    size_t tet_mesh_resolution = 0;
    std::array<double, 3> tet_mesh_bbox_min{};
    std::array<double, 3> tet_mesh_bbox_max{};
};

// Throws std::runtime_error when the file does not exist or lacks a key.
// Relative paths are taken relative to the configuration file's folder.
Config parse_config_file(const std::string& filename);

bool load_tet_mesh(const std::string& filename, std::vector<std::array<double, 3>>& pts,
                   std::vector<std::array<size_t, 4>>& tets);

bool generate_tet_mesh(size_t resolution, const std::array<double, 3>& bbox_min,
                       const std::array<double, 3>& bbox_max, std::vector<std::array<double, 3>>& pts,
                       std::vector<std::array<size_t, 4>>& tets);

bool save_result(const std::string& filename, const std::vector<std::array<double, 3>>& mesh_pts,
                 const std::vector<PolygonFace>& mesh_faces, const std::vector<std::vector<size_t>>& patches,
                 const std::vector<size_t>& patch_function_label, const std::vector<Edge>& edges,
                 const std::vector<std::vector<size_t>>& chains,
                 const std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                 const std::vector<std::vector<size_t>>& shells, const std::vector<std::vector<size_t>>& cells,
                 const std::vector<std::vector<bool>>& cell_function_label);

bool save_result_MI(const std::string& filename, const std::vector<std::array<double, 3>>& mesh_pts,
                    const std::vector<PolygonFace>& mesh_faces, const std::vector<std::vector<size_t>>& patches,
                    const std::vector<std::pair<size_t, size_t>>& patch_function_label,
                    const std::vector<Edge>& edges, const std::vector<std::vector<size_t>>& chains,
                    const std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                    const std::vector<std::vector<size_t>>& shells, const std::vector<std::vector<size_t>>& cells,
                    const std::vector<size_t>& cell_function_label);

bool save_timings(const std::string& filename, const std::vector<std::string>& timing_labels,
                  const std::vector<double>& timings);

bool save_statistics(const std::string& filename, const std::vector<std::string>& stats_labels,
                     const std::vector<size_t>& stats);
