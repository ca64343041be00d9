// Test-only: a caller shaped like the reference's command-line drivers
// (app/implicit_arrangement.cpp:11-127, app/material_interface.cpp:15-131),
// compiled unchanged in structure against this build's headers only:
// configuration, lookup-table switch, mesh loading or generation,
// load_functions, the drop-in entry point, and the result / timing / stats
// files.  Usage: app_caller ia|mi <config.json> [-R].  The last stdout line
// is a JSON summary for the test.
#include <simplicial_arrangement/lookup_table.h>

#include <Eigen/Core>
#include <cstring>
#include <iostream>

#include "implicit_functions.h"
#include "rin/implicit_arrangement.h"
#include "rin/io.h"

using namespace simplicial_arrangement;

int main(int argc, const char* argv[]) {
    if (argc < 3) return 3;
    const bool mi = std::strcmp(argv[1], "mi") == 0;
    const bool robust_test = argc > 3 && std::strcmp(argv[3], "-R") == 0;

    // parse configure file
    Config config = parse_config_file(argv[2]);

    if (config.use_lookup) {
        std::cout << "load table ..." << std::endl;
        bool loaded = mi ? load_lookup_table(simplicial_arrangement::MATERIAL_INTERFACE) : load_lookup_table();
        if (loaded) {
            std::cout << "loading finished." << std::endl;
        } else {
            std::cout << "loading failed." << std::endl;
            return -1;
        }
    } else {
        disable_lookup_table();
        config.use_secondary_lookup = false;
    }

    // load tet mesh
    std::vector<std::array<double, 3>> pts;
    std::vector<std::array<size_t, 4>> tets;
    if (config.tet_mesh_file != "") {
        std::cout << "load mesh file " << config.tet_mesh_file << std::endl;
        load_tet_mesh(config.tet_mesh_file, pts, tets);
    } else {
        std::cout << "generating mesh with resolution " << config.tet_mesh_resolution << std::endl;
        generate_tet_mesh(config.tet_mesh_resolution, config.tet_mesh_bbox_min, config.tet_mesh_bbox_max, pts, tets);
    }

    // load implicit functions and compute function values at vertices
    Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> funcVals;
    if (load_functions(config.func_file, pts, funcVals)) {
        std::cout << "function loading finished." << std::endl;
    } else {
        std::cout << "function loading failed." << std::endl;
        return -2;
    }

    std::vector<std::array<double, 3>> out_pts;
    std::vector<PolygonFace> out_faces;
    std::vector<std::vector<size_t>> patches;
    std::vector<size_t> patch_function_label;
    std::vector<std::pair<size_t, size_t>> patch_material_label;
    std::vector<Edge> out_edges;
    std::vector<std::vector<size_t>> chains;
    std::vector<std::vector<size_t>> non_manifold_edges_of_vert;
    std::vector<std::vector<size_t>> shells;
    std::vector<std::vector<size_t>> cells;
    std::vector<std::vector<bool>> cell_function_label;
    std::vector<size_t> cell_material_label;
    std::vector<std::string> timing_labels;
    std::vector<double> timings;
    std::vector<std::string> stats_labels;
    std::vector<size_t> stats;

    bool ok;
    if (mi)
        ok = material_interface(robust_test, config.use_lookup, config.use_secondary_lookup,
                                config.use_topo_ray_shooting, pts, tets, funcVals, out_pts, out_faces, patches,
                                patch_material_label, out_edges, chains, non_manifold_edges_of_vert, shells, cells,
                                cell_material_label, timing_labels, timings, stats_labels, stats);
    else
        ok = implicit_arrangement(robust_test, config.use_lookup, config.use_secondary_lookup,
                                  config.use_topo_ray_shooting, pts, tets, funcVals, out_pts, out_faces, patches,
                                  patch_function_label, out_edges, chains, non_manifold_edges_of_vert, shells, cells,
                                  cell_function_label, timing_labels, timings, stats_labels, stats);
    if (!ok) {
        std::cout << "{\"ok\":false}" << std::endl;
        return -1;
    }
    if (!robust_test) {
        if (mi)
            save_result_MI(config.output_dir + "/mesh.json", out_pts, out_faces, patches, patch_material_label,
                           out_edges, chains, non_manifold_edges_of_vert, shells, cells, cell_material_label);
        else
            save_result(config.output_dir + "/mesh.json", out_pts, out_faces, patches, patch_function_label,
                        out_edges, chains, non_manifold_edges_of_vert, shells, cells, cell_function_label);
    }
    save_timings(config.output_dir + "/timings.json", timing_labels, timings);
    save_statistics(config.output_dir + "/stats.json", stats_labels, stats);
    std::cout << "{\"ok\":true,\"num_pts\":" << pts.size() << ",\"num_tets\":" << tets.size() << "}" << std::endl;
    return 0;
}
