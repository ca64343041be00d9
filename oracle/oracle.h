// ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product never links it.
//
// A CPU restatement, written independently of the product's HIP path, of the
// reference's per-tetrahedron hot path:
//   generate_tet_mesh            src/io.cpp:95-152
//   function evaluation          app/implicit_arrangement.cpp:59 (external
//                                implicit_functions, formulas fixed by this
//                                project, sign conventions pinned by
//                                tests/test_implicit_networks.cpp known answers)
//   IA signs + filter            src/implicit_arrangement.cpp:53-116
//   MI highest material + filter src/material_interface.cpp:53-152
//   per-tet cut                  src/implicit_arrangement.cpp:244-306 and
//                                src/material_interface.cpp:288-351 (the
//                                external simplicial_arrangement library;
//                                restated as incremental insertion with exact
//                                predicates)
//   extract_iso_mesh             src/extract_mesh.cpp:10-265
//   extract_MI_mesh              src/extract_mesh.cpp:569-986
//   compute_iso_vert_xyz         src/extract_mesh.cpp:1446-1538
//   compute_MI_vert_xyz          src/extract_mesh.cpp:1541-1637
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace oracle {

constexpr int NONE = -1;

// ---------------------------------------------------------------- per-tet IA
struct IAFace {
    int plane;                 // supporting plane (0-3 tet boundary, 4+ implicit)
    int pos, neg;              // incident cells; neg == NONE on the tet boundary
    std::vector<int> vs;       // vertex cycle, normal points to the pos side
    std::vector<int> es;       // es[i]: the other plane through edge vs[i]->vs[i+1]
};
struct IACut {
    int nplanes = 4;
    std::vector<std::array<int, 3>> verts;   // sorted plane triples
    std::vector<IAFace> faces;
    int ncells = 1;
    bool dup = false;                        // some plane coincides with another
    std::vector<int> grp;                    // representative plane per plane
    std::vector<int> ori;                    // 1: same positive side as representative
};
// planes: k rows of 4 corner values, in insertion order
IACut ia_cut(const std::vector<std::array<double, 4>>& planes);

// ---------------------------------------------------------------- per-tet MI
struct MIFace {
    int pos, neg;              // labels; boundary faces have pos in 0..3
    std::vector<int> vs;       // cycle, normal points to the pos side
    std::vector<int> es;       // third label meeting at edge vs[i]->vs[i+1]
};
struct MICut {
    int nlabels = 4;
    std::vector<std::array<int, 4>> verts;   // sorted label quadruples
    std::vector<MIFace> faces;
    std::vector<int> cells;                  // material label per cell
    bool dup = false;
    std::vector<int> grp;                    // representative label per label
};
MICut mi_cut(const std::vector<std::array<double, 4>>& materials);

// ---------------------------------------------------------------- pipeline
struct Func {
    int type;          // 0 plane, 1 sphere, 2 cylinder, 3 torus, 4 zero
    double p[12];
};
void grid(int res, const double bmin[3], const double bmax[3],
          std::vector<std::array<double, 3>>& pts, std::vector<std::array<uint64_t, 4>>& tets);
// row-major V x F
void eval(const std::vector<Func>& funcs, const std::vector<std::array<double, 3>>& pts,
          std::vector<double>& vals);

struct IsoVert {
    uint64_t tet, tet_vert;
    int simplex_size;
    uint64_t sv[4];
    uint64_t fn[4];        // IA: 3 used; MI: 4 used
};
struct IsoFace {
    std::vector<uint64_t> vs;
    std::vector<std::pair<uint64_t, uint64_t>> tet_faces;
    uint64_t f0, f1;
};

struct Result {
    bool mi = false;
    uint64_t nv = 0, nt = 0, nf = 0;
    // filter
    // IA: zero (vertex, function) entries; MI: vertices with tied maxima
    uint64_t num_degenerate = 0, num_intersecting = 0;
    std::vector<uint64_t> func_in_tet, start_index;
    // MI highest material
    std::vector<uint64_t> highest;
    std::vector<uint8_t> degenerate_vertex;
    // cut buckets (IA: 1,2,>=3 ; MI: 2,3,>=4)
    uint64_t nb[3] = {0, 0, 0};
    std::vector<int64_t> cut_index;      // per tet, -1 if empty
    std::vector<IACut> ia_cuts;
    std::vector<MICut> mi_cuts;
    // extraction
    std::vector<IsoVert> verts;
    std::vector<IsoFace> faces;
    std::vector<std::array<double, 3>> pts;
    // stage seconds, in the reference's timing label order
    std::vector<std::pair<const char*, double>> timings;
};

// F_LOOKUP: one-function IA tets and two-material MI tets take their complex
// from a table indexed by corner signs (the reference's default lookup-table
// path, app/implicit_arrangement.cpp:29-34, src/implicit_arrangement.cpp:276-283;
// src/material_interface.cpp:318-328).  Used by the CPU baseline; the tests
// check it against the general cut.
enum Flags { F_KEEP_CUTS = 1, F_REVERSE = 2, F_LOOKUP = 4 };

// The table: the complex of one plane (IA) or two materials (MI) depends only
// on the corner signs of the plane (of material 1 minus material 0), so the
// 81 patterns are cut once with rows of -1/0/+1.
const IACut& ia_table1(int pattern);
const MICut& mi_table2(int pattern);

void run_ia(const std::vector<std::array<double, 3>>& pts,
            const std::vector<std::array<uint64_t, 4>>& tets,
            const std::vector<double>& vals, uint64_t F, int flags, Result& r);
void run_mi(const std::vector<std::array<double, 3>>& pts,
            const std::vector<std::array<uint64_t, 4>>& tets,
            const std::vector<double>& vals, uint64_t F, int flags, Result& r);

// record serialisation in the layout documented in include/rin.h
void serialize_ia(const IACut& c, std::vector<uint16_t>& out);
void serialize_mi(const MICut& c, std::vector<uint16_t>& out);

}  // namespace oracle
