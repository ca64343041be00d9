#include "cut_view.h"

namespace rin {

CutView decode_cut(const uint16_t* r) {
    CutView C;
    if (r[0] != 1 && r[0] != 2) throw std::runtime_error("rin: bad cut record kind");
    C.mi = r[0] == 2;
    C.nlabels = r[1];
    const size_t nv = r[2], nf = r[3], nc = r[4];
    const bool dup = r[5] & 1;
    size_t p = RIN_REC_HEADER;
    C.vertices.resize(nv);
    for (size_t v = 0; v < nv; ++v, p += 4)
        for (int q = 0; q < 4; ++q) C.vertices[v][q] = (r[p + q] == RIN_REC_NONE) ? Mesh_None : r[p + q];
    C.faces.resize(nf);
    for (size_t f = 0; f < nf; ++f) {
        CutFace& F = C.faces[f];
        uint16_t a = r[p], b = r[p + 1], c = r[p + 2], n = r[p + 3];
        p += 4;
        if (C.mi) {
            F.positive_material_label = a;
            F.negative_material_label = b;
        } else {
            F.supporting_plane = a;
            F.positive_cell = b;
            F.negative_cell = (c == RIN_REC_NONE) ? Mesh_None : c;
        }
        F.vertices.assign(r + p, r + p + n);
        p += n;
    }
    C.cells.resize(nc);
    if (C.mi) {
        for (size_t c = 0; c < nc; ++c) C.cells[c].material_label = r[p + c];
        p += nc;
        for (size_t c = 0; c < nc; ++c) {
            size_t lab = C.cells[c].material_label;
            for (size_t f = 0; f < nf; ++f)
                if (C.faces[f].positive_material_label == lab || C.faces[f].negative_material_label == lab)
                    C.cells[c].faces.push_back(f);
        }
    } else {
        for (size_t f = 0; f < nf; ++f) {
            if (C.faces[f].positive_cell < nc) C.cells[C.faces[f].positive_cell].faces.push_back(f);
            if (C.faces[f].negative_cell < nc) C.cells[C.faces[f].negative_cell].faces.push_back(f);
        }
        // a cell lists its faces in ascending face order
        for (auto& cell : C.cells) {
            auto& fs = cell.faces;
            for (size_t i = 1; i < fs.size(); ++i)
                for (size_t j = i; j > 0 && fs[j - 1] > fs[j]; --j) std::swap(fs[j - 1], fs[j]);
        }
    }
    if (dup) {
        const size_t L = C.nlabels;
        std::vector<size_t> group_of_rep(L, Mesh_None);
        C.unique_plane_indices.assign(L, Mesh_None);
        for (size_t l = 0; l < L; ++l) {
            size_t rep = r[p + l];
            if (rep >= L) throw std::runtime_error("rin: bad group in cut record");
            if (group_of_rep[rep] == Mesh_None) {
                group_of_rep[rep] = C.unique_planes.size();
                C.unique_planes.emplace_back();
            }
            C.unique_plane_indices[l] = group_of_rep[rep];
            C.unique_planes[group_of_rep[rep]].push_back(l);
        }
        p += L;
        if (!C.mi) {
            C.unique_plane_orientations.resize(L);
            for (size_t l = 0; l < L; ++l) C.unique_plane_orientations[l] = r[p + l] != 0;
            p += L;
        }
    }
    if (p != r[6]) throw std::runtime_error("rin: cut record length mismatch");
    return C;
}

}  // namespace rin
