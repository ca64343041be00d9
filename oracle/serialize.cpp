// ORACLE — test infrastructure only.
// Writes the oracle's per-tet complexes in the record layout the product's
// host graph reads (layout documented in include/rin/cut_record.h).
#include <stdexcept>

#include "oracle.h"

namespace oracle {

static void finish(std::vector<uint16_t>& out, size_t start) {
    size_t len = out.size() - start;
    if (len > 0xFFFF) throw std::runtime_error("oracle: record too long");
    out[start + 6] = (uint16_t)len;
}

void serialize_ia(const IACut& c, std::vector<uint16_t>& out) {
    size_t s = out.size();
    out.insert(out.end(), {1, (uint16_t)c.nplanes, (uint16_t)c.verts.size(), (uint16_t)c.faces.size(),
                           (uint16_t)c.ncells, (uint16_t)(c.dup ? 1 : 0), 0, 0});
    for (const auto& v : c.verts) out.insert(out.end(), {(uint16_t)v[0], (uint16_t)v[1], (uint16_t)v[2], 0xFFFF});
    for (const auto& f : c.faces) {
        out.insert(out.end(), {(uint16_t)f.plane, (uint16_t)f.pos, (uint16_t)(f.neg < 0 ? 0xFFFF : f.neg),
                               (uint16_t)f.vs.size()});
        for (int v : f.vs) out.push_back((uint16_t)v);
    }
    if (c.dup) {
        for (int p = 0; p < c.nplanes; ++p) out.push_back((uint16_t)c.grp[p]);
        for (int p = 0; p < c.nplanes; ++p) out.push_back((uint16_t)c.ori[p]);
    }
    finish(out, s);
}

void serialize_mi(const MICut& c, std::vector<uint16_t>& out) {
    size_t s = out.size();
    out.insert(out.end(), {2, (uint16_t)c.nlabels, (uint16_t)c.verts.size(), (uint16_t)c.faces.size(),
                           (uint16_t)c.cells.size(), (uint16_t)(c.dup ? 1 : 0), 0, 0});
    for (const auto& v : c.verts)
        out.insert(out.end(), {(uint16_t)v[0], (uint16_t)v[1], (uint16_t)v[2], (uint16_t)v[3]});
    for (const auto& f : c.faces) {
        out.insert(out.end(), {(uint16_t)f.pos, (uint16_t)f.neg, 0xFFFF, (uint16_t)f.vs.size()});
        for (int v : f.vs) out.push_back((uint16_t)v);
    }
    for (int x : c.cells) out.push_back((uint16_t)x);
    if (c.dup)
        for (int l = 0; l < c.nlabels; ++l) out.push_back((uint16_t)c.grp[l]);
    finish(out, s);
}

}  // namespace oracle
