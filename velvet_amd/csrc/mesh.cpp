// Constraint generation for an arbitrary triangle mesh (host only; no device work).
//
// The reference generates constraints from grid indices only -- structural and shear springs,
// one bending constraint per quad with rest angle 0 (Velvet/VtClothObjectGPU.hpp:75-132, with a
// TODO for real rest angles at 128-129). Here any triangle list gives:
//   * stretch: every unique edge, rest = current length;
//   * bending: every edge shared by exactly two triangles, as (p0, p1, p2, p3) with (p2, p3) the
//     edge in the first triangle's winding, p0 that triangle's third vertex and p1 the other
//     triangle's third vertex, rest = the current angle between the two triangle normals
//     c1 = (p2-p0) x (p3-p0), c2 = (p3-p1) x (p2-p1) -- the quantity the solver's dihedral
//     constraint drives (Velvet/VtClothSolverGPU.cu:128-177) -- so the given shape is at rest.
// Edges are numbered in order of first appearance (triangle order, then edges (0,1), (1,2),
// (2,0)). Edges of more than two triangles get no bending constraint; degenerate triangles
// (zero normal) give none either.
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace vtx {

struct MeshEdge {
    uint32_t a, b;       // in the first triangle's winding
    uint32_t opp0;       // third vertex of the first triangle
    uint32_t opp1;       // third vertex of the second triangle
    uint32_t tris;       // triangles sharing the edge
};

static double rest_angle(const float* x, uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, bool* ok) {
    auto P = [&](uint32_t i, int c) { return (double)x[3 * (size_t)i + c]; };
    double a[3], b[3], c[3], d[3];
    for (int k = 0; k < 3; k++) {
        a[k] = P(p2, k) - P(p0, k);
        b[k] = P(p3, k) - P(p0, k);
        c[k] = P(p3, k) - P(p1, k);
        d[k] = P(p2, k) - P(p1, k);
    }
    const double c1[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double c2[3] = {c[1] * d[2] - c[2] * d[1], c[2] * d[0] - c[0] * d[2], c[0] * d[1] - c[1] * d[0]};
    const double l1 = std::sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    const double l2 = std::sqrt(c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2]);
    *ok = l1 > 0.0 && l2 > 0.0;
    if (!*ok) return 0.0;
    double cd = (c1[0] * c2[0] + c1[1] * c2[1] + c1[2] * c2[2]) / (l1 * l2);
    cd = cd < -1.0 ? -1.0 : cd > 1.0 ? 1.0 : cd;
    return std::acos(cd);
}

void mesh_constraints(const float* x, uint32_t n, const uint32_t* tris, uint32_t ntri,
                      std::vector<int32_t>& st_ij, std::vector<float>& st_rest,
                      std::vector<uint32_t>& bd_idx, std::vector<float>& bd_rest) {
    VTX_REQUIRE(n == 0 || x, VTX_EINVAL, "positions are NULL");
    VTX_REQUIRE(ntri == 0 || tris, VTX_EINVAL, "triangles are NULL");
    std::vector<MeshEdge> edges;
    std::unordered_map<uint64_t, uint32_t> index;
    index.reserve((size_t)ntri * 2);
    for (uint32_t t = 0; t < ntri; t++) {
        const uint32_t* v = tris + 3 * (size_t)t;
        for (int k = 0; k < 3; k++)
            VTX_REQUIRE(v[k] < n, VTX_EINVAL, "triangle index out of range");
        for (int k = 0; k < 3; k++) {
            const uint32_t a = v[k], b = v[(k + 1) % 3], c = v[(k + 2) % 3];
            if (a == b) continue;
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = index.find(key);
            if (it == index.end()) {
                index.emplace(key, (uint32_t)edges.size());
                edges.push_back(MeshEdge{a, b, c, 0u, 1u});
            } else {
                MeshEdge& e = edges[it->second];
                if (e.tris == 1) e.opp1 = c;
                e.tris++;
            }
        }
    }
    st_ij.clear(); st_rest.clear(); bd_idx.clear(); bd_rest.clear();
    for (const MeshEdge& e : edges) {
        double l2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double d = (double)x[3 * (size_t)e.a + k] - (double)x[3 * (size_t)e.b + k];
            l2 += d * d;
        }
        st_ij.push_back((int32_t)e.a);
        st_ij.push_back((int32_t)e.b);
        st_rest.push_back((float)std::sqrt(l2));
    }
    for (const MeshEdge& e : edges) {
        if (e.tris != 2 || e.opp0 == e.opp1) continue;
        bool ok = false;
        const double r = rest_angle(x, e.opp0, e.opp1, e.a, e.b, &ok);
        if (!ok) continue;
        bd_idx.insert(bd_idx.end(), {e.opp0, e.opp1, e.a, e.b});
        bd_rest.push_back((float)r);
    }
}

}  // namespace vtx
