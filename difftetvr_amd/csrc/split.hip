// Coarse-to-fine refinement of a tet mesh (host C++): the vertices with the largest gradient magnitudes are
// subdivided one after another.  Subdividing vertex a inserts the midpoint of every edge incident to a (position
// and colour interpolated at t = 0.5) and replaces each tet (a, b, c, d) incident to a by the corner tet
// (a, ab, ac, ad) and the prism (b, c, d | ab, ac, ad), split into three tets.
//
// Follows TetMesh::splitByLargestGradientMagnitudes / subdivideVertices / subdivideAtVertex
// (src/Tet/TetMesh.cpp:1786-1879, :1757-1784, :1374-1640).  The reference chooses the prism diagonals by a
// constraint solver over the incident prisms (FlipSolver, PRISM_TO_TET_TABLE :1333-1350).  Here every quad
// face takes the diagonal through its vertex of smallest index (Dompierre, Labbe, Vallet, Camarero, "How to
// subdivide pyramids, prisms and hexahedra into tetrahedra", IMR 1999): two prisms sharing a quad see the same
// four indices and so agree, and the rule never yields the one cyclic (unsplittable) pattern.  Each new tet takes
// the orientation sign of the tet it replaces.
//
// Ordering: tets that survive keep their order, new tets are appended in creation order (per incident tet in
// increasing tet index: the corner tet, then the three prism tets), new vertices are appended per split in
// increasing index of the edge's other endpoint.  The reference's orders come from OpenVolumeMesh's incidence
// iterators, so the element order may differ from it while the refined mesh is the same up to the diagonals.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/difftetvr_amd.h"
#include "dtv_kernels.h"

struct dtv_split_result {
    std::vector<std::array<int32_t, 4>> cells;
    std::vector<float> positions;  // [V,3]
    std::vector<float> colors;     // [V,4]
};

namespace {

using Tet = std::array<int32_t, 4>;

struct Refiner {
    std::vector<Tet> tets;                  // every tet ever created, in creation order
    std::vector<uint8_t> alive;
    std::vector<std::vector<int32_t>> inc;  // vertex -> tets containing it (alive or not; filtered on use)
    std::vector<float> pos, col;

    int32_t num_vertices() const { return (int32_t)(pos.size() / 3); }

    double orient(const Tet& t) const {
        double p[4][3];
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) p[i][k] = pos[3 * (size_t)t[i] + k];
        const double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        const double e3[3] = {p[3][0] - p[0][0], p[3][1] - p[0][1], p[3][2] - p[0][2]};
        const double cx = e1[1] * e2[2] - e1[2] * e2[1];
        const double cy = e1[2] * e2[0] - e1[0] * e2[2];
        const double cz = e1[0] * e2[1] - e1[1] * e2[0];
        return cx * e3[0] + cy * e3[1] + cz * e3[2];
    }

    void add_tet(Tet t, double parent_sign) {
        if ((orient(t) < 0.0) != (parent_sign < 0.0)) std::swap(t[2], t[3]);
        const int32_t id = (int32_t)tets.size();
        tets.push_back(t);
        alive.push_back(1);
        for (int32_t v : t) inc[(size_t)v].push_back(id);
    }

    int32_t add_midpoint(int32_t a, int32_t b) {
        // glm::mix(x, y, 0.5) = x * 0.5 + y * 0.5, the same in either order
        for (int k = 0; k < 3; k++) pos.push_back(pos[3 * (size_t)a + k] * 0.5f + pos[3 * (size_t)b + k] * 0.5f);
        for (int k = 0; k < 4; k++) col.push_back(col[4 * (size_t)a + k] * 0.5f + col[4 * (size_t)b + k] * 0.5f);
        inc.emplace_back();
        return num_vertices() - 1;
    }

    // The prism (P0 P1 P2 | P3 P4 P5), Pi and Pi+3 joined by an edge, as three tets.  m is its vertex of
    // smallest index: both quads through m are cut by diagonals from m, so the tet of m and the opposite
    // triangle separates, leaving the pyramid (apex m) over the third quad, whose diagonal again goes through
    // that quad's smallest vertex.
    void split_prism(const int32_t P[6], double sign) {
        int mi = 0;
        for (int i = 1; i < 6; i++)
            if (P[i] < P[mi]) mi = i;
        const int tri = mi < 3 ? 0 : 3;         // the triangle holding m
        const int oth = 3 - tri;                // the opposite triangle
        const int k = mi - tri;                 // position of m in its triangle
        const int32_t m = P[mi], mp = P[oth + k];
        const int32_t x = P[tri + (k + 1) % 3], y = P[tri + (k + 2) % 3];
        const int32_t xp = P[oth + (k + 1) % 3], yp = P[oth + (k + 2) % 3];
        add_tet({m, mp, xp, yp}, sign);
        const int32_t qmin = std::min(std::min(x, y), std::min(xp, yp));
        if (qmin == x || qmin == yp) {          // diagonal x-y'
            add_tet({m, x, y, yp}, sign);
            add_tet({m, x, yp, xp}, sign);
        } else {                                // diagonal y-x'
            add_tet({m, x, y, xp}, sign);
            add_tet({m, y, yp, xp}, sign);
        }
    }

    void subdivide_at_vertex(int32_t a) {
        std::vector<int32_t> cells;
        for (int32_t t : inc[(size_t)a])
            if (alive[(size_t)t]) cells.push_back(t);
        if (cells.empty()) return;
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        std::vector<int32_t> nbrs;
        for (int32_t t : cells)
            for (int32_t v : tets[(size_t)t])
                if (v != a) nbrs.push_back(v);
        std::sort(nbrs.begin(), nbrs.end());
        nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
        std::vector<int32_t> mid(nbrs.size());
        for (size_t i = 0; i < nbrs.size(); i++) mid[i] = add_midpoint(a, nbrs[i]);
        auto mid_of = [&](int32_t v) {
            return mid[(size_t)(std::lower_bound(nbrs.begin(), nbrs.end(), v) - nbrs.begin())];
        };
        for (int32_t t : cells) {
            const Tet old = tets[(size_t)t];
            const double sign = orient(old);
            alive[(size_t)t] = 0;
            int32_t o[3], n = 0;
            for (int32_t v : old)
                if (v != a) o[n++] = v;
            const int32_t P[6] = {o[0], o[1], o[2], mid_of(o[0]), mid_of(o[1]), mid_of(o[2])};
            add_tet({a, P[3], P[4], P[5]}, sign);
            split_prism(P, sign);
        }
    }
};

int check_mesh(const char* f, const int32_t* cells, int64_t T, const float* pos, const float* col, int64_t V) {
    if (!cells || !pos || !col) return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": NULL mesh array");
    if (T < 0 || V <= 0 || V > (int64_t)1 << 30 || T > (int64_t)1 << 30)
        return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": bad mesh size");
    for (int64_t i = 0; i < 4 * T; i++)
        if (cells[i] < 0 || cells[i] >= V)
            return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": cell index out of range");
    return DTV_OK;
}

int refine(const int32_t* cells, int64_t T, const float* pos, const float* col, int64_t V, const float* mags,
           int64_t num_splits, dtv_split_result** out) {
    // descending (magnitude, index), as std::sort over the pairs read back to front (TetMesh.cpp:1769-1773)
    // (a NaN magnitude sorts last: the comparison needs a strict weak order)
    std::vector<std::pair<float, int32_t>> order((size_t)V);
    for (int64_t i = 0; i < V; i++) order[(size_t)i] = {std::isnan(mags[i]) ? -INFINITY : mags[i], (int32_t)i};
    std::sort(order.rbegin(), order.rend());
    num_splits = std::min<int64_t>(std::max<int64_t>(num_splits, 0), V);

    Refiner R;
    R.pos.assign(pos, pos + 3 * V);
    R.col.assign(col, col + 4 * V);
    R.inc.resize((size_t)V);
    R.tets.reserve((size_t)T + 16 * (size_t)num_splits);
    for (int64_t t = 0; t < T; t++) {
        const Tet c = {cells[4 * t], cells[4 * t + 1], cells[4 * t + 2], cells[4 * t + 3]};
        R.tets.push_back(c);
        R.alive.push_back(1);
        for (int32_t v : c) R.inc[(size_t)v].push_back((int32_t)t);
    }
    for (int64_t i = 0; i < num_splits; i++) R.subdivide_at_vertex(order[(size_t)i].second);

    auto res = std::make_unique<dtv_split_result>();
    for (size_t t = 0; t < R.tets.size(); t++)
        if (R.alive[t]) res->cells.push_back(R.tets[t]);
    res->positions = std::move(R.pos);
    res->colors = std::move(R.col);
    *out = res.release();
    return DTV_OK;
}

}  // namespace

extern "C" {

int dtv_subdivide_vertices(const int32_t* cells, int64_t num_cells, const float* positions, const float* colors,
                           int64_t num_vertices, const float* gradient_magnitudes, int64_t num_splits,
                           dtv_split_result** out) {
    const char* f = "dtv_subdivide_vertices";
    if (!gradient_magnitudes || !out) return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": NULL argument");
    const int rc = check_mesh(f, cells, num_cells, positions, colors, num_vertices);
    if (rc != DTV_OK) return rc;
    return refine(cells, num_cells, positions, colors, num_vertices, gradient_magnitudes, num_splits, out);
}

int dtv_split_by_largest_gradient_magnitudes(const int32_t* cells, int64_t num_cells, const float* positions,
                                             const float* colors, int64_t num_vertices, const float* gradients,
                                             int32_t gradient_components, float splits_ratio,
                                             dtv_split_result** out) {
    const char* f = "dtv_split_by_largest_gradient_magnitudes";
    if (!gradients || !out) return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": NULL argument");
    if (gradient_components != 3 && gradient_components != 4)
        return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": gradients must have 3 or 4 components");
    if (!(splits_ratio >= 0.0f)) return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, std::string(f) + ": negative splits_ratio");
    const int rc = check_mesh(f, cells, num_cells, positions, colors, num_vertices);
    if (rc != DTV_OK) return rc;
    // computeVectorMagnitudeField: glm::length of each row
    std::vector<float> mags((size_t)num_vertices);
    for (int64_t i = 0; i < num_vertices; i++) {
        float s = 0.0f;
        for (int k = 0; k < gradient_components; k++) {
            const float g = gradients[i * gradient_components + k];
            s += g * g;
        }
        mags[(size_t)i] = std::sqrt(s);
    }
    const int64_t num_splits = (int64_t)std::ceil((double)splits_ratio * (double)num_vertices);
    return refine(cells, num_cells, positions, colors, num_vertices, mags.data(), num_splits, out);
}

int dtv_split_result_sizes(const dtv_split_result* r, int64_t* num_cells, int64_t* num_vertices) {
    if (!r || !num_cells || !num_vertices)
        return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, "dtv_split_result_sizes: NULL argument");
    *num_cells = (int64_t)r->cells.size();
    *num_vertices = (int64_t)(r->positions.size() / 3);
    return DTV_OK;
}

int dtv_split_result_copy(const dtv_split_result* r, int32_t* cells, float* positions, float* colors) {
    if (!r || !cells || !positions || !colors)
        return dtv::report_error(DTV_ERR_INVALID_ARGUMENT, "dtv_split_result_copy: NULL argument");
    if (!r->cells.empty()) std::memcpy(cells, r->cells.data(), r->cells.size() * sizeof(Tet));
    std::memcpy(positions, r->positions.data(), r->positions.size() * sizeof(float));
    std::memcpy(colors, r->colors.data(), r->colors.size() * sizeof(float));
    return DTV_OK;
}

int dtv_split_result_destroy(dtv_split_result* r) {
    delete r;
    return DTV_OK;
}

}  // extern "C"
