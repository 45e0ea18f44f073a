// The topology builder's cell graph and its per-cell passes, shared by the host builder and the
// device build kernels (build_device.hip), so that both evaluate the same operations in the
// same order (no floating-point contraction on either side) and produce bit-identical arrays.
//
// Cells are numbered leaf * 512 + i + 8 j + 64 k (before the final brick order: the builder's
// leaf order; after it: the brick order, i.e. the dof).  A leaf brick's same-level face
// neighbour comes from its level's brick grid; across a coarse/fine face the neighbour is the
// covering coarse cell or the four fine cells.
#pragma once

#include "solid.hpp"

namespace pcfd {

// brick status in a level's brick grid: a leaf index >= 0, or one of these
constexpr int ST_ABSENT = -2, ST_REFINED = -1, ST_EMPTY = -3;

// error codes of the per-cell passes (raised as BuildError by the host)
enum GraphError : int {
    GE_NONE = 0,
    GE_ABSENT_LEVEL0 = 1,      // internal: an absent level-0 brick
    GE_NESTING = 2,            // internal: AMR nesting violated
    GE_CUT_AT_BOUNDARY = 3,    // a cut cell touches the domain boundary
    GE_COINCIDENT = 4,         // a compact face between coincident centroids
};
inline const char* graph_error_text(int e) {
    switch (e) {
        case GE_ABSENT_LEVEL0: return "internal: absent level-0 brick";
        case GE_NESTING: return "internal: AMR nesting violated";
        case GE_CUT_AT_BOUNDARY: return "a cut cell touches the domain boundary; enlarge the domain";
        case GE_COINCIDENT: return "compact face between coincident centroids";
        default: return "internal: unknown graph error";
    }
}

struct BuildView {
    int levels = 1;
    double lo[3] = {0, 0, 0};
    double h[4] = {0, 0, 0, 0};
    int nb[4][3] = {};             // brick grid of each level
    const int* status[4] = {};     // per level: status of each brick (x fastest)
    const int* leaf_level = nullptr;
    const int* leaf_coord = nullptr;   // 3 per leaf
    const int* cut = nullptr;          // per cell: index into cuts, or -1
    const CutData* cuts = nullptr;
    const uint8_t* active = nullptr;   // per cell
};

// A face neighbour: its leaf and local cell, the open fraction of the shared face, its level and
// the centroid of the face's open part (the open-sample mean of the cut cell that sampled the
// face, else the face centre).
struct Neighbor {
    int leaf;
    int local;
    double ap;
    int level;
    double xf[3];
};

PCFD_HD inline int brick_status(const BuildView& V, int l, int x, int y, int z) {
    return V.status[l][(static_cast<int64_t>(z) * V.nb[l][1] + y) * V.nb[l][0] + x];
}

// Neighbours of cell (leaf, local) across face dir (0..5: -x, +x, -y, +y, -z, +z): one at the
// same or the coarser level, up to four at the finer one.  Returns their number, or -GraphError;
// boundary is set when the face lies on the domain boundary.
PCFD_HD inline int graph_neighbors(const BuildView& V, int leaf, int local, int dir, Neighbor out[4], bool& boundary) {
    boundary = false;
    const int l = V.leaf_level[leaf];
    const int axis = dir >> 1, sign = (dir & 1) ? 1 : -1;
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const int* lc = V.leaf_coord + 3 * leaf;
    const int I[3] = {lc[0] * 8 + (local & 7), lc[1] * 8 + ((local >> 3) & 7), lc[2] * 8 + (local >> 6)};
    int J[3] = {I[0], I[1], I[2]};
    J[axis] += sign;
    if (J[axis] < 0 || J[axis] >= V.nb[l][axis] * 8) {
        boundary = true;
        return 0;
    }
    const int st = brick_status(V, l, J[0] >> 3, J[1] >> 3, J[2] >> 3);
    // the centroid of the open part of a face on this cell's face plane: the mean of the open
    // samples of cut `cut` on its face `cdir` when the face is partly open, else the centre of
    // the face of the cell at integer position C with size hc
    const double plane = V.lo[axis] + (I[axis] + (sign > 0 ? 1 : 0)) * V.h[l];
    auto face_centroid = [&](double* xf, int cut, int cdir, const int* C, double hc) {
        xf[axis] = plane;
        if (cut >= 0 && V.cuts[cut].ap[cdir] > 0.0 && V.cuts[cut].ap[cdir] < 1.0) {
            xf[t1] = V.cuts[cut].fc[cdir][0];
            xf[t2] = V.cuts[cut].fc[cdir][1];
        } else {
            xf[t1] = V.lo[t1] + (C[t1] + 0.5) * hc;
            xf[t2] = V.lo[t2] + (C[t2] + 0.5) * hc;
        }
    };
    auto loc = [](const int c[3]) { return (c[0] & 7) + 8 * (c[1] & 7) + 64 * (c[2] & 7); };
    const int ci = V.cut[static_cast<int64_t>(leaf) * 512 + local];
    if (st >= 0) {
        // same level: both sides sampled the same face points
        const int jq = loc(J);
        const int cj = V.cut[static_cast<int64_t>(st) * 512 + jq];
        out[0].leaf = st;
        out[0].local = jq;
        out[0].level = l;
        out[0].ap = ci >= 0 ? V.cuts[ci].ap[dir] : (cj >= 0 ? V.cuts[cj].ap[dir ^ 1] : 1.0);
        if (ci >= 0) face_centroid(out[0].xf, ci, dir, I, V.h[l]);
        else face_centroid(out[0].xf, cj, dir ^ 1, I, V.h[l]);
        return 1;
    }
    if (st == ST_EMPTY) return 0;
    if (st == ST_REFINED) {
        // four finer cells across the face; each sampled its own (fine) face
        int n = 0;
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
                int K[3];
                K[axis] = sign > 0 ? 2 * J[axis] : 2 * J[axis] + 1;
                K[t1] = 2 * J[t1] + a;
                K[t2] = 2 * J[t2] + b;
                const int fs = brick_status(V, l + 1, K[0] >> 3, K[1] >> 3, K[2] >> 3);
                if (fs < 0) continue;  // empty brick
                const int fq = loc(K);
                const int cf = V.cut[static_cast<int64_t>(fs) * 512 + fq];
                out[n].leaf = fs;
                out[n].local = fq;
                out[n].level = l + 1;
                out[n].ap = cf >= 0 ? V.cuts[cf].ap[dir ^ 1] : 1.0;
                face_centroid(out[n].xf, cf, dir ^ 1, K, V.h[l + 1]);
                ++n;
            }
        return n;
    }
    // absent at level l: the coarser leaf at level l-1 covers it; this (fine) cell sampled the face
    if (l == 0) return -GE_ABSENT_LEVEL0;
    const int K[3] = {J[0] >> 1, J[1] >> 1, J[2] >> 1};
    const int cs = brick_status(V, l - 1, K[0] >> 3, K[1] >> 3, K[2] >> 3);
    if (cs == ST_EMPTY) return 0;
    if (cs < 0) return -GE_NESTING;
    out[0].leaf = cs;
    out[0].local = loc(K);
    out[0].level = l - 1;
    out[0].ap = ci >= 0 ? V.cuts[ci].ap[dir] : 1.0;
    face_centroid(out[0].xf, ci, dir, I, V.h[l]);
    return 1;
}

// Morton order of a cell or brick position (21 bits per axis)
PCFD_HD inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
PCFD_HD inline uint64_t morton(int x, int y, int z) {
    return spread3(static_cast<uint64_t>(x)) | (spread3(static_cast<uint64_t>(y)) << 1) |
           (spread3(static_cast<uint64_t>(z)) << 2);
}

// Multigrid cell key: size exponent (cell size = finest h * 2^e) above the Morton order of the
// cell's position at that size.  Sorting the keys orders a coarse level.
PCFD_HD inline uint64_t mg_key(int e, int x, int y, int z) { return (static_cast<uint64_t>(e) << 60) | morton(x, y, z); }

// the coarsest multigrid level is solved exactly (dense inverse) below this many cells
constexpr int kCoarseMax = 2048;

PCFD_HD inline double clamp01(double w) { return w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w); }
PCFD_HD inline bool finite_value(double v) { return v == v && v - v == 0.0; }

// Weight of the upper cell when a cell value is interpolated to a face at x_f between the
// centroids c_l and c_u (composite_face_interpolation_weight, amr_advection.cu:732-753): the
// face's position between the centroids along the axis when it lies between them (within
// 1e-8), else the distance weight |x_f - c_l| / (|x_f - c_l| + |x_f - c_u|); clamped to [0, 1].
// err = GE_COINCIDENT for coincident centroids.
PCFD_HD inline double face_weight(const double* cl, const double* cu, const double* xf, int axis, int& err) {
    double d2 = 0;
    for (int r = 0; r < 3; ++r) d2 += (cu[r] - cl[r]) * (cu[r] - cl[r]);
    const double dist = std::sqrt(d2), normal = cu[axis] - cl[axis];
    if (!(dist > 0) || !finite_value(dist)) {
        err = GE_COINCIDENT;
        return 0.5;
    }
    if ((normal < 0 ? -normal : normal) > 1e-12 * dist) {
        const double w = (xf[axis] - cl[axis]) / normal;
        if (finite_value(w) && w >= -1e-8 && w <= 1 + 1e-8) return clamp01(w);
    }
    double dl = 0, du = 0;
    for (int r = 0; r < 3; ++r) {
        dl += (xf[r] - cl[r]) * (xf[r] - cl[r]);
        du += (xf[r] - cu[r]) * (xf[r] - cu[r]);
    }
    dl = std::sqrt(dl);
    du = std::sqrt(du);
    const double total = dl + du;
    if (!(total > 1e-14 * (dist > 1.0 ? dist : 1.0)) || !finite_value(total)) return 0.5;
    return clamp01(dl / total);
}

// Cell a = brick * 512 + q after the final brick order (leaf = brick).  Volume, centroid and
// the embedded flag: a partial volume, or a partially open face to an active neighbour (the
// endpoints of the reference's embedded connections, amr_advection.cu:799).
struct CellGeometry {
    double volume;
    double centroid[3];
    uint8_t embedded;
};
PCFD_HD inline int cell_geometry(const BuildView& V, int b, int q, CellGeometry& g) {
    const int l = V.leaf_level[b];
    const double h = V.h[l];
    const int* lc = V.leaf_coord + 3 * b;
    const int I[3] = {lc[0] * 8 + (q & 7), lc[1] * 8 + ((q >> 3) & 7), lc[2] * 8 + (q >> 6)};
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const int ci = V.cut[a];
    const bool act = V.active[a] != 0;
    g.volume = act ? (ci >= 0 ? V.cuts[ci].f : 1.0) * h * h * h : 0.0;
    for (int ax = 0; ax < 3; ++ax) g.centroid[ax] = act && ci >= 0 ? V.cuts[ci].c[ax] : V.lo[ax] + (I[ax] + 0.5) * h;
    g.embedded = 0;
    if (!act) return GE_NONE;
    if (g.volume != h * h * h) g.embedded = 1;
    for (int dir = 0; dir < 6; ++dir) {
        Neighbor nbs[4];
        bool bnd;
        const int n = graph_neighbors(V, b, q, dir, nbs, bnd);
        if (n < 0) return -n;
        for (int k = 0; k < n; ++k)
            if (V.active[static_cast<int64_t>(nbs[k].leaf) * 512 + nbs[k].local] && nbs[k].ap > 0.0 && nbs[k].ap < 1.0)
                g.embedded = 1;
    }
    return GE_NONE;
}

// One compact edge as the face pass creates it
struct EdgeRecord {
    int a, b;          // lower and upper cell along the axis
    int axis;
    double area, coef, dcoef, w;
    double x[3];       // centroid of the face's open part
};

// Faces of active cell a: mask bits, diffusion conductance sum, closure area vector S and the
// first moment fm = sum over open faces of (x_f - c_a) (x) (A n)_f (row-major), and the compact
// edges the cell creates (same level: from the lower cell; coarse/fine: from the fine cell),
// passed to emit(const EdgeRecord&) in face, then neighbour order.  centroid and embedded are
// the per-cell arrays of cell_geometry.
struct CellFaces {
    uint8_t mask;
    uint8_t has_compact;
    double diffusion_sum;
    double S[3];
    double fm[9];
};
template <class Emit>
PCFD_HD inline int cell_faces(const BuildView& V, const double* centroid, const uint8_t* embedded, int b, int q,
                              CellFaces& f, Emit&& emit) {
    constexpr uint8_t kActive = 64, kPX = 1, kMX = 8;
    const int l = V.leaf_level[b];
    const double h = V.h[l];
    const int* lc = V.leaf_coord + 3 * b;
    const int I[3] = {lc[0] * 8 + (q & 7), lc[1] * 8 + ((q >> 3) & 7), lc[2] * 8 + (q >> 6)};
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    f.mask = kActive;
    f.has_compact = 0;
    f.diffusion_sum = 0;
    for (int r = 0; r < 3; ++r) f.S[r] = 0;
    for (int r = 0; r < 9; ++r) f.fm[r] = 0;
    const int ci = V.cut[a];
    const double* ca = centroid + 3 * a;
    for (int dir = 0; dir < 6; ++dir) {
        const int axis = dir >> 1, up = dir & 1;
        const double nsign = up ? 1.0 : -1.0;
        Neighbor nbs[4];
        bool bnd;
        const int n = graph_neighbors(V, b, q, dir, nbs, bnd);
        if (n < 0) return -n;
        double open = 0;
        if (bnd) {
            if (ci >= 0 && V.cuts[ci].ap[dir] < 1.0) return GE_CUT_AT_BOUNDARY;
            open = h * h;
            for (int r = 0; r < 3; ++r) {
                const double xf = r == axis ? V.lo[r] + (I[r] + up) * h : V.lo[r] + (I[r] + 0.5) * h;
                f.fm[3 * r + axis] += (xf - ca[r]) * nsign * open;
            }
        }
        for (int k = 0; k < n; ++k) {
            const Neighbor& nb2 = nbs[k];
            const int64_t other = static_cast<int64_t>(nb2.leaf) * 512 + nb2.local;
            if (!V.active[other] || nb2.ap <= 0) continue;
            const double hn = V.h[nb2.level];
            const double hf = h < hn ? h : hn;
            const double area = nb2.ap * hf * hf;
            open += area;
            for (int r = 0; r < 3; ++r) f.fm[3 * r + axis] += (nb2.xf[r] - ca[r]) * nsign * area;
            const double coef = area / (0.5 * (h + hn));
            const bool same = nb2.level == l;
            if (same && nb2.ap >= 1.0 && !embedded[a] && !embedded[other]) {
                f.mask |= up ? (kPX << axis) : (kMX << axis);
                f.diffusion_sum += coef;
                continue;
            }
            const int64_t lower = up ? a : other, upper = up ? other : a;
            const double* cl = centroid + 3 * lower;
            const double* cu = centroid + 3 * upper;
            // diffusion conductance: open area |dx_axis| / |dx|^2 on the centroids for a fully
            // open same-level face (the reference's perimeter conductance, amr_advection.cu:1180),
            // the operator's coefficient on coarse/fine and partially open faces
            // (connection_conductance, :1183)
            double dcoef = coef;
            if (same && nb2.ap >= 1.0) {
                double d2 = 0;
                for (int r = 0; r < 3; ++r) d2 += (cu[r] - cl[r]) * (cu[r] - cl[r]);
                const double dd = cu[axis] - cl[axis], dn = dd < 0 ? -dd : dd;
                if (!(d2 > 0) || !(dn > 0)) return GE_COINCIDENT;
                dcoef = area * dn / d2;
            }
            f.diffusion_sum += dcoef;
            f.has_compact = 1;
            // compact edge, created once: same level from the lower cell, c/f from the fine cell
            const bool create = same ? up != 0 : nb2.level < l;
            if (!create) continue;
            EdgeRecord e;
            e.a = static_cast<int>(lower);
            e.b = static_cast<int>(upper);
            e.axis = axis;
            e.area = area;
            e.coef = coef;
            e.dcoef = dcoef;
            int err = GE_NONE;
            e.w = face_weight(cl, cu, nb2.xf, axis, err);
            if (err != GE_NONE) return err;
            for (int r = 0; r < 3; ++r) e.x[r] = nb2.xf[r];
            emit(e);
        }
        f.S[axis] -= nsign * open;
    }
    return GE_NONE;
}

// |S| of a cell and the wall test (|S| > 1e-12 h^2)
PCFD_HD inline double closure_norm(const double S[3]) { return std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]); }

}  // namespace pcfd
