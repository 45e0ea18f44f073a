// Non-orthogonal correction of the composite pressure operator: per-face two-point factor
// from the control-volume centroids, the affine correction vector, and the weighted
// least-squares pressure-gradient stencils it needs (finalize_nonorthogonal_pressure_topology,
// amr_pressure.cpp:63-190).  Host preprocessing; the device applies it in the projection.
//
// For a face between cells L (lower) and U (upper) along axis a with centroid displacement
// d = c_U - c_L, the normal derivative is written as  dp/dn = (p_U - p_L) / d_a + k . g_f
// with k = e_a - d / d_a (k_a = 0); for a linear p both terms together are exact.  The
// orthogonal operator already carries the first term with the factor 1/h (structured faces)
// or coef/area (compact edges); the record stores the difference `delta` of the factors and k.
// g_f interpolates the two cell gradients with the weight of the face's position between the
// centroids along the axis.  Thresholds follow the reference's production build (FP32 fields):
// corrections of magnitude <= 16 FLT_EPSILON are dropped, and the stencil's normal matrix keeps
// eigenvalues above 1e-4 of the largest.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "topology.hpp"

namespace pcfd {
namespace {

struct Sample { int nbr; double w; };

// Pseudo-inverse of a symmetric 3x3 matrix by cyclic Jacobi rotations; returns the rank
// (eigenvalues above rel * largest).  inv may be null.
int sym_pinv3(const double* in, double* inv, double rel) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = in[3 * r + c];
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        double diag = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
        if (off <= 1e-15 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                // rotation that zeroes a[p][q] (Golub & Van Loan, symmetric Schur 2x2)
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {  // A <- A J
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- J^T A
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    const double top = std::max({std::fabs(a[0][0]), std::fabs(a[1][1]), std::fabs(a[2][2])});
    int rank = 0;
    if (inv) std::fill(inv, inv + 9, 0.0);
    for (int m = 0; m < 3; ++m) {
        const double lam = a[m][m];
        if (!(lam > rel * top)) continue;
        ++rank;
        if (inv)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) inv[3 * r + c] += v[r][m] * v[c][m] / lam;
    }
    return rank;
}

struct Geometry {
    const Topology& T;
    std::vector<int> special_of;  // dof -> special index (only for special dofs)
    explicit Geometry(const Topology& t) : T(t) {
        special_of.assign(T.ndof, -1);
        for (size_t s = 0; s < T.special_dof.size(); ++s) special_of[T.special_dof[s]] = static_cast<int>(s);
    }
    double cen(int64_t a, int ax) const { return T.centroid[3 * a + ax]; }
    double h_of(int64_t a) const { return T.level_h[T.brick_level[a >> 9]]; }
    // structured neighbour across the +axis (dir 1) or -axis (dir 0) face, or -1
    int64_t structured(int64_t a, int axis, int dir) const {
        const uint8_t m = T.mask[a];
        if (!(m & ((dir ? CM_PX : CM_MX) << axis))) return -1;
        const int b = static_cast<int>(a >> 9), q = static_cast<int>(a & 511);
        int loc[3] = {q & 7, (q >> 3) & 7, q >> 6};
        int nb = b;
        loc[axis] += dir ? 1 : -1;
        if (loc[axis] < 0 || loc[axis] > 7) {
            nb = T.brick_nbr[6 * b + 2 * axis + dir];
            if (nb < 0) return -1;
            loc[axis] = dir ? 0 : 7;
        }
        return static_cast<int64_t>(nb) * 512 + loc[0] + 8 * loc[1] + 64 * loc[2];
    }
    // every open connection of cell a: structured faces (area h^2) and compact edges
    template <class F>
    void for_each_conn(int64_t a, F&& fn) const {
        const double h = h_of(a);
        for (int axis = 0; axis < 3; ++axis)
            for (int dir = 0; dir < 2; ++dir) {
                const int64_t b = structured(a, axis, dir);
                if (b >= 0 && (T.mask[b] & CM_ACTIVE)) fn(b, h * h);
            }
        const int s = special_of[a];
        if (s < 0) return;
        for (int k = T.special_inc_off[s]; k < T.special_inc_off[s + 1]; ++k) {
            const int inc = T.special_inc[k];
            const int e = inc >= 0 ? inc : ~inc;
            const int64_t other = inc >= 0 ? T.edge_b[e] : T.edge_a[e];
            if (T.mask[other] & CM_ACTIVE) fn(other, T.edge_area[e]);
        }
    }
    double dist2(int64_t a, int64_t b) const {
        double s = 0;
        for (int ax = 0; ax < 3; ++ax) { const double d = cen(b, ax) - cen(a, ax); s += d * d; }
        return s;
    }
};

}  // namespace

Topology::NonOrth compute_nonorthogonal(const Topology& T) {
    if (T.instances() > 1) throw BuildError("non-orthogonal correction: single-instance topologies only");
    Topology::NonOrth N;
    const Geometry G(T);
    const double resolution = 16.0 * static_cast<double>(FLT_EPSILON);

    // one face record; returns false when the face needs no correction
    auto add_face = [&](int64_t lo, int64_t up, int axis, double area, double base, int edge) {
        const double h = G.h_of(lo);
        const int b = static_cast<int>(lo >> 9), q = static_cast<int>(lo & 511);
        const int I = T.brick_coord[3 * b + axis] * 8 + (axis == 0 ? (q & 7) : axis == 1 ? ((q >> 3) & 7) : (q >> 6));
        const double plane = T.lo[axis] + (I + 1) * h;  // the lower cell's +axis face
        double d[3];
        for (int ax = 0; ax < 3; ++ax) d[ax] = G.cen(up, ax) - G.cen(lo, ax);
        const double eps = 1e-10 * (edge >= 0 ? std::max(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), 1e-12) : h);
        const double ld = plane - G.cen(lo, axis), ud = G.cen(up, axis) - plane;
        if (!(d[axis] > 2 * eps && ld > eps && ud > eps))
            throw BuildError("non-orthogonal correction: a pressure face is not bracketed by its control-volume centroids (dofs " +
                             std::to_string(lo) + ", " + std::to_string(up) + ")");
        const double alpha = 1.0 / d[axis];
        double k[3];
        for (int ax = 0; ax < 3; ++ax) k[ax] = (ax == axis ? 1.0 : 0.0) - d[ax] * alpha;
        const double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (kk > 0 && kk <= resolution * resolution) {
            k[0] = k[1] = k[2] = 0.0;
            ++N.below_resolution;
        }
        const double delta = alpha - base;
        if (edge < 0 && std::fabs(delta) * h <= 1e-12 && k[0] * k[0] + k[1] * k[1] + k[2] * k[2] <= 1e-24) return;
        N.f_lower.push_back(static_cast<int>(lo));
        N.f_upper.push_back(static_cast<int>(up));
        N.f_edge.push_back(edge >= 0 ? edge : -1 - axis);
        N.f_area.push_back(area);
        N.f_delta.push_back(delta);
        N.f_k.insert(N.f_k.end(), k, k + 3);
        N.f_w.push_back(ld / (ld + ud));
        if (edge >= 0) ++N.compact_faces; else ++N.regular_faces;
    };
    // compact edges (coarse/fine and partially open faces): factor coef/area in the operator
    for (size_t e = 0; e < T.edge_a.size(); ++e) {
        if (!(T.mask[T.edge_a[e]] & CM_ACTIVE) || !(T.mask[T.edge_b[e]] & CM_ACTIVE)) continue;
        add_face(T.edge_a[e], T.edge_b[e], T.edge_axis[e], T.edge_area[e], T.edge_coef[e] / T.edge_area[e], static_cast<int>(e));
    }
    // structured faces next to a cell whose centroid is off its cell centre (special cells);
    // each face once, from its lower cell unless the lower cell is itself special
    for (size_t s = 0; s < T.special_dof.size(); ++s) {
        const int64_t a = T.special_dof[s];
        if (!(T.mask[a] & CM_ACTIVE)) continue;
        const double h = G.h_of(a);
        for (int axis = 0; axis < 3; ++axis) {
            const int64_t up = G.structured(a, axis, 1);
            if (up >= 0 && (T.mask[up] & CM_ACTIVE)) add_face(a, up, axis, h * h, 1.0 / h, -1);
            const int64_t dn = G.structured(a, axis, 0);
            if (dn >= 0 && (T.mask[dn] & CM_ACTIVE) && G.special_of[dn] < 0) add_face(dn, a, axis, h * h, 1.0 / h, -1);
        }
    }
    const int nf = static_cast<int>(N.f_lower.size());

    // gradient nodes: both cells of every face whose correction vector is not zero, and of
    // every structured record (amr_pressure.cpp:147-151)
    std::unordered_map<int, int> node_of;
    auto node = [&](int dof) {
        auto it = node_of.emplace(dof, static_cast<int>(N.g_dof.size()));
        if (it.second) N.g_dof.push_back(dof);
        return it.first->second;
    };
    N.f_lnode.assign(nf, -1);
    N.f_unode.assign(nf, -1);
    for (int f = 0; f < nf; ++f) {
        const double* k = &N.f_k[3 * f];
        if (N.f_edge[f] >= 0 && k[0] * k[0] + k[1] * k[1] + k[2] * k[2] <= 1e-24) continue;
        N.f_lnode[f] = node(N.f_lower[f]);
        N.f_unode[f] = node(N.f_upper[f]);
    }
    const int nn = static_cast<int>(N.g_dof.size());

    // samples: every open connection, weight area / |d|^2 (amr_pressure.cpp:112-118)
    std::vector<std::vector<Sample>> samples(nn);
    for (int n = 0; n < nn; ++n) {
        const int64_t c = N.g_dof[n];
        G.for_each_conn(c, [&](int64_t b, double area) {
            const double d2 = G.dist2(c, b);
            if (d2 > 1e-24 && area > 0) samples[n].push_back({static_cast<int>(b), area / d2});
        });
    }
    auto normal_rank = [&](int n, double* inv) {
        const int64_t c = N.g_dof[n];
        double A[9] = {};
        for (const Sample& s : samples[n]) {
            double d[3];
            for (int ax = 0; ax < 3; ++ax) d[ax] = G.cen(s.nbr, ax) - G.cen(c, ax);
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) A[3 * r + cc] += s.w * d[r] * d[cc];
        }
        return sym_pinv3(A, inv, 1e-4);
    };
    // widen a rank-deficient stencil by one ring of open connections: a neighbour's
    // neighbour joins with weight min(first-hop area, second-hop area) / |d|^2
    // (amr_pressure.cpp:125-141)
    N.g_ring.assign(nn, 1);
    auto extend = [&](const std::vector<int>& deficient) {
        std::unordered_map<int, std::vector<std::pair<int, double>>> via;  // middle -> (node, path area)
        for (int n : deficient) {
            const int64_t c = N.g_dof[n];
            for (const Sample& s : samples[n]) via[s.nbr].push_back({n, s.w * G.dist2(c, s.nbr)});
        }
        std::vector<std::vector<Sample>> add(nn);
        for (const auto& [middle, srcs] : via) {
            G.for_each_conn(middle, [&](int64_t other, double area) {
                for (const auto& [n, path] : srcs) {
                    const int64_t c = N.g_dof[n];
                    if (other == c) continue;
                    bool present = false;
                    for (const Sample& s : samples[n]) if (s.nbr == other) { present = true; break; }
                    if (present) continue;
                    const double d2 = G.dist2(c, other);
                    if (!(d2 > 1e-24)) continue;
                    const double w = std::min(path, area) / d2;
                    auto& cand = add[n];
                    auto it = std::find_if(cand.begin(), cand.end(), [&](const Sample& s) { return s.nbr == other; });
                    if (it == cand.end()) cand.push_back({static_cast<int>(other), w});
                    else it->w = std::max(it->w, w);
                }
            });
        }
        for (int n : deficient) {
            auto& cand = add[n];
            std::sort(cand.begin(), cand.end(), [](const Sample& a, const Sample& b) { return a.nbr < b.nbr; });
            samples[n].insert(samples[n].end(), cand.begin(), cand.end());
        }
    };
    std::vector<int> deficient;
    for (int n = 0; n < nn; ++n) if (normal_rank(n, nullptr) < 3) deficient.push_back(n);
    for (int ring = 2; ring <= 3 && !deficient.empty(); ++ring) {
        for (int n : deficient) N.g_ring[n] = static_cast<uint8_t>(ring);
        extend(deficient);
        std::vector<int> rest;
        for (int n : deficient) if (normal_rank(n, nullptr) < 3) rest.push_back(n);
        deficient.swap(rest);
    }
    deficient.clear();
    for (int n = 0; n < nn; ++n) if (normal_rank(n, nullptr) == 2) { deficient.push_back(n); N.g_ring[n] = 4; }
    if (!deficient.empty()) extend(deficient);

    // stencil weights gw_q = M^+ (w_q d_q) and each node's affine response R = sum gw_q d_q^T
    std::vector<double> R(9 * static_cast<size_t>(nn), 0.0);
    N.g_off.assign(1, 0);
    N.g_rank.assign(nn, 0);
    for (int n = 0; n < nn; ++n) {
        double inv[9];
        N.g_rank[n] = static_cast<uint8_t>(normal_rank(n, inv));
        const int64_t c = N.g_dof[n];
        for (const Sample& s : samples[n]) {
            double d[3], g[3] = {0, 0, 0};
            for (int ax = 0; ax < 3; ++ax) d[ax] = G.cen(s.nbr, ax) - G.cen(c, ax);
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) g[r] += inv[3 * r + cc] * s.w * d[cc];
            N.g_nbr.push_back(s.nbr);
            N.g_w.insert(N.g_w.end(), g, g + 3);
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) R[9 * n + 3 * r + cc] += g[r] * d[cc];
        }
        N.g_off.push_back(static_cast<int>(N.g_nbr.size()));
    }
    // every correction must be represented by its interpolated stencils:
    // ((1-w) R_L + w R_U)^T k == k (amr_pressure.cpp:161-186)
    for (int f = 0; f < nf; ++f) {
        const double* k = &N.f_k[3 * f];
        const double mag = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        if (mag <= 1e-12) continue;
        const int l = N.f_lnode[f], u = N.f_unode[f];
        if (l < 0 || u < 0) throw BuildError("non-orthogonal correction without gradient nodes at face " + std::to_string(f));
        const double w = N.f_w[f];
        double err2 = 0;
        for (int cc = 0; cc < 3; ++cc) {
            double rep = 0;
            for (int r = 0; r < 3; ++r) rep += ((1 - w) * R[9 * l + 3 * r + cc] + w * R[9 * u + 3 * r + cc]) * k[r];
            err2 += (rep - k[cc]) * (rep - k[cc]);
        }
        if (std::sqrt(err2) > 1e-8 + 1e-3 * mag)
            throw BuildError("unsupported non-orthogonal pressure correction at dofs " + std::to_string(N.f_lower[f]) + ", " +
                             std::to_string(N.f_upper[f]) + " (gradient stencil ranks " + std::to_string(N.g_rank[l]) +
                             ", " + std::to_string(N.g_rank[u]) + ")");
    }
    // per-dof incidence, dofs ascending, faces in order: deterministic gathers on the device
    std::vector<std::pair<int, int>> inc;
    inc.reserve(2 * static_cast<size_t>(nf));
    for (int f = 0; f < nf; ++f) {
        inc.push_back({N.f_lower[f], f});
        inc.push_back({N.f_upper[f], ~f});
    }
    std::stable_sort(inc.begin(), inc.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    N.inc_off.assign(1, 0);
    for (size_t i = 0; i < inc.size(); ++i) {
        if (i == 0 || inc[i].first != inc[i - 1].first) {
            if (i > 0) N.inc_off.push_back(static_cast<int>(i));
            N.inc_dof.push_back(inc[i].first);
        }
        N.inc_face.push_back(inc[i].second);
    }
    if (!inc.empty()) N.inc_off.push_back(static_cast<int>(inc.size()));
    N.built = true;
    return N;
}

void build_nonorthogonal(Topology& T) {
    if (!T.nonorth.built) T.nonorth = compute_nonorthogonal(T);
}

}  // namespace pcfd
