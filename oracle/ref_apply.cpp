// Test infrastructure: a driver that evaluates the reference's own composite pressure
// operator (CompositeAmrPressureSystem::apply_cpu / diagonal_cpu, amr_pressure.cpp:498-543)
// on this project's topology, so the tests can check the oracle's operator against the
// reference code itself.  It is built only from the reference sources where they lie
// (oracle/Makefile, target ref) into oracle/_ref/, and only tests load it.
//
// The topology maps onto the reference's structures one to one: the stored bricks of each
// level become that level's BrickMetadata (same 8^3 brick numbering, dof = level_offset +
// brick * 512 + i + 8 j + 64 k, amr_pressure.cpp:213-216), a cell whose +axis face is not a
// structured face gets the cut bit of that axis, and every compact edge becomes an
// `embedded` connection with open area A and gradient factor coef / A.  pcfdref_operator gives
// no pressure-gradient nodes, so the non-orthogonal terms of apply_cpu vanish and it evaluates
// the orthogonal operator; pcfdref_full_operator adds the non-orthogonal records.
#include "core/fluid/amr_grid.h"
#include "core/fluid/amr_pressure.h"

#include <cstdint>
#include <exception>
#include <vector>

using namespace paracfd::core;

namespace {

struct Inputs {
    int nlevels;
    const double* level_h;      // nlevels
    const int* level_first;     // nlevels + 1: first stored brick of each level
    const int* brick_nbr;       // 6 per brick: -x, +x, -y, +y, -z, +z (global brick index or -1)
    const std::uint8_t* brick_flags;
    int outlet;
    const std::uint8_t* mask;   // per dof: +x/+y/+z structured bits 1/2/4, active 64
    int nedges;
    const int* ea;
    const int* eb;
    const double* earea;
    const double* ecoef;
};

void build(const Inputs& in, AmrHierarchy& H, CompositeAmrPressureSystem& S) {
    const int nb = in.level_first[in.nlevels];
    Aabb3d dom;
    dom.lo = {0.0, 0.0, 0.0};
    dom.hi = {8 * in.level_h[0], 8 * in.level_h[0], 8 * in.level_h[0]};
    H = AmrHierarchy::uniform(dom, in.level_h[0], 8, 1);
    auto& levels = H.levels();
    levels.assign(in.nlevels, AmrLevel{});
    S.level_offset.assign(in.nlevels, 0);
    for (int l = 0; l < in.nlevels; ++l) {
        AmrLevel& L = levels[l];
        L.level = l;
        L.h = in.level_h[l];
        const int b0 = in.level_first[l], b1 = in.level_first[l + 1];
        S.level_offset[l] = b0 * 512;
        L.bricks.resize(b1 - b0);
        for (int b = b0; b < b1; ++b) {
            BrickMetadata& m = L.bricks[b - b0];
            m.level = l;
            m.h = L.h;
            m.flags = in.brick_flags[b];
            for (int f = 0; f < 6; ++f) {
                const int n = in.brick_nbr[6 * static_cast<std::size_t>(b) + f];
                m.same_level_neighbor[f] = n >= 0 ? n - b0 : -1;
            }
        }
        L.active_bricks = L.bricks.size();
    }
    S.hierarchy = &H;
    S.brick_size = 8;
    S.storage_size = nb * 512;
    S.pressure_outlet_xmax = in.outlet != 0;
    S.active.assign(S.storage_size, 0);
    S.cut_face_mask.assign(S.storage_size, 0);
    for (int a = 0; a < S.storage_size; ++a) {
        const std::uint8_t m = in.mask[a];
        if (!(m & 64)) continue;
        S.active[a] = 1;
        for (int axis = 0; axis < 3; ++axis)
            if (!(m & (1u << axis))) S.cut_face_mask[a] |= static_cast<std::uint8_t>(1u << axis);
    }
    S.embedded.resize(in.nedges);
    for (int e = 0; e < in.nedges; ++e) {
        CoarseFinePressureConnection& c = S.embedded[e];
        c.coarse_dof = in.ea[e];
        c.fine_dof = in.eb[e];
        c.open_area = in.earea[e];
        c.orthogonal_gradient_factor = in.ecoef[e] / in.earea[e];
    }
}

// The non-orthogonal records of this project's topology (nf_* faces, ng_* gradient nodes)
// placed into the reference's structures: a compact face becomes the `embedded` connection of
// its edge with factor 1/d_axis, correction vector, weight and nodes; a structured face a
// RegularPressureCorrection; the stencils the pressure_gradient_* arrays.  apply_cpu then
// evaluates the full operator (amr_pressure.cpp:516-519).
struct NonOrthInputs {
    int nfaces;
    const int* lower;
    const int* upper;
    const int* edge;
    const double* area;
    const double* delta;
    const double* k;
    const double* w;
    const int* lnode;
    const int* unode;
    int nnodes;
    const int* gdof;
    const int* goff;
    const int* gnbr;
    const double* gw;
};

void add_nonorth(const NonOrthInputs& n, CompositeAmrPressureSystem& S) {
    for (int f = 0; f < n.nfaces; ++f) {
        const Vec3d k{n.k[3 * f], n.k[3 * f + 1], n.k[3 * f + 2]};
        if (n.edge[f] >= 0) {
            CoarseFinePressureConnection& c = S.embedded[n.edge[f]];
            c.coarse_dof = n.lower[f];
            c.fine_dof = n.upper[f];
            c.direction = 1;
            c.orthogonal_gradient_factor += n.delta[f];
            c.nonorthogonal_correction = k;
            c.upper_gradient_weight = n.w[f];
            c.lower_gradient_node = n.lnode[f];
            c.upper_gradient_node = n.unode[f];
        } else {
            RegularPressureCorrection r;
            r.lower_dof = n.lower[f];
            r.upper_dof = n.upper[f];
            r.axis = static_cast<std::int8_t>(-1 - n.edge[f]);
            r.open_area = n.area[f];
            r.two_point_delta = n.delta[f];
            r.nonorthogonal_correction = k;
            r.upper_gradient_weight = n.w[f];
            r.lower_gradient_node = n.lnode[f];
            r.upper_gradient_node = n.unode[f];
            S.regular_pressure_corrections.push_back(r);
        }
    }
    S.pressure_gradient_dof.assign(n.gdof, n.gdof + n.nnodes);
    S.pressure_gradient_offset.assign(n.goff, n.goff + n.nnodes + 1);
    const int entries = n.goff[n.nnodes];
    S.pressure_gradient_neighbour.assign(n.gnbr, n.gnbr + entries);
    S.pressure_gradient_weight.resize(entries);
    for (int q = 0; q < entries; ++q) S.pressure_gradient_weight[q] = Vec3d{n.gw[3 * q], n.gw[3 * q + 1], n.gw[3 * q + 2]};
}

}  // namespace

extern "C" __attribute__((visibility("default"))) int pcfdref_full_operator(
    int nlevels, const double* level_h, const int* level_first, const int* brick_nbr, const std::uint8_t* brick_flags,
    int outlet, const std::uint8_t* mask, int nedges, const int* ea, const int* eb, const double* earea,
    const double* ecoef, int nfaces, const int* flower, const int* fupper, const int* fedge, const double* farea,
    const double* fdelta, const double* fk, const double* fw, const int* flnode, const int* funode, int nnodes,
    const int* gdof, const int* goff, const int* gnbr, const double* gw, const double* p, double* ap) {
    try {
        const Inputs in{nlevels, level_h, level_first, brick_nbr, brick_flags, outlet, mask, nedges, ea, eb, earea, ecoef};
        const NonOrthInputs no{nfaces, flower, fupper, fedge, farea, fdelta, fk, fw, flnode, funode, nnodes, gdof, goff, gnbr, gw};
        AmrHierarchy H;
        CompositeAmrPressureSystem S;
        build(in, H, S);
        add_nonorth(no, S);
        std::vector<double> pv(p, p + S.storage_size), out;
        S.apply_cpu(pv, out);
        for (int a = 0; a < S.storage_size; ++a) ap[a] = out[a];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

extern "C" __attribute__((visibility("default"))) int pcfdref_operator(
    int nlevels, const double* level_h, const int* level_first, const int* brick_nbr, const std::uint8_t* brick_flags,
    int outlet, const std::uint8_t* mask, int nedges, const int* ea, const int* eb, const double* earea,
    const double* ecoef, const double* p, double* ap, double* diag) {
    try {
        const Inputs in{nlevels, level_h, level_first, brick_nbr, brick_flags, outlet, mask, nedges, ea, eb, earea, ecoef};
        AmrHierarchy H;
        CompositeAmrPressureSystem S;
        build(in, H, S);
        std::vector<double> pv(p, p + S.storage_size), out;
        S.apply_cpu(pv, out);
        for (int a = 0; a < S.storage_size; ++a) ap[a] = out[a];
        S.diagonal_cpu(out);
        for (int a = 0; a < S.storage_size; ++a) diag[a] = out[a];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
