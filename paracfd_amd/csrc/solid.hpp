// Point-in-solid tests of the analytic bodies and the cut-cell sub-sampler, shared by the host
// builder and the device sampler.  Every function here uses only +, -, *, /, sqrt, comparisons
// and integer arithmetic, and both sides are built without floating-point contraction, so a cell
// sampled on the device is bit-identical to the same cell sampled on the host (IEEE-754 fixes
// each of those operations to the correctly rounded result).
//
// Sub-sampling of one cell (the reference's cut-cell classification, amr_eb.cpp:82-375, by
// sub-sample counting on the analytic body instead of triangle clipping): ns^3 volume samples at
// the cell's sub-cell centres and ns^2 samples on each of its six faces.  The fluid fraction and
// centroid, the wall estimate (fluid samples next to a solid sample) and each face's open fraction
// and open centroid are sums over those samples in one fixed order.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define PCFD_HD __host__ __device__
#else
#define PCFD_HD
#endif

namespace pcfd {

enum SolidKind : int { SOLID_NONE = 0, SOLID_SPHERE = 1, SOLID_NACA_WING = 2, SOLID_PARAGLIDER = 3 };

// Flat description of an analytic body; the section polygon is referenced by plain pointers
// (host vectors for the host builder, device copies for the device sampler).
struct SolidDesc {
    int kind = SOLID_NONE;
    double c[3] = {0, 0, 0}, r = 0;              // sphere
    double le[3] = {0, 0, 0}, ca = 1, sa = 0;    // wings: leading edge (pitch axis), cos / sin of the pitch
    double chord = 1, half_span = 0;             // straight wing
    double b = 0, c0 = 0, R = 0, tsw = 0;        // paraglider: span, root chord, arc radius, tan(sweep)
    int flat = 1;                                // paraglider without anhedral
    // closed section polygon in chord units with a bucketed vertical-ray crossing test: bucket k
    // of nb over [xmin, xmax] lists the segments bseg[boff[k] .. boff[k+1]) whose x-range meets it
    const double* px = nullptr;
    const double* pz = nullptr;
    int np = 0;
    double xmin = 0, xmax = 0;
    int nb = 1;
    const int* boff = nullptr;
    const int* bseg = nullptr;
};

PCFD_HD inline int poly_bucket(const SolidDesc& G, double x) {
    const int k = static_cast<int>((x - G.xmin) / (G.xmax - G.xmin) * G.nb);
    return k < 0 ? 0 : (k > G.nb - 1 ? G.nb - 1 : k);
}

PCFD_HD inline bool poly_inside(const SolidDesc& G, double x, double z) {
    if (x < G.xmin || x > G.xmax) return false;
    const int k = poly_bucket(G, x);
    int crossings = 0;
    for (int e = G.boff[k]; e < G.boff[k + 1]; ++e) {
        const int s = G.bseg[e], t = s + 1 == G.np ? 0 : s + 1;
        const double xs = G.px[s], xt = G.px[t];
        if ((xs <= x && x < xt) || (xt <= x && x < xs)) {
            const double zc = G.pz[s] + (G.pz[t] - G.pz[s]) * (x - xs) / (xt - xs);
            if (zc > z) ++crossings;
        }
    }
    return (crossings & 1) != 0;
}

// atan(v) for |v| < 0.2: the Taylor series to v^25 (the first omitted term is below 1e-20 v)
PCFD_HD inline double atan_series(double v) {
    const double w = v * v;
    double s = 1.0 / 25.0;
    for (int k = 11; k >= 0; --k) s = ((k & 1) ? -1.0 : 1.0) / (2 * k + 1) + w * s;
    return v * s;
}

// atan2 from +, -, *, /, sqrt only (so the host and the device agree bit for bit): reduction to
// t = min/max in [0, 1], t > tan(pi/8) mapped by atan t = pi/4 + atan((t - 1) / (t + 1)), one
// half-angle step atan u = 2 atan(u / (1 + sqrt(1 + u^2))) to |v| < 0.2, then atan_series.
PCFD_HD inline double exact_ops_atan2(double y, double x) {
    const double kPi = 3.14159265358979323846;
    const double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    if (ax == 0 && ay == 0) return 0;
    const bool swap = ay > ax;
    double t = swap ? ax / ay : ay / ax, base = 0;
    if (t > 0.41421356237309503) {
        t = (t - 1) / (t + 1);
        base = 0.25 * kPi;
    }
    const double v = t / (1 + std::sqrt(1 + t * t));
    double r = base + 2 * atan_series(v);
    if (swap) r = 0.5 * kPi - r;
    if (x < 0) r = kPi - r;
    return y < 0 ? -r : r;
}

// paraglider planform: elliptic chord along the span arc and the leading edge's x
PCFD_HD inline double glider_chord(const SolidDesc& G, double s) {
    const double e = 2.0 * s / G.b, q = 1.0 - e * e;
    return G.c0 * std::sqrt(q > 0.0 ? q : 0.0);
}
PCFD_HD inline double glider_x_le(const SolidDesc& G, double s) {
    return 0.25 * (G.c0 - glider_chord(G, s)) + (s < 0 ? -s : s) * G.tsw;
}

PCFD_HD inline bool solid_at(const SolidDesc& G, double x, double y, double z) {
    switch (G.kind) {
        case SOLID_SPHERE: {
            const double dx = x - G.c[0], dy = y - G.c[1], dz = z - G.c[2];
            return dx * dx + dy * dy + dz * dz < G.r * G.r;
        }
        case SOLID_NACA_WING: {
            const double dy = y - G.le[1];
            if ((dy < 0 ? -dy : dy) > G.half_span) return false;
            const double dx = x - G.le[0], dz = z - G.le[2];
            const double xl = (dx * G.ca - dz * G.sa) / G.chord, zl = (dx * G.sa + dz * G.ca) / G.chord;
            return poly_inside(G, xl, zl);
        }
        case SOLID_PARAGLIDER: {
            // invert the span arc: angle about the arc centre gives the span station s, radial
            // distance the offset normal to the section plane
            const double dx = x - G.le[0], dz = z - G.le[2], yw = y - G.le[1];
            const double xw = dx * G.ca - dz * G.sa, zw = dx * G.sa + dz * G.ca;
            double s, off;
            if (G.flat) {
                s = yw;
                off = zw;
            } else {
                const double zc = zw + G.R;
                s = G.R * exact_ops_atan2(yw, zc);
                off = std::sqrt(yw * yw + zc * zc) - G.R;
            }
            if ((s < 0 ? -s : s) > 0.49 * G.b) return false;
            const double c = glider_chord(G, s);
            if (!(c > 0)) return false;
            return poly_inside(G, (xw - glider_x_le(G, s)) / c, off / c);
        }
        default: return false;
    }
}

// Cut-cell data of one sampled cell
struct CutData {
    double f = 1;                       // fluid volume fraction
    double c[3] = {0, 0, 0};            // fluid centroid
    double ap[6] = {1, 1, 1, 1, 1, 1};  // face open fractions -x, +x, -y, +y, -z, +z
    double fc[6][2] = {};               // mean of the open face samples' tangential coordinates (axis+1, axis+2)
    double wall[3] = {0, 0, 0};         // mean of the fluid samples next to a solid sample
    int wall_n = 0;
};

// the most sub-samples per axis the device sampler stages (ns^3 + 6 ns^2 sample bits per cell)
constexpr int kMaxDeviceSubsamples = 16;

PCFD_HD constexpr int cell_samples(int ns) { return ns * ns * ns + 6 * ns * ns; }

// coordinate of sub-sample a of ns along one axis of cell I (size h) of a grid at lo
PCFD_HD inline double sample_coord(double lo, int I, int a, int ns, double h) { return lo + (I + (a + 0.5) / ns) * h; }

// Solid bit of sample s of cell I: s < ns^3 is the volume sample (a, b, c) at (c ns + b) ns + a;
// then face dir's sample (a, b) at ns^3 + (dir ns + b) ns + a, on the face plane, a along the
// tangential axis axis+1 and b along axis+2.
PCFD_HD inline bool sample_solid(const SolidDesc& G, const double lo[3], const int I[3], double h, int ns, int s) {
    const int nv = ns * ns * ns;
    double p[3];
    if (s < nv) {
        const int a = s % ns, b = (s / ns) % ns, c = s / (ns * ns);
        p[0] = sample_coord(lo[0], I[0], a, ns, h);
        p[1] = sample_coord(lo[1], I[1], b, ns, h);
        p[2] = sample_coord(lo[2], I[2], c, ns, h);
    } else {
        const int f = s - nv, dir = f / (ns * ns), r = f % (ns * ns), a = r % ns, b = r / ns;
        const int axis = dir >> 1, side = dir & 1, t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
        p[axis] = lo[axis] + (I[axis] + side) * h;
        p[t1] = sample_coord(lo[t1], I[t1], a, ns, h);
        p[t2] = sample_coord(lo[t2], I[t2], b, ns, h);
    }
    return solid_at(G, p[0], p[1], p[2]);
}

// Part `part` of a sampled cell's data from its sample bits: 0 the fluid fraction and centroid,
// 1 the wall estimate, 2 + dir face dir's open fraction and open centroid.  Each part writes
// only its own fields (all of them) and sums its samples in index order.
PCFD_HD inline void cut_part(const uint8_t* sol, const double lo[3], const int I[3], double h, int ns, int part,
                             CutData& cd) {
    const int nv = ns * ns * ns;
    if (part == 0) {
        int fluid = 0;
        double cx = 0, cy = 0, cz = 0;
        for (int c = 0; c < ns; ++c)
            for (int b = 0; b < ns; ++b)
                for (int a = 0; a < ns; ++a) {
                    if (sol[(c * ns + b) * ns + a]) continue;
                    ++fluid;
                    cx += sample_coord(lo[0], I[0], a, ns, h);
                    cy += sample_coord(lo[1], I[1], b, ns, h);
                    cz += sample_coord(lo[2], I[2], c, ns, h);
                }
        cd.f = static_cast<double>(fluid) / nv;
        cd.c[0] = fluid > 0 ? cx / fluid : 0.0;
        cd.c[1] = fluid > 0 ? cy / fluid : 0.0;
        cd.c[2] = fluid > 0 ? cz / fluid : 0.0;
    } else if (part == 1) {
        int n = 0;
        double w[3] = {0, 0, 0};
        for (int c = 0; c < ns; ++c)
            for (int b = 0; b < ns; ++b)
                for (int a = 0; a < ns; ++a) {
                    if (sol[(c * ns + b) * ns + a]) continue;
                    const bool touches = (a > 0 && sol[(c * ns + b) * ns + a - 1]) ||
                                         (a + 1 < ns && sol[(c * ns + b) * ns + a + 1]) ||
                                         (b > 0 && sol[(c * ns + b - 1) * ns + a]) ||
                                         (b + 1 < ns && sol[(c * ns + b + 1) * ns + a]) ||
                                         (c > 0 && sol[((c - 1) * ns + b) * ns + a]) ||
                                         (c + 1 < ns && sol[((c + 1) * ns + b) * ns + a]);
                    if (!touches) continue;
                    w[0] += sample_coord(lo[0], I[0], a, ns, h);
                    w[1] += sample_coord(lo[1], I[1], b, ns, h);
                    w[2] += sample_coord(lo[2], I[2], c, ns, h);
                    ++n;
                }
        cd.wall_n = n;
        for (int r = 0; r < 3; ++r) cd.wall[r] = n > 0 ? w[r] / n : 0.0;
    } else {
        const int dir = part - 2, axis = dir >> 1, t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
        int open = 0;
        double s1 = 0, s2 = 0;
        for (int b = 0; b < ns; ++b)
            for (int a = 0; a < ns; ++a) {
                if (sol[nv + (dir * ns + b) * ns + a]) continue;
                ++open;
                s1 += sample_coord(lo[t1], I[t1], a, ns, h);
                s2 += sample_coord(lo[t2], I[t2], b, ns, h);
            }
        cd.ap[dir] = static_cast<double>(open) / (ns * ns);
        cd.fc[dir][0] = open > 0 ? s1 / open : 0.0;
        cd.fc[dir][1] = open > 0 ? s2 / open : 0.0;
    }
}

}  // namespace pcfd
