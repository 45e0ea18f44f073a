// Host preprocessing: geometry, nested AMR bricks, cut cells, compact edges, surface
// patches and the geometric multigrid hierarchy.  See topology.hpp.
#include "topology.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <thread>

namespace pcfd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body) {
    unsigned hw = std::thread::hardware_concurrency();
    int threads = static_cast<int>(std::max(1u, std::min(hw ? hw : 1u, 16u)));
    if (n < 4096 || threads == 1) { body(0, n); return; }
    std::vector<std::thread> pool;
    std::atomic<int64_t> next{0};
    const int64_t chunk = std::max<int64_t>(1024, n / (threads * 8));
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&] {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                body(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}


// ------------------------------------------------------------------------------------------
// analytic solids
// ------------------------------------------------------------------------------------------
class Geometry {
public:
    Geometry() = default;
    Geometry(const Geometry&) = delete;
    Geometry& operator=(const Geometry&) = delete;
    virtual ~Geometry() = default;
    bool solid(double x, double y, double z) const { return solid_at(desc, x, y, z); }
    virtual void surface(double spacing, std::vector<double>& pts) const = 0;
    // the body as the shared point-in-solid test sees it (polygon pointers into this object)
    SolidDesc desc;
    double aabb_lo[3] = {0, 0, 0}, aabb_hi[3] = {-1, -1, -1};
    double ref_area = 0;
    bool empty() const { return aabb_hi[0] < aabb_lo[0]; }
};

class NoGeometry final : public Geometry {
public:
    void surface(double, std::vector<double>&) const override {}
};

class Sphere final : public Geometry {
public:
    Sphere(const double c[3], double r) : r_(r) {
        for (int a = 0; a < 3; ++a) { c_[a] = c[a]; aabb_lo[a] = c[a] - r; aabb_hi[a] = c[a] + r; }
        ref_area = kPi * r * r;
        desc.kind = SOLID_SPHERE;
        for (int a = 0; a < 3; ++a) desc.c[a] = c[a];
        desc.r = r;
    }
    void surface(double spacing, std::vector<double>& pts) const override {
        const int n = std::max(64, static_cast<int>(4 * kPi * r_ * r_ / (spacing * spacing)) + 1);
        const double golden = kPi * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < n; ++i) {
            const double z = 1.0 - 2.0 * (i + 0.5) / n, rr = std::sqrt(std::max(0.0, 1 - z * z));
            const double t = golden * i;
            pts.push_back(c_[0] + r_ * rr * std::cos(t));
            pts.push_back(c_[1] + r_ * rr * std::sin(t));
            pts.push_back(c_[2] + r_ * z);
        }
    }

private:
    double c_[3], r_;
};

// 2-D closed polygon with a bucketed vertical-ray crossing test (the test itself is
// poly_inside, shared with the device sampler).
class Polygon2 {
public:
    void set(std::vector<double> x, std::vector<double> z) {
        x_ = std::move(x); z_ = std::move(z);
        xmin_ = *std::min_element(x_.begin(), x_.end());
        xmax_ = *std::max_element(x_.begin(), x_.end());
        nb_ = 512;
        SolidDesc G;
        describe(G);
        std::vector<std::vector<int>> buckets(nb_);
        const int n = static_cast<int>(x_.size());
        for (int s = 0; s < n; ++s) {
            const int t = (s + 1) % n;
            const double lo = std::min(x_[s], x_[t]), hi = std::max(x_[s], x_[t]);
            const int b0 = poly_bucket(G, lo), b1 = poly_bucket(G, hi);
            for (int b = b0; b <= b1; ++b) buckets[b].push_back(s);
        }
        off_.assign(nb_ + 1, 0);
        seg_.clear();
        for (int b = 0; b < nb_; ++b) {
            seg_.insert(seg_.end(), buckets[b].begin(), buckets[b].end());
            off_[b + 1] = static_cast<int>(seg_.size());
        }
    }
    // the polygon fields of a solid description (pointers into this polygon)
    void describe(SolidDesc& G) const {
        G.px = x_.data(); G.pz = z_.data(); G.np = static_cast<int>(x_.size());
        G.xmin = xmin_; G.xmax = xmax_; G.nb = nb_;
        G.boff = off_.data(); G.bseg = seg_.data();
    }
    bool inside(double x, double z) const {
        SolidDesc G;
        describe(G);
        return poly_inside(G, x, z);
    }
    const std::vector<double>& xs() const { return x_; }
    const std::vector<double>& zs() const { return z_; }
    const std::vector<int>& bucket_off() const { return off_; }
    const std::vector<int>& bucket_seg() const { return seg_; }
    double xmin() const { return xmin_; }
    double xmax() const { return xmax_; }

private:
    std::vector<double> x_, z_;
    double xmin_ = 0, xmax_ = 0;
    int nb_ = 1;
    std::vector<int> off_ = {0, 0}, seg_;
};

// NACA 4-digit section in chord units: camber line and closed-trailing-edge thickness with
// the -0.1036 quartic coefficient (naca_step.cpp:85-92), thickness applied normal to the
// camber line.  Returns the closed polygon TE -> upper -> LE -> lower.
void naca4_polygon(int digits, int n, std::vector<double>& px, std::vector<double>& pz) {
    const double m = (digits / 1000) / 100.0, p = ((digits / 100) % 10) / 10.0, t = (digits % 100) / 100.0;
    std::vector<double> xu(n + 1), zu(n + 1), xl(n + 1), zl(n + 1);
    for (int i = 0; i <= n; ++i) {
        const double beta = kPi * i / n, x = 0.5 * (1 - std::cos(beta));
        const double yt = 5 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                                   0.1036 * x * x * x * x);
        double yc = 0, slope = 0;
        if (m > 0) {
            if (x < p) { yc = m / (p * p) * (2 * p * x - x * x); slope = 2 * m / (p * p) * (p - x); }
            else { yc = m / ((1 - p) * (1 - p)) * ((1 - 2 * p) + 2 * p * x - x * x); slope = 2 * m / ((1 - p) * (1 - p)) * (p - x); }
        }
        const double th = std::atan(slope);
        xu[i] = x - yt * std::sin(th); zu[i] = yc + yt * std::cos(th);
        xl[i] = x + yt * std::sin(th); zl[i] = yc - yt * std::cos(th);
    }
    px.clear(); pz.clear();
    for (int i = n; i >= 0; --i) { px.push_back(xu[i]); pz.push_back(zu[i]); }
    for (int i = 1; i < n; ++i) { px.push_back(xl[i]); pz.push_back(zl[i]); }
}

class NacaWing final : public Geometry {
public:
    NacaWing(const pcfd_geometry_desc& g) {
        std::vector<double> px, pz;
        naca4_polygon(g.naca_digits, 400, px, pz);
        poly_.set(px, pz);
        for (int a = 0; a < 3; ++a) le_[a] = g.center[a];
        chord_ = g.chord; half_span_ = 0.5 * g.span;
        const double al = g.alpha_deg * kPi / 180.0;
        ca_ = std::cos(al); sa_ = std::sin(al);
        aabb_lo[0] = aabb_lo[2] = 1e300; aabb_hi[0] = aabb_hi[2] = -1e300;
        for (size_t i = 0; i < px.size(); ++i) {
            double w[3]; to_world(px[i], pz[i], 0, w);
            aabb_lo[0] = std::min(aabb_lo[0], w[0]); aabb_hi[0] = std::max(aabb_hi[0], w[0]);
            aabb_lo[2] = std::min(aabb_lo[2], w[2]); aabb_hi[2] = std::max(aabb_hi[2], w[2]);
        }
        aabb_lo[1] = le_[1] - half_span_; aabb_hi[1] = le_[1] + half_span_;
        ref_area = g.chord * g.span;
        desc.kind = SOLID_NACA_WING;
        for (int a = 0; a < 3; ++a) desc.le[a] = le_[a];
        desc.ca = ca_; desc.sa = sa_; desc.chord = chord_; desc.half_span = half_span_;
        poly_.describe(desc);
    }
    void to_world(double xl, double zl, double y, double w[3]) const {
        w[0] = le_[0] + chord_ * (xl * ca_ + zl * sa_);
        w[1] = le_[1] + y;
        w[2] = le_[2] + chord_ * (-xl * sa_ + zl * ca_);
    }
    void surface(double spacing, std::vector<double>& pts) const override {
        // resample the section outline by arc length
        const auto& px = poly_.xs(); const auto& pz = poly_.zs();
        std::vector<double> sx, sz;
        const int n = static_cast<int>(px.size());
        const double ds = spacing / chord_;
        for (int s = 0; s < n; ++s) {
            const int t = (s + 1) % n;
            const double len = std::hypot(px[t] - px[s], pz[t] - pz[s]);
            const int k = std::max(1, static_cast<int>(std::ceil(len / ds)));
            for (int q = 0; q < k; ++q) {
                const double f = static_cast<double>(q) / k;
                sx.push_back(px[s] + f * (px[t] - px[s]));
                sz.push_back(pz[s] + f * (pz[t] - pz[s]));
            }
        }
        const int ns = std::max(2, static_cast<int>(std::ceil(2 * half_span_ / spacing)) + 1);
        for (int j = 0; j < ns; ++j) {
            const double y = -half_span_ + 2 * half_span_ * j / (ns - 1);
            for (size_t q = 0; q < sx.size(); ++q) {
                double w[3]; to_world(sx[q], sz[q], y, w);
                pts.insert(pts.end(), w, w + 3);
            }
        }
        // tip caps
        for (double xl = poly_.xmin(); xl <= poly_.xmax(); xl += ds)
            for (double zl = -0.3; zl <= 0.3; zl += ds)
                if (poly_.inside(xl, zl))
                    for (double y : {-half_span_, half_span_}) {
                        double w[3]; to_world(xl, zl, y, w);
                        pts.insert(pts.end(), w, w + 3);
                    }
    }

private:
    Polygon2 poly_;
    double le_[3], chord_, half_span_, ca_, sa_;
};

// Synthetic paraglider-like wing.  The span coordinate s in [-b/2, b/2] runs along a circular
// arc in the y-z plane that droops to the tips: each half of the arc turns through
// anhedral_deg (the tip's span tangent droops anhedral_deg below horizontal), so its radius is
// R = (b/2) / anhedral.  The chord is elliptic,
// c(s) = c0 sqrt(1 - (2s/b)^2), with root chord c0 (aspect ratio 5, c0 = 4 b / (5 pi), when the
// given chord is 0), truncated at |2s/b| = 0.98; the leading edge follows the ellipse's
// quarter-chord line plus
// |s| tan(sweep).  Sections are the NACA 4-digit section of naca_digits with the trailing 30 %
// of the camber line turned up by reflex * ((x - 0.7) / 0.3)^2 (reflex: upturn at the
// trailing edge in chords), set normal to the arc, then the whole wing is pitched by alpha
// about +y through center (the root leading edge).  The point-in-solid test inverts the arc
// exactly: radial distance and angle about the arc centre give (s, section normal offset).
class Paraglider final : public Geometry {
public:
    Paraglider(const pcfd_geometry_desc& g) {
        b_ = g.span;
        c0_ = g.chord > 0 ? g.chord : 4.0 * b_ / (5.0 * kPi);
        const double area = 0.25 * kPi * c0_ * b_;  // planform area along the arc
        const double half = g.anhedral_deg * kPi / 180.0;
        flat_ = !(half > 1e-6);
        R_ = flat_ ? 0.0 : 0.5 * b_ / half;
        tsw_ = std::tan(g.sweep_deg * kPi / 180.0);
        for (int a = 0; a < 3; ++a) le_[a] = g.center[a];
        const double al = g.alpha_deg * kPi / 180.0;
        ca_ = std::cos(al); sa_ = std::sin(al);
        // section: NACA mean line + reflex upturn, thickness normal to the camber line
        const int digits = g.naca_digits;
        const double m = (digits / 1000) / 100.0, p = ((digits / 100) % 10) / 10.0, t = (digits % 100) / 100.0;
        const double rf = g.reflex;
        const int n = 400;
        std::vector<double> xu(n + 1), zu(n + 1), xl(n + 1), zl(n + 1);
        for (int i = 0; i <= n; ++i) {
            const double beta = kPi * i / n, x = 0.5 * (1 - std::cos(beta));
            const double yt = 5 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                                       0.1036 * x * x * x * x);
            double yc = 0, slope = 0;
            if (m > 0) {
                if (x < p) { yc = m / (p * p) * (2 * p * x - x * x); slope = 2 * m / (p * p) * (p - x); }
                else { yc = m / ((1 - p) * (1 - p)) * ((1 - 2 * p) + 2 * p * x - x * x); slope = 2 * m / ((1 - p) * (1 - p)) * (p - x); }
            }
            if (x > 0.7) {
                const double u = (x - 0.7) / 0.3;
                yc += rf * u * u;
                slope += rf * 2 * u / 0.3;
            }
            const double th = std::atan(slope);
            xu[i] = x - yt * std::sin(th); zu[i] = yc + yt * std::cos(th);
            xl[i] = x + yt * std::sin(th); zl[i] = yc - yt * std::cos(th);
        }
        std::vector<double> px, pz;
        for (int i = n; i >= 0; --i) { px.push_back(xu[i]); pz.push_back(zu[i]); }
        for (int i = 1; i < n; ++i) { px.push_back(xl[i]); pz.push_back(zl[i]); }
        poly_.set(px, pz);
        zmin_ = *std::min_element(pz.begin(), pz.end());
        zmax_ = *std::max_element(pz.begin(), pz.end());
        ref_area = area;
        desc.kind = SOLID_PARAGLIDER;
        for (int a = 0; a < 3; ++a) desc.le[a] = le_[a];
        desc.ca = ca_; desc.sa = sa_;
        desc.b = b_; desc.c0 = c0_; desc.R = R_; desc.tsw = tsw_; desc.flat = flat_ ? 1 : 0;
        poly_.describe(desc);
        // bounding box from a dense sample of the surface
        std::vector<double> pts;
        surface(0.02 * c0_, pts);
        aabb_lo[0] = aabb_lo[1] = aabb_lo[2] = 1e300;
        aabb_hi[0] = aabb_hi[1] = aabb_hi[2] = -1e300;
        for (size_t i = 0; i < pts.size(); i += 3)
            for (int a = 0; a < 3; ++a) {
                aabb_lo[a] = std::min(aabb_lo[a], pts[i + a] - 0.02 * c0_);
                aabb_hi[a] = std::max(aabb_hi[a], pts[i + a] + 0.02 * c0_);
            }
    }
    double chord(double s) const { return glider_chord(desc, s); }
    double x_le(double s) const { return glider_x_le(desc, s); }
    // wing frame (before pitch) -> world
    void to_world(double xw, double yw, double zw, double w[3]) const {
        w[0] = le_[0] + xw * ca_ + zw * sa_;
        w[1] = le_[1] + yw;
        w[2] = le_[2] - xw * sa_ + zw * ca_;
    }
    // section point (xi along the chord, zeta normal, in chords) at span station s
    void section_point(double s, double xi, double zeta, double w[3]) const {
        const double c = chord(s);
        double y, z;
        if (flat_) { y = s; z = c * zeta; }
        else {
            const double th = s / R_, rad = R_ + c * zeta;
            y = rad * std::sin(th);
            z = rad * std::cos(th) - R_;
        }
        to_world(x_le(s) + c * xi, y, z, w);
    }
    void surface(double spacing, std::vector<double>& pts) const override {
        const auto& px = poly_.xs(); const auto& pz = poly_.zs();
        const int n = static_cast<int>(px.size());
        const double smax = 0.49 * b_;
        const int ns = std::max(2, static_cast<int>(std::ceil(2 * smax / spacing)) + 1);
        for (int j = 0; j < ns; ++j) {
            const double s = -smax + 2 * smax * j / (ns - 1), c = chord(s);
            const double ds = spacing / std::max(c, 1e-9);
            for (int e = 0; e < n; ++e) {
                const int f = (e + 1) % n;
                const double len = std::hypot(px[f] - px[e], pz[f] - pz[e]);
                const int k = std::max(1, static_cast<int>(std::ceil(len / ds)));
                for (int q = 0; q < k; ++q) {
                    const double u = static_cast<double>(q) / k;
                    double w[3];
                    section_point(s, px[e] + u * (px[f] - px[e]), pz[e] + u * (pz[f] - pz[e]), w);
                    pts.insert(pts.end(), w, w + 3);
                }
            }
        }
        // tip caps
        for (double sgn : {-1.0, 1.0}) {
            const double c = chord(sgn * smax), d = spacing / std::max(c, 1e-9);
            for (double xi = poly_.xmin(); xi <= poly_.xmax(); xi += d)
                for (double ze = zmin_; ze <= zmax_; ze += d)
                    if (poly_.inside(xi, ze)) {
                        double w[3];
                        section_point(sgn * smax, xi, ze, w);
                        pts.insert(pts.end(), w, w + 3);
                    }
        }
    }

private:
    Polygon2 poly_;
    double b_, c0_, R_, tsw_, ca_, sa_, zmin_, zmax_;
    double le_[3];
    bool flat_;
};

// Uniform hash grid of surface points for "is any point within r of this box" queries.
class PointGrid {
public:
    void build(const std::vector<double>& pts, double cell) {
        g_ = cell;
        const int64_t n = static_cast<int64_t>(pts.size() / 3);
        if (n == 0) { dim_[0] = dim_[1] = dim_[2] = 0; return; }
        for (int a = 0; a < 3; ++a) { lo_[a] = 1e300; double hi = -1e300;
            for (int64_t i = 0; i < n; ++i) { lo_[a] = std::min(lo_[a], pts[3 * i + a]); hi = std::max(hi, pts[3 * i + a]); }
            lo_[a] -= cell; dim_[a] = static_cast<int>((hi - lo_[a]) / cell) + 2; }
        const int64_t cells = static_cast<int64_t>(dim_[0]) * dim_[1] * dim_[2];
        off_.assign(cells + 1, 0);
        std::vector<int64_t> key(n);
        for (int64_t i = 0; i < n; ++i) {
            int c[3];
            for (int a = 0; a < 3; ++a) c[a] = std::clamp(static_cast<int>((pts[3 * i + a] - lo_[a]) / cell), 0, dim_[a] - 1);
            key[i] = (static_cast<int64_t>(c[2]) * dim_[1] + c[1]) * dim_[0] + c[0];
            ++off_[key[i] + 1];
        }
        for (int64_t c = 0; c < cells; ++c) off_[c + 1] += off_[c];
        pts_.resize(3 * n);
        std::vector<int64_t> fill(off_.begin(), off_.end() - 1);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t d = fill[key[i]]++;
            for (int a = 0; a < 3; ++a) pts_[3 * d + a] = pts[3 * i + a];
        }
    }
    bool near_box(const double blo[3], const double bhi[3], double r) const {
        if (dim_[0] == 0) return false;
        int c0[3], c1[3];
        for (int a = 0; a < 3; ++a) {
            c0[a] = static_cast<int>(std::floor((blo[a] - r - lo_[a]) / g_));
            c1[a] = static_cast<int>(std::floor((bhi[a] + r - lo_[a]) / g_));
            if (c1[a] < 0 || c0[a] >= dim_[a]) return false;
            c0[a] = std::max(c0[a], 0); c1[a] = std::min(c1[a], dim_[a] - 1);
        }
        const double r2 = r * r;
        for (int z = c0[2]; z <= c1[2]; ++z)
            for (int y = c0[1]; y <= c1[1]; ++y)
                for (int x = c0[0]; x <= c1[0]; ++x) {
                    const int64_t c = (static_cast<int64_t>(z) * dim_[1] + y) * dim_[0] + x;
                    for (int64_t q = off_[c]; q < off_[c + 1]; ++q) {
                        double d2 = 0;
                        for (int a = 0; a < 3; ++a) {
                            const double p = pts_[3 * q + a];
                            const double e = p < blo[a] ? blo[a] - p : (p > bhi[a] ? p - bhi[a] : 0.0);
                            d2 += e * e;
                        }
                        if (d2 <= r2) return true;
                    }
                }
        return false;
    }

private:
    double lo_[3] = {0, 0, 0}, g_ = 1;
    int dim_[3] = {0, 0, 0};
    std::vector<int64_t> off_;
    std::vector<double> pts_;
};

// ------------------------------------------------------------------------------------------
// builder
// ------------------------------------------------------------------------------------------
struct LevelGrid {
    int nb[3] = {0, 0, 0};
    double h = 0;
    std::vector<int> status;
    int64_t idx(int x, int y, int z) const { return (static_cast<int64_t>(z) * nb[1] + y) * nb[0] + x; }
    bool inside(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < nb[0] && y < nb[1] && z < nb[2]; }
};

// A leaf brick of the build (its cells are leaf * 512 + q in the flat per-cell arrays)
struct Leaf {
    int level;
    int coord[3];
};

class Builder {
public:
    Builder(const pcfd_geometry_desc& g, const pcfd_grid_desc& d, BuildAccelerator* accel)
        : gd_(g), dd_(d), accel_(accel) {}
    Topology run();

private:
    const pcfd_geometry_desc gd_;
    const pcfd_grid_desc dd_;
    BuildAccelerator* accel_;
    std::chrono::steady_clock::time_point t_start_, t_last_;
    // wall time of the stage that ends now (milliseconds since the previous stage ended)
    void stage_done(int stage) {
        const auto now = std::chrono::steady_clock::now();
        T_.build_ms[stage] += std::chrono::duration<double, std::milli>(now - t_last_).count();
        t_last_ = now;
    }
    std::unique_ptr<Geometry> geo_;
    Topology T_;
    int L_ = 1;
    std::vector<LevelGrid> grid_;
    std::vector<Leaf> leaves_;
    std::vector<CutData> cuts_;
    // per cell (leaf * 512 + q): 0 solid, 1 full fluid (far from the surface), 2 sampled; the
    // index into cuts_ (or -1); active fluid
    std::vector<uint8_t> state_;
    std::vector<int> cut_;
    std::vector<uint8_t> active_;
    std::vector<int> leaf_level_, leaf_coord_;  // leaves_ as flat tables (graph view)
    int64_t total_cells_[3];

    void setup_geometry();
    void setup_domain();
    void refine();
    void classify();
    // the cell graph over the current leaf order (host pointers; refreshed by sync_leaves)
    BuildView view_;
    void sync_leaves();
    // neighbours of (leaf, local) in direction dir (0..5): same level (1 entry), coarser (1), finer (4)
    int neighbors(int leaf, int local, int dir, Neighbor out[4], bool& boundary) const {
        const int n = graph_neighbors(view_, leaf, local, dir, out, boundary);
        if (n < 0) throw BuildError(graph_error_text(-n));
        return n;
    }
    bool is_active(int leaf, int local) const { return active_[static_cast<int64_t>(leaf) * kBC + local] != 0; }
    void components();
    void host_cell_passes(int nb, CellPasses& P) const;
    void finalize();
    void build_multigrid();
    void host_multigrid();
    void coarse_inverse();
};

void Builder::sync_leaves() {
    const size_t nl = leaves_.size();
    leaf_level_.resize(nl);
    leaf_coord_.resize(3 * nl);
    for (size_t lb = 0; lb < nl; ++lb) {
        leaf_level_[lb] = leaves_[lb].level;
        for (int a = 0; a < 3; ++a) leaf_coord_[3 * lb + a] = leaves_[lb].coord[a];
    }
    view_.levels = L_;
    for (int a = 0; a < 3; ++a) view_.lo[a] = T_.lo[a];
    for (int l = 0; l < L_; ++l) {
        view_.h[l] = grid_[l].h;
        for (int a = 0; a < 3; ++a) view_.nb[l][a] = grid_[l].nb[a];
        view_.status[l] = grid_[l].status.data();
    }
    view_.leaf_level = leaf_level_.data();
    view_.leaf_coord = leaf_coord_.data();
    view_.cut = cut_.data();
    view_.cuts = cuts_.data();
    view_.active = active_.data();
}

void Builder::setup_geometry() {
    switch (gd_.kind) {
        case PCFD_GEOM_NONE: geo_ = std::make_unique<NoGeometry>(); break;
        case PCFD_GEOM_SPHERE:
            if (!(gd_.radius > 0)) throw BuildError("sphere radius must be positive");
            geo_ = std::make_unique<Sphere>(gd_.center, gd_.radius); break;
        case PCFD_GEOM_NACA4_WING:
            if (!(gd_.chord > 0) || !(gd_.span > 0) || gd_.naca_digits <= 0 || gd_.naca_digits > 9999)
                throw BuildError("invalid NACA wing description");
            geo_ = std::make_unique<NacaWing>(gd_); break;
        case PCFD_GEOM_PARAGLIDER:
            if (!(gd_.span > 0) || !(gd_.chord >= 0) || gd_.naca_digits <= 0 || gd_.naca_digits > 9999 ||
                !(gd_.anhedral_deg >= 0 && gd_.anhedral_deg < 90) || !(std::abs(gd_.sweep_deg) < 60) ||
                !(std::abs(gd_.reflex) < 0.2))
                throw BuildError("invalid paraglider description");
            geo_ = std::make_unique<Paraglider>(gd_); break;
        default: throw BuildError("unsupported geometry kind");
    }
}

void Builder::setup_domain() {
    L_ = dd_.levels;
    if (L_ < 1 || L_ > 4) throw BuildError("levels must be in 1..4");
    if (dd_.base_cells_x < kBS) throw BuildError("base_cells_x must be at least one brick");
    for (int a = 0; a < 3; ++a)
        if (!(dd_.hi[a] > dd_.lo[a])) throw BuildError("empty domain");
    const double h0 = (dd_.hi[0] - dd_.lo[0]) / dd_.base_cells_x;
    T_.h0 = h0;
    T_.levels = L_;
    T_.outlet = dd_.outlet_xmax != 0;
    // x: pad downstream to whole bricks; y, z: pad symmetrically (amr_grid.cpp:66-96)
    int cells[3];
    cells[0] = (dd_.base_cells_x + kBS - 1) / kBS * kBS;
    T_.lo[0] = dd_.lo[0];
    for (int a = 1; a < 3; ++a) {
        const double len = dd_.hi[a] - dd_.lo[a];
        cells[a] = static_cast<int>(std::ceil(len / h0 - 1e-9));
        cells[a] = (cells[a] + kBS - 1) / kBS * kBS;
        const double centre = 0.5 * (dd_.lo[a] + dd_.hi[a]);
        T_.lo[a] = centre - 0.5 * cells[a] * h0;
    }
    for (int a = 0; a < 3; ++a) T_.base_bricks[a] = cells[a] / kBS;
    grid_.resize(L_);
    T_.level_h.resize(L_);
    for (int l = 0; l < L_; ++l) {
        for (int a = 0; a < 3; ++a) grid_[l].nb[a] = T_.base_bricks[a] << l;
        grid_[l].h = h0 / (1 << l);
        T_.level_h[l] = grid_[l].h;
        grid_[l].status.assign(static_cast<size_t>(grid_[l].nb[0]) * grid_[l].nb[1] * grid_[l].nb[2], ST_ABSENT);
    }
    T_.finest_h = grid_[L_ - 1].h;
}

void Builder::refine() {
    // desired refinement by distance band, then proper nesting from fine to coarse
    std::vector<std::vector<uint8_t>> R(L_);
    for (int l = 0; l + 1 < L_; ++l) {
        LevelGrid& G = grid_[l];
        R[l].assign(G.status.size(), 0);
        if (geo_->empty()) continue;
        const double band = dd_.refine_band[l];
        const double spacing = 0.5 * grid_[l + 1].h;
        std::vector<double> pts;
        geo_->surface(spacing, pts);
        PointGrid pg;
        pg.build(pts, std::max(band, 8 * G.h) * 0.5);
        const double bh = kBS * G.h;
        parallel_for(static_cast<int64_t>(G.status.size()), [&](int64_t b0, int64_t b1) {
            for (int64_t b = b0; b < b1; ++b) {
                const int x = static_cast<int>(b % G.nb[0]), y = static_cast<int>((b / G.nb[0]) % G.nb[1]),
                          z = static_cast<int>(b / (static_cast<int64_t>(G.nb[0]) * G.nb[1]));
                double lo[3] = {T_.lo[0] + x * bh, T_.lo[1] + y * bh, T_.lo[2] + z * bh};
                double hi[3] = {lo[0] + bh, lo[1] + bh, lo[2] + bh};
                R[l][b] = pg.near_box(lo, hi, band + spacing) ? 1 : 0;
            }
        });
    }
    // nesting: a refined level-l brick needs all 26 level-l neighbours to exist
    for (int l = L_ - 2; l >= 1; --l) {
        LevelGrid& G = grid_[l];
        LevelGrid& P = grid_[l - 1];
        for (int z = 0; z < G.nb[2]; ++z)
            for (int y = 0; y < G.nb[1]; ++y)
                for (int x = 0; x < G.nb[0]; ++x) {
                    if (!R[l][G.idx(x, y, z)]) continue;
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                const int nx = x + dx, ny = y + dy, nz = z + dz;
                                if (!G.inside(nx, ny, nz)) continue;
                                R[l - 1][P.idx(nx >> 1, ny >> 1, nz >> 1)] = 1;
                            }
                }
    }
    // existence top-down
    for (int l = 0; l < L_; ++l) {
        LevelGrid& G = grid_[l];
        for (int z = 0; z < G.nb[2]; ++z)
            for (int y = 0; y < G.nb[1]; ++y)
                for (int x = 0; x < G.nb[0]; ++x) {
                    bool exists = l == 0 || grid_[l - 1].status[grid_[l - 1].idx(x >> 1, y >> 1, z >> 1)] == ST_REFINED;
                    if (!exists) continue;
                    const int64_t b = G.idx(x, y, z);
                    if (l + 1 < L_ && R[l][b]) { G.status[b] = ST_REFINED; continue; }
                    G.status[b] = static_cast<int>(leaves_.size());
                    Leaf leaf;
                    leaf.level = l; leaf.coord[0] = x; leaf.coord[1] = y; leaf.coord[2] = z;
                    leaves_.push_back(std::move(leaf));
                }
    }
}

void Builder::classify() {
    const int ns = std::max(1, dd_.subsamples);
    std::vector<double> pts;
    PointGrid pg;
    double spacing = 0.5 * T_.finest_h;
    if (!geo_->empty()) {
        geo_->surface(spacing, pts);
        pg.build(pts, 2 * T_.finest_h);
    }
    // cells within `spacing` of a surface point are sub-sampled; the others are fluid or solid
    // by their centre (whole bricks at once when the brick is not near the surface)
    state_.assign(leaves_.size() * kBC, 1);
    cut_.assign(leaves_.size() * kBC, -1);
    parallel_for(static_cast<int64_t>(leaves_.size()), [&](int64_t b0, int64_t b1) {
        for (int64_t lb = b0; lb < b1; ++lb) {
            const Leaf& leaf = leaves_[lb];
            uint8_t* state = &state_[lb * kBC];
            const double h = grid_[leaf.level].h, bh = kBS * h;
            double blo[3], bhi[3];
            for (int a = 0; a < 3; ++a) { blo[a] = T_.lo[a] + leaf.coord[a] * bh; bhi[a] = blo[a] + bh; }
            if (!pg.near_box(blo, bhi, spacing)) {
                const bool s = geo_->solid(0.5 * (blo[0] + bhi[0]), 0.5 * (blo[1] + bhi[1]), 0.5 * (blo[2] + bhi[2]));
                if (s) std::fill(state, state + kBC, 0);
                continue;
            }
            for (int q = 0; q < kBC; ++q) {
                const int I[3] = {leaf.coord[0] * kBS + (q & 7), leaf.coord[1] * kBS + ((q >> 3) & 7),
                                  leaf.coord[2] * kBS + (q >> 6)};
                double clo[3], chi[3];
                for (int a = 0; a < 3; ++a) { clo[a] = T_.lo[a] + I[a] * h; chi[a] = clo[a] + h; }
                if (pg.near_box(clo, chi, spacing)) state[q] = 2;
                else if (geo_->solid(0.5 * (clo[0] + chi[0]), 0.5 * (clo[1] + chi[1]), 0.5 * (clo[2] + chi[2])))
                    state[q] = 0;
            }
        }
    });
    // the cut list in leaf, then cell order: integer position in the level's cell grid and level
    std::vector<int> cells;
    for (size_t lb = 0; lb < leaves_.size(); ++lb)
        for (int q = 0; q < kBC; ++q)
            if (state_[lb * kBC + q] == 2) {
                const Leaf& leaf = leaves_[lb];
                cut_[lb * kBC + q] = static_cast<int>(cells.size() / 4);
                cells.insert(cells.end(), {leaf.coord[0] * kBS + (q & 7), leaf.coord[1] * kBS + ((q >> 3) & 7),
                                           leaf.coord[2] * kBS + (q >> 6), leaf.level});
            }
    stage_done(BUILD_CANDIDATES);
    const int64_t ncut = static_cast<int64_t>(cells.size() / 4);
    cuts_.assign(ncut, CutData{});
    if (accel_) {
        accel_->sample(geo_->desc, T_.lo, T_.level_h, ns, cells, cuts_);
    } else {
        parallel_for(ncut, [&](int64_t c0, int64_t c1) {
            std::vector<uint8_t> sol(cell_samples(ns));
            for (int64_t i = c0; i < c1; ++i) {
                const int* I = &cells[4 * i];
                const double h = T_.level_h[I[3]];
                for (int s = 0; s < cell_samples(ns); ++s) sol[s] = sample_solid(geo_->desc, T_.lo, I, h, ns, s);
                for (int part = 0; part < 8; ++part) cut_part(sol.data(), T_.lo, I, h, ns, part, cuts_[i]);
            }
        });
    }
    stage_done(BUILD_SAMPLING);
    // small cut cells become solid (min_volume_fraction)
    const double fmin = dd_.min_volume_fraction;
    active_.assign(leaves_.size() * kBC, 0);
    for (size_t c = 0; c < active_.size(); ++c)
        active_[c] = state_[c] == 1 || (state_[c] == 2 && cuts_[cut_[c]].f >= fmin && cuts_[cut_[c]].f > 0);
    sync_leaves();
}

void Builder::components() {
    // keep only fluid connected to the outlet (or the largest component without an outlet)
    if (accel_) {
        accel_->connected_fluid(view_, static_cast<int64_t>(leaves_.size()), T_.outlet, active_);
        return;
    }
    const int64_t ncell = static_cast<int64_t>(leaves_.size()) * kBC;
    std::vector<int> comp(ncell, -1);
    std::vector<int64_t> sizes;
    std::vector<uint8_t> touches_outlet;
    std::vector<int64_t> stack;
    for (int64_t c = 0; c < ncell; ++c) {
        if (!active_[c] || comp[c] >= 0) continue;
        const int id = static_cast<int>(sizes.size());
        sizes.push_back(0);
        touches_outlet.push_back(0);
        stack.clear();
        stack.push_back(c);
        comp[c] = id;
        while (!stack.empty()) {
            const int64_t cc = stack.back();
            stack.pop_back();
            ++sizes[id];
            for (int dir = 0; dir < 6; ++dir) {
                Neighbor nb[4]; bool bnd;
                const int n = neighbors(static_cast<int>(cc / kBC), static_cast<int>(cc % kBC), dir, nb, bnd);
                if (bnd && dir == 1) touches_outlet[id] = 1;
                for (int k = 0; k < n; ++k) {
                    const int64_t o = static_cast<int64_t>(nb[k].leaf) * kBC + nb[k].local;
                    if (nb[k].ap <= 0 || !active_[o] || comp[o] >= 0) continue;
                    comp[o] = id;
                    stack.push_back(o);
                }
            }
        }
    }
    std::vector<uint8_t> keep(sizes.size(), 0);
    if (T_.outlet) {
        for (size_t c = 0; c < sizes.size(); ++c) keep[c] = touches_outlet[c];
    } else if (!sizes.empty()) {
        keep[std::max_element(sizes.begin(), sizes.end()) - sizes.begin()] = 1;
    }
    for (int64_t c = 0; c < ncell; ++c)
        if (active_[c] && !keep[comp[c]]) active_[c] = 0;
}

// finalize's per-cell passes on the host, in dof order
void Builder::host_cell_passes(int nb, CellPasses& P) const {
    const int64_t ndof = static_cast<int64_t>(nb) * kBC;
    P.volume.assign(ndof, 0.0);
    P.centroid.assign(3 * ndof, 0.0);
    P.embedded.assign(ndof, 0);
    for (int64_t a = 0; a < ndof; ++a) {
        CellGeometry g;
        const int err = cell_geometry(view_, static_cast<int>(a / kBC), static_cast<int>(a % kBC), g);
        if (err != GE_NONE) throw BuildError(graph_error_text(err));
        P.volume[a] = g.volume;
        for (int r = 0; r < 3; ++r) P.centroid[3 * a + r] = g.centroid[r];
        P.embedded[a] = g.embedded;
    }
    P.mask.assign(ndof, 0);
    P.has_compact.assign(ndof, 0);
    P.diffusion_sum.assign(ndof, 0.0);
    P.S.assign(3 * ndof, 0.0);
    P.wall_dof.clear();
    P.wall_moment.clear();
    P.edges.clear();
    for (int64_t a = 0; a < ndof; ++a) {
        if (!active_[a]) continue;
        CellFaces f;
        const int err = cell_faces(view_, P.centroid.data(), P.embedded.data(), static_cast<int>(a / kBC),
                                   static_cast<int>(a % kBC), f, [&](const EdgeRecord& e) { P.edges.push_back(e); });
        if (err != GE_NONE) throw BuildError(graph_error_text(err));
        P.mask[a] = f.mask;
        P.has_compact[a] = f.has_compact;
        P.diffusion_sum[a] = f.diffusion_sum;
        for (int r = 0; r < 3; ++r) P.S[3 * a + r] = f.S[r];
        const double h = grid_[leaves_[a / kBC].level].h;
        if (closure_norm(f.S) > 1e-12 * h * h) {
            P.wall_dof.push_back(a);
            P.wall_moment.insert(P.wall_moment.end(), f.fm, f.fm + 9);
        }
    }
}

void Builder::finalize() {
    // drop empty leaves; order kept leaves by level then Morton order
    const int nl = static_cast<int>(leaves_.size());
    std::vector<int> order;
    for (int lb = 0; lb < nl; ++lb) {
        bool any = false;
        for (int q = 0; q < kBC && !any; ++q) any = active_[static_cast<int64_t>(lb) * kBC + q];
        const Leaf& lf = leaves_[lb];
        if (!any) grid_[lf.level].status[grid_[lf.level].idx(lf.coord[0], lf.coord[1], lf.coord[2])] = ST_EMPTY;
        else order.push_back(lb);
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const Leaf& A = leaves_[a]; const Leaf& B = leaves_[b];
        if (A.level != B.level) return A.level < B.level;
        return morton(A.coord[0], A.coord[1], A.coord[2]) < morton(B.coord[0], B.coord[1], B.coord[2]);
    });
    std::vector<int> newid(nl, -1);
    for (size_t i = 0; i < order.size(); ++i) newid[order[i]] = static_cast<int>(i);
    const int nb = static_cast<int>(order.size());
    T_.nbricks = nb;
    T_.ndof = static_cast<int64_t>(nb) * kBC;
    T_.brick_level.resize(nb);
    T_.brick_coord.resize(3 * nb);
    T_.brick_flags.assign(nb, 0);
    T_.brick_nbr.assign(6 * nb, -1);
    T_.level_brick_begin.assign(L_ + 1, 0);
    for (int b = 0; b < nb; ++b) {
        const Leaf& lf = leaves_[order[b]];
        T_.brick_level[b] = lf.level;
        for (int a = 0; a < 3; ++a) T_.brick_coord[3 * b + a] = lf.coord[a];
        const LevelGrid& G = grid_[lf.level];
        uint8_t fl = 0;
        if (lf.coord[0] == 0) fl |= BF_XMIN;
        if (lf.coord[0] == G.nb[0] - 1) fl |= BF_XMAX;
        if (lf.coord[1] == 0) fl |= BF_YMIN;
        if (lf.coord[1] == G.nb[1] - 1) fl |= BF_YMAX;
        if (lf.coord[2] == 0) fl |= BF_ZMIN;
        if (lf.coord[2] == G.nb[2] - 1) fl |= BF_ZMAX;
        T_.brick_flags[b] = fl;
        for (int dir = 0; dir < 6; ++dir) {
            int c[3] = {lf.coord[0], lf.coord[1], lf.coord[2]};
            c[dir >> 1] += (dir & 1) ? 1 : -1;
            if (!G.inside(c[0], c[1], c[2])) continue;
            const int st = G.status[G.idx(c[0], c[1], c[2])];
            if (st >= 0) T_.brick_nbr[6 * b + dir] = newid[st];
        }
    }
    for (int l = 0; l <= L_; ++l)
        T_.level_brick_begin[l] = static_cast<int>(std::lower_bound(T_.brick_level.begin(), T_.brick_level.end(), l) - T_.brick_level.begin());
    // leaf status now holds the final brick index
    for (int l = 0; l < L_; ++l)
        for (auto& s : grid_[l].status) if (s >= 0) s = newid[s];
    {
        std::vector<Leaf> nl2(nb);
        std::vector<uint8_t> na(static_cast<size_t>(nb) * kBC);
        std::vector<int> nc(static_cast<size_t>(nb) * kBC);
        for (int b = 0; b < nb; ++b) {
            nl2[b] = leaves_[order[b]];
            const int64_t src = static_cast<int64_t>(order[b]) * kBC, dst = static_cast<int64_t>(b) * kBC;
            std::copy(active_.begin() + src, active_.begin() + src + kBC, na.begin() + dst);
            std::copy(cut_.begin() + src, cut_.begin() + src + kBC, nc.begin() + dst);
        }
        leaves_ = std::move(nl2);
        active_ = std::move(na);
        cut_ = std::move(nc);
        state_.clear();
        state_.shrink_to_fit();
    }
    sync_leaves();

    // volumes and centroids of every cell, the embedded nodes (every same-level face of an
    // embedded node is a compact edge, as the reference's EB momentum regular connections are,
    // amr_advection.cu:800-812, so that its transport, reconstruction and traction carry the
    // centroid corrections); then the masks, compact edges, closure area vectors and the first
    // moment of each cell's open faces about its centroid (the wall's moment follows from it)
    CellPasses P;
    if (accel_) accel_->cell_passes(view_, nb, P);
    else host_cell_passes(nb, P);
    T_.mask = std::move(P.mask);
    T_.centroid = std::move(P.centroid);
    const std::vector<double>& volume = P.volume;
    const std::vector<double>& S = P.S;
    const std::vector<uint8_t>& has_compact = P.has_compact;
    const std::vector<double>& diffusion_sum = P.diffusion_sum;
    const std::vector<int64_t>& wall_dof = P.wall_dof;
    const size_t ne_created = P.edges.size();
    T_.edge_a.resize(ne_created); T_.edge_b.resize(ne_created); T_.edge_axis.resize(ne_created);
    T_.edge_area.resize(ne_created); T_.edge_coef.resize(ne_created); T_.edge_dcoef.resize(ne_created);
    T_.edge_w.resize(ne_created); T_.edge_x.resize(3 * ne_created);
    for (size_t e = 0; e < ne_created; ++e) {
        const EdgeRecord& r = P.edges[e];
        T_.edge_a[e] = r.a; T_.edge_b[e] = r.b; T_.edge_axis[e] = static_cast<int8_t>(r.axis);
        T_.edge_area[e] = r.area; T_.edge_coef[e] = r.coef; T_.edge_dcoef[e] = r.dcoef; T_.edge_w[e] = r.w;
        for (int k = 0; k < 3; ++k) T_.edge_x[3 * e + k] = r.x[k];
    }
    // boundary-face diffusion does not exist; the outlet adds nothing either (zero flux)
    // special cells and patches
    int64_t n_active = 0;
    double max_rate = 0;
    size_t wall_next = 0;
    T_.brick_special_off.assign(nb + 1, 0);
    std::vector<int> special_of(T_.ndof, -1);
    for (int b = 0; b < nb; ++b) {
        T_.brick_special_off[b] = static_cast<int>(T_.special_dof.size());
        const double h = grid_[leaves_[b].level].h;
        for (int q = 0; q < kBC; ++q) {
            const int64_t a = static_cast<int64_t>(b) * kBC + q;
            if (!(T_.mask[a] & CM_ACTIVE)) continue;
            ++n_active;
            max_rate = std::max(max_rate, diffusion_sum[a] / volume[a]);
            const double smag = std::sqrt(S[3 * a] * S[3 * a] + S[3 * a + 1] * S[3 * a + 1] + S[3 * a + 2] * S[3 * a + 2]);
            const bool wall = smag > 1e-12 * h * h;
            const bool partial = volume[a] != h * h * h;
            if (!(wall || partial || has_compact[a])) continue;
            T_.mask[a] |= CM_SPECIAL;
            special_of[a] = static_cast<int>(T_.special_dof.size());
            T_.special_dof.push_back(static_cast<int>(a));
            T_.special_volume.push_back(volume[a]);
            if (wall) {
                T_.special_patch.push_back(static_cast<int>(T_.patch_dof.size()));
                T_.patch_dof.push_back(static_cast<int>(a));
                for (int ax = 0; ax < 3; ++ax) T_.patch_S.push_back(S[3 * a + ax]);
                // first moment of the wall about the centroid, D = int_W (x - c) (x) n dA, from the
                // divergence theorem over the control volume: int_dV (x - c) (x) n dA = V I, so
                // D = V I - sum over the open faces of (x_f - c) (x) (A n)_f, as the area vector
                // S = -sum (A n)_f closes the faces' zeroth moment
                if (wall_next >= wall_dof.size() || wall_dof[wall_next] != a) throw BuildError("internal: wall moment order");
                const double* fm = &P.wall_moment[9 * wall_next++];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) T_.patch_D.push_back((r == c ? volume[a] : 0.0) - fm[3 * r + c]);
                const int ci = cut_[a];
                if (ci >= 0 && cuts_[ci].wall_n > 0) {
                    for (int ax = 0; ax < 3; ++ax) T_.patch_x.push_back(cuts_[ci].wall[ax]);
                } else {
                    for (int ax = 0; ax < 3; ++ax) T_.patch_x.push_back(T_.centroid[3 * a + ax] + 0.5 * h * S[3 * a + ax] / smag);
                }
            } else {
                T_.special_patch.push_back(-1);
            }
        }
    }
    T_.brick_special_off[nb] = static_cast<int>(T_.special_dof.size());
    T_.n_active = n_active;
    T_.max_diffusion_coefficient = max_rate;
    // incidence CSR of special cells, edges in ascending order
    const int nsp = static_cast<int>(T_.special_dof.size());
    T_.special_inc_off.assign(nsp + 1, 0);
    const int ne = static_cast<int>(T_.edge_a.size());
    for (int e = 0; e < ne; ++e) {
        ++T_.special_inc_off[special_of[T_.edge_a[e]] + 1];
        ++T_.special_inc_off[special_of[T_.edge_b[e]] + 1];
    }
    for (int s = 0; s < nsp; ++s) T_.special_inc_off[s + 1] += T_.special_inc_off[s];
    T_.special_inc.assign(T_.special_inc_off[nsp], 0);
    {
        std::vector<int> fill(T_.special_inc_off.begin(), T_.special_inc_off.end() - 1);
        for (int e = 0; e < ne; ++e) {
            T_.special_inc[fill[special_of[T_.edge_a[e]]]++] = e;
            T_.special_inc[fill[special_of[T_.edge_b[e]]]++] = ~e;
        }
    }
    // momentum redistribution groups (amr_advection.cu:1108-1178): an active cut cell whose fluid
    // fraction is below merge_volume_fraction, or with an axis whose open area (both faces) lies
    // in (0, merge_volume_fraction h^2), is united with the neighbour across its largest compact
    // edge (the first in edge order, as the reference's largest donor); the groups are the
    // components of these unions, ordered by their lowest cell
    T_.redist_group.assign(nsp, -1);
    T_.redist_off.assign(1, 0);
    T_.redist_member.clear();
    T_.redist_volume.clear();
    if (dd_.merge_volume_fraction > 0) {
        const double fm = dd_.merge_volume_fraction;
        std::vector<int> parent(nsp);
        std::iota(parent.begin(), parent.end(), 0);
        auto root = [&](int s) {
            while (parent[s] != s) s = parent[s] = parent[parent[s]];
            return s;
        };
        std::vector<uint8_t> seed(nsp, 0);
        for (int s = 0; s < nsp; ++s) {
            const int64_t a = T_.special_dof[s];
            const double h = grid_[leaves_[a / kBC].level].h;
            if (!(volume[a] < h * h * h)) continue;
            double dir[3] = {0, 0, 0}, best = 0;
            int donor = -1;
            for (int k = T_.special_inc_off[s]; k < T_.special_inc_off[s + 1]; ++k) {
                const int raw = T_.special_inc[k], e = raw >= 0 ? raw : ~raw;
                dir[T_.edge_axis[e]] += T_.edge_area[e];
                if (T_.edge_area[e] > best) {
                    best = T_.edge_area[e];
                    donor = special_of[raw >= 0 ? T_.edge_b[e] : T_.edge_a[e]];
                }
            }
            bool weak = volume[a] < fm * h * h * h;
            for (int ax = 0; ax < 3; ++ax) weak = weak || (dir[ax] > 0 && dir[ax] < fm * h * h);
            if (!weak || donor < 0) continue;
            int ra = root(s), rb = root(donor);
            if (ra != rb) {
                if (ra > rb) std::swap(ra, rb);
                parent[rb] = ra;
            }
            seed[s] = 1;
        }
        std::vector<uint8_t> grouped(nsp, 0);
        for (int s = 0; s < nsp; ++s) if (seed[s]) grouped[root(s)] = 1;
        std::vector<int> group_of_root(nsp, -1);
        std::vector<std::vector<int>> members;
        for (int s = 0; s < nsp; ++s) {
            const int r = root(s);
            if (!grouped[r]) continue;
            if (group_of_root[r] < 0) {
                group_of_root[r] = static_cast<int>(members.size());
                members.emplace_back();
            }
            T_.redist_group[s] = group_of_root[r];
            members[group_of_root[r]].push_back(s);
        }
        for (const auto& g : members) {
            double v = 0;
            for (int s : g) v += volume[T_.special_dof[s]];
            T_.redist_member.insert(T_.redist_member.end(), g.begin(), g.end());
            T_.redist_off.push_back(static_cast<int>(T_.redist_member.size()));
            T_.redist_volume.push_back(v);
        }
    }
    // solid volume of the body: every leaf cell's non-fluid part that lies inside the geometry
    double body = 0;
    for (int l = 0; l < L_; ++l) {
        const LevelGrid& G = grid_[l];
        const double h3 = G.h * G.h * G.h;
        for (size_t s = 0; s < G.status.size(); ++s) {
            const int st = G.status[s];
            if (st == ST_EMPTY) body += kBC * h3;
        }
    }
    for (int b = 0; b < nb; ++b) {
        const double h3 = std::pow(grid_[leaves_[b].level].h, 3);
        for (int q = 0; q < kBC; ++q) {
            const int64_t a = static_cast<int64_t>(b) * kBC + q;
            body += h3 - volume[a];
        }
    }
    T_.body_volume = body;
    T_.reference_area = geo_->ref_area;
}

// Geometric multigrid: at every level each cell is merged with its 2x2x2 siblings (cells of
// every AMR level double in size together, so neighbours stay within 2:1); coarse couplings
// are re-discretised as open area / centre distance.  Transfers are piecewise constant
// (R = P^T).
void Builder::build_multigrid() {
    T_.mg.clear();
    if (!accel_ || !accel_->multigrid(T_, L_, kCoarseMax, T_.mg)) host_multigrid();
    coarse_inverse();
}

void Builder::host_multigrid() {
    MGLevel L0;
    L0.n = static_cast<int>(T_.ndof);
    L0.exponent.assign(T_.ndof, 0);
    // per-cell integer coordinates at its own exponent
    std::vector<int> coord(3 * T_.ndof, 0);
    for (int b = 0; b < T_.nbricks; ++b) {
        const int l = T_.brick_level[b];
        for (int q = 0; q < kBC; ++q) {
            const int64_t a = static_cast<int64_t>(b) * kBC + q;
            L0.exponent[a] = L_ - 1 - l;
            coord[3 * a] = T_.brick_coord[3 * b] * kBS + (q & 7);
            coord[3 * a + 1] = T_.brick_coord[3 * b + 1] * kBS + ((q >> 3) & 7);
            coord[3 * a + 2] = T_.brick_coord[3 * b + 2] * kBS + (q >> 6);
        }
    }
    // level-0 faces as (a, b, axis, area) records, outlet areas per cell
    struct Face { int a, b; int axis; double area; };
    std::vector<Face> faces;
    std::vector<double> outlet(T_.ndof, 0.0);
    for (int b = 0; b < T_.nbricks; ++b) {
        const double h = T_.level_h[T_.brick_level[b]];
        for (int q = 0; q < kBC; ++q) {
            const int64_t a = static_cast<int64_t>(b) * kBC + q;
            const uint8_t m = T_.mask[a];
            if (!(m & CM_ACTIVE)) continue;
            for (int axis = 0; axis < 3; ++axis) {
                if (!(m & (CM_PX << axis))) continue;
                int c[3] = {q & 7, (q >> 3) & 7, q >> 6};
                int nbk = b;
                if (++c[axis] == kBS) { c[axis] = 0; nbk = T_.brick_nbr[6 * b + 2 * axis + 1]; }
                faces.push_back({static_cast<int>(a), nbk * kBC + c[0] + 8 * c[1] + 64 * c[2], axis, h * h});
            }
            if (T_.outlet && (T_.brick_flags[b] & BF_XMAX) && (q & 7) == 7) outlet[a] = h * h;
        }
    }
    for (size_t e = 0; e < T_.edge_a.size(); ++e)
        faces.push_back({T_.edge_a[e], T_.edge_b[e], T_.edge_axis[e], T_.edge_area[e]});
    std::vector<uint8_t> act(T_.ndof);
    for (int64_t a = 0; a < T_.ndof; ++a) act[a] = (T_.mask[a] & CM_ACTIVE) ? 1 : 0;

    T_.mg.push_back(std::move(L0));
    const double hf = T_.finest_h;
    const int coarse_max = kCoarseMax;
    int m = 0;
    std::vector<int> cur_exp = T_.mg[0].exponent;
    std::vector<int> cur_coord = coord;
    std::vector<uint8_t> cur_act = act;
    std::vector<double> cur_outlet = outlet;
    int64_t cur_n = T_.ndof;
    while (true) {
        // count active cells at the current level
        int64_t nact = 0;
        for (int64_t a = 0; a < cur_n; ++a) nact += cur_act[a];
        if (m >= 1 && (nact <= coarse_max || m >= 10)) break;
        const int mm = m + 1;
        // Parent keys: every cell proposes its 2x2x2 parent; a proposal is blocked when a
        // smaller cell lies inside the parent's footprint (a finer AMR region that has not
        // caught up yet), and the cell is then carried over unchanged.  Blocked parents never
        // overlap an existing cell, so the coarse cells tile the fluid without overlap.
        auto key_of = mg_key;
        int max_e = 0;
        for (int64_t a = 0; a < cur_n; ++a) if (cur_act[a]) max_e = std::max(max_e, cur_exp[a]);
        std::vector<uint64_t> blocked;
        for (int64_t a = 0; a < cur_n; ++a) {
            if (!cur_act[a]) continue;
            const int e = cur_exp[a];
            for (int E = e + 2; E <= max_e + 1; ++E) {
                const int s = E - e;
                blocked.push_back(key_of(E, cur_coord[3 * a] >> s, cur_coord[3 * a + 1] >> s, cur_coord[3 * a + 2] >> s));
            }
        }
        std::sort(blocked.begin(), blocked.end());
        blocked.erase(std::unique(blocked.begin(), blocked.end()), blocked.end());
        std::vector<std::pair<uint64_t, int64_t>> keys;
        std::vector<uint8_t> grows(cur_n, 0);
        keys.reserve(nact);
        for (int64_t a = 0; a < cur_n; ++a) {
            if (!cur_act[a]) continue;
            const int e = cur_exp[a];
            uint64_t key = key_of(e + 1, cur_coord[3 * a] >> 1, cur_coord[3 * a + 1] >> 1, cur_coord[3 * a + 2] >> 1);
            if (std::binary_search(blocked.begin(), blocked.end(), key)) {
                key = key_of(e, cur_coord[3 * a], cur_coord[3 * a + 1], cur_coord[3 * a + 2]);
            } else {
                grows[a] = 1;
            }
            keys.push_back({key, a});
        }
        std::sort(keys.begin(), keys.end());
        MGLevel C;
        std::vector<int>& parent = T_.mg[m].parent;
        parent.assign(cur_n, -1);
        std::vector<int> nexp, ncoord;
        int nc = -1;
        uint64_t last = ~0ULL;
        for (auto& [key, a] : keys) {
            if (key != last) {
                ++nc; last = key;
                const int s = grows[a] ? 1 : 0;
                nexp.push_back(cur_exp[a] + s);
                ncoord.push_back(cur_coord[3 * a] >> s);
                ncoord.push_back(cur_coord[3 * a + 1] >> s);
                ncoord.push_back(cur_coord[3 * a + 2] >> s);
            }
            parent[a] = nc;
        }
        C.n = nc + 1;
        if (C.n >= nact) {  // nothing merged: stop at the current level
            parent.assign(cur_n, -1);
            break;
        }
        C.exponent = nexp;
        // children CSR
        C.child_off.assign(C.n + 1, 0);
        for (int64_t a = 0; a < cur_n; ++a) if (parent[a] >= 0) ++C.child_off[parent[a] + 1];
        for (int c = 0; c < C.n; ++c) C.child_off[c + 1] += C.child_off[c];
        C.child.assign(C.child_off[C.n], 0);
        {
            std::vector<int> fill(C.child_off.begin(), C.child_off.end() - 1);
            for (int64_t a = 0; a < cur_n; ++a) if (parent[a] >= 0) C.child[fill[parent[a]]++] = static_cast<int>(a);
        }
        // coarse faces: each face has a unique owner slot (finer side, or the lower side for equal sizes)
        std::vector<int> slot_nbr(6 * static_cast<size_t>(C.n), -1);
        std::vector<double> slot_area(6 * static_cast<size_t>(C.n), 0.0);
        for (const Face& f : faces) {
            const int pa = parent[f.a], pb = parent[f.b];
            if (pa < 0 || pb < 0 || pa == pb) continue;
            int owner, dir, other;
            if (nexp[pa] <= nexp[pb]) { owner = pa; dir = 2 * f.axis + 1; other = pb; }
            else { owner = pb; dir = 2 * f.axis; other = pa; }
            if (nexp[pa] == nexp[pb]) { owner = pa; dir = 2 * f.axis + 1; other = pb; }
            const size_t s = 6 * static_cast<size_t>(owner) + dir;
            if (slot_nbr[s] >= 0 && slot_nbr[s] != other) throw BuildError("internal: inconsistent coarse face");
            slot_nbr[s] = other;
            slot_area[s] += f.area;
        }
        std::vector<double> coutlet(C.n, 0.0);
        for (int64_t a = 0; a < cur_n; ++a) if (parent[a] >= 0) coutlet[parent[a]] += cur_outlet[a];
        // CSR (symmetric)
        std::vector<int> deg(C.n, 0);
        for (int c = 0; c < C.n; ++c)
            for (int d = 0; d < 6; ++d)
                if (slot_nbr[6 * c + d] >= 0) { ++deg[c]; ++deg[slot_nbr[6 * c + d]]; }
        C.row_off.assign(C.n + 1, 0);
        for (int c = 0; c < C.n; ++c) C.row_off[c + 1] = C.row_off[c] + deg[c];
        C.col.assign(C.row_off[C.n], 0);
        C.coef.assign(C.row_off[C.n], 0.0);
        C.diag.assign(C.n, 0.0);
        {
            std::vector<int> fill(C.row_off.begin(), C.row_off.end() - 1);
            for (int c = 0; c < C.n; ++c)
                for (int d = 0; d < 6; ++d) {
                    const int o = slot_nbr[6 * c + d];
                    if (o < 0) continue;
                    const double sc = hf * std::ldexp(1.0, nexp[c]), so = hf * std::ldexp(1.0, nexp[o]);
                    const double k = slot_area[6 * c + d] / (0.5 * (sc + so));
                    C.col[fill[c]] = o; C.coef[fill[c]++] = k;
                    C.col[fill[o]] = c; C.coef[fill[o]++] = k;
                    C.diag[c] += k; C.diag[o] += k;
                }
            // deterministic row order: sort each row by column
            for (int c = 0; c < C.n; ++c) {
                std::vector<std::pair<int, double>> row;
                for (int q = C.row_off[c]; q < C.row_off[c + 1]; ++q) row.push_back({C.col[q], C.coef[q]});
                std::sort(row.begin(), row.end());
                for (size_t q = 0; q < row.size(); ++q) { C.col[C.row_off[c] + q] = row[q].first; C.coef[C.row_off[c] + q] = row[q].second; }
            }
        }
        for (int c = 0; c < C.n; ++c) {
            const double sc = hf * std::ldexp(1.0, nexp[c]);
            C.diag[c] += coutlet[c] / (0.5 * sc);
        }
        // faces of the new level for the next coarsening
        std::vector<Face> nfaces;
        nfaces.reserve(C.row_off[C.n] / 2);
        for (int c = 0; c < C.n; ++c)
            for (int d = 0; d < 6; ++d) {
                const int o = slot_nbr[6 * c + d];
                if (o < 0) continue;
                if (d & 1) nfaces.push_back({c, o, d >> 1, slot_area[6 * c + d]});
                else nfaces.push_back({o, c, d >> 1, slot_area[6 * c + d]});
            }
        faces = std::move(nfaces);
        cur_exp = nexp;
        cur_coord = ncoord;
        cur_act.assign(C.n, 1);
        cur_outlet = coutlet;
        cur_n = C.n;
        T_.mg.push_back(std::move(C));
        m = mm;
    }
    T_.mg.back().parent.assign(T_.mg.back().n, -1);
}

// dense inverse of the coarsest operator (Cholesky); skipped when singular
void Builder::coarse_inverse() {
    MGLevel& K = T_.mg.back();
    if (T_.mg.size() > 1 && K.n <= 4096) {
        const int n = K.n;
        std::vector<double> A(static_cast<size_t>(n) * n, 0.0);
        for (int r = 0; r < n; ++r) {
            A[static_cast<size_t>(r) * n + r] = K.diag[r];
            for (int q = K.row_off[r]; q < K.row_off[r + 1]; ++q) A[static_cast<size_t>(r) * n + K.col[q]] -= K.coef[q];
        }
        const unsigned hw = std::thread::hardware_concurrency();
        const int nt = std::max(1, std::min(static_cast<int>(hw ? hw : 1u), std::min(16, n / 64)));
        auto on_threads = [&](int count, const std::function<void(int, int)>& body) {
            if (nt == 1 || count < 2 * nt) { body(0, count); return; }
            std::vector<std::thread> pool;
            for (int t = 0; t < nt; ++t) pool.emplace_back(body, count * t / nt, count * (t + 1) / nt);
            for (auto& th : pool) th.join();
        };
        auto at = [&](int r, int c) -> double& { return A[static_cast<size_t>(r) * n + c]; };
        // Cholesky A = L L^T in place (lower), by blocks of 64 columns: each entry is
        // A_ij - sum_{k<j} L_ik L_jk (k ascending) divided by L_jj, as in the unblocked order;
        // the rows below a block are independent and run on host threads
        bool ok = true;
        constexpr int kNB = 64;
        for (int j0 = 0; j0 < n && ok; j0 += kNB) {
            const int j1 = std::min(n, j0 + kNB);
            for (int j = j0; j < j1 && ok; ++j) {  // the diagonal block
                double d = at(j, j);
                for (int k = 0; k < j; ++k) d -= at(j, k) * at(j, k);
                if (!(d > 1e-14 * K.diag[j])) { ok = false; break; }
                d = std::sqrt(d);
                at(j, j) = d;
                for (int i = j + 1; i < j1; ++i) {
                    double v = at(i, j);
                    for (int k = 0; k < j; ++k) v -= at(i, k) * at(j, k);
                    at(i, j) = v / d;
                }
            }
            if (!ok || j1 == n) break;
            on_threads(n - j1, [&](int r0, int r1) {
                for (int i = j1 + r0; i < j1 + r1; ++i)
                    for (int j = j0; j < j1; ++j) {
                        double v = at(i, j);
                        for (int k = 0; k < j; ++k) v -= at(i, k) * at(j, k);
                        at(i, j) = v / at(j, j);
                    }
            });
        }
        if (ok) {
            // inverse column by column (columns are independent: host threads); the backward
            // substitution reads L^T row by row
            std::vector<double> U(static_cast<size_t>(n) * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int k = 0; k <= i; ++k) U[static_cast<size_t>(k) * n + i] = at(i, k);
            K.coarse_inv.assign(static_cast<size_t>(n) * n, 0.0);
            on_threads(n, [&](int c0, int c1) {
                std::vector<double> y(n);
                for (int c = c0; c < c1; ++c) {
                    for (int i = 0; i < n; ++i) {
                        double v = i == c ? 1.0 : 0.0;
                        for (int k = 0; k < i; ++k) v -= A[static_cast<size_t>(i) * n + k] * y[k];
                        y[i] = v / A[static_cast<size_t>(i) * n + i];
                    }
                    for (int i = n - 1; i >= 0; --i) {
                        double v = y[i];
                        for (int k = i + 1; k < n; ++k) v -= U[static_cast<size_t>(i) * n + k] * y[k];
                        y[i] = v / U[static_cast<size_t>(i) * n + i];
                    }
                    for (int i = 0; i < n; ++i) K.coarse_inv[static_cast<size_t>(i) * n + c] = y[i];
                }
            });
        }
    }
}

Topology Builder::run() {
    t_start_ = t_last_ = std::chrono::steady_clock::now();
    setup_geometry();
    setup_domain();
    stage_done(BUILD_GEOMETRY);
    refine();
    stage_done(BUILD_REFINE);
    classify();
    components();
    stage_done(BUILD_COMPONENTS);
    finalize();
    stage_done(BUILD_FINALIZE);
    build_multigrid();
    stage_done(BUILD_MULTIGRID);
    T_.build_ms[BUILD_TOTAL] = std::chrono::duration<double, std::milli>(t_last_ - t_start_).count();
    if (accel_) {
        T_.sampled_on_device = accel_->device();
        T_.build_ms[BUILD_DEVICE_KERNEL] = accel_->kernel_ms();
    }
    return std::move(T_);
}

}  // namespace

// instance tables of a single case
static void set_single_instance(Topology& T) {
    T.inst_brick_off = {0, T.nbricks};
    T.inst_patch_off = {0, static_cast<int>(T.patch_dof.size())};
    T.inst_active = {T.n_active};
    T.inst_reference_area = {T.reference_area};
    T.inst_body_volume = {T.body_volume};
    T.inst_max_diffusion = {T.max_diffusion_coefficient};
    T.mg_inst_off.assign(T.mg.size(), {});
    for (size_t m = 1; m < T.mg.size(); ++m) T.mg_inst_off[m] = {0, T.mg[m].n};
    const int last = static_cast<int>(T.mg.size()) - 1;
    if (last >= 1 && !T.mg.back().coarse_inv.empty()) {
        T.dense_level = {last};
        T.dense_off = {0};
        T.dense_n = {T.mg.back().n};
        T.dense_inv_off = {0};
        T.dense_inv = T.mg.back().coarse_inv;
    } else {
        T.dense_level = {-1};
        T.dense_off = {0};
        T.dense_n = {0};
        T.dense_inv_off = {0};
    }
}

Topology build_topology(const pcfd_geometry_desc& geometry, const pcfd_grid_desc& grid, BuildAccelerator* accel) {
    Builder b(geometry, grid, accel);
    Topology T = b.run();
    set_single_instance(T);
    return T;
}

Topology merge_topologies(const std::vector<const Topology*>& parts) {
    if (parts.empty()) throw BuildError("merge: no topology");
    const Topology& P0 = *parts[0];
    Topology T;
    for (int a = 0; a < 3; ++a) { T.lo[a] = P0.lo[a]; T.base_bricks[a] = P0.base_bricks[a]; }
    T.h0 = P0.h0;
    T.levels = P0.levels;
    T.outlet = P0.outlet;
    T.level_h = P0.level_h;
    T.finest_h = P0.finest_h;
    size_t depth = 0;
    for (const Topology* p : parts) {
        if (p->levels != P0.levels || p->level_h != P0.level_h || p->outlet != P0.outlet)
            throw BuildError("merge: instances need the same AMR levels, cell sizes and outlet");
        if (p->instances() != 1) throw BuildError("merge: parts must be single cases");
        if (p->dense_level[0] < 0) throw BuildError("merge: every instance needs a dense coarse inverse");
        depth = std::max(depth, p->mg.size());
    }
    T.mg.resize(depth);
    T.mg_inst_off.assign(depth, {});
    for (size_t m = 1; m < depth; ++m) T.mg_inst_off[m].push_back(0);
    T.inst_brick_off.push_back(0);
    T.inst_patch_off.push_back(0);
    T.level_brick_begin.assign(T.levels + 1, 0);
    int64_t dof0 = 0;
    int brick0 = 0, special0 = 0, edge0 = 0, patch0 = 0, inc0 = 0;
    std::vector<int> cell0(depth, 0), nnz0(depth, 0), child0(depth, 0);
    auto shift = [](int v, int64_t by) { return v >= 0 ? static_cast<int>(v + by) : v; };
    for (size_t i = 0; i < parts.size(); ++i) {
        const Topology& P = *parts[i];
        // bricks and cells
        T.brick_level.insert(T.brick_level.end(), P.brick_level.begin(), P.brick_level.end());
        T.brick_coord.insert(T.brick_coord.end(), P.brick_coord.begin(), P.brick_coord.end());
        T.brick_flags.insert(T.brick_flags.end(), P.brick_flags.begin(), P.brick_flags.end());
        for (int v : P.brick_nbr) T.brick_nbr.push_back(shift(v, brick0));
        T.mask.insert(T.mask.end(), P.mask.begin(), P.mask.end());
        T.centroid.insert(T.centroid.end(), P.centroid.begin(), P.centroid.end());
        // special cells
        for (int v : P.special_dof) T.special_dof.push_back(static_cast<int>(v + dof0));
        T.special_volume.insert(T.special_volume.end(), P.special_volume.begin(), P.special_volume.end());
        if (T.special_inc_off.empty()) T.special_inc_off.push_back(0);
        for (size_t k = 1; k < P.special_inc_off.size(); ++k) T.special_inc_off.push_back(P.special_inc_off[k] + inc0);
        for (int raw : P.special_inc) T.special_inc.push_back(raw >= 0 ? raw + edge0 : ~(~raw + edge0));
        for (int v : P.special_patch) T.special_patch.push_back(shift(v, patch0));
        if (T.brick_special_off.empty()) T.brick_special_off.push_back(0);
        for (size_t b = 1; b < P.brick_special_off.size(); ++b) T.brick_special_off.push_back(P.brick_special_off[b] + special0);
        // compact edges
        for (int v : P.edge_a) T.edge_a.push_back(static_cast<int>(v + dof0));
        for (int v : P.edge_b) T.edge_b.push_back(static_cast<int>(v + dof0));
        T.edge_axis.insert(T.edge_axis.end(), P.edge_axis.begin(), P.edge_axis.end());
        T.edge_area.insert(T.edge_area.end(), P.edge_area.begin(), P.edge_area.end());
        T.edge_coef.insert(T.edge_coef.end(), P.edge_coef.begin(), P.edge_coef.end());
        T.edge_w.insert(T.edge_w.end(), P.edge_w.begin(), P.edge_w.end());
        T.edge_dcoef.insert(T.edge_dcoef.end(), P.edge_dcoef.begin(), P.edge_dcoef.end());
        T.edge_x.insert(T.edge_x.end(), P.edge_x.begin(), P.edge_x.end());
        // surface patches
        for (int v : P.patch_dof) T.patch_dof.push_back(static_cast<int>(v + dof0));
        T.patch_S.insert(T.patch_S.end(), P.patch_S.begin(), P.patch_S.end());
        T.patch_x.insert(T.patch_x.end(), P.patch_x.begin(), P.patch_x.end());
        T.patch_D.insert(T.patch_D.end(), P.patch_D.begin(), P.patch_D.end());
        // redistribution groups
        const int group0 = static_cast<int>(T.redist_volume.size());
        for (int v : P.redist_group) T.redist_group.push_back(shift(v, group0));
        if (T.redist_off.empty()) T.redist_off.push_back(0);
        const int member0 = static_cast<int>(T.redist_member.size());
        for (size_t g = 1; g < P.redist_off.size(); ++g) T.redist_off.push_back(P.redist_off[g] + member0);
        for (int v : P.redist_member) T.redist_member.push_back(v + special0);
        T.redist_volume.insert(T.redist_volume.end(), P.redist_volume.begin(), P.redist_volume.end());
        // multigrid: level 0 carries the parent map and the cell exponents
        const int pdepth = static_cast<int>(P.mg.size());
        for (int v : P.mg[0].parent) T.mg[0].parent.push_back(pdepth > 1 ? shift(v, cell0[1]) : -1);
        T.mg[0].exponent.insert(T.mg[0].exponent.end(), P.mg[0].exponent.begin(), P.mg[0].exponent.end());
        for (size_t m = 1; m < depth; ++m) {
            MGLevel& D = T.mg[m];
            if (static_cast<int>(m) < pdepth) {
                const MGLevel& S = P.mg[m];
                if (D.row_off.empty()) D.row_off.push_back(0);
                for (int r = 0; r < S.n; ++r) D.row_off.push_back(S.row_off[r + 1] + nnz0[m]);
                for (int v : S.col) D.col.push_back(v + cell0[m]);
                D.coef.insert(D.coef.end(), S.coef.begin(), S.coef.end());
                D.diag.insert(D.diag.end(), S.diag.begin(), S.diag.end());
                for (int v : S.parent) D.parent.push_back(static_cast<int>(m) + 1 < pdepth ? shift(v, cell0[m + 1]) : -1);
                if (D.child_off.empty()) D.child_off.push_back(0);
                for (int c = 0; c < S.n; ++c) D.child_off.push_back(S.child_off[c + 1] + child0[m]);
                const int below = m == 1 ? 0 : cell0[m - 1];
                for (int v : S.child) D.child.push_back(m == 1 ? static_cast<int>(v + dof0) : v + below);
                D.exponent.insert(D.exponent.end(), S.exponent.begin(), S.exponent.end());
                D.n += S.n;
                nnz0[m] += static_cast<int>(S.col.size());
                child0[m] += static_cast<int>(S.child.size());
            }
            T.mg_inst_off[m].push_back(D.n);
        }
        // dense coarse block of this instance
        T.dense_level.push_back(P.dense_level[0]);
        T.dense_off.push_back(T.mg_inst_off[P.dense_level[0]][i]);
        T.dense_n.push_back(P.dense_n[0]);
        T.dense_inv_off.push_back(static_cast<int64_t>(T.dense_inv.size()));
        T.dense_inv.insert(T.dense_inv.end(), P.dense_inv.begin(), P.dense_inv.end());
        // per-instance facts
        T.inst_active.push_back(P.n_active);
        T.inst_reference_area.push_back(P.reference_area);
        T.inst_body_volume.push_back(P.body_volume);
        T.inst_max_diffusion.push_back(P.max_diffusion_coefficient);
        T.n_active += P.n_active;
        T.body_volume += P.body_volume;
        T.max_diffusion_coefficient = std::max(T.max_diffusion_coefficient, P.max_diffusion_coefficient);
        // advance the offsets
        for (size_t m = 1; m < depth; ++m) cell0[m] = T.mg_inst_off[m].back();
        dof0 += P.ndof;
        brick0 += P.nbricks;
        special0 += static_cast<int>(P.special_dof.size());
        edge0 += static_cast<int>(P.edge_a.size());
        patch0 += static_cast<int>(P.patch_dof.size());
        inc0 += static_cast<int>(P.special_inc.size());
        T.inst_brick_off.push_back(brick0);
        T.inst_patch_off.push_back(patch0);
    }
    T.nbricks = brick0;
    T.ndof = dof0;
    T.mg[0].n = static_cast<int>(dof0);
    T.reference_area = P0.reference_area;
    if (T.special_inc_off.empty()) T.special_inc_off.push_back(0);
    if (T.redist_off.empty()) T.redist_off.push_back(0);
    for (size_t m = 1; m < depth; ++m) {
        if (T.mg[m].row_off.empty()) T.mg[m].row_off.push_back(0);
        if (T.mg[m].child_off.empty()) T.mg[m].child_off.push_back(0);
    }
    return T;
}

}  // namespace pcfd
