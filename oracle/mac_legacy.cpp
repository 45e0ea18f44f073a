// ORACLE - TEST INFRASTRUCTURE ONLY.  Restatement of the reference's uniform FP64 MAC
// primitives whose outputs are frozen in tests/golden/*.f64 (written by
// tools/parity_probe.cu:102-178).  Uniform spacing only (the golden grid is 16^3, h=0.05).
//
//   poisson_rhs / apply / residual / jacobi / gs_band / restrict / prolong_add /
//   subtract_gradient / max_abs_divergence    project.cu:22-140, CPU twins :280-383
//   mac_advect (MacCormack, RK2 backtrace)    advect.cu:13-176, mac_grid.h:150-291
//   smagorinsky_nut / mac_diffuse             turbulence.cu:17-159
// Boundary conditions: the default BC of mac_grid.h (x no-slip, y free-slip, z-min no-slip,
// z-max moving lid at lid_u = 1).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace {

struct Grid {
    int nx, ny, nz;
    double h;
    int p(int i, int j, int k) const { return (k * ny + j) * nx + i; }
    int u(int i, int j, int k) const { return (k * ny + j) * (nx + 1) + i; }
    int v(int i, int j, int k) const { return (k * (ny + 1) + j) * nx + i; }
    int w(int i, int j, int k) const { return (k * ny + j) * nx + i; }
};

enum { NOSLIP = 0, FREESLIP = 1, MOVLID = 2 };
struct Walls { int xmin = NOSLIP, xmax = NOSLIP, ymin = FREESLIP, ymax = FREESLIP, zmin = NOSLIP, zmax = MOVLID; double lid = 1.0; };
const Walls kWalls{};

// ghost-aware reads: the normal index clamps onto the stored wall face, tangential
// out-of-range indices mirror with the wall's sign (and the lid's tangential speed for u)
double get_u(const double* f, const Grid& g, int i, int j, int k) {
    double s = 1.0, add = 0.0;
    i = std::clamp(i, 0, g.nx);
    if (j < 0) { if (kWalls.ymin == NOSLIP) s = -s; j = 0; }
    else if (j > g.ny - 1) { if (kWalls.ymax == NOSLIP) s = -s; j = g.ny - 1; }
    if (k < 0) { if (kWalls.zmin == NOSLIP) s = -s; k = 0; }
    else if (k > g.nz - 1) {
        if (kWalls.zmax == MOVLID) { add += 2.0 * kWalls.lid; s = -s; }
        else if (kWalls.zmax == NOSLIP) s = -s;
        k = g.nz - 1;
    }
    return s * f[g.u(i, j, k)] + add;
}
double get_v(const double* f, const Grid& g, int i, int j, int k) {
    double s = 1.0;
    j = std::clamp(j, 0, g.ny);
    if (i < 0) { if (kWalls.xmin == NOSLIP) s = -s; i = 0; }
    else if (i > g.nx - 1) { if (kWalls.xmax == NOSLIP) s = -s; i = g.nx - 1; }
    if (k < 0) { if (kWalls.zmin == NOSLIP) s = -s; k = 0; }
    else if (k > g.nz - 1) { if (kWalls.zmax == MOVLID || kWalls.zmax == NOSLIP) s = -s; k = g.nz - 1; }
    return s * f[g.v(i, j, k)];
}
double get_w(const double* f, const Grid& g, int i, int j, int k) {
    double s = 1.0;
    k = std::clamp(k, 0, g.nz);
    if (i < 0) { if (kWalls.xmin == NOSLIP) s = -s; i = 0; }
    else if (i > g.nx - 1) { if (kWalls.xmax == NOSLIP) s = -s; i = g.nx - 1; }
    if (j < 0) { if (kWalls.ymin == NOSLIP) s = -s; j = 0; }
    else if (j > g.ny - 1) { if (kWalls.ymax == NOSLIP) s = -s; j = g.ny - 1; }
    return s * f[g.w(i, j, k)];
}
double get(int comp, const double* f, const Grid& g, int i, int j, int k) {
    return comp == 0 ? get_u(f, g, i, j, k) : (comp == 1 ? get_v(f, g, i, j, k) : get_w(f, g, i, j, k));
}

// trilinear sample of a staggered component at a physical point; optional 8-corner bounds
double sample(int comp, const double* f, const Grid& g, double x, double y, double z, double* lo, double* hi) {
    // face-centred along the component's own axis, cell-centred along the others
    double t[3] = {x / g.h - 0.5, y / g.h - 0.5, z / g.h - 0.5};
    t[comp] += 0.5;
    int b[3];
    double fr[3];
    for (int a = 0; a < 3; ++a) { b[a] = static_cast<int>(std::floor(t[a])); fr[a] = t[a] - b[a]; }
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb)
            for (int cc = 0; cc < 2; ++cc) c[a][bb][cc] = get(comp, f, g, b[0] + a, b[1] + bb, b[2] + cc);
    if (lo && hi) {
        double mn = c[0][0][0], mx = c[0][0][0];
        for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb)
                for (int cc = 0; cc < 2; ++cc) { mn = std::min(mn, c[a][bb][cc]); mx = std::max(mx, c[a][bb][cc]); }
        *lo = mn; *hi = mx;
    }
    const double c00 = c[0][0][0] * (1 - fr[0]) + c[1][0][0] * fr[0];
    const double c10 = c[0][1][0] * (1 - fr[0]) + c[1][1][0] * fr[0];
    const double c01 = c[0][0][1] * (1 - fr[0]) + c[1][0][1] * fr[0];
    const double c11 = c[0][1][1] * (1 - fr[0]) + c[1][1][1] * fr[0];
    const double c0 = c00 * (1 - fr[1]) + c10 * fr[1];
    const double c1 = c01 * (1 - fr[1]) + c11 * fr[1];
    return c0 * (1 - fr[2]) + c1 * fr[2];
}

void keep_inside(const Grid& g, double& x, double& y, double& z) {
    const double e = 1e-9;
    x = std::clamp(x, e, g.nx * g.h - e);
    y = std::clamp(y, e, g.ny * g.h - e);
    z = std::clamp(z, e, g.nz * g.h - e);
}

void velocity(const double* u, const double* v, const double* w, const Grid& g, double x, double y, double z, double out[3]) {
    out[0] = sample(0, u, g, x, y, z, nullptr, nullptr);
    out[1] = sample(1, v, g, x, y, z, nullptr, nullptr);
    out[2] = sample(2, w, g, x, y, z, nullptr, nullptr);
}

// midpoint (RK2) trace over dt; dt > 0 goes upstream
void trace(const double* u, const double* v, const double* w, const Grid& g, double dt, const double p[3], double q[3]) {
    double a[3];
    velocity(u, v, w, g, p[0], p[1], p[2], a);
    double m[3] = {p[0] - 0.5 * dt * a[0], p[1] - 0.5 * dt * a[1], p[2] - 0.5 * dt * a[2]};
    keep_inside(g, m[0], m[1], m[2]);
    velocity(u, v, w, g, m[0], m[1], m[2], a);
    q[0] = p[0] - dt * a[0]; q[1] = p[1] - dt * a[1]; q[2] = p[2] - dt * a[2];
    keep_inside(g, q[0], q[1], q[2]);
}

void extent(int comp, const Grid& g, int n[3]) {
    n[0] = g.nx; n[1] = g.ny; n[2] = g.nz;
    n[comp] += 1;
}
int index_of(int comp, const Grid& g, int i, int j, int k) {
    return comp == 0 ? g.u(i, j, k) : (comp == 1 ? g.v(i, j, k) : g.w(i, j, k));
}
bool interior(int comp, const Grid& g, int i, int j, int k) {
    const int c[3] = {i, j, k};
    int n[3];
    extent(comp, g, n);
    for (int a = 0; a < 3; ++a) {
        const int lo = a == comp ? 1 : 0, hi = a == comp ? n[a] - 2 : n[a] - 1;
        if (c[a] < lo || c[a] > hi) return false;
    }
    return true;
}
void node_position(int comp, const Grid& g, int i, int j, int k, double p[3]) {
    p[0] = (i + 0.5) * g.h; p[1] = (j + 0.5) * g.h; p[2] = (k + 0.5) * g.h;
    p[comp] -= 0.5 * g.h;
}

void advect_component(int comp, const double* f, const double* u, const double* v, const double* w, double* out,
                      const Grid& g, double dt, int band) {
    int n[3];
    extent(comp, g, n);
    const int total = n[0] * n[1] * n[2];
    std::vector<double> hat(total);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int id = index_of(comp, g, i, j, k);
                if (!interior(comp, g, i, j, k)) { hat[id] = f[id]; continue; }
                double p[3], q[3];
                node_position(comp, g, i, j, k, p);
                trace(u, v, w, g, dt, p, q);
                hat[id] = sample(comp, f, g, q[0], q[1], q[2], nullptr, nullptr);
            }
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int id = index_of(comp, g, i, j, k);
                if (!interior(comp, g, i, j, k)) { out[id] = f[id]; continue; }
                const bool wall = i < band || i > n[0] - 1 - band || j < band || j > n[1] - 1 - band || k < band ||
                                  k > n[2] - 1 - band;
                if (wall) { out[id] = hat[id]; continue; }
                double p[3], q[3], r[3], lo, hi;
                node_position(comp, g, i, j, k, p);
                trace(u, v, w, g, dt, p, q);
                sample(comp, f, g, q[0], q[1], q[2], &lo, &hi);
                trace(u, v, w, g, -dt, p, r);
                const double back = sample(comp, hat.data(), g, r[0], r[1], r[2], nullptr, nullptr);
                out[id] = std::clamp(hat[id] + 0.5 * (f[id] - back), lo, hi);
            }
}

// 7-point Neumann stencil on the uniform box: returns (sum of weights, weighted neighbour sum)
void stencil(const double* p, const Grid& g, int i, int j, int k, double& diag, double& nb) {
    const double w = 1.0 / (g.h * g.h);
    diag = 0; nb = 0;
    if (i > 0) { diag += w; nb += w * p[g.p(i - 1, j, k)]; }
    if (i < g.nx - 1) { diag += w; nb += w * p[g.p(i + 1, j, k)]; }
    if (j > 0) { diag += w; nb += w * p[g.p(i, j - 1, k)]; }
    if (j < g.ny - 1) { diag += w; nb += w * p[g.p(i, j + 1, k)]; }
    if (k > 0) { diag += w; nb += w * p[g.p(i, j, k - 1)]; }
    if (k < g.nz - 1) { diag += w; nb += w * p[g.p(i, j, k + 1)]; }
}
double divergence(const double* u, const double* v, const double* w, const Grid& g, int i, int j, int k) {
    return (u[g.u(i + 1, j, k)] - u[g.u(i, j, k)]) / g.h + (v[g.v(i, j + 1, k)] - v[g.v(i, j, k)]) / g.h +
           (w[g.w(i, j, k + 1)] - w[g.w(i, j, k)]) / g.h;
}
// cell-centred factor-2 linear transfer weights along one axis
void transfer(int fi, int nc, int& c0, int& c1, double& w0, double& w1) {
    const double pos = (fi + 0.5) * 0.5 - 0.5;
    const int b = static_cast<int>(std::floor(pos));
    const double t = pos - b;
    c0 = std::clamp(b, 0, nc - 1);
    c1 = std::clamp(b + 1, 0, nc - 1);
    w0 = 1.0 - t; w1 = t;
}

double centred(int comp, const double* f, const Grid& g, int i, int j, int k) {
    int o[3] = {i, j, k};
    const double a = get(comp, f, g, o[0], o[1], o[2]);
    o[comp] += 1;
    return 0.5 * (a + get(comp, f, g, o[0], o[1], o[2]));
}

}  // namespace

extern "C" {

void legacy_fill(int64_t count, uint32_t seed, double lo, double hi, double* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> dist(lo, hi);
    for (int64_t i = 0; i < count; ++i) out[i] = dist(rng);
}

void legacy_poisson_rhs(int nx, int ny, int nz, double h, const double* u, const double* v, const double* w, double rho,
                        double dt, double* rhs) {
    Grid g{nx, ny, nz, h};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) rhs[g.p(i, j, k)] = -(rho / dt) * divergence(u, v, w, g, i, j, k);
}
void legacy_poisson_apply(int nx, int ny, int nz, double h, const double* p, double* out) {
    Grid g{nx, ny, nz, h};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double d, nb;
                stencil(p, g, i, j, k, d, nb);
                out[g.p(i, j, k)] = d * p[g.p(i, j, k)] - nb;
            }
}
void legacy_poisson_residual(int nx, int ny, int nz, double h, const double* p, const double* rhs, double* out) {
    Grid g{nx, ny, nz, h};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double d, nb;
                stencil(p, g, i, j, k, d, nb);
                out[g.p(i, j, k)] = rhs[g.p(i, j, k)] - (d * p[g.p(i, j, k)] - nb);
            }
}
void legacy_jacobi(int nx, int ny, int nz, double h, double* p, const double* rhs, double omega, int sweeps) {
    Grid g{nx, ny, nz, h};
    std::vector<double> next(static_cast<size_t>(nx) * ny * nz);
    for (int s = 0; s < sweeps; ++s) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int c = g.p(i, j, k);
                    double d, nb;
                    stencil(p, g, i, j, k, d, nb);
                    next[c] = d == 0.0 ? p[c] : p[c] + omega * (rhs[c] - (d * p[c] - nb)) / d;
                }
        std::copy(next.begin(), next.end(), p);
    }
}
void legacy_gs_band(int nx, int ny, int nz, double h, double* p, const double* rhs, int band, int sweeps, int forward) {
    Grid g{nx, ny, nz, h};
    for (int s = 0; s < sweeps; ++s)
        for (int pass = 0; pass < 2; ++pass) {
            const int colour = forward ? pass : 1 - pass;
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        if (((i + j + k) & 1) != colour) continue;
                        const bool in = i < band || i >= nx - band || j < band || j >= ny - band || k < band || k >= nz - band;
                        if (!in) continue;
                        double d, nb;
                        stencil(p, g, i, j, k, d, nb);
                        if (d != 0.0) p[g.p(i, j, k)] = (rhs[g.p(i, j, k)] + nb) / d;
                    }
        }
}
void legacy_restrict(int nx, int ny, int nz, double h, const double* fine, double* coarse) {
    Grid f{nx, ny, nz, h}, c{nx / 2, ny / 2, nz / 2, 2 * h};
    std::fill(coarse, coarse + static_cast<size_t>(c.nx) * c.ny * c.nz, 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int x[2], y[2], z[2];
                double wx[2], wy[2], wz[2];
                transfer(i, c.nx, x[0], x[1], wx[0], wx[1]);
                transfer(j, c.ny, y[0], y[1], wy[0], wy[1]);
                transfer(k, c.nz, z[0], z[1], wz[0], wz[1]);
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int d = 0; d < 2; ++d)
                            coarse[c.p(x[a], y[b], z[d])] += fine[f.p(i, j, k)] * 0.125 * wx[a] * wy[b] * wz[d];
            }
}
void legacy_prolong_add(int nx, int ny, int nz, double h, const double* coarse, double* fine) {
    Grid f{nx, ny, nz, h}, c{nx / 2, ny / 2, nz / 2, 2 * h};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int x[2], y[2], z[2];
                double wx[2], wy[2], wz[2];
                transfer(i, c.nx, x[0], x[1], wx[0], wx[1]);
                transfer(j, c.ny, y[0], y[1], wy[0], wy[1]);
                transfer(k, c.nz, z[0], z[1], wz[0], wz[1]);
                double s = 0;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int d = 0; d < 2; ++d) s += wx[a] * wy[b] * wz[d] * coarse[c.p(x[a], y[b], z[d])];
                fine[f.p(i, j, k)] += s;
            }
}
void legacy_subtract_gradient(int nx, int ny, int nz, double h, double* u, double* v, double* w, const double* p,
                              double rho, double dt) {
    Grid g{nx, ny, nz, h};
    const double s = dt / rho;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 1; i < nx; ++i) u[g.u(i, j, k)] -= s * (p[g.p(i, j, k)] - p[g.p(i - 1, j, k)]) / h;
    for (int k = 0; k < nz; ++k)
        for (int j = 1; j < ny; ++j)
            for (int i = 0; i < nx; ++i) v[g.v(i, j, k)] -= s * (p[g.p(i, j, k)] - p[g.p(i, j - 1, k)]) / h;
    for (int k = 1; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) w[g.w(i, j, k)] -= s * (p[g.p(i, j, k)] - p[g.p(i, j, k - 1)]) / h;
}
double legacy_max_abs_divergence(int nx, int ny, int nz, double h, const double* u, const double* v, const double* w) {
    Grid g{nx, ny, nz, h};
    double m = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) m = std::max(m, std::abs(divergence(u, v, w, g, i, j, k)));
    return m;
}
void legacy_mac_advect(int nx, int ny, int nz, double h, const double* u, const double* v, const double* w, double* uo,
                       double* vo, double* wo, double dt, int band) {
    Grid g{nx, ny, nz, h};
    advect_component(0, u, u, v, w, uo, g, dt, band);
    advect_component(1, v, u, v, w, vo, g, dt, band);
    advect_component(2, w, u, v, w, wo, g, dt, band);
}
void legacy_smagorinsky_nut(int nx, int ny, int nz, double h, const double* u, const double* v, const double* w,
                            double cs, double* nut) {
    Grid g{nx, ny, nz, h};
    const double* f[3] = {u, v, w};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int c[3] = {i, j, k};
                double grad[3][3];  // grad[comp][axis]
                for (int comp = 0; comp < 3; ++comp)
                    for (int axis = 0; axis < 3; ++axis) {
                        if (comp == axis) {
                            int o[3] = {i, j, k};
                            const double a = get(comp, f[comp], g, o[0], o[1], o[2]);
                            o[axis] += 1;
                            grad[comp][axis] = (get(comp, f[comp], g, o[0], o[1], o[2]) - a) / h;
                        } else {
                            int m[3] = {c[0], c[1], c[2]}, p[3] = {c[0], c[1], c[2]};
                            m[axis] -= 1; p[axis] += 1;
                            grad[comp][axis] = (centred(comp, f[comp], g, p[0], p[1], p[2]) -
                                                centred(comp, f[comp], g, m[0], m[1], m[2])) / (2 * h);
                        }
                    }
                const double sxy = 0.5 * (grad[0][1] + grad[1][0]), sxz = 0.5 * (grad[0][2] + grad[2][0]),
                             syz = 0.5 * (grad[1][2] + grad[2][1]);
                const double s2 = grad[0][0] * grad[0][0] + grad[1][1] * grad[1][1] + grad[2][2] * grad[2][2] +
                                  2.0 * (sxy * sxy + sxz * sxz + syz * syz);
                const double l = cs * std::cbrt(h * h * h);
                nut[g.p(i, j, k)] = l * l * std::sqrt(2.0 * s2);
            }
}
void legacy_mac_diffuse(int nx, int ny, int nz, double h, const double* u, const double* v, const double* w,
                        const double* nut, double dt, double nu, double* uo, double* vo, double* wo) {
    Grid g{nx, ny, nz, h};
    const double* in[3] = {u, v, w};
    double* out[3] = {uo, vo, wo};
    for (int comp = 0; comp < 3; ++comp) {
        int n[3];
        extent(comp, g, n);
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    const int id = index_of(comp, g, i, j, k);
                    if (!interior(comp, g, i, j, k)) { out[comp][id] = in[comp][id]; continue; }
                    const double c = in[comp][id];
                    double lap = 0;
                    for (int axis = 0; axis < 3; ++axis) {
                        int m[3] = {i, j, k}, p[3] = {i, j, k};
                        m[axis] -= 1; p[axis] += 1;
                        const double fm = get(comp, in[comp], g, m[0], m[1], m[2]);
                        const double fp = get(comp, in[comp], g, p[0], p[1], p[2]);
                        lap += 2.0 / (h + h) * ((fp - c) / h + (fm - c) / h);
                    }
                    double nf = 0.0;
                    if (nut) {
                        int m[3] = {i, j, k};
                        m[comp] -= 1;
                        nf = 0.5 * (nut[g.p(m[0], m[1], m[2])] + nut[g.p(i, j, k)]);
                    }
                    out[comp][id] = c + dt * (nu + nf) * lap;
                }
    }
}

}  // extern "C"
