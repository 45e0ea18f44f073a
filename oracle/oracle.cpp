// ORACLE - TEST INFRASTRUCTURE ONLY.  Serial FP64 restatement of one ParaCFD solver step on
// the shared brick/cut-cell topology.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker / reported CPU baseline.
//
// Each stage follows the reference's CPU twin or kernel named in its comment:
//   outflow_rate      max_outflow_rate            amr_advection.cu:169-186,1316-1320
//   advect            transport_regular/compact   amr_advection.cu:62-112, CPU twin :874-888
//   diffuse           diffuse_*_cell_momentum      amr_advection.cu:233-252, :913-919
//   reconstruct       reconstruct_*_cell_flux      amr_advection.cu:257-293, :1069
//   rhs               structured_integrated_flux + normalize_divergence_rhs  amr_pressure.cu:255-284
//   apply             structured_apply + irregular_orthogonal_gather         amr_pressure.cu:72-134
//   pcg               solve_orthogonal            amr_pressure.cu:1267-1353
//   correct           structured/special_flux_correction                     amr_pressure.cu:285-330
//   impulse           pressure_*_cell_impulse     amr_advection.cu:380-405, CPU twin :979-1053
//   smooth_wall       apply_composite_cell_smooth_fabric_wall_cpu  amr_advection.cu:920-942,
//                     friction velocity amr_pressure.h:41-53
//   loads             compute_pressure_loads      aero_loads.cpp:35-47,
//                     accumulate_viscous_loads    aero_loads.cpp:49-56
// The V-cycle is this project's geometric multigrid (the reference uses aggregation AMG,
// amr_pressure.cu:1103-1154); its primitives are pinned by the reference golden vectors.
//
// Parity status: the V-cycle primitives (apply, residual, Jacobi) are pinned by
// tests/golden/*.f64 through mac_legacy.cpp; transport, reconstruction, projection,
// impulse and loads are pinned by the physics contracts in tests/test_oracle_contracts.py
// (closed forms, invariants, brute force).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/paracfd.h"

namespace {

constexpr int BS = 8, BC = 512;
enum : uint8_t { BF_XMIN = 1, BF_XMAX = 2, BF_YMIN = 4, BF_YMAX = 8, BF_ZMIN = 16, BF_ZMAX = 32 };
enum : uint8_t { CM_PX = 1, CM_MX = 8, CM_ACTIVE = 64, CM_SPECIAL = 128 };

thread_local std::string g_error;

}  // namespace

extern "C" {

typedef struct oracle_topology {
    int64_t nbricks, ndof, nspecial, nedges, npatches;
    int levels, outlet;
    const double* level_h;
    const int* brick_level;
    const uint8_t* brick_flags;
    const int* brick_nbr;
    const uint8_t* mask;
    const int* special_dof;
    const double* special_volume;
    const int* special_patch;
    const int* edge_a;
    const int* edge_b;
    const int8_t* edge_axis;
    const double* edge_area;
    const double* edge_coef;
    const double* edge_w;
    const int* patch_dof;
    const double* patch_S;
    const double* patch_x;
    const double* patch_D;           /* 9 per patch: int_W (x - c) (x) n dA */
    const double* edge_x;            /* 3 per edge: face centroid of the open part */
    const double* edge_dcoef;        /* diffusion conductance */
    double finest_h;
    double max_diffusion_coefficient;
    int mg_levels;
    const int64_t* mg_n;             /* cells per level, mg_n[0] = ndof */
    const int* const* mg_parent;     /* per level (last is unused) */
    const int* const* mg_row_off;    /* per level, entry 0 unused */
    const int* const* mg_col;
    const double* const* mg_coef;
    const double* const* mg_diag;
    const double* coarse_inv;        /* may be NULL */
    const double* centroid;          /* 3 per dof: fluid centroid (Smagorinsky gradients) */
    int64_t ngroups;                 /* momentum redistribution groups */
    const int* redist_group;         /* per special cell: group or -1 (may be NULL when ngroups = 0) */
    const int* redist_off;           /* groups + 1 */
    const int* redist_member;        /* special ranks */
    const double* redist_volume;     /* per group */
} oracle_topology;

}

namespace {

struct Oracle {
    // topology
    int64_t nb = 0, n = 0;
    int levels = 1;
    bool outlet = true;
    std::vector<double> level_h;
    std::vector<int> brick_level, brick_nbr;
    std::vector<uint8_t> brick_flags, mask;
    std::vector<int> special_of;     // dof -> special or -1
    std::vector<int64_t> sdof;       // special -> dof
    std::vector<double> special_volume;
    std::vector<int> special_patch;
    std::vector<int> ea, eb;
    std::vector<int8_t> eaxis;
    std::vector<double> earea, ecoef, ew;
    std::vector<int> patch_dof;
    std::vector<double> patch_S, patch_x, patch_D;
    std::vector<double> edcoef;      // diffusion conductance per edge
    std::vector<double> edisp;       // 6 per edge: x_f - c_a, x_f - c_b
    std::vector<double> centroid;
    // momentum redistribution groups: special rank -> group (or -1), members, volumes
    std::vector<int> rgroup, roff, rmember;
    std::vector<double> rvolume;
    double finest_h = 0, max_diff_coef = 0;
    // Smagorinsky least-squares gradient stencils of the special cells (built on first use)
    std::vector<int> sg_off, sg_nbr;
    std::vector<double> sg_wd, sg_inv;
    struct Level { int64_t n; std::vector<int> parent, row_off, col; std::vector<double> coef, diag; };
    std::vector<Level> mg;
    std::vector<double> coarse_inv;
    // options
    pcfd_options opt{};
    double ref_area = 0, origin[3] = {0, 0, 0};
    // state
    std::vector<double> vx, vy, vz, fx, fy, fz, fe, p;
    double time = 0;
    bool initialized = false;
    std::vector<double> diag0;  // composite diagonal
    // Planted mistake for the contract tests (oracle_plant): 0 none; 1 CFL rate without the
    // x-min inflow face; 2 structured diffusion conductance h^2 instead of h; 3 diffusion
    // substeps rounded down; 4 RHS without the 1/dt; 5 outlet correction over h instead of h/2;
    // 6 wall force p_a S + g.(x_p - c_a) S (the point form at the patch point, without the wall
    // moment); 7 compact face pressure without the skew term; 8 compact reconstruction without
    // the transverse correction; 9 compact transport donor-only; 10 compact transport slope
    // without the (1 - C) factor; 11 gradient stencil weight area/|d| instead of area/|d|^2;
    // 12 wall traction in the impulse without D^T g (the loads keep it); 13 a redistribution
    // group's increment divided by each member's own volume instead of the group's.
    int defect = 0;

    double h_of(int64_t a) const { return level_h[brick_level[a / BC]]; }
    bool in_group(int64_t a) const {
        const int s = special_of[a];
        return s >= 0 && !rgroup.empty() && rgroup[s] >= 0;
    }
    // Momentum increments d (per cell, 3 components) applied as v += d / V, except on the members
    // of a redistribution group: they all take the group's summed increment over the group's
    // volume (apply_or_accumulate_redistributed_momentum + finalize_redistributed_state,
    // amr_advection.cu:118-149, 1251-1268).
    void apply_increments(const std::vector<double> (&d)[3]) {
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a) || in_group(a)) continue;
            const double iv = 1.0 / vol(a);
            vx[a] += d[0][a] * iv; vy[a] += d[1][a] * iv; vz[a] += d[2][a] * iv;
        }
        for (size_t g = 0; g + 1 < roff.size(); ++g) {
            double S[3] = {0, 0, 0};
            for (int k = roff[g]; k < roff[g + 1]; ++k)
                for (int c = 0; c < 3; ++c) S[c] += d[c][sdof[rmember[k]]];
            for (int k = roff[g]; k < roff[g + 1]; ++k) {
                const int64_t a = sdof[rmember[k]];
                const double iv = 1.0 / (defect == 13 ? vol(a) : rvolume[g]);
                vx[a] += S[0] * iv; vy[a] += S[1] * iv; vz[a] += S[2] * iv;
            }
        }
    }
    double vol(int64_t a) const {
        const int s = special_of[a];
        if (s >= 0) return special_volume[s];
        const double h = h_of(a);
        return h * h * h;
    }
    bool active(int64_t a) const { return mask[a] & CM_ACTIVE; }
    // structured neighbour in direction (axis, +1/-1); requires the mask bit
    int64_t snbr(int64_t a, int axis, int sign) const {
        const uint8_t bit = sign > 0 ? (CM_PX << axis) : (CM_MX << axis);
        if (!(mask[a] & bit)) return -1;
        const int64_t b = a / BC;
        const int q = static_cast<int>(a % BC);
        int c[3] = {q & 7, (q >> 3) & 7, q >> 6};
        int64_t bb = b;
        c[axis] += sign;
        if (c[axis] < 0 || c[axis] >= BS) {
            bb = brick_nbr[6 * b + 2 * axis + (sign > 0)];
            c[axis] = sign > 0 ? 0 : BS - 1;
        }
        return bb * BC + c[0] + 8 * c[1] + 64 * c[2];
    }
    bool at_xmin(int64_t a) const { return (a % BC & 7) == 0 && (brick_flags[a / BC] & BF_XMIN); }
    bool at_xmax(int64_t a) const { return (a % BC & 7) == 7 && (brick_flags[a / BC] & BF_XMAX); }
    double& face(int axis, int64_t a) { return axis == 0 ? fx[a] : (axis == 1 ? fy[a] : fz[a]); }
    double comp(int c, int64_t a) const { return c == 0 ? vx[a] : (c == 1 ? vy[a] : vz[a]); }

    // ---- transport ------------------------------------------------------------------------
    static double minmod(double a, double b) {
        if (a * b <= 0) return 0.0;
        return a >= 0 ? std::min(a, b) : std::max(a, b);
    }
    double outflow_rate() const {
        std::vector<double> out(n, 0.0);
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a), A = h * h;
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double f = A * (axis == 0 ? fx[a] : (axis == 1 ? fy[a] : fz[a]));
                out[a] += std::max(f, 0.0);
                out[b] += std::max(-f, 0.0);
            }
            if (at_xmin(a) && defect != 1) out[a] += std::max(-A * opt.speed, 0.0);
            if (outlet && at_xmax(a)) out[a] += std::max(A * fx[a], 0.0);
        }
        for (size_t e = 0; e < ea.size(); ++e) {
            const double f = earea[e] * fe[e];
            out[ea[e]] += std::max(f, 0.0);
            out[eb[e]] += std::max(-f, 0.0);
        }
        double rate = 0;
        for (int64_t a = 0; a < n; ++a) if (active(a)) rate = std::max(rate, out[a] / vol(a));
        return rate;
    }
    void advect(double dt) {
        std::vector<double> d[3] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a), A = h * h;
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double un = axis == 0 ? fx[a] : (axis == 1 ? fy[a] : fz[a]);
                const double swept = dt * A * un, courant = std::abs(un) * dt / h;
                const int64_t lower = snbr(a, axis, -1), upper = snbr(b, axis, 1);
                const double pred = 0.5 * (1.0 - std::min(1.0, courant));
                for (int c = 0; c < 3; ++c) {
                    const double va = comp(c, a), vb = comp(c, b);
                    double fv;
                    if (swept >= 0) fv = lower < 0 ? va : va + pred * minmod(va - comp(c, lower), vb - va);
                    else fv = upper < 0 ? vb : vb - pred * minmod(vb - va, comp(c, upper) - vb);
                    d[c][a] -= swept * fv;
                    d[c][b] += swept * fv;
                }
            }
            if (at_xmin(a)) {
                const double swept = dt * A * opt.speed;
                if (swept >= 0) d[0][a] += swept * opt.speed;
                else for (int c = 0; c < 3; ++c) d[c][a] += swept * comp(c, a);
            }
            if (outlet && at_xmax(a)) {
                const double swept = dt * A * fx[a];
                if (swept >= 0) for (int c = 0; c < 3; ++c) d[c][a] -= swept * comp(c, a);
                else d[0][a] -= swept * opt.speed;
            }
        }
        // compact faces: the donor's value reconstructed to the face centroid with its
        // least-squares gradient, scaled by 1 - C (C = min(1, |swept| / V_donor)) and clamped
        // to the range of the two cell values (compact_muscl_cell_face_value,
        // amr_advection.cu:94-105); gradients of the velocity before the transport
        std::vector<double> g;
        velocity_gradients(g);
        for (size_t e = 0; e < ea.size(); ++e) {
            const double swept = dt * earea[e] * fe[e];
            const bool lower = swept >= 0;
            const int64_t donor = lower ? ea[e] : eb[e];
            const int sp = special_of[donor];
            const double C = std::min(1.0, std::abs(swept) / vol(donor));
            const double* dsp = &edisp[6 * e + (lower ? 0 : 3)];
            for (int c = 0; c < 3; ++c) {
                const double va = comp(c, ea[e]), vb = comp(c, eb[e]);
                double fv = lower ? va : vb;
                if (sp >= 0 && defect != 9) {
                    const double* gc = &g[9 * static_cast<size_t>(sp) + 3 * c];
                    const double slope = defect == 10 ? 1.0 : 1.0 - C;
                    const double rec = fv + slope * (gc[0] * dsp[0] + gc[1] * dsp[1] + gc[2] * dsp[2]);
                    fv = std::max(std::min(va, vb), std::min(std::max(va, vb), rec));
                }
                const double f = swept * fv;
                d[c][ea[e]] -= f;
                d[c][eb[e]] += f;
            }
        }
        apply_increments(d);
    }
    // ---- diffusion and the Smagorinsky model ------------------------------------------------
    // Symmetric 3x3 pseudo-inverse: cyclic Jacobi eigen-decomposition (Golub & Van Loan,
    // Alg. 8.4.3), keeping the modes whose eigenvalue exceeds 1e-4 of the largest (the rank
    // rule of amr_advection.cu:822-831).
    static void pinv3(const double* A, double* out) {
        double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) a[r][c] = A[3 * r + c];
        for (int sweep = 0; sweep < 64; ++sweep) {
            const double off = std::abs(a[0][1]) + std::abs(a[0][2]) + std::abs(a[1][2]);
            const double dia = std::abs(a[0][0]) + std::abs(a[1][1]) + std::abs(a[2][2]);
            if (!(off > 1e-300) || off <= 1e-17 * dia) break;
            for (int p = 0; p < 2; ++p)
                for (int q = p + 1; q < 3; ++q) {
                    if (a[p][q] == 0.0) continue;
                    const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
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
        const double lmax = std::max({std::abs(a[0][0]), std::abs(a[1][1]), std::abs(a[2][2])});
        for (int k = 0; k < 9; ++k) out[k] = 0.0;
        for (int m = 0; m < 3; ++m) {
            const double l = a[m][m];
            if (!(l > 1e-4 * lmax)) continue;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) out[3 * r + c] += v[r][m] * v[c][m] / l;
        }
    }
    // Least-squares gradient stencils of the special cells (amr_advection.cu:834):
    // neighbours across the open structured faces (area h^2) and the compact edges (open
    // area), weight area / |x_b - x_a|^2; the normal matrix is pseudo-inverted.
    void build_gradient_stencils() {
        const int ns = static_cast<int>(special_volume.size());
        sg_off.assign(ns + 1, 0);
        sg_nbr.clear(); sg_wd.clear(); sg_inv.assign(9 * static_cast<size_t>(ns), 0.0);
        std::vector<std::vector<int>> inc(ns);
        for (size_t e = 0; e < ea.size(); ++e) {
            if (special_of[ea[e]] >= 0) inc[special_of[ea[e]]].push_back(static_cast<int>(e));
            if (special_of[eb[e]] >= 0) inc[special_of[eb[e]]].push_back(~static_cast<int>(e));
        }
        for (int sp = 0; sp < ns; ++sp) {
            const int64_t a = sdof[sp];
            double N[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            auto add = [&](int64_t b, double area) {
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = centroid[3 * b + c] - centroid[3 * a + c];
                const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (!(area > 0) || !(d2 > 1e-20)) throw std::runtime_error("invalid gradient stencil geometry");
                const double w = defect == 11 ? area / std::sqrt(d2) : area / d2;
                sg_nbr.push_back(static_cast<int>(b));
                for (int c = 0; c < 3; ++c) sg_wd.push_back(w * d[c]);
                for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) N[3 * r + c] += w * d[r] * d[c];
            };
            const double h = h_of(a);
            for (int axis = 0; axis < 3; ++axis)
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                    const int64_t b = snbr(a, axis, sgn);
                    if (b >= 0) add(b, h * h);
                }
            for (int raw : inc[sp]) {
                const int e = raw >= 0 ? raw : ~raw;
                add(raw >= 0 ? eb[e] : ea[e], earea[e]);
            }
            sg_off[sp + 1] = static_cast<int>(sg_nbr.size());
            pinv3(N, &sg_inv[9 * static_cast<size_t>(sp)]);
        }
    }
    // Least-squares gradient at every special cell of the current velocity (9 per cell,
    // d v_c / d x_r at 3 c + r; accumulate_special_cell_gradient + solve_special_cell_gradient,
    // amr_advection.cu:206-216) or of the pressure (3 per cell; the same stencil applied to p,
    // as reconstruct_scalar_pressure_gradient applies the pressure topology's, :400-405).
    void velocity_gradients(std::vector<double>& g) const {
        const int ns = static_cast<int>(special_volume.size());
        g.assign(9 * static_cast<size_t>(ns), 0.0);
        for (int sp = 0; sp < ns; ++sp) {
            const int64_t a = sdof[sp];
            double rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = sg_off[sp]; k < sg_off[sp + 1]; ++k)
                for (int c = 0; c < 3; ++c) {
                    const double dv = comp(c, sg_nbr[k]) - comp(c, a);
                    for (int r = 0; r < 3; ++r) rhs[3 * c + r] += sg_wd[3 * k + r] * dv;
                }
            const double* M = &sg_inv[9 * static_cast<size_t>(sp)];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    g[9 * static_cast<size_t>(sp) + 3 * c + r] = M[3 * r] * rhs[3 * c] + M[3 * r + 1] * rhs[3 * c + 1] + M[3 * r + 2] * rhs[3 * c + 2];
        }
    }
    void pressure_gradients(std::vector<double>& g) const {
        const int ns = static_cast<int>(special_volume.size());
        g.assign(3 * static_cast<size_t>(ns), 0.0);
        for (int sp = 0; sp < ns; ++sp) {
            const int64_t a = sdof[sp];
            double rhs[3] = {0, 0, 0};
            for (int k = sg_off[sp]; k < sg_off[sp + 1]; ++k) {
                const double dp = p[sg_nbr[k]] - p[a];
                for (int r = 0; r < 3; ++r) rhs[r] += sg_wd[3 * k + r] * dp;
            }
            const double* M = &sg_inv[9 * static_cast<size_t>(sp)];
            for (int r = 0; r < 3; ++r) g[3 * static_cast<size_t>(sp) + r] = M[3 * r] * rhs[0] + M[3 * r + 1] * rhs[1] + M[3 * r + 2] * rhs[2];
        }
    }
    static double strain_rate(const double* g) {  // sqrt(2 S:S), S = sym(grad v)
        const double sxx = g[0], syy = g[4], szz = g[8];
        const double sxy = 0.5 * (g[1] + g[3]), sxz = 0.5 * (g[2] + g[6]), syz = 0.5 * (g[5] + g[7]);
        return std::sqrt(std::max(0.0, 2.0 * (sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz))));
    }
    // Cell viscosity nu_mol + (Cs V^(1/3))^2 |S| (amr_advection.cu:198-204 regular cells,
    // central or one-sided differences; :219-223 special cells, least squares)
    void viscosity(std::vector<double>& nu) {
        nu.assign(n, 0.0);
        const double cs = opt.smagorinsky_cs;
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            nu[a] = opt.nu;
            if (!(cs > 0)) continue;
            double g[9];
            const int sp = special_of[a];
            if (sp < 0) {
                const double h = h_of(a);
                for (int c = 0; c < 3; ++c)
                    for (int axis = 0; axis < 3; ++axis) {
                        const int64_t lo = snbr(a, axis, -1), up = snbr(a, axis, 1);
                        double v = 0.0;
                        if (lo >= 0 && up >= 0) v = (comp(c, up) - comp(c, lo)) / (2.0 * h);
                        else if (up >= 0) v = (comp(c, up) - comp(c, a)) / h;
                        else if (lo >= 0) v = (comp(c, a) - comp(c, lo)) / h;
                        g[3 * c + axis] = v;
                    }
            } else {
                double rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int k = sg_off[sp]; k < sg_off[sp + 1]; ++k)
                    for (int c = 0; c < 3; ++c) {
                        const double dv = comp(c, sg_nbr[k]) - comp(c, a);
                        for (int axis = 0; axis < 3; ++axis) rhs[3 * c + axis] += sg_wd[3 * k + axis] * dv;
                    }
                const double* M = &sg_inv[9 * static_cast<size_t>(sp)];
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                        g[3 * c + r] = M[3 * r] * rhs[3 * c] + M[3 * r + 1] * rhs[3 * c + 1] + M[3 * r + 2] * rhs[3 * c + 2];
            }
            const double L = cs * std::cbrt(vol(a));
            nu[a] += L * L * strain_rate(g);
        }
    }
    // max over cells of sum_faces conductance (nu_a + nu_b)/2 / V (amr_advection.cu:1343)
    double diffusion_rate(const std::vector<double>& nu) const {
        std::vector<double> rate(n, 0.0);
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a);
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double k = h * 0.5 * (nu[a] + nu[b]);
                rate[a] += k; rate[b] += k;
            }
        }
        for (size_t e = 0; e < ea.size(); ++e) {
            const double k = edcoef[e] * 0.5 * (nu[ea[e]] + nu[eb[e]]);
            rate[ea[e]] += k; rate[eb[e]] += k;
        }
        double mx = 0.0;
        for (int64_t a = 0; a < n; ++a) if (active(a)) mx = std::max(mx, rate[a] / vol(a));
        return mx;
    }
    int diffusion_substeps(double dt) {
        if (opt.smagorinsky_cs > 0) {
            viscosity(nu_field);
            return std::max(1, static_cast<int>(std::ceil(dt * diffusion_rate(nu_field) / 0.7)));
        }
        if (!(opt.nu > 0)) return 0;
        if (defect == 3) return std::max(1, static_cast<int>(std::floor(dt * opt.nu * max_diff_coef / 0.7)));
        return std::max(1, static_cast<int>(std::ceil(dt * opt.nu * max_diff_coef / 0.7)));
    }
    std::vector<double> nu_field;
    // explicit diffusion; with Cs > 0 the face coefficient is (nu_a + nu_b)/2 of the field
    // computed once per step from the transported velocity (amr_advection.cu:1342-1344)
    void diffuse(double dt) {
        const int s = diffusion_substeps(dt);
        if (s == 0) return;
        const bool var = opt.smagorinsky_cs > 0;
        const double dts = dt / s;
        for (int it = 0; it < s; ++it) {
            std::vector<double> d[3] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
            for (int64_t a = 0; a < n; ++a) {
                if (!active(a)) continue;
                const double h = h_of(a);
                for (int axis = 0; axis < 3; ++axis) {
                    const int64_t b = snbr(a, axis, 1);
                    if (b < 0) continue;
                    const double k = dts * (var ? 0.5 * (nu_field[a] + nu_field[b]) : opt.nu) * (defect == 2 ? h * h : h);
                    for (int c = 0; c < 3; ++c) {
                        const double m = k * (comp(c, b) - comp(c, a));
                        d[c][a] += m; d[c][b] -= m;
                    }
                }
            }
            for (size_t e = 0; e < ea.size(); ++e) {
                const double k = dts * (var ? 0.5 * (nu_field[ea[e]] + nu_field[eb[e]]) : opt.nu) * edcoef[e];
                for (int c = 0; c < 3; ++c) {
                    const double m = k * (comp(c, eb[e]) - comp(c, ea[e]));
                    d[c][ea[e]] += m; d[c][eb[e]] -= m;
                }
            }
            apply_increments(d);
        }
    }
    void reconstruct() {
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b >= 0) face(axis, a) = 0.5 * (comp(axis, a) + comp(axis, b));
            }
            if (outlet && at_xmax(a)) fx[a] = vx[a];
        }
        // compact faces: the weighted mean plus the transverse correction both cells' gradients
        // support, minmod(g_a . s, g_b . s) with the skew s = x_f - ((1-w) c_a + w c_b)
        // (reconstruct_compact_cell_flux_kernel, amr_advection.cu:268-293)
        std::vector<double> g;
        velocity_gradients(g);
        for (size_t e = 0; e < ea.size(); ++e) {
            const int ax = eaxis[e];
            const double w = ew[e];
            const double* da = &edisp[6 * e];
            const double* db = da + 3;
            const double skew[3] = {(1.0 - w) * da[0] + w * db[0], (1.0 - w) * da[1] + w * db[1], (1.0 - w) * da[2] + w * db[2]};
            double corr[2] = {0.0, 0.0};
            for (int side = 0; side < 2; ++side) {
                const int sp = special_of[side ? eb[e] : ea[e]];
                if (sp < 0) continue;
                const double* gc = &g[9 * static_cast<size_t>(sp) + 3 * ax];
                corr[side] = gc[0] * skew[0] + gc[1] * skew[1] + gc[2] * skew[2];
            }
            fe[e] = (1.0 - w) * comp(ax, ea[e]) + w * comp(ax, eb[e]) + (defect == 8 ? 0.0 : minmod(corr[0], corr[1]));
        }
    }

    // ---- projection -----------------------------------------------------------------------
    void integrated_divergence(std::vector<double>& div) const {
        div.assign(n, 0.0);
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a), A = h * h;
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double f = A * (axis == 0 ? fx[a] : (axis == 1 ? fy[a] : fz[a]));
                div[a] += f; div[b] -= f;
            }
            if (at_xmin(a)) div[a] -= A * opt.speed;
            if (outlet && at_xmax(a)) div[a] += A * fx[a];
        }
        for (size_t e = 0; e < ea.size(); ++e) {
            const double f = earea[e] * fe[e];
            div[ea[e]] += f; div[eb[e]] -= f;
        }
    }
    void apply(const std::vector<double>& x, std::vector<double>& y) const {
        y.assign(n, 0.0);
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a);
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double f = h * (x[a] - x[b]);
                y[a] += f; y[b] -= f;
            }
            if (outlet && at_xmax(a)) y[a] += 2.0 * h * x[a];
        }
        for (size_t e = 0; e < ea.size(); ++e) {
            const double f = ecoef[e] * (x[ea[e]] - x[eb[e]]);
            y[ea[e]] += f; y[eb[e]] -= f;
        }
    }
    void composite_diagonal() {
        diag0.assign(n, 0.0);
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a);
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                diag0[a] += h; diag0[b] += h;
            }
            if (outlet && at_xmax(a)) diag0[a] += 2.0 * h;
        }
        for (size_t e = 0; e < ea.size(); ++e) { diag0[ea[e]] += ecoef[e]; diag0[eb[e]] += ecoef[e]; }
    }
    double dot(const std::vector<double>& a, const std::vector<double>& b) const {
        double s = 0;
        for (int64_t i = 0; i < static_cast<int64_t>(a.size()); ++i) s += a[i] * b[i];
        return s;
    }
    // coarse CSR level operator
    void apply_level(int m, const std::vector<double>& x, std::vector<double>& y) const {
        if (m == 0) { apply(x, y); return; }
        const Level& L = mg[m];
        y.assign(L.n, 0.0);
        for (int64_t r = 0; r < L.n; ++r) {
            double s = L.diag[r] * x[r];
            for (int q = L.row_off[r]; q < L.row_off[r + 1]; ++q) s -= L.coef[q] * x[L.col[q]];
            y[r] = s;
        }
    }
    const std::vector<double>& level_diag(int m) const { return m == 0 ? diag0 : mg[m].diag; }
    bool level_active(int m, int64_t r) const { return m == 0 ? active(r) : true; }
    // damped Jacobi sweeps; sweep s has weight omega for even s and omega2 for odd s
    void jacobi(int m, std::vector<double>& x, const std::vector<double>& b, int sweeps, double omega, bool zero_start,
                double omega2 = -1.0) const {
        const std::vector<double>& D = level_diag(m);
        const int64_t nn = static_cast<int64_t>(b.size());
        std::vector<double> Ax;
        for (int s = 0; s < sweeps; ++s) {
            const double w = (s & 1) && omega2 > 0.0 ? omega2 : omega;
            if (s == 0 && zero_start) {
                for (int64_t r = 0; r < nn; ++r) x[r] = (level_active(m, r) && D[r] > 0) ? w * b[r] / D[r] : 0.0;
                continue;
            }
            apply_level(m, x, Ax);
            for (int64_t r = 0; r < nn; ++r)
                if (level_active(m, r) && D[r] > 0) x[r] += w * (b[r] - Ax[r]) / D[r];
        }
    }
    void vcycle(int m, const std::vector<double>& b, std::vector<double>& x) const {
        const int64_t nn = static_cast<int64_t>(b.size());
        const int last = static_cast<int>(mg.size()) - 1;
        x.assign(nn, 0.0);
        if (m == last && m > 0) {
            if (!coarse_inv.empty()) {
                for (int64_t r = 0; r < nn; ++r) {
                    double s = 0;
                    for (int64_t c = 0; c < nn; ++c) s += coarse_inv[r * nn + c] * b[c];
                    x[r] = s;
                }
            } else {
                jacobi(m, x, b, 64, 2.0 / 3.0, true);  // fallback without a factorisation
            }
            return;
        }
        jacobi(m, x, b, opt.pre_smooth, opt.smoother_weight, true, opt.smoother_weight2);
        std::vector<double> Ax, r(nn), bc(mg[m + 1].n, 0.0), xc;
        apply_level(m, x, Ax);
        for (int64_t i = 0; i < nn; ++i) r[i] = level_active(m, i) ? b[i] - Ax[i] : 0.0;
        for (int64_t i = 0; i < nn; ++i) if (mg[m].parent[i] >= 0) bc[mg[m].parent[i]] += r[i];
        vcycle(m + 1, bc, xc);
        for (int64_t i = 0; i < nn; ++i) if (mg[m].parent[i] >= 0) x[i] += xc[mg[m].parent[i]];
        jacobi(m, x, b, opt.post_smooth, opt.smoother_weight, false, opt.smoother_weight2);
    }
    void precondition(const std::vector<double>& r, std::vector<double>& z) const {
        if (opt.preconditioner == PCFD_PC_JACOBI || mg.size() < 2) {
            z.assign(n, 0.0);
            for (int64_t a = 0; a < n; ++a) z[a] = (active(a) && diag0[a] > 0) ? r[a] / diag0[a] : 0.0;
            return;
        }
        vcycle(0, r, z);
    }
    struct SolveResult { int iterations; double relres; bool converged; };
    SolveResult pcg(const std::vector<double>& b, std::vector<double>& x, bool warm) const {
        SolveResult res{0, 0.0, false};
        const double bnorm = std::sqrt(dot(b, b));
        if (!warm) std::fill(x.begin(), x.end(), 0.0);
        if (!(bnorm > 0)) {
            std::fill(x.begin(), x.end(), 0.0);
            res.converged = true;
            return res;
        }
        std::vector<double> r(n), z, d, q;
        apply(x, q);
        for (int64_t a = 0; a < n; ++a) r[a] = active(a) ? b[a] - q[a] : 0.0;
        double rr = dot(r, r);
        res.relres = std::sqrt(rr) / bnorm;
        if (res.relres <= opt.rtol) { res.converged = true; return res; }
        precondition(r, z);
        double rz = dot(r, z);
        d = z;
        for (int it = 1; it <= opt.max_iterations; ++it) {
            apply(d, q);
            const double pAp = dot(d, q);
            if (!(pAp > 0) || !std::isfinite(pAp)) break;
            const double alpha = rz / pAp;
            for (int64_t a = 0; a < n; ++a) { x[a] += alpha * d[a]; r[a] -= alpha * q[a]; }
            rr = dot(r, r);
            res.iterations = it;
            res.relres = std::sqrt(rr) / bnorm;
            if (res.relres <= opt.rtol) { res.converged = true; break; }
            precondition(r, z);
            const double rz_new = dot(r, z);
            const double beta = rz_new / rz;
            rz = rz_new;
            for (int64_t a = 0; a < n; ++a) d[a] = z[a] + beta * d[a];
        }
        return res;
    }
    void correct(double dt) {
        const double s = dt / opt.rho;
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a);
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b >= 0) face(axis, a) -= s * (p[b] - p[a]) / h;
            }
            if (outlet && at_xmax(a)) fx[a] += s * p[a] / (defect == 5 ? h : 0.5 * h);
        }
        for (size_t e = 0; e < ea.size(); ++e) fe[e] -= s * (p[eb[e]] - p[ea[e]]) * (ecoef[e] / earea[e]);
    }
    double max_abs_divergence() const {
        std::vector<double> div;
        integrated_divergence(div);
        double m = 0;
        for (int64_t a = 0; a < n; ++a) if (active(a)) m = std::max(m, std::abs(div[a] / vol(a)));
        return m;
    }
    // RHS = -(rho/dt) * integrated divergence (normalize_divergence_rhs, amr_pressure.cu:255-284)
    void rhs(double dt, std::vector<double>& b) const {
        std::vector<double> div;
        integrated_divergence(div);
        b.assign(n, 0.0);
        const double scale = defect == 4 ? opt.rho : opt.rho / dt;
        for (int64_t a = 0; a < n; ++a) b[a] = active(a) ? -scale * div[a] : 0.0;
    }
    SolveResult project(double dt, bool warm, double* maxdiv) {
        std::vector<double> b;
        rhs(dt, b);
        SolveResult r = pcg(b, p, warm);
        correct(dt);
        if (maxdiv) *maxdiv = max_abs_divergence();
        return r;
    }
    // wall force of patch pch for the pressure p_a + g.(x - c_a) on the wall: p_a S + D^T g
    // (D = int_W (x - c_a) (x) n dA; for a planar patch with centroid x_p this is the
    // reference's (p_a + g.(x_p - c_a)) S, amr_advection.cu:391-399, aero_loads.cpp:38-40)
    void wall_force_p(int pch, const std::vector<double>& gp, double pref, double* F) const {
        const int64_t a = patch_dof[pch];
        const int sp = special_of[a];
        const double* g = &gp[3 * static_cast<size_t>(sp)];
        const double* D = &patch_D[9 * static_cast<size_t>(pch)];
        if (defect == 6) {
            double pw = p[a] - pref;
            for (int r = 0; r < 3; ++r) pw += g[r] * (patch_x[3 * pch + r] - centroid[3 * a + r]);
            for (int c = 0; c < 3; ++c) F[c] = pw * patch_S[3 * pch + c];
            return;
        }
        for (int c = 0; c < 3; ++c)
            F[c] = (p[a] - pref) * patch_S[3 * pch + c] + D[c] * g[0] + D[3 + c] * g[1] + D[6 + c] * g[2];
    }
    // Smooth-wall law (Spalding, kappa 0.41, B 5.2) inverted for the friction velocity by 36
    // bisections of u+ in [0, 80] with u+ y+(u+) = u_t y / nu (amr_pressure.h:41-53)
    static double friction_velocity(double ut, double y, double nu) {
        if (!(ut > 0.0) || !(y > 0.0) || !(nu > 0.0)) return 0.0;
        const double re = ut * y / nu, kappa = 0.41, c = std::exp(-kappa * 5.2);
        double lo = 0.0, hi = 80.0;
        for (int it = 0; it < 36; ++it) {
            const double up = 0.5 * (lo + hi), x = kappa * up, x2 = x * x;
            const double rem = x < 0.5 ? x2 * x2 * (1.0 / 24.0 + x * (1.0 / 120.0 + x * (1.0 / 720.0)))
                                       : std::exp(x) - 1.0 - x - 0.5 * x2 - x2 * x * (1.0 / 6.0);
            if (up * (up + c * rem) < re) lo = up;
            else hi = up;
        }
        return ut / std::max(0.5 * (lo + hi), 1e-12);
    }
    // wall shear, implicit per wall cell (amr_advection.cu:920-942): per patch k = A u_tau^2 /
    // |v_t| from the current velocity, (I + dt/V sum k (I - n n^T)) v' = v by Cholesky, patch
    // force rho k (v' - (v'.n) n).  Patches of a cell in patch order.
    std::vector<double> wall_force;
    bool wall_applied = false;
    void smooth_wall(double dt) {
        const size_t np = patch_dof.size();
        wall_force.assign(3 * np, 0.0);
        std::vector<int> order(np);
        for (size_t q = 0; q < np; ++q) order[q] = static_cast<int>(q);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return patch_dof[x] < patch_dof[y]; });
        std::vector<double> kp(np, 0.0);
        std::vector<double> dw[3] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
        for (size_t j0 = 0; j0 < np;) {
            const int64_t a = patch_dof[order[j0]];
            size_t j1 = j0;
            while (j1 < np && patch_dof[order[j1]] == a) ++j1;
            const double v[3] = {vx[a], vy[a], vz[a]};
            double K[6] = {0, 0, 0, 0, 0, 0};
            for (size_t j = j0; j < j1; ++j) {
                const int q = order[j];
                const double* S = &patch_S[3 * q];
                const double area = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
                const double nn[3] = {S[0] / area, S[1] / area, S[2] / area};
                double y = 0.0;
                for (int c = 0; c < 3; ++c) y += (centroid[3 * a + c] - patch_x[3 * q + c]) * S[c] / area;
                y = std::abs(y);
                // fluid centroid on the patch's plane: the slab estimate V / (2A)
                if (!(y > std::max(1e-14, std::cbrt(vol(a)) * 1e-10))) y = 0.5 * vol(a) / area;
                const double un = v[0] * nn[0] + v[1] * nn[1] + v[2] * nn[2];
                const double t[3] = {v[0] - un * nn[0], v[1] - un * nn[1], v[2] - un * nn[2]};
                const double speed = std::sqrt(std::max(0.0, t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
                double k = 0.0;
                if (speed > 0.0) {
                    const double ut = friction_velocity(speed, y, opt.nu);
                    k = area * ut * ut / speed;
                }
                kp[q] = k;
                K[0] += k * (1.0 - nn[0] * nn[0]); K[1] -= k * nn[0] * nn[1]; K[2] -= k * nn[0] * nn[2];
                K[3] += k * (1.0 - nn[1] * nn[1]); K[4] -= k * nn[1] * nn[2]; K[5] += k * (1.0 - nn[2] * nn[2]);
            }
            const double s = dt / vol(a);
            const double a00 = 1.0 + s * K[0], a01 = s * K[1], a02 = s * K[2], a11 = 1.0 + s * K[3], a12 = s * K[4],
                         a22 = 1.0 + s * K[5];
            const double l00 = std::sqrt(a00), l10 = a01 / l00, l20 = a02 / l00;
            const double l11 = std::sqrt(a11 - l10 * l10), l21 = (a12 - l20 * l10) / l11;
            const double l22 = std::sqrt(a22 - l20 * l20 - l21 * l21);
            const double y0 = v[0] / l00, y1 = (v[1] - l10 * y0) / l11, y2 = (v[2] - l20 * y0 - l21 * y1) / l22;
            const double x2 = y2 / l22, x1 = (y1 - l21 * x2) / l11, x0 = (y0 - l10 * x1 - l20 * x2) / l00;
            if (in_group(a)) {  // the group shares the increment (amr_advection.cu:1279-1293)
                dw[0][a] = vol(a) * (x0 - v[0]); dw[1][a] = vol(a) * (x1 - v[1]); dw[2][a] = vol(a) * (x2 - v[2]);
            } else {
                vx[a] = x0; vy[a] = x1; vz[a] = x2;
            }
            for (size_t j = j0; j < j1; ++j) {
                const int q = order[j];
                const double* S = &patch_S[3 * q];
                const double area = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
                const double nn[3] = {S[0] / area, S[1] / area, S[2] / area};
                const double un = x0 * nn[0] + x1 * nn[1] + x2 * nn[2];
                const double xv[3] = {x0, x1, x2};
                for (int c = 0; c < 3; ++c) wall_force[3 * q + c] = opt.rho * kp[q] * (xv[c] - un * nn[c]);
            }
            j0 = j1;
        }
        if (!roff.empty() && roff.size() > 1) {
            // members only: the others were set above and carry no increment here
            for (size_t g = 0; g + 1 < roff.size(); ++g) {
                double S[3] = {0, 0, 0};
                for (int k = roff[g]; k < roff[g + 1]; ++k)
                    for (int c = 0; c < 3; ++c) S[c] += dw[c][sdof[rmember[k]]];
                for (int k = roff[g]; k < roff[g + 1]; ++k) {
                    const int64_t a = sdof[rmember[k]];
                    const double iv = 1.0 / (defect == 13 ? vol(a) : rvolume[g]);
                    vx[a] += S[0] * iv; vy[a] += S[1] * iv; vz[a] += S[2] * iv;
                }
            }
        }
        wall_applied = true;
    }
    void impulse(double dt) {
        const double s = dt / opt.rho;
        std::vector<double> d[3] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            const double h = h_of(a), A = h * h;
            for (int axis = 0; axis < 3; ++axis) {
                const int64_t b = snbr(a, axis, 1);
                if (b < 0) continue;
                const double imp = s * A * 0.5 * (p[a] + p[b]);
                d[axis][a] -= imp; d[axis][b] += imp;
            }
            const int64_t q = a % BC;
            const int i = q & 7, j = (q >> 3) & 7, k = static_cast<int>(q >> 6);
            const uint8_t f = brick_flags[a / BC];
            const double bnd = s * A * p[a];
            if (i == 0 && (f & BF_XMIN)) d[0][a] += bnd;
            if (i == 7 && (f & BF_XMAX) && !outlet) d[0][a] -= bnd;
            if (j == 0 && (f & BF_YMIN)) d[1][a] += bnd;
            if (j == 7 && (f & BF_YMAX)) d[1][a] -= bnd;
            if (k == 0 && (f & BF_ZMIN)) d[2][a] += bnd;
            if (k == 7 && (f & BF_ZMAX)) d[2][a] -= bnd;
        }
        // compact faces: the weighted face pressure plus skew . ((1-w) g_a + w g_b)
        // (pressure_compact_cell_impulse_kernel, amr_advection.cu:376-390); walls: p S + D^T g
        std::vector<double> gp;
        pressure_gradients(gp);
        for (size_t e = 0; e < ea.size(); ++e) {
            const double w = ew[e];
            double pf = (1.0 - w) * p[ea[e]] + w * p[eb[e]];
            const int sa = special_of[ea[e]], sb = special_of[eb[e]];
            if (sa >= 0 && sb >= 0 && defect != 7) {
                const double* da = &edisp[6 * e];
                const double* db = da + 3;
                for (int c = 0; c < 3; ++c)
                    pf += ((1.0 - w) * da[c] + w * db[c]) * ((1.0 - w) * gp[3 * static_cast<size_t>(sa) + c] + w * gp[3 * static_cast<size_t>(sb) + c]);
            }
            const double imp = s * earea[e] * pf;
            d[eaxis[e]][ea[e]] -= imp; d[eaxis[e]][eb[e]] += imp;
        }
        for (size_t pch = 0; pch < patch_dof.size(); ++pch) {
            double F[3];
            wall_force_p(static_cast<int>(pch), gp, 0.0, F);
            if (defect == 12)
                for (int c = 0; c < 3; ++c) F[c] = p[patch_dof[pch]] * patch_S[3 * pch + c];
            for (int c = 0; c < 3; ++c) d[c][patch_dof[pch]] -= s * F[c];
        }
        apply_increments(d);
    }
    void loads(double pref, pcfd_loads* out) const {
        std::memset(out, 0, sizeof(*out));
        std::vector<double> gp;
        pressure_gradients(gp);
        for (size_t pch = 0; pch < patch_dof.size(); ++pch) {
            double Fp[3];
            wall_force_p(static_cast<int>(pch), gp, pref, Fp);
            double r[3];
            for (int c = 0; c < 3; ++c) r[c] = patch_x[3 * pch + c] - origin[c];
            for (int part = 0; part < 2; ++part) {
                double F[3];
                for (int c = 0; c < 3; ++c)
                    F[c] = part == 0 ? Fp[c] : (wall_force.empty() ? 0.0 : wall_force[3 * pch + c]);
                double* f = part == 0 ? out->force : out->viscous_force;
                double* m = part == 0 ? out->moment : out->viscous_moment;
                for (int c = 0; c < 3; ++c) f[c] += F[c];
                m[0] += r[1] * F[2] - r[2] * F[1];
                m[1] += r[2] * F[0] - r[0] * F[2];
                m[2] += r[0] * F[1] - r[1] * F[0];
            }
        }
        out->viscous_valid = wall_applied ? 1 : 0;
        for (int c = 0; c < 3; ++c) {
            out->total_force[c] = out->force[c] + out->viscous_force[c];
            out->total_moment[c] = out->moment[c] + out->viscous_moment[c];
        }
        out->drag = out->force[0]; out->side = out->force[1]; out->lift = out->force[2];
        out->viscous_drag = out->viscous_force[0]; out->viscous_side = out->viscous_force[1]; out->viscous_lift = out->viscous_force[2];
        out->total_drag = out->total_force[0]; out->total_side = out->total_force[1]; out->total_lift = out->total_force[2];
        const double q = 0.5 * opt.rho * opt.speed * opt.speed;
        if (q > 0 && ref_area > 0) {
            const double qs = q * ref_area;
            out->coefficients_valid = 1;
            out->cd = out->drag / qs; out->cs = out->side / qs; out->cl = out->lift / qs;
            out->cd_viscous = out->viscous_drag / qs; out->cs_viscous = out->viscous_side / qs; out->cl_viscous = out->viscous_lift / qs;
            out->total_cd = out->total_drag / qs; out->total_cs = out->total_side / qs; out->total_cl = out->total_lift / qs;
        }
    }
    // max |v| over active cells; a non-finite component makes it infinite (std::max alone would
    // drop a NaN)
    double max_abs_velocity() const {
        double m = 0;
        for (int64_t a = 0; a < n; ++a) {
            if (!active(a)) continue;
            if (!std::isfinite(vx[a]) || !std::isfinite(vy[a]) || !std::isfinite(vz[a])) return INFINITY;
            m = std::max({m, std::abs(vx[a]), std::abs(vy[a]), std::abs(vz[a])});
        }
        return m;
    }
};

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

void* oracle_create(const oracle_topology* t, const pcfd_options* o) {
    try {
        auto* s = new Oracle();
        s->nb = t->nbricks; s->n = t->ndof; s->levels = t->levels; s->outlet = t->outlet != 0;
        s->level_h.assign(t->level_h, t->level_h + t->levels);
        s->brick_level.assign(t->brick_level, t->brick_level + t->nbricks);
        s->brick_flags.assign(t->brick_flags, t->brick_flags + t->nbricks);
        s->brick_nbr.assign(t->brick_nbr, t->brick_nbr + 6 * t->nbricks);
        s->mask.assign(t->mask, t->mask + t->ndof);
        s->special_of.assign(t->ndof, -1);
        for (int64_t i = 0; i < t->nspecial; ++i) s->special_of[t->special_dof[i]] = static_cast<int>(i);
        s->sdof.assign(t->special_dof, t->special_dof + t->nspecial);
        s->special_volume.assign(t->special_volume, t->special_volume + t->nspecial);
        s->special_patch.assign(t->special_patch, t->special_patch + t->nspecial);
        s->ea.assign(t->edge_a, t->edge_a + t->nedges);
        s->eb.assign(t->edge_b, t->edge_b + t->nedges);
        s->eaxis.assign(t->edge_axis, t->edge_axis + t->nedges);
        s->earea.assign(t->edge_area, t->edge_area + t->nedges);
        s->ecoef.assign(t->edge_coef, t->edge_coef + t->nedges);
        s->ew.assign(t->edge_w, t->edge_w + t->nedges);
        s->patch_dof.assign(t->patch_dof, t->patch_dof + t->npatches);
        s->patch_S.assign(t->patch_S, t->patch_S + 3 * t->npatches);
        s->patch_x.assign(t->patch_x, t->patch_x + 3 * t->npatches);
        s->patch_D.assign(t->patch_D, t->patch_D + 9 * t->npatches);
        s->edcoef.assign(t->edge_dcoef, t->edge_dcoef + t->nedges);
        if (t->ngroups > 0) {
            s->rgroup.assign(t->redist_group, t->redist_group + t->nspecial);
            s->roff.assign(t->redist_off, t->redist_off + t->ngroups + 1);
            s->rmember.assign(t->redist_member, t->redist_member + t->redist_off[t->ngroups]);
            s->rvolume.assign(t->redist_volume, t->redist_volume + t->ngroups);
        }
        s->finest_h = t->finest_h;
        s->max_diff_coef = t->max_diffusion_coefficient;
        if (!t->centroid) throw std::invalid_argument("the oracle needs the cell centroids");
        s->centroid.assign(t->centroid, t->centroid + 3 * t->ndof);
        s->edisp.assign(6 * t->nedges, 0.0);
        for (int64_t e = 0; e < t->nedges; ++e)
            for (int c = 0; c < 3; ++c) {
                s->edisp[6 * e + c] = t->edge_x[3 * e + c] - s->centroid[3 * static_cast<int64_t>(t->edge_a[e]) + c];
                s->edisp[6 * e + 3 + c] = t->edge_x[3 * e + c] - s->centroid[3 * static_cast<int64_t>(t->edge_b[e]) + c];
            }
        s->build_gradient_stencils();
        s->mg.resize(t->mg_levels);
        for (int m = 0; m < t->mg_levels; ++m) {
            auto& L = s->mg[m];
            L.n = t->mg_n[m];
            if (m + 1 < t->mg_levels) L.parent.assign(t->mg_parent[m], t->mg_parent[m] + L.n);
            if (m >= 1) {
                L.row_off.assign(t->mg_row_off[m], t->mg_row_off[m] + L.n + 1);
                L.col.assign(t->mg_col[m], t->mg_col[m] + L.row_off.back());
                L.coef.assign(t->mg_coef[m], t->mg_coef[m] + L.row_off.back());
                L.diag.assign(t->mg_diag[m], t->mg_diag[m] + L.n);
            }
        }
        if (t->coarse_inv && t->mg_levels > 1) {
            const int64_t nc = t->mg_n[t->mg_levels - 1];
            s->coarse_inv.assign(t->coarse_inv, t->coarse_inv + nc * nc);
        }
        s->opt = *o;
        const int64_t n = s->n;
        s->vx.assign(n, 0); s->vy.assign(n, 0); s->vz.assign(n, 0);
        s->fx.assign(n, 0); s->fy.assign(n, 0); s->fz.assign(n, 0);
        s->fe.assign(t->nedges, 0); s->p.assign(n, 0);
        s->composite_diagonal();
        return s;
    } catch (const std::exception& e) {
        g_error = e.what();
        return nullptr;
    }
}

void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }

void oracle_set_reference(void* h, double area, const double origin[3]) {
    auto* s = static_cast<Oracle*>(h);
    s->ref_area = area;
    for (int c = 0; c < 3; ++c) s->origin[c] = origin[c];
}

static void fill_stats(Oracle* s, pcfd_step_stats* st, double dt, const Oracle::SolveResult& r, double maxdiv, int subs) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->dt = dt; st->iterations = r.iterations; st->converged = r.converged; st->relative_residual = r.relres;
    st->max_abs_divergence = maxdiv; st->diffusion_substeps = subs; st->physical_time = s->time;
    st->max_abs_velocity = s->max_abs_velocity();
}

int oracle_initialize(void* h, pcfd_step_stats* st) {
    auto* s = static_cast<Oracle*>(h);
    for (int64_t a = 0; a < s->n; ++a) {
        const bool act = s->active(a);
        s->vx[a] = act ? s->opt.speed : 0.0; s->vy[a] = 0; s->vz[a] = 0;
    }
    s->reconstruct();
    const double dt = s->opt.cfl * s->finest_h / std::max(1e-9, std::abs(s->opt.speed));
    double md = 0;
    std::fill(s->p.begin(), s->p.end(), 0.0);
    auto r = s->project(dt, false, &md);
    if (r.converged) s->impulse(dt);
    s->initialized = r.converged;
    fill_stats(s, st, dt, r, md, 0);
    return PCFD_OK;
}

int oracle_advect(void* h, double dt, double* dt_used, int* substeps) {
    auto* s = static_cast<Oracle*>(h);
    if (!(dt > 0)) {
        const double rate = std::max({1e-9, std::abs(s->opt.speed) / s->finest_h, s->outflow_rate()});
        dt = s->opt.cfl / rate;
    }
    s->advect(dt);
    const int subs = s->diffusion_substeps(dt);
    s->diffuse(dt);
    if (s->opt.smooth_wall) s->smooth_wall(dt);
    s->reconstruct();
    if (dt_used) *dt_used = dt;
    if (substeps) *substeps = subs;
    return PCFD_OK;
}

int oracle_project(void* h, double dt, int warm, pcfd_step_stats* st) {
    auto* s = static_cast<Oracle*>(h);
    double md = 0;
    auto r = s->project(dt, warm != 0, &md);
    fill_stats(s, st, dt, r, md, 0);
    return PCFD_OK;
}

int oracle_impulse(void* h, double dt) { static_cast<Oracle*>(h)->impulse(dt); return PCFD_OK; }

int oracle_step(void* h, pcfd_step_stats* st) {
    auto* s = static_cast<Oracle*>(h);
    if (!s->initialized) { g_error = "step before a converged initialization"; return PCFD_ESTATE; }
    const double vmax = s->max_abs_velocity();
    if (!std::isfinite(vmax)) { g_error = "non-finite velocity before timestep"; return PCFD_ENONFINITE; }
    const double rate = s->outflow_rate();
    const double dt = s->opt.cfl / std::max({1e-9, std::abs(s->opt.speed) / s->finest_h, rate});
    s->advect(dt);
    const int subs = s->diffusion_substeps(dt);
    s->diffuse(dt);
    if (s->opt.smooth_wall) s->smooth_wall(dt);
    s->reconstruct();
    double md = 0;
    auto r = s->project(dt, true, &md);
    if (r.converged) { s->impulse(dt); s->time += dt; }
    fill_stats(s, st, dt, r, md, subs);
    if (st) st->cfl_rate = rate;
    return PCFD_OK;
}

int oracle_loads(void* h, double pref, pcfd_loads* out) { static_cast<Oracle*>(h)->loads(pref, out); return PCFD_OK; }

// the smooth-wall pieces on their own, for the tests that pin them
double oracle_wall_friction(double ut, double y, double nu) { return Oracle::friction_velocity(ut, y, nu); }
int oracle_smooth_wall(void* h, double dt, double* force) {
    auto* s = static_cast<Oracle*>(h);
    if (s->centroid.empty()) { g_error = "the smooth-wall model needs cell centroids"; return PCFD_EINVAL; }
    s->smooth_wall(dt);
    if (force) std::copy(s->wall_force.begin(), s->wall_force.end(), force);
    return PCFD_OK;
}

int oracle_get_state(void* h, double* x, double* y, double* z, double* p) {
    auto* s = static_cast<Oracle*>(h);
    const size_t n = s->n * sizeof(double);
    if (x) std::memcpy(x, s->vx.data(), n);
    if (y) std::memcpy(y, s->vy.data(), n);
    if (z) std::memcpy(z, s->vz.data(), n);
    if (p) std::memcpy(p, s->p.data(), n);
    return PCFD_OK;
}
int oracle_set_state(void* h, const double* x, const double* y, const double* z, const double* p) {
    auto* s = static_cast<Oracle*>(h);
    const size_t n = s->n * sizeof(double);
    if (x) std::memcpy(s->vx.data(), x, n);
    if (y) std::memcpy(s->vy.data(), y, n);
    if (z) std::memcpy(s->vz.data(), z, n);
    if (p) std::memcpy(s->p.data(), p, n);
    return PCFD_OK;
}
int oracle_get_faces(void* h, double* fx, double* fy, double* fz, double* fe) {
    auto* s = static_cast<Oracle*>(h);
    const size_t n = s->n * sizeof(double);
    if (fx) std::memcpy(fx, s->fx.data(), n);
    if (fy) std::memcpy(fy, s->fy.data(), n);
    if (fz) std::memcpy(fz, s->fz.data(), n);
    if (fe) std::memcpy(fe, s->fe.data(), s->fe.size() * sizeof(double));
    return PCFD_OK;
}
int oracle_set_faces(void* h, const double* fx, const double* fy, const double* fz, const double* fe) {
    auto* s = static_cast<Oracle*>(h);
    const size_t n = s->n * sizeof(double);
    if (fx) std::memcpy(s->fx.data(), fx, n);
    if (fy) std::memcpy(s->fy.data(), fy, n);
    if (fz) std::memcpy(s->fz.data(), fz, n);
    if (fe) std::memcpy(s->fe.data(), fe, s->fe.size() * sizeof(double));
    return PCFD_OK;
}
int oracle_mark_initialized(void* h, double time) {
    auto* s = static_cast<Oracle*>(h);
    s->initialized = true; s->time = time;
    return PCFD_OK;
}
/* op: 0 apply, 1 residual b - A x, 2 Jacobi sweeps (x in/out), 3 V-cycle on level `level` */
int oracle_mg_primitive(void* h, int op, int level, const double* x, const double* b, double* out, double omega, int sweeps) {
    auto* s = static_cast<Oracle*>(h);
    if (level < 0 || level >= static_cast<int>(s->mg.size())) { g_error = "bad level"; return PCFD_EINVAL; }
    const int64_t nn = level == 0 ? s->n : s->mg[level].n;
    std::vector<double> vx(x ? x : out, (x ? x : out) + nn), vb, y;
    if (b) vb.assign(b, b + nn);
    switch (op) {
        case 0: s->apply_level(level, vx, y); break;
        case 1: s->apply_level(level, vx, y); for (int64_t i = 0; i < nn; ++i) y[i] = s->level_active(level, i) ? vb[i] - y[i] : 0.0; break;
        case 2: s->jacobi(level, vx, vb, sweeps, omega, false); y = vx; break;
        case 3: s->vcycle(level, vb, y); break;
        default: g_error = "bad op"; return PCFD_EINVAL;
    }
    std::memcpy(out, y.data(), nn * sizeof(double));
    return PCFD_OK;
}
/* Dense assembly of the composite operator for brute-force checks on tiny grids. */
int oracle_dense_operator(void* h, double* A) {
    auto* s = static_cast<Oracle*>(h);
    const int64_t n = s->n;
    std::vector<double> e(n, 0.0), y;
    for (int64_t c = 0; c < n; ++c) {
        e[c] = 1.0;
        s->apply(e, y);
        for (int64_t r = 0; r < n; ++r) A[r * n + c] = y[r];
        e[c] = 0.0;
    }
    return PCFD_OK;
}
double oracle_outflow_rate(void* h) { return static_cast<Oracle*>(h)->outflow_rate(); }
/* Cell viscosity nu_mol + (Cs V^(1/3))^2 |S| of the current velocity (0 on inactive cells). */
int oracle_viscosity(void* h, double* nu) {
    try {
        auto* s = static_cast<Oracle*>(h);
        std::vector<double> v;
        s->viscosity(v);
        std::memcpy(nu, v.data(), v.size() * sizeof(double));
        return PCFD_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return PCFD_EINVAL;
    }
}
/* Plant mistake `defect` (see Oracle::defect; 0 removes it) for the contract tests. */
int oracle_plant(void* h, int defect) {
    auto* s = static_cast<Oracle*>(h);
    s->defect = defect;
    s->build_gradient_stencils();
    return PCFD_OK;
}
/* Face correction u -= (dt/rho) dp/dn with the current pressure (no solve). */
int oracle_correct(void* h, double dt) { static_cast<Oracle*>(h)->correct(dt); return PCFD_OK; }
/* The projection's right-hand side -(rho/dt) * integrated divergence of the current faces. */
int oracle_rhs(void* h, double dt, double* b) {
    auto* s = static_cast<Oracle*>(h);
    std::vector<double> v;
    s->rhs(dt, v);
    std::memcpy(b, v.data(), v.size() * sizeof(double));
    return PCFD_OK;
}
/* Least-squares gradients at the special cells of the current velocity (9 per cell) and
 * pressure (3 per cell). */
int oracle_gradients(void* h, double* gv, double* gp) {
    auto* s = static_cast<Oracle*>(h);
    std::vector<double> g;
    if (gv) { s->velocity_gradients(g); std::memcpy(gv, g.data(), g.size() * sizeof(double)); }
    if (gp) { s->pressure_gradients(g); std::memcpy(gp, g.data(), g.size() * sizeof(double)); }
    return PCFD_OK;
}
int oracle_divergence(void* h, double* div) {
    auto* s = static_cast<Oracle*>(h);
    std::vector<double> d;
    s->integrated_divergence(d);
    std::memcpy(div, d.data(), d.size() * sizeof(double));
    return PCFD_OK;
}
/* Time `steps` oracle steps; returns seconds (std::chrono::steady_clock). */
double oracle_time_steps(void* h, int steps) {
    auto* s = static_cast<Oracle*>(h);
    const auto t0 = std::chrono::steady_clock::now();
    pcfd_step_stats st;
    for (int i = 0; i < steps; ++i) oracle_step(s, &st);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
