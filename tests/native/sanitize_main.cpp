// Host-side sanitizer driver (test infrastructure, built by tests/test_sanitizers.py with
// AddressSanitizer and UndefinedBehaviorSanitizer): builds topologies through the topology C ABI
// (AMR bricks, cut cells, compact edges, wall moments, multigrid, dense coarse inverse, merged
// batches, the non-orthogonal records) and runs the FP64 oracle on them, so that every host
// code path the GPU solver's input depends on runs under both sanitizers.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../oracle/oracle.cpp"

namespace {

void need(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "sanitize driver: %s failed: %s\n", what, pcfd_last_error());
        std::exit(2);
    }
}

template <class T>
std::vector<T> exported(const pcfd_topology* t, int id) {
    const int64_t n = pcfd_topology_array_size(t, id);
    std::vector<T> v(n > 0 ? static_cast<size_t>(n) : 0);
    if (n > 0) need(pcfd_topology_export(t, id, v.data(), v.size() * sizeof(T)) == PCFD_OK, "export");
    return v;
}

pcfd_topology* build(const pcfd_geometry_desc& g, const pcfd_grid_desc& d) {
    pcfd_topology* t = nullptr;
    need(pcfd_topology_build(&g, &d, &t) == PCFD_OK, "pcfd_topology_build");
    return t;
}

pcfd_grid_desc grid(double lo0, double lo1, double lo2, double hi0, double hi1, double hi2, int cells_x, int levels,
                    double band0, double band1) {
    pcfd_grid_desc d{};
    d.lo[0] = lo0; d.lo[1] = lo1; d.lo[2] = lo2;
    d.hi[0] = hi0; d.hi[1] = hi1; d.hi[2] = hi2;
    d.base_cells_x = cells_x;
    d.levels = levels;
    d.refine_band[0] = band0;
    d.refine_band[1] = band1;
    d.min_volume_fraction = 0.1;
    d.merge_volume_fraction = 0.25;
    d.subsamples = 4;
    d.outlet_xmax = 1;
    return d;
}

// the oracle on a topology: initialize and `steps` steps, then the loads
void run_oracle(const pcfd_topology* t, int steps, double cs, int smooth_wall) {
    pcfd_topology_info info{};
    need(pcfd_topology_get_info(t, &info) == PCFD_OK, "get_info");
    auto level_h = exported<double>(t, PCFD_ARR_LEVEL_H);
    auto brick_level = exported<int>(t, PCFD_ARR_BRICK_LEVEL);
    auto brick_flags = exported<uint8_t>(t, PCFD_ARR_BRICK_FLAGS);
    auto brick_nbr = exported<int>(t, PCFD_ARR_BRICK_NBR);
    auto mask = exported<uint8_t>(t, PCFD_ARR_MASK);
    auto special_dof = exported<int>(t, PCFD_ARR_SPECIAL_DOF);
    auto special_volume = exported<double>(t, PCFD_ARR_SPECIAL_VOLUME);
    auto special_patch = exported<int>(t, PCFD_ARR_SPECIAL_PATCH);
    auto edge_a = exported<int>(t, PCFD_ARR_EDGE_A);
    auto edge_b = exported<int>(t, PCFD_ARR_EDGE_B);
    auto edge_axis = exported<int8_t>(t, PCFD_ARR_EDGE_AXIS);
    auto edge_area = exported<double>(t, PCFD_ARR_EDGE_AREA);
    auto edge_coef = exported<double>(t, PCFD_ARR_EDGE_COEF);
    auto edge_w = exported<double>(t, PCFD_ARR_EDGE_W);
    auto edge_x = exported<double>(t, PCFD_ARR_EDGE_X);
    auto edge_dcoef = exported<double>(t, PCFD_ARR_EDGE_DCOEF);
    auto patch_dof = exported<int>(t, PCFD_ARR_PATCH_DOF);
    auto patch_S = exported<double>(t, PCFD_ARR_PATCH_S);
    auto patch_x = exported<double>(t, PCFD_ARR_PATCH_X);
    auto patch_D = exported<double>(t, PCFD_ARR_PATCH_D);
    auto centroid = exported<double>(t, PCFD_ARR_CELL_CENTROID);
    auto rgroup = exported<int>(t, PCFD_ARR_REDIST_GROUP);
    auto roff = exported<int>(t, PCFD_ARR_REDIST_OFF);
    auto rmember = exported<int>(t, PCFD_ARR_REDIST_MEMBER);
    auto rvolume = exported<double>(t, PCFD_ARR_REDIST_VOLUME);
    const int nl = static_cast<int>(info.mg_levels);
    std::vector<std::vector<int>> parent(nl), row_off(nl), col(nl);
    std::vector<std::vector<double>> coef(nl), diag(nl);
    std::vector<int64_t> mg_n(nl);
    std::vector<const int*> pp(nl), pr(nl), pc(nl);
    std::vector<const double*> pcf(nl), pd(nl);
    for (int m = 0; m < nl; ++m) {
        parent[m] = exported<int>(t, PCFD_ARR_MG_PARENT + 8 * m);
        if (m >= 1) {
            row_off[m] = exported<int>(t, PCFD_ARR_MG_ROW_OFF + 8 * m);
            col[m] = exported<int>(t, PCFD_ARR_MG_COL + 8 * m);
            coef[m] = exported<double>(t, PCFD_ARR_MG_COEF + 8 * m);
            diag[m] = exported<double>(t, PCFD_ARR_MG_DIAG + 8 * m);
        }
        mg_n[m] = m == 0 ? info.dofs : static_cast<int64_t>(diag[m].size());
        pp[m] = parent[m].data(); pr[m] = row_off[m].data(); pc[m] = col[m].data();
        pcf[m] = coef[m].data(); pd[m] = diag[m].data();
    }
    auto cinv = exported<double>(t, PCFD_ARR_MG_COARSE_INV + 8 * (nl - 1));
    oracle_topology ot{};
    ot.nbricks = info.bricks; ot.ndof = info.dofs; ot.nspecial = info.special_cells; ot.nedges = info.compact_edges;
    ot.npatches = info.patches; ot.levels = info.levels; ot.outlet = info.outlet;
    ot.level_h = level_h.data(); ot.brick_level = brick_level.data(); ot.brick_flags = brick_flags.data();
    ot.brick_nbr = brick_nbr.data(); ot.mask = mask.data(); ot.special_dof = special_dof.data();
    ot.special_volume = special_volume.data(); ot.special_patch = special_patch.data();
    ot.edge_a = edge_a.data(); ot.edge_b = edge_b.data(); ot.edge_axis = edge_axis.data(); ot.edge_area = edge_area.data();
    ot.edge_coef = edge_coef.data(); ot.edge_w = edge_w.data(); ot.patch_dof = patch_dof.data();
    ot.patch_S = patch_S.data(); ot.patch_x = patch_x.data(); ot.patch_D = patch_D.data(); ot.edge_x = edge_x.data();
    ot.edge_dcoef = edge_dcoef.data(); ot.finest_h = level_h.back(); ot.max_diffusion_coefficient = info.max_diffusion_coefficient;
    ot.mg_levels = nl; ot.mg_n = mg_n.data(); ot.mg_parent = pp.data(); ot.mg_row_off = pr.data(); ot.mg_col = pc.data();
    ot.mg_coef = pcf.data(); ot.mg_diag = pd.data(); ot.coarse_inv = cinv.empty() ? nullptr : cinv.data();
    ot.centroid = centroid.data();
    ot.ngroups = static_cast<int64_t>(rvolume.size());
    ot.redist_group = rgroup.data(); ot.redist_off = roff.data(); ot.redist_member = rmember.data();
    ot.redist_volume = rvolume.data();
    pcfd_options o{};
    o.rho = 1.0; o.nu = 1e-3; o.speed = 1.0; o.cfl = 0.5; o.rtol = 1e-8; o.max_iterations = 200;
    o.pre_smooth = 2; o.post_smooth = 2; o.check_interval = 2;
    o.smoother_weight = 1.0 / (1.2 + 0.8 / std::sqrt(2.0));
    o.smoother_weight2 = 1.0 / (1.2 - 0.8 / std::sqrt(2.0));
    o.smagorinsky_cs = cs;
    o.smooth_wall = smooth_wall;
    void* h = oracle_create(&ot, &o);
    need(h != nullptr, "oracle_create");
    const double origin[3] = {0, 0, 0};
    oracle_set_reference(h, info.reference_area, origin);
    pcfd_step_stats st{};
    need(oracle_initialize(h, &st) == PCFD_OK && st.converged, "oracle_initialize");
    for (int s = 0; s < steps; ++s) need(oracle_step(h, &st) == PCFD_OK && st.converged, "oracle_step");
    pcfd_loads ld{};
    need(oracle_loads(h, 0.0, &ld) == PCFD_OK && std::isfinite(ld.force[0]), "oracle_loads");
    std::vector<double> gv(9 * special_dof.size() + 1), gp(3 * special_dof.size() + 1);
    need(oracle_gradients(h, gv.data(), gp.data()) == PCFD_OK, "oracle_gradients");
    oracle_destroy(h);
    std::printf("oracle: %lld cells, %d steps, drag %.6g\n", static_cast<long long>(info.active_cells), steps, ld.drag);
}

}  // namespace

int main() {
    pcfd_geometry_desc sphere{};
    sphere.kind = PCFD_GEOM_SPHERE;
    sphere.center[0] = 1.0; sphere.center[1] = 1.0; sphere.center[2] = 1.0;
    sphere.radius = 0.25;
    pcfd_topology* s2 = build(sphere, grid(0, 0, 0, 4, 2, 2, 32, 2, 0.3, 0.15));
    need(pcfd_topology_build_nonorthogonal(s2) == PCFD_OK, "pcfd_topology_build_nonorthogonal");
    run_oracle(s2, 2, 0.17, 1);

    pcfd_geometry_desc wing{};
    wing.kind = PCFD_GEOM_NACA4_WING;
    wing.naca_digits = 2412; wing.chord = 1.0; wing.span = 2.03; wing.alpha_deg = 4.0;
    pcfd_topology* w3 = build(wing, grid(-3, -3, -2, 6, 3, 3.03, 24, 3, 1.0, 0.25));
    run_oracle(w3, 1, 0.0, 0);
    wing.naca_digits = 4412;
    wing.alpha_deg = -4.0;
    pcfd_topology* wa = build(wing, grid(-3, -3, -2, 6, 3, 3.03, 24, 2, 0.3, 0.0));
    wing.alpha_deg = 10.4;
    pcfd_topology* wb = build(wing, grid(-3, -3, -2, 6, 3, 3.03, 24, 2, 0.3, 0.0));
    const pcfd_topology* parts[2] = {wa, wb};
    pcfd_topology* merged = nullptr;
    need(pcfd_topology_merge(parts, 2, &merged) == PCFD_OK, "pcfd_topology_merge");
    need(pcfd_topology_build_nonorthogonal(merged) == PCFD_EINVAL, "refusal of a batch's non-orthogonal records");

    pcfd_geometry_desc glider{};
    glider.kind = PCFD_GEOM_PARAGLIDER;
    glider.naca_digits = 2412; glider.span = 8.0; glider.alpha_deg = 6.0; glider.sweep_deg = 8.0;
    glider.anhedral_deg = 25.0; glider.reflex = 0.02;
    const double c0 = 4.0 * 8.0 / (5.0 * M_PI), half = 25.0 * M_PI / 180.0, drop = (4.0 / half) * (1 - std::cos(half));
    pcfd_topology* g2 = build(glider, grid(-3 * c0, -6.0, -drop - 2.5 * c0, 6 * c0, 6.0, 2.5 * c0, 40, 2, c0, 0.0));
    run_oracle(g2, 1, 0.0, 0);

    for (pcfd_topology* t : {s2, w3, wa, wb, merged, g2}) pcfd_topology_free(t);
    std::printf("sanitize ok\n");
    return 0;
}
