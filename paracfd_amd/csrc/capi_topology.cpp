// C ABI: error reporting and topology construction / export.
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paracfd.h"
#include "topology.hpp"

namespace pcfd {
thread_local std::string g_last_error;
void set_error(const std::string& m) { g_last_error = m; }
}  // namespace pcfd

struct pcfd_topology {
    pcfd::Topology t;
};

namespace pcfd {
pcfd_status build_topology_handle(const char* where, const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid,
                                  BuildAccelerator* accel, pcfd_topology** out) {
    if (!geometry || !grid || !out) { set_error(std::string(where) + ": null argument"); return PCFD_EINVAL; }
    *out = nullptr;
    try {
        *out = new pcfd_topology{build_topology(*geometry, *grid, accel)};
        return PCFD_OK;
    } catch (const BuildError& e) {
        set_error(e.what());
        return PCFD_EINVAL;
    } catch (const DeviceError& e) {
        set_error(std::string(where) + ": " + e.what());
        return PCFD_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error(std::string(where) + ": out of host memory");
        return PCFD_ENOMEM;
    } catch (const std::exception& e) {
        set_error(std::string(where) + ": " + e.what());
        return PCFD_EINVAL;
    }
}
}  // namespace pcfd

extern "C" {

const char* pcfd_last_error(void) { return pcfd::g_last_error.c_str(); }
int pcfd_abi_version(void) { return PCFD_ABI_VERSION; }

pcfd_status pcfd_topology_build_nonorthogonal(pcfd_topology* topology) {
    if (!topology) { pcfd::set_error("pcfd_topology_build_nonorthogonal: null topology"); return PCFD_EINVAL; }
    try {
        pcfd::build_nonorthogonal(topology->t);
        return PCFD_OK;
    } catch (const pcfd::BuildError& e) {
        pcfd::set_error(e.what());
        return PCFD_EINVAL;
    } catch (const std::bad_alloc&) {
        pcfd::set_error("pcfd_topology_build_nonorthogonal: out of host memory");
        return PCFD_ENOMEM;
    } catch (const std::exception& e) {
        pcfd::set_error(std::string("pcfd_topology_build_nonorthogonal: ") + e.what());
        return PCFD_EINVAL;
    }
}

pcfd_status pcfd_topology_build(const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid, pcfd_topology** out) {
    return pcfd::build_topology_handle("pcfd_topology_build", geometry, grid, nullptr, out);
}

pcfd_status pcfd_topology_build_times(const pcfd_topology* topology, double ms[9], int* device) {
    if (!topology || !ms || !device) { pcfd::set_error("pcfd_topology_build_times: null argument"); return PCFD_EINVAL; }
    for (int k = 0; k < 9; ++k) ms[k] = topology->t.build_ms[k];
    *device = topology->t.sampled_on_device;
    return PCFD_OK;
}

void pcfd_topology_free(pcfd_topology* topology) { delete topology; }

pcfd_status pcfd_topology_get_info(const pcfd_topology* topology, pcfd_topology_info* info) {
    if (!topology || !info) { pcfd::set_error("pcfd_topology_get_info: null argument"); return PCFD_EINVAL; }
    const pcfd::Topology& t = topology->t;
    std::memset(info, 0, sizeof(*info));
    info->bricks = t.nbricks;
    info->dofs = t.ndof;
    info->active_cells = t.n_active;
    info->special_cells = static_cast<int64_t>(t.special_dof.size());
    info->compact_edges = static_cast<int64_t>(t.edge_a.size());
    info->patches = static_cast<int64_t>(t.patch_dof.size());
    info->mg_levels = static_cast<int64_t>(t.mg.size());
    for (size_t m = 0; m < t.mg.size() && m < 12; ++m) info->mg_cells[m] = m == 0 ? t.n_active : t.mg[m].n;
    info->levels = t.levels;
    info->h0 = t.h0;
    for (int a = 0; a < 3; ++a) { info->lo[a] = t.lo[a]; info->base_bricks[a] = t.base_bricks[a]; }
    info->body_volume = t.body_volume;
    info->reference_area = t.reference_area;
    info->finest_h = t.finest_h;
    info->max_diffusion_coefficient = t.max_diffusion_coefficient;
    info->outlet = t.outlet ? 1 : 0;
    info->instances = t.instances();
    return PCFD_OK;
}

pcfd_status pcfd_topology_merge(const pcfd_topology* const* parts, int count, pcfd_topology** out) {
    if (!parts || count < 1 || !out) { pcfd::set_error("pcfd_topology_merge: invalid argument"); return PCFD_EINVAL; }
    *out = nullptr;
    try {
        std::vector<const pcfd::Topology*> v;
        for (int i = 0; i < count; ++i) {
            if (!parts[i]) { pcfd::set_error("pcfd_topology_merge: null part"); return PCFD_EINVAL; }
            v.push_back(&parts[i]->t);
        }
        *out = new pcfd_topology{pcfd::merge_topologies(v)};
        return PCFD_OK;
    } catch (const pcfd::BuildError& e) {
        pcfd::set_error(std::string("pcfd_topology_merge: ") + e.what());
        return PCFD_EINVAL;
    } catch (const std::bad_alloc&) {
        pcfd::set_error("pcfd_topology_merge: out of host memory");
        return PCFD_ENOMEM;
    } catch (const std::exception& e) {
        pcfd::set_error(std::string("pcfd_topology_merge: ") + e.what());
        return PCFD_EINVAL;
    }
}

pcfd_status pcfd_topology_instance(const pcfd_topology* topology, int instance, pcfd_instance_info* info) {
    if (!topology || !info) { pcfd::set_error("pcfd_topology_instance: null argument"); return PCFD_EINVAL; }
    const pcfd::Topology& t = topology->t;
    if (instance < 0 || instance >= t.instances()) { pcfd::set_error("pcfd_topology_instance: no such instance"); return PCFD_EINVAL; }
    info->brick_begin = t.inst_brick_off[instance];
    info->brick_end = t.inst_brick_off[instance + 1];
    info->dof_begin = info->brick_begin * 512;
    info->dof_end = info->brick_end * 512;
    info->patch_begin = t.inst_patch_off[instance];
    info->patch_end = t.inst_patch_off[instance + 1];
    info->active_cells = t.inst_active[instance];
    info->reference_area = t.inst_reference_area[instance];
    info->body_volume = t.inst_body_volume[instance];
    return PCFD_OK;
}

namespace {
struct ArrayRef { const void* data; int64_t count; size_t elem; };

bool lookup(const pcfd::Topology& t, int id, ArrayRef& r) {
    auto set = [&](const auto& v, int64_t count = -1) {
        r.data = v.data();
        r.count = count >= 0 ? count : static_cast<int64_t>(v.size());
        r.elem = sizeof(typename std::decay_t<decltype(v)>::value_type);
        return true;
    };
    switch (id) {
        case PCFD_ARR_BRICK_LEVEL: return set(t.brick_level);
        case PCFD_ARR_BRICK_COORD: return set(t.brick_coord);
        case PCFD_ARR_BRICK_FLAGS: return set(t.brick_flags);
        case PCFD_ARR_BRICK_NBR: return set(t.brick_nbr);
        case PCFD_ARR_MASK: return set(t.mask);
        case PCFD_ARR_SPECIAL_DOF: return set(t.special_dof);
        case PCFD_ARR_SPECIAL_VOLUME: return set(t.special_volume);
        case PCFD_ARR_SPECIAL_INC_OFF: return set(t.special_inc_off);
        case PCFD_ARR_SPECIAL_INC: return set(t.special_inc);
        case PCFD_ARR_SPECIAL_PATCH: return set(t.special_patch);
        case PCFD_ARR_BRICK_SPECIAL_OFF: return set(t.brick_special_off);
        case PCFD_ARR_EDGE_A: return set(t.edge_a);
        case PCFD_ARR_EDGE_B: return set(t.edge_b);
        case PCFD_ARR_EDGE_AXIS: return set(t.edge_axis);
        case PCFD_ARR_EDGE_AREA: return set(t.edge_area);
        case PCFD_ARR_EDGE_COEF: return set(t.edge_coef);
        case PCFD_ARR_EDGE_W: return set(t.edge_w);
        case PCFD_ARR_PATCH_DOF: return set(t.patch_dof);
        case PCFD_ARR_PATCH_S: return set(t.patch_S);
        case PCFD_ARR_PATCH_X: return set(t.patch_x);
        case PCFD_ARR_PATCH_D: return set(t.patch_D);
        case PCFD_ARR_EDGE_X: return set(t.edge_x);
        case PCFD_ARR_EDGE_DCOEF: return set(t.edge_dcoef);
        case PCFD_ARR_REDIST_GROUP: return set(t.redist_group);
        case PCFD_ARR_REDIST_OFF: return set(t.redist_off);
        case PCFD_ARR_REDIST_MEMBER: return set(t.redist_member);
        case PCFD_ARR_REDIST_VOLUME: return set(t.redist_volume);
        case PCFD_ARR_CELL_CENTROID: return set(t.centroid);
        case PCFD_ARR_LEVEL_H: return set(t.level_h);
        default: break;
    }
    if (id >= PCFD_ARR_NF_LOWER && id <= PCFD_ARR_NG_RING) {
        const auto& N = t.nonorth;
        if (!N.built) return false;
        switch (id) {
            case PCFD_ARR_NF_LOWER: return set(N.f_lower);
            case PCFD_ARR_NF_UPPER: return set(N.f_upper);
            case PCFD_ARR_NF_EDGE: return set(N.f_edge);
            case PCFD_ARR_NF_AREA: return set(N.f_area);
            case PCFD_ARR_NF_DELTA: return set(N.f_delta);
            case PCFD_ARR_NF_K: return set(N.f_k);
            case PCFD_ARR_NF_W: return set(N.f_w);
            case PCFD_ARR_NF_LNODE: return set(N.f_lnode);
            case PCFD_ARR_NF_UNODE: return set(N.f_unode);
            case PCFD_ARR_NG_DOF: return set(N.g_dof);
            case PCFD_ARR_NG_OFF: return set(N.g_off);
            case PCFD_ARR_NG_NBR: return set(N.g_nbr);
            case PCFD_ARR_NG_W: return set(N.g_w);
            case PCFD_ARR_NG_RANK: return set(N.g_rank);
            case PCFD_ARR_NG_RING: return set(N.g_ring);
            default: return false;
        }
    }
    if (id >= PCFD_ARR_MG_ROW_OFF) {
        const int m = (id - PCFD_ARR_MG_ROW_OFF) / 8, kind = (id - PCFD_ARR_MG_ROW_OFF) % 8;
        if (m < 0 || m >= static_cast<int>(t.mg.size())) return false;
        const pcfd::MGLevel& L = t.mg[m];
        switch (kind) {
            case 0: return m >= 1 && set(L.row_off);
            case 1: return m >= 1 && set(L.col);
            case 2: return m >= 1 && set(L.coef);
            case 3: return m >= 1 && set(L.diag);
            case 4: return set(L.parent);
            case 5: return set(L.coarse_inv);
            default: return false;
        }
    }
    return false;
}
}  // namespace

int64_t pcfd_topology_array_size(const pcfd_topology* topology, int array_id) {
    if (!topology) return -1;
    ArrayRef r{};
    if (!lookup(topology->t, array_id, r)) return -1;
    return r.count;
}

pcfd_status pcfd_topology_export(const pcfd_topology* topology, int array_id, void* dst, size_t bytes) {
    if (!topology) { pcfd::set_error("pcfd_topology_export: null topology"); return PCFD_EINVAL; }
    ArrayRef r{};
    if (!lookup(topology->t, array_id, r)) { pcfd::set_error("pcfd_topology_export: unknown array id"); return PCFD_EINVAL; }
    if (bytes != static_cast<size_t>(r.count) * r.elem) { pcfd::set_error("pcfd_topology_export: size mismatch"); return PCFD_EINVAL; }
    if (bytes) std::memcpy(dst, r.data, bytes);
    return PCFD_OK;
}

}  // extern "C"

const pcfd::Topology& pcfd_topology_ref(const pcfd_topology* t) { return t->t; }
