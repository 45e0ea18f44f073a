// Static preprocessing for one wing/sphere case: 2:1 nested AMR bricks of 8^3 cubic cells,
// sub-sampled cut cells, compact (coarse/fine and partially open) face records, surface
// patches and the geometric multigrid hierarchy.  Host code; the cut-cell sub-sampling can be
// handed to a CellSampler (the device sampler runs it on a GPU, bit-identical to the host's).
//
// Plays the role of AmrHierarchy::build_static (src/core/fluid/amr_grid.cpp:103-142),
// build_amr_embedded_boundary_atlas (amr_eb.cpp:82-375) and
// build_composite_amr_pressure_system (amr_pressure.cpp:317-445).  The cut-cell geometry
// is this project's own analytic-solid sub-sampler; the reference's BSP builder
// (solid_embedded_boundary.cpp:566) needs OCCT and is not reproduced.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/paracfd.h"
#include "build_graph.hpp"

namespace pcfd {

constexpr int kBS = 8;          // brick edge in cells
constexpr int kBC = 512;        // cells per brick

// Domain-boundary bits of a brick (BRICK_XMIN.. in amr_grid.h).
enum BrickFlag : uint8_t { BF_XMIN = 1, BF_XMAX = 2, BF_YMIN = 4, BF_YMAX = 8, BF_ZMIN = 16, BF_ZMAX = 32 };

// Per-cell mask.  Bits 0-2: the +x/+y/+z face is a structured face (same level, fully open,
// both cells active); bits 3-5: the same for the -x/-y/-z face; bit 6: active fluid cell;
// bit 7: special cell (has compact edges, an embedded-boundary patch or a partial volume).
enum CellMask : uint8_t {
    CM_PX = 1, CM_PY = 2, CM_PZ = 4, CM_MX = 8, CM_MY = 16, CM_MZ = 32, CM_ACTIVE = 64, CM_SPECIAL = 128
};

struct MGLevel {
    int n = 0;
    std::vector<int> row_off, col;      // CSR of off-diagonal couplings (symmetric)
    std::vector<double> coef;           // positive coupling: (A x)_a = diag_a x_a - sum coef x_b
    std::vector<double> diag;
    std::vector<int> parent;            // index in level m+1, -1 on the coarsest
    std::vector<int> child_off, child;  // children in level m-1 (for m >= 1)
    std::vector<int> exponent;          // cell size = h_finest * 2^exponent
    std::vector<double> coarse_inv;     // dense inverse on the coarsest level (row-major)
};

// Stages of a topology build, as timed into Topology::build_ms (wall milliseconds)
enum BuildStage : int {
    BUILD_GEOMETRY = 0,    // body and domain set-up
    BUILD_REFINE = 1,      // refinement bands and 2:1 nesting
    BUILD_CANDIDATES = 2,  // cells near the surface (the cut list) and full / solid cells
    BUILD_SAMPLING = 3,    // cut-cell sub-sampling
    BUILD_COMPONENTS = 4,  // small-cell removal and fluid connected to the outlet
    BUILD_FINALIZE = 5,    // brick order, masks, compact edges, wall patches
    BUILD_MULTIGRID = 6,   // coarse levels and the dense coarse inverse
    BUILD_TOTAL = 7,
    BUILD_DEVICE_KERNEL = 8  // the device kernels alone (parts of the stages above)
};

struct Topology {
    // wall time of each build stage (BuildStage) in milliseconds; zero for a merged batch
    double build_ms[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int sampled_on_device = -1;  // device that ran the accelerated build stages, or -1 (host)

    // grid
    double lo[3] = {0, 0, 0};
    double h0 = 0;
    int levels = 1;
    int base_bricks[3] = {0, 0, 0};
    bool outlet = true;
    std::vector<double> level_h;

    // leaf, non-empty bricks sorted by level then Morton order
    int nbricks = 0;
    std::vector<int> brick_level;
    std::vector<int> brick_coord;       // 3 per brick
    std::vector<uint8_t> brick_flags;
    std::vector<int> brick_nbr;         // 6 per brick: -x,+x,-y,+y,-z,+z same-level neighbour
    std::vector<int> level_brick_begin; // levels+1

    // cells (dof = brick*512 + i + 8 j + 64 k)
    int64_t ndof = 0;
    int64_t n_active = 0;
    std::vector<uint8_t> mask;
    std::vector<double> centroid;       // 3 per dof

    // special cells in dof order
    std::vector<int> special_dof;
    std::vector<double> special_volume;
    std::vector<int> special_inc_off;
    std::vector<int> special_inc;       // e >= 0: lower side of edge e; ~e: upper side
    std::vector<int> special_patch;
    std::vector<int> brick_special_off; // nbricks+1

    // compact edges: coarse/fine faces, partially open faces and the fully open same-level
    // faces of embedded nodes (cells with a partial volume or a partially open face)
    std::vector<int> edge_a, edge_b;
    std::vector<int8_t> edge_axis;
    std::vector<double> edge_area, edge_coef, edge_w;
    std::vector<double> edge_dcoef;   // diffusion conductance
    std::vector<double> edge_x;       // 3 per edge: centroid of the face's open part

    // embedded-boundary surface patches (one per cell with a wall)
    std::vector<int> patch_dof;
    std::vector<double> patch_S, patch_x;  // 3 per patch
    std::vector<double> patch_D;           // 9 per patch: int_W (x - c) (x) n dA, row-major [r][c]

    // momentum redistribution groups of small and directionally weak cut cells: group g's
    // members are the special ranks redist_member[redist_off[g] .. redist_off[g+1]) (ascending)
    std::vector<int> redist_group;         // per special cell: its group or -1
    std::vector<int> redist_off, redist_member;
    std::vector<double> redist_volume;     // per group: sum of its members' volumes (member order)

    // multigrid: mg[0] is the composite level (only parent/exponent are filled there)
    std::vector<MGLevel> mg;

    double body_volume = 0;
    double reference_area = 0;
    double max_diffusion_coefficient = 0; // max over cells of sum(coef)/V (for diffusion substeps)
    double finest_h = 0;

    // Batch of independent instances (a single case is a batch of one).  Instance i owns the
    // bricks [inst_brick_off[i], inst_brick_off[i+1]) and, contiguously, their dofs, special
    // cells, compact edges and patches; no face, edge or stencil crosses instances.
    std::vector<int> inst_brick_off;        // ninst+1
    std::vector<int> inst_patch_off;        // ninst+1
    std::vector<int64_t> inst_active;       // active cells per instance
    std::vector<double> inst_reference_area, inst_body_volume, inst_max_diffusion;
    // multigrid level m >= 1: instance i's cells are [mg_inst_off[m][i], mg_inst_off[m][i+1])
    std::vector<std::vector<int>> mg_inst_off;
    // coarsest level of each instance with its dense inverse (level -1: none)
    std::vector<int> dense_level, dense_off, dense_n;
    std::vector<int64_t> dense_inv_off;
    std::vector<double> dense_inv;
    int instances() const { return static_cast<int>(inst_brick_off.size()) - 1; }

    // Non-orthogonal pressure correction (built on request by build_nonorthogonal).  Face f
    // between cells lower < upper along `axis` adds to the composite operator the flux
    //   F = -area (delta (p_upper - p_lower) + k . g_f),  g_f = (1-w) g[lnode] + w g[unode],
    // out[lower] += F, out[upper] -= F, where g[n] = sum_q gw[q] (p[gnbr[q]] - p[gdof[n]]) is
    // the weighted least-squares gradient at node n.  nf_edge: compact edge e, or -1-axis for
    // a structured face (its face velocity is f[axis][lower]).
    struct NonOrth {
        bool built = false;
        std::vector<int> f_lower, f_upper, f_edge, f_lnode, f_unode;
        std::vector<double> f_area, f_delta, f_k, f_w;   // f_k: 3 per face
        std::vector<int> g_dof, g_off, g_nbr;
        std::vector<double> g_w;                        // 3 per stencil entry
        std::vector<uint8_t> g_rank, g_ring;
        std::vector<int> inc_dof, inc_off, inc_face;    // per touched dof: f (lower side) or ~f
        int64_t regular_faces = 0, compact_faces = 0, below_resolution = 0;
    } nonorth;
};

// Build the non-orthogonal correction of the pressure operator (idempotent; throws
// BuildError for a face whose centroids do not bracket it or a correction the gradient
// stencils cannot represent).  Single-instance topologies only.
void build_nonorthogonal(Topology& T);
Topology::NonOrth compute_nonorthogonal(const Topology& T);

// A geometry, grid or topology the builders cannot represent: a bad argument of the caller
// (reported as PCFD_EINVAL wherever it crosses the C ABI, also from pcfd_create).
struct BuildError : std::invalid_argument {
    explicit BuildError(const std::string& m) : std::invalid_argument(m) {}
};
// A HIP runtime failure inside a build step that runs on a device (PCFD_EDEVICE).
struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

// Outputs of the builder's per-cell passes over the brick-ordered cell graph (dof = brick * 512 + q)
struct CellPasses {
    std::vector<double> volume;         // per dof (0 for inactive cells)
    std::vector<double> centroid;       // 3 per dof
    std::vector<uint8_t> embedded;      // per dof
    std::vector<uint8_t> mask;          // per dof: active and structured-face bits
    std::vector<uint8_t> has_compact;   // per dof
    std::vector<double> diffusion_sum;  // per dof
    std::vector<double> S;              // 3 per dof: closure area vector
    std::vector<int64_t> wall_dof;      // active cells with |S| > 1e-12 h^2, ascending
    std::vector<double> wall_moment;    // 9 per wall cell: cell_faces' fm
    std::vector<EdgeRecord> edges;      // compact edges in creation order (dof, then face order)
};

// Stages of the build that can run elsewhere than on the host threads (the device build does
// them on a GPU).  Each must give exactly what the host builder gives:
//   sample: the cut list's cells (4 ints each: integer position in its level's cell grid, then
//     the level) -> one CutData per cell, as sample_solid / cut_part give it;
//   connected_fluid: clear every active cell that is not in a component of the graph (faces with
//     an open fraction > 0 between active cells) touching the x-max boundary (outlet), or,
//     without an outlet, not in the largest component (ties: the one holding the lowest cell);
//   cell_passes: cell_geometry on every cell, then cell_faces on every active cell, in dof order;
//   multigrid: the coarse levels (see Builder::host_multigrid).
class BuildAccelerator {
public:
    virtual ~BuildAccelerator() = default;
    virtual void sample(const SolidDesc& solid, const double lo[3], const std::vector<double>& level_h, int ns,
                        const std::vector<int>& cells, std::vector<CutData>& out) = 0;
    virtual void connected_fluid(const BuildView& view, int64_t nleaves, bool outlet, std::vector<uint8_t>& active) = 0;
    virtual void cell_passes(const BuildView& view, int nbricks, CellPasses& out) = 0;
    // the multigrid levels (mg[0] the composite level's exponent and parent, then every coarse
    // level) from the finished composite topology, as the host's coarsening builds them (the
    // dense coarse inverse excluded); false to leave them to the host
    virtual bool multigrid(const Topology& T, int levels, int coarse_max, std::vector<MGLevel>& mg) = 0;
    virtual int device() const { return -1; }       // the device it runs on, or -1
    virtual double kernel_ms() const { return 0.0; } // device kernel time so far
};

// Build a topology; the stages of `accel` run there when given, else on the host.
Topology build_topology(const pcfd_geometry_desc& geometry, const pcfd_grid_desc& grid,
                        BuildAccelerator* accel = nullptr);

// Build a C ABI topology handle (the sampler as for build_topology), mapping a BuildError to
// PCFD_EINVAL, a DeviceError to PCFD_EDEVICE and host allocation failure to PCFD_ENOMEM.
pcfd_status build_topology_handle(const char* where, const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid,
                                  BuildAccelerator* accel, pcfd_topology** out);

// Concatenate independent topologies into one batch (same AMR levels and cell sizes; each
// part needs its dense coarse inverse).  Every index is offset into the merged numbering.
Topology merge_topologies(const std::vector<const Topology*>& parts);

}  // namespace pcfd
