/*
 * paracfd.h - C ABI of the MI355X solver step for ParaCFD's external-aerodynamics core.
 *
 * One solver step = conservative MUSCL cell-momentum transport, MGPCG pressure projection
 * to a relative residual tolerance, and the surface-load integral, on static Cartesian AMR
 * bricks (8^3 cells) with cut-cell embedded boundaries.
 *
 * The reference exposes this path as the C++ class ExternalAeroCore
 * (src/core/fluid/external_aero_core.h:124-182); it has no C ABI.  Each entry point below
 * names the reference member it replaces.  Conventions:
 *   - every function returns a pcfd_status; PCFD_OK == 0.  The reference throws
 *     std::invalid_argument / std::runtime_error (src/core/fluid/amr_fields.cu:53,
 *     src/core/fluid/external_aero_core.cpp:242,306,308); here the same conditions return
 *     PCFD_EINVAL / PCFD_EDEVICE / PCFD_ENONFINITE / PCFD_ESTATE with the message in
 *     pcfd_last_error().
 *   - a projection that does not reach the tolerance is a status, not an error
 *     (external_aero_core.cpp:327): pcfd_step returns PCFD_OK with stats->converged == 0
 *     and physical time does not advance.
 *   - a solver handle owns all of its device memory and one HIP stream.  Calls on one
 *     handle are not thread-safe; distinct handles are independent.
 *   - host arrays are plain pointers with explicit element counts; nothing here needs a
 *     torch or HIP type.
 */
#ifndef PARACFD_H
#define PARACFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCFD_ABI_VERSION 5

typedef enum pcfd_status {
    PCFD_OK = 0,
    PCFD_EINVAL = 1,          /* bad argument (reference: std::invalid_argument) */
    PCFD_EDEVICE = 2,         /* HIP runtime failure (reference: check() -> runtime_error) */
    PCFD_ENOTCONVERGED = 3,   /* reserved: projection failed and the caller asked for an error */
    PCFD_ENONFINITE = 4,      /* non-finite state before a step (external_aero_core.cpp:308) */
    PCFD_ESTATE = 5,          /* wrong call order, e.g. step before initialize (:306) */
    PCFD_ENOMEM = 6           /* device or host allocation failed / over budget (:91-95) */
} pcfd_status;

typedef enum pcfd_geometry_kind {
    PCFD_GEOM_NONE = 0,        /* empty box */
    PCFD_GEOM_SPHERE = 1,      /* sphere: center, radius */
    PCFD_GEOM_NACA4_WING = 2,  /* straight extruded NACA 4-digit wing, span along y */
    PCFD_GEOM_PARAGLIDER = 3   /* arched, swept, elliptic-planform reflexed wing */
} pcfd_geometry_kind;

typedef struct pcfd_geometry_desc {
    int kind;                 /* pcfd_geometry_kind */
    double center[3];         /* sphere centre; wing: leading edge of the root section (pitch pivot) */
    double radius;            /* sphere radius */
    int naca_digits;          /* 4-digit code, e.g. 2412 (wing) */
    double chord;             /* wing root chord [m] (paraglider: 0 = aspect ratio 5 from the span) */
    double span;              /* wing span along y [m], centred on center[1] */
    double alpha_deg;         /* pitch about +y through center; positive = nose up */
    double sweep_deg;         /* paraglider: leading-edge sweep added to the elliptic quarter-chord line */
    double anhedral_deg;      /* paraglider: turn of each half of the circular span arc (tip droop angle) */
    double reflex;            /* paraglider: upturn of the camber line at the trailing edge (chords),
                                 quadratic over the last 30 % of the chord */
} pcfd_geometry_desc;

typedef struct pcfd_grid_desc {
    double lo[3], hi[3];      /* requested domain box; y/z are padded symmetrically and x downstream
                                 to whole 8^3 bricks of cubic cells (amr_grid.cpp:66-96) */
    int base_cells_x;         /* level-0 cells along x; h0 = (hi[0]-lo[0]) / base_cells_x */
    int levels;               /* AMR levels, 1..4; level l has h = h0 / 2^l */
    double refine_band[4];    /* level-l bricks within refine_band[l] of the surface are refined */
    double min_volume_fraction; /* cut cells below this fluid fraction become solid */
    int subsamples;           /* sub-samples per axis for cut-cell volume/aperture */
    int outlet_xmax;          /* 1: Dirichlet p = 0 at x-max (pressure_outlet_xmax); 0: all Neumann */
    double merge_volume_fraction; /* active cut cells below this fluid fraction, or with an axis whose
                                 open area is below it times h^2, share their momentum updates with the
                                 neighbour across their largest open face (redistribution groups, the
                                 reference's same-side merge threshold, paraglider_config.h:37, and
                                 momentum redistribution, amr_advection.cu:1108-1178); 0: none */
} pcfd_grid_desc;

typedef struct pcfd_topology pcfd_topology;
typedef struct pcfd_solver pcfd_solver;

typedef struct pcfd_topology_info {
    int64_t bricks;           /* stored (leaf, non-empty) bricks */
    int64_t dofs;             /* bricks * 512 slots, brick-major */
    int64_t active_cells;     /* N: the cells in cell-steps/s */
    int64_t special_cells;    /* cells touching a compact edge, cut cells, EB patches */
    int64_t compact_edges;    /* coarse/fine and partially open faces */
    int64_t patches;          /* embedded-boundary surface patches */
    int64_t mg_levels;        /* multigrid levels including the composite level */
    int64_t mg_cells[12];     /* cells per multigrid level */
    int levels;
    double h0;
    double lo[3];             /* padded domain origin */
    int base_bricks[3];       /* level-0 brick grid */
    double body_volume;       /* solid volume represented by the cut cells */
    double reference_area;    /* wing planform area or sphere frontal area */
    double finest_h;          /* cell size of the finest level */
    double max_diffusion_coefficient; /* max over cells of sum(open area / distance) / volume */
    int outlet;               /* 1 when x-max is a Dirichlet outlet */
    int instances;            /* independent cases in this topology (1 unless merged) */
} pcfd_topology_info;

/* One instance of a merged topology: its contiguous ranges in the merged numbering. */
typedef struct pcfd_instance_info {
    int64_t brick_begin, brick_end;
    int64_t dof_begin, dof_end;   /* = 512 * bricks */
    int64_t patch_begin, patch_end;
    int64_t active_cells;
    double reference_area;
    double body_volume;
} pcfd_instance_info;

/* Topology arrays exported for host-side checks (oracle, tests). */
typedef enum pcfd_array_id {
    PCFD_ARR_BRICK_LEVEL = 0,      /* int32[bricks] */
    PCFD_ARR_BRICK_COORD = 1,      /* int32[bricks*3] brick coordinates in its level's brick grid */
    PCFD_ARR_BRICK_FLAGS = 2,      /* uint8[bricks] domain-boundary bits */
    PCFD_ARR_BRICK_NBR = 3,        /* int32[bricks*6] same-level neighbour brick or -1 */
    PCFD_ARR_MASK = 4,             /* uint8[dofs] cell mask bits */
    PCFD_ARR_SPECIAL_DOF = 5,      /* int32[special] */
    PCFD_ARR_SPECIAL_VOLUME = 6,   /* double[special] fluid volume */
    PCFD_ARR_SPECIAL_INC_OFF = 7,  /* int32[special+1] */
    PCFD_ARR_SPECIAL_INC = 8,      /* int32[inc]: edge e if the cell is the lower side, ~e if upper */
    PCFD_ARR_SPECIAL_PATCH = 9,    /* int32[special] patch index or -1 */
    PCFD_ARR_BRICK_SPECIAL_OFF = 10, /* int32[bricks+1] */
    PCFD_ARR_EDGE_A = 11,          /* int32[edges] lower cell along the axis */
    PCFD_ARR_EDGE_B = 12,          /* int32[edges] upper cell */
    PCFD_ARR_EDGE_AXIS = 13,       /* int8[edges] */
    PCFD_ARR_EDGE_AREA = 14,       /* double[edges] open area */
    PCFD_ARR_EDGE_COEF = 15,       /* double[edges] open area / centre distance */
    PCFD_ARR_EDGE_W = 16,          /* double[edges] interpolation weight of the upper cell at the face
                                      centroid (composite_face_interpolation_weight,
                                      amr_advection.cu:732-753) */
    PCFD_ARR_PATCH_DOF = 17,       /* int32[patches] */
    PCFD_ARR_PATCH_S = 18,         /* double[patches*3] area vector, fluid -> solid */
    PCFD_ARR_PATCH_X = 19,         /* double[patches*3] patch centroid */
    PCFD_ARR_PATCH_D = 20,         /* double[patches*9] first moment of the wall about the cell
                                      centroid, int_W (x - c) (x) n dA, row-major: the wall force of
                                      a pressure p + g.(x - c) is p S + D^T g */
    PCFD_ARR_EDGE_X = 21,          /* double[edges*3] centroid of the face's open part */
    PCFD_ARR_EDGE_DCOEF = 22,      /* double[edges] diffusion conductance */
    PCFD_ARR_REDIST_GROUP = 40,    /* int32[special] redistribution group of each special cell or -1 */
    PCFD_ARR_REDIST_OFF = 41,      /* int32[groups+1] */
    PCFD_ARR_REDIST_MEMBER = 42,   /* int32[members] special ranks of each group's cells, ascending */
    PCFD_ARR_REDIST_VOLUME = 43,   /* double[groups] fluid volume of each group */
    PCFD_ARR_CELL_CENTROID = 23,   /* double[dofs*3] fluid centroid (cell centre for full cells) */
    PCFD_ARR_LEVEL_H = 24,         /* double[levels] */
    /* non-orthogonal correction (after pcfd_topology_build_nonorthogonal; size -1 before) */
    PCFD_ARR_NF_LOWER = 25,        /* int32[faces] lower cell along the face's axis */
    PCFD_ARR_NF_UPPER = 26,        /* int32[faces] upper cell */
    PCFD_ARR_NF_EDGE = 27,         /* int32[faces] compact edge, or -1-axis for a structured face */
    PCFD_ARR_NF_AREA = 28,         /* double[faces] open area */
    PCFD_ARR_NF_DELTA = 29,        /* double[faces] 1/d_axis minus the orthogonal factor */
    PCFD_ARR_NF_K = 30,            /* double[faces*3] correction vector e_axis - d/d_axis */
    PCFD_ARR_NF_W = 31,            /* double[faces] gradient weight of the upper cell */
    PCFD_ARR_NF_LNODE = 32,        /* int32[faces] gradient node of the lower cell or -1 */
    PCFD_ARR_NF_UNODE = 33,        /* int32[faces] gradient node of the upper cell or -1 */
    PCFD_ARR_NG_DOF = 34,          /* int32[nodes] cell of each gradient node */
    PCFD_ARR_NG_OFF = 35,          /* int32[nodes+1] */
    PCFD_ARR_NG_NBR = 36,          /* int32[entries] stencil cells */
    PCFD_ARR_NG_W = 37,            /* double[entries*3] g = sum w (p_nbr - p_node) */
    PCFD_ARR_NG_RANK = 38,         /* uint8[nodes] rank of the stencil's normal matrix */
    PCFD_ARR_NG_RING = 39,         /* uint8[nodes] 1 + rings added for a deficient stencil */
    PCFD_ARR_MG_ROW_OFF = 100,     /* + 8*m: int32[n_m+1] CSR of multigrid level m (m >= 1) */
    PCFD_ARR_MG_COL = 101,         /* + 8*m: int32[nnz] */
    PCFD_ARR_MG_COEF = 102,        /* + 8*m: double[nnz] off-diagonal couplings (positive) */
    PCFD_ARR_MG_DIAG = 103,        /* + 8*m: double[n_m] */
    PCFD_ARR_MG_PARENT = 104,      /* + 8*m: int32[n_m] (m=0: dofs) index in level m+1 or -1 */
    PCFD_ARR_MG_COARSE_INV = 105   /* + 8*m: double[n_m*n_m] dense inverse, coarsest level only */
} pcfd_array_id;

typedef enum pcfd_precision { PCFD_FP64 = 0, PCFD_FP32 = 1, PCFD_BF16 = 2 } pcfd_precision;
typedef enum pcfd_preconditioner {
    PCFD_PC_MG = 0,       /* geometric V-cycle (plain MGPCG) */
    PCFD_PC_JACOBI = 1,   /* diagonal */
    PCFD_PC_LEARNED = 2   /* learned conv-net correction inside flexible PCG, MG fallback */
} pcfd_preconditioner;

typedef struct pcfd_options {
    double rho;               /* density (freestream.rho) */
    double nu;                /* molecular kinematic viscosity */
    double speed;             /* freestream speed along +x */
    double cfl;               /* solver.cfl */
    double smagorinsky_cs;    /* 0 disables the Smagorinsky model */
    double rtol;              /* projection_tolerance, relative to ||rhs|| */
    int max_iterations;       /* projection_max_iterations */
    int field_precision;      /* PCFD_FP64 or PCFD_FP32 velocity / face fields (Real) */
    int vcycle_precision;     /* PCFD_FP64, PCFD_FP32 or PCFD_BF16 (storage) V-cycle */
    int preconditioner;       /* pcfd_preconditioner */
    int pre_smooth, post_smooth; /* V-cycle Jacobi sweeps */
    int check_interval;       /* Krylov iterations between host convergence reads (>=1);
                                 results do not depend on it */
    double smoother_weight;   /* damped-Jacobi weight of the first sweep of each pair of V-cycle
                                 sweeps (sweep s uses smoother_weight for even s, smoother_weight2
                                 for odd s); default 1/(1.2 + 0.8/sqrt 2) = 0.5664 (the reference's
                                 algebraic MG uses 2/3 for every sweep, amr_pressure.cu:1103) */
    int learned_bricks;       /* PCFD_PC_LEARNED: bricks the network corrects: 0 all bricks,
                                 1 bricks holding special cells (cut cells, coarse/fine faces)
                                 and their face neighbours, 2 bricks holding cut cells and
                                 their face neighbours */
    int learned_capacity;     /* PCFD_PC_LEARNED: recorded residuals kept for training */
    int smooth_wall;          /* 1: smooth-wall shear model on the surface patches after the
                                 diffusion of every step (Spalding's law, implicit per wall cell;
                                 reference: ExternalAeroExecutionOptions::smooth_fabric_wall,
                                 external_aero_core.h:112, on by default there;
                                 DeviceCompositeCellMomentumTransport::apply_smooth_fabric_wall_model,
                                 amr_advection.cu:1347-1365); 0 (default here): off */
    double smoother_weight2;  /* weight of the second sweep of each pair; default
                                 1/(1.2 - 0.8/sqrt 2) = 1.5765: with smoother_weight the roots of the
                                 degree-2 Chebyshev polynomial on [0.4, 2] of D^-1 A.  Set both equal
                                 for plain damped Jacobi.  The pair must satisfy
                                 |(1 - w1 l)(1 - w2 l)| < 1 on (0, 2] */
    int nonorthogonal;        /* 1: the projection solves the full operator with the
                                 non-orthogonal correction (centroid two-point factors plus
                                 the least-squares correction term, amr_pressure.cu:831-834)
                                 by flexible GMRES right-preconditioned with the V-cycle of the
                                 orthogonal operator, and the face correction carries the same
                                 derivative (amr_pressure.cu:2047-2069).  The reference uses
                                 GCRO-LGMRES with recycling for the same system
                                 (amr_pressure.cu:1390-1698).  Single cases with the MG
                                 preconditioner only.  0 (default): the orthogonal operator and
                                 PCG */
    int gmres_restart;        /* Krylov vectors per GMRES cycle (default 24, as the reference;
                                 clamped to 2..30) */
    double memory_budget_bytes; /* device memory the solver may allocate; pcfd_create and any later
                                 allocation beyond it fail with PCFD_ENOMEM before calling the
                                 allocator (reference: ExternalAeroExecutionOptions::
                                 gpu_memory_budget_bytes, external_aero_core.cpp:91-95); 0: no limit */
    int gmres_recycle;        /* non-orthogonal solve: GCRO recycling (amr_pressure.cu:1390-1698)
                                 keeps this many realised corrections U and their orthonormal
                                 images C = A U across cycles and projections (3 is the reference's
                                 capacity, amr_pressure.h:396; at most 8).  Default 0, plain flexible
                                 GMRES: with this V-cycle the recycled space leaves the iteration
                                 counts unchanged and only adds vector traffic */
} pcfd_options;

typedef struct pcfd_step_stats {
    double dt;
    double physical_time;
    int iterations;           /* Krylov iterations of this projection */
    int converged;
    double relative_residual; /* ||r|| / ||b|| at exit */
    double max_abs_velocity;
    double cfl_rate;          /* max outflow rate used for dt */
    double max_abs_divergence;/* after projection, integrated divergence / volume (FP64) */
    int diffusion_substeps;
    int fallback_used;        /* learned preconditioner fell back to plain MG */
    int learned_used;         /* the learned network ran in this projection */
    int learned_rejected;     /* validation after training found the network slower than plain
                                 MGPCG and switched it off (stays set until the next training) */
    double advect_ms, project_ms, step_ms; /* device-event timings */
} pcfd_step_stats;

typedef struct pcfd_loads {
    double force[3];          /* pressure force on the body [N] */
    double moment[3];         /* about the moment origin */
    double drag, side, lift;  /* force x, y, z */
    double cd, cs, cl;        /* coefficients, valid when q*S_ref > 0 */
    int coefficients_valid;
    /* wall shear of the smooth-wall model's last application (reference:
     * accumulate_viscous_loads, aero_loads.cpp:49-56); zero and viscous_valid = 0 while the
     * model is off or has not run */
    int viscous_valid;
    double viscous_force[3], viscous_moment[3];
    double viscous_drag, viscous_side, viscous_lift;
    double cd_viscous, cs_viscous, cl_viscous;
    /* pressure + viscous (reference: ExternalAeroCore::aerodynamic_loads) */
    double total_force[3], total_moment[3];
    double total_drag, total_side, total_lift;
    double total_cd, total_cs, total_cl;
} pcfd_loads;

/* Last error message of the calling thread (reference: exception what()). */
const char* pcfd_last_error(void);
int pcfd_abi_version(void);

/* Static preprocessing: AMR bricks, cut cells, compact edges, surface patches and the
 * multigrid hierarchy.  Replaces the ExternalAeroCore constructor's preprocessing
 * (external_aero_core.cpp:45-236: AmrHierarchy::build_static, build_amr_embedded_boundary_atlas,
 * build_composite_amr_pressure_system). Host only; needs no GPU. */
pcfd_status pcfd_topology_build(const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid,
                                pcfd_topology** out);
/* The same build with the cut-cell sub-sampling (the reference's cut-cell classification,
 * amr_eb.cpp:82-375, here by sub-sample counting on the analytic body) run on HIP device
 * `device`: every exported array is bit-identical to pcfd_topology_build's.  Needs
 * grid->subsamples <= 16.  Device failures are PCFD_EDEVICE. */
pcfd_status pcfd_topology_build_on_device(const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid,
                                          int device, pcfd_topology** out);
void pcfd_topology_free(pcfd_topology* topology);
/* Wall time of each build stage in milliseconds (no counterpart in the reference): 0 body and
 * domain, 1 refinement and nesting, 2 cut list, 3 cut-cell sub-sampling (device: including
 * the transfers), 4 small-cell removal and connected fluid, 5 bricks / masks / compact edges /
 * wall patches, 6 multigrid, 7 total, 8 the device sampling kernel alone (0 on the host);
 * *device = the device that sub-sampled, or -1.  Zero for a merged batch. */
pcfd_status pcfd_topology_build_times(const pcfd_topology* topology, double ms[9], int* device);
pcfd_status pcfd_topology_get_info(const pcfd_topology* topology, pcfd_topology_info* info);
/* Build the non-orthogonal correction of the pressure operator into the topology
 * (finalize_nonorthogonal_pressure_topology, amr_pressure.cpp:63-190): per-face centroid
 * factors and correction vectors, least-squares gradient stencils with ring extension of
 * rank-deficient ones, and the check that every correction is represented.  Idempotent;
 * single-instance topologies only.  Solvers created with options.nonorthogonal = 1 call it. */
pcfd_status pcfd_topology_build_nonorthogonal(pcfd_topology* topology);
/* Element count of an exported array (see pcfd_array_id), or -1. */
int64_t pcfd_topology_array_size(const pcfd_topology* topology, int array_id);
/* Copy an exported array into dst (bytes must equal count * element size). */
pcfd_status pcfd_topology_export(const pcfd_topology* topology, int array_id, void* dst, size_t bytes);

/* Batching (no counterpart in the reference, which runs one ExternalAeroCore per case):
 * concatenate `count` single-case topologies with the same AMR levels and cell sizes into
 * one batch topology.  A solver created on it advances every instance in the same kernel
 * launches and replayed graphs; each instance keeps its own timestep, Krylov scalars,
 * convergence and loads, and computes exactly what it computes alone.  The parts may be
 * freed afterwards. */
pcfd_status pcfd_topology_merge(const pcfd_topology* const* parts, int count, pcfd_topology** out);
pcfd_status pcfd_topology_instance(const pcfd_topology* topology, int instance, pcfd_instance_info* info);

/* Device solver on the current HIP device.  Replaces ExternalAeroCore's device objects
 * (DeviceAmrFields, DeviceCompositeAmrProjection, DeviceCompositeCellMomentumTransport;
 * external_aero_core.cpp:77-200). The topology may be freed after creation. */
pcfd_status pcfd_create(const pcfd_topology* topology, const pcfd_options* options, pcfd_solver** out);
void pcfd_destroy(pcfd_solver* solver);
void pcfd_default_options(pcfd_options* options);

/* ExternalAeroCore::initialize (external_aero_core.cpp:245-302): freestream seed, face
 * reconstruction, first projection with dt = cfl*h/U, first pressure impulse. */
pcfd_status pcfd_initialize(pcfd_solver* solver, pcfd_step_stats* stats);

/* ExternalAeroCore::step (external_aero_core.cpp:304-328), cell-momentum branch:
 * CFL dt -> advect -> diffuse -> reconstruct face fluxes -> project -> pressure impulse. */
pcfd_status pcfd_step(pcfd_solver* solver, pcfd_step_stats* stats);

/* The three stages of one step, callable separately.
 * pcfd_advect: DeviceCompositeCellMomentumTransport::step + diffuse_smagorinsky +
 *   reconstruct_fluxes (amr_advection.cu:1322-1326,1339-1345,1423-1456) and the external
 *   boundary faces (amr_fields.cu:86-118).  dt <= 0 selects the CFL timestep
 *   (external_aero_core.cpp:309); the timestep used is written to *dt_used. */
pcfd_status pcfd_advect(pcfd_solver* solver, double dt, double* dt_used, pcfd_step_stats* stats);
/* DeviceCompositeAmrProjection::project(rho, dt, tol, maxit, warm) (amr_pressure.h:478,
 * amr_pressure.cu:2085-2095): RHS, PCG solve, face-flux correction, divergence check. */
pcfd_status pcfd_project(pcfd_solver* solver, double dt, int warm_start, pcfd_step_stats* stats);
/* apply_projected_pressure_gradient (amr_advection.cu:1412-1421). */
pcfd_status pcfd_apply_pressure_impulse(pcfd_solver* solver, double dt);
/* ExternalAeroCore::pressure_loads(pref) -> compute_pressure_loads (external_aero_core.cpp:330-337,
 * aero_loads.cpp:35-47), evaluated on the device in FP64. S_ref and the moment origin come
 * from pcfd_set_reference. */
pcfd_status pcfd_integrate_loads(pcfd_solver* solver, double pref, pcfd_loads* loads);
/* Sets S_ref and the moment origin of every instance. */
pcfd_status pcfd_set_reference(pcfd_solver* solver, double area, const double origin[3]);

/* Batches (a solver on a merged topology, pcfd_topology_merge).  pcfd_step / pcfd_initialize
 * / pcfd_project advance every instance; their stats then report the worst instance
 * (iterations, residual, divergence), all-instance convergence and the smallest dt, and
 * pcfd_instance_stats gives each instance's own stats of the last call.  Each instance
 * advances its own physical time only when its projection converged.  Loads are per
 * instance (pcfd_integrate_loads is for single cases).  No counterpart in the reference. */
pcfd_status pcfd_instance_stats(const pcfd_solver* solver, int instance, pcfd_step_stats* stats);
pcfd_status pcfd_integrate_loads_instance(pcfd_solver* solver, int instance, double pref, pcfd_loads* loads);
/* Loads of every instance (loads[count], count == instances) with one device synchronisation. */
pcfd_status pcfd_integrate_loads_batch(pcfd_solver* solver, double pref, pcfd_loads* loads, int count);
pcfd_status pcfd_set_reference_instance(pcfd_solver* solver, int instance, double area, const double origin[3]);

/* Launch-shape tuning.  No counterpart in the reference (its kernels have fixed launch
 * shapes).  The tunables decide which workgroup computes which brick or coarse-level row, or
 * whether the composite-level V-cycle runs as fused kernels; none changes the arithmetic, so
 * every configuration gives bit-identical results.  pcfd_autotune times `repetitions` full
 * steps per candidate from one snapshot of the state (greedy: XCD chunk, then coarse
 * workgroup size, then fused or plain V-cycle), keeps the fastest configuration and restores
 * the state and physical time, so the solution is unchanged.  Needs a converged
 * initialization. */
typedef struct pcfd_tuning {
    int xcd_chunk_log2;     /* bricks go to the 8 XCDs in chunks of 2^k (-1: one range per XCD) */
    int coarse_block;       /* workgroup size of the coarse-level kernels: 64, 128, 256 or 512 */
    int fused_vcycle;       /* 1: fused V-cycle kernels (composite level; residual + restriction on
                               coarse levels), 0: the plain sequence */
    int graph_replay;       /* 1: PCG iterations replayed from a captured HIP graph, 0: launched one
                               by one (not searched by pcfd_autotune) */
    int candidates;         /* configurations timed by the last pcfd_autotune */
    double step_ms;         /* fastest step of the chosen configuration (last pcfd_autotune) */
    double default_step_ms; /* fastest step of the configuration it started from */
} pcfd_tuning;
pcfd_status pcfd_autotune(pcfd_solver* solver, int repetitions, pcfd_tuning* result);
pcfd_status pcfd_get_tuning(const pcfd_solver* solver, pcfd_tuning* tuning);
/* Applies the first four fields of a tuning (e.g. one stored from an earlier autotune). */
pcfd_status pcfd_set_tuning(pcfd_solver* solver, const pcfd_tuning* tuning);

/* State I/O (upload_state/download_state amr_advection.cu:1301-1309, upload_pressure
 * amr_pressure.cu:2098-2108, download_pressure).  Arrays are FP64 host arrays of length
 * dofs (velocity: three arrays).  Face velocities: +x/+y/+z face of every dof, then one
 * value per compact edge. */
pcfd_status pcfd_get_velocity(pcfd_solver* solver, double* x, double* y, double* z);
pcfd_status pcfd_set_velocity(pcfd_solver* solver, const double* x, const double* y, const double* z);
pcfd_status pcfd_get_pressure(pcfd_solver* solver, double* p);
pcfd_status pcfd_set_pressure(pcfd_solver* solver, const double* p);
pcfd_status pcfd_get_faces(pcfd_solver* solver, double* fx, double* fy, double* fz, double* edge);
pcfd_status pcfd_set_faces(pcfd_solver* solver, const double* fx, const double* fy, const double* fz,
                           const double* edge);
/* Mark the state as initialized after pcfd_set_* (checkpoint resume). */
pcfd_status pcfd_mark_initialized(pcfd_solver* solver, double physical_time);
int64_t pcfd_device_bytes(const pcfd_solver* solver);

/* Field of one instance on the plane x[axis] = coord, as an image for a viewer (no counterpart
 * on the AMR path of the reference, which samples its AMR state on the host,
 * core/fluid/amr_sampling.h, and colours a uniform MAC state on the device,
 * gui/slice_field.h:20-31).  out[i + nu j] is the value at the centre of pixel (i, j) of the
 * window lo..hi in the plane's two other axes (in increasing order): the cell-centred value of
 * the finest leaf cell containing the point, NaN where it is solid or outside the grid.
 * field: 0 |v|, 1 u, 2 v, 3 w, 4 pressure, 5 Cp = p / (rho U^2 / 2).  out is host memory
 * (nu * nv floats); only the image crosses PCIe. */
pcfd_status pcfd_sample_slice(pcfd_solver* solver, int instance, int field, int axis, double coord, const double lo[2],
                              const double hi[2], int nu, int nv, float* out);
/* pcfd_sample_slice without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (a viewer's render loop, code under HIP graph capture, harnesses
 * that forbid hipMalloc).  The image is staged in a fixed device array of 2^20 floats per GPU,
 * held until the process ends; nu * nv above 2^20 is PCFD_EINVAL.  Calls on one GPU share that
 * array, so calls that overlap on one GPU (other threads, other solvers) race.  Same arguments,
 * results and errors as pcfd_sample_slice otherwise. */
pcfd_status pcfd_sample_slice_static(pcfd_solver* solver, int instance, int field, int axis, double coord,
                                     const double lo[2], const double hi[2], int nu, int nv, float* out);

/* Multigrid primitives on level `level` (0 = composite level), host arrays in, host array out.
 * op: 0 apply A x, 1 residual b - A x, 2 damped-Jacobi sweeps on x (omega, sweeps),
 *     3 one full V-cycle z = M^-1 b.  The uniform legacy primitives project.cu:286-328
 *     (poisson_apply_cpu, poisson_residual_cpu, jacobi_smooth_cpu) are the special case
 *     of a single-level box without outlet, scaled by h^3. */
pcfd_status pcfd_mg_primitive(pcfd_solver* solver, int op, int level, const double* x, const double* b,
                              double* out, double omega, int sweeps);

/* Measurement hook with no counterpart in the reference (which has no kernel-level timing):
 * average duration of one launch of a hot kernel on the solver's stream, timed with HIP
 * events over `reps` launches, and its algorithmic HBM bytes per launch.
 * kernel: 0 composite operator apply + dot (CG), 1 composite-level V-cycle sweep,
 *         2 MUSCL transport, 3 one full V-cycle (bytes reported as 0), 4 learned correction
 *         (learned solvers only), 5 CG direction update + operator apply + dot
 *         (k_cg_dir_apply), 6 CG x/r update + two pre-smoothing sweeps (k_vc_update_pre2; its
 *         x is a scratch vector, so the state is unchanged).
 * Leaves the Krylov loop flags in the "finished" state; call between steps. */
pcfd_status pcfd_time_kernel(pcfd_solver* solver, int kernel, int reps, double* avg_ms, double* bytes_per_launch);

/* In-step launch timing, no counterpart in the reference: between pcfd_probe_start and
 * pcfd_probe_stop every launch of the kernels in `kernel_mask` (bit k for kernel id k of
 * pcfd_time_kernel; ids 1, 2, 5 and 6) made by the calls in between (pcfd_step, ...) is
 * bracketed by a pair of HIP events on the solver's stream.  While probing, the Krylov
 * iterations are launched one by one instead of replayed from captured graphs (same
 * arithmetic, same results).  pcfd_probe_stop returns, per kernel id 0..7, the number of
 * timed launches and their summed duration in ms (arrays of 8, either may be NULL). */
pcfd_status pcfd_probe_start(pcfd_solver* solver, unsigned kernel_mask);
pcfd_status pcfd_probe_stop(pcfd_solver* solver, int64_t* launches, double* total_ms);

/* ---- Learned preconditioner (PCFD_PC_LEARNED) --------------------------------------------
 * No counterpart in the reference, whose projection runs plain (GCRO / PCG) Krylov with an
 * algebraic V-cycle (amr_pressure.cu:1103-1353).  A 4-layer 3^3 conv net per brick (16
 * channels, one weight set per AMR level, BF16 MFMA, FP32 accumulate; its input reads a
 * one-cell halo from the 26 neighbouring bricks) corrects the V-cycle inside flexible PCG:
 * z = M_mg(r) + (s_b/h_b) N_level(r'/s_b) with r' = r - A M_mg(r) and s_b the brick's RMS of r'.
 * It is trained to minimise the energy norm of the error it leaves, dz^T A dz - 2 dz^T r'.
 * A projection whose learned solve stalls or fails restarts from its iterate with plain
 * MGPCG (pcfd_step_stats.fallback_used).  The network starts disabled; it is enabled by
 * pcfd_learned_train or pcfd_learned_set_weights.  After pcfd_learned_train, the next three
 * steps validate it: if their mean projection time exceeds that of the last (up to four)
 * projections without the network, the network is switched off (learned_rejected). */
#define PCFD_NN_WEIGHTS 110592  /* 4 sets x 4 layers x 27 taps x 16 cout x 16 cin,
                                   [set][layer][tap][cout][cin]; set = AMR level (levels >= 3: set 3) */
#define PCFD_NN_BIASES 256       /* 4 sets x 4 layers x 16 */

typedef struct pcfd_learned_stats {
    int pairs;                /* stored training pairs */
    int train_steps;          /* Adam steps taken so far */
    double loss_first;        /* mean loss of the first / last step of the last train call */
    double loss_last;
    double collect_ms, train_ms;
    int bricks;               /* bricks the network corrects */
    int64_t cells;            /* their active cells */
} pcfd_learned_stats;

/* Advance a copy of the current state by one transport step, solve its projection with plain
 * MGPCG warm-started from the current pressure (as a step does), and store the post-V-cycle
 * residuals r'_k = r_k - A M_mg(r_k) of its first
 * `max_pairs` iterations (with their normalisation sum r'^2/h).  The solver state is left
 * unchanged.  Residuals beyond learned_capacity replace the oldest. */
pcfd_status pcfd_learned_collect(pcfd_solver* solver, int max_pairs, pcfd_learned_stats* stats);
/* `steps` Adam steps (lr, betas 0.9/0.999), each over `batch` stored residuals taken in turn
 * (loss: energy of the remaining error divided by sum r'^2/h, averaged over the batch);
 * enables the network. */
pcfd_status pcfd_learned_train(pcfd_solver* solver, int steps, int batch, double learning_rate,
                               pcfd_learned_stats* stats);
/* FP32 master weights [PCFD_NN_WEIGHTS] and biases [PCFD_NN_BIASES]; set enables the network. */
pcfd_status pcfd_learned_get_weights(pcfd_solver* solver, float* weights, float* biases);
pcfd_status pcfd_learned_set_weights(pcfd_solver* solver, const float* weights, const float* biases);
/* 1 enables, 0 disables the network (the solver then runs plain MGPCG in flexible PCG). */
pcfd_status pcfd_learned_enable(pcfd_solver* solver, int on);
/* Verification hooks: raw network output y[dofs] and brick scales s[bricks] for an input
 * residual r (every brick, the halo from every neighbour); loss and gradient (weights then
 * biases) of one post-V-cycle residual over the network's bricks with the training kernels
 * (input read as FP32), without an update. */
pcfd_status pcfd_learned_forward(pcfd_solver* solver, const double* r, float* y, float* s);
pcfd_status pcfd_learned_gradient(pcfd_solver* solver, const double* residual, double* loss, float* grad_weights,
                                  float* grad_biases);

#ifdef __cplusplus
}
#endif
#endif /* PARACFD_H */
