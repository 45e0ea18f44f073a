// Cell kernels of one solver step on the brick layout: CFL outflow rate, MUSCL transport,
// explicit diffusion, face-velocity reconstruction, divergence / RHS, face correction,
// pressure impulse and the surface-load integral.  R is the field type (double or float);
// pressure and every reduction stay in FP64.
//
// All cell kernels run one 512-lane workgroup per brick and evaluate each face from both
// sides with the same arithmetic (gather form), so transport and impulse are conservative
// without atomics and bitwise reproducible.
#pragma once

#include "device.hpp"

namespace pcfd {

template <class R>
struct FieldsView {
    R* v[3];        // cell velocity
    R* f[3];        // velocity on the +x/+y/+z face of each cell (and the x-max outlet face)
    R* fe;          // compact-edge face velocity (normal component along the edge axis)
};

template <class R>
__device__ __forceinline__ R minmod(R a, R b) {
    if (a * b <= R(0)) return R(0);
    return a >= R(0) ? fmin(a, b) : fmax(a, b);
}

// Upwind MUSCL-minmod face value with the Courant-dependent slope factor
// 0.5 (1 - min(1, C)) (amr_advection.cu:62-75).
template <class R>
__device__ __forceinline__ R muscl_value(R swept, R pred, R va, R vb, R vl, bool has_l, R vu, bool has_u) {
    // upwind side selected before the one minmod: up ? va + pred mm(va - vl, vb - va)
    // : vb - pred mm(vb - va, vu - vb), with the same operations on the selected operands
    const bool up = swept >= R(0);
    const R dab = vb - va;
    const R s1 = up ? va - vl : dab, s2 = up ? dab : vu - vb;
    const R base = up ? va : vb, p = up ? pred : -pred;
    return (up ? has_l : has_u) ? base + p * minmod(s1, s2) : base;
}

// true where 1/h is exact (h a power of two, h and 1/h normal): x * (1/h) == x / h then
template <class R>
__device__ __forceinline__ bool exact_reciprocal(R h) {
    if constexpr (sizeof(R) == 4) {
        const uint32_t u = __float_as_uint(h);
        const uint32_t e = (u >> 23) & 0xffu;
        return (u & 0x807fffffu) == 0u && e > 1u && e < 253u;
    } else {
        const uint64_t u = static_cast<uint64_t>(__double_as_longlong(h));
        const uint64_t e = (u >> 52) & 0x7ffu;
        return (u & 0x800fffffffffffffull) == 0ull && e > 1u && e < 2045u;
    }
}

__device__ __forceinline__ double cell_h(const MeshView& M, int b) { return M.h[M.brick_level[b]]; }

// dof across face (axis, sign) of cell q of brick b from the brick record (a scalar load), -1
// where the brick has no same-level neighbour on that side
__device__ __forceinline__ int64_t face_dof(const MeshView& M, int b, int q, int axis, int sign) {
    const int stride = 1 << (3 * axis);
    const int c = ((q >> (3 * axis)) & 7) + sign;
    const int nb = M.meta[b].nbr[2 * axis + (sign > 0 ? 1 : 0)];
    const int64_t across = nb >= 0 ? static_cast<int64_t>(nb) * 512 + q - sign * 7 * stride : -1;
    return c >= 0 && c < 8 ? static_cast<int64_t>(b) * 512 + q + sign * stride : across;
}

// the six face velocities of cell q: its +x/+y/+z faces and the +faces of its -x/-y/-z
// neighbours (0 where there is no neighbour brick); loaded without waiting for the mask
template <class R>
struct FaceVel {
    R up[3], dn[3];
};
template <class R>
__device__ __forceinline__ FaceVel<R> load_faces(const MeshView& M, const FieldsView<R>& F, int b, int q) {
    FaceVel<R> f;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        f.up[axis] = F.f[axis][a];
        const int64_t o = face_dof(M, b, q, axis, -1);
        f.dn[axis] = load_if(F.f[axis], o >= 0, o, a, R(0));
    }
    return f;
}

// the signed compact-edge fluxes area * fe of a special cell's incidences, summed in
// incidence order (outgoing positive): groups of four, the edge ids of a group loaded
// together, then their areas and face velocities; `op(acc, flux)` folds one term
template <class R, class Op>
__device__ __forceinline__ void for_each_edge_flux(const MeshView& M, const FieldsView<R>& F, int sp, Op visit) {
    const int k0 = M.special_inc_off[sp], k1 = M.special_inc_off[sp + 1];
    for (int k = k0; k < k1; k += 4) {
        int raw[4];
        double ar[4], fv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) raw[j] = load_if(M.special_inc, k + j < k1, k + j, k, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = raw[j] >= 0 ? raw[j] : ~raw[j];
            ar[j] = load_if(M.earea, k + j < k1, e, e, 0.0);
            fv[j] = double(load_if(F.fe, k + j < k1, e, e, R(0)));
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < k1) {
                const double flux = ar[j] * fv[j];
                visit(raw[j] >= 0 ? flux : -flux);
            }
    }
}

template <class R, class Op>
__device__ __forceinline__ void fold_edge_fluxes(const MeshView& M, const FieldsView<R>& F, int sp, double& acc, Op op) {
    for_each_edge_flux<R>(M, F, sp, [&](double x) { acc = op(acc, x); });
}

// A member of a momentum redistribution group hands its increment to the group (k_redistribute)
// instead of applying it (apply_or_accumulate_redistributed_momentum, amr_advection.cu:118-129).
__device__ __forceinline__ bool redistributed(const MeshView& M, int sp) {
    return M.ngroups > 0 && sp >= 0 && M.rgroup[sp] >= 0;
}
template <class T>
__device__ __forceinline__ void hand_over(const MeshView& M, int sp, const T d[3]) {
    for (int c = 0; c < 3; ++c) M.rdelta[3 * static_cast<int64_t>(sp) + c] = double(d[c]);
}

// max |component| of a velocity, infinite where a component is not finite
__device__ __forceinline__ double abs_max3(double x, double y, double z) {
    const double w = fmax(fabs(x), fmax(fabs(y), fabs(z)));
    return (!isfinite(x) || !isfinite(y) || !isfinite(z)) ? INFINITY : w;
}

// max of non-negative doubles through their bit patterns (ordered like the values); the result
// does not depend on the order of the updates
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// slot[key] = max(slot[key], w) for every lane with key >= 0 (w >= 0): the lanes of a wave
// that share a key combine first, one atomic per distinct key (special cells are ordered by
// brick, so a wave holds one to a few).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_max_to_slots(double* slot, int key, double w) {
    for (;;) {
        const unsigned long long pending = __ballot(key >= 0);
        if (pending == 0) break;
        const int lead = __ffsll(static_cast<long long>(pending)) - 1;
        const int k = __shfl(key, lead);
        double v = key == k ? w : 0.0;
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
        if (static_cast<int>(__lane_id()) == lead) atomic_max_nonneg(slot + k, v);
        if (key == k) key = -1;
    }
}

// Every member of group g takes the group's summed increment (member order) over the group's
// volume (finalize_redistributed_state_kernel, amr_advection.cu:142-149); out holds the
// members' values before the increment.
// With `vmax_part`, each member's |v| also raises its brick's slot (after k_impulse stored it).
template <class R>
__global__ __launch_bounds__(256) void k_redistribute(MeshView M, R* __restrict__ ox, R* __restrict__ oy,
                                                      R* __restrict__ oz, double* vmax_part) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= M.ngroups) return;
    double S[3] = {0.0, 0.0, 0.0};
    for (int k = M.roff[g]; k < M.roff[g + 1]; ++k) {
        const int64_t s = M.rmember[k];
        for (int c = 0; c < 3; ++c) S[c] += M.rdelta[3 * s + c];
    }
    const double iv = 1.0 / M.rvolume[g];
    R* out[3] = {ox, oy, oz};
    for (int k = M.roff[g]; k < M.roff[g + 1]; ++k) {
        const int64_t a = M.special_dof[M.rmember[k]];
        R vn[3];
        for (int c = 0; c < 3; ++c) out[c][a] = vn[c] = R(double(out[c][a]) + S[c] * iv);
        if (vmax_part) atomic_max_nonneg(vmax_part + (a >> 9), abs_max3(double(vn[0]), double(vn[1]), double(vn[2])));
    }
}

// Least-squares gradient stencils of the special cells: neighbours across every open face
// (structured faces, area h^2, and compact edges, open area), weight area / |d|^2 on the
// fluid centroids, the pseudo-inverted 3x3 normal matrix (amr_advection.cu:822-865).
// g = inv * sum_k wd_k (f(nbr_k) - f(cell)).
struct SmagView {
    const int* off;       // nspecial + 1
    const int* nbr;       // neighbour dof per stencil entry
    const double* wd;     // 3 per entry: weight * displacement
    const double* inv;    // 9 per special cell
};

// velocity gradient at every special cell, 9 per cell (d v_c / d x_r at 3 c + r), one lane per
// cell (accumulate_special_cell_gradient + solve_special_cell_gradient, amr_advection.cu:206-216)
template <class R>
__global__ __launch_bounds__(256) void k_grad_v(MeshView M, FieldsView<R> F, SmagView G, R* __restrict__ g) {
    const int sp = blockIdx.x * blockDim.x + threadIdx.x;
    if (sp >= M.nspecial) return;
    const int64_t a = M.special_dof[sp];
    const int k0 = G.off[sp], k1 = G.off[sp + 1];
    const double va[3] = {double(F.v[0][a]), double(F.v[1][a]), double(F.v[2][a])};
    double rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // stencil entries in groups of four: indices and weights of a group loaded together,
    // then the neighbours' values, summed in stencil order
#pragma unroll 2
    for (int k = k0; k < k1; k += 4) {
        int64_t o[4];
        double w[4][3], v[4][3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = k + j < k1;
            o[j] = load_if(G.nbr, ok, k + j, k, 0);
            o[j] = ok ? o[j] : a;
            for (int r = 0; r < 3; ++r) w[j][r] = load_if(G.wd, ok, 3 * (k + j) + r, 3 * k + r, 0.0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            for (int c = 0; c < 3; ++c) v[j][c] = double(F.v[c][o[j]]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < k1)
                for (int c = 0; c < 3; ++c) {
                    const double dv = v[j][c] - va[c];
                    rhs[3 * c] += w[j][0] * dv; rhs[3 * c + 1] += w[j][1] * dv; rhs[3 * c + 2] += w[j][2] * dv;
                }
    }
    const double* I = G.inv + 9 * static_cast<int64_t>(sp);
    R* out = g + 9 * static_cast<int64_t>(sp);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            out[3 * c + r] = R(I[3 * r] * rhs[3 * c] + I[3 * r + 1] * rhs[3 * c + 1] + I[3 * r + 2] * rhs[3 * c + 2]);
}

// pressure gradient at every special cell, 3 per cell, with the same stencils
// (reconstruct_scalar_pressure_gradient_kernel, amr_advection.cu:400-405)
__global__ __launch_bounds__(256) void k_grad_p(MeshView M, SmagView G, const double* __restrict__ p,
                                                double* __restrict__ g) {
    const int sp = blockIdx.x * blockDim.x + threadIdx.x;
    if (sp >= M.nspecial) return;
    const int64_t a = M.special_dof[sp];
    const int k0 = G.off[sp], k1 = G.off[sp + 1];
    const double pa = p[a];
    double rhs[3] = {0, 0, 0};
#pragma unroll 2
    for (int k = k0; k < k1; k += 4) {  // groups of four, as k_grad_v
        int64_t o[4];
        double w[4][3], v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = k + j < k1;
            o[j] = load_if(G.nbr, ok, k + j, k, 0);
            o[j] = ok ? o[j] : a;
            for (int r = 0; r < 3; ++r) w[j][r] = load_if(G.wd, ok, 3 * (k + j) + r, 3 * k + r, 0.0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = p[o[j]];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < k1) {
                const double dp = v[j] - pa;
                rhs[0] += w[j][0] * dp; rhs[1] += w[j][1] * dp; rhs[2] += w[j][2] * dp;
            }
    }
    const double* I = G.inv + 9 * static_cast<int64_t>(sp);
    for (int r = 0; r < 3; ++r)
        g[3 * static_cast<int64_t>(sp) + r] = I[3 * r] * rhs[0] + I[3 * r + 1] * rhs[1] + I[3 * r + 2] * rhs[2];
}

// ---------------------------------------------------------------------------------------
// CFL: max over cells of sum(max(outflow, 0)) / V  (amr_advection.cu:169-186) -> dt
// ---------------------------------------------------------------------------------------
// outflow of a cell through its structured faces and the domain's inflow / outlet faces
template <class R>
__device__ __forceinline__ double cell_outflow(const MeshView& M, const FaceVel<R>& f, uint16_t bf, int q, uint8_t m,
                                               double A) {
    double out = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) out += fmax(A * double(f.up[axis]), 0.0);
        if (m & (kMX << axis)) out += fmax(-A * double(f.dn[axis]), 0.0);
    }
    if ((q & 7) == 0 && (bf & kXMIN)) out += fmax(-A * M.speed, 0.0);
    if (M.outlet && (q & 7) == 7 && (bf & kXMAX)) out += fmax(A * double(f.up[0]), 0.0);
    return out;
}

// Two cells per lane (kBL lanes per brick).  With `vmax`, also max |v| over the active cells
// (a non-finite component makes it infinite) into partials[nbricks + b].
template <class R>
__global__ __launch_bounds__(kBL) void k_outflow_rate(MeshView M, FieldsView<R> F, double* partials, int vmax) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const double h = M.meta[b].h, A = h * h;
    const uint16_t bf = M.meta[b].flags;
    uint8_t m[2];
    FaceVel<R> cf[2];
    R v[2][3];
    int sp[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        m[k] = M.mask[base + q];
        cf[k] = load_faces<R>(M, F, b, q);
        for (int c = 0; c < 3; ++c) v[k][c] = vmax ? F.v[c][base + q] : R(0);
    }
    for (int k = 0; k < 2; ++k) sp[k] = special_rank(M, b, t + kBL * k, (m[k] & kSpecial) != 0);
    double rate = 0.0, vm = 0.0;
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        if (!(m[k] & kActive)) continue;
        double out = cell_outflow<R>(M, cf[k], bf, q, m[k], A);
        double V = h * h * h;
        if (sp[k] >= 0) {
            V = M.special_volume[sp[k]];
            fold_edge_fluxes<R>(M, F, sp[k], out, [](double acc, double f) { return acc + fmax(f, 0.0); });
        }
        rate = fmax(rate, out / V);
        vm = fmax(vm, abs_max3(double(v[k][0]), double(v[k][1]), double(v[k][2])));
    }
    store_partial(partials, b, block_max<kBL>(rate));
    if (vmax) {
        const double mv = block_max<kBL>(vm);
        if (threadIdx.x == 0) partials[M.nbricks + b] = mv;
    }
}

// ---------------------------------------------------------------------------------------
// Conservative MUSCL transport of the cell momentum (amr_advection.cu:82-112, 113-118)
// ---------------------------------------------------------------------------------------
// Compact faces (coarse/fine, partially open, and the fully open faces of embedded nodes) carry
// the donor's value reconstructed to the face centroid with its least-squares gradient g
// (velocity before the transport), scaled by 1 - C, C = min(1, |swept| / V_donor), and clamped
// to the two cells' values (compact_muscl_cell_face_value, amr_advection.cu:94-105).
// Velocity tile of a brick with a two-cell face halo for the MUSCL stencil: slot
// (k+2) 144 + (j+2) 12 + (i+2) of a 12^3 box.  The brick's cells and its six face slabs, two
// cells deep, are filled (zero where there is no same-level neighbour brick); edges and corners
// are never read.  A slab cell is read only behind a structured face, whose neighbour exists.
constexpr int kTile2 = 1728;
__device__ __forceinline__ int tile2_index(int i, int j, int k) { return (k + 2) * 144 + (j + 2) * 12 + (i + 2); }

// every lane: its own cell and, for the 768 halo slots, halo slot q and q + 512; wave w serves
// one face and layer, so the neighbour brick is a scalar read of the brick record
template <class R>
__device__ __forceinline__ void stage_tile2(const MeshView& M, const FieldsView<R>& F, int b, R (*vt)[kTile2],
                                            uint8_t* mt, uint8_t m_own, R own[3]) {
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    for (int c = 0; c < 3; ++c) own[c] = F.v[c][a];
    R hv[2][3];
    uint8_t hm[2] = {0, 0};
    int hidx[2] = {-1, -1};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        for (int c = 0; c < 3; ++c) hv[r][c] = R(0);
        const int h = q + 512 * r;
        const bool slot = h < 768;  // uniform over a wave: the other waves skip these loads
        if (!slot) continue;
        const int face = __builtin_amdgcn_readfirstlane(h >> 7), layer = (h >> 6) & 1, t = h & 63;
        const int axis = face >> 1, up = face & 1, a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        int cc[3];
        cc[a1] = t & 7;
        cc[a2] = t >> 3;
        cc[axis] = up ? 8 + layer : -1 - layer;
        hidx[r] = tile2_index(cc[0], cc[1], cc[2]);
        const int nb = M.meta[b].nbr[face];
        if (nb >= 0) {
            cc[axis] = up ? layer : 7 - layer;
            const int64_t o = static_cast<int64_t>(nb) * 512 + cc[0] + 8 * cc[1] + 64 * cc[2];
            for (int c = 0; c < 3; ++c) hv[r][c] = F.v[c][o];
            hm[r] = M.mask[o];
        }
    }
    const int ti = tile2_index(q & 7, (q >> 3) & 7, q >> 6);
    for (int c = 0; c < 3; ++c) vt[c][ti] = own[c];
    mt[ti] = m_own;
#pragma unroll
    for (int r = 0; r < 2; ++r)
        if (hidx[r] >= 0) {
            for (int c = 0; c < 3; ++c) vt[c][hidx[r]] = hv[r][c];
            mt[hidx[r]] = hm[r];
        }
    __syncthreads();
}

// Regular faces read the MUSCL stencil (two cells each side) from the LDS tile.  Each face
// value is computed once: every cell reconstructs its own +x/+y/+z faces (where it or its
// neighbour inside the brick has that face), and waves 0-2 reconstruct the 64 -x, -y and -z
// faces on the brick's low sides (one axis per wave, lane t: cell (t & 7, t >> 3) of the face).
// The face values then go to LDS over the velocity tile, and every cell sums its faces in the
// order +x, -x, +y, -y, +z, -z with the same expressions as a cell reconstructing all six itself
// (bit-identical).  Special cells hand their sum over to k_advect_compact, which adds their
// compact-face fluxes.
template <class R>
__global__ __launch_bounds__(512) void k_advect(MeshView M, FieldsView<R> F, R* __restrict__ ox, R* __restrict__ oy,
                                                R* __restrict__ oz, const DevScalars* S, R* __restrict__ dsp) {
    const int b = brick_of_block(M.xcd_chunk), q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    // every independent load before the tile barrier: the mask, the cell's +x/+y/+z face
    // velocities and those of its -x/-y/-z neighbours (their +faces), the face velocity of the
    // low-side face this lane reconstructs, the instance's timestep and cell size
    const uint8_t m = M.mask[a];
    R fo[3], fn[3];
    for (int axis = 0; axis < 3; ++axis) {
        const int stride = 1 << (3 * axis);
        fo[axis] = F.f[axis][a];
        int64_t o = a - stride;
        if (axis_coord(q, axis) == 0) {
            const int nb = M.meta[b].nbr[2 * axis];
            o = nb >= 0 ? static_cast<int64_t>(nb) * 512 + q + 7 * stride : -1;
        }
        fn[axis] = o >= 0 ? F.f[axis][o] : R(0);
    }
    const bool low = q < 192;  // waves 0-2
    const int baxis = __builtin_amdgcn_readfirstlane(q >> 6), bt = q & 63;
    int tb = 0;
    R ub = R(0);
    if (low) {
        const int a1 = baxis == 0 ? 1 : 0, a2 = baxis == 2 ? 1 : 2;
        int cc[3];
        cc[a1] = bt & 7;
        cc[a2] = bt >> 3;
        cc[baxis] = 0;
        tb = tile2_index(cc[0], cc[1], cc[2]);
        const int nb = M.meta[b].nbr[2 * baxis];
        if (nb >= 0) ub = F.f[baxis][static_cast<int64_t>(nb) * 512 + cc[0] + 8 * cc[1] + 64 * cc[2] + 7 * (1 << (3 * baxis))];
    }
    const double dt_inst = S[M.meta[b].inst].dt;
    const R h = R(M.meta[b].h), A = h * h;
    const uint16_t bf = M.meta[b].flags;
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    __shared__ R vt[3][kTile2];
    __shared__ uint8_t mt[kTile2];
    R va[3];
    stage_tile2<R>(M, F, b, vt, mt, m, va);
    const R dt = R(dt_inst);
    const int ta = tile2_index(q & 7, (q >> 3) & 7, q >> 6);
    R fp[3][3];
    R fb[3] = {R(0), R(0), R(0)};
    const R ih = R(1) / h;
    // P2: the Courant fraction |u| dt / h as a product with the exact 1/h (same value)
    auto reconstruct = [&](auto P2) {
        auto courant = [&](R un) -> R {
            if constexpr (decltype(P2)::value) return fabs(un) * dt * ih;
            else return fabs(un) * dt / h;
        };
        // own +faces
        for (int axis = 0; axis < 3; ++axis) {
            const int st = axis == 0 ? 1 : (axis == 1 ? 12 : 144);
            const int b1 = ta + st;
            const uint8_t mb = mt[b1];
            for (int c = 0; c < 3; ++c) fp[axis][c] = R(0);
            const bool inner = axis_coord(q, axis) < 7;
            if ((m & (kPX << axis)) || (inner && (mb & (kMX << axis)))) {
                const bool has_l = (m & (kMX << axis)) != 0, has_u = (mb & (kPX << axis)) != 0;
                const int l = has_l ? ta - st : ta;
                const int u = has_u ? b1 + st : b1;
                const R un = fo[axis];
                const R swept = dt * A * un, pred = R(0.5) * (R(1) - fmin(R(1), courant(un)));
                for (int c = 0; c < 3; ++c)
                    fp[axis][c] = muscl_value(swept, pred, va[c], vt[c][b1], vt[c][l], has_l, vt[c][u], has_u);
            }
        }
        // low-side faces (b0 in the neighbour brick, cell tb)
        if (low) {
            const int st = baxis == 0 ? 1 : (baxis == 1 ? 12 : 144);
            const uint8_t mc = mt[tb];
            if (mc & (kMX << baxis)) {
                const int b0 = tb - st;
                const uint8_t m0 = mt[b0];
                const bool has_l = (m0 & (kMX << baxis)) != 0, has_u = (mc & (kPX << baxis)) != 0;
                const int l = has_l ? b0 - st : b0;
                const int u = has_u ? tb + st : tb;
                const R swept = dt * A * ub, pred = R(0.5) * (R(1) - fmin(R(1), courant(ub)));
                for (int c = 0; c < 3; ++c)
                    fb[c] = muscl_value(swept, pred, vt[c][b0], vt[c][tb], vt[c][l], has_l, vt[c][u], has_u);
            }
        }
    };
    if (exact_reciprocal(h)) reconstruct(std::true_type{});
    else reconstruct(std::false_type{});
    __syncthreads();  // every tile read done: the face values go over the tile
    R* fv = &vt[0][0];  // +faces at (3 axis + c) 512 + cell, low-side faces at 4608 + (3 axis + c) 64 + t
    for (int axis = 0; axis < 3; ++axis)
        for (int c = 0; c < 3; ++c) fv[(3 * axis + c) * 512 + q] = fp[axis][c];
    if (low)
        for (int c = 0; c < 3; ++c) fv[4608 + (3 * baxis + c) * 64 + bt] = fb[c];
    __syncthreads();
    R* out[3] = {ox, oy, oz};
    if (!(m & kActive)) {
        for (int c = 0; c < 3; ++c) out[c][a] = va[c];
        return;
    }
    R d[3] = {R(0), R(0), R(0)};
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) {  // face (a, b1)
            const R swept = dt * A * fo[axis];
            for (int c = 0; c < 3; ++c) d[c] -= swept * fp[axis][c];
        }
        if (m & (kMX << axis)) {  // face (b0, a)
            const int ca = axis_coord(q, axis);
            const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
            const int slot = ca > 0 ? (3 * axis) * 512 + q - (1 << (3 * axis))
                                    : 4608 + (3 * axis) * 64 + axis_coord(q, a1) + 8 * axis_coord(q, a2);
            const int cstride = ca > 0 ? 512 : 64;
            const R swept = dt * A * fn[axis];
            for (int c = 0; c < 3; ++c) d[c] += swept * fv[slot + c * cstride];
        }
    }
    if ((q & 7) == 0 && (bf & kXMIN)) {
        const R U = R(M.speed), swept = dt * A * U;
        if (swept >= R(0)) d[0] += swept * U;
        else for (int c = 0; c < 3; ++c) d[c] += swept * va[c];
    }
    if (M.outlet && (q & 7) == 7 && (bf & kXMAX)) {
        const R swept = dt * A * fo[0];
        if (swept >= R(0)) for (int c = 0; c < 3; ++c) d[c] -= swept * va[c];
        else d[0] -= swept * R(M.speed);
    }
    if (sp >= 0) {  // compact faces, volume and redistribution in k_advect_compact
        for (int c = 0; c < 3; ++c) dsp[3 * static_cast<int64_t>(sp) + c] = d[c];
        return;
    }
    const R V = h * h * h;
    const R iv = R(1) / V;
    for (int c = 0; c < 3; ++c) out[c][a] = va[c] + d[c] * iv;
}

// Compact faces of the active special cells, one lane per cell, continuing the sum that
// k_advect handed over: the donor's value reconstructed to the face centroid with its
// least-squares gradient, scaled by 1 - C and clamped to the two cells' values; incidences in
// pairs (indices, then records, then values), summed in incidence order.
template <class R>
__global__ __launch_bounds__(256) void k_advect_compact(MeshView M, FieldsView<R> F, R* __restrict__ ox,
                                                        R* __restrict__ oy, R* __restrict__ oz, const DevScalars* S,
                                                        const R* __restrict__ G, const R* __restrict__ dsp) {
    const int sp = blockIdx.x * blockDim.x + threadIdx.x;
    if (sp >= M.nspecial) return;
    const int64_t a = M.special_dof[sp];
    const uint8_t m = M.mask[a];
    const int k0 = M.special_inc_off[sp], k1 = M.special_inc_off[sp + 1];
    const R dt = R(S[M.meta[a >> 9].inst].dt);
    const R V = R(M.special_volume[sp]);
    R va[3], d[3];
    for (int c = 0; c < 3; ++c) {
        va[c] = F.v[c][a];
        d[c] = dsp[3 * static_cast<int64_t>(sp) + c];
    }
    if (!(m & kActive)) return;  // k_advect wrote the cell
    for (int k = k0; k < k1; k += 2) {
        int raw[2], e[2];
        for (int j = 0; j < 2; ++j) {
            raw[j] = load_if(M.special_inc, k + j < k1, k + j, k, 0);
            e[j] = raw[j] >= 0 ? raw[j] : ~raw[j];
        }
        R swept[2];
        int64_t ia[2], ib[2];
        int es[2][2];
        for (int j = 0; j < 2; ++j) {
            swept[j] = dt * R(M.earea[e[j]]) * F.fe[e[j]];
            ia[j] = M.ea[e[j]];
            ib[j] = M.eb[e[j]];
            es[j][0] = M.esp[2 * e[j]];
            es[j][1] = M.esp[2 * e[j] + 1];
        }
        R vA[2][3], vB[2][3], g[2][9], C[2], dx[2][3];
        int ds[2];
        for (int j = 0; j < 2; ++j) {
            const bool lower = swept[j] >= R(0);
            ds[j] = es[j][lower ? 0 : 1];
            for (int c = 0; c < 3; ++c) {
                vA[j][c] = F.v[c][ia[j]];
                vB[j][c] = F.v[c][ib[j]];
            }
            C[j] = R(1);
            for (int r = 0; r < 3; ++r) dx[j][r] = R(0);
            for (int r = 0; r < 9; ++r) g[j][r] = R(0);
            if (ds[j] >= 0) {
                C[j] = fmin(R(1), fabs(swept[j]) / R(M.special_volume[ds[j]]));
                const double* dd = M.edisp + 6 * static_cast<int64_t>(e[j]) + (lower ? 0 : 3);
                for (int r = 0; r < 3; ++r) dx[j][r] = R(dd[r]);
                for (int r = 0; r < 9; ++r) g[j][r] = G[9 * static_cast<int64_t>(ds[j]) + r];
            }
        }
        for (int j = 0; j < 2; ++j) {
            if (k + j >= k1) break;
            const bool lower = swept[j] >= R(0);
            for (int c = 0; c < 3; ++c) {
                const R fa = vA[j][c], fb = vB[j][c];
                R fv = lower ? fa : fb;
                if (ds[j] >= 0) {
                    const R* gc = g[j] + 3 * c;
                    const R rec = fv + (R(1) - C[j]) * (gc[0] * dx[j][0] + gc[1] * dx[j][1] + gc[2] * dx[j][2]);
                    fv = fmax(fmin(fa, fb), fmin(fmax(fa, fb), rec));
                }
                const R flux = swept[j] * fv;
                d[c] += raw[j] >= 0 ? -flux : flux;
            }
        }
    }
    R* out[3] = {ox, oy, oz};
    if (redistributed(M, sp)) {
        hand_over(M, sp, d);
        for (int c = 0; c < 3; ++c) out[c][a] = va[c];
        return;
    }
    const R iv = R(1) / V;
    for (int c = 0; c < 3; ++c) out[c][a] = va[c] + d[c] * iv;
}

// ---------------------------------------------------------------------------------------
// Explicit diffusion substep (amr_advection.cu:233-241): conductance h on structured faces,
// open area / distance on compact edges.
// ---------------------------------------------------------------------------------------
// Two cells per lane; the three velocity components with a one-cell face halo in LDS; compact
// edges in groups of four incidences.
template <class R>
__global__ __launch_bounds__(kBL) void k_diffuse(MeshView M, const R* __restrict__ vx, const R* __restrict__ vy,
                                                 const R* __restrict__ vz, R* __restrict__ ox, R* __restrict__ oy,
                                                 R* __restrict__ oz, const DevScalars* S, int step) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const R* in[3] = {vx, vy, vz};
    R* out[3] = {ox, oy, oz};
    uint8_t m[2];
    for (int k = 0; k < 2; ++k) m[k] = M.mask[base + t + kBL * k];
    TileFetch2<R> tf[3];
    for (int c = 0; c < 3; ++c) tf[c] = fetch_tile_x2<R>(M, b, [&](int64_t i) -> R { return in[c][i]; });
    const DevScalars& I = S[M.meta[b].inst];
    const double dts = I.dts;
    const R h = R(M.meta[b].h);
    int sp[2];
    for (int k = 0; k < 2; ++k) sp[k] = special_rank(M, b, t + kBL * k, (m[k] & kSpecial) != 0);
    if (step >= I.subs) {  // this instance needs fewer substeps (the whole brick)
        for (int k = 0; k < 2; ++k) {
            for (int c = 0; c < 3; ++c) out[c][base + t + kBL * k] = tf[c].own[k];
            if (redistributed(M, sp[k])) {
                const double zero[3] = {0.0, 0.0, 0.0};
                hand_over(M, sp[k], zero);
            }
        }
        return;
    }
    __shared__ R tv[3][kTileN];
    for (int c = 0; c < 3; ++c) stage_tile_x2<R>(tv[c], tf[c], false);
    __syncthreads();
    const R kf = R(dts * M.nu) * h;
    for (int kc = 0; kc < 2; ++kc) {
        const int q = t + kBL * kc;
        const int64_t a = base + q;
        R va[3];
        for (int c = 0; c < 3; ++c) va[c] = tf[c].own[kc];
        if (!(m[kc] & kActive)) {
            for (int c = 0; c < 3; ++c) out[c][a] = va[c];
            continue;
        }
        R d[3] = {R(0), R(0), R(0)};
        const int ta = tile_index(q & 7, (q >> 3) & 7, q >> 6);
        for (int axis = 0; axis < 3; ++axis)
            for (int sg = -1; sg <= 1; sg += 2) {
                if (!(m[kc] & ((sg > 0 ? kPX : kMX) << axis))) continue;
                const int o = ta + sg * (axis == 0 ? 1 : (axis == 1 ? 10 : 100));
                for (int c = 0; c < 3; ++c) d[c] += kf * (tv[c][o] - va[c]);
            }
        R V = h * h * h;
        const int spc = sp[kc];
        if (spc >= 0) {
            V = R(M.special_volume[spc]);
            const int j0 = M.special_inc_off[spc], j1 = M.special_inc_off[spc + 1];
            for (int j = j0; j < j1; j += 4) {
                int raw[4];
                int64_t o[4];
                double dc[4];
                R vo[4][3];
#pragma unroll
                for (int u = 0; u < 4; ++u) raw[u] = load_if(M.special_inc, j + u < j1, j + u, j, 0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = raw[u] >= 0 ? raw[u] : ~raw[u];
                    const bool ok = j + u < j1;
                    const int64_t eb = M.eb[e], ea = M.ea[e];
                    o[u] = ok ? (raw[u] >= 0 ? eb : ea) : a;
                    dc[u] = load_if(M.edcoef, ok, e, e, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    for (int c = 0; c < 3; ++c) vo[u][c] = in[c][o[u]];
                for (int u = 0; u < 4; ++u) {
                    if (j + u >= j1) break;
                    const R ke = R(dts * M.nu * dc[u]);
                    for (int c = 0; c < 3; ++c) d[c] += ke * (vo[u][c] - va[c]);
                }
            }
        }
        if (redistributed(M, spc)) {
            hand_over(M, spc, d);
            for (int c = 0; c < 3; ++c) out[c][a] = va[c];
            continue;
        }
        const R iv = R(1) / V;
        for (int c = 0; c < 3; ++c) out[c][a] = va[c] + d[c] * iv;
    }
}

// ---------------------------------------------------------------------------------------
// Smagorinsky model (amr_advection.cu:198-252, 1339-1345): cell viscosity
// nu = nu_mol + (Cs V^(1/3))^2 sqrt(2 S:S) from the transported velocity, once per step;
// regular cells differentiate across their open structured faces (central, or one-sided
// where one neighbour is missing), special cells use a least-squares gradient whose
// pseudo-inverted normal matrix and weighted displacements are precomputed per special cell
// (stencil CSR sg_off / sg_nbr / sg_wd, inverses sg_inv).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double strain_rate(const double* g) {
    const double sxx = g[0], syy = g[4], szz = g[8];
    const double sxy = 0.5 * (g[1] + g[3]), sxz = 0.5 * (g[2] + g[6]), syz = 0.5 * (g[5] + g[7]);
    return sqrt(fmax(0.0, 2.0 * (sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz))));
}

template <class R>
__global__ __launch_bounds__(512) void k_viscosity(MeshView M, FieldsView<R> F, SmagView G, double nu_mol, double cs,
                                                   R* __restrict__ nu) {
    const int b = brick_of_block(M.xcd_chunk), q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    if (!(m & kActive)) {
        nu[a] = R(0);
        return;
    }
    const double h = cell_h(M, b);
    double g[9];
    if (sp < 0) {
        for (int axis = 0; axis < 3; ++axis) {
            const int64_t lo = (m & (kMX << axis)) ? neighbor_dof(M, b, q, axis, -1) : -1;
            const int64_t up = (m & (kPX << axis)) ? neighbor_dof(M, b, q, axis, 1) : -1;
            for (int c = 0; c < 3; ++c) {
                const R* v = F.v[c];
                double d = 0.0;
                if (lo >= 0 && up >= 0) d = (double(v[up]) - double(v[lo])) / (2.0 * h);
                else if (up >= 0) d = (double(v[up]) - double(v[a])) / h;
                else if (lo >= 0) d = (double(v[a]) - double(v[lo])) / h;
                g[3 * c + axis] = d;
            }
        }
    } else {
        double rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = G.off[sp]; k < G.off[sp + 1]; ++k) {
            const int64_t o = G.nbr[k];
            for (int c = 0; c < 3; ++c) {
                const double dv = double(F.v[c][o]) - double(F.v[c][a]);
                for (int axis = 0; axis < 3; ++axis) rhs[3 * c + axis] += G.wd[3 * k + axis] * dv;
            }
        }
        const double* I = G.inv + 9 * static_cast<int64_t>(sp);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) g[3 * c + r] = I[3 * r] * rhs[3 * c] + I[3 * r + 1] * rhs[3 * c + 1] + I[3 * r + 2] * rhs[3 * c + 2];
    }
    const double V = sp >= 0 ? M.special_volume[sp] : h * h * h;
    const double L = cs * cbrt(V);
    nu[a] = R(nu_mol + L * L * strain_rate(g));
}

// per-cell explicit diffusion rate sum_faces conductance (nu_a + nu_b)/2 / V, block maxima
template <class R>
__global__ __launch_bounds__(512) void k_diff_rate(MeshView M, const R* __restrict__ nu, double* partials) {
    const int b = brick_of_block(M.xcd_chunk), q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    double rate = 0.0;
    if (m & kActive) {
        const double h = cell_h(M, b), na = double(nu[a]);
        double s = 0.0;
        for (int axis = 0; axis < 3; ++axis)
            for (int sg = -1; sg <= 1; sg += 2) {
                if (!(m & ((sg > 0 ? kPX : kMX) << axis))) continue;
                s += h * 0.5 * (na + double(nu[neighbor_dof(M, b, q, axis, sg)]));
            }
        double V = h * h * h;
        if (sp >= 0) {
            V = M.special_volume[sp];
            for (int j = M.special_inc_off[sp]; j < M.special_inc_off[sp + 1]; ++j) {
                const int raw = M.special_inc[j];
                const int e = raw >= 0 ? raw : ~raw;
                const int64_t o = raw >= 0 ? M.eb[e] : M.ea[e];
                s += M.edcoef[e] * 0.5 * (na + double(nu[o]));
            }
        }
        rate = s / V;
    }
    store_partial(partials, b, block_max<512>(rate));
}

// one explicit diffusion substep with the cell viscosity field (face coefficient mean of
// the two cells' viscosities)
template <class R>
__global__ __launch_bounds__(512) void k_diffuse_var(MeshView M, const R* __restrict__ vx, const R* __restrict__ vy,
                                                     const R* __restrict__ vz, R* __restrict__ ox, R* __restrict__ oy,
                                                     R* __restrict__ oz, const R* __restrict__ nu, const DevScalars* S,
                                                     int step) {
    const int b = brick_of_block(M.xcd_chunk), q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    const R* in[3] = {vx, vy, vz};
    R* out[3] = {ox, oy, oz};
    const DevScalars& I = S[M.inst[b]];
    const double dts = I.dts;
    if (!(m & kActive) || step >= I.subs) {
        for (int c = 0; c < 3; ++c) out[c][a] = in[c][a];
        if (redistributed(M, sp)) {
            const double zero[3] = {0.0, 0.0, 0.0};
            hand_over(M, sp, zero);
        }
        return;
    }
    const R h = R(cell_h(M, b));
    const R half = R(0.5), na = nu[a];
    R va[3], d[3] = {R(0), R(0), R(0)};
    for (int c = 0; c < 3; ++c) va[c] = in[c][a];
    for (int axis = 0; axis < 3; ++axis)
        for (int s = -1; s <= 1; s += 2) {
            if (!(m & ((s > 0 ? kPX : kMX) << axis))) continue;
            const int64_t o = neighbor_dof(M, b, q, axis, s);
            const R k = R(dts) * (half * (na + nu[o])) * h;
            for (int c = 0; c < 3; ++c) d[c] += k * (in[c][o] - va[c]);
        }
    R V = h * h * h;
    if (sp >= 0) {
        V = R(M.special_volume[sp]);
        for (int j = M.special_inc_off[sp]; j < M.special_inc_off[sp + 1]; ++j) {
            const int raw = M.special_inc[j];
            const int e = raw >= 0 ? raw : ~raw;
            const int64_t o = raw >= 0 ? M.eb[e] : M.ea[e];
            const R ke = R(dts) * (half * (na + nu[o])) * R(M.edcoef[e]);
            for (int c = 0; c < 3; ++c) d[c] += ke * (in[c][o] - va[c]);
        }
    }
    if (redistributed(M, sp)) {
        hand_over(M, sp, d);
        for (int c = 0; c < 3; ++c) out[c][a] = va[c];
        return;
    }
    const R iv = R(1) / V;
    for (int c = 0; c < 3; ++c) out[c][a] = va[c] + d[c] * iv;
}

// ---------------------------------------------------------------------------------------
// Face reconstruction (amr_advection.cu:257-293): structured face = mean of the two cells'
// normal components; x-max outlet face = interior cell (zero gradient); compact edges use
// the distance weight.
// ---------------------------------------------------------------------------------------
// two cells per lane; the cell's and its +x/+y/+z neighbours' velocities are loaded before
// the mask is tested
template <class R>
__global__ __launch_bounds__(kBL) void k_reconstruct(MeshView M, FieldsView<R> F) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const uint16_t bf = M.meta[b].flags;
    uint8_t m[2];
    R va[2][3], vn[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int64_t a = base + q;
        m[k] = M.mask[a];
        for (int axis = 0; axis < 3; ++axis) {
            va[k][axis] = F.v[axis][a];
            const int64_t o = face_dof(M, b, q, axis, 1);
            vn[k][axis] = load_if(F.v[axis], o >= 0, o, a, R(0));
        }
    }
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int64_t a = base + q;
        if (!(m[k] & kActive)) continue;
        for (int axis = 0; axis < 3; ++axis)
            if (m[k] & (kPX << axis)) F.f[axis][a] = R(0.5) * (va[k][axis] + vn[k][axis]);
        if (M.outlet && (q & 7) == 7 && (bf & kXMAX)) F.f[0][a] = va[k][0];
    }
}

// compact faces: the weighted mean of the two cells' normal components plus the transverse
// correction both cells' gradients support, minmod(g_a . s, g_b . s), s = x_f - ((1-w) c_a + w c_b)
// (reconstruct_compact_cell_flux_kernel, amr_advection.cu:268-293)
template <class R>
__global__ void k_reconstruct_edges(MeshView M, FieldsView<R> F, const R* __restrict__ G) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M.nedges) return;
    const int ax = M.eaxis[e];
    const R w = R(M.ew[e]);
    const double* da = M.edisp + 6 * static_cast<int64_t>(e);
    R skew[3];
    for (int r = 0; r < 3; ++r) skew[r] = (R(1) - w) * R(da[r]) + w * R(da[3 + r]);
    R corr[2] = {R(0), R(0)};
    for (int side = 0; side < 2; ++side) {
        const int sp = M.esp[2 * e + side];
        if (sp < 0) continue;
        const R* gc = G + 9 * static_cast<int64_t>(sp) + 3 * ax;
        corr[side] = gc[0] * skew[0] + gc[1] * skew[1] + gc[2] * skew[2];
    }
    F.fe[e] = (R(1) - w) * F.v[ax][M.ea[e]] + w * F.v[ax][M.eb[e]] + minmod(corr[0], corr[1]);
}

// integrated divergence of the face fluxes around cell a (FP64 accumulation)
template <class R>
__device__ __forceinline__ double cell_divergence_faces(const MeshView& M, const FaceVel<R>& f, uint16_t bf, int q,
                                                        uint8_t m, double h) {
    const double A = h * h;
    double div = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) div += A * double(f.up[axis]);
        if (m & (kMX << axis)) div -= A * double(f.dn[axis]);
    }
    if ((q & 7) == 0 && (bf & kXMIN)) div -= A * M.speed;
    if (M.outlet && (q & 7) == 7 && (bf & kXMAX)) div += A * double(f.up[0]);
    return div;
}

template <class R>
__device__ __forceinline__ double cell_divergence(const MeshView& M, const FieldsView<R>& F, const FaceVel<R>& f,
                                                  uint16_t bf, int q, uint8_t m, int sp, double h) {
    double div = cell_divergence_faces<R>(M, f, bf, q, m, h);
    if (sp >= 0) fold_edge_fluxes<R>(M, F, sp, div, [](double acc, double x) { return acc + x; });
    return div;
}

// RHS = -(rho/dt) * integrated divergence (amr_pressure.cu:255-284); ||b|| reduction; two
// cells per lane
template <class R>
__global__ __launch_bounds__(kBL) void k_rhs(MeshView M, FieldsView<R> F, double* __restrict__ rhs, double* partials,
                                             const DevScalars* S) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const double h = M.meta[b].h;
    const uint16_t bf = M.meta[b].flags;
    const double dt = S[M.meta[b].inst].dt;
    uint8_t m[2];
    FaceVel<R> cf[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        m[k] = M.mask[base + t + kBL * k];
        cf[k] = load_faces<R>(M, F, b, t + kBL * k);
    }
    double sum = 0.0;
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int sp = special_rank(M, b, q, (m[k] & kSpecial) != 0);
        double val = 0.0;
        if (m[k] & kActive) val = -(M.rho / dt) * cell_divergence<R>(M, F, cf[k], bf, q, m[k], sp, h);
        rhs[base + q] = val;
        sum += val * val;
    }
    store_partial(partials, b, block_sum<kBL>(sum));
}

// max |integrated divergence| / V after the correction (FP64 check); two cells per lane
// With `rate_part`, also the outflow rate of k_outflow_rate from the same (final) faces into
// rate_part[b]: the next step's timestep without another pass over the faces.
template <class R>
__global__ __launch_bounds__(kBL) void k_divcheck(MeshView M, FieldsView<R> F, double* partials, double* rate_part) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const double h = M.meta[b].h, A = h * h;
    const uint16_t bf = M.meta[b].flags;
    uint8_t m[2];
    FaceVel<R> cf[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        m[k] = M.mask[base + t + kBL * k];
        cf[k] = load_faces<R>(M, F, b, t + kBL * k);
    }
    double val = 0.0, rate = 0.0;
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int sp = special_rank(M, b, q, (m[k] & kSpecial) != 0);
        if (m[k] & kActive) {
            const double V = sp >= 0 ? M.special_volume[sp] : h * h * h;
            // one walk over the compact edges for both sums (each in incidence order)
            double div = cell_divergence_faces<R>(M, cf[k], bf, q, m[k], h);
            double out = cell_outflow<R>(M, cf[k], bf, q, m[k], A);
            if (sp >= 0)
                for_each_edge_flux<R>(M, F, sp, [&](double x) {
                    div += x;
                    out += fmax(x, 0.0);
                });
            val = fmax(val, fabs(div / V));
            if (rate_part) rate = fmax(rate, out / V);
        }
    }
    store_partial(partials, b, block_max<kBL>(val));
    if (rate_part) store_partial(rate_part, b, block_max<kBL>(rate));
}

// The projection's right-hand side b = -(rho/dt) div (as k_rhs, amr_pressure.cu:255-284) and
// the initial residual r = b - A x (as k_cg_residual) in one pass; b is not stored.  Partials: ||b||^2 at b's slot,
// ||r||^2 at nbricks + b (FIN_BNORM_RESID).  Runs before the solve's flags are reset, so it
// does not test them.
template <class R>
__global__ __launch_bounds__(kBL) void k_rhs_residual(MeshView M, FieldsView<R> F, const double* __restrict__ x,
                                                      double* __restrict__ r, double* partials, const DevScalars* S) {
    PCFD_BRICK2_BEGIN;
    const uint16_t bf = M.meta[b].flags;
    const double dt = S[M.meta[b].inst].dt;
    FaceVel<R> fv[2];
    for (int k = 0; k < 2; ++k) fv[k] = load_faces<R>(M, F, b, t + kBL * k);
    auto val = [&](int64_t c) -> double { return x[c]; };
    const TileFetch2<double> tf = fetch_tile_x2<double>(M, b, val);
    int sp[2];
    CompactFetch<double> cf[2];
    for (int k = 0; k < 2; ++k) {
        sp[k] = special_rank(M, b, t + kBL * k, (m[k] & kSpecial) != 0);
        cf[k] = fetch_compact<double>(M, sp[k], val);
    }
    __shared__ double tile[kTileN];
    stage_tile_x2<double>(tile, tf);
    double sb = 0.0, sr = 0.0;
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        double rb = 0.0, ra = 0.0;
        if (m[k] & kActive) {
            rb = -(M.rho / dt) * cell_divergence<R>(M, F, fv[k], bf, q, m[k], sp[k], bs.h);
            asm volatile("" : "+v"(rb));  // b rounded as stored, not contracted into r = b - A x
            double dg;
            ra = rb - row0f<double>(M, tile, cf[k], val, bs, q, m[k], tf.own[k], dg);
        }
        r[base + q] = ra;
        sb += rb * rb;
        sr += ra * ra;
    }
    store_partial(partials, b, block_sum<kBL>(sb));
    store_partial(partials, M.nbricks + b, block_sum<kBL>(sr));
}

// face correction u -= (dt/rho) dp/dn (amr_pressure.cu:285-310); outlet p = 0 half a cell away
// two cells per lane; the pressures of the cell and its +x/+y/+z neighbours and the face
// velocities are loaded before the mask is tested
template <class R>
__global__ __launch_bounds__(kBL) void k_correct(MeshView M, FieldsView<R> F, const double* __restrict__ p,
                                                 const DevScalars* S) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const double s = S[M.meta[b].inst].dt / M.rho, h = M.meta[b].h;
    const uint16_t bf = M.meta[b].flags;
    uint8_t m[2];
    double pa[2], pn[2][3];
    R fu[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int64_t a = base + q;
        m[k] = M.mask[a];
        pa[k] = p[a];
        for (int axis = 0; axis < 3; ++axis) {
            fu[k][axis] = F.f[axis][a];
            const int64_t o = face_dof(M, b, q, axis, 1);
            pn[k][axis] = o >= 0 ? p[o] : 0.0;
        }
    }
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int64_t a = base + q;
        if (!(m[k] & kActive)) continue;
        R f0 = fu[k][0];
        for (int axis = 0; axis < 3; ++axis) {
            if (!(m[k] & (kPX << axis))) continue;
            const R v = R(double(fu[k][axis]) - s * (pn[k][axis] - pa[k]) / h);
            F.f[axis][a] = v;
            if (axis == 0) f0 = v;
        }
        if (M.outlet && (q & 7) == 7 && (bf & kXMAX)) F.f[0][a] = R(double(f0) + s * pa[k] / (0.5 * h));
    }
}

template <class R>
__global__ void k_correct_edges(MeshView M, FieldsView<R> F, const double* __restrict__ p, const DevScalars* S) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M.nedges) return;
    const double s = S[M.inst[M.ea[e] >> 9]].dt / M.rho;
    F.fe[e] = R(double(F.fe[e]) - s * (p[M.eb[e]] - p[M.ea[e]]) * (M.ecoef[e] / M.earea[e]));
}

// wall force of patch pch for the pressure p_a + g.(x - c_a) on its wall: (p_a - pref) S + D^T g
// with D = int_W (x - c_a) (x) n dA; for a planar patch with centroid x_p this is the
// reference's (p_a + g.(x_p - c_a)) S (amr_advection.cu:391-399, aero_loads.cpp:38-40)
__device__ __forceinline__ void wall_force_p(const MeshView& M, const double* p, const double* gp, int pch, int sp,
                                             double pref, double* F) {
    const double pa = p[M.patch_dof[pch]] - pref;
    const double* D = M.patch_D + 9 * static_cast<int64_t>(pch);
    const double* g = gp + 3 * static_cast<int64_t>(sp);
    for (int c = 0; c < 3; ++c) F[c] = pa * M.patch_S[3 * pch + c] + D[c] * g[0] + D[3 + c] * g[1] + D[6 + c] * g[2];
}

// ---------------------------------------------------------------------------------------
// Pressure impulse on the control volumes (amr_advection.cu:380-405): internal faces carry
// the face pressure, walls the cell pressure, the outlet p = 0, embedded-boundary patches
// the reconstructed wall pressure.  u += dm / V.
// ---------------------------------------------------------------------------------------
// dt: the given value when positive, else the instance's own step dt; with only_converged,
// instances whose projection did not converge are left unchanged.
// Two cells per lane; the pressures of the cell and its six neighbours and the cell velocity
// are loaded before the mask is tested; compact faces in groups of two incidences.
template <class R>
__global__ __launch_bounds__(kBL) void k_impulse(MeshView M, FieldsView<R> F, const double* __restrict__ p,
                                                 const DevScalars* S, double dt_given, int only_converged,
                                                 double* __restrict__ dsp, double* vmax_part) {
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;
    const int64_t base = static_cast<int64_t>(b) * 512;
    const DevScalars& I = S[M.meta[b].inst];
    const double h = M.meta[b].h, A = h * h;
    const uint16_t bf = M.meta[b].flags;
    uint8_t m[2];
    double pa[2], pp[2][3], pm[2][3];
    R v0[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = t + kBL * k;
        const int64_t a = base + q;
        m[k] = M.mask[a];
        pa[k] = p[a];
        for (int axis = 0; axis < 3; ++axis) {
            const int64_t ou = face_dof(M, b, q, axis, 1), od = face_dof(M, b, q, axis, -1);
            pp[k][axis] = load_if(p, ou >= 0, ou, a, 0.0);
            pm[k][axis] = load_if(p, od >= 0, od, a, 0.0);
            v0[k][axis] = F.v[axis][a];
        }
    }
    int sp[2];
    for (int k = 0; k < 2; ++k) sp[k] = special_rank(M, b, t + kBL * k, (m[k] & kSpecial) != 0);
    if (only_converged && !I.converged) {  // block-uniform; k_impulse_compact hands over zeros
        if (vmax_part) {  // the velocities stay: |v|max of every active cell
            double vm = 0.0;
            for (int k = 0; k < 2; ++k)
                if (m[k] & kActive) vm = fmax(vm, abs_max3(double(v0[k][0]), double(v0[k][1]), double(v0[k][2])));
            store_partial(vmax_part, b, block_max<kBL>(vm));
        }
        return;
    }
    const double dt = dt_given > 0.0 ? dt_given : I.dt;
    double vm = 0.0;  // non-special cells; the special ones update the slot atomically later
    const double s = dt / M.rho;
    for (int kc = 0; kc < 2; ++kc) {
        const int q = t + kBL * kc;
        const int64_t a = base + q;
        if (!(m[kc] & kActive)) continue;
        double d[3] = {0.0, 0.0, 0.0};
        for (int axis = 0; axis < 3; ++axis) {
            if (m[kc] & (kPX << axis)) d[axis] -= s * A * 0.5 * (pa[kc] + pp[kc][axis]);
            if (m[kc] & (kMX << axis)) d[axis] += s * A * 0.5 * (pm[kc][axis] + pa[kc]);
        }
        const int i = q & 7, j = (q >> 3) & 7, kz = q >> 6;
        const double bnd = s * A * pa[kc];
        if (i == 0 && (bf & kXMIN)) d[0] += bnd;
        if (i == 7 && (bf & kXMAX) && !M.outlet) d[0] -= bnd;
        if (j == 0 && (bf & kYMIN)) d[1] += bnd;
        if (j == 7 && (bf & kYMAX)) d[1] -= bnd;
        if (kz == 0 && (bf & kZMIN)) d[2] += bnd;
        if (kz == 7 && (bf & kZMAX)) d[2] -= bnd;
        if (sp[kc] >= 0) {  // compact faces, wall, volume and redistribution in k_impulse_compact
            for (int c = 0; c < 3; ++c) dsp[3 * static_cast<int64_t>(sp[kc]) + c] = d[c];
            continue;
        }
        const double iv = 1.0 / (h * h * h);
        R vn[3];
        for (int c = 0; c < 3; ++c) F.v[c][a] = vn[c] = R(double(v0[kc][c]) + d[c] * iv);
        vm = fmax(vm, abs_max3(double(vn[0]), double(vn[1]), double(vn[2])));
    }
    if (vmax_part) store_partial(vmax_part, b, block_max<kBL>(vm));
}

// The special cells' part of the impulse, one lane per cell, continuing the structured-face
// sum that k_impulse handed over: compact faces (weighted face pressure plus
// skew . ((1-w) g_a + w g_b), pressure_compact_cell_impulse_kernel, amr_advection.cu:376-390),
// the wall traction of the cell's patch, the fluid volume and the redistribution hand-over.
template <class R>
__device__ __forceinline__ void impulse_special_cell(const MeshView& M, const FieldsView<R>& F, const double* __restrict__ p,
                                                     const double* __restrict__ gp, const DevScalars* S, double dt_given,
                                                     int only_converged, const double* __restrict__ dsp, int spc,
                                                     int& key, double& w) {
    const int64_t a = M.special_dof[spc];
    const uint8_t m = M.mask[a];
    const DevScalars& I = S[M.meta[a >> 9].inst];
    R v0[3];
    double d[3];
    for (int c = 0; c < 3; ++c) {
        v0[c] = F.v[c][a];
        d[c] = dsp[3 * static_cast<int64_t>(spc) + c];
    }
    const double V = M.special_volume[spc];
    if (only_converged && !I.converged) {
        if (redistributed(M, spc)) {
            const double zero[3] = {0.0, 0.0, 0.0};
            hand_over(M, spc, zero);
        }
        return;
    }
    if (!(m & kActive)) return;
    const double dt = dt_given > 0.0 ? dt_given : I.dt;
    const double s = dt / M.rho;
    // compact faces: weighted face pressure plus skew . ((1-w) g_a + w g_b)
    // (pressure_compact_cell_impulse_kernel, amr_advection.cu:376-390)
    const int t0 = M.special_inc_off[spc], t1 = M.special_inc_off[spc + 1];
    for (int tt = t0; tt < t1; tt += 2) {
        int raw[2], e[2], ia[2], ib[2], sa[2], sb[2], ax[2];
        double w[2], ar[2], pA[2], pB[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            raw[u] = load_if(M.special_inc, tt + u < t1, tt + u, tt, 0);
            e[u] = raw[u] >= 0 ? raw[u] : ~raw[u];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const bool ok = tt + u < t1;
            w[u] = load_if(M.ew, ok, e[u], e[u], 0.0);
            ar[u] = load_if(M.earea, ok, e[u], e[u], 0.0);
            ia[u] = load_if(M.ea, ok, e[u], e[u], 0);
            ib[u] = load_if(M.eb, ok, e[u], e[u], 0);
            sa[u] = load_if(M.esp, ok, 2 * e[u], 2 * e[u], -1);
            sb[u] = load_if(M.esp, ok, 2 * e[u] + 1, 2 * e[u] + 1, -1);
            ax[u] = load_if(M.eaxis, ok, e[u], e[u], int8_t(0));
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            pA[u] = load_if(p, tt + u < t1, ia[u], ia[u], 0.0);
            pB[u] = load_if(p, tt + u < t1, ib[u], ib[u], 0.0);
        }
        for (int u = 0; u < 2; ++u) {
            if (tt + u >= t1) break;
            double pf = (1.0 - w[u]) * pA[u] + w[u] * pB[u];
            if (sa[u] >= 0 && sb[u] >= 0) {
                const double* da = M.edisp + 6 * static_cast<int64_t>(e[u]);
                for (int c = 0; c < 3; ++c)
                    pf += ((1.0 - w[u]) * da[c] + w[u] * da[3 + c]) *
                          ((1.0 - w[u]) * gp[3 * sa[u] + c] + w[u] * gp[3 * sb[u] + c]);
            }
            const double imp = s * ar[u] * pf;
            d[ax[u]] += raw[u] >= 0 ? -imp : imp;
        }
    }
    const int pch = M.special_patch[spc];
    if (pch >= 0) {
        double Fw[3];
        wall_force_p(M, p, gp, pch, spc, 0.0, Fw);
        for (int c = 0; c < 3; ++c) d[c] -= s * Fw[c];
    }
    if (redistributed(M, spc)) {  // k_redistribute writes the velocity (and its |v|)
        hand_over(M, spc, d);
        return;
    }
    const double iv = 1.0 / V;
    R vn[3];
    for (int c = 0; c < 3; ++c) F.v[c][a] = vn[c] = R(double(v0[c]) + d[c] * iv);
    key = static_cast<int>(a >> 9);
    w = abs_max3(double(vn[0]), double(vn[1]), double(vn[2]));
}

template <class R>
__global__ __launch_bounds__(256) void k_impulse_compact(MeshView M, FieldsView<R> F, const double* __restrict__ p,
                                                         const double* __restrict__ gp, const DevScalars* S,
                                                         double dt_given, int only_converged,
                                                         const double* __restrict__ dsp, double* vmax_part) {
    const int spc = blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    double w = 0.0;
    if (spc < M.nspecial) impulse_special_cell<R>(M, F, p, gp, S, dt_given, only_converged, dsp, spc, key, w);
    if (vmax_part) wave_max_to_slots(vmax_part, key, w);  // every lane of the wave reaches it
}

// ---------------------------------------------------------------------------------------
// Surface-load integral (aero_loads.cpp:35-47) on the device in FP64: F = sum (p_w - pref) S,
// M = sum (x - origin) x F; deterministic block partials.
// ---------------------------------------------------------------------------------------
// pressure force and moment of each patch, then the wall-shear force of the smooth-wall model
// (wall_force, zero while the model is off) and its moment: 12 partial arrays
__global__ __launch_bounds__(256) void k_loads(MeshView M, const double* __restrict__ p, const double* __restrict__ gp,
                                               const double* __restrict__ wall_force, double pref, double ox,
                                               double oy, double oz, int pbegin, int pend, double* partials) {
    const int pch = pbegin + blockIdx.x * blockDim.x + threadIdx.x;
    double v[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (pch < pend) {
        double Fp[3];
        wall_force_p(M, p, gp, pch, M.patch_sp[pch], pref, Fp);
        const double rx = M.patch_x[3 * pch] - ox, ry = M.patch_x[3 * pch + 1] - oy, rz = M.patch_x[3 * pch + 2] - oz;
        for (int part = 0; part < 2; ++part) {
            double F[3];
            for (int c = 0; c < 3; ++c) F[c] = part == 0 ? Fp[c] : wall_force[3 * pch + c];
            double* o = v + 6 * part;
            o[0] = F[0]; o[1] = F[1]; o[2] = F[2];
            o[3] = ry * F[2] - rz * F[1]; o[4] = rz * F[0] - rx * F[2]; o[5] = rx * F[1] - ry * F[0];
        }
    }
    for (int c = 0; c < 12; ++c) {
        const double part = block_sum<256>(v[c]);
        if (threadIdx.x == 0) partials[c * gridDim.x + blockIdx.x] = part;
    }
}

// Field sampling on an axis-aligned plane for a viewer (gui/slice_field.h:20-31 fields; the
// reference samples its uniform MAC state there and the AMR state on the host,
// core/fluid/amr_sampling.h).  Pixel (i, j) of an nu x nv image is the point at the centre of
// the pixel's rectangle in the plane; its value is the cell-centred value of the finest leaf
// cell containing it (no interpolation), NaN where the point is solid or outside the grid.
// The leaf is found from per-level maps of brick coordinates to stored bricks, finest first.
struct SliceMap {
    const int* map;      // per level: dims x*y*z brick ids (-1: none), level l at off[l]
    int64_t off[8];
    int dims[8][3];
    int levels;
    double lo[3];
    double h[8];
};

template <class R>
__global__ __launch_bounds__(256) void k_slice(MeshView M, FieldsView<R> F, const double* __restrict__ p, SliceMap G,
                                               int field, int axis, double coord, double u0, double u1, double v0,
                                               double v1, int nu, int nv, double qinf, float* __restrict__ out) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= nu * nv) return;
    const int i = pix % nu, j = pix / nu;
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    double x[3];
    x[axis] = coord;
    x[a1] = u0 + (i + 0.5) * (u1 - u0) / nu;
    x[a2] = v0 + (j + 0.5) * (v1 - v0) / nv;
    float val = __builtin_nanf("");
    for (int l = G.levels - 1; l >= 0; --l) {
        int c[3], bc[3];
        bool in = true;
        for (int d = 0; d < 3; ++d) {
            const double t = floor((x[d] - G.lo[d]) / G.h[l]);
            in = in && t >= 0.0 && t < 8.0 * G.dims[l][d];
            c[d] = in ? static_cast<int>(t) : 0;
            bc[d] = c[d] >> 3;
        }
        if (!in) continue;
        const int b = G.map[G.off[l] + (static_cast<int64_t>(bc[2]) * G.dims[l][1] + bc[1]) * G.dims[l][0] + bc[0]];
        if (b < 0) continue;
        const int64_t a = static_cast<int64_t>(b) * 512 + (c[0] & 7) + 8 * (c[1] & 7) + 64 * (c[2] & 7);
        if (M.mask[a] & kActive) {
            const double vx = double(F.v[0][a]), vy = double(F.v[1][a]), vz = double(F.v[2][a]);
            double r = 0.0;
            switch (field) {
                case 0: r = sqrt(vx * vx + vy * vy + vz * vz); break;
                case 1: r = vx; break;
                case 2: r = vy; break;
                case 3: r = vz; break;
                case 4: r = p[a]; break;
                default: r = p[a] / qinf; break;
            }
            val = static_cast<float>(r);
        }
        break;
    }
    out[pix] = val;
}

// Smooth-wall law (Spalding): y+ = u+ + e^(-kappa B) (e^(kappa u+) - 1 - kappa u+ - (kappa u+)^2/2
// - (kappa u+)^3/6), kappa = 0.41, B = 5.2.  With Re_y = u_t y / nu = u+ y+ the friction
// velocity is u_t / u+; u+ by 36 bisections of [0, 80] (amr_pressure.h:41-53).  Below
// kappa u+ = 0.5 the bracket is its Taylor series x^4/24 + x^5/120 + x^6/720 (no cancellation).
__host__ __device__ inline double smooth_wall_friction_velocity(double ut, double y, double nu) {
    if (!(ut > 0.0) || !(y > 0.0) || !(nu > 0.0)) return 0.0;
    const double re = ut * y / nu, kappa = 0.41, c = exp(-kappa * 5.2);
    double lo = 0.0, hi = 80.0;
    for (int it = 0; it < 36; ++it) {
        const double up = 0.5 * (lo + hi), x = kappa * up, x2 = x * x;
        const double rem = x < 0.5 ? x2 * x2 * (1.0 / 24.0 + x * (1.0 / 120.0 + x * (1.0 / 720.0)))
                                   : exp(x) - 1.0 - x - 0.5 * x2 - x2 * x * (1.0 / 6.0);
        const double yp = up + c * rem;
        if (up * yp < re) lo = up;
        else hi = up;
    }
    return ut / fmax(0.5 * (lo + hi), 1e-12);
}

// Wall shear of one wall cell (all its patches), implicit in the new velocity
// (amr_advection.cu:421-437, CPU twin :920-942): per patch with unit normal n, area A and wall
// distance y, the tangential speed |v_t| of the current velocity gives k = A u_tau^2 / |v_t|;
// (I + dt/V sum k (I - n n^T)) v' = v is solved by a 3x3 Cholesky factorisation, and the
// patch force is rho k (v' - (v'.n) n).  One lane per wall cell, patches in a fixed order, so
// the result does not depend on scheduling (the reference accumulates with atomics).
template <class R>
__global__ __launch_bounds__(256) void k_smooth_wall(MeshView M, FieldsView<R> F, int nwall,
                                                     const int* __restrict__ wall_cell, const int* __restrict__ wall_off,
                                                     const int* __restrict__ wall_patch, const double* __restrict__ wall_vol,
                                                     const double* __restrict__ wall_dist, double nu, double rho,
                                                     double* __restrict__ wall_k, double* __restrict__ wall_force,
                                                     const DevScalars* S) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwall) return;
    const int a = wall_cell[w];
    const double dt = S[M.meta[a >> 9].inst].dt;
    const double v[3] = {double(F.v[0][a]), double(F.v[1][a]), double(F.v[2][a])};
    double K[6] = {0, 0, 0, 0, 0, 0};  // symmetric: 00 01 02 11 12 22
    for (int j = wall_off[w]; j < wall_off[w + 1]; ++j) {
        const int pch = wall_patch[j];
        const double Sx = M.patch_S[3 * pch], Sy = M.patch_S[3 * pch + 1], Sz = M.patch_S[3 * pch + 2];
        const double area = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
        const double n[3] = {Sx / area, Sy / area, Sz / area};
        const double un = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
        const double t[3] = {v[0] - un * n[0], v[1] - un * n[1], v[2] - un * n[2]};
        const double speed = sqrt(fmax(0.0, t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
        double k = 0.0;
        if (speed > 0.0) {
            const double ut = smooth_wall_friction_velocity(speed, wall_dist[pch], nu);
            k = area * ut * ut / speed;
        }
        wall_k[pch] = k;
        K[0] += k * (1.0 - n[0] * n[0]); K[1] -= k * n[0] * n[1]; K[2] -= k * n[0] * n[2];
        K[3] += k * (1.0 - n[1] * n[1]); K[4] -= k * n[1] * n[2]; K[5] += k * (1.0 - n[2] * n[2]);
    }
    const double s = dt / wall_vol[w];
    const double a00 = 1.0 + s * K[0], a01 = s * K[1], a02 = s * K[2], a11 = 1.0 + s * K[3], a12 = s * K[4],
                 a22 = 1.0 + s * K[5];
    const double l00 = sqrt(a00), l10 = a01 / l00, l20 = a02 / l00;
    const double l11 = sqrt(a11 - l10 * l10), l21 = (a12 - l20 * l10) / l11;
    const double l22 = sqrt(a22 - l20 * l20 - l21 * l21);
    const double y0 = v[0] / l00, y1 = (v[1] - l10 * y0) / l11, y2 = (v[2] - l20 * y0 - l21 * y1) / l22;
    const double x2 = y2 / l22, x1 = (y1 - l21 * x2) / l11, x0 = (y0 - l10 * x1 - l20 * x2) / l00;
    const int sp = M.patch_sp[wall_patch[wall_off[w]]];
    if (redistributed(M, sp)) {  // the group shares the increment (amr_advection.cu:1279-1293)
        const double dm[3] = {wall_vol[w] * (x0 - v[0]), wall_vol[w] * (x1 - v[1]), wall_vol[w] * (x2 - v[2])};
        hand_over(M, sp, dm);
    } else {
        F.v[0][a] = R(x0); F.v[1][a] = R(x1); F.v[2][a] = R(x2);
    }
    // the force from the stored (storage-precision) velocity, as the reference reads it back
    const double vn[3] = {double(R(x0)), double(R(x1)), double(R(x2))};
    for (int j = wall_off[w]; j < wall_off[w + 1]; ++j) {
        const int pch = wall_patch[j];
        const double Sx = M.patch_S[3 * pch], Sy = M.patch_S[3 * pch + 1], Sz = M.patch_S[3 * pch + 2];
        const double area = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
        const double n[3] = {Sx / area, Sy / area, Sz / area};
        const double un = vn[0] * n[0] + vn[1] * n[1] + vn[2] * n[2];
        const double k = wall_k[pch];
        for (int c = 0; c < 3; ++c) wall_force[3 * pch + c] = rho * k * (vn[c] - un * n[c]);
    }
}

// max |v| over active cells; a non-finite component makes the maximum infinite
template <class R>
__global__ __launch_bounds__(512) void k_vmax(MeshView M, FieldsView<R> F, double* partials) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * 512 + threadIdx.x;
    double v = 0.0;
    if (M.mask[a] & kActive) {
        const double x = double(F.v[0][a]), y = double(F.v[1][a]), z = double(F.v[2][a]);
        v = fmax(fabs(x), fmax(fabs(y), fabs(z)));
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) v = INFINITY;
    }
    store_partial(partials, static_cast<int>(blockIdx.x), block_max<512>(v));
}

// the same timestep for every instance (an explicit dt given by the caller)
__global__ void k_set_dt(DevScalars* S, int ninst, double dt) {
    const int i = threadIdx.x;
    if (i < ninst) S[i].dt = dt;
}

// diffusion substeps of each instance from its timestep: rate nu * max(sum conductance / V)
// (constant viscosity) or the reduced Smagorinsky rate; substeps ceil(dt rate / 0.7)
// (amr_advection.cu:1343), dts = dt / substeps; no diffusion when the rate is zero
__global__ void k_diffusion_substeps(DevScalars* S, int ninst, double nu, const double* __restrict__ maxdiff,
                                     int use_rate) {
    const int i = threadIdx.x;
    if (i >= ninst) return;
    const double dt = S[i].dt;
    int subs = 0;
    if (use_rate) subs = max(1, static_cast<int>(ceil(dt * S[i].diff_rate / 0.7)));
    else if (nu > 0.0) subs = max(1, static_cast<int>(ceil(dt * nu * maxdiff[i] / 0.7)));
    S[i].subs = subs;
    S[i].dts = subs > 0 ? dt / subs : 0.0;
}

template <class R>
__global__ void k_init_velocity(MeshView M, FieldsView<R> F) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= M.ndof) return;
    const bool act = (M.mask[a] & kActive) != 0;
    F.v[0][a] = act ? R(M.speed) : R(0);
    F.v[1][a] = R(0);
    F.v[2][a] = R(0);
    F.f[0][a] = R(0); F.f[1][a] = R(0); F.f[2][a] = R(0);
}

}  // namespace pcfd
