// Device-side views and helpers shared by the gfx950 kernels.
//
// Layout in HBM (brick-major, ghost-free): dof = brick * 512 + i + 8 j + 64 k.  One
// workgroup of 512 lanes (8 wave64) owns one 8^3 brick in the cell kernels, so a brick's
// cells are read as 512 consecutive elements and cross-brick neighbours come from the
// six same-level neighbour bricks (brick_nbr).  Special cells (cut cells, coarse/fine and
// partially open faces) are ranked inside the workgroup with a wave ballot and own a small
// incidence list of compact edges; regular cells carry only one mask byte.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pcfd {

constexpr int kBrickThreads = 512;
constexpr uint8_t kPX = 1, kMX = 8, kActive = 64, kSpecial = 128;
constexpr uint8_t kXMIN = 1, kXMAX = 2, kYMIN = 4, kYMAX = 8, kZMIN = 16, kZMAX = 32;

// Scalars the Krylov loop keeps on the device (no host round trip per iteration).
struct DevScalars {
    double rz, rz_prev, pAp, alpha, beta, rr, bnorm, relres;
    double dt, rate, maxdiv, vmax, rpz;
    double loads[6];
    int iter, done, converged, breakdown, zero_rhs, pad0;
    unsigned int ticket[16];
};

struct MeshView {
    int nbricks;
    int64_t ndof;
    const uint8_t* mask;
    const int* brick_level;
    const uint8_t* brick_flags;
    const int* brick_nbr;
    double h[4];
    int outlet;
    const int* brick_special_off;
    const double* special_volume;
    const int* special_inc_off;
    const int* special_inc;
    const int* special_patch;
    int nedges;
    const int* ea;
    const int* eb;
    const int8_t* eaxis;
    const double* earea;
    const double* ecoef;
    const double* ew;
    int npatches;
    const int* patch_dof;
    const double* patch_S;
    const double* patch_x;
    const int* lsq_off;
    const int* lsq_dof;
    const double* lsq_beta;
    double speed, rho, cfl, nu, rtol;
    double finest_h;
    int max_iterations;
};

struct CsrView {
    int n;
    const int* row_off;
    const int* col;
    const void* coef;   // double or float
    const void* diag;   // double or float
    const int* parent;  // into the next coarser level
    const int* child_off;
    const int* child;
};

__device__ __forceinline__ int axis_coord(int q, int axis) { return (q >> (3 * axis)) & 7; }

// neighbour dof across face (axis, sign) of local cell q in brick b; the caller has checked
// the structured-face bit, so a cross-brick neighbour exists
__device__ __forceinline__ int64_t neighbor_dof(const MeshView& M, int b, int q, int axis, int sign) {
    const int stride = 1 << (3 * axis);
    const int c = axis_coord(q, axis) + sign;
    if (c >= 0 && c < 8) return static_cast<int64_t>(b) * 512 + q + sign * stride;
    const int nb = M.brick_nbr[6 * b + 2 * axis + (sign > 0 ? 1 : 0)];
    return static_cast<int64_t>(nb) * 512 + q - sign * 7 * stride;
}

__device__ __forceinline__ int64_t neighbor_dof(const MeshView& M, int64_t a, int axis, int sign) {
    return neighbor_dof(M, static_cast<int>(a >> 9), static_cast<int>(a & 511), axis, sign);
}

// Rank of a special cell inside its brick: ballot per wave, wave counts through LDS.
// Block-collective: every lane of the 512-lane workgroup must call it.
__device__ __forceinline__ int special_rank(const MeshView& M, int b, int q, bool sp) {
    __shared__ int wave_count[8];
    const unsigned long long bal = __ballot(sp);
    const int wave = q >> 6, lane = q & 63;
    if (lane == 0) wave_count[wave] = __popcll(bal);
    __syncthreads();
    if (!sp) return -1;
    int base = M.brick_special_off[b];
    for (int w = 0; w < wave; ++w) base += wave_count[w];
    const unsigned long long below = lane == 0 ? 0ull : (bal & ((~0ull) >> (64 - lane)));
    return base + __popcll(below);
}

template <int BLOCK>
__device__ __forceinline__ double block_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ double ws[BLOCK / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x < 64) {
        r = threadIdx.x < BLOCK / 64 ? ws[threadIdx.x] : 0.0;
        for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    }
    return r;  // valid in thread 0
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double wm[BLOCK / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) wm[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x < 64) {
        r = threadIdx.x < BLOCK / 64 ? wm[threadIdx.x] : 0.0;
        for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    }
    return r;
}

// Reductions are two-stage and deterministic: every workgroup stores its partial with a
// plain store (partials[array * nparts + block]); a one-workgroup finalize kernel launched
// next on the same stream sums them in block order.  (A "last workgroup finalises" scheme
// needs an agent-scope release fence per workgroup, which writes back L2 and cost ~0.4 ms
// per launch at 8k workgroups on gfx950.)
__device__ __forceinline__ void store_partial(double* partials, double value) {
    if (threadIdx.x == 0) partials[blockIdx.x] = value;
}

// Sum (or max) of n partials in fixed order; result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double gather_partials(const double* partials, int n, bool take_max) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        const double v = partials[i];
        acc = take_max ? fmax(acc, v) : acc + v;
    }
    return take_max ? block_max<BLOCK>(acc) : block_sum<BLOCK>(acc);
}

}  // namespace pcfd
