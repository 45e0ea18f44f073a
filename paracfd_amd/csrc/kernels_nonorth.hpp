// Non-orthogonal correction of the pressure operator on the device, and the vector kernels
// of the flexible GMRES outer solve that it needs (the full operator is not symmetric).
//
// Operator (amr_pressure.cu:95-178, 831-834): gradient at the correction nodes
//   g[n] = sum_q w_q (p[nbr_q] - p[dof_n]),
// then for every correction face f (lower L, upper U) the flux
//   F_f = -area_f (delta_f (p_U - p_L) + k_f . ((1-w_f) g[lnode] + w_f g[unode]))
// gathered per touched cell (out[L] += F, out[U] -= F) from a host-built incidence list, so
// every cell has one owner thread and the sum order is fixed.  Face correction
// (amr_pressure.cu:2047-2069): u_f -= (dt/rho) (delta (p_U - p_L) + k . g_f) on the face's
// structured face velocity f[axis][L] or compact edge.  All FP64 except the stored face
// velocities.  Memory-bound on small index lists; one thread per node / cell / face.
#pragma once

#include "device.hpp"

namespace pcfd {

struct NonOrthView {
    int nf = 0, nn = 0, ninc = 0;
    const int *lower = nullptr, *upper = nullptr, *edge = nullptr, *lnode = nullptr, *unode = nullptr;
    const double *area = nullptr, *delta = nullptr, *k = nullptr, *w = nullptr;
    const int *gdof = nullptr, *goff = nullptr, *gnbr = nullptr;
    const double* gw = nullptr;
    const int *inc_dof = nullptr, *inc_off = nullptr, *inc_face = nullptr;
    double* grad = nullptr;  // 3 per node
};

__global__ __launch_bounds__(256) void k_no_grad(NonOrthView V, const double* __restrict__ p) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V.nn) return;
    const double pc = p[V.gdof[n]];
    double g0 = 0, g1 = 0, g2 = 0;
    for (int q = V.goff[n]; q < V.goff[n + 1]; ++q) {
        const double d = p[V.gnbr[q]] - pc;
        g0 += V.gw[3 * q] * d;
        g1 += V.gw[3 * q + 1] * d;
        g2 += V.gw[3 * q + 2] * d;
    }
    V.grad[3 * n] = g0;
    V.grad[3 * n + 1] = g1;
    V.grad[3 * n + 2] = g2;
}

// delta (p_U - p_L) + k . g_f of face f
__device__ __forceinline__ double no_extra(const NonOrthView& V, int f, const double* __restrict__ p) {
    double e = V.delta[f] * (p[V.upper[f]] - p[V.lower[f]]);
    const int ln = V.lnode[f], un = V.unode[f];
    if (ln >= 0 && un >= 0) {
        const double w = V.w[f];
        for (int c = 0; c < 3; ++c) e += V.k[3 * f + c] * ((1.0 - w) * V.grad[3 * ln + c] + w * V.grad[3 * un + c]);
    }
    return e;
}

// out[cell] += sign * (sum of the cell's correction fluxes)
__global__ __launch_bounds__(256) void k_no_apply(NonOrthView V, const double* __restrict__ p, double* __restrict__ out,
                                                  double sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V.ninc) return;
    double s = 0;
    for (int j = V.inc_off[i]; j < V.inc_off[i + 1]; ++j) {
        const int fe = V.inc_face[j];
        const int f = fe >= 0 ? fe : ~fe;
        const double F = -V.area[f] * no_extra(V, f, p);
        s += fe >= 0 ? F : -F;
    }
    out[V.inc_dof[i]] += sign * s;
}

template <class R>
__global__ __launch_bounds__(256) void k_no_correct(NonOrthView V, FieldsView<R> F, const double* __restrict__ p,
                                                    const DevScalars* S, double rho) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= V.nf) return;
    const double s = S[0].dt / rho;
    const double e = no_extra(V, f, p);
    const int ed = V.edge[f];
    if (ed >= 0) {
        F.fe[ed] = R(double(F.fe[ed]) - s * e);
    } else {
        R* fa = F.f[-1 - ed];
        const int l = V.lower[f];
        fa[l] = R(double(fa[l]) - s * e);
    }
}

// ---------------------------------------------------------------------------------------
// Flexible GMRES vector kernels.  Dot products are two-stage and deterministic: block b of
// grid column v writes partials[v * G + b]; k_mdot_final reduces them in a fixed order.
// Only active cells enter a dot.
// ---------------------------------------------------------------------------------------
constexpr int kGmresBlocks = 1024;

template <int BLOCK>
__device__ __forceinline__ double block_reduce_sum(double v) {
    __shared__ double ws[BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x < 64) {
        r = threadIdx.x < BLOCK / 64 ? ws[threadIdx.x] : 0.0;
        for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    }
    return r;  // valid in thread 0
}

constexpr int kGmresMaxDots = 32;

// partials[v * gridDim.x + block] = sum over active i of V[v][i] * w[i] for v < nv: w and the
// mask are read once, the nv sums stay in registers (nv <= kGmresMaxDots)
__global__ __launch_bounds__(256) void k_mdot(int64_t n, const double* __restrict__ Vb, int64_t ldv, int nv,
                                              const double* __restrict__ w, const uint8_t* __restrict__ mask,
                                              double* __restrict__ partials) {
    double acc[kGmresMaxDots];
#pragma unroll
    for (int v = 0; v < kGmresMaxDots; ++v) acc[v] = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!(mask[i] & kActive)) continue;
        const double wi = w[i];
#pragma unroll
        for (int v = 0; v < kGmresMaxDots; ++v)
            if (v < nv) acc[v] += Vb[v * ldv + i] * wi;
    }
#pragma unroll
    for (int v = 0; v < kGmresMaxDots; ++v) {
        if (v < nv) {  // nv is uniform over the block, so the barriers below are too
            const double s = block_reduce_sum<256>(acc[v]);
            if (threadIdx.x == 0) partials[static_cast<int64_t>(v) * gridDim.x + blockIdx.x] = s;
            __syncthreads();  // the reduction's shared slots are reused by the next v
        }
    }
}

__global__ __launch_bounds__(256) void k_mdot_final(int nparts, const double* __restrict__ partials, double* __restrict__ out) {
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += partials[static_cast<int64_t>(blockIdx.x) * nparts + i];
    s = block_reduce_sum<256>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// w -= sum_v h[v] V[v] on active cells; partials[block] = that block's sum of the new w^2
__global__ __launch_bounds__(256) void k_gs_update(int64_t n, const double* __restrict__ Vb, int64_t ldv, int nv,
                                                   const double* __restrict__ h, double* __restrict__ w,
                                                   const uint8_t* __restrict__ mask, double* __restrict__ partials) {
    double ss = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!(mask[i] & kActive)) { w[i] = 0.0; continue; }
        double s = w[i];
        for (int v = 0; v < nv; ++v) s -= h[v] * Vb[v * ldv + i];
        w[i] = s;
        ss += s * s;
    }
    ss = block_reduce_sum<256>(ss);
    if (threadIdx.x == 0) partials[blockIdx.x] = ss;
}

// GCRO recycle update: w -= sum_v a[v] C[v] and, when z is given, z += zsign sum_v a[v] U[v]
// (the same coefficients keep A z = w when A U = C); partials[block] = that block's sum of the
// new w^2
__global__ __launch_bounds__(256) void k_recycle_update(int64_t n, const double* __restrict__ Cb,
                                                        const double* __restrict__ Ub, int64_t ld, int nv,
                                                        const double* __restrict__ a, double* __restrict__ w,
                                                        double* __restrict__ z, double zsign,
                                                        const uint8_t* __restrict__ mask, double* __restrict__ partials) {
    double ss = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!(mask[i] & kActive)) {
            w[i] = 0.0;
            if (z) z[i] = 0.0;
            continue;
        }
        double s = w[i];
        for (int v = 0; v < nv; ++v) s -= a[v] * Cb[v * ld + i];
        w[i] = s;
        ss += s * s;
        if (z) {
            double t = 0;
            for (int v = 0; v < nv; ++v) t += a[v] * Ub[v * ld + i];
            z[i] += zsign * t;
        }
    }
    ss = block_reduce_sum<256>(ss);
    if (threadIdx.x == 0) partials[blockIdx.x] = ss;
}

// y += a x on active cells; partials[block] = that block's sum of the new y^2
__global__ __launch_bounds__(256) void k_axpy_norm(int64_t n, double* __restrict__ y, double a,
                                                   const double* __restrict__ x, const uint8_t* __restrict__ mask,
                                                   double* __restrict__ partials) {
    double ss = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!(mask[i] & kActive)) continue;
        const double s = y[i] + a * x[i];
        y[i] = s;
        ss += s * s;
    }
    ss = block_reduce_sum<256>(ss);
    if (threadIdx.x == 0) partials[blockIdx.x] = ss;
}

// dst = a * src on active cells, 0 elsewhere
// (src may be dst)
__global__ __launch_bounds__(256) void k_scale(int64_t n, const double* src, double a, double* dst,
                                               const uint8_t* __restrict__ mask) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
        dst[i] = (mask[i] & kActive) ? a * src[i] : 0.0;
}

// x += sum_j y[j] Z[j] on active cells
__global__ __launch_bounds__(256) void k_combine(int64_t n, const double* __restrict__ Zb, int64_t ldz, int nz,
                                                 const double* __restrict__ y, double* __restrict__ x,
                                                 const uint8_t* __restrict__ mask) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!(mask[i] & kActive)) continue;
        double s = x[i];
        for (int j = 0; j < nz; ++j) s += y[j] * Zb[j * ldz + i];
        x[i] = s;
    }
}

// final statistics of the outer solve into the instance scalars
__global__ void k_no_status(DevScalars* S, int iterations, int converged, double relres) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    S[0].iter = iterations;
    S[0].converged = converged;
    S[0].relres = relres;
    S[0].done = 1;
}

}  // namespace pcfd
