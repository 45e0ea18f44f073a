// Projection solve kernels: FP64 preconditioned conjugate gradients with device-resident
// scalars and a device convergence flag, and the geometric V-cycle in FP64 / FP32 / BF16
// storage (VT).  The composite level uses the brick layout; coarser levels are CSR.
//
// The Krylov scalars never leave the device: each reduction kernel's last workgroup
// finalises alpha / beta / the relative residual and raises S->done, and every kernel of
// the loop returns immediately once it is set, so the host may enqueue several iterations
// between convergence reads without changing the result.
#pragma once

#include "device.hpp"

namespace pcfd {

// BF16 storage with FP32 arithmetic (round to nearest even on store)
struct bf16 {
    uint16_t bits;
};
__device__ __forceinline__ float to_f(float x) { return x; }
__device__ __forceinline__ double to_f(double x) { return x; }
__device__ __forceinline__ float to_f(bf16 x) { return __uint_as_float(static_cast<unsigned>(x.bits) << 16); }
template <class T> struct Arith { using type = T; };
template <> struct Arith<bf16> { using type = float; };
template <class T>
__device__ __forceinline__ T from_f(typename Arith<T>::type x) { return T(x); }
template <>
__device__ __forceinline__ bf16 from_f<bf16>(float x) {
    const unsigned u = __float_as_uint(x);
    const unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    return bf16{static_cast<uint16_t>(r >> 16)};
}

// (A x)_a and the diagonal on the composite level (amr_pressure.cu:72-93 structured part,
// :105-134 compact gather, outlet 2h p at x-max)
template <class A, class T>
__device__ __forceinline__ A apply_row0(const MeshView& M, const T* __restrict__ x, int b, int q, uint8_t m, int sp,
                                        A h, A& diag) {
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const A xa = to_f(x[a]);
    A acc = A(0), dg = A(0);
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) { acc += h * (xa - A(to_f(x[neighbor_dof(M, b, q, axis, 1)]))); dg += h; }
        if (m & (kMX << axis)) { acc += h * (xa - A(to_f(x[neighbor_dof(M, b, q, axis, -1)]))); dg += h; }
    }
    if (M.outlet && (q & 7) == 7 && (M.brick_flags[b] & kXMAX)) { acc += A(2) * h * xa; dg += A(2) * h; }
    if (sp >= 0)
        for (int k = M.special_inc_off[sp]; k < M.special_inc_off[sp + 1]; ++k) {
            const int raw = M.special_inc[k];
            const int e = raw >= 0 ? raw : ~raw;
            const int64_t o = raw >= 0 ? M.eb[e] : M.ea[e];
            const A c = A(M.ecoef[e]);
            acc += c * (xa - A(to_f(x[o])));
            dg += c;
        }
    diag = dg;
    return acc;
}

// ---------------------------------------------------------------------------------------
// PCG (amr_pressure.cu:1267-1353 solve_orthogonal): r = b - A x, alpha, x/r update, beta
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_cg_residual(MeshView M, const double* __restrict__ x,
                                                     const double* __restrict__ rhs, double* __restrict__ r,
                                                     double* partials, DevScalars* S) {
    if (S->done) return;
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    double ra = 0.0;
    if (m & kActive) {
        double dg;
        ra = rhs[a] - apply_row0<double>(M, x, b, q, m, sp, M.h[M.brick_level[b]], dg);
    }
    r[a] = ra;
    store_partial(partials, block_sum<512>(ra * ra));
}

__global__ __launch_bounds__(512) void k_cg_apply_dot(MeshView M, const double* __restrict__ d, double* __restrict__ Ad,
                                                      double* partials, DevScalars* S) {
    if (S->done) return;
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    double y = 0.0;
    if (m & kActive) {
        double dg;
        y = apply_row0<double>(M, d, b, q, m, sp, M.h[M.brick_level[b]], dg);
    }
    Ad[a] = y;
    store_partial(partials, block_sum<512>(y * d[a]));
}

__global__ __launch_bounds__(256) void k_cg_update(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ rprev, const double* __restrict__ d,
                                                   const double* __restrict__ Ad, double* partials, const DevScalars* S) {
    if (S->done) return;
    const double alpha = S->alpha;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        x[i] += alpha * d[i];
        const double ri = r[i];
        if (rprev) rprev[i] = ri;
        const double rn = ri - alpha * Ad[i];
        r[i] = rn;
        acc += rn * rn;
    }
    store_partial(partials, block_sum<256>(acc));
}

// Scalar finalisation of each reduction (one workgroup, partials summed in block order).
enum FinalizeOp : int {
    FIN_RATE = 0,     // max outflow rate -> dt = cfl / max(U/h, rate)
    FIN_BNORM = 1,    // ||b||, resets the Krylov flags; zero RHS ends the solve
    FIN_MAXDIV = 2,   // max |div| / V
    FIN_VMAX = 3,     // max |v|
    FIN_RESID = 4,    // initial ||r||, convergence test
    FIN_APPLY = 5,    // pAp -> alpha, breakdown guard
    FIN_UPDATE = 6,   // ||r||, iteration count, convergence / iteration limit
    FIN_RZ = 7,       // r.z (and r_prev.z) -> beta, rz
    FIN_RZ_INITIAL = 8,
    FIN_LOADS = 9     // six load sums
};

__global__ __launch_bounds__(1024) void k_finalize(int op, const double* __restrict__ partials, int nparts,
                                                   DevScalars* S, double cfl, double rate_floor, double rtol,
                                                   int max_iterations, int flexible) {
    if ((op == FIN_RESID || op == FIN_APPLY || op == FIN_UPDATE || op == FIN_RZ || op == FIN_RZ_INITIAL) && S->done) return;
    const bool take_max = op == FIN_RATE || op == FIN_MAXDIV || op == FIN_VMAX;
    const int arrays = op == FIN_LOADS ? 6 : ((op == FIN_RZ || op == FIN_RZ_INITIAL) && flexible ? 2 : 1);
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < arrays; ++c) t[c] = gather_partials<1024>(partials + c * nparts, nparts, take_max);
    if (threadIdx.x != 0) return;
    switch (op) {
        case FIN_RATE:
            S->rate = t[0];
            S->dt = cfl / fmax(rate_floor, t[0]);
            break;
        case FIN_BNORM:
            S->bnorm = sqrt(t[0]);
            S->iter = 0; S->done = 0; S->converged = 0; S->breakdown = 0; S->zero_rhs = 0;
            if (!(t[0] > 0.0)) { S->done = 1; S->converged = 1; S->zero_rhs = 1; S->relres = 0.0; }
            break;
        case FIN_MAXDIV: S->maxdiv = t[0]; break;
        case FIN_VMAX: S->vmax = t[0]; break;
        case FIN_RESID:
            S->rr = t[0];
            S->relres = sqrt(t[0]) / S->bnorm;
            if (S->relres <= rtol) { S->done = 1; S->converged = 1; }
            break;
        case FIN_APPLY:
            S->pAp = t[0];
            if (!(t[0] > 0.0) || isinf(t[0])) { S->done = 1; S->breakdown = 1; }
            else S->alpha = S->rz / t[0];
            break;
        case FIN_UPDATE:
            S->rr = t[0];
            S->iter += 1;
            S->relres = sqrt(t[0]) / S->bnorm;
            if (S->relres <= rtol) { S->done = 1; S->converged = 1; }
            else if (S->iter >= max_iterations || !(t[0] == t[0])) S->done = 1;
            break;
        case FIN_RZ:
            // Fletcher-Reeves for a fixed preconditioner; Polak-Ribiere (flexible CG) with r_prev
            S->beta = (t[0] - t[1]) / S->rz;
            S->rz = t[0];
            break;
        case FIN_RZ_INITIAL: S->rz = t[0]; break;
        case FIN_LOADS:
            for (int c = 0; c < 6; ++c) S->loads[c] = t[c];
            break;
        default: break;
    }
}

__global__ __launch_bounds__(256) void k_cg_direction(int64_t n, const double* __restrict__ z, double* __restrict__ d,
                                                      const DevScalars* S, int initial) {
    if (S->done) return;
    const double beta = initial ? 0.0 : S->beta;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
        d[i] = initial ? z[i] : z[i] + beta * d[i];
}

// ---------------------------------------------------------------------------------------
// V-cycle, composite level: damped Jacobi (project.cu:76-83 restated with the composite
// operator).  mode 0: x = w b / D; 1: x += w (b - A x) / D; 2: last sweep, writes z (FP64)
// and the r.z (and r_prev.z for flexible CG) reductions that finalise beta.
// ---------------------------------------------------------------------------------------
template <class A>
__device__ __forceinline__ A diag_row0(const MeshView& M, int b, int q, uint8_t m, int sp, A h) {
    A dg = A(0);
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) dg += h;
        if (m & (kMX << axis)) dg += h;
    }
    if (M.outlet && (q & 7) == 7 && (M.brick_flags[b] & kXMAX)) dg += A(2) * h;
    if (sp >= 0)
        for (int k = M.special_inc_off[sp]; k < M.special_inc_off[sp + 1]; ++k) {
            const int raw = M.special_inc[k];
            dg += A(M.ecoef[raw >= 0 ? raw : ~raw]);
        }
    return dg;
}

// mode 2 stores two partial arrays of gridDim.x entries: r.z then r_prev.z
template <class VT>
__global__ __launch_bounds__(512) void k_smooth0(MeshView M, int mode, const VT* __restrict__ xin, VT* __restrict__ xout,
                                                 const double* __restrict__ rhs, double* __restrict__ zout,
                                                 const double* __restrict__ rprev, double omega, double* partials,
                                                 const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    A val = A(0);
    if (m & kActive) {
        const A h = A(M.h[M.brick_level[b]]);
        if (mode == 0 || xin == nullptr) {
            val = A(omega) * A(rhs[a]) / diag_row0<A>(M, b, q, m, sp, h);
        } else {
            A dg;
            const A Ax = apply_row0<A, VT>(M, xin, b, q, m, sp, h, dg);
            val = A(to_f(xin[a])) + A(omega) * (A(rhs[a]) - Ax) / dg;
        }
    }
    if (mode != 2) {
        xout[a] = from_f<VT>(val);
        return;
    }
    const double z = double(val);
    zout[a] = z;
    const double s1 = block_sum<512>(rhs[a] * z);
    const double s2 = rprev ? block_sum<512>(rprev[a] * z) : 0.0;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}

template <class VT>
__global__ __launch_bounds__(512) void k_residual0(MeshView M, const VT* __restrict__ x, const double* __restrict__ rhs,
                                                   VT* __restrict__ res, const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    A val = A(0);
    if (m & kActive) {
        A dg;
        val = A(rhs[a]) - apply_row0<A, VT>(M, x, b, q, m, sp, A(M.h[M.brick_level[b]]), dg);
    }
    res[a] = from_f<VT>(val);
}

// coarse b = sum of the children's residuals (R = P^T, piecewise-constant P)
template <class VT>
__global__ void k_restrict(CsrView C, const VT* __restrict__ fine, VT* __restrict__ coarse, const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n) return;
    A s = A(0);
    for (int k = C.child_off[c]; k < C.child_off[c + 1]; ++k) s += A(to_f(fine[C.child[k]]));
    coarse[c] = from_f<VT>(s);
}

template <class VT>
__global__ void k_prolong(int64_t n, const int* __restrict__ parent, VT* __restrict__ fine, const VT* __restrict__ coarse,
                          const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int pc = parent[i];
    if (pc >= 0) fine[i] = from_f<VT>(A(to_f(fine[i])) + A(to_f(coarse[pc])));
}

// CSR level: mode 0 first sweep from zero, 1 sweep, residual when res != nullptr
template <class VT, class CT>
__global__ void k_smooth_csr(CsrView C, int mode, const VT* __restrict__ xin, VT* __restrict__ xout,
                             const VT* __restrict__ rhs, double omega, const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= C.n) return;
    const CT* coef = static_cast<const CT*>(C.coef);
    const CT* diag = static_cast<const CT*>(C.diag);
    const A D = A(diag[r]);
    const A br = A(to_f(rhs[r]));
    if (mode == 0) { xout[r] = from_f<VT>(A(omega) * br / D); return; }
    A Ax = D * A(to_f(xin[r]));
    for (int k = C.row_off[r]; k < C.row_off[r + 1]; ++k) Ax -= A(coef[k]) * A(to_f(xin[C.col[k]]));
    xout[r] = from_f<VT>(A(to_f(xin[r])) + A(omega) * (br - Ax) / D);
}

template <class VT, class CT>
__global__ void k_residual_csr(CsrView C, const VT* __restrict__ x, const VT* __restrict__ rhs, VT* __restrict__ res,
                               const DevScalars* S) {
    if (S->done) return;
    using A = typename Arith<VT>::type;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= C.n) return;
    const CT* coef = static_cast<const CT*>(C.coef);
    const CT* diag = static_cast<const CT*>(C.diag);
    A Ax = A(diag[r]) * A(to_f(x[r]));
    for (int k = C.row_off[r]; k < C.row_off[r + 1]; ++k) Ax -= A(coef[k]) * A(to_f(x[C.col[k]]));
    res[r] = from_f<VT>(A(to_f(rhs[r])) - Ax);
}

// coarsest level: x = A^-1 b with the dense inverse, one workgroup per row
template <class VT, class CT>
__global__ __launch_bounds__(256) void k_dense_solve(int n, const CT* __restrict__ inv, const VT* __restrict__ rhs,
                                                     VT* __restrict__ x, const DevScalars* S) {
    if (S->done) return;
    const int r = blockIdx.x;
    double acc = 0.0;
    for (int c = threadIdx.x; c < n; c += 256) acc += double(inv[static_cast<int64_t>(r) * n + c]) * double(to_f(rhs[c]));
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0) x[r] = from_f<VT>(static_cast<typename Arith<VT>::type>(acc));
}

// convert an FP64 vector to the V-cycle type (used by the multigrid primitive entry point)
template <class VT>
__global__ void k_convert(int64_t n, const double* __restrict__ in, VT* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = from_f<VT>(static_cast<typename Arith<VT>::type>(in[i]));
}
template <class VT>
__global__ void k_to_double(int64_t n, const VT* __restrict__ in, double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = double(to_f(in[i]));
}

}  // namespace pcfd
