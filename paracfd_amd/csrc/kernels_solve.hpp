// Projection solve kernels: FP64 preconditioned conjugate gradients with device-resident
// scalars and a device convergence flag, and the geometric V-cycle in FP64 / FP32 / BF16
// storage (VT).  The composite level uses the brick layout with an LDS tile per brick;
// coarser levels are SELL-64 (slices of 64 rows, slot-major).
//
// The Krylov scalars never leave the device: a finalize kernel after each reduction (one
// workgroup per instance of a batch) computes alpha / beta / the relative residual and
// raises the instance's done flag, and every composite-level kernel of the loop returns
// immediately for the bricks of a finished instance, so the host may enqueue (or replay as
// a graph) several iterations between convergence reads without changing the result.
// Coarse-level kernels test the batch aggregate flag only: until every instance is done
// they also rewrite the finished instances' scratch levels, which nothing reads.
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
template <class VT>
__device__ __forceinline__ typename Arith<VT>::type round_vt(typename Arith<VT>::type v) {
    return to_f(from_f<VT>(v));
}

__device__ __forceinline__ double level_h(const MeshView& M, int b) { return M.meta[b].h; }

// composite-level diagonal (sum of face conductances + outlet + compact edges)
template <class A>
__device__ __forceinline__ A diag_row0(const MeshView& M, int b, int q, uint8_t m, int sp, A h) {
    A dg = A(0);
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) dg += h;
        if (m & (kMX << axis)) dg += h;
    }
    if (M.outlet && (q & 7) == 7 && (M.meta[b].flags & kXMAX)) dg += A(2) * h;
    if (sp >= 0)
        for (int k = M.special_inc_off[sp]; k < M.special_inc_off[sp + 1]; ++k) dg += A(M.sp_coef[k]);
    return dg;
}

// (A y)_a on the composite level (amr_pressure.cu:72-93 structured part, :105-134 compact
// gather, outlet 2h p at x-max) with y already in the brick's LDS tile; val gives y at any
// dof for the compact edges.  Returns the row and its diagonal.
template <class A, class Fn>
__device__ __forceinline__ A row0(const MeshView& M, const A* tile, Fn val, const BrickScalars& bs, int q, uint8_t m, int sp,
                                  A ya, A& dg) {
    A acc = tile_apply<A>(tile, q, m, A(bs.h), bs.outlet, ya, dg);
    if (sp >= 0) compact_apply<A>(M, sp, val, ya, acc, dg);
    return acc;
}

// Brick kernels issue every independent load first (brick record: one scalar load; cell,
// halo and right-hand-side values: vector loads) and only then test the instance's done
// flag and stage the tile, so one workgroup waits for about two memory latencies instead of
// one per dependent step.
// the same with the compact-edge terms fetched before the tile barrier
template <class A, class Fn>
__device__ __forceinline__ A row0f(const MeshView& M, const A* tile, const CompactFetch<A>& cf, Fn val,
                                   const BrickScalars& bs, int q, uint8_t m, A ya, A& dg) {
    A acc = tile_apply<A>(tile, q, m, A(bs.h), bs.outlet, ya, dg);
    if (m & kSpecial) compact_apply_fetched<A>(M, cf, val, ya, acc, dg);
    return acc;
}

#define PCFD_BRICK_BEGIN                                           \
    const int b = brick_of_block(M.xcd_chunk), q = threadIdx.x;               \
    const int64_t a = static_cast<int64_t>(b) * 512 + q;           \
    const BrickScalars bs = brick_scalars(M, b);                   \
    const uint8_t m = M.mask[a]

// two cells per lane (kBL lanes per brick): cell q = t + kBL h, h = 0, 1
#define PCFD_BRICK2_BEGIN                                                     \
    const int b = brick_of_block(M.xcd_chunk), t = threadIdx.x;               \
    const int64_t base = static_cast<int64_t>(b) * 512;                       \
    const BrickScalars bs = brick_scalars(M, b);                              \
    uint8_t m[2];                                                             \
    for (int h = 0; h < 2; ++h) m[h] = M.mask[base + t + kBL * h]

#define PCFD_BRICK_PROLOGUE \
    PCFD_BRICK_BEGIN;       \
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0)

__device__ __forceinline__ bool brick_done(const MeshView& M, const DevScalars* S, int b) {
    return S[M.meta[b].inst].done != 0;
}

// ---------------------------------------------------------------------------------------
// PCG (amr_pressure.cu:1267-1353 solve_orthogonal)
// ---------------------------------------------------------------------------------------
// r = b - A x and the ||r||^2 partial
__global__ __launch_bounds__(kBL, 8) void k_cg_residual(MeshView M, const double* __restrict__ x,
                                                     const double* __restrict__ rhs, double* __restrict__ r,
                                                     double* partials, const DevScalars* S) {
    PCFD_BRICK2_BEGIN;
    auto val = [&](int64_t c) -> double { return x[c]; };
    const TileFetch2<double> tf = fetch_tile_x2<double>(M, b, val);
    double rb[2];
    for (int h = 0; h < 2; ++h) rb[h] = rhs[base + t + kBL * h];
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<double> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<double>(M, sp[h], val);
    }
    if (done) return;
    __shared__ double tile[kTileN];
    stage_tile_x2<double>(tile, tf);
    double sum = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        double ra = 0.0;
        if (m[h] & kActive) {
            double dg;
            ra = rb[h] - row0f<double>(M, tile, cf[h], val, bs, q, m[h], tf.own[h], dg);
        }
        r[base + q] = ra;
        sum += ra * ra;
    }
    store_partial(partials, b, block_sum<kBL>(sum));
}

// A d and the d.Ad partial (used by the measurement hook)
__global__ __launch_bounds__(kBL) void k_cg_apply_dot(MeshView M, const double* __restrict__ d, double* __restrict__ Ad,
                                                      double* partials, const DevScalars* S) {
    PCFD_BRICK2_BEGIN;
    auto val = [&](int64_t c) -> double { return d[c]; };
    const TileFetch2<double> tf = fetch_tile_x2<double>(M, b, val);
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<double> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<double>(M, sp[h], val);
    }
    if (done) return;
    __shared__ double tile[kTileN];
    stage_tile_x2<double>(tile, tf);
    double sum = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        double y = 0.0;
        if (m[h] & kActive) {
            double dg;
            y = row0f<double>(M, tile, cf[h], val, bs, q, m[h], tf.own[h], dg);
        }
        Ad[base + q] = y;
        sum += y * tf.own[h];
    }
    store_partial(partials, b, block_sum<kBL>(sum));
}

// CG direction update fused into the operator apply: d_new = z + beta d_old (first
// iteration: z), Ad = A d_new and the d_new.Ad partial; d is double-buffered
template <class ZT>
struct ZD {
    ZT z;
    double d;
};
template <class ZT, bool INITIAL>
__global__ __launch_bounds__(kBL, 6) void k_cg_dir_apply(MeshView M, const ZT* __restrict__ z,
                                                      const double* __restrict__ dold, double* __restrict__ dnew,
                                                      double* __restrict__ Ad, double* partials, const DevScalars* S) {
    PCFD_BRICK2_BEGIN;
    const DevScalars& I = S[M.meta[b].inst];
    const double beta = INITIAL ? 0.0 : I.beta;
    auto ld = [&](int64_t c) {
        if constexpr (INITIAL) return z[c];
        else return ZD<ZT>{z[c], dold[c]};
    };
    auto cb = [&](const auto& r) -> double {
        if constexpr (INITIAL) return double(r);
        else return double(r.z) + beta * r.d;
    };
    auto dn = [&](int64_t c) -> double { return cb(ld(c)); };
    const TileFetch2<double> tf = fetch_tile_x2<double>(M, b, ld, NoFill{}, cb);
    const bool done = I.done != 0;
    int sp[2];
    CompactFetch<double> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<double>(M, sp[h], ld, NoFill{}, cb);
    }
    if (done) return;
    __shared__ double tile[kTileN];
    stage_tile_x2<double>(tile, tf);
    double sum = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        const double da = tf.own[h];
        double y = 0.0;
        if (m[h] & kActive) {
            double dg;
            y = row0f<double>(M, tile, cf[h], dn, bs, q, m[h], da, dg);
        }
        dnew[base + q] = da;
        Ad[base + q] = y;
        sum += y * da;
    }
    store_partial(partials, b, block_sum<kBL>(sum));
}

// x += alpha d, r -= alpha Ad (r_prev keeps the old r for flexible CG) and the ||r||^2
// partial, one brick per workgroup with the brick's instance's alpha
__global__ __launch_bounds__(kBL) void k_cg_update(MeshView M, double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ rprev, const double* __restrict__ d,
                                                   const double* __restrict__ Ad, double* partials, const DevScalars* S) {
    const int b = brick_of_block(M.xcd_chunk);
    const int64_t base = static_cast<int64_t>(b) * 512 + threadIdx.x;
    double xi[2], di[2], ri[2], adi[2];
    for (int h = 0; h < 2; ++h) {
        const int64_t i = base + kBL * h;
        xi[h] = x[i]; di[h] = d[i]; ri[h] = r[i]; adi[h] = Ad[i];
    }
    const DevScalars& I = S[M.meta[b].inst];
    if (I.done) return;
    const double alpha = I.alpha;
    double sum = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int64_t i = base + kBL * h;
        x[i] = xi[h] + alpha * di[h];
        if (rprev) rprev[i] = ri[h];
        const double rn = ri[h] - alpha * adi[h];
        r[i] = rn;
        sum += rn * rn;
    }
    store_partial(partials, b, block_sum<kBL>(sum));
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
    FIN_LOADS = 9,    // twelve load sums (pressure and wall-shear force and moment)
    FIN_DIFFRATE = 10, // max explicit diffusion rate (Smagorinsky substeps)
    FIN_RATE_VMAX = 11, // FIN_RATE and FIN_VMAX from two partial arrays
    FIN_BNORM_RESID = 12 // FIN_BNORM, then FIN_RESID, from two partial arrays (||b||^2, ||r||^2)
};

// The instance's scalars that an op reads (done flag, rz, bnorm, iteration count) are loaded
// before the partials are gathered, so the kernel waits for one latency for both.  Thread 0
// applies the op unless `skip`; returns the instance's done flag after the op (thread 0).
__device__ __forceinline__ int finalize_one(int op, const double* __restrict__ partials, int nparts, int beg, int end,
                                            DevScalars* S, bool krylov, double cfl, double rate_floor, double rtol,
                                            int max_iterations, int flexible) {
    int done = S->done;
    const double rz = S->rz, bnorm = S->bnorm;
    const int iter = S->iter;
    const bool take_max = op == FIN_RATE || op == FIN_MAXDIV || op == FIN_VMAX || op == FIN_DIFFRATE || op == FIN_RATE_VMAX;
    const int arrays = op == FIN_LOADS ? 12
                       : (((op == FIN_RZ || op == FIN_RZ_INITIAL) && flexible) || op == FIN_RATE_VMAX || op == FIN_BNORM_RESID ? 2
                                                                                                                           : 1);
    double t[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (arrays == 2) gather_partials_n<1024, 2, 20>(partials + beg, nparts, end - beg, take_max, t);
    else if (arrays == 1) gather_partials_n<1024, 1, 40>(partials + beg, nparts, end - beg, take_max, t);
    else
        for (int c = 0; c < arrays; ++c) t[c] = gather_partials<1024>(partials + c * nparts + beg, end - beg, take_max);
    if (threadIdx.x != 0 || (krylov && done)) return done;
    switch (op) {
        case FIN_RATE:
            S->rate = t[0];
            S->dt = cfl / fmax(rate_floor, t[0]);
            break;
        case FIN_RATE_VMAX:
            S->rate = t[0];
            S->dt = cfl / fmax(rate_floor, t[0]);
            S->vmax = t[1];
            break;
        case FIN_BNORM:
        case FIN_BNORM_RESID: {
            const double bn = sqrt(t[0]);
            S->bnorm = bn;
            S->iter = 0; S->converged = 0; S->breakdown = 0; S->zero_rhs = 0;
            done = 0;
            if (!(t[0] > 0.0)) { done = 1; S->converged = 1; S->zero_rhs = 1; S->relres = 0.0; }
            S->done = done;
            if (op == FIN_BNORM_RESID && !done) {  // FIN_RESID with the new ||b||
                S->rr = t[1];
                const double rel = sqrt(t[1]) / bn;
                S->relres = rel;
                if (rel <= rtol) { done = 1; S->done = 1; S->converged = 1; }
            }
            break;
        }
        case FIN_MAXDIV: S->maxdiv = t[0]; break;
        case FIN_VMAX: S->vmax = t[0]; break;
        case FIN_RESID: {
            S->rr = t[0];
            const double rel = sqrt(t[0]) / bnorm;
            S->relres = rel;
            if (rel <= rtol) { done = 1; S->done = 1; S->converged = 1; }
            break;
        }
        case FIN_APPLY:
            S->pAp = t[0];
            if (!(t[0] > 0.0) || isinf(t[0])) { done = 1; S->done = 1; S->breakdown = 1; }
            else S->alpha = rz / t[0];
            break;
        case FIN_UPDATE: {
            S->rr = t[0];
            S->iter = iter + 1;
            const double rel = sqrt(t[0]) / bnorm;
            S->relres = rel;
            if (rel <= rtol) { done = 1; S->done = 1; S->converged = 1; }
            else if (iter + 1 >= max_iterations || !(t[0] == t[0])) { done = 1; S->done = 1; }
            break;
        }
        case FIN_RZ:
            // Fletcher-Reeves for a fixed preconditioner; Polak-Ribiere (flexible CG) with r_prev
            S->beta = (t[0] - t[1]) / rz;
            S->rz = t[0];
            break;
        case FIN_RZ_INITIAL: S->rz = t[0]; break;
        case FIN_LOADS:
            for (int c = 0; c < 12; ++c) S->loads[c] = t[c];
            break;
        case FIN_DIFFRATE: S->diff_rate = t[0]; break;
        default: break;
    }
    return done;
}

// One workgroup per instance: partial slots [seg[i], seg[i+1]) of each array (arrays are
// nparts apart; seg == nullptr: all nparts slots), scalars of instance i in S[i].  With
// `agg`, agg->done is set when every instance is done: directly for one instance, else by
// the last workgroup to finish.
__global__ __launch_bounds__(1024) void k_finalize(int op, const double* __restrict__ partials, int nparts,
                                                   const int* __restrict__ seg, DevScalars* Sbase, DevScalars* agg,
                                                   double cfl, double rate_floor, double rtol, int max_iterations,
                                                   int flexible) {
    const int inst = blockIdx.x;
    DevScalars* S = Sbase + inst;
    const bool krylov = op == FIN_RESID || op == FIN_APPLY || op == FIN_UPDATE || op == FIN_RZ || op == FIN_RZ_INITIAL;
    const int done = finalize_one(op, partials, nparts, seg ? seg[inst] : 0, seg ? seg[inst + 1] : nparts, S, krylov,
                                  cfl, rate_floor, rtol, max_iterations, flexible);
    if (agg && threadIdx.x == 0) {
        if (gridDim.x == 1) {
            agg->done = done;
            return;
        }
        __threadfence();
        const unsigned int ticket = atomicAdd(&agg->ticket[0], 1u);
        if (ticket == gridDim.x - 1) {  // every instance's update is visible
            __threadfence();
            int all = 1;
            for (unsigned int i = 0; i < gridDim.x; ++i)
                all &= __hip_atomic_load(&Sbase[i].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 1 : 0;
            agg->done = all;
            agg->ticket[0] = 0u;
        }
    }
}

// ---------------------------------------------------------------------------------------
// V-cycle, composite level: damped Jacobi (project.cu:76-83 restated with the composite
// operator).  Every kernel reproduces the arithmetic (including the rounding to the
// storage type VT between stages) of the plain sequence "sweep, sweep, residual, restrict,
// prolong, sweep, sweep"; the fused ones only remove HBM round trips.
// ---------------------------------------------------------------------------------------

// composite-level diagonal in the V-cycle arithmetic type (setup, once)
template <class DT>
__global__ __launch_bounds__(512) void k_diag0(MeshView M, DT* __restrict__ diag) {
    PCFD_BRICK_PROLOGUE;
    diag[a] = (m & kActive) ? diag_row0<DT>(M, b, q, m, sp, DT(level_h(M, b))) : DT(1);
}

// One sweep.  mode 0: x = w b / D (from zero); 1: x += w (b - A x) / D; 2: last sweep,
// writes z (ZT: FP64, or FP32 when the sweep's arithmetic is FP32 and z is exact in it) and
// the r.z (and r_prev.z) partials (two arrays of gridDim.x entries).
template <class VT, class ZT = double>
__global__ __launch_bounds__(kBL) void k_smooth0(MeshView M, int mode, const VT* __restrict__ xin, VT* __restrict__ xout,
                                                 const double* __restrict__ rhs, ZT* __restrict__ zout,
                                                 const double* __restrict__ rprev, double omega, double* partials,
                                                 const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK2_BEGIN;
    const bool from_zero = mode == 0 || xin == nullptr;
    // from zero every neighbour value is 0: the loads read a valid array of at least n values
    // and are discarded, so that no load sits behind a branch
    const VT* xsrc = from_zero ? reinterpret_cast<const VT*>(rhs) : xin;
    auto ld = [&](int64_t c) -> VT { return xsrc[c]; };
    auto cb = [&](VT v) -> A { return from_zero ? A(0) : A(to_f(v)); };
    auto val = [&](int64_t c) -> A { return cb(ld(c)); };
    const TileFetch2<A> tf = fetch_tile_x2<A>(M, b, ld, NoFill{}, cb);
    double rb[2], rp[2];
    for (int h = 0; h < 2; ++h) {
        rb[h] = rhs[base + t + kBL * h];
        rp[h] = (mode == 2 && rprev) ? rprev[base + t + kBL * h] : 0.0;
    }
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<A> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<A>(M, sp[h], ld, NoFill{}, cb);
    }
    if (done) return;
    __shared__ A tile[kTileN];
    stage_tile_x2<A>(tile, tf);
    double s1 = 0.0, s2 = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        A v = A(0);
        if (m[h] & kActive) {
            const A hb = A(bs.h);
            if (from_zero) {
                v = A(omega) * A(rb[h]) / diag_row0<A>(M, b, q, m[h], sp[h], hb);
            } else {
                A dg;
                const A xa = tf.own[h];
                const A Ax = row0f<A>(M, tile, cf[h], val, bs, q, m[h], xa, dg);
                v = xa + A(omega) * (A(rb[h]) - Ax) / dg;
            }
        }
        if (mode != 2) {
            xout[base + q] = from_f<VT>(v);
            continue;
        }
        const double z = double(v);
        zout[base + q] = static_cast<ZT>(v);
        s1 += rb[h] * z;
        s2 += rp[h] * z;
    }
    if (mode != 2) return;
    s1 = block_sum<kBL>(s1);
    s2 = rprev ? block_sum<kBL>(s2) : 0.0;
    if (threadIdx.x == 0) {
        partials[b] = s1;
        partials[M.nbricks + b] = s2;
    }
}

// residual b - A x on the composite level (unfused path and the primitive entry point)
template <class VT>
__global__ __launch_bounds__(512) void k_residual0(MeshView M, const VT* __restrict__ x, const double* __restrict__ rhs,
                                                   VT* __restrict__ res, const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK_BEGIN;
    auto val = [&](int64_t c) -> A { return A(to_f(x[c])); };
    const TileFetch<A> tf = fetch_tile<A>(M, b, val);
    const double rb = rhs[a];
    const bool done = brick_done(M, S, b);
    const int sp = special_rank(M, b, q, (m & kSpecial) != 0);
    const CompactFetch<A> cf = fetch_compact<A>(M, sp, val);
    if (done) return;
    __shared__ A tile[kTileN];
    stage_tile<A>(tile, tf);
    A v = A(0);
    if (m & kActive) {
        A dg;
        v = A(rb) - row0f<A>(M, tile, cf, val, bs, q, m, tf.own, dg);
    }
    res[a] = from_f<VT>(v);
}

// two pre-smoothing sweeps from x = 0:  x1 = w1 r / D,  x2 = x1 + w2 (r - A x1) / D
template <class VT>
__global__ __launch_bounds__(kBL, sizeof(VT) == 8 ? 6 : 8) void k_vc_pre2(MeshView M, const double* __restrict__ rhs,
                                                 const typename Arith<VT>::type* __restrict__ diag, VT* __restrict__ xout,
                                                 double omega, double omega2, const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK2_BEGIN;
    const A w = A(omega), w2 = A(omega2);
    struct Raw { double r; A dg; };
    auto ld = [&](int64_t c) { return Raw{rhs[c], diag[c]}; };
    auto cb = [&](const Raw& v) -> A { return round_vt<VT>(w * A(v.r) / v.dg); };
    auto x1 = [&](int64_t c) -> A { return cb(ld(c)); };
    const TileFetch2<A> tf = fetch_tile_x2<A>(M, b, ld, NoFill{}, cb);
    double rb[2];
    for (int h = 0; h < 2; ++h) rb[h] = rhs[base + t + kBL * h];
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<A> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<A>(M, sp[h], ld, NoFill{}, cb);
    }
    if (done) return;
    __shared__ A tile[kTileN];
    stage_tile_x2<A>(tile, tf);
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        A v = A(0);
        if (m[h] & kActive) {
            A dg;
            const A xa = tf.own[h];
            const A Ax = row0f<A>(M, tile, cf[h], x1, bs, q, m[h], xa, dg);
            v = xa + w2 * (A(rb[h]) - Ax) / dg;
        }
        xout[base + q] = from_f<VT>(v);
    }
}

// The CG update fused into the first V-cycle pass: x += alpha d, r = rin - alpha Ad (written
// to rout, never in place: neighbour bricks read rin for their halo) and the ||r||^2
// partial, then the two pre-smoothing sweeps from zero on the new r, whose halo and
// compact-neighbour values are formed on the fly with the same arithmetic.  The results are
// those of k_cg_update followed by k_vc_pre2.
template <class VT>
__global__ __launch_bounds__(kBL, sizeof(VT) == 8 ? 5 : 7) void k_vc_update_pre2(MeshView M, double* __restrict__ x, const double* __restrict__ rin,
                                                        double* __restrict__ rout, const double* __restrict__ d,
                                                        const double* __restrict__ Ad,
                                                        const typename Arith<VT>::type* __restrict__ diag,
                                                        VT* __restrict__ xout, double omega, double omega2,
                                                        double* partials, const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK2_BEGIN;
    const DevScalars& I = S[M.meta[b].inst];
    const double alpha = I.alpha;
    const A w = A(omega), w2 = A(omega2);
    struct Raw { double r, ad; A dg; };
    auto ld = [&](int64_t c) { return Raw{rin[c], Ad[c], diag[c]}; };
    auto cb = [&](const Raw& v) -> A { return round_vt<VT>(w * A(v.r - alpha * v.ad) / v.dg); };
    auto x1 = [&](int64_t c) -> A { return cb(ld(c)); };
    Raw own[2];
    const TileFetch2<A> tf = fetch_tile_x2<A>(M, b, ld, NoFill{}, cb, own);
    double xi[2], di[2], ri[2], adi[2];
    for (int h = 0; h < 2; ++h) {
        const int64_t a = base + t + kBL * h;
        xi[h] = x[a]; di[h] = d[a]; ri[h] = own[h].r; adi[h] = own[h].ad;
    }
    const bool done = I.done != 0;
    int sp[2];
    CompactFetch<A> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<A>(M, sp[h], ld, NoFill{}, cb);
    }
    if (done) return;
    double rn[2], sum = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int64_t a = base + t + kBL * h;
        rn[h] = ri[h] - alpha * adi[h];
        x[a] = xi[h] + alpha * di[h];
        rout[a] = rn[h];
        sum += rn[h] * rn[h];
    }
    __shared__ A tile[kTileN];
    stage_tile_x2<A>(tile, tf);
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        A v = A(0);
        if (m[h] & kActive) {
            A dg;
            const A xa = tf.own[h];
            const A Ax = row0f<A>(M, tile, cf[h], x1, bs, q, m[h], xa, dg);
            v = xa + w2 * (A(rn[h]) - Ax) / dg;
        }
        xout[base + q] = from_f<VT>(v);
    }
    store_partial(partials, b, block_sum<kBL>(sum));
}

// residual r - A x summed over each 2x2x2 group into its level-1 parent (children of one
// parent always lie in one brick); cells that were not merged map one-to-one
template <class VT>
__global__ __launch_bounds__(kBL) void k_vc_restrict0(MeshView M, const VT* __restrict__ x, const double* __restrict__ rhs,
                                                      const int* __restrict__ parent0, VT* __restrict__ coarse,
                                                      const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK2_BEGIN;
    auto ld = [&](int64_t c) -> VT { return x[c]; };
    auto cb = [&](VT v) -> A { return A(to_f(v)); };
    auto val = [&](int64_t c) -> A { return cb(ld(c)); };
    const TileFetch2<A> tf = fetch_tile_x2<A>(M, b, ld, NoFill{}, cb);
    double rb[2];
    int pa[2];
    for (int h = 0; h < 2; ++h) {
        rb[h] = rhs[base + t + kBL * h];
        pa[h] = parent0[base + t + kBL * h];
    }
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<A> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<A>(M, sp[h], ld, NoFill{}, cb);
    }
    if (done) return;
    __shared__ A tile[kTileN];
    __shared__ A sres[512];
    __shared__ int spar[512];
    stage_tile_x2<A>(tile, tf);
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        A v = A(0);
        if (m[h] & kActive) {
            A dg;
            v = A(rb[h]) - row0f<A>(M, tile, cf[h], val, bs, q, m[h], tf.own[h], dg);
        }
        sres[q] = round_vt<VT>(v);
        spar[q] = pa[h];
    }
    __syncthreads();
    if ((t & 0x49) != 0) return;  // leaders have even i, j, k (lane t leads cells t and t + 256)
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        const int kids[8] = {q, q + 1, q + 8, q + 9, q + 64, q + 65, q + 72, q + 73};
        int P = -1;
        bool shared = true;
        for (int k = 0; k < 8; ++k) {
            const int pc = spar[kids[k]];
            if (pc < 0) continue;
            if (P < 0) P = pc;
            else if (pc != P) shared = false;
        }
        if (P < 0) continue;
        if (shared) {
            A sum = A(0);
            for (int k = 0; k < 8; ++k) if (spar[kids[k]] >= 0) sum += sres[kids[k]];
            coarse[P] = from_f<VT>(sum);
        } else {
            for (int k = 0; k < 8; ++k) {
                const int pc = spar[kids[k]];
                if (pc >= 0) coarse[pc] = from_f<VT>(A(0) + sres[kids[k]]);
            }
        }
    }
}

// prolongation x += e[parent] fused with one post-smoothing sweep; `last` writes z (ZT, as
// k_smooth0) and the r.z / r_prev.z partials like k_smooth0 mode 2
template <class VT, class ZT = double>
__global__ __launch_bounds__(kBL) void k_vc_post_prolong(MeshView M, int last, const VT* __restrict__ x,
                                                         const VT* __restrict__ e1, const int* __restrict__ parent0,
                                                         const double* __restrict__ rhs, VT* __restrict__ xout,
                                                         ZT* __restrict__ zout, const double* __restrict__ rprev,
                                                         double omega, double* partials, const DevScalars* S) {
    using A = typename Arith<VT>::type;
    PCFD_BRICK2_BEGIN;
    // x and the parent index, then the parent's correction (index 0 for unmerged cells,
    // discarded), then the sum
    struct Raw { VT x; int pc; VT e; };
    auto ld = [&](int64_t c) { return Raw{x[c], parent0[c], VT{}}; };
    auto fl = [&](Raw& r) { r.e = e1[r.pc >= 0 ? r.pc : 0]; };
    auto cb = [&](const Raw& r) -> A {
        const A xc = A(to_f(r.x));
        return r.pc >= 0 ? round_vt<VT>(xc + A(to_f(r.e))) : xc;
    };
    auto y = [&](int64_t c) -> A {
        Raw r = ld(c);
        fl(r);
        return cb(r);
    };
    const TileFetch2<A> tf = fetch_tile_x2<A>(M, b, ld, fl, cb);
    double rb[2], rp[2];
    for (int h = 0; h < 2; ++h) {
        rb[h] = rhs[base + t + kBL * h];
        rp[h] = (last && rprev) ? rprev[base + t + kBL * h] : 0.0;
    }
    const bool done = brick_done(M, S, b);
    int sp[2];
    CompactFetch<A> cf[2];
    for (int h = 0; h < 2; ++h) {
        sp[h] = special_rank(M, b, t + kBL * h, (m[h] & kSpecial) != 0);
        cf[h] = fetch_compact<A>(M, sp[h], ld, fl, cb);
    }
    if (done) return;
    __shared__ A tile[kTileN];
    stage_tile_x2<A>(tile, tf);
    double s1 = 0.0, s2 = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        A v = A(0);
        if (m[h] & kActive) {
            A dg;
            const A ya = tf.own[h];
            const A Ay = row0f<A>(M, tile, cf[h], y, bs, q, m[h], ya, dg);
            v = ya + A(omega) * (A(rb[h]) - Ay) / dg;
        }
        if (!last) {
            xout[base + q] = from_f<VT>(v);
            continue;
        }
        const double z = double(v);
        zout[base + q] = static_cast<ZT>(v);
        s1 += rb[h] * z;
        s2 += rp[h] * z;
    }
    if (!last) return;
    s1 = block_sum<kBL>(s1);
    s2 = rprev ? block_sum<kBL>(s2) : 0.0;
    if (threadIdx.x == 0) {
        partials[b] = s1;
        partials[M.nbricks + b] = s2;
    }
}

// unfused transfers (primitive entry point and pre/post counts other than the fused ones)
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

// ---------------------------------------------------------------------------------------
// Coarse levels
// ---------------------------------------------------------------------------------------
// (A y)_r = D_r y_r - sum coef y_col, entries in CSR order (the SELL slots hold them in
// order, then zero padding)
template <class A, class CT, class Fn>
__device__ __forceinline__ A ell_row(const CsrView& C, int r, Fn y, A yr) {
    const CT* coef = static_cast<const CT*>(C.sl_coef);
    const int s0 = C.sl_off[r >> 6], s1 = C.sl_off[(r >> 6) + 1];
    A Ax = A(static_cast<const CT*>(C.diag)[r]) * yr;
    const CT u = static_cast<const CT*>(C.sl_uni)[r >> 6];
    if (u == u) {  // slice-uniform: the same coefficient in every slot, columns only
        const A ua = A(u);
        for (int s = s0 + (r & 63); s < s1; s += 512) {
            int c[8];
            A v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) c[j] = s + 64 * j < s1 ? C.sl_col[s + 64 * j] : r;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = y(c[j]);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (s + 64 * j < s1) Ax -= ua * v[j];
        }
        return Ax;
    }
    // eight slots per round (most coarse rows have at most eight): their columns and
    // coefficients, then their values, are loaded before any is used; subtracted in slot order
    for (int s = s0 + (r & 63); s < s1; s += 512) {
        int c[8];
        A k[8], v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool ok = s + 64 * j < s1;
            c[j] = ok ? C.sl_col[s + 64 * j] : r;
            k[j] = ok ? A(coef[s + 64 * j]) : A(0);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = y(c[j]);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (s + 64 * j < s1) Ax -= k[j] * v[j];
    }
    return Ax;
}

// CSR forms (multigrid primitive entry point; same arithmetic as the ELL kernels)
template <class VT, class CT>
__global__ void k_smooth_csr(CsrView C, int mode, const VT* __restrict__ xin, VT* __restrict__ xout,
                             const VT* __restrict__ rhs, double omega) {
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
__global__ void k_residual_csr(CsrView C, const VT* __restrict__ x, const VT* __restrict__ rhs, VT* __restrict__ res) {
    using A = typename Arith<VT>::type;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= C.n) return;
    const CT* coef = static_cast<const CT*>(C.coef);
    const CT* diag = static_cast<const CT*>(C.diag);
    A Ax = A(diag[r]) * A(to_f(x[r]));
    for (int k = C.row_off[r]; k < C.row_off[r + 1]; ++k) Ax -= A(coef[k]) * A(to_f(x[C.col[k]]));
    res[r] = from_f<VT>(A(to_f(rhs[r])) - Ax);
}

// Coarse-level rows: one lane per row, each operation a per-row device function.

// ELL sweep: mode 0 from zero, 1 damped Jacobi
template <class VT, class CT>
__device__ __forceinline__ void ell_smooth_row(const CsrView& C, int r, int mode, const VT* __restrict__ xin,
                                               VT* __restrict__ xout, const VT* __restrict__ rhs, double omega) {
    using A = typename Arith<VT>::type;
    const A D = A(static_cast<const CT*>(C.diag)[r]);
    const A br = A(to_f(rhs[r]));
    if (mode == 0) { xout[r] = from_f<VT>(A(omega) * br / D); return; }
    const A xr = A(to_f(xin[r]));
    const A Ax = ell_row<A, CT>(C, r, [&](int c) -> A { return A(to_f(xin[c])); }, xr);
    xout[r] = from_f<VT>(xr + A(omega) * (br - Ax) / D);
}

// two sweeps from zero in one pass (x1 = w1 b / D at the neighbours on the fly, then weight w2)
template <class VT, class CT>
__device__ __forceinline__ void ell_pre2_row(const CsrView& C, int r, const VT* __restrict__ rhs, VT* __restrict__ xout,
                                             double omega, double omega2) {
    using A = typename Arith<VT>::type;
    const CT* diag = static_cast<const CT*>(C.diag);
    const A w = A(omega);
    auto x1 = [&](int c) -> A { return round_vt<VT>(w * A(to_f(rhs[c])) / A(diag[c])); };
    const A xr = x1(r);
    const A Ax = ell_row<A, CT>(C, r, x1, xr);
    xout[r] = from_f<VT>(xr + A(omega2) * (A(to_f(rhs[r])) - Ax) / A(diag[r]));
}

// residual b - A x
template <class VT, class CT>
__device__ __forceinline__ void ell_residual_row(const CsrView& C, int r, const VT* __restrict__ x,
                                                 const VT* __restrict__ rhs, VT* __restrict__ res) {
    using A = typename Arith<VT>::type;
    const A xr = A(to_f(x[r]));
    const A Ax = ell_row<A, CT>(C, r, [&](int j) -> A { return A(to_f(x[j])); }, xr);
    res[r] = from_f<VT>(A(to_f(rhs[r])) - Ax);
}

// restriction from a coarse level (m >= 1): the children of a coarse cell are a contiguous
// ascending range of the finer level's rows (Morton order), summed in that order
template <class VT>
__device__ __forceinline__ void restrict_range_row(const CsrView& Cc, int c, const VT* __restrict__ fine,
                                                   VT* __restrict__ coarse) {
    using A = typename Arith<VT>::type;
    const int k0 = Cc.child_off[c], k1 = Cc.child_off[c + 1], f0 = Cc.child_first[c];
    A s = A(0);
    // eight children per round, all loaded before the first is added
    for (int k = k0; k < k1; k += 8) {
        A v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = k + j < k1 ? A(to_f(fine[f0 + (k + j - k0)])) : A(0);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k + j < k1) s += v[j];
    }
    coarse[c] = from_f<VT>(s);
}

// prolongation from level m+1 fused with one sweep on level m
template <class VT, class CT>
__device__ __forceinline__ void ell_post_prolong_row(const CsrView& C, int r, const VT* __restrict__ x,
                                                     const VT* __restrict__ e, const VT* __restrict__ rhs,
                                                     VT* __restrict__ xout, double omega) {
    using A = typename Arith<VT>::type;
    auto y = [&](int c) -> A { return round_vt<VT>(A(to_f(x[c])) + A(to_f(e[C.parent[c]]))); };
    const A yr = y(r);
    const A Ay = ell_row<A, CT>(C, r, y, yr);
    xout[r] = from_f<VT>(yr + A(omega) * (A(to_f(rhs[r])) - Ay) / A(static_cast<const CT*>(C.diag)[r]));
}

template <class VT, class CT>
__global__ void k_ell_smooth(CsrView C, int mode, const VT* __restrict__ xin, VT* __restrict__ xout,
                             const VT* __restrict__ rhs, double omega) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < C.n) ell_smooth_row<VT, CT>(C, r, mode, xin, xout, rhs, omega);
}

template <class VT, class CT>
__global__ void k_ell_pre2(CsrView C, const VT* __restrict__ rhs, VT* __restrict__ xout, double omega, double omega2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < C.n) ell_pre2_row<VT, CT>(C, r, rhs, xout, omega, omega2);
}

template <class VT, class CT>
__global__ void k_ell_residual(CsrView C, const VT* __restrict__ x, const VT* __restrict__ rhs, VT* __restrict__ res) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < C.n) ell_residual_row<VT, CT>(C, r, x, rhs, res);
}

template <class VT>
__global__ void k_restrict_range(CsrView Cc, const VT* __restrict__ fine, VT* __restrict__ coarse) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < Cc.n) restrict_range_row<VT>(Cc, c, fine, coarse);
}

// residual of level m fused with its restriction to level m+1 where the children of
// consecutive coarse rows lie far apart: eight lanes per coarse row (at most eight children, a
// contiguous range), lane j computes child j's residual (rounded to VT, as stored by the
// unfused pair), the group's first lane sums them in order.  Every lane reaches the shuffles.
template <class VT, class CT>
__global__ __launch_bounds__(256) void k_ell_residual_restrict_rows(CsrView C, CsrView NC, const VT* __restrict__ x,
                                                                    const VT* __restrict__ rhs, VT* __restrict__ coarse) {
    using A = typename Arith<VT>::type;
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 3, j = threadIdx.x & 7;
    const int lane = static_cast<int>(__lane_id()), g0 = lane & ~7;
    int k0 = 0, n = 0, f0 = 0;
    if (c < NC.n) {
        k0 = NC.child_off[c];
        n = NC.child_off[c + 1] - k0;
        f0 = NC.child_first[c];
    }
    A v = A(0);
    if (j < n) {
        const int r = f0 + j;
        const A xr = A(to_f(x[r]));
        const A Ax = ell_row<A, CT>(C, r, [&](int q) -> A { return A(to_f(x[q])); }, xr);
        v = A(to_f(from_f<VT>(A(to_f(rhs[r])) - Ax)));
    }
    A s = A(0);
    for (int k = 0; k < 8; ++k) {
        const A u = __shfl(v, g0 + k);
        if (k < n) s += u;
    }
    if (j == 0 && c < NC.n) coarse[c] = from_f<VT>(s);
}

// residual of level m fused with its restriction to level m+1: a workgroup takes
// kRestrictGroup consecutive rows of level m+1, whose children form one window of at most 256
// consecutive rows of level m (checked on the host); each lane computes one row's residual
// (rounded to VT, as stored by the unfused pair), then each coarse row sums its children in
// order from LDS.  Rows of the window without a parent in the group are computed and unused.
constexpr int kRestrictGroup = 32;
template <class VT, class CT>
__global__ __launch_bounds__(256) void k_ell_residual_restrict(CsrView C, CsrView NC, const VT* __restrict__ x,
                                                               const VT* __restrict__ rhs, VT* __restrict__ coarse) {
    using A = typename Arith<VT>::type;
    __shared__ A res[256];
    const int c0 = blockIdx.x * kRestrictGroup, c1 = min(c0 + kRestrictGroup, NC.n);
    const int f0 = NC.child_first[c0];
    const int fend = NC.child_first[c1 - 1] + (NC.child_off[c1] - NC.child_off[c1 - 1]);
    const int t = threadIdx.x, r = f0 + t;
    if (r < fend) {
        const A xr = A(to_f(x[r]));
        const A Ax = ell_row<A, CT>(C, r, [&](int j) -> A { return A(to_f(x[j])); }, xr);
        res[t] = round_vt<VT>(A(to_f(rhs[r])) - Ax);
    }
    __syncthreads();
    if (t < c1 - c0) {
        const int c = c0 + t;
        const int n = NC.child_off[c + 1] - NC.child_off[c], k = NC.child_first[c] - f0;
        A s = A(0);
        for (int j = 0; j < n; ++j) s += res[k + j];
        coarse[c] = from_f<VT>(s);
    }
}

template <class VT, class CT>
__global__ void k_ell_post_prolong(CsrView C, const VT* __restrict__ x, const VT* __restrict__ e, const VT* __restrict__ rhs,
                                   VT* __restrict__ xout, double omega) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < C.n) ell_post_prolong_row<VT, CT>(C, r, x, e, rhs, xout, omega);
}

// coarsest level of each instance: x = A^-1 b with its dense inverse; block (row, k) solves
// row `row` of dense block k (cells [off[k], off[k] + n[k]) of the level, inverse at ioff[k]):
// lane t sums the columns t, t + 256, ... in that order, then the block sum
template <class VT, class CT>
__global__ __launch_bounds__(256) void k_dense_solve(const int* __restrict__ off, const int* __restrict__ nrow,
                                                     const int64_t* __restrict__ ioff, const CT* __restrict__ inv,
                                                     const VT* __restrict__ rhs, VT* __restrict__ x) {
    const int k = blockIdx.y, r = blockIdx.x, n = nrow[k];
    if (r >= n) return;  // block-uniform
    const CT* row = inv + ioff[k] + static_cast<int64_t>(r) * n;
    const VT* b = rhs + off[k];
    double acc = 0.0;
    // eight columns per lane and round, loaded before they are accumulated in column order
    for (int c0 = threadIdx.x; c0 < n; c0 += 8 * 256) {
        double w[8], v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = c0 + 256 * j;
            w[j] = c < n ? double(row[c]) : 0.0;
            v[j] = c < n ? double(to_f(b[c])) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (c0 + 256 * j < n) acc += w[j] * v[j];
    }
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0) x[off[k] + r] = from_f<VT>(static_cast<typename Arith<VT>::type>(acc));
}

#undef PCFD_BRICK_PROLOGUE
#undef PCFD_BRICK_BEGIN

}  // namespace pcfd
