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
#include <type_traits>
#include <utility>

namespace pcfd {

constexpr int kBrickThreads = 512;
constexpr uint8_t kPX = 1, kMX = 8, kActive = 64, kSpecial = 128;
constexpr uint8_t kXMIN = 1, kXMAX = 2, kYMIN = 4, kYMAX = 8, kZMIN = 16, kZMAX = 32;

// Scalars the Krylov loop keeps on the device (no host round trip per iteration), one
// record per instance of a batch; record [ninst] is the batch aggregate (done when every
// instance is done), read by the kernels that do not know their instance.
struct DevScalars {
    double rz, rz_prev, pAp, alpha, beta, rr, bnorm, relres;
    double dt, rate, maxdiv, vmax, diff_rate;
    double dts;       // diffusion substep of this instance
    double loads[12];  // pressure force, moment; wall-shear force, moment
    int iter, done, converged, breakdown, zero_rhs, subs;
    unsigned int ticket[16];
};

// Per-brick record, 64 bytes, read by a brick kernel with scalar loads before any cell
// data: neighbours for the halo, the special-cell ranks, instance, cell size, boundary bits.
struct alignas(64) BrickMeta {
    int nbr[6];            // same-level neighbour brick per face (-x, +x, -y, +y, -z, +z) or -1
    int special_off;       // rank of the brick's first special cell
    int inst;              // instance of the batch
    double h;              // cell size of the brick's level
    uint16_t flags;        // domain-boundary bits (kXMIN .. kZMAX)
    uint16_t level;
    uint16_t wave_pre[8];  // special cells of the brick in the waves before wave w
    int pad;
};
static_assert(sizeof(BrickMeta) == 64, "BrickMeta is one 64-byte record");

struct MeshView {
    const BrickMeta* meta;
    const int* sp_other;     // per special incidence (special_inc order): the other cell's dof
    const double* sp_coef;   // and the edge conductance
    const int* sp4_other;    // the first four incidences of special cell s at 4 s + j (padding:
    const double* sp4_coef;  //   the cell itself with conductance 0)
    int xcd_chunk;           // log2 of the bricks per XCD chunk of the block -> brick order (-1: one range per XCD)
    int nbricks;
    int64_t ndof;
    int ninst;               // instances of the batch (1 for a single solver)
    const int* inst;         // instance of each brick
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
    const double* edcoef;    // diffusion conductance of each edge
    const double* edisp;     // 6 per edge: face centroid minus the centroid of a, then of b
    const int* esp;          // 2 per edge: special ranks of a and b
    const int* special_dof;  // dof of each special cell
    int nspecial;
    int npatches;
    const int* patch_dof;
    const int* patch_sp;     // special rank of the patch's cell
    const double* patch_S;
    const double* patch_x;
    const double* patch_D;   // 9 per patch: int_W (x - c) (x) n dA, row-major
    int ngroups;             // momentum redistribution groups
    const int* rgroup;       // per special cell: its group or -1
    const int* roff;         // groups + 1
    const int* rmember;      // special ranks of each group's members
    const double* rvolume;   // per group: fluid volume
    double* rdelta;          // 3 per special cell: a member's momentum increment (scratch)
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
    const int* child_first;  // child[child_off[c]] (or -1 without children): one load, no chain
    // the same rows in SELL-64: slices of 64 rows, slice s holds its rows slot-major at
    // sl_off[s] + 64 k + (r & 63) with k < (sl_off[s+1] - sl_off[s]) / 64, padded with
    // (column r, coefficient 0); coalesced, and only as wide as its longest row
    const int* sl_off;
    const int* sl_col;
    const void* sl_coef;
    // per slice: the coefficient every slot of the slice holds, when the slice has no padding and
    // one coefficient throughout (regular coarse cells), else NaN; its rows load no coefficients
    const void* sl_uni;
};

// LDS tile of one brick with a one-cell halo: 10 x 10 x 10 values
constexpr int kTileN = 1000;
__device__ __forceinline__ int tile_index(int i, int j, int k) { return (k + 1) * 100 + (j + 1) * 10 + (i + 1); }

// Values of one lane for a brick tile: its own cell and, for lanes q < 384, one face-halo
// cell of a same-level neighbour brick (zero where there is none).  The loads are issued
// here and consumed by stage_tile, so a kernel can issue its other per-cell loads (and test
// its done flag) while they are in flight.  Wave w < 6 serves face w, so the neighbour index
// is a scalar read of the brick record.
template <class A>
struct TileFetch {
    A own, halo;
    int hidx;
};

template <class A, class Fn>
__device__ __forceinline__ TileFetch<A> fetch_tile(const MeshView& M, int b, Fn val) {
    const int q = threadIdx.x;
    TileFetch<A> f;
    f.own = val(static_cast<int64_t>(b) * 512 + q);
    f.halo = A(0);
    f.hidx = 0;
    if (q < 384) {
        const int face = __builtin_amdgcn_readfirstlane(q >> 6), t = q & 63, axis = face >> 1, up = face & 1;
        const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        int c[3];
        c[a1] = t & 7;
        c[a2] = t >> 3;
        c[axis] = up ? 8 : -1;
        f.hidx = tile_index(c[0], c[1], c[2]);
        const int nb = M.meta[b].nbr[face];
        if (nb >= 0) {
            c[axis] = up ? 0 : 7;
            f.halo = val(static_cast<int64_t>(nb) * 512 + c[0] + 8 * c[1] + 64 * c[2]);
        }
    }
    return f;
}

// Write the fetched values into the tile and synchronise.  Every lane must call it.
template <class A>
__device__ __forceinline__ void stage_tile(A* tile, const TileFetch<A>& f) {
    const int q = threadIdx.x;
    tile[tile_index(q & 7, (q >> 3) & 7, q >> 6)] = f.own;
    if (q < 384) tile[f.hidx] = f.halo;
    __syncthreads();
}

// The projection's composite-level kernels run kBL = 256 lanes per brick, lane t owning
// cells t and t + 256, so that each wave has the loads of two cells in flight: these kernels
// wait on memory latency, not bandwidth (a Jacobi sweep of the 19.8 M-cell wing: 155 us with
// one cell per lane, 106 us with two, 120 us with four).  The tile of a brick is the same; a
// lane stages its two cells and up to two of the 384 face-halo cells (slots t and t + 256;
// waves 0-3 serve faces 0-3, waves 0-1 faces 4-5 as well).
constexpr int kBL = 256;

template <class A>
struct TileFetch2 {
    A own[2], halo[2];
    int hidx[2];
};

// Values gathered in three phases so that every load of a phase is in flight before the
// first is used: `load(c)` reads cell c's raw values (no arithmetic on them), `fill(raw)` the
// loads that depend on them (an index read in `load`), `comb(raw)` forms the value.  With a
// `comb`, the halo slots are filled without branches: a lane whose slot has no neighbour brick
// (or that has no slot) reads its own cell and keeps 0, so the loads of one phase form one
// basic block.  Plain loads (no fill, no comb) skip the missing slots instead.
struct NoFill {
    template <class T>
    __device__ __forceinline__ void operator()(T&) const {}
};
struct Same {
    template <class T>
    __device__ __forceinline__ T operator()(const T& v) const { return v; }
};

// `keep` (if not null) receives the raw values of the lane's two cells
template <class A, class Ld, class Fl, class Cb, class Raw = decltype(std::declval<Ld>()(int64_t(0)))>
__device__ __forceinline__ TileFetch2<A> fetch_tile_x2(const MeshView& M, int b, Ld load, Fl fill, Cb comb,
                                                       Raw* keep = nullptr) {
    const int t = threadIdx.x;
    TileFetch2<A> f;
    const int64_t base = static_cast<int64_t>(b) * 512;
    constexpr bool kPure = std::is_same<Fl, NoFill>::value && std::is_same<Cb, Same>::value;
    Raw ro[2], rh[2];
    bool hv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) ro[h] = load(base + t + kBL * h);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int s = t + kBL * r;
        const bool slot = s < 384;  // uniform over a wave
        const int face = __builtin_amdgcn_readfirstlane(slot ? s >> 6 : 0), tt = s & 63, axis = face >> 1, up = face & 1;
        const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        int c[3];
        c[a1] = tt & 7;
        c[a2] = tt >> 3;
        c[axis] = up ? 8 : -1;
        f.hidx[r] = slot ? tile_index(c[0], c[1], c[2]) : -1;
        const int nb = M.meta[b].nbr[face];
        hv[r] = slot && nb >= 0;
        c[axis] = up ? 0 : 7;
        const int64_t src = static_cast<int64_t>(nb) * 512 + c[0] + 8 * c[1] + 64 * c[2];
        if constexpr (kPure) {
            // plain loads: nothing uses them before the other loads are issued, so the
            // lanes without a neighbour value skip the read
            rh[r] = Raw{};
            if (hv[r]) rh[r] = load(src);
        } else {
            rh[r] = load(hv[r] ? src : base + t);
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) fill(ro[h]);
#pragma unroll
    for (int r = 0; r < 2; ++r) fill(rh[r]);
#pragma unroll
    for (int h = 0; h < 2; ++h) f.own[h] = comb(ro[h]);
#pragma unroll
    for (int r = 0; r < 2; ++r) f.halo[r] = hv[r] ? comb(rh[r]) : A(0);
    if (keep) {
        keep[0] = ro[0];
        keep[1] = ro[1];
    }
    return f;
}

// one-function form: `val(c)` loads and forms the value of cell c
template <class A, class Fn>
__device__ __forceinline__ TileFetch2<A> fetch_tile_x2(const MeshView& M, int b, Fn val) {
    return fetch_tile_x2<A>(M, b, val, NoFill{}, Same{});
}

// Write the fetched values into the tile and synchronise (unless `sync` is false: the caller
// stages more tiles and synchronises once).  Every lane must call it.
template <class A>
__device__ __forceinline__ void stage_tile_x2(A* tile, const TileFetch2<A>& f, bool sync = true) {
    const int t = threadIdx.x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = t + kBL * h;
        tile[tile_index(q & 7, (q >> 3) & 7, q >> 6)] = f.own[h];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
        if (f.hidx[r] >= 0) tile[f.hidx[r]] = f.halo[r];
    if (sync) __syncthreads();
}

// The first four compact-edge terms of a special cell, fetched before the tile barrier (the
// other cells' values and the conductances), and its incidence range for the rest.
template <class A>
struct CompactFetch {
    A c[4], v[4];
    int k0, k1;
};

template <class A, class Ld, class Fl, class Cb>
__device__ __forceinline__ CompactFetch<A> fetch_compact(const MeshView& M, int sp, Ld load, Fl fill, Cb comb) {
    CompactFetch<A> f;
    f.k0 = f.k1 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) { f.c[j] = A(0); f.v[j] = A(0); }
    if (sp >= 0) {
        f.k0 = M.special_inc_off[sp];
        f.k1 = M.special_inc_off[sp + 1];
        int o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            o[j] = M.sp4_other[4 * static_cast<int64_t>(sp) + j];
            f.c[j] = A(M.sp4_coef[4 * static_cast<int64_t>(sp) + j]);
        }
        using Raw = decltype(load(int64_t(0)));
        Raw rv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) rv[j] = load(static_cast<int64_t>(o[j]));
#pragma unroll
        for (int j = 0; j < 4; ++j) fill(rv[j]);
#pragma unroll
        for (int j = 0; j < 4; ++j) f.v[j] = comb(rv[j]);
    }
    return f;
}

template <class A, class Fn>
__device__ __forceinline__ CompactFetch<A> fetch_compact(const MeshView& M, int sp, Fn val) {
    return fetch_compact<A>(M, sp, val, NoFill{}, Same{});
}

// compact-edge part of (A y)_a from a fetch (in incidence order; incidences past the fourth
// are gathered here)
template <class A, class Fn>
__device__ __forceinline__ void compact_apply_fetched(const MeshView& M, const CompactFetch<A>& f, Fn val, A ya, A& acc,
                                                      A& dg) {
    const int n = f.k1 - f.k0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < n) {
            acc += f.c[j] * (ya - f.v[j]);
            dg += f.c[j];
        }
    for (int k = f.k0 + 4; k < f.k1; ++k) {
        const A c = A(M.sp_coef[k]);
        acc += c * (ya - val(static_cast<int64_t>(M.sp_other[k])));
        dg += c;
    }
}

// Fill the tile with val(dof) for the brick's 512 cells and the 6 x 64 face halo cells
// (zero where there is no neighbour).  Every lane of the 512-lane workgroup must call it.
template <class A, class Fn>
__device__ __forceinline__ void load_tile(const MeshView& M, int b, A* tile, Fn val) {
    const TileFetch<A> f = fetch_tile<A>(M, b, val);
    const int q = threadIdx.x;
    tile[tile_index(q & 7, (q >> 3) & 7, q >> 6)] = f.own;
    if (q < 384) tile[f.hidx] = f.halo;
    __syncthreads();
}

// The brick's cell size and outlet flag (x-max brick of a case with an outlet), read with the
// brick's first loads: after the tile barrier (a memory fence) a read of the record would be
// issued again and waited for.
struct BrickScalars {
    double h;
    bool outlet;
};
__device__ __forceinline__ BrickScalars brick_scalars(const MeshView& M, int b) {
    return BrickScalars{M.meta[b].h, M.outlet && (M.meta[b].flags & kXMAX) != 0};
}

// structured part of (A y)_a from the tile: faces in the order +x, -x, +y, -y, +z, -z,
// then the x-max outlet (2h y_a)
template <class A>
__device__ __forceinline__ A tile_apply(const A* tile, int q, uint8_t m, A h, bool outlet, A ya, A& dg) {
    const int c = tile_index(q & 7, (q >> 3) & 7, q >> 6);
    const int step[3] = {1, 10, 100};
    A acc = A(0), d = A(0);
    for (int axis = 0; axis < 3; ++axis) {
        if (m & (kPX << axis)) { acc += h * (ya - tile[c + step[axis]]); d += h; }
        if (m & (kMX << axis)) { acc += h * (ya - tile[c - step[axis]]); d += h; }
    }
    if (outlet && (q & 7) == 7) { acc += A(2) * h * ya; d += A(2) * h; }
    dg = d;
    return acc;
}

// compact-edge part of (A y)_a for a special cell: its incidences in groups of four, the
// other cells and conductances of a group loaded together, then their values, summed in
// incidence order
template <class A, class Fn>
__device__ __forceinline__ void compact_apply(const MeshView& M, int sp, Fn val, A ya, A& acc, A& dg) {
    const int k0 = M.special_inc_off[sp], k1 = M.special_inc_off[sp + 1];
    for (int k = k0; k < k1; k += 4) {
        int o[4];
        A c[4], v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool ok = k + j < k1;
            o[j] = ok ? M.sp_other[k + j] : 0;
            c[j] = ok ? A(M.sp_coef[k + j]) : A(0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = k + j < k1 ? val(static_cast<int64_t>(o[j])) : A(0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (k + j < k1) {
                acc += c[j] * (ya - v[j]);
                dg += c[j];
            }
    }
}

__device__ __forceinline__ int axis_coord(int q, int axis) { return (q >> (3 * axis)) & 7; }

// p[i] where ok, else dflt, read without a branch (from p[safe], a valid index, and
// discarded), so that the loads of a phase stay in one basic block and are all in flight
// before the first use
template <class T>
__device__ __forceinline__ T load_if(const T* __restrict__ p, bool ok, int64_t i, int64_t safe, T dflt) {
    const T v = p[ok ? i : safe];
    return ok ? v : dflt;
}

// XCD-aware brick order: workgroups are dispatched round-robin over the 8 XCDs of the
// MI355X, so block g runs on XCD g % 8.  The bricks (Morton order) are dealt to the XCDs in
// chunks of 2^lc consecutive bricks, so a brick and most of its neighbours (halo reads)
// share one L2 while bricks of one region (the cut cells of a body) spread over all XCDs;
// lc < 0 gives every XCD one contiguous range.  Bricks past the last whole round of 8
// chunks are split into per-XCD ranges.
__device__ __forceinline__ int brick_of(int lc, int g, int G) {
    const int full = lc >= 0 ? (G >> (lc + 3)) << (lc + 3) : 0;
    if (g < full) {
        const int x = g & 7, idx = g >> 3;
        return ((((idx >> lc) << 3) + x) << lc) + (idx & ((1 << lc) - 1));
    }
    const int t = g - full, T = G - full;
    const int x = t & 7, idx = t >> 3, per = T >> 3, rem = T & 7;
    return full + x * per + (x < rem ? x : rem) + idx;
}

__device__ __forceinline__ int brick_of_block(int lc = -1) {
    return brick_of(lc, static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
}

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

// Rank of a special cell inside its brick: the brick's first rank, the specials of the
// earlier waves (brick record) and a ballot of this wave.  No barrier; all lanes of the wave
// must call it.
__device__ __forceinline__ int special_rank(const MeshView& M, int b, int q, bool sp) {
    const int wave = __builtin_amdgcn_readfirstlane(q >> 6), lane = q & 63;
    // wave_pre read as the 32-bit word holding it (a scalar load)
    const uint32_t pair = reinterpret_cast<const uint32_t*>(M.meta[b].wave_pre)[wave >> 1];
    const int base = M.meta[b].special_off + static_cast<int>((pair >> (16 * (wave & 1))) & 0xffffu);
    const unsigned long long bal = __ballot(sp);
    if (!sp) return -1;
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
// plain store at its brick's slot (partials[array * nparts + brick]); a finalize kernel
// launched next on the same stream, one workgroup per instance, sums the slots of that
// instance's bricks in brick order, so a batched instance reduces exactly as it does alone.
// (A "last workgroup finalises" scheme needs an agent-scope release fence per workgroup,
// which writes back L2 and cost ~0.4 ms per launch at 8k workgroups on gfx950.)
__device__ __forceinline__ void store_partial(double* partials, int slot, double value) {
    if (threadIdx.x == 0) partials[slot] = value;
}

// Sum (or max) of n partials in fixed order; result valid in thread 0.  Thread t folds
// slots t, t + BLOCK, t + 2 BLOCK, ... in that order; the slots are loaded 16 at a time
// before any is folded, so a batch costs one memory latency, not sixteen.
template <int BLOCK>
__device__ __forceinline__ double gather_partials(const double* partials, int n, bool take_max) {
    constexpr int U = 16;
    double acc = 0.0;
    for (int base = threadIdx.x; base < n; base += U * BLOCK) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = base + u * BLOCK < n ? partials[base + u * BLOCK] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * BLOCK < n) acc = take_max ? fmax(acc, v[u]) : acc + v[u];
    }
    return take_max ? block_max<BLOCK>(acc) : block_sum<BLOCK>(acc);
}

// The same for NA arrays at once (a, a + stride, ...), U slots of each per round: one memory
// latency per round for all of them (a brick-level grid of 40 k slots is one round).  Thread t
// folds slots t, t + BLOCK, ... of each array in that order, as gather_partials does, so the
// sums are the same bit for bit.
template <int BLOCK, int NA, int U>
__device__ __forceinline__ void gather_partials_n(const double* a, int stride, int n, bool take_max, double* out) {
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    for (int base = threadIdx.x; base < n; base += U * BLOCK) {
        double v[NA][U];
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u) v[k][u] = base + u * BLOCK < n ? a[k * stride + base + u * BLOCK] : 0.0;
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (base + u * BLOCK < n) acc[k] = take_max ? fmax(acc[k], v[k][u]) : acc[k] + v[k][u];
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) out[k] = take_max ? block_max<BLOCK>(acc[k]) : block_sum<BLOCK>(acc[k]);
}

}  // namespace pcfd
