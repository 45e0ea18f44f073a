// Learned preconditioner: a 4-layer 3x3x3 convolutional network (16 channels, BF16 MFMA,
// FP32 accumulate) applied per 8^3 brick, one weight set per AMR level (cell size h_b),
// correcting the geometric V-cycle inside flexible PCG:
//     z  = M_mg(r)                      (V-cycle)
//     r' = r - A z                      (what the V-cycle left; k_nn_resid)
//     z += (s_b / h_b) N_level(r' / s_b, geometry)     on the network's bricks
// s_b is the brick's RMS of r' over active cells (the correction is positively homogeneous
// in r'), and the composite operator scales like h.  The network is trained on the device to
// minimise the energy norm of the error left after the correction,
//     L(dz) = dz^T A dz - 2 dz^T r'  (= ||A^-1 r' - dz||_A^2 - ||A^-1 r'||_A^2),
// over post-V-cycle residuals r' recorded from the solver's own MGPCG iterations; the
// gradient 2 (A dz - r') needs one operator apply, no exact solution.  Plain MGPCG is the
// fallback.
//
// Brick set: the network's bricks, position p -> brick list[p] (no list: p = brick), with a
// neighbour table nbtab[27 p + d] = position of the same-level brick at offset d (reached face
// by face in x, y, z order) or -1 (no such brick, or not in the set).  r' lives in a compact
// FP32 store of 512 values per position.
//
// Input features per tile position (16 channels; the brick's 8^3 cells and the one-cell halo
// from its 26 neighbours): r'/s_b, 1 (active), special-cell flag, the parities of the cell's
// i, j, k in its brick as +-1, the fraction of its six structured faces that are open, the six
// structured-face bits, 0, 0, 0; inactive positions and positions outside the set carry
// zeros.  Layers: 16 -> 16 channels; layer 0 reads the halo, the hidden layers pad with zeros;
// ReLU after layers 0..2, outputs of inactive cells forced to zero; the last layer's channel 0
// is the network output.  Weights [set][layer][tap][cout][cin] BF16 (and a tap-flipped,
// cin/cout-transposed copy for the input gradient), biases [set][layer][cout] FP32, FP32
// master weights for Adam.  All network kernels run one 512-lane workgroup (8 waves) per
// brick, thread q owning cell q.  An activation tile is 10^3 positions x 16 channels x 2
// bytes = 32 KB; with one layer's weights (13.5 KB) two bricks share a CU.
//
// MFMA v_mfma_f32_16x16x32_bf16 lane layout (checked on gfx950): A[m][k] lane l holds
// m = l%16, k = 8(l/16)+j; B[k][n]: n = l%16, same k; D[m][n]: m = 4(l/16)+i, n = l%16.
#pragma once

#include "device.hpp"
#include "kernels_solve.hpp"

namespace pcfd {

constexpr int kNNLayers = 4;
constexpr int kNNC = 16;
constexpr int kNNTaps = 27;
constexpr int kNNSets = 4;                         // weight sets: AMR levels 0, 1, 2, >= 3
constexpr int kNNWLayer = kNNTaps * kNNC * kNNC;   // weights per layer
constexpr int kNNWSet = kNNLayers * kNNWLayer;     // weights per set
constexpr int kNNW = kNNSets * kNNWSet;            // all weights
constexpr int kNNBSet = kNNLayers * kNNC;          // biases per set
constexpr int kNNB = kNNSets * kNNBSet;            // all biases
constexpr int kNNTile = kTileN * kNNC;             // bf16 values in one activation tile
constexpr int kNNStore = 3 * 512 * kNNC;           // per-brick training scratch (inputs of layers 1..3)
constexpr int kNNMT = 4;                           // 16-cell M-tiles per wave (32 tiles, 8 waves)
constexpr int kNNKSteps = (kNNTaps * kNNC + 31) / 32;  // 32-wide K steps of one convolution (two taps each)
constexpr int kNNWUnits = kNNWLayer / 8;           // 16-byte units of one layer's weights

typedef __bf16 nn_bf16x8 __attribute__((ext_vector_type(8)));
typedef float nn_floatx4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t nn_f2bf(float x) {
    unsigned u;
    __builtin_memcpy(&u, &x, 4);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float nn_bf2f(uint16_t b) { return __uint_as_float(static_cast<unsigned>(b) << 16); }

// tap t = (dz+1)*9 + (dy+1)*3 + (dx+1); offset in the 10^3 tile; offset(26 - t) = -offset(t)
__device__ __forceinline__ int nn_tap_offset(int t) { return (t / 9 - 1) * 100 + ((t / 3) % 3 - 1) * 10 + (t % 3 - 1); }
__device__ __forceinline__ int nn_pos(int cell) { return tile_index(cell & 7, (cell >> 3) & 7, cell >> 6); }
__device__ __forceinline__ bool nn_halo_pos(int p) {
    const int ti = p % 10, tj = (p / 10) % 10, tk = p / 100;
    return ti == 0 || ti == 9 || tj == 0 || tj == 9 || tk == 0 || tk == 9;
}

struct NNView {
    const uint16_t* w;    // [set][L][27][cout][cin]
    const uint16_t* wt;   // [set][L][27][cin][cout], taps flipped: wt[s][l][t][i][o] = w[s][l][26-t][o][i]
    const float* bias;    // [set][L][16]
};

// the network's brick set
struct NNSetView {
    const int* list;      // position -> brick (null: position = brick)
    const int* nbtab;     // 27 per position: neighbour's position or -1
    const uint8_t* masks; // cell masks per position (512 each; the mesh's own masks when position = brick)
};

__device__ __forceinline__ int nn_set(const MeshView& M, int b) {
    const int l = M.meta[b].level;
    return l < kNNSets ? l : kNNSets - 1;
}

// LDS of the network kernels (one brick): two activation tiles, one layer's weights, flags.
// The last layer's FP32 output (512 values) is written over T0.
struct NNShared {
    uint16_t T0[kNNTile];
    uint16_t T1[kNNTile];
    uint16_t WL[kNNWLayer];
    float red[16];
    uint8_t act[512];
};

// One 3x3x3 convolution 16 -> 16 channels over the brick: 32 M-tiles of 16 cells, 4 per wave.
// K = 27 taps x 16 channels, two taps per 32-wide MFMA step (14 steps, the last half zero);
// N = 16 output channels.  Output to a tile (BF16) or, when out_tile == nullptr, channel 0
// into yout (FP32, 512).  wl: [27][cout][cin] in LDS.
__device__ __forceinline__ void nn_conv(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wl,
                                        const float* bias, bool relu, const uint8_t* act,
                                        uint16_t* __restrict__ out_tile, float* __restrict__ yout) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, n = lane & 15, half = (g & 1) * 8;
    nn_floatx4 acc[kNNMT];
    int pos[kNNMT];
    for (int mt = 0; mt < kNNMT; ++mt) {
        acc[mt] = nn_floatx4{0.f, 0.f, 0.f, 0.f};
        pos[mt] = nn_pos((wave * kNNMT + mt) * 16 + n);
    }
#pragma unroll 2
    for (int s = 0; s < kNNKSteps; ++s) {
        const int t = 2 * s + (g >> 1);
        const bool valid = t < kNNTaps;
        nn_bf16x8 bfrag = {};
        int off = 0;
        if (valid) {
            bfrag = *reinterpret_cast<const nn_bf16x8*>(wl + (t * kNNC + n) * kNNC + half);
            off = nn_tap_offset(t);
        }
        for (int mt = 0; mt < kNNMT; ++mt) {
            nn_bf16x8 afrag = {};
            if (valid) afrag = *reinterpret_cast<const nn_bf16x8*>(in + (pos[mt] + off) * kNNC + half);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag, acc[mt], 0, 0, 0);
        }
    }
    const float bn = bias ? bias[n] : 0.f;
    for (int mt = 0; mt < kNNMT; ++mt)
        for (int i = 0; i < 4; ++i) {
            const int cell = (wave * kNNMT + mt) * 16 + 4 * g + i;
            float v = acc[mt][i] + bn;
            if (relu) v = fmaxf(v, 0.f);
            if (!act[cell]) v = 0.f;
            if (out_tile) out_tile[nn_pos(cell) * kNNC + n] = nn_f2bf(v);
            else if (n == 0) yout[cell] = v;
        }
}

// Weight gradient of one layer, dW[t][cout][cin] += sum_cells g[cell][cout] a[cell+off(t)][cin]:
// one tap per MFMA (M = cout, N = cin, K = 512 cells in 16 steps); wave w owns taps w, w+8,
// w+16, w+24.  Bias gradient db[cout] += sum_cells g.
__device__ __forceinline__ void nn_weight_grad(const uint16_t* __restrict__ gt, const uint16_t* __restrict__ at,
                                               float* __restrict__ gw, float* __restrict__ gb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, n = lane & 15;
    for (int t = wave; t < kNNTaps; t += 8) {
        const int off = nn_tap_offset(t);
        nn_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < 16; ++s) {
            nn_bf16x8 afrag, bfrag;
            for (int j = 0; j < 8; ++j) {
                const int p = nn_pos(32 * s + 8 * g + j);
                afrag[j] = __builtin_bit_cast(__bf16, gt[p * kNNC + n]);
                bfrag[j] = __builtin_bit_cast(__bf16, at[(p + off) * kNNC + n]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag, acc, 0, 0, 0);
        }
        for (int i = 0; i < 4; ++i) atomicAdd(gw + (t * kNNC + 4 * g + i) * kNNC + n, acc[i]);
    }
    if (threadIdx.x < kNNC) {
        float s = 0.f;
        for (int cell = 0; cell < 512; ++cell) s += nn_bf2f(gt[nn_pos(cell) * kNNC + threadIdx.x]);
        atomicAdd(gb + threadIdx.x, s);
    }
}

__device__ __forceinline__ void nn_zero_tile(uint16_t* tile) {
    for (int i = threadIdx.x; i < kNNTile / 8; i += blockDim.x) reinterpret_cast<uint4*>(tile)[i] = uint4{0u, 0u, 0u, 0u};
}
__device__ __forceinline__ void nn_zero_halo(uint16_t* tile) {
    for (int p = threadIdx.x; p < kTileN; p += blockDim.x)
        if (nn_halo_pos(p)) {
            reinterpret_cast<uint4*>(tile)[2 * p] = uint4{0u, 0u, 0u, 0u};
            reinterpret_cast<uint4*>(tile)[2 * p + 1] = uint4{0u, 0u, 0u, 0u};
        }
}

// One layer's weights (kNNWUnits units of 16 bytes) through registers into LDS: fetch issues
// the global load, put stores it, so that the next layer's load can be in flight during the
// current convolution.
struct NNWRegs { uint4 v[2]; };
__device__ __forceinline__ void nn_fetch_weights(const uint16_t* __restrict__ src, NNWRegs& R) {
    const int t = threadIdx.x;
    if (t < kNNWUnits) R.v[0] = reinterpret_cast<const uint4*>(src)[t];
    if (t + 512 < kNNWUnits) R.v[1] = reinterpret_cast<const uint4*>(src)[t + 512];
}
__device__ __forceinline__ void nn_put_weights(const NNWRegs& R, uint16_t* dst) {
    const int t = threadIdx.x;
    if (t < kNNWUnits) reinterpret_cast<uint4*>(dst)[t] = R.v[0];
    if (t + 512 < kNNWUnits) reinterpret_cast<uint4*>(dst)[t + 512] = R.v[1];
}
__device__ __forceinline__ void nn_load_weights(const uint16_t* __restrict__ src, uint16_t* dst) {
    NNWRegs R;
    nn_fetch_weights(src, R);
    nn_put_weights(R, dst);
}

// Brick scale s_b (RMS of v over active cells; 0 when that is zero) and the active flags of
// the brick's cells into L.act.  v is cell q's input value, m its mask.  Block-collective.
__device__ __forceinline__ float nn_scale(float v, uint8_t m, NNShared& L) {
    const int q = threadIdx.x;
    const bool act = (m & kActive) != 0;
    L.act[q] = act ? 1 : 0;
    float sq = act ? v * v : 0.f, cnt = act ? 1.f : 0.f;
    for (int o = 32; o > 0; o >>= 1) { sq += __shfl_xor(sq, o, 64); cnt += __shfl_xor(cnt, o, 64); }
    if ((q & 63) == 0) { L.red[2 * (q >> 6)] = sq; L.red[2 * (q >> 6) + 1] = cnt; }
    __syncthreads();
    float S = 0.f, N = 0.f;
    for (int w = 0; w < 8; ++w) { S += L.red[2 * w]; N += L.red[2 * w + 1]; }
    return N > 0.f ? sqrtf(S / N) : 0.f;
}

// Features of every position of the 10^3 tile of the brick at position `pos` (its cells and
// the halo from the neighbours in the set) from the residual store, scaled by inv = 1/s_b.
// Writes every position (zeros where there is no input).
__device__ __forceinline__ void nn_stage_features(const NNSetView& S, int pos,
                                                  const float* __restrict__ store, float inv, uint16_t* tile) {
    constexpr uint32_t one = 0x3f80, minus_one = 0xbf80;
    for (int p = threadIdx.x; p < kTileN; p += blockDim.x) {
        const int ti = p % 10, tj = (p / 10) % 10, tk = p / 100;
        const int d = (ti == 0 ? 0 : ti == 9 ? 2 : 1) + 3 * (tj == 0 ? 0 : tj == 9 ? 2 : 1) + 9 * (tk == 0 ? 0 : tk == 9 ? 2 : 1);
        const int np = S.nbtab[27 * pos + d];
        uint4 f0 = {0u, 0u, 0u, 0u}, f1 = {0u, 0u, 0u, 0u};
        if (np >= 0) {
            const int ci = (ti + 7) & 7, cj = (tj + 7) & 7, ck = (tk + 7) & 7;
            const int64_t c = static_cast<int64_t>(np) * 512 + ci + 8 * cj + 64 * ck;
            const uint32_t m = S.masks[c];
            const float v = store[c];
            if (m & kActive) {
                auto bit = [&](int k) { return ((m >> k) & 1u) ? one : 0u; };
                f0.x = nn_f2bf(v * inv) | (one << 16);
                f0.y = ((m & kSpecial) ? one : 0u) | ((ci & 1 ? one : minus_one) << 16);
                f0.z = (cj & 1 ? one : minus_one) | ((ck & 1 ? one : minus_one) << 16);
                f0.w = nn_f2bf(static_cast<float>(__popc(m & 63u)) / 6.f) | (bit(0) << 16);
                f1.x = bit(1) | (bit(2) << 16);
                f1.y = bit(3) | (bit(4) << 16);
                f1.z = bit(5);
            }
        }
        reinterpret_cast<uint4*>(tile)[2 * p] = f0;
        reinterpret_cast<uint4*>(tile)[2 * p + 1] = f1;
    }
}

// Layers 0..3 of weight set `set` from the feature tile T0 (T1 zeroed) into the FP32 output
// over T0 (returned).  When `store` is set, the inputs of layers 1..3 are written there
// (training).  T0's halo is cleared once layer 0 has read it (hidden layers pad with zeros).
__device__ __forceinline__ const float* nn_layers(const NNView& W, int set, NNShared& L, uint16_t* store,
                                                  const NNWRegs* first = nullptr) {
    const uint16_t* wset = W.w + static_cast<size_t>(set) * kNNWSet;
    uint16_t* in = L.T0;
    uint16_t* out = L.T1;
    NNWRegs R;
    if (first) R = *first;
    else nn_fetch_weights(wset, R);
    for (int l = 0; l < kNNLayers; ++l) {
        nn_put_weights(R, L.WL);
        if (l + 1 < kNNLayers) nn_fetch_weights(wset + static_cast<size_t>(l + 1) * kNNWLayer, R);
        __syncthreads();
        if (store && l > 0) {
            const uint4* src = reinterpret_cast<const uint4*>(in) + 2 * nn_pos(threadIdx.x);
            uint4* dst = reinterpret_cast<uint4*>(store + (static_cast<size_t>(l - 1) * 512 + threadIdx.x) * kNNC);
            dst[0] = src[0];
            dst[1] = src[1];
        }
        if (l == 1) nn_zero_halo(out);   // T0: layer 0 has read its halo; layer 1 writes its interior
        const bool last = l == kNNLayers - 1;
        nn_conv(in, L.WL, W.bias + set * kNNBSet + l * kNNC, !last, L.act, last ? nullptr : out,
                last ? reinterpret_cast<float*>(out) : nullptr);
        __syncthreads();
        uint16_t* t = in; in = out; out = t;
    }
    return reinterpret_cast<const float*>(in);  // the last layer wrote over T0
}

// brick and list position of this workgroup: with a list, block p handles list[p];
// without one, the XCD-aware brick order (position = brick)
struct NNBrick { int b, pos; };
__device__ __forceinline__ NNBrick nn_brick(const int* list) {
    if (list) return NNBrick{list[blockIdx.x], static_cast<int>(blockIdx.x)};
    const int b = brick_of_block();
    return NNBrick{b, b};
}

// r' = r - A z on the network's bricks into the FP32 store (0 on inactive cells), after the
// V-cycle's last kernel wrote z.  S / off may be null (no done / switched-off test).
template <class ZT>
__global__ __launch_bounds__(512) void k_nn_resid(MeshView M, const int* __restrict__ list, const double* __restrict__ r,
                                                  const ZT* __restrict__ z, float* __restrict__ store,
                                                  const DevScalars* S, const int* __restrict__ off) {
    __shared__ double ztile[kTileN];
    const NNBrick B = nn_brick(list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    auto zv = [&](int64_t c) -> double { return double(z[c]); };
    const TileFetch<double> tf = fetch_tile<double>(M, B.b, zv);
    const double ra = r[a];
    const bool skip = (S && S->done) || (off && *off);
    const int sp = special_rank(M, B.b, q, (m & kSpecial) != 0);
    const CompactFetch<double> cf = fetch_compact<double>(M, sp, zv);
    if (skip) return;
    stage_tile<double>(ztile, tf);
    __syncthreads();
    double rp = 0.0;
    if (m & kActive) {
        double dg;
        rp = ra - row0f<double>(M, ztile, cf, zv, brick_scalars(M, B.b), q, m, tf.own, dg);
    }
    store[static_cast<int64_t>(B.pos) * 512 + q] = static_cast<float>(rp);
}

// Inference inside PCG, after k_nn_resid: the network on the brick from the residual store,
// dz = (s_b / h_b) y into dz[pos] (applied to z by k_nn_add, so that no block reads a z that
// another block has changed), and the brick's partials increased by r.dz and r_prev.dz.
__global__ __launch_bounds__(512, 4) void k_nn_apply(MeshView M, NNView W, NNSetView NS, const float* __restrict__ store,
                                                  const double* __restrict__ r, double* __restrict__ dz,
                                                  const double* __restrict__ rprev, double* partials,
                                                  int nparts, const DevScalars* S, const int* __restrict__ off) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    const NNBrick B = nn_brick(NS.list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    const float v = store[static_cast<int64_t>(B.pos) * 512 + q];
    const double ra = r[a];
    const double rpa = rprev ? rprev[a] : 0.0;
    const int set = nn_set(M, B.b);
    NNWRegs W0;
    nn_fetch_weights(W.w + static_cast<size_t>(set) * kNNWSet, W0);
    if (S->done || *off) return;
    const float s = nn_scale(v, m, L);
    double d = 0.0;
    if (s > 0.f) {  // block-uniform
        nn_zero_tile(L.T1);
        nn_stage_features(NS, B.pos, store, 1.f / s, L.T0);
        const float* y = nn_layers(W, set, L, nullptr, &W0);
        if (m & kActive) d = double(s) / level_h(M, B.b) * double(y[q]);
    }
    dz[static_cast<int64_t>(B.pos) * 512 + q] = d;
    const double t1 = block_sum<512>(ra * d);
    const double t2 = rprev ? block_sum<512>(rpa * d) : 0.0;
    if (threadIdx.x == 0) {
        partials[B.b] += t1;
        partials[nparts + B.b] += t2;
    }
}

// z += dz on the network's bricks
__global__ __launch_bounds__(512) void k_nn_add(const int* __restrict__ list, const double* __restrict__ dz,
                                                double* __restrict__ z, const DevScalars* S, const int* __restrict__ off) {
    if (S && (S->done || *off)) return;
    const NNBrick B = nn_brick(list);
    z[static_cast<int64_t>(B.b) * 512 + threadIdx.x] += dz[static_cast<int64_t>(B.pos) * 512 + threadIdx.x];
}

// FP64 vector -> FP32 store of the set's bricks (position-major)
__global__ __launch_bounds__(512) void k_nn_gather(const int* __restrict__ list, const double* __restrict__ v,
                                                   float* __restrict__ store) {
    const int p = blockIdx.x;
    const int b = list ? list[p] : p;
    store[static_cast<int64_t>(p) * 512 + threadIdx.x] = static_cast<float>(v[static_cast<int64_t>(b) * 512 + threadIdx.x]);
}

// Raw network output of every brick of the set (block = position): y[pos*512+q] (0 on
// inactive cells and zero-input bricks) and s[pos], input from the residual store.
__global__ __launch_bounds__(512, 4) void k_nn_forward_raw(MeshView M, NNView W, NNSetView NS, const float* __restrict__ store,
                                                        float* __restrict__ yout, float* __restrict__ sout) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    const int pos = blockIdx.x, q = threadIdx.x;
    const int b = NS.list ? NS.list[pos] : pos;
    const uint8_t m = M.mask[static_cast<int64_t>(b) * 512 + q];
    const float s = nn_scale(store[static_cast<int64_t>(pos) * 512 + q], m, L);
    float y = 0.f;
    if (s > 0.f) {
        nn_zero_tile(L.T1);
        nn_stage_features(NS, pos, store, 1.f / s, L.T0);
        y = nn_layers(W, nn_set(M, b), L, nullptr)[q];
    }
    yout[static_cast<int64_t>(pos) * 512 + q] = y;
    if (q == 0) sout[pos] = s;
}

// ---- training (one recorded post-V-cycle residual per pass, in the FP32 store) -------------
// forward over the set: layer inputs to scratch[pos], dz = (s_b/h_b) y into the full-length
// vector dfull (its other entries stay zero), s_b into sb[pos]
__global__ __launch_bounds__(512, 4) void k_nn_train_fwd(MeshView M, NNView W, NNSetView NS, const float* __restrict__ store,
                                                      uint16_t* __restrict__ scratch, double* __restrict__ dfull,
                                                      float* __restrict__ sb) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    const NNBrick B = nn_brick(NS.list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    const float s = nn_scale(store[static_cast<int64_t>(B.pos) * 512 + q], m, L);
    double d = 0.0;
    if (s > 0.f) {
        nn_zero_tile(L.T1);
        nn_stage_features(NS, B.pos, store, 1.f / s, L.T0);
        const float* y = nn_layers(W, nn_set(M, B.b), L, scratch + static_cast<size_t>(blockIdx.x) * kNNStore);
        if (m & kActive) d = double(s) / level_h(M, B.b) * double(y[q]);
    }
    dfull[a] = d;
    if (q == 0) sb[blockIdx.x] = s;
}

// energy loss and its gradient on the set's bricks:
// gy = 2 (A dz - r') inv_norm,  loss += inv_norm * sum dz (A dz - 2 r')
__global__ __launch_bounds__(512) void k_nn_energy(MeshView M, const int* __restrict__ list, const double* __restrict__ dfull,
                                                   const float* __restrict__ store, double* __restrict__ gy,
                                                   double* __restrict__ loss, double inv_norm) {
    __shared__ double tile[kTileN];
    const NNBrick B = nn_brick(list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int sp = special_rank(M, B.b, q, (m & kSpecial) != 0);
    auto dv = [&](int64_t c) -> double { return dfull[c]; };
    load_tile<double>(M, B.b, tile, dv);
    double g = 0.0, l = 0.0;
    if (m & kActive) {
        double dg;
        const double rp = store[static_cast<int64_t>(B.pos) * 512 + q];
        const double da = tile[nn_pos(q)];
        const double Ad = row0<double>(M, tile, dv, brick_scalars(M, B.b), q, m, sp, da, dg);
        g = 2.0 * (Ad - rp) * inv_norm;
        l = da * (Ad - 2.0 * rp) * inv_norm;
    }
    gy[static_cast<int64_t>(blockIdx.x) * 512 + q] = g;
    const double t = block_sum<512>(l);
    if (q == 0) atomicAdd(loss, t);
}

// backward over the set: output gradient (s_b/h_b) gy (BF16), MFMA weight gradients and
// transposed-convolution input gradients through the ReLU masks, accumulated into
// grad = [dW (kNNW) | db (kNNB)] of the brick's weight set with FP32 atomics
__global__ __launch_bounds__(512) void k_nn_train_bwd(MeshView M, NNView W, NNSetView NS, const float* __restrict__ store,
                                                      const double* __restrict__ gy, const float* __restrict__ sb,
                                                      const uint16_t* __restrict__ scratch, float* __restrict__ grad) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    __shared__ __attribute__((aligned(16))) uint16_t T2[kNNTile];
    const NNBrick B = nn_brick(NS.list);
    const float s = sb[blockIdx.x];
    if (!(s > 0.f)) return;  // block-uniform
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    const int set = nn_set(M, B.b);
    const uint16_t* saved = scratch + static_cast<size_t>(blockIdx.x) * kNNStore;
    nn_scale(store[static_cast<int64_t>(B.pos) * 512 + q], m, L);   // active flags
    nn_zero_tile(L.T0);
    nn_zero_tile(T2);
    // layer-0 input (features with the neighbour halo) into T1
    nn_stage_features(NS, B.pos, store, 1.f / s, L.T1);
    __syncthreads();
    // output gradient into channel 0 of T0
    if (m & kActive) {
        const float gq = static_cast<float>(double(s) / level_h(M, B.b) * gy[static_cast<int64_t>(blockIdx.x) * 512 + q]);
        L.T0[nn_pos(q) * kNNC] = nn_f2bf(gq);
    }
    __syncthreads();
    // Three tiles: gcur (gradient w.r.t. layer l's pre-activation output), gnext and T1 (the
    // features, input of layer 0).  The input of layer l > 0 is staged from scratch into
    // gnext, its ReLU pattern is kept in registers, and gnext is then overwritten by the
    // transposed convolution.  The halos of gcur and gnext stay zero.
    uint16_t* gcur = L.T0;
    uint16_t* gnext = T2;
    float* gw = grad + static_cast<size_t>(set) * kNNWSet;
    float* gb = grad + kNNW + set * kNNBSet;
    for (int l = kNNLayers - 1; l >= 0; --l) {
        uint16_t* ain = L.T1;
        if (l > 0) {
            ain = gnext;
            const uint4* src = reinterpret_cast<const uint4*>(saved + (static_cast<size_t>(l - 1) * 512 + q) * kNNC);
            uint4* dst = reinterpret_cast<uint4*>(ain) + 2 * nn_pos(q);
            dst[0] = src[0];
            dst[1] = src[1];
        }
        __syncthreads();
        nn_weight_grad(gcur, ain, gw + static_cast<size_t>(l) * kNNWLayer, gb + l * kNNC);
        if (l > 0) {
            // ReLU pattern of layer l-1's output (this layer's input) at cell q
            uint32_t pmask = 0;
            for (int c = 0; c < kNNC; ++c)
                if (nn_bf2f(ain[nn_pos(q) * kNNC + c]) > 0.f) pmask |= 1u << c;
            nn_load_weights(W.wt + static_cast<size_t>(set) * kNNWSet + static_cast<size_t>(l) * kNNWLayer, L.WL);
            __syncthreads();
            nn_conv(gcur, L.WL, nullptr, false, L.act, gnext, nullptr);
            __syncthreads();
            for (int c = 0; c < kNNC; ++c)
                if (!((pmask >> c) & 1u)) gnext[nn_pos(q) * kNNC + c] = 0;
            __syncthreads();
            uint16_t* t = gcur; gcur = gnext; gnext = t;
        }
        __syncthreads();
    }
}

// BF16 copies (both layouts) of the FP32 master weights.
__global__ void k_nn_pack(const float* __restrict__ w32, uint16_t* __restrict__ w, uint16_t* __restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNNW) return;
    const int lset = i / kNNWLayer, rem = i % kNNWLayer;   // lset = set * kNNLayers + layer
    const int t = rem / (kNNC * kNNC), cout = (rem / kNNC) % kNNC, cin = rem % kNNC;
    const uint16_t bits = nn_f2bf(w32[i]);
    w[i] = bits;
    wt[static_cast<size_t>(lset) * kNNWLayer + ((26 - t) * kNNC + cin) * kNNC + cout] = bits;
}

// Adam on p = [weights | biases] (FP32), gradient cleared after use.
__global__ void k_nn_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, float* __restrict__ g,
                          float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNNW + kNNB) return;
    const float gi = g[i];
    m[i] = beta1 * m[i] + (1.f - beta1) * gi;
    v[i] = beta2 * v[i] + (1.f - beta2) * gi * gi;
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    g[i] = 0.f;
}

// per-position partials of sum_active v^2 / h_b over the set's bricks (normalisation of a
// recorded residual in the FP32 store)
__global__ __launch_bounds__(512) void k_nn_norm(MeshView M, const int* __restrict__ list, const float* __restrict__ store,
                                                 double* partials) {
    const int p = blockIdx.x, q = threadIdx.x;
    const int b = list ? list[p] : p;
    const double x = (M.mask[static_cast<int64_t>(b) * 512 + q] & kActive) ? double(store[static_cast<int64_t>(p) * 512 + q]) : 0.0;
    const double t = block_sum<512>(x * x / level_h(M, b));
    if (q == 0) partials[p] = t;
}

}  // namespace pcfd
