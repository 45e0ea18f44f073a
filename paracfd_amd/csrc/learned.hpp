// Learned preconditioner: a 4-layer 3x3x3 convolutional network (8 channels, BF16 MFMA,
// FP32 accumulate) applied per 8^3 brick (every AMR level, cell size h_b), correcting the
// geometric V-cycle inside flexible PCG:
//     z  = M_mg(r)                      (V-cycle)
//     r' = r - A z                      (what the V-cycle left, computed in the kernel)
//     z += (s_b / h_b) N(r' / s_b, geometry)     on the network's bricks
// s_b is the brick's RMS of r' over active cells (the correction is positively homogeneous
// in r'), and the composite operator scales like h.  The network is trained on the device to
// minimise the energy norm of the error left after the correction,
//     L(dz) = dz^T A dz - 2 dz^T r'  (= ||A^-1 r' - dz||_A^2 - ||A^-1 r'||_A^2),
// over post-V-cycle residuals r' recorded from the solver's own MGPCG iterations; the
// gradient 2 (A dz - r') needs one operator apply, no exact solution.  Plain MGPCG is the
// fallback.
//
// Input features per active cell (8 channels): r'/s_b, 1 (active), special-cell flag, the
// parities of i, j, k as +-1, the fraction of the six structured faces that are open, and 0.
// Inactive cells carry zeros.  Layers: 8 -> 8 channels, zero padding at the brick boundary,
// ReLU after layers 0..2, outputs of inactive cells forced to zero; the last layer's channel
// 0 is the network output.  Weights [layer][tap][cout][cin] BF16 (and a tap-flipped,
// cin/cout-transposed copy for the input gradient), biases FP32, FP32 master weights for
// Adam.  All network kernels run one 512-lane workgroup (8 waves) per brick, thread q owning
// cell q.  An activation tile is 10^3 positions x 8 channels x 2 bytes = 16 KB, so four
// bricks share a CU.
//
// MFMA v_mfma_f32_16x16x32_bf16 lane layout (checked on gfx950): A[m][k] lane l holds
// m = l%16, k = 8(l/16)+j; B[k][n]: n = l%16, same k; D[m][n]: m = 4(l/16)+i, n = l%16.
#pragma once

#include "device.hpp"
#include "kernels_solve.hpp"

namespace pcfd {

constexpr int kNNLayers = 4;
constexpr int kNNC = 8;
constexpr int kNNTaps = 27;
constexpr int kNNWLayer = kNNTaps * kNNC * kNNC;  // weights per layer
constexpr int kNNW = kNNLayers * kNNWLayer;        // all weights
constexpr int kNNB = kNNLayers * kNNC;             // all biases
constexpr int kNNTile = kTileN * kNNC;             // bf16 values in one activation tile
constexpr int kNNStore = 3 * 512 * kNNC;           // per-brick training scratch (inputs of layers 1..3)
constexpr int kNNMT = 4;                           // 16-cell M-tiles per wave (32 tiles, 8 waves)

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

struct NNView {
    const uint16_t* w;    // [L][27][cout][cin]
    const uint16_t* wt;   // [L][27][cin][cout], taps flipped: wt[l][t][i][o] = w[l][26-t][o][i]
    const float* bias;    // [L][8]
};

// LDS of the network kernels (one brick): two activation tiles, one layer's weights, flags
struct NNShared {
    uint16_t T0[kNNTile];
    uint16_t T1[kNNTile];
    uint16_t WL[kNNWLayer];
    float y[512];
    float red[16];
    uint8_t act[512];
};

// One 3x3x3 convolution 8 -> 8 channels over the brick: 32 M-tiles of 16 cells, 4 per wave.
// K = 27 taps x 8 channels, four taps per 32-wide MFMA step (7 steps, the last tap slot
// zero); N = 16 with output channels 8..15 zero.  Output to a tile (BF16) or, when
// out_tile == nullptr, channel 0 into yout (FP32, 512).  wl: [27][cout][cin] in LDS.
__device__ __forceinline__ void nn_conv(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wl,
                                        const float* bias, bool relu, const uint8_t* act,
                                        uint16_t* __restrict__ out_tile, float* __restrict__ yout) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, n = lane & 15;
    nn_floatx4 acc[kNNMT];
    int pos[kNNMT];
    for (int mt = 0; mt < kNNMT; ++mt) {
        acc[mt] = nn_floatx4{0.f, 0.f, 0.f, 0.f};
        pos[mt] = nn_pos((wave * kNNMT + mt) * 16 + n);
    }
    for (int s = 0; s < 7; ++s) {
        const int t = 4 * s + g;
        const bool valid = t < kNNTaps;
        nn_bf16x8 bfrag = {};
        int off = 0;
        if (valid) {
            if (n < kNNC) bfrag = *reinterpret_cast<const nn_bf16x8*>(wl + (t * kNNC + n) * kNNC);
            off = nn_tap_offset(t);
        }
        for (int mt = 0; mt < kNNMT; ++mt) {
            nn_bf16x8 afrag = {};
            if (valid) afrag = *reinterpret_cast<const nn_bf16x8*>(in + (pos[mt] + off) * kNNC);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag, acc[mt], 0, 0, 0);
        }
    }
    if (n >= kNNC) return;
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

// Weight gradient of one layer, dW[t][cout][cin] += sum_cells g[cell][cout] a[cell+off(t)][cin],
// two taps per MFMA: M = cout (rows 8..15 zero), N = 2 taps x 8 cin, K = 512 cells in 16
// steps; wave w owns tap pairs w, w+8 (14 pairs).  Bias gradient db[cout] += sum_cells g.
__device__ __forceinline__ void nn_weight_grad(const uint16_t* __restrict__ gt, const uint16_t* __restrict__ at,
                                               float* __restrict__ gw, float* __restrict__ gb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, n = lane & 15;
    for (int pr = wave; pr < 14; pr += 8) {
        const int t = 2 * pr + (n >> 3), cin = n & 7;
        const bool valid = t < kNNTaps;
        const int off = valid ? nn_tap_offset(t) : 0;
        nn_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < 16; ++s) {
            nn_bf16x8 afrag = {}, bfrag = {};
            for (int j = 0; j < 8; ++j) {
                const int p = nn_pos(32 * s + 8 * g + j);
                if (n < kNNC) afrag[j] = __builtin_bit_cast(__bf16, gt[p * kNNC + n]);
                if (valid) bfrag[j] = __builtin_bit_cast(__bf16, at[(p + off) * kNNC + cin]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag, acc, 0, 0, 0);
        }
        if (valid)
            for (int i = 0; i < 4; ++i) {
                const int cout = 4 * g + i;
                if (cout < kNNC) atomicAdd(gw + (t * kNNC + cout) * kNNC + cin, acc[i]);
            }
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

// One layer's weights (kNNWLayer / 8 = 216 units of 16 bytes) through registers into LDS:
// fetch issues the global load, put stores it, so that the next layer's load can be in
// flight during the current convolution.
struct NNWRegs { uint4 v; };
__device__ __forceinline__ void nn_fetch_weights(const uint16_t* __restrict__ src, NNWRegs& R) {
    if (threadIdx.x < kNNWLayer / 8) R.v = reinterpret_cast<const uint4*>(src)[threadIdx.x];
}
__device__ __forceinline__ void nn_put_weights(const NNWRegs& R, uint16_t* dst) {
    if (threadIdx.x < kNNWLayer / 8) reinterpret_cast<uint4*>(dst)[threadIdx.x] = R.v;
}
__device__ __forceinline__ void nn_load_weights(const uint16_t* __restrict__ src, uint16_t* dst) {
    NNWRegs R;
    nn_fetch_weights(src, R);
    nn_put_weights(R, dst);
}

// Brick scale s_b (RMS of v over active cells; 0 when that is zero) and, when s_b > 0, the
// features of cell q = threadIdx.x into the zeroed tile.  v is cell q's input value, m its
// mask.  Block-collective (512 lanes).
__device__ __forceinline__ float nn_features(float v, uint8_t m, uint16_t* tile, NNShared& L) {
    const int q = threadIdx.x;
    const bool act = (m & kActive) != 0;
    L.act[q] = act ? 1 : 0;
    float sq = act ? v * v : 0.f, cnt = act ? 1.f : 0.f;
    for (int o = 32; o > 0; o >>= 1) { sq += __shfl_xor(sq, o, 64); cnt += __shfl_xor(cnt, o, 64); }
    if ((q & 63) == 0) { L.red[2 * (q >> 6)] = sq; L.red[2 * (q >> 6) + 1] = cnt; }
    __syncthreads();
    float S = 0.f, N = 0.f;
    for (int w = 0; w < 8; ++w) { S += L.red[2 * w]; N += L.red[2 * w + 1]; }
    const float s = N > 0.f ? sqrtf(S / N) : 0.f;
    if (act && s > 0.f) {
        constexpr uint16_t one = 0x3f80, minus_one = 0xbf80;
        const int open = __popc(static_cast<unsigned>(m) & 63u);
        uint4 f;
        f.x = nn_f2bf(v * (1.f / s)) | (unsigned(one) << 16);
        f.y = ((m & kSpecial) ? one : 0u) | (unsigned((q & 1) ? one : minus_one) << 16);
        f.z = ((q >> 3) & 1 ? one : minus_one) | (unsigned((q >> 6) & 1 ? one : minus_one) << 16);
        f.w = nn_f2bf(static_cast<float>(open) / 6.f);
        reinterpret_cast<uint4*>(tile)[nn_pos(q)] = f;
    }
    __syncthreads();
    return s;
}

// Layers 0..3 from the feature tile T0 (T1 zeroed, halos of both zero) into L.y.  When
// `store` is set, the inputs of layers 1..3 are written there (training).
__device__ __forceinline__ void nn_layers(const NNView& W, NNShared& L, uint16_t* store,
                                          const NNWRegs* first = nullptr) {
    uint16_t* in = L.T0;
    uint16_t* out = L.T1;
    NNWRegs R;
    if (first) R = *first;
    else nn_fetch_weights(W.w, R);
    for (int l = 0; l < kNNLayers; ++l) {
        nn_put_weights(R, L.WL);
        if (l + 1 < kNNLayers) nn_fetch_weights(W.w + static_cast<size_t>(l + 1) * kNNWLayer, R);
        __syncthreads();
        if (store && l > 0)
            *reinterpret_cast<uint4*>(store + (static_cast<size_t>(l - 1) * 512 + threadIdx.x) * kNNC) =
                reinterpret_cast<const uint4*>(in)[nn_pos(threadIdx.x)];
        const bool last = l == kNNLayers - 1;
        nn_conv(in, L.WL, W.bias + l * kNNC, !last, L.act, last ? nullptr : out, last ? L.y : nullptr);
        __syncthreads();
        uint16_t* t = in; in = out; out = t;
    }
}

// brick and list position of this workgroup: with a list, block p handles list[p];
// without one, the XCD-aware brick order (position = brick)
struct NNBrick { int b, pos; };
__device__ __forceinline__ NNBrick nn_brick(const int* list) {
    if (list) return NNBrick{list[blockIdx.x], static_cast<int>(blockIdx.x)};
    const int b = brick_of_block();
    return NNBrick{b, b};
}

// Inference inside PCG, after the V-cycle's last kernel wrote z and the r.z / r_prev.z
// partials (at brick slots): r' = r - A z on the brick, the network, dz = (s_b / h_b) y into
// dz[pos] (applied to z by k_nn_add, so that no block reads a z that another block has
// changed), and the brick's partials increased by r.dz and r_prev.dz.
__global__ __launch_bounds__(512) void k_nn_apply(MeshView M, NNView W, const int* __restrict__ list,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ z, double* __restrict__ dz,
                                                  const double* __restrict__ rprev, double* partials,
                                                  int nparts, const DevScalars* S, const int* __restrict__ off) {
    // every independent load first (brick, mask, z tile and halo, r, compact neighbours, the
    // first layer's weights), then the flags, then the tile barrier
    __shared__ __attribute__((aligned(16))) NNShared L;
    const NNBrick B = nn_brick(list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    double* ztile = reinterpret_cast<double*>(L.T1);   // the z tile lives in T1 until the layers
    auto zv = [&](int64_t c) -> double { return z[c]; };
    const TileFetch<double> tf = fetch_tile<double>(M, B.b, zv);
    const double ra = r[a];
    NNWRegs W0;
    nn_fetch_weights(W.w, W0);
    const bool skip = S->done || *off;
    const int sp = special_rank(M, B.b, q, (m & kSpecial) != 0);
    const CompactFetch<double> cf = fetch_compact<double>(M, sp, zv);
    if (skip) return;
    stage_tile<double>(ztile, tf);
    double rp = 0.0;
    if (m & kActive) {
        double dg;
        rp = ra - row0f<double>(M, ztile, cf, zv, B.b, q, m, tf.own, dg);
    }
    __syncthreads();
    nn_zero_tile(L.T0);
    nn_zero_tile(L.T1);
    __syncthreads();
    const float s = nn_features(static_cast<float>(rp), m, L.T0, L);
    double d = 0.0;
    if (s > 0.f) {  // block-uniform
        nn_layers(W, L, nullptr, &W0);
        if (m & kActive) d = double(s) / level_h(M, B.b) * double(L.y[q]);
    }
    dz[static_cast<int64_t>(B.pos) * 512 + q] = d;
    const double t1 = block_sum<512>(r[a] * d);
    const double t2 = rprev ? block_sum<512>(rprev[a] * d) : 0.0;
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

// Raw network output for every brick (block = brick): y[dof] (0 on inactive cells and
// zero-input bricks) and s[brick], for input v.
__global__ __launch_bounds__(512) void k_nn_forward_raw(MeshView M, NNView W, const double* __restrict__ v,
                                                        float* __restrict__ yout, float* __restrict__ sout) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    nn_zero_tile(L.T0);
    nn_zero_tile(L.T1);
    __syncthreads();
    const uint8_t m = M.mask[a];
    const float s = nn_features(static_cast<float>(v[a]), m, L.T0, L);
    if (s > 0.f) nn_layers(W, L, nullptr);
    yout[a] = s > 0.f ? L.y[q] : 0.f;
    if (q == 0) sout[b] = s;
}

// ---- training (one recorded post-V-cycle residual rp per pass) ----------------------------
// forward over the listed bricks: layer inputs to scratch[pos], dz = (s_b/h_b) y into the
// full-length vector dfull (its other entries stay zero), s_b into sb[pos]
__global__ __launch_bounds__(512) void k_nn_train_fwd(MeshView M, NNView W, const int* __restrict__ list,
                                                      const double* __restrict__ rp, uint16_t* __restrict__ scratch,
                                                      double* __restrict__ dfull, float* __restrict__ sb) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    const NNBrick B = nn_brick(list);
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    nn_zero_tile(L.T0);
    nn_zero_tile(L.T1);
    __syncthreads();
    const uint8_t m = M.mask[a];
    const float s = nn_features(static_cast<float>(rp[a]), m, L.T0, L);
    if (s > 0.f) nn_layers(W, L, scratch + static_cast<size_t>(blockIdx.x) * kNNStore);
    dfull[a] = (s > 0.f && (m & kActive)) ? double(s) / level_h(M, B.b) * double(L.y[q]) : 0.0;
    if (q == 0) sb[blockIdx.x] = s;
}

// energy loss and its gradient on the listed bricks:
// gy = 2 (A dz - r') inv_norm,  loss += inv_norm * sum dz (A dz - 2 r')
__global__ __launch_bounds__(512) void k_nn_energy(MeshView M, const int* __restrict__ list, const double* __restrict__ dfull,
                                                   const double* __restrict__ rp, double* __restrict__ gy,
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
        const double da = tile[nn_pos(q)];
        const double Ad = row0<double>(M, tile, dv, B.b, q, m, sp, da, dg);
        g = 2.0 * (Ad - rp[a]) * inv_norm;
        l = da * (Ad - 2.0 * rp[a]) * inv_norm;
    }
    gy[static_cast<int64_t>(blockIdx.x) * 512 + q] = g;
    const double t = block_sum<512>(l);
    if (q == 0) atomicAdd(loss, t);
}

// backward over the listed bricks: output gradient (s_b/h_b) gy (BF16), MFMA weight
// gradients and transposed-convolution input gradients through the ReLU masks, accumulated
// into grad = [dW (kNNW) | db (kNNB)] with FP32 atomics
__global__ __launch_bounds__(512) void k_nn_train_bwd(MeshView M, NNView W, const int* __restrict__ list,
                                                      const double* __restrict__ rp, const double* __restrict__ gy,
                                                      const float* __restrict__ sb, const uint16_t* __restrict__ scratch,
                                                      float* __restrict__ grad) {
    __shared__ __attribute__((aligned(16))) NNShared L;
    __shared__ __attribute__((aligned(16))) uint16_t T2[kNNTile];
    const NNBrick B = nn_brick(list);
    const float s = sb[blockIdx.x];
    if (!(s > 0.f)) return;  // block-uniform
    const int q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(B.b) * 512 + q;
    const uint8_t m = M.mask[a];
    const uint16_t* store = scratch + static_cast<size_t>(blockIdx.x) * kNNStore;
    nn_zero_tile(L.T0);
    nn_zero_tile(L.T1);
    nn_zero_tile(T2);
    __syncthreads();
    // layer-0 input (features) into T1, output gradient into channel 0 of T0
    nn_features(static_cast<float>(rp[a]), m, L.T1, L);
    if (m & kActive) {
        const float gq = static_cast<float>(double(s) / level_h(M, B.b) * gy[static_cast<int64_t>(blockIdx.x) * 512 + q]);
        L.T0[nn_pos(q) * kNNC] = nn_f2bf(gq);
    }
    __syncthreads();
    // Three tiles: gcur (gradient w.r.t. layer l's pre-activation output), gnext and T1 (the
    // features, input of layer 0).  The input of layer l > 0 is staged from scratch into
    // gnext, its ReLU pattern is kept in registers, and gnext is then overwritten by the
    // transposed convolution.
    uint16_t* gcur = L.T0;
    uint16_t* gnext = T2;
    for (int l = kNNLayers - 1; l >= 0; --l) {
        uint16_t* ain = L.T1;
        if (l > 0) {
            ain = gnext;
            reinterpret_cast<uint4*>(ain)[nn_pos(q)] =
                *reinterpret_cast<const uint4*>(store + (static_cast<size_t>(l - 1) * 512 + q) * kNNC);
        }
        __syncthreads();
        nn_weight_grad(gcur, ain, grad + static_cast<size_t>(l) * kNNWLayer, grad + kNNW + l * kNNC);
        if (l > 0) {
            // ReLU pattern of layer l-1's output (this layer's input) at cell q
            uint32_t pmask = 0;
            for (int c = 0; c < kNNC; ++c)
                if (nn_bf2f(ain[nn_pos(q) * kNNC + c]) > 0.f) pmask |= 1u << c;
            nn_load_weights(W.wt + static_cast<size_t>(l) * kNNWLayer, L.WL);
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
    const int l = i / kNNWLayer, rem = i % kNNWLayer, t = rem / (kNNC * kNNC), cout = (rem / kNNC) % kNNC, cin = rem % kNNC;
    const uint16_t bits = nn_f2bf(w32[i]);
    w[i] = bits;
    wt[static_cast<size_t>(l) * kNNWLayer + ((26 - t) * kNNC + cin) * kNNC + cout] = bits;
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

// per-brick partials of sum_active v^2 / h_b (normalisation of a recorded residual)
__global__ __launch_bounds__(512) void k_nn_norm(MeshView M, const double* __restrict__ v, double* partials) {
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const double x = (M.mask[a] & kActive) ? v[a] : 0.0;
    const double t = block_sum<512>(x * x / level_h(M, b));
    if (q == 0) partials[b] = t;
}

}  // namespace pcfd
