// Device-side topology preprocessing (pcfd_topology_build_on_device): the per-cell stages of
// the topology builder run on a GPU, bit-identical to the host builder.
//
//   sample           one wave64 per cut-list cell, four cells per 256-lane workgroup: the lanes
//                    evaluate the ns^3 + 6 ns^2 point-in-solid samples into LDS, then lanes 0..7
//                    each reduce one part (volume, wall, six faces) with the host's code and
//                    summation order (solid.hpp).
//   connected_fluid  lock-free union-find over the cell graph: every active cell links its +x,
//                    +y, +z open faces (graph_neighbors, build_graph.hpp) by compare-and-swap on
//                    roots, a larger root always under a smaller one, with path halving in the
//                    finds; then every cell points at its root (the lowest cell of its
//                    component), the roots touching the x-max boundary (or the largest root) are
//                    kept.  The components, hence the kept cells, are those of the host's flood
//                    fill.
//   cell_passes      one 512-lane workgroup per brick: cell_geometry on every cell, then
//                    cell_faces on every active cell twice, first counting each cell's compact
//                    edges and wall moment (a block scan gives the cell's offset in its brick, a
//                    one-workgroup scan the brick's), then writing them at those offsets: the
//                    host's dof-then-face order.
//   multigrid        the coarse levels (build_multigrid.hip).
// The surrounding stages (refinement bands, the cut list, brick ordering, special cells, the
// dense coarse inverse) stay on the host.
//
// Built with -ffp-contract=off: no multiply-add is fused, as on the host.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "topology.hpp"

namespace pcfd {
void set_error(const std::string& m);
namespace devbuild {
double multigrid(hipStream_t stream, hipEvent_t e0, hipEvent_t e1, const Topology& T, int levels, int coarse_max,
                 std::vector<MGLevel>& mg);
}
namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string("device build: ") + what + ": " + hipGetErrorString(e));
}

constexpr int kCellsPerBlock = 4;
constexpr int kBitsPerCell = cell_samples(kMaxDeviceSubsamples);

// ------------------------------------------------------------------------------------------
// cut-cell sub-sampling
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sample_cells(SolidDesc G, double lo0, double lo1, double lo2,
                                                      const double* __restrict__ level_h, int ns,
                                                      const int* __restrict__ cells, int64_t n,
                                                      CutData* __restrict__ out) {
    __shared__ uint8_t bits[kCellsPerBlock][kBitsPerCell];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kCellsPerBlock + w;
    const bool live = i < n;
    const double lo[3] = {lo0, lo1, lo2};
    int I[3] = {0, 0, 0};
    double h = 0;
    if (live) {
        for (int a = 0; a < 3; ++a) I[a] = cells[4 * i + a];
        h = level_h[cells[4 * i + 3]];
        const int total = cell_samples(ns);
        for (int s = lane; s < total; s += 64) bits[w][s] = sample_solid(G, lo, I, h, ns, s) ? 1 : 0;
    }
    __syncthreads();
    if (live && lane < 8) cut_part(bits[w], lo, I, h, ns, lane, out[i]);
}

// ------------------------------------------------------------------------------------------
// connected fluid
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// root of x; every parent is an ancestor with a smaller index, so halving a path keeps the tree
__device__ int find_root(int* parent, int x) {
    int px = ld(parent + x);
    while (px != x) {
        const int ppx = ld(parent + px);
        if (ppx != px) st(parent + x, ppx);
        x = px;
        px = ppx;
    }
    return x;
}

__device__ void unite(int* parent, int a, int b) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    while (a != b) {
        if (a < b) { const int t = a; a = b; b = t; }
        const int old = atomicCAS(parent + a, a, b);
        if (old == a) return;
        a = find_root(parent, old);
        b = find_root(parent, b);
    }
}

__device__ __forceinline__ void raise_error(int* err, int code) { atomicMax(err, code); }

__global__ __launch_bounds__(256) void k_cc_init(int* parent, int n) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < n) parent[c] = c;
}

__global__ __launch_bounds__(256) void k_cc_link(BuildView V, int* parent, int n, int* err) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n || !V.active[c]) return;
    for (int dir = 1; dir < 6; dir += 2) {
        Neighbor nbs[4];
        bool bnd;
        const int m = graph_neighbors(V, c >> 9, c & 511, dir, nbs, bnd);
        if (m < 0) { raise_error(err, -m); return; }
        for (int k = 0; k < m; ++k) {
            const int o = nbs[k].leaf * 512 + nbs[k].local;
            if (nbs[k].ap > 0 && V.active[o]) unite(parent, c, o);
        }
    }
}

__global__ __launch_bounds__(256) void k_cc_compress(int* parent, int n) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < n) parent[c] = find_root(parent, c);
}

// outlet: flag the roots of components with a cell on the x-max boundary; else count each
// root's cells
__global__ __launch_bounds__(256) void k_cc_mark(BuildView V, const int* __restrict__ parent, int n, int outlet,
                                                 uint8_t* flag, int* size, int* err) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n || !V.active[c]) return;
    if (outlet) {
        Neighbor nbs[4];
        bool bnd;
        const int m = graph_neighbors(V, c >> 9, c & 511, 1, nbs, bnd);
        if (m < 0) { raise_error(err, -m); return; }
        if (bnd) flag[parent[c]] = 1;
    } else {
        atomicAdd(size + parent[c], 1);
    }
}

// the largest component, ties to the lowest root: max of (size, ~root)
__global__ __launch_bounds__(256) void k_cc_largest(const uint8_t* __restrict__ active, const int* __restrict__ parent,
                                                    const int* __restrict__ size, int n, unsigned long long* best) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n || !active[c] || parent[c] != c) return;
    const unsigned long long key = (static_cast<unsigned long long>(size[c]) << 32) | (0xffffffffu - static_cast<unsigned>(c));
    atomicMax(best, key);
}

__global__ __launch_bounds__(256) void k_cc_apply(uint8_t* active, const int* __restrict__ parent, int n, int outlet,
                                                  const uint8_t* __restrict__ flag, const unsigned long long* best) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n || !active[c]) return;
    const bool keep = outlet ? flag[parent[c]] != 0 : parent[c] == static_cast<int>(0xffffffffu - static_cast<unsigned>(*best));
    if (!keep) active[c] = 0;
}

// ------------------------------------------------------------------------------------------
// finalize's per-cell passes
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_cell_geometry(BuildView V, double* __restrict__ volume,
                                                       double* __restrict__ centroid, uint8_t* __restrict__ embedded,
                                                       int* err) {
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    CellGeometry g;
    const int e = cell_geometry(V, b, q, g);
    if (e != GE_NONE) raise_error(err, e);
    volume[a] = g.volume;
    for (int r = 0; r < 3; ++r) centroid[3 * a + r] = g.centroid[r];
    embedded[a] = g.embedded;
}

// exclusive scan of two counts over the 512 lanes of a workgroup; returns the totals
__device__ void block_scan2(int& x, int& y, int& tx, int& ty) {
    __shared__ int wx[8], wy[8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int ix = x, iy = y;
    for (int d = 1; d < 64; d <<= 1) {
        const int ux = __shfl_up(ix, d, 64), uy = __shfl_up(iy, d, 64);
        if (lane >= d) { ix += ux; iy += uy; }
    }
    if (lane == 63) { wx[w] = ix; wy[w] = iy; }
    __syncthreads();
    int bx = 0, by = 0;
    tx = 0;
    ty = 0;
    for (int k = 0; k < 8; ++k) {
        if (k < w) { bx += wx[k]; by += wy[k]; }
        tx += wx[k];
        ty += wy[k];
    }
    x = bx + ix - x;
    y = by + iy - y;
}

// mode 0: per-cell outputs, each cell's compact-edge and wall counts scanned within the brick
// (local offsets and brick totals); mode 1: the edges and wall moments at their offsets
__global__ __launch_bounds__(512) void k_cell_faces(BuildView V, const double* __restrict__ centroid,
                                                    const uint8_t* __restrict__ embedded, int mode,
                                                    uint8_t* __restrict__ mask, uint8_t* __restrict__ has_compact,
                                                    double* __restrict__ diffusion_sum, double* __restrict__ S,
                                                    int* __restrict__ local_off, int64_t* __restrict__ brick_count,
                                                    const int64_t* __restrict__ brick_base, EdgeRecord* __restrict__ edges,
                                                    int64_t* __restrict__ wall_dof, double* __restrict__ wall_moment,
                                                    int* err) {
    const int b = blockIdx.x, q = threadIdx.x;
    const int64_t a = static_cast<int64_t>(b) * 512 + q;
    const bool act = V.active[a] != 0;
    CellFaces f;
    int ne = 0;
    bool wall = false;
    if (act) {
        int e;
        if (mode == 0) {
            e = cell_faces(V, centroid, embedded, b, q, f, [&](const EdgeRecord&) { ++ne; });
        } else {
            const int64_t base = brick_base[2 * b] + local_off[2 * a];
            e = cell_faces(V, centroid, embedded, b, q, f, [&](const EdgeRecord& r) { edges[base + ne++] = r; });
        }
        if (e != GE_NONE) raise_error(err, e);
        const double h = V.h[V.leaf_level[b]];
        wall = closure_norm(f.S) > 1e-12 * h * h;
    }
    if (mode == 0) {
        mask[a] = act ? f.mask : 0;
        has_compact[a] = act ? f.has_compact : 0;
        diffusion_sum[a] = act ? f.diffusion_sum : 0.0;
        for (int r = 0; r < 3; ++r) S[3 * a + r] = act ? f.S[r] : 0.0;
        int x = ne, y = wall ? 1 : 0, tx, ty;
        block_scan2(x, y, tx, ty);
        local_off[2 * a] = x;
        local_off[2 * a + 1] = y;
        if (q == 0) {
            brick_count[2 * b] = tx;
            brick_count[2 * b + 1] = ty;
        }
    } else if (wall) {
        const int64_t slot = brick_base[2 * b + 1] + local_off[2 * a + 1];
        wall_dof[slot] = a;
        for (int r = 0; r < 9; ++r) wall_moment[9 * slot + r] = f.fm[r];
    }
}

// exclusive scan of the bricks' two counts (one workgroup); base[2 nb], base[2 nb + 1] = totals
__global__ __launch_bounds__(1024) void k_scan_bricks(const int64_t* __restrict__ count, int nb, int64_t* __restrict__ base) {
    __shared__ int64_t part[2][1024];
    __shared__ int64_t carry[2];
    if (threadIdx.x < 2) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int64_t x = i < nb ? count[2 * i] : 0, y = i < nb ? count[2 * i + 1] : 0;
        part[0][threadIdx.x] = x;
        part[1][threadIdx.x] = y;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int64_t ux = threadIdx.x >= d ? part[0][threadIdx.x - d] : 0;
            const int64_t uy = threadIdx.x >= d ? part[1][threadIdx.x - d] : 0;
            __syncthreads();
            part[0][threadIdx.x] += ux;
            part[1][threadIdx.x] += uy;
            __syncthreads();
        }
        if (i < nb) {
            base[2 * i] = carry[0] + part[0][threadIdx.x] - x;
            base[2 * i + 1] = carry[1] + part[1][threadIdx.x] - y;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            carry[0] += part[0][1023];
            carry[1] += part[1][1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        base[2 * nb] = carry[0];
        base[2 * nb + 1] = carry[1];
    }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
// device allocation freed with its owner
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void* alloc(size_t bytes) {
        release();
        if (bytes) check(hipMalloc(&p_, bytes), "hipMalloc");
        return p_;
    }
    template <class T>
    T* upload(const T* src, size_t count, hipStream_t s) {
        T* d = static_cast<T*>(alloc(count * sizeof(T)));
        if (count) check(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        return d;
    }
    template <class T>
    T* get() const { return static_cast<T*>(p_); }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
    }

private:
    void* p_ = nullptr;
};

class DeviceBuild final : public BuildAccelerator {
public:
    explicit DeviceBuild(int device) : device_(device) {
        check(hipGetDevice(&prev_), "hipGetDevice");
        try {
            check(hipSetDevice(device_), "hipSetDevice");
            check(hipStreamCreate(&stream_), "hipStreamCreate");
            for (auto& e : ev_) check(hipEventCreate(&e), "hipEventCreate");
            check(hipMalloc(reinterpret_cast<void**>(&err_), 16), "hipMalloc");
        } catch (...) {
            release();
            throw;
        }
    }
    ~DeviceBuild() override { release(); }
    int device() const override { return device_; }
    double kernel_ms() const override { return kernel_ms_; }

    void sample(const SolidDesc& solid, const double lo[3], const std::vector<double>& level_h, int ns,
                const std::vector<int>& cells, std::vector<CutData>& out) override {
        if (ns < 1 || ns > kMaxDeviceSubsamples)
            throw BuildError("pcfd_topology_build_on_device: subsamples must be in 1..16");
        const int64_t n = static_cast<int64_t>(cells.size() / 4);
        out.resize(n);
        ncuts_ = n;
        if (n == 0) return;
        DevBuf px, pz, boff, bseg, h, c;
        SolidDesc G = solid;
        if (solid.np > 0) {
            G.px = px.upload(solid.px, solid.np, stream_);
            G.pz = pz.upload(solid.pz, solid.np, stream_);
            G.boff = boff.upload(solid.boff, static_cast<size_t>(solid.nb) + 1, stream_);
            G.bseg = bseg.upload(solid.bseg, static_cast<size_t>(solid.boff[solid.nb]), stream_);
        }
        const double* dh = h.upload(level_h.data(), level_h.size(), stream_);
        const int* dc = c.upload(cells.data(), cells.size(), stream_);
        CutData* d = static_cast<CutData*>(cuts_.alloc(n * sizeof(CutData)));
        const int64_t blocks = (n + kCellsPerBlock - 1) / kCellsPerBlock;
        if (blocks > 0x7fffffff) throw BuildError("pcfd_topology_build_on_device: cut list too long");
        begin();
        k_sample_cells<<<static_cast<unsigned>(blocks), 256, 0, stream_>>>(G, lo[0], lo[1], lo[2], dh, ns, dc, n, d);
        check(hipGetLastError(), "k_sample_cells");
        end();
        check(hipMemcpyAsync(out.data(), d, n * sizeof(CutData), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
        check(hipStreamSynchronize(stream_), "k_sample_cells");
        accumulate();
    }

    void connected_fluid(const BuildView& view, int64_t nleaves, bool outlet, std::vector<uint8_t>& active) override {
        const int64_t n64 = nleaves * 512;
        if (n64 > 0x7fffffff) throw BuildError("pcfd_topology_build_on_device: too many cells");
        const int n = static_cast<int>(n64);
        if (n == 0) return;
        Graph g;
        BuildView V = upload_graph(view, nleaves, g);
        DevBuf parent, flag, size, best;
        int* p = static_cast<int*>(parent.alloc(n * sizeof(int)));
        uint8_t* fl = static_cast<uint8_t*>(flag.alloc(n));
        int* sz = static_cast<int*>(size.alloc(n * sizeof(int)));
        auto* bst = static_cast<unsigned long long*>(best.alloc(sizeof(unsigned long long)));
        check(hipMemsetAsync(fl, 0, n, stream_), "hipMemsetAsync");
        check(hipMemsetAsync(sz, 0, n * sizeof(int), stream_), "hipMemsetAsync");
        check(hipMemsetAsync(bst, 0, sizeof(unsigned long long), stream_), "hipMemsetAsync");
        check(hipMemsetAsync(err_, 0, sizeof(int), stream_), "hipMemsetAsync");
        const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
        uint8_t* act = const_cast<uint8_t*>(V.active);
        begin();
        k_cc_init<<<blocks, 256, 0, stream_>>>(p, n);
        k_cc_link<<<blocks, 256, 0, stream_>>>(V, p, n, err_);
        k_cc_compress<<<blocks, 256, 0, stream_>>>(p, n);
        k_cc_mark<<<blocks, 256, 0, stream_>>>(V, p, n, outlet ? 1 : 0, fl, sz, err_);
        if (!outlet) k_cc_largest<<<blocks, 256, 0, stream_>>>(act, p, sz, n, bst);
        k_cc_apply<<<blocks, 256, 0, stream_>>>(act, p, n, outlet ? 1 : 0, fl, bst);
        check(hipGetLastError(), "connected fluid kernels");
        end();
        check(hipMemcpyAsync(active.data(), act, n, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
        raise_device_error();
        accumulate();
    }

    void cell_passes(const BuildView& view, int nbricks, CellPasses& P) override {
        const int64_t ndof = static_cast<int64_t>(nbricks) * 512;
        P = CellPasses{};
        if (nbricks == 0) return;
        Graph g;
        BuildView V = upload_graph(view, nbricks, g);
        DevBuf vol, cen, emb, msk, hc, ds, S, loff, bcount, bbase;
        double* dvol = static_cast<double*>(vol.alloc(ndof * sizeof(double)));
        double* dcen = static_cast<double*>(cen.alloc(3 * ndof * sizeof(double)));
        uint8_t* demb = static_cast<uint8_t*>(emb.alloc(ndof));
        uint8_t* dmsk = static_cast<uint8_t*>(msk.alloc(ndof));
        uint8_t* dhc = static_cast<uint8_t*>(hc.alloc(ndof));
        double* dds = static_cast<double*>(ds.alloc(ndof * sizeof(double)));
        double* dS = static_cast<double*>(S.alloc(3 * ndof * sizeof(double)));
        int* dloff = static_cast<int*>(loff.alloc(2 * ndof * sizeof(int)));
        int64_t* dbc = static_cast<int64_t*>(bcount.alloc(2 * static_cast<size_t>(nbricks) * sizeof(int64_t)));
        int64_t* dbb = static_cast<int64_t*>(bbase.alloc((2 * static_cast<size_t>(nbricks) + 2) * sizeof(int64_t)));
        check(hipMemsetAsync(err_, 0, sizeof(int), stream_), "hipMemsetAsync");
        begin();
        k_cell_geometry<<<nbricks, 512, 0, stream_>>>(V, dvol, dcen, demb, err_);
        k_cell_faces<<<nbricks, 512, 0, stream_>>>(V, dcen, demb, 0, dmsk, dhc, dds, dS, dloff, dbc, nullptr, nullptr,
                                                   nullptr, nullptr, err_);
        k_scan_bricks<<<1, 1024, 0, stream_>>>(dbc, nbricks, dbb);
        check(hipGetLastError(), "cell pass kernels");
        end();
        int64_t totals[2] = {0, 0};
        check(hipMemcpyAsync(totals, dbb + 2 * static_cast<size_t>(nbricks), sizeof(totals), hipMemcpyDeviceToHost, stream_),
              "hipMemcpyAsync");
        raise_device_error();
        accumulate();
        DevBuf edges, wdof, wmom;
        EdgeRecord* de = static_cast<EdgeRecord*>(edges.alloc(totals[0] * sizeof(EdgeRecord)));
        int64_t* dwd = static_cast<int64_t*>(wdof.alloc(totals[1] * sizeof(int64_t)));
        double* dwm = static_cast<double*>(wmom.alloc(9 * totals[1] * sizeof(double)));
        begin();
        k_cell_faces<<<nbricks, 512, 0, stream_>>>(V, dcen, demb, 1, nullptr, nullptr, nullptr, nullptr, dloff, nullptr,
                                                   dbb, de, dwd, dwm, err_);
        check(hipGetLastError(), "k_cell_faces");
        end();
        P.volume.resize(ndof);
        P.centroid.resize(3 * ndof);
        P.mask.resize(ndof);
        P.has_compact.resize(ndof);
        P.diffusion_sum.resize(ndof);
        P.S.resize(3 * ndof);
        P.edges.resize(totals[0]);
        P.wall_dof.resize(totals[1]);
        P.wall_moment.resize(9 * totals[1]);
        download(P.volume.data(), dvol, ndof * sizeof(double));
        download(P.centroid.data(), dcen, 3 * ndof * sizeof(double));
        download(P.mask.data(), dmsk, ndof);
        download(P.has_compact.data(), dhc, ndof);
        download(P.diffusion_sum.data(), dds, ndof * sizeof(double));
        download(P.S.data(), dS, 3 * ndof * sizeof(double));
        download(P.edges.data(), de, totals[0] * sizeof(EdgeRecord));
        download(P.wall_dof.data(), dwd, totals[1] * sizeof(int64_t));
        download(P.wall_moment.data(), dwm, 9 * totals[1] * sizeof(double));
        raise_device_error();
        accumulate();
    }

    bool multigrid(const Topology& T, int levels, int coarse_max, std::vector<MGLevel>& mg) override {
        kernel_ms_ += devbuild::multigrid(stream_, ev_[0], ev_[1], T, levels, coarse_max, mg);
        return true;
    }

private:
    void release() {
        for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
        for (auto& e : ev_) e = nullptr;
        if (err_) (void)hipFree(err_);
        err_ = nullptr;
        cuts_.release();
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        (void)hipSetDevice(prev_);
    }

    struct Graph {
        DevBuf status[4], level, coord, cut, active;
    };

    // the cell graph on the device (the cut data stay there from sample)
    BuildView upload_graph(const BuildView& host, int64_t nleaves, Graph& g) {
        BuildView V = host;
        for (int l = 0; l < host.levels; ++l) {
            const size_t nb = static_cast<size_t>(host.nb[l][0]) * host.nb[l][1] * host.nb[l][2];
            V.status[l] = g.status[l].upload(host.status[l], nb, stream_);
        }
        V.leaf_level = g.level.upload(host.leaf_level, nleaves, stream_);
        V.leaf_coord = g.coord.upload(host.leaf_coord, 3 * nleaves, stream_);
        V.cut = g.cut.upload(host.cut, 512 * nleaves, stream_);
        V.active = g.active.upload(host.active, 512 * nleaves, stream_);
        V.cuts = cuts_.get<CutData>();
        if (ncuts_ > 0 && !V.cuts) throw DeviceError("device build: cut data not on the device");
        return V;
    }
    void download(void* dst, const void* src, size_t bytes) {
        if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    }
    void raise_device_error() {
        int e = 0;
        check(hipMemcpyAsync(&e, err_, sizeof(int), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
        check(hipStreamSynchronize(stream_), "device build");
        if (e != GE_NONE) throw BuildError(graph_error_text(e));
    }
    void begin() { check(hipEventRecord(ev_[0], stream_), "hipEventRecord"); }
    void end() { check(hipEventRecord(ev_[1], stream_), "hipEventRecord"); }
    void accumulate() {
        check(hipStreamSynchronize(stream_), "device build");
        float ms = 0;
        check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
        kernel_ms_ += ms;
    }

    int device_, prev_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[2] = {};
    int* err_ = nullptr;
    DevBuf cuts_;
    int64_t ncuts_ = 0;
    double kernel_ms_ = 0;
};

}  // namespace
}  // namespace pcfd

extern "C" pcfd_status pcfd_topology_build_on_device(const pcfd_geometry_desc* geometry, const pcfd_grid_desc* grid,
                                                     int device, pcfd_topology** out) {
    if (out) *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        pcfd::set_error("pcfd_topology_build_on_device: no such HIP device");
        return PCFD_EDEVICE;
    }
    try {
        pcfd::DeviceBuild accel(device);
        return pcfd::build_topology_handle("pcfd_topology_build_on_device", geometry, grid, &accel, out);
    } catch (const pcfd::DeviceError& e) {
        pcfd::set_error(std::string("pcfd_topology_build_on_device: ") + e.what());
        return PCFD_EDEVICE;
    } catch (const std::bad_alloc&) {
        pcfd::set_error("pcfd_topology_build_on_device: out of host memory");
        return PCFD_ENOMEM;
    }
}
