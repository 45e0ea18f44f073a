// Device build of the geometric multigrid hierarchy (part of pcfd_topology_build_on_device),
// bit-identical to the host builder's coarsening (topology.cpp, Builder::host_multigrid).
//
// Each coarsening runs the host's steps as data-parallel passes:
//   * blocked parents: every active cell of size exponent e lists its ancestors' keys at
//     exponents e+2 .. max+1 (mg_key: exponent over Morton order); radix sort, unique;
//   * parent keys: every active cell proposes its 2x2x2 parent unless that key is blocked (then
//     it is carried over); the (key, cell) pairs are radix-sorted (stable, so equal keys keep the
//     cells in ascending order, as the host's sort of pairs does), heads of equal keys numbered
//     by a scan: parents, children lists and the coarse cells' exponents and positions;
//   * coarse face slots: every fine face between two coarse cells names its slot (the finer
//     side's, or the lower side's for equal sizes, per direction); (slot, face) pairs are
//     stable-sorted by slot so that each slot sums its faces' areas in the host's face order;
//   * rows: each slot gives an entry to both its cells' rows; the entries stable-sorted by row
//     keep the slot order, in which each row sums its diagonal before sorting itself by
//     (column, coefficient) -- the host's insertion order and row sort;
//   * the next level's faces are the slots in slot order.
// Composite-level faces (structured faces per active cell in dof then axis order, then the
// compact edges) are laid out with a scan.  The dense coarse inverse stays on the host.
//
// Built with -ffp-contract=off: no multiply-add is fused, as on the host.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>
#include <string>
#include <vector>

#include "topology.hpp"

namespace pcfd {
namespace devbuild {
namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string("device multigrid: ") + what + ": " + hipGetErrorString(e));
}

// device array that only grows (regrowing waits for the device: no kernel still uses the old one)
template <class T>
class Buf {
public:
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p_) (void)hipFree(p_);
    }
    T* get(size_t count) {
        if (count > cap_) {
            check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            if (p_) (void)hipFree(p_);
            p_ = nullptr;
            cap_ = 0;
            check(hipMalloc(reinterpret_cast<void**>(&p_), (count ? count : 1) * sizeof(T)), "hipMalloc");
            cap_ = count;
        }
        return p_;
    }
    T* ptr() const { return p_; }

private:
    T* p_ = nullptr;
    size_t cap_ = 0;
};

constexpr int kT = 256;
unsigned grid(int64_t n) { return static_cast<unsigned>((n + kT - 1) / kT); }
int bits_for(int64_t v) {
    int b = 1;
    while (b < 62 && (int64_t(1) << b) <= v) ++b;
    return b;
}

// ---- composite level ---------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_l0_cells(const uint8_t* __restrict__ mask, const int* __restrict__ blevel,
                                                 const int* __restrict__ bcoord, const uint8_t* __restrict__ bflags,
                                                 const double* __restrict__ level_h, int levels, int outlet, int64_t ndof,
                                                 int* __restrict__ exp, int* __restrict__ coord, uint8_t* __restrict__ act,
                                                 double* __restrict__ out_area, int64_t* __restrict__ fcount) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (a >= ndof) return;
    const int b = static_cast<int>(a >> 9), q = static_cast<int>(a & 511), l = blevel[b];
    exp[a] = levels - 1 - l;
    coord[3 * a] = bcoord[3 * b] * 8 + (q & 7);
    coord[3 * a + 1] = bcoord[3 * b + 1] * 8 + ((q >> 3) & 7);
    coord[3 * a + 2] = bcoord[3 * b + 2] * 8 + (q >> 6);
    const uint8_t m = mask[a];
    const bool ac = (m & 64) != 0;
    act[a] = ac ? 1 : 0;
    const double h = level_h[l];
    out_area[a] = ac && outlet && (bflags[b] & 2) && (q & 7) == 7 ? h * h : 0.0;
    fcount[a] = ac ? ((m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1)) : 0;
}

__global__ __launch_bounds__(kT) void k_l0_faces(const uint8_t* __restrict__ mask, const int* __restrict__ blevel,
                                                 const int* __restrict__ bnbr, const double* __restrict__ level_h,
                                                 int64_t ndof, const int64_t* __restrict__ fpos, int* __restrict__ fa,
                                                 int* __restrict__ fb, int* __restrict__ fax, double* __restrict__ far) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (a >= ndof) return;
    const uint8_t m = mask[a];
    if (!(m & 64)) return;
    const int b = static_cast<int>(a >> 9), q = static_cast<int>(a & 511);
    const double h = level_h[blevel[b]];
    int64_t p = fpos[a];
    for (int axis = 0; axis < 3; ++axis) {
        if (!(m & (1 << axis))) continue;
        int c[3] = {q & 7, (q >> 3) & 7, q >> 6};
        int nbk = b;
        if (++c[axis] == 8) {
            c[axis] = 0;
            nbk = bnbr[6 * b + 2 * axis + 1];
        }
        fa[p] = static_cast<int>(a);
        fb[p] = nbk * 512 + c[0] + 8 * c[1] + 64 * c[2];
        fax[p] = axis;
        far[p] = h * h;
        ++p;
    }
}

// ---- one coarsening ----------------------------------------------------------------------
// active count and largest exponent (block reduction, one atomic per block)
__global__ __launch_bounds__(kT) void k_stats(int64_t n, const uint8_t* __restrict__ act, const int* __restrict__ exp,
                                              unsigned long long* nact, int* maxe) {
    __shared__ int sc[kT], se[kT];
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    const bool ac = a < n && act[a];
    sc[threadIdx.x] = ac ? 1 : 0;
    se[threadIdx.x] = ac ? exp[a] : 0;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sc[threadIdx.x] += sc[threadIdx.x + s];
            se[threadIdx.x] = se[threadIdx.x] > se[threadIdx.x + s] ? se[threadIdx.x] : se[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(nact, static_cast<unsigned long long>(sc[0]));
        atomicMax(maxe, se[0]);
    }
}

__global__ __launch_bounds__(kT) void k_blocked_count(int64_t n, const uint8_t* __restrict__ act, const int* __restrict__ exp,
                                                      int maxe, int64_t* __restrict__ cnt) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (a >= n) return;
    const int k = maxe - exp[a];
    cnt[a] = act[a] && k > 0 ? k : 0;
}

__global__ __launch_bounds__(kT) void k_blocked_write(int64_t n, const uint8_t* __restrict__ act, const int* __restrict__ exp,
                                                      const int* __restrict__ coord, int maxe,
                                                      const int64_t* __restrict__ pos, uint64_t* __restrict__ keys) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (a >= n || !act[a]) return;
    const int e = exp[a];
    int64_t p = pos[a];
    for (int E = e + 2; E <= maxe + 1; ++E) {
        const int s = E - e;
        keys[p++] = mg_key(E, coord[3 * a] >> s, coord[3 * a + 1] >> s, coord[3 * a + 2] >> s);
    }
}

__device__ bool sorted_contains(const uint64_t* v, int64_t n, uint64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (v[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && v[lo] == key;
}

__global__ __launch_bounds__(kT) void k_parent_keys(int64_t n, const uint8_t* __restrict__ act, const int* __restrict__ exp,
                                                    const int* __restrict__ coord, const uint64_t* __restrict__ blocked,
                                                    const int64_t* nblocked, const int64_t* __restrict__ apos,
                                                    uint64_t* __restrict__ keys, int* __restrict__ vals,
                                                    uint8_t* __restrict__ grows) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (a >= n || !act[a]) return;
    const int e = exp[a], x = coord[3 * a], y = coord[3 * a + 1], z = coord[3 * a + 2];
    uint64_t key = mg_key(e + 1, x >> 1, y >> 1, z >> 1);
    uint8_t g = 1;
    if (sorted_contains(blocked, *nblocked, key)) {
        key = mg_key(e, x, y, z);
        g = 0;
    }
    grows[a] = g;
    const int64_t p = apos[a];
    keys[p] = key;
    vals[p] = static_cast<int>(a);
}

__global__ __launch_bounds__(kT) void k_heads(int64_t m, const uint64_t* __restrict__ sk, int64_t* __restrict__ head) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < m) head[i] = i == 0 || sk[i] != sk[i - 1] ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_coarse_cells(int64_t m, const int* __restrict__ sv, const int64_t* __restrict__ idp1,
                                                     const int* __restrict__ exp, const int* __restrict__ coord,
                                                     const uint8_t* __restrict__ grows, int* __restrict__ parent,
                                                     int* __restrict__ child, int* __restrict__ child_off,
                                                     int* __restrict__ nexp, int* __restrict__ ncoord) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= m) return;
    const int a = sv[i];
    const int id = static_cast<int>(idp1[i] - 1);
    parent[a] = id;
    child[i] = a;
    if (i == 0 || idp1[i - 1] != idp1[i]) {
        child_off[id] = static_cast<int>(i);
        const int s = grows[a];
        nexp[id] = exp[a] + s;
        for (int r = 0; r < 3; ++r) ncoord[3 * id + r] = coord[3 * static_cast<int64_t>(a) + r] >> s;
    }
}

__global__ __launch_bounds__(kT) void k_coarse_outlet(int nc, const int* __restrict__ child_off, const int* __restrict__ child,
                                                      const double* __restrict__ outlet, double* __restrict__ cout) {
    const int c = blockIdx.x * kT + threadIdx.x;
    if (c >= nc) return;
    double s = 0.0;
    for (int i = child_off[c]; i < child_off[c + 1]; ++i) s += outlet[child[i]];
    cout[c] = s;
}

// the slot of a fine face between coarse cells pa != pb: the finer side's (the lower side's for
// equal sizes) face towards the other
__device__ __forceinline__ void face_slot(int pa, int pb, int axis, const int* nexp, int& slot, int& other) {
    int owner, dir;
    if (nexp[pa] <= nexp[pb]) {
        owner = pa; dir = 2 * axis + 1; other = pb;
    } else {
        owner = pb; dir = 2 * axis; other = pa;
    }
    slot = 6 * owner + dir;
}

__global__ __launch_bounds__(kT) void k_slot_flags(int64_t nf, const int* __restrict__ fa, const int* __restrict__ fb,
                                                   const int* __restrict__ parent, int64_t* __restrict__ flag) {
    const int64_t f = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (f >= nf) return;
    const int pa = parent[fa[f]], pb = parent[fb[f]];
    flag[f] = pa >= 0 && pb >= 0 && pa != pb ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_slot_pairs(int64_t nf, const int* __restrict__ fa, const int* __restrict__ fb,
                                                   const int* __restrict__ fax, const int* __restrict__ parent,
                                                   const int* __restrict__ nexp, const int64_t* __restrict__ flag,
                                                   const int64_t* __restrict__ pos, int* __restrict__ sk, int* __restrict__ sv) {
    const int64_t f = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (f >= nf || !flag[f]) return;
    int slot, other;
    face_slot(parent[fa[f]], parent[fb[f]], fax[f], nexp, slot, other);
    sk[pos[f]] = slot;
    sv[pos[f]] = static_cast<int>(f);
}

__global__ __launch_bounds__(kT) void k_slot_sum(int64_t np, const int* __restrict__ sk, const int* __restrict__ sv,
                                                 const int* __restrict__ fa, const int* __restrict__ fb,
                                                 const int* __restrict__ fax, const double* __restrict__ far,
                                                 const int* __restrict__ parent, const int* __restrict__ nexp,
                                                 int* __restrict__ slot_nbr, double* __restrict__ slot_area, int* err) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i >= np || (i > 0 && sk[i] == sk[i - 1])) return;
    const int s = sk[i];
    double area = 0.0;
    int nbr = -1;
    for (int64_t j = i; j < np && sk[j] == s; ++j) {
        const int f = sv[j];
        int slot, other;
        face_slot(parent[fa[f]], parent[fb[f]], fax[f], nexp, slot, other);
        if (nbr >= 0 && nbr != other) atomicMax(err, 1);
        nbr = other;
        area += far[f];
    }
    slot_nbr[s] = nbr;
    slot_area[s] = area;
}

__global__ __launch_bounds__(kT) void k_slot_used(int64_t ns, const int* __restrict__ slot_nbr, int64_t* __restrict__ flag) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (s < ns) flag[s] = slot_nbr[s] >= 0 ? 1 : 0;
}

// the two row entries of every used slot (entry 2j: the slot's own cell, 2j+1: its neighbour)
// and the next level's face (lower cell, upper cell, axis, area) in slot order
__global__ __launch_bounds__(kT) void k_entries(int64_t ns, const int* __restrict__ slot_nbr,
                                                const double* __restrict__ slot_area, const int* __restrict__ nexp,
                                                double hf, const int64_t* __restrict__ rank, int* __restrict__ erow,
                                                int* __restrict__ eidx, int* __restrict__ ecol, double* __restrict__ ek,
                                                int* __restrict__ deg, int* __restrict__ nfa, int* __restrict__ nfb,
                                                int* __restrict__ nfax, double* __restrict__ nfar) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (s >= ns) return;
    const int o = slot_nbr[s];
    if (o < 0) return;
    const int c = static_cast<int>(s / 6), d = static_cast<int>(s % 6);
    const int64_t j = rank[s];
    const double sc = hf * ldexp(1.0, nexp[c]), so = hf * ldexp(1.0, nexp[o]);
    const double k = slot_area[s] / (0.5 * (sc + so));
    erow[2 * j] = c;
    ecol[2 * j] = o;
    ek[2 * j] = k;
    eidx[2 * j] = static_cast<int>(2 * j);
    erow[2 * j + 1] = o;
    ecol[2 * j + 1] = c;
    ek[2 * j + 1] = k;
    eidx[2 * j + 1] = static_cast<int>(2 * j + 1);
    atomicAdd(deg + c, 1);
    atomicAdd(deg + o, 1);
    nfa[j] = (d & 1) ? c : o;
    nfb[j] = (d & 1) ? o : c;
    nfax[j] = d >> 1;
    nfar[j] = slot_area[s];
}

// each row: the diagonal summed in slot order plus the outlet term, then the row sorted by
// (column, coefficient)
__global__ __launch_bounds__(kT) void k_rows(int nc, const int64_t* __restrict__ row_off, const int* __restrict__ sidx,
                                             const int* __restrict__ ecol, const double* __restrict__ ek,
                                             const double* __restrict__ cout, const int* __restrict__ nexp, double hf,
                                             int* __restrict__ col, double* __restrict__ coef, double* __restrict__ diag) {
    const int c = blockIdx.x * kT + threadIdx.x;
    if (c >= nc) return;
    const int64_t r0 = row_off[c], r1 = row_off[c + 1];
    double dg = 0.0;
    for (int64_t j = r0; j < r1; ++j) {
        const int e = sidx[j];
        dg += ek[e];
        col[j] = ecol[e];
        coef[j] = ek[e];
    }
    const double sc = hf * ldexp(1.0, nexp[c]);
    diag[c] = dg + cout[c] / (0.5 * sc);
    for (int64_t j = r0 + 1; j < r1; ++j) {
        const int cj = col[j];
        const double kj = coef[j];
        int64_t i = j;
        while (i > r0 && (col[i - 1] > cj || (col[i - 1] == cj && coef[i - 1] > kj))) {
            col[i] = col[i - 1];
            coef[i] = coef[i - 1];
            --i;
        }
        col[i] = cj;
        coef[i] = kj;
    }
}

__global__ void k_set(int* p, int v) { *p = v; }

__global__ __launch_bounds__(kT) void k_to_int(int64_t n, const int64_t* __restrict__ in, int* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) out[i] = static_cast<int>(in[i]);
}

// ---- host orchestration --------------------------------------------------------------------
class Builder {
public:
    Builder(hipStream_t s, hipEvent_t e0, hipEvent_t e1) : s_(s), e0_(e0), e1_(e1) {}
    double kernel_ms = 0;

    void run(const Topology& T, int levels, int coarse_max, std::vector<MGLevel>& mg) {
        const int64_t ndof = T.ndof, nb = T.nbricks;
        if (ndof > 0x7fffffff) throw BuildError("device multigrid: too many cells");
        // composite level
        Buf<uint8_t> mask, bflags;
        Buf<int> blevel, bcoord, bnbr;
        Buf<double> level_h;
        up(mask.get(ndof), T.mask.data(), ndof);
        up(blevel.get(nb), T.brick_level.data(), nb);
        up(bcoord.get(3 * nb), T.brick_coord.data(), 3 * nb);
        up(bnbr.get(6 * nb), T.brick_nbr.data(), 6 * nb);
        up(bflags.get(nb), T.brick_flags.data(), nb);
        up(level_h.get(T.level_h.size()), T.level_h.data(), T.level_h.size());
        int64_t* fcount = c64a_.get(ndof + 1);
        begin();
        k_l0_cells<<<grid(ndof), kT, 0, s_>>>(mask.ptr(), blevel.ptr(), bcoord.ptr(), bflags.ptr(), level_h.ptr(), levels,
                                              T.outlet ? 1 : 0, ndof, exp_.get(ndof), coord_.get(3 * ndof), act_.get(ndof),
                                              outlet_.get(ndof), fcount);
        check(hipGetLastError(), "k_l0_cells");
        scan(fcount, c64b_.get(ndof + 1), ndof);
        end();
        const int64_t nstruct = read64(c64b_.ptr() + ndof);
        const int64_t ne = static_cast<int64_t>(T.edge_a.size());
        nf_ = nstruct + ne;
        begin();
        k_l0_faces<<<grid(ndof), kT, 0, s_>>>(mask.ptr(), blevel.ptr(), bnbr.ptr(), level_h.ptr(), ndof, c64b_.ptr(),
                                              fa_.get(nf_), fb_.get(nf_), fax_.get(nf_), far_.get(nf_));
        check(hipGetLastError(), "k_l0_faces");
        end();
        if (ne > 0) {
            std::vector<int> ax(T.edge_axis.begin(), T.edge_axis.end());
            up(fa_.ptr() + nstruct, T.edge_a.data(), ne);
            up(fb_.ptr() + nstruct, T.edge_b.data(), ne);
            up(fax_.ptr() + nstruct, ax.data(), ne);
            up(far_.ptr() + nstruct, T.edge_area.data(), ne);
            sync();
        }
        MGLevel L0;
        L0.n = static_cast<int>(ndof);
        L0.exponent.resize(ndof);
        down(L0.exponent.data(), exp_.ptr(), ndof);
        mg.push_back(std::move(L0));
        n_ = ndof;
        const double hf = T.finest_h;
        int m = 0;
        while (true) {
            int64_t nact = 0;
            int maxe = 0;
            stats(nact, maxe);
            if (m >= 1 && (nact <= coarse_max || m >= 10)) break;
            if (!coarsen(m, nact, maxe, hf, mg)) break;
            ++m;
        }
        mg.back().parent.assign(mg.back().n, -1);
    }

private:
    template <class T>
    void up(T* dst, const T* src, int64_t n) {
        if (n > 0) check(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s_), "hipMemcpyAsync");
    }
    template <class T>
    void down(T* dst, const T* src, int64_t n) {
        if (n > 0) check(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s_), "hipMemcpyAsync");
        sync();
    }
    void sync() { check(hipStreamSynchronize(s_), "device multigrid"); }
    int64_t read64(const int64_t* p) {
        int64_t v = 0;
        down(&v, p, 1);
        return v;
    }
    void begin() { check(hipEventRecord(e0_, s_), "hipEventRecord"); }
    void end() {
        check(hipEventRecord(e1_, s_), "hipEventRecord");
        sync();
        float ms = 0;
        check(hipEventElapsedTime(&ms, e0_, e1_), "hipEventElapsedTime");
        kernel_ms += ms;
    }
    void* temp(size_t bytes) { return tmp_.get(bytes ? bytes : 1); }
    // exclusive sum of n counts into out[0..n], out[n] = total
    void scan(const int64_t* in, int64_t* out, int64_t n) {
        size_t bytes = 0;
        check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(n + 1), s_), "scan");
        // the count array is read one past its end: give it a zero there
        check(hipMemsetAsync(const_cast<int64_t*>(in) + n, 0, sizeof(int64_t), s_), "hipMemsetAsync");
        check(hipcub::DeviceScan::ExclusiveSum(temp(bytes), bytes, in, out, static_cast<int>(n + 1), s_), "scan");
    }
    void stats(int64_t& nact, int& maxe) {
        unsigned long long* d = u64_.get(1);
        int* e = i32_.get(1);
        check(hipMemsetAsync(d, 0, sizeof(unsigned long long), s_), "hipMemsetAsync");
        check(hipMemsetAsync(e, 0, sizeof(int), s_), "hipMemsetAsync");
        begin();
        k_stats<<<grid(n_), kT, 0, s_>>>(n_, act_.ptr(), exp_.ptr(), d, e);
        check(hipGetLastError(), "k_stats");
        end();
        unsigned long long hn = 0;
        down(&hn, d, 1);
        down(&maxe, e, 1);
        nact = static_cast<int64_t>(hn);
    }

    // level m (n_ cells in exp_/coord_/act_/outlet_, nf_ faces) -> level m+1; false: stop
    bool coarsen(int m, int64_t nact, int maxe, double hf, std::vector<MGLevel>& mg) {
        const int64_t n = n_;
        // blocked parents
        int64_t* cnt = c64a_.get(n + 1);
        int64_t* pos = c64b_.get(n + 1);
        begin();
        k_blocked_count<<<grid(n), kT, 0, s_>>>(n, act_.ptr(), exp_.ptr(), maxe, cnt);
        check(hipGetLastError(), "k_blocked_count");
        scan(cnt, pos, n);
        end();
        const int64_t nbk = read64(pos + n);
        uint64_t* bk = k64a_.get(nbk + 1);
        uint64_t* bks = k64b_.get(nbk + 1);
        int64_t* nuniq = c64c_.get(1);
        check(hipMemsetAsync(nuniq, 0, sizeof(int64_t), s_), "hipMemsetAsync");
        begin();
        if (nbk > 0) {
            k_blocked_write<<<grid(n), kT, 0, s_>>>(n, act_.ptr(), exp_.ptr(), coord_.ptr(), maxe, pos, bk);
            check(hipGetLastError(), "k_blocked_write");
            size_t bytes = 0;
            check(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, bk, bks, static_cast<int>(nbk), 0, 64, s_), "sort");
            check(hipcub::DeviceRadixSort::SortKeys(temp(bytes), bytes, bk, bks, static_cast<int>(nbk), 0, 64, s_), "sort");
            bytes = 0;
            check(hipcub::DeviceSelect::Unique(nullptr, bytes, bks, bk, nuniq, static_cast<int>(nbk), s_), "unique");
            check(hipcub::DeviceSelect::Unique(temp(bytes), bytes, bks, bk, nuniq, static_cast<int>(nbk), s_), "unique");
        }
        // parent keys of the active cells in cell order, then sorted
        int64_t* acnt = c64a_.get(n + 1);
        k_act_count(n, acnt);
        scan(acnt, pos, n);
        uint64_t* pk = k64b_.get(nact + 1);
        int* pv = i32a_.get(nact + 1);
        uint8_t* grows = u8a_.get(n);
        k_parent_keys<<<grid(n), kT, 0, s_>>>(n, act_.ptr(), exp_.ptr(), coord_.ptr(), bk, nuniq, pos, pk, pv, grows);
        check(hipGetLastError(), "k_parent_keys");
        uint64_t* sk = k64c_.get(nact + 1);
        int* sv = i32b_.get(nact + 1);
        if (nact > 0) {
            size_t bytes = 0;
            check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, pk, sk, pv, sv, static_cast<int>(nact), 0, 64, s_), "sort");
            check(hipcub::DeviceRadixSort::SortPairs(temp(bytes), bytes, pk, sk, pv, sv, static_cast<int>(nact), 0, 64, s_),
                  "sort");
        }
        int64_t* head = c64a_.get(nact + 1);
        int64_t* idp1 = c64c_.get(nact + 1);
        int64_t nc = 0;
        if (nact > 0) {
            k_heads<<<grid(nact), kT, 0, s_>>>(nact, sk, head);
            check(hipGetLastError(), "k_heads");
            size_t bytes = 0;
            check(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, head, idp1, static_cast<int>(nact), s_), "scan");
            check(hipcub::DeviceScan::InclusiveSum(temp(bytes), bytes, head, idp1, static_cast<int>(nact), s_), "scan");
        }
        end();
        if (nact > 0) nc = read64(idp1 + nact - 1);
        // parents, children, coarse exponents and positions
        int* parent = i32c_.get(n);
        int* child = i32d_.get(nact + 1);
        int* child_off = i32e_.get(nc + 1);
        int* nexp = i32f_.get(nc + 1);
        int* ncoord = i32g_.get(3 * nc + 3);
        check(hipMemsetAsync(parent, 0xff, n * sizeof(int), s_), "hipMemsetAsync");
        begin();
        if (nact > 0) {
            k_coarse_cells<<<grid(nact), kT, 0, s_>>>(nact, sv, idp1, exp_.ptr(), coord_.ptr(), grows, parent, child,
                                                      child_off, nexp, ncoord);
            check(hipGetLastError(), "k_coarse_cells");
        }
        k_set<<<1, 1, 0, s_>>>(child_off + nc, static_cast<int>(nact));
        check(hipGetLastError(), "k_set");
        end();
        MGLevel& fine = mg[m];
        if (nc >= nact) {
            fine.parent.assign(n, -1);
            return false;
        }
        fine.parent.resize(n);
        down(fine.parent.data(), parent, n);
        MGLevel C;
        C.n = static_cast<int>(nc);
        C.exponent.resize(nc);
        down(C.exponent.data(), nexp, nc);
        C.child_off.resize(nc + 1);
        down(C.child_off.data(), child_off, nc + 1);
        C.child.resize(nact);
        down(C.child.data(), child, nact);
        // outlet areas and the coarse face slots
        double* cout = f64a_.get(nc);
        int64_t* flag = c64a_.get(nf_ + 1);
        int64_t* fpos = c64b_.get(nf_ + 1);
        const int64_t ns = 6 * nc;
        int* slot_nbr = i32h_.get(ns);
        double* slot_area = f64b_.get(ns);
        int* err = i32_.get(1);
        check(hipMemsetAsync(err, 0, sizeof(int), s_), "hipMemsetAsync");
        check(hipMemsetAsync(slot_nbr, 0xff, ns * sizeof(int), s_), "hipMemsetAsync");
        check(hipMemsetAsync(slot_area, 0, ns * sizeof(double), s_), "hipMemsetAsync");
        begin();
        k_coarse_outlet<<<grid(nc), kT, 0, s_>>>(static_cast<int>(nc), child_off, child, outlet_.ptr(), cout);
        k_slot_flags<<<grid(nf_), kT, 0, s_>>>(nf_, fa_.ptr(), fb_.ptr(), parent, flag);
        check(hipGetLastError(), "k_slot_flags");
        scan(flag, fpos, nf_);
        end();
        const int64_t np = read64(fpos + nf_);
        int* spk = i32a_.get(np + 1);
        int* spv = i32b_.get(np + 1);
        int* spks = i32d_.get(np + 1);
        int* spvs = i32i_.get(np + 1);
        begin();
        if (np > 0) {
            k_slot_pairs<<<grid(nf_), kT, 0, s_>>>(nf_, fa_.ptr(), fb_.ptr(), fax_.ptr(), parent, nexp, flag, fpos, spk, spv);
            check(hipGetLastError(), "k_slot_pairs");
            size_t bytes = 0;
            const int bits = bits_for(ns);
            check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, spk, spks, spv, spvs, static_cast<int>(np), 0, bits, s_),
                  "sort");
            check(hipcub::DeviceRadixSort::SortPairs(temp(bytes), bytes, spk, spks, spv, spvs, static_cast<int>(np), 0, bits,
                                                     s_),
                  "sort");
            k_slot_sum<<<grid(np), kT, 0, s_>>>(np, spks, spvs, fa_.ptr(), fb_.ptr(), fax_.ptr(), far_.ptr(), parent, nexp,
                                                slot_nbr, slot_area, err);
            check(hipGetLastError(), "k_slot_sum");
        }
        // row entries and the next level's faces in slot order
        int64_t* used = c64a_.get(ns + 1);
        int64_t* rank = c64b_.get(ns + 1);
        k_slot_used<<<grid(ns), kT, 0, s_>>>(ns, slot_nbr, used);
        scan(used, rank, ns);
        end();
        int hostErr = 0;
        down(&hostErr, err, 1);
        if (hostErr) throw BuildError("internal: inconsistent coarse face");
        const int64_t nused = read64(rank + ns);
        const int64_t nent = 2 * nused;
        int* erow = i32a_.get(nent + 1);
        int* eidx = i32b_.get(nent + 1);
        int* ecol = i32d_.get(nent + 1);
        double* ek = f64c_.get(nent + 1);
        int* deg = i32i_.get(nc + 1);
        int* nfa = i32j_.get(nused + 1);
        int* nfb = i32k_.get(nused + 1);
        int* nfax = i32l_.get(nused + 1);
        double* nfar = f64d_.get(nused + 1);
        check(hipMemsetAsync(deg, 0, (nc + 1) * sizeof(int), s_), "hipMemsetAsync");
        begin();
        k_entries<<<grid(ns), kT, 0, s_>>>(ns, slot_nbr, slot_area, nexp, hf, rank, erow, eidx, ecol, ek, deg, nfa, nfb,
                                           nfax, nfar);
        check(hipGetLastError(), "k_entries");
        int* erows = i32c_.get(nent + 1);   // parent no longer needed on the device
        int* eidxs = i32h_.get(nent + 1);   // nor the slot neighbours
        if (nent > 0) {
            size_t bytes = 0;
            const int bits = bits_for(nc);
            check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, erow, erows, eidx, eidxs, static_cast<int>(nent), 0, bits,
                                                     s_),
                  "sort");
            check(hipcub::DeviceRadixSort::SortPairs(temp(bytes), bytes, erow, erows, eidx, eidxs, static_cast<int>(nent), 0,
                                                     bits, s_),
                  "sort");
        }
        int64_t* deg64 = c64a_.get(nc + 1);
        int64_t* row_off = c64c_.get(nc + 1);
        k_int_to64(nc, deg, deg64);
        scan(deg64, row_off, nc);
        int* col = i32e_.get(nent + 1);    // child_off already downloaded
        double* coef = f64b_.get(nent + 1);  // slot areas consumed by k_entries
        double* diag = f64e_.get(nc + 1);
        k_rows<<<grid(nc), kT, 0, s_>>>(static_cast<int>(nc), row_off, eidxs, ecol, ek, cout, nexp, hf, col, coef, diag);
        check(hipGetLastError(), "k_rows");
        int* row_off32 = i32d_.get(nc + 1);  // ecol consumed by k_rows
        k_to_int<<<grid(nc + 1), kT, 0, s_>>>(nc + 1, row_off, row_off32);
        end();
        C.row_off.resize(nc + 1);
        down(C.row_off.data(), row_off32, nc + 1);
        C.col.resize(nent);
        down(C.col.data(), col, nent);
        C.coef.resize(nent);
        down(C.coef.data(), coef, nent);
        C.diag.resize(nc);
        down(C.diag.data(), diag, nc);
        mg.push_back(std::move(C));
        // the next level's state: its cells all active
        check(hipMemcpyAsync(exp_.get(nc), nexp, nc * sizeof(int), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemcpyAsync(coord_.get(3 * nc), ncoord, 3 * nc * sizeof(int), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemsetAsync(act_.get(nc), 1, nc, s_), "hipMemsetAsync");
        check(hipMemcpyAsync(outlet_.get(nc), cout, nc * sizeof(double), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemcpyAsync(fa_.get(nused), nfa, nused * sizeof(int), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemcpyAsync(fb_.get(nused), nfb, nused * sizeof(int), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemcpyAsync(fax_.get(nused), nfax, nused * sizeof(int), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        check(hipMemcpyAsync(far_.get(nused), nfar, nused * sizeof(double), hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync");
        sync();
        n_ = nc;
        nf_ = nused;
        return true;
    }
    void k_act_count(int64_t n, int64_t* cnt);
    void k_int_to64(int64_t n, const int* in, int64_t* out);

    hipStream_t s_;
    hipEvent_t e0_, e1_;
    int64_t n_ = 0, nf_ = 0;
    // the current level: cells and faces
    Buf<int> exp_, coord_;
    Buf<uint8_t> act_;
    Buf<double> outlet_;
    Buf<int> fa_, fb_, fax_;
    Buf<double> far_;
    // scratch
    Buf<int64_t> c64a_, c64b_, c64c_;
    Buf<uint64_t> k64a_, k64b_, k64c_;
    Buf<int> i32_, i32a_, i32b_, i32c_, i32d_, i32e_, i32f_, i32g_, i32h_, i32i_, i32j_, i32k_, i32l_;
    Buf<uint8_t> u8a_;
    Buf<double> f64a_, f64b_, f64c_, f64d_, f64e_;
    Buf<unsigned long long> u64_;
    Buf<char> tmp_;
};

__global__ __launch_bounds__(kT) void k_act_to64(int64_t n, const uint8_t* __restrict__ act, int64_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) out[i] = act[i] ? 1 : 0;
}
__global__ __launch_bounds__(kT) void k_i32_to64(int64_t n, const int* __restrict__ in, int64_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) out[i] = in[i];
}
void Builder::k_act_count(int64_t n, int64_t* cnt) {
    k_act_to64<<<grid(n), kT, 0, s_>>>(n, act_.ptr(), cnt);
    check(hipGetLastError(), "k_act_to64");
}
void Builder::k_int_to64(int64_t n, const int* in, int64_t* out) {
    k_i32_to64<<<grid(n), kT, 0, s_>>>(n, in, out);
    check(hipGetLastError(), "k_i32_to64");
}

}  // namespace

// the multigrid levels of a finished composite topology on the current device (see above)
double multigrid(hipStream_t stream, hipEvent_t e0, hipEvent_t e1, const Topology& T, int levels, int coarse_max,
                 std::vector<MGLevel>& mg) {
    Builder b(stream, e0, e1);
    mg.clear();
    b.run(T, levels, coarse_max, mg);
    return b.kernel_ms;
}

}  // namespace devbuild
}  // namespace pcfd
