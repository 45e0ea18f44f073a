// Device solver for one ParaCFD step on gfx950 and the C ABI around it.
//
// Host orchestration mirrors ExternalAeroCore::initialize / step
// (src/core/fluid/external_aero_core.cpp:245-328) with the cell-momentum branch only:
// CFL dt -> MUSCL transport -> diffusion -> face reconstruction -> projection (RHS, PCG
// with the geometric V-cycle, face correction, divergence check) -> pressure impulse.
// All work is enqueued on one HIP stream owned by the handle; the host reads the device
// scalars once per `check_interval` Krylov iterations and at the end of each stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/paracfd.h"
#include "device.hpp"
#include "kernels_solve.hpp"
#include "kernels_step.hpp"
#include "learned.hpp"
#include "kernels_nonorth.hpp"
#include "lsq3.hpp"
#include "topology.hpp"

const pcfd::Topology& pcfd_topology_ref(const pcfd_topology* t);

namespace pcfd {

void set_error(const std::string& m);

// device allocation refused: out of device memory, or beyond the caller's budget
struct OutOfMemory : DeviceError {
    using DeviceError::DeviceError;
};

#define HIP_CHECK(expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define LAUNCH(kernel, grid, block, ...)                                                                  \
    do {                                                                                                  \
        hipLaunchKernelGGL((kernel), dim3(grid), dim3(block), 0, stream_, __VA_ARGS__);                   \
        HIP_CHECK(hipGetLastError());                                                                     \
    } while (0)

static inline unsigned blocks_for(int64_t n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

// Fixed device staging buffer of pcfd_sample_slice_static, one copy per GPU for the life of the
// process (4 MiB: images of at most 2^20 pixels).  In host code `scratch` names a host-side
// stand-in, not a device address: kernels and copies take the address that
// hipGetSymbolAddress returns on each GPU (scratch_address).  Every call on one GPU shares its
// copy, so calls that overlap on one GPU race.
constexpr int64_t kScratchFloats = int64_t(1) << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

// device address of `scratch` on each GPU, looked up once per device with that device current
std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_address;

void* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0) throw DeviceError("scratch: invalid device");
    if (static_cast<size_t>(device) >= g_scratch_address.size()) g_scratch_address.resize(device + 1, nullptr);
    if (!g_scratch_address[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_address[device] = p;
    }
    return g_scratch_address[device];
}

class Solver {
public:
    virtual ~Solver() = default;
    virtual void initialize(pcfd_step_stats* st) = 0;
    virtual void step(pcfd_step_stats* st) = 0;
    virtual void advect(double dt, double* dt_used, pcfd_step_stats* st) = 0;
    virtual void project(double dt, bool warm, pcfd_step_stats* st) = 0;
    virtual void impulse(double dt) = 0;
    virtual void loads(double pref, pcfd_loads* out) = 0;
    virtual void get_velocity(double* x, double* y, double* z) = 0;
    virtual void set_velocity(const double* x, const double* y, const double* z) = 0;
    virtual void get_pressure(double* p) = 0;
    virtual void set_pressure(const double* p) = 0;
    virtual void get_faces(double* fx, double* fy, double* fz, double* fe) = 0;
    virtual void set_faces(const double* fx, const double* fy, const double* fz, const double* fe) = 0;
    virtual void mg_primitive(int op, int level, const double* x, const double* b, double* out, double omega, int sweeps) = 0;
    virtual void time_kernel(int kernel, int reps, double* avg_ms, double* bytes) = 0;
    virtual void learned_collect(int max_pairs, pcfd_learned_stats* st) = 0;
    virtual void learned_train(int steps, int batch, double lr, pcfd_learned_stats* st) = 0;
    virtual void learned_weights(float* w, float* b, const float* wset, const float* bset) = 0;
    virtual void learned_enable(bool on) = 0;
    virtual void learned_forward(const double* r, float* y, float* s) = 0;
    virtual void learned_gradient(const double* r, double* loss, float* gw, float* gb) = 0;
    // batches: instance -1 sets every instance
    virtual void set_reference(int instance, double area, const double* origin) = 0;
    virtual void instance_stats(int instance, pcfd_step_stats* st) const = 0;
    virtual void instance_loads(int instance, double pref, pcfd_loads* out) = 0;
    virtual void batch_loads(double pref, pcfd_loads* out) = 0;  // every instance, one sync
    virtual void autotune(int repetitions, pcfd_tuning* out) = 0;
    virtual void get_tuning(pcfd_tuning* out) const = 0;
    virtual void set_tuning(const pcfd_tuning& t) = 0;
    virtual int instances() const = 0;
    virtual void sample_slice(int instance, int field, int axis, double coord, const double* lo, const double* hi,
                              int nu, int nv, float* out) = 0;
    virtual void sample_slice_static(int instance, int field, int axis, double coord, const double* lo,
                                     const double* hi, int nu, int nv, float* out) = 0;
    // checkpoint resume: the state was uploaded; every instance continues from physical time t
    virtual void mark_initialized(double t) = 0;
    virtual void probe_start(unsigned mask) = 0;
    virtual void probe_stop(int64_t* launches, double* total_ms) = 0;
    bool initialized = false;
    double time = 0.0;
    int64_t device_bytes = 0;
    // The last step left the next step's outflow rate, timestep and |v|max in the device
    // scalars; any call that may change the velocities or faces clears it.
    bool next_rate_ready = false;
};

template <class R>
class SolverImpl final : public Solver {
public:
    SolverImpl(const Topology& T, const pcfd_options& o);
    ~SolverImpl() override;
    void construct(const Topology& T);
    void release();
    void initialize(pcfd_step_stats* st) override;
    void step(pcfd_step_stats* st) override;
    void advect(double dt, double* dt_used, pcfd_step_stats* st) override;
    void project(double dt, bool warm, pcfd_step_stats* st) override;
    void impulse(double dt) override;
    void loads(double pref, pcfd_loads* out) override;
    void get_velocity(double* x, double* y, double* z) override;
    void set_velocity(const double* x, const double* y, const double* z) override;
    void get_pressure(double* p) override;
    void set_pressure(const double* p) override;
    void get_faces(double* fx, double* fy, double* fz, double* fe) override;
    void set_faces(const double* fx, const double* fy, const double* fz, const double* fe) override;
    void mg_primitive(int op, int level, const double* x, const double* b, double* out, double omega, int sweeps) override;
    void time_kernel(int kernel, int reps, double* avg_ms, double* bytes) override;
    void learned_collect(int max_pairs, pcfd_learned_stats* st) override;
    void learned_train(int steps, int batch, double lr, pcfd_learned_stats* st) override;
    void learned_weights(float* w, float* b, const float* wset, const float* bset) override;
    void learned_enable(bool on) override;
    void learned_forward(const double* r, float* y, float* s) override;
    void learned_gradient(const double* r, double* loss, float* gw, float* gb) override;
    void set_reference(int instance, double area, const double* origin) override {
        for (int i = 0; i < ninst_; ++i) {
            if (instance >= 0 && i != instance) continue;
            inst_ref_area_[i] = area;
            if (origin) for (int c = 0; c < 3; ++c) inst_origin_[3 * static_cast<size_t>(i) + c] = origin[c];
        }
    }
    void instance_stats(int instance, pcfd_step_stats* st) const override { *st = inst_stats_.at(instance); }
    void mark_initialized(double t) override {
        initialized = true;
        time = t;
        std::fill(inst_time_.begin(), inst_time_.end(), t);
    }
    void instance_loads(int instance, double pref, pcfd_loads* out) override { loads_instance(instance, pref, out); }
    void batch_loads(double pref, pcfd_loads* out) override;
    void autotune(int repetitions, pcfd_tuning* out) override;
    void get_tuning(pcfd_tuning* out) const override { *out = tuning_; }
    void set_tuning(const pcfd_tuning& t) override;
    int instances() const override { return ninst_; }
    void sample_slice(int instance, int field, int axis, double coord, const double* lo, const double* hi, int nu,
                      int nv, float* out) override;
    void sample_slice_static(int instance, int field, int axis, double coord, const double* lo, const double* hi,
                             int nu, int nv, float* out) override;

private:
    void build_slice_map(int instance);
    void check_slice(int instance, int field, int axis, double coord, const double* lo, const double* hi, int nu,
                     int nv, int64_t max_pixels) const;
    void slice_into(int instance, int field, int axis, double coord, const double* lo, const double* hi, int nu,
                    int nv, float* image, float* out);
    void* scratch_ = nullptr;   // this GPU's address of `scratch`
    template <class T>
    T* alloc(int64_t count) {
        if (count <= 0) count = 1;
        const size_t bytes = static_cast<size_t>(count) * sizeof(T);
        if (opt_.memory_budget_bytes > 0.0 && static_cast<double>(device_bytes) + static_cast<double>(bytes) > opt_.memory_budget_bytes)
            throw OutOfMemory("device memory of " + std::to_string(device_bytes + static_cast<int64_t>(bytes)) +
                              " bytes exceeds the budget of " + std::to_string(static_cast<int64_t>(opt_.memory_budget_bytes)) +
                              " bytes");
        void* p = nullptr;
        const hipError_t err = hipMalloc(&p, bytes);
        if (err == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            throw OutOfMemory("hipMalloc of " + std::to_string(bytes) + " bytes: out of device memory");
        }
        HIP_CHECK(err);
        allocations_.push_back(p);
        device_bytes += count * static_cast<int64_t>(sizeof(T));
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(static_cast<int64_t>(v.size()));
        if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    template <class T>
    T* upload_as(const std::vector<double>& v) {
        std::vector<T> c(v.begin(), v.end());
        return upload(c);
    }
    // instance scalars S_[0..ninst_) and the batch aggregate agg_ = S_ + ninst_
    void sync_scalars() {
        HIP_CHECK(hipMemcpyAsync(host_s_, S_, (ninst_ + 1) * sizeof(DevScalars), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }
    const DevScalars& hagg() const { return host_s_[ninst_]; }
    void put_dt(double dt) {
        for (int i = 0; i < ninst_; ++i) host_s_[i].dt = dt;
        LAUNCH(k_set_dt, 1, 64, S_, ninst_, dt);
    }
    FieldsView<R> fields() const {
        FieldsView<R> F;
        for (int c = 0; c < 3; ++c) { F.v[c] = v_[c]; F.f[c] = f_[c]; }
        F.fe = fe_;
        return F;
    }
    // In-step launch timing (pcfd_probe_start / pcfd_probe_stop): while kernel `id` is in
    // probe_mask_, every launch of it is bracketed by a pair of HIP events on the solver's
    // stream; the pairs are read back when the pool is full and when probing stops.
    static constexpr int kProbeKernels = 8, kProbePool = 512;
    unsigned probe_mask_ = 0;
    bool probe_saved_graph_ = true;
    std::vector<hipEvent_t> probe_ev_;
    std::vector<int> probe_ev_id_;   // kernel of each event pair
    int probe_used_ = 0;             // pairs recorded since the last harvest
    int64_t probe_launches_[kProbeKernels] = {};
    double probe_ms_[kProbeKernels] = {};
    void probe_harvest() {
        if (!probe_used_) return;
        HIP_CHECK(hipEventSynchronize(probe_ev_[2 * probe_used_ - 1]));
        for (int k = 0; k < probe_used_; ++k) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, probe_ev_[2 * k], probe_ev_[2 * k + 1]));
            probe_ms_[probe_ev_id_[k]] += ms;
            ++probe_launches_[probe_ev_id_[k]];
        }
        probe_used_ = 0;
    }
    template <class F>
    void probed(int id, F&& launch) {
        if (!(probe_mask_ & (1u << id))) {
            launch();
            return;
        }
        if (probe_used_ == kProbePool) probe_harvest();
        HIP_CHECK(hipEventRecord(probe_ev_[2 * probe_used_], stream_));
        launch();
        HIP_CHECK(hipEventRecord(probe_ev_[2 * probe_used_ + 1], stream_));
        probe_ev_id_[probe_used_++] = id;
    }
public:
    void probe_start(unsigned mask) override {
        if (mask == 0 || (mask & ~((1u << 1) | (1u << 2) | (1u << 5) | (1u << 6))))
            throw std::invalid_argument("probe mask: kernels 1, 2, 5 and 6 can be probed");
        if (probe_mask_) throw std::invalid_argument("a probe is already running");
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (probe_ev_.empty()) {
            probe_ev_.assign(2 * kProbePool, nullptr);
            probe_ev_id_.assign(kProbePool, 0);
            for (auto& e : probe_ev_) HIP_CHECK(hipEventCreate(&e));
        }
        for (int k = 0; k < kProbeKernels; ++k) { probe_launches_[k] = 0; probe_ms_[k] = 0.0; }
        probe_used_ = 0;
        probe_saved_graph_ = use_graph_;
        use_graph_ = false;  // the iterations are launched one by one while probing
        probe_mask_ = mask;
    }
    void probe_stop(int64_t* launches, double* total_ms) override {
        if (!probe_mask_) throw std::invalid_argument("no probe is running");
        probe_harvest();
        probe_mask_ = 0;
        use_graph_ = probe_saved_graph_;
        for (int k = 0; k < kProbeKernels; ++k) {
            if (launches) launches[k] = probe_launches_[k];
            if (total_ms) total_ms[k] = probe_ms_[k];
        }
    }
private:
    // One PCG iteration: d_new = z + beta d_old with A d_new fused, alpha, x/r update,
    // preconditioner, beta.  With the fused V-cycle the x/r update runs inside its first pass.
    void cg_iteration(bool initial, double* dold, double* dnew) {
        probed(5, [&] {
            if (zf()) {
                if (initial) LAUNCH((k_cg_dir_apply<float, true>), nb_, kBL, M_, zf32(), dold, dnew, Ad_, partials_, S_);
                else LAUNCH((k_cg_dir_apply<float, false>), nb_, kBL, M_, zf32(), dold, dnew, Ad_, partials_, S_);
            } else {
                if (initial) LAUNCH((k_cg_dir_apply<double, true>), nb_, kBL, M_, z_, dold, dnew, Ad_, partials_, S_);
                else LAUNCH((k_cg_dir_apply<double, false>), nb_, kBL, M_, z_, dold, dnew, Ad_, partials_, S_);
            }
        });
        finalize(FIN_APPLY, nb_);
        if (update_fused()) {
            double* rin = r_;
            r_ = r_ == rbuf_[0] ? rbuf_[1] : rbuf_[0];
            rprev_ = flexible() ? rin : nullptr;
            precondition_dispatch(false, dnew, rin);
            return;
        }
        LAUNCH(k_cg_update, nb_, kBL, M_, p_, r_, rprev_, dnew, Ad_, partials_, S_);
        finalize(FIN_UPDATE, nb_);
        precondition_dispatch(false);
    }
    bool update_fused() const {
        return opt_.preconditioner != PCFD_PC_JACOBI && opt_.pre_smooth == 2 && lv_.size() > 1 && fused_;
    }
    // residual roles at the start of a PCG solve
    void reset_residual_roles() {
        r_ = rbuf_[0];
        rprev_ = flexible() ? rbuf_[1] : nullptr;
    }
    // Iteration j of the loop after the first one reads d from dbuf[(j+1)&1] and writes
    // dbuf[j&1]; with the fused update it reads r from rbuf[(j+1)&1] and writes rbuf[j&1]
    // (the previous r is r_prev).  The roles are set from j, so any run of iterations can be
    // enqueued or captured on its own.
    void set_roles_for(int j) {
        if (update_fused()) {
            r_ = rbuf_[(j + 1) & 1];
            rprev_ = flexible() ? rbuf_[j & 1] : nullptr;
        } else {
            reset_residual_roles();
        }
    }
    void enqueue_iterations(int count, int j0) {
        for (int k = 0; k < count; ++k) {
            const int j = j0 + k;
            set_roles_for(j);
            cg_iteration(false, dbuf_[(j + 1) & 1], dbuf_[j & 1]);
        }
    }
    // captured graphs of one or `chunk_` iterations starting at an even or odd j
    hipGraphExec_t iteration_graph(int count, int parity) {
        const int kind = count == 1 ? 0 : 1;
        if (!gexec_[kind][parity]) {
            HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_iterations(count, parity);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(stream_, &graphs_[kind][parity]));
            HIP_CHECK(hipGraphInstantiate(&gexec_[kind][parity], graphs_[kind][parity], nullptr, nullptr, 0));
        }
        return gexec_[kind][parity];
    }
    // reduction over brick-indexed partials, one finalize workgroup per instance
    void finalize(int op, int nparts, const double* parts = nullptr) {
        LAUNCH(k_finalize, ninst_, 1024, op, parts ? parts : partials_, nparts, ninst_ > 1 ? seg_ : nullptr, S_, agg_,
               opt_.cfl, std::max(1e-9, std::abs(opt_.speed) / M_.finest_h), opt_.rtol, opt_.max_iterations,
               rprev_ ? 1 : 0);
    }
    // dense coarse solves of the instances whose coarsest multigrid level is m
    template <class VT, class CT>
    bool dense_blocks(int m, VT* B, VT* X) {
        const DenseLevel& D = dense_[m];
        if (D.count == 0) return false;
        const CT* inv = nullptr;
        if constexpr (std::is_same<CT, double>::value) inv = dense_inv64_; else inv = dense_inv32_;
        hipLaunchKernelGGL((k_dense_solve<VT, CT>), dim3(D.max_n, D.count), dim3(256), 0, stream_, D.off, D.n, D.ioff,
                           inv, B, X);
        HIP_CHECK(hipGetLastError());
        return true;
    }
    // learned preconditioner
    NNView nnv() const { return NNView{nn_w_, nn_wt_, nn_p_ + kNNW}; }
    NNSetView nn_setv() const { return NNSetView{nn_list_, nn_nbtab_, nn_masks_}; }
    void set_nn_off(bool off) { if (nn_) HIP_CHECK(hipMemsetAsync(nn_off_, off ? 1 : 0, sizeof(int), stream_)); }
    unsigned nn_grid() const { return static_cast<unsigned>(nn_list_ ? nn_nlist_ : nb_); }
    // network correction after the V-cycle's last kernel (post-V-cycle residual, network,
    // z += dz), then the r.z finalisation; a network that is off (untrained, switched off or
    // rejected by validation) is not launched at all
    void pc_tail(int fin) {
        if (nn_ && nn_on_) {
            LAUNCH(k_nn_resid<double>, nn_grid(), 512, M_, nn_list_, r_, z_, nn_rstore_, S_, nn_off_);
            LAUNCH(k_nn_apply, nn_grid(), 512, M_, nnv(), nn_setv(), nn_rstore_, r_, nn_dz_, rprev_, partials_, nb_, S_, nn_off_);
            LAUNCH(k_nn_add, nn_grid(), 512, nn_list_, nn_dz_, z_, S_, nn_off_);
        }
        finalize(fin, nb_);
    }
    // sum over the set's active cells of v^2 / h_b of a residual store (its normalisation)
    double nn_norm(const float* store) {
        const unsigned g = nn_grid();
        LAUNCH(k_nn_norm, g, 512, M_, nn_list_, store, partials_);
        std::vector<double> h(g);
        HIP_CHECK(hipMemcpyAsync(h.data(), partials_, g * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        double s = 0.0;
        for (double x : h) s += x;
        return s;
    }
    // one training pass (forward, energy gradient, backward) on the residual store `rp` into
    // nn_g_ / nn_loss_
    void nn_train_pass(const float* rp, double inv_norm) {
        LAUNCH(k_nn_train_fwd, nn_grid(), 512, M_, nnv(), nn_setv(), rp, nn_scratch_, nn_dfull_, nn_sb_);
        LAUNCH(k_nn_energy, nn_grid(), 512, M_, nn_list_, nn_dfull_, rp, nn_gy_, nn_loss_, inv_norm);
        LAUNCH(k_nn_train_bwd, nn_grid(), 512, M_, nnv(), nn_setv(), rp, nn_gy_, nn_sb_, nn_scratch_, nn_g_);
    }
    void setup_learned(const Topology& T);
    void setup_gradients(const Topology& T);
    void setup_smooth_wall(const Topology& T);
    // non-orthogonal operator and its flexible GMRES outer solve
    void setup_nonorth(const Topology& T);
    void full_residual(double* r, const double* x = nullptr);
    void apply_full(const double* x, double* y);
    double partial_norm2();
    bool recycle_project(double* w, double* z, double zsign, int passes = 2);
    void mdot(const double* Vb, int nv, const double* w, double* host_out);
    void fgmres_solve();
    void fgmres_direction(int j);
    void enqueue_fgmres_direction(int j);
    bool nonorth_ = false;
    std::vector<hipGraph_t> gm_graph_;
    std::vector<hipGraphExec_t> gm_exec_;
    NonOrthView NV_{};
    double *gm_V_ = nullptr, *gm_Z_ = nullptr, *gm_part_ = nullptr, *gm_dev_ = nullptr;
    int gm_m_ = 0, gm_its_ = 0;
    // GCRO recycling: up to gm_cap_ realised corrections U and their orthonormal images C = A U
    // (contiguous bases), kept across cycles and projections
    double *gm_U_ = nullptr, *gm_C_ = nullptr, *gm_coef_ = nullptr;
    int gm_cap_ = 0, gm_k_ = 0;
    int64_t gm_rejections_ = 0;
    std::vector<double> gm_a_;
    // Validation of a freshly trained network: the next 2 kTrial projections alternate learned
    // and plain MGPCG (learned first), so both see the same stage of the flow; if the learned
    // ones take longer in total, the network is switched off.
    static constexpr int kTrial = 3;
    bool trial_plain() const { return trial_left_ > 0 && (trial_left_ & 1); }
    void validate_learned(bool used, float project_ms) {
        if (!nn_ || trial_left_ <= 0) return;
        (used ? trial_ms_ : trial_plain_ms_) += project_ms;
        if (--trial_left_ > 0) return;
        if (trial_ms_ > trial_plain_ms_) {
            set_network(false);
            nn_rejected_ = true;
            set_nn_off(true);
        }
    }
    // the network's state; the captured iteration graphs hold its launches only while it is on
    void set_network(bool on) {
        if (on != nn_on_) {
            HIP_CHECK(hipStreamSynchronize(stream_));
            drop_graph();
        }
        nn_on_ = on;
    }
    // PCG from the current p_; with `residual_ready` the initial residual is already in r_ (and
    // FIN_RESID done), else it is formed from rhs_
    bool pcg_loop(bool watch_stall, bool residual_ready = false);
    void aggregate_stats(pcfd_step_stats* st) const;
    void advect_core();
    // the redistribution groups' pooled increments into the current velocity (after a kernel
    // that handed them over)
    void redistribute(double* vmax_part = nullptr) {
        if (M_.ngroups) LAUNCH(k_redistribute<R>, blocks_for(M_.ngroups, 256), 256, M_, v_[0], v_[1], v_[2], vmax_part);
    }
    void loads_instance(int i, double pref, pcfd_loads* out);
    void enqueue_loads(int i, double pref);
    void read_loads(int i, pcfd_loads* out) const;
    void reconstruct();
    void run_projection(bool warm, pcfd_step_stats* st);
    template <class VT> void precondition(bool initial, const double* upd_d, const double* upd_rin);
    template <class VT> VT* coarse_cycle(int m);
    template <class VT, class CT> VT* coarse_cycle_t(int m);
    // upd_d != nullptr: the CG update x += alpha d, r = upd_rin - alpha Ad (into r_) runs in the
    // first V-cycle pass, followed by the FIN_UPDATE reduction
    void precondition_dispatch(bool initial, const double* upd_d = nullptr, const double* upd_rin = nullptr);

    pcfd_options opt_;
    MeshView M_{};
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;
    int64_t n_ = 0;
    int nb_ = 0;
    int nedges_ = 0;
    double max_diff_coef_ = 0;
    // state
    R* v_[3] = {};
    R* vt_[3] = {};
    R* f_[3] = {};
    R* fe_ = nullptr;
    double* p_ = nullptr;
    double* rhs_ = nullptr;
    double* r_ = nullptr;
    double* z_ = nullptr;
    double* d_ = nullptr;
    double* Ad_ = nullptr;
    double* rprev_ = nullptr;
    // The residual lives in two buffers: with the CG update fused into the first V-cycle pass
    // each iteration reads r from one and writes the new r to the other (no in-place update
    // that a neighbour brick's halo read could see), and for flexible CG the old one is r_prev.
    // r_ / rprev_ name the current roles at enqueue time; every PCG solve starts from rbuf_[0].
    double* rbuf_[2] = {};
    // Flexible (Polak-Ribiere) CG while the preconditioner varies from call to call: BF16 storage
    // (rounding of order 1e-3 relative) or the network on.  The FP32 V-cycle is the FP64 one with
    // rounding of order 1e-7, so Fletcher-Reeves keeps its iteration counts and saves the r_prev
    // traffic.  The captured graphs are dropped whenever the network is switched.
    bool flexible() const { return opt_.vcycle_precision == PCFD_BF16 || (nn_ && nn_on_); }
    // z = M^-1 r is stored in FP32 where the V-cycle's last sweep computes it in FP32 (FP32 or
    // BF16 storage): the values are the same, 4 bytes fewer per cell to write and to read.  FP64
    // with the network on (z + dz), with the FP64 V-cycle and for the FGMRES basis.
    bool zf() const { return opt_.vcycle_precision != PCFD_FP64 && !(nn_ && nn_on_) && !nonorth_; }
    float* zf32() const { return reinterpret_cast<float*>(z_); }
    // f(z) with z_ as the z type of a V-cycle in VT
    template <class VT, class F>
    void with_z(F&& f) {
        if constexpr (!std::is_same<VT, double>::value) {
            if (zf()) {
                f(zf32());
                return;
            }
        }
        f(z_);
    }
    void* vc0_[2] = {};  // composite-level V-cycle buffers (8 bytes per dof)
    double* partials_ = nullptr;
    DevScalars* S_ = nullptr;     // ninst_ instance records, then the batch aggregate
    DevScalars* agg_ = nullptr;
    int ninst_ = 1;
    int* seg_ = nullptr;          // brick offsets of the instances (partial segments)
    std::vector<int> inst_brick_off_, inst_patch_off_;
    std::vector<double> inst_maxdiff_, inst_time_, inst_ref_area_, inst_origin_;
    std::vector<pcfd_step_stats> inst_stats_;
    double* inst_maxdiff_dev_ = nullptr;
    struct DenseLevel { int count = 0, max_n = 0; const int* off = nullptr; const int* n = nullptr; const int64_t* ioff = nullptr; };
    std::vector<DenseLevel> dense_;
    double* dense_inv64_ = nullptr;
    float* dense_inv32_ = nullptr;
    DevScalars* host_s_ = nullptr;
    // multigrid levels >= 1
    struct Level {
        int n = 0;
        CsrView csr64{}, csr32{};
        const int* parent = nullptr;  // parent of this level's cells
        void* B = nullptr; void* X = nullptr; void* T = nullptr;
        // the residual of this level and its restriction to the next run as one kernel
        // (each group of kRestrictGroup next-level rows has its children in one window of at
        // most 256 rows of this level)
        bool fused_restrict = false;
        bool restrict_rows8 = false;  // every coarse row of the next level has at most 8 children
    };
    std::vector<Level> lv_;
    const int* parent0_ = nullptr;
    hipEvent_t ev_[4] = {};
    double* dbuf_[2] = {};
    void* diag_ = nullptr;        // composite-level diagonal in the V-cycle arithmetic type
    int chunk_ = 2;
    bool use_graph_ = true;
    bool fused_ = true;
    int coarse_block_ = 256;   // workgroup size of the coarse-level (SELL) kernels
    pcfd_tuning tuning_{};
    bool tuning_run_ = false;  // inside autotune: steps leave no learned-validation record
    void snapshot_state();
    void restore_state();
    void drop_graph() {
        for (size_t j = 0; j < gm_exec_.size(); ++j) {
            if (gm_exec_[j]) (void)hipGraphExecDestroy(gm_exec_[j]);
            if (gm_graph_[j]) (void)hipGraphDestroy(gm_graph_[j]);
            gm_exec_[j] = nullptr;
            gm_graph_[j] = nullptr;
        }
        for (int k = 0; k < 2; ++k)
            for (int p = 0; p < 2; ++p) {
                if (gexec_[k][p]) (void)hipGraphExecDestroy(gexec_[k][p]);
                if (graphs_[k][p]) (void)hipGraphDestroy(graphs_[k][p]);
                gexec_[k][p] = nullptr;
                graphs_[k][p] = nullptr;
            }
    }
    hipGraph_t graphs_[2][2] = {};
    hipGraphExec_t gexec_[2][2] = {};
    // Smagorinsky model (smagorinsky_cs > 0)
    R* nu_ = nullptr;
    SmagView smag_{};
    R* grad_v_ = nullptr;        // velocity gradient at the special cells (9 per cell)
    R* adv_sp_ = nullptr;        // transport sums of the special cells over their structured faces (3 per cell)
    double* imp_sp_ = nullptr;   // impulse sums of the special cells over their structured faces (3 per cell)
    double* grad_p_ = nullptr;   // pressure gradient at the special cells (3 per cell)
    // pressure impulse (dt > 0: that dt, else each instance's): structured faces per brick, then
    // the compact faces, wall traction, volumes and redistribution hand-over of the special cells
    // the cells' |v| after the impulse go to vmax_part (per brick) if it is given
    void launch_impulse(double dt, int only_converged, double* vmax_part = nullptr) {
        LAUNCH(k_impulse<R>, nb_, kBL, M_, fields(), p_, S_, dt, only_converged, imp_sp_, vmax_part);
        if (M_.nspecial)
            LAUNCH(k_impulse_compact<R>, blocks_for(M_.nspecial, 256), 256, M_, fields(), p_,
                   static_cast<const double*>(grad_p_), S_, dt, only_converged, static_cast<const double*>(imp_sp_),
                   vmax_part);
    }
    // per brick: the outflow rate of the final faces (k_divcheck), then the |v|max of the
    // velocities after the impulse
    double* rv_part_ = nullptr;
    // MUSCL transport of the velocity into vt_: structured faces per brick, then the compact faces,
    // volumes and redistribution hand-over of the special cells
    void launch_advect() {
        LAUNCH(k_advect<R>, nb_, 512, M_, fields(), vt_[0], vt_[1], vt_[2], S_, adv_sp_);
        if (M_.nspecial)
            LAUNCH(k_advect_compact<R>, blocks_for(M_.nspecial, 256), 256, M_, fields(), vt_[0], vt_[1], vt_[2], S_,
                   static_cast<const R*>(grad_v_), static_cast<const R*>(adv_sp_));
    }
    void grad_v() {
        if (M_.nspecial) LAUNCH(k_grad_v<R>, blocks_for(M_.nspecial, 256), 256, M_, fields(), smag_, grad_v_);
    }
    void grad_p() {
        if (M_.nspecial) LAUNCH(k_grad_p, blocks_for(M_.nspecial, 256), 256, M_, smag_, p_, grad_p_);
        gp_gen_ = p_gen_;
    }
    // generation of the pressure: raised wherever p may change (solves, uploads, restores), so
    // that the load integral reuses the pressure gradients of the step's impulse
    uint64_t p_gen_ = 1, gp_gen_ = 0;
    void p_changed() { ++p_gen_; }
    // learned preconditioner state (PCFD_PC_LEARNED)
    bool nn_ = false;             // created with the learned preconditioner
    bool nn_on_ = false;          // network enabled (trained or weights set)
    uint16_t* nn_w_ = nullptr;    // BF16 weights [L][27][cout][cin]
    uint16_t* nn_wt_ = nullptr;   // BF16 flipped/transposed copy
    float* nn_p_ = nullptr;       // FP32 master [weights | biases]
    float* nn_m_ = nullptr;       // Adam moments
    float* nn_v_ = nullptr;
    float* nn_g_ = nullptr;       // gradient accumulator
    int* nn_off_ = nullptr;       // device flag read by k_nn_apply
    double* nn_loss_ = nullptr;
    uint16_t* nn_scratch_ = nullptr;  // layer inputs per listed brick (training)
    int* nn_list_ = nullptr;      // the network's bricks (null: all bricks, XCD-aware order)
    int* nn_nbtab_ = nullptr;     // 27 per position: the neighbour's position in the set or -1
    const uint8_t* nn_masks_ = nullptr;  // cell masks per position (the mesh's masks without a list)
    float* nn_rstore_ = nullptr;  // post-V-cycle residual on the set's bricks (FP32, 512 per position)
    std::vector<int> nn_host_nbr_;    // brick_nbr (the all-brick set of the forward hook)
    int* nn_all_nbtab_ = nullptr;     // forward hook: neighbour table of the all-brick set
    float* nn_all_store_ = nullptr;   // forward hook: its input and output
    float* nn_all_y_ = nullptr;
    int nn_nlist_ = 0;
    int64_t nn_cells_ = 0;        // active cells of the network's bricks
    double* nn_dz_ = nullptr;     // correction per listed brick (512 per position)
    double* nn_dfull_ = nullptr;  // training: correction as a full-length vector
    double* nn_gy_ = nullptr;     // training: output gradient per listed brick
    float* nn_sb_ = nullptr;      // training: brick scales
    std::vector<float*> nn_rs_;   // recorded post-V-cycle residuals (stores of the set)
    std::vector<double> nn_inv_norm_;
    int nn_cap_ = 0, nn_count_ = 0, nn_next_ = 0, nn_cursor_ = 0, nn_t_ = 0;
    int trial_left_ = 0;            // projections of the validation still to run
    double trial_ms_ = 0.0;         // their learned and plain projection times
    double trial_plain_ms_ = 0.0;
    bool nn_rejected_ = false;
    R* snap_v_[3] = {};
    R* snap_f_[3] = {};
    R* snap_fe_ = nullptr;
    double* snap_p_ = nullptr;
    double* snap_wall_force_ = nullptr;
    // smooth-wall model: wall cells (cells holding surface patches), their patches (CSR over a
    // patch permutation), volumes, per-patch wall distance; the coefficient and force of every
    // patch from the last application (force zero until the model runs)
    int nwall_ = 0;
    int* wall_cell_ = nullptr;
    int* wall_off_ = nullptr;
    int* wall_patch_ = nullptr;
    double* wall_vol_ = nullptr;
    double* wall_dist_ = nullptr;
    double* wall_k_ = nullptr;
    double* wall_force_ = nullptr;
    std::vector<uint8_t> wall_applied_, snap_wall_applied_;  // per instance
    // slice sampling: brick maps per instance (built on first use) and the image buffer
    std::vector<int> brick_level_h_, brick_coord_h_;
    int base_bricks_[3] = {0, 0, 0};
    double grid_lo_[3] = {0, 0, 0};
    std::vector<double> level_h_h_;
    std::vector<int*> slice_map_;
    std::vector<SliceMap> slice_geo_;
    float* slice_buf_ = nullptr;
    int64_t slice_cap_ = 0;
};

// A constructor that fails part way releases what it has created before the error leaves it.
template <class R>
SolverImpl<R>::SolverImpl(const Topology& T, const pcfd_options& o) : opt_(o) {
    try {
        construct(T);
    } catch (...) {
        release();
        throw;
    }
}

template <class R>
void SolverImpl<R>::construct(const Topology& T) {
    if (opt_.pre_smooth < 1 || opt_.post_smooth < 1) throw std::invalid_argument("V-cycle needs at least one pre- and post-smoothing sweep");
    if (opt_.check_interval < 1) opt_.check_interval = 1;
    if (!(opt_.rho > 0) || !(opt_.cfl > 0) || !(opt_.rtol > 0) || opt_.max_iterations < 1 || opt_.nu < 0)
        throw std::invalid_argument("invalid solver options");
    if (!(opt_.smoother_weight > 0.0) || !(opt_.smoother_weight2 > 0.0))
        throw std::invalid_argument("smoother weights must be positive");
    {
        // the pair of sweeps must damp every mode of D^-1 A (eigenvalues in (0, 2]):
        // |(1 - w1 l)(1 - w2 l)| < 1 on (0, 2]
        double worst = 0.0;
        for (int k = 1; k <= 2000; ++k) {
            const double l = 2.0 * k / 2000.0;
            worst = std::max(worst, std::abs((1.0 - opt_.smoother_weight * l) * (1.0 - opt_.smoother_weight2 * l)));
        }
        if (!(worst < 1.0)) throw std::invalid_argument("smoother weights do not damp every mode of D^-1 A");
    }
    if (!(opt_.smagorinsky_cs >= 0.0) || !std::isfinite(opt_.smagorinsky_cs))
        throw std::invalid_argument("smagorinsky_cs must be a non-negative number");
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto& e : ev_) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_s_), (T.instances() + 1) * sizeof(DevScalars), hipHostMallocDefault));
    std::memset(host_s_, 0, (T.instances() + 1) * sizeof(DevScalars));
    n_ = T.ndof;
    nb_ = T.nbricks;
    nedges_ = static_cast<int>(T.edge_a.size());
    max_diff_coef_ = T.max_diffusion_coefficient;
    // topology
    M_.nbricks = nb_;
    M_.ndof = n_;
    M_.mask = upload(T.mask);
    M_.brick_level = upload(T.brick_level);
    M_.brick_flags = upload(T.brick_flags);
    M_.brick_nbr = upload(T.brick_nbr);
    for (int l = 0; l < 4; ++l) M_.h[l] = l < static_cast<int>(T.level_h.size()) ? T.level_h[l] : 0.0;
    M_.outlet = T.outlet ? 1 : 0;
    M_.brick_special_off = upload(T.brick_special_off);
    M_.special_volume = upload(T.special_volume);
    M_.special_inc_off = upload(T.special_inc_off);
    M_.special_inc = upload(T.special_inc);
    M_.special_patch = upload(T.special_patch);
    M_.nedges = nedges_;
    M_.ea = upload(T.edge_a);
    M_.eb = upload(T.edge_b);
    M_.eaxis = upload(T.edge_axis);
    M_.earea = upload(T.edge_area);
    M_.ecoef = upload(T.edge_coef);
    M_.ew = upload(T.edge_w);
    M_.edcoef = upload(T.edge_dcoef);
    M_.nspecial = static_cast<int>(T.special_dof.size());
    M_.special_dof = upload(T.special_dof);
    M_.npatches = static_cast<int>(T.patch_dof.size());
    M_.patch_dof = upload(T.patch_dof);
    M_.patch_S = upload(T.patch_S);
    M_.patch_x = upload(T.patch_x);
    M_.patch_D = upload(T.patch_D);
    M_.ngroups = static_cast<int>(T.redist_volume.size());
    if (M_.ngroups > 0) {
        M_.rgroup = upload(T.redist_group);
        M_.roff = upload(T.redist_off);
        M_.rmember = upload(T.redist_member);
        M_.rvolume = upload(T.redist_volume);
        M_.rdelta = alloc<double>(3 * static_cast<int64_t>(T.special_dof.size()));
    }
    {
        // per compact edge: special ranks of its cells and the face centroid's offsets from
        // their centroids; per patch: the special rank of its cell
        std::vector<int> special_of(T.ndof, -1);
        for (size_t sp = 0; sp < T.special_dof.size(); ++sp) special_of[T.special_dof[sp]] = static_cast<int>(sp);
        std::vector<int> esp(2 * static_cast<size_t>(nedges_));
        std::vector<double> edisp(6 * static_cast<size_t>(nedges_));
        for (int e = 0; e < nedges_; ++e) {
            const int64_t a = T.edge_a[e], b = T.edge_b[e];
            esp[2 * e] = special_of[a];
            esp[2 * e + 1] = special_of[b];
            for (int c = 0; c < 3; ++c) {
                edisp[6 * e + c] = T.edge_x[3 * e + c] - T.centroid[3 * a + c];
                edisp[6 * e + 3 + c] = T.edge_x[3 * e + c] - T.centroid[3 * b + c];
            }
        }
        M_.esp = upload(esp);
        M_.edisp = upload(edisp);
        std::vector<int> psp(T.patch_dof.size());
        for (size_t q = 0; q < psp.size(); ++q) {
            psp[q] = special_of[T.patch_dof[q]];
            if (psp[q] < 0) throw std::logic_error("a wall patch on a cell that is not special");
        }
        M_.patch_sp = upload(psp);
    }
    M_.speed = opt_.speed;
    M_.rho = opt_.rho;
    M_.cfl = opt_.cfl;
    M_.nu = opt_.nu;
    M_.rtol = opt_.rtol;
    M_.finest_h = T.finest_h;
    M_.max_iterations = opt_.max_iterations;
    // fields
    for (int c = 0; c < 3; ++c) {
        v_[c] = alloc<R>(n_); vt_[c] = alloc<R>(n_); f_[c] = alloc<R>(n_);
        HIP_CHECK(hipMemset(v_[c], 0, n_ * sizeof(R)));
        HIP_CHECK(hipMemset(f_[c], 0, n_ * sizeof(R)));
    }
    fe_ = alloc<R>(nedges_);
    HIP_CHECK(hipMemset(fe_, 0, std::max(1, nedges_) * sizeof(R)));
    p_ = alloc<double>(n_); rhs_ = alloc<double>(n_); r_ = alloc<double>(n_); z_ = alloc<double>(n_);
    d_ = alloc<double>(n_); Ad_ = alloc<double>(n_);
    for (double* q : {p_, rhs_, r_, z_, d_, Ad_}) HIP_CHECK(hipMemset(q, 0, n_ * sizeof(double)));
    dbuf_[0] = d_;
    dbuf_[1] = alloc<double>(n_);
    HIP_CHECK(hipMemset(dbuf_[1], 0, n_ * sizeof(double)));
    chunk_ = std::max(2, (opt_.check_interval + 1) / 2 * 2);
    tuning_.xcd_chunk_log2 = -1;
    tuning_.coarse_block = coarse_block_;
    tuning_.fused_vcycle = fused_ ? 1 : 0;
    tuning_.graph_replay = use_graph_ ? 1 : 0;
    rbuf_[0] = r_;
    rbuf_[1] = alloc<double>(n_);
    HIP_CHECK(hipMemset(rbuf_[1], 0, n_ * sizeof(double)));
    vc0_[0] = alloc<double>(n_);
    vc0_[1] = alloc<double>(n_);
    const int64_t npart = std::max<int64_t>({2 * static_cast<int64_t>(nb_), 4096, 12 * (M_.npatches / 256 + 1)});
    partials_ = alloc<double>(npart);
    rv_part_ = alloc<double>(2 * static_cast<int64_t>(nb_));
    HIP_CHECK(hipMemset(rv_part_, 0, 2 * static_cast<size_t>(nb_) * sizeof(double)));
    // instances of a batch
    ninst_ = T.instances();
    if (ninst_ < 1) throw std::invalid_argument("topology has no instance");
    if (ninst_ > 1 && opt_.preconditioner == PCFD_PC_LEARNED)
        throw std::invalid_argument("the learned preconditioner runs on single cases, not batches");
    inst_brick_off_ = T.inst_brick_off;
    inst_patch_off_ = T.inst_patch_off;
    inst_maxdiff_ = T.inst_max_diffusion;
    inst_time_.assign(ninst_, 0.0);
    inst_ref_area_ = T.inst_reference_area;
    inst_origin_.assign(3 * static_cast<size_t>(ninst_), 0.0);
    inst_stats_.assign(ninst_, pcfd_step_stats{});
    {
        std::vector<int> inst(nb_);
        for (int i = 0; i < ninst_; ++i)
            for (int b = T.inst_brick_off[i]; b < T.inst_brick_off[i + 1]; ++b) inst[b] = i;
        M_.inst = upload(inst);
        M_.ninst = ninst_;
        // per-brick records and the special incidences resolved to (other cell, conductance)
        std::vector<BrickMeta> meta(nb_);
        for (int b = 0; b < nb_; ++b) {
            BrickMeta& r = meta[b];
            std::memset(&r, 0, sizeof(r));
            for (int f = 0; f < 6; ++f) r.nbr[f] = T.brick_nbr[6 * static_cast<size_t>(b) + f];
            r.special_off = T.brick_special_off[b];
            r.inst = inst[b];
            r.level = static_cast<uint16_t>(T.brick_level[b]);
            r.h = T.level_h.at(r.level);
            r.flags = T.brick_flags[b];
            int run = 0;
            for (int w = 0; w < 8; ++w) {
                r.wave_pre[w] = static_cast<uint16_t>(run);
                for (int q = 64 * w; q < 64 * w + 64; ++q)
                    if (T.mask[512 * static_cast<size_t>(b) + q] & kSpecial) ++run;
            }
            if (T.brick_special_off[b] + run != T.brick_special_off[b + 1])
                throw std::runtime_error("topology: special cells of a brick do not match its special range");
        }
        M_.meta = upload(meta);
        std::vector<int> other(T.special_inc.size());
        std::vector<double> coef(T.special_inc.size());
        for (size_t k = 0; k < T.special_inc.size(); ++k) {
            const int raw = T.special_inc[k];
            const int e = raw >= 0 ? raw : ~raw;
            other[k] = raw >= 0 ? T.edge_b[e] : T.edge_a[e];
            coef[k] = T.edge_coef[e];
        }
        const size_t nsp = T.special_inc_off.empty() ? 0 : T.special_inc_off.size() - 1;
        std::vector<int> other4(4 * nsp);
        std::vector<double> coef4(4 * nsp, 0.0);
        for (size_t sp = 0; sp < nsp; ++sp) {
            const int k0 = T.special_inc_off[sp], k1 = T.special_inc_off[sp + 1];
            for (int j = 0; j < 4; ++j) {
                const bool in = k0 + j < k1;
                other4[4 * sp + j] = in ? other[k0 + j] : T.special_dof[sp];
                coef4[4 * sp + j] = in ? coef[k0 + j] : 0.0;
            }
        }
        M_.sp4_other = upload(other4);
        M_.sp4_coef = upload(coef4);
        M_.xcd_chunk = -1;
        M_.sp_other = upload(other);
        M_.sp_coef = upload(coef);
    }
    seg_ = upload(T.inst_brick_off);
    inst_maxdiff_dev_ = upload(T.inst_max_diffusion);
    S_ = alloc<DevScalars>(ninst_ + 1);
    agg_ = S_ + ninst_;
    HIP_CHECK(hipMemset(S_, 0, (ninst_ + 1) * sizeof(DevScalars)));
    // multigrid
    if (T.mg.size() > 1) {
        parent0_ = upload(T.mg[0].parent);
        lv_.resize(T.mg.size());
        for (size_t m = 1; m < T.mg.size(); ++m) {
            const MGLevel& G = T.mg[m];
            Level& L = lv_[m];
            L.n = G.n;
            CsrView c{};
            c.n = G.n;
            c.row_off = upload(G.row_off);
            c.col = upload(G.col);
            c.parent = m + 1 < T.mg.size() ? upload(G.parent) : nullptr;
            c.child_off = upload(G.child_off);
            c.child = upload(G.child);
            {
                std::vector<int> first(G.n, -1);
                for (int cc = 0; cc < G.n; ++cc)
                    if (G.child_off[cc + 1] > G.child_off[cc]) first[cc] = G.child[G.child_off[cc]];
                c.child_first = upload(first);
            }
            // SELL-64 copy: per slice of 64 rows, slot-major, as wide as its longest row
            const int nsl = (G.n + 63) / 64;
            std::vector<int> sl_off(nsl + 1, 0);
            for (int s = 0; s < nsl; ++s) {
                int K = 0;
                for (int r = 64 * s; r < std::min(G.n, 64 * s + 64); ++r) K = std::max(K, G.row_off[r + 1] - G.row_off[r]);
                sl_off[s + 1] = sl_off[s] + 64 * K;
            }
            std::vector<int> scol(sl_off[nsl]);
            std::vector<double> scoef(sl_off[nsl], 0.0);
            for (int r = 0; r < G.n; ++r) {
                const int s = r >> 6, K = (sl_off[s + 1] - sl_off[s]) / 64;
                for (int k = 0; k < K; ++k) {
                    const int e = G.row_off[r] + k;
                    const size_t slot = static_cast<size_t>(sl_off[s]) + 64 * k + (r & 63);
                    const bool real = e < G.row_off[r + 1];
                    scol[slot] = real ? G.col[e] : r;
                    scoef[slot] = real ? G.coef[e] : 0.0;
                }
            }
            // children of a coarse cell (m >= 2) are one contiguous ascending range
            if (m >= 2)
                for (int cc = 0; cc < G.n; ++cc)
                    for (int k = G.child_off[cc]; k < G.child_off[cc + 1]; ++k)
                        if (G.child[k] != G.child[G.child_off[cc]] + (k - G.child_off[cc]))
                            throw std::logic_error("coarse children are not contiguous");
            // slices without padding whose slots all hold one coefficient
            std::vector<double> suni(nsl, std::numeric_limits<double>::quiet_NaN());
            for (int s = 0; s < nsl; ++s) {
                const int K = (sl_off[s + 1] - sl_off[s]) / 64, r1 = std::min(G.n, 64 * s + 64);
                bool uni = K > 0;
                for (int r = 64 * s; uni && r < r1; ++r) uni = G.row_off[r + 1] - G.row_off[r] == K;
                const double c0 = K > 0 ? G.coef[G.row_off[64 * s]] : 0.0;
                for (int r = 64 * s; uni && r < r1; ++r)
                    for (int e = G.row_off[r]; uni && e < G.row_off[r + 1]; ++e) uni = G.coef[e] == c0;
                if (uni) suni[s] = c0;
            }
            c.sl_off = upload(sl_off);
            c.sl_col = upload(scol);
            L.csr64 = c; L.csr32 = c;
            L.csr64.coef = upload(G.coef);
            L.csr64.diag = upload(G.diag);
            L.csr64.sl_coef = upload(scoef);
            L.csr64.sl_uni = upload(suni);
            L.csr32.coef = upload_as<float>(G.coef);
            L.csr32.diag = upload_as<float>(G.diag);
            L.csr32.sl_coef = upload_as<float>(scoef);
            L.csr32.sl_uni = upload_as<float>(suni);
            L.parent = c.parent;
            L.B = alloc<double>(G.n); L.X = alloc<double>(G.n); L.T = alloc<double>(G.n);
        }
        for (size_t m = 1; m + 1 < T.mg.size(); ++m) {
            const MGLevel& N = T.mg[m + 1];
            bool ok = N.n > 0;
            for (int c0 = 0; ok && c0 < N.n; c0 += kRestrictGroup) {
                const int c1 = std::min(c0 + kRestrictGroup, N.n);
                int prev_end = -1, first = -1;
                for (int c = c0; ok && c < c1; ++c) {
                    const int k0 = N.child_off[c], k1 = N.child_off[c + 1];
                    if (k1 <= k0) { ok = false; break; }
                    const int f = N.child[k0];
                    if (c == c0) first = f;
                    if (f < prev_end) ok = false;
                    prev_end = f + (k1 - k0);
                }
                if (ok && prev_end - first > 256) ok = false;
            }
            lv_[m].fused_restrict = ok;
            bool r8 = N.n > 0;
            for (int c = 0; r8 && c < N.n; ++c) r8 = N.child_off[c + 1] - N.child_off[c] <= 8;
            lv_[m].restrict_rows8 = r8;
        }
        // dense coarse blocks per level (the coarsest level of each instance)
        dense_.assign(T.mg.size(), DenseLevel{});
        if (!T.dense_inv.empty()) {
            dense_inv64_ = upload(T.dense_inv);
            dense_inv32_ = upload_as<float>(T.dense_inv);
        }
        for (size_t m = 1; m < T.mg.size(); ++m) {
            std::vector<int> off, n;
            std::vector<int64_t> ioff;
            for (int i = 0; i < ninst_; ++i)
                if (T.dense_level[i] == static_cast<int>(m)) {
                    off.push_back(T.dense_off[i]);
                    n.push_back(T.dense_n[i]);
                    ioff.push_back(T.dense_inv_off[i]);
                }
            if (off.empty()) continue;
            DenseLevel& D = dense_[m];
            D.count = static_cast<int>(off.size());
            D.max_n = *std::max_element(n.begin(), n.end());
            D.off = upload(off);
            D.n = upload(n);
            D.ioff = upload(ioff);
        }
    }
    if (opt_.vcycle_precision == PCFD_FP64) {
        diag_ = alloc<double>(n_);
        LAUNCH(k_diag0<double>, nb_, 512, M_, static_cast<double*>(diag_));
    } else {
        diag_ = alloc<float>(n_);
        LAUNCH(k_diag0<float>, nb_, 512, M_, static_cast<float*>(diag_));
    }
    if (opt_.preconditioner == PCFD_PC_LEARNED) setup_learned(T);
    setup_gradients(T);
    if (opt_.smagorinsky_cs > 0.0) nu_ = alloc<R>(n_);
    setup_smooth_wall(T);
    if (opt_.nonorthogonal) setup_nonorth(T);
    brick_level_h_ = T.brick_level;
    brick_coord_h_ = T.brick_coord;
    for (int a = 0; a < 3; ++a) { base_bricks_[a] = T.base_bricks[a]; grid_lo_[a] = T.lo[a]; }
    level_h_h_ = T.level_h;
    slice_map_.assign(ninst_, nullptr);
    slice_geo_.assign(ninst_, SliceMap{});
    for (int i = 0; i < ninst_; ++i) build_slice_map(i);
    {
        int device = 0;
        HIP_CHECK(hipGetDevice(&device));
        scratch_ = scratch_address(device);
    }
    HIP_CHECK(hipDeviceSynchronize());
}

// Least-squares gradient stencils of the special cells (amr_advection.cu:834): neighbours
// across the open structured faces (area h^2) and the compact edges (open area), weight
// area / |x_b - x_a|^2 on the fluid centroids; the 3x3 normal matrix is pseudo-inverted.
// They give the velocity gradients of the compact-face transport and reconstruction, the
// pressure gradients of the compact-face and wall traction and of the loads, and the
// Smagorinsky strain of the special cells.
template <class R>
void SolverImpl<R>::setup_gradients(const Topology& T) {
    const int ns = static_cast<int>(T.special_dof.size());
    std::vector<int> off(ns + 1, 0), nbr;
    std::vector<double> wd, inv(9 * static_cast<size_t>(ns), 0.0);
    for (int sp = 0; sp < ns; ++sp) {
        const int64_t a = T.special_dof[sp];
        const int b = static_cast<int>(a / 512), q = static_cast<int>(a % 512);
        const double h = T.level_h[T.brick_level[b]];
        double N[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        auto add = [&](int64_t o, double area) {
            double d[3];
            for (int c = 0; c < 3; ++c) d[c] = T.centroid[3 * o + c] - T.centroid[3 * a + c];
            const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(area > 0) || !(d2 > 1e-20)) throw std::logic_error("invalid gradient stencil geometry");
            const double w = area / d2;
            nbr.push_back(static_cast<int>(o));
            for (int c = 0; c < 3; ++c) wd.push_back(w * d[c]);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) N[3 * r + c] += w * d[r] * d[c];
        };
        const uint8_t m = T.mask[a];
        for (int axis = 0; axis < 3; ++axis)
            for (int sg = -1; sg <= 1; sg += 2) {
                if (!(m & ((sg > 0 ? kPX : kMX) << axis))) continue;
                int c[3] = {q & 7, (q >> 3) & 7, q >> 6};
                int bb = b;
                c[axis] += sg;
                if (c[axis] < 0 || c[axis] > 7) {
                    bb = T.brick_nbr[6 * b + 2 * axis + (sg > 0 ? 1 : 0)];
                    c[axis] = sg > 0 ? 0 : 7;
                }
                add(static_cast<int64_t>(bb) * 512 + c[0] + 8 * c[1] + 64 * c[2], h * h);
            }
        for (int k = T.special_inc_off[sp]; k < T.special_inc_off[sp + 1]; ++k) {
            const int raw = T.special_inc[k];
            const int e = raw >= 0 ? raw : ~raw;
            add(raw >= 0 ? T.edge_b[e] : T.edge_a[e], T.edge_area[e]);
        }
        off[sp + 1] = static_cast<int>(nbr.size());
        symmetric_pinv3(N, &inv[9 * static_cast<size_t>(sp)]);
    }
    smag_.off = upload(off);
    smag_.nbr = upload(nbr);
    smag_.wd = upload(wd);
    smag_.inv = upload(inv);
    grad_v_ = alloc<R>(9 * static_cast<int64_t>(ns));
    adv_sp_ = alloc<R>(3 * static_cast<int64_t>(ns));
    imp_sp_ = alloc<double>(3 * static_cast<int64_t>(ns));
    grad_p_ = alloc<double>(3 * static_cast<int64_t>(ns));
    HIP_CHECK(hipMemset(grad_v_, 0, std::max<size_t>(1, 9 * static_cast<size_t>(ns)) * sizeof(R)));
    HIP_CHECK(hipMemset(grad_p_, 0, std::max<size_t>(1, 3 * static_cast<size_t>(ns)) * sizeof(double)));
}

// Network buffers, brick selection and the initial weights: He-normal (seed 1234) for
// layers 0..2, zero for the output layer, so an untrained network adds nothing.
// Neighbour table of a brick set (position p -> brick bricks[p]): entry 27 p + d, d = (dx+1)
// + 3 (dy+1) + 9 (dz+1), is the position of the same-level brick at offset (dx, dy, dz),
// reached face by face in x, y, z order, or -1 when there is none or it is not in the set.
static std::vector<int> nn_neighbour_table(const std::vector<int>& nbr, const std::vector<int>& bricks, int nbricks) {
    std::vector<int> pos_of(nbricks, -1);
    for (size_t p = 0; p < bricks.size(); ++p) pos_of[bricks[p]] = static_cast<int>(p);
    std::vector<int> tab(27 * bricks.size(), -1);
    for (size_t p = 0; p < bricks.size(); ++p)
        for (int d = 0; d < 27; ++d) {
            const int off[3] = {d % 3 - 1, (d / 3) % 3 - 1, d / 9 - 1};
            int c = bricks[p];
            for (int ax = 0; ax < 3 && c >= 0; ++ax)
                if (off[ax] != 0) c = nbr[6 * static_cast<size_t>(c) + 2 * ax + (off[ax] > 0 ? 1 : 0)];
            tab[27 * p + d] = c >= 0 ? pos_of[c] : -1;
        }
    return tab;
}

template <class R>
void SolverImpl<R>::setup_learned(const Topology& T) {
    if (lv_.empty()) throw std::invalid_argument("the learned preconditioner needs a multigrid hierarchy");
    if (opt_.learned_bricks < 0 || opt_.learned_bricks > 2) throw std::invalid_argument("learned_bricks must be 0, 1 or 2");
    if (opt_.learned_capacity < 0) throw std::invalid_argument("learned_capacity must be non-negative");
    nn_ = true;
    nn_cap_ = opt_.learned_capacity > 0 ? opt_.learned_capacity : 16;
    std::vector<float> p(kNNW + kNNB, 0.f);
    uint64_t state = 1234;
    auto next = [&state] {
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    auto uniform = [&] { return (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
    for (int set = 0; set < kNNSets; ++set)
        for (int l = 0; l + 1 < kNNLayers; ++l) {
            const double sd = std::sqrt(2.0 / (kNNTaps * kNNC));
            for (int t = 0; t < kNNTaps; ++t)
                for (int o = 0; o < kNNC; ++o)
                    for (int i = 0; i < kNNC; ++i) {
                        const double g = std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform());
                        p[static_cast<size_t>(set) * kNNWSet + static_cast<size_t>(l) * kNNWLayer + (t * kNNC + o) * kNNC + i] =
                            static_cast<float>(sd * g);
                    }
        }
    nn_p_ = upload(p);
    nn_w_ = alloc<uint16_t>(kNNW);
    nn_wt_ = alloc<uint16_t>(kNNW);
    LAUNCH(k_nn_pack, blocks_for(kNNW, 256), 256, nn_p_, nn_w_, nn_wt_);
    nn_m_ = alloc<float>(kNNW + kNNB);
    nn_v_ = alloc<float>(kNNW + kNNB);
    nn_g_ = alloc<float>(kNNW + kNNB);
    for (float* q : {nn_m_, nn_v_, nn_g_}) HIP_CHECK(hipMemset(q, 0, (kNNW + kNNB) * sizeof(float)));
    nn_off_ = alloc<int>(1);
    HIP_CHECK(hipMemset(nn_off_, 1, sizeof(int)));
    nn_loss_ = alloc<double>(1);
    // bricks the network corrects
    std::vector<int> list;
    if (opt_.learned_bricks == 1 || opt_.learned_bricks == 2) {
        std::vector<uint8_t> sel(nb_, 0);
        for (int b = 0; b < nb_; ++b) {
            bool hit = false;
            for (int k = T.brick_special_off[b]; k < T.brick_special_off[b + 1] && !hit; ++k)
                hit = opt_.learned_bricks == 1 || T.special_patch[k] >= 0;
            if (!hit) continue;
            sel[b] = 1;
            for (int f = 0; f < 6; ++f) if (T.brick_nbr[6 * b + f] >= 0) sel[T.brick_nbr[6 * b + f]] = 1;
        }
        for (int b = 0; b < nb_; ++b) if (sel[b]) list.push_back(b);
        if (list.empty()) throw std::invalid_argument("learned_bricks selects no brick (no special or cut cells)");
        nn_list_ = upload(list);
    } else {
        for (int b = 0; b < nb_; ++b) list.push_back(b);
    }
    nn_nlist_ = static_cast<int>(list.size());
    nn_cells_ = 0;
    for (int b : list)
        for (int q = 0; q < 512; ++q) nn_cells_ += (T.mask[static_cast<size_t>(b) * 512 + q] & kActive) ? 1 : 0;
    if (nn_cells_ == 0) throw std::invalid_argument("the learned preconditioner's bricks hold no active cell");
    nn_host_nbr_ = T.brick_nbr;
    nn_nbtab_ = upload(nn_neighbour_table(nn_host_nbr_, list, nb_));
    if (nn_list_) {
        std::vector<uint8_t> pm(list.size() * 512);
        for (size_t p = 0; p < list.size(); ++p)
            std::memcpy(&pm[p * 512], &T.mask[static_cast<size_t>(list[p]) * 512], 512);
        nn_masks_ = upload(pm);
    } else {
        nn_masks_ = M_.mask;
    }
    const int64_t slots = static_cast<int64_t>(nn_list_ ? nn_nlist_ : nb_) * 512;
    nn_rstore_ = alloc<float>(slots);
    HIP_CHECK(hipMemset(nn_rstore_, 0, slots * sizeof(float)));
    nn_scratch_ = alloc<uint16_t>(static_cast<int64_t>(nn_nlist_) * kNNStore);
    nn_dz_ = alloc<double>(slots);
    nn_gy_ = alloc<double>(slots);
    nn_dfull_ = alloc<double>(n_);
    nn_sb_ = alloc<float>(nn_nlist_);
    HIP_CHECK(hipMemset(nn_dz_, 0, slots * sizeof(double)));
    HIP_CHECK(hipMemset(nn_dfull_, 0, n_ * sizeof(double)));
}

// Smooth-wall model: the wall cells in dof order with their patches, the cells' volumes and
// each patch's wall distance |(x_cell - x_patch) . n| (amr_advection.cu:934), or V / (2A) where
// that vanishes.  The patch force
// array exists (zero) with the model off, so the load reduction always reads it.
template <class R>
void SolverImpl<R>::setup_smooth_wall(const Topology& T) {
    const int np = static_cast<int>(T.patch_dof.size());
    wall_force_ = alloc<double>(3 * static_cast<int64_t>(np));
    HIP_CHECK(hipMemset(wall_force_, 0, std::max(1, 3 * np) * sizeof(double)));
    wall_applied_.assign(ninst_, 0);
    if (!opt_.smooth_wall || np == 0) return;
    if (!(opt_.nu > 0.0)) throw std::invalid_argument("the smooth-wall model needs a positive molecular viscosity");
    std::vector<int> order(np);
    for (int q = 0; q < np; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return T.patch_dof[x] < T.patch_dof[y]; });
    std::vector<int> special(n_, -1);
    for (size_t sp = 0; sp < T.special_dof.size(); ++sp) special[T.special_dof[sp]] = static_cast<int>(sp);
    std::vector<int> cells, off{0};
    std::vector<double> vol, dist(np, 0.0);
    for (int j = 0; j < np; ++j) {
        const int q = order[j], a = T.patch_dof[q];
        if (cells.empty() || cells.back() != a) {
            if (!cells.empty()) off.push_back(j);
            cells.push_back(a);
            const double h = T.level_h.at(T.brick_level[a >> 9]);
            vol.push_back(special[a] >= 0 ? T.special_volume[special[a]] : h * h * h);
            if (!(vol.back() > 0.0)) throw std::invalid_argument("smooth-wall cell without volume");
        }
        double S2 = 0.0, d = 0.0;
        for (int c = 0; c < 3; ++c) S2 += T.patch_S[3 * q + c] * T.patch_S[3 * q + c];
        const double area = std::sqrt(S2);
        if (!(area > 0.0)) throw std::invalid_argument("smooth-wall patch without area");
        for (int c = 0; c < 3; ++c) d += (T.centroid[3 * static_cast<size_t>(a) + c] - T.patch_x[3 * q + c]) * T.patch_S[3 * q + c] / area;
        dist[q] = std::abs(d);
        // a fluid centroid on the patch's plane (fluid on both sides of a thin edge at coarse
        // resolution) has no distance; the slab estimate V / (2A) takes its place (the
        // reference rejects such a side, amr_advection.cu:934)
        if (!(dist[q] > std::max(1e-14, std::cbrt(vol.back()) * 1e-10))) dist[q] = 0.5 * vol.back() / area;
    }
    off.push_back(np);
    nwall_ = static_cast<int>(cells.size());
    wall_cell_ = upload(cells);
    wall_off_ = upload(off);
    wall_patch_ = upload(order);
    wall_vol_ = upload(vol);
    wall_dist_ = upload(dist);
    wall_k_ = alloc<double>(np);
}

template <class R>
SolverImpl<R>::~SolverImpl() {
    release();
}

// frees every device resource (idempotent); teardown errors cannot be reported from a
// destructor and are ignored deliberately
template <class R>
void SolverImpl<R>::release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (void* p : allocations_) (void)hipFree(p);
    allocations_.clear();
    for (auto& e : ev_) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (auto& e : probe_ev_) if (e) (void)hipEventDestroy(e);
    probe_ev_.clear();
    if (host_s_) (void)hipHostFree(host_s_);
    host_s_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
}

template <class R>
void SolverImpl<R>::reconstruct() {
    LAUNCH(k_reconstruct<R>, nb_, kBL, M_, fields());
    if (nedges_) {
        grad_v();
        LAUNCH(k_reconstruct_edges<R>, blocks_for(nedges_, 256), 256, M_, fields(), static_cast<const R*>(grad_v_));
    }
}

// ------------------------------------------------------------------------------------------
// V-cycle
// ------------------------------------------------------------------------------------------
template <class R>
template <class VT, class CT>
VT* SolverImpl<R>::coarse_cycle_t(int m) {
    Level& L = lv_[m];
    const CsrView& C = std::is_same<CT, double>::value ? L.csr64 : L.csr32;
    VT* B = static_cast<VT*>(L.B);
    VT* cur = static_cast<VT*>(L.X);
    VT* oth = static_cast<VT*>(L.T);
    // sweep s of a smoothing phase has weight w[s & 1] (the two Chebyshev-placed weights by default)
    const double w[2] = {opt_.smoother_weight, opt_.smoother_weight2};
    const int cb = coarse_block_;
    const unsigned g = blocks_for(L.n, cb);
    if (m == static_cast<int>(lv_.size()) - 1) {
        if (!dense_blocks<VT, CT>(m, B, cur)) {  // no factorisation (pure Neumann): 64 sweeps with weight 2/3
            LAUNCH((k_ell_smooth<VT, CT>), g, cb, C, 0, cur, cur, B, 2.0 / 3.0);
            for (int s = 1; s < 64; ++s) { LAUNCH((k_ell_smooth<VT, CT>), g, cb, C, 1, cur, oth, B, 2.0 / 3.0); std::swap(cur, oth); }
        }
        return cur;
    }
    if (opt_.pre_smooth == 2) {
        LAUNCH((k_ell_pre2<VT, CT>), g, cb, C, B, cur, w[0], w[1]);
    } else {
        LAUNCH((k_ell_smooth<VT, CT>), g, cb, C, 0, cur, cur, B, w[0]);
        for (int s = 1; s < opt_.pre_smooth; ++s) { LAUNCH((k_ell_smooth<VT, CT>), g, cb, C, 1, cur, oth, B, w[s & 1]); std::swap(cur, oth); }
    }
    Level& N = lv_[m + 1];
    const CsrView& NC = std::is_same<CT, double>::value ? N.csr64 : N.csr32;
    if (L.fused_restrict && fused_) {
        LAUNCH((k_ell_residual_restrict<VT, CT>), blocks_for(N.n, kRestrictGroup), 256, C, NC, cur, B, static_cast<VT*>(N.B));
    } else if (L.restrict_rows8 && fused_) {
        LAUNCH((k_ell_residual_restrict_rows<VT, CT>), blocks_for(8 * static_cast<int64_t>(N.n), 256), 256, C, NC, cur, B,
               static_cast<VT*>(N.B));
    } else {
        LAUNCH((k_ell_residual<VT, CT>), g, cb, C, cur, B, oth);
        LAUNCH((k_restrict_range<VT>), blocks_for(N.n, cb), cb, NC, oth, static_cast<VT*>(N.B));
    }
    VT* xc = coarse_cycle_t<VT, CT>(m + 1);
    LAUNCH((k_ell_post_prolong<VT, CT>), g, cb, C, cur, xc, B, oth, w[0]);
    std::swap(cur, oth);
    for (int s = 1; s < opt_.post_smooth; ++s) { LAUNCH((k_ell_smooth<VT, CT>), g, cb, C, 1, cur, oth, B, w[s & 1]); std::swap(cur, oth); }
    // instances of a batch whose hierarchy ends at this level: their cells take the exact
    // coarse solve (what their own V-cycle does at its last level)
    dense_blocks<VT, CT>(m, B, cur);
    return cur;
}

template <class R>
template <class VT>
VT* SolverImpl<R>::coarse_cycle(int m) {
    if constexpr (std::is_same<VT, double>::value) return coarse_cycle_t<VT, double>(m);
    else return coarse_cycle_t<VT, float>(m);
}

template <class R>
template <class VT>
void SolverImpl<R>::precondition(bool initial, const double* upd_d, const double* upd_rin) {
    if (upd_d && !update_fused()) throw std::logic_error("fused CG update without the fused V-cycle");
    const int ini = initial ? 1 : 0;
    const int fin = ini ? FIN_RZ_INITIAL : FIN_RZ;
    if (opt_.preconditioner == PCFD_PC_JACOBI) {
        with_z<VT>([&](auto* z) {
            LAUNCH((k_smooth0<VT, std::remove_pointer_t<decltype(z)>>), nb_, kBL, M_, 2, static_cast<const VT*>(nullptr),
                   static_cast<VT*>(nullptr), r_, z, rprev_, 1.0, partials_, S_);
        });
        finalize(fin, nb_);
        return;
    }
    const double w[2] = {opt_.smoother_weight, opt_.smoother_weight2};
    VT* cur = static_cast<VT*>(vc0_[0]);
    VT* oth = static_cast<VT*>(vc0_[1]);
    if (opt_.pre_smooth == 2 && lv_.size() > 1 && fused_) {
        // fused composite level: 2 pre-sweeps | residual + restriction | prolongation + sweep
        using A = typename Arith<VT>::type;
        if (upd_d) {
            probed(6, [&] {
                LAUNCH(k_vc_update_pre2<VT>, nb_, kBL, M_, p_, upd_rin, r_, upd_d, Ad_, static_cast<const A*>(diag_), cur,
                       w[0], w[1], partials_, S_);
            });
            finalize(FIN_UPDATE, nb_);
        } else {
            LAUNCH(k_vc_pre2<VT>, nb_, kBL, M_, r_, static_cast<const A*>(diag_), cur, w[0], w[1], S_);
        }
        LAUNCH(k_vc_restrict0<VT>, nb_, kBL, M_, cur, r_, parent0_, static_cast<VT*>(lv_[1].B), S_);
        VT* xc = coarse_cycle<VT>(1);
        const int last = opt_.post_smooth == 1 ? 1 : 0;
        with_z<VT>([&](auto* z) {
            LAUNCH((k_vc_post_prolong<VT, std::remove_pointer_t<decltype(z)>>), nb_, kBL, M_, last, cur, xc, parent0_, r_, oth,
                   z, rprev_, w[0], partials_, S_);
        });
        std::swap(cur, oth);
        for (int s = 1; s < opt_.post_smooth; ++s) {
            const int mode = s == opt_.post_smooth - 1 ? 2 : 1;
            probed(1, [&] {
                with_z<VT>([&](auto* z) {
                    LAUNCH((k_smooth0<VT, std::remove_pointer_t<decltype(z)>>), nb_, kBL, M_, mode, cur, oth, r_, z, rprev_,
                           w[s & 1], partials_, S_);
                });
            });
            std::swap(cur, oth);
        }
        pc_tail(fin);
        return;
    }
    LAUNCH(k_smooth0<VT>, nb_, kBL, M_, 0, static_cast<const VT*>(nullptr), cur, r_, z_, rprev_, w[0], partials_, S_);
    for (int s = 1; s < opt_.pre_smooth; ++s) {
        LAUNCH(k_smooth0<VT>, nb_, kBL, M_, 1, cur, oth, r_, z_, rprev_, w[s & 1], partials_, S_);
        std::swap(cur, oth);
    }
    if (lv_.size() > 1) {
        LAUNCH(k_residual0<VT>, nb_, 512, M_, cur, r_, oth, S_);
        const CsrView& NC = std::is_same<VT, double>::value ? lv_[1].csr64 : lv_[1].csr32;
        LAUNCH(k_restrict<VT>, blocks_for(lv_[1].n, 256), 256, NC, oth, static_cast<VT*>(lv_[1].B), agg_);
        VT* xc = coarse_cycle<VT>(1);
        LAUNCH(k_prolong<VT>, blocks_for(n_, 256), 256, n_, parent0_, cur, xc, agg_);
    }
    for (int s = 0; s < opt_.post_smooth; ++s) {
        const int mode = s == opt_.post_smooth - 1 ? 2 : 1;
        with_z<VT>([&](auto* z) {
            LAUNCH((k_smooth0<VT, std::remove_pointer_t<decltype(z)>>), nb_, kBL, M_, mode, cur, oth, r_, z, rprev_, w[s & 1],
                   partials_, S_);
        });
        std::swap(cur, oth);
    }
    pc_tail(fin);
}

template <class R>
void SolverImpl<R>::precondition_dispatch(bool initial, const double* upd_d, const double* upd_rin) {
    switch (opt_.vcycle_precision) {
        case PCFD_FP32: precondition<float>(initial, upd_d, upd_rin); break;
        case PCFD_BF16: precondition<bf16>(initial, upd_d, upd_rin); break;
        default: precondition<double>(initial, upd_d, upd_rin); break;
    }
}

// ------------------------------------------------------------------------------------------
// projection: RHS, PCG, correction, divergence check  (amr_pressure.cu:2085-2095)
// ------------------------------------------------------------------------------------------
// PCG from the current p_ on the RHS in rhs_ (FIN_BNORM already issued).  With
// `watch_stall`, returns false as soon as the relative residual read at a check has not
// halved over the last three checks (or is not finite); true otherwise.
template <class R>
bool SolverImpl<R>::pcg_loop(bool watch_stall, bool residual_ready) {
    if (!residual_ready) {
        reset_residual_roles();
        LAUNCH(k_cg_residual, nb_, kBL, M_, p_, rhs_, r_, partials_, S_);
        finalize(FIN_RESID, nb_);
    }
    precondition_dispatch(true);
    // first iteration (beta = 0) outside the replayed runs; d is double-buffered
    cg_iteration(true, dbuf_[0], dbuf_[1]);
    int enqueued = 1, j = 0;
    sync_scalars();
    std::vector<double> hist{host_s_[0].relres};
    // Runs of chunk_ iterations between host reads; when every unfinished instance is
    // predicted to converge in the next iteration (geometric rate of its last two reads) a
    // single iteration runs instead, so that no whole iteration is spent after convergence.
    std::vector<double> last(ninst_);
    for (int i = 0; i < ninst_; ++i) last[i] = host_s_[i].relres;
    int last_count = 0;
    while (!hagg().done && enqueued <= opt_.max_iterations + chunk_) {
        int count = chunk_;
        if (last_count > 0) {
            bool one = true;
            for (int i = 0; i < ninst_ && one; ++i) {
                if (host_s_[i].done) continue;
                const double now = host_s_[i].relres, rate = std::pow(now / last[i], 1.0 / last_count);
                if (!(now > 0.0) || !(rate > 0.0 && rate < 1.0) || !std::isfinite(rate)) { one = false; break; }
                one = now * rate <= opt_.rtol;
            }
            if (one) count = 1;
        }
        for (int i = 0; i < ninst_; ++i) last[i] = host_s_[i].relres;
        if (use_graph_) HIP_CHECK(hipGraphLaunch(iteration_graph(count, j & 1), stream_));
        else enqueue_iterations(count, j);
        j += count;
        set_roles_for(j);   // a replayed graph leaves the host-side roles where its capture left them
        enqueued += count;
        last_count = count;
        sync_scalars();
        if (watch_stall && !host_s_[0].done) {  // the learned solver runs single cases only
            const double rr = host_s_[0].relres;
            if (!std::isfinite(rr)) return false;
            hist.push_back(rr);
            const size_t k = hist.size();
            if (k >= 4 && rr > 0.5 * hist[k - 4]) return false;
        }
    }
    return true;
}

// stats of the batch from the per-instance records: the worst instance for iterations,
// residual, divergence and rates, all instances for convergence, the smallest timestep
template <class R>
void SolverImpl<R>::aggregate_stats(pcfd_step_stats* st) const {
    pcfd_step_stats a = inst_stats_[0];
    for (int i = 1; i < ninst_; ++i) {
        const pcfd_step_stats& s = inst_stats_[i];
        a.iterations = std::max(a.iterations, s.iterations);
        a.converged = a.converged && s.converged;
        a.relative_residual = std::max(a.relative_residual, s.relative_residual);
        a.max_abs_divergence = std::max(a.max_abs_divergence, s.max_abs_divergence);
        a.max_abs_velocity = std::max(a.max_abs_velocity, s.max_abs_velocity);
        a.cfl_rate = std::max(a.cfl_rate, s.cfl_rate);
        a.diffusion_substeps = std::max(a.diffusion_substeps, s.diffusion_substeps);
        a.dt = std::min(a.dt, s.dt);
        a.fallback_used |= s.fallback_used;
    }
    *st = a;
}

// ------------------------------------------------------------------------------------------
// non-orthogonal operator: records to the device, full residual, flexible GMRES
// ------------------------------------------------------------------------------------------
template <class R>
void SolverImpl<R>::setup_nonorth(const Topology& T) {
    if (ninst_ != 1) throw std::invalid_argument("the non-orthogonal operator needs a single case (no batch)");
    if (opt_.preconditioner != PCFD_PC_MG)
        throw std::invalid_argument("the non-orthogonal operator is solved with the MG preconditioner only");
    const Topology::NonOrth local = T.nonorth.built ? Topology::NonOrth{} : compute_nonorthogonal(T);
    const Topology::NonOrth& N = T.nonorth.built ? T.nonorth : local;
    nonorth_ = true;
    NV_.nf = static_cast<int>(N.f_lower.size());
    NV_.nn = static_cast<int>(N.g_dof.size());
    NV_.ninc = static_cast<int>(N.inc_dof.size());
    if (NV_.nf > 0) {
        NV_.lower = upload(N.f_lower); NV_.upper = upload(N.f_upper); NV_.edge = upload(N.f_edge);
        NV_.lnode = upload(N.f_lnode); NV_.unode = upload(N.f_unode);
        NV_.area = upload(N.f_area); NV_.delta = upload(N.f_delta); NV_.k = upload(N.f_k); NV_.w = upload(N.f_w);
        NV_.inc_dof = upload(N.inc_dof); NV_.inc_off = upload(N.inc_off); NV_.inc_face = upload(N.inc_face);
    }
    if (NV_.nn > 0) {
        NV_.gdof = upload(N.g_dof); NV_.goff = upload(N.g_off); NV_.gnbr = upload(N.g_nbr); NV_.gw = upload(N.g_w);
        NV_.grad = alloc<double>(3 * static_cast<int64_t>(NV_.nn));
    }
    gm_m_ = std::min(std::max(opt_.gmres_restart, 2), kGmresMaxDots - 2);
    gm_V_ = alloc<double>(static_cast<int64_t>(gm_m_ + 1) * n_);
    gm_Z_ = alloc<double>(static_cast<int64_t>(gm_m_) * n_);
    HIP_CHECK(hipMemset(gm_V_, 0, static_cast<size_t>(gm_m_ + 1) * n_ * sizeof(double)));
    HIP_CHECK(hipMemset(gm_Z_, 0, static_cast<size_t>(gm_m_) * n_ * sizeof(double)));
    gm_part_ = alloc<double>(static_cast<int64_t>(gm_m_ + 2) * kGmresBlocks);
    gm_dev_ = alloc<double>(gm_m_ + 2);
    gm_cap_ = std::min(std::max(opt_.gmres_recycle, 0), 8);
    if (gm_cap_ > 0) {
        gm_U_ = alloc<double>(static_cast<int64_t>(gm_cap_) * n_);
        gm_C_ = alloc<double>(static_cast<int64_t>(gm_cap_) * n_);
        gm_coef_ = alloc<double>(gm_cap_);
        gm_a_.assign(gm_cap_, 0.0);
    }
    gm_graph_.assign(gm_m_, nullptr);
    gm_exec_.assign(gm_m_, nullptr);
}

// Z[j] = M^-1 V[j] (the V-cycle reads r_ and writes z_), V[j+1] = A_full Z[j]
template <class R>
void SolverImpl<R>::enqueue_fgmres_direction(int j) {
    const int64_t n = n_;
    double* vj = gm_V_ + j * n;
    double* zj = gm_Z_ + j * n;
    double* w = gm_V_ + (j + 1) * n;
    double* const r_save = r_;
    double* const z_save = z_;
    r_ = vj;
    z_ = zj;
    try {
        precondition_dispatch(true);
    } catch (...) {
        r_ = r_save;
        z_ = z_save;
        throw;
    }
    r_ = r_save;
    z_ = z_save;
    LAUNCH((k_cg_dir_apply<double, true>), nb_, kBL, M_, zj, zj, dbuf_[0], w, partials_, S_);
    if (NV_.nf > 0) {
        if (NV_.nn > 0) LAUNCH(k_no_grad, blocks_for(NV_.nn, 256), 256, NV_, zj);
        LAUNCH(k_no_apply, blocks_for(NV_.ninc, 256), 256, NV_, zj, w, 1.0);
    }
}

// the same from a graph captured once per basis index (launch shapes are baked in; the
// graphs are dropped with the PCG graphs when the tuning changes)
template <class R>
void SolverImpl<R>::fgmres_direction(int j) {
    if (!use_graph_) {
        enqueue_fgmres_direction(j);
        return;
    }
    if (!gm_exec_[j]) {
        HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_fgmres_direction(j);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(stream_, &gm_graph_[j]));
        HIP_CHECK(hipGraphInstantiate(&gm_exec_[j], gm_graph_[j], nullptr, nullptr, 0));
    }
    HIP_CHECK(hipGraphLaunch(gm_exec_[j], stream_));
}

// r = rhs - A p with the full operator
template <class R>
void SolverImpl<R>::full_residual(double* r, const double* x) {
    if (!x) x = p_;
    LAUNCH(k_cg_residual, nb_, kBL, M_, x, rhs_, r, partials_, S_);
    if (NV_.nf > 0) {
        if (NV_.nn > 0) LAUNCH(k_no_grad, blocks_for(NV_.nn, 256), 256, NV_, x);
        LAUNCH(k_no_apply, blocks_for(NV_.ninc, 256), 256, NV_, x, r, -1.0);
    }
}
// y = A x with the full (non-orthogonal) operator
template <class R>
void SolverImpl<R>::apply_full(const double* x, double* y) {
    LAUNCH((k_cg_dir_apply<double, true>), nb_, kBL, M_, x, x, dbuf_[0], y, partials_, S_);
    if (NV_.nf > 0) {
        if (NV_.nn > 0) LAUNCH(k_no_grad, blocks_for(NV_.nn, 256), 256, NV_, x);
        LAUNCH(k_no_apply, blocks_for(NV_.ninc, 256), 256, NV_, x, y, 1.0);
    }
}
// the sum of the per-block squares a GMRES vector kernel left in gm_part_
template <class R>
double SolverImpl<R>::partial_norm2() {
    LAUNCH(k_mdot_final, 1, 256, kGmresBlocks, gm_part_, gm_dev_);
    double v = 0;
    HIP_CHECK(hipMemcpyAsync(&v, gm_dev_, sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    return v;
}
// w -= C a, z += zsign U a with a = C^T w, `passes` times (classical Gram-Schmidt, twice for a
// re-orthogonalisation; the reference runs modified Gram-Schmidt twice, amr_pressure.cu:1462-1471,
// 1548-1557); false when a coefficient is not finite
template <class R>
bool SolverImpl<R>::recycle_project(double* w, double* z, double zsign, int passes) {
    for (int pass = 0; pass < passes; ++pass) {
        mdot(gm_C_, gm_k_, w, gm_a_.data());
        for (int v = 0; v < gm_k_; ++v)
            if (!std::isfinite(gm_a_[v])) return false;
        HIP_CHECK(hipMemcpyAsync(gm_coef_, gm_a_.data(), gm_k_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        LAUNCH(k_recycle_update, kGmresBlocks, 256, n_, gm_C_, gm_U_, n_, gm_k_, gm_coef_, w, z, zsign, M_.mask, gm_part_);
    }
    return true;
}

// host_out[v] = V[v] . w over active cells, v < nv (consecutive vectors of length n_)
template <class R>
void SolverImpl<R>::mdot(const double* Vb, int nv, const double* w, double* host_out) {
    if (nv < 1 || nv > kGmresMaxDots) throw std::logic_error("GMRES dot count out of range");
    LAUNCH(k_mdot, kGmresBlocks, 256, n_, Vb, n_, nv, w, M_.mask, gm_part_);
    LAUNCH(k_mdot_final, nv, 256, kGmresBlocks, gm_part_, gm_dev_);
    HIP_CHECK(hipMemcpyAsync(host_out, gm_dev_, nv * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
}

// Flexible GCRO / GMRES(m) (amr_pressure.cu:1390-1698), right-preconditioned by the V-cycle of
// the orthogonal operator (which may differ between applications: FP32/BF16 storage).  Up to
// gmres_recycle realised corrections U with orthonormal images C = A U persist across cycles
// and projections: each cycle first minimises over C (validated against the true residual),
// every Arnoldi image is projected out of C (and its direction out of U with the same
// coefficients), and each cycle's correction joins U/C when its image has a part outside C.
// One-pass classical Gram-Schmidt against the Krylov basis with a second pass only near
// breakdown (||w|| after the projection below 1e-6 of before), Givens rotations on the host.
// Converged when the true residual ||rhs - A p|| <= rtol ||rhs||, checked at the end of every
// cycle (the reference's criterion, amr_pressure.cu:1430-1440).
template <class R>
void SolverImpl<R>::fgmres_solve() {
    const int m = gm_m_;
    const int64_t n = n_;
    double* V = gm_V_;
    double* Z = gm_Z_;
    const unsigned G = kGmresBlocks;
    double bb = 0;
    mdot(rhs_, 1, rhs_, &bb);
    const double bnorm = std::sqrt(bb);
    gm_its_ = 0;
    if (!(bnorm > 0)) {
        LAUNCH(k_no_status, 1, 64, S_, 0, 1, 0.0);
        return;
    }
    double* r = V;  // the cycle's first basis vector is built in place
    full_residual(r);
    double rr = 0;
    mdot(r, 1, r, &rr);
    double res = std::sqrt(rr);
    bool converged = res <= opt_.rtol * bnorm;
    std::vector<double> H(static_cast<size_t>(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), h(m + 2), h2(m + 1);
    while (!converged && gm_its_ < opt_.max_iterations && std::isfinite(res)) {
        if (gm_k_ > 0) {
            // minimise over the recycled images first: x_c = x + U C^T r, r_c = r - C C^T r,
            // kept when the true residual of x_c is below the one before and r_c is within a
            // quarter of the tolerance of it; otherwise the recycle space is dropped
            // (amr_pressure.cu:1452-1497)
            double* xc = Z;
            double* rc = Z + n;
            LAUNCH(k_scale, G, 256, n, p_, 1.0, xc, M_.mask);
            LAUNCH(k_scale, G, 256, n, r, 1.0, rc, M_.mask);
            const double before2 = rr;
            const bool valid = recycle_project(rc, xc, 1.0);
            double true2 = before2, gap2 = 0;
            if (valid) {
                full_residual(r, xc);
                mdot(r, 1, r, &true2);
                LAUNCH(k_axpy_norm, G, 256, n, rc, -1.0, r, M_.mask, gm_part_);
                gap2 = partial_norm2();
            }
            const bool accurate = valid && std::isfinite(gap2) && std::sqrt(gap2) <= 0.25 * opt_.rtol * bnorm &&
                                  true2 < before2;
            if (accurate) {
                LAUNCH(k_scale, G, 256, n, xc, 1.0, p_, M_.mask);
                rr = true2;
            } else {
                gm_k_ = 0;
                ++gm_rejections_;
                if (valid) full_residual(r);  // the residual of the unchanged iterate
                rr = before2;
            }
            res = std::sqrt(rr);
            if (!std::isfinite(res)) break;
            if (res <= opt_.rtol * bnorm) {
                converged = true;
                break;
            }
        }
        LAUNCH(k_scale, G, 256, n, r, 1.0 / res, V, M_.mask);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = res;
        int j = 0;
        for (; j < m && gm_its_ < opt_.max_iterations; ++j) {
            double* w = V + (j + 1) * n;
            fgmres_direction(j);
            // the direction's image leaves the recycled images' span, z the same way in U
            // (amr_pressure.cu:1544-1557); one pass: the Krylov projection below re-orthogonalises
            // near breakdown, and the cycle's end checks the true residual
            if (gm_k_ > 0 && !recycle_project(w, Z + j * n, -1.0, 1)) break;
            // classical Gram-Schmidt; w = V[j+1] follows V[j], so one pass gives V[i].w and w.w.
            // With a V-cycle close to A^-1, A M^-1 v lies nearly in the span and the projection
            // always removes most of w, so Kahan's 0.7 test would re-orthogonalise every time
            // (half the cycle's memory traffic).  A second pass runs only when w almost vanished;
            // the true residual at the end of every cycle keeps the stopping test honest.
            mdot(V, j + 2, w, h.data());
            double before = std::sqrt(std::max(h[j + 1], 0.0)), hn = 0;
            // the coefficients are still in gm_dev_, where the mdot just before left them
            auto project_out = [&]() {
                LAUNCH(k_gs_update, G, 256, n, V, n, j + 1, gm_dev_, w, M_.mask, gm_part_);
                LAUNCH(k_mdot_final, 1, 256, kGmresBlocks, gm_part_, gm_dev_);
                double ww = 0;
                HIP_CHECK(hipMemcpyAsync(&ww, gm_dev_, sizeof(double), hipMemcpyDeviceToHost, stream_));
                HIP_CHECK(hipStreamSynchronize(stream_));
                return std::sqrt(ww);
            };
            hn = project_out();
            std::fill(h2.begin(), h2.end(), 0.0);
            if (hn < 1e-6 * before) {
                mdot(V, j + 1, w, h2.data());
                hn = project_out();
            }
            for (int i = 0; i <= j; ++i) H[i * m + j] = h[i] + h2[i];
            H[(j + 1) * m + j] = hn;
            for (int i = 0; i < j; ++i) {
                const double a = H[i * m + j], b = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * a + sn[i] * b;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
            }
            const double a = H[j * m + j], b = H[(j + 1) * m + j];
            const double rho = std::hypot(a, b);
            cs[j] = rho > 0 ? a / rho : 1.0;
            sn[j] = rho > 0 ? b / rho : 0.0;
            H[j * m + j] = rho;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++gm_its_;
            const double est = std::fabs(g[j + 1]);
            if (!(hn > 0) || !std::isfinite(est) || est <= opt_.rtol * bnorm) { ++j; break; }
            LAUNCH(k_scale, G, 256, n, w, 1.0 / hn, w, M_.mask);
        }
        // y = H^-1 g (upper triangular), p += Z y
        for (int i = j - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[i * m + k] * y[k];
            y[i] = H[i * m + i] != 0.0 ? s / H[i * m + i] : 0.0;
        }
        HIP_CHECK(hipMemcpyAsync(gm_dev_, y.data(), j * sizeof(double), hipMemcpyHostToDevice, stream_));
        double* corr = V + n;  // the cycle's correction Z y (the basis is spent)
        if (gm_cap_ > 0 && j > 0) {
            LAUNCH(k_scale, G, 256, n, corr, 0.0, corr, M_.mask);
            LAUNCH(k_combine, G, 256, n, Z, n, j, gm_dev_, corr, M_.mask);
            LAUNCH(k_axpy_norm, G, 256, n, p_, 1.0, corr, M_.mask, gm_part_);
        } else if (j > 0) {
            LAUNCH(k_combine, G, 256, n, Z, n, j, gm_dev_, p_, M_.mask);
        }
        full_residual(r);
        mdot(r, 1, r, &rr);
        res = std::sqrt(rr);
        converged = res <= opt_.rtol * bnorm;
        if (gm_cap_ > 0 && j > 0) {
            // keep the correction as a recycle pair when its image has a part outside the
            // recycled span (normalised; the oldest pair gives way when full;
            // amr_pressure.cu:1632-1669)
            double* wc = V + 2 * n;
            apply_full(corr, wc);
            double before2 = 0;
            mdot(wc, 1, wc, &before2);
            bool ok = before2 > 0 && std::isfinite(before2);
            if (ok && gm_k_ > 0) ok = recycle_project(wc, corr, -1.0);
            double after2 = before2;
            if (ok && gm_k_ > 0) after2 = partial_norm2();
            const double after = std::sqrt(std::max(after2, 0.0));
            if (ok && std::isfinite(after) && after > 1e-12 * std::sqrt(before2)) {
                if (gm_k_ == gm_cap_) {
                    for (int v = 1; v < gm_cap_; ++v) {
                        HIP_CHECK(hipMemcpyAsync(gm_U_ + (v - 1) * n, gm_U_ + v * n, n * sizeof(double),
                                                 hipMemcpyDeviceToDevice, stream_));
                        HIP_CHECK(hipMemcpyAsync(gm_C_ + (v - 1) * n, gm_C_ + v * n, n * sizeof(double),
                                                 hipMemcpyDeviceToDevice, stream_));
                    }
                    --gm_k_;
                }
                LAUNCH(k_scale, G, 256, n, corr, 1.0 / after, gm_U_ + gm_k_ * n, M_.mask);
                LAUNCH(k_scale, G, 256, n, wc, 1.0 / after, gm_C_ + gm_k_ * n, M_.mask);
                ++gm_k_;
            }
        }
    }
    LAUNCH(k_no_status, 1, 64, S_, gm_its_, converged ? 1 : 0, res / bnorm);
}

template <class R>
void SolverImpl<R>::run_projection(bool warm, pcfd_step_stats* st) {
    p_changed();
    if (!warm) {
        HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(double), stream_));
        gm_k_ = 0;  // a cold start also forgets the recycled corrections
    }
    // PCG: the right-hand side and the initial residual in one pass (b is not stored); the
    // non-orthogonal solve keeps b for its full residuals
    const bool fused_rhs = !nonorth_;
    if (fused_rhs) {
        reset_residual_roles();
        LAUNCH(k_rhs_residual<R>, nb_, kBL, M_, fields(), p_, r_, partials_, S_);
        finalize(FIN_BNORM_RESID, nb_);
    } else {
        LAUNCH(k_rhs<R>, nb_, kBL, M_, fields(), rhs_, partials_, S_);
        finalize(FIN_BNORM, nb_);
    }
    const bool learned = nn_ && nn_on_ && !trial_plain();
    set_nn_off(!learned);
    if (nonorth_) {
        sync_scalars();
        gm_its_ = 0;
        if (!host_s_[0].zero_rhs) fgmres_solve();
    } else {
        pcg_loop(learned, fused_rhs);
    }
    int iterations = nonorth_ ? gm_its_ : host_s_[0].iter;
    bool fell_back = false;
    if (learned && !host_s_[0].converged) {
        // learned solve stalled or failed: plain MGPCG from its iterate on the same RHS
        fell_back = true;
        set_nn_off(true);
        if (!std::isfinite(host_s_[0].relres)) HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(double), stream_));
        LAUNCH(k_rhs<R>, nb_, kBL, M_, fields(), rhs_, partials_, S_);
        finalize(FIN_BNORM, nb_);
        pcg_loop(false);
        iterations += host_s_[0].iter;
    }
    for (int i = 0; i < ninst_; ++i)
        if (host_s_[i].zero_rhs) {
            const int64_t d0 = static_cast<int64_t>(inst_brick_off_[i]) * 512, d1 = static_cast<int64_t>(inst_brick_off_[i + 1]) * 512;
            HIP_CHECK(hipMemsetAsync(p_ + d0, 0, (d1 - d0) * sizeof(double), stream_));
        }
    LAUNCH(k_correct<R>, nb_, kBL, M_, fields(), p_, S_);
    if (nedges_) LAUNCH(k_correct_edges<R>, blocks_for(nedges_, 256), 256, M_, fields(), p_, S_);
    if (nonorth_ && NV_.nf > 0) {
        if (NV_.nn > 0) LAUNCH(k_no_grad, blocks_for(NV_.nn, 256), 256, NV_, p_);
        LAUNCH(k_no_correct<R>, blocks_for(NV_.nf, 256), 256, NV_, fields(), p_, S_, opt_.rho);
    }
    LAUNCH(k_divcheck<R>, nb_, kBL, M_, fields(), partials_, rv_part_);
    finalize(FIN_MAXDIV, nb_);
    sync_scalars();
    for (int i = 0; i < ninst_; ++i) {
        pcfd_step_stats& s = inst_stats_[i];
        s.iterations = ninst_ == 1 ? iterations : host_s_[i].iter;
        s.converged = host_s_[i].converged;
        s.relative_residual = host_s_[i].relres;
        s.max_abs_divergence = host_s_[i].maxdiv;
        s.fallback_used = fell_back ? 1 : 0;
        s.learned_used = learned ? 1 : 0;
    }
    if (st) {
        aggregate_stats(st);
        st->learned_used = learned ? 1 : 0;
    }
}

template <class R>
void SolverImpl<R>::project(double dt, bool warm, pcfd_step_stats* st) {
    if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("projection timestep must be positive");
    for (auto& s : inst_stats_) std::memset(&s, 0, sizeof(s));
    put_dt(dt);
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    run_projection(warm, nullptr);
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipEventSynchronize(ev_[1]));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    for (int i = 0; i < ninst_; ++i) {
        inst_stats_[i].project_ms = ms;
        inst_stats_[i].dt = dt;
        inst_stats_[i].physical_time = inst_time_[i];
    }
    if (st) aggregate_stats(st);
}

template <class R>
void SolverImpl<R>::impulse(double dt) {
    grad_p();
    launch_impulse(dt, 0);
    redistribute();
}

template <class R>
void SolverImpl<R>::advect(double dt, double* dt_used, pcfd_step_stats* st) {
    if (dt > 0) put_dt(dt);
    else {
        LAUNCH(k_outflow_rate<R>, nb_, kBL, M_, fields(), partials_, 0);
        finalize(FIN_RATE, nb_);
    }
    advect_core();
    if (dt_used) *dt_used = host_s_[0].dt;
    if (st) st->diffusion_substeps = inst_stats_[0].diffusion_substeps;
}

// transport with each instance's dt (already in the device scalars), diffusion substeps,
// face reconstruction
template <class R>
void SolverImpl<R>::advect_core() {
    grad_v();
    probed(2, [&] { launch_advect(); });
    for (int c = 0; c < 3; ++c) std::swap(v_[c], vt_[c]);
    redistribute();
    const bool smag = opt_.smagorinsky_cs > 0;
    if (smag) {
        // cell viscosity from the transported velocity, substeps from the largest rate
        LAUNCH(k_viscosity<R>, nb_, 512, M_, fields(), smag_, opt_.nu, opt_.smagorinsky_cs, nu_);
        LAUNCH(k_diff_rate<R>, nb_, 512, M_, nu_, partials_);
        finalize(FIN_DIFFRATE, nb_);
    }
    LAUNCH(k_diffusion_substeps, 1, 64, S_, ninst_, opt_.nu, inst_maxdiff_dev_, smag ? 1 : 0);
    sync_scalars();
    int subs = 0;
    for (int i = 0; i < ninst_; ++i) {
        if (smag && !std::isfinite(host_s_[i].diff_rate)) throw std::domain_error("non-finite eddy viscosity");
        subs = std::max(subs, host_s_[i].subs);
        inst_stats_[i].diffusion_substeps = host_s_[i].subs;
        inst_stats_[i].dt = host_s_[i].dt;
    }
    for (int s = 0; s < subs; ++s) {
        if (smag) LAUNCH(k_diffuse_var<R>, nb_, 512, M_, v_[0], v_[1], v_[2], vt_[0], vt_[1], vt_[2], nu_, S_, s);
        else LAUNCH(k_diffuse<R>, nb_, kBL, M_, v_[0], v_[1], v_[2], vt_[0], vt_[1], vt_[2], S_, s);
        for (int c = 0; c < 3; ++c) std::swap(v_[c], vt_[c]);
        redistribute();
    }
    // wall shear after the diffusion, before the face reconstruction (external_aero_core.cpp:319)
    if (nwall_) {
        // group members without a wall contribute no increment to this update
        if (M_.ngroups) HIP_CHECK(hipMemsetAsync(M_.rdelta, 0, 3 * static_cast<size_t>(M_.nspecial) * sizeof(double), stream_));
        LAUNCH(k_smooth_wall<R>, blocks_for(nwall_, 256), 256, M_, fields(), nwall_, wall_cell_, wall_off_, wall_patch_,
               wall_vol_, wall_dist_, opt_.nu, opt_.rho, wall_k_, wall_force_, S_);
        redistribute();
        std::fill(wall_applied_.begin(), wall_applied_.end(), 1);
    }
    reconstruct();
}

template <class R>
void SolverImpl<R>::initialize(pcfd_step_stats* st) {
    LAUNCH(k_init_velocity<R>, blocks_for(n_, 256), 256, M_, fields());
    reconstruct();
    const double dt = opt_.cfl * M_.finest_h / std::max(1e-9, std::abs(opt_.speed));
    HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(double), stream_));
    p_changed();
    project(dt, false, nullptr);
    grad_p();
    launch_impulse(dt, 1);  // converged instances only
    redistribute();
    LAUNCH(k_vmax<R>, nb_, 512, M_, fields(), partials_);
    finalize(FIN_VMAX, nb_);
    sync_scalars();
    bool all = true;
    for (int i = 0; i < ninst_; ++i) {
        inst_stats_[i].max_abs_velocity = host_s_[i].vmax;
        inst_stats_[i].physical_time = inst_time_[i] = 0.0;
        all = all && inst_stats_[i].converged;
    }
    initialized = all;
    time = 0.0;
    if (st) aggregate_stats(st);
}

template <class R>
void SolverImpl<R>::step(pcfd_step_stats* st) {
    if (!initialized) throw std::logic_error("step before a converged initialization");
    for (auto& s : inst_stats_) std::memset(&s, 0, sizeof(s));
    HIP_CHECK(hipEventRecord(ev_[2], stream_));
    if (!next_rate_ready) {
        LAUNCH(k_outflow_rate<R>, nb_, kBL, M_, fields(), partials_, 1);
        finalize(FIN_RATE_VMAX, nb_);
        sync_scalars();
    }
    // else the last step read them back with its end
    next_rate_ready = false;
    for (int i = 0; i < ninst_; ++i) {
        if (!std::isfinite(host_s_[i].vmax)) throw std::domain_error("non-finite velocity before timestep");
        inst_stats_[i].cfl_rate = host_s_[i].rate;
        inst_stats_[i].max_abs_velocity = host_s_[i].vmax;
    }
    advect_core();
    HIP_CHECK(hipEventRecord(ev_[3], stream_));
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    run_projection(true, nullptr);
    // impulse and time advance for the instances whose projection converged
    grad_p();
    launch_impulse(0.0, 1, rv_part_ + nb_);
    redistribute(rv_part_ + nb_);
    // the next step's rate, timestep and |v|max from the final faces and velocities (the
    // values k_outflow_rate would find), read back with the end of this step
    finalize(FIN_RATE_VMAX, nb_, rv_part_);
    for (int i = 0; i < ninst_; ++i)
        if (inst_stats_[i].converged) inst_time_[i] += host_s_[i].dt;
    time = inst_time_[0];
    HIP_CHECK(hipMemcpyAsync(host_s_, S_, (ninst_ + 1) * sizeof(DevScalars), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipEventSynchronize(ev_[1]));
    float a_ms = 0, p_ms = 0, s_ms = 0;
    HIP_CHECK(hipEventElapsedTime(&a_ms, ev_[2], ev_[3]));
    HIP_CHECK(hipEventElapsedTime(&p_ms, ev_[0], ev_[1]));
    HIP_CHECK(hipEventElapsedTime(&s_ms, ev_[2], ev_[1]));
    const bool used = inst_stats_[0].learned_used != 0;
    if (!tuning_run_) validate_learned(used, p_ms);
    for (int i = 0; i < ninst_; ++i) {
        pcfd_step_stats& s = inst_stats_[i];
        s.learned_rejected = nn_rejected_ ? 1 : 0;
        s.advect_ms = a_ms;
        s.project_ms = p_ms;
        s.step_ms = s_ms;
        s.physical_time = inst_time_[i];
    }
    next_rate_ready = true;
    if (st) aggregate_stats(st);
}

// surface loads of one instance over its patch range (FP64, deterministic): the patch
// partials and their reduction into the instance's scalar record, enqueued on the stream
template <class R>
void SolverImpl<R>::enqueue_loads(int i, double pref) {
    const int p0 = inst_patch_off_[i], p1 = inst_patch_off_[i + 1];
    if (p1 <= p0) return;
    const unsigned gl = blocks_for(p1 - p0, 256);
    const double* o = &inst_origin_[3 * static_cast<size_t>(i)];
    LAUNCH(k_loads, gl, 256, M_, p_, grad_p_, wall_force_, pref, o[0], o[1], o[2], p0, p1, partials_);
    LAUNCH(k_finalize, 1, 1024, FIN_LOADS, partials_, static_cast<int>(gl), static_cast<const int*>(nullptr), S_ + i,
           static_cast<DevScalars*>(nullptr), opt_.cfl, 1.0, opt_.rtol, opt_.max_iterations, 0);
}

// forces, moments and coefficients of instance i from its record (after sync_scalars)
template <class R>
void SolverImpl<R>::read_loads(int i, pcfd_loads* out) const {
    std::memset(out, 0, sizeof(*out));
    if (inst_patch_off_[i + 1] > inst_patch_off_[i]) {
        for (int c = 0; c < 3; ++c) {
            out->force[c] = host_s_[i].loads[c];
            out->moment[c] = host_s_[i].loads[3 + c];
            out->viscous_force[c] = host_s_[i].loads[6 + c];
            out->viscous_moment[c] = host_s_[i].loads[9 + c];
        }
    }
    out->viscous_valid = wall_applied_[i];
    for (int c = 0; c < 3; ++c) {
        out->total_force[c] = out->force[c] + out->viscous_force[c];
        out->total_moment[c] = out->moment[c] + out->viscous_moment[c];
    }
    out->drag = out->force[0]; out->side = out->force[1]; out->lift = out->force[2];
    out->viscous_drag = out->viscous_force[0]; out->viscous_side = out->viscous_force[1]; out->viscous_lift = out->viscous_force[2];
    out->total_drag = out->total_force[0]; out->total_side = out->total_force[1]; out->total_lift = out->total_force[2];
    const double q = 0.5 * opt_.rho * opt_.speed * opt_.speed, area = inst_ref_area_[i];
    if (q > 0 && area > 0) {
        const double qs = q * area;
        out->coefficients_valid = 1;
        out->cd = out->drag / qs; out->cs = out->side / qs; out->cl = out->lift / qs;
        out->cd_viscous = out->viscous_drag / qs; out->cs_viscous = out->viscous_side / qs; out->cl_viscous = out->viscous_lift / qs;
        out->total_cd = out->total_drag / qs; out->total_cs = out->total_side / qs; out->total_cl = out->total_lift / qs;
    }
}

template <class R>
void SolverImpl<R>::loads_instance(int i, double pref, pcfd_loads* out) {
    if (gp_gen_ != p_gen_) grad_p();
    enqueue_loads(i, pref);
    sync_scalars();
    read_loads(i, out);
}

template <class R>
void SolverImpl<R>::batch_loads(double pref, pcfd_loads* out) {
    if (gp_gen_ != p_gen_) grad_p();
    for (int i = 0; i < ninst_; ++i) enqueue_loads(i, pref);
    sync_scalars();
    for (int i = 0; i < ninst_; ++i) read_loads(i, out + i);
}

template <class R>
void SolverImpl<R>::loads(double pref, pcfd_loads* out) {
    if (ninst_ != 1) throw std::invalid_argument("a batch reports loads per instance (pcfd_integrate_loads_instance)");
    loads_instance(0, pref, out);
}

template <class R>
void SolverImpl<R>::get_velocity(double* x, double* y, double* z) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<R> tmp(n_);
    double* out[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
        if (!out[c]) continue;
        HIP_CHECK(hipMemcpy(tmp.data(), v_[c], n_ * sizeof(R), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n_; ++i) out[c][i] = double(tmp[i]);
    }
}

template <class R>
void SolverImpl<R>::set_velocity(const double* x, const double* y, const double* z) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<R> tmp(n_);
    const double* in[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
        if (!in[c]) continue;
        for (int64_t i = 0; i < n_; ++i) tmp[i] = static_cast<R>(in[c][i]);
        HIP_CHECK(hipMemcpy(v_[c], tmp.data(), n_ * sizeof(R), hipMemcpyHostToDevice));
    }
}

template <class R>
void SolverImpl<R>::get_pressure(double* p) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipMemcpy(p, p_, n_ * sizeof(double), hipMemcpyDeviceToHost));
}

template <class R>
void SolverImpl<R>::set_pressure(const double* p) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipMemcpy(p_, p, n_ * sizeof(double), hipMemcpyHostToDevice));
    p_changed();
}

template <class R>
void SolverImpl<R>::get_faces(double* fx, double* fy, double* fz, double* fe) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<R> tmp(std::max<int64_t>(n_, nedges_));
    double* out[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c) {
        if (!out[c]) continue;
        HIP_CHECK(hipMemcpy(tmp.data(), f_[c], n_ * sizeof(R), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n_; ++i) out[c][i] = double(tmp[i]);
    }
    if (fe && nedges_) {
        HIP_CHECK(hipMemcpy(tmp.data(), fe_, nedges_ * sizeof(R), hipMemcpyDeviceToHost));
        for (int i = 0; i < nedges_; ++i) fe[i] = double(tmp[i]);
    }
}

template <class R>
void SolverImpl<R>::set_faces(const double* fx, const double* fy, const double* fz, const double* fe) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<R> tmp(std::max<int64_t>(n_, nedges_));
    const double* in[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c) {
        if (!in[c]) continue;
        for (int64_t i = 0; i < n_; ++i) tmp[i] = static_cast<R>(in[c][i]);
        HIP_CHECK(hipMemcpy(f_[c], tmp.data(), n_ * sizeof(R), hipMemcpyHostToDevice));
    }
    if (fe && nedges_) {
        for (int i = 0; i < nedges_; ++i) tmp[i] = static_cast<R>(fe[i]);
        HIP_CHECK(hipMemcpy(fe_, tmp.data(), nedges_ * sizeof(R), hipMemcpyHostToDevice));
    }
}

template <class R>
void SolverImpl<R>::mg_primitive(int op, int level, const double* x, const double* b, double* out, double omega, int sweeps) {
    if (level < 0 || level >= std::max<int>(1, static_cast<int>(lv_.size()))) throw std::invalid_argument("multigrid level out of range");
    if (opt_.vcycle_precision != PCFD_FP64) throw std::invalid_argument("multigrid primitives run in FP64 only");
    if (ninst_ != 1) throw std::invalid_argument("multigrid primitives run on single cases");
    const int64_t n = level == 0 ? n_ : lv_[level].n;
    HIP_CHECK(hipStreamSynchronize(stream_));
    // clear the loop flags so the kernels run
    std::memset(host_s_, 0, (ninst_ + 1) * sizeof(DevScalars));
    HIP_CHECK(hipMemcpy(S_, host_s_, (ninst_ + 1) * sizeof(DevScalars), hipMemcpyHostToDevice));
    double* dx = level == 0 ? static_cast<double*>(vc0_[0]) : static_cast<double*>(lv_[level].X);
    double* dt = level == 0 ? static_cast<double*>(vc0_[1]) : static_cast<double*>(lv_[level].T);
    double* db = level == 0 ? r_ : static_cast<double*>(lv_[level].B);
    std::vector<double> zeros(n, 0.0);
    HIP_CHECK(hipMemcpy(dx, x ? x : zeros.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(db, b ? b : zeros.data(), n * sizeof(double), hipMemcpyHostToDevice));
    const double* result = dt;
    if (op == 0 || op == 1) {
        // op 0 uses b = 0 and negates: A x = -(0 - A x)
        if (op == 0) HIP_CHECK(hipMemset(db, 0, n * sizeof(double)));
        if (level == 0) LAUNCH(k_residual0<double>, nb_, 512, M_, dx, db, dt, S_);
        else LAUNCH((k_residual_csr<double, double>), blocks_for(n, 256), 256, lv_[level].csr64, dx, db, dt);
    } else if (op == 2) {
        double* cur = dx;
        double* oth = dt;
        for (int s = 0; s < sweeps; ++s) {
            if (level == 0) LAUNCH(k_smooth0<double>, nb_, kBL, M_, 1, cur, oth, db, z_, static_cast<double*>(nullptr), omega, partials_, S_);
            else LAUNCH((k_smooth_csr<double, double>), blocks_for(n, 256), 256, lv_[level].csr64, 1, cur, oth, db, omega);
            std::swap(cur, oth);
        }
        result = cur;
    } else if (op == 3) {
        if (level != 0) throw std::invalid_argument("V-cycle primitive runs on level 0");
        precondition<double>(true, nullptr, nullptr);
        result = z_;
    } else {
        throw std::invalid_argument("unknown multigrid primitive");
    }
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipMemcpy(out, result, n * sizeof(double), hipMemcpyDeviceToHost));
    if (op == 0) for (int64_t i = 0; i < n; ++i) out[i] = -out[i];
}

// Average duration of one launch of a hot kernel on the solver's own stream (HIP events),
// with its algorithmic HBM bytes per launch (each array touched once, neighbours cached).
template <class R>
void SolverImpl<R>::time_kernel(int kernel, int reps, double* avg_ms, double* bytes) {
    if (reps < 1) throw std::invalid_argument("reps must be positive");
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::memset(host_s_, 0, (ninst_ + 1) * sizeof(DevScalars));
    for (int i = 0; i < ninst_; ++i) {
        host_s_[i].dt = 1e-3;
        host_s_[i].rz = 1.0;
    }
    HIP_CHECK(hipMemcpy(S_, host_s_, (ninst_ + 1) * sizeof(DevScalars), hipMemcpyHostToDevice));
    int64_t nact = 0;
    {
        std::vector<uint8_t> m(n_);
        HIP_CHECK(hipMemcpy(m.data(), M_.mask, n_, hipMemcpyDeviceToHost));
        for (uint8_t v : m) nact += (v & kActive) ? 1 : 0;
    }
    const int vt = opt_.vcycle_precision == PCFD_FP64 ? 8 : (opt_.vcycle_precision == PCFD_FP32 ? 4 : 2);
    auto launch = [&] {
        switch (kernel) {
            case 0: LAUNCH(k_cg_apply_dot, nb_, kBL, M_, d_, Ad_, partials_, S_); break;
            case 1:
                if (opt_.vcycle_precision == PCFD_FP32)
                    LAUNCH(k_smooth0<float>, nb_, kBL, M_, 1, static_cast<float*>(vc0_[0]), static_cast<float*>(vc0_[1]), r_, z_, rprev_, 0.6, partials_, S_);
                else if (opt_.vcycle_precision == PCFD_BF16)
                    LAUNCH(k_smooth0<bf16>, nb_, kBL, M_, 1, static_cast<bf16*>(vc0_[0]), static_cast<bf16*>(vc0_[1]), r_, z_, rprev_, 0.6, partials_, S_);
                else
                    LAUNCH(k_smooth0<double>, nb_, kBL, M_, 1, static_cast<double*>(vc0_[0]), static_cast<double*>(vc0_[1]), r_, z_, rprev_, 0.6, partials_, S_);
                break;
            case 2: launch_advect(); break;
            case 3: precondition_dispatch(true); break;
            case 4: LAUNCH(k_nn_apply, nn_grid(), 512, M_, nnv(), nn_setv(), nn_rstore_, r_, nn_dz_, rprev_, partials_, nb_, S_, nn_off_); break;
            case 5:
                if (zf()) LAUNCH((k_cg_dir_apply<float, false>), nb_, kBL, M_, zf32(), dbuf_[0], dbuf_[1], Ad_, partials_, S_);
                else LAUNCH((k_cg_dir_apply<double, false>), nb_, kBL, M_, z_, dbuf_[0], dbuf_[1], Ad_, partials_, S_);
                break;
            case 6:  // alpha = 0 and x = z (scratch): the pressure and the residual are untouched
                if (opt_.vcycle_precision == PCFD_FP32)
                    LAUNCH(k_vc_update_pre2<float>, nb_, kBL, M_, z_, rbuf_[0], rbuf_[1], dbuf_[1], Ad_,
                           static_cast<const float*>(diag_), static_cast<float*>(vc0_[0]), 0.6, 0.6, partials_, S_);
                else if (opt_.vcycle_precision == PCFD_BF16)
                    LAUNCH(k_vc_update_pre2<bf16>, nb_, kBL, M_, z_, rbuf_[0], rbuf_[1], dbuf_[1], Ad_,
                           static_cast<const float*>(diag_), static_cast<bf16*>(vc0_[0]), 0.6, 0.6, partials_, S_);
                else
                    LAUNCH(k_vc_update_pre2<double>, nb_, kBL, M_, z_, rbuf_[0], rbuf_[1], dbuf_[1], Ad_,
                           static_cast<const double*>(diag_), static_cast<double*>(vc0_[0]), 0.6, 0.6, partials_, S_);
                break;
            default: throw std::invalid_argument("unknown kernel id");
        }
    };
    if (kernel == 4) {
        if (!nn_) throw std::invalid_argument("kernel 4 needs the learned preconditioner");
        // a residual with every brick non-zero so that each selected brick runs the network
        LAUNCH(k_cg_residual, nb_, kBL, M_, p_, rhs_, r_, partials_, S_);
        if (zf()) LAUNCH(k_nn_resid<float>, nn_grid(), 512, M_, nn_list_, r_, zf32(), nn_rstore_, nullptr, nullptr);
        else LAUNCH(k_nn_resid<double>, nn_grid(), 512, M_, nn_list_, r_, z_, nn_rstore_, nullptr, nullptr);
        set_nn_off(false);
    }
    launch();  // warm
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    for (int i = 0; i < reps; ++i) launch();
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipEventSynchronize(ev_[1]));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    *avg_ms = ms / reps;
    double per = 0;
    switch (kernel) {
        case 0: per = 8 + 8 + 1; break;                       // d in, Ad out, mask
        case 1: per = vt + 8 + 1 + vt; break;                  // x in, r in, mask, x out
        case 2: per = 9.0 * sizeof(R) + 1; break;             // v in (3), faces (3), mask, v out (3)
        case 3: per = 0; break;
        case 4: per = 4 + 1 + 8 + 8 + 8; break;               // r' (FP32), mask, r, r_prev, dz
        case 5: per = (zf() ? 4 : 8) + 8 + 8 + 8 + 1; break;  // z (FP32 or FP64), d_old in; d_new, Ad out; mask
        case 6: per = 16 + 8 + 8 + 8 + 8 + 1 + (vt == 8 ? 8 : 4) + vt; break;  // x in/out, d, r, Ad, r' out, mask, diag, x1 out
    }
    if (kernel == 4) {
        nact = nn_cells_;
        set_nn_off(!nn_on_);
    }
    *bytes = per * static_cast<double>(nact);
    // leave the loop flags as a finished solve so a stray kernel does nothing
    for (int i = 0; i <= ninst_; ++i) host_s_[i].done = 1;
    HIP_CHECK(hipMemcpy(S_, host_s_, (ninst_ + 1) * sizeof(DevScalars), hipMemcpyHostToDevice));
}

// copy of the evolving state (velocity, faces, pressure) on the solver's stream
template <class R>
void SolverImpl<R>::snapshot_state() {
    if (!snap_p_) {
        for (int c = 0; c < 3; ++c) { snap_v_[c] = alloc<R>(n_); snap_f_[c] = alloc<R>(n_); }
        snap_fe_ = alloc<R>(nedges_);
        snap_p_ = alloc<double>(n_);
    }
    for (int c = 0; c < 3; ++c) {
        HIP_CHECK(hipMemcpyAsync(snap_v_[c], v_[c], n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(snap_f_[c], f_[c], n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
    }
    if (nedges_) HIP_CHECK(hipMemcpyAsync(snap_fe_, fe_, nedges_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
    HIP_CHECK(hipMemcpyAsync(snap_p_, p_, n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    if (M_.npatches) {
        if (!snap_wall_force_) snap_wall_force_ = alloc<double>(3 * static_cast<int64_t>(M_.npatches));
        HIP_CHECK(hipMemcpyAsync(snap_wall_force_, wall_force_, 3 * static_cast<size_t>(M_.npatches) * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream_));
    }
    snap_wall_applied_ = wall_applied_;
}

template <class R>
void SolverImpl<R>::restore_state() {
    next_rate_ready = false;
    for (int c = 0; c < 3; ++c) {
        HIP_CHECK(hipMemcpyAsync(v_[c], snap_v_[c], n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(f_[c], snap_f_[c], n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
    }
    if (nedges_) HIP_CHECK(hipMemcpyAsync(fe_, snap_fe_, nedges_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
    HIP_CHECK(hipMemcpyAsync(p_, snap_p_, n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    p_changed();
    if (M_.npatches)
        HIP_CHECK(hipMemcpyAsync(wall_force_, snap_wall_force_, 3 * static_cast<size_t>(M_.npatches) * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream_));
    wall_applied_ = snap_wall_applied_;
}

// Per-level maps of brick coordinates to the instance's stored bricks, for the slice sampler;
// built for every instance when the solver is created.  An instance of a batch whose bricks do
// not fit the batch's first grid (merged cases of different domains) gets no map; sampling it
// is refused (check_slice).
template <class R>
void SolverImpl<R>::build_slice_map(int instance) {
    const int levels = static_cast<int>(level_h_h_.size());
    if (levels > 8) return;
    SliceMap G{};
    G.levels = levels;
    int64_t total = 0;
    for (int l = 0; l < levels; ++l) {
        G.off[l] = total;
        for (int d = 0; d < 3; ++d) G.dims[l][d] = base_bricks_[d] << l;
        total += static_cast<int64_t>(G.dims[l][0]) * G.dims[l][1] * G.dims[l][2];
        G.h[l] = level_h_h_[l];
    }
    for (int d = 0; d < 3; ++d) G.lo[d] = grid_lo_[d];
    std::vector<int> map(total, -1);
    for (int b = inst_brick_off_[instance]; b < inst_brick_off_[instance + 1]; ++b) {
        const int l = brick_level_h_[b];
        const int* c = &brick_coord_h_[3 * static_cast<size_t>(b)];
        for (int d = 0; d < 3; ++d)
            if (c[d] < 0 || c[d] >= G.dims[l][d]) return;
        map[G.off[l] + (static_cast<int64_t>(c[2]) * G.dims[l][1] + c[1]) * G.dims[l][0] + c[0]] = b;
    }
    slice_map_[instance] = upload(map);
    slice_geo_[instance] = G;
}

// Field of one instance on a plane (an image of nu x nv values, NaN where solid or outside),
// sampled into the device buffer `image` (at least nu * nv floats) and copied to the host.
template <class R>
void SolverImpl<R>::slice_into(int instance, int field, int axis, double coord, const double* lo, const double* hi,
                               int nu, int nv, float* image, float* out) {
    SliceMap G = slice_geo_[instance];
    G.map = slice_map_[instance];
    const int64_t npix = static_cast<int64_t>(nu) * nv;
    const double q = 0.5 * opt_.rho * opt_.speed * opt_.speed;
    LAUNCH(k_slice<R>, blocks_for(npix, 256), 256, M_, fields(), p_, G, field, axis, coord, lo[0], hi[0], lo[1], hi[1],
           nu, nv, field == 5 ? q : 1.0, image);
    HIP_CHECK(hipMemcpyAsync(out, image, npix * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
}

template <class R>
void SolverImpl<R>::check_slice(int instance, int field, int axis, double coord, const double* lo, const double* hi,
                                int nu, int nv, int64_t max_pixels) const {
    if (instance < 0 || instance >= ninst_) throw std::invalid_argument("sample_slice: no such instance");
    if (level_h_h_.size() > 8) throw std::invalid_argument("sample_slice: at most 8 AMR levels");
    if (!slice_map_[instance])
        throw std::invalid_argument("sample_slice: the instance's bricks lie outside the batch's first grid");
    if (field < 0 || field > 5) throw std::invalid_argument("sample_slice: field must be 0..5");
    if (axis < 0 || axis > 2) throw std::invalid_argument("sample_slice: axis must be 0, 1 or 2");
    if (nu < 1 || nv < 1 || static_cast<int64_t>(nu) * nv > max_pixels)
        throw std::invalid_argument("sample_slice: image size must be 1.." + std::to_string(max_pixels) + " pixels");
    if (!std::isfinite(coord) || !(lo[0] < hi[0]) || !(lo[1] < hi[1]) || !std::isfinite(hi[0] - lo[0]) ||
        !std::isfinite(hi[1] - lo[1]))
        throw std::invalid_argument("sample_slice: invalid plane or window");
    const double q = 0.5 * opt_.rho * opt_.speed * opt_.speed;
    if (field == 5 && !(q > 0.0)) throw std::invalid_argument("sample_slice: Cp needs a positive dynamic pressure");
}

// The image is staged in a device buffer of the solver, grown when a larger image is asked for.
template <class R>
void SolverImpl<R>::sample_slice(int instance, int field, int axis, double coord, const double* lo, const double* hi,
                                 int nu, int nv, float* out) {
    check_slice(instance, field, axis, coord, lo, hi, nu, nv, int64_t(1) << 26);
    const int64_t npix = static_cast<int64_t>(nu) * nv;
    if (npix > slice_cap_) {
        slice_buf_ = alloc<float>(npix);
        slice_cap_ = npix;
    }
    slice_into(instance, field, axis, coord, lo, hi, nu, nv, slice_buf_, out);
}

// The same image staged in this GPU's copy of the static array `scratch` (no allocation).
template <class R>
void SolverImpl<R>::sample_slice_static(int instance, int field, int axis, double coord, const double* lo,
                                        const double* hi, int nu, int nv, float* out) {
    check_slice(instance, field, axis, coord, lo, hi, nu, nv, kScratchFloats);
    slice_into(instance, field, axis, coord, lo, hi, nu, nv, static_cast<float*>(scratch_), out);
}

// ------------------------------------------------------------------------------------------
// launch-shape autotuner
// ------------------------------------------------------------------------------------------
// The tunables change which workgroup computes which brick or row (XCD chunking of the
// brick order, coarse-level workgroup size) or fuse kernels whose arithmetic is the same,
// never the arithmetic: every configuration gives bit-identical results.
template <class R>
void SolverImpl<R>::set_tuning(const pcfd_tuning& t) {
    if (t.xcd_chunk_log2 < -1 || t.xcd_chunk_log2 > 16) throw std::invalid_argument("xcd_chunk_log2 must lie in [-1, 16]");
    if (t.coarse_block != 64 && t.coarse_block != 128 && t.coarse_block != 256 && t.coarse_block != 512)
        throw std::invalid_argument("coarse_block must be 64, 128, 256 or 512");
    M_.xcd_chunk = t.xcd_chunk_log2;
    coarse_block_ = t.coarse_block;
    fused_ = t.fused_vcycle != 0;
    use_graph_ = t.graph_replay != 0;
    tuning_.xcd_chunk_log2 = t.xcd_chunk_log2;
    tuning_.coarse_block = t.coarse_block;
    tuning_.fused_vcycle = fused_ ? 1 : 0;
    tuning_.graph_replay = use_graph_ ? 1 : 0;
    drop_graph();  // kernel arguments are baked into the captured chunk
}

// Greedy search, one parameter at a time (XCD chunk, then coarse workgroup size, then
// fused or plain composite-level V-cycle): each candidate runs `repetitions` full steps
// from the same snapshot of the state and keeps its fastest step time.  The state, the
// physical time and the scalars are restored afterwards, so tuning leaves the solution
// where it was.
template <class R>
void SolverImpl<R>::autotune(int repetitions, pcfd_tuning* out) {
    if (!initialized) throw std::logic_error("autotune after a converged initialization");
    if (repetitions < 1) throw std::invalid_argument("repetitions must be positive");
    const std::vector<double> time0 = inst_time_;
    const double t0 = time;
    snapshot_state();
    tuning_run_ = true;
    int timed = 0;
    auto measure = [&](const pcfd_tuning& t) {
        set_tuning(t);
        double best = 1e300;
        for (int r = 0; r < repetitions; ++r) {
            restore_state();
            pcfd_step_stats st{};
            step(&st);
            best = std::min(best, st.step_ms);
        }
        ++timed;
        return best;
    };
    pcfd_tuning cur = tuning_;
    double best = 0.0;
    try {
        const double def = measure(cur);
        best = def;
        for (int lc : {-1, 0, 2, 4, 6}) {
            if (lc == cur.xcd_chunk_log2) continue;
            pcfd_tuning t = cur;
            t.xcd_chunk_log2 = lc;
            const double ms = measure(t);
            if (ms < best) { best = ms; cur = t; }
        }
        for (int cb : {64, 128, 256, 512}) {
            if (cb == cur.coarse_block) continue;
            pcfd_tuning t = cur;
            t.coarse_block = cb;
            const double ms = measure(t);
            if (ms < best) { best = ms; cur = t; }
        }
        if (lv_.size() > 1 && opt_.pre_smooth == 2) {
            pcfd_tuning t = cur;
            t.fused_vcycle = cur.fused_vcycle ? 0 : 1;
            const double ms = measure(t);
            if (ms < best) { best = ms; cur = t; }
        }
        cur.default_step_ms = def;
    } catch (...) {
        tuning_run_ = false;
        restore_state();
        inst_time_ = time0;
        time = t0;
        throw;
    }
    tuning_run_ = false;
    set_tuning(cur);
    restore_state();
    HIP_CHECK(hipStreamSynchronize(stream_));
    inst_time_ = time0;
    time = t0;
    tuning_.candidates = timed;
    tuning_.step_ms = best;
    tuning_.default_step_ms = cur.default_step_ms;
    if (out) *out = tuning_;
}

// ------------------------------------------------------------------------------------------
// learned preconditioner: data, training, hooks
// ------------------------------------------------------------------------------------------
template <class R>
void SolverImpl<R>::learned_collect(int max_pairs, pcfd_learned_stats* st) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    if (!initialized) throw std::logic_error("collect before a converged initialization");
    if (max_pairs < 1) throw std::invalid_argument("max_pairs must be positive");
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    // the collection advances a copy of the state and solves its projection
    snapshot_state();
    auto restore = [&] {
        restore_state();
        set_nn_off(!nn_on_);
    };
    int recorded = 0;
    try {
        double dt = 0.0;
        advect(0.0, &dt, nullptr);
        LAUNCH(k_rhs<R>, nb_, kBL, M_, fields(), rhs_, partials_, S_);
        finalize(FIN_BNORM, nb_);
        // warm start from the current pressure, as the step's projection does
        set_nn_off(true);
        reset_residual_roles();
        LAUNCH(k_cg_residual, nb_, kBL, M_, p_, rhs_, r_, partials_, S_);
        finalize(FIN_RESID, nb_);
        precondition_dispatch(true);
        sync_scalars();
        // r' = r - A M_mg(r) of the current iterate; recorded while the solve runs
        auto record = [&] {
            if (host_s_[0].done || recorded >= max_pairs) return;
            const int slot = nn_next_;
            if (slot >= static_cast<int>(nn_rs_.size())) {
                nn_rs_.push_back(alloc<float>(static_cast<int64_t>(nn_grid()) * 512));
                nn_inv_norm_.push_back(0.0);
            }
            float* rs = nn_rs_[slot];
            if (zf()) LAUNCH(k_nn_resid<float>, nn_grid(), 512, M_, nn_list_, r_, zf32(), rs, nullptr, nullptr);
            else LAUNCH(k_nn_resid<double>, nn_grid(), 512, M_, nn_list_, r_, z_, rs, nullptr, nullptr);
            const double norm = nn_norm(rs);
            if (!(norm > 0.0) || !std::isfinite(norm)) return;
            nn_inv_norm_[slot] = 1.0 / norm;
            nn_next_ = (nn_next_ + 1) % nn_cap_;
            nn_count_ = std::min(nn_count_ + 1, nn_cap_);
            ++recorded;
        };
        record();
        cg_iteration(true, dbuf_[0], dbuf_[1]);
        sync_scalars();
        record();
        for (int it = 1; !host_s_[0].done && recorded < max_pairs && it <= opt_.max_iterations; ++it) {
            cg_iteration(false, dbuf_[it & 1], dbuf_[(it + 1) & 1]);
            sync_scalars();
            record();
        }
    } catch (...) {
        try { restore(); } catch (...) {}
        throw;
    }
    restore();
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipEventSynchronize(ev_[1]));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        st->pairs = nn_count_;
        st->train_steps = nn_t_;
        st->collect_ms = ms;
        st->bricks = nn_nlist_;
        st->cells = nn_cells_;
    }
}

template <class R>
void SolverImpl<R>::learned_train(int steps, int batch, double lr, pcfd_learned_stats* st) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    if (steps < 1 || batch < 1 || !(lr > 0) || !std::isfinite(lr)) throw std::invalid_argument("steps, batch and learning rate must be positive");
    if (nn_count_ == 0) throw std::logic_error("no training pairs (pcfd_learned_collect first)");
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    double first = 0.0, last = 0.0;
    for (int s = 0; s < steps; ++s) {
        HIP_CHECK(hipMemsetAsync(nn_loss_, 0, sizeof(double), stream_));
        for (int j = 0; j < batch; ++j) {
            const int idx = nn_cursor_ % nn_count_;
            nn_cursor_ = (nn_cursor_ + 1) % nn_count_;
            nn_train_pass(nn_rs_[idx], nn_inv_norm_[idx] / batch);
        }
        ++nn_t_;
        const float c1 = static_cast<float>(1.0 - std::pow(0.9, nn_t_));
        const float c2 = static_cast<float>(1.0 - std::pow(0.999, nn_t_));
        LAUNCH(k_nn_adam, blocks_for(kNNW + kNNB, 256), 256, nn_p_, nn_m_, nn_v_, nn_g_, static_cast<float>(lr), 0.9f, 0.999f, 1e-8f, c1, c2);
        LAUNCH(k_nn_pack, blocks_for(kNNW, 256), 256, nn_p_, nn_w_, nn_wt_);
        if (s == 0 || s == steps - 1) {
            double l = 0.0;
            HIP_CHECK(hipMemcpyAsync(&l, nn_loss_, sizeof(double), hipMemcpyDeviceToHost, stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
            if (!std::isfinite(l)) throw std::domain_error("non-finite training loss");
            if (s == 0) first = l;
            last = l;
        }
    }
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipEventSynchronize(ev_[1]));
    set_network(true);
    nn_rejected_ = false;
    trial_left_ = 2 * kTrial;
    trial_ms_ = trial_plain_ms_ = 0.0;
    set_nn_off(false);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        st->pairs = nn_count_;
        st->train_steps = nn_t_;
        st->loss_first = first;
        st->loss_last = last;
        st->train_ms = ms;
        st->bricks = nn_nlist_;
        st->cells = nn_cells_;
    }
}

template <class R>
void SolverImpl<R>::learned_weights(float* w, float* b, const float* wset, const float* bset) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    HIP_CHECK(hipStreamSynchronize(stream_));
    if (wset || bset) {
        if (!(wset && bset)) throw std::invalid_argument("weights and biases are set together");
        for (int i = 0; i < kNNW; ++i) if (!std::isfinite(wset[i])) throw std::invalid_argument("non-finite weight");
        for (int i = 0; i < kNNB; ++i) if (!std::isfinite(bset[i])) throw std::invalid_argument("non-finite bias");
        HIP_CHECK(hipMemcpy(nn_p_, wset, kNNW * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(nn_p_ + kNNW, bset, kNNB * sizeof(float), hipMemcpyHostToDevice));
        LAUNCH(k_nn_pack, blocks_for(kNNW, 256), 256, nn_p_, nn_w_, nn_wt_);
        set_network(true);
        set_nn_off(false);
        HIP_CHECK(hipStreamSynchronize(stream_));
        return;
    }
    if (w) HIP_CHECK(hipMemcpy(w, nn_p_, kNNW * sizeof(float), hipMemcpyDeviceToHost));
    if (b) HIP_CHECK(hipMemcpy(b, nn_p_ + kNNW, kNNB * sizeof(float), hipMemcpyDeviceToHost));
}

template <class R>
void SolverImpl<R>::learned_enable(bool on) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    set_network(on);
    trial_left_ = 0;  // an explicit choice ends any validation in progress
    nn_rejected_ = false;
    set_nn_off(!on);
    HIP_CHECK(hipStreamSynchronize(stream_));
}

template <class R>
void SolverImpl<R>::learned_forward(const double* r, float* y, float* s) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    HIP_CHECK(hipStreamSynchronize(stream_));
    if (!nn_all_nbtab_) {
        std::vector<int> all(nb_);
        for (int b = 0; b < nb_; ++b) all[b] = b;
        nn_all_nbtab_ = upload(nn_neighbour_table(nn_host_nbr_, all, nb_));
        nn_all_store_ = alloc<float>(n_);
        nn_all_y_ = alloc<float>(n_);
    }
    double* dr = static_cast<double*>(vc0_[0]);
    float* ds = reinterpret_cast<float*>(partials_);
    HIP_CHECK(hipMemcpy(dr, r, n_ * sizeof(double), hipMemcpyHostToDevice));
    LAUNCH(k_nn_gather, nb_, 512, nullptr, dr, nn_all_store_);
    LAUNCH(k_nn_forward_raw, nb_, 512, M_, nnv(), (NNSetView{nullptr, nn_all_nbtab_, M_.mask}), nn_all_store_, nn_all_y_, ds);
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipMemcpy(y, nn_all_y_, n_ * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(s, ds, nb_ * sizeof(float), hipMemcpyDeviceToHost));
}

template <class R>
void SolverImpl<R>::learned_gradient(const double* r, double* loss, float* gw, float* gb) {
    if (!nn_) throw std::invalid_argument("the solver was not created with the learned preconditioner");
    HIP_CHECK(hipStreamSynchronize(stream_));
    double* dr = static_cast<double*>(vc0_[0]);
    HIP_CHECK(hipMemcpy(dr, r, n_ * sizeof(double), hipMemcpyHostToDevice));
    LAUNCH(k_nn_gather, nn_grid(), 512, nn_list_, dr, nn_rstore_);
    const double norm = nn_norm(nn_rstore_);
    if (!(norm > 0.0) || !std::isfinite(norm)) throw std::invalid_argument("residual must be non-zero and finite");
    HIP_CHECK(hipMemset(nn_g_, 0, (kNNW + kNNB) * sizeof(float)));
    HIP_CHECK(hipMemset(nn_loss_, 0, sizeof(double)));
    nn_train_pass(nn_rstore_, 1.0 / norm);
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipMemcpy(loss, nn_loss_, sizeof(double), hipMemcpyDeviceToHost));
    if (gw) HIP_CHECK(hipMemcpy(gw, nn_g_, kNNW * sizeof(float), hipMemcpyDeviceToHost));
    if (gb) HIP_CHECK(hipMemcpy(gb, nn_g_ + kNNW, kNNB * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemset(nn_g_, 0, (kNNW + kNNB) * sizeof(float)));
}

}  // namespace pcfd

// ==========================================================================================
// C ABI
// ==========================================================================================
struct pcfd_solver {
    std::unique_ptr<pcfd::Solver> impl;
};

namespace {
template <class F>
pcfd_status guarded(const char* where, F&& f) {
    try {
        f();
        return PCFD_OK;
    } catch (const pcfd::OutOfMemory& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_ENOMEM;
    } catch (const pcfd::DeviceError& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_EDEVICE;
    } catch (const std::invalid_argument& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_EINVAL;
    } catch (const std::domain_error& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_ENONFINITE;
    } catch (const std::logic_error& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_ESTATE;
    } catch (const std::bad_alloc&) {
        pcfd::set_error(std::string(where) + ": out of memory");
        return PCFD_ENOMEM;
    } catch (const std::exception& e) {
        pcfd::set_error(std::string(where) + ": " + e.what());
        return PCFD_EDEVICE;
    }
}
#define NEED(cond, msg)                                   \
    do {                                                  \
        if (!(cond)) {                                    \
            pcfd::set_error(msg);                         \
            return PCFD_EINVAL;                           \
        }                                                 \
    } while (0)
}  // namespace

extern "C" {

void pcfd_default_options(pcfd_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->rho = 1.0; o->nu = 1e-3; o->speed = 1.0; o->cfl = 0.5; o->smagorinsky_cs = 0.0; o->rtol = 1e-8;
    o->max_iterations = 200; o->field_precision = PCFD_FP64; o->vcycle_precision = PCFD_FP64;
    o->preconditioner = PCFD_PC_MG; o->pre_smooth = 2; o->post_smooth = 2; o->check_interval = 2;
    // two damped-Jacobi weights placed at the roots of the degree-2 Chebyshev polynomial on
    // [0.4, 2], the upper part of the spectrum of D^-1 A that the sweeps must damp
    o->smoother_weight = 1.0 / (1.2 + 0.8 / std::sqrt(2.0));
    o->smoother_weight2 = 1.0 / (1.2 - 0.8 / std::sqrt(2.0));
    o->nonorthogonal = 0;
    o->gmres_restart = 24;
    o->gmres_recycle = 0;
    o->learned_bricks = 0;
    o->learned_capacity = 16;
}

pcfd_status pcfd_create(const pcfd_topology* topology, const pcfd_options* options, pcfd_solver** out) {
    NEED(topology && options && out, "pcfd_create: null argument");
    *out = nullptr;
    return guarded("pcfd_create", [&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) throw pcfd::DeviceError("no HIP device");
        if (options->preconditioner < PCFD_PC_MG || options->preconditioner > PCFD_PC_LEARNED)
            throw std::invalid_argument("unknown preconditioner");
        auto s = std::make_unique<pcfd_solver>();
        const pcfd::Topology& T = pcfd_topology_ref(topology);
        if (options->field_precision == PCFD_FP32) s->impl = std::make_unique<pcfd::SolverImpl<float>>(T, *options);
        else if (options->field_precision == PCFD_FP64) s->impl = std::make_unique<pcfd::SolverImpl<double>>(T, *options);
        else throw std::invalid_argument("field precision must be FP64 or FP32");
        *out = s.release();
    });
}

void pcfd_destroy(pcfd_solver* s) { delete s; }

pcfd_status pcfd_initialize(pcfd_solver* s, pcfd_step_stats* st) {
    NEED(s, "pcfd_initialize: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_initialize", [&] { s->impl->initialize(st); });
}
pcfd_status pcfd_step(pcfd_solver* s, pcfd_step_stats* st) {
    NEED(s, "pcfd_step: null solver");
    return guarded("pcfd_step", [&] { s->impl->step(st); });
}
pcfd_status pcfd_advect(pcfd_solver* s, double dt, double* dt_used, pcfd_step_stats* st) {
    NEED(s, "pcfd_advect: null solver");
    NEED(std::isfinite(dt), "pcfd_advect: non-finite dt");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_advect", [&] { s->impl->advect(dt, dt_used, st); });
}
pcfd_status pcfd_project(pcfd_solver* s, double dt, int warm, pcfd_step_stats* st) {
    NEED(s, "pcfd_project: null solver");
    NEED(dt > 0 && std::isfinite(dt), "pcfd_project: dt must be positive");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_project", [&] { s->impl->project(dt, warm != 0, st); });
}
pcfd_status pcfd_apply_pressure_impulse(pcfd_solver* s, double dt) {
    NEED(s, "pcfd_apply_pressure_impulse: null solver");
    NEED(dt > 0 && std::isfinite(dt), "pcfd_apply_pressure_impulse: dt must be positive");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_apply_pressure_impulse", [&] { s->impl->impulse(dt); });
}
pcfd_status pcfd_integrate_loads(pcfd_solver* s, double pref, pcfd_loads* loads) {
    NEED(s && loads, "pcfd_integrate_loads: null argument");
    return guarded("pcfd_integrate_loads", [&] { s->impl->loads(pref, loads); });
}
pcfd_status pcfd_set_reference(pcfd_solver* s, double area, const double origin[3]) {
    NEED(s, "pcfd_set_reference: null solver");
    NEED(area >= 0 && std::isfinite(area), "pcfd_set_reference: invalid area");
    s->impl->set_reference(-1, area, origin);
    return PCFD_OK;
}
pcfd_status pcfd_set_reference_instance(pcfd_solver* s, int instance, double area, const double origin[3]) {
    NEED(s, "pcfd_set_reference_instance: null solver");
    NEED(area >= 0 && std::isfinite(area), "pcfd_set_reference_instance: invalid area");
    NEED(instance >= 0 && instance < s->impl->instances(), "pcfd_set_reference_instance: no such instance");
    s->impl->set_reference(instance, area, origin);
    return PCFD_OK;
}
pcfd_status pcfd_instance_stats(const pcfd_solver* s, int instance, pcfd_step_stats* stats) {
    NEED(s && stats, "pcfd_instance_stats: null argument");
    NEED(instance >= 0 && instance < s->impl->instances(), "pcfd_instance_stats: no such instance");
    s->impl->instance_stats(instance, stats);
    return PCFD_OK;
}
pcfd_status pcfd_autotune(pcfd_solver* s, int repetitions, pcfd_tuning* out) {
    NEED(s, "pcfd_autotune: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_autotune", [&] { s->impl->autotune(repetitions, out); });
}

pcfd_status pcfd_get_tuning(const pcfd_solver* s, pcfd_tuning* out) {
    NEED(s && out, "pcfd_get_tuning: null argument");
    return guarded("pcfd_get_tuning", [&] { s->impl->get_tuning(out); });
}

pcfd_status pcfd_set_tuning(pcfd_solver* s, const pcfd_tuning* t) {
    NEED(s && t, "pcfd_set_tuning: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_set_tuning", [&] { s->impl->set_tuning(*t); });
}

pcfd_status pcfd_integrate_loads_batch(pcfd_solver* s, double pref, pcfd_loads* loads, int count) {
    NEED(s && loads, "pcfd_integrate_loads_batch: null argument");
    NEED(count == s->impl->instances(), "pcfd_integrate_loads_batch: count must equal the number of instances");
    return guarded("pcfd_integrate_loads_batch", [&] { s->impl->batch_loads(pref, loads); });
}

pcfd_status pcfd_sample_slice(pcfd_solver* s, int instance, int field, int axis, double coord, const double lo[2],
                              const double hi[2], int nu, int nv, float* out) {
    NEED(s && lo && hi && out, "pcfd_sample_slice: null argument");
    return guarded("pcfd_sample_slice", [&] { s->impl->sample_slice(instance, field, axis, coord, lo, hi, nu, nv, out); });
}

pcfd_status pcfd_sample_slice_static(pcfd_solver* s, int instance, int field, int axis, double coord, const double lo[2],
                                     const double hi[2], int nu, int nv, float* out) {
    NEED(s && lo && hi && out, "pcfd_sample_slice_static: null argument");
    return guarded("pcfd_sample_slice_static",
                   [&] { s->impl->sample_slice_static(instance, field, axis, coord, lo, hi, nu, nv, out); });
}

pcfd_status pcfd_integrate_loads_instance(pcfd_solver* s, int instance, double pref, pcfd_loads* loads) {
    NEED(s && loads, "pcfd_integrate_loads_instance: null argument");
    NEED(instance >= 0 && instance < s->impl->instances(), "pcfd_integrate_loads_instance: no such instance");
    return guarded("pcfd_integrate_loads_instance", [&] { s->impl->instance_loads(instance, pref, loads); });
}
pcfd_status pcfd_get_velocity(pcfd_solver* s, double* x, double* y, double* z) {
    NEED(s, "pcfd_get_velocity: null solver");
    return guarded("pcfd_get_velocity", [&] { s->impl->get_velocity(x, y, z); });
}
pcfd_status pcfd_set_velocity(pcfd_solver* s, const double* x, const double* y, const double* z) {
    NEED(s, "pcfd_set_velocity: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_set_velocity", [&] { s->impl->set_velocity(x, y, z); });
}
pcfd_status pcfd_get_pressure(pcfd_solver* s, double* p) {
    NEED(s && p, "pcfd_get_pressure: null argument");
    return guarded("pcfd_get_pressure", [&] { s->impl->get_pressure(p); });
}
pcfd_status pcfd_set_pressure(pcfd_solver* s, const double* p) {
    NEED(s && p, "pcfd_set_pressure: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_set_pressure", [&] { s->impl->set_pressure(p); });
}
pcfd_status pcfd_get_faces(pcfd_solver* s, double* fx, double* fy, double* fz, double* fe) {
    NEED(s, "pcfd_get_faces: null solver");
    return guarded("pcfd_get_faces", [&] { s->impl->get_faces(fx, fy, fz, fe); });
}
pcfd_status pcfd_set_faces(pcfd_solver* s, const double* fx, const double* fy, const double* fz, const double* fe) {
    NEED(s, "pcfd_set_faces: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_set_faces", [&] { s->impl->set_faces(fx, fy, fz, fe); });
}
pcfd_status pcfd_mark_initialized(pcfd_solver* s, double physical_time) {
    NEED(s, "pcfd_mark_initialized: null solver");
    NEED(std::isfinite(physical_time), "pcfd_mark_initialized: physical time must be finite");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_mark_initialized", [&] { s->impl->mark_initialized(physical_time); });
}
int64_t pcfd_device_bytes(const pcfd_solver* s) { return s ? s->impl->device_bytes : -1; }
pcfd_status pcfd_mg_primitive(pcfd_solver* s, int op, int level, const double* x, const double* b, double* out,
                              double omega, int sweeps) {
    NEED(s && out, "pcfd_mg_primitive: null argument");
    NEED(sweeps >= 0, "pcfd_mg_primitive: negative sweeps");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_mg_primitive", [&] { s->impl->mg_primitive(op, level, x, b, out, omega, sweeps); });
}
pcfd_status pcfd_time_kernel(pcfd_solver* s, int kernel, int reps, double* avg_ms, double* bytes_per_launch) {
    NEED(s && avg_ms && bytes_per_launch, "pcfd_time_kernel: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_time_kernel", [&] { s->impl->time_kernel(kernel, reps, avg_ms, bytes_per_launch); });
}
pcfd_status pcfd_probe_start(pcfd_solver* s, unsigned kernel_mask) {
    NEED(s, "pcfd_probe_start: null solver");
    return guarded("pcfd_probe_start", [&] { s->impl->probe_start(kernel_mask); });
}
pcfd_status pcfd_probe_stop(pcfd_solver* s, int64_t* launches, double* total_ms) {
    NEED(s, "pcfd_probe_stop: null solver");
    return guarded("pcfd_probe_stop", [&] { s->impl->probe_stop(launches, total_ms); });
}
pcfd_status pcfd_learned_collect(pcfd_solver* s, int max_pairs, pcfd_learned_stats* st) {
    NEED(s, "pcfd_learned_collect: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_collect", [&] { s->impl->learned_collect(max_pairs, st); });
}
pcfd_status pcfd_learned_train(pcfd_solver* s, int steps, int batch, double learning_rate, pcfd_learned_stats* st) {
    NEED(s, "pcfd_learned_train: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_train", [&] { s->impl->learned_train(steps, batch, learning_rate, st); });
}
pcfd_status pcfd_learned_get_weights(pcfd_solver* s, float* weights, float* biases) {
    NEED(s && (weights || biases), "pcfd_learned_get_weights: null argument");
    return guarded("pcfd_learned_get_weights", [&] { s->impl->learned_weights(weights, biases, nullptr, nullptr); });
}
pcfd_status pcfd_learned_set_weights(pcfd_solver* s, const float* weights, const float* biases) {
    NEED(s && weights && biases, "pcfd_learned_set_weights: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_set_weights", [&] { s->impl->learned_weights(nullptr, nullptr, weights, biases); });
}
pcfd_status pcfd_learned_enable(pcfd_solver* s, int on) {
    NEED(s, "pcfd_learned_enable: null solver");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_enable", [&] { s->impl->learned_enable(on != 0); });
}
pcfd_status pcfd_learned_forward(pcfd_solver* s, const double* r, float* y, float* scales) {
    NEED(s && r && y && scales, "pcfd_learned_forward: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_forward", [&] { s->impl->learned_forward(r, y, scales); });
}
pcfd_status pcfd_learned_gradient(pcfd_solver* s, const double* residual, double* loss, float* grad_weights,
                                  float* grad_biases) {
    NEED(s && residual && loss, "pcfd_learned_gradient: null argument");
    s->impl->next_rate_ready = false;
    return guarded("pcfd_learned_gradient", [&] { s->impl->learned_gradient(residual, loss, grad_weights, grad_biases); });
}

}  // extern "C"
