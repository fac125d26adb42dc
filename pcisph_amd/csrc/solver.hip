// solver.hip - host orchestration of the PCISPH step and its C ABI (include/pcisph.h).
//
// Replaces sim::Fluid (PCISPH/src/sim/Fluid.h:13-102, Fluid.cpp:32-466).  One HIP
// stream per handle; every launch of a step is enqueued on it.  The correction
// loop never waits on the stream: the pressure-force kernel of iteration k copies
// that iteration's max density variation into host-visible memory when it starts,
// and the host, which keeps one iteration enqueued ahead, reads it from there to
// decide whether to enqueue another (the reference reads the whole dv buffer back
// and waits on it every iteration, Fluid.cpp:283-306).  Iterations enqueued past
// the last one find the loop finished and return at once.
#include <hip/hip_runtime.h>
#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pcisph.h"
#include "mlp.h"
#include "sph_launch.h"

namespace pcisph {

constexpr float kPi = 3.14159265358979323846f;  // PCISPH/src/utils/constants.h:4

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Error(e_ == hipErrorOutOfMemory ? PCISPH_ERR_OUT_OF_MEMORY : PCISPH_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Fluid::update_deduced_attributes (Fluid.cpp:377-443): the same float operations
// with C++11 <cmath> overloads (pow(float, float) is float, pow(float, int) double).
void derive_params(Simulation_Params& p) {
    p.kernel_radius = 4.f * p.particle_radius;
    p.kernel_radius2 = std::pow(p.kernel_radius, 2.f);
    p.particle_mass = (float)(p.rest_density / std::pow(1.f / (2.f * p.particle_radius), 3));
    uint32_t buckets = p.fluid_count / 2;
    buckets -= buckets % 64;  // a multiple of 64 keeps the 27 neighbour keys distinct
    p.bucket_count = std::max(64u, buckets);
    p.cell_size = p.kernel_radius;
    p.poly6_normalization = 315.f / (64.f * kPi * std::pow(p.kernel_radius, 9.f));
    p.poly6_d1_normalization = -945.f / (32.f * kPi * std::pow(p.kernel_radius, 9.f));
    p.viscosity_d2_normalization = 45.f / (kPi * std::pow(p.kernel_radius, 6.0f));
    p.spiky_d1_normalization = -45.f / (kPi * std::pow(p.kernel_radius, 6.0f));
    p.surface_tension_normalization = 32.f / (kPi * std::pow(p.kernel_radius, 9.f));
    p.surface_tension_term = std::pow(p.kernel_radius, 6.f) / 64.f;
    // PCISPH scaling factor delta from a prototype particle with a full lattice
    // neighbourhood at spacing 2r (Solenthaler & Pajarola 2009, eq. 8).
    const float spacing = 2.f * p.particle_radius;
    if (spacing == 0.f) {  // the lattice loop below would not advance (the reference hangs)
        p.density_variation_scaling_factor = 0.f;
        return;
    }
    const float beta = p.delta_t * p.delta_t * p.particle_mass * p.particle_mass * 2.f /
                       (p.rest_density * p.rest_density);
    float grad_sum[3] = {0.f, 0.f, 0.f};
    float grad_sq_sum = 0.f;
    const float extent = p.kernel_radius + spacing;
    for (float z = -extent; z <= extent; z += 2.f * p.particle_radius)
        for (float y = -extent; y <= extent; y += 2.f * p.particle_radius)
            for (float x = -extent; x <= extent; x += 2.f * p.particle_radius) {
                const float r = std::sqrt(x * x + y * y + z * z);
                const float r2 = r * r;
                if (!(r2 < p.kernel_radius2)) continue;
                const float f = (float)(p.poly6_d1_normalization * std::pow(p.kernel_radius2 - r2, 2));
                const float g[3] = {-f * x, -f * y, -f * z};
                for (int i = 0; i < 3; i++) {
                    grad_sum[i] += g[i];
                    grad_sq_sum += g[i] * g[i];
                }
            }
    float sum_sq = 0.f;
    for (int i = 0; i < 3; i++) sum_sq += grad_sum[i] * grad_sum[i];
    p.density_variation_scaling_factor = -1.f / (beta * (-sum_sq - grad_sq_sum));
}

template <typename R>
SphParams<R> make_sph_params(const Simulation_Params& p, float threshold, const pcisph_solver_options& o) {
    SphParams<R> P;
    std::memset(&P, 0, sizeof(P));  // padding too: scene tables are compared bytewise
    P.h = (R)p.kernel_radius;
    P.h2 = (R)p.kernel_radius2;
    P.cell = (R)p.cell_size;
    P.rho0 = (R)p.rest_density;
    P.m = (R)p.particle_mass;
    P.dt = (R)p.delta_t;
    P.delta = (R)p.density_variation_scaling_factor;
    P.threshold = (R)threshold;
    P.density_scale = (R)p.particle_mass * (R)p.poly6_normalization;
    P.normal_scale = (R)p.kernel_radius * (R)p.particle_mass * (R)p.poly6_d1_normalization;
    P.visc_scale = (R)p.viscosity_constant * (R)p.particle_mass * (R)p.viscosity_d2_normalization;
    P.coh_scale = -(R)p.surface_tension_coefficient * (R)p.particle_mass * (R)p.particle_mass *
                  (R)p.surface_tension_normalization;
    P.curv_scale = -(R)p.surface_tension_coefficient * (R)p.particle_mass;
    P.pf_scale = -(R)p.spiky_d1_normalization * (R)p.particle_mass * (R)p.particle_mass;
    P.st_term = (R)p.surface_tension_term;
    P.gravity_force = (R)p.particle_mass * (R)p.gravity;
    P.rho0_sq = (R)p.rest_density * (R)p.rest_density;
    P.grid = make_hash_grid(p.bucket_count);
    P.n = p.fluid_count;
    P.nb = p.boundary_count;
    P.min_iterations = o.min_iterations;
    P.max_iterations = o.max_iterations;
    return P;
}

uint32_t key_bits_for(uint32_t buckets) {
    uint32_t b = 0;
    while (b < 32 && (uint64_t(1) << b) < buckets) ++b;
    return b;
}

// HIP-event timing of the step's stages.  Event pairs are recorded around each
// stage's launches and resolved lazily (at read time, or when many are pending at
// the end of a step), so timing never waits on the stream inside a step.  Timed
// steps launch kernel by kernel (events cannot be read back from HIP-graph
// replays); with a sampling interval n > 1 only every n-th step is timed and the
// others replay their graphs.
class StageProfiler {
   public:
    ~StageProfiler() {
        for (auto& p : pending_) free_.push_back(p.a), free_.push_back(p.b);
        for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    }

    // every = 0: off; n: time every n-th step from the next one
    void enable(uint32_t every, hipStream_t s) {
        if (!on_ && !every) return;
        resolve(s);
        on_ = every != 0;
        every_ = every ? every : 1;
        first_ = step_;
        for (auto& v : ms_) v = 0;
        for (auto& v : launches_) v = 0;
    }

    // Whether the step about to run is timed.
    bool begin_step() {
        active_ = on_ && (step_ - first_) % every_ == 0;
        return active_;
    }

    template <typename F>
    void timed(int stage, hipStream_t s, int iter, F&& fn) {
        if (!active_) {
            fn();
            return;
        }
        hipEvent_t a = get(), b = get();
        HIP_CHECK(hipEventRecord(a, s));
        fn();
        HIP_CHECK(hipEventRecord(b, s));
        pending_.push_back({stage, iter, step_, a, b, false});
    }

    // Iterations after `last_iter` of this step were enqueued ahead and skipped.
    void end_step(int last_iter) {
        if (active_)
            for (auto& p : pending_)
                if (p.step == step_ && p.iter > last_iter) p.skipped = true;
        active_ = false;
        ++step_;
    }

    bool on() const { return on_; }

    void maybe_resolve(hipStream_t s) {
        if (pending_.size() > 8192) resolve(s);
    }

    void read(hipStream_t s, pcisph_stage_times* out) {
        resolve(s);
        std::memset(out, 0, sizeof(*out));
        for (int i = 0; i < PCISPH_STAGE_COUNT; i++) {
            out->ms[i] = ms_[i];
            out->launches[i] = launches_[i];
        }
    }

   private:
    struct Pending {
        int stage, iter;
        uint64_t step;
        hipEvent_t a, b;
        bool skipped;
    };

    hipEvent_t get() {
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        return e;
    }

    void resolve(hipStream_t s) {
        if (pending_.empty()) return;
        HIP_CHECK(hipStreamSynchronize(s));
        for (auto& p : pending_) {
            if (!p.skipped) {
                float ms = 0;
                HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
                ms_[p.stage] += ms;
                launches_[p.stage] += 1;
            }
            free_.push_back(p.a);
            free_.push_back(p.b);
        }
        pending_.clear();
    }

    bool on_ = false, active_ = false;
    uint64_t step_ = 0, first_ = 0;
    uint32_t every_ = 1;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> free_;
    double ms_[PCISPH_STAGE_COUNT] = {};
    uint64_t launches_[PCISPH_STAGE_COUNT] = {};
};

struct StepResult {
    uint32_t iterations = 0;
    uint32_t converged = 0;
    double max_dv = 0, max_expansion = 0;
    uint64_t fluid_candidates = 0, boundary_candidates = 0;
    bool warm = false, warm_rejected = false;
    uint32_t list_waves = 0;
    double first_dv = 0;  // the plain start's max density variation after iteration 0
};

// What a step does for the warm-start network: use it (params != null) and/or
// append (features, converged pressure) samples of every stride-th particle of
// every scene.
struct WarmArgs {
    const float* params = nullptr;
    uint32_t collect_stride = 0, collect_phase = 0;
    float* store_x = nullptr;
    float* store_t = nullptr;
    uint64_t store_base = 0, store_cap = 0;
};

// One scene of a batch as the engine sees it: its parameters (derived), the
// density-variation threshold of its loop.
struct SceneInput {
    const Simulation_Params* params;
    float threshold;
    bool warm = false;  // the step starts this scene from the network
};

class EngineBase {
   public:
    virtual ~EngineBase() = default;
    // fluid / boundary counts of every scene; keeps each scene's positions,
    // velocities and boundary positions (up to the smaller of old and new count)
    virtual void set_counts(const std::vector<uint32_t>& n, const std::vector<uint32_t>& nb) = 0;
    // bucket tables and sort scratch for the scenes' bucket counts
    virtual void realloc_internal(const std::vector<SceneInput>& scenes) = 0;
    virtual void zero_velocities(uint32_t scene) = 0;
    virtual void upload(uint32_t scene, int field, const void* src, size_t count, bool device) = 0;
    virtual void download(uint32_t scene, int field, void* dst, size_t count, bool device, bool f64) = 0;
    virtual void step(const std::vector<SceneInput>& scenes, const pcisph_solver_options& o, bool sort_boundary,
                      bool fast, const WarmArgs& w, std::vector<StepResult>& out, uint64_t* samples) = 0;
    virtual std::vector<uint32_t> nonfinite() = 0;
    virtual size_t field_size(uint32_t scene, int field, const Simulation_Params& p) const = 0;
    virtual void profile_enable(uint32_t every) = 0;
    // workgroup size of the per-particle kernels for option value `requested` (0: by size)
    virtual uint32_t block_for(uint32_t requested) const = 0;
    virtual void profile_read(pcisph_stage_times* out) = 0;
    virtual void translate_boundary(uint32_t scene, float dx, float dy, float dz) = 0;
    // every scene's walls at once: shifts[3 s .. 3 s + 2] for scene s
    virtual void translate_boundaries(const float* shifts, uint32_t block) = 0;
    // device-side copy of the state a step starts from (positions, velocities with the
    // previous pressures, boundary positions), for trial runs of the autotuner
    virtual void snapshot() = 0;
    virtual void restore() = 0;
};

size_t field_size_for(int field, uint32_t n, uint32_t nb, uint32_t B) {
    switch (field) {
        case PCISPH_FLUID_POSITIONS:
        case PCISPH_FLUID_VELOCITIES:
        case PCISPH_FLUID_NORMALS:
        case PCISPH_FLUID_PREDICTED_POSITIONS:
        case PCISPH_FLUID_OTHER_FORCES:
        case PCISPH_FLUID_PRESSURE_FORCES: return 3 * (size_t)n;
        case PCISPH_FLUID_DENSITIES:
        case PCISPH_FLUID_PRESSURES:
        case PCISPH_FLUID_DENSITY_VARIATIONS:
        case PCISPH_FLUID_KEYS:
        case PCISPH_FLUID_SRC_LOCATIONS: return n;
        case PCISPH_BOUNDARY_POSITIONS: return 3 * (size_t)nb;
        case PCISPH_BOUNDARY_PRESSURES:
        case PCISPH_BOUNDARY_KEYS:
        case PCISPH_BOUNDARY_SRC_LOCATIONS:
        case PCISPH_BOUNDARY_INIT_PRED_DENSITIES: return nb;
        case PCISPH_FLUID_CELL_OFFSETS: return 2 * (size_t)B;
        case PCISPH_BOUNDARY_CELL_OFFSETS: return nb ? 2 * (size_t)B : 0;
        case PCISPH_FLUID_NEIGHBOR_COUNTS: return 4 * (size_t)n;
    }
    return (size_t)-1;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <typename T>
void dalloc(T*& p, size_t count, hipStream_t s) {
    dfree(p);
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    HIP_CHECK(hipMalloc((void**)&p, bytes));
    HIP_CHECK(hipMemsetAsync((void*)p, 0, bytes, s));
}

// Exclusive prefix sums of `v` (size + 1 entries).
std::vector<uint32_t> offsets_of(const std::vector<uint32_t>& v) {
    std::vector<uint32_t> o(v.size() + 1, 0);
    for (size_t i = 0; i < v.size(); ++i) {
        const uint64_t next = (uint64_t)o[i] + v[i];
        if (next > 0xFFFFFFFFull - 1024) throw Error(PCISPH_ERR_SIZE, "more than 2^32 particles or buckets in a batch");
        o[i + 1] = (uint32_t)next;
    }
    return o;
}

template <typename R>
class Engine final : public EngineBase {
   public:
    Engine(int device, hipStream_t stream) : device_(device), stream_(stream) {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipHostMalloc((void**)&host_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
        std::memset((void*)host_, 0, sizeof(HostMirror));
        HIP_CHECK(hipHostGetDevicePointer((void**)&dev_host_, (void*)host_, 0));
        // the correction loop kernel's barrier gives up after 5 s of waiting
        int khz = 0;
        HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
        sync_timeout_ = 5000ull * (unsigned long long)(khz > 0 ? khz : 100000);
    }

    ~Engine() override {
        (void)hipStreamSynchronize(stream_);
        drop_graphs();
        dfree(ctl_);
        free_public();
        free_internal();
        dfree(staging_);
        dfree(feat_);
        dfree(warm_rows_);
        drop_snapshot();
        dfree(A_.NL.f);
        dfree(A_.NL.b);
        dfree(A_.NL.bf);
        dfree(A_.NL.fcount);
        dfree(A_.NL.bcount);
        dfree(A_.NL.bfcount);
        dfree(A_.NL.disp[0]);
        dfree(A_.NL.disp[1]);
        for (int k = 0; k < 2; ++k) {
            dfree(okeys_[k]);
            dfree(ovals_[k]);
        }
        dfree(D_.xp0);
        dfree(D_.xf0);
        dfree(D_.p0);
        dfree(D_.dv0);
        dfree(D_.bp0);
        dfree(D_.bxf0);
        dfree(scene_table_);
        dfree(block_table_);
        dfree(running_table_);
        dfree(sync_);
        dfree(flags_dev_);
        dfree(states_);
        dfree(shift_dev_);
        if (shift_host_) (void)hipHostFree(shift_host_);
        if (shift_event_) (void)hipEventDestroy(shift_event_);
        if (results_) (void)hipHostFree((void*)results_);
        if (host_) (void)hipHostFree((void*)host_);
    }

    void free_public() {
        for (auto** q : {&A_.pos_in, &A_.vel_in, &A_.pos, &A_.vel, &A_.xr, &A_.nrm, &A_.fo, &A_.fp, &A_.xp,
                         &A_.xf, &A_.bpos, &A_.bpos_tmp, &A_.bxf})
            dfree(*q);
        for (auto** q : {&A_.p, &A_.dv, &A_.bp, &A_.binit}) dfree(*q);
        dfree(A_.counts);
        for (int i = 0; i < 2; i++) {
            dfree(A_.keys[i]);
            dfree(A_.vals[i]);
            dfree(A_.bkeys[i]);
            dfree(A_.bvals[i]);
        }
    }

    void free_internal() {
        dfree(canon_);
        dfree(stor_);
        dfree(srcc_);
        dfree(code_);
        dfree(ccells_);
        dfree(A_.cells);
        dfree(A_.bcells);
        dfree(A_.fgrid);
        dfree(A_.bgrid);
        dfree(ext_dev_);
        dfree(sort_scratch_);
        for (int k = 0; k < 2; ++k) {
            dfree(stage_keys_[k]);
            dfree(stage_vals_[k]);
        }
        dfree(stage_rows_);
        dfree(stage_src_);
        dfree(stage_info_);
        dfree(stage_brows_);
        dfree(stage_binfo_);
        for (int k = 0; k < 2; ++k) {
            dfree(bstage_keys_[k]);
            dfree(bstage_vals_[k]);
        }
        stage_n_ = stage_tiles_ = stage_nb_ = 0;
        bwork_key_.clear();
    }

    // New arrays of `count` elements holding each scene's old contents (at most the
    // smaller of its old and new count) at the scene's new offset.
    template <typename T>
    void regrow(T*& p, uint32_t count, const std::vector<uint32_t>& old_off, const std::vector<uint32_t>& new_off) {
        T* fresh = nullptr;
        dalloc(fresh, count, stream_);
        if (p) {
            const size_t scenes = std::min(old_off.size(), new_off.size()) - 1;
            for (size_t s = 0; s < scenes; ++s) {
                const uint32_t keep = std::min(old_off[s + 1] - old_off[s], new_off[s + 1] - new_off[s]);
                if (keep)
                    HIP_CHECK(hipMemcpyAsync(fresh + new_off[s], p + old_off[s], sizeof(T) * keep,
                                             hipMemcpyDeviceToDevice, stream_));
            }
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
        dfree(p);
        p = fresh;
    }

    void set_counts(const std::vector<uint32_t>& n, const std::vector<uint32_t>& nb) override {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        const std::vector<uint32_t> f_off = offsets_of(n), b_off = offsets_of(nb);
        const uint32_t N = f_off.back(), Nb = b_off.back();
        if (f_off != f_off_) {
            canonicalize();  // scenes keep their first particles in the reference's order
            regrow(A_.pos_in, N, f_off_, f_off);
            regrow(A_.vel_in, N, f_off_, f_off);
            dalloc(canon_, N, stream_);
            dalloc(stor_, N, stream_);
            fill_iota(Launch{stream_, 256}, canon_, N);
            fill_iota(Launch{stream_, 256}, stor_, N);
            for (auto** q : {&A_.pos, &A_.vel, &A_.xr, &A_.nrm, &A_.fo, &A_.fp, &A_.xp, &A_.xf}) dalloc(*q, N, stream_);
            dalloc(A_.p, N, stream_);
            dalloc(A_.dv, N, stream_);
            dalloc(A_.counts, N, stream_);
            for (int i = 0; i < 2; i++) {
                dalloc(A_.keys[i], N, stream_);
                dalloc(A_.vals[i], N, stream_);
            }
            A_.n = N;
            fluid_sorted_ = 0;
        }
        if (b_off != b_off_) {
            regrow(A_.bpos, Nb, b_off_, b_off);
            regrow(A_.binit, Nb, b_off_, b_off);
            for (auto** q : {&A_.bpos_tmp, &A_.bxf}) dalloc(*q, Nb, stream_);
            dalloc(A_.bp, Nb, stream_);
            for (int i = 0; i < 2; i++) {
                dalloc(A_.bkeys[i], Nb, stream_);
                dalloc(A_.bvals[i], Nb, stream_);
            }
            A_.nb = Nb;
            boundary_sorted_ = 0;
        }
        n_ = n;
        nb_ = nb;
        f_off_ = f_off;
        b_off_ = b_off;
        uint32_t largest = 1;
        for (size_t s = 0; s < n.size(); ++s) largest = std::max({largest, n[s], nb[s]});
        const size_t stage = 4 * (size_t)largest;
        if (stage > staging_words_) {
            dalloc(staging_, stage * 2, stream_);  // room for double results too
            staging_words_ = stage;
        }
        if (states_n_ != n.size()) {
            dalloc(states_, std::max<size_t>(n.size(), 1), stream_);
            if (results_) (void)hipHostFree((void*)results_);
            results_ = nullptr;
            HIP_CHECK(hipHostMalloc((void**)&results_, sizeof(SceneResult) * std::max<size_t>(n.size(), 1),
                                    hipHostMallocMapped | hipHostMallocCoherent));
            std::memset((void*)results_, 0, sizeof(SceneResult) * std::max<size_t>(n.size(), 1));
            HIP_CHECK(hipHostGetDevicePointer((void**)&dev_results_, (void*)results_, 0));
            states_n_ = (uint32_t)n.size();
        }
        blocks_for_ = 0;  // block table out of date
        boxes_dirty_ = true;
        near_.assign(n.size(), 0);
        near_steps_.assign(n.size(), 0);
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void realloc_internal(const std::vector<SceneInput>& scenes) override {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        std::vector<uint32_t> buckets(scenes.size());
        for (size_t s = 0; s < scenes.size(); ++s) buckets[s] = scenes[s].params->bucket_count;
        const std::vector<uint32_t> base = offsets_of(buckets);
        if (base.back() != A_.B || !A_.cells) {
            dalloc(A_.cells, base.back(), stream_);
            dalloc(A_.bcells, base.back(), stream_);
            A_.B = base.back();
        }
        buckets_ = buckets;
        bucket_base_ = base;
        boxes_dirty_ = true;
        const size_t words = std::max(radix_sort_scratch_words(A_.n, 8), radix_sort_scratch_words(A_.nb, 8));
        if (words > sort_scratch_words_) {
            dalloc(sort_scratch_, words, stream_);
            sort_scratch_words_ = words;
        }
        // Uninitialised in the reference (allocated, never written, Fluid.cpp:451):
        // defined here as zero.
        if (A_.nb) HIP_CHECK(hipMemsetAsync(A_.binit, 0, sizeof(R) * A_.nb, stream_));
    }

    void zero_velocities(uint32_t scene) override {
        if (n_[scene])
            HIP_CHECK(hipMemsetAsync(A_.vel_in + f_off_[scene], 0, sizeof(V4<R>) * n_[scene], stream_));
    }

    void translate_boundary(uint32_t scene, float dx, float dy, float dz) override {
        DataLaunch<R>::translate(Launch{stream_, 256}, A_.bpos + b_off_[scene], nb_[scene], dx, dy, dz);
        HIP_CHECK(hipGetLastError());
    }

    void translate_boundaries(const float* shifts, uint32_t block) override {
        prepare_blocks(block, blocks_for_ & 0x80000000u);
        const size_t S = n_.size();
        if (shift_cap_ < S) {
            HIP_CHECK(hipStreamSynchronize(stream_));
            if (shift_host_) (void)hipHostFree(shift_host_);
            shift_host_ = nullptr;
            HIP_CHECK(hipHostMalloc((void**)&shift_host_, sizeof(float) * 3 * S, hipHostMallocDefault));
            dalloc(shift_dev_, 3 * S, stream_);
            if (!shift_event_) HIP_CHECK(hipEventCreateWithFlags(&shift_event_, hipEventDisableTiming));
            shift_cap_ = S;
        } else {
            HIP_CHECK(hipEventSynchronize(shift_event_));  // the last copy out of shift_host_ is done
        }
        std::memcpy(shift_host_, shifts, sizeof(float) * 3 * S);
        HIP_CHECK(hipMemcpyAsync(shift_dev_, shift_host_, sizeof(float) * 3 * S, hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipEventRecord(shift_event_, stream_));
        Batch<R> B{};
        B.blocks = block_table_;
        B.fluid_blocks = fluid_blocks_;
        B.boundary_blocks = boundary_blocks_;
        DataLaunch<R>::translate_scenes(Launch{stream_, block}, B, shift_dev_, A_.bpos);
        HIP_CHECK(hipGetLastError());
    }

    void snapshot() override {
        canonicalize();
        dalloc(snap_pos_, std::max(A_.n, 1u), stream_);
        dalloc(snap_vel_, std::max(A_.n, 1u), stream_);
        dalloc(snap_b_, std::max(A_.nb, 1u), stream_);
        HIP_CHECK(hipMemcpyAsync(snap_pos_, A_.pos_in, sizeof(V4<R>) * A_.n, hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(snap_vel_, A_.vel_in, sizeof(V4<R>) * A_.n, hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(snap_b_, A_.bpos, sizeof(V4<R>) * A_.nb, hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void restore() override {
        if (!snap_pos_) throw Error(PCISPH_ERR_INTERNAL, "no snapshot");
        HIP_CHECK(hipMemcpyAsync(A_.pos_in, snap_pos_, sizeof(V4<R>) * A_.n, hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(A_.vel_in, snap_vel_, sizeof(V4<R>) * A_.n, hipMemcpyDeviceToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(A_.bpos, snap_b_, sizeof(V4<R>) * A_.nb, hipMemcpyDeviceToDevice, stream_));
        if (layout_active_) {  // the snapshot is in canonical order
            fill_iota(Launch{stream_, 256}, canon_, A_.n);
            fill_iota(Launch{stream_, 256}, stor_, A_.n);
            HIP_CHECK(hipGetLastError());
            layout_active_ = false;
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
        boxes_dirty_ = true;
    }

    void drop_snapshot() {
        dfree(snap_pos_);
        dfree(snap_vel_);
        dfree(snap_b_);
    }

    size_t field_size(uint32_t, int field, const Simulation_Params& p) const override {
        return field_size_for(field, p.fluid_count, p.boundary_count, p.bucket_count);
    }

    struct FieldView {
        V4<R>* vec = nullptr;    // vector field
        R* scalar = nullptr;     // scalar field base
        uint32_t stride = 1, offset = 0, n = 0;
        const uint32_t* raw = nullptr;  // uint32 field
        size_t raw_words = 0;
        uint32_t raw_bias = 0;          // subtracted from every word of a host download
        bool raw_pairs = false;         // (start, end) pairs: (0, 0) is an empty bucket
        const uint32_t* map = nullptr;  // element i of the public array is device element map[i]
        uint32_t raw_item_words = 0;    // uint32 words per element of a mapped uint32 field
    };

    // The canonical bucket ranges (the reference's cell table) of the last sort, from
    // its keys; a layout step keeps only the ranges in slots.
    const uint2* canonical_cells() {
        if (ccells_b_ != A_.B || !ccells_) {
            dalloc(ccells_, std::max(A_.B, 1u), stream_);
            ccells_b_ = A_.B;
        }
        const Launch L{stream_, 256};
        fill_u32(L, (uint32_t*)ccells_, 2 * (size_t)A_.B, 0u);
        cells_from_keys(L, A_.n, A_.keys[fluid_sorted_], ccells_);
        HIP_CHECK(hipGetLastError());
        return ccells_;
    }

    FieldView view(uint32_t s, int field) {
        FieldView v;
        const uint32_t f = f_off_[s], b = b_off_[s], n = n_[s], nb = nb_[s];
        const uint32_t B = s < buckets_.size() ? buckets_[s] : 0, base = s < buckets_.size() ? bucket_base_[s] : 0;
        // fluid per-particle arrays: through the slot map while the arrays are in a layout
        const uint32_t fo = layout_active_ ? 0u : f;
        if (layout_active_) v.map = stor_ + f;
        switch (field) {
            case PCISPH_FLUID_POSITIONS: v.vec = A_.pos_in + fo; v.n = n; break;
            case PCISPH_FLUID_VELOCITIES: v.vec = A_.vel_in + fo; v.n = n; break;
            case PCISPH_FLUID_NORMALS: v.vec = A_.nrm + fo; v.n = n; break;
            case PCISPH_FLUID_PREDICTED_POSITIONS: v.vec = A_.xp + fo; v.n = n; break;
            case PCISPH_FLUID_OTHER_FORCES: v.vec = A_.fo + fo; v.n = n; break;
            case PCISPH_FLUID_PRESSURE_FORCES: v.vec = A_.fp + fo; v.n = n; break;
            case PCISPH_FLUID_DENSITIES: v.scalar = (R*)(A_.xr + fo); v.stride = 4; v.offset = 3; v.n = n; break;
            case PCISPH_FLUID_PRESSURES: v.scalar = A_.p + fo; v.n = n; break;
            case PCISPH_FLUID_DENSITY_VARIATIONS: v.scalar = A_.dv + fo; v.n = n; break;
            case PCISPH_BOUNDARY_POSITIONS: v.vec = A_.bpos + b; v.n = nb; v.map = nullptr; break;
            case PCISPH_BOUNDARY_PRESSURES: v.scalar = A_.bp + b; v.n = nb; v.map = nullptr; break;
            case PCISPH_BOUNDARY_INIT_PRED_DENSITIES: v.scalar = A_.binit + b; v.n = nb; v.map = nullptr; break;
            // sort products are in canonical order
            case PCISPH_FLUID_KEYS:
                v.raw = A_.keys[fluid_sorted_] + f; v.raw_words = n; v.raw_bias = base; v.map = nullptr;
                break;
            case PCISPH_FLUID_SRC_LOCATIONS:
                v.raw = (sorted_layout_ ? srcc_ : A_.vals[fluid_sorted_]) + f; v.raw_words = n; v.raw_bias = f;
                v.map = nullptr;
                break;
            case PCISPH_FLUID_CELL_OFFSETS:
                v.raw = (const uint32_t*)((sorted_layout_ ? canonical_cells() : A_.cells) + base);
                v.raw_words = 2 * (size_t)B; v.raw_bias = f; v.raw_pairs = true; v.map = nullptr;
                break;
            case PCISPH_BOUNDARY_KEYS:
                v.raw = A_.bkeys[boundary_sorted_] + b; v.raw_words = nb; v.raw_bias = base; v.map = nullptr;
                break;
            case PCISPH_BOUNDARY_SRC_LOCATIONS:
                v.raw = A_.bvals[boundary_sorted_] + b; v.raw_words = nb; v.raw_bias = b; v.map = nullptr;
                break;
            case PCISPH_BOUNDARY_CELL_OFFSETS:
                v.raw = (const uint32_t*)(A_.bcells + base); v.raw_words = nb ? 2 * (size_t)B : 0; v.raw_bias = b;
                v.raw_pairs = true; v.map = nullptr;
                break;
            case PCISPH_FLUID_NEIGHBOR_COUNTS:
                v.raw = (const uint32_t*)(A_.counts + fo); v.raw_words = 4 * (size_t)n; v.raw_item_words = 4;
                break;
            default: throw Error(PCISPH_ERR_INVALID_ARGUMENT, "unknown field " + std::to_string(field));
        }
        return v;
    }

    void upload(uint32_t s, int field, const void* src, size_t count, bool device) override {
        HIP_CHECK(hipSetDevice(device_));
        FieldView v = view(s, field);
        if (v.raw) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "field is read-only (a product of the sort)");
        if (!count) return;
        const Launch L{stream_, 256};
        const float* from = (const float*)src;
        if (!device) {
            HIP_CHECK(hipMemcpyAsync(staging_, src, count * sizeof(float), hipMemcpyHostToDevice, stream_));
            from = (const float*)staging_;
        }
        if (v.vec) DataLaunch<R>::unpack_vec3(L, from, v.vec, v.n, v.map);
        else DataLaunch<R>::unpack_scalar(L, from, v.scalar, v.n, v.stride, v.offset, v.map);
        if (field == PCISPH_FLUID_POSITIONS || field == PCISPH_BOUNDARY_POSITIONS) boxes_dirty_ = true;
        HIP_CHECK(hipGetLastError());
        if (!device) HIP_CHECK(hipStreamSynchronize(stream_));  // staging reuse + host buffer lifetime
    }

    void download(uint32_t s, int field, void* dst, size_t count, bool device, bool f64) override {
        HIP_CHECK(hipSetDevice(device_));
        FieldView v = view(s, field);
        if (!count) return;
        const Launch L{stream_, 256};
        if (v.raw) {
            if (f64) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "integer field has no FP64 view");
            if (device && v.raw_bias)
                throw Error(PCISPH_ERR_INVALID_ARGUMENT,
                            "device download of a sort product is available for the first scene only");
            const uint32_t* from = v.raw;
            if (v.map) {  // a per-particle field in a layout: gathered into canonical order first
                gather_words(L, v.raw_item_words ? (uint32_t)(count / v.raw_item_words) : 0u, v.raw_item_words,
                             v.map, v.raw, (uint32_t*)staging_);
                HIP_CHECK(hipGetLastError());
                from = (const uint32_t*)staging_;
            }
            HIP_CHECK(hipMemcpyAsync(dst, from, count * 4, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
            if (v.raw_bias) {  // batch-wide indices and keys to the scene's own
                uint32_t* w = (uint32_t*)dst;
                if (v.raw_pairs) {
                    for (size_t k = 0; k + 1 < count; k += 2)
                        if (w[k] | w[k + 1]) w[k] -= v.raw_bias, w[k + 1] -= v.raw_bias;
                } else {
                    for (size_t k = 0; k < count; ++k) w[k] -= v.raw_bias;
                }
            }
            return;
        }
        void* to = device ? dst : (void*)staging_;
        if (f64) {
            if (v.vec) DataLaunch<R>::pack_vec3_f64(L, v.vec, (double*)to, v.n, v.map);
            else DataLaunch<R>::pack_scalar_f64(L, v.scalar, (double*)to, v.n, v.stride, v.offset, v.map);
        } else {
            if (v.vec) DataLaunch<R>::pack_vec3(L, v.vec, (float*)to, v.n, v.map);
            else DataLaunch<R>::pack_scalar(L, v.scalar, (float*)to, v.n, v.stride, v.offset, v.map);
        }
        HIP_CHECK(hipGetLastError());
        if (!device)
            HIP_CHECK(hipMemcpyAsync(dst, staging_, count * (f64 ? 8 : 4), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void profile_enable(uint32_t every) override { prof_.enable(every, stream_); }

    // Without a block-size option: 256 threads once the batch's fluid particles fill
    // 256 such workgroups, else 128 or 64, so small scenes still spread over the CUs
    // (measured: the 16k dam break 0.49 -> 0.40 ms per step at 64 threads; 64k and
    // larger fastest at 256).
    uint32_t block_for(uint32_t requested) const override {
        if (requested) return requested;
        return A_.n >= 65536 ? 256u : (A_.n >= 32768 ? 128u : 64u);
    }
    void profile_read(pcisph_stage_times* out) override { prof_.read(stream_, out); }

    // Every scene's non-finite flag: gathered into one word array on the device and
    // copied at once (a strided copy of one word per scene took milliseconds for
    // hundreds of scenes).
    std::vector<uint32_t> nonfinite() override {
        std::vector<uint32_t> v(n_.size(), 0);
        if (!v.empty()) {
            if (flags_cap_ < v.size()) {
                dalloc(flags_dev_, v.size(), stream_);
                flags_cap_ = v.size();
            }
            nonfinite_flags(Launch{stream_, 256}, states_, (uint32_t)v.size(), flags_dev_);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(v.data(), flags_dev_, sizeof(uint32_t) * v.size(), hipMemcpyDeviceToHost,
                                     stream_));
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
        return v;
    }

    void wait_for_iteration(uint32_t k, unsigned int tag) {
        auto t0 = std::chrono::steady_clock::now();
        uint64_t spins = 0;
        while (__atomic_load_n(&host_->tag[k], __ATOMIC_ACQUIRE) != tag) {
            _mm_pause();
            if ((++spins & 0x3FF) == 0) {
                const hipError_t q = hipStreamQuery(stream_);
                if (q != hipSuccess && q != hipErrorNotReady)
                    throw Error(PCISPH_ERR_DEVICE, std::string("stream failed: ") + hipGetErrorString(q));
                if (q == hipSuccess && __atomic_load_n(&host_->tag[k], __ATOMIC_ACQUIRE) != tag)
                    throw Error(PCISPH_ERR_INTERNAL,
                                "correction iteration " + std::to_string(k) + " finished without a result");
                const double s =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (s > 120.0) throw Error(PCISPH_ERR_DEVICE, "timed out waiting for the correction loop");
            }
        }
    }

    // The block table for workgroups of `block` threads: every scene's fluid
    // particles in ranges of `block`, then every scene's boundary particles.
    //
    // With `xcd` the fluid ranges are placed for the 8 XCDs, to which workgroups are
    // dispatched round-robin (workgroup w runs on XCD w mod 8): each run of 8 x 16
    // workgroups deals 8 consecutive chunks of 16 ranges to the 8 XCDs, so an XCD's
    // L2 sees compact parts of the spatially coherent order while the work stays
    // interleaved over the XCDs (one contiguous eighth per XCD measured slower: the
    // parts finish at different times).
    void prepare_blocks(uint32_t block, bool xcd) {
        const uint32_t want = block | (xcd ? 0x80000000u : 0u);
        if (blocks_for_ == want) return;
        std::vector<BlockRange> t;
        // fluid ranges carry their index in this order (the staging index of sph_staged.h)
        for (uint32_t s = 0; s < n_.size(); ++s)
            for (uint32_t b = f_off_[s]; b < f_off_[s + 1]; b += block)
                t.push_back({s, b, std::min(b + block, f_off_[s + 1]),
                             (b == f_off_[s] ? kRangeFirstOfScene : 0u) | ((uint32_t)t.size() << kRangeTileShift)});
        const uint32_t fluid_blocks = (uint32_t)t.size();
        if (xcd && fluid_blocks > 8) {
            constexpr uint32_t kXcds = 8, kChunk = 16, kGroup = kXcds * kChunk;
            std::vector<BlockRange> placed(t.begin(), t.end());
            const uint32_t full = fluid_blocks / kGroup * kGroup;  // the tail keeps its order
            for (uint32_t p = 0; p < full; ++p) {
                const uint32_t g = p / kGroup, r = p % kGroup;
                placed[p] = t[g * kGroup + (r % kXcds) * kChunk + r / kXcds];
            }
            std::copy(placed.begin(), placed.end(), t.begin());
        }
        for (uint32_t s = 0; s < nb_.size(); ++s)
            for (uint32_t b = b_off_[s]; b < b_off_[s + 1]; b += block)
                t.push_back({s, b, std::min(b + block, b_off_[s + 1]), kRangeBoundary});
        HIP_CHECK(hipStreamSynchronize(stream_));  // kernels enqueued ahead may still read the old table
        dalloc(block_table_, std::max<size_t>(t.size(), 1), stream_);
        dalloc(running_table_, std::max<size_t>(t.size(), 1), stream_);
        blocks_host_ = t;
        if (!t.empty())
            HIP_CHECK(hipMemcpyAsync(block_table_, t.data(), sizeof(BlockRange) * t.size(), hipMemcpyHostToDevice,
                                     stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        fluid_blocks_ = fluid_blocks;
        boundary_blocks_ = (uint32_t)t.size() - fluid_blocks;
        blocks_for_ = want;
    }

    // The scene table of this step; uploaded when it differs from the last one.
    void prepare_scenes(const std::vector<SceneInput>& scenes, const pcisph_solver_options& o) {
        std::vector<SphParams<R>> t(scenes.size());
        for (size_t s = 0; s < scenes.size(); ++s) {
            const Simulation_Params& p = *scenes[s].params;
            SphParams<R> P = make_sph_params<R>(p, scenes[s].threshold, o);
            P.f_off = f_off_[s];
            P.b_off = b_off_[s];
            P.bucket_base = bucket_base_[s];
            P.p_ref = (float)(P.delta * P.rho0 / R(100));
            P.div_scale = (float)((double)p.particle_mass * 45.0 /
                                  (3.14159265358979323846 * std::pow((double)P.h, 6)) * (double)P.dt * 100.0);
            const float skin = scene_skin(s);
            const float h = (float)P.h, radius = (1.f + skin) * h;
            P.list_r2 = skin > 0.f ? radius * radius : 0.f;
            // near lists serve the passes at x only: no displacement tracking or tests
            P.disp_limit = s < near_.size() && near_[s] ? 0.f : 0.9f * skin * h;
            if (s < boxes_.size()) {
                for (int a = 0; a < 3; ++a) {
                    P.box_lo[a] = boxes_[s].lo[a];
                    P.box_dim[a] = boxes_[s].dim[a];
                }
                P.box_base = boxes_[s].base;
            }
            P.warm = scenes[s].warm ? 1u : 0u;
            t[s] = P;
        }
        if (t.size() == scene_host_.size() &&
            std::memcmp(t.data(), scene_host_.data(), sizeof(SphParams<R>) * t.size()) == 0)
            return;
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (t.size() != scene_host_.size()) dalloc(scene_table_, std::max<size_t>(t.size(), 1), stream_);
        HIP_CHECK(hipMemcpyAsync(scene_table_, t.data(), sizeof(SphParams<R>) * t.size(), hipMemcpyHostToDevice,
                                 stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        scene_host_ = t;
    }

    // The cell boxes of the dense range grids (CellTable): each scene's box covers the
    // cells of its fluid and boundary particles with a margin of two cells -- the
    // boundary's alone when that is larger than kBoxCellsPerParticle cells per
    // particle, none when that is too.  Measured once after positions are set (upload,
    // restore, new counts); particles that later leave the box look their
    // neighbourhoods up in the bucket tables.  Needs the block and scene tables.
    static constexpr uint64_t kBoxCellsPerParticle = 8;
    void rebuild_boxes(const Launch& L, size_t S) {
        std::vector<int> ext(12 * S);
        for (size_t s = 0; s < S; ++s)
            for (int a = 0; a < 3; ++a) {
                ext[12 * s + a] = ext[12 * s + 6 + a] = INT32_MAX;
                ext[12 * s + 3 + a] = ext[12 * s + 9 + a] = INT32_MIN;
            }
        if (ext_cap_ < ext.size()) {
            dalloc(ext_dev_, ext.size(), stream_);
            ext_cap_ = ext.size();
        }
        if (S) {
            HIP_CHECK(hipMemcpyAsync(ext_dev_, ext.data(), sizeof(int) * ext.size(), hipMemcpyHostToDevice, stream_));
            Batch<R> B{};
            B.scenes = scene_table_;
            B.blocks = block_table_;
            B.fluid_blocks = fluid_blocks_;
            B.boundary_blocks = boundary_blocks_;
            DataLaunch<R>::extents(L, B, A_.pos_in, A_.bpos, ext_dev_);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(ext.data(), ext_dev_, sizeof(int) * ext.size(), hipMemcpyDeviceToHost, stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
        }
        boxes_.assign(S, SceneBox{});
        uint64_t total = 0;
        uint32_t largest = 0;
        for (size_t s = 0; s < S; ++s) {
            const uint64_t limit = kBoxCellsPerParticle * ((uint64_t)n_[s] + nb_[s]) + 4096;
            auto fit = [&](bool fluid, SceneBox& b) {
                int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
                bool any = false;
                for (int part = fluid ? 0 : 1; part < 2; ++part) {
                    if ((part == 0 ? n_[s] : nb_[s]) == 0) continue;
                    any = true;
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min<int64_t>(lo[a], ext[12 * s + 6 * part + a]);
                        hi[a] = std::max<int64_t>(hi[a], ext[12 * s + 6 * part + 3 + a]);
                    }
                }
                if (!any) return false;
                uint64_t cells = 1;
                for (int a = 0; a < 3; ++a) {
                    const uint64_t d = (uint64_t)(hi[a] - lo[a] + 5);
                    cells *= d;
                    if (cells > limit) return false;
                    b.lo[a] = (uint32_t)(lo[a] - 2);  // cell coordinates as 32-bit words
                    b.dim[a] = (uint32_t)d;
                }
                return total + cells <= 0xFFFFFFFFull;
            };
            SceneBox b{};
            if (fit(true, b) || fit(false, b)) {
                const uint64_t cells = (uint64_t)b.dim[0] * b.dim[1] * b.dim[2];
                b.base = (uint32_t)total;
                total += cells;
                largest = std::max(largest, (uint32_t)cells);
                boxes_[s] = b;
            }
        }
        if (total > box_cap_) {
            HIP_CHECK(hipStreamSynchronize(stream_));
            dalloc(A_.fgrid, total, stream_);
            dalloc(A_.bgrid, total, stream_);
            box_cap_ = total;
        }
        A_.box_cells = (uint32_t)total;
        box_max_ = largest;
        boxes_dirty_ = false;
        bgrid_stale_ = true;
    }

    // Captures the launches `fn` makes on the stream into an executable graph.
    template <typename F>
    hipGraphExec_t capture(F&& fn) {
        HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            fn();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t e = nullptr;
        const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_CHECK(rc);
        return e;
    }

    void drop_graphs() {
        if (prologue_graph_) (void)hipGraphExecDestroy(prologue_graph_);
        prologue_graph_ = nullptr;
        for (hipGraphExec_t e : loop_graphs_)
            if (e) (void)hipGraphExecDestroy(e);
        loop_graphs_.clear();
        graph_key_.clear();
    }

    // Every value a recorded launch bakes in; graphs are re-captured when it changes.
    std::vector<uint64_t> launch_key(const Launch& L, bool fast, bool warm, bool features, uint32_t digit,
                                     uint32_t key_bits, uint32_t maxit, uint32_t chunk, uint32_t layout_bits,
                                     uint32_t morton_bits, const float* mlp, bool staged,
                                     bool all_features) const {
        std::vector<uint64_t> k;
        auto put = [&](uint64_t v) { k.push_back(v); };
        auto ptr = [&](const void* p) { k.push_back((uint64_t)(uintptr_t)p); };
        put(L.block), put(L.onesweep), put(fast), put(warm), put(features), put(all_features), put(digit), put(key_bits), put(maxit),
            put(chunk);
        put(staged), ptr(A_.stage.work), ptr(A_.stage.rows), ptr(A_.stage.src), ptr(A_.stage.info);
        ptr(A_.stage.bwork), ptr(A_.stage.brows), ptr(A_.stage.binfo);
        put(A_.stage.cap);
        put(layout_bits), put(morton_bits);
        put(A_.tile_pressure);
        put(A_.n), put(A_.nb), put(A_.B);
        for (const void* p : {(const void*)A_.pos_in, (const void*)A_.vel_in, (const void*)A_.pos,
                              (const void*)A_.vel, (const void*)A_.xr, (const void*)A_.nrm, (const void*)A_.fo,
                              (const void*)A_.fp, (const void*)A_.xp, (const void*)A_.xf, (const void*)A_.p,
                              (const void*)A_.dv, (const void*)A_.counts, (const void*)A_.keys[0],
                              (const void*)A_.keys[1], (const void*)A_.vals[0], (const void*)A_.vals[1],
                              (const void*)A_.cells, (const void*)A_.bpos, (const void*)A_.bxf, (const void*)A_.bp,
                              (const void*)A_.binit, (const void*)A_.bcells, (const void*)sort_scratch_,
                              (const void*)feat_, (const void*)warm_rows_, (const void*)mlp, (const void*)ctl_,
                              (const void*)dev_host_,
                              (const void*)dev_results_, (const void*)D_.xp0, (const void*)D_.xf0,
                              (const void*)D_.p0, (const void*)D_.dv0, (const void*)D_.bp0, (const void*)D_.bxf0})
            ptr(p);
        const Batch<R>& B = A_.batch;
        ptr(B.scenes), ptr(B.blocks), ptr(B.states), ptr(okeys_[0]), ptr(okeys_[1]), ptr(ovals_[0]),
            ptr(ovals_[1]), ptr(canon_), ptr(stor_), ptr(srcc_), ptr(code_);
        put(B.scene_count), put(B.fluid_blocks), put(B.boundary_blocks), put(B.n_total), put(B.nb_total);
        const NeighborLists& NL = A_.NL;
        ptr(NL.f), ptr(NL.b), ptr(NL.bf), ptr(NL.fcount), ptr(NL.bcount), ptr(NL.bfcount);
        ptr(NL.disp[0]), ptr(NL.disp[1]), ptr(D_.disp_next);
        ptr(B.fgrid), ptr(B.bgrid), put(A_.box_cells), put(box_max_);
        put(NL.cap_f), put(NL.cap_b), put(NL.cap_bf), put(NL.on), put(NL.buckets), put(D_.buckets);
        return k;
    }

    void step(const std::vector<SceneInput>& scenes, const pcisph_solver_options& o, bool sort_boundary, bool fast,
              const WarmArgs& w, std::vector<StepResult>& out, uint64_t* samples) override {
        HIP_CHECK(hipSetDevice(device_));
        const uint32_t S = (uint32_t)scenes.size();
        const Launch L{stream_, block_for(o.block_size), (o.work_order & 128) != 0, &dev_host_->sort_abort};
        const MlpLaunch ML{stream_};
        const uint32_t digit = o.sort_digit_bits ? o.sort_digit_bits : 8u;
        const uint32_t key_bits = key_bits_for(A_.B);
        const unsigned int tag = ++tag_ ? tag_ : ++tag_;
        const bool warm = w.params != nullptr;
        const bool features = warm || w.collect_stride;
        const uint32_t maxit = o.max_iterations, minit = std::max(o.min_iterations, 1u);
        prepare_blocks(L.block, (o.work_order & 2) != 0);
        setup_lists(o.neighbor_skin);
        prepare_scenes(scenes, o);
        if (boxes_dirty_ || boxes_.size() != S) {
            rebuild_boxes(L, S);
            prepare_scenes(scenes, o);
        }
        // Order codes (order_code): `obits` bits per axis, the outside-the-box bit below
        // the scene index.  The Morton layout's bucket codes take cells outside the box
        // modulo 2^obits; the staged pressure pass's work order takes them in blocks of
        // 8 cells along each axis (2^(obits+2) cells either side of the box), which keeps
        // clumps thrown clear of the scene together at the same key width
        const uint32_t sbits = key_bits_for(S);
        const uint32_t obits = std::min(7u, sbits < 31 ? (31u - sbits) / 3u : 0u);
        const uint32_t order_bits = obits | (obits << 8) | (3u << 16);
        // the layout pays for its second sort only where the data outgrows the caches:
        // measured 15% slower at 64k fluid particles, 3% faster at 256k, 20% at 1M
        const bool layout =
            (o.work_order & 1) && obits >= 2 && A_.n > 1 && (A_.n >= kLayoutMinParticles || (o.work_order & 8));
        const uint32_t layout_bits = layout ? 3 * obits + 1 + sbits : 0;
        if (!layout) canonicalize();
        if (layout) ensure_order_buffers();
        A_.batch = Batch<R>{scene_table_, block_table_, states_, S, fluid_blocks_, boundary_blocks_,
                            A_.n, A_.nb, 0, A_.box_cells ? A_.fgrid : nullptr,
                            A_.box_cells ? A_.bgrid : nullptr};
        A_.tile_pressure = (o.work_order & 4) != 0;
        // The FP32 fluid pressure pass from the step's LDS staging: where no scene's
        // pressure passes may walk their lists this step (work_order bit 4 forces it,
        // bit 5 switches it off)
        bool all_buckets = true;
        for (size_t s = 0; s < scene_host_.size(); ++s)
            all_buckets = all_buckets && (!A_.NL.on || !(scene_host_[s].disp_limit > 0.f) || !n_[s]);
        const bool staged = std::is_same<R, float>::value && A_.n > 0 && !(o.work_order & 32) &&
                            ((o.work_order & 16) || all_buckets);
        const uint32_t stage_bits = 3 * obits + 1 + sbits;
        A_.staged_pressure = staged;
        if (staged) {
            ensure_stage_buffers(fluid_blocks_);
            A_.stage = StageMaps{stage_vals_[radix_sort_result(A_.n, stage_bits, digit)],
                                 stage_rows_,
                                 stage_src_,
                                 stage_info_,
                                 stage_brows_,
                                 stage_binfo_,
                                 bstage_vals_[radix_sort_result(A_.nb, stage_bits, digit)],
                                 o.stage_cap ? o.stage_cap : kStageCap};
        }
        if (features) ensure_warm_buffers(warm);
        if (warm) upload_warm_rows(scenes);
        D_.disp_next = A_.NL.on ? A_.NL.disp[1] : nullptr;
        D_.buckets = A_.NL.buckets;
        if (!ctl_) dalloc(ctl_, 64, stream_);
        // graphs record launches; a timed step launches them one by one (and keeps the
        // graphs for the untimed ones)
        const bool profiled = prof_.begin_step();
        const uint32_t chunk = profiled ? 0u : o.graph_iterations;
        // the correction loop as one cooperative launch (work_order bit 6): not in a
        // step the stage profiler times pass by pass, nor with the network's dual first
        // iteration or the tiled pressure pass
        const bool persistent = (o.work_order & 64) && !profiled && !warm && !A_.tile_pressure;
        if (chunk) {
            std::vector<uint64_t> key =
                launch_key(L, fast, warm, features, digit, key_bits, maxit, chunk, layout_bits, order_bits, w.params, staged,
                           w.collect_stride != 0);
            if (key != graph_key_) {
                drop_graphs();
                graph_key_ = key;
            }
        } else if (!o.graph_iterations && !graph_key_.empty()) {
            drop_graphs();
        }
        // this step's loop records and tag (outside any graph: the tag changes per step)
        fill_u64(L, (unsigned long long*)states_, sizeof(LoopState) / 8 * (size_t)S, 0ull);
        fill_u32(L, ctl_, 1, tag);
        if (sort_boundary && A_.nb) {
            prof_.timed(PCISPH_STAGE_BOUNDARY_SORT, stream_, -1, [&] {
                fill_u32(L, (uint32_t*)A_.bcells, 2 * (size_t)A_.B, 0u);
                DataLaunch<R>::hash(L, A_.batch, true, A_.bpos, A_.bkeys[0], A_.bvals[0]);
                boundary_sorted_ = radix_sort(L, A_.bkeys, A_.bvals, A_.nb, key_bits, digit, sort_scratch_);
                if (boundary_sorted_ < 0) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "too many boundary particles");
                DataLaunch<R>::reorder(L, A_.nb, A_.bkeys[boundary_sorted_], A_.bvals[boundary_sorted_],
                                       A_.bpos, nullptr, A_.bpos_tmp, nullptr, A_.bcells);
                HIP_CHECK(hipMemcpyAsync(A_.bpos, A_.bpos_tmp, sizeof(V4<R>) * A_.nb, hipMemcpyDeviceToDevice,
                                         stream_));
                if (o.launch_boundary_init) launch(fast).boundary_init(L, A_);
            });
        }
        if (A_.box_cells && A_.nb && (bgrid_stale_ || sort_boundary)) {
            DataLaunch<R>::fill_box(L, scene_table_, S, box_max_, A_.bcells, A_.bgrid, A_.bpos);
            HIP_CHECK(hipGetLastError());
            bgrid_stale_ = false;
        }
        if (sort_boundary) bwork_key_.clear();
        if (staged && A_.nb) {
            const std::vector<uint64_t> bk{digit, stage_bits, A_.nb, (uint64_t)(uintptr_t)A_.stage.bwork};
            if (bk != bwork_key_) {
                const int r = DataLaunch<R>::boundary_order(L, A_, order_bits, stage_bits, digit, sort_scratch_,
                                                            bstage_keys_, bstage_vals_);
                if (r < 0 || bstage_vals_[r] != A_.stage.bwork)
                    throw Error(PCISPH_ERR_INTERNAL, "boundary work order not where expected");
                bwork_key_ = bk;
            }
        }
        const LaunchTable ops = launch(fast);
        // everything before the correction loop
        auto prologue = [&] {
            prof_.timed(PCISPH_STAGE_HASH, stream_, -1, [&] {
                fill_u32(L, (uint32_t*)A_.cells, 2 * (size_t)A_.B, 0u);
                DataLaunch<R>::hash(L, A_.batch, false, A_.pos_in, A_.keys[0], A_.vals[0], layout ? canon_ : nullptr);
            });
            prof_.timed(PCISPH_STAGE_SORT, stream_, -1, [&] {
                fluid_sorted_ = radix_sort(L, A_.keys, A_.vals, A_.n, key_bits, digit, sort_scratch_);
            });
            if (fluid_sorted_ < 0) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "too many particles for one sort");
            prof_.timed(PCISPH_STAGE_REORDER, stream_, -1, [&] {
                if (layout) {
                    LayoutArgs<R> a{};
                    a.K = A_.keys[fluid_sorted_];
                    a.S = A_.vals[fluid_sorted_];
                    a.canon = canon_;
                    a.stor = stor_;
                    a.srcc = srcc_;
                    a.code = code_;
                    for (int k = 0; k < 2; ++k) {
                        a.okeys[k] = okeys_[k];
                        a.ovals[k] = ovals_[k];
                    }
                    a.bits = obits | (obits << 8);  // outside the box: modulo 2^obits
                    a.key_bits = layout_bits;
                    a.digit_bits = digit;
                    a.scratch = sort_scratch_;
                    a.pin = A_.pos_in;
                    a.vin = A_.vel_in;
                    a.pout = A_.pos;
                    a.vout = A_.vel;
                    a.cells = A_.cells;
                    if (DataLaunch<R>::layout(L, A_.batch, a) < 0)
                        throw Error(PCISPH_ERR_INTERNAL, "layout sort failed");
                } else {
                    DataLaunch<R>::reorder(L, A_.n, A_.keys[fluid_sorted_], A_.vals[fluid_sorted_], A_.pos_in,
                                           A_.vel_in, A_.pos, A_.vel, A_.cells);
                }
                if (A_.box_cells) DataLaunch<R>::fill_box(L, scene_table_, S, box_max_, A_.cells, A_.fgrid);
                if (staged)
                    DataLaunch<R>::stage_setup(L, A_, order_bits, stage_bits, digit, sort_scratch_, stage_keys_,
                                               stage_vals_);
            });
            prof_.timed(PCISPH_STAGE_DENSITY, stream_, -1, [&] { ops.density(L, A_); });
            prof_.timed(PCISPH_STAGE_NORMAL, stream_, -1, [&] { ops.normal(L, A_); });
            FeatureArgs F{features ? feat_ : nullptr, A_.counts, w.collect_stride ? 1u : 0u};
            prof_.timed(PCISPH_STAGE_FORCE_INIT, stream_, -1, [&] { ops.force_init(L, A_, F); });
            if (warm) {
                // network start: p from the features, its pressure force, and a prediction
                // with it; the plain prediction is kept for the dual first iteration
                prof_.timed(PCISPH_STAGE_WARM_START, stream_, -1, [&] {
                    mlp_warm_start<R>(ML, w.params, feat_, warm_rows_, A_.n, A_.p);
                    DataLaunch<R>::warm_apply(L, A_);
                    ops.pressure_force(L, A_, 0, dev_host_, dev_results_, ctl_, 1, D_.xp0);
                });
            }
        };
        // Batches: once at least half of the workgroups belong to scenes whose loops
        // have ended (they return at once, but every launch still dispatches them), the
        // iterations enqueued from then on are launched one by one over a block table
        // of the running scenes alone (AR), re-cut when it halves again.
        const DeviceArrays<R>* AI = &A_;
        DeviceArrays<R> AR = A_;
        size_t table_blocks = blocks_host_.size();
        bool recut = false;
        auto keep_running = [&]() {
            // the host copies of earlier tables: the stream has run past their copies
            // once an iteration of this step has been seen
            if (!recut) running_host_.clear();
            recut = true;
            std::vector<BlockRange> t;
            uint32_t nf = 0;
            for (const BlockRange& b : blocks_host_)
                if (results_[b.scene].running) {
                    t.push_back(b);
                    nf += (b.flags & kRangeBoundary) ? 0u : 1u;
                }
            if (t.empty() || t.size() * 2 > table_blocks) return;
            // stream-ordered: iterations enqueued before read the previous table first;
            // the host copy stays alive until the copy has run
            running_host_.push_back(std::move(t));
            const std::vector<BlockRange>& h = running_host_.back();
            HIP_CHECK(hipMemcpyAsync(running_table_, h.data(), sizeof(BlockRange) * h.size(),
                                     hipMemcpyHostToDevice, stream_));
            AR.batch.blocks = running_table_;
            AR.batch.fluid_blocks = nf;
            AR.batch.boundary_blocks = (uint32_t)h.size() - nf;
            AI = &AR;
            table_blocks = h.size();
        };
        auto iteration = [&](uint32_t it) {
            const DeviceArrays<R>& A = *AI;
            prof_.timed(PCISPH_STAGE_PRESSURE, stream_, (int)it, [&] {
                // iteration 0 of a warm-started step: the dual pass for the scenes started
                // from the network, the plain pass for the others
                if (warm && it == 0) ops.pressure_dual(L, A, D_);
                ops.pressure(L, A, it);
            });
            prof_.timed(PCISPH_STAGE_PRESSURE_FORCE, stream_, (int)it, [&] {
                ops.pressure_force(L, A, it, dev_host_, dev_results_, ctl_, 0, nullptr);
            });
        };
        if (chunk) {
            if (!prologue_graph_) {
                prologue_graph_ = capture(prologue);
                prologue_sorted_ = fluid_sorted_;
            }
            fluid_sorted_ = prologue_sorted_;
            HIP_CHECK(hipGraphLaunch(prologue_graph_, stream_));
        } else {
            prologue();
        }
        HIP_CHECK(hipGetLastError());
        layout_active_ = layout;
        sorted_layout_ = layout;

        // Iterations are enqueued `step` at a time (one graph, or one iteration), keeping
        // at least one iteration queued beyond the one the host waits for.
        const uint32_t first = std::min(minit - 1, maxit - 1);
        const uint32_t lookahead = 1;
        const uint32_t step_iters = chunk ? chunk : 1u;
        uint32_t enq = 0;
        auto enqueue = [&]() {
            if (chunk && AI == &A_) {
                const uint32_t c = enq / chunk;
                if (loop_graphs_.size() <= c) loop_graphs_.resize(c + 1, nullptr);
                if (!loop_graphs_[c])
                    loop_graphs_[c] = capture([&] {
                        for (uint32_t it = enq; it < std::min(enq + chunk, maxit); ++it) iteration(it);
                    });
                HIP_CHECK(hipGraphLaunch(loop_graphs_[c], stream_));
                enq = std::min(enq + step_iters, maxit);
            } else {
                iteration(enq);
                ++enq;
            }
        };
        if (persistent) {
            if (!sync_) dalloc(sync_, 4, stream_);
            fill_u32(L, sync_, 4, 0u);
            const hipError_t e =
                ops.correction_loop(L, A_, maxit, dev_host_, dev_results_, ctl_, GridSync{sync_, sync_timeout_});
            if (e != hipSuccess)
                throw Error(PCISPH_ERR_DEVICE, std::string("correction loop launch: ") + hipGetErrorString(e));
            enq = maxit;  // every iteration is in that launch
        }
        while (enq < maxit && enq <= first + lookahead) enqueue();
        HIP_CHECK(hipGetLastError());
        uint32_t k = first;
        for (;;) {
            wait_for_iteration(k, tag);
            if (host_->running[k] == 0) break;
            if (k + 1 >= maxit) throw Error(PCISPH_ERR_INTERNAL, "scenes still running after the last iteration");
            if (S > 1 && !persistent) keep_running();
            ++k;
            if (enq < maxit && enq <= k + lookahead) enqueue();
        }
        HIP_CHECK(hipGetLastError());
        if (host_->sort_abort) {  // the sorts ran before the loop the host has just seen end
            host_->sort_abort = 0;
            throw Error(PCISPH_ERR_DEVICE, "one-sweep radix sort: a look-back wait expired; the step is invalid");
        }
        uint64_t collected = 0;
        if (w.collect_stride) {
            const uint32_t phase = w.collect_phase % w.collect_stride;
            uint64_t base = w.store_base;
            for (uint32_t s = 0; s < S; ++s) {
                const uint32_t n = n_[s];
                if (!n || base >= w.store_cap) continue;
                const float p_ref = scene_host_[s].p_ref;
                mlp_collect<R>(ML, feat_ + (size_t)f_off_[s] * 8, A_.p + f_off_[s], n, w.collect_stride, phase, p_ref,
                               w.store_x, w.store_t, base, w.store_cap);
                const uint64_t c = n > phase ? (n - phase + w.collect_stride - 1) / w.collect_stride : 0;
                base += c;
                collected += c;
            }
            HIP_CHECK(hipGetLastError());
        }
        if (samples) *samples = collected;
        prof_.end_step((int)k);
        prof_.maybe_resolve(stream_);
        out.assign(S, StepResult{});
        for (uint32_t s = 0; s < S; ++s) {
            if (!n_[s]) continue;
            const volatile SceneResult& q = results_[s];
            StepResult& r = out[s];
            r.iterations = q.iterations;
            const R mdv = Bits<R>::to(q.max_dv);
            const R rho0 = scene_host_[s].rho0;
            r.max_dv = (double)mdv;
            r.max_expansion = (double)Bits<R>::to(q.max_exp);
            r.converged = (r.iterations >= minit) && (mdv / rho0 < scene_host_[s].threshold);
            r.fluid_candidates = q.candidates[0];
            r.boundary_candidates = q.candidates[1];
            r.warm = warm && scenes[s].warm;
            r.warm_rejected = r.warm && q.warm_rejected;
            r.first_dv = (double)Bits<R>::to(q.first_dv);
            r.list_waves = q.list_waves;
        }
        update_list_radius(out);
    }

    // The list radius of each scene.  Lists serve the passes at x whatever their
    // radius, and the pressure passes only where particles stay within the skin; a
    // scene whose pressure passes walked their buckets (lists in under 1% of their
    // fluid wavefronts) in a step builds lists of radius (1 + kNearSkin) h -- fewer entries
    // to write and walk -- and neither tracks displacements nor tests its lists in
    // the pressure passes, until a probe step every kProbeEvery steps with the full
    // radius finds them used again.  Results do not depend on the radius.
    static constexpr float kNearSkin = 0.01f;
    static constexpr uint32_t kProbeEvery = 32;
    float scene_skin(size_t s) const {
        if (!(list_skin_ > 0.f)) return 0.f;
        return s < near_.size() && near_[s] ? std::min(list_skin_, kNearSkin) : list_skin_;
    }
    void update_list_radius(const std::vector<StepResult>& out) {
        near_.resize(out.size(), 0);
        near_steps_.resize(out.size(), 0);
        if (!A_.NL.on) return;
        for (size_t s = 0; s < out.size(); ++s) {
            if (!n_[s] || !out[s].iterations) continue;
            const uint64_t waves = (uint64_t)(n_[s] + 63) / 64 * (uint64_t)out[s].iterations;
            const bool used = 100ull * out[s].list_waves >= waves;
            if (!near_[s]) {
                if (!used && list_skin_ > kNearSkin) {
                    near_[s] = 1;
                    near_steps_[s] = 0;
                }
            } else if (++near_steps_[s] >= kProbeEvery) {
                near_[s] = 0;  // the next step probes the full radius
            }
        }
    }

    // Neighbour-list storage (capacities: fluid 128, boundary 96 per fluid particle,
    // fluid 96 per boundary particle, in chunks of four 24-bit entries = three words);
    // each scene's list radius and validity limit follow from its kernel radius (scene
    // table).  Batches of 2^24 or more fluid or boundary particles step without lists
    // (entries would not fit; results do not depend on lists).
    void setup_lists(float skin) {
        NeighborLists& NL = A_.NL;
        list_skin_ = skin > 0.f ? skin : 0.f;
        if (!(skin > 0.f) || A_.n >= (1u << 24) || A_.nb >= (1u << 24)) {
            NL.on = 0;
            return;
        }
        constexpr uint32_t kCapF = 128, kCapB = 96, kCapBF = 96;
        auto words = [](uint32_t cap, uint32_t n) { return (size_t)cap / 4 * 3 * std::max(n, 1u); };
        if (list_n_ != A_.n || list_nb_ != A_.nb || !NL.f) {
            dalloc(NL.f, words(kCapF, A_.n), stream_);
            dalloc(NL.b, words(kCapB, A_.n), stream_);
            dalloc(NL.bf, words(kCapBF, A_.nb), stream_);
            dalloc(NL.fcount, std::max(A_.n, 1u), stream_);
            dalloc(NL.bcount, std::max(A_.n, 1u), stream_);
            dalloc(NL.bfcount, std::max(A_.nb, 1u), stream_);
            NL.cap_f = kCapF;
            NL.cap_b = kCapB;
            NL.cap_bf = kCapBF;
            list_n_ = A_.n;
            list_nb_ = A_.nb;
        }
        if (list_buckets_ != A_.B || !NL.disp[0]) {
            dalloc(NL.disp[0], std::max(A_.B, 1u), stream_);
            dalloc(NL.disp[1], std::max(A_.B, 1u), stream_);
            list_buckets_ = A_.B;
        }
        NL.buckets = A_.B;
        NL.on = 1;
    }

    static constexpr uint32_t kLayoutMinParticles = 1u << 17;

    // Buffers of the pressure pass's staging for `tiles` fluid ranges.
    void ensure_stage_buffers(uint32_t tiles) {
        if (stage_n_ != A_.n || !stage_rows_) {
            for (int k = 0; k < 2; ++k) {
                dalloc(stage_keys_[k], std::max(A_.n, 1u), stream_);
                dalloc(stage_vals_[k], std::max(A_.n, 1u), stream_);
            }
            dalloc(stage_rows_, 9 * (size_t)std::max(A_.n, 1u), stream_);
            dalloc(stage_brows_, 9 * (size_t)std::max(A_.n, 1u), stream_);
            stage_n_ = A_.n;
        }
        if (stage_tiles_ < tiles || !stage_src_) {
            dalloc(stage_src_, (size_t)std::max(tiles, 1u) * kStageCap, stream_);
            dalloc(stage_info_, std::max(tiles, 1u), stream_);
            dalloc(stage_binfo_, std::max(tiles, 1u), stream_);
            stage_tiles_ = tiles;
        }
        if (stage_nb_ != A_.nb || !bstage_vals_[0]) {
            for (int k = 0; k < 2; ++k) {
                dalloc(bstage_keys_[k], std::max(A_.nb, 1u), stream_);
                dalloc(bstage_vals_[k], std::max(A_.nb, 1u), stream_);
            }
            stage_nb_ = A_.nb;
            bwork_key_.clear();
        }
    }

    // Buffers of the Morton layout: its sort, the bucket codes, the reference's
    // source locations.  The slot maps canon_ / stor_ live with the fluid arrays.
    void ensure_order_buffers() {
        if (order_n_ != A_.n || !okeys_[0]) {
            for (int k = 0; k < 2; ++k) {
                dalloc(okeys_[k], std::max(A_.n, 1u), stream_);
                dalloc(ovals_[k], std::max(A_.n, 1u), stream_);
            }
            dalloc(srcc_, std::max(A_.n, 1u), stream_);
            order_n_ = A_.n;
        }
        if (code_b_ != A_.B || !code_) {
            dalloc(code_, std::max(A_.B, 1u), stream_);
            code_b_ = A_.B;
        }
    }

    // The public fluid arrays (and this step's per-particle results) back in
    // canonical order, identity slot maps.
    void canonicalize() {
        if (!layout_active_) return;
        const Launch L{stream_, 256};
        const uint32_t n = A_.n;
        auto permute = [&](void* p, uint32_t words) {
            if (!p) return;
            gather_words(L, n, words, stor_, (const uint32_t*)p, (uint32_t*)A_.pos);  // A_.pos: scratch here
            HIP_CHECK(hipMemcpyAsync(p, A_.pos, (size_t)n * words * 4, hipMemcpyDeviceToDevice, stream_));
        };
        constexpr uint32_t v4 = sizeof(V4<R>) / 4, r1 = sizeof(R) / 4;
        for (V4<R>* q : {A_.pos_in, A_.vel_in, A_.xr, A_.nrm, A_.fo, A_.fp, A_.xp, A_.xf}) permute(q, v4);
        for (R* q : {A_.p, A_.dv}) permute(q, r1);
        permute(A_.counts, 4);
        fill_iota(L, canon_, n);
        fill_iota(L, stor_, n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(stream_));
        layout_active_ = false;
    }

    void ensure_warm_buffers(bool dual) {
        const uint32_t n = A_.n, nb = A_.nb;
        if (feat_n_ != n) {
            dalloc(feat_, 8 * (size_t)std::max(n, 1u), stream_);
            feat_n_ = n;
        }
        if (dual && (dual_n_ != n || dual_nb_ != nb)) {
            dalloc(D_.xp0, n, stream_);
            dalloc(D_.xf0, n, stream_);
            dalloc(D_.p0, n, stream_);
            dalloc(D_.dv0, n, stream_);
            dalloc(D_.bp0, std::max(nb, 1u), stream_);
            dalloc(D_.bxf0, std::max(nb, 1u), stream_);
            dual_n_ = n;
            dual_nb_ = nb;
        }
    }

    // The fluid rows of the scenes the step starts from the network, as the inference
    // kernel reads them (mlp.h): adjacent scenes merged into one range; uploaded when
    // they differ from the last step's.
    void upload_warm_rows(const std::vector<SceneInput>& scenes) {
        std::vector<uint32_t> t{0, 0};
        for (size_t s = 0; s < scenes.size() && s < n_.size(); ++s) {
            if (!scenes[s].warm || !n_[s]) continue;
            const uint32_t b = f_off_[s], e = b + n_[s];
            if (t.size() > 2 && t[t.size() - 2] == b) {
                t[t.size() - 2] = e;
            } else {
                t.insert(t.end(), {b, e, 0u});
                ++t[0];
            }
        }
        uint32_t tiles = 0;
        for (uint32_t k = 0; k < t[0]; ++k) {
            t[2 + 3 * k + 2] = tiles;
            tiles += (t[2 + 3 * k + 1] - t[2 + 3 * k] + 15) / 16;
        }
        t[1] = tiles;
        if (t == warm_rows_host_) return;
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (t.size() > warm_rows_cap_) {
            dalloc(warm_rows_, t.size(), stream_);
            warm_rows_cap_ = t.size();
        }
        HIP_CHECK(hipMemcpyAsync(warm_rows_, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        warm_rows_host_ = t;
    }

   private:
    struct LaunchTable {
        void (*density)(const Launch&, const DeviceArrays<R>&);
        void (*normal)(const Launch&, const DeviceArrays<R>&);
        void (*force_init)(const Launch&, const DeviceArrays<R>&, const FeatureArgs&);
        void (*pressure)(const Launch&, const DeviceArrays<R>&, uint32_t);
        void (*pressure_force)(const Launch&, const DeviceArrays<R>&, uint32_t, HostMirror*, SceneResult*,
                               const unsigned int*, int, V4<R>*);
        void (*boundary_init)(const Launch&, const DeviceArrays<R>&);
        void (*pressure_dual)(const Launch&, const DeviceArrays<R>&, const DualArrays<R>&);
        hipError_t (*correction_loop)(const Launch&, const DeviceArrays<R>&, uint32_t, HostMirror*, SceneResult*,
                                      const unsigned int*, const GridSync&);
    };
    static LaunchTable launch(bool fast);

    int device_;
    hipStream_t stream_;
    DeviceArrays<R> A_;
    DualArrays<R> D_{};
    // scene layout: counts, first particle, bucket counts and first bucket per scene
    std::vector<uint32_t> n_{0}, nb_{0}, f_off_{0, 0}, b_off_{0, 0}, buckets_, bucket_base_{0};
    std::vector<SphParams<R>> scene_host_;
    SphParams<R>* scene_table_ = nullptr;
    BlockRange* block_table_ = nullptr;
    BlockRange* running_table_ = nullptr;  // the ranges of the scenes still running (batches)
    unsigned int* sync_ = nullptr;          // the correction loop kernel's barrier words
    unsigned long long sync_timeout_ = 0;   // its longest wait, wall-clock ticks
    std::vector<std::vector<BlockRange>> running_host_;  // its uploads of the current step
    uint32_t* flags_dev_ = nullptr;  // per-scene non-finite flags, gathered
    size_t flags_cap_ = 0;
    std::vector<BlockRange> blocks_host_;
    uint32_t blocks_for_ = 0, fluid_blocks_ = 0, boundary_blocks_ = 0;
    LoopState* states_ = nullptr;
    uint32_t states_n_ = 0;
    volatile SceneResult* results_ = nullptr;
    SceneResult* dev_results_ = nullptr;
    int fluid_sorted_ = 0, boundary_sorted_ = 0;
    struct SceneBox {
        uint32_t lo[3], dim[3], base;
    };
    std::vector<SceneBox> boxes_;           // per scene; dim[0] == 0: no box
    bool boxes_dirty_ = true, bgrid_stale_ = true;
    int* ext_dev_ = nullptr;
    size_t ext_cap_ = 0;
    uint64_t box_cap_ = 0;                  // grid entries allocated
    uint32_t box_max_ = 0;                  // cells of the largest box
    uint32_t* sort_scratch_ = nullptr;
    size_t sort_scratch_words_ = 0;
    float* staging_ = nullptr;
    size_t staging_words_ = 0;
    float* feat_ = nullptr;
    uint32_t* warm_rows_ = nullptr;  // the inference kernel's row ranges
    size_t warm_rows_cap_ = 0;
    std::vector<uint32_t> warm_rows_host_;
    uint32_t feat_n_ = 0, dual_n_ = 0, dual_nb_ = 0, list_n_ = 0, list_nb_ = 0;
    float list_skin_ = 0.f;
    std::vector<uint8_t> near_;          // per scene: lists of radius (1 + kNearSkin) h this step
    std::vector<uint32_t> near_steps_;   // steps since the scene went to near lists
    // staging of the FP32 pressure pass (sph_staged.h): work-order sort, row
    // descriptors [9][n] (fluid and boundary rows), staging maps [fluid ranges][kStageCap], staged counts
    uint32_t* stage_keys_[2] = {nullptr, nullptr};
    uint32_t* stage_vals_[2] = {nullptr, nullptr};
    uint32_t* stage_rows_ = nullptr;
    uint32_t* stage_src_ = nullptr;
    uint32_t* stage_info_ = nullptr;
    uint32_t* stage_brows_ = nullptr;
    uint32_t* stage_binfo_ = nullptr;
    // the boundary work order of the staged pass, sorted when the boundary is (or the
    // sort's digit width or code bits change): keys and values of its sort
    uint32_t* bstage_keys_[2] = {nullptr, nullptr};
    uint32_t* bstage_vals_[2] = {nullptr, nullptr};
    std::vector<uint64_t> bwork_key_;
    uint32_t stage_n_ = 0, stage_tiles_ = 0, stage_nb_ = 0;
    uint32_t* okeys_[2] = {nullptr, nullptr};  // layout sort: bucket codes
    uint32_t* ovals_[2] = {nullptr, nullptr};  // and canonical indices (the result is the layout)
    uint32_t* canon_ = nullptr;                // slot -> canonical index of the public fluid arrays
    uint32_t* stor_ = nullptr;                 // canonical index -> slot
    uint32_t* srcc_ = nullptr;                 // the reference's source locations of the last layout step
    uint32_t* code_ = nullptr;                 // layout code of every bucket of the batch
    uint2* ccells_ = nullptr;                  // canonical bucket ranges, built for downloads
    uint32_t code_b_ = 0, ccells_b_ = 0;
    bool layout_active_ = false;               // the public fluid arrays are in a Morton layout
    bool sorted_layout_ = false;               // the last sort's products come from a layout step
    uint32_t order_n_ = 0;
    uint32_t list_buckets_ = 0;
    uint32_t* ctl_ = nullptr;  // device word: the tag of the step in flight
    hipGraphExec_t prologue_graph_ = nullptr;
    int prologue_sorted_ = 0;
    std::vector<hipGraphExec_t> loop_graphs_;  // [c]: iterations c*chunk .. (c+1)*chunk-1
    std::vector<uint64_t> graph_key_;
    float* shift_host_ = nullptr;  // pinned staging of pcisph_translate_boundaries
    float* shift_dev_ = nullptr;
    size_t shift_cap_ = 0;
    hipEvent_t shift_event_ = nullptr;
    V4<R>* snap_pos_ = nullptr;
    V4<R>* snap_vel_ = nullptr;
    V4<R>* snap_b_ = nullptr;
    volatile HostMirror* host_ = nullptr;
    HostMirror* dev_host_ = nullptr;
    unsigned int tag_ = 0;

   public:
    StageProfiler prof_;
};

template <>
Engine<float>::LaunchTable Engine<float>::launch(bool fast) {
    if (fast)
        return {SphLaunch<float, true>::density, SphLaunch<float, true>::normal, SphLaunch<float, true>::force_init,
                SphLaunch<float, true>::pressure, SphLaunch<float, true>::pressure_force,
                SphLaunch<float, true>::boundary_init, SphLaunch<float, true>::pressure_dual,
                SphLaunch<float, true>::correction_loop};
    return {SphLaunch<float, false>::density, SphLaunch<float, false>::normal, SphLaunch<float, false>::force_init,
            SphLaunch<float, false>::pressure, SphLaunch<float, false>::pressure_force,
            SphLaunch<float, false>::boundary_init, SphLaunch<float, false>::pressure_dual,
            SphLaunch<float, false>::correction_loop};
}

template <>
Engine<double>::LaunchTable Engine<double>::launch(bool) {
    return {SphLaunch<double, false>::density, SphLaunch<double, false>::normal,
            SphLaunch<double, false>::force_init, SphLaunch<double, false>::pressure,
            SphLaunch<double, false>::pressure_force, SphLaunch<double, false>::boundary_init,
            SphLaunch<double, false>::pressure_dual, SphLaunch<double, false>::correction_loop};
}

}  // namespace pcisph

using namespace pcisph;

// Arithmetic of the step, chosen with pcisph_set_precision before uploading.
enum { kPrecisionFast = 0, kPrecisionIeee = 1, kPrecisionFp64 = 2 };

struct pcisph_solver {
    // One scene of the batch: what a sim::Fluid holds (Fluid.h:13-102) -- its
    // parameters, whether a setter changed them since the last step (Fluid.cpp:109-112),
    // and its loop's density-variation threshold.
    struct Scene {
        Simulation_Params params{};
        bool params_changed = true;
        bool boundary_updated = true;
        float threshold = 0.01f;
        bool warm_skip = false;  // left out of the next step's warm start (warm_start_eligible)
    };
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<Scene> scenes = std::vector<Scene>(1);
    uint32_t current = 0;  // the scene that setters, uploads and downloads address
    pcisph_solver_options opt{3, 7, 0, 0, 0, 0, 0.2f, 4, 3, 0};
    int precision = kPrecisionFast;
    std::unique_ptr<EngineBase> eng;
    std::string err;
    // warm-start network and its training data (device memory, allocated on first use)
    float* mlp = nullptr;        // FP32 parameters
    float* mlp_m = nullptr;      // Adam moments
    float* mlp_v = nullptr;
    float* mlp_grad = nullptr;
    float* mlp_loss = nullptr;
    uint32_t adam_step = 0;
    bool warm_start = false;
    float* samples_x = nullptr;  // [cap][8]
    float* samples_t = nullptr;  // [cap]
    uint64_t samples_cap = 0, samples_count = 0;
    uint32_t collect_stride = 0, collect_phase = 0;

    ~pcisph_solver() {
        for (float* q : {mlp, mlp_m, mlp_v, mlp_grad, mlp_loss, samples_x, samples_t})
            if (q) (void)hipFree(q);
    }

    Scene& cur() { return scenes[current]; }
    const Scene& cur() const { return scenes[current]; }

    void push_counts() {
        std::vector<uint32_t> n(scenes.size()), nb(scenes.size());
        for (size_t i = 0; i < scenes.size(); ++i) {
            n[i] = scenes[i].params.fluid_count;
            nb[i] = scenes[i].params.boundary_count;
        }
        eng->set_counts(n, nb);
    }

    uint64_t total_fluid() const {
        uint64_t t = 0;
        for (const Scene& sc : scenes) t += sc.params.fluid_count;
        return t;
    }

    void ensure_mlp() {
        if (mlp) return;
        HIP_CHECK(hipSetDevice(device));
        for (float** q : {&mlp, &mlp_m, &mlp_v, &mlp_grad}) {
            HIP_CHECK(hipMalloc((void**)q, sizeof(float) * kMlpParams));
            HIP_CHECK(hipMemsetAsync(*q, 0, sizeof(float) * kMlpParams, stream));
        }
        HIP_CHECK(hipMalloc((void**)&mlp_loss, sizeof(float) * 4));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
};

static thread_local std::string g_create_error;

// ------------------------------------------------------------------ gradients without allocation
// The scratch of pcisph_mlp_gradients_noalloc: the caller's rows staged on the device
// (x, then t) and the gradient and loss accumulators, for at most kNoAllocRows rows.
// Each GPU holds its own copy of the array until the process ends.  Its device
// address is taken once per device, at pcisph_create with the device current
// (hipGetSymbolAddress), and host code uses only that address, never the array's name.
constexpr size_t kNoAllocRows = 65536;
constexpr size_t kNoAllocGrad = (kNoAllocRows * (kMlpIn + 1) + 63) / 64 * 64;  // accumulators' offset
constexpr size_t kNoAllocWords = kNoAllocGrad + (kMlpParams + 1 + 63) / 64 * 64;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocWords];
static float* g_scratch[kNoAllocDevices];
static std::mutex g_scratch_mu;

// Once per device, with that device current: the device address of `scratch`.
static void init_scratch(int device) {
    if (device < 0 || device >= kNoAllocDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = (float*)p;
}

template <typename Fn>
static int guarded(pcisph_solver* s, Fn&& fn) {
    if (!s) return PCISPH_ERR_INVALID_ARGUMENT;
    try {
        s->err.clear();
        fn();
        return PCISPH_OK;
    } catch (const Error& e) {
        s->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        s->err = "host allocation failed";
        return PCISPH_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        s->err = e.what();
        return PCISPH_ERR_INTERNAL;
    }
}

static void make_engine(pcisph_solver* s) {
    if (s->precision == kPrecisionFp64) s->eng.reset(new Engine<double>(s->device, s->stream));
    else s->eng.reset(new Engine<float>(s->device, s->stream));
    s->push_counts();
    for (auto& sc : s->scenes) sc.params_changed = true;
}

extern "C" {

int pcisph_create(int device_ordinal, void* hip_stream, pcisph_solver** out) {
    if (!out) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_create_error = "no HIP device";
        return PCISPH_ERR_NO_DEVICE;
    }
    if (device_ordinal < 0 || device_ordinal >= count) {
        g_create_error = "device ordinal out of range";
        return PCISPH_ERR_INVALID_ARGUMENT;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("device is not gfx950: ") + prop.gcnArchName;
        return PCISPH_ERR_NO_DEVICE;
    }
    auto* s = new pcisph_solver();
    s->device = device_ordinal;
    int rc = guarded(s, [&] {
        HIP_CHECK(hipSetDevice(device_ordinal));
        init_scratch(device_ordinal);
        if (hip_stream) {
            s->stream = (hipStream_t)hip_stream;
        } else {
            HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            s->own_stream = true;
        }
        make_engine(s);
    });
    if (rc) {
        g_create_error = s->err;
        pcisph_destroy(s);
        return rc;
    }
    *out = s;
    return PCISPH_OK;
}

int pcisph_destroy(pcisph_solver* s) {
    if (!s) return PCISPH_ERR_INVALID_ARGUMENT;
    s->eng.reset();
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return PCISPH_OK;
}

const char* pcisph_last_error(const pcisph_solver* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

#define PCISPH_SETTER(name, type, member, resize)                                      \
    int pcisph_set_##name(pcisph_solver* s, type v) {                                  \
        return guarded(s, [&] {                                                        \
            s->cur().params.member = v;                                                \
            s->cur().params_changed = true;                                            \
            if (resize) s->push_counts();                                              \
        });                                                                            \
    }
PCISPH_SETTER(boundary_count, uint32_t, boundary_count, true)
PCISPH_SETTER(fluid_count, uint32_t, fluid_count, true)
PCISPH_SETTER(delta_t, float, delta_t, false)
PCISPH_SETTER(rest_density, float, rest_density, false)
PCISPH_SETTER(particle_radius, float, particle_radius, false)
PCISPH_SETTER(gravity, float, gravity, false)
PCISPH_SETTER(viscosity, float, viscosity_constant, false)
PCISPH_SETTER(surface_tension, float, surface_tension_coefficient, false)
#undef PCISPH_SETTER

int pcisph_set_density_variation_threshold(pcisph_solver* s, float threshold) {
    return guarded(s, [&] { s->cur().threshold = threshold; });
}

int pcisph_get_params(const pcisph_solver* s, Simulation_Params* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = s->cur().params;
    return PCISPH_OK;
}

int pcisph_set_solver_options(pcisph_solver* s, const pcisph_solver_options* opt) {
    return guarded(s, [&] {
        if (!opt) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null options");
        if (opt->max_iterations == 0 || opt->max_iterations > (uint32_t)kMaxIterationsCap)
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "max_iterations must be in 1..256");
        if (opt->block_size && (opt->block_size % 64 || opt->block_size > 512))
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64, at most 512");
        if (opt->sort_digit_bits && (opt->sort_digit_bits < 4 || opt->sort_digit_bits > 8))
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "sort_digit_bits must be in 4..8");
        if (!(opt->neighbor_skin >= 0.f && opt->neighbor_skin <= 2.f))
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "neighbor_skin must be in [0, 2]");
        if (opt->work_order > 255) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "work_order must be in 0..255");
        if (opt->stage_cap && (opt->stage_cap < 64 || opt->stage_cap > kStageCap))
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "stage_cap must be 0 or in 64..2048");
        s->opt = *opt;
    });
}

int pcisph_get_solver_options(const pcisph_solver* s, pcisph_solver_options* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = s->opt;
    return PCISPH_OK;
}

int pcisph_set_precision(pcisph_solver* s, int precision) {
    return guarded(s, [&] {
        if (precision < kPrecisionFast || precision > kPrecisionFp64)
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "precision must be 0 (fast FP32), 1 (IEEE FP32) or 2 (FP64)");
        const bool was64 = s->precision == kPrecisionFp64, is64 = precision == kPrecisionFp64;
        s->precision = precision;
        if (was64 != is64) make_engine(s);
    });
}

int pcisph_check_buffers_consistent(const pcisph_solver* s) {
    if (!s) return PCISPH_ERR_INVALID_ARGUMENT;
    return PCISPH_OK;  // the solver owns its buffers and sizes them with the counts
}

int pcisph_field_size(const pcisph_solver* s, int field, size_t* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    const size_t n = s->eng->field_size(s->current, field, s->cur().params);
    if (n == (size_t)-1) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = n;
    return PCISPH_OK;
}

static void check_count(pcisph_solver* s, int field, size_t count) {
    const Simulation_Params& p = s->cur().params;
    const size_t n = s->eng->field_size(s->current, field, p);
    if (n == (size_t)-1) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "unknown field " + std::to_string(field));
    if (count != n)
        throw Error(PCISPH_ERR_SIZE, "Inconsistent size (" + std::to_string(count) + ") of field " +
                                         std::to_string(field) + " for " + std::to_string(p.fluid_count) +
                                         " fluid / " + std::to_string(p.boundary_count) +
                                         " boundary particles (expected " + std::to_string(n) + ")");
}

static int upload_impl(pcisph_solver* s, int field, const void* p, size_t count, bool device) {
    return guarded(s, [&] {
        if (!p && count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer");
        check_count(s, field, count);
        s->eng->upload(s->current, field, p, count, device);
        if (field == PCISPH_BOUNDARY_POSITIONS) s->cur().boundary_updated = true;
    });
}

static int download_impl(pcisph_solver* s, int field, void* p, size_t count, bool device, bool f64) {
    return guarded(s, [&] {
        if (!p && count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer");
        check_count(s, field, count);
        s->eng->download(s->current, field, p, count, device, f64);
    });
}

int pcisph_upload(pcisph_solver* s, int field, const void* host, size_t count) {
    return upload_impl(s, field, host, count, false);
}
int pcisph_download(pcisph_solver* s, int field, void* host, size_t count) {
    return download_impl(s, field, host, count, false, false);
}
int pcisph_upload_device(pcisph_solver* s, int field, const void* device_ptr, size_t count) {
    return upload_impl(s, field, device_ptr, count, true);
}
int pcisph_download_device(pcisph_solver* s, int field, void* device_ptr, size_t count) {
    return download_impl(s, field, device_ptr, count, true, false);
}
int pcisph_download_f64(pcisph_solver* s, int field, double* host, size_t count) {
    return download_impl(s, field, host, count, false, true);
}

// update_deduced_attributes for every scene whose parameters changed
// (Fluid.cpp:109-112, 377-466): derived parameters, bucket tables, and the scene's
// velocities zeroed (Fluid.cpp:462-463).
static void apply_param_changes(pcisph_solver* s) {
    bool any = false;
    for (auto& sc : s->scenes) any = any || sc.params_changed;
    if (!any) return;
    std::vector<pcisph::SceneInput> in;
    for (auto& sc : s->scenes) {
        if (sc.params_changed) derive_params(sc.params);
        in.push_back({&sc.params, sc.threshold});
    }
    s->eng->realloc_internal(in);
    for (uint32_t i = 0; i < s->scenes.size(); ++i) {
        auto& sc = s->scenes[i];
        if (!sc.params_changed) continue;
        s->eng->zero_velocities(i);
        sc.boundary_updated = true;
        sc.params_changed = false;
    }
}

// Whether the next step may start a scene from the network, from its last step.  Not
// after a step that ran to the iteration cap without converging: a start cannot
// shorten a loop the cap ends.  Nor after a step that converged at the minimum
// iteration count unless the network's start was kept there and the plain start's
// first iteration was above the bound: no start can go below the minimum, and the
// warm start then costs its inference, network pressure force and dual first
// iteration for nothing.  A scene left out runs plain steps until one needs more
// than the minimum.
static bool warm_start_eligible(const StepResult& r, const pcisph_solver_options& o, double rest_density,
                                double threshold) {
    const uint32_t floor_it = std::max(o.min_iterations, 1u);
    if (r.iterations >= o.max_iterations && !r.converged) return false;
    if (!(r.converged && r.iterations <= floor_it)) return true;
    const bool kept = r.warm && !r.warm_rejected;
    return kept && !(r.first_dv / rest_density < threshold);
}

static pcisph_step_stats to_stats(const pcisph_solver* s, const StepResult& r, const Simulation_Params& p) {
    pcisph_step_stats st{};
    if (!p.fluid_count) return st;
    st.warm_start = (r.warm ? 1u : 0u) | (r.warm_rejected ? 2u : 0u);
    st.max_density_expansion = (float)r.max_expansion;
    st.list_waves = r.list_waves;
    st.iterations = r.iterations;
    st.converged = r.converged;
    st.max_density_variation = (float)r.max_dv;
    st.max_density_error = (float)(r.max_dv / (double)p.rest_density);
    if (s->precision != kPrecisionFp64) st.max_density_error = (float)r.max_dv / p.rest_density;
    st.fluid_candidates = r.fluid_candidates;
    st.boundary_candidates = r.boundary_candidates;
    st.first_density_error = (float)(r.first_dv / (double)p.rest_density);
    return st;
}

// The batch's statistics as one step's: the largest iteration count, error and
// expansion, converged if every scene converged, candidates summed.
static pcisph_step_stats combine(const std::vector<pcisph_step_stats>& v) {
    pcisph_step_stats a{};
    if (v.empty()) return a;
    a.converged = 1;
    bool any = false;
    for (const auto& st : v) {
        if (!st.iterations) continue;
        any = true;
        a.iterations = std::max(a.iterations, st.iterations);
        a.converged = a.converged && st.converged;
        a.max_density_variation = std::max(a.max_density_variation, st.max_density_variation);
        a.max_density_error = std::max(a.max_density_error, st.max_density_error);
        a.max_density_expansion = std::max(a.max_density_expansion, st.max_density_expansion);
        a.nonfinite |= st.nonfinite;
        a.warm_start |= st.warm_start;
        a.fluid_candidates += st.fluid_candidates;
        a.boundary_candidates += st.boundary_candidates;
        a.list_waves = std::max(a.list_waves, st.list_waves);
        a.first_density_error = std::max(a.first_density_error, st.first_density_error);
    }
    if (!any) a.converged = 0;
    return a;
}

// One step of every scene; per-scene statistics into `out` (resized to the scene count).
static void step_impl(pcisph_solver* s, std::vector<pcisph_step_stats>* out) {
    apply_param_changes(s);
    if (out) out->assign(s->scenes.size(), pcisph_step_stats{});
    if (!s->total_fluid()) return;
    // A scene left out of the warm start (warm_start_eligible) runs exactly the plain
    // step, and none of the inference, network pressure force and dual first iteration;
    // the step leaves them out altogether when every scene is left out.
    WarmArgs w;
    bool any_warm = false;
    for (const auto& sc : s->scenes) any_warm = any_warm || (sc.params.fluid_count && !sc.warm_skip);
    if (s->warm_start && s->mlp && any_warm) w.params = s->mlp;
    if (s->collect_stride && s->samples_x && s->samples_count < s->samples_cap) {
        w.collect_stride = s->collect_stride;
        w.collect_phase = s->collect_phase++;
        w.store_x = s->samples_x;
        w.store_t = s->samples_t;
        w.store_base = s->samples_count;
        w.store_cap = s->samples_cap;
    }
    std::vector<pcisph::SceneInput> in;
    bool sort_boundary = s->opt.resort_boundary != 0;
    for (auto& sc : s->scenes) {
        in.push_back({&sc.params, sc.threshold, w.params != nullptr && !sc.warm_skip});
        sort_boundary = sort_boundary || sc.boundary_updated;
    }
    // the fast kernels' clamped poly6 term needs h^2 <= 1 (kernel radius at most 1 m);
    // a batch with a larger one runs the IEEE FP32 kernels
    bool fast = s->precision == kPrecisionFast;
    for (const auto& sc : s->scenes) fast = fast && sc.params.kernel_radius2 <= 1.f;
    std::vector<StepResult> res;
    uint64_t samples = 0;
    s->eng->step(in, s->opt, sort_boundary, fast, w, res, &samples);
    s->samples_count = std::min(s->samples_cap, s->samples_count + samples);
    for (auto& sc : s->scenes) sc.boundary_updated = false;
    for (size_t i = 0; i < res.size() && i < s->scenes.size(); ++i)
        s->scenes[i].warm_skip =
            !warm_start_eligible(res[i], s->opt, s->scenes[i].params.rest_density, s->scenes[i].threshold);
    if (out)
        for (size_t i = 0; i < res.size(); ++i) {
            (*out)[i] = to_stats(s, res[i], s->scenes[i].params);
            if (s->scenes[i].params.fluid_count && s->scenes[i].warm_skip) (*out)[i].warm_start |= 4u;
        }
}

static void mark_nonfinite(pcisph_solver* s, std::vector<pcisph_step_stats>& v) {
    if (!s->total_fluid()) return;
    const std::vector<uint32_t> bad = s->eng->nonfinite();
    for (size_t i = 0; i < v.size() && i < bad.size(); ++i) v[i].nonfinite = bad[i];
}

int pcisph_step(pcisph_solver* s, pcisph_step_stats* stats) {
    return guarded(s, [&] {
        std::vector<pcisph_step_stats> v;
        step_impl(s, stats ? &v : nullptr);
        if (!stats) return;
        mark_nonfinite(s, v);
        *stats = combine(v);
    });
}

int pcisph_run(pcisph_solver* s, uint32_t n_steps, pcisph_step_stats* stats) {
    return guarded(s, [&] {
        std::vector<pcisph_step_stats> v;
        for (uint32_t i = 0; i < n_steps; i++) {
            step_impl(s, stats ? &v : nullptr);
            if (!stats) continue;
            if (i + 1 == n_steps) mark_nonfinite(s, v);
            stats[i] = combine(v);
        }
    });
}

int pcisph_step_scenes(pcisph_solver* s, pcisph_step_stats* stats, size_t count) {
    return guarded(s, [&] {
        if (stats && count != s->scenes.size())
            throw Error(PCISPH_ERR_SIZE, "expected statistics for " + std::to_string(s->scenes.size()) + " scenes");
        std::vector<pcisph_step_stats> v;
        step_impl(s, stats ? &v : nullptr);
        if (!stats) return;
        mark_nonfinite(s, v);
        std::copy(v.begin(), v.end(), stats);
    });
}

int pcisph_set_scene_count(pcisph_solver* s, uint32_t count) {
    return guarded(s, [&] {
        if (count == 0 || count > (1u << 20)) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "scene count must be in 1..2^20");
        if (count == s->scenes.size()) return;
        s->scenes.resize(count);
        if (s->current >= count) s->current = 0;
        for (auto& sc : s->scenes) sc.params_changed = true;  // bucket layout of the batch
        s->push_counts();
    });
}

int pcisph_get_scene_count(const pcisph_solver* s, uint32_t* count) {
    if (!s || !count) return PCISPH_ERR_INVALID_ARGUMENT;
    *count = (uint32_t)s->scenes.size();
    return PCISPH_OK;
}

int pcisph_select_scene(pcisph_solver* s, uint32_t index) {
    return guarded(s, [&] {
        if (index >= s->scenes.size())
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "scene " + std::to_string(index) + " out of range");
        s->current = index;
    });
}

int pcisph_prepare(pcisph_solver* s) {
    return guarded(s, [&] { apply_param_changes(s); });
}

int pcisph_sync(pcisph_solver* s) {
    return guarded(s, [&] { HIP_CHECK(hipStreamSynchronize(s->stream)); });
}

// ------------------------------------------------------------------ warm start
struct DeviceTemp {
    void* p = nullptr;
    explicit DeviceTemp(size_t bytes) { HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    ~DeviceTemp() { (void)hipFree(p); }
    float* f() const { return (float*)p; }
};

int pcisph_mlp_set_params(pcisph_solver* s, const float* params, size_t count) {
    return guarded(s, [&] {
        if (!params || count != (size_t)kMlpParams)
            throw Error(PCISPH_ERR_SIZE, "expected " + std::to_string(kMlpParams) + " parameters");
        s->ensure_mlp();
        HIP_CHECK(hipMemcpyAsync(s->mlp, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s->stream));
        HIP_CHECK(hipMemsetAsync(s->mlp_m, 0, sizeof(float) * kMlpParams, s->stream));
        HIP_CHECK(hipMemsetAsync(s->mlp_v, 0, sizeof(float) * kMlpParams, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        s->adam_step = 0;
    });
}

int pcisph_mlp_get_params(pcisph_solver* s, float* params, size_t count) {
    return guarded(s, [&] {
        if (!params || count != (size_t)kMlpParams)
            throw Error(PCISPH_ERR_SIZE, "expected " + std::to_string(kMlpParams) + " parameters");
        s->ensure_mlp();
        HIP_CHECK(hipMemcpyAsync(params, s->mlp, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int pcisph_mlp_forward(pcisph_solver* s, const float* x, float* y, size_t n) {
    return guarded(s, [&] {
        if ((!x || !y) && n) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer");
        s->ensure_mlp();
        DeviceTemp dx(n * kMlpIn * sizeof(float)), dy(n * sizeof(float));
        HIP_CHECK(hipMemcpyAsync(dx.p, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, s->stream));
        mlp_forward(MlpLaunch{s->stream}, s->mlp, dx.f(), dy.f(), (uint32_t)n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(y, dy.p, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int pcisph_mlp_gradients(pcisph_solver* s, const float* x, const float* t, size_t n, float* grad, float* loss) {
    return guarded(s, [&] {
        if (!x || !t || !grad || !loss || !n) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer or empty batch");
        s->ensure_mlp();
        DeviceTemp dx(n * kMlpIn * sizeof(float)), dt(n * sizeof(float));
        HIP_CHECK(hipMemcpyAsync(dx.p, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIP_CHECK(hipMemcpyAsync(dt.p, t, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
        mlp_gradients(MlpLaunch{s->stream}, s->mlp, dx.f(), dt.f(), (uint32_t)n, s->mlp_grad, s->mlp_loss);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(grad, s->mlp_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipMemcpyAsync(loss, s->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

// pcisph_mlp_gradients with its device scratch in the static array above: no device
// memory is allocated or freed.  Every call on one GPU shares that GPU's array (the
// accumulators are cleared by kernels at each call); calls that overlap on one GPU
// race.
int pcisph_mlp_gradients_noalloc(pcisph_solver* s, const float* x, const float* t, size_t n, float* grad,
                                 float* loss) {
    return guarded(s, [&] {
        if (!x || !t || !grad || !loss || !n) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer or empty batch");
        if (n > kNoAllocRows)
            throw Error(PCISPH_ERR_SIZE, "at most " + std::to_string(kNoAllocRows) + " rows per call");
        if (!s->mlp) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "no network parameters (pcisph_mlp_set_params first)");
        float* base = s->device >= 0 && s->device < kNoAllocDevices ? g_scratch[s->device] : nullptr;
        if (!base) throw Error(PCISPH_ERR_INTERNAL, "no scratch address for this device");
        HIP_CHECK(hipSetDevice(s->device));
        float* dx = base;
        float* dt = base + kNoAllocRows * kMlpIn;
        float* dg = base + kNoAllocGrad;
        float* dl = dg + kMlpParams;
        HIP_CHECK(hipMemcpyAsync(dx, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIP_CHECK(hipMemcpyAsync(dt, t, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
        mlp_gradients(MlpLaunch{s->stream}, s->mlp, dx, dt, (uint32_t)n, dg, dl);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(grad, dg, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipMemcpyAsync(loss, dl, sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int pcisph_samples_reserve(pcisph_solver* s, size_t capacity) {
    return guarded(s, [&] {
        HIP_CHECK(hipSetDevice(s->device));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        for (float** q : {&s->samples_x, &s->samples_t})
            if (*q) {
                (void)hipFree(*q);
                *q = nullptr;
            }
        s->samples_cap = s->samples_count = 0;
        if (!capacity) return;
        HIP_CHECK(hipMalloc((void**)&s->samples_x, capacity * kMlpIn * sizeof(float)));
        HIP_CHECK(hipMalloc((void**)&s->samples_t, capacity * sizeof(float)));
        s->samples_cap = capacity;
    });
}

int pcisph_samples_collect(pcisph_solver* s, uint32_t stride) {
    return guarded(s, [&] {
        if (stride && !s->samples_x) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "reserve a sample store first");
        if (stride && s->precision == kPrecisionFp64)
            throw Error(PCISPH_ERR_INVALID_ARGUMENT, "samples are collected from FP32 solvers");
        s->collect_stride = stride;
    });
}

int pcisph_samples_count(const pcisph_solver* s, size_t* count) {
    if (!s || !count) return PCISPH_ERR_INVALID_ARGUMENT;
    *count = (size_t)s->samples_count;
    return PCISPH_OK;
}

int pcisph_samples_download(pcisph_solver* s, float* x, float* t, size_t count) {
    return guarded(s, [&] {
        if (count > s->samples_count) throw Error(PCISPH_ERR_SIZE, "more samples requested than stored");
        if (!count) return;
        HIP_CHECK(hipStreamSynchronize(s->stream));
        HIP_CHECK(hipMemcpy(x, s->samples_x, count * kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(t, s->samples_t, count * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int pcisph_samples_upload(pcisph_solver* s, const float* x, const float* t, size_t count) {
    return guarded(s, [&] {
        if (count > s->samples_cap) throw Error(PCISPH_ERR_SIZE, "sample store too small");
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (count) {
            HIP_CHECK(hipMemcpy(s->samples_x, x, count * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(s->samples_t, t, count * sizeof(float), hipMemcpyHostToDevice));
        }
        s->samples_count = count;
    });
}

int pcisph_mlp_train(pcisph_solver* s, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* last_loss) {
    return guarded(s, [&] {
        if (!s->samples_count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "no samples to train on");
        if (!batch) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "batch must be positive");
        s->ensure_mlp();
        const MlpLaunch ML{s->stream};
        DeviceTemp bx((size_t)batch * kMlpIn * sizeof(float)), bt((size_t)batch * sizeof(float));
        for (uint32_t i = 0; i < steps; ++i) {
            mlp_gather_batch(ML, s->samples_x, s->samples_t, s->samples_count, seed, s->adam_step, batch, bx.f(),
                             bt.f());
            mlp_gradients(ML, s->mlp, bx.f(), bt.f(), batch, s->mlp_grad, s->mlp_loss);
            mlp_adam(ML, s->mlp, s->mlp_grad, s->mlp_m, s->mlp_v, ++s->adam_step, lr, 0.9f, 0.999f, 1e-8f);
        }
        HIP_CHECK(hipGetLastError());
        float loss = 0.f;
        HIP_CHECK(hipMemcpyAsync(&loss, s->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (last_loss) *last_loss = loss / (float)batch;
    });
}

int pcisph_translate_boundary(pcisph_solver* s, float dx, float dy, float dz) {
    return guarded(s, [&] {
        s->eng->translate_boundary(s->current, dx, dy, dz);
        s->cur().boundary_updated = true;
    });
}

int pcisph_translate_boundaries(pcisph_solver* s, const float* shifts, size_t count) {
    return guarded(s, [&] {
        if (!shifts || count != 3 * s->scenes.size())
            throw Error(PCISPH_ERR_SIZE, "expected 3 shifts per scene (" + std::to_string(3 * s->scenes.size()) + ")");
        s->eng->translate_boundaries(shifts, s->eng->block_for(s->opt.block_size));
        for (uint32_t i = 0; i < s->scenes.size(); ++i)
            if (shifts[3 * i] != 0.f || shifts[3 * i + 1] != 0.f || shifts[3 * i + 2] != 0.f)
                s->scenes[i].boundary_updated = true;
    });
}

int pcisph_set_warm_start(pcisph_solver* s, int enable) {
    return guarded(s, [&] {
        if (enable) s->ensure_mlp();
        s->warm_start = enable != 0;
        for (auto& sc : s->scenes) sc.warm_skip = false;
    });
}

// ------------------------------------------------------------------ autotuner
// Times `trial_steps` steps of the current scene for each block size of the particle
// kernels, then each radix digit width, then each tiling of the pressure pass
// (work_order bits 4 and 5), from the same starting state each time (a device
// snapshot restored after every trial), and keeps the fastest options.  The state after the call is the state before it: particle state,
// pending boundary sorts, each scene's warm-start eligibility and
// sample collection; on an error the options are the caller's again too.
int pcisph_autotune(pcisph_solver* s, uint32_t trial_steps, pcisph_solver_options* chosen) {
    return guarded(s, [&] {
        if (!s->total_fluid()) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "no particles to tune on");
        if (!trial_steps) trial_steps = 3;
        apply_param_changes(s);
        // the boundary sort a pending update would do is part of every trial's first step
        std::vector<bool> boundary_pending, warm_skip;
        for (auto& sc : s->scenes) {
            boundary_pending.push_back(sc.boundary_updated);
            warm_skip.push_back(sc.warm_skip);
        }
        auto reset_boundary = [&] {
            for (size_t i = 0; i < s->scenes.size(); ++i) {
                s->scenes[i].boundary_updated = boundary_pending[i];
                s->scenes[i].warm_skip = warm_skip[i];
            }
        };
        s->eng->snapshot();
        const pcisph_solver_options base = s->opt;
        const uint32_t collect = s->collect_stride, phase = s->collect_phase;
        s->collect_stride = 0;
        // puts the caller's state back however the trials end
        struct Guard {
            std::function<void()> fn;
            ~Guard() {
                try {
                    fn();
                } catch (...) {
                }
            }
        } guard{[&] {
            s->opt = base;
            s->collect_stride = collect;
            s->collect_phase = phase;
            reset_boundary();
            s->eng->restore();
        }};
        double best = 1e300;
        pcisph_solver_options best_opt = base;
        auto trial = [&](uint32_t block, uint32_t digit, uint32_t work_order) {
            s->opt = base;
            s->opt.block_size = block;
            s->opt.sort_digit_bits = digit;
            s->opt.work_order = work_order;
            s->eng->restore();
            reset_boundary();
            step_impl(s, nullptr);  // warm-up: caches, code objects
            s->eng->restore();
            reset_boundary();
            HIP_CHECK(hipStreamSynchronize(s->stream));
            const auto t0 = std::chrono::steady_clock::now();
            for (uint32_t i = 0; i < trial_steps; ++i) step_impl(s, nullptr);
            HIP_CHECK(hipStreamSynchronize(s->stream));
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (dt < best) {
                best = dt;
                best_opt = s->opt;
            }
        };
        const uint32_t wo = base.work_order & ~48u;
        for (uint32_t block : {64u, 128u, 256u, 512u}) trial(block, 8u, base.work_order);
        const uint32_t block = best_opt.block_size;
        for (uint32_t digit : {5u, 6u, 7u}) trial(block, digit, base.work_order);
        // the pressure pass's tiles: LDS-staged workgroup tiles where no lists serve it
        // (the default), always (bit 4), or never (bit 5: every lane walks its buckets)
        const uint32_t digit = best_opt.sort_digit_bits;
        for (uint32_t tiles : {0u, 16u, 32u})
            if ((wo | tiles) != base.work_order) trial(block, digit, wo | tiles);
        guard.fn();
        guard.fn = [] {};
        s->opt = best_opt;
        if (chosen) *chosen = best_opt;
    });
}

int pcisph_snapshot(pcisph_solver* s) {
    return guarded(s, [&] {
        apply_param_changes(s);
        s->eng->snapshot();
    });
}

int pcisph_restore(pcisph_solver* s) {
    return guarded(s, [&] { s->eng->restore(); });
}

int pcisph_profile_enable(pcisph_solver* s, int enable) {
    return guarded(s, [&] {
        if (enable < 0) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "negative profiling interval");
        s->eng->profile_enable((uint32_t)enable);
    });
}

int pcisph_profile_read(pcisph_solver* s, pcisph_stage_times* out) {
    return guarded(s, [&] {
        if (!out) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null output");
        s->eng->profile_read(out);
    });
}

int pcisph_compute_params(float particle_radius, float rest_density, uint32_t boundary_count,
                          uint32_t fluid_count, float viscosity, float surface_tension, float gravity,
                          float delta_t, Simulation_Params* out) {
    if (!out) return PCISPH_ERR_INVALID_ARGUMENT;
    Simulation_Params p{};
    p.particle_radius = particle_radius;
    p.rest_density = rest_density;
    p.boundary_count = boundary_count;
    p.fluid_count = fluid_count;
    p.viscosity_constant = viscosity;
    p.surface_tension_coefficient = surface_tension;
    p.gravity = gravity;
    p.delta_t = delta_t;
    derive_params(p);
    *out = p;
    return PCISPH_OK;
}

}  // extern "C"
