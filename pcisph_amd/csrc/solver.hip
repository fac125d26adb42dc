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
#include <memory>
#include <stdexcept>
#include <string>
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
// the end of a step), so timing never waits on the stream inside a step.
class StageProfiler {
   public:
    ~StageProfiler() {
        for (auto& p : pending_) free_.push_back(p.a), free_.push_back(p.b);
        for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    }

    void enable(bool on, hipStream_t s) {
        if (!on_ && !on) return;
        resolve(s);
        on_ = on;
        for (auto& v : ms_) v = 0;
        for (auto& v : launches_) v = 0;
    }

    template <typename F>
    void timed(int stage, hipStream_t s, int iter, F&& fn) {
        if (!on_) {
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
        if (!on_) return;
        for (auto& p : pending_)
            if (p.step == step_ && p.iter > last_iter) p.skipped = true;
        ++step_;
    }

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

    bool on_ = false;
    uint64_t step_ = 0;
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
    uint64_t samples = 0;
    uint32_t list_passes = 0;
};

// What a step does for the warm-start network: use it (params != null) and/or
// append (features, converged pressure) samples of every stride-th particle.
struct WarmArgs {
    const float* params = nullptr;
    uint32_t collect_stride = 0, collect_phase = 0;
    float* store_x = nullptr;
    float* store_t = nullptr;
    uint64_t store_base = 0, store_cap = 0;
};

class EngineBase {
   public:
    virtual ~EngineBase() = default;
    virtual void resize_public(uint32_t n, uint32_t nb) = 0;
    virtual void realloc_internal(const Simulation_Params& p) = 0;
    virtual void zero_velocities() = 0;
    virtual void upload(int field, const void* src, size_t count, bool device) = 0;
    virtual void download(int field, void* dst, size_t count, bool device, bool f64) = 0;
    virtual StepResult step(const Simulation_Params& p, float threshold, const pcisph_solver_options& o,
                            bool sort_boundary, bool fast, const WarmArgs& w) = 0;
    virtual uint32_t nonfinite() = 0;
    virtual size_t field_size(int field, const Simulation_Params& p) const = 0;
    virtual void profile_enable(bool on) = 0;
    virtual void profile_read(pcisph_stage_times* out) = 0;
    // device-side copy of the state a step starts from (positions, velocities with the
    // previous pressures, boundary positions), for trial runs of the autotuner
    virtual void translate_boundary(float dx, float dy, float dz) = 0;
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

template <typename R>
class Engine final : public EngineBase {
   public:
    Engine(int device, hipStream_t stream) : device_(device), stream_(stream) {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipHostMalloc((void**)&host_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
        std::memset((void*)host_, 0, sizeof(HostMirror));
        HIP_CHECK(hipHostGetDevicePointer((void**)&dev_host_, (void*)host_, 0));
        dalloc(A_.state, 1, stream_);
    }

    ~Engine() override {
        (void)hipStreamSynchronize(stream_);
        free_public();
        free_internal();
        dfree(A_.state);
        dfree(staging_);
        dfree(feat_);
        drop_snapshot();
        dfree(A_.NL.f);
        dfree(A_.NL.b);
        dfree(A_.NL.bf);
        dfree(A_.NL.fcount);
        dfree(A_.NL.bcount);
        dfree(A_.NL.bfcount);
        dfree(D_.xp0);
        dfree(D_.xf0);
        dfree(D_.p0);
        dfree(D_.dv0);
        dfree(D_.bp0);
        dfree(D_.bxf0);
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
        dfree(A_.cells);
        dfree(A_.bcells);
        dfree(sort_scratch_);
    }

    void resize_public(uint32_t n, uint32_t nb) override {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (n != A_.n) {
            for (auto** q : {&A_.pos_in, &A_.vel_in, &A_.pos, &A_.vel, &A_.xr, &A_.nrm, &A_.fo, &A_.fp, &A_.xp,
                             &A_.xf})
                dalloc(*q, n, stream_);
            dalloc(A_.p, n, stream_);
            dalloc(A_.dv, n, stream_);
            dalloc(A_.counts, n, stream_);
            for (int i = 0; i < 2; i++) {
                dalloc(A_.keys[i], n, stream_);
                dalloc(A_.vals[i], n, stream_);
            }
            A_.n = n;
            fluid_sorted_ = 0;
        }
        if (nb != A_.nb) {
            for (auto** q : {&A_.bpos, &A_.bpos_tmp, &A_.bxf}) dalloc(*q, nb, stream_);
            dalloc(A_.bp, nb, stream_);
            dalloc(A_.binit, nb, stream_);
            for (int i = 0; i < 2; i++) {
                dalloc(A_.bkeys[i], nb, stream_);
                dalloc(A_.bvals[i], nb, stream_);
            }
            A_.nb = nb;
            boundary_sorted_ = 0;
        }
        const size_t stage = std::max<size_t>(4 * (size_t)std::max(n, nb), 1);
        if (stage > staging_words_) {
            dalloc(staging_, stage * 2, stream_);  // room for double results too
            staging_words_ = stage;
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void realloc_internal(const Simulation_Params& p) override {
        HIP_CHECK(hipSetDevice(device_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (p.bucket_count != A_.B || !A_.cells) {
            dalloc(A_.cells, p.bucket_count, stream_);
            dalloc(A_.bcells, p.bucket_count, stream_);
            A_.B = p.bucket_count;
        }
        const size_t words = std::max(radix_sort_scratch_words(A_.n, 8), radix_sort_scratch_words(A_.nb, 8));
        if (words > sort_scratch_words_) {
            dalloc(sort_scratch_, words, stream_);
            sort_scratch_words_ = words;
        }
        // Uninitialised in the reference (allocated, never written, Fluid.cpp:451):
        // defined here as zero.
        if (A_.nb) HIP_CHECK(hipMemsetAsync(A_.binit, 0, sizeof(R) * A_.nb, stream_));
    }

    void zero_velocities() override {
        if (A_.n) HIP_CHECK(hipMemsetAsync(A_.vel_in, 0, sizeof(V4<R>) * A_.n, stream_));
    }

    void translate_boundary(float dx, float dy, float dz) override {
        DataLaunch<R>::translate(Launch{stream_, 256}, A_.bpos, A_.nb, dx, dy, dz);
        HIP_CHECK(hipGetLastError());
    }

    void snapshot() override {
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
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void drop_snapshot() {
        dfree(snap_pos_);
        dfree(snap_vel_);
        dfree(snap_b_);
    }

    size_t field_size(int field, const Simulation_Params& p) const override {
        return field_size_for(field, p.fluid_count, p.boundary_count, p.bucket_count);
    }

    struct FieldView {
        V4<R>* vec = nullptr;    // vector field
        R* scalar = nullptr;     // scalar field base
        uint32_t stride = 1, offset = 0, n = 0;
        const void* raw = nullptr;  // uint32 field
        size_t raw_words = 0;
        bool boundary = false;
    };

    FieldView view(int field) {
        FieldView v;
        switch (field) {
            case PCISPH_FLUID_POSITIONS: v.vec = A_.pos_in; v.n = A_.n; break;
            case PCISPH_FLUID_VELOCITIES: v.vec = A_.vel_in; v.n = A_.n; break;
            case PCISPH_FLUID_NORMALS: v.vec = A_.nrm; v.n = A_.n; break;
            case PCISPH_FLUID_PREDICTED_POSITIONS: v.vec = A_.xp; v.n = A_.n; break;
            case PCISPH_FLUID_OTHER_FORCES: v.vec = A_.fo; v.n = A_.n; break;
            case PCISPH_FLUID_PRESSURE_FORCES: v.vec = A_.fp; v.n = A_.n; break;
            case PCISPH_FLUID_DENSITIES: v.scalar = (R*)A_.xr; v.stride = 4; v.offset = 3; v.n = A_.n; break;
            case PCISPH_FLUID_PRESSURES: v.scalar = A_.p; v.n = A_.n; break;
            case PCISPH_FLUID_DENSITY_VARIATIONS: v.scalar = A_.dv; v.n = A_.n; break;
            case PCISPH_BOUNDARY_POSITIONS: v.vec = A_.bpos; v.n = A_.nb; v.boundary = true; break;
            case PCISPH_BOUNDARY_PRESSURES: v.scalar = A_.bp; v.n = A_.nb; v.boundary = true; break;
            case PCISPH_BOUNDARY_INIT_PRED_DENSITIES: v.scalar = A_.binit; v.n = A_.nb; v.boundary = true; break;
            case PCISPH_FLUID_KEYS: v.raw = A_.keys[fluid_sorted_]; v.raw_words = A_.n; break;
            case PCISPH_FLUID_SRC_LOCATIONS: v.raw = A_.vals[fluid_sorted_]; v.raw_words = A_.n; break;
            case PCISPH_FLUID_CELL_OFFSETS: v.raw = A_.cells; v.raw_words = 2 * (size_t)A_.B; break;
            case PCISPH_BOUNDARY_KEYS: v.raw = A_.bkeys[boundary_sorted_]; v.raw_words = A_.nb; break;
            case PCISPH_BOUNDARY_SRC_LOCATIONS: v.raw = A_.bvals[boundary_sorted_]; v.raw_words = A_.nb; break;
            case PCISPH_BOUNDARY_CELL_OFFSETS: v.raw = A_.bcells; v.raw_words = A_.nb ? 2 * (size_t)A_.B : 0; break;
            case PCISPH_FLUID_NEIGHBOR_COUNTS: v.raw = A_.counts; v.raw_words = 4 * (size_t)A_.n; break;
            default: throw Error(PCISPH_ERR_INVALID_ARGUMENT, "unknown field " + std::to_string(field));
        }
        return v;
    }

    void upload(int field, const void* src, size_t count, bool device) override {
        HIP_CHECK(hipSetDevice(device_));
        FieldView v = view(field);
        if (v.raw) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "field is read-only (a product of the sort)");
        if (!count) return;
        const Launch L{stream_, 256};
        const float* from = (const float*)src;
        if (!device) {
            HIP_CHECK(hipMemcpyAsync(staging_, src, count * sizeof(float), hipMemcpyHostToDevice, stream_));
            from = (const float*)staging_;
        }
        if (v.vec) DataLaunch<R>::unpack_vec3(L, from, v.vec, v.n);
        else DataLaunch<R>::unpack_scalar(L, from, v.scalar, v.n, v.stride, v.offset);
        HIP_CHECK(hipGetLastError());
        if (!device) HIP_CHECK(hipStreamSynchronize(stream_));  // staging reuse + host buffer lifetime
    }

    void download(int field, void* dst, size_t count, bool device, bool f64) override {
        HIP_CHECK(hipSetDevice(device_));
        FieldView v = view(field);
        if (!count) return;
        const Launch L{stream_, 256};
        if (v.raw) {
            if (f64) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "integer field has no FP64 view");
            HIP_CHECK(hipMemcpyAsync(dst, v.raw, count * 4, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     stream_));
            HIP_CHECK(hipStreamSynchronize(stream_));
            return;
        }
        void* to = device ? dst : (void*)staging_;
        if (f64) {
            if (v.vec) DataLaunch<R>::pack_vec3_f64(L, v.vec, (double*)to, v.n);
            else DataLaunch<R>::pack_scalar_f64(L, v.scalar, (double*)to, v.n, v.stride, v.offset);
        } else {
            if (v.vec) DataLaunch<R>::pack_vec3(L, v.vec, (float*)to, v.n);
            else DataLaunch<R>::pack_scalar(L, v.scalar, (float*)to, v.n, v.stride, v.offset);
        }
        HIP_CHECK(hipGetLastError());
        if (!device)
            HIP_CHECK(hipMemcpyAsync(dst, staging_, count * (f64 ? 8 : 4), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void profile_enable(bool on) override { prof_.enable(on, stream_); }
    void profile_read(pcisph_stage_times* out) override { prof_.read(stream_, out); }

    uint32_t nonfinite() override {
        unsigned int v = 0;
        HIP_CHECK(hipMemcpyAsync(&v, &A_.state->nonfinite, sizeof(v), hipMemcpyDeviceToHost, stream_));
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

    StepResult step(const Simulation_Params& p, float threshold, const pcisph_solver_options& o,
                    bool sort_boundary, bool fast, const WarmArgs& w) override {
        HIP_CHECK(hipSetDevice(device_));
        const Launch L{stream_, o.block_size ? o.block_size : 256u};
        const MlpLaunch ML{stream_};
        const uint32_t digit = o.sort_digit_bits ? o.sort_digit_bits : 8u;
        const SphParams<R> P = make_sph_params<R>(p, threshold, o);
        const uint32_t key_bits = key_bits_for(p.bucket_count);
        const unsigned int tag = ++tag_ ? tag_ : ++tag_;
        const bool warm = w.params != nullptr;
        const bool features = warm || w.collect_stride;
        if (features) ensure_warm_buffers(warm);
        setup_lists(o.neighbor_skin, P);
        fill_u64(L, (unsigned long long*)A_.state, sizeof(LoopState) / 8, 0ull);
        if (sort_boundary && A_.nb) {
            prof_.timed(PCISPH_STAGE_BOUNDARY_SORT, stream_, -1, [&] {
                fill_u32(L, (uint32_t*)A_.bcells, 2 * (size_t)A_.B, 0u);
                DataLaunch<R>::hash(L, A_.bpos, A_.nb, P.cell, P.grid, A_.bkeys[0], A_.bvals[0]);
                boundary_sorted_ = radix_sort(L, A_.bkeys, A_.bvals, A_.nb, key_bits, digit, sort_scratch_);
                if (boundary_sorted_ < 0) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "too many boundary particles");
                DataLaunch<R>::reorder(L, A_.nb, A_.bkeys[boundary_sorted_], A_.bvals[boundary_sorted_],
                                       A_.bpos, nullptr, A_.bpos_tmp, nullptr, A_.bcells);
                std::swap(A_.bpos, A_.bpos_tmp);
                if (o.launch_boundary_init) launch(fast).boundary_init(L, P, A_);
            });
        }
        prof_.timed(PCISPH_STAGE_HASH, stream_, -1, [&] {
            fill_u32(L, (uint32_t*)A_.cells, 2 * (size_t)A_.B, 0u);
            DataLaunch<R>::hash(L, A_.pos_in, A_.n, P.cell, P.grid, A_.keys[0], A_.vals[0]);
        });
        prof_.timed(PCISPH_STAGE_SORT, stream_, -1, [&] {
            fluid_sorted_ = radix_sort(L, A_.keys, A_.vals, A_.n, key_bits, digit, sort_scratch_);
        });
        if (fluid_sorted_ < 0) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "too many particles for one sort");
        prof_.timed(PCISPH_STAGE_REORDER, stream_, -1, [&] {
            DataLaunch<R>::reorder(L, A_.n, A_.keys[fluid_sorted_], A_.vals[fluid_sorted_], A_.pos_in,
                                   A_.vel_in, A_.pos, A_.vel, A_.cells);
        });
        const LaunchTable ops = launch(fast);
        prof_.timed(PCISPH_STAGE_DENSITY, stream_, -1, [&] { ops.density(L, P, A_); });
        prof_.timed(PCISPH_STAGE_NORMAL, stream_, -1, [&] { ops.normal(L, P, A_); });
        const float p_ref = (float)(P.delta * P.rho0 / R(100));
        FeatureArgs F{features ? feat_ : nullptr, A_.counts, p_ref,
                      (float)((double)p.particle_mass * 45.0 / (3.14159265358979323846 * std::pow((double)P.h, 6)) *
                              (double)P.dt * 100.0)};
        prof_.timed(PCISPH_STAGE_FORCE_INIT, stream_, -1, [&] { ops.force_init(L, P, A_, F); });
        if (warm) {
            // network start: p from the features, its pressure force, and a prediction
            // with it; the plain prediction is kept for the dual first iteration
            prof_.timed(PCISPH_STAGE_WARM_START, stream_, -1, [&] {
                mlp_warm_start<R, V4<R>>(ML, w.params, feat_, A_.n, p_ref, A_.xr, A_.p, A_.xf);
                ops.pressure_force(L, P, A_, 0, dev_host_, tag, 1, D_.xp0);
            });
        }
        HIP_CHECK(hipGetLastError());

        const uint32_t maxit = o.max_iterations, minit = std::max(o.min_iterations, 1u);
        const uint32_t first = std::min(minit - 1, maxit - 1);
        const uint32_t lookahead = 1;
        uint32_t enq = 0;
        auto enqueue = [&]() {
            prof_.timed(PCISPH_STAGE_PRESSURE, stream_, (int)enq, [&] {
                if (warm && enq == 0) ops.pressure_dual(L, P, A_, D_);
                else ops.pressure(L, P, A_, enq);
            });
            prof_.timed(PCISPH_STAGE_PRESSURE_FORCE, stream_, (int)enq, [&] {
                ops.pressure_force(L, P, A_, enq, dev_host_, tag, 0, nullptr);
            });
            ++enq;
        };
        while (enq < maxit && enq <= first + lookahead) enqueue();
        HIP_CHECK(hipGetLastError());
        uint32_t k = first;
        for (;;) {
            wait_for_iteration(k, tag);
            if (iteration_is_last<R>((const unsigned long long*)host_->max_dv, k, minit, maxit, P.rho0,
                                     P.threshold))
                break;
            ++k;
            if (enq < maxit) enqueue();
        }
        HIP_CHECK(hipGetLastError());
        StepResult r;
        if (w.collect_stride) {
            const uint32_t phase = w.collect_phase % w.collect_stride;
            mlp_collect<R>(ML, feat_, A_.p, A_.n, w.collect_stride, phase, p_ref, w.store_x, w.store_t,
                           w.store_base, w.store_cap);
            r.samples = A_.n > phase ? (A_.n - phase + w.collect_stride - 1) / w.collect_stride : 0;
        }
        prof_.end_step((int)k);
        prof_.maybe_resolve(stream_);
        r.iterations = k + 1;
        const R mdv = Bits<R>::to(host_->max_dv[k]);
        r.max_dv = (double)mdv;
        r.max_expansion = (double)Bits<R>::to(host_->max_exp[k]);
        r.converged = (k + 1 >= minit) && (mdv / P.rho0 < P.threshold);
        r.fluid_candidates = host_->candidates[0];
        r.boundary_candidates = host_->candidates[1];
        r.warm = warm;
        r.warm_rejected = warm && host_->warm_rejected;
        r.list_passes = host_->list_passes;
        return r;
    }

    // Neighbour-list storage (capacities: fluid 128, boundary 96 per fluid particle,
    // fluid 96 per boundary particle) and this step's radius and validity limit.
    void setup_lists(float skin, const SphParams<R>& P) {
        NeighborLists& NL = A_.NL;
        if (!(skin > 0.f)) {
            NL.r2 = 0.f;
            return;
        }
        constexpr uint32_t kCapF = 128, kCapB = 96, kCapBF = 96;
        if (list_n_ != A_.n || list_nb_ != A_.nb || !NL.f) {
            dalloc(NL.f, (size_t)kCapF * std::max(A_.n, 1u), stream_);
            dalloc(NL.b, (size_t)kCapB * std::max(A_.n, 1u), stream_);
            dalloc(NL.bf, (size_t)kCapBF * std::max(A_.nb, 1u), stream_);
            dalloc(NL.fcount, std::max(A_.n, 1u), stream_);
            dalloc(NL.bcount, std::max(A_.n, 1u), stream_);
            dalloc(NL.bfcount, std::max(A_.nb, 1u), stream_);
            NL.cap_f = kCapF;
            NL.cap_b = kCapB;
            NL.cap_bf = kCapBF;
            list_n_ = A_.n;
            list_nb_ = A_.nb;
        }
        const float radius = (1.f + skin) * (float)P.h;
        NL.r2 = radius * radius;
        NL.disp_limit = 0.45f * skin * (float)P.h;
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

   private:
    struct LaunchTable {
        void (*density)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&);
        void (*normal)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&);
        void (*force_init)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&, const FeatureArgs&);
        void (*pressure)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&, uint32_t);
        void (*pressure_force)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&, uint32_t, HostMirror*,
                               unsigned int, int, V4<R>*);
        void (*boundary_init)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&);
        void (*pressure_dual)(const Launch&, const SphParams<R>&, const DeviceArrays<R>&, const DualArrays<R>&);
    };
    static LaunchTable launch(bool fast);

    int device_;
    hipStream_t stream_;
    DeviceArrays<R> A_;
    DualArrays<R> D_{};
    int fluid_sorted_ = 0, boundary_sorted_ = 0;
    uint32_t* sort_scratch_ = nullptr;
    size_t sort_scratch_words_ = 0;
    float* staging_ = nullptr;
    size_t staging_words_ = 0;
    float* feat_ = nullptr;
    uint32_t feat_n_ = 0, dual_n_ = 0, dual_nb_ = 0, list_n_ = 0, list_nb_ = 0;
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
                SphLaunch<float, true>::boundary_init, SphLaunch<float, true>::pressure_dual};
    return {SphLaunch<float, false>::density, SphLaunch<float, false>::normal, SphLaunch<float, false>::force_init,
            SphLaunch<float, false>::pressure, SphLaunch<float, false>::pressure_force,
            SphLaunch<float, false>::boundary_init, SphLaunch<float, false>::pressure_dual};
}

template <>
Engine<double>::LaunchTable Engine<double>::launch(bool) {
    return {SphLaunch<double, false>::density, SphLaunch<double, false>::normal,
            SphLaunch<double, false>::force_init, SphLaunch<double, false>::pressure,
            SphLaunch<double, false>::pressure_force, SphLaunch<double, false>::boundary_init,
            SphLaunch<double, false>::pressure_dual};
}

}  // namespace pcisph

using namespace pcisph;

// Arithmetic of the step, chosen with pcisph_set_precision before uploading.
enum { kPrecisionFast = 0, kPrecisionIeee = 1, kPrecisionFp64 = 2 };

struct pcisph_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    Simulation_Params params{};
    bool params_changed = true;
    bool boundary_updated = true;
    float threshold = 0.01f;
    pcisph_solver_options opt{3, 7, 0, 0, 0, 0, 0.2f, 0};
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
    s->eng->resize_public(s->params.fluid_count, s->params.boundary_count);
    s->params_changed = true;
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
            s->params.member = v;                                                      \
            s->params_changed = true;                                                  \
            if (resize) s->eng->resize_public(s->params.fluid_count, s->params.boundary_count); \
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
    return guarded(s, [&] { s->threshold = threshold; });
}

int pcisph_get_params(const pcisph_solver* s, Simulation_Params* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = s->params;
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
    const size_t n = s->eng->field_size(field, s->params);
    if (n == (size_t)-1) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = n;
    return PCISPH_OK;
}

static void check_count(pcisph_solver* s, int field, size_t count) {
    const size_t n = s->eng->field_size(field, s->params);
    if (n == (size_t)-1) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "unknown field " + std::to_string(field));
    if (count != n)
        throw Error(PCISPH_ERR_SIZE, "Inconsistent size (" + std::to_string(count) + ") of field " +
                                         std::to_string(field) + " for " + std::to_string(s->params.fluid_count) +
                                         " fluid / " + std::to_string(s->params.boundary_count) +
                                         " boundary particles (expected " + std::to_string(n) + ")");
}

static int upload_impl(pcisph_solver* s, int field, const void* p, size_t count, bool device) {
    return guarded(s, [&] {
        if (!p && count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer");
        check_count(s, field, count);
        s->eng->upload(field, p, count, device);
        if (field == PCISPH_BOUNDARY_POSITIONS) s->boundary_updated = true;
    });
}

static int download_impl(pcisph_solver* s, int field, void* p, size_t count, bool device, bool f64) {
    return guarded(s, [&] {
        if (!p && count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "null pointer");
        check_count(s, field, count);
        s->eng->download(field, p, count, device, f64);
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

static void apply_param_changes(pcisph_solver* s) {
    if (!s->params_changed) return;
    derive_params(s->params);
    s->eng->realloc_internal(s->params);
    s->eng->zero_velocities();  // Fluid.cpp:462-463
    s->boundary_updated = true;
    s->params_changed = false;
}

static void step_impl(pcisph_solver* s, pcisph_step_stats* stats) {
    apply_param_changes(s);
    pcisph_step_stats st{};
    if (s->params.fluid_count) {
        WarmArgs w;
        if (s->warm_start && s->mlp) w.params = s->mlp;
        if (s->collect_stride && s->samples_x && s->samples_count < s->samples_cap) {
            w.collect_stride = s->collect_stride;
            w.collect_phase = s->collect_phase++;
            w.store_x = s->samples_x;
            w.store_t = s->samples_t;
            w.store_base = s->samples_count;
            w.store_cap = s->samples_cap;
        }
        const StepResult r = s->eng->step(s->params, s->threshold, s->opt,
                                          s->boundary_updated || s->opt.resort_boundary,
                                          s->precision == kPrecisionFast, w);
        s->samples_count = std::min(s->samples_cap, s->samples_count + r.samples);
        st.warm_start = (r.warm ? 1u : 0u) | (r.warm_rejected ? 2u : 0u);
        st.max_density_expansion = (float)r.max_expansion;
        st.list_passes = r.list_passes;
        s->boundary_updated = false;
        st.iterations = r.iterations;
        st.converged = r.converged;
        st.max_density_variation = (float)r.max_dv;
        st.max_density_error = (float)(r.max_dv / (double)s->params.rest_density);
        if (s->precision != kPrecisionFp64)
            st.max_density_error = (float)r.max_dv / s->params.rest_density;
        st.fluid_candidates = r.fluid_candidates;
        st.boundary_candidates = r.boundary_candidates;
    }
    if (stats) *stats = st;
}

int pcisph_step(pcisph_solver* s, pcisph_step_stats* stats) {
    return guarded(s, [&] {
        step_impl(s, stats);
        if (stats && s->params.fluid_count) stats->nonfinite = s->eng->nonfinite();
    });
}

int pcisph_run(pcisph_solver* s, uint32_t n_steps, pcisph_step_stats* stats) {
    return guarded(s, [&] {
        for (uint32_t i = 0; i < n_steps; i++) step_impl(s, stats ? stats + i : nullptr);
        if (stats && n_steps && s->params.fluid_count) stats[n_steps - 1].nonfinite = s->eng->nonfinite();
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
        s->eng->translate_boundary(dx, dy, dz);
        s->boundary_updated = true;
    });
}

int pcisph_set_warm_start(pcisph_solver* s, int enable) {
    return guarded(s, [&] {
        if (enable) s->ensure_mlp();
        s->warm_start = enable != 0;
    });
}

// ------------------------------------------------------------------ autotuner
// Times `trial_steps` steps of the current scene for each block size of the particle
// kernels and each radix digit width, from the same starting state each time (a
// device snapshot restored after every trial), then keeps the fastest pair in the
// solver options.  The state after the call is the state before it.
int pcisph_autotune(pcisph_solver* s, uint32_t trial_steps, pcisph_solver_options* chosen) {
    return guarded(s, [&] {
        if (!s->params.fluid_count) throw Error(PCISPH_ERR_INVALID_ARGUMENT, "no particles to tune on");
        if (!trial_steps) trial_steps = 3;
        apply_param_changes(s);
        // the boundary sort a pending update would do is part of every trial's first step
        const bool boundary_pending = s->boundary_updated;
        s->eng->snapshot();
        const pcisph_solver_options base = s->opt;
        const uint32_t collect = s->collect_stride;
        s->collect_stride = 0;
        double best = 1e300;
        pcisph_solver_options best_opt = base;
        auto trial = [&](uint32_t block, uint32_t digit) {
            s->opt = base;
            s->opt.block_size = block;
            s->opt.sort_digit_bits = digit;
            s->eng->restore();
            s->boundary_updated = boundary_pending;
            step_impl(s, nullptr);  // warm-up: caches, code objects
            s->eng->restore();
            s->boundary_updated = boundary_pending;
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
        for (uint32_t block : {64u, 128u, 256u, 512u}) trial(block, 8u);
        const uint32_t block = best_opt.block_size;
        for (uint32_t digit : {5u, 6u, 7u}) trial(block, digit);
        s->opt = best_opt;
        s->collect_stride = collect;
        s->eng->restore();
        s->boundary_updated = boundary_pending;
        if (chosen) *chosen = best_opt;
    });
}

int pcisph_profile_enable(pcisph_solver* s, int enable) {
    return guarded(s, [&] { s->eng->profile_enable(enable != 0); });
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
