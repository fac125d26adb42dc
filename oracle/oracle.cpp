/*
 * oracle.cpp - CPU oracle of the PCISPH step.  TEST INFRASTRUCTURE ONLY: used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker;
 * never linked into or called by the product library.
 *
 * A plain restatement of the reference, function by function, in the reference's
 * order of floating-point operations and summation (boundary buckets before fluid
 * buckets, the 27 bucket offsets x-fastest, ascending particle index inside a
 * bucket: PCISPH/data/kernels/grid_utils.cl:65-79).  Templated on the scalar type:
 * `float` is the parity oracle, `double` is the FP64 reference run.  Per-particle
 * loops run under OpenMP; each particle's result depends only on its own loop, so
 * the thread count never changes a bit.
 *
 * Pinning (each function is checked by tests/test_oracle.py against something
 * other than itself):
 *   derive_params     -> float64 closed forms of the constants and of the PCISPH
 *                        delta lattice sum (tests/test_oracle.py::test_params_*)
 *   cell_pos/hash_key -> Python big-int known answers, distinctness of the 27
 *                        neighbour keys when B % 64 == 0, NaN/saturation cases
 *   sort + cell table -> numpy stable argsort and brute-force run detection
 *   update_density    -> isolated particle rho = m K6 h^6, lattice interior
 *                        rho/rho0 = 315*330/(64*pi*512), O(N^2) brute force
 *   update_normal     -> zero at a symmetric lattice centre, two-particle form
 *   force_init        -> isolated particle F = (0, m g, 0), zero viscosity for a
 *                        uniform velocity, antisymmetric cohesion pair forces;
 *                        magnitudes: float64 two-particle closed forms of viscosity
 *                        and of cohesion + curvature (both C(r) branches, K_ij) and a
 *                        brute-force cluster sum from the published formulas
 *                        (tests/test_oracle_forces.py)
 *   integrate         -> free-fall closed form
 *   update_pressure   -> dv = max(0, rho* - rho0) and p += delta*dv by hand on
 *                        a two-particle scene; boundary-mode pass with init term
 *   pressure_force    -> fluid-only scenes: sum of forces ~ 0 (antisymmetry);
 *                        magnitudes: float64 brute force of the spiky-gradient sum
 *                        fluid-fluid, and fluid-boundary with rho_b = rho0 (with and
 *                        without the boundary-boundary pass) (tests/test_oracle_forces.py)
 *   correction loop   -> min/max iteration semantics on tiny scenes
 * OpenCL built-ins are restated as IEEE operations: distance()/fast_length() as
 * sqrt(dot), `/` correctly rounded, no FP contraction (built with
 * -ffp-contract=off).  The reference leaves these implementation-defined, so
 * their exact bits are "parity unpinned" beyond those tolerance checks.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr float kPi = 3.14159265358979323846f;  // PCISPH/src/utils/constants.h:4

struct I3 {
    int32_t x, y, z;
};

// convert_int3_sat_rtn(pos / cell_size): round toward -inf, saturate, NaN -> 0
// (PCISPH/data/kernels/grid_utils.cl:9-11).
template <typename T>
int32_t sat_floor(T v) {
    if (std::isnan(v)) return 0;
    T f = std::floor(v);
    if (f >= (T)2147483648.0) return INT32_MAX;
    if (f < (T)-2147483648.0) return INT32_MIN;
    return (int32_t)f;
}

template <typename T>
I3 cell_pos(T x, T y, T z, T cell) {
    return {sat_floor<T>(x / cell), sat_floor<T>(y / cell), sat_floor<T>(z / cell)};
}

// get_hash_key (grid_utils.cl:12-28), active `#if 1` variant, in 32-bit
// two's-complement wrap arithmetic.
uint32_t hash_key(I3 c, uint32_t B) {
    const uint32_t x = (uint32_t)c.x, y = (uint32_t)c.y, z = (uint32_t)c.z;
    const uint32_t g = (x * 19349663u) ^ (y * 73856093u) ^ (z * 83492791u);
    const uint32_t v = (g << 6) + (x & 63u) + 4u * (y & 63u) + 16u * (z & 63u);
    return (v & 0x7FFFFFFFu) % B;
}

// 27 neighbour offsets, i -> ((i%3)-1, ((i/3)%3)-1, ((i/9)%3)-1) (grid_utils.cl:69)
inline I3 neighbour_cell(I3 c, int i) {
    return {(int32_t)((uint32_t)c.x + (uint32_t)(i % 3 - 1)),
            (int32_t)((uint32_t)c.y + (uint32_t)((i / 3) % 3 - 1)),
            (int32_t)((uint32_t)c.z + (uint32_t)((i / 9) % 3 - 1))};
}

// Fluid.cpp:125-133
uint32_t sort_bit_count(uint32_t elements) {
    uint32_t max_value = 0;
    for (int i = 0; i < 32; i++) {
        max_value = (max_value << 1) | 1;
        if (max_value >= elements - 1) return (uint32_t)(i + 1);
    }
    throw std::runtime_error("sort_bit_count failed");
}

// Fluid::update_deduced_attributes (Fluid.cpp:377-443), with C++11 <cmath>
// overload resolution: pow(float, float) is the float overload, pow(float, int)
// promotes to double.
void derive_params(Simulation_Params& p) {
    p.kernel_radius = 4.f * p.particle_radius;
    p.kernel_radius2 = std::pow(p.kernel_radius, 2.f);
    p.particle_mass = (float)(p.rest_density / std::pow(1.f / (2.f * p.particle_radius), 3));
    p.bucket_count = p.fluid_count / 2;
    p.bucket_count -= p.bucket_count % 64;
    p.bucket_count = std::max(64u, p.bucket_count);
    p.cell_size = p.kernel_radius;
    p.poly6_normalization = 315.f / (64.f * kPi * std::pow(p.kernel_radius, 9.f));
    p.poly6_d1_normalization = -945.f / (32.f * kPi * std::pow(p.kernel_radius, 9.f));
    p.viscosity_d2_normalization = 45.f / (kPi * std::pow(p.kernel_radius, 6.0f));
    p.spiky_d1_normalization = -45.f / (kPi * std::pow(p.kernel_radius, 6.0f));
    p.surface_tension_normalization = 32.f / (kPi * std::pow(p.kernel_radius, 9.f));
    p.surface_tension_term = std::pow(p.kernel_radius, 6.f) / 64.f;
    const float particle_size = 2.f * p.particle_radius;
    if (particle_size == 0.f) {  // the lattice loop would not advance (the reference hangs)
        p.density_variation_scaling_factor = 0.f;
        return;
    }
    const float beta = p.delta_t * p.delta_t * p.particle_mass * p.particle_mass * 2.f /
                       (p.rest_density * p.rest_density);
    float sum[3] = {0.f, 0.f, 0.f};
    float dot_sum = 0.f;
    const float lo = -p.kernel_radius - particle_size, hi = p.kernel_radius + particle_size;
    for (float z = lo; z <= hi; z += 2.f * p.particle_radius)
        for (float y = lo; y <= hi; y += 2.f * p.particle_radius)
            for (float x = lo; x <= hi; x += 2.f * p.particle_radius) {
                const float r = std::sqrt(x * x + y * y + z * z);
                const float r2 = r * r;
                if (r2 < p.kernel_radius2) {
                    const float f = (float)(p.poly6_d1_normalization *
                                            std::pow(p.kernel_radius2 - r2, 2));
                    const float g[3] = {-f * x, -f * y, -f * z};
                    for (int i = 0; i < 3; i++) {
                        sum[i] += g[i];
                        dot_sum += g[i] * g[i];
                    }
                }
            }
    float sum_dot = 0.f;
    for (int i = 0; i < 3; i++) sum_dot += sum[i] * sum[i];
    p.density_variation_scaling_factor = -1.f / (beta * (-sum_dot - dot_sum));
}

// ---------------------------------------------------------------- kernels
template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
inline V3<T> ld3(const std::vector<T>& a, size_t i) {
    return {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
}
template <typename T>
inline void st3(std::vector<T>& a, size_t i, V3<T> v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}
template <typename T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T>
inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// smoothing kernels (pcisph.cl:8-61)
template <typename T>
inline T poly6(T r2, T h2) {
    if (r2 > h2) return T(0);
    const T dr = h2 - r2;
    return dr * dr * dr;
}
template <typename T>
inline V3<T> poly6_d1(V3<T> d, T h2) {
    const T r2 = dot(d, d);
    if (r2 > h2) return {T(0), T(0), T(0)};
    return ((r2 - h2) * (r2 - h2)) * d;
}
template <typename T>
inline V3<T> spiky_d1(V3<T> d, T h) {
    const T r = std::sqrt(dot(d, d));  // fast_length
    if (r > h || r < T(0.0001f)) return {T(0), T(0), T(0)};
    const T dr = h - r;
    return d * dr * dr / r;
}
template <typename T>
inline T viscosity_d2(T r, T h) {
    return r > h ? T(0) : h - r;
}
template <typename T>
inline T surface_tension(T r, T h, T st_term) {
    if ((double)r <= 0.5 * (double)h) {
        const T t1 = h - r;
        const T t2 = (t1 * t1 * t1) * (r * r * r);
        return T(2) * t2 - st_term;
    } else if (r <= h) {
        const T t1 = h - r;
        return (t1 * t1 * t1) * (r * r * r);
    }
    return T(0);
}

struct CellTable {
    std::vector<uint32_t> offsets;  // (start, end) per bucket
};

template <typename T>
struct Sim {
    Simulation_Params params{};
    bool params_changed = true;
    bool boundary_updated = true;
    float threshold = 0.01f;
    pcisph_solver_options opt{3, 7, 0, 0, 0, 0, 0.2f, 0, 0, 0};  // launch options unused: nothing is launched
    // public particle state
    std::vector<T> b_pos, b_p;
    std::vector<T> f_pos, f_n, f_pred, f_rho, f_fo, f_v, f_p, f_fp, f_dv;
    // internal buffers (Fluid.h:88-99)
    std::vector<uint32_t> b_cells, b_keys, b_src, f_cells, f_keys, f_src, f_counts;
    std::vector<T> b_init;

    uint32_t N() const { return params.fluid_count; }
    uint32_t Nb() const { return params.boundary_count; }

    void resize_public() {
        b_pos.assign(3 * (size_t)Nb(), T(0));
        b_p.assign(Nb(), T(0));
        for (auto* v : {&f_pos, &f_n, &f_pred, &f_fo, &f_v, &f_fp}) v->assign(3 * (size_t)N(), T(0));
        for (auto* v : {&f_rho, &f_p, &f_dv}) v->assign(N(), T(0));
    }

    void update_deduced_attributes() {
        derive_params(params);
        const size_t B = params.bucket_count;
        if (Nb() > 0) {
            b_cells.assign(2 * B, 0);
            b_keys.assign(Nb(), 0);
            b_src.assign(Nb(), 0);
            b_init.assign(Nb(), T(0));  // never written by the reference: defined as zero
        }
        f_cells.assign(2 * B, 0);
        f_keys.assign(N(), 0);
        f_src.assign(N(), 0);
        f_dv.assign(N(), T(0));
        f_counts.assign(4 * (size_t)N(), 0);
        std::fill(f_v.begin(), f_v.end(), T(0));  // Fluid.cpp:462-463
        boundary_updated = true;
    }

    // sort_utils.cl:7-78 plus the stable key/value sort (clogs postcondition,
    // radixsort.h:182-184)
    void sort(uint32_t count, std::vector<T>& pos, std::vector<T>* vel, std::vector<uint32_t>& cells,
              std::vector<uint32_t>& keys, std::vector<uint32_t>& src) {
        const uint32_t B = params.bucket_count;
        std::fill(cells.begin(), cells.end(), 0u);
        const T cs = (T)params.cell_size;
        for (uint32_t i = 0; i < count; i++) {
            V3<T> p = ld3(pos, i);
            keys[i] = hash_key(cell_pos(p.x, p.y, p.z, cs), B);
        }
        std::vector<uint32_t> order(count);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
        std::vector<uint32_t> sk(count);
        for (uint32_t i = 0; i < count; i++) sk[i] = keys[order[i]];
        keys.swap(sk);
        src = order;
        std::vector<T> tmp = pos;
        for (uint32_t d = 0; d < count; d++) st3(pos, d, ld3(tmp, src[d]));
        if (vel) {
            std::vector<T> tv = *vel;
            for (uint32_t d = 0; d < count; d++) st3(*vel, d, ld3(tv, src[d]));
        }
        for (uint32_t d = 0; d < count; d++) {
            const uint32_t k = keys[d];
            if (d == 0 || k != keys[d - 1]) cells[2 * k] = d;
            if (d == count - 1 || k != keys[d + 1]) cells[2 * k + 1] = d + 1;
        }
    }

    template <typename F>
    void foreach_neighbor(const std::vector<uint32_t>& cells, V3<T> pos, F&& body) const {
        const I3 c = cell_pos(pos.x, pos.y, pos.z, (T)params.cell_size);
        for (int i = 0; i < 27; i++) {
            const uint32_t k = hash_key(neighbour_cell(c, i), params.bucket_count);
            const uint32_t start = cells[2 * k], end = cells[2 * k + 1];
            for (uint32_t j = start; j < end; j++) body(j);
        }
    }

    // pcisph.cl:65-92 (plus the neighbour statistics of PCISPH_FLUID_NEIGHBOR_COUNTS)
    void update_density() {
        const T h2 = (T)params.kernel_radius2;
        const T scale = (T)params.particle_mass * (T)params.poly6_normalization;
        const bool has_b = Nb() > 0;
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            const V3<T> self = ld3(f_pos, i);
            T density = T(0);
            uint32_t cf = 0, cb = 0, nf = 0, nb = 0;
            if (has_b)
                foreach_neighbor(b_cells, self, [&](uint32_t j) {
                    const V3<T> d = self - ld3(b_pos, j);
                    const T r2 = dot(d, d);
                    density += poly6(r2, h2);
                    cb++;
                    nb += !(r2 > h2);
                });
            foreach_neighbor(f_cells, self, [&](uint32_t j) {
                const V3<T> d = self - ld3(f_pos, j);
                const T r2 = dot(d, d);
                density += poly6(r2, h2);
                cf++;
                nf += !(r2 > h2);
            });
            density *= scale;
            f_rho[i] = density;
            f_counts[4 * i + 0] = cf;
            f_counts[4 * i + 1] = cb;
            f_counts[4 * i + 2] = nf;
            f_counts[4 * i + 3] = nb;
        }
    }

    // pcisph.cl:94-112
    void update_normal() {
        const T h2 = (T)params.kernel_radius2;
        const T scale = (T)params.kernel_radius * (T)params.particle_mass * (T)params.poly6_d1_normalization;
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            const V3<T> self = ld3(f_pos, i);
            V3<T> n{T(0), T(0), T(0)};
            foreach_neighbor(f_cells, self, [&](uint32_t j) {
                const V3<T> d = self - ld3(f_pos, j);
                n = n + poly6_d1(d, h2) / f_rho[j];
            });
            st3(f_n, i, n * scale);
        }
    }

    // pcisph.cl:121-173
    void force_initialization() {
        const T h = (T)params.kernel_radius;
        const T rho0 = (T)params.rest_density;
        const T m = (T)params.particle_mass;
        const T st_term = (T)params.surface_tension_term;
        const T visc_scale = (T)params.viscosity_constant * m * (T)params.viscosity_d2_normalization;
        const T coh_scale = -(T)params.surface_tension_coefficient * m * m * (T)params.surface_tension_normalization;
        const T curv_scale = -(T)params.surface_tension_coefficient * m;
        const V3<T> gravity{T(0), m * (T)params.gravity, T(0)};
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            const V3<T> self = ld3(f_pos, i), sv = ld3(f_v, i), sn = ld3(f_n, i);
            const T srho = f_rho[i];
            V3<T> visc{T(0), T(0), T(0)}, coh{T(0), T(0), T(0)}, curv{T(0), T(0), T(0)};
            foreach_neighbor(f_cells, self, [&](uint32_t j) {
                const V3<T> op = ld3(f_pos, j), ov = ld3(f_v, j), on = ld3(f_n, j);
                const T orho = f_rho[j];
                const V3<T> d = self - op;
                const T dist = std::sqrt(dot(d, d));
                if (orho > T(0.0001f)) visc = visc + (ov - sv) * (viscosity_d2(dist, h) / orho);
                if (dist > T(0.0001f) && dist < h) {
                    const T corr = T(2) * rho0 / (srho + orho);
                    const T stk = surface_tension(dist, h, st_term);
                    const V3<T> dir = (self - op) / dist;
                    coh = coh + (corr * stk) * dir;
                    curv = curv + corr * (sn - on);
                }
            });
            visc = visc * visc_scale;
            coh = coh * coh_scale;
            curv = curv * curv_scale;
            const V3<T> st = coh + curv;
            st3(f_fo, i, (gravity + visc) + st);
            f_p[i] = T(0);
            st3(f_fp, i, V3<T>{T(0), T(0), T(0)});
        }
    }

    // pcisph.cl:175-193; predict writes only positions (vel_out == NULL)
    void integrate(std::vector<T>& pos_out, bool write_vel) {
        const T m = (T)params.particle_mass, dt = (T)params.delta_t;
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            V3<T> p = ld3(f_pos, i), v = ld3(f_v, i);
            const V3<T> f = ld3(f_fo, i) + ld3(f_fp, i);
            const V3<T> a = f / m;
            v = v + a * dt;
            p = p + v * dt;
            st3(pos_out, i, p);
            if (write_vel) st3(f_v, i, v);
        }
    }

    // pcisph.cl:195-210 (opt-in; never launched by the reference)
    void init_boundary_boundary() {
        const T h2 = (T)params.kernel_radius2;
#pragma omp parallel for schedule(static)
        for (int64_t bb = 0; bb < (int64_t)Nb(); bb++) {
            const V3<T> self = ld3(b_pos, (size_t)bb);
            T r = T(0);
            foreach_neighbor(b_cells, self, [&](uint32_t j) {
                const V3<T> d = self - ld3(b_pos, j);
                r += poly6(dot(d, d), h2);
            });
            b_init[(size_t)bb] = r;
        }
    }

    // pcisph.cl:212-263, boundary mode then fluid mode
    void update_pressure() {
        const T h2 = (T)params.kernel_radius2;
        const T scale = (T)params.particle_mass * (T)params.poly6_normalization;
        const T rho0 = (T)params.rest_density;
        const T delta = (T)params.density_variation_scaling_factor;
        if (Nb() > 0) {
#pragma omp parallel for schedule(static)
            for (int64_t bb = 0; bb < (int64_t)Nb(); bb++) {
                const size_t b = (size_t)bb;
                const V3<T> self = ld3(b_pos, b);
                T pd = T(0);
                pd += b_init[b];
                foreach_neighbor(f_cells, self, [&](uint32_t j) {
                    const V3<T> d = self - ld3(f_pred, j);
                    pd += poly6(dot(d, d), h2);
                });
                pd *= scale;
                const T dv = std::fmax(T(0), pd - rho0);
                if (dv == T(0)) continue;
                b_p[b] += dv * delta;
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            const V3<T> self = ld3(f_pos, i), pred = ld3(f_pred, i);
            T pd = T(0);
            if (Nb() > 0)
                foreach_neighbor(b_cells, self, [&](uint32_t j) {
                    const V3<T> d = pred - ld3(b_pos, j);
                    pd += poly6(dot(d, d), h2);
                });
            foreach_neighbor(f_cells, self, [&](uint32_t j) {
                const V3<T> d = pred - ld3(f_pred, j);
                pd += poly6(dot(d, d), h2);
            });
            pd *= scale;
            const T dv = std::fmax(T(0), pd - rho0);
            f_dv[i] = dv;
            if (dv == T(0)) continue;
            f_p[i] += dv * delta;
        }
    }

    // pcisph.cl:265-300
    void update_pressure_force() {
        const T h = (T)params.kernel_radius;
        const T rho0 = (T)params.rest_density;
        const T m = (T)params.particle_mass;
        const T scale = -(T)params.spiky_d1_normalization * m * m;
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N(); ii++) {
            const size_t i = (size_t)ii;
            const V3<T> self = ld3(f_pos, i);
            const T sp = f_p[i], srho = f_rho[i];
            const T sf = sp / (srho * srho);
            V3<T> pf{T(0), T(0), T(0)};
            if (Nb() > 0)
                foreach_neighbor(b_cells, self, [&](uint32_t j) {
                    const T of = b_p[j] / (rho0 * rho0);
                    pf = pf + spiky_d1(self - ld3(b_pos, j), h) * (sf + of);
                });
            foreach_neighbor(f_cells, self, [&](uint32_t j) {
                const T orho = f_rho[j];
                const T of = f_p[j] / (orho * orho);
                pf = pf + spiky_d1(self - ld3(f_pos, j), h) * (sf + of);
            });
            st3(f_fp, i, pf * scale);
        }
    }

    // Fluid::update (Fluid.cpp:108-327)
    void step(pcisph_step_stats* stats) {
        if (params_changed) {
            update_deduced_attributes();
            params_changed = false;
        }
        pcisph_step_stats st{};
        if (N() == 0) {
            if (stats) *stats = st;
            return;
        }
        if ((boundary_updated || opt.resort_boundary) && Nb() > 0) {
            sort(Nb(), b_pos, nullptr, b_cells, b_keys, b_src);
            if (opt.launch_boundary_init) init_boundary_boundary();
            boundary_updated = false;
        }
        sort(N(), f_pos, &f_v, f_cells, f_keys, f_src);
        update_density();
        update_normal();
        if (Nb() > 0) std::fill(b_p.begin(), b_p.end(), T(0));
        force_initialization();
        T max_dv = T(0);
        uint32_t it = 0;
        bool converged = false;
        for (uint32_t i = 0; i < opt.max_iterations; i++) {
            integrate(f_pred, false);
            update_pressure();
            update_pressure_force();
            it = i + 1;
            max_dv = *std::max_element(f_dv.begin(), f_dv.end());
            if (i + 1 >= opt.min_iterations) {
                if (max_dv / (T)params.rest_density < (T)threshold) {
                    converged = true;
                    break;
                }
            }
        }
        integrate(f_pos, true);
        st.iterations = it;
        st.converged = converged ? 1u : 0u;
        st.max_density_variation = (float)max_dv;
        st.max_density_error = (float)(max_dv / (T)params.rest_density);
        bool bad = false;
        for (T v : f_pos) bad |= !std::isfinite(v);
        for (T v : f_v) bad |= !std::isfinite(v);
        st.nonfinite = bad ? 1u : 0u;
        uint64_t cf = 0, cb = 0;
        for (size_t i = 0; i < N(); i++) {
            cf += f_counts[4 * i];
            cb += f_counts[4 * i + 1];
        }
        st.fluid_candidates = cf;
        st.boundary_candidates = cb;
        if (stats) *stats = st;
    }

    size_t field_size(int field) const {
        const size_t n = N(), nb = Nb(), B = params.bucket_count;
        switch (field) {
            case PCISPH_FLUID_POSITIONS:
            case PCISPH_FLUID_VELOCITIES:
            case PCISPH_FLUID_NORMALS:
            case PCISPH_FLUID_PREDICTED_POSITIONS:
            case PCISPH_FLUID_OTHER_FORCES:
            case PCISPH_FLUID_PRESSURE_FORCES: return 3 * n;
            case PCISPH_FLUID_DENSITIES:
            case PCISPH_FLUID_PRESSURES:
            case PCISPH_FLUID_DENSITY_VARIATIONS:
            case PCISPH_FLUID_KEYS:
            case PCISPH_FLUID_SRC_LOCATIONS: return n;
            case PCISPH_BOUNDARY_POSITIONS: return 3 * nb;
            case PCISPH_BOUNDARY_PRESSURES:
            case PCISPH_BOUNDARY_KEYS:
            case PCISPH_BOUNDARY_SRC_LOCATIONS:
            case PCISPH_BOUNDARY_INIT_PRED_DENSITIES: return nb;
            case PCISPH_FLUID_CELL_OFFSETS: return 2 * B;
            case PCISPH_BOUNDARY_CELL_OFFSETS: return nb ? 2 * B : 0;
            case PCISPH_FLUID_NEIGHBOR_COUNTS: return 4 * n;
        }
        return (size_t)-1;
    }

    std::vector<T>* float_field(int field) {
        switch (field) {
            case PCISPH_FLUID_POSITIONS: return &f_pos;
            case PCISPH_FLUID_VELOCITIES: return &f_v;
            case PCISPH_FLUID_NORMALS: return &f_n;
            case PCISPH_FLUID_PREDICTED_POSITIONS: return &f_pred;
            case PCISPH_FLUID_DENSITIES: return &f_rho;
            case PCISPH_FLUID_OTHER_FORCES: return &f_fo;
            case PCISPH_FLUID_PRESSURES: return &f_p;
            case PCISPH_FLUID_PRESSURE_FORCES: return &f_fp;
            case PCISPH_FLUID_DENSITY_VARIATIONS: return &f_dv;
            case PCISPH_BOUNDARY_POSITIONS: return &b_pos;
            case PCISPH_BOUNDARY_PRESSURES: return &b_p;
            case PCISPH_BOUNDARY_INIT_PRED_DENSITIES: return &b_init;
        }
        return nullptr;
    }
    std::vector<uint32_t>* uint_field(int field) {
        switch (field) {
            case PCISPH_FLUID_KEYS: return &f_keys;
            case PCISPH_FLUID_SRC_LOCATIONS: return &f_src;
            case PCISPH_FLUID_CELL_OFFSETS: return &f_cells;
            case PCISPH_BOUNDARY_KEYS: return &b_keys;
            case PCISPH_BOUNDARY_SRC_LOCATIONS: return &b_src;
            case PCISPH_BOUNDARY_CELL_OFFSETS: return &b_cells;
            case PCISPH_FLUID_NEIGHBOR_COUNTS: return &f_counts;
        }
        return nullptr;
    }
};

struct Base {
    virtual ~Base() = default;
};
template <typename T>
struct Holder : Base {
    Sim<T> sim;
};

}  // namespace

struct oracle_sim {
    int fp64 = 0;
    Holder<float> f32;
    Holder<double> f64;
};

template <typename Fn>
static int dispatch(oracle_sim* s, Fn&& fn) {
    if (!s) return PCISPH_ERR_INVALID_ARGUMENT;
    try {
        if (s->fp64) return fn(s->f64.sim);
        return fn(s->f32.sim);
    } catch (const std::exception&) {
        return PCISPH_ERR_INTERNAL;
    }
}

extern "C" {

oracle_sim* oracle_create(int fp64) {
    auto* s = new oracle_sim();
    s->fp64 = fp64 ? 1 : 0;
    return s;
}
void oracle_destroy(oracle_sim* s) { delete s; }

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

#define ORACLE_SETTER(name, type, field, resize)                        \
    int oracle_set_##name(oracle_sim* s, type v) {                      \
        return dispatch(s, [&](auto& sim) {                             \
            sim.params.field = v;                                       \
            sim.params_changed = true;                                  \
            if (resize) sim.resize_public();                            \
            return 0;                                                   \
        });                                                             \
    }
ORACLE_SETTER(boundary_count, uint32_t, boundary_count, true)
ORACLE_SETTER(fluid_count, uint32_t, fluid_count, true)
ORACLE_SETTER(delta_t, float, delta_t, false)
ORACLE_SETTER(rest_density, float, rest_density, false)
ORACLE_SETTER(particle_radius, float, particle_radius, false)
ORACLE_SETTER(gravity, float, gravity, false)
ORACLE_SETTER(viscosity, float, viscosity_constant, false)
ORACLE_SETTER(surface_tension, float, surface_tension_coefficient, false)
#undef ORACLE_SETTER

int oracle_set_density_variation_threshold(oracle_sim* s, float v) {
    return dispatch(s, [&](auto& sim) {
        sim.threshold = v;
        return 0;
    });
}

int oracle_get_params(const oracle_sim* s, Simulation_Params* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = s->fp64 ? s->f64.sim.params : s->f32.sim.params;
    return 0;
}

int oracle_set_solver_options(oracle_sim* s, const pcisph_solver_options* o) {
    if (!o || o->max_iterations == 0) return PCISPH_ERR_INVALID_ARGUMENT;
    return dispatch(s, [&](auto& sim) {
        sim.opt = *o;
        return 0;
    });
}

int oracle_field_size(const oracle_sim* s, int field, size_t* out) {
    if (!s || !out) return PCISPH_ERR_INVALID_ARGUMENT;
    size_t n = s->fp64 ? s->f64.sim.field_size(field) : s->f32.sim.field_size(field);
    if (n == (size_t)-1) return PCISPH_ERR_INVALID_ARGUMENT;
    *out = n;
    return 0;
}

int oracle_upload(oracle_sim* s, int field, const void* host, size_t count) {
    return dispatch(s, [&](auto& sim) {
        using T = typename std::remove_reference<decltype(sim.f_pos)>::type::value_type;
        auto* v = sim.float_field(field);
        if (!v || !host) return (int)PCISPH_ERR_INVALID_ARGUMENT;
        if (count != sim.field_size(field)) return (int)PCISPH_ERR_SIZE;
        const float* h = (const float*)host;
        v->resize(count);
        for (size_t i = 0; i < count; i++) (*v)[i] = (T)h[i];
        if (field == PCISPH_BOUNDARY_POSITIONS) sim.boundary_updated = true;
        return 0;
    });
}

int oracle_download(oracle_sim* s, int field, void* host, size_t count) {
    return dispatch(s, [&](auto& sim) {
        if (!host) return (int)PCISPH_ERR_INVALID_ARGUMENT;
        if (count != sim.field_size(field)) return (int)PCISPH_ERR_SIZE;
        if (auto* v = sim.float_field(field)) {
            float* h = (float*)host;
            for (size_t i = 0; i < count; i++) h[i] = i < v->size() ? (float)(*v)[i] : 0.f;
            return 0;
        }
        if (auto* u = sim.uint_field(field)) {
            uint32_t* h = (uint32_t*)host;
            for (size_t i = 0; i < count; i++) h[i] = i < u->size() ? (*u)[i] : 0u;
            return 0;
        }
        return (int)PCISPH_ERR_INVALID_ARGUMENT;
    });
}

int oracle_download_f64(oracle_sim* s, int field, double* host, size_t count) {
    return dispatch(s, [&](auto& sim) {
        auto* v = sim.float_field(field);
        if (!v || !host) return (int)PCISPH_ERR_INVALID_ARGUMENT;
        if (count != sim.field_size(field)) return (int)PCISPH_ERR_SIZE;
        for (size_t i = 0; i < count; i++) host[i] = i < v->size() ? (double)(*v)[i] : 0.0;
        return 0;
    });
}

int oracle_step(oracle_sim* s, pcisph_step_stats* stats) {
    return dispatch(s, [&](auto& sim) {
        sim.step(stats);
        return 0;
    });
}

uint32_t oracle_hash_key(int32_t cx, int32_t cy, int32_t cz, uint32_t B) { return hash_key({cx, cy, cz}, B); }

void oracle_cell_pos(float x, float y, float z, float cell, int32_t out[3]) {
    I3 c = cell_pos<float>(x, y, z, cell);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

uint32_t oracle_sort_bit_count(uint32_t n) {
    try {
        return sort_bit_count(n);
    } catch (...) {
        return 0;
    }
}

int oracle_compute_params(float particle_radius, float rest_density, uint32_t boundary_count,
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
    return 0;
}

}  // extern "C"
