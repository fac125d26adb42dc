// sph_launch.h - host-side launchers of the step's kernels.  The kernels are
// instantiated in sph_ieee.hip (IEEE FP32 and FP64) and sph_fast.hip (fast FP32);
// the solver calls them only through these declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_kernels.h"
#include "sph_persistent.h"
#include "sph_staged.h"

namespace pcisph {

// Device arrays of one batch of scenes (precision R).  Fluid arrays hold the
// batch's N fluid particles, boundary arrays its Nb boundary particles, cell tables
// its B buckets (uint2 start/end); `batch` holds the scene and block tables.
template <typename R>
struct DeviceArrays {
    uint32_t n = 0, nb = 0, B = 0;
    Batch<R> batch{};
    // fluid: public state in the previous step's order, and this step's sorted copies
    V4<R>* pos_in = nullptr;
    V4<R>* vel_in = nullptr;
    V4<R>* pos = nullptr;
    V4<R>* vel = nullptr;
    V4<R>* xr = nullptr;
    V4<R>* nrm = nullptr;
    V4<R>* fo = nullptr;
    V4<R>* fp = nullptr;
    V4<R>* xp = nullptr;
    V4<R>* xf = nullptr;
    R* p = nullptr;
    R* dv = nullptr;
    uint4* counts = nullptr;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint2* cells = nullptr;
    // boundary
    V4<R>* bpos = nullptr;
    V4<R>* bpos_tmp = nullptr;
    V4<R>* bxf = nullptr;
    R* bp = nullptr;
    R* binit = nullptr;
    uint32_t* bkeys[2] = {nullptr, nullptr};
    uint32_t* bvals[2] = {nullptr, nullptr};
    uint2* bcells = nullptr;
    // dense range grids of the scenes' cell boxes (CellTable), `box_cells` entries each
    uint2* fgrid = nullptr;
    uint2* bgrid = nullptr;
    uint32_t box_cells = 0;
    NeighborLists NL;  // on == 0: no lists this step
    bool tile_pressure = false;  // FP32 fluid pressure passes from LDS tiles (256-thread workgroups)
    // FP32 fluid pressure passes from the per-step LDS staging (sph_staged.h)
    bool staged_pressure = false;
    StageMaps stage{};
};

// `block` must be the workgroup size the batch's block table was built for.
struct Launch {
    hipStream_t stream;
    uint32_t block;
    bool onesweep = false;               // radix sorts: the one-sweep variant (radix_onesweep.hip)
    unsigned int* sort_abort = nullptr;  // host-visible flag its bounded look-back raises on expiry
};

// Each launch covers the whole batch: fluid-only kernels run the fluid ranges of the
// block table, the others fluid and boundary ranges.
template <typename R, bool FAST>
struct SphLaunch {
    static void density(const Launch& L, const DeviceArrays<R>& A);
    static void normal(const Launch& L, const DeviceArrays<R>& A);
    static void force_init(const Launch& L, const DeviceArrays<R>& A, const FeatureArgs& F);
    static void pressure(const Launch& L, const DeviceArrays<R>& A, uint32_t iter);
    static void pressure_force(const Launch& L, const DeviceArrays<R>& A, uint32_t iter, HostMirror* host,
                               SceneResult* res, const unsigned int* tag, int warm_pass, V4<R>* xp_save);
    // the whole correction loop as one cooperative launch (k_correction_loop, sph_persistent.h):
    // the grid is the block table's range count, at most what the device holds resident
    static hipError_t correction_loop(const Launch& L, const DeviceArrays<R>& A, uint32_t maxit, HostMirror* host,
                                      SceneResult* res, const unsigned int* tag, const GridSync& g);
    // iteration 0 of the scenes a warm-started step starts from the network (k_pressure_dual,
    // k_select_start, k_adopt_plain_start); the pressure pass runs it for the others
    static void pressure_dual(const Launch& L, const DeviceArrays<R>& A, const DualArrays<R>& D);
    static void boundary_init(const Launch& L, const DeviceArrays<R>& A);
};

// The Morton layout of a step (k_layout_codes, k_layout_keys, k_layout_reorder):
// from the sorted keys K and source slots S of the reference's sort, the new slot
// order of the fluid particles, their positions and velocities moved into it, both
// index maps and the bucket ranges in slots.
template <typename R>
struct LayoutArgs {
    const uint32_t* K;
    const uint32_t* S;
    uint32_t* canon;  // slot -> canonical index: read (previous layout), then rewritten
    uint32_t* stor;   // canonical index -> slot: rewritten
    uint32_t* srcc;   // out: the reference's source locations (canonical indices before the sort)
    uint32_t* code;   // bucket codes, one word per bucket of the batch
    uint32_t* okeys[2];
    uint32_t* ovals[2];
    uint32_t bits, key_bits, digit_bits;  // bits: order_code's packed widths (inside | outside << 8)
    uint32_t* scratch;
    const V4<R>* pin;
    const V4<R>* vin;
    V4<R>* pout;
    V4<R>* vout;
    uint2* cells;  // out: bucket ranges in slots (cleared by the caller)
};

// Precision-only (no arithmetic mode) helpers, instantiated in sph_ieee.hip.
template <typename R>
struct DataLaunch {
    // keys = bucket_base + hash of every fluid (boundary = false) or boundary particle,
    // stored at the particle's canonical index when `canon` maps slots to them
    static void hash(const Launch& L, const Batch<R>& B, bool boundary, const V4<R>* pos, uint32_t* keys,
                     uint32_t* vals, const uint32_t* canon = nullptr);
    // the Morton layout (LayoutArgs); returns the buffer index of the layout sort or -1
    static int layout(const Launch& L, const Batch<R>& B, const LayoutArgs<R>& a);
    // dst[i] = src[map[i]]
    static void gather_v4(const Launch& L, uint32_t n, const uint32_t* map, const V4<R>* src, V4<R>* dst);
    // p = max(0, previous pressure + network output * p_ref of the scene), xf = (x, p / rho^2)
    static void warm_apply(const Launch& L, const DeviceArrays<R>& A);
    static void reorder(const Launch& L, uint32_t n, const uint32_t* keys, const uint32_t* src,
                        const V4<R>* pin, const V4<R>* vin, V4<R>* pout, V4<R>* vout, uint2* cells);
    // float[3n] <-> V4<R> (w set to 0 on the way in); element i of the public array
    // is device element map[i] (i without a map)
    static void unpack_vec3(const Launch& L, const float* src, V4<R>* dst, uint32_t n, const uint32_t* map);
    static void pack_vec3(const Launch& L, const V4<R>* src, float* dst, uint32_t n, const uint32_t* map);
    static void pack_vec3_f64(const Launch& L, const V4<R>* src, double* dst, uint32_t n, const uint32_t* map);
    // float[n] <-> every `stride`-th R starting at `offset`
    static void unpack_scalar(const Launch& L, const float* src, R* dst, uint32_t n, uint32_t stride,
                              uint32_t offset, const uint32_t* map);
    static void pack_scalar(const Launch& L, const R* src, float* dst, uint32_t n, uint32_t stride,
                            uint32_t offset, const uint32_t* map);
    static void pack_scalar_f64(const Launch& L, const R* src, double* dst, uint32_t n, uint32_t stride,
                                uint32_t offset, const uint32_t* map);
    // p[i].xyz += (dx, dy, dz) in precision R
    static void translate(const Launch& L, V4<R>* p, uint32_t n, float dx, float dy, float dz);
    // every scene's boundary positions += shift[3 s .. 3 s + 2] (device array)
    static void translate_scenes(const Launch& L, const Batch<R>& B, const float* shift, V4<R>* bpos);
    // the range grids of all `scene_count` scenes' boxes (the largest has `max_cells`
    // cells) from a bucket table; with `far_pos`, the positions the table's ranges
    // index, entries of far buckets are flagged (kFarBucket) and cells whose whole
    // neighbourhood is empty or far are marked (kQuietCell)
    static void fill_box(const Launch& L, const SphParams<R>* scenes, uint32_t scene_count, uint32_t max_cells,
                         const uint2* cells, uint2* grid, const V4<R>* far_pos = nullptr);
    // cell extents of every scene's fluid and boundary particles into ext[12 S]
    // (k_extents; ext prepared by the caller)
    static void extents(const Launch& L, const Batch<R>& B, const V4<R>* pos, const V4<R>* bpos, int* ext);
    // the per-step staging of the pressure pass (sph_staged.h): the work order -- the
    // fluid slots sorted stably by the order code of their own cell (order_code, `bits`
    // its packed widths) through keys/vals -- then each fluid range's staging map and
    // row descriptors into A.stage (whose `work` must be vals[radix_sort_result(...)])
    static void stage_setup(const Launch& L, const DeviceArrays<R>& A, uint32_t bits, uint32_t key_bits,
                            uint32_t digit_bits, uint32_t* scratch, uint32_t* keys[2], uint32_t* vals[2]);
    // the boundary work order of the staged pressure pass: the boundary slots sorted
    // stably by the Morton code of their own cell (as stage_setup's work order) through
    // keys/vals; returns the index of the buffer pair holding it (radix_sort)
    static int boundary_order(const Launch& L, const DeviceArrays<R>& A, uint32_t bits, uint32_t key_bits,
                              uint32_t digit_bits, uint32_t* scratch, uint32_t* keys[2], uint32_t* vals[2]);
};

void fill_u32(const Launch& L, uint32_t* p, size_t n, uint32_t value);
void fill_iota(const Launch& L, uint32_t* p, uint32_t n);  // p[i] = i
// out[s] = states[s].nonfinite: every scene's flag in one contiguous word array
void nonfinite_flags(const Launch& L, const LoopState* states, uint32_t count, uint32_t* out);
// dst[i] = item map[i] of src, `words` uint32 per item
void gather_words(const Launch& L, uint32_t n, uint32_t words, const uint32_t* map, const uint32_t* src,
                  uint32_t* dst);
// bucket ranges (c, c + 1 transitions) of sorted keys into cells (cleared by the caller)
void cells_from_keys(const Launch& L, uint32_t n, const uint32_t* keys, uint2* cells);
void fill_u64(const Launch& L, unsigned long long* p, size_t n, unsigned long long value);

// Stable LSD radix sort of (key, value) pairs on keys < 2^key_bits.  Sorts from
// buffer 0 through buffer 1 alternately; returns the index (0/1) of the buffer that
// holds the result.  `scratch` must hold radix_sort_scratch_words(n, digit_bits) words.
size_t radix_sort_scratch_words(uint32_t n, uint32_t digit_bits);
int radix_sort(const Launch& L, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
               uint32_t digit_bits, uint32_t* scratch);
// The buffer index radix_sort returns for these sizes (known before it is enqueued).
int radix_sort_result(uint32_t n, uint32_t key_bits, uint32_t digit_bits);
// The one-sweep variant radix_sort runs for L.onesweep (same contract and result
// index); its scratch need, included in radix_sort_scratch_words.
size_t radix_onesweep_scratch_words(uint32_t n, uint32_t digit_bits);
int radix_sort_onesweep(const Launch& L, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
                        uint32_t digit_bits, uint32_t* scratch);

}  // namespace pcisph
