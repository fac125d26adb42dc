// sph_ieee.hip - IEEE instantiations of the step kernels (built with
// -ffp-contract=off): FP32 in the reference's order of operations, and FP64.
// Also the precision-only data kernels (hash, reorder, layout conversion).
#include <algorithm>

#include "sph_launch_impl.h"

namespace pcisph {

namespace {

// Layout conversion between the public float[3n] / float[n] arrays (element i) and
// the device arrays (element map[i], or i without a map).
template <typename R>
__global__ void k_unpack_vec3(const float* __restrict__ src, V4<R>* __restrict__ dst, uint32_t n,
                              const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V4<R> v;
    v.x = (R)src[3 * (size_t)i];
    v.y = (R)src[3 * (size_t)i + 1];
    v.z = (R)src[3 * (size_t)i + 2];
    v.w = 0;
    dst[map ? map[i] : i] = v;
}

template <typename R, typename O>
__global__ void k_pack_vec3(const V4<R>* __restrict__ src, O* __restrict__ dst, uint32_t n,
                            const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V4<R> v = src[map ? map[i] : i];
    dst[3 * (size_t)i] = (O)v.x;
    dst[3 * (size_t)i + 1] = (O)v.y;
    dst[3 * (size_t)i + 2] = (O)v.z;
}

template <typename R>
__global__ void k_unpack_scalar(const float* __restrict__ src, R* __restrict__ dst, uint32_t n,
                                uint32_t stride, uint32_t offset, const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[(size_t)(map ? map[i] : i) * stride + offset] = (R)src[i];
}

template <typename R, typename O>
__global__ void k_pack_scalar(const R* __restrict__ src, O* __restrict__ dst, uint32_t n, uint32_t stride,
                              uint32_t offset, const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = (O)src[(size_t)(map ? map[i] : i) * stride + offset];
}

// dst[i] = item map[i] of src (`words` uint32 per item)
__global__ void k_gather_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t n,
                               uint32_t words, const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t w = 0; w < words; ++w) dst[(size_t)i * words + w] = src[(size_t)map[i] * words + w];
}

template <typename R>
__global__ void k_gather_v4(const V4<R>* __restrict__ src, V4<R>* __restrict__ dst, uint32_t n,
                            const uint32_t* __restrict__ map) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_nonfinite_flags(const LoopState* __restrict__ states, uint32_t count, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < count) out[s] = states[s].nonfinite;
}

// Bucket ranges in canonical indices from the sorted keys (the reference's cell table).
__global__ void k_cells_from_keys(uint32_t n, const uint32_t* __restrict__ keys, uint2* __restrict__ cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const uint32_t k = keys[c];
    if (c == 0 || k != keys[c - 1]) cells[k].x = c;
    if (c == n - 1 || k != keys[c + 1]) cells[k].y = c + 1;
}

template <typename R>
__global__ void k_translate(V4<R>* __restrict__ p, uint32_t n, R dx, R dy, R dz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V4<R> v = p[i];
    v.x = v.x + dx;
    v.y = v.y + dy;
    v.z = v.z + dz;
    p[i] = v;
}


// The Morton layout of the fluid particles.  The reference's arrays are in bucket
// order (sort_utils.cl:25-78); this build stores each bucket's particles contiguously
// and in the reference's order, but orders the buckets along a Morton curve, so the
// candidates a neighbourhood gathers lie together in memory.  Kernels index storage
// slots; `canon` (slot -> canonical index, the index in the reference's arrays) and
// `stor` (its inverse) translate at the public interface.
//
// Part 1, over canonical indices c of the sorted keys K (S[c] = the particle's slot
// before this step): the reference's source location of c (the canonical index the
// particle had before the sort) and, at the first c of each bucket, the bucket's
// layout code: the order code of that particle's cell (order_code, `bits` its packed
// widths).
template <typename R>
__global__ void k_layout_codes(const SphParams<R>* __restrict__ scenes, const BlockRange* __restrict__ blocks,
                               const uint32_t* __restrict__ K, const uint32_t* __restrict__ S,
                               const uint32_t* __restrict__ canon, const V4<R>* __restrict__ pos, uint32_t bits,
                               uint32_t* __restrict__ code, uint32_t* __restrict__ srcc) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t c = rg.begin + threadIdx.x;
    if (c >= rg.end) return;
    const uint32_t s = S[c];
    srcc[c] = canon[s];
    const uint32_t k = K[c];
    if (c == 0 || K[c - 1] != k) {
        const SphParams<R>& P = scenes[rg.scene];
        code[k] = order_code<R>(P, rg.scene, cell_of4<R>(pos[s], P.cell), bits);
    }
}

// Part 2: every canonical index takes its bucket's code; sorted stably, the values
// give the new layout (slot -> canonical index) with buckets whole and in the
// reference's order inside.
__global__ void k_layout_keys(uint32_t n, const uint32_t* __restrict__ K, const uint32_t* __restrict__ code,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    keys[c] = code[K[c]];
    vals[c] = c;
}

// Part 3: slot t takes canonical index order[t] -- position and velocity from the
// particle's previous slot S[order[t]] -- records both maps, and the bucket ranges
// in slots.
template <typename R>
__global__ void k_layout_reorder(uint32_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ K,
                                 const uint32_t* __restrict__ S, const V4<R>* __restrict__ pin,
                                 const V4<R>* __restrict__ vin, V4<R>* __restrict__ pout, V4<R>* __restrict__ vout,
                                 uint32_t* __restrict__ canon, uint32_t* __restrict__ stor, uint2* __restrict__ cells) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = order[t];
    const uint32_t s = S[c];
    pout[t] = pin[s];
    vout[t] = vin[s];
    canon[t] = c;
    stor[c] = t;
    const uint32_t k = K[c];
    if (t == 0 || K[order[t - 1]] != k) cells[k].x = t;
    if (t == n - 1 || K[order[t + 1]] != k) cells[k].y = t + 1;
}

// Every boundary range moved by its scene's shift[3 * scene ..].
template <typename R>
__global__ void k_translate_scenes(const BlockRange* __restrict__ blocks, const float* __restrict__ shift,
                                   V4<R>* __restrict__ p) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t i = rg.begin + threadIdx.x;
    if (i >= rg.end) return;
    const R dx = (R)shift[3 * rg.scene], dy = (R)shift[3 * rg.scene + 1], dz = (R)shift[3 * rg.scene + 2];
    V4<R> v = p[i];
    v.x = v.x + dx;
    v.y = v.y + dy;
    v.z = v.z + dz;
    p[i] = v;
}

}  // namespace

template <typename R>
int DataLaunch<R>::layout(const Launch& L, const Batch<R>& B, const LayoutArgs<R>& a) {
    const uint32_t n = B.n_total;
    if (!n || !B.fluid_blocks) return 0;
    hipLaunchKernelGGL((k_layout_codes<R>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B.scenes, B.blocks,
                       a.K, a.S, a.canon, a.pin, a.bits, a.code, a.srcc);
    hipLaunchKernelGGL(k_layout_keys, dim3(grid_for(n, L.block)), dim3(L.block), 0, L.stream, n, a.K, a.code,
                       a.okeys[0], a.ovals[0]);
    uint32_t* keys[2] = {a.okeys[0], a.okeys[1]};
    uint32_t* vals[2] = {a.ovals[0], a.ovals[1]};
    const int res = radix_sort(L, keys, vals, n, a.key_bits, a.digit_bits, a.scratch);
    if (res < 0) return res;
    hipLaunchKernelGGL((k_layout_reorder<R>), dim3(grid_for(n, L.block)), dim3(L.block), 0, L.stream, n,
                       vals[res], a.K, a.S, a.pin, a.vin, a.pout, a.vout, a.canon, a.stor, a.cells);
    return res;
}

template <typename R>
void DataLaunch<R>::gather_v4(const Launch& L, uint32_t n, const uint32_t* map, const V4<R>* src, V4<R>* dst) {
    if (!n) return;
    hipLaunchKernelGGL((k_gather_v4<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n, map);
}

void gather_words(const Launch& L, uint32_t n, uint32_t words, const uint32_t* map, const uint32_t* src,
                  uint32_t* dst) {
    if (!n) return;
    hipLaunchKernelGGL(k_gather_words, dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n, words, map);
}

void nonfinite_flags(const Launch& L, const LoopState* states, uint32_t count, uint32_t* out) {
    if (!count) return;
    hipLaunchKernelGGL(k_nonfinite_flags, dim3(grid_for(count, 256)), dim3(256), 0, L.stream, states, count, out);
}

void fill_iota(const Launch& L, uint32_t* p, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256)), dim3(256), 0, L.stream, p, n);
}

void cells_from_keys(const Launch& L, uint32_t n, const uint32_t* keys, uint2* cells) {
    if (!n) return;
    hipLaunchKernelGGL(k_cells_from_keys, dim3(grid_for(n, 256)), dim3(256), 0, L.stream, n, keys, cells);
}

template <typename R>
void DataLaunch<R>::translate_scenes(const Launch& L, const Batch<R>& B, const float* shift, V4<R>* bpos) {
    if (!B.boundary_blocks) return;
    hipLaunchKernelGGL((k_translate_scenes<R>), dim3(B.boundary_blocks), dim3(L.block), 0, L.stream,
                       B.blocks + B.fluid_blocks, shift, bpos);
}

template <typename R>
void DataLaunch<R>::fill_box(const Launch& L, const SphParams<R>* scenes, uint32_t scene_count, uint32_t max_cells,
                             const uint2* cells, uint2* grid, const V4<R>* far_pos) {
    if (!scene_count || !max_cells) return;
    const uint32_t blocks = std::min(grid_for(max_cells, 256), 1024u);
    hipLaunchKernelGGL((k_fill_box<R>), dim3(blocks, scene_count), dim3(256), 0, L.stream, scenes, cells, grid,
                       far_pos);
    if (far_pos)
        hipLaunchKernelGGL((k_mark_quiet<R>), dim3(blocks, scene_count), dim3(256), 0, L.stream, scenes, grid);
}

template <typename R>
void DataLaunch<R>::extents(const Launch& L, const Batch<R>& B, const V4<R>* pos, const V4<R>* bpos, int* ext) {
    const uint32_t grid = B.fluid_blocks + B.boundary_blocks;
    if (!grid) return;
    hipLaunchKernelGGL((k_extents<R>), dim3(grid), dim3(L.block), 0, L.stream, B.scenes, B.blocks, pos, bpos, ext);
}

template <typename R>
void DataLaunch<R>::translate(const Launch& L, V4<R>* p, uint32_t n, float dx, float dy, float dz) {
    if (!n) return;
    hipLaunchKernelGGL((k_translate<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, p, n, (R)dx, (R)dy, (R)dz);
}

template <typename R>
void DataLaunch<R>::hash(const Launch& L, const Batch<R>& B, bool boundary, const V4<R>* pos, uint32_t* keys,
                         uint32_t* vals, const uint32_t* canon) {
    const uint32_t grid = boundary ? B.boundary_blocks : B.fluid_blocks;
    if (!grid) return;
    hipLaunchKernelGGL((k_hash<R>), dim3(grid), dim3(L.block), 0, L.stream, B.scenes,
                       B.blocks + (boundary ? B.fluid_blocks : 0u), pos, keys, vals, canon);
}

template <typename R>
void DataLaunch<R>::stage_setup(const Launch& L, const DeviceArrays<R>& A, uint32_t bits, uint32_t key_bits,
                                uint32_t digit_bits, uint32_t* scratch, uint32_t* keys[2], uint32_t* vals[2]) {
    const Batch<R>& B = A.batch;
    if (!B.fluid_blocks) return;
    hipLaunchKernelGGL((k_stage_codes<R>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B.scenes, B.blocks,
                       A.pos, bits, keys[0], vals[0]);
    radix_sort(L, keys, vals, B.n_total, key_bits, digit_bits, scratch);
    hipLaunchKernelGGL((k_stage_setup<R>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B, A.cells, A.bcells,
                       A.pos, A.stage);
}

template <typename R>
int DataLaunch<R>::boundary_order(const Launch& L, const DeviceArrays<R>& A, uint32_t bits, uint32_t key_bits,
                                  uint32_t digit_bits, uint32_t* scratch, uint32_t* keys[2], uint32_t* vals[2]) {
    const Batch<R>& B = A.batch;
    if (!B.boundary_blocks) return 0;
    hipLaunchKernelGGL((k_stage_codes<R>), dim3(B.boundary_blocks), dim3(L.block), 0, L.stream, B.scenes,
                       B.blocks + B.fluid_blocks, A.bpos, bits, keys[0], vals[0]);
    return radix_sort(L, keys, vals, B.nb_total, key_bits, digit_bits, scratch);
}

template <typename R>
void DataLaunch<R>::warm_apply(const Launch& L, const DeviceArrays<R>& A) {
    if (!A.batch.fluid_blocks) return;
    hipLaunchKernelGGL((k_warm_apply<R>), dim3(A.batch.fluid_blocks), dim3(L.block), 0, L.stream, A.batch, A.xr,
                       A.vel, A.p, A.xf);
}

template <typename R>
void DataLaunch<R>::reorder(const Launch& L, uint32_t n, const uint32_t* keys, const uint32_t* src,
                            const V4<R>* pin, const V4<R>* vin, V4<R>* pout, V4<R>* vout, uint2* cells) {
    if (!n) return;
    hipLaunchKernelGGL((k_reorder<R>), dim3(grid_for(n, L.block)), dim3(L.block), 0, L.stream, n, keys, src,
                       pin, vin, pout, vout, cells);
}

template <typename R>
void DataLaunch<R>::unpack_vec3(const Launch& L, const float* src, V4<R>* dst, uint32_t n, const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_unpack_vec3<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n, map);
}

template <typename R>
void DataLaunch<R>::pack_vec3(const Launch& L, const V4<R>* src, float* dst, uint32_t n, const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_vec3<R, float>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n, map);
}

template <typename R>
void DataLaunch<R>::pack_vec3_f64(const Launch& L, const V4<R>* src, double* dst, uint32_t n,
                                  const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_vec3<R, double>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n, map);
}

template <typename R>
void DataLaunch<R>::unpack_scalar(const Launch& L, const float* src, R* dst, uint32_t n, uint32_t stride,
                                  uint32_t offset, const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_unpack_scalar<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n,
                       stride, offset, map);
}

template <typename R>
void DataLaunch<R>::pack_scalar(const Launch& L, const R* src, float* dst, uint32_t n, uint32_t stride,
                                uint32_t offset, const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_scalar<R, float>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n,
                       stride, offset, map);
}

template <typename R>
void DataLaunch<R>::pack_scalar_f64(const Launch& L, const R* src, double* dst, uint32_t n, uint32_t stride,
                                    uint32_t offset, const uint32_t* map) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_scalar<R, double>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst,
                       n, stride, offset, map);
}

template struct DataLaunch<float>;
template struct DataLaunch<double>;
template struct SphLaunch<float, false>;
template struct SphLaunch<double, false>;

}  // namespace pcisph
