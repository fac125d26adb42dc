// sph_ieee.hip - IEEE instantiations of the step kernels (built with
// -ffp-contract=off): FP32 in the reference's order of operations, and FP64.
// Also the precision-only data kernels (hash, reorder, layout conversion).
#include <algorithm>

#include "sph_launch_impl.h"

namespace pcisph {

namespace {

template <typename R>
__global__ void k_unpack_vec3(const float* __restrict__ src, V4<R>* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V4<R> v;
    v.x = (R)src[3 * (size_t)i];
    v.y = (R)src[3 * (size_t)i + 1];
    v.z = (R)src[3 * (size_t)i + 2];
    v.w = 0;
    dst[i] = v;
}

template <typename R, typename O>
__global__ void k_pack_vec3(const V4<R>* __restrict__ src, O* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V4<R> v = src[i];
    dst[3 * (size_t)i] = (O)v.x;
    dst[3 * (size_t)i + 1] = (O)v.y;
    dst[3 * (size_t)i + 2] = (O)v.z;
}

template <typename R>
__global__ void k_unpack_scalar(const float* __restrict__ src, R* __restrict__ dst, uint32_t n,
                                uint32_t stride, uint32_t offset) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[(size_t)i * stride + offset] = (R)src[i];
}

template <typename R, typename O>
__global__ void k_pack_scalar(const R* __restrict__ src, O* __restrict__ dst, uint32_t n, uint32_t stride,
                              uint32_t offset) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = (O)src[(size_t)i * stride + offset];
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

// The low `bits` bits of v spread to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint32_t spread3(uint32_t v, uint32_t bits) {
    uint32_t out = 0;
    for (uint32_t k = 0; k < bits; ++k) out |= ((v >> k) & 1u) << (3 * k);
    return out;
}

// Work-order keys of the sorted fluid particles: scene index above the Morton code
// of the particle's cell (coordinates taken modulo 2^bits), value = the slot.
template <typename R>
__global__ void k_morton(const SphParams<R>* __restrict__ scenes, const BlockRange* __restrict__ blocks,
                         const V4<R>* __restrict__ pos, uint32_t bits, uint32_t* __restrict__ keys,
                         uint32_t* __restrict__ vals) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t t = rg.begin + threadIdx.x;
    if (t >= rg.end) return;
    const int3 c = cell_of4<R>(pos[t], scenes[rg.scene].cell);
    const uint32_t code = spread3((uint32_t)c.x, bits) | (spread3((uint32_t)c.y, bits) << 1) |
                          (spread3((uint32_t)c.z, bits) << 2);
    keys[t] = (rg.scene << (3 * bits)) | code;
    vals[t] = t;
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
void DataLaunch<R>::morton(const Launch& L, const Batch<R>& B, const V4<R>* pos, uint32_t bits, uint32_t* keys,
                           uint32_t* vals) {
    if (!B.fluid_blocks) return;
    hipLaunchKernelGGL((k_morton<R>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B.scenes, B.blocks, pos, bits,
                       keys, vals);
}

template <typename R>
void DataLaunch<R>::translate_scenes(const Launch& L, const Batch<R>& B, const float* shift, V4<R>* bpos) {
    if (!B.boundary_blocks) return;
    hipLaunchKernelGGL((k_translate_scenes<R>), dim3(B.boundary_blocks), dim3(L.block), 0, L.stream,
                       B.blocks + B.fluid_blocks, shift, bpos);
}

template <typename R>
void DataLaunch<R>::fill_box(const Launch& L, const SphParams<R>* scenes, uint32_t scene_count, uint32_t max_cells,
                             const uint2* cells, uint2* grid) {
    if (!scene_count || !max_cells) return;
    const uint32_t blocks = std::min(grid_for(max_cells, 256), 1024u);
    hipLaunchKernelGGL((k_fill_box<R>), dim3(blocks, scene_count), dim3(256), 0, L.stream, scenes, cells, grid);
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
                         uint32_t* vals) {
    const uint32_t grid = boundary ? B.boundary_blocks : B.fluid_blocks;
    if (!grid) return;
    hipLaunchKernelGGL((k_hash<R>), dim3(grid), dim3(L.block), 0, L.stream, B.scenes,
                       B.blocks + (boundary ? B.fluid_blocks : 0u), pos, keys, vals);
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
void DataLaunch<R>::unpack_vec3(const Launch& L, const float* src, V4<R>* dst, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL((k_unpack_vec3<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n);
}

template <typename R>
void DataLaunch<R>::pack_vec3(const Launch& L, const V4<R>* src, float* dst, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_vec3<R, float>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n);
}

template <typename R>
void DataLaunch<R>::pack_vec3_f64(const Launch& L, const V4<R>* src, double* dst, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_vec3<R, double>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n);
}

template <typename R>
void DataLaunch<R>::unpack_scalar(const Launch& L, const float* src, R* dst, uint32_t n, uint32_t stride,
                                  uint32_t offset) {
    if (!n) return;
    hipLaunchKernelGGL((k_unpack_scalar<R>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n,
                       stride, offset);
}

template <typename R>
void DataLaunch<R>::pack_scalar(const Launch& L, const R* src, float* dst, uint32_t n, uint32_t stride,
                                uint32_t offset) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_scalar<R, float>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst, n,
                       stride, offset);
}

template <typename R>
void DataLaunch<R>::pack_scalar_f64(const Launch& L, const R* src, double* dst, uint32_t n, uint32_t stride,
                                    uint32_t offset) {
    if (!n) return;
    hipLaunchKernelGGL((k_pack_scalar<R, double>), dim3(grid_for(n, 256)), dim3(256), 0, L.stream, src, dst,
                       n, stride, offset);
}

template struct DataLaunch<float>;
template struct DataLaunch<double>;
template struct SphLaunch<float, false>;
template struct SphLaunch<double, false>;

}  // namespace pcisph
