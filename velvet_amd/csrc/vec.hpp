// Small float3 helper for the kernels. Operations are spelled out component by component so
// the order of evaluation is the one written.
#pragma once
#include <hip/hip_runtime.h>

namespace vtx {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 xyz(const float4& a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ float4 f4(const V3& a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
__device__ __forceinline__ V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
// glm::normalize: v * inversesqrt(dot(v, v)); a zero vector gives NaN as in glm.
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }

// Row-major 3x4 affine transform of a point / of a vector by the upper 3x3.
__device__ __forceinline__ V3 xform_point(const float* m, V3 p) {
    return V3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
              m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
__device__ __forceinline__ V3 xform_vec(const float* m, V3 p) {
    return V3{m[0] * p.x + m[1] * p.y + m[2] * p.z,
              m[4] * p.x + m[5] * p.y + m[6] * p.z,
              m[8] * p.x + m[9] * p.y + m[10] * p.z};
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (workgroup b runs on XCD
// b % 8, each XCD with its own L2). Consecutive tiles are spatial neighbours that share halo,
// cell-table and gather data, so hand each XCD a contiguous run of tiles: block b of nb
// processes tile xcd_tile(b, nb), a bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + i;
}

// The tile the current workgroup of a tile-indexed launch processes (remap = 0: its own index).
__device__ __forceinline__ uint32_t tile_of(uint32_t remap) {
    return remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
}

// A launch whose TILE-particle tiles are each covered by TILE / BS workgroups of BS threads:
// the tile of the current workgroup and the current thread's lane within it (the sub-blocks of
// one tile stay consecutive, so the XCD remap keeps them on one XCD).
template <int TILE, int BS>
__device__ __forceinline__ void tile_lane(uint32_t remap, uint32_t& tile, uint32_t& lane) {
    constexpr uint32_t S = TILE / BS;
    const uint32_t b = tile_of(remap);
    tile = b / S;
    lane = (b % S) * BS + threadIdx.x;
}

}  // namespace vtx
