// Host-side launchers of the HIP kernels (one per pass); all are asynchronous on the stream.
#pragma once
#include <hip/hip_runtime.h>

#include "vtx_device.hpp"

namespace vtx {

enum { EPILOGUE_NONE = 0, EPILOGUE_FINALIZE = 1, EPILOGUE_FINALIZE_PREDICT = 2 };

hipError_t launch_frame_begin(const DevState& d, float frame_dt, float h, hipStream_t s);
// block: threads per workgroup of the gather kernels (64, 128 or 256; same results)
hipError_t launch_collide(const DevState& d, int in, float h, int block, hipStream_t s);
// lra_batch: dense attachment distances loaded per batch (4, 16 or 64; same results)
hipError_t launch_iterate(const DevState& d, int in, float h, int epilogue, bool constraints, int lra_batch,
                          hipStream_t s);
hipError_t launch_normals(const DevState& d, hipStream_t s);
hipError_t launch_check_finite(const DevState& d, uint32_t* flags, uint32_t nseg, hipStream_t s);

hipError_t launch_hash_keys(const DevState& d, int in, hipStream_t s);
hipError_t launch_sort_pass(const DevState& d, int src, int shift, hipStream_t s);
hipError_t launch_scan(uint32_t* a, uint32_t n, uint32_t* part, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* a, uint32_t n, uint32_t v, hipStream_t s);
hipError_t launch_clear_cells(const DevState& d, hipStream_t s);
hipError_t launch_hash_small(const DevState& d, int in, int dst, hipStream_t s);
hipError_t launch_cell_bounds(const DevState& d, int in, int src, hipStream_t s);
// Thread -> particle order of the neighbour search (same tables for every order).
enum { NBR_PARTICLE = 0, NBR_SORTED = 1, NBR_PATCH = 2 };
hipError_t launch_neighbors(const DevState& d, int in, int order, int block, hipStream_t s);

// HashObjects of the kernel-launch tier (Velvet/SpatialHashGPU.cuh:7-25) on raw device arrays.
struct HashObjectsArgs {
    uint32_t* particle_hash;         // n: sorted hashes on return
    uint32_t* particle_index;        // n: object indices in sorted order on return
    uint32_t* cell_start;            // M
    uint32_t* cell_end;              // M
    uint32_t* neighbors;             // max_num_neighbors x n, column-major
    const float* positions;          // n xyz
    const float* original_positions; // n xyz
    uint32_t num_objects, max_num_neighbors, table_size;
    float cell_spacing, cell_spacing2, particle_diameter2;
};
// Workspace bytes hash_objects needs for n objects and M cells (256-B aligned parts).
size_t hash_workspace_bytes(uint32_t n, uint32_t M);
hipError_t hash_objects(const HashObjectsArgs& a, void* workspace, hipStream_t s);
// hash_objects with its workspace in a fixed per-device array of VTX_HASH_SCRATCH_BYTES (no
// allocation); hipErrorInvalidValue when hash_workspace_bytes exceeds it.
#define VTX_HASH_SCRATCH_BYTES (64u << 20)
hipError_t hash_objects_noalloc(const HashObjectsArgs& a, hipStream_t s);

}  // namespace vtx
