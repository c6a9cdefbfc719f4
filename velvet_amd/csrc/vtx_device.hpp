// Device-side data layout shared by the HIP kernels and the host engine.
//
// Particles of all segments are packed back to back in float4 arrays (xyz + one payload lane):
//   X  positions          w = inverse mass
//   V  velocities         w unused
//   P  predicted (x2)     w = inverse mass, ping-ponged between Jacobi passes
//   X0 initial positions  collision filter snapshot (Velvet/SpatialHashGPU.hpp:32-39)
//   N  normals
// Work is cut into particle tiles of VTX_TILE consecutive particles of one segment, one tile
// per workgroup, so every per-segment value a kernel needs is wave-uniform (scalar loads).
// Incident constraints are stored per tile in sliced-ELL form: entry k of lane l sits at
// [tile offset + k * VTX_TILE + l], so the lanes of a wave read consecutive words.
#pragma once
#include <stdint.h>

#define VTX_TILE 256
#define VTX_SORT_ITEMS 8                       // keys per thread in a sort tile
#define VTX_SORT_TILE (VTX_TILE * VTX_SORT_ITEMS)
#define VTX_SCAN_BLOCK 1024                     // elements per block of the device scan
#define VTX_INVALID 0xFFFFFFFFu
#define VTX_EPSILON 1e-6f                       // Velvet/Common.cuh:21
#define VTX_MAX_NEIGHBORS 64

#define VTX_SEG_SELFCOL 1u
#define VTX_SEG_GRID 2u          // Jacobi pass runs on 2-D grid tiles (stencil layout)
#define VTX_SEG_BEND_REST0 4u    // every bending rest angle is 0 (no rest-angle loads)
// Per bending slot type (X, R, D): every geometrically possible slot is present (presence is
// computed from the cell position) or none is; otherwise presence is read per cell.
#define VTX_SEG_FULL_X 8u
#define VTX_SEG_FULL_R 16u
#define VTX_SEG_FULL_D 32u
#define VTX_SEG_NONE_X 64u
#define VTX_SEG_NONE_R 128u
#define VTX_SEG_NONE_D 256u
// Attachments are "every particle to every slot, in slot order" (Velvet/VtClothObjectGPU.hpp:134-148):
// only the distances are stored, slot-major: lra[lbase + slot * n + particle].
#define VTX_SEG_LRA_DENSE 512u
// The segment's spatial hash is built by one workgroup in LDS (k_hash_small): self-collision,
// at most VTX_HASH_LDS_N particles and a table of at most 2^16 cells. The device-wide hash
// kernels skip it.
#define VTX_SEG_HASH_LDS 1024u
#define VTX_HASH_LDS_N 16384
// The segment hashes its cell coordinates after adding VTX_HASH_SHIFT to each (32-bit
// wrap-around): a tuned hash grid. HashCoords gives cells (x, y, z) and (-x, y, -z) one bucket
// (with odd x and z the two products' two's-complement negations cancel in the XOR, and abs()
// folds the remaining cases), so a cloth centred on the origin shares every bucket with its
// mirror image; shifted coordinates have no such partner nearby.
#define VTX_SEG_HASH_SHIFT 2048u
#define VTX_HASH_SHIFT 1000003u
// The segment's neighbour table holds 16-bit entries (0xffff the sentinel): self-collision
// segments of at most VTX_NBR16_N particles, whose indices and sentinel fit in 16 bits. Chunks
// of four entries are then 8 bytes: half the table traffic of the search's stores and of the
// contact pass's loads.
#define VTX_SEG_NBR16 4096u
#define VTX_NBR16_N 65535u
// A particle's attachment terms are summed in four parts by slot index -- slots
// [q*Qs, (q+1)*Qs) with Qs = ceil(scount / 4) -- each part from zero in slot order, and the parts
// are added to the particle's delta in order q = 0..3. One lane can run all four parts, or four
// lanes one part each (stencil segments with at least VTX_LRA_SPLIT_MIN dense slots), with the
// same result.
#define VTX_LRA_PARTS 4
// Most streams a frame's segment groups run on (vtx_tuning.streams).
#define VTX_MAX_STREAMS 4
#define VTX_LRA_SPLIT_MIN 32u

// 2-D grid tiles of the stencil path: VTX_GT x VTX_GT particles, halo of 2 rows/cols.
#define VTX_GT 16
#define VTX_GH 2
#define VTX_GP (VTX_GT + 2 * VTX_GH)      // 20: staged positions per side
// Bending slot types of a grid cell (a, b), each with its canonical vertex order
// (p0, p1, p2, p3) as (row, col) offsets from the cell:
//   X: diagonal (a+1,b)-(a,b+1)    (0,0) (1,1) (0,1) (1,0)   -- Velvet/VtClothObjectGPU.hpp:121-130
//   R: row edge (a,b)-(a,b+1)      (-1,1) (1,0) (0,0) (0,1)
//   D: column edge (a,b)-(a+1,b)   (1,-1) (0,1) (1,0) (0,0)
#define VTX_BX 1u
#define VTX_BR 2u
#define VTX_BD 4u

namespace vtx {

struct alignas(16) SegDev {
    uint32_t pbase, n, tbase, M;           // particle base / count, hash-table base / size
    uint32_t nbase_lo, nbase_hi, sbase, scount;   // neighbour-table base, attach-slot base/count
    uint32_t cbase, ccount, flags, max_nbr;       // collider base/count, flags, neighbour cap
    float diameter, cell, cell2, diam2;
    float max_speed, damping, omega, lrs;
    float bend_compliance, margin, friction, stiffness;
    float gx, gy, gz, pad0;
    uint32_t mod_lo, mod_hi;               // ceil(2^64 / M): exact remainder by multiplication
    uint32_t rows, cols;                   // grid segments: vertex grid shape
    uint32_t bbase, trbase, lbase, pad2;   // bend-record / triangle-record / dense-LRA bases
};

struct alignas(16) TileDev {
    uint32_t seg, begin, count, pad0;      // begin = global index of the tile's first particle
    uint32_t st_off, st_k, at_off, at_k;   // stretch / attach ELL offsets (entries) and widths
    uint32_t bd_off, bd_k, tr_off, tr_k;   // bend / triangle ELL offsets and widths
};

struct alignas(16) SortTileDev {
    uint32_t seg, begin, count, hbase;     // hbase: histogram base of the segment
    uint32_t ntiles, tidx, bias, pad1;     // sort tiles of the segment, index of this one;
                                           // bias = pbase - (particles of earlier segments), mod 2^32
};

struct alignas(16) GridTileDev {
    uint32_t seg, a0, b0, pad0;            // segment, first row / column of the tile
    uint32_t at_off, at_k, pad1, pad2;     // attach ELL of the tile's lanes
};

// Collider with its transforms as 3x4 row-major affine matrices:
//   inv = inverse(cur), rel = last * inverse(cur) (Velvet/VtClothSolverGPU.cuh:91-96),
//   cur3 = upper-left 3x3 of cur (used by the cube's correction, cuh:86).
struct alignas(16) ColliderDev {
    int32_t type;
    float dt;
    float pad0, pad1;
    float pos[4];
    float scale[4];
    float cur3[12];
    float inv[12];
    float rel[12];
};

// Pointers and sizes every kernel may need; passed by value (kernel-argument segment).
struct DevState {
    const SegDev* segs;
    const TileDev* tiles;
    const SortTileDev* stiles;
    float4* X;
    float4* V;
    float4* P[2];
    const float4* X0;
    float4* Nrm;
    // ELL entries and records hold segment-local indices; kernels add the segment's bases
    const uint2* st;          // stretch ELL: (other particle, rest bits)
    const uint2* at;          // attach ELL: (slot, distance bits)
    const uint32_t* bd;       // bend ELL: bend id | role << 30
    const uint4* brec;        // bend records: 4 particle indices
    const float* brest;       // bend rest angles
    const uint32_t* tr;       // triangle ELL: triangle id
    const uint4* trec;        // triangle records: 3 particle indices
    const float4* slots;      // attach slot positions
    const float* lra;         // dense long-range attachment distances
    const ColliderDev* cols;
    uint32_t* keys[2];        // sort keys (hash) ping-pong
    uint32_t* vals[2];        // sort values (local particle index) ping-pong
    uint32_t* hist;           // per (segment, digit, sort tile) counts, scanned in place
    uint32_t* part;           // scan partials
    uint2* cells;             // per hash cell: (start, end) in sorted order, start ~0u = empty
    float4* sPX;              // per sorted position, 32 B: the predicted position with w = local
                              // particle index, then the initial position (one cache line holds both)
    uint32_t* nbr;            // neighbour tables per segment, chunks of four entries (16 or 8 bytes)
    const TileDev* itiles;    // particle tiles of segments on the general (ELL) Jacobi path
    const GridTileDev* gtiles;
    const GridTileDev* ltiles;   // grid tiles whose dense attachments run on four lanes per particle
    const uint32_t* hsegs;       // segments whose hash is built in LDS (VTX_SEG_HASH_LDS)
    const float4* rest4;      // grid: stretch rest lengths owned by a vertex (R, D, X-diag, anti-diag)
    const uint32_t* gflags;   // grid: bending slots present at a cell (VTX_BX | VTX_BR | VTX_BD)
    const float4* brest4;     // grid: bending rest angles of a cell (X, R, D)
    uint32_t ntiles, nstiles, nhist, ncells;
    uint32_t hbase0, nltiles, nhsegs, pad3;  // first histogram entry of this table set
    uint32_t nparticles, nitiles, ngtiles, remap;   // remap: workgroups -> tiles through xcd_tile
};

}  // namespace vtx
