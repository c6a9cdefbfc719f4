// Spatial-hash build for gfx950: cell hash, stable LSD radix sort, cell start/end and the
// cached neighbour table (Velvet/SpatialHashGPU.cu:13-196).
//
// All segments are sorted in one pass sequence. Keys are the per-segment hash (< M of that
// segment), values the local particle index. Sort tiles never straddle segments and the
// digit histograms are laid out (segment, digit, tile), so one exclusive scan over the whole
// histogram gives every (segment, digit, tile) its global output offset: the segments come
// out sorted independently, in place, without sorting on segment bits. Ranking inside a tile
// is stable (each wave a contiguous chunk, in wave order; inside a chunk 64 keys at a time by
// wave ballots, in order), so the result equals a stable sort by hash within each segment -- the
// contract of the reference's
// cub::DeviceRadixSort::SortPairs call (Velvet/SpatialHashGPU.cu:133-157).
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.hpp"
#include "vec.hpp"
#include "vtx_device.hpp"

namespace vtx {


// HashCoords (Velvet/SpatialHashGPU.cu:18-22): 32-bit wrap-around products, then
// abs(h % M). |h % M| == |h| mod M for C's truncating remainder, and the unsigned remainder is
// taken by multiplication with c = ceil(2^64 / M): ((c * a) mod 2^64) * M >> 64 == a mod M for
// all 32-bit a and M (Lemire, Kaser, Kurz, "Faster remainder by direct computation", 2019).
__device__ __forceinline__ uint32_t hash_coords(int32_t x, int32_t y, int32_t z, uint32_t M, uint64_t c) {
    const int32_t h = (int32_t)(((uint32_t)x * 92837111u) ^ ((uint32_t)y * 689287499u) ^
                                ((uint32_t)z * 283923481u));
    const uint32_t a = h < 0 ? (uint32_t)(-(int64_t)h) : (uint32_t)h;
    const uint64_t low = c * (uint64_t)a;
    return (uint32_t)__umul64hi(low, (uint64_t)M);
}

__device__ __forceinline__ uint64_t seg_mod(const SegDev& sg) { return (uint64_t)sg.mod_hi << 32 | sg.mod_lo; }

// ComputeIntCoord (Velvet/SpatialHashGPU.cu:13-16): floor of a correctly rounded binary32
// quotient (the library is built with -fhip-fp32-correctly-rounded-divide-sqrt).
__device__ __forceinline__ int32_t int_coord(float v, float cell) {
    return (int32_t)floorf(__fdiv_rn(v, cell));
}

// Cell coordinate as hashed: ComputeIntCoord plus the segment's grid shift (0 on the reference
// grid, VTX_HASH_SHIFT on a tuned one), with 32-bit wrap-around.
__device__ __forceinline__ int32_t hash_coord(float v, float cell, uint32_t shift) {
    return (int32_t)((uint32_t)int_coord(v, cell) + shift);
}

__device__ __forceinline__ uint32_t seg_shift(const SegDev& sg) {
    return (sg.flags & VTX_SEG_HASH_SHIFT) ? VTX_HASH_SHIFT : 0u;
}

// ComputeParticleHash_Kernel (Velvet/SpatialHashGPU.cu:34-42) over predicted positions.
__global__ __launch_bounds__(VTX_TILE) void k_hash_keys(DevState d, int in) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + threadIdx.x;
    if (sg.flags & VTX_SEG_HASH_LDS) return;
    d.vals[0][g] = g - sg.pbase;
    if (!(sg.flags & VTX_SEG_SELFCOL)) {
        d.keys[0][g] = 0;      // segment without self-collision: sorted as one cell, unused
        return;
    }
    const float4 p = d.P[in][g];
    const uint32_t sh = seg_shift(sg);
    d.keys[0][g] = hash_coords(hash_coord(p.x, sg.cell, sh), hash_coord(p.y, sg.cell, sh),
                               hash_coord(p.z, sg.cell, sh), sg.M, seg_mod(sg));
}

// Digit histogram of one sort tile (count keys from `keys`): the count of digit g goes to
// out[g * stride].
__device__ __forceinline__ void tile_hist(const uint32_t* __restrict__ keys, uint32_t count, int shift,
                                          uint32_t* __restrict__ out, uint32_t stride) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t r = 0; r < VTX_SORT_ITEMS; r++) {
        const uint32_t i = r * VTX_TILE + threadIdx.x;
        if (i < count) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    out[threadIdx.x * stride] = cnt[threadIdx.x];
}

// Per-tile digit histogram into hist[hbase + digit * ntiles + tidx].
__global__ __launch_bounds__(VTX_TILE) void k_sort_hist(DevState d, int src, int shift) {
    const SortTileDev t = d.stiles[blockIdx.x];
    tile_hist(d.keys[src] + t.begin, t.count, shift, d.hist + t.hbase + t.tidx, t.ntiles);
}

// Wave-level exclusive scan helpers (wave64).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Block (256 threads) exclusive scan; returns the exclusive prefix, *total the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* sh) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < VTX_TILE / 64; w++) {
        const uint32_t s = sh[w];
        off += (w < wave) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// Device-wide exclusive scan of hist[0..n) in two launches: block sums, then each block adds
// the sums of the blocks before it (a block reduction over at most a few hundred sums) and
// scans its own elements.
__global__ __launch_bounds__(VTX_TILE) void k_scan_reduce(uint32_t* __restrict__ a, uint32_t n,
                                                          uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[VTX_TILE / 64];
    const uint32_t base = blockIdx.x * VTX_SCAN_BLOCK;
    uint32_t s = 0;
    for (uint32_t r = 0; r < VTX_SCAN_BLOCK / VTX_TILE; r++) {
        const uint32_t i = base + r * VTX_TILE + threadIdx.x;
        if (i < n) s += a[i];
    }
    uint32_t total;
    block_excl_scan(s, &total, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ __launch_bounds__(VTX_TILE) void k_scan_apply(uint32_t* __restrict__ a, uint32_t n,
                                                         const uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[VTX_TILE / 64];
    const uint32_t base = blockIdx.x * VTX_SCAN_BLOCK;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < blockIdx.x; b0 += VTX_TILE) {   // sums of the earlier blocks
        const uint32_t b = b0 + threadIdx.x;
        uint32_t total;
        block_excl_scan(b < blockIdx.x ? part[b] : 0u, &total, sh);
        carry += total;
    }
    for (uint32_t r = 0; r < VTX_SCAN_BLOCK / VTX_TILE; r++) {
        const uint32_t i = base + r * VTX_TILE + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_excl_scan(v, &total, sh);
        if (i < n) a[i] = carry + ex;
        carry += total;
    }
}

// Stable scatter of one sort tile (count keys and values from kin / vin). Wave w owns the tile's
// contiguous chunk w of count / 4 keys: it counts its chunk's digits with LDS atomics (order does
// not matter), each digit's start then runs through the waves in order from the tile's base
// (digit g of the tile starts at output position scanned[g * stride] + bias), and each wave
// scatters its chunk in order, 64 keys at a time ranked by ballots over the 8 digit bits, moving
// its own per-digit offsets on: chunks in wave order and keys in chunk order keep the sort
// stable, with two workgroup barriers per tile.
__device__ __forceinline__ void tile_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                             uint32_t count, int shift, const uint32_t* __restrict__ scanned,
                                             uint32_t stride, uint32_t bias, uint32_t* __restrict__ kout,
                                             uint32_t* __restrict__ vout) {
    constexpr uint32_t NW = VTX_TILE / 64;
    __shared__ uint32_t wcnt[NW][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t chunk = (count + NW - 1) / NW;
    const uint32_t c0 = min(wave * chunk, count), c1 = min(c0 + chunk, count);
    uint32_t* __restrict__ hw = wcnt[wave];
    for (uint32_t k = lane; k < 256; k += 64) hw[k] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = c0 + lane; i < c1; i += 64) atomicAdd(&hw[(kin[i] >> shift) & 255u], 1u);
    __syncthreads();
    {   // thread tid = digit: the waves' starts in wave order
        uint32_t run = scanned[tid * stride] + bias;
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) {
            const uint32_t c = wcnt[w][tid];
            wcnt[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t r0 = c0; r0 < c1; r0 += 64) {
        const uint32_t i = r0 + lane;
        const bool valid = i < c1;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t dg = (key >> shift) & 255u;
        uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t m = (uint32_t)((int32_t)(dg << (31 - b)) >> 31);   // bit b as 0 / ~0
            const uint64_t bal = __builtin_amdgcn_ballot_w64(m != 0u);
            peers &= ~(bal ^ ((uint64_t)m << 32 | m));                        // lanes whose bit b is this lane's
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t at = hw[dg];
        if (valid) {
            kout[at + rank] = key;
            vout[at + rank] = val;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) hw[dg] = at + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ __launch_bounds__(VTX_TILE) void k_sort_scatter(DevState d, int src, int shift) {
    const SortTileDev t = d.stiles[blockIdx.x];
    tile_scatter(d.keys[src] + t.begin, d.vals[src] + t.begin, t.count, shift, d.hist + t.hbase + t.tidx, t.ntiles,
                 t.bias, d.keys[src ^ 1], d.vals[src ^ 1]);
}

__global__ void k_fill_u32(uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// Mark every cell of every self-colliding segment empty: lane i of a segment clears cells
// i, i + n, i + 2n, ... < M (the reference's memset, Velvet/SpatialHashGPU.cu:185).
__global__ __launch_bounds__(VTX_TILE) void k_clear_cells(DevState d) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    if (!(sg.flags & VTX_SEG_SELFCOL) || (sg.flags & VTX_SEG_HASH_LDS)) return;
    uint2* __restrict__ cells = d.cells + sg.tbase;
    for (uint32_t i = t.begin - sg.pbase + threadIdx.x; i < sg.M; i += sg.n) cells[i] = make_uint2(VTX_INVALID, VTX_INVALID);
}

// FindCellStart_Kernel (Velvet/SpatialHashGPU.cu:44-77) per segment with local sorted
// positions (cells were filled with ~0u: empty), plus the inverse permutation (particle ->
// sorted position) in the sort's free key buffer, from which k_sorted_records writes the
// sorted records in particle order.
__global__ __launch_bounds__(VTX_TILE) void k_cell_bounds(DevState d, int in, int src) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    if (!(sg.flags & VTX_SEG_SELFCOL) || (sg.flags & VTX_SEG_HASH_LDS)) return;
    const uint32_t g = t.begin + threadIdx.x;
    const uint32_t local = g - sg.pbase;
    const uint32_t* __restrict__ sk = d.keys[src];
    const uint32_t key = sk[g];
    uint2* __restrict__ cells = d.cells + sg.tbase;
    if (local == 0) {
        cells[key].x = 0;
    } else {
        const uint32_t prev = sk[g - 1];
        if (key != prev) {
            cells[key].x = local;
            cells[prev].y = local;
        }
    }
    if (local == sg.n - 1) cells[key].y = sg.n;
    d.keys[src ^ 1][sg.pbase + d.vals[src][g]] = local;
}

// The sorted copies for the neighbour search, written from particle order: coalesced reads of the
// predicted and initial positions, one scattered 32-B record per particle (a gather in sorted
// order reads two scattered 16-B halves instead). w of the position copy carries the particle's
// local index (the inverse mass is not needed there). inv = the inverse permutation k_cell_bounds
// left in the sort's free key buffer.
__global__ __launch_bounds__(VTX_TILE) void k_sorted_records(DevState d, int in, int src) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    if (!(sg.flags & VTX_SEG_SELFCOL) || (sg.flags & VTX_SEG_HASH_LDS)) return;
    const uint32_t g = t.begin + threadIdx.x;
    const size_t o = 2 * (size_t)(sg.pbase + d.keys[src ^ 1][g]);
    const float4 q = d.P[in][g];
    d.sPX[o] = make_float4(q.x, q.y, q.z, __uint_as_float(g - sg.pbase));
    d.sPX[o + 1] = d.X0[g];
}

// The whole spatial-hash build of one small segment in one workgroup of VTX_HASH_LDS_BT threads:
// keys of the predicted positions (k_hash_keys), a stable LSD radix sort in LDS on key << 16 |
// local index (8 bits per pass; each wave counts its own contiguous chunk with LDS atomics and
// then scatters it, 64 keys at a time ranked by wave ballots over the pass's digit bits, so only
// three workgroup barriers per pass), the cell table (k_clear_cells + k_cell_bounds) and the
// sorted records. The sort runs on one CU and is bound by its VALU issue (a wave64 instruction
// takes four cycles of a 16-lane SIMD), hence the atomic counts: the ballot match is needed only
// for the ranks of the stable scatter. The records are written from particle order through the
// inverse permutation (kept in the free sort buffer): the positions and initial positions are
// read coalesced and each 32-B record is one scattered write, where the sorted-order form gathers
// two scattered 16-B reads per particle (batch of 512 cloths: 99.6 -> 79.2 us per build).
// Same keys, same stable order, so the same table and records as the device-wide path, in one
// launch for every small segment instead of a dozen over all of them.
// NMAX: particles per segment at most (16384: 144 KB of LDS, one workgroup per CU). Global
// loads are issued eight elements at a time so that their latencies overlap.
#define VTX_HASH_LDS_BT 1024
template <uint32_t NMAX>
__global__ __launch_bounds__(VTX_HASH_LDS_BT) void k_hash_small(DevState d, const uint32_t* __restrict__ list, int in,
                                                                int dst) {
    constexpr uint32_t NW = VTX_HASH_LDS_BT / 64;
    __shared__ uint32_t buf[2][NMAX];                // key << 16 | local index
    __shared__ uint32_t wcnt[NW][256];               // per wave: digit counts, then running offsets
    __shared__ uint32_t wsum[4];
    const SegDev sg = d.segs[list[blockIdx.x]];
    const uint32_t n = sg.n, M = sg.M, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint2* __restrict__ cells = d.cells + sg.tbase;
    for (uint32_t i = tid; i < M; i += VTX_HASH_LDS_BT) cells[i] = make_uint2(VTX_INVALID, VTX_INVALID);
    const uint64_t mc = seg_mod(sg);
    const uint32_t sh = seg_shift(sg);
    const float4* __restrict__ P = d.P[in] + sg.pbase;
    // the predicted positions stay in registers (NMAX / STEP rounds of U per thread) until the
    // sorted records are written: the workgroup runs alone on its CU, so the 48 extra VGPRs cost
    // no occupancy, and the records then read only the initial positions (75.0 against 78.7 us for
    // the batch of 512 cloths)
    constexpr uint32_t U = 8, STEP = U * VTX_HASH_LDS_BT, NS = (NMAX + STEP - 1) / STEP;
    float kx[NS][U], ky[NS][U], kz[NS][U];
#pragma unroll
    for (uint32_t s0 = 0; s0 < NS; s0++) {
        const uint32_t i0 = tid + s0 * STEP;
        float4 p[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * VTX_HASH_LDS_BT;
            p[u] = i < n ? P[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * VTX_HASH_LDS_BT;
            kx[s0][u] = p[u].x;
            ky[s0][u] = p[u].y;
            kz[s0][u] = p[u].z;
            if (i < n)
                buf[0][i] = hash_coords(hash_coord(p[u].x, sg.cell, sh), hash_coord(p[u].y, sg.cell, sh),
                                        hash_coord(p[u].z, sg.cell, sh), M, mc) << 16 | i;
        }
    }
    __syncthreads();
    const uint32_t bits = M > 1 ? 32u - (uint32_t)__clz(M - 1u) : 0u;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // wave w owns the contiguous chunk [c0, c1) of the array: the chunks in wave order and the
    // keys in each chunk in order make the sort stable
    const uint32_t chunk = (n + NW - 1) / NW;
    const uint32_t c0 = min(wave * chunk, n), c1 = min(c0 + chunk, n);
    uint32_t* __restrict__ hw = wcnt[wave];
    uint32_t cur = 0;
    for (uint32_t shift = 16; shift < 16 + bits; shift += 8) {
        const int nb = (int)min(8u, 16u + bits - shift);   // digit bits of this pass (the key has `bits`)
        // 1. per-wave digit counts (LDS atomics on the wave's own counters: order does not matter)
        for (uint32_t k = lane; k < 256; k += 64) hw[k] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = c0 + lane; i < c1; i += 64) atomicAdd(&hw[(buf[cur][i] >> shift) & 255u], 1u);
        __syncthreads();
        // 2. each digit's start: digits in order, then waves in order (first 256 threads)
        uint32_t tot = 0, incl = 0;
        if (tid < 256) {
            for (uint32_t w = 0; w < NW; w++) tot += wcnt[w][tid];
            incl = wave_incl_scan(tot);
            if (lane == 63) wsum[wave] = incl;
        }
        __syncthreads();
        if (tid < 256) {
            uint32_t run = incl - tot;
            for (uint32_t w = 0; w < wave; w++) run += wsum[w];
            for (uint32_t w = 0; w < NW; w++) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = run;
                run += c;
            }
        }
        __syncthreads();
        // 3. each wave scatters its chunk in order
        for (uint32_t r0 = c0; r0 < c1; r0 += 64) {
            const uint32_t i = r0 + lane;
            const bool valid = i < c1;
            const uint32_t e = valid ? buf[cur][i] : 0u;
            const uint32_t dg = (e >> shift) & 255u;
            // lanes with this lane's digit: per digit bit, the ballot of the bit or its complement
            // (invalid lanes hold digit 0 and are cleared by the first mask)
            uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                if (b >= nb) break;                 // bits above the key are zero in every digit
                const uint32_t m = (uint32_t)((int32_t)(dg << (31 - b)) >> 31);   // bit b as 0 / ~0
                const uint64_t bal = __builtin_amdgcn_ballot_w64(m != 0u);
                peers &= ~(bal ^ ((uint64_t)m << 32 | m));   // lanes whose bit b equals this lane's
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            const uint32_t at = hw[dg];
            if (valid) buf[cur ^ 1][at + rank] = e;
            __builtin_amdgcn_wave_barrier();
            if (valid && rank == 0) hw[dg] = at + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        cur ^= 1;
    }
    // sorted keys and values, cell start / end (as k_cell_bounds); the inverse permutation
    uint32_t* __restrict__ ko = d.keys[dst] + sg.pbase;
    uint32_t* __restrict__ vo = d.vals[dst] + sg.pbase;
    uint32_t* __restrict__ inv = buf[cur ^ 1];
    for (uint32_t g = tid; g < n; g += VTX_HASH_LDS_BT) {
        const uint32_t e = buf[cur][g], key = e >> 16, id = e & 0xFFFFu;
        inv[id] = g;
        ko[g] = key;
        vo[g] = id;
        if (g == 0) {
            cells[key].x = 0;
        } else {
            const uint32_t prev = buf[cur][g - 1] >> 16;
            if (key != prev) {
                cells[key].x = g;
                cells[prev].y = g;
            }
        }
        if (g == n - 1) cells[key].y = n;
    }
    __syncthreads();
    // sorted records written from particle order: coalesced reads, scattered 32-B records
#pragma unroll
    for (uint32_t s0 = 0; s0 < NS; s0++) {
        const uint32_t i0 = tid + s0 * STEP;
        float4 x0[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * VTX_HASH_LDS_BT;
            x0[u] = i < n ? d.X0[sg.pbase + i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * VTX_HASH_LDS_BT;
            if (i >= n) continue;
            const size_t o = 2 * (size_t)(sg.pbase + inv[i]);
            d.sPX[o] = make_float4(kx[s0][u], ky[s0][u], kz[s0][u], __uint_as_float(i));
            d.sPX[o + 1] = x0[u];
        }
    }
}

// CacheNeighbors_Kernel (Velvet/SpatialHashGPU.cu:79-130). Cells are visited x-, y-, z-nested;
// inside a cell in sorted order, at most max_nbr entries per cell and in total; the list ends
// with ~0u unless full. Squared distances are fmaf(dz,dz,fmaf(dy,dy,dx*dx)).
// Thread -> particle mapping (ORDER), all giving the same table:
//   NBR_PARTICLE: index order over the segment's particle tiles;
//   NBR_SORTED:   hash-sorted order, so the lanes of one cell walk identical candidate ranges and
//                 their loads merge; after the sweep, lanes re-dealt by candidate total;
//   NBR_PATCH:    16 x 16 vertex patches of a grid cloth (its stencil tiles): a wave is 4 rows x
//                 16 columns, so lanes share cells and candidates, and the 16 lanes of a patch
//                 row hold consecutive particles, whose chunks are consecutive in the table.
// A first sweep over the 27 cells finds the non-empty ones (bit q of mask) and the total
// candidate count, and keeps the ranges of the first VTX_NB_CACHE non-empty cells in this
// lane's LDS column, packed as start | length << 25 (self-collision segments have < 2^25
// particles, lengths are capped at max_nbr <= 64); a lane with more non-empty cells recomputes
// the rest. The walk is one flat loop over the cells' concatenation in the same order, with the
// next candidate's position and initial position loaded while the current one is tested.
#define VTX_NB_CACHE 16
#ifndef VTX_NB_BATCH
#define VTX_NB_BATCH 3   // cell ranges of the first sweep loaded per batch (27 = 9 batches of 3)
#endif

// The float4 at record i < 2^25 of the 32-B sorted records from a wave-uniform base (the
// position half or, one float4 on, the initial-position half): a 32-bit byte offset lets the
// load use the scalar-base + 32-bit vector-offset form.
__device__ __forceinline__ float4 ld_sorted(const float4* base, uint32_t i) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (size_t)((i & 0x1FFFFFFu) << 5));
}

// What the first sweep over the 27 cells leaves for the walk: the non-empty cells (bit q of
// mask, q = 9 dx + 3 dy + dz) and the candidate total.
struct NbCells {
    uint32_t mask, total;
};

// The first sweep of the 27-cell walk (CacheNeighbors_Kernel, Velvet/SpatialHashGPU.cu:94-129) for
// a particle at p: keeps the ranges of the first VTX_NB_CACHE non-empty cells in LDS column col.
template <int BS, int NB = VTX_NB_BATCH>
__device__ __forceinline__ NbCells neighbor_cells(const float4 p, const uint2* __restrict__ cells, const float cell,
                                                  const uint32_t M, const uint64_t mc, const uint32_t K,
                                                  const uint32_t shift, uint32_t (&crange)[VTX_NB_CACHE][BS],
                                                  const uint32_t col) {
    const int32_t ix = hash_coord(p.x, cell, shift), iy = hash_coord(p.y, cell, shift),
                  iz = hash_coord(p.z, cell, shift);
    uint32_t mask = 0, total = 0, nz = 0;
    // NB cell ranges in flight at a time, each one 8-byte load at a 32-bit byte offset
    // (tables have M < 2^28 cells)
#pragma unroll
    for (int q0 = 0; q0 < 27; q0 += NB) {
        uint64_t cv[NB];
#pragma unroll
        for (int r = 0; r < NB; r++) {
            const int q = q0 + r;
            const uint32_t hq = hash_coords(ix - 1 + q / 9, iy - 1 + (q / 3) % 3, iz - 1 + q % 3, M, mc);
            cv[r] = *reinterpret_cast<const uint64_t*>(reinterpret_cast<const char*>(cells) + (hq << 3));
        }
#pragma unroll
        for (int r = 0; r < NB; r++) {
            const int q = q0 + r;
            const uint32_t cx = (uint32_t)cv[r], cy = (uint32_t)(cv[r] >> 32);
            const uint32_t len = cx == VTX_INVALID ? 0u : min(cy, cx + K) - cx;
            if (len && nz < VTX_NB_CACHE) crange[nz][col] = cx | len << 25;
            nz += len ? 1u : 0u;
            mask |= (len ? 1u : 0u) << q;
            total += len;
        }
    }
    return NbCells{mask, total};
}

// The candidate walk after the sweep c (same function, same particle p / o / id; crange column col
// from the sweep): sP the 32-B sorted records (position with w = index at 2i, initial position at
// 2i + 1), cells the (start, end) table of M cells, shift the grid shift of the cell coordinates.
// The list goes to nb, the particle's column of an n-particle table (entry offsets):
//   CHUNK = false: entry k at offset k * n (column-major, the kernel-launch tier's layout);
//   CHUNK = true:  entries in chunks of four, chunk c = entries 4c..4c+3 at offset 4 (c * n)
//                  (nb at the table's entry 4 id, chunk-aligned): a lane's four consecutive
//                  entries share one 16-byte (32-bit entries) or 8-byte (16-bit entries) run, and
//                  a reader loads four at once.
// Entries are 32-bit or, for a segment with VTX_SEG_NBR16, 16-bit (the sentinel truncated to
// 0xffff); nb.put(k, v) stores entry offset k.
struct NbOut {
    char* base;   // the particle's first entry
    bool w16;     // 16-bit entries (uniform over the workgroup's segment)
    __device__ __forceinline__ void put(uint32_t k, uint32_t v) const {
        if (w16) *reinterpret_cast<uint16_t*>(base + ((size_t)k << 1)) = (uint16_t)v;
        else *reinterpret_cast<uint32_t*>(base + ((size_t)k << 2)) = v;
    }
};

template <int BS, bool CHUNK>
__device__ __forceinline__ void neighbor_list(NbCells c, const float4 p, const float4 o, const uint32_t id,
                                              const float4* __restrict__ sP, const uint2* __restrict__ cells,
                                              const NbOut nb, const uint32_t n, const float cell,
                                              const float cell2, const float diam2, const uint32_t M,
                                              const uint64_t mc, const uint32_t K, const uint32_t shift,
                                              const uint32_t (&crange)[VTX_NB_CACHE][BS], const uint32_t col) {
#pragma clang fp contract(off)
    const float4* __restrict__ sX0 = sP + 1;
    const uint32_t total = c.total;
    uint32_t mask = c.mask;
    if (total == 0) {
        nb.put(0, VTX_INVALID);
        return;
    }
    // position of the next candidate (called only while candidates remain): s runs through the
    // current cell's range up to e; there the walk moves on to the next non-empty cell, from
    // the LDS cache or, beyond it, from the cell table again
    uint32_t z = 0, s = 0, e = 0;
    auto next = [&]() -> uint32_t {
        if (s == e) {
            uint32_t st, ln;
            if (z < VTX_NB_CACHE) {
                const uint32_t r = crange[z][col];
                st = r & 0x1FFFFFFu;
                ln = r >> 25;
            } else {
                const int q = __builtin_ctz(mask);
                const uint2 r = cells[hash_coords(hash_coord(p.x, cell, shift) - 1 + q / 9,
                                                  hash_coord(p.y, cell, shift) - 1 + (q / 3) % 3,
                                                  hash_coord(p.z, cell, shift) - 1 + q % 3, M, mc)];
                st = r.x;
                ln = min(r.y, r.x + K) - r.x;
            }
            mask &= mask - 1u;
            z++;
            s = st;
            e = st + ln;
        }
        return s++;
    };
    // candidate test in walk order, one predicate and one predicated store at the running
    // offset koff of entry k; a full list takes no more entries
    uint32_t k = 0, koff = 0;
    auto take = [&](float4 qp, float4 r) {
        const uint32_t j = __float_as_uint(qp.w);
        const float dx = p.x - qp.x, dy = p.y - qp.y, dz = p.z - qp.z;
        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const float ex = o.x - r.x, ey = o.y - r.y, ez = o.z - r.z;
        const float e2 = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
        const bool acc = (j != id) & (d2 < cell2) & (e2 > diam2) & (k < K);
        if (acc) nb.put(koff, j);
        koff += !acc ? 0u : !CHUNK ? n : (k & 3u) == 3u ? 4 * n - 3u : 1u;
        k += acc ? 1u : 0u;
    };
    // two candidates per trip with alternating load buffers: the loads of the next candidate
    // are in flight while the current one is tested
    uint32_t cur = next();
    float4 qa = ld_sorted(sP, cur), ra = ld_sorted(sX0, cur), qb, rb;
    for (uint32_t it = 0; it < total; it += 2) {
        const bool two = it + 1 < total;
        if (two) {
            cur = next();
            qb = ld_sorted(sP, cur);
            rb = ld_sorted(sX0, cur);
        }
        take(qa, ra);
        if (!two) break;
        if (it + 2 < total) {
            cur = next();
            qa = ld_sorted(sP, cur);
            ra = ld_sorted(sX0, cur);
        }
        take(qb, rb);
    }
    if (k < K) nb.put(koff, VTX_INVALID);                        // a full list has no sentinel
}

// The whole 27-cell walk of one particle, its crange column the thread's own.
template <int BS, bool CHUNK>
__device__ __forceinline__ void neighbor_walk(const float4 p, const float4 o, const uint32_t id,
                                              const float4* __restrict__ sP, const uint2* __restrict__ cells,
                                              const NbOut nb, const uint32_t n, const float cell,
                                              const float cell2, const float diam2, const uint32_t M,
                                              const uint64_t mc, const uint32_t K, const uint32_t shift,
                                              uint32_t (&crange)[VTX_NB_CACHE][BS]) {
    const NbCells c = neighbor_cells<BS>(p, cells, cell, M, mc, K, shift, crange, threadIdx.x);
    neighbor_list<BS, CHUNK>(c, p, o, id, sP, cells, nb, n, cell, cell2, diam2, M, mc, K, shift, crange,
                             threadIdx.x);
}

// In hash-sorted order (NBR_SORTED) the workgroup's threads take its particles again after the
// sweep, in descending order of their candidate totals (buckets of 4 candidates, ties in any
// order): the walk of a wave lasts as long as its longest list, and lanes of similar totals
// (among them the lanes of one cell) waste fewer trips. A thread reads its new particle's sweep
// from the LDS column of the thread that swept it. The lists do not depend on which thread walks
// them. (Batch of 512 draped cloths: 289.6 -> 282.4 us per search; the patch and particle orders
// keep their lanes, whose table stores and loads are the ones that pay there: on one 1024 x 1024
// cloth the re-dealt patch order measured 180.2 -> 185.1 us.)
#define VTX_NB_BUCKETS 64
template <int ORDER, int BS>
__global__ __launch_bounds__(BS) void k_neighbors(DevState d, const TileDev* __restrict__ tiles,
                                                  const GridTileDev* __restrict__ gtiles, int in) {
    constexpr bool REDEAL = ORDER == NBR_SORTED;
    constexpr int RB = REDEAL ? BS : 1;
    static_assert(BS >= VTX_NB_BUCKETS, "one wave scans the bucket counts");
    __shared__ uint32_t crange[VTX_NB_CACHE][BS];   // column per sweeping thread
    __shared__ uint32_t bcount[REDEAL ? VTX_NB_BUCKETS : 1];
    __shared__ uint32_t bpart[RB], bmask[RB], btotal[RB];
    uint32_t seg, i, tix, lane;
    bool live;
    tile_lane<VTX_TILE, BS>(d.remap, tix, lane);
    if (ORDER == NBR_PATCH) {
        const GridTileDev t = gtiles[tix];
        seg = t.seg;
        const SegDev& s0 = d.segs[seg];
        const uint32_t a = t.a0 + lane / VTX_GT, b = t.b0 + lane % VTX_GT;
        live = a < s0.rows && b < s0.cols;
        i = a * s0.cols + b;
    } else {
        const TileDev t = tiles[tix];
        live = lane < t.count;
        seg = t.seg;
        i = t.begin + lane - d.segs[seg].pbase;                 // sorted position or particle
    }
    if (!REDEAL && !live) return;
    const SegDev sg = d.segs[seg];                             // one segment per workgroup
    if (!(sg.flags & VTX_SEG_SELFCOL)) return;
    const uint32_t n = sg.n, M = sg.M, K = sg.max_nbr, shift = seg_shift(sg);
    const uint64_t mc = seg_mod(sg);
    const float4* __restrict__ sP = d.sPX + 2 * (size_t)sg.pbase;   // position of sorted i at 2i, initial at 2i+1
    const uint2* __restrict__ cells = d.cells + sg.tbase;
    char* const table = reinterpret_cast<char*>(d.nbr + ((uint64_t)sg.nbase_hi << 32 | sg.nbase_lo));
    const bool n16 = sg.flags & VTX_SEG_NBR16;
    const uint32_t esh = n16 ? 1u : 2u;   // log2 of the entry size
    if constexpr (!REDEAL) {
        const float4 p = d.P[in][sg.pbase + i], o = d.X0[sg.pbase + i];
        neighbor_walk<BS, true>(p, o, i, sP, cells, NbOut{table + ((size_t)4 * i << esh), n16}, n, sg.cell,
                                sg.cell2, sg.diam2, M, mc, K, shift, crange);
    } else {
        NbCells c{0u, 0u};
        // nine cell ranges in flight per batch: this path holds fewer registers than the others
        // (batches of 3 measured 261.8 against 258.4 us on the batch of 512 cloths)
        if (live) c = neighbor_cells<BS, 9>(sP[2 * i], cells, sg.cell, M, mc, K, shift, crange, threadIdx.x);
        // descending totals: bucket VTX_NB_BUCKETS - 1 - total / 4, idle threads last
        const uint32_t t = threadIdx.x;
        if (t < VTX_NB_BUCKETS) bcount[t] = 0;
        __syncthreads();
        const uint32_t key =
            live ? VTX_NB_BUCKETS - 1u - min(c.total >> 2, VTX_NB_BUCKETS - 1u) : VTX_NB_BUCKETS - 1u;
        const uint32_t slot = atomicAdd(&bcount[key], 1u);
        __syncthreads();
        if (t < VTX_NB_BUCKETS) {
            const uint32_t v = bcount[t];
            bcount[t] = wave_incl_scan(v) - v;
        }
        __syncthreads();
        const uint32_t at = bcount[key] + slot;
        bpart[at] = live ? i : VTX_INVALID;
        bmask[at] = c.mask;
        btotal[at] = c.total | t << 16;                         // totals <= 27 * 64
        __syncthreads();
        i = bpart[t];
        if (i == VTX_INVALID) return;
        const uint32_t tc = btotal[t];
        const float4 p = sP[2 * i], o = sP[2 * i + 1];
        const uint32_t id = __float_as_uint(p.w);
        neighbor_list<BS, true>(NbCells{bmask[t], tc & 0xFFFFu}, p, o, id, sP, cells,
                                NbOut{table + ((size_t)4 * id << esh), n16}, n, sg.cell, sg.cell2, sg.diam2, M, mc,
                                K, shift, crange, tc >> 16);
    }
}

// ---------------------------------------------------------------- kernel-launch tier
// HashObjects on raw device arrays (Velvet/SpatialHashGPU.cu:159-196, SpatialHashGPU.cuh:17-25):
// one object set, positions and initial positions as packed xyz triples (glm::vec3), the
// sort's key/value ping-pong partly in the caller's particle_hash / particle_index and partly in
// a workspace, which also holds the interleaved cell table and the 32-B sorted records that
// the walk reads (one cache line per candidate).

__device__ __forceinline__ float4 ld_xyz(const float* __restrict__ a, uint32_t i, float w) {
    return make_float4(a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2], w);
}

// ComputeParticleHash_Kernel (Velvet/SpatialHashGPU.cu:34-42).
__global__ __launch_bounds__(VTX_TILE) void k_raw_keys(const float* __restrict__ pos, uint32_t n, float cell,
                                                       uint32_t M, uint64_t mc, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * VTX_TILE + threadIdx.x;
    if (i >= n) return;
    const float4 p = ld_xyz(pos, i, 0.f);
    keys[i] = hash_coords(int_coord(p.x, cell), int_coord(p.y, cell), int_coord(p.z, cell), M, mc);
    vals[i] = i;
}

// Sort tile b covers keys [b * VTX_SORT_TILE, ...); its digit g count sits at hist[g * ntiles + b].
__global__ __launch_bounds__(VTX_TILE) void k_raw_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                            uint32_t* __restrict__ hist) {
    const uint32_t b = blockIdx.x, begin = b * VTX_SORT_TILE;
    tile_hist(keys + begin, min(n - begin, (uint32_t)VTX_SORT_TILE), shift, hist + b, gridDim.x);
}

__global__ __launch_bounds__(VTX_TILE) void k_raw_sort_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, uint32_t n,
                                                               int shift, const uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    const uint32_t b = blockIdx.x, begin = b * VTX_SORT_TILE;
    tile_scatter(kin + begin, vin + begin, min(n - begin, (uint32_t)VTX_SORT_TILE), shift, hist + b, gridDim.x, 0u,
                 kout, vout);
}

// FindCellStart_Kernel (Velvet/SpatialHashGPU.cu:44-77) into the interleaved table (cells were
// filled with ~0u), plus the sorted records.
__global__ __launch_bounds__(VTX_TILE) void k_raw_cell_bounds(const uint32_t* __restrict__ skeys,
                                                              const uint32_t* __restrict__ svals, uint32_t n,
                                                              const float* __restrict__ pos,
                                                              const float* __restrict__ pos0,
                                                              uint2* __restrict__ cells, float4* __restrict__ rec) {
    const uint32_t g = blockIdx.x * VTX_TILE + threadIdx.x;
    if (g >= n) return;
    const uint32_t key = skeys[g];
    if (g == 0) {
        cells[key].x = 0;
    } else {
        const uint32_t prev = skeys[g - 1];
        if (key != prev) {
            cells[key].x = g;
            cells[prev].y = g;
        }
    }
    if (g == n - 1) cells[key].y = n;
    const uint32_t id = svals[g];
    rec[2 * (size_t)g] = ld_xyz(pos, id, __uint_as_float(id));
    rec[2 * (size_t)g + 1] = ld_xyz(pos0, id, 0.f);
}

// CacheNeighbors_Kernel (Velvet/SpatialHashGPU.cu:79-130), threads in sorted order.
template <int BS>
__global__ __launch_bounds__(BS) void k_raw_neighbors(const float4* __restrict__ rec, const uint2* __restrict__ cells,
                                                      uint32_t* __restrict__ nbr, uint32_t n, float cell, float cell2,
                                                      float diam2, uint32_t M, uint64_t mc, uint32_t K) {
    __shared__ uint32_t crange[VTX_NB_CACHE][BS];
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    const float4 p = rec[2 * (size_t)i], o = rec[2 * (size_t)i + 1];
    const uint32_t id = __float_as_uint(p.w);
    neighbor_walk<BS, false>(p, o, id, rec, cells, NbOut{reinterpret_cast<char*>(nbr + id), false}, n, cell, cell2,
                             diam2, M, mc, K, 0u, crange);
}

// The interleaved table to the caller's cell_start / cell_end (empty cells: ~0u in both).
__global__ __launch_bounds__(VTX_TILE) void k_raw_split_cells(const uint2* __restrict__ cells, uint32_t M,
                                                              uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    for (uint32_t i = blockIdx.x * VTX_TILE + threadIdx.x; i < M; i += gridDim.x * VTX_TILE) {
        const uint2 c = cells[i];
        start[i] = c.x;
        end[i] = c.y;
    }
}

static uint64_t mod_constant(uint32_t M) {   // ceil(2^64 / M) mod 2^64 (M = 1 gives 0: h % 1 = 0)
    return M > 1 ? ~0ull / M + 1ull : 0ull;
}

size_t hash_workspace_bytes(uint32_t n, uint32_t M) {
    const size_t tiles = (n + VTX_SORT_TILE - 1) / VTX_SORT_TILE;
    const size_t nhist = 256 * tiles, nparts = (nhist + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up(32 * (size_t)n) + up(8 * (size_t)M) + up(8 * (size_t)n) + up(4 * nhist) + up(4 * nparts);
}

hipError_t hash_objects(const HashObjectsArgs& a, void* workspace, hipStream_t s) {
    const uint32_t n = a.num_objects, M = a.table_size;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* w = static_cast<char*>(workspace);
    float4* rec = reinterpret_cast<float4*>(w);
    w += up(32 * (size_t)n);
    uint2* cells = reinterpret_cast<uint2*>(w);
    w += up(8 * (size_t)M);
    uint32_t* kalt = reinterpret_cast<uint32_t*>(w);
    uint32_t* valt = kalt + n;
    w += up(8 * (size_t)n);
    const uint32_t tiles = (n + VTX_SORT_TILE - 1) / VTX_SORT_TILE;
    uint32_t* hist = reinterpret_cast<uint32_t*>(w);
    w += up(4 * (size_t)256 * tiles);
    uint32_t* part = reinterpret_cast<uint32_t*>(w);

    // stable LSD radix sort on ceil(log2 M) bits (Velvet/SpatialHashGPU.cu:179), 8 per pass;
    // the keys start on the side that makes the last pass end in the caller's arrays
    const uint32_t bits = M > 1 ? 32u - (uint32_t)__builtin_clz(M - 1u) : 0u;
    const int passes = (int)((bits + 7) / 8);
    uint32_t* k[2] = {a.particle_hash, kalt};
    uint32_t* v[2] = {a.particle_index, valt};
    int src = passes & 1;
    const uint32_t blocks = (n + VTX_TILE - 1) / VTX_TILE;
    const uint64_t mc = mod_constant(M);
    if (n) hipLaunchKernelGGL(k_raw_keys, dim3(blocks), dim3(VTX_TILE), 0, s, a.positions, n, a.cell_spacing, M, mc,
                              k[src], v[src]);
    for (int pass = 0; pass < passes && n; pass++) {
        hipLaunchKernelGGL(k_raw_sort_hist, dim3(tiles), dim3(VTX_TILE), 0, s, (const uint32_t*)k[src], n, 8 * pass,
                           hist);
        hipError_t e = launch_scan(hist, 256 * tiles, part, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_raw_sort_scatter, dim3(tiles), dim3(VTX_TILE), 0, s, (const uint32_t*)k[src],
                           (const uint32_t*)v[src], n, 8 * pass, (const uint32_t*)hist, k[src ^ 1], v[src ^ 1]);
        src ^= 1;
    }
    hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(cells), 2 * M, VTX_INVALID, s);
    if (e != hipSuccess) return e;
    if (n) {
        hipLaunchKernelGGL(k_raw_cell_bounds, dim3(blocks), dim3(VTX_TILE), 0, s, (const uint32_t*)a.particle_hash,
                           (const uint32_t*)a.particle_index, n, a.positions, a.original_positions, cells, rec);
        hipLaunchKernelGGL((k_raw_neighbors<128>), dim3((n + 127) / 128), dim3(128), 0, s, (const float4*)rec,
                           (const uint2*)cells, a.neighbors, n, a.cell_spacing, a.cell_spacing2, a.particle_diameter2,
                           M, mc, a.max_num_neighbors);
    }
    uint32_t sb = (M + VTX_TILE - 1) / VTX_TILE;
    if (sb > 4096) sb = 4096;
    if (sb) hipLaunchKernelGGL(k_raw_split_cells, dim3(sb), dim3(VTX_TILE), 0, s, (const uint2*)cells, M, a.cell_start,
                               a.cell_end);
    return hipGetLastError();
}

// Workspace of hash_objects_noalloc: a fixed device array, one copy per GPU, alive until the
// process ends. Host code never passes `scratch` itself to a kernel (on the host the name is a
// shadow, not a device address): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current, and only that address is used.
static __device__ __align__(256) unsigned char scratch[VTX_HASH_SCRATCH_BYTES];

#define VTX_SCRATCH_MAX_DEVICES 64
static std::mutex scratch_mu;
static void* scratch_ptr[VTX_SCRATCH_MAX_DEVICES];

static hipError_t scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= VTX_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

hipError_t hash_objects_noalloc(const HashObjectsArgs& a, hipStream_t s) {
    if (hash_workspace_bytes(a.num_objects, a.table_size) > (size_t)VTX_HASH_SCRATCH_BYTES) return hipErrorInvalidValue;
    void* ws = nullptr;
    const hipError_t e = scratch_address(&ws);
    if (e != hipSuccess) return e;
    // every workspace part is written before it is read (records, cell table fill, sort
    // keys, histograms, scan partials), so nothing is cleared between calls
    return hash_objects(a, ws, s);
}

// ---------------------------------------------------------------- host launchers
hipError_t launch_hash_keys(const DevState& d, int in, hipStream_t s) {
    hipLaunchKernelGGL(k_hash_keys, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, in);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* a, uint32_t n, uint32_t* part, hipStream_t s) {
    const uint32_t nb = (n + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK;
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(VTX_TILE), 0, s, a, n, part);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(VTX_TILE), 0, s, a, n, (const uint32_t*)part);
    return hipGetLastError();
}

hipError_t launch_sort_pass(const DevState& d, int src, int shift, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_hist, dim3(d.nstiles), dim3(VTX_TILE), 0, s, d, src, shift);
    hipError_t e = launch_scan(d.hist + d.hbase0, d.nhist, d.part, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sort_scatter, dim3(d.nstiles), dim3(VTX_TILE), 0, s, d, src, shift);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* a, uint32_t n, uint32_t v, hipStream_t s) {
    if (!n) return hipSuccess;
    uint32_t blocks = (n + 1023) / 1024;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(1024), 0, s, a, n, v);
    return hipGetLastError();
}

hipError_t launch_hash_small(const DevState& d, int in, int dst, hipStream_t s) {
    // one launch, largest segments first: workgroups of small segments fill the CUs that the
    // large ones leave
    if (d.nhsegs)
        hipLaunchKernelGGL((k_hash_small<VTX_HASH_LDS_N>), dim3(d.nhsegs), dim3(VTX_HASH_LDS_BT), 0, s, d, d.hsegs, in,
                           dst);
    return hipGetLastError();
}

hipError_t launch_clear_cells(const DevState& d, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_cells, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_cell_bounds(const DevState& d, int in, int src, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_bounds, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, in, src);
    hipLaunchKernelGGL(k_sorted_records, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, in, src);
    return hipGetLastError();
}

template <int BS>
static void launch_neighbors_bs(const DevState& d, int in, int order, hipStream_t s) {
    constexpr uint32_t S = VTX_TILE / BS;
    const dim3 bl(BS);
    if (order == NBR_PARTICLE) {
        hipLaunchKernelGGL((k_neighbors<NBR_PARTICLE, BS>), dim3(d.ntiles * S), bl, 0, s, d, d.tiles, d.gtiles, in);
    } else if (order == NBR_SORTED) {
        hipLaunchKernelGGL((k_neighbors<NBR_SORTED, BS>), dim3(d.ntiles * S), bl, 0, s, d, d.tiles, d.gtiles, in);
    } else {   // grid segments by patches (both grid tile lists), the others in sorted order
        if (d.ngtiles)
            hipLaunchKernelGGL((k_neighbors<NBR_PATCH, BS>), dim3(d.ngtiles * S), bl, 0, s, d, d.itiles, d.gtiles, in);
        if (d.nltiles)
            hipLaunchKernelGGL((k_neighbors<NBR_PATCH, BS>), dim3(d.nltiles * S), bl, 0, s, d, d.itiles, d.ltiles, in);
        if (d.nitiles)
            hipLaunchKernelGGL((k_neighbors<NBR_SORTED, BS>), dim3(d.nitiles * S), bl, 0, s, d, d.itiles, d.gtiles, in);
    }
}

hipError_t launch_neighbors(const DevState& d, int in, int order, int block, hipStream_t s) {
    if (block == 64) launch_neighbors_bs<64>(d, in, order, s);
    else if (block == 128) launch_neighbors_bs<128>(d, in, order, s);
    else launch_neighbors_bs<256>(d, in, order, s);
    return hipGetLastError();
}

}  // namespace vtx
