/*
 * binary32 restatement of the reference's spatial-hash stage -- TEST INFRASTRUCTURE ONLY.
 *
 * Used to check the device hash stage bit for bit: cell coordinates are
 * (int)floorf(x / cell) with IEEE binary32 division (Velvet/SpatialHashGPU.cu:13-16), the hash
 * multiplies wrap in 32 bits and the remainder truncates toward zero
 * (Velvet/SpatialHashGPU.cu:18-22), the sort is stable by hash (the reference's
 * cub::DeviceRadixSort::SortPairs, Velvet/SpatialHashGPU.cu:133-157), and the neighbour filter
 * is Velvet/SpatialHashGPU.cu:79-130.
 *
 * Squared distances are evaluated as fmaf(dz, dz, fmaf(dy, dy, dx * dx)): glm's dot(v, v) is
 * (x*x + y*y) + z*z, and this is its contraction into fused multiply-adds. The device kernel
 * evaluates the same expression, so both sides round identically. This file must be compiled
 * with -ffp-contract=off so that the compiler adds no other contraction.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int32_t h32_hash_coords(int32_t x, int32_t y, int32_t z, int32_t M) {
    uint32_t hx = (uint32_t)x * 92837111u, hy = (uint32_t)y * 689287499u, hz = (uint32_t)z * 283923481u;
    int32_t h = (int32_t)(hx ^ hy ^ hz);
    int32_t r = h % M;
    return r < 0 ? -r : r;
}

int32_t h32_int_coord(float value, float cell) {
    volatile float q = value / cell;     /* one correctly rounded binary32 division */
    return (int32_t)floorf(q);
}

/* the cell coordinate as hashed: shifted by `shift` cells with 32-bit wrap-around */
static inline int32_t hashed_coord(float value, float cell, int32_t shift) {
    return (int32_t)((uint32_t)h32_int_coord(value, cell) + (uint32_t)shift);
}

static inline float dist2(const float* a, const float* b) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

int h32_hash(int32_t n, const float* p, const float* x0, float cell, float diameter, int32_t M,
             int32_t max_nbr, int32_t shift, uint32_t* particle_hash, uint32_t* sorted_hash,
             uint32_t* sorted_index, uint32_t* cell_start, uint32_t* cell_end, uint32_t* nbr) {
    uint32_t* cnt = (uint32_t*)calloc((size_t)M + 1, sizeof(uint32_t));
    if (!cnt) return -1;
    for (int32_t i = 0; i < n; i++) {
        int32_t ix = hashed_coord(p[3 * i], cell, shift), iy = hashed_coord(p[3 * i + 1], cell, shift),
                iz = hashed_coord(p[3 * i + 2], cell, shift);
        particle_hash[i] = (uint32_t)h32_hash_coords(ix, iy, iz, M);
        cnt[particle_hash[i]]++;
    }
    uint32_t run = 0;
    for (int32_t h = 0; h < M; h++) {
        uint32_t c = cnt[h];
        cell_start[h] = c ? run : 0xffffffffu;
        cell_end[h] = c ? run + c : 0xffffffffu;
        cnt[h] = run;
        run += c;
    }
    for (int32_t i = 0; i < n; i++) {
        uint32_t pos = cnt[particle_hash[i]]++;
        sorted_index[pos] = (uint32_t)i;
        sorted_hash[pos] = particle_hash[i];
    }
    const float cell2 = cell * cell;              /* Velvet/SpatialHashGPU.hpp:44 */
    const float diam2 = diameter * diameter;      /* Velvet/SpatialHashGPU.hpp:47 */
    for (int32_t id = 0; id < n; id++) {
        const float* pos = p + 3 * id;
        const float* org = x0 + 3 * id;
        int32_t ix = hashed_coord(pos[0], cell, shift), iy = hashed_coord(pos[1], cell, shift),
                iz = hashed_coord(pos[2], cell, shift);
        int32_t k = 0;
        int full = 0;
        for (int32_t xx = ix - 1; xx <= ix + 1 && !full; xx++)
            for (int32_t yy = iy - 1; yy <= iy + 1 && !full; yy++)
                for (int32_t zz = iz - 1; zz <= iz + 1 && !full; zz++) {
                    int32_t h = h32_hash_coords(xx, yy, zz, M);
                    uint32_t start = cell_start[h];
                    if (start == 0xffffffffu) continue;
                    uint32_t end = cell_end[h];
                    if (end > start + (uint32_t)max_nbr) end = start + (uint32_t)max_nbr;
                    for (uint32_t s = start; s < end; s++) {
                        uint32_t j = sorted_index[s];
                        if (j != (uint32_t)id && dist2(pos, p + 3 * j) < cell2 &&
                            dist2(org, x0 + 3 * j) > diam2) {
                            nbr[(size_t)k * n + id] = j;
                            k++;
                            if (k >= max_nbr) { full = 1; break; }
                        }
                    }
                }
        if (k < max_nbr) nbr[(size_t)k * n + id] = 0xffffffffu;
    }
    free(cnt);
    return 0;
}
