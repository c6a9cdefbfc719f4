/*
 * binary32 restatement of the reference's CPU cloth solver -- TEST INFRASTRUCTURE ONLY.
 *
 * It is the CPU timing baseline: VtClothSolverCPU (Velvet/VtClothSolverCPU.hpp:17-434) with
 * SpatialHashCPU (Velvet/SpatialHashCPU.hpp:11-124) and the CPU collider SDFs
 * (Velvet/Collider.hpp:43-77), single-threaded as in the reference. It is a different method
 * (Gauss-Seidel, unilateral stretch, hard pins) and is never compared for parity.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define K_EPS 1e-6f   /* Velvet/VtClothSolverCPU.hpp:422 */

typedef struct { float x, y, z; } f3;
static inline f3 F(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 fadd(f3 a, f3 b) { return F(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 fsub(f3 a, f3 b) { return F(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 fmulv(f3 a, float s) { return F(a.x * s, a.y * s, a.z * s); }
static inline f3 fdivv(f3 a, float s) { return F(a.x / s, a.y / s, a.z / s); }
static inline float fdot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float flen(f3 a) { return sqrtf(fdot(a, a)); }
static inline f3 fcross(f3 a, f3 b) { return F(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline f3 fnorm(f3 a) { return fmulv(a, 1.0f / sqrtf(fdot(a, a))); }

struct cpu_solver {
    int32_t n, res1, T;
    f3 *x, *p, *v, *normals;
    float* w;
    int32_t S;
    int32_t* st;
    float* st_rest;
    int32_t B;
    int32_t* bd;
    int32_t nat;
    int32_t* at;
    f3* at_pos;
    uint32_t* tris;
    float diameter;
    int32_t nsph, has_plane;
    f3* sc;
    float* sr;
    int32_t substeps, iterations;
    f3 g;
    float damping, friction, margin, bend_compliance;
    /* SpatialHashCPU */
    int32_t M;
    int32_t* cell_start;
    int32_t* cell_entries;
    int32_t** nb;
    int32_t* nb_n;
    int32_t* nb_cap;
};

/* SpatialHashCPU::ComputeIntCoord / HashCoords (Velvet/SpatialHashCPU.hpp:63-72). */
static int32_t int_coord(float v, float s) { return (int32_t)floorf(v / s); }
static int32_t hash_xyz(int32_t x, int32_t y, int32_t z, int32_t M) {
    uint32_t hx = (uint32_t)x * 92837111u, hy = (uint32_t)y * 689287499u, hz = (uint32_t)z * 283923481u;
    int32_t h = (int32_t)(hx ^ hy ^ hz);
    int32_t r = h % M;
    return r < 0 ? -r : r;
}
static int32_t hash_pos(const cpu_solver* s, f3 p) {
    return hash_xyz(int_coord(p.x, s->diameter), int_coord(p.y, s->diameter), int_coord(p.z, s->diameter), s->M);
}

/* SpatialHashCPU::HashObjects + CacheNeighbors (Velvet/SpatialHashCPU.hpp:23-53,86-123). */
static void hash_objects(cpu_solver* s, const f3* pos) {
    memset(s->cell_start, 0, sizeof(int32_t) * (size_t)(s->M + 1));
    for (int32_t i = 0; i < s->n; i++) s->cell_start[hash_pos(s, pos[i])]++;
    int32_t start = 0;
    for (int32_t i = 0; i < s->M; i++) {
        start += s->cell_start[i];
        s->cell_start[i] = start;
    }
    s->cell_start[s->M] = start;
    for (int32_t i = 0; i < s->n; i++) {
        int32_t h = hash_pos(s, pos[i]);
        s->cell_start[h]--;
        s->cell_entries[s->cell_start[h]] = i;
    }
    for (int32_t i = 0; i < s->n; i++) {
        s->nb_n[i] = 0;
        int32_t ix = int_coord(pos[i].x, s->diameter), iy = int_coord(pos[i].y, s->diameter),
                iz = int_coord(pos[i].z, s->diameter);
        for (int32_t xx = ix - 1; xx <= ix + 1; xx++)
            for (int32_t yy = iy - 1; yy <= iy + 1; yy++)
                for (int32_t zz = iz - 1; zz <= iz + 1; zz++) {
                    int32_t h = hash_xyz(xx, yy, zz, s->M);
                    for (int32_t k = s->cell_start[h]; k < s->cell_start[h + 1]; k++) {
                        if (s->nb_n[i] == s->nb_cap[i]) {
                            s->nb_cap[i] = s->nb_cap[i] ? 2 * s->nb_cap[i] : 32;
                            s->nb[i] = (int32_t*)realloc(s->nb[i], sizeof(int32_t) * (size_t)s->nb_cap[i]);
                        }
                        s->nb[i][s->nb_n[i]++] = s->cell_entries[k];
                    }
                }
    }
}

/* Collider::ComputeSDF for plane and sphere (Velvet/Collider.hpp:43-77). */
static f3 friction_of(const cpu_solver* s, f3 corr, f3 rel) {
    f3 f = F(0, 0, 0);
    float cl = flen(corr);
    if (s->friction > 0 && cl > 0) {
        f3 nrm = fdivv(corr, cl);
        f3 tan = fsub(rel, fmulv(nrm, fdot(rel, nrm)));
        float tl = flen(tan);
        float m = cl * s->friction / tl;
        f = fmulv(tan, -(m < 1.0f ? m : 1.0f));
    }
    return f;
}

/* VtClothSolverCPU::CollideSDF (Velvet/VtClothSolverCPU.hpp:274-290). */
static void collide_sdf(cpu_solver* s, f3* pos) {
    for (int32_t i = 0; i < s->n; i++) {
        int32_t nc = s->nsph + (s->has_plane ? 1 : 0);
        for (int32_t c = 0; c < nc; c++) {
            f3 q = pos[i];
            f3 corr = F(0, 0, 0);
            if (c < s->nsph) {
                float radius = s->sr[c] + s->margin;
                f3 diff = fsub(q, s->sc[c]);
                float d = flen(diff);
                if (d < radius) corr = fmulv(fdivv(diff, d), radius - d);
            } else if (q.y < s->margin) {
                corr = F(0, s->margin - q.y, 0);
            }
            pos[i] = fadd(pos[i], corr);
            f3 rel = fsub(pos[i], s->x[i]);
            pos[i] = fadd(pos[i], friction_of(s, corr, rel));
        }
    }
}

static void predict(cpu_solver* s, float dt) {
    for (int32_t i = 0; i < s->n; i++) {
        s->v[i] = fadd(s->v[i], fmulv(s->g, dt));
        s->p[i] = fadd(s->x[i], fmulv(s->v[i], dt));
    }
}

/* Gauss-Seidel passes (Velvet/VtClothSolverCPU.hpp:194-349). */
static void solve_stretch(cpu_solver* s) {
    for (int32_t c = 0; c < s->S; c++) {
        int32_t a = s->st[2 * c], b = s->st[2 * c + 1];
        f3 diff = fsub(s->p[a], s->p[b]);
        float d = flen(diff);
        float w1 = s->w[a], w2 = s->w[b];
        if (d > s->st_rest[c] && w1 + w2 > 0) {
            f3 g = fdivv(diff, d + K_EPS);
            float lambda = (d - s->st_rest[c]) / (w1 + w2);
            s->p[a] = fsub(s->p[a], fmulv(g, w1 * lambda));
            s->p[b] = fadd(s->p[b], fmulv(g, w2 * lambda));
        }
    }
}

static float glm_clamp01(float d) {
    float m = (d < 0.0f) ? 0.0f : d;     /* glm max(x, 0): (x < 0) ? 0 : x keeps NaN */
    return (1.0f < m) ? 1.0f : m;
}

static void solve_bending(cpu_solver* s, float dt) {
    float xpbd = s->bend_compliance / dt / dt;
    for (int32_t c = 0; c < s->B; c++) {
        int32_t i1 = s->bd[4 * c + 2], i2 = s->bd[4 * c + 1], i3 = s->bd[4 * c], i4 = s->bd[4 * c + 3];
        float w1 = s->w[i1], w2 = s->w[i2], w3 = s->w[i3], w4 = s->w[i4];
        f3 p1 = s->p[i1];
        f3 p2 = fsub(s->p[i2], p1), p3 = fsub(s->p[i3], p1), p4 = fsub(s->p[i4], p1);
        f3 n1 = fnorm(fcross(p2, p3));
        f3 n2 = fnorm(fcross(p2, p4));
        float d = glm_clamp01(fdot(n1, n2));
        float angle = acosf(d);
        if (angle < K_EPS || isnan(d)) continue;
        float l23 = flen(fcross(p2, p3)) + K_EPS, l24 = flen(fcross(p2, p4)) + K_EPS;
        f3 q3 = fdivv(fadd(fcross(p2, n2), fmulv(fcross(n1, p2), d)), l23);
        f3 q4 = fdivv(fadd(fcross(p2, n1), fmulv(fcross(n2, p2), d)), l24);
        f3 q2 = fsub(fmulv(fdivv(fadd(fcross(p3, n2), fmulv(fcross(n1, p3), d)), l23), -1.0f),
                     fdivv(fadd(fcross(p4, n1), fmulv(fcross(n2, p4), d)), l24));
        f3 q1 = fsub(fsub(fmulv(q2, -1.0f), q3), q4);
        float denom = xpbd + (w1 * fdot(q1, q1) + w2 * fdot(q2, q2) + w3 * fdot(q3, q3) + w4 * fdot(q4, q4));
        if (denom < K_EPS) continue;
        float lambda = sqrtf(1.0f - d * d) * (angle - 0.0f) / denom;
        s->p[i1] = fadd(s->p[i1], fmulv(q1, w1 * lambda));
        s->p[i2] = fadd(s->p[i2], fmulv(q2, w2 * lambda));
        s->p[i3] = fadd(s->p[i3], fmulv(q3, w3 * lambda));
        s->p[i4] = fadd(s->p[i4], fmulv(q4, w4 * lambda));
    }
}

static void collide_particles(cpu_solver* s) {
    for (int32_t i = 0; i < s->n; i++) {
        for (int32_t k = 0; k < s->nb_n[i]; k++) {
            int32_t j = s->nb[i][k];
            if (i >= j) continue;
            f3 diff = fsub(s->p[i], s->p[j]);
            float d = flen(diff);
            float w1 = s->w[i], w2 = s->w[j];
            if (d < s->diameter && w1 + w2 > 0) {
                f3 g = fdivv(diff, d + K_EPS);
                float lambda = (d - s->diameter) / (w1 + w2);
                f3 common = fmulv(g, lambda);
                s->p[i] = fsub(s->p[i], fmulv(common, w1));
                s->p[j] = fadd(s->p[j], fmulv(common, w2));
                f3 rel = fsub(fsub(s->p[i], s->x[i]), fsub(s->p[j], s->x[j]));
                f3 f = friction_of(s, common, rel);
                s->p[i] = fadd(s->p[i], fmulv(f, w1));
                s->p[j] = fsub(s->p[j], fmulv(f, w2));
            }
        }
    }
}

static void solve_attachment(cpu_solver* s) {
    for (int32_t a = 0; a < s->nat; a++) s->p[s->at[a]] = s->at_pos[a];
}

static void finalize(cpu_solver* s, float dt) {
    for (int32_t i = 0; i < s->n; i++) {
        s->v[i] = fmulv(fdivv(fsub(s->p[i], s->x[i]), dt), 1.0f - s->damping * dt);
        s->x[i] = s->p[i];
    }
}

static void compute_normals(cpu_solver* s) {
    memset(s->normals, 0, sizeof(f3) * (size_t)s->n);
    for (int32_t t = 0; t < s->T; t++) {
        uint32_t a = s->tris[3 * t], b = s->tris[3 * t + 1], c = s->tris[3 * t + 2];
        f3 nr = fcross(fsub(s->x[b], s->x[a]), fsub(s->x[c], s->x[a]));
        s->normals[a] = fadd(s->normals[a], nr);
        s->normals[b] = fadd(s->normals[b], nr);
        s->normals[c] = fadd(s->normals[c], nr);
    }
    for (int32_t i = 0; i < s->n; i++) s->normals[i] = fnorm(s->normals[i]);
}

cpu_solver* cpu_solver_create(int32_t res1, const float* x, int32_t n, const uint32_t* tris, int32_t T,
                              const int32_t* attached, int32_t nattached, const float* sphere_c,
                              const float* sphere_r, int32_t nspheres, int32_t has_plane, int32_t substeps,
                              int32_t iterations, const float gravity[3], float damping, float friction,
                              float margin, float bend_compliance) {
    cpu_solver* s = (cpu_solver*)calloc(1, sizeof(cpu_solver));
    s->n = n;
    s->res1 = res1;
    s->T = T;
    s->x = (f3*)malloc(sizeof(f3) * (size_t)n);
    s->p = (f3*)calloc((size_t)n, sizeof(f3));
    s->v = (f3*)calloc((size_t)n, sizeof(f3));
    s->normals = (f3*)calloc((size_t)n, sizeof(f3));
    s->w = (float*)malloc(sizeof(float) * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        s->x[i] = F(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        s->w[i] = 1.0f;
    }
    s->tris = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)T);
    memcpy(s->tris, tris, sizeof(uint32_t) * 3 * (size_t)T);
    s->diameter = flen(fsub(s->x[0], s->x[res1]));     /* Velvet/VtClothSolverCPU.hpp:61 */
    /* GenerateStretch (Velvet/VtClothSolverCPU.hpp:109-151) */
    int32_t r = res1 - 1;
    s->st = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)n);
    s->st_rest = (float*)malloc(sizeof(float) * 4 * (size_t)n);
    for (int32_t xx = 0; xx < res1; xx++)
        for (int32_t yy = 0; yy < res1; yy++) {
            int32_t pairs[4][2];
            int32_t np = 0;
            if (yy != r) { pairs[np][0] = xx * res1 + yy; pairs[np][1] = xx * res1 + yy + 1; np++; }
            if (xx != r) { pairs[np][0] = xx * res1 + yy; pairs[np][1] = (xx + 1) * res1 + yy; np++; }
            if (yy != r && xx != r) {
                pairs[np][0] = xx * res1 + yy; pairs[np][1] = (xx + 1) * res1 + yy + 1; np++;
                pairs[np][0] = xx * res1 + yy + 1; pairs[np][1] = (xx + 1) * res1 + yy; np++;
            }
            for (int32_t k = 0; k < np; k++) {
                s->st[2 * s->S] = pairs[k][0];
                s->st[2 * s->S + 1] = pairs[k][1];
                s->st_rest[s->S] = flen(fsub(s->x[pairs[k][0]], s->x[pairs[k][1]]));
                s->S++;
            }
        }
    /* GenerateAttachment: hard pins (Velvet/VtClothSolverCPU.hpp:153-160) */
    s->nat = nattached;
    s->at = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nattached > 0 ? nattached : 1));
    s->at_pos = (f3*)malloc(sizeof(f3) * (size_t)(nattached > 0 ? nattached : 1));
    for (int32_t a = 0; a < nattached; a++) {
        s->at[a] = attached[a];
        s->at_pos[a] = s->x[attached[a]];
        s->w[attached[a]] = 0.0f;
    }
    /* GenerateBending (Velvet/VtClothSolverCPU.hpp:162-176): (i0, i1, i2, i5) per 6 indices */
    s->B = T / 2;
    s->bd = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(s->B > 0 ? s->B : 1));
    for (int32_t q = 0; q < s->B; q++) {
        s->bd[4 * q] = (int32_t)tris[6 * q];
        s->bd[4 * q + 1] = (int32_t)tris[6 * q + 1];
        s->bd[4 * q + 2] = (int32_t)tris[6 * q + 2];
        s->bd[4 * q + 3] = (int32_t)tris[6 * q + 5];
    }
    s->nsph = nspheres;
    s->has_plane = has_plane;
    s->sc = (f3*)malloc(sizeof(f3) * (size_t)(nspheres > 0 ? nspheres : 1));
    s->sr = (float*)malloc(sizeof(float) * (size_t)(nspheres > 0 ? nspheres : 1));
    for (int32_t c = 0; c < nspheres; c++) {
        s->sc[c] = F(sphere_c[3 * c], sphere_c[3 * c + 1], sphere_c[3 * c + 2]);
        s->sr[c] = sphere_r[c];
    }
    s->substeps = substeps;
    s->iterations = iterations;
    s->g = F(gravity[0], gravity[1], gravity[2]);
    s->damping = damping;
    s->friction = friction;
    s->margin = margin;
    s->bend_compliance = bend_compliance;
    s->M = 2 * n;
    s->cell_start = (int32_t*)calloc((size_t)s->M + 1, sizeof(int32_t));
    s->cell_entries = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    s->nb = (int32_t**)calloc((size_t)n, sizeof(int32_t*));
    s->nb_n = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    s->nb_cap = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    return s;
}

/* VtClothSolverCPU::Simulate (Velvet/VtClothSolverCPU.hpp:69-100), GL upload excluded. */
void cpu_solver_simulate(cpu_solver* s, float frame_dt) {
    float h = frame_dt / (float)s->substeps;
    collide_sdf(s, s->x);
    predict(s, frame_dt);
    hash_objects(s, s->p);
    for (int32_t sub = 0; sub < s->substeps; sub++) {
        predict(s, h);
        for (int32_t it = 0; it < s->iterations; it++) {
            solve_stretch(s);
            solve_bending(s, h);
            collide_particles(s);
            collide_sdf(s, s->p);
            solve_attachment(s);
        }
        finalize(s, h);
    }
    compute_normals(s);
}

void cpu_solver_positions(const cpu_solver* s, float* out) {
    for (int32_t i = 0; i < s->n; i++) {
        out[3 * i] = s->x[i].x;
        out[3 * i + 1] = s->x[i].y;
        out[3 * i + 2] = s->x[i].z;
    }
}

/* Test access: replace positions and velocities (n xyz triples each; NULL keeps them), read the
 * velocities, and read each particle's neighbour-list length from the frame's hash (every entry
 * of the 27 probed cells, itself and duplicates included, Velvet/SpatialHashCPU.hpp:96-123). */
void cpu_solver_set_state(cpu_solver* s, const float* x, const float* v) {
    for (int32_t i = 0; i < s->n; i++) {
        if (x) s->x[i] = F(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        if (v) s->v[i] = F(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
}

void cpu_solver_velocities(const cpu_solver* s, float* out) {
    for (int32_t i = 0; i < s->n; i++) {
        out[3 * i] = s->v[i].x;
        out[3 * i + 1] = s->v[i].y;
        out[3 * i + 2] = s->v[i].z;
    }
}

void cpu_solver_neighbor_counts(const cpu_solver* s, int32_t* out) {
    for (int32_t i = 0; i < s->n; i++) out[i] = s->nb_n[i];
}

void cpu_solver_destroy(cpu_solver* s) {
    if (!s) return;
    for (int32_t i = 0; i < s->n; i++) free(s->nb[i]);
    free(s->nb); free(s->nb_n); free(s->nb_cap); free(s->cell_start); free(s->cell_entries);
    free(s->x); free(s->p); free(s->v); free(s->normals); free(s->w); free(s->st); free(s->st_rest);
    free(s->bd); free(s->at); free(s->at_pos); free(s->tris); free(s->sc); free(s->sr);
    free(s);
}
