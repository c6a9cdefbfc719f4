/*
 * FP64 restatement of the reference GPU XPBD method -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Every stage follows one reference kernel; the citation is on each function. Arithmetic is
 * double precision on the binary32 inputs widened to double. Jacobi semantics are kept: every
 * constraint pass reads the positions of the start of the pass, accumulates corrections and
 * counts, and applies delta / count * omega afterwards (Velvet/VtClothSolverGPU.cu:253-264).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EPSILON ((double)1e-6f)   /* Velvet/Common.cuh:21 */

typedef struct { double x, y, z; } d3;

static inline d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 ld(const double* a, int64_t i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
static inline void st(double* a, int64_t i, d3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
static inline d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 mul(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline d3 dv(d3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double len(d3 a) { return sqrt(dot(a, a)); }
static inline d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 vabs(d3 a) { return mk(fabs(a.x), fabs(a.y), fabs(a.z)); }

/* 4x4 column-major helpers: m[col*4+row]. */
static d3 xform_point(const double* m, d3 p) {
    return mk(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12],
              m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]);
}
static d3 xform_vec3(const double* m, d3 p) {   /* glm::mat3(mat4) * v */
    return mk(m[0] * p.x + m[4] * p.y + m[8] * p.z,
              m[1] * p.x + m[5] * p.y + m[9] * p.z,
              m[2] * p.x + m[6] * p.y + m[10] * p.z);
}
static void mat4_mul(const double* a, const double* b, double* out) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a[k * 4 + r] * b[c * 4 + k];
            out[c * 4 + r] = s;
        }
}
static void mat4_inverse(const double* m, double* inv) {
    /* Gauss-Jordan on a row-major copy. */
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int k = 0; k < 8; k++) { double t = a[c][k]; a[c][k] = a[piv][k]; a[piv][k] = t; }
        double d = a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] /= d;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            double f = a[r][c];
            for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) inv[c * 4 + r] = a[r][c + 4];
}

static double sgn(double v) { return (v > 0) ? 1.0 : (v < 0 ? -1.0 : 0.0); }

/* SDFCollider::ComputeSDF (Velvet/VtClothSolverGPU.cuh:22-89). */
static d3 compute_sdf(const orc_collider* c, const double* inv_cur, d3 p, double margin) {
    if (c->type == 1) { /* plane: only position.y is used (cuh:24-31) */
        double offset = p.y - (c->position[1] + margin);
        if (offset < 0) return mk(0, -offset, 0);
    } else if (c->type == 0) { /* sphere: radius scale.x + margin (cuh:32-43) */
        double radius = c->scale[0] + margin;
        d3 diff = sub(p, mk(c->position[0], c->position[1], c->position[2]));
        double distance = len(diff);
        double offset = distance - radius;
        if (offset < 0) {
            d3 direction = dv(diff, distance);
            return mul(direction, -offset);
        }
    } else if (c->type == 2) { /* cube with rounded corners (cuh:44-87) */
        d3 correction = mk(0, 0, 0);
        d3 local = xform_point(inv_cur, p);
        d3 size = mk(0.5 + margin / c->scale[0], 0.5 + margin / c->scale[1], 0.5 + margin / c->scale[2]);
        d3 la = vabs(local);
        d3 offset = sub(la, size);
        double maxv = fmax(offset.x, fmax(offset.y, offset.z));
        double minv = fmin(offset.x, fmin(offset.y, offset.z));
        double midv = offset.x + offset.y + offset.z - maxv - minv;
        double scalar = 1.0;
        if (maxv < 0) {
            double m = (double)0.03f;
            if (midv > -m) scalar = (double)0.2f;
            if (minv > -m) {
                d3 mask = mk(offset.x < 0 ? sgn(local.x) : 0, offset.y < 0 ? sgn(local.y) : 0,
                             offset.z < 0 ? sgn(local.z) : 0);
                d3 vec = add(offset, mk(m, m, m));
                double l = len(vec);
                if (l < m) {
                    d3 nv = dv(vec, l);
                    correction = mk(mask.x * nv.x * (m - l), mask.y * nv.y * (m - l), mask.z * nv.z * (m - l));
                }
            } else if (offset.x == maxv) {
                correction = mk(copysign(-offset.x, local.x), 0, 0);
            } else if (offset.y == maxv) {
                correction = mk(0, copysign(-offset.y, local.y), 0);
            } else if (offset.z == maxv) {
                correction = mk(0, 0, copysign(-offset.z, local.z));
            }
        }
        return xform_vec3(c->cur, mul(correction, scalar));
    }
    return mk(0, 0, 0);
}

/* ComputeFriction (Velvet/VtClothSolverGPU.cu:272-287). */
static d3 compute_friction(d3 correction, d3 rel, double friction) {
    d3 f = mk(0, 0, 0);
    double cl = len(correction);
    if (friction > 0 && cl > 0) {
        d3 nrm = dv(correction, cl);
        d3 tan = sub(rel, mul(nrm, dot(rel, nrm)));
        double tl = len(tan);
        double maxt = cl * friction;
        double s = maxt / tl;
        if (!(s < 1.0)) s = 1.0;   /* min(maxTan / tanLen, 1): inf and 1 both give 1 */
        f = mul(tan, -s);
    }
    return f;
}

/* PredictPositions_Kernel (Velvet/VtClothSolverGPU.cu:42-53): every particle, pinned included. */
void orc_predict(int32_t n, const double* x, double* v, double* p, const double g[3], double h) {
    for (int32_t i = 0; i < n; i++) {
        d3 vi = add(ld(v, i), mul(mk(g[0], g[1], g[2]), h));
        st(v, i, vi);
        st(p, i, add(ld(x, i), mul(vi, h)));
    }
}

/* CollideSDF_Kernel (Velvet/VtClothSolverGPU.cu:289-314), colliders in order; friction only
 * when the correction is nonzero; relative displacement pred - pos - VelocityAt(pred) * dt,
 * VelocityAt = (p - last * inv(cur) * p) / collider dt (Velvet/VtClothSolverGPU.cuh:91-96). */
void orc_collide_sdf(int32_t n, double* p, const double* x, const orc_collider* cols, int32_t C,
                     double margin, double friction, double dt) {
    if (C <= 0) return;
    double* inv = (double*)malloc(sizeof(double) * 16 * C);
    double* rel = (double*)malloc(sizeof(double) * 16 * C);
    for (int32_t c = 0; c < C; c++) {
        mat4_inverse(cols[c].cur, inv + 16 * c);
        mat4_mul(cols[c].last, inv + 16 * c, rel + 16 * c);
    }
    for (int32_t i = 0; i < n; i++) {
        d3 pos = ld(x, i);
        d3 pred = ld(p, i);
        for (int32_t c = 0; c < C; c++) {
            d3 corr = compute_sdf(&cols[c], inv + 16 * c, pred, margin);
            pred = add(pred, corr);
            if (dot(corr, corr) > 0) {
                d3 last = xform_point(rel + 16 * c, pred);
                d3 vel = dv(sub(pred, last), cols[c].delta_time);
                d3 relv = sub(sub(pred, pos), mul(vel, dt));
                pred = add(pred, compute_friction(corr, relv, friction));
            }
        }
        st(p, i, pred);
    }
    free(inv);
    free(rel);
}

/* One Jacobi iteration: SolveStretch (cu:65-102), SolveAttachment (cu:205-235),
 * SolveBending (cu:117-189), then ApplyDeltas (cu:253-264). */
void orc_constraint_iteration(const orc_params* prm, const orc_scene* sc, double* p, double h) {
    const int32_t n = sc->n;
    double* delta = (double*)calloc((size_t)3 * n, sizeof(double));
    int32_t* count = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    const double* w = sc->w;

    for (int32_t c = 0; c < sc->S; c++) {
        int32_t i1 = sc->st[2 * c], i2 = sc->st[2 * c + 1];
        double rest = sc->st_rest[c];
        d3 diff = sub(ld(p, i1), ld(p, i2));
        double distance = len(diff);
        double w1 = w[i1], w2 = w[i2];
        if (distance != rest && w1 + w2 > 0) {         /* bilateral (cu:85) */
            d3 grad = dv(diff, distance + EPSILON);
            double lambda = (distance - rest) / (w1 + w2) * prm->stretch_stiffness;
            d3 common = mul(grad, lambda);
            st(delta, i1, add(ld(delta, i1), mul(common, -w1)));
            st(delta, i2, add(ld(delta, i2), mul(common, w2)));
            count[i1]++;
            count[i2]++;
        }
    }
    for (int32_t c = 0; c < sc->A; c++) {
        int32_t pid = sc->at_pid[c];
        d3 slot = ld(sc->slots, sc->at_slot[c]);
        double target = sc->at_dist[c] * prm->long_range_stretchiness;
        if (w[pid] == 0 && target > 0) continue;       /* cu:222 */
        d3 diff = sub(ld(p, pid), slot);
        double dist = len(diff);
        if (dist > target) {                           /* one-sided (cu:228) */
            d3 corr = add(mul(diff, -1.0), mul(dv(diff, dist), target));
            st(delta, pid, add(ld(delta, pid), corr));
            count[pid]++;
        }
    }
    double xpbd_bend = prm->bend_compliance / h / h;   /* cu:173 */
    for (int32_t c = 0; c < sc->B; c++) {
        uint32_t i0 = sc->bd[4 * c], i1 = sc->bd[4 * c + 1], i2 = sc->bd[4 * c + 2], i3 = sc->bd[4 * c + 3];
        double rest = sc->bd_rest[c];
        double w0 = w[i0], w1 = w[i1], w2 = w[i2], w3 = w[i3];
        d3 p0 = ld(p, i0), p1 = ld(p, i1), p2 = ld(p, i2), p3 = ld(p, i3);
        d3 e = sub(p3, p2);
        double elen = len(e);
        if (elen < EPSILON) continue;
        double inv_elen = 1.0 / elen;
        d3 n1 = cross(sub(p2, p0), sub(p3, p0));
        n1 = dv(n1, dot(n1, n1));
        d3 n2 = cross(sub(p3, p1), sub(p2, p1));
        n2 = dv(n2, dot(n2, n2));
        d3 d0 = mul(n1, elen);
        d3 d1 = mul(n2, elen);
        d3 d2 = add(mul(n1, dot(sub(p0, p3), e) * inv_elen), mul(n2, dot(sub(p1, p3), e) * inv_elen));
        d3 dd3 = add(mul(n1, dot(sub(p2, p0), e) * inv_elen), mul(n2, dot(sub(p2, p1), e) * inv_elen));
        n1 = dv(n1, len(n1));
        n2 = dv(n2, len(n2));
        double cd = dot(n1, n2);
        if (cd < -1.0) cd = -1.0;
        if (cd > 1.0) cd = 1.0;
        double phi = acos(cd);
        double lambda = w0 * dot(d0, d0) + w1 * dot(d1, d1) + w2 * dot(d2, d2) + w3 * dot(dd3, dd3);
        if (lambda < EPSILON) continue;
        lambda = (phi - rest) / (lambda + xpbd_bend);
        if (dot(cross(n1, n2), e) > 0.0) lambda = -lambda;
        st(delta, i0, add(ld(delta, i0), mul(d0, -w0 * lambda)));
        st(delta, i1, add(ld(delta, i1), mul(d1, -w1 * lambda)));
        st(delta, i2, add(ld(delta, i2), mul(d2, -w2 * lambda)));
        st(delta, i3, add(ld(delta, i3), mul(dd3, -w3 * lambda)));
        count[i0]++;
        count[i1]++;
        count[i2]++;
        count[i3]++;
    }
    for (int32_t i = 0; i < n; i++) {
        if (count[i] > 0)
            st(p, i, add(ld(p, i), mul(dv(ld(delta, i), (double)count[i]), prm->relaxation_factor)));
    }
    free(delta);
    free(count);
}

/* Finalize_Kernel (Velvet/VtClothSolverGPU.cu:388-407). */
void orc_finalize(int32_t n, double* x, double* v, const double* p, double h, double max_speed,
                  double damping) {
    for (int32_t i = 0; i < n; i++) {
        d3 xi = ld(x, i);
        d3 np = ld(p, i);
        d3 raw = dv(sub(np, xi), h);
        double l = len(raw);
        if (l > max_speed) {
            raw = mul(dv(raw, l), max_speed);
            np = add(xi, mul(raw, h));
        }
        st(v, i, mul(raw, 1.0 - damping * h));
        st(x, i, np);
    }
}

/* ComputeNormal (Velvet/VtClothSolverGPU.cu:419-465): area-weighted sum, then normalize. */
void orc_normals(int32_t n, const double* x, int32_t T, const uint32_t* tris, double* normals) {
    memset(normals, 0, sizeof(double) * 3 * (size_t)n);
    for (int32_t t = 0; t < T; t++) {
        uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        d3 p1 = ld(x, a), p2 = ld(x, b), p3 = ld(x, c);
        d3 nr = cross(sub(p2, p1), sub(p3, p1));
        st(normals, a, add(ld(normals, a), nr));
        st(normals, b, add(ld(normals, b), nr));
        st(normals, c, add(ld(normals, c), nr));
    }
    for (int32_t i = 0; i < n; i++) {
        d3 s = ld(normals, i);
        st(normals, i, dv(s, len(s)));   /* zero sum -> NaN, as glm::normalize */
    }
}

/* HashCoords with int32 wrap-around and C remainder (Velvet/SpatialHashGPU.cu:18-22). */
static int32_t hash_coords(int32_t x, int32_t y, int32_t z, int32_t M) {
    uint32_t hx = (uint32_t)x * 92837111u, hy = (uint32_t)y * 689287499u, hz = (uint32_t)z * 283923481u;
    int32_t h = (int32_t)(hx ^ hy ^ hz);
    int32_t r = h % M;
    return r < 0 ? -r : r;
}

/* ComputeParticleHash + stable sort + FindCellStart + CacheNeighbors in FP64
 * (Velvet/SpatialHashGPU.cu:34-130,159-196). */
static int32_t coord64(double v, double cell, int32_t shift) {
    return (int32_t)((uint32_t)(int32_t)floor(v / cell) + (uint32_t)shift);
}

int orc_hash64(int32_t n, const double* p, const double* x0, double cell, double diameter,
               int32_t M, int32_t max_nbr, int32_t shift, uint32_t* cell_start, uint32_t* cell_end,
               uint32_t* sorted_index, uint32_t* nbr) {
    int32_t* hash = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* cnt = (uint32_t*)calloc((size_t)M + 1, sizeof(uint32_t));
    if (!hash || !cnt) return -1;
    for (int32_t i = 0; i < n; i++) {
        int32_t ix = coord64(p[3 * i], cell, shift);
        int32_t iy = coord64(p[3 * i + 1], cell, shift);
        int32_t iz = coord64(p[3 * i + 2], cell, shift);
        hash[i] = hash_coords(ix, iy, iz, M);
        cnt[hash[i]]++;
    }
    /* counting sort in index order == stable sort by hash */
    uint32_t run = 0;
    for (int32_t h = 0; h < M; h++) {
        uint32_t c = cnt[h];
        cell_start[h] = c ? run : 0xffffffffu;
        cell_end[h] = c ? run + c : 0xffffffffu;
        cnt[h] = run;
        run += c;
    }
    for (int32_t i = 0; i < n; i++) sorted_index[cnt[hash[i]]++] = (uint32_t)i;

    const double cell2 = cell * cell, diam2 = diameter * diameter;
    for (int32_t id = 0; id < n; id++) {
        d3 pos = ld(p, id), org = ld(x0, id);
        int32_t ix = coord64(pos.x, cell, shift), iy = coord64(pos.y, cell, shift),
                iz = coord64(pos.z, cell, shift);
        int32_t k = 0;
        int full = 0;
        for (int32_t xx = ix - 1; xx <= ix + 1 && !full; xx++)
            for (int32_t yy = iy - 1; yy <= iy + 1 && !full; yy++)
                for (int32_t zz = iz - 1; zz <= iz + 1 && !full; zz++) {
                    int32_t h = hash_coords(xx, yy, zz, M);
                    uint32_t start = cell_start[h];
                    if (start == 0xffffffffu) continue;
                    uint32_t end = cell_end[h];
                    if (end > start + (uint32_t)max_nbr) end = start + (uint32_t)max_nbr;
                    for (uint32_t s = start; s < end; s++) {
                        uint32_t j = sorted_index[s];
                        if (j != (uint32_t)id) {
                            d3 dp = sub(pos, ld(p, j)), d0 = sub(org, ld(x0, j));
                            if (dot(dp, dp) < cell2 && dot(d0, d0) > diam2) {
                                nbr[(size_t)k * n + id] = j;
                                k++;
                                if (k >= max_nbr) { full = 1; break; }
                            }
                        }
                    }
                }
        if (k < max_nbr) nbr[(size_t)k * n + id] = 0xffffffffu;
    }
    free(hash);
    free(cnt);
    return 0;
}

/* CollideParticles_Kernel + ApplyDeltas (Velvet/VtClothSolverGPU.cu:329-386): Jacobi contact
 * over the cached neighbour column; writes (not accumulates) delta and count. */
void orc_collide_particles(int32_t n, double* p, const double* x, const double* w,
                           const uint32_t* nbr, int32_t max_nbr, double diameter,
                           double friction, double omega) {
    double* newp = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    for (int32_t id = 0; id < n; id++) {
        d3 delta = mk(0, 0, 0);
        int32_t cnt = 0;
        d3 pi = ld(p, id);
        d3 vel_i = sub(pi, ld(x, id));
        double wi = w[id];
        for (int32_t k = 0; k < max_nbr; k++) {
            uint32_t j = nbr[(size_t)k * n + id];
            if (j >= (uint32_t)n) break;
            double denom = wi + w[j];
            if (denom <= 0) continue;
            d3 pj = ld(p, j);
            d3 diff = sub(pi, pj);
            double dist = len(diff);
            if (dist >= diameter) continue;
            d3 grad = dv(diff, dist + EPSILON);
            double lambda = (dist - diameter) / denom;
            d3 common = mul(grad, lambda);
            cnt++;
            delta = sub(delta, mul(common, wi));
            d3 relv = sub(vel_i, sub(pj, ld(x, j)));
            delta = add(delta, mul(compute_friction(common, relv, friction), wi));
        }
        if (cnt > 0) pi = add(pi, mul(dv(delta, (double)cnt), omega));
        st(newp, id, pi);
    }
    memcpy(p, newp, sizeof(double) * 3 * (size_t)n);
    free(newp);
}

/* VtClothSolverGPU::Simulate (Velvet/VtClothSolverGPU.hpp:56-111). */
int orc_simulate(const orc_params* prm, orc_scene* sc, int32_t frames, int32_t max_substeps) {
    const int32_t n = sc->n;
    const int32_t M = prm->hash_table_size > 0 ? prm->hash_table_size : 2 * n;
    const int32_t K = prm->max_num_neighbors;
    double* p = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    uint32_t* cs = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)M);
    uint32_t* ce = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)M);
    uint32_t* si = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* nbr = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K * (size_t)(n > 0 ? n : 1));
    if (!p || !cs || !ce || !si || !nbr) return -1;
    for (size_t i = 0; i < (size_t)K * (size_t)n; i++) nbr[i] = 0xffffffffu;
    int32_t done = 0;
    const double frame_dt = prm->frame_dt;
    const double h = frame_dt / prm->num_substeps;
    for (int32_t f = 0; f < frames; f++) {
        /* pre-stabilisation: predicted aliases positions, frame dt (hpp:73) */
        orc_collide_sdf(n, sc->x, sc->x, sc->cols, sc->C, prm->collision_margin, prm->friction, frame_dt);
        for (int32_t s = 0; s < prm->num_substeps; s++) {
            if (max_substeps >= 0 && done >= max_substeps) goto out;
            orc_predict(n, sc->x, sc->v, p, prm->gravity, h);
            if (prm->enable_self_collision) {
                if (s % prm->interleaved_hash == 0)
                    orc_hash64(n, p, sc->x0, prm->hash_cell_size, prm->particle_diameter, M, K, prm->hash_shift, cs,
                               ce, si, nbr);
                orc_collide_particles(n, p, sc->x, sc->w, nbr, K, prm->particle_diameter, prm->friction,
                                      prm->relaxation_factor);
            }
            orc_collide_sdf(n, p, sc->x, sc->cols, sc->C, prm->collision_margin, prm->friction, h);
            for (int32_t it = 0; it < prm->num_iterations; it++) orc_constraint_iteration(prm, sc, p, h);
            orc_finalize(n, sc->x, sc->v, p, h, prm->max_speed, prm->damping);
            done++;
        }
        if (sc->normals) orc_normals(n, sc->x, sc->T, sc->tris, sc->normals);
    }
out:
    free(p);
    free(cs);
    free(ce);
    free(si);
    free(nbr);
    return 0;
}
