// XPBD solver passes for gfx950.
//
// One workgroup = one particle tile of one segment (VTX_TILE threads, one particle each).
// Every constraint pass is an atomic-free gather: a particle walks its own incident
// constraints (sliced ELL), evaluates each one in full from the pass's input positions and
// keeps only its own correction, so the reference's Jacobi scatter-with-atomics
// (Velvet/VtClothSolverGPU.cu:13-21, 65-251) becomes a fixed-order, deterministic sum and a
// pass reads P[in] and writes P[out]. The reference's ApplyDeltas (cu:253-270) is fused
// into the end of each pass; Finalize (cu:388-417) and the next substep's PredictPositions
// (cu:42-63) are fused into the last pass of a substep.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "vec.hpp"
#include "vtx_device.hpp"

namespace vtx {

// Hardware reciprocal / reciprocal square root / square root (1 ulp) for the solver passes.
// (The library is built with correctly rounded division and square root, which the hash needs
// for bit-exact cell coordinates; the solver passes are held to the FP64 oracle's tolerance.)
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float flen(V3 a) { return fsqrt(dot(a, a)); }

// arccos on [-1, 1]: the minimax form sqrt(1 - x) * (a0 + a1 x + ... + a7 x^7) of Abramowitz &
// Stegun, Handbook of Mathematical Functions, 4.4.46 (|error| <= 2e-8 on [0, 1]), with
// acos(x) = pi - acos(-x) below 0. In binary32 it stays within 4.3e-7 of the exact value
// (checked against double-precision arccos over [-1, 1]), 2.3e-8 near 1 where bending angles
// are small.
__device__ __forceinline__ float facos(float x) {
    const float ax = fabsf(x);
    float p = -0.0012624911f;
    p = fmaf(p, ax, 0.0066700901f);
    p = fmaf(p, ax, -0.0170881256f);
    p = fmaf(p, ax, 0.0308918810f);
    p = fmaf(p, ax, -0.0501743046f);
    p = fmaf(p, ax, 0.0889789874f);
    p = fmaf(p, ax, -0.2145988016f);
    p = fmaf(p, ax, 1.5707963050f);
    const float r = fsqrt(1.0f - ax) * p;
    return x < 0.f ? 3.14159265358979f - r : r;
}

// ComputeFriction (Velvet/VtClothSolverGPU.cu:272-287). tl == 0 gives mt * inf = inf, clamped
// to 1 as the reference's mt / 0 is.
__device__ __forceinline__ V3 friction_of(V3 correction, V3 rel, float friction) {
    V3 f = v3(0.f, 0.f, 0.f);
    float cl = flen(correction);
    if (friction > 0.f && cl > 0.f) {
        V3 nrm = correction * frcp(cl);
        V3 tan = rel - nrm * dot(rel, nrm);
        float tl = flen(tan);
        float mt = cl * friction;
        f = -tan * fminf(mt * frcp(tl), 1.0f);
    }
    return f;
}

__device__ __forceinline__ float sgnf(float v) { return (v > 0.f) ? 1.0f : (v < 0.f ? -1.0f : 0.0f); }

// SDFCollider::ComputeSDF (Velvet/VtClothSolverGPU.cuh:22-89).
__device__ V3 compute_sdf(const ColliderDev& c, V3 p, float margin) {
    if (c.type == 1) {
        float offset = p.y - (c.pos[1] + margin);
        if (offset < 0.f) return v3(0.f, -offset, 0.f);
    } else if (c.type == 0) {
        float radius = c.scale[0] + margin;
        V3 diff = p - v3(c.pos[0], c.pos[1], c.pos[2]);
        float distance = flen(diff);
        float offset = distance - radius;
        if (offset < 0.f) {
            V3 direction = diff * frcp(distance);
            return -offset * direction;
        }
    } else if (c.type == 2) {
        V3 correction = v3(0.f, 0.f, 0.f);
        V3 local = xform_point(c.inv, p);
        V3 size = v3(0.5f + margin * frcp(c.scale[0]), 0.5f + margin * frcp(c.scale[1]),
                     0.5f + margin * frcp(c.scale[2]));
        V3 offset = vabs(local) - size;
        float maxv = fmaxf(offset.x, fmaxf(offset.y, offset.z));
        float minv = fminf(offset.x, fminf(offset.y, offset.z));
        float midv = offset.x + offset.y + offset.z - maxv - minv;
        float scalar = 1.0f;
        if (maxv < 0.f) {
            const float m = 0.03f;
            if (midv > -m) scalar = 0.2f;
            if (minv > -m) {
                V3 mask = v3(offset.x < 0.f ? sgnf(local.x) : 0.f, offset.y < 0.f ? sgnf(local.y) : 0.f,
                             offset.z < 0.f ? sgnf(local.z) : 0.f);
                V3 vec = offset + v3(m, m, m);
                float l = flen(vec);
                if (l < m) {
                    V3 nv = vec * frsq(dot(vec, vec));   // glm::normalize
                    correction = v3(mask.x * nv.x * (m - l), mask.y * nv.y * (m - l), mask.z * nv.z * (m - l));
                }
            } else if (offset.x == maxv) {
                correction = v3(copysignf(-offset.x, local.x), 0.f, 0.f);
            } else if (offset.y == maxv) {
                correction = v3(0.f, copysignf(-offset.y, local.y), 0.f);
            } else if (offset.z == maxv) {
                correction = v3(0.f, 0.f, copysignf(-offset.z, local.z));
            }
        }
        return xform_vec(c.cur3, correction * scalar);
    }
    return v3(0.f, 0.f, 0.f);
}

// CollideSDF_Kernel body (Velvet/VtClothSolverGPU.cu:289-314) for one particle: colliders in
// order, friction only for a nonzero correction, collider velocity from VelocityAt
// (Velvet/VtClothSolverGPU.cuh:91-96).
__device__ V3 collide_sdf(V3 pred, V3 pos, const ColliderDev* cols, uint32_t cb, uint32_t cc,
                          float margin, float friction, float dt) {
    for (uint32_t c = cb; c < cb + cc; c++) {
        const ColliderDev& col = cols[c];
        V3 corr = compute_sdf(col, pred, margin);
        pred += corr;
        if (dot(corr, corr) > 0.f) {
            V3 lastp = xform_point(col.rel, pred);
            V3 vel = (pred - lastp) * frcp(col.dt);
            V3 relv = pred - pos - vel * dt;
            pred += friction_of(corr, relv, friction);
        }
    }
    return pred;
}

// Pre-stabilisation SDF on positions with the frame time (Velvet/VtClothSolverGPU.hpp:73) and
// the first substep's PredictPositions (cu:42-53). The predicted velocity is not stored: every
// substep ends in Finalize, which derives the velocity from the positions alone and stores it, so
// between frames (where the host and the failure check read it) V holds Finalize's value, as in
// the reference, and nothing reads V inside a frame.
__global__ __launch_bounds__(VTX_TILE) void k_frame_begin(DevState d, float frame_dt, float h) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + threadIdx.x;
    float4 x4 = d.X[g];
    V3 x = xyz(x4);
    if (sg.ccount) {
        x = collide_sdf(x, x, d.cols, sg.cbase, sg.ccount, sg.margin, sg.friction, frame_dt);
        d.X[g] = f4(x, x4.w);
    }
    const V3 v = xyz(d.V[g]) + v3(sg.gx, sg.gy, sg.gz) * h;
    d.P[0][g] = f4(x + v * h, x4.w);
}

// CollideParticles_Kernel + ApplyDeltas (Velvet/VtClothSolverGPU.cu:329-386) followed by the
// substep's CollideSDF (cu:289-314, hpp:87).
template <int BS>
__global__ __launch_bounds__(BS) void k_collide(DevState d, int in, float h) {
    uint32_t tix, lane;
    tile_lane<VTX_TILE, BS>(d.remap, tix, lane);
    const TileDev t = d.tiles[tix];
    if (lane >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + lane;
    const float4* __restrict__ P = d.P[in];
    const float4 p4 = P[g];
    V3 p = xyz(p4);
    const float w = p4.w;
    const V3 x = xyz(d.X[g]);
    if (sg.flags & VTX_SEG_SELFCOL) {
        const uint32_t n = sg.n;
        // the list in chunks of four entries: chunk c at nb4[c * n] (32-bit entries) or nb2[c * n]
        // (16-bit entries, 0xffff >= n the sentinel)
        const uint32_t* __restrict__ table = d.nbr + ((uint64_t)sg.nbase_hi << 32 | sg.nbase_lo);
        const uint4* __restrict__ nb4 = reinterpret_cast<const uint4*>(table) + (g - sg.pbase);
        const uint2* __restrict__ nb2 = reinterpret_cast<const uint2*>(table) + (g - sg.pbase);
        const bool n16 = sg.flags & VTX_SEG_NBR16;
        V3 delta = v3(0.f, 0.f, 0.f);
        int cnt = 0;
        const V3 vel_i = p - x;
        // neighbours four at a time (one chunk): while a group's positions load, the next chunk
        // loads too; the four terms are then added in list order. An entry counts only if every
        // entry before it is a particle index (the list ends at the first sentinel, later rows are
        // stale). A neighbour's start-of-substep position enters only the friction of a contact,
        // so it is loaded for contacts alone: most listed neighbours of a resting cloth lie
        // beyond the diameter, and the pass is bound by the cache lines its gathers touch.
        bool live = true;
        const uint32_t K = sg.max_nbr;
        uint32_t jn[4];
        auto chunk = [&](uint32_t k0) {   // entries k0..k0+3, those at K or beyond as sentinels
            uint4 c;
            if (n16) {
                const uint2 h = nb2[(size_t)(k0 >> 2) * n];
                c = make_uint4(h.x & 0xFFFFu, h.x >> 16, h.y & 0xFFFFu, h.y >> 16);
            } else {
                c = nb4[(size_t)(k0 >> 2) * n];
            }
            jn[0] = c.x;
            jn[1] = k0 + 1 < K ? c.y : VTX_INVALID;
            jn[2] = k0 + 2 < K ? c.z : VTX_INVALID;
            jn[3] = k0 + 3 < K ? c.w : VTX_INVALID;
        };
        chunk(0);
        for (uint32_t k0 = 0; k0 < K && live; k0 += 4) {
            bool ok[4];
            uint32_t jg[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                live = live && jn[u] < n;
                ok[u] = live;
                jg[u] = sg.pbase + (ok[u] ? jn[u] : 0u);
            }
            float4 pj4[4], xj4[4];
#pragma unroll
            for (int u = 0; u < 4; u++) pj4[u] = P[jg[u]];
            if (live && k0 + 4 < K) chunk(k0 + 4);
            bool hit[4];
            float dist[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                dist[u] = flen(p - xyz(pj4[u]));
                hit[u] = ok[u] && !(w + pj4[u].w <= 0.f) && !(dist[u] >= sg.diameter);
                if (hit[u]) xj4[u] = d.X[jg[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (!hit[u]) continue;
                const float denom = w + pj4[u].w;
                const V3 pj = xyz(pj4[u]);
                const V3 diff = p - pj;
                const V3 grad = diff * frcp(dist[u] + VTX_EPSILON);
                const float lambda = (dist[u] - sg.diameter) * frcp(denom);
                const V3 common = lambda * grad;
                cnt++;
                delta -= w * common;
                const V3 relv = vel_i - (pj - xyz(xj4[u]));
                delta += w * friction_of(common, relv, sg.friction);
            }
        }
        if (cnt > 0) p += delta * (frcp((float)cnt) * sg.omega);
    }
    if (sg.ccount) p = collide_sdf(p, x, d.cols, sg.cbase, sg.ccount, sg.margin, sg.friction, h);
    d.P[in ^ 1][g] = f4(p, w);
}

// Epilogue modes of the Jacobi pass.
enum { EPI_NONE = 0, EPI_FINALIZE = 1, EPI_FINALIZE_PREDICT = 2 };

// One stretch constraint seen from one endpoint (Velvet/VtClothSolverGPU.cu:80-98): bilateral,
// compliance 0, eps in the gradient denominator. q = other endpoint (xyz, inverse mass);
// nws = -w * stiffness, the endpoint's factor of every one of its terms.
__device__ __forceinline__ void stretch_term(V3 p, float w, float nws, float4 q, float rest, V3& delta, int& cnt) {
    const V3 diff = p - xyz(q);
    const float dist = fsqrt(dot(diff, diff));
    const float wsum = w + q.w;
    if (dist != rest && wsum > 0.f) {
        delta += ((dist - rest) * nws * frcp(wsum * (dist + VTX_EPSILON))) * diff;
        cnt++;
    }
}

// One long-range attachment (Velvet/VtClothSolverGPU.cu:218-234): with t the target distance,
// a particle farther than t from its slot moves by -diff + diff / |diff| * t, which is
// diff * (t / |diff| - 1); "|diff| > t" is the same test as "t / |diff| - 1 < 0", so one
// reciprocal square root serves both. Branch-free (selects): the slot position is used on every
// path, so its loads can be issued ahead of the term instead of inside a conditional block.
__device__ __forceinline__ void attach_term(V3 p, float w, float4 slot, float dist0, float lrs, V3& delta,
                                            int& cnt) {
    const float target = dist0 * lrs;
    const V3 diff = p - xyz(slot);
    const float f = __builtin_fmaf(target, frsq(dot(diff, diff)), -1.0f);
    // the reference returns early for a pinned particle with a positive target
    const bool act = !(w == 0.f && target > 0.f) && f < 0.f;
    const float k = act ? f : 0.f;
    delta.x = __builtin_fmaf(diff.x, k, delta.x);
    delta.y = __builtin_fmaf(diff.y, k, delta.y);
    delta.z = __builtin_fmaf(diff.z, k, delta.z);
    cnt += act ? 1 : 0;
}

// Slot range [s0, s1) of part q of a segment's attachments (see VTX_LRA_PARTS).
__device__ __forceinline__ void lra_part_range(const SegDev& sg, uint32_t q, uint32_t& s0, uint32_t& s1) {
    const uint32_t qs = (sg.scount + VTX_LRA_PARTS - 1) / VTX_LRA_PARTS;
    s0 = min(q * qs, sg.scount);
    s1 = min(s0 + qs, sg.scount);
}

// Dense attachments of one particle to slots [s0, s1), summed from zero in slot order.
// Distances are fetched LB slots at a time so that several loads are in flight per lane (a large
// cloth may give only one wave per SIMD); every LB gives the same sum.
template <int LB>
__device__ __forceinline__ void attach_dense(const DevState& d, const SegDev& sg, uint32_t local, uint32_t s0,
                                             uint32_t s1, V3 p, float w, V3& part, int& cnt) {
    const float* __restrict__ lra = d.lra + sg.lbase + local;
    const float4* __restrict__ slot = d.slots + sg.sbase;
    const size_t n = sg.n;
    uint32_t s = s0;
    for (; s + LB <= s1; s += LB) {
        float r[LB];
#pragma unroll
        for (int u = 0; u < LB; u++) r[u] = lra[(size_t)(s + u) * n];
        // slot positions (wave-uniform) four at a time, so that few are live at once
#pragma unroll
        for (int g = 0; g < LB; g += 4) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; u++) q[u] = slot[s + g + u];
#pragma unroll
            for (int u = 0; u < 4; u++) attach_term(p, w, q[u], r[g + u], sg.lrs, part, cnt);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (; s < s1; s++) attach_term(p, w, slot[s], lra[(size_t)s * n], sg.lrs, part, cnt);
}

// All attachments of one particle, added to delta part by part (VTX_LRA_PARTS): dense
// (slot-major distances, the slot position is the same for every lane) or from the tile's attach
// ELL in registration order, where a part ends wherever the slot's part changes -- so attachments
// registered particle-major in slot order sum exactly as the dense form does.
template <int LB>
__device__ __forceinline__ void attach_terms(const DevState& d, const SegDev& sg, uint32_t at_off, uint32_t at_k,
                                             uint32_t lane, uint32_t local, V3 p, float w, V3& delta, int& cnt) {
    if (sg.flags & VTX_SEG_LRA_DENSE) {
        if (sg.scount <= VTX_LRA_PARTS) {   // every slot is a part of its own
            const float* __restrict__ lra = d.lra + sg.lbase + local;
            const float4* __restrict__ slot = d.slots + sg.sbase;
            for (uint32_t s = 0; s < sg.scount; s++) {
                V3 part = v3(0.f, 0.f, 0.f);
                attach_term(p, w, slot[s], lra[(size_t)s * sg.n], sg.lrs, part, cnt);
                delta += part;
            }
            return;
        }
#pragma unroll 1
        for (uint32_t q = 0; q < VTX_LRA_PARTS; q++) {
            uint32_t s0, s1;
            lra_part_range(sg, q, s0, s1);
            V3 part = v3(0.f, 0.f, 0.f);
            attach_dense<LB>(d, sg, local, s0, s1, p, w, part, cnt);
            delta += part;
        }
        return;
    }
    const uint32_t qs = (sg.scount + VTX_LRA_PARTS - 1) / VTX_LRA_PARTS;
    const uint2* __restrict__ at = d.at + at_off + lane;
    static_assert(VTX_LRA_PARTS == 4, "the part of a listed slot is found with three compares");
    V3 part = v3(0.f, 0.f, 0.f);
    uint32_t cur = 0;
    for (uint32_t k = 0; k < at_k; k++) {
        const uint2 e = at[(size_t)k * VTX_TILE];
        if (e.x == VTX_INVALID) continue;
        const uint32_t q = (e.x >= qs ? 1u : 0u) + (e.x >= 2 * qs ? 1u : 0u) + (e.x >= 3 * qs ? 1u : 0u);
        if (q != cur) {
            delta += part;
            part = v3(0.f, 0.f, 0.f);
            cur = q;
        }
        attach_term(p, w, d.slots[sg.sbase + e.x], __uint_as_float(e.y), sg.lrs, part, cnt);
    }
    delta += part;
}

// The dihedral bending constraint (Velvet/VtClothSolverGPU.cu:128-177), evaluated through the
// edge vector e = p3 - p2, a = p2 - p0 and b = p3 - p1:
//   c1 = (p2-p0) x (p3-p0) = a x e,   c2 = (p3-p1) x (p2-p1) = e x b,
//   (p0-p3).e = -(a.e) - |e|^2,        (p1-p3).e = -(b.e),
//   (c1 x c2).e = |e|^2 (a.c2)         (since e.c2 = 0), so the sign rule needs one dot product.
// With r1 = 1/|c1| and r2 = 1/|c2| (one reciprocal square root each), the reference's
// n1 = c1 r1^2, n2 = c2 r2^2 and normalize(n1).normalize(n2) = (c1.c2) r1 r2, so every gradient is
// a combination of c1 and c2: d0 = k0 c1, d1 = k1 c2 (k0 = |e| r1^2, k1 = |e| r2^2) and
// d2 = s1 c1 + s2 c2, d3 = -(d0 + d1 + d2) = t1 c1 + t2 c2 (the four gradients of a dihedral
// angle sum to zero; the reference's d3 = (p2-p0).e/|e| n1 + (p2-p1).e/|e| n2 is the same).
// Returns false on the reference's early-outs; else lam carries the sign rule and the correction
// of vertex k is -w_k * lam * d_k.
struct BendGrad {
    V3 c1, c2, d2;   // d0 = k0 c1, d1 = k1 c2, d3 = -(d0 + d1 + d2)
    float k0, k1;
};

__device__ __forceinline__ bool bend_eval(float4 q0, float4 q1, float4 q2, float4 q3, float rest, float xpbd,
                                          float& lam, BendGrad& g) {
    const V3 p0 = xyz(q0), p1 = xyz(q1), p2 = xyz(q2), p3 = xyz(q3);
    const V3 e = p3 - p2;
    const float ee = dot(e, e);
    if (ee < VTX_EPSILON * VTX_EPSILON) return false;   // |e| < eps
    const float inv_elen = frsq(ee);                      // |e| and 1/|e| from one rsq
    const float elen = ee * inv_elen;
    const V3 a = p2 - p0, b = p3 - p1;
    g.c1 = cross(a, e);
    g.c2 = cross(e, b);
    const float g11 = dot(g.c1, g.c1), g22 = dot(g.c2, g.c2), g12 = dot(g.c1, g.c2);
    const float r1 = frsq(g11), r2 = frsq(g22);
    const float r1s = r1 * r1, r2s = r2 * r2;
    g.k0 = elen * r1s;
    g.k1 = elen * r2s;
    const float s1 = ((-dot(a, e) - ee) * inv_elen) * r1s, s2 = (-dot(b, e) * inv_elen) * r2s;
    g.d2 = s1 * g.c1 + s2 * g.c2;
    const float cd = fminf(fmaxf(g12 * (r1 * r2), -1.0f), 1.0f);
    const float phi = facos(cd);
    // |d0|^2 = |e|^2 r1^2, |d1|^2 = |e|^2 r2^2, and d3 = t1 c1 + t2 c2 from the Gram entries:
    // |d3|^2 = t1 (t1 g11 + 2 t2 g12) + t2^2 g22
    const float t1 = -(g.k0 + s1), t2 = -(g.k1 + s2), t2g = t2 * g12;
    const float d3s = __builtin_fmaf(t1, __builtin_fmaf(t1, g11, t2g + t2g), t2 * (t2 * g22));
    float l = ee * (q0.w * r1s + q1.w * r2s) + q2.w * dot(g.d2, g.d2) + q3.w * d3s;
    if (l < VTX_EPSILON) return false;
    l = (phi - rest) * frcp(l + xpbd);
    if (dot(a, g.c2) > 0.0f) l = -l;      // (n1 x n2).e > 0
    lam = l;
    return true;
}

// Finalize (Velvet/VtClothSolverGPU.cu:396-406) and optionally the next substep's Predict, whose
// velocity is not stored (the next Finalize overwrites it; see k_frame_begin).
template <int EPI>
__device__ __forceinline__ void epilogue(const DevState& d, const SegDev& sg, uint32_t g, int in, V3 p, float w,
                                         float h) {
    if (EPI == EPI_NONE) {
        d.P[in ^ 1][g] = f4(p, w);
    } else {
        const float4 x4 = d.X[g];
        const V3 x = xyz(x4);
        V3 raw = (p - x) * frcp(h);
        const float rl = flen(raw);
        if (rl > sg.max_speed) {
            raw = raw * (frcp(rl) * sg.max_speed);
            p = x + raw * h;
        }
        V3 v = raw * (1.0f - sg.damping * h);
        d.X[g] = f4(p, x4.w);
        if (EPI == EPI_FINALIZE_PREDICT) {
            v = v + v3(sg.gx, sg.gy, sg.gz) * h;
            d.P[in ^ 1][g] = f4(p + v * h, x4.w);
        } else {
            d.V[g] = f4(v, 0.f);
        }
    }
}

// One Jacobi pass on the general path: SolveStretch (cu:65-102) + SolveAttachment
// (cu:205-235) + SolveBending (cu:117-189) gathered per particle over its incident
// constraints (sliced ELL), then ApplyDeltas (cu:253-264) and the epilogue.
template <int EPI, bool CONSTRAINTS, int LB>
__global__ __launch_bounds__(VTX_TILE) void k_iterate(DevState d, const TileDev* __restrict__ tiles, int in,
                                                      float h) {
    const TileDev t = tiles[tile_of(d.remap)];
    const uint32_t lane = threadIdx.x;
    if (lane >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + lane;
    const float4* __restrict__ P = d.P[in];
    const float4 p4 = P[g];
    V3 p = xyz(p4);
    const float w = p4.w;

    if (CONSTRAINTS) {
        V3 delta = v3(0.f, 0.f, 0.f);
        int cnt = 0;
        const uint2* __restrict__ st = d.st + t.st_off + lane;
        const float nws = -w * sg.stiffness;
        for (uint32_t k = 0; k < t.st_k; k++) {
            const uint2 e = st[(size_t)k * VTX_TILE];
            if (e.x == VTX_INVALID) continue;
            stretch_term(p, w, nws, P[sg.pbase + e.x], __uint_as_float(e.y), delta, cnt);
        }
        attach_terms<LB>(d, sg, t.at_off, t.at_k, lane, g - sg.pbase, p, w, delta, cnt);
        const uint32_t* __restrict__ bd = d.bd + t.bd_off + lane;
        const float xpbd_bend = sg.bend_compliance * frcp(h) * frcp(h);   // compliance / h^2 (cu:173)
        for (uint32_t k = 0; k < t.bd_k; k++) {
            const uint32_t e = bd[(size_t)k * VTX_TILE];
            if (e == VTX_INVALID) continue;
            const uint32_t bid = sg.bbase + (e & 0x3FFFFFFFu);
            const uint32_t role = e >> 30;
            const uint4 ix = d.brec[bid];
            const uint32_t pb = sg.pbase;
            float lam;
            BendGrad bg;
            if (!bend_eval(P[pb + ix.x], P[pb + ix.y], P[pb + ix.z], P[pb + ix.w], d.brest[bid], xpbd_bend, lam, bg))
                continue;
            const V3 dk = role == 0 ? bg.k0 * bg.c1 : role == 1 ? bg.k1 * bg.c2 : role == 2 ? bg.d2
                                                                              : -(bg.k0 * bg.c1 + bg.k1 * bg.c2 + bg.d2);
            delta += (-w * lam) * dk;
            cnt++;
        }
        if (cnt > 0) p += delta * (frcp((float)cnt) * sg.omega);
    }
    epilogue<EPI>(d, sg, g, in, p, w, h);
}

// ---------------------------------------------------------------- grid stencil path
// A segment whose cloth is an R x C vertex grid with the reference's topology (structural +
// shear stretch, bending on the X/R/D slots of the cells) runs its Jacobi pass on 16 x 16
// tiles. The tile's positions plus a halo of 2 are staged in LDS; every bending slot touching
// the tile is evaluated once (not once per member vertex), its four corrections are kept in
// LDS, and each vertex gathers its 8 stretch terms, its attachments and its 12 bending terms in
// a fixed order. The three slot types are evaluated one after another through the same LDS
// area (X, then R, then D), which keeps the workgroup at ~18 KB of LDS (8 workgroups per CU).
// Slot owner ranges relative to the tile origin (a0, b0): X rows/cols [-1, 15] (17 x 17),
// R rows [-1, 16] cols [-1, 15] (18 x 17), D rows [-1, 15] cols [-1, 16] (17 x 18).
#define VTX_NSX (17 * 17)
#define VTX_NSR (18 * 17)
#define VTX_NSD (17 * 18)
#define VTX_NSMAX VTX_NSR
#define VTX_RC 18              // owner cells of any type per tile side: -1..16

template <int TYPE>
__device__ __forceinline__ int slot_index(int ra, int rb) {
    return TYPE == 2 ? (ra + 1) * 18 + (rb + 1) : (ra + 1) * 17 + (rb + 1);
}

// Owner cells of one slot type that can hold a slot in a tile, tile-relative: rows
// [ra0, ra0 + h), columns [rb0, rb0 + w). The type's owner range of the tile (X: rows and columns
// -1..15, R: rows -1..16, D: columns -1..16) clipped to the cells where the slot is geometrically
// possible (X: a <= R-2, b <= C-2; R: also a >= 1; D: also b >= 1). h or w <= 0: none.
template <int TYPE>
__device__ __forceinline__ void slot_rect(int a0, int b0, int rows, int cols, int& ra0, int& rb0, int& h, int& w) {
    constexpr int RA1 = TYPE == 1 ? 16 : 15, RB1 = TYPE == 2 ? 16 : 15;
    ra0 = max(-1, (TYPE == 1 ? 1 : 0) - a0);
    rb0 = max(-1, (TYPE == 2 ? 1 : 0) - b0);
    h = min(RA1, rows - 2 - a0) - ra0 + 1;
    w = min(RB1, cols - 2 - b0) - rb0 + 1;
}

// Evaluate the bending slot owned at tile-relative cell (ra, rb) of one type (a cell of
// slot_rect's rectangle) into LDS at its row-major index j in the type's range; act (cleared
// before) marks the slots that contribute. The slot keeps g_k = -lam d_k for roles 0..2 (planes
// 3k..3k+2); role 3's is -(g_0 + g_1 + g_2), as the four gradients of a dihedral angle sum to
// zero, and each vertex scales the sum of its terms by its own inverse mass. SREST: the rest
// angles come from srest, the tile's owner cells (rows and columns -1..16, row-major by 18) staged
// in LDS with the positions, instead of from brest4.
template <int TYPE, bool SREST>
__device__ __forceinline__ void eval_slot(const DevState& d, const SegDev& sg, const float4* sp,
                                          const float4* srest, float (*corr)[VTX_NSMAX], uint8_t* act, int a0,
                                          int b0, int cols, int ra, int rb, float xpbd) {
    constexpr int W = TYPE == 2 ? 18 : 17;                                  // owner columns of the range
    const uint32_t j = (uint32_t)((ra + 1) * W + (rb + 1));                 // row-major storage index
    const int a = a0 + ra, b = b0 + rb;
    // in the rectangle every slot of a FULL type is present; otherwise the cell's flags tell
    bool on = ((sg.flags >> (3 + TYPE)) & 1u) || ((d.gflags[sg.pbase + (uint32_t)(a * cols + b)] >> TYPE) & 1u);
    if (on) {
        const int o = (ra + VTX_GH) * VTX_GP + (rb + VTX_GH);    // owner in staged coordinates
        // canonical vertex order per type (vtx_device.hpp)
        constexpr int i0 = TYPE == 0 ? 0 : TYPE == 1 ? -VTX_GP + 1 : VTX_GP - 1;
        constexpr int i1 = TYPE == 0 ? VTX_GP + 1 : TYPE == 1 ? VTX_GP : 1;
        constexpr int i2 = TYPE == 0 ? 1 : TYPE == 1 ? 0 : VTX_GP;
        constexpr int i3 = TYPE == 0 ? VTX_GP : TYPE == 1 ? 1 : 0;
        const float4 q0 = sp[o + i0], q1 = sp[o + i1], q2 = sp[o + i2], q3 = sp[o + i3];
        float rest = 0.f;
        if (!(sg.flags & VTX_SEG_BEND_REST0)) {
            const float4 r4 = SREST ? srest[(ra + 1) * VTX_RC + (rb + 1)] : d.brest4[sg.pbase + (uint32_t)(a * cols + b)];
            rest = TYPE == 0 ? r4.x : TYPE == 1 ? r4.y : r4.z;
        }
        float lam;
        BendGrad bg;
        on = bend_eval(q0, q1, q2, q3, rest, xpbd, lam, bg);
        if (on) {
            const V3 g0 = (-lam * bg.k0) * bg.c1, g1 = (-lam * bg.k1) * bg.c2, g2 = -lam * bg.d2;
            corr[0][j] = g0.x; corr[1][j] = g0.y; corr[2][j] = g0.z;
            corr[3][j] = g1.x; corr[4][j] = g1.y; corr[5][j] = g1.z;
            corr[6][j] = g2.x; corr[7][j] = g2.y; corr[8][j] = g2.z;
            act[j] = 1;
        }
    }
}

// One slot type: the workgroup's threads take the cells of the type's rectangle densely (thread
// t: row t / w, column t % w), so a tile cut by the cloth's edge runs no lanes on cells that
// cannot hold a slot; then every tile vertex adds its four memberships of this type in role
// order to bend (gradients times -lam; the vertex's inverse mass is applied to the sum).
template <int TYPE, int BT, bool SREST>
__device__ __forceinline__ void bend_type_pass(const DevState& d, const SegDev& sg, const float4* sp,
                                               const float4* srest, float (*corr)[VTX_NSMAX], uint8_t* act,
                                               int a0, int b0,
                                               int rows, int cols, float xpbd, bool inside, int ty, int tx,
                                               V3& bend, int& cnt) {
    const uint32_t tid = threadIdx.x;
    if (!((sg.flags >> (6 + TYPE)) & 1u)) {   // the type is present at all
        int ra0, rb0, h, w;
        slot_rect<TYPE>(a0, b0, rows, cols, ra0, rb0, h, w);
        if (h > 0 && w > 0) {
            const uint32_t n = (uint32_t)(h * w), m = (65536u + (uint32_t)w - 1u) / (uint32_t)w;   // t / w = t m >> 16 (t < 1024, w <= 18)
            static_assert(2 * BT >= VTX_NSMAX, "a slot type fits two rounds of the workgroup");
#pragma unroll
            for (uint32_t r = 0; r < (BT >= VTX_NSMAX ? 1u : 2u); r++) {
                const uint32_t t = tid + r * BT;
                if (t < n) {
                    const uint32_t q = (t * m) >> 16;
                    eval_slot<TYPE, SREST>(d, sg, sp, srest, corr, act, a0, b0, cols, ra0 + (int)q, rb0 + (int)(t - q * (uint32_t)w), xpbd);
                }
            }
        }
    }
    __syncthreads();
    if (inside) {
        // role r of the slot owned at (ty + dy_r, tx + dx_r), per type:
        //   X: (0,0) (-1,-1) (0,-1) (-1,0);  R: (1,-1) (-1,0) (0,0) (0,-1);  D: (-1,1) (0,-1) (-1,0) (0,0)
        int sidx[4];
        if (TYPE == 0) {
            sidx[0] = slot_index<0>(ty, tx); sidx[1] = slot_index<0>(ty - 1, tx - 1);
            sidx[2] = slot_index<0>(ty, tx - 1); sidx[3] = slot_index<0>(ty - 1, tx);
        } else if (TYPE == 1) {
            sidx[0] = slot_index<1>(ty + 1, tx - 1); sidx[1] = slot_index<1>(ty - 1, tx);
            sidx[2] = slot_index<1>(ty, tx); sidx[3] = slot_index<1>(ty, tx - 1);
        } else {
            sidx[0] = slot_index<2>(ty - 1, tx + 1); sidx[1] = slot_index<2>(ty, tx - 1);
            sidx[2] = slot_index<2>(ty - 1, tx); sidx[3] = slot_index<2>(ty, tx);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j = sidx[r];
            if (act[j]) {
                if (r < 3) {
                    bend += v3(corr[3 * r][j], corr[3 * r + 1][j], corr[3 * r + 2][j]);
                } else {
                    bend -= v3(corr[0][j], corr[1][j], corr[2][j]) + v3(corr[3][j], corr[4][j], corr[5][j]) +
                            v3(corr[6][j], corr[7][j], corr[8][j]);
                }
                cnt++;
            }
        }
    }
    __syncthreads();   // the next type reuses corr / act
}

__device__ __forceinline__ float4 nan4() {
    const float q = __int_as_float(0x7FC00000);
    return make_float4(q, q, q, q);
}

// SPLIT = 1: one thread per vertex. SPLIT = VTX_LRA_PARTS: a workgroup of four threads per
// vertex for segments with many dense attachments (a 256 x 256 cloth with 256 slots gives one
// tile per CU, and a single wave per SIMD would walk 256 slots per vertex): thread q of a vertex
// sums attachment part q, the first thread adds the parts in order, and all threads share the
// bending slots (one round per type). Both forms add the same terms in the same order.
// The split form (one workgroup per CU on a large cloth, so little hides a load's latency) also
// stages the tile's bending rest angles in LDS with the positions instead of loading one per slot
// evaluation.
template <int EPI, int LB, int SPLIT>
__global__ __launch_bounds__(VTX_TILE * SPLIT) void k_iterate_grid(DevState d, const GridTileDev* __restrict__ tiles,
                                                                   int in, float h) {
    constexpr int BT = VTX_TILE * SPLIT;
    __shared__ float4 sp[VTX_GP * VTX_GP];
    __shared__ float corr[9][VTX_NSMAX];           // plane r*3+c: component c of role r's g_r
    __shared__ uint32_t actw[(VTX_NSX + VTX_NSR + VTX_NSD + 3) / 4];   // per type: slot contributes
    constexpr bool PRE = SPLIT > 1;
    __shared__ float4 srest_s[PRE ? VTX_RC * VTX_RC : 1];
    uint8_t* const act = reinterpret_cast<uint8_t*>(actw);
    const GridTileDev t = tiles[tile_of(d.remap)];
    const SegDev sg = d.segs[t.seg];
    const int rows = (int)sg.rows, cols = (int)sg.cols;
    const int a0 = (int)t.a0, b0 = (int)t.b0;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = SPLIT == 1 ? tid : tid % VTX_TILE;
    // part of the split form: wave-uniform (VTX_TILE is a multiple of the wave), so the slot
    // positions below are scalar loads
    const uint32_t qpart = SPLIT == 1 ? 0u : __builtin_amdgcn_readfirstlane(tid / VTX_TILE);
    const float4* __restrict__ P = d.P[in];
    const int ty = (int)(lane / VTX_GT), tx = (int)(lane % VTX_GT);
    const int a = a0 + ty, b = b0 + tx;
    const bool in_cloth = a < rows && b < cols;
    const bool inside = in_cloth && qpart == 0;        // the thread that owns the vertex
    const uint32_t g = in_cloth ? sg.pbase + (uint32_t)(a * cols + b) : 0u;
    // stretch rest lengths: R own/left, D own/up, X-diagonal own/up-left, anti-diagonal of left/up,
    // issued before the staging loads and always (from the vertex itself where a neighbour is
    // missing): a load that a lane may skip leaves the outstanding count unknown to the compiler,
    // which then waits for every load before the first use
    const float4 r_own = d.rest4[g];
    float4 r_left = d.rest4[inside && b > 0 ? g - 1 : g];
    float4 r_up = d.rest4[inside && a > 0 ? g - cols : g];
    float4 r_ul = d.rest4[inside && a > 0 && b > 0 ? g - cols - 1 : g];
    if (!(b > 0)) r_left = nan4();
    if (!(a > 0)) r_up = nan4();
    if (!(a > 0 && b > 0)) r_ul = nan4();

    // clear the slot marks; stage positions of rows [a0-2, a0+18) x cols [b0-2, b0+18)
    for (uint32_t e = tid; e < (VTX_NSX + VTX_NSR + VTX_NSD + 3) / 4; e += BT) actw[e] = 0u;
    for (uint32_t e = tid; e < VTX_GP * VTX_GP; e += BT) {
        const int sa = a0 - VTX_GH + (int)(e / VTX_GP), sb = b0 - VTX_GH + (int)(e % VTX_GP);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (sa >= 0 && sa < rows && sb >= 0 && sb < cols) v = P[sg.pbase + (uint32_t)(sa * cols + sb)];
        sp[e] = v;
    }
    if (PRE && !(sg.flags & VTX_SEG_BEND_REST0)) {   // rest angles of owner cells rows/cols -1..16
        for (uint32_t e = tid; e < VTX_RC * VTX_RC; e += BT) {
            const int sa = a0 - 1 + (int)(e / VTX_RC), sb = b0 - 1 + (int)(e % VTX_RC);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sa >= 0 && sa < rows && sb >= 0 && sb < cols) v = d.brest4[sg.pbase + (uint32_t)(sa * cols + sb)];
            srest_s[e] = v;
        }
    }

    __syncthreads();

    const int sc = (ty + VTX_GH) * VTX_GP + (tx + VTX_GH);    // staged index of this vertex
    const float4 p4 = sp[sc];
    V3 p = xyz(p4);
    const float w = p4.w;
    V3 delta = v3(0.f, 0.f, 0.f);
    int cnt = 0;

    if (inside) {
        const float k = -w * sg.stiffness;
        if (r_own.x == r_own.x) stretch_term(p, w, k, sp[sc + 1], r_own.x, delta, cnt);
        if (r_left.x == r_left.x) stretch_term(p, w, k, sp[sc - 1], r_left.x, delta, cnt);
        if (r_own.y == r_own.y) stretch_term(p, w, k, sp[sc + VTX_GP], r_own.y, delta, cnt);
        if (r_up.y == r_up.y) stretch_term(p, w, k, sp[sc - VTX_GP], r_up.y, delta, cnt);
        if (r_own.z == r_own.z) stretch_term(p, w, k, sp[sc + VTX_GP + 1], r_own.z, delta, cnt);
        if (r_ul.z == r_ul.z) stretch_term(p, w, k, sp[sc - VTX_GP - 1], r_ul.z, delta, cnt);
        if (r_left.w == r_left.w) stretch_term(p, w, k, sp[sc + VTX_GP - 1], r_left.w, delta, cnt);
        if (r_up.w == r_up.w) stretch_term(p, w, k, sp[sc - VTX_GP + 1], r_up.w, delta, cnt);
        if constexpr (SPLIT == 1) attach_terms<LB>(d, sg, t.at_off, t.at_k, tid, (uint32_t)(a * cols + b), p, w, delta, cnt);
    }
    if constexpr (SPLIT > 1) {   // only dense segments come here (VTX_SEG_LRA_DENSE)
        __shared__ float4 lpart[SPLIT - 1][VTX_TILE];
        V3 part = v3(0.f, 0.f, 0.f);
        int pc = 0;
        if (in_cloth) {
            uint32_t s0, s1;
            lra_part_range(sg, qpart, s0, s1);
            attach_dense<LB>(d, sg, g - sg.pbase, s0, s1, p, w, part, pc);
            if (qpart > 0) lpart[qpart - 1][lane] = f4(part, __int_as_float(pc));
        }
        __syncthreads();
        if (inside) {
            delta += part;
            cnt += pc;
#pragma unroll
            for (int q = 0; q < SPLIT - 1; q++) {
                const float4 e = lpart[q][lane];
                delta += xyz(e);
                cnt += __float_as_int(e.w);
            }
        }
    }
    // bending: X roles 0..3, R roles 0..3, D roles 0..3
    const float xpbd_bend = sg.bend_compliance * frcp(h) * frcp(h);   // compliance / h^2 (cu:173)
    V3 bend = v3(0.f, 0.f, 0.f);
    const float4* const srest = srest_s;   // read by the split form (PRE) only, when staged
    bend_type_pass<0, BT, PRE>(d, sg, sp, srest, corr, act, a0, b0, rows, cols, xpbd_bend, inside, ty, tx, bend, cnt);
    bend_type_pass<1, BT, PRE>(d, sg, sp, srest, corr, act + VTX_NSX, a0, b0, rows, cols, xpbd_bend, inside, ty, tx, bend,
                          cnt);
    bend_type_pass<2, BT, PRE>(d, sg, sp, srest, corr, act + VTX_NSX + VTX_NSR, a0, b0, rows, cols, xpbd_bend, inside, ty,
                          tx, bend, cnt);
    if (!inside) return;
    delta += w * bend;
    if (cnt > 0) p += delta * (frcp((float)cnt) * sg.omega);
    epilogue<EPI>(d, sg, g, in, p, w, h);
}

// ComputeTriangleNormals + ComputeVertexNormals (Velvet/VtClothSolverGPU.cu:419-450) as a
// per-vertex gather over incident triangles (segments off the grid path).
__global__ __launch_bounds__(VTX_TILE) void k_normals(DevState d, const TileDev* __restrict__ tiles) {
    const TileDev t = tiles[tile_of(d.remap)];
    const uint32_t lane = threadIdx.x;
    if (lane >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + lane;
    V3 s = v3(0.f, 0.f, 0.f);
    const uint32_t* __restrict__ tr = d.tr + t.tr_off + lane;
    for (uint32_t k = 0; k < t.tr_k; k++) {
        const uint32_t e = tr[(size_t)k * VTX_TILE];
        if (e == VTX_INVALID) continue;
        const uint4 ix = d.trec[sg.trbase + e];
        const uint32_t pb = sg.pbase;
        const V3 a = xyz(d.X[pb + ix.x]), b = xyz(d.X[pb + ix.y]), c = xyz(d.X[pb + ix.z]);
        s += cross(b - a, c - a);
    }
    d.Nrm[g] = f4(normalize(s), 0.f);
}

// The same for a grid segment's 16 x 16 tile: positions of the tile and a one-vertex halo staged
// in LDS, one read per vertex instead of the incidence gather. Quad (a, b) holds triangles
// (A, B, C) and (C, B, D) with A = (a, b), B = (a+1, b), C = (a, b+1), D = (a+1, b+1) (the grid
// the stencil path accepts), so a vertex's incident triangles in triangle-index order are T1 of
// quad (r-1, c-1), T0 and T1 of (r-1, c), T0 and T1 of (r, c-1), T0 of (r, c), each crossed as
// cross(b - a, c - a) in its stored vertex order: the terms and the order of the general gather.
__global__ __launch_bounds__(VTX_TILE) void k_normals_grid(DevState d, const GridTileDev* __restrict__ tiles) {
    constexpr int GS = VTX_GT + 2;                 // staged side: tile + one-vertex halo
    __shared__ float4 sx[GS * GS];
    const GridTileDev t = tiles[tile_of(d.remap)];
    const SegDev sg = d.segs[t.seg];
    const int rows = (int)sg.rows, cols = (int)sg.cols, a0 = (int)t.a0, b0 = (int)t.b0;
    for (uint32_t e = threadIdx.x; e < GS * GS; e += VTX_TILE) {
        const int a = a0 - 1 + (int)(e / GS), b = b0 - 1 + (int)(e % GS);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (a >= 0 && a < rows && b >= 0 && b < cols) v = d.X[sg.pbase + (uint32_t)(a * cols + b)];
        sx[e] = v;
    }
    __syncthreads();
    const int ty = (int)(threadIdx.x / VTX_GT), tx = (int)(threadIdx.x % VTX_GT);
    const int r = a0 + ty, c = b0 + tx;
    if (r >= rows || c >= cols) return;
    auto at = [&](int dr, int dc) { return xyz(sx[(ty + 1 + dr) * GS + (tx + 1 + dc)]); };
    const V3 v = at(0, 0);
    V3 s = v3(0.f, 0.f, 0.f);
    if (r >= 1 && c >= 1) {                        // quad (r-1, c-1): v = D
        const V3 B = at(0, -1), C = at(-1, 0);
        s += cross(B - C, v - C);
    }
    if (r >= 1 && c + 1 < cols) {                  // quad (r-1, c): v = B
        const V3 A = at(-1, 0), C = at(-1, 1), D = at(0, 1);
        s += cross(v - A, C - A);
        s += cross(v - C, D - C);
    }
    if (r + 1 < rows && c >= 1) {                  // quad (r, c-1): v = C
        const V3 A = at(0, -1), B = at(1, -1), D = at(1, 0);
        s += cross(B - A, v - A);
        s += cross(B - v, D - v);
    }
    if (r + 1 < rows && c + 1 < cols) {            // quad (r, c): v = A
        const V3 B = at(1, 0), C = at(0, 1);
        s += cross(B - v, C - v);
    }
    d.Nrm[sg.pbase + (uint32_t)(r * cols + c)] = f4(normalize(s), 0.f);
}

// Failure detection: flags[t.seg] = 1 when a particle of the segment has a non-finite position
// or velocity (the reference has no such check; its only net is the speed clamp, cu:399-404).
__global__ void k_clear_flags(uint32_t* flags, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = 0;
}

__global__ __launch_bounds__(VTX_TILE) void k_check_finite(DevState d, uint32_t* flags) {
    const TileDev t = d.tiles[tile_of(d.remap)];
    bool bad = false;
    if (threadIdx.x < t.count) {
        const uint32_t g = t.begin + threadIdx.x;
        const float4 x = d.X[g], v = d.V[g];
        bad = !(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z));
    }
    if (__any(bad) && (threadIdx.x & 63u) == 0) flags[t.seg] = 1u;   // same value from any wave: no race on the result
}

// ---------------------------------------------------------------- host launchers
static inline hipError_t last_launch() { return hipGetLastError(); }

hipError_t launch_frame_begin(const DevState& d, float frame_dt, float h, hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    hipLaunchKernelGGL(k_frame_begin, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, frame_dt, h);
    return last_launch();
}

hipError_t launch_collide(const DevState& d, int in, float h, int block, hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    if (block == 64) hipLaunchKernelGGL(k_collide<64>, dim3(d.ntiles * 4), dim3(64), 0, s, d, in, h);
    else if (block == 128) hipLaunchKernelGGL(k_collide<128>, dim3(d.ntiles * 2), dim3(128), 0, s, d, in, h);
    else hipLaunchKernelGGL(k_collide<256>, dim3(d.ntiles), dim3(256), 0, s, d, in, h);
    return last_launch();
}

template <int LB, int SPLIT>
static void launch_iterate_grid(const DevState& d, const GridTileDev* tiles, uint32_t ntiles, int in, float h,
                                int epilogue, hipStream_t s) {
    if (!ntiles) return;
    const dim3 gg(ntiles), bl(VTX_TILE * SPLIT);
    if (epilogue == EPI_NONE) hipLaunchKernelGGL((k_iterate_grid<EPI_NONE, LB, SPLIT>), gg, bl, 0, s, d, tiles, in, h);
    else if (epilogue == EPI_FINALIZE) hipLaunchKernelGGL((k_iterate_grid<EPI_FINALIZE, LB, SPLIT>), gg, bl, 0, s, d, tiles, in, h);
    else hipLaunchKernelGGL((k_iterate_grid<EPI_FINALIZE_PREDICT, LB, SPLIT>), gg, bl, 0, s, d, tiles, in, h);
}

template <int LB>
static void launch_iterate_lb(const DevState& d, int in, float h, int epilogue, hipStream_t s) {
    dim3 bl(VTX_TILE);
    launch_iterate_grid<LB, 1>(d, d.gtiles, d.ngtiles, in, h, epilogue, s);
    // four threads per particle leave 128 registers per thread: batches of 64 would spill
    launch_iterate_grid<(LB > 16 ? 16 : LB), VTX_LRA_PARTS>(d, d.ltiles, d.nltiles, in, h, epilogue, s);
    if (d.nitiles) {
        dim3 gr(d.nitiles);
        if (epilogue == EPI_NONE) hipLaunchKernelGGL((k_iterate<EPI_NONE, true, LB>), gr, bl, 0, s, d, d.itiles, in, h);
        else if (epilogue == EPI_FINALIZE) hipLaunchKernelGGL((k_iterate<EPI_FINALIZE, true, LB>), gr, bl, 0, s, d, d.itiles, in, h);
        else hipLaunchKernelGGL((k_iterate<EPI_FINALIZE_PREDICT, true, LB>), gr, bl, 0, s, d, d.itiles, in, h);
    }
}

hipError_t launch_iterate(const DevState& d, int in, float h, int epilogue, bool constraints, int lra_batch,
                          hipStream_t s) {
    if (!constraints) {   // no Jacobi pass: only the epilogue, over every particle
        if (!d.ntiles) return hipSuccess;
        dim3 gr(d.ntiles), bl(VTX_TILE);
        if (epilogue == EPI_NONE) hipLaunchKernelGGL((k_iterate<EPI_NONE, false, 1>), gr, bl, 0, s, d, d.tiles, in, h);
        else if (epilogue == EPI_FINALIZE) hipLaunchKernelGGL((k_iterate<EPI_FINALIZE, false, 1>), gr, bl, 0, s, d, d.tiles, in, h);
        else hipLaunchKernelGGL((k_iterate<EPI_FINALIZE_PREDICT, false, 1>), gr, bl, 0, s, d, d.tiles, in, h);
        return last_launch();
    }
    if (lra_batch >= 64) launch_iterate_lb<64>(d, in, h, epilogue, s);
    else if (lra_batch >= 16) launch_iterate_lb<16>(d, in, h, epilogue, s);
    else launch_iterate_lb<4>(d, in, h, epilogue, s);
    return last_launch();
}

hipError_t launch_check_finite(const DevState& d, uint32_t* flags, uint32_t nseg, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_flags, dim3((nseg + 255) / 256), dim3(256), 0, s, flags, nseg);
    if (d.ntiles) hipLaunchKernelGGL(k_check_finite, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, flags);
    return last_launch();
}

hipError_t launch_normals(const DevState& d, hipStream_t s) {
    if (d.nitiles) hipLaunchKernelGGL(k_normals, dim3(d.nitiles), dim3(VTX_TILE), 0, s, d, d.itiles);
    if (d.ngtiles) hipLaunchKernelGGL(k_normals_grid, dim3(d.ngtiles), dim3(VTX_TILE), 0, s, d, d.gtiles);
    if (d.nltiles) hipLaunchKernelGGL(k_normals_grid, dim3(d.nltiles), dim3(VTX_TILE), 0, s, d, d.ltiles);
    return last_launch();
}

}  // namespace vtx
