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

// ComputeFriction (Velvet/VtClothSolverGPU.cu:272-287).
__device__ __forceinline__ V3 friction_of(V3 correction, V3 rel, float friction) {
    V3 f = v3(0.f, 0.f, 0.f);
    float cl = length(correction);
    if (friction > 0.f && cl > 0.f) {
        V3 nrm = correction / cl;
        V3 tan = rel - nrm * dot(rel, nrm);
        float tl = length(tan);
        float mt = cl * friction;
        f = -tan * fminf(mt / tl, 1.0f);
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
        float distance = length(diff);
        float offset = distance - radius;
        if (offset < 0.f) {
            V3 direction = diff / distance;
            return -offset * direction;
        }
    } else if (c.type == 2) {
        V3 correction = v3(0.f, 0.f, 0.f);
        V3 local = xform_point(c.inv, p);
        V3 size = v3(0.5f + margin / c.scale[0], 0.5f + margin / c.scale[1], 0.5f + margin / c.scale[2]);
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
                float l = length(vec);
                if (l < m) {
                    V3 nv = normalize(vec);
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
            V3 vel = (pred - lastp) / col.dt;
            V3 relv = pred - pos - vel * dt;
            pred += friction_of(corr, relv, friction);
        }
    }
    return pred;
}

// Pre-stabilisation SDF on positions with the frame time (Velvet/VtClothSolverGPU.hpp:73) and
// the first substep's PredictPositions (cu:42-53).
__global__ __launch_bounds__(VTX_TILE) void k_frame_begin(DevState d, float frame_dt, float h) {
    const TileDev t = d.tiles[blockIdx.x];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + threadIdx.x;
    float4 x4 = d.X[g];
    V3 x = xyz(x4);
    if (sg.ccount) {
        x = collide_sdf(x, x, d.cols, sg.cbase, sg.ccount, sg.margin, sg.friction, frame_dt);
        d.X[g] = f4(x, x4.w);
    }
    V3 v = xyz(d.V[g]) + v3(sg.gx, sg.gy, sg.gz) * h;
    d.V[g] = f4(v, 0.f);
    d.P[0][g] = f4(x + v * h, x4.w);
}

// CollideParticles_Kernel + ApplyDeltas (Velvet/VtClothSolverGPU.cu:329-386) followed by the
// substep's CollideSDF (cu:289-314, hpp:87).
__global__ __launch_bounds__(VTX_TILE) void k_collide(DevState d, int in, float h) {
    const TileDev t = d.tiles[blockIdx.x];
    if (threadIdx.x >= t.count) return;
    const SegDev sg = d.segs[t.seg];
    const uint32_t g = t.begin + threadIdx.x;
    const float4* __restrict__ P = d.P[in];
    const float4 p4 = P[g];
    V3 p = xyz(p4);
    const float w = p4.w;
    const V3 x = xyz(d.X[g]);
    if (sg.flags & VTX_SEG_SELFCOL) {
        const uint32_t n = sg.n;
        const uint32_t* __restrict__ nb =
            d.nbr + ((uint64_t)sg.nbase_hi << 32 | sg.nbase_lo) + (g - sg.pbase);
        V3 delta = v3(0.f, 0.f, 0.f);
        int cnt = 0;
        const V3 vel_i = p - x;
        for (uint32_t k = 0; k < sg.max_nbr; k++) {
            const uint32_t j = nb[(size_t)k * n];
            if (j >= n) break;
            const uint32_t jg = sg.pbase + j;
            const float4 pj4 = P[jg];
            const float denom = w + pj4.w;
            if (denom <= 0.f) continue;
            const V3 pj = xyz(pj4);
            const V3 diff = p - pj;
            const float dist = length(diff);
            if (dist >= sg.diameter) continue;
            const V3 grad = diff / (dist + VTX_EPSILON);
            const float lambda = (dist - sg.diameter) / denom;
            const V3 common = lambda * grad;
            cnt++;
            delta -= w * common;
            const V3 relv = vel_i - (pj - xyz(d.X[jg]));
            delta += w * friction_of(common, relv, sg.friction);
        }
        if (cnt > 0) p += delta / (float)cnt * sg.omega;
    }
    if (sg.ccount) p = collide_sdf(p, x, d.cols, sg.cbase, sg.ccount, sg.margin, sg.friction, h);
    d.P[in ^ 1][g] = f4(p, w);
}

// Epilogue modes of the Jacobi pass.
enum { EPI_NONE = 0, EPI_FINALIZE = 1, EPI_FINALIZE_PREDICT = 2 };

// One Jacobi pass: SolveStretch (cu:65-102) + SolveAttachment (cu:205-235) +
// SolveBending (cu:117-189) gathered per particle, then ApplyDeltas (cu:253-264); with an
// epilogue, Finalize (cu:388-407) and optionally the next substep's Predict (cu:42-53).
template <int EPI, bool CONSTRAINTS>
__global__ __launch_bounds__(VTX_TILE) void k_iterate(DevState d, int in, float h) {
    const TileDev t = d.tiles[blockIdx.x];
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
        // stretch: bilateral distance constraint, compliance 0 (cu:85-98)
        const uint2* __restrict__ st = d.st + t.st_off + lane;
        for (uint32_t k = 0; k < t.st_k; k++) {
            const uint2 e = st[(size_t)k * VTX_TILE];
            if (e.x == VTX_INVALID) continue;
            const float4 pj4 = P[e.x];
            const float rest = __uint_as_float(e.y);
            const V3 diff = p - xyz(pj4);
            const float dist = length(diff);
            const float wsum = w + pj4.w;
            if (dist != rest && wsum > 0.f) {
                const V3 grad = diff / (dist + VTX_EPSILON);
                const float lambda = (dist - rest) / wsum * sg.stiffness;
                delta += -w * (lambda * grad);
                cnt++;
            }
        }
        // long-range attachment: one-sided, pinned particles skip non-zero targets (cu:218-234)
        const uint2* __restrict__ at = d.at + t.at_off + lane;
        for (uint32_t k = 0; k < t.at_k; k++) {
            const uint2 e = at[(size_t)k * VTX_TILE];
            if (e.x == VTX_INVALID) continue;
            const float target = __uint_as_float(e.y) * sg.lrs;
            if (w == 0.f && target > 0.f) continue;
            const V3 diff = p - xyz(d.slots[e.x]);
            const float dist = length(diff);
            if (dist > target) {
                delta += -diff + diff / dist * target;
                cnt++;
            }
        }
        // dihedral bending (cu:128-188); this particle takes the correction of its role
        const uint32_t* __restrict__ bd = d.bd + t.bd_off + lane;
        const float xpbd_bend = sg.bend_compliance / h / h;
        for (uint32_t k = 0; k < t.bd_k; k++) {
            const uint32_t e = bd[(size_t)k * VTX_TILE];
            if (e == VTX_INVALID) continue;
            const uint32_t bid = e & 0x3FFFFFFFu;
            const uint32_t role = e >> 30;
            const uint4 ix = d.brec[bid];
            const float rest = d.brest[bid];
            const float4 q0 = P[ix.x], q1 = P[ix.y], q2 = P[ix.z], q3 = P[ix.w];
            const V3 p0 = xyz(q0), p1 = xyz(q1), p2 = xyz(q2), p3 = xyz(q3);
            const V3 e3 = p3 - p2;
            const float elen = length(e3);
            if (elen < VTX_EPSILON) continue;
            const float inv_elen = 1.0f / elen;
            V3 n1 = cross(p2 - p0, p3 - p0);
            n1 = n1 / dot(n1, n1);
            V3 n2 = cross(p3 - p1, p2 - p1);
            n2 = n2 / dot(n2, n2);
            const V3 d0 = elen * n1;
            const V3 d1 = elen * n2;
            const V3 d2 = dot(p0 - p3, e3) * inv_elen * n1 + dot(p1 - p3, e3) * inv_elen * n2;
            const V3 d3 = dot(p2 - p0, e3) * inv_elen * n1 + dot(p2 - p1, e3) * inv_elen * n2;
            n1 = normalize(n1);
            n2 = normalize(n2);
            const float cd = fminf(fmaxf(dot(n1, n2), -1.0f), 1.0f);
            const float phi = acosf(cd);
            float lambda = q0.w * dot(d0, d0) + q1.w * dot(d1, d1) + q2.w * dot(d2, d2) + q3.w * dot(d3, d3);
            if (lambda < VTX_EPSILON) continue;
            lambda = (phi - rest) / (lambda + xpbd_bend);
            if (dot(cross(n1, n2), e3) > 0.0f) lambda = -lambda;
            const V3 dk = role == 0 ? d0 : role == 1 ? d1 : role == 2 ? d2 : d3;
            delta += -w * lambda * dk;
            cnt++;
        }
        if (cnt > 0) p += delta / (float)cnt * sg.omega;
    }

    if (EPI == EPI_NONE) {
        d.P[in ^ 1][g] = f4(p, w);
    } else {
        // Finalize (cu:396-406)
        const float4 x4 = d.X[g];
        const V3 x = xyz(x4);
        V3 raw = (p - x) / h;
        const float rl = length(raw);
        if (rl > sg.max_speed) {
            raw = raw / rl * sg.max_speed;
            p = x + raw * h;
        }
        V3 v = raw * (1.0f - sg.damping * h);
        d.X[g] = f4(p, x4.w);
        if (EPI == EPI_FINALIZE_PREDICT) {
            v = v + v3(sg.gx, sg.gy, sg.gz) * h;
            d.P[in ^ 1][g] = f4(p + v * h, x4.w);
        }
        d.V[g] = f4(v, 0.f);
    }
}

// ComputeTriangleNormals + ComputeVertexNormals (Velvet/VtClothSolverGPU.cu:419-450) as a
// per-vertex gather over incident triangles.
__global__ __launch_bounds__(VTX_TILE) void k_normals(DevState d) {
    const TileDev t = d.tiles[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    if (lane >= t.count) return;
    const uint32_t g = t.begin + lane;
    V3 s = v3(0.f, 0.f, 0.f);
    const uint32_t* __restrict__ tr = d.tr + t.tr_off + lane;
    for (uint32_t k = 0; k < t.tr_k; k++) {
        const uint32_t e = tr[(size_t)k * VTX_TILE];
        if (e == VTX_INVALID) continue;
        const uint4 ix = d.trec[e];
        const V3 a = xyz(d.X[ix.x]), b = xyz(d.X[ix.y]), c = xyz(d.X[ix.z]);
        s += cross(b - a, c - a);
    }
    d.Nrm[g] = f4(normalize(s), 0.f);
}

// ---------------------------------------------------------------- host launchers
static inline hipError_t last_launch() { return hipGetLastError(); }

hipError_t launch_frame_begin(const DevState& d, float frame_dt, float h, hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    hipLaunchKernelGGL(k_frame_begin, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, frame_dt, h);
    return last_launch();
}

hipError_t launch_collide(const DevState& d, int in, float h, hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    hipLaunchKernelGGL(k_collide, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d, in, h);
    return last_launch();
}

hipError_t launch_iterate(const DevState& d, int in, float h, int epilogue, bool constraints,
                          hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    dim3 gr(d.ntiles), bl(VTX_TILE);
    if (constraints) {
        if (epilogue == EPI_NONE) hipLaunchKernelGGL((k_iterate<EPI_NONE, true>), gr, bl, 0, s, d, in, h);
        else if (epilogue == EPI_FINALIZE) hipLaunchKernelGGL((k_iterate<EPI_FINALIZE, true>), gr, bl, 0, s, d, in, h);
        else hipLaunchKernelGGL((k_iterate<EPI_FINALIZE_PREDICT, true>), gr, bl, 0, s, d, in, h);
    } else {
        if (epilogue == EPI_NONE) hipLaunchKernelGGL((k_iterate<EPI_NONE, false>), gr, bl, 0, s, d, in, h);
        else if (epilogue == EPI_FINALIZE) hipLaunchKernelGGL((k_iterate<EPI_FINALIZE, false>), gr, bl, 0, s, d, in, h);
        else hipLaunchKernelGGL((k_iterate<EPI_FINALIZE_PREDICT, false>), gr, bl, 0, s, d, in, h);
    }
    return last_launch();
}

hipError_t launch_normals(const DevState& d, hipStream_t s) {
    if (!d.ntiles) return hipSuccess;
    hipLaunchKernelGGL(k_normals, dim3(d.ntiles), dim3(VTX_TILE), 0, s, d);
    return last_launch();
}

}  // namespace vtx
