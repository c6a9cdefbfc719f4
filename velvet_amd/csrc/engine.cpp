// Host engine: registration mirrors VtClothSolverGPU (Velvet/VtClothSolverGPU.hpp:114-205),
// commit() packs every segment into the device layout of vtx_device.hpp, simulate() records
// one frame (Velvet/VtClothSolverGPU.hpp:56-111) and replays it as a hipGraph.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>

#include "launch.hpp"

namespace vtx {

static const float kFixedDt = 1.0f / 60.0f;   // Velvet/Timer.hpp:235

Engine::Engine(const vtx_params& p) {
    VTX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
    add_segment(p);
}

Engine::~Engine() {
    if (stream_) hipStreamSynchronize(stream_);
    if (gexec_) hipGraphExecDestroy(gexec_);
    for (auto& st : stages_) {
        hipEventDestroy(st.a);
        hipEventDestroy(st.b);
    }
    free_device();
    if (own_stream_ && stream_) hipStreamDestroy(stream_);
}

void Engine::check_seg(int32_t s) const {
    VTX_REQUIRE(s >= 0 && s < (int32_t)segs_.size(), VTX_EINVAL, "segment index out of range");
}

Segment& Engine::seg(int32_t i) {
    check_seg(i);
    return segs_[(size_t)i];
}

static void validate_params(const vtx_params& p) {
    VTX_REQUIRE(p.num_substeps >= 1, VTX_EINVAL, "num_substeps must be >= 1");
    VTX_REQUIRE(p.num_iterations >= 0, VTX_EINVAL, "num_iterations must be >= 0");
    VTX_REQUIRE(p.max_num_neighbors >= 1 && p.max_num_neighbors <= VTX_MAX_NEIGHBORS, VTX_EINVAL,
                "max_num_neighbors must be in [1, 64]");
    VTX_REQUIRE(p.interleaved_hash >= 1, VTX_EINVAL, "interleaved_hash must be >= 1");
    VTX_REQUIRE(p.hash_table_size >= 0, VTX_EINVAL, "hash_table_size must be >= 0");
}

int32_t Engine::add_segment(const vtx_params& p) {
    validate_params(p);
    pull_state();
    Segment s;
    s.prm = p;
    segs_.push_back(std::move(s));
    mark_dirty();
    return (int32_t)segs_.size() - 1;
}

void Engine::set_params(int32_t s, const vtx_params& p) {
    validate_params(p);
    pull_state();
    seg(s).prm = p;
    mark_dirty();
}

// Device -> host copy of the particle state before a structural change.
void Engine::pull_state() {
    if (!dev_live_ || dirty_) return;
    VTX_HIP(hipStreamSynchronize(stream_));
    for (size_t k = 0; k < segs_.size(); k++) {
        Segment& sg = segs_[k];
        const uint32_t n = sg.n();
        if (!n) continue;
        std::vector<float4> tmp(n);
        const uint32_t base = hseg_[k].pbase;
        VTX_HIP(hipMemcpy(tmp.data(), ds_.X + base, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) {
            sg.x[3 * i] = tmp[i].x;
            sg.x[3 * i + 1] = tmp[i].y;
            sg.x[3 * i + 2] = tmp[i].z;
            sg.w[i] = tmp[i].w;
        }
        VTX_HIP(hipMemcpy(tmp.data(), ds_.V + base, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) {
            sg.v[3 * i] = tmp[i].x;
            sg.v[3 * i + 1] = tmp[i].y;
            sg.v[3 * i + 2] = tmp[i].z;
        }
    }
}

void Engine::mark_dirty() {
    dirty_ = true;
    if (gexec_) {
        hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
}

// AddCloth (Velvet/VtClothSolverGPU.hpp:114-156).
int32_t Engine::add_cloth(int32_t s, const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                          const float* model, float diameter) {
    Segment& sg = seg(s);
    VTX_REQUIRE(n == 0 || xyz, VTX_EINVAL, "positions are NULL");
    VTX_REQUIRE(ntri == 0 || tris, VTX_EINVAL, "triangles are NULL");
    for (uint32_t t = 0; t < 3 * ntri; t++) VTX_REQUIRE(tris[t] < n, VTX_EINVAL, "triangle index out of range");
    pull_state();
    const uint32_t prev = sg.n();
    sg.prm.particle_diameter = diameter;
    sg.prm.max_speed = 2.0f * diameter / kFixedDt * (float)sg.prm.num_substeps;
    for (uint32_t t = 0; t < 3 * ntri; t++) sg.tris.push_back(tris[t] + prev);
    for (uint32_t i = 0; i < n; i++) {
        float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
        if (model) {   // InitializePositions_Kernel: x = M * vec4(x, 1) (Velvet/VtClothSolverGPU.cu:30-34)
            const float* m = model;
            float ox = m[0] * px + m[4] * py + m[8] * pz + m[12];
            float oy = m[1] * px + m[5] * py + m[9] * pz + m[13];
            float oz = m[2] * px + m[6] * py + m[10] * pz + m[14];
            px = ox;
            py = oy;
            pz = oz;
        }
        sg.x.push_back(px);
        sg.x.push_back(py);
        sg.x.push_back(pz);
        sg.v.insert(sg.v.end(), {0.f, 0.f, 0.f});
        sg.w.push_back(1.0f);
    }
    sg.cloth_starts.push_back(prev);
    sg.x0 = sg.x;   // SetInitialPositions snapshot of every particle (Velvet/SpatialHashGPU.hpp:32-39)
    mark_dirty();
    return (int32_t)prev;
}

void Engine::add_stretch(int32_t s, const int32_t* ij, const float* rest, uint32_t count) {
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (ij && rest), VTX_EINVAL, "stretch arrays are NULL");
    const int32_t n = (int32_t)sg.n();
    for (uint32_t c = 0; c < 2 * count; c++)
        VTX_REQUIRE(ij[c] >= 0 && ij[c] < n, VTX_EINVAL, "stretch index out of range");
    pull_state();
    sg.st.insert(sg.st.end(), ij, ij + 2 * count);
    sg.st_rest.insert(sg.st_rest.end(), rest, rest + count);
    mark_dirty();
}

int32_t Engine::add_attach_slot(int32_t s, const float* pos) {
    Segment& sg = seg(s);
    VTX_REQUIRE(pos, VTX_EINVAL, "slot position is NULL");
    pull_state();
    sg.slots.insert(sg.slots.end(), pos, pos + 3);
    mark_dirty();
    return (int32_t)(sg.slots.size() / 3 - 1);
}

void Engine::set_attach_slot_position(int32_t s, int32_t slot, const float* pos) {
    Segment& sg = seg(s);
    VTX_REQUIRE(pos, VTX_EINVAL, "slot position is NULL");
    VTX_REQUIRE(slot >= 0 && (size_t)slot < sg.slots.size() / 3, VTX_EINVAL, "slot index out of range");
    std::copy(pos, pos + 3, sg.slots.begin() + 3 * slot);
    if (!dirty_) {
        float4 v = make_float4(pos[0], pos[1], pos[2], 0.f);
        VTX_HIP(hipMemcpyAsync((float4*)ds_.slots + hseg_[(size_t)s].sbase + slot, &v, sizeof(float4),
                               hipMemcpyHostToDevice, stream_));
        VTX_HIP(hipStreamSynchronize(stream_));
    }
}

// AddAttach (Velvet/VtClothSolverGPU.hpp:170-176): distance 0 pins the particle.
void Engine::add_attach(int32_t s, const int32_t* pid, const int32_t* slot, const float* dist, uint32_t count) {
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (pid && slot && dist), VTX_EINVAL, "attach arrays are NULL");
    const int32_t n = (int32_t)sg.n();
    const int32_t L = (int32_t)(sg.slots.size() / 3);
    for (uint32_t c = 0; c < count; c++) {
        VTX_REQUIRE(pid[c] >= 0 && pid[c] < n, VTX_EINVAL, "attach particle out of range");
        VTX_REQUIRE(slot[c] >= 0 && slot[c] < L, VTX_EINVAL, "attach slot out of range");
    }
    pull_state();
    for (uint32_t c = 0; c < count; c++) {
        if (dist[c] == 0.0f) sg.w[(size_t)pid[c]] = 0.0f;
        sg.at_pid.push_back(pid[c]);
        sg.at_slot.push_back(slot[c]);
        sg.at_dist.push_back(dist[c]);
    }
    mark_dirty();
}

void Engine::add_bend(int32_t s, const uint32_t* idx, const float* angle, uint32_t count) {
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (idx && angle), VTX_EINVAL, "bend arrays are NULL");
    VTX_REQUIRE(sg.bd.size() / 4 + count < (1u << 28), VTX_ECAPACITY, "too many bend constraints");
    for (uint32_t c = 0; c < 4 * count; c++) VTX_REQUIRE(idx[c] < sg.n(), VTX_EINVAL, "bend index out of range");
    pull_state();
    sg.bd.insert(sg.bd.end(), idx, idx + 4 * count);
    sg.bd_rest.insert(sg.bd_rest.end(), angle, angle + count);
    mark_dirty();
}

void Engine::set_colliders(int32_t s, const vtx_collider* cols, uint32_t n) {
    Segment& sg = seg(s);
    VTX_REQUIRE(n == 0 || cols, VTX_EINVAL, "colliders are NULL");
    for (uint32_t c = 0; c < n; c++)
        VTX_REQUIRE(cols[c].type >= 0 && cols[c].type <= 2, VTX_EINVAL, "unknown collider type");
    const bool same_count = sg.cols.size() == n;
    sg.cols.assign(cols, cols + n);
    if (same_count && !dirty_) {
        upload_colliders();   // the graph reads the collider table; no re-record needed
    } else {
        pull_state();
        mark_dirty();
    }
}

// ---------------------------------------------------------------------------- colliders
static void inverse4(const double* m, double* inv) {   // column-major Gauss-Jordan
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
        double d = a[c][c];
        if (d == 0.0) d = 1e-300;
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

static ColliderDev pack_collider(const vtx_collider& c) {
    ColliderDev d{};
    d.type = c.type;
    d.dt = c.delta_time;
    for (int k = 0; k < 3; k++) {
        d.pos[k] = c.position[k];
        d.scale[k] = c.scale[k];
    }
    double cur[16], last[16], inv[16], rel[16];
    for (int k = 0; k < 16; k++) {
        cur[k] = c.cur[k];
        last[k] = c.last[k];
    }
    inverse4(cur, inv);
    for (int col = 0; col < 4; col++)
        for (int r = 0; r < 4; r++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += last[k * 4 + r] * inv[col * 4 + k];
            rel[col * 4 + r] = s;
        }
    for (int r = 0; r < 3; r++)
        for (int col = 0; col < 4; col++) {
            d.cur3[r * 4 + col] = col < 3 ? (float)cur[col * 4 + r] : 0.f;
            d.inv[r * 4 + col] = (float)inv[col * 4 + r];
            d.rel[r * 4 + col] = (float)rel[col * 4 + r];
        }
    return d;
}

void Engine::upload_colliders() {
    std::vector<ColliderDev> h;
    for (auto& sg : segs_)
        for (auto& c : sg.cols) h.push_back(pack_collider(c));
    if (h.empty()) return;
    VTX_HIP(hipMemcpyAsync((void*)ds_.cols, h.data(), sizeof(ColliderDev) * h.size(), hipMemcpyHostToDevice,
                           stream_));
    VTX_HIP(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------- device layout
template <class T>
T* Engine::dalloc(size_t count) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(sizeof(T) * count, 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        hipGetLastError();
        throw Error(VTX_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    allocs_.push_back(p);
    return (T*)p;
}

template <class T>
T* Engine::dupload(const std::vector<T>& h) {
    T* p = dalloc<T>(h.size());
    if (!h.empty()) VTX_HIP(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return p;
}

void Engine::free_device() {
    for (void* p : allocs_) hipFree(p);
    allocs_.clear();
    ds_ = DevState{};
}

static uint32_t ceil_log2(uint32_t v) {
    uint32_t b = 0;
    while ((1ull << b) < v) b++;
    return b;
}

// Sliced-ELL packing of per-particle lists for one tile: entries [off + k * VTX_TILE + lane].
template <class E>
static void pack_tile(std::vector<E>& out, const std::vector<uint32_t>& start, const std::vector<E>& items,
                      uint32_t first, uint32_t count, E pad, uint32_t* off, uint32_t* width) {
    uint32_t K = 0;
    for (uint32_t l = 0; l < count; l++) K = std::max(K, start[first + l + 1] - start[first + l]);
    *off = (uint32_t)out.size();
    *width = K;
    out.resize(out.size() + (size_t)K * VTX_TILE, pad);
    for (uint32_t l = 0; l < count; l++) {
        const uint32_t b = start[first + l], e = start[first + l + 1];
        for (uint32_t k = 0; k < e - b; k++) out[*off + (size_t)k * VTX_TILE + l] = items[b + k];
    }
}

// Sliced-ELL packing with an explicit lane -> particle map (-1: lane has no particle).
template <class E>
static void pack_lanes(std::vector<E>& out, const std::vector<uint32_t>& start, const std::vector<E>& items,
                       const std::vector<int64_t>& lanes, E pad, uint32_t* off, uint32_t* width) {
    uint32_t K = 0;
    for (int64_t i : lanes)
        if (i >= 0) K = std::max(K, start[(size_t)i + 1] - start[(size_t)i]);
    *off = (uint32_t)out.size();
    *width = K;
    out.resize(out.size() + (size_t)K * VTX_TILE, pad);
    for (size_t l = 0; l < lanes.size(); l++) {
        if (lanes[l] < 0) continue;
        const uint32_t b = start[(size_t)lanes[l]], e = start[(size_t)lanes[l] + 1];
        for (uint32_t k = 0; k < e - b; k++) out[*off + (size_t)k * VTX_TILE + l] = items[b + k];
    }
}

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// ---------------------------------------------------------------------------- grid stencil
// A segment takes the stencil path when it holds one cloth that is an R x C vertex grid with
// the triangulation of GenerateClothMesh (Velvet/Scene.hpp:153-165), every stretch constraint
// joins grid neighbours along a row, a column or a quad diagonal (at most once per pair) and
// every bending constraint is the dihedral of an interior edge with its two opposite vertices
// (the per-quad form of Velvet/VtClothObjectGPU.hpp:121-130 or any interior edge). The
// constraint set and its values are kept exactly; only the storage becomes a stencil.
struct GridLayout {
    uint32_t rows = 0, cols = 0;
    std::vector<float4> rest4;     // per vertex: R, D, X-diagonal, anti-diagonal rest (NaN = none)
    std::vector<uint32_t> flags;   // per cell: VTX_BX | VTX_BR | VTX_BD
    std::vector<float4> brest4;    // per cell: X, R, D rest angles
    bool rest0 = true;
};

static bool detect_grid(const Segment& sg, GridLayout& G) {
    if (sg.cloth_starts.size() != 1) return false;
    const uint32_t n = sg.n();
    const size_t T = sg.tris.size() / 3;
    if (n < 4 || T < 2) return false;
    const uint32_t cols = sg.tris[1];
    if (cols < 2 || n % cols) return false;
    const uint32_t rows = n / cols;
    if (rows < 2 || T != 2ull * (rows - 1) * (cols - 1)) return false;
    size_t t = 0;
    for (uint32_t a = 0; a + 1 < rows; a++)
        for (uint32_t b = 0; b + 1 < cols; b++) {
            const uint32_t A = a * cols + b, B = A + cols, C = A + 1, D = B + 1;
            const uint32_t* q = &sg.tris[6 * t / 2];
            if (q[0] != A || q[1] != B || q[2] != C || q[3] != C || q[4] != B || q[5] != D) return false;
            t += 2;
        }
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    G.rows = rows;
    G.cols = cols;
    G.rest4.assign(n, make_float4(qnan, qnan, qnan, qnan));
    const size_t S = sg.st_rest.size();
    for (size_t c = 0; c < S; c++) {
        const int64_t i = sg.st[2 * c], j = sg.st[2 * c + 1];
        const int64_t ai = i / cols, bi = i % cols, aj = j / cols, bj = j % cols;
        const int64_t da = aj - ai, db = bj - bi;
        int64_t owner;
        int comp;
        if (da == 0 && db == 1) { owner = i; comp = 0; }
        else if (da == 0 && db == -1) { owner = j; comp = 0; }
        else if (da == 1 && db == 0) { owner = i; comp = 1; }
        else if (da == -1 && db == 0) { owner = j; comp = 1; }
        else if (da == 1 && db == 1) { owner = i; comp = 2; }
        else if (da == -1 && db == -1) { owner = j; comp = 2; }
        else if (da == 1 && db == -1) { owner = ai * cols + (bi - 1); comp = 3; }   // (a,b+1)-(a+1,b) owned by (a,b)
        else if (da == -1 && db == 1) { owner = aj * cols + (bj - 1); comp = 3; }
        else return false;
        float* slot = &G.rest4[(size_t)owner].x + comp;
        if (*slot == *slot) return false;   // a second constraint on the same pair
        *slot = sg.st_rest[c];
    }
    G.flags.assign(n, 0u);
    G.brest4.assign(n, make_float4(0.f, 0.f, 0.f, 0.f));
    auto rc = [cols](uint32_t v, int64_t& a, int64_t& b) { a = v / cols; b = v % cols; };
    auto same_pair = [](int64_t x0, int64_t x1, int64_t y0, int64_t y1) {
        return (x0 == y0 && x1 == y1) || (x0 == y1 && x1 == y0);
    };
    const size_t B = sg.bd_rest.size();
    for (size_t c = 0; c < B; c++) {
        const uint32_t* q = &sg.bd[4 * c];
        int64_t au, bu, av, bv;
        rc(q[2], au, bu);
        rc(q[3], av, bv);
        int64_t a, b, o0, o1;
        int type;
        if (au == av && (bu - bv == 1 || bv - bu == 1)) {
            type = 1; a = au; b = std::min(bu, bv);
            if (a < 1 || a + 1 >= rows) return false;
            o0 = (a + 1) * cols + b; o1 = (a - 1) * cols + b + 1;
        } else if (bu == bv && (au - av == 1 || av - au == 1)) {
            type = 2; a = std::min(au, av); b = bu;
            if (b < 1 || b + 1 >= cols) return false;
            o0 = a * cols + b + 1; o1 = (a + 1) * cols + b - 1;
        } else if ((au - av == 1 && bv - bu == 1) || (av - au == 1 && bu - bv == 1)) {
            type = 0; a = std::min(au, av); b = std::min(bu, bv);
            o0 = a * cols + b; o1 = (a + 1) * cols + b + 1;
        } else {
            return false;
        }
        if (!same_pair(q[0], q[1], o0, o1)) return false;
        uint32_t& f = G.flags[(size_t)(a * cols + b)];
        if (f & (1u << type)) return false;
        f |= 1u << type;
        (&G.brest4[(size_t)(a * cols + b)].x)[type] = sg.bd_rest[c];
        if (sg.bd_rest[c] != 0.f) G.rest0 = false;
    }
    return true;
}

void Engine::set_stencil_path(bool on) {
    if (on == use_stencil_) return;
    pull_state();
    use_stencil_ = on;
    mark_dirty();
}

int32_t Engine::stencil_segments() {
    ensure_committed();
    int32_t k = 0;
    for (uint8_t g : seg_grid_) k += g;
    return k;
}

void Engine::commit() {
    VTX_HIP(hipStreamSynchronize(stream_));
    if (gexec_) {
        hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
    free_device();
    hseg_.clear();
    htile_.clear();

    const vtx_params& p0 = segs_[0].prm;
    substeps_ = p0.num_substeps;
    iterations_ = p0.num_iterations;
    interleave_ = p0.interleaved_hash;
    for (auto& sg : segs_) {
        VTX_REQUIRE(sg.prm.num_substeps == substeps_ && sg.prm.num_iterations == iterations_ &&
                        sg.prm.interleaved_hash == interleave_ &&
                        sg.prm.max_num_neighbors == p0.max_num_neighbors,
                    VTX_EINVAL, "segments of one solver must share substeps, iterations, interleave and max neighbours");
    }

    std::vector<float4> hx, hv, hx0;
    std::vector<uint2> est, eat;
    std::vector<uint32_t> ebd, etr;
    std::vector<uint4> brec, trec;
    std::vector<float> brest;
    std::vector<float4> hslots;
    std::vector<ColliderDev> hcols;
    std::vector<SortTileDev> hst;
    std::vector<TileDev> hitile;
    std::vector<GridTileDev> hgtile;
    std::vector<float4> hrest4, hbrest4;
    std::vector<uint32_t> hgflags;
    seg_grid_.assign(segs_.size(), 0);

    uint32_t pbase = 0, tbase = 0, sbase = 0, cbase = 0, bbase = 0, trbase = 0, hbase = 0;
    uint64_t nbase = 0;
    uint32_t maxbits = 0;
    any_selfcol_ = false;
    need_collide_ = false;

    for (size_t si = 0; si < segs_.size(); si++) {
        const Segment& sg = segs_[si];
        const vtx_params& p = sg.prm;
        const uint32_t n = sg.n();
        SegDev d{};
        d.pbase = pbase;
        d.n = n;
        const bool selfcol = p.enable_self_collision && n > 0;
        d.M = selfcol ? sg.table_size() : 0;
        VTX_REQUIRE(!selfcol || d.M >= 1, VTX_EINVAL, "hash table size must be >= 1");
        VTX_REQUIRE(!selfcol || n < (1u << 25), VTX_ECAPACITY,
                    "self collision supports at most 2^25 particles per segment");
        d.tbase = tbase;
        d.nbase_lo = (uint32_t)(nbase & 0xFFFFFFFFu);
        d.nbase_hi = (uint32_t)(nbase >> 32);
        d.sbase = sbase;
        d.scount = (uint32_t)(sg.slots.size() / 3);
        d.cbase = cbase;
        d.ccount = (uint32_t)sg.cols.size();
        d.flags = selfcol ? VTX_SEG_SELFCOL : 0u;
        d.max_nbr = (uint32_t)p.max_num_neighbors;
        // binary32 products exactly as the reference forms them (Velvet/SpatialHashGPU.hpp:21,44,47)
        d.diameter = p.particle_diameter;
        d.cell = p.particle_diameter * p.hash_cell_size_scalar;
        d.cell2 = d.cell * d.cell;
        d.diam2 = p.particle_diameter * p.particle_diameter;
        d.max_speed = p.max_speed;
        d.damping = p.damping;
        d.omega = p.relaxation_factor;
        d.lrs = p.long_range_stretchiness;
        d.bend_compliance = p.bend_compliance;
        d.margin = p.collision_margin;
        d.friction = p.friction;
        d.stiffness = p.stretch_stiffness;
        d.gx = p.gravity[0];
        d.gy = p.gravity[1];
        d.gz = p.gravity[2];
        if (selfcol) {
            const uint64_t c = UINT64_MAX / d.M + 1;   // ceil(2^64 / M); wraps to 0 for M = 1
            d.mod_lo = (uint32_t)(c & 0xFFFFFFFFu);
            d.mod_hi = (uint32_t)(c >> 32);
        }
        GridLayout G;
        const bool grid = use_stencil_ && detect_grid(sg, G);
        seg_grid_[si] = grid ? 1 : 0;
        if (grid) {
            d.flags |= VTX_SEG_GRID | (G.rest0 ? VTX_SEG_BEND_REST0 : 0u);
            // per slot type: all geometrically possible slots present, or none
            for (int type = 0; type < 3; type++) {
                bool full = true, none = true;
                for (uint32_t a = 0; a < G.rows && (full || none); a++)
                    for (uint32_t b = 0; b < G.cols; b++) {
                        const bool possible = a + 1 < G.rows && b + 1 < G.cols && (type != 1 || a >= 1) &&
                                              (type != 2 || b >= 1);
                        const bool present = (G.flags[(size_t)a * G.cols + b] >> type) & 1u;
                        if (present != possible) full = false;
                        if (present) none = false;
                    }
                if (full) d.flags |= VTX_SEG_FULL_X << type;
                if (none) d.flags |= VTX_SEG_NONE_X << type;
            }
            d.rows = G.rows;
            d.cols = G.cols;
            hrest4.insert(hrest4.end(), G.rest4.begin(), G.rest4.end());
            hgflags.insert(hgflags.end(), G.flags.begin(), G.flags.end());
            hbrest4.insert(hbrest4.end(), G.brest4.begin(), G.brest4.end());
        } else {
            hrest4.resize(hrest4.size() + n, make_float4(0.f, 0.f, 0.f, 0.f));
            hgflags.resize(hgflags.size() + n, 0u);
            hbrest4.resize(hbrest4.size() + n, make_float4(0.f, 0.f, 0.f, 0.f));
        }
        hseg_.push_back(d);
        if (selfcol) {
            any_selfcol_ = true;
            maxbits = std::max(maxbits, std::max(1u, ceil_log2(d.M)));
        }
        if (selfcol || d.ccount) need_collide_ = true;

        for (uint32_t i = 0; i < n; i++) {
            hx.push_back(make_float4(sg.x[3 * i], sg.x[3 * i + 1], sg.x[3 * i + 2], sg.w[i]));
            hv.push_back(make_float4(sg.v[3 * i], sg.v[3 * i + 1], sg.v[3 * i + 2], 0.f));
            hx0.push_back(make_float4(sg.x0[3 * i], sg.x0[3 * i + 1], sg.x0[3 * i + 2], 0.f));
        }
        for (uint32_t k = 0; k < d.scount; k++)
            hslots.push_back(make_float4(sg.slots[3 * k], sg.slots[3 * k + 1], sg.slots[3 * k + 2], 0.f));
        for (auto& c : sg.cols) hcols.push_back(pack_collider(c));

        // incidence lists (CSR over local particles, registration order)
        std::vector<uint32_t> s_st(n + 1, 0), s_at(n + 1, 0), s_bd(n + 1, 0), s_tr(n + 1, 0);
        const size_t S = sg.st_rest.size(), A = sg.at_pid.size(), B = sg.bd_rest.size(), T = sg.tris.size() / 3;
        for (size_t c = 0; c < S; c++) {
            s_st[(size_t)sg.st[2 * c] + 1]++;
            s_st[(size_t)sg.st[2 * c + 1] + 1]++;
        }
        for (size_t c = 0; c < A; c++) s_at[(size_t)sg.at_pid[c] + 1]++;
        for (size_t c = 0; c < B; c++)
            for (int r = 0; r < 4; r++) s_bd[(size_t)sg.bd[4 * c + r] + 1]++;
        for (size_t c = 0; c < T; c++)
            for (int r = 0; r < 3; r++) s_tr[(size_t)sg.tris[3 * c + r] + 1]++;
        for (uint32_t i = 0; i < n; i++) {
            s_st[i + 1] += s_st[i];
            s_at[i + 1] += s_at[i];
            s_bd[i + 1] += s_bd[i];
            s_tr[i + 1] += s_tr[i];
        }
        std::vector<uint2> l_st(s_st[n]), l_at(s_at[n]);
        std::vector<uint32_t> l_bd(s_bd[n]), l_tr(s_tr[n]);
        {
            std::vector<uint32_t> f(s_st.begin(), s_st.end() - 1);
            for (size_t c = 0; c < S; c++) {
                const uint32_t a = (uint32_t)sg.st[2 * c], b = (uint32_t)sg.st[2 * c + 1];
                const uint32_t r = fbits(sg.st_rest[c]);
                l_st[f[a]++] = make_uint2(pbase + b, r);
                l_st[f[b]++] = make_uint2(pbase + a, r);
            }
        }
        {
            std::vector<uint32_t> f(s_at.begin(), s_at.end() - 1);
            for (size_t c = 0; c < A; c++)
                l_at[f[(size_t)sg.at_pid[c]]++] = make_uint2(sbase + (uint32_t)sg.at_slot[c], fbits(sg.at_dist[c]));
        }
        {
            std::vector<uint32_t> f(s_bd.begin(), s_bd.end() - 1);
            for (size_t c = 0; c < B; c++)
                for (uint32_t r = 0; r < 4; r++)
                    l_bd[f[sg.bd[4 * c + r]]++] = (bbase + (uint32_t)c) | (r << 30);
        }
        {
            std::vector<uint32_t> f(s_tr.begin(), s_tr.end() - 1);
            for (size_t c = 0; c < T; c++)
                for (int r = 0; r < 3; r++) l_tr[f[sg.tris[3 * c + r]]++] = trbase + (uint32_t)c;
        }
        for (size_t c = 0; c < B; c++) {
            brec.push_back(make_uint4(pbase + sg.bd[4 * c], pbase + sg.bd[4 * c + 1], pbase + sg.bd[4 * c + 2],
                                      pbase + sg.bd[4 * c + 3]));
            brest.push_back(sg.bd_rest[c]);
        }
        for (size_t c = 0; c < T; c++)
            trec.push_back(make_uint4(pbase + sg.tris[3 * c], pbase + sg.tris[3 * c + 1], pbase + sg.tris[3 * c + 2], 0));

        for (uint32_t first = 0; first < n; first += VTX_TILE) {
            TileDev t{};
            t.seg = (uint32_t)si;
            t.begin = pbase + first;
            t.count = std::min<uint32_t>(VTX_TILE, n - first);
            if (!grid) {   // the stencil path keeps its own attach lists and needs no stretch/bend lists
                pack_tile(est, s_st, l_st, first, t.count, make_uint2(VTX_INVALID, 0u), &t.st_off, &t.st_k);
                pack_tile(eat, s_at, l_at, first, t.count, make_uint2(VTX_INVALID, 0u), &t.at_off, &t.at_k);
                pack_tile(ebd, s_bd, l_bd, first, t.count, VTX_INVALID, &t.bd_off, &t.bd_k);
            }
            pack_tile(etr, s_tr, l_tr, first, t.count, VTX_INVALID, &t.tr_off, &t.tr_k);
            htile_.push_back(t);
            if (!grid) hitile.push_back(t);
        }
        if (grid) {
            for (uint32_t a0 = 0; a0 < G.rows; a0 += VTX_GT)
                for (uint32_t b0 = 0; b0 < G.cols; b0 += VTX_GT) {
                    GridTileDev gt{};
                    gt.seg = (uint32_t)si;
                    gt.a0 = a0;
                    gt.b0 = b0;
                    std::vector<int64_t> lanes(VTX_TILE, -1);
                    for (uint32_t l = 0; l < VTX_TILE; l++) {
                        const uint32_t a = a0 + l / VTX_GT, b = b0 + l % VTX_GT;
                        if (a < G.rows && b < G.cols) lanes[l] = (int64_t)a * G.cols + b;
                    }
                    pack_lanes(eat, s_at, l_at, lanes, make_uint2(VTX_INVALID, 0u), &gt.at_off, &gt.at_k);
                    hgtile.push_back(gt);
                }
        }
        const uint32_t nst = (n + VTX_SORT_TILE - 1) / VTX_SORT_TILE;
        for (uint32_t k = 0; k < nst; k++) {
            SortTileDev st{};
            st.seg = (uint32_t)si;
            st.begin = pbase + k * VTX_SORT_TILE;
            st.count = std::min<uint32_t>(VTX_SORT_TILE, n - k * VTX_SORT_TILE);
            st.hbase = hbase;
            st.ntiles = nst;
            st.tidx = k;
            hst.push_back(st);
        }
        hbase += 256u * nst;
        pbase += n;
        tbase += d.M;
        nbase += selfcol ? (uint64_t)d.max_nbr * n : 0;
        sbase += d.scount;
        cbase += d.ccount;
        bbase += (uint32_t)B;
        trbase += (uint32_t)T;
    }
    passes_ = std::max(1, (int)((maxbits + 7) / 8));

    DevState d{};
    d.nparticles = pbase;
    d.ntiles = (uint32_t)htile_.size();
    d.nitiles = (uint32_t)hitile.size();
    d.ngtiles = (uint32_t)hgtile.size();
    d.itiles = dupload(hitile);
    d.gtiles = dupload(hgtile);
    d.rest4 = dupload(hrest4);
    d.gflags = dupload(hgflags);
    d.brest4 = dupload(hbrest4);
    d.nstiles = (uint32_t)hst.size();
    d.nhist = hbase;
    d.ncells = tbase;
    d.segs = dupload(hseg_);
    d.tiles = dupload(htile_);
    d.stiles = dupload(hst);
    d.X = dupload(hx);
    d.V = dupload(hv);
    d.X0 = dupload(hx0);
    d.P[0] = dupload(hx);
    d.P[1] = dupload(hx);
    d.Nrm = dalloc<float4>(pbase);
    VTX_HIP(hipMemset(d.Nrm, 0, sizeof(float4) * std::max<uint32_t>(pbase, 1)));
    d.st = dupload(est);
    d.at = dupload(eat);
    d.bd = dupload(ebd);
    d.brec = dupload(brec);
    d.brest = dupload(brest);
    d.tr = dupload(etr);
    d.trec = dupload(trec);
    d.slots = dupload(hslots);
    d.cols = dupload(hcols);
    if (any_selfcol_) {
        for (int k = 0; k < 2; k++) {
            d.keys[k] = dalloc<uint32_t>(pbase);
            d.vals[k] = dalloc<uint32_t>(pbase);
        }
        d.hist = dalloc<uint32_t>(hbase);
        d.part = dalloc<uint32_t>((hbase + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK + 1);
        d.cells = dalloc<uint2>(tbase);
        VTX_HIP(hipMemset(d.cells, 0xFF, sizeof(uint2) * std::max<uint32_t>(tbase, 1)));
        d.sP = dalloc<float4>(pbase);
        d.sX0 = dalloc<float4>(pbase);
        d.nbr = dalloc<uint32_t>((size_t)nbase);
        VTX_HIP(hipMemset(d.nbr, 0xFF, sizeof(uint32_t) * std::max<uint64_t>(nbase, 1)));
    }
    VTX_REQUIRE((d.nhist + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK <= 65536u, VTX_ECAPACITY, "histogram too large");
    VTX_HIP(hipDeviceSynchronize());
    ds_ = d;
    dirty_ = false;
}

void Engine::ensure_committed() {
    if (dirty_) commit();
}

// ---------------------------------------------------------------------------- frames
void Engine::stage_begin(const char* label, hipStream_t s) {
    if (!timing_) return;
    if (stage_used_ == stages_.size()) {
        StageTime st;
        st.label = label;
        VTX_HIP(hipEventCreate(&st.a));
        VTX_HIP(hipEventCreate(&st.b));
        stages_.push_back(st);
    }
    stages_[stage_used_].label = label;
    VTX_HIP(hipEventRecord(stages_[stage_used_].a, s));
}

void Engine::stage_end(hipStream_t s) {
    if (!timing_) return;
    VTX_HIP(hipEventRecord(stages_[stage_used_].b, s));
    stage_used_++;
}

// SpatialHashGPU::Hash -> HashObjects (Velvet/SpatialHashGPU.hpp:41-52, Velvet/SpatialHashGPU.cu:159-196).
void Engine::hash_build(int in, hipStream_t s) {
    stage_begin("k_hash_keys", s);
    VTX_HIP(launch_hash_keys(ds_, in, s));
    stage_end(s);
    stage_begin("hash_sort", s);
    int src = 0;
    for (int pass = 0; pass < passes_; pass++) {
        VTX_HIP(launch_sort_pass(ds_, src, 8 * pass, s));
        src ^= 1;
    }
    stage_end(s);
    stage_begin("hash_cells", s);
    VTX_HIP(launch_fill_u32((uint32_t*)ds_.cells, 2 * ds_.ncells, VTX_INVALID, s));
    VTX_HIP(launch_cell_bounds(ds_, in, src, s));
    stage_end(s);
    stage_begin("k_neighbors", s);
    VTX_HIP(launch_neighbors(ds_, in, src, s));
    stage_end(s);
    sorted_src_ = src;
}

// VtClothSolverGPU::Simulate stage order (Velvet/VtClothSolverGPU.hpp:56-111). In timing mode
// every launch group is bracketed by an event pair labelled with its kernel.
void Engine::record_frame(float frame_dt, hipStream_t s) {
    const float h = frame_dt / (float)substeps_;
    stage_begin("k_frame_begin", s);
    VTX_HIP(launch_frame_begin(ds_, frame_dt, h, s));
    stage_end(s);
    int cur = 0;
    for (int sub = 0; sub < substeps_; sub++) {
        if (any_selfcol_ && sub % interleave_ == 0) hash_build(cur, s);
        if (need_collide_) {
            stage_begin("k_collide", s);
            VTX_HIP(launch_collide(ds_, cur, h, s));
            stage_end(s);
            cur ^= 1;
        }
        const bool last_sub = sub == substeps_ - 1;
        const int epi_last = last_sub ? EPILOGUE_FINALIZE : EPILOGUE_FINALIZE_PREDICT;
        if (iterations_ == 0) {
            stage_begin("k_iterate", s);
            VTX_HIP(launch_iterate(ds_, cur, h, epi_last, false, s));
            stage_end(s);
            cur ^= 1;
        }
        for (int it = 0; it < iterations_; it++) {
            const int epi = it == iterations_ - 1 ? epi_last : EPILOGUE_NONE;
            stage_begin("k_iterate", s);
            VTX_HIP(launch_iterate(ds_, cur, h, epi, true, s));
            stage_end(s);
            cur ^= 1;
        }
    }
    stage_begin("k_normals", s);
    VTX_HIP(launch_normals(ds_, s));
    stage_end(s);
}

void Engine::simulate(float frame_dt) {
    VTX_REQUIRE(frame_dt > 0.f, VTX_EINVAL, "frame_dt must be positive");
    ensure_committed();
    if (ds_.nparticles == 0) return;
    if (timing_) {
        stage_used_ = 0;
        record_frame(frame_dt, stream_);
        VTX_HIP(hipStreamSynchronize(stream_));
        for (size_t k = 0; k < stage_used_; k++) {
            float ms = 0.f;
            VTX_HIP(hipEventElapsedTime(&ms, stages_[k].a, stages_[k].b));
            timing_ms_[stages_[k].label] += ms;
            timing_n_[stages_[k].label] += 1;
            timing_ms_["total"] += ms;
            timing_n_["total"] += 1;
        }
    } else if (use_graph_) {
        if (!gexec_ || graph_dt_ != frame_dt) {
            if (gexec_) {
                hipGraphExecDestroy(gexec_);
                gexec_ = nullptr;
            }
            hipGraph_t graph = nullptr;
            VTX_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                record_frame(frame_dt, stream_);
            } catch (...) {
                hipStreamEndCapture(stream_, &graph);
                if (graph) hipGraphDestroy(graph);
                throw;
            }
            VTX_HIP(hipStreamEndCapture(stream_, &graph));
            hipError_t e = hipGraphInstantiate(&gexec_, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            VTX_HIP(e);
            graph_dt_ = frame_dt;
        }
        VTX_HIP(hipGraphLaunch(gexec_, stream_));
    } else {
        record_frame(frame_dt, stream_);
    }
    dev_live_ = true;
}

void Engine::synchronize() { VTX_HIP(hipStreamSynchronize(stream_)); }

void Engine::set_use_graph(bool on) {
    use_graph_ = on;
    if (!on && gexec_) {
        hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
}

void Engine::set_stream(hipStream_t s) {
    synchronize();
    if (own_stream_ && stream_) hipStreamDestroy(stream_);
    if (s) {
        stream_ = s;
        own_stream_ = false;
    } else {
        VTX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        own_stream_ = true;
    }
    if (gexec_) {
        hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
}

void Engine::set_timing(bool on) {
    timing_ = on;
    timing_ms_.clear();
    timing_n_.clear();
}

std::string Engine::timing_report() {
    std::ostringstream o;
    o << "{";
    bool first = true;
    o.precision(9);
    for (auto& kv : timing_ms_) {
        o << (first ? "" : ", ") << "\"" << kv.first << "\": {\"ms\": " << kv.second
          << ", \"launches\": " << timing_n_[kv.first] << "}";
        first = false;
    }
    o << "}";
    return o.str();
}

// ---------------------------------------------------------------------------- state access
void Engine::read_particles(int32_t s, int what, float* out) {
    check_seg(s);
    VTX_REQUIRE(out, VTX_EINVAL, "output is NULL");
    ensure_committed();
    synchronize();
    const SegDev& d = hseg_[(size_t)s];
    if (!d.n) return;
    std::vector<float4> tmp(d.n);
    const float4* src = what == 0 ? ds_.X : what == 1 ? ds_.V : what == 2 ? ds_.Nrm : ds_.X;
    VTX_HIP(hipMemcpy(tmp.data(), src + d.pbase, sizeof(float4) * d.n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < d.n; i++) {
        if (what == 3) {
            out[i] = tmp[i].w;
        } else {
            out[3 * i] = tmp[i].x;
            out[3 * i + 1] = tmp[i].y;
            out[3 * i + 2] = tmp[i].z;
        }
    }
}

float4* Engine::device_positions(int32_t s, uint32_t* n) {
    check_seg(s);
    ensure_committed();
    if (n) *n = hseg_[(size_t)s].n;
    return ds_.X + hseg_[(size_t)s].pbase;
}

float4* Engine::device_normals(int32_t s, uint32_t* n) {
    check_seg(s);
    ensure_committed();
    if (n) *n = hseg_[(size_t)s].n;
    return ds_.Nrm + hseg_[(size_t)s].pbase;
}

// MouseGrabber-style edit (Velvet/MouseGrabber.hpp:51-52,60,74-75).
void Engine::set_particle(int32_t s, int32_t i, const float* pos, const float* vel, float inv_mass) {
    Segment& sg = seg(s);
    VTX_REQUIRE(i >= 0 && (uint32_t)i < sg.n(), VTX_EINVAL, "particle index out of range");
    if (dirty_ || !dev_live_) {
        if (pos) std::copy(pos, pos + 3, sg.x.begin() + 3 * i);
        if (vel) std::copy(vel, vel + 3, sg.v.begin() + 3 * i);
        if (inv_mass >= 0.f) sg.w[(size_t)i] = inv_mass;
        if (!dirty_) mark_dirty();
        return;
    }
    synchronize();
    const uint32_t g = hseg_[(size_t)s].pbase + (uint32_t)i;
    float4 x;
    VTX_HIP(hipMemcpy(&x, ds_.X + g, sizeof(float4), hipMemcpyDeviceToHost));
    if (pos) {
        x.x = pos[0];
        x.y = pos[1];
        x.z = pos[2];
    }
    if (inv_mass >= 0.f) x.w = inv_mass;
    VTX_HIP(hipMemcpy(ds_.X + g, &x, sizeof(float4), hipMemcpyHostToDevice));
    if (vel) {
        float4 v = make_float4(vel[0], vel[1], vel[2], 0.f);
        VTX_HIP(hipMemcpy(ds_.V + g, &v, sizeof(float4), hipMemcpyHostToDevice));
    }
}

void Engine::write_predicted(int32_t s, const float* xyz) {
    check_seg(s);
    VTX_REQUIRE(xyz, VTX_EINVAL, "positions are NULL");
    ensure_committed();
    synchronize();
    const SegDev& d = hseg_[(size_t)s];
    if (!d.n) return;
    std::vector<float4> tmp(d.n);
    VTX_HIP(hipMemcpy(tmp.data(), ds_.X + d.pbase, sizeof(float4) * d.n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < d.n; i++) {
        tmp[i].x = xyz[3 * i];
        tmp[i].y = xyz[3 * i + 1];
        tmp[i].z = xyz[3 * i + 2];
    }
    VTX_HIP(hipMemcpy(ds_.P[0] + d.pbase, tmp.data(), sizeof(float4) * d.n, hipMemcpyHostToDevice));
}

void Engine::run_hash() {
    ensure_committed();
    VTX_REQUIRE(any_selfcol_, VTX_ESTATE, "no segment has self collision enabled");
    hash_build(0, stream_);
    synchronize();
}

uint32_t Engine::hash_table_size(int32_t s) {
    check_seg(s);
    ensure_committed();
    return hseg_[(size_t)s].M;
}

void Engine::read_hash(int32_t s, uint32_t* cs, uint32_t* ce, uint32_t* sh, uint32_t* si, uint32_t* nbr) {
    check_seg(s);
    ensure_committed();
    synchronize();
    const SegDev& d = hseg_[(size_t)s];
    VTX_REQUIRE(d.flags & VTX_SEG_SELFCOL, VTX_ESTATE, "segment has no spatial hash");
    if (cs || ce) {
        std::vector<uint2> tmp(d.M);
        VTX_HIP(hipMemcpy(tmp.data(), ds_.cells + d.tbase, sizeof(uint2) * d.M, hipMemcpyDeviceToHost));
        for (uint32_t h = 0; h < d.M; h++) {
            if (cs) cs[h] = tmp[h].x;
            if (ce) ce[h] = tmp[h].y;
        }
    }
    if (sh) VTX_HIP(hipMemcpy(sh, ds_.keys[sorted_src_] + d.pbase, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost));
    if (si) VTX_HIP(hipMemcpy(si, ds_.vals[sorted_src_] + d.pbase, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost));
    if (nbr) {
        const uint64_t nb = (uint64_t)d.nbase_hi << 32 | d.nbase_lo;
        VTX_HIP(hipMemcpy(nbr, ds_.nbr + nb, sizeof(uint32_t) * (size_t)d.max_nbr * d.n, hipMemcpyDeviceToHost));
    }
}

}  // namespace vtx
