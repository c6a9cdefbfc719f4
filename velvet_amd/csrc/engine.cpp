// Host engine: registration mirrors VtClothSolverGPU (Velvet/VtClothSolverGPU.hpp:114-205).
// Segments are prepared on the host with segment-local indices, uploaded into ranges of
// device pools (so segments can join and leave a running solver without touching the others),
// and every frame (Velvet/VtClothSolverGPU.hpp:56-111) is recorded once and replayed as a
// hipGraph until the set of active segments changes.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <sstream>

#include "launch.hpp"

namespace vtx {

static void fill_state(const Segment& sg, Prepared& pr);

static const float kFixedDt = 1.0f / 60.0f;   // Velvet/Timer.hpp:235

// ---------------------------------------------------------------------------- ranges
bool RangeAlloc::take(size_t n, size_t* off) {
    if (n == 0) {
        *off = 0;
        return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= n) {
            *off = it->first;
            const size_t rest = it->second - n, o = it->first + n;
            free_.erase(it);
            if (rest) free_[o] = rest;
            used_ += n;
            return true;
        }
    }
    return false;
}

void RangeAlloc::give(size_t off, size_t n) {
    if (n == 0) return;
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

void RangeAlloc::grow(size_t new_cap) {
    if (new_cap <= cap) return;
    const size_t add = new_cap - cap;
    used_ += add;            // give() subtracts it again: the new space starts free
    give(cap, add);
    cap = new_cap;
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

// ---------------------------------------------------------------------------- lifecycle
Engine::Engine(const vtx_params& p) {
    VTX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
    add_segment(p);
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (hipStream_t x : xstreams_) (void)hipStreamSynchronize(x);
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    for (auto& st : stages_) {
        (void)hipEventDestroy(st.a);
        (void)hipEventDestroy(st.b);
    }
    free_tables();
    free_pools();
    if (hist_) (void)hipFree(hist_);
    if (part_) (void)hipFree(part_);
    if (flags_) (void)hipFree(flags_);
    for (hipStream_t x : xstreams_) (void)hipStreamDestroy(x);
    if (cstream_) {
        (void)hipStreamSynchronize(cstream_);
        (void)hipStreamDestroy(cstream_);
    }
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    for (hipEvent_t e : ev_join_) (void)hipEventDestroy(e);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Engine::check_seg(int32_t s) const {
    VTX_REQUIRE(s >= 0 && s < (int32_t)segs_.size() && segs_[(size_t)s], VTX_EINVAL, "no such segment");
}

Segment& Engine::seg(int32_t i) {
    check_seg(i);
    return *segs_[(size_t)i];
}

int32_t Engine::num_segments() const {
    int32_t k = 0;
    for (auto& p : segs_) k += p ? 1 : 0;
    return k;
}

static void validate_params(const vtx_params& p) {
    VTX_REQUIRE(p.num_substeps >= 1, VTX_EINVAL, "num_substeps must be >= 1");
    VTX_REQUIRE(p.num_iterations >= 0, VTX_EINVAL, "num_iterations must be >= 0");
    VTX_REQUIRE(p.max_num_neighbors >= 1 && p.max_num_neighbors <= VTX_MAX_NEIGHBORS, VTX_EINVAL,
                "max_num_neighbors must be in [1, 64]");
    VTX_REQUIRE(p.interleaved_hash >= 1, VTX_EINVAL, "interleaved_hash must be >= 1");
    VTX_REQUIRE(p.hash_table_size >= 0, VTX_EINVAL, "hash_table_size must be >= 0");
    VTX_REQUIRE(p.tune_hash_grid == 0 || p.tune_hash_grid == 1, VTX_EINVAL, "tune_hash_grid must be 0 or 1");
}

int32_t Engine::add_segment(const vtx_params& p, bool active) {
    validate_params(p);
    auto s = std::make_unique<Segment>();
    s->prm = p;
    s->active = active;
    segs_.push_back(std::move(s));
    if (active) tables_dirty_ = true;
    return (int32_t)segs_.size() - 1;
}

void Engine::remove_segment(int32_t s) {
    join();
    Segment& sg = seg(s);
    if (sg.uploaded) {
        synchronize();
        release(sg);
    }
    if (sg.active) tables_dirty_ = true;
    segs_[(size_t)s].reset();
}

void Engine::set_active(int32_t s, bool active) {
    join();
    Segment& sg = seg(s);
    if (sg.active == active) return;
    sg.active = active;
    tables_dirty_ = true;
}

// Deactivate a segment and give its device memory back; its state goes to the host and its
// prepared layout is kept, so re-activation only uploads.
void Engine::park(int32_t s) {
    join();
    Segment& sg = seg(s);
    if (sg.uploaded) {
        pull_state(sg);
        if (sg.prepared) fill_state(sg, sg.prep);
        release(sg);
        if (sg.active) tables_dirty_ = true;
    }
    if (sg.active) {
        sg.active = false;
        tables_dirty_ = true;
    }
}

namespace {
// Snapshot blob: header, then n float4 positions (w = inverse mass), n float4 velocities,
// n float4 initial positions and L float4 slot positions.
struct SnapHeader {
    uint32_t magic, version, n, nslots;
    uint64_t bytes, reserved;
};
constexpr uint32_t kSnapMagic = 0x53585456u;   // "VTXS"
}  // namespace

size_t Engine::snapshot_bytes(int32_t s) {
    Segment& sg = seg(s);
    return sizeof(SnapHeader) + sizeof(float4) * (3 * (size_t)sg.n() + sg.slots.size() / 3);
}

void Engine::snapshot(int32_t s, void* buf, size_t bytes) {
    join();
    Segment& sg = seg(s);
    const size_t need = snapshot_bytes(s);
    VTX_REQUIRE(buf && bytes >= need, VTX_EINVAL, "snapshot buffer too small");
    const uint32_t n = sg.n(), L = (uint32_t)(sg.slots.size() / 3);
    SnapHeader h{kSnapMagic, 1u, n, L, (uint64_t)need, 0};
    std::memcpy(buf, &h, sizeof h);
    float4* out = reinterpret_cast<float4*>(static_cast<char*>(buf) + sizeof h);
    if (sg.uploaded) {
        synchronize();
        if (n) {
            VTX_HIP(hipMemcpy(out, pl_.X + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
            VTX_HIP(hipMemcpy(out + n, pl_.V + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
            VTX_HIP(hipMemcpy(out + 2 * (size_t)n, pl_.X0 + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
        }
        if (L) VTX_HIP(hipMemcpy(out + 3 * (size_t)n, pl_.slots + sg.rg.slot, sizeof(float4) * L, hipMemcpyDeviceToHost));
    } else {
        Prepared tmp;
        fill_state(sg, tmp);
        std::copy(tmp.x.begin(), tmp.x.end(), out);
        std::copy(tmp.v.begin(), tmp.v.end(), out + n);
        std::copy(tmp.x0.begin(), tmp.x0.end(), out + 2 * (size_t)n);
        std::copy(tmp.slots.begin(), tmp.slots.end(), out + 3 * (size_t)n);
    }
}

void Engine::restore(int32_t s, const void* buf, size_t bytes) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(buf && bytes >= sizeof(SnapHeader), VTX_EINVAL, "snapshot buffer too small");
    SnapHeader h;
    std::memcpy(&h, buf, sizeof h);
    const uint32_t n = sg.n(), L = (uint32_t)(sg.slots.size() / 3);
    VTX_REQUIRE(h.magic == kSnapMagic && h.version == 1u, VTX_EINVAL, "not a segment snapshot");
    VTX_REQUIRE(h.n == n && h.nslots == L, VTX_EINVAL, "snapshot is of a segment with other particle or slot counts");
    VTX_REQUIRE(h.bytes == snapshot_bytes(s) && bytes >= h.bytes, VTX_EINVAL, "snapshot size mismatch");
    const float4* in = reinterpret_cast<const float4*>(static_cast<const char*>(buf) + sizeof h);
    for (uint32_t i = 0; i < n; i++) {
        const float4 x = in[i], v = in[n + i], x0 = in[2 * (size_t)n + i];
        sg.x[3 * i] = x.x; sg.x[3 * i + 1] = x.y; sg.x[3 * i + 2] = x.z;
        sg.w[i] = x.w;
        sg.v[3 * i] = v.x; sg.v[3 * i + 1] = v.y; sg.v[3 * i + 2] = v.z;
        sg.x0[3 * i] = x0.x; sg.x0[3 * i + 1] = x0.y; sg.x0[3 * i + 2] = x0.z;
    }
    for (uint32_t k = 0; k < L; k++) {
        const float4 q = in[3 * (size_t)n + k];
        sg.slots[3 * k] = q.x; sg.slots[3 * k + 1] = q.y; sg.slots[3 * k + 2] = q.z;
    }
    if (sg.prepared) fill_state(sg, sg.prep);
    if (sg.uploaded) {
        synchronize();
        if (n) {
            VTX_HIP(hipMemcpy(pl_.X + sg.rg.p, in, sizeof(float4) * n, hipMemcpyHostToDevice));
            VTX_HIP(hipMemcpy(pl_.V + sg.rg.p, in + n, sizeof(float4) * n, hipMemcpyHostToDevice));
            VTX_HIP(hipMemcpy(pl_.X0 + sg.rg.p, in + 2 * (size_t)n, sizeof(float4) * n, hipMemcpyHostToDevice));
            VTX_HIP(hipMemcpy(pl_.P0 + sg.rg.p, in, sizeof(float4) * n, hipMemcpyHostToDevice));
            VTX_HIP(hipMemcpy(pl_.P1 + sg.rg.p, in, sizeof(float4) * n, hipMemcpyHostToDevice));
        }
        if (L) VTX_HIP(hipMemcpy(pl_.slots + sg.rg.slot, in + 3 * (size_t)n, sizeof(float4) * L, hipMemcpyHostToDevice));
    }
}

Readback::~Readback() {
    for (int k = 0; k < kBuffers; k++) {
        if (ev[k]) {
            (void)hipEventSynchronize(ev[k]);
            (void)hipEventDestroy(ev[k]);
        }
        if (host[k]) (void)hipHostFree(host[k]);
    }
}

void Engine::readback_request(int32_t s) {
    join();
    Segment& sg = seg(s);
    ensure_committed();
    VTX_REQUIRE(sg.active && sg.uploaded, VTX_ESTATE, "segment is not on the device");
    const uint32_t n = sg.n();
    if (!n) return;
    if (!sg.rb || sg.rb->n != n) {
        sg.rb.reset(new Readback);
        for (int k = 0; k < Readback::kBuffers; k++) {
            VTX_HIP(hipHostMalloc((void**)&sg.rb->host[k], sizeof(float4) * 2 * (size_t)n, hipHostMallocDefault));
            VTX_HIP(hipEventCreateWithFlags(&sg.rb->ev[k], hipEventDisableTiming));
        }
        sg.rb->n = n;
    }
    Readback& rb = *sg.rb;
    int k = -1;
    for (int t = 0; t < Readback::kBuffers; t++)
        if (t != rb.handed && (k < 0 || rb.seq[t] < rb.seq[k])) k = t;
    VTX_HIP(hipMemcpyAsync(rb.host[k], pl_.X + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost, stream_));
    VTX_HIP(hipMemcpyAsync(rb.host[k] + n, pl_.N + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost, stream_));
    VTX_HIP(hipEventRecord(rb.ev[k], stream_));
    rb.frame[k] = frames_;
    rb.pending[k] = true;
    rb.seq[k] = ++rb.nseq;
}

bool Engine::readback_latest(int32_t s, const float4** pos, const float4** nrm, uint32_t* n, uint64_t* frame) {
    Segment& sg = seg(s);
    if (!sg.rb) return false;
    Readback& rb = *sg.rb;
    int order[Readback::kBuffers];
    for (int t = 0; t < Readback::kBuffers; t++) order[t] = t;
    std::sort(order, order + Readback::kBuffers, [&](int a, int b) { return rb.seq[a] > rb.seq[b]; });
    for (int k : order) {                        // newest request first
        if (!rb.pending[k]) continue;
        const hipError_t e = hipEventQuery(rb.ev[k]);
        if (e == hipSuccess) {
            *pos = rb.host[k];
            *nrm = rb.host[k] + rb.n;
            *n = rb.n;
            *frame = rb.frame[k];
            rb.handed = k;
            return true;
        }
        if (e != hipErrorNotReady) VTX_HIP(e);
    }
    return false;
}

void Engine::invalidate_graph() {
    if (gexec_) {
        (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
}

// Device -> host copy of one segment's particle state.
void Engine::pull_state(Segment& sg) {
    if (!sg.uploaded) return;
    sync_stream();
    const uint32_t n = sg.n();
    if (!n) return;
    std::vector<float4> tmp(n);
    VTX_HIP(hipMemcpy(tmp.data(), pl_.X + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
        sg.x[3 * i] = tmp[i].x;
        sg.x[3 * i + 1] = tmp[i].y;
        sg.x[3 * i + 2] = tmp[i].z;
        sg.w[i] = tmp[i].w;
    }
    VTX_HIP(hipMemcpy(tmp.data(), pl_.V + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
        sg.v[3 * i] = tmp[i].x;
        sg.v[3 * i + 1] = tmp[i].y;
        sg.v[3 * i + 2] = tmp[i].z;
    }
}

// A structural change of one segment: keep its simulated state on the host, drop its device
// ranges and prepared layout; other segments are untouched.
void Engine::touch(int32_t s) {
    Segment& sg = seg(s);
    if (sg.uploaded) {
        pull_state(sg);
        release(sg);
        if (sg.active) tables_dirty_ = true;
    }
    sg.prepared = false;
}

void Engine::set_params(int32_t s, const vtx_params& p) {
    join();
    validate_params(p);
    touch(s);
    seg(s).prm = p;
}

// AddCloth (Velvet/VtClothSolverGPU.hpp:114-156).
int32_t Engine::add_cloth(int32_t s, const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                          const float* model, float diameter) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(n == 0 || xyz, VTX_EINVAL, "positions are NULL");
    VTX_REQUIRE(ntri == 0 || tris, VTX_EINVAL, "triangles are NULL");
    for (uint32_t t = 0; t < 3 * ntri; t++) VTX_REQUIRE(tris[t] < n, VTX_EINVAL, "triangle index out of range");
    touch(s);
    const uint32_t prev = sg.n();
    sg.prm.particle_diameter = diameter;
    sg.prm.max_speed = 2.0f * diameter / kFixedDt * (float)sg.prm.num_substeps;
    for (uint32_t t = 0; t < 3 * ntri; t++) sg.tris.push_back(tris[t] + prev);
    for (uint32_t i = 0; i < n; i++) {
        float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
        if (model) {   // InitializePositions_Kernel: x = M * vec4(x, 1) (Velvet/VtClothSolverGPU.cu:30-34)
            const float* m = model;
            const float ox = m[0] * px + m[4] * py + m[8] * pz + m[12];
            const float oy = m[1] * px + m[5] * py + m[9] * pz + m[13];
            const float oz = m[2] * px + m[6] * py + m[10] * pz + m[14];
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
    return (int32_t)prev;
}

void Engine::add_stretch(int32_t s, const int32_t* ij, const float* rest, uint32_t count) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (ij && rest), VTX_EINVAL, "stretch arrays are NULL");
    const int32_t n = (int32_t)sg.n();
    for (uint32_t c = 0; c < 2 * count; c++)
        VTX_REQUIRE(ij[c] >= 0 && ij[c] < n, VTX_EINVAL, "stretch index out of range");
    touch(s);
    sg.st.insert(sg.st.end(), ij, ij + 2 * count);
    sg.st_rest.insert(sg.st_rest.end(), rest, rest + count);
}

int32_t Engine::add_attach_slot(int32_t s, const float* pos) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(pos, VTX_EINVAL, "slot position is NULL");
    touch(s);
    sg.slots.insert(sg.slots.end(), pos, pos + 3);
    return (int32_t)(sg.slots.size() / 3 - 1);
}

void Engine::set_attach_slot_position(int32_t s, int32_t slot, const float* pos) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(pos, VTX_EINVAL, "slot position is NULL");
    VTX_REQUIRE(slot >= 0 && (size_t)slot < sg.slots.size() / 3, VTX_EINVAL, "slot index out of range");
    std::copy(pos, pos + 3, sg.slots.begin() + 3 * slot);
    const float4 v = make_float4(pos[0], pos[1], pos[2], 0.f);
    if (sg.prepared) sg.prep.slots[(size_t)slot] = v;
    if (sg.uploaded) {   // the graph reads the slot table: no re-record
        VTX_HIP(hipMemcpyAsync(pl_.slots + sg.rg.slot + slot, &v, sizeof(float4), hipMemcpyHostToDevice, stream_));
        sync_stream();
    }
}

// AddAttach (Velvet/VtClothSolverGPU.hpp:170-176): distance 0 pins the particle.
void Engine::add_attach(int32_t s, const int32_t* pid, const int32_t* slot, const float* dist, uint32_t count) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (pid && slot && dist), VTX_EINVAL, "attach arrays are NULL");
    const int32_t n = (int32_t)sg.n();
    const int32_t L = (int32_t)(sg.slots.size() / 3);
    for (uint32_t c = 0; c < count; c++) {
        VTX_REQUIRE(pid[c] >= 0 && pid[c] < n, VTX_EINVAL, "attach particle out of range");
        VTX_REQUIRE(slot[c] >= 0 && slot[c] < L, VTX_EINVAL, "attach slot out of range");
    }
    touch(s);
    for (uint32_t c = 0; c < count; c++)
        if (dist[c] == 0.0f) sg.w[(size_t)pid[c]] = 0.0f;
    sg.at_pid.insert(sg.at_pid.end(), pid, pid + count);
    sg.at_slot.insert(sg.at_slot.end(), slot, slot + count);
    sg.at_dist.insert(sg.at_dist.end(), dist, dist + count);
}

void Engine::add_bend(int32_t s, const uint32_t* idx, const float* angle, uint32_t count) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(count == 0 || (idx && angle), VTX_EINVAL, "bend arrays are NULL");
    VTX_REQUIRE(sg.bd.size() / 4 + count < (1u << 28), VTX_ECAPACITY, "too many bend constraints");
    for (uint32_t c = 0; c < 4 * count; c++) VTX_REQUIRE(idx[c] < sg.n(), VTX_EINVAL, "bend index out of range");
    touch(s);
    sg.bd.insert(sg.bd.end(), idx, idx + 4 * count);
    sg.bd_rest.insert(sg.bd_rest.end(), angle, angle + count);
}

void Engine::add_mesh_constraints(int32_t s, int32_t offset, const uint32_t* tris, uint32_t ntri) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(ntri == 0 || tris, VTX_EINVAL, "triangles are NULL");
    VTX_REQUIRE(offset >= 0 && (uint32_t)offset <= sg.n(), VTX_EINVAL, "cloth offset out of range");
    if (sg.uploaded) pull_state(sg);
    std::vector<uint32_t> g((size_t)3 * ntri);
    for (size_t t = 0; t < g.size(); t++) {
        VTX_REQUIRE((uint64_t)tris[t] + (uint64_t)offset < sg.n(), VTX_EINVAL, "triangle index out of range");
        g[t] = tris[t] + (uint32_t)offset;
    }
    std::vector<int32_t> ij;
    std::vector<float> sr, br;
    std::vector<uint32_t> bi;
    mesh_constraints(sg.x.data(), sg.n(), g.data(), ntri, ij, sr, bi, br);
    if (!sr.empty()) add_stretch(s, ij.data(), sr.data(), (uint32_t)sr.size());
    if (!br.empty()) add_bend(s, bi.data(), br.data(), (uint32_t)br.size());
}

void Engine::set_colliders(int32_t s, const vtx_collider* cols, uint32_t n) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(n == 0 || cols, VTX_EINVAL, "colliders are NULL");
    for (uint32_t c = 0; c < n; c++)
        VTX_REQUIRE(cols[c].type >= 0 && cols[c].type <= 2, VTX_EINVAL, "unknown collider type");
    if (sg.cols.size() == n && sg.prepared) {   // same count: rewrite in place, no re-record
        sg.cols.assign(cols, cols + n);
        for (uint32_t c = 0; c < n; c++) sg.prep.cols[c] = pack_collider(cols[c]);
        if (sg.uploaded && n) {
            VTX_HIP(hipMemcpyAsync(pl_.cols + sg.rg.col, sg.prep.cols.data(), sizeof(ColliderDev) * n,
                                   hipMemcpyHostToDevice, stream_));
            sync_stream();
        }
        return;
    }
    touch(s);
    sg.cols.assign(cols, cols + n);
}

// ---------------------------------------------------------------------------- preparation
// The simulated state of a segment in device format: positions with inverse masses,
// velocities, initial positions and slot positions.
static void fill_state(const Segment& sg, Prepared& pr) {
    const uint32_t n = sg.n();
    pr.x.resize(n);
    pr.v.resize(n);
    pr.x0.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        pr.x[i] = make_float4(sg.x[3 * i], sg.x[3 * i + 1], sg.x[3 * i + 2], sg.w[i]);
        pr.v[i] = make_float4(sg.v[3 * i], sg.v[3 * i + 1], sg.v[3 * i + 2], 0.f);
        pr.x0[i] = make_float4(sg.x0[3 * i], sg.x0[3 * i + 1], sg.x0[3 * i + 2], 0.f);
    }
    const uint32_t L = (uint32_t)(sg.slots.size() / 3);
    pr.slots.resize(L);
    for (uint32_t k = 0; k < L; k++)
        pr.slots[k] = make_float4(sg.slots[3 * k], sg.slots[3 * k + 1], sg.slots[3 * k + 2], 0.f);
}

// Host packing of one segment into its position-independent device layout.
void Engine::prepare(int32_t si) { prepare_segment(seg(si)); }

void Engine::prepare_segment(Segment& sg) {
    if (sg.prepared) return;
    Prepared pr;
    const uint32_t n = sg.n();
    VTX_REQUIRE(!sg.selfcol() || n < (1u << 25), VTX_ECAPACITY,
                "self collision supports at most 2^25 particles per segment");
    VTX_REQUIRE(!sg.selfcol() || sg.table_size() < (1u << 28), VTX_ECAPACITY, "hash table size must be < 2^28");
    fill_state(sg, pr);
    const uint32_t L = (uint32_t)(sg.slots.size() / 3);
    for (auto& c : sg.cols) pr.cols.push_back(pack_collider(c));

    GridLayout G;
    pr.grid = use_stencil_ && detect_grid(sg, G);
    if (pr.grid) {
        pr.grows = G.rows;
        pr.gcols = G.cols;
        pr.grid_flags = VTX_SEG_GRID | (G.rest0 ? VTX_SEG_BEND_REST0 : 0u);
        for (int type = 0; type < 3; type++) {   // all geometrically possible slots present, or none
            bool full = true, none = true;
            for (uint32_t a = 0; a < G.rows && (full || none); a++)
                for (uint32_t b = 0; b < G.cols; b++) {
                    const bool possible = a + 1 < G.rows && b + 1 < G.cols && (type != 1 || a >= 1) && (type != 2 || b >= 1);
                    const bool present = (G.flags[(size_t)a * G.cols + b] >> type) & 1u;
                    if (present != possible) full = false;
                    if (present) none = false;
                }
            if (full) pr.grid_flags |= VTX_SEG_FULL_X << type;
            if (none) pr.grid_flags |= VTX_SEG_NONE_X << type;
        }
        pr.rest4 = std::move(G.rest4);
        pr.gflags = std::move(G.flags);
        pr.brest4 = std::move(G.brest4);
    }

    // incidence lists (CSR over local particles, registration order); the stencil path reads
    // neither stretch / bend lists nor triangle lists (its normals come from the grid), so a grid
    // segment builds only the attachment lists
    std::vector<uint32_t> s_st(n + 1, 0), s_at(n + 1, 0), s_bd(n + 1, 0), s_tr(n + 1, 0);
    const size_t S = sg.st_rest.size(), A = sg.at_pid.size(), B = sg.bd_rest.size();
    const size_t T = pr.grid ? 0 : sg.tris.size() / 3;
    if (!pr.grid) {
        for (size_t c = 0; c < S; c++) {
            s_st[(size_t)sg.st[2 * c] + 1]++;
            s_st[(size_t)sg.st[2 * c + 1] + 1]++;
        }
        for (size_t c = 0; c < B; c++)
            for (int r = 0; r < 4; r++) s_bd[(size_t)sg.bd[4 * c + r] + 1]++;
    }
    for (size_t c = 0; c < A; c++) s_at[(size_t)sg.at_pid[c] + 1]++;
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
    if (!pr.grid) {
        std::vector<uint32_t> f(s_st.begin(), s_st.end() - 1);
        for (size_t c = 0; c < S; c++) {
            const uint32_t a = (uint32_t)sg.st[2 * c], b = (uint32_t)sg.st[2 * c + 1];
            const uint32_t r = fbits(sg.st_rest[c]);
            l_st[f[a]++] = make_uint2(b, r);
            l_st[f[b]++] = make_uint2(a, r);
        }
        std::vector<uint32_t> fb(s_bd.begin(), s_bd.end() - 1);
        for (size_t c = 0; c < B; c++)
            for (uint32_t r = 0; r < 4; r++) l_bd[fb[sg.bd[4 * c + r]]++] = (uint32_t)c | (r << 30);
        for (size_t c = 0; c < B; c++) {
            pr.brec.push_back(make_uint4(sg.bd[4 * c], sg.bd[4 * c + 1], sg.bd[4 * c + 2], sg.bd[4 * c + 3]));
            pr.brest.push_back(sg.bd_rest[c]);
        }
    }
    // dense long-range attachments: entry c = slot * n + particle for every slot and particle
    pr.lra_dense = n > 0 && L > 0 && A == (size_t)L * n;
    for (size_t c = 0; c < A && pr.lra_dense; c++)
        pr.lra_dense = (size_t)sg.at_pid[c] == c % n && (size_t)sg.at_slot[c] == c / n;
    if (pr.lra_dense) {
        pr.lra = sg.at_dist;   // already slot-major
        std::fill(s_at.begin(), s_at.end(), 0u);
        l_at.clear();
    } else {
        std::vector<uint32_t> f(s_at.begin(), s_at.end() - 1);
        for (size_t c = 0; c < A; c++)
            l_at[f[(size_t)sg.at_pid[c]]++] = make_uint2((uint32_t)sg.at_slot[c], fbits(sg.at_dist[c]));
    }
    {
        std::vector<uint32_t> f(s_tr.begin(), s_tr.end() - 1);
        for (size_t c = 0; c < T; c++)
            for (int r = 0; r < 3; r++) l_tr[f[sg.tris[3 * c + r]]++] = (uint32_t)c;
        for (size_t c = 0; c < T; c++)
            pr.trec.push_back(make_uint4(sg.tris[3 * c], sg.tris[3 * c + 1], sg.tris[3 * c + 2], 0));
    }
    for (uint32_t first = 0; first < n; first += VTX_TILE) {
        TileDev t{};
        t.begin = first;                   // local; made absolute in the tables
        t.count = std::min<uint32_t>(VTX_TILE, n - first);
        if (!pr.grid) {   // the stencil path keeps its own attach lists and needs no stretch/bend lists
            pack_tile(pr.est, s_st, l_st, first, t.count, make_uint2(VTX_INVALID, 0u), &t.st_off, &t.st_k);
            pack_tile(pr.eat, s_at, l_at, first, t.count, make_uint2(VTX_INVALID, 0u), &t.at_off, &t.at_k);
            pack_tile(pr.ebd, s_bd, l_bd, first, t.count, VTX_INVALID, &t.bd_off, &t.bd_k);
        }
        if (!pr.grid) pack_tile(pr.etr, s_tr, l_tr, first, t.count, VTX_INVALID, &t.tr_off, &t.tr_k);
        pr.tiles.push_back(t);
    }
    if (pr.grid) {
        for (uint32_t a0 = 0; a0 < pr.grows; a0 += VTX_GT)
            for (uint32_t b0 = 0; b0 < pr.gcols; b0 += VTX_GT) {
                GridTileDev gt{};
                gt.a0 = a0;
                gt.b0 = b0;
                std::vector<int64_t> lanes(VTX_TILE, -1);
                for (uint32_t l = 0; l < VTX_TILE; l++) {
                    const uint32_t a = a0 + l / VTX_GT, b = b0 + l % VTX_GT;
                    if (a < pr.grows && b < pr.gcols) lanes[l] = (int64_t)a * pr.gcols + b;
                }
                pack_lanes(pr.eat, s_at, l_at, lanes, make_uint2(VTX_INVALID, 0u), &gt.at_off, &gt.at_k);
                pr.gtiles.push_back(gt);
            }
    }
    sg.prep = std::move(pr);
    sg.prepared = true;
}

static Ranges needs_of(const Segment& sg) {
    const Prepared& pr = sg.prep;
    Ranges r;
    r.n_p = sg.n();
    r.n_cell = sg.cell_capacity();
    // neighbour lists in chunks of four entries (the cap rounded up to whole chunks), 32-bit
    // entries, or 16-bit ones (VTX_SEG_NBR16) in a range rounded up to whole 16-byte units
    const size_t k4n = (size_t)((sg.prm.max_num_neighbors + 3) & ~3) * sg.n();
    r.n_nbr = !sg.selfcol() ? 0 : sg.n() <= VTX_NBR16_N ? ((k4n / 2 + 3) & ~size_t(3)) : k4n;
    r.n_st = pr.est.size();
    r.n_at = pr.eat.size();
    r.n_bd = pr.ebd.size();
    r.n_tr = pr.etr.size();
    r.n_brec = pr.brec.size();
    r.n_trec = pr.trec.size();
    r.n_slot = pr.slots.size();
    r.n_col = pr.cols.size();
    r.n_lra = pr.lra.size();
    return r;
}

static size_t bytes_of(const Ranges& r) {
    return r.n_p * (10 * sizeof(float4) + 5 * sizeof(uint32_t)) + r.n_cell * sizeof(uint2) + r.n_nbr * 4 +
           (r.n_st + r.n_at) * sizeof(uint2) + (r.n_bd + r.n_tr) * 4 + r.n_brec * (sizeof(uint4) + 4) +
           r.n_trec * sizeof(uint4) + r.n_slot * sizeof(float4) + r.n_col * sizeof(ColliderDev) + r.n_lra * 4;
}

size_t Engine::bytes_estimate(const Segment& sg) { return bytes_of(needs_of(sg)); }

size_t Engine::device_bytes_estimate(int32_t s) {
    prepare(s);
    return bytes_of(needs_of(seg(s)));
}

size_t Engine::pool_bytes() const {
    Ranges c;
    c.n_p = ap_.cap;
    c.n_cell = acell_.cap;
    c.n_nbr = anbr_.cap;
    c.n_st = ast_.cap;
    c.n_at = aat_.cap;
    c.n_bd = abd_.cap;
    c.n_tr = atr_.cap;
    c.n_brec = abrec_.cap;
    c.n_trec = atrec_.cap;
    c.n_slot = aslot_.cap;
    c.n_col = acol_.cap;
    c.n_lra = alra_.cap;
    return bytes_of(c);
}

// ---------------------------------------------------------------------------- pools
template <class T>
static void regrow(T*& p, size_t old_cap, size_t new_cap) {
    T* q = nullptr;
    const size_t bytes = std::max<size_t>(sizeof(T) * new_cap, 256);
    hipError_t e = hipMalloc((void**)&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(VTX_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    if (p && old_cap) VTX_HIP(hipMemcpy(q, p, sizeof(T) * old_cap, hipMemcpyDeviceToDevice));
    if (p) VTX_HIP(hipFree(p));
    p = q;
}

static size_t next_cap(const RangeAlloc& a, size_t need_more) {
    const size_t want = a.used() + need_more;
    size_t c = std::max<size_t>(a.cap, 1);
    while (c < want) c = c + c / 2 + 1;
    return c;
}

// Make every pool large enough for `need` more entries (existing contents are kept in place).
void Engine::grow_pools(const Ranges& need) {
    auto grow = [&](RangeAlloc& a, size_t more, auto&& move) {
        // first-fit may fail on fragmentation even when used + more <= cap: grow by `more` then
        size_t off;
        if (a.take(more, &off)) {
            a.give(off, more);
            return;
        }
        const size_t nc = std::max(next_cap(a, more), a.cap + more);
        move(a.cap, nc);
        a.grow(nc);
        invalidate_graph();
        tables_dirty_ = true;
    };
    grow(ap_, need.n_p, [&](size_t o, size_t c) {
        regrow(pl_.X, o, c); regrow(pl_.V, o, c); regrow(pl_.P0, o, c); regrow(pl_.P1, o, c);
        regrow(pl_.X0, o, c); regrow(pl_.N, o, c); regrow(pl_.sPX, 2 * o, 2 * c);
        regrow(pl_.rest4, o, c); regrow(pl_.brest4, o, c); regrow(pl_.k0, o, c); regrow(pl_.k1, o, c);
        regrow(pl_.v0, o, c); regrow(pl_.v1, o, c); regrow(pl_.gflags, o, c);
    });
    grow(acell_, need.n_cell, [&](size_t o, size_t c) { regrow(pl_.cells, o, c); });
    grow(anbr_, need.n_nbr, [&](size_t o, size_t c) { regrow(pl_.nbr, o, c); });
    grow(ast_, need.n_st, [&](size_t o, size_t c) { regrow(pl_.st, o, c); });
    grow(aat_, need.n_at, [&](size_t o, size_t c) { regrow(pl_.at, o, c); });
    grow(abd_, need.n_bd, [&](size_t o, size_t c) { regrow(pl_.bd, o, c); });
    grow(atr_, need.n_tr, [&](size_t o, size_t c) { regrow(pl_.tr, o, c); });
    grow(abrec_, need.n_brec, [&](size_t o, size_t c) { regrow(pl_.brec, o, c); regrow(pl_.brest, o, c); });
    grow(atrec_, need.n_trec, [&](size_t o, size_t c) { regrow(pl_.trec, o, c); });
    grow(aslot_, need.n_slot, [&](size_t o, size_t c) { regrow(pl_.slots, o, c); });
    grow(acol_, need.n_col, [&](size_t o, size_t c) { regrow(pl_.cols, o, c); });
    grow(alra_, need.n_lra, [&](size_t o, size_t c) { regrow(pl_.lra, o, c); });
}

// Whether every pool can take `need` without growing.
bool Engine::pools_fit(const Ranges& need) {
    auto fits = [](RangeAlloc& a, size_t n) {
        size_t off;
        if (!n) return true;
        if (!a.take(n, &off)) return false;
        a.give(off, n);
        return true;
    };
    return fits(ap_, need.n_p) && fits(acell_, need.n_cell) && fits(anbr_, need.n_nbr) && fits(ast_, need.n_st) &&
           fits(aat_, need.n_at) && fits(abd_, need.n_bd) && fits(atr_, need.n_tr) && fits(abrec_, need.n_brec) &&
           fits(atrec_, need.n_trec) && fits(aslot_, need.n_slot) && fits(acol_, need.n_col) && fits(alra_, need.n_lra);
}

void Engine::stage(int32_t s) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(!sg.active, VTX_ESTATE, "only an inactive segment can be staged");
    if (sg.uploaded) return;
    prepare(s);
    if (!pools_fit(needs_of(sg))) sync_stream();   // growth moves every pool
    if (!cstream_) VTX_HIP(hipStreamCreateWithFlags(&cstream_, hipStreamNonBlocking));
    upload(sg, cstream_);
}

void Engine::reserve(uint64_t particles) {
    join();
    if (!particles) return;
    sync_stream();
    Ranges need;
    need.n_p = particles;
    need.n_cell = 2 * particles;                              // default table size 2 n
    need.n_nbr = 64 * particles;                              // default neighbour cap
    need.n_tr = 8 * particles;                                // triangle lists, 6 per vertex + padding
    need.n_trec = 2 * particles;                              // 2 triangles per vertex
    grow_pools(need);
}

void Engine::free_pools() {
    void* ptrs[] = {pl_.X, pl_.V, pl_.P0, pl_.P1, pl_.X0, pl_.N, pl_.sPX, pl_.rest4, pl_.brest4,
                    pl_.k0, pl_.k1, pl_.v0, pl_.v1, pl_.gflags, pl_.cells, pl_.nbr, pl_.st, pl_.at, pl_.bd,
                    pl_.tr, pl_.brec, pl_.trec, pl_.brest, pl_.slots, pl_.cols, pl_.lra};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    pl_ = Pools{};
}

template <class T>
static void h2d(T* dst, const std::vector<T>& src, hipStream_t s) {
    if (!src.empty()) VTX_HIP(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, s));
}

// Copy one prepared segment into freshly taken pool ranges.
void Engine::upload(Segment& sg, hipStream_t s) {
    const Ranges need = needs_of(sg);
    grow_pools(need);
    Ranges r = need;
    auto take = [](RangeAlloc& a, size_t n, size_t* off) {
        if (!a.take(n, off)) throw Error(VTX_ENOMEM, "pool range allocation failed");
    };
    take(ap_, need.n_p, &r.p);
    take(acell_, need.n_cell, &r.cell);
    take(anbr_, need.n_nbr, &r.nbr);
    // every neighbour range is a multiple of four words, so every offset is one too: the contact
    // pass's chunk loads stay aligned
    if (r.nbr % 4) throw Error(VTX_ESTATE, "neighbour table range not 16-byte aligned");
    take(ast_, need.n_st, &r.st);
    take(aat_, need.n_at, &r.at);
    take(abd_, need.n_bd, &r.bd);
    take(atr_, need.n_tr, &r.tr);
    take(abrec_, need.n_brec, &r.brec);
    take(atrec_, need.n_trec, &r.trec);
    take(aslot_, need.n_slot, &r.slot);
    take(acol_, need.n_col, &r.col);
    take(alra_, need.n_lra, &r.lra);
    VTX_REQUIRE(r.lra + r.n_lra < (1ull << 32), VTX_ECAPACITY, "dense attachment pool exceeds 2^32 entries");
    sg.rg = r;
    // a failed copy gives the ranges back, so that a later attempt starts clean
    try {
        const Prepared& pr = sg.prep;
        const vtx_params& p = sg.prm;
        const uint32_t n = sg.n();
        SegDev d{};
        d.pbase = (uint32_t)r.p;
        d.n = n;
        const bool selfcol = sg.selfcol();
        d.M = selfcol ? sg.table_size() : 0;
        VTX_REQUIRE(!selfcol || d.M >= 1, VTX_EINVAL, "hash table size must be >= 1");
        VTX_REQUIRE(!selfcol || d.M < (1u << 28), VTX_ECAPACITY, "hash table size must be < 2^28");
        d.tbase = (uint32_t)r.cell;
        d.nbase_lo = (uint32_t)(r.nbr & 0xFFFFFFFFu);
        d.nbase_hi = (uint32_t)((uint64_t)r.nbr >> 32);
        d.sbase = (uint32_t)r.slot;
        d.scount = (uint32_t)pr.slots.size();
        d.cbase = (uint32_t)r.col;
        d.ccount = (uint32_t)pr.cols.size();
        d.bbase = (uint32_t)r.brec;
        d.trbase = (uint32_t)r.trec;
        d.flags = (selfcol ? VTX_SEG_SELFCOL : 0u) | (pr.grid ? pr.grid_flags : 0u) |
                  (pr.lra_dense ? VTX_SEG_LRA_DENSE : 0u) | (selfcol && n <= VTX_NBR16_N ? VTX_SEG_NBR16 : 0u);
        d.lbase = (uint32_t)r.lra;
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
        d.rows = pr.grows;
        d.cols = pr.gcols;
        sg.dev = d;

        // Host-to-device bytes are the cost of admitting a segment: copies that repeat data already
        // on the device are made there (predicted positions, and initial positions when they equal
        // the positions), zero velocities are cleared, and all-zero rest angles are not read.
        const size_t pbytes = sizeof(float4) * n;
        h2d(pl_.X + r.p, pr.x, s);
        if (n) {
            VTX_HIP(hipMemcpyAsync(pl_.P0 + r.p, pl_.X + r.p, pbytes, hipMemcpyDeviceToDevice, s));
            VTX_HIP(hipMemcpyAsync(pl_.P1 + r.p, pl_.X + r.p, pbytes, hipMemcpyDeviceToDevice, s));
            // kernels read only xyz of the initial positions, so equal xyz lets X stand in
            bool same = pr.x0.size() == pr.x.size();
            for (uint32_t i = 0; same && i < n; i++) same = std::memcmp(&pr.x0[i], &pr.x[i], 12) == 0;
            if (same)
                VTX_HIP(hipMemcpyAsync(pl_.X0 + r.p, pl_.X + r.p, pbytes, hipMemcpyDeviceToDevice, s));
            else
                h2d(pl_.X0 + r.p, pr.x0, s);
            const bool v0 = std::all_of(pr.v.begin(), pr.v.end(), [](const float4& v) {
                return v.x == 0.f && v.y == 0.f && v.z == 0.f && v.w == 0.f;
            });
            if (v0) VTX_HIP(hipMemsetAsync(pl_.V + r.p, 0, pbytes, s));
            else h2d(pl_.V + r.p, pr.v, s);
            VTX_HIP(hipMemsetAsync(pl_.N + r.p, 0, pbytes, s));
        }
        if (pr.grid) {
            h2d(pl_.rest4 + r.p, pr.rest4, s);
            h2d(pl_.gflags + r.p, pr.gflags, s);
            if (!(pr.grid_flags & VTX_SEG_BEND_REST0)) h2d(pl_.brest4 + r.p, pr.brest4, s);
        }
        if (r.n_cell) VTX_HIP(hipMemsetAsync(pl_.cells + r.cell, 0xFF, sizeof(uint2) * r.n_cell, s));
        if (r.n_nbr) VTX_HIP(hipMemsetAsync(pl_.nbr + r.nbr, 0xFF, sizeof(uint32_t) * r.n_nbr, s));
        h2d(pl_.st + r.st, pr.est, s);
        h2d(pl_.at + r.at, pr.eat, s);
        h2d(pl_.bd + r.bd, pr.ebd, s);
        h2d(pl_.tr + r.tr, pr.etr, s);
        h2d(pl_.brec + r.brec, pr.brec, s);
        h2d(pl_.brest + r.brec, pr.brest, s);
        h2d(pl_.trec + r.trec, pr.trec, s);
        h2d(pl_.slots + r.slot, pr.slots, s);
        h2d(pl_.cols + r.col, pr.cols, s);
        h2d(pl_.lra + r.lra, pr.lra, s);
        VTX_HIP(hipStreamSynchronize(s));
    } catch (...) {
        sg.uploaded = true;
        release(sg);
        throw;
    }
    sg.uploaded = true;
}

void Engine::release(Segment& sg) {
    if (!sg.uploaded) return;
    const Ranges& r = sg.rg;
    ap_.give(r.p, r.n_p);
    acell_.give(r.cell, r.n_cell);
    anbr_.give(r.nbr, r.n_nbr);
    ast_.give(r.st, r.n_st);
    aat_.give(r.at, r.n_at);
    abd_.give(r.bd, r.n_bd);
    atr_.give(r.tr, r.n_tr);
    abrec_.give(r.brec, r.n_brec);
    atrec_.give(r.trec, r.n_trec);
    aslot_.give(r.slot, r.n_slot);
    acol_.give(r.col, r.n_col);
    alra_.give(r.lra, r.n_lra);
    sg.uploaded = false;
}

void Engine::free_tables() {
    if (tab_dev_) (void)hipFree(tab_dev_);
    if (tab_host_) (void)hipHostFree(tab_host_);
    tab_dev_ = tab_host_ = nullptr;
    tab_cap_ = 0;
}

// The launch tables live in one device arena, filled from one pinned staging buffer by one
// asynchronous copy on the solver's stream; both are reused across rebuilds and only grow.
// Callers have synchronized the stream, so neither buffer is still in use.
struct TableSet {
    std::vector<std::pair<const void*, size_t>> src;
    std::vector<size_t> off;
    size_t total = 0;
    template <class T>
    size_t add(const std::vector<T>& h) {
        const size_t o = (total + 255) & ~(size_t)255;
        src.emplace_back(h.data(), sizeof(T) * h.size());
        off.push_back(o);
        total = o + sizeof(T) * h.size();
        return src.size() - 1;
    }
};

void Engine::upload_tables(const TableSet& t) {
    const size_t bytes = std::max<size_t>(t.total, 256);
    if (bytes > tab_cap_) {
        free_tables();
        const size_t cap = bytes + bytes / 2;
        if (hipMalloc((void**)&tab_dev_, cap) != hipSuccess) {
            (void)hipGetLastError();
            tab_dev_ = nullptr;
            throw Error(VTX_ENOMEM, "hipMalloc of the launch tables failed");
        }
        if (hipHostMalloc((void**)&tab_host_, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            tab_host_ = nullptr;
            throw Error(VTX_ENOMEM, "pinned staging of the launch tables failed");
        }
        tab_cap_ = cap;
    }
    for (size_t k = 0; k < t.src.size(); k++)
        if (t.src[k].second) std::memcpy(tab_host_ + t.off[k], t.src[k].first, t.src[k].second);
    VTX_HIP(hipMemcpyAsync(tab_dev_, tab_host_, t.total, hipMemcpyHostToDevice, stream_));
}

static uint32_t ceil_log2(uint32_t v) {
    uint32_t b = 0;
    while ((1ull << b) < v) b++;
    return b;
}

// Segment, tile and sort tables of the active segments.
void Engine::rebuild_tables() {
    sync_stream();
    invalidate_graph();
    std::vector<SegDev> hseg;
    std::vector<TileDev> htile, hitile;
    std::vector<GridTileDev> hgtile, hltile;
    std::vector<uint32_t> hsmall;                // table indices of segments hashed in LDS
    std::vector<SortTileDev> hst;
    table_seg_.clear();
    any_selfcol_ = false;
    any_lra_dense_ = false;
    any_lra_split_ = false;
    any_hash_lds_ = false;
    any_grid_tunable_ = false;
    need_collide_ = false;
    live_particles_ = 0;
    uint32_t maxbits = 0, hbase = 0, compact = 0;
    bool first = true;
    // two groups (two streams): the first segments up to half of the particles, then the rest;
    // each group sorts on its own, so its scan positions restart at zero
    uint64_t total_n = 0;
    uint32_t nlive = 0;
    for (auto& sp : segs_)
        if (sp && sp->active && sp->uploaded && sp->n()) {
            total_n += sp->n();
            nlive++;
        }
    // G groups of about total/G particles each (a prefix of the segment table per group); timing
    // mode keeps one group so that per-launch figures cover the whole workload
    const uint32_t G = timing_ ? 1u : std::max(1u, std::min<uint32_t>((uint32_t)tun_.streams, nlive));
    struct Cut { size_t t, i, g, l, s, m; uint32_t h, p; };
    std::vector<Cut> cuts;                     // where groups 1..G-1 begin
    uint64_t cum = 0;
    for (size_t si = 0; si < segs_.size(); si++) {
        Segment* sgp = segs_[si].get();
        if (!sgp || !sgp->active || !sgp->uploaded || sgp->n() == 0) continue;
        Segment& sg = *sgp;
        const vtx_params& p = sg.prm;
        if (first) {
            substeps_ = p.num_substeps;
            iterations_ = p.num_iterations;
            interleave_ = p.interleaved_hash;
            first = false;
        }
        VTX_REQUIRE(p.num_substeps == substeps_ && p.num_iterations == iterations_ && p.interleaved_hash == interleave_,
                    VTX_EINVAL, "active segments of one solver must share substeps, iterations and interleave");
        if (cuts.size() + 1 < G && cum > 0 && cum * G >= total_n * (cuts.size() + 1)) {
            cuts.push_back(Cut{htile.size(), hitile.size(), hgtile.size(), hltile.size(), hst.size(), hsmall.size(),
                               hbase, (uint32_t)cum});
            compact = 0;                       // each group sorts on its own
        }
        cum += sg.n();
        const uint32_t ti = (uint32_t)hseg.size();
        table_seg_.push_back((int32_t)si);
        if (sg.dev.flags & VTX_SEG_SELFCOL) {   // hash grid of this table set (Tuning::hash_grid)
            const bool shifted = sg.prm.tune_hash_grid && tun_.hash_grid > 0;
            any_grid_tunable_ = any_grid_tunable_ || sg.prm.tune_hash_grid;
            sg.dev.M = sg.grid_table_size(shifted ? tun_.hash_grid : 0);
            const uint64_t c = UINT64_MAX / sg.dev.M + 1;   // ceil(2^64 / M); wraps to 0 for M = 1
            sg.dev.mod_lo = (uint32_t)(c & 0xFFFFFFFFu);
            sg.dev.mod_hi = (uint32_t)(c >> 32);
            sg.dev.flags = (sg.dev.flags & ~VTX_SEG_HASH_SHIFT) | (shifted ? VTX_SEG_HASH_SHIFT : 0u);
        }
        const bool hlds = (sg.dev.flags & VTX_SEG_SELFCOL) && sg.dev.n <= VTX_HASH_LDS_N && sg.dev.M <= (1u << 16);
        any_hash_lds_ = any_hash_lds_ || hlds;
        const bool in_lds = hlds && tun_.hash_lds;
        hseg.push_back(sg.dev);
        if (in_lds) {
            hseg.back().flags |= VTX_SEG_HASH_LDS;
            hsmall.push_back(ti);
        }
        const SegDev& d = sg.dev;
        live_particles_ += d.n;
        if (d.flags & VTX_SEG_SELFCOL) {
            any_selfcol_ = true;
            if (!in_lds) maxbits = std::max(maxbits, std::max(1u, ceil_log2(d.M)));
        }
        if ((d.flags & VTX_SEG_SELFCOL) || d.ccount) need_collide_ = true;
        if (d.flags & VTX_SEG_LRA_DENSE) any_lra_dense_ = true;
        const bool lsplit = sg.prep.grid && (d.flags & VTX_SEG_LRA_DENSE) && d.scount >= VTX_LRA_SPLIT_MIN;
        any_lra_split_ = any_lra_split_ || lsplit;
        for (TileDev t : sg.prep.tiles) {
            t.seg = ti;
            t.begin += d.pbase;
            t.st_off += (uint32_t)sg.rg.st;
            t.at_off += (uint32_t)sg.rg.at;
            t.bd_off += (uint32_t)sg.rg.bd;
            t.tr_off += (uint32_t)sg.rg.tr;
            htile.push_back(t);
            if (!sg.prep.grid) hitile.push_back(t);
        }
        for (GridTileDev gt : sg.prep.gtiles) {
            gt.seg = ti;
            gt.at_off += (uint32_t)sg.rg.at;
            (lsplit && tun_.lra_split ? hltile : hgtile).push_back(gt);
        }
        const uint32_t nst = in_lds ? 0u : (d.n + VTX_SORT_TILE - 1) / VTX_SORT_TILE;
        for (uint32_t k = 0; k < nst; k++) {
            SortTileDev st{};
            st.seg = ti;
            st.begin = d.pbase + k * VTX_SORT_TILE;
            st.count = std::min<uint32_t>(VTX_SORT_TILE, d.n - k * VTX_SORT_TILE);
            st.hbase = hbase;
            st.ntiles = nst;
            st.tidx = k;
            st.bias = d.pbase - compact;   // scan positions are compact; outputs go to the segment's range
            hst.push_back(st);
        }
        hbase += 256u * nst;
        if (!in_lds) compact += d.n;
    }
    passes_ = std::max(1, (int)((maxbits + 7) / 8));
    VTX_REQUIRE((hbase + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK <= 65536u, VTX_ECAPACITY, "histogram too large");
    const size_t npart = VTX_MAX_STREAMS * ((hbase + VTX_SCAN_BLOCK - 1) / VTX_SCAN_BLOCK + 1);   // a share per group
    if (hbase > hist_cap_) {
        if (hist_) (void)hipFree(hist_);
        hist_cap_ = hbase + hbase / 2;
        VTX_HIP(hipMalloc((void**)&hist_, sizeof(uint32_t) * std::max<size_t>(hist_cap_, 64)));
    }
    if (npart > part_cap_) {
        if (part_) (void)hipFree(part_);
        part_cap_ = npart + npart / 2;
        VTX_HIP(hipMalloc((void**)&part_, sizeof(uint32_t) * std::max<size_t>(part_cap_, 64)));
    }
    if (hseg.size() > flag_cap_) {
        if (flags_) (void)hipFree(flags_);
        flag_cap_ = hseg.size() + hseg.size() / 2 + 16;
        VTX_HIP(hipMalloc((void**)&flags_, sizeof(uint32_t) * flag_cap_));
    }

    DevState d{};
    d.nparticles = (uint32_t)total_n;
    d.ntiles = (uint32_t)htile.size();
    d.nitiles = (uint32_t)hitile.size();
    d.ngtiles = (uint32_t)hgtile.size();
    d.nltiles = (uint32_t)hltile.size();
    d.nhsegs = (uint32_t)hsmall.size();
    d.nstiles = (uint32_t)hst.size();
    d.nhist = hbase;
    d.ncells = (uint32_t)acell_.cap;
    {   // largest small segments first, within each group's range, so the longest builds start first
        std::vector<size_t> edge{0};
        for (const Cut& c : cuts) edge.push_back(c.m);
        edge.push_back(hsmall.size());
        for (size_t k = 0; k + 1 < edge.size(); k++)
            std::stable_sort(hsmall.begin() + (ptrdiff_t)edge[k], hsmall.begin() + (ptrdiff_t)edge[k + 1],
                             [&](uint32_t a, uint32_t b) { return hseg[a].n > hseg[b].n; });
    }
    TableSet ts;
    const size_t k_seg = ts.add(hseg), k_tile = ts.add(htile), k_itile = ts.add(hitile), k_gtile = ts.add(hgtile),
                 k_ltile = ts.add(hltile), k_st = ts.add(hst), k_hs = ts.add(hsmall);
    upload_tables(ts);
    d.segs = (const SegDev*)(tab_dev_ + ts.off[k_seg]);
    d.tiles = (const TileDev*)(tab_dev_ + ts.off[k_tile]);
    d.itiles = (const TileDev*)(tab_dev_ + ts.off[k_itile]);
    d.gtiles = (const GridTileDev*)(tab_dev_ + ts.off[k_gtile]);
    d.ltiles = (const GridTileDev*)(tab_dev_ + ts.off[k_ltile]);
    d.stiles = (const SortTileDev*)(tab_dev_ + ts.off[k_st]);
    d.hsegs = (const uint32_t*)(tab_dev_ + ts.off[k_hs]);
    d.X = pl_.X;
    d.V = pl_.V;
    d.P[0] = pl_.P0;
    d.P[1] = pl_.P1;
    d.X0 = pl_.X0;
    d.Nrm = pl_.N;
    d.sPX = pl_.sPX;
    d.rest4 = pl_.rest4;
    d.brest4 = pl_.brest4;
    d.gflags = pl_.gflags;
    d.keys[0] = pl_.k0;
    d.keys[1] = pl_.k1;
    d.vals[0] = pl_.v0;
    d.vals[1] = pl_.v1;
    d.cells = pl_.cells;
    d.nbr = pl_.nbr;
    d.st = pl_.st;
    d.at = pl_.at;
    d.bd = pl_.bd;
    d.tr = pl_.tr;
    d.brec = pl_.brec;
    d.brest = pl_.brest;
    d.trec = pl_.trec;
    d.slots = pl_.slots;
    d.lra = pl_.lra;
    d.remap = tun_.remap ? 1u : 0u;
    d.cols = pl_.cols;
    d.hist = hist_;
    d.part = part_;
    d.hbase0 = 0;
    ds_ = d;
    groups_.clear();
    cuts.push_back(Cut{htile.size(), hitile.size(), hgtile.size(), hltile.size(), hst.size(), hsmall.size(), hbase,
                       d.nparticles});
    Cut lo{0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t k = 0; k < cuts.size(); k++) {
        const Cut& hi = cuts[k];
        DevState g = d;
        g.tiles = d.tiles + lo.t;
        g.itiles = d.itiles + lo.i;
        g.gtiles = d.gtiles + lo.g;
        g.ltiles = d.ltiles + lo.l;
        g.stiles = d.stiles + lo.s;
        g.hsegs = d.hsegs + lo.m;
        g.nhsegs = (uint32_t)(hi.m - lo.m);
        g.ntiles = (uint32_t)(hi.t - lo.t);
        g.nitiles = (uint32_t)(hi.i - lo.i);
        g.ngtiles = (uint32_t)(hi.g - lo.g);
        g.nltiles = (uint32_t)(hi.l - lo.l);
        g.nstiles = (uint32_t)(hi.s - lo.s);
        g.hbase0 = lo.h;
        g.nhist = hi.h - lo.h;
        g.nparticles = hi.p - lo.p;
        g.part = part_ + k * (part_cap_ / VTX_MAX_STREAMS);
        groups_.push_back(g);
        lo = hi;
    }
    if (groups_.size() > 1) {
        while (xstreams_.size() + 1 < groups_.size()) {
            hipStream_t x = nullptr;
            hipEvent_t e = nullptr;
            VTX_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
            xstreams_.push_back(x);
            VTX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev_join_.push_back(e);
        }
        if (!ev_fork_) VTX_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    }
    tables_dirty_ = false;
}

void Engine::ensure_committed() {
    // size the pools once for everything that is about to be uploaded
    Ranges need;
    bool any = false;
    for (size_t si = 0; si < segs_.size(); si++) {
        Segment* sg = segs_[si].get();
        if (!sg || !sg->active || sg->uploaded) continue;
        prepare((int32_t)si);
        const Ranges r = needs_of(*sg);
        need.n_p += r.n_p; need.n_cell += r.n_cell; need.n_nbr += r.n_nbr; need.n_st += r.n_st;
        need.n_at += r.n_at; need.n_bd += r.n_bd; need.n_tr += r.n_tr; need.n_brec += r.n_brec;
        need.n_trec += r.n_trec; need.n_slot += r.n_slot; need.n_col += r.n_col; need.n_lra += r.n_lra;
        any = true;
    }
    if (any) {
        sync_stream();
        grow_pools(need);
        for (auto& sg : segs_)
            if (sg && sg->active && !sg->uploaded) upload(*sg, stream_);
        tables_dirty_ = true;
    }
    if (tables_dirty_) rebuild_tables();
}

// ---------------------------------------------------------------------------- grid path switch
void Engine::set_stencil_path(bool on) {
    join();
    if (on == use_stencil_) return;
    for (size_t si = 0; si < segs_.size(); si++)
        if (segs_[si]) touch((int32_t)si);
    use_stencil_ = on;
}

int32_t Engine::stencil_segments() {
    ensure_committed();
    int32_t k = 0;
    for (int32_t si : table_seg_) k += segs_[(size_t)si]->prep.grid ? 1 : 0;
    return k;
}

// ---------------------------------------------------------------------------- frames
void Engine::stage_begin(const char* kernel, const char* label, hipStream_t s) {
    if (!timing_) return;
    if (stage_used_ == stages_.size()) {
        StageTime st;
        VTX_HIP(hipEventCreate(&st.a));
        VTX_HIP(hipEventCreate(&st.b));
        stages_.push_back(st);
    }
    stages_[stage_used_].kernel = kernel;
    stages_[stage_used_].label = label;
    VTX_HIP(hipEventRecord(stages_[stage_used_].a, s));
}

void Engine::stage_end(hipStream_t s) {
    if (!timing_) return;
    VTX_HIP(hipEventRecord(stages_[stage_used_].b, s));
    stage_used_++;
}

// SpatialHashGPU::Hash -> HashObjects (Velvet/SpatialHashGPU.hpp:41-52, Velvet/SpatialHashGPU.cu:159-196).
void Engine::hash_build(const DevState& d, int in, hipStream_t s) {
    // small segments: the whole build in one workgroup each, into the buffer the device-wide
    // sort of the other segments ends in
    stage_begin("k_hash_small", "Solver_HashParticle+Solver_HashSort+Solver_HashBuildCell", s);
    VTX_HIP(launch_hash_small(d, in, passes_ & 1, s));
    stage_end(s);
    int src = passes_ & 1;
    if (d.nstiles) {   // segments on the device-wide path (every segment has sort tiles but those above)
        stage_begin("k_hash_keys", "Solver_HashParticle", s);
        VTX_HIP(launch_hash_keys(d, in, s));
        stage_end(s);
        stage_begin("k_sort_hist+k_scan_reduce+k_scan_apply+k_sort_scatter", "Solver_HashSort", s);
        src = 0;
        for (int pass = 0; pass < passes_; pass++) {
            VTX_HIP(launch_sort_pass(d, src, 8 * pass, s));
            src ^= 1;
        }
        stage_end(s);
        stage_begin("k_clear_cells+k_cell_bounds+k_sorted_records", "Solver_HashBuildCell", s);
        VTX_HIP(launch_clear_cells(d, s));
        VTX_HIP(launch_cell_bounds(d, in, src, s));
        stage_end(s);
    }
    stage_begin("k_neighbors", "Solver_HashCache", s);
    VTX_HIP(launch_neighbors(d, in, tun_.nbr_order, tun_.gather_block, s));
    stage_end(s);
    sorted_src_ = src;
}

// VtClothSolverGPU::Simulate stage order (Velvet/VtClothSolverGPU.hpp:56-111) for one table set.
// In timing mode every launch group is bracketed by an event pair labelled with its kernel.
void Engine::record_group(const DevState& d, float frame_dt, hipStream_t s, hipEvent_t after_hash) {
    const float h = frame_dt / (float)substeps_;
    stage_begin("k_frame_begin", "Solver_CollideSDFs+Solver_Predict", s);
    VTX_HIP(launch_frame_begin(d, frame_dt, h, s));
    stage_end(s);
    if (after_hash && !any_selfcol_) VTX_HIP(hipEventRecord(after_hash, s));
    int cur = 0;
    for (int sub = 0; sub < substeps_; sub++) {
        if (any_selfcol_ && sub % interleave_ == 0) {
            hash_build(d, cur, s);
            if (after_hash && sub == 0) VTX_HIP(hipEventRecord(after_hash, s));
        }
        if (need_collide_) {
            stage_begin("k_collide", "Solver_CollideParticles+Solver_ApplyDeltas+Solver_CollideSDFs", s);
            VTX_HIP(launch_collide(d, cur, h, tun_.gather_block, s));
            stage_end(s);
            cur ^= 1;
        }
        const bool last_sub = sub == substeps_ - 1;
        const int epi_last = last_sub ? EPILOGUE_FINALIZE : EPILOGUE_FINALIZE_PREDICT;
        if (iterations_ == 0) {
            stage_begin("k_iterate", last_sub ? "Solver_Finalize" : "Solver_Finalize+Solver_Predict", s);
            VTX_HIP(launch_iterate(d, cur, h, epi_last, false, tun_.lra_batch, s));
            stage_end(s);
            cur ^= 1;
        }
        for (int it = 0; it < iterations_; it++) {
            const int epi = it == iterations_ - 1 ? epi_last : EPILOGUE_NONE;
            stage_begin("k_iterate",
                        epi == EPILOGUE_NONE
                            ? "Solver_SolveStretch+Solver_SolveAttach+Solver_SolveBending+Solver_ApplyDeltas"
                        : epi == EPILOGUE_FINALIZE
                            ? "Solver_SolveStretch+Solver_SolveAttach+Solver_SolveBending+Solver_ApplyDeltas+"
                              "Solver_Finalize"
                            : "Solver_SolveStretch+Solver_SolveAttach+Solver_SolveBending+Solver_ApplyDeltas+"
                              "Solver_Finalize+Solver_Predict",
                        s);
            VTX_HIP(launch_iterate(d, cur, h, epi, true, tun_.lra_batch, s));
            stage_end(s);
            cur ^= 1;
        }
    }
    stage_begin("k_normals", "Solver_UpdateNormals", s);
    VTX_HIP(launch_normals(d, s));
    stage_end(s);
}

// One frame: with several table sets, group k > 0 runs on stream k of its own. Joined (graph
// capture, timing): forked from and joined back into the solver's stream, so every later
// operation on it sees all groups. Lazy: the extra streams start once group 0 has built its
// hash, so that their hash and neighbour phases overlap group 0's iterations rather than its
// own, and the solver's stream does not wait for them at the end of the frame; join() adds
// that wait, and every operation other than the next frame that touches device state calls it.
void Engine::record_frame(float frame_dt, hipStream_t s, bool lazy) {
    if (groups_.size() < 2) {
        record_group(groups_.empty() ? ds_ : groups_[0], frame_dt, s);
        return;
    }
    if (lazy) {
        record_group(groups_[0], frame_dt, s, ev_fork_);
    } else {
        VTX_HIP(hipEventRecord(ev_fork_, s));
        record_group(groups_[0], frame_dt, s);
    }
    pending_join_ = groups_.size() - 1;
    try {
        for (size_t k = 1; k < groups_.size(); k++) {
            VTX_HIP(hipStreamWaitEvent(xstreams_[k - 1], ev_fork_, 0));
            record_group(groups_[k], frame_dt, xstreams_[k - 1]);
            VTX_HIP(hipEventRecord(ev_join_[k - 1], xstreams_[k - 1]));
        }
    } catch (...) {
        // a launch failed part-way: whatever each stream holds is still joined later
        if (lazy)
            for (size_t k = 1; k < groups_.size(); k++) (void)hipEventRecord(ev_join_[k - 1], xstreams_[k - 1]);
        throw;
    }
    if (!lazy) join();
}

void Engine::join() {
    for (size_t k = 0; k < pending_join_; k++) VTX_HIP(hipStreamWaitEvent(stream_, ev_join_[k], 0));
    pending_join_ = 0;
}

void Engine::simulate(float frame_dt) {
    VTX_REQUIRE(frame_dt > 0.f, VTX_EINVAL, "frame_dt must be positive");
    ensure_committed();
    if (ds_.nparticles == 0) return;
    frames_++;
    if (timing_) {
        stage_used_ = 0;
        join();
        record_frame(frame_dt, stream_, false);
        sync_stream();
        for (size_t k = 0; k < stage_used_; k++) {
            float ms = 0.f;
            VTX_HIP(hipEventElapsedTime(&ms, stages_[k].a, stages_[k].b));
            timing_ms_[stages_[k].label] += ms;
            timing_n_[stages_[k].label] += 1;
            timing_kernel_[stages_[k].label] = stages_[k].kernel;
            timing_ms_["total"] += ms;
            timing_n_["total"] += 1;
        }
    } else if (tun_.graph) {
        if (!gexec_ || graph_dt_ != frame_dt) {
            invalidate_graph();
            hipGraph_t graph = nullptr;
            join();
            VTX_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                record_frame(frame_dt, stream_, false);
            } catch (...) {
                (void)hipStreamEndCapture(stream_, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            VTX_HIP(hipStreamEndCapture(stream_, &graph));
            hipError_t e = hipGraphInstantiate(&gexec_, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            VTX_HIP(e);
            graph_dt_ = frame_dt;
        }
        join();
        VTX_HIP(hipGraphLaunch(gexec_, stream_));
    } else {
        // a caller's stream keeps its order: work queued on it after the frame sees every group
        record_frame(frame_dt, stream_, own_stream_);
    }
}

void Engine::sync_stream() {
    join();
    VTX_HIP(hipStreamSynchronize(stream_));
}

void Engine::synchronize() { sync_stream(); }

void Engine::set_use_graph(bool on) {
    join();
    tun_.graph = on;
    invalidate_graph();
}

void Engine::set_tuning(const Tuning& t) {
    join();
    VTX_REQUIRE(t.lra_batch == 4 || t.lra_batch == 16 || t.lra_batch == 64, VTX_EINVAL,
                "lra_batch must be 4, 16 or 64");
    VTX_REQUIRE(t.nbr_order >= 0 && t.nbr_order <= 2, VTX_EINVAL, "neighbor_order must be 0, 1 or 2");
    VTX_REQUIRE(t.gather_block == 64 || t.gather_block == 128 || t.gather_block == 256, VTX_EINVAL,
                "gather_block must be 64, 128 or 256");
    VTX_REQUIRE(t.streams >= 1 && t.streams <= VTX_MAX_STREAMS, VTX_EINVAL, "streams must be 1 to 4");
    VTX_REQUIRE(t.hash_grid >= 0 && t.hash_grid <= 3, VTX_EINVAL, "hash_grid must be 0 to 3");
    if (t.streams != tun_.streams || t.lra_split != tun_.lra_split || t.hash_lds != tun_.hash_lds ||
        t.hash_grid != tun_.hash_grid)
        tables_dirty_ = true;
    tun_ = t;
    ds_.remap = t.remap ? 1u : 0u;
    for (DevState& g : groups_) g.remap = ds_.remap;
    invalidate_graph();
}

// Shape class of the live set: log2 buckets of particles and segments, whether the hash,
// the stencil path and dense attachments are present, and the substep / iteration counts.
std::string Engine::shape_class() {
    ensure_committed();
    uint32_t nseg = 0, ngrid = 0;
    for (auto& sg : segs_)
        if (sg && sg->active && sg->uploaded && sg->n()) {
            nseg++;
            if (sg->prep.grid) ngrid++;
        }
    std::ostringstream o;
    const uint32_t np = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, live_particles_), UINT32_MAX);
    o << "p" << ceil_log2(np) << ".s" << ceil_log2(std::max(1u, nseg))
      << ".h" << (any_selfcol_ ? (any_grid_tunable_ ? 2 : 1) : 0) << ".g" << (ngrid ? 1 : 0) << ".l"
      << (any_lra_split_ ? 2 : any_lra_dense_ ? 1 : 0) << "."
      << substeps_ << "x" << iterations_;
    return o.str();
}

namespace {
struct TunedClass {
    Tuning t;
    double ms;
};
std::mutex g_tuned_mu;
std::map<std::string, TunedClass> g_tuned;   // shape class -> tuning, shared by the process's solvers
}  // namespace

Tuning Engine::autotune(int frames, float frame_dt, bool use_cache, double* frame_ms) {
    join();
    VTX_REQUIRE(frames >= 1 && frame_dt > 0.f, VTX_EINVAL, "autotune needs frames >= 1 and frame_dt > 0");
    ensure_committed();
    if (frame_ms) *frame_ms = 0.0;
    if (ds_.nparticles == 0) return tun_;
    const std::string key = shape_class();
    if (use_cache) {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            set_tuning(it->second.t);
            return tun_;
        }
    }
    const size_t n = ap_.cap;
    float4 *sx = nullptr, *sv = nullptr, *sn = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timing = timing_;
    const uint64_t frames_before = frames_;
    timing_ = false;
    auto cleanup = [&] {
        for (size_t k = 0; k < pending_join_; k++) (void)hipStreamSynchronize(xstreams_[k]);
        pending_join_ = 0;
        (void)hipStreamSynchronize(stream_);
        if (sx) (void)hipFree(sx);
        if (sv) (void)hipFree(sv);
        if (sn) (void)hipFree(sn);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        timing_ = timing;
        frames_ = frames_before;
    };
    Tuning best = tun_;
    double bms = 0.0;
    try {
        VTX_HIP(hipMalloc((void**)&sx, n * sizeof(float4)));
        VTX_HIP(hipMalloc((void**)&sv, n * sizeof(float4)));
        VTX_HIP(hipMalloc((void**)&sn, n * sizeof(float4)));
        VTX_HIP(hipEventCreate(&e0));
        VTX_HIP(hipEventCreate(&e1));
        // positions and velocities are the state a frame starts from; normals are its output
        VTX_HIP(hipMemcpyAsync(sx, pl_.X, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
        VTX_HIP(hipMemcpyAsync(sv, pl_.V, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
        VTX_HIP(hipMemcpyAsync(sn, pl_.N, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
        auto restore = [&] {
            join();
            VTX_HIP(hipMemcpyAsync(pl_.X, sx, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
            VTX_HIP(hipMemcpyAsync(pl_.V, sv, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
            VTX_HIP(hipMemcpyAsync(pl_.N, sn, n * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
        };
        auto trial = [&](const Tuning& t) {
            set_tuning(t);
            simulate(frame_dt);                 // warm-up, records the graph
            restore();
            VTX_HIP(hipEventRecord(e0, stream_));
            for (int f = 0; f < frames; f++) simulate(frame_dt);
            join();
            VTX_HIP(hipEventRecord(e1, stream_));
            restore();
            sync_stream();
            float ms = 0.f;
            VTX_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / frames;
        };
        bms = trial(best);
        // one knob at a time, every other value of it against the best so far; the hash grid
        // (opted-in segments only) goes before the knobs of the hash and neighbour phases
        static const int order[] = {0, 8, 1, 2, 3, 4, 5, 6, 7};
        for (int knob : order) {
            if (knob == 8 && !any_grid_tunable_) continue;
            if (knob == 1 && !any_selfcol_) continue;
            if (knob == 2 && !any_lra_dense_) continue;
            if (knob == 4 && !need_collide_) continue;
            if (knob == 5 && table_seg_.size() < 2) continue;
            if (knob == 6 && !any_lra_split_) continue;
            if (knob == 7 && !any_hash_lds_) continue;
            const int nval = knob == 5 || knob == 8 ? 4 : knob == 1 || knob == 2 || knob == 4 ? 3 : 2;
            const Tuning base = best;
            for (int v = 0; v < nval; v++) {
                Tuning t = base;
                if (knob == 0) t.graph = v != 0;
                if (knob == 1) t.nbr_order = v;
                if (knob == 2) t.lra_batch = v == 0 ? 4 : v == 1 ? 16 : 64;
                if (knob == 3) t.remap = v != 0;
                if (knob == 4) t.gather_block = 64 << v;
                if (knob == 5) t.streams = v + 1;
                if (knob == 6) t.lra_split = v != 0;
                if (knob == 7) t.hash_lds = v != 0;
                if (knob == 8) t.hash_grid = v;
                if (knob == 8 ? t.hash_grid == base.hash_grid
                    : knob == 0 ? t.graph == base.graph : knob == 1 ? t.nbr_order == base.nbr_order
                                : knob == 2 ? t.lra_batch == base.lra_batch
                                : knob == 3 ? t.remap == base.remap
                                : knob == 4 ? t.gather_block == base.gather_block
                                : knob == 5 ? t.streams == base.streams
                                : knob == 6 ? t.lra_split == base.lra_split : t.hash_lds == base.hash_lds)
                    continue;
                const double ms = trial(t);
                if (ms < 0.99 * bms) {
                    best = t;
                    bms = ms;
                }
            }
        }
        set_tuning(best);
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        g_tuned[key] = TunedClass{best, bms};
    }
    if (frame_ms) *frame_ms = bms;
    return tun_;
}

void Engine::set_stream(hipStream_t s) {
    join();
    synchronize();
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    if (s) {
        stream_ = s;
        own_stream_ = false;
    } else {
        VTX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        own_stream_ = true;
    }
    invalidate_graph();
}

void Engine::set_timing(bool on) {
    join();
    if (on != timing_) tables_dirty_ = true;   // timing mode runs all segments as one table set
    timing_ = on;
    timing_ms_.clear();
    timing_n_.clear();
    timing_kernel_.clear();
}

std::string Engine::timing_report() {
    std::ostringstream o;
    o << "{";
    bool first = true;
    o.precision(9);
    for (auto& kv : timing_ms_) {
        o << (first ? "" : ", ") << "\"" << kv.first << "\": {\"ms\": " << kv.second
          << ", \"launches\": " << timing_n_[kv.first];
        auto k = timing_kernel_.find(kv.first);
        if (k != timing_kernel_.end()) o << ", \"kernel\": \"" << k->second << "\"";
        o << "}";
        first = false;
    }
    o << "}";
    return o.str();
}

// ---------------------------------------------------------------------------- state access
void Engine::read_particles(int32_t s, int what, float* out) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(out, VTX_EINVAL, "output is NULL");
    ensure_committed();
    const uint32_t n = sg.n();
    if (!n) return;
    if (!sg.uploaded) {   // pending segment: host state
        for (uint32_t i = 0; i < n; i++) {
            if (what == 3) out[i] = sg.w[i];
            else
                for (int k = 0; k < 3; k++)
                    out[3 * i + k] = what == 0 ? sg.x[3 * i + k] : what == 1 ? sg.v[3 * i + k] : 0.f;
        }
        return;
    }
    synchronize();
    std::vector<float4> tmp(n);
    const float4* src = what == 0 ? pl_.X : what == 1 ? pl_.V : what == 2 ? pl_.N : pl_.X;
    VTX_HIP(hipMemcpy(tmp.data(), src + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
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
    join();
    Segment& sg = seg(s);
    ensure_committed();
    VTX_REQUIRE(sg.uploaded, VTX_ESTATE, "segment is not resident on the device");
    if (n) *n = sg.n();
    return pl_.X + sg.rg.p;
}

float4* Engine::device_normals(int32_t s, uint32_t* n) {
    join();
    Segment& sg = seg(s);
    ensure_committed();
    VTX_REQUIRE(sg.uploaded, VTX_ESTATE, "segment is not resident on the device");
    if (n) *n = sg.n();
    return pl_.N + sg.rg.p;
}

// MouseGrabber-style edit (Velvet/MouseGrabber.hpp:51-52,60,74-75).
void Engine::set_particle(int32_t s, int32_t i, const float* pos, const float* vel, float inv_mass) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(i >= 0 && (uint32_t)i < sg.n(), VTX_EINVAL, "particle index out of range");
    if (!sg.uploaded) {
        if (pos) std::copy(pos, pos + 3, sg.x.begin() + 3 * i);
        if (vel) std::copy(vel, vel + 3, sg.v.begin() + 3 * i);
        if (inv_mass >= 0.f) sg.w[(size_t)i] = inv_mass;
        sg.prepared = false;
        return;
    }
    synchronize();
    const size_t g = sg.rg.p + (size_t)i;
    float4 x;
    VTX_HIP(hipMemcpy(&x, pl_.X + g, sizeof(float4), hipMemcpyDeviceToHost));
    if (pos) {
        x.x = pos[0];
        x.y = pos[1];
        x.z = pos[2];
    }
    if (inv_mass >= 0.f) x.w = inv_mass;
    VTX_HIP(hipMemcpy(pl_.X + g, &x, sizeof(float4), hipMemcpyHostToDevice));
    if (vel) {
        const float4 v = make_float4(vel[0], vel[1], vel[2], 0.f);
        VTX_HIP(hipMemcpy(pl_.V + g, &v, sizeof(float4), hipMemcpyHostToDevice));
    }
}

void Engine::write_predicted(int32_t s, const float* xyz) {
    join();
    Segment& sg = seg(s);
    VTX_REQUIRE(xyz, VTX_EINVAL, "positions are NULL");
    ensure_committed();
    VTX_REQUIRE(sg.uploaded, VTX_ESTATE, "segment is not resident on the device");
    synchronize();
    const uint32_t n = sg.n();
    if (!n) return;
    std::vector<float4> tmp(n);
    VTX_HIP(hipMemcpy(tmp.data(), pl_.X + sg.rg.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) {
        tmp[i].x = xyz[3 * i];
        tmp[i].y = xyz[3 * i + 1];
        tmp[i].z = xyz[3 * i + 2];
    }
    VTX_HIP(hipMemcpy(pl_.P0 + sg.rg.p, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
}

void Engine::run_hash() {
    join();
    ensure_committed();
    VTX_REQUIRE(any_selfcol_, VTX_ESTATE, "no active segment has self collision enabled");
    for (const DevState& g : groups_) hash_build(g, 0, stream_);
    synchronize();
}

uint32_t Engine::hash_table_size(int32_t s) {
    Segment& sg = seg(s);
    ensure_committed();
    return sg.uploaded ? sg.dev.M : 0;
}

void Engine::read_hash(int32_t s, uint32_t* cs, uint32_t* ce, uint32_t* sh, uint32_t* si, uint32_t* nbr) {
    join();
    Segment& sg = seg(s);
    ensure_committed();
    synchronize();
    VTX_REQUIRE(sg.uploaded && (sg.dev.flags & VTX_SEG_SELFCOL), VTX_ESTATE, "segment has no spatial hash");
    const SegDev& d = sg.dev;
    if (cs || ce) {
        std::vector<uint2> tmp(d.M);
        VTX_HIP(hipMemcpy(tmp.data(), pl_.cells + d.tbase, sizeof(uint2) * d.M, hipMemcpyDeviceToHost));
        for (uint32_t h = 0; h < d.M; h++) {
            if (cs) cs[h] = tmp[h].x;
            if (ce) ce[h] = tmp[h].y;
        }
    }
    const uint32_t* keys = sorted_src_ ? pl_.k1 : pl_.k0;
    const uint32_t* vals = sorted_src_ ? pl_.v1 : pl_.v0;
    if (sh) VTX_HIP(hipMemcpy(sh, keys + d.pbase, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost));
    if (si) VTX_HIP(hipMemcpy(si, vals + d.pbase, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost));
    if (nbr) {   // device chunks (entries 4c..4c+3 of particle i at 4 (c n + i)) to column-major
        const uint64_t nb = (uint64_t)d.nbase_hi << 32 | d.nbase_lo;
        const size_t K = d.max_nbr, n = d.n, K4 = (K + 3) & ~size_t(3);
        if (d.flags & VTX_SEG_NBR16) {
            std::vector<uint16_t> tmp(K4 * n);
            VTX_HIP(hipMemcpy(tmp.data(), pl_.nbr + nb, sizeof(uint16_t) * tmp.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < K; k++)
                for (size_t i = 0; i < n; i++) {
                    const uint16_t v = tmp[4 * ((k >> 2) * n + i) + (k & 3)];
                    nbr[k * n + i] = v == 0xFFFFu ? VTX_INVALID : v;
                }
        } else {
            std::vector<uint32_t> tmp(K4 * n);
            VTX_HIP(hipMemcpy(tmp.data(), pl_.nbr + nb, sizeof(uint32_t) * tmp.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < K; k++)
                for (size_t i = 0; i < n; i++) nbr[k * n + i] = tmp[4 * ((k >> 2) * n + i) + (k & 3)];
        }
    }
}

std::vector<int32_t> Engine::nonfinite_segments() {
    join();
    ensure_committed();
    std::vector<int32_t> bad;
    if (table_seg_.empty()) return bad;
    VTX_HIP(launch_check_finite(ds_, flags_, (uint32_t)table_seg_.size(), stream_));
    std::vector<uint32_t> h(table_seg_.size());
    VTX_HIP(hipMemcpyAsync(h.data(), flags_, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (size_t k = 0; k < h.size(); k++)
        if (h[k]) bad.push_back(table_seg_[k]);
    return bad;
}

}  // namespace vtx
