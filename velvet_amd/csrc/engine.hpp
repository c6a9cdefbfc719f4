// Host engine behind the C ABI: registration, device pools, frame recording, graphs.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vtx.h"
#include "vtx_device.hpp"

namespace vtx {

struct TableSet;   // host images of the launch tables (engine.cpp)

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define VTX_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::vtx::Error(VTX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define VTX_REQUIRE(cond, code, msg)                    \
    do {                                                \
        if (!(cond)) throw ::vtx::Error((code), (msg)); \
    } while (0)

// First-fit allocator of index ranges [0, cap) with coalescing frees.
class RangeAlloc {
public:
    size_t cap = 0;
    bool take(size_t n, size_t* off);          // false if no free range is large enough
    void give(size_t off, size_t n);
    void grow(size_t new_cap);                 // extends the free space at the end
    size_t used() const { return used_; }
private:
    std::map<size_t, size_t> free_;            // offset -> length
    size_t used_ = 0;
};

// A segment's device data, prepared on the host with segment-local indices and offsets, so it
// can be uploaded into any ranges of the pools without rewriting.
struct Prepared {
    bool grid = false;
    uint32_t grows = 0, gcols = 0, grid_flags = 0;
    std::vector<float4> x, v, x0;
    std::vector<uint2> est, eat;
    std::vector<uint32_t> ebd, etr;
    std::vector<uint4> brec, trec;
    std::vector<float> brest;
    std::vector<float4> slots, rest4, brest4;
    std::vector<uint32_t> gflags;
    std::vector<ColliderDev> cols;
    std::vector<float> lra;                     // dense attachments: distances, slot-major
    bool lra_dense = false;
    std::vector<TileDev> tiles;                 // ELL offsets relative to the segment's ranges
    std::vector<GridTileDev> gtiles;
};

// Ranges a segment occupies in the device pools (offset, length).
struct Ranges {
    size_t p = 0, cell = 0, nbr = 0, st = 0, at = 0, bd = 0, tr = 0, brec = 0, trec = 0, slot = 0, col = 0, lra = 0;
    size_t n_p = 0, n_cell = 0, n_nbr = 0, n_st = 0, n_at = 0, n_bd = 0, n_tr = 0, n_brec = 0, n_trec = 0,
           n_slot = 0, n_col = 0, n_lra = 0;
};

// One independent simulation: the state of one reference VtClothSolverGPU
// (Velvet/VtClothSolverGPU.hpp:207-231).
// Device -> pinned-host copies of a segment's positions and normals for a viewer, in three
// buffers: each request fills the least recently requested buffer other than the one last
// handed to the viewer, after the work queued so far, so the newest request and the buffer the
// viewer holds are never overwritten.
struct Readback {
    static constexpr int kBuffers = 3;
    float4* host[kBuffers] = {};             // n positions (w = inverse mass), then n normals
    hipEvent_t ev[kBuffers] = {};
    uint64_t frame[kBuffers] = {};
    uint64_t seq[kBuffers] = {};             // request order (0 = never requested)
    bool pending[kBuffers] = {};
    uint32_t n = 0;
    uint64_t nseq = 0;
    int handed = -1;                         // buffer returned by the last readback_latest
    Readback() = default;
    Readback(const Readback&) = delete;
    Readback& operator=(const Readback&) = delete;
    ~Readback();
};

struct Segment {
    vtx_params prm{};
    std::vector<float> x, v, w, x0;          // 3n, 3n, n, 3n (host copy; see `uploaded`)
    std::vector<uint32_t> tris;              // 3T, local indices
    std::vector<int32_t> st;                 // 2S
    std::vector<float> st_rest;              // S
    std::vector<uint32_t> bd;                // 4B
    std::vector<float> bd_rest;              // B
    std::vector<int32_t> at_pid, at_slot;    // A
    std::vector<float> at_dist;              // A
    std::vector<float> slots;                // 3L
    std::vector<vtx_collider> cols;
    std::vector<uint32_t> cloth_starts;      // first particle of each AddCloth
    bool active = true;                      // simulated by the solver's frames
    bool uploaded = false;                   // device ranges hold the current state
    bool prepared = false;                   // `prep` matches the host data
    Prepared prep;
    Ranges rg;
    SegDev dev{};                            // absolute bases filled at upload
    std::unique_ptr<Readback> rb;            // viewer copies, created on first request
    uint32_t n() const { return (uint32_t)w.size(); }
    uint32_t table_size() const { return prm.hash_table_size > 0 ? (uint32_t)prm.hash_table_size : 2u * n(); }
    bool selfcol() const { return prm.enable_self_collision && n() > 0; }
    // Hash table cells of the grid for a Tuning::hash_grid value: the parameters' table size,
    // enlarged 2x / 4x at values 2 / 3 for segments that opted in (kept below 2^28).
    uint32_t grid_table_size(int hash_grid) const {
        uint64_t m = table_size();
        if (prm.tune_hash_grid && hash_grid >= 2) m <<= (hash_grid - 1);
        while (m >= (1u << 28) && m > table_size()) m >>= 1;
        return (uint32_t)m;
    }
    // cells reserved in the pool: room for every hash_grid value
    uint32_t cell_capacity() const { return selfcol() ? grid_table_size(3) : 0u; }
};

// Launch variants the tuner chooses among. Every combination gives bit-identical results: they
// change which workgroup takes which tile, the order threads take particles in, how many loads
// are in flight and whether launches are replayed from a graph, never the arithmetic.
struct Tuning {
    bool graph = true;        // replay frames from a hipGraph
    int nbr_order = 2;        // neighbour search thread order: 0 particle, 1 sorted, 2 grid patches
    int lra_batch = 16;       // dense attachment distances loaded per batch: 4, 16 or 64
    bool remap = true;        // XCD-aware workgroup -> tile remap
    int gather_block = 256;   // threads per workgroup of the neighbour search and contact pass
    int streams = 1;          // k > 1: segments split into k particle-balanced groups whose frames
                              // run on k streams, overlapping latency- and VALU-bound stages
    bool lra_split = true;    // grid segments with >= VTX_LRA_SPLIT_MIN dense slots: four threads
                              // per particle in the Jacobi pass
    bool hash_lds = true;     // small self-colliding segments build their hash in one workgroup
    int hash_grid = 0;        // hash-grid resolution of segments with tune_hash_grid: 0 the
                              // reference grid; 1..3 origin shifted, table 1x / 2x / 4x. The only
                              // field that changes results (list order, bucket sharing)
};

// A timed launch group: its kernel and the reference's timer labels of the stages it runs
// (Velvet/Timer.hpp:242-260 ScopedTimerGPU labels, Velvet/GUI.cpp:32-51), '+'-joined where one
// kernel fuses several of them.
struct StageTime {
    const char* kernel;
    const char* label;
    hipEvent_t a, b;
};

// Stretch (every unique edge) and bending (every edge of exactly two triangles, rest angle from
// the given positions) constraints of a triangle mesh; host only (mesh.cpp).
void mesh_constraints(const float* x, uint32_t n, const uint32_t* tris, uint32_t ntri,
                      std::vector<int32_t>& st_ij, std::vector<float>& st_rest,
                      std::vector<uint32_t>& bd_idx, std::vector<float>& bd_rest);

class Engine {
public:
    explicit Engine(const vtx_params& p);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int32_t add_segment(const vtx_params& p, bool active = true);
    void remove_segment(int32_t s);
    void set_active(int32_t s, bool active);
    void park(int32_t s);                    // deactivate and free its device memory, keep its state
    // Whole simulated state of a segment (positions + inverse masses, velocities, initial
    // positions, slot positions) to / from a self-describing host blob.
    size_t snapshot_bytes(int32_t s);
    void snapshot(int32_t s, void* buf, size_t bytes);
    void restore(int32_t s, const void* buf, size_t bytes);
    int32_t num_segments() const;
    Segment& seg(int32_t i);

    void set_params(int32_t s, const vtx_params& p);
    int32_t add_cloth(int32_t s, const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                      const float* model, float diameter);
    void add_stretch(int32_t s, const int32_t* ij, const float* rest, uint32_t count);
    int32_t add_attach_slot(int32_t s, const float* pos);
    void set_attach_slot_position(int32_t s, int32_t slot, const float* pos);
    void add_attach(int32_t s, const int32_t* pid, const int32_t* slot, const float* dist, uint32_t count);
    void add_bend(int32_t s, const uint32_t* idx, const float* angle, uint32_t count);
    // Stretch and bending constraints of the cloth added at `offset`, from its current positions.
    void add_mesh_constraints(int32_t s, int32_t offset, const uint32_t* tris, uint32_t ntri);
    void set_colliders(int32_t s, const vtx_collider* cols, uint32_t n);

    void prepare(int32_t s);                 // host packing ahead of upload (optional)
    // The same packing of one segment object; touches nothing else of the engine (it reads only
    // the stencil-path switch), so another thread may run it on a segment that is neither
    // active nor being registered while this engine runs frames.
    void prepare_segment(Segment& sg);
    static size_t bytes_estimate(const Segment& sg);   // device bytes of a prepared segment
    size_t device_bytes_estimate(int32_t s);
    void simulate(float frame_dt);
    void synchronize();
    // Make the solver's stream wait for the work the stream groups have queued.
    void join();
    void set_use_graph(bool on);
    const Tuning& tuning() const { return tun_; }
    void set_tuning(const Tuning& t);
    // Times the variants (coordinate descent, `frames` timed frames each, the state restored
    // after every trial so the simulation does not advance), applies the fastest and caches it
    // for the current shape class; with use_cache a class tuned before in this process is applied
    // without timing. Returns the applied tuning; *frame_ms = its frame time (0 from the cache
    // when it was not measured here).
    Tuning autotune(int frames, float frame_dt, bool use_cache, double* frame_ms);
    std::string shape_class();
    void set_stencil_path(bool on);
    int32_t stencil_segments();
    void set_stream(hipStream_t s);
    hipStream_t stream() const { return stream_; }
    void set_timing(bool on);
    std::string timing_report();

    void read_particles(int32_t s, int what, float* out);   // 0 x, 1 v, 2 normals, 3 inv mass
    float4* device_positions(int32_t s, uint32_t* n);
    float4* device_normals(int32_t s, uint32_t* n);
    void set_particle(int32_t s, int32_t i, const float* pos, const float* vel, float inv_mass);
    // Queue a copy of segment s's positions and normals into the older of its two pinned
    // buffers, after the work queued so far; the newest completed copy, without waiting
    // (false if none has completed).
    void readback_request(int32_t s);
    bool readback_latest(int32_t s, const float4** pos, const float4** nrm, uint32_t* n, uint64_t* frame);

    void write_predicted(int32_t s, const float* xyz);
    void run_hash();
    uint32_t hash_table_size(int32_t s);
    void read_hash(int32_t s, uint32_t* cs, uint32_t* ce, uint32_t* sh, uint32_t* si, uint32_t* nbr);

    // Segments with a non-finite position or velocity after the last frame.
    std::vector<int32_t> nonfinite_segments();
    uint64_t live_particles() const { return live_particles_; }
    size_t pool_bytes() const;
    // Upload active segments not yet on the device and rebuild the launch tables, now rather
    // than at the next frame (returns when the uploads are done).
    void commit() { ensure_committed(); }
    // Upload an inactive segment into free pool ranges on a copy stream of its own, so that it
    // can overlap a frame of the active segments (their ranges are disjoint); returns when the
    // copies are done. If the pools must grow first, the running frame is waited for (growth
    // moves every pool). The segment stays inactive; activating it later only rebuilds tables.
    void stage(int32_t s);
    // Grow the pools so that segments totalling `particles` particles (grid cloths, default hash
    // table and neighbour cap) fit without a later growth.
    void reserve(uint64_t particles);

private:
    void pull_state(Segment& sg);
    void touch(int32_t s);                   // a structural change of segment s
    void invalidate_graph();
    void sync_stream();                      // join(), then wait for the solver's stream
    void ensure_committed();
    void upload(Segment& sg, hipStream_t s);
    bool pools_fit(const Ranges& need);
    void release(Segment& sg);
    void rebuild_tables();
    void upload_tables(const TableSet& t);
    void grow_pools(const Ranges& need);
    void record_frame(float frame_dt, hipStream_t s, bool lazy);
    void hash_build(const DevState& d, int in, hipStream_t s);
    void record_group(const DevState& d, float frame_dt, hipStream_t s, hipEvent_t after_hash = nullptr);
    void stage_begin(const char* kernel, const char* label, hipStream_t s);
    void stage_end(hipStream_t s);
    void check_seg(int32_t s) const;
    void free_tables();
    void free_pools();

    std::vector<std::unique_ptr<Segment>> segs_;
    bool tables_dirty_ = true;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;

    // device pools: every array of a pool is indexed by that pool's allocator
    RangeAlloc ap_, acell_, anbr_, ast_, aat_, abd_, atr_, abrec_, atrec_, aslot_, acol_, alra_;
    struct Pools {
        float4 *X = nullptr, *V = nullptr, *P0 = nullptr, *P1 = nullptr, *X0 = nullptr, *N = nullptr,
               *sPX = nullptr, *rest4 = nullptr, *brest4 = nullptr;
        uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *gflags = nullptr;
        uint2* cells = nullptr;
        uint32_t* nbr = nullptr;
        uint2 *st = nullptr, *at = nullptr;
        uint32_t *bd = nullptr, *tr = nullptr;
        uint4 *brec = nullptr, *trec = nullptr;
        float* brest = nullptr;
        float4* slots = nullptr;
        ColliderDev* cols = nullptr;
        float* lra = nullptr;
    } pl_;
    uint8_t* tab_dev_ = nullptr;    // launch tables (segments, tiles, sort tiles), one arena
    uint8_t* tab_host_ = nullptr;   // pinned staging of the same size
    size_t tab_cap_ = 0;
    size_t hist_cap_ = 0, part_cap_ = 0, flag_cap_ = 0;
    uint32_t* hist_ = nullptr;
    uint32_t* part_ = nullptr;
    uint32_t* flags_ = nullptr;

    DevState ds_{};                          // every live segment
    std::vector<DevState> groups_;           // the frame's table sets, one per stream
    std::vector<hipStream_t> xstreams_;      // streams of groups 1.. (created on demand)
    hipStream_t cstream_ = nullptr;   // uploads of staged segments
    hipEvent_t ev_fork_ = nullptr;
    std::vector<hipEvent_t> ev_join_;        // one per extra stream
    size_t pending_join_ = 0;                // extra streams whose last frame the solver's stream has not joined
    std::vector<int32_t> table_seg_;         // table index -> segment id
    int passes_ = 1;
    int sorted_src_ = 0;
    bool any_selfcol_ = false;
    bool need_collide_ = false;
    int substeps_ = 1, iterations_ = 0, interleave_ = 1;
    uint64_t live_particles_ = 0;
    uint64_t frames_ = 0;                    // frames simulated
    Tuning tun_;
    bool any_lra_dense_ = false;
    bool any_lra_split_ = false;   // a stencil segment has >= VTX_LRA_SPLIT_MIN dense slots
    bool any_hash_lds_ = false;    // a self-colliding segment is small enough for k_hash_small
    bool any_grid_tunable_ = false;   // a live self-colliding segment has tune_hash_grid
    bool use_stencil_ = true;
    hipGraphExec_t gexec_ = nullptr;
    float graph_dt_ = -1.f;

    bool timing_ = false;
    std::vector<StageTime> stages_;
    size_t stage_used_ = 0;
    std::map<std::string, double> timing_ms_;
    std::map<std::string, long> timing_n_;
    std::map<std::string, std::string> timing_kernel_;
};

}  // namespace vtx
