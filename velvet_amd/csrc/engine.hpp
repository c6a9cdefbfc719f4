// Host engine behind the C ABI: registration, device layout, frame recording, graphs.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vtx.h"
#include "vtx_device.hpp"

namespace vtx {

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

#define VTX_REQUIRE(cond, code, msg)                  \
    do {                                              \
        if (!(cond)) throw ::vtx::Error((code), (msg)); \
    } while (0)

// One independent simulation: the state of one reference VtClothSolverGPU
// (Velvet/VtClothSolverGPU.hpp:207-231), kept on the host until it is committed.
struct Segment {
    vtx_params prm{};
    std::vector<float> x, v, w, x0;          // 3n, 3n, n, 3n
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
    uint32_t n() const { return (uint32_t)w.size(); }
    uint32_t table_size() const {
        return prm.hash_table_size > 0 ? (uint32_t)prm.hash_table_size : 2u * n();
    }
};

struct StageTime {
    std::string label;
    hipEvent_t a, b;
};

class Engine {
public:
    explicit Engine(const vtx_params& p);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int32_t add_segment(const vtx_params& p);
    int32_t num_segments() const { return (int32_t)segs_.size(); }
    Segment& seg(int32_t i);

    void set_params(int32_t s, const vtx_params& p);
    int32_t add_cloth(int32_t s, const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                      const float* model, float diameter);
    void add_stretch(int32_t s, const int32_t* ij, const float* rest, uint32_t count);
    int32_t add_attach_slot(int32_t s, const float* pos);
    void set_attach_slot_position(int32_t s, int32_t slot, const float* pos);
    void add_attach(int32_t s, const int32_t* pid, const int32_t* slot, const float* dist, uint32_t count);
    void add_bend(int32_t s, const uint32_t* idx, const float* angle, uint32_t count);
    void set_colliders(int32_t s, const vtx_collider* cols, uint32_t n);

    void simulate(float frame_dt);
    void synchronize();
    void set_use_graph(bool on);
    void set_stencil_path(bool on);
    int32_t stencil_segments();
    void set_stream(hipStream_t s);
    void set_timing(bool on);
    std::string timing_report();

    void read_particles(int32_t s, int what, float* out);   // 0 x, 1 v, 2 normals, 3 inv mass
    float4* device_positions(int32_t s, uint32_t* n);
    float4* device_normals(int32_t s, uint32_t* n);
    void set_particle(int32_t s, int32_t i, const float* pos, const float* vel, float inv_mass);

    void write_predicted(int32_t s, const float* xyz);
    void run_hash();
    uint32_t hash_table_size(int32_t s);
    void read_hash(int32_t s, uint32_t* cs, uint32_t* ce, uint32_t* sh, uint32_t* si, uint32_t* nbr);

    // layout facts for measurement
    uint32_t total_particles() const { return ds_.nparticles; }
    uint32_t sort_passes() const { return (uint32_t)passes_; }

private:
    void pull_state();
    void mark_dirty();
    void ensure_committed();
    void commit();
    void free_device();
    void upload_colliders();
    void record_frame(float frame_dt, hipStream_t s);
    void hash_build(int in, hipStream_t s);
    void stage_begin(const char* label, hipStream_t s);
    void stage_end(hipStream_t s);
    void check_seg(int32_t s) const;
    template <class T>
    T* dalloc(size_t count);
    template <class T>
    T* dupload(const std::vector<T>& h);

    std::vector<Segment> segs_;
    bool dirty_ = true;        // host layout changed since the last commit
    bool dev_live_ = false;    // device holds the authoritative particle state
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;

    DevState ds_{};
    std::vector<void*> allocs_;
    std::vector<SegDev> hseg_;
    std::vector<TileDev> htile_;
    int passes_ = 1;
    int sorted_src_ = 0;
    bool any_selfcol_ = false;
    bool need_collide_ = false;
    int substeps_ = 1, iterations_ = 0, interleave_ = 1;

    bool use_graph_ = true;
    bool use_stencil_ = true;
    std::vector<uint8_t> seg_grid_;
    hipGraphExec_t gexec_ = nullptr;
    float graph_dt_ = -1.f;

    bool timing_ = false;
    std::vector<StageTime> stages_;
    size_t stage_used_ = 0;
    std::map<std::string, double> timing_ms_;
    std::map<std::string, long> timing_n_;
};

}  // namespace vtx
