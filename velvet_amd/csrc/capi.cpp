// C ABI of the solver (declared in include/vtx.h): argument checks, error codes and the
// per-thread error message around the C++ engine.
#include <cstring>
#include <new>
#include <string>

#include "../../include/vtx.h"
#include "engine.hpp"
#include "launch.hpp"
#include "server.hpp"

struct vtx_solver {
    vtx::Engine engine;
    explicit vtx_solver(const vtx_params& p) : engine(p) {}
};

struct vtx_server {
    vtx_solver solver;
    vtx::Server server;
    vtx_server(const vtx_params& p, const vtx_server_config& c) : solver(p), server(solver.engine, c) {}
};

namespace {
thread_local std::string g_error;
void set_error(const char* m) { g_error = m ? m : ""; }
}  // namespace

#define API_BEGIN \
    try {
#define API_END                                   \
    }                                             \
    catch (const vtx::Error& e) {                 \
        set_error(e.what());                      \
        return e.code;                            \
    }                                             \
    catch (const std::bad_alloc&) {               \
        set_error("host allocation failed");      \
        return VTX_ENOMEM;                        \
    }                                             \
    catch (const std::exception& e) {             \
        set_error(e.what());                      \
        return VTX_EINVAL;                        \
    }                                             \
    return VTX_OK;

#define NEED(p)                                  \
    do {                                         \
        if (!(p)) {                              \
            set_error(#p " is NULL");            \
            return VTX_EINVAL;                   \
        }                                        \
    } while (0)

extern "C" {

void vtx_params_default(vtx_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->num_substeps = 2;          // Velvet/Common.hpp:21-47
    p->num_iterations = 4;
    p->max_num_neighbors = 64;
    p->max_speed = 50.f;
    p->gravity[0] = 0.f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.f;
    p->bend_compliance = 0.f;
    p->damping = 0.25f;
    p->relaxation_factor = 1.f;
    p->long_range_stretchiness = 1.2f;
    p->collision_margin = 0.06f;
    p->friction = 0.1f;
    p->enable_self_collision = 1;
    p->interleaved_hash = 3;
    p->particle_diameter = 0.f;
    p->particle_diameter_scalar = 1.5f;
    p->hash_cell_size_scalar = 1.5f;
    p->hash_table_size = 0;
    p->stretch_stiffness = 1.f;
    p->tune_hash_grid = 0;
}

const char* vtx_last_error(void) { return g_error.c_str(); }
const char* vtx_version(void) { return "vtx 0.2 gfx950"; }
int vtx_abi_version(void) { return VTX_ABI_VERSION; }

int vtx_solver_create(const vtx_params* params, vtx_solver** out) {
    NEED(out);
    API_BEGIN
    vtx_params p;
    if (params) p = *params;
    else vtx_params_default(&p);
    *out = new vtx_solver(p);
    API_END
}

int vtx_solver_destroy(vtx_solver* s) {
    API_BEGIN
    delete s;
    API_END
}

int vtx_set_params(vtx_solver* s, int32_t seg, const vtx_params* params) {
    NEED(s); NEED(params);
    API_BEGIN s->engine.set_params(seg, *params); API_END
}

int vtx_get_params(vtx_solver* s, int32_t seg, vtx_params* out) {
    NEED(s); NEED(out);
    API_BEGIN *out = s->engine.seg(seg).prm; API_END
}

int vtx_add_segment(vtx_solver* s, const vtx_params* params, int32_t* seg_out) {
    NEED(s); NEED(params);
    API_BEGIN
    int32_t k = s->engine.add_segment(*params);
    if (seg_out) *seg_out = k;
    API_END
}

int vtx_remove_segment(vtx_solver* s, int32_t seg) {
    NEED(s);
    API_BEGIN s->engine.remove_segment(seg); API_END
}

int vtx_set_segment_active(vtx_solver* s, int32_t seg, int32_t active) {
    NEED(s);
    API_BEGIN s->engine.set_active(seg, active != 0); API_END
}

int vtx_num_segments(vtx_solver* s, int32_t* out) {
    NEED(s); NEED(out);
    API_BEGIN *out = s->engine.num_segments(); API_END
}

int vtx_add_cloth(vtx_solver* s, int32_t seg, const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                  const float* model, float diameter, int32_t* offset) {
    NEED(s);
    API_BEGIN
    int32_t o = s->engine.add_cloth(seg, xyz, n, tris, ntri, model, diameter);
    if (offset) *offset = o;
    API_END
}

int vtx_add_stretch(vtx_solver* s, int32_t seg, int32_t i, int32_t j, float rest) {
    NEED(s);
    int32_t ij[2] = {i, j};
    API_BEGIN s->engine.add_stretch(seg, ij, &rest, 1); API_END
}

int vtx_add_stretch_n(vtx_solver* s, int32_t seg, const int32_t* ij, const float* rest, uint32_t count) {
    NEED(s);
    API_BEGIN s->engine.add_stretch(seg, ij, rest, count); API_END
}

int vtx_add_attach_slot(vtx_solver* s, int32_t seg, const float pos[3], int32_t* slot) {
    NEED(s);
    API_BEGIN
    int32_t k = s->engine.add_attach_slot(seg, pos);
    if (slot) *slot = k;
    API_END
}

int vtx_set_attach_slot_position(vtx_solver* s, int32_t seg, int32_t slot, const float pos[3]) {
    NEED(s);
    API_BEGIN s->engine.set_attach_slot_position(seg, slot, pos); API_END
}

int vtx_add_attach(vtx_solver* s, int32_t seg, int32_t particle, int32_t slot, float dist) {
    NEED(s);
    API_BEGIN s->engine.add_attach(seg, &particle, &slot, &dist, 1); API_END
}

int vtx_add_attach_n(vtx_solver* s, int32_t seg, const int32_t* particle, const int32_t* slot, const float* dist,
                     uint32_t count) {
    NEED(s);
    API_BEGIN s->engine.add_attach(seg, particle, slot, dist, count); API_END
}

int vtx_add_bend(vtx_solver* s, int32_t seg, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, float angle) {
    NEED(s);
    uint32_t idx[4] = {i0, i1, i2, i3};
    API_BEGIN s->engine.add_bend(seg, idx, &angle, 1); API_END
}

int vtx_add_bend_n(vtx_solver* s, int32_t seg, const uint32_t* idx, const float* angle, uint32_t count) {
    NEED(s);
    API_BEGIN s->engine.add_bend(seg, idx, angle, count); API_END
}

int vtx_set_colliders(vtx_solver* s, int32_t seg, const vtx_collider* cols, uint32_t n) {
    NEED(s);
    API_BEGIN s->engine.set_colliders(seg, cols, n); API_END
}

int vtx_simulate(vtx_solver* s, float frame_dt) {
    NEED(s);
    API_BEGIN s->engine.simulate(frame_dt); API_END
}

int vtx_synchronize(vtx_solver* s) {
    NEED(s);
    API_BEGIN s->engine.synchronize(); API_END
}

int vtx_set_use_graph(vtx_solver* s, int32_t use_graph) {
    NEED(s);
    API_BEGIN s->engine.set_use_graph(use_graph != 0); API_END
}

static void to_abi(const vtx::Tuning& t, double ms, vtx_tuning* out) {
    out->use_graph = t.graph ? 1 : 0;
    out->neighbor_order = t.nbr_order;
    out->lra_batch = t.lra_batch;
    out->xcd_remap = t.remap ? 1 : 0;
    out->frame_ms = (float)ms;
    out->gather_block = t.gather_block;
    out->streams = t.streams;
    out->lra_split = t.lra_split ? 1 : 0;
    out->hash_lds = t.hash_lds ? 1 : 0;
    out->hash_grid = t.hash_grid;
}

int vtx_get_tuning(vtx_solver* s, vtx_tuning* out) {
    NEED(s); NEED(out);
    API_BEGIN to_abi(s->engine.tuning(), 0.0, out); API_END
}

int vtx_set_tuning(vtx_solver* s, const vtx_tuning* t) {
    NEED(s); NEED(t);
    API_BEGIN
    vtx::Tuning k;
    k.graph = t->use_graph != 0;
    k.nbr_order = t->neighbor_order;
    k.lra_batch = t->lra_batch;
    k.remap = t->xcd_remap != 0;
    k.gather_block = t->gather_block;
    k.streams = t->streams;
    k.lra_split = t->lra_split != 0;
    k.hash_grid = t->hash_grid;
    k.hash_lds = t->hash_lds != 0;
    s->engine.set_tuning(k);
    API_END
}

int vtx_autotune(vtx_solver* s, int32_t frames, float frame_dt, int32_t use_cache, vtx_tuning* out) {
    NEED(s);
    API_BEGIN
    double ms = 0.0;
    const vtx::Tuning t = s->engine.autotune(frames, frame_dt, use_cache != 0, &ms);
    if (out) to_abi(t, ms, out);
    API_END
}

int vtx_set_stencil_path(vtx_solver* s, int32_t enable) {
    NEED(s);
    API_BEGIN s->engine.set_stencil_path(enable != 0); API_END
}

int vtx_stencil_segments(vtx_solver* s, int32_t* count) {
    NEED(s); NEED(count);
    API_BEGIN *count = s->engine.stencil_segments(); API_END
}

int vtx_set_stream(vtx_solver* s, void* hip_stream) {
    NEED(s);
    API_BEGIN s->engine.set_stream((hipStream_t)hip_stream); API_END
}

int vtx_num_particles(vtx_solver* s, int32_t seg, uint32_t* n) {
    NEED(s); NEED(n);
    API_BEGIN *n = s->engine.seg(seg).n(); API_END
}

int vtx_read_positions(vtx_solver* s, int32_t seg, float* xyz) {
    NEED(s); NEED(xyz);
    API_BEGIN s->engine.read_particles(seg, 0, xyz); API_END
}

int vtx_read_velocities(vtx_solver* s, int32_t seg, float* xyz) {
    NEED(s); NEED(xyz);
    API_BEGIN s->engine.read_particles(seg, 1, xyz); API_END
}

int vtx_read_normals(vtx_solver* s, int32_t seg, float* xyz) {
    NEED(s); NEED(xyz);
    API_BEGIN s->engine.read_particles(seg, 2, xyz); API_END
}

int vtx_read_inv_masses(vtx_solver* s, int32_t seg, float* w) {
    NEED(s); NEED(w);
    API_BEGIN s->engine.read_particles(seg, 3, w); API_END
}

int vtx_device_positions(vtx_solver* s, int32_t seg, void** ptr, uint32_t* n) {
    NEED(s); NEED(ptr);
    API_BEGIN *ptr = (void*)s->engine.device_positions(seg, n); API_END
}

int vtx_device_normals(vtx_solver* s, int32_t seg, void** ptr, uint32_t* n) {
    NEED(s); NEED(ptr);
    API_BEGIN *ptr = (void*)s->engine.device_normals(seg, n); API_END
}

int vtx_mesh_constraints(const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                         int32_t* stretch_ij, float* stretch_rest, uint32_t stretch_cap, uint32_t* stretch_count,
                         uint32_t* bend_idx, float* bend_rest, uint32_t bend_cap, uint32_t* bend_count) {
    NEED(stretch_count); NEED(bend_count);
    API_BEGIN
    std::vector<int32_t> ij;
    std::vector<float> sr, br;
    std::vector<uint32_t> bi;
    vtx::mesh_constraints(xyz, n, tris, ntri, ij, sr, bi, br);
    *stretch_count = (uint32_t)sr.size();
    *bend_count = (uint32_t)br.size();
    if (stretch_ij || stretch_rest) {
        VTX_REQUIRE(stretch_cap >= sr.size(), VTX_ECAPACITY, "stretch arrays too small");
        if (stretch_ij) std::copy(ij.begin(), ij.end(), stretch_ij);
        if (stretch_rest) std::copy(sr.begin(), sr.end(), stretch_rest);
    }
    if (bend_idx || bend_rest) {
        VTX_REQUIRE(bend_cap >= br.size(), VTX_ECAPACITY, "bend arrays too small");
        if (bend_idx) std::copy(bi.begin(), bi.end(), bend_idx);
        if (bend_rest) std::copy(br.begin(), br.end(), bend_rest);
    }
    API_END
}

int vtx_add_mesh_constraints(vtx_solver* s, int32_t seg, int32_t offset, const uint32_t* tris, uint32_t ntri) {
    NEED(s);
    API_BEGIN s->engine.add_mesh_constraints(seg, offset, tris, ntri); API_END
}

int vtx_readback_request(vtx_solver* s, int32_t seg) {
    NEED(s);
    API_BEGIN s->engine.readback_request(seg); API_END
}

int vtx_readback_latest(vtx_solver* s, int32_t seg, const float** positions, const float** normals, uint32_t* n,
                        uint64_t* frame) {
    NEED(s); NEED(positions); NEED(normals); NEED(n); NEED(frame);
    API_BEGIN
    const float4 *p = nullptr, *q = nullptr;
    uint32_t m = 0;
    uint64_t f = 0;
    if (s->engine.readback_latest(seg, &p, &q, &m, &f)) {
        *positions = reinterpret_cast<const float*>(p);
        *normals = reinterpret_cast<const float*>(q);
        *n = m;
        *frame = f;
    } else {
        *positions = nullptr;
        *normals = nullptr;
        *n = 0;
        *frame = 0;
    }
    API_END
}

int vtx_snapshot_size(vtx_solver* s, int32_t seg, uint64_t* bytes) {
    NEED(s); NEED(bytes);
    API_BEGIN *bytes = (uint64_t)s->engine.snapshot_bytes(seg); API_END
}

int vtx_snapshot(vtx_solver* s, int32_t seg, void* buf, uint64_t bytes) {
    NEED(s); NEED(buf);
    API_BEGIN s->engine.snapshot(seg, buf, (size_t)bytes); API_END
}

int vtx_restore(vtx_solver* s, int32_t seg, const void* buf, uint64_t bytes) {
    NEED(s); NEED(buf);
    API_BEGIN s->engine.restore(seg, buf, (size_t)bytes); API_END
}

int vtx_set_particle(vtx_solver* s, int32_t seg, int32_t i, const float* pos, const float* vel, float inv_mass) {
    NEED(s);
    API_BEGIN s->engine.set_particle(seg, i, pos, vel, inv_mass); API_END
}

int vtx_write_predicted(vtx_solver* s, int32_t seg, const float* xyz) {
    NEED(s);
    API_BEGIN s->engine.write_predicted(seg, xyz); API_END
}

int vtx_run_hash(vtx_solver* s) {
    NEED(s);
    API_BEGIN s->engine.run_hash(); API_END
}

int vtx_hash_table_size(vtx_solver* s, int32_t seg, uint32_t* M) {
    NEED(s); NEED(M);
    API_BEGIN *M = s->engine.hash_table_size(seg); API_END
}

int vtx_read_hash(vtx_solver* s, int32_t seg, uint32_t* cell_start, uint32_t* cell_end, uint32_t* sorted_hash,
                  uint32_t* sorted_index, uint32_t* nbr) {
    NEED(s);
    API_BEGIN s->engine.read_hash(seg, cell_start, cell_end, sorted_hash, sorted_index, nbr); API_END
}

// ---------------------------------------------------------------- kernel-launch tier
static void hash_args(vtx::HashObjectsArgs& a, uint32_t* particle_hash, uint32_t* particle_index,
                      uint32_t* cell_start, uint32_t* cell_end, uint32_t* neighbors, const float* positions,
                      const float* original_positions, const vtx_hash_params* p) {
    VTX_REQUIRE(p, VTX_EINVAL, "params is NULL");
    VTX_REQUIRE(p->max_num_neighbors >= 1 && p->max_num_neighbors <= VTX_MAX_NEIGHBORS, VTX_EINVAL,
                "max_num_neighbors must be in [1, 64]");
    VTX_REQUIRE(p->table_size >= 1 && p->table_size < (1 << 28), VTX_EINVAL, "table_size must be in [1, 2^28)");
    VTX_REQUIRE(p->num_objects < (1u << 25), VTX_ECAPACITY, "num_objects must be below 2^25");
    VTX_REQUIRE(cell_start && cell_end, VTX_EINVAL, "cell_start / cell_end are NULL");
    VTX_REQUIRE(p->num_objects == 0 || (particle_hash && particle_index && neighbors && positions && original_positions),
                VTX_EINVAL, "an object array is NULL");
    VTX_REQUIRE(p->cell_spacing > 0.f, VTX_EINVAL, "cell_spacing must be positive");
    a.particle_hash = particle_hash;
    a.particle_index = particle_index;
    a.cell_start = cell_start;
    a.cell_end = cell_end;
    a.neighbors = neighbors;
    a.positions = positions;
    a.original_positions = original_positions;
    a.num_objects = p->num_objects;
    a.max_num_neighbors = p->max_num_neighbors;
    a.table_size = (uint32_t)p->table_size;
    a.cell_spacing = p->cell_spacing;
    a.cell_spacing2 = p->cell_spacing2;
    a.particle_diameter2 = p->particle_diameter2;
}

int vtx_hash_objects(uint32_t* particle_hash, uint32_t* particle_index, uint32_t* cell_start, uint32_t* cell_end,
                     uint32_t* neighbors, const float* positions, const float* original_positions,
                     const vtx_hash_params* params, void* hip_stream) {
    API_BEGIN
    vtx::HashObjectsArgs a{};
    hash_args(a, particle_hash, particle_index, cell_start, cell_end, neighbors, positions, original_positions, params);
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    void* ws = nullptr;
    const size_t bytes = vtx::hash_workspace_bytes(a.num_objects, a.table_size);
    const hipError_t e = hipMallocAsync(&ws, bytes, st);
    VTX_REQUIRE(e == hipSuccess, VTX_ENOMEM, "hipMallocAsync of the hash workspace failed");
    const hipError_t r = vtx::hash_objects(a, ws, st);
    const hipError_t f = hipFreeAsync(ws, st);
    VTX_HIP(r);
    VTX_HIP(f);
    API_END
}

static_assert(VTX_HASH_SCRATCH_BYTES == VTX_HASH_SCRATCH_LIMIT, "scratch size in launch.hpp and vtx.h differ");
static_assert(VTX_HASH_SHIFT == VTX_HASH_GRID_SHIFT, "grid shift in vtx_device.hpp and vtx.h differ");

int vtx_hash_objects_noalloc(uint32_t* particle_hash, uint32_t* particle_index, uint32_t* cell_start,
                             uint32_t* cell_end, uint32_t* neighbors, const float* positions,
                             const float* original_positions, const vtx_hash_params* params, void* hip_stream) {
    API_BEGIN
    vtx::HashObjectsArgs a{};
    hash_args(a, particle_hash, particle_index, cell_start, cell_end, neighbors, positions, original_positions, params);
    VTX_REQUIRE(vtx::hash_workspace_bytes(a.num_objects, a.table_size) <= (size_t)VTX_HASH_SCRATCH_BYTES,
                VTX_ECAPACITY,
                "num_objects / table_size need more workspace than the fixed scratch array holds "
                "(40.5 B per object + 8 B per cell, at most 64 MiB)");
    VTX_HIP(vtx::hash_objects_noalloc(a, static_cast<hipStream_t>(hip_stream)));
    API_END
}

int vtx_set_timing(vtx_solver* s, int32_t on) {
    NEED(s);
    API_BEGIN s->engine.set_timing(on != 0); API_END
}

int vtx_read_timing(vtx_solver* s, char* buf, size_t len) {
    NEED(s); NEED(buf);
    API_BEGIN
    std::string r = s->engine.timing_report();
    if (len) {
        std::strncpy(buf, r.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    API_END
}

int vtx_server_create(const vtx_server_config* cfg, vtx_server** out) {
    NEED(cfg); NEED(out);
    API_BEGIN
    vtx_params p;
    vtx_params_default(&p);
    *out = new vtx_server(p, *cfg);
    API_END
}

int vtx_server_destroy(vtx_server* sv) {
    API_BEGIN delete sv; API_END
}

int vtx_server_solver(vtx_server* sv, vtx_solver** out) {
    NEED(sv); NEED(out);
    API_BEGIN *out = &sv->solver; API_END
}

int vtx_server_new_job(vtx_server* sv, const vtx_params* params, int32_t frames, int32_t* job, int32_t* seg) {
    NEED(sv); NEED(params); NEED(job);
    API_BEGIN *job = sv->server.new_job(*params, frames, seg); API_END
}

int vtx_server_submit(vtx_server* sv, int32_t job) {
    NEED(sv);
    API_BEGIN sv->server.submit(job); API_END
}

int vtx_server_step(vtx_server* sv, float frame_dt, vtx_server_stats* out) {
    NEED(sv);
    API_BEGIN sv->server.step(frame_dt, out); API_END
}

int vtx_server_job_status(vtx_server* sv, int32_t job, int32_t* status, int32_t* frames_done) {
    NEED(sv);
    API_BEGIN
    const vtx::Job& j = sv->server.job(job);
    if (status) *status = j.status;
    if (frames_done) *frames_done = j.done;
    API_END
}

int vtx_server_read_result(vtx_server* sv, int32_t job, float* xyz, uint32_t n) {
    NEED(sv); NEED(xyz);
    API_BEGIN
    const vtx::Job& j = sv->server.job(job);
    if (j.status != VTX_JOB_DONE) throw vtx::Error(VTX_ESTATE, "job has no result");
    if (n != j.n) throw vtx::Error(VTX_EINVAL, "particle count mismatch");
    std::memcpy(xyz, j.result.data(), sizeof(float) * 3 * (size_t)n);
    API_END
}

int vtx_server_preempt(vtx_server* sv, int32_t job) {
    NEED(sv);
    API_BEGIN sv->server.preempt(job); API_END
}

int vtx_server_release_job(vtx_server* sv, int32_t job) {
    NEED(sv);
    API_BEGIN sv->server.release(job); API_END
}

int vtx_server_drain(vtx_server* sv) {
    NEED(sv);
    API_BEGIN sv->server.drain(); API_END
}

}  // extern "C"
