/*
 * vtx -- MI355X-native XPBD cloth solver, C ABI.
 *
 * The drop-in boundary for the reference's GPU cloth path. Every entry point below names the
 * reference interface it replaces (file:line in vitalight/Velvet). Conventions:
 *
 *  - Handle based, no globals: a vtx_solver owns one HIP stream and all its device buffers.
 *    Calls on different handles may run from different threads; calls on one handle may not.
 *  - Every function returns 0 on success or a negative VTX_E* code; vtx_last_error() gives
 *    the message (per calling thread). Nothing calls exit().
 *  - Vectors cross the boundary as packed xyz float triples on the host.
 *  - A solver holds one or more *segments*. A segment is one independent simulation with its
 *    own parameters, particles, constraints, colliders and spatial-hash grid -- exactly the
 *    state of one reference VtClothSolverGPU. Segment 0 exists from creation, so a caller that
 *    mirrors the reference uses segment 0 only. More segments are packed into the same
 *    launches (batching); they never interact.
 *  - Particle and slot indices are local to their segment.
 */
#ifndef VTX_H
#define VTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VTX_OK 0
#define VTX_EINVAL (-1)     /* bad argument */
#define VTX_EHIP (-2)       /* HIP runtime error */
#define VTX_ENOMEM (-3)     /* device or host allocation failed */
#define VTX_ESTATE (-4)     /* call not valid in the current state */
#define VTX_ECAPACITY (-5)  /* exceeds a stated maximum */

#define VTX_COLLIDER_SPHERE 0   /* Velvet/Common.hpp:113-118 (ColliderType order) */
#define VTX_COLLIDER_PLANE 1
#define VTX_COLLIDER_CUBE 2

/* Solver parameters. Replaces VtSimParams (Velvet/Common.hpp:19-48); defaults from
 * vtx_params_default() are the reference's. num_substeps, num_iterations, interleaved_hash and
 * max_num_neighbors must agree across the segments of one solver. */
typedef struct vtx_params {
    int32_t num_substeps;            /* 2 */
    int32_t num_iterations;          /* 4 */
    int32_t max_num_neighbors;       /* 64 (at most 64) */
    float max_speed;                 /* 50; add_cloth sets 2 d / (1/60) * num_substeps */
    float gravity[3];                /* (0, -9.8, 0) */
    float bend_compliance;           /* 0 */
    float damping;                   /* 0.25 */
    float relaxation_factor;         /* 1 (SOR omega of the averaged apply) */
    float long_range_stretchiness;   /* 1.2 */
    float collision_margin;          /* 0.06 */
    float friction;                  /* 0.1 */
    int32_t enable_self_collision;   /* 1 */
    int32_t interleaved_hash;        /* 3: rebuild neighbours every n substeps */
    float particle_diameter;         /* set by add_cloth (last cloth wins) */
    float particle_diameter_scalar;  /* 1.5 (used by mesh-side generators) */
    float hash_cell_size_scalar;     /* 1.5: cell = diameter * scalar */
    int32_t hash_table_size;         /* extension: 0 = 2 x particles (reference); < 2^28 */
    float stretch_stiffness;         /* extension: lambda multiplier; 1 = reference */
    int32_t tune_hash_grid;          /* extension: 0 = the reference's hash grid, never changed;
                                        1 = vtx_tuning.hash_grid (set by hand or picked by
                                        vtx_autotune) may shift this segment's grid origin and
                                        enlarge its table (see vtx_tuning) */
} vtx_params;

/* Collider. Replaces the host Collider -> SDFCollider packing
 * (Velvet/VtClothSolverGPU.hpp:187-205, Velvet/VtClothSolverGPU.cuh:8-18). Matrices are 4x4
 * column-major (glm layout): cur is this frame's transform, last the previous frame's; a
 * static collider passes last == cur. delta_time is the frame time the motion spans. */
typedef struct vtx_collider {
    int32_t type;
    float position[3];
    float scale[3];
    float cur[16];
    float last[16];
    float delta_time;
} vtx_collider;

typedef struct vtx_solver vtx_solver;

/* Layout version of the structs above and below; vtx_abi_version() returns the library's, so a
 * binding can check at load time that it was written against the same layouts. */
#define VTX_ABI_VERSION 5

void vtx_params_default(vtx_params* p);
const char* vtx_last_error(void);
const char* vtx_version(void);
int vtx_abi_version(void);

/* Create / destroy. Replaces VtClothSolverGPU construction and Start()
 * (Velvet/VtClothSolverGPU.hpp:27-33) and OnDestroy (hpp:50-54). */
int vtx_solver_create(const vtx_params* params, vtx_solver** out);
int vtx_solver_destroy(vtx_solver* s);

/* Parameters of a segment. Replaces writes to Global::simParams (Velvet/Global.hpp:24). */
int vtx_set_params(vtx_solver* s, int32_t seg, const vtx_params* params);
int vtx_get_params(vtx_solver* s, int32_t seg, vtx_params* out);

/* Batching extension: add an independent segment; returns its index (indices are never
 * reused). A removed segment frees its device ranges; an inactive segment keeps its state
 * resident but is skipped by vtx_simulate. Other segments are not moved or re-uploaded. */
int vtx_add_segment(vtx_solver* s, const vtx_params* params, int32_t* seg_out);
int vtx_remove_segment(vtx_solver* s, int32_t seg);
int vtx_set_segment_active(vtx_solver* s, int32_t seg, int32_t active);
int vtx_num_segments(vtx_solver* s, int32_t* out);

/* Replaces int AddCloth(shared_ptr<Mesh>, glm::mat4, float) (Velvet/VtClothSolverGPU.hpp:114-156)
 * including InitializePositions (Velvet/VtClothSolverGPU.cu:30-40): positions are transformed
 * by model (column-major 4x4, may be NULL for identity), inverse masses start at 1, triangle
 * indices are offset, the segment's diameter and max speed are set, and the collision-filter
 * snapshot of initial positions is retaken for the whole segment. *offset receives the index
 * of the cloth's first particle. */
int vtx_add_cloth(vtx_solver* s, int32_t seg, const float* xyz, uint32_t n, const uint32_t* tris,
                  uint32_t ntri, const float* model, float diameter, int32_t* offset);

/* Replaces AddStretch (Velvet/VtClothSolverGPU.hpp:158-163). ij holds 2*count indices. */
int vtx_add_stretch(vtx_solver* s, int32_t seg, int32_t i, int32_t j, float rest);
int vtx_add_stretch_n(vtx_solver* s, int32_t seg, const int32_t* ij, const float* rest, uint32_t count);

/* Replaces AddAttachSlot (Velvet/VtClothSolverGPU.hpp:165-168); *slot receives its index. */
int vtx_add_attach_slot(vtx_solver* s, int32_t seg, const float pos[3], int32_t* slot);
/* Replaces host writes to attachSlotPositions (e.g. Velvet/main.cpp:326-331). */
int vtx_set_attach_slot_position(vtx_solver* s, int32_t seg, int32_t slot, const float pos[3]);

/* Replaces AddAttach (Velvet/VtClothSolverGPU.hpp:170-176): distance 0 pins the particle. */
int vtx_add_attach(vtx_solver* s, int32_t seg, int32_t particle, int32_t slot, float dist);
int vtx_add_attach_n(vtx_solver* s, int32_t seg, const int32_t* particle, const int32_t* slot,
                     const float* dist, uint32_t count);

/* Replaces AddBend (Velvet/VtClothSolverGPU.hpp:178-185). idx holds 4*count indices. */
int vtx_add_bend(vtx_solver* s, int32_t seg, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3,
                 float angle);
int vtx_add_bend_n(vtx_solver* s, int32_t seg, const uint32_t* idx, const float* angle, uint32_t count);

/* Replaces UpdateColliders (Velvet/VtClothSolverGPU.hpp:187-205). Only enabled colliders are
 * passed (the reference keeps stale slots for disabled ones). */
int vtx_set_colliders(vtx_solver* s, int32_t seg, const vtx_collider* cols, uint32_t n);

/* Replaces Simulate() (Velvet/VtClothSolverGPU.hpp:56-111): one frame of every segment.
 * Asynchronous on the solver's stream; vtx_synchronize waits for it. On a stream bound with
 * vtx_set_stream, work queued on that stream after the call sees the whole frame. On the
 * solver's own stream the extra stream groups (vtx_tuning.streams) may run into the next frame;
 * every other call that touches device state waits for them first. */
int vtx_simulate(vtx_solver* s, float frame_dt);
int vtx_synchronize(vtx_solver* s);

/* Launch-mode switches. use_graph: record each frame's launches into a hipGraph and replay
 * it (default 1). */
int vtx_set_use_graph(vtx_solver* s, int32_t use_graph);
/* Extension: run the Jacobi pass of segments whose cloth is a regular vertex grid with the
 * reference topology on 2-D LDS tiles (default 1); other segments use per-particle incidence
 * lists. vtx_stencil_segments reports how many segments took the stencil layout. */
int vtx_set_stencil_path(vtx_solver* s, int32_t enable);
int vtx_stencil_segments(vtx_solver* s, int32_t* count);
/* Extension (no reference counterpart; the reference has fixed launch shapes,
 * Velvet/Common.cuh:23-31, and a fixed hash grid, Velvet/SpatialHashGPU.hpp:21-22): launch
 * variants of one frame and the hash-grid resolution of opted-in segments. Every combination of
 * the launch variants gives bit-identical results; only speed differs. hash_grid is the
 * exception (see below). */
typedef struct vtx_tuning {
    int32_t use_graph;         /* replay frames from a hipGraph */
    int32_t neighbor_order;    /* neighbour search thread order: 0 particle index, 1 hash-sorted,
                                  2 vertex patches for grid cloths (sorted for other meshes) */
    int32_t lra_batch;         /* dense attachment distances loaded per batch: 4, 16 or 64 */
    int32_t xcd_remap;         /* hand each XCD a contiguous run of tiles */
    float frame_ms;            /* vtx_autotune: the chosen variant's frame time (0 from the cache) */
    int32_t gather_block;      /* threads per workgroup of the neighbour search and contact pass:
                                  64, 128 or 256 (each 256-particle tile split across workgroups) */
    int32_t streams;           /* k = 1..4: the segments split into k particle-balanced groups
                                  whose frames run concurrently on k streams */
    int32_t lra_split;         /* 1: grid cloths with 32 or more dense attachment slots run their
                                  Jacobi pass with four threads per particle; 0: one thread */
    int32_t hash_lds;          /* 1: self-colliding segments of at most 16384 particles (table of
                                  at most 2^16 cells) build their spatial hash in one workgroup;
                                  0: every segment takes the device-wide sort */
    int32_t hash_grid;         /* hash-grid resolution of segments with tune_hash_grid = 1 (others
                                  keep the reference grid): 0 = the reference grid (table of
                                  hash_table_size cells, cell coordinates hashed as they are);
                                  1, 2, 3 = cell coordinates shifted by VTX_HASH_GRID_SHIFT in
                                  x, y and z before hashing, and a table of 1x, 2x, 4x
                                  hash_table_size cells (kept below 2^28). The reference hash
                                  maps cells (x, y, z) and (-x, y, -z) to one bucket, so a cloth
                                  centred on the origin has every bucket shared by two distant
                                  grid cells; the shift removes that. Unlike the other fields,
                                  this one changes the hash tables and the order of neighbour
                                  lists (not the set of neighbours within reach, except where
                                  buckets collide), so results differ at rounding level. */
} vtx_tuning;
#define VTX_HASH_GRID_SHIFT 1000003   /* cells, added with 32-bit wrap-around; odd, so that no table
                                         size (powers of two included) keeps the mirror pairing */
int vtx_get_tuning(vtx_solver* s, vtx_tuning* out);   /* vtx_set_tuning ignores frame_ms */
int vtx_set_tuning(vtx_solver* s, const vtx_tuning* t);
/* Time the variants on the current live set (coordinate descent, `frames` timed frames each,
 * positions and velocities restored after every trial, so the simulation does not advance),
 * apply the fastest and cache it for the live set's shape class (log2 buckets of particles and
 * segments, hash / stencil / dense-attachment presence, substeps x iterations). use_cache = 1
 * applies a class tuned earlier in this process without timing. hash_grid is tried only when a
 * live segment has tune_hash_grid = 1. */
int vtx_autotune(vtx_solver* s, int32_t frames, float frame_dt, int32_t use_cache, vtx_tuning* out);
/* Bind the solver to an existing stream (hipStream_t), or NULL for the solver's own. */
int vtx_set_stream(vtx_solver* s, void* hip_stream);

/* Replaces the public device buffers positions / normals / velocities / invMasses
 * (Velvet/VtClothSolverGPU.hpp:207-231): D2H reads of a segment's particles as xyz. */
int vtx_num_particles(vtx_solver* s, int32_t seg, uint32_t* n);
int vtx_read_positions(vtx_solver* s, int32_t seg, float* xyz);
int vtx_read_velocities(vtx_solver* s, int32_t seg, float* xyz);
int vtx_read_normals(vtx_solver* s, int32_t seg, float* xyz);
int vtx_read_inv_masses(vtx_solver* s, int32_t seg, float* w);
/* Zero-copy view of device positions / normals: float4 per particle (xyz, w = inverse mass
 * for positions), *ptr points at the segment's first particle. Valid until the next
 * registration call. Replaces the CUDA-GL interop mirror (Velvet/VtBuffer.hpp:188-236). */
int vtx_device_positions(vtx_solver* s, int32_t seg, void** ptr, uint32_t* n);
int vtx_device_normals(vtx_solver* s, int32_t seg, void** ptr, uint32_t* n);

/* Extension: constraints of an arbitrary triangle mesh. The reference generates them from grid
 * indices only, with bending rest angle 0 (Velvet/VtClothObjectGPU.hpp:75-132, TODO at 128-129).
 * Stretch: every unique edge, rest = current length, as index pairs. Bending: every edge of
 * exactly two triangles, as (p0, p1, p2, p3) = (third vertex of the first triangle, third vertex
 * of the second, the edge in the first triangle's winding), rest = the current angle between the
 * two triangle normals, so the given shape is at rest. Edges in order of first appearance.
 * vtx_mesh_constraints is host only (no solver, no GPU): counts are always written; arrays are
 * written when non-NULL and must hold stretch_cap pairs / bend_cap quadruples (else
 * VTX_ECAPACITY). vtx_add_mesh_constraints generates them from the current positions of a cloth
 * added at `offset` (triangle indices local to that cloth) and registers them on `seg`. */
int vtx_mesh_constraints(const float* xyz, uint32_t n, const uint32_t* tris, uint32_t ntri,
                         int32_t* stretch_ij, float* stretch_rest, uint32_t stretch_cap, uint32_t* stretch_count,
                         uint32_t* bend_idx, float* bend_rest, uint32_t bend_cap, uint32_t* bend_count);
int vtx_add_mesh_constraints(vtx_solver* s, int32_t seg, int32_t offset, const uint32_t* tris, uint32_t ntri);

/* Extension (the reference has no checkpointing; "reset" rebuilds scenes, Velvet/VtEngine.cpp:85-107):
 * the whole simulated state of a segment -- positions with inverse masses, velocities, initial
 * positions, slot positions -- to and from a self-describing host blob. Restore needs a segment
 * with the same particle and slot counts; registration data (constraints, colliders) is not in
 * the blob. */
int vtx_snapshot_size(vtx_solver* s, int32_t seg, uint64_t* bytes);
int vtx_snapshot(vtx_solver* s, int32_t seg, void* buf, uint64_t bytes);
int vtx_restore(vtx_solver* s, int32_t seg, const void* buf, uint64_t bytes);

/* Extension replacing the reference's CUDA-GL interop (Velvet/VtBuffer.hpp:167-229) for a viewer
 * that runs at its own pace: vtx_readback_request queues a copy of a segment's positions
 * (float4, w = inverse mass) and normals (float4) into one of three pinned host buffers (the
 * least recently requested one that vtx_readback_latest did not hand out last), after the work
 * queued so far; vtx_readback_latest returns the newest completed copy without waiting
 * (*positions = NULL if none has completed) and the number of frames simulated before it was
 * requested. A returned buffer stays valid until the next vtx_readback_latest call on the
 * segment: no request writes into it before then. */
int vtx_readback_request(vtx_solver* s, int32_t seg);
int vtx_readback_latest(vtx_solver* s, int32_t seg, const float** positions, const float** normals, uint32_t* n,
                        uint64_t* frame);

/* Host edit of one particle, the MouseGrabber effect (Velvet/MouseGrabber.hpp:51-52,60,74-75).
 * pos/vel may be NULL to keep them; inv_mass < 0 keeps it. */
int vtx_set_particle(vtx_solver* s, int32_t seg, int32_t i, const float* pos, const float* vel,
                     float inv_mass);

/* Spatial-hash stage on explicit positions, for inspection. Replaces
 * SpatialHashGPU::Hash (Velvet/SpatialHashGPU.hpp:41-52) / HashObjects
 * (Velvet/SpatialHashGPU.cu:159-196). vtx_write_predicted sets a segment's predicted
 * positions; vtx_run_hash hashes every segment's predicted positions; vtx_read_hash returns
 * the segment's table (cell_start / cell_end, M entries), its sorted hashes and particle
 * indices (n each) and its neighbour table (max_num_neighbors x n, column-major, entries
 * after a 0xffffffff sentinel are undefined). Any output pointer may be NULL. */
int vtx_write_predicted(vtx_solver* s, int32_t seg, const float* xyz);
int vtx_run_hash(vtx_solver* s);
int vtx_hash_table_size(vtx_solver* s, int32_t seg, uint32_t* M);
int vtx_read_hash(vtx_solver* s, int32_t seg, uint32_t* cell_start, uint32_t* cell_end,
                  uint32_t* sorted_hash, uint32_t* sorted_index, uint32_t* nbr);

/* Kernel-launch tier of the spatial hash. Replaces HashParams and HashObjects
 * (Velvet/SpatialHashGPU.cuh:7-25, Velvet/SpatialHashGPU.cu:159-196), the free function the
 * reference's SpatialHashGPU::Hash calls on raw device pointers. Every array pointer is device
 * memory: positions / original_positions hold num_objects packed xyz floats (glm::vec3); on
 * return particle_hash / particle_index hold the hashes sorted stably and the object indices in
 * that order (the reference sorts them in place), cell_start / cell_end (table_size entries
 * each) the first / one-past-last sorted position of every non-empty cell and 0xffffffff in
 * both for an empty one, and neighbors (max_num_neighbors x num_objects, entry k of object i at
 * neighbors[k * num_objects + i]) each object's list, ended by 0xffffffff unless full -- the
 * same tables as vtx_read_hash. Accepted neighbours j != i have |p_i - p_j|^2 < cell_spacing2
 * and |x0_i - x0_j|^2 > particle_diameter2. Limits: 1 <= max_num_neighbors <= 64,
 * 1 <= table_size < 2^28, num_objects < 2^25. Asynchronous on hip_stream (a hipStream_t; NULL =
 * the null stream); the input arrays must stay unchanged until the work is done.
 * vtx_hash_objects takes its sort and walk workspace with hipMallocAsync on the stream and
 * frees it there (hipFreeAsync), as the reference resizes its static sort storage. */
typedef struct vtx_hash_params {
    uint32_t num_objects;
    uint32_t max_num_neighbors;
    float cell_spacing;
    float cell_spacing2;
    int32_t table_size;
    float particle_diameter2;
} vtx_hash_params;
int vtx_hash_objects(uint32_t* particle_hash, uint32_t* particle_index, uint32_t* cell_start, uint32_t* cell_end,
                     uint32_t* neighbors, const float* positions, const float* original_positions,
                     const vtx_hash_params* params, void* hip_stream);

/* vtx_hash_objects without any allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (work recorded under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, limits and results as
 * vtx_hash_objects; the sort and walk workspace is a fixed device array of
 * VTX_HASH_SCRATCH_LIMIT bytes, one per GPU, held from the first call on that GPU until the
 * process ends. It needs 40.5 B per object plus 8 B per table cell (each part rounded up to
 * 256 B): e.g. 1,048,576 objects with 2^21 cells fit; larger calls return VTX_ECAPACITY before
 * any device work. The workspace is that of the current device (hipGetDevice), so hip_stream
 * must belong to it. Calls on one GPU share its array: two calls whose work overlaps on that
 * GPU (different streams, or host threads without ordering) race and give undefined tables;
 * calls queued in order on one stream are safe. Nothing in the array needs clearing between
 * calls. */
#define VTX_HASH_SCRATCH_LIMIT (64u << 20)
int vtx_hash_objects_noalloc(uint32_t* particle_hash, uint32_t* particle_index, uint32_t* cell_start,
                             uint32_t* cell_end, uint32_t* neighbors, const float* positions,
                             const float* original_positions, const vtx_hash_params* params, void* hip_stream);

/* Per-stage GPU time under the reference's ScopedTimerGPU labels (Velvet/Timer.hpp:242-260,
 * Velvet/GUI.cpp:32-51). With timing on, frames are launched without the graph and every launch
 * group is bracketed by a HIP event pair on the solver's stream; vtx_read_timing returns JSON
 * {label: {"ms": total, "launches": count, "kernel": name}} accumulated since timing was switched
 * on, plus "total". A label is the reference's stage label ("Solver_HashCache", ...), or several
 * joined by '+' where one kernel runs those stages fused ("Solver_CollideParticles+
 * Solver_ApplyDeltas+Solver_CollideSDFs" for k_collide); "kernel" names the kernel(s). */
int vtx_set_timing(vtx_solver* s, int32_t on);
int vtx_read_timing(vtx_solver* s, char* buf, size_t len);

/* ---------------------------------------------------------------------------------------
 * Serving extension (no counterpart in the reference): continuous batching of simulation
 * jobs. A job is a pending segment of the server's solver with a frame horizon; the caller
 * registers its cloths / constraints / colliders on that segment through the calls above
 * (vtx_server_solver gives the solver handle), then submits it. Each vtx_server_step admits
 * queued jobs first-come-first-served while the device-memory budget, the live-particle cap
 * and the job cap allow, advances every live job one frame in the same launches, and retires
 * finished jobs (their final positions stay readable until vtx_server_release_job).
 * ------------------------------------------------------------------------------------- */
#define VTX_JOB_PENDING 0    /* created, not submitted */
#define VTX_JOB_QUEUED 1
#define VTX_JOB_RUNNING 2
#define VTX_JOB_DONE 3
#define VTX_JOB_FAILED 4     /* evicted: non-finite state detected */
#define VTX_JOB_REJECTED 5   /* larger than the whole budget, or beyond a capacity limit of the
                                solver (seen at the job's first admission attempt or status query) */

typedef struct vtx_server_config {
    uint64_t hbm_budget_bytes;     /* device bytes jobs may occupy; 0 = free HBM at creation */
    uint64_t hbm_reserve_bytes;    /* kept free when the budget comes from free HBM */
    uint64_t max_live_particles;   /* latency budget; 0 = none */
    uint32_t max_live_jobs;        /* 0 = none */
    int32_t use_graph;             /* replay frames as hipGraphs (re-recorded on membership change) */
    int32_t check_every;           /* non-finite check period in steps; 0 = never */
    float step_budget_ms;          /* admit while the predicted frame time stays within it (cost
                                      per particle-substep measured on the live steps); 0 = none */
    uint64_t reserve_particles;    /* size the device pools at creation for this many live
                                      particles of grid cloths, so admission never waits for a
                                      pool to grow; 0 = grow on demand */
    int32_t pipelined;             /* 1: vtx_server_step returns once the frame is launched; the
                                      next step (or vtx_server_drain) waits for it and retires
                                      finished jobs, so the caller's work between steps (new
                                      jobs, polling) overlaps the frame on the GPU. 0: each step
                                      waits for its own frame */
} vtx_server_config;

typedef struct vtx_server_stats {
    uint32_t live_jobs, queued_jobs, admitted, completed, failed, pad;
    uint64_t live_particles;
    uint64_t particle_substeps;    /* work of the frame completed by this step */
    double step_ms;                /* wall time of this step (admission + frame + retirement);
                                      pipelined: from the previous step's frame launch to this
                                      step's (or to this step's end if it launches none), the
                                      period at which live jobs advance */
    double admit_ms;               /* uploading admitted jobs and rebuilding the tables */
    double frame_ms;               /* GPU time of the frame completed by this step (pipelined:
                                      the previous step's frame) */
    double retire_ms;              /* reading back and removing finished jobs */
} vtx_server_stats;

typedef struct vtx_server vtx_server;

int vtx_server_create(const vtx_server_config* cfg, vtx_server** out);
int vtx_server_destroy(vtx_server* sv);
int vtx_server_solver(vtx_server* sv, vtx_solver** out);
int vtx_server_new_job(vtx_server* sv, const vtx_params* params, int32_t frames, int32_t* job, int32_t* seg);
/* Queue a job. Its host packing runs on the server's worker thread while the caller goes on
 * (registering more jobs, stepping); the job's first admission attempt or status query waits for
 * that packing. A submitted job's segment must not be changed any more. */
int vtx_server_submit(vtx_server* sv, int32_t job);
int vtx_server_step(vtx_server* sv, float frame_dt, vtx_server_stats* out);
int vtx_server_job_status(vtx_server* sv, int32_t job, int32_t* status, int32_t* frames_done);
int vtx_server_read_result(vtx_server* sv, int32_t job, float* xyz, uint32_t n);
int vtx_server_release_job(vtx_server* sv, int32_t job);
/* Take a running job off the device (state kept on the host, device memory freed) and queue it
 * again behind the waiting jobs; it resumes from the frame it reached with the same result as
 * an uninterrupted run. */
int vtx_server_preempt(vtx_server* sv, int32_t job);
/* Pipelined servers: wait for the frame in flight and retire the jobs it finished (their
 * statistics go to the next step's). A no-op when no frame is in flight. */
int vtx_server_drain(vtx_server* sv);

#ifdef __cplusplus
}
#endif
#endif
