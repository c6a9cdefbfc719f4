/*
 * Oracle for the XPBD cloth path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a CPU restatement of the reference's algorithms. It is linked only by the
 * test-suite, by the smoke check and by the benchmark's CPU-baseline leg; the solver library
 * never calls it.
 *
 *   xpbd64.c      FP64 restatement of the reference GPU method (Velvet/VtClothSolverGPU.cu,
 *                 Velvet/VtClothSolverGPU.hpp:56-111, Velvet/SpatialHashGPU.cu): the parity
 *                 oracle for trajectories.
 *   hash32.c      binary32 restatement of the spatial-hash stage (Velvet/SpatialHashGPU.cu:13-196)
 *                 for bit-exact checks of cell tables and neighbour lists.
 *   cpu_solver.c  binary32 restatement of the reference's single-threaded CPU solver
 *                 (Velvet/VtClothSolverCPU.hpp + Velvet/SpatialHashCPU.hpp): the CPU timing
 *                 baseline, not a parity oracle.
 */
#ifndef VTX_ORACLE_H
#define VTX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Solver parameters (Velvet/Common.hpp:19-48) in double. hash_cell_size is the binary32
 * product diameter * hashCellSizeScalar (Velvet/SpatialHashGPU.hpp:21) widened to double. */
typedef struct orc_params {
    int32_t num_substeps;
    int32_t num_iterations;
    int32_t max_num_neighbors;
    int32_t enable_self_collision;
    int32_t interleaved_hash;
    int32_t hash_table_size;          /* 0 -> 2 * n (Velvet/SpatialHashGPU.hpp:22) */
    int32_t hash_shift;               /* added to every cell coordinate before hashing (32-bit
                                         wrap-around); 0 is the reference grid */
    double max_speed;
    double gravity[3];
    double bend_compliance;
    double damping;
    double relaxation_factor;
    double long_range_stretchiness;
    double collision_margin;
    double friction;
    double particle_diameter;
    double hash_cell_size;
    double stretch_stiffness;         /* extension; 1.0 reproduces the reference */
    double frame_dt;
} orc_params;

/* Collider (Velvet/VtClothSolverGPU.cuh:8-18). Matrices are column-major like glm:
 * m[col * 4 + row]. type: 0 sphere, 1 plane, 2 cube (Velvet/Common.hpp:113-118). */
typedef struct orc_collider {
    int32_t type;
    int32_t pad;
    double position[3];
    double scale[3];
    double cur[16];
    double last[16];
    double delta_time;
} orc_collider;

/* One independent simulation (the reference solver's whole state). Vectors are xyz-packed. */
typedef struct orc_scene {
    int32_t n;
    double* x;                 /* positions            3n, in/out */
    double* v;                 /* velocities           3n, in/out */
    double* w;                 /* inverse masses       n  */
    const double* x0;          /* initial positions    3n (collision filter) */
    int32_t S;
    const int32_t* st;         /* 2S */
    const double* st_rest;     /* S  */
    int32_t B;
    const uint32_t* bd;        /* 4B */
    const double* bd_rest;     /* B  */
    int32_t A;
    const int32_t* at_pid;     /* A */
    const int32_t* at_slot;    /* A */
    const double* at_dist;     /* A */
    int32_t L;
    const double* slots;       /* 3L */
    int32_t C;
    const orc_collider* cols;  /* C */
    int32_t T;
    const uint32_t* tris;      /* 3T */
    double* normals;           /* 3n out */
} orc_scene;

/* ---- FP64 stages (xpbd64.c) ---- */
void orc_predict(int32_t n, const double* x, double* v, double* p, const double g[3], double h);
void orc_collide_sdf(int32_t n, double* p, const double* x, const orc_collider* cols, int32_t C,
                     double margin, double friction, double dt);
/* Jacobi pass: stretch + attach + bend accumulated into deltas, then averaged apply. */
void orc_constraint_iteration(const orc_params* prm, const orc_scene* sc, double* p, double h);
void orc_finalize(int32_t n, double* x, double* v, const double* p, double h, double max_speed,
                  double damping);
void orc_normals(int32_t n, const double* x, int32_t T, const uint32_t* tris, double* normals);
/* Spatial hash in FP64: cell tables and column-major neighbour lists (sentinel 0xffffffff). */
int orc_hash64(int32_t n, const double* p, const double* x0, double cell, double diameter,
               int32_t M, int32_t max_nbr, int32_t shift, uint32_t* cell_start, uint32_t* cell_end,
               uint32_t* sorted_index, uint32_t* nbr);
void orc_collide_particles(int32_t n, double* p, const double* x, const double* w,
                           const uint32_t* nbr, int32_t max_nbr, double diameter,
                           double friction, double omega);
/* Whole frames (Velvet/VtClothSolverGPU.hpp:56-111). max_substeps < 0: no limit; otherwise
 * stop after that many substeps in total (the frame's remaining stages are skipped). */
int orc_simulate(const orc_params* prm, orc_scene* sc, int32_t frames, int32_t max_substeps);

/* ---- binary32 hash restatement (hash32.c) ---- */
int32_t h32_hash_coords(int32_t x, int32_t y, int32_t z, int32_t M);
int32_t h32_int_coord(float value, float cell);
int h32_hash(int32_t n, const float* p, const float* x0, float cell, float diameter, int32_t M,
             int32_t max_nbr, int32_t shift, uint32_t* particle_hash, uint32_t* sorted_hash,
             uint32_t* sorted_index, uint32_t* cell_start, uint32_t* cell_end, uint32_t* nbr);

/* ---- CPU baseline (cpu_solver.c) ---- */
typedef struct cpu_solver cpu_solver;
cpu_solver* cpu_solver_create(int32_t resolution_plus_one, const float* x, int32_t n,
                              const uint32_t* tris, int32_t T, const int32_t* attached,
                              int32_t nattached, const float* sphere_c, const float* sphere_r,
                              int32_t nspheres, int32_t has_plane, int32_t substeps,
                              int32_t iterations, const float gravity[3], float damping,
                              float friction, float margin, float bend_compliance);
void cpu_solver_simulate(cpu_solver* s, float frame_dt);
void cpu_solver_positions(const cpu_solver* s, float* out);
void cpu_solver_set_state(cpu_solver* s, const float* x, const float* v);
void cpu_solver_velocities(const cpu_solver* s, float* out);
void cpu_solver_neighbor_counts(const cpu_solver* s, int32_t* out);
void cpu_solver_destroy(cpu_solver* s);

#ifdef __cplusplus
}
#endif
#endif
