/*
 * nbody_amd.h -- C ABI of the MI355X Barnes-Hut step pipeline.
 *
 * Every entry point takes plain pointers and sizes.  Arrays named d_* live in
 * device (HBM) memory; arrays named h_* live in host memory.  Every function
 * returns an nbh_status; on failure nbh_last_error_message() describes the
 * failure and nbh_last_device_code() holds the device-detected error code
 * (the first-wins error word of the failing kernel), or 0.
 *
 * The C++ drop-in headers (amsc-nbody_amd/cpp/nbody/) rebuild the
 * reference's classes and free functions on top of this ABI and rethrow each
 * status as the exception class the reference throws.
 *
 * Morton key width is a runtime parameter `bits_per_axis`:
 *   10 -> 30-bit keys stored in uint32_t (reference layout, 2 padding bits,
 *         octree depth <= 10; src/radix_to_octree.hpp:10-19),
 *   21 -> 63-bit keys stored in uint64_t (1 padding bit, depth <= 21).
 * Functions that take `const void* keys` read uint32_t keys for 10 and
 * uint64_t keys for 21.
 *
 * Streams: every call is ordered on the stream of its nbh_context (the null
 * stream when the context was created with stream == NULL) and returns after
 * the work is complete, as the reference's wrappers do
 * (each ends in cudaDeviceSynchronize).  The simulation object runs a whole
 * step with a single host synchronisation.
 */
#ifndef NBODY_AMD_H
#define NBODY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum nbh_status {
    NBH_OK = 0,
    NBH_INVALID_ARGUMENT = -1, /* std::invalid_argument in the reference */
    NBH_LOGIC_ERROR = -2,      /* std::logic_error                      */
    NBH_RUNTIME_ERROR = -3,    /* std::runtime_error (device codes)     */
    NBH_DOMAIN_ERROR = -4,     /* std::domain_error (non-finite input)  */
    NBH_OVERFLOW_ERROR = -5,   /* std::overflow_error                   */
    NBH_HIP_ERROR = -6,        /* HIP runtime failure (runtime_error)   */
    NBH_OUT_OF_MEMORY = -7     /* allocation failure (runtime_error)    */
} nbh_status;

const char* nbh_last_error_message(void);
int nbh_last_device_code(void);
const char* nbh_version(void);

/* --------------------------------------------------------------- context */
/* A context binds a stream and a grow-only device workspace (sort buffers,
 * scan carries, error words).  Not thread-safe: one context per host thread. */
typedef struct nbh_context nbh_context;

int nbh_context_create(nbh_context** out, void* hip_stream);
int nbh_context_destroy(nbh_context* ctx);

/* Synchronous copy of `bytes` from device memory to host memory (waits for
 * all work on the device first). */
int nbh_copy_to_host(void* h_dst, const void* d_src, size_t bytes);

/* --------------------------------------------------------- bounding box */
/* Replaces Bbox::compute_box (gpu/cuda_kernels/globalbounding.cu:94-144).
 * Writes [centre_x, centre_y, centre_z, side] to d_box (device) and, when
 * h_box != NULL, to h_box.  count == 0 -> NBH_INVALID_ARGUMENT; a non-finite
 * coordinate -> NBH_DOMAIN_ERROR. */
int nbh_bbox_compute(nbh_context* ctx, const double* d_x, const double* d_y,
                     const double* d_z, size_t count, double* d_box,
                     double* h_box);

/* ------------------------------------------------------------ Morton keys */
/* Replaces MortonKeys::build (gpu/cuda_kernels/morton.cu:160-225): normalise,
 * quantise, interleave, then a stable sort of (key, index) pairs.
 * d_box is the 4-double device box of nbh_bbox_compute.  Keys are written in
 * ascending order to d_keys and the matching original indices to d_indices. */
int nbh_morton_keys(nbh_context* ctx, const double* d_x, const double* d_y,
                    const double* d_z, size_t count, const double* d_box,
                    int bits_per_axis, void* d_keys, uint32_t* d_indices);

/* Stable LSD radix sort of (key, value) pairs in place (the primitive behind
 * thrust::stable_sort_by_key at morton.cu:221-224).  key_bits is the number
 * of significant low key bits (keys must be < 2^key_bits); key_bytes 4 or 8. */
int nbh_radix_sort_pairs(nbh_context* ctx, void* d_keys, uint32_t* d_values,
                         size_t count, int key_bytes, int key_bits);

/* ------------------------------------------------------ Morton leaf groups */
/* Replaces buildMortonLeafGroups (gpu/cuda_kernels/morton_leaf_groups.cu:58-103):
 * run-length encode sorted keys into unique keys and counts, exclusive scan of
 * the counts into firstParticle.  key_bytes 4 or 8.  *h_n_groups receives the
 * group count. */
int nbh_leaf_groups_build(nbh_context* ctx, const void* d_sorted_keys,
                          int key_bytes, int n_particles, int capacity,
                          void* d_unique_keys, int* d_first_particle,
                          int* d_particle_count, int* h_n_groups);

/* The same build without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scan's scratch words (look-back status words, tile ticket, and the group
 * count when d_n_groups is NULL) live in a file-scope device array sized for
 * NBH_LEAF_GROUPS_NOALLOC_MAX particles (about 1 MiB, held on every device
 * the library runs on until the process ends); larger n_particles ->
 * NBH_INVALID_ARGUMENT.  Work is ordered on hip_stream (NULL: the null
 * stream) with that stream's device current; the call makes no hipMalloc,
 * hipFree or memset, and clears the status words with a kernel.  The group
 * count goes to d_n_groups (device int) when given; with h_n_groups the call
 * also synchronises the stream, validates the count and returns it there,
 * without it the call returns as soon as the work is queued.  Every call on
 * a device shares that device's array, which keeps the last call's
 * contents: calls that overlap on one device race. */
enum { NBH_LEAF_GROUPS_NOALLOC_MAX = 1 << 28 };
int nbh_leaf_groups_build_noalloc(void* hip_stream, const void* d_sorted_keys,
                                  int key_bytes, int n_particles, int capacity,
                                  void* d_unique_keys, int* d_first_particle,
                                  int* d_particle_count, int* d_n_groups,
                                  int* h_n_groups);

/* ------------------------------------------------------- Karras radix tree */
typedef struct nbh_tree_view {
    int n_leaves;
    int n_internal_nodes;
    int* left;          /* n_leaves-1 */
    int* right;         /* n_leaves-1 */
    int* parent;        /* 2*n_leaves-1 */
    int* range_first;   /* n_leaves-1 */
    int* range_last;    /* n_leaves-1 */
    int* prefix_length; /* n_leaves-1 */
    double* mass;       /* 2*n_leaves-1 (legacy centre-of-mass API only) */
    double* com_x;
    double* com_y;
    double* com_z;
    int* visit_count;   /* n_leaves-1 (legacy centre-of-mass API only) */
} nbh_tree_view;

/* Replaces buildTree (gpu/cuda_kernels/tree_builder.cu:303-352): Karras 2012
 * binary radix tree over sorted keys; leaves are 0..n-1, internal node i is
 * n+i, the root is node n.  key_bytes 4 (delta = clz32, ties 32+clz(i^j)) or 8
 * (delta = clz64, ties 64+clz(i^j)). */
int nbh_tree_build(nbh_context* ctx, const nbh_tree_view* tree,
                   const void* d_sorted_keys, int key_bytes, int n_leaves);

/* Replace computeCenterOfMass / computeGroupedCenterOfMass
 * (gpu/cuda_kernels/tree_builder.cu:374-503), the binary-tree mass and centre
 * of mass kept for API completeness.  d_first/d_count NULL -> one particle per
 * leaf. */
int nbh_tree_center_of_mass(nbh_context* ctx, const nbh_tree_view* tree,
                            const int* d_first_particle,
                            const int* d_particle_count,
                            const uint32_t* d_sorted_indices,
                            const double* d_mass, const double* d_x,
                            const double* d_y, const double* d_z);

/* ----------------------------------------------- radix -> octree planning */
/* Replaces buildRadixToOctreePlan (gpu/cuda_kernels/radix_to_octree.cu:174-244).
 * n_radix_nodes = 2*n_leaves-1.  Writes radix level, edge node count and the
 * exclusive scan of the counts; *h_n_octree_nodes = 1 + sum(counts). */
int nbh_radix_to_octree_plan(nbh_context* ctx, const nbh_tree_view* tree,
                             const void* d_unique_keys, int bits_per_axis,
                             int* d_radix_level, int* d_edge_node_count,
                             int* d_edge_node_offset, int* h_n_octree_nodes);

/* ----------------------------------------------------------------- octree */
typedef struct nbh_octree_view {
    int n_particles;
    int n_leaves;
    int n_nodes;
    int root;
    int* children;        /* 8*n_nodes, node-major: children[node*8+octant] */
    int* parent;
    int* level;
    void* prefix;         /* uint32_t (10 bits/axis) or uint64_t (21) */
    int* first_particle;
    int* particle_count;
    double* mass;
    double* com_x;
    double* com_y;
    double* com_z;
    double* center_x;
    double* center_y;
    double* center_z;
    double* half_size;
    int* pending_children;
    int* child_count;
} nbh_octree_view;

/* Replaces buildSparseOctreeTopology (gpu/cuda_kernels/octree_builder.cu:254-362). */
int nbh_octree_build_topology(nbh_context* ctx, const nbh_octree_view* octree,
                              const nbh_tree_view* tree,
                              const void* d_unique_keys,
                              const int* d_group_first,
                              const int* d_group_count, int n_groups,
                              int n_particles, int bits_per_axis,
                              const int* d_radix_level,
                              const int* d_edge_node_count,
                              const int* d_edge_node_offset);

/* Replaces computeOctreeGeometry (gpu/cuda_kernels/octree_geometry.cu:137-173).
 * h_box = host [centre_x, centre_y, centre_z, side]. */
int nbh_octree_geometry(nbh_context* ctx, const nbh_octree_view* octree,
                        int bits_per_axis, const double* h_box);

/* Replaces computeOctreeMassAndCenterOfMass
 * (gpu/cuda_kernels/octree_physics.cu:283-337). Particle arrays in original
 * order, d_sorted_indices maps sorted position -> original index. */
int nbh_octree_mass_com(nbh_context* ctx, const nbh_octree_view* octree,
                        const uint32_t* d_sorted_indices, const double* d_mass,
                        const double* d_x, const double* d_y,
                        const double* d_z);

/* ------------------------------------------------------------ Barnes-Hut */
enum { NBH_ORDER_ORIGINAL = 0, NBH_ORDER_MORTON = 1 };

/* Replaces computeBarnesHutAcceleration (gpu/cuda_kernels/barnes_hut.cu:325-391):
 * FP64 stackless traversal, MAC d > s/theta + delta, exact leaves, octants
 * visited 7 -> 0.  particle_order NBH_ORDER_MORTON means m/x/y/z are already
 * in sorted (Morton) order.  Accelerations are written in original order. */
int nbh_bh_acceleration(nbh_context* ctx, const nbh_octree_view* octree,
                        const uint32_t* d_sorted_indices, const double* d_mass,
                        const double* d_x, const double* d_y,
                        const double* d_z, double* d_ax, double* d_ay,
                        double* d_az, size_t particle_count, double theta,
                        double softening, double gravitational_constant,
                        int particle_order);

/* ------------------------------------------------------------- leapfrog */
/* Replace leapfrogKickDrift / leapfrogKick (gpu/cuda_kernels/integrator.cu:109-145). */
int nbh_leapfrog_kick_drift(nbh_context* ctx, double* d_x, double* d_y,
                            double* d_z, double* d_vx, double* d_vy,
                            double* d_vz, const double* d_ax,
                            const double* d_ay, const double* d_az,
                            size_t count, double dt);
int nbh_leapfrog_kick(nbh_context* ctx, double* d_vx, double* d_vy,
                      double* d_vz, const double* d_ax, const double* d_ay,
                      const double* d_az, size_t count, double dt);

/* Replaces gatherMortonOrderedParticleData (gpu/cuda_kernels/simulation.cu:33-107). */
int nbh_gather_morton(nbh_context* ctx, const uint32_t* d_sorted_indices,
                      const double* d_mass, const double* d_x,
                      const double* d_y, const double* d_z, double* d_out_mass,
                      double* d_out_x, double* d_out_y, double* d_out_z,
                      size_t count);

/* ------------------------------------------------------------ simulation */
enum { NBH_PRECISION_FP64 = 0, NBH_PRECISION_MIXED = 1 };

typedef struct nbh_sim_config {
    double time_step;             /* SimulationConfig::timeStep  */
    double theta;                 /* SimulationConfig::theta     */
    double softening;             /* SimulationConfig::softening */
    double gravitational_constant;/* default 6.67430e-11 (src/particle.hpp:16) */
    int precision;                /* NBH_PRECISION_FP64 or _MIXED */
    int bits_per_axis;            /* 10 or 21 */
} nbh_sim_config;

typedef struct nbh_sim nbh_sim;

typedef struct nbh_particles_view {
    size_t count;
    double* mass;
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
} nbh_particles_view;

/* Replaces Simulation::Simulation (src/simulation.cpp:21-43,70-74): validates the
 * config and copies the particle state (host arrays h_* of length count) into
 * device memory owned by the simulation. */
int nbh_sim_create(nbh_sim** out, const nbh_sim_config* config, size_t count,
                   const double* h_mass, const double* h_x, const double* h_y,
                   const double* h_z, const double* h_vx, const double* h_vy,
                   const double* h_vz, void* hip_stream);
/* The same from device arrays (copied; the caller keeps ownership). */
int nbh_sim_create_device(nbh_sim** out, const nbh_sim_config* config,
                          size_t count, const double* d_mass, const double* d_x,
                          const double* d_y, const double* d_z,
                          const double* d_vx, const double* d_vy,
                          const double* d_vz, void* hip_stream);
int nbh_sim_destroy(nbh_sim* sim);
/* Simulation::initialize / step / run (src/simulation.cpp:78-127). */
int nbh_sim_initialize(nbh_sim* sim);
int nbh_sim_step(nbh_sim* sim);
int nbh_sim_run(nbh_sim* sim, size_t steps);
double nbh_sim_time(const nbh_sim* sim);
size_t nbh_sim_step_number(const nbh_sim* sim);
/* Simulation::particles (src/simulation.cpp:133-135): device pointers. */
int nbh_sim_particles(nbh_sim* sim, nbh_particles_view* out);
/* Current accelerations (original order) of the last force evaluation. */
int nbh_sim_accelerations(nbh_sim* sim, double** d_ax, double** d_ay,
                          double** d_az);
/* Structures of the last rebuild, for parity checks: sorted keys/indices,
 * leaf groups, radix tree and octree (device pointers owned by the sim). */
typedef struct nbh_sim_structure {
    int n_particles;
    int n_groups;
    int n_octree_nodes;
    int bits_per_axis;
    const void* sorted_keys;
    const uint32_t* sorted_indices;
    const void* unique_keys;
    const int* group_first;
    const int* group_count;
    nbh_tree_view tree;
    nbh_octree_view octree;
    const int* radix_level;
    const int* edge_node_count;
    const int* edge_node_offset;
    const double* box; /* device [cx, cy, cz, side] */
    /* The sorted indices in the traversal's packet order, along a 3-D
     * Hilbert curve through the grid positions of the keys (no counterpart
     * in the reference, which walks one target per thread). */
    const uint32_t* target_order;
} nbh_sim_structure;
int nbh_sim_structure_view(nbh_sim* sim, nbh_sim_structure* out);

/* Phase timings of the last step in milliseconds (HIP events on the sim's
 * stream), phase k named by nbh_phase_name(k). Returns the phase count. */
int nbh_phase_count(void);
const char* nbh_phase_name(int phase);
int nbh_sim_phase_times(nbh_sim* sim, double* ms, int capacity);
int nbh_sim_set_timing(nbh_sim* sim, int enabled);
/* Recompute the current forces once more and count the particle-particle and
 * particle-node interactions of that evaluation (not part of a step), and,
 * when wave_visits != NULL, the node visits of all wavefronts (mixed
 * precision only; interactions / (64 * visits) is the packet efficiency). */
int nbh_sim_count_interactions(nbh_sim* sim, uint64_t* pp, uint64_t* pn,
                               uint64_t* wave_visits);

/* ------------------------------------------------------- energy, adjoint */
/* No reference entry point: the reference evaluates energies only inside its
 * tests (tests/unit/cuda/simulation_test.cu, two-body orbit). Both functions
 * walk the current tree in FP64 (precision NBH_PRECISION_FP64 only): the
 * potential of particle i is -sum G m_j (r^2 + eps^2)^{-1/2} over the same
 * interaction list as its force.  kinetic = 1/2 sum m v^2,
 * potential = 1/2 sum m phi.  d_phi: device array of count, original order. */
int nbh_sim_energy(nbh_sim* sim, double* kinetic, double* potential);
int nbh_sim_potential(nbh_sim* sim, double* d_phi);

/* Force vector-Jacobian product at the current tree: g = (da/dx)^T abar
 * and, when d_gm != NULL, gm = (da/dm)^T abar, the tree (interaction lists)
 * held fixed.  A source node's mass and centre of mass are differentiated
 * through to its particles.  An FP64 simulation differentiates its FP64
 * forces; a mixed one its mixed forces, along the mixed walk (FP32
 * interactions on fixed-point separations, compensated target sums, FP64
 * source and node accumulation and push-down).  All arrays are device
 * arrays of count in original order.  Pairs with nbh_sim_accelerations as
 * the forward. */
int nbh_sim_force_vjp(nbh_sim* sim, const double* d_abx, const double* d_aby,
                      const double* d_abz, double* d_gx, double* d_gy,
                      double* d_gz, double* d_gm);

/* Reverse mode of `steps` kick-drift-kick steps (the FP64 or the mixed
 * pipeline, as config->precision says; tree fixed within each force
 * evaluation).  Forward takes the initial state (device
 * arrays, original order), stores (x, v) every checkpoint_every steps and
 * returns the final state; backward takes the cotangents of the final
 * positions and velocities and returns those of the initial positions,
 * velocities and, when d_m0bar != NULL, masses, recomputing each segment
 * from its checkpoint. */
typedef struct nbh_adjoint nbh_adjoint;
int nbh_adjoint_create(nbh_adjoint** out, const nbh_sim_config* config,
                       size_t count, int steps, int checkpoint_every,
                       void* hip_stream);
int nbh_adjoint_destroy(nbh_adjoint* adjoint);
int nbh_adjoint_forward(nbh_adjoint* adjoint, const double* d_mass,
                        const double* d_x, const double* d_y, const double* d_z,
                        const double* d_vx, const double* d_vy,
                        const double* d_vz, double* d_x_out, double* d_y_out,
                        double* d_z_out, double* d_vx_out, double* d_vy_out,
                        double* d_vz_out);
/* Number of nbh_adjoint_forward calls made on `adjoint` so far.  A backward
 * replays the checkpoints and masses of the most recent forward only, so a
 * caller holding the result of an earlier forward compares the generation it
 * recorded after that forward with this one before calling backward.  No
 * reference counterpart (the reference has no adjoint). */
int nbh_adjoint_generation(const nbh_adjoint* adjoint, uint64_t* generation);
int nbh_adjoint_backward(nbh_adjoint* adjoint, const double* d_xbar,
                         const double* d_ybar, const double* d_zbar,
                         const double* d_vxbar, const double* d_vybar,
                         const double* d_vzbar, double* d_x0bar,
                         double* d_y0bar, double* d_z0bar, double* d_vx0bar,
                         double* d_vy0bar, double* d_vz0bar, double* d_m0bar);

/* ------------------------------------------------------------ workloads */
/* Seeded synthetic particle sets (host arrays of length count).
 * kind: 0 Plummer (a=1, M=1, truncated at 10a, equal masses)
 *       1 uniform cube [-1,1]^3, lognormal(0,0.5) masses normalised to 1,
 *         Gaussian velocities sigma 0.1
 *       2 exponential disk (80%) + Hernquist bulge (20%)
 *       3 256 Gaussian clumps in [-10,10]^3, cold start
 *       4 Gaussian blob N(0,1)^3, velocities sigma 0.01, equal masses */
int nbh_generate(int kind, size_t count, uint64_t seed, double* h_mass,
                 double* h_x, double* h_y, double* h_z, double* h_vx,
                 double* h_vy, double* h_vz);

/* Planar binary particle files (src/particle.hpp:19-77,145-224): a uint64 or
 * uint32 particle count, seven blocks of `count` doubles (mass, x, y, z, vx,
 * vy, vz) and optionally `count` uint8 particle types; every value finite,
 * masses non-negative, the file size exact.  info reports the count and
 * whether a type block is present; read fills host arrays of at least
 * `capacity` elements (h_type may be NULL) and sets *count.  Format and
 * validation failures -> NBH_RUNTIME_ERROR with the reference's messages. */
int nbh_particle_file_info(const char* path, size_t* count, int* has_types);
int nbh_particle_file_read(const char* path, size_t capacity, double* h_mass,
                           double* h_x, double* h_y, double* h_z, double* h_vx,
                           double* h_vy, double* h_vz, uint8_t* h_type,
                           size_t* count);

/* FP64 direct sum on `n_targets` targets (indices into the particle arrays),
 * softened, on the device: the accuracy reference for large runs. */
int nbh_direct_sum(nbh_context* ctx, const double* d_mass, const double* d_x,
                   const double* d_y, const double* d_z, size_t count,
                   const uint32_t* d_targets, size_t n_targets,
                   double softening, double gravitational_constant,
                   double* d_ax, double* d_ay, double* d_az);

/* ------------------------------------------------------- frame capture */
/* Sample j of `sample_count` (1 <= sample_count <= count) reads particle
 * j * count / sample_count. */

/* Replaces gatherPositionBounds, the host bound reduction and
 * gatherQuantizedPositionSamples of DeviceFrameWriter::write
 * (src/streaming/device_capture.cu:27-137,244-313).  Writes the sampled
 * per-axis bounds {min x, min y, min z, max x, max y, max z} to h_bounds and
 * 3 * sample_count interleaved uint16 positions to d_positions:
 *   q = 0 when max == min,
 *   q = (uint16)fma(clamp((p - min) / (max - min), 0, 1), 65535, 0.5) otherwise.
 * A non-finite sampled coordinate -> NBH_RUNTIME_ERROR ("Cannot quantize a
 * graphical frame with non-finite positions"). */
int nbh_frame_capture(nbh_context* ctx, const double* d_x, const double* d_y,
                      const double* d_z, size_t count, size_t sample_count,
                      uint16_t* d_positions, double* h_bounds);

/* Replaces gatherParticleTypeSamples and the type validation of the
 * DeviceFrameWriter constructor (src/streaming/device_capture.cu:139-153,
 * 186-214): d_types_out[j] = d_type[j * count / sample_count]; a value above
 * ParticleType::Asteroid -> NBH_RUNTIME_ERROR ("Cannot capture a frame with
 * invalid particle types"). */
int nbh_frame_sample_types(nbh_context* ctx, const uint8_t* d_type, size_t count,
                           size_t sample_count, uint8_t* d_types_out);

/* The .nbsnap frame header (src/streaming/frame.hpp:25-44); the little-endian
 * file layout is given in amsc-nbody_amd/cpp/nbody/streaming/frame.hpp. */
typedef struct nbh_frame_header {
    uint32_t version;      /* 1, 2 or 3 */
    uint32_t header_bytes; /* 72, 80 or 120 (version 2), 128 */
    uint64_t sequence;
    uint64_t simulation_step;
    double simulation_time;
    uint64_t source_particle_count;
    uint64_t particle_count;
    uint32_t scalar_bytes; /* 4 (float32) or 2 (quantized) */
    uint32_t components;   /* 3 */
    uint64_t payload_bytes;
    uint64_t total_steps;  /* 80-byte layout only */
    double bounds_min[3];  /* quantized layouts */
    double bounds_max[3];
    uint64_t particle_type_bytes;
} nbh_frame_header;

/* FrameSpool (src/streaming/frame.cpp:180-505) for C callers.  session_id
 * NULL -> "<microseconds since the epoch>-<pid>" (letters, digits, '-' and
 * '_' only).  Each call publishes
 * <directory>/run-<session>-frame-<20-digit sequence>.nbsnap through a ".part"
 * file and a rename, refuses to overwrite a published frame, and copies the
 * published path into path_out (may be NULL).  write_quantized: version 3
 * with h_types, version 2 (120-byte header) without; bounds as written by
 * nbh_frame_capture.  write_float: version 2 (80-byte header, total_steps)
 * with h_types, version 1 without.  source_count 0 means `count`.  Argument
 * errors -> NBH_INVALID_ARGUMENT, file errors -> NBH_RUNTIME_ERROR, with the
 * reference's messages. */
int nbh_frame_write_quantized(const char* directory, const char* session_id,
                              uint64_t sequence, uint64_t step, double time,
                              size_t count, size_t source_count,
                              const double* bounds, const uint16_t* h_positions,
                              const uint8_t* h_types, char* path_out,
                              size_t path_capacity);
int nbh_frame_write_float(const char* directory, const char* session_id,
                          uint64_t sequence, uint64_t step, double time,
                          uint64_t total_steps, size_t count, size_t source_count,
                          const float* h_positions, const uint8_t* h_types,
                          char* path_out, size_t path_capacity);
/* FrameSpool::markComplete: publishes run-<session>.complete holding the
 * session id and a newline; idempotent. */
int nbh_frame_mark_complete(const char* directory, const char* session_id,
                            char* path_out, size_t path_capacity);
/* readFrameHeader, readFramePositions, readFrameParticleTypes
 * (src/streaming/frame.cpp:507-720): every layout, validated against the
 * file size; quantized positions are decoded to float32 with the bounds.
 * h_types (may be NULL) reads as ParticleType::Unknown for frames without
 * a type block. */
int nbh_frame_read_header(const char* path, nbh_frame_header* out);
int nbh_frame_read(const char* path, size_t capacity, float* h_positions,
                   uint8_t* h_types, size_t* count);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_AMD_H */
