/* rds.h -- C ABI of the MI355X engine for PyRDSim's time-step loop.
 *
 * The reference (juliankappler/PyRDSim) exposes its step loop only through the
 * Python class `pyrdsim.system` (pyrdsim/__init__.py:1); it has no FFI.  The
 * engine therefore sits behind a Python facade with the same surface
 * (pyrdsim_amd.system), and this ABI is what that facade binds with ctypes.
 * Each entry point names the reference code it replaces.
 *
 * Conventions
 *   - every function returns RDS_OK (0) or a negative RDS_E_* code; the text of
 *     the last error of the calling thread is returned by rds_last_error();
 *   - all array arguments of the state/stream calls are DEVICE pointers owned by
 *     the caller (PyTorch-ROCm tensors in the facade); the engine owns only its
 *     internal work buffers, its HIP stream and its graphs;
 *   - calls are asynchronous on the context's stream unless stated otherwise;
 *   - one host thread per context.
 *
 * Particle order: the caller's order ("pid", 0..N_total-1, replicas
 * concatenated) is the order of every array passed in or out.  Internally the
 * engine keeps particles sorted by cell.
 */
#ifndef RDS_H
#define RDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RDS_ABI_VERSION 4

#define RDS_MAX_TYPES 8
#define RDS_MAX_TERMS 32
#define RDS_MAX_WALLS 8
#define RDS_MAX_REACTIONS 8

enum {
  RDS_OK = 0,
  RDS_E_ARG = -1,      /* invalid argument or configuration */
  RDS_E_HIP = -2,      /* HIP runtime error */
  RDS_E_BOUNDS = -3,   /* a coordinate left the box (pyrdsim.pyx:2286-2292) */
  RDS_E_STATE = -4,    /* call out of order (e.g. run before set_state) */
  RDS_E_STREAM = -5,   /* injected random stream exhausted */
  RDS_E_ANGLE = -6     /* an angle's cosine below -1.0000001 (pyrdsim.pyx:1364-1370) */
};

/* Precision.  RDS_FP64: FP64 pair forces and positions.  RDS_MIXED: FP32
 * pair forces, FP64 positions and update.  RDS_COMPENSATED: FP32 pair forces
 * and compensated FP32 positions -- each coordinate an unevaluated sum of two
 * floats, differenced and updated with float arithmetic only (two-sum), read
 * back as the exact FP64 value of the pair; cell search only (skin = 0). */
enum { RDS_FP64 = 0, RDS_MIXED = 1, RDS_COMPENSATED = 2 };

/* source of the Gaussian noise and of the reaction uniforms */
enum { RDS_RNG_PHILOX = 0, RDS_RNG_INJECTED = 1 };

/* reaction probability per step: the reference's p = dt*k (pyrdsim.pyx:2214)
 * or p = 1 - exp(-dt*k) */
enum { RDS_PROB_LINEAR = 0, RDS_PROB_EXP = 1 };

/* pair force laws, pyrdsim.pyx:1068-1130 */
enum { RDS_LJ = 1, RDS_TANH = 2, RDS_YUKAWA = 3, RDS_HARMONIC = 4, RDS_PIECEWISE_LINEAR = 5 };

/* reaction kinds, pyrdsim.pyx:2109-2206 */
enum { RDS_RX_LINEAR = 1, RDS_RX_CATALYZED = 2, RDS_RX_INHIBITED = 3 };

/* One `[interaction]` section of nonbonded_interactions.txt with a pair
 * (pyrdsim.pyx:1037-1130).  type_a <= type_b as written in the file: the
 * reference only matches sorted type pairs (pyrdsim.pyx:1999-2007). */
typedef struct {
  int32_t kind;
  int32_t type_a, type_b;
  int32_t n;            /* lj exponent */
  double u0, r0, rc, a, dr;
} rds_pair_term;

/* A `type = wall` section (pyrdsim.pyx:1132-1145, 1979-1986). */
typedef struct {
  int32_t particle_type, direction;
  double u0, r0, a;
} rds_wall_term;

/* A `[reaction]` section of reactions.txt (pyrdsim.pyx:2070-2222).  Unused
 * species fields are -1. */
typedef struct {
  int32_t kind, reactant, product, catalyst, inhibitor, inhibitor2;
  double rate, r0, rate_decrease;
} rds_reaction_def;

/* A model: everything that the reference reads from the interaction and
 * reaction files, in file order. */
typedef struct {
  int32_t n_types, n_terms, n_walls, n_reactions;
  rds_pair_term terms[RDS_MAX_TERMS];
  rds_wall_term walls[RDS_MAX_WALLS];
  rds_reaction_def reactions[RDS_MAX_REACTIONS];
} rds_model;

/* One independent system (a "replica").  A plain simulation is one replica.
 * force_prefac / noise_prefac are the reference's force_prefacs and
 * noise_prefacs (pyrdsim.pyx:488-496), computed by the caller so that they are
 * bit-identical to the reference's. */
typedef struct {
  int32_t model;
  int32_t n_particles;
  double L[3];          /* box; L[2] is read in three dimensions only */
  int32_t periodic[3];  /* 1 periodic, 0 reflecting (reference: periodic < 0) */
  double force_prefac[RDS_MAX_TYPES];
  double noise_prefac[RDS_MAX_TYPES];
  double rate[RDS_MAX_REACTIONS];  /* rate of each reaction of the model */
  double dt;
  uint64_t seed;        /* Philox key */
} rds_replica;

typedef struct {
  int32_t n_models, n_replicas;
  const rds_model* models;      /* host */
  const rds_replica* replicas;  /* host */
  int32_t precision;            /* RDS_FP64 | RDS_MIXED | RDS_COMPENSATED */
  int32_t rng;                  /* RDS_RNG_PHILOX | RDS_RNG_INJECTED */
  int32_t reaction_prob;        /* RDS_PROB_LINEAR | RDS_PROB_EXP */
  int32_t has_molecules;        /* set_state passes molecule numbers */
  int32_t tile_particles;       /* target particles per force tile, 0 = 56 */
  int32_t cells_per_cutoff;     /* 1 -> 3x3 stencil, 2 -> 5x5; 0 = 1 */
  int32_t graph_steps;          /* steps per captured hipGraph, 0 = no graph */
  int32_t tile_rows;            /* cell rows per force/reaction tile: 1, 2, 4, 8 (0 = 1) */
  /* harmonic bonds (pyrdsim.pyx:1303-1320, 2037-2047), global particle
   * indices, host arrays; n_bonds may be 0 */
  int64_t n_bonds;
  const int32_t* bond_i;
  const int32_t* bond_j;
  const double* bond_r0;
  const double* bond_u0;
  /* Neighbour search.  skin = 0: every step re-sorts the particles by cell and
   * the force kernel scans the 3x3 cell stencil.  skin > 0 (Verlet lists, as
   * the reference's create_neighbor_list with rs = range + skin,
   * pyrdsim.pyx:1448-1587): cells are at least (range + skin) wide, and each
   * particle keeps a list of the partners within range + skin.  The sort and
   * the lists are rebuilt in the step after some particle of the context has
   * moved more than skin/2 from where it was at the last rebuild, in place of
   * the reference's fixed neighbor_list_update_frequency; every pair within
   * range is therefore always listed.  list_capacity: entries per particle
   * (0 = 64); a particle with more partners falls back to the stencil scan. */
  double skin;
  int32_t list_capacity;
  /* 2 (0 means 2) or 3.  Three dimensions (pyrdsim.pyx:1159-1185,
   * 1999-2067 loop over N_dim) run the cell search (skin = 0) in FP64 or
   * mixed precision. */
  int32_t n_dim;
  /* harmonic angles (pyrdsim.pyx:1322-1407, 2052-2065) in definition order,
   * global particle indices, host arrays; n_angles may be 0.  angle_j is the
   * vertex.  theta0 in rad and k0 per rad^2, as the reference holds them
   * after parsing (pyrdsim.pyx:987-988: theta0 = pi deg/180,
   * k0 = k (180/pi)^2), so that U = (k0/2)(theta - theta0)^2.  Every angle
   * takes the straight-angle series (pyrdsim.pyx:1378-1386) when the LAST
   * angle defined has theta0 close to pi (np.isclose, atol 1e-5), because the
   * reference's angle closures read that flag when they are called.
   * angle_threshold is the reference's angle_pi_approximation_threshold
   * (0.1, pyrdsim.pyx:354).  The legs (i, j) and (j, k) count as bonded
   * partners of the catalysis search, like bonds (pyrdsim.pyx:1405-1406). */
  int64_t n_angles;
  const int32_t* angle_i;
  const int32_t* angle_j;
  const int32_t* angle_k;
  const double* angle_theta0;
  const double* angle_k0;
  double angle_threshold;
} rds_config;

/* Result of the last rds_sync. */
typedef struct {
  int64_t steps_done;       /* steps completed since set_state's step0 */
  int32_t bounds_error;     /* 1 if a coordinate left [0, L] */
  int32_t stream_error;     /* 1 if an injected stream ran out */
  int64_t error_step;       /* first failing step (reference step number) */
  int32_t error_particle;   /* pid of that failure */
  int32_t error_axis;
  double error_value;       /* coordinate after the boundary conditions */
  int64_t reaction_events;  /* type changes since set_state */
  int32_t integrity;        /* nonzero: an internal cell-list invariant failed */
  int32_t angle_error;      /* 1 if an angle's cosine fell below -1.0000001
                               (error_value holds it when bounds_error is 0) */
  int64_t rebuilds;         /* list mode: sorts + list builds since set_state */
  int64_t list_overflow;    /* list mode: particles that exceeded list_capacity */
  int64_t list_entries;     /* list mode: sum of list lengths over all builds */
} rds_status;

typedef struct rds_ctx rds_ctx;

int rds_abi_version(void);
const char* rds_last_error(void);

/* Builds cell grids, tile lists and device tables; allocates work buffers.
 * Replaces the setup half of simulate(): create_arrays,
 * set_up_neighbor_list_cells, create_bond_and_angle_force_functions
 * (pyrdsim.pyx:2385-2389). */
int rds_ctx_create(const rds_config* cfg, rds_ctx** out);
int rds_ctx_destroy(rds_ctx* ctx);

/* Loads positions [N][2] f64, types [N] i32 and, with has_molecules, molecule
 * numbers [N] i32 (device pointers), and bins them into cells.  step0 is the
 * number of steps already completed (pyrdsim.pyx:2377-2383, 2398). */
int rds_set_state(rds_ctx* ctx, const double* x, const int32_t* types,
                  const int32_t* molecules, int64_t step0);

/* Writes the current positions and types in caller order (the live x and
 * particle_types of simulate()'s result, pyrdsim.pyx:2446-2456). */
int rds_get_state(rds_ctx* ctx, double* x, int32_t* types);

/* Injected streams (device pointers): normals [n_steps][N][2], u1 and u2
 * [n_steps][N], for steps first_step .. first_step+n_steps-1.  u2 may be
 * NULL when no reaction can fire. */
int rds_set_injected(rds_ctx* ctx, const double* normals, const double* u1,
                     const double* u2, int64_t first_step, int64_t n_steps);

/* include_random_force of simulate(random_force=...) (pyrdsim.pyx:2369-2372) */
int rds_set_random_force(rds_ctx* ctx, int32_t include);

/* Runs n_steps integration steps (pyrdsim.pyx:2410-2426: integration_step =
 * update_positions, apply_boundary_conditions, cell rebuild, reaction_step,
 * update_drifts).  Asynchronous. */
int rds_run(rds_ctx* ctx, int64_t n_steps);

/* Waits for the stream and reports the status; returns RDS_E_BOUNDS,
 * RDS_E_ANGLE or RDS_E_STREAM when the status holds such an error. */
int rds_sync(rds_ctx* ctx, rds_status* status);

/* Drift of the current state, [N][2] f64 in caller order (update_drifts,
 * pyrdsim.pyx:1941-2067: pair forces, walls, bonds and angles). */
int rds_compute_forces(rds_ctx* ctx, double* forces);

/* Cell of each particle (global cell index) in caller order, and the grid:
 * cells_xy [n_replicas][2] = (lx, ly) and cell_len [n_replicas][2] (host). */
int rds_cell_ids(rds_ctx* ctx, int32_t* cells);
int rds_grid(rds_ctx* ctx, int32_t* cells_xy, double* cell_len, int32_t* cell_base);
/* The same for any n_dim: cells_xyz and cell_len [n_replicas][3] (host);
 * lz = 1 in two dimensions. */
int rds_grid3(rds_ctx* ctx, int32_t* cells_xyz, double* cell_len, int32_t* cell_base);

/* Number of other particles (different molecule) within distance r, caller
 * order (the neighbour sets of create_neighbor_list, pyrdsim.pyx:1448-1587). */
int rds_neighbor_counts(rds_ctx* ctx, double r, int32_t* counts);

/* Analysis of the current state on the device.  The reference computes none
 * of these: it writes trajectory.h5, and its examples post-process that file
 * on the host (create_video.py).  These replace that host work; r_max and
 * r_link are at most half of every periodic box length.
 *
 * rds_pair_histogram: per replica, the unordered pairs of particles whose
 * types are in type_mask (bit t = type t; 0 = every type) at minimum-image
 * distance r <= r_max, counted in bin floor(r / dr) when that is below n_bins
 * (1..4096).  hist: device [n_replicas][n_bins] uint64, overwritten.  The
 * RDF is these counts over the ideal-gas shell counts. */
int rds_pair_histogram(rds_ctx* ctx, double r_max, double dr, int32_t n_bins,
                       uint32_t type_mask, uint64_t* hist);

/* Particles of each type per replica: device [n_replicas][RDS_MAX_TYPES]
 * int64, overwritten (activated fraction = count of the active type / n). */
int rds_type_counts(rds_ctx* ctx, int64_t* counts);

/* Clusters: connected components of the particles whose types are in
 * type_mask (0 = every type), two particles linked when r <= r_link
 * (minimum image).  labels: device [N] int32 in caller order, the smallest
 * particle index of the particle's cluster, or -1 for a particle of another
 * type.  Waits for the stream (one host check per propagation pass). */
int rds_cluster_labels(rds_ctx* ctx, double r_link, uint32_t type_mask, int32_t* labels);

/* Largest particle count (n_total) of rds_cluster_labels_static. */
#define RDS_STATIC_LABELS_MAX 8388608

/* rds_cluster_labels without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc, set-up code around graph capture).
 * Same result as rds_cluster_labels; the label array lives in a file-scope
 * device array of RDS_STATIC_LABELS_MAX int32 entries (32 MiB per GPU, held
 * until the process ends) instead of memory allocated on first use, and
 * contexts with more particles get RDS_E_ARG.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other host threads, other
 * contexts) race and must be serialised by the caller. */
int rds_cluster_labels_static(rds_ctx* ctx, double r_link, uint32_t type_mask,
                              int32_t* labels);

/* Times n_steps eager steps with HIP events around every launch on the
 * context's stream; avg_ms[k] for k = 0 force+integrate, 1 scan, 2 scatter,
 * 3 reaction, 4 bonds, 5 list mode's binning + list build (the conditional
 * kernels are timed on every step, also when they have nothing to do);
 * counts[k] = launches timed. */
int rds_time_kernels(rds_ctx* ctx, int64_t n_steps, double* avg_ms, int64_t* counts);

/* Test hook with no counterpart in the reference: fills the LDS of many
 * workgroups with a 32-bit pattern, so that a following launch that read
 * LDS it never wrote would see that pattern (NaN bits make it visible). */
int rds_debug_fill_lds(rds_ctx* ctx, uint32_t pattern);

/* Re-tiles a context (kernel work decomposition only; the state is kept):
 * tile_rows cell rows per tile (1, 2, 4, 8) and about tile_particles particles
 * per tile row strip (0: 56 in two dimensions; in three, whose tiles are
 * always 2 x 2 rows in y and z and where tile_rows is checked and reported,
 * not used, 128).  No counterpart in the reference; used by the autotuner.
 * rds_get_tiling reports the current choice and LDS capacity. */
int rds_set_tiling(rds_ctx* ctx, int32_t tile_rows, int32_t tile_particles);
int rds_get_tiling(rds_ctx* ctx, int32_t* tile_rows, int32_t* tile_particles, int32_t* lds_cap);

/* Sizes: n_total particles, n_cells, n_tiles. */
int rds_sizes(rds_ctx* ctx, int64_t* n_total, int64_t* n_cells, int64_t* n_tiles);

/* The context's HIP stream (hipStream_t). */
int rds_stream(rds_ctx* ctx, void** stream);

#ifdef __cplusplus
}
#endif

#endif /* RDS_H */
