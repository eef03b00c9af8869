/*
 * sphgpu.h -- C-ABI of the MI355X-native SPH per-step hot path.
 *
 * One opaque context owns the device-resident particle state (SoA in HBM),
 * the octree, the neighbour lists and one HIP stream.  Every entry point is
 * extern "C", takes plain pointers and sizes, returns 0 on success or a
 * negative SPHGPU_E* code (text via sphgpu_last_error), never throws, and is
 * blocking on return (stream-ordered work is complete).  One host thread per
 * context; a context is not thread-safe.
 *
 * The entry points replace, one for one, the calls the reference driver makes
 * on its Module objects (reference: include/module.hpp:10-14,
 * src/solver.cpp:941-978):
 *
 *   sphgpu_build_tree       <- Simulation::make_tree / BHTree::make
 *                              (src/simulation.cpp:40-43, src/bhtree.cpp:59-115)
 *   sphgpu_pre_interaction  <- {,disph::,gsph::}PreInteraction::calculation
 *                              (src/pre_interaction.cpp:39-169,
 *                               src/disph/d_pre_interaction.cpp:21-162,
 *                               src/gsph/g_pre_interaction.cpp:27-144)
 *   sphgpu_fluid_force      <- {,disph::,gsph::}FluidForce::calculation
 *                              (src/fluid_force.cpp:26-86,
 *                               src/disph/d_fluid_force.cpp:26-86,
 *                               src/gsph/g_fluid_force.cpp:44-171)
 *   sphgpu_gravity          <- GravityForce::calculation (src/gravity_force.cpp:52-89)
 *   sphgpu_timestep         <- TimeStep::calculation (src/timestep.cpp:18-38)
 *   sphgpu_predict/correct  <- Solver::predict / Solver::correct
 *                              (src/solver.cpp:980-1062, 1064-1113)
 *   sphgpu_energy_momentum  <- Solver::compute_total_energies (src/solver.cpp:1149-1175)
 *
 * Host particle data crosses the boundary as one FP64 block of
 * SPHGPU_NFIELDS rows of n values (row-major, [field][particle], particle
 * index == SPHParticle::id), selected by a field bit mask.  The layout is the
 * reference SPHParticle (include/particle.hpp:8-43) plus the GSPH MUSCL
 * gradient arrays (src/solver.cpp:707-719) flattened to scalar rows.
 */
#ifndef SPHGPU_H
#define SPHGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums mirroring include/parameters.hpp:8-27 ------------------------ */
/* SPHType (include/parameters.hpp:9-15); GDISPH = DISPH pre-interaction and
 * pressure-energy force with the GSPH MUSCL/HLL p* (src/gdisph/) */
enum { SPHGPU_SSPH = 0, SPHGPU_DISPH = 1, SPHGPU_GSPH = 2, SPHGPU_GDISPH = 3 };
enum { SPHGPU_CUBIC_SPLINE = 0, SPHGPU_WENDLAND = 1, SPHGPU_GAUSSIAN = 2 };
enum { SPHGPU_RIEMANN_HLL = 0, SPHGPU_RIEMANN_ITERATIVE = 1 };
/* pair-loop arithmetic: FP32 (compensated sums) or FP64 (reference-equivalent) */
enum { SPHGPU_PAIR_FP32 = 0, SPHGPU_PAIR_FP64 = 1 };
/* Riemann starting guess: the reference's linear impedance guess, or the MLP */
enum { SPHGPU_GUESS_PLAIN = 0, SPHGPU_GUESS_LEARNED = 1 };

/* ---- error codes --------------------------------------------------------- */
enum {
    SPHGPU_OK = 0,
    SPHGPU_E_ARG = -1,          /* bad argument / unsupported parameter      */
    SPHGPU_E_HIP = -2,          /* HIP runtime error                         */
    SPHGPU_E_NEIGHBOR_OVERFLOW = -3, /* bhtree.cpp:129-135,300-302           */
    SPHGPU_E_NO_FREE_NODE = -4, /* bhtree.cpp:222-224                        */
    SPHGPU_E_STATE = -5,        /* call order violated (e.g. no tree)        */
    SPHGPU_E_NOMEM = -6
};

/*
 * Parameters: the keys and defaults of the reference parameter JSON
 * (src/solver.cpp:299-438), as consumed by SPHParameters
 * (include/parameters.hpp:28-97).  sphgpu_params_default() fills the
 * reference's code defaults.
 */
typedef struct sphgpu_params {
    int32_t dim;                 /* 1, 2 or 3 (SPH_DIM)                         */
    int32_t sph_type;            /* SPHType: ssph | disph | gsph                */
    int32_t kernel;              /* KernelType                                  */
    int32_t riemann_solver;      /* gsph.riemann_solver ("hll" | "iterative")   */
    int32_t neighbor_number;     /* neighborNumber (50)                         */
    int32_t leaf_particle_num;   /* leafParticleNumber (1)                      */
    int32_t max_tree_level;      /* maxTreeLevel (20)                           */
    int32_t iterative_sml;       /* iterativeSmoothingLength (1)                */
    int32_t use_balsara_switch;  /* useBalsaraSwitch (1)                        */
    int32_t use_time_dependent_av; /* useTimeDependentAV (0)                    */
    int32_t use_ac;              /* useArtificialConductivity (0)               */
    int32_t periodic;            /* periodic (0)                                */
    int32_t use_gravity;         /* useGravity (0)                              */
    int32_t gsph_2nd_order;      /* use2ndOrderGSPH (1)                         */
    int32_t pair_precision;      /* SPHGPU_PAIR_* (build-specific)              */
    int32_t riemann_guess;       /* SPHGPU_GUESS_* (build-specific)             */
    double gamma;                /* gamma (5/3)                                 */
    double cfl_sound;            /* cflSound (0.3)                              */
    double cfl_force;            /* cflForce (0.125)                            */
    double av_alpha;             /* avAlpha (1)                                 */
    double alpha_max;            /* alphaMax (2)                                */
    double alpha_min;            /* alphaMin (0.1)                              */
    double epsilon_av;           /* epsilonAV (0.2)                             */
    double alpha_ac;             /* alphaAC (1)                                 */
    double gravity_constant;     /* G (1)                                       */
    double theta;                /* theta (0.5)                                 */
    double range_min[3];         /* rangeMin                                    */
    double range_max[3];         /* rangeMax                                    */
} sphgpu_params;

/* ---- host field block layout: rows of n doubles ------------------------- */
enum {
    SPHGPU_F_POS = 0,       /* 0..2   pos                                     */
    SPHGPU_F_VEL = 3,       /* 3..5   vel                                     */
    SPHGPU_F_VEL_P = 6,     /* 6..8   vel_p                                   */
    SPHGPU_F_ACC = 9,       /* 9..11  acc                                     */
    SPHGPU_F_MASS = 12,
    SPHGPU_F_DENS = 13,
    SPHGPU_F_PRES = 14,
    SPHGPU_F_ENE = 15,
    SPHGPU_F_ENE_P = 16,
    SPHGPU_F_DENE = 17,
    SPHGPU_F_SML = 18,
    SPHGPU_F_SOUND = 19,
    SPHGPU_F_BALSARA = 20,
    SPHGPU_F_ALPHA = 21,
    SPHGPU_F_GRADH = 22,
    SPHGPU_F_PHI = 23,
    SPHGPU_F_ID = 24,       /* integer-valued                                 */
    SPHGPU_F_NEIGHBOR = 25, /* integer-valued                                 */
    SPHGPU_F_GRAD_D = 26,   /* 26..28 grad_density                            */
    SPHGPU_F_GRAD_P = 29,   /* 29..31 grad_pressure                           */
    SPHGPU_F_GRAD_V = 32,   /* 32..40 grad_velocity_k[axis] at 32 + 3k + axis */
    SPHGPU_NFIELDS = 41
};
#define SPHGPU_MASK(f) (((uint64_t)1) << (f))
#define SPHGPU_MASK_ALL ((((uint64_t)1) << SPHGPU_NFIELDS) - 1)

typedef struct sphgpu_ctx sphgpu_ctx;

/* Counters of the last module calls (device-side sums). */
typedef struct sphgpu_stats {
    int64_t pre_candidates;      /* sum over i of candidate-list length       */
    int64_t pre_newton_iters;    /* sum over i of h Newton iterations         */
    int64_t pre_newton_fail;     /* particles whose h iteration did not converge */
    int64_t force_pairs;         /* accepted force pairs (directed)           */
    int64_t force_candidates;    /* force candidate-list entries              */
    int64_t riemann_iters;       /* sum of Newton iterations in the Riemann solver */
    int64_t riemann_wave_max_iters; /* sum over wavefronts of max iterations  */
    int64_t riemann_calls;       /* Riemann problems solved                   */
    int64_t riemann_fallbacks;   /* learned guesses rejected for the plain one */
    int64_t riemann_learned;     /* pairs that used the learned guess         */
    int64_t gravity_interactions;/* particle-particle + particle-node terms    */
    int64_t nodes;               /* tree nodes (root included)                */
    int64_t leaves;
    int64_t search_node_visits;  /* tree nodes visited by neighbour-search wavefronts */
    int64_t search_member_tests; /* leaf members tested per wavefront (summed) */
    int32_t max_depth;           /* deepest node level (root = 1)             */
    int32_t pad;
    int64_t search_windows;      /* 64-node windows loaded by search wavefronts */
    int64_t search_leaves;       /* leaves whose members a wavefront tested   */
    int64_t search_exact;        /* candidates decided by the FP64 predicate  */
    int64_t flist_h_beyond;      /* particles whose h reached the pre-interaction
                                    search radius (force lists then repaired)  */
    int64_t flist_other;         /* particles that made a derived force list
                                    unusable (over 256 candidates, r = 0 pairs,
                                    list capacity)                             */
    int64_t flist_searches;      /* FluidForce calls that ran the is_ij search */
    int64_t search_reach_tests;  /* members of the leaves each particle's own
                                    search sphere reaches (the candidate tests
                                    of a per-particle tree walk), summed      */
} sphgpu_stats;

/* ---- lifecycle ------------------------------------------------------------ */
void        sphgpu_params_default(sphgpu_params *p, int dim);
int         sphgpu_create(const sphgpu_params *p, int n, sphgpu_ctx **out);
int         sphgpu_destroy(sphgpu_ctx *ctx);
const char *sphgpu_last_error(const sphgpu_ctx *ctx);
/* the parameters the context was created with and its particle count */
int         sphgpu_get_params(const sphgpu_ctx *ctx, sphgpu_params *p, int32_t *n);
int         sphgpu_device_info(char *name, int name_len, int *cu_count, double *hbm_gb);

/* ---- state exchange (fields: [SPHGPU_NFIELDS][n] FP64, index = id) ------- */
/* An upload that includes the mass row also records, from a host scan of
 * that row, whether every mass is equal; the density sums then take the
 * common mass instead of loading each neighbour's (same results). */
int sphgpu_upload(sphgpu_ctx *ctx, const double *fields, uint64_t mask);
int sphgpu_download(sphgpu_ctx *ctx, double *fields, uint64_t mask);
int sphgpu_set_dt(sphgpu_ctx *ctx, double dt);
int sphgpu_get_dt(sphgpu_ctx *ctx, double *dt);
int sphgpu_set_time(sphgpu_ctx *ctx, double t);
int sphgpu_get_time(sphgpu_ctx *ctx, double *t);
int sphgpu_get_h_per_v_sig(sphgpu_ctx *ctx, double *h_per_v_sig);
/* Simulation::set_h_per_v_sig (include/simulation.hpp:31), for drivers that
 * run PreInteraction elsewhere and TimeStep here. */
int sphgpu_set_h_per_v_sig(sphgpu_ctx *ctx, double h_per_v_sig);

/* ---- modules (drop-ins for Module::calculation) -------------------------- */
int sphgpu_build_tree(sphgpu_ctx *ctx);
/* BHTree::set_kernel (src/bhtree.cpp:117-120): kernel_size of every node from
 * the current smoothing lengths; PreInteraction calls it itself. */
int sphgpu_set_kernel(sphgpu_ctx *ctx);
int sphgpu_pre_interaction(sphgpu_ctx *ctx);   /* includes set_kernel       */
int sphgpu_fluid_force(sphgpu_ctx *ctx);
int sphgpu_gravity(sphgpu_ctx *ctx);
int sphgpu_timestep(sphgpu_ctx *ctx);
int sphgpu_predict(sphgpu_ctx *ctx);
int sphgpu_correct(sphgpu_ctx *ctx);

/* Solver::initialize tail (src/solver.cpp:941-963): alpha, balsara, sound,
 * tree, pre-interaction, fluid force, gravity at dt = 0. */
int sphgpu_initialize(sphgpu_ctx *ctx);
/* nsteps x Solver::integrate + Simulation::update_time (src/solver.cpp:966-978).
 * The step's PreInteraction searches its candidates within 1.05 h_guess and
 * repeats itself at the reference's 1.2 h_guess when some h reaches that
 * radius: the fields are the reference's either way. */
int sphgpu_integrate(sphgpu_ctx *ctx, int nsteps);
/* out[0..6] = kinetic, thermal, potential, total, P_x, P_y, P_z (FP64) */
int sphgpu_energy_momentum(sphgpu_ctx *ctx, double out[7]);
int sphgpu_get_stats(sphgpu_ctx *ctx, sphgpu_stats *out);
int sphgpu_reset_stats(sphgpu_ctx *ctx);

/* ---- parity hooks -------------------------------------------------------- */
/* BHTree::neighbor_search for every particle with caller-supplied smoothing
 * lengths h[n] (index = id).  is_ij selects the FluidForce variant.  The sets
 * are returned in id space, unsorted: counts[n], lists[n*cap].  For is_ij the
 * set is the one the force loop accepts (r < max(h_i,h_j), r != 0).
 * is_ij = 2 returns, without searching, the force pair sets the last
 * sphgpu_pre_interaction derived from its own lists (h is ignored; error
 * SPHGPU_E_STATE when none is current). */
int sphgpu_neighbor_sets(sphgpu_ctx *ctx, const double *h, int is_ij, int cap,
                         int32_t *counts, int32_t *lists);
/* Tree topology after sphgpu_build_tree: per node (pre-order) level, particle
 * count, first-member offset into members[], is_leaf, centre[3], edge, mass,
 * com[3], kernel_size.  members[] holds ids in key order.  Buffers sized by
 * the caller from n_nodes (call with NULL buffers to query). */
int sphgpu_tree_dump(sphgpu_ctx *ctx, int64_t *n_nodes, int32_t *level,
                     int32_t *count, int32_t *first, int32_t *is_leaf,
                     double *center, double *edge, double *mass, double *com,
                     double *kernel_size, int32_t *members);
/* Batched Riemann problems, states [u, rho, p, c] x n for left and right;
 * solver = SPHGPU_RIEMANN_*; precision = SPHGPU_PAIR_*; iters may be NULL.
 * With the FP32 iterative solver, a context whose riemann_guess is
 * SPHGPU_GUESS_LEARNED and that holds a network starts the pairs the GSPH
 * force kernel would hand to the network from the learned guess;
 * learned_used (may be NULL) then reports 0 = plain guess, 1 = learned guess
 * accepted, 2 = learned guess rejected for the plain one. */
int sphgpu_riemann_batch(sphgpu_ctx *ctx, int solver, int precision, int n,
                         const double *left, const double *right,
                         double *pstar, double *vstar, int32_t *iters,
                         int32_t *learned_used);
/* sphgpu_riemann_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * run (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The states are staged in a static device area of the
 * library (one per GPU, held until the process ends) instead of buffers
 * allocated per call, so n is limited to SPHGPU_RIEMANN_STATIC_MAX
 * (SPHGPU_E_ARG beyond it).  All calls on one GPU share that area: calls that
 * overlap on the same GPU (several host threads or contexts) race. */
#define SPHGPU_RIEMANN_STATIC_MAX 32768
int sphgpu_riemann_batch_static(sphgpu_ctx *ctx, int solver, int precision, int n,
                                const double *left, const double *right,
                                double *pstar, double *vstar, int32_t *iters,
                                int32_t *learned_used);

/* ---- learned Riemann starting guess ------------------------------------- */
/* Pair-state log written by the GSPH force kernel: rows of 7 floats
 * (u, rho, p)_L, (u, rho, p)_R, converged p*.  One row in `stride` pairs is
 * kept, up to `capacity` rows (ring buffer). */
int sphgpu_riemann_log_enable(sphgpu_ctx *ctx, int64_t capacity, int stride);
int sphgpu_riemann_log_read(sphgpu_ctx *ctx, float *rows, int64_t max_rows,
                            int64_t *n_rows);
/* MLP 6 -> H -> H -> 1 (H = 64); parameters FP32 master copy, bf16 compute. */
int sphgpu_mlp_init(sphgpu_ctx *ctx, uint64_t seed);
int sphgpu_mlp_set_params(sphgpu_ctx *ctx, const float *params, int64_t n);
int sphgpu_mlp_get_params(sphgpu_ctx *ctx, float *params, int64_t n);
int64_t sphgpu_mlp_param_count(void);
/* One Adam step on a batch of logged rows (device-resident after log_read or
 * given here as host rows); returns the batch loss. */
int sphgpu_mlp_train(sphgpu_ctx *ctx, const float *rows, int64_t n_rows,
                     int epochs, int batch, float lr, float *final_loss);
/* Forward and backward of the mean-squared loss on host rows, for gradient
 * checks: loss and dL/dparams (FP32). */
int sphgpu_mlp_loss_grad(sphgpu_ctx *ctx, const float *rows, int64_t n_rows,
                         float *loss, float *grad);
int sphgpu_mlp_predict(sphgpu_ctx *ctx, const float *rows, int64_t n_rows,
                       float *pstar_guess);

/* ---- timing ------------------------------------------------------------- */
/* Per-kernel averages (ms) of launches since the last reset, measured with
 * HIP events on the context's stream.  name_buf receives "name;name;...". */
int sphgpu_kernel_times(sphgpu_ctx *ctx, char *name_buf, int buf_len,
                        double *avg_ms, int64_t *count, int max_kernels,
                        int *n_kernels);
/* on = 0: no events; 1: every kernel group; 2: only the pair kernels
 * (neighbour searches, PreInteraction, FluidForce, gravity walk), so that a
 * timed run carries few events. */
int sphgpu_enable_timing(sphgpu_ctx *ctx, int on);

/* Measurement hook (no reference counterpart): copies nbytes (rounded down
 * to 16 B) between two device buffers with a plain 16-byte-per-lane copy
 * kernel (k_copy_probe) on the context's stream and returns its duration.
 * Its HBM counter readings calibrate the FETCH_SIZE / WRITE_SIZE counters,
 * and nbytes / ms is the attainable copy bandwidth. */
int sphgpu_copy_probe(sphgpu_ctx *ctx, int64_t nbytes, double *ms);

/* ---- snapshot and energy files (the reference's on-disk layout) ----------- */
/* OutputMetadata (include/output_metadata.hpp; src/output_metadata.cpp:8-37)
 * as the CSV header prints it.  Strings are NUL-terminated; the unit fields
 * carry UnitSystem's names and CGS factors (src/units.cpp). */
typedef struct sphgpu_output_meta {
    int32_t format_version;          /* 1                                     */
    int32_t step;                    /* snapshot count                        */
    char    timestamp[32];           /* "%Y-%m-%dT%H:%M:%SZ" (UTC)            */
    char    simulation_name[64];     /* "SPH Simulation"                      */
    char    unit_type[16];           /* "Code", "Galactic", "SI", "CGS"       */
    char    length_unit[32], mass_unit[32], time_unit[32], velocity_unit[32], energy_unit[32];
    double  length_to_cgs, mass_to_cgs, time_to_cgs, velocity_to_cgs;
    double  time_code, time_physical;
    int32_t particle_count;
    int32_t neighbor_number;
    int32_t sph_type;                /* SPHGPU_SSPH / _DISPH / _GSPH          */
    int32_t kernel;                  /* SPHGPU_CUBIC_SPLINE / _WENDLAND       */
    int32_t use_gravity, use_balsara, use_time_dependent_av;
    int32_t is_relaxation, relaxation_step, relaxation_total_steps;
    double  gamma, gravitational_constant;
    double  accumulated_time, alpha_scaling, rho_center, K, R, M_total;
    double  kinetic_energy, thermal_energy, potential_energy, total_energy;
    int32_t precision;               /* significant digits after the point (16) */
    int32_t pad;
} sphgpu_output_meta;

/* OutputManager::build_metadata (src/output_manager.cpp:254-300): code units,
 * the context's parameters and time, the current UTC time, and the energies
 * of the device state (OutputManager::compute_energies, :302-327). */
int sphgpu_output_meta_init(sphgpu_ctx *ctx, int32_t step, sphgpu_output_meta *meta);
/* OutputManager::generate_filename (src/output_manager.cpp:329-338):
 * "<dir>/<prefix>_%04d<ext>" into buf (returns SPHGPU_E_ARG if it does not fit). */
int sphgpu_snapshot_filename(const char *dir, const char *prefix, int32_t count, const char *ext,
                             char *buf, int32_t buf_len);
/* CSVWriter::open + write_particles (src/writers/csv_writer.cpp:17-165): the
 * 50-line header (59 with the relaxation block), the column line and one row
 * per particle in id order.  Axes beyond the context's dimension are written
 * as 0 (the reference reads past its vec_t there). */
int sphgpu_write_snapshot_csv(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta);
/* The same file from a host field block ([SPHGPU_NFIELDS][n], index = id),
 * for callers that hold the state on the host (no context needed). */
int sphgpu_write_csv_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                            const double *fields);
/* CSVWriter::read_metadata + read_particles (src/writers/csv_writer.cpp:174-371)
 * and OutputManager::load_for_resume (src/output_manager.cpp:163-233): the
 * header's metadata into meta (may be NULL), the 21 columns of every row into
 * the context (ids must be 0..n-1 for the context's n; fields the file does
 * not hold are set to 0), and the context time to "Time (code)".  A file
 * that the reference would reject (fewer than 51 lines, a row without 21
 * columns, a field that does not parse, no particle) returns SPHGPU_E_ARG. */
int sphgpu_read_snapshot_csv(sphgpu_ctx *ctx, const char *path, sphgpu_output_meta *meta);
/* The same parse into a host field block ([SPHGPU_NFIELDS][n], index = id;
 * rows must number n), and the row count of a snapshot file. */
int sphgpu_read_csv_fields(const char *path, sphgpu_output_meta *meta, int32_t dim, int32_t n, double *fields);
int sphgpu_csv_row_count(const char *path, int32_t *rows);
/* OutputManager::initialize, energy part (src/output_manager.cpp:55-73):
 * creates/truncates path and writes the 6 comment lines. */
int sphgpu_energy_file_begin(const char *path, const sphgpu_output_meta *meta);
/* OutputManager::write_energy (src/output_manager.cpp:235-252): appends
 * "time  kinetic  thermal  potential  total" (total = the sum of the three). */
int sphgpu_energy_file_write(const char *path, double time, double kinetic, double thermal, double potential,
                             int32_t precision);
/* The same row for the device state: the context's time and the energies of
 * sphgpu_energy_momentum. */
int sphgpu_energy_file_append(sphgpu_ctx *ctx, const char *path, int32_t precision);
/* VTKWriter::open + write_particles (src/writers/vtk_writer.cpp:17-331): a
 * legacy-VTK UNSTRUCTURED_GRID with one vertex per particle, positions,
 * density, pressure, mass, internal_energy, smoothing_length, sound_speed,
 * particle_id, velocity and acceleration, as big-endian binary (binary != 0,
 * the reference's default) or ASCII.  An empty set writes the header and
 * returns SPHGPU_E_ARG, as the reference's write_particles fails. */
int sphgpu_write_vtk_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                            const double *fields, int32_t binary);
int sphgpu_write_snapshot_vtk(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta, int32_t binary);
/* OutputMetadata::to_json (which = 0) or UnitSystem::to_json (which = 1)
 * as nlohmann::json::dump(2) prints them (src/output_metadata.cpp:67-116,
 * src/units.cpp:206-231): the text the HDF5 writer stores in its
 * "metadata_json" and "unit_system" attributes.  *needed gets the size with
 * the terminating NUL; buf_len 0 only queries it. */
int sphgpu_output_meta_json(const sphgpu_output_meta *meta, int32_t which, char *buf, int32_t buf_len,
                            int32_t *needed);
/* OutputMetadata::from_json (src/output_metadata.cpp:118-168): absent keys
 * take the reference's defaults; unit names follow units.type_enum.  Text
 * that is not a JSON object, or a key of the wrong JSON type, returns
 * SPHGPU_E_ARG. */
int sphgpu_output_meta_from_json(const char *text, sphgpu_output_meta *meta);

#ifdef __cplusplus
}
#endif
#endif /* SPHGPU_H */
