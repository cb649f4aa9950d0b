/*
 * molflow_gpu.h -- C ABI of the MI355X test-particle Monte Carlo tracer.
 *
 * This is the drop-in boundary for MolFlow+'s CPU ParticleTracer. A host
 * "simulation unit" (the role of MolflowSimulation : Simulation_Abstract,
 * reference src/Simulation/MolflowSimulation.h:22-45) loads a model through
 * mfg_load_model, builds the acceleration structure, runs desorptions and
 * adds the device tallies into caller-owned buffers laid out like
 * GlobalSimuState (reference src/Simulation/MolflowSimGeom.h:56-154).
 *
 * Conventions
 *  - Every function returns an int status (MFG_OK = 0); no C++ exception
 *    crosses the ABI. mfg_last_error(ctx) holds the text of the last error.
 *  - The caller owns host memory, the library owns device memory.
 *  - One context per host thread; calls on one context are not re-entrant.
 *  - Units follow the reference: lengths in cm, speeds in m/s, time in s,
 *    outgassing in Pa*m3/s, gas mass in g/mol, temperatures in K.
 *  - All integers are fixed width; structs are plain C, naturally aligned,
 *    and start with a struct_size field checked by the library.
 */
#ifndef MOLFLOW_GPU_H
#define MOLFLOW_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFG_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
enum {
    MFG_OK = 0,
    MFG_ERR_INVALID = 1,     /* bad argument or model (SanityCheck role) */
    MFG_ERR_DEVICE = 2,      /* HIP runtime error */
    MFG_ERR_STATE = 3,       /* call out of order (e.g. run before load) */
    MFG_ERR_NOMEM = 4,       /* device allocation failed */
    MFG_ERR_DESORPTION = 5,  /* no desorbing facet (MCStepResult::DesorptionError) */
    MFG_ERR_LIMIT = 6        /* size above a documented maximum */
};

/* ---- enums restated from reference src/MolflowTypes.h:17-47 ------------ */
enum { MFG_DES_NONE = 0, MFG_DES_UNIFORM = 1, MFG_DES_COSINE = 2, MFG_DES_COSINE_N = 3, MFG_DES_ANGLEMAP = 4 };
enum {
    MFG_PROFILE_NONE = 0, MFG_PROFILE_U = 1, MFG_PROFILE_V = 2, MFG_PROFILE_ANGULAR = 3,
    MFG_PROFILE_VELOCITY = 4, MFG_PROFILE_ORT_VELOCITY = 5, MFG_PROFILE_TAN_VELOCITY = 6
};
#define MFG_PROFILE_SIZE 100   /* PROFILE_SIZE; fixture B01 Profile size="100" */

/* ---- model description -------------------------------------------------- */

/* One facet's parameters. Field meaning follows the reference XML loader
 * (src/IO/LoaderXML.cpp:831-1126) and the facet "sh" struct it fills. */
typedef struct mfg_facet_desc {
    double sticking;           /* constant sticking factor (Sticking constValue) */
    double opacity;            /* constant opacity, clamped to [0,1] (MolflowSimGeom.cpp:150) */
    double temperature;        /* K */
    double accomodation;       /* accFactor: 1 = full thermalisation (Particle.cpp:1138-1163) */
    double outgassing;         /* Pa*m3/s, constant outgassing */
    double desorb_exponent;    /* N of DES_COSINE_N */
    double diffuse_part;       /* ReflectionParam (MolflowTypes.h:230-241) */
    double specular_part;
    double cosine_exponent;
    double sojourn_freq;       /* Hz, wall sojourn (Particle.cpp:772-776) */
    double sojourn_e;          /* J/mol */
    double tex_width_precise;  /* texDimX: cells along U (texWidth = ceil) */
    double tex_height_precise; /* texDimY */
    double am_theta_limit;     /* angle-map theta split (AnglemapParams) */
    double fh_distance_binsize, fh_distance_max;  /* facet histograms */
    double fh_time_binsize, fh_time_max;
    double og_ratio_u, og_ratio_v;                 /* outgassing-map samples per cm */
    double total_outgassing;   /* Pa*m3/s summed over the outgassing map */
    int32_t sticking_param;    /* index of a time-dependent parameter, -1 = constant */
    int32_t opacity_param;
    int32_t outgassing_param;
    int32_t temperature_param;
    int32_t desorb_type;       /* MFG_DES_* */
    int32_t is_2sided;
    int32_t super_idx;         /* structure, 0-based; -1 = in all structures */
    int32_t super_dest;        /* link facet: destination structure + 1; 0 = none */
    int32_t teleport_dest;     /* destination facet + 1; 0 = none; -1 = back to origin */
    int32_t profile_type;      /* MFG_PROFILE_* */
    int32_t is_moving;
    int32_t enable_sojourn;
    int32_t is_textured;       /* hasMesh */
    int32_t count_des, count_abs, count_refl, count_trans, count_dir;
    int32_t am_record;         /* record the incident-angle map */
    uint32_t am_phi_width, am_theta_lower_res, am_theta_higher_res;
    int32_t fh_record_bounce, fh_record_distance, fh_record_time;
    uint32_t fh_bounce_binsize, fh_bounce_max;
    int32_t use_outgassing_map;
    uint32_t og_map_width, og_map_height;  /* cells of the outgassing map */
    int32_t reserved0;
} mfg_facet_desc;

/* Global simulation settings (reference sp / otfParams, LoaderXML.cpp:96-202). */
typedef struct mfg_settings {
    double gas_mass;              /* g/mol */
    double half_life;             /* s */
    double time_window;           /* timeWindowSize (latestMoment = window/2 with no moments) */
    double low_flux_cutoff;
    double motion_vector1[3];     /* rotation base point */
    double motion_vector2[3];     /* translation velocity or rotation axis*omega */
    double torque_ref[3];
    double scatter_mass_ratio;
    double scatter_mean_free_path_cm;
    double scatter_cutoff_speed;
    double scatter_rho;
    double gh_distance_binsize, gh_distance_max;   /* global histograms */
    double gh_time_binsize, gh_time_max;
    int32_t enable_decay;
    int32_t use_maxwell;
    int32_t calc_constant_flow;
    int32_t low_flux_mode;
    int32_t motion_type;          /* 0 none, 1 translation, 2 rotation */
    int32_t enable_force;         /* enableForceMeasurement */
    int32_t scatter_enabled;
    int32_t scatter_enable_cutoff;
    int32_t gh_record_bounce, gh_record_distance, gh_record_time;
    uint32_t gh_bounce_binsize, gh_bounce_max;
    int32_t log_enabled;          /* particle log (Particle.cpp:1087-1102) */
    int32_t log_facet_id;
    uint32_t log_limit;
} mfg_settings;

/* The whole model as flat arrays. All pointers are read during
 * mfg_load_model only; the library copies what it needs. */
typedef struct mfg_model_view {
    uint32_t struct_size;         /* sizeof(mfg_model_view) */
    uint32_t abi_version;         /* MFG_ABI_VERSION */
    uint32_t n_vertices;
    uint32_t n_facets;
    uint32_t n_structures;
    uint32_t n_params;            /* time-dependent parameters */
    uint32_t n_moments;           /* user moments after parsing/sorting */
    uint32_t reserved0;
    const double* vertices;               /* 3*n_vertices, cm */
    const uint32_t* facet_vertex_offset;  /* n_facets+1 */
    const uint32_t* facet_vertex_index;   /* facet_vertex_offset[n_facets] */
    const mfg_facet_desc* facets;         /* n_facets */
    const mfg_settings* settings;
    const uint32_t* param_offset;         /* n_params+1 */
    const double* param_x;                /* param_offset[n_params] */
    const double* param_y;
    const uint8_t* param_log_x;           /* n_params */
    const uint8_t* param_log_y;
    const double* moment_time;            /* n_moments, centre of window */
    const double* moment_window;          /* n_moments */
    const uint64_t* og_map_offset;        /* n_facets+1 into og_map_values, nullable */
    const double* og_map_values;          /* Pa*m3/s per cell */
    const uint64_t* am_pdf_offset;        /* n_facets+1 into am_pdf (DES_ANGLEMAP), nullable */
    const uint64_t* am_pdf;               /* recorded incident-angle counts of each DES_ANGLEMAP
                                             facet, (am_theta_lower_res + am_theta_higher_res)
                                             lines of am_phi_width counts (LoaderXML.cpp:1043-1073;
                                             sampled as in AnglemapGeneration.cpp:15-410) */
} mfg_model_view;

/* ---- tallies (the output contract, GlobalSimuState) -------------------- */

/* GlobalHitBuffer subset (Particle.cpp:33-49). */
typedef struct mfg_global_counters {
    uint64_t nb_mc_hit;
    double nb_hit_equiv;
    uint64_t nb_desorbed;
    double nb_abs_equiv;
    uint64_t nb_leak_total;
    double dist_traveled_total;
    double dist_traveled_full_hits_only;
    uint64_t nb_decayed;          /* decay + overtime + brownian-cutoff terminations */
} mfg_global_counters;

/* FacetHitBuffer (fields evidenced at Particle.cpp:1193-1230). 64 bytes. */
typedef struct mfg_facet_counters {
    uint64_t nb_mc_hit;
    double nb_hit_equiv;
    uint64_t nb_desorbed;
    double nb_abs_equiv;
    double sum_1_per_ort_velocity;
    double sum_v_ort;
    double sum_1_per_velocity;
    double reserved;
} mfg_facet_counters;

/* TextureCell (MolflowTypes.h:137-152). */
typedef struct mfg_texture_cell {
    double count_equiv;
    double sum_v_ort_per_area;
    double sum_1_per_ort_velocity;
} mfg_texture_cell;

/* ProfileSlice (MolflowTypes.h:118-134). */
typedef struct mfg_profile_slice {
    double count_equiv;
    double sum_v_ort;
    double sum_1_per_ort_velocity;
} mfg_profile_slice;

/* DirectionCell: summed velocity vectors and a count. */
typedef struct mfg_direction_cell {
    double dir[3];
    uint64_t count;
} mfg_direction_cell;

/* ParticleLoggerItem (Particle.cpp:1087-1102). */
typedef struct mfg_log_entry {
    double u, v;
    double hit_theta, hit_phi;
    double ori_ratio;
    double particle_decay_moment;
    double time;
    double velocity;
} mfg_log_entry;

/* Sizes of every tally array, all per moment slot unless noted.
 * A slot is 0 = steady state, 1..n_moments = time moments.
 * Layout rules (both the library and any checker follow them):
 *   facets:      [slot][facet]
 *   forces:      [slot][facet][9]  (impulse, impulse_square, impulse_momentum)
 *   textures:    [slot][cell]; facet f owns cells tex_offset[f] ..+ w*h,
 *                w = ceil(tex_width_precise), h = ceil(tex_height_precise),
 *                for facets with is_textured, in facet order
 *   directions:  [slot][cell] with the same cells, for is_textured && count_dir
 *   profiles:    [slot][slice]; 100 slices per facet with profile_type != 0
 *   angle maps:  [cell] (not per slot); phi_width*(lower+higher) per am_record facet
 *   global hist: [slot][bins]: bounce, then distance, then time bins
 *   facet hist:  [slot][bins]: per facet with a histogram, bounce, distance, time
 *   bounce bins = max/binsize + 2, distance/time bins = (size_t)(max/binsize) + 1
 */
typedef struct mfg_tally_layout {
    uint32_t struct_size;
    uint32_t n_facets;
    uint32_t n_slots;
    uint32_t n_structures;
    uint64_t n_tex_cells;
    uint64_t n_dir_cells;
    uint64_t n_profile_slices;
    uint64_t n_angle_cells;
    uint64_t n_global_hist_bins;
    uint64_t n_facet_hist_bins;
    uint64_t log_capacity;
    uint64_t leak_capacity;
} mfg_tally_layout;

/* Caller-owned buffers that mfg_read_tallies ADDS into (merge semantics of
 * UpdateMCHits, Particle.cpp:21-73). Any pointer may be NULL to skip. */
typedef struct mfg_tally_view {
    uint32_t struct_size;
    uint32_t reserved0;
    mfg_global_counters* global;         /* 1 */
    mfg_facet_counters* facets;          /* n_slots*n_facets */
    double* forces;                      /* n_slots*n_facets*9 */
    mfg_texture_cell* textures;          /* n_slots*n_tex_cells */
    mfg_direction_cell* directions;      /* n_slots*n_dir_cells */
    mfg_profile_slice* profiles;         /* n_slots*n_profile_slices */
    uint64_t* angle_maps;                /* n_angle_cells */
    double* global_hist;                 /* n_slots*n_global_hist_bins */
    double* facet_hist;                  /* n_slots*n_facet_hist_bins */
    mfg_log_entry* log;                  /* log_capacity entries (overwritten, not added) */
    uint64_t* log_count;                 /* entries written into log */
    double* leak_pos_dir;                /* leak_capacity*6 (overwritten) */
    uint64_t* leak_count;
} mfg_tally_view;

/* Result of one mfg_run (MCStepResult, Particle.h:24-28, plus counts). */
typedef struct mfg_step_result {
    int32_t status;               /* 0 Success, 1 DesorptionError, 2 MaxReached */
    int32_t stopped;              /* 1 if mfg_request_stop ended the run early,
                                     2 if the launch time limit did */
    uint64_t desorbed;            /* particles started by this call */
    uint64_t hits;                /* nbMCHit added by this call */
    double kernel_ms;             /* device time of the trace kernel(s) */
} mfg_step_result;

/* ---- context lifecycle ----------------------------------------------- */
typedef struct mfg_context mfg_context;

/* Create a context on a HIP device. Replaces ConstructParticleTracers
 * (MolflowSimulation.cpp:44-55): one context drives every lane of the GPU. */
int mfg_create(mfg_context** out, int device);
void mfg_destroy(mfg_context* ctx);
const char* mfg_last_error(const mfg_context* ctx);
int mfg_abi_version(void);

/* Load a model: facet frames, polygons, texture increments, Maxwell CDF,
 * integrated desorption, moments and the source table are derived here.
 * Replaces LoadSimulation + PrepareToRun (MolflowSimulation.cpp:92-154,
 * MolflowSimGeom.cpp:194-253). Tallies are resized and zeroed. */
int mfg_load_model(mfg_context* ctx, const mfg_model_view* model);

/* Build the per-structure BVH. Replaces RebuildAccelStructure /
 * BuildAccelStructure (MolflowSimulation.cpp:75-88, MolflowSimGeom.cpp:124-184).
 * width = children per node (4 or 8; 0 = default 8; nodes are one 128 B
 * line with 8-bit quantised child boxes either way), leaf_size = max facets
 * per leaf (1..16; 0 = default 2, the reference's setting). MFG_ERR_LIMIT
 * when the leaves hold 2^27 facet records or more. */
int mfg_build_accel(mfg_context* ctx, int width, int leaf_size);

/* Zero tallies and restart the particle stream for a seed. Replaces
 * ResetSimulation (MolflowSimulation.cpp:164-178). particle_offset lets a
 * resumed run keep RNG streams distinct (set it to the loaded nbDesorbed). */
int mfg_reset(mfg_context* ctx, uint64_t seed, uint64_t particle_offset);

/* Desorb n_des new particles and trace all of them until absorption, leak
 * or a volume event. Replaces the controller loop around SimulationMCStep
 * (Particle.cpp:207-396) with a desorption limit. max_events bounds the
 * events of any one particle (0 = unbounded). Synchronous. */
int mfg_run(mfg_context* ctx, uint64_t n_des, uint64_t max_events, mfg_step_result* result);

/* Ask a running mfg_run (from another thread) to stop desorbing; in-flight
 * particles finish. The exitRequested analogue (Particle.h:122). */
int mfg_request_stop(mfg_context* ctx);

/* Wall-clock limit of one trace launch (default 60 s, 0 < seconds <= 86400).
 * A wave that reaches it starts no particle, parks its particles in flight
 * (carry enabled) or ends them where they are (carry disabled), and the run
 * returns with stopped = 2. Parked particles the launch never resumed stay
 * parked. It bounds every launch, also one whose particles
 * can never be absorbed and whose max_events is 0. No counterpart in the
 * reference, whose SimulationMCStep returns after a given number of events
 * (Particle.cpp:207-396) instead. */
int mfg_set_time_limit(mfg_context* ctx, double seconds);

/* Sizes and per-facet offsets of the tally arrays. kind: 0 texture cell
 * offset, 1 direction cell offset, 2 profile slice offset, 3 angle-map cell
 * offset, 4 facet-histogram bin offset; -1 where a facet has none. */
int mfg_get_layout(const mfg_context* ctx, mfg_tally_layout* layout);
int mfg_get_facet_offsets(const mfg_context* ctx, int kind, int64_t* out, size_t n);

/* Add the device tallies into caller buffers (UpdateMCHits + GlobalSimuState
 * +=, Particle.cpp:21-73, MolflowSimGeom.cpp:452-458). Device tallies are
 * not cleared (mfg_take_tallies clears them, mfg_reset restarts the run). */
int mfg_read_tallies(mfg_context* ctx, const mfg_tally_view* out);

/* Periodic report: add the device tallies into caller buffers, then zero
 * them, keeping the particles parked between runs, the particle counter
 * and the random streams. The UpdateHitsAndLog role (Particle.cpp:1294-1312:
 * globalState += tmpState, then tmpState->Reset()), so k runs with a take
 * after each add up to one uninterrupted run. The log and leak buffers are
 * copied (not added) and restart empty. One model only (MFG_ERR_STATE with
 * variants: read each variant, then call mfg_clear_tallies). */
int mfg_take_tallies(mfg_context* ctx, const mfg_tally_view* out);
/* Zero the device tallies of every loaded variant, keeping parked
 * particles and streams (tmpState->Reset(), Particle.cpp:1309). */
int mfg_clear_tallies(mfg_context* ctx);
/* Empty the particle log (ReinitializeParticleLog,
 * MolflowSimulation.cpp:15-35). */
int mfg_reinit_particle_log(mfg_context* ctx);

/* Check a model view without loading it (SanityCheckModel,
 * MolflowSimulation.cpp:57-71, MolflowSimGeom.cpp:337-404): runs the same
 * preparation as mfg_load_model plus the reference's own checks, writes one
 * line per problem into out (cap bytes, NUL-terminated, may be NULL) and
 * their count into *n_problems. Needs no device. Returns MFG_OK whenever
 * the check ran; n_problems == 0 means the model loads. */
int mfg_sanity_check(const mfg_model_view* model, char* out, size_t cap, uint32_t* n_problems);

/* The model totals the physical quantities are scaled with
 * (CalcTotalOutgassing, MolflowSimGeom.cpp:290-335; read by
 * GetMoleculesPerTP, CSVExporter.cpp:59-79): out[0] final outgassing rate
 * (molecules/s), out[1] the same in Pa m3/s, out[2] total desorbed
 * molecules up to the latest moment, out[3] the latest moment (s). Needs no
 * device. Returns the error mfg_load_model would give for the view. */
int mfg_model_totals(const mfg_model_view* model, double out[4]);

/* Bytes of one full tally state (GetHitsSize, MolflowSimulation.cpp:156-160). */
int mfg_hits_size(const mfg_context* ctx, uint64_t* bytes);

/* ---- fixed ray batches (intersection parity) --------------------------- */
/* First hard hit of n rays in structure `structure`, skipping facet
 * last_hit[i] (-1 = none). The role of BVHAccel::Intersect (called at
 * Particle.cpp:245) with opacity treated as 1. Outputs facet id (-1 = miss),
 * distance t (cm) and facet coordinates (u,v). Host buffers. */
int mfg_intersect_batch(mfg_context* ctx, uint64_t n, const double* origins, const double* dirs,
                        const int32_t* last_hit, int32_t structure,
                        int32_t* out_facet, double* out_t, double* out_u, double* out_v);

/* mfg_intersect_batch without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as mfg_intersect_batch; the rays
 * and results are staged through a fixed device array that the library
 * holds on every GPU it runs on (about 5 MB, for MFG_STATIC_MAX_RAYS rays),
 * so n above MFG_STATIC_MAX_RAYS is refused with MFG_ERR_LIMIT. Calls on
 * one GPU share that array: calls that overlap in time on one GPU (from
 * several contexts or threads) race and must be serialised by the caller. */
#define MFG_STATIC_MAX_RAYS 65536
int mfg_intersect_batch_static(mfg_context* ctx, uint64_t n, const double* origins, const double* dirs,
                               const int32_t* last_hit, int32_t structure,
                               int32_t* out_facet, double* out_t, double* out_u, double* out_v);

/* ---- derived tables, exposed for checking ------------------------------ */
/* Maxwell wall-collision speed CDF at 1 K (CDFGeneration.cpp:55-83):
 * 100 (velocity, cumulative probability) pairs. */
int mfg_maxwell_cdf(double gas_mass, double* out_velocity, double* out_cumulative, size_t n);
/* Parse "begin,spacing,end" (or a single time) into moments
 * (TimeMoments.cpp:58-83). Writes up to cap times; *n gets the count. */
int mfg_parse_user_moment(const char* content, double window, double* out_times, size_t cap, size_t* n);
/* Integrated desorption table of a time-dependent outgassing parameter
 * (IDGeneration.cpp:45-206) given latestMoment. Writes up to cap points. */
int mfg_integrated_desorption(const double* x, const double* y, size_t n_points, int log_x, int log_y,
                              double latest_moment, double* out_t, double* out_cum, size_t cap, size_t* n);

/* ---- per-particle event trace (testing aid) ---------------------------- */
/* Record, for particles with index < n_particles, the facet id and event
 * kind of their first k events during later mfg_run calls. kind: 0 desorb,
 * 1 bounce, 2 absorb, 3 transparent pass, 4 teleport, 5 link, 6 leak,
 * 7 overtime/decay, 8 scatter. n_particles = 0 disables tracing. */
int mfg_set_trace(mfg_context* ctx, uint32_t n_particles, uint32_t k);
int mfg_get_trace(mfg_context* ctx, int32_t* facet, int32_t* kind, size_t n);

/* ---- batched variants (many geometry/sticking variants, one launch) ---- */
/* Variant v is a whole model; all variants are traced by one persistent
 * launch that takes (variant, particle) work from a global queue.
 * n_des[v] desorptions per variant, seeds[v] its RNG seed. Tallies of
 * variant v are read with mfg_read_variant_tallies. */
int mfg_load_variants(mfg_context* ctx, const mfg_model_view* const* models, uint32_t n_variants);
int mfg_run_variants(mfg_context* ctx, const uint64_t* n_des, const uint64_t* seeds,
                     uint64_t max_events, mfg_step_result* result);
int mfg_get_variant_layout(const mfg_context* ctx, uint32_t variant, mfg_tally_layout* layout);
int mfg_read_variant_tallies(mfg_context* ctx, uint32_t variant, const mfg_tally_view* out);
/* Launch shape used by runs: threads per block (a multiple of 64 up to
 * 256, 768 or 1024) and blocks per CU (0 = defaults / autotuned). 768 and
 * 1024 run one block per CU at 3 and 4 waves per SIMD, whose LDS copy of
 * the top of the BVH is then shared by the whole CU; they apply to models
 * without transparency, moments, volume events and the other features of
 * the full kernel (those runs take 256-thread blocks). MFG_ERR_INVALID for
 * other sizes. */
int mfg_set_launch(mfg_context* ctx, uint32_t block_threads, uint32_t blocks_per_cu);
/* Particles kept in flight between runs (the particle state that persists
 * between SimulationMCStep calls, reference Particle.h:95-122 and
 * MolflowSimulation.cpp SimulationMCStep loop).
 * enable = 1: when a run's desorption queue is empty, the particles still in
 *   flight are parked on the device and the next run resumes them before
 *   desorbing new ones, so a run never waits for the longest trajectories.
 *   Tallies then hold every event traced so far; the counts of a resumed
 *   particle continue its own random stream, so results do not depend on how
 *   the work is split into runs.
 * enable = 0 (default): every run traces all its particles, and all parked
 *   ones, to their end. A run of 0 desorptions with carry disabled drains.
 * Not available together with the adjoint (MFG_ERR_STATE). mfg_reset and
 * mfg_load_* drop parked particles. */
int mfg_set_carry(mfg_context* ctx, int enable);
/* Number of particles parked on the device after the last run. */
int mfg_in_flight(const mfg_context* ctx, uint64_t* n);

/* Occupancy bound of the trace kernel: waves per SIMD (1..5; 0 = default).
 * Higher bounds cap registers (cold state spills) to hide more latency. */
int mfg_set_kernel_variant(mfg_context* ctx, uint32_t waves_per_simd);

/* Launch autotuner: times short probe runs of n_probe desorptions for every
 * BVH shape (8-wide with 2 or 1 facets per leaf, 4-wide with 2), compiled
 * occupancy bound and block size (128, 256, and 768 and 1024 where the
 * model's kernel has them) on the loaded model(s) (particles stay in
 * flight between probes), keeps the
 * fastest, then zeroes the tallies and puts the particle streams back where
 * they were, so the probes leave no trace in later results. */
typedef struct mfg_tune_result {
    uint32_t waves;          /* chosen waves per SIMD */
    uint32_t block;          /* threads per block */
    uint32_t blocks_per_cu;  /* resident blocks per CU at that shape */
    uint32_t n_hot;          /* facets whose slot-0 counters are summed per block in LDS
                                (those with >= 0.5 % of the probes' counter updates) */
    double hits_per_s;       /* probe throughput of the choice */
    uint32_t bvh_width;      /* children per BVH node of the choice (the BVH is rebuilt with it) */
    uint32_t leaf_size;      /* facets per BVH leaf of the choice */
} mfg_tune_result;
int mfg_autotune(mfg_context* ctx, uint64_t n_probe, mfg_tune_result* out);

/* Capture the variant run in a hipGraph and replay it on later calls. */
int mfg_enable_graph(mfg_context* ctx, int enable);

/* ---- sticking adjoint --------------------------------------------------- */
/* Declare sticking parameter groups and target facets. group_of_facet[f]
 * is the group (0..n_groups-1) whose sticking value facet f takes, or -1.
 * Requires low-flux mode (the path does not depend on sticking). During
 * later mfg_run calls the forward pass accumulates, for every target t,
 * J0[t] = sum of hit weights (nbHitEquiv) and the path-score Jacobian
 * J[t][g] = d J0[t] / d s_g in FP64. Each group starts from the constant
 * sticking its facets share (MFG_ERR_INVALID if they differ or one has a
 * time-dependent sticking); if that value is not inside (0,1), runs are
 * refused (MFG_ERR_STATE) until mfg_adjoint_set_sticking sets one. */
int mfg_adjoint_setup(mfg_context* ctx, const int32_t* group_of_facet, uint32_t n_groups,
                      const int32_t* target_facets, uint32_t n_targets);
/* Set the sticking value of each group (overrides facet sticking). */
int mfg_adjoint_set_sticking(mfg_context* ctx, const double* sticking, uint32_t n_groups);
/* Read J0 (n_targets) and J (n_targets*n_groups, row-major). */
int mfg_adjoint_read(mfg_context* ctx, double* j0, double* jac);
/* Backward: grad[g] = sum_t upstream[t] * J[t][g] / n_des, computed on the
 * device from the accumulated Jacobian. */
int mfg_adjoint_backward(mfg_context* ctx, const double* upstream, double* grad_out);

#ifdef __cplusplus
}
#endif

#endif /* MOLFLOW_GPU_H */
