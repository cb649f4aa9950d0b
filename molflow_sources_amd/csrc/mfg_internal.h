// mfg_internal.h -- device data layout of the MI355X tracer, shared by the
// host runtime (mfg_host.cpp) and the HIP kernels (mfg_trace.hip).
//
// HBM layout (all variants concatenated; a plain model is one variant):
//   DFacet[nF_total]      128 B, one cache line per ray/facet test
//   DPhys[nF_total]       physics and recording parameters, read on a hit
//   double2 poly[]        normalised (u,v) polygon vertices
//   BvhNode[]             up to 8-wide nodes, 128 B each, 8-bit quantised boxes
//   LeafFacet[]           the ray/facet test data in leaf order (128 B) and
//   LeafUV[]              triangle/quad (u,v) polygons (64 B): a leaf is a
//                         contiguous run, no index indirection
//   DVariant[nV]          per-variant bases and settings
//   tallies               [variant][slot][facet] 64 B counters, textures,
//                         profiles, ... (layout of include/molflow_gpu.h)
#pragma once
#include <stdint.h>

#include "../../include/molflow_gpu.h"

#define MFG_ADJ_SEG 32            // target scores kept per lane for the adjoint's cutoff term
#define MFG_BVH_WIDTH 8          // children per node (a 4-wide build uses 4 of the 8 slots)
#define MFG_MAX_STACK 64
#ifndef MFG_LDS_STACK
#define MFG_LDS_STACK 16         // traversal stack entries per lane kept in LDS by the tracer
#endif
#define MFG_MAX_GROUPS 64
#define MFG_MAXWELL_GUIDE 128
#define MFG_MAX_HOT 32           // facets whose counters are privatised per block

// feature bits of a trace-kernel instantiation / of a loaded model
enum : uint32_t {
    FEAT_TRANSPARENT = 1u << 0,   // opacity < 1 somewhere (coin + pass traversal)
    FEAT_MOMENTS = 1u << 1,       // time moments
    FEAT_TDPARAM = 1u << 2,       // time-dependent sticking/opacity/temperature/outgassing
    FEAT_VOLUME = 1u << 3,        // overtime, decay or scattering
    FEAT_TELEPORT = 1u << 4,      // teleport or link facets
    FEAT_MOVING = 1u << 5,
    FEAT_SOJOURN = 1u << 6,
    FEAT_LOWFLUX = 1u << 7,
    FEAT_FORCE = 1u << 8,
    FEAT_HIST = 1u << 9,
    FEAT_LOG = 1u << 10,
    FEAT_OGMAP = 1u << 11,
    FEAT_TRACE = 1u << 12,
    FEAT_AMSRC = 1u << 13,        // DES_ANGLEMAP sources
    FEAT_ADJOINT = 1u << 14,
    FEAT_ALL = (1u << 14) - 1,
};

// facet flags (DFacet::flags, DPhys::flags)
enum : uint32_t {
    FF_2SIDED = 1u << 0,
    FF_OPAQUE = 1u << 1,        // constant opacity 1
    FF_TRANSP = 1u << 2,        // constant opacity 0
    FF_TD_OPACITY = 1u << 3,
    FF_TEXTURE = 1u << 4,
    FF_COUNT_DES = 1u << 5,
    FF_COUNT_ABS = 1u << 6,
    FF_COUNT_REFL = 1u << 7,
    FF_COUNT_TRANS = 1u << 8,
    FF_COUNT_DIR = 1u << 9,
    FF_PROFILE = 1u << 10,
    FF_ANGLEMAP = 1u << 11,
    FF_MOVING = 1u << 12,
    FF_SOJOURN = 1u << 13,
    FF_FHIST = 1u << 14,
    FF_TELEPORT = 1u << 15,
    FF_LINK = 1u << 16,
    FF_DIFFUSE = 1u << 17,      // diffuse_part > 0.999999
    FF_FULL_ACC = 1u << 18,     // accomodation > 0.9999
    FF_OGMAP = 1u << 19,
};

struct alignas(16) DFacet {
    double O[3];
    double N[3];
    double U[3];
    double V[3];
    double invU2, invV2;
    uint32_t poly_off;
    uint32_t poly_n;
    uint32_t flags;
    float opacity;
};
static_assert(sizeof(DFacet) == 128, "DFacet must be one 128-byte line");

struct alignas(16) DPhys {
    double nU[3], nV[3];
    double sticking, opacity, temperature, sqrt_temp, accomodation;
    double diffuse, specular, cos_exp, des_exp;
    double sojourn_freq, sojourn_e;
    double tex_wp, tex_hp;          // precise texture size
    double max_speed;
    double am_theta_limit;
    double fh_dist_bin, fh_time_bin;
    double og_wp, og_hp;            // outgassing map precise size
    double centre[3];
    int64_t tex_off, dir_off;       // cell offsets within the variant, -1 none
    int32_t prof_off, am_off;       // -1 none
    int32_t fh_off;                 // facet histogram offset, -1 none
    int32_t tex_w, tex_h;
    uint32_t flags;
    int32_t profile_type;
    int32_t super_idx, super_dest, teleport_dest;
    int32_t sticking_param, opacity_param, temperature_param, outgassing_param;
    int32_t id_index;               // integrated-desorption table (variant-local), -1
    int32_t local_id;               // facet index within its variant
    uint32_t am_phi, am_lo, am_hi;
    uint32_t fh_bounce_bin, fh_nb_bounce, fh_nb_dist, fh_nb_time;
    int32_t adj_group;              // sticking group, -1
    int32_t adj_target;             // adjoint target index, -1
    uint32_t og_w, og_h;
    int64_t og_off;                 // into og cdf array (global), -1
    int32_t desorb_type;
    int32_t pad_;
    int64_t am_src_off;             // DES_ANGLEMAP source table in the pool (global), -1
    int32_t hot_idx;                // slot in the per-block hot-facet counter table, -1
    int32_t pad2_;
};

// One BVH node of up to 8 children in one 128 B line. Child boxes are
// quantised to 8-bit steps per plane relative to the node's origin with a
// power-of-two step per axis: lo = o + qlo * 2^e, hi = o + qhi * 2^e in FP32
// (the builder checks that these FP32 values enclose the padded FP64 box).
// The steps are stored as the FP16 values of those integers (exact), which
// the kernel converts and scales in one v_fma_mix_f32. Children are found
// without per-child words: a node's inner children are consecutive nodes
// from node_base in slot order, and the leaf records of its leaf children
// are consecutive from leaf_base in slot order, each leaf holding 1 + its
// 4-bit count facets.
struct alignas(128) BvhNode {
    float ox, oy, oz;           // quantisation origin
    int8_t ex, ey, ez;          // per-axis step exponents
    uint8_t leaf_mask;          // bit c set when child slot c is a leaf
    int32_t node_base;          // first inner child (0 without inner children)
    int32_t leaf_base;          // first leaf record of the first leaf child (0 without leaves)
    uint32_t counts;            // bits 4c..4c+3: facets - 1 of a leaf child in slot c
    uint8_t inner_mask;         // bit c set when child slot c is an inner node
    uint8_t pad[3];
    uint16_t qlo[3][8];         // [axis][child], FP16 bit patterns of 0..255
    uint16_t qhi[3][8];
};
static_assert(sizeof(BvhNode) == 128, "BvhNode must be 128 bytes");

// Exact (FP64) ray/facet test data of one leaf entry, stored in leaf order
// so that a leaf's facets are contiguous (no index indirection). One record
// is one 128 B line, loaded as eight 16 B loads issued together.
struct alignas(128) LeafFacet {
    int32_t id;            // global facet index
    uint32_t flags;        // FF_* of the facet
    uint32_t poly_off;     // first (u,v) pair in poly (polygons of > 3 vertices)
    uint32_t poly_n;
    double O[3], N[3], U[3], V[3];
    double invU2, invV2;
};
static_assert(sizeof(LeafFacet) == 128, "LeafFacet must be one 128 B line");
// (u,v) of the vertices of a triangle or quadrilateral, parallel to
// LeafFacet (read only when the ray meets the facet's (u,v) rectangle);
// larger polygons are read from poly
struct alignas(64) LeafUV {
    double uv[8];
};
static_assert(sizeof(LeafUV) == 64, "LeafUV layout");

// global counters of one variant (mfg_global_counters)
struct DGlobal {
    unsigned long long nb_mc_hit;
    double nb_hit_equiv;
    unsigned long long nb_desorbed;
    double nb_abs_equiv;
    unsigned long long nb_leak;
    double dist_total;
    double dist_full;
    unsigned long long nb_decayed;
};

struct DVariant {
    // geometry
    uint32_t facet_base, n_facets;
    uint32_t root_base, n_struct;      // into root table (one BVH root per structure)
    // source table
    uint32_t src_base, src_n;          // cumulative weights / facet ids
    double total_des;
    // settings
    double latest, gas_mass, half_life, low_flux_cutoff, c_factor;
    double scatter_mr, scatter_mfp, scatter_cut, scatter_rho;
    double motion1[3], motion2[3], torque[3];
    double gh_dist_bin, gh_time_bin;
    uint32_t flags;                    // FEAT_* actually used by this variant
    int32_t use_maxwell, calc_const_flow, enable_decay, scatter_enabled, scatter_cutoff, low_flux, motion_type;
    int32_t enable_force;
    uint32_t gh_bounce_bin, gh_nb_bounce, gh_nb_dist, gh_nb_time;
    int32_t log_facet;                 // global facet id or -1
    uint32_t log_limit;
    // tables (global offsets)
    uint32_t maxwell_base;             // 100 (v,p) pairs
    uint32_t maxwell_guide;            // MFG_MAXWELL_GUIDE entries into `guides`
    uint32_t param_base, n_params;     // into param table headers
    uint32_t id_base, n_ids;           // into ID table headers
    uint32_t mom_base, n_moments;      // into interval arrays
    // tally bases
    uint64_t fc_base;                  // facet counters, [slot][facet]
    uint64_t force_base;               // doubles
    uint64_t tex_base, n_tex;          // texture cells per slot
    uint64_t dir_base, n_dir;
    uint64_t prof_base, n_prof;
    uint64_t am_base;
    uint64_t gh_base, n_gh;
    uint64_t fh_base, n_fh;
    uint32_t n_slots, glob_index;
    // run
    uint64_t seed;
    uint64_t pid_base;                 // particle index offset of this run
};

struct DTable {                        // a (x,y) table in the global table pool
    uint32_t off, n;
    int32_t log_x, log_y;
};

// Everything a trace launch needs; passed by value (kernarg segment).
// One particle still in flight at the end of a launch, kept on the device
// and resumed first by the next launch (the particle state that persists
// between the reference's SimulationMCStep calls, Particle.h:95-122).
struct CarryRec {
    double pos[3], dir[3];
    double time, vel, ori, gen_time, dist, decay_t, scatter_path, hu, hv;
    uint64_t pid, events_left;
    uint32_t k0, k1, blk, b0, b1, b2, b3, avail;   // random stream position
    int32_t structure, last_hit, tele_from;
    uint32_t nb_bounces, nb_events, variant, tr_n;
};
static_assert(sizeof(CarryRec) == 200, "CarryRec layout");

struct TraceParams {
    const DFacet* facets;
    const DPhys* phys;
    const double* poly;               // (u,v) pairs
    const BvhNode* nodes;
    const LeafFacet* leaves;
    const LeafUV* leaf_uv;
    const int32_t* roots;             // per (variant, structure): root node or -1
    const DVariant* variants;
    const uint64_t* var_prefix;       // n_variants+1 prefix of this run's desorptions
    uint32_t n_variants;
    uint32_t stack_size;              // LDS stack entries per lane
    uint32_t stack_total;             // entries a walk may need; the ones past stack_size
    int32_t* stack_spill;             // live in this per-lane global area (lane-major)
    const double* src_cum;            // cumulative source weights
    const int32_t* src_facet;
    const double* og_cdf;
    const uint32_t* guides;           // guide tables of the Maxwell CDFs
    const double* tab_x;              // table pool (maxwell CDFs, params, IDs)
    const double* tab_y;
    const DTable* params;
    const DTable* ids;
    const double* mom_start;
    const double* mom_end;
    // DES_ANGLEMAP source tables, one block per facet:
    // [W, L, H, theta limit, lower ratio, total count, line sums (T),
    //  theta CDF (T), phi CDFs (T*W), phi pdfs (T*W)], T = L + H
    const double* am_src;
    // hot facets: their slot-0 counter updates are summed in LDS per block
    // and added to the global records once, when the block ends
    const uint64_t* hot_fc;           // index into fc of each hot facet's slot-0 record
    uint32_t n_hot;
    // tallies
    mfg_facet_counters* fc;
    double* force;
    mfg_texture_cell* tex;
    const double* tex_inc;            // per texture cell (same layout as slot 0)
    mfg_direction_cell* dir;
    mfg_profile_slice* prof;
    unsigned long long* am;
    double* gh;
    double* fh;
    DGlobal* glob;
    mfg_log_entry* log;
    unsigned long long* log_count;
    double* leaks;
    unsigned long long* leak_count;
    uint32_t leak_cap;
    // run control
    unsigned long long* next_particle; // work queue head
    uint64_t total;                    // desorptions in this run
    uint64_t max_events;
    volatile int32_t* stop_flag;       // host-mapped
    // a wave that has run for time_limit ticks of the device wall clock
    // starts nothing more, parks (carry) or ends its particles, sets
    // *expired and leaves: every launch ends even if no particle can
    uint64_t time_limit;
    unsigned long long* expired;
    // particles kept between launches: two halves of carry_cap records;
    // carry_ctl[0..1] count the records of each half, carry_ctl[2] names the
    // half this launch resumes from (the other half receives its leftovers)
    CarryRec* carry;
    unsigned long long* carry_ctl;
    uint64_t carry_cap;
    uint32_t carry_keep;               // 1: leave particles in flight when the queue runs dry
    // nodes [0, lds_nodes) (the top of the breadth-first node array) are
    // read from a copy each block keeps in LDS
    uint32_t lds_nodes;
    // trace
    int32_t* trace_facet;
    int32_t* trace_kind;
    uint32_t trace_n, trace_k;
    // adjoint
    double* adj_j0;
    double* adj_jac;
    const double* adj_sticking;
    uint32_t* adj_counts;             // per lane: sticky collisions per group
    // per lane: the target scores recorded since the particle last survived
    // a sticky collision (MFG_ADJ_SEG entries), for the cutoff term; a lane
    // with more sets *adj_seg_overflow
    int32_t* adj_seg_t;
    double* adj_seg_x;
    unsigned long long* adj_seg_overflow;
    uint32_t adj_groups, adj_targets;
};

struct IsectParams {
    const DFacet* facets;
    const double* poly;
    const BvhNode* nodes;
    const LeafFacet* leaves;
    const LeafUV* leaf_uv;
    int32_t root;
    uint32_t stack_size;
    uint64_t n;
    const double* origins;
    const double* dirs;
    const int32_t* last_hit;
    int32_t* out_facet;
    double* out_t;
    double* out_u;
    double* out_v;
    uint32_t facet_base;
};

// Staging layout of the allocation-free intersect entry: the library's
// static device array `scratch` (mfg_trace.hip) holds one of these
#define MFG_MAX_DEVICES 64
struct StaticRays {
    double origins[3 * MFG_STATIC_MAX_RAYS];
    double dirs[3 * MFG_STATIC_MAX_RAYS];
    double t[MFG_STATIC_MAX_RAYS], u[MFG_STATIC_MAX_RAYS], v[MFG_STATIC_MAX_RAYS];
    int32_t last_hit[MFG_STATIC_MAX_RAYS];
    int32_t facet[MFG_STATIC_MAX_RAYS];
};
// device address of `scratch` on `device`, which must be the current
// device; looked up once per device (hipGetSymbolAddress), 0 on success
int mfg_scratch_address(int device, StaticRays** out);

// kernel launchers (mfg_trace.hip)
int mfg_launch_trace(const TraceParams& p, uint32_t features, int waves, int grid, int block, void* stream);
int mfg_launch_intersect(const IsectParams& p, void* stream);
// after a carry launch: append the parked particles the launch did not
// claim (time limit) to the output half, then swap the halves
int mfg_launch_carry_flip(unsigned long long* carry_ctl, const unsigned long long* claimed, CarryRec* carry,
                          uint64_t cap, unsigned long long* overflow, void* stream);
int mfg_launch_zero(void* ptr, uint64_t bytes, void* stream);
int mfg_launch_backward(const double* jac, const double* up, double* grad, uint32_t n_targets, uint32_t n_groups,
                        double scale, void* stream);
uint32_t mfg_trace_kernel_features(uint32_t model_features);
int mfg_trace_occupancy(uint32_t features, int waves, int block, size_t lds_bytes);
int mfg_trace_variants(uint32_t features, int* waves_out, int cap);
// 1 when the feature set has a one-block-per-CU kernel of `block` lanes
// (768: 3 waves per SIMD, 1024: 4)
int mfg_trace_has_cu_block(uint32_t features, int block);
