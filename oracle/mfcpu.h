/*
 * mfcpu.h -- CPU oracle for the MolFlow+ test-particle tracer.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain FP64 C++ restatement of the
 * reference's CPU ParticleTracer (src/Simulation/Particle.cpp and callees),
 * used by tests/ as the parity checker and by bench.py as the timed CPU
 * baseline. The product library (molflow_sources_amd) never links, loads or
 * calls it.
 *
 * It reads the same mfg_model_view as the GPU library and fills tallies in
 * the layout documented in include/molflow_gpu.h.
 */
#ifndef MFCPU_H
#define MFCPU_H

#include "../include/molflow_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfcpu_model mfcpu_model;

/* Prepare a model (frames, polygons, tables, BVH). NULL on error; the
 * message is copied into err (cap bytes). */
mfcpu_model* mfcpu_load(const mfg_model_view* view, char* err, size_t cap);
void mfcpu_free(mfcpu_model* m);
int mfcpu_layout(const mfcpu_model* m, mfg_tally_layout* layout);
int mfcpu_facet_offsets(const mfcpu_model* m, int kind, int64_t* out, size_t n);

/* Trace particles first_particle .. first_particle+n_des-1 (RNG streams are
 * keyed by seed and particle index), adding into `out`. threads <= 0 uses
 * all OpenMP threads. Returns 0 on success. */
int mfcpu_run(mfcpu_model* m, uint64_t seed, uint64_t first_particle, uint64_t n_des,
              uint64_t max_events, int threads, const mfg_tally_view* out);

/* Per-particle event trace of particles first..first+n-1, k events each
 * (same kind codes as mfg_set_trace). */
int mfcpu_trace(mfcpu_model* m, uint64_t seed, uint64_t first_particle, uint32_t n, uint32_t k,
                int32_t* facet, int32_t* kind);

/* First hard hit of each ray; brute != 0 tests every facet, else the BVH. */
int mfcpu_intersect_batch(mfcpu_model* m, uint64_t n, const double* origins, const double* dirs,
                          const int32_t* last_hit, int32_t structure, int brute,
                          int32_t* out_facet, double* out_t, double* out_u, double* out_v);

/* Mean BVH nodes and facet tests visited per ray for a batch (roofline
 * accounting of algorithmic bytes). */
int mfcpu_bvh_stats(mfcpu_model* m, uint64_t n, const double* origins, const double* dirs,
                    const int32_t* last_hit, int32_t structure, double* mean_nodes, double* mean_prims);

/* Philox4x32-10 block (for known-answer tests). */
void mfcpu_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Derived tables. */
int mfcpu_maxwell_cdf(double gas_mass, double* v, double* p, size_t n);
int mfcpu_parse_user_moment(const char* content, double window, double* out, size_t cap, size_t* n);
int mfcpu_integrated_desorption(const double* x, const double* y, size_t n_points, int log_x, int log_y,
                                double latest_moment, double* out_t, double* out_cum, size_t cap, size_t* n);
double mfcpu_facet_area(const mfcpu_model* m, uint32_t facet);
int mfcpu_facet_frame(const mfcpu_model* m, uint32_t facet, double* o_u_v_n /* 12 */);
int mfcpu_texture_increments(const mfcpu_model* m, uint32_t facet, double* out, size_t n);
double mfcpu_total_desorbed_molecules(const mfcpu_model* m);
double mfcpu_final_outgassing_rate(const mfcpu_model* m);
double mfcpu_latest_moment(const mfcpu_model* m);

/* Test hook of the hard-sphere scattering law: one scatter of each of n
 * particles moving along +z at unit speed; writes the new direction's z
 * component and the new speed. */
int mfcpu_scatter_sample(double mass_ratio, double rho, uint64_t n, uint64_t seed, double* out_cos,
                         double* out_speed);

/* Sticking adjoint (low-flux mode): J0[t] and J[t][g] accumulated over
 * particles first..first+n_des-1, with group sticking values overriding
 * the facet sticking of grouped facets. */
int mfcpu_adjoint_run(mfcpu_model* m, uint64_t seed, uint64_t first_particle, uint64_t n_des,
                      const int32_t* group_of_facet, uint32_t n_groups, const double* group_sticking,
                      const int32_t* targets, uint32_t n_targets, int threads,
                      double* j0, double* jac, uint64_t max_events);

#ifdef __cplusplus
}
#endif
#endif
