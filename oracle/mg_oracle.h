/*
 * mg_oracle.h — C-ABI of the CPU oracle for the Morphogenesis SPH step.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain C++ restatement of the reference
 * step (src/fluid_system_cuda.cu, src/fluid_system.cpp, src/host_CUDA.cpp) used by
 * tests/ as the checker and by bench.py as the CPU baseline. The product path
 * (morphogenesis_amd) never links or calls it.
 *
 * The entry points mirror include/mg_sph.h one for one (mgo_* vs mg_*), so a test can
 * drive both with the same call sequence. `precision` selects FP32 (32) or FP64 (64)
 * arithmetic; the FP64 variant is used for drift statements only.
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H

#include "../include/mg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgo_handle mgo_handle;

int mgo_create(mgo_handle** out, int max_points, int bonds_per_particle, int precision);
int mgo_destroy(mgo_handle* h);
int mgo_set_params(mgo_handle* h, const MgParams* p);
int mgo_set_grid(mgo_handle* h, const MgGrid* g);
int mgo_set_genome(mgo_handle* h, const MgGenome* g);
int mgo_set_freeze(mgo_handle* h, int freeze);
int mgo_set_threads(mgo_handle* h, int nthreads);
int mgo_upload(mgo_handle* h, int buf_id, const void* src, size_t bytes);
int mgo_download(mgo_handle* h, int buf_id, void* dst, size_t bytes);
int mgo_get_counts(mgo_handle* h, uint32_t* num_points, uint32_t* active_points);

/* phases (reference names in comments) */
int mgo_insert_particles(mgo_handle* h);              /* InsertParticlesCL   */
int mgo_prefix_sum_cells(mgo_handle* h);              /* PrefixSumCellsCL(1) */
int mgo_counting_sort_full(mgo_handle* h);            /* CountingSortFullCL  */
int mgo_compute_pressure(mgo_handle* h);              /* ComputePressureCL   */
int mgo_compute_force(mgo_handle* h);                 /* ComputeForceCL      */
int mgo_advance(mgo_handle* h, float time, float dt, float ss); /* AdvanceCL  */
int mgo_compute_diffusion(mgo_handle* h);             /* ComputeDiffusionCL  */
int mgo_compute_genes(mgo_handle* h);                 /* ComputeGenesCL      */
int mgo_zero_vel(mgo_handle* h);                      /* ZeroVelCL           */
int mgo_initialize_bonds(mgo_handle* h);              /* InitializeBondsCL   */
float mgo_acosf(float x);                             /* the bond-angle acos */
int mgo_compute_bond_changes(mgo_handle* h, uint32_t steps); /* ComputeBondChangesCL */
int mgo_prefix_sum_changes(mgo_handle* h, int zero_offsets); /* PrefixSumChangesCL   */
int mgo_counting_sort_changes(mgo_handle* h);         /* CountingSortChangesCL */
int mgo_compute_particle_changes(mgo_handle* h, uint64_t seed); /* ComputeParticleChangesCL (heal) */
float mgo_tanhf(float x);                             /* the remodelling tanh   */
float mgo_expf(float x);                              /* its exp                */
uint32_t mgo_heal_word(uint64_t seed, uint32_t pid);  /* heal's random word     */

/* composites */
int mgo_physical_sort(mgo_handle* h);                 /* Run2PhysicalSort    */
int mgo_inner_physical_loop(mgo_handle* h, float time, float dt); /* Run2InnerPhysicalLoop */
int mgo_gene_action(mgo_handle* h);                   /* Run2GeneAction      */
/* n steps of [inner loop, (gene action), physical sort]; time advances by dt each step */
int mgo_step(mgo_handle* h, int nsteps, float dt, unsigned flags);

#ifdef __cplusplus
}
#endif
#endif
