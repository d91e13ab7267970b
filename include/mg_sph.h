/*
 * mg_sph.h — C-ABI of the MI355X (gfx950) Morphogenesis SPH step library (libmg_sph.so).
 *
 * Plain pointers and sizes only; no device pointers cross this boundary. A handle owns
 * all device memory and one HIP stream. Buffers are exchanged in the reference layout
 * (src/fluid.h:138-206): mg_upload / mg_download take host memory laid out exactly like
 * the reference's host mirrors m_Fluid.mcpu[buf_id]. Every function returns MG_OK (0) or
 * a negative MG_ERR_* code; mg_last_error() describes the last failure. Step functions
 * are asynchronous on the handle's stream; mg_download, mg_get_counts and mg_sync wait.
 * A handle is not thread-safe.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef MG_SPH_H
#define MG_SPH_H

#include "mg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mg_handle mg_handle;

/* Allocation for max_points particle slots with B bonds per particle.
 * Replaces FluidSystem::AllocateParticles / CreateFluidBuffers
 * (src/fluid_system.cpp:771-845, src/host_CL.cpp:1521-1606). */
int mg_create(mg_handle** out, int max_points, int bonds_per_particle);
/* Replaces FluidSystem::Exit (src/fluid_system.cpp:640-650) without calling exit(). */
int mg_destroy(mg_handle* h);
/* Text of the last error on this thread ("" if none). */
const char* mg_last_error(void);

/* Grid set-up. Replaces FluidSetupCL(num, gsrch, res, size, delta, gmin, gmax, total, chk)
 * (src/host_CL.cpp:39-174) and allocates the grid buffers (AllocateGrid,
 * src/fluid_system.cpp:872-928). */
int mg_fluid_setup(mg_handle* h, int num, int gsrch, const int32_t res[3], const float size[3],
                   const float delta[3], const float gmin[3], const float gmax[3], int total,
                   int chk);
/* Same, from a prepared MgGrid. */
int mg_set_grid(mg_handle* h, const MgGrid* g);
/* Parameters. Replaces FluidParamCL(ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff,
 * visc, surface_t, damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl, a_f, a_p)
 * (src/fluid_system.cpp:521-610). */
int mg_fluid_param(mg_handle* h, float ss, float sr, float pr, float mass, float rest,
                   const float bmin[3], const float bmax[3], float estiff, float istiff,
                   float visc, float surface_t, float damp, float fmin, float fmax, float ffreq,
                   float gslope, float gx, float gy, float gz, float al, float vl, float a_f,
                   float a_p);
/* Same, from a prepared MgParams (including the diff_scale / secrete_scale additions). */
int mg_set_params(mg_handle* h, const MgParams* p);
/* Host-only derivations (no device work, callable without a GPU): the MgParams that
 * FluidParamCL (src/fluid_system.cpp:521-610) derives from its argument list, exactly as
 * mg_fluid_param stores them, and the MgGrid of SetupGrid (src/fluid_system.cpp:2073-2131)
 * followed by FluidSetupCL (src/host_CL.cpp:39-106). FFI callers use these instead of
 * restating the float32 host arithmetic. */
int mg_derive_params(MgParams* out, float ss, float sr, float pr, float mass, float rest,
                     const float bmin[3], const float bmax[3], float estiff, float istiff,
                     float visc, float surface_t, float damp, float fmin, float fmax, float ffreq,
                     float gslope, float gx, float gy, float gz, float al, float vl, float a_f,
                     float a_p);
int mg_derive_grid(MgGrid* out, const float vmin[3], const float vmax[3], float sim_scale,
                   float cell_size, float smooth_radius);
/* The parameters and grid a handle holds: the reference's host copies m_FParams and
 * m_GridRes / m_GridMin / ... (src/fluid_system.h:741, 756-760). */
int mg_get_params(mg_handle* h, MgParams* out);
int mg_get_grid(mg_handle* h, MgGrid* out);
/* Replaces UpdateGenome (src/fluid_system.cpp:494-502); MgGenome has FGenome's layout. */
int mg_set_genome(mg_handle* h, const MgGenome* g);
/* Replaces setFreeze (src/fluid_system.cpp:2020-2035). */
int mg_set_freeze(mg_handle* h, int freeze);
/* Selects the SPH kernels of pressure and force for the handle's phase calls and steps
 * (MG_SPH_REFERENCE, the default: the reference's live Wendland-C2 density and (sr - r)^3
 * force; MG_SPH_POLY6_SPIKY: Mueller 2003 poly6 density, spiky pressure and viscosity
 * Laplacian force). No counterpart entry point in the reference: the constants come from
 * FluidParamCL (src/fluid_system.cpp:550, 568-569, 587) and the forms from the reference's
 * commented kernels (src/fluid_system_cuda.cu:668-672, 695, 3140-3160, 3213-3220).
 * MG_ERR_UNSUPPORTED for a handle in a batch (batches step the reference kernels). */
int mg_set_sph_kernels(mg_handle* h, int mode);
int mg_get_sph_kernels(mg_handle* h, int* mode);

/* Host -> device copy of one buffer in the reference layout; bytes must equal
 * mg_buffer_bytes(buf_id, N, C, B). Replaces TransferToCL (src/host_CL.cpp:1608-1634). */
int mg_upload(mg_handle* h, int buf_id, const void* src, size_t bytes);
/* Device -> host copy of one buffer in the reference layout. Replaces TransferFromCL
 * (src/host_CL.cpp:1636-1668). Waits for the stream. */
int mg_download(mg_handle* h, int buf_id, void* dst, size_t bytes);
/* mNumPoints and mActivePoints as the last sort left them (src/host_CL.cpp:1072-1093). */
int mg_get_counts(mg_handle* h, uint32_t* num_points, uint32_t* active_points);
/* Wait for all work queued on the handle. */
int mg_sync(mg_handle* h);

/* ---- phases (src/fluid_system.h:619-651) ---- */
/* InsertParticlesCL(gcell, gndx, gcnt) (src/host_CL.cpp:228-426); non-null pointers receive
 * FGCELL, FGNDX (N entries each) and FBIN_COUNT (C entries), as the reference's optional
 * readback does. */
int mg_insert_particles(mg_handle* h, uint32_t* gcell, uint32_t* gndx, uint32_t* gcnt);
/* PrefixSumCellsCL(zero_offsets) (src/host_CL.cpp:429-863; src/host_CUDA.cpp:307-400). */
int mg_prefix_sum_cells(mg_handle* h, int zero_offsets);
/* CountingSortFullCL(ppos) (src/host_CL.cpp:1064-1413; src/host_CUDA.cpp:498-580),
 * including the dense gene lists. */
int mg_counting_sort_full(mg_handle* h);
/* ComputePressureCL (src/host_CL.cpp:1918-1934). */
int mg_compute_pressure(mg_handle* h);
/* ComputeForceCL (src/host_CL.cpp:1950-1963). */
int mg_compute_force(mg_handle* h);
/* TransferPosVelVeval + AdvanceCL(time, dt, ss) + TransferPosVelVevalFromTemp
 * (src/host_CL.cpp:199-226, 2182-2216). */
int mg_advance(mg_handle* h, float time, float dt, float ss);
/* ComputeDiffusionCL (src/host_CL.cpp:1936-1948). */
int mg_compute_diffusion(mg_handle* h);
/* ComputeGenesCL (src/host_CL.cpp:1965-2019). */
int mg_compute_genes(mg_handle* h);
/* ZeroVelCL (src/host_CL.cpp:2175-2180). */
int mg_zero_vel(mg_handle* h);

/* No counterpart in the reference: mg_prefix_sum_cells' cell scan (the exclusive prefix sum
 * FBIN_OFFSET of FBIN_COUNT) for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). It allocates
 * nothing: cnt (n counts) and off (n offsets) are device buffers of the caller, the scan's
 * scratch is a fixed array in the library (one per GPU), and the work is queued on `stream`
 * (a hipStream_t; null = the null stream). n <= 67,108,864 (16,384 tiles of 4,096), else
 * MG_ERR_SIZE. report (device, may be null) receives {sum of the counts, nonzero if the
 * scan's bounded wait expired}. Calls that overlap on one GPU share the scratch and race. */
int mg_prefix_sum_cells_noalloc(const uint32_t* cnt, uint32_t* off, int n, uint32_t* report, void* stream);

/* InitializeBondsCL (src/host_CUDA.cpp:650-660, src/host_CL.cpp:1884-1916): the
 * initialize_bonds kernel (src/fluid_system_cuda.cu:1938-2044) over gene 1's dense list. Needs
 * the sorted state (after a physical sort or a step). */
int mg_initialize_bonds(mg_handle* h);

/* ---- remodelling: Run2Remodelling (src/fluid_system.cpp:2037-2071) is the four calls below
 * in order (its AssembleFibresCL launches nothing). They work on the slots, grid and gene lists
 * of the last sort, as the reference does between its inner loops and the next sort. ---- */
/* ComputeBondChangesCL(steps_per_InnerPhysicalLoop) (src/host_CL.cpp:2057-2086,
 * src/host_CUDA.cpp:744-761): clears the change counts and offsets, then computeBondChanges
 * (src/fluid_system_cuda.cu:1065-1330) over gene 2's dense list: strain integrators averaged
 * over steps_per_inner_loop (> 0) and cleared, rest length and modulus remodelled by the
 * genome's tanh parameters, change words cleared, and each particle's first broken or empty
 * slot counted into its cell's heal count. */
int mg_compute_bond_changes(mg_handle* h, uint32_t steps_per_inner_loop);
/* PrefixSumChangesCL(zero_offsets) (src/host_CL.cpp:865-1062) with tally_denselist_lengths
 * (src/fluid_system_cuda.cu:185-194): exclusive per-cell offsets of the nine change lists and
 * their lengths. The reference always passes 1; the offsets are exclusive for any value. */
int mg_prefix_sum_changes(mg_handle* h, int zero_offsets);
/* CountingSortChangesCL (src/host_CL.cpp:1415-1519, src/host_CUDA.cpp:604-648):
 * countingSortChanges (src/fluid_system_cuda.cu:406-618) writes the (slot, bond) pairs of each
 * change list in cell order, slot order within a cell. */
int mg_counting_sort_changes(mg_handle* h);
/* ComputeParticleChangesCL (src/host_CL.cpp:2088-2173, src/host_CUDA.cpp:763-841): heal
 * (src/fluid_system_cuda.cu:2047-2107) over list 0, its random shift drawn from (seed, particle
 * ID) instead of the reference's curand states. computeBondChanges fills no other list (its code
 * for them is commented out in the reference), so their kernels have nothing to do; a
 * non-empty one (possible only by uploading counts) is rejected with MG_ERR_UNSUPPORTED before
 * anything runs. Returns with the incoming-bond lists rebuilt. */
int mg_compute_particle_changes(mg_handle* h, uint64_t seed);

/* ---- composites (src/fluid_system.cpp:1919-2014) ---- */
int mg_physical_sort(mg_handle* h);                          /* Run2PhysicalSort      */
int mg_inner_physical_loop(mg_handle* h, float time, float dt); /* Run2InnerPhysicalLoop */
int mg_gene_action(mg_handle* h);                            /* Run2GeneAction        */
/* nsteps x [Run2InnerPhysicalLoop, Run2GeneAction if flags & MG_STEP_GENES,
 * Run2PhysicalSort] — the hot loop of Run2Simulation (src/fluid_system.cpp:2671-2684)
 * with steps_per_InnerPhysicalLoop = 1. Internal time advances by dt per step. */
int mg_step(mg_handle* h, int nsteps, float dt, unsigned flags);

/* ---- measurement (no reference counterpart; for bench.py and profiling) ----
 * Like mg_step, with a HIP event pair around every phase launch on the handle's stream.
 * ms_out[phase] accumulates milliseconds and calls_out[phase] the launches, for phase ids
 * MG_PH_*; both arrays need MG_PH_COUNT entries (either may be null). */
enum {
    MG_PH_PRESSURE = 0, MG_PH_FORCE, MG_PH_ADVANCE, MG_PH_DIFFUSION, MG_PH_GENES, MG_PH_INSERT,
    MG_PH_SCAN, MG_PH_PLACE, MG_PH_SORT, MG_PH_LISTS, MG_PH_FILL, MG_PH_LEARN, MG_PH_COUNT
};
int mg_step_timed(mg_handle* h, int nsteps, float dt, unsigned flags, double* ms_out,
                  int* calls_out);

/* Test hook (no reference counterpart): sets the cell scan's 64-bit tile ticket, so a test
 * can run scans across the point where a 32-bit ticket would wrap. The value must be a
 * multiple of the scan's tile count, ceil(C / 4096), as it always is between launches. */
int mg_debug_set_scan_ticket(mg_handle* h, uint64_t ticket);

/* ---- batched instances (no counterpart in the reference, which steps one FluidSystem) ----
 * A batch steps many independent handles (each with its own buffers, grid, parameters and
 * genome) together: by default one persistent launch per step covers every phase of every
 * instance (resident workgroups, per-phase work queues, grid barriers); the alternative
 * dispatch form runs one launch per phase and size class. Every instance evolves bit for bit
 * as under mg_step. While a handle belongs to a batch its work
 * is queued on the batch's stream (uploads, downloads and mg_step keep working); destroy the
 * batch before any of its handles. Handles must be in the sorted state (after
 * mg_physical_sort or a step) when mg_batch_step is called. */
typedef struct mg_batch mg_batch;
int mg_batch_create(mg_batch** out, mg_handle* const* handles, int n);
int mg_batch_destroy(mg_batch* b);
/* nsteps fused steps of every instance; flags as mg_step (MG_STEP_GRAPH replays two
 * captured steps at a time). */
int mg_batch_step(mg_batch* b, int nsteps, float dt, unsigned flags);
/* Wait for the batch's stream; fails if a kernel of any instance gave up a bounded wait. */
int mg_batch_sync(mg_batch* b);
/* Size classes (by particle count: < 8192, < 24576, larger): *nclass, and per class the
 * number of instances and the block size of its particle and cell launches (arrays of at
 * least *nclass entries, either may be null). */
int mg_batch_classes(mg_batch* b, int* nclass, int* members, unsigned* threads);
/* Measures `steps` steps of the dispatch form at each block size (256, 128, 64) and keeps,
 * per size class, the one with the least kernel time (threads_out, one entry per class, may be
 * null, receives it); then, at those block sizes, the pressure sweep's tile shape per class
 * (candidates tested at a time 8 or 4, staged band of 128 or 64 slots per wavefront) with the
 * least pressure time; then times `steps` whole steps of that dispatch form and of the
 * persistent form at each workgroup count per CU that stays resident, and keeps the fastest.
 * The steps are real steps of the batch. */
int mg_batch_autotune(mg_batch* b, int steps, float dt, unsigned flags, unsigned* threads_out);
/* The dispatch form's configuration of size class cls (0 .. nclass - 1): its instances, block
 * size, and the pressure sweep's candidate batch and band capacity (any output may be null). */
int mg_batch_class_config(mg_batch* b, int cls, int* instances, unsigned* threads, int* cand_batch,
                          int* band_cap);
/* Sets size class cls's pressure sweep tile shape for the dispatch form: cand_batch 8 or 4,
 * band_cap 128 or 64 (MG_ERR_ARG otherwise). Every shape gives the same bits. */
int mg_batch_set_tile_shape(mg_batch* b, int cls, int cand_batch, int band_cap);
/* The launch form in use: *persistent = 1 for one persistent launch per step (with
 * *workgroups workgroups of 256 threads), 0 for the dispatch form. Either may be null. */
int mg_batch_launch_config(mg_batch* b, int* persistent, unsigned* workgroups);
/* Chooses the form: on = 1 persistent (workgroups_per_cu workgroups on each CU, 0 = all that
 * stay resident; batches whose instances exceed 65,535 blocks of 256 use the dispatch form),
 * on = 0 the dispatch form. */
int mg_batch_set_persistent(mg_batch* b, int on, unsigned workgroups_per_cu);
/* Like mg_batch_step without graph replay, with HIP events around every launch group:
 * ms_out[phase] accumulates milliseconds summed over the size classes and calls_out[phase]
 * the launch groups, phase ids MG_PH_* (MG_PH_COUNT entries each; either may be null). */
int mg_batch_step_timed(mg_batch* b, int nsteps, float dt, unsigned flags, double* ms_out,
                        int* calls_out);

/* ---- learned coarse step (no counterpart in the reference) ----
 * An MLP (53 inputs: 8 bonded neighbours' relative position and velocity, density, pressure
 * and the coarse step's acceleration; two hidden layers of 128 SiLU units; 6 outputs)
 * predicts for each particle two velocity corrections for one coarse step, each the value
 * four fine steps of dt / 4 reach minus the plain coarse step's: the drift c_d (the position
 * moves by (v_next + c_d) dt) and veval c_e ((v + v_next) / 2 + c_e is stored); the velocity
 * stays v_next. mg_step_learned runs the step's pressure and force passes, the network, then
 * the advance. BF16 weights, FP32 accumulation on MFMA. Parameter vector (fp32,
 * mg_net_nparams() floats, row-major [in][out]): W1[64][128] b1[128] W2[128][128] b2[128]
 * W3[128][16] b3[16]; inputs 53..63 and outputs 6..15 are zero padding. */
typedef struct mg_net mg_net;
size_t mg_net_nparams(void);
/* Xavier-uniform weights from a seed, zero biases, identity normalisation. */
int mg_net_create(mg_net** out, uint64_t seed);
int mg_net_destroy(mg_net* n);
int mg_net_get_params(mg_net* n, float* out, size_t count);
int mg_net_set_params(mg_net* n, const float* in, size_t count);
/* Input i is used as (x_i - mu[i]) * inv_sigma[i] (64 entries each); output c is
 * y_c * out_scale[c] (6 entries). A particle whose normalised input exceeds in_limit, whose
 * |y_c| exceeds out_limit, or whose drift correction would move it more than 1e-3 (world
 * units) in the step takes the plain step (counted by mg_learned_fallbacks). */
int mg_net_set_norm(mg_net* n, const float* mu, const float* inv_sigma, const float* out_scale,
                    float in_limit, float out_limit);
/* Training set: X rows x 64 raw inputs, T rows x 6 targets (in output units, i.e. already
 * divided by out_scale), mask rows (0 = ignore the row). */
int mg_net_load_data(mg_net* n, const float* X, const float* T, const float* mask, size_t rows);
/* `steps` Adam updates (lr, beta 0.9 / 0.999) on minibatches of `batch` consecutive rows of
 * the training set, mean-squared error; loss_out (steps entries, may be null). */
int mg_net_train(mg_net* n, int steps, size_t batch, float lr, uint64_t seed, float* loss_out);
/* One forward and backward pass over training rows [row0, row0 + rows): gradients
 * (mg_net_nparams() floats) and the mean loss; no update. */
int mg_net_grad(mg_net* n, size_t row0, size_t rows, float* grads_out, float* loss_out);
/* Network outputs y (6 per row, before out_scale) of training rows [row0, row0 + rows). */
int mg_net_forward(mg_net* n, size_t row0, size_t rows, float* y_out);
/* Raw inputs of every slot of the handle's current sorted state (N x 64; zero rows for
 * inactive slots), after the pressure and force passes a learned step runs first (the force
 * pass changes no bond record; the state stays as a step finds it). Needs the sorted state. */
int mg_net_features(mg_handle* h, float* X_out);
/* The corrections a learned step of length dt would apply to the current sorted state:
 * N x 8 floats (c_d, used, c_e, 0). */
int mg_net_infer(mg_handle* h, mg_net* n, float dt, float* DV_out);
/* nsteps learned steps of dt (the coarse step): mg_step with each active particle's
 * corrections applied in its advance. Needs the sorted state. */
int mg_step_learned(mg_handle* h, mg_net* n, int nsteps, float dt, unsigned flags);
/* Particles that took the plain step inside learned steps since the last call. */
int mg_learned_fallbacks(mg_handle* h, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif
