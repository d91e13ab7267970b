/*
 * fluid_system.cpp — C++ FluidSystem facade over the C-ABI (see fluid_system.hpp).
 *
 * Also exports mgfs_* C entry points so the facade can be exercised from any FFI.
 */
#include "fluid_system.hpp"

#include <cstring>

namespace mg {

namespace {
const int kParticleBufs[] = {MG_FPOS,      MG_FVEL,         MG_FVEVAL,      MG_FFORCE,
                             MG_FPRESS,    MG_FDENSITY,     MG_FAGE,        MG_FCOLOR,
                             MG_FGCELL,    MG_FGNDX,        MG_FELASTIDX,   MG_FNERVEIDX,
                             MG_FCONC,     MG_FEPIGEN,      MG_FBIN,        MG_FPARTICLEIDX,
                             MG_FPARTICLE_ID, MG_FMASS_RADIUS};
}

FluidSystem::FluidSystem(int max_points, int bonds_per_particle)
    : mMaxPoints(max_points), mBonds(bonds_per_particle)
{
    check(mg_create(&h_, max_points, bonds_per_particle), "mg_create");
    mcpu_.resize(MG_MAX_BUF);
    for (int id : kParticleBufs) mcpu_[id].assign(bufBytes(id), 0);
}

FluidSystem::~FluidSystem()
{
    if (h_) mg_destroy(h_);
}

void FluidSystem::check(int r, const char* what)
{
    if (r != MG_OK) throw Error(r, std::string(what) + ": " + mg_last_error());
}

size_t FluidSystem::bufBytes(int id) const
{
    return mg_buffer_bytes(id, mMaxPoints, mGridTotal, mBonds);
}

char* FluidSystem::bufC(int id)
{
    if (id < 0 || id >= MG_MAX_BUF) throw Error(MG_ERR_ARG, "bufC: bad buffer id");
    if (!mg_buffer_known(id)) throw Error(MG_ERR_UNSUPPORTED, "bufC: buffer not used by the step");
    size_t n = bufBytes(id);
    if (mcpu_[id].size() != n) mcpu_[id].assign(n, 0);
    return mcpu_[id].data();
}

void FluidSystem::SetupGrid(const float vmin[3], const float vmax[3], float sim_scale,
                            float cell_size, float smooth_radius)
{
    MgGrid g;
    check(mg_grid_setup(&g, vmin, vmax, sim_scale, cell_size, smooth_radius), "SetupGrid");
    check(mg_set_grid(h_, &g), "mg_set_grid");
    mGridTotal = g.total;
}

void FluidSystem::FluidSetupCL(int num, int gsrch, const int32_t res[3], const float size[3],
                               const float delta[3], const float gmin[3], const float gmax[3],
                               int total, int chk)
{
    check(mg_fluid_setup(h_, num, gsrch, res, size, delta, gmin, gmax, total, chk), "FluidSetupCL");
    mGridTotal = total;
}

void FluidSystem::FluidParamCL(float ss, float sr, float pr, float mass, float rest,
                               const float bmin[3], const float bmax[3], float estiff,
                               float istiff, float visc, float surface_tension, float damp,
                               float fmin, float fmax, float ffreq, float gslope, float gx,
                               float gy, float gz, float al, float vl, float a_f, float a_p)
{
    check(mg_fluid_param(h_, ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff, visc,
                         surface_tension, damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl,
                         a_f, a_p),
          "FluidParamCL");
}

void FluidSystem::SetParams(const MgParams& p) { check(mg_set_params(h_, &p), "SetParams"); }
void FluidSystem::UpdateGenome() { check(mg_set_genome(h_, &m_FGenome), "UpdateGenome"); }
void FluidSystem::setFreeze(bool freeze) { check(mg_set_freeze(h_, freeze ? 1 : 0), "setFreeze"); }

void FluidSystem::SetSPHKernels(int mode) { check(mg_set_sph_kernels(h_, mode), "SetSPHKernels"); }

void FluidSystem::TransferToCL(int id)
{
    check(mg_upload(h_, id, bufC(id), bufBytes(id)), "TransferToCL");
}

void FluidSystem::TransferFromCL(int id)
{
    check(mg_download(h_, id, bufC(id), bufBytes(id)), "TransferFromCL");
}

void FluidSystem::TransferToCL()
{
    for (int id : kParticleBufs) TransferToCL(id);
}

void FluidSystem::TransferFromCL()
{
    for (int id : kParticleBufs) TransferFromCL(id);
}

void FluidSystem::InsertParticlesCL(uint32_t* gcell, uint32_t* gndx, uint32_t* gcnt)
{
    check(mg_insert_particles(h_, gcell, gndx, gcnt), "InsertParticlesCL");
}
void FluidSystem::PrefixSumCellsCL(int zero_offsets)
{
    check(mg_prefix_sum_cells(h_, zero_offsets), "PrefixSumCellsCL");
}
void FluidSystem::CountingSortFullCL(float*) { check(mg_counting_sort_full(h_), "CountingSortFullCL"); }
void FluidSystem::ComputePressureCL() { check(mg_compute_pressure(h_), "ComputePressureCL"); }
void FluidSystem::ComputeForceCL() { check(mg_compute_force(h_), "ComputeForceCL"); }
void FluidSystem::ComputeDiffusionCL() { check(mg_compute_diffusion(h_), "ComputeDiffusionCL"); }
void FluidSystem::ComputeGenesCL() { check(mg_compute_genes(h_), "ComputeGenesCL"); }
void FluidSystem::AdvanceCL(float time, float dt, float ss)
{
    check(mg_advance(h_, time, dt, ss), "AdvanceCL");
}
void FluidSystem::ZeroVelCL() { check(mg_zero_vel(h_), "ZeroVelCL"); }
void FluidSystem::InitializeBondsCL() { check(mg_initialize_bonds(h_), "InitializeBondsCL"); }
void FluidSystem::ComputeBondChangesCL(uint32_t steps)
{
    check(mg_compute_bond_changes(h_, steps), "ComputeBondChangesCL");
}
void FluidSystem::PrefixSumChangesCL(int zero_offsets)
{
    check(mg_prefix_sum_changes(h_, zero_offsets), "PrefixSumChangesCL");
}
void FluidSystem::CountingSortChangesCL() { check(mg_counting_sort_changes(h_), "CountingSortChangesCL"); }
void FluidSystem::ComputeParticleChangesCL(uint64_t seed)
{
    check(mg_compute_particle_changes(h_, seed), "ComputeParticleChangesCL");
}
void FluidSystem::Run2Remodelling(uint32_t steps)
{
    ComputeBondChangesCL(steps);
    PrefixSumChangesCL(1);
    CountingSortChangesCL();
    ComputeParticleChangesCL(heal_calls_++);
}

void FluidSystem::Run2PhysicalSort() { check(mg_physical_sort(h_), "Run2PhysicalSort"); }
void FluidSystem::Run2InnerPhysicalLoop()
{
    check(mg_inner_physical_loop(h_, (float)m_Time, (float)m_DT), "Run2InnerPhysicalLoop");
    AdvanceTime();
}
void FluidSystem::Run2GeneAction() { check(mg_gene_action(h_), "Run2GeneAction"); }

void FluidSystem::RunSteps(int nsteps, bool gene_activity, bool graph)
{
    unsigned flags = (gene_activity ? MG_STEP_GENES : 0u) | (graph ? MG_STEP_GRAPH : 0u);
    check(mg_step(h_, nsteps, (float)m_DT, flags), "RunSteps");
    m_Time += nsteps * m_DT;
}

int FluidSystem::NumPoints()
{
    uint32_t n = 0;
    check(mg_get_counts(h_, &n, nullptr), "NumPoints");
    return (int)n;
}

int FluidSystem::ActivePoints()
{
    uint32_t a = 0;
    check(mg_get_counts(h_, nullptr, &a), "ActivePoints");
    return (int)a;
}

} // namespace mg

/* ---- C entry points for the facade (FFI callers and tests) ---- */
extern "C" {

struct mgfs_handle;

int mgfs_create(mgfs_handle** out, int max_points, int bonds_per_particle)
{
    try {
        *out = reinterpret_cast<mgfs_handle*>(new mg::FluidSystem(max_points, bonds_per_particle));
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

int mgfs_destroy(mgfs_handle* f)
{
    delete reinterpret_cast<mg::FluidSystem*>(f);
    return MG_OK;
}

/* copy `bytes` into the facade's host mirror of buffer `id` (m_Fluid.mcpu[id]) */
int mgfs_set_host(mgfs_handle* f, int id, const void* src, size_t bytes)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        if (bytes != fs->bufBytes(id)) return MG_ERR_SIZE;
        std::memcpy(fs->bufC(id), src, bytes);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

int mgfs_get_host(mgfs_handle* f, int id, void* dst, size_t bytes)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        if (bytes != fs->bufBytes(id)) return MG_ERR_SIZE;
        std::memcpy(dst, fs->bufC(id), bytes);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

int mgfs_setup(mgfs_handle* f, const MgGrid* g, const MgParams* p, const MgGenome* genome, double dt)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        fs->FluidSetupCL(fs->MaxPoints(), g->srch, g->res, g->size, g->delta, g->gmin, g->gmax,
                         g->total, g->chk);
        fs->SetParams(*p);
        fs->SetGenome(*genome);
        fs->UpdateGenome();
        fs->SetDT(dt);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

/* the facade's SetupGrid (src/fluid_system.cpp:2073-2131, then FluidSetupCL) */
int mgfs_setup_grid(mgfs_handle* f, const float vmin[3], const float vmax[3], float sim_scale,
                    float cell_size, float smooth_radius)
{
    try {
        reinterpret_cast<mg::FluidSystem*>(f)->SetupGrid(vmin, vmax, sim_scale, cell_size, smooth_radius);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

/* the facade's FluidParamCL (src/fluid_system.cpp:521-610) */
int mgfs_fluid_param(mgfs_handle* f, float ss, float sr, float pr, float mass, float rest,
                     const float bmin[3], const float bmax[3], float estiff, float istiff,
                     float visc, float surface_t, float damp, float fmin, float fmax, float ffreq,
                     float gslope, float gx, float gy, float gz, float al, float vl, float a_f,
                     float a_p)
{
    try {
        reinterpret_cast<mg::FluidSystem*>(f)->FluidParamCL(ss, sr, pr, mass, rest, bmin, bmax,
                                                            estiff, istiff, visc, surface_t, damp,
                                                            fmin, fmax, ffreq, gslope, gx, gy, gz,
                                                            al, vl, a_f, a_p);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

/* the genome and time step without touching the grid or parameters */
int mgfs_set_genome_dt(mgfs_handle* f, const MgGenome* genome, double dt)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        fs->SetGenome(*genome);
        fs->UpdateGenome();
        fs->SetDT(dt);
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

/* the C-ABI handle the facade drives (for mg_get_params / mg_get_grid / mg_download) */
mg_handle* mgfs_handle_of(mgfs_handle* f) { return reinterpret_cast<mg::FluidSystem*>(f)->handle(); }

int mgfs_transfer_to(mgfs_handle* f)
{
    try { reinterpret_cast<mg::FluidSystem*>(f)->TransferToCL(); return MG_OK; }
    catch (const mg::Error& e) { return e.code; }
}

int mgfs_transfer_from(mgfs_handle* f)
{
    try { reinterpret_cast<mg::FluidSystem*>(f)->TransferFromCL(); return MG_OK; }
    catch (const mg::Error& e) { return e.code; }
}

/* Run2Simulation's hot loop: initial Run2PhysicalSort, then nsteps x
 * [Run2InnerPhysicalLoop, Run2GeneAction (if genes), Run2PhysicalSort] */
int mgfs_run(mgfs_handle* f, int nsteps, int genes)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        fs->Run2PhysicalSort();
        for (int s = 0; s < nsteps; s++) {
            fs->Run2InnerPhysicalLoop();
            if (genes) fs->Run2GeneAction();
            fs->Run2PhysicalSort();
        }
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

/* Run2Simulation's loop with remodelling (src/fluid_system.cpp:2671-2684): a sort, then
 * `cycles` x [steps x Run2InnerPhysicalLoop, Run2GeneAction (genes), Run2Remodelling(steps),
 * Run2PhysicalSort]. */
int mgfs_run_remodel(mgfs_handle* f, int cycles, int steps, int genes)
{
    auto* fs = reinterpret_cast<mg::FluidSystem*>(f);
    try {
        fs->Run2PhysicalSort();
        for (int c = 0; c < cycles; c++) {
            for (int s = 0; s < steps; s++) fs->Run2InnerPhysicalLoop();
            if (genes) fs->Run2GeneAction();
            fs->Run2Remodelling((uint32_t)steps);
            fs->Run2PhysicalSort();
        }
        return MG_OK;
    } catch (const mg::Error& e) {
        return e.code;
    }
}

} // extern "C"
