/*
 * fluid_system.hpp — C++ facade with the reference's FluidSystem step API over the C-ABI
 * of include/mg_sph.h.
 *
 * Method names and argument meanings follow src/fluid_system.h:577-651. Host mirrors
 * (the reference's m_Fluid.mcpu[]) hold every buffer in the reference layout; TransferToCL
 * / TransferFromCL move them to and from the device. Errors throw mg::Error instead of
 * calling exit() (src/fluid_system.cpp:447-452).
 */
#ifndef MG_FLUID_SYSTEM_HPP
#define MG_FLUID_SYSTEM_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mg_sph.h"

namespace mg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

class FluidSystem {
public:
    FluidSystem(int max_points, int bonds_per_particle);
    ~FluidSystem();
    FluidSystem(const FluidSystem&) = delete;
    FluidSystem& operator=(const FluidSystem&) = delete;

    /* set-up: SetupGrid (src/fluid_system.cpp:2073-2131), FluidSetupCL (src/host_CL.cpp:39-174),
     * FluidParamCL (src/fluid_system.cpp:521-610), SetGenome/UpdateGenome, setFreeze */
    void SetupGrid(const float vmin[3], const float vmax[3], float sim_scale, float cell_size,
                   float smooth_radius);
    void FluidSetupCL(int num, int gsrch, const int32_t res[3], const float size[3],
                      const float delta[3], const float gmin[3], const float gmax[3], int total,
                      int chk);
    void FluidParamCL(float ss, float sr, float pr, float mass, float rest, const float bmin[3],
                      const float bmax[3], float estiff, float istiff, float visc,
                      float surface_tension, float damp, float fmin, float fmax, float ffreq,
                      float gslope, float gx, float gy, float gz, float al, float vl, float a_f,
                      float a_p);
    void SetParams(const MgParams& p);
    void SetGenome(const MgGenome& g) { m_FGenome = g; }
    MgGenome GetGenome() const { return m_FGenome; }
    void UpdateGenome();
    void setFreeze(bool freeze);
    /* SPH kernels of pressure and force: MG_SPH_REFERENCE (default) or MG_SPH_POLY6_SPIKY
     * (Mueller 2003 poly6 / spiky / viscosity Laplacian; mg_set_sph_kernels) */
    void SetSPHKernels(int mode);

    /* host mirrors in the reference layout (bufC/bufF/bufI of src/fluid.h:244-253) */
    char* bufC(int id);
    float* bufF(int id) { return reinterpret_cast<float*>(bufC(id)); }
    uint32_t* bufI(int id) { return reinterpret_cast<uint32_t*>(bufC(id)); }
    size_t bufBytes(int id) const;

    /* src/host_CL.cpp:1608-1668 */
    void TransferToCL();
    void TransferFromCL();
    void TransferToCL(int id);
    void TransferFromCL(int id);

    /* step phases, src/fluid_system.h:619-651 */
    void InsertParticlesCL(uint32_t* gcell, uint32_t* gndx, uint32_t* gcnt);
    void PrefixSumCellsCL(int zero_offsets);
    void CountingSortFullCL(float* ppos);
    void ComputePressureCL();
    void ComputeForceCL();
    void ComputeDiffusionCL();
    void ComputeGenesCL();
    void TransferPosVelVeval() {}          /* folded into AdvanceCL */
    void AdvanceCL(float time, float dt, float ss);
    void TransferPosVelVevalFromTemp() {}  /* folded into AdvanceCL */
    void ZeroVelCL();
    void InitializeBondsCL(); /* src/host_CUDA.cpp:650-660 */
    /* remodelling (src/host_CL.cpp:865-1062, 1415-1519, 2057-2173) */
    void ComputeBondChangesCL(uint32_t steps_per_InnerPhysicalLoop);
    void PrefixSumChangesCL(int zero_offsets);
    void CountingSortChangesCL();
    void ComputeParticleChangesCL(uint64_t seed); /* heal's random shifts from (seed, ID) */
    /* src/fluid_system.cpp:2037-2071; seed = the number of earlier calls */
    void Run2Remodelling(uint32_t steps_per_InnerPhysicalLoop);
    void AdvanceTime() { m_Time += m_DT; }

    /* src/fluid_system.cpp:1919-2014 */
    void Run2PhysicalSort();
    void Run2InnerPhysicalLoop();
    void Run2GeneAction();
    /* the hot loop of Run2Simulation (src/fluid_system.cpp:2671-2684) with
     * steps_per_InnerPhysicalLoop = 1: nsteps x [inner loop, gene action?, sort] */
    void RunSteps(int nsteps, bool gene_activity, bool graph = false);

    int NumPoints();
    int ActivePoints();
    int MaxPoints() const { return mMaxPoints; }
    int BondsPerParticle() const { return mBonds; }
    double GetDT() const { return m_DT; }
    void SetDT(double dt) { m_DT = dt; }
    double GetTime() const { return m_Time; }
    mg_handle* handle() { return h_; }

private:
    void check(int r, const char* what);
    mg_handle* h_ = nullptr;
    uint64_t heal_calls_ = 0; /* Run2Remodelling's heal seed */
    int mMaxPoints = 0, mBonds = 0, mGridTotal = 0;
    double m_Time = 0.0, m_DT = 0.003;
    MgGenome m_FGenome{};
    std::vector<std::vector<char>> mcpu_;
};

} // namespace mg

#endif
