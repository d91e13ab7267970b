/*
 * mg_oracle.cpp — CPU oracle for the Morphogenesis per-timestep SPH step.
 *
 * TEST INFRASTRUCTURE: the checker for tests/ and the CPU baseline for bench.py.
 * Nothing in the product path (morphogenesis_amd/) links or calls this file.
 *
 * A plain restatement of the reference step, templated on the arithmetic type
 * (float = the FP32 parity oracle, double = the FP64 drift oracle). It works on
 * buffers in the reference layout (src/fluid.h:138-206) and follows the reference
 * kernels expression by expression; each function cites the lines it restates.
 *
 * Deterministic choices (the reference is order-nondeterministic in these places):
 *  - insert runs particles in ascending index order, so the atomicAdd slot FGNDX of
 *    src/fluid_system_cuda.cu:81 becomes the stable rank of the particle in its cell;
 *  - the bond-force scatter (atomicAdd into FFORCE[j], src/fluid_system_cuda.cu:3311-3313,
 *    3354-3356) is applied as an ordered sum: FFORCE[i] += own total, then += each
 *    incoming bond's term in (owner FPARTICLE_ID, owner home index, slot) order, where
 *    "home index" is the slot a particle occupied when the oracle was created and only
 *    breaks ties between equal IDs. With unique IDs the order is a function of the state
 *    buffers alone, so a state downloaded and uploaded again continues bit for bit;
 *  - diffusion is a Jacobi update (reads the pre-step concentrations) instead of the
 *    in-place race of src/fluid_system_cuda.cu:3102-3104;
 *  - gene action uses float FCONC in AoS layout and applies the tally once per living
 *    particle (see compute_genes for the reference's type punning and repeated tally).
 *
 * Expression-level choices that pin bits: pow(x,3) and pow(x,2) of the force kernel are
 * x*x*x and x*x; float3/float is multiplication by the reciprocal (src/cutil_math.h:389-393).
 * Built with -ffp-contract=off (no FMA contraction) so FP32 results are reproducible.
 */
#include "mg_oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename R> struct V3 {
    R x, y, z;
};
template <typename R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return mk(a.x * s, a.y * s, a.z * s); }
/* src/cutil_math.h:389-393: division by a scalar multiplies by its reciprocal */
template <typename R> inline V3<R> divs(V3<R> a, R s) { R inv = R(1) / s; return a * inv; }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline bool has_nan(V3<R> a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

/* acos in float by the Cephes single-precision method (S. L. Moshier, Cephes Math Library,
 * asinf.c / acosf.c), the operations of mg_acosf in morphogenesis_amd/csrc/mg_device.h one
 * for one: |x| <= 0.5 -> pi/2 - asin(x); |x| > 0.5 -> 2 asin(sqrt((1 - |x|) / 2)), mirrored for
 * x < 0; asin(a), a <= 0.5, by the published odd polynomial. NaN outside [-1, 1].
 * Used for the bond-angle score of initialize_bonds (the reference calls CUDA's acos there). */
inline float asinf_half(float a)
{
    if (a < 1.0e-4f) return a;
    const float z = a * a;
    float p = 4.2163199048e-2f;
    p = p * z + 2.4181311049e-2f;
    p = p * z + 4.5470025998e-2f;
    p = p * z + 7.4953002686e-2f;
    p = p * z + 1.6666752422e-1f;
    return p * z * a + a;
}
inline float acos_f(float x)
{
    const float pio2 = 1.5707963267948966f, pi = 3.141592653589793f;
    if (!(x >= -1.0f && x <= 1.0f)) return std::nanf("");
    if (x < -0.5f) return pi - 2.0f * asinf_half(std::sqrt(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asinf_half(std::sqrt(0.5f * (1.0f - x)));
    return x < 0.0f ? pio2 + asinf_half(-x) : pio2 - asinf_half(x);
}
inline double acos_r(double x) { return std::acos(x); }
inline float acos_r(float x) { return acos_f(x); }

/* exp and tanh in float by the Cephes single-precision method (S. L. Moshier, Cephes Math
 * Library, expf.c / tanhf.c), the operations of mg_expf / mg_tanhf in
 * morphogenesis_amd/csrc/mg_device.h one for one: exp(x) = 2^n exp(r) with n = floor(x log2(e)
 * + 1/2), ln2 split in two parts and the published degree-5 polynomial; tanh(x) =
 * sign(x) (1 - 2 / (exp(2|x|) + 1)) for |x| >= 0.625, +-1 above 44.36, the published odd
 * polynomial below; evaluated on |x| with the sign copied back. Used by the bond remodelling
 * (the reference calls CUDA's tanh there). */
inline float exp_f(float x)
{
    if (x > 88.72283905206835f) return HUGE_VALF;
    if (x < -103.278929903431851103f) return 0.0f;
    const float z = std::floor(1.44269504088896341f * x + 0.5f);
    x = x - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    const int n = (int)z;
    const float x2 = x * x;
    float p = 1.9875691500e-4f;
    p = p * x + 1.3981999507e-3f;
    p = p * x + 8.3334519073e-3f;
    p = p * x + 4.1665795894e-2f;
    p = p * x + 1.6666665459e-1f;
    p = p * x + 5.0000001201e-1f;
    return std::ldexp(p * x2 + x + 1.0f, n);
}
inline float tanh_f(float x)
{
    const float a = std::fabs(x);
    float r;
    if (a > 44.361419526034175f) {
        r = 1.0f;
    } else if (a >= 0.625f) {
        r = 1.0f - 2.0f / (exp_f(a + a) + 1.0f);
    } else {
        const float s = a * a;
        float p = -5.70498872745e-3f;
        p = p * s + 2.06390887954e-2f;
        p = p * s + -5.37397155531e-2f;
        p = p * s + 1.33314422036e-1f;
        p = p * s + -3.33332819422e-1f;
        r = p * s * a + a;
    }
    return std::copysign(r, x);
}
inline double tanh_r(double x) { return std::tanh(x); }
inline float tanh_r(float x) { return tanh_f(x); }

/* heal's random word: the SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014) of
 * seed + golden * (particle ID + 1), low 32 bits; mg_heal_word in mg_device.h. */
inline uint32_t heal_word(uint64_t seed, uint32_t pid)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)pid + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

/* float -> int truncation; out-of-range and NaN map to INT_MIN, which fails every
 * acceptance test exactly as the hardware conversions do. */
template <typename R> inline int trunc_i(R v)
{
    if (!(v > R(-2147483648.0) && v < R(2147483647.0))) return INT_MIN;
    return (int)v;
}

inline bool is_float_slot(int k)
{
    return k == MG_B_LIMIT || k == MG_B_REST || k == MG_B_MODULUS || k == MG_B_DAMP ||
           k == MG_B_STRAIN2 || k == MG_B_STRAIN;
}

struct Base {
    virtual ~Base() {}
    virtual int upload(int buf, const void* src, size_t bytes) = 0;
    virtual int download(int buf, void* dst, size_t bytes) = 0;
    virtual void insert() = 0;
    virtual void prefix_sum() = 0;
    virtual void sort() = 0;
    virtual void pressure() = 0;
    virtual void force() = 0;
    virtual void advance(float time, float dt, float ss) = 0;
    virtual void diffusion() = 0;
    virtual void genes() = 0;
    virtual void zero_vel() = 0;
    virtual void initialize_bonds() = 0;
    virtual void compute_bond_changes(uint32_t steps) = 0;
    virtual void prefix_sum_changes() = 0;
    virtual void counting_sort_changes() = 0;
    virtual void heal(uint64_t seed) = 0;
    uint32_t clen[MG_NUM_CHANGES] = {};
    virtual void advance_special(float time, float dt, float ss, float zoff) = 0;
    MgParams p;
    MgGrid g;
    MgGenome gen;
    bool have_params = false, have_grid = false;
    uint32_t num_points = 0, active_points = 0;
    int nthreads = 0;
    double time = 0.0;
    int sph = MG_SPH_REFERENCE; /* SPH kernels of pressure and force */
};

template <typename R> struct Oracle : Base {
    static constexpr bool F32 = std::is_same<R, float>::value;
    int N, B;
    int C = 0;
    std::vector<R> pos, vel, veval, frc;    /* [N][3] */
    std::vector<R> press, dens;              /* [N]    */
    std::vector<uint32_t> age, color, gcell, gndx, nerve, bin, pid, mrad; /* [N] */
    std::vector<uint32_t> bu;                /* [N][B][9] raw record words           */
    std::vector<double> bd;                  /* [N][B][9] float slots (FP64 oracle)   */
    std::vector<uint32_t> pidx;              /* [N][B][2]                              */
    std::vector<R> conc;                     /* [N][16]                                */
    std::vector<uint32_t> epigen;            /* [16][N]                                */
    std::vector<uint32_t> home;              /* [N] creation slot of the occupant      */
    std::vector<uint32_t> cnt, off, gcnt, goff, lists;
    uint32_t lens[MG_NUM_GENES];
    /* change lists: counts / offsets [9][C], pairs [9][2][N], heal slot per sorted slot */
    std::vector<uint32_t> ccnt, coff, clists, hs;

    Oracle(int n, int b) : N(n), B(b)
    {
        pos.assign((size_t)N * 3, R(0)); vel = pos; veval = pos; frc = pos;
        press.assign(N, R(0)); dens.assign(N, R(0));
        age.assign(N, 0); color = age; gcell.assign(N, MG_GRID_UNDEF); gndx.assign(N, 0);
        nerve = age; bin = age; pid.assign(N, MG_NONE); mrad = age;
        bu.assign((size_t)N * B * MG_DATA_PER_BOND, MG_NONE);
        if (!F32) bd.assign((size_t)N * B * MG_DATA_PER_BOND, 0.0);
        pidx.assign((size_t)N * B * 2, MG_NONE);
        conc.assign((size_t)N * MG_NUM_TF, R(0));
        epigen.assign((size_t)N * MG_NUM_GENES, 0);
        home.resize(N);
        for (int i = 0; i < N; i++) home[i] = (uint32_t)i;
        lists.assign((size_t)MG_NUM_GENES * N, MG_NONE);
        std::memset(lens, 0, sizeof(lens));
        clists.assign((size_t)MG_NUM_CHANGES * 2 * N, MG_NONE);
        hs.assign(N, MG_NONE);
        num_points = active_points = (uint32_t)N;
    }

    void set_grid_arrays()
    {
        C = g.total;
        cnt.assign(C, 0); off.assign(C, 0);
        gcnt.assign((size_t)MG_NUM_GENES * C, 0); goff = gcnt;
        ccnt.assign((size_t)MG_NUM_CHANGES * C, 0); coff = ccnt;
    }

    /* ---- bond record float-slot access ---- */
    inline R bf(size_t k) const
    {
        if constexpr (F32) return u2f(bu[k]);
        else return (R)bd[k];
    }
    inline void sbf(size_t k, R v)
    {
        if constexpr (F32) bu[k] = f2u(v);
        else bd[k] = (double)v;
    }
    inline size_t brec(size_t i, int a) const { return (i * B + a) * MG_DATA_PER_BOND; }
    inline V3<R> P(size_t i) const { return mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); }
    inline V3<R> VE(size_t i) const { return mk(veval[3 * i], veval[3 * i + 1], veval[3 * i + 2]); }

    /* ---------------------------------------------------------------- I/O */
    int upload(int buf, const void* src, size_t bytes) override
    {
        if (!mg_buffer_known(buf)) return MG_ERR_UNSUPPORTED;
        size_t need = mg_buffer_bytes(buf, N, C, B);
        if (bytes != need) return MG_ERR_SIZE;
        if (need == 0) return MG_OK;
        const float* f = (const float*)src;
        const uint32_t* u = (const uint32_t*)src;
        auto v4 = [&](std::vector<R>& dst) {
            for (int i = 0; i < N; i++)
                for (int k = 0; k < 3; k++) dst[3 * (size_t)i + k] = (R)f[4 * (size_t)i + k];
        };
        switch (buf) {
        case MG_FPOS: v4(pos); break;
        case MG_FVEL: v4(vel); break;
        case MG_FVEVAL: v4(veval); break;
        case MG_FFORCE: v4(frc); break;
        case MG_FPRESS: for (int i = 0; i < N; i++) press[i] = (R)f[i]; break;
        case MG_FDENSITY: for (int i = 0; i < N; i++) dens[i] = (R)f[i]; break;
        case MG_FAGE: std::memcpy(age.data(), u, bytes); break;
        case MG_FCOLOR: std::memcpy(color.data(), u, bytes); break;
        case MG_FGCELL: std::memcpy(gcell.data(), u, bytes); break;
        case MG_FGNDX: std::memcpy(gndx.data(), u, bytes); break;
        case MG_FNERVEIDX: std::memcpy(nerve.data(), u, bytes); break;
        case MG_FBIN: std::memcpy(bin.data(), u, bytes); break;
        case MG_FPARTICLE_ID: std::memcpy(pid.data(), u, bytes); break;
        case MG_FMASS_RADIUS: std::memcpy(mrad.data(), u, bytes); break;
        case MG_FELASTIDX:
            std::memcpy(bu.data(), u, bytes);
            if (!F32)
                for (size_t k = 0; k < bu.size(); k++)
                    if (is_float_slot((int)(k % MG_DATA_PER_BOND))) bd[k] = (double)u2f(bu[k]);
            break;
        case MG_FPARTICLEIDX: std::memcpy(pidx.data(), u, bytes); break;
        case MG_FCONC: for (size_t k = 0; k < conc.size(); k++) conc[k] = (R)f[k]; break;
        case MG_FEPIGEN: std::memcpy(epigen.data(), u, bytes); break;
        case MG_FBIN_COUNT: std::memcpy(cnt.data(), u, bytes); break;
        case MG_FBIN_OFFSET: std::memcpy(off.data(), u, bytes); break;
        case MG_FBIN_COUNT_ACTIVE_GENES: std::memcpy(gcnt.data(), u, bytes); break;
        case MG_FBIN_OFFSET_ACTIVE_GENES: std::memcpy(goff.data(), u, bytes); break;
        case MG_FDENSE_LIST_LENGTHS: std::memcpy(lens, u, bytes); break;
        case MG_FDENSE_LISTS: std::memcpy(lists.data(), u, bytes); break;
        case MG_FBIN_COUNT_CHANGES: std::memcpy(ccnt.data(), u, bytes); break;
        case MG_FBIN_OFFSET_CHANGES: std::memcpy(coff.data(), u, bytes); break;
        case MG_FDENSE_LIST_LENGTHS_CHANGES: std::memcpy(clen, u, bytes); break;
        case MG_FDENSE_LISTS_CHANGES: std::memcpy(clists.data(), u, bytes); break;
        default: return MG_ERR_UNSUPPORTED;
        }
        return MG_OK;
    }

    int download(int buf, void* dst, size_t bytes) override
    {
        if (!mg_buffer_known(buf)) return MG_ERR_UNSUPPORTED;
        size_t need = mg_buffer_bytes(buf, N, C, B);
        if (bytes != need) return MG_ERR_SIZE;
        if (need == 0) return MG_OK;
        float* f = (float*)dst;
        uint32_t* u = (uint32_t*)dst;
        auto v4 = [&](const std::vector<R>& s) {
            for (int i = 0; i < N; i++) {
                for (int k = 0; k < 3; k++) f[4 * (size_t)i + k] = (float)s[3 * (size_t)i + k];
                f[4 * (size_t)i + 3] = 0.0f;
            }
        };
        switch (buf) {
        case MG_FPOS: v4(pos); break;
        case MG_FVEL: v4(vel); break;
        case MG_FVEVAL: v4(veval); break;
        case MG_FFORCE: v4(frc); break;
        case MG_FPRESS: for (int i = 0; i < N; i++) f[i] = (float)press[i]; break;
        case MG_FDENSITY: for (int i = 0; i < N; i++) f[i] = (float)dens[i]; break;
        case MG_FAGE: std::memcpy(u, age.data(), bytes); break;
        case MG_FCOLOR: std::memcpy(u, color.data(), bytes); break;
        case MG_FGCELL: std::memcpy(u, gcell.data(), bytes); break;
        case MG_FGNDX: std::memcpy(u, gndx.data(), bytes); break;
        case MG_FNERVEIDX: std::memcpy(u, nerve.data(), bytes); break;
        case MG_FBIN: std::memcpy(u, bin.data(), bytes); break;
        case MG_FPARTICLE_ID: std::memcpy(u, pid.data(), bytes); break;
        case MG_FMASS_RADIUS: std::memcpy(u, mrad.data(), bytes); break;
        case MG_FELASTIDX:
            std::memcpy(u, bu.data(), bytes);
            if (!F32)
                for (size_t k = 0; k < bu.size(); k++)
                    if (is_float_slot((int)(k % MG_DATA_PER_BOND)) && bu[k] != MG_NONE)
                        u[k] = f2u((float)bd[k]);
            break;
        case MG_FPARTICLEIDX: std::memcpy(u, pidx.data(), bytes); break;
        case MG_FCONC: for (size_t k = 0; k < conc.size(); k++) f[k] = (float)conc[k]; break;
        case MG_FEPIGEN: std::memcpy(u, epigen.data(), bytes); break;
        case MG_FBIN_COUNT: std::memcpy(u, cnt.data(), bytes); break;
        case MG_FBIN_OFFSET: std::memcpy(u, off.data(), bytes); break;
        case MG_FBIN_COUNT_ACTIVE_GENES: std::memcpy(u, gcnt.data(), bytes); break;
        case MG_FBIN_OFFSET_ACTIVE_GENES: std::memcpy(u, goff.data(), bytes); break;
        case MG_FDENSE_LIST_LENGTHS: std::memcpy(u, lens, bytes); break;
        case MG_FDENSE_LISTS: std::memcpy(u, lists.data(), bytes); break;
        case MG_FBIN_COUNT_CHANGES: std::memcpy(u, ccnt.data(), bytes); break;
        case MG_FBIN_OFFSET_CHANGES: std::memcpy(u, coff.data(), bytes); break;
        case MG_FDENSE_LIST_LENGTHS_CHANGES: std::memcpy(u, clen, bytes); break;
        case MG_FDENSE_LISTS_CHANGES: std::memcpy(u, clists.data(), bytes); break;
        default: return MG_ERR_UNSUPPORTED;
        }
        return MG_OK;
    }

    /* ------------------------------------------------------- insert (a4)
     * src/fluid_system_cuda.cu:53-100 with the host memsets of src/host_CUDA.cpp:249-253.
     * Launched over all mMaxPoints slots (src/host_CUDA.cpp:258). */
    void insert() override
    {
        std::fill(cnt.begin(), cnt.end(), 0u);
        std::fill(off.begin(), off.end(), 0u);
        std::fill(gcnt.begin(), gcnt.end(), 0u);
        std::fill(goff.begin(), goff.end(), 0u);
        const R mnx = (R)g.gmin[0], mny = (R)g.gmin[1], mnz = (R)g.gmin[2];
        const R dx = (R)g.delta[0], dy = (R)g.delta[1], dz = (R)g.delta[2];
        for (int i = 0; i < N; i++) {
            R gfx = (pos[3 * (size_t)i] - mnx) * dx;
            R gfy = (pos[3 * (size_t)i + 1] - mny) * dy;
            R gfz = (pos[3 * (size_t)i + 2] - mnz) * dz;
            int cx = trunc_i(gfx), cy = trunc_i(gfy), cz = trunc_i(gfz);
            if (cx >= 1 && cx <= g.scan_max[0] && cy >= 1 && cy <= g.scan_max[1] && cz >= 1 &&
                cz <= g.scan_max[2]) {
                int gs = (cy * g.res[2] + cz) * g.res[0] + cx;
                gcell[i] = (uint32_t)gs;
                gndx[i] = cnt[gs]++; /* atomicAdd in ascending i = stable rank */
                for (int gene = 0; gene < MG_NUM_GENES; gene++)
                    if (epigen[(size_t)gene * N + i] > 0) gcnt[(size_t)gene * C + gs]++;
            } else {
                gcell[i] = MG_GRID_UNDEF;
            }
        }
    }

    /* ----------------------------------------------------- prefix sum (a5, a6)
     * prefixSum/prefixFixup with zero_offsets=1 (src/fluid_system_cuda.cu:136-183) as driven by
     * src/host_CUDA.cpp:307-381 compute the exclusive scan; tally_denselist_lengths
     * (src/fluid_system_cuda.cu:185-194) gives the gene list lengths. */
    void prefix_sum() override
    {
        uint32_t run = 0;
        for (int c = 0; c < C; c++) { off[c] = run; run += cnt[c]; }
        for (int gene = 0; gene < MG_NUM_GENES; gene++) {
            uint32_t r = 0;
            for (int c = 0; c < C; c++) {
                goff[(size_t)gene * C + c] = r;
                r += gcnt[(size_t)gene * C + c];
            }
            lens[gene] = gcnt[(size_t)gene * C + C - 1] + goff[(size_t)gene * C + C - 1];
        }
    }

    /* ------------------------------------------------ counting sort (a7, a8)
     * Host side src/host_CL.cpp:1064-1206 / src/host_CUDA.cpp:498-580: read mNumPoints =
     * FBIN_OFFSET[C-1] and mActivePoints = FBIN_OFFSET[grid_ScanMax], copy fbuf -> ftemp,
     * fill FPOS with max(VOLMAX) and reset FELASTIDX/FPARTICLEIDX/FPARTICLE_ID to the
     * sentinel and FFORCE/FCONC/FEPIGEN to zero (whole buffers, the intent of the
     * "mMaxPoints*3 ie num elements" memset), then scatter (src/fluid_system_cuda.cu:196-296)
     * and build the dense gene lists (src/fluid_system_cuda.cu:329-404). */
    void sort() override
    {
        const int gsm = (g.scan_max[1] * g.res[2] + g.scan_max[2]) * g.res[0] + g.scan_max[0];
        num_points = off[C - 1];
        active_points = off[gsm];
        const uint32_t pnum = (uint32_t)N; /* countingSortFull is launched with mMaxPoints */

        /* ftemp copies */
        std::vector<R> t_pos = pos, t_vel = vel, t_vev = veval, t_press = press, t_dens = dens;
        std::vector<uint32_t> t_age = age, t_color = color, t_gcell = gcell, t_gndx = gndx;
        std::vector<uint32_t> t_bu = bu, t_pidx = pidx, t_pid = pid, t_mrad = mrad, t_nerve = nerve;
        std::vector<double> t_bd = bd;
        std::vector<R> t_conc = conc;
        std::vector<uint32_t> t_epi = epigen, t_home = home;

        float vmax = std::max(std::max(g.gmax[0], g.gmax[1]), g.gmax[2]);
        std::fill(pos.begin(), pos.end(), (R)vmax);
        std::fill(bu.begin(), bu.end(), MG_NONE);
        if (!F32) std::fill(bd.begin(), bd.end(), (double)u2f(MG_NONE));
        std::fill(pidx.begin(), pidx.end(), MG_NONE);
        std::fill(pid.begin(), pid.end(), MG_NONE);
        std::fill(frc.begin(), frc.end(), R(0));
        std::fill(conc.begin(), conc.end(), R(0));
        std::fill(epigen.begin(), epigen.end(), 0u);

        for (uint32_t i = 0; i < pnum; i++) {
            uint32_t icell = t_gcell[i];
            if (icell == MG_GRID_UNDEF) continue;
            uint32_t indx = t_gndx[i];
            uint32_t s = off[icell] + indx;
            bin[s] = s;
            for (int k = 0; k < 3; k++) {
                pos[3 * (size_t)s + k] = t_pos[3 * (size_t)i + k];
                vel[3 * (size_t)s + k] = t_vel[3 * (size_t)i + k];
                veval[3 * (size_t)s + k] = t_vev[3 * (size_t)i + k];
                frc[3 * (size_t)s + k] = R(0);
            }
            press[s] = t_press[i];
            dens[s] = t_dens[i];
            age[s] = t_age[i];
            color[s] = t_color[i];
            gcell[s] = icell;
            gndx[s] = indx;
            home[s] = t_home[i];
            for (int a = 0; a < B; a++) {
                size_t src = brec(i, a), dst = brec(s, a);
                uint32_t j = t_bu[src + MG_B_INDEX];
                uint32_t js = MG_NONE;
                if (j < pnum) {
                    uint32_t jcell = t_gcell[j];
                    if (jcell != MG_GRID_UNDEF) {
                        uint32_t jndx = t_gndx[j];
                        if (off[jcell] + jndx < pnum) js = off[jcell] + jndx;
                    }
                }
                bu[dst + MG_B_INDEX] = js;
                for (int k = 1; k < MG_DATA_PER_BOND; k++) {
                    bu[dst + k] = t_bu[src + k];
                    if (!F32) bd[dst + k] = t_bd[src + k];
                }
            }
            for (int a = 0; a < B; a++) {
                uint32_t k = t_pidx[((size_t)i * B + a) * 2];
                uint32_t b = t_pidx[((size_t)i * B + a) * 2 + 1];
                uint32_t ks = MG_NONE;
                if (k < pnum) {
                    uint32_t kcell = t_gcell[k];
                    if (kcell != MG_GRID_UNDEF) ks = off[kcell] + t_gndx[k];
                }
                pidx[((size_t)s * B + a) * 2] = ks;
                pidx[((size_t)s * B + a) * 2 + 1] = b;
            }
            pid[s] = t_pid[i];
            mrad[s] = t_mrad[i];
            nerve[s] = t_nerve[i];
            for (int a = 0; a < MG_NUM_GENES; a++) epigen[(size_t)pnum * a + s] = t_epi[(size_t)pnum * a + i];
            for (int a = 0; a < MG_NUM_TF; a++) conc[(size_t)s * MG_NUM_TF + a] = t_conc[(size_t)i * MG_NUM_TF + a];
        }
        dense_lists();
    }

    /* countingSortDenseLists (src/fluid_system_cuda.cu:329-404): per bin, per particle in
     * ascending order, per gene with nonzero FEPIGEN, append to the gene's list. */
    void dense_lists()
    {
        std::fill(lists.begin(), lists.end(), MG_NONE);
        for (int bin_ = 0; bin_ < C; bin_++) {
            uint32_t count = cnt[bin_];
            if (count == 0) continue;
            uint32_t grdoffset = off[bin_];
            if (grdoffset + count > (uint32_t)N) continue; /* overflow guard, :360 */
            uint32_t ctr[MG_NUM_GENES] = {0};
            for (uint32_t particle = grdoffset; particle < grdoffset + count; particle++)
                for (int gene = 0; gene < MG_NUM_GENES; gene++)
                    if (epigen[particle + (size_t)N * gene] > 0) {
                        uint32_t at = goff[(size_t)gene * C + bin_] + ctr[gene];
                        if (at < (uint32_t)N) lists[(size_t)gene * N + at] = particle;
                        ctr[gene]++;
                    }
        }
    }

    /* --------------------------------------------------------- pressure (a9)
     * contributePressure/computePressure, src/fluid_system_cuda.cu:620-712. */
    void pressure() override
    {
        const int na = (int)active_points;
        const R r2 = (R)p.r2, sr = (R)p.psmoothradius, d2 = (R)p.d2;
        const bool poly6 = sph == MG_SPH_POLY6_SPIKY;
        const int nadj = (1 * g.res[2] + 1) * g.res[0] + 1;
#pragma omp parallel for schedule(static) num_threads(nth())
        for (int i = 0; i < na; i++) {
            uint32_t gc = gcell[i];
            if (gc == MG_GRID_UNDEF) continue;
            gc -= nadj;
            V3<R> pi = P(i);
            R sum = R(0);
            for (int c = 0; c < g.adj_cnt; c++) {
                int cell = (int)gc + g.adj[c];
                R csum = R(0);
                if (cnt[cell] != 0) {
                    uint32_t clast = off[cell] + cnt[cell];
                    for (uint32_t cndx = off[cell]; cndx < clast; cndx++) {
                        uint32_t pndx = bin[cndx];
                        V3<R> dist = pi - P(pndx);
                        R dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
                        if (dsq < r2 && dsq > R(0)) {
                            if (poly6) {
                                /* Mueller 2003 poly6 (h^2 - r^2)^3, as the reference's commented
                                 * sum at src/fluid_system_cuda.cu:668-672 */
                                R c = (r2 - dsq) * d2;
                                csum += c * c * c;
                            } else {
                                R r = std::sqrt(dsq);
                                R q = r / sr;
                                R b = R(1) - q;
                                b *= b;
                                b *= b;
                                csum += b * (R(4) * q + R(1));
                            }
                        }
                    }
                }
                sum += csum;
            }
            sum = sum * (R)p.pmass * (R)(poly6 ? p.poly6kern : p.wendlandC2kern); /* :695 */
            if (sum == R(0)) sum = R(1);
            press[i] = (sum - (R)p.prest_dens) * (R)p.pintstiff;
            dens[i] = R(1) / sum;
        }
    }

    /* ------------------------------------------------------------ force (a10)
     * contributeForce/computeForce, src/fluid_system_cuda.cu:3167-3358. Pass 1 computes each
     * particle's own total (bond terms subtracted in slot order, then the per-cell fluid sums
     * times pmass) and every bond term; pass 2 adds them to FFORCE in canonical order. */
    void force() override
    {
        const int na = (int)active_points;
        const uint32_t pnum = (uint32_t)na;
        const R sr = (R)p.psmoothradius, r2 = (R)p.r2, d2 = (R)p.d2;
        const bool poly6 = sph == MG_SPH_POLY6_SPIKY;
        const int nadj = (1 * g.res[2] + 1) * g.res[0] + 1;
        const bool freeze = p.freeze != 0;
        std::vector<R> own((size_t)N * 3, R(0));
        std::vector<R> et((size_t)N * B * 3, R(0));
        std::vector<uint8_t> ok((size_t)N * B, 0); /* bond term was applied */

#pragma omp parallel for schedule(static) num_threads(nth())
        for (int i = 0; i < na; i++) {
            uint32_t gc = gcell[i];
            if (gc == MG_GRID_UNDEF) continue;
            gc -= nadj;
            V3<R> force = mk(R(0), R(0), R(0));
            V3<R> ipos = P(i);
            V3<R> pvel = VE(i);
            for (int a = 0; a < B; a++) {
                size_t bond = brec(i, a);
                uint32_t j = bu[bond + MG_B_INDEX];
                R restlength = bf(bond + MG_B_REST);
                if (j >= pnum || (double)restlength < 0.000000001) continue;
                R elastic_limit = bf(bond + MG_B_LIMIT);
                R modulus = bf(bond + MG_B_MODULUS);
                R damping_coeff = bf(bond + MG_B_DAMP);
                V3<R> j_pos = P(j);
                V3<R> dist = ipos - j_pos;
                R dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
                R abs_dist = std::sqrt(dsq) + (R)FLT_MIN;
                V3<R> rel_vel = VE(j) - pvel;
                R spring_strain = std::fmax(R(0), (abs_dist - restlength) / restlength);
                sbf(bond + MG_B_STRAIN2, bf(bond + MG_B_STRAIN2) + spring_strain * spring_strain);
                sbf(bond + MG_B_STRAIN, bf(bond + MG_B_STRAIN) + spring_strain);
                R fac = (R)(abs_dist < elastic_limit ? 1 : 0);
                V3<R> e1 = divs(dist, abs_dist) * spring_strain * modulus;
                V3<R> e2 = damping_coeff * rel_vel;
                V3<R> eterm = divs(fac * (e1 - e2), (R)p.pmass);
                if (!has_nan(eterm)) {
                    force = force - eterm;
                    size_t e = ((size_t)i * B + a) * 3;
                    et[e] = eterm.x; et[e + 1] = eterm.y; et[e + 2] = eterm.z;
                    ok[(size_t)i * B + a] = 1;
                }
                if (abs_dist >= elastic_limit && !freeze) sbf(bond + MG_B_REST, R(0));
            }
            V3<R> ivev = VE(i);
            R ipress = press[i], idens = dens[i];
            V3<R> fsum = mk(R(0), R(0), R(0));
            for (int c = 0; c < g.adj_cnt; c++) {
                int cell = (int)gc + g.adj[c];
                V3<R> cf = mk(R(0), R(0), R(0));
                if (cnt[cell] != 0) {
                    uint32_t clast = off[cell] + cnt[cell];
                    for (uint32_t cndx = off[cell]; cndx < clast; cndx++) {
                        uint32_t j = bin[cndx];
                        V3<R> dist = ipos - P(j);
                        R dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
                        if (poly6) {
                            /* Mueller 2003 spiky pressure and viscosity Laplacian terms in the
                             * arrangement of the reference's comments
                             * (src/fluid_system_cuda.cu:3140-3160, 3213-3220), over the support
                             * r < sr */
                            if (dsq < r2 && dsq > R(0)) {
                                R sdist = std::sqrt(dsq * d2);
                                R c = sr - sdist;
                                R pterm = (R)p.psimscale * R(-0.5) * c * (R)p.spikykern * (ipress + press[j]) / sdist;
                                R dterm = c * idens * dens[j];
                                V3<R> delta_v = VE(j) - ivev;
                                cf = cf + (dist * pterm + delta_v * (R)p.vterm) * dterm;
                            }
                            continue;
                        }
                        R r = std::sqrt(dsq);
                        if (dsq < R(1) && dsq > R(0)) {
                            R t = sr - r;
                            R kern = t * t * t;
                            R press2 = R(100) * (ipress + press[j]);
                            R s = idens * dens[j] * R(100);
                            V3<R> pterm = (s * divs(dist, r)) * (press2 * kern - (R)p.psurface_t * (t * t));
                            V3<R> delta_v = VE(j) - ivev;
                            V3<R> vterm = (R(100000) * delta_v) * kern;
                            cf = cf + (pterm + vterm);
                        }
                    }
                }
                fsum = fsum + cf;
            }
            force = force + fsum * (R)p.pmass;
            own[3 * (size_t)i] = force.x; own[3 * (size_t)i + 1] = force.y; own[3 * (size_t)i + 2] = force.z;
        }

        /* canonical incoming order: owners by (particle ID, home index), then slot */
        std::vector<uint32_t> owners((size_t)na);
        for (int s = 0; s < na; s++) owners[s] = (uint32_t)s;
        std::sort(owners.begin(), owners.end(), [&](uint32_t x, uint32_t y) {
            return pid[x] != pid[y] ? pid[x] < pid[y] : home[x] < home[y];
        });
        std::vector<uint32_t> icnt(na + 1, 0);
        for (uint32_t k : owners) {
            for (int a = 0; a < B; a++)
                if (ok[(size_t)k * B + a]) icnt[bu[brec(k, a) + MG_B_INDEX] + 1]++;
        }
        for (int j = 0; j < na; j++) icnt[j + 1] += icnt[j];
        std::vector<uint32_t> ient(icnt[na]);
        std::vector<uint32_t> fill(icnt.begin(), icnt.end() - 1);
        for (uint32_t k : owners) {
            for (int a = 0; a < B; a++)
                if (ok[(size_t)k * B + a]) ient[fill[bu[brec(k, a) + MG_B_INDEX]]++] = k * B + a;
        }
#pragma omp parallel for schedule(static) num_threads(nth())
        for (int j = 0; j < na; j++) {
            if (gcell[j] == MG_GRID_UNDEF) continue;
            V3<R> F = mk(frc[3 * (size_t)j], frc[3 * (size_t)j + 1], frc[3 * (size_t)j + 2]);
            F = F + mk(own[3 * (size_t)j], own[3 * (size_t)j + 1], own[3 * (size_t)j + 2]);
            for (uint32_t e = icnt[j]; e < icnt[j + 1]; e++) {
                size_t q = (size_t)ient[e] * 3;
                F = F + mk(et[q], et[q + 1], et[q + 2]);
            }
            frc[3 * (size_t)j] = F.x; frc[3 * (size_t)j + 1] = F.y; frc[3 * (size_t)j + 2] = F.z;
        }
    }

    /* ---------------------------------------------------------- advance (a11)
     * advanceParticles, src/fluid_system_cuda.cu:3470-3614, with the fbuf->ftemp->fbuf
     * copies of src/host_CL.cpp:199-226 folded in: a GRID_UNDEF particle ends the step
     * unchanged (its parked position is overwritten by the copy back). */
    void advance(float time_, float dt_, float ss) override
    {
        (void)time_; (void)ss;
        const int na = (int)active_points;
        const R dt = (R)dt_;
        const R eps = (R)MG_EPSILON;
        const R prad = (R)p.pradius, slope = (R)p.pground_slope;
        const R estiff = (R)p.pextstiff, damp = (R)p.pdamp;
        const V3<R> bmin = mk((R)p.pboundmin[0], (R)p.pboundmin[1], (R)p.pboundmin[2]);
        const V3<R> bmax = mk((R)p.pboundmax[0], (R)p.pboundmax[1], (R)p.pboundmax[2]);
        const R ny = (R)(1.0 - (double)p.pground_slope);
#pragma omp parallel for schedule(static) num_threads(nth())
        for (int i = 0; i < na; i++) {
            if (gcell[i] == MG_GRID_UNDEF) continue;
            V3<R> pos_ = P(i);
            V3<R> vev = VE(i);
            V3<R> accel = mk(frc[3 * (size_t)i], frc[3 * (size_t)i + 1], frc[3 * (size_t)i + 2]);
            V3<R> norm;
            R diff, adj, speed;
            diff = prad - (pos_.y - (bmin.y + (pos_.x - bmin.x) * slope));
            if (diff > eps) {
                norm = mk(-slope, ny, R(0));
                adj = estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            diff = prad - (bmax.y - pos_.y);
            if (diff > eps) {
                norm = mk(R(0), R(-1), R(0));
                adj = estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            diff = prad - (pos_.x - bmin.x);
            if (diff > eps) {
                norm = mk(R(1), R(0), R(0));
                adj = ((R)p.pforce_min + R(1)) * estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            diff = prad - (bmax.x - pos_.x);
            if (diff > eps) {
                norm = mk(R(-1), R(0), R(0));
                adj = ((R)p.pforce_max + R(1)) * estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            diff = prad - (pos_.z - bmin.z);
            if (diff > eps) {
                norm = mk(R(0), R(0), R(1));
                adj = estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            diff = prad - (bmax.z - pos_.z);
            if (diff > eps) {
                norm = mk(R(0), R(0), R(-1));
                adj = estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            V3<R> dist = bmax - pos_;
            diff = R(2) * prad - (dist.x + dist.y + dist.z);
            if (diff > eps) {
                norm = mk(R(1), R(1), R(1));
                adj = estiff * diff - damp * dot(norm, vev);
                norm = norm * adj; accel = accel + norm;
            }
            accel = accel + mk((R)p.pgravity[0], (R)p.pgravity[1], (R)p.pgravity[2]);
            speed = accel.x * accel.x + accel.y * accel.y + accel.z * accel.z;
            if (speed > (R)p.AL2) accel = accel * ((R)p.AL / std::sqrt(speed));
            V3<R> v = mk(vel[3 * (size_t)i], vel[3 * (size_t)i + 1], vel[3 * (size_t)i + 2]);
            speed = v.x * v.x + v.y * v.y + v.z * v.z;
            if (speed > (R)p.VL2) {
                speed = (R)p.VL2;
                v = v * ((R)p.VL / std::sqrt(speed)); /* no-op limit, src/fluid_system_cuda.cu:3577-3579 */
            }
            V3<R> vnext = accel * dt + v;
            V3<R> ve = (v + vnext) * R(0.5);
            V3<R> np = pos_ + vnext * dt;
            for (int k = 0; k < 3; k++) {
                veval[3 * (size_t)i + k] = (&ve.x)[k];
                vel[3 * (size_t)i + k] = (&vnext.x)[k];
                pos[3 * (size_t)i + k] = (&np.x)[k];
            }
        }
    }

    /* advance followed by SpecialParticlesCL (src/host_CUDA.cpp:869-898): externalActuation
     * (src/fluid_system_cuda.cu:3617-3692) gives gene 12's list members below mActivePoints
     * their pre-advance state with z + zoff, then fixedParticles (:3694-3705) gives gene 11's
     * members below mNumPoints their pre-advance state. */
    void advance_special(float time_, float dt_, float ss, float zoff) override
    {
        const std::vector<R> p0 = pos, v0 = vel, e0 = veval;
        advance(time_, dt_, ss);
        auto restore = [&](uint32_t i) {
            for (int k = 0; k < 3; k++) {
                pos[3 * (size_t)i + k] = p0[3 * (size_t)i + k];
                vel[3 * (size_t)i + k] = v0[3 * (size_t)i + k];
                veval[3 * (size_t)i + k] = e0[3 * (size_t)i + k];
            }
        };
        for (uint32_t li = 0; li < lens[12] && li < (uint32_t)N; li++) {
            const uint32_t i = lists[(size_t)12 * N + li];
            if (i >= active_points || gcell[i] == MG_GRID_UNDEF) continue;
            restore(i);
            pos[3 * (size_t)i + 2] = pos[3 * (size_t)i + 2] + (R)zoff;
        }
        for (uint32_t li = 0; li < lens[11] && li < (uint32_t)N; li++) {
            const uint32_t i = lists[(size_t)11 * N + li];
            if (i >= num_points) continue;
            restore(i);
        }
    }

    /* -------------------------------------------------------- diffusion (a12)
     * contributeDiffusion/computeDiffusion, src/fluid_system_cuda.cu:3026-3105, as a Jacobi
     * update. The uint diffusability is scaled by diff_scale (1.0 = reference). */
    void diffusion() override
    {
        const int na = (int)active_points;
        const R d2 = (R)p.psimscale * (R)p.psimscale;
        const R r2 = (R)p.r2 / d2;
        const int nadj = (1 * g.res[2] + 1) * g.res[0] + 1;
        std::vector<R> old = conc;
        R D[MG_NUM_TF], K[MG_NUM_TF];
        for (int t = 0; t < MG_NUM_TF; t++) {
            D[t] = (R)gen.tf_diffusability[t] * (R)p.diff_scale;
            K[t] = (R)gen.tf_breakdown_rate[t];
        }
#pragma omp parallel for schedule(static) num_threads(nth())
        for (int i = 0; i < na; i++) {
            R cur[MG_NUM_TF], nw[MG_NUM_TF];
            for (int t = 0; t < MG_NUM_TF; t++) { cur[t] = old[(size_t)i * MG_NUM_TF + t]; nw[t] = R(0); }
            uint32_t gc = gcell[i];
            if (gc == MG_GRID_UNDEF) continue;
            gc -= nadj;
            V3<R> pi = P(i);
            for (int c = 0; c < g.adj_cnt; c++) {
                int cell = (int)gc + g.adj[c];
                if (cnt[cell] == 0) continue;
                uint32_t clast = off[cell] + cnt[cell];
                for (uint32_t cndx = off[cell]; cndx < clast; cndx++) {
                    uint32_t pndx = bin[cndx];
                    V3<R> dist = pi - P(pndx);
                    R dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
                    if (dsq < r2 && dsq > R(0)) {
                        R cc = (r2 - dsq) * d2;
                        for (int t = 0; t < MG_NUM_TF; t++)
                            if (gen.tf_diffusability[t])
                                nw[t] += D[t] * cc * (old[(size_t)pndx * MG_NUM_TF + t] - cur[t]);
                    }
                }
            }
            for (int t = 0; t < MG_NUM_TF; t++) conc[(size_t)i * MG_NUM_TF + t] = K[t] * (cur[t] + nw[t]);
        }
    }

    /* ------------------------------------------------------------ genes (a13)
     * computeGeneAction + tallyGeneAction, src/fluid_system_cuda.cu:714-772, driven per gene
     * by src/host_CUDA.cpp:681-710. Fixed semantics of this build:
     *  - activity reads FCONC as float in its AoS layout (the reference reads the float bits
     *    as uint at the SoA index p+maxPoints*tf, :739);
     *  - secretion and activation accumulate into per-particle float / uint deltas (the
     *    reference atomically adds a float-to-uint conversion into ftemp, :747, :754);
     *  - the countdown only touches gene+1 when gene+1 < NUM_GENES (the reference reads one
     *    gene past the end for gene 15, :727);
     *  - the tally applies the deltas once to each particle of the "living" list 2 with index
     *    below mActivePoints (the reference launches it once per gene over the first len[gene]
     *    entries of list 2, applying it up to 16 times, src/host_CUDA.cpp:699-709).
     * The secretion rate is scaled by secrete_scale (1.0 = reference). */
    void genes() override
    {
        const uint32_t pnum = (uint32_t)N;
        std::vector<R> dconc((size_t)N * MG_NUM_TF, R(0));
        std::vector<uint32_t> depi((size_t)N * MG_NUM_GENES, 0);
        for (int gene = 0; gene < MG_NUM_GENES; gene++) {
            uint32_t len = lens[gene];
            if (len > pnum) len = pnum;
            for (uint32_t li = 0; li < len; li++) {
                uint32_t pi = lists[(size_t)gene * N + li];
                if (pi >= pnum) continue;
                uint32_t& e = epigen[(size_t)gene * pnum + pi];
                int delay = (int)e;
                if (0 < delay && delay < MG_INT_MAX) {
                    e--;
                    if (delay == 1 && gene + 1 < MG_NUM_GENES && epigen[(size_t)(gene + 1) * pnum + pi])
                        epigen[(size_t)(gene + 1) * pnum + pi] = gen.delay[gene + 1];
                }
                R activity = R(0);
                for (int tf = 0; tf < MG_NUM_TF; tf++)
                    if (gen.sensitivity[gene][tf] != 0)
                        activity += (R)gen.sensitivity[gene][tf] * conc[(size_t)pi * MG_NUM_TF + tf];
                int numTF = gen.secrete[gene][2 * MG_NUM_TF];
                for (int j = 0; j < numTF && j < MG_NUM_TF; j++) {
                    int tf = gen.secrete[gene][j * 2];
                    int rate = gen.secrete[gene][j * 2 + 1];
                    if (tf < 0 || tf >= MG_NUM_TF) continue;
                    dconc[(size_t)pi * MG_NUM_TF + tf] += ((R)rate * activity) * (R)p.secrete_scale;
                }
                int numL = gen.activate[gene][2 * MG_NUM_GENES];
                for (int j = 0; j < numL && j < MG_NUM_GENES; j++) {
                    int other = gen.activate[gene][j * 2];
                    int threshold = gen.activate[gene][j * 2 + 1];
                    if (other < 0 || other >= MG_NUM_GENES) continue;
                    if ((R)threshold < activity) depi[(size_t)other * pnum + pi] += 1u;
                }
            }
        }
        uint32_t len2 = std::min(lens[2], pnum);
        for (uint32_t li = 0; li < len2; li++) {
            uint32_t i = lists[(size_t)2 * N + li];
            if (i >= active_points) continue;
            for (int j = 0; j < MG_NUM_TF; j++)
                conc[(size_t)i * MG_NUM_TF + j] += dconc[(size_t)i * MG_NUM_TF + j];
            for (int j = 0; j < MG_NUM_GENES; j++)
                epigen[(size_t)j * pnum + i] += depi[(size_t)j * pnum + i];
        }
    }

    /* ZeroVelCL (src/host_CL.cpp:2175-2180), zeroing the whole FVEL and FVEVAL buffers. */
    void zero_vel() override
    {
        std::fill(vel.begin(), vel.end(), R(0));
        std::fill(veval.begin(), veval.end(), R(0));
    }

    /* ------------------------------------------------ bond formation
     * initialize_bonds + contribFindBonds, src/fluid_system_cuda.cu:1938-2044, over gene 1's
     * dense list below mActivePoints. Per bond slot: candidates j of the cells gc + gridAdj[c]
     * (no (1,1,1) offset is subtracted in the reference here) with ID != 0 and
     * 0 < dsq < rd2, scored by sum_k |1.91 - acos(u.v / (|u| |v|))| over the particle's current
     * bonds k (FLT_MAX when j is one of them; the sum runs in double as the reference's
     * double literal 1.91 makes it), or by dsq for slot 0; the first strictly best candidate
     * below mActivePoints gets the slot, with the genome's default elastLim, rest length,
     * modulus and damping of the slot's bond type (src/fluid_system_cuda.cu:1975-2043).
     * FPARTICLEIDX is not updated, as in the reference. The candidate index is the slot
     * itself (FBIN is the identity after a sort, src/fluid_system_cuda.cu:221). */
    /* contribFindBonds (src/fluid_system_cuda.cu:1938-1976) over the cells gc + gridAdj[c]
     * (no (1,1,1) offset, as in the reference) for bond slot `bond` of particle i, seen from x:
     * candidates with a non-zero ID and 0 < |x - x_j|^2 < rd2, scored by the sum over i's
     * current bonds k of |1.91 - acos(u.v / (|u| |v|))| (FLT_MAX when j is already bonded) or by
     * |u|^2 for slot 0; the first strictly best, or MG_NONE. */
    uint32_t find_bond(uint32_t i, uint32_t gc, V3<R> x, int bond) const
    {
        const R rd2 = (R)p.rd2;
        R best = (R)FLT_MAX;
        uint32_t to = MG_NONE;
        for (int c = 0; c < g.adj_cnt; c++) {
            const uint32_t cell = gc + (uint32_t)g.adj[c];
            if (cell >= (uint32_t)C || cnt[cell] == 0) continue;
            const uint32_t c0 = off[cell];
            for (uint32_t j = c0; j < c0 + cnt[cell] && j < (uint32_t)N; j++) {
                if (pid[j] == 0u) continue;
                const V3<R> u = x - P(j);
                const R dsq = dot(u, u);
                if (!(dsq < rd2 && dsq > R(0))) continue;
                R theta = R(0);
                for (int b = 0; b < B; b++) {
                    const uint32_t k = bu[brec(i, b) + MG_B_INDEX];
                    if (j == k) {
                        theta = (R)FLT_MAX;
                        break;
                    }
                    if (k < (uint32_t)N) {
                        const V3<R> v = x - P(k);
                        const R lu = std::sqrt(dot(u, u)), lv = std::sqrt(dot(v, v));
                        const R cosang = dot(u, v) / (lu * lv);
                        theta = (R)((double)theta + std::fabs(1.91 - (double)acos_r(cosang)));
                    }
                }
                if (bond == 0) theta = dsq;
                if (theta < best) {
                    best = theta;
                    to = j;
                }
            }
        }
        return to;
    }

    /* bond types of slots 0-3 by tissue gene (src/fluid_system_cuda.cu:1133-1139, :2003-2009):
     * bone 9 -> apatite in all four, tendon 6 / muscle 7 / elastic ligament 10 -> collagen in
     * slot 0, elastin otherwise (the cartilage branch re-tests gene 6 and is never taken) */
    void bond_types(uint32_t i, int bt4[4]) const
    {
        auto on = [&](int gi) { return epigen[(size_t)gi * N + i] > 0u; };
        bt4[0] = bt4[1] = bt4[2] = bt4[3] = 0;
        if (on(9)) bt4[0] = bt4[1] = bt4[2] = bt4[3] = 2;
        else if (on(6) || on(7) || on(10)) bt4[0] = 1;
    }

    /* initialize_bonds (src/fluid_system_cuda.cu:1978-2044) over gene 1's dense list below
     * mActivePoints: each slot in turn takes find_bond's candidate with the slot type's
     * default parameters. */
    void initialize_bonds() override
    {
        const uint32_t na = active_points, len = lens[1];
        for (uint32_t li = 0; li < len && li < (uint32_t)N; li++) {
            const uint32_t i = lists[(size_t)N + li];
            if (i >= na) continue;
            const uint32_t gc = gcell[i];
            if (gc == MG_GRID_UNDEF) continue;
            int bt4[4];
            bond_types(i, bt4);
            for (int bond = 0; bond < B; bond++) {
                const uint32_t to = find_bond(i, gc, P(i), bond);
                if (to < na) make_bond(i, bond, to, bond < 4 ? bt4[bond] : 0);
            }
        }
    }

    /* the record words 0-7 written by initialize_bonds and makeBond
     * (src/fluid_system_cuda.cu:1573-1596): partner, the bond type's default limit / rest
     * length / modulus / damping, partner ID, words 6-7 zero (makeBond also clears word 8) */
    void make_bond(uint32_t i, int a, uint32_t to, int bt)
    {
        const size_t r = brec(i, a);
        bu[r + MG_B_INDEX] = to;
        sbf(r + MG_B_LIMIT, (R)gen.param[bt][8]);
        sbf(r + MG_B_REST, (R)gen.param[bt][9]);
        sbf(r + MG_B_MODULUS, (R)gen.param[bt][10]);
        sbf(r + MG_B_DAMP, (R)gen.param[bt][11]);
        bu[r + MG_B_ID] = pid[to];
        sbf(r + MG_B_STRAIN2, R(0));
        sbf(r + MG_B_STRAIN, R(0));
    }

    /* ----------------------------------------------------------- remodelling
     * computeBondChanges (src/fluid_system_cuda.cu:1065-1330) with the host clears of
     * ComputeBondChangesCL (src/host_CL.cpp:2057-2086), over gene 2's dense list below
     * mActivePoints: per slot in order, rest length > 0 -> integrators averaged over `steps`
     * and cleared, rest *= l_a + l_b tanh(l_c <strain> - l_d), modulus *= s_a + s_b
     * tanh(s_c <strain^2> - s_d); change word cleared; the first broken (rest 0) or empty slot
     * is counted for heal in the particle's cell, gets change bit 0 and ends the loop. */
    void compute_bond_changes(uint32_t steps) override
    {
        std::fill(ccnt.begin(), ccnt.end(), 0u);
        std::fill(coff.begin(), coff.end(), 0u);
        std::fill(hs.begin(), hs.end(), MG_NONE);
        const uint32_t na = active_points, len = lens[2];
        const R fs = (R)steps;
        for (uint32_t li = 0; li < len && li < (uint32_t)N; li++) {
            const uint32_t i = lists[(size_t)2 * N + li];
            if (i >= na) continue;
            int bt4[4];
            bond_types(i, bt4);
            for (int a = 0; a < B; a++) {
                const size_t r = brec(i, a);
                R rest = bf(r + MG_B_REST);
                if (rest > R(0)) {
                    const R si = bf(r + MG_B_STRAIN) / fs, ssi = bf(r + MG_B_STRAIN2) / fs;
                    sbf(r + MG_B_STRAIN, R(0));
                    sbf(r + MG_B_STRAIN2, R(0));
                    const float* t = gen.tanh_param[a < 4 ? bt4[a] : 0];
                    rest = rest * ((R)t[0] + (R)t[1] * tanh_r((R)t[2] * si - (R)t[3]));
                    sbf(r + MG_B_REST, rest);
                    sbf(r + MG_B_MODULUS,
                        bf(r + MG_B_MODULUS) * ((R)t[4] + (R)t[5] * tanh_r((R)t[6] * ssi - (R)t[7])));
                }
                bu[r + MG_B_CHANGE] = 0;
                if (rest == R(0) || bu[r + MG_B_INDEX] == MG_NONE) {
                    const uint32_t gc = gcell[i];
                    if (gc < (uint32_t)C) {
                        ccnt[gc]++;
                        hs[i] = (uint32_t)a;
                    }
                    bu[r + MG_B_CHANGE] = 1;
                    break;
                }
            }
        }
    }

    /* PrefixSumChangesCL + tally_denselist_lengths (src/host_CL.cpp:865-1062,
     * src/fluid_system_cuda.cu:185-194): exclusive offsets per list, length = last count +
     * last offset */
    void prefix_sum_changes() override
    {
        for (int l = 0; l < MG_NUM_CHANGES; l++) {
            uint32_t run = 0;
            for (int c = 0; c < C; c++) {
                coff[(size_t)l * C + c] = run;
                run += ccnt[(size_t)l * C + c];
            }
            clen[l] = C > 0 ? ccnt[(size_t)l * C + C - 1] + coff[(size_t)l * C + C - 1] : 0u;
        }
    }

    /* countingSortChanges (src/fluid_system_cuda.cu:406-618): per cell, its slots in order and
     * their bond slots in order; list 0 takes the pairs compute_bond_changes counted, lists
     * 1-8 the slots whose change word has the list's bit; at most the cell's count per list;
     * cells whose range ends beyond mActivePoints are skipped */
    void counting_sort_changes() override
    {
        std::fill(clists.begin(), clists.end(), MG_NONE);
        const uint32_t na = active_points;
        for (int c = 0; c < C; c++) {
            const uint32_t count = cnt[c];
            if (count == 0) continue;
            const uint32_t o = off[c];
            if (o + count > na) continue;
            uint32_t n[MG_NUM_CHANGES] = {};
            auto put = [&](int l, uint32_t s, uint32_t a) {
                if (n[l] >= ccnt[(size_t)l * C + c]) return;
                const uint32_t at = coff[(size_t)l * C + c] + n[l]++;
                if (at >= (uint32_t)N) return;
                clists[(size_t)l * 2 * N + at] = s;
                clists[(size_t)l * 2 * N + N + at] = a;
            };
            for (uint32_t s = o; s < o + count; s++)
                for (int a = 0; a < B; a++) {
                    if (hs[s] == (uint32_t)a) put(0, s, (uint32_t)a);
                    const uint32_t ch = bu[brec(s, a) + MG_B_CHANGE];
                    for (int l = 1; l < MG_NUM_CHANGES; l++)
                        if ((ch >> l) & 1u) put(l, s, (uint32_t)a);
                }
        }
    }

    /* heal (src/fluid_system_cuda.cu:2047-2107) over list 0: the position shifted by
     * -sqrt(rd2) / (4 + r) on each axis whose sign bit of the random word is set, find_bond
     * from there for the listed slot, and a found candidate bonded as elastin (makeBond). */
    void heal(uint64_t seed) override
    {
        const uint32_t na = active_points, len = std::min<uint32_t>(clen[0], (uint32_t)N);
        const R max_len = std::sqrt((R)p.rd2);
        for (uint32_t k = 0; k < len; k++) {
            const uint32_t i = clists[k];
            if (i >= na) continue;
            const uint32_t bond = clists[(size_t)N + k];
            if (bond >= (uint32_t)B) continue;
            const uint32_t gc = gcell[i];
            if (gc == MG_GRID_UNDEF) continue;
            const uint32_t r = heal_word(seed, pid[i]);
            V3<R> x = P(i);
            x.x = x.x + (max_len / (R)(4u + (r & 7u))) * (R(-1) * (R)(1u & (r >> 3)));
            x.y = x.y + (max_len / (R)(4u + ((r >> 4) & 7u))) * (R(-1) * (R)(1u & (r >> 7)));
            x.z = x.z + (max_len / (R)(4u + ((r >> 8) & 7u))) * (R(-1) * (R)(1u & (r >> 11)));
            const uint32_t to = find_bond(i, gc, x, (int)bond);
            if (to < (uint32_t)N) {
                make_bond(i, (int)bond, to, 0);
                bu[brec(i, (int)bond) + MG_B_CHANGE] = 0;
            }
        }
    }

    int nth() const
    {
#ifdef _OPENMP
        return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
        return 1;
#endif
    }
};

} // namespace

struct mgo_handle {
    Base* o;
    int N, B;
};

template <typename F> static int with_grid(mgo_handle* h, F fn)
{
    if (!h || !h->o->have_grid || !h->o->have_params) return MG_ERR_STATE;
    fn();
    return MG_OK;
}

extern "C" {

int mgo_create(mgo_handle** out, int max_points, int bonds_per_particle, int precision)
{
    if (!out || max_points <= 0 || bonds_per_particle < 0 || bonds_per_particle > 64) return MG_ERR_ARG;
    if (precision != 32 && precision != 64) return MG_ERR_ARG;
    mgo_handle* h = new mgo_handle;
    h->N = max_points;
    h->B = bonds_per_particle;
    if (precision == 32) h->o = new Oracle<float>(max_points, bonds_per_particle);
    else h->o = new Oracle<double>(max_points, bonds_per_particle);
    *out = h;
    return MG_OK;
}

int mgo_destroy(mgo_handle* h)
{
    if (!h) return MG_ERR_ARG;
    delete h->o;
    delete h;
    return MG_OK;
}

int mgo_set_params(mgo_handle* h, const MgParams* p)
{
    if (!h || !p) return MG_ERR_ARG;
    h->o->p = *p;
    h->o->have_params = true;
    return MG_OK;
}

int mgo_set_grid(mgo_handle* h, const MgGrid* g)
{
    if (!h || !g || g->total <= 0 || g->adj_cnt <= 0 || g->adj_cnt > MG_MAX_ADJ) return MG_ERR_ARG;
    h->o->g = *g;
    h->o->have_grid = true;
    if (auto* o = dynamic_cast<Oracle<float>*>(h->o)) o->set_grid_arrays();
    if (auto* o = dynamic_cast<Oracle<double>*>(h->o)) o->set_grid_arrays();
    return MG_OK;
}

int mgo_set_genome(mgo_handle* h, const MgGenome* g)
{
    if (!h || !g) return MG_ERR_ARG;
    h->o->gen = *g;
    return MG_OK;
}

int mgo_set_freeze(mgo_handle* h, int freeze)
{
    if (!h) return MG_ERR_ARG;
    h->o->p.freeze = freeze ? 1 : 0;
    return MG_OK;
}

int mgo_set_sph_kernels(mgo_handle* h, int mode)
{
    if (!h || (mode != MG_SPH_REFERENCE && mode != MG_SPH_POLY6_SPIKY)) return MG_ERR_ARG;
    h->o->sph = mode;
    return MG_OK;
}

int mgo_set_threads(mgo_handle* h, int nthreads)
{
    if (!h) return MG_ERR_ARG;
    h->o->nthreads = nthreads;
    return MG_OK;
}

int mgo_upload(mgo_handle* h, int buf_id, const void* src, size_t bytes)
{
    if (!h || (!src && bytes)) return MG_ERR_ARG;
    if (!h->o->have_grid && mg_bytes_per_particle(buf_id, h->B) == 0) return MG_ERR_STATE;
    return h->o->upload(buf_id, src, bytes);
}

int mgo_download(mgo_handle* h, int buf_id, void* dst, size_t bytes)
{
    if (!h || (!dst && bytes)) return MG_ERR_ARG;
    if (!h->o->have_grid && mg_bytes_per_particle(buf_id, h->B) == 0) return MG_ERR_STATE;
    return h->o->download(buf_id, dst, bytes);
}

int mgo_get_counts(mgo_handle* h, uint32_t* num_points, uint32_t* active_points)
{
    if (!h) return MG_ERR_ARG;
    if (num_points) *num_points = h->o->num_points;
    if (active_points) *active_points = h->o->active_points;
    return MG_OK;
}

int mgo_insert_particles(mgo_handle* h) { return with_grid(h, [&] { h->o->insert(); }); }
int mgo_prefix_sum_cells(mgo_handle* h) { return with_grid(h, [&] { h->o->prefix_sum(); }); }
int mgo_counting_sort_full(mgo_handle* h) { return with_grid(h, [&] { h->o->sort(); }); }
int mgo_compute_pressure(mgo_handle* h) { return with_grid(h, [&] { h->o->pressure(); }); }
int mgo_compute_force(mgo_handle* h) { return with_grid(h, [&] { h->o->force(); }); }
int mgo_advance(mgo_handle* h, float time, float dt, float ss)
{
    return with_grid(h, [&] { h->o->advance(time, dt, ss); });
}
int mgo_compute_diffusion(mgo_handle* h) { return with_grid(h, [&] { h->o->diffusion(); }); }
int mgo_compute_genes(mgo_handle* h) { return with_grid(h, [&] { h->o->genes(); }); }
int mgo_zero_vel(mgo_handle* h) { return with_grid(h, [&] { h->o->zero_vel(); }); }
int mgo_initialize_bonds(mgo_handle* h) { return with_grid(h, [&] { h->o->initialize_bonds(); }); }
int mgo_compute_bond_changes(mgo_handle* h, uint32_t steps)
{
    if (steps == 0) return MG_ERR_ARG;
    return with_grid(h, [&] { h->o->compute_bond_changes(steps); });
}
int mgo_prefix_sum_changes(mgo_handle* h, int zero_offsets)
{
    (void)zero_offsets;
    return with_grid(h, [&] { h->o->prefix_sum_changes(); });
}
int mgo_counting_sort_changes(mgo_handle* h) { return with_grid(h, [&] { h->o->counting_sort_changes(); }); }
int mgo_compute_particle_changes(mgo_handle* h, uint64_t seed)
{
    if (h && h->o)
        for (int l = 1; l < MG_NUM_CHANGES; l++)
            if (h->o->clen[l] != 0) return MG_ERR_UNSUPPORTED;
    return with_grid(h, [&] { h->o->heal(seed); });
}
float mgo_tanhf(float x) { return tanh_f(x); }
float mgo_expf(float x) { return exp_f(x); }
uint32_t mgo_heal_word(uint64_t seed, uint32_t pid) { return heal_word(seed, pid); }
float mgo_acosf(float x) { return acos_f(x); }

int mgo_physical_sort(mgo_handle* h)
{
    return with_grid(h, [&] {
        h->o->insert();
        h->o->prefix_sum();
        h->o->sort();
    });
}

int mgo_inner_physical_loop(mgo_handle* h, float time, float dt)
{
    return with_grid(h, [&] {
        h->o->pressure();
        h->o->force();
        h->o->advance(time, dt, h->o->p.psimscale);
    });
}

int mgo_gene_action(mgo_handle* h)
{
    return with_grid(h, [&] {
        h->o->diffusion();
        h->o->genes();
    });
}

int mgo_step(mgo_handle* h, int nsteps, float dt, unsigned flags)
{
    if (nsteps < 0) return MG_ERR_ARG;
    return with_grid(h, [&] {
        for (int s = 0; s < nsteps; s++) {
            h->o->pressure();
            h->o->force();
            if (flags & MG_STEP_SPECIAL) {
                /* actuation offset for a step starting at this time, evaluated as the library's
                 * host does (src/fluid_system_cuda.cu:3682) */
                const float t = (float)h->o->time;
                const float zoff = h->o->p.actuation_factor * sinf(t / (h->o->p.actuation_period * dt));
                h->o->advance_special(t, dt, h->o->p.psimscale, zoff);
            } else {
                h->o->advance((float)h->o->time, dt, h->o->p.psimscale);
            }
            h->o->time += dt;
            if (flags & MG_STEP_GENES) {
                h->o->diffusion();
                h->o->genes();
            }
            h->o->insert();
            h->o->prefix_sum();
            h->o->sort();
        }
    });
}

} // extern "C"
