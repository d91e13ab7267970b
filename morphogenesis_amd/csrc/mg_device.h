/*
 * mg_device.h — device-side state and arithmetic shared by the step kernels.
 *
 * HBM layout (internal; the reference layout exists only at mg_upload/mg_download):
 *   sorted, ping-ponged per step (index = current sorted slot s):
 *     P[s]    float4 (pos.xyz, pressure)        V[s]   float4 (veval.xyz, 1/density)
 *     VEL[s]  float4 (vel.xyz, 0)               CONC[s][16] float
 *     HOME[s] creation slot of the occupant     GCELL[s], GNDX[s], GMASK[s] (gene bits)
 *   home-indexed (index = creation slot h, never moved by the sort):
 *     SLOT[h] current sorted slot or MG_NONE    EPIGEN[g][h]
 *     BPARTNER[a][h] partner's home (slot-major) BPAR[a][h] float4 (limit, rest, modulus, damping)
 *     BINT[a][h] float2 (strain^2, strain)      BCOLD[a][h] uint2 (partner ID, change bits)
 *     PIDX[a][h] uint2 (incoming owner's home, owner slot), AGE/COLOR/PID/MRAD/NERVE[h]
 *     INC_OFF[h]..INC_OFF[h+1]: incoming bonds (owner_home*B + slot), ascending.
 */
#ifndef MG_DEVICE_H
#define MG_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

#include "../../include/mg_types.h"

/* device counters */
enum {
    MG_C_NUM_POINTS = 0,  /* mNumPoints of the last sort                              */
    MG_C_ACTIVE = 1,      /* mActivePoints of the last sort                           */
    MG_C_TAIL = 2,        /* slots >= this hold the sort's reset tail                 */
    MG_C_HWM = 3,         /* FBIN[s] == s for s below this                            */
    MG_C_LISTS = 4,       /* 1 once a sort has built the gene membership              */
    MG_C_SCAN_NUM = 5,    /* FBIN_OFFSET[C-1] of the last scan (committed by the sort) */
    MG_C_SCAN_ACTIVE = 6, /* FBIN_OFFSET[grid_ScanMax] of the last scan               */
    MG_C_ERR = 8,         /* nonzero: a kernel gave up a bounded wait (result invalid)  */
    MG_C_SKIN = 9,        /* nonzero: a particle of this fused step moved farther than
                           * the hit masks' skin allows (the diffusion sweep tests its
                           * candidates itself)                                        */
    MG_C_COUNT = 16
};

/* Hit-mask skin. The pressure sweep records the candidates with dsq < MG_SKIN_LIM, i.e.
 * within 1 + 0.01 of the particle, while the cut-off of the pressure, force and diffusion
 * sums is 1. The pressure and force sweeps run on the same positions; the diffusion sweep
 * runs after the advance, and a pair outside the skin can only come within the cut-off if
 * the two particles moved 0.01 together. The fused force-advance raises MG_C_SKIN when any
 * particle moves by more than sqrt(MG_SKIN_DISP2) = 0.004 in the step (two of them: 0.008,
 * leaving 0.002 for rounding, far above the float rounding of positions and squared
 * distances), and the diffusion sweep then tests its candidates instead of using the masks. */
#define MG_SKIN_LIM 1.0201f
#define MG_SKIN_DISP2 1.6e-5f

struct DS {
    /* sorted, current */
    float4 *P, *V, *VEL, *F;
    float* CONC;
    uint32_t *HOME, *GCELL, *GNDX, *GMASK;
    /* sorted, other half of the ping-pong */
    float4 *P2, *V2, *VEL2;
    float* CONC2;
    uint32_t *HOME2, *GCELL2, *GNDX2, *GMASK2;
    /* home-indexed */
    uint32_t *SLOT, *EPIGEN, *BPARTNER;
    float4* BPAR;
    float2* BINT;
    uint32_t *INC_OFF, *INC;
    uint64_t* LIVE; /* per home: bit a set when bond slot a holds a partner (B <= 64) */
    /* grid (current) and the grid the fused step bins into */
    uint32_t *CNT, *OFF, *GCNT, *GOFF, *PERM, *LISTS, *LENS;
    uint32_t *CNTN, *OFFN;
    /* post-gene membership recorded for the coming sort (fused step), or null */
    uint32_t* GMN;
    /* cleared by k_pressure when set (zero_n words) */
    uint32_t* ZERO;
    uint32_t zero_n;
    /* 1 when OFF is the exclusive prefix sum of CNT (the sorted state) and the adjacency
     * list is made of rows of g.srch consecutive cells: sweeps then walk one slot range per
     * row. 0 selects the per-cell walk (any CNT / OFF contents). */
    int rows;
    /* 1 when the slots are sorted by cell (CNT / OFF / GCELL of the last sort) and the
     * positions have not changed since: sweeps may skip cells out of reach */
    int cull;
    /* learned coarse step: per slot two float4 corrections (drift c_d with the use flag in w,
     * veval c_e) applied by the advance; null on plain steps */
    const float4* DVc;
    /* Hit masks handed from the pressure sweep to the force sweep and the diffusion sweep of
     * the same step: row s of slot i at [s * N + i], bit k = slot OFF[c0] + k is within the
     * skin (dsq < MG_SKIN_LIM at the pressure sweep's positions); ~0 = the row's window
     * exceeds 64 slots (the reader tests that row itself). Readers test dsq against their own
     * cut-off for every recorded candidate. Null when not in use. */
    uint64_t* HMW;
    const uint64_t* HMR;
    /* Bond records handed from the pressure pass to the force pass of the same positions (the
     * hit masks' rule): for live bond slot a of home h, at [a * N + h], the owner's bond term
     * (eterm.xyz) and its strain in w, with the sign bit of w set when the bond breaks this
     * step (abs_dist >= elastic limit); w is NaN for a slot the force skips (partner inactive
     * or rest length below 1e-9). The force adds its own slots' records and those of its
     * incoming bonds instead of evaluating each bond twice. Null when not in use. */
    float4* ETW;
    const float4* ETR;
    /* 1: a force pass that changes no bond record (no integrators, no breaking), for the
     * learned step's training inputs */
    int dry;
    uint32_t* ctr;
    int N, B, C;
};

/* ---------------------------------------------------------------- arithmetic
 * The expression trees below follow the reference kernels operation by operation; the
 * oracle (oracle/mg_oracle.cpp) evaluates the same trees. Compiled with
 * -ffp-contract=off and correctly rounded division/sqrt, so FP32 results are the
 * IEEE results of these trees. */

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mk3(a.x * s, a.y * s, a.z * s); }
/* src/cutil_math.h:389-393: float3 / float multiplies by the reciprocal */
__device__ __forceinline__ f3 divs(f3 a, float s)
{
    float inv = 1.0f / s;
    return a * inv;
}
/* Correctly rounded sqrtf and 1/x in their short forms. With IEEE division and square root
 * (-fhip-fp32-correctly-rounded-divide-sqrt) the compiler expands sqrtf(x) to: scale x < 2^-96
 * by 2^32; v_sqrt_f32; step the result one ulp down if fma(-(s - ulp), s, x) <= 0, one ulp up if
 * fma(-(s + ulp), s, x) > 0; unscale; pass +-0 and +inf through. On [2^-96, +inf] the scaling and
 * the pass-through do nothing, so the middle steps alone give the same bits. 1/x expands to
 * v_div_scale, v_rcp_f32, a refinement by FMAs, v_div_fmas and v_div_fixup; for
 * 2^-64 <= |x| <= 2^64 the scalings are the identity (numerator 1, no denormal reciprocal or
 * quotient), v_div_fmas is a plain fma and v_div_fixup returns its input, so the FMAs alone give
 * the same bits; tests/test_gpu_rounding.py compares every float of both ranges with sqrtf and
 * 1.0f / x on the GPU. */
__device__ __forceinline__ float sqrt_rn_short(float x) /* 2^-96 <= x <= +inf */
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    float r = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : r;
}
__device__ __forceinline__ float rcp_rn_short(float b) /* 2^-64 <= |b| <= 2^64 */
{
    float r = __builtin_amdgcn_rcpf(b);
    r = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
    float q = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
    return __builtin_fmaf(__builtin_fmaf(-b, q, 1.0f), r, q);
}
/* true when every active lane's x lies in [lo, hi] (wave-uniform): the short forms apply */
__device__ __forceinline__ bool wave_all_in(float x, float lo, float hi)
{
    return __ballot(!(x >= lo && x <= hi)) == 0ull;
}
/* r = sqrtf(dsq), inv = 1.0f / r, for hits with dsq < 2^96 */
__device__ __forceinline__ void sqrt_rcp_rn(float dsq, float& r, float& inv)
{
    if (wave_all_in(dsq, 0x1p-96f, 0x1p96f)) {
        r = sqrt_rn_short(dsq);
        inv = rcp_rn_short(r);
    } else {
        r = sqrtf(dsq);
        inv = 1.0f / r;
    }
}
__device__ __forceinline__ float sqrt_rn(float x)
{
    return wave_all_in(x, 0x1p-96f, __builtin_inff()) ? sqrt_rn_short(x) : sqrtf(x);
}
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ bool has_nan3(f3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

/* One elastic bond seen from its owner o towards partner p
 * (src/fluid_system_cuda.cu:3263-3314). Returns the acceleration term eterm and the
 * strain; `apply` is false when eterm has a NaN component (the reference skips it). */
__device__ __forceinline__ f3 bond_term(f3 opos, f3 oveval, f3 ppos, f3 pveval, float4 bp,
                                        float pmass, float* strain_out, float* abs_dist_out)
{
    const float elastic_limit = bp.x, restlength = bp.y, modulus = bp.z, damping_coeff = bp.w;
    f3 dist = opos - ppos;
    float dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
    float abs_dist, inv_dist;
    if (wave_all_in(dsq, 0x1p-96f, 0x1p96f)) {
        abs_dist = sqrt_rn_short(dsq) + FLT_MIN;
        inv_dist = rcp_rn_short(abs_dist);
    } else {
        abs_dist = sqrtf(dsq) + FLT_MIN;
        inv_dist = 1.0f / abs_dist;
    }
    f3 rel_vel = pveval - oveval;
    float spring_strain = fmaxf(0.0f, (abs_dist - restlength) / restlength);
    float fac = (float)(abs_dist < elastic_limit ? 1 : 0);
    f3 e1 = dist * inv_dist * spring_strain * modulus;
    f3 e2 = damping_coeff * rel_vel;
    f3 eterm = divs(fac * (e1 - e2), pmass);
    *strain_out = spring_strain;
    *abs_dist_out = abs_dist;
    return eterm;
}

/* Fluid pair term for one neighbour within dsq < 1 (src/fluid_system_cuda.cu:3203-3221). */
/* The selectable Mueller (2003) kernels (MG_SPH_POLY6_SPIKY): the pair force in the arrangement
 * of Fluids v3, which the reference keeps in comments (src/fluid_system_cuda.cu:3140-3160,
 * 3213-3220) with the constants FluidParamCL computes (src/fluid_system.cpp:568-569, 587):
 * sdist = sqrt(dsq d2), c = sr - sdist,
 * pterm = simscale (-1/2) c spikykern (p_i + p_j) / sdist (spikykern = -45 / (pi sr^6)),
 * (pterm dist + vterm (v_j - v_i)) c (1/rho_i) (1/rho_j), vterm = lapkern * visc: the
 * pressure term -(p_i + p_j) / 2 grad W_spiky and the viscosity term mu (v_j - v_i)
 * lap W_visc, both over rho_i rho_j (times pmass where the per-cell sums are added). */
__device__ __forceinline__ f3 fluid_term_poly6(f3 dist, float dsq, const MgParams& p, float ipress, float jpress,
                                               float idens, float jdens, f3 ivev, f3 jvev)
{
    const float sdist = sqrtf(dsq * p.d2);
    const float c = p.psmoothradius - sdist;
    const float pterm = p.psimscale * -0.5f * c * p.spikykern * (ipress + jpress) / sdist;
    const float dterm = c * idens * jdens;
    return (dist * pterm + (jvev - ivev) * p.vterm) * dterm;
}

__device__ __forceinline__ f3 fluid_term(f3 dist, float r, float inv_r, float sr, float ipress,
                                         float jpress, float idens, float jdens, f3 ivev, f3 jvev,
                                         float surface_t)
{
    float t = sr - r;
    float kern = t * t * t;
    float press = 100.0f * (ipress + jpress);
    float s = idens * jdens * 100.0f;
    f3 pterm = (s * (dist * inv_r)) * (press * kern - surface_t * (t * t)); /* dist / r: inv_r = 1 / r */
    f3 delta_v = jvev - ivev;
    f3 vterm = (100000.0f * delta_v) * kern;
    return pterm + vterm;
}

/* acos in float by the Cephes single-precision method (S. L. Moshier, Cephes Math Library,
 * asinf.c / acosf.c): |x| <= 0.5 -> pi/2 - asin(x); |x| > 0.5 -> 2 asin(sqrt((1 - |x|) / 2)),
 * mirrored for x < 0; asin(a) for a <= 0.5 by its degree-11 odd minimax polynomial
 * (published coefficients). Relative error a few 1e-7. NaN outside [-1, 1]. The oracle evaluates
 * the same operations (oracle/mg_oracle.cpp, mg_acosf), so both sides produce the same bits. */
__host__ __device__ __forceinline__ float mg_asinf_half(float a) /* 0 <= a <= 0.5 */
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
__host__ __device__ __forceinline__ float mg_acosf(float x)
{
    const float pio2 = 1.5707963267948966f, pi = 3.141592653589793f;
    if (!(x >= -1.0f && x <= 1.0f)) return __builtin_nanf("");
    if (x < -0.5f) return pi - 2.0f * mg_asinf_half(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * mg_asinf_half(sqrtf(0.5f * (1.0f - x)));
    return x < 0.0f ? pio2 + mg_asinf_half(-x) : pio2 - mg_asinf_half(x);
}

/* exp and tanh in float by the Cephes single-precision method (S. L. Moshier, Cephes Math
 * Library, expf.c / tanhf.c), every operation in float:
 *   exp(x) = 2^n exp(r), n = floor(x log2(e) + 1/2), r = x - n ln2 with ln2 split in two
 *   parts (0.693359375 and -2.12194440e-4), exp(r) by the published degree-5 polynomial;
 *   tanh(x) = sign(x) (1 - 2 / (exp(2|x|) + 1)) for |x| >= 0.625, +-1 above 44.36, and the
 *   published odd polynomial below 0.625; evaluated on |x| with the sign copied back, so
 *   tanh(-x) = -tanh(x) exactly, signed zeros included.
 * Max error about 1.2 ulp against the exact function. Used by the bond remodelling
 * (the reference calls CUDA's tanh there); the oracle evaluates the same operations. */
__host__ __device__ __forceinline__ float mg_expf(float x)
{
    if (x > 88.72283905206835f) return __builtin_inff();
    if (x < -103.278929903431851103f) return 0.0f;
    const float z = floorf(1.44269504088896341f * x + 0.5f);
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
    return ldexpf(p * x2 + x + 1.0f, n);
}
__host__ __device__ __forceinline__ float mg_tanhf(float x)
{
    const float a = fabsf(x);
    float r;
    if (a > 44.361419526034175f) {
        r = 1.0f;
    } else if (a >= 0.625f) {
        r = 1.0f - 2.0f / (mg_expf(a + a) + 1.0f);
    } else {
        const float s = a * a;
        float p = -5.70498872745e-3f;
        p = p * s + 2.06390887954e-2f;
        p = p * s + -5.37397155531e-2f;
        p = p * s + 1.33314422036e-1f;
        p = p * s + -3.33332819422e-1f;
        r = p * s * a + a;
    }
    return copysignf(r, x);
}

/* The random word heal draws for a particle: the SplitMix64 output function (G. Steele,
 * D. Lea, C. Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) of
 * seed + golden * (particle ID + 1), low 32 bits. The reference draws from per-particle
 * curand states instead; this word depends only on (seed, ID), not on the launch. */
__host__ __device__ __forceinline__ uint32_t mg_heal_word(uint64_t seed, uint32_t pid)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)pid + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

/* XCD-aware block index: the dispatcher places consecutive workgroups on different XCDs
 * (round robin over 8); remap so each XCD walks one contiguous range of particles and
 * neighbour data stays in that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t per = nb / 8u, rem = nb % 8u;
    const uint32_t x = b % 8u, k = b / 8u;
    /* XCD x owns per (+1 if x < rem) consecutive logical blocks */
    return x * per + (x < rem ? x : rem) + k;
}

#endif
