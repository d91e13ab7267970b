/*
 * mg_mlp.hip — the learned coarse-step correction (no counterpart in the reference).
 *
 * A per-particle MLP 64 -> 128 -> 128 -> 16 (SiLU hidden layers; 53 real inputs, 6 real
 * outputs, the rest zero padding to MFMA tile sizes) predicts two velocity corrections for
 * one coarse step of 4 dt, each the fine segment's value minus the plain coarse step's:
 *   drift c_d: the position moves by (v_next + c_d) 4 dt, as four fine steps move it;
 *   veval c_e: veval is the plain step's (v + v_next) / 2 + c_e, the fine segment's last.
 * The plain coarse step overshoots four fine steps by about 6 dt^2 a per step in position,
 * because the leapfrog drift uses the end-of-step velocity, and its veval lags by about
 * 1.5 dt a; the stored velocity itself agrees with the fine one to a difference that the
 * local state does not predict, so it keeps the plain step's value.
 * Inputs, from the particle's sorted state after the step's pressure and force passes:
 *   8 bonded neighbours (live own bond slots in slot order, then incoming bonds in CSR
 *   order): (x_j - x_i) and (v_j - v_i), 48 values; density and pressure; the coarse step's
 *   acceleration (FFORCE with walls, gravity and the acceleration limit, as the advance
 *   applies it), 3 values; zeros to 64.
 * Each input is normalised as (x - mu) * inv_sigma; output c is y_c * out_scale[c].
 *
 * Weights are bf16, accumulation fp32, on gfx950 MFMA (v_mfma_f32_16x16x32_bf16): for a
 * 16 x 32 A tile lane l holds A[l % 16][8 (l / 16) + e], for a 32 x 16 B tile lane l holds
 * B[8 (l / 16) + e][l % 16] (e = 0..7), and the 16 x 16 fp32 result lane l holds
 * D[4 (l / 16) + r][l % 16] (r = 0..3).
 *
 * Training uses a generic bf16 GEMM with split-K partials reduced in a fixed order, and
 * fp32 Adam on fp32 master weights; inference is one fused kernel per step.
 */

#define MLP_IN 64
#define MLP_IN_REAL 53
/* a correction that would move a particle farther than this in one step (world units, a
 * thousandth of the unit smoothing radius) is refused: the particle takes the plain step */
#define MLP_MAX_SHIFT 1e-3f
#define MLP_H 128
#define MLP_OUT 16
#define MLP_OUT_REAL 6
#define MLP_NBR 8

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

__device__ __forceinline__ f32x4_t mfma16(bf16x8_t a, bf16x8_t b, f32x4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

/* sigmoid with the hardware exp2 and reciprocal (a few ulp; the network runs in bf16). The
 * same function serves training and inference. */
__device__ __forceinline__ float sigmoid_fast(float z)
{
    const float e = __builtin_amdgcn_exp2f(-1.44269504f * z);
    return __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }

/* d silu / dz = s (1 + z (1 - s)), s = sigmoid(z) */
__device__ __forceinline__ float silu_grad(float z)
{
    const float s = sigmoid_fast(z);
    return s * (1.0f + z * (1.0f - s));
}

/* Everything the inference kernel reads, device resident. Weight fragments are laid out so
 * that each lane's 8 bf16 of a B operand are contiguous:
 *   W?f[ntile][kchunk][col 0..15][k 0..31] = W[kchunk * 32 + k][ntile * 16 + col]. */
struct MlpDev {
    __bf16 w1f[(MLP_IN / 32) * (MLP_H / 16) * 16 * 32];
    __bf16 w2f[(MLP_H / 32) * (MLP_H / 16) * 16 * 32];
    __bf16 w3f[(MLP_H / 32) * (MLP_OUT / 16) * 16 * 32];
    float b1[MLP_H], b2[MLP_H], b3[MLP_OUT];
    float mu[MLP_IN], inv_sigma[MLP_IN];
    float out_scale[MLP_OUT];
    float in_limit;  /* |normalised input| above this -> fallback (plain coarse step) */
    float out_limit; /* |y| above this -> fallback */
};

/* ------------------------------------------------------------------ features
 * The 64 raw inputs of active particle i (sorted state after the pressure and force passes;
 * FFORCE in d.F). Neighbours: the live own bond slots in slot order, then the incoming bonds
 * in CSR order, skipping partners outside the active range and broken bonds (rest length
 * below 1e-9), up to MLP_NBR. Candidates are examined MLP_NBR at a time, each group's
 * records loaded together (partner, rest length, then slot), and the chosen neighbours'
 * positions and velocities loaded together at the end. */
__device__ __forceinline__ void mlp_features(const DS& d, const MgParams& p, uint32_t i, uint32_t na,
                                             float x[MLP_IN])
{
    const int N = d.N, B = d.B;
    const float* BPARf = (const float*)d.BPAR;
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) x[k] = 0.0f;
    const uint32_t h = d.HOME[i];
    const float4 Pi = d.P[i], Vi = d.VEL[i];
    uint32_t sel[MLP_NBR];
    int m = 0;
    auto push = [&](uint32_t j) {
#pragma unroll
        for (int q = 0; q < MLP_NBR; q++)
            if (q == m) sel[q] = j;
        m++;
    };
    /* own live slots (LIVE: partner present; B <= 64) */
    uint64_t live = d.LIVE[h];
    while (live != 0ull && m < MLP_NBR) {
        uint32_t a[MLP_NBR], jh[MLP_NBR], jj[MLP_NBR];
        float rest[MLP_NBR];
        bool has[MLP_NBR];
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++) {
            has[k] = live != 0ull;
            a[k] = has[k] ? (uint32_t)__builtin_ctzll(live) : 0u;
            live &= live - 1ull;
        }
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++) {
            jh[k] = MG_NONE;
            rest[k] = 0.0f;
            if (has[k]) {
                const size_t bi = (size_t)a[k] * N + h;
                jh[k] = d.BPARTNER[bi];
                rest[k] = BPARf[bi * 4 + 1];
            }
        }
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++) jj[k] = (has[k] && jh[k] < (uint32_t)N) ? d.SLOT[jh[k]] : MG_NONE;
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++)
            if (m < MLP_NBR && jj[k] < na && !((double)rest[k] < 0.000000001)) push(jj[k]);
    }
    /* incoming bonds */
    const uint32_t e0 = d.INC_OFF[h], e1 = d.INC_OFF[h + 1];
    for (uint32_t e = e0; e < e1 && m < MLP_NBR; e += MLP_NBR) {
        uint32_t kh[MLP_NBR], kk[MLP_NBR];
        float rest[MLP_NBR];
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++) {
            kh[k] = MG_NONE;
            rest[k] = 0.0f;
            if (e + (uint32_t)k < e1) {
                const uint32_t packed = d.INC[e + (uint32_t)k];
                kh[k] = packed / (uint32_t)B;
                const uint32_t b = packed - kh[k] * (uint32_t)B;
                rest[k] = BPARf[((size_t)b * N + kh[k]) * 4 + 1];
            }
        }
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++) kk[k] = kh[k] < (uint32_t)N ? d.SLOT[kh[k]] : MG_NONE;
#pragma unroll
        for (int k = 0; k < MLP_NBR; k++)
            if (m < MLP_NBR && kk[k] < na && !((double)rest[k] < 0.000000001)) push(kk[k]);
    }
#pragma unroll
    for (int q = 0; q < MLP_NBR; q++) {
        if (q < m) {
            const float4 Pj = d.P[sel[q]], Vj = d.VEL[sel[q]];
            x[6 * q + 0] = Pj.x - Pi.x;
            x[6 * q + 1] = Pj.y - Pi.y;
            x[6 * q + 2] = Pj.z - Pi.z;
            x[6 * q + 3] = Vj.x - Vi.x;
            x[6 * q + 4] = Vj.y - Vi.y;
            x[6 * q + 5] = Vj.z - Vi.z;
        }
    }
    const float4 Vi_eval = d.V[i];
    x[48] = 1.0f / Vi_eval.w; /* density */
    x[49] = Pi.w;             /* pressure */
    const f3 a = advance_accel(p, xyz(Pi), xyz(Vi_eval), xyz(d.F[i]));
    x[50] = a.x;
    x[51] = a.y;
    x[52] = a.z;
}

/* raw features of every active slot (rows of inactive slots are zero), for training data */
__global__ LAUNCH_BOUNDS_256 void k_mlp_features(DS d, MgParams p, float* __restrict__ X)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)d.N) return;
    float x[MLP_IN];
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    if (i < na && d.GCELL[i] != MG_GRID_UNDEF) {
        mlp_features(d, p, i, na, x);
    } else {
#pragma unroll
        for (int k = 0; k < MLP_IN; k++) x[k] = 0.0f;
    }
    float4* o = (float4*)(X + (size_t)i * MLP_IN);
#pragma unroll
    for (int k = 0; k < MLP_IN / 4; k++) o[k] = make_float4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
}

/* --------------------------------------------------------------- inference
 * Two kernels per learned step.
 *
 * k_mlp_gather, one thread per slot: the particle's inputs, normalised and rounded to bf16,
 * into X[i][0..63] (128 bytes), and flags[i] = active | (every normalised input within
 * in_limit) << 1.
 *
 * k_mlp_apply, one wavefront per group of 32 particles (two 16-particle tiles): the three
 * layers as transposed products, H^T = W^T X^T, so the weights are the MFMA A operand and the
 * particles the B operand. The 16 x 16 result of a layer leaves lane l holding neurons
 * 16 mt + 4 (l / 16) + r of particle l % 16, which is a B operand of the next layer for the
 * K order nu(32 kc + 8 q + e) = 32 kc + (e < 4 ? 4 q + e : 16 + 4 q + e - 4) (q = l / 16);
 * W2 and W3 are packed with their input rows in that order (k_pack_frag with perm), so the
 * activations never leave registers. Weight fragments and biases are staged in LDS once per
 * block. DV[2i] = (c_d, 1), DV[2i + 1] = (c_e, 0) for corrected particles, zeros for inactive
 * or fallback ones (a normalised input beyond in_limit, an output beyond out_limit, or a
 * drift correction that would move the particle more than MLP_MAX_SHIFT in the step of
 * length dt); fallbacks of active particles are counted in *fallbacks. */
__global__ LAUNCH_BOUNDS_256 void k_mlp_gather(DS d, MgParams p, const MlpDev* __restrict__ net,
                                               uint4* __restrict__ X, uint32_t* __restrict__ flags)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)d.N) return;
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    const bool act = i < na && d.GCELL[i] != MG_GRID_UNDEF;
    float x[MLP_IN];
    if (act) {
        mlp_features(d, p, i, na, x);
    } else {
#pragma unroll
        for (int k = 0; k < MLP_IN; k++) x[k] = 0.0f;
    }
    bool ok = act;
    uint32_t w[MLP_IN / 2];
#pragma unroll
    for (int k = 0; k < MLP_IN; k += 2) {
        const float v0 = act ? (x[k] - net->mu[k]) * net->inv_sigma[k] : 0.0f;
        const float v1 = act ? (x[k + 1] - net->mu[k + 1]) * net->inv_sigma[k + 1] : 0.0f;
        if (!(fabsf(v0) <= net->in_limit) || !(fabsf(v1) <= net->in_limit)) ok = false;
        const __bf16 b0 = to_bf16(v0), b1 = to_bf16(v1);
        w[k / 2] = (uint32_t)__builtin_bit_cast(uint16_t, b0) | ((uint32_t)__builtin_bit_cast(uint16_t, b1) << 16);
    }
    uint4* o = X + (size_t)i * (MLP_IN / 8);
#pragma unroll
    for (int k = 0; k < MLP_IN / 8; k++) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
    flags[i] = (act ? 1u : 0u) | (ok ? 2u : 0u);
}

#define MLP_WAVES 4
struct MlpLds {
    __bf16 w1f[(MLP_IN / 32) * (MLP_H / 16) * 16 * 32];
    __bf16 w2f[(MLP_H / 32) * (MLP_H / 16) * 16 * 32];
    __bf16 w3f[(MLP_H / 32) * (MLP_OUT / 16) * 16 * 32];
    float b1[MLP_H], b2[MLP_H], b3[MLP_OUT], out_scale[MLP_OUT];
};

__device__ __forceinline__ bf16x8_t lds8(const __bf16* p) { return *(const bf16x8_t*)p; }

/* bias + SiLU of a layer's results (neuron 16 mt + 4 q + r in acc[mt][r]) packed as the next
 * layer's B operands: op[kc] holds neurons nu(32 kc + 8 q + e), e = 0..7 */
__device__ __forceinline__ void silu_pack(const f32x4_t acc[MLP_H / 16], const float* bias, int q,
                                          bf16x8_t op[MLP_H / 32])
{
#pragma unroll
    for (int kc = 0; kc < MLP_H / 32; kc++) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int mt = 2 * kc + (e >> 2), r = e & 3;
            op[kc][e] = to_bf16(silu(acc[mt][r] + bias[16 * mt + 4 * q + r]));
        }
    }
}

__global__ __launch_bounds__(256) void k_mlp_apply(const MlpDev* __restrict__ net, const uint4* __restrict__ X,
                                                   const uint32_t* __restrict__ flags, float4* __restrict__ DV,
                                                   uint32_t* __restrict__ fallbacks, float dt, uint32_t N)
{
    extern __shared__ float4 mg_lds[];
    MlpLds& L = *(MlpLds*)mg_lds;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    {
        const uint4* s1 = (const uint4*)net->w1f;
        uint4* d1 = (uint4*)L.w1f;
        for (int k = t; k < (int)(sizeof(L.w1f) / 16); k += 256) d1[k] = s1[k];
        const uint4* s2 = (const uint4*)net->w2f;
        uint4* d2 = (uint4*)L.w2f;
        for (int k = t; k < (int)(sizeof(L.w2f) / 16); k += 256) d2[k] = s2[k];
        const uint4* s3 = (const uint4*)net->w3f;
        uint4* d3 = (uint4*)L.w3f;
        for (int k = t; k < (int)(sizeof(L.w3f) / 16); k += 256) d3[k] = s3[k];
        for (int k = t; k < MLP_H; k += 256) {
            L.b1[k] = net->b1[k];
            L.b2[k] = net->b2[k];
        }
        if (t < MLP_OUT) {
            L.b3[t] = net->b3[t];
            L.out_scale[t] = net->out_scale[t];
        }
    }
    __syncthreads();
    const int q = l >> 4, col = l & 15;
    const uint32_t ngroups = (N + 31u) / 32u;
    uint32_t nfall = 0;
    for (uint32_t grp = blockIdx.x * MLP_WAVES + (uint32_t)w; grp < ngroups; grp += gridDim.x * MLP_WAVES) {
        /* layer-1 B operands: particle col of tile tt, inputs 32 kc + 8 q .. + 7 */
        bf16x8_t xb[2][MLP_IN / 32];
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
            const uint32_t i = grp * 32u + (uint32_t)(16 * tt + col);
#pragma unroll
            for (int kc = 0; kc < MLP_IN / 32; kc++) {
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (i < N) v = X[(size_t)i * (MLP_IN / 8) + (size_t)(4 * kc + q)];
                xb[tt][kc] = __builtin_bit_cast(bf16x8_t, v);
            }
        }
        /* layer 1: [128 x 64] x [64 x 16] per tile */
        f32x4_t acc[2][MLP_H / 16];
#pragma unroll
        for (int mt = 0; mt < MLP_H / 16; mt++) {
            acc[0][mt] = acc[1][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < MLP_IN / 32; kc++) {
                const bf16x8_t a = lds8(&L.w1f[((mt * (MLP_IN / 32) + kc) * 16 + col) * 32 + 8 * q]);
                acc[0][mt] = mfma16(a, xb[0][kc], acc[0][mt]);
                acc[1][mt] = mfma16(a, xb[1][kc], acc[1][mt]);
            }
            __builtin_amdgcn_sched_barrier(0); /* one tile's weight fragments live at a time */
        }
        bf16x8_t hb[2][MLP_H / 32];
        silu_pack(acc[0], L.b1, q, hb[0]);
        silu_pack(acc[1], L.b1, q, hb[1]);
        /* layer 2: [128 x 128] x [128 x 16] per tile */
#pragma unroll
        for (int mt = 0; mt < MLP_H / 16; mt++) {
            acc[0][mt] = acc[1][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < MLP_H / 32; kc++) {
                const bf16x8_t a = lds8(&L.w2f[((mt * (MLP_H / 32) + kc) * 16 + col) * 32 + 8 * q]);
                acc[0][mt] = mfma16(a, hb[0][kc], acc[0][mt]);
                acc[1][mt] = mfma16(a, hb[1][kc], acc[1][mt]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        silu_pack(acc[0], L.b2, q, hb[0]);
        silu_pack(acc[1], L.b2, q, hb[1]);
        /* layer 3: [16 x 128] x [128 x 16] per tile */
        f32x4_t y[2] = {f32x4_t{0.f, 0.f, 0.f, 0.f}, f32x4_t{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int kc = 0; kc < MLP_H / 32; kc++) {
            const bf16x8_t a = lds8(&L.w3f[(kc * 16 + col) * 32 + 8 * q]);
            y[0] = mfma16(a, hb[0][kc], y[0]);
            y[1] = mfma16(a, hb[1][kc], y[1]);
        }
        /* lane l holds outputs 4 q + r of particle col; lanes q = 0 take outputs 4 and 5 from
         * lane l + 16 and write the particle's corrections */
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
            float o[4];
            bool lim = true;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int c = 4 * q + r;
                const float v = y[tt][r] + L.b3[c];
                if (c < MLP_OUT_REAL && !(fabsf(v) <= net->out_limit)) lim = false;
                o[r] = v * L.out_scale[c];
            }
            const float o4 = __shfl(o[0], l + 16), o5 = __shfl(o[1], l + 16);
            const bool lim_hi = __shfl((int)lim, l + 16) != 0;
            if (q == 0) {
                const uint32_t i = grp * 32u + (uint32_t)(16 * tt + col);
                if (i < N) {
                    const uint32_t fl = flags[i];
                    const f3 cd = mk3(o[0], o[1], o[2]), ce = mk3(o[3], o4, o5);
                    /* the residual the network may apply: a shift of at most MLP_MAX_SHIFT */
                    const bool use = (fl & 2u) != 0u && lim && lim_hi && sqrtf(dot3(cd, cd)) * dt <= MLP_MAX_SHIFT;
                    if (use) {
                        DV[2 * (size_t)i] = make_float4(cd.x, cd.y, cd.z, 1.0f);
                        DV[2 * (size_t)i + 1] = make_float4(ce.x, ce.y, ce.z, 0.0f);
                    } else {
                        DV[2 * (size_t)i] = make_float4(0.f, 0.f, 0.f, 0.f);
                        DV[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
                    }
                    if ((fl & 1u) && !use) nfall++;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) nfall += (uint32_t)__shfl_xor((int)nfall, o);
    if (l == 0 && nfall) atomicAdd(fallbacks, nfall);
}

/* ------------------------------------------------------------ training GEMM
 * C[s][m][n] = sum over k in split s of A(m, k) * B(k, n), A(m, k) = A[m * sam + k * sak],
 * B(k, n) = B[k * sbk + n * sbn] (bf16), fp32 partials per split; k_gemm_reduce adds the
 * splits in order (deterministic). 64 x 64 block tile, four 32 x 32 wave tiles, K step 32. */
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const __bf16* __restrict__ A, long sam,
                                                   long sak, const __bf16* __restrict__ Bm, long sbk, long sbn,
                                                   float* __restrict__ C, int kper)
{
    __shared__ __bf16 As[64][32 + 8];
    __shared__ __bf16 Bs[64][32 + 8]; /* transposed: Bs[n][k] */
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int k_lo = blockIdx.z * kper, k_hi = min(K, k_lo + kper);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    /* each thread stages one 8-element chunk of A and one of B per K step: 8 consecutive k of
     * a row when the operand is contiguous in k, else 8 consecutive m (n) of a k when it is
     * contiguous in m (n), with one 16-byte load when the chunk is in range and aligned */
    const bool a_k = sak == 1, a_m = !a_k && sam == 1;
    const bool b_k = sbk == 1, b_n = !b_k && sbn == 1;
    auto aligned = [](const __bf16* q) { return ((uintptr_t)q & 15u) == 0u; };
    for (int k0 = k_lo; k0 < k_hi; k0 += 32) {
        if (a_k || a_m) {
            const int r = a_k ? t >> 2 : (t & 7) * 8, c = a_k ? (t & 3) * 8 : t >> 3;
            const int m = m0 + r, k = k0 + c;
            const __bf16* q = A + (long)m * sam + (long)k * sak;
            const bool full = a_k ? (m < M && k + 8 <= k_hi) : (k < k_hi && m + 8 <= M);
            if (full && aligned(q)) {
                const bf16x8_t v = *(const bf16x8_t*)q;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    if (a_k) As[r][c + e] = v[e];
                    else As[r + e][c] = v[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int mm = a_k ? m : m + e, kk = a_k ? k + e : k;
                    As[a_k ? r : r + e][a_k ? c + e : c] =
                        (mm < M && kk < k_hi) ? A[(long)mm * sam + (long)kk * sak] : to_bf16(0.0f);
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int idx = t * 8 + e;
                const int r = idx >> 5, c = idx & 31;
                const int m = m0 + r, k = k0 + c;
                As[r][c] = (m < M && k < k_hi) ? A[(long)m * sam + (long)k * sak] : to_bf16(0.0f);
            }
        }
        if (b_k || b_n) {
            const int r = b_k ? t >> 2 : (t & 7) * 8, c = b_k ? (t & 3) * 8 : t >> 3;
            const int n = n0 + r, k = k0 + c;
            const __bf16* q = Bm + (long)k * sbk + (long)n * sbn;
            const bool full = b_k ? (n < N && k + 8 <= k_hi) : (k < k_hi && n + 8 <= N);
            if (full && aligned(q)) {
                const bf16x8_t v = *(const bf16x8_t*)q;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    if (b_k) Bs[r][c + e] = v[e];
                    else Bs[r + e][c] = v[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int nn = b_k ? n : n + e, kk = b_k ? k + e : k;
                    Bs[b_k ? r : r + e][b_k ? c + e : c] =
                        (nn < N && kk < k_hi) ? Bm[(long)kk * sbk + (long)nn * sbn] : to_bf16(0.0f);
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int idx = t * 8 + e;
                const int r = idx >> 5, c = idx & 31;
                const int n = n0 + r, k = k0 + c;
                Bs[r][c] = (n < N && k < k_hi) ? Bm[(long)k * sbk + (long)n * sbn] : to_bf16(0.0f);
            }
        }
        __syncthreads();
        const int row = l & 15, kq = l >> 4;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const bf16x8_t av = lds8(&As[wm + a * 16 + row][8 * kq]);
                const bf16x8_t bv = lds8(&Bs[wn + b * 16 + row][8 * kq]);
                acc[a][b] = mfma16(av, bv, acc[a][b]);
            }
        __syncthreads();
    }
    float* Cs = C + (size_t)blockIdx.z * M * N;
    const int kq = l >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + a * 16 + 4 * kq + r, n = n0 + wn + b * 16 + (l & 15);
                if (m < M && n < N) Cs[(size_t)m * N + n] = acc[a][b][r];
            }
}

__global__ LAUNCH_BOUNDS_256 void k_gemm_reduce(const float* __restrict__ part, int splits, size_t mn,
                                                float* __restrict__ C)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mn) return;
    float s = part[i];
    for (int k = 1; k < splits; k++) s += part[(size_t)k * mn + i];
    C[i] = s;
}

/* ---------------------------------------------------------- elementwise steps */
__global__ LAUNCH_BOUNDS_256 void k_to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

/* rows of X (raw features) normalised and converted to bf16 */
__global__ LAUNCH_BOUNDS_256 void k_normalize_rows(const float* __restrict__ X, const float* __restrict__ mu,
                                                   const float* __restrict__ inv_sigma, __bf16* __restrict__ Xn,
                                                   size_t rows)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * MLP_IN) return;
    const int k = (int)(i % MLP_IN);
    Xn[i] = to_bf16((X[i] - mu[k]) * inv_sigma[k]);
}

/* Z += bias (per column); A = silu(Z) in fp32 and bf16 */
__global__ LAUNCH_BOUNDS_256 void k_bias_silu(float* __restrict__ Z, const float* __restrict__ bias, int cols,
                                              size_t rows, __bf16* __restrict__ Ab, int apply_silu)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const float z = Z[i] + bias[i % cols];
    Z[i] = z;
    if (Ab) Ab[i] = to_bf16(apply_silu ? silu(z) : z);
}

/* loss = sum over rows with mask of |Y[:, :3] - T|^2 / count; dY = 2 (Y - T) / count */
__global__ LAUNCH_BOUNDS_256 void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                        const float* __restrict__ mask, size_t rows, float inv_count,
                                        float* __restrict__ dY, __bf16* __restrict__ dYb, float* __restrict__ lpart)
{
    __shared__ float red[256];
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (r < rows) {
        const float mk = mask[r];
        for (int c = 0; c < MLP_OUT; c++) {
            float g = 0.f;
            if (c < MLP_OUT_REAL && mk != 0.f) {
                const float e = Y[r * MLP_OUT + c] - T[r * MLP_OUT_REAL + c];
                l += e * e;
                g = 2.0f * e * inv_count;
            }
            dY[r * MLP_OUT + c] = g;
            dYb[r * MLP_OUT + c] = to_bf16(g);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0] * inv_count;
}

/* dZ = dA * silu'(Z) (fp32 and bf16) */
__global__ LAUNCH_BOUNDS_256 void k_silu_back(const float* __restrict__ dA, const float* __restrict__ Z, size_t n,
                                              float* __restrict__ dZ, __bf16* __restrict__ dZb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dA[i] * silu_grad(Z[i]);
    dZ[i] = g;
    dZb[i] = to_bf16(g);
}

/* column sums of G [rows x cols], one block per column, fixed-order tree: deterministic */
__global__ LAUNCH_BOUNDS_256 void k_colsum(const float* __restrict__ G, size_t rows, int cols, float* __restrict__ out)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (size_t r = threadIdx.x; r < rows; r += 256) s += G[r * cols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

/* Adam (Kingma & Ba, 2015) on fp32 master weights */
__global__ LAUNCH_BOUNDS_256 void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                         float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                                         float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* W [K x N] fp32 row-major -> fragment layout (ntile, kchunk, col, k) in bf16 */
/* Fragment kk of chunk kc, column col of tile nt takes W[k][n] with k = 32 kc + kk, or with
 * perm the row whose neuron the previous layer's result leaves at that position:
 * nu(32 kc + 8 q + e) = 32 kc + (e < 4 ? 4 q + e : 16 + 4 q + e - 4). */
__global__ LAUNCH_BOUNDS_256 void k_pack_frag(const float* __restrict__ W, int K, int N, __bf16* __restrict__ F,
                                              int perm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * N) return;
    const int k = i / N, n = i % N;
    const int nt = n / 16, col = n % 16, kc = k / 32;
    int kk = k % 32;
    if (perm) {
        /* the inverse of nu inside the chunk: neuron offset o = kk -> position 8 q + e */
        const int o = kk, hi = o >= 16 ? 1 : 0, rem = o - 16 * hi;
        kk = 8 * (rem / 4) + 4 * hi + rem % 4;
    }
    F[(((size_t)nt * (K / 32) + kc) * 16 + col) * 32 + kk] = to_bf16(W[i]);
}
