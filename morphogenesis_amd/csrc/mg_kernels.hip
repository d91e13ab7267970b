/*
 * mg_kernels.hip — gfx950 kernels of the Morphogenesis per-timestep SPH step.
 *
 * One thread per particle (or per cell) with wave64-aware access patterns; every kernel
 * reads the live particle counts from device counters, so a step needs no host sync.
 * The arithmetic follows the reference kernels in src/fluid_system_cuda.cu (cited per
 * kernel) with the deterministic order described in mg_device.h.
 *
 * Two launch sequences share the per-particle device functions below, so they produce
 * identical bits:
 *   phase kernels (k_insert, k_scan, k_place, k_sort_gather, k_pressure,
 *   k_force, k_advance, k_diffusion, k_genes), one per reference phase, used by the
 *   FluidSystem-style entry points;
 *   fused kernels (k_force_advance, k_diffuse_genes_insert), the hot loop of mg_step:
 *   pressure -> force+advance -> diffusion+genes+insert -> scan -> place -> rank+gather.
 */
#include "mg_device.h"

#include <cfloat>
#include <climits>

#define LAUNCH_BOUNDS_256 __launch_bounds__(256)
/* sweep kernels: at most 128 VGPRs, so four 256-thread blocks fit on a CU (every wave of a
 * 262k-particle launch resident at once) */
#define LAUNCH_BOUNDS_SWEEP __launch_bounds__(256, 4)

/* orders the wavefront's LDS writes before its later LDS reads (and the reverse) */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* ------------------------------------------------------------------ fills */
__global__ LAUNCH_BOUNDS_256 void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

/* -------------------------------------------------------------- insert (a4)
 * insertParticles, src/fluid_system_cuda.cu:53-100, for the particle in slot s (all lanes of
 * the wavefront must call it). Lanes of one wavefront that fall in the same cell (the input
 * is cell-sorted from the previous step, so they are contiguous) are counted with one
 * atomicAdd per run: the run head adds the run length and broadcasts the base. FGNDX
 * receives a slot in [0, count) as the reference's atomicAdd does; the sort later replaces
 * it with the stable rank (k_sort_gather). */
__device__ __forceinline__ void insert_one(const DS& d, const MgGrid& g, uint32_t s, bool valid,
                                           float4 p, uint32_t* CNT, uint32_t* GCNT)
{
    const uint32_t lane = __lane_id();
    uint32_t key = 0xFFFFFFFFu;
    if (valid) {
        float gx = (p.x - g.gmin[0]) * g.delta[0];
        float gy = (p.y - g.gmin[1]) * g.delta[1];
        float gz = (p.z - g.gmin[2]) * g.delta[2];
        int cx = __float2int_rz(gx), cy = __float2int_rz(gy), cz = __float2int_rz(gz);
        if (cx >= 1 && cx <= g.scan_max[0] && cy >= 1 && cy <= g.scan_max[1] && cz >= 1 &&
            cz <= g.scan_max[2]) {
            key = (uint32_t)((cy * g.res[2] + cz) * g.res[0] + cx);
        } else {
            d.GCELL[s] = MG_GRID_UNDEF;
            if (s < d.ctr[MG_C_TAIL]) d.SLOT[d.HOME[s]] = MG_NONE; /* particle leaves the model */
        }
    }
    uint32_t prev = __shfl_up(key, 1);
    bool head = (lane == 0) || (prev != key);
    uint64_t heads = __ballot(head);
    uint64_t le = heads & ((2ull << lane) - 1ull);
    uint32_t start = 63u - (uint32_t)__clzll((long long)le);
    uint64_t gt = heads & ~((2ull << lane) - 1ull);
    uint32_t next = gt ? (uint32_t)(__ffsll((long long)gt) - 1) : 64u;
    uint32_t base = 0;
    if (head && key != 0xFFFFFFFFu) base = atomicAdd(&CNT[key], next - start);
    base = __shfl(base, (int)start);
    if (key != 0xFFFFFFFFu) {
        d.GCELL[s] = key;
        d.GNDX[s] = base + (lane - start);
        if (GCNT) {
            const uint32_t h = d.HOME[s];
            for (int gene = 0; gene < MG_NUM_GENES; gene++)
                if (d.EPIGEN[(size_t)gene * d.N + h] > 0u)
                    atomicAdd(&GCNT[(size_t)gene * d.C + key], 1u);
        }
    }
}

__global__ LAUNCH_BOUNDS_256 void k_insert(DS d, MgGrid g, int gene_counts)
{
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = s < (uint32_t)d.N;
    float4 p = valid ? d.P[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    insert_one(d, g, s, valid, p, d.CNT, gene_counts ? d.GCNT : nullptr);
}

/* ----------------------------------------------------------- scan (a5, a6)
 * Exclusive prefix sum (prefixSum/prefixFixup with zero_offsets=1,
 * src/fluid_system_cuda.cu:136-183, driven by src/host_CUDA.cpp:307-381). One 1024-thread
 * block per segment (blockIdx.x); each thread owns a contiguous run of the segment.
 * Segment 0 with `cell_counters` set also records OFF[C-1] and OFF[grid_ScanMax], which the
 * sort commits as mNumPoints and mActivePoints (src/host_CUDA.cpp:501-511). `lens`, when
 * given, receives count[C-1] + offset[C-1] per segment (tally_denselist_lengths,
 * src/fluid_system_cuda.cu:185-194). */
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* in, uint32_t* out, int n,
                                               uint32_t* ctr, int cell_counters, int gsm,
                                               uint32_t* lens)
{
    __shared__ uint32_t wsum[16];
    const int seg = blockIdx.x;
    in += (size_t)seg * n;
    out += (size_t)seg * n;
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, t * per);
    const int b1 = min(n, b0 + per);
    uint32_t local = 0;
    for (int i = b0; i < b1; i++) local += in[i];
    const int lane = t & 63, w = t >> 6;
    uint32_t x = local;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 16; k++) { uint32_t v = wsum[k]; wsum[k] = run; run += v; }
    }
    __syncthreads();
    uint32_t run = wsum[w] + x - local;
    for (int i = b0; i < b1; i++) { uint32_t v = in[i]; out[i] = run; run += v; }
    __syncthreads();
    if (t == 0) {
        if (lens) lens[seg] = in[n - 1] + out[n - 1];
        if (cell_counters) {
            ctr[MG_C_SCAN_NUM] = out[n - 1];
            ctr[MG_C_SCAN_ACTIVE] = out[gsm];
        }
    }
}

/* Single-pass exclusive scan with decoupled look-back (Merrill & Garland, "Single-pass
 * Parallel Prefix Scan with Decoupled Look-back", 2016) for the cell counts of the fused
 * step. Tiles of 4096 counts (256 threads x 16); each tile publishes its aggregate, then
 * walks back over predecessors' published aggregates / inclusive prefixes until it meets an
 * inclusive one. Tile ids come from a monotonic 64-bit ticket taken at block start, so a tile
 * only ever waits on tiles that started before it (no deadlock), and the launch number
 * (ticket / tiles) tags every status word, so status words from earlier launches read as
 * "not ready" without any reset (also under hipGraph replay). The ticket is 64 bits wide so
 * that it never wraps: a 32-bit ticket would, after 2^32 tiles, split a launch over two
 * epochs whenever the tile count does not divide 2^32. The tag keeps the launch number's low
 * 30 bits; a reader only ever meets the previous launch's words or its own launch's, since
 * every launch rewrites every tile's word.
 * status word: [63:34] launch epoch, [33:32] flag (1 aggregate, 2 inclusive), [31:0] value. */
#define SCAN_TILE 4096
__device__ __forceinline__ void scan_lookback_body(const uint32_t* __restrict__ in,
                                                   uint32_t* __restrict__ out, int n,
                                                   unsigned long long* status,
                                                   unsigned long long* ticket, uint32_t* ctr,
                                                   int gsm, uint32_t tiles)
{
    __shared__ unsigned long long s_ticket;
    __shared__ uint32_t s_excl;
    __shared__ uint32_t wsum[4];
    const int t = threadIdx.x;
    if (t == 0) s_ticket = atomicAdd(ticket, 1ull);
    __syncthreads();
    const uint32_t tile = (uint32_t)(s_ticket % tiles);
    const unsigned long long epoch = (s_ticket / tiles + 1ull) & 0x3FFFFFFFull;
    const int base = (int)tile * SCAN_TILE + t * 16;
    uint32_t v[16];
    if (base + 15 < n) {
        const uint4* q = (const uint4*)(in + base);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint4 w = q[k];
            v[4 * k] = w.x; v[4 * k + 1] = w.y; v[4 * k + 2] = w.z; v[4 * k + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = (base + k < n) ? in[base + k] : 0u;
    }
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) local += v[k];
    const int lane = t & 63, w = t >> 6;
    uint32_t x = local;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for (int k = 0; k < 4; k++) {
        if (k < w) wbase += wsum[k];
        total += wsum[k];
    }
    /* Look-back by the first wavefront: lane l reads the status word of tile (top - l), so one
     * round trip covers 64 predecessors. The nearest inclusive word ends the walk; every word
     * between it and this tile must be published (aggregate or inclusive), else the window is
     * read again. A window without an inclusive word adds its 64 aggregates and moves down. */
    if (w == 0) {
        unsigned long long* st = status + tile;
        if (lane == 0)
            __hip_atomic_store(st, (epoch << 34) | ((tile == 0 ? 2ull : 1ull) << 32) | total,
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t excl = 0;
        int top = (int)tile - 1;
        uint32_t spins = 0;
        while (top >= 0) {
            const int pidx = top - lane;
            bool ready = true, incl = true;
            uint32_t val = 0;
            if (pidx >= 0) {
                const unsigned long long wd = __hip_atomic_load(status + pidx, __ATOMIC_ACQUIRE,
                                                                __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long flag = (wd >> 32) & 3ull;
                ready = (wd >> 34) == epoch && flag != 0ull;
                incl = ready && flag == 2ull;
                val = (uint32_t)wd;
            }
            const unsigned long long im = __ballot(incl);
            const int stop = im ? __builtin_ctzll(im) : 64;
            const unsigned long long need = stop >= 63 ? ~0ull : (2ull << stop) - 1ull;
            if (__ballot(!ready) & need) {
                /* bounded wait: predecessors always publish, but never spin forever */
                if (++spins > (1u << 24)) {
                    if (lane == 0) atomicOr(&ctr[MG_C_ERR], 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            uint32_t part = lane <= stop ? val : 0u;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o);
            excl += part;
            if (stop < 64) break;
            top -= 64;
        }
        if (lane == 0) {
            if (tile != 0)
                __hip_atomic_store(st, (epoch << 34) | (2ull << 32) | (uint32_t)(excl + total),
                                   __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            s_excl = excl;
        }
    }
    __syncthreads();
    uint32_t run = s_excl + wbase + x - local;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int idx = base + k;
        if (idx < n) {
            out[idx] = run;
            if (idx == n - 1) ctr[MG_C_SCAN_NUM] = run;
            if (idx == gsm) ctr[MG_C_SCAN_ACTIVE] = run;
        }
        run += v[k];
    }
}

__global__ __launch_bounds__(256) void k_scan_lookback(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, int n,
                                                       unsigned long long* status,
                                                       unsigned long long* ticket, uint32_t* ctr,
                                                       int gsm)
{
    scan_lookback_body(in, out, n, status, ticket, ctr, gsm, gridDim.x);
}

/* ------------------------------------------------ two-level exclusive scan
 * For arrays the cell scans do not cover (incoming-bond and change-list offsets), one segment
 * of n words per blockIdx.y: k_tile_sums reduces tiles of 4096 words (and, when `maxw` is
 * given, folds their maximum into *maxw), k_scan scans each segment's tile sums in one block,
 * k_tile_scan writes each tile's exclusive prefix. */
__global__ __launch_bounds__(256) void k_tile_sums(const uint32_t* __restrict__ in, int n,
                                                   uint32_t* __restrict__ sums, uint32_t* maxw)
{
    __shared__ uint32_t ws[4], wm[4];
    in += (size_t)blockIdx.y * n;
    sums += (size_t)blockIdx.y * gridDim.x;
    const int t = threadIdx.x, base = (int)blockIdx.x * SCAN_TILE + t * 16;
    uint32_t s = 0, m = 0;
    for (int k = 0; k < 16; k++) {
        const uint32_t v = base + k < n ? in[base + k] : 0u;
        s += v;
        m = max(m, v);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        m = max(m, (uint32_t)__shfl_xor(m, o));
    }
    if ((t & 63) == 0) { ws[t >> 6] = s; wm[t >> 6] = m; }
    __syncthreads();
    if (t == 0) {
        sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
        if (maxw) atomicMax(maxw, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
    }
}

__global__ __launch_bounds__(256) void k_tile_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   int n, const uint32_t* __restrict__ tile_excl)
{
    __shared__ uint32_t wsum[4];
    in += (size_t)blockIdx.y * n;
    out += (size_t)blockIdx.y * n;
    tile_excl += (size_t)blockIdx.y * gridDim.x;
    const int t = threadIdx.x, base = (int)blockIdx.x * SCAN_TILE + t * 16;
    uint32_t v[16], local = 0;
    for (int k = 0; k < 16; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        local += v[k];
    }
    const int lane = t & 63, w = t >> 6;
    uint32_t x = local;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t run = tile_excl[blockIdx.x] + x - local;
    for (int k = 0; k < w; k++) run += wsum[k];
    for (int k = 0; k < 16; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

/* ------------------------------------------------ incoming-bond lists
 * INC_OFF[j] .. INC_OFF[j + 1] lists the bonds whose partner is home j as owner_home * B + slot,
 * ordered by (owner's particle ID, owner home, slot), so the force sums a particle's incoming
 * terms in an order that is a property of the state, not of the launch. k_inc_count counts
 * per partner, keeps each bond's place among its partner's (the count's atomic, any order) and
 * records each owner's live-slot mask; after the scan k_inc_fill scatters the entries to those
 * places and k_inc_order puts each partner's run in key order:
 * one wavefront per four partners (their runs are adjacent), each entry's position is the
 * number of smaller keys in its run (keys are distinct), runs of up to 256 entries (longer
 * runs take the host build). */
__global__ LAUNCH_BOUNDS_256 void k_inc_count(const uint32_t* __restrict__ BPARTNER, uint32_t* __restrict__ cnt,
                                              uint32_t* __restrict__ POS, uint64_t* __restrict__ LIVE, int N, int B)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= (uint32_t)N) return;
    uint64_t live = 0;
    for (int a = 0; a < B; a++) {
        const size_t bi = (size_t)a * N + h;
        const uint32_t j = BPARTNER[bi];
        if (j < (uint32_t)N) {
            POS[bi] = atomicAdd(&cnt[j], 1u); /* the bond's place in its partner's run */
            if (a < 64) live |= 1ull << a;
        }
    }
    LIVE[h] = live;
}

__global__ LAUNCH_BOUNDS_256 void k_inc_fill(const uint32_t* __restrict__ BPARTNER, const uint32_t* __restrict__ OFF,
                                             const uint32_t* __restrict__ POS, uint32_t* __restrict__ ENT, int N, int B)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= (uint32_t)N) return;
    for (int a = 0; a < B; a++) {
        const size_t bi = (size_t)a * N + h;
        const uint32_t j = BPARTNER[bi];
        if (j < (uint32_t)N) ENT[OFF[j] + POS[bi]] = h * (uint32_t)B + (uint32_t)a;
    }
}

#define INC_RUN_MAX 256
#define INC_PER_WAVE 4 /* partners per wavefront: their runs are one contiguous range of ENT */
__global__ __launch_bounds__(256) void k_inc_order(const uint32_t* __restrict__ OFF, uint32_t* __restrict__ ENT,
                                                   const uint32_t* __restrict__ PID, int N, int B)
{
    __shared__ unsigned long long keys[4][INC_PER_WAVE * INC_RUN_MAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t j0 = ((uint32_t)blockIdx.x * 4u + (uint32_t)w) * INC_PER_WAVE;
    if (j0 >= (uint32_t)N) return;
    uint32_t st[INC_PER_WAVE + 1]; /* the runs' bounds relative to the first run's start */
    const uint32_t o = OFF[j0];
#pragma unroll
    for (int k = 0; k <= INC_PER_WAVE; k++) st[k] = OFF[min(j0 + (uint32_t)k, (uint32_t)N)] - o;
    const uint32_t tot = st[INC_PER_WAVE];
    if (tot > INC_PER_WAVE * INC_RUN_MAX) return; /* k_inc_order_long takes this group */
    for (uint32_t t = (uint32_t)lane; t < tot; t += 64) {
        const uint32_t e = ENT[o + t];
        keys[w][t] = ((unsigned long long)PID[e / (uint32_t)B] << 32) | e;
    }
    wave_lds_sync();
    for (uint32_t t = (uint32_t)lane; t < tot; t += 64) {
        uint32_t rs = 0, re = st[1];
#pragma unroll
        for (int k = 1; k < INC_PER_WAVE; k++)
            if (t >= st[k]) {
                rs = st[k];
                re = st[k + 1];
            }
        const unsigned long long key = keys[w][t];
        uint32_t rank = 0;
        for (uint32_t q = rs; q < re; q++) rank += keys[w][q] < key ? 1u : 0u;
        ENT[o + rs + rank] = (uint32_t)key;
    }
}

/* The groups of INC_PER_WAVE partners whose runs together exceed one wavefront's stage (a hub
 * particle with hundreds of incoming bonds): one workgroup per group, the group's keys (up to
 * INC_LONG_MAX) in LDS, each entry written at its rank among its run's keys. */
#define INC_LONG_MAX 6144
__global__ __launch_bounds__(256) void k_inc_order_long(const uint32_t* __restrict__ OFF, uint32_t* __restrict__ ENT,
                                                        const uint32_t* __restrict__ PID, int N, int B)
{
    __shared__ unsigned long long keys[INC_LONG_MAX];
    const uint32_t j0 = (uint32_t)blockIdx.x * INC_PER_WAVE;
    if (j0 >= (uint32_t)N) return;
    uint32_t st[INC_PER_WAVE + 1];
    const uint32_t o = OFF[j0];
#pragma unroll
    for (int k = 0; k <= INC_PER_WAVE; k++) st[k] = OFF[min(j0 + (uint32_t)k, (uint32_t)N)] - o;
    const uint32_t tot = st[INC_PER_WAVE];
    if (tot <= INC_PER_WAVE * INC_RUN_MAX || tot > INC_LONG_MAX) return; /* k_inc_order's, or the host's */
    for (uint32_t t = threadIdx.x; t < tot; t += blockDim.x) {
        const uint32_t e = ENT[o + t];
        keys[t] = ((unsigned long long)PID[e / (uint32_t)B] << 32) | e;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < tot; t += blockDim.x) {
        uint32_t rs = 0, re = st[1];
#pragma unroll
        for (int k = 1; k < INC_PER_WAVE; k++)
            if (t >= st[k]) {
                rs = st[k];
                re = st[k + 1];
            }
        const unsigned long long key = keys[t];
        uint32_t rank = 0;
        for (uint32_t q = rs; q < re; q++) rank += keys[q] < key ? 1u : 0u;
        ENT[o + rs + rank] = (uint32_t)key;
    }
}

/* ------------------------------------------------------ placement
 * PERM[OFF[cell] + slot] = s: each cell's run of PERM holds its particles in insertion order;
 * the gather ranks them by pre-sort index (the stable order) and FGNDX becomes that rank. */
__device__ __forceinline__ void place_body(const DS& d, uint32_t bid)
{
    const uint32_t s = bid * blockDim.x + threadIdx.x;
    if (s >= (uint32_t)d.N) return;
    const uint32_t c = d.GCELL[s];
    if (c == MG_GRID_UNDEF) return;
    d.PERM[d.OFF[c] + d.GNDX[s]] = s;
}

__global__ LAUNCH_BOUNDS_256 void k_place(DS d) { place_body(d, blockIdx.x); }

/* ------------------------------------------------------- counting sort (a7)
 * countingSortFull (src/fluid_system_cuda.cu:196-296) with the host resets of
 * src/host_CL.cpp:1156-1206: slot dd < mNumPoints takes the particle of rank dd - OFF[c] in
 * its cell c (ranked by pre-sort index, the stable order); slots past mNumPoints get the
 * reset values (FPOS filled with max(VOLMAX), FCONC zero) and keep the stale ones the
 * reference leaves in place. Bond records and incoming lists do not move: they are addressed
 * by home index, and the partner indices the reference rewrites here are resolved through
 * SLOT. Reads P/V/VEL/CONC/HOME/GCELL/GNDX, writes their "2" halves. The gene membership
 * comes from GMN (precomputed) when given, else from EPIGEN. Block 0 commits the counters.
 *
 * PERM holds each cell's run in insertion order (k_place), d.CNT / d.OFF are the grid k_place
 * used. Thread p takes the particle s = PERM[p], counts the smaller pre-sort indices in its
 * cell's run (its rank r: the stable order) and moves s to dd = OFF[c] + r. The destinations
 * of a wavefront are a permutation of its own slots within their cells' runs, so the stores
 * stay as contiguous as a gather's; the runs it re-reads are a few cache lines per cell. */
__device__ __forceinline__ void sort_gather_body(const DS& d, float vmax, uint32_t bid)
{
    const uint32_t p = bid * blockDim.x + threadIdx.x;
    const uint32_t nump = d.ctr[MG_C_SCAN_NUM];
    if (bid == 0 && threadIdx.x == 0) {
        d.ctr[MG_C_NUM_POINTS] = nump;
        d.ctr[MG_C_ACTIVE] = d.ctr[MG_C_SCAN_ACTIVE];
        d.ctr[MG_C_TAIL] = nump;
        if (nump > d.ctr[MG_C_HWM]) d.ctr[MG_C_HWM] = nump;
        d.ctr[MG_C_LISTS] = 1u;
    }
    const uint32_t N = (uint32_t)d.N;
    /* the last cell's particles (possible when scan_max reaches the grid's edge) are placed
     * past mNumPoints = OFF[C - 1]: ranked as well, they move nowhere */
    const uint32_t lastn = d.CNT[d.C - 1];
    const bool placed = p < nump || (p < N && p - nump < lastn);
    const uint32_t s = placed ? d.PERM[p] : 0u;
    uint32_t dd = p, rank = 0, cell = 0;
    if (placed) {
        cell = d.GCELL[s];
        if (cell < (uint32_t)d.C) {
            const uint32_t o = min(d.OFF[cell], N);
            const uint32_t n = min(d.CNT[cell], N - o);
            const uint32_t* run = d.PERM + o;
            for (uint32_t k0 = 0; k0 < n; k0 += 8) {
                uint32_t v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) v[j] = run[min(k0 + (uint32_t)j, n - 1u)];
#pragma unroll
                for (int j = 0; j < 8; j++) rank += (k0 + (uint32_t)j < n && v[j] < s) ? 1u : 0u;
            }
            if (p < nump) dd = min(o + rank, N - 1u); /* s is in its run: rank < n */
        }
        /* a reset-tail slot keeps the pre-sort FGNDX of its own index, which for a placed
         * particle is its rank (the reference's insert slot, taken in index order) */
        if (s >= nump && s < N) d.GNDX2[s] = rank;
    }
    {
        /* the 64-byte concentration rows of the wavefront's 64 slots, copied by the whole
         * wavefront in four passes of 64 consecutive float4 (16 rows each), so every store
         * writes whole rows; row q takes its source slot and destination from lane q, the
         * reset tail takes zeros */
        const uint32_t lane = __lane_id(), wbase = p - lane;
        const float4* csrc = (const float4*)d.CONC;
        float4* cdst = (float4*)d.CONC2;
#pragma unroll
        for (uint32_t k = 0; k < 4u; k++) {
            const uint32_t f = k * 64u + lane, q = f >> 2;
            const uint32_t sq = (uint32_t)__shfl((int)s, (int)q);
            const uint32_t dq = (uint32_t)__shfl((int)dd, (int)q);
            if (wbase + q < N)
                cdst[(size_t)dq * 4u + (f & 3u)] =
                    wbase + q < nump ? csrc[(size_t)sq * 4u + (f & 3u)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    if (p >= N) return;
    if (p < nump) {
        const uint32_t h = d.HOME[s];
        d.SLOT[h] = dd;
        d.P2[dd] = d.P[s];
        d.V2[dd] = d.V[s];
        d.VEL2[dd] = d.VEL[s];
        d.HOME2[dd] = h;
        d.GCELL2[dd] = cell;
        d.GNDX2[dd] = rank;
        uint32_t m = 0;
        if (d.GMN) {
            m = d.GMN[s];
        } else {
            for (int gene = 0; gene < MG_NUM_GENES; gene++)
                if (d.EPIGEN[(size_t)gene * d.N + h] > 0u) m |= 1u << gene;
        }
        d.GMASK2[dd] = m;
    } else {
        float4 q = d.P[p];
        d.P2[p] = make_float4(vmax, vmax, vmax, q.w);
        d.V2[p] = d.V[p];
        d.VEL2[p] = d.VEL[p];
        d.HOME2[p] = d.HOME[p];
        const uint32_t gc = d.GCELL[p];
        d.GCELL2[p] = gc;
        if (gc >= (uint32_t)d.C) d.GNDX2[p] = d.GNDX[p]; /* placed: its rank, above */
        d.GMASK2[p] = 0u;
    }
}

__global__ LAUNCH_BOUNDS_256 void k_sort_gather(DS d, float vmax) { sort_gather_body(d, vmax, xcd_block()); }

/* countingSortDenseLists (src/fluid_system_cuda.cu:329-404): per bin, particles in order,
 * genes with a nonzero epigenetic state append the particle to the gene's list. */
__global__ LAUNCH_BOUNDS_256 void k_dense_lists(DS d)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (uint32_t)d.C) return;
    const uint32_t n = d.CNT[c];
    if (n == 0) return;
    const uint32_t base = d.OFF[c];
    if (base + n > (uint32_t)d.N) return;
    uint32_t ctrg[MG_NUM_GENES];
    for (int g = 0; g < MG_NUM_GENES; g++) ctrg[g] = 0;
    for (uint32_t q = base; q < base + n; q++) {
        const uint32_t m = d.GMASK[q];
        for (int g = 0; g < MG_NUM_GENES; g++)
            if (m & (1u << g)) {
                uint32_t at = d.GOFF[(size_t)g * d.C + c] + ctrg[g]++;
                if (at < (uint32_t)d.N) d.LISTS[(size_t)g * d.N + at] = q;
            }
    }
}

/* Gene counts per cell from the sorted state (rebuilds the gene grid after fused steps). */
__global__ LAUNCH_BOUNDS_256 void k_gene_counts_sorted(DS d)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.ctr[MG_C_NUM_POINTS]) return;
    const uint32_t c = d.GCELL[q];
    if (c == MG_GRID_UNDEF) return;
    const uint32_t m = d.GMASK[q];
    for (int g = 0; g < MG_NUM_GENES; g++)
        if (m & (1u << g)) atomicAdd(&d.GCNT[(size_t)g * d.C + c], 1u);
}

/* ------------------------------------------------------------ neighbour sweep
 * The reference visits the srch^3 cells gc + gridAdj[c] in gridAdj order and, inside a cell,
 * the slots OFF[cell] .. OFF[cell] + CNT[cell] - 1 in order. gridAdj is built y, z, x with
 * x fastest, so the cells come in srch^2 rows of srch consecutive cell indices; in the
 * sorted state (d.rows) cell k of a row starting at cell c0 holds the slots
 * [OFF[c0 + k], OFF[c0 + k + 1]), so one row needs srch + 1 offsets. A sweep calls
 *   visit(j, P[j], x_i - x_j, |x_i - x_j|^2, x[])  for candidates with 0 < dsq < lim,
 *   end_cell()                                     after each cell,
 * in exactly the reference's order, fetching up to K candidates of a cell at a time so their
 * loads are in flight together; x[] holds the caller's NX extra float4 fields of slot j
 * (velocity, concentrations), fetched for hits only. Callers use end_cell() to add a
 * per-cell partial sum: for a cell without hits it adds +0, a no-op for sums that start at
 * +0, so cells may also be skipped.
 *
 * With d.cull (slots sorted by cell and positions unchanged since the binning), cells whose
 * box lies farther than sqrt(lim) from x_i are skipped: no candidate in them can hit. The
 * box distance is shrunk by 1e-3 cell per axis, far above the rounding of the binning and
 * of dsq, so the skip never drops a hit.
 *
 * P arrays carry MG_P_PAD slots past N, so a batch reads P[j0 .. j0 + K - 1] unclamped. */
#define MG_P_PAD 16

struct RowBounds {
    uint32_t b[5]; /* cell k of the row: [b[k], b[k + 1]) */
};

__device__ __forceinline__ RowBounds row_bounds(const DS& d, const MgGrid& g, uint32_t gc0, int s)
{
    RowBounds r;
    if (d.rows) {
        const uint32_t c0 = gc0 + (uint32_t)g.adj[s * g.srch];
#pragma unroll
        for (int k = 0; k < 5; k++) r.b[k] = (k <= g.srch) ? d.OFF[c0 + (uint32_t)k] : 0u;
    } else {
        const uint32_t c = gc0 + (uint32_t)g.adj[s];
        r.b[0] = d.OFF[c];
        r.b[1] = r.b[0] + d.CNT[c];
        r.b[2] = r.b[3] = r.b[4] = 0u;
    }
    /* never read past the particle arrays, whatever CNT / OFF hold */
#pragma unroll
    for (int k = 0; k < 5; k++) r.b[k] = min(r.b[k], (uint32_t)d.N);
    return r;
}

/* squared distance from x_i to the box of the neighbour cell at offset (ox, oy, oz) from the
 * particle's own cell, in world units, each axis gap shrunk by 1e-3 cell; fr = the particle's
 * position inside its own cell in cell units */
__device__ __forceinline__ float axis_gap(float fr, int o, float inv_delta)
{
    const float gap = o < 0 ? fr - (float)(o + 1) : (o > 0 ? (float)o - fr : 0.0f);
    return fmaxf(gap - 1e-3f, 0.0f) * inv_delta;
}

/* Wave-staged bands. The 64 particles of a wavefront sit in a narrow range of cells
 * [cmin, cmax] (the slots are cell-sorted), so for one row of the neighbourhood all their
 * candidates lie in the slot band [OFF[cmin + a], OFF[cmax + a + srch]) (a = the row's
 * gridAdj offset; OFF is non-decreasing in the sorted state). The wave loads the band once,
 * coalesced, into LDS (positions and NX extra float4 fields per slot) and every lane then
 * reads its candidates from LDS instead of issuing dependent global loads. The band of row
 * s + 1 is fetched into registers while row s is swept. A band wider than BAND_CAP slots,
 * or any sweep outside the sorted state, reads global memory directly. */
#define BAND_CAP 128
#define BAND_PAD 8 /* a batch of up to 8 candidates may read past the band's end */

template <int NX, int CAP = BAND_CAP>
struct Band {
    float4 p[CAP + BAND_PAD];
    float4 x[NX > 0 ? NX : 1][CAP + BAND_PAD];
};


__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

/* The candidates of one cell, slots [lo, hi), K at a time. getp(j) / getx(j, q) fetch the
 * position / extra field q of slot j; extras are fetched for the hits of a batch only. */
template <int K, int NX, class GetP, class GetX, class Visit>
__device__ __forceinline__ void cell_candidates(uint32_t lo, uint32_t hi, f3 pi, float lim, GetP getp,
                                                GetX getx, Visit visit)
{
    constexpr int NXX = NX > 0 ? NX : 1;
    for (uint32_t j0 = lo; j0 < hi; j0 += K) {
        float4 pj[K];
#pragma unroll
        for (int q = 0; q < K; q++) pj[q] = getp(j0 + (uint32_t)q);
        bool hit[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const f3 dist = pi - xyz(pj[q]);
            const float dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
            hit[q] = (j0 + (uint32_t)q < hi) && dsq < lim && dsq > 0.0f;
        }
        float4 ex[K][NXX];
#pragma unroll
        for (int q = 0; q < K; q++)
#pragma unroll
            for (int x = 0; x < NX; x++)
                if (hit[q]) ex[q][x] = getx(j0 + (uint32_t)q, x);
#pragma unroll
        for (int q = 0; q < K; q++) {
            if (hit[q]) {
                /* recomputed (same operations) rather than kept in registers */
                const f3 dist = pi - xyz(pj[q]);
                const float dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
                visit(j0 + (uint32_t)q, pj[q], dist, dsq, ex[q]);
            }
        }
    }
}

/* Must be called by all 64 lanes of the wavefront (`valid` selects the lanes that sweep);
 * xload(j, q) is the global load of extra field q of slot j. K candidates are tested at a
 * time; bands of up to CAP (<= 128) slots are staged. */
template <int K, int NX, int HM = 0, int CAP = BAND_CAP, class XLoad, class Visit, class EndCell>
__device__ __forceinline__ void sweep(const DS& d, const MgGrid& g, Band<NX, CAP>& wb, bool valid,
                                      uint32_t gc0, f3 pi, float lim, XLoad xload, Visit visit,
                                      EndCell end_cell, uint32_t self = 0u,
                                      float rec_lim = 0.0f, bool hm_ok = true)
{
    constexpr int NXX = NX > 0 ? NX : 1;
    const uint32_t lane = __lane_id();
    const bool rows = d.rows != 0;
    const int nseg = rows ? g.srch * g.srch : g.adj_cnt;
    const int ncell = rows ? g.srch : 1;
    const bool cull = rows && d.cull;
    const uint32_t cmin = wave_min(valid ? gc0 : 0xFFFFFFFFu);
    const uint32_t cmax = wave_max(valid ? gc0 : 0u);
    if (cmin == 0xFFFFFFFFu) return; /* no lane sweeps */
    const uint32_t N = (uint32_t)d.N;

    /* position of x_i inside its own cell (cell units) and cell sizes, for the culling */
    f3 fr = mk3(0.f, 0.f, 0.f), inv = mk3(0.f, 0.f, 0.f);
    if (cull && valid) {
        const uint32_t own = gc0 + (uint32_t)((1 * g.res[2] + 1) * g.res[0] + 1);
        const uint32_t cx = own % (uint32_t)g.res[0];
        const uint32_t cz = (own / (uint32_t)g.res[0]) % (uint32_t)g.res[2];
        const uint32_t cy = own / ((uint32_t)g.res[0] * (uint32_t)g.res[2]);
        fr = mk3((pi.x - g.gmin[0]) * g.delta[0] - (float)cx, (pi.y - g.gmin[1]) * g.delta[1] - (float)cy,
                 (pi.z - g.gmin[2]) * g.delta[2] - (float)cz);
        inv = mk3(1.0f / g.delta[0], 1.0f / g.delta[1], 1.0f / g.delta[2]);
    }

    /* band of segment s (wave-uniform); fits == false selects the global path */
    auto band_of = [&](int s, uint32_t& blo, uint32_t& bhi) -> bool {
        if (!rows) return false;
        const uint32_t a = (uint32_t)g.adj[s * g.srch];
        blo = min(d.OFF[cmin + a], N);
        bhi = min(d.OFF[cmax + a + (uint32_t)g.srch], N);
        static_assert(CAP <= 128, "two slots per lane");
        return bhi >= blo && bhi - blo <= (uint32_t)CAP;
    };
    /* the next band, two slots per lane (lane and lane + 64) */
    float4 rp0 = make_float4(0.f, 0.f, 0.f, 0.f), rp1 = rp0, rx0[NXX], rx1[NXX];
#pragma unroll
    for (int x = 0; x < NXX; x++) rx0[x] = rx1[x] = rp0;
    auto fetch = [&](bool fits, uint32_t blo, uint32_t bhi) {
        if (!fits) return;
        const uint32_t j0 = blo + lane, j1 = j0 + 64u;
        if (j0 < bhi) {
            rp0 = d.P[j0];
#pragma unroll
            for (int x = 0; x < NX; x++) rx0[x] = xload(j0, x);
        }
        if (j1 < bhi) {
            rp1 = d.P[j1];
#pragma unroll
            for (int x = 0; x < NX; x++) rx1[x] = xload(j1, x);
        }
    };

    uint32_t blo = 0, bhi = 0;
    bool fits = band_of(0, blo, bhi);
    fetch(fits, blo, bhi);
    RowBounds nxt = row_bounds(d, g, valid ? gc0 : cmin, 0);
    /* HM == 1: record the candidates of each row within rec_lim (>= lim) and visit those
     * within lim (pressure); HM == 2: test only the recorded candidates (force, diffusion) */
    const bool hm_w = HM == 1 && rows && d.HMW != nullptr;
    const bool hm_r = HM == 2 && rows && d.HMR != nullptr && hm_ok;
    const float tlim = hm_w ? fmaxf(rec_lim, lim) : lim; /* candidate test and culling */
    uint64_t hm_next = ~0ull;
    if (hm_r && valid) hm_next = d.HMR[self];
    for (int s = 0; s < nseg; s++) {
        const RowBounds rb = nxt;
        const uint64_t hm_row = hm_next;
        if (hm_r && valid && s + 1 < nseg) hm_next = d.HMR[(size_t)(s + 1) * N + self];
        const uint32_t cblo = blo, cbhi = bhi;
        const bool cfits = fits;
        if (cfits) {
            wave_lds_sync(); /* the previous row's reads are done */
            if (cblo + lane < cbhi) {
                wb.p[lane] = rp0;
#pragma unroll
                for (int x = 0; x < NX; x++) wb.x[x][lane] = rx0[x];
            }
            if (cblo + lane + 64u < cbhi) {
                wb.p[lane + 64u] = rp1;
#pragma unroll
                for (int x = 0; x < NX; x++) wb.x[x][lane + 64u] = rx1[x];
            }
            wave_lds_sync();
        }
        if (s + 1 < nseg) {
            fits = band_of(s + 1, blo, bhi);
            fetch(fits, blo, bhi);
            nxt = row_bounds(d, g, valid ? gc0 : cmin, s + 1);
        }
        if (!valid) continue;
        /* the first cell boundary of the row above slot j */
        auto next_bound = [&](uint32_t j) -> uint32_t {
            uint32_t c = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 3; k >= 1; k--)
                if (k < ncell && j < rb.b[k]) c = rb.b[k];
            return c;
        };
        /* the recorded candidates m of this row (bit k = slot b[0] + k), ascending; a cell's
         * partial sum is flushed where the first hit past its end is visited and after the row
         * (the skipped flushes add +0 to sums that start at +0). One loop for the staged band
         * and one for global memory, chosen per row, so that the band's reads are LDS reads (a
         * select between the two would make them flat loads).
         * Rows where some lane holds more than 12 recorded candidates (compressed states)
         * issue the next candidate's reads before visiting the current one, so their latency
         * overlaps the visit; shorter rows take the plain loop, where that bookkeeping costs
         * more than it hides. */
#define MG_WALK_MASK(M, GETP, GETX)                                                            \
    do {                                                                                       \
        uint64_t m_ = (M);                                                                     \
        uint32_t cnext_ = next_bound(rb.b[0]);                                                 \
        while (m_ != 0ull) {                                                                   \
            const uint32_t j = rb.b[0] + (uint32_t)__builtin_ctzll(m_);                        \
            m_ &= m_ - 1ull;                                                                   \
            if (j >= cnext_) {                                                                 \
                end_cell();                                                                    \
                cnext_ = next_bound(j);                                                        \
            }                                                                                  \
            const float4 pj = GETP(j);                                                         \
            const f3 dist = pi - xyz(pj);                                                      \
            const float dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;             \
            if (!(dsq < lim && dsq > 0.0f)) continue; /* recorded within the skin only */      \
            float4 ex[NXX];                                                                    \
            _Pragma("unroll") for (int x = 0; x < NX; x++) ex[x] = GETX(j, x);                \
            visit(j, pj, dist, dsq, ex);                                                       \
        }                                                                                      \
        end_cell();                                                                            \
    } while (0)
#define MG_WALK_MASK_AHEAD(M, GETP, GETX)                                                      \
    do {                                                                                       \
        uint64_t m_ = (M);                                                                     \
        uint32_t cnext_ = next_bound(rb.b[0]);                                                 \
        bool has_ = m_ != 0ull;                                                                \
        uint32_t jn_ = 0u;                                                                     \
        float4 pn_ = make_float4(0.f, 0.f, 0.f, 0.f), xn_[NXX];                                \
        _Pragma("unroll") for (int x = 0; x < NXX; x++) xn_[x] = pn_;                          \
        if (has_) {                                                                            \
            jn_ = rb.b[0] + (uint32_t)__builtin_ctzll(m_);                                     \
            m_ &= m_ - 1ull;                                                                   \
            pn_ = GETP(jn_);                                                                   \
            _Pragma("unroll") for (int x = 0; x < NX; x++) xn_[x] = GETX(jn_, x);             \
        }                                                                                      \
        while (has_) {                                                                         \
            const uint32_t j = jn_;                                                            \
            const float4 pj = pn_;                                                             \
            float4 ex[NXX];                                                                    \
            _Pragma("unroll") for (int x = 0; x < NXX; x++) ex[x] = xn_[x];                    \
            has_ = m_ != 0ull;                                                                 \
            if (has_) {                                                                        \
                jn_ = rb.b[0] + (uint32_t)__builtin_ctzll(m_);                                 \
                m_ &= m_ - 1ull;                                                               \
                pn_ = GETP(jn_);                                                               \
                _Pragma("unroll") for (int x = 0; x < NX; x++) xn_[x] = GETX(jn_, x);         \
            }                                                                                  \
            if (j >= cnext_) {                                                                 \
                end_cell();                                                                    \
                cnext_ = next_bound(j);                                                        \
            }                                                                                  \
            const f3 dist = pi - xyz(pj);                                                      \
            const float dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;             \
            if (dsq < lim && dsq > 0.0f) visit(j, pj, dist, dsq, ex); /* recorded in the skin */ \
        }                                                                                      \
        end_cell();                                                                            \
    } while (0)
#define MG_BAND_P(j) wb.p[(j) - cblo]
#define MG_BAND_X(j, x) wb.x[x][(j) - cblo]
#define MG_GLOBAL_P(j) d.P[j]
#define MG_GLOBAL_X(j, x) xload(j, x)
        auto walk = [&](uint64_t m) {
            const bool ahead = __ballot(__builtin_popcountll(m) > 12) != 0ull;
            if (cfits) {
                if (ahead)
                    MG_WALK_MASK_AHEAD(m, MG_BAND_P, MG_BAND_X);
                else
                    MG_WALK_MASK(m, MG_BAND_P, MG_BAND_X);
            } else {
                if (ahead)
                    MG_WALK_MASK_AHEAD(m, MG_GLOBAL_P, MG_GLOBAL_X);
                else
                    MG_WALK_MASK(m, MG_GLOBAL_P, MG_GLOBAL_X);
            }
        };
#undef MG_WALK_MASK
#undef MG_WALK_MASK_AHEAD
#undef MG_BAND_P
#undef MG_BAND_X
#undef MG_GLOBAL_P
#undef MG_GLOBAL_X
        if (hm_r && hm_row != ~0ull) {
            walk(hm_row); /* the pressure sweep's record */
            continue;
        }
        /* window of the row for the hit record: [b[0], b[ncell]) */
        const uint32_t wlo = rb.b[0];
        const bool wide = !hm_w || (ncell >= 4 ? rb.b[4] : ncell == 3 ? rb.b[3] : ncell == 2 ? rb.b[2] : rb.b[1]) - wlo > 64u;
        uint64_t rec = 0ull;
        auto visit_rec = [&](uint32_t j, float4 pj, f3 dist, float dsq, const float4* x) {
            if (!wide) rec |= 1ull << ((j - wlo) & 63u);
            if (dsq < lim) visit(j, pj, dist, dsq, x);
        };
        float gyz = 0.0f;
        if (cull) {
            const float gy = axis_gap(fr.y, s / g.srch - 1, inv.y);
            const float gz = axis_gap(fr.z, s % g.srch - 1, inv.z);
            gyz = gy * gy + gz * gz;
        }
        auto live_of = [&](int k) -> bool {
            bool live = k < ncell;
            if (cull && live) {
                const float gx = axis_gap(fr.x, k - 1, inv.x);
                live = gyz + gx * gx < tlim;
            }
            return live;
        };
        /* sweeps that read recorded masks take the candidate path only for unrecorded rows,
         * and always per cell */
        bool over = false;
        if (HM != 2) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < ncell) over |= rb.b[k + 1] - rb.b[k] > 8u;
        }
        if (HM != 2 && __ballot(over) != 0ull) {
            /* a lane of the wavefront has a cell of more than 8 candidates: every lane walks
             * its live cells as one slot range, K candidates at a time, and flushes a cell's
             * partial sum at the first candidate past the cell's end and after the range
             * (fewer partly filled batches once cell occupancies vary; in a lattice of 8 per
             * cell the per-cell batches below fill exactly). The live cells are contiguous:
             * a cell's gap grows away from the own column (offset 0, gap 0). */
            bool any = false;
            uint32_t flo = 0u, fhi = 0u;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (!live_of(k)) continue;
                if (!any) flo = rb.b[k];
                any = true;
                fhi = max(rb.b[k + 1], rb.b[k]);
            }
            if (any) {
                const uint32_t lo = flo, hi = max(fhi, flo);
                uint32_t cnext = next_bound(lo);
                auto visit_flat = [&](uint32_t j, float4 pj, f3 dist, float dsq, const float4* x) {
                    if (j >= cnext) {
                        end_cell();
                        cnext = next_bound(j);
                    }
                    visit_rec(j, pj, dist, dsq, x);
                };
                if (cfits) {
                    cell_candidates<K, NX>(
                        lo, hi, pi, tlim, [&](uint32_t j) { return wb.p[j - cblo]; },
                        [&](uint32_t j, int x) { return wb.x[x][j - cblo]; }, visit_flat);
                } else {
                    cell_candidates<K, NX>(
                        lo, hi, pi, tlim, [&](uint32_t j) { return d.P[j]; },
                        [&](uint32_t j, int x) { return xload(j, x); }, visit_flat);
                }
                end_cell();
            }
            if (hm_w) d.HMW[(size_t)s * N + self] = wide ? ~0ull : rec;
            continue;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!live_of(k)) continue;
            const uint32_t lo = rb.b[k], hi = max(rb.b[k + 1], lo);
            if (cfits) {
                cell_candidates<K, NX>(
                    lo, hi, pi, tlim, [&](uint32_t j) { return wb.p[j - cblo]; },
                    [&](uint32_t j, int x) { return wb.x[x][j - cblo]; }, visit_rec);
            } else {
                cell_candidates<K, NX>(
                    lo, hi, pi, tlim, [&](uint32_t j) { return d.P[j]; },
                    [&](uint32_t j, int x) { return xload(j, x); }, visit_rec);
            }
            end_cell();
        }
        if (hm_w) d.HMW[(size_t)s * N + self] = wide ? ~0ull : rec;
    }
}

/* The bond terms of active slot i's live bond slots, each as the owner's force pass would
 * evaluate it (same operands: the step's positions and veval), recorded for that pass
 * (d.ETW; the record layout is described at DS::ETW). Slots in groups of 4, ascending, the
 * partners' slots of a group fetched together. */
__device__ __forceinline__ void record_bond_terms(const DS& d, const MgParams& p, uint32_t i)
{
    const uint32_t N = (uint32_t)d.N, na = d.ctr[MG_C_ACTIVE];
    const uint32_t h = d.HOME[i];
    const float4 P_i = d.P[i], V_i = d.V[i];
    const f3 ipos = xyz(P_i), ivev = xyz(V_i);
    uint64_t live = d.LIVE[h];
    while (live) {
        const int q = (__ffsll((long long)live) - 1) & ~3;
        const uint32_t g4 = (uint32_t)(live >> q) & 0xFu;
        live &= ~(0xFull << q);
        uint32_t jc[4];
        float4 bp[4], pj[4], vj[4];
        bool use[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const size_t bi = (size_t)(q + k) * N + h;
            jc[k] = (g4 & (1u << k)) ? d.SLOT[d.BPARTNER[bi]] : MG_NONE;
            bp[k] = (g4 & (1u << k)) ? d.BPAR[bi] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            use[k] = jc[k] < na && !((double)bp[k].y < 0.000000001);
            pj[k] = vj[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (use[k]) {
                pj[k] = d.P[jc[k]];
                vj[k] = d.V[jc[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!(g4 & (1u << k))) continue;
            float4 r = make_float4(0.f, 0.f, 0.f, __builtin_nanf(""));
            if (use[k]) {
                float strain, abs_dist;
                const f3 eterm = bond_term(ipos, ivev, xyz(pj[k]), xyz(vj[k]), bp[k], p.pmass, &strain, &abs_dist);
                const uint32_t brk = abs_dist >= bp[k].x ? 0x80000000u : 0u;
                r = make_float4(eterm.x, eterm.y, eterm.z, __uint_as_float(__float_as_uint(strain) | brk));
            }
            d.ETW[(size_t)(q + k) * N + h] = r;
        }
    }
}

/* ------------------------------------------------------------ pressure (a9)
 * contributePressure/computePressure, src/fluid_system_cuda.cu:620-712. Per cell partial
 * sums are added to the running total in gridAdj order. Writes FPRESS into P.w and
 * FDENSITY (= 1/density) into V.w. Also clears d.ZERO (the next grid's counters) when set. */
template <int K = 8, int CAP = BAND_CAP, int MODE = MG_SPH_REFERENCE>
__device__ __forceinline__ void pressure_body(const DS& d, const MgParams& p, const MgGrid& g,
                                              Band<0, CAP>* bands, uint32_t bid, uint32_t nblk)
{
    if (d.ZERO) {
        for (uint32_t z = bid * blockDim.x + threadIdx.x; z < d.zero_n; z += nblk * blockDim.x)
            d.ZERO[z] = 0u;
    }
    /* a new set of masks: the force-advance of this step raises the flag again if needed */
    if (d.HMW && bid == 0 && threadIdx.x == 0) d.ctr[MG_C_SKIN] = 0u;
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    uint32_t gc = (i < d.ctr[MG_C_ACTIVE]) ? d.GCELL[i] : MG_GRID_UNDEF;
    const bool valid = gc != MG_GRID_UNDEF;
    gc -= (uint32_t)((1 * g.res[2] + 1) * g.res[0] + 1);
    const f3 pi = valid ? xyz(d.P[i]) : mk3(0.f, 0.f, 0.f);
    const float sr = p.psmoothradius;
    float sum = 0.0f, csum = 0.0f;
    auto flush = [&]() {
        sum += csum;
        csum = 0.0f;
    };
    auto none = [](uint32_t, int) { return make_float4(0.f, 0.f, 0.f, 0.f); };
    /* r / sr for a power-of-two sr is r * (1 / sr) exactly (a scaling by 2^-k), so that case
     * skips the correctly rounded division; the sums are the same bits either way */
    const uint32_t srb = __float_as_uint(sr);
    const bool sr_pow2 = (srb & 0x007FFFFFu) == 0u && ((srb >> 23) & 0xFFu) != 0u && ((srb >> 23) & 0xFFu) != 0xFFu;
    if (MODE == MG_SPH_POLY6_SPIKY) {
        /* the poly6 sum the reference keeps in comments (src/fluid_system_cuda.cu:668-672, 695) */
        const float r2 = p.r2, d2 = p.d2;
        sweep<K, 0, 1, CAP>(
            d, g, bands[threadIdx.x >> 6], valid, gc, pi, p.r2, none,
            [&](uint32_t, float4, f3, float dsq, const float4*) {
                const float c = (r2 - dsq) * d2;
                csum += c * c * c;
            },
            flush, i, MG_SKIN_LIM);
    } else if (sr_pow2) {
        const float inv_sr = 1.0f / sr;
        sweep<K, 0, 1, CAP>(
            d, g, bands[threadIdx.x >> 6], valid, gc, pi, p.r2, none,
            [&](uint32_t, float4, f3, float dsq, const float4*) {
                float r = sqrt_rn(dsq);
                float q = r * inv_sr;
                float b = 1.0f - q;
                b *= b;
                b *= b;
                csum += b * (4.0f * q + 1.0f);
            },
            flush, i, MG_SKIN_LIM);
    } else {
        sweep<K, 0, 1, CAP>(
            d, g, bands[threadIdx.x >> 6], valid, gc, pi, p.r2, none,
            [&](uint32_t, float4, f3, float dsq, const float4*) {
                float r = sqrt_rn(dsq);
                float q = r / sr;
                float b = 1.0f - q;
                b *= b;
                b *= b;
                csum += b * (4.0f * q + 1.0f);
            },
            flush, i, MG_SKIN_LIM);
    }
    if (!valid) return;
    sum = sum * p.pmass * (MODE == MG_SPH_POLY6_SPIKY ? p.poly6kern : p.wendlandC2kern);
    if (sum == 0.0f) sum = 1.0f;
    ((float*)&d.P[i])[3] = (sum - p.prest_dens) * p.pintstiff;
    ((float*)&d.V[i])[3] = 1.0f / sum;
    if (d.ETW) record_bond_terms(d, p, i);
}

/* the sweep kernels take (blockDim.x / 64) Band<NX> of dynamic LDS, one per wavefront;
 * MODE selects the SPH kernels (MG_SPH_*) */
template <int MODE>
__global__ LAUNCH_BOUNDS_SWEEP void k_pressure(DS d, MgParams p, MgGrid g)
{
    extern __shared__ float4 mg_lds[];
    pressure_body<8, BAND_CAP, MODE>(d, p, g, (Band<0>*)mg_lds, xcd_block(), gridDim.x);
}

/* --------------------------------------------------------------- force (a10)
 * contributeForce/computeForce, src/fluid_system_cuda.cu:3167-3358, for active particle i in
 * cell gc (not GRID_UNDEF). The owner subtracts its bond terms in slot order and adds pmass *
 * (sum over cells of the per-cell fluid sums); the atomicAdd scatter of bond terms to
 * partners is replaced by each particle gathering the terms of its incoming bonds (same
 * expression, same operands, hence the same bits) in the incoming CSR's order (owner particle
 * ID, owner home, slot):
 *   FFORCE[i] = ((F0 + own) + in_1) + in_2 ...
 * Updates the bond integrators and breaks bonds as the reference does. The live bond slots
 * (LIVE, kept with the bond records) are fetched 4 at a time in slot order; incoming bonds
 * are fetched 4 at a time; in both loops the next group's records are fetched while the
 * current group's partner positions load. */
template <int MODE = MG_SPH_REFERENCE>
__device__ __forceinline__ f3 force_one(const DS& d, const MgParams& p, const MgGrid& g, Band<1>& wb,
                                        bool valid, uint32_t i, uint32_t na, uint32_t gc, f3 F0)
{
    gc -= (uint32_t)((1 * g.res[2] + 1) * g.res[0] + 1);
    const int N = d.N, B = d.B;
    uint32_t h = 0, e0 = 0, e1 = 0;
    float4 P_i = make_float4(0.f, 0.f, 0.f, 0.f), V_i = P_i;
    if (valid) {
        h = d.HOME[i];
        e0 = d.INC_OFF[h];
        e1 = d.INC_OFF[h + 1];
        P_i = d.P[i];
        V_i = d.V[i];
    }
    const f3 ipos = xyz(P_i), ivev = xyz(V_i);
    const float ipress = P_i.w, idens = V_i.w;
    const bool freeze = p.freeze != 0;

    f3 force = mk3(0.0f, 0.0f, 0.0f);
    /* live slots in groups of 4, ascending; a group's partner slots and records are fetched
     * while the previous group's partner positions load (software pipeline). The own bond
     * terms are summed after the fluid sweep (into their own total, as before it), and the
     * first group's partner slots are fetched before the sweep. */
    uint64_t live = valid ? d.LIVE[h] : 0ull;
    int q = 0;
    uint32_t g4 = 0, j[4];
    auto next_group = [&]() {
        g4 = 0;
        if (live) {
            q = (__ffsll((long long)live) - 1) & ~3;
            g4 = (uint32_t)(live >> q) & 0xFu;
            live &= ~(0xFull << q);
        }
        if (d.ETR) return; /* the records replace the partners' positions */
#pragma unroll
        for (int k = 0; k < 4; k++) j[k] = (g4 & (1u << k)) ? d.SLOT[d.BPARTNER[(size_t)(q + k) * N + h]] : MG_NONE;
    };
    next_group();

    uint32_t first[4] = {0u, 0u, 0u, 0u}; /* the first 4 incoming entries */
    if (valid && e0 < e1)
#pragma unroll
        for (int k = 0; k < 4; k++) first[k] = d.INC[min(e0 + (uint32_t)k, e1 - 1u)];
    /* fluid terms: per-cell sums added in gridAdj order */
    const float sr = p.psmoothradius, st = p.psurface_t;
    f3 fsum = mk3(0.0f, 0.0f, 0.0f), cf = mk3(0.0f, 0.0f, 0.0f);
    /* the reference's force tests dsq < 1 (src/fluid_system_cuda.cu:3203); the Mueller kernels
     * have the smoothing radius as support */
    const float flim = MODE == MG_SPH_POLY6_SPIKY ? p.r2 : 1.0f;
    sweep<4, 1, 2>(
        d, g, wb, valid, gc, ipos, flim, [&](uint32_t j, int) { return d.V[j]; },
        [&](uint32_t, float4 Pj, f3 dist, float dsq, const float4* x) {
            const float4 Vj = x[0];
            if (MODE == MG_SPH_POLY6_SPIKY) {
                cf = cf + fluid_term_poly6(dist, dsq, p, ipress, Pj.w, idens, Vj.w, ivev, xyz(Vj));
            } else {
                float r, inv_r;
                sqrt_rcp_rn(dsq, r, inv_r);
                cf = cf + fluid_term(dist, r, inv_r, sr, ipress, Pj.w, idens, Vj.w, ivev, xyz(Vj), st);
            }
        },
        [&]() {
            fsum = fsum + cf;
            cf = mk3(0.0f, 0.0f, 0.0f);
        },
        i);

    float4 bp[4];
    float2 in[4];
    if (d.ETR) {
        /* the pressure pass's records of the own live slots: integrators, force, breaking in
         * the order below (the same operations on the same values as the evaluation) */
        while (g4) {
            const int qc = q;
            const uint32_t gc4 = g4;
            float4 r[4];
            float2 inc[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                r[k] = make_float4(0.f, 0.f, 0.f, __builtin_nanf(""));
                inc[k] = make_float2(0.f, 0.f);
                if (gc4 & (1u << k)) {
                    const size_t bi = (size_t)(qc + k) * N + h;
                    r[k] = d.ETR[bi];
                    inc[k] = d.BINT[bi];
                }
            }
            next_group();
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (r[k].w != r[k].w) continue; /* not a live slot, or one the force skips */
                const size_t bi = (size_t)(qc + k) * N + h;
                const uint32_t wbits = __float_as_uint(r[k].w);
                const float strain = __uint_as_float(wbits & 0x7FFFFFFFu);
                if (!d.dry) {
                    float2 v = inc[k];
                    v.x = v.x + strain * strain;
                    v.y = v.y + strain;
                    d.BINT[bi] = v;
                }
                const f3 eterm = mk3(r[k].x, r[k].y, r[k].z);
                if (!has_nan3(eterm)) force = force - eterm;
                if ((wbits >> 31) && !freeze && !d.dry) ((float*)&d.BPAR[bi])[1] = 0.0f;
            }
        }
    }
    auto records = [&]() {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bp[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            in[k] = make_float2(0.f, 0.f);
            if (g4 & (1u << k)) {
                const size_t bi = (size_t)(q + k) * N + h;
                bp[k] = d.BPAR[bi];
                in[k] = d.BINT[bi];
            }
        }
    };
    records();
    while (g4) {
        const int qc = q;
        uint32_t jc[4];
        float4 bpc[4];
        float2 inc[4];
        bool use[4];
        float4 pj[4], vj[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            jc[k] = j[k];
            bpc[k] = bp[k];
            inc[k] = in[k];
            use[k] = jc[k] < na && !((double)bpc[k].y < 0.000000001);
            pj[k] = vj[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (use[k]) {
                pj[k] = d.P[jc[k]];
                vj[k] = d.V[jc[k]];
            }
        }
        next_group();
        records();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!use[k]) continue;
            const size_t bi = (size_t)(qc + k) * N + h;
            float strain, abs_dist;
            f3 eterm = bond_term(ipos, ivev, xyz(pj[k]), xyz(vj[k]), bpc[k], p.pmass, &strain, &abs_dist);
            if (!d.dry) {
                float2 v = inc[k];
                v.x = v.x + strain * strain;
                v.y = v.y + strain;
                d.BINT[bi] = v;
            }
            if (!has_nan3(eterm)) force = force - eterm;
            if (abs_dist >= bpc[k].x && !freeze && !d.dry) ((float*)&d.BPAR[bi])[1] = 0.0f;
        }
    }
    force = force + fsum * p.pmass;

    f3 F = F0 + force;
    /* incoming bonds 4 at a time, the next 4 fetched while this group's positions load (the
     * first 4 entries are read before the fluid sweep) */
    uint32_t kk[4];
    auto fetch_in = [&](const uint32_t packed[4]) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t kh = packed[k] / (uint32_t)B, b = packed[k] - kh * (uint32_t)B;
            kk[k] = d.SLOT[kh];
            bp[k] = d.BPAR[(size_t)b * N + kh];
        }
    };
    auto entries = [&](uint32_t e, uint32_t packed[4]) {
#pragma unroll
        for (int k = 0; k < 4; k++) packed[k] = d.INC[min(e + (uint32_t)k, e1 - 1u)];
    };
    if (d.ETR) {
        /* the owners' records of the incoming bonds (an inactive owner's record is stale: its
         * slot decides, as in the evaluation) */
        uint32_t packed[4] = {first[0], first[1], first[2], first[3]};
        for (uint32_t e = e0; valid && e < e1; e += 4) {
            uint32_t ks[4];
            float4 r[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t kh = packed[k] / (uint32_t)B, b = packed[k] - kh * (uint32_t)B;
                ks[k] = d.SLOT[kh];
                r[k] = d.ETR[(size_t)b * N + kh];
            }
            if (e + 4u < e1) entries(e + 4u, packed);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (!(e + (uint32_t)k < e1 && ks[k] < na && r[k].w == r[k].w)) continue;
                const f3 eterm = mk3(r[k].x, r[k].y, r[k].z);
                if (!has_nan3(eterm)) F = F + eterm;
            }
        }
        return F;
    }
    if (valid && e0 < e1) fetch_in(first);
    for (uint32_t e = e0; valid && e < e1; e += 4) {
        bool use[4];
        float4 bpc[4], pk[4], vk[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bpc[k] = bp[k];
            use[k] = e + (uint32_t)k < e1 && kk[k] < na && !((double)bpc[k].y < 0.000000001);
            pk[k] = vk[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (use[k]) {
                pk[k] = d.P[kk[k]];
                vk[k] = d.V[kk[k]];
            }
        }
        if (e + 4u < e1) {
            uint32_t packed[4];
            entries(e + 4u, packed);
            fetch_in(packed);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!use[k]) continue;
            float strain, abs_dist;
            f3 eterm = bond_term(xyz(pk[k]), xyz(vk[k]), ipos, ivev, bpc[k], p.pmass, &strain, &abs_dist);
            if (!has_nan3(eterm)) F = F + eterm;
        }
    }
    return F;
}

/* ComputeForceCL: FFORCE[i] += force (accumulate) or = 0 + force. */
template <int MODE>
__global__ LAUNCH_BOUNDS_SWEEP void k_force(DS d, MgParams p, MgGrid g, int accumulate)
{
    extern __shared__ float4 mg_lds[];
    Band<1>* bands = (Band<1>*)mg_lds;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    const uint32_t gc = (i < na) ? d.GCELL[i] : MG_GRID_UNDEF;
    const bool valid = gc != MG_GRID_UNDEF;
    const f3 F0 = (valid && accumulate) ? xyz(d.F[i]) : mk3(0.0f, 0.0f, 0.0f);
    const f3 F = force_one<MODE>(d, p, g, bands[threadIdx.x >> 6], valid, i, na, gc, F0);
    if (valid) d.F[i] = make_float4(F.x, F.y, F.z, 0.0f);
}

/* ------------------------------------------------------------- advance (a11)
 * advanceParticles, src/fluid_system_cuda.cu:3470-3614, for one particle, in two parts:
 * advance_accel is the acceleration the step applies (FFORCE plus the wall penalties, gravity
 * and the acceleration limit, :3497-3576), advance_one the (no-op) velocity limit and the
 * leapfrog update that return the new position, veval and velocity. */
__device__ __forceinline__ f3 advance_accel(const MgParams& p, f3 pos, f3 vev, f3 accel)
{
    const float eps = MG_EPSILON, prad = p.pradius, slope = p.pground_slope;
    const float estiff = p.pextstiff, damp = p.pdamp;
    const f3 bmin = mk3(p.pboundmin[0], p.pboundmin[1], p.pboundmin[2]);
    const f3 bmax = mk3(p.pboundmax[0], p.pboundmax[1], p.pboundmax[2]);
    f3 norm;
    float diff, adj, speed;
    diff = prad - (pos.y - (bmin.y + (pos.x - bmin.x) * slope));
    if (diff > eps) {
        norm = mk3(-slope, (float)(1.0 - (double)slope), 0.0f);
        adj = estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    diff = prad - (bmax.y - pos.y);
    if (diff > eps) {
        norm = mk3(0.0f, -1.0f, 0.0f);
        adj = estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    diff = prad - (pos.x - bmin.x);
    if (diff > eps) {
        norm = mk3(1.0f, 0.0f, 0.0f);
        adj = (p.pforce_min + 1.0f) * estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    diff = prad - (bmax.x - pos.x);
    if (diff > eps) {
        norm = mk3(-1.0f, 0.0f, 0.0f);
        adj = (p.pforce_max + 1.0f) * estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    diff = prad - (pos.z - bmin.z);
    if (diff > eps) {
        norm = mk3(0.0f, 0.0f, 1.0f);
        adj = estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    diff = prad - (bmax.z - pos.z);
    if (diff > eps) {
        norm = mk3(0.0f, 0.0f, -1.0f);
        adj = estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    f3 dist = bmax - pos;
    diff = 2.0f * prad - (dist.x + dist.y + dist.z);
    if (diff > eps) {
        norm = mk3(1.0f, 1.0f, 1.0f);
        adj = estiff * diff - damp * dot3(norm, vev);
        norm = norm * adj; accel = accel + norm;
    }
    accel = accel + mk3(p.pgravity[0], p.pgravity[1], p.pgravity[2]);
    speed = accel.x * accel.x + accel.y * accel.y + accel.z * accel.z;
    if (speed > p.AL2) accel = accel * (p.AL / sqrtf(speed));
    return accel;
}

/* The rest of advanceParticles (src/fluid_system_cuda.cu:3577-3600). A learned step's
 * corrections (corr[0].w != 0; two float4: drift c_d, veval c_e) move the position by
 * (v_next + c_d) dt and store veval (v + v_next) / 2 + c_e; the velocity stays v_next. */
__device__ __forceinline__ void advance_one(const MgParams& p, float dt, f3 pos, f3 vev, f3 v,
                                            f3 accel, f3* np_out, f3* ve_out, f3* vnext_out,
                                            const float4* corr = nullptr)
{
    accel = advance_accel(p, pos, vev, accel);
    float speed = v.x * v.x + v.y * v.y + v.z * v.z;
    if (speed > p.VL2) {
        speed = p.VL2;
        v = v * (p.VL / sqrtf(speed)); /* no-op limit, src/fluid_system_cuda.cu:3577-3579 */
    }
    const f3 vnext = accel * dt + v;
    if (corr && corr[0].w != 0.0f) { /* learned step */
        *np_out = pos + (vnext + xyz(corr[0])) * dt;
        *ve_out = (v + vnext) * 0.5f + xyz(corr[1]);
        *vnext_out = vnext;
        return;
    }
    *ve_out = (v + vnext) * 0.5f;
    *np_out = pos + vnext * dt;
    *vnext_out = vnext;
}

/* AdvanceCL in place. With the copies of src/host_CL.cpp:199-226 around it, a GRID_UNDEF
 * particle ends unchanged. */
/* externalActuation + fixedParticles (src/fluid_system_cuda.cu:3617-3705) on one advanced
 * particle: gene 12's list members keep their pre-advance state with z moved by zoff, gene
 * 11's (applied last, so they win) keep it unchanged. */
__device__ __forceinline__ void special_one(const DS& d, uint32_t i, float zoff, f3 pos, f3 vev, f3 v,
                                            f3* np, f3* ve, f3* vn)
{
    const uint32_t m = d.GMASK[i];
    if (m & (1u << 12)) {
        *np = mk3(pos.x, pos.y, pos.z + zoff);
        *ve = vev;
        *vn = v;
    }
    if (m & (1u << 11)) {
        *np = pos;
        *ve = vev;
        *vn = v;
    }
}

__global__ LAUNCH_BOUNDS_256 void k_advance(DS d, MgParams p, float dt, int special, float zoff)
{
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= d.ctr[MG_C_ACTIVE]) return;
    if (d.GCELL[i] == MG_GRID_UNDEF) return;
    const float4 P_i = d.P[i], V_i = d.V[i];
    f3 np, ve, vn;
    advance_one(p, dt, xyz(P_i), xyz(V_i), xyz(d.VEL[i]), xyz(d.F[i]), &np, &ve, &vn);
    if (special) special_one(d, i, zoff, xyz(P_i), xyz(V_i), xyz(d.VEL[i]), &np, &ve, &vn);
    d.P[i] = make_float4(np.x, np.y, np.z, P_i.w);
    d.V[i] = make_float4(ve.x, ve.y, ve.z, V_i.w);
    d.VEL[i] = make_float4(vn.x, vn.y, vn.z, 0.0f);
}

/* Advance slot i with acceleration F into the other half of the ping-pong (P2/V2/VEL2), so
 * neighbours keep reading the pre-step state; an inactive slot copies through. With d.DVc (a
 * learned step) the slot's corrections, when marked used, enter the update. With
 * d.HMR (hit masks in use) a displacement beyond the masks' skin raises MG_C_SKIN. */
__device__ __forceinline__ void advance_other_half(const DS& d, const MgParams& p, float dt, uint32_t i,
                                                   bool valid, f3 F, int special, float zoff)
{
    const float4 P_i = d.P[i], V_i = d.V[i], VEL_i = d.VEL[i];
    if (!valid) {
        d.P2[i] = P_i;
        d.V2[i] = V_i;
        d.VEL2[i] = VEL_i;
        return;
    }
    f3 np, ve, vn;
    if (d.DVc) {
        const float4 c[2] = {d.DVc[2 * (size_t)i], d.DVc[2 * (size_t)i + 1]};
        advance_one(p, dt, xyz(P_i), xyz(V_i), xyz(VEL_i), F, &np, &ve, &vn, c);
    } else {
        advance_one(p, dt, xyz(P_i), xyz(V_i), xyz(VEL_i), F, &np, &ve, &vn);
    }
    if (special) special_one(d, i, zoff, xyz(P_i), xyz(V_i), xyz(VEL_i), &np, &ve, &vn);
    if (d.HMR) {
        /* the masks serve the diffusion sweep on the advanced positions only while every
         * particle stays inside half the skin (NaN and inf displacements raise the flag too) */
        const f3 dx = np - xyz(P_i);
        if (!(dot3(dx, dx) <= MG_SKIN_DISP2)) atomicOr(&d.ctr[MG_C_SKIN], 1u);
    }
    d.P2[i] = make_float4(np.x, np.y, np.z, P_i.w);
    d.V2[i] = make_float4(ve.x, ve.y, ve.z, V_i.w);
    d.VEL2[i] = make_float4(vn.x, vn.y, vn.z, 0.0f);
}

/* The learned coarse step's advance: FFORCE from this step's force pass (k_force), the
 * network's corrections (DVc), the advanced state into the other half. */
__global__ LAUNCH_BOUNDS_256 void k_advance_learned(DS d, MgParams p, float dt, int special, float zoff)
{
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)d.N) return;
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    const bool valid = i < na && d.GCELL[i] != MG_GRID_UNDEF;
    advance_other_half(d, p, dt, i, valid, valid ? xyz(d.F[i]) : mk3(0.f, 0.f, 0.f), special, zoff);
}

/* Fused ComputeForceCL + AdvanceCL for a step that starts from a sort (FFORCE = 0): the force
 * stays in registers and the advanced state goes to the other half of the ping-pong
 * (P2/V2/VEL2) so neighbours keep reading the pre-step state. Inactive slots copy through. */
template <int MODE = MG_SPH_REFERENCE>
__device__ __forceinline__ void force_advance_body(const DS& d, const MgParams& p, const MgGrid& g,
                                                   Band<1>* bands, float dt, uint32_t bid,
                                                   int special = 0, float zoff = 0.0f)
{
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    const uint32_t gc = (i < na) ? d.GCELL[i] : MG_GRID_UNDEF;
    const bool valid = gc != MG_GRID_UNDEF;
    const f3 F = force_one<MODE>(d, p, g, bands[threadIdx.x >> 6], valid, i, na, gc, mk3(0.0f, 0.0f, 0.0f));
    if (i >= (uint32_t)d.N) return;
    advance_other_half(d, p, dt, i, valid, F, special, zoff);
}

template <int MODE>
__global__ LAUNCH_BOUNDS_SWEEP void k_force_advance(DS d, MgParams p, MgGrid g, float dt, int special,
                                                    float zoff)
{
    extern __shared__ float4 mg_lds[];
    force_advance_body<MODE>(d, p, g, (Band<1>*)mg_lds, dt, xcd_block(), special, zoff);
}

/* ----------------------------------------------------------- diffusion (a12)
 * contributeDiffusion/computeDiffusion, src/fluid_system_cuda.cu:3026-3105, as a Jacobi update
 * of particle i (cell gc, not GRID_UNDEF) from the concentrations CONC (read-only). */
__device__ __forceinline__ void load16(const float* src, float v[16])
{
    const float4* s4 = (const float4*)src;
    float4 a = s4[0], b = s4[1], c = s4[2], e = s4[3];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    v[8] = c.x; v[9] = c.y; v[10] = c.z; v[11] = c.w;
    v[12] = e.x; v[13] = e.y; v[14] = e.z; v[15] = e.w;
}

__device__ __forceinline__ void store16(float* dst, const float v[16])
{
    float4* d4 = (float4*)dst;
    d4[0] = make_float4(v[0], v[1], v[2], v[3]);
    d4[1] = make_float4(v[4], v[5], v[6], v[7]);
    d4[2] = make_float4(v[8], v[9], v[10], v[11]);
    d4[3] = make_float4(v[12], v[13], v[14], v[15]);
}

/* The diffusion sweep touches only the float4 groups of FCONC that hold a diffusing TF. The
 * host prepares, from the genome and diff_scale:
 *   ng        groups used (0 when no TF diffuses; 3 is padded to all 4); a kernel
 *             instantiated for NG >= ng groups also runs the plan (slots q >= ng hold no
 *             diffusing TF and are ignored),
 *   grp[q]    the FCONC float4 group of slot q,
 *   dm        bit 4q+e set when TF 4*grp[q]+e diffuses,
 *   D[4q+e]   (float)tf_diffusability[4*grp[q]+e] * diff_scale. */
struct DiffPlan {
    int ng;
    int grp[4];
    uint32_t dm;
    float D[MG_NUM_TF];
};

/* nw[t] for the TFs of the plan's groups: sum over neighbours j within r2 (slot order) of
 * D * ((r2 - dsq) * simscale^2) * (conc_j - conc_i). Called by all lanes of the wavefront;
 * lanes with valid == false return zeros. */
template <int NG>
__device__ __forceinline__ void diffuse_sweep(const DS& d, const MgParams& p, const MgGrid& g,
                                              const DiffPlan& dp, Band<NG>& wb, bool valid, uint32_t i,
                                              uint32_t gc, float nw[4 * NG])
{
    gc -= (uint32_t)((1 * g.res[2] + 1) * g.res[0] + 1);
    const float d2 = p.psimscale * p.psimscale;
    const float r2 = p.r2 / d2;
    float cur[4 * NG];
#pragma unroll
    for (int t = 0; t < 4 * NG; t++) cur[t] = 0.0f;
    f3 pi = mk3(0.f, 0.f, 0.f);
    if (valid) {
        const float4* ci = (const float4*)(d.CONC + (size_t)i * MG_NUM_TF);
#pragma unroll
        for (int q = 0; q < NG; q++) {
            const float4 v = ci[dp.grp[q]];
            cur[4 * q] = v.x; cur[4 * q + 1] = v.y; cur[4 * q + 2] = v.z; cur[4 * q + 3] = v.w;
        }
        pi = xyz(d.P[i]);
    }
#pragma unroll
    for (int t = 0; t < 4 * NG; t++) nw[t] = 0.0f;
    /* the step's hit masks (d.HMR, recorded within the skin before the advance) stand for
     * the candidate tests unless a particle moved too far (MG_C_SKIN): either way the visits
     * are the candidates within r2 at the advanced positions, in the same order */
    const bool hm_ok = d.HMR != nullptr && d.ctr[MG_C_SKIN] == 0u && r2 == 1.0f;
    sweep<(NG >= 4 ? 1 : (NG >= 2 ? 2 : 4)), NG, 2>(
        d, g, wb, valid, gc, pi, r2,
        [&](uint32_t j, int q) { return ((const float4*)(d.CONC + (size_t)j * MG_NUM_TF))[dp.grp[q]]; },
        [&](uint32_t, float4, f3, float dsq, const float4* x) {
            const float cc = (r2 - dsq) * d2;
            const float* cn = (const float*)x;
#pragma unroll
            for (int t = 0; t < 4 * NG; t++)
                if (dp.dm & (1u << t)) nw[t] += dp.D[t] * cc * (cn[t] - cur[t]);
        },
        []() {}, i, 0.0f, hm_ok);
}

/* contributeDiffusion/computeDiffusion, src/fluid_system_cuda.cu:3026-3105, as a Jacobi update
 * of particle i (cell gc, not GRID_UNDEF) from the concentrations CONC (read-only):
 *   out[t] = tf_breakdown_rate[t] * (cur[t] + nw[t]),  nw[t] = +0 for TFs that do not diffuse.
 * Called by all lanes of the wavefront; out is written for valid lanes only. */
template <int NG, class Gen>
__device__ __forceinline__ void diffuse_one(const DS& d, const MgParams& p, const MgGrid& g,
                                            const Gen& gen, const DiffPlan& dp, Band<NG>& wb,
                                            bool valid, uint32_t i, uint32_t gc, float out[MG_NUM_TF])
{
    float nwf[MG_NUM_TF];
#pragma unroll
    for (int t = 0; t < MG_NUM_TF; t++) nwf[t] = 0.0f;
    if (dp.ng > 0) {
        float nw[4 * NG];
        diffuse_sweep<NG>(d, p, g, dp, wb, valid, i, gc, nw);
#pragma unroll
        for (int q = 0; q < NG; q++)
#pragma unroll
            for (int t = 0; t < MG_NUM_TF; t++)
                if (q < dp.ng && (t >> 2) == dp.grp[q]) nwf[t] = nw[4 * q + (t & 3)];
    }
    if (!valid) return;
    float cur[MG_NUM_TF];
    load16(d.CONC + (size_t)i * MG_NUM_TF, cur);
#pragma unroll
    for (int t = 0; t < MG_NUM_TF; t++) out[t] = (float)gen.tf_breakdown_rate[t] * (cur[t] + nwf[t]);
}

/* ComputeDiffusionCL: CONC -> CONC2 for every slot (slots that do not diffuse copy through). */
template <int NG>
__global__ LAUNCH_BOUNDS_SWEEP void k_diffusion(DS d, MgParams p, MgGrid g, MgGenome gen, DiffPlan dp)
{
    extern __shared__ float4 mg_lds[];
    Band<NG>* bands = (Band<NG>*)mg_lds;
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t gc = (i < d.ctr[MG_C_ACTIVE]) ? d.GCELL[i] : MG_GRID_UNDEF;
    const bool valid = gc != MG_GRID_UNDEF;
    float out[MG_NUM_TF];
    diffuse_one<NG>(d, p, g, gen, dp, bands[threadIdx.x >> 6], valid, i, gc, out);
    if (i >= (uint32_t)d.N) return;
    if (!valid) load16(d.CONC + (size_t)i * MG_NUM_TF, out);
    store16(d.CONC2 + (size_t)i * MG_NUM_TF, out);
}

/* --------------------------------------------------------------- genes (a13)
 * computeGeneAction + tallyGeneAction (src/fluid_system_cuda.cu:714-772) for one particle with
 * sort-time membership m (the dense lists): genes in order 0..15, secretion/activation deltas
 * in registers, then the tally when the particle is a member of the living list (gene 2)
 * below mActivePoints. Fixed semantics are those of the oracle's genes() (float FCONC in
 * AoS, one tally, bounded countdown). Updates epi[] and conc[] in place. */
template <class Gen>
__device__ __forceinline__ void genes_one(const MgParams& p, const Gen& gen, uint32_t m,
                                          bool tally, uint32_t epi[MG_NUM_GENES],
                                          float conc[MG_NUM_TF])
{
    float dconc[MG_NUM_TF];
    uint32_t depi[MG_NUM_GENES];
#pragma unroll
    for (int t = 0; t < MG_NUM_TF; t++) { dconc[t] = 0.0f; depi[t] = 0u; }
    for (int gene = 0; gene < MG_NUM_GENES; gene++) {
        if (!(m & (1u << gene))) continue;
        const int delay = (int)epi[gene];
        if (0 < delay && delay < MG_INT_MAX) {
            epi[gene]--;
            if (delay == 1 && gene + 1 < MG_NUM_GENES && epi[gene + 1]) epi[gene + 1] = gen.delay[gene + 1];
        }
        float activity = 0.0f;
        for (int tf = 0; tf < MG_NUM_TF; tf++)
            if (gen.sensitivity[gene][tf] != 0u) activity += (float)gen.sensitivity[gene][tf] * conc[tf];
        const int numTF = gen.secrete[gene][2 * MG_NUM_TF];
        for (int j = 0; j < numTF && j < MG_NUM_TF; j++) {
            const int tf = gen.secrete[gene][j * 2];
            const int rate = gen.secrete[gene][j * 2 + 1];
            if (tf < 0 || tf >= MG_NUM_TF) continue;
            dconc[tf] += ((float)rate * activity) * p.secrete_scale;
        }
        const int numL = gen.activate[gene][2 * MG_NUM_GENES];
        for (int j = 0; j < numL && j < MG_NUM_GENES; j++) {
            const int other = gen.activate[gene][j * 2];
            const int threshold = gen.activate[gene][j * 2 + 1];
            if (other < 0 || other >= MG_NUM_GENES) continue;
            if ((float)threshold < activity) depi[other] += 1u;
        }
    }
    if (tally) {
        for (int t = 0; t < MG_NUM_TF; t++) conc[t] = conc[t] + dconc[t];
        for (int g = 0; g < MG_NUM_GENES; g++) epi[g] += depi[g];
    }
}

/* ComputeGenesCL in place (EPIGEN by home, CONC by slot). */
__global__ LAUNCH_BOUNDS_256 void k_genes(DS d, MgParams p, MgGenome gen)
{
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    if (!d.ctr[MG_C_LISTS]) return;
    if (s >= d.ctr[MG_C_NUM_POINTS]) return;
    const uint32_t m = d.GMASK[s];
    if (m == 0) return;
    const uint32_t h = d.HOME[s];
    uint32_t epi[MG_NUM_GENES];
#pragma unroll
    for (int g = 0; g < MG_NUM_GENES; g++) epi[g] = d.EPIGEN[(size_t)g * d.N + h];
    float conc[MG_NUM_TF];
    float* cp = d.CONC + (size_t)s * MG_NUM_TF;
    load16(cp, conc);
    const bool tally = (m & 4u) && s < d.ctr[MG_C_ACTIVE];
    uint32_t was[MG_NUM_GENES];
#pragma unroll
    for (int g = 0; g < MG_NUM_GENES; g++) was[g] = epi[g];
    genes_one(p, gen, m, tally, epi, conc);
    if (tally) store16(cp, conc);
    /* most states do not change in a step: store only those that did */
#pragma unroll
    for (int g = 0; g < MG_NUM_GENES; g++)
        if (epi[g] != was[g]) d.EPIGEN[(size_t)g * d.N + h] = epi[g];
}

/* Fused ComputeDiffusionCL + ComputeGenesCL + InsertParticlesCL of mg_step. Reads positions
 * after the advance (d.P), the pre-step concentrations (d.CONC) and the pre-step grid
 * (d.CNT / d.OFF); writes CONC2, updates EPIGEN, records the post-gene membership in GMN for
 * the coming sort and bins the particle into the next grid (d.CNTN). */
template <int NG, class Gen>
__device__ __forceinline__ void diffuse_genes_insert_body(const DS& d, const MgParams& p, const MgGrid& g,
                                                          const Gen& gen, const DiffPlan& dp,
                                                          int genes, Band<NG>* bands, uint32_t bid)
{
    const uint32_t i = bid * blockDim.x + threadIdx.x;
    const bool valid = i < (uint32_t)d.N;
    const uint32_t gc = (i < d.ctr[MG_C_ACTIVE]) ? d.GCELL[i] : MG_GRID_UNDEF;
    const bool diffuses = genes && gc != MG_GRID_UNDEF;
    float out[MG_NUM_TF];
    if (genes) diffuse_one<NG>(d, p, g, gen, dp, bands[threadIdx.x >> 6], diffuses, i, gc, out);
    float4 P_i = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        P_i = d.P[i];
        if (!diffuses) load16(d.CONC + (size_t)i * MG_NUM_TF, out);
        uint32_t mnext = 0;
        if (i < d.ctr[MG_C_NUM_POINTS]) {
            const uint32_t h = d.HOME[i];
            uint32_t epi[MG_NUM_GENES];
#pragma unroll
            for (int k = 0; k < MG_NUM_GENES; k++) epi[k] = d.EPIGEN[(size_t)k * d.N + h];
            const uint32_t m = d.GMASK[i];
            if (genes && m != 0 && d.ctr[MG_C_LISTS]) {
                const bool tally = (m & 4u) && i < d.ctr[MG_C_ACTIVE];
                uint32_t was[MG_NUM_GENES];
#pragma unroll
                for (int k = 0; k < MG_NUM_GENES; k++) was[k] = epi[k];
                genes_one(p, gen, m, tally, epi, out);
                /* most states do not change in a step: store only those that did */
#pragma unroll
                for (int k = 0; k < MG_NUM_GENES; k++)
                    if (epi[k] != was[k]) d.EPIGEN[(size_t)k * d.N + h] = epi[k];
            }
#pragma unroll
            for (int k = 0; k < MG_NUM_GENES; k++)
                if (epi[k] > 0u) mnext |= 1u << k;
        }
        d.GMN[i] = mnext;
        store16(d.CONC2 + (size_t)i * MG_NUM_TF, out);
    }
    insert_one(d, g, i, valid, P_i, d.CNTN, nullptr);
}

template <int NG>
__global__ LAUNCH_BOUNDS_SWEEP void k_diffuse_genes_insert(DS d, MgParams p, MgGrid g, MgGenome gen,
                                                           DiffPlan dp, int genes)
{
    extern __shared__ float4 mg_lds[];
    diffuse_genes_insert_body<NG>(d, p, g, gen, dp, genes, (Band<NG>*)mg_lds, xcd_block());
}

/* -------------------------------------------------------- bond formation
 * contribFindBonds (src/fluid_system_cuda.cu:1938-1976) over the cells gc + gridAdj[c] for one
 * bond slot of the particle at slot i (home h), seen from the point x (the particle's position
 * for initialize_bonds, a shifted point for heal). The reference does not subtract the (1,1,1)
 * offset here, so the block is the one above and right of the particle's cell. Every candidate
 * j with a non-zero particle ID and 0 < |x - x_j|^2 < rd2 is scored by the sum over the
 * particle's current bonds k of |1.91 - acos(u.v / (|u| |v|))| (u = x - x_j, v = x - x_k;
 * FLT_MAX when j is already bonded), or by |u|^2 for slot 0; returns the first strictly best
 * candidate, or MG_NONE. */
__device__ __forceinline__ uint32_t find_bond(const DS& d, const MgGrid& g, const uint32_t* __restrict__ PID,
                                              uint32_t gc, uint32_t h, f3 x, int bond, float rd2)
{
    const uint32_t N = (uint32_t)d.N, C = (uint32_t)d.C;
    const int B = d.B;
    float best = FLT_MAX;
    uint32_t to = MG_NONE;
    for (int c = 0; c < g.adj_cnt; c++) {
        const uint32_t cell = gc + (uint32_t)g.adj[c];
        if (cell >= C) continue;
        const uint32_t cn = d.CNT[cell];
        if (cn == 0u) continue;
        const uint32_t c0 = d.OFF[cell];
        for (uint32_t j = c0; j < c0 + cn && j < N; j++) {
            if (PID[d.HOME[j]] == 0u) continue;
            const f3 u = x - xyz(d.P[j]);
            const float dsq = u.x * u.x + u.y * u.y + u.z * u.z;
            if (!(dsq < rd2 && dsq > 0.0f)) continue;
            float theta = dsq; /* slot 0's score whatever the angle sum */
            if (bond != 0) {
                theta = 0.0f;
                for (int b = 0; b < B; b++) {
                    const uint32_t kh = d.BPARTNER[(size_t)b * N + h];
                    const uint32_t k = kh == MG_NONE ? MG_NONE : d.SLOT[kh];
                    if (j == k) {
                        theta = FLT_MAX;
                        break;
                    }
                    if (k < N) {
                        const f3 v = x - xyz(d.P[k]);
                        const float lu = sqrtf(u.x * u.x + u.y * u.y + u.z * u.z);
                        const float lv = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                        const float cosang = (u.x * v.x + u.y * v.y + u.z * v.z) / (lu * lv);
                        theta = (float)((double)theta + fabs(1.91 - (double)mg_acosf(cosang)));
                    }
                }
            }
            if (theta < best) {
                best = theta;
                to = j;
            }
        }
    }
    return to;
}

/* The searches stage the particle's current bonds in LDS, once per particle and shared by the S
 * lanes of its group: kb[b] holds slot b's partner slot (MG_NONE when empty or unplaced) and
 * vb[b] = (x - x_k, |x - x_k|) for the search point x. Each angle term then reads LDS instead of
 * walking the partner home -> slot -> position chain; slot 0's score is dsq whatever the angle
 * sum, so its loop is skipped. 20 B per slot and group; 64-thread blocks. */
#define BOND_STAGE_W 64
#define BOND_STAGE_MAX_B 192 /* at 8 lanes per particle: 8 groups x 192 slots x 20 B = 30 KB */
__device__ __forceinline__ void stage_bond(const DS& d, f3 x, uint32_t k, int b, float4* vb, uint32_t* kb)
{
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (k < (uint32_t)d.N) {
        const f3 w = x - xyz(d.P[k]);
        v = make_float4(w.x, w.y, w.z, sqrtf(w.x * w.x + w.y * w.y + w.z * w.z));
    }
    kb[b] = k;
    vb[b] = v;
}
/* lanes sub, sub + S, ... of the group stage slots sub, sub + S, ...; the caller syncs */
template <int S>
__device__ __forceinline__ void stage_bonds(const DS& d, uint32_t h, f3 x, uint32_t sub, float4* vb, uint32_t* kb)
{
    const uint32_t N = (uint32_t)d.N;
    for (int b = (int)sub; b < d.B; b += S) {
        const uint32_t kh = d.BPARTNER[(size_t)b * N + h];
        stage_bond(d, x, kh == MG_NONE ? MG_NONE : d.SLOT[kh], b, vb, kb);
    }
}
/* find_bond over the staged bonds, split over the S consecutive lanes of a particle's group (S a
 * power of two, all S lanes call it together): the candidates are numbered q = 0, 1, ... in the
 * reference's visit order (cells in gridAdj order, slots ascending) and lane `sub` scores those
 * with q = sub mod S, keeping its first strictly best; the group then takes the smallest score,
 * ties to the smallest q. That is the candidate the serial loop keeps (the first one reaching
 * the minimum; FLT_MAX and NaN scores never win), and each score is computed by the same
 * operations in the same order, so the result is bit-equal to find_bond. */
template <int S>
__device__ __forceinline__ uint32_t find_bond_split(const DS& d, const MgGrid& g, const uint32_t* __restrict__ PID,
                                                    uint32_t gc, f3 x, int bond, float rd2, const float4* vb,
                                                    const uint32_t* kb, uint32_t sub)
{
    const uint32_t N = (uint32_t)d.N, C = (uint32_t)d.C;
    const int B = d.B;
    float best = FLT_MAX;
    uint32_t bq = 0xFFFFFFFFu, to = MG_NONE, qbase = 0;
    for (int c = 0; c < g.adj_cnt; c++) {
        const uint32_t cell = gc + (uint32_t)g.adj[c];
        if (cell >= C) continue;
        const uint32_t cn = d.CNT[cell];
        if (cn == 0u) continue;
        const uint32_t c0 = d.OFF[cell];
        const uint32_t c1 = min(c0 + cn, N);
        if (c0 >= c1) continue;
        for (uint32_t j = c0 + ((sub + (uint32_t)S - qbase % (uint32_t)S) % (uint32_t)S); j < c1; j += (uint32_t)S) {
            if (PID[d.HOME[j]] == 0u) continue;
            const f3 u = x - xyz(d.P[j]);
            const float dsq = u.x * u.x + u.y * u.y + u.z * u.z;
            if (!(dsq < rd2 && dsq > 0.0f)) continue;
            float theta = dsq;
            if (bond != 0) {
                theta = 0.0f;
                const float lu = sqrtf(dsq);
                for (int b = 0; b < B; b++) {
                    const uint32_t k = kb[b];
                    if (j == k) {
                        theta = FLT_MAX;
                        break;
                    }
                    if (k < N) {
                        const float4 v = vb[b];
                        const float cosang = (u.x * v.x + u.y * v.y + u.z * v.z) / (lu * v.w);
                        theta = (float)((double)theta + fabs(1.91 - (double)mg_acosf(cosang)));
                    }
                }
            }
            if (theta < best) {
                best = theta;
                bq = qbase + (j - c0);
                to = j;
            }
        }
        qbase += c1 - c0;
    }
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
        const float ot = __shfl_xor(best, o);
        const uint32_t oq = (uint32_t)__shfl_xor((int)bq, o);
        const uint32_t oj = (uint32_t)__shfl_xor((int)to, o);
        if (ot < best || (ot == best && oq < bq)) {
            best = ot;
            bq = oq;
            to = oj;
        }
    }
    return to;
}

/* Bond types of slots 0-3 by tissue gene (src/fluid_system_cuda.cu:1133-1139, :2003-2009):
 * bone 9 -> apatite (2) in all four; tendon 6 / muscle 7 / elastic ligament 10 -> collagen (1)
 * in slot 0; elastin (0) otherwise. The reference's cartilage branch tests gene 6 again after
 * the tendon branch and is never taken. */
__device__ __forceinline__ void bond_types(const DS& d, uint32_t h, int bt4[4])
{
    auto gene_on = [&](int gidx) { return d.EPIGEN[(size_t)gidx * d.N + h] > 0u; };
    bt4[0] = bt4[1] = bt4[2] = bt4[3] = 0;
    if (gene_on(9)) {
        bt4[0] = bt4[1] = bt4[2] = bt4[3] = 2;
    } else if (gene_on(6) || gene_on(7) || gene_on(10)) {
        bt4[0] = 1;
    }
}

/* initialize_bonds (src/fluid_system_cuda.cu:1978-2044) for the particles of gene 1's dense
 * list ("solid") below mActivePoints: for each bond slot in turn the best candidate of
 * find_bond is written into the slot with the genome's default parameters of the slot's bond
 * type. Each particle writes only its own records. */
template <int S>
__global__ __launch_bounds__(BOND_STAGE_W) void k_initialize_bonds(DS d, MgParams p, MgGrid g, MgGenome gen,
                                                                   const uint32_t* __restrict__ PID, uint2* BCOLD,
                                                                   int staged)
{
    extern __shared__ float4 mg_lds[];
    const uint32_t li = (blockIdx.x * blockDim.x + threadIdx.x) / (uint32_t)S;
    const uint32_t sub = threadIdx.x % (uint32_t)S, grp = threadIdx.x / (uint32_t)S;
    const uint32_t N = (uint32_t)d.N;
    const int B = d.B;
    if (li >= N || li >= d.LENS[1]) return; /* whole groups leave together */
    const uint32_t i = d.LISTS[(size_t)1 * N + li];
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    if (i >= na) return;
    const uint32_t gc = d.GCELL[i];
    if (gc == MG_GRID_UNDEF) return;
    const uint32_t h = d.HOME[i];
    const f3 ipos = xyz(d.P[i]);
    float4* vb = mg_lds + (size_t)grp * B;
    uint32_t* kb = (uint32_t*)(mg_lds + (size_t)(BOND_STAGE_W / S) * B) + (size_t)grp * B;
    if (staged) {
        stage_bonds<S>(d, h, ipos, sub, vb, kb);
        wave_lds_sync();
    }
    int bt4[4];
    bond_types(d, h, bt4);
    for (int bond = 0; bond < B; bond++) {
        const uint32_t to = staged ? find_bond_split<S>(d, g, PID, gc, ipos, bond, p.rd2, vb, kb, sub)
                                   : find_bond(d, g, PID, gc, h, ipos, bond, p.rd2);
        if (to < na && sub == 0u) {
            const int bt = bond < 4 ? bt4[bond] : 0;
            const size_t bi = (size_t)bond * N + h;
            const uint32_t th = d.HOME[to];
            d.BPARTNER[bi] = th;
            d.BPAR[bi] = make_float4(gen.param[bt][8], gen.param[bt][9], gen.param[bt][10], gen.param[bt][11]);
            d.BINT[bi] = make_float2(0.0f, 0.0f);
            BCOLD[bi].x = PID[th];
            if (staged) stage_bond(d, ipos, to, bond, vb, kb); /* the slot's new partner */
        }
        if (staged) wave_lds_sync();
    }
}

/* Bond formation with the angle terms kept from slot to slot (at most 32 bond slots). At slot
 * s a candidate's score sums |1.91 - acos(u.v / (|u| |v|))| over the particle's bonds in slot
 * order, and from slot s to s + 1 only slot s's partner changes. So 32 lanes serve a particle,
 * lane q holding candidate q of the search (visit order) with e[b] = the double term
 * |1.91 - acos| of its angle to slot b's current partner, and pre = the float sum over the
 * slots before s, which no later slot changes. Each score step is find_bond's
 * (float)((double)theta + term). An empty slot holds +0, which leaves theta's bits as they are
 * (find_bond skips it); a slot whose partner is the candidate holds +inf, which makes the score
 * +inf or NaN, so, like find_bond's FLT_MAX, it never wins. The smallest score over the lanes,
 * ties to the smallest q, is then find_bond's candidate. A particle with more than 32
 * candidates runs find_bond_split over its staged bonds instead. */
#define BOND_TAB_S 32
#define BOND_TAB_B 32
#define BT_PARTNER_TERM __builtin_inf()
typedef double d16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bond_angle(float4 u, float4 v) /* w: the vector's length */
{
    return mg_acosf((u.x * v.x + u.y * v.y + u.z * v.z) / (u.w * v.w));
}
__device__ __forceinline__ double bond_term(float angle) { return fabs(1.91 - (double)angle); }

/* the candidates of find_bond for the 32-lane group `grp` of the wavefront, in visit order:
 * the first 32 into cu (x - x_j and |x - x_j|) and cj (j); returns how many there are */
__device__ __forceinline__ uint32_t gather_bond_candidates(const DS& d, const MgGrid& g,
                                                           const uint32_t* __restrict__ PID, uint32_t gc, f3 x,
                                                           float rd2, uint32_t sub, uint32_t grp, float4* cu,
                                                           uint32_t* cj)
{
    const uint32_t N = (uint32_t)d.N, C = (uint32_t)d.C;
    uint32_t q = 0;
    for (int c = 0; c < g.adj_cnt; c++) {
        const uint32_t cell = gc + (uint32_t)g.adj[c];
        if (cell >= C) continue;
        const uint32_t cn = d.CNT[cell];
        if (cn == 0u) continue;
        const uint32_t c0 = d.OFF[cell];
        const uint32_t c1 = min(c0 + cn, N);
        for (uint32_t base = c0; base < c1; base += BOND_TAB_S) {
            const uint32_t j = base + sub;
            bool ok = false;
            f3 u = mk3(0.0f, 0.0f, 0.0f);
            float dsq = 0.0f;
            if (j < c1 && PID[d.HOME[j]] != 0u) {
                u = x - xyz(d.P[j]);
                dsq = u.x * u.x + u.y * u.y + u.z * u.z;
                ok = dsq < rd2 && dsq > 0.0f;
            }
            /* lanes of the other group, or not in this loop, contribute no bits */
            const uint32_t m = (uint32_t)(__ballot(ok) >> (grp * BOND_TAB_S));
            if (ok) {
                const uint32_t at = q + (uint32_t)__popc(m & ((1u << sub) - 1u));
                if (at < BOND_TAB_S) {
                    cu[at] = make_float4(u.x, u.y, u.z, sqrtf(dsq));
                    cj[at] = j;
                }
            }
            q += (uint32_t)__popc(m);
        }
    }
    return q;
}

/* The same candidates with the 27 cells' counts and offsets loaded at once (lane c: cell c,
 * at most 32 cells) and their slots taken as one concatenated range, 32 per round: lane `sub`
 * of a round finds its cell by binary search over the cells' exclusive starts in LDS (cs: start,
 * cb: first slot). The concatenation keeps the visit order. */
__device__ __forceinline__ uint32_t gather_bond_candidates_flat(const DS& d, const MgGrid& g,
                                                                const uint32_t* __restrict__ PID, uint32_t gc, f3 x,
                                                                float rd2, uint32_t sub, uint32_t grp, float4* cu,
                                                                uint32_t* cj, uint32_t* cs, uint32_t* cb)
{
    const uint32_t N = (uint32_t)d.N, C = (uint32_t)d.C;
    const uint32_t nc = (uint32_t)g.adj_cnt; /* <= 32 */
    uint32_t len = 0, c0 = 0;
    if (sub < nc) {
        const uint32_t cell = gc + (uint32_t)g.adj[sub];
        if (cell < C) {
            const uint32_t cn = d.CNT[cell];
            if (cn != 0u) {
                c0 = d.OFF[cell];
                const uint32_t c1 = min(c0 + cn, N);
                len = c1 > c0 ? c1 - c0 : 0u;
            }
        }
    }
    uint32_t incl = len;
#pragma unroll
    for (int o = 1; o < BOND_TAB_S; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o, BOND_TAB_S);
        if (sub >= (uint32_t)o) incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, BOND_TAB_S - 1, BOND_TAB_S);
    cs[sub] = incl - len;
    cb[sub] = c0;
    wave_lds_sync();
    uint32_t q = 0;
    for (uint32_t base = 0; base < total; base += BOND_TAB_S) {
        const uint32_t gi = base + sub;
        bool ok = false;
        f3 u = mk3(0.0f, 0.0f, 0.0f);
        float dsq = 0.0f;
        uint32_t j = 0;
        if (gi < total) {
            /* the last cell whose start is <= gi: a cell of length > 0 */
            uint32_t c = 0;
#pragma unroll
            for (uint32_t step = BOND_TAB_S / 2; step > 0; step >>= 1)
                if (c + step < nc && cs[c + step] <= gi) c += step;
            j = cb[c] + (gi - cs[c]);
            if (PID[d.HOME[j]] != 0u) {
                u = x - xyz(d.P[j]);
                dsq = u.x * u.x + u.y * u.y + u.z * u.z;
                ok = dsq < rd2 && dsq > 0.0f;
            }
        }
        const uint32_t m = (uint32_t)(__ballot(ok) >> (grp * BOND_TAB_S));
        if (ok) {
            const uint32_t at = q + (uint32_t)__popc(m & ((1u << sub) - 1u));
            if (at < BOND_TAB_S) {
                cu[at] = make_float4(u.x, u.y, u.z, sqrtf(dsq));
                cj[at] = j;
            }
        }
        q += (uint32_t)__popc(m);
    }
    return q;
}

/* minimum of a 64-bit key over the 32 lanes of a group, in every lane of it: DPP exchanges
 * inside 16-lane rows (lane ^ 1, lane ^ 2 in quads; the mirrored lane in the half row, then in
 * the row), one swap of the two rows */
template <int CTRL>
__device__ __forceinline__ uint64_t key_min_dpp(uint64_t k)
{
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(k >> 32), CTRL, 0xF, 0xF, false);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)k, CTRL, 0xF, 0xF, false);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    return o < k ? o : k;
}
__device__ __forceinline__ uint64_t group_key_min(uint64_t k)
{
    k = key_min_dpp<0xB1>(k);  /* quad_perm [1,0,3,2] */
    k = key_min_dpp<0x4E>(k);  /* quad_perm [2,3,0,1] */
    k = key_min_dpp<0x141>(k); /* row_half_mirror */
    k = key_min_dpp<0x140>(k); /* row_mirror */
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), 16, BOND_TAB_S);
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)k, 16, BOND_TAB_S);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    return o < k ? o : k;
}

__device__ __forceinline__ void write_bond(const DS& d, const MgGenome& gen, const int bt4[4], int bond, uint32_t h,
                                           uint32_t to, const uint32_t* __restrict__ PID, uint2* BCOLD)
{
    const int bt = bond < 4 ? bt4[bond] : 0;
    const size_t bi = (size_t)bond * d.N + h;
    const uint32_t th = d.HOME[to];
    d.BPARTNER[bi] = th;
    d.BPAR[bi] = make_float4(gen.param[bt][8], gen.param[bt][9], gen.param[bt][10], gen.param[bt][11]);
    d.BINT[bi] = make_float2(0.0f, 0.0f);
    BCOLD[bi].x = PID[th];
}

/* LDS per 32-lane group: B staged bonds (float4 + uint), 32 candidates (float4 + uint) and
 * the 32 cells' starts and first slots (uint, uint) */
__global__ __launch_bounds__(BOND_STAGE_W) void k_initialize_bonds_tab(DS d, MgParams p, MgGrid g, MgGenome gen,
                                                                       const uint32_t* __restrict__ PID,
                                                                       uint2* BCOLD)
{
    extern __shared__ float4 mg_lds[];
    constexpr uint32_t G = BOND_STAGE_W / BOND_TAB_S;
    const uint32_t li = (blockIdx.x * blockDim.x + threadIdx.x) / BOND_TAB_S;
    const uint32_t sub = threadIdx.x % BOND_TAB_S, grp = threadIdx.x / BOND_TAB_S;
    const uint32_t N = (uint32_t)d.N;
    const int B = d.B;
    if (li >= N || li >= d.LENS[1]) return; /* whole groups leave together */
    const uint32_t i = d.LISTS[(size_t)1 * N + li];
    const uint32_t na = d.ctr[MG_C_ACTIVE];
    if (i >= na) return;
    const uint32_t gc = d.GCELL[i];
    if (gc == MG_GRID_UNDEF) return;
    const uint32_t h = d.HOME[i];
    const f3 ipos = xyz(d.P[i]);
    float4* vb = mg_lds + (size_t)grp * B;
    float4* cu = mg_lds + (size_t)G * B + (size_t)grp * BOND_TAB_S;
    uint32_t* kb = (uint32_t*)(mg_lds + (size_t)G * (B + BOND_TAB_S)) + (size_t)grp * B;
    uint32_t* cj = (uint32_t*)(mg_lds + (size_t)G * (B + BOND_TAB_S)) + (size_t)G * B + (size_t)grp * BOND_TAB_S;
    uint32_t* cs = cj + (size_t)G * BOND_TAB_S;
    uint32_t* cb = cs + (size_t)G * BOND_TAB_S;
    stage_bonds<BOND_TAB_S>(d, h, ipos, sub, vb, kb);
    const uint32_t nq = g.adj_cnt <= BOND_TAB_S
                            ? gather_bond_candidates_flat(d, g, PID, gc, ipos, p.rd2, sub, grp, cu, cj, cs, cb)
                            : gather_bond_candidates(d, g, PID, gc, ipos, p.rd2, sub, grp, cu, cj);
    wave_lds_sync();
    int bt4[4];
    bond_types(d, h, bt4);
    if (nq > BOND_TAB_S) { /* dense neighbourhood */
        for (int bond = 0; bond < B; bond++) {
            const uint32_t to = find_bond_split<BOND_TAB_S>(d, g, PID, gc, ipos, bond, p.rd2, vb, kb, sub);
            if (to < na && sub == 0u) {
                write_bond(d, gen, bt4, bond, h, to, PID, BCOLD);
                stage_bond(d, ipos, to, bond, vb, kb);
            }
            wave_lds_sync();
        }
        return;
    }
    const bool have = sub < nq;
    float4 u = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    uint32_t j = MG_NONE;
    if (have) {
        u = cu[sub];
        j = cj[sub];
    }
    const float dsq = u.x * u.x + u.y * u.y + u.z * u.z; /* the candidate test's value, bit for bit */
    /* e[b] in two vectors of 16 indexed by a wave-uniform slot number (register indexing) */
    d16 e0, e1;
#pragma unroll
    for (int b = 0; b < BOND_TAB_B; b++) {
        double t = 0.0;
        if (b < B) {
            const uint32_t k = kb[b];
            if (have && k == j)
                t = BT_PARTNER_TERM;
            else if (k < N)
                t = bond_term(bond_angle(u, vb[b]));
        }
        if (b < 16)
            e0[b] = t;
        else
            e1[b - 16] = t;
    }
    float pre = 0.0f;         /* the score over slots 0 .. bond-1 */
    uint32_t win = MG_NONE;   /* lane b: slot b's new partner */
    const int B0 = min(B, 16);
    for (int bond = 0; bond < B; bond++) {
        float th = dsq; /* slot 0's score whatever the angle sum */
        if (bond != 0) {
            th = pre;
            for (int b = bond; b < B0; b++) th = (float)((double)th + e0[b]);
            for (int b = max(bond, 16); b < B; b++) th = (float)((double)th + e1[b - 16]);
        }
        /* (score, q) as one key: scores that can win are finite and >= 0, whose bits order
         * as the values do; the smallest key is the smallest score, ties to the smallest q */
        uint64_t key = ~0ull;
        if (have && th < FLT_MAX) key = ((uint64_t)__float_as_uint(th) << 32) | sub;
        key = group_key_min(key);
        if (key != ~0ull) {
            const uint32_t bq = (uint32_t)key, to = cj[bq];
            if (to < na) {
                if (sub == (uint32_t)bond) win = to;
                /* the winner's (x - x_to, |x - x_to|) is what stage_bond computes for the new partner */
                const double t = j == to ? BT_PARTNER_TERM : bond_term(bond_angle(u, cu[bq]));
                if (bond < 16)
                    e0[bond] = t;
                else
                    e1[bond - 16] = t;
            }
        }
        pre = (float)((double)pre + (bond < 16 ? e0[bond] : e1[bond - 16]));
    }
    /* each slot is written once, after the searches (none of them reads the records) */
    if ((int)sub < B && win < na) write_bond(d, gen, bt4, (int)sub, h, win, PID, BCOLD);
}

/* ------------------------------------------------------------- remodelling
 * Change lists (src/fluid.h:103): list l of cell c holds the (slot, bond) pairs of the cell's
 * particles whose change word (bond record word 8) has bit l set; list 0 is heal. */

/* computeBondChanges (src/fluid_system_cuda.cu:1065-1330) for the particles of gene 2's dense
 * list ("living") below mActivePoints. Per bond slot in order: a slot with rest length > 0 has
 * its strain integrators averaged over `steps` and cleared, its rest length multiplied by
 * l_a + l_b tanh(l_c <strain> - l_d) and its modulus by s_a + s_b tanh(s_c <strain^2> - s_d)
 * (tanh_param row of the slot's bond type); the change word is cleared; the first slot that
 * is broken (rest length 0) or empty is counted into the cell's heal count, gets change bit 0,
 * is recorded in HS[slot] (MG_NONE otherwise) and ends the particle's loop. Slots after it
 * keep their change words. */
struct TanhTab {
    float t[3][8];
};
__global__ LAUNCH_BOUNDS_256 void k_bond_changes(DS d, TanhTab tt, uint32_t steps, uint32_t* __restrict__ CCNT,
                                                 uint32_t* __restrict__ HS, uint2* BCOLD)
{
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = (uint32_t)d.N;
    const int B = d.B;
    if (li >= N || li >= d.LENS[2]) return;
    const uint32_t i = d.LISTS[(size_t)2 * N + li];
    if (i >= d.ctr[MG_C_ACTIVE]) return;
    const uint32_t h = d.HOME[i];
    int bt4[4];
    bond_types(d, h, bt4);
    const float fs = (float)steps;
    for (int a = 0; a < B; a++) {
        const size_t bi = (size_t)a * N + h;
        float4 bp = d.BPAR[bi];
        if (bp.y > 0.0f) {
            const float2 in = d.BINT[bi];
            const float si = in.y / fs, ssi = in.x / fs;
            d.BINT[bi] = make_float2(0.0f, 0.0f);
            const float* t = tt.t[a < 4 ? bt4[a] : 0];
            bp.y = bp.y * (t[0] + t[1] * mg_tanhf(t[2] * si - t[3]));
            bp.z = bp.z * (t[4] + t[5] * mg_tanhf(t[6] * ssi - t[7]));
            d.BPAR[bi] = bp;
        }
        BCOLD[bi].y = 0u;
        if (bp.y == 0.0f || d.BPARTNER[bi] == MG_NONE) {
            const uint32_t gc = d.GCELL[i];
            if (gc < (uint32_t)d.C) {
                atomicAdd(&CCNT[gc], 1u);
                HS[i] = (uint32_t)a;
            }
            BCOLD[bi].y = 1u;
            break;
        }
    }
}

/* countingSortChanges (src/fluid_system_cuda.cu:406-618), one thread per cell: the cell's
 * particles in slot order, their bond slots in order, and for each set change bit l the pair
 * (slot, bond) goes to list l at COFF[l][c] + the cell's running count. List 0 (heal) takes
 * exactly the pairs computeBondChanges counted (HS): the reference re-reads change bit 0,
 * which slots after a particle's counted one can still carry from earlier calls, and then
 * writes past the cell's range into the next cell's. Every list is also capped at CCNT[l][c]
 * pairs per cell. A cell whose range ends beyond mActivePoints is skipped, as in the
 * reference. CL holds list l's slots at [l][0][*] and bond indices at [l][1][*]. */
__global__ LAUNCH_BOUNDS_256 void k_sort_changes(DS d, const uint32_t* __restrict__ CCNT,
                                                 const uint32_t* __restrict__ COFF, const uint32_t* __restrict__ HS,
                                                 uint32_t* __restrict__ CL, const uint2* __restrict__ BCOLD)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t C = (uint32_t)d.C, N = (uint32_t)d.N;
    if (c >= C) return;
    uint32_t cap[MG_NUM_CHANGES], others = 0;
    for (int l = 0; l < MG_NUM_CHANGES; l++) {
        cap[l] = CCNT[(size_t)l * C + c];
        if (l > 0) others |= cap[l];
    }
    if ((cap[0] | others) == 0u) return; /* nothing can be written */
    const uint32_t count = d.CNT[c];
    if (count == 0u) return;
    const uint32_t off = d.OFF[c];
    if (off + count > d.ctr[MG_C_ACTIVE]) return;
    uint32_t base[MG_NUM_CHANGES], n[MG_NUM_CHANGES];
    for (int l = 0; l < MG_NUM_CHANGES; l++) {
        base[l] = COFF[(size_t)l * C + c];
        n[l] = 0;
    }
    auto put = [&](int l, uint32_t s, uint32_t a) {
        if (n[l] >= cap[l]) return;
        const uint32_t at = base[l] + n[l]++;
        if (at >= N) return;
        CL[(size_t)l * 2 * N + at] = s;
        CL[(size_t)l * 2 * N + N + at] = a;
    };
    if (others == 0u) { /* the usual case: only the heal list, no change words to read */
        for (uint32_t s = off; s < off + count; s++) {
            const uint32_t hs = HS[s];
            if (hs < (uint32_t)d.B) put(0, s, hs);
        }
        return;
    }
    for (uint32_t s = off; s < off + count; s++) {
        const uint32_t hs = HS[s];
        const uint32_t h = d.HOME[s];
        for (int a = 0; a < d.B; a++) {
            if (hs == (uint32_t)a) put(0, s, (uint32_t)a);
            const uint32_t ch = BCOLD[(size_t)a * N + h].y;
            for (int l = 1; l < MG_NUM_CHANGES; l++)
                if ((ch >> l) & 1u) put(l, s, (uint32_t)a);
        }
    }
}

/* heal (src/fluid_system_cuda.cu:2047-2107) over list 0: the particle's position is shifted
 * by -sqrt(rd2) / (4 + r) per axis where the axis's sign bit of the random word is set
 * (r = 3 bits of the word; x: bits 0-2 and 3, y: 4-6 and 7, z: 8-10 and 11), find_bond runs
 * from the shifted point for the listed slot, and a candidate found is bonded with elastin's
 * default parameters (makeBond, :1573-1596; the partner's incoming list is derived from the
 * owners afterwards). Each listed particle appears once, so each thread writes only its own
 * record. */
template <int S>
__global__ __launch_bounds__(BOND_STAGE_W) void k_heal(DS d, MgParams p, MgGrid g, float4 elastin,
                                                       const uint32_t* __restrict__ CL,
                                                       const uint32_t* __restrict__ CLEN,
                                                       const uint32_t* __restrict__ PID, uint2* BCOLD,
                                                       unsigned long long seed, int staged)
{
    extern __shared__ float4 mg_lds[];
    const uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) / (uint32_t)S;
    const uint32_t sub = threadIdx.x % (uint32_t)S;
    const uint32_t N = (uint32_t)d.N;
    if (k >= N || k >= CLEN[0]) return;
    const uint32_t i = CL[k];
    if (i >= d.ctr[MG_C_ACTIVE]) return;
    const uint32_t bond = CL[N + k];
    if (bond >= (uint32_t)d.B) return;
    const uint32_t gc = d.GCELL[i];
    if (gc == MG_GRID_UNDEF) return;
    const uint32_t h = d.HOME[i];
    const uint32_t r = mg_heal_word(seed, PID[h]);
    const float max_len = sqrtf(p.rd2);
    f3 x = xyz(d.P[i]);
    x.x = x.x + (max_len / (float)(4u + (r & 7u))) * (-1.0f * (float)(1u & (r >> 3)));
    x.y = x.y + (max_len / (float)(4u + ((r >> 4) & 7u))) * (-1.0f * (float)(1u & (r >> 7)));
    x.z = x.z + (max_len / (float)(4u + ((r >> 8) & 7u))) * (-1.0f * (float)(1u & (r >> 11)));
    uint32_t to;
    if (staged) {
        const uint32_t grp = threadIdx.x / (uint32_t)S;
        float4* vb = mg_lds + (size_t)grp * d.B;
        uint32_t* kb = (uint32_t*)(mg_lds + (size_t)(BOND_STAGE_W / S) * d.B) + (size_t)grp * d.B;
        if (bond != 0u) {
            stage_bonds<S>(d, h, x, sub, vb, kb);
            wave_lds_sync();
        }
        to = find_bond_split<S>(d, g, PID, gc, x, (int)bond, p.rd2, vb, kb, sub);
    } else {
        to = find_bond(d, g, PID, gc, h, x, (int)bond, p.rd2);
    }
    if (to < N && sub == 0u) {
        const size_t bi = (size_t)bond * N + h;
        const uint32_t th = d.HOME[to];
        d.BPARTNER[bi] = th;
        d.BPAR[bi] = elastin;
        d.BINT[bi] = make_float2(0.0f, 0.0f);
        BCOLD[bi] = make_uint2(PID[th], 0u);
    }
}

/* heal as k_heal with 32 lanes per listed particle: the candidates gathered in one pass
 * (gather_bond_candidates_flat), lane q scoring candidate q over the bonds staged for the
 * shifted point in slot order with find_bond's operations, and the smallest (score, q) key of
 * the group taking the slot. More than 32 candidates (or more than 32 search cells): the
 * staged search of find_bond_split instead. LDS as k_initialize_bonds_tab. */
__global__ __launch_bounds__(BOND_STAGE_W) void k_heal_flat(DS d, MgParams p, MgGrid g, float4 elastin,
                                                            const uint32_t* __restrict__ CL,
                                                            const uint32_t* __restrict__ CLEN,
                                                            const uint32_t* __restrict__ PID, uint2* BCOLD,
                                                            unsigned long long seed)
{
    extern __shared__ float4 mg_lds[];
    constexpr uint32_t G = BOND_STAGE_W / BOND_TAB_S;
    const uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) / BOND_TAB_S;
    const uint32_t sub = threadIdx.x % BOND_TAB_S, grp = threadIdx.x / BOND_TAB_S;
    const uint32_t N = (uint32_t)d.N;
    const int B = d.B;
    if (k >= N || k >= CLEN[0]) return; /* whole groups leave together */
    const uint32_t i = CL[k];
    if (i >= d.ctr[MG_C_ACTIVE]) return;
    const uint32_t bond = CL[N + k];
    if (bond >= (uint32_t)B) return;
    const uint32_t gc = d.GCELL[i];
    if (gc == MG_GRID_UNDEF) return;
    const uint32_t h = d.HOME[i];
    const uint32_t r = mg_heal_word(seed, PID[h]);
    const float max_len = sqrtf(p.rd2);
    f3 x = xyz(d.P[i]);
    x.x = x.x + (max_len / (float)(4u + (r & 7u))) * (-1.0f * (float)(1u & (r >> 3)));
    x.y = x.y + (max_len / (float)(4u + ((r >> 4) & 7u))) * (-1.0f * (float)(1u & (r >> 7)));
    x.z = x.z + (max_len / (float)(4u + ((r >> 8) & 7u))) * (-1.0f * (float)(1u & (r >> 11)));
    float4* vb = mg_lds + (size_t)grp * B;
    float4* cu = mg_lds + (size_t)G * B + (size_t)grp * BOND_TAB_S;
    uint32_t* kb = (uint32_t*)(mg_lds + (size_t)G * (B + BOND_TAB_S)) + (size_t)grp * B;
    uint32_t* cj = (uint32_t*)(mg_lds + (size_t)G * (B + BOND_TAB_S)) + (size_t)G * B + (size_t)grp * BOND_TAB_S;
    uint32_t* cs = cj + (size_t)G * BOND_TAB_S;
    uint32_t* cb = cs + (size_t)G * BOND_TAB_S;
    if (bond != 0u) stage_bonds<BOND_TAB_S>(d, h, x, sub, vb, kb); /* slot 0 scores by distance */
    const uint32_t nq = g.adj_cnt <= BOND_TAB_S
                            ? gather_bond_candidates_flat(d, g, PID, gc, x, p.rd2, sub, grp, cu, cj, cs, cb)
                            : gather_bond_candidates(d, g, PID, gc, x, p.rd2, sub, grp, cu, cj);
    wave_lds_sync();
    uint32_t to = MG_NONE;
    if (nq > BOND_TAB_S) {
        to = find_bond_split<BOND_TAB_S>(d, g, PID, gc, x, (int)bond, p.rd2, vb, kb, sub);
    } else {
        uint64_t key = ~0ull;
        if (sub < nq) {
            const float4 u = cu[sub];
            const uint32_t j = cj[sub];
            float th = u.x * u.x + u.y * u.y + u.z * u.z; /* slot 0: the candidate test's dsq */
            if (bond != 0u) {
                th = 0.0f;
                for (int b = 0; b < B; b++) {
                    const uint32_t kk = kb[b];
                    if (j == kk) {
                        th = FLT_MAX;
                        break;
                    }
                    if (kk < N) th = (float)((double)th + bond_term(bond_angle(u, vb[b])));
                }
            }
            if (th < FLT_MAX) key = ((uint64_t)__float_as_uint(th) << 32) | sub;
        }
        key = group_key_min(key);
        if (key != ~0ull) to = cj[(uint32_t)key];
    }
    if (to < N && sub == 0u) {
        const size_t bi = (size_t)bond * N + h;
        const uint32_t th = d.HOME[to];
        d.BPARTNER[bi] = th;
        d.BPAR[bi] = elastin;
        d.BINT[bi] = make_float2(0.0f, 0.0f);
        BCOLD[bi] = make_uint2(PID[th], 0u);
    }
}

/* ZeroVelCL (src/host_CL.cpp:2175-2180) over the whole FVEL / FVEVAL buffers. */
__global__ LAUNCH_BOUNDS_256 void k_zero_vel(DS d)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)d.N) return;
    d.VEL[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 v = d.V[i];
    d.V[i] = make_float4(0.f, 0.f, 0.f, v.w);
}

/* -------------------------------------------------------------- batched step
 * One launch per phase steps many independent instances (own buffers, grid, parameters and
 * genome). Block b of a launch belongs to instance map[b], which owns the launch's blocks
 * [b0, b0 + nb) for that phase; each instance then computes exactly what its own fused step
 * computes (the same bodies with the same block-local indices), so a batched instance
 * matches the instance stepped alone bit for bit. */
struct BatchInst {
    DS pressure, force, diffuse, place, sort; /* the views of the fused step's launches */
    MgParams p;
    MgGrid g;
    DiffPlan dp;
    const MgGenome* gen;
    const uint32_t* scan_in;
    uint32_t* scan_out;
    unsigned long long* scan_status;
    unsigned long long* scan_ticket;
    uint32_t* ctr;
    int C, gsm, genes;
    float vmax, dt;
    uint32_t pb0, pnb; /* particle launches */
    uint32_t tb0, tnb; /* scan tiles */
};

/* The instance table's buffers are device allocations. The batched kernels load their
 * pointers from the table, so the compiler cannot tell their address space and issues flat
 * accesses, which also count against the LDS counter the sweeps wait on. A local copy of the
 * view whose pointers pass through the global address space (by value, which the compiler
 * cannot fold back into the generic pointer) lets it address them as global memory. */
template <class T>
__device__ __forceinline__ T* as_global(T* q)
{
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)q;
}

__device__ __forceinline__ DS global_view(const DS& in)
{
    DS d = in;
#define MG_GLOBAL(f) d.f = as_global(d.f)
    MG_GLOBAL(P); MG_GLOBAL(V); MG_GLOBAL(VEL); MG_GLOBAL(F); MG_GLOBAL(CONC); MG_GLOBAL(HOME);
    MG_GLOBAL(GCELL); MG_GLOBAL(GNDX); MG_GLOBAL(GMASK); MG_GLOBAL(P2); MG_GLOBAL(V2);
    MG_GLOBAL(VEL2); MG_GLOBAL(CONC2); MG_GLOBAL(HOME2); MG_GLOBAL(GCELL2); MG_GLOBAL(GNDX2);
    MG_GLOBAL(GMASK2); MG_GLOBAL(SLOT); MG_GLOBAL(EPIGEN); MG_GLOBAL(BPARTNER); MG_GLOBAL(BPAR);
    MG_GLOBAL(BINT); MG_GLOBAL(INC_OFF); MG_GLOBAL(INC); MG_GLOBAL(LIVE); MG_GLOBAL(CNT);
    MG_GLOBAL(OFF); MG_GLOBAL(GCNT); MG_GLOBAL(GOFF); MG_GLOBAL(PERM); MG_GLOBAL(LISTS);
    MG_GLOBAL(LENS); MG_GLOBAL(CNTN); MG_GLOBAL(OFFN); MG_GLOBAL(GMN); MG_GLOBAL(ZERO);
    MG_GLOBAL(DVc); MG_GLOBAL(HMW); MG_GLOBAL(HMR); MG_GLOBAL(ETW); MG_GLOBAL(ETR); MG_GLOBAL(ctr);
#undef MG_GLOBAL
    return d;
}

/* the pressure sweep at candidate batch K and band capacity CAP (the autotuner's tile shapes) */
template <int K, int CAP>
__global__ LAUNCH_BOUNDS_SWEEP void kb_pressure(const BatchInst* __restrict__ I, const uint16_t* __restrict__ map)
{
    extern __shared__ float4 mg_lds[];
    const uint32_t b = xcd_block();
    const BatchInst& in = I[map[b]];
    pressure_body<K, CAP>(global_view(in.pressure), in.p, in.g, (Band<0, CAP>*)mg_lds, b - in.pb0, in.pnb);
}

__global__ LAUNCH_BOUNDS_SWEEP void kb_force_advance(const BatchInst* __restrict__ I,
                                                     const uint16_t* __restrict__ map)
{
    extern __shared__ float4 mg_lds[];
    const uint32_t b = xcd_block();
    const BatchInst& in = I[map[b]];
    force_advance_body(global_view(in.force), in.p, in.g, (Band<1>*)mg_lds, in.dt, b - in.pb0);
}

/* An instance's genome read through the constant address space: nothing writes it while a
 * step runs, so its wavefront-uniform reads become scalar loads (through the table's generic
 * pointer they are flat loads, and every gene waits on memory). The gene and diffusion bodies
 * take the genome's type as a template parameter so the reads keep this address space. */
typedef const __attribute__((address_space(4))) MgGenome ConstGenome;
__device__ __forceinline__ ConstGenome& const_genome(const MgGenome* g) { return *(ConstGenome*)g; }

template <int NG>
__global__ LAUNCH_BOUNDS_SWEEP void kb_diffuse_genes_insert(const BatchInst* __restrict__ I,
                                                            const uint16_t* __restrict__ map)
{
    extern __shared__ float4 mg_lds[];
    const uint32_t b = xcd_block();
    const BatchInst& in = I[map[b]];
    diffuse_genes_insert_body<NG>(global_view(in.diffuse), in.p, in.g, const_genome(in.gen), in.dp, in.genes,
                                  (Band<NG>*)mg_lds, b - in.pb0);
}

__global__ __launch_bounds__(256) void kb_scan(const BatchInst* __restrict__ I, const uint16_t* __restrict__ map)
{
    const BatchInst& in = I[map[blockIdx.x]];
    scan_lookback_body(as_global(in.scan_in), as_global(in.scan_out), in.C, as_global(in.scan_status),
                       as_global(in.scan_ticket), as_global(in.ctr), in.gsm, in.tnb);
}

__global__ LAUNCH_BOUNDS_256 void kb_place(const BatchInst* __restrict__ I, const uint16_t* __restrict__ map)
{
    const BatchInst& in = I[map[blockIdx.x]];
    place_body(global_view(in.place), blockIdx.x - in.pb0);
}

__global__ LAUNCH_BOUNDS_256 void kb_sort_gather(const BatchInst* __restrict__ I, const uint16_t* __restrict__ map)
{
    const uint32_t b = xcd_block();
    const BatchInst& in = I[map[b]];
    sort_gather_body(global_view(in.sort), in.vmax, b - in.pb0);
}

/* --------------------------------------------------- persistent batched step
 * One launch per step for every instance of a batch: a grid of workgroups that stay resident
 * (sized from the occupancy of this kernel) walks the six phases of the fused step, each
 * phase a queue of (instance, block) items taken with one atomic per item (largest instances
 * first), with a grid barrier between phases. Each item runs the same body with the same
 * block-local index as the instance's own launch, so every instance evolves bit for bit as
 * it would alone.
 *
 * Barrier: sense reversal. The last workgroup to arrive resets the arrival count and the
 * queue of the phase that just ended (every workgroup has left it), then flips `sense`; the
 * others wait for the flip with an agent-scope acquire, so the phase's writes, released by
 * every arriving workgroup, are visible to all XCDs after it. Queue 5 (the last phase) is
 * reset at the first barrier of the next launch. Nothing needs clearing between launches,
 * so the launch replays from a graph as it is. A wait that exceeds its bound raises `err`
 * and gives up (the step's results are then invalid) instead of spinning forever, e.g.
 * when the GPU cannot keep the whole grid resident. */
#define PERSIST_PHASES 6
struct PersistCtl {
    unsigned count, sense, err, pad;
    unsigned q[PERSIST_PHASES + 1];
};

struct PersistPlan {
    const BatchInst* I;
    const uint32_t* items[PERSIST_PHASES]; /* (instance << 16) | block, per phase */
    uint32_t nitems[PERSIST_PHASES];
    PersistCtl* ctl;
};

__device__ __forceinline__ void persist_barrier(PersistCtl* ctl, unsigned& sense, int ended)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        sense ^= 1u;
        __atomic_thread_fence(__ATOMIC_RELEASE);
        const unsigned arrived =
            __hip_atomic_fetch_add(&ctl->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == gridDim.x) {
            __hip_atomic_store(&ctl->q[ended], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ended == 0) __hip_atomic_store(&ctl->q[PERSIST_PHASES - 1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->sense, sense, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&ctl->sense, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != sense) {
                if (++spins > (1u << 22)) {
                    atomicOr(&ctl->err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
    __syncthreads();
}

template <int NG>
__global__ __launch_bounds__(256) void kb_step(PersistPlan P)
{
    extern __shared__ float4 mg_lds[];
    __shared__ uint32_t s_item;
    unsigned sense = 0;
    if (threadIdx.x == 0) sense = __hip_atomic_load(&P.ctl->sense, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    for (int ph = 0; ph < PERSIST_PHASES; ph++) {
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(&P.ctl->q[ph], 1u);
            __syncthreads();
            const uint32_t k = s_item;
            __syncthreads();
            if (k >= P.nitems[ph]) break;
            const uint32_t it = P.items[ph][k];
            const BatchInst& in = P.I[it >> 16];
            const uint32_t bid = it & 0xFFFFu;
            switch (ph) {
            case 0: pressure_body(global_view(in.pressure), in.p, in.g, (Band<0>*)mg_lds, bid, in.pnb); break;
            case 1: force_advance_body(global_view(in.force), in.p, in.g, (Band<1>*)mg_lds, in.dt, bid); break;
            case 2:
                diffuse_genes_insert_body<NG>(global_view(in.diffuse), in.p, in.g, const_genome(in.gen), in.dp, in.genes,
                                              (Band<NG>*)mg_lds, bid);
                break;
            case 3:
                scan_lookback_body(as_global(in.scan_in), as_global(in.scan_out), in.C, as_global(in.scan_status),
                                   as_global(in.scan_ticket), as_global(in.ctr), in.gsm, in.tnb);
                break;
            case 4: place_body(global_view(in.place), bid); break;
            default: sort_gather_body(global_view(in.sort), in.vmax, bid); break;
            }
            __syncthreads(); /* the bodies' LDS is reused by the next item */
        }
        if (ph + 1 < PERSIST_PHASES) persist_barrier(P.ctl, sense, ph);
    }
}

#include "mg_mlp.hip"
