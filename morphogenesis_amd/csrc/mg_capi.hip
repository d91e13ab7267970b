/*
 * mg_capi.hip — host side of libmg_sph.so: the C-ABI of include/mg_sph.h.
 *
 * A handle owns every device buffer and one HIP stream (default priority). Buffers cross
 * the boundary only in the reference layout; the conversion kernels below translate to the
 * internal HBM layout described in mg_device.h. Nothing here falls back to the CPU: a
 * missing GPU or a failed HIP call is reported as MG_ERR_HIP.
 */
#include "mg_kernels.hip"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mg_sph.h"

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);  \
            return MG_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)

#define CHK(x)                                                                                \
    do {                                                                                      \
        int r_ = (x);                                                                         \
        if (r_ != MG_OK) return r_;                                                           \
    } while (0)

inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

/* ------------------------------------------------ layout conversion kernels */
__global__ void k_set_xyz(float4* dst, const float4* src, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 s = src[i], d = dst[i];
    dst[i] = make_float4(s.x, s.y, s.z, d.w);
}
__global__ void k_get_xyz(float4* dst, const float4* src, int n, int zero)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 s = src[i];
    dst[i] = zero ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(s.x, s.y, s.z, 0.0f);
}
__global__ void k_set_w(float4* dst, const float* src, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ((float*)&dst[i])[3] = src[i];
}
__global__ void k_get_w(float* dst, const float4* src, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[i].w;
}
__global__ void k_iota(uint32_t* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
/* ---------------------------------------------------------- re-homing
 * A permutation sigma of the home indices: every home-indexed record moves from h to
 * sigma[h] and every stored home value v becomes sigma[v]. */
template <typename T>
__global__ void k_scatter_home(T* dst, const T* src, const uint32_t* sigma, int n, int rows)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (h >= n || r >= rows) return;
    dst[(size_t)r * n + sigma[h]] = src[(size_t)r * n + h];
}
__global__ void k_rehome_bonds(uint32_t* nPartner, uint2* nPidx, const uint32_t* Partner, const uint2* Pidx,
                               const uint32_t* sigma, int n)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (h >= n) return;
    const size_t o = (size_t)a * n + h, no = (size_t)a * n + sigma[h];
    const uint32_t jh = Partner[o];
    nPartner[no] = (jh < (uint32_t)n) ? sigma[jh] : jh;
    const uint2 pi = Pidx[o];
    nPidx[no] = make_uint2(pi.x < (uint32_t)n ? sigma[pi.x] : pi.x, pi.y);
}
__global__ void k_map_values(uint32_t* v, const uint32_t* sigma, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t x = v[s];
    if (x < (uint32_t)n) v[s] = sigma[x];
}
/* home-indexed u32 field -> sorted order; slots >= tail get `reset` unless keep_stale */
__global__ void k_gather_home(uint32_t* dst, const uint32_t* src, const uint32_t* HOME,
                              const uint32_t* ctr, int n, int keep_stale, uint32_t reset)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[s] = (keep_stale || (uint32_t)s < ctr[MG_C_TAIL]) ? src[HOME[s]] : reset;
}
__global__ void k_gather_epigen(uint32_t* dst, const uint32_t* src, const uint32_t* HOME,
                                const uint32_t* ctr, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int g = blockIdx.y;
    if (s >= n) return;
    dst[(size_t)g * n + s] = ((uint32_t)s < ctr[MG_C_TAIL]) ? src[(size_t)g * n + HOME[s]] : 0u;
}
/* home-indexed bond storage -> bond storage indexed by sorted slot (normalisation) */
__global__ void k_norm_bonds(uint32_t* nPartner, float4* nPar, float2* nInt, uint2* nCold,
                             uint2* nPidx, const uint32_t* Partner, const float4* Par,
                             const float2* Int, const uint2* Cold, const uint2* Pidx,
                             const uint32_t* HOME, const uint32_t* SLOT, const uint32_t* ctr,
                             int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (s >= n) return;
    size_t o = (size_t)a * n + s;
    if ((uint32_t)s < ctr[MG_C_TAIL]) {
        size_t hi = (size_t)a * n + HOME[s];
        uint32_t jh = Partner[hi];
        nPartner[o] = (jh == MG_NONE) ? MG_NONE : SLOT[jh];
        nPar[o] = Par[hi];
        nInt[o] = Int[hi];
        nCold[o] = Cold[hi];
        uint2 pi = Pidx[hi];
        nPidx[o] = make_uint2(pi.x == MG_NONE ? MG_NONE : SLOT[pi.x], pi.y);
    } else {
        float nf = __uint_as_float(MG_NONE);
        nPartner[o] = MG_NONE;
        nPar[o] = make_float4(nf, nf, nf, nf);
        nInt[o] = make_float2(nf, nf);
        nCold[o] = make_uint2(MG_NONE, MG_NONE);
        nPidx[o] = make_uint2(MG_NONE, MG_NONE);
    }
}
/* reference bond records [N][B][9] -> internal (after normalisation HOME is identity) */
__global__ void k_upload_bonds(uint32_t* Partner, float4* Par, float2* Int, uint2* Cold,
                               const uint32_t* rec, int n, int B)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (s >= n) return;
    const uint32_t* w = rec + ((size_t)s * B + a) * MG_DATA_PER_BOND;
    size_t o = (size_t)a * n + s;
    Partner[o] = (w[0] < (uint32_t)n) ? w[0] : MG_NONE;
    Par[o] = make_float4(__uint_as_float(w[1]), __uint_as_float(w[2]), __uint_as_float(w[3]),
                         __uint_as_float(w[4]));
    Int[o] = make_float2(__uint_as_float(w[6]), __uint_as_float(w[7]));
    Cold[o] = make_uint2(w[5], w[8]);
}
__global__ void k_download_bonds(uint32_t* rec, const uint32_t* Partner, const float4* Par,
                                 const float2* Int, const uint2* Cold, const uint32_t* HOME,
                                 const uint32_t* SLOT, const uint32_t* ctr, int n, int B)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (s >= n) return;
    uint32_t* w = rec + ((size_t)s * B + a) * MG_DATA_PER_BOND;
    if ((uint32_t)s >= ctr[MG_C_TAIL]) {
        for (int k = 0; k < MG_DATA_PER_BOND; k++) w[k] = MG_NONE;
        return;
    }
    size_t hi = (size_t)a * n + HOME[s];
    uint32_t jh = Partner[hi];
    float4 bp = Par[hi];
    float2 bi = Int[hi];
    uint2 bc = Cold[hi];
    w[0] = (jh == MG_NONE) ? MG_NONE : SLOT[jh];
    w[1] = __float_as_uint(bp.x);
    w[2] = __float_as_uint(bp.y);
    w[3] = __float_as_uint(bp.z);
    w[4] = __float_as_uint(bp.w);
    w[5] = bc.x;
    w[6] = __float_as_uint(bi.x);
    w[7] = __float_as_uint(bi.y);
    w[8] = bc.y;
}
__global__ void k_upload_pidx(uint2* Pidx, const uint32_t* rec, int n, int B)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (s >= n) return;
    const uint32_t* w = rec + ((size_t)s * B + a) * 2;
    Pidx[(size_t)a * n + s] = make_uint2(w[0] < (uint32_t)n ? w[0] : MG_NONE, w[1]);
}
__global__ void k_download_pidx(uint32_t* rec, const uint2* Pidx, const uint32_t* HOME,
                                const uint32_t* SLOT, const uint32_t* ctr, int n, int B)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int a = blockIdx.y;
    if (s >= n) return;
    uint32_t* w = rec + ((size_t)s * B + a) * 2;
    if ((uint32_t)s >= ctr[MG_C_TAIL]) {
        w[0] = MG_NONE;
        w[1] = MG_NONE;
        return;
    }
    uint2 pi = Pidx[(size_t)a * n + HOME[s]];
    w[0] = pi.x == MG_NONE ? MG_NONE : SLOT[pi.x];
    w[1] = pi.y;
}
__global__ void k_download_fbin(uint32_t* dst, const uint32_t* FBIN, const uint32_t* ctr, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[s] = ((uint32_t)s < ctr[MG_C_HWM]) ? (uint32_t)s : FBIN[s];
}

} // namespace

/* ================================================================== handle */
struct mg_handle {
    int N = 0, B = 0, C = 0;
    int sph = MG_SPH_REFERENCE; /* SPH kernels of pressure and force (MG_SPH_*) */
    MgParams p;
    MgGrid g;
    MgGenome gen;
    bool have_params = false, have_grid = false;
    hipStream_t st = nullptr;     /* the stream work is queued on (a batch's while in one) */
    hipStream_t own_st = nullptr; /* the handle's own stream */
    struct mg_batch* batch = nullptr;
    /* sorted ping-pong halves; current-half indices: kcur for P/V/VEL, icur for
     * HOME/GCELL/GNDX/GMASK, ccur for CONC (each group flips on its own) */
    float4 *P[2] = {}, *V[2] = {}, *VEL[2] = {};
    float* CONC[2] = {};
    uint32_t *HOME[2] = {}, *GCELL[2] = {}, *GNDX[2] = {}, *GMASK[2] = {};
    int kcur = 0, icur = 0, ccur = 0, gcur = 0;
    float4* F = nullptr;
    bool force_zero = true;
    /* learned coarse steps: per-slot velocity correction and the fallback counter */
    float4* DV = nullptr;
    uint4* XB = nullptr;    /* learned step: normalised bf16 inputs, N x 64 */
    uint32_t* MF = nullptr; /* learned step: per-slot flags (active, inputs in range) */
    uint32_t* fallbacks = nullptr;
    /* home-indexed */
    uint32_t *SLOT = nullptr, *EPIGEN = nullptr, *BPARTNER = nullptr;
    float4* BPAR = nullptr;
    float2* BINT = nullptr;
    uint2 *BCOLD = nullptr, *PIDX = nullptr;
    uint32_t *AGE = nullptr, *COLOR = nullptr, *PID = nullptr, *MRAD = nullptr, *NERVE = nullptr;
    uint32_t* FBIN = nullptr;
    uint32_t *INC_OFF = nullptr, *INC = nullptr;
    size_t inc_cap = 0;
    uint32_t* incw = nullptr;   /* device build of INC: counts [N], tile sums, maximum */
    uint32_t* incpos = nullptr; /* device build of INC: each bond's place in its partner's run [B][N] */
    uint64_t* LIVE = nullptr; /* per home: bit a set when bond slot a holds a partner */
    uint64_t* HM = nullptr;   /* hit masks, pressure -> force of a fused step: [row][slot] */
    int hm_rows = 0;          /* rows HM holds per slot */
    float4* ET = nullptr;     /* bond records, pressure -> force beside the masks: [bond slot][home] */
    /* grid */
    uint32_t *CNT[2] = {}, *OFF[2] = {};  /* gcur is current; the fused step bins into 1-gcur */
    uint32_t *GCNT = nullptr, *GOFF = nullptr, *LENS = nullptr;
    uint32_t *LISTS = nullptr, *PERM = nullptr, *GMN = nullptr;
    bool gene_grid_stale = false;         /* gene counts/offsets/lists not rebuilt since a fused step */
    /* remodelling change lists, allocated on first use: counts and offsets [9][C], lengths [9],
     * (slot, bond) pairs [9][2][N], and per slot the bond computeBondChanges counted for heal */
    uint32_t *CCNT = nullptr, *COFF = nullptr, *CLEN = nullptr, *CL = nullptr, *HS = nullptr;
    uint32_t* tsum = nullptr; /* tile sums and prefixes of the segmented scans [2][16][tiles] */
    /* OFF[gcur] is the exclusive prefix sum of CNT[gcur] (set by a scan, cleared by an insert
     * or an upload of either); with a row-structured gridAdj the sweeps walk rows */
    bool grid_scanned = false;
    bool adj_rows = false;
    /* slots sorted by cell by the last sort and positions unchanged since (no advance, no
     * upload): sweeps may skip neighbour cells out of reach */
    bool cells_fresh = false;
    bool pos_binned = false; /* GCELL / CNT come from the current positions (insert since the last move) */
    uint32_t* ctr = nullptr;
    unsigned long long* scan_status = nullptr;  /* decoupled look-back status, one word per tile */
    unsigned long long* scan_ticket = nullptr; /* 64-bit: never wraps (see k_scan_lookback) */
    void* staging = nullptr;
    size_t staging_bytes = 0;
    bool normalized = true;
    /* re-homing: due after a sort outside the fused step (the homes no longer follow the
     * slots) and every kRehomeSteps fused steps */
    bool rehome_due = false;
    int steps_since_rehome = 0;
    double time = 0.0;
    std::vector<void*> allocs;
    /* phase timing (mg_step_timed) */
    bool timing = false;
    struct Mark {
        hipEvent_t a, b;
        int ph;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    /* two captured fused steps (mg_step with MG_STEP_GRAPH) and what they were captured for */
    hipGraphExec_t graph = nullptr;
    uint64_t epoch = 1;  /* bumped whenever kernel arguments or buffer addresses change */
    uint64_t graph_epoch = 0;
    float graph_dt = 0.f;
    unsigned graph_flags = 0;
    int graph_parity = -1;
    const void* graph_net = nullptr; /* the net of a learned-step graph, with its version */
    uint64_t graph_net_version = 0;
};

namespace {

template <typename T> int dalloc(mg_handle* h, T** p, size_t count)
{
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIPCHK(hipMalloc(&q, bytes));
    h->allocs.push_back(q);
    *p = (T*)q;
    return MG_OK;
}

void dfree(mg_handle* h, void* p)
{
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(p);
}

hipEvent_t pool_event(mg_handle* h)
{
    if (h->pool_used == h->pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->pool.push_back(e);
    }
    return h->pool[h->pool_used++];
}

/* Brackets the launches of one phase with an event pair when the handle is timing. */
struct PhaseScope {
    mg_handle* h;
    int ph;
    hipEvent_t a = nullptr;
    PhaseScope(mg_handle* h_, int ph_) : h(h_), ph(ph_)
    {
        if (h->timing && (a = pool_event(h))) (void)hipEventRecord(a, h->st);
    }
    ~PhaseScope()
    {
        if (!a) return;
        hipEvent_t b = pool_event(h);
        if (!b) return;
        (void)hipEventRecord(b, h->st);
        h->marks.push_back({a, b, ph});
    }
};

int fill_u32(mg_handle* h, void* p, uint32_t v, size_t n)
{
    if (n == 0) return MG_OK;
    unsigned blocks = (unsigned)std::min<size_t>(nblk(n), 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, h->st, (uint32_t*)p, v, n);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int staging(mg_handle* h, size_t bytes)
{
    if (bytes <= h->staging_bytes) return MG_OK;
    dfree(h, h->staging);
    h->staging = nullptr;
    h->staging_bytes = 0;
    CHK(dalloc(h, (uint8_t**)&h->staging, bytes));
    h->staging_bytes = bytes;
    h->epoch++;
    return MG_OK;
}

/* Kernel view of the handle. `kin` / `conc` select which half is read as current (0) or
 * as the other half (1): the fused step reads the advanced kinematics and the diffused
 * concentrations from the "other" halves. */
DS make_ds(mg_handle* h, int kin = 0, int conc = 0)
{
    DS d;
    const int kc = kin ? 1 - h->kcur : h->kcur, ko = 1 - kc;
    const int cc = conc ? 1 - h->ccur : h->ccur;
    const int ic = h->icur, io = 1 - h->icur;
    d.P = h->P[kc]; d.V = h->V[kc]; d.VEL = h->VEL[kc]; d.F = h->F;
    d.P2 = h->P[ko]; d.V2 = h->V[ko]; d.VEL2 = h->VEL[ko];
    d.CONC = h->CONC[cc]; d.CONC2 = h->CONC[1 - cc];
    d.HOME = h->HOME[ic]; d.GCELL = h->GCELL[ic]; d.GNDX = h->GNDX[ic]; d.GMASK = h->GMASK[ic];
    d.HOME2 = h->HOME[io]; d.GCELL2 = h->GCELL[io]; d.GNDX2 = h->GNDX[io]; d.GMASK2 = h->GMASK[io];
    d.SLOT = h->SLOT; d.EPIGEN = h->EPIGEN; d.BPARTNER = h->BPARTNER;
    d.BPAR = h->BPAR; d.BINT = h->BINT;
    d.INC_OFF = h->INC_OFF; d.INC = h->INC; d.LIVE = h->LIVE;
    d.CNT = h->CNT[h->gcur]; d.OFF = h->OFF[h->gcur];
    d.CNTN = h->CNT[1 - h->gcur]; d.OFFN = h->OFF[1 - h->gcur];
    d.GCNT = h->GCNT; d.GOFF = h->GOFF;
    d.PERM = h->PERM; d.LISTS = h->LISTS; d.LENS = h->LENS;
    d.GMN = nullptr;
    d.ZERO = nullptr;
    d.zero_n = 0;
    d.rows = (h->grid_scanned && h->adj_rows) ? 1 : 0;
    d.cull = (d.rows && h->cells_fresh) ? 1 : 0;
    d.DVc = nullptr;
    d.HMW = nullptr;
    d.HMR = nullptr;
    d.ETW = nullptr;
    d.ETR = nullptr;
    d.dry = 0;
    d.ctr = h->ctr;
    d.N = h->N; d.B = h->B; d.C = h->C;
    return d;
}

int ready(mg_handle* h)
{
    if (!h) { set_err("null handle"); return MG_ERR_ARG; }
    if (!h->have_grid || !h->have_params) {
        set_err("grid and parameters must be set before stepping");
        return MG_ERR_STATE;
    }
    return MG_OK;
}

/* Rebuild the incoming-bond CSR from BPARTNER (home space): entries of particle j are the
 * (owner home, slot) pairs whose partner is j, ordered by the owner's particle ID, then its
 * home, then the slot. The ID order makes the force's summation order a property of the
 * state itself: homes are renumbered by uploads and re-homing, IDs travel with particles.
 * rebuild_inc builds it on the device (runs of up to INC_LONG_MAX / INC_PER_WAVE entries);
 * this host build serves particles with more incoming bonds than that. */
int rebuild_inc_host(mg_handle* h)
{
    const int N = h->N, B = h->B;
    std::vector<uint32_t> partner((size_t)N * B);
    std::vector<uint32_t> owners((size_t)N);
    for (int k = 0; k < N; k++) owners[k] = (uint32_t)k;
    if (B > 0) {
        std::vector<uint32_t> pid((size_t)N);
        HIPCHK(hipMemcpyAsync(partner.data(), h->BPARTNER, partner.size() * 4, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipMemcpyAsync(pid.data(), h->PID, pid.size() * 4, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        if (!std::is_sorted(pid.begin(), pid.end()))
            std::stable_sort(owners.begin(), owners.end(),
                             [&](uint32_t x, uint32_t y) { return pid[x] < pid[y]; });
    }
    std::vector<uint32_t> off((size_t)N + 1, 0);
    std::vector<uint64_t> live((size_t)N, 0);
    for (int a = 0; a < B; a++)
        for (int k = 0; k < N; k++) {
            uint32_t j = partner[(size_t)a * N + k];
            if (j < (uint32_t)N) {
                off[j + 1]++;
                live[k] |= 1ull << a;
            }
        }
    for (int j = 0; j < N; j++) off[j + 1] += off[j];
    std::vector<uint32_t> ent(std::max<uint32_t>(off[N], 1));
    std::vector<uint32_t> fillp(off.begin(), off.end() - 1);
    for (uint32_t k : owners)
        for (int a = 0; a < B; a++) {
            uint32_t j = partner[(size_t)a * N + k];
            if (j < (uint32_t)N) ent[fillp[j]++] = (uint32_t)k * (uint32_t)B + (uint32_t)a;
        }
    if (ent.size() > h->inc_cap) {
        dfree(h, h->INC);
        h->INC = nullptr;
        CHK(dalloc(h, &h->INC, ent.size()));
        h->inc_cap = ent.size();
    }
    h->epoch++;
    HIPCHK(hipMemcpyAsync(h->INC_OFF, off.data(), off.size() * 4, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->INC, ent.data(), ent.size() * 4, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->LIVE, live.data(), live.size() * 8, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

int rebuild_inc(mg_handle* h)
{
    const int N = h->N, B = h->B;
    if (B == 0) return rebuild_inc_host(h);
    const unsigned tiles = (unsigned)((N + SCAN_TILE - 1) / SCAN_TILE);
    if (!h->incw) CHK(dalloc(h, &h->incw, (size_t)N + 2 * tiles + 2));
    if (!h->incpos) CHK(dalloc(h, &h->incpos, (size_t)N * B));
    uint32_t *cnt = h->incw, *sums = cnt + N, *excl = sums + tiles, *stat = excl + tiles;
    CHK(fill_u32(h, cnt, 0, (size_t)N));
    CHK(fill_u32(h, stat, 0, 1));
    hipLaunchKernelGGL(k_inc_count, dim3(nblk(N)), dim3(256), 0, h->st, h->BPARTNER, cnt, h->incpos, h->LIVE, N, B);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(256), 0, h->st, cnt, N, sums, stat);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, h->st, sums, excl, (int)tiles, h->ctr, 0, 0, h->INC_OFF + N);
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(256), 0, h->st, cnt, h->INC_OFF, N, excl);
    HIPCHK(hipGetLastError());
    uint32_t total = 0, longest = 0;
    HIPCHK(hipMemcpyAsync(&total, h->INC_OFF + N, 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(&longest, stat, 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    /* a group of INC_PER_WAVE partners holds at most INC_PER_WAVE * longest entries */
    if ((size_t)longest * INC_PER_WAVE > INC_LONG_MAX) return rebuild_inc_host(h);
    if (std::max<size_t>(total, 1) > h->inc_cap) {
        dfree(h, h->INC);
        h->INC = nullptr;
        CHK(dalloc(h, &h->INC, std::max<size_t>(total, 1)));
        h->inc_cap = std::max<size_t>(total, 1);
    }
    h->epoch++;
    if (total == 0) return MG_OK;
    hipLaunchKernelGGL(k_inc_fill, dim3(nblk(N)), dim3(256), 0, h->st, h->BPARTNER, h->INC_OFF, h->incpos, h->INC, N,
                       B);
    const unsigned blocks = (unsigned)(((size_t)N + 4 * INC_PER_WAVE - 1) / (4 * INC_PER_WAVE));
    hipLaunchKernelGGL(k_inc_order, dim3(blocks), dim3(256), 0, h->st, h->INC_OFF, h->INC, h->PID, N, B);
    if (longest > INC_RUN_MAX) {
        const unsigned groups = (unsigned)(((size_t)N + INC_PER_WAVE - 1) / INC_PER_WAVE);
        hipLaunchKernelGGL(k_inc_order_long, dim3(groups), dim3(256), 0, h->st, h->INC_OFF, h->INC, h->PID, N, B);
    }
    HIPCHK(hipGetLastError());
    return MG_OK;
}

/* Make the home index equal to the sorted slot again (HOME = SLOT = identity), materialising
 * the sort's reset tail, so uploads can address home-indexed storage by slot. */
int normalize(mg_handle* h)
{
    if (h->normalized) return MG_OK;
    const int N = h->N, B = h->B;
    const uint32_t* HOME = h->HOME[h->icur];
    auto regather = [&](uint32_t** field, int keep_stale, uint32_t reset) -> int {
        uint32_t* nf = nullptr;
        CHK(dalloc(h, &nf, N));
        hipLaunchKernelGGL(k_gather_home, dim3(nblk(N)), dim3(256), 0, h->st, nf, *field, HOME,
                           h->ctr, N, keep_stale, reset);
        HIPCHK(hipGetLastError());
        dfree(h, *field);
        *field = nf;
        return MG_OK;
    };
    CHK(regather(&h->AGE, 1, 0));
    CHK(regather(&h->COLOR, 1, 0));
    CHK(regather(&h->MRAD, 1, 0));
    CHK(regather(&h->NERVE, 1, 0));
    CHK(regather(&h->PID, 0, MG_NONE));
    {
        uint32_t* ne = nullptr;
        CHK(dalloc(h, &ne, (size_t)MG_NUM_GENES * N));
        hipLaunchKernelGGL(k_gather_epigen, dim3(nblk(N), MG_NUM_GENES), dim3(256), 0, h->st, ne,
                           h->EPIGEN, HOME, h->ctr, N);
        HIPCHK(hipGetLastError());
        dfree(h, h->EPIGEN);
        h->EPIGEN = ne;
    }
    if (B > 0) {
        uint32_t* nP = nullptr;
        float4* nPar = nullptr;
        float2* nInt = nullptr;
        uint2 *nCold = nullptr, *nPidx = nullptr;
        CHK(dalloc(h, &nP, (size_t)B * N));
        CHK(dalloc(h, &nPar, (size_t)B * N));
        CHK(dalloc(h, &nInt, (size_t)B * N));
        CHK(dalloc(h, &nCold, (size_t)B * N));
        CHK(dalloc(h, &nPidx, (size_t)B * N));
        hipLaunchKernelGGL(k_norm_bonds, dim3(nblk(N), B), dim3(256), 0, h->st, nP, nPar, nInt,
                           nCold, nPidx, h->BPARTNER, h->BPAR, h->BINT, h->BCOLD, h->PIDX, HOME,
                           h->SLOT, h->ctr, N);
        HIPCHK(hipGetLastError());
        dfree(h, h->BPARTNER); dfree(h, h->BPAR); dfree(h, h->BINT); dfree(h, h->BCOLD); dfree(h, h->PIDX);
        h->BPARTNER = nP; h->BPAR = nPar; h->BINT = nInt; h->BCOLD = nCold; h->PIDX = nPidx;
    }
    hipLaunchKernelGGL(k_iota, dim3(nblk(N)), dim3(256), 0, h->st, h->HOME[h->icur], N);
    hipLaunchKernelGGL(k_iota, dim3(nblk(N)), dim3(256), 0, h->st, h->SLOT, N);
    HIPCHK(hipGetLastError());
    uint32_t tail = (uint32_t)N;
    HIPCHK(hipMemcpyAsync(h->ctr + MG_C_TAIL, &tail, 4, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->normalized = true;
    h->epoch++;
    return rebuild_inc(h);
}

/* Re-homing: after a sort the particles' records stay at the homes they had at upload, so a
 * wavefront's bond-record reads scatter. rehome() renumbers the homes to the current sorted
 * slots (the particles outside the sorted region take the remaining numbers in order), moves
 * every home-indexed record, maps every stored home value, and relocates each particle's
 * incoming-bond row keeping its entries in their order, so the force gathers the same terms
 * in the same order: no result changes. */
const int kRehomeSteps = 2048;

int rehome(mg_handle* h)
{
    const int N = h->N, B = h->B;
    std::vector<uint32_t> slot(N);
    HIPCHK(hipMemcpyAsync(slot.data(), h->SLOT, (size_t)N * 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    std::vector<uint32_t> sigma(N, MG_NONE);
    std::vector<char> used(N, 0);
    for (int k = 0; k < N; k++) {
        const uint32_t sl = slot[k];
        if (sl >= (uint32_t)N) continue;
        if (used[sl]) return MG_OK; /* not a permutation: leave the homes as they are */
        used[sl] = 1;
        sigma[k] = sl;
    }
    uint32_t next = 0;
    bool ident = true;
    for (int k = 0; k < N; k++) {
        if (sigma[k] == MG_NONE) {
            while (used[next]) next++;
            sigma[k] = next;
            used[next] = 1;
        }
        if (sigma[k] != (uint32_t)k) ident = false;
    }
    h->rehome_due = false;
    h->steps_since_rehome = 0;
    if (ident) return MG_OK;
    uint32_t* dsig = nullptr;
    CHK(dalloc(h, &dsig, N));
    HIPCHK(hipMemcpyAsync(dsig, sigma.data(), (size_t)N * 4, hipMemcpyHostToDevice, h->st));
    const unsigned nb = nblk(N);
    auto move = [&](auto** field, int rows) -> int {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*field)>>;
        T* nf = nullptr;
        CHK(dalloc(h, &nf, (size_t)rows * N));
        hipLaunchKernelGGL(k_scatter_home<T>, dim3(nb, rows), dim3(256), 0, h->st, nf, (const T*)*field, dsig, N, rows);
        HIPCHK(hipGetLastError());
        dfree(h, *field);
        *field = nf;
        return MG_OK;
    };
    CHK(move(&h->SLOT, 1));
    CHK(move(&h->AGE, 1));
    CHK(move(&h->COLOR, 1));
    CHK(move(&h->PID, 1));
    CHK(move(&h->MRAD, 1));
    CHK(move(&h->NERVE, 1));
    CHK(move(&h->LIVE, 1));
    CHK(move(&h->EPIGEN, MG_NUM_GENES));
    if (B > 0) {
        CHK(move(&h->BPAR, B));
        CHK(move(&h->BINT, B));
        CHK(move(&h->BCOLD, B));
        uint32_t* nP = nullptr;
        uint2* nPidx = nullptr;
        CHK(dalloc(h, &nP, (size_t)B * N));
        CHK(dalloc(h, &nPidx, (size_t)B * N));
        hipLaunchKernelGGL(k_rehome_bonds, dim3(nb, B), dim3(256), 0, h->st, nP, nPidx, h->BPARTNER, h->PIDX, dsig, N);
        HIPCHK(hipGetLastError());
        dfree(h, h->BPARTNER);
        dfree(h, h->PIDX);
        h->BPARTNER = nP;
        h->PIDX = nPidx;
    }
    for (int k = 0; k < 2; k++)
        hipLaunchKernelGGL(k_map_values, dim3(nb), dim3(256), 0, h->st, h->HOME[k], dsig, N);
    HIPCHK(hipGetLastError());
    /* incoming-bond rows follow their particle, entries in the same order */
    std::vector<uint32_t> off((size_t)N + 1);
    HIPCHK(hipMemcpyAsync(off.data(), h->INC_OFF, off.size() * 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    std::vector<uint32_t> ent(std::max<uint32_t>(off[N], 1));
    if (off[N]) HIPCHK(hipMemcpyAsync(ent.data(), h->INC, (size_t)off[N] * 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    std::vector<uint32_t> inv(N);
    for (int k = 0; k < N; k++) inv[sigma[k]] = (uint32_t)k;
    std::vector<uint32_t> noff((size_t)N + 1, 0), nent(ent.size());
    for (int q = 0; q < N; q++) {
        const uint32_t k = inv[q];
        uint32_t o = noff[q];
        for (uint32_t e = off[k]; e < off[k + 1]; e++) {
            const uint32_t packed = ent[e];
            const uint32_t kh = packed / (uint32_t)B, b = packed - kh * (uint32_t)B;
            nent[o++] = sigma[kh] * (uint32_t)B + b;
        }
        noff[q + 1] = o;
    }
    HIPCHK(hipMemcpyAsync(h->INC_OFF, noff.data(), noff.size() * 4, hipMemcpyHostToDevice, h->st));
    if (off[N]) HIPCHK(hipMemcpyAsync(h->INC, nent.data(), (size_t)off[N] * 4, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    dfree(h, dsig);
    h->epoch++;
    return MG_OK;
}

int maybe_rehome(mg_handle* h)
{
    if (h->force_zero && (h->rehome_due || h->steps_since_rehome >= kRehomeSteps)) return rehome(h);
    return MG_OK;
}

/* dynamic LDS of a sweep kernel: one Band<NX, CAP> per wavefront of the block */
template <int NX, int CAP = BAND_CAP> constexpr size_t band_lds(unsigned threads)
{
    return (threads / 64) * sizeof(Band<NX, CAP>);
}

/* the pressure and force launches for the handle's SPH kernels */
void launch_pressure(mg_handle* h, unsigned nb, const DS& d)
{
    if (h->sph == MG_SPH_POLY6_SPIKY)
        hipLaunchKernelGGL(k_pressure<MG_SPH_POLY6_SPIKY>, dim3(nb), dim3(256), band_lds<0>(256), h->st, d, h->p, h->g);
    else
        hipLaunchKernelGGL(k_pressure<MG_SPH_REFERENCE>, dim3(nb), dim3(256), band_lds<0>(256), h->st, d, h->p, h->g);
}

void launch_force(mg_handle* h, unsigned nb, const DS& d, int accumulate)
{
    if (h->sph == MG_SPH_POLY6_SPIKY)
        hipLaunchKernelGGL(k_force<MG_SPH_POLY6_SPIKY>, dim3(nb), dim3(256), band_lds<1>(256), h->st, d, h->p, h->g,
                           accumulate);
    else
        hipLaunchKernelGGL(k_force<MG_SPH_REFERENCE>, dim3(nb), dim3(256), band_lds<1>(256), h->st, d, h->p, h->g,
                           accumulate);
}

/* The FCONC float4 groups the diffusion sweep reads and the per-TF coefficients
 * tf_diffusability * diff_scale (the same float operations as the reference's loop). */
DiffPlan diff_plan(const mg_handle* h)
{
    DiffPlan dp;
    std::memset(&dp, 0, sizeof(dp));
    int used[4], n = 0;
    for (int q = 0; q < 4; q++) {
        bool any = false;
        for (int e = 0; e < 4; e++) any = any || h->gen.tf_diffusability[4 * q + e] != 0;
        if (any) used[n++] = q;
    }
    if (n == 3) {
        n = 4;
        for (int q = 0; q < 4; q++) used[q] = q;
    }
    dp.ng = n;
    for (int q = 0; q < 4; q++) dp.grp[q] = q < n ? used[q] : 0;
    for (int q = 0; q < n; q++)
        for (int e = 0; e < 4; e++) {
            const int t = 4 * used[q] + e;
            const uint32_t df = h->gen.tf_diffusability[t];
            dp.D[4 * q + e] = (float)df * h->p.diff_scale;
            if (df) dp.dm |= 1u << (4 * q + e);
        }
    return dp;
}

/* --------------------------------------------------------------- phases */
int do_insert(mg_handle* h, int gene_counts)
{
    DS d = make_ds(h);
    {
        PhaseScope ps(h, MG_PH_FILL);
        CHK(fill_u32(h, h->CNT[h->gcur], 0, h->C));
        CHK(fill_u32(h, h->OFF[h->gcur], 0, h->C));
        if (gene_counts) {
            CHK(fill_u32(h, h->GCNT, 0, (size_t)MG_NUM_GENES * h->C));
            CHK(fill_u32(h, h->GOFF, 0, (size_t)MG_NUM_GENES * h->C));
        }
    }
    PhaseScope ps(h, MG_PH_INSERT);
    hipLaunchKernelGGL(k_insert, dim3(nblk(h->N)), dim3(256), 0, h->st, d, h->g, gene_counts);
    HIPCHK(hipGetLastError());
    h->grid_scanned = false;
    h->pos_binned = true;
    h->cells_fresh = false;
    return MG_OK;
}

int grid_scan_max_index(const MgGrid& g)
{
    return (g.scan_max[1] * g.res[2] + g.scan_max[2]) * g.res[0] + g.scan_max[0];
}

/* Exclusive scans of `segs` (<= 16) consecutive segments of C words, in -> out, each segment's
 * total into lens[seg] (count[C-1] + offset[C-1], tally_denselist_lengths): tile sums, one block
 * per segment scanning its tile sums, tile prefixes. */
int seg_scan(mg_handle* h, const uint32_t* in, uint32_t* out, int segs, uint32_t* lens)
{
    const int C = h->C;
    const unsigned tiles = (unsigned)((C + SCAN_TILE - 1) / SCAN_TILE);
    if (!h->tsum) CHK(dalloc(h, &h->tsum, (size_t)2 * MG_NUM_GENES * tiles));
    uint32_t *sums = h->tsum, *excl = sums + (size_t)MG_NUM_GENES * tiles;
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles, segs), dim3(256), 0, h->st, in, C, sums, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_scan, dim3(segs), dim3(1024), 0, h->st, sums, excl, (int)tiles, h->ctr, 0, 0, lens);
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles, segs), dim3(256), 0, h->st, in, out, C, excl);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int do_scan(mg_handle* h, int genes)
{
    PhaseScope ps(h, MG_PH_SCAN);
    const unsigned tiles = (unsigned)((h->C + SCAN_TILE - 1) / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_lookback, dim3(tiles), dim3(256), 0, h->st, h->CNT[h->gcur], h->OFF[h->gcur], h->C,
                       h->scan_status, h->scan_ticket, h->ctr, grid_scan_max_index(h->g));
    HIPCHK(hipGetLastError());
    h->grid_scanned = true;
    if (genes) CHK(seg_scan(h, h->GCNT, h->GOFF, MG_NUM_GENES, h->LENS));
    return MG_OK;
}

int do_sort(mg_handle* h, int lists)
{
    DS d = make_ds(h);
    {
        PhaseScope ps(h, MG_PH_PLACE);
        hipLaunchKernelGGL(k_place, dim3(nblk(h->N)), dim3(256), 0, h->st, d);
    }
    {
        PhaseScope ps(h, MG_PH_SORT);
        const float vmax = std::max(std::max(h->g.gmax[0], h->g.gmax[1]), h->g.gmax[2]);
        hipLaunchKernelGGL(k_sort_gather, dim3(nblk(h->N)), dim3(256), 0, h->st, d, vmax);
    }
    HIPCHK(hipGetLastError());
    h->kcur = 1 - h->kcur;
    h->icur = 1 - h->icur;
    h->ccur = 1 - h->ccur;
    h->cells_fresh = h->pos_binned;
    h->rehome_due = true;
    h->force_zero = true;
    h->normalized = false;
    if (lists) {
        PhaseScope ps(h, MG_PH_LISTS);
        DS d2 = make_ds(h);
        CHK(fill_u32(h, h->LISTS, MG_NONE, (size_t)MG_NUM_GENES * h->N));
        hipLaunchKernelGGL(k_dense_lists, dim3(nblk(h->C)), dim3(256), 0, h->st, d2);
        HIPCHK(hipGetLastError());
    }
    return MG_OK;
}

/* The change-list buffers, allocated and cleared (pairs to MG_NONE) on first use. */
int ensure_changes(mg_handle* h)
{
    const size_t nc = (size_t)MG_NUM_CHANGES * h->C;
    if (!h->CCNT) { CHK(dalloc(h, &h->CCNT, nc)); CHK(fill_u32(h, h->CCNT, 0, nc)); }
    if (!h->COFF) { CHK(dalloc(h, &h->COFF, nc)); CHK(fill_u32(h, h->COFF, 0, nc)); }
    if (!h->CLEN) { CHK(dalloc(h, &h->CLEN, MG_NUM_CHANGES)); CHK(fill_u32(h, h->CLEN, 0, MG_NUM_CHANGES)); }
    const size_t np = (size_t)MG_NUM_CHANGES * 2 * h->N;
    if (!h->CL) { CHK(dalloc(h, &h->CL, np)); CHK(fill_u32(h, h->CL, MG_NONE, np)); }
    if (!h->HS) { CHK(dalloc(h, &h->HS, h->N)); CHK(fill_u32(h, h->HS, MG_NONE, h->N)); }
    return MG_OK;
}

/* Rebuild FBIN_COUNT_ACTIVE_GENES / FBIN_OFFSET_ACTIVE_GENES / FDENSE_LIST_LENGTHS /
 * FDENSE_LISTS from the sorted state after fused steps and Run2PhysicalSort (which skip them):
 * the values the reference's insert + scan + dense-list kernels leave after a sort. */
int materialize_gene_grid(mg_handle* h)
{
    if (!h->gene_grid_stale) return MG_OK;
    PhaseScope ps(h, MG_PH_LISTS);
    DS d = make_ds(h);
    CHK(fill_u32(h, h->GCNT, 0, (size_t)MG_NUM_GENES * h->C));
    hipLaunchKernelGGL(k_gene_counts_sorted, dim3(nblk(h->N)), dim3(256), 0, h->st, d);
    CHK(seg_scan(h, h->GCNT, h->GOFF, MG_NUM_GENES, h->LENS));
    CHK(fill_u32(h, h->LISTS, MG_NONE, (size_t)MG_NUM_GENES * h->N));
    hipLaunchKernelGGL(k_dense_lists, dim3(nblk(h->C)), dim3(256), 0, h->st, d);
    HIPCHK(hipGetLastError());
    h->gene_grid_stale = false;
    return MG_OK;
}

/* One step of the hot loop with the fused kernels. Starts and ends in the sorted state with
 * FFORCE = 0 (as after CountingSortFull). */
} // namespace

/* learned steps (mg_learn.hip) */
struct mg_net;
int launch_mlp_infer(mg_handle* h, const mg_net* net, float dt);
uint64_t net_version(const mg_net* net);

namespace {

/* The z displacement of externally actuated particles for a step starting at the handle's
 * time: actuation_factor * sin(time / (actuation_period * dt)) in float
 * (src/fluid_system_cuda.cu:3682), evaluated once on the host. */
float actuation_offset(const mg_handle* h, float dt)
{
    const float t = (float)h->time;
    return h->p.actuation_factor * sinf(t / (h->p.actuation_period * dt));
}

/* Graph replay bakes kernel arguments in: a step whose actuation offset changes with time
 * runs eagerly. */
bool graph_ok(const mg_handle* h, unsigned flags)
{
    return !(flags & MG_STEP_SPECIAL) || h->p.actuation_factor == 0.0f;
}

/* Hit masks for the fused step: the force and diffusion sweeps reuse the pressure sweep's
 * neighbour sets (recorded within a skin of 0.01 around the cut-off) when the sweeps walk rows
 * of the sorted grid (rows of srch <= 4 cells) and share the cut-off (pressure: dsq < r2,
 * force: dsq < 1, diffusion: dsq < r2 at the advanced positions, while no particle moved
 * farther than the skin allows). ensure_hit_masks allocates the buffer outside
 * any stream capture (a new buffer bumps the epoch, so captured graphs are rebuilt);
 * hit_masks returns it when the current state may use it. */
bool hm_eligible(const mg_handle* h) { return h->adj_rows && h->p.r2 == 1.0f && h->g.srch <= 4; }

void ensure_hit_masks(mg_handle* h)
{
    const int nseg = h->g.srch * h->g.srch;
    if (!hm_eligible(h) || h->hm_rows >= nseg) return;
    if (h->HM) dfree(h, h->HM);
    h->HM = nullptr;
    h->hm_rows = 0;
    h->epoch++;
    if (dalloc(h, &h->HM, (size_t)nseg * h->N) != MG_OK) {
        h->HM = nullptr; /* no masks: the force sweep tests its candidates itself */
        return;
    }
    h->hm_rows = nseg;
    if (h->ET) dfree(h, h->ET);
    h->ET = nullptr;
    if (h->B > 0 && dalloc(h, &h->ET, (size_t)h->B * h->N) != MG_OK)
        h->ET = nullptr; /* no records: the force evaluates its bonds itself */
}

uint64_t* hit_masks(const mg_handle* h)
{
    return (h->grid_scanned && hm_eligible(h) && h->hm_rows >= h->g.srch * h->g.srch) ? h->HM : nullptr;
}

/* the bond records go with the hit masks: written by a pressure pass that records masks and
 * read by the force pass that visits them (the same positions) */
float4* bond_records(const mg_handle* h, const uint64_t* hm) { return hm ? h->ET : nullptr; }

int do_fused_step(mg_handle* h, float dt, unsigned flags, const mg_net* net = nullptr)
{
    const int genes = (flags & MG_STEP_GENES) ? 1 : 0;
    const unsigned nb = nblk(h->N);
    uint64_t* hm = hit_masks(h);
    {
        PhaseScope ps(h, MG_PH_PRESSURE);
        DS d = make_ds(h);
        d.ZERO = h->CNT[1 - h->gcur];
        d.zero_n = (uint32_t)h->C;
        d.HMW = hm;
        d.ETW = bond_records(h, hm);
        launch_pressure(h, nb, d);
    }
    const int special = (flags & MG_STEP_SPECIAL) ? 1 : 0;
    const float zoff = special ? actuation_offset(h, dt) : 0.0f;
    if (net) {
        /* learned coarse step: the network reads the step's acceleration, so the force pass
         * (FFORCE = 0 + force), the inference and the advance are separate launches */
        {
            PhaseScope ps(h, MG_PH_FORCE);
            DS d = make_ds(h);
            d.HMR = hm;
            d.ETR = bond_records(h, hm);
            launch_force(h, nb, d, 0);
        }
        {
            PhaseScope ps(h, MG_PH_LEARN);
            CHK(launch_mlp_infer(h, net, dt));
        }
        {
            PhaseScope ps(h, MG_PH_ADVANCE);
            DS d = make_ds(h);
            d.DVc = h->DV;
            d.HMR = hm;
            hipLaunchKernelGGL(k_advance_learned, dim3(nb), dim3(256), 0, h->st, d, h->p, dt, special, zoff);
        }
    } else {
        PhaseScope ps(h, MG_PH_FORCE);
        DS d = make_ds(h);
        d.HMR = hm;
        d.ETR = bond_records(h, hm);
        if (h->sph == MG_SPH_POLY6_SPIKY)
            hipLaunchKernelGGL(k_force_advance<MG_SPH_POLY6_SPIKY>, dim3(nb), dim3(256), band_lds<1>(256), h->st, d,
                               h->p, h->g, dt, special, zoff);
        else
            hipLaunchKernelGGL(k_force_advance<MG_SPH_REFERENCE>, dim3(nb), dim3(256), band_lds<1>(256), h->st, d,
                               h->p, h->g, dt, special, zoff);
    }
    {
        PhaseScope ps(h, MG_PH_DIFFUSION);
        DS d = make_ds(h, 1, 0);
        d.GMN = h->GMN;
        d.cull = 0; /* positions moved since the binning */
        d.HMR = hm; /* within the skin, unless the force-advance raised MG_C_SKIN */
        const DiffPlan dp = diff_plan(h);
        if (dp.ng <= 1)
            hipLaunchKernelGGL(k_diffuse_genes_insert<1>, dim3(nb), dim3(256), band_lds<1>(256), h->st, d, h->p, h->g, h->gen, dp, genes);
        else if (dp.ng == 2)
            hipLaunchKernelGGL(k_diffuse_genes_insert<2>, dim3(nb), dim3(256), band_lds<2>(256), h->st, d, h->p, h->g, h->gen, dp, genes);
        else
            hipLaunchKernelGGL(k_diffuse_genes_insert<4>, dim3(nb), dim3(256), band_lds<4>(256), h->st, d, h->p, h->g, h->gen, dp, genes);
    }
    {
        PhaseScope ps(h, MG_PH_SCAN);
        const unsigned tiles = (unsigned)((h->C + SCAN_TILE - 1) / SCAN_TILE);
        hipLaunchKernelGGL(k_scan_lookback, dim3(tiles), dim3(256), 0, h->st, h->CNT[1 - h->gcur],
                           h->OFF[1 - h->gcur], h->C, h->scan_status, h->scan_ticket, h->ctr,
                           grid_scan_max_index(h->g));
    }
    {
        PhaseScope ps(h, MG_PH_PLACE);
        DS d = make_ds(h);
        d.CNT = d.CNTN;
        d.OFF = d.OFFN;
        hipLaunchKernelGGL(k_place, dim3(nb), dim3(256), 0, h->st, d);
    }
    {
        PhaseScope ps(h, MG_PH_SORT);
        DS d = make_ds(h, 1, 1);
        d.GMN = h->GMN;
        d.CNT = d.CNTN; /* the grid k_place binned into */
        d.OFF = d.OFFN;
        const float vmax = std::max(std::max(h->g.gmax[0], h->g.gmax[1]), h->g.gmax[2]);
        hipLaunchKernelGGL(k_sort_gather, dim3(nb), dim3(256), 0, h->st, d, vmax);
    }
    HIPCHK(hipGetLastError());
    h->icur = 1 - h->icur;
    h->gcur = 1 - h->gcur;
    h->force_zero = true;
    h->normalized = false;
    h->gene_grid_stale = true;
    h->grid_scanned = true;
    h->pos_binned = true;
    h->cells_fresh = true;
    return MG_OK;
}

/* Two fused steps captured once into a hipGraph and replayed; the step has a two-step
 * period in its buffer roles (icur, gcur flip every step), so a two-step graph leaves the
 * handle's indices where they started. Recaptured when dt, flags, the parities or the
 * handle's epoch change. */
int graph_steps(mg_handle* h, int pairs, float dt, unsigned flags, const mg_net* net = nullptr)
{
    ensure_hit_masks(h); /* allocates (and bumps the epoch) before any capture, not inside it */
    const int parity = ((h->cells_fresh && h->grid_scanned) << 4) | (h->kcur << 3) | (h->icur << 2) |
                       (h->ccur << 1) | h->gcur;
    const uint64_t nv = net ? net_version(net) : 0;
    if (!h->graph || h->graph_epoch != h->epoch || h->graph_dt != dt || h->graph_flags != flags ||
        h->graph_parity != parity || h->graph_net != (const void*)net || h->graph_net_version != nv) {
        if (h->graph) {
            (void)hipGraphExecDestroy(h->graph);
            h->graph = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        int r1 = do_fused_step(h, dt, flags, net);
        int r2 = (r1 == MG_OK) ? do_fused_step(h, dt, flags, net) : r1;
        hipError_t ec = hipStreamEndCapture(h->st, &g);
        if (r1 != MG_OK || r2 != MG_OK) return r1 != MG_OK ? r1 : r2;
        if (ec != hipSuccess) {
            set_err("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
            return MG_ERR_HIP;
        }
        hipError_t ei = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            h->graph = nullptr;
            set_err("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
            return MG_ERR_HIP;
        }
        h->graph_epoch = h->epoch;
        h->graph_dt = dt;
        h->graph_flags = flags;
        h->graph_parity = parity;
        h->graph_net = net;
        h->graph_net_version = nv;
    }
    for (int k = 0; k < pairs; k++) HIPCHK(hipGraphLaunch(h->graph, h->st));
    h->force_zero = true;
    h->normalized = false;
    h->gene_grid_stale = true;
    h->grid_scanned = true;
    h->pos_binned = true;
    h->cells_fresh = true;
    return MG_OK;
}

int ensure_force(mg_handle* h)
{
    if (h->force_zero) {
        PhaseScope ps(h, MG_PH_FILL);
        CHK(fill_u32(h, h->F, 0, (size_t)h->N * 4));
        h->force_zero = false;
    }
    return MG_OK;
}

/* hm: hit masks to record (pressure) or to visit (force); only a pressure and a force of the
 * same positions and grid may share them, as inside one inner loop */
int do_pressure(mg_handle* h, uint64_t* hm = nullptr)
{
    PhaseScope ps(h, MG_PH_PRESSURE);
    DS d = make_ds(h);
    d.HMW = hm;
    d.ETW = bond_records(h, hm);
    launch_pressure(h, nblk(h->N), d);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int do_force(mg_handle* h, const uint64_t* hm = nullptr)
{
    CHK(ensure_force(h));
    PhaseScope ps(h, MG_PH_FORCE);
    DS d = make_ds(h);
    d.HMR = hm;
    d.ETR = bond_records(h, hm);
    launch_force(h, nblk(h->N), d, 1);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int do_advance(mg_handle* h, float dt, unsigned flags = 0)
{
    CHK(ensure_force(h));
    PhaseScope ps(h, MG_PH_ADVANCE);
    const int special = (flags & MG_STEP_SPECIAL) ? 1 : 0;
    const float zoff = special ? actuation_offset(h, dt) : 0.0f;
    hipLaunchKernelGGL(k_advance, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h), h->p, dt, special, zoff);
    HIPCHK(hipGetLastError());
    h->cells_fresh = false;
    h->pos_binned = false;
    return MG_OK;
}

int do_diffusion(mg_handle* h)
{
    PhaseScope ps(h, MG_PH_DIFFUSION);
    const DiffPlan dp = diff_plan(h);
    const DS d = make_ds(h);
    if (dp.ng <= 1)
        hipLaunchKernelGGL(k_diffusion<1>, dim3(nblk(h->N)), dim3(256), band_lds<1>(256), h->st, d, h->p, h->g, h->gen, dp);
    else if (dp.ng == 2)
        hipLaunchKernelGGL(k_diffusion<2>, dim3(nblk(h->N)), dim3(256), band_lds<2>(256), h->st, d, h->p, h->g, h->gen, dp);
    else
        hipLaunchKernelGGL(k_diffusion<4>, dim3(nblk(h->N)), dim3(256), band_lds<4>(256), h->st, d, h->p, h->g, h->gen, dp);
    HIPCHK(hipGetLastError());
    h->ccur = 1 - h->ccur;
    return MG_OK;
}

int do_genes(mg_handle* h)
{
    PhaseScope ps(h, MG_PH_GENES);
    hipLaunchKernelGGL(k_genes, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h), h->p, h->gen);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

} // namespace

/* ================================================================== C-ABI */
extern "C" {

const char* mg_last_error(void) { return g_err.c_str(); }

int mg_create(mg_handle** out, int max_points, int bonds_per_particle)
{
    if (!out || max_points <= 0 || bonds_per_particle < 0 || bonds_per_particle > 64) {
        set_err("mg_create: bad arguments (max_points=%d, B=%d)", max_points, bonds_per_particle);
        return MG_ERR_ARG;
    }
    if ((size_t)max_points * (size_t)(bonds_per_particle > 0 ? bonds_per_particle : 1) >= 0xFFFFFFFFull) {
        set_err("mg_create: max_points * B must fit in 32 bits");
        return MG_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_err("mg_create: no HIP device");
        return MG_ERR_HIP;
    }
    mg_handle* h = new mg_handle;
    h->N = max_points;
    h->B = bonds_per_particle;
    std::memset(&h->p, 0, sizeof(h->p));
    std::memset(&h->g, 0, sizeof(h->g));
    std::memset(&h->gen, 0, sizeof(h->gen));
    auto fail = [&](int r) { mg_destroy(h); return r; };
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        set_err("mg_create: hipStreamCreate failed");
        delete h;
        return MG_ERR_HIP;
    }
    h->own_st = h->st;
    const size_t N = (size_t)max_points, B = (size_t)bonds_per_particle;
    int r = MG_OK;
    for (int k = 0; k < 2 && r == MG_OK; k++) {
        r = dalloc(h, &h->P[k], N + MG_P_PAD);
        if (r == MG_OK) r = dalloc(h, &h->V[k], N);
        if (r == MG_OK) r = dalloc(h, &h->VEL[k], N);
        if (r == MG_OK) r = dalloc(h, &h->CONC[k], N * MG_NUM_TF);
        if (r == MG_OK) r = dalloc(h, &h->HOME[k], N);
        if (r == MG_OK) r = dalloc(h, &h->GCELL[k], N);
        if (r == MG_OK) r = dalloc(h, &h->GNDX[k], N);
        if (r == MG_OK) r = dalloc(h, &h->GMASK[k], N);
    }
    if (r == MG_OK) r = dalloc(h, &h->F, N);
    if (r == MG_OK) r = dalloc(h, &h->SLOT, N);
    if (r == MG_OK) r = dalloc(h, &h->EPIGEN, N * MG_NUM_GENES);
    if (r == MG_OK) r = dalloc(h, &h->BPARTNER, std::max<size_t>(B * N, 1));
    if (r == MG_OK) r = dalloc(h, &h->BPAR, std::max<size_t>(B * N, 1));
    if (r == MG_OK) r = dalloc(h, &h->BINT, std::max<size_t>(B * N, 1));
    if (r == MG_OK) r = dalloc(h, &h->BCOLD, std::max<size_t>(B * N, 1));
    if (r == MG_OK) r = dalloc(h, &h->PIDX, std::max<size_t>(B * N, 1));
    if (r == MG_OK) r = dalloc(h, &h->AGE, N);
    if (r == MG_OK) r = dalloc(h, &h->COLOR, N);
    if (r == MG_OK) r = dalloc(h, &h->PID, N);
    if (r == MG_OK) r = dalloc(h, &h->MRAD, N);
    if (r == MG_OK) r = dalloc(h, &h->NERVE, N);
    if (r == MG_OK) r = dalloc(h, &h->FBIN, N);
    if (r == MG_OK) r = dalloc(h, &h->INC_OFF, N + 1);
    if (r == MG_OK) r = dalloc(h, &h->LIVE, N);
    if (r == MG_OK) r = dalloc(h, &h->PERM, N);
    if (r == MG_OK) r = dalloc(h, &h->GMN, N);
    if (r == MG_OK) r = dalloc(h, &h->LISTS, N * MG_NUM_GENES);
    if (r == MG_OK) r = dalloc(h, &h->LENS, MG_NUM_GENES);
    if (r == MG_OK) r = dalloc(h, &h->ctr, MG_C_COUNT);
    if (r != MG_OK) return fail(r);
    /* initial state: what the reference's host allocation leaves (zeroed buffers, sentinel
     * bonds), every slot its own home, no sort yet */
    const int n = max_points;
    for (int k = 0; k < 2; k++) {
        fill_u32(h, h->P[k], 0, (N + MG_P_PAD) * 4);
        fill_u32(h, h->V[k], 0, N * 4);
        fill_u32(h, h->VEL[k], 0, N * 4);
        fill_u32(h, h->CONC[k], 0, N * MG_NUM_TF);
        fill_u32(h, h->GCELL[k], MG_GRID_UNDEF, N);
        fill_u32(h, h->GNDX[k], 0, N);
        fill_u32(h, h->GMASK[k], 0, N);
    }
    fill_u32(h, h->F, 0, N * 4);
    fill_u32(h, h->EPIGEN, 0, N * MG_NUM_GENES);
    fill_u32(h, h->BPARTNER, MG_NONE, B * N);
    fill_u32(h, h->BPAR, MG_NONE, B * N * 4);
    fill_u32(h, h->BINT, MG_NONE, B * N * 2);
    fill_u32(h, h->BCOLD, MG_NONE, B * N * 2);
    fill_u32(h, h->PIDX, MG_NONE, B * N * 2);
    fill_u32(h, h->AGE, 0, N);
    fill_u32(h, h->COLOR, 0, N);
    fill_u32(h, h->PID, MG_NONE, N);
    fill_u32(h, h->MRAD, 0, N);
    fill_u32(h, h->NERVE, 0, N);
    fill_u32(h, h->FBIN, 0, N);
    fill_u32(h, h->INC_OFF, 0, N + 1);
    fill_u32(h, h->LIVE, 0, N * 2);
    fill_u32(h, h->LISTS, MG_NONE, N * MG_NUM_GENES);
    fill_u32(h, h->LENS, 0, MG_NUM_GENES);
    hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, h->st, h->HOME[0], n);
    hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, h->st, h->SLOT, n);
    uint32_t c0[MG_C_COUNT] = {0};
    c0[MG_C_NUM_POINTS] = (uint32_t)n;
    c0[MG_C_ACTIVE] = (uint32_t)n;
    c0[MG_C_TAIL] = (uint32_t)n;
    if (hipMemcpyAsync(h->ctr, c0, sizeof(c0), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
        set_err("mg_create: initialisation failed");
        return fail(MG_ERR_HIP);
    }
    *out = h;
    return MG_OK;
}

int mg_destroy(mg_handle* h)
{
    if (!h) return MG_ERR_ARG;
    if (h->batch) {
        set_err("mg_destroy: the handle belongs to a batch; destroy the batch first");
        return MG_ERR_STATE;
    }
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (void* q : h->allocs) (void)hipFree(q);
    h->allocs.clear();
    for (hipEvent_t e : h->pool) (void)hipEventDestroy(e);
    h->pool.clear();
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
    return MG_OK;
}

int mg_set_grid(mg_handle* h, const MgGrid* g)
{
    if (!h || !g || g->total <= 0 || g->adj_cnt <= 0 || g->adj_cnt > MG_MAX_ADJ) {
        set_err("mg_set_grid: bad grid");
        return MG_ERR_ARG;
    }
    for (int k = 0; k < 3; k++)
        if (g->scan_max[k] < 1 || g->scan_max[k] + g->srch > g->res[k]) {
            set_err("mg_set_grid: scan window outside the grid");
            return MG_ERR_ARG;
        }
    if (g->res[0] * g->res[1] * g->res[2] != g->total) {
        set_err("mg_set_grid: total != res.x*res.y*res.z");
        return MG_ERR_ARG;
    }
    {
        /* every neighbour cell of an in-window particle must be a grid cell */
        const long long gc0max = ((long long)(g->scan_max[1] - 1) * g->res[2] + (g->scan_max[2] - 1)) * g->res[0] +
                                 (g->scan_max[0] - 1);
        for (int c = 0; c < g->adj_cnt; c++)
            if (g->adj[c] < 0 || gc0max + g->adj[c] >= (long long)g->total) {
                set_err("mg_set_grid: gridAdj[%d] = %d leaves the grid", c, g->adj[c]);
                return MG_ERR_ARG;
            }
    }
    if (h->have_grid && g->total != h->C) {
        for (int k = 0; k < 2; k++) { dfree(h, h->CNT[k]); dfree(h, h->OFF[k]); }
        dfree(h, h->GCNT); dfree(h, h->GOFF);
        dfree(h, h->CCNT); dfree(h, h->COFF); dfree(h, h->tsum);
        h->CCNT = h->COFF = h->tsum = nullptr;
        h->have_grid = false;
    }
    if (!h->have_grid) {
        h->C = g->total;
        for (int k = 0; k < 2; k++) {
            CHK(dalloc(h, &h->CNT[k], h->C));
            CHK(dalloc(h, &h->OFF[k], h->C));
            CHK(fill_u32(h, h->CNT[k], 0, h->C));
            CHK(fill_u32(h, h->OFF[k], 0, h->C));
        }
        if (h->scan_status) { dfree(h, h->scan_status); h->scan_status = nullptr; }
        if (!h->scan_ticket) CHK(dalloc(h, &h->scan_ticket, 1));
        CHK(dalloc(h, &h->scan_status, (size_t)(h->C + SCAN_TILE - 1) / SCAN_TILE));
        CHK(fill_u32(h, h->scan_status, 0, 2 * ((size_t)(h->C + SCAN_TILE - 1) / SCAN_TILE)));
        CHK(fill_u32(h, h->scan_ticket, 0, 2));
        CHK(dalloc(h, &h->GCNT, (size_t)MG_NUM_GENES * h->C));
        CHK(dalloc(h, &h->GOFF, (size_t)MG_NUM_GENES * h->C));
        CHK(fill_u32(h, h->GCNT, 0, (size_t)MG_NUM_GENES * h->C));
        CHK(fill_u32(h, h->GOFF, 0, (size_t)MG_NUM_GENES * h->C));
    }
    h->g = *g;
    h->have_grid = true;
    /* row walks need gridAdj[(y * srch + z) * srch + x] == (y * res.z + z) * res.x + x */
    h->adj_rows = g->srch >= 1 && g->srch <= 4 && g->adj_cnt == g->srch * g->srch * g->srch;
    for (int y = 0; h->adj_rows && y < g->srch; y++)
        for (int z = 0; z < g->srch; z++)
            for (int x = 0; x < g->srch; x++)
                if (g->adj[(y * g->srch + z) * g->srch + x] != (y * g->res[2] + z) * g->res[0] + x)
                    h->adj_rows = false;
    h->grid_scanned = false;
    h->epoch++;
    return MG_OK;
}

int mg_fluid_setup(mg_handle* h, int num, int gsrch, const int32_t res[3], const float size[3],
                   const float delta[3], const float gmin[3], const float gmax[3], int total, int chk)
{
    if (!h || !res || !size || !delta || !gmin || !gmax) { set_err("mg_fluid_setup: null argument"); return MG_ERR_ARG; }
    if (num != h->N) { set_err("mg_fluid_setup: num (%d) != max_points (%d)", num, h->N); return MG_ERR_ARG; }
    MgGrid g;
    if (mg_grid_from_setup(&g, gsrch, res, size, delta, gmin, gmax, total, chk) != MG_OK) {
        set_err("mg_fluid_setup: bad gsrch");
        return MG_ERR_ARG;
    }
    return mg_set_grid(h, &g);
}

int mg_set_params(mg_handle* h, const MgParams* p)
{
    if (!h || !p) { set_err("mg_set_params: null argument"); return MG_ERR_ARG; }
    h->p = *p;
    h->have_params = true;
    h->epoch++;
    return MG_OK;
}

int mg_fluid_param(mg_handle* h, float ss, float sr, float pr, float mass, float rest,
                   const float bmin[3], const float bmax[3], float estiff, float istiff, float visc,
                   float surface_t, float damp, float fmin, float fmax, float ffreq, float gslope,
                   float gx, float gy, float gz, float al, float vl, float a_f, float a_p)
{
    if (!h || !bmin || !bmax) { set_err("mg_fluid_param: null argument"); return MG_ERR_ARG; }
    MgParams p;
    mg_params_from_fluid_param(&p, ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff, visc,
                               surface_t, damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl, a_f, a_p);
    if (h->have_params) {
        p.freeze = h->p.freeze;
        p.diff_scale = h->p.diff_scale;
        p.secrete_scale = h->p.secrete_scale;
    }
    return mg_set_params(h, &p);
}

int mg_derive_params(MgParams* out, float ss, float sr, float pr, float mass, float rest,
                     const float bmin[3], const float bmax[3], float estiff, float istiff,
                     float visc, float surface_t, float damp, float fmin, float fmax, float ffreq,
                     float gslope, float gx, float gy, float gz, float al, float vl, float a_f,
                     float a_p)
{
    if (!out || !bmin || !bmax) { set_err("mg_derive_params: null argument"); return MG_ERR_ARG; }
    mg_params_from_fluid_param(out, ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff, visc,
                               surface_t, damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl, a_f,
                               a_p);
    return MG_OK;
}

int mg_derive_grid(MgGrid* out, const float vmin[3], const float vmax[3], float sim_scale,
                   float cell_size, float smooth_radius)
{
    if (!out || !vmin || !vmax) { set_err("mg_derive_grid: null argument"); return MG_ERR_ARG; }
    if (!(sim_scale > 0.0f) || !(cell_size > 0.0f) || !(smooth_radius > 0.0f)) {
        set_err("mg_derive_grid: sim_scale, cell_size and smooth_radius must be positive");
        return MG_ERR_ARG;
    }
    if (mg_grid_setup(out, vmin, vmax, sim_scale, cell_size, smooth_radius) != MG_OK) {
        set_err("mg_derive_grid: the search width exceeds %d cells", MG_MAX_ADJ);
        return MG_ERR_ARG;
    }
    return MG_OK;
}

int mg_get_params(mg_handle* h, MgParams* out)
{
    if (!h || !out) { set_err("mg_get_params: null argument"); return MG_ERR_ARG; }
    if (!h->have_params) { set_err("mg_get_params: no parameters set"); return MG_ERR_STATE; }
    *out = h->p;
    return MG_OK;
}

int mg_get_grid(mg_handle* h, MgGrid* out)
{
    if (!h || !out) { set_err("mg_get_grid: null argument"); return MG_ERR_ARG; }
    if (!h->have_grid) { set_err("mg_get_grid: no grid set"); return MG_ERR_STATE; }
    *out = h->g;
    return MG_OK;
}

int mg_debug_set_scan_ticket(mg_handle* h, uint64_t ticket)
{
    if (!h) { set_err("mg_debug_set_scan_ticket: null handle"); return MG_ERR_ARG; }
    if (!h->scan_ticket || !h->have_grid) { set_err("mg_debug_set_scan_ticket: no grid set"); return MG_ERR_STATE; }
    /* every scan launch takes exactly `tiles` tickets, so the ticket is always a multiple of
     * the tile count between launches; another value would split a launch over two epochs */
    const uint64_t tiles = (uint64_t)(h->C + SCAN_TILE - 1) / SCAN_TILE;
    if (ticket % tiles != 0) {
        set_err("mg_debug_set_scan_ticket: %llu is not a multiple of the %llu scan tiles",
                (unsigned long long)ticket, (unsigned long long)tiles);
        return MG_ERR_ARG;
    }
    unsigned long long t = ticket;
    HIPCHK(hipMemcpyAsync(h->scan_ticket, &t, sizeof(t), hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

int mg_set_genome(mg_handle* h, const MgGenome* g)
{
    if (!h || !g) { set_err("mg_set_genome: null argument"); return MG_ERR_ARG; }
    h->gen = *g;
    h->epoch++;
    return MG_OK;
}

int mg_set_sph_kernels(mg_handle* h, int mode)
{
    if (!h) { set_err("mg_set_sph_kernels: null handle"); return MG_ERR_ARG; }
    if (mode != MG_SPH_REFERENCE && mode != MG_SPH_POLY6_SPIKY) {
        set_err("mg_set_sph_kernels: unknown mode %d", mode);
        return MG_ERR_ARG;
    }
    if (h->batch && mode != MG_SPH_REFERENCE) {
        set_err("mg_set_sph_kernels: the handle is in a batch (batches step the reference kernels)");
        return MG_ERR_UNSUPPORTED;
    }
    h->sph = mode;
    h->epoch++; /* captured graphs hold the old kernels */
    return MG_OK;
}

int mg_get_sph_kernels(mg_handle* h, int* mode)
{
    if (!h || !mode) return MG_ERR_ARG;
    *mode = h->sph;
    return MG_OK;
}

int mg_set_freeze(mg_handle* h, int freeze)
{
    if (!h) { set_err("mg_set_freeze: null handle"); return MG_ERR_ARG; }
    h->p.freeze = freeze ? 1 : 0;
    h->epoch++;
    return MG_OK;
}

int mg_upload(mg_handle* h, int buf_id, const void* src, size_t bytes)
{
    if (!h || (!src && bytes)) { set_err("mg_upload: null argument"); return MG_ERR_ARG; }
    const int N = h->N, B = h->B;
    if (!h->have_grid && mg_bytes_per_particle(buf_id, B) == 0) {
        set_err("mg_upload: grid buffer %d before mg_set_grid", buf_id);
        return MG_ERR_STATE;
    }
    if (!mg_buffer_known(buf_id)) { set_err("mg_upload: unsupported buffer id %d", buf_id); return MG_ERR_UNSUPPORTED; }
    const size_t need = mg_buffer_bytes(buf_id, N, h->C, B);
    if (bytes != need) { set_err("mg_upload: buffer %d needs %zu bytes, got %zu", buf_id, need, bytes); return MG_ERR_SIZE; }
    if (need == 0) return MG_OK;
    CHK(normalize(h));
    h->cells_fresh = false;
    h->pos_binned = false;
    if (buf_id >= MG_FBIN_COUNT_ACTIVE_GENES && buf_id <= MG_FDENSE_LISTS) CHK(materialize_gene_grid(h));
    if (buf_id >= MG_FBIN_OFFSET_CHANGES && buf_id <= MG_FDENSE_LISTS_CHANGES) CHK(ensure_changes(h));
    CHK(staging(h, bytes));
    HIPCHK(hipMemcpyAsync(h->staging, src, bytes, hipMemcpyHostToDevice, h->st));
    const int c = h->kcur, ic = h->icur;
    const unsigned nb = nblk(N);
    void* stg = h->staging;
    switch (buf_id) {
    case MG_FPOS: hipLaunchKernelGGL(k_set_xyz, dim3(nb), dim3(256), 0, h->st, h->P[c], (const float4*)stg, N); break;
    case MG_FVEL: hipLaunchKernelGGL(k_set_xyz, dim3(nb), dim3(256), 0, h->st, h->VEL[c], (const float4*)stg, N); break;
    case MG_FVEVAL: hipLaunchKernelGGL(k_set_xyz, dim3(nb), dim3(256), 0, h->st, h->V[c], (const float4*)stg, N); break;
    case MG_FFORCE:
        hipLaunchKernelGGL(k_set_xyz, dim3(nb), dim3(256), 0, h->st, h->F, (const float4*)stg, N);
        h->force_zero = false;
        break;
    case MG_FPRESS: hipLaunchKernelGGL(k_set_w, dim3(nb), dim3(256), 0, h->st, h->P[c], (const float*)stg, N); break;
    case MG_FDENSITY: hipLaunchKernelGGL(k_set_w, dim3(nb), dim3(256), 0, h->st, h->V[c], (const float*)stg, N); break;
    case MG_FAGE: HIPCHK(hipMemcpyAsync(h->AGE, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FCOLOR: HIPCHK(hipMemcpyAsync(h->COLOR, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FGCELL: HIPCHK(hipMemcpyAsync(h->GCELL[ic], stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FGNDX: HIPCHK(hipMemcpyAsync(h->GNDX[ic], stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FNERVEIDX: HIPCHK(hipMemcpyAsync(h->NERVE, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FPARTICLE_ID:
        HIPCHK(hipMemcpyAsync(h->PID, stg, bytes, hipMemcpyDeviceToDevice, h->st));
        if (B > 0) CHK(rebuild_inc(h)); /* the incoming-bond order follows the owners' IDs */
        break;
    case MG_FMASS_RADIUS: HIPCHK(hipMemcpyAsync(h->MRAD, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FBIN: {
        HIPCHK(hipMemcpyAsync(h->FBIN, stg, bytes, hipMemcpyDeviceToDevice, h->st));
        uint32_t z = 0;
        HIPCHK(hipMemcpyAsync(h->ctr + MG_C_HWM, &z, 4, hipMemcpyHostToDevice, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        break;
    }
    case MG_FCONC: HIPCHK(hipMemcpyAsync(h->CONC[h->ccur], stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FEPIGEN: HIPCHK(hipMemcpyAsync(h->EPIGEN, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FELASTIDX:
        if (B > 0) {
            hipLaunchKernelGGL(k_upload_bonds, dim3(nb, B), dim3(256), 0, h->st, h->BPARTNER, h->BPAR,
                               h->BINT, h->BCOLD, (const uint32_t*)stg, N, B);
            HIPCHK(hipGetLastError());
            CHK(rebuild_inc(h));
        }
        break;
    case MG_FPARTICLEIDX:
        if (B > 0)
            hipLaunchKernelGGL(k_upload_pidx, dim3(nb, B), dim3(256), 0, h->st, h->PIDX, (const uint32_t*)stg, N, B);
        break;
    case MG_FBIN_COUNT:
        HIPCHK(hipMemcpyAsync(h->CNT[h->gcur], stg, bytes, hipMemcpyDeviceToDevice, h->st));
        h->grid_scanned = false;
        break;
    case MG_FBIN_OFFSET:
        HIPCHK(hipMemcpyAsync(h->OFF[h->gcur], stg, bytes, hipMemcpyDeviceToDevice, h->st));
        h->grid_scanned = false;
        break;
    case MG_FBIN_COUNT_ACTIVE_GENES: HIPCHK(hipMemcpyAsync(h->GCNT, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FBIN_OFFSET_ACTIVE_GENES: HIPCHK(hipMemcpyAsync(h->GOFF, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FDENSE_LIST_LENGTHS: HIPCHK(hipMemcpyAsync(h->LENS, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FDENSE_LISTS: HIPCHK(hipMemcpyAsync(h->LISTS, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FBIN_COUNT_CHANGES: HIPCHK(hipMemcpyAsync(h->CCNT, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FBIN_OFFSET_CHANGES: HIPCHK(hipMemcpyAsync(h->COFF, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FDENSE_LIST_LENGTHS_CHANGES: HIPCHK(hipMemcpyAsync(h->CLEN, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    case MG_FDENSE_LISTS_CHANGES: HIPCHK(hipMemcpyAsync(h->CL, stg, bytes, hipMemcpyDeviceToDevice, h->st)); break;
    default: set_err("mg_upload: unsupported buffer id %d", buf_id); return MG_ERR_UNSUPPORTED;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

int mg_download(mg_handle* h, int buf_id, void* dst, size_t bytes)
{
    if (!h || (!dst && bytes)) { set_err("mg_download: null argument"); return MG_ERR_ARG; }
    const int N = h->N, B = h->B;
    if (!h->have_grid && mg_bytes_per_particle(buf_id, B) == 0) {
        set_err("mg_download: grid buffer %d before mg_set_grid", buf_id);
        return MG_ERR_STATE;
    }
    if (!mg_buffer_known(buf_id)) { set_err("mg_download: unsupported buffer id %d", buf_id); return MG_ERR_UNSUPPORTED; }
    const size_t need = mg_buffer_bytes(buf_id, N, h->C, B);
    if (bytes != need) { set_err("mg_download: buffer %d needs %zu bytes, got %zu", buf_id, need, bytes); return MG_ERR_SIZE; }
    if (need == 0) return MG_OK;
    if (buf_id >= MG_FBIN_COUNT_ACTIVE_GENES && buf_id <= MG_FDENSE_LISTS) CHK(materialize_gene_grid(h));
    if (buf_id >= MG_FBIN_OFFSET_CHANGES && buf_id <= MG_FDENSE_LISTS_CHANGES) CHK(ensure_changes(h));
    CHK(staging(h, bytes));
    const int c = h->kcur, ic = h->icur;
    const unsigned nb = nblk(N);
    void* stg = h->staging;
    const uint32_t* HOME = h->HOME[ic];
    const void* direct = nullptr;
    switch (buf_id) {
    case MG_FPOS: hipLaunchKernelGGL(k_get_xyz, dim3(nb), dim3(256), 0, h->st, (float4*)stg, h->P[c], N, 0); break;
    case MG_FVEL: hipLaunchKernelGGL(k_get_xyz, dim3(nb), dim3(256), 0, h->st, (float4*)stg, h->VEL[c], N, 0); break;
    case MG_FVEVAL: hipLaunchKernelGGL(k_get_xyz, dim3(nb), dim3(256), 0, h->st, (float4*)stg, h->V[c], N, 0); break;
    case MG_FFORCE: hipLaunchKernelGGL(k_get_xyz, dim3(nb), dim3(256), 0, h->st, (float4*)stg, h->F, N, h->force_zero ? 1 : 0); break;
    case MG_FPRESS: hipLaunchKernelGGL(k_get_w, dim3(nb), dim3(256), 0, h->st, (float*)stg, h->P[c], N); break;
    case MG_FDENSITY: hipLaunchKernelGGL(k_get_w, dim3(nb), dim3(256), 0, h->st, (float*)stg, h->V[c], N); break;
    case MG_FAGE: hipLaunchKernelGGL(k_gather_home, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->AGE, HOME, h->ctr, N, 1, 0u); break;
    case MG_FCOLOR: hipLaunchKernelGGL(k_gather_home, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->COLOR, HOME, h->ctr, N, 1, 0u); break;
    case MG_FNERVEIDX: hipLaunchKernelGGL(k_gather_home, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->NERVE, HOME, h->ctr, N, 1, 0u); break;
    case MG_FMASS_RADIUS: hipLaunchKernelGGL(k_gather_home, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->MRAD, HOME, h->ctr, N, 1, 0u); break;
    case MG_FPARTICLE_ID: hipLaunchKernelGGL(k_gather_home, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->PID, HOME, h->ctr, N, 0, MG_NONE); break;
    case MG_FGCELL: direct = h->GCELL[ic]; break;
    case MG_FGNDX: direct = h->GNDX[ic]; break;
    case MG_FBIN: hipLaunchKernelGGL(k_download_fbin, dim3(nb), dim3(256), 0, h->st, (uint32_t*)stg, h->FBIN, h->ctr, N); break;
    case MG_FCONC: direct = h->CONC[h->ccur]; break;
    case MG_FEPIGEN:
        hipLaunchKernelGGL(k_gather_epigen, dim3(nb, MG_NUM_GENES), dim3(256), 0, h->st, (uint32_t*)stg, h->EPIGEN, HOME, h->ctr, N);
        break;
    case MG_FELASTIDX:
        if (B > 0)
            hipLaunchKernelGGL(k_download_bonds, dim3(nb, B), dim3(256), 0, h->st, (uint32_t*)stg, h->BPARTNER,
                               h->BPAR, h->BINT, h->BCOLD, HOME, h->SLOT, h->ctr, N, B);
        break;
    case MG_FPARTICLEIDX:
        if (B > 0)
            hipLaunchKernelGGL(k_download_pidx, dim3(nb, B), dim3(256), 0, h->st, (uint32_t*)stg, h->PIDX, HOME,
                               h->SLOT, h->ctr, N, B);
        break;
    case MG_FBIN_COUNT: direct = h->CNT[h->gcur]; break;
    case MG_FBIN_OFFSET: direct = h->OFF[h->gcur]; break;
    case MG_FBIN_COUNT_ACTIVE_GENES: direct = h->GCNT; break;
    case MG_FBIN_OFFSET_ACTIVE_GENES: direct = h->GOFF; break;
    case MG_FDENSE_LIST_LENGTHS: direct = h->LENS; break;
    case MG_FDENSE_LISTS: direct = h->LISTS; break;
    case MG_FBIN_COUNT_CHANGES: direct = h->CCNT; break;
    case MG_FBIN_OFFSET_CHANGES: direct = h->COFF; break;
    case MG_FDENSE_LIST_LENGTHS_CHANGES: direct = h->CLEN; break;
    case MG_FDENSE_LISTS_CHANGES: direct = h->CL; break;
    default: set_err("mg_download: unsupported buffer id %d", buf_id); return MG_ERR_UNSUPPORTED;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst, direct ? direct : stg, bytes, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

int mg_get_counts(mg_handle* h, uint32_t* num_points, uint32_t* active_points)
{
    if (!h) { set_err("mg_get_counts: null handle"); return MG_ERR_ARG; }
    uint32_t c[MG_C_COUNT];
    HIPCHK(hipMemcpyAsync(c, h->ctr, sizeof(c), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (c[MG_C_ERR]) { set_err("device error flag set (a bounded wait expired)"); return MG_ERR_STATE; }
    if (num_points) *num_points = c[MG_C_NUM_POINTS];
    if (active_points) *active_points = c[MG_C_ACTIVE];
    return MG_OK;
}

int mg_sync(mg_handle* h)
{
    if (!h) { set_err("mg_sync: null handle"); return MG_ERR_ARG; }
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, h->ctr + MG_C_ERR, 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (err) { set_err("device error flag set (a bounded wait expired)"); return MG_ERR_STATE; }
    return MG_OK;
}

int mg_insert_particles(mg_handle* h, uint32_t* gcell, uint32_t* gndx, uint32_t* gcnt)
{
    CHK(ready(h));
    CHK(materialize_gene_grid(h));
    CHK(do_insert(h, 1));
    if (gcell) HIPCHK(hipMemcpyAsync(gcell, h->GCELL[h->icur], (size_t)h->N * 4, hipMemcpyDeviceToHost, h->st));
    if (gndx) HIPCHK(hipMemcpyAsync(gndx, h->GNDX[h->icur], (size_t)h->N * 4, hipMemcpyDeviceToHost, h->st));
    if (gcnt) HIPCHK(hipMemcpyAsync(gcnt, h->CNT[h->gcur], (size_t)h->C * 4, hipMemcpyDeviceToHost, h->st));
    if (gcell || gndx || gcnt) HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

int mg_prefix_sum_cells(mg_handle* h, int zero_offsets)
{
    CHK(ready(h));
    if (zero_offsets != 1) {
        set_err("mg_prefix_sum_cells: only zero_offsets=1 (exclusive offsets) is used by the step");
        return MG_ERR_UNSUPPORTED;
    }
    CHK(materialize_gene_grid(h));
    return do_scan(h, 1);
}

/* ---- the cell scan without allocation (mg_prefix_sum_cells_noalloc) ----
 * The look-back scan's scratch lives in one file-scope device array instead of a handle's
 * allocations. Layout (64-bit words): [0] the tile ticket, [1, 9) the scan counters (16
 * words of 32 bits, MG_C_*), [9, 9 + kScanNoAllocTiles) one status word per tile. Each call
 * clears the ticket, counters and the status words of its tiles before the scan (a fill
 * kernel, so a captured call holds no memset node), so every call starts from zero; calls
 * that overlap on one GPU share the array and race. Every GPU the library runs on holds its
 * own copy until the process ends; host code reaches it only through the address
 * hipGetSymbolAddress returns for the current device, looked up once per device. */
namespace {
constexpr unsigned kScanNoAllocTiles = 16384; /* up to 16384 x 4096 = 67,108,864 cells */
constexpr unsigned kScanNoAllocHead = 9;       /* ticket + 8 words of counters */
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};
} // namespace

static __device__ __align__(256) unsigned long long scratch[kScanNoAllocHead + kScanNoAllocTiles];

/* the device address of `scratch` on the current device (set-up on the first call there) */
static int scratch_on_device(unsigned long long** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_err("mg_prefix_sum_cells_noalloc: device %d outside the supported range", dev);
        return MG_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_err("mg_prefix_sum_cells_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return MG_ERR_HIP;
        }
        g_scratch_ptr[dev] = (unsigned long long*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return MG_OK;
}

/* report[0] = the total (the last cell's exclusive offset plus its count), report[1] = the
 * bounded-wait flag */
__global__ void k_scan_report(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, int n,
                              const uint32_t* __restrict__ ctr, uint32_t* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[0] = off[n - 1] + cnt[n - 1];
        out[1] = ctr[MG_C_ERR];
    }
}

int mg_prefix_sum_cells_noalloc(const uint32_t* cnt, uint32_t* off, int n, uint32_t* report, void* stream)
{
    if (!cnt || !off || n <= 0) {
        set_err("mg_prefix_sum_cells_noalloc: null buffer or n <= 0");
        return MG_ERR_ARG;
    }
    const unsigned tiles = (unsigned)(((size_t)n + SCAN_TILE - 1) / SCAN_TILE);
    if (tiles > kScanNoAllocTiles) {
        set_err("mg_prefix_sum_cells_noalloc: n = %d exceeds the maximum of %u cells", n,
                kScanNoAllocTiles * (unsigned)SCAN_TILE);
        return MG_ERR_SIZE;
    }
    unsigned long long* base = nullptr;
    CHK(scratch_on_device(&base));
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* ticket = base;
    uint32_t* ctr = (uint32_t*)(base + 1);
    unsigned long long* status = base + kScanNoAllocHead;
    const size_t words = 2 * ((size_t)kScanNoAllocHead + tiles);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((words + 255) / 256, 1024)), dim3(256), 0, st,
                       (uint32_t*)base, 0u, words);
    hipLaunchKernelGGL(k_scan_lookback, dim3(tiles), dim3(256), 0, st, cnt, off, n, status, ticket, ctr, -1);
    if (report) hipLaunchKernelGGL(k_scan_report, dim3(1), dim3(64), 0, st, cnt, off, n, ctr, report);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int mg_counting_sort_full(mg_handle* h)
{
    CHK(ready(h));
    CHK(materialize_gene_grid(h));
    return do_sort(h, 1);
}

int mg_compute_pressure(mg_handle* h) { CHK(ready(h)); return do_pressure(h); }
int mg_compute_force(mg_handle* h) { CHK(ready(h)); return do_force(h); }
int mg_advance(mg_handle* h, float time, float dt, float ss)
{
    (void)time; (void)ss;
    CHK(ready(h));
    return do_advance(h, dt);
}
int mg_compute_diffusion(mg_handle* h) { CHK(ready(h)); return do_diffusion(h); }
int mg_compute_genes(mg_handle* h) { CHK(ready(h)); return do_genes(h); }
int mg_zero_vel(mg_handle* h)
{
    CHK(ready(h));
    hipLaunchKernelGGL(k_zero_vel, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h));
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int mg_initialize_bonds(mg_handle* h)
{
    CHK(ready(h));
    if (h->B == 0) return MG_OK;
    if (!h->force_zero) {
        set_err("mg_initialize_bonds: the handle must be in the sorted state (after a physical sort or a step)");
        return MG_ERR_STATE;
    }
    CHK(materialize_gene_grid(h));
    /* up to 32 slots: 32 lanes per listed particle, one candidate each, angle terms kept from
     * slot to slot; else eight lanes share its candidate scan (one lane when unstaged) */
    const int staged = h->B <= BOND_STAGE_MAX_B;
    const size_t lds = staged ? (size_t)h->B * (BOND_STAGE_W / 8) * (sizeof(float4) + sizeof(uint32_t)) : 0;
    if (h->B <= BOND_TAB_B)
        hipLaunchKernelGGL(k_initialize_bonds_tab, dim3(nblk((size_t)h->N * BOND_TAB_S, BOND_STAGE_W)),
                           dim3(BOND_STAGE_W),
                           (size_t)(BOND_STAGE_W / BOND_TAB_S) *
                               ((h->B + BOND_TAB_S) * (sizeof(float4) + sizeof(uint32_t)) + 2 * BOND_TAB_S * sizeof(uint32_t)),
                           h->st, make_ds(h), h->p, h->g, h->gen, h->PID, h->BCOLD);
    else if (staged)
        hipLaunchKernelGGL(k_initialize_bonds<8>, dim3(nblk((size_t)h->N * 8, BOND_STAGE_W)), dim3(BOND_STAGE_W), lds,
                           h->st, make_ds(h), h->p, h->g, h->gen, h->PID, h->BCOLD, 1);
    else
        hipLaunchKernelGGL(k_initialize_bonds<1>, dim3(nblk(h->N, BOND_STAGE_W)), dim3(BOND_STAGE_W), 0, h->st,
                           make_ds(h), h->p, h->g, h->gen, h->PID, h->BCOLD, 0);
    HIPCHK(hipGetLastError());
    return rebuild_inc(h); /* new partners: incoming-bond CSR and live-slot masks */
}

int mg_compute_bond_changes(mg_handle* h, uint32_t steps_per_inner_loop)
{
    CHK(ready(h));
    if (steps_per_inner_loop == 0) {
        set_err("mg_compute_bond_changes: steps_per_inner_loop must be > 0");
        return MG_ERR_ARG;
    }
    CHK(ensure_changes(h));
    CHK(materialize_gene_grid(h));
    CHK(fill_u32(h, h->CCNT, 0, (size_t)MG_NUM_CHANGES * h->C));
    CHK(fill_u32(h, h->COFF, 0, (size_t)MG_NUM_CHANGES * h->C));
    CHK(fill_u32(h, h->HS, MG_NONE, h->N));
    if (h->B == 0) return MG_OK;
    TanhTab tt;
    std::memcpy(tt.t, h->gen.tanh_param, sizeof(tt.t));
    hipLaunchKernelGGL(k_bond_changes, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h), tt,
                       steps_per_inner_loop, h->CCNT, h->HS, h->BCOLD);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int mg_prefix_sum_changes(mg_handle* h, int zero_offsets)
{
    (void)zero_offsets;
    CHK(ready(h));
    CHK(ensure_changes(h));
    CHK(seg_scan(h, h->CCNT, h->COFF, MG_NUM_CHANGES, h->CLEN));
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int mg_counting_sort_changes(mg_handle* h)
{
    CHK(ready(h));
    CHK(ensure_changes(h));
    CHK(fill_u32(h, h->CL, MG_NONE, (size_t)MG_NUM_CHANGES * 2 * h->N));
    if (h->B == 0) return MG_OK;
    hipLaunchKernelGGL(k_sort_changes, dim3(nblk(h->C)), dim3(256), 0, h->st, make_ds(h), h->CCNT, h->COFF, h->HS,
                       h->CL, h->BCOLD);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

int mg_compute_particle_changes(mg_handle* h, uint64_t seed)
{
    CHK(ready(h));
    CHK(ensure_changes(h));
    uint32_t len[MG_NUM_CHANGES];
    HIPCHK(hipMemcpyAsync(len, h->CLEN, sizeof(len), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int l = 1; l < MG_NUM_CHANGES; l++)
        if (len[l] != 0) {
            set_err("mg_compute_particle_changes: change list %d holds %u entries; only heal (list 0) is built", l,
                    len[l]);
            return MG_ERR_UNSUPPORTED;
        }
    if (h->B == 0 || len[0] == 0) return MG_OK;
    const float4 elastin = make_float4(h->gen.param[0][8], h->gen.param[0][9], h->gen.param[0][10], h->gen.param[0][11]);
    const uint32_t n = std::min<uint32_t>(len[0], (uint32_t)h->N);
    /* 32 lanes per listed particle, one candidate each (one lane when the bonds are not staged) */
    const int staged = h->B <= BOND_STAGE_MAX_B;
    const size_t lds = (size_t)(BOND_STAGE_W / BOND_TAB_S) *
                       ((h->B + BOND_TAB_S) * (sizeof(float4) + sizeof(uint32_t)) + 2 * BOND_TAB_S * sizeof(uint32_t));
    if (staged)
        hipLaunchKernelGGL(k_heal_flat, dim3(nblk((size_t)n * BOND_TAB_S, BOND_STAGE_W)), dim3(BOND_STAGE_W), lds,
                           h->st, make_ds(h), h->p, h->g, elastin, h->CL, h->CLEN, h->PID, h->BCOLD,
                           (unsigned long long)seed);
    else
        hipLaunchKernelGGL(k_heal<1>, dim3(nblk(n, BOND_STAGE_W)), dim3(BOND_STAGE_W), 0, h->st, make_ds(h), h->p,
                           h->g, elastin, h->CL, h->CLEN, h->PID, h->BCOLD, (unsigned long long)seed, 0);
    HIPCHK(hipGetLastError());
    return rebuild_inc(h); /* new partners: incoming-bond CSR and live-slot masks */
}

int mg_physical_sort(mg_handle* h)
{
    CHK(ready(h));
    /* the gene counts, offsets and dense lists are observable only through downloads and the
     * calls that read them, which rebuild them from the sorted state (materialize_gene_grid) */
    CHK(do_insert(h, 0));
    CHK(do_scan(h, 0));
    CHK(do_sort(h, 0));
    h->gene_grid_stale = true;
    return MG_OK;
}

int mg_inner_physical_loop(mg_handle* h, float time, float dt)
{
    (void)time;
    CHK(ready(h));
    ensure_hit_masks(h);
    uint64_t* hm = hit_masks(h); /* pressure and force of the same positions: share the masks */
    CHK(do_pressure(h, hm));
    CHK(do_force(h, hm));
    return do_advance(h, dt);
}

int mg_gene_action(mg_handle* h)
{
    CHK(ready(h));
    CHK(do_diffusion(h));
    return do_genes(h);
}

int mg_step(mg_handle* h, int nsteps, float dt, unsigned flags)
{
    CHK(ready(h));
    if (nsteps < 0) { set_err("mg_step: nsteps < 0"); return MG_ERR_ARG; }
    CHK(maybe_rehome(h));
    h->steps_since_rehome += nsteps;
    ensure_hit_masks(h);
    int s = 0;
    if ((flags & MG_STEP_GRAPH) && h->force_zero && nsteps >= 2 && graph_ok(h, flags)) {
        const int pairs = nsteps / 2;
        CHK(graph_steps(h, pairs, dt, flags & (MG_STEP_GENES | MG_STEP_SPECIAL)));
        h->time += 2.0 * pairs * dt;
        s = 2 * pairs;
    }
    for (; s < nsteps; s++) {
        if (h->force_zero) {
            /* the usual case: the previous operation was a sort */
            CHK(do_fused_step(h, dt, flags));
            h->time += dt;
            continue;
        }
        CHK(materialize_gene_grid(h));
        uint64_t* hm = hit_masks(h);
        CHK(do_pressure(h, hm));
        CHK(do_force(h, hm));
        CHK(do_advance(h, dt, flags));
        h->time += dt;
        if (flags & MG_STEP_GENES) {
            CHK(do_diffusion(h));
            CHK(do_genes(h));
        }
        CHK(do_insert(h, 0));
        CHK(do_scan(h, 0));
        CHK(do_sort(h, 0));
        h->gene_grid_stale = true;
    }
    return MG_OK;
}

int mg_step_timed(mg_handle* h, int nsteps, float dt, unsigned flags, double* ms_out, int* calls_out)
{
    CHK(ready(h));
    h->timing = true;
    h->marks.clear();
    h->pool_used = 0;
    int r = mg_step(h, nsteps, dt, flags & ~MG_STEP_GRAPH);
    h->timing = false;
    if (r != MG_OK) return r;
    HIPCHK(hipStreamSynchronize(h->st));
    for (const auto& m : h->marks) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, m.a, m.b));
        if (ms_out) ms_out[m.ph] += ms;
        if (calls_out) calls_out[m.ph] += 1;
    }
    h->marks.clear();
    return MG_OK;
}

} // extern "C"

#include "mg_batch.hip"
#include "mg_learn.hip"
