/*
 * mg_learn.hip — host side of the learned coarse step (included by mg_capi.hip; kernels in
 * mg_mlp.hip). No counterpart in the reference.
 *
 * mg_net owns fp32 master weights with Adam state, their bf16 copies for the training GEMMs,
 * the packed inference parameters (MlpDev) and an uploaded training set. Parameter vector
 * layout (fp32, row-major [in][out]):
 *   W1 [64][128] | b1 [128] | W2 [128][128] | b2 [128] | W3 [128][16] | b3 [16]
 * Rows of W1 past input 50 and columns of W3 past output 3 are padding: they start at zero
 * and receive zero gradient (their inputs / output gradients are zero).
 */

namespace {
const size_t kOffW1 = 0, kOffB1 = kOffW1 + MLP_IN * MLP_H, kOffW2 = kOffB1 + MLP_H;
const size_t kOffB2 = kOffW2 + MLP_H * MLP_H, kOffW3 = kOffB2 + MLP_H, kOffB3 = kOffW3 + MLP_H * MLP_OUT;
const size_t kNParams = kOffB3 + MLP_OUT;
} // namespace

struct mg_net {
    hipStream_t st = nullptr;
    MlpDev* dev = nullptr;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    __bf16* wb = nullptr; /* bf16 copy of the whole parameter vector (only the W parts are read) */
    int adam_t = 0;
    uint64_t version = 1; /* bumped on every parameter / normalisation change */
    float mu[MLP_IN], inv_sigma[MLP_IN];
    float out_scale[MLP_OUT] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
    float in_limit = 1e30f, out_limit = 1e30f;
    /* training set */
    float *X = nullptr, *T = nullptr, *mask = nullptr;
    size_t rows = 0;
    std::vector<uint64_t> used_prefix; /* rows with a nonzero mask before each row (rows + 1) */
    float* loss_dev = nullptr;         /* per-step losses of mg_net_train, downloaded at its end */
    size_t loss_cap = 0;
    /* work buffers for minibatches of up to cap rows */
    size_t cap = 0;
    __bf16 *Xn = nullptr, *A1b = nullptr, *A2b = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *dY = nullptr, *dA2 = nullptr, *dZ2 = nullptr;
    float *dA1 = nullptr, *dZ1 = nullptr, *part = nullptr, *lpart = nullptr;
    size_t part_cap = 0;
    std::vector<void*> allocs;
};

namespace {

template <typename T> int nalloc(mg_net* n, T** p, size_t count)
{
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)));
    n->allocs.push_back(q);
    *p = (T*)q;
    return MG_OK;
}

unsigned gblk(size_t n) { return (unsigned)((n + 255) / 256); }

/* bf16 copies and packed inference parameters from the fp32 master weights */
int net_sync(mg_net* n)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(gblk(kNParams)), dim3(256), 0, n->st, n->params, n->wb, kNParams);
    /* W1 in input order; W2 and W3 in the order the previous layer's results sit in registers */
    hipLaunchKernelGGL(k_pack_frag, dim3(gblk(MLP_IN * MLP_H)), dim3(256), 0, n->st, n->params + kOffW1, MLP_IN,
                       MLP_H, n->dev->w1f, 0);
    hipLaunchKernelGGL(k_pack_frag, dim3(gblk(MLP_H * MLP_H)), dim3(256), 0, n->st, n->params + kOffW2, MLP_H,
                       MLP_H, n->dev->w2f, 1);
    hipLaunchKernelGGL(k_pack_frag, dim3(gblk(MLP_H * MLP_OUT)), dim3(256), 0, n->st, n->params + kOffW3, MLP_H,
                       MLP_OUT, n->dev->w3f, 1);
    HIPCHK(hipMemcpyAsync(n->dev->b1, n->params + kOffB1, MLP_H * 4, hipMemcpyDeviceToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->dev->b2, n->params + kOffB2, MLP_H * 4, hipMemcpyDeviceToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->dev->b3, n->params + kOffB3, MLP_OUT * 4, hipMemcpyDeviceToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->dev->mu, n->mu, sizeof(n->mu), hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->dev->inv_sigma, n->inv_sigma, sizeof(n->inv_sigma), hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->dev->out_scale, n->out_scale, sizeof(n->out_scale), hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(&n->dev->in_limit, &n->in_limit, 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(&n->dev->out_limit, &n->out_limit, 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipStreamSynchronize(n->st)); /* the host-side sources above must stay valid */
    n->version++;
    return MG_OK;
}

int net_reserve(mg_net* n, size_t S)
{
    if (S <= n->cap) return MG_OK;
    CHK(nalloc(n, &n->Xn, S * MLP_IN));
    CHK(nalloc(n, &n->A1b, S * MLP_H));
    CHK(nalloc(n, &n->A2b, S * MLP_H));
    CHK(nalloc(n, &n->dYb, S * MLP_OUT));
    CHK(nalloc(n, &n->dZ2b, S * MLP_H));
    CHK(nalloc(n, &n->dZ1b, S * MLP_H));
    CHK(nalloc(n, &n->Z1, S * MLP_H));
    CHK(nalloc(n, &n->Z2, S * MLP_H));
    CHK(nalloc(n, &n->Y, S * MLP_OUT));
    CHK(nalloc(n, &n->dY, S * MLP_OUT));
    CHK(nalloc(n, &n->dA2, S * MLP_H));
    CHK(nalloc(n, &n->dZ2, S * MLP_H));
    CHK(nalloc(n, &n->dA1, S * MLP_H));
    CHK(nalloc(n, &n->dZ1, S * MLP_H));
    CHK(nalloc(n, &n->lpart, gblk(S)));
    n->cap = S;
    return MG_OK;
}

/* C[M x N] = A(m,k) B(k,n), K split into `splits` chunks reduced in order */
int gemm(mg_net* n, int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn,
         float* C, int splits)
{
    splits = std::max(1, std::min(splits, (K + 31) / 32));
    const int kper = ((K + splits - 1) / splits + 31) / 32 * 32;
    splits = (K + kper - 1) / kper;
    const size_t mn = (size_t)M * N;
    if (splits > 1 && mn * splits > n->part_cap) {
        CHK(nalloc(n, &n->part, mn * splits));
        n->part_cap = mn * splits;
    }
    float* out = splits > 1 ? n->part : C;
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 63) / 64, (M + 63) / 64, splits), dim3(256), 0, n->st, M, N, K, A, sam,
                       sak, B, sbk, sbn, out, kper);
    if (splits > 1) hipLaunchKernelGGL(k_gemm_reduce, dim3(gblk(mn)), dim3(256), 0, n->st, n->part, splits, mn, C);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

/* forward (and with grads, backward) over rows [r0, r0 + S) of the training set */
/* one thread: the blocks' loss partials summed in order in double, as a float (the host's
 * sum of the same partials gives the same bits) */
__global__ void k_sum_partials(const float* __restrict__ part, unsigned count, float* __restrict__ out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (unsigned k = 0; k < count; k++) s += part[k];
    *out = (float)s;
}

/* forward (and backward) over rows [r0, r0 + S); the mean loss goes to the host (loss_host)
 * or stays on the device (loss_dev), either may be null */
int net_pass(mg_net* n, size_t r0, size_t S, bool backward, float* loss_host, float* loss_dev = nullptr)
{
    CHK(net_reserve(n, S));
    const float* X = n->X + r0 * MLP_IN;
    float *mu_d = n->dev->mu, *is_d = n->dev->inv_sigma;
    const __bf16 *W1 = n->wb + kOffW1, *W2 = n->wb + kOffW2, *W3 = n->wb + kOffW3;
    hipLaunchKernelGGL(k_normalize_rows, dim3(gblk(S * MLP_IN)), dim3(256), 0, n->st, X, mu_d, is_d, n->Xn, S);
    CHK(gemm(n, (int)S, MLP_H, MLP_IN, n->Xn, MLP_IN, 1, W1, MLP_H, 1, n->Z1, 1));
    hipLaunchKernelGGL(k_bias_silu, dim3(gblk(S * MLP_H)), dim3(256), 0, n->st, n->Z1, n->params + kOffB1, MLP_H, S,
                       n->A1b, 1);
    CHK(gemm(n, (int)S, MLP_H, MLP_H, n->A1b, MLP_H, 1, W2, MLP_H, 1, n->Z2, 1));
    hipLaunchKernelGGL(k_bias_silu, dim3(gblk(S * MLP_H)), dim3(256), 0, n->st, n->Z2, n->params + kOffB2, MLP_H, S,
                       n->A2b, 1);
    CHK(gemm(n, (int)S, MLP_OUT, MLP_H, n->A2b, MLP_H, 1, W3, MLP_OUT, 1, n->Y, 1));
    hipLaunchKernelGGL(k_bias_silu, dim3(gblk(S * MLP_OUT)), dim3(256), 0, n->st, n->Y, n->params + kOffB3, MLP_OUT,
                       S, (__bf16*)nullptr, 0);
    if (!backward) return MG_OK;
    /* count of masked rows in this window, for the mean (from the prefix taken at load) */
    const double cnt = (double)(n->used_prefix[r0 + S] - n->used_prefix[r0]);
    const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    const unsigned lb = gblk(S);
    hipLaunchKernelGGL(k_mse, dim3(lb), dim3(256), 0, n->st, n->Y, n->T + r0 * MLP_OUT_REAL, n->mask + r0, S, inv, n->dY,
                       n->dYb, n->lpart);
    float* g = n->grads;
    const int ks = (int)std::min<size_t>(64, std::max<size_t>(1, S / 1024));
    /* layer 3 */
    CHK(gemm(n, MLP_H, MLP_OUT, (int)S, n->A2b, 1, MLP_H, n->dYb, MLP_OUT, 1, g + kOffW3, ks));
    hipLaunchKernelGGL(k_colsum, dim3(MLP_OUT), dim3(256), 0, n->st, n->dY, S, MLP_OUT, g + kOffB3);
    CHK(gemm(n, (int)S, MLP_H, MLP_OUT, n->dYb, MLP_OUT, 1, W3, 1, MLP_OUT, n->dA2, 1));
    hipLaunchKernelGGL(k_silu_back, dim3(gblk(S * MLP_H)), dim3(256), 0, n->st, n->dA2, n->Z2, S * MLP_H, n->dZ2, n->dZ2b);
    /* layer 2 */
    CHK(gemm(n, MLP_H, MLP_H, (int)S, n->A1b, 1, MLP_H, n->dZ2b, MLP_H, 1, g + kOffW2, ks));
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, n->st, n->dZ2, S, MLP_H, g + kOffB2);
    CHK(gemm(n, (int)S, MLP_H, MLP_H, n->dZ2b, MLP_H, 1, W2, 1, MLP_H, n->dA1, 1));
    hipLaunchKernelGGL(k_silu_back, dim3(gblk(S * MLP_H)), dim3(256), 0, n->st, n->dA1, n->Z1, S * MLP_H, n->dZ1, n->dZ1b);
    /* layer 1 */
    CHK(gemm(n, MLP_IN, MLP_H, (int)S, n->Xn, 1, MLP_IN, n->dZ1b, MLP_H, 1, g + kOffW1, ks));
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, n->st, n->dZ1, S, MLP_H, g + kOffB1);
    HIPCHK(hipGetLastError());
    if (loss_dev) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, n->st, n->lpart, lb, loss_dev);
        HIPCHK(hipGetLastError());
    }
    if (loss_host) {
        std::vector<float> lp(lb);
        HIPCHK(hipMemcpyAsync(lp.data(), n->lpart, lb * 4, hipMemcpyDeviceToHost, n->st));
        HIPCHK(hipStreamSynchronize(n->st));
        double s = 0;
        for (float x : lp) s += x;
        *loss_host = (float)s;
    }
    return MG_OK;
}

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

} // namespace

uint64_t net_version(const mg_net* net) { return net->version; }

/* the correction buffer and fallback counter of learned steps; allocated outside any capture */
int ensure_learn_buffers(mg_handle* h)
{
    if (h->DV) return MG_OK;
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, (size_t)h->N * (MLP_IN / 8) * sizeof(uint4)));
    h->allocs.push_back(q);
    h->XB = (uint4*)q;
    HIPCHK(hipMalloc(&q, (size_t)h->N * sizeof(uint32_t)));
    h->allocs.push_back(q);
    h->MF = (uint32_t*)q;
    HIPCHK(hipMalloc(&q, (size_t)h->N * 2 * sizeof(float4)));
    h->allocs.push_back(q);
    h->DV = (float4*)q;
    HIPCHK(hipMalloc(&q, 256));
    h->allocs.push_back(q);
    h->fallbacks = (uint32_t*)q;
    HIPCHK(hipMemsetAsync(h->fallbacks, 0, 4, h->st));
    HIPCHK(hipMemsetAsync(h->DV, 0, (size_t)h->N * 2 * sizeof(float4), h->st));
    h->epoch++;
    return MG_OK;
}

/* the MLP pass of a learned step of length dt (called from do_fused_step after the force
 * pass has written FFORCE) */
int launch_mlp_infer(mg_handle* h, const mg_net* net, float dt)
{
    if (!h->DV) {
        set_err("learned step: correction buffer missing");
        return MG_ERR_STATE;
    }
    hipLaunchKernelGGL(k_mlp_gather, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h), h->p, net->dev, h->XB,
                       h->MF);
    HIPCHK(hipGetLastError());
    /* groups of 32 particles, 4 per block at a time; at most 512 blocks, so the weights are
     * staged a bounded number of times */
    const unsigned groups = (unsigned)((h->N + 31) / 32);
    const unsigned blocks = std::max(1u, std::min((groups + MLP_WAVES - 1) / MLP_WAVES, 512u));
    hipLaunchKernelGGL(k_mlp_apply, dim3(blocks), dim3(256), sizeof(MlpLds), h->st, net->dev, h->XB, h->MF, h->DV,
                       h->fallbacks, dt, (uint32_t)h->N);
    HIPCHK(hipGetLastError());
    return MG_OK;
}

extern "C" {

size_t mg_net_nparams(void) { return kNParams; }

int mg_net_create(mg_net** out, uint64_t seed)
{
    if (!out) return MG_ERR_ARG;
    mg_net* n = new mg_net;
    auto fail = [&](int r) {
        for (void* q : n->allocs) (void)hipFree(q);
        if (n->st) (void)hipStreamDestroy(n->st);
        delete n;
        return r;
    };
    if (hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) {
        set_err("mg_net_create: hipStreamCreate failed");
        return fail(MG_ERR_HIP);
    }
    int r = nalloc(n, &n->dev, 1);
    if (r == MG_OK) r = nalloc(n, &n->params, kNParams);
    if (r == MG_OK) r = nalloc(n, &n->grads, kNParams);
    if (r == MG_OK) r = nalloc(n, &n->m, kNParams);
    if (r == MG_OK) r = nalloc(n, &n->v, kNParams);
    if (r == MG_OK) r = nalloc(n, &n->wb, kNParams);
    if (r != MG_OK) return fail(r);
    if (hipFuncSetAttribute((const void*)k_mlp_apply, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(MlpLds)) != hipSuccess) {
        set_err("mg_net_create: the inference kernel's LDS (%zu bytes) is not available", sizeof(MlpLds));
        return fail(MG_ERR_HIP);
    }
    /* Xavier-uniform (Glorot & Bengio, 2010) over the real fan-in / fan-out; padding zero */
    std::vector<float> p(kNParams, 0.f);
    uint64_t s = seed;
    auto unif = [&](float lim) { return (float)(((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * lim; };
    const float l1 = std::sqrt(6.f / (MLP_IN_REAL + MLP_H)), l2 = std::sqrt(6.f / (2 * MLP_H));
    const float l3 = std::sqrt(6.f / (MLP_H + MLP_OUT_REAL));
    for (int i = 0; i < MLP_IN_REAL; i++)
        for (int j = 0; j < MLP_H; j++) p[kOffW1 + i * MLP_H + j] = unif(l1);
    for (int i = 0; i < MLP_H; i++)
        for (int j = 0; j < MLP_H; j++) p[kOffW2 + i * MLP_H + j] = unif(l2);
    for (int i = 0; i < MLP_H; i++)
        for (int j = 0; j < MLP_OUT_REAL; j++) p[kOffW3 + i * MLP_OUT + j] = unif(l3);
    for (int k = 0; k < MLP_IN; k++) n->mu[k] = 0.f, n->inv_sigma[k] = 1.f;
    if (hipMemcpy(n->params, p.data(), kNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, kNParams * 4) != hipSuccess || hipMemset(n->v, 0, kNParams * 4) != hipSuccess ||
        hipMemset(n->grads, 0, kNParams * 4) != hipSuccess) {
        set_err("mg_net_create: initial upload failed");
        return fail(MG_ERR_HIP);
    }
    r = net_sync(n);
    if (r != MG_OK) return fail(r);
    *out = n;
    return MG_OK;
}

int mg_net_destroy(mg_net* n)
{
    if (!n) return MG_ERR_ARG;
    (void)hipStreamSynchronize(n->st);
    for (void* q : n->allocs) (void)hipFree(q);
    (void)hipStreamDestroy(n->st);
    delete n;
    return MG_OK;
}

int mg_net_get_params(mg_net* n, float* out, size_t count)
{
    if (!n || !out || count != kNParams) { set_err("mg_net_get_params: need %zu floats", kNParams); return MG_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(out, n->params, kNParams * 4, hipMemcpyDeviceToHost, n->st));
    HIPCHK(hipStreamSynchronize(n->st));
    return MG_OK;
}

int mg_net_set_params(mg_net* n, const float* in, size_t count)
{
    if (!n || !in || count != kNParams) { set_err("mg_net_set_params: need %zu floats", kNParams); return MG_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(n->params, in, kNParams * 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipStreamSynchronize(n->st));
    return net_sync(n);
}

int mg_net_set_norm(mg_net* n, const float* mu, const float* inv_sigma, const float* out_scale,
                    float in_limit, float out_limit)
{
    if (!n || !mu || !inv_sigma || !out_scale) return MG_ERR_ARG;
    std::memcpy(n->mu, mu, sizeof(n->mu));
    std::memcpy(n->inv_sigma, inv_sigma, sizeof(n->inv_sigma));
    for (int c = 0; c < MLP_OUT; c++) n->out_scale[c] = c < MLP_OUT_REAL ? out_scale[c] : 0.0f;
    n->in_limit = in_limit;
    n->out_limit = out_limit;
    return net_sync(n);
}

int mg_net_load_data(mg_net* n, const float* X, const float* T, const float* mask, size_t rows)
{
    if (!n || !X || !T || !mask || rows == 0) return MG_ERR_ARG;
    CHK(nalloc(n, &n->X, rows * MLP_IN));
    CHK(nalloc(n, &n->T, rows * MLP_OUT_REAL));
    CHK(nalloc(n, &n->mask, rows));
    HIPCHK(hipMemcpyAsync(n->X, X, rows * MLP_IN * 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->T, T, rows * MLP_OUT_REAL * 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipMemcpyAsync(n->mask, mask, rows * 4, hipMemcpyHostToDevice, n->st));
    HIPCHK(hipStreamSynchronize(n->st));
    n->used_prefix.assign(rows + 1, 0);
    for (size_t r = 0; r < rows; r++) n->used_prefix[r + 1] = n->used_prefix[r] + (mask[r] != 0.f ? 1u : 0u);
    n->rows = rows;
    return MG_OK;
}

/* one forward + backward over data rows [row0, row0 + rows); gradients (kNParams floats) and
 * the mean loss to the host; no update */
int mg_net_grad(mg_net* n, size_t row0, size_t rows, float* grads_out, float* loss_out)
{
    if (!n || rows == 0 || row0 + rows > n->rows) { set_err("mg_net_grad: rows outside the data"); return MG_ERR_ARG; }
    CHK(net_pass(n, row0, rows, true, loss_out));
    if (grads_out) {
        HIPCHK(hipMemcpyAsync(grads_out, n->grads, kNParams * 4, hipMemcpyDeviceToHost, n->st));
    }
    HIPCHK(hipStreamSynchronize(n->st));
    return MG_OK;
}

/* network outputs y (before out_scale; MLP_OUT_REAL per row) for data rows [row0, row0 + rows) */
int mg_net_forward(mg_net* n, size_t row0, size_t rows, float* y_out)
{
    if (!n || !y_out || rows == 0 || row0 + rows > n->rows) { set_err("mg_net_forward: rows outside the data"); return MG_ERR_ARG; }
    CHK(net_pass(n, row0, rows, false, nullptr));
    std::vector<float> y(rows * MLP_OUT);
    HIPCHK(hipMemcpyAsync(y.data(), n->Y, rows * MLP_OUT * 4, hipMemcpyDeviceToHost, n->st));
    HIPCHK(hipStreamSynchronize(n->st));
    for (size_t r = 0; r < rows; r++)
        for (int c = 0; c < MLP_OUT_REAL; c++) y_out[r * MLP_OUT_REAL + c] = y[r * MLP_OUT + c];
    return MG_OK;
}

/* Adam over minibatches: `steps` updates of `batch` consecutive rows (windows walk the data
 * cyclically from a seeded start); loss_out (steps entries, may be null) gets each
 * minibatch's mean loss */
int mg_net_train(mg_net* n, int steps, size_t batch, float lr, uint64_t seed, float* loss_out)
{
    if (!n || steps < 0 || batch == 0 || batch > n->rows) { set_err("mg_net_train: bad arguments"); return MG_ERR_ARG; }
    uint64_t s = seed;
    const size_t windows = n->rows / batch;
    size_t w = (size_t)(splitmix(s) % std::max<size_t>(windows, 1));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    if (loss_out && (size_t)steps > n->loss_cap) {
        CHK(nalloc(n, &n->loss_dev, (size_t)steps));
        n->loss_cap = (size_t)steps;
    }
    /* no host synchronisation inside the loop: the losses stay on the device until the end */
    for (int k = 0; k < steps; k++) {
        const size_t r0 = (w % std::max<size_t>(windows, 1)) * batch;
        w++;
        CHK(net_pass(n, r0, batch, true, nullptr, loss_out ? n->loss_dev + k : nullptr));
        n->adam_t++;
        const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(gblk(kNParams)), dim3(256), 0, n->st, n->params, n->grads, n->m, n->v,
                           kNParams, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_to_bf16, dim3(gblk(kNParams)), dim3(256), 0, n->st, n->params, n->wb, kNParams);
        HIPCHK(hipGetLastError());
    }
    if (loss_out && steps > 0)
        HIPCHK(hipMemcpyAsync(loss_out, n->loss_dev, (size_t)steps * 4, hipMemcpyDeviceToHost, n->st));
    return net_sync(n); /* synchronises the stream */
}

/* Raw MLP inputs of every slot of h's current sorted state (N x 64 floats; inactive rows
 * zero), from the pressure and force passes a learned step would run first. The force pass
 * changes no bond record (dry) and FFORCE keeps counting as zero, so the state is left as a
 * step would find it. */
int mg_net_features(mg_handle* h, float* X_out)
{
    CHK(ready(h));
    if (!X_out) return MG_ERR_ARG;
    if (!h->force_zero) {
        set_err("mg_net_features: the handle must be in the sorted state (after a physical sort or a step)");
        return MG_ERR_STATE;
    }
    uint64_t* hm = hit_masks(h);
    CHK(do_pressure(h, hm));
    {
        DS d = make_ds(h);
        d.HMR = hm;
        d.ETR = bond_records(h, hm);
        d.dry = 1;
        launch_force(h, nblk(h->N), d, 0);
        HIPCHK(hipGetLastError());
    }
    float* d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)h->N * MLP_IN * 4));
    hipLaunchKernelGGL(k_mlp_features, dim3(nblk(h->N)), dim3(256), 0, h->st, make_ds(h), h->p, d);
    hipError_t e = hipMemcpyAsync(X_out, d, (size_t)h->N * MLP_IN * 4, hipMemcpyDeviceToHost, h->st);
    hipError_t e2 = hipStreamSynchronize(h->st);
    (void)hipFree(d);
    HIPCHK(e);
    HIPCHK(e2);
    return MG_OK;
}

/* The corrections a learned step of length dt would apply to h's current sorted state:
 * DV_out (N x 8 floats: c_d, used, c_e, 0), after the same pressure and (dry) force passes
 * as mg_net_features. */
int mg_net_infer(mg_handle* h, mg_net* n, float dt, float* DV_out)
{
    CHK(ready(h));
    if (!n || !DV_out) return MG_ERR_ARG;
    if (!h->force_zero) {
        set_err("mg_net_infer: the handle must be in the sorted state (after a physical sort or a step)");
        return MG_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(n->st));
    CHK(ensure_learn_buffers(h));
    uint64_t* hm = hit_masks(h);
    CHK(do_pressure(h, hm));
    {
        DS d = make_ds(h);
        d.HMR = hm;
        d.ETR = bond_records(h, hm);
        d.dry = 1;
        launch_force(h, nblk(h->N), d, 0);
        HIPCHK(hipGetLastError());
    }
    CHK(launch_mlp_infer(h, n, dt));
    HIPCHK(hipMemcpyAsync(DV_out, h->DV, (size_t)h->N * 32, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return MG_OK;
}

/* nsteps learned coarse steps of dt each: the fused step with the MLP correction added to
 * every active particle's velocity in the advance (particles whose inputs or outputs leave
 * the trained range take the plain step). flags as mg_step. */
int mg_step_learned(mg_handle* h, mg_net* n, int nsteps, float dt, unsigned flags)
{
    CHK(ready(h));
    if (!n || nsteps < 0) { set_err("mg_step_learned: bad arguments"); return MG_ERR_ARG; }
    if (!h->force_zero) {
        set_err("mg_step_learned: the handle must be in the sorted state (after a physical sort or a step)");
        return MG_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(n->st));
    CHK(ensure_learn_buffers(h));
    CHK(maybe_rehome(h));
    h->steps_since_rehome += nsteps;
    ensure_hit_masks(h);
    int s = 0;
    if ((flags & MG_STEP_GRAPH) && nsteps >= 2 && graph_ok(h, flags)) {
        const int pairs = nsteps / 2;
        CHK(graph_steps(h, pairs, dt, flags & (MG_STEP_GENES | MG_STEP_SPECIAL), n));
        h->time += 2.0 * pairs * dt;
        s = 2 * pairs;
    }
    for (; s < nsteps; s++) {
        CHK(do_fused_step(h, dt, flags, n));
        h->time += dt;
    }
    return MG_OK;
}

/* particles that took the plain step inside learned steps since the last call (then reset) */
int mg_learned_fallbacks(mg_handle* h, uint64_t* count)
{
    if (!h || !count) return MG_ERR_ARG;
    *count = 0;
    if (!h->fallbacks) return MG_OK;
    uint32_t c = 0;
    HIPCHK(hipMemcpyAsync(&c, h->fallbacks, 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemsetAsync(h->fallbacks, 0, 4, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    *count = c;
    return MG_OK;
}

} // extern "C"
