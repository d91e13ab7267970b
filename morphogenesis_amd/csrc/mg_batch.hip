/*
 * mg_batch.hip — many independent handles stepped together (included by mg_capi.hip).
 *
 * Instances are grouped into size classes by particle count; each class runs the fused step
 * as seven launches (pressure, force + advance, diffusion + genes + insert, scan, place,
 * cellsort, gather) that cover all of its instances, so a step of 256 instances costs
 * 7 x (number of classes) dispatches instead of 7 x 256. The kernels look up their instance
 * in a device table (BatchInst) built from each handle exactly as its own fused step builds
 * its kernel arguments, so every instance evolves bit for bit as it would alone. Two tables
 * per class cover the two buffer parities of consecutive steps; two consecutive steps are
 * captured into one hipGraph and replayed. The block size of each class's particle and cell
 * launches is chosen by mg_batch_autotune from measured per-class kernel time.
 */

struct mg_batch {
    struct Class {
        std::vector<int> members;     /* handle indices, largest first */
        unsigned threads = 256;       /* block size of the particle / cell launches */
        int ng = 1;                   /* diffusion sweep template: 1, 2 or 4 groups */
        BatchInst* I[2] = {};         /* tables for the start parity and the next one */
        uint16_t *pmap = nullptr, *cmap = nullptr, *tmap = nullptr;
        uint32_t pblocks = 0, cblocks = 0, tblocks = 0;
        double ms = 0.0;              /* accumulated kernel time (autotuning) */
    };
    std::vector<mg_handle*> hs;
    std::vector<Class> cls;
    std::vector<MgGenome*> dgen;      /* device copies of the handles' genomes */
    std::vector<void*> allocs;        /* table allocations (freed on rebuild) */
    hipStream_t st = nullptr;
    /* what the tables were built for */
    bool built = false;
    std::vector<uint64_t> sig;
    float sig_dt = 0.f;
    unsigned sig_flags = 0;
    /* graphs of two steps, by start parity (0: the parity the tables were built at) */
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    /* per-class timing during autotuning */
    bool timing = false;
    std::vector<hipEvent_t> ev;
    struct Mark {
        int cls, ph;
        size_t ev; /* index of the start event; the end event follows it */
    };
    std::vector<Mark> marks;
};

namespace {

/* size classes by particle count */
const int kBatchClassBound[] = {8192, 24576};
const int kBatchClasses = 3;

uint64_t handle_sig(const mg_handle* h)
{
    return (h->epoch << 8) | ((uint64_t)h->kcur << 5) | ((uint64_t)h->ccur << 4) |
           ((uint64_t)h->cells_fresh << 3) | ((uint64_t)h->grid_scanned << 2) | ((uint64_t)h->icur << 1) |
           (uint64_t)h->gcur;
}

template <typename T> int balloc(mg_batch* b, T** p, size_t count)
{
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)));
    b->allocs.push_back(q);
    *p = (T*)q;
    return MG_OK;
}

void batch_free_tables(mg_batch* b)
{
    for (int k = 0; k < 2; k++)
        if (b->graph[k]) {
            (void)hipGraphExecDestroy(b->graph[k]);
            b->graph[k] = nullptr;
        }
    for (void* q : b->allocs) (void)hipFree(q);
    b->allocs.clear();
    for (auto& c : b->cls) {
        c.I[0] = c.I[1] = nullptr;
        c.pmap = c.cmap = c.tmap = nullptr;
    }
    b->built = false;
}

/* the kernel views a fused step of handle h uses (do_fused_step) */
void fill_inst(BatchInst& t, mg_handle* h, const MgGenome* dgen, float dt, unsigned flags)
{
    t.pressure = make_ds(h);
    t.pressure.ZERO = h->CNT[1 - h->gcur];
    t.pressure.zero_n = (uint32_t)h->C;
    t.pressure.HMW = hit_masks(h);
    t.force = make_ds(h);
    t.force.HMR = t.pressure.HMW;
    t.diffuse = make_ds(h, 1, 0);
    t.diffuse.GMN = h->GMN;
    t.diffuse.cull = 0;
    t.diffuse.HMR = t.pressure.HMW;
    t.place = make_ds(h);
    t.place.CNT = t.place.CNTN;
    t.place.OFF = t.place.OFFN;
    t.sort = make_ds(h, 1, 1);
    t.sort.GMN = h->GMN;
    t.p = h->p;
    t.g = h->g;
    t.dp = diff_plan(h);
    t.gen = dgen;
    t.scan_in = h->CNT[1 - h->gcur];
    t.scan_out = h->OFF[1 - h->gcur];
    t.scan_status = h->scan_status;
    t.scan_ticket = h->scan_ticket;
    t.ctr = h->ctr;
    t.C = h->C;
    t.gsm = grid_scan_max_index(h->g);
    t.genes = (flags & MG_STEP_GENES) ? 1 : 0;
    t.vmax = std::max(std::max(h->g.gmax[0], h->g.gmax[1]), h->g.gmax[2]);
    t.dt = dt;
}

int batch_build(mg_batch* b, float dt, unsigned flags)
{
    batch_free_tables(b);
    const int n = (int)b->hs.size();
    for (int i = 0; i < n; i++) {
        mg_handle* h = b->hs[i];
        HIPCHK(hipMemcpyAsync(b->dgen[i], &h->gen, sizeof(MgGenome), hipMemcpyHostToDevice, b->st));
    }
    for (auto& c : b->cls) {
        if (c.members.empty()) continue;
        const unsigned T = c.threads;
        std::vector<uint16_t> pm, cm, tm;
        std::vector<BatchInst> tab[2];
        c.ng = 1;
        for (int q = 0; q < 2; q++) tab[q].resize(c.members.size());
        for (size_t k = 0; k < c.members.size(); k++) {
            mg_handle* h = b->hs[c.members[k]];
            const uint32_t pnb = nblk(h->N, T), cnb = nblk(h->C, T);
            const uint32_t tnb = (uint32_t)((h->C + SCAN_TILE - 1) / SCAN_TILE);
            const uint32_t pb0 = (uint32_t)pm.size(), cb0 = (uint32_t)cm.size(), tb0 = (uint32_t)tm.size();
            pm.insert(pm.end(), pnb, (uint16_t)k);
            cm.insert(cm.end(), cnb, (uint16_t)k);
            tm.insert(tm.end(), tnb, (uint16_t)k);
            for (int q = 0; q < 2; q++) {
                /* q = 1: the state one fused step later */
                const int ic = h->icur, gc = h->gcur;
                const bool gs = h->grid_scanned, pb = h->pos_binned, cf = h->cells_fresh;
                if (q == 1) {
                    h->icur = 1 - ic;
                    h->gcur = 1 - gc;
                    h->grid_scanned = h->pos_binned = h->cells_fresh = true;
                }
                BatchInst& t = tab[q][k];
                fill_inst(t, h, b->dgen[c.members[k]], dt, flags);
                t.pb0 = pb0; t.pnb = pnb;
                t.cb0 = cb0; t.cnb = cnb;
                t.tb0 = tb0; t.tnb = tnb;
                h->icur = ic;
                h->gcur = gc;
                h->grid_scanned = gs;
                h->pos_binned = pb;
                h->cells_fresh = cf;
            }
            const int ng = tab[0][k].dp.ng;
            c.ng = std::max(c.ng, ng <= 1 ? 1 : (ng == 2 ? 2 : 4));
        }
        c.pblocks = (uint32_t)pm.size();
        c.cblocks = (uint32_t)cm.size();
        c.tblocks = (uint32_t)tm.size();
        for (int q = 0; q < 2; q++) {
            CHK(balloc(b, &c.I[q], tab[q].size()));
            HIPCHK(hipMemcpyAsync(c.I[q], tab[q].data(), tab[q].size() * sizeof(BatchInst),
                                  hipMemcpyHostToDevice, b->st));
        }
        CHK(balloc(b, &c.pmap, pm.size()));
        CHK(balloc(b, &c.cmap, cm.size()));
        CHK(balloc(b, &c.tmap, tm.size()));
        HIPCHK(hipMemcpyAsync(c.pmap, pm.data(), pm.size() * 2, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(c.cmap, cm.data(), cm.size() * 2, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(c.tmap, tm.data(), tm.size() * 2, hipMemcpyHostToDevice, b->st));
        /* the host copies go away with this scope: wait for the uploads */
        HIPCHK(hipStreamSynchronize(b->st));
    }
    HIPCHK(hipStreamSynchronize(b->st));
    b->sig.resize(b->hs.size());
    for (size_t i = 0; i < b->hs.size(); i++) b->sig[i] = handle_sig(b->hs[i]);
    b->sig_dt = dt;
    b->sig_flags = flags;
    b->built = true;
    return MG_OK;
}

/* start parity of the next step relative to the tables: 0, 1, or -1 (rebuild needed) */
int batch_parity(const mg_batch* b, float dt, unsigned flags)
{
    if (!b->built || b->sig_dt != dt || b->sig_flags != flags) return -1;
    bool same = true, next = true;
    for (size_t i = 0; i < b->hs.size(); i++) {
        const mg_handle* h = b->hs[i];
        if (handle_sig(h) != b->sig[i]) same = false;
        /* one fused step after the build state: fresh flags set, icur and gcur flipped */
        const uint64_t s0 = b->sig[i];
        const uint64_t nxt = (s0 & ~0xFull) | (1ull << 3) | (1ull << 2) | ((~s0) & 3ull);
        if (handle_sig(h) != nxt) next = false;
    }
    if (same) return 0;
    if (next) return 1;
    return -1;
}

void batch_mark(mg_batch* b, int cls, int ph, bool start)
{
    if (!b->timing) return;
    if (start) {
        while (b->ev.size() < b->marks.size() * 2 + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            b->ev.push_back(e);
        }
        const size_t k = b->marks.size() * 2;
        (void)hipEventRecord(b->ev[k], b->st);
        b->marks.push_back({cls, ph, k});
    } else {
        (void)hipEventRecord(b->ev[b->marks.back().ev + 1], b->st);
    }
}

/* one fused step of every instance, tables of parity q */
int batch_launch(mg_batch* b, int q)
{
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_PRESSURE, true);
        hipLaunchKernelGGL(kb_pressure, dim3(c.pblocks), dim3(c.threads), band_lds<0>(c.threads), b->st, c.I[q], c.pmap);
        batch_mark(b, (int)ci, MG_PH_PRESSURE, false);
    }
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_FORCE, true);
        hipLaunchKernelGGL(kb_force_advance, dim3(c.pblocks), dim3(c.threads), band_lds<1>(c.threads), b->st, c.I[q],
                           c.pmap);
        batch_mark(b, (int)ci, MG_PH_FORCE, false);
    }
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_DIFFUSION, true);
        if (c.ng <= 1)
            hipLaunchKernelGGL(kb_diffuse_genes_insert<1>, dim3(c.pblocks), dim3(c.threads), band_lds<1>(c.threads),
                               b->st, c.I[q], c.pmap);
        else if (c.ng == 2)
            hipLaunchKernelGGL(kb_diffuse_genes_insert<2>, dim3(c.pblocks), dim3(c.threads), band_lds<2>(c.threads),
                               b->st, c.I[q], c.pmap);
        else
            hipLaunchKernelGGL(kb_diffuse_genes_insert<4>, dim3(c.pblocks), dim3(c.threads), band_lds<4>(c.threads),
                               b->st, c.I[q], c.pmap);
        batch_mark(b, (int)ci, MG_PH_DIFFUSION, false);
    }
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_SCAN, true);
        hipLaunchKernelGGL(kb_scan, dim3(c.tblocks), dim3(256), 0, b->st, c.I[q], c.tmap);
        batch_mark(b, (int)ci, MG_PH_SCAN, false);
    }
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_PLACE, true);
        hipLaunchKernelGGL(kb_place, dim3(c.pblocks), dim3(c.threads), 0, b->st, c.I[q], c.pmap);
        hipLaunchKernelGGL(kb_cellsort, dim3(c.cblocks), dim3(c.threads), 0, b->st, c.I[q], c.cmap);
        batch_mark(b, (int)ci, MG_PH_PLACE, false);
    }
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_SORT, true);
        hipLaunchKernelGGL(kb_sort_gather, dim3(c.pblocks), dim3(c.threads), 0, b->st, c.I[q], c.pmap);
        batch_mark(b, (int)ci, MG_PH_SORT, false);
    }
    HIPCHK(hipGetLastError());
    return MG_OK;
}

/* the host-side state a fused step leaves in each handle */
void batch_after_steps(mg_batch* b, int nsteps, float dt)
{
    for (mg_handle* h : b->hs) {
        if (nsteps & 1) {
            h->icur = 1 - h->icur;
            h->gcur = 1 - h->gcur;
        }
        if (nsteps > 0) {
            h->force_zero = true;
            h->normalized = false;
            h->gene_grid_stale = true;
            h->grid_scanned = true;
            h->pos_binned = true;
            h->cells_fresh = true;
        }
        h->time += (double)nsteps * dt;
    }
}

int batch_check(mg_batch* b)
{
    for (mg_handle* h : b->hs) {
        CHK(ready(h));
        if (!h->force_zero) {
            set_err("mg_batch_step: every handle must be in the sorted state (after a physical sort "
                    "or a step)");
            return MG_ERR_STATE;
        }
    }
    return MG_OK;
}

/* nsteps fused steps; with use_graph, pairs of steps replay a captured graph */
int batch_steps(mg_batch* b, int nsteps, float dt, unsigned flags, bool use_graph)
{
    if (flags & MG_STEP_SPECIAL) {
        set_err("mg_batch_step: MG_STEP_SPECIAL is not supported for batches");
        return MG_ERR_UNSUPPORTED;
    }
    CHK(batch_check(b));
    for (mg_handle* h : b->hs) {
        CHK(maybe_rehome(h));
        h->steps_since_rehome += nsteps;
        ensure_hit_masks(h);
    }
    flags &= MG_STEP_GENES;
    int q = batch_parity(b, dt, flags);
    if (q < 0) {
        CHK(batch_build(b, dt, flags));
        q = 0;
    }
    int s = 0;
    if (use_graph && nsteps >= 2) {
        if (!b->graph[q]) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            int r1 = batch_launch(b, q);
            int r2 = (r1 == MG_OK) ? batch_launch(b, 1 - q) : r1;
            hipError_t ec = hipStreamEndCapture(b->st, &g);
            if (r1 != MG_OK || r2 != MG_OK) return r1 != MG_OK ? r1 : r2;
            if (ec != hipSuccess) {
                set_err("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
                return MG_ERR_HIP;
            }
            hipError_t ei = hipGraphInstantiate(&b->graph[q], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                b->graph[q] = nullptr;
                set_err("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
                return MG_ERR_HIP;
            }
        }
        for (; s + 2 <= nsteps; s += 2) HIPCHK(hipGraphLaunch(b->graph[q], b->st));
    }
    for (; s < nsteps; s++) {
        CHK(batch_launch(b, q));
        q = 1 - q;
    }
    batch_after_steps(b, nsteps, dt);
    return MG_OK;
}

} // namespace

extern "C" {

int mg_batch_create(mg_batch** out, mg_handle* const* handles, int n)
{
    if (!out || !handles || n <= 0 || n > 65535) {
        set_err("mg_batch_create: need 1..65535 handles");
        return MG_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        if (!handles[i]) { set_err("mg_batch_create: null handle %d", i); return MG_ERR_ARG; }
        if (handles[i]->batch) { set_err("mg_batch_create: handle %d is already in a batch", i); return MG_ERR_STATE; }
        for (int j = 0; j < i; j++)
            if (handles[j] == handles[i]) { set_err("mg_batch_create: handle %d repeated", i); return MG_ERR_ARG; }
    }
    mg_batch* b = new mg_batch;
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        set_err("mg_batch_create: hipStreamCreate failed");
        return MG_ERR_HIP;
    }
    b->hs.assign(handles, handles + n);
    b->cls.resize(kBatchClasses);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->hs[x]->N > b->hs[y]->N; });
    for (int i : order) {
        int c = 0;
        while (c < kBatchClasses - 1 && b->hs[i]->N >= kBatchClassBound[c]) c++;
        b->cls[c].members.push_back(i);
    }
    b->dgen.resize(n, nullptr);
    for (int i = 0; i < n; i++) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(MgGenome)) != hipSuccess) {
            for (int k = 0; k < i; k++) (void)hipFree(b->dgen[k]);
            (void)hipStreamDestroy(b->st);
            delete b;
            set_err("mg_batch_create: hipMalloc failed");
            return MG_ERR_HIP;
        }
        b->dgen[i] = (MgGenome*)q;
    }
    /* from here on the handles queue their work on the batch's stream */
    for (mg_handle* h : b->hs) {
        (void)hipStreamSynchronize(h->st);
        h->st = b->st;
        h->batch = b;
    }
    *out = b;
    return MG_OK;
}

int mg_batch_destroy(mg_batch* b)
{
    if (!b) return MG_ERR_ARG;
    (void)hipStreamSynchronize(b->st);
    batch_free_tables(b);
    for (MgGenome* g : b->dgen) (void)hipFree(g);
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    for (mg_handle* h : b->hs) {
        h->st = h->own_st;
        h->batch = nullptr;
    }
    (void)hipStreamDestroy(b->st);
    delete b;
    return MG_OK;
}

int mg_batch_step(mg_batch* b, int nsteps, float dt, unsigned flags)
{
    if (!b || nsteps < 0) { set_err("mg_batch_step: bad arguments"); return MG_ERR_ARG; }
    return batch_steps(b, nsteps, dt, flags, (flags & MG_STEP_GRAPH) != 0);
}

int mg_batch_sync(mg_batch* b)
{
    if (!b) return MG_ERR_ARG;
    HIPCHK(hipStreamSynchronize(b->st));
    for (mg_handle* h : b->hs) {
        uint32_t err = 0;
        HIPCHK(hipMemcpy(&err, h->ctr + MG_C_ERR, 4, hipMemcpyDeviceToHost));
        if (err) {
            set_err("mg_batch_sync: a kernel gave up a bounded wait; results are invalid");
            return MG_ERR_HIP;
        }
    }
    return MG_OK;
}

int mg_batch_classes(mg_batch* b, int* nclass, int* members, unsigned* threads)
{
    if (!b || !nclass) return MG_ERR_ARG;
    *nclass = (int)b->cls.size();
    for (size_t c = 0; c < b->cls.size(); c++) {
        if (members) members[c] = (int)b->cls[c].members.size();
        if (threads) threads[c] = b->cls[c].threads;
    }
    return MG_OK;
}

int mg_batch_autotune(mg_batch* b, int steps, float dt, unsigned flags, unsigned* threads_out)
{
    if (!b || steps < 1) { set_err("mg_batch_autotune: bad arguments"); return MG_ERR_ARG; }
    const unsigned cand[] = {256, 128, 64};
    const int nc = (int)b->cls.size();
    std::vector<double> best(nc, 1e300);
    std::vector<unsigned> pick(nc, 256);
    for (unsigned T : cand) {
        for (auto& c : b->cls) c.threads = T;
        b->built = false; /* block geometry changed */
        b->timing = true;
        b->marks.clear();
        int r = batch_steps(b, steps, dt, flags, false);
        b->timing = false;
        if (r != MG_OK) return r;
        HIPCHK(hipStreamSynchronize(b->st));
        std::vector<double> ms(nc, 0.0);
        for (auto& m : b->marks) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, b->ev[m.ev], b->ev[m.ev + 1]));
            ms[m.cls] += t;
        }
        b->marks.clear();
        for (int c = 0; c < nc; c++)
            if (!b->cls[c].members.empty() && ms[c] < best[c]) {
                best[c] = ms[c];
                pick[c] = T;
            }
    }
    for (int c = 0; c < nc; c++) {
        b->cls[c].threads = pick[c];
        b->cls[c].ms = b->cls[c].members.empty() ? 0.0 : best[c] / steps;
        if (threads_out) threads_out[c] = pick[c];
    }
    b->built = false;
    return MG_OK;
}

/* Like mg_batch_step without graph replay, with a HIP event pair around every launch group;
 * ms_out[phase] (MG_PH_* ids, MG_PH_COUNT entries) accumulates milliseconds over all
 * classes and calls_out[phase] the launch groups (either may be null). */
int mg_batch_step_timed(mg_batch* b, int nsteps, float dt, unsigned flags, double* ms_out, int* calls_out)
{
    if (!b || nsteps < 0) { set_err("mg_batch_step_timed: bad arguments"); return MG_ERR_ARG; }
    b->timing = true;
    b->marks.clear();
    int r = batch_steps(b, nsteps, dt, flags, false);
    b->timing = false;
    if (r != MG_OK) return r;
    HIPCHK(hipStreamSynchronize(b->st));
    for (auto& m : b->marks) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, b->ev[m.ev], b->ev[m.ev + 1]));
        if (ms_out) ms_out[m.ph] += t;
        if (calls_out) calls_out[m.ph] += 1;
    }
    b->marks.clear();
    return MG_OK;
}

} // extern "C"
