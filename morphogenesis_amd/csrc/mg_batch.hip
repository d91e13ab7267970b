/*
 * mg_batch.hip — many independent handles stepped together (included by mg_capi.hip).
 *
 * A step of every instance is one persistent launch (kb_step): resident workgroups walk the
 * six phases of the fused step (pressure, force + advance, diffusion + genes + insert,
 * scan, place, rank + gather) over per-phase queues of (instance, block) items, with grid
 * barriers between phases. The alternative is the dispatch form: instances grouped into
 * size classes by particle count, each class running the six phases as launches that cover
 * all of its instances (6 x classes dispatches per step). Both look up their instance in a
 * device table (BatchInst) built from each handle exactly as its own fused step builds its
 * kernel arguments, so every instance evolves bit for bit as it would alone. Tables cover the
 * two buffer parities of consecutive steps; two consecutive steps are captured into one
 * hipGraph and replayed. mg_batch_autotune measures the dispatch form at each block size per
 * class and the persistent form at each workgroup count per CU, and keeps the fastest.
 */

struct mg_batch {
    struct Class {
        std::vector<int> members;     /* handle indices, largest first */
        unsigned threads = 256;       /* block size of the particle / cell launches */
        int pshape = 0;               /* pressure sweep tile shape (kPShapes) */
        int ng = 1;                   /* diffusion sweep template: 1, 2 or 4 groups */
        BatchInst* I[2] = {};         /* tables for the start parity and the next one */
        uint16_t *pmap = nullptr, *tmap = nullptr;
        uint32_t pblocks = 0, tblocks = 0;
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
    /* the persistent form: one kb_step launch per step (when every instance fits its
     * 16-bit block and instance fields) */
    bool persistent = true;
    bool persist_ok = false;
    int persist_ng = 1;
    unsigned persist_per_cu = 0;      /* resident workgroups per CU (occupancy) */
    unsigned persist_use_per_cu = 0;  /* chosen workgroups per CU (0: all resident) */
    int cus = 0;
    BatchInst* PI[2] = {};            /* every instance, by parity */
    uint32_t* pitems[PERSIST_PHASES] = {};
    uint32_t pnitems[PERSIST_PHASES] = {};
    PersistCtl* pctl = nullptr;
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

/* tile shapes of the pressure sweep the autotuner chooses from: candidates tested at a time
 * and the slots of a wavefront's staged band (the force and diffusion sweeps mostly visit the
 * pressure sweep's recorded hits, so their shapes stay fixed) */
struct PShape {
    int K, cap;
};
const PShape kPShapes[] = {{8, 128}, {4, 128}, {8, 64}, {4, 64}};
const int kNPShapes = 4;

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
    b->PI[0] = b->PI[1] = nullptr;
    for (int k = 0; k < PERSIST_PHASES; k++) {
        b->pitems[k] = nullptr;
        b->pnitems[k] = 0;
    }
    b->pctl = nullptr;
    b->persist_ok = false;
    for (auto& c : b->cls) {
        c.I[0] = c.I[1] = nullptr;
        c.pmap = c.tmap = nullptr;
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
    t.pressure.ETW = bond_records(h, t.pressure.HMW);
    t.force = make_ds(h);
    t.force.HMR = t.pressure.HMW;
    t.force.ETR = t.pressure.ETW;
    t.diffuse = make_ds(h, 1, 0);
    t.diffuse.GMN = h->GMN;
    t.diffuse.cull = 0;
    t.diffuse.HMR = t.pressure.HMW;
    t.place = make_ds(h);
    t.place.CNT = t.place.CNTN;
    t.place.OFF = t.place.OFFN;
    t.sort = make_ds(h, 1, 1);
    t.sort.GMN = h->GMN;
    t.sort.CNT = t.sort.CNTN; /* the grid k_place bins into */
    t.sort.OFF = t.sort.OFFN;
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

template <int NG> size_t persist_lds() { return band_lds<(NG > 1 ? NG : 1)>(256); }

template <int NG> int persist_occupancy(unsigned* per_cu)
{
    int nb = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)kb_step<NG>, 256, persist_lds<NG>()));
    *per_cu = (unsigned)std::max(nb, 0);
    return MG_OK;
}

/* Tables and item lists of the persistent form: every instance (largest first), its blocks of
 * 256 particles / cells and its scan tiles, for both parities. */
int persist_build(mg_batch* b, float dt, unsigned flags)
{
    const int n = (int)b->hs.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->hs[x]->N > b->hs[y]->N; });
    std::vector<BatchInst> tab[2];
    std::vector<uint32_t> items[PERSIST_PHASES];
    int ng = 1;
    bool fits = n <= 65535;
    for (int k = 0; k < n && fits; k++) {
        mg_handle* h = b->hs[order[k]];
        const uint32_t pnb = nblk(h->N, 256);
        const uint32_t tnb = (uint32_t)((h->C + SCAN_TILE - 1) / SCAN_TILE);
        if (pnb > 65536 || tnb > 65536) { fits = false; break; }
        for (int q = 0; q < 2; q++) {
            const int ic = h->icur, gc = h->gcur;
            const bool gs = h->grid_scanned, pb = h->pos_binned, cf = h->cells_fresh;
            if (q == 1) {
                h->icur = 1 - ic;
                h->gcur = 1 - gc;
                h->grid_scanned = h->pos_binned = h->cells_fresh = true;
            }
            BatchInst t;
            fill_inst(t, h, b->dgen[order[k]], dt, flags);
            t.pb0 = 0; t.pnb = pnb;
            t.tb0 = 0; t.tnb = tnb;
            tab[q].push_back(t);
            h->icur = ic;
            h->gcur = gc;
            h->grid_scanned = gs;
            h->pos_binned = pb;
            h->cells_fresh = cf;
        }
        const int g = tab[0].back().dp.ng;
        ng = std::max(ng, g <= 1 ? 1 : (g == 2 ? 2 : 4));
        const uint32_t base = (uint32_t)k << 16;
        for (uint32_t x = 0; x < pnb; x++)
            for (int ph : {0, 1, 2, 4, 5}) items[ph].push_back(base | x);
        for (uint32_t x = 0; x < tnb; x++) items[3].push_back(base | x);
    }
    b->persist_ok = false;
    if (!fits) return MG_OK; /* the dispatch form serves this batch */
    b->persist_ng = ng;
    if (ng <= 1) CHK(persist_occupancy<1>(&b->persist_per_cu));
    else if (ng == 2) CHK(persist_occupancy<2>(&b->persist_per_cu));
    else CHK(persist_occupancy<4>(&b->persist_per_cu));
    if (b->persist_per_cu == 0) return MG_OK;
    if (!b->cus) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    for (int q = 0; q < 2; q++) {
        CHK(balloc(b, &b->PI[q], tab[q].size()));
        HIPCHK(hipMemcpyAsync(b->PI[q], tab[q].data(), tab[q].size() * sizeof(BatchInst), hipMemcpyHostToDevice, b->st));
    }
    for (int ph = 0; ph < PERSIST_PHASES; ph++) {
        CHK(balloc(b, &b->pitems[ph], items[ph].size()));
        HIPCHK(hipMemcpyAsync(b->pitems[ph], items[ph].data(), items[ph].size() * 4, hipMemcpyHostToDevice, b->st));
        b->pnitems[ph] = (uint32_t)items[ph].size();
    }
    CHK(balloc(b, &b->pctl, 1));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, b->st, (uint32_t*)b->pctl, 0u,
                       sizeof(PersistCtl) / sizeof(uint32_t));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->st)); /* the host tables go away with this scope */
    b->persist_ok = true;
    return MG_OK;
}

unsigned persist_workgroups(const mg_batch* b)
{
    const unsigned per_cu = b->persist_use_per_cu ? std::min(b->persist_use_per_cu, b->persist_per_cu)
                                                  : b->persist_per_cu;
    return per_cu * (unsigned)b->cus;
}

int persist_launch(mg_batch* b, int q)
{
    PersistPlan P;
    P.I = b->PI[q];
    for (int ph = 0; ph < PERSIST_PHASES; ph++) {
        P.items[ph] = b->pitems[ph];
        P.nitems[ph] = b->pnitems[ph];
    }
    P.ctl = b->pctl;
    const unsigned G = persist_workgroups(b);
    if (b->persist_ng <= 1)
        hipLaunchKernelGGL(kb_step<1>, dim3(G), dim3(256), persist_lds<1>(), b->st, P);
    else if (b->persist_ng == 2)
        hipLaunchKernelGGL(kb_step<2>, dim3(G), dim3(256), persist_lds<2>(), b->st, P);
    else
        hipLaunchKernelGGL(kb_step<4>, dim3(G), dim3(256), persist_lds<4>(), b->st, P);
    HIPCHK(hipGetLastError());
    return MG_OK;
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
        std::vector<uint16_t> pm, tm;
        std::vector<BatchInst> tab[2];
        c.ng = 1;
        for (int q = 0; q < 2; q++) tab[q].resize(c.members.size());
        for (size_t k = 0; k < c.members.size(); k++) {
            mg_handle* h = b->hs[c.members[k]];
            const uint32_t pnb = nblk(h->N, T);
            const uint32_t tnb = (uint32_t)((h->C + SCAN_TILE - 1) / SCAN_TILE);
            const uint32_t pb0 = (uint32_t)pm.size(), tb0 = (uint32_t)tm.size();
            pm.insert(pm.end(), pnb, (uint16_t)k);
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
        c.tblocks = (uint32_t)tm.size();
        for (int q = 0; q < 2; q++) {
            CHK(balloc(b, &c.I[q], tab[q].size()));
            HIPCHK(hipMemcpyAsync(c.I[q], tab[q].data(), tab[q].size() * sizeof(BatchInst),
                                  hipMemcpyHostToDevice, b->st));
        }
        CHK(balloc(b, &c.pmap, pm.size()));
        CHK(balloc(b, &c.tmap, tm.size()));
        HIPCHK(hipMemcpyAsync(c.pmap, pm.data(), pm.size() * 2, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(c.tmap, tm.data(), tm.size() * 2, hipMemcpyHostToDevice, b->st));
        /* the host copies go away with this scope: wait for the uploads */
        HIPCHK(hipStreamSynchronize(b->st));
    }
    HIPCHK(hipStreamSynchronize(b->st));
    CHK(persist_build(b, dt, flags));
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

/* one fused step of every instance, tables of parity q: the persistent launch, or the dispatch
 * form (always when timing per phase) */
int batch_launch(mg_batch* b, int q)
{
    if (b->persistent && b->persist_ok && !b->timing) return persist_launch(b, q);
    for (size_t ci = 0; ci < b->cls.size(); ci++) {
        auto& c = b->cls[ci];
        if (c.members.empty()) continue;
        batch_mark(b, (int)ci, MG_PH_PRESSURE, true);
        switch (c.pshape) {
        case 1:
            hipLaunchKernelGGL((kb_pressure<4, 128>), dim3(c.pblocks), dim3(c.threads), (band_lds<0, 128>(c.threads)),
                               b->st, c.I[q], c.pmap);
            break;
        case 2:
            hipLaunchKernelGGL((kb_pressure<8, 64>), dim3(c.pblocks), dim3(c.threads), (band_lds<0, 64>(c.threads)),
                               b->st, c.I[q], c.pmap);
            break;
        case 3:
            hipLaunchKernelGGL((kb_pressure<4, 64>), dim3(c.pblocks), dim3(c.threads), (band_lds<0, 64>(c.threads)),
                               b->st, c.I[q], c.pmap);
            break;
        default:
            hipLaunchKernelGGL((kb_pressure<8, 128>), dim3(c.pblocks), dim3(c.threads), (band_lds<0, 128>(c.threads)),
                               b->st, c.I[q], c.pmap);
            break;
        }
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
        if (handles[i]->sph != MG_SPH_REFERENCE) {
            set_err("mg_batch_create: handle %d uses the poly6/spiky kernels; batches step the reference kernels", i);
            return MG_ERR_UNSUPPORTED;
        }
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
    if (b->pctl) {
        uint32_t err = 0;
        HIPCHK(hipMemcpy(&err, &b->pctl->err, 4, hipMemcpyDeviceToHost));
        if (err) {
            set_err("mg_batch_sync: a grid barrier of the persistent step gave up its bounded wait; results are invalid");
            return MG_ERR_HIP;
        }
    }
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

/* `steps` dispatch-form steps with HIP events per launch group; ms[class] accumulates the
 * milliseconds of the launches of phase `ph` (every phase when ph < 0) */
static int batch_time_classes(mg_batch* b, int steps, float dt, unsigned flags, int ph,
                              std::vector<double>& ms)
{
    b->timing = true;
    b->marks.clear();
    int r = batch_steps(b, steps, dt, flags, false);
    b->timing = false;
    if (r != MG_OK) return r;
    HIPCHK(hipStreamSynchronize(b->st));
    ms.assign(b->cls.size(), 0.0);
    for (auto& m : b->marks) {
        if (ph >= 0 && m.ph != ph) continue;
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, b->ev[m.ev], b->ev[m.ev + 1]));
        ms[m.cls] += t;
    }
    b->marks.clear();
    return MG_OK;
}

int mg_batch_autotune(mg_batch* b, int steps, float dt, unsigned flags, unsigned* threads_out)
{
    if (!b || steps < 1) { set_err("mg_batch_autotune: bad arguments"); return MG_ERR_ARG; }
    const unsigned cand[] = {256, 128, 64};
    const int nc = (int)b->cls.size();
    std::vector<double> best(nc, 1e300), ms;
    std::vector<unsigned> pick(nc, 256);
    /* block size per class, at the default tile shape */
    for (auto& c : b->cls) c.pshape = 0;
    for (unsigned T : cand) {
        for (auto& c : b->cls) c.threads = T;
        b->built = false; /* block geometry changed */
        CHK(batch_time_classes(b, steps, dt, flags, -1, ms));
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
    /* the pressure sweep's tile shape per class, at the chosen block sizes */
    std::vector<double> pbest(nc, 1e300);
    std::vector<int> pshape(nc, 0);
    for (int sh = 0; sh < kNPShapes; sh++) {
        for (auto& c : b->cls) c.pshape = sh;
        CHK(batch_time_classes(b, steps, dt, flags, MG_PH_PRESSURE, ms));
        for (int c = 0; c < nc; c++)
            if (!b->cls[c].members.empty() && ms[c] < pbest[c]) {
                pbest[c] = ms[c];
                pshape[c] = sh;
            }
    }
    for (int c = 0; c < nc; c++) b->cls[c].pshape = pshape[c];
    /* the dispatch form with these block sizes against the persistent form at each workgroup
     * count per CU: whole steps between two events */
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto time_steps = [&](double* ms) -> int {
        HIPCHK(hipEventRecord(e0, b->st));
        CHK(batch_steps(b, steps, dt, flags, false));
        HIPCHK(hipEventRecord(e1, b->st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        return MG_OK;
    };
    double best_ms = 1e300;
    bool best_persistent = false;
    unsigned best_per_cu = 0;
    b->persistent = false;
    int r = time_steps(&best_ms);
    if (r == MG_OK && b->persist_ok) {
        const unsigned maxc = b->persist_per_cu;
        b->persistent = true;
        for (unsigned per_cu = 1; per_cu <= maxc && r == MG_OK; per_cu++) {
            b->persist_use_per_cu = per_cu;
            double ms = 0.0;
            r = time_steps(&ms);
            if (r == MG_OK) r = mg_batch_sync(b);
            if (r == MG_OK && ms < best_ms) {
                best_ms = ms;
                best_persistent = true;
                best_per_cu = per_cu;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->persistent = best_persistent;
    b->persist_use_per_cu = best_per_cu;
    return r;
}

int mg_batch_class_config(mg_batch* b, int cls, int* instances, unsigned* threads, int* cand_batch, int* band_cap)
{
    if (!b || cls < 0 || cls >= (int)b->cls.size()) return MG_ERR_ARG;
    const auto& c = b->cls[cls];
    if (instances) *instances = (int)c.members.size();
    if (threads) *threads = c.threads;
    if (cand_batch) *cand_batch = kPShapes[c.pshape].K;
    if (band_cap) *band_cap = kPShapes[c.pshape].cap;
    return MG_OK;
}

int mg_batch_set_tile_shape(mg_batch* b, int cls, int cand_batch, int band_cap)
{
    if (!b || cls < 0 || cls >= (int)b->cls.size()) {
        set_err("mg_batch_set_tile_shape: bad arguments");
        return MG_ERR_ARG;
    }
    for (int sh = 0; sh < kNPShapes; sh++)
        if (kPShapes[sh].K == cand_batch && kPShapes[sh].cap == band_cap) {
            b->cls[cls].pshape = sh;
            return MG_OK;
        }
    set_err("mg_batch_set_tile_shape: no pressure sweep of candidate batch %d and band %d", cand_batch, band_cap);
    return MG_ERR_ARG;
}

int mg_batch_launch_config(mg_batch* b, int* persistent, unsigned* workgroups)
{
    if (!b) return MG_ERR_ARG;
    const bool on = b->persistent && b->persist_ok;
    if (persistent) *persistent = on ? 1 : 0;
    if (workgroups) *workgroups = on ? persist_workgroups(b) : 0u;
    return MG_OK;
}

int mg_batch_set_persistent(mg_batch* b, int on, unsigned workgroups_per_cu)
{
    if (!b) return MG_ERR_ARG;
    b->persistent = on != 0;
    b->persist_use_per_cu = workgroups_per_cu;
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
