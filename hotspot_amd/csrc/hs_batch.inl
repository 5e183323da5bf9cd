// Batched engine: many models advanced by shared launches.  Jobs (a steady
// solve, a transient sequence, or a steady solve followed by a transient
// sequence from its block temperatures) are admitted first come, first
// served while their device footprint fits the engine's budget.  Resident
// jobs form groups by phase and kernel form.  Each group advances in chunks
// (a few steady iterations, or a slice of transient intervals) on its own
// stream, and the host takes a group's results once the event recorded
// after its chunk has completed, independently of the other groups: a
// steady solve that converges is reported while the transient groups keep
// running, and a job that arrives joins its group at the group's next chunk.
// Each problem's arithmetic is the same as when it is solved alone
// (per-problem scalars, per-problem stopping, reductions in a fixed slot
// order), so a job's result does not depend on what it is batched with or
// when it arrived.  Backward-Euler transients (the integrator of the
// reference's SuperLU build) advance one implicit solve per interval: the
// solves of every such job run in shared launches, and the host moves a job
// to its next interval when its solve has converged.  Included at the end of
// hs_engine.inl.
#include <chrono>
#include <deque>
#include <map>
#include <thread>

namespace hs {

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum JobPhase { J_QUEUED = 0, J_STEADY = 1, J_TRANS = 2, J_DONE = 3, J_FAILED = 4 };

struct EngJob {
    int id = -1;
    int phase = J_QUEUED;
    Model *M = nullptr;
    std::unique_ptr<DevModel> D;
    std::vector<double> power;          // rows x nblk
    int rows = 1;
    std::vector<double> init;           // start temperatures (empty: steady first)
    int nsteps = 0;
    double dt = 0.0;
    int nblk = 0;
    bool hetero = false, adv = false;   // heterogeneous coefficients; advection (non-symmetric G)
    size_t footprint = 0;               // estimate used for admission
    size_t held = 0;                    // device bytes while resident
    std::vector<double> result;         // final block vector
    std::vector<double> seq;            // nsteps x nblk
    double t_submit = 0, t_admit = 0, t_done = 0;
    int steady_iters = 0, refinements = 0, rk4_calls = 0;
    long taylor_sweeps = 0;
    double taylor_node_bytes = 0;
    int step_seen = 0;                  // intervals done at the last look at its group's flags
    int holds = 0;                      // Taylor: intervals held and run again
    int kplan = 0;                      // Taylor: Horner steps per sub-step of its plan
    int kmin = 0;                       // Taylor: fewest steps its group launched per sub-step (0: the plan's)
    unsigned long stamp = 0;            // the model's coefficients when its device state was built
    bool use_mg = false;                // steady phase: multigrid-preconditioned solve
    double setup_s = 0.0;               // host time of the admission (device state, inputs)
    bool resident = false;              // admitted on its model's resident device state
    // the temperature-leakage fixed point of the steady phase (the model's
    // leakage_used, temperature.c:708-746): each solve's power adds the
    // leakage at the last solve's block temperatures, the first at zero
    bool leak = false;
    LeakBlocks lb;                      // the power-layer blocks, copied at submission
    std::vector<double> leak_old, leak_power;
    int leak_iter = 0;                  // steady solves done (leakage or natural-convection fixed point)
    long be_iters = 0;                  // backward Euler: solver iterations over its intervals
    // a steady solve under the package model (package.c:353-397): the
    // sink's r_convec from the fan and sink (forced convection, one solve)
    // or the natural-convection fixed point of hotspot.c:667-682, each
    // re-solve warm-started from the last with r_convec from the average
    // sink temperature; the device state's convection is then the job's own
    bool package = false, natural = false;
    package_config_t pkg{};
    SinkModel sink;
    double r_convec = 0.0;              // of the solve in progress, then of the model after the last
    // a finished transient's sequence on its way to the host: pinned
    // staging and the event after the copy (the job is done when it completes)
    double *h_seq = nullptr;
    size_t dl_n = 0;                    // doubles in flight
    bool dl_seq = false;                // the sequence (else the final block vector)
    hipEvent_t ev_dl = nullptr;
    std::string error;                  // J_FAILED: the reason
};

// Kernel forms of the transient groups: homogeneous problems whose FP32
// Taylor sweeps take float4 tiles, other homogeneous problems, heterogeneous
// problems.  A group's launches take one form for all its members, and the
// forms round differently, so a problem always goes to the group of its own
// form: its result does not depend on what else is resident.
constexpr int kTransForms = 3;
static int trans_form(const EngJob &j);
// Transient parts of an engine: each transient form is split into this many
// groups, one per stream.  While one
// part runs the latency-bound kernels of its interval (power map, first
// term, finishes), the other's Horner sweeps use the memory bandwidth.
constexpr int kTransParts = 2;
// device time a transient chunk aims at: long enough to amortise the host's
// look at the flags, short enough that an arriving job waits little to join
constexpr double kSliceMs = 20.0;
// ... and while steady solves are resident (they go first)
constexpr double kSliceYieldMs = 2.0;
// longest a tick waits for some group's chunk to end
constexpr double kWaitMs = 2.0;
// intervals of a transient group's first chunk, which times them
constexpr int kFirstChunk = 2;
// streams of the steady groups: the kernels of a steady round are small and
// latency-bound, so groups of different shapes overlap on the device; a
// group's chunk takes a stream that no other steady chunk is running on
constexpr int kSteadyStreams = 6;
// inner iterations of a multigrid-preconditioned CG round: an FP32 inner
// solve gains its four digits in about 10 of them, and a refinement step
// takes place at the start of a round, so a solve takes one round per
// refinement (rounds of 3 at the default 128 steady iterations per chunk)
constexpr int kMgRoundIters = 12;

// A group of resident jobs in one phase with one kernel form (homogeneous or
// heterogeneous coefficients): descriptor array and work-lists on the device.
struct BatchGroup {
    Group g;
    std::vector<EngJob *> members;
    bool dirty = true;
    DevPool *pool = nullptr, *hpool = nullptr;
    hipStream_t stream = nullptr;       // the stream its chunks run on
    size_t cap_desc = 0, cap_work = 0, cap_gwork = 0, cap_flags = 0, cap_trec = 0, cap_twork = 0;
    // a chunk's results, written by its last kernel straight into pinned
    // host memory (mapped): no copy-engine transfer that would queue behind
    // the chunk and hold up other streams' copies meanwhile
    double *h_flags = nullptr, *m_flags = nullptr;   // host / device address
    int *h_kmax = nullptr, *m_kmax = nullptr;         // Taylor: the term counts of the last chunk
    // interval bodies captured per count of Horner steps launched (kgraph)
    Replayer rp[kTayMaxK + 1];
    SeqForm form{};
    // multigrid steady groups: per level, the descriptors and column work-lists
    std::vector<MgGroup> lv;
    std::vector<void *> lv_bufs;
    // set-up copies of the last membership change: pinned staging blocks in
    // flight on the set-up stream, and the event after them that the
    // group's next chunk waits on
    std::vector<void *> staged;
    hipEvent_t ev_sync = nullptr;
    bool synced = false;                // a set-up since the last chunk: the chunk waits on ev_sync
    // the chunk in flight
    bool busy = false;
    int ran = 0;                        // members launched in it (the first `ran`; later ones joined since)
    int chunk = 0;                      // transient intervals in it
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // device time per interval and node of the last chunk (0: none yet), and
    // the nodes of the members that ran it: a chunk's intervals are sized by
    // the nodes its members hold now
    double ms_per_node_interval = 0.0;
    double chunk_nodes = 0.0;

    // a membership change: each interval body is captured again at its
    // next use, and its executable graph updated in place from the capture
    void reset_graphs() {
        for (Replayer &r : rp) r.stale = true;
    }
    void unstage() {
        for (void *p : staged) hfree(hpool, p);
        staged.clear();
    }
    void release() {
        for (Replayer &r : rp) r.reset();
        unstage();
        if (ev_sync) (void)hipEventDestroy(ev_sync);
        ev_sync = nullptr;
        for (void *p : lv_bufs) dfree(pool, p);
        lv_bufs.clear();
        lv.clear();
        dfree(pool, g.d_desc);
        dfree(pool, g.d_work);
        dfree(pool, g.d_gwork);
        hfree(hpool, h_flags);
        hfree(hpool, h_kmax);
        dfree(pool, g.d_kmax);
        dfree(pool, g.d_trec);
        dfree(pool, g.d_twork);
        if (ev_a) (void)hipEventDestroy(ev_a);
        if (ev_b) (void)hipEventDestroy(ev_b);
        ev_a = ev_b = nullptr;
        g = Group{};
        h_flags = m_flags = nullptr;
        h_kmax = m_kmax = nullptr;
        cap_desc = cap_work = cap_gwork = cap_flags = cap_trec = cap_twork = 0;
    }
    ~BatchGroup() { release(); }
    void remove(EngJob *j) {
        members.erase(std::find(members.begin(), members.end(), j));
        dirty = true;
    }
};

// The device state of a model whose job has ended, kept for the model's next
// job (the RC model stays resident: coefficients, maps, multigrid levels and
// GMRES buffers are not built again), while its coefficients are unchanged.
struct Resident {
    const Model *M = nullptr;
    unsigned long stamp = 0;
    std::unique_ptr<DevModel> D;
    double last_use = 0.0;
};

// The read-only device data of a model (maps, block ranges, work-lists,
// heterogeneous coefficients), built once while its coefficients are
// unchanged; every job on the model takes a device model of its own state
// that shares it (dev_clone), so that a job on a model already seen only
// allocates its vectors.
struct ModelConst {
    const Model *M = nullptr;
    unsigned long stamp = 0;
    std::shared_ptr<DevModel> tpl;
    double last_use = 0.0;
};

struct Engine {
    DevPool dpool, hpool;               // declared first: destroyed after the groups that use them
    hs_options_t opt{};
    size_t budget = 0, used = 0;
    int next_id = 0;
    int tick_intervals = 32;            // most transient intervals per chunk
    int tick_iters = 128;               // steady iterations per chunk
    // s_setup: admission, phase changes and results (never waits on a chunk)
    hipStream_t s_setup = nullptr, s_steady[kSteadyStreams] = {}, s_trans[kTransParts] = {}, s_be = nullptr;
    std::deque<EngJob *> queue;
    std::map<int, std::unique_ptr<EngJob>> jobs;
    std::vector<Resident> resident;     // idle device models of ended jobs
    size_t resident_bytes = 0;
    std::vector<ModelConst> consts;     // read-only device data per model
    size_t const_bytes = 0;
    std::vector<EngJob *> downloading;  // finished transients whose sequences are being copied
    hs_engine_stats_t stats{};
    // steady groups by key: kind (0 homogeneous CG, 1 heterogeneous CG,
    // 2 non-symmetric GMRES) + 4 x multigrid levels (0: Jacobi);
    // transient: by kernel form (trans_form), each in parts on the transient
    // streams, [form + kTransForms * part]
    std::map<int, BatchGroup> steady;
    BatchGroup trans[kTransForms * kTransParts];
    // backward-Euler transients: the interval solves, keyed as steady groups
    std::map<int, BatchGroup> be;
    ~Engine() {
        for (hipStream_t s : streams())
            if (s) (void)hipStreamSynchronize(s);
        for (EngJob *j : downloading) {
            hfree(&hpool, j->h_seq);
            (void)hipEventDestroy(j->ev_dl);
        }
        for (auto &kv : jobs) kv.second->D.reset();
        resident.clear();
        consts.clear();
        for (auto &kv : steady) kv.second.release();
        for (auto &kv : be) kv.second.release();
        for (auto &g : trans) g.release();
        for (hipStream_t s : streams())
            if (s) (void)hipStreamDestroy(s);
    }
    std::vector<hipStream_t> streams() const {
        std::vector<hipStream_t> v{s_setup};
        v.insert(v.end(), s_steady, s_steady + kSteadyStreams);
        v.insert(v.end(), s_trans, s_trans + kTransParts);
        v.push_back(s_be);
        return v;
    }
};

static void start_download(Engine &E, EngJob &j, bool seq = false);

// Device bytes dev_build allocates for a model, plus the sequence buffers
// and, for a backward-Euler transient (be), the shifted operator's arrays.
static size_t dev_footprint(const Model &M, int nsteps, bool mg, bool be) {
    const grid_model_t &g = *M.g;
    ProbDesc t{};
    t.nl = g.n_layers;
    t.nr = g.rows;
    t.nc = g.cols;
    set_tiling(t);
    const size_t ncell = (size_t)M.ncell(), nblk = (size_t)g.total_n_blocks + M.nx;
    const size_t vlen = ((ncell + M.nx + 63) / 64) * 64;
    size_t b = vlen * (NB64 * sizeof(double) + NB32 * sizeof(float));
    for (const BlockMap &bm : M.maps)
        b += bm.cell_ptr.size() * 4 + bm.ent_unit.size() * 4 + bm.ent_occ.size() * 8;
    b += (size_t)g.total_n_blocks * (8 + 16 + 4) + 2 * nblk * 8;
    b += (size_t)(t.ntiles + 1) * (kRed * 8 + 8) + ((size_t)g.total_n_blocks / 4 + 2) * 8;
    b += NSCAL * 8 + sizeof(ProbDesc) + 64;
    if (M.detailed) b += ncell * 176 + (size_t)g.n_layers * 40;   // assembly inputs + per-cell coefficients
    b += 2 * (size_t)nsteps * nblk * 8;
    bool adv = false;
    for (auto *m : M.mc) adv |= m != nullptr;
    if (mg)    // coarse levels: 3 FP32 vectors, 4 coefficient arrays; a third of the cells in all
        b += (ncell / 3 + 64 * 8) * 7 * 4 + (size_t)g.n_layers * 4096;
    if (adv)   // GMRES state, partials and one basis (FP32 or FP64)
        b += GM_SIZE * 8 + (size_t)(t.ntiles + 1) * (kGmM + 1) * 8 + (size_t)(kGmM + 1) * vlen * 8;
    if (be && M.detailed) b += ncell * 16;   // per-cell shift and shifted Jacobi diagonal
    if (be && adv)   // the shifted multigrid levels: per-cell shift (FP64) and 1/diag (FP32)
        b += (ncell + ncell / 3 + 64 * 8) * 12 + (ncell / 3 + 64 * 8) * 7 * 4 + (size_t)g.n_layers * 4096;
    // pool size classes round each allocation up by at most an eighth
    return b + b / 8 + (64 << 10);
}

template <typename T> static void ensure_dev(DevPool *pool, T *&p, size_t &cap, size_t n) {
    if (n <= cap) return;
    dfree(pool, p);
    cap = std::max(n, cap * 2);
    p = (T *)dmalloc(pool, cap * sizeof(T));
}

// A copy of the group's set-up through pinned staging on stream s, without
// a host wait (the staging block is returned at the group's next set-up,
// after the chunk that waited for the copy).
static void stage_h2d(BatchGroup &B, void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    void *p = hmalloc(B.hpool, bytes);
    memcpy(p, src, bytes);
    B.staged.push_back(p);
    HIP_OK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s));
}

// Uploads the group's descriptors and work-lists after a membership change
// (on the set-up stream: the group's own stream may carry another group's
// chunk, and this group is between chunks; the group's next chunk waits for
// the copies on the device).  A backward-Euler group (be) takes each
// member's shifted descriptor and shifted multigrid levels.
static void group_sync(Engine &E, BatchGroup &B, bool het, bool trans, bool be = false) {
    if (!B.dirty) return;
    B.unstage();
    struct Timed {
        double &acc;
        double t0 = now_s();
        ~Timed() { acc += now_s() - t0; }
    } timed{E.stats.sync_s};
    hipStream_t s = E.s_setup;
    const Upload staged = [&B](void *dst, const void *src, size_t bytes, hipStream_t st) {
        stage_h2d(B, dst, src, bytes, st);
    };
    const int n = (int)B.members.size();
    std::vector<ProbDesc> desc(n);
    std::vector<WorkItem> w, gw;
    std::vector<double> rec;   // Taylor records (step factors), transient Taylor groups
    int m = 1, K = 1;
    bool v4 = true;
    for (int i = 0; i < n; i++) {
        const DevModel &D = *B.members[i]->D;
        desc[i] = be ? D.be_h : D.h;
        v4 = v4 && vec4_ok(D.h);
        for (int t = 0; t <= D.h.ntiles; t++) w.push_back(WorkItem{i, t});
        for (int u = 0; u < D.h.nblocks; u += kBlock / kWave) gw.push_back(WorkItem{i, u});
        gw.push_back(WorkItem{i, -1});
        if (trans && E.opt.transient == HS_TRANS_TAYLOR) {
            const TaylorPlan p = taylor_plan(D, B.members[i]->dt, E.opt.taylor_tol);
            m = std::max(m, p.m);
            K = std::max(K, p.K);
            rec.resize((size_t)kTrec * n);
            taylor_record(p, &rec[(size_t)kTrec * i]);
        }
    }
    B.reset_graphs();
    ensure_dev(B.pool, B.g.d_desc, B.cap_desc, (size_t)n);
    ensure_dev(B.pool, B.g.d_work, B.cap_work, w.size());
    ensure_dev(B.pool, B.g.d_gwork, B.cap_gwork, gw.size());
    if (!B.g.d_kmax) B.g.d_kmax = (int *)dmalloc(B.pool, 2 * sizeof(int));
    if (!B.h_kmax) {
        B.h_kmax = (int *)hmalloc(B.hpool, 2 * sizeof(int));
        HIP_OK(hipHostGetDevicePointer((void **)&B.m_kmax, B.h_kmax, 0));
    }
    const int zero[2] = {0, 0};
    stage_h2d(B, B.g.d_kmax, zero, sizeof zero, s);
    B.h_kmax[0] = B.h_kmax[1] = 0;
    B.g.kgraph = 0;
    ensure_dev(B.pool, B.g.d_trec, B.cap_trec, (size_t)kTrec * std::max(n, 1));
    if (!rec.empty()) stage_h2d(B, B.g.d_trec, rec.data(), rec.size() * sizeof(double), s);
    if ((size_t)n * 8 > B.cap_flags) {
        hfree(B.hpool, B.h_flags);
        B.cap_flags = std::max((size_t)n * 8, B.cap_flags * 2);
        B.h_flags = (double *)hmalloc(B.hpool, B.cap_flags * sizeof(double));
        HIP_OK(hipHostGetDevicePointer((void **)&B.m_flags, B.h_flags, 0));
    }
    stage_h2d(B, B.g.d_desc, desc.data(), n * sizeof(ProbDesc), s);
    stage_h2d(B, B.g.d_work, w.data(), w.size() * sizeof(WorkItem), s);
    stage_h2d(B, B.g.d_gwork, gw.data(), gw.size() * sizeof(WorkItem), s);
    B.g.nprob = n;
    B.g.nwork = (int)w.size();
    B.g.ngwork = (int)gw.size();
    B.g.stream = B.stream;
    B.g.hetero = het;
    B.g.vec4 = v4;
    B.g.nper = (n > 0 && (int)w.size() == n * (desc[0].ntiles + 1)) ? desc[0].ntiles + 1 : 0;
    for (int i = 0; i < n && B.g.nper; i++)
        if (desc[i].ntiles != desc[0].ntiles) B.g.nper = 0;
    B.form = SeqForm{E.opt.precision, E.opt.transient, m, K};
    // the float4 term sweep's work-list; its tiling timed per size class for
    // FP32 Taylor groups
    if (trans && v4 && E.opt.transient == HS_TRANS_TAYLOR && E.opt.precision == HS_PREC_MIXED) {
        tune_term_tiling(B.g, desc, B.cap_twork, B.pool, s, staged);
    } else {
        std::vector<WorkItem> tw;
        const int per = term_items(desc, tw);
        upload_term_items(B.g, tw, per, B.cap_twork, B.pool, s, staged);
        B.g.tz = term_form(desc);
        B.g.tg = term_groups(desc);
        B.g.zfirst = false;
        B.g.zpow = false;
    }
    // multigrid levels of a steady group (every member has the same count),
    // or of a backward-Euler group of non-symmetric steps
    for (void *p : B.lv_bufs) dfree(B.pool, p);
    B.lv_bufs.clear();
    B.lv.clear();
    const int L = !n ? 0
                  : be ? (B.members[0]->adv ? (int)B.members[0]->D->be_mg.size() : 0)
                       : (!trans && B.members[0]->use_mg) ? (int)B.members[0]->D->mg.size() : 0;
    for (int l = 0; l < L; l++) {
        std::vector<ProbDesc> ld(n);
        std::vector<WorkItem> lw;
        int nlmax = 1;
        for (int i = 0; i < n; i++) {
            const DevModel &D = *B.members[i]->D;
            // level 0 is the model itself, but for the upwind advection of
            // a non-symmetric model's preconditioner; the shifted levels of
            // a backward-Euler step start from the upwind level 0
            ld[i] = be ? D.be_mg[l].h : l == 0 && !D.h.has_adv ? D.h : D.mg[l].h;
            nlmax = std::max(nlmax, ld[i].nl);
            auto ci = column_items(ld[i], i);
            lw.insert(lw.end(), ci.begin(), ci.end());
        }
        ProbDesc *dd = (ProbDesc *)dmalloc(B.pool, n * sizeof(ProbDesc));
        WorkItem *dw = (WorkItem *)dmalloc(B.pool, lw.size() * sizeof(WorkItem));
        B.lv_bufs.push_back(dd);
        B.lv_bufs.push_back(dw);
        stage_h2d(B, dd, ld.data(), n * sizeof(ProbDesc), s);
        stage_h2d(B, dw, lw.data(), lw.size() * sizeof(WorkItem), s);
        B.lv.push_back(MgGroup{dd, dw, (int)lw.size(), nlmax, l > 0 || het || be});
    }
    if (!B.ev_sync) HIP_OK(hipEventCreateWithFlags(&B.ev_sync, hipEventDisableTiming));
    HIP_OK(hipEventRecord(B.ev_sync, s));
    B.synced = true;
    B.dirty = false;
}

// the group's chunk waits on the device for the copies of its last set-up
static void wait_sync(BatchGroup &B) {
    if (!B.synced) return;
    HIP_OK(hipStreamWaitEvent(B.stream, B.ev_sync, 0));
    B.synced = false;
}

// writes the given scalar slots of every member into the pinned flags, and
// for a Taylor group its term counts (after the chunk, on the group's stream)
static void gather(BatchGroup &B, std::initializer_list<int> slots, bool kmax = false) {
    Slots sl{};
    for (int v : slots) sl.s[sl.n++] = v;
    hipLaunchKernelGGL(k_gather, dim3(small_grid(B.g.nprob)), dim3(64), 0, B.g.stream,
                       (const ProbDesc *)B.g.d_desc, B.g.nprob, sl, B.m_flags, kmax ? B.g.d_kmax : nullptr,
                       kmax ? B.m_kmax : nullptr);
}

// A job's device model goes back to the resident set, its sequence buffers
// to the pool.
static void release_device(Engine &E, EngJob &j) {
    E.used -= j.held;
    j.held = 0;
    DevModel &D = *j.D;
    D.release(D.h.seq_in);
    D.release(D.h.seq_out);
    D.h.seq_in = D.h.seq_out = nullptr;
    D.h.seq_one = 0;
    bool shared_data_kept = !D.base;
    for (const ModelConst &c : E.consts) shared_data_kept |= c.tpl == D.base;
    // not kept: coefficients of the job's own r_convec (not the model's), or
    // read-only data that the engine no longer counts against its budget
    if (j.package || !shared_data_kept) {
        j.D.reset();
        return;
    }
    // kept under the stamp of the coefficients it was built from (the model
    // itself may be gone by now: a later model, or this one repopulated,
    // carries another stamp and never matches)
    E.resident_bytes += D.bytes;
    E.resident.push_back(Resident{j.M, j.stamp, std::move(j.D), now_s()});
}

// Zeroes the state of a resident device model for its next job, as a fresh
// build leaves it; the coefficients, maps and multigrid levels stay.
static void dev_reset_state(DevModel &D) {
    hipStream_t s = D.stream;
    ProbDesc &h = D.h;
    for (int b = 0; b < NB64; b++) dzero(h.b64[b], D.vlen * sizeof(double), s);
    for (int b = 0; b < NB32; b++) dzero(h.b32[b], D.vlen * sizeof(float), s);
    dzero(h.lzero, kMaxLayers * sizeof(int), s);
    dzero(h.partial, (size_t)slot_count(h) * kRed * sizeof(double), s);
    dzero(h.counter, sizeof(unsigned), s);
    dzero(h.scal, NSCAL * sizeof(double), s);
    dzero(D.grp.d_kmax, 2 * sizeof(int), s);
    dzero(D.grp.d_trec, kTrec * sizeof(double), s);
    dzero(h.blk_in, D.nblk * sizeof(double), s);
    dzero(h.blk_out, D.nblk * sizeof(double), s);
    for (size_t l = 1; l < D.mg.size(); l++) {
        const ProbDesc &c = D.mg[l].h;
        for (int b : {(int)F_R, (int)F_D0, (int)F_D1}) dzero(c.b32[b], c.vlen * sizeof(float), s);
        dzero(c.partial, (size_t)slot_count(c) * kRed * sizeof(double), s);
    }
    if (h.gm) {
        dzero(h.gm, GM_SIZE * sizeof(double), s);
        dzero(h.gpart, (size_t)(h.ntiles + 1) * (kGmM + 1) * sizeof(double), s);
    }
    if (h.gmv32) dzero(h.gmv32, (size_t)(kGmM + 1) * D.vlen * sizeof(float), s);
    if (h.gmv64) dzero(h.gmv64, (size_t)(kGmM + 1) * D.vlen * sizeof(double), s);
    h2d(D.d_desc, &h, sizeof(ProbDesc), s);
}

// The resident device model of M's current coefficients, if any; resident
// models of a changed (or deleted and replaced) model are dropped.
static DevModel *take_resident(Engine &E, const Model &M) {
    DevModel *found = nullptr;
    for (size_t k = 0; k < E.resident.size();) {
        Resident &r = E.resident[k];
        if (r.M != &M || (r.stamp == M.stamp && found)) {
            k++;
            continue;
        }
        // M's: the first of its current coefficients is taken, stale ones dropped
        E.resident_bytes -= r.D->bytes;
        if (r.stamp == M.stamp) found = r.D.release();
        E.resident.erase(E.resident.begin() + k);
    }
    return found;
}

// Drops the least recently used resident models, then the least recently
// used read-only model data, until `need` more bytes fit the budget beside
// the resident jobs (their blocks go to the pool; a model's read-only data
// stays on the device while a job's device model uses it).
static void evict_resident(Engine &E, size_t need) {
    while (!E.resident.empty() && E.used + E.resident_bytes + E.const_bytes + need > E.budget) {
        size_t k = 0;
        for (size_t i = 1; i < E.resident.size(); i++)
            if (E.resident[i].last_use < E.resident[k].last_use) k = i;
        E.resident_bytes -= E.resident[k].D->bytes;
        E.resident.erase(E.resident.begin() + k);
    }
    while (!E.consts.empty() && E.used + E.resident_bytes + E.const_bytes + need > E.budget) {
        size_t k = 0;
        for (size_t i = 1; i < E.consts.size(); i++)
            if (E.consts[i].last_use < E.consts[k].last_use) k = i;
        E.const_bytes -= E.consts[k].tpl->bytes;
        E.consts.erase(E.consts.begin() + k);
    }
}

// The read-only device data of M's current coefficients, built on first use
// (entries of M's earlier coefficients are dropped).
static std::shared_ptr<DevModel> model_const(Engine &E, Model &M) {
    for (size_t k = 0; k < E.consts.size();) {
        ModelConst &c = E.consts[k];
        if (c.M == &M && c.stamp != M.stamp) {
            E.const_bytes -= c.tpl->bytes;
            E.consts.erase(E.consts.begin() + k);
            continue;
        }
        if (c.M == &M) {
            c.last_use = now_s();
            return c.tpl;
        }
        k++;
    }
    std::shared_ptr<DevModel> t(dev_build(M, E.s_setup, &E.dpool, &E.hpool, false));
    E.consts.push_back(ModelConst{&M, M.stamp, t, now_s()});
    E.const_bytes += t->bytes;
    return t;
}

static int steady_key(const EngJob &j) {
    return (j.adv ? 2 : j.hetero ? 1 : 0) + 4 * (j.use_mg ? (int)j.D->mg.size() : 0);
}

static int trans_form(const EngJob &j) { return j.hetero ? 2 : vec4_ok(j.D->h) ? 0 : 1; }

// ---- backward Euler (RCutil.c:268-428 as compute_temp_grid calls it,
// temperature_grid.c:3632-3662): interval k solves
//   (C/h + G) y' = C y / h + P_k + sources
// from y, in place (y stays in B_Y0: the step never flips the state's
// parity), with the solver of the single-model API: FP64 Jacobi CG for
// symmetric G, GMRES right-preconditioned by the shifted upwind multigrid
// V-cycle inside FP64 refinement for coolant advection; relative residual
// 1e-2 x the steady tolerance.
static double be_rtol(const Engine &E) { return 1e-2 * E.opt.steady_rtol; }

// Starts the solve of the job's current interval (S_STEP): its power, the
// right-hand side from the state, the solver's scalars.  On the set-up
// stream without a host wait: the job's group is between chunks, and its
// next chunk waits for the set-up stream on the device.
static void be_begin(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    run_b2g(D.grp, V_POWER, B_RHS, 1, 1);
    if (j.hetero)
        launch_named("k_be_rhs", k_be_rhs<true>, D.grp.nwork, D.stream, (const ProbDesc *)D.d_desc,
                     (const WorkItem *)D.d_work, (int)B_Y0, (int)B_RHS, (int)B_T1, 1.0 / j.dt);
    else
        launch_named("k_be_rhs", k_be_rhs<false>, D.grp.nwork, D.stream, (const ProbDesc *)D.d_desc,
                     (const WorkItem *)D.d_work, (int)B_Y0, (int)B_RHS, (int)B_T1, 1.0 / j.dt);
    if (j.adv) {
        set_scal(D, kMixedStart);
    } else {
        set_scal(D, {{S_DONE, 0.0}});   // the CG start skips a problem still marked done
        const double r = be_rtol(E);
        if (j.hetero) jpcg64_start<true>(D.be_grp, r * r, E.opt.max_iters, B_Y0, B_T1);
        else jpcg64_start<false>(D.be_grp, r * r, E.opt.max_iters, B_Y0, B_T1);
    }
}

static int be_key(const EngJob &j) { return (j.adv ? 2 : j.hetero ? 1 : 0) + 4 * (j.adv ? (int)j.D->be_mg.size() : 0); }

// The shifted operator of the job's step h (once per job: h is the job's
// interval), then its first interval's solve; the job joins its group.
static void start_be(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    if (j.adv) ensure_gmres(D, true);   // the shifted descriptor copies its buffers
    build_be(D, j.dt, j.adv ? (int)B_Y0 : -1, j.adv ? (int)B_T1 : -1);
    if (j.adv) be_levels(D, 1.0 / j.dt);
    // device memory held, the shifted operator's arrays included
    E.used += D.bytes - j.held;
    j.held = D.bytes;
    be_begin(E, j);
    BatchGroup &B = E.be[be_key(j)];
    B.members.push_back(&j);
    B.dirty = true;
}

static void start_transient(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    auto kv = seq_slots(j.nsteps, j.dt);
    kv.push_back({S_YPAR, 0.0});
    if (E.opt.transient == HS_TRANS_BE) {
        set_scal(D, kv);
        j.phase = J_TRANS;
        j.step_seen = 0;
        start_be(E, j);
        return;
    }
    if (E.opt.transient == HS_TRANS_TAYLOR) {
        const TaylorPlan p = taylor_plan(D, j.dt, E.opt.taylor_tol);
        const auto t = taylor_slots(p, D.rho);
        kv.insert(kv.end(), t.begin(), t.end());
        j.kplan = p.K;
    }
    set_scal(D, kv);
    j.phase = J_TRANS;
    j.step_seen = 0;
    // the part with the fewest members; one part only while a kernel timer
    // runs (launches of the parts would overlap in its event spans)
    const int parts = g_timer.cap > 0 ? 1 : kTransParts;
    const int form = trans_form(j);
    int best = form;
    for (int q = 1; q < parts; q++)
        if (E.trans[form + kTransForms * q].members.size() < E.trans[best].members.size())
            best = form + kTransForms * q;
    BatchGroup &B = E.trans[best];
    B.members.push_back(&j);
    B.dirty = true;
}

// Starts a steady solve from the state in B_STEADY (the ambient fill of the
// admission, or the last solve of a leakage fixed point), with the solver
// of the single-model API for the engine's form, and puts the job into its
// steady group.
static void start_steady(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    if (j.adv) {
        // non-symmetric: GMRES with the upwind multigrid preconditioner
        // in every form (Jacobi does not converge at channel Peclet numbers)
        ensure_gmres(D, true);
        build_mg(D);
        set_scal(D, kMixedStart);
    } else if (j.use_mg) {
        build_mg(D);
        set_scal(D, kMixedStart);
    } else if (E.opt.precision == HS_PREC_MIXED) {
        set_scal(D, kMixedStart);
    } else {
        set_scal(D, {{S_DONE, 0.0}});
        const double rtol2 = E.opt.steady_rtol * E.opt.steady_rtol;
        if (j.hetero) jpcg64_start<true>(D.grp, rtol2, E.opt.max_iters);
        else jpcg64_start<false>(D.grp, rtol2, E.opt.max_iters);
    }
    // no host wait: joining the group changes its membership, and its next
    // chunk waits on the device for the set-up stream's work so far
    j.phase = J_STEADY;
    BatchGroup &B = E.steady[steady_key(j)];
    B.members.push_back(&j);
    B.dirty = true;
}

// A job that cannot finish: its device state is released and the reason
// kept for hs_engine_poll.
static void fail_job(Engine &E, EngJob &j, const std::string &why) {
    j.error = why;
    j.phase = J_FAILED;
    j.t_done = now_s();
    release_device(E, j);
}

static void admit(Engine &E, EngJob &j) {
    Model &M = *j.M;
    j.t_admit = now_s();
    j.stamp = M.stamp;
    if (DevModel *R = take_resident(E, M)) {
        j.D.reset(R);
        j.resident = true;
        dev_reset_state(*R);
    } else {
        j.D.reset(dev_clone(model_const(E, M), E.s_setup, &E.dpool, &E.hpool));
    }
    DevModel &D = *j.D;
    const size_t nb = (size_t)D.nblk;
    if (j.nsteps > 0) {
        // interval powers: one row for all, one per interval, or the steady row first
        const int first = j.rows == j.nsteps + 1 ? 1 : 0;
        const size_t in_rows = j.rows == 1 ? 1 : (size_t)j.nsteps;
        D.h.seq_in = D.alloc<double>(in_rows * nb);
        D.h.seq_out = D.alloc<double>((size_t)j.nsteps * nb);
        D.h.seq_one = j.rows == 1 ? 1 : 0;
        h2d(D.h.seq_in, j.power.data() + first * nb, in_rows * nb * sizeof(double), D.stream);
        h2d(D.d_desc, &D.h, sizeof(ProbDesc), D.stream);
    }
    j.held = D.bytes;
    E.used += j.held;
    if (j.package) {
        double rz;
        package_RC_t pk;
        sink_resistances(j.sink, j.r_convec, rz, pk);
        dev_set_convec(D, j.sink.hsidx, rz, pk);
    }
    if (j.init.empty()) {
        if (j.leak) {
            // the first solve's power: the leakage at zero block temperatures
            std::vector<double> zero(nb, 0.0);
            j.leak_old.assign(nb, 0.0);
            j.leak_power.assign(nb, 0.0);
            leakage_step(j.lb, j.power.data(), zero.data(), j.leak_power.data(), j.leak_old.data());
            upload_blocks(D, j.leak_power.data());
        } else {
            upload_blocks(D, j.power.data());
        }
        run_b2g(D.grp, V_POWER, B_RHS, 1);
        const int nbk = (int)((D.vlen + kBlock - 1) / kBlock);
        launch(k_fill, nbk, D.stream, D.h.b64[B_STEADY], D.vlen, M.g->config.ambient);
        j.use_mg = j.adv || E.opt.steady_solver == HS_STEADY_MGPCG;
        start_steady(E, j);
    } else {
        upload_blocks(D, j.init.data());
        run_b2g(D.grp, V_TEMP, B_Y0, 0);
        HIP_OK(hipStreamSynchronize(D.stream));
        start_transient(E, j);
    }
    // device memory held, multigrid levels and GMRES state included
    E.used += D.bytes - j.held;
    j.held = D.bytes;
    j.setup_s = now_s() - j.t_admit;
}

// A finished steady solve: block temperatures, then either the result or
// the start of the job's transient sequence from them.
static void steady_done(Engine &E, EngJob &j, const double *flags) {
    DevModel &D = *j.D;
    const bool refined = E.opt.precision == HS_PREC_MIXED || j.adv || E.opt.steady_solver == HS_STEADY_MGPCG;
    j.steady_iters = (int)flags[1];
    j.refinements = refined ? (int)flags[2] : 0;
    run_g2b(D.grp, B_STEADY);
    if (j.leak) {
        // one step of the temperature-leakage fixed point (temperature.c:
        // 708-746): converged, failed, or solved again from this state
        std::vector<double> t(D.nblk);
        download_blocks(D, t.data());
        const double dmax = leakage_rise(j.lb, t.data(), j.leak_old.data());
        const int it = j.leak_iter++;
        if (!(dmax < kLeakTol)) {
            if (dmax > kLeakTempHigh && it > 0) {
                fail_job(E, j, "temperature is too high, possible thermal runaway. Double-check power inputs and "
                               "package settings.\n");
                return;
            }
            if (it + 1 > kLeakMaxIter) {
                fail_job(E, j, "too many iterations before temperature-leakage convergence -- possible thermal "
                               "runaway\n");
                return;
            }
            leakage_step(j.lb, j.power.data(), t.data(), j.leak_power.data(), j.leak_old.data());
            upload_blocks(D, j.leak_power.data());
            run_b2g(D.grp, V_POWER, B_RHS, 1);
            start_steady(E, j);
            return;
        }
    }
    if (j.natural) {
        // one step of the natural-convection fixed point: r_convec at the
        // average sink temperature (+ kSmallForConvec); solved again, from
        // this state, until r_convec moves by less than kNaturalConvecTol
        std::vector<double> t(D.nblk);
        download_blocks(D, t.data());
        j.leak_iter++;
        double r = 0.0, sink = 0.0;
        try {
            sink = sink_temp(j.sink, t.data()) + kSmallForConvec;
            convection_t p{};
            calc_natural_convec(&p, &j.pkg, &j.sink.c, sink);
            r = p.r_th;
        } catch (const Error &e) {
            fail_job(E, j, e.msg);
            return;
        }
        const double r_old = j.r_convec;
        j.r_convec = r;
        if (sink > kMaxSinkTemp) {
            fail_job(E, j, "too high power for a natural convection package -- possible thermal runaway\n");
            return;
        }
        if (!(std::fabs(r - r_old) < kNaturalConvecTol)) {
            double rz;
            package_RC_t pk;
            sink_resistances(j.sink, r, rz, pk);
            dev_set_convec(D, j.sink.hsidx, rz, pk);
            upload_blocks(D, j.power.data());
            run_b2g(D.grp, V_POWER, B_RHS, 1);
            start_steady(E, j);
            return;
        }
        j.result = t;
        j.phase = J_DONE;
        j.t_done = now_s();
        release_device(E, j);
        return;
    }
    if (j.nsteps == 0) {
        start_download(E, j);   // done when the block vector has arrived
        return;
    }
    run_b2g(D.grp, V_TEMP, B_Y0, 0, 2);
    start_transient(E, j);
}

// A finished job's block vector (or sequence) goes to pinned host memory
// without a host wait; the job is done once the copy has completed
// (finish_downloads), and its device state is released then.
static void start_download(Engine &E, EngJob &j, bool seq) {
    DevModel &D = *j.D;
    j.dl_seq = seq;
    j.dl_n = seq ? (size_t)j.nsteps * D.nblk : (size_t)D.nblk;
    const size_t bytes = j.dl_n * sizeof(double);
    j.h_seq = (double *)hmalloc(&E.hpool, bytes);
    HIP_OK(hipMemcpyAsync(j.h_seq, seq ? D.h.seq_out : D.h.blk_out, bytes, hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipEventCreateWithFlags(&j.ev_dl, hipEventDisableTiming));
    HIP_OK(hipEventRecord(j.ev_dl, D.stream));
    E.downloading.push_back(&j);
}

static void finish_downloads(Engine &E) {
    for (size_t k = 0; k < E.downloading.size();) {
        EngJob &j = *E.downloading[k];
        const hipError_t r = hipEventQuery(j.ev_dl);
        if (r == hipErrorNotReady) {
            k++;
            continue;
        }
        HIP_OK(r);
        if (j.dl_seq) {
            j.seq.assign(j.h_seq, j.h_seq + j.dl_n);
            j.result.assign(j.seq.end() - j.D->nblk, j.seq.end());
        } else {
            j.result.assign(j.h_seq, j.h_seq + j.dl_n);
        }
        hfree(&E.hpool, j.h_seq);
        HIP_OK(hipEventDestroy(j.ev_dl));
        j.h_seq = nullptr;
        j.ev_dl = nullptr;
        j.phase = J_DONE;
        j.t_done = now_s();
        release_device(E, j);
        E.downloading.erase(E.downloading.begin() + k);
    }
}

// A finished transient's sequence, copied to the host straight into the
// job's result (a sequence is megabytes: staging it through pinned memory
// per job, from an engine's fresh pool, costs more than the copy)
static void finish_sequence(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    const size_t n = (size_t)j.nsteps * D.nblk;
    j.seq.resize(n);
    d2h(j.seq.data(), D.h.seq_out, n * sizeof(double), D.stream);
    j.result.assign(j.seq.end() - D.nblk, j.seq.end());
    j.phase = J_DONE;
    j.t_done = now_s();
    release_device(E, j);
}

static void trans_done(Engine &E, EngJob &j, const double *flags) {
    j.rk4_calls = E.opt.transient == HS_TRANS_RK4 ? (int)flags[1] : 0;
    j.taylor_sweeps = E.opt.transient == HS_TRANS_TAYLOR ? (long)flags[2] : 0;
    j.taylor_node_bytes = E.opt.transient == HS_TRANS_TAYLOR ? flags[3] : 0.0;
    j.holds = (int)flags[4];
    finish_sequence(E, j);
}

// A converged interval solve of a backward-Euler job (flags: S_DONE,
// iterations, refinements, r.r, b.b): its block temperatures go into the
// sequence, and the job starts its next interval or finishes.  Returns
// whether the job leaves its group.
static bool be_step_done(Engine &E, EngJob &j, const double *flags) {
    DevModel &D = *j.D;
    j.be_iters += (long)flags[1];
    // a step stopped at its iteration limits far above its tolerance
    const double rel = flags[4] > 0 ? std::sqrt(flags[3] / flags[4]) : 0.0;
    if (!(rel <= std::max(100.0 * be_rtol(E), 1e-9))) {
        char msg[160];
        snprintf(msg, sizeof msg, "backward Euler step did not converge (relative residual %.2e after %d iterations)\n",
                 rel, (int)flags[1]);
        fail_job(E, j, msg);
        return true;
    }
    run_g2b(D.grp, B_Y0, 1);
    j.step_seen++;
    set_scal(D, {{S_STEP, (double)j.step_seen}});
    if (j.step_seen < j.nsteps) {
        be_begin(E, j);
        return false;
    }
    finish_sequence(E, j);
    return true;
}

// One round of a steady group: the unit that a chunk replays (captured once
// per membership as a graph).
template <bool HET>
static void steady_round(Engine &E, const Group &G, bool adv, const std::vector<MgGroup> &lv) {
    const double rtol2 = E.opt.steady_rtol * E.opt.steady_rtol;
    if (!lv.empty() && adv) {   // multigrid-preconditioned GMRES cycles
        gm_round_mg<true>(G, lv, rtol2, 1e-8, kGmM);
    } else if (!lv.empty()) {   // multigrid-preconditioned CG rounds
        mixed_round_mg<HET>(G, lv, rtol2, E.opt.max_iters, kMgRoundIters);
    } else if (adv) {   // GMRES cycles inside FP64 refinement
        if (E.opt.precision == HS_PREC_MIXED) gm_round<float, true>(G, rtol2, 1e-8, kGmM);
        else gm_round<double, true>(G, rtol2, 1e-12, kGmM);
    } else if (E.opt.precision == HS_PREC_MIXED) {
        mixed_round<HET>(G, rtol2, E.opt.max_iters, 64);
    } else {
        cg_iterations<double, HET>(G, 64, B_STEADY, B_R, B_Z, B_P0, B_Q, S_DONE, S_TOL2, E.opt.max_iters);
    }
}

// One round of a backward-Euler group, as the single-model step iterates
// (kind 2: non-symmetric GMRES cycles with the shifted multigrid levels)
template <bool HET>
static void be_round(Engine &E, const Group &G, int kind, const std::vector<MgGroup> &lv) {
    const double r = be_rtol(E);
    if (kind == 2) gm_round_mg<true>(G, lv, r * r, 1e-8, kGmM);
    else cg_iterations<double, HET>(G, 64, B_Y0, B_R, B_Z, B_P0, B_Q, S_DONE, S_TOL2, E.opt.max_iters);
}

// rounds of a chunk: one per 64 of the steady iterations per chunk, one per
// kGmM for GMRES cycles without multigrid, and for multigrid-preconditioned
// CG one per kMgRoundIters + 28 (a V-cycle costs several plain iterations)
static int steady_rounds(const Engine &E, bool adv, bool mg) {
    if (adv && !mg) return std::max(1, E.tick_iters / kGmM);
    if (mg && !adv) return std::max(1, E.tick_iters / (kMgRoundIters + 28));
    return std::max(1, E.tick_iters / 64);
}

static void group_events(BatchGroup &B) {
    if (!B.ev_a) HIP_OK(hipEventCreate(&B.ev_a));
    if (!B.ev_b) HIP_OK(hipEventCreate(&B.ev_b));
}

// one chunk of every group that is between chunks
static void launch_chunks(Engine &E) {
    // steady streams that a chunk still runs on
    bool taken[kSteadyStreams] = {};
    for (auto &kv : E.steady)
        for (int k = 0; k < kSteadyStreams; k++)
            if (kv.second.busy && kv.second.stream == E.s_steady[k]) taken[k] = true;
    for (auto &kv : E.steady) {
        BatchGroup &B = kv.second;
        if (B.busy || B.members.empty()) continue;
        const int kind = kv.first % 4;
        // a stream of its own, so that a group's chunk never queues behind
        // another group's (all taken: the group waits for the next tick)
        int k = 0;
        while (k < kSteadyStreams && taken[k]) k++;
        if (k == kSteadyStreams) continue;
        taken[k] = true;
        B.pool = &E.dpool;
        B.hpool = &E.hpool;
        B.stream = E.s_steady[k];
        group_sync(E, B, kind > 0, false);
        B.g.stream = B.stream;   // the launches of the chunk (and its capture) on the chosen stream
        group_events(B);
        wait_sync(B);
        HIP_OK(hipEventRecord(B.ev_a, B.stream));
        const int rounds = steady_rounds(E, kind == 2, !B.lv.empty());
        B.rp[0].run(B.stream, rounds, [&] {
            if (kind) steady_round<true>(E, B.g, kind == 2, B.lv);
            else steady_round<false>(E, B.g, false, B.lv);
        });
        const bool refined = E.opt.precision == HS_PREC_MIXED || kind == 2 || !B.lv.empty();
        gather(B, {S_DONE, refined ? S_TITERS : S_ITERS, S_NREFINE});
        HIP_OK(hipEventRecord(B.ev_b, B.stream));
        E.stats.steady_chunks++;
        B.ran = B.g.nprob;
        B.busy = true;
    }
    // backward-Euler solves on their own stream
    for (auto &kv : E.be) {
        BatchGroup &B = kv.second;
        if (B.busy || B.members.empty()) continue;
        const int kind = kv.first % 4;
        B.pool = &E.dpool;
        B.hpool = &E.hpool;
        B.stream = E.s_be;
        group_sync(E, B, kind > 0, false, true);
        group_events(B);
        wait_sync(B);
        HIP_OK(hipEventRecord(B.ev_a, B.stream));
        B.rp[0].run(B.stream, steady_rounds(E, kind == 2, !B.lv.empty()), [&] {
            if (kind) be_round<true>(E, B.g, kind, B.lv);
            else be_round<false>(E, B.g, kind, B.lv);
        });
        gather(B, {S_DONE, kind == 2 ? S_TITERS : S_ITERS, S_NREFINE, S_RR, S_BB});
        HIP_OK(hipEventRecord(B.ev_b, B.stream));
        E.stats.transient_chunks++;
        B.ran = B.g.nprob;
        B.busy = true;
    }
    // Steady solves go first: a steady round is a long chain of small,
    // latency-bound kernels, each of which would wait for compute units
    // behind the transient sweeps.  While a steady group has members, a
    // transient chunk starts only when no steady chunk runs, and lasts at
    // most kSliceYieldMs, so that steady work arriving meanwhile waits little.
    bool steady_resident = false, steady_busy = false;
    for (auto &kv : E.steady) {
        steady_resident |= !kv.second.members.empty();
        steady_busy |= kv.second.busy;
    }
    for (int h = 0; h < kTransForms * kTransParts; h++) {
        BatchGroup &B = E.trans[h];
        if (B.busy || B.members.empty() || steady_busy) continue;
        B.pool = &E.dpool;
        B.hpool = &E.hpool;
        B.stream = E.s_trans[h / kTransForms];
        group_sync(E, B, h % kTransForms == 2, true);
        group_events(B);
        wait_sync(B);
        // the intervals left to the member furthest from its end, within the
        // tick's intervals and the device-time slice
        int rem = 1;
        for (const EngJob *j : B.members) rem = std::max(rem, j->nsteps - j->step_seen);
        int n = std::min(E.tick_intervals, rem);
        const double slice = steady_resident ? kSliceYieldMs : kSliceMs;
        double nodes = 0.0;
        for (const EngJob *j : B.members) nodes += (double)j->D->vlen;
        // a group's first chunk runs a few intervals to time them
        if (B.ms_per_node_interval > 0)
            n = std::min(n, std::max(1, (int)(slice / (B.ms_per_node_interval * nodes))));
        else
            n = std::min(n, kFirstChunk);
        B.chunk = n;
        B.chunk_nodes = nodes;
        const int kl = B.g.kgraph > 0 ? std::min(B.g.kgraph, B.form.K) : B.form.K;
        for (EngJob *j : B.members)
            if (j->kplan && kl < j->kplan && (j->kmin == 0 || kl < j->kmin)) j->kmin = kl;
        HIP_OK(hipEventRecord(B.ev_a, B.stream));
        B.rp[B.g.kgraph].run(B.stream, n, [&] { seq_body(B.g, B.form); });
        // with the term counts of this chunk (the largest since cleared starts again)
        gather(B, {S_STEP, S_TCALLS, S_TTERMS, S_TBYTES, S_NHOLD}, B.form.transient == HS_TRANS_TAYLOR);
        HIP_OK(hipEventRecord(B.ev_b, B.stream));
        E.stats.transient_chunks++;
        B.ran = B.g.nprob;
        B.busy = true;
    }
}

static bool chunk_done(BatchGroup &B) {
    if (!B.busy) return false;
    const hipError_t r = hipEventQuery(B.ev_b);
    if (r == hipErrorNotReady) return false;
    HIP_OK(r);
    return true;
}

// waits until some group's chunk has ended, or kWaitMs
static void wait_chunk(Engine &E) {
    const double t0 = now_s();
    for (;;) {
        bool busy = false, done = false;
        for (auto &kv : E.steady) {
            busy |= kv.second.busy;
            done |= chunk_done(kv.second);
        }
        for (auto &B : E.trans) {
            busy |= B.busy;
            done |= chunk_done(B);
        }
        for (auto &kv : E.be) {
            busy |= kv.second.busy;
            done |= chunk_done(kv.second);
        }
        for (EngJob *j : E.downloading) {
            busy = true;
            done |= hipEventQuery(j->ev_dl) != hipErrorNotReady;
        }
        if (!busy || done || now_s() - t0 > kWaitMs * 1e-3) return;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// results of every group whose chunk has ended: finished jobs leave their
// groups (a steady solve with intervals to follow joins a transient group)
static void harvest(Engine &E) {
    finish_downloads(E);
    for (auto &kv : E.be) {
        BatchGroup &B = kv.second;
        if (!chunk_done(B)) continue;
        B.busy = false;
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, B.ev_a, B.ev_b));
        E.stats.transient_chunk_s += 1e-3 * ms;
        std::vector<EngJob *> fin;
        bool begun = false;   // a member started its next interval on the set-up stream
        for (int i = 0; i < B.ran; i++)
            if (B.h_flags[i * 8] != 0.0) {
                if (be_step_done(E, *B.members[i], B.h_flags + i * 8)) fin.push_back(B.members[i]);
                else begun = true;
            }
        for (EngJob *j : fin) B.remove(j);
        if (begun) {   // the group's next chunk waits for those starts on the device
            if (!B.ev_sync) HIP_OK(hipEventCreateWithFlags(&B.ev_sync, hipEventDisableTiming));
            HIP_OK(hipEventRecord(B.ev_sync, E.s_setup));
            B.synced = true;
        }
    }
    for (auto &kv : E.steady) {
        BatchGroup &B = kv.second;
        if (!chunk_done(B)) continue;
        B.busy = false;
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, B.ev_a, B.ev_b));
        E.stats.steady_chunk_s += 1e-3 * ms;
        std::vector<std::pair<EngJob *, std::vector<double>>> fin;
        for (int i = 0; i < B.ran; i++)
            if (B.h_flags[i * 8] != 0.0)
                fin.push_back({B.members[i], std::vector<double>(B.h_flags + i * 8, B.h_flags + i * 8 + 8)});
        for (auto &f : fin) {
            B.remove(f.first);
            steady_done(E, *f.first, f.second.data());
        }
    }
    for (int h = 0; h < kTransForms * kTransParts; h++) {
        BatchGroup &B = E.trans[h];
        if (!chunk_done(B)) continue;
        B.busy = false;
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, B.ev_a, B.ev_b));
        B.ms_per_node_interval = (double)ms / std::max(1, B.chunk) / std::max(1.0, B.chunk_nodes);
        E.stats.transient_chunk_s += 1e-3 * ms;
        // Taylor groups of one sub-step per interval launch the Horner steps
        // that the last chunk's largest term count needs instead of every
        // step of the plan, so no launch skips every problem.  A problem that
        // needs more is held (its interval not advanced) until the next chunk
        // raises the count; each count's interval body is captured once.
        if (B.form.transient == HS_TRANS_TAYLOR && B.form.m == 1 && B.h_kmax[1] > 0) {
            const int want = std::min(B.form.K, std::max(2, B.h_kmax[1]));
            B.g.kgraph = want >= B.form.K ? 0 : want;
        }
        std::vector<std::pair<EngJob *, std::vector<double>>> fin;
        for (int i = 0; i < B.ran; i++) {
            EngJob *j = B.members[i];
            const double *f = B.h_flags + i * 8;
            j->step_seen = (int)f[0];
            j->holds = (int)f[4];
            if (f[0] >= (double)j->nsteps) fin.push_back({j, std::vector<double>(f, f + 8)});
        }
        for (auto &f : fin) {
            B.remove(f.first);
            trans_done(E, *f.first, f.second.data());
        }
    }
}

// ---- entry points (wrapped by the C ABI) -----------------------------------------
Engine *eng_create(int precision, int transient) {
    std::unique_ptr<Engine> E(new Engine());
    E->hpool.host = true;
    E->opt.precision = precision;
    // the mixed (served) form solves steady states with multigrid-preconditioned CG
    E->opt.steady_solver = precision == HS_PREC_MIXED ? HS_STEADY_MGPCG : HS_STEADY_JPCG;
    E->opt.transient = transient;
    E->opt.steady_rtol = 1e-12;
    E->opt.max_iters = 200000;
    E->opt.taylor_tol = 1e-7;
    E->opt.steady_warm_start = 0;
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    E->budget = (size_t)(0.8 * (double)fr);
    HIP_OK(hipStreamCreateWithFlags(&E->s_setup, hipStreamNonBlocking));
    for (auto &s : E->s_steady) HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &s : E->s_trans) HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&E->s_be, hipStreamNonBlocking));
    return E.release();
}

void eng_destroy(Engine *E) { delete E; }

int eng_active(Engine &E) {
    int n = (int)E.queue.size();
    for (auto &kv : E.steady) n += (int)kv.second.members.size();
    for (auto &g : E.trans) n += (int)g.members.size();
    for (auto &kv : E.be) n += (int)kv.second.members.size();
    return n + (int)E.downloading.size();
}

void eng_set_options(Engine &E, const hs_options_t &o) {
    if (eng_active(E)) fatal("engine options can change only while no job is queued or running\n");
    E.opt = o;
}

void eng_set_tick(Engine &E, int intervals, int iters) {
    if (intervals <= 0 || iters <= 0) fatal("tick sizes must be positive\n");
    E.tick_intervals = intervals;
    E.tick_iters = ((iters + 63) / 64) * 64;
}

void eng_set_budget(Engine &E, double bytes) {
    if (!(bytes > 0)) fatal("engine budget must be positive\n");
    E.budget = (size_t)bytes;
}

int eng_submit(Engine &E, Model &M, const double *power, int rows, const double *init, int nsteps, double dt) {
    grid_model_t &g = *M.g;
    if (!g.r_ready) fatal("hs_engine_submit: R model not ready\n");
    if (nsteps < 0) fatal("hs_engine_submit: negative n_steps\n");
    if (nsteps > 0 && (!g.c_ready || !(dt > 0))) fatal("hs_engine_submit: transient job needs the C model and dt > 0\n");
    if (!power) fatal("hs_engine_submit: no power\n");
    if (!(rows == 1 || (nsteps > 0 && (rows == nsteps || rows == nsteps + 1))))
        fatal("hs_engine_submit: power rows must be 1, n_steps or n_steps + 1\n");
    if (nsteps == 0 && init) fatal("hs_engine_submit: a steady job takes no start temperatures\n");
    bool adv = false;
    for (auto *m : M.mc) adv |= m != nullptr;
    if (g.n_layers > kMaxLayers) fatal("too many layers for the device model\n");
    std::unique_ptr<EngJob> j(new EngJob());
    j->id = E.next_id++;
    j->M = &M;
    j->nblk = g.total_n_blocks + M.nx;
    j->rows = rows;
    j->power.assign(power, power + (size_t)rows * j->nblk);
    // the package-node entries of a power vector carry no power (set_internal_power_grid)
    for (int r = 0; r < rows; r++)
        for (int k = 0; k < M.nx; k++) j->power[(size_t)r * j->nblk + g.total_n_blocks + k] = 0.0;
    if (init) j->init.assign(init, init + j->nblk);
    j->nsteps = nsteps;
    j->dt = dt;
    j->hetero = M.detailed;
    j->adv = adv;
    if (g.config.leakage_used && !init) {
        j->leak = true;
        j->lb = leak_blocks(M);
    }
    j->footprint = dev_footprint(M, nsteps, !init && (E.opt.steady_solver == HS_STEADY_MGPCG || adv),
                                 nsteps > 0 && E.opt.transient == HS_TRANS_BE);
    if (j->footprint > E.budget)
        fatal("hs_engine_submit: the job needs " + std::to_string(j->footprint) +
              " bytes of device memory, above the engine budget of " + std::to_string(E.budget) + "\n");
    j->t_submit = now_s();
    E.queue.push_back(j.get());
    const int id = j->id;
    E.jobs[id] = std::move(j);
    return id;
}

// A steady job under the package model: r_convec from the fan and sink
// (forced convection), or the natural-convection fixed point starting at
// the ambient temperature + kSmallForConvec (hotspot.c:447-455, :667-682).
int eng_submit_package(Engine &E, Model &M, const double *power, const package_config_t &pkg) {
    if (M.detailed) fatal("hs_engine_submit_package: a detailed-3D model's convection cannot change on the device\n");
    for (auto *m : M.mc)
        if (m) fatal("hs_engine_submit_package: a microchannel model's convection cannot change on the device\n");
    if (M.g->config.leakage_used) fatal("hs_engine_submit_package: the model uses leakage\n");
    const int id = eng_submit(E, M, power, 1, nullptr, 0, 0.0);
    EngJob &j = *E.jobs[id];
    j.package = true;
    j.natural = pkg.natural_convec != 0;
    j.pkg = pkg;
    j.sink = sink_model(M);
    convection_t p{};
    if (j.natural) calc_natural_convec(&p, &j.pkg, &j.sink.c, j.sink.c.ambient + kSmallForConvec);
    else calculate_flow(&p, &j.pkg, &j.sink.c);
    j.r_convec = p.r_th;
    if (!(j.r_convec > 0)) {
        E.queue.pop_back();
        E.jobs.erase(id);
        fatal("hs_engine_submit_package: the package model gives no positive r_convec\n");
    }
    return id;
}

// admission: first come, first served while the footprint fits
int eng_admit(Engine &E) {
    while (!E.queue.empty()) {
        EngJob &j = *E.queue.front();
        if (E.used + j.footprint > E.budget && E.used > 0) break;
        // resident models and the pool's idle blocks count against the
        // budget until dropped or returned to the runtime
        evict_resident(E, j.footprint);
        if (E.used + E.resident_bytes + E.const_bytes + E.dpool.idle_bytes + j.footprint > E.budget) E.dpool.trim();
        E.queue.pop_front();
        const double t0 = now_s();
        admit(E, j);
        E.stats.admit_s += now_s() - t0;
        E.stats.resident_hits += j.resident ? 1 : 0;
    }
    return (int)E.queue.size();
}

int eng_tick(Engine &E) {
    eng_admit(E);
    double t0 = now_s();
    const double s0 = E.stats.sync_s;
    launch_chunks(E);
    double t1 = now_s();
    E.stats.launch_s += (t1 - t0) - (E.stats.sync_s - s0);
    wait_chunk(E);
    t0 = now_s();
    E.stats.wait_s += t0 - t1;
    harvest(E);
    E.stats.harvest_s += now_s() - t0;
    return eng_active(E);
}

void eng_get_stats(Engine &E, hs_engine_stats_t *st) { *st = E.stats; }

static EngJob &job_of(Engine &E, int id) {
    auto it = E.jobs.find(id);
    if (it == E.jobs.end()) fatal("unknown engine job\n");
    return *it->second;
}

int eng_poll(Engine &E, int id, double *out) {
    EngJob &j = job_of(E, id);
    if (j.phase == J_FAILED) fatal(j.error);
    if (j.phase != J_DONE) return 0;
    if (out) memcpy(out, j.result.data(), j.result.size() * sizeof(double));
    return 1;
}

int eng_result_seq(Engine &E, int id, double *out) {
    EngJob &j = job_of(E, id);
    if (j.phase == J_FAILED) fatal(j.error);
    if (j.phase != J_DONE) return 0;
    if (j.nsteps == 0) fatal("hs_engine_result_seq: a steady job has no sequence\n");
    if (out) memcpy(out, j.seq.data(), j.seq.size() * sizeof(double));
    return 1;
}

void eng_job_info(Engine &E, int id, hs_job_info_t *info) {
    EngJob &j = job_of(E, id);
    info->state = j.phase;
    info->steady_iters = j.steady_iters;
    info->steady_refinements = j.refinements;
    info->rk4_calls = j.rk4_calls;
    info->taylor_sweeps = (double)j.taylor_sweeps;
    info->taylor_node_bytes = j.taylor_node_bytes;
    info->wait_s = j.phase == J_QUEUED ? now_s() - j.t_submit : j.t_admit - j.t_submit;
    info->latency_s = j.phase >= J_DONE ? j.t_done - j.t_submit : now_s() - j.t_submit;
    info->device_bytes = (double)(j.held ? j.held : j.footprint);
    info->held_intervals = j.holds;
    info->horner_steps_min = j.kmin;
    info->horner_steps_plan = j.kplan;
    info->setup_s = j.setup_s;
    info->resident_model = j.resident ? 1 : 0;
    info->steady_solves = (j.leak || j.natural) ? j.leak_iter : (j.phase >= J_TRANS && j.init.empty() ? 1 : 0);
    info->r_convec = j.package ? j.r_convec : 0.0;
    info->be_iters = (double)j.be_iters;
}

void eng_release(Engine &E, int id) {
    EngJob &j = job_of(E, id);
    if (j.phase != J_DONE && j.phase != J_FAILED) fatal("hs_engine_release: the job has not finished\n");
    E.jobs.erase(id);
}

}  // namespace hs
