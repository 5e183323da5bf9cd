// Batched engine: many models advanced by shared launches.  Jobs (a steady
// solve, a transient sequence, or a steady solve followed by a transient
// sequence from its block temperatures) are admitted first come, first
// served while their device footprint fits the engine's budget; every tick
// runs a chunk of iterations for all jobs in their steady phase and a chunk
// of intervals for all jobs in their transient phase, on two streams, then
// moves jobs between phases by their device-side flags.  Each problem's
// arithmetic is the same as when it is solved alone (per-problem scalars,
// per-problem stopping, reductions in a fixed slot order), so a job's result
// does not depend on what it is batched with.  Included at the end of
// hs_engine.inl.
#include <chrono>
#include <deque>
#include <map>

namespace hs {

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum JobPhase { J_QUEUED = 0, J_STEADY = 1, J_TRANS = 2, J_DONE = 3 };

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
};

// A group of resident jobs in one phase with one kernel form (homogeneous or
// heterogeneous coefficients): descriptor array and work-lists on the device.
struct BatchGroup {
    Group g;
    std::vector<EngJob *> members;
    bool dirty = true;
    size_t cap_desc = 0, cap_work = 0, cap_gwork = 0, cap_flags = 0, cap_trec = 0, cap_twork = 0, cap_pwork = 0;
    double *d_flags = nullptr, *h_flags = nullptr;
    Replayer rp;
    SeqForm form{};
    // multigrid steady groups: per level, the descriptors and column work-lists
    std::vector<MgGroup> lv;
    std::vector<void *> lv_bufs;

    void release() {
        rp.reset();
        for (void *p : lv_bufs) (void)hipFree(p);
        lv_bufs.clear();
        lv.clear();
        if (g.d_desc) (void)hipFree(g.d_desc);
        if (g.d_work) (void)hipFree(g.d_work);
        if (g.d_gwork) (void)hipFree(g.d_gwork);
        if (d_flags) (void)hipFree(d_flags);
        if (h_flags) (void)hipHostFree(h_flags);
        if (g.d_kmax) (void)hipFree(g.d_kmax);
        if (g.d_trec) (void)hipFree(g.d_trec);
        if (g.d_twork) (void)hipFree(g.d_twork);
        if (g.d_pwork) (void)hipFree(g.d_pwork);
        g = Group{};
        cap_trec = cap_twork = cap_pwork = 0;
        d_flags = h_flags = nullptr;
        cap_desc = cap_work = cap_gwork = cap_flags = 0;
    }
    ~BatchGroup() { release(); }
    void remove(EngJob *j) {
        members.erase(std::find(members.begin(), members.end(), j));
        dirty = true;
    }
};

// Transient streams of an engine; its transient groups are split into
// HS_TRANS_STREAMS parts (default 2, at most kTransStreams), one per stream.
constexpr int kTransStreams = 3;
static int trans_streams() {
    static const int n = [] {
        const char *e = getenv("HS_TRANS_STREAMS");
        return e ? std::max(1, std::min(kTransStreams, atoi(e))) : 2;
    }();
    return n;
}

struct Engine {
    hs_options_t opt{};
    size_t budget = 0, used = 0;
    int next_id = 0;
    int tick_intervals = 32;            // transient intervals per tick
    int tick_iters = 128;               // steady iterations per tick
    hipStream_t s_steady = nullptr, s_trans[kTransStreams] = {};
    std::deque<EngJob *> queue;
    std::map<int, std::unique_ptr<EngJob>> jobs;
    // steady groups by key: kind (0 homogeneous CG, 1 heterogeneous CG,
    // 2 non-symmetric GMRES) + 4 x multigrid levels (0: Jacobi);
    // transient: homogeneous, heterogeneous, each in parts on the transient
    // streams, [hetero + 2 * part]
    std::map<int, BatchGroup> steady;
    BatchGroup trans[2 * kTransStreams];
    ~Engine() {
        for (auto &kv : jobs) kv.second->D.reset();
        for (auto &kv : steady) kv.second.release();
        for (auto &g : trans) g.release();
        if (s_steady) (void)hipStreamDestroy(s_steady);
        for (auto s : s_trans)
            if (s) (void)hipStreamDestroy(s);
    }
};

// Device bytes dev_build allocates for a model, plus the sequence buffers.
static size_t dev_footprint(const Model &M, int nsteps, bool mg) {
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
    return b + b / 50 + (64 << 10);
}

template <typename T> static void ensure_dev(T *&p, size_t &cap, size_t n) {
    if (n <= cap) return;
    if (p) HIP_OK(hipFree(p));
    cap = std::max(n, cap * 2);
    HIP_OK(hipMalloc((void **)&p, cap * sizeof(T)));
}

// Uploads the group's descriptors and work-lists after a membership change.
static void group_sync(Engine &E, BatchGroup &B, hipStream_t s, bool het, bool trans) {
    if (!B.dirty) return;
    const int n = (int)B.members.size();
    std::vector<ProbDesc> desc(n);
    std::vector<WorkItem> w, gw;
    std::vector<double> rec;   // Taylor records (step factors), transient Taylor groups
    int m = 1, K = 1;
    bool v4 = true;
    for (int i = 0; i < n; i++) {
        const DevModel &D = *B.members[i]->D;
        desc[i] = D.h;
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
    B.rp.reset();
    ensure_dev(B.g.d_desc, B.cap_desc, (size_t)n);
    ensure_dev(B.g.d_work, B.cap_work, w.size());
    ensure_dev(B.g.d_gwork, B.cap_gwork, gw.size());
    if (!B.g.d_kmax) HIP_OK(hipMalloc((void **)&B.g.d_kmax, 2 * sizeof(int)));
    HIP_OK(hipMemset(B.g.d_kmax, 0, 2 * sizeof(int)));
    B.g.kgraph = 0;
    ensure_dev(B.g.d_trec, B.cap_trec, (size_t)kTrec * n);
    if (!rec.empty()) HIP_OK(hipMemcpy(B.g.d_trec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
    if ((size_t)n * 8 > B.cap_flags) {
        if (B.d_flags) HIP_OK(hipFree(B.d_flags));
        if (B.h_flags) HIP_OK(hipHostFree(B.h_flags));
        B.cap_flags = std::max((size_t)n * 8, B.cap_flags * 2);
        HIP_OK(hipMalloc((void **)&B.d_flags, B.cap_flags * sizeof(double)));
        HIP_OK(hipHostMalloc((void **)&B.h_flags, B.cap_flags * sizeof(double), 0));
    }
    HIP_OK(hipMemcpy(B.g.d_desc, desc.data(), n * sizeof(ProbDesc), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(B.g.d_work, w.data(), w.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(B.g.d_gwork, gw.data(), gw.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    B.g.nprob = n;
    B.g.nwork = (int)w.size();
    B.g.ngwork = (int)gw.size();
    B.g.stream = s;
    B.g.hetero = het;
    B.g.vec4 = v4;
    B.g.nper = (n > 0 && (int)w.size() == n * (desc[0].ntiles + 1)) ? desc[0].ntiles + 1 : 0;
    for (int i = 0; i < n && B.g.nper; i++)
        if (desc[i].ntiles != desc[0].ntiles) B.g.nper = 0;
    B.form = SeqForm{E.opt.precision, E.opt.transient, m, K};
    // the paired sweeps' work-list (patches through all layers)
    {
        std::vector<WorkItem> pw;
        B.g.pairs = trans && !het && E.opt.transient == HS_TRANS_TAYLOR && E.opt.precision == HS_PREC_MIXED &&
                    pairs_enabled();
        B.g.pnper = pair_items(desc, pw);
        ensure_dev(B.g.d_pwork, B.cap_pwork, pw.size());
        HIP_OK(hipMemcpy(B.g.d_pwork, pw.data(), pw.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
        B.g.npwork = (int)pw.size();
    }
    // the float4 term sweep's work-list; its tiling timed for FP32 Taylor groups
    if (trans && v4 && E.opt.transient == HS_TRANS_TAYLOR && E.opt.precision == HS_PREC_MIXED) {
        tune_term_tiling(B.g, desc, B.cap_twork);
    } else {
        std::vector<WorkItem> tw;
        const int per = term_items(desc, tw);
        upload_term_items(B.g, tw, per, B.cap_twork);
    }
    // multigrid levels of a steady group (every member has the same count)
    for (void *p : B.lv_bufs) HIP_OK(hipFree(p));
    B.lv_bufs.clear();
    B.lv.clear();
    const int L = (!trans && n) ? (int)B.members[0]->D->mg.size() : 0;
    for (int l = 0; l < L; l++) {
        std::vector<ProbDesc> ld(n);
        std::vector<WorkItem> lw;
        int nlmax = 1;
        for (int i = 0; i < n; i++) {
            const DevModel &D = *B.members[i]->D;
            // level 0 is the model itself, but for the upwind advection of
            // a non-symmetric model's preconditioner
            ld[i] = l == 0 && !D.h.has_adv ? D.h : D.mg[l].h;
            nlmax = std::max(nlmax, ld[i].nl);
            auto ci = column_items(ld[i], i);
            lw.insert(lw.end(), ci.begin(), ci.end());
        }
        ProbDesc *dd = nullptr;
        WorkItem *dw = nullptr;
        HIP_OK(hipMalloc((void **)&dd, n * sizeof(ProbDesc)));
        HIP_OK(hipMalloc((void **)&dw, lw.size() * sizeof(WorkItem)));
        B.lv_bufs.push_back(dd);
        B.lv_bufs.push_back(dw);
        HIP_OK(hipMemcpy(dd, ld.data(), n * sizeof(ProbDesc), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dw, lw.data(), lw.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
        B.lv.push_back(MgGroup{dd, dw, (int)lw.size(), nlmax, l > 0 || het});
    }
    B.dirty = false;
}

static void gather(BatchGroup &B, std::initializer_list<int> slots) {
    Slots sl{};
    for (int v : slots) sl.s[sl.n++] = v;
    hipLaunchKernelGGL(k_gather, dim3(small_grid(B.g.nprob)), dim3(64), 0, B.g.stream,
                       (const ProbDesc *)B.g.d_desc, B.g.nprob, sl, B.d_flags);
    HIP_OK(hipMemcpyAsync(B.h_flags, B.d_flags, (size_t)B.g.nprob * 8 * sizeof(double), hipMemcpyDeviceToHost,
                          B.g.stream));
}

static void release_device(Engine &E, EngJob &j) {
    E.used -= j.held;
    j.held = 0;
    j.D.reset();
}

static int steady_key(const EngJob &j) {
    return (j.adv ? 2 : j.hetero ? 1 : 0) + 4 * (int)j.D->mg.size();
}

static void start_transient(Engine &E, EngJob &j) {
    DevModel &D = *j.D;
    auto kv = seq_slots(j.nsteps, j.dt);
    kv.push_back({S_YPAR, 0.0});
    if (E.opt.transient == HS_TRANS_TAYLOR) {
        const auto t = taylor_slots(taylor_plan(D, j.dt, E.opt.taylor_tol), D.rho);
        kv.insert(kv.end(), t.begin(), t.end());
    }
    write_scal(D, kv);
    j.phase = J_TRANS;
    // the part with the fewest members: while one part runs the latency-
    // bound kernels of its interval (power map, first term, finishes), the
    // others' Horner sweeps use the memory bandwidth.  One part only while a
    // kernel timer runs (launches of the parts would overlap in its event
    // spans).
    const int parts = g_timer.cap > 0 ? 1 : trans_streams();
    int best = j.hetero;
    for (int q = 1; q < parts; q++)
        if (E.trans[j.hetero + 2 * q].members.size() < E.trans[best].members.size()) best = j.hetero + 2 * q;
    BatchGroup &B = E.trans[best];
    B.members.push_back(&j);
    B.dirty = true;
}

static void admit(Engine &E, EngJob &j) {
    Model &M = *j.M;
    j.D.reset(dev_build(M, E.s_steady));
    DevModel &D = *j.D;
    j.t_admit = now_s();
    const size_t nb = (size_t)D.nblk;
    if (j.nsteps > 0) {
        D.h.seq_in = D.alloc<double>((size_t)j.nsteps * nb);
        D.h.seq_out = D.alloc<double>((size_t)j.nsteps * nb);
        // interval powers: one row for all, one per interval, or the steady row first
        const int first = j.rows == j.nsteps + 1 ? 1 : 0;
        if (j.rows == 1) {
            std::vector<double> rep((size_t)j.nsteps * nb);
            for (int k = 0; k < j.nsteps; k++) std::copy(j.power.begin(), j.power.end(), rep.begin() + k * nb);
            HIP_OK(hipMemcpy(D.h.seq_in, rep.data(), rep.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            HIP_OK(hipMemcpyAsync(D.h.seq_in, j.power.data() + first * nb, (size_t)j.nsteps * nb * sizeof(double),
                                  hipMemcpyHostToDevice, D.stream));
        }
        HIP_OK(hipMemcpyAsync(D.d_desc, &D.h, sizeof(ProbDesc), hipMemcpyHostToDevice, D.stream));
        HIP_OK(hipStreamSynchronize(D.stream));
    }
    j.held = D.bytes;
    E.used += j.held;
    if (j.init.empty()) {
        upload_blocks(D, j.power.data());
        run_b2g(D.grp, V_POWER, B_RHS, 1);
        const int nbk = (int)((D.vlen + kBlock - 1) / kBlock);
        launch(k_fill, nbk, D.stream, D.h.b64[B_STEADY], D.vlen, M.g->config.ambient);
        if (j.adv) {
            // non-symmetric: GMRES with the upwind multigrid preconditioner
            // in every form (Jacobi does not converge at channel Peclet numbers)
            ensure_gmres(D, true);
            build_mg(D);
            write_scal(D, kMixedStart);
        } else if (E.opt.steady_solver == HS_STEADY_MGPCG) {
            build_mg(D);
            write_scal(D, kMixedStart);
        } else if (E.opt.precision == HS_PREC_MIXED) {
            write_scal(D, kMixedStart);
        } else {
            write_scal(D, {{S_DONE, 0.0}});
            const double rtol2 = E.opt.steady_rtol * E.opt.steady_rtol;
            if (j.hetero) jpcg64_start<true>(D.grp, rtol2, E.opt.max_iters);
            else jpcg64_start<false>(D.grp, rtol2, E.opt.max_iters);
        }
        j.phase = J_STEADY;
        BatchGroup &B = E.steady[steady_key(j)];
        B.members.push_back(&j);
        B.dirty = true;
    } else {
        upload_blocks(D, j.init.data());
        run_b2g(D.grp, V_TEMP, B_Y0, 0);
        HIP_OK(hipStreamSynchronize(D.stream));
        start_transient(E, j);
    }
    // device memory held, multigrid levels and GMRES state included
    E.used += D.bytes - j.held;
    j.held = D.bytes;
}

// A finished steady solve: block temperatures, then either the result or
// the start of the job's transient sequence from them.
static void steady_done(Engine &E, EngJob &j, const double *flags) {
    DevModel &D = *j.D;
    const bool refined = E.opt.precision == HS_PREC_MIXED || j.adv || E.opt.steady_solver == HS_STEADY_MGPCG;
    j.steady_iters = (int)flags[1];
    j.refinements = refined ? (int)flags[2] : 0;
    run_g2b(D.grp, B_STEADY);
    if (j.nsteps == 0) {
        j.result.resize(D.nblk);
        download_blocks(D, j.result.data());
        j.phase = J_DONE;
        j.t_done = now_s();
        release_device(E, j);
        return;
    }
    run_b2g(D.grp, V_TEMP, B_Y0, 0, 2);
    HIP_OK(hipStreamSynchronize(D.stream));
    start_transient(E, j);
}

static void trans_done(Engine &E, EngJob &j, const double *flags) {
    DevModel &D = *j.D;
    j.rk4_calls = E.opt.transient == HS_TRANS_RK4 ? (int)flags[1] : 0;
    j.taylor_sweeps = E.opt.transient == HS_TRANS_TAYLOR ? (long)flags[2] : 0;
    j.taylor_node_bytes = E.opt.transient == HS_TRANS_TAYLOR ? flags[3] : 0.0;
    const size_t n = (size_t)j.nsteps * D.nblk;
    j.seq.resize(n);
    HIP_OK(hipMemcpyAsync(j.seq.data(), D.h.seq_out, n * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
    j.result.assign(j.seq.end() - D.nblk, j.seq.end());
    j.phase = J_DONE;
    j.t_done = now_s();
    release_device(E, j);
}

template <bool HET>
static void steady_chunk(Engine &E, const Group &G, bool adv, const std::vector<MgGroup> &lv) {
    const double rtol2 = E.opt.steady_rtol * E.opt.steady_rtol;
    if (!lv.empty() && adv) {   // multigrid-preconditioned GMRES cycles
        for (int done = 0; done < E.tick_iters; done += 64) gm_round_mg<true>(G, lv, rtol2, 1e-8, kGmM);
    } else if (!lv.empty()) {   // multigrid-preconditioned CG rounds
        for (int done = 0; done < E.tick_iters; done += 64) mixed_round_mg<HET>(G, lv, rtol2, E.opt.max_iters, 16);
    } else if (adv) {   // GMRES cycles inside FP64 refinement
        const int cycles = std::max(1, E.tick_iters / kGmM);
        for (int c = 0; c < cycles; c++) {
            if (E.opt.precision == HS_PREC_MIXED) gm_round<float, true>(G, rtol2, 1e-8, kGmM);
            else gm_round<double, true>(G, rtol2, 1e-12, kGmM);
        }
    } else if (E.opt.precision == HS_PREC_MIXED) {
        for (int done = 0; done < E.tick_iters; done += 64) mixed_round<HET>(G, rtol2, E.opt.max_iters, 64);
    } else {
        cg_iterations<double, HET>(G, E.tick_iters, B_STEADY, B_R, B_Z, B_P0, B_Q, S_DONE, S_TOL2,
                                   E.opt.max_iters);
    }
}

// ---- entry points (wrapped by the C ABI) -----------------------------------------
Engine *eng_create(int precision, int transient) {
    std::unique_ptr<Engine> E(new Engine());
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
    HIP_OK(hipStreamCreateWithFlags(&E->s_steady, hipStreamNonBlocking));
    for (auto &s : E->s_trans) HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return E.release();
}

void eng_destroy(Engine *E) { delete E; }

int eng_active(Engine &E) {
    int n = (int)E.queue.size();
    for (auto &kv : E.steady) n += (int)kv.second.members.size();
    for (auto &g : E.trans) n += (int)g.members.size();
    return n;
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
    j->footprint = dev_footprint(M, nsteps, !init && (E.opt.steady_solver == HS_STEADY_MGPCG || adv));
    if (j->footprint > E.budget)
        fatal("hs_engine_submit: the job needs " + std::to_string(j->footprint) +
              " bytes of device memory, above the engine budget of " + std::to_string(E.budget) + "\n");
    j->t_submit = now_s();
    E.queue.push_back(j.get());
    const int id = j->id;
    E.jobs[id] = std::move(j);
    return id;
}

// admission: first come, first served while the footprint fits
int eng_admit(Engine &E) {
    while (!E.queue.empty()) {
        EngJob &j = *E.queue.front();
        if (E.used + j.footprint > E.budget && E.used > 0) break;
        E.queue.pop_front();
        admit(E, j);
    }
    return (int)E.queue.size();
}

int eng_tick(Engine &E) {
    eng_admit(E);
    // members launched this tick per group: only their gathered flags are
    // fresh (a group may gain members below, or be empty this tick)
    std::map<int, int> ran_steady;
    int ran_trans[2 * kTransStreams] = {};
    for (auto &kv : E.steady) {
        BatchGroup &B = kv.second;
        const int kind = kv.first % 4;
        if (B.members.empty()) continue;
        group_sync(E, B, E.s_steady, kind > 0, false);
        ran_steady[kv.first] = B.g.nprob;
        if (kind) steady_chunk<true>(E, B.g, kind == 2, B.lv);
        else steady_chunk<false>(E, B.g, false, B.lv);
        const bool refined = E.opt.precision == HS_PREC_MIXED || kind == 2 || !B.lv.empty();
        gather(B, {S_DONE, refined ? S_TITERS : S_ITERS, S_NREFINE});
    }
    for (int h = 0; h < 2 * kTransStreams; h++) {
        BatchGroup &B = E.trans[h];
        if (B.members.empty()) continue;
        group_sync(E, B, E.s_trans[h / 2], h % 2, true);
        ran_trans[h] = B.g.nprob;
        B.rp.run(E.s_trans[h / 2], E.tick_intervals, [&] { seq_body(B.g, B.form); });
        gather(B, {S_STEP, S_TCALLS, S_TTERMS, S_TBYTES});
    }
    HIP_OK(hipStreamSynchronize(E.s_steady));
    for (auto st : E.s_trans) HIP_OK(hipStreamSynchronize(st));
    // Taylor groups of one sub-step per interval launch the Horner steps
    // that the tick's largest term count needs, plus one: fewer launches that
    // skip every problem.  A problem that needs more is held (its interval
    // not advanced) until the next tick raises the count.  A change
    // recaptures the group's graph; the count falls only by two or more.
    for (int h = 0; h < 2 * kTransStreams; h++) {
        BatchGroup &B = E.trans[h];
        if (!ran_trans[h] || B.form.transient != HS_TRANS_TAYLOR || B.form.m != 1 || !kgraph_enabled()) continue;
        int kt[2] = {0, 0};
        HIP_OK(hipMemcpy(kt, B.g.d_kmax, sizeof kt, hipMemcpyDeviceToHost));
        HIP_OK(hipMemsetAsync(B.g.d_kmax + 1, 0, sizeof(int), E.s_trans[h / 2]));
        if (kt[1] <= 0) continue;
        const int cur = B.g.kgraph > 0 ? B.g.kgraph : B.form.K;
        static const int margin = getenv("HS_KGRAPH_MARGIN") ? std::max(0, atoi(getenv("HS_KGRAPH_MARGIN"))) : 1;
        const int want = std::min(B.form.K, std::max(2, kt[1] + margin));
        if (kt[1] > cur || want < cur - 1) {
            B.g.kgraph = want >= B.form.K ? 0 : want;
            B.rp.reset();
        }
    }
    for (auto &kv : E.steady) {
        BatchGroup &B = kv.second;
        std::vector<std::pair<EngJob *, std::vector<double>>> fin;
        for (int i = 0; i < ran_steady[kv.first]; i++)
            if (B.h_flags[i * 8] != 0.0)
                fin.push_back({B.members[i], std::vector<double>(B.h_flags + i * 8, B.h_flags + i * 8 + 8)});
        for (auto &f : fin) {
            B.remove(f.first);
            steady_done(E, *f.first, f.second.data());
        }
    }
    for (int h = 0; h < 2 * kTransStreams; h++) {
        BatchGroup &B = E.trans[h];
        std::vector<std::pair<EngJob *, std::vector<double>>> fin;
        for (int i = 0; i < ran_trans[h]; i++)
            if (B.h_flags[i * 8] >= (double)B.members[i]->nsteps)
                fin.push_back({B.members[i], std::vector<double>(B.h_flags + i * 8, B.h_flags + i * 8 + 8)});
        for (auto &f : fin) {
            B.remove(f.first);
            trans_done(E, *f.first, f.second.data());
        }
    }
    return eng_active(E);
}

static EngJob &job_of(Engine &E, int id) {
    auto it = E.jobs.find(id);
    if (it == E.jobs.end()) fatal("unknown engine job\n");
    return *it->second;
}

int eng_poll(Engine &E, int id, double *out) {
    EngJob &j = job_of(E, id);
    if (j.phase != J_DONE) return 0;
    if (out) memcpy(out, j.result.data(), j.result.size() * sizeof(double));
    return 1;
}

int eng_result_seq(Engine &E, int id, double *out) {
    EngJob &j = job_of(E, id);
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
    info->latency_s = j.phase == J_DONE ? j.t_done - j.t_submit : now_s() - j.t_submit;
    info->device_bytes = (double)(j.held ? j.held : j.footprint);
}

void eng_release(Engine &E, int id) {
    EngJob &j = job_of(E, id);
    if (j.phase != J_DONE) fatal("hs_engine_release: the job has not finished\n");
    E.jobs.erase(id);
}

}  // namespace hs
