// Host side of the device engine: per-model device state, descriptor
// building, heterogeneous assembly, and the solve drivers.  Included at the
// end of hs_device.hip (same translation unit as the kernels).
#include <array>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <unordered_map>

namespace hs {

// Device (or pinned host) memory kept for reuse instead of returned to the
// runtime.  hipFree waits for every stream of the device, so in the engine a
// job that ends would stall the launches of every other resident job; its
// blocks go back to the pool instead and serve the next job of a similar
// size.  Blocks are reused by size class (steps of 1/8 between powers of
// two, 256 B below 4 KiB).
struct DevPool {
    bool host = false;
    std::multimap<size_t, void *> idle;
    std::unordered_map<void *, size_t> cls;
    size_t idle_bytes = 0, total_bytes = 0;
    static size_t size_class(size_t n) {
        if (n <= 4096) return (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
        size_t p = 4096;
        while (p * 2 <= n) p *= 2;
        const size_t step = p / 8;
        return (n + step - 1) / step * step;
    }
    void *get(size_t n, size_t *got) {
        const size_t s = size_class(n);
        void *p = nullptr;
        auto it = idle.find(s);
        if (it != idle.end()) {
            p = it->second;
            idle.erase(it);
            idle_bytes -= s;
        } else {
            if (host) HIP_OK(hipHostMalloc(&p, s, 0));
            else HIP_OK(hipMalloc(&p, s));
            cls[p] = s;
            total_bytes += s;
        }
        if (got) *got = s;
        return p;
    }
    void put(void *p) {
        const size_t s = cls.at(p);
        idle.insert({s, p});
        idle_bytes += s;
    }
    // returns the idle blocks to the runtime (hipFree waits for the device)
    void trim() {
        for (auto &kv : idle) {
            if (host) (void)hipHostFree(kv.second);
            else (void)hipFree(kv.second);
            cls.erase(kv.second);
            total_bytes -= kv.first;
        }
        idle.clear();
        idle_bytes = 0;
    }
    ~DevPool() { trim(); }
};

static void *dmalloc(DevPool *pool, size_t bytes, size_t *got = nullptr) {
    if (pool) return pool->get(bytes, got);
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(bytes, 1)));
    if (got) *got = std::max<size_t>(bytes, 1);
    return p;
}
static void dfree(DevPool *pool, void *p) {
    if (!p) return;
    if (pool) pool->put(p);
    else (void)hipFree(p);
}
static void *hmalloc(DevPool *pool, size_t bytes) {
    if (pool) return pool->get(bytes, nullptr);
    void *p = nullptr;
    HIP_OK(hipHostMalloc(&p, std::max<size_t>(bytes, 1), 0));
    return p;
}
static void hfree(DevPool *pool, void *p) {
    if (!p) return;
    if (pool) pool->put(p);
    else (void)hipHostFree(p);
}

// Copies ordered on a stream (never the null stream, whose work can queue
// behind other streams' launches on a shared hardware queue); they return
// once the copy is done, so the host source may go out of scope.
static void h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
}
static void d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
}
static void dzero(void *p, size_t bytes, hipStream_t s) { HIP_OK(hipMemsetAsync(p, 0, bytes, s)); }

// A set of problems advanced by shared launches: the descriptor array, the
// cell-tile work-list (tiles of every problem, then its package item) and the
// grid->block work-list.  One model is a group of one.
struct Group {
    ProbDesc *d_desc = nullptr;
    int nprob = 0;
    WorkItem *d_work = nullptr;
    int nwork = 0;
    WorkItem *d_gwork = nullptr;
    int ngwork = 0;
    hipStream_t stream = nullptr;
    bool hetero = false;
    bool vec4 = false;   // every problem takes float4 term sweeps (vec4_ok)
    int *d_kmax = nullptr;   // Taylor: largest term count of the current sub-step, and since last cleared
    int kgraph = 0;          // Taylor: Horner steps launched per sub-step, below the plan (0: the plan's)
    double *d_trec = nullptr;   // Taylor: per problem {tau, terms} of the current sub-step
    int nper = 0;   // work items per problem when every problem has the same count, else 0
    // work-list of the float4 Taylor sweep over the term tiling (t_*)
    WorkItem *d_twork = nullptr;
    int ntwork = 0, tnper = 0;
    int tz = 0;   // its tiles: 0 per layer for every problem, 1 z-marching for every problem, 2 mixed
    int tg = 2;   // most row groups of 4 rows per thread among its problems (t_rpt / 4)
    bool zfirst = false;   // the first term z-marches over the term work-list (k_tay_first4z)
    bool zpow = false;     // ... and maps the cell power of sequence intervals itself (zpow_ok)
};

// A homogeneous problem whose tiles are at least 16 columns wide with at
// least 4 rows per thread runs its FP32 term sweeps with float4 accesses,
// when its rows hold whole tiles of both its tilings (the float4 term tiling
// and the per-layer tiling of the double2 first term); partial row tiles are
// fine (rows past the grid are skipped), partial column tiles are not.
static bool vec4_ok(const ProbDesc &h) {
    return !h.hetero && h.t_rpt >= 4 && h.t_shift >= 4 && h.nc % (1 << h.t_shift) == 0 &&
           h.nc % (1 << h.tx_shift) == 0;
}

// One multigrid level of one problem: its descriptor (the finest level is
// the problem's own) and the column work-list (2D tiles, all layers per
// thread, then the package item).
struct MgLevel {
    ProbDesc h{};
    ProbDesc *d_desc = nullptr;
    WorkItem *d_cwork = nullptr;
    int ncwork = 0;
};

struct DevModel {
    Group grp;                          // this model alone
    std::vector<MgLevel> mg;            // multigrid levels (built on first use); [0] is this model
    ProbDesc h{};                       // host copy of the descriptor
    ProbDesc *d_desc = nullptr;
    WorkItem *d_work = nullptr;         // cell tiles + package item
    int nwork = 0;
    WorkItem *d_gwork = nullptr;        // grid -> block items
    int ngwork = 0;
    std::vector<void *> allocs;
    size_t bytes = 0;                   // device memory held
    hipStream_t stream = nullptr;
    bool own_stream = true;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *h_scal = nullptr;           // pinned copy of the scalars
    double *h_blk = nullptr;            // pinned block vector staging
    int nblk = 0;
    long vlen = 0;
    double rho = 0.0;                   // bound on ||C^-1 G||_inf (transient)
    bool hetero = false, has_c = false;
    std::vector<double> trec_plan;      // Taylor record of grp.d_trec as last uploaded (step factors)
    size_t cap_twork = 0;               // grp.d_twork capacity (its own allocation)
    bool term_tuned = false;            // term tiling chosen by timing
    // backward Euler: descriptor of the shifted operator G + C/h (its group
    // shares the work-lists), per-cell shift and Jacobi diagonal (hetero)
    ProbDesc *d_be_desc = nullptr;
    ProbDesc be_h{};                    // host copy of *d_be_desc
    Group be_grp;
    double *be_shift = nullptr, *be_dinv = nullptr;
    std::vector<MgLevel> be_mg;         // multigrid levels of the shifted step (non-symmetric steps)
    double be_sigma = 0.0;              // the shift 1/h they hold

    DevPool *pool = nullptr, *hpool = nullptr;   // an engine's pools (nullptr: the runtime's)
    // the device model whose read-only data (maps, block ranges, work-lists,
    // heterogeneous coefficients) this one uses instead of its own copy
    std::shared_ptr<DevModel> base;

    std::vector<size_t> alloc_sz;       // bytes of each allocation in allocs

    // zeroed, whether fresh or a pool block of an earlier job
    template <typename T> T *alloc(size_t n) {
        size_t got = 0;
        void *p = dmalloc(pool, std::max<size_t>(n, 1) * sizeof(T), &got);
        allocs.push_back(p);
        alloc_sz.push_back(got);
        bytes += got;
        dzero(p, got, stream);
        return (T *)p;
    }
    template <typename T> T *upload(const T *src, size_t n) {
        T *p = alloc<T>(n);
        h2d(p, src, n * sizeof(T), stream);
        return p;
    }
    // returns one allocation of alloc() (e.g. a job's sequence buffers)
    void release(void *p) {
        for (size_t k = 0; k < allocs.size(); k++)
            if (allocs[k] == p) {
                bytes -= alloc_sz[k];
                allocs.erase(allocs.begin() + k);
                alloc_sz.erase(alloc_sz.begin() + k);
                dfree(pool, p);
                return;
            }
    }
    ~DevModel() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) dfree(pool, p);
        dfree(pool, grp.d_twork);
        hfree(hpool, h_scal);
        hfree(hpool, h_blk);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }
};

void dev_release(Model &M) {
    delete M.dev;
    M.dev = nullptr;
}

// Optional timing of one kernel family (capped): the live measurement
// behind the roofline figures.  Consecutive launches of the family on one
// stream form a run, timed by one pair of events: the start event before its
// first launch and the end event when the run closes (at the next launch of
// another kernel through launch_named, a launch on another stream, or the
// read), so no event sits between the run's kernels.  The mean per launch is
// the run's span over its launch count, the gaps between the launches
// included.
struct KernelTimer {
    struct Run {
        hipEvent_t a, b;
        int n;
    };
    std::string name;
    int cap = 0;
    std::vector<Run> ev;                                 // pending (plain launches)
    bool open = false;                                   // ev.back() still takes launches (no end event yet)
    hipStream_t stream = nullptr;                        // stream of the open run
    int launched = 0;
    double graph_ms = 0.0;                               // sampled from graph replays
    int graph_n = 0;
    int taken() const { return launched + graph_n; }
    void close() {
        if (!open) return;
        HIP_OK(hipEventCreate(&ev.back().b));
        HIP_OK(hipEventRecord(ev.back().b, stream));
        open = false;
    }
};
static KernelTimer g_timer;

template <typename K, typename... A>
static inline void launch_named(const char *name, K kern, int nblocks, hipStream_t s, A... args) {
    if (nblocks <= 0) return;
    const bool timed = g_timer.cap > 0 && g_timer.taken() < g_timer.cap && g_timer.name == name;
    if (!timed || s != g_timer.stream) g_timer.close();
    if (timed && !g_timer.open) {
        hipEvent_t a = nullptr;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventRecord(a, s));
        g_timer.ev.push_back({a, nullptr, 0});
        g_timer.open = true;
        g_timer.stream = s;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kBlock), 0, s, args...);
    if (timed) {
        g_timer.ev.back().n++;
        g_timer.launched++;
    }
}
#define launch(kern, ...) launch_named(#kern, kern, __VA_ARGS__)

static int ilog2(int v) {
    int s = 0;
    while ((1 << (s + 1)) <= v) s++;
    return s;
}

// Tile shape: TX = min(nc, 64) consecutive columns (one wavefront reads a
// contiguous row segment), TY = 256 / TX thread rows, each thread owning rpt
// grid rows.  rpt grows (up to 8) while a layer keeps at least 8 tiles
// along its rows: fewer, longer-lived workgroups amortise each workgroup's
// fixed cost (work item, descriptor and flag reads) over more cells.
static int rows_per_thread(int nr, int ty) {
    int r = 1;
    while (r < 8 && nr >= ty * r * 2 * 8) r *= 2;
    return r;
}

// tiling of the float4 Taylor sweep: tile width 1 << shift (at most nc),
// rpt rows per thread; z: column tiles through every layer instead of one
// tile per layer
static void set_term_tiling(ProbDesc &h, int shift, int rpt, int z = 0) {
    const int ty = kBlock >> shift;
    h.t_shift = shift;
    h.t_rpt = rpt;
    h.t_nti = (h.nr + ty * rpt - 1) / (ty * rpt);
    h.t_ntiles = h.nl * h.t_nti * (h.nc >> shift);
    h.t_z = z;
}
// cell items of the float4 sweep's work-list for a problem
static int term_cell_items(const ProbDesc &h) { return h.t_z ? h.t_ntiles / h.nl : h.t_ntiles; }

static void set_tiling(ProbDesc &h) {
    h.tx_shift = ilog2(std::min(h.nc, 64));
    const int tx = 1 << h.tx_shift, ty = kBlock >> h.tx_shift;
    h.rpt = rows_per_thread(h.nr, ty);
    h.ntj = (h.nc + tx - 1) / tx;
    h.nti = (h.nr + ty * h.rpt - 1) / (ty * h.rpt);
    h.ntiles = h.nl * h.nti * h.ntj;
    h.nti2 = (h.nr + ty - 1) / ty;
    set_term_tiling(h, h.tx_shift, h.rpt);
}

// Work-list of the float4 term sweep: (problem, tile) over each problem's
// term tiling, its package item last.  Returns the items per problem when
// every problem has the same count, else 0.
// the tile form of a launch over these problems (Group::tz, Group::tg)
static int term_form(const std::vector<ProbDesc> &desc) {
    int z = 0, l = 0;
    for (const ProbDesc &h : desc) (h.t_z ? z : l)++;
    return l == 0 && z > 0 ? 1 : z == 0 ? 0 : 2;
}
static int term_groups(const std::vector<ProbDesc> &desc) {
    int g = 1;
    for (const ProbDesc &h : desc) g = std::max(g, h.t_rpt / 4);
    return g;
}
static int slot_count(const ProbDesc &h);
// every problem z-marches and its column tiles and package item fit its
// reduction slots: the first term can z-march too
static bool first_z_ok(const std::vector<ProbDesc> &desc) {
    for (const ProbDesc &h : desc)
        if (!h.t_z || h.t_ntiles / h.nl + 1 > slot_count(h)) return false;
    return !desc.empty();
}

// Whether k_tay_first4z can map the cell power itself: every problem has at
// most kZPowTab blocks, one-block codes on every layer and cells of one to
// four blocks (blk_w set).
static bool zpow_ok(const std::vector<ProbDesc> &desc) {
    for (const ProbDesc &h : desc)
        if (h.nblocks > kZPowTab || !h.blk_w) return false;
    return !desc.empty();
}

// the float4 Horner step over the group's term work-list, in the kernel of
// its tile form (fin: the step adds its result into the FP64 state)
static void launch_term4(const Group &G, hipStream_t s, int e1b, int d0b, int d1b, int j, int fin, bool named) {
    auto go = [&](auto kern) {
        if (named)
            launch_named("k_tay_term", kern, G.ntwork, s, (const ProbDesc *)G.d_desc, (const WorkItem *)G.d_twork,
                         G.tnper, e1b, d0b, d1b, j, (const int *)G.d_kmax, (const double *)G.d_trec);
        else
            hipLaunchKernelGGL(kern, dim3(G.ntwork), dim3(kBlock), 0, s, (const ProbDesc *)G.d_desc,
                               (const WorkItem *)G.d_twork, G.tnper, e1b, d0b, d1b, j, (const int *)G.d_kmax,
                               (const double *)G.d_trec);
    };
    auto by_fin = [&](auto k0, auto k1) { fin ? go(k1) : go(k0); };
    const bool one = G.tg == 1;
    if (G.tz == 1)
        one ? by_fin(k_tay_term4<1, 1, 0>, k_tay_term4<1, 1, 1>) : by_fin(k_tay_term4<1, 2, 0>, k_tay_term4<1, 2, 1>);
    else if (G.tz == 0)
        one ? by_fin(k_tay_term4<0, 1, 0>, k_tay_term4<0, 1, 1>) : by_fin(k_tay_term4<0, 2, 0>, k_tay_term4<0, 2, 1>);
    else
        one ? by_fin(k_tay_term4<2, 1, 0>, k_tay_term4<2, 1, 1>) : by_fin(k_tay_term4<2, 2, 0>, k_tay_term4<2, 2, 1>);
}

static int term_items(const std::vector<ProbDesc> &desc, std::vector<WorkItem> &w) {
    w.clear();
    int per = desc.empty() ? 0 : term_cell_items(desc[0]) + 1;
    for (int i = 0; i < (int)desc.size(); i++) {
        const int n = term_cell_items(desc[i]);
        for (int t = 0; t <= n; t++) w.push_back(WorkItem{i, t});
        if (n + 1 != per) per = 0;
    }
    return per;
}

// Tile shape of the float4 term sweep, picked by timing per size class
// (layers, rows, columns, and the power of two below the number of the
// group's problems of that shape: how many share a launch changes what the
// caches hold and how the workgroups fill the device).  Candidates: tile
// widths of 16 to 256 columns (at most nc) with 4 or 8 rows per thread, per
// layer or z-marching.  Each is timed over the Horner steps of an interval
// of the class's problems (five steps reading h and E_1 and writing a term
// buffer, then the last step, which adds into the state: the term buffers
// are free between intervals and the state is redirected to free FP64
// buffers while timing), and the fastest kept.  The arithmetic of a cell is
// the same under every tiling, so the choice changes only the speed.
// Choices are remembered per class, so a group whose membership changes
// times only the classes it has not met; a shape pinned by
// hs_pin_sweep_tiling takes its pin whatever the count.
struct TermShape {
    int shift = 0, rpt = 0, z = 0;
};
using ShapeKey = std::array<int, 4>;
using GridKey = std::array<int, 3>;
static std::map<ShapeKey, TermShape> &term_shape_cache() {
    static std::map<ShapeKey, TermShape> m;
    return m;
}
static std::map<GridKey, TermShape> &term_shape_pins() {
    static std::map<GridKey, TermShape> m;
    return m;
}
static GridKey grid_key(const ProbDesc &h) { return GridKey{h.nl, h.nr, h.nc}; }

// Host-to-device copy of a set-up: by default ordered on s and waited for
// (h2d); an engine group passes one that stages the copy without a wait.
using Upload = std::function<void(void *, const void *, size_t, hipStream_t)>;
static void h2d_wait(void *dst, const void *src, size_t bytes, hipStream_t s) { h2d(dst, src, bytes, s); }

static void upload_term_items(Group &G, const std::vector<WorkItem> &w, int per, size_t &cap, DevPool *pool,
                              hipStream_t s, const Upload &up = h2d_wait) {
    if (w.size() > cap) {
        dfree(pool, G.d_twork);
        cap = std::max(w.size(), 2 * cap);
        G.d_twork = (WorkItem *)dmalloc(pool, cap * sizeof(WorkItem));
    }
    up(G.d_twork, w.data(), w.size() * sizeof(WorkItem), s);
    G.ntwork = (int)w.size();
    G.tnper = per;
}

// Times the candidates on stream s (the group's own stream may be busy
// with another group's launches) and applies each problem's class shape.
static void tune_term_tiling(Group &G, std::vector<ProbDesc> &desc, size_t &cap, DevPool *pool, hipStream_t s,
                             const Upload &up = h2d_wait) {
    auto &cache = term_shape_cache();
    const auto &pins = term_shape_pins();
    const int n = (int)desc.size();
    std::map<GridKey, int> count;
    for (const ProbDesc &h : desc) count[grid_key(h)]++;
    auto key_of = [&](const ProbDesc &h) {
        const GridKey g = grid_key(h);
        return ShapeKey{g[0], g[1], g[2], ilog2(count[g])};
    };
    // a class (grid shape, problem count) not timed yet takes the choice of
    // the nearest count of its grid shape, if any: a served stream forms
    // groups of ever new compositions, and the choice changes little with
    // the count; only grid shapes never seen are timed
    auto resolve = [&](const ShapeKey &k) {
        if (cache.count(k)) return k;
        ShapeKey best = k;
        int gap = 1 << 30;
        for (auto it = cache.lower_bound(ShapeKey{k[0], k[1], k[2], -1});
             it != cache.end() && it->first[0] == k[0] && it->first[1] == k[1] && it->first[2] == k[2]; ++it)
            if (std::abs(it->first[3] - k[3]) < gap) {
                gap = std::abs(it->first[3] - k[3]);
                best = it->first;
            }
        return best;
    };
    std::vector<ShapeKey> todo;
    for (const ProbDesc &h : desc) {
        const ShapeKey k = key_of(h);
        if (!pins.count(grid_key(h)) && !cache.count(resolve(k)) && std::find(todo.begin(), todo.end(), k) == todo.end())
            todo.push_back(k);
    }
    if (!todo.empty()) {
        // every problem active for steps j = 2 of a three-term series: h
        // from one term buffer, E_1 from the accumulator, the result to the
        // other term buffer, or (the last step) into the state, here the RK4
        // work buffers; the group's Taylor records are restored after timing
        std::vector<double> saved((size_t)kTrec * n), rec((size_t)kTrec * n, 0.0);
        d2h(saved.data(), G.d_trec, saved.size() * sizeof(double), s);
        for (int i = 0; i < n; i++) {
            rec[(size_t)kTrec * i] = 1e-9;
            rec[(size_t)kTrec * i + 1] = 3.0;
            rec[(size_t)kTrec * i + 4] = 1e-9 / 2;   // s_2 (unshifted)
            ((float *)&rec[(size_t)kTrec * i + kTrecF])[4] = (float)(1e-9 / 2);
        }
        const int kt[2] = {3, 0};
        h2d(G.d_trec, rec.data(), rec.size() * sizeof(double), s);
        h2d(G.d_kmax, kt, sizeof kt, s);
        std::vector<ProbDesc> tdesc(desc);
        for (ProbDesc &h : tdesc) {
            h.b64[B_Y0] = h.b64[B_KA];
            h.b64[B_Y1] = h.b64[B_KB];
        }
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        for (const ShapeKey &k : todo) {
            std::vector<int> mem;
            for (int i = 0; i < n; i++)
                if (key_of(desc[i]) == k) mem.push_back(i);
            TermShape best;
            float best_ms = 1e30f;
            // tile width at most 256: a tile row's 64 threads stay within one
            // wavefront (the z-marching sweep shuffles along the row)
            for (int shift = 4; shift <= std::min(8, ilog2(k[2])); shift++)
                for (int rpt : {4, 8})
                if (k[2] % (1 << shift) == 0)
                for (int z : {0, 1}) {
                    for (int i : mem) set_term_tiling(tdesc[i], shift, rpt, z);
                    h2d(G.d_desc, tdesc.data(), tdesc.size() * sizeof(ProbDesc), s);
                    std::vector<WorkItem> w;
                    for (int i : mem)
                        for (int t = 0; t <= term_cell_items(tdesc[i]); t++) w.push_back(WorkItem{i, t});
                    upload_term_items(G, w, 0, cap, pool, s);
                    constexpr int kWarm = 1, kTimed = 2, kSteps = 5;
                    // the launch covers this class's problems only
                    G.tz = z;
                    G.tg = rpt / 4;
                    for (int r = 0; r < kWarm + kTimed; r++) {
                        if (r == kWarm) HIP_OK(hipEventRecord(e0, s));
                        for (int q = 0; q < kSteps; q++)
                            launch_term4(G, s, (int)F_ACC, (int)F_D0, (int)F_D1, 2, 0, false);
                        launch_term4(G, s, (int)F_ACC, (int)F_D0, (int)F_D1, 2, 1, false);
                    }
                    HIP_OK(hipEventRecord(e1, s));
                    HIP_OK(hipEventSynchronize(e1));
                    float ms = 0;
                    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = TermShape{shift, rpt, z};
                    }
                }
            cache[k] = best;
        }
        HIP_OK(hipEventDestroy(e0));
        HIP_OK(hipEventDestroy(e1));
        const int zero[2] = {0, 0};
        h2d(G.d_trec, saved.data(), saved.size() * sizeof(double), s);
        h2d(G.d_kmax, zero, sizeof zero, s);
    }
    for (ProbDesc &h : desc) {
        const auto pin = pins.find(grid_key(h));
        const TermShape t = pin != pins.end() ? pin->second : cache.at(resolve(key_of(h)));
        set_term_tiling(h, std::min(t.shift, ilog2(h.nc)), t.rpt, t.z);
    }
    up(G.d_desc, desc.data(), desc.size() * sizeof(ProbDesc), s);
    std::vector<WorkItem> w;
    const int per = term_items(desc, w);
    upload_term_items(G, w, per, cap, pool, s, up);
    G.tz = term_form(desc);
    G.tg = term_groups(desc);
    G.zfirst = first_z_ok(desc);
    G.zpow = G.zfirst && zpow_ok(desc);
}

// Pins the float4 sweep tiling of the model's grid shape (layers, rows,
// columns) in place of the tuner's choice, for the groups set up after the
// call; shift 0 drops the pin, and the next setup times the candidates.
static void pin_term_tiling(const Model &M, int shift, int rpt, int z) {
    const grid_model_t &g = *M.g;
    const GridKey k{g.n_layers, g.rows, g.cols};
    if (shift == 0) {
        term_shape_pins().erase(k);
        return;
    }
    if (shift < 4 || shift > 8 || g.cols % (1 << shift) != 0)
        fatal("sweep tiling: the tile width 1 << shift must be 16 to 256 and divide the grid's columns\n");
    if (rpt != 4 && rpt != 8) fatal("sweep tiling: rows per thread must be 4 or 8\n");
    if (z != 0 && z != 1) fatal("sweep tiling: z must be 0 (per-layer tiles) or 1 (z-marching column tiles)\n");
    term_shape_pins()[k] = TermShape{shift, rpt, z};
}


// reduction slots a problem needs: layer tiles or column tiles, + package item
static int slot_count(const ProbDesc &h) { return std::max(h.ntiles, h.nti2 * h.ntj) + 1; }

// The sink's package nodes' conductances to ambient (the sink's r_convec
// split by area; single_iteration_steady_pack :2120-2228).
static void sink_ambient(ProbDesc &h, const package_RC_t &pk) {
    for (int o : {SINK_N, SINK_S, SINK_W, SINK_E}) h.x_gamb[o] = 1.0 / (pk.r_hs_per + pk.r_amb_per);
    h.x_gamb[SINK_C_N] = h.x_gamb[SINK_C_S] = 1.0 / (pk.r_hs_c_per_y + pk.r_amb_c_per_y);
    h.x_gamb[SINK_C_W] = h.x_gamb[SINK_C_E] = 1.0 / (pk.r_hs_c_per_x + pk.r_amb_c_per_x);
}

// Layer and package-node coefficients of the descriptor from the host model.
static void fill_coefficients(Model &M, ProbDesc &h) {
    grid_model_t &g = *M.g;
    const package_RC_t &pk = g.pack;
    const int nl = g.n_layers, nr = g.rows, nc = g.cols;
    const double cw = g.width / nc, ch = g.height / nr;
    h.nl = nl;
    h.nr = nr;
    h.nc = nc;
    h.nx = M.nx;
    h.ambient = g.config.ambient;
    h.cell_area = (g.width * g.height) / (nc * nr);
    for (int n = 0; n < kMaxLayers; n++)
        for (int e = 0; e < 4; e++) h.ring_node[n][e] = -1;
    for (int n = 0; n < nl; n++) {
        const layer_t &L = g.layers[n];
        // row neighbours use rx, column neighbours ry; the vertical link
        // between n and n+1 uses the upper layer's full rz (find_res :116-132)
        h.gx[n] = 1.0 / L.rx;
        h.gy[n] = 1.0 / L.ry;
        h.gz[n] = n < nl - 1 ? 1.0 / L.rz : 0.0;
        h.cinv[n] = g.c_ready ? 1.0 / L.c : 0.0;
        h.gamb[n] = 0.0;
    }
    h.gamb[M.hsidx] = 1.0 / g.layers[M.hsidx].rz;
    if (g.config.model_secondary)
        h.gamb[M.pcbidx] = 1.0 / (g.config.r_convec_sec * (g.config.s_pcb * g.config.s_pcb) / (cw * ch));
    // package rings: an edge cell reaches its node through half its own
    // lateral resistance plus the periphery resistance shared by nc (nr) cells
    auto ring = [&](int n, int nn, int ns, int ne, int nw, double r1y, double r1x) {
        const layer_t &L = g.layers[n];
        h.ring_gns[n] = 1.0 / (L.ry / 2.0 + nc * r1y);
        h.ring_gew[n] = 1.0 / (L.rx / 2.0 + nr * r1x);
        h.ring_node[n][0] = nn;
        h.ring_node[n][1] = ns;
        h.ring_node[n][2] = ne;
        h.ring_node[n][3] = nw;
        const int nodes[4] = {nn, ns, ne, nw};
        for (int e = 0; e < 4; e++) {
            h.x_ring_layer[nodes[e]] = n;
            h.x_ring_edge[nodes[e]] = e;
            h.x_ring_g[nodes[e]] = e < 2 ? h.ring_gns[n] : h.ring_gew[n];
        }
    };
    for (int k = 0; k < kMaxExtra; k++) {
        h.x_ring_layer[k] = -1;
        h.x_ring_edge[k] = 0;
        h.x_ring_g[k] = 0.0;
        h.x_ncpl[k] = 0;
        h.x_gamb[k] = 0.0;
        h.x_cinv[k] = 0.0;
    }
    ring(M.spidx, SP_N, SP_S, SP_E, SP_W, pk.r_sp1_y, pk.r_sp1_x);
    ring(M.hsidx, SINK_C_N, SINK_C_S, SINK_C_E, SINK_C_W, pk.r_hs1_y, pk.r_hs1_x);
    auto cpl = [&](int a, int b, double gval, int steady_only_a = 0) {
        int q = h.x_ncpl[a]++;
        h.x_cpl_node[a][q] = b;
        h.x_cpl_g[a][q] = gval;
        h.x_cpl_steady_only[a][q] = steady_only_a;
        q = h.x_ncpl[b]++;
        h.x_cpl_node[b][q] = a;
        h.x_cpl_g[b][q] = gval;
        h.x_cpl_steady_only[b][q] = 0;
    };
    // package network (single_iteration_steady_pack :2120-2228)
    sink_ambient(h, pk);
    cpl(SINK_N, SINK_C_N, 1.0 / (pk.r_hs2_y + pk.r_hs));
    cpl(SINK_S, SINK_C_S, 1.0 / (pk.r_hs2_y + pk.r_hs));
    cpl(SINK_W, SINK_C_W, 1.0 / (pk.r_hs2_x + pk.r_hs));
    cpl(SINK_E, SINK_C_E, 1.0 / (pk.r_hs2_x + pk.r_hs));
    cpl(SINK_C_N, SP_N, 1.0 / pk.r_sp_per_y);
    cpl(SINK_C_S, SP_S, 1.0 / pk.r_sp_per_y);
    cpl(SINK_C_W, SP_W, 1.0 / pk.r_sp_per_x);
    cpl(SINK_C_E, SP_E, 1.0 / pk.r_sp_per_x);
    if (g.config.model_secondary) {
        ring(M.subidx, SUB_N, SUB_S, SUB_E, SUB_W, pk.r_sub1_y, pk.r_sub1_x);
        ring(M.solderidx, SOLDER_N, SOLDER_S, SOLDER_E, SOLDER_W, pk.r_solder1_y, pk.r_solder1_x);
        ring(M.pcbidx, PCB_C_N, PCB_C_S, PCB_C_E, PCB_C_W, pk.r_pcb1_y, pk.r_pcb1_x);
        for (int o : {PCB_N, PCB_S, PCB_W, PCB_E}) h.x_gamb[o] = 1.0 / pk.r_amb_sec_per;
        cpl(PCB_N, PCB_C_N, 1.0 / (pk.r_pcb2_y + pk.r_pcb));
        cpl(PCB_S, PCB_C_S, 1.0 / (pk.r_pcb2_y + pk.r_pcb));
        cpl(PCB_W, PCB_C_W, 1.0 / (pk.r_pcb2_x + pk.r_pcb));
        cpl(PCB_E, PCB_C_E, 1.0 / (pk.r_pcb2_x + pk.r_pcb));
        h.x_gamb[PCB_C_N] = h.x_gamb[PCB_C_S] = 1.0 / pk.r_amb_sec_c_per_y;
        h.x_gamb[PCB_C_W] = h.x_gamb[PCB_C_E] = 1.0 / pk.r_amb_sec_c_per_x;
        cpl(PCB_C_N, SOLDER_N, 1.0 / pk.r_pcb_c_per_y);
        cpl(PCB_C_S, SOLDER_S, 1.0 / pk.r_pcb_c_per_y);
        cpl(PCB_C_W, SOLDER_W, 1.0 / pk.r_pcb_c_per_x);
        cpl(PCB_C_E, SOLDER_E, 1.0 / pk.r_pcb_c_per_x);
        // the solder nodes' transient equations leave out the substrate link
        // that their steady equations include (slope_fn_pack :3034-3070 vs
        // single_iteration_steady_pack :2314-2352): steady-only on that side
        cpl(SOLDER_N, SUB_N, 1.0 / pk.r_solder_per_y, 1);
        cpl(SOLDER_S, SUB_S, 1.0 / pk.r_solder_per_y, 1);
        cpl(SOLDER_W, SUB_W, 1.0 / pk.r_solder_per_x, 1);
        cpl(SOLDER_E, SUB_E, 1.0 / pk.r_solder_per_x, 1);
    }
    if (g.c_ready) {
        // package-node capacitances (slope_fn_pack :2884-3091)
        double c[kMaxExtra] = {0};
        c[SINK_N] = c[SINK_S] = c[SINK_W] = c[SINK_E] = pk.c_hs_per + pk.c_amb_per;
        c[SINK_C_N] = c[SINK_C_S] = pk.c_hs_c_per_y + pk.c_amb_c_per_y;
        c[SINK_C_W] = c[SINK_C_E] = pk.c_hs_c_per_x + pk.c_amb_c_per_x;
        c[SP_N] = c[SP_S] = pk.c_sp_per_y;
        c[SP_W] = c[SP_E] = pk.c_sp_per_x;
        if (g.config.model_secondary) {
            c[PCB_N] = c[PCB_S] = c[PCB_W] = c[PCB_E] = pk.c_pcb_per + pk.c_amb_sec_per;
            c[PCB_C_N] = c[PCB_C_S] = pk.c_pcb_c_per_y + pk.c_amb_sec_c_per_y;
            c[PCB_C_W] = c[PCB_C_E] = pk.c_pcb_c_per_x + pk.c_amb_sec_c_per_x;
            c[SOLDER_N] = c[SOLDER_S] = pk.c_solder_per_y;
            c[SOLDER_W] = c[SOLDER_E] = pk.c_solder_per_x;
            c[SUB_N] = c[SUB_S] = pk.c_sub_per_y;
            c[SUB_W] = c[SUB_E] = pk.c_sub_per_x;
        }
        for (int k = 0; k < M.nx; k++) h.x_cinv[k] = 1.0 / c[k];
    }
}

// Per-cell resistances of a detailed-3D layer: the list head's values where
// the cell has them, else the layer's (find_res_3D :144-173, find_cap_3D
// :176-186).
static void cell_resistances(const Model &M, int n, double *rx, double *ry, double *rz, double *cap) {
    const grid_model_t &g = *M.g;
    const layer_t &L = g.layers[n];
    const long plane = (long)g.rows * g.cols;
    const int mi = M.layer_map[n];
    const BlockMap &bm = M.maps[mi];
    const bool cellwise = M.detailed && !M.map_pkg[mi];
    for (long c = 0; c < plane; c++) {
        const bool hr = cellwise && bm.has_res[c];
        rx[c] = hr ? bm.rx[c] : L.rx;
        ry[c] = hr ? bm.ry[c] : L.ry;
        rz[c] = hr ? bm.rz[c] : L.rz;
        cap[c] = (cellwise && bm.lock[c]) ? bm.cap[c] : L.c;
    }
}

// The pressure network of one microchannel layer on the device (k_pr_*):
// CG iterations in groups of 32 launches between checks of the done flag.
static void pressure_on_device(hipStream_t stream, DevPool *pool, int nr, int nc, const std::vector<int> &types,
                               double pump, std::vector<double> &p) {
    const long n = (long)nr * nc;
    const int nb = (int)((n + kBlock - 1) / kBlock);
    std::vector<void *> bufs;
    auto alloc = [&](size_t bytes) {
        void *q = dmalloc(pool, bytes);
        bufs.push_back(q);
        return q;
    };
    struct Free {
        std::vector<void *> &b;
        DevPool *pool;
        hipStream_t s;
        ~Free() {
            (void)hipStreamSynchronize(s);
            for (void *q : b) dfree(pool, q);
        }
    } release{bufs, pool, stream};
    int *t = (int *)alloc(n * sizeof(int));
    double *x = (double *)alloc(n * sizeof(double)), *d = (double *)alloc(n * sizeof(double));
    double *r = (double *)alloc(n * sizeof(double)), *z = (double *)alloc(n * sizeof(double));
    double *s = (double *)alloc(n * sizeof(double)), *q = (double *)alloc(n * sizeof(double));
    double *part = (double *)alloc(2 * nb * sizeof(double));
    PrState *st = (PrState *)alloc(sizeof(PrState));
    HIP_OK(hipMemcpyAsync(t, types.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemsetAsync(q, 0, n * sizeof(double), stream));
    launch(k_pr_setup, nb, stream, (const int *)t, nr, nc, pump, x, d, r, z, s, part);
    launch(k_pr_scalars, 1, stream, (const double *)part, nb, 0, st);
    PrState hs{};
    for (int it = 0; it < 100000; it += 32) {
        for (int k = 0; k < 32; k++) {
            launch(k_pr_apply, nb, stream, (const int *)t, nr, nc, (const double *)d, (const double *)s, q, part,
                   (const PrState *)st);
            launch(k_pr_scalars, 1, stream, (const double *)part, nb, 1, st);
            launch(k_pr_update, nb, stream, (const int *)t, n, (const double *)d, (const double *)s,
                   (const double *)q, x, r, z, part, (const PrState *)st);
            launch(k_pr_scalars, 1, stream, (const double *)part, nb, 2, st);
            launch(k_pr_dir, nb, stream, (const int *)t, n, (const double *)z, s, (const PrState *)st);
        }
        HIP_OK(hipMemcpyAsync(&hs, st, sizeof hs, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        if (hs.done) break;
    }
    if (!hs.done) fatal("the microchannel pressure network did not converge\n");
    p.assign(n, 0.0);
    HIP_OK(hipMemcpyAsync(p.data(), x, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
}

static void assemble_hetero(Model &M, DevModel &D) {
    grid_model_t &g = *M.g;
    const int nl = g.n_layers, nr = g.rows, nc = g.cols;
    const long plane = (long)nr * nc, ncell = plane * nl;
    std::vector<double> rx(ncell), ry(ncell), rz(ncell), cap(ncell), press(ncell, 0.0);
    std::vector<int> fluid(ncell, 0), types(ncell, 0);
    std::vector<double> htc(nl, 0.0), thick(nl), hydro(nl, 0.0), cv(nl, 0.0), tin(nl, 0.0);
    bool adv = false;
    for (int n = 0; n < nl; n++) {
        cell_resistances(M, n, &rx[n * plane], &ry[n * plane], &rz[n * plane], &cap[n * plane]);
        thick[n] = g.layers[n].thickness;
        if (M.mc[n]) {
            Microchannel &mc = *M.mc[n];
            // solved once per model, on its first device build
            if (mc.pressure.empty())
                pressure_on_device(D.stream, D.pool, nr, nc, mc.types, mc.cfg.pumping_pressure, mc.pressure);
            htc[n] = mc.cfg.htc;
            hydro[n] = mc.hydro_c;
            cv[n] = mc.cfg.coolant_capac;
            tin[n] = mc.cfg.inlet_temperature;
            for (long c = 0; c < plane; c++) {
                const int t = mc.types[c];
                types[n * plane + c] = t;
                fluid[n * plane + c] = (t == CELL_FLUID || t == CELL_INLET || t == CELL_OUTLET);
                press[n * plane + c] = mc.pressure[c];
            }
            adv = true;
        }
    }
    AsmIn in;
    in.rx = D.upload(rx.data(), ncell);
    in.ry = D.upload(ry.data(), ncell);
    in.rz = D.upload(rz.data(), ncell);
    in.cap = D.upload(cap.data(), ncell);
    in.fluid = D.upload(fluid.data(), ncell);
    in.types = D.upload(types.data(), ncell);
    in.press = D.upload(press.data(), ncell);
    in.htc = D.upload(htc.data(), nl);
    in.thick = D.upload(thick.data(), nl);
    in.hydro = D.upload(hydro.data(), nl);
    in.cv = D.upload(cv.data(), nl);
    in.tin = D.upload(tin.data(), nl);
    in.nl = nl;
    in.nr = nr;
    in.nc = nc;
    in.detailed = M.detailed;
    in.superlu_outlets = M.superlu_outlets ? 1 : 0;
    in.cw = g.width / nc;
    in.ch = g.height / nr;
    ProbDesc &h = D.h;
    double *gE = D.alloc<double>(ncell), *gS = D.alloc<double>(ncell), *gD = D.alloc<double>(ncell);
    double *ci = D.alloc<double>(ncell), *di = D.alloc<double>(ncell);
    double *aN = D.alloc<double>(ncell), *aS = D.alloc<double>(ncell), *aE = D.alloc<double>(ncell);
    double *aW = D.alloc<double>(ncell), *aSelf = D.alloc<double>(ncell), *bc = D.alloc<double>(ncell);
    const int nb = (int)((ncell + kBlock - 1) / kBlock);
    launch(k_assemble, nb, D.stream, in, gE, gS, gD, ci, aN, aS, aE, aW, aSelf, bc);
    h.gE = gE;
    h.gS = gS;
    h.gD = gD;
    h.cinvc = ci;
    h.dinvc = di;
    h.aN = aN;
    h.aS = aS;
    h.aE = aE;
    h.aW = aW;
    h.aSelf = aSelf;
    h.bconst = adv ? bc : nullptr;
    h.has_adv = adv;
    float *fE = D.alloc<float>(ncell), *fS = D.alloc<float>(ncell), *fD = D.alloc<float>(ncell);
    float *fci = D.alloc<float>(ncell), *fdi = D.alloc<float>(ncell);
    float *faN = D.alloc<float>(ncell), *faS = D.alloc<float>(ncell), *faE = D.alloc<float>(ncell);
    float *faW = D.alloc<float>(ncell), *faSelf = D.alloc<float>(ncell);
    h.fE = fE;
    h.fS = fS;
    h.fD = fD;
    h.fcinv = fci;
    h.fdinv = fdi;
    h.faN = faN;
    h.faS = faS;
    h.faE = faE;
    h.faW = faW;
    h.faSelf = faSelf;
    HIP_OK(hipMemcpyAsync(D.d_desc, &h, sizeof(ProbDesc), hipMemcpyHostToDevice, D.stream));
    launch(k_coef_finish, nb, D.stream, (const ProbDesc *)D.d_desc, ncell, fE, fS, fD, fci, faN, faS,
           faE, faW, faSelf, di, fdi);
    HIP_OK(hipStreamSynchronize(D.stream));
}

// ||C^-1 G||_inf bound: each row's |diag| + sum|off-diag| over its capacitance
__global__ void k_rho(const ProbDesc *__restrict__ D, long ncell, unsigned long long *out) {
    const ProbDesc &d = D[0];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (c < ncell) {
        const long plane = (long)d.nr * d.nc;
        const int n = (int)(c / plane);
        const long cc = c - n * plane;
        const int i = (int)(cc / d.nc), j = (int)(cc % d.nc);
        double off = 0.0;
        if (i > 0) off += d.gS[c - d.nc] + fabs(d.aN[c]);
        if (i < d.nr - 1) off += d.gS[c] + fabs(d.aS[c]);
        if (j < d.nc - 1) off += d.gE[c] + fabs(d.aE[c]);
        if (j > 0) off += d.gE[c - 1] + fabs(d.aW[c]);
        if (n > 0) off += d.gD[c - plane];
        if (n < d.nl - 1) off += d.gD[c];
        const int *rn = d.ring_node[n];
        if (i == 0 && rn[0] >= 0) off += d.ring_gns[n];
        if (i == d.nr - 1 && rn[1] >= 0) off += d.ring_gns[n];
        if (j == d.nc - 1 && rn[2] >= 0) off += d.ring_gew[n];
        if (j == 0 && rn[3] >= 0) off += d.ring_gew[n];
        v = (fabs(1.0 / d.dinvc[c]) + off) * d.cinvc[c];
    }
    v = wave_max(v);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

static double host_rho(const Model &M, const ProbDesc &h) {
    // homogeneous layers: the largest row sum is at a ring corner cell
    double rho = 0.0;
    for (int n = 0; n < h.nl; n++) {
        double diag = 2 * h.gx[n] + 2 * h.gy[n] + h.gamb[n];
        if (n > 0) diag += h.gz[n - 1];
        if (n < h.nl - 1) diag += h.gz[n];
        if (h.ring_node[n][0] >= 0) diag += h.ring_gns[n] + h.ring_gew[n];
        rho = std::max(rho, 2 * diag * h.cinv[n]);
    }
    return rho;
}

static double package_rho(const Model &M, const ProbDesc &h) {
    double rho = 0.0;
    for (int k = 0; k < M.nx; k++) {
        double diag = h.x_gamb[k];
        if (h.x_ring_layer[k] >= 0) diag += h.x_ring_g[k] * (h.x_ring_edge[k] < 2 ? h.nc : h.nr);
        for (int q = 0; q < h.x_ncpl[k]; q++) diag += h.x_cpl_g[k][q];
        rho = std::max(rho, 2 * diag * h.x_cinv[k]);
    }
    return rho;
}

// The per-job state of a device model: the FP64 and FP32 vectors, the
// block staging vectors, reductions, scalars, the descriptor on the device
// and the single-model group's Taylor buffers (every allocation zeroed).
static void alloc_state(DevModel &D) {
    ProbDesc &h = D.h;
    for (int b = 0; b < NB64; b++) h.b64[b] = D.alloc<double>(D.vlen);
    for (int b = 0; b < NB32; b++) h.b32[b] = D.alloc<float>(D.vlen);
    h.lzero = D.alloc<int>(kMaxLayers);
    h.blk_in = D.alloc<double>(D.nblk);
    h.blk_out = D.alloc<double>(D.nblk);
    D.h_blk = (double *)hmalloc(D.hpool, D.nblk * sizeof(double));
    D.h_scal = (double *)hmalloc(D.hpool, NSCAL * sizeof(double));
    h.partial = D.alloc<double>((size_t)slot_count(h) * kRed);
    h.counter = D.alloc<unsigned>(1);
    h.scal = D.alloc<double>(NSCAL);
    h.seq_in = h.seq_out = nullptr;
    h.seq_one = 0;
    h.gm = h.gpart = nullptr;
    h.gmv32 = nullptr;
    h.gmv64 = nullptr;
}

// The single-model group of a device model (its descriptor and work-lists)
// with its Taylor buffers and the float4 term work-list.
static void single_group(DevModel &D) {
    D.grp.d_desc = D.d_desc;
    D.grp.nprob = 1;
    D.grp.d_work = D.d_work;
    D.grp.nwork = D.nwork;
    D.grp.d_gwork = D.d_gwork;
    D.grp.ngwork = D.ngwork;
    D.grp.stream = D.stream;
    D.grp.hetero = D.hetero;
    D.grp.vec4 = vec4_ok(D.h);
    D.grp.d_kmax = D.alloc<int>(2);
    D.grp.d_trec = D.alloc<double>(kTrec);
    D.grp.nper = D.nwork;
    std::vector<WorkItem> tw;
    const int per = term_items({D.h}, tw);
    upload_term_items(D.grp, tw, per, D.cap_twork, D.pool, D.stream);
}

// Device state of a model: coefficients, maps, work-lists and the solver
// buffers.  With `shared` the launches go to that stream (the engine's);
// with pools, device and pinned host memory come from them.  Without
// with_state only the read-only data is built (a template for dev_clone).
static DevModel *dev_build(Model &M, hipStream_t shared, DevPool *pool = nullptr, DevPool *hpool = nullptr,
                           bool with_state = true) {
    grid_model_t &g = *M.g;
    if (g.n_layers > kMaxLayers) fatal("too many layers for the device model\n");
    // the kernels address a vector through 32-bit byte offsets (8 B a node)
    if ((double)M.ncell() + M.nx + 64 >= (double)(1u << 29))
        fatal("grid too large for the device model: layers x rows x columns must stay below 2^29 nodes\n");
    std::unique_ptr<DevModel> own(new DevModel());
    DevModel *D = own.get();
    D->pool = pool;
    D->hpool = hpool;
    if (shared) {
        D->stream = shared;
        D->own_stream = false;
    } else {
        HIP_OK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
    }
    HIP_OK(hipEventCreate(&D->ev0));
    HIP_OK(hipEventCreate(&D->ev1));
    ProbDesc &h = D->h;
    memset(&h, 0, sizeof h);
    fill_coefficients(M, h);
    set_tiling(h);
    D->hetero = M.detailed;
    D->has_c = g.c_ready;
    h.hetero = D->hetero;
    h.map_mode = g.map_mode;
    const long ncell = M.ncell();
    D->vlen = ((ncell + M.nx + 63) / 64) * 64;
    h.vlen = (unsigned)D->vlen;
    D->nblk = g.total_n_blocks + M.nx;
    if (with_state) alloc_state(*D);
    // block <-> grid maps
    std::vector<int> base;
    layer_bases(M, base);
    std::vector<const int *> dptr(M.maps.size()), dunit(M.maps.size());
    std::vector<const double *> docc(M.maps.size());
    for (size_t mi = 0; mi < M.maps.size(); mi++) {
        const BlockMap &bm = M.maps[mi];
        dptr[mi] = D->upload(bm.cell_ptr.data(), bm.cell_ptr.size());
        // one spare entry: k_b2g reads a cell's first entry even when it has none
        std::vector<int> u(bm.ent_unit);
        std::vector<double> o(bm.ent_occ);
        u.push_back(0);
        o.push_back(0.0);
        dunit[mi] = D->upload(u.data(), u.size());
        docc[mi] = D->upload(o.data(), o.size());
    }
    std::vector<int4> rng;
    std::vector<int> blayer;
    std::vector<const int4 *> dcu(M.maps.size(), nullptr);
    std::vector<const unsigned short *> dcode(M.maps.size(), nullptr);
    std::vector<const float *> dw(M.maps.size(), nullptr);
    std::vector<const int *> dmoff(M.maps.size(), nullptr), dmcell(M.maps.size(), nullptr);
    std::vector<const int4 *> dmuw(M.maps.size(), nullptr);
    // per map: the weights of one-block cells (host copy), and whether some
    // cell is covered by no block or by five or more
    std::vector<std::vector<float>> hw(M.maps.size());
    std::vector<char> odd(M.maps.size(), 0);
    for (int n = 0; n < g.n_layers; n++) {
        const int mi = M.layer_map[n];
        h.map_ptr[n] = dptr[mi];
        h.map_unit[n] = dunit[mi];
        h.map_occ[n] = docc[mi];
        h.layer_base[n] = base[n];
        const flp_t *f = g.layers[n].flp;
        std::vector<double> area(f->n_units);
        for (int u = 0; u < f->n_units; u++) {
            area[u] = f->units[u].width * f->units[u].height;
            const glist_t &r = M.maps[mi].ranges[u];
            rng.push_back(make_int4(r.i1, r.i2, r.j1, r.j2));
            blayer.push_back(n);
        }
        h.unit_area[n] = D->upload(area.data(), area.size());
        if (!dcu[mi]) {   // layers on one map share its floorplan, hence the areas
            const BlockMap &bm = M.maps[mi];
            const size_t plane = bm.cell_ptr.size() - 1;
            // per cell {units}, {weights}: up to four blocks, unused slots
            // repeat the first block with weight 0; a cell of five or more
            // blocks has unit -1 and takes its map entries
            std::vector<int4> uw(2 * plane, make_int4(-1, -1, -1, -1));
            auto bits = [](float f) {
                int b;
                memcpy(&b, &f, sizeof b);
                return b;
            };
            for (size_t c = 0; c < plane; c++) {
                const int e = bm.cell_ptr[c], cnt = bm.cell_ptr[c + 1] - e;
                if (cnt < 1 || cnt > 4) {
                    odd[mi] = 1;
                    continue;
                }
                int u[4];
                float w[4];
                for (int q = 0; q < 4; q++) {
                    u[q] = bm.ent_unit[e + (q < cnt ? q : 0)];
                    w[q] = q < cnt ? (float)(bm.ent_occ[e + q] * h.cell_area / area[u[q]]) : 0.f;
                }
                uw[2 * c] = make_int4(u[0], u[1], u[2], u[3]);
                uw[2 * c + 1] = make_int4(bits(w[0]), bits(w[1]), bits(w[2]), bits(w[3]));
            }
            dcu[mi] = D->upload(uw.data(), uw.size());
            if (f->n_units < (int)kCellMixed) {
                std::vector<unsigned short> code(plane, kCellMixed);
                for (size_t c = 0; c < plane; c++) {
                    const int e = bm.cell_ptr[c];
                    if (bm.cell_ptr[c + 1] - e == 1 && bm.ent_occ[e] == 1.0)
                        code[c] = (unsigned short)bm.ent_unit[e];
                }
                // the weight of a cell inside one block, as cell_uw has it
                std::vector<float> w(f->n_units);
                for (int u = 0; u < f->n_units; u++) w[u] = (float)(1.0 * h.cell_area / area[u]);
                dcode[mi] = D->upload(code.data(), code.size());
                dw[mi] = D->upload(w.data(), w.size());
                hw[mi] = w;
                // the boundary cells by tile of the standard tiling (rows of
                // TY * rpt, columns of TX; tile ti * ntj + tj)
                const int tx = 1 << h.tx_shift, rows = (kBlock >> h.tx_shift) * h.rpt;
                std::vector<int> off(1, 0), cell;
                std::vector<int4> rec;
                for (int ti = 0; ti < h.nti; ti++)
                    for (int tj = 0; tj < h.ntj; tj++) {
                        for (int i = ti * rows; i < std::min(g.rows, (ti + 1) * rows); i++)
                            for (int j = tj * tx; j < std::min(g.cols, (tj + 1) * tx); j++) {
                                const size_t c = (size_t)i * g.cols + j;
                                if (code[c] != kCellMixed) continue;
                                cell.push_back((int)c);
                                rec.push_back(uw[2 * c]);
                                rec.push_back(uw[2 * c + 1]);
                            }
                        off.push_back((int)cell.size());
                    }
                cell.push_back(0);   // no empty uploads
                rec.push_back(make_int4(0, 0, 0, 0));
                dmoff[mi] = D->upload(off.data(), off.size());
                dmcell[mi] = D->upload(cell.data(), cell.size());
                dmuw[mi] = D->upload(rec.data(), rec.size());
            }
        }
        h.cell_uw[n] = dcu[mi];
        h.cell_u16[n] = dcode[mi];
        h.unit_w[n] = dw[mi];
        h.mix_off[n] = dmoff[mi];
        h.mix_cell[n] = dmcell[mi];
        h.mix_uw[n] = dmuw[mi];
    }
    h.nblocks = g.total_n_blocks;
    h.blk_range = D->upload(rng.data(), rng.size());
    h.blk_layer = D->upload(blayer.data(), blayer.size());
    {
        std::vector<float> bw(g.total_n_blocks);
        bool ok = true;
        for (int n = 0; n < g.n_layers && ok; n++) {
            const int mi = M.layer_map[n];
            ok = dcode[mi] && !odd[mi] && base[n] + hw[mi].size() <= bw.size();
            if (ok) std::copy(hw[mi].begin(), hw[mi].end(), bw.begin() + base[n]);
        }
        h.blk_w = ok && !bw.empty() ? D->upload(bw.data(), bw.size()) : nullptr;
    }
    D->d_desc = D->alloc<ProbDesc>(1);
    h2d(D->d_desc, &h, sizeof(ProbDesc), D->stream);
    // work lists
    std::vector<WorkItem> w(h.ntiles + 1);
    for (int t = 0; t <= h.ntiles; t++) w[t] = WorkItem{0, t};
    D->nwork = h.ntiles + 1;
    D->d_work = D->upload(w.data(), w.size());
    std::vector<WorkItem> gw;
    for (int u = 0; u < g.total_n_blocks; u += kBlock / kWave) gw.push_back(WorkItem{0, u});
    gw.push_back(WorkItem{0, -1});
    D->ngwork = (int)gw.size();
    D->d_gwork = D->upload(gw.data(), gw.size());
    if (with_state) single_group(*D);
    if (D->hetero) assemble_hetero(M, *D);
    // transient step bound
    if (g.c_ready) {
        if (D->hetero) {
            unsigned long long *r = D->alloc<unsigned long long>(1);
            dzero(r, sizeof *r, D->stream);
            launch(k_rho, (int)((ncell + kBlock - 1) / kBlock), D->stream, (const ProbDesc *)D->d_desc,
                   ncell, r);
            unsigned long long bits = 0;
            d2h(&bits, r, sizeof bits, D->stream);
            double v;
            memcpy(&v, &bits, sizeof v);
            D->rho = v;
        } else {
            D->rho = host_rho(M, h);
        }
        D->rho = std::max(D->rho, package_rho(M, h));
    }
    h2d(D->d_desc, &h, sizeof(ProbDesc), D->stream);
    return own.release();
}

// A device model of the template's model that shares the template's
// read-only data and allocates only its own state: what a job on a model
// already resident on the device needs.
static DevModel *dev_clone(const std::shared_ptr<DevModel> &tpl, hipStream_t shared, DevPool *pool,
                           DevPool *hpool) {
    std::unique_ptr<DevModel> own(new DevModel());
    DevModel *D = own.get();
    D->base = tpl;
    D->pool = pool;
    D->hpool = hpool;
    D->stream = shared;
    D->own_stream = false;
    HIP_OK(hipEventCreate(&D->ev0));
    HIP_OK(hipEventCreate(&D->ev1));
    D->h = tpl->h;
    D->hetero = tpl->hetero;
    D->has_c = tpl->has_c;
    D->vlen = tpl->vlen;
    D->nblk = tpl->nblk;
    D->rho = tpl->rho;
    D->d_work = tpl->d_work;
    D->nwork = tpl->nwork;
    D->d_gwork = tpl->d_gwork;
    D->ngwork = tpl->ngwork;
    alloc_state(*D);
    D->d_desc = D->alloc<ProbDesc>(1);
    single_group(*D);
    h2d(D->d_desc, &D->h, sizeof(ProbDesc), D->stream);
    return own.release();
}

static void dev_prepare(Model &M) {
    grid_model_t &g = *M.g;
    if (!g.r_ready) fatal("R model not ready\n");
    if (M.dev && !M.dev_dirty && M.dev->has_c == (bool)g.c_ready) return;
    // keep the device state across a rebuild (warm starts)
    std::vector<double> keep_steady, keep_trans;
    if (M.dev && M.steady_valid) {
        keep_steady.resize(M.vec_len());
        dev_get_grid(M, 0, keep_steady.data());
    }
    if (M.dev && M.trans_valid) {
        keep_trans.resize(M.vec_len());
        dev_get_grid(M, 1, keep_trans.data());
    }
    dev_release(M);
    M.dev = dev_build(M, nullptr);
    M.dev_dirty = false;
    if (!keep_steady.empty()) dev_set_grid(M, 0, keep_steady.data());
    if (!keep_trans.empty()) dev_set_grid(M, 1, keep_trans.data());
}

static void read_scal(DevModel &D) {
    HIP_OK(hipMemcpyAsync(D.h_scal, D.h.scal, NSCAL * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

static void write_scal(DevModel &D, const std::vector<std::pair<int, double>> &kv) {
    read_scal(D);
    for (auto &p : kv) D.h_scal[p.first] = p.second;
    HIP_OK(hipMemcpyAsync(D.h.scal, D.h_scal, NSCAL * sizeof(double), hipMemcpyHostToDevice, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

// the given scalar slots set by a launch on the model's stream, without a
// host wait (the engine's phase changes)
static void set_scal(DevModel &D, const std::vector<std::pair<int, double>> &kv) {
    for (size_t k = 0; k < kv.size(); k += 16) {
        ScalSet s{};
        for (size_t q = k; q < std::min(kv.size(), k + 16); q++) {
            s.slot[s.n] = kv[q].first;
            s.val[s.n++] = kv[q].second;
        }
        hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(64), 0, D.stream, D.h.scal, s);
    }
}

static void upload_blocks(DevModel &D, const double *v) {
    memcpy(D.h_blk, v, D.nblk * sizeof(double));
    HIP_OK(hipMemcpyAsync(D.h.blk_in, D.h_blk, D.nblk * sizeof(double), hipMemcpyHostToDevice, D.stream));
}

static void download_blocks(DevModel &D, double *v) {
    HIP_OK(hipMemcpyAsync(D.h_blk, D.h.blk_out, D.nblk * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
    memcpy(v, D.h_blk, D.nblk * sizeof(double));
}

static inline int small_grid(int n) { return (n + 63) / 64; }

static void run_b2g(const Group &G, int type, int out, int sources, int from_seq = 0) {
    launch(k_b2g, G.nwork, G.stream, (const ProbDesc *)G.d_desc, (const WorkItem *)G.d_work, type, out,
           sources, from_seq);
}

// The power of an interval for the FP32 Taylor sweeps of homogeneous float4
// groups: cell power alone, in FP32 (F_PW), layers without power left
// unwritten; k_tay_first2 adds the ambient sources itself.
static bool power32(const Group &G, bool f32, bool het) { return f32 && !het && G.vec4; }
static void run_b2g_power32(const Group &G, int from_seq = 0) {
    launch(k_b2g_power32, G.nwork, G.stream, (const ProbDesc *)G.d_desc, (const WorkItem *)G.d_work, from_seq);
}

static void run_g2b(const Group &G, int in, int to_seq = 0) {
    launch(k_g2b, G.ngwork, G.stream, (const ProbDesc *)G.d_desc, (const WorkItem *)G.d_gwork, in, to_seq);
}

// ---- steady ----------------------------------------------------------------
static void finish(const Group &G, int op, const FinArgs &a) {
    launch(k_finish, G.nprob, G.stream, (const ProbDesc *)G.d_desc, op, a);
}

template <typename T, bool HET>
static void cg_iterations(const Group &G, int n, int xb, int rb, int zb, int p0b, int qb, int done_slot,
                          int tol_slot, int max_iters) {
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    FinArgs a{};
    a.done_slot = done_slot;
    a.tol_slot = tol_slot;
    a.max_iters = max_iters;
    for (int k = 0; k < n; k++) {
        launch_named("k_cg_dir", k_cg_dir<T, HET>, G.nwork, G.stream, dd, w, zb, p0b, qb, done_slot);
        finish(G, FIN_ALPHA, a);
        launch_named("k_cg_upd", k_cg_upd<T, HET>, G.nwork, G.stream, dd, w, xb, rb, zb, p0b, qb, done_slot,
                     tol_slot, max_iters, 0);
        finish(G, FIN_BETA, a);
    }
}

// FP64 Jacobi-PCG: r = rhs - G x, z, and the CG scalars of every problem
template <bool HET>
static void jpcg64_start(const Group &G, double rtol2, int max_iters, int xb = B_STEADY, int rhsb = B_RHS) {
    launch_named("k_cg_init", k_cg_init<double, HET>, G.nwork, G.stream, (const ProbDesc *)G.d_desc,
                 (const WorkItem *)G.d_work, xb, (int)B_R, (int)B_Z, rhsb, rtol2, max_iters);
    FinArgs a{};
    a.rtol2 = rtol2;
    a.max_iters = max_iters;
    finish(G, FIN_CG_INIT, a);
}

// GMRES state and basis of a problem, allocated on its first non-symmetric
// steady solve in the given precision.
static void ensure_gmres(DevModel &D, bool f32) {
    bool changed = false;
    if (!D.h.gm) {
        D.h.gm = D.alloc<double>(GM_SIZE);
        D.h.gpart = D.alloc<double>((size_t)(D.h.ntiles + 1) * (kGmM + 1));
        changed = true;
    }
    if (f32 && !D.h.gmv32) {
        D.h.gmv32 = D.alloc<float>((size_t)(kGmM + 1) * D.vlen);
        changed = true;
    }
    if (!f32 && !D.h.gmv64) {
        D.h.gmv64 = D.alloc<double>((size_t)(kGmM + 1) * D.vlen);
        changed = true;
    }
    if (changed) h2d(D.d_desc, &D.h, sizeof(ProbDesc), D.stream);
}

// One refinement round of the non-symmetric solve: problems whose GMRES
// cycle has finished take its correction and start the next cycle from a
// fresh FP64 residual (or stop); then `iters` Arnoldi iterations.
template <typename T, bool HET>
static void gm_round(const Group &G, double rtol2, double inner2, int iters) {
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    const bool f32 = sizeof(T) == 4;
    const int zb = f32 ? F_Z : B_Z, wb = f32 ? F_Q : B_Q;
    launch_named("k_gm_correct", k_gm_correct<T, HET>, G.nwork, G.stream, dd, w, -1);
    launch_named("k_gm_resid", k_gm_resid<T, HET>, G.nwork, G.stream, dd, w, zb, 0);
    FinArgs a{};
    a.rtol2 = rtol2;
    a.inner_rtol2 = inner2;
    a.max_refine = 2000;
    finish(G, FIN_GM_IR, a);
    for (int k = 0; k < iters; k++) {
        launch_named("k_gm_apply", k_gm_apply<T, HET>, G.nwork, G.stream, dd, w, zb, wb);
        launch_named("k_gm_dots", k_gm_dots<T>, G.nwork, G.stream, dd, w, wb);
        finish(G, FIN_GM_H, a);
        launch_named("k_gm_orth", k_gm_orth<T, HET>, G.nwork, G.stream, dd, w, wb, zb, 0);
        finish(G, FIN_GM_NORM, a);
    }
}


constexpr double kInnerRtol2 = 1e-8;   // each FP32 solve gains 4 digits
constexpr int kMaxRefine = 30;

// ---- multigrid -------------------------------------------------------------
constexpr float kMgOmega = 0.8f;   // line-Jacobi damping
constexpr int kMgCoarseSweeps = 8;

static std::vector<WorkItem> column_items(const ProbDesc &h, int prob) {
    std::vector<WorkItem> w;
    for (int t = 0; t < h.nti2 * h.ntj; t++) w.push_back(WorkItem{prob, t});
    w.push_back(WorkItem{prob, h.nti2 * h.ntj});
    return w;
}

// Coarse levels down to about 4 x 4 cells per layer; coefficients built on
// the device from the level above.
static void build_mg(DevModel &D) {
    if (!D.mg.empty()) return;
    MgLevel l0;
    l0.h = D.h;
    l0.d_desc = D.d_desc;
    if (D.h.has_adv) {
        // the preconditioner's finest level: the upwind advection (k_upwind_coef)
        const long ncell = (long)D.h.nl * D.h.nr * D.h.nc;
        float *u[6];
        for (float *&p : u) p = D.alloc<float>(ncell);
        hipLaunchKernelGGL(k_upwind_coef, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, D.stream,
                           (const ProbDesc *)D.d_desc, ncell, u[0], u[1], u[2], u[3], u[4], u[5]);
        l0.h.faN = u[0];
        l0.h.faS = u[1];
        l0.h.faE = u[2];
        l0.h.faW = u[3];
        l0.h.faSelf = u[4];
        l0.h.fdinv = u[5];
        l0.d_desc = D.upload(&l0.h, 1);
    }
    auto w0 = column_items(D.h, 0);
    l0.d_cwork = D.upload(w0.data(), w0.size());
    l0.ncwork = (int)w0.size();
    D.mg.push_back(l0);
    bool prev_het = D.hetero;
    // coarsen while both sizes are even (always, for power-of-two grids):
    // every fine cell has its parent (i/2, j/2)
    while (D.mg.back().h.nr >= 8 && D.mg.back().h.nc >= 8 && D.mg.back().h.nr % 2 == 0 &&
           D.mg.back().h.nc % 2 == 0) {
        const ProbDesc f = D.mg.back().h;
        MgLevel L;
        ProbDesc &h = L.h;
        h = f;
        h.nr = f.nr / 2;
        h.nc = f.nc / 2;
        set_tiling(h);
        h.hetero = 1;
        for (int n = 0; n < h.nl; n++) {
            h.gamb[n] = f.gamb[n] * 4.0;
            h.ring_gns[n] = f.ring_gns[n] * 2.0;
            h.ring_gew[n] = f.ring_gew[n] * 2.0;
        }
        for (int k = 0; k < kMaxExtra; k++) h.x_ring_g[k] = f.x_ring_g[k] * 2.0;
        const size_t ncell = (size_t)h.nl * h.nr * h.nc, vlen = ((ncell + h.nx + 63) / 64) * 64;
        h.vlen = (unsigned)vlen;
        for (int b = 0; b < NB64; b++) h.b64[b] = nullptr;
        for (int b = 0; b < NB32; b++) h.b32[b] = nullptr;
        for (int b : {(int)F_R, (int)F_D0, (int)F_D1}) {
            h.b32[b] = D.alloc<float>(vlen);
            dzero(h.b32[b], vlen * sizeof(float), D.stream);
        }
        h.gE = h.gS = h.gD = h.cinvc = h.dinvc = nullptr;
        h.aN = h.aS = h.aE = h.aW = h.aSelf = nullptr;
        h.bconst = nullptr;
        h.fcinv = nullptr;
        h.fE = D.alloc<float>(ncell);
        h.fS = D.alloc<float>(ncell);
        h.fD = D.alloc<float>(ncell);
        h.fdinv = D.alloc<float>(ncell);
        if (h.has_adv) {
            h.faN = D.alloc<float>(ncell);
            h.faS = D.alloc<float>(ncell);
            h.faE = D.alloc<float>(ncell);
            h.faW = D.alloc<float>(ncell);
            h.faSelf = D.alloc<float>(ncell);
        } else {
            h.faN = h.faS = h.faE = h.faW = h.faSelf = nullptr;
        }
        h.partial = D.alloc<double>((size_t)slot_count(h) * kRed);
        h.gm = h.gpart = nullptr;
        h.gmv32 = nullptr;
        h.gmv64 = nullptr;
        h.seq_in = h.seq_out = nullptr;
        L.d_desc = D.upload(&h, 1);
        auto w = column_items(h, 0);
        L.d_cwork = D.upload(w.data(), w.size());
        L.ncwork = (int)w.size();
        hipLaunchKernelGGL(k_mg_coarsen, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, D.stream,
                           (const ProbDesc *)D.mg.back().d_desc, (const ProbDesc *)L.d_desc, (int)prev_het);
        prev_het = true;
        D.mg.push_back(L);
    }
    HIP_OK(hipStreamSynchronize(D.stream));
}

// The sink's convection resistance of a homogeneous model's device state
// set to the one behind rz (the sink layer's vertical resistance to
// ambient) and pk (the package resistances), as populate_R_model with that
// r_convec would give them: the sink layer's ambient conductance, the sink
// nodes' ones, and the multigrid levels coarsened again from them (coarse
// ambient conductances sum the four children's, k_mg_coarsen recomputes the
// coarse faces and diagonals).  The steady state stays: a re-solve starts
// from it.
static void dev_set_convec(DevModel &D, int hsidx, double rz, const package_RC_t &pk) {
    if (D.hetero || D.h.has_adv) fatal("the sink's convection is set on the device for homogeneous models only\n");
    ProbDesc &h = D.h;
    h.gamb[hsidx] = 1.0 / rz;
    if (hsidx < h.nl - 1) h.gz[hsidx] = 1.0 / rz;
    sink_ambient(h, pk);
    h2d(D.d_desc, &h, sizeof(ProbDesc), D.stream);
    if (!D.mg.empty()) {
        D.mg[0].h = h;
        for (size_t l = 1; l < D.mg.size(); l++) {
            const ProbDesc &f = D.mg[l - 1].h;
            ProbDesc &c = D.mg[l].h;
            for (int n = 0; n < c.nl; n++) c.gamb[n] = f.gamb[n] * 4.0;
            for (int k = 0; k < kMaxExtra; k++) c.x_gamb[k] = f.x_gamb[k];
            h2d(D.mg[l].d_desc, &c, sizeof(ProbDesc), D.stream);
            const size_t ncell = (size_t)c.nl * c.nr * c.nc;
            hipLaunchKernelGGL(k_mg_coarsen, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, D.stream,
                               (const ProbDesc *)D.mg[l - 1].d_desc, (const ProbDesc *)D.mg[l].d_desc, l > 1 ? 1 : 0);
        }
    }
    HIP_OK(hipStreamSynchronize(D.stream));
}

// The levels of a group: descriptor arrays and column work-lists per level.
struct MgGroup {
    const ProbDesc *d;
    const WorkItem *w;
    int nw, nl_max;
    bool het;
};

static void mg_smooth(const MgGroup &L, const MgGroup *C, hipStream_t s, int xin, int xcb, int xout, int bb,
                      int when, int dot) {
    const size_t lds = (size_t)2 * L.nl_max * kBlock * sizeof(float);
    const ProbDesc *dc = C ? C->d : L.d;
    if (L.het)
        hipLaunchKernelGGL(k_mg_smooth<true>, dim3(L.nw), dim3(kBlock), lds, s, L.d, dc, L.w, xin, xcb, xout, bb,
                           kMgOmega, when, dot);
    else
        hipLaunchKernelGGL(k_mg_smooth<false>, dim3(L.nw), dim3(kBlock), lds, s, L.d, dc, L.w, xin, xcb, xout, bb,
                           kMgOmega, when, dot);
}

// z = M b on the finest level (b0, z0 are FP32 buffer ids there): one V-cycle,
// one line-Jacobi sweep before and after each coarse correction; with dot,
// the finest sweep leaves the partial (b, z) in the reduction slots.
static void mg_xline(const MgGroup &L, hipStream_t s, int xin, int xout, int bb, int when) {
    if (L.het)
        hipLaunchKernelGGL(k_mg_xline<true>, dim3(L.nw), dim3(kBlock), 0, s, L.d, L.w, xin, xout, bb, kMgOmega, when);
    else
        hipLaunchKernelGGL(k_mg_xline<false>, dim3(L.nw), dim3(kBlock), 0, s, L.d, L.w, xin, xout, bb, kMgOmega,
                           when);
}

// xz: advection-dominated problems smooth with an x-line sweep before the
// z-line sweep (and after it on the way up).
static void vcycle(const std::vector<MgGroup> &lv, hipStream_t s, int when, int b0, int z0, bool dot,
                   bool xz = false) {
    const int L = (int)lv.size() - 1;
    std::vector<int> cur(L + 1);
    auto other = [](int b) { return b == F_D0 ? (int)F_D1 : (int)F_D0; };
    // pre-smoothing of level l from zero (x then z with xz)
    auto pre = [&](int l, int b) {
        if (xz) {
            mg_xline(lv[l], s, -1, F_D0, b, when);
            mg_smooth(lv[l], nullptr, s, F_D0, -1, F_D1, b, when, -1);
            cur[l] = F_D1;
        } else {
            mg_smooth(lv[l], nullptr, s, -1, -1, F_D0, b, when, -1);
            cur[l] = F_D0;
        }
    };
    for (int l = 0; l < L; l++) {
        const int b = l == 0 ? b0 : (int)F_R;
        pre(l, b);
        const MgGroup &f = lv[l], &c = lv[l + 1];
        if (f.het)
            hipLaunchKernelGGL(k_mg_restrict<true>, dim3(c.nw), dim3(kBlock), 0, s, f.d, c.d, c.w, cur[l], b,
                               (int)F_R, when);
        else
            hipLaunchKernelGGL(k_mg_restrict<false>, dim3(c.nw), dim3(kBlock), 0, s, f.d, c.d, c.w, cur[l], b,
                               (int)F_R, when);
    }
    const int bL = L == 0 ? b0 : (int)F_R;
    pre(L, bL);
    for (int k = 1; k < kMgCoarseSweeps; k++) {
        if (xz) {
            mg_xline(lv[L], s, cur[L], other(cur[L]), bL, when);
            cur[L] = other(cur[L]);
        }
        mg_smooth(lv[L], nullptr, s, cur[L], -1, other(cur[L]), bL, when, -1);
        cur[L] = other(cur[L]);
    }
    if (L == 0) {   // a single level: the sweeps are the preconditioner
        mg_smooth(lv[0], nullptr, s, cur[0], -1, z0, b0, when, dot ? 1 : -1);
        return;
    }
    for (int l = L - 1; l >= 0; l--) {
        const int b = l == 0 ? b0 : (int)F_R;
        if (xz) {   // prolong + z-line, then x-line
            mg_smooth(lv[l], &lv[l + 1], s, cur[l], cur[l + 1], other(cur[l]), b, when, -1);
            cur[l] = other(cur[l]);
            const int out = l == 0 ? z0 : other(cur[l]);
            mg_xline(lv[l], s, cur[l], out, b, when);
            cur[l] = out;
            continue;
        }
        const int out = l == 0 ? z0 : other(cur[l]);
        mg_smooth(lv[l], &lv[l + 1], s, cur[l], cur[l + 1], out, b, when, (l == 0 && dot) ? 1 : -1);
        cur[l] = out;
    }
}

static std::vector<MgGroup> mg_levels_of(const DevModel &D) {
    std::vector<MgGroup> lv;
    for (size_t l = 0; l < D.mg.size(); l++)
        lv.push_back(MgGroup{D.mg[l].d_desc, D.mg[l].d_cwork, D.mg[l].ncwork, D.h.nl, l > 0 || D.hetero});
    return lv;
}

// Multigrid-preconditioned CG (FP32) inside FP64 refinement, one round.
template <bool HET>
static void mixed_round_mg(const Group &G, const std::vector<MgGroup> &lv, double rtol2, int max_iters,
                           int chunk) {
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    launch(k_ir_correct, G.nwork, G.stream, dd, w);
    launch_named("k_ir_resid", k_ir_resid<HET>, G.nwork, G.stream, dd, w, rtol2, kInnerRtol2, kMaxRefine);
    FinArgs a{};
    a.rtol2 = rtol2;
    a.inner_rtol2 = kInnerRtol2;
    a.max_refine = kMaxRefine;
    a.mg = 1;
    finish(G, FIN_IR, a);
    vcycle(lv, G.stream, MG_INIT, F_R, F_Z, true);
    finish(G, FIN_MG_RZ0, a);
    a.done_slot = S_IDONE;
    a.tol_slot = S_ITOL2;
    a.max_iters = max_iters;
    for (int k = 0; k < chunk; k++) {
        launch_named("k_cg_dir", k_cg_dir<float, HET>, G.nwork, G.stream, dd, w, (int)F_Z, (int)F_P0, (int)F_Q,
                     (int)S_IDONE);
        finish(G, FIN_ALPHA, a);
        launch_named("k_cg_upd", k_cg_upd<float, HET>, G.nwork, G.stream, dd, w, (int)F_X, (int)F_R, (int)F_Z,
                     (int)F_P0, (int)F_Q, (int)S_IDONE, (int)S_ITOL2, max_iters, 1);
        finish(G, FIN_MG_RR, a);
        vcycle(lv, G.stream, MG_RUNNING, F_R, F_Z, true);
        finish(G, FIN_MG_BETA, a);
    }
}

// The same with the multigrid V-cycle (x- and z-line smoothing) as the right
// preconditioner, FP32: v_j goes to F_Z, M v_j to F_X; the cycle's
// correction is M (V y), added to x by k_ir_correct.
template <bool HET>
static void gm_round_mg(const Group &G, const std::vector<MgGroup> &lv, double rtol2, double inner2, int iters) {
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    launch_named("k_gm_correct", k_gm_correct<float, HET>, G.nwork, G.stream, dd, w, (int)F_Z);
    vcycle(lv, G.stream, MG_CORRECT, F_Z, F_X, false, true);
    launch(k_ir_correct, G.nwork, G.stream, dd, w);
    launch_named("k_gm_resid", k_gm_resid<float, HET>, G.nwork, G.stream, dd, w, (int)F_Z, 1);
    FinArgs a{};
    a.rtol2 = rtol2;
    a.inner_rtol2 = inner2;
    a.max_refine = 2000;
    a.mg = 1;
    finish(G, FIN_GM_IR, a);
    vcycle(lv, G.stream, MG_INIT, F_Z, F_X, true, true);
    finish(G, FIN_MG_RZ0, a);
    for (int k = 0; k < iters; k++) {
        launch_named("k_gm_apply", k_gm_apply<float, HET>, G.nwork, G.stream, dd, w, (int)F_X, (int)F_Q);
        launch_named("k_gm_dots", k_gm_dots<float>, G.nwork, G.stream, dd, w, (int)F_Q);
        finish(G, FIN_GM_H, a);
        launch_named("k_gm_orth", k_gm_orth<float, HET>, G.nwork, G.stream, dd, w, (int)F_Q, (int)F_Z, 1);
        finish(G, FIN_GM_NORM, a);
        vcycle(lv, G.stream, MG_RUNNING, F_Z, F_X, false, true);
    }
}

// Mixed-precision refinement, one round: problems whose inner FP32 solve has
// finished take its correction and start the next one from a fresh FP64
// residual (or stop); then `chunk` inner iterations for every problem.

template <bool HET>
static void mixed_round(const Group &G, double rtol2, int max_iters, int chunk) {
    launch(k_ir_correct, G.nwork, G.stream, (const ProbDesc *)G.d_desc, (const WorkItem *)G.d_work);
    launch_named("k_ir_resid", k_ir_resid<HET>, G.nwork, G.stream, (const ProbDesc *)G.d_desc,
                 (const WorkItem *)G.d_work, rtol2, kInnerRtol2, kMaxRefine);
    FinArgs a{};
    a.rtol2 = rtol2;
    a.inner_rtol2 = kInnerRtol2;
    a.max_refine = kMaxRefine;
    finish(G, FIN_IR, a);
    cg_iterations<float, HET>(G, chunk, F_X, F_R, F_Z, F_P0, F_Q, S_IDONE, S_ITOL2, max_iters);
}

// scalar slots that start a mixed solve
static const std::vector<std::pair<int, double>> kMixedStart = {
    {S_NREFINE, 0.0}, {S_DONE, 0.0}, {S_IDONE, 1.0}, {S_CORR, 0.0}, {S_TITERS, 0.0}, {S_ITERS, 0.0}};

template <bool HET>
static void steady_jpcg64(Model &M, DevModel &D) {
    const double rtol2 = M.opt.steady_rtol * M.opt.steady_rtol;
    jpcg64_start<HET>(D.grp, rtol2, M.opt.max_iters);
    int chunk = 8;
    for (;;) {
        read_scal(D);
        if (D.h_scal[S_DONE] != 0.0) break;
        cg_iterations<double, HET>(D.grp, chunk, B_STEADY, B_R, B_Z, B_P0, B_Q, S_DONE, S_TOL2, M.opt.max_iters);
        chunk = std::min(chunk * 2, 64);
    }
    M.stats.steady_iters = (int)D.h_scal[S_ITERS];
    M.stats.steady_refinements = 0;
    M.stats.steady_rel_resid = D.h_scal[S_BB] > 0 ? std::sqrt(D.h_scal[S_RR] / D.h_scal[S_BB]) : 0.0;
}

template <bool HET>
static void steady_mixed(Model &M, DevModel &D) {
    const double rtol2 = M.opt.steady_rtol * M.opt.steady_rtol;
    write_scal(D, kMixedStart);
    const bool mg = M.opt.steady_solver == HS_STEADY_MGPCG;
    std::vector<MgGroup> lv;
    if (mg) {
        build_mg(D);
        lv = mg_levels_of(D);
    }
    int chunk = 8;
    for (;;) {
        if (mg) mixed_round_mg<HET>(D.grp, lv, rtol2, M.opt.max_iters, chunk);
        else mixed_round<HET>(D.grp, rtol2, M.opt.max_iters, chunk);
        read_scal(D);
        if (D.h_scal[S_DONE] != 0.0) break;
        // a new inner solve starts with short chunks
        chunk = D.h_scal[S_IDONE] != 0.0 ? 8 : std::min(chunk * 2, 64);
    }
    M.stats.steady_iters = (int)D.h_scal[S_TITERS];
    M.stats.steady_refinements = (int)D.h_scal[S_NREFINE];
    M.stats.steady_rel_resid = D.h_scal[S_BB] > 0 ? std::sqrt(D.h_scal[S_RR] / D.h_scal[S_BB]) : 0.0;
}

// Non-symmetric G (advection): FP64 refinement around GMRES(kGmM) cycles in
// FP64 (each to a 1e-6 relative residual) or FP32 (1e-4).
template <typename T, bool HET>
static void steady_gmres(Model &M, DevModel &D) {
    ensure_gmres(D, sizeof(T) == 4);
    const double rtol2 = M.opt.steady_rtol * M.opt.steady_rtol;
    const double inner2 = sizeof(T) == 4 ? 1e-8 : 1e-12;
    write_scal(D, kMixedStart);
    for (int guard = 0;; guard++) {
        gm_round<T, HET>(D.grp, rtol2, inner2, kGmM);
        read_scal(D);
        if (D.h_scal[S_DONE] != 0.0) break;
        if (guard > 100000) fatal("non-symmetric steady solve did not finish\n");
    }
    M.stats.steady_iters = (int)D.h_scal[S_TITERS];
    M.stats.steady_refinements = (int)D.h_scal[S_NREFINE];
    M.stats.steady_rel_resid = D.h_scal[S_BB] > 0 ? std::sqrt(D.h_scal[S_RR] / D.h_scal[S_BB]) : 0.0;
}

// Non-symmetric G with the multigrid preconditioner: FP32 GMRES cycles
// (each to a 1e-4 relative residual) inside FP64 refinement.
static void steady_gmres_mg(Model &M, DevModel &D) {
    ensure_gmres(D, true);
    build_mg(D);
    const std::vector<MgGroup> lv = mg_levels_of(D);
    const double rtol2 = M.opt.steady_rtol * M.opt.steady_rtol;
    write_scal(D, kMixedStart);
    for (int guard = 0;; guard++) {
        gm_round_mg<true>(D.grp, lv, rtol2, 1e-8, kGmM);
        read_scal(D);
        if (D.h_scal[S_DONE] != 0.0) break;
        if (guard > 100000) fatal("non-symmetric steady solve did not finish\n");
    }
    M.stats.steady_iters = (int)D.h_scal[S_TITERS];
    M.stats.steady_refinements = (int)D.h_scal[S_NREFINE];
    M.stats.steady_rel_resid = D.h_scal[S_BB] > 0 ? std::sqrt(D.h_scal[S_RR] / D.h_scal[S_BB]) : 0.0;
}

void dev_steady(Model &M, const double *block_power, double *block_temp) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    HIP_OK(hipEventRecord(D.ev0, D.stream));
    upload_blocks(D, block_power);
    run_b2g(D.grp, V_POWER, B_RHS, 1);
    if (!M.steady_valid || !M.opt.steady_warm_start) {
        const int nb = (int)((D.vlen + kBlock - 1) / kBlock);
        launch(k_fill, nb, D.stream, D.h.b64[B_STEADY], D.vlen, M.g->config.ambient);
    }
    write_scal(D, {{S_DONE, 0.0}});
    if (D.h.has_adv) {
        // non-symmetric: GMRES with the upwind multigrid preconditioner in
        // every form; Jacobi-preconditioned GMRES does not converge at the
        // cell Peclet numbers of real channels
        steady_gmres_mg(M, D);
    } else if (M.opt.precision == HS_PREC_MIXED || M.opt.steady_solver == HS_STEADY_MGPCG) {
        if (D.hetero) steady_mixed<true>(M, D);
        else steady_mixed<false>(M, D);
    } else {
        if (D.hetero) steady_jpcg64<true>(M, D);
        else steady_jpcg64<false>(M, D);
    }
    run_g2b(D.grp, B_STEADY);
    HIP_OK(hipEventRecord(D.ev1, D.stream));
    download_blocks(D, block_temp);
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    M.stats.last_solve_ms = ms;
    // a solve that stopped at its iteration limits far from the tolerance
    const double rel = M.stats.steady_rel_resid;
    if (!(rel <= std::max(100.0 * M.opt.steady_rtol, 1e-6))) {
        char msg[160];
        snprintf(msg, sizeof msg, "steady solve did not converge (relative residual %.2e after %d iterations)\n",
                 rel, M.stats.steady_iters);
        fatal(msg);
    }
}

// ---- transient ---------------------------------------------------------------
template <bool HET>
static void rk4_attempt(const Group &G) {
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    auto st = [&](int mode, int base, int kin, double a, int kout, int out, double hc, int cmp) {
        launch_named("k_rk_stage", k_rk_stage<HET>, G.nwork, G.stream, dd, w, mode, base, kin, a, kout, out, hc,
                     cmp);
    };
    const int Y = 100, YT = 101;
    st(ST_K, Y, -1, 0.0, B_K1, -1, 0.0, -1);                 // k1 = f(y), first attempt only
    st(ST_FIRST, Y, B_K1, 0.5, B_KA, -1, 0.0, -1);           // full step
    st(ST_MID, Y, B_KA, 0.5, B_KB, -1, 0.0, -1);
    st(ST_LAST, Y, B_KB, 1.0, -1, YT, 1.0, -1);
    st(ST_FIRST, Y, B_K1, 0.25, B_KA, -1, 0.0, -1);          // first half step
    st(ST_MID, Y, B_KA, 0.25, B_KB, -1, 0.0, -1);
    st(ST_LAST, Y, B_KB, 0.5, -1, B_T1, 0.5, -1);
    st(ST_K, B_T1, -1, 0.0, B_K1, -1, 0.0, -1);              // k1 = f(t1)
    st(ST_FIRST, B_T1, B_K1, 0.25, B_KA, -1, 0.0, -1);       // second half step
    st(ST_MID, B_T1, B_KA, 0.25, B_KB, -1, 0.0, -1);
    st(ST_LAST_ERR, B_T1, B_KB, 0.5, -1, -1, 0.5, YT);
    launch(k_rk_control, G.nprob, G.stream, dd);
}

template <bool HET>
static void transient_rk4(Model &M, DevModel &D, double dt) {
    write_scal(D, {{S_T, 0.0}, {S_DT, dt}, {S_H, kMinStep}, {S_STATUS, (double)RK_NEED_K1}, {S_CALLS, 0.0}});
    int batch = 4;
    for (int guard = 0;; guard++) {
        for (int k = 0; k < batch; k++) rk4_attempt<HET>(D.grp);
        read_scal(D);
        if (D.h_scal[S_STATUS] == (double)RK_DONE) break;
        batch = 2;
        if (guard > 10000000) fatal("transient integration did not finish\n");
    }
    M.stats.rk4_calls = (int)D.h_scal[S_CALLS];
}

// Taylor set-up: sub-steps keep z = rho*tau <= 4; the term count is the
// a-priori bound for a 1000 K change per sub-step, and the device stops each
// series early once its actual tail bound is below the tolerance.
//
// The sub-step's increment is tau phi_1(-tau A) f(y), A = C^-1 G, phi_1(x) =
// (e^x - 1)/x.  Unshifted, the series is Taylor's about x = 0 and its terms
// shrink like z^k/(k+1)!.  When every row of A is weakly diagonally dominant
// (conduction, no coolant advection), the series is expanded about the
// centre of the spectrum instead, x0 = -sigma with sigma = z/2:
//   phi_1(X) = sum_k c_k (X + sigma I)^k,  c_k = (1/k!) int_0^1 s^k e^{-sigma s} ds,
// and ||X + sigma I||_inf <= z/2 (row i: |sigma - a_ii tau| + sum_j |a_ij| tau
// <= z/2 when a_ii + sum_j |a_ij| <= rho and sum_j |a_ij| <= a_ii).  Since
// c_k <= c_0/(k+1)! (Chebyshev's integral inequality: s^k rises where
// e^{-sigma s} falls), the tail after K terms is bounded by the unshifted
// rule with z replaced by z/2, so a series needs about half the terms' worth
// of z: 8-9 instead of 12-13 Horner steps at z = 2.3.  In Horner form with
// E_1 = c_0 tau f(y) (S_TAU1) and h_k scaled by c_k/c_0, step j (k = j - 2)
// is h <- E_1 + r_k (tau M h + sigma h), r_k = c_{k+1}/c_k: the Taylor record
// holds s_j = r_k tau and t_j = r_k sigma.  Unshifted, r_k = 1/(k+2) (s_j =
// tau/j, t_j = 0) and c_0 = 1.
struct TaylorPlan {
    int m = 1, K = 1;
    double tau = 0.0, tol = 0.0;
    double sigma = 0.0, tau1 = 0.0;   // shift of the series; factor of the slope in E_1
};
// k! c_k = sum_i (-sigma)^i / (i! (k + i + 1)); sigma <= 2, so 40 terms reach
// the double rounding
static double shifted_phi_coef(int k, double sigma) {
    double term = 1.0, sum = 0.0;
    for (int i = 0; i < 40; i++) {
        sum += term / (k + i + 1);
        term *= -sigma / (i + 1);
    }
    return sum;
}
static TaylorPlan taylor_plan(const DevModel &D, double dt, double taylor_tol) {
    TaylorPlan p;
    const double zmax = 4.0;
    p.m = std::max(1, (int)std::ceil(D.rho * dt / zmax));
    p.tau = dt / p.m;
    const bool shift = !D.h.has_adv;
    p.sigma = shift ? 0.5 * D.rho * p.tau : 0.0;
    p.tau1 = shift ? p.tau * shifted_phi_coef(0, p.sigma) : p.tau;
    const double z = shift ? p.sigma : D.rho * p.tau;
    p.tol = taylor_tol / p.m;
    p.K = 1;
    double term = 1000.0;
    while (p.K < 60) {
        term *= z / (p.K + 1);
        const double q = z / (p.K + 2);
        if (term * (q < 1 ? 1.0 / (1.0 - q) : 1e30) <= p.tol) break;
        p.K++;
    }
    return p;
}

// scalar slots of a problem's Taylor plan
static std::vector<std::pair<int, double>> taylor_slots(const TaylorPlan &p, double rho) {
    return {{S_TAU, p.tau}, {S_TK, (double)p.m}, {S_RHO, rho}, {S_TTOL, p.tol}, {S_KPLAN, (double)p.K},
            {S_TAU1, p.tau1}, {S_TSIG, p.sigma}, {S_THOLD, 0.0}};
}

// step factors of a problem's Taylor record (kTrec doubles; [0] and [1] are
// the device's)
static void taylor_record(const TaylorPlan &p, double *rec) {
    std::fill(rec, rec + kTrec, 0.0);
    float *rf = (float *)(rec + kTrecF);
    double ck = shifted_phi_coef(0, p.sigma);
    for (int j = 2; j <= std::min(p.K, kTayMaxK); j++) {
        if (p.sigma == 0.0) {
            rec[2 * j] = p.tau / j;
        } else {
            const int k = j - 2;
            const double ck1 = shifted_phi_coef(k + 1, p.sigma);
            const double r = ck1 / (ck * (k + 1));
            rec[2 * j] = r * p.tau;
            rec[2 * j + 1] = r * p.sigma;
            ck = ck1;
        }
        rf[2 * j] = (float)rec[2 * j];
        rf[2 * j + 1] = (float)rec[2 * j + 1];
    }
}

// a single model's record, uploaded when its plan changes
static void upload_taylor_record(DevModel &D, const TaylorPlan &p) {
    std::vector<double> rec(kTrec);
    taylor_record(p, rec.data());
    if (rec == D.trec_plan) return;
    D.trec_plan = rec;
    HIP_OK(hipMemcpyAsync(D.grp.d_trec, D.trec_plan.data(), kTrec * sizeof(double), hipMemcpyHostToDevice,
                          D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

// One interval of every problem: m sub-steps of at most K terms (the largest
// plan of the group).  E_1 goes to the accumulator buffer; the Horner steps
// j = K..2 alternate between the two term buffers, and each problem takes the
// last of them that its own term count needs.
template <typename TD, bool HET>
static void taylor_interval(const Group &G, int m, int Kplan, int seq) {
    // the steps launched: the plan's, or fewer where the engine has set
    // kgraph (a problem needing more is held, see FIN_TAY)
    const int K = G.kgraph > 0 ? std::min(Kplan, G.kgraph) : Kplan;
    const ProbDesc *dd = G.d_desc;
    const WorkItem *w = G.d_work;
    const bool f32 = sizeof(TD) == 4;
    const int db0 = f32 ? F_D0 : B_KA, db1 = f32 ? F_D1 : B_KB, accb = f32 ? F_ACC : B_ACC;
    for (int s = 0; s < m; s++) {
        const bool zf = f32 && !HET && G.vec4 && G.zfirst;
        if (zf) {
            auto k = G.zpow && seq == 1 ? (G.tg == 1 ? k_tay_first4z<1, true> : k_tay_first4z<2, true>)
                                        : (G.tg == 1 ? k_tay_first4z<1, false> : k_tay_first4z<2, false>);
            launch_named("k_tay_first", k, G.ntwork, G.stream, dd, (const WorkItem *)G.d_twork, G.tnper, accb, s,
                         seq, G.d_kmax);
        } else if (f32 && !HET && G.vec4)
            launch_named("k_tay_first", k_tay_first2, G.nwork, G.stream, dd, w, accb, s, seq, G.d_kmax, 1);
        else
            launch_named("k_tay_first", k_tay_first<TD, HET>, G.nwork, G.stream, dd, w, accb, s, seq, G.d_kmax);
        // float4 sweeps finish the sub-step in their last launch (j = 2)
        const bool fused = f32 && !HET && G.vec4 && K >= 2;
        FinArgs a{};
        a.sub = s;
        a.seq = seq;
        a.kmax = G.d_kmax;
        a.trec = G.d_trec;
        a.kg = K;
        a.fused = fused ? 1 : 0;
        a.zslots = zf ? 1 : 0;
        finish(G, FIN_TAY, a);
        const int *km = G.d_kmax;
        const double *tr = G.d_trec;
        for (int j = K; j >= 2; j--) {
            if (f32 && !HET && G.vec4) {
                launch_term4(G, G.stream, accb, db0, db1, j, (fused && j == 2) ? 1 : 0, true);
            } else {
                launch_named("k_tay_term", k_tay_term<TD, HET>, G.nwork, G.stream, dd, w, G.nper, accb, db0, db1,
                             j, km, tr);
            }
        }
        if (!fused)
            launch_named("k_tay_final", k_tay_final<TD>, G.nwork, G.stream, dd, w, G.nper, accb, db0, db1, tr);
    }
}

template <typename TD, bool HET>
static void transient_taylor(Model &M, DevModel &D, double dt) {
    const TaylorPlan p = taylor_plan(D, dt, M.opt.taylor_tol);
    write_scal(D, taylor_slots(p, D.rho));
    upload_taylor_record(D, p);
    taylor_interval<TD, HET>(D.grp, p.m, p.K, 0);
    M.stats.taylor_terms = p.K;
    M.stats.taylor_substeps = p.m;
}

// Backward Euler (RCutil.c:268-428 as compute_temp_grid calls it,
// temperature_grid.c:3632-3662): one implicit step of the whole interval h,
//   (C/h + G) y' = C y / h + P + sources,
// G the assembled matrix of the SuperLU path (build_transient_grid_matrix),
// whose secondary package links are the steady ones (symmetric); the
// secondary package nodes carry no capacitance in it (build_diagonal_matrix
// :6750-6768).  FP64 Jacobi-preconditioned CG from y, to the steady
// tolerance; the shift C/h goes into the ambient conductances of a second
// descriptor (per layer, or per cell for heterogeneous models).
// With xb/rhsb given, the shifted descriptor's B_STEADY and B_RHS slots point
// at those buffers (the GMRES kernels solve into B_STEADY from B_RHS).
static void build_be(DevModel &D, double h, int xb = -1, int rhsb = -1) {
    const double sigma = 1.0 / h;
    ProbDesc b = D.h;
    if (xb >= 0) b.b64[B_STEADY] = D.h.b64[xb];
    if (rhsb >= 0) b.b64[B_RHS] = D.h.b64[rhsb];
    if (!D.hetero)
        for (int n = 0; n < b.nl; n++) b.gamb[n] += sigma / b.cinv[n];
    for (int k = 0; k < std::min(b.nx, (int)EXTRA); k++) b.x_gamb[k] += sigma / b.x_cinv[k];
    if (D.hetero) {
        const long ncell = (long)b.nl * b.nr * b.nc;
        if (!D.be_shift) {
            D.be_shift = D.alloc<double>(ncell);
            D.be_dinv = D.alloc<double>(ncell);
        }
        hipLaunchKernelGGL(k_be_coef, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, D.stream,
                           (const ProbDesc *)D.d_desc, ncell, sigma, D.be_shift, D.be_dinv);
        b.bshift = D.be_shift;
        b.dinvc = D.be_dinv;
    }
    if (!D.d_be_desc) D.d_be_desc = D.alloc<ProbDesc>(1);
    D.be_h = b;
    HIP_OK(hipMemcpyAsync(D.d_be_desc, &D.be_h, sizeof(ProbDesc), hipMemcpyHostToDevice, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
    D.be_grp = D.grp;
    D.be_grp.d_desc = D.d_be_desc;
}

// The multigrid levels of the backward-Euler step's operator C/h + G for
// non-symmetric (advection) models: the steady hierarchy's coefficients
// (upwind advection, aggregated faces) with the step's shift added per cell
// on every level (k_mg_shift) and to the primary package nodes, as build_be
// adds it to the fine operator.  Rebuilt when h changes; D.be_shift must
// hold the fine shift (build_be).
static std::vector<MgGroup> be_levels(DevModel &D, double sigma) {
    build_mg(D);
    if (D.be_mg.empty()) {
        for (const MgLevel &m : D.mg) {
            MgLevel L = m;
            const long ncell = (long)L.h.nl * L.h.nr * L.h.nc;
            L.h.bshift = D.alloc<double>(ncell);
            L.h.fdinv = D.alloc<float>(ncell);
            L.d_desc = D.alloc<ProbDesc>(1);
            D.be_mg.push_back(L);
        }
        D.be_sigma = 0.0;
    }
    if (sigma != D.be_sigma) {
        for (size_t l = 0; l < D.be_mg.size(); l++) {
            ProbDesc &h = D.be_mg[l].h;
            for (int k = 0; k < std::min(h.nx, (int)EXTRA); k++) h.x_gamb[k] = D.mg[l].h.x_gamb[k] + sigma / h.x_cinv[k];
            h2d(D.be_mg[l].d_desc, &h, sizeof(ProbDesc), D.stream);
            const long ncell = (long)h.nl * h.nr * h.nc;
            hipLaunchKernelGGL(k_mg_shift, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, D.stream,
                               l == 0 ? (const ProbDesc *)nullptr : (const ProbDesc *)D.be_mg[l - 1].d_desc,
                               (const ProbDesc *)D.be_mg[l].d_desc, (const double *)D.be_shift,
                               (const float *)D.mg[l].h.fdinv);
        }
        HIP_OK(hipStreamSynchronize(D.stream));
        D.be_sigma = sigma;
    }
    std::vector<MgGroup> lv;
    for (const MgLevel &L : D.be_mg) lv.push_back(MgGroup{L.d_desc, L.d_cwork, L.ncwork, D.h.nl, true});
    return lv;
}

template <bool HET>
static void transient_be(Model &M, DevModel &D, double dt) {
    read_scal(D);
    const int yb = B_Y0 + (int)D.h_scal[S_YPAR];
    const bool adv = D.h.has_adv;
    if (adv) ensure_gmres(D, true);   // the shifted descriptor copies its buffers
    build_be(D, dt, adv ? yb : -1, adv ? (int)B_T1 : -1);
    launch_named("k_be_rhs", k_be_rhs<HET>, D.grp.nwork, D.stream, (const ProbDesc *)D.d_desc,
                 (const WorkItem *)D.d_work, yb, (int)B_RHS, (int)B_T1, 1.0 / dt);
    // C y / h dominates the right-hand side's norm, so the relative residual
    // goes 100x below the steady tolerance to keep every node as close
    const double rtol = 1e-2 * M.opt.steady_rtol, rtol2 = rtol * rtol;
    if (adv) {
        // coolant advection makes the system non-symmetric: FP32 GMRES
        // cycles from y, right-preconditioned by the multigrid V-cycle of
        // the shifted upwind operator, inside FP64 refinement (Jacobi-
        // preconditioned GMRES stalls on these systems as on the steady ones)
        const std::vector<MgGroup> lv = be_levels(D, 1.0 / dt);
        write_scal(D, kMixedStart);
        for (int guard = 0;; guard++) {
            gm_round_mg<true>(D.be_grp, lv, rtol2, 1e-8, kGmM);
            read_scal(D);
            if (D.h_scal[S_DONE] != 0.0) break;
            if (guard > 100000) fatal("backward Euler: the non-symmetric solve did not finish\n");
        }
        M.stats.be_iters = (int)D.h_scal[S_TITERS];
    } else {
        write_scal(D, {{S_DONE, 0.0}});   // the CG start skips a problem still marked done
        jpcg64_start<HET>(D.be_grp, rtol2, M.opt.max_iters, yb, B_T1);
        int chunk = 8;
        for (;;) {
            read_scal(D);
            if (D.h_scal[S_DONE] != 0.0) break;
            cg_iterations<double, HET>(D.be_grp, chunk, yb, B_R, B_Z, B_P0, B_Q, S_DONE, S_TOL2, M.opt.max_iters);
            chunk = std::min(chunk * 2, 64);
        }
        M.stats.be_iters = (int)D.h_scal[S_ITERS];
    }
    // a step stopped at its iteration limits far above its tolerance
    const double rel = D.h_scal[S_BB] > 0 ? std::sqrt(D.h_scal[S_RR] / D.h_scal[S_BB]) : 0.0;
    if (!(rel <= std::max(100.0 * rtol, 1e-9))) {
        char msg[160];
        snprintf(msg, sizeof msg, "backward Euler step did not converge (relative residual %.2e after %d iterations)\n",
                 rel, M.stats.be_iters);
        fatal(msg);
    }
}

// the float4 term tiling of a single model, timed on its first FP32 Taylor use
static void ensure_term_tuned(const Model &M, DevModel &D) {
    if (D.term_tuned || !D.grp.vec4 || M.opt.transient != HS_TRANS_TAYLOR || M.opt.precision != HS_PREC_MIXED)
        return;
    std::vector<ProbDesc> d{D.h};
    tune_term_tiling(D.grp, d, D.cap_twork, D.pool, D.stream);
    D.h = d[0];
    D.term_tuned = true;
}

void dev_transient(Model &M, const double *block_power, const double *block_temp_in, double dt,
                   double *block_temp_out) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    ensure_term_tuned(M, D);
    M.stats.be_iters = 0;
    HIP_OK(hipEventRecord(D.ev0, D.stream));
    upload_blocks(D, block_power);
    if (M.opt.transient == HS_TRANS_TAYLOR && power32(D.grp, M.opt.precision == HS_PREC_MIXED, D.hetero))
        run_b2g_power32(D.grp);
    else
        run_b2g(D.grp, V_POWER, B_RHS, 1);
    if (block_temp_in) {
        HIP_OK(hipStreamSynchronize(D.stream));
        upload_blocks(D, block_temp_in);
        run_b2g(D.grp, V_TEMP, B_Y0, 0);
        write_scal(D, {{S_YPAR, 0.0}});
    }
    if (M.opt.transient == HS_TRANS_TAYLOR) {
        if (M.opt.precision == HS_PREC_MIXED) {
            if (D.hetero) transient_taylor<float, true>(M, D, dt);
            else transient_taylor<float, false>(M, D, dt);
        } else {
            if (D.hetero) transient_taylor<double, true>(M, D, dt);
            else transient_taylor<double, false>(M, D, dt);
        }
    } else if (M.opt.transient == HS_TRANS_BE) {
        if (D.hetero) transient_be<true>(M, D, dt);
        else transient_be<false>(M, D, dt);
    } else {
        if (D.hetero) transient_rk4<true>(M, D, dt);
        else transient_rk4<false>(M, D, dt);
    }
    read_scal(D);
    const int ypar = (int)D.h_scal[S_YPAR];
    run_g2b(D.grp, B_Y0 + ypar);
    HIP_OK(hipEventRecord(D.ev1, D.stream));
    download_blocks(D, block_temp_out);
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    M.stats.last_solve_ms = ms;
}

// ---- device-resident sequences of intervals -----------------------------------
// Captures one interval's launches once as a hipGraph and replays it for
// every interval (the kernels take the interval index from device memory, so
// the same graph serves all of them).
static bool graphs_enabled() { return g_timer.cap == 0; }

struct Replayer {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool stale = false;   // the body's launches changed (problems, buffers): capture again
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        stale = false;
    }
    ~Replayer() { reset(); }
    // runs body `steps` times on the stream: plain launches while a kernel
    // timer or profiler wants them, else a graph captured on first use.  A
    // stale graph is captured again and its executable updated in place
    // when the new capture has the same nodes (only their launch parameters
    // differ: grid sizes, buffers), which costs less than instantiating it.
    template <class F> void run(hipStream_t s, int steps, F body) {
        if (!graphs_enabled()) {
            for (int k = 0; k < steps; k++) body();
            return;
        }
        if (!exec || stale) {
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            body();
            HIP_OK(hipStreamEndCapture(s, &g));
            bool updated = false;
            if (exec) {
                hipGraphExecUpdateResult why;
                hipGraphNode_t bad = nullptr;
                updated = hipGraphExecUpdate(exec, g, &bad, &why) == hipSuccess && why == hipGraphExecUpdateSuccess;
                (void)hipGetLastError();
            }
            if (!updated) {
                if (exec) (void)hipGraphExecDestroy(exec);
                exec = nullptr;
                HIP_OK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
            }
            if (graph) (void)hipGraphDestroy(graph);
            graph = g;
            stale = false;
        }
        for (int k = 0; k < steps; k++) HIP_OK(hipGraphLaunch(exec, s));
    }
};

// RK4 interval of every problem of the group with `slots` attempts; an
// interval that needs more continues in the next launch of the body.
template <bool HET>
static void rk4_seq_body(const Group &G, int slots) {
    run_b2g(G, V_POWER, B_RHS, 1, 1);
    hipLaunchKernelGGL(k_rk_begin, dim3(small_grid(G.nprob)), dim3(64), 0, G.stream, (const ProbDesc *)G.d_desc,
                       G.nprob);
    for (int a = 0; a < slots; a++) rk4_attempt<HET>(G);
    hipLaunchKernelGGL(k_rk_end, dim3(small_grid(G.nprob)), dim3(64), 0, G.stream, (const ProbDesc *)G.d_desc,
                       G.nprob);
    run_g2b(G, -1, 1);
    hipLaunchKernelGGL(k_step_advance, dim3(small_grid(G.nprob)), dim3(64), 0, G.stream,
                       (const ProbDesc *)G.d_desc, G.nprob);
}

template <typename TD, bool HET>
static void taylor_seq_body(const Group &G, int m, int K) {
    if (power32(G, sizeof(TD) == 4, HET)) {
        // a z-marching first term with zpow maps the power itself
        if (!(G.zfirst && G.zpow)) run_b2g_power32(G, 1);
    } else {
        run_b2g(G, V_POWER, B_RHS, 1, 1);
    }
    taylor_interval<TD, HET>(G, m, K, 1);
    run_g2b(G, -1, 1);
    hipLaunchKernelGGL(k_step_advance, dim3(small_grid(G.nprob)), dim3(64), 0, G.stream,
                       (const ProbDesc *)G.d_desc, G.nprob);
}

// The interval body of a group for the given form (precision, integrator).
struct SeqForm {
    int precision, transient, m, K;
    static constexpr int kSlots = 5;   // RK4 attempts per interval launch
};
static void seq_body(const Group &G, const SeqForm &f) {
    if (f.transient == HS_TRANS_TAYLOR) {
        const bool f32 = f.precision == HS_PREC_MIXED;
        if (f32 && G.hetero) taylor_seq_body<float, true>(G, f.m, f.K);
        else if (f32) taylor_seq_body<float, false>(G, f.m, f.K);
        else if (G.hetero) taylor_seq_body<double, true>(G, f.m, f.K);
        else taylor_seq_body<double, false>(G, f.m, f.K);
    } else {
        if (G.hetero) rk4_seq_body<true>(G, SeqForm::kSlots);
        else rk4_seq_body<false>(G, SeqForm::kSlots);
    }
}

// scalar slots that start a sequence of n intervals of length dt
static std::vector<std::pair<int, double>> seq_slots(int n, double dt) {
    return {{S_STEP, 0.0}, {S_NSTEPS, (double)n}, {S_TCALLS, 0.0}, {S_CALLS, 0.0},
            {S_STATUS, (double)RK_DONE}, {S_DT, dt}, {S_TTERMS, 0.0}, {S_TBYTES, 0.0}, {S_NHOLD, 0.0}};
}

void dev_transient_run(Model &M, const double *power_seq, int steps, double dt, const double *init_temp,
                       double *temp_seq) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    ensure_term_tuned(M, D);
    if (!init_temp && !M.trans_valid) fatal("hs_transient_run: no transient state to continue from\n");
    if (M.opt.transient == HS_TRANS_BE) {
        // one implicit solve per interval, each converged under host control
        for (int k = 0; k < steps; k++)
            dev_transient(M, power_seq + (size_t)k * D.nblk, k == 0 ? init_temp : nullptr, dt,
                          temp_seq ? temp_seq + (size_t)k * D.nblk : nullptr);
        return;
    }
    const size_t n = (size_t)steps * D.nblk;
    double *seq_in = nullptr, *seq_out = nullptr;
    HIP_OK(hipMalloc((void **)&seq_in, n * sizeof(double)));
    struct Free {
        std::vector<double *> p;
        ~Free() {
            for (double *x : p) (void)hipFree(x);
        }
    } tmp;
    tmp.p.push_back(seq_in);
    HIP_OK(hipMalloc((void **)&seq_out, n * sizeof(double)));
    tmp.p.push_back(seq_out);
    D.h.seq_in = seq_in;
    D.h.seq_out = seq_out;
    h2d(D.d_desc, &D.h, sizeof(ProbDesc), D.stream);
    HIP_OK(hipEventRecord(D.ev0, D.stream));
    HIP_OK(hipMemcpyAsync(seq_in, power_seq, n * sizeof(double), hipMemcpyHostToDevice, D.stream));
    if (init_temp) {
        upload_blocks(D, init_temp);
        run_b2g(D.grp, V_TEMP, B_Y0, 0);
        HIP_OK(hipStreamSynchronize(D.stream));
        write_scal(D, {{S_YPAR, 0.0}});
    }
    write_scal(D, seq_slots(steps, dt));
    SeqForm f{M.opt.precision, M.opt.transient, 1, 1};
    if (M.opt.transient == HS_TRANS_TAYLOR) {
        const TaylorPlan p = taylor_plan(D, dt, M.opt.taylor_tol);
        write_scal(D, taylor_slots(p, D.rho));
        upload_taylor_record(D, p);
        f.m = p.m;
        f.K = p.K;
        M.stats.taylor_terms = p.K;
        M.stats.taylor_substeps = p.m;
    }
    {
        Replayer rp;
        int todo = steps;
        for (int guard = 0; todo > 0; guard++) {
            rp.run(D.stream, todo, [&] { seq_body(D.grp, f); });
            read_scal(D);
            todo = steps - (int)D.h_scal[S_STEP];
            if (guard > 1000000) fatal("transient sequence did not finish\n");
        }
        HIP_OK(hipStreamSynchronize(D.stream));
    }
    if (M.opt.transient == HS_TRANS_RK4) M.stats.rk4_calls = (int)D.h_scal[S_TCALLS];
    HIP_OK(hipEventRecord(D.ev1, D.stream));
    if (temp_seq)
        HIP_OK(hipMemcpyAsync(temp_seq, seq_out, n * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    M.stats.last_solve_ms = ms;
    D.h.seq_in = D.h.seq_out = nullptr;
    h2d(D.d_desc, &D.h, sizeof(ProbDesc), D.stream);
}

// ---- grid-level access ---------------------------------------------------------
void dev_get_grid(Model &M, int which, double *out) {
    if (!M.dev) fatal("no device state\n");
    DevModel &D = *M.dev;
    int b = B_STEADY;
    if (which == 1) {
        read_scal(D);
        b = B_Y0 + (int)D.h_scal[S_YPAR];
    }
    HIP_OK(hipMemcpyAsync(out, D.h.b64[b], M.vec_len() * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

void dev_set_grid(Model &M, int which, const double *in) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    int b = B_STEADY;
    if (which == 1) {
        read_scal(D);
        b = B_Y0 + (int)D.h_scal[S_YPAR];
    }
    HIP_OK(hipMemcpyAsync(D.h.b64[b], in, M.vec_len() * sizeof(double), hipMemcpyHostToDevice, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

void dev_slope(Model &M, const double *block_power, const double *grid_temp, double *out) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    std::vector<double> p(block_power, block_power + D.nblk);
    for (int k = 0; k < M.nx; k++) p[M.g->total_n_blocks + k] = 0.0;
    upload_blocks(D, p.data());
    run_b2g(D.grp, V_POWER, B_RHS, 1);
    HIP_OK(hipMemcpyAsync(D.h.b64[B_TMP], grid_temp, M.vec_len() * sizeof(double), hipMemcpyHostToDevice,
                          D.stream));
    if (D.hetero)
        launch_named("k_slope", k_slope<true>, D.nwork, D.stream, (const ProbDesc *)D.d_desc, (const WorkItem *)D.d_work,
               (int)B_TMP, (int)B_KA);
    else
        launch_named("k_slope", k_slope<false>, D.nwork, D.stream, (const ProbDesc *)D.d_desc, (const WorkItem *)D.d_work,
               (int)B_TMP, (int)B_KA);
    HIP_OK(hipMemcpyAsync(out, D.h.b64[B_KA], M.vec_len() * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

void dev_b2g(Model &M, const double *block_vec, double *grid_vec, int type) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    upload_blocks(D, block_vec);
    run_b2g(D.grp, type, B_TMP, 0);
    HIP_OK(hipMemcpyAsync(grid_vec, D.h.b64[B_TMP], M.vec_len() * sizeof(double), hipMemcpyDeviceToHost,
                          D.stream));
    HIP_OK(hipStreamSynchronize(D.stream));
}

void dev_g2b(Model &M, const double *grid_vec, double *block_vec) {
    dev_prepare(M);
    DevModel &D = *M.dev;
    HIP_OK(hipMemcpyAsync(D.h.b64[B_TMP], grid_vec, M.vec_len() * sizeof(double), hipMemcpyHostToDevice,
                          D.stream));
    run_g2b(D.grp, B_TMP);
    download_blocks(D, block_vec);
}

}  // namespace hs

#include "hs_batch.inl"

extern "C" {
void hs_kernel_timer(const char *name, int max_samples) {
    using namespace hs;
    g_timer.close();
    for (auto &e : g_timer.ev) {
        (void)hipEventSynchronize(e.b);
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    g_timer.ev.clear();
    g_timer.launched = 0;
    g_timer.graph_ms = 0.0;
    g_timer.graph_n = 0;
    g_timer.name = name ? name : "";
    g_timer.cap = name ? max_samples : 0;
}

int hs_kernel_timer_read(double *avg_ms, double *total_ms) {
    using namespace hs;
    double tot = g_timer.graph_ms;
    int n = g_timer.graph_n;
    g_timer.close();
    for (auto &e : g_timer.ev) {
        if (hipEventSynchronize(e.b) != hipSuccess) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            tot += ms;
            n += e.n;
        }
    }
    if (avg_ms) *avg_ms = n ? tot / n : 0.0;
    if (total_ms) *total_ms = tot;
    return n;
}

}

// ---- the operator with no device allocation ----------------------------------------
namespace hs {

// Device scratch of hs_grid_slope_noalloc: the operator's descriptor (pointing
// at the caller's buffers) in the first 256-byte-aligned bytes, then the
// cell power map (one FP64 grid vector).  Every GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` only names the
// symbol: its device address, per device, comes from hipGetSymbolAddress.
constexpr size_t kScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

static unsigned char *scratch_on_current_device() {
    static std::mutex mu;
    static std::vector<unsigned char *> ptr;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fatal(std::string("hs_grid_slope_noalloc: no address for the device scratch: ") + hipGetErrorString(e) +
                  "\n");
        ptr[dev] = (unsigned char *)p;
    }
    return ptr[dev];
}

static void slope_noalloc(Model &M, const double *d_power, const double *d_temp, double *d_slope) {
    const size_t desc_bytes = (sizeof(ProbDesc) + 255) / 256 * 256;
    const size_t vlen = ((size_t)M.ncell() + M.nx + 63) / 64 * 64;
    const size_t need = desc_bytes + vlen * sizeof(double);
    if (need > kScratchBytes)
        fatal("hs_grid_slope_noalloc: the model needs " + std::to_string(need) + " bytes of scratch, above its " +
              std::to_string(kScratchBytes) + "\n");
    if (!M.dev || M.dev_dirty || M.dev->has_c != (bool)M.g->c_ready)
        fatal("hs_grid_slope_noalloc: the model has no device state (build it first, e.g. with hs_grid_slope)\n");
    if (!d_power || !d_temp || !d_slope) fatal("hs_grid_slope_noalloc: null device buffer\n");
    DevModel &D = *M.dev;
    unsigned char *base = scratch_on_current_device();
    ProbDesc h = D.h;
    h.blk_in = const_cast<double *>(d_power);
    h.b64[B_RHS] = (double *)(base + desc_bytes);
    h.b64[B_TMP] = const_cast<double *>(d_temp);
    h.b64[B_KA] = d_slope;
    ProbDesc *dd = (ProbDesc *)base;
    HIP_OK(hipMemcpyAsync(dd, &h, sizeof h, hipMemcpyHostToDevice, D.stream));
    launch(k_b2g, D.nwork, D.stream, (const ProbDesc *)dd, (const WorkItem *)D.d_work, (int)V_POWER, (int)B_RHS, 1, 0);
    if (D.hetero)
        launch(k_slope<true>, D.nwork, D.stream, (const ProbDesc *)dd, (const WorkItem *)D.d_work, (int)B_TMP,
               (int)B_KA);
    else
        launch(k_slope<false>, D.nwork, D.stream, (const ProbDesc *)dd, (const WorkItem *)D.d_work, (int)B_TMP,
               (int)B_KA);
    HIP_OK(hipStreamSynchronize(D.stream));
}

}  // namespace hs

extern "C" int hs_grid_slope_noalloc(RC_model_t *model, const double *d_block_power, const double *d_grid_temp,
                                     double *d_grid_slope) {
    HS_BEGIN
    if (!model || model->type != GRID_MODEL) hs::fatal("hs_grid_slope_noalloc: not a grid model\n");
    hs::Model &M = *hs::model_of(model);
    if (!M.g->r_ready || !M.g->c_ready) hs::fatal("grid model not ready\n");
    hs::slope_noalloc(M, d_block_power, d_grid_temp, d_grid_slope);
    return 0;
    HS_END(-1)
}

extern "C" int hs_pin_sweep_tiling(RC_model_t *model, int shift, int rows_per_thread, int z) {
    HS_BEGIN
    if (!model || model->type != GRID_MODEL) hs::fatal("hs_pin_sweep_tiling: not a grid model\n");
    hs::pin_term_tiling(*hs::model_of(model), shift, rows_per_thread, z);
    return 0;
    HS_END(-1)
}
