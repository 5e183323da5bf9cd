// Device solver of the block thermal model (temperature_block.c) for gfx950.
//
// The block model is a small dense system: n = 4 n_units + 12 nodes, with
//   steady:     B T = P                                  (lusolve)
//   transient:  dT/dt = A^-1 P - C T,  C = A^-1 B         (slope_fn_block)
// B is factorised once per populate_R (lupdcmp, RCutil.c:42-105), the steady
// state is a forward and a backward substitution (lusolve, RCutil.c:124-183),
// and compute_temp integrates with the reference's adaptive RK4 (rk4 and
// rk4_core, RCutil.c:440-640; the interval loop of compute_temp_block,
// temperature_block.c:449-486).
//
// Every step of these is latency-bound (a 132-node ev6 model has 17 k matrix
// entries), so one workgroup owns one model and runs a whole factorisation,
// a whole solve or a whole RK4 attempt in one launch, synchronising with
// workgroup barriers instead of kernel boundaries.
//
// Arithmetic follows the reference's expressions operation by operation,
// unfused, and its sums in the reference's order, so the LU factors, the
// forward substitution and every RK4 attempt are the reference's to the
// bit; with the step-size controller on the host (the C library's pow), a
// transient is bit-identical to the reference's.  Only the backward
// substitution sums in another order (within 1e-12 K).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hs_internal.h"

#pragma clang fp contract(off)

namespace hs {

#define BLK_HIP_OK(x)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            fatal(std::string("HIP error ") + hipGetErrorString(e_) + " in " #x "\n");   \
    } while (0)

constexpr int kBW = 64;             // wavefront width
constexpr int kSolveThreads = 256;
constexpr int kRkThreads = 256;
constexpr int kMaxBlockNodes = NL * 1024 + EXTRA;   // 1024 units

__device__ __forceinline__ double blk_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kBW);
    return v;
}

// ---------------------------------------------------------------------------
// LU factorisation with partial pivoting, in place (lupdcmp): at step k the
// pivot is the first row i >= k of the largest |a[i][k]|; rows k and pivot
// swap whole; the multipliers a[i][k] /= a[k][k] and the rank-1 update of the
// trailing block follow.  A pivot below 1e-6 in magnitude is the reference's
// singular-matrix error (status 1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_blk_lu(double *__restrict__ a, int *__restrict__ perm, int n,
                                                 int *__restrict__ status) {
    __shared__ double s_val[1024 / kBW];
    __shared__ int s_idx[1024 / kBW];
    __shared__ int s_piv;
    const int tid = threadIdx.x, T = blockDim.x, lane = tid & (kBW - 1), wave = tid / kBW, nw = T / kBW;
    for (int i = tid; i < n; i += T) perm[i] = i;
    for (int k = 0; k < n - 1; k++) {
        double best = 0.0;
        int bi = n;
        for (int i = k + tid; i < n; i += T) {
            const double v = fabs(a[(size_t)i * n + k]);
            if (v > best) best = v, bi = i;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(best, o, kBW);
            const int oi = __shfl_xor(bi, o, kBW);
            if (ov > best || (ov == best && oi < bi)) best = ov, bi = oi;
        }
        if (lane == 0) s_val[wave] = best, s_idx[wave] = bi;
        __syncthreads();
        if (tid == 0) {
            double bv = s_val[0];
            int bx = s_idx[0];
            for (int w = 1; w < nw; w++)
                if (s_val[w] > bv || (s_val[w] == bv && s_idx[w] < bx)) bv = s_val[w], bx = s_idx[w];
            s_piv = fabs(bv - 0.0) < 1.0e-6 ? -1 : bx;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv < 0) {
            if (tid == 0) *status = 1;
            return;
        }
        if (piv != k) {
            for (int j = tid; j < n; j += T) {
                const double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)piv * n + j];
                a[(size_t)piv * n + j] = t;
            }
            if (tid == 0) {
                const int t = perm[k];
                perm[k] = perm[piv];
                perm[piv] = t;
            }
            __syncthreads();
        }
        const double akk = a[(size_t)k * n + k];
        for (int i = k + 1 + tid; i < n; i += T) a[(size_t)i * n + k] /= akk;
        __syncthreads();
        for (int i = k + 1 + wave; i < n; i += nw) {
            const double lik = a[(size_t)i * n + k];
            double *ri = a + (size_t)i * n;
            const double *rk = a + (size_t)k * n;
            for (int j = k + 1 + lane; j < n; j += kBW) ri[j] -= lik * rk[j];
        }
        __syncthreads();
    }
    if (tid == 0) *status = 0;
}

// ---------------------------------------------------------------------------
// Forward and backward substitution (lusolve): L y = P b (L unit lower),
// U x = y.  Thread t owns rows t, t + T, ... and keeps their partial sums in
// registers; at step j the owner of row j finishes it and broadcasts it
// through LDS (two slots alternate, so one barrier per step suffices), and
// every thread folds it into its rows.  Forward sums accumulate in the
// reference's order (increasing j); backward ones in decreasing j.  The
// factor entries of a step are loaded before its barrier.
// ---------------------------------------------------------------------------
template <int RQ>
__global__ void __launch_bounds__(kSolveThreads) k_blk_solve(const double *__restrict__ lu,
                                                             const int *__restrict__ perm, int n,
                                                             const double *__restrict__ rhs,
                                                             double *__restrict__ x) {
    __shared__ double s_v[2];
    const int tid = threadIdx.x, T = blockDim.x;
    double acc[RQ], l[RQ];
#pragma unroll
    for (int q = 0; q < RQ; q++) acc[q] = 0.0;
    for (int j = 0; j < n; j++) {
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            const int i = tid + q * T;
            l[q] = (i > j && i < n) ? lu[(size_t)i * n + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RQ; q++)
            if (tid + q * T == j) {
                const double yj = rhs[perm[j]] - acc[q];
                x[j] = yj;
                s_v[j & 1] = yj;
            }
        __syncthreads();
        const double yj = s_v[j & 1];
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            const int i = tid + q * T;
            if (i > j && i < n) acc[q] += yj * l[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RQ; q++) acc[q] = 0.0;
    for (int j = n - 1; j >= 0; j--) {
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            const int i = tid + q * T;
            l[q] = i < j ? lu[(size_t)i * n + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RQ; q++)
            if (tid + q * T == j) {
                const double xj = (x[j] - acc[q]) / lu[(size_t)j * n + j];
                x[j] = xj;
                s_v[j & 1] = xj;
            }
        __syncthreads();
        const double xj = s_v[j & 1];
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            const int i = tid + q * T;
            if (i < j) acc[q] += xj * l[q];
        }
    }
}

// ---------------------------------------------------------------------------
// compute_temp of the block model: one attempt of the adaptive RK4 (rk4,
// RCutil.c:544-640) per launch.  The attempt takes a full step of size h and
// two half steps from y, and reports the largest difference of the two
// results; the step-size controller runs on the host with the C library's
// pow, the reference's, so the step sequence and every rounding are the
// reference's.  (Over long intervals the step sequence depends on the last
// bits of each error estimate, so a controller whose pow rounds differently
// would follow another, equally valid, sequence.)
//
// The slope k = A^-1 P - C v forms every row's sum in the reference's order
// (increasing column, matvectmult RCutil.c:685-701): thread i owns row i and
// walks the columns of C stored transposed (coalesced across the threads),
// reading v from LDS.
// ---------------------------------------------------------------------------
struct BlkAttempt {
    const double *ct;     // n x n, C transposed: ct[j * n + i] = C[i][j]
    const double *inva;   // n
    const double *power;  // n (package entries already set)
    double *y;            // n, the state
    double *w;            // 9 n work vectors
    double *err;          // out: max |full step - two half steps|
    double h;
    int n;
    int fresh;            // first attempt of the interval: form A^-1 P
    int commit;           // the previous attempt was accepted: y = its full step
    int first;            // first attempt of an rk4 call: k1 = f(y)
};

__global__ void __launch_bounds__(kRkThreads) k_blk_attempt(BlkAttempt A) {
    extern __shared__ double s_v[];   // n: the vector the slope is taken at
    __shared__ double s_red[kRkThreads / kBW];
    const int n = A.n, tid = threadIdx.x, T = blockDim.x;
    const int lane = tid & (kBW - 1), wave = tid / kBW, nw = T / kBW;
    double *p = A.w, *k1 = p + n, *k2 = k1 + n, *k3 = k2 + n, *k4 = k3 + n;
    double *tv = k4 + n, *yt = tv + n, *t1 = yt + n, *t2 = t1 + n;
    double *y = A.y;
    const double *ct = A.ct;
    const double h = A.h;
    // element-wise phases give element i to thread i mod T, as the slope
    // gives it row i: a thread reads back only what it wrote itself
    if (A.fresh)
        for (int i = tid; i < n; i += T) p[i] = A.inva[i] * A.power[i];
    if (A.commit)
        for (int i = tid; i < n; i += T) y[i] = yt[i];

    // k = A^-1 P - C v  (slope_fn_block, temperature_block.c:416-446)
    auto slope = [&](const double *v, double *k) {
        __syncthreads();   // every v[i] written, the previous s_v read
        for (int i = tid; i < n; i += T) s_v[i] = v[i];
        __syncthreads();
        for (int i = tid; i < n; i += T) {
            double s = 0.0;
            for (int j = 0; j < n; j++) s += ct[(size_t)j * n + i] * s_v[j];
            k[i] = p[i] - s;
        }
    };
    // one classic RK4 step of size hh from yv with the given first slope
    // (rk4_core, RCutil.c:440-531)
    auto core = [&](const double *yv, const double *ka, double hh, double *out) {
        for (int i = tid; i < n; i += T) tv[i] = yv[i] + hh / 2.0 * ka[i];
        slope(tv, k2);
        for (int i = tid; i < n; i += T) tv[i] = yv[i] + hh / 2.0 * k2[i];
        slope(tv, k3);
        for (int i = tid; i < n; i += T) tv[i] = yv[i] + hh * k3[i];
        slope(tv, k4);
        for (int i = tid; i < n; i += T) out[i] = yv[i] + hh * (ka[i] + 2 * k2[i] + 2 * k3[i] + k4[i]) / 6.0;
    };

    if (A.first) slope(y, k1);
    core(y, k1, h, yt);
    core(y, k1, h / 2.0, t1);
    slope(t1, k1);   // k1 as a retry of this step reuses it (RCutil.c:568)
    core(t1, k1, h / 2.0, t2);
    double e = 0.0;
    for (int i = tid; i < n; i += T) e = fmax(e, fabs(yt[i] - t2[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, kBW));
    if (lane == 0) s_red[wave] = e;
    __syncthreads();
    if (tid == 0) {
        double mx = s_red[0];
        for (int w = 1; w < nw; w++) mx = fmax(mx, s_red[w]);
        *A.err = mx;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct BlockDev {
    int cap = 0;
    double *lu = nullptr, *c = nullptr, *inva = nullptr, *pw = nullptr, *y = nullptr, *w = nullptr;
    double *scal = nullptr;
    int *perm = nullptr, *status = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static void free_buffers(BlockDev &D) {
    for (void *p : {(void *)D.lu, (void *)D.c, (void *)D.inva, (void *)D.pw, (void *)D.y, (void *)D.w,
                    (void *)D.scal, (void *)D.perm, (void *)D.status})
        if (p) (void)hipFree(p);
    D.lu = D.c = D.inva = D.pw = D.y = D.w = D.scal = nullptr;
    D.perm = D.status = nullptr;
    D.cap = 0;
}

void blk_dev_release(BlockModel &B) {
    if (!B.dev) return;
    BlockDev &D = *B.dev;
    free_buffers(D);
    if (D.ev0) (void)hipEventDestroy(D.ev0);
    if (D.ev1) (void)hipEventDestroy(D.ev1);
    if (D.stream) (void)hipStreamDestroy(D.stream);
    delete B.dev;
    B.dev = nullptr;
}

template <typename T>
static T *dalloc(size_t count) {
    void *p = nullptr;
    BLK_HIP_OK(hipMalloc(&p, count * sizeof(T)));
    return (T *)p;
}

// device buffers for the model's current node count
static BlockDev &ensure(BlockModel &B) {
    const int n = B.m->n_nodes;
    if (n > kMaxBlockNodes)
        fatal("block model too large for the device solver (more than 1024 units)\n");
    if (!B.dev) {
        B.dev = new BlockDev();
        BlockDev &D = *B.dev;
        BLK_HIP_OK(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));
        BLK_HIP_OK(hipEventCreate(&D.ev0));
        BLK_HIP_OK(hipEventCreate(&D.ev1));
    }
    BlockDev &D = *B.dev;
    if (D.cap >= n) return D;
    free_buffers(D);
    const size_t nn = (size_t)n * n;
    D.lu = dalloc<double>(nn);
    D.c = dalloc<double>(nn);
    D.inva = dalloc<double>(n);
    D.pw = dalloc<double>(n);
    D.y = dalloc<double>(n);
    D.w = dalloc<double>((size_t)9 * n);
    D.scal = dalloc<double>(4);
    D.perm = dalloc<int>(n);
    D.status = dalloc<int>(1);
    D.cap = n;
    B.lu_stale = B.c_stale = true;
    return D;
}

static void put(void *dst, const void *src, size_t bytes, hipStream_t s) {
    BLK_HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}
static void get(void *dst, const void *src, size_t bytes, hipStream_t s) {
    BLK_HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
}

static void launch_solve(BlockDev &D, int n) {
    const int rq = (n + kSolveThreads - 1) / kSolveThreads;
    dim3 g(1), b(kSolveThreads);
    if (rq <= 1) k_blk_solve<1><<<g, b, 0, D.stream>>>(D.lu, D.perm, n, D.pw, D.y);
    else if (rq <= 2) k_blk_solve<2><<<g, b, 0, D.stream>>>(D.lu, D.perm, n, D.pw, D.y);
    else if (rq <= 4) k_blk_solve<4><<<g, b, 0, D.stream>>>(D.lu, D.perm, n, D.pw, D.y);
    else if (rq <= 8) k_blk_solve<8><<<g, b, 0, D.stream>>>(D.lu, D.perm, n, D.pw, D.y);
    else if (rq <= 17) k_blk_solve<17><<<g, b, 0, D.stream>>>(D.lu, D.perm, n, D.pw, D.y);
    else fatal("block model too large for the device solver (more than 1024 units)\n");
    BLK_HIP_OK(hipGetLastError());
}

bool blk_dev_present() {
    static int present = -1;
    if (present < 0) {
        int count = 0;
        present = hipGetDeviceCount(&count) == hipSuccess && count > 0 ? 1 : 0;
    }
    return present == 1;
}

void blk_dev_factor(BlockModel &B) {
    BlockDev &D = ensure(B);
    block_model_t &m = *B.m;
    const int n = m.n_nodes;
    const size_t nn = (size_t)n * n;
    put(D.lu, m.b[0], nn * sizeof(double), D.stream);
    const int threads = n <= 256 ? 256 : n <= 512 ? 512 : 1024;
    k_blk_lu<<<1, threads, 0, D.stream>>>(D.lu, D.perm, n, D.status);
    BLK_HIP_OK(hipGetLastError());
    int st = 0;
    get(&st, D.status, sizeof(int), D.stream);
    BLK_HIP_OK(hipStreamSynchronize(D.stream));
    if (st) fatal("singular matrix in lupdcmp\n");
    // host copies of the factors, as the reference keeps them in the model
    get(m.lu[0], D.lu, nn * sizeof(double), D.stream);
    get(m.p, D.perm, n * sizeof(int), D.stream);
    BLK_HIP_OK(hipStreamSynchronize(D.stream));
    B.lu_stale = false;
}

void blk_dev_steady(BlockModel &B, const double *power, double *temp) {
    BlockDev &D = ensure(B);
    block_model_t &m = *B.m;
    const int n = m.n_nodes;
    BLK_HIP_OK(hipEventRecord(D.ev0, D.stream));
    if (B.lu_stale) blk_dev_factor(B);
    put(D.pw, power, n * sizeof(double), D.stream);
    launch_solve(D, n);
    BLK_HIP_OK(hipEventRecord(D.ev1, D.stream));
    get(temp, D.y, n * sizeof(double), D.stream);
    BLK_HIP_OK(hipStreamSynchronize(D.stream));
    float ms = 0;
    BLK_HIP_OK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    B.stats.last_solve_ms = ms;
}

void blk_dev_transient(BlockModel &B, const double *power, double *temp, double dt) {
    BlockDev &D = ensure(B);
    block_model_t &m = *B.m;
    const int n = m.n_nodes;
    if (B.c_stale) {
        // C transposed, so that the slope's row sums read it coalesced
        std::vector<double> ct((size_t)n * n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) ct[(size_t)j * n + i] = m.c[i][j];
        put(D.c, ct.data(), ct.size() * sizeof(double), D.stream);
        put(D.inva, m.inva, n * sizeof(double), D.stream);
        BLK_HIP_OK(hipStreamSynchronize(D.stream));
        B.c_stale = false;
    }
    put(D.pw, power, n * sizeof(double), D.stream);
    put(D.y, temp, n * sizeof(double), D.stream);
    BLK_HIP_OK(hipEventRecord(D.ev0, D.stream));
    BlkAttempt a{D.c, D.inva, D.pw, D.y, D.w, D.scal, 0.0, n, 1, 0, 0};
    const size_t lds = (size_t)n * sizeof(double);
    // the interval loop of compute_temp_block (temperature_block.c:475-481)
    // and rk4's controller (RCutil.c:598-621)
    int calls = 0;
    double t = 0, new_h = kMinStep;
    while (t < dt && new_h >= kMinStep * kDelta) {
        double h = new_h, nt = h;
        a.first = 1;
        do {
            h = nt;
            a.h = h;
            k_blk_attempt<<<1, kRkThreads, lds, D.stream>>>(a);
            BLK_HIP_OK(hipGetLastError());
            a.fresh = a.commit = a.first = 0;
            double mx = 0;
            get(&mx, D.scal, sizeof mx, D.stream);
            BLK_HIP_OK(hipStreamSynchronize(D.stream));
            if (mx <= 0.01) {
                nt = 0.95 * h * std::pow(std::fabs(0.01 / mx), 0.2);
                if (nt > 5.0 * h) nt = 5.0 * h;
            } else {
                nt = 0.95 * h * std::pow(std::fabs(0.01 / mx), 0.25);
                if (nt < h / 10.0) nt = h / 10.0;
            }
        } while (nt < h);
        a.commit = 1;   // the accepted full step becomes y at the next launch
        const double rest = dt - t - h;
        new_h = nt < rest ? nt : rest;
        t += h;
        calls++;
    }
    BLK_HIP_OK(hipEventRecord(D.ev1, D.stream));
    // the last accepted full step (or the unchanged state when none was taken)
    get(temp, a.commit ? D.w + (size_t)6 * n : D.y, n * sizeof(double), D.stream);
    BLK_HIP_OK(hipStreamSynchronize(D.stream));
    float ms = 0;
    BLK_HIP_OK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    B.stats.last_solve_ms = ms;
    B.stats.rk4_calls = calls;
}

}  // namespace hs
