// Kernels and solve drivers of the grid thermal model for gfx950.
//
// Discrete model (the reference's per-cell form, temperature_grid.c:26-136,
// :2426-2608, :2857-3601): for every grid cell and package node c
//
//   flux_c = sum_nbr g (T_nbr - T_c) [+ advection]      (conduction in)
//   G T    = gam_c T_c - flux_c                          (linear operator)
//   rhs_c  = P_c + gam_c * T_ambient [+ inlet source]
//   steady:    G T = rhs                 transient: dT/dt = (rhs - G T) / C
//
// with g the face conductances (1/rx for row neighbours, 1/ry for column
// neighbours, 1/rz[n] between layers n and n+1, per-cell values in detailed-3D
// and microchannel models), gam the ambient conductance (sink and PCB cells),
// and the package-node rings coupling boundary rows to the 12 (or 28) package
// nodes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "hs_kernels.h"

namespace hs {

#define HIP_OK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            fatal(std::string("HIP error ") + hipGetErrorString(e_) + " in " #x "\n");   \
    } while (0)

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Sums (or maxes) N values over the workgroup; the result is valid in thread 0.
template <int N, bool MAX = false>
__device__ __forceinline__ void block_reduce(double (&v)[N]) {
    __shared__ double red[kBlock / kWave][N];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < N; q++) v[q] = MAX ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < N; q++) red[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            double a = red[0][q];
            for (int k = 1; k < kBlock / kWave; k++) a = MAX ? fmax(a, red[k][q]) : a + red[k][q];
            v[q] = a;
        }
    }
    __syncthreads();
}

// Each workgroup stores its partials (valid in thread 0) in its own slot with
// plain stores; a finish kernel launched next combines the slots of each
// problem in slot order.  No device-scope fences in the hot kernels (on a
// multi-XCD part such a fence writes back the XCD's L2) and no contended
// atomics: the kernel boundary makes the slots visible.
template <int N>
__device__ __forceinline__ void write_partials(const ProbDesc &d, int slot, const double (&v)[N]) {
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < N; q++) d.partial[slot * kRed + q] = v[q];
}

// Whole workgroup: combine a problem's slots in slot order (thread 0 result).
template <int N, bool MAX = false>
__device__ __forceinline__ void reduce_slots(const ProbDesc &d, double (&v)[N], int nslots = -1) {
    if (nslots < 0) nslots = d.ntiles + 1;
#pragma unroll
    for (int q = 0; q < N; q++) v[q] = 0.0;
    for (int s = threadIdx.x; s < nslots; s += kBlock)
#pragma unroll
        for (int q = 0; q < N; q++) {
            const double x = d.partial[s * kRed + q];
            v[q] = MAX ? fmax(v[q], x) : v[q] + x;
        }
    block_reduce<N, MAX>(v);
}

// Workgroup -> work item, XCD-aware: the hardware hands workgroup b to XCD
// b % 8, so give each XCD a contiguous run of work items (neighbouring tiles,
// whose halos then stay in that XCD's L2 from one kernel to the next).
__device__ __forceinline__ int work_index() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct Cell {
    int n, i, j;
    unsigned c;   // node index within the problem (< 2^29 for every supported grid)
    bool active;
};

// Element access through a 32-bit byte offset from a wave-uniform base: the
// address is an SGPR pair plus one VGPR (global_load ... saddr), half the
// registers of a 64-bit address per load.  Every buffer these take is device
// memory, so the pointers are cast to the global address space: loads through
// generic pointers read from descriptors would otherwise be flat loads, which
// also count against the LDS/scalar counter and so wait behind each other.
#define HS_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ T ld_at(const T *base, unsigned i) {
    return *(const HS_GLOBAL T *)((const HS_GLOBAL char *)base + i * (unsigned)sizeof(T));
}
template <typename T> __device__ __forceinline__ void st_at(T *base, unsigned i, T v) {
    *(HS_GLOBAL T *)((HS_GLOBAL char *)base + i * (unsigned)sizeof(T)) = v;
}

// A tile is TX columns x (TY * rpt) rows of one layer; thread (tx, ty) owns
// the rpt consecutive rows starting at ty * rpt of its column.  Cells beyond
// the grid or beyond rpt are inactive; their index is clamped to a valid cell
// so that loads stay in bounds.
__device__ __forceinline__ Cell tile_cell(const ProbDesc &d, int tile, int r) {
    const int tpl = d.nti * d.ntj;
    const int n = tile / tpl;
    const int rem = tile - n * tpl;
    const int ti = rem / d.ntj, tj = rem - ti * d.ntj;
    const int sh = d.tx_shift;
    const int tx = threadIdx.x & ((1 << sh) - 1), ty = threadIdx.x >> sh;
    const int i = ti * ((kBlock >> sh) * d.rpt) + ty * d.rpt + r, j = (tj << sh) + tx;
    Cell q;
    q.active = r < d.rpt && i < d.nr && j < d.nc;
    q.n = n;
    q.i = min(i, d.nr - 1);
    q.j = min(j, d.nc - 1);
    q.c = ((unsigned)n * d.nr + q.i) * d.nc + q.j;
    return q;
}

__device__ __forceinline__ int tile_layer(const ProbDesc &d, int tile) { return tile / (d.nti * d.ntj); }

// Runs a kernel body over the cells a thread owns in a tile, RM cells at a
// time: load(q) performs every global read of a cell and returns the values,
// fin(q, v) does the arithmetic and the stores.  All reads of a group are
// issued before its first store, so each thread keeps RM cells' reads in
// flight instead of waiting on one cell at a time.
template <int RM, class Load, class Fin>
__device__ __forceinline__ void tile_loop(const ProbDesc &d, int tile, Load load, Fin fin) {
    using V = decltype(load(Cell{}));
    for (int r0 = 0; r0 < d.rpt; r0 += RM) {
        Cell q[RM];
        V v[RM];
#pragma unroll
        for (int k = 0; k < RM; k++) {
            q[k] = tile_cell(d, tile, r0 + k);
            if (r0 + k < d.rpt) v[k] = load(q[k]);
        }
#pragma unroll
        for (int k = 0; k < RM; k++)
            if (q[k].active) fin(q[k], v[k]);
    }
}

// cells in flight per thread: FP32 homogeneous 8, FP64 or heterogeneous 4,
// FP64 heterogeneous 2 (register budget)
template <typename T, bool HET> struct InFlight {
    static constexpr int value = (sizeof(T) == 4 ? 8 : 4) / (HET ? 2 : 1);
};

template <typename T> struct Coef;
template <> struct Coef<double> {
    __device__ static const double *E(const ProbDesc &d) { return d.gE; }
    __device__ static const double *S(const ProbDesc &d) { return d.gS; }
    __device__ static const double *D(const ProbDesc &d) { return d.gD; }
    __device__ static const double *cinv(const ProbDesc &d) { return d.cinvc; }
    __device__ static const double *dinv(const ProbDesc &d) { return d.dinvc; }
    __device__ static const double *aN(const ProbDesc &d) { return d.aN; }
    __device__ static const double *aS(const ProbDesc &d) { return d.aS; }
    __device__ static const double *aE(const ProbDesc &d) { return d.aE; }
    __device__ static const double *aW(const ProbDesc &d) { return d.aW; }
    __device__ static const double *aSelf(const ProbDesc &d) { return d.aSelf; }
};
template <> struct Coef<float> {
    __device__ static const float *E(const ProbDesc &d) { return d.fE; }
    __device__ static const float *S(const ProbDesc &d) { return d.fS; }
    __device__ static const float *D(const ProbDesc &d) { return d.fD; }
    __device__ static const float *cinv(const ProbDesc &d) { return d.fcinv; }
    __device__ static const float *dinv(const ProbDesc &d) { return d.fdinv; }
    __device__ static const float *aN(const ProbDesc &d) { return d.faN; }
    __device__ static const float *aS(const ProbDesc &d) { return d.faS; }
    __device__ static const float *aE(const ProbDesc &d) { return d.faE; }
    __device__ static const float *aW(const ProbDesc &d) { return d.faW; }
    __device__ static const float *aSelf(const ProbDesc &d) { return d.faSelf; }
};

// input loaders: plain vector, or an RK stage base + a*k
template <typename T> struct LdPlain {
    const T *__restrict__ x;
    __device__ __forceinline__ T operator()(unsigned c) const { return ld_at(x, c); }
};
template <typename T> struct LdStage {
    const T *__restrict__ b;
    const T *__restrict__ k;
    T a;
    __device__ __forceinline__ T operator()(unsigned c) const { return ld_at(b, c) + a * ld_at(k, c); }
};
template <typename T> struct LdDir {   // CG direction p = z + beta * p_old
    const T *__restrict__ z;
    const T *__restrict__ p;
    T beta;
    __device__ __forceinline__ T operator()(unsigned c) const {
        return beta == T(0) ? ld_at(z, c) : ld_at(z, c) + beta * ld_at(p, c);
    }
};

// The values around a cell that the operator reads: the state at the cell
// and its six neighbours (the cell itself where a neighbour does not exist)
// and, for heterogeneous models, the face conductances and advection
// coefficients.
template <typename T, bool HET> struct Pts {
    T c, n, s, e, w, u, dn;
    T gn, gs, ge, gw, gu, gd;
    T an, as, ae, aw, self;
};

template <typename T, bool HET, class L>
__device__ __forceinline__ void load_pts(const ProbDesc &d, const L &ld, const Cell &q, Pts<T, HET> &p) {
    const int nr = d.nr, nc = d.nc, nl = d.nl, n = q.n, i = q.i, j = q.j;
    const unsigned c = q.c, plane = (unsigned)nr * nc;
    const unsigned cn = i > 0 ? c - nc : c, cs = i < nr - 1 ? c + nc : c;
    const unsigned ce = j < nc - 1 ? c + 1 : c, cw = j > 0 ? c - 1 : c;
    const unsigned cu = n > 0 ? c - plane : c, cd = n < nl - 1 ? c + plane : c;
    p.c = ld(c);
    p.n = ld(cn);
    p.s = ld(cs);
    p.e = ld(ce);
    p.w = ld(cw);
    p.u = ld(cu);
    p.dn = ld(cd);
    if (HET) {
        const T *gE = Coef<T>::E(d), *gS = Coef<T>::S(d), *gD = Coef<T>::D(d);
        p.gn = ld_at(gS, cn);
        p.gs = ld_at(gS, c);
        p.ge = ld_at(gE, c);
        p.gw = ld_at(gE, cw);
        p.gu = ld_at(gD, cu);
        p.gd = ld_at(gD, c);
        if (d.has_adv) {
            p.an = ld_at(Coef<T>::aN(d), c);
            p.as = ld_at(Coef<T>::aS(d), c);
            p.ae = ld_at(Coef<T>::aE(d), c);
            p.aw = ld_at(Coef<T>::aW(d), c);
            p.self = ld_at(Coef<T>::aSelf(d), c);
        }
    }
}

// The state of the package nodes a layer's edges couple to (N, S, E, W),
// read once per tile (every cell of a tile is in one layer).
template <typename T, class L>
__device__ __forceinline__ void load_ring(const ProbDesc &d, const L &ld, int n, T (&ring)[4]) {
    const unsigned ncell = (unsigned)d.nl * d.nr * d.nc;
#pragma unroll
    for (int e = 0; e < 4; e++) ring[e] = ld(ncell + max(d.ring_node[n][e], 0));
}

// Conduction (and advection) into the cell, the diagonal of G at the cell
// and its ambient conductance, from the loaded values.  A missing neighbour
// has zero conductance and reads the cell itself, so it adds exactly zero.
template <typename T, bool HET>
__device__ __forceinline__ void stencil(const ProbDesc &d, const Cell &q, const Pts<T, HET> &p,
                                        const T (&ring)[4], T &flux, T &diag, T &gam) {
    const int nr = d.nr, nc = d.nc, nl = d.nl, n = q.n, i = q.i, j = q.j;
    const bool hn = i > 0, hs = i < nr - 1, he = j < nc - 1, hw = j > 0, hu = n > 0, hd = n < nl - 1;
    const T xc = p.c;
    T gn, gs, ge, gw, gu, gd;
    if (!HET) {
        const T gx = (T)d.gx[n], gy = (T)d.gy[n];
        gn = hn ? gx : T(0);
        gs = hs ? gx : T(0);
        ge = he ? gy : T(0);
        gw = hw ? gy : T(0);
        gu = hu ? (T)d.gz[max(n - 1, 0)] : T(0);
        gd = hd ? (T)d.gz[n] : T(0);
    } else {
        gn = hn ? p.gn : T(0);
        gs = hs ? p.gs : T(0);
        ge = he ? p.ge : T(0);
        gw = hw ? p.gw : T(0);
        gu = hu ? p.gu : T(0);
        gd = hd ? p.gd : T(0);
    }
    flux = gn * (p.n - xc);
    flux += gs * (p.s - xc);
    flux += ge * (p.e - xc);
    flux += gw * (p.w - xc);
    flux += gu * (p.u - xc);
    flux += gd * (p.dn - xc);
    diag = gn + gs + ge + gw + gu + gd;
    if (HET && d.has_adv) {
        // coolant advection Cv Q(nbr->c) (T_nbr + T_c)/2 and the outlet sink
        // (temperature_grid.c:3256-3590); a* hold Cv Q / 2 per face
        const T xn = hn ? p.n : T(0), xs = hs ? p.s : T(0), xe = he ? p.e : T(0), xw = hw ? p.w : T(0);
        flux += p.an * (xn + xc) + p.as * (xs + xc) + p.ae * (xe + xc) + p.aw * (xw + xc) + p.self * xc;
        diag -= p.an + p.as + p.ae + p.aw + p.self;
    }
    const int *rn = d.ring_node[n];
    const T rgn = (T)d.ring_gns[n], rge = (T)d.ring_gew[n];
    if (i == 0 && rn[0] >= 0) { flux += rgn * (ring[0] - xc); diag += rgn; }
    if (i == nr - 1 && rn[1] >= 0) { flux += rgn * (ring[1] - xc); diag += rgn; }
    if (j == nc - 1 && rn[2] >= 0) { flux += rge * (ring[2] - xc); diag += rge; }
    if (j == 0 && rn[3] >= 0) { flux += rge * (ring[3] - xc); diag += rge; }
    gam = (T)d.gamb[n];
    if (HET && d.bshift) gam += (T)d.bshift[q.c];
    diag += gam;
}

// Diagonal of G at a cell without reading the state (pointwise kernels).
template <typename T, bool HET>
__device__ __forceinline__ T cell_dinv(const ProbDesc &d, const Cell &q) {
    if (HET) return Coef<T>::dinv(d)[q.c];
    const int nr = d.nr, nc = d.nc, nl = d.nl, n = q.n, i = q.i, j = q.j;
    const T gx = (T)d.gx[n], gy = (T)d.gy[n];
    T diag = (T)d.gamb[n];
    if (i > 0) diag += gx;
    if (i < nr - 1) diag += gx;
    if (j < nc - 1) diag += gy;
    if (j > 0) diag += gy;
    if (n > 0) diag += (T)d.gz[n - 1];
    if (n < nl - 1) diag += (T)d.gz[n];
    const int *rn = d.ring_node[n];
    if (i == 0 && rn[0] >= 0) diag += (T)d.ring_gns[n];
    if (i == nr - 1 && rn[1] >= 0) diag += (T)d.ring_gns[n];
    if (j == nc - 1 && rn[2] >= 0) diag += (T)d.ring_gew[n];
    if (j == 0 && rn[3] >= 0) diag += (T)d.ring_gew[n];
    return T(1) / diag;
}

// Package nodes: flux into node k from its ring of boundary cells and from
// the package nodes it couples to (single_iteration_steady_pack :2090-2393,
// slope_fn_pack :2857-3093); results in shared memory.  All threads call it.
template <typename T, class L>
__device__ __forceinline__ void package_terms(const ProbDesc &d, const L &ld, bool transient,
                                              T *xflux, T *xdiag) {
    const int nr = d.nr, nc = d.nc;
    const long plane = (long)nr * nc, ncell = plane * d.nl;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    for (int k = w; k < d.nx; k += kBlock / kWave) {
        const int L_ = d.x_ring_layer[k];
        const T xk = ld(ncell + k);
        double s = 0.0;
        int m = 0;
        if (L_ >= 0) {
            const int e = d.x_ring_edge[k];
            m = (e < 2) ? nc : nr;
            const long b = (long)L_ * plane;
            for (int t = lane; t < m; t += kWave) {
                long c;
                if (e == 0) c = b + t;
                else if (e == 1) c = b + (long)(nr - 1) * nc + t;
                else if (e == 2) c = b + (long)t * nc + (nc - 1);
                else c = b + (long)t * nc;
                s += (double)(ld(c) - xk);
            }
            s = wave_sum(s);
        }
        if (lane == 0) {
            const T gr = L_ >= 0 ? (T)d.x_ring_g[k] : T(0);
            T flux = gr * (T)s, diag = gr * (T)m;
            for (int q = 0; q < d.x_ncpl[k]; q++) {
                if (transient && d.x_cpl_steady_only[k][q]) continue;
                const T g = (T)d.x_cpl_g[k][q];
                flux += g * (ld(ncell + d.x_cpl_node[k][q]) - xk);
                diag += g;
            }
            diag += (T)d.x_gamb[k];
            xflux[k] = flux;
            xdiag[k] = diag;
        }
    }
    __syncthreads();
}

// diagonal of G at package node k: ambient, ring and coupling conductances
template <typename T> __device__ __forceinline__ T pkg_diag(const ProbDesc &d, int k) {
    const int L_ = d.x_ring_layer[k];
    T diag = (T)d.x_gamb[k];
    if (L_ >= 0) diag += (T)d.x_ring_g[k] * (T)(d.x_ring_edge[k] < 2 ? d.nc : d.nr);
    for (int q = 0; q < d.x_ncpl[k]; q++) diag += (T)d.x_cpl_g[k][q];
    return diag;
}

__device__ __forceinline__ bool problem_skip(const ProbDesc &d, int flag_slot) {
    return flag_slot >= 0 && d.scal[flag_slot] != 0.0;
}

// sequence mode: a problem whose intervals are all done takes no part
__device__ __forceinline__ bool seq_finished(const ProbDesc &d, int seq) {
    return seq && d.scal[S_STEP] >= d.scal[S_NSTEPS];
}

// ---------------------------------------------------------------------------
// block <-> grid translation (xlate_vector_b2g :1888-1929, xlate_temp_g2b
// :1932-2015)
// ---------------------------------------------------------------------------
// Power of plane cell cc of layer n for the FP32 Taylor groups, from its
// cell_uw record (units u, weights w) and the layer's block values vb: the
// power density of each block times its occupancy, times the cell area
// (blist_avg x area, temperature_grid.c:1908-1910), with the weights of up to
// four blocks in FP32.  k_b2g and k_tay_first4z both map power with it.
__device__ __forceinline__ double four_block_power(const int4 &w, double v0, double v1, double v2, double v3) {
#pragma clang fp contract(on)
    return (double)__int_as_float(w.x) * v0 + (double)__int_as_float(w.y) * v1 +
           (double)__int_as_float(w.z) * v2 + (double)__int_as_float(w.w) * v3;
}
__device__ __forceinline__ double cell_power_uw(const ProbDesc &d, int n, unsigned cc, const int4 &u,
                                                const int4 &w, const double *vb) {
    if (u.x >= 0) return four_block_power(w, vb[u.x], vb[u.y], vb[u.z], vb[u.w]);
    const int *ptr = d.map_ptr[n], *unit = d.map_unit[n];
    const double *occ = d.map_occ[n], *area = d.unit_area[n];
    double val = 0.0;
    for (int e = ptr[cc]; e < ptr[cc + 1]; e++) val += (occ[e] * vb[unit[e]]) / area[unit[e]];
    return val * d.cell_area;
}
constexpr int kPowTab = 1024;   // blocks of a layer whose powers k_b2g tabulates in LDS

// out = sum over the cell's blocks of occ * v_u (/ area_u * cell area for
// power); with add_sources the ambient / inlet constants are added (rhs).
// from_seq: 1 reads the block vector of the current sequence step, 2 the
// block output vector (a steady result handed to a transient), 0 the staging
// vector.
//
// With power32 (V_POWER, no sources), the cell power goes to the FP32 buffer
// out_buf; the tiles of a layer whose blocks all carry zero power write
// nothing and mark the layer in lzero instead (every workgroup of the layer
// writes the same mark, which the next kernel reads).
__device__ __forceinline__ const double *b2g_source(const ProbDesc &d, int from_seq) {
    return from_seq == 1 ? d.seq_in + (d.seq_one ? 0L : (long)d.scal[S_STEP] * (d.nblocks + d.nx))
           : from_seq == 2 ? d.blk_out : d.blk_in;
}
// whole workgroup: whether some block of layer n carries power in vb
__device__ __forceinline__ bool layer_has_power(const ProbDesc &d, int n, const double *vb) {
    __shared__ int nonzero;
    const int nu = (n + 1 < d.nl ? d.layer_base[n + 1] : d.nblocks) - d.layer_base[n];
    if (threadIdx.x == 0) nonzero = 0;
    __syncthreads();
    for (int u = threadIdx.x; u < nu; u += kBlock)
        if (vb[u] != 0.0) nonzero = 1;
    __syncthreads();
    return nonzero != 0;
}
__global__ void __launch_bounds__(kBlock) k_b2g(const ProbDesc *__restrict__ D,
                                                const WorkItem *__restrict__ W, int type,
                                                int out_buf, int add_sources, int from_seq) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (seq_finished(d, from_seq == 1)) return;
    double *out = d.b64[out_buf];
    auto put = [&](unsigned c, double v) { st_at(out, c, v); };
    const double *b = b2g_source(d, from_seq);
    const long ncell = (long)d.nl * d.nr * d.nc;
    if (wi.tile < d.ntiles) {
        const int n = tile_layer(d, wi.tile);
        const int *ptr = d.map_ptr[n];
        const int *unit = d.map_unit[n];
        const double *occ = d.map_occ[n];
        const double *vb = b + d.layer_base[n];
        const double *area = d.unit_area[n];
        const double src = d.gamb[n] * d.ambient;
        // A layer whose blocks all carry zero power (bond, spreader and sink
        // layers of a chip stack) maps to zero without reading its map.
        const bool mapped = type != V_POWER || layer_has_power(d, n, vb);
        if (!mapped) {
            tile_loop<8>(d, wi.tile, [&](const Cell &) { return 0; }, [&](const Cell &q, int) {
                double val = 0.0;
                if (add_sources) {
                    val += src;
                    if (d.bconst) val += ld_at(d.bconst, q.c);
                }
                st_at(out, q.c, val);
            });
            return;
        }
        // a cell's first two map entries are gathered for 8 cells at a time
        // (three dependent loads: entry range, entries, block values); cells
        // covered by more blocks finish the rest in a loop
        struct Ent {
            int e0, e1;
            int u0, u1;
            double o0, o1;
        };
        tile_loop<8>(
            d, wi.tile,
            [&](const Cell &q) {
                Ent t;
                const unsigned cc = (unsigned)q.i * d.nc + q.j;
                t.e0 = ld_at(ptr, cc);
                t.e1 = ld_at(ptr, cc + 1);
                const unsigned f = (unsigned)t.e0, g = (unsigned)max(min(t.e0 + 1, t.e1 - 1), t.e0);
                t.u0 = ld_at(unit, f);
                t.o0 = ld_at(occ, f);
                t.u1 = ld_at(unit, g);
                t.o1 = ld_at(occ, g);
                return t;
            },
            [&](const Cell &q, const Ent &t) {
                double val = 0.0;
                const int cnt = t.e1 - t.e0;
                if (type == V_POWER) {
                    // power density of each block times its occupancy, times the cell area
                    if (cnt > 0) val += (t.o0 * vb[t.u0]) / area[t.u0];
                    if (cnt > 1) val += (t.o1 * vb[t.u1]) / area[t.u1];
                    for (int e = t.e0 + 2; e < t.e1; e++) val += (occ[e] * vb[unit[e]]) / area[unit[e]];
                    val *= d.cell_area;
                } else {
                    if (cnt > 0) val += t.o0 * vb[t.u0];
                    if (cnt > 1) val += t.o1 * vb[t.u1];
                    for (int e = t.e0 + 2; e < t.e1; e++) val += occ[e] * vb[unit[e]];
                }
                if (add_sources) {
                    val += src;
                    if (d.bconst) val += ld_at(d.bconst, q.c);
                }
                put(q.c, val);
            });
    } else {
        const int k = threadIdx.x;
        if (k < d.nx) {
            double v = (type == V_POWER) ? 0.0 : b[d.nblocks + k];
            if (add_sources) v += d.x_gamb[k] * d.ambient;
            out[ncell + k] = v;
        }
    }
}

// The cell power alone in FP32 for the FP32 Taylor groups (F_PW): the tiles
// of a layer whose blocks all carry zero power write nothing and mark the
// layer in lzero instead (every workgroup of the layer writes the same mark,
// which the next kernel reads); the package nodes take none.  A layer with
// one-block codes reads a cell's code and, at a block boundary, its record,
// with the layer's block powers and the power of a cell inside each block
// tabulated in LDS; otherwise every cell reads its record.
__global__ void __launch_bounds__(kBlock) k_b2g_power32(const ProbDesc *__restrict__ D,
                                                        const WorkItem *__restrict__ W, int from_seq) {
    __shared__ float cpt[kPowTab];
    __shared__ double bpt[kPowTab];
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (seq_finished(d, from_seq == 1) || wi.tile >= d.ntiles) return;
    const int n = tile_layer(d, wi.tile);
    const double *vb = b2g_source(d, from_seq) + d.layer_base[n];
    const bool mapped = layer_has_power(d, n, vb);
    if (threadIdx.x == 0) d.lzero[n] = mapped ? 0 : 1;
    if (!mapped) return;
    float *out32 = d.b32[F_PW];
    // a cell of up to four blocks: their weights times their power; a cell
    // of five or more: its map entries
    const int4 *uw = d.cell_uw[n];
    const int nu = (n + 1 < d.nl ? d.layer_base[n + 1] : d.nblocks) - d.layer_base[n];
    if (d.cell_u16[n] && nu <= kPowTab) {
        for (int u = threadIdx.x; u < nu; u += kBlock) {
            cpt[u] = (float)((double)d.unit_w[n][u] * vb[u]);
            bpt[u] = vb[u];
        }
        __syncthreads();
        // the cells inside one block, then the tile's boundary cells from
        // its list
        const unsigned short *code = d.cell_u16[n];
        tile_loop<8>(
            d, wi.tile, [&](const Cell &q) { return (unsigned)code[(unsigned)q.i * d.nc + q.j]; },
            [&](const Cell &q, unsigned k) {
                if (k != kCellMixed) st_at(out32, q.c, cpt[k]);
            });
        const int rem = wi.tile - n * d.nti * d.ntj;
        const int e0 = d.mix_off[n][rem], e1 = d.mix_off[n][rem + 1];
        const unsigned lo = (unsigned)n * d.nr * d.nc;
        for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
            const unsigned cc = (unsigned)d.mix_cell[n][e];
            const int4 u = d.mix_uw[n][2 * e], w = d.mix_uw[n][2 * e + 1];
            const float p = u.x >= 0 ? (float)four_block_power(w, bpt[u.x], bpt[u.y], bpt[u.z], bpt[u.w])
                                     : (float)cell_power_uw(d, n, cc, u, w, vb);
            st_at(out32, lo + cc, p);
        }
        return;
    }
    struct Rec {
        int4 u, w;
    };
    tile_loop<4>(
        d, wi.tile,
        [&](const Cell &q) {
            const unsigned cc = (unsigned)q.i * d.nc + q.j;
            return Rec{uw[2 * cc], uw[2 * cc + 1]};
        },
        [&](const Cell &q, const Rec &r) {
            st_at(out32, q.c, (float)cell_power_uw(d, n, (unsigned)q.i * d.nc + q.j, r.u, r.w, vb));
        });
}

// One wavefront per block of the vector; work items: (problem, first block).
// in_buf < 0: the current transient state; to_seq: write the current step of
// the output sequence instead of the staging vector.
__global__ void __launch_bounds__(kBlock) k_g2b(const ProbDesc *__restrict__ D,
                                                const WorkItem *__restrict__ W, int in_buf,
                                                int to_seq) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (seq_finished(d, to_seq)) return;
    const double *g = d.b64[in_buf >= 0 ? in_buf : B_Y0 + (int)d.scal[S_YPAR]];
    double *bo = to_seq ? d.seq_out + (long)d.scal[S_STEP] * (d.nblocks + d.nx) : d.blk_out;
    const long plane = (long)d.nr * d.nc, ncell = plane * d.nl;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int u = wi.tile + w;
    if (wi.tile == -1) {   // package nodes
        if (threadIdx.x < d.nx) bo[d.nblocks + threadIdx.x] = g[ncell + threadIdx.x];
        return;
    }
    if (u >= d.nblocks) return;
    const int n = d.blk_layer[u];
    const int4 r = d.blk_range[u];
    const double *gl = g + n * plane;
    double res;
    if (d.map_mode == GRID_CENTER) {
        const int ci1 = (r.x + r.y) / 2, cj1 = (r.z + r.w) / 2;
        const int ci2 = ci1 - !((r.y - r.x) % 2), cj2 = cj1 - !((r.w - r.z) % 2);
        res = (gl[(long)ci1 * d.nc + cj1] + gl[(long)ci2 * d.nc + cj1] + gl[(long)ci1 * d.nc + cj2] +
               gl[(long)ci2 * d.nc + cj2]) / 4;
    } else {
        const int bw = r.w - r.z, cnt = (r.y - r.x) * bw;
        double s = 0.0, mn = gl[(long)r.x * d.nc + r.z], mx = mn;
        for (int t = lane; t < cnt; t += kWave) {
            const double v = gl[(long)(r.x + t / bw) * d.nc + r.z + t % bw];
            s += v;
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
        s = wave_sum(s);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, o, kWave));
            mx = fmax(mx, __shfl_xor(mx, o, kWave));
        }
        res = d.map_mode == GRID_AVG ? s / cnt : d.map_mode == GRID_MIN ? mn : mx;
    }
    if (lane == 0) bo[u] = res;
}

// ---------------------------------------------------------------------------
// steady: preconditioned CG (Jacobi), FP64 or FP32 (inner solve of the mixed
// iterative refinement)
// ---------------------------------------------------------------------------
// r = rhs - G x; z = D^-1 r; partials r.z, r.r, rhs.rhs.  The last workgroup
// initialises the problem's CG scalars.
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_cg_init(const ProbDesc *__restrict__ D,
                                                    const WorkItem *__restrict__ W, int xb,
                                                    int rb, int zb, int rhsb, double rtol2,
                                                    int max_iters) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, S_DONE)) return;
    T *x = sizeof(T) == 8 ? (T *)d.b64[xb] : (T *)d.b32[xb];
    T *r = sizeof(T) == 8 ? (T *)d.b64[rb] : (T *)d.b32[rb];
    T *z = sizeof(T) == 8 ? (T *)d.b64[zb] : (T *)d.b32[zb];
    const T *rhs = sizeof(T) == 8 ? (const T *)d.b64[rhsb] : (const T *)d.b32[rhsb];
    LdPlain<T> ld{x};
    double v[3] = {0, 0, 0};
    if (wi.tile < d.ntiles) {
        T ring[4];
        load_ring<T>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<T, HET> p;
            T b;
        };
        tile_loop<InFlight<T, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V a{};
                load_pts<T, HET>(d, ld, q, a.p);
                a.b = ld_at(rhs, q.c);
                return a;
            },
            [&](const Cell &q, const V &a) {
                T flux, diag, gam;
                stencil<T, HET>(d, q, a.p, ring, flux, diag, gam);
                const T rr = a.b + flux - gam * a.p.c;
                const T zz = rr / diag;
                st_at(r, q.c, rr);
                st_at(z, q.c, zz);
                v[0] += (double)rr * zz;
                v[1] += (double)rr * rr;
                v[2] += (double)a.b * a.b;
            });
    } else {
        __shared__ T xf[kMaxExtra], xd[kMaxExtra];
        package_terms<T>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            const T xc = x[c];
            const T rr = rhs[c] + xf[k] - (T)d.x_gamb[k] * xc;
            const T zz = rr / xd[k];
            r[c] = rr;
            z[c] = zz;
            v[0] = (double)rr * zz;
            v[1] = (double)rr * rr;
            v[2] = (double)rhs[c] * rhs[c];
        }
    }
    block_reduce<3>(v);
    write_partials<3>(d, wi.tile, v);
}

// p = z + beta p_old;  q = G p;  partial p.q -> alpha = rz / pq
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_cg_dir(const ProbDesc *__restrict__ D,
                                                   const WorkItem *__restrict__ W, int zb,
                                                   int p0b, int qb, int done_slot) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, done_slot) || (done_slot != S_DONE && problem_skip(d, S_DONE))) return;
    const int par = (int)d.scal[S_PPAR];
    auto buf = [&](int b) -> T * { return sizeof(T) == 8 ? (T *)d.b64[b] : (T *)d.b32[b]; };
    const T *z = buf(zb), *pold = buf(p0b + par);
    T *pnew = buf(p0b + 1 - par), *qv = buf(qb);
    LdDir<T> ld{z, pold, (T)d.scal[S_BETA]};
    double v[1] = {0};
    if (wi.tile < d.ntiles) {
        T ring[4];
        load_ring<T>(d, ld, tile_layer(d, wi.tile), ring);
        tile_loop<InFlight<T, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                Pts<T, HET> p{};
                load_pts<T, HET>(d, ld, q, p);
                return p;
            },
            [&](const Cell &q, const Pts<T, HET> &p) {
                T flux, diag, gam;
                stencil<T, HET>(d, q, p, ring, flux, diag, gam);
                const T Gp = gam * p.c - flux;
                st_at(pnew, q.c, p.c);
                st_at(qv, q.c, Gp);
                v[0] += (double)p.c * Gp;
            });
    } else {
        __shared__ T xf[kMaxExtra], xd[kMaxExtra];
        package_terms<T>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            const T pc = ld(c);
            const T Gp = (T)d.x_gamb[k] * pc - xf[k];
            pnew[c] = pc;
            qv[c] = Gp;
            v[0] = (double)pc * Gp;
        }
    }
    block_reduce<1>(v);
    write_partials<1>(d, wi.tile, v);
}

// x += alpha p; r -= alpha q; z = D^-1 r; partials r.z, r.r -> beta, done
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_cg_upd(const ProbDesc *__restrict__ D,
                                                   const WorkItem *__restrict__ W, int xb,
                                                   int rb, int zb, int p0b, int qb, int done_slot,
                                                   int tol_slot, int max_iters, int mg) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, done_slot) || (done_slot != S_DONE && problem_skip(d, S_DONE))) return;
    const int par = (int)d.scal[S_PPAR];
    auto buf = [&](int b) -> T * { return sizeof(T) == 8 ? (T *)d.b64[b] : (T *)d.b32[b]; };
    T *x = buf(xb), *r = buf(rb), *z = buf(zb);   // mg: z comes from a V-cycle afterwards
    const T *p = buf(p0b + 1 - par), *qv = buf(qb);
    const T alpha = (T)d.scal[S_ALPHA];
    double v[2] = {0, 0};
    auto upd = [&](unsigned c, T xc, T pc, T rc, T qc, T dinv) {
        st_at(x, c, xc + alpha * pc);
        const T rr = rc - alpha * qc;
        st_at(r, c, rr);
        if (!mg) {
            const T zz = rr * dinv;
            st_at(z, c, zz);
            v[0] += (double)rr * zz;
        }
        v[1] += (double)rr * rr;
    };
    if (wi.tile < d.ntiles) {
        struct V {
            T x, p, r, q, di;
        };
        tile_loop<8>(
            d, wi.tile,
            [&](const Cell &q) {
                V a{};
                a.x = ld_at(x, q.c);
                a.p = ld_at(p, q.c);
                a.r = ld_at(r, q.c);
                a.q = ld_at(qv, q.c);
                if (HET) a.di = ld_at(Coef<T>::dinv(d), q.c);
                return a;
            },
            [&](const Cell &q, const V &a) { upd(q.c, a.x, a.p, a.r, a.q, HET ? a.di : cell_dinv<T, false>(d, q)); });
    } else {
        if (threadIdx.x < d.nx) {
            // diagonal of a package node: ring and coupling conductances + ambient
            const int k = threadIdx.x;
            const long c = (long)d.nl * d.nr * d.nc + k;
            upd(c, x[c], p[c], r[c], qv[c], T(1) / pkg_diag<T>(d, k));
        }
    }
    block_reduce<2>(v);
    write_partials<2>(d, wi.tile, v);
}

// Mixed precision: FP64 residual r = rhs - G x, its FP32 copy as the inner
// right-hand side, e = 0, inner CG initialised.  Outer convergence in FP64.
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_ir_resid(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, double rtol2,
                                                     double inner_rtol2, int max_refine) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, S_DONE) || d.scal[S_IDONE] == 0.0) return;   // converged / inner solve running
    const double *x = d.b64[B_STEADY], *rhs = d.b64[B_RHS];
    float *r32 = d.b32[F_R], *z32 = d.b32[F_Z], *e32 = d.b32[F_X];
    LdPlain<double> ld{x};
    double v[3] = {0, 0, 0};
    auto put = [&](unsigned c, double rr, double diag, double rhsc) {
        const float rf = (float)rr;
        const float zf = (float)(rr / diag);
        st_at(r32, c, rf);
        st_at(z32, c, zf);
        st_at(e32, c, 0.0f);
        v[0] += (double)rf * zf;
        v[1] += rr * rr;
        v[2] += rhsc * rhsc;
    };
    if (wi.tile < d.ntiles) {
        double ring[4];
        load_ring<double>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<double, HET> p;
            double b;
        };
        tile_loop<InFlight<double, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V a{};
                load_pts<double, HET>(d, ld, q, a.p);
                a.b = ld_at(rhs, q.c);
                return a;
            },
            [&](const Cell &q, const V &a) {
                double flux, diag, gam;
                stencil<double, HET>(d, q, a.p, ring, flux, diag, gam);
                put(q.c, a.b + flux - gam * a.p.c, diag, a.b);
            });
    } else {
        __shared__ double xf[kMaxExtra], xd[kMaxExtra];
        package_terms<double>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            put(c, rhs[c] + xf[k] - d.x_gamb[k] * x[c], xd[k], rhs[c]);
        }
    }
    block_reduce<3>(v);
    write_partials<3>(d, wi.tile, v);
}

// diagonal of G at a package node: ring and coupling conductances + ambient

template <typename T> __device__ __forceinline__ T *vec_of(const ProbDesc &d, int b) {
    return sizeof(T) == 8 ? (T *)d.b64[b] : (T *)d.b32[b];
}

// x (FP64) += e (FP32), for problems whose inner solve finished with a
// correction pending
__global__ void __launch_bounds__(kBlock) k_ir_correct(const ProbDesc *__restrict__ D,
                                                       const WorkItem *__restrict__ W) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, S_DONE) || d.scal[S_IDONE] == 0.0 || d.scal[S_CORR] == 0.0) return;
    double *x = d.b64[B_STEADY];
    const float *e = d.b32[F_X];
    if (wi.tile < d.ntiles) {
        struct V {
            double x;
            float e;
        };
        tile_loop<8>(
            d, wi.tile,
            [&](const Cell &q) {
                V a;
                a.x = ld_at(x, q.c);
                a.e = ld_at(e, q.c);
                return a;
            },
            [&](const Cell &q, const V &a) { st_at(x, q.c, a.x + (double)a.e); });
    } else if (threadIdx.x < d.nx) {
        const long c = (long)d.nl * d.nr * d.nc + threadIdx.x;
        x[c] += (double)e[c];
    }
}

// ---------------------------------------------------------------------------
// transient: the reference's adaptive RK4 (RCutil.c:440-640) with the step
// controller on the device
// ---------------------------------------------------------------------------
enum StageMode { ST_K = 0, ST_FIRST = 1, ST_MID = 2, ST_LAST = 3, ST_LAST_ERR = 4 };

// One slope evaluation k = (rhs - G s) / C at s = base + a*h*kin, fused with
// the RK4 combination selected by mode:
//   ST_K        kout = k
//   ST_FIRST    kout = k, acc = kin + 2k                (k2 stage)
//   ST_MID      kout = k, acc += 2k                     (k3 stage)
//   ST_LAST     out = base + hc*h*(acc + k)/6           (k4 stage)
//   ST_LAST_ERR as ST_LAST but only err = max|out - cmp| is kept
// Buffers are FP64 ids; base/out/cmp ids >= 100 mean "state y, parity-relative"
// (100 = current y, 101 = the other one).
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_rk_stage(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, int mode,
                                                     int base_b, int kin_b, double a_mul,
                                                     int kout_b, int out_b, double hc_mul,
                                                     int cmp_b) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    const double st = d.scal[S_STATUS];
    if (st == (double)RK_DONE) return;
    if (mode == ST_K && kin_b < 0 && st != (double)RK_NEED_K1 && base_b >= 100) return;   // k1 reuse
    const int ypar = (int)d.scal[S_YPAR];
    auto id = [&](int b) { return b >= 100 ? B_Y0 + ((b - 100) ^ ypar) : b; };
    const double h = d.scal[S_H];
    const double *base = d.b64[id(base_b)];
    const double *kin = kin_b >= 0 ? d.b64[kin_b] : base;
    const double a = kin_b >= 0 ? a_mul * h : 0.0;
    double *kout = kout_b >= 0 ? d.b64[kout_b] : nullptr;
    double *acc = d.b64[B_ACC];
    double *out = out_b >= 0 ? d.b64[id(out_b)] : nullptr;
    const double *cmp = cmp_b >= 0 ? d.b64[id(cmp_b)] : nullptr;
    const double hc = hc_mul * h;
    const double *rhs = d.b64[B_RHS];
    LdStage<double> ld{base, kin, a};
    double v[1] = {0};
    // k and the raw values the mode combines it with (kin, acc, base, cmp at c)
    auto fin = [&](unsigned c, double k, double kinc, double accc, double basec, double cmpc) {
        switch (mode) {
            case ST_K: st_at(kout, c, k); break;
            case ST_FIRST: st_at(kout, c, k); st_at(acc, c, kinc + 2 * k); break;
            case ST_MID: st_at(kout, c, k); st_at(acc, c, accc + 2 * k); break;
            case ST_LAST: st_at(out, c, basec + hc * (accc + k) / 6.0); break;
            default: {
                const double o = basec + hc * (accc + k) / 6.0;
                v[0] = fmax(v[0], fabs(cmpc - o));
            }
        }
    };
    const bool need_kin = mode == ST_FIRST, need_acc = mode >= ST_MID;
    const bool need_base = mode >= ST_LAST, need_cmp = mode == ST_LAST_ERR;
    if (wi.tile < d.ntiles) {
        double ring[4];
        load_ring<double>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<double, HET> p;
            double b, kin, acc, base, cmp, ci;
        };
        tile_loop<InFlight<double, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V t{};
                load_pts<double, HET>(d, ld, q, t.p);
                t.b = ld_at(rhs, q.c);
                if (need_kin) t.kin = ld_at(kin, q.c);
                if (need_acc) t.acc = ld_at(acc, q.c);
                if (need_base) t.base = ld_at(base, q.c);
                if (need_cmp) t.cmp = ld_at(cmp, q.c);
                if (HET) t.ci = ld_at(Coef<double>::cinv(d), q.c);
                return t;
            },
            [&](const Cell &q, const V &t) {
                double flux, diag, gam;
                stencil<double, HET>(d, q, t.p, ring, flux, diag, gam);
                const double ci = HET ? t.ci : d.cinv[q.n];
                fin(q.c, ci * (t.b + flux - gam * t.p.c), t.kin, t.acc, t.base, t.cmp);
            });
    } else {
        __shared__ double xf[kMaxExtra], xd[kMaxExtra];
        package_terms<double>(d, ld, true, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            fin(c, d.x_cinv[k] * (rhs[c] + xf[k] - d.x_gamb[k] * ld(c)), need_kin ? kin[c] : 0.0,
                need_acc ? acc[c] : 0.0, need_base ? base[c] : 0.0, need_cmp ? cmp[c] : 0.0);
        }
    }
    if (mode != ST_LAST_ERR) return;
    block_reduce<1, true>(v);
    write_partials<1>(d, wi.tile, v);
}

// Step-size control of one attempt (rk4, RCutil.c:598-621) and the interval
// loop of compute_temp_grid (temperature_grid.c:3676-3692); one workgroup
// per problem, which first combines the step-doubling error of the attempt.
__global__ void __launch_bounds__(kBlock) k_rk_control(const ProbDesc *__restrict__ D) {
    const ProbDesc &d = D[blockIdx.x];
    double *s = d.scal;
    if (s[S_STATUS] == (double)RK_DONE) return;
    double v[1];
    reduce_slots<1, true>(d, v);
    if (threadIdx.x != 0) return;
    const double h = s[S_H], mx = v[0];
    s[S_ERR] = mx;
    double nh;
    if (mx <= 0.01) {
        nh = 0.95 * h * pow(fabs(0.01 / mx), 0.2);
        if (nh > 5.0 * h) nh = 5.0 * h;
    } else {
        nh = 0.95 * h * pow(fabs(0.01 / mx), 0.25);
        if (nh < h / 10.0) nh = h / 10.0;
    }
    if (nh < h) {          // retry with the smaller step, k1 as overwritten
        s[S_H] = nh;
        s[S_STATUS] = RK_ATTEMPT;
        return;
    }
    // accept: y <- ytemp (the other parity buffer)
    s[S_YPAR] = 1.0 - s[S_YPAR];
    s[S_CALLS] += 1.0;
    const double t = s[S_T];
    nh = fmin(nh, s[S_DT] - t - h);
    s[S_T] = t + h;
    if (s[S_T] < s[S_DT] && nh >= kMinStep * kDelta) {
        s[S_H] = nh;
        s[S_STATUS] = RK_NEED_K1;
    } else {
        s[S_STATUS] = RK_DONE;
    }
}

// Finish kernels of the CG / refinement / Taylor reductions: one workgroup
// per problem, same skip condition as the producer kernel.
enum FinOp { FIN_CG_INIT = 0, FIN_ALPHA, FIN_BETA, FIN_IR, FIN_TAY, FIN_GM_IR, FIN_GM_H, FIN_GM_NORM,
             FIN_MG_RZ0, FIN_MG_RR, FIN_MG_BETA };
struct FinArgs {
    double rtol2, inner_rtol2;
    int max_iters, max_refine, done_slot, tol_slot, sub, seq, mg;
    int *kmax;     // FIN_TAY: [0] largest term count of the group's sub-step, [1] the same since the host cleared it
    double *trec;  // FIN_TAY: per problem {tau, terms (0: no sub-step)}
    int kg;        // FIN_TAY: the group's plan (launches of the sub-step)
    int fused;     // FIN_TAY: the last step adds its result into the state (k_tay_term4)
    int zslots;    // FIN_TAY: the first term z-marched (k_tay_first4z): t_ntiles / nl + 1 slots
};

// ---------------------------------------------------------------------------
// geometric multigrid (FP32 preconditioner of the inner solves).  Levels
// halve rows and columns and keep the layers and the package nodes.  Coarse
// operators aggregate the fine ones: vertical, ambient, ring and advection
// couplings sum over the fine faces of a coarse face; lateral conduction sums
// and halves (the coarse cell centres are twice as far apart), which is the
// rediscretised operator for homogeneous layers.  Internal advection of an
// aggregate moves into its self term.  Smoother: z-line Jacobi (a tridiagonal
// solve over the layers of each column: thin bond layers make the vertical
// coupling far stronger than the lateral one).  Transfer: restriction sums
// the four fine residuals, prolongation is piecewise constant (its transpose).
// ---------------------------------------------------------------------------

// fine-level face and advection coefficients (per layer for homogeneous
// levels, per cell otherwise); 0 where the face does not exist
struct FineCoef {
    const ProbDesc &d;
    bool het;
    __device__ unsigned idx(int n, int i, int j) const { return ((unsigned)n * d.nr + i) * d.nc + j; }
    __device__ float gS(int n, int i, int j) const {
        if (i >= d.nr - 1) return 0.f;
        return het ? d.fS[idx(n, i, j)] : (float)d.gx[n];
    }
    __device__ float gE(int n, int i, int j) const {
        if (j >= d.nc - 1) return 0.f;
        return het ? d.fE[idx(n, i, j)] : (float)d.gy[n];
    }
    __device__ float gD(int n, int i, int j) const {
        if (n >= d.nl - 1) return 0.f;
        return het ? d.fD[idx(n, i, j)] : (float)d.gz[n];
    }
    __device__ bool adv() const { return het && d.has_adv; }
    __device__ float aN(int n, int i, int j) const { return adv() ? d.faN[idx(n, i, j)] : 0.f; }
    __device__ float aS(int n, int i, int j) const { return adv() ? d.faS[idx(n, i, j)] : 0.f; }
    __device__ float aE(int n, int i, int j) const { return adv() ? d.faE[idx(n, i, j)] : 0.f; }
    __device__ float aW(int n, int i, int j) const { return adv() ? d.faW[idx(n, i, j)] : 0.f; }
    __device__ float self(int n, int i, int j) const { return adv() ? d.faSelf[idx(n, i, j)] : 0.f; }
};

// coarse faces from the fine level
__device__ __forceinline__ float cg_S(const FineCoef &f, int n, int I, int J, int NR) {
    return I < NR - 1 ? 0.5f * (f.gS(n, 2 * I + 1, 2 * J) + f.gS(n, 2 * I + 1, 2 * J + 1)) : 0.f;
}
__device__ __forceinline__ float cg_E(const FineCoef &f, int n, int I, int J, int NC) {
    return J < NC - 1 ? 0.5f * (f.gE(n, 2 * I, 2 * J + 1) + f.gE(n, 2 * I + 1, 2 * J + 1)) : 0.f;
}
__device__ __forceinline__ float cg_D(const FineCoef &f, int n, int I, int J) {
    return f.gD(n, 2 * I, 2 * J) + f.gD(n, 2 * I, 2 * J + 1) + f.gD(n, 2 * I + 1, 2 * J) +
           f.gD(n, 2 * I + 1, 2 * J + 1);
}

// coarse-level coefficients and Jacobi diagonal, one thread per coarse cell
__global__ void k_mg_coarsen(const ProbDesc *__restrict__ F, const ProbDesc *__restrict__ Cp, int fine_het) {
    const ProbDesc &d = *F;
    const ProbDesc &c = *Cp;
    const unsigned ncell = (unsigned)c.nl * c.nr * c.nc;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const unsigned plane = (unsigned)c.nr * c.nc;
    const int n = t / plane, rem = t - n * plane, I = rem / c.nc, J = rem - I * c.nc;
    const FineCoef f{d, fine_het != 0};
    const float gs = cg_S(f, n, I, J, c.nr), ge = cg_E(f, n, I, J, c.nc), gd = n < c.nl - 1 ? cg_D(f, n, I, J) : 0.f;
    const float gn = I > 0 ? cg_S(f, n, I - 1, J, c.nr) : 0.f, gw = J > 0 ? cg_E(f, n, I, J - 1, c.nc) : 0.f;
    const float gu = n > 0 ? cg_D(f, n - 1, I, J) : 0.f;
    float an = 0, as = 0, ae = 0, aw = 0, self = 0;
    if (f.adv()) {
        const int i0 = 2 * I, j0 = 2 * J;
        // Galerkin sums across the coarse faces, halved like the conduction;
        // the self term keeps the Galerkin row sum (internal advection of
        // the aggregate twice, plus the halved-away external part)
        const float sn = f.aN(n, i0, j0) + f.aN(n, i0, j0 + 1), ss = f.aS(n, i0 + 1, j0) + f.aS(n, i0 + 1, j0 + 1);
        const float se = f.aE(n, i0, j0 + 1) + f.aE(n, i0 + 1, j0 + 1), sw = f.aW(n, i0, j0) + f.aW(n, i0 + 1, j0);
        an = 0.5f * sn;
        as = 0.5f * ss;
        ae = 0.5f * se;
        aw = 0.5f * sw;
        self = f.self(n, i0, j0) + f.self(n, i0, j0 + 1) + f.self(n, i0 + 1, j0) + f.self(n, i0 + 1, j0 + 1) +
               2.f * (f.aS(n, i0, j0) + f.aE(n, i0, j0) + f.aS(n, i0, j0 + 1) + f.aW(n, i0, j0 + 1) +
                      f.aN(n, i0 + 1, j0) + f.aE(n, i0 + 1, j0) + f.aN(n, i0 + 1, j0 + 1) + f.aW(n, i0 + 1, j0 + 1)) +
               (sn + ss + se + sw);
        ((float *)c.faN)[t] = an;
        ((float *)c.faS)[t] = as;
        ((float *)c.faE)[t] = ae;
        ((float *)c.faW)[t] = aw;
        ((float *)c.faSelf)[t] = self;
    }
    ((float *)c.fS)[t] = gs;
    ((float *)c.fE)[t] = ge;
    ((float *)c.fD)[t] = gd;
    float diag = gn + gs + ge + gw + gu + gd - (an + as + ae + aw + self) + (float)c.gamb[n];
    const int *rn = c.ring_node[n];
    if (I == 0 && rn[0] >= 0) diag += (float)c.ring_gns[n];
    if (I == c.nr - 1 && rn[1] >= 0) diag += (float)c.ring_gns[n];
    if (J == c.nc - 1 && rn[2] >= 0) diag += (float)c.ring_gew[n];
    if (J == 0 && rn[3] >= 0) diag += (float)c.ring_gew[n];
    ((float *)c.fdinv)[t] = 1.f / diag;
}

// Upwind form of the coolant advection, for the multigrid preconditioner of
// the non-symmetric solves.  The operator advects with central averages,
// Cv Q (T_nbr + T) / 2 on every fluid face (temperature_grid.c:3256-3590),
// which leaves a fluid row without diagonal dominance at high cell Peclet
// numbers, where relaxation stalls.  The V-cycle works on the upwind operator
// instead -- heat enters at the upstream cell's temperature and leaves at the
// cell's own, an M-matrix -- and GMRES corrects the difference.  In the
// stencil's form sum a (T_nbr + T) + self T: a' = 2a on inflow faces, 0 on
// outflow faces, self' = self - 2 sum |a|.
__global__ void k_upwind_coef(const ProbDesc *__restrict__ D, long ncell, float *uN, float *uS, float *uE,
                              float *uW, float *uSelf, float *udinv) {
    const ProbDesc &d = *D;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const double a[4] = {d.aN[c], d.aS[c], d.aE[c], d.aW[c]};
    double u[4], sa = 0.0, su = 0.0, sabs = 0.0;
    for (int k = 0; k < 4; k++) {
        u[k] = a[k] > 0.0 ? 2.0 * a[k] : 0.0;
        sa += a[k];
        su += u[k];
        sabs += fabs(a[k]);
    }
    const double self = d.aSelf[c], uself = self - 2.0 * sabs;
    uN[c] = (float)u[0];
    uS[c] = (float)u[1];
    uE[c] = (float)u[2];
    uW[c] = (float)u[3];
    uSelf[c] = (float)uself;
    udinv[c] = (float)(1.0 / (1.0 / d.dinvc[c] + (sa + self) - (su + uself)));
}

// Backward-Euler levels of the multigrid preconditioner of a non-symmetric
// implicit step: the step's diagonal shift sigma C per cell (k_be_coef) goes
// into each level's ambient term, a coarse cell's shift being the sum of its
// four children's, and into the level's Jacobi diagonal (base_dinv: the
// level's own, unshifted).  One thread per cell of level C; F is the level
// above (nullptr: C is the finest level and takes fine_shift itself).
__global__ void k_mg_shift(const ProbDesc *__restrict__ F, const ProbDesc *__restrict__ Cp,
                           const double *__restrict__ fine_shift, const float *__restrict__ base_dinv) {
    const ProbDesc &c = *Cp;
    const unsigned ncell = (unsigned)c.nl * c.nr * c.nc;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    double s;
    if (!F) {
        s = fine_shift[t];
    } else {
        const ProbDesc &f = *F;
        const unsigned plane = (unsigned)c.nr * c.nc;
        const unsigned n = t / plane, rem = t - n * plane, I = rem / c.nc, J = rem - I * c.nc;
        const unsigned f0 = (n * f.nr + 2 * I) * f.nc + 2 * J;
        s = (f.bshift[f0] + f.bshift[f0 + 1]) + (f.bshift[f0 + f.nc] + f.bshift[f0 + f.nc + 1]);
    }
    ((double *)c.bshift)[t] = s;
    ((float *)c.fdinv)[t] = 1.f / (1.f / base_dinv[t] + (float)s);
}

// when a multigrid kernel runs for a problem (flags of the fine problem,
// whose scalars every level shares)
enum MgWhen { MG_RUNNING = 0, MG_INIT = 1, MG_CORRECT = 2 };
__device__ __forceinline__ bool mg_skip(const ProbDesc &d, int when) {
    const double *s = d.scal;
    if (s[S_DONE] != 0.0) return true;
    if (when == MG_INIT) return s[S_MGINIT] == 0.0;
    if (when == MG_CORRECT) return s[S_IDONE] == 0.0 || s[S_CORR] == 0.0;
    return s[S_IDONE] != 0.0;
}

// column of a 2D tile: TX = min(nc, 64) columns x 256/TX rows, all layers
struct Col {
    int i, j;
    bool active;
};
__device__ __forceinline__ Col col_of(const ProbDesc &d, int tile) {
    const int sh = d.tx_shift;
    const int ti = tile / d.ntj, tj = tile - ti * d.ntj;
    const int i = ti * (kBlock >> sh) + (threadIdx.x >> sh), j = (tj << sh) + (threadIdx.x & ((1 << sh) - 1));
    return Col{min(i, d.nr - 1), min(j, d.nc - 1), i < d.nr && j < d.nc};
}

// state loader of a smoothing sweep: x (0 when zero), plus the coarse
// correction of the parent cell when prolonging
struct LdMg {
    const float *x;          // nullptr: zero
    const float *xc;         // coarse solution (prolongation) or nullptr
    unsigned ncell, cncell;  // node counts of this and the coarse level
    int nr, nc, cnr, cnc;
    __device__ __forceinline__ float operator()(unsigned c) const {
        float v = x ? ld_at(x, c) : 0.f;
        if (xc) {
            unsigned p;
            if (c >= ncell) {
                p = cncell + (c - ncell);
            } else {
                const unsigned plane = (unsigned)nr * nc;
                const unsigned n = c / plane, rem = c - n * plane, i = rem / nc, j = rem - i * nc;
                p = (n * cnr + (i >> 1)) * cnc + (j >> 1);
            }
            v += ld_at(xc, p);
        }
        return v;
    }
};

// One z-line Jacobi sweep x_out = x + omega T^-1 (b - A x) (T: the
// tridiagonal vertical part of A), x as given by the loader.  Package nodes:
// point Jacobi.  With dot >= 0, the partial (b, x_out) goes to the problem's
// reduction slot (the last sweep of a V-cycle at the finest level).
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_mg_smooth(const ProbDesc *__restrict__ D,
                                                      const ProbDesc *__restrict__ DC,
                                                      const WorkItem *__restrict__ W, int xin, int xcb,
                                                      int xout, int bb, float omega, int when, int dot) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (mg_skip(d, when)) return;
    const unsigned ncell = (unsigned)d.nl * d.nr * d.nc;
    LdMg ld{xin >= 0 ? d.b32[xin] : nullptr, nullptr, ncell, 0, d.nr, d.nc, 0, 0};
    if (xcb >= 0) {
        const ProbDesc &c = DC[wi.prob];
        ld.xc = c.b32[xcb];
        ld.cncell = (unsigned)c.nl * c.nr * c.nc;
        ld.cnr = c.nr;
        ld.cnc = c.nc;
    }
    const float *b = d.b32[bb];
    float *xo = d.b32[xout];
    double acc[1] = {0.0};
    const int ntiles2 = d.nti2 * d.ntj;
    if (wi.tile < ntiles2) {
        extern __shared__ float lds[];   // [2][nl][kBlock]: Thomas coefficients
        float *cp = lds, *dp = lds + d.nl * kBlock;
        const Col col = col_of(d, wi.tile);
        const unsigned plane = (unsigned)d.nr * d.nc;
        const unsigned c0 = (unsigned)col.i * d.nc + col.j;
        float ring[4];
        float lower = 0.f;
        for (int n = 0; n < d.nl; n++) {
            load_ring<float>(d, ld, n, ring);
            Cell q{n, col.i, col.j, n * plane + c0, col.active};
            Pts<float, HET> p{};
            load_pts<float, HET>(d, ld, q, p);
            float flux, diag, gam;
            stencil<float, HET>(d, q, p, ring, flux, diag, gam);
            const float res = ld_at(b, q.c) - (gam * p.c - flux);
            // vertical couplings of the tridiagonal part
            const float gd = n < d.nl - 1 ? (HET ? ld_at(Coef<float>::D(d), q.c) : (float)d.gz[n]) : 0.f;
            const float up = -gd;   // coefficient of x(n+1) in row n
            const float m = n == 0 ? diag : diag - lower * cp[(n - 1) * kBlock + threadIdx.x];
            cp[n * kBlock + threadIdx.x] = up / m;
            dp[n * kBlock + threadIdx.x] = (res - (n == 0 ? 0.f : lower * dp[(n - 1) * kBlock + threadIdx.x])) / m;
            lower = -gd;            // coefficient of x(n) in row n+1
        }
        float dx = 0.f;
        for (int n = d.nl - 1; n >= 0; n--) {
            dx = dp[n * kBlock + threadIdx.x] - (n == d.nl - 1 ? 0.f : cp[n * kBlock + threadIdx.x] * dx);
            const unsigned c = n * plane + c0;
            const float xn = ld(c) + omega * dx;
            if (col.active) {
                st_at(xo, c, xn);
                if (dot) acc[0] += (double)ld_at(b, c) * xn;
            }
        }
    } else {
        __shared__ float xf[kMaxExtra], xd[kMaxExtra];
        package_terms<float>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const unsigned c = ncell + k;
            const float xc = ld(c);
            const float res = b[c] - ((float)d.x_gamb[k] * xc - xf[k]);
            const float xn = xc + omega * res / xd[k];
            xo[c] = xn;
            if (dot) acc[0] += (double)b[c] * xn;
        }
    }
    if (dot < 0) return;
    block_reduce<1>(acc);
    write_partials<1>(d, wi.tile, acc);   // column tiles, then the package item at nti2 * ntj
}

// One x-line Jacobi sweep x_out = x + omega T^-1 (b - A x), T the tridiagonal
// part of A along the columns of a row segment of TX = min(nc, 64) cells
// (couplings across segment ends dropped), for every layer of the thread's
// column.  The segments are solved by parallel cyclic reduction across the
// lanes of the wavefront (TX divides 64, so a wavefront holds whole
// segments).  Package nodes: point Jacobi.  For advection-dominated layers,
// where the coolant carries heat along the rows.
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_mg_xline(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, int xin, int xout, int bb,
                                                     float omega, int when) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (mg_skip(d, when)) return;
    const unsigned ncell = (unsigned)d.nl * d.nr * d.nc;
    LdMg ld{xin >= 0 ? d.b32[xin] : nullptr, nullptr, ncell, 0, d.nr, d.nc, 0, 0};
    const float *b = d.b32[bb];
    float *xo = d.b32[xout];
    const int ntiles2 = d.nti2 * d.ntj;
    if (wi.tile < ntiles2) {
        const Col col = col_of(d, wi.tile);
        const int seg = 1 << d.tx_shift, ls = threadIdx.x & (seg - 1);
        const unsigned plane = (unsigned)d.nr * d.nc;
        float ring[4];
        for (int n = 0; n < d.nl; n++) {
            load_ring<float>(d, ld, n, ring);
            Cell q{n, col.i, col.j, n * plane + (unsigned)col.i * d.nc + col.j, col.active};
            Pts<float, HET> p{};
            load_pts<float, HET>(d, ld, q, p);
            float flux, diag, gam;
            stencil<float, HET>(d, q, p, ring, flux, diag, gam);
            float rhs = ld_at(b, q.c) - (gam * p.c - flux);
            // row of T: coefficients of x(j-1), x(j), x(j+1) in (A x)_c
            float lo, up;
            if (HET) {
                const float gw = col.j > 0 ? p.gw : 0.f, ge = col.j < d.nc - 1 ? p.ge : 0.f;
                lo = -gw - ((d.has_adv && col.j > 0) ? p.aw : 0.f);
                up = -ge - ((d.has_adv && col.j < d.nc - 1) ? p.ae : 0.f);
            } else {
                lo = col.j > 0 ? -(float)d.gy[n] : 0.f;
                up = col.j < d.nc - 1 ? -(float)d.gy[n] : 0.f;
            }
            if (ls == 0) lo = 0.f;
            if (ls == seg - 1) up = 0.f;
            float bd = diag;
            if (!col.active) {   // an idle lane solves x = 0
                lo = up = rhs = 0.f;
                bd = 1.f;
            }
            // parallel cyclic reduction within the segment
            for (int s = 1; s < seg; s <<= 1) {
                const float lo_m = __shfl_up(lo, s, kWave), bd_m = __shfl_up(bd, s, kWave);
                const float up_m = __shfl_up(up, s, kWave), r_m = __shfl_up(rhs, s, kWave);
                const float lo_p = __shfl_down(lo, s, kWave), bd_p = __shfl_down(bd, s, kWave);
                const float up_p = __shfl_down(up, s, kWave), r_p = __shfl_down(rhs, s, kWave);
                const bool hm = ls - s >= 0, hp = ls + s < seg;
                const float al = hm ? -lo / bd_m : 0.f, ga = hp ? -up / bd_p : 0.f;
                const float nlo = hm ? al * lo_m : 0.f, nup = hp ? ga * up_p : 0.f;
                bd = bd + (hm ? al * up_m : 0.f) + (hp ? ga * lo_p : 0.f);
                rhs = rhs + (hm ? al * r_m : 0.f) + (hp ? ga * r_p : 0.f);
                lo = nlo;
                up = nup;
            }
            if (col.active) st_at(xo, q.c, p.c + omega * (rhs / bd));
        }
    } else {
        __shared__ float xf[kMaxExtra], xd[kMaxExtra];
        package_terms<float>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const unsigned c = ncell + k;
            const float xc = ld(c);
            const float res = b[c] - ((float)d.x_gamb[k] * xc - xf[k]);
            xo[c] = xc + omega * res / xd[k];
        }
    }
}

// coarse right-hand side: sum of the four fine residuals b - A x
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_mg_restrict(const ProbDesc *__restrict__ DF,
                                                        const ProbDesc *__restrict__ DC,
                                                        const WorkItem *__restrict__ W, int xf_b, int bf_b,
                                                        int bc_b, int when) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &f = DF[wi.prob];
    const ProbDesc &c = DC[wi.prob];
    if (mg_skip(f, when)) return;
    LdPlain<float> ld{f.b32[xf_b]};
    const float *bf = f.b32[bf_b];
    float *bc = c.b32[bc_b];
    const unsigned fncell = (unsigned)f.nl * f.nr * f.nc, cncell = (unsigned)c.nl * c.nr * c.nc;
    const int ntiles2 = c.nti2 * c.ntj;
    if (wi.tile < ntiles2) {
        const Col col = col_of(c, wi.tile);
        const unsigned fplane = (unsigned)f.nr * f.nc, cplane = (unsigned)c.nr * c.nc;
        float ring[4];
        for (int n = 0; n < c.nl; n++) {
            load_ring<float>(f, ld, n, ring);
            float s = 0.f;
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int i = 2 * col.i + a, j = 2 * col.j + e;
                    Cell q{n, i, j, n * fplane + (unsigned)i * f.nc + j, true};
                    Pts<float, HET> p{};
                    load_pts<float, HET>(f, ld, q, p);
                    float flux, diag, gam;
                    stencil<float, HET>(f, q, p, ring, flux, diag, gam);
                    s += ld_at(bf, q.c) - (gam * p.c - flux);
                }
            if (col.active) st_at(bc, n * cplane + (unsigned)col.i * c.nc + col.j, s);
        }
    } else {
        __shared__ float xf[kMaxExtra], xd[kMaxExtra];
        package_terms<float>(f, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < f.nx) bc[cncell + k] = bf[fncell + k] - ((float)f.x_gamb[k] * ld(fncell + k) - xf[k]);
    }
}

// ---------------------------------------------------------------------------
// steady, non-symmetric G (coolant advection): restarted GMRES(kGmM), right-
// preconditioned with the Jacobi diagonal, as the inner solve of FP64
// iterative refinement (one GMRES cycle from e = 0 per refinement, on the
// FP64 residual).  The basis is kept unnormalised: v_i = scale_i * vhat_i.
// Per iteration j: w = G z_j / scale_j (z_j = diag^-1 v_j); h_ij = (w, v_i) /
// scale_i; v_j+1 = w - sum_i h_ij v_i / scale_i, z_j+1 = diag^-1 v_j+1,
// h_j+1,j = ||v_j+1||; Givens rotations and the least-squares solve run in
// the problem's finish workgroup.  Per-problem state (H, rotations, g, y,
// scales, j) lives in ProbDesc::gm.
// ---------------------------------------------------------------------------
enum GmState {
    GM_J = 0, GM_K, GM_TOL,
    GM_SCALE = 4,                        // [kGmM + 1]
    GM_G = GM_SCALE + kGmM + 1,          // [kGmM + 1]
    GM_CS = GM_G + kGmM + 1,             // [kGmM]
    GM_SN = GM_CS + kGmM,                // [kGmM]
    GM_Y = GM_SN + kGmM,                 // [kGmM]
    GM_H = GM_Y + kGmM,                  // column-major [(kGmM + 1) x kGmM]
    GM_SIZE = GM_H + (kGmM + 1) * kGmM
};

template <typename T> __device__ __forceinline__ T *gm_vec(const ProbDesc &d, int i) {
    const size_t off = (size_t)i * d.vlen;
    return sizeof(T) == 8 ? (T *)(d.gmv64 + off) : (T *)(d.gmv32 + off);
}

__device__ __forceinline__ bool gm_skip(const ProbDesc &d) {
    return problem_skip(d, S_DONE) || d.scal[S_IDONE] != 0.0;
}

template <typename T, bool HET> __device__ __forceinline__ T node_dinv(const ProbDesc &d, const Cell &q) {
    return HET ? ld_at(Coef<T>::dinv(d), q.c) : cell_dinv<T, false>(d, q);
}

// FP64 residual r = rhs - G x -> v_0 = r, z = diag^-1 r (T); partials
// (T(r), T(r)), (r, r), (rhs, rhs).
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_gm_resid(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, int zb, int mg) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, S_DONE) || d.scal[S_IDONE] == 0.0) return;   // converged / cycle running
    const double *x = d.b64[B_STEADY], *rhs = d.b64[B_RHS];
    T *v0 = gm_vec<T>(d, 0), *z = vec_of<T>(d, zb);
    LdPlain<double> ld{x};
    double w[3] = {0, 0, 0};
    auto put = [&](unsigned c, double rr, T dinv, double bc) {
        const T rt = (T)rr;
        st_at(v0, c, rt);
        st_at(z, c, mg ? rt : rt * dinv);   // mg: z is the multigrid input (v_0)
        w[0] += (double)rt * rt;
        w[1] += rr * rr;
        w[2] += bc * bc;
    };
    if (wi.tile < d.ntiles) {
        double ring[4];
        load_ring<double>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<double, HET> p;
            double b;
            T di;
        };
        tile_loop<InFlight<double, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V a{};
                load_pts<double, HET>(d, ld, q, a.p);
                a.b = ld_at(rhs, q.c);
                a.di = node_dinv<T, HET>(d, q);
                return a;
            },
            [&](const Cell &q, const V &a) {
                double flux, diag, gam;
                stencil<double, HET>(d, q, a.p, ring, flux, diag, gam);
                put(q.c, a.b + flux - gam * a.p.c, a.di, a.b);
            });
    } else {
        __shared__ double xf[kMaxExtra], xd[kMaxExtra];
        package_terms<double>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const unsigned c = (unsigned)d.nl * d.nr * d.nc + k;
            put(c, rhs[c] + xf[k] - d.x_gamb[k] * x[c], T(1) / pkg_diag<T>(d, k), rhs[c]);
        }
    }
    block_reduce<3>(w);
    write_partials<3>(d, wi.tile, w);
}

// w = G z / scale_j
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_gm_apply(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, int zb, int wb) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (gm_skip(d)) return;
    const T *z = vec_of<T>(d, zb);
    T *out = vec_of<T>(d, wb);
    const T inv = (T)(1.0 / d.gm[GM_SCALE + (int)d.gm[GM_J]]);
    LdPlain<T> ld{z};
    if (wi.tile < d.ntiles) {
        T ring[4];
        load_ring<T>(d, ld, tile_layer(d, wi.tile), ring);
        tile_loop<InFlight<T, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                Pts<T, HET> p{};
                load_pts<T, HET>(d, ld, q, p);
                return p;
            },
            [&](const Cell &q, const Pts<T, HET> &p) {
                T flux, diag, gam;
                stencil<T, HET>(d, q, p, ring, flux, diag, gam);
                st_at(out, q.c, (gam * p.c - flux) * inv);
            });
    } else {
        __shared__ T xf[kMaxExtra], xd[kMaxExtra];
        package_terms<T>(d, ld, false, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const unsigned c = (unsigned)d.nl * d.nr * d.nc + k;
            out[c] = ((T)d.x_gamb[k] * z[c] - xf[k]) * inv;
        }
    }
}

// partials (w, v_i), i = 0..j, into the problem's GMRES slots
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gm_dots(const ProbDesc *__restrict__ D,
                                                    const WorkItem *__restrict__ W, int wb) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (gm_skip(d)) return;
    const int j = (int)d.gm[GM_J];
    const T *w = vec_of<T>(d, wb);
    T wr[8];
    unsigned cr[8];
    bool ok[8];
    const bool pkg = wi.tile >= d.ntiles;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        if (pkg) {
            ok[r] = r == 0 && threadIdx.x < d.nx;
            cr[r] = (unsigned)d.nl * d.nr * d.nc + (ok[r] ? threadIdx.x : 0);
        } else {
            const Cell q = tile_cell(d, wi.tile, r);
            ok[r] = q.active;
            cr[r] = q.c;
        }
        wr[r] = ok[r] ? ld_at(w, cr[r]) : T(0);
    }
    __shared__ double red[kBlock / kWave][kGmM + 1];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (int i = 0; i <= j; i++) {
        const T *vi = gm_vec<T>(d, i);
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++)
            if (ok[r]) a += (double)wr[r] * ld_at(vi, cr[r]);
        a = wave_sum(a);
        if (lane == 0) red[wv][i] = a;
    }
    __syncthreads();
    if (threadIdx.x <= j) {
        double s = 0.0;
        for (int q = 0; q < kBlock / kWave; q++) s += red[q][threadIdx.x];
        d.gpart[(size_t)wi.tile * (kGmM + 1) + threadIdx.x] = s;
    }
}

// v_j+1 = w - sum_i h_ij v_i / scale_i, z = diag^-1 v_j+1; partial ||v_j+1||^2
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_gm_orth(const ProbDesc *__restrict__ D,
                                                    const WorkItem *__restrict__ W, int wb, int zb, int mg) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (gm_skip(d)) return;
    const int j = (int)d.gm[GM_J];
    const T *w = vec_of<T>(d, wb);
    T *vn = gm_vec<T>(d, j + 1), *z = vec_of<T>(d, zb);
    double acc[1] = {0.0};
    auto put = [&](unsigned c, T dinv) {
        T x = ld_at(w, c);
        for (int i = 0; i <= j; i++) {
            const T hc = (T)(d.gm[GM_H + j * (kGmM + 1) + i] / d.gm[GM_SCALE + i]);
            x -= hc * ld_at(gm_vec<T>(d, i), c);
        }
        st_at(vn, c, x);
        st_at(z, c, mg ? x : x * dinv);   // mg: z is the multigrid input (v_j+1)
        acc[0] += (double)x * x;
    };
    if (wi.tile < d.ntiles) {
        tile_loop<1>(
            d, wi.tile, [&](const Cell &q) { return node_dinv<T, HET>(d, q); },
            [&](const Cell &q, T dinv) { put(q.c, dinv); });
    } else if (threadIdx.x < d.nx) {
        const int k = threadIdx.x;
        put((unsigned)d.nl * d.nr * d.nc + k, T(1) / pkg_diag<T>(d, k));
    }
    block_reduce<1>(acc);
    if (threadIdx.x == 0) d.gpart[(size_t)wi.tile * (kGmM + 1)] = acc[0];
}

// x (FP64) += diag^-1 sum_i y_i v_i / scale_i, for problems whose cycle
// finished with a correction pending
template <typename T, bool HET>
__global__ void __launch_bounds__(kBlock) k_gm_correct(const ProbDesc *__restrict__ D,
                                                       const WorkItem *__restrict__ W, int mgout) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (problem_skip(d, S_DONE) || d.scal[S_IDONE] == 0.0 || d.scal[S_CORR] == 0.0) return;
    const int k = (int)d.gm[GM_K];
    double *x = d.b64[B_STEADY];
    T *u = mgout >= 0 ? vec_of<T>(d, mgout) : nullptr;
    auto put = [&](unsigned c, T dinv) {
        T e = T(0);
        for (int i = 0; i < k; i++)
            e += (T)(d.gm[GM_Y + i] / d.gm[GM_SCALE + i]) * ld_at(gm_vec<T>(d, i), c);
        if (u) st_at(u, c, e);   // multigrid: the V-cycle of V y follows
        else st_at(x, c, ld_at(x, c) + (double)(e * dinv));
    };
    if (wi.tile < d.ntiles) {
        tile_loop<1>(
            d, wi.tile, [&](const Cell &q) { return node_dinv<T, HET>(d, q); },
            [&](const Cell &q, T dinv) { put(q.c, dinv); });
    } else if (threadIdx.x < d.nx) {
        const int kk = threadIdx.x;
        put((unsigned)d.nl * d.nr * d.nc + kk, T(1) / pkg_diag<T>(d, kk));
    }
}

// per-problem GMRES bookkeeping (one thread): start of a cycle, the
// Hessenberg column of an iteration, its rotation and the stopping test
__device__ void gm_finish_resid(const ProbDesc &d, const double (&v)[3], const FinArgs &a) {
    double *s = d.scal, *g = d.gm;
    s[S_RR] = v[1];
    s[S_BB] = v[2];
    s[S_TOL2] = a.rtol2 * v[2];
    s[S_TITERS] += s[S_ITERS];
    s[S_ITERS] = 0.0;
    const double nv = sqrt(v[0]);
    const bool done = v[1] <= s[S_TOL2] || s[S_NREFINE] >= a.max_refine || !(v[1] == v[1]) || nv == 0.0;
    s[S_DONE] = done ? 1.0 : 0.0;
    s[S_IDONE] = done ? 1.0 : 0.0;
    s[S_CORR] = done ? 0.0 : 1.0;
    s[S_NREFINE] += 1.0;
    s[S_MGINIT] = (a.mg && !done) ? 1.0 : 0.0;
    g[GM_J] = 0.0;
    g[GM_K] = 0.0;
    g[GM_TOL] = sqrt(a.inner_rtol2) * nv;
    g[GM_SCALE] = nv > 0.0 ? nv : 1.0;
    for (int i = 0; i <= kGmM; i++) g[GM_G + i] = 0.0;
    g[GM_G] = nv;
}

__device__ void gm_finish_norm(const ProbDesc &d, double hn) {
    double *s = d.scal, *g = d.gm;
    const int j = (int)g[GM_J];
    double *h = g + GM_H + j * (kGmM + 1);
    h[j + 1] = hn;
    g[GM_SCALE + j + 1] = hn > 0.0 ? hn : 1.0;
    for (int i = 0; i < j; i++) {
        const double c = g[GM_CS + i], sn = g[GM_SN + i];
        const double a0 = h[i], a1 = h[i + 1];
        h[i] = c * a0 + sn * a1;
        h[i + 1] = -sn * a0 + c * a1;
    }
    const double rr = hypot(h[j], h[j + 1]);
    const double c = rr > 0.0 ? h[j] / rr : 1.0, sn = rr > 0.0 ? h[j + 1] / rr : 0.0;
    g[GM_CS + j] = c;
    g[GM_SN + j] = sn;
    h[j] = rr;
    h[j + 1] = 0.0;
    g[GM_G + j + 1] = -sn * g[GM_G + j];
    g[GM_G + j] = c * g[GM_G + j];
    s[S_ITERS] += 1.0;
    const int k = j + 1;
    g[GM_J] = k;
    if (fabs(g[GM_G + k]) <= g[GM_TOL] || k == kGmM || hn == 0.0 || !(hn == hn)) {
        // least squares: R y = g (upper triangular, k x k)
        for (int i = k - 1; i >= 0; i--) {
            double t = g[GM_G + i];
            for (int l = i + 1; l < k; l++) t -= g[GM_H + l * (kGmM + 1) + i] * g[GM_Y + l];
            const double dii = g[GM_H + i * (kGmM + 1) + i];
            g[GM_Y + i] = dii != 0.0 ? t / dii : 0.0;
        }
        g[GM_K] = k;
        s[S_IDONE] = 1.0;
    }
}

// Whole workgroup: sum of one GMRES partial over a problem's slots, in slot order
__device__ __forceinline__ void reduce_gslots(const ProbDesc &d, int i, double (&v)[1]) {
    const int nslots = d.ntiles + 1;
    v[0] = 0.0;
    for (int t = threadIdx.x; t < nslots; t += kBlock) v[0] += d.gpart[(size_t)t * (kGmM + 1) + i];
    block_reduce<1>(v);
}

__global__ void __launch_bounds__(kBlock) k_finish(const ProbDesc *__restrict__ D, int op, FinArgs a) {
    const ProbDesc &d = D[blockIdx.x];
    double *s = d.scal;
    double v[3];
    switch (op) {
        case FIN_CG_INIT: {
            if (s[S_DONE] != 0.0) return;
            reduce_slots<3>(d, v);
            if (threadIdx.x) return;
            s[S_RZ] = v[0];
            s[S_RR] = v[1];
            s[S_BB] = v[2];
            s[S_TOL2] = a.rtol2 * v[2];
            s[S_BETA] = 0.0;
            s[S_ITERS] = 0.0;
            s[S_PPAR] = 0.0;
            s[S_DONE] = (v[1] <= s[S_TOL2] || a.max_iters <= 0) ? 1.0 : 0.0;
            return;
        }
        case FIN_ALPHA: {
            if (s[a.done_slot] != 0.0 || s[S_DONE] != 0.0) return;
            double w[1];
            reduce_slots<1>(d, w);
            if (threadIdx.x) return;
            s[S_PQ] = w[0];
            s[S_ALPHA] = w[0] != 0.0 ? s[S_RZ] / w[0] : 0.0;
            return;
        }
        case FIN_BETA: {
            if (s[a.done_slot] != 0.0 || s[S_DONE] != 0.0) return;
            double w[2];
            reduce_slots<2>(d, w);
            if (threadIdx.x) return;
            s[S_BETA] = s[S_RZ] != 0.0 ? w[0] / s[S_RZ] : 0.0;
            s[S_RZ] = w[0];
            s[S_RR] = w[1];
            s[S_PPAR] = 1.0 - s[S_PPAR];
            s[S_ITERS] += 1.0;
            if (w[1] <= s[a.tol_slot] || s[S_ITERS] >= a.max_iters || !(w[1] == w[1])) s[a.done_slot] = 1.0;
            return;
        }
        case FIN_IR: {
            if (s[S_DONE] != 0.0 || s[S_IDONE] == 0.0) return;
            reduce_slots<3>(d, v);
            if (threadIdx.x) return;
            s[S_RR] = v[1];
            s[S_BB] = v[2];
            s[S_TOL2] = a.rtol2 * v[2];
            s[S_RZ] = v[0];
            s[S_RR0] = v[1];
            s[S_ITOL2] = a.inner_rtol2 * v[1];
            s[S_BETA] = 0.0;
            s[S_TITERS] += s[S_ITERS];
            s[S_ITERS] = 0.0;
            s[S_PPAR] = 0.0;
            const bool done = v[1] <= s[S_TOL2] || s[S_NREFINE] >= a.max_refine || !(v[1] == v[1]);
            s[S_DONE] = done ? 1.0 : 0.0;
            s[S_IDONE] = done ? 1.0 : 0.0;
            s[S_CORR] = done ? 0.0 : 1.0;
            s[S_NREFINE] += 1.0;
            s[S_MGINIT] = (a.mg && !done) ? 1.0 : 0.0;
            return;
        }
        case FIN_MG_RZ0: {   // first V-cycle of an inner solve done: z = M r
            if (s[S_DONE] != 0.0 || s[S_MGINIT] == 0.0) return;
            double w[1];
            reduce_slots<1>(d, w, d.nti2 * d.ntj + 1);
            if (threadIdx.x) return;
            s[S_RZ] = w[0];
            s[S_BETA] = 0.0;
            s[S_PPAR] = 0.0;
            s[S_MGINIT] = 0.0;
            return;
        }
        case FIN_MG_RR: {    // after the update of x and r: stop test
            if (s[a.done_slot] != 0.0 || s[S_DONE] != 0.0) return;
            double w[2];
            reduce_slots<2>(d, w);
            if (threadIdx.x) return;
            s[S_RR] = w[1];
            s[S_ITERS] += 1.0;
            if (w[1] <= s[a.tol_slot] || s[S_ITERS] >= a.max_iters || !(w[1] == w[1])) s[a.done_slot] = 1.0;
            return;
        }
        case FIN_MG_BETA: {  // after z = M r: beta and the next direction's parity
            if (s[a.done_slot] != 0.0 || s[S_DONE] != 0.0) return;
            double w[1];
            reduce_slots<1>(d, w, d.nti2 * d.ntj + 1);
            if (threadIdx.x) return;
            s[S_BETA] = s[S_RZ] != 0.0 ? w[0] / s[S_RZ] : 0.0;
            s[S_RZ] = w[0];
            s[S_PPAR] = 1.0 - s[S_PPAR];
            return;
        }
        case FIN_GM_IR: {
            if (s[S_DONE] != 0.0 || s[S_IDONE] == 0.0) return;
            reduce_slots<3>(d, v);
            if (threadIdx.x) return;
            gm_finish_resid(d, v, a);
            return;
        }
        case FIN_GM_H: {
            if (gm_skip(d)) return;
            const int j = (int)d.gm[GM_J];
            for (int i = 0; i <= j; i++) {
                double w[1];
                reduce_gslots(d, i, w);
                if (threadIdx.x == 0) d.gm[GM_H + j * (kGmM + 1) + i] = w[0] / d.gm[GM_SCALE + i];
            }
            return;
        }
        case FIN_GM_NORM: {
            if (gm_skip(d)) return;
            double w[1];
            reduce_gslots(d, 0, w);
            if (threadIdx.x) return;
            gm_finish_norm(d, sqrt(w[0]));
            return;
        }
        case FIN_TAY: {
            // number of terms: the smallest K whose geometric tail bound from
            // max|E_1| (||E_{k+1}|| <= ||E_k|| z/(k+1)) is under tol, with
            // z = rho tau, or sigma = rho tau / 2 for a shifted series
            if (a.sub >= (int)s[S_TK] || seq_finished(d, a.seq)) {
                if (threadIdx.x == 0) {
                    a.trec[kTrec * blockIdx.x] = 0.0;
                    a.trec[kTrec * blockIdx.x + 1] = 0.0;
                }
                return;
            }
            double w[1];
            reduce_slots<1, true>(d, w, a.zslots ? d.t_ntiles / d.nl + 1 : -1);
            if (threadIdx.x) return;
            const double z = s[S_TSIG] > 0.0 ? s[S_TSIG] : s[S_RHO] * s[S_TAU];
            const int kplan = (int)s[S_KPLAN];
            double term = w[0];
            int K = 1;
            while (K < kplan) {
                term *= z / (K + 1);
                const double q = z / (K + 2);
                if (q < 1.0 && term / (1.0 - q) <= s[S_TTOL]) break;
                K++;
            }
            atomicMax(a.kmax + 1, K);   // largest since the host last cleared it, held problems included
            s[S_TMAX] = w[0];
            s[S_KNEED] = K;
            // A group launching fewer steps than this problem needs (kg < K,
            // single sub-step groups only) holds the problem: no step takes
            // it, its state stays, and its interval is not advanced, so it
            // runs the same interval again once the host has raised kg.
            const bool hold = K > a.kg;
            s[S_THOLD] = hold ? 1.0 : 0.0;
            if (hold) {
                s[S_NHOLD] += 1.0;
                a.trec[kTrec * blockIdx.x] = s[S_TAU];
                a.trec[kTrec * blockIdx.x + 1] = 0.0;
                return;
            }
            s[S_TTERMS] += K - 1;
            // algorithmic bytes per node of the dominant sweeps: 12 (read h
            // and E_1, write h), 8 when h is E_1 itself (a problem's first)
            if (a.fused) {
                // ... and the last step (j = 2) reads the FP64 state and
                // writes it back instead of writing h: +16 B; a problem with
                // one term adds E_1 to the state in that launch (20 B)
                s[S_TBYTES] += K >= 3 ? 8.0 + 12.0 * (K - 3) + 24.0 : K >= 1 ? 20.0 : 0.0;
            } else {
                s[S_TBYTES] += K >= 2 ? 8.0 + 12.0 * (K - 2) : 0.0;
            }
            a.trec[kTrec * blockIdx.x] = s[S_TAU];
            a.trec[kTrec * blockIdx.x + 1] = K;
            atomicMax(a.kmax, K);
            return;
        }
    }
}

// Intervals of a device-resident sequence.  An interval that needs more RK4
// attempts than one launch sequence holds is continued by the next one: a
// problem starts a new interval only when its previous one is done, and its
// step counter moves only then.
__global__ void k_rk_begin(const ProbDesc *__restrict__ D, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    double *s = D[p].scal;
    if (seq_finished(D[p], 1) || s[S_STATUS] != (double)RK_DONE) return;
    s[S_T] = 0.0;
    s[S_H] = kMinStep;
    s[S_STATUS] = RK_NEED_K1;
    s[S_CALLS] = 0.0;
}

__global__ void k_rk_end(const ProbDesc *__restrict__ D, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    double *s = D[p].scal;
    if (seq_finished(D[p], 1) || s[S_STATUS] != (double)RK_DONE) return;
    s[S_TCALLS] += s[S_CALLS];
    s[S_CALLS] = 0.0;
}

__global__ void k_step_advance(const ProbDesc *__restrict__ D, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    if (seq_finished(D[p], 1) || D[p].scal[S_STATUS] != (double)RK_DONE || D[p].scal[S_THOLD] != 0.0) return;
    D[p].scal[S_STEP] += 1.0;
}

// copies up to 8 scalar slots of every problem into out[p * 8 + i]; with
// kmax, also the group's two term counts into kmax_out, and clears the
// second (the largest since the last gather)
struct Slots {
    int n;
    int s[8];
};
__global__ void k_gather(const ProbDesc *__restrict__ D, int nprob, Slots sl, double *out, int *kmax,
                         int *kmax_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (kmax && p == 0) {
        kmax_out[0] = kmax[0];
        kmax_out[1] = kmax[1];
        kmax[1] = 0;
    }
    if (p >= nprob) return;
    for (int i = 0; i < sl.n; i++) out[p * 8 + i] = D[p].scal[sl.s[i]];
}

// sets scalar slots of one problem from the launch's arguments (no host
// copy: the values travel with the launch)
struct ScalSet {
    int n;
    int slot[16];
    double val[16];
};
__global__ void k_set_scal(double *scal, ScalSet s) {
    if ((int)threadIdx.x < s.n) scal[s.slot[threadIdx.x]] = s.val[threadIdx.x];
}

// writes scalar slots of every problem: vals[p * 8 + i] into slot sl.s[i]
__global__ void k_scatter(const ProbDesc *__restrict__ D, int nprob, Slots sl, const double *vals) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    for (int i = 0; i < sl.n; i++) D[p].scal[sl.s[i]] = vals[p * 8 + i];
}

// ---------------------------------------------------------------------------
// transient: exponential (Taylor) integrator of the linear ODE over a
// sub-step tau (P constant within an interval):
//   y(t + tau) = y + sum_{k=1..K} E_k,  E_1 = tau f(y),
//   E_{k+1} = tau/(k+1) M E_k,  M = -C^-1 G (the operator without sources).
// The sum is evaluated in nested (Horner) form,
//   h = E_1 + (tau M/2)(E_1 + (tau M/3)(E_1 + ... + (tau M/K) E_1)),
// one sweep per step h <- E_1 + (tau/j) M h for j = K, ..., 2: a sweep reads
// h (with its halo) and E_1 and writes h, 12 bytes per FP32 node instead of
// the 16 of a forward sweep that also carries the running sum.  Conduction
// problems expand the series about the centre of M's spectrum instead (the
// host's plan gives the factors): E_1 = c_0 tau f(y) and the step is
// h <- E_1 + s_j M h + t_j h, with s_j, t_j from the problem's Taylor record
// (tau/j and 0 unshifted).  E_1 comes
// from the FP64 residual; h and E_1 are kept in TD (FP32 in the mixed form:
// they are temperature changes, not temperatures).  K comes from max|E_1|
// and the geometric tail bound (||C^-1 G||_inf <= rho), so the number of
// steps is known before the first one.  A group launches the steps of its
// largest plan Kg as j = Kg, ..., 2; a problem with K terms takes the steps
// j <= K and skips the others, and a launch with j above every problem's K
// ends at its first load.
// ---------------------------------------------------------------------------
template <typename TD, bool HET>
__global__ void __launch_bounds__(kBlock) k_tay_first(const ProbDesc *__restrict__ D,
                                                      const WorkItem *__restrict__ W, int e1b, int sub,
                                                      int seq, int *__restrict__ kmax) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kmax = 0;   // raised by the finish of each problem
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (sub >= (int)d.scal[S_TK] || seq_finished(d, seq)) return;
    const int ypar = (int)d.scal[S_YPAR];
    const double *y = d.b64[B_Y0 + ypar], *rhs = d.b64[B_RHS];
    TD *e1 = sizeof(TD) == 8 ? (TD *)d.b64[e1b] : (TD *)d.b32[e1b];
    const double tau = d.scal[S_TAU1];
    LdPlain<double> ld{y};
    double v[1] = {0};
    auto fin = [&](unsigned c, double f) {
        const TD e = (TD)(tau * f);
        st_at(e1, c, e);
        v[0] = fmax(v[0], fabs((double)e));
    };
    if (wi.tile < d.ntiles) {
        double ring[4];
        load_ring<double>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<double, HET> p;
            double b, ci;
        };
        tile_loop<InFlight<double, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V t{};
                load_pts<double, HET>(d, ld, q, t.p);
                t.b = ld_at(rhs, q.c);
                if (HET) t.ci = ld_at(Coef<double>::cinv(d), q.c);
                return t;
            },
            [&](const Cell &q, const V &t) {
                double flux, diag, gam;
                stencil<double, HET>(d, q, t.p, ring, flux, diag, gam);
                const double ci = HET ? t.ci : d.cinv[q.n];
                fin(q.c, ci * (t.b + flux - gam * t.p.c));
            });
    } else {
        __shared__ double xf[kMaxExtra], xd[kMaxExtra];
        package_terms<double>(d, ld, true, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            fin(c, d.x_cinv[k] * (rhs[c] + xf[k] - d.x_gamb[k] * y[c]));
        }
    }
    block_reduce<1, true>(v);
    write_partials<1>(d, wi.tile, v);
}

typedef double hs_d2 __attribute__((ext_vector_type(2)));
typedef float hs_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ hs_d2 ld2(const double *base, unsigned i) {
    return *(const HS_GLOBAL hs_d2 *)((const HS_GLOBAL char *)base + i * 8u);
}

// Coefficients of one layer for the first term of the FP32 series.
struct FirstLayer {
    double gx, gy, gu, gd, ci, gam, rgn, rge, src;
    const int *rn;
};
__device__ __forceinline__ FirstLayer first_layer(const ProbDesc &d, int n) {
    FirstLayer L;
    L.gx = d.gx[n];
    L.gy = d.gy[n];
    L.gu = n > 0 ? d.gz[max(n - 1, 0)] : 0.0;
    L.gd = n < d.nl - 1 ? d.gz[n] : 0.0;
    L.ci = d.cinv[n];
    L.gam = d.gamb[n];
    L.rgn = d.ring_gns[n];
    L.rge = d.ring_gew[n];
    L.src = L.gam * d.ambient;
    L.rn = d.ring_node[n];
    return L;
}

// E_1 = tau C^-1 (b + flux - g_amb x) of cell (i, j) in FP64, rounded to
// FP32, from the cell (xc), its neighbours in the plane (n, s, w, e) and the
// layers (u, d), and b = P + gamb T_ambient: the arithmetic of stencil() and
// k_tay_first in the same order.  A missing neighbour has zero conductance.
// Both first-term kernels compute a cell with this one function, so E_1
// does not depend on the kernel or its tiling.
__device__ __forceinline__ float first_cell(const ProbDesc &d, const FirstLayer &L, int i, int j, double xc,
                                            double xn, double xs, double xw, double xe, double xu, double xd,
                                            double b, double tau, const double (&ring)[4]) {
#pragma clang fp contract(on)
    const int nr = d.nr, nc = d.nc;
    const double gn = i > 0 ? L.gx : 0.0, gs = i < nr - 1 ? L.gx : 0.0;
    const double ge = j < nc - 1 ? L.gy : 0.0, gw = j > 0 ? L.gy : 0.0;
    double flux = gn * (xn - xc);
    flux += gs * (xs - xc);
    flux += ge * (xe - xc);
    flux += gw * (xw - xc);
    flux += L.gu * (xu - xc);
    flux += L.gd * (xd - xc);
    if (i == 0 && L.rn[0] >= 0) flux += L.rgn * (ring[0] - xc);
    if (i == nr - 1 && L.rn[1] >= 0) flux += L.rgn * (ring[1] - xc);
    if (j == nc - 1 && L.rn[2] >= 0) flux += L.rge * (ring[2] - xc);
    if (j == 0 && L.rn[3] >= 0) flux += L.rge * (ring[3] - xc);
    return (float)(tau * (L.ci * (b + flux - L.gam * xc)));
}

// The package nodes' E_1 (the package item of both first-term kernels).
__device__ __forceinline__ void first_package(const ProbDesc &d, const double *y, const double *rhs, float *e1,
                                              double tau, int power32, double (&v)[1]) {
    __shared__ double xf[kMaxExtra], xd[kMaxExtra];
    LdPlain<double> ld{y};
    package_terms<double>(d, ld, true, xf, xd);
    const int k = threadIdx.x;
    if (k < d.nx) {
        const long c = (long)d.nl * d.nr * d.nc + k;
        const double b = power32 ? 0.0 + d.x_gamb[k] * d.ambient : rhs[c];
        const float e = (float)(tau * (d.x_cinv[k] * (b + xf[k] - d.x_gamb[k] * y[c])));
        e1[c] = e;
        v[0] = fmax(v[0], fabs((double)e));
    }
}


// E_1 of a homogeneous problem for the FP32 series with double2 accesses:
// within the standard tile, each thread owns 2 consecutive columns x rpt/2
// rows, one row at a time.
//
// With power32 the right-hand side is the FP32 cell power of F_PW (none for
// a layer marked in lzero) plus the ambient sources, formed here: P + gamb
// T_ambient on cells, gamb T_ambient on the package nodes, as k_b2g adds
// them otherwise.
__global__ void __launch_bounds__(kBlock) k_tay_first2(const ProbDesc *__restrict__ D,
                                                       const WorkItem *__restrict__ W, int e1b, int sub, int seq,
                                                       int *__restrict__ kmax, int power32) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kmax = 0;   // raised by the finish of each problem
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    if (sub >= (int)d.scal[S_TK] || seq_finished(d, seq)) return;
    const int ypar = (int)d.scal[S_YPAR];
    const double *y = d.b64[B_Y0 + ypar], *rhs = d.b64[B_RHS];
    const float *pw = d.b32[F_PW];
    float *e1 = d.b32[e1b];
    const double tau = d.scal[S_TAU1];
    double v[1] = {0};
    if (wi.tile < d.ntiles) {
        const int tile = wi.tile, tpl = d.nti * d.ntj;
        const int n = tile / tpl, rem = tile - n * tpl;
        const int ti = rem / d.ntj, tj = rem - ti * d.ntj;
        const int sh = d.tx_shift;
        const int tx2 = threadIdx.x & ((1 << (sh - 1)) - 1), ty2 = threadIdx.x >> (sh - 1);
        const int G = d.rpt >> 1;
        const int rows = (kBlock >> sh) * d.rpt;
        const int nr = d.nr, nc = d.nc, nl = d.nl;
        const unsigned plane = (unsigned)nr * nc, ncell = plane * nl;
        const int j0 = (tj << sh) + 2 * tx2;
        const FirstLayer L = first_layer(d, n);
        const bool pz = power32 && d.lzero[n];
        double ring[4];
        for (int e = 0; e < 4; e++) ring[e] = ld_at(y, ncell + max(L.rn[e], 0));
        for (int g = 0; g < G; g++) {
            const int I = ti * rows + ty2 * G + g;
            const int i = min(I, nr - 1);
            const unsigned c = ((unsigned)n * nr + i) * nc + j0;
            const hs_d2 C = ld2(y, c);
            const hs_d2 N = ld2(y, i > 0 ? c - nc : c);
            const hs_d2 S = ld2(y, i < nr - 1 ? c + nc : c);
            const hs_d2 U = ld2(y, n > 0 ? c - plane : c);
            const hs_d2 Dn = ld2(y, n < nl - 1 ? c + plane : c);
            const double Wv = ld_at(y, j0 > 0 ? c - 1 : c);
            const double Ev = ld_at(y, j0 + 2 < nc ? c + 2 : c);
            hs_d2 B;
            if (!power32) {
                B = ld2(rhs, c);
            } else if (pz) {
                B = hs_d2{0.0 + L.src, 0.0 + L.src};
            } else {
                const hs_f2 p = *(const HS_GLOBAL hs_f2 *)((const HS_GLOBAL char *)pw + c * 4u);
                B = hs_d2{(double)p.x + L.src, (double)p.y + L.src};
            }
            if (I >= nr) continue;
            float e[2];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const double xw = k == 0 ? Wv : C[0];
                const double xe = k == 1 ? Ev : C[1];
                e[k] = first_cell(d, L, i, j0 + k, C[k], N[k], S[k], xw, xe, U[k], Dn[k], B[k], tau, ring);
                v[0] = fmax(v[0], fabs((double)e[k]));
            }
            const hs_f2 ev = {e[0], e[1]};
            *(HS_GLOBAL hs_f2 *)((HS_GLOBAL char *)e1 + c * 4u) = ev;
        }
    } else {
        first_package(d, y, rhs, e1, tau, power32, v);
    }
    block_reduce<1, true>(v);
    write_partials<1>(d, wi.tile, v);
}

// Buffers of Horner step j of a problem with K = S_KNEED terms: the first
// step (j = K) reads h = E_1; the steps then alternate d0 -> d1 -> d0.
struct HornerBufs {
    int in, out;
};
__device__ __forceinline__ HornerBufs horner_bufs(int K, int j, int e1b, int d0b, int d1b) {
    const int q = K - j;
    return HornerBufs{q == 0 ? e1b : ((q & 1) ? d0b : d1b), (q & 1) ? d1b : d0b};
}
// the buffer that holds the sum after the last step
__device__ __forceinline__ int horner_result(int K, int e1b, int d0b, int d1b) {
    return K <= 1 ? e1b : ((K & 1) ? d1b : d0b);
}

// Problem and tile of this workgroup in a Taylor launch, and the problem's
// {tau, terms} record of the current sub-step (terms 0: nothing to do).  A
// group whose problems all have the same tile count (nper > 0) maps the
// workgroup arithmetically, so the record and the descriptor are the first
// loads; otherwise through the work-list.
struct TayItem {
    int prob, tile, K;
    double tau;
    const double *rec;   // the problem's record (step factors at [2 j], [2 j + 1])
    const float *recf;   // the same in FP32
};
__device__ __forceinline__ TayItem tay_item_at(int b, const WorkItem *__restrict__ W, int nper,
                                               const double *__restrict__ trec) {
    TayItem t;
    if (nper > 0) {
        t.prob = b / nper;
        t.tile = b - t.prob * nper;
    } else {
        const WorkItem wi = W[b];
        t.prob = wi.prob;
        t.tile = wi.tile;
    }
    t.rec = trec + kTrec * t.prob;
    t.recf = (const float *)(t.rec + kTrecF);
    t.tau = t.rec[0];
    t.K = (int)t.rec[1];
    return t;
}
__device__ __forceinline__ TayItem tay_item(const WorkItem *__restrict__ W, int nper,
                                            const double *__restrict__ trec) {
    return tay_item_at(work_index(), W, nper, trec);
}

// FP32 Horner step of a homogeneous problem with float4 accesses: within the
// same tile, each thread owns 4 consecutive columns x rpt/4 rows (rpt >= 4,
// tile width >= 16), so a cell costs 2 load instructions instead of 8.  The
// arithmetic per cell is that of stencil() in the same order.
typedef float hs_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float *base, unsigned i) {
    const hs_f4 v = *(const HS_GLOBAL hs_f4 *)((const HS_GLOBAL char *)base + i * 4u);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(float *base, unsigned i, float4 v) {
    const hs_f4 w = {v.x, v.y, v.z, v.w};
    *(HS_GLOBAL hs_f4 *)((HS_GLOBAL char *)base + i * 4u) = w;
}
__device__ __forceinline__ float comp(const float4 &v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// The float4 sweeps take problems whose rows hold whole tiles (tile width
// 1 << t_shift, 16 to 256, dividing nc), so every float4 group lies inside
// its row.
__device__ __forceinline__ int term_ntj(const ProbDesc &d) { return d.nc >> d.t_shift; }
__device__ __forceinline__ int term_layer(const ProbDesc &d, int tile) { return tile / (d.t_nti * term_ntj(d)); }

__device__ __forceinline__ void st2(double *base, unsigned i, hs_d2 v) {
    *(HS_GLOBAL hs_d2 *)((HS_GLOBAL char *)base + i * 8u) = v;
}
// y[c..c+3] += v (the FP64 state takes the sub-step's increment)
__device__ __forceinline__ void add_state4(double *y, unsigned c, const float (&v)[4]) {
    hs_d2 a = ld2(y, c), b = ld2(y, c + 2);
    a.x += (double)v[0];
    a.y += (double)v[1];
    b.x += (double)v[2];
    b.y += (double)v[3];
    st2(y, c, a);
    st2(y, c + 2, b);
}

// Coefficients of one layer for the FP32 Horner steps.
struct LayerF {
    float gx, gy, gu, gd, ci, gam, rgn, rge;
    const int *rn;
};
__device__ __forceinline__ LayerF layer_f(const ProbDesc &d, int n) {
    LayerF L;
    L.gx = (float)d.gx[n];
    L.gy = (float)d.gy[n];
    L.gu = n > 0 ? (float)d.gz[max(n - 1, 0)] : 0.f;
    L.gd = n < d.nl - 1 ? (float)d.gz[n] : 0.f;
    L.ci = (float)d.cinv[n];
    L.gam = (float)d.gamb[n];
    L.rn = d.ring_node[n];
    L.rgn = (float)d.ring_gns[n];
    L.rge = (float)d.ring_gew[n];
    return L;
}

// One Horner step on 4 cells of row i, columns j0..j0+3 of a layer, from the
// cells (C), the rows above and below (N, S), the layers above and below (U,
// D), the columns left of j0 and right of j0+3 (xw0, xe3) and E_1 (A):
//   e = E_1 + s C^-1 (flux - g_amb x) + t x.
// A missing neighbour has zero conductance (its value is the cell's own).
// Both float4 sweeps (per layer and z-marching) compute a cell with this one
// function, so its arithmetic does not depend on the sweep or its tiling.
__device__ __forceinline__ void horner_row4(const ProbDesc &d, const LayerF &L, int i, int j0, const float4 &C,
                                            const float4 &N, const float4 &S, const float4 &U, const float4 &D,
                                            float xw0, float xe3, const float4 &A, float s, float sft,
                                            const float (&ring)[4], float (&e)[4]) {
    // multiply-adds fuse only within a statement, as written, wherever the
    // function is inlined
#pragma clang fp contract(on)
    const int nr = d.nr, nc = d.nc;
    if (i > 0 && i < nr - 1 && j0 > 0 && j0 + 4 < nc) {
        // away from the edges: every neighbour exists and no ring term
        // applies (the same operations as below with those known)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float xc = comp(C, k);
            const float xw = k == 0 ? xw0 : comp(C, k - 1);
            const float xe = k == 3 ? xe3 : comp(C, k + 1);
            float flux = L.gx * (comp(N, k) - xc);
            flux += L.gx * (comp(S, k) - xc);
            flux += L.gy * (xe - xc);
            flux += L.gy * (xw - xc);
            flux += L.gu * (comp(U, k) - xc);
            flux += L.gd * (comp(D, k) - xc);
            e[k] = comp(A, k) + s * (L.ci * (flux - L.gam * xc)) + sft * xc;
        }
        return;
    }
    const float gn = i > 0 ? L.gx : 0.f, gs = i < nr - 1 ? L.gx : 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int j = j0 + k;
        const float xc = comp(C, k);
        const float xw = k == 0 ? xw0 : comp(C, k - 1);
        const float xe = k == 3 ? xe3 : comp(C, k + 1);
        const float ge = j < nc - 1 ? L.gy : 0.f, gw = j > 0 ? L.gy : 0.f;
        float flux = gn * (comp(N, k) - xc);
        flux += gs * (comp(S, k) - xc);
        flux += ge * (xe - xc);
        flux += gw * (xw - xc);
        flux += L.gu * (comp(U, k) - xc);
        flux += L.gd * (comp(D, k) - xc);
        if (i == 0 && L.rn[0] >= 0) flux += L.rgn * (ring[0] - xc);
        if (i == nr - 1 && L.rn[1] >= 0) flux += L.rgn * (ring[1] - xc);
        if (j == nc - 1 && L.rn[2] >= 0) flux += L.rge * (ring[2] - xc);
        if (j == 0 && L.rn[3] >= 0) flux += L.rge * (ring[3] - xc);
        e[k] = comp(A, k) + s * (L.ci * (flux - L.gam * xc)) + sft * xc;
    }
}

// With y given (the last step, j = 2), the result is the sub-step's
// increment and goes into the FP64 state instead of hout.
template <int GM>
__device__ __forceinline__ void term_vec4(const ProbDesc &d, int tile, const float *hin, float *hout,
                                          const float *e1, float s, float sft, const float (&ring)[4], double *y) {
    const int ntj = term_ntj(d), tpl = d.t_nti * ntj;
    const int n = tile / tpl, rem = tile - n * tpl;
    const int ti = rem / ntj, tj = rem - ti * ntj;
    const int sh = d.t_shift;
    const int tx4 = threadIdx.x & ((1 << (sh - 2)) - 1), ty4 = threadIdx.x >> (sh - 2);
    const int G = d.t_rpt >> 2;
    const int rows = (kBlock >> sh) * d.t_rpt;
    const int nr = d.nr, nc = d.nc, nl = d.nl;
    const unsigned plane = (unsigned)nr * nc;
    const int j0 = (tj << sh) + 4 * tx4;
    const LayerF L = layer_f(d, n);
    float4 C[GM], N[GM], S[GM], U[GM], Dn[GM], A[GM];
    float Wv[GM], Ev[GM];
    int I[GM];
    unsigned cc[GM];
#pragma unroll
    for (int g = 0; g < GM; g++) {
        I[g] = ti * rows + ty4 * G + g;
        if (g < G) {
            const int i = min(I[g], nr - 1);
            const unsigned c = ((unsigned)n * nr + i) * nc + j0;
            cc[g] = c;
            C[g] = ld4(hin, c);
            N[g] = ld4(hin, i > 0 ? c - nc : c);
            S[g] = ld4(hin, i < nr - 1 ? c + nc : c);
            U[g] = ld4(hin, n > 0 ? c - plane : c);
            Dn[g] = ld4(hin, n < nl - 1 ? c + plane : c);
            Wv[g] = ld_at(hin, j0 > 0 ? c - 1 : c);
            Ev[g] = ld_at(hin, j0 + 4 < nc ? c + 4 : c);
            A[g] = ld4(e1, c);
        }
    }
#pragma unroll
    for (int g = 0; g < GM; g++) {
        if (g >= G || I[g] >= nr) continue;
        float e[4];
        horner_row4(d, L, I[g], j0, C[g], N[g], S[g], U[g], Dn[g], Wv[g], Ev[g], A[g], s, sft, ring, e);
        if (y) add_state4(y, cc[g], e);
        else st4(hout, cc[g], make_float4(e[0], e[1], e[2], e[3]));
    }
}

// The same step z-marching: the workgroup takes a column tile (the term
// tiling's columns and rows) through every layer from the top, so each
// thread keeps the layers above and below its cells in registers and reads
// each plane of h and E_1 once.  The neighbours in the plane come from the
// other threads: the columns left and right by a shuffle within the tile
// row, the rows above and below from the thread's own registers or, at its
// first and last row, from the threads above and below through LDS (one
// barrier per layer, two LDS buffers).  Only the tile's outer halo is read
// from memory, by the threads at the tile's edge; a missing neighbour takes
// the cell's own value, as the clamped loads of the per-layer sweep do.
template <int GM>
__device__ __forceinline__ void term_vec4z(const ProbDesc &d, int ztile, const float *hin, float *hout,
                                           const float *e1, float s, float sft, double *yb) {
    __shared__ float4 edge[2][2][kBlock];   // [layer parity][first row, last row][thread]
    const int ntj = term_ntj(d);
    const int ti = ztile / ntj, tj = ztile - ti * ntj;
    const int sh = d.t_shift;
    const int tw = 1 << (sh - 2);           // threads along a tile row
    const int tx4 = threadIdx.x & (tw - 1), ty4 = threadIdx.x >> (sh - 2);
    const int tyn = kBlock >> (sh - 2);     // thread rows
    const int G = d.t_rpt >> 2;
    const int rows = tyn * G;
    const int nr = d.nr, nc = d.nc, nl = d.nl;
    const unsigned plane = (unsigned)nr * nc;
    const unsigned ncell = plane * (unsigned)nl;
    const int j0 = (tj << sh) + 4 * tx4;
    const bool top = ty4 == 0, bottom = ty4 == tyn - 1;
    const bool left = tx4 == 0, right = tx4 == tw - 1;
    int I[GM], ic[GM];
    float4 U[GM], C[GM], Dn[GM];
#pragma unroll
    for (int g = 0; g < GM; g++) {
        I[g] = ti * rows + ty4 * G + g;
        ic[g] = min(I[g], nr - 1);
        if (g < G) {
            C[g] = ld4(hin, (unsigned)ic[g] * nc + j0);
            U[g] = C[g];
        }
    }
#pragma unroll 1
    for (int n = 0; n < nl; n++) {
        const unsigned off = (unsigned)n * plane;
        const int p = n & 1;
        const float4 Cf = C[0];
        float4 Cl = C[0];
#pragma unroll
        for (int g = 1; g < GM; g++)
            if (g == G - 1) Cl = C[g];
        edge[p][0][threadIdx.x] = Cf;
        edge[p][1][threadIdx.x] = Cl;
        float4 N0 = Cf, SG = Cl, A[GM];
        float Wv[GM], Ev[GM];
        {
            const unsigned cf = off + (unsigned)ic[0] * nc + j0;
            const unsigned cl = off + (unsigned)ic[G - 1] * nc + j0;
            if (top && ic[0] > 0) N0 = ld4(hin, cf - nc);
            if (bottom && ic[G - 1] < nr - 1) SG = ld4(hin, cl + nc);
        }
#pragma unroll
        for (int g = 0; g < GM; g++) {
            if (g >= G) continue;
            const unsigned c = off + (unsigned)ic[g] * nc + j0;
            Dn[g] = n < nl - 1 ? ld4(hin, c + plane) : C[g];
            A[g] = ld4(e1, c);
            Wv[g] = __shfl_up(C[g].w, 1, tw);
            Ev[g] = __shfl_down(C[g].x, 1, tw);
            if (left) Wv[g] = j0 > 0 ? ld_at(hin, c - 1) : C[g].x;
            if (right) Ev[g] = j0 + 4 < nc ? ld_at(hin, c + 4) : C[g].x;
        }
        float ring[4];
#pragma unroll
        for (int q = 0; q < 4; q++) ring[q] = ld_at(hin, ncell + max(d.ring_node[n][q], 0));
        const LayerF L = layer_f(d, n);
        __syncthreads();
        if (!top) N0 = edge[p][1][threadIdx.x - tw];
        if (!bottom) SG = edge[p][0][threadIdx.x + tw];
#pragma unroll
        for (int g = 0; g < GM; g++) {
            if (g >= G || I[g] >= nr) continue;
            // the rows next to row g inside the thread's own rows are its
            // registers
            const float4 N = g == 0 ? N0 : C[g - 1];
            const float4 S = g == G - 1 ? SG : C[g + 1];
            const unsigned c = off + (unsigned)ic[g] * nc + j0;
            float e[4];
            horner_row4(d, L, I[g], j0, C[g], N, S, U[g], Dn[g], Wv[g], Ev[g], A[g], s, sft, ring, e);
            if (yb) add_state4(yb, c, e);
            else st4(hout, c, make_float4(e[0], e[1], e[2], e[3]));
        }
#pragma unroll
        for (int g = 0; g < GM; g++) {
            U[g] = C[g];
            C[g] = Dn[g];
        }
    }
}

// A problem whose series has one term: y += E_1 over the term tile.
__device__ __forceinline__ void add_e1_vec4(const ProbDesc &d, int tile, const float *e1, double *y) {
    const int ntj = term_ntj(d), tpl = d.t_nti * ntj;
    const int n = tile / tpl, rem = tile - n * tpl;
    const int ti = rem / ntj, tj = rem - ti * ntj;
    const int sh = d.t_shift;
    const int tx4 = threadIdx.x & ((1 << (sh - 2)) - 1), ty4 = threadIdx.x >> (sh - 2);
    const int G = d.t_rpt >> 2;
    const int rows = (kBlock >> sh) * d.t_rpt;
    const int j0 = (tj << sh) + 4 * tx4;
    for (int g = 0; g < G; g++) {
        const int i = ti * rows + ty4 * G + g;
        if (i >= d.nr) continue;
        const unsigned c = ((unsigned)n * d.nr + i) * d.nc + j0;
        const float4 a = ld4(e1, c);
        const float v[4] = {a.x, a.y, a.z, a.w};
        add_state4(y, c, v);
    }
}

// E_1 of a homogeneous problem z-marching over the float4 term tiling's
// column tiles (groups whose Horner steps z-march): each thread owns 4
// consecutive columns x t_rpt/4 rows through every layer and keeps the
// layers above and below in registers; its neighbours in the plane come as
// in term_vec4z (a shuffle along the tile row, LDS across thread rows, the
// tile's outer halo from memory).  A cell's E_1 is first_cell's, as in
// k_tay_first2, with the FP32 cell power of F_PW (none for a layer marked in
// lzero).
//
// With ZP the kernel maps the power of the current sequence step itself, as
// k_b2g_power32 would (problems of at most kZPowTab blocks with blk_w): the
// block powers, the power of a cell inside each block (by block index) and
// which layers carry power are tabulated in LDS; a cell reads its code and,
// at a block boundary, its record of up to four blocks.  The values are
// those k_b2g_power32 writes to F_PW.
__device__ __forceinline__ double comp2(const hs_d2 &a, const hs_d2 &b, int k) { return k < 2 ? a[k] : b[k - 2]; }
constexpr int kZPowTab = 1024;   // blocks of a problem whose cell powers k_tay_first4z tabulates in LDS
typedef unsigned short hs_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ hs_u4 ld_u16x4(const unsigned short *base, unsigned i) {
    return *(const HS_GLOBAL hs_u4 *)((const HS_GLOBAL char *)base + i * 2u);
}
template <int GM, bool ZP>
__device__ __forceinline__ void first_vec4z(const ProbDesc &d, int ztile, const double *y, float *e1, double tau,
                                            double (&v)[1]) {
    // [layer parity][first row, last row (one row a thread: the same)][columns 0-1, 2-3][thread]
    __shared__ hs_d2 edge[2][GM][2][kBlock];
    __shared__ float zcp[ZP ? kZPowTab : 1];    // power of a cell inside block b (ZP)
    __shared__ double zbp[ZP ? kZPowTab : 1];   // power of block b (ZP)
    __shared__ int zpl[ZP ? kMaxLayers : 1];    // layer n carries power (ZP)
    if (ZP) {
        const double *vb = b2g_source(d, 1);
        if ((int)threadIdx.x < d.nl) zpl[threadIdx.x] = 0;
        __syncthreads();
        for (int b = threadIdx.x; b < d.nblocks; b += kBlock) {
            const double p = vb[b];
            zcp[b] = (float)((double)d.blk_w[b] * p);
            zbp[b] = p;
            if (p != 0.0) zpl[d.blk_layer[b]] = 1;
        }
        __syncthreads();
    }
    const int ntj = term_ntj(d);
    const int ti = ztile / ntj, tj = ztile - ti * ntj;
    const int sh = d.t_shift;
    const int tw = 1 << (sh - 2);             // threads along a tile row
    const int tx4 = threadIdx.x & (tw - 1), ty4 = threadIdx.x >> (sh - 2);
    const int tyn = kBlock >> (sh - 2);       // thread rows
    const int G = d.t_rpt >> 2;
    const int rows = tyn * G;
    const int nr = d.nr, nc = d.nc, nl = d.nl;
    const unsigned plane = (unsigned)nr * nc;
    const unsigned ncell = plane * (unsigned)nl;
    const int j0 = (tj << sh) + 4 * tx4;
    const bool top = ty4 == 0, bottom = ty4 == tyn - 1;
    const bool left = tx4 == 0, right = tx4 == tw - 1;
    const float *pw = d.b32[F_PW];
    int I[GM], ic[GM];
    hs_d2 Ua[GM], Ub[GM], Ca[GM], Cb[GM], Da[GM], Db[GM];
#pragma unroll
    for (int g = 0; g < GM; g++) {
        I[g] = ti * rows + ty4 * G + g;
        ic[g] = min(I[g], nr - 1);
        if (g < G) {
            const unsigned c = (unsigned)ic[g] * nc + j0;
            Ca[g] = ld2(y, c);
            Cb[g] = ld2(y, c + 2);
            Ua[g] = Ca[g];
            Ub[g] = Cb[g];
        }
    }
#pragma unroll 1
    for (int n = 0; n < nl; n++) {
        const unsigned off = (unsigned)n * plane;
        const int p = n & 1;
        const FirstLayer L = first_layer(d, n);
        const bool pz = ZP ? zpl[n] == 0 : d.lzero[n] != 0;
        hs_d2 Fa = Ca[0], Fb = Cb[0], La = Ca[0], Lb = Cb[0];
#pragma unroll
        for (int g = 1; g < GM; g++)
            if (g == G - 1) {
                La = Ca[g];
                Lb = Cb[g];
            }
        edge[p][0][0][threadIdx.x] = Fa;
        edge[p][0][1][threadIdx.x] = Fb;
        edge[p][GM - 1][0][threadIdx.x] = La;
        edge[p][GM - 1][1][threadIdx.x] = Lb;
        hs_d2 Na = Fa, Nb = Fb, Sa = La, Sb = Lb;
        {
            const unsigned cf = off + (unsigned)ic[0] * nc + j0;
            const unsigned cl = off + (unsigned)ic[G - 1] * nc + j0;
            if (top && ic[0] > 0) {
                Na = ld2(y, cf - nc);
                Nb = ld2(y, cf - nc + 2);
            }
            if (bottom && ic[G - 1] < nr - 1) {
                Sa = ld2(y, cl + nc);
                Sb = ld2(y, cl + nc + 2);
            }
        }
        double Wv[GM], Ev[GM];
        float4 P[GM];   // cell power
#pragma unroll
        for (int g = 0; g < GM; g++) {
            if (g >= G) continue;
            const unsigned c = off + (unsigned)ic[g] * nc + j0;
            if (n < nl - 1) {
                Da[g] = ld2(y, c + plane);
                Db[g] = ld2(y, c + plane + 2);
            } else {
                Da[g] = Ca[g];
                Db[g] = Cb[g];
            }
            Wv[g] = __shfl_up(Cb[g][1], 1, tw);
            Ev[g] = __shfl_down(Ca[g][0], 1, tw);
            if (left) Wv[g] = j0 > 0 ? ld_at(y, c - 1) : Ca[g][0];
            if (right) Ev[g] = j0 + 4 < nc ? ld_at(y, c + 4) : Ca[g][0];
            if (pz) {
                P[g] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else if (ZP) {
                // one-block cells from the table; the records of boundary
                // cells are all loaded before any is used
                const hs_u4 k4 = ld_u16x4(d.cell_u16[n], c - off);
                const int lb = d.layer_base[n];
                const int4 *uw = d.cell_uw[n] + 2 * (c - off);
                float q[4];
                int4 ru[4], rw[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const unsigned u = k4[k];
                    q[k] = u != kCellMixed ? zcp[lb + u] : 0.f;
                    if (u == kCellMixed) {
                        ru[k] = uw[2 * k];
                        rw[k] = uw[2 * k + 1];
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k4[k] == kCellMixed)
                        q[k] = (float)four_block_power(rw[k], zbp[lb + ru[k].x], zbp[lb + ru[k].y],
                                                       zbp[lb + ru[k].z], zbp[lb + ru[k].w]);
                P[g] = make_float4(q[0], q[1], q[2], q[3]);
            } else {
                P[g] = ld4(pw, c);
            }
        }
        double ring[4];
#pragma unroll
        for (int q = 0; q < 4; q++) ring[q] = ld_at(y, ncell + max(L.rn[q], 0));
        __syncthreads();
        if (!top) {
            Na = edge[p][GM - 1][0][threadIdx.x - tw];
            Nb = edge[p][GM - 1][1][threadIdx.x - tw];
        }
        if (!bottom) {
            Sa = edge[p][0][0][threadIdx.x + tw];
            Sb = edge[p][0][1][threadIdx.x + tw];
        }
#pragma unroll
        for (int g = 0; g < GM; g++) {
            if (g >= G || I[g] >= nr) continue;
            // the rows next to row g inside the thread's own rows are its
            // registers
            const hs_d2 na = g == 0 ? Na : Ca[g - 1], nb = g == 0 ? Nb : Cb[g - 1];
            const hs_d2 sa = g == G - 1 ? Sa : Ca[g + 1], sb = g == G - 1 ? Sb : Cb[g + 1];
            float e[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double xc = comp2(Ca[g], Cb[g], k);
                const double xw = k == 0 ? Wv[g] : comp2(Ca[g], Cb[g], k - 1);
                const double xe = k == 3 ? Ev[g] : comp2(Ca[g], Cb[g], k + 1);
                const double b = (double)comp(P[g], k) + L.src;
                e[k] = first_cell(d, L, I[g], j0 + k, xc, comp2(na, nb, k), comp2(sa, sb, k), xw, xe,
                                  comp2(Ua[g], Ub[g], k), comp2(Da[g], Db[g], k), b, tau, ring);
                v[0] = fmax(v[0], fabs((double)e[k]));
            }
            st4(e1, off + (unsigned)ic[g] * nc + j0, make_float4(e[0], e[1], e[2], e[3]));
        }
#pragma unroll
        for (int g = 0; g < GM; g++) {
            Ua[g] = Ca[g];
            Ub[g] = Cb[g];
            Ca[g] = Da[g];
            Cb[g] = Db[g];
        }
    }
}

// The first term over the z-marching term work-list of a FP32 Taylor group
// (power32 1): the partial max of a column tile goes to its slot, the
// package item's to the slot after the last column tile (the finish reduces
// t_ntiles / nl + 1 slots).  ZP (sequence intervals, seq 1): the cell power
// of the current sequence step is mapped here, with no k_b2g_power32 before.
// Three waves per SIMD (at most 168 VGPRs; the ZP form takes 170 unbounded).
template <int GM, bool ZP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_tay_first4z(const ProbDesc *__restrict__ D, const WorkItem *__restrict__ W, int nper, int e1b, int sub, int seq,
              int *__restrict__ kmax) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kmax = 0;   // raised by the finish of each problem
    const int b = work_index();
    int prob, ztile;
    if (nper > 0) {
        prob = b / nper;
        ztile = b - prob * nper;
    } else {
        const WorkItem wi = W[b];
        prob = wi.prob;
        ztile = wi.tile;
    }
    const ProbDesc &d = D[prob];
    if (sub >= (int)d.scal[S_TK] || seq_finished(d, seq)) return;
    const double *y = d.b64[B_Y0 + (int)d.scal[S_YPAR]], *rhs = d.b64[B_RHS];
    float *e1 = d.b32[e1b];
    const double tau = d.scal[S_TAU1];
    double v[1] = {0};
    if (ztile < d.t_ntiles / d.nl) first_vec4z<GM, ZP>(d, ztile, y, e1, tau, v);
    else first_package(d, y, rhs, e1, tau, 1, v);
    block_reduce<1, true>(v);
    write_partials<1>(d, ztile, v);
}

// Horner step j: h <- E_1 + (tau/j) M h
template <typename TD, bool HET>
__global__ void __launch_bounds__(kBlock) k_tay_term(const ProbDesc *__restrict__ D,
                                                     const WorkItem *__restrict__ W, int nper, int e1b,
                                                     int d0b, int d1b, int j, const int *__restrict__ kmax,
                                                     const double *__restrict__ trec) {
    if (j > *kmax) return;
    const TayItem it = tay_item(W, nper, trec);
    if (j > it.K) return;
    const ProbDesc &d = D[it.prob];
    const HornerBufs hb = horner_bufs(it.K, j, e1b, d0b, d1b);
    auto buf = [&](int b) -> TD * { return sizeof(TD) == 8 ? (TD *)d.b64[b] : (TD *)d.b32[b]; };
    const TD *hin = buf(hb.in), *e1 = buf(e1b);
    TD *hout = buf(hb.out);
    const TD s = sizeof(TD) == 4 ? (TD)it.recf[2 * j] : (TD)it.rec[2 * j];
    const TD sft = sizeof(TD) == 4 ? (TD)it.recf[2 * j + 1] : (TD)it.rec[2 * j + 1];
    LdPlain<TD> ld{hin};
    const WorkItem wi{it.prob, it.tile};
    if (wi.tile < d.ntiles) {
        TD ring[4];
        load_ring<TD>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<TD, HET> p;
            TD e1, ci;
        };
        tile_loop<InFlight<TD, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V t{};
                load_pts<TD, HET>(d, ld, q, t.p);
                t.e1 = ld_at(e1, q.c);
                if (HET) t.ci = ld_at(Coef<TD>::cinv(d), q.c);
                return t;
            },
            [&](const Cell &q, const V &t) {
                TD flux, diag, gam;
                stencil<TD, HET>(d, q, t.p, ring, flux, diag, gam);
                const TD ci = HET ? t.ci : (TD)d.cinv[q.n];
                st_at(hout, q.c, t.e1 + s * (ci * (flux - gam * t.p.c)) + sft * t.p.c);
            });
    } else {
        __shared__ TD xf[kMaxExtra], xd[kMaxExtra];
        package_terms<TD>(d, ld, true, xf, xd);
        const int kk = threadIdx.x;
        if (kk < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + kk;
            hout[c] = e1[c] + s * ((TD)d.x_cinv[kk] * (xf[kk] - (TD)d.x_gamb[kk] * hin[c])) + sft * hin[c];
        }
    }
}

// k_tay_term<float, false> for groups whose problems all take float4 tiles
// (vec4_ok), over the term tiling (t_*) and its work-list: per layer tiles,
// or (t_z) column tiles through every layer; the package item is the scalar
// kernel's.
//
// With fin set (the launch of step j = 2, the last of every series), the
// step's result -- the sub-step's increment -- is added to the FP64 state
// in place of k_tay_final, and a problem whose series has one term adds
// E_1 in the same launch.
// ZF: 0 every problem of the launch takes per-layer tiles, 1 every problem
// z-marches, 2 either by its t_z; GM: most row groups per thread (t_rpt / 4)
// of the launch's problems.  Each form compiled alone keeps its own register
// budget.
template <int ZF, int GM, int FIN>
__device__ __forceinline__ void term4_item(const ProbDesc *__restrict__ D, const TayItem &it, int e1b, int d0b,
                                           int d1b, int j) {
    constexpr bool fin = FIN != 0;
    if (it.K == 0 || (!fin && j > it.K)) return;
    const ProbDesc &d = D[it.prob];
    const bool z = ZF == 2 ? d.t_z != 0 : ZF == 1;
    const int tpl = d.t_ntiles / d.nl;                 // term tiles per layer
    const int ntile = z ? tpl : d.t_ntiles;            // cell items of the problem
    const float *e1 = d.b32[e1b];
    double *y = fin ? d.b64[B_Y0 + (int)d.scal[S_YPAR]] : nullptr;
    if (j > it.K) {   // fin, one term
        if (it.tile < ntile) {
            if (z)
                for (int n = 0; n < d.nl; n++) add_e1_vec4(d, n * tpl + it.tile, e1, y);
            else
                add_e1_vec4(d, it.tile, e1, y);
        } else if ((int)threadIdx.x < d.nx) {
            const unsigned c = (unsigned)d.nl * d.nr * d.nc + threadIdx.x;
            y[c] += (double)e1[c];
        }
        return;
    }
    const HornerBufs hb = horner_bufs(it.K, j, e1b, d0b, d1b);
    const float *hin = d.b32[hb.in];
    float *hout = d.b32[hb.out];
    const float s = it.recf[2 * j], sft = it.recf[2 * j + 1];
    LdPlain<float> ld{hin};
    if (it.tile < ntile) {
        if (z) {
            term_vec4z<GM>(d, it.tile, hin, hout, e1, s, sft, y);
        } else {
            float ring[4];
            load_ring<float>(d, ld, term_layer(d, it.tile), ring);
            term_vec4<GM>(d, it.tile, hin, hout, e1, s, sft, ring, y);
        }
    } else {
        __shared__ float xf[kMaxExtra], xd[kMaxExtra];
        package_terms<float>(d, ld, true, xf, xd);
        const int kk = threadIdx.x;
        if (kk < d.nx) {
            const unsigned c = (unsigned)d.nl * d.nr * d.nc + kk;
            const float v = e1[c] + s * ((float)d.x_cinv[kk] * (xf[kk] - (float)d.x_gamb[kk] * hin[c])) + sft * hin[c];
            if (fin) y[c] += (double)v;
            else hout[c] = v;
        }
    }
}

// FIN: the launch of step j = 2, which adds into the FP64 state
template <int ZF, int GM, int FIN>
__global__ void __launch_bounds__(kBlock) k_tay_term4(const ProbDesc *__restrict__ D,
                                                      const WorkItem *__restrict__ W, int nper, int e1b,
                                                      int d0b, int d1b, int j, const int *__restrict__ kmax,
                                                      const double *__restrict__ trec) {
    if (!FIN && j > *kmax) return;
    term4_item<ZF, GM, FIN>(D, tay_item(W, nper, trec), e1b, d0b, d1b, j);
}

// y += h (the sum of the sub-step's terms)
template <typename TD>
__global__ void __launch_bounds__(kBlock) k_tay_final(const ProbDesc *__restrict__ D,
                                                      const WorkItem *__restrict__ W, int nper, int e1b,
                                                      int d0b, int d1b, const double *__restrict__ trec) {
    const TayItem it = tay_item(W, nper, trec);
    if (it.K == 0) return;
    const ProbDesc &d = D[it.prob];
    const WorkItem wi{it.prob, it.tile};
    const int ypar = (int)d.scal[S_YPAR];
    double *y = d.b64[B_Y0 + ypar];
    const int hb = horner_result(it.K, e1b, d0b, d1b);
    const TD *acc = sizeof(TD) == 8 ? (const TD *)d.b64[hb] : (const TD *)d.b32[hb];
    if (wi.tile < d.ntiles) {
        struct V {
            double y;
            TD a;
        };
        tile_loop<8>(
            d, wi.tile,
            [&](const Cell &q) {
                V t;
                t.y = ld_at(y, q.c);
                t.a = ld_at(acc, q.c);
                return t;
            },
            [&](const Cell &q, const V &t) { st_at(y, q.c, t.y + (double)t.a); });
    } else if (threadIdx.x < d.nx) {
        const long c = (long)d.nl * d.nr * d.nc + threadIdx.x;
        y[c] += (double)acc[c];
    }
}

// slope on a given state (operator-level queries): out = (rhs - G_t x)/C
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_slope(const ProbDesc *__restrict__ D,
                                                  const WorkItem *__restrict__ W, int xb, int outb) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    const double *x = d.b64[xb], *rhs = d.b64[B_RHS];
    double *out = d.b64[outb];
    LdPlain<double> ld{x};
    if (wi.tile < d.ntiles) {
        double ring[4];
        load_ring<double>(d, ld, tile_layer(d, wi.tile), ring);
        struct V {
            Pts<double, HET> p;
            double b, ci;
        };
        tile_loop<InFlight<double, HET>::value>(
            d, wi.tile,
            [&](const Cell &q) {
                V t{};
                load_pts<double, HET>(d, ld, q, t.p);
                t.b = ld_at(rhs, q.c);
                if (HET) t.ci = ld_at(Coef<double>::cinv(d), q.c);
                return t;
            },
            [&](const Cell &q, const V &t) {
                double flux, diag, gam;
                stencil<double, HET>(d, q, t.p, ring, flux, diag, gam);
                const double ci = HET ? t.ci : d.cinv[q.n];
                st_at(out, q.c, ci * (t.b + flux - gam * t.p.c));
            });
    } else {
        __shared__ double xf[kMaxExtra], xd[kMaxExtra];
        package_terms<double>(d, ld, true, xf, xd);
        const int k = threadIdx.x;
        if (k < d.nx) {
            const long c = (long)d.nl * d.nr * d.nc + k;
            out[c] = d.x_cinv[k] * (rhs[c] + xf[k] - d.x_gamb[k] * x[c]);
        }
    }
}

// ---------------------------------------------------------------------------
// conductance assembly of heterogeneous models (find_res :26-136,
// find_res_3D :144-173, find_cap_3D :176-186) from per-cell resistances
// ---------------------------------------------------------------------------
struct AsmIn {
    const double *rx, *ry, *rz, *cap;   // per cell (detailed values or layer values)
    const int *fluid;                   // per cell: 1 for coolant cells
    const double *htc;                  // per layer (0 if not microchannel)
    const double *thick;                // per layer
    const double *press;                // per cell pressure (microchannel layers)
    const int *types;                   // per cell network cell type
    const double *hydro, *cv, *tin;     // per layer: hydraulic conductance, coolant capacity, inlet T
    int nl, nr, nc, detailed;
    int superlu_outlets;                // outlet cells as the SuperLU build's matrix has them
    double cw, ch;
};

// ---------------------------------------------------------------------------
// Pressure network of a microchannel layer (microchannel.c:378-565).  FLUID
// cells satisfy sum over fluid-type neighbours (p - p_nbr) = 0 (the uniform
// hydraulic conductance of the cells cancels), with p the pumping pressure at
// INLET cells and 0 at OUTLET cells.  The reference assembles the dense
// system over the fluid cells and solves it by Gauss-Jordan (gaussj,
// util.c:621-699), O(n_fluid^3).  Here the same system, symmetric positive
// definite once the fixed pressures move to the right-hand side, is solved
// by Jacobi-preconditioned CG over the plane (every other cell held fixed)
// to a 1e-15 relative residual.  Scalars stay on the device: partial sums
// per workgroup, one workgroup folding them in a fixed order (so the result
// does not depend on scheduling), and a done flag that freezes every kernel
// once the residual is reached.
// ---------------------------------------------------------------------------
struct PrState {
    double rz, bb, rr, alpha, beta;
    int done, iters;
};

__device__ __forceinline__ bool pr_fluid_type(int t) { return t == CELL_FLUID || t == CELL_INLET || t == CELL_OUTLET; }

// fluid-type neighbours of cell c = i*nc + j, in the order N, S, W, E
template <typename F>
__device__ __forceinline__ void pr_nbrs(int nr, int nc, int i, int j, long c, F &&fn) {
    if (i > 0) fn(c - nc);
    if (i < nr - 1) fn(c + nc);
    if (j > 0) fn(c - 1);
    if (j < nc - 1) fn(c + 1);
}

// x: fixed values (pump at inlets) and 0; d: fluid-type neighbour count of
// FLUID cells; r = b (fixed neighbours' pressure), z = r/d, s = z;
// partials {r.z, b.b}
__global__ void __launch_bounds__(kBlock) k_pr_setup(const int *__restrict__ types, int nr, int nc, double pump,
                                                     double *__restrict__ x, double *__restrict__ d,
                                                     double *__restrict__ r, double *__restrict__ z,
                                                     double *__restrict__ s, double *__restrict__ part) {
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (c < (long)nr * nc) {
        const int t = types[c], i = (int)(c / nc), j = (int)(c % nc);
        double dc = 0.0, bc = 0.0;
        if (t == CELL_FLUID)
            pr_nbrs(nr, nc, i, j, c, [&](long o) {
                const int to = types[o];
                if (!pr_fluid_type(to)) return;
                dc += 1.0;
                if (to == CELL_INLET) bc += pump;
            });
        x[c] = t == CELL_INLET ? pump : 0.0;
        d[c] = dc;
        r[c] = bc;
        const double zc = dc > 0.0 ? bc / dc : 0.0;
        z[c] = zc;
        s[c] = zc;
        v[0] = bc * zc;
        v[1] = bc * bc;
    }
    block_reduce<2>(v);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = v[0], part[2 * blockIdx.x + 1] = v[1];
}

// q = A s over the FLUID cells; partial s.q
__global__ void __launch_bounds__(kBlock) k_pr_apply(const int *__restrict__ types, int nr, int nc,
                                                     const double *__restrict__ d, const double *__restrict__ s,
                                                     double *__restrict__ q, double *__restrict__ part,
                                                     const PrState *__restrict__ st) {
    if (st->done) return;
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (c < (long)nr * nc && types[c] == CELL_FLUID) {
        const int i = (int)(c / nc), j = (int)(c % nc);
        double acc = d[c] * s[c];
        pr_nbrs(nr, nc, i, j, c, [&](long o) {
            if (types[o] == CELL_FLUID) acc -= s[o];
        });
        q[c] = acc;
        v[0] = s[c] * acc;
    }
    block_reduce<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// x += alpha s, r -= alpha q, z = r/d; partials {r.r, r.z}
__global__ void __launch_bounds__(kBlock) k_pr_update(const int *__restrict__ types, long n,
                                                      const double *__restrict__ d, const double *__restrict__ s,
                                                      const double *__restrict__ q, double *__restrict__ x,
                                                      double *__restrict__ r, double *__restrict__ z,
                                                      double *__restrict__ part, const PrState *__restrict__ st) {
    if (st->done) return;
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (c < n && types[c] == CELL_FLUID) {
        const double a = st->alpha;
        x[c] += a * s[c];
        const double rc = r[c] - a * q[c];
        r[c] = rc;
        const double zc = rc / d[c];
        z[c] = zc;
        v[0] = rc * rc;
        v[1] = rc * zc;
    }
    block_reduce<2>(v);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = v[0], part[2 * blockIdx.x + 1] = v[1];
}

// s = z + beta s
__global__ void __launch_bounds__(kBlock) k_pr_dir(const int *__restrict__ types, long n,
                                                   const double *__restrict__ z, double *__restrict__ s,
                                                   const PrState *__restrict__ st) {
    if (st->done) return;
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    if (c < n && types[c] == CELL_FLUID) s[c] = z[c] + st->beta * s[c];
}

// the scalars of one phase from the partial sums (one workgroup, fixed order):
// phase 0 after setup, 1 after apply (alpha), 2 after update (beta, stop)
__global__ void __launch_bounds__(kBlock) k_pr_scalars(const double *__restrict__ part, int nparts, int phase,
                                                       PrState *__restrict__ st) {
    if (phase > 0 && st->done) return;
    const int w = phase == 1 ? 1 : 2;
    double v[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < nparts; k += kBlock)
        for (int q = 0; q < w; q++) v[q] += part[(long)w * k + q];
    block_reduce<2>(v);
    if (threadIdx.x != 0) return;
    if (phase == 0) {
        st->rz = v[0];
        st->bb = v[1];
        st->iters = 0;
        st->done = v[1] == 0.0;
    } else if (phase == 1) {
        st->alpha = st->rz / v[0];
    } else {
        st->rr = v[0];
        st->iters += 1;
        if (v[0] <= 1e-30 * st->bb) {
            st->done = 1;
        } else {
            st->beta = v[1] / st->rz;
            st->rz = v[1];
        }
    }
}

// ---------------------------------------------------------------------------
// backward Euler (the SuperLU path's integrator, RCutil.c:268-428 as called
// by compute_temp_grid :3632-3662): one implicit step over the interval,
//   (C/h + G) y' = C y / h + P + sources,
// solved by CG with a descriptor whose ambient conductances carry the C/h
// shift.  G is the assembled matrix of that path, whose secondary package
// links are the steady ones (build_transient_grid_matrix, both directions of
// the solder-substrate link), and the secondary-path package nodes have no
// capacitance in it (build_diagonal_matrix :6750-6768).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double be_cap(const ProbDesc &d, const Cell &q, bool het) {
    return het ? 1.0 / d.cinvc[q.c] : 1.0 / d.cinv[q.n];
}

// out = rhs + sigma C y (descriptor without the shift)
template <bool HET>
__global__ void __launch_bounds__(kBlock) k_be_rhs(const ProbDesc *__restrict__ D,
                                                   const WorkItem *__restrict__ W, int yb, int rhsb, int outb,
                                                   double sigma) {
    const WorkItem wi = W[work_index()];
    const ProbDesc &d = D[wi.prob];
    const double *y = d.b64[yb], *rhs = d.b64[rhsb];
    double *out = d.b64[outb];
    if (wi.tile < d.ntiles) {
        tile_loop<8>(
            d, wi.tile, [&](const Cell &q) { return ld_at(rhs, q.c) + sigma * be_cap(d, q, HET) * ld_at(y, q.c); },
            [&](const Cell &q, double v) { st_at(out, q.c, v); });
    } else if ((int)threadIdx.x < d.nx) {
        const int k = threadIdx.x;
        const unsigned c = (unsigned)d.nl * d.nr * d.nc + k;
        out[c] = rhs[c] + (k < EXTRA ? sigma / d.x_cinv[k] : 0.0) * y[c];
    }
}

// per-cell shift sigma C and the shifted Jacobi diagonal of a heterogeneous model
__global__ void k_be_coef(const ProbDesc *__restrict__ D, long ncell, double sigma, double *shift,
                          double *dinv) {
    const ProbDesc &d = *D;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const double s = sigma / d.cinvc[c];
    shift[c] = s;
    dinv[c] = 1.0 / (1.0 / d.dinvc[c] + s);
}

__global__ void k_assemble(AsmIn in, double *gE, double *gS, double *gD, double *cinv,
                           double *aN, double *aS, double *aE, double *aW, double *aSelf,
                           double *bconst) {
    const long plane = (long)in.nr * in.nc;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane * in.nl) return;
    const int n = (int)(c / plane);
    const long cc = c - n * plane;
    const int i = (int)(cc / in.nc), j = (int)(cc % in.nc);
    const int f = in.fluid[c];
    // lateral faces: detailed half + half, or solid half + 1/(htc A) for a
    // coolant/wall pair (the fluid layer's htc, face area cw*t or ch*t)
    auto lat = [&](long o, const double *r, double area_len) {
        const int fo = in.fluid[o];
        if (f != fo) {
            const double rs = f ? r[o] : r[c];
            return 1.0 / (rs / 2.0 + 1.0 / (in.htc[n] * area_len * in.thick[n]));
        }
        return 1.0 / (r[c] / 2.0 + r[o] / 2.0);
    };
    gE[c] = j < in.nc - 1 ? lat(c + 1, in.ry, in.ch) : 0.0;
    gS[c] = i < in.nr - 1 ? lat(c + in.nc, in.rx, in.cw) : 0.0;
    if (n < in.nl - 1) {
        const long o = c + plane;
        const int fo = in.fluid[o];
        double r;
        if (f != fo) {
            const double h = in.htc[n + 1] != 0.0 ? in.htc[n + 1] : in.htc[n];
            if (n == 0) r = in.rz[c] + 1.0 / (h * in.ch * in.cw);
            else r = (f ? in.rz[o] : in.rz[c]) / 2.0 + 1.0 / (h * in.ch * in.cw);
        } else if (in.detailed) {
            r = n == 0 ? in.rz[c] + in.rz[o] / 2.0 : in.rz[c] / 2.0 + in.rz[o] / 2.0;
        } else {
            r = in.rz[c];
        }
        gD[c] = 1.0 / r;
    } else {
        gD[c] = 0.0;
    }
    cinv[c] = 1.0 / in.cap[c];
    // advection (temperature_grid.c:3256-3590): a_dir = Cv Q(nbr->c) / 2
    double an = 0, as = 0, ae = 0, aw = 0, self = 0, b = 0;
    if (f && in.hydro[n] != 0.0) {
        const double hC = in.hydro[n], Cv = in.cv[n];
        const double pc = in.press[c];
        double inflow = 0.0;
        auto q = [&](long o) { return (in.press[o] - pc) * hC; };
        if (i > 0 && in.fluid[c - in.nc]) { const double Q = q(c - in.nc); an = Cv * Q / 2.0; inflow += Q; }
        if (i < in.nr - 1 && in.fluid[c + in.nc]) { const double Q = q(c + in.nc); as = Cv * Q / 2.0; inflow += Q; }
        if (j > 0 && in.fluid[c - 1]) { const double Q = q(c - 1); aw = Cv * Q / 2.0; inflow += Q; }
        if (j < in.nc - 1 && in.fluid[c + 1]) { const double Q = q(c + 1); ae = Cv * Q / 2.0; inflow += Q; }
        const bool edge = i == 0 || i == in.nr - 1 || j == 0 || j == in.nc - 1;
        const int t = in.types[c];
        if (t == CELL_INLET && edge) b = Cv * (-inflow) * in.tin[n];
        else if (t == CELL_OUTLET && edge) self = Cv * (-inflow);
        if (t == CELL_OUTLET && in.superlu_outlets) {
            // the SuperLU build's matrix (build_transient_grid_matrix,
            // temperature_grid.c:5313-5345) takes an outlet's advection from
            // its first fluid neighbour alone, in the order N, S, W, E, as
            // -Cv Q/2 to that neighbour and +Cv Q/2 on the diagonal (at any
            // outlet, on the edge or not)
            const int first = (i > 0 && in.fluid[c - in.nc]) ? 0
                              : (i < in.nr - 1 && in.fluid[c + in.nc]) ? 1
                              : (j > 0 && in.fluid[c - 1]) ? 2
                              : (j < in.nc - 1 && in.fluid[c + 1]) ? 3 : -1;
            const double a = first == 0 ? an : first == 1 ? as : first == 2 ? aw : first == 3 ? ae : 0.0;
            an = first == 0 ? a : 0.0;
            as = first == 1 ? a : 0.0;
            aw = first == 2 ? a : 0.0;
            ae = first == 3 ? a : 0.0;
            self = first >= 0 ? -2.0 * a : 0.0;
        }
    }
    aN[c] = an;
    aS[c] = as;
    aE[c] = ae;
    aW[c] = aw;
    aSelf[c] = self;
    bconst[c] = b;
}

// FP32 copies of the coefficient arrays and the inverse diagonal of G
__global__ void k_coef_finish(const ProbDesc *__restrict__ D, long ncell, float *fE, float *fS,
                              float *fD, float *fcinv, float *faN, float *faS, float *faE,
                              float *faW, float *faSelf, double *dinv, float *fdinv) {
    const ProbDesc &d = D[0];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const long plane = (long)d.nr * d.nc;
    Cell q;
    q.n = (int)(c / plane);
    const long cc = c - q.n * plane;
    q.i = (int)(cc / d.nc);
    q.j = (int)(cc % d.nc);
    q.c = c;
    q.active = true;
    fE[c] = (float)d.gE[c];
    fS[c] = (float)d.gS[c];
    fD[c] = (float)d.gD[c];
    fcinv[c] = (float)d.cinvc[c];
    faN[c] = (float)d.aN[c];
    faS[c] = (float)d.aS[c];
    faE[c] = (float)d.aE[c];
    faW[c] = (float)d.aW[c];
    faSelf[c] = (float)d.aSelf[c];
    // diagonal of G: conduction + ring + ambient - advection self terms
    double diag = d.gamb[q.n];
    const int nr = d.nr, nc = d.nc, nl = d.nl, i = q.i, j = q.j, n = q.n;
    if (i > 0) diag += d.gS[c - nc];
    if (i < nr - 1) diag += d.gS[c];
    if (j < nc - 1) diag += d.gE[c];
    if (j > 0) diag += d.gE[c - 1];
    if (n > 0) diag += d.gD[c - plane];
    if (n < nl - 1) diag += d.gD[c];
    const int *rn = d.ring_node[n];
    if (i == 0 && rn[0] >= 0) diag += d.ring_gns[n];
    if (i == nr - 1 && rn[1] >= 0) diag += d.ring_gns[n];
    if (j == nc - 1 && rn[2] >= 0) diag += d.ring_gew[n];
    if (j == 0 && rn[3] >= 0) diag += d.ring_gew[n];
    diag -= d.aN[c] + d.aS[c] + d.aE[c] + d.aW[c] + d.aSelf[c];
    dinv[c] = 1.0 / diag;
    fdinv[c] = (float)(1.0 / diag);
}

__global__ void k_fill(double *x, long n, double v) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) x[c] = v;
}

}  // namespace hs

#include "hs_engine.inl"
