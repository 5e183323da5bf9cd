// Block thermal model (temperature_block.c): set-up, dispatch and text
// formats on the host.  The model has four node layers per floorplan unit
// (silicon, interface, spreader, sink: node layer*n_units + unit) and the 12
// package nodes (NL*n_units + k).  The conductance matrix B (B T = P in the
// steady state) and the capacitances A are assembled here from the
// floorplan's geometry; the LU factorisation, the steady solve and the
// adaptive-RK4 transient run on the device (hs_block.hip).
//
// Host mirrors keep the reference's struct fields filled (b, a, inva, c, g,
// len, lateral conductances, borders; lu and p once factorised), in util.c's
// dmatrix layout: row pointers into one contiguous block.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>

#include <algorithm>
#include <string>

#include "hs_internal.h"

namespace hs {

static_assert(offsetof(block_model_t, config) == 8, "block_model_t (temperature_block.h:16-67)");
static_assert(offsetof(block_model_t, pack) == 8 + sizeof(thermal_config_t) + 6 * 8,
              "block_model_t (temperature_block.h:16-67)");

// ---- util.c:159-226 matrix layout -------------------------------------------
template <typename T>
static T **alloc_mat(int nr, int nc) {
    T **m = (T **)calloc(nr > 0 ? nr : 1, sizeof(T *));
    if (!m) fatal("memory allocation error\n");
    m[0] = (T *)calloc((size_t)(nr > 0 ? nr : 1) * (nc > 0 ? nc : 1), sizeof(T));
    if (!m[0]) fatal("memory allocation error\n");
    for (int i = 1; i < nr; i++) m[i] = m[0] + (size_t)nc * i;
    return m;
}
template <typename T>
static void free_mat(T **m) {
    if (!m) return;
    free(m[0]);
    free(m);
}
// resize_dmatrix: the same storage re-strided to nc columns
static void restride(double **m, int nr, int nc) {
    for (int i = 1; i < nr; i++) m[i] = m[0] + (size_t)nc * i;
}
static double *alloc_vec(int n) {
    double *v = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    if (!v) fatal("memory allocation error\n");
    return v;
}

BlockModel *block_of(block_model_t *b) {
    if (!b) fatal("null model\n");
    BlockHandle *h = reinterpret_cast<BlockHandle *>(reinterpret_cast<char *>(b) - offsetof(BlockHandle, pub));
    return h->impl;
}

// ---- floorplan adjacency (flp.c:1404-1499) ------------------------------------
namespace {
struct Box {
    double x1, x2, y1, y2;
};
Box box_of(const unit_t &u) { return {u.leftx, u.leftx + u.width, u.bottomy, u.bottomy + u.height}; }
// touching at a corner only does not make two units neighbours
bool corner_only(const Box &a, const Box &b) {
    return (eq(a.x2, b.x1) && eq(a.y2, b.y1)) || (eq(a.x1, b.x2) && eq(a.y1, b.y2)) ||
           (eq(a.x2, b.x1) && eq(a.y1, b.y2)) || (eq(a.x1, b.x2) && eq(a.y2, b.y1));
}
// the closed intervals [a1,a2] and [b1,b2] meet
bool meet(double a1, double a2, double b1, double b2) {
    return (b1 >= a1 && b1 <= a2) || (b2 >= a1 && b2 <= a2) || (a1 >= b1 && a1 <= b2) ||
           (a2 >= b1 && a2 <= b2);
}
// side by side: a vertical edge in common (is_horiz_adj)
bool side_by_side(const Box &a, const Box &b) {
    return !corner_only(a, b) && (eq(a.x1, b.x2) || eq(a.x2, b.x1)) && meet(a.y1, a.y2, b.y1, b.y2);
}
// one above the other: a horizontal edge in common (is_vert_adj)
bool stacked(const Box &a, const Box &b) {
    return !corner_only(a, b) && (eq(a.y1, b.y2) || eq(a.y2, b.y1)) && meet(a.x1, a.x2, b.x1, b.x2);
}
// length of the common edge (get_shared_len); 0 for a unit and itself
double common_edge(const Box &a, const Box &b, bool same) {
    if (same) return 0.0;
    double p11 = 0, p12 = 0, p21 = 0, p22 = 0;
    if (side_by_side(a, b)) p11 = a.y1, p12 = a.y2, p21 = b.y1, p22 = b.y2;
    if (stacked(a, b)) p11 = a.x1, p12 = a.x2, p21 = b.x1, p22 = b.x2;
    return std::min(p12, p22) - std::max(p11, p21);
}

void check_floorplan(const block_model_t &m, const flp_t *flp, const char *msg) {
    if (m.flp != flp || m.n_units != flp->n_units || m.n_nodes != NL * flp->n_units + EXTRA) fatal(msg);
}

// package nodes' ambient power (set_internal_power_block, temperature_block.c:382-389):
// interface and spreader nodes carry none, sink and package nodes the
// ambient's injection through their ambient conductance
void internal_power(const block_model_t &m, double *power) {
    const int n = m.n_units;
    for (int i = 0; i < n; i++) power[IFACE * n + i] = 0.0, power[HSP * n + i] = 0.0;
    for (int i = 0; i < n + EXTRA; i++) power[HSINK * n + i] = m.config.ambient * m.g_amb[i];
}

FILE *open_out(const char *file) {
    FILE *fp = !strcasecmp(file, "stdout") ? stdout : !strcasecmp(file, "stderr") ? stderr : fopen(file, "w");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for writing\n");
    return fp;
}
void close_out(FILE *fp) {
    if (fp != stdout && fp != stderr) fclose(fp);
}
}  // namespace

// ---- construction (temperature_block.c:19-56, temperature.c:616-633) ----------
static BlockHandle *block_new(thermal_config_t *config, flp_t *placeholder) {
    if (!placeholder) fatal("the block model needs a floorplan\n");
    BlockHandle *h = (BlockHandle *)calloc(1, sizeof(BlockHandle));
    if (!h) fatal("memory allocation error\n");
    try {
        h->impl = new BlockModel();
        h->impl->m = &h->pub;
        block_model_t &m = h->pub;
        const int n = placeholder->n_units, N = NL * n + EXTRA;
        m.config = *config;
        m.n_units = m.base_n_units = n;
        m.n_nodes = N;
        m.border = alloc_mat<int>(n, 4);
        m.len = alloc_mat<double>(n, n);
        m.g = alloc_mat<double>(N, N);
        double **lat[8] = {&m.gx, &m.gy, &m.gx_int, &m.gy_int, &m.gx_sp, &m.gy_sp, &m.gx_hs, &m.gy_hs};
        for (double **v : lat) *v = alloc_vec(n);
        m.g_amb = alloc_vec(n + EXTRA);
        m.t_vector = alloc_vec(N);
        m.p = (int *)calloc(N, sizeof(int));
        if (!m.p) fatal("memory allocation error\n");
        m.a = alloc_vec(N);
        m.inva = alloc_vec(N);
        m.b = alloc_mat<double>(N, N);
        m.c = alloc_mat<double>(N, N);
        m.lu = alloc_mat<double>(N, N);
        m.flp = placeholder;
        h->rc.block = &h->pub;
        h->rc.type = BLOCK_MODEL;
        h->rc.config = &h->pub.config;
        return h;
    } catch (...) {
        block_model_t &m = h->pub;
        free_mat(m.border), free_mat(m.len), free_mat(m.g), free_mat(m.b), free_mat(m.c), free_mat(m.lu);
        for (double *v : {m.gx, m.gy, m.gx_int, m.gy_int, m.gx_sp, m.gy_sp, m.gx_hs, m.gy_hs, m.g_amb,
                          m.t_vector, m.a, m.inva})
            free(v);
        free(m.p);
        delete h->impl;
        free(h);
        throw;
    }
}

RC_model_t *block_rc_alloc(thermal_config_t *config, flp_t *placeholder) {
    return &block_new(config, placeholder)->rc;
}

void block_steady(block_model_t *model, double *power, double *temp) {
    BlockModel &B = *block_of(model);
    if (!model->r_ready) fatal("R model not ready\n");
    internal_power(*model, power);
    blk_dev_steady(B, power, temp);
}

}  // namespace hs

using namespace hs;

extern "C" {

block_model_t *alloc_block_model(thermal_config_t *config, flp_t *placeholder) {
    HS_BEGIN
    return &block_new(config, placeholder)->pub;
    HS_END(nullptr)
}

void delete_block_model(block_model_t *model) {
    HS_BEGIN
    if (!model) return;
    BlockModel *B = block_of(model);
    blk_dev_release(*B);
    block_model_t &m = *model;
    free_mat(m.border), free_mat(m.len), free_mat(m.g), free_mat(m.b), free_mat(m.c), free_mat(m.lu);
    for (double *v : {m.gx, m.gy, m.gx_int, m.gy_int, m.gx_sp, m.gy_sp, m.gx_hs, m.gy_hs, m.g_amb,
                      m.t_vector, m.a, m.inva})
        free(v);
    free(m.p);
    delete B;
    free(reinterpret_cast<char *>(model) - offsetof(BlockHandle, pub));
    HS_END()
}

// ---- conductances (temperature_block.c:64-296) --------------------------------
void populate_R_model_block(block_model_t *model, flp_t *flp) {
    HS_BEGIN
    BlockModel &B = *block_of(model);
    block_model_t &m = *model;
    const thermal_config_t &c = m.config;
    const double w_chip = total_width(flp), l_chip = total_height(flp);
    if (w_chip > c.s_sink || l_chip > c.s_sink || w_chip > c.s_spreader || l_chip > c.s_spreader)
        fatal("inordinate floorplan size!\n");
    check_floorplan(m, flp, "mismatch between the floorplan and the thermal model\n");
    const int n = flp->n_units, N = m.n_nodes;
    std::vector<Box> box(n);
    for (int i = 0; i < n; i++) box[i] = box_of(flp->units[i]);

    // centre-to-edge lateral conductances of every unit in each node layer:
    // across half the width (x) and half the height (y)
    struct Lat {
        double *gx, *gy, k, t;
    } lat[NL] = {{m.gx, m.gy, c.k_chip, c.t_chip},
                 {m.gx_int, m.gy_int, c.k_interface, c.t_interface},
                 {m.gx_sp, m.gy_sp, c.k_spreader, c.t_spreader},
                 {m.gx_hs, m.gy_hs, c.k_sink, c.t_sink}};
    for (int i = 0; i < n; i++) {
        const unit_t &u = flp->units[i];
        for (int l = 0; l < NL; l++) {
            if (l == 0 && c.block_omit_lateral) {
                lat[l].gx[i] = lat[l].gy[i] = 0;
                continue;
            }
            lat[l].gx[i] = 1.0 / getr(lat[l].k, u.width / 2.0, u.height * lat[l].t);
            lat[l].gy[i] = 1.0 / getr(lat[l].k, u.height / 2.0, u.width * lat[l].t);
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) m.len[i][j] = m.len[j][i] = common_edge(box[i], box[j], i == j);

    package_R(m.pack, c, w_chip, l_chip);
    const package_RC_t &pk = m.pack;

    // units on the chip's west/east/north/south edges (border columns 0-3) and
    // the parallel sums of their spreader and sink conductances to that edge
    double sp_edge[4] = {0, 0, 0, 0}, hs_edge[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        const Box &b = box[i];
        const bool on[4] = {eq(b.x1, 0), eq(b.x2, w_chip), eq(b.y2, l_chip), eq(b.y1, 0)};
        for (int e = 0; e < 4; e++) {
            m.border[i][e] = on[e] ? 1 : 0;
            if (!on[e]) continue;
            sp_edge[e] += e < 2 ? m.gx_sp[i] : m.gy_sp[i];
            hs_edge[e] += e < 2 ? m.gx_hs[i] : m.gy_hs[i];
        }
    }

    // g[a][b]: conductance of node a's half of the a-b link; the link's
    // conductance is the series combination of both halves (links given
    // whole are entered as 2x on both sides)
    std::fill(m.g[0], m.g[0] + (size_t)N * N, 0.0);
    std::fill(m.g_amb, m.g_amb + n + EXTRA, 0.0);
    double **g = m.g;
    auto link = [&](int a, int b, double v) { g[a][b] = g[b][a] = v; };
    const int X = NL * n;   // first package node
    for (int i = 0; i < n; i++) {
        const unit_t &u = flp->units[i];
        const double area = u.height * u.width;
        for (int j = 0; j < n; j++) {
            double part[NL] = {0, 0, 0, 0};
            if (side_by_side(box[i], box[j])) {
                for (int l = 0; l < NL; l++) part[l] = lat[l].gx[i] / u.height;
            } else if (stacked(box[i], box[j])) {
                for (int l = 0; l < NL; l++) part[l] = lat[l].gy[i] / u.width;
            }
            for (int l = 0; l < NL; l++) g[l * n + i][l * n + j] = part[l] * m.len[i][j];
        }
        // vertical links through the silicon, interface and spreader
        link(i, IFACE * n + i, 2.0 / getr(c.k_chip, c.t_chip, area));
        link(IFACE * n + i, HSP * n + i, 2.0 / getr(c.k_interface, c.t_interface, area));
        link(HSP * n + i, HSINK * n + i, 2.0 / getr(c.k_spreader, c.t_spreader, area));
        // sink to ambient: r_convec shared in proportion to the area
        const double r_amb = c.r_convec * (c.s_sink * c.s_sink) / area;
        m.g_amb[i] = 1.0 / (getr(c.k_sink, c.t_sink, area) + r_amb);
        // edge units to the spreader's and the sink's peripheral nodes
        const int *bd = m.border[i];
        link(HSP * n + i, X + SP_N, 2.0 * bd[2] / ((1.0 / m.gy_sp[i]) + pk.r_sp1_y * sp_edge[2] / m.gy_sp[i]));
        link(HSP * n + i, X + SP_S, 2.0 * bd[3] / ((1.0 / m.gy_sp[i]) + pk.r_sp1_y * sp_edge[3] / m.gy_sp[i]));
        link(HSP * n + i, X + SP_E, 2.0 * bd[1] / ((1.0 / m.gx_sp[i]) + pk.r_sp1_x * sp_edge[1] / m.gx_sp[i]));
        link(HSP * n + i, X + SP_W, 2.0 * bd[0] / ((1.0 / m.gx_sp[i]) + pk.r_sp1_x * sp_edge[0] / m.gx_sp[i]));
        link(HSINK * n + i, X + SINK_C_N,
             2.0 * bd[2] / ((1.0 / m.gy_hs[i]) + pk.r_hs1_y * hs_edge[2] / m.gy_hs[i]));
        link(HSINK * n + i, X + SINK_C_S,
             2.0 * bd[3] / ((1.0 / m.gy_hs[i]) + pk.r_hs1_y * hs_edge[3] / m.gy_hs[i]));
        link(HSINK * n + i, X + SINK_C_E,
             2.0 * bd[1] / ((1.0 / m.gx_hs[i]) + pk.r_hs1_x * hs_edge[1] / m.gx_hs[i]));
        link(HSINK * n + i, X + SINK_C_W,
             2.0 * bd[0] / ((1.0 / m.gx_hs[i]) + pk.r_hs1_x * hs_edge[0] / m.gx_hs[i]));
    }
    // peripheral spreader nodes down to the sink's inner ring, inner ring out
    // to the outer ring, both rings to ambient
    link(X + SP_N, X + SINK_C_N, 2.0 / pk.r_sp_per_y);
    link(X + SP_S, X + SINK_C_S, 2.0 / pk.r_sp_per_y);
    link(X + SP_E, X + SINK_C_E, 2.0 / pk.r_sp_per_x);
    link(X + SP_W, X + SINK_C_W, 2.0 / pk.r_sp_per_x);
    link(X + SINK_C_N, X + SINK_N, 2.0 / (pk.r_hs + pk.r_hs2_y));
    link(X + SINK_C_S, X + SINK_S, 2.0 / (pk.r_hs + pk.r_hs2_y));
    link(X + SINK_C_E, X + SINK_E, 2.0 / (pk.r_hs + pk.r_hs2_x));
    link(X + SINK_C_W, X + SINK_W, 2.0 / (pk.r_hs + pk.r_hs2_x));
    m.g_amb[n + SINK_C_N] = m.g_amb[n + SINK_C_S] = 1.0 / (pk.r_hs_c_per_y + pk.r_amb_c_per_y);
    m.g_amb[n + SINK_C_E] = m.g_amb[n + SINK_C_W] = 1.0 / (pk.r_hs_c_per_x + pk.r_amb_c_per_x);
    m.g_amb[n + SINK_N] = m.g_amb[n + SINK_S] = m.g_amb[n + SINK_E] = m.g_amb[n + SINK_W] =
        1.0 / (pk.r_hs_per + pk.r_amb_per);

    // B: off-diagonal -(series conductance) of each link, diagonal the sum of
    // the node's links plus its ambient conductance (sink-layer and sink
    // ring nodes)
    double **b = m.b;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < i; j++)
            b[i][j] = b[j][i] = (g[i][j] == 0.0 || g[j][i] == 0.0) ? 0.0 : -1.0 / ((1.0 / g[i][j]) + (1.0 / g[j][i]));
    for (int i = 0; i < N; i++) {
        double d = (i >= HSINK * n && i < X) ? m.g_amb[i % n] : (i >= X + SINK_C_W) ? m.g_amb[n + i - X] : 0.0;
        for (int j = 0; j < N; j++)
            if (j != i) d -= b[i][j];
        b[i][i] = d;
    }
    m.flp = flp;
    m.r_ready = 1;
    B.lu_stale = true;
    // factorised here, as the reference does, where a GPU is visible (a
    // singular B stops populate_R); elsewhere at the first steady solve
    if (blk_dev_present()) blk_dev_factor(B);
    HS_END()
}

// ---- capacitances (temperature_block.c:303-376) --------------------------------
void populate_C_model_block(block_model_t *model, flp_t *flp) {
    HS_BEGIN
    BlockModel &B = *block_of(model);
    block_model_t &m = *model;
    const thermal_config_t &c = m.config;
    if (!m.r_ready) fatal("R model not ready\n");
    check_floorplan(m, flp, "different floorplans for R and C models!\n");
    const int n = flp->n_units, N = m.n_nodes, X = NL * n;
    package_C(m.pack, c, total_width(flp), total_height(flp));
    const package_RC_t &pk = m.pack;
    for (int i = 0; i < n; i++) {
        const double area = flp->units[i].height * flp->units[i].width;
        m.a[i] = getcap(c.p_chip, c.t_chip, area);
        m.a[IFACE * n + i] = getcap(c.p_interface, c.t_interface, area);
        m.a[HSP * n + i] = getcap(c.p_spreader, c.t_spreader, area);
        // the sink's share of c_convec, in proportion to the area
        const double c_amb = kCFactor * c.c_convec / (c.s_sink * c.s_sink) * area;
        m.a[HSINK * n + i] = getcap(c.p_sink, c.t_sink, area) + c_amb;
    }
    m.a[X + SP_N] = m.a[X + SP_S] = pk.c_sp_per_y;
    m.a[X + SP_E] = m.a[X + SP_W] = pk.c_sp_per_x;
    m.a[X + SINK_C_N] = m.a[X + SINK_C_S] = pk.c_hs_c_per_y + pk.c_amb_c_per_y;
    m.a[X + SINK_C_E] = m.a[X + SINK_C_W] = pk.c_hs_c_per_x + pk.c_amb_c_per_x;
    m.a[X + SINK_N] = m.a[X + SINK_S] = m.a[X + SINK_E] = m.a[X + SINK_W] = pk.c_hs_per + pk.c_amb_per;
    for (int i = 0; i < N; i++) m.inva[i] = 1.0 / m.a[i];
    // the transient system dT/dt = A^-1 P - C T with C = A^-1 B
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) m.c[i][j] = m.inva[i] * m.b[i][j];
    m.c_ready = 1;
    B.c_stale = true;
    HS_END()
}

// ---- solves (temperature_block.c:398-486) ---------------------------------------
void steady_state_temp_block(block_model_t *model, double *power, double *temp) {
    HS_BEGIN
    block_steady(model, power, temp);
    HS_END()
}

void compute_temp_block(block_model_t *model, double *power, double *temp, double time_elapsed) {
    HS_BEGIN
    BlockModel &B = *block_of(model);
    if (!model->r_ready || !model->c_ready) fatal("block model not ready\n");
    if (temp == model->t_vector) fatal("output same as scratch pad\n");
    if (!temp) fatal("the block model's compute_temp needs a temperature vector\n");
    internal_power(*model, power);
    blk_dev_transient(B, power, temp, time_elapsed);
    HS_END()
}

// ---- vectors (temperature_block.c:489-530, :577-580) -----------------------------
double *hotspot_vector_block(block_model_t *model) {
    HS_BEGIN
    return alloc_vec(model->n_nodes);
    HS_END(nullptr)
}

void trim_hotspot_vector_block(block_model_t *model, double *dst, double *src, int at, int size) {
    HS_BEGIN
    const int N = model->n_nodes;
    for (int i = 0; i < at && i < N; i++) dst[i] = src[i];
    for (int i = at + size; i < N; i++) dst[i - size] = src[i];
    HS_END()
}

void resize_thermal_model_block(block_model_t *model, int n_units) {
    HS_BEGIN
    BlockModel &B = *block_of(model);
    if (n_units > model->base_n_units) fatal("resizing block model to more than the allocated space\n");
    model->n_units = n_units;
    model->n_nodes = NL * n_units + EXTRA;
    restride(model->len, n_units, n_units);
    for (double **mat : {model->g, model->b, model->c, model->lu}) restride(mat, model->n_nodes, model->n_nodes);
    B.lu_stale = B.c_stale = true;
    HS_END()
}

void set_temp_block(block_model_t *model, double *temp, double val) {
    HS_BEGIN
    for (int i = 0; i < model->n_nodes; i++) temp[i] = val;
    HS_END()
}

void copy_temp_block(block_model_t *model, double *dst, double *src) {
    HS_BEGIN
    memcpy(dst, src, sizeof(double) * (NL * model->flp->n_units + EXTRA));
    HS_END()
}

// ---- text formats (temperature_block.c:533-764) --------------------------------
static const char *const kLayerPrefix[NL] = {"", "iface_", "hsp_", "hsink_"};

void dump_temp_block(block_model_t *model, double *temp, char *file) {
    HS_BEGIN
    const flp_t *flp = model->flp;
    const int n = flp->n_units;
    FILE *fp = open_out(file);
    for (int l = 0; l < NL; l++)
        for (int i = 0; i < n; i++) fprintf(fp, "%s%s\t%.2f\n", kLayerPrefix[l], flp->units[i].name, temp[l * n + i]);
    for (int i = 0; i < EXTRA; i++) fprintf(fp, "inode_%d\t%.2f\n", i, temp[NL * n + i]);
    close_out(fp);
    HS_END()
}

void read_temp_block(block_model_t *model, double *temp, char *file, int clip) {
    HS_BEGIN
    flp_t *flp = model->flp;
    const int n = flp->n_units;
    const auto lines = read_lines(file);
    size_t li = 0;
    // next line that is neither blank nor a comment
    auto next = [&]() -> const std::string & {
        std::string tok;
        while (li < lines.size())
            if (first_token(lines[li++].c_str(), tok)) return lines[li - 1];
        fatal("not enough lines in temperature file\n");
    };
    char name[STR_SIZE];
    double val, mx = 0;
    for (int l = 0; l < NL; l++) {
        const std::string fmt = std::string(kLayerPrefix[l]) + "%511s%lf";
        for (int i = 0; i < n; i++) {
            if (sscanf(next().c_str(), fmt.c_str(), name, &val) != 2) fatal("invalid temperature file format\n");
            const int idx = get_blk_index(flp, name);
            if (idx < 0) fatal("unit in temperature file not found in floorplan\n");
            temp[idx + l * n] = val;
            if (l == 0 && temp[idx] > mx) mx = temp[idx];
        }
    }
    for (int i = 0; i < EXTRA; i++) {
        if (sscanf(next().c_str(), "%511s%lf", name, &val) != 2) fatal("invalid temperature file format\n");
        char want[32];
        snprintf(want, sizeof want, "inode_%d", i);
        if (strcasecmp(want, name)) fatal("invalid temperature file format\n");
        temp[i + NL * n] = val;
    }
    if (li < lines.size()) fatal("too many lines in temperature file\n");
    const double thr = model->config.thermal_threshold, amb = model->config.ambient;
    if (clip && mx > thr) {
        // scale every rise over ambient so that the hottest unit sits at the threshold
        const double factor = (thr - amb) / (mx - amb);
        for (int i = 0; i < NL * n + EXTRA; i++) temp[i] = (temp[i] - amb) * factor + amb;
    }
    HS_END()
}

void dump_power_block(block_model_t *model, double *power, char *file) {
    HS_BEGIN
    const flp_t *flp = model->flp;
    FILE *fp = open_out(file);
    for (int i = 0; i < flp->n_units; i++) fprintf(fp, "%s\t%.6f\n", flp->units[i].name, power[i]);
    close_out(fp);
    HS_END()
}

void read_power_block(block_model_t *model, double *power, char *file) {
    HS_BEGIN
    char name[STR_SIZE];
    double val;
    for (const std::string &l : read_lines(file)) {
        std::string tok;
        if (!first_token(l.c_str(), tok)) continue;
        if (sscanf(l.c_str(), "%511s%lf", name, &val) != 2) fatal("invalid power file format\n");
        const int idx = get_blk_index(model->flp, name);
        if (idx < 0) fatal("unit in power file not found in floorplan\n");
        power[idx] = val;
    }
    HS_END()
}

double find_max_temp_block(block_model_t *model, double *temp) {
    HS_BEGIN
    double mx = 0.0;
    for (int i = 0; i < model->n_units; i++) {
        if (temp[i] < 0) fatal("negative temperature!\n");
        mx = std::max(mx, temp[i]);
    }
    return mx;
    HS_END(0.0)
}

double find_avg_temp_block(block_model_t *model, double *temp) {
    HS_BEGIN
    double sum = 0.0;
    for (int i = 0; i < model->n_units; i++) {
        if (temp[i] < 0) fatal("negative temperature!\n");
        sum += temp[i];
    }
    return sum / model->n_units;
    HS_END(0.0)
}

// area-weighted mean temperature of the sink: the units' sink nodes, then the
// inner ring (W, E, N, S) and the outer ring over the rest of the sink
double calc_sink_temp_block(block_model_t *model, double *temp, thermal_config_t *config) {
    HS_BEGIN
    const flp_t *flp = model->flp;
    const int n = flp->n_units, X = NL * n;
    const double width = total_width(flp), height = total_height(flp);
    const double spr = config->s_spreader * config->s_spreader, sink = config->s_sink * config->s_sink;
    auto at = [&](int k) {
        if (temp[k] < 0) fatal("negative temperature!\n");
        return temp[k];
    };
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += at(HSINK * n + i) * (flp->units[i].width * flp->units[i].height);
    for (int k = SINK_C_W; k <= SINK_C_E; k++)
        sum += at(X + k) * 0.25 * (config->s_spreader + height) * (config->s_spreader - width);
    for (int k = SINK_C_N; k <= SINK_C_S; k++)
        sum += at(X + k) * 0.25 * (config->s_spreader + width) * (config->s_spreader - height);
    for (int k = SINK_W; k <= SINK_S; k++) sum += at(X + k) * 0.25 * (sink - spr);
    return sum / sink;
    HS_END(0.0)
}

}  // extern "C"
