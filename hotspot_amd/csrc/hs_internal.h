// Internal declarations shared by the host model (hs_host.cpp) and the
// device engine (hs_device.hip).  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hotspot_amd.h"

namespace hs {

// ---- constants of the grid model (temperature.h:29-119, util.h:18-19) ----
constexpr double kDelta = 1.0e-6;
constexpr double kLargeNum = 1.0e100;
constexpr double kCFactor = 0.333;
constexpr double kMinStep = 1e-7;
constexpr double kOccupancyThreshold = 0.95;
constexpr int kMaxLayers = 32;
constexpr int kMaxExtra = EXTRA + EXTRA_SEC;

enum ExtraNode {
    SP_W = 0, SP_E, SP_N, SP_S, SINK_C_W, SINK_C_E, SINK_C_N, SINK_C_S,
    SINK_W, SINK_E, SINK_N, SINK_S,
    SUB_W, SUB_E, SUB_N, SUB_S, SOLDER_W, SOLDER_E, SOLDER_N, SOLDER_S,
    PCB_C_W, PCB_C_E, PCB_C_N, PCB_C_S, PCB_W, PCB_E, PCB_N, PCB_S
};
enum MapMode { GRID_AVG = 0, GRID_MIN = 1, GRID_MAX = 2, GRID_CENTER = 3 };
enum CellType { CELL_TSV = -1, CELL_SOLID = 0, CELL_FLUID = 1, CELL_INLET = 2, CELL_OUTLET = 3 };

// error raised inside the library; converted at the C boundary
struct Error {
    std::string msg;
};
[[noreturn]] void fatal(const std::string &msg);
void warning(const std::string &msg);
// at the C boundary: print and exit (the reference's fatal) or record for
// hs_last_error, by hs_set_fatal_mode
void report(const std::string &msg);

// body of a C entry point: errors raised inside become report()
#define HS_BEGIN try {
#define HS_END(ret)                                   \
    }                                                 \
    catch (const hs::Error &e) {                      \
        hs::report(e.msg);                            \
        return ret;                                   \
    }                                                 \
    catch (const std::bad_alloc &) {                  \
        hs::report("memory allocation error\n");      \
        return ret;                                   \
    }                                                 \
    catch (const std::exception &e) {                 \
        hs::report(std::string(e.what()) + "\n");     \
        return ret;                                   \
    }

// ---- small numerics and host helpers (util.c:17-20, RCutil.c:20-31; hs_host.cpp)
inline bool eq(double x, double y) { return (x - y < 0 ? y - x : x - y) < kDelta; }
inline double getr(double k, double t, double area) { return t / (k * area); }
inline double getcap(double sp, double t, double area) { return kCFactor * sp * t * area; }
// first whitespace-delimited token of a line; false for blank/comment lines
bool first_token(const char *line, std::string &tok);
std::vector<std::string> read_lines(const char *file);
double total_width(const flp_t *f);    // get_total_width, flp.c:1501-1515
double total_height(const flp_t *f);   // get_total_height, flp.c:1517-1531
// populate_package_R / populate_package_C (temperature.c:458-569)
void package_R(package_RC_t &p, const thermal_config_t &c, double w, double h);
void package_C(package_RC_t &p, const thermal_config_t &c, double w, double h);

// ---- block<->grid map of one distinct floorplan ---------------------------
struct BlockMap {
    // per cell (row-major i*nc+j): the blocks overlapping it, in block order
    std::vector<int> cell_ptr;        // nr*nc + 1
    std::vector<int> ent_unit;
    std::vector<double> ent_occ;
    std::vector<glist_t> ranges;      // per block: [i1,i2) x [j1,j2)
    // detailed-3D values of each cell's list head (blist_t fields)
    std::vector<double> rx, ry, rz, cap;
    std::vector<uint8_t> lock, has_res;
};

// per-layer microchannel state (flow field of one network)
struct Microchannel {
    microchannel_config_t cfg;        // owned copy (cell_types etc. allocated)
    std::vector<int> types;           // nr*nc cell types
    std::vector<double> pressure;     // nr*nc (0 for solid cells); empty until the first device build
    double hydro_c = 0.0;             // hydraulic conductance of one cell
};

// ---- the host model ------------------------------------------------------
struct DevModel;  // device state, defined in hs_device.hip

struct Model {
    grid_model_t *g = nullptr;        // the public struct
    std::vector<layer_t> layers;
    std::vector<int> layer_map;       // layer -> index into maps
    std::vector<BlockMap> maps;
    std::vector<int> map_owner;       // map -> layer whose floorplan/props built it
    std::vector<uint8_t> map_pkg;     // map belongs to package layers (hasRes cleared)
    std::vector<flp_t *> owned_flps;
    std::vector<Microchannel *> mc;   // per layer (nullptr if not microchannel)
    int nx = EXTRA;
    int spidx = -1, hsidx = -1, subidx = -1, solderidx = -1, pcbidx = -1;
    bool detailed = false;
    // coolant advection at outlet cells as the reference's SuperLU build
    // assembles it (only the first fluid neighbour, N, S, W, E)
    bool superlu_outlets = false;
    hs_options_t opt{};
    hs_stats_t stats{};
    DevModel *dev = nullptr;
    bool dev_dirty = true;            // host coefficients changed since upload
    unsigned long stamp = 0;          // process-unique, renewed whenever the coefficients change
                                      // (an engine keeps a model's device state while it matches)
    int steady_valid = 0;             // last_steady holds a solution
    int trans_valid = 0;              // last_trans holds a state
    int mirror_steady = 0, mirror_trans = 0;  // host mirrors up to date

    int nr() const { return g->rows; }
    int nc() const { return g->cols; }
    int nl() const { return g->n_layers; }
    long ncell() const { return (long)g->n_layers * g->rows * g->cols; }
    long vec_len() const { return ncell() + nx; }
};

// handle layout: what alloc_RC_model returns points at .rc
struct Handle {
    RC_model_t rc;
    grid_model_t pub;
    Model *impl;
};
Model *model_of(RC_model_t *m);
Model *model_of(grid_model_t *g);

// host helpers used by the device side
void layer_bases(const Model &M, std::vector<int> &base);   // block-vector offset per layer
// The temperature-leakage fixed point of the grid model (temperature.c:
// 708-746), one step at a time: leakage_step writes power + the leakage of
// each power-layer block at temp into power_new and copies temp into
// temp_old; leakage_rise is the largest rise of a power-layer block since
// temp_old.  The loop stops when the rise is below kLeakTol, and fails above
// kLeakMaxIter steps or when a rise after the first exceeds kLeakTempHigh.
constexpr double kLeakTol = 0.01, kLeakTempHigh = 500.0;
constexpr int kLeakMaxIter = 100;
struct LeakBlocks {      // the power-layer blocks: block-vector index, height, width
    int mode = 0;        // leakage_mode
    std::vector<int> idx;
    std::vector<double> h, w;
};
LeakBlocks leak_blocks(const Model &M);
void leakage_step(const LeakBlocks &b, const double *power, const double *temp, double *power_new,
                  double *temp_old);
double leakage_rise(const LeakBlocks &b, const double *temp, const double *temp_old);

// The natural-convection fixed point over the steady solve (hotspot.c:
// 667-682): what it reads of a model, copied so that a caller may delete the
// model once it has handed it over.  sink_temp is calc_sink_temp's grid
// branch (temperature.c:993-1031) on a block vector; sink_resistances gives
// the sink layer's rz (populate_R_model_grid, temperature_grid.c:1042-1059)
// and the package resistances (populate_package_R) at r_convec r.
// calc_natural_convec's sink temperature is the average plus
// kSmallForConvec; the loop stops when r_convec moves by less than
// kNaturalConvecTol (package.h:31-33) and fails above kMaxSinkTemp.
constexpr double kSmallForConvec = 0.01, kNaturalConvecTol = 0.01, kMaxSinkTemp = 1000.0;
struct SinkModel {
    thermal_config_t c{};              // the configuration (r_convec, ambient, s_sink, s_spreader, ...)
    double width = 0, height = 0;      // of the grid
    int rows = 0, cols = 0;
    int hsidx = 0;                     // the sink layer
    double k = 0, thickness = 0;       // its conductivity and thickness
    int base = 0;                      // its blocks' offset in the block vector
    std::vector<double> area;          // their width x height
    int total_n_blocks = 0;
};
SinkModel sink_model(const Model &M);
double sink_temp(const SinkModel &s, const double *temp);
void sink_resistances(const SinkModel &s, double r, double &rz, package_RC_t &pk);

// ---- device engine entry points (hs_device.hip) ---------------------------
void dev_release(Model &M);
// steady: block power in (package entries already zero), block temps out
void dev_steady(Model &M, const double *block_power, double *block_temp);
// transient: block power; block temp (nullptr = continue from last_trans)
void dev_transient(Model &M, const double *block_power, const double *block_temp_in,
                   double dt, double *block_temp_out);
void dev_transient_run(Model &M, const double *power_seq, int steps, double dt,
                       const double *init_temp, double *temp_seq);
void dev_get_grid(Model &M, int which, double *out);
void dev_set_grid(Model &M, int which, const double *in);
void dev_slope(Model &M, const double *block_power, const double *grid_temp, double *out);
void dev_b2g(Model &M, const double *block_vec, double *grid_vec, int type);
void dev_g2b(Model &M, const double *grid_vec, double *block_vec);

// ---- block model (hs_block.cpp host side, hs_block.hip device side) -------
struct BlockDev;  // device state, defined in hs_block.hip

struct BlockModel {
    block_model_t *m = nullptr;       // the public struct
    BlockDev *dev = nullptr;
    bool lu_stale = true;             // b changed since the device factorisation
    bool c_stale = true;              // c / inva changed since their upload
    hs_stats_t stats{};
};

// handle layout: alloc_RC_model returns .rc, alloc_block_model returns .pub
struct BlockHandle {
    RC_model_t rc;
    block_model_t pub;
    BlockModel *impl;
};
BlockModel *block_of(block_model_t *b);
RC_model_t *block_rc_alloc(thermal_config_t *config, flp_t *placeholder);
// steady_state_temp_block without the C-boundary error conversion
void block_steady(block_model_t *model, double *power, double *temp);

// device entry points (hs_block.hip)
bool blk_dev_present();                    // a GPU is visible
void blk_dev_release(BlockModel &B);
// LU factors of b on the device (lupdcmp); host copies into lu and p
void blk_dev_factor(BlockModel &B);
// B T = P by the LU factors (refactorised on the device when b changed)
void blk_dev_steady(BlockModel &B, const double *power, double *temp);
// compute_temp_block's adaptive RK4 over dt, temp in and out
void blk_dev_transient(BlockModel &B, const double *power, double *temp, double dt);

// batched engine (hs_batch.inl)
struct Engine;
Engine *eng_create(int precision, int transient);
void eng_destroy(Engine *E);
void eng_set_options(Engine &E, const hs_options_t &o);
void eng_set_tick(Engine &E, int intervals, int iters);
void eng_set_budget(Engine &E, double bytes);
int eng_submit(Engine &E, Model &M, const double *power, int rows, const double *init, int nsteps, double dt);
int eng_submit_package(Engine &E, Model &M, const double *power, const package_config_t &pkg);
int eng_admit(Engine &E);
int eng_tick(Engine &E);
int eng_poll(Engine &E, int id, double *out);
int eng_result_seq(Engine &E, int id, double *out);
void eng_job_info(Engine &E, int id, hs_job_info_t *info);
void eng_release(Engine &E, int id);
int eng_active(Engine &E);
void eng_get_stats(Engine &E, hs_engine_stats_t *st);

}  // namespace hs
