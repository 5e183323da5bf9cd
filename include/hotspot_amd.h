/*
 * hotspot_amd.h -- C ABI of the MI355X grid thermal model.
 *
 * Drop-in for HotSpot's thermal models (grid and block): the entry points keep the names,
 * signatures and argument meaning of the reference library, and the structs
 * keep the reference *implementation* layouts (temperature.h, flp.h,
 * temperature_grid.h, microchannel.h, materials.h), so code compiled against
 * the reference headers that reads model->grid->n_layers,
 * layers[i].has_power, layers[i].flp->n_units or has_lcf
 * (hotspot.c:546-549, :578-587, :720) works unchanged.  The stale copy of
 * these structs in the reference's hotspot-iface.h (no per-unit detailed-3D
 * fields in unit_t, grid_transient_file placed after steady_file) is *not*
 * followed; the offsets are pinned by static asserts in hs_host.cpp.
 *
 * Errors: the reference prints "error: ..." and calls exit(1) (util.c:32-36).
 * This library does the same by default; hs_set_fatal_mode(0) makes it
 * record the error (hs_last_error) and return instead.
 *
 * Everything below "extensions" has no counterpart in the reference.
 */
#ifndef HOTSPOT_AMD_H
#define HOTSPOT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STR_SIZE 512
#define MAX_UNITS 8192
#define EXTRA 12
#define EXTRA_SEC 16
#define BLOCK_MODEL 0
#define GRID_MODEL 1
#define V_POWER 0
#define V_TEMP 1

/* ---- floorplan: flp.h:378-400 ---------------------------------------- */
typedef struct unit_t_st {
    char name[STR_SIZE];
    double width;
    double height;
    double leftx;
    double bottomy;
    double specificheat;   /* detailed 3D: per-unit volumetric heat capacity */
    double resistivity;    /* detailed 3D: per-unit resistivity */
    int hasRes;
    int hasSh;
} unit_t;

typedef struct flp_t_st {
    unit_t *units;
    int n_units;
    double **wire_density;
} flp_t;

/* ---- configuration: temperature.h:141-221 ------------------------------ */
typedef struct thermal_config_t_st {
    double t_chip, k_chip, p_chip, thermal_threshold;
    double c_convec, r_convec, s_sink, t_sink, k_sink, p_sink;
    double s_spreader, t_spreader, k_spreader, p_spreader;
    double t_interface, k_interface, p_interface;
    int model_secondary;
    double r_convec_sec, c_convec_sec;
    int n_metal;
    double t_metal, t_c4, s_c4;
    int n_c4;
    double s_sub, t_sub, s_solder, t_solder, s_pcb, t_pcb;
    double ambient;
    char init_file[STR_SIZE];
    double init_temp;
    char steady_file[STR_SIZE];
    double sampling_intvl;
    double base_proc_freq;
    int dtm_used;
    char model_type[STR_SIZE];
    int leakage_used;
    int leakage_mode;
    int package_model_used;
    char package_config_file[STR_SIZE];
    int block_omit_lateral;
    int grid_rows;
    int grid_cols;
    char grid_layer_file[STR_SIZE];
    char grid_steady_file[STR_SIZE];
    char grid_map_mode[STR_SIZE];
    char grid_transient_file[STR_SIZE];
    int detailed_3D_used;
} thermal_config_t;

/* ---- microchannels and materials: microchannel.h:29-90, materials.h ---- */
typedef struct microchannel_config_t_st {
    double cell_width, cell_height, cell_thickness;
    double pumping_pressure, pump_internal_res, inlet_temperature;
    double coolant_capac, coolant_res, coolant_visc;
    double wall_capac, wall_res, htc;
    char network_file[STR_SIZE];
    char floorplan_file[STR_SIZE + 3];
    int num_rows, num_columns, n_fluid_cells;
    int **cell_types;
    int **mapping;
    double **A;
    double *b;
    int nnz;
} microchannel_config_t;

typedef struct material_t_st {
    int material_type;
    double thermal_conductivity;
    double volumetric_heat_capacity;
    double dynamic_viscosity;
} material_t;

typedef struct materials_list_t_st {
    int size;
    char **names;
    material_t *material_properties;
} materials_list_t;

/* ---- package R/C: temperature.h:237-334 (52 doubles) ------------------- */
typedef struct package_RC_t_st {
    double r_sp1_x, r_sp1_y, r_hs1_x, r_hs1_y, r_hs2_x, r_hs2_y, r_hs;
    double r_sp_per_x, r_sp_per_y, r_hs_c_per_x, r_hs_c_per_y, r_hs_per;
    double c_sp_per_x, c_sp_per_y, c_hs_c_per_x, c_hs_c_per_y, c_hs_per;
    double r_amb_c_per_x, c_amb_c_per_x, r_amb_c_per_y, c_amb_c_per_y;
    double r_amb_per, c_amb_per;
    double r_sub1_x, r_sub1_y, r_solder1_x, r_solder1_y;
    double r_pcb1_x, r_pcb1_y, r_pcb2_x, r_pcb2_y, r_pcb;
    double r_sub_per_x, r_sub_per_y, r_solder_per_x, r_solder_per_y;
    double r_pcb_c_per_x, r_pcb_c_per_y, r_pcb_per;
    double c_sub_per_x, c_sub_per_y, c_solder_per_x, c_solder_per_y;
    double c_pcb_c_per_x, c_pcb_c_per_y, c_pcb_per;
    double r_amb_sec_c_per_x, c_amb_sec_c_per_x, r_amb_sec_c_per_y, c_amb_sec_c_per_y;
    double r_amb_sec_per, c_amb_sec_per;
} package_RC_t;

/* ---- grid model: temperature_grid.h:63-191 (SUPERLU=0 layout) ---------- */
typedef struct blist_t_st {
    int idx;
    double occupancy;
    struct blist_t_st *next;
    int lock, hasRes, hasCap;
    double rx, ry, rz;
    double capacitance;
} blist_t;

typedef struct glist_t_st {
    int i1, i2, j1, j2;
} glist_t;

typedef struct layer_t_st {
    flp_t *flp;
    int no;
    int has_lateral;
    int has_power;
    double k;
    double thickness;
    double sp;
    int is_microchannel;
    microchannel_config_t *microchannel_config;
    double rx, ry, rz;
    double c;
    blist_t ***b2gmap;   /* not populated by this library (NULL) */
    glist_t *g2bmap;     /* block -> [i1,i2) x [j1,j2) cell ranges */
} layer_t;

typedef struct grid_model_vector_t_st {
    double ***cuboid;
    double *extra;
} grid_model_vector_t;

typedef struct grid_model_t_st {
    thermal_config_t config;
    layer_t *layers;
    int n_layers;
    int rows;
    int cols;
    double width;
    double height;
    package_RC_t pack;
    int total_n_blocks;
    int map_mode;
    int r_ready;
    int c_ready;
    int has_lcf;
    grid_model_vector_t *last_steady;   /* host mirror, refreshed on demand */
    grid_model_vector_t *last_trans;    /* host mirror, refreshed on demand */
    double *last_temp;
    int base_n_units;
    int use_microchannels;
    microchannel_config_t *default_microchannel_config;
} grid_model_t;

/* ---- block model: temperature_block.h:5-67 ------------------------------
 * Four node layers per floorplan unit (silicon, interface, spreader, sink),
 * then the 12 package nodes: node ids layer*n_units + unit, NL*n_units + k. */
#ifndef NL
#define NL 4
#endif
#ifndef IFACE
#define IFACE 1
#endif
#ifndef HSP
#define HSP 2
#endif
#ifndef HSINK
#define HSINK 3
#endif

typedef struct block_model_t_st {
    flp_t *flp;
    thermal_config_t config;
    double **b;          /* conductance matrix, n_nodes x n_nodes */
    double **lu;         /* its LU factors (device-computed copy) */
    int *p;              /* row permutation of the factorisation */
    double *a;           /* node capacitances */
    double *inva;        /* 1 / a */
    double **c;          /* inva * b */
    package_RC_t pack;
    double *gx, *gy;
    double *gx_int, *gy_int;
    double *gx_sp, *gy_sp;
    double *gx_hs, *gy_hs;
    double *g_amb;
    double *t_vector;
    double **len, **g;
    int **border;
    int n_nodes;
    int n_units;
    int base_n_units;
    int r_ready;
    int c_ready;
} block_model_t;

typedef struct RC_model_t_st {
    union {
        struct block_model_t_st *block;
        struct grid_model_t_st *grid;
    };
    int type;
    thermal_config_t *config;
} RC_model_t;

/* ---- reference entry points ------------------------------------------- */
/* flp.c:1175-1218 */
flp_t *read_flp(char *file, int read_connects, int initialize_connects);
/* flp.c:1258 */
void free_flp(flp_t *flp, int compacted, int free_connects);
/* flp.c:1385-1402 */
int get_blk_index(flp_t *flp, char *name);
/* temperature.c:19-109 */
thermal_config_t default_thermal_config(void);
/* microchannel.c:17-43 */
microchannel_config_t default_microchannel_config(void);
/* materials.c:13-15 */
void default_materials(materials_list_t *materials_list);
/* materials.c:17-155 */
void materials_add_from_file(materials_list_t *materials_list, char *materials_filename);
/* materials.c:152-161 */
void free_materials(materials_list_t *materials_list);
/* temperature.c:616-633 (model_type "grid" or "block") */
RC_model_t *alloc_RC_model(thermal_config_t *config, flp_t *placeholder,
                           microchannel_config_t *microchannel_config,
                           materials_list_t *materials_list, int do_detailed_3D,
                           int use_microchannels);
/* temperature.c:930-938 */
void delete_RC_model(RC_model_t *model);
/* temperature.c:636-643 -> temperature_grid.c:950-1063 */
void populate_R_model(RC_model_t *model, flp_t *flp);
/* temperature.c:646-653 -> temperature_grid.c:1065-1114 */
void populate_C_model(RC_model_t *model, flp_t *flp);
/* temperature.c:764-772 -> temperature_grid.c:1182-1194 */
double *hotspot_vector(RC_model_t *model);
/* util.c:89 */
void free_dvector(double *v);
/* temperature.c:799-806 -> temperature_grid.c:1245-1259 */
void set_temp(RC_model_t *model, double *temp, double val);
/* temperature.c:830-837 -> temperature_grid.c:1449-1455 */
void copy_temp(RC_model_t *model, double *dst, double *src);
/* temperature.c:809-816 -> temperature_grid.c:1322-1447 */
void dump_temp(RC_model_t *model, double *temp, char *file);
/* temperature.c:844-851 -> temperature_grid.c:1462-1642 */
void read_temp(RC_model_t *model, double *temp, char *file, int clip);
/* temperature.c:854-861 -> temperature_grid.c:1645-1682 */
void dump_power(RC_model_t *model, double *power, char *file);
/* temperature.c:867-874 -> temperature_grid.c:1684-1764 */
void read_power(RC_model_t *model, double *power, char *file);
/* temperature.c:877-885 -> temperature_grid.c:1766-1790 */
double find_max_temp(RC_model_t *model, double *temp);
/* temperature.c:888-896 -> temperature_grid.c:1792-1813 */
double find_avg_temp(RC_model_t *model, double *temp);
/* temperature.c:819-827 -> temperature_grid.c:1816-1856 */
double calc_sink_temp(RC_model_t *model, double *temp);
/* temperature.c:656-751 (grid branch :707-749, leakage loop included) */
void steady_state_temp(RC_model_t *model, double *power, double *temp);
/* temperature.c:754-761 -> temperature_grid.c:3603-3704 */
void compute_temp(RC_model_t *model, double *power, double *temp, double time_elapsed);
/* temperature.c:778-787 -> temperature_grid.c:1203-1220 */
void trim_hotspot_vector(RC_model_t *model, double *dst, double *src, int at, int size);
/* temperature.c:789-797 -> temperature_grid.c:1223-1242 */
void resize_thermal_model(RC_model_t *model, int n_units);
/* temperature.c:910-927 */
double calc_leakage(int mode, double h, double w, double temp);
/* temperature_grid.c:1262-1295 */
void dump_steady_temp_grid(grid_model_t *model, char *file);

/* ---- block model (temperature_block.c), also reached through the RC_model_t
 * entry points above when model_type is "block".  The factorisation, the
 * steady solve and the transient integration run on the device. */
/* temperature_block.c:19-56 */
block_model_t *alloc_block_model(thermal_config_t *config, flp_t *placeholder);
/* temperature_block.c:833-856 */
void delete_block_model(block_model_t *model);
/* temperature_block.c:64-296 */
void populate_R_model_block(block_model_t *model, flp_t *flp);
/* temperature_block.c:303-376 */
void populate_C_model_block(block_model_t *model, flp_t *flp);
/* temperature_block.c:398-411 */
void steady_state_temp_block(block_model_t *model, double *power, double *temp);
/* temperature_block.c:449-486 */
void compute_temp_block(block_model_t *model, double *power, double *temp, double time_elapsed);
/* temperature_block.c:489-492 */
double *hotspot_vector_block(block_model_t *model);
/* temperature_block.c:498-507 */
void trim_hotspot_vector_block(block_model_t *model, double *dst, double *src, int at, int size);
/* temperature_block.c:510-522 */
void resize_thermal_model_block(block_model_t *model, int n_units);
/* temperature_block.c:525-530 */
void set_temp_block(block_model_t *model, double *temp, double val);
/* temperature_block.c:533-575 */
void dump_temp_block(block_model_t *model, double *temp, char *file);
/* temperature_block.c:577-580 */
void copy_temp_block(block_model_t *model, double *dst, double *src);
/* temperature_block.c:587-695 */
void read_temp_block(block_model_t *model, double *temp, char *file, int clip);
/* temperature_block.c:698-720 */
void dump_power_block(block_model_t *model, double *power, char *file);
/* temperature_block.c:726-764 */
void read_power_block(block_model_t *model, double *power, char *file);
/* temperature_block.c:766-778 */
double find_max_temp_block(block_model_t *model, double *temp);
/* temperature_block.c:780-792 */
double find_avg_temp_block(block_model_t *model, double *temp);
/* temperature_block.c:795-831 */
double calc_sink_temp_block(block_model_t *model, double *temp, thermal_config_t *config);
/* temperature_grid.c:1297-1319 */
void dump_transient_temp_grid(grid_model_t *model, int trace_num, double sampling_intvl,
                              char *filename);

/* ---- configuration tables: util.h:92-117, util.c:322-439 --------------- */
typedef struct str_pair_st {
    char name[STR_SIZE];
    char value[STR_SIZE];
} str_pair;

/* util.c:322-353: "-name value" lines of a file ('#' comments); returns the count */
int read_str_pairs(str_pair *table, int max_entries, char *file);
/* util.c:359-376: "-name value" pairs of a command line */
int parse_cmdline(str_pair *table, int max_entries, int argc, char **argv);
/* util.c:379-391: writes "<prefix>name\tvalue" lines */
void dump_str_pairs(str_pair *table, int size, char *file, char *prefix);
/* util.c:394-404: case-insensitive lookup, -1 if absent */
int get_str_index(str_pair *table, int size, char *str);
/* util.c:426-439: keeps the first of entries with the same name; returns the new size */
int str_pairs_remove_duplicates(str_pair *table, int size);
/* temperature.c:115-339: overrides config fields named in the table, then validates */
void thermal_config_add_from_strs(thermal_config_t *config, materials_list_t *materials_list,
                                  str_pair *table, int size);
/* temperature.c:345-402: the 51 fields as name/value pairs; returns 51 */
int thermal_config_to_strs(thermal_config_t *config, str_pair *table, int max_entries);
/* microchannel.c:49-108 */
void microchannel_config_add_from_strs(microchannel_config_t *config, materials_list_t *materials_list,
                                       str_pair *table, int size);
/* ---- package model: package.h:46-115, package.c ----------------------- */
typedef struct package_config_t_st {
    int natural_convec;    /* 0: forced convection, 1: natural convection */
    int flow_type;         /* 0: lateral air flow from the sink side, 1: impinging from the top */
    int sink_type;         /* 0: fin-channel sink, 1: pin-fin sink */
    double fin_height, fin_width, channel_width;
    double pin_height, pin_diam, pin_dist;
    double fan_radius, motor_radius;
    int rpm;
} package_config_t;

typedef struct convection_t_st {
    double n_fin, sur_area_fin;
    double n_pin, sur_area_pin;
    double reynolds, nusselt, h_coeff, v;
    double r_th;           /* lumped convection resistance of the sink (K/W) */
} convection_t;

/* package.c:18-51 */
package_config_t default_package_config(void);
/* package.c:54-108 */
void package_config_add_from_strs(package_config_t *config, str_pair *table, int size);
/* package.c:110-140: returns 12 */
int package_config_to_strs(package_config_t *config, str_pair *table, int max_entries);
/* package.c:159-278: forced convection */
void calculate_flow(convection_t *p, package_config_t *config, thermal_config_t *thermal_config);
/* package.c:281-335: natural convection at an average sink temperature */
void calc_natural_convec(convection_t *p, package_config_t *config, thermal_config_t *thermal_config,
                         double sink_temp);
/* package.c:353-397: sets thermal_config->r_convec from the package file named in the
 * table (package_config_file); returns natural_convec */
int package_model(thermal_config_t *thermal_config, str_pair *table, int size, double sink_temp);

/* microchannel.c:114-156: writes 17 name/value pairs and returns 16, as the
 * reference does (the last pair, n_fluid_cells, is not counted) */
int microchannel_config_to_strs(microchannel_config_t *config, str_pair *table, int max_entries);

/* ---- extensions --------------------------------------------------------
 * None of these has a counterpart in the reference.
 */

/* error channel: 0 = no error since the last call; the message is copied. */
int hs_last_error(char *buf, int len);
/* 1 (default): print "error: ..." and exit(1) like the reference; 0: record and return */
void hs_set_fatal_mode(int exit_on_error);

/* precision of the solves */
#define HS_PREC_FP64  0   /* every sweep in FP64 */
#define HS_PREC_MIXED 1   /* FP32 sweeps inside FP64 refinement / FP64 residuals */
/* steady solvers */
#define HS_STEADY_JPCG  0 /* Jacobi-preconditioned CG (G symmetric) */
#define HS_STEADY_MGPCG 1 /* FP32 CG preconditioned by a geometric-multigrid V-cycle (z-line
                           * Jacobi smoothing, coarsening rows and columns), inside FP64
                           * iterative refinement, for either precision option.
                           * Where coolant advection (microchannel layers) makes G
                           * non-symmetric, either option solves by restarted GMRES(32)
                           * cycles inside FP64 refinement, right-preconditioned by the
                           * V-cycle of the upwind form of the advection.  A steady solve
                           * that stops at its iteration limits above max(100 steady_rtol,
                           * 1e-6) relative residual is an error. */
/* transient integrators */
#define HS_TRANS_RK4    0 /* the reference's adaptive step-doubling RK4 */
#define HS_TRANS_TAYLOR 1 /* exponential (Taylor) integrator of the linear ODE, error bounded a priori */
#define HS_TRANS_BE     2 /* the reference's SuperLU-build integrator: one backward-Euler step per
                           * interval, (C/h + G) T' = C T / h + P (RCutil.c:268-428), solved by FP64
                           * Jacobi CG (FP64 GMRES for microchannel models, whose G is
                           * non-symmetric); secondary package nodes without capacitance */

typedef struct hs_options_st {
    int precision;        /* HS_PREC_* */
    int steady_solver;    /* HS_STEADY_* */
    int transient;        /* HS_TRANS_* */
    double steady_rtol;   /* relative residual target of the steady solve (FP64 norm) */
    int max_iters;        /* per steady solve */
    double taylor_tol;    /* bound on the Taylor truncation error per interval (K) */
    int steady_warm_start; /* 1: start a steady solve from the previous solution (0: from ambient) */
} hs_options_t;

typedef struct hs_stats_st {
    int steady_iters;         /* CG (or BiCGStab) iterations of the last steady solve */
    int steady_refinements;   /* FP64 refinement steps (mixed precision) */
    double steady_rel_resid;  /* final ||P + b - G T||_2 / ||P + b||_2 in FP64 */
    int rk4_calls;            /* RK4 step attempts of the last compute_temp */
    int taylor_terms;         /* terms per sub-step of the last Taylor interval */
    int taylor_substeps;
    double last_solve_ms;     /* device time of the last steady/compute_temp call */
    int be_iters;             /* CG iterations of the last backward-Euler interval */
} hs_stats_t;

/* Coolant advection at outlet cells: 0 (default) as the reference's slope
 * (slope_fn_grid, temperature_grid.c:3256-3590) has it, every fluid
 * neighbour's flow plus the outflow at an edge outlet; 1 as the matrix of
 * the reference's SuperLU build (build_transient_grid_matrix,
 * temperature_grid.c:5313-5345) has it, the first fluid neighbour alone in
 * the order N, S, W, E, at every outlet.  The two agree where each outlet
 * has one fluid neighbour (examples/example5); they differ where outlets
 * touch (channels more than one cell wide), whose outflow the SuperLU build
 * takes from an outlet neighbour carrying no flow.  Applies to every solve
 * of the model.  No counterpart in the reference, which picks one by its
 * build (SUPERLU). */
void hs_set_outlet_rule(RC_model_t *model, int superlu);
void hs_get_options(RC_model_t *model, hs_options_t *opt);
void hs_set_options(RC_model_t *model, const hs_options_t *opt);
void hs_get_stats(RC_model_t *model, hs_stats_t *stats);

/* Pins the tiling of the FP32 Taylor sweep for every model of this model's
 * grid shape (layers x rows x columns) set up after the call, in place of
 * the tiling the library times on first use of a shape.  No counterpart in
 * the reference: the tiling changes speed only, never results, and the pin
 * is for reproducible timings and for tests of every tiling.  shift: tile
 * width 1 << shift, 4..8, dividing the grid's columns (0 drops the pin);
 * rows_per_thread 4 or 8; z 0 per-layer tiles, 1 column tiles marched
 * through the layers.  Problems that do not take float4 sweeps (layers of
 * several materials, grids narrower than 16 columns) ignore it.  Not to be
 * called while a solve or an engine tick runs.  Returns 0, or -1 on error
 * (hs_last_error). */
int hs_pin_sweep_tiling(RC_model_t *model, int shift, int rows_per_thread, int z);

/* Grid-level state in the reference's flat layout [n_layers*rows*cols + extras]
 * (A3D, temperature_grid.c:2854).  which: 0 = last steady, 1 = last transient. */
int hs_grid_size(RC_model_t *model);
void hs_get_grid_temp(RC_model_t *model, int which, double *out);
void hs_set_grid_temp(RC_model_t *model, int which, const double *in);
/* one application of the transient right-hand side f(T) = (P + b - G T)/C on
 * the device, P given per block (package-node entries ignored), T and dT in
 * the flat grid layout; for operator-level checks. */
void hs_grid_slope(RC_model_t *model, const double *block_power, const double *grid_temp,
                   double *grid_slope);
/* hs_grid_slope on device buffers, allocating no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, allocation-free serving loops).
 * d_block_power (total blocks + package nodes; package entries ignored),
 * d_grid_temp and d_grid_slope (hs_grid_size doubles each) are device
 * pointers the caller owns.  The descriptor and the cell power map go to a
 * fixed 32 MiB library scratch array per GPU, so grids of more than about
 * 4.19 million nodes are refused, and the model's device state must exist
 * already (built by an earlier call such as hs_grid_slope or a solve), else
 * the call fails.  All calls on one GPU share its scratch: calls that
 * overlap in time on one GPU race.  Returns 0, or -1 on error
 * (hs_last_error). */
int hs_grid_slope_noalloc(RC_model_t *model, const double *d_block_power, const double *d_grid_temp,
                          double *d_grid_slope);
/* block -> grid translation on the device (type V_POWER or V_TEMP) */
void hs_block_to_grid(RC_model_t *model, const double *block_vec, double *grid_vec, int type);
/* grid -> block translation on the device (the model's grid_map_mode) */
void hs_grid_to_block(RC_model_t *model, const double *grid_vec, double *block_vec);

/* Runs n_steps transient intervals of dt seconds on the device without a host
 * round trip per interval: equivalent to n_steps calls
 * compute_temp(model, power_seq + k*n, NULL, dt) with n = total blocks + package
 * nodes (the package-node entries of each power vector are ignored).
 * init_temp (a block vector) starts the state like compute_temp's first call;
 * NULL continues from the current transient state.  The block temperatures of
 * every interval are written to temp_seq (n_steps*n doubles) when non-NULL. */
void hs_transient_run(RC_model_t *model, const double *power_seq, int n_steps, double dt,
                      const double *init_temp, double *temp_seq);

/* Times the launches of one kernel family (e.g. "k_tay_term", "k_rk_stage",
 * "k_cg_dir") with device events, up to max_samples launches; consecutive
 * launches of the family share one pair of events around their run.  NULL
 * stops. */
void hs_kernel_timer(const char *name, int max_samples);
/* Number of timed launches; their mean and total duration in ms. */
int hs_kernel_timer_read(double *avg_ms, double *total_ms);

/* ---- batched engine: many models advanced by shared launches -------------
 * Jobs are admitted first come, first served while their device footprint
 * fits the budget (default: 80 % of the free device memory at creation).
 * Resident jobs form groups by phase (steady, transient) and kernel form;
 * each group advances in chunks of steady iterations or transient intervals
 * (one launch sequence per group) on its own stream, and the engine takes a
 * group's results when its chunk ends, independently of the other groups:
 * a finished steady solve is reported without waiting for the transient
 * groups, and an arriving job joins its group at the group's next chunk.
 * A job's result is the same as that of the same solve run alone.  The
 * engine's options (hs_engine_set_options) apply to every job; a model must
 * stay allocated and unchanged until its jobs are admitted.  transient is
 * HS_TRANS_RK4, HS_TRANS_TAYLOR or HS_TRANS_BE; with HS_TRANS_BE every
 * interval of a job is one implicit solve, the solves of all such jobs
 * sharing launches, and a job moves to its next interval when its solve has
 * converged (a solve stopped far above its tolerance fails the job). */
typedef struct hs_engine_st hs_engine_t;
hs_engine_t *hs_engine_create(int precision, int transient);
void hs_engine_destroy(hs_engine_t *e);
/* only while no job is queued or running */
void hs_engine_set_options(hs_engine_t *e, const hs_options_t *opt);
/* most transient intervals and the steady iterations of one chunk (defaults
 * 32, 128); a transient chunk also stops at about 20 ms of device time */
void hs_engine_set_tick(hs_engine_t *e, int intervals, int steady_iters);
/* bytes of device memory the engine may use for resident problems */
void hs_engine_set_budget(hs_engine_t *e, double bytes);
/* Submits a job on a populated model.  power holds power_rows block vectors
 * (total blocks + package nodes each; package-node entries are ignored):
 * 1 (the same power throughout), n_steps (one per interval, the steady solve
 * uses the first) or n_steps + 1 (the steady power, then one per interval).
 * n_steps == 0: a steady solve (steady_state_temp); n_steps > 0: n_steps
 * intervals of dt (compute_temp) from init_temp (a block vector) or, when
 * init_temp is NULL, from the block temperatures of a steady solve run
 * first.  Inputs are copied.  On a model with leakage_used the steady
 * solve is the temperature-leakage fixed point of steady_state_temp
 * (temperature.c:708-746) from zero block temperatures (the command line's
 * fresh temperature vector), each solve warm-started from the last.
 * Returns the job id, or -1 on error (e.g. a footprint above the budget). */
int hs_engine_submit(hs_engine_t *e, RC_model_t *model, const double *power, int power_rows,
                     const double *init_temp, int n_steps, double dt);
/* Submits a steady job under the package model (package.c:353-397):
 * r_convec from the fan and sink for forced convection, one solve; with
 * package->natural_convec, the natural-convection fixed point of the
 * command line (hotspot.c:667-682): r_convec at the ambient temperature +
 * 0.01 K, then, after each solve, at its average sink temperature
 * (calc_sink_temp + 0.01 K), solved again from the last solution until
 * r_convec moves by less than 0.01 K/W; the job fails above a 1000 K sink.
 * The result is the last solve's block vector, hs_job_info_t's r_convec
 * the last r_convec and steady_solves the number of solves.  Homogeneous
 * models without leakage only (no detailed-3D or microchannel layers); the
 * model itself is not changed.  Returns the job id, or -1 on error.  No
 * counterpart in the reference, whose command line runs the loop around
 * steady_state_temp and populate_R_model. */
int hs_engine_submit_package(hs_engine_t *e, RC_model_t *model, const double *power,
                             const package_config_t *package);
/* Admits queued jobs while they fit (their inputs are then resident on the
 * device); returns the number still queued.  hs_engine_tick admits too. */
int hs_engine_admit(hs_engine_t *e);
/* Admits what fits, starts a chunk for every group that is between chunks,
 * waits up to 2 ms for some chunk to end and takes the results of every
 * group whose chunk has ended; returns the number of jobs queued or running.
 * Call it in a loop (submitting arrivals in between) until it returns 0. */
int hs_engine_tick(hs_engine_t *e);
/* 1 when finished (the final block vector is copied into out, may be NULL),
 * 0 while queued or running, -1 for an unknown job. */
int hs_engine_poll(hs_engine_t *e, int job, double *out);
/* the block temperatures of every interval (n_steps block vectors) of a
 * finished transient job; 1 when copied, 0 while it runs, -1 on error */
int hs_engine_result_seq(hs_engine_t *e, int job, double *out);
typedef struct hs_job_info_st {
    int state;                /* 0 queued, 1 steady phase, 2 transient phase, 3 done, 4 failed
                               * (hs_engine_poll reports the reason) */
    int steady_iters;         /* CG iterations (inner ones for mixed precision) */
    int steady_refinements;
    int rk4_calls;            /* RK4 step attempts over the whole sequence */
    double taylor_sweeps;     /* Taylor term sweeps E_k -> E_k+1 over the whole sequence */
    double wait_s;            /* submission -> admission */
    double latency_s;         /* submission -> result (so far, while running) */
    double device_bytes;      /* device memory held while resident */
    double taylor_node_bytes; /* algorithmic bytes per node moved by the dominant Taylor sweep
                               * kernel over the whole sequence (throughput accounting) */
    int held_intervals;       /* Taylor: intervals held and run again because its group launched
                               * fewer Horner steps than the interval needed */
    int horner_steps_min;     /* Taylor: fewest Horner steps per sub-step its group launched while
                               * it ran, when below its plan (0: the plan's every time) */
    int horner_steps_plan;    /* Taylor: Horner steps per sub-step of its a-priori plan */
    double setup_s;           /* host time of its admission: device state and inputs */
    int resident_model;       /* 1: admitted on the device state its model kept from an earlier job
                               * of this engine (coefficients, maps, multigrid levels not rebuilt) */
    int steady_solves;        /* steady solves of its steady phase: those of the temperature-leakage
                               * fixed point for a model with leakage_used, else 1 (0: none yet) */
    double be_iters;          /* backward Euler: solver iterations over its intervals so far */
    double r_convec;          /* package-model jobs: the sink's r_convec after its last solve
                               * (package_model at that solve's average sink temperature) */
} hs_job_info_t;
int hs_engine_job_info(hs_engine_t *e, int job, hs_job_info_t *info);
/* frees a finished job's results */
void hs_engine_release(hs_engine_t *e, int job);
int hs_engine_active(hs_engine_t *e);
/* where the engine's host thread has spent its time since creation, and
 * the chunks it has run (serving observability) */
typedef struct hs_engine_stats_st {
    double admit_s;           /* admissions: device state, inputs */
    double sync_s;            /* uploads of group descriptors after membership changes, tuning */
    double launch_s;          /* launching chunks */
    double wait_s;            /* waiting for a chunk to end */
    double harvest_s;         /* results: phase changes, downloads */
    long steady_chunks, transient_chunks;
    long resident_hits;       /* admissions on a model's resident device state */
    double steady_chunk_s;    /* device time of the steady chunks (first launch to last, summed) */
    double transient_chunk_s; /* the same for the transient chunks */
} hs_engine_stats_t;
void hs_engine_get_stats(hs_engine_t *e, hs_engine_stats_t *stats);

#ifdef __cplusplus
}
#endif
#endif
