// Host side of the MI355X grid thermal model: configuration, floorplan /
// layer-file / materials / microchannel-network parsing, the layer stack,
// block<->grid maps, layer and package R/C, the text file formats, and the
// reference's C entry points.  The solves run in hs_device.hip.
//
// Each function cites the reference code whose behaviour it keeps.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <strings.h>

#include "hs_internal.h"

// ---- ABI layout pins (reference implementation headers) -------------------
static_assert(sizeof(unit_t) == 512 + 6 * 8 + 2 * 4, "unit_t layout (flp.h:378-391)");
static_assert(offsetof(thermal_config_t, model_secondary) == 17 * 8, "thermal_config_t");
static_assert(offsetof(thermal_config_t, grid_transient_file) + STR_SIZE + 4 ==
                  offsetof(thermal_config_t, detailed_3D_used) + 4,
              "grid_transient_file precedes detailed_3D_used (temperature.h:217-220)");
static_assert(sizeof(package_RC_t) == 52 * 8, "package_RC_t (temperature.h:237-334)");
static_assert(offsetof(layer_t, rx) == 64, "layer_t (temperature_grid.h:101-126)");
static_assert(offsetof(grid_model_t, layers) == sizeof(thermal_config_t), "grid_model_t");
static_assert(offsetof(RC_model_t, type) == 8 && offsetof(RC_model_t, config) == 16, "RC_model_t");

namespace hs {

// ---------------------------------------------------------------------------
// errors (util.c:32-41)
// ---------------------------------------------------------------------------
static int g_exit_on_error = 1;
static thread_local std::string g_last_error;
static thread_local int g_has_error = 0;

void fatal(const std::string &msg) { throw Error{msg}; }
void warning(const std::string &msg) { fprintf(stderr, "warning: %s", msg.c_str()); }

void report(const std::string &msg) {
    if (g_exit_on_error) {
        fprintf(stderr, "error: %s", msg.c_str());
        exit(1);
    }
    g_last_error = msg;
    g_has_error = 1;
}

Model *model_of(RC_model_t *m) {
    if (!m) fatal("null model\n");
    if (m->type == BLOCK_MODEL) fatal("this call needs a grid model (model_type grid)\n");
    if (m->type != GRID_MODEL) fatal("unknown model type\n");
    return reinterpret_cast<Handle *>(m)->impl;
}

Model *model_of(grid_model_t *g) {
    if (!g) fatal("null model\n");
    Handle *h = reinterpret_cast<Handle *>(reinterpret_cast<char *>(g) - offsetof(Handle, pub));
    return h->impl;
}

// ---------------------------------------------------------------------------
// small numerics (util.c:17-20, :57-77; RCutil.c:20-31)
// ---------------------------------------------------------------------------
static int tolerant_ceil(double v) {
    double nearest = std::floor(v + 0.5);
    return eq(v, nearest) ? (int)nearest : (int)std::ceil(v);
}
static int tolerant_floor(double v) {
    double nearest = std::floor(v + 0.5);
    return eq(v, nearest) ? (int)nearest : (int)std::floor(v);
}

// ---------------------------------------------------------------------------
// text input helpers
// ---------------------------------------------------------------------------
// First whitespace-delimited token of a line, "" for blank/comment lines.
bool first_token(const char *line, std::string &tok) {
    const char *p = line;
    while (*p && strchr(" \r\t\n", *p)) p++;
    if (!*p || *p == '#') return false;
    const char *q = p;
    while (*q && !strchr(" \r\t\n", *q)) q++;
    tok.assign(p, q - p);
    return true;
}

std::vector<std::string> read_lines(const char *file) {
    FILE *fp = strcasecmp(file, "stdin") ? fopen(file, "r") : stdin;
    if (!fp) fatal(std::string("error opening file ") + file + "\n");
    std::vector<std::string> out;
    std::vector<char> buf(65536);
    while (fgets(buf.data(), (int)buf.size(), fp)) out.emplace_back(buf.data());
    if (fp != stdin) fclose(fp);
    return out;
}

// ---------------------------------------------------------------------------
// floorplans (flp.c:1004-1218, :334-350, :1385-1402, :1501-1531)
// ---------------------------------------------------------------------------
static flp_t *flp_from_lines(const std::vector<std::string> &lines, int read_connects,
                             int init_connects) {
    int count = 0;
    char name[STR_SIZE], name2[STR_SIZE];
    for (auto &l : lines) {
        std::string tok;
        if (!first_token(l.c_str(), tok)) continue;
        double a[6];
        if (sscanf(l.c_str(), "%511s%lf%lf%lf%lf%lf%lf", name, &a[0], &a[1], &a[2], &a[3],
                   &a[4], &a[5]) > 4)
            count++;
    }
    if (!count) fatal("no units specified in the floorplan file\n");
    flp_t *flp = (flp_t *)calloc(1, sizeof(flp_t));
    flp->units = (unit_t *)calloc(count, sizeof(unit_t));
    if (!flp || !flp->units) fatal("memory allocation error\n");
    flp->n_units = count;
    if (read_connects || init_connects) {
        flp->wire_density = (double **)calloc(count, sizeof(double *));
        for (int i = 0; i < count; i++) flp->wire_density[i] = (double *)calloc(count, sizeof(double));
    }
    int i = 0;
    for (auto &l : lines) {
        std::string tok;
        if (!first_token(l.c_str(), tok)) continue;
        double v[7];
        int got = sscanf(l.c_str(), "%511s%lf%lf%lf%lf%lf%lf%lf", name, &v[0], &v[1], &v[2],
                         &v[3], &v[4], &v[5], &v[6]);
        if (got > 7) fatal("invalid floorplan file format\n");
        if (got == 7 || got == 5) {
            if (i >= count) fatal("mismatch of number of units\n");
            unit_t &u = flp->units[i++];
            snprintf(u.name, STR_SIZE, "%s", name);
            u.width = v[0];
            u.height = v[1];
            u.leftx = v[2];
            u.bottomy = v[3];
            u.specificheat = got == 7 ? v[4] : 0.0;
            u.resistivity = got == 7 ? v[5] : 0.0;
            u.hasRes = u.hasSh = got == 7;
        } else {
            double wd;
            if (sscanf(l.c_str(), "%511s%511s%lf", name, name2, &wd) != 3)
                fatal("invalid floorplan file format\n");
        }
    }
    if (i != count) fatal("mismatch of number of units\n");
    if (read_connects) {
        for (auto &l : lines) {
            std::string tok;
            if (!first_token(l.c_str(), tok)) continue;
            double f[6];
            int ti;
            if (sscanf(l.c_str(), "%511s%lf%lf%lf%lf%lf%lf", name, &f[0], &f[1], &f[2], &f[3],
                       &f[4], &f[5]) == 7 ||
                sscanf(l.c_str(), "%511s%lf%lf%lf%d", name, &f[0], &f[1], &f[2], &ti) == 5)
                continue;
            double wd;
            if (sscanf(l.c_str(), "%511s%511s%lf", name, name2, &wd) == 3) {
                int x = get_blk_index(flp, name), y = get_blk_index(flp, name2);
                if (x < 0 || y < 0) fatal("block not found\n");
                if (x == y) fatal("block connected to itself?\n");
                double &a = flp->wire_density[x][y], &b = flp->wire_density[y][x];
                if (!a && !b) {
                    a = b = wd;
                } else if (a != b || a != wd) {
                    fatal(std::string("wrong connectivity information for blocks ") + name +
                          " and " + name2 + "\n");
                }
            } else {
                fatal("invalid floorplan file format\n");
            }
        }
    } else if (init_connects) {
        for (int x = 0; x < count; x++)
            for (int y = 0; y < count; y++) flp->wire_density[x][y] = 1.0;
    }
    // move the origin to (0,0) (flp.c:334-350)
    double minx = flp->units[0].leftx, miny = flp->units[0].bottomy;
    for (int k = 1; k < count; k++) {
        minx = std::min(minx, flp->units[k].leftx);
        miny = std::min(miny, flp->units[k].bottomy);
    }
    for (int k = 0; k < count; k++) {
        flp->units[k].leftx += (0.0 - minx);
        flp->units[k].bottomy += (0.0 - miny);
    }
    return flp;
}

double total_width(const flp_t *f) {
    double mn = f->units[0].leftx, mx = f->units[0].leftx + f->units[0].width;
    for (int i = 1; i < f->n_units; i++) {
        mn = std::min(mn, f->units[i].leftx);
        mx = std::max(mx, f->units[i].leftx + f->units[i].width);
    }
    return mx - mn;
}
double total_height(const flp_t *f) {
    double mn = f->units[0].bottomy, mx = f->units[0].bottomy + f->units[0].height;
    for (int i = 1; i < f->n_units; i++) {
        mn = std::min(mn, f->units[i].bottomy);
        mx = std::max(mx, f->units[i].bottomy + f->units[i].height);
    }
    return mx - mn;
}

static void free_flp_impl(flp_t *flp) {
    if (!flp) return;
    if (flp->wire_density) {
        for (int i = 0; i < flp->n_units; i++) free(flp->wire_density[i]);
        free(flp->wire_density);
    }
    free(flp->units);
    free(flp);
}

// ---------------------------------------------------------------------------
// materials (materials.c:17-205)
// ---------------------------------------------------------------------------
static bool find_material(const materials_list_t *ml, const char *name, material_t &out) {
    if (!ml) return false;
    for (int i = 0; i < ml->size; i++)
        if (!strncmp(ml->names[i], name, STR_SIZE)) {
            out = ml->material_properties[i];
            return true;
        }
    return false;
}

// ---------------------------------------------------------------------------
// package R and C (temperature.c:458-569)
// ---------------------------------------------------------------------------
constexpr double K_METAL = 1.0 / 0.0025, K_C4 = 1.0 / 0.8, K_SUB = 1.0 / 0.5;
constexpr double K_SOLDER = 1.0 / 0.06, K_PCB = 1.0 / 0.333;
constexpr double SPEC_HEAT_METAL = 3.55e6, SPEC_HEAT_C4 = 1.65e6, SPEC_HEAT_SUB = 1.6e6;
constexpr double SPEC_HEAT_SOLDER = 2.1e6, SPEC_HEAT_PCB = 1.32e6;

void package_R(package_RC_t &p, const thermal_config_t &c, double w, double h) {
    const double ssp = c.s_spreader, tsp = c.t_spreader, shs = c.s_sink, ths = c.t_sink;
    const double khs = c.k_sink, ksp = c.k_spreader, rcv = c.r_convec;
    const double ssub = c.s_sub, tsub = c.t_sub, ssol = c.s_solder, tsol = c.t_solder;
    const double spcb = c.s_pcb, tpcb = c.t_pcb, rsec = c.r_convec_sec;
    // spreader / sink lateral
    p.r_sp1_x = getr(ksp, (ssp - w) / 4.0, (ssp + 3 * h) / 4.0 * tsp);
    p.r_sp1_y = getr(ksp, (ssp - h) / 4.0, (ssp + 3 * w) / 4.0 * tsp);
    p.r_hs1_x = getr(khs, (ssp - w) / 4.0, (ssp + 3 * h) / 4.0 * ths);
    p.r_hs1_y = getr(khs, (ssp - h) / 4.0, (ssp + 3 * w) / 4.0 * ths);
    p.r_hs2_x = getr(khs, (ssp - w) / 4.0, (3 * ssp + h) / 4.0 * ths);
    p.r_hs2_y = getr(khs, (ssp - h) / 4.0, (3 * ssp + w) / 4.0 * ths);
    p.r_hs = getr(khs, (shs - ssp) / 4.0, (shs + 3 * ssp) / 4.0 * ths);
    // spreader / sink vertical
    p.r_sp_per_x = getr(ksp, tsp, (ssp + h) * (ssp - w) / 4.0);
    p.r_sp_per_y = getr(ksp, tsp, (ssp + w) * (ssp - h) / 4.0);
    p.r_hs_c_per_x = getr(khs, ths, (ssp + h) * (ssp - w) / 4.0);
    p.r_hs_c_per_y = getr(khs, ths, (ssp + w) * (ssp - h) / 4.0);
    p.r_hs_per = getr(khs, ths, (shs * shs - ssp * ssp) / 4.0);
    // to ambient, r_convec split by area
    p.r_amb_c_per_x = rcv * (shs * shs) / ((ssp + h) * (ssp - w) / 4.0);
    p.r_amb_c_per_y = rcv * (shs * shs) / ((ssp + w) * (ssp - h) / 4.0);
    p.r_amb_per = rcv * (shs * shs) / ((shs * shs - ssp * ssp) / 4.0);
    // secondary path lateral
    p.r_sub1_x = getr(K_SUB, (ssub - w) / 4.0, (ssub + 3 * h) / 4.0 * tsub);
    p.r_sub1_y = getr(K_SUB, (ssub - h) / 4.0, (ssub + 3 * w) / 4.0 * tsub);
    p.r_solder1_x = getr(K_SOLDER, (ssol - w) / 4.0, (ssol + 3 * h) / 4.0 * tsol);
    p.r_solder1_y = getr(K_SOLDER, (ssol - h) / 4.0, (ssol + 3 * w) / 4.0 * tsol);
    p.r_pcb1_x = getr(K_PCB, (ssol - w) / 4.0, (ssol + 3 * h) / 4.0 * tpcb);
    p.r_pcb1_y = getr(K_PCB, (ssol - h) / 4.0, (ssol + 3 * w) / 4.0 * tpcb);
    p.r_pcb2_x = getr(K_PCB, (ssol - w) / 4.0, (3 * ssol + h) / 4.0 * tpcb);
    p.r_pcb2_y = getr(K_PCB, (ssol - h) / 4.0, (3 * ssol + w) / 4.0 * tpcb);
    p.r_pcb = getr(K_PCB, (spcb - ssol) / 4.0, (spcb + 3 * ssol) / 4.0 * tpcb);
    // secondary path vertical
    p.r_sub_per_x = getr(K_SUB, tsub, (ssub + h) * (ssub - w) / 4.0);
    p.r_sub_per_y = getr(K_SUB, tsub, (ssub + w) * (ssub - h) / 4.0);
    p.r_solder_per_x = getr(K_SOLDER, tsol, (ssol + h) * (ssol - w) / 4.0);
    p.r_solder_per_y = getr(K_SOLDER, tsol, (ssol + w) * (ssol - h) / 4.0);
    p.r_pcb_c_per_x = getr(K_PCB, tpcb, (ssol + h) * (ssol - w) / 4.0);
    p.r_pcb_c_per_y = getr(K_PCB, tpcb, (ssol + w) * (ssol - h) / 4.0);
    p.r_pcb_per = getr(K_PCB, tpcb, (spcb * spcb - ssol * ssol) / 4.0);
    p.r_amb_sec_c_per_x = rsec * (spcb * spcb) / ((ssol + h) * (ssol - w) / 4.0);
    p.r_amb_sec_c_per_y = rsec * (spcb * spcb) / ((ssol + w) * (ssol - h) / 4.0);
    p.r_amb_sec_per = rsec * (spcb * spcb) / ((spcb * spcb - ssol * ssol) / 4.0);
}

void package_C(package_RC_t &p, const thermal_config_t &c, double w, double h) {
    const double ssp = c.s_spreader, tsp = c.t_spreader, shs = c.s_sink, ths = c.t_sink;
    const double ccv = c.c_convec, phs = c.p_sink, psp = c.p_spreader;
    const double ssub = c.s_sub, tsub = c.t_sub, ssol = c.s_solder, tsol = c.t_solder;
    const double spcb = c.s_pcb, tpcb = c.t_pcb, csec = c.c_convec_sec;
    p.c_sp_per_x = getcap(psp, tsp, (ssp + h) * (ssp - w) / 4.0);
    p.c_sp_per_y = getcap(psp, tsp, (ssp + w) * (ssp - h) / 4.0);
    p.c_hs_c_per_x = getcap(phs, ths, (ssp + h) * (ssp - w) / 4.0);
    p.c_hs_c_per_y = getcap(phs, ths, (ssp + w) * (ssp - h) / 4.0);
    p.c_hs_per = getcap(phs, ths, (shs * shs - ssp * ssp) / 4.0);
    p.c_amb_c_per_x = kCFactor * ccv / (shs * shs) * ((ssp + h) * (ssp - w) / 4.0);
    p.c_amb_c_per_y = kCFactor * ccv / (shs * shs) * ((ssp + w) * (ssp - h) / 4.0);
    p.c_amb_per = kCFactor * ccv / (shs * shs) * ((shs * shs - ssp * ssp) / 4.0);
    p.c_sub_per_x = getcap(SPEC_HEAT_SUB, tsub, (ssub + h) * (ssub - w) / 4.0);
    p.c_sub_per_y = getcap(SPEC_HEAT_SUB, tsub, (ssub + w) * (ssub - h) / 4.0);
    p.c_solder_per_x = getcap(SPEC_HEAT_SOLDER, tsol, (ssol + h) * (ssol - w) / 4.0);
    p.c_solder_per_y = getcap(SPEC_HEAT_SOLDER, tsol, (ssol + w) * (ssol - h) / 4.0);
    p.c_pcb_c_per_x = getcap(SPEC_HEAT_PCB, tpcb, (ssol + h) * (ssol - w) / 4.0);
    p.c_pcb_c_per_y = getcap(SPEC_HEAT_PCB, tpcb, (ssol + w) * (ssol - h) / 4.0);
    p.c_pcb_per = getcap(SPEC_HEAT_PCB, tpcb, (spcb * spcb - ssol * ssol) / 4.0);
    p.c_amb_sec_c_per_x = kCFactor * csec / (spcb * spcb) * ((ssol + h) * (ssol - w) / 4.0);
    p.c_amb_sec_c_per_y = kCFactor * csec / (spcb * spcb) * ((ssol + w) * (ssol - h) / 4.0);
    p.c_amb_sec_per = kCFactor * csec / (spcb * spcb) * ((spcb * spcb - ssol * ssol) / 4.0);
}

// ---------------------------------------------------------------------------
// microchannel networks (microchannel.c:258-376)
// ---------------------------------------------------------------------------
static double hydro_c(double h, double w, double L, double visc) {
    if (h == w) return (0.42229 * std::pow(h, 4)) / (12 * visc * L);
    if (h > w) return ((1 - 0.63 * (w / h)) * std::pow(w, 3) * h) / (12 * visc * L);
    return ((1 - 0.63 * (h / w)) * std::pow(h, 3) * w) / (12 * visc * L);
}

static inline bool is_fluid(int t) { return t == CELL_FLUID || t == CELL_INLET || t == CELL_OUTLET; }

// Parses the CSV network, writes the per-cell floorplan next to it exactly as
// the reference does and returns the layer's floorplan, whose values carry
// the same '%e' rounding as the reference's (which reads the file back).
static flp_t *build_network(Microchannel &mc, int nr, int nc) {
    microchannel_config_t &c = mc.cfg;
    FILE *fp = fopen(c.network_file, "r");
    if (!fp) fatal(std::string("Unable to open ") + c.network_file + "\n");
    mc.types.assign((size_t)nr * nc, 0);
    std::vector<char> line(4096);
    int i = 0;
    // a row per line read; a final line without a newline is dropped, as in
    // the reference's fgets/feof loop (microchannel.c:288-324)
    while (fgets(line.data(), (int)line.size(), fp)) {
        if (feof(fp)) break;
        int j = 0;
        char *save = nullptr;
        for (char *cell = strtok_r(line.data(), " ,\n", &save); cell;
             cell = strtok_r(nullptr, " ,\n", &save)) {
            if (i >= nr) {
                fclose(fp);
                fatal("Microchannel has more rows than num_rows\n");
            }
            if (j >= nc) {
                fclose(fp);
                fatal("Microchannel row has more cells than num_columns\n");
            }
            mc.types[(size_t)i * nc + j] = atoi(cell);
            j++;
        }
        i++;
        if (i > nr) {
            fclose(fp);
            fatal("Microchannel has more rows than num_rows\n");
        }
    }
    fclose(fp);
    c.num_rows = nr;
    c.num_columns = nc;
    c.cell_types = (int **)calloc(nr, sizeof(int *));
    for (int r = 0; r < nr; r++) {
        c.cell_types[r] = (int *)calloc(nc, sizeof(int));
        for (int q = 0; q < nc; q++) c.cell_types[r][q] = mc.types[(size_t)r * nc + q];
    }
    // generated floorplan (microchannel.c:326-351): one unit per cell, written
    // with "%e" and used as read back.  Only the row and column positions and
    // six other values are distinct, so each is formatted (and read back) once.
    snprintf(c.floorplan_file, sizeof(c.floorplan_file), "%s", c.network_file);
    char *ext = strstr(c.floorplan_file, ".csv");
    if (ext) strcpy(ext, ".flp");
    auto fmt = [](double v) {
        char b[64];
        snprintf(b, sizeof b, "%e", v);
        return std::string(b);
    };
    std::vector<std::string> sx(nc), sy(nr);
    for (int q = 0; q < nc; q++) sx[q] = fmt(q * c.cell_width);
    for (int r = 0; r < nr; r++) sy[r] = fmt((nr - r - 1) * c.cell_height);
    const std::string sw = fmt(c.cell_width), sh = fmt(c.cell_height);
    const std::string scv[2] = {fmt(c.wall_capac), fmt(c.coolant_capac)};
    const std::string srs[2] = {fmt(c.wall_res), fmt(c.coolant_res)};
    auto val = [](const std::string &t) { return strtod(t.c_str(), nullptr); };
    flp_t *flp = (flp_t *)calloc(1, sizeof(flp_t));
    if (!flp) fatal("memory allocation error\n");
    flp->n_units = nr * nc;
    flp->units = (unit_t *)calloc((size_t)nr * nc, sizeof(unit_t));
    if (!flp->units) {
        free(flp);
        fatal("memory allocation error\n");
    }
    FILE *out = fopen(c.floorplan_file, "w");
    if (out) fputs("# Name\tw\th\tx\ty\tc_v\tp\n", out);
    std::string text;
    for (int r = 0; r < nr; r++)
        for (int q = 0; q < nc; q++) {
            const int f = is_fluid(mc.types[(size_t)r * nc + q]) ? 1 : 0;
            unit_t &u = flp->units[(size_t)r * nc + q];
            snprintf(u.name, STR_SIZE, "Cell_%d_%d", r, q);
            u.width = val(sw);
            u.height = val(sh);
            u.leftx = val(sx[q]);
            u.bottomy = val(sy[r]);
            u.specificheat = val(scv[f]);
            u.resistivity = val(srs[f]);
            u.hasRes = u.hasSh = 1;
            if (out) {
                text.assign(u.name);
                for (const std::string *t : {&sw, &sh, (const std::string *)&sx[q], (const std::string *)&sy[r],
                                             &scv[f], &srs[f]})
                    text += '\t', text += *t;
                text += '\n';
                fputs(text.c_str(), out);
            }
        }
    if (out) fclose(out);
    // pressure network (microchannel.c:378-565)
    int nfl = 0;
    for (int t : mc.types) nfl += is_fluid(t);
    c.n_fluid_cells = nfl;
    if (c.pump_internal_res != 0.0)
        fatal("a pump with internal resistance is not supported (ideal pump only)\n");
    // the pressure network is solved on the device at the model's first
    // device build (pressure_on_device); a FLUID cell without a fluid-type
    // neighbour would leave it singular
    for (int r = 0; r < nr; r++)
        for (int q = 0; q < nc; q++) {
            if (mc.types[(size_t)r * nc + q] != CELL_FLUID) continue;
            const bool linked = (r > 0 && is_fluid(mc.types[(size_t)(r - 1) * nc + q])) ||
                                (r < nr - 1 && is_fluid(mc.types[(size_t)(r + 1) * nc + q])) ||
                                (q > 0 && is_fluid(mc.types[(size_t)r * nc + q - 1])) ||
                                (q < nc - 1 && is_fluid(mc.types[(size_t)r * nc + q + 1]));
            if (!linked) fatal("isolated fluid cell in the microchannel network\n");
        }
    mc.pressure.clear();
    mc.hydro_c = hydro_c(c.cell_thickness, c.cell_width, c.cell_height, c.coolant_visc);
    return flp;
}

// A second layer on the same network: its cell types, counts and per-cell
// floorplan are those of the first (whose floorplan file is already written).
static flp_t *copy_network(Microchannel &mc, const Microchannel &from, const flp_t *from_flp, int nr, int nc) {
    microchannel_config_t &c = mc.cfg;
    mc.types = from.types;
    c.num_rows = nr;
    c.num_columns = nc;
    c.n_fluid_cells = from.cfg.n_fluid_cells;
    snprintf(c.floorplan_file, sizeof(c.floorplan_file), "%s", from.cfg.floorplan_file);
    if (c.pump_internal_res != 0.0)
        fatal("a pump with internal resistance is not supported (ideal pump only)\n");
    c.cell_types = (int **)calloc(nr, sizeof(int *));
    for (int r = 0; r < nr; r++) {
        c.cell_types[r] = (int *)calloc(nc, sizeof(int));
        for (int q = 0; q < nc; q++) c.cell_types[r][q] = mc.types[(size_t)r * nc + q];
    }
    mc.pressure.clear();
    mc.hydro_c = hydro_c(c.cell_thickness, c.cell_width, c.cell_height, c.coolant_visc);
    flp_t *f = (flp_t *)calloc(1, sizeof(flp_t));
    if (!f) fatal("memory allocation error\n");
    f->n_units = from_flp->n_units;
    f->units = (unit_t *)malloc(sizeof(unit_t) * (size_t)f->n_units);
    if (!f->units) fatal("memory allocation error\n");
    memcpy(f->units, from_flp->units, sizeof(unit_t) * (size_t)f->n_units);
    return f;
}

// ---------------------------------------------------------------------------
// layer stack (temperature_grid.c:459-948)
// ---------------------------------------------------------------------------
static void set_layer(layer_t &l, int no, int lateral, int power, double k, double t, double sp,
                      flp_t *flp) {
    l.no = no;
    l.has_lateral = lateral;
    l.has_power = power;
    l.k = k;
    l.thickness = t;
    l.sp = sp;
    l.flp = flp;
}

struct LcfRecord {
    int no, lateral, power;
    double sp, k, thickness;
    std::string flp;
};

// 7-field records, one field per line (temperature_grid.c:612-760)
static std::vector<LcfRecord> parse_lcf(const char *file, const materials_list_t *ml) {
    auto lines = read_lines(file);
    std::vector<LcfRecord> recs;
    LcfRecord r{};
    int field = 0;
    for (auto &l : lines) {
        std::string tok;
        if (!first_token(l.c_str(), tok)) continue;
        double dv;
        switch (field) {
            case 0:
                if (sscanf(tok.c_str(), "%d", &r.no) != 1) fatal("invalid layer number\n");
                field = 1;
                break;
            case 1:
                if (tok[0] == 'Y' || tok[0] == 'y') r.lateral = 1;
                else if (tok[0] == 'N' || tok[0] == 'n') r.lateral = 0;
                else fatal("invalid layer heat flow indicator\n");
                field = 2;
                break;
            case 2:
                if (tok[0] == 'Y' || tok[0] == 'y') r.power = 1;
                else if (tok[0] == 'N' || tok[0] == 'n') r.power = 0;
                else fatal("invalid layer power dissipation indicator\n");
                field = 3;
                break;
            case 3:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) {
                    material_t m{};
                    double k = -1, cv = -1;
                    if (find_material(ml, tok.c_str(), m)) {
                        k = m.thermal_conductivity;
                        cv = m.volumetric_heat_capacity;
                    }
                    r.sp = cv;
                    r.k = k;
                    field = 5;
                } else {
                    r.sp = dv;
                    field = 4;
                }
                break;
            case 4:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) fatal("invalid resistivity\n");
                r.k = 1.0 / dv;
                field = 5;
                break;
            case 5:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) fatal("invalid thickness\n");
                r.thickness = dv;
                field = 6;
                break;
            case 6:
                r.flp = tok;
                recs.push_back(r);
                r = LcfRecord{};
                field = 0;
                break;
        }
    }
    return recs;
}

static void build_layers(Model &M, flp_t *flp_default, materials_list_t *ml) {
    grid_model_t &g = *M.g;
    const bool sec = g.config.model_secondary;
    std::vector<int> chip_idx;  // indices of chip (non-package) layers
    int base = sec ? 3 : 0;
    if (g.has_lcf) {
        auto recs = parse_lcf(g.config.grid_layer_file, ml);
        int inner = (int)recs.size();
        g.n_layers = inner + 2 + (sec ? 3 : 0);
        M.layers.assign(g.n_layers, layer_t{});
        M.mc.assign(g.n_layers, nullptr);
        std::vector<int> seen(inner, 0);
        for (size_t q = 0; q < recs.size(); q++) {
            const LcfRecord &r = recs[q];
            if (r.no >= inner || r.no < 0)
                fatal("layer number must be >= 0 and < no. of layers specified\n");
            if (seen[r.no]) fatal("layer numbers must be unique\n");
            seen[r.no] = 1;
            int i = base + r.no;
            layer_t &L = M.layers[i];
            set_layer(L, i, r.lateral, r.power, r.k, r.thickness, r.sp, nullptr);
            bool csv = r.flp.find(".csv") != std::string::npos;
            if (csv && g.use_microchannels) {
                if (q == 0) fatal("a microchannel layer cannot be the first layer\n");
                auto *mc = new Microchannel();
                if (g.default_microchannel_config) mc->cfg = *g.default_microchannel_config;
                else mc->cfg = default_microchannel_config();
                mc->cfg.cell_types = nullptr;
                mc->cfg.mapping = nullptr;
                mc->cfg.A = nullptr;
                mc->cfg.b = nullptr;
                mc->cfg.cell_width = g.width / g.cols;
                mc->cfg.cell_height = g.height / g.rows;
                mc->cfg.cell_thickness = r.thickness;
                snprintf(mc->cfg.network_file, STR_SIZE, "%s", r.flp.c_str());
                // a network already built for an earlier layer with the same
                // file and cell geometry is the same network: copy it
                const Microchannel *same = nullptr;
                for (int k = 0; k < i && !same; k++) {
                    const Microchannel *o = M.mc[k];
                    if (o && !strcmp(o->cfg.network_file, mc->cfg.network_file) &&
                        o->cfg.cell_width == mc->cfg.cell_width && o->cfg.cell_height == mc->cfg.cell_height &&
                        o->cfg.cell_thickness == mc->cfg.cell_thickness)
                        same = o, L.flp = copy_network(*mc, *o, M.layers[k].flp, g.rows, g.cols);
                }
                if (!same) L.flp = build_network(*mc, g.rows, g.cols);
                L.is_microchannel = 1;
                L.microchannel_config = &mc->cfg;
                M.mc[i] = mc;
            } else if (csv) {
                fatal("Floorplan file has microchannel extension but use_microchannels = 0\n");
            } else {
                L.flp = flp_from_lines(read_lines(r.flp.c_str()), 0, 0);
            }
            M.owned_flps.push_back(L.flp);
            if (q == 0) {
                g.width = total_width(L.flp);
                g.height = total_height(L.flp);
            } else if (!eq(g.width, total_width(L.flp)) || !eq(g.height, total_height(L.flp))) {
                fatal("width and height differ across layers\n");
            }
        }
        for (int q = 0; q < inner; q++) chip_idx.push_back(base + q);
    } else {
        if (!flp_default) fatal("Either LCF or FLP file must be specified\n");
        g.n_layers = 2 + 2 + (sec ? 5 : 0);
        M.layers.assign(g.n_layers, layer_t{});
        M.mc.assign(g.n_layers, nullptr);
        int si = sec ? 5 : 0, tim = sec ? 6 : 1;
        set_layer(M.layers[si], si, 1, 1, g.config.k_chip, g.config.t_chip, g.config.p_chip, flp_default);
        set_layer(M.layers[tim], tim, 1, 0, g.config.k_interface, g.config.t_interface,
                  g.config.p_interface, flp_default);
        if (sec) {
            set_layer(M.layers[4], 4, 1, 0, K_METAL, g.config.t_metal, SPEC_HEAT_METAL, flp_default);
            set_layer(M.layers[3], 3, 1, 0, K_C4, g.config.t_c4, SPEC_HEAT_C4, flp_default);
            chip_idx = {3, 4, 5, 6};
        } else {
            chip_idx = {0, 1};
        }
    }
    const int nl = g.n_layers;
    M.spidx = nl - 2;
    M.hsidx = nl - 1;
    flp_t *last = M.layers[M.spidx - 1].flp;
    set_layer(M.layers[M.spidx], M.spidx, 1, 0, g.config.k_spreader, g.config.t_spreader,
              g.config.p_spreader, last);
    set_layer(M.layers[M.hsidx], M.hsidx, 1, 0, g.config.k_sink, g.config.t_sink, g.config.p_sink,
              last);
    if (sec) {
        M.subidx = 2;
        M.solderidx = 1;
        M.pcbidx = 0;
        flp_t *top = M.layers[g.has_lcf ? 3 : 5].flp;
        set_layer(M.layers[2], 2, 1, 0, K_SUB, g.config.t_sub, SPEC_HEAT_SUB, top);
        set_layer(M.layers[1], 1, 1, 0, K_SOLDER, g.config.t_solder, SPEC_HEAT_SOLDER, top);
        set_layer(M.layers[0], 0, 1, 0, K_PCB, g.config.t_pcb, SPEC_HEAT_PCB, top);
    }
    M.nx = EXTRA + (sec ? EXTRA_SEC : 0);
    // block-grid maps: one per distinct floorplan, package layers share the
    // map of the layer whose floorplan they borrow, except in detailed-3D
    // mode where they get their own (with per-cell values disabled)
    // (temperature_grid.c:480-495, :532-606, :755-760, :1009-1039)
    M.layer_map.assign(nl, -1);
    auto new_map = [&](int owner, bool pkg) {
        M.maps.emplace_back();
        M.map_owner.push_back(owner);
        M.map_pkg.push_back(pkg);
        return (int)M.maps.size() - 1;
    };
    if (g.has_lcf) {
        for (int i : chip_idx) M.layer_map[i] = new_map(i, false);
    } else {
        int m = new_map(sec ? 5 : 0, false);
        for (int i : chip_idx) M.layer_map[i] = m;
    }
    if (M.detailed) {
        int m = new_map(M.spidx, true);
        M.layer_map[M.spidx] = M.layer_map[M.hsidx] = m;
    } else {
        M.layer_map[M.spidx] = M.layer_map[M.hsidx] = M.layer_map[M.spidx - 1];
    }
    if (sec) {
        int top = g.has_lcf ? 3 : 5;
        if (M.detailed) {
            int m = new_map(M.subidx, true);
            M.layer_map[0] = M.layer_map[1] = M.layer_map[2] = m;
        } else {
            M.layer_map[0] = M.layer_map[1] = M.layer_map[2] = M.layer_map[top];
        }
    }
    g.layers = M.layers.data();
    g.total_n_blocks = 0;
    for (int i = 0; i < nl; i++) g.total_n_blocks += g.layers[i].flp->n_units;
}

// set_bgmap (temperature_grid.c:341-456) with the detailed-3D list-head
// rules of new_blist / blist_append (:193-224, :286-320)
static void build_map(Model &M, int mi) {
    grid_model_t &g = *M.g;
    const layer_t &L = M.layers[M.map_owner[mi]];
    BlockMap &bm = M.maps[mi];
    const int nr = g.rows, nc = g.cols;
    const double cw = g.width / nc, ch = g.height / nr, H = g.height;
    const bool det = M.detailed;
    const flp_t *f = L.flp;
    std::vector<std::vector<std::pair<int, double>>> ent((size_t)nr * nc);
    bm.ranges.assign(f->n_units, glist_t{});
    bm.rx.assign((size_t)nr * nc, 0.0);
    bm.ry = bm.rz = bm.cap = bm.rx;
    bm.lock.assign((size_t)nr * nc, 0);
    bm.has_res.assign((size_t)nr * nc, 0);
    for (int u = 0; u < f->n_units; u++) {
        const unit_t &un = f->units[u];
        const double lu = un.leftx, ru = lu + un.width, bu = un.bottomy, tu = bu + un.height;
        const int i1 = nr - tolerant_ceil(tu / ch), i2 = nr - tolerant_floor(bu / ch);
        const int j1 = tolerant_floor(lu / cw), j2 = tolerant_ceil(ru / cw);
        if (i1 < 0 || j1 < 0) fatal("negative grid cell start index!\n");
        if (i2 > nr || j2 > nc) fatal("grid cell end index out of bounds!\n");
        if (i1 >= i2 || j1 >= j2) fatal("invalid floorplan spec or grid resolution\n");
        bm.ranges[u] = glist_t{i1, i2, j1, j2};
        const double res = (un.hasRes && det) ? un.resistivity : 1.0 / L.k;
        const double sh = (un.hasSh && det) ? un.specificheat : L.sp;
        const double t = L.thickness;
        for (int i = i1; i < i2; i++)
            for (int j = j1; j < j2; j++) {
                double occ;
                if (i > i1 && i < i2 - 1 && j > j1 && j < j2 - 1) {
                    occ = 1.0;
                } else {
                    const double lc = j * cw, rc = (j + 1) * cw;
                    const double tc = H - i * ch, bc = H - (i + 1) * ch;
                    double oh = std::min(tu, tc) - std::max(bu, bc);
                    double ow = std::min(ru, rc) - std::max(lu, lc);
                    if (eq(oh / ch, 0)) oh = 0;
                    else if (eq(oh / ch, 1)) oh = ch;
                    if (eq(ow / cw, 0)) ow = 0;
                    else if (eq(ow / cw, 1)) ow = cw;
                    occ = (oh * ow) / (ch * cw);
                    if (oh < 0 || ow < 0) fatal("negative overlap!\n");
                }
                const size_t c = (size_t)i * nc + j;
                if (det) {
                    const double rxu = getr(1.0 / res, cw, ch * t), ryu = getr(1.0 / res, ch, cw * t);
                    const double rzu = getr(1.0 / res, t, cw * ch), cu = getcap(sh, t, cw * ch);
                    if (ent[c].empty()) {
                        if (occ >= kOccupancyThreshold) {
                            bm.rx[c] = rxu; bm.ry[c] = ryu; bm.rz[c] = rzu; bm.lock[c] = 1;
                        } else {
                            bm.rx[c] = 1 / ((1 / rxu) * occ);
                            bm.ry[c] = 1 / ((1 / ryu) * occ);
                            bm.rz[c] = 1 / ((1 / rzu) * occ);
                            bm.lock[c] = 0;
                        }
                        bm.cap[c] = cu;
                    } else if (!bm.lock[c]) {
                        if (occ >= kOccupancyThreshold) {
                            bm.rx[c] = rxu; bm.ry[c] = ryu; bm.rz[c] = rzu; bm.cap[c] = cu;
                            bm.lock[c] = 1;
                        } else {
                            bm.rx[c] = 1 / ((1 / bm.rx[c]) + (1 / rxu) * occ);
                            bm.ry[c] = 1 / ((1 / bm.ry[c]) + (1 / ryu) * occ);
                            bm.rz[c] = 1 / ((1 / bm.rz[c]) + (1 / rzu) * occ);
                            bm.lock[c] = 0;
                        }
                    }
                }
                if (ent[c].empty()) bm.has_res[c] = 1;
                ent[c].emplace_back(u, occ);
            }
    }
    bm.cell_ptr.assign((size_t)nr * nc + 1, 0);
    bm.ent_unit.clear();
    bm.ent_occ.clear();
    for (size_t c = 0; c < ent.size(); c++) {
        if (ent[c].empty())
            fatal("grid cell not covered by any floorplan block\n");
        for (auto &e : ent[c]) {
            bm.ent_unit.push_back(e.first);
            bm.ent_occ.push_back(e.second);
        }
        bm.cell_ptr[c + 1] = (int)bm.ent_unit.size();
    }
    if (M.map_pkg[mi]) {
        std::fill(bm.has_res.begin(), bm.has_res.end(), 0);
        std::fill(bm.lock.begin(), bm.lock.end(), 0);
    }
}

void layer_bases(const Model &M, std::vector<int> &base) {
    base.assign(M.nl(), 0);
    int b = 0;
    for (int n = 0; n < M.nl(); n++) {
        base[n] = b;
        b += M.layers[n].flp->n_units;
    }
}

static void alloc_grid_vector(grid_model_vector_t *&v, const Model &M) {
    v = (grid_model_vector_t *)calloc(1, sizeof(grid_model_vector_t));
    const int nl = M.nl(), nr = M.nr(), nc = M.nc();
    v->cuboid = (double ***)calloc(nl, sizeof(double **));
    v->cuboid[0] = (double **)calloc((size_t)nl * nr, sizeof(double *));
    v->cuboid[0][0] = (double *)calloc((size_t)nl * nr * nc + M.nx, sizeof(double));
    for (int n = 1; n < nl; n++) v->cuboid[n] = v->cuboid[0] + (size_t)nr * n;
    for (int n = 0; n < nl; n++)
        for (int i = 0; i < nr; i++) v->cuboid[n][i] = v->cuboid[0][0] + (size_t)nr * nc * n + (size_t)nc * i;
    v->extra = v->cuboid[0][0] + (size_t)nl * nr * nc;
}

static void free_grid_vector(grid_model_vector_t *v) {
    if (!v) return;
    free(v->cuboid[0][0]);
    free(v->cuboid[0]);
    free(v->cuboid);
    free(v);
}

static void refresh_mirror(Model &M, int which) {
    if (which == 0 && !M.mirror_steady && M.steady_valid) {
        dev_get_grid(M, 0, M.g->last_steady->cuboid[0][0]);
        M.mirror_steady = 1;
    }
    if (which == 1 && !M.mirror_trans && M.trans_valid) {
        dev_get_grid(M, 1, M.g->last_trans->cuboid[0][0]);
        M.mirror_trans = 1;
    }
}

// stamps of model coefficients: unique over the process, so a device state
// kept for a model is recognised as stale after any change (or when another
// model takes the same address)
static unsigned long next_stamp() {
    static unsigned long s = 0;
    return ++s;
}

static void delete_model(Model *M) {
    if (!M) return;
    dev_release(*M);
    for (auto *f : M->owned_flps) free_flp_impl(f);
    for (auto *mc : M->mc)
        if (mc) {
            if (mc->cfg.cell_types) {
                for (int i = 0; i < mc->cfg.num_rows; i++) free(mc->cfg.cell_types[i]);
                free(mc->cfg.cell_types);
            }
            delete mc;
        }
    for (auto &L : M->layers) free(L.g2bmap);
    free_grid_vector(M->g->last_steady);
    free_grid_vector(M->g->last_trans);
    delete M;
}

static int total_nodes(const Model &M) { return M.g->total_n_blocks + M.nx; }

static const char *layer_prefix(const Model &M, int n, char *buf) {
    const grid_model_t &g = *M.g;
    const bool sec = g.config.model_secondary;
    if (n == M.spidx) return "hsp_";
    if (n == M.hsidx) return "hsink_";
    if (sec && n == M.subidx) return "sub_";
    if (sec && n == M.solderidx) return "solder_";
    if (sec && n == M.pcbidx) return "pcb_";
    if (g.has_lcf) {
        snprintf(buf, 32, "layer_%d_", n);
        return buf;
    }
    const int si = sec ? 5 : 0, tim = sec ? 6 : 1;
    if (n == si) return "";
    if (n == tim) return "iface_";
    if (sec && n == 4) return "metal_";
    if (sec && n == 3) return "c4_";
    fatal("unknown layer\n");
}

}  // namespace hs

using namespace hs;

extern "C" {

int hs_last_error(char *buf, int len) {
    int had = g_has_error;
    if (buf && len > 0) snprintf(buf, len, "%s", had ? g_last_error.c_str() : "");
    g_has_error = 0;
    g_last_error.clear();
    return had;
}

void hs_set_fatal_mode(int exit_on_error) { g_exit_on_error = exit_on_error ? 1 : 0; }

// ---- configuration (temperature.c:19-109; microchannel.c:17-43) ----------
thermal_config_t default_thermal_config(void) {
    thermal_config_t c;
    memset(&c, 0, sizeof c);
    c.t_chip = 0.15e-3;
    c.k_chip = 100.0;
    c.p_chip = 1.75e6;
    c.thermal_threshold = 81.8 + 273.15;
    c.c_convec = 140.4;
    c.r_convec = 0.1;
    c.s_sink = 60e-3;
    c.t_sink = 6.9e-3;
    c.k_sink = 400.0;
    c.p_sink = 3.55e6;
    c.s_spreader = 30e-3;
    c.t_spreader = 1e-3;
    c.k_spreader = 400.0;
    c.p_spreader = 3.55e6;
    c.t_interface = 20e-6;
    c.k_interface = 4.0;
    c.p_interface = 4.0e6;
    c.model_secondary = 0;
    c.r_convec_sec = 1.0;
    c.c_convec_sec = 140.4;
    c.n_metal = 8;
    c.t_metal = 10.0e-6;
    c.t_c4 = 0.0001;
    c.s_c4 = 20.0e-6;
    c.n_c4 = 400;
    c.s_sub = 0.021;
    c.t_sub = 0.001;
    c.s_solder = 0.021;
    c.t_solder = 0.00094;
    c.s_pcb = 0.1;
    c.t_pcb = 0.002;
    c.ambient = 45 + 273.15;
    strcpy(c.init_file, "(null)");
    c.init_temp = 60 + 273.15;
    strcpy(c.steady_file, "(null)");
    strcpy(c.grid_transient_file, "(null)");
    c.sampling_intvl = 3.333e-6;
    c.base_proc_freq = 3e9;
    c.dtm_used = 0;
    c.leakage_used = 0;
    c.leakage_mode = 0;
    c.package_model_used = 0;
    strcpy(c.package_config_file, "(null)");
    strcpy(c.model_type, "block");
    c.block_omit_lateral = 0;
    c.grid_rows = 64;
    c.grid_cols = 64;
    strcpy(c.grid_layer_file, "(null)");
    strcpy(c.grid_steady_file, "(null)");
    strcpy(c.grid_map_mode, "center");
    c.detailed_3D_used = 0;
    return c;
}

microchannel_config_t default_microchannel_config(void) {
    microchannel_config_t c;
    memset(&c, 0, sizeof c);
    c.cell_width = 100e-6;
    c.cell_height = 100e-6;
    c.cell_thickness = 100e-6;
    c.pumping_pressure = 5000;
    c.pump_internal_res = 0;
    c.inlet_temperature = 300;
    c.coolant_capac = 4172638;
    c.coolant_res = 1.647717911;
    c.coolant_visc = 0.000889;
    c.wall_capac = 1635660;
    c.wall_res = 0.0076923077;
    c.htc = 27132;
    c.num_rows = -1;
    c.num_columns = -1;
    c.n_fluid_cells = -1;
    return c;
}

void default_materials(materials_list_t *ml) {
    if (ml) {
        ml->size = 0;
        ml->names = nullptr;
        ml->material_properties = nullptr;
    }
}

void materials_add_from_file(materials_list_t *ml, char *file) {
    HS_BEGIN
    auto lines = read_lines(file);
    std::vector<std::string> names;
    std::vector<material_t> props;
    int field = 0;
    material_t m{};
    std::string name, kind;
    for (auto &l : lines) {
        std::string tok;
        if (!first_token(l.c_str(), tok)) continue;
        double dv = 0;
        switch (field) {
            case 0: name = tok; field = 1; break;
            case 1:
                kind = tok;
                if (kind == "solid") m.material_type = 0;
                else if (kind == "fluid") m.material_type = 1;
                else fatal("invalid material type: must be solid or fluid\n");
                field = 2;
                break;
            case 2:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) fatal("invalid material thermal conductivity\n");
                if (dv < 0) fatal("thermal conductivities must be nonnegative\n");
                m.thermal_conductivity = dv;
                field = 3;
                break;
            case 3:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) fatal("invalid material volumetric heat capacity\n");
                if (dv < 0) fatal("volumetric heat capacities must be nonnegative\n");
                m.volumetric_heat_capacity = dv;
                if (kind == "solid") {
                    m.dynamic_viscosity = 0.0;
                    names.push_back(name);
                    props.push_back(m);
                    field = 0;
                } else {
                    field = 4;
                }
                break;
            case 4:
                if (sscanf(tok.c_str(), "%lf", &dv) != 1) fatal("invalid material dynamic viscosity\n");
                if (dv < 0) fatal("dynamic viscosities must be nonnegative\n");
                m.dynamic_viscosity = dv;
                names.push_back(name);
                props.push_back(m);
                field = 0;
                break;
        }
    }
    ml->size = (int)names.size();
    ml->names = (char **)calloc(names.size() ? names.size() : 1, sizeof(char *));
    ml->material_properties = (material_t *)calloc(names.size() ? names.size() : 1, sizeof(material_t));
    for (size_t i = 0; i < names.size(); i++) {
        ml->names[i] = (char *)calloc(1, STR_SIZE);
        snprintf(ml->names[i], STR_SIZE, "%s", names[i].c_str());
        ml->material_properties[i] = props[i];
    }
    HS_END()
}

void free_materials(materials_list_t *ml) {
    if (!ml) return;
    for (int i = 0; i < ml->size; i++) free(ml->names[i]);
    free(ml->names);
    free(ml->material_properties);
    ml->names = nullptr;
    ml->material_properties = nullptr;
    ml->size = 0;
}

// ---- floorplans -----------------------------------------------------------
flp_t *read_flp(char *file, int read_connects, int initialize_connects) {
    HS_BEGIN
    return flp_from_lines(read_lines(file), read_connects, initialize_connects);
    HS_END(nullptr)
}

void free_flp(flp_t *flp, int compacted, int free_connects) {
    (void)compacted;
    (void)free_connects;
    free_flp_impl(flp);
}

int get_blk_index(flp_t *flp, char *name) {
    HS_BEGIN
    if (!flp) fatal("null pointer in get_blk_index\n");
    for (int i = 0; i < flp->n_units; i++)
        if (!strcasecmp(name, flp->units[i].name)) return i;
    fatal(std::string("block ") + name + " not found\n");
    HS_END(-1)
}

// ---- model lifecycle (temperature.c:616-653, temperature_grid.c:896-1114) --
RC_model_t *alloc_RC_model(thermal_config_t *config, flp_t *placeholder,
                           microchannel_config_t *microchannel_config,
                           materials_list_t *materials_list, int do_detailed_3D,
                           int use_microchannels) {
    Handle *h = nullptr;
    Model *M = nullptr;
    HS_BEGIN
    if (!strcasecmp(config->model_type, "block")) return block_rc_alloc(config, placeholder);
    if (strcasecmp(config->model_type, "grid")) fatal("unknown model type\n");
    // Any grid size, as the reference's SuperLU build takes (its iterative
    // build asks for powers of two, temperature_grid.c:901-905, for its own
    // multigrid): the tile kernels cover partial tiles, and multigrid
    // coarsens while both sizes stay even.
    if (config->grid_rows <= 0 || config->grid_cols <= 0)
        fatal("grid rows and columns must be positive\n");
    h = (Handle *)calloc(1, sizeof(Handle));
    M = new Model();
    M->stamp = next_stamp();
    h->impl = M;
    M->g = &h->pub;
    grid_model_t &g = h->pub;
    g.config = *config;
    g.rows = config->grid_rows;
    g.cols = config->grid_cols;
    g.use_microchannels = use_microchannels;
    g.default_microchannel_config = microchannel_config;
    if (do_detailed_3D) g.config.detailed_3D_used = 1;
    M->detailed = g.config.detailed_3D_used != 0;
    const char *mm = g.config.grid_map_mode;
    if (!strcasecmp(mm, "avg")) g.map_mode = GRID_AVG;
    else if (!strcasecmp(mm, "min")) g.map_mode = GRID_MIN;
    else if (!strcasecmp(mm, "max")) g.map_mode = GRID_MAX;
    else if (!strcasecmp(mm, "center")) g.map_mode = GRID_CENTER;
    else fatal("unknown mapping mode\n");
    g.has_lcf = strcmp(g.config.grid_layer_file, "(null)") != 0;
    if (!g.has_lcf) {
        if (!placeholder) fatal("Either LCF or FLP file must be specified\n");
        g.base_n_units = placeholder->n_units;
    }
    if (M->detailed && !g.has_lcf)
        fatal("-do_detailed_3D can only be used in 3D mode (if a grid_layer_file is specified)\n");
    if (use_microchannels && !M->detailed)
        fatal("-use_microchannels requires -model_type grid and do_detailed_3D on options\n");
    build_layers(*M, placeholder, materials_list);
    alloc_grid_vector(g.last_steady, *M);
    alloc_grid_vector(g.last_trans, *M);
    M->opt.precision = HS_PREC_FP64;
    M->opt.steady_solver = HS_STEADY_JPCG;
    M->opt.transient = HS_TRANS_RK4;
    M->opt.steady_rtol = 1e-12;
    M->opt.max_iters = 200000;
    M->opt.taylor_tol = 1e-7;
    M->opt.steady_warm_start = 0;
    h->rc.grid = &h->pub;
    h->rc.type = GRID_MODEL;
    h->rc.config = &h->pub.config;
    return &h->rc;
    } catch (const hs::Error &e) {
        if (M) delete M;
        free(h);
        hs::report(e.msg);
        return nullptr;
    }
}

void delete_RC_model(RC_model_t *model) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        delete_block_model(model->block);
        return;
    }
    if (!model) return;
    Model *M = model_of(model);
    delete_model(M);
    free(reinterpret_cast<Handle *>(model));
    HS_END()
}

void populate_R_model(RC_model_t *model, flp_t *flp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        populate_R_model_block(model->block, flp);
        return;
    }
    Model &M = *model_of(model);
    grid_model_t &g = *M.g;
    const int nl = g.n_layers;
    if (!g.has_lcf) {
        int si = g.config.model_secondary ? 5 : 0;
        if (flp != g.layers[si].flp) fatal("mismatch between the floorplan and the thermal model\n");
        g.width = total_width(flp);
        g.height = total_height(flp);
    }
    for (size_t mi = 0; mi < M.maps.size(); mi++) build_map(M, (int)mi);
    for (int n = 0; n < nl; n++) {
        const BlockMap &bm = M.maps[M.layer_map[n]];
        free(g.layers[n].g2bmap);
        g.layers[n].g2bmap = (glist_t *)calloc(bm.ranges.size(), sizeof(glist_t));
        std::copy(bm.ranges.begin(), bm.ranges.end(), g.layers[n].g2bmap);
    }
    if (g.width > g.config.s_sink || g.height > g.config.s_sink || g.width > g.config.s_spreader ||
        g.height > g.config.s_spreader)
        fatal("inordinate floorplan size!\n");
    const double cw = g.width / g.cols, ch = g.height / g.rows;
    package_R(g.pack, g.config, g.width, g.height);
    for (int n = 0; n < nl; n++) {
        layer_t &L = g.layers[n];
        if (L.has_lateral) {
            L.rx = getr(L.k, cw, ch * L.thickness);
            L.ry = getr(L.k, ch, cw * L.thickness);
        } else {
            L.rx = L.ry = kLargeNum;
        }
        L.rz = getr(L.k, L.thickness, cw * ch);
        if (n == M.hsidx) L.rz += g.config.r_convec * (g.config.s_sink * g.config.s_sink) / (cw * ch);
    }
    g.r_ready = 1;
    M.dev_dirty = true;
    M.stamp = next_stamp();
    HS_END()
}

void populate_C_model(RC_model_t *model, flp_t *flp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        populate_C_model_block(model->block, flp);
        return;
    }
    Model &M = *model_of(model);
    grid_model_t &g = *M.g;
    if (!g.r_ready) fatal("R model not ready\n");
    if (!g.has_lcf && flp != g.layers[g.config.model_secondary ? 5 : 0].flp)
        fatal("different floorplans for R and C models!\n");
    const double cw = g.width / g.cols, ch = g.height / g.rows;
    package_C(g.pack, g.config, g.width, g.height);
    for (int n = 0; n < g.n_layers; n++)
        g.layers[n].c = getcap(g.layers[n].sp, g.layers[n].thickness, cw * ch);
    g.layers[M.hsidx].c += kCFactor * g.config.c_convec * (cw * ch) / (g.config.s_sink * g.config.s_sink);
    if (g.config.model_secondary)
        g.layers[M.pcbidx].c += kCFactor * g.config.c_convec_sec * (cw * ch) / (g.config.s_pcb * g.config.s_pcb);
    g.c_ready = 1;
    M.dev_dirty = true;
    M.stamp = next_stamp();
    HS_END()
}

// ---- vectors (temperature_grid.c:1182-1259, :1449-1455) --------------------
double *hotspot_vector(RC_model_t *model) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) return hotspot_vector_block(model->block);
    Model &M = *model_of(model);
    if (M.g->total_n_blocks <= 0) fatal("total_n_blocks is not greater than zero\n");
    double *v = (double *)calloc(total_nodes(M), sizeof(double));
    if (!v) fatal("memory allocation error\n");
    return v;
    HS_END(nullptr)
}

void free_dvector(double *v) { free(v); }

void set_temp(RC_model_t *model, double *temp, double val) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        set_temp_block(model->block, temp, val);
        return;
    }
    Model &M = *model_of(model);
    for (int i = 0; i < total_nodes(M); i++) temp[i] = val;
    HS_END()
}

void copy_temp(RC_model_t *model, double *dst, double *src) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        copy_temp_block(model->block, dst, src);
        return;
    }
    Model &M = *model_of(model);
    memcpy(dst, src, sizeof(double) * total_nodes(M));
    HS_END()
}

void trim_hotspot_vector(RC_model_t *model, double *dst, double *src, int at, int size) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        trim_hotspot_vector_block(model->block, dst, src, at, size);
        return;
    }
    Model &M = *model_of(model);
    if (M.g->has_lcf) fatal("trim_hotspot_vector_grid called with lcf file\n");
    const int tn = total_nodes(M);
    for (int i = 0; i < at && i < tn; i++) dst[i] = src[i];
    for (int i = at + size; i < tn; i++) dst[i - size] = src[i];
    HS_END()
}

void resize_thermal_model(RC_model_t *model, int n_units) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        resize_thermal_model_block(model->block, n_units);
        return;
    }
    Model &M = *model_of(model);
    if (M.g->has_lcf) fatal("resize_thermal_model_grid called with lcf file\n");
    if (n_units > M.g->base_n_units) fatal("resizing grid model to more than the allocated space\n");
    M.g->total_n_blocks = 0;
    for (int i = 0; i < M.nl(); i++) M.g->total_n_blocks += M.layers[i].flp->n_units;
    M.dev_dirty = true;
    M.stamp = next_stamp();
    HS_END()
}

// ---- text formats (temperature_grid.c:1262-1764) ---------------------------
void dump_temp(RC_model_t *model, double *temp, char *file) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        dump_temp_block(model->block, temp, file);
        return;
    }
    Model &M = *model_of(model);
    FILE *fp = !strcasecmp(file, "stdout") ? stdout : !strcasecmp(file, "stderr") ? stderr : fopen(file, "w");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for writing\n");
    int base = 0;
    char buf[32];
    for (int n = 0; n < M.nl(); n++) {
        const char *pre = layer_prefix(M, n, buf);
        const flp_t *f = M.layers[n].flp;
        for (int i = 0; i < f->n_units; i++) fprintf(fp, "%s%s\t%.2f\n", pre, f->units[i].name, temp[base + i]);
        base += f->n_units;
    }
    for (int i = 0; i < M.nx; i++) fprintf(fp, "inode_%d\t%.2f\n", i, temp[base + i]);
    if (fp != stdout && fp != stderr) fclose(fp);
    HS_END()
}

void read_temp(RC_model_t *model, double *temp, char *file, int clip) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        read_temp_block(model->block, temp, file, clip);
        return;
    }
    Model &M = *model_of(model);
    grid_model_t &g = *M.g;
    auto lines = read_lines(file);
    size_t li = 0;
    auto next = [&](std::string &out) {
        while (li < lines.size()) {
            std::string tok;
            const std::string &l = lines[li++];
            if (first_token(l.c_str(), tok)) {
                out = l;
                return;
            }
        }
        fatal("not enough lines in temperature file\n");
    };
    int base = 0;
    double mx = 0;
    char buf[32], name[STR_SIZE];
    for (int n = 0; n < M.nl(); n++) {
        std::string fmt = std::string(layer_prefix(M, n, buf)) + "%511s%lf";
        flp_t *f = M.layers[n].flp;
        for (int i = 0; i < f->n_units; i++) {
            std::string l;
            next(l);
            double val;
            if (sscanf(l.c_str(), fmt.c_str(), name, &val) != 2) fatal("invalid temperature file format\n");
            int idx = get_blk_index(f, name);
            if (idx < 0) fatal("unit in temperature file not found in floorplan\n");
            temp[base + idx] = val;
            // the reference tracks the maximum over layer 0 (read_temp_grid :1594-1595)
            if (n == 0 && temp[idx] > mx) mx = temp[idx];
        }
        base += f->n_units;
    }
    for (int i = 0; i < M.nx; i++) {
        std::string l;
        next(l);
        double val;
        if (sscanf(l.c_str(), "%511s%lf", name, &val) != 2) fatal("invalid temperature file format\n");
        char want[32];
        snprintf(want, sizeof want, "inode_%d", i);
        if (strcasecmp(want, name)) fatal("invalid temperature file format\n");
        temp[base + i] = val;
    }
    for (; li < lines.size(); li++) {
        if (!lines[li].empty()) fatal("too many lines in temperature file\n");
    }
    if (clip && mx > g.config.thermal_threshold) {
        double factor = (g.config.thermal_threshold - g.config.ambient) / (mx - g.config.ambient);
        for (int i = 0; i < total_nodes(M); i++) temp[i] = (temp[i] - g.config.ambient) * factor + g.config.ambient;
    }
    HS_END()
}

void dump_power(RC_model_t *model, double *power, char *file) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        dump_power_block(model->block, power, file);
        return;
    }
    Model &M = *model_of(model);
    FILE *fp = !strcasecmp(file, "stdout") ? stdout : !strcasecmp(file, "stderr") ? stderr : fopen(file, "w");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for writing\n");
    int base = 0;
    for (int n = 0; n < M.nl(); n++) {
        const flp_t *f = M.layers[n].flp;
        if (M.layers[n].has_power)
            for (int i = 0; i < f->n_units; i++) {
                if (M.g->has_lcf) fprintf(fp, "layer_%d_%s\t%.6f\n", n, f->units[i].name, power[base + i]);
                else fprintf(fp, "%s\t%.6f\n", f->units[i].name, power[base + i]);
            }
        base += f->n_units;
    }
    if (fp != stdout && fp != stderr) fclose(fp);
    HS_END()
}

void read_power(RC_model_t *model, double *power, char *file) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        read_power_block(model->block, power, file);
        return;
    }
    Model &M = *model_of(model);
    auto lines = read_lines(file);
    char name[STR_SIZE];
    double val;
    if (M.g->has_lcf) {
        size_t li = 0;
        int base = 0;
        for (int n = 0; n < M.nl(); n++) {
            flp_t *f = M.layers[n].flp;
            if (M.layers[n].has_power) {
                char fmt[64];
                snprintf(fmt, sizeof fmt, "layer_%d_%%511s%%lf", n);
                for (int i = 0; i < f->n_units; i++) {
                    std::string tok;
                    while (li < lines.size() && !first_token(lines[li].c_str(), tok)) li++;
                    if (li >= lines.size()) fatal("not enough lines in power file\n");
                    if (sscanf(lines[li++].c_str(), fmt, name, &val) != 2) fatal("invalid power file format\n");
                    int idx = get_blk_index(f, name);
                    if (idx < 0) fatal("unit in power file not found in floorplan\n");
                    power[base + idx] = val;
                }
            }
            base += f->n_units;
        }
        for (; li < lines.size(); li++)
            if (!lines[li].empty()) fatal("too many lines in power file\n");
    } else {
        // default stack: names resolve against layer 0's floorplan and land at
        // that index of the vector (read_power_grid :1749-1760)
        for (auto &l : lines) {
            std::string tok;
            if (!first_token(l.c_str(), tok)) continue;
            if (sscanf(l.c_str(), "%511s%lf", name, &val) != 2) fatal("invalid power file format\n");
            int idx = get_blk_index(M.layers[0].flp, name);
            if (idx < 0) fatal("unit in power file not found in floorplan\n");
            power[idx] = val;
        }
    }
    HS_END()
}

double find_max_temp(RC_model_t *model, double *temp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) return find_max_temp_block(model->block, temp);
    Model &M = *model_of(model);
    int si = !M.g->config.model_secondary ? 0 : (M.g->has_lcf ? 3 : 5);
    double mx = 0.0;
    // the reference scans the first n_units entries (those of layer 0)
    for (int i = 0; i < M.layers[si].flp->n_units; i++) {
        if (temp[i] < 0) fatal("negative temperature!\n");
        if (mx < temp[i]) mx = temp[i];
    }
    return mx;
    HS_END(0.0)
}

double find_avg_temp(RC_model_t *model, double *temp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) return find_avg_temp_block(model->block, temp);
    Model &M = *model_of(model);
    int base = 0, count = 0;
    double sum = 0.0;
    for (int n = 0; n < M.nl(); n++) {
        const int nu = M.layers[n].flp->n_units;
        if (M.layers[n].has_power) {
            for (int i = 0; i < nu; i++) {
                if (temp[base + i] < 0) fatal("negative temperature!\n");
                sum += temp[base + i];
            }
            count += nu;
        }
        base += nu;
    }
    if (!count) fatal("no power dissipating units?!\n");
    return sum / count;
    HS_END(0.0)
}

double calc_sink_temp(RC_model_t *model, double *temp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) return calc_sink_temp_block(model->block, temp, model->config);
    return sink_temp(sink_model(*model_of(model)), temp);
    HS_END(0.0)
}

// leakage power of a block (calc_leakage, temperature.c:910-930); throws
// inside the library's own loops, so the error reaches their caller
static double leakage_of(int mode, double h, double w, double temp) {
    if (mode) fatal("HotLeakage currently is not implemented in this release of HotSpot, please check back later.\n");
    return 1.5e+4 * h * w * std::exp(0.036 * (temp - 383.15));
}

double calc_leakage(int mode, double h, double w, double temp) {
    HS_BEGIN
    return leakage_of(mode, h, w, temp);
    HS_END(0.0)
}

void dump_steady_temp_grid(grid_model_t *model, char *file) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (!M.g->r_ready) fatal("R model not ready\n");
    refresh_mirror(M, 0);
    FILE *fp = !strcasecmp(file, "stdout") ? stdout : !strcasecmp(file, "stderr") ? stderr : fopen(file, "w");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for writing\n");
    for (int n = 0; n < M.nl(); n++) {
        fprintf(fp, "Layer %d:\n", n);
        for (int i = 0; i < M.nr(); i++)
            for (int j = 0; j < M.nc(); j++)
                fprintf(fp, "%d\t%.2f\n", i * M.nc() + j, M.g->last_steady->cuboid[n][i][j]);
    }
    if (fp != stdout && fp != stderr) fclose(fp);
    HS_END()
}

void dump_transient_temp_grid(grid_model_t *model, int trace_num, double sampling_intvl, char *filename) {
    HS_BEGIN
    Model &M = *model_of(model);
    refresh_mirror(M, 1);
    FILE *fp = fopen(filename, "a");
    if (!fp) fatal(std::string("Unable to open ") + filename + " for appending\n");
    fprintf(fp, "t = %lf\n", trace_num * sampling_intvl);
    for (int n = 0; n < M.nl(); n++) {
        fprintf(fp, "Layer %d:\n", n);
        for (int i = 0; i < M.nr(); i++)
            for (int j = 0; j < M.nc(); j++)
                fprintf(fp, "%d\t%.2f\n", i * M.nc() + j, M.g->last_trans->cuboid[n][i][j]);
    }
    fclose(fp);
    HS_END()
}

// ---- solves (temperature.c:656-761) ---------------------------------------
static void steady_grid(Model &M, double *power, double *temp) {
    if (!M.g->r_ready) fatal("R model not ready\n");
    // package nodes carry no power: the caller's array is written
    // (set_internal_power_grid, temperature_grid.c:2018-2024)
    for (int i = 0; i < M.nx; i++) power[M.g->total_n_blocks + i] = 0.0;
    dev_steady(M, power, temp);
    M.steady_valid = 1;
    M.mirror_steady = 0;
}

// temperature-leakage fixed point of the block model (temperature.c:670-705)
static void steady_block(block_model_t *b, double *power, double *temp) {
    const thermal_config_t &c = b->config;
    if (!c.leakage_used) {
        block_steady(b, power, temp);
        return;
    }
    const int n = b->flp->n_units;
    std::vector<double> temp_old(b->n_nodes, 0.0), power_new(b->n_nodes, 0.0);
    bool conv = false;
    for (int it = 0; !conv && it <= 100; it++) {
        for (int i = 0; i < n; i++) {
            const unit_t &u = b->flp->units[i];
            power_new[i] = power[i] + leakage_of(c.leakage_mode, u.height, u.width, temp[i]);
            temp_old[i] = temp[i];
        }
        block_steady(b, power_new.data(), temp);
        double dmax = 0.0;
        for (int i = 0; i < n; i++) dmax = std::max(dmax, temp[i] - temp_old[i]);
        if (dmax < 0.01) conv = true;
        if (dmax > 500.0 && it > 1)
            fatal("temperature is too high, possible thermal runaway. Double-check power inputs and package settings.\n");
    }
    if (!conv)
        fatal("too many iterations before temperature-leakage convergence -- possible thermal runaway\n");
}

void steady_state_temp(RC_model_t *model, double *power, double *temp) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        steady_block(model->block, power, temp);
        return;
    }
    Model &M = *model_of(model);
    grid_model_t &g = *M.g;
    if (!g.config.leakage_used) {
        steady_grid(M, power, temp);
        return;
    }
    // temperature-leakage fixed point (temperature.c:708-746); each re-solve
    // starts from the previous one
    const int tn = total_nodes(M);
    const int warm = M.opt.steady_warm_start;
    M.opt.steady_warm_start = 1;
    struct Restore {
        Model &m;
        int w;
        ~Restore() { m.opt.steady_warm_start = w; }
    } restore{M, warm};
    std::vector<double> temp_old(tn, 0.0), power_new(tn, 0.0);
    const LeakBlocks lb = leak_blocks(M);
    bool conv = false;
    for (int it = 0; !conv && it <= kLeakMaxIter; it++) {
        leakage_step(lb, power, temp, power_new.data(), temp_old.data());
        steady_grid(M, power_new.data(), temp);
        const double dmax = leakage_rise(lb, temp, temp_old.data());
        if (dmax < kLeakTol) conv = true;
        if (dmax > kLeakTempHigh && it > 0)
            fatal("temperature is too high, possible thermal runaway. Double-check power inputs and package settings.\n");
    }
    if (!conv)
        fatal("too many iterations before temperature-leakage convergence -- possible thermal runaway\n");
    HS_END()
}

extern "C++" {
namespace hs {

SinkModel sink_model(const Model &M) {
    const grid_model_t &g = *M.g;
    SinkModel s;
    s.c = g.config;
    s.width = g.width;
    s.height = g.height;
    s.rows = g.rows;
    s.cols = g.cols;
    s.hsidx = M.hsidx;
    s.k = g.layers[M.hsidx].k;
    s.thickness = g.layers[M.hsidx].thickness;
    for (int n = 0; n < M.hsidx; n++) s.base += M.layers[n].flp->n_units;
    const flp_t *f = M.layers[M.hsidx].flp;
    for (int i = 0; i < f->n_units; i++) s.area.push_back(f->units[i].width * f->units[i].height);
    s.total_n_blocks = g.total_n_blocks;
    return s;
}

// area-weighted average over the sink layer's blocks and the sink's package
// nodes, over the sink's area
double sink_temp(const SinkModel &s, const double *temp) {
    const thermal_config_t &c = s.c;
    const double spr = c.s_spreader * c.s_spreader, sink = c.s_sink * c.s_sink;
    double sum = 0.0;
    for (size_t q = 0; q < s.area.size(); q++) {
        if (temp[s.base + q] < 0) fatal("negative temperature!\n");
        sum += temp[s.base + q] * s.area[q];
    }
    const int base = s.total_n_blocks;
    for (int i = SINK_C_W; i <= SINK_C_E; i++) {
        if (temp[i + base] < 0) fatal("negative temperature!\n");
        sum += temp[i + base] * 0.25 * (c.s_spreader + s.height) * (c.s_spreader - s.width);
    }
    for (int i = SINK_C_N; i <= SINK_C_S; i++) {
        if (temp[i + base] < 0) fatal("negative temperature!\n");
        sum += temp[i + base] * 0.25 * (c.s_spreader + s.width) * (c.s_spreader - s.height);
    }
    for (int i = SINK_W; i <= SINK_S; i++) {
        if (temp[i + base] < 0) fatal("negative temperature!\n");
        sum += temp[i + base] * 0.25 * (sink - spr);
    }
    return sum / sink;
}

// the same expressions as populate_R_model's for the sink layer and the package
void sink_resistances(const SinkModel &s, double r, double &rz, package_RC_t &pk) {
    thermal_config_t c = s.c;
    c.r_convec = r;
    package_R(pk, c, s.width, s.height);
    const double cw = s.width / s.cols, ch = s.height / s.rows;
    rz = getr(s.k, s.thickness, cw * ch);
    rz += c.r_convec * (c.s_sink * c.s_sink) / (cw * ch);
}

LeakBlocks leak_blocks(const Model &M) {
    LeakBlocks b;
    b.mode = M.g->config.leakage_mode;
    int base = 0;
    for (int k = 0; k < M.nl(); k++) {
        const flp_t *f = M.layers[k].flp;
        if (M.layers[k].has_power)
            for (int j = 0; j < f->n_units; j++) {
                b.idx.push_back(base + j);
                b.h.push_back(f->units[j].height);
                b.w.push_back(f->units[j].width);
            }
        base += f->n_units;
    }
    return b;
}

void leakage_step(const LeakBlocks &b, const double *power, const double *temp, double *power_new,
                  double *temp_old) {
    for (size_t q = 0; q < b.idx.size(); q++) {
        const int i = b.idx[q];
        power_new[i] = power[i] + leakage_of(b.mode, b.h[q], b.w[q], temp[i]);
        temp_old[i] = temp[i];
    }
}

double leakage_rise(const LeakBlocks &b, const double *temp, const double *temp_old) {
    double dmax = 0.0;
    for (int i : b.idx) dmax = std::max(dmax, temp[i] - temp_old[i]);
    return dmax;
}
}  // namespace hs
}

void compute_temp(RC_model_t *model, double *power, double *temp, double time_elapsed) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        compute_temp_block(model->block, power, temp, time_elapsed);
        return;
    }
    Model &M = *model_of(model);
    grid_model_t &g = *M.g;
    if (!g.r_ready || !g.c_ready) fatal("grid model not ready\n");
    for (int i = 0; i < M.nx; i++) power[g.total_n_blocks + i] = 0.0;
    if (temp) g.last_temp = temp;
    if (!g.last_temp) fatal("compute_temp: no temperature vector to continue from\n");
    if (!temp && !M.trans_valid) fatal("compute_temp: no transient state to continue from\n");
    dev_transient(M, power, temp, time_elapsed, g.last_temp);
    M.trans_valid = 1;
    M.mirror_trans = 0;
    HS_END()
}

// ---- extensions -------------------------------------------------------------
void hs_get_options(RC_model_t *model, hs_options_t *opt) {
    HS_BEGIN
    *opt = model_of(model)->opt;
    HS_END()
}

void hs_set_options(RC_model_t *model, const hs_options_t *opt) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (opt->precision != HS_PREC_FP64 && opt->precision != HS_PREC_MIXED) fatal("unknown precision\n");
    if (opt->steady_solver != HS_STEADY_JPCG && opt->steady_solver != HS_STEADY_MGPCG)
        fatal("unknown steady solver\n");
    if (opt->transient != HS_TRANS_RK4 && opt->transient != HS_TRANS_TAYLOR && opt->transient != HS_TRANS_BE)
        fatal("unknown transient integrator\n");
    if (!(opt->steady_rtol > 0) || opt->max_iters <= 0 || !(opt->taylor_tol > 0)) fatal("invalid solver tolerances\n");
    M.opt = *opt;
    HS_END()
}

void hs_set_outlet_rule(RC_model_t *model, int superlu) {
    HS_BEGIN
    if (!model || model->type == BLOCK_MODEL) fatal("the outlet rule applies to grid models\n");
    Model &M = *model_of(model);
    if (M.superlu_outlets == (superlu != 0)) return;
    M.superlu_outlets = superlu != 0;
    M.dev_dirty = true;
    M.stamp = next_stamp();
    HS_END()
}

void hs_get_stats(RC_model_t *model, hs_stats_t *stats) {
    HS_BEGIN
    if (model && model->type == BLOCK_MODEL) {
        *stats = block_of(model->block)->stats;
        return;
    }
    *stats = model_of(model)->stats;
    HS_END()
}

int hs_grid_size(RC_model_t *model) {
    HS_BEGIN
    return (int)model_of(model)->vec_len();
    HS_END(-1)
}

void hs_get_grid_temp(RC_model_t *model, int which, double *out) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (which != 0 && which != 1) fatal("which must be 0 (steady) or 1 (transient)\n");
    if (which == 0 && !M.steady_valid) fatal("no steady state has been computed\n");
    if (which == 1 && !M.trans_valid) fatal("no transient state has been computed\n");
    dev_get_grid(M, which, out);
    HS_END()
}

void hs_set_grid_temp(RC_model_t *model, int which, const double *in) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (which != 0 && which != 1) fatal("which must be 0 (steady) or 1 (transient)\n");
    dev_set_grid(M, which, in);
    if (which == 0) { M.steady_valid = 1; M.mirror_steady = 0; }
    else { M.trans_valid = 1; M.mirror_trans = 0; }
    HS_END()
}

void hs_transient_run(RC_model_t *model, const double *power_seq, int n_steps, double dt,
                      const double *init_temp, double *temp_seq) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (!M.g->r_ready || !M.g->c_ready) fatal("grid model not ready\n");
    if (n_steps <= 0) fatal("hs_transient_run: n_steps must be positive\n");
    dev_transient_run(M, power_seq, n_steps, dt, init_temp, temp_seq);
    M.trans_valid = 1;
    M.mirror_trans = 0;
    HS_END()
}

void hs_grid_slope(RC_model_t *model, const double *block_power, const double *grid_temp, double *grid_slope) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (!M.g->r_ready || !M.g->c_ready) fatal("grid model not ready\n");
    dev_slope(M, block_power, grid_temp, grid_slope);
    HS_END()
}

void hs_block_to_grid(RC_model_t *model, const double *block_vec, double *grid_vec, int type) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (!M.g->r_ready) fatal("R model not ready\n");
    if (type != V_POWER && type != V_TEMP) fatal("unknown vector type\n");
    dev_b2g(M, block_vec, grid_vec, type);
    HS_END()
}

void hs_grid_to_block(RC_model_t *model, const double *grid_vec, double *block_vec) {
    HS_BEGIN
    Model &M = *model_of(model);
    if (!M.g->r_ready) fatal("R model not ready\n");
    dev_g2b(M, grid_vec, block_vec);
    HS_END()
}


// ---- batched engine ---------------------------------------------------------
struct hs_engine_st {
    hs::Engine *impl;
};

static hs::Engine &engine_of(hs_engine_t *e) {
    if (!e || !e->impl) fatal("null engine\n");
    return *e->impl;
}

hs_engine_t *hs_engine_create(int precision, int transient) {
    HS_BEGIN
    if (precision != HS_PREC_FP64 && precision != HS_PREC_MIXED) fatal("unknown precision\n");
    if (transient != HS_TRANS_RK4 && transient != HS_TRANS_TAYLOR && transient != HS_TRANS_BE)
        fatal("unknown transient integrator\n");
    std::unique_ptr<hs_engine_st> e(new hs_engine_st{nullptr});
    e->impl = eng_create(precision, transient);
    return e.release();
    HS_END(nullptr)
}

void hs_engine_destroy(hs_engine_t *e) {
    HS_BEGIN
    if (!e) return;
    eng_destroy(e->impl);
    delete e;
    HS_END()
}

void hs_engine_set_options(hs_engine_t *e, const hs_options_t *opt) {
    HS_BEGIN
    if (!opt) fatal("null options\n");
    if (opt->precision != HS_PREC_FP64 && opt->precision != HS_PREC_MIXED) fatal("unknown precision\n");
    if (opt->steady_solver != HS_STEADY_JPCG && opt->steady_solver != HS_STEADY_MGPCG)
        fatal("unknown steady solver\n");
    if (opt->transient != HS_TRANS_RK4 && opt->transient != HS_TRANS_TAYLOR && opt->transient != HS_TRANS_BE)
        fatal("unknown transient integrator\n");
    if (!(opt->steady_rtol > 0) || opt->max_iters <= 0 || !(opt->taylor_tol > 0)) fatal("invalid solver tolerances\n");
    eng_set_options(engine_of(e), *opt);
    HS_END()
}

void hs_engine_set_tick(hs_engine_t *e, int intervals, int steady_iters) {
    HS_BEGIN
    eng_set_tick(engine_of(e), intervals, steady_iters);
    HS_END()
}

void hs_engine_set_budget(hs_engine_t *e, double bytes) {
    HS_BEGIN
    eng_set_budget(engine_of(e), bytes);
    HS_END()
}

int hs_engine_submit(hs_engine_t *e, RC_model_t *model, const double *power, int power_rows,
                     const double *init_temp, int n_steps, double dt) {
    HS_BEGIN
    return eng_submit(engine_of(e), *model_of(model), power, power_rows, init_temp, n_steps, dt);
    HS_END(-1)
}

int hs_engine_submit_package(hs_engine_t *e, RC_model_t *model, const double *power,
                             const package_config_t *package) {
    HS_BEGIN
    if (!package) fatal("null package configuration\n");
    return eng_submit_package(engine_of(e), *model_of(model), power, *package);
    HS_END(-1)
}

int hs_engine_admit(hs_engine_t *e) {
    HS_BEGIN
    return eng_admit(engine_of(e));
    HS_END(-1)
}

int hs_engine_tick(hs_engine_t *e) {
    HS_BEGIN
    return eng_tick(engine_of(e));
    HS_END(-1)
}

int hs_engine_poll(hs_engine_t *e, int job, double *out) {
    HS_BEGIN
    return eng_poll(engine_of(e), job, out);
    HS_END(-1)
}

int hs_engine_result_seq(hs_engine_t *e, int job, double *out) {
    HS_BEGIN
    return eng_result_seq(engine_of(e), job, out);
    HS_END(-1)
}

int hs_engine_job_info(hs_engine_t *e, int job, hs_job_info_t *info) {
    HS_BEGIN
    if (!info) fatal("null job info\n");
    eng_job_info(engine_of(e), job, info);
    return 0;
    HS_END(-1)
}

void hs_engine_release(hs_engine_t *e, int job) {
    HS_BEGIN
    eng_release(engine_of(e), job);
    HS_END()
}

int hs_engine_active(hs_engine_t *e) {
    HS_BEGIN
    return eng_active(engine_of(e));
    HS_END(-1)
}

void hs_engine_get_stats(hs_engine_t *e, hs_engine_stats_t *stats) {
    HS_BEGIN
    if (!stats) fatal("null engine stats\n");
    eng_get_stats(engine_of(e), stats);
    HS_END()
}

}  // extern "C"
