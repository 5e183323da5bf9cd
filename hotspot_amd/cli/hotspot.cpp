// hotspot -- the reference's command-line thermal simulator (hotspot.c:368-735)
// for the grid and the block model, over libhotspot_amd (every solve runs on
// the GPU).
//
//   hotspot -f <flp> -p <ptrace> [-o <ttrace>] [-c <config>] [-d <dump>] [-<name> <value>]...
//
// Same options, files and output formats as the reference: the power trace is
// read, each row is one sampling interval of the transient (written to the
// -o trace with "%.2f" temperatures), and the average power of all rows gives
// the steady state (-steady_file, -grid_steady_file).  Without
// -grid_transient_file the whole trace runs as one device-resident sequence
// (hs_transient_run, equal to one compute_temp call per row); with it, row by
// row so the grid state can be written after each interval.
//
// With -package_model_used 1 the heat sink's convection resistance comes
// from the package model (its parameters in -package_config_file); with
// natural convection it is updated from the average sink temperature before
// every interval and iterated with the steady solve until it settles
// (hotspot.c:447-455, :558-566, :624-641).  The block model (model_type
// block) runs row by row with compute_temp, as in the reference.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>
#include <vector>

#include "hotspot_amd.h"

namespace {

constexpr int kMaxEntries = 512;   // util.h:30
// package.h:31-33: starting sink-to-ambient difference, r_convec tolerance of
// the natural-convection iteration, sink temperature of a thermal runaway
constexpr double kSmallForConvec = 0.01, kNaturalConvecTol = 0.01, kMaxSinkTemp = 1000;
const char *kNull = "(null)";

[[noreturn]] void die(const std::string &msg) {
    fprintf(stderr, "error: %s", msg.c_str());
    exit(1);
}

void usage(const char *prog) {
    printf("Usage: %s -f <file> -p <file> [-o <file>] [-c <file>] [-d <file>] [options]\n", prog);
    printf("Thermal simulation of a power trace on the GPU (grid or block model).\n");
    printf("   -f <file>\tfloorplan (e.g. ev6.flp); a -grid_layer_file overrides it\n");
    printf("   -p <file>\tpower trace input\n");
    printf("  [-o <file>]\ttransient temperature trace output; without it only the steady\n");
    printf("            \tstate is computed\n");
    printf("  [-c <file>]\tconfiguration parameters (e.g. hotspot.config)\n");
    printf("  [-d <file>]\twrite the configuration parameters in use to a file\n");
    printf("  [-materials_file <file>]\tmaterial properties\n");
    printf("  [-detailed_3D <on/off>]\theterogeneous per-cell R and C (needs a layer file)\n");
    printf("  [-use_microchannels <0/1>]\tmicrochannel cooling layers (needs detailed_3D on)\n");
    printf("  [options]\t\"-<name> <value>\" pairs override the configuration file,\n");
    printf("           \te.g. \"-model_type block\", \"-grid_rows 128\"\n");
    printf("  [-transient_solver <rk4|be|taylor>]\tgrid transient integrator: the reference's adaptive\n");
    printf("           \tRK4 (default), the backward-Euler step of its SuperLU build, or the Taylor series\n");
    printf("  [-precision <fp64|mixed>]\tgrid solves in FP64 (default) or FP32 inside FP64 refinement\n");
}


// command-line-only parameters (hotspot.c:94-147)
struct Global {
    std::string flp, ptrace, ttrace, config, dump, detailed_3d, materials;
    int microchannels = 0;
};

std::string value_of(str_pair *t, int n, const char *name, const char *dflt, const char *what) {
    const int i = get_str_index(t, n, const_cast<char *>(name));
    if (i < 0) return dflt;
    char tok[STR_SIZE];
    if (sscanf(t[i].value, "%511s", tok) != 1)
        die(std::string("invalid format for configuration  parameter ") + what + "\n");
    return tok;
}

Global global_from(str_pair *t, int n) {
    Global g;
    g.flp = value_of(t, n, "f", kNull, "flp_file");
    if (get_str_index(t, n, const_cast<char *>("p")) < 0)
        die("required parameter p_infile missing. check usage\n");
    g.ptrace = value_of(t, n, "p", kNull, "p_infile");
    g.ttrace = value_of(t, n, "o", kNull, "t_outfile");
    g.config = value_of(t, n, "c", kNull, "config");
    g.dump = value_of(t, n, "d", kNull, "dump_config");
    g.detailed_3d = value_of(t, n, "detailed_3D", "off", "lc");
    g.materials = value_of(t, n, "materials_file", kNull, "materials_file");
    const int i = get_str_index(t, n, const_cast<char *>("use_microchannels"));
    if (i >= 0 && sscanf(t[i].value, "%d", &g.microchannels) != 1)
        die("invalid format for configuration  parameter use_microchannels\n");
    return g;
}

int global_to(const Global &g, str_pair *t) {
    const std::pair<const char *, std::string> kv[] = {
        {"f", g.flp}, {"p", g.ptrace}, {"o", g.ttrace}, {"c", g.config}, {"d", g.dump},
        {"detailed_3D", g.detailed_3d}, {"use_microchannels", std::to_string(g.microchannels)},
        {"materials_file", g.materials}};
    int n = 0;
    for (auto &p : kv) {
        snprintf(t[n].name, STR_SIZE, "%s", p.first);
        snprintf(t[n].value, STR_SIZE, "%s", p.second.c_str());
        n++;
    }
    return n;
}

// power / temperature traces: a line of unit names, then one line of values
// per interval, whitespace separated; blank lines are skipped (hotspot.c:182-248)
bool next_line(FILE *fp, std::string &line) {
    line.clear();
    int ch;
    bool any = false;
    while ((ch = fgetc(fp)) != EOF) {
        any = true;
        if (ch == '\n') return true;
        line.push_back((char)ch);
    }
    return any;   // a last line without a newline
}

std::vector<std::string> split(const std::string &line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i])) i++;
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j])) j++;
        if (j > i) out.push_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

std::vector<std::string> read_names(FILE *fp) {
    std::string line;
    while (next_line(fp, line)) {
        auto names = split(line);
        if (names.empty()) continue;
        if ((int)names.size() > MAX_UNITS) die("no. of units exceeded limit\n");
        return names;
    }
    die("not enough names in trace file\n");
}

// values of the next non-blank line; empty at the end of the file
std::vector<double> read_vals(FILE *fp) {
    std::string line;
    while (next_line(fp, line)) {
        auto toks = split(line);
        if (toks.empty()) continue;
        std::vector<double> v(toks.size());
        for (size_t i = 0; i < toks.size(); i++) {
            char *end = nullptr;
            v[i] = strtod(toks[i].c_str(), &end);
            if (end == toks[i].c_str()) die("invalid format of values\n");
        }
        return v;
    }
    return {};
}

void write_row(FILE *fp, const std::vector<std::string> &s) {
    for (size_t i = 0; i < s.size(); i++) fprintf(fp, "%s%c", s[i].c_str(), i + 1 < s.size() ? '\t' : '\n');
}
void write_row(FILE *fp, const std::vector<double> &v) {
    for (size_t i = 0; i < v.size(); i++) fprintf(fp, "%.2f%c", v[i], i + 1 < v.size() ? '\t' : '\n');
}

// grid solver options of this library (hs_options_t); the reference's
// parsers ignore these names
void apply_solver_options(RC_model_t *model, str_pair *t, int n) {
    hs_options_t o;
    hs_get_options(model, &o);
    const std::string ts = value_of(t, n, "transient_solver", "rk4", "transient_solver");
    if (ts == "rk4") o.transient = HS_TRANS_RK4;
    else if (ts == "be") o.transient = HS_TRANS_BE;
    else if (ts == "taylor") o.transient = HS_TRANS_TAYLOR;
    else die("invalid value for -transient_solver (rk4, be or taylor)\n");
    const std::string pr = value_of(t, n, "precision", "fp64", "precision");
    if (pr == "fp64") o.precision = HS_PREC_FP64;
    else if (pr == "mixed") o.precision = HS_PREC_MIXED;
    else die("invalid value for -precision (fp64 or mixed)\n");
    hs_set_options(model, &o);
}

}  // namespace

int main(int argc, char **argv) {
    if (!(argc >= 5 && argc % 2)) {
        usage(argv[0]);
        return 1;
    }
    printf("Parsing input files...\n");
    static str_pair table[kMaxEntries];
    int size = parse_cmdline(table, kMaxEntries, argc, argv);
    Global g = global_from(table, size);
    const bool transient = g.ttrace != kNull;
    // command-line entries come first, so they win over the config file's
    if (g.config != kNull) size += read_str_pairs(&table[size], kMaxEntries - size, &g.config[0]);
    size = str_pairs_remove_duplicates(table, size);
    const bool detailed = g.detailed_3d == "on";

    materials_list_t materials;
    default_materials(&materials);
    if (g.materials != kNull) materials_add_from_file(&materials, &g.materials[0]);
    thermal_config_t config = default_thermal_config();
    thermal_config_add_from_strs(&config, &materials, table, size);
    microchannel_config_t channel_cfg = default_microchannel_config();
    microchannel_config_t *channel = nullptr;
    if (g.microchannels) {
        microchannel_config_add_from_strs(&channel_cfg, &materials, table, size);
        channel = &channel_cfg;
    }
    // package model: r_convec from the sink and fan (hotspot.c:447-455)
    int natural = 0;
    if (get_str_index(table, size, const_cast<char *>("package_model_used")) >= 0 && config.package_model_used) {
        natural = package_model(&config, table, size, config.ambient + kSmallForConvec);
        if (config.r_convec < 0.01 || config.r_convec > 50.0)
            printf("Warning: Heatsink convection resistance is not realistic, double-check your package settings...\n");
    }
    const bool block = !strcasecmp(config.model_type, "block");

    if (g.dump != kNull) {
        static str_pair out[kMaxEntries];
        int n = global_to(g, out);
        n += thermal_config_to_strs(&config, &out[n], kMaxEntries - n);
        if (channel) n += microchannel_config_to_strs(channel, &out[n], kMaxEntries - n);
        dump_str_pairs(out, n, &g.dump[0], const_cast<char *>("-"));
    }

    flp_t *flp = nullptr;
    if (block) {
        if (g.flp == kNull) die("the block model needs a floorplan file (-f)\n");
        flp = read_flp(&g.flp[0], 0, 0);
    } else if (strcmp(config.grid_layer_file, kNull)) {
        if (g.flp != kNull)
            fprintf(stderr, "Warning: Layer Configuration File %s specified. Overriding floorplan file %s\n",
                    config.grid_layer_file, g.flp.c_str());
    } else if (g.flp != kNull) {
        flp = read_flp(&g.flp[0], 0, 0);
    } else {
        die("Either LCF or FLP file must be specified\n");
    }
    RC_model_t *model = alloc_RC_model(&config, flp, channel, &materials, detailed, g.microchannels);
    if (!model) die("could not create the thermal model\n");
    grid_model_t *grid = block ? nullptr : model->grid;
    if (grid && !grid->has_lcf && detailed)
        die("-do_detailed_3D can only be used in 3D mode (if a grid_layer_file is specified)\n");
    if (g.microchannels && !detailed)
        die("-use_microchannels requires -model_type grid and do_detailed_3D on options\n");

    if (grid) apply_solver_options(model, table, size);

    printf("Creating thermal circuit...\n");
    populate_R_model(model, flp);
    if (transient) populate_C_model(model, flp);

    double *temp = transient ? hotspot_vector(model) : nullptr;
    double *power = hotspot_vector(model);
    double *steady = hotspot_vector(model);
    double *overall = hotspot_vector(model);
    if (transient) {
        if (strcmp(model->config->init_file, kNull))
            read_temp(model, temp, model->config->init_file, model->config->dtm_used ? 1 : 0);
        else
            set_temp(model, temp, model->config->init_temp);
    }

    // trace column -> block-vector entry, over the power-dissipating layers
    // (the block model: the floorplan's units)
    std::vector<int> slot;
    int n_power = block ? flp->n_units : 0;
    for (int l = 0; grid && l < grid->n_layers; l++)
        if (grid->layers[l].has_power) n_power += grid->layers[l].flp->n_units;
    FILE *pin = fopen(g.ptrace.c_str(), "r");
    if (!pin) die("unable to open power trace input file\n");
    FILE *tout = nullptr;
    if (transient && !(tout = fopen(g.ttrace.c_str(), "w"))) die("unable to open temperature trace file for output\n");
    std::vector<std::string> names = read_names(pin);
    if ((int)names.size() != n_power) die("no. of units in floorplan and trace file differ\n");
    if (block)
        for (int k = 0; k < n_power; k++) slot.push_back(get_blk_index(flp, &names[k][0]));
    for (int l = 0, base = 0, col = 0; grid && l < grid->n_layers; l++) {
        flp_t *f = grid->layers[l].flp;
        if (grid->layers[l].has_power) {
            for (int j = 0; j < f->n_units; j++)
                slot.push_back(base + get_blk_index(f, &names[col + j][0]));
            col += f->n_units;
        }
        base += f->n_units;
    }
    if (transient) write_row(tout, names);

    const int nvec = block ? model->block->n_nodes
                           : grid->total_n_blocks + (config.model_secondary ? EXTRA + EXTRA_SEC : EXTRA);
    const double dt = model->config->sampling_intvl;
    // row by row for the block model, when the grid state is written after
    // each interval, or when the convection resistance follows the sink
    // temperature
    const bool grid_trace = !block && strcmp(model->config->grid_transient_file, kNull) != 0;
    const bool per_row = block || grid_trace || natural;
    std::vector<double> seq;   // power rows, block-vector order
    int rows = 0;
    for (std::vector<double> v; !(v = read_vals(pin)).empty(); rows++) {
        if ((int)v.size() != n_power) die("invalid trace file format\n");
        for (int k = 0; k < n_power; k++) power[slot[k]] = v[k];
        for (int k = 0; k < n_power; k++) overall[slot[k]] += power[slot[k]];
        if (!transient) continue;
        if (per_row) {
            if (natural) {   // hotspot.c:590-595
                natural = package_model(model->config, table, size, calc_sink_temp(model, temp));
                populate_R_model(model, flp);
            }
            printf("Computing temperatures for t = %e...\n", rows * dt);
            compute_temp(model, power, block || rows == 0 ? temp : nullptr, dt);
            if (grid_trace) dump_transient_temp_grid(grid, rows, dt, model->config->grid_transient_file);
            std::vector<double> out(n_power);
            for (int k = 0; k < n_power; k++) out[k] = temp[slot[k]];
            write_row(tout, out);
        } else {
            seq.insert(seq.end(), power, power + nvec);
        }
    }
    if (!rows) die("no power numbers in trace file\n");
    if (transient && !per_row) {
        printf("Computing temperatures for %d intervals of %e s...\n", rows, dt);
        std::vector<double> out_seq((size_t)rows * nvec);
        hs_transient_run(model, seq.data(), rows, dt, temp, out_seq.data());
        std::vector<double> out(n_power);
        for (int r = 0; r < rows; r++) {
            for (int k = 0; k < n_power; k++) out[k] = out_seq[(size_t)r * nvec + slot[k]];
            write_row(tout, out);
        }
    }
    for (int k = 0; k < n_power; k++) overall[slot[k]] /= rows;

    if (natural) {   // r_convec iterated with the steady state (hotspot.c:624-638)
        for (bool converged = false; !converged;) {
            const double r_old = model->config->r_convec;
            steady_state_temp(model, overall, steady);
            const double sink = calc_sink_temp(model, steady) + kSmallForConvec;
            natural = package_model(model->config, table, size, sink);
            populate_R_model(model, flp);
            if (sink > kMaxSinkTemp)
                die("too high power for a natural convection package -- possible thermal runaway\n");
            converged = std::fabs(model->config->r_convec - r_old) < kNaturalConvecTol;
        }
    } else {
        fprintf(stderr, "Computing steady-state temperatures...\n");
        steady_state_temp(model, overall, steady);
    }
    if (strcmp(model->config->steady_file, kNull)) dump_temp(model, steady, model->config->steady_file);
    if (grid && strcmp(model->config->grid_steady_file, kNull))
        dump_steady_temp_grid(grid, model->config->grid_steady_file);

    fclose(pin);
    if (tout) fclose(tout);
    delete_RC_model(model);
    if (flp) free_flp(flp, 0, 0);
    if (temp) free_dvector(temp);
    free_materials(&materials);
    free_dvector(power);
    free_dvector(steady);
    free_dvector(overall);
    printf("Simulation complete.\n");
    return 0;
}
