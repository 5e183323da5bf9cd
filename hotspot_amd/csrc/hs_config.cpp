// Configuration tables of the reference's command line and config files:
// name/value pairs (util.c:322-439), the thermal configuration
// (temperature.c:115-402) and the microchannel configuration
// (microchannel.c:49-156).  Host code only.
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <strings.h>

#include "hs_internal.h"

namespace hs {
namespace {

constexpr int kLineSize = 65536;   // util.h:29

// one field of a config struct that a name/value pair can set
enum Kind { K_DBL, K_INT, K_STR };
struct Field {
    const char *name;
    Kind kind;
    size_t off;
    bool parsed;   // read by *_add_from_strs (detailed_3D_used is only written)
};

#define TF(f, k) Field{#f, k, offsetof(thermal_config_t, f), true}
// the order of thermal_config_to_strs (temperature.c:350-400)
const Field kThermal[] = {
    TF(t_chip, K_DBL), TF(k_chip, K_DBL), TF(p_chip, K_DBL), TF(thermal_threshold, K_DBL),
    TF(c_convec, K_DBL), TF(r_convec, K_DBL), TF(s_sink, K_DBL), TF(t_sink, K_DBL),
    TF(k_sink, K_DBL), TF(p_sink, K_DBL), TF(s_spreader, K_DBL), TF(t_spreader, K_DBL),
    TF(k_spreader, K_DBL), TF(p_spreader, K_DBL), TF(t_interface, K_DBL), TF(k_interface, K_DBL),
    TF(p_interface, K_DBL), TF(model_secondary, K_INT), TF(r_convec_sec, K_DBL), TF(c_convec_sec, K_DBL),
    TF(n_metal, K_INT), TF(t_metal, K_DBL), TF(t_c4, K_DBL), TF(s_c4, K_DBL),
    TF(n_c4, K_INT), TF(s_sub, K_DBL), TF(t_sub, K_DBL), TF(s_solder, K_DBL),
    TF(t_solder, K_DBL), TF(s_pcb, K_DBL), TF(t_pcb, K_DBL), TF(ambient, K_DBL),
    TF(init_file, K_STR), TF(init_temp, K_DBL), TF(steady_file, K_STR), TF(sampling_intvl, K_DBL),
    TF(base_proc_freq, K_DBL), TF(dtm_used, K_INT), TF(model_type, K_STR), TF(leakage_used, K_INT),
    TF(leakage_mode, K_INT), TF(package_model_used, K_INT), TF(package_config_file, K_STR),
    TF(block_omit_lateral, K_INT), TF(grid_rows, K_INT), TF(grid_cols, K_INT),
    TF(grid_layer_file, K_STR), TF(grid_steady_file, K_STR), TF(grid_map_mode, K_STR),
    TF(grid_transient_file, K_STR),
    Field{"detailed_3D_used", K_INT, offsetof(thermal_config_t, detailed_3D_used), false},
};
#undef TF

#define MF(f, k, p) Field{#f, k, offsetof(microchannel_config_t, f), p}
// the order of microchannel_config_to_strs (microchannel.c:119-152)
const Field kChannel[] = {
    MF(cell_width, K_DBL, false), MF(cell_height, K_DBL, false), MF(cell_thickness, K_DBL, false),
    MF(pumping_pressure, K_DBL, true), MF(pump_internal_res, K_DBL, true),
    MF(inlet_temperature, K_DBL, true), MF(coolant_capac, K_DBL, true), MF(coolant_res, K_DBL, true),
    MF(coolant_visc, K_DBL, true), MF(wall_capac, K_DBL, true), MF(wall_res, K_DBL, true),
    MF(htc, K_DBL, true), MF(network_file, K_STR, true), MF(floorplan_file, K_STR, false),
    MF(num_rows, K_INT, false), MF(num_columns, K_INT, false), MF(n_fluid_cells, K_INT, false),
};
#undef MF

// sscanf-style read of one value (leading token for strings) into a field
void parse_field(void *base, const Field &f, const char *value) {
    char *p = static_cast<char *>(base) + f.off;
    bool ok = false;
    switch (f.kind) {
        case K_DBL: ok = sscanf(value, "%lf", reinterpret_cast<double *>(p)) == 1; break;
        case K_INT: ok = sscanf(value, "%d", reinterpret_cast<int *>(p)) == 1; break;
        case K_STR: {
            char tok[STR_SIZE];
            ok = sscanf(value, "%511s", tok) == 1;
            if (ok) snprintf(p, STR_SIZE, "%s", tok);
            break;
        }
    }
    if (!ok) fatal(std::string("invalid format for configuration  parameter ") + f.name + "\n");
}

void format_field(const void *base, const Field &f, const char *dfmt, str_pair &out) {
    const char *p = static_cast<const char *>(base) + f.off;
    snprintf(out.name, STR_SIZE, "%s", f.name);
    switch (f.kind) {
        case K_DBL: snprintf(out.value, STR_SIZE, dfmt, *reinterpret_cast<const double *>(p)); break;
        case K_INT: snprintf(out.value, STR_SIZE, "%d", *reinterpret_cast<const int *>(p)); break;
        case K_STR: snprintf(out.value, STR_SIZE, "%s", p); break;
    }
}

int lookup(const str_pair *table, int size, const char *name) {
    if (!table) fatal("null pointer in get_str_index\n");
    for (int i = 0; i < size; i++)
        if (!strcasecmp(name, table[i].name)) return i;
    return -1;
}

// material property lookups (materials.c:177-205): exact name match, -1 if absent
const material_t *material(const materials_list_t *ml, const char *name) {
    if (!ml) return nullptr;
    for (int i = 0; i < ml->size; i++)
        if (!strncmp(ml->names[i], name, STR_SIZE)) return &ml->material_properties[i];
    return nullptr;
}

// a material_<x> entry sets the conductivity and heat capacity of <x>
void apply_material(const materials_list_t *ml, const str_pair *table, int size, const char *entry,
                    double &k, double &p) {
    const int idx = lookup(table, size, entry);
    if (idx < 0) return;
    char name[STR_SIZE];
    if (sscanf(table[idx].value, "%511s", name) != 1)
        fatal(std::string("invalid format for configuration parameter ") + entry + "\n");
    const material_t *m = material(ml, name);
    k = m ? m->thermal_conductivity : -1.0;
    p = m ? m->volumetric_heat_capacity : -1.0;
    if (k < 0 || p < 0)
        fatal(std::string("material name specified in configuration parameter ") + entry + " not found\n");
}

}  // namespace
}  // namespace hs

using namespace hs;

extern "C" {

int read_str_pairs(str_pair *table, int max_entries, char *file) {
    HS_BEGIN
    FILE *fp = fopen(file, "r");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for reading\n");
    std::string line(kLineSize, '\0');
    int n = 0;
    // a last line without its newline is dropped, as fgets + feof do in the reference
    while (n < max_entries && fgets(&line[0], kLineSize, fp) && !feof(fp)) {
        char name[STR_SIZE], value[STR_SIZE];
        char first[2] = {0, 0};
        if (sscanf(line.c_str(), " %1s", first) != 1 || first[0] == '#') continue;   // blank or comment
        if (sscanf(line.c_str(), "%511s%511s", name, value) != 2 || name[0] != '-') {
            fclose(fp);
            fatal("invalid file format\n");
        }
        snprintf(table[n].name, STR_SIZE, "%s", name + 1);
        snprintf(table[n].value, STR_SIZE, "%s", value);
        n++;
    }
    fclose(fp);
    return n;
    HS_END(0)
}

int parse_cmdline(str_pair *table, int max_entries, int argc, char **argv) {
    HS_BEGIN
    int n = 0;
    for (int i = 1; i < argc && n < max_entries; i += 2) {
        if (argv[i][0] != '-') fatal("invalid command line. check usage\n");
        snprintf(table[n].name, STR_SIZE, "%s", argv[i] + 1);
        if (i + 1 >= argc) break;   // a trailing name without a value is dropped
        snprintf(table[n].value, STR_SIZE, "%s", argv[i + 1]);
        n++;
    }
    return n;
    HS_END(0)
}

void dump_str_pairs(str_pair *table, int size, char *file, char *prefix) {
    HS_BEGIN
    FILE *fp = fopen(file, "w");
    if (!fp) fatal(std::string("error: ") + file + " could not be opened for writing\n");
    for (int i = 0; i < size; i++) fprintf(fp, "%s%s\t%s\n", prefix, table[i].name, table[i].value);
    fclose(fp);
    HS_END()
}

int get_str_index(str_pair *table, int size, char *str) {
    HS_BEGIN
    return lookup(table, size, str);
    HS_END(-1)
}

int str_pairs_remove_duplicates(str_pair *table, int size) {
    int kept = 0;
    for (int i = 0; i < size; i++) {
        bool seen = false;
        for (int k = 0; k < kept && !seen; k++) seen = !strcasecmp(table[k].name, table[i].name);
        if (seen) continue;
        if (kept != i) table[kept] = table[i];
        kept++;
    }
    return kept;
}

void thermal_config_add_from_strs(thermal_config_t *config, materials_list_t *materials_list, str_pair *table,
                                  int size) {
    HS_BEGIN
    thermal_config_t &c = *config;
    for (const Field &f : kThermal) {
        if (!f.parsed) continue;
        const int idx = lookup(table, size, f.name);
        if (idx >= 0) parse_field(config, f, table[idx].value);
    }
    // validation (temperature.c:269-297)
    if (c.t_chip <= 0 || c.s_sink <= 0 || c.t_sink <= 0 || c.s_spreader <= 0 || c.t_spreader <= 0 ||
        c.t_interface <= 0)
        fatal("chip and package dimensions should be greater than zero\n");
    if (c.t_metal <= 0 || c.n_metal <= 0 || c.t_c4 <= 0 || c.s_c4 <= 0 || c.n_c4 <= 0 || c.s_sub <= 0 ||
        c.t_sub <= 0 || c.s_solder <= 0 || c.t_solder <= 0 || c.s_pcb <= 0 || c.r_convec_sec <= 0 ||
        c.c_convec_sec <= 0)
        fatal("secondary heat tranfer layer dimensions should be greater than zero\n");
    if (c.leakage_used == 1) {
        printf("Warning: transient leakage iteration is not supported in this release...\n");
        printf(" ...all transient results are without thermal-leakage loop.\n");
    }
    if (c.model_secondary == 1 && !strcasecmp(c.model_type, "block"))
        fatal("secondary heat tranfer path is supported only in the grid mode\n");
    if (c.thermal_threshold < 0 || c.c_convec < 0 || c.r_convec < 0 || c.ambient < 0 || c.base_proc_freq <= 0 ||
        c.sampling_intvl <= 0)
        fatal("invalid thermal simulation parameters\n");
    if (strcasecmp(c.model_type, "block") && strcasecmp(c.model_type, "grid"))
        fatal("invalid model type. use 'block' or 'grid'\n");
    if (c.grid_rows <= 0 || c.grid_cols <= 0) fatal("grid rows and columns should both be greater than zero\n");
    if (strcasecmp(c.grid_map_mode, "avg") && strcasecmp(c.grid_map_mode, "min") &&
        strcasecmp(c.grid_map_mode, "max") && strcasecmp(c.grid_map_mode, "center"))
        fatal("invalid mapping mode. use 'avg', 'min', 'max' or 'center'\n");
    // named materials override the numeric values (temperature.c:299-338)
    apply_material(materials_list, table, size, "material_chip", c.k_chip, c.p_chip);
    apply_material(materials_list, table, size, "material_sink", c.k_sink, c.p_sink);
    apply_material(materials_list, table, size, "material_spreader", c.k_spreader, c.p_spreader);
    apply_material(materials_list, table, size, "material_interface", c.k_interface, c.p_interface);
    HS_END()
}

int thermal_config_to_strs(thermal_config_t *config, str_pair *table, int max_entries) {
    HS_BEGIN
    const int n = (int)(sizeof kThermal / sizeof kThermal[0]);
    if (max_entries < n) fatal("not enough entries in table\n");
    for (int i = 0; i < n; i++) format_field(config, kThermal[i], "%lg", table[i]);
    return n;
    HS_END(0)
}

void microchannel_config_add_from_strs(microchannel_config_t *config, materials_list_t *materials_list,
                                       str_pair *table, int size) {
    HS_BEGIN
    for (const Field &f : kChannel) {
        if (!f.parsed) continue;
        const int idx = lookup(table, size, f.name);
        if (idx >= 0) parse_field(config, f, table[idx].value);
    }
    // wall_material / coolant_material (microchannel.c:84-107): resistivities
    // are the inverse conductivities, the coolant also takes its viscosity
    for (const char *entry : {"wall_material", "coolant_material"}) {
        const int idx = lookup(table, size, entry);
        if (idx < 0) continue;
        char name[STR_SIZE];
        if (sscanf(table[idx].value, "%511s", name) != 1)
            fatal(std::string("invalid format for configuration parameter ") + entry + "\n");
        const material_t *m = material(materials_list, name);
        const double k = m ? m->thermal_conductivity : -1.0;
        const double cap = m ? m->volumetric_heat_capacity : -1.0;
        const double res = 1.0 / k;
        if (entry[0] == 'w') {
            config->wall_res = res;
            config->wall_capac = cap;
            if (res < 0 || cap < 0)
                fatal("material name specified in configuration parameter wall_material not found\n");
        } else {
            config->coolant_res = res;
            config->coolant_capac = cap;
            config->coolant_visc = m ? m->dynamic_viscosity : -1.0;
            if (res < 0 || cap < 0 || config->coolant_visc < 0)
                fatal("material name specified in configuration parameter coolant_material not found\n");
        }
    }
    HS_END()
}

int microchannel_config_to_strs(microchannel_config_t *config, str_pair *table, int max_entries) {
    HS_BEGIN
    const int n = (int)(sizeof kChannel / sizeof kChannel[0]);
    if (max_entries < n) fatal("not enough entries in table\n");
    for (int i = 0; i < n; i++) format_field(config, kChannel[i], "%e", table[i]);
    return n - 1;   // the reference counts 16 of its 17 entries
    HS_END(0)
}

}  // extern "C"
