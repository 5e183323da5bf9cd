// Package model: the heat sink's convection resistance r_convec from the
// sink and fan geometry, for forced convection (lateral or impinging air
// flow through a fin-channel or pin-fin sink) or natural convection at a
// given sink temperature (package.c:18-397).  The correlations, their
// constants and the order of their terms follow the reference, since
// r_convec must come out the same.  Host code only.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <strings.h>

#include "hs_internal.h"

namespace hs {
namespace {

// physical constants and correlation limits (package.h:12-45)
constexpr double kAirDensity = 1.059, kAirCond = 0.028, kAirDynVisc = 2.0e-5, kAirKinVisc = 1.9e-5;
constexpr double kPrandtl = 0.73, kReynoldsLaminar = 3500;
constexpr double kVolExp = 3.0e-3, kGravity = 9.8, kStefan = 5.67e-8, kEmissivity = 0.95;
constexpr double kPi = 3.1416;   // the reference's value of pi
constexpr double kFanAlpha = 1, kFanBeta = 4, kRpmToRad = 0.105;

// fins or pins that fit on the sink base and their wetted area
struct SinkShape {
    double n_fin, area_fin, n_pin, area_pin;
};
SinkShape sink_shape(const package_config_t &c, double s_sink) {
    SinkShape s;
    const double per_side = (s_sink - c.pin_diam) / (c.pin_diam + c.pin_dist);
    s.n_fin = std::ceil((s_sink - c.fin_width) / (c.fin_width + c.channel_width) - 0.5);
    s.area_fin = s_sink * (s_sink + 2.0 * s.n_fin * c.fin_height);
    s.n_pin = std::ceil(per_side * per_side - 0.5);
    s.area_pin = s_sink * s_sink + kPi * c.pin_diam * c.pin_height * s.n_pin;
    return s;
}

// resistance of a fin-channel sink with heat transfer coefficient h
double fin_channel_r(const package_config_t &c, double h, double k_sink, double s_sink, double n_fin) {
    const double m = std::sqrt(2 * h / (k_sink * c.fin_width));   // fin parameter
    const double eta = std::tanh(m * s_sink) / (m * s_sink);        // fin efficiency
    return 1.0 / (h * (c.channel_width + 2 * eta * s_sink) * n_fin * s_sink);
}

#define PF(f, k) Field{#f, k, offsetof(package_config_t, f)}
enum Kind { K_DBL, K_INT };
struct Field {
    const char *name;
    Kind kind;
    size_t off;
};
// the order of package_config_to_strs (package.c:115-140)
const Field kPackage[] = {PF(natural_convec, K_INT), PF(flow_type, K_INT), PF(sink_type, K_INT),
                          PF(fin_height, K_DBL),     PF(fin_width, K_DBL), PF(channel_width, K_DBL),
                          PF(pin_height, K_DBL),     PF(pin_diam, K_DBL),  PF(pin_dist, K_DBL),
                          PF(fan_radius, K_DBL),     PF(motor_radius, K_DBL), PF(rpm, K_INT)};
#undef PF

}  // namespace
}  // namespace hs

using namespace hs;

extern "C" {

package_config_t default_package_config(void) {
    package_config_t c;
    memset(&c, 0, sizeof c);
    c.natural_convec = 0;
    c.flow_type = 0;
    c.sink_type = 0;
    c.fin_height = 0.03;
    c.fin_width = 0.001;
    c.channel_width = 0.002;
    c.pin_height = 0.02;
    c.pin_diam = 0.002;
    c.pin_dist = 0.005;
    c.fan_radius = 0.03;
    c.motor_radius = 0.01;
    c.rpm = 1000;
    return c;
}

void package_config_add_from_strs(package_config_t *config, str_pair *table, int size) {
    HS_BEGIN
    for (const Field &f : kPackage) {
        const int idx = get_str_index(table, size, const_cast<char *>(f.name));
        if (idx < 0) continue;
        char *p = reinterpret_cast<char *>(config) + f.off;
        const bool ok = f.kind == K_DBL ? sscanf(table[idx].value, "%lf", reinterpret_cast<double *>(p)) == 1
                                        : sscanf(table[idx].value, "%d", reinterpret_cast<int *>(p)) == 1;
        if (!ok) fatal(std::string("invalid format for heatsink configuration parameter ") + f.name + "\n");
    }
    const package_config_t &c = *config;
    if (c.fin_height <= 0 || c.fin_width <= 0 || c.channel_width <= 0 || c.pin_height <= 0 || c.pin_diam <= 0 ||
        c.pin_dist <= 0 || c.fan_radius <= 0 || c.motor_radius <= 0 || c.rpm <= 0)
        fatal("heatsink/fan dimensions and fan speed should be greater than zero\n");
    if (c.natural_convec != 0 && c.natural_convec != 1) fatal("invalid convection mode\n");
    if (c.flow_type != 0 && c.flow_type != 1) fatal("invalid air flow type\n");
    if (c.sink_type != 0 && c.sink_type != 1) fatal("invalid heatsink type\n");
    HS_END()
}

int package_config_to_strs(package_config_t *config, str_pair *table, int max_entries) {
    HS_BEGIN
    const int n = (int)(sizeof kPackage / sizeof kPackage[0]);
    if (max_entries < n) fatal("not enough entries in table\n");
    for (int i = 0; i < n; i++) {
        const Field &f = kPackage[i];
        const char *p = reinterpret_cast<const char *>(config) + f.off;
        snprintf(table[i].name, STR_SIZE, "%s", f.name);
        if (f.kind == K_DBL) snprintf(table[i].value, STR_SIZE, "%lg", *reinterpret_cast<const double *>(p));
        else snprintf(table[i].value, STR_SIZE, "%d", *reinterpret_cast<const int *>(p));
    }
    return n;
    HS_END(0)
}

// forced convection (package.c:159-278): fan volume flow from a spindle
// motor model, air speed through the sink, Reynolds number, then the
// Nusselt correlation of the flow and sink type
void calculate_flow(convection_t *p, package_config_t *config, thermal_config_t *thermal_config) {
    const package_config_t &c = *config;
    const double s_sink = thermal_config->s_sink, k_sink = thermal_config->k_sink;
    const SinkShape sh = sink_shape(c, s_sink);
    const double rpm = c.rpm;
    const double dr = std::sqrt(c.fan_radius * c.fan_radius - c.motor_radius * c.motor_radius);
    const double r_mean = c.motor_radius + dr;
    const double a_fan = kPi * dr * dr;
    const double vol_v =
        a_fan * std::sqrt(0.25 * kPi * std::pow(r_mean, kFanBeta - 1) * std::pow(rpm * kRpmToRad, kFanAlpha) / a_fan);
    double dh, v;
    if (c.flow_type == 0) {   // lateral flow through the channels
        dh = 2.0 * c.channel_width * s_sink / (c.channel_width + s_sink);
        v = c.sink_type == 0 ? vol_v / ((sh.n_fin - 1) * c.channel_width * c.fin_height)
                             : vol_v / ((std::sqrt(sh.n_pin) - 1) * c.pin_dist * c.pin_height);
    } else {                  // impinging from the top (open area between the pins)
        dh = 2.0 * s_sink / std::sqrt(kPi);
        v = vol_v / (s_sink * s_sink - sh.n_pin * kPi * (c.pin_diam * 0.5) * (c.pin_diam * 0.5));
    }
    const double re = kAirDensity * v * dh / kAirDynVisc;
    double nu, h, r;
    if (c.flow_type == 0 && c.sink_type == 0) {
        if (re <= kReynoldsLaminar) {   // developing laminar flow between plates
            const double re_s = kAirDensity * v * c.channel_width * c.channel_width / (kAirDynVisc * s_sink);
            const double t1 = std::pow(re_s * kPrandtl * 0.5, -3);
            const double t2 = std::pow(0.664 * std::sqrt(re_s) * std::pow(kPrandtl, 0.33) * std::sqrt(1 + 3.65 / std::sqrt(re_s)), -3);
            nu = std::pow(t1 + t2, -0.33);
            h = nu * kAirCond / c.channel_width;
        } else {                        // turbulent: Petukhov friction factor
            const double f = 1.0 / std::pow((0.79 * std::log(re) - 1.64), 2.0);
            nu = 0.125 * f * re * std::pow(kPrandtl, 0.33);
            h = nu * kAirCond / dh;
        }
        r = fin_channel_r(c, h, k_sink, s_sink, sh.n_fin);
    } else if (c.flow_type == 0) {      // pin bank
        const double c1 = 0.6;
        const double rows = std::sqrt(sh.n_pin) - 1;
        const double c2 = 0.8 + 0.8333 * std::sqrt(rows) / std::pow((1 + 0.03 * std::pow(rows, 2.0)), 0.25);
        nu = 0.4 + 0.9 * (1.25 * c2 * c1 * std::sqrt(re) + 0.001 * re) * std::pow(kPrandtl, 0.33);
        h = nu * kAirCond / c.pin_diam;
        r = 1.0 / (h * sh.area_pin);
    } else {                            // impinging jet
        const double a_hs = c.sink_type == 0 ? sh.area_fin : sh.area_pin;
        const double d_jet = 2.0 * s_sink / std::sqrt(kPi);
        nu = 1.92 * std::pow(re, 0.716) * std::pow(kPrandtl, 0.4) * std::pow(a_hs / (s_sink * s_sink), -0.698);
        h = nu * kAirCond / d_jet;
        r = 1.0 / (h * a_hs);
    }
    p->n_fin = sh.n_fin;
    p->sur_area_fin = sh.area_fin;
    p->n_pin = sh.n_pin;
    p->sur_area_pin = sh.area_pin;
    p->reynolds = re;
    p->nusselt = nu;
    p->h_coeff = h;
    p->v = v;
    p->r_th = r;
}

// natural convection at a sink temperature, in parallel with radiation
// (package.c:281-335)
void calc_natural_convec(convection_t *p, package_config_t *config, thermal_config_t *thermal_config,
                         double sink_temp) {
    const package_config_t &c = *config;
    const double s_sink = thermal_config->s_sink, ambient = thermal_config->ambient;
    const SinkShape sh = sink_shape(c, s_sink);
    const double area = c.sink_type == 0 ? sh.area_fin : sh.area_pin;
    const double w = c.sink_type == 0 ? c.channel_width : c.pin_dist;
    const double ra = kGravity * kVolExp * (sink_temp - ambient) * std::pow(w, 3.0) * kPrandtl / (kAirKinVisc * kAirKinVisc);
    const double g = ra * w / s_sink;
    const double nu = std::pow((576 / (g * g) + 2.873 / std::sqrt(g)), -0.5);
    const double h = nu * kAirCond / w;
    const double r_conv = 1.0 / (h * area);
    const double r_rad =
        (sink_temp - ambient) / (kEmissivity * kStefan * (std::pow(sink_temp, 4.0) - std::pow(ambient, 4.0)) * area);
    p->n_fin = sh.n_fin;
    p->sur_area_fin = sh.area_fin;
    p->n_pin = sh.n_pin;
    p->sur_area_pin = sh.area_pin;
    p->nusselt = nu;
    p->h_coeff = h;
    p->r_th = r_conv * r_rad / (r_conv + r_rad);
}

// package_config_file from the table, the package file's parameters, and
// r_convec from the forced or natural model; returns natural_convec
// (package.c:353-397)
int package_model(thermal_config_t *thermal_config, str_pair *table, int size, double sink_temp) {
    HS_BEGIN
    static str_pair ptable[512];
    package_config_t pc = default_package_config();
    const int idx = get_str_index(table, size, const_cast<char *>("package_config_file"));
    if (idx >= 0 && sscanf(table[idx].value, "%511s", thermal_config->package_config_file) != 1)
        fatal("invalid format for configuration  parameter package_config_file\n");
    int n = 0;
    if (strcmp(thermal_config->package_config_file, "(null)"))
        n = read_str_pairs(ptable, 512, thermal_config->package_config_file);
    package_config_add_from_strs(&pc, ptable, n);
    convection_t p;
    memset(&p, 0, sizeof p);
    if (!pc.natural_convec) calculate_flow(&p, &pc, thermal_config);
    else calc_natural_convec(&p, &pc, thermal_config, sink_temp);
    thermal_config->r_convec = p.r_th;
    return pc.natural_convec;
    HS_END(0)
}

}  // extern "C"
