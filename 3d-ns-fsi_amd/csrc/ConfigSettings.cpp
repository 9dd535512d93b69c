/*
 * ConfigSettings.cpp — the reference's ConfigFile format (libconfig syntax) parsed without libconfig.
 *
 * Replaces ConfigSettings::loadSettings / readSettingValues / setDerivedParameters
 * (src/ConfigSettings.cpp:18-163 of the reference). Same keys, same strict typing as libconfig without
 * auto-conversion (an integer key written "48." or a float key written "1" is a type error), same
 * derived machinePrecisionBuffer. Extension keys (all optional) cover what the reference hard-codes:
 * boundary types per face, body type/centre/radius/axis, initial condition, perturbation seed and
 * amplitude, inlet ramp time, precision.
 *
 * Grammar accepted: settings "name = value" or "name : value" with optional ';' or ','; values are
 * integers (decimal or 0x hex, optional L/LL suffix), floats, booleans, "strings" (adjacent literals
 * concatenate), arrays [..], lists (..) and groups {..}; comments #, // and C-style block comments.
 */
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ns_fsi.h"
#include "ns_internal.h"

namespace {

struct Value {
    enum Kind { Int, Float, Bool, String, Array, Group } kind = Int;
    long long i = 0;
    double f = 0;
    bool b = false;
    std::string s;
    std::vector<Value> items;         // array elements, or group members (names below)
    std::vector<std::string> names;   // group member names, in file order
};

struct ParseError : std::runtime_error { using std::runtime_error::runtime_error; };

class Parser {
public:
    explicit Parser(const char* text) : p(text), line(1) {}

    std::map<std::string, Value> parseTop() {
        std::map<std::string, Value> out;
        skip();
        while (*p) {
            parseSetting(out);
            skip();
        }
        return out;
    }

private:
    const char* p;
    int line;

    [[noreturn]] void fail(const std::string& m) { throw ParseError("line " + std::to_string(line) + ": " + m); }

    void skip() {
        for (;;) {
            if (*p == '\n') { ++line; ++p; }
            else if (std::isspace((unsigned char)*p)) ++p;
            else if (*p == '#' || (p[0] == '/' && p[1] == '/')) { while (*p && *p != '\n') ++p; }
            else if (p[0] == '/' && p[1] == '*') {
                p += 2;
                while (*p && !(p[0] == '*' && p[1] == '/')) { if (*p == '\n') ++line; ++p; }
                if (!*p) fail("unterminated comment");
                p += 2;
            } else break;
        }
    }

    void parseSetting(std::map<std::string, Value>& out) {
        if (!(std::isalpha((unsigned char)*p) || *p == '*')) fail("expected setting name");
        const char* s = p;
        while (std::isalnum((unsigned char)*p) || *p == '_' || *p == '-' || *p == '*') ++p;
        std::string name(s, p);
        skip();
        if (*p != '=' && *p != ':') fail("expected '=' or ':' after " + name);
        ++p;
        skip();
        Value v = parseValue();
        skip();
        if (*p == ';' || *p == ',') ++p;
        if (out.count(name)) fail("duplicate setting " + name);
        out[name] = v;
    }

    Value parseValue() {
        Value v;
        if (*p == '"') {
            v.kind = Value::String;
            while (*p == '"') {
                ++p;
                while (*p && *p != '"') {
                    if (*p == '\\') {
                        ++p;
                        char c = *p;
                        v.s += c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c == 'f' ? '\f' : c;
                    } else {
                        if (*p == '\n') ++line;
                        v.s += *p;
                    }
                    if (*p) ++p;
                }
                if (*p != '"') fail("unterminated string");
                ++p;
                skip();
            }
            return v;
        }
        if (*p == '[' || *p == '(') {
            char close = *p == '[' ? ']' : ')';
            v.kind = Value::Array;
            ++p;
            skip();
            while (*p && *p != close) {
                v.items.push_back(parseValue());
                skip();
                if (*p == ',') { ++p; skip(); }
                else if (*p != close) fail("expected ',' or closing bracket");
            }
            if (*p != close) fail("unterminated array");
            ++p;
            return v;
        }
        if (*p == '{') {
            v.kind = Value::Group;
            ++p;
            skip();
            std::map<std::string, Value> inner;
            while (*p && *p != '}') { parseSetting(inner); skip(); }
            if (*p != '}') fail("unterminated group");
            ++p;
            for (auto& kv : inner) {  // the reference reads top-level keys only; groups serve the extension keys
                v.names.push_back(kv.first);
                v.items.push_back(kv.second);
            }
            return v;
        }
        // booleans
        auto word = [&](const char* w) {
            size_t n = std::strlen(w);
            for (size_t q = 0; q < n; ++q)
                if (std::tolower((unsigned char)p[q]) != w[q]) return false;
            return !std::isalnum((unsigned char)p[n]) && p[n] != '_';
        };
        if (word("true")) { v.kind = Value::Bool; v.b = true; p += 4; return v; }
        if (word("false")) { v.kind = Value::Bool; v.b = false; p += 5; return v; }
        // numbers
        const char* s = p;
        if (*p == '+' || *p == '-') ++p;
        if (p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) {
            p += 2;
            while (std::isxdigit((unsigned char)*p)) ++p;
            v.kind = Value::Int;
            v.i = (long long)std::strtoull(s, nullptr, 16);
            if (*s == '-') v.i = -(long long)std::strtoull(s + 1, nullptr, 16);
            while (*p == 'L') ++p;
            return v;
        }
        bool isFloat = false;
        while (std::isdigit((unsigned char)*p)) ++p;
        if (*p == '.') { isFloat = true; ++p; while (std::isdigit((unsigned char)*p)) ++p; }
        if (*p == 'e' || *p == 'E') {
            isFloat = true;
            ++p;
            if (*p == '+' || *p == '-') ++p;
            if (!std::isdigit((unsigned char)*p)) fail("bad exponent");
            while (std::isdigit((unsigned char)*p)) ++p;
        }
        if (p == s || (p == s + 1 && (*s == '+' || *s == '-'))) fail("expected a value");
        std::string num(s, p);
        if (isFloat) { v.kind = Value::Float; v.f = std::strtod(num.c_str(), nullptr); }
        else {
            v.kind = Value::Int;
            v.i = std::strtoll(num.c_str(), nullptr, 10);
            while (*p == 'L') ++p;
        }
        return v;
    }
};

class Reader {
public:
    Reader(const std::map<std::string, Value>& m) : m(m) {}
    bool has(const char* k) const { return m.count(k) != 0; }
    const Value& get(const char* k) const {
        auto it = m.find(k);
        if (it == m.end()) throw ParseError(std::string("ConfigFile setting not found: ") + k);
        return it->second;
    }
    long long integer(const char* k) const {
        const Value& v = get(k);
        if (v.kind != Value::Int) throw ParseError(std::string("ConfigFile setting type mismatch: ") + k + " must be an integer");
        return v.i;
    }
    double real(const char* k) const {
        const Value& v = get(k);
        if (v.kind != Value::Float) throw ParseError(std::string("ConfigFile setting type mismatch: ") + k + " must be a floating-point value");
        return v.f;
    }
    bool boolean(const char* k) const {
        const Value& v = get(k);
        if (v.kind != Value::Bool) throw ParseError(std::string("ConfigFile setting type mismatch: ") + k + " must be a boolean");
        return v.b;
    }
    std::string str(const char* k) const {
        const Value& v = get(k);
        if (v.kind != Value::String) throw ParseError(std::string("ConfigFile setting type mismatch: ") + k + " must be a string");
        return v.s;
    }
    const std::vector<Value>& array(const char* k) const {
        const Value& v = get(k);
        if (v.kind != Value::Array) throw ParseError(std::string("ConfigFile setting type mismatch: ") + k + " must be an array");
        return v.items;
    }

private:
    const std::map<std::string, Value>& m;
};

// One entry of the extraBodies list: { type = "sphere" | "cylinder"; axis = "x" | "y" | "z"; centre = [x, y, z];
// radius = r; } (axis only for a cylinder).
ns_body bodyFromGroup(const Value& g, int q) {
    const std::string where = "extraBodies entry " + std::to_string(q) + ": ";
    if (g.kind != Value::Group) throw ParseError(where + "must be a group { ... }");
    ns_body b{};
    b.type = -1;
    bool haveCentre = false, haveRadius = false;
    for (size_t m = 0; m < g.names.size(); ++m) {
        const std::string& n = g.names[m];
        const Value& v = g.items[m];
        if (n == "type") {
            if (v.kind != Value::String) throw ParseError(where + "type must be a string");
            if (v.s == "cylinder") b.type = NS_BODY_CYLINDER;
            else if (v.s == "sphere") b.type = NS_BODY_SPHERE;
            else throw ParseError(where + "type must be cylinder or sphere");
        } else if (n == "axis") {
            if (v.kind != Value::String || v.s.size() != 1 || v.s.find_first_of("xyz") != 0)
                throw ParseError(where + "axis must be x, y or z");
            b.axis = v.s[0] - 'x';
        } else if (n == "centre") {
            if (v.kind != Value::Array || v.items.size() != 3) throw ParseError(where + "centre needs 3 floats");
            for (int a = 0; a < 3; ++a) {
                if (v.items[a].kind != Value::Float) throw ParseError(where + "centre must hold floats");
                b.centre[a] = v.items[a].f;
            }
            haveCentre = true;
        } else if (n == "radius") {
            if (v.kind != Value::Float) throw ParseError(where + "radius must be a floating-point value");
            b.radius = v.f;
            haveRadius = true;
        } else {
            throw ParseError(where + "unknown member " + n);
        }
    }
    if (b.type < 0 || !haveCentre || !haveRadius) throw ParseError(where + "needs type, centre and radius");
    return b;
}

int bcFromString(const std::string& s) {
    if (s == "inlet") return NS_BC_INLET;
    if (s == "outlet") return NS_BC_OUTLET;
    if (s == "periodic") return NS_BC_PERIODIC;
    if (s == "symmetry") return NS_BC_SYMMETRY;
    throw ParseError("boundaryConditions entries must be inlet, outlet, periodic or symmetry");
}

}  // namespace

namespace nsfsi {

void setDerived(ns_config& c) {
    // ConfigSettings::setDerivedParameters, src/ConfigSettings.cpp:160-163
    c.machine_precision_buffer = std::numeric_limits<double>::epsilon() * (c.L_x + c.L_y + c.L_z);
}

std::string validateConfig(const ns_config& c) {
    if (c.NI < 3 || c.NJ < 3 || c.NK < 3) return "NI, NJ and NK must be at least 3";
    for (int a = 0; a < 3; ++a) {
        int tmin = c.bc[2 * a], tmax = c.bc[2 * a + 1];
        for (int t : {tmin, tmax})
            if (t < 0 || t > 3) return "unknown boundary type";
        if ((tmin == NS_BC_PERIODIC) != (tmax == NS_BC_PERIODIC)) return "periodic faces must come in pairs";
        int n = a == 0 ? c.NI : a == 1 ? c.NJ : c.NK;
        int gmin = (tmin == NS_BC_PERIODIC || tmin == NS_BC_SYMMETRY) ? 1 : 0;
        int gmax = tmax == NS_BC_SYMMETRY ? 1 : 0;
        if (n - 1 - gmin - gmax < 1) return "too few nodes along an axis for its boundary types";
    }
    if (c.body_type < 0 || c.body_type > 2) return "unknown body type";
    if (c.body_type == NS_BODY_CYLINDER && (c.body_axis < 0 || c.body_axis > 2)) return "bodyAxis must be x, y or z";
    if (c.body_type != NS_BODY_NONE && !(c.body_radius > 0)) return "bodyRadius must be positive";
    if (c.n_extra_bodies < 0 || c.n_extra_bodies > NS_MAX_BODIES - 1) return "at most 7 extra bodies";
    if (c.n_extra_bodies > 0 && c.body_type == NS_BODY_NONE) return "extra bodies need a first body (bodyType)";
    for (int q = 0; q < c.n_extra_bodies; ++q) {
        const ns_body& b = c.extra_bodies[q];
        if (b.type != NS_BODY_CYLINDER && b.type != NS_BODY_SPHERE) return "extra body type must be cylinder or sphere";
        if (b.type == NS_BODY_CYLINDER && (b.axis < 0 || b.axis > 2)) return "extra cylinder axis must be x, y or z";
        if (!(b.radius > 0)) return "extra body radius must be positive";
    }
    if (c.ic_type < 0 || c.ic_type > 3) return "unknown initial condition";
    if (c.precision != NS_FP64 && c.precision != NS_FP32) return "precision must be fp64 or fp32";
    if (c.stop_criterion != 0 && c.stop_criterion != 1) return "stopCriterion must be timesteps or end_time";
    if (c.conv_history < 0 || c.conv_history > 2) return "saveConvergenceHistory must be none, density or all";
    if (c.n_filter_intervals < 0 || c.n_filter_intervals > NS_MAX_FILTER || c.n_filter_change_times < 0 ||
        c.n_filter_change_times > NS_MAX_FILTER)
        return "too many filter intervals";
    // src/ConfigSettings.cpp:116-121
    if (c.n_filter_change_times > c.n_filter_intervals || c.n_filter_change_times < c.n_filter_intervals - 1)
        return "filterIntervalChangeTimes must be the same size or one element smaller than filterIntervals.";
    if (!(c.inlet_ramp_time > 0)) return "inletRampTime must be positive";
    return "";
}

}  // namespace nsfsi

extern "C" void ns_config_default(ns_config* c) {
    std::memset(c, 0, sizeof(*c));
    // The reference's shipped ConfigFile (ConfigFile:22-125)
    c->NI = 161; c->NJ = 83; c->NK = 3;
    c->stop_criterion = 1;
    c->stop_time_level = 20;
    c->t_end = 200.;
    c->conv_stability_limit = 0.9;
    c->visc_stability_limit = -0.68;
    c->n_filter_intervals = 3;
    c->filter_intervals[0] = 50; c->filter_intervals[1] = 200; c->filter_intervals[2] = 1000;
    c->n_filter_change_times = 3;
    c->filter_change_times[0] = 30.; c->filter_change_times[1] = 60.; c->filter_change_times[2] = 100.;
    c->status_report_interval = 2.;
    c->save_density = 1; c->save_momentum = 0; c->save_energy = 0; c->save_velocity = 1;
    c->save_pressure = 1; c->save_temperature = 1; c->save_viscosity = 0; c->save_thermal_cond = 0;
    c->save_ic = 0; c->save_final = 0; c->save_intervals = 1;
    c->save_period = 5.; c->save_intervals_start_time = 0.; c->save_intervals_end_time = 0.;
    c->conv_history = 2;
    c->gamma = 1.4; c->Pr = 0.72; c->R = 287.; c->Re = 400.; c->sutherlands_C2 = 110.4;
    c->M_0 = 0.25; c->T_0 = 300.; c->L_x = 2.; c->L_y = 1.; c->L_z = 0.05;
    // Hard-coded layout of src/Mesh.cpp:33-55: periodic z, symmetry y, inlet x-min, outlet x-max, and a
    // z-axis cylinder at (L_x/4, L_y/2) of radius L_y/10.
    c->bc[NS_XMIN] = NS_BC_INLET; c->bc[NS_XMAX] = NS_BC_OUTLET;
    c->bc[NS_YMIN] = NS_BC_SYMMETRY; c->bc[NS_YMAX] = NS_BC_SYMMETRY;
    c->bc[NS_ZMIN] = NS_BC_PERIODIC; c->bc[NS_ZMAX] = NS_BC_PERIODIC;
    c->body_type = NS_BODY_CYLINDER;
    c->body_axis = 2;
    c->body_centre[0] = c->L_x / 4; c->body_centre[1] = c->L_y / 2; c->body_centre[2] = 0;
    c->body_radius = c->L_y / 10;
    c->ic_type = NS_IC_STAGNATION;  // src/Solver.cpp:26
    c->ic_seed = 0;
    c->ic_amplitude = 0;
    c->inlet_ramp_time = 10.;  // src/Boundary.cpp:267
    c->precision = NS_FP64;
    c->kahan = 1;
    nsfsi::setDerived(*c);
}

extern "C" ns_status ns_config_parse(const char* text, ns_config* out, char* errbuf, size_t errbuf_len) {
    auto report = [&](const std::string& m) {
        if (errbuf && errbuf_len) std::snprintf(errbuf, errbuf_len, "%s", m.c_str());
        return NS_ERR_CONFIG;
    };
    if (!text || !out) return report("null argument");
    ns_config c;
    ns_config_default(&c);
    try {
        Parser parser(text);
        std::map<std::string, Value> m = parser.parseTop();
        Reader r(m);
        // readSettingValues, src/ConfigSettings.cpp:77-156
        c.NI = (int32_t)(unsigned)r.integer("NI");
        c.NJ = (int32_t)(unsigned)r.integer("NJ");
        c.NK = (int32_t)(unsigned)r.integer("NK");
        std::string sc = r.str("stopCriterion");
        if (sc == "timesteps") { c.stop_criterion = 0; c.stop_time_level = (unsigned)r.integer("stopTimeLevel"); }
        else if (sc == "end_time") { c.stop_criterion = 1; c.t_end = r.real("t_end"); }
        else throw ParseError("Error when reading ConfigFile. 'stopCriterion' needs to be 'timesteps' or 'end_time'.");
        c.conv_stability_limit = r.real("convStabilityLimit");
        c.visc_stability_limit = r.real("viscStabilityLimit");
        const auto& fi = r.array("filterIntervals");
        if (fi.size() > NS_MAX_FILTER) throw ParseError("too many filterIntervals");
        c.n_filter_intervals = (int32_t)fi.size();
        for (size_t q = 0; q < fi.size(); ++q) {
            if (fi[q].kind != Value::Int) throw ParseError("ConfigFile setting type mismatch: filterIntervals must hold integers");
            c.filter_intervals[q] = (uint32_t)fi[q].i;
        }
        const auto& ft = r.array("filterIntervalChangeTimes");
        if (ft.size() > NS_MAX_FILTER) throw ParseError("too many filterIntervalChangeTimes");
        c.n_filter_change_times = (int32_t)ft.size();
        for (size_t q = 0; q < ft.size(); ++q) {
            if (ft[q].kind != Value::Float) throw ParseError("ConfigFile setting type mismatch: filterIntervalChangeTimes must hold floats");
            c.filter_change_times[q] = ft[q].f;
        }
        c.status_report_interval = r.real("statusReportInterval");
        c.save_density = r.boolean("saveDensity");
        c.save_momentum = r.boolean("saveMomentum");
        c.save_energy = r.boolean("saveEnergy");
        c.save_velocity = r.boolean("saveVelocity");
        c.save_pressure = r.boolean("savePressure");
        c.save_temperature = r.boolean("saveTemperature");
        c.save_viscosity = r.boolean("saveViscosity");
        c.save_thermal_cond = r.boolean("saveThermalCond");
        c.save_ic = r.boolean("saveIC");
        c.save_final = r.boolean("saveFinal");
        c.save_intervals = r.boolean("saveIntervals");
        c.save_period = r.real("savePeriod");
        c.save_intervals_start_time = r.real("saveIntervalsStartTime");
        c.save_intervals_end_time = r.real("saveIntervalsEndTime");
        std::string ch = r.str("saveConvergenceHistory");
        if (ch == "none") c.conv_history = 0;
        else if (ch == "density") c.conv_history = 1;
        else if (ch == "all") c.conv_history = 2;
        else throw ParseError("saveConvergenceHistory must be 'none', 'density' or 'all'.");
        c.gamma = r.real("Gamma");
        c.Pr = r.real("Pr");
        c.R = r.real("R");
        c.Re = r.real("Re");
        c.sutherlands_C2 = r.real("sutherlands_C2");
        c.M_0 = r.real("M_0");
        c.T_0 = r.real("T_0");
        c.L_x = r.real("L_x");
        c.L_y = r.real("L_y");
        c.L_z = r.real("L_z");
        // ---- extension keys
        c.body_centre[0] = c.L_x / 4; c.body_centre[1] = c.L_y / 2; c.body_centre[2] = 0;
        c.body_radius = c.L_y / 10;
        if (r.has("boundaryConditions")) {
            const auto& b = r.array("boundaryConditions");
            if (b.size() != 6) throw ParseError("boundaryConditions needs 6 entries: x-min, x-max, y-min, y-max, z-min, z-max");
            for (int q = 0; q < 6; ++q) {
                if (b[q].kind != Value::String) throw ParseError("boundaryConditions entries must be strings");
                c.bc[q] = bcFromString(b[q].s);
            }
        }
        if (r.has("bodyType")) {
            std::string bt = r.str("bodyType");
            if (bt == "none") c.body_type = NS_BODY_NONE;
            else if (bt == "cylinder") c.body_type = NS_BODY_CYLINDER;
            else if (bt == "sphere") c.body_type = NS_BODY_SPHERE;
            else throw ParseError("bodyType must be none, cylinder or sphere");
        }
        if (r.has("bodyAxis")) {
            std::string ax = r.str("bodyAxis");
            if (ax == "x") c.body_axis = 0; else if (ax == "y") c.body_axis = 1; else if (ax == "z") c.body_axis = 2;
            else throw ParseError("bodyAxis must be x, y or z");
        }
        if (r.has("bodyCentre")) {
            const auto& bc = r.array("bodyCentre");
            if (bc.size() != 3) throw ParseError("bodyCentre needs 3 floats");
            for (int q = 0; q < 3; ++q) {
                if (bc[q].kind != Value::Float) throw ParseError("ConfigFile setting type mismatch: bodyCentre must hold floats");
                c.body_centre[q] = bc[q].f;
            }
        }
        if (r.has("bodyRadius")) c.body_radius = r.real("bodyRadius");
        if (r.has("extraBodies")) {
            const auto& eb = r.array("extraBodies");
            if (eb.size() > NS_MAX_BODIES - 1) throw ParseError("extraBodies holds at most 7 bodies");
            c.n_extra_bodies = (int32_t)eb.size();
            for (size_t q = 0; q < eb.size(); ++q) c.extra_bodies[q] = bodyFromGroup(eb[q], (int)q);
        }
        if (r.has("initialCondition")) {
            std::string ic = r.str("initialCondition");
            if (ic == "stagnation") c.ic_type = NS_IC_STAGNATION;
            else if (ic == "uniform_oblique") c.ic_type = NS_IC_UNIFORM_OBLIQUE;
            else if (ic == "freestream_gaussian") c.ic_type = NS_IC_FREESTREAM_GAUSSIAN;
            else if (ic == "freestream_uniform") c.ic_type = NS_IC_FREESTREAM_UNIFORM;
            else throw ParseError("initialCondition must be stagnation, uniform_oblique, freestream_gaussian or freestream_uniform");
        }
        if (r.has("perturbationSeed")) c.ic_seed = (uint64_t)r.integer("perturbationSeed");
        if (r.has("perturbationAmplitude")) c.ic_amplitude = r.real("perturbationAmplitude");
        if (r.has("inletRampTime")) c.inlet_ramp_time = r.real("inletRampTime");
        if (r.has("precision")) {
            std::string pr = r.str("precision");
            if (pr == "fp64") c.precision = NS_FP64; else if (pr == "fp32") c.precision = NS_FP32;
            else throw ParseError("precision must be fp64 or fp32");
        }
        if (r.has("kahan")) c.kahan = r.boolean("kahan");
        nsfsi::setDerived(c);
        std::string v = nsfsi::validateConfig(c);
        if (!v.empty()) throw ParseError(v);
    } catch (const std::exception& e) {
        return report(e.what());
    }
    *out = c;
    return NS_OK;
}
