/*
 * sanitize_check.cpp — TEST INFRASTRUCTURE ONLY. Drives the serial oracle (ns_oracle.cpp) and the host-side
 * ConfigFile parser of the device library (3d-ns-fsi_amd/csrc/ConfigSettings.cpp) through every entry point
 * on the geometries the parity suite uses, built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`). A memory error or undefined behaviour aborts the run with the sanitizer's
 * report; a clean run prints "sanitize ok". The reference's own latent defects this must not reproduce:
 * the sphere augmentation iterator (src/Boundary.cpp:1041-1042) and the size_t wrap of body boxes near
 * index 0 (src/Boundary.cpp:945-963, :1061-1070), both replaced by checked code in the oracle.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/ns_fsi.h"
#include "../3d-ns-fsi_amd/csrc/ns_internal.h"

extern "C" {
int ns_oracle_create(const ns_config* cfg, void** out);
int ns_oracle_initialize(void* h);
int ns_oracle_march(void* h, int nsteps);
int ns_oracle_run(void* h);
void ns_oracle_state(void* h, double* t, uint64_t* level, double* dt);
void ns_oracle_geometry(void* h, double* spacing, double* offset);
int ns_oracle_get_field(void* h, int which, double* out);
int ns_oracle_set_field(void* h, int which, const double* in);
int ns_oracle_derive_all(void* h);
int ns_oracle_slopes(void* h, double* out);
int ns_oracle_apply_bcs(void* h, double tnow);
int ns_oracle_get_node_types(void* h, uint8_t* out);
void ns_oracle_counts(void* h, int64_t* nGhost, int64_t* nFilter, int64_t* nInterior);
int ns_oracle_get_ghosts(void* h, int32_t* ijk, double* bi, double* ip, int32_t* box);
int ns_oracle_get_filter_nodes(void* h, int32_t* ijk);
int64_t ns_oracle_norm_history(void* h, int var, double* out, int64_t n);
const char* ns_oracle_last_error(void* h);
void ns_oracle_destroy(void* h);
int ns_oracle_set_time(void* h, double t, uint64_t level);
int ns_oracle_filter_now(void* h);
int ns_oracle_rho_filter(void* h);
double ns_oracle_norm_of_change(void* h, int q);
double ns_oracle_update_dt(void* h, double tnow);
void ns_oracle_lu8_solve(const double* A, const double* b, double* x);
double ns_oracle_uniform01(uint64_t seed, uint64_t counter);
int ns_oracle_tl_create(const ns_config* cfg, void** out);
int ns_oracle_tl_initialize(void* h);
int ns_oracle_tl_set_state(void* h, const double* val, const double* tan);
int ns_oracle_tl_get_state(void* h, double* val, double* tan);
int ns_oracle_tl_set_dt_schedule(void* h, const double* dts, int64_t n);
int ns_oracle_tl_march(void* h, int nsteps);
int ns_oracle_tl_slopes(void* h, double* val, double* tan);
void ns_oracle_tl_state(void* h, double* t, uint64_t* level, double* dt);
const char* ns_oracle_tl_last_error(void* h);
void ns_oracle_tl_destroy(void* h);
}

static int failures = 0;
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "check failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

static ns_config sideWalls(int NI, int NJ, int NK, double L) {
    ns_config c;
    ns_config_default(&c);
    c.NI = NI; c.NJ = NJ; c.NK = NK;
    c.L_x = c.L_y = c.L_z = L;
    c.bc[NS_XMIN] = NS_BC_INLET; c.bc[NS_XMAX] = NS_BC_OUTLET;
    for (int f = 2; f < 6; ++f) c.bc[f] = NS_BC_SYMMETRY;
    c.body_type = NS_BODY_SPHERE;
    c.body_centre[0] = c.body_centre[1] = c.body_centre[2] = L / 2;
    c.body_radius = 0.5;
    c.M_0 = 0.3; c.Re = 333.3; c.conv_stability_limit = 0.5;
    c.ic_type = NS_IC_FREESTREAM_GAUSSIAN; c.ic_amplitude = 1e-3; c.ic_seed = 1;
    c.stop_criterion = 0; c.stop_time_level = 4;
    c.n_filter_intervals = 1; c.filter_intervals[0] = 0; c.n_filter_change_times = 0;
    c.conv_history = 2;
    nsfsi::setDerived(c);
    return c;
}

static void exercise(const char* name, const ns_config& c, int steps) {
    void* h = nullptr;
    CHECK(ns_oracle_create(&c, &h) == NS_OK);
    if (!h) return;
    int s = ns_oracle_initialize(h);
    if (s != NS_OK) {
        std::fprintf(stderr, "%s: initialize %d %s\n", name, s, ns_oracle_last_error(h));
        ++failures;
        ns_oracle_destroy(h);
        return;
    }
    const size_t N = (size_t)c.NI * c.NJ * c.NK;
    std::vector<double> f(N), K(5 * N);
    std::vector<uint8_t> nt(N);
    int64_t ng = 0, nf = 0, ni = 0;
    ns_oracle_counts(h, &ng, &nf, &ni);
    std::vector<int32_t> ijk(3 * ng + 3), box(3 * ng + 3), fijk(3 * nf + 3);
    std::vector<double> bi(3 * ng + 3), ip(3 * ng + 3);
    CHECK(ns_oracle_get_ghosts(h, ijk.data(), bi.data(), ip.data(), box.data()) == NS_OK);
    CHECK(ns_oracle_get_filter_nodes(h, fijk.data()) == NS_OK);
    CHECK(ns_oracle_get_node_types(h, nt.data()) == NS_OK);
    CHECK(ns_oracle_slopes(h, K.data()) == NS_OK);
    CHECK(ns_oracle_apply_bcs(h, 0.25) == NS_OK);
    CHECK(ns_oracle_march(h, steps) == NS_OK);
    for (int w = 0; w < 17; ++w) CHECK(ns_oracle_get_field(h, w, f.data()) == NS_OK);
    CHECK(ns_oracle_set_field(h, 16, f.data()) == NS_OK);
    CHECK(ns_oracle_derive_all(h) == NS_OK);
    for (int q = 0; q < 5; ++q) CHECK(std::isfinite(ns_oracle_norm_of_change(h, q)));
    std::vector<double> hist(steps + 1);
    for (int q = 0; q < 5; ++q) CHECK(ns_oracle_norm_history(h, q, hist.data(), steps) >= 0);
    CHECK(ns_oracle_set_time(h, 1.0, 3) == NS_OK);
    CHECK(ns_oracle_filter_now(h) >= 0);
    CHECK(ns_oracle_rho_filter(h) >= 0);
    CHECK(ns_oracle_update_dt(h, 1.0) > 0);
    double t, dt, sp[3], of[3];
    uint64_t lv;
    ns_oracle_state(h, &t, &lv, &dt);
    ns_oracle_geometry(h, sp, of);
    ns_oracle_destroy(h);
    // the tangent-linear form: same set-up, a tangent on every state value, two steps at the first dt
    void* g = nullptr;
    CHECK(ns_oracle_tl_create(&c, &g) == NS_OK);
    if (g) {
        CHECK(ns_oracle_tl_initialize(g) == NS_OK);
        std::vector<double> v(5 * N), d(5 * N);
        CHECK(ns_oracle_tl_get_state(g, v.data(), d.data()) == NS_OK);
        for (size_t x = 0; x < d.size(); ++x) d[x] = (double)((x * 2654435761u) % 1000) * 1e-3 - 0.5;
        CHECK(ns_oracle_tl_set_state(g, v.data(), d.data()) == NS_OK);
        double tt, dd;
        uint64_t ll;
        ns_oracle_tl_state(g, &tt, &ll, &dd);
        const double dts[2] = {dd, dd};
        CHECK(ns_oracle_tl_set_dt_schedule(g, dts, 2) == NS_OK);
        CHECK(ns_oracle_tl_march(g, 2) == NS_OK);
        CHECK(ns_oracle_tl_slopes(g, v.data(), d.data()) == NS_OK);
        CHECK(ns_oracle_tl_last_error(g) != nullptr);
        ns_oracle_tl_destroy(g);
    }
    std::printf("%-28s ghosts %6lld filter %6lld interior %8lld  dt %.6e\n", name, (long long)ng, (long long)nf,
                (long long)ni, dt);
    (void)t;
    (void)lv;
}

int main(int argc, char** argv) {
    // 1. sphere with slip side walls (48^3 workload at 24^3)
    exercise("sphere 24^3", sideWalls(24, 24, 24, 4.0), 6);
    // 2. z-cylinder, off-centre sphere, x-cylinder with the box at the mesh's low corner
    {
        ns_config c = sideWalls(64, 64, 64, 8.0);
        c.body_type = NS_BODY_CYLINDER; c.body_axis = 2; c.body_centre[2] = 0;
        nsfsi::setDerived(c);
        exercise("z-cylinder 64^3", c, 2);
        ns_config d = sideWalls(33, 29, 31, 3.0);
        d.L_y = 2.5; d.L_z = 2.7; d.body_centre[0] = 1.31; d.body_centre[1] = 1.27; d.body_centre[2] = 1.33;
        d.body_radius = 0.61;
        nsfsi::setDerived(d);
        exercise("off-centre sphere", d, 4);
        ns_config e = sideWalls(30, 26, 28, 3.0);
        e.L_y = 2.0; e.L_z = 2.2; e.body_type = NS_BODY_CYLINDER; e.body_axis = 0;
        e.body_centre[0] = 0.0; e.body_centre[1] = 1.03; e.body_centre[2] = 1.11; e.body_radius = 0.41;
        nsfsi::setDerived(e);
        exercise("x-cylinder", e, 4);
    }
    // 3. the reference's shipped case with its periodic density filter (fires at step 50)
    {
        ns_config c;
        ns_config_default(&c);
        c.stop_criterion = 0; c.stop_time_level = 52;
        exercise("reference default 161x83x3", c, 52);
        c.stop_criterion = 1; c.t_end = 0.05;
        void* h = nullptr;
        CHECK(ns_oracle_create(&c, &h) == NS_OK);
        CHECK(ns_oracle_initialize(h) == NS_OK);
        CHECK(ns_oracle_run(h) == NS_OK);
        double t;
        ns_oracle_state(h, &t, nullptr, nullptr);
        CHECK(t == 0.05);
        ns_oracle_destroy(h);
    }
    // 4. inflow along y, and periodic x with inflow along z; a body box that leaves the mesh must be an error
    {
        ns_config c = sideWalls(30, 34, 26, 3.0);
        for (int f = 0; f < 6; ++f) c.bc[f] = NS_BC_SYMMETRY;
        c.bc[NS_YMIN] = NS_BC_INLET; c.bc[NS_YMAX] = NS_BC_OUTLET;
        c.body_radius = 0.45;
        nsfsi::setDerived(c);
        exercise("y-inlet sphere", c, 3);
        ns_config d = sideWalls(28, 26, 36, 3.0);
        d.bc[NS_XMIN] = d.bc[NS_XMAX] = NS_BC_PERIODIC;
        d.bc[NS_ZMIN] = NS_BC_INLET; d.bc[NS_ZMAX] = NS_BC_OUTLET;
        d.body_type = NS_BODY_CYLINDER; d.body_axis = 0; d.body_radius = 0.4;
        nsfsi::setDerived(d);
        exercise("periodic-x z-inlet", d, 3);
        ns_config e = sideWalls(20, 20, 20, 2.0);
        e.body_centre[0] = 0.05;
        nsfsi::setDerived(e);
        void* h = nullptr;
        CHECK(ns_oracle_create(&e, &h) == NS_OK);
        CHECK(ns_oracle_initialize(h) == NS_ERR_GEOMETRY);
        ns_oracle_destroy(h);
    }
    // 4b. three bodies (two spheres and a z-cylinder), each with its own ghost map and filter list
    {
        ns_config c = sideWalls(48, 40, 26, 4.8);
        c.L_y = 4.0; c.L_z = 2.6;
        c.body_centre[0] = 1.0; c.body_centre[1] = 1.0; c.body_centre[2] = 1.3; c.body_radius = 0.35;
        c.n_extra_bodies = 2;
        c.extra_bodies[0] = ns_body{NS_BODY_SPHERE, 0, {1.0, 3.0, 1.3}, 0.35};
        c.extra_bodies[1] = ns_body{NS_BODY_CYLINDER, 2, {3.4, 2.0, 0.0}, 0.4};
        nsfsi::setDerived(c);
        exercise("three bodies", c, 3);
    }
    // 5. fully periodic box, filter every 3 steps
    {
        ns_config c = sideWalls(14, 15, 16, 1.0);
        for (int f = 0; f < 6; ++f) c.bc[f] = NS_BC_PERIODIC;
        c.body_type = NS_BODY_NONE;
        c.filter_intervals[0] = 3;
        nsfsi::setDerived(c);
        exercise("periodic box, filter", c, 7);
    }
    // 6. 8x8 solves and the generator
    {
        double A[64], b[8], x[8];
        for (int r = 0; r < 8; ++r) {
            b[r] = ns_oracle_uniform01(9, 100 + r);
            for (int q = 0; q < 8; ++q) A[r * 8 + q] = ns_oracle_uniform01(9, r * 8 + q) + (r == q ? 2 : 0);
        }
        ns_oracle_lu8_solve(A, b, x);
        for (int r = 0; r < 8; ++r) {
            double s = 0;
            for (int q = 0; q < 8; ++q) s += A[r * 8 + q] * x[q];
            CHECK(std::fabs(s - b[r]) < 1e-12);
        }
        std::memset(A, 0, sizeof(A));
        ns_oracle_lu8_solve(A, b, x);  // singular: zeros, no division by zero
        for (double v : x) CHECK(v == 0.0);
    }
    // 7. the host ConfigFile parser on the fixture and on malformed input
    if (argc > 1) {
        FILE* fp = std::fopen(argv[1], "rb");
        CHECK(fp != nullptr);
        if (fp) {
            std::string text;
            char buf[4096];
            size_t n;
            while ((n = std::fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, n);
            std::fclose(fp);
            ns_config c, d;
            char err[256];
            CHECK(ns_config_parse(text.c_str(), &c, err, sizeof(err)) == NS_OK);
            ns_config_default(&d);
            CHECK(c.NI == d.NI && c.Re == d.Re && c.n_filter_intervals == d.n_filter_intervals);
            const char* bad[] = {"NI = 161.", "Gamma = ", "x = [1, 2", "/* open comment", "s = \"unterminated",
                                 "NI = 0x", "a = {b = (1, 2.5, \"c\")}", "NI = 99999999999999999999", "", "=", ";;;"};
            for (const char* t : bad) {
                std::string s = text + "\n" + t + "\n";
                ns_config_parse(s.c_str(), &c, err, sizeof(err));
                ns_config_parse(t, &c, err, sizeof(err));
                ns_config_parse(t, &c, nullptr, 0);
            }
            for (size_t cut = 0; cut < text.size(); cut += 37) {
                std::string s = text.substr(0, cut);
                ns_config_parse(s.c_str(), &c, err, sizeof(err));
            }
        }
    }
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
