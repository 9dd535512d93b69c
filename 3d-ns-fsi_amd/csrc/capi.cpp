/*
 * capi.cpp — the extern "C" boundary (include/ns_fsi.h). Converts exceptions into status codes and
 * keeps the last error text per handle; nothing throws across this boundary.
 */
#include <cstring>
#include <memory>
#include <string>

#include "../../include/ns_fsi.h"
#include "case.h"

struct ns_handle {
    std::unique_ptr<nsfsi::Case> c;
    std::string err;
};

namespace {

template <class F>
ns_status guarded(ns_handle* h, F f) {
    if (!h) return NS_ERR_ARGUMENT;
    try {
        f();
        return NS_OK;
    } catch (const nsfsi::ConfigError& e) {
        h->err = e.what();
        return NS_ERR_CONFIG;
    } catch (const nsfsi::GeometryError& e) {
        h->err = e.what();
        return NS_ERR_GEOMETRY;
    } catch (const nsfsi::HipError& e) {
        h->err = e.what();
        return NS_ERR_HIP;
    } catch (const nsfsi::StateError& e) {
        h->err = e.what();
        return NS_ERR_STATE;
    } catch (const nsfsi::Unsupported& e) {
        h->err = e.what();
        return NS_ERR_UNSUPPORTED;
    } catch (const std::invalid_argument& e) {
        h->err = e.what();
        return NS_ERR_ARGUMENT;
    } catch (const std::exception& e) {
        h->err = e.what();
        return NS_ERR_HIP;
    }
}

nsfsi::Case& live(ns_handle* h) {
    if (!h->c) throw nsfsi::StateError("handle has no case");
    return *h->c;
}

}  // namespace

extern "C" {

int32_t ns_abi_version(void) { return NS_FSI_ABI_VERSION; }

ns_status ns_create(const ns_config* cfg, ns_handle** out) {
    if (!cfg || !out) return NS_ERR_ARGUMENT;
    *out = nullptr;
    auto* h = new ns_handle();
    ns_status s = guarded(h, [&] { h->c = std::make_unique<nsfsi::Case>(*cfg); });
    if (s != NS_OK) {
        // keep the handle so the caller can read the error text; it owns no device memory
        *out = h;
        return s;
    }
    *out = h;
    return NS_OK;
}

ns_status ns_init(ns_handle* h) {
    return guarded(h, [&] { live(h).init(); });
}

ns_status ns_set_initial_field(ns_handle* h, const double* U5) {
    return guarded(h, [&] {
        if (!U5) throw std::invalid_argument("null field");
        auto& c = live(h);
        if (!c.initialized) throw nsfsi::StateError("set_initial_field before init");
        c.setInitialField(U5);
    });
}

ns_status ns_step(ns_handle* h, int32_t nsteps, double* t, uint64_t* level, double* dt) {
    ns_status s = guarded(h, [&] {
        if (nsteps < 0) throw std::invalid_argument("negative step count");
        auto& c = live(h);
        c.step(nsteps);
        if (t) *t = c.t;
        if (level) *level = c.level;
        if (dt) *dt = c.dt;
    });
    if (s == NS_OK && h->c->nonFinite) {
        h->err = "Non-finite value found in the conserved variables";
        return NS_ERR_NONFINITE;
    }
    return s;
}

ns_status ns_run(ns_handle* h, double* t, uint64_t* level) {
    ns_status s = guarded(h, [&] {
        auto& c = live(h);
        if (!c.initialized) throw nsfsi::StateError("run before init");
        // Simulation::run / checkStoppingCriterion, src/Simulation.cpp:19-52
        if (c.cfg.stop_criterion == 0) {
            if (c.level < c.cfg.stop_time_level) c.step((int)(c.cfg.stop_time_level - c.level));
        } else {
            while (c.t < c.cfg.t_end && !c.nonFinite) c.step(1);
        }
        if (t) *t = c.t;
        if (level) *level = c.level;
    });
    if (s == NS_OK && h->c->nonFinite) {
        h->err = "Non-finite value found in the conserved variables";
        return NS_ERR_NONFINITE;
    }
    return s;
}

ns_status ns_get_time(ns_handle* h, double* t, uint64_t* level, double* dt) {
    return guarded(h, [&] {
        auto& c = live(h);
        if (t) *t = c.t;
        if (level) *level = c.level;
        if (dt) *dt = c.dt;
    });
}

ns_status ns_get_field(ns_handle* h, int32_t which, double* out) {
    return guarded(h, [&] {
        if (!out || which < 0 || which > 11) throw std::invalid_argument("bad field id or null output");
        live(h).getField(which, out);
    });
}

ns_status ns_get_node_types(ns_handle* h, uint8_t* out) {
    return guarded(h, [&] {
        if (!out) throw std::invalid_argument("null output");
        live(h).nodeTypes(out);
    });
}

ns_status ns_get_counts(ns_handle* h, int64_t* n_ghosts, int64_t* n_filter, int64_t* n_interior) {
    return guarded(h, [&] {
        auto& c = live(h);
        if (!c.initialized) throw nsfsi::StateError("counts before init");
        if (n_ghosts) *n_ghosts = c.nGhost;
        if (n_filter) *n_filter = c.nFilter;
        if (n_interior) *n_interior = c.interiorCount();
    });
}

ns_status ns_get_ghosts(ns_handle* h, int32_t* ijk, double* bi, double* ip, int32_t* box) {
    return guarded(h, [&] { live(h).ghosts(ijk, bi, ip, box); });
}

ns_status ns_get_filter_nodes(ns_handle* h, int32_t* ijk) {
    return guarded(h, [&] {
        if (!ijk) throw std::invalid_argument("null output");
        live(h).filterNodes(ijk);
    });
}

ns_status ns_norm_history(ns_handle* h, int32_t var, double* out, int64_t n, int64_t* count) {
    return guarded(h, [&] {
        if (var < 0 || var > 4) throw std::invalid_argument("var must be 0..4");
        auto& c = live(h);
        c.readScalars();
        const auto& v = c.normsHost[var];
        int64_t avail = (int64_t)v.size();
        bool stored = c.cfg.conv_history == 2 || (c.cfg.conv_history == 1 && var == 0);
        if (!stored) avail = 0;
        if (count) *count = avail;
        for (int64_t q = 0; out && q < n && q < avail; ++q) out[q] = v[q];
    });
}

ns_status ns_force_history(ns_handle* h, double* out, int64_t n, int64_t* count) {
    return guarded(h, [&] {
        auto& c = live(h);
        c.readScalars();
        const int64_t avail = (int64_t)(c.forceHost.size() / 4);
        if (count) *count = avail;
        for (int64_t q = 0; out && q < n && q < avail; ++q)
            for (int v = 0; v < 4; ++v) out[4 * q + v] = c.forceHost[4 * q + v];
    });
}

ns_status ns_totals(ns_handle* h, double out[5]) {
    return guarded(h, [&] {
        if (!out) throw std::invalid_argument("null output");
        auto& c = live(h);
        if (!c.initialized) throw nsfsi::StateError("totals before init");
        c.totals(out);
    });
}

ns_status ns_get_geometry(ns_handle* h, double spacing[3], double offset[3]) {
    return guarded(h, [&] {
        auto& c = live(h);
        if (!c.initialized) throw nsfsi::StateError("geometry before init");
        if (spacing) { spacing[0] = c.P.hx; spacing[1] = c.P.hy; spacing[2] = c.P.hz; }
        if (offset) { offset[0] = c.P.ox; offset[1] = c.P.oy; offset[2] = c.P.oz; }
    });
}

ns_status ns_profile(ns_handle* h, int32_t nsteps, double ms[4], int32_t launches[4]) {
    return guarded(h, [&] {
        if (!ms || !launches || nsteps < 1) throw std::invalid_argument("bad profile arguments");
        int l[4];
        live(h).profile(nsteps, ms, l);
        for (int g = 0; g < 4; ++g) launches[g] = l[g];
    });
}

ns_status ns_set_option(ns_handle* h, const char* name, int64_t value) {
    return guarded(h, [&] {
        if (!name) throw std::invalid_argument("null option name");
        live(h).setOption(name, value);
    });
}

ns_status ns_get_tile(ns_handle* h, int32_t* ty, int32_t* kc) {
    return guarded(h, [&] {
        auto& c = live(h);
        if (ty) *ty = c.tile.ty;
        if (kc) *kc = c.tile.kc;
    });
}

const char* ns_last_error(ns_handle* h) {
    if (!h) return "null handle";
    return h->err.c_str();
}

void ns_destroy(ns_handle* h) { delete h; }

}  // extern "C"
