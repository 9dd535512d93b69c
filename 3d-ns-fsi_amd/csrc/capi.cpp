/*
 * capi.cpp — the extern "C" boundary (include/ns_fsi.h). Converts exceptions into status codes and
 * keeps the last error text per handle; nothing throws across this boundary.
 */
#include <cstring>
#include <memory>
#include <string>

#include "../../include/ns_fsi.h"
#include "batch_host.h"
#include "case.h"

struct ns_handle {
    std::unique_ptr<nsfsi::Case> c;
    std::string err;
};

struct ns_batch {
    std::unique_ptr<nsfsi::Batch> b;
    std::string err;
};

namespace {

template <class H, class F>
ns_status guarded(H* h, F f) {
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

nsfsi::Batch& live(ns_batch* h) {
    if (!h->b) throw nsfsi::StateError("handle has no batch");
    return *h->b;
}

nsfsi::Case& batchCase(ns_batch* h, int32_t c) {
    auto& b = live(h);
    if (c < 0 || c >= (int32_t)b.cases.size()) throw std::invalid_argument("case index out of range");
    return *b.cases[c];
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

ns_status ns_set_dt_schedule(ns_handle* h, const double* dts, int64_t n) {
    return guarded(h, [&] { live(h).setDtSchedule(dts, n); });
}

ns_status ns_set_inflow_profile(ns_handle* h, const double* q, int64_t n) {
    return guarded(h, [&] { live(h).setInflowProfile(q, n); });
}

ns_status ns_adjoint_drag(ns_handle* h, int32_t nsteps, int32_t n0, int32_t n1, int32_t checkpoint, double* J,
                          double* grad_u0, double* grad_inflow, double* grad_re, double* dts) {
    return guarded(h, [&] { live(h).adjointDrag(nsteps, n0, n1, checkpoint, J, grad_u0, grad_inflow, grad_re, dts); });
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

ns_status ns_save(ns_handle* h, const char* path) {
    return guarded(h, [&] {
        if (!path) throw std::invalid_argument("null path");
        live(h).save(path);
    });
}

ns_status ns_restore(ns_handle* h, const char* path) {
    return guarded(h, [&] {
        if (!path) throw std::invalid_argument("null path");
        live(h).restore(path);
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

ns_status ns_batch_create(const ns_config* cfgs, int32_t n, ns_batch** out) {
    if (!cfgs || !out || n < 1) return NS_ERR_ARGUMENT;
    auto* h = new ns_batch();
    *out = h;
    return guarded(h, [&] { h->b = std::make_unique<nsfsi::Batch>(cfgs, n); });
}

ns_status ns_batch_init(ns_batch* h) {
    return guarded(h, [&] { live(h).init(); });
}

ns_status ns_batch_step(ns_batch* h, int32_t nsteps) {
    ns_status s = guarded(h, [&] {
        if (nsteps < 0) throw std::invalid_argument("negative step count");
        live(h).step(nsteps);
    });
    if (s == NS_OK)
        for (size_t c = 0; c < h->b->cases.size(); ++c)
            if (h->b->cases[c]->nonFinite) {
                h->err = "Non-finite value found in the conserved variables of case " + std::to_string(c);
                return NS_ERR_NONFINITE;
            }
    return s;
}

ns_status ns_batch_get_time(ns_batch* h, int32_t c, double* t, uint64_t* level, double* dt) {
    return guarded(h, [&] {
        auto& k = batchCase(h, c);
        if (t) *t = k.t;
        if (level) *level = k.level;
        if (dt) *dt = k.dt;
    });
}

ns_status ns_batch_get_field(ns_batch* h, int32_t c, int32_t which, double* out) {
    return guarded(h, [&] {
        if (!out || which < 0 || which > 11) throw std::invalid_argument("bad field id or null output");
        batchCase(h, c).getField(which, out);
    });
}

ns_status ns_batch_norm_history(ns_batch* h, int32_t c, int32_t var, double* out, int64_t n, int64_t* count) {
    return guarded(h, [&] {
        if (var < 0 || var > 4) throw std::invalid_argument("var must be 0..4");
        auto& k = batchCase(h, c);
        const auto& v = k.normsHost[var];
        int64_t avail = (int64_t)v.size();
        bool stored = k.cfg.conv_history == 2 || (k.cfg.conv_history == 1 && var == 0);
        if (!stored) avail = 0;
        if (count) *count = avail;
        for (int64_t q = 0; out && q < n && q < avail; ++q) out[q] = v[q];
    });
}

ns_status ns_batch_set_option(ns_batch* h, const char* name, int64_t value) {
    return guarded(h, [&] {
        if (!name) throw std::invalid_argument("null option name");
        live(h).setOption(name, value);
    });
}

ns_status ns_batch_get_tile(ns_batch* h, int32_t* ty, int32_t* kc) {
    return guarded(h, [&] {
        auto& b = live(h);
        if (ty) *ty = b.tile.ty;
        if (kc) *kc = b.tile.kc;
    });
}

const char* ns_batch_last_error(ns_batch* h) {
    if (!h) return "null handle";
    return h->err.c_str();
}

void ns_batch_destroy(ns_batch* h) { delete h; }

}  // extern "C"
