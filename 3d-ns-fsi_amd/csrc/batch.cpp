/*
 * batch.cpp — many independent cases (own grid size, body, Mach, Re, initial field and time step)
 * advanced together on one GPU: every kernel of the time step runs once for all cases (the fused stage
 * over the union of the cases' tiles, the boundary kernels over the union of their face, filter and
 * ghost nodes, one finalising block per case), so a step costs ~30 launches whatever the case count,
 * and the step replays as one captured HIP graph. Each case keeps its own device dt, t and timeLevel
 * (src/Solver.cpp:80-126 per case), so cases with different stability limits advance at their own pace.
 *
 * The cases are ordinary single-case objects (classification, ghost geometry, initial field and first dt
 * exactly as for one case); the batch reads their device buffers through a table built once.
 */
#include "batch_host.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace nsfsi {

namespace {
constexpr int kRoleU = 0, kRoleA = 1, kRoleS1 = 2, kRoleS2 = 3;
}

Batch::Batch(const ns_config* cfgs, int n) {
    if (n <= 0) throw std::invalid_argument("batch needs at least one case");
    for (int c = 0; c < n; ++c) {
        const ns_config& k = cfgs[c];
        if (k.precision != NS_FP32) throw Unsupported("batched cases run in FP32 (precision = fp32)");
        if (k.kahan != cfgs[0].kahan) throw Unsupported("all batched cases need the same kahan setting");
        if (k.n_extra_bodies != 0) throw Unsupported("batched cases hold at most one immersed body");
        cases.push_back(std::make_unique<Case>(k));
    }
    kahan = cfgs[0].kahan != 0;
    hipCheck(hipStreamCreate(&stream), "stream");
}

Batch::~Batch() {
    // a destructor cannot throw: queued work is waited for, failed releases are reported on stderr
    auto report = [](hipError_t e, const char* what) {
        if (e != hipSuccess) std::fprintf(stderr, "nsfsi: %s failed in the batch destructor: %s\n", what, hipGetErrorString(e));
    };
    if (stream) report(hipStreamSynchronize(stream), "hipStreamSynchronize");
    for (int q = 0; q < 2; ++q) {
        if (graphExec[q]) report(hipGraphExecDestroy(graphExec[q]), "hipGraphExecDestroy");
        graphExec[q] = nullptr;
    }
    for (void* p : allocations) report(hipFree(p), "hipFree");
    if (stream) report(hipStreamDestroy(stream), "hipStreamDestroy");
}

void* Batch::alloc(size_t bytes) {
    void* p = nullptr;
    hipCheck(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc");
    allocations.push_back(p);
    return p;
}

void Batch::dropGraphs() {
    if (graphExec[0] || graphExec[1]) hipCheck(hipStreamSynchronize(stream), "sync before dropping graphs");
    for (int q = 0; q < 2; ++q) {
        if (graphExec[q]) (void)hipGraphExecDestroy(graphExec[q]);
        graphExec[q] = nullptr;
    }
}

void Batch::init() {
    for (auto& c : cases) {
        c->init();
        if (c->nFixed > 0)
            throw Unsupported("a body whose solid nodes reach a symmetry face is not available in batched runs");
    }
    const int n = (int)cases.size();
    // device views (buffers in their initial roles; parity 0)
    std::vector<CaseView<float>> views(n);
    std::vector<int64_t> fstart[6], filt(n + 1, 0), gh(n + 1, 0);
    for (auto& v : fstart) v.assign(n + 1, 0);
    std::vector<int> bb(n + 1, 0), ib(n + 1, 0);
    for (int c = 0; c < n; ++c) {
        Case& k = *cases[c];
        CaseView<float>& v = views[c];
        std::memset(&v, 0, sizeof v);
        v.L = k.L;
        v.P = k.P;
        v.W = k.W;
        v.faceStart[0] = 0;
        for (int f = 0; f < 6; ++f) {
            v.faces[f] = k.faces[f];
            const EdgeFace& e = k.faces[f];
            int64_t cnt = std::max<int64_t>(0, (int64_t)(e.i1 - e.i0 + 1)) * std::max<int64_t>(0, (int64_t)(e.j1 - e.j0 + 1)) *
                          std::max<int64_t>(0, (int64_t)(e.k1 - e.k0 + 1));
            v.faceStart[f + 1] = v.faceStart[f] + cnt;
            // the edge-face pass skips the outlet face that the case's stage kernel applies itself
            fstart[f][c + 1] = fstart[f][c] + (k.stageOutlet() == &k.faces[f] ? 0 : cnt);
        }
        v.cls = k.cls;
        v.buf[0] = (float*)k.U;
        v.buf[1] = (float*)k.A;
        v.buf[2] = (float*)k.S1;
        v.buf[3] = (float*)k.S2;
        v.rhoPre = (float*)k.rhoPre;
        v.rhoEPre = (float*)k.rhoEPre;
        v.ftmp = (float*)k.ftmp;
        v.filterList = k.filterList;
        v.nFilter = k.nFilter;
        v.G = k.G;
        v.nGhost = k.nGhost;
        v.inactiveList = k.inactiveList;
        v.inactiveIC = k.inactiveIC;
        v.fs = FilterSchedule{};
        v.fs.n_intervals = k.cfg.n_filter_intervals;
        v.fs.n_change = k.cfg.n_filter_change_times;
        for (int q = 0; q < NS_MAX_FILTER; ++q) {
            v.fs.intervals[q] = k.cfg.filter_intervals[q];
            v.fs.change_times[q] = k.cfg.filter_change_times[q];
        }
        v.nInactive = k.nInactive;
        v.scal = k.scal;
        v.partials = k.partials;
        v.hist = k.hist;
        v.histCap = k.histCap;
        v.convHistory = k.cfg.conv_history;
        v.stopEndTime = k.cfg.stop_criterion == 1;
        v.tEnd = k.cfg.t_end;
        v.cfl = k.cfg.conv_stability_limit;
        v.sigma = k.cfg.visc_stability_limit;
        v.norms = k.cfg.conv_history > 0;
        const int64_t nb = v.faceStart[6] + v.nGhost + v.nFilter;
        // nStage4 (set with the tile) + these must fit the case's partial slots; checked in buildStageTables
        v.nBoundaryBlocks = (int)std::min<int64_t>(64, std::max<int64_t>(1, (nb + 1023) / 1024));
        v.nInactiveBlocks = v.nInactive > 0 ? inactiveReduceBlocks(v.nInactive) : 0;
        filt[c + 1] = filt[c] + v.nFilter;
        gh[c + 1] = gh[c] + v.nGhost;
        bb[c + 1] = bb[c] + v.nBoundaryBlocks;
        ib[c + 1] = ib[c] + v.nInactiveBlocks;
    }
    hostViews = views;
    // cases with a periodic density filter, and their nodes (x fastest) as one thread range
    std::vector<int> fids;
    std::vector<int64_t> fstartN(1, 0);
    for (int c = 0; c < n; ++c)
        if (!cases[c]->filterDisabled()) {
            fids.push_back(c);
            fstartN.push_back(fstartN.back() + (int64_t)cases[c]->L.nodes());
        }
    nFiltCases = (int)fids.size();
    filtNodeTotal = fstartN.back();
    if (nFiltCases > 0) {
        devFiltIds = (int*)alloc(fids.size() * sizeof(int));
        devFiltStart = (int64_t*)alloc(fstartN.size() * sizeof(int64_t));
        hipCheck(hipMemcpy(devFiltIds, fids.data(), fids.size() * sizeof(int), hipMemcpyHostToDevice), "copy");
        hipCheck(hipMemcpy(devFiltStart, fstartN.data(), fstartN.size() * sizeof(int64_t), hipMemcpyHostToDevice), "copy");
    }
    devViews = (CaseView<float>*)alloc(n * sizeof(CaseView<float>));
    for (int f = 0; f < 6; ++f) {
        devFaceStart[f] = (int64_t*)alloc((n + 1) * sizeof(int64_t));
        faceTotal[f] = fstart[f][n];
        hipCheck(hipMemcpy(devFaceStart[f], fstart[f].data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice), "copy");
    }
    // faces in application order, grouped greedily: a face joins the current launch when it is independent
    // of every face already in it in every case
    faceGroups.clear();
    EdgeFaceListB g{};
    for (int f = 0; f < 6; ++f) {
        if (faceTotal[f] == 0) continue;
        bool ok = true;
        for (int m = 0; m < g.n && ok; ++m)
            for (int c = 0; c < n && ok; ++c) ok = edgeFacesIndependent(views[c].L, views[c].faces[g.face[m]], views[c].faces[f]);
        if (!ok) {
            faceGroups.push_back(g);
            g = EdgeFaceListB{};
        }
        g.face[g.n] = f;
        g.start[g.n] = devFaceStart[f];
        g.off[g.n + 1] = g.off[g.n] + faceTotal[f];
        ++g.n;
    }
    if (g.n > 0) faceGroups.push_back(g);
    devFilterStart = (int64_t*)alloc((n + 1) * sizeof(int64_t));
    devGhostStart = (int64_t*)alloc((n + 1) * sizeof(int64_t));
    devBBStart = (int*)alloc((n + 1) * sizeof(int));
    devIBStart = (int*)alloc((n + 1) * sizeof(int));
    filterTotal = filt[n];
    ghostTotal = gh[n];
    totalBB = bb[n];
    totalIB = ib[n];
    hipCheck(hipMemcpy(devFilterStart, filt.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice), "copy");
    hipCheck(hipMemcpy(devGhostStart, gh.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice), "copy");
    hipCheck(hipMemcpy(devBBStart, bb.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "copy");
    hipCheck(hipMemcpy(devIBStart, ib.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "copy");
    const size_t ab = fusedArgsBytes();
    for (int p = 0; p < 2; ++p)
        for (int s = 0; s < 4; ++s) devArgs[p][s] = alloc(n * ab);
    devStageStart = (int*)alloc((n + 1) * sizeof(int));
    if (!devCtr) {
        devCtr = (unsigned*)alloc(9 * sizeof(unsigned));
        hipCheck(hipMemsetAsync(devCtr, 0, 9 * sizeof(unsigned), stream), "memset");
    }
    buildStageTables();
    initialized = true;
}

// Stage arguments of every case for both buffer parities and the tile shape in use; stage-4 block
// counts go into the case views (their partial slots precede the boundary partials).
void Batch::buildStageTables() {
    const int n = (int)cases.size();
    const size_t ab = fusedArgsBytes();
    std::vector<char> h(n * ab);
    std::vector<int> start(n + 1, 0);
    for (int p = 0; p < 2; ++p)
        for (int st = 1; st <= 4; ++st) {
            for (int c = 0; c < n; ++c) {
                Case& k = *cases[c];
                const CaseView<float>& v = hostViews[c];
                float* U = v.buf[p];
                float* A = v.buf[1 - p];
                float* in[4] = {U, v.buf[2], v.buf[3], v.buf[2]};
                float* out[4] = {v.buf[2], v.buf[3], v.buf[2], A};
                // stage 1 of a case with the density filter takes the filtered density from S2 when it fired
                const bool filt = st == 1 && !k.filterDisabled();
                int nb = buildFusedArgs(h.data() + c * ab, st, k.L, k.P, k.W, in[st - 1], U, out[st - 1], A, k.cls,
                                        (const float*)k.rhoPre, (const float*)k.rhoEPre, &k.scal->dt, tile,
                                        st == 4 ? k.partials : nullptr, k.cfg.conv_history > 0,
                                        kahan ? k.comp : nullptr, k.stageOutlet(), filt ? v.buf[3] : nullptr,
                                        filt ? &k.scal->filterNow : nullptr);
                start[c + 1] = start[c] + nb;
                hostViews[c].nStage4 = nb;
                const size_t need = (size_t)nb + hostViews[c].nBoundaryBlocks + hostViews[c].nInactiveBlocks;
                if (need > k.partialSlots) throw std::invalid_argument("tile too small for the scan partial buffer");
            }
            hipCheck(hipMemcpy(devArgs[p][st - 1], h.data(), n * ab, hipMemcpyHostToDevice), "copy args");
        }
    totalTiles = start[n];
    hipCheck(hipMemcpy(devStageStart, start.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "copy");
    hipCheck(hipMemcpy(devViews, hostViews.data(), n * sizeof(CaseView<float>), hipMemcpyHostToDevice), "copy views");
    dropGraphs();
}

void Batch::applyBCs(int role, double frac) {
    const int n = (int)cases.size();
    for (const EdgeFaceListB& g : faceGroups) launchEdgeFacesBatched<float>(stream, devViews, g, n, role, parity, frac);
    launchFilterComputeBatched<float>(stream, devViews, devFilterStart, filterTotal, n, role, parity);
    launchFilterWriteBatched<float>(stream, devViews, devFilterStart, filterTotal, n, role, parity);
    launchGhostApplyBatched<float>(stream, devViews, devGhostStart, ghostTotal, n, role, parity);
}

// One RK4 step of every case (Solver::marchTimeStep, src/Solver.cpp:130-167, without the density filter).
void Batch::stepOnce() {
    const int n = (int)cases.size();
    const int outRole[4] = {kRoleS1, kRoleS2, kRoleS1, kRoleA};
    const double frac[4] = {0.5, 0.5, 1.0, 1.0};
    // periodic density filter (src/Solver.cpp:134), as the single-case step: decided on the device, the
    // filtered density in S2 for stage 1, committed to U after it
    if (nFiltCases > 0) {
        launchFilterFlagBatched(stream, devViews, devFiltIds, nFiltCases);
        launchRhoFilterBatched(stream, devViews, devFiltIds, devFiltStart, nFiltCases, filtNodeTotal, parity);
    }
    for (int st = 1; st <= 4; ++st) {
        launchStageBatched(stream, st, tile.ty, true, kahan, devArgs[parity][st - 1], devStageStart, n, totalTiles,
                           persistent ? devCtr : nullptr);
        if (st == 1 && nFiltCases > 0)
            launchRhoCommitBatched(stream, devViews, devFiltIds, devFiltStart, nFiltCases, filtNodeTotal, parity);
        applyBCs(outRole[st - 1], frac[st - 1]);
    }
    launchBoundaryReduceBatched<float>(stream, devViews, devBBStart, totalBB, n, parity);
    launchInactiveReduceBatched<float>(stream, devViews, devIBStart, totalIB, n, parity);
    launchFinalizeBatched<float>(stream, devViews, n);
    (void)kRoleU;
}

void Batch::advanceParity() {
    parity ^= 1;
    for (auto& c : cases) std::swap(c->U, c->A);  // the cases' own view of which buffer holds U
}

void Batch::step(int nsteps) {
    if (!initialized) throw StateError("batch step before init");
    int done = 0;
    const int cap = cases[0]->histCap;
    while (done < nsteps) {
        const int chunk = std::min(nsteps - done, cap);  // the cases' norm rings hold cap steps
        for (int q = 0; q < chunk; ++q) {
            if (useGraph) {
                if (!graphExec[parity]) {
                    hipGraph_t g;
                    hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
                    stepOnce();
                    hipCheck(hipStreamEndCapture(stream, &g), "end capture");
                    hipCheck(hipGraphInstantiate(&graphExec[parity], g, nullptr, nullptr, 0), "instantiate");
                    (void)hipGraphDestroy(g);
                }
                hipCheck(hipGraphLaunch(graphExec[parity], stream), "graph launch");
            } else {
                stepOnce();
            }
            advanceParity();
        }
        hipCheck(hipGetLastError(), "batch kernels");
        hipCheck(hipStreamSynchronize(stream), "batch sync");
        for (auto& c : cases) c->readScalars();
        done += chunk;
    }
}

double Batch::autotune(int reps) {
    if (!initialized) throw StateError("autotune before init");
    int nh = 0;
    const int* tys = fusedTileHeights(&nh);
    int nk = 0;
    const int* kcs = fusedChunkPlanes(&nk);
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "event");
    hipCheck(hipEventCreate(&e1), "event");
    const int n = (int)cases.size();
    FusedTile best = tile;
    double bestMs = 1e30;
    tuneTimes.clear();
    for (int a = 0; a < nh; ++a)
        for (int b = 0; b < nk; ++b) {
            tile.ty = tys[a];
            tile.kc = kcs[b];
            buildStageTables();
            // stages 1-3 write S1, S2 and the accumulator's interior, which the next step rewrites in the same
            // order before reading them (U, the accumulator's other nodes and the carry are untouched)
            auto run = [&](int count) {
                for (int r = 0; r < count; ++r)
                    for (int st = 1; st <= 3; ++st)
                        launchStageBatched(stream, st, tile.ty, false, kahan, devArgs[parity][st - 1], devStageStart, n,
                                           totalTiles, persistent ? devCtr : nullptr);
            };
            run(1);
            hipCheck(hipEventRecord(e0, stream), "record");
            run(reps);
            hipCheck(hipEventRecord(e1, stream), "record");
            hipCheck(hipEventSynchronize(e1), "sync");
            float ms = 0;
            hipCheck(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            const double per = ms / reps;
            tuneTimes.push_back(per);
            if (per < bestMs) {
                bestMs = per;
                best = tile;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tile = best;
    buildStageTables();
    return bestMs;
}

void Batch::setOption(const std::string& name, int64_t value) {
    dropGraphs();
    if (name == "graph") {
        useGraph = value != 0;
    } else if (name == "tile_ty") {
        int nh = 0;
        const int* v = fusedTileHeights(&nh);
        bool ok = false;
        for (int q = 0; q < nh; ++q) ok = ok || v[q] == value;
        if (!ok) throw std::invalid_argument("tile_ty must be one of 4, 6, 8, 14");
        tile.ty = (int)value;
        if (initialized) buildStageTables();
    } else if (name == "tile_kc") {
        if (value < kMinTileKc || value > 4096) throw std::invalid_argument("tile_kc out of range (8..4096)");
        tile.kc = (int)value;
        if (initialized) buildStageTables();
    } else if (name == "persistent") {
        persistent = value != 0;
    } else if (name == "autotune") {
        autotune(value > 0 ? (int)value : 3);
    } else {
        throw std::invalid_argument("unknown batch option: " + name);
    }
}

}  // namespace nsfsi
