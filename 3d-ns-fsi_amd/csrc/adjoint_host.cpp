/*
 * adjoint_host.cpp — discrete adjoint of the drag objective over a run of time steps.
 *
 * J = 1 / (n1 - n0 + 1) * sum_{L = n0..n1} C_D(U^L),  C_D = Fx / (0.5 M_0^2 A_ref), Fx the force on the
 * body of ns_force_history. The step sizes are those of the forward run, held fixed (the adjoint is that
 * of the map at a frozen step sequence; finite differences must use the same sequence, which the run
 * returns). Gradients: with respect to the initial conserved state, the inlet profile factors and Re.
 *
 * Memory: checkpoints of the full state every `ckpt` steps; per segment a recomputation that records
 * the four stage states of every step (tape), then the reverse steps of the segment. An FP32 case runs the
 * forward and the recomputation in FP32 (its step, compensated combine included; forces and step sizes
 * reduced in FP64) and the reverse sweep in FP64 on the widened stage states: the adjoint of the FP64 map
 * linearised along the FP32 trajectory. Reverse step m
 * (forward U^{m-1} -> U^m) consumes the adjoint of U^m and of its pre-filter tables and produces that of
 * U^{m-1}; stage 1 of step m copies its non-interior nodes from U^{m-2} (the reference's stale values),
 * so it also produces an adjoint of U^{m-2} that reverse step m-1 adds to its result.
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "case.h"

namespace nsfsi {

double Case::referenceArea() const {
    const BodyDesc& B = bodies.b[0];
    if (bodies.n != 1) throw Unsupported("the drag objective covers one immersed body");
    if (B.type == NS_BODY_SPHERE) return M_PI * B.r * B.r;
    if (B.type == NS_BODY_CYLINDER) {
        const int n = B.axis == 0 ? L.NI : (B.axis == 1 ? L.NJ : L.NK);
        const double h = B.axis == 0 ? P.hx : (B.axis == 1 ? P.hy : P.hz);
        return 2 * B.r * (n - 2) * h;
    }
    throw std::invalid_argument("no body");
}

void Case::adjointDrag(int nsteps, int n0, int n1, int ckpt, double* J, double* gradU0, double* gradInflow,
                       double* gradRe, double* dtsOut) {
    if (!initialized) throw StateError("adjoint before init");
    if (level != 0) throw StateError("the adjoint starts from the initial state (time level 0)");
    if (!filterDisabled()) throw Unsupported("the adjoint does not cover the periodic density filter");
    if (nGhost == 0) throw Unsupported("the drag objective needs an immersed body");
    if (bodies.n != 1) throw Unsupported("the drag objective covers one immersed body");
    if (nFixed > 0) throw Unsupported("the adjoint does not cover bodies whose solid nodes reach a symmetry face");
    if (nsteps < 1 || n0 < 0 || n1 < n0 || n1 > nsteps || ckpt < 1)
        throw std::invalid_argument("need 1 <= nsteps, 0 <= n0 <= n1 <= nsteps, ckpt >= 1");
    checkSchedule(nsteps);  // a prescribed schedule must cover the forward run
    dropGraphs();
    const bool graphWas = useGraph, forceWas = bodyForce;
    useGraph = false;
    const size_t F5 = 5 * L.fstride, WN = std::max<size_t>(1, W.size());
    const size_t fb = F5 * sizeof(double), wb = WN * sizeof(double);
    // the state buffers and tables in the forward's precision, and the compensated combine's carry
    const size_t es = fp32 ? sizeof(float) : sizeof(double), fbT = F5 * es, wbT = WN * es;
    const size_t cb = comp ? F5 * sizeof(int8_t) : 0;
    std::vector<void*> mine;
    auto dalloc = [&](size_t bytes) {
        void* p = nullptr;
        hipCheck(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc (adjoint)");
        mine.push_back(p);
        return (double*)p;
    };
    auto zero = [&](void* p, size_t bytes) { hipCheck(hipMemsetAsync(p, 0, bytes, stream), "memset"); };
    auto copy = [&](void* d, const void* s, size_t bytes) {
        hipCheck(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, stream), "copy");
    };
    try {
        // ---- forward pass with checkpoints; records the step sizes and the forces
        const int nck = nsteps / ckpt + 1;
        // U64 / P64 / PE64: the state and tables in FP64 for the reverse sweep (for an FP64 case, U / P / PE)
        struct Ck { void *U, *A, *P, *PE; int8_t* C; double *U64, *P64, *PE64; StepScalars s; };
        std::vector<Ck> cks(nck + 1);
        std::vector<double> tAt(nsteps + 1), dtAt(nsteps + 1);
        auto save = [&](Ck& c) {
            if (!c.U) {
                c.U = dalloc(fbT);
                c.A = dalloc(fbT);
                c.P = dalloc(wbT);
                c.PE = dalloc(wbT);
                if (comp) c.C = (int8_t*)dalloc(cb);
                if (fp32) {
                    c.U64 = dalloc(fb);
                    c.P64 = dalloc(wb);
                    c.PE64 = dalloc(wb);
                } else {
                    c.U64 = (double*)c.U;
                    c.P64 = (double*)c.P;
                    c.PE64 = (double*)c.PE;
                }
            }
            copy(c.U, U, fbT);
            copy(c.A, A, fbT);
            copy(c.P, rhoPre, wbT);
            copy(c.PE, rhoEPre, wbT);
            if (comp) copy(c.C, comp, cb);
            if (fp32) {
                launchWiden(stream, (const float*)U, c.U64, F5);
                launchWiden(stream, (const float*)rhoPre, c.P64, WN);
                launchWiden(stream, (const float*)rhoEPre, c.PE64, WN);
            }
            hipCheck(hipMemcpyAsync(&c.s, scal, sizeof(StepScalars), hipMemcpyDeviceToHost, stream), "copy");
            hipCheck(hipStreamSynchronize(stream), "sync");
        };
        auto restore = [&](const Ck& c) {
            copy(U, c.U, fbT);
            copy(A, c.A, fbT);
            copy(rhoPre, c.P, wbT);
            copy(rhoEPre, c.PE, wbT);
            if (comp) copy(comp, c.C, cb);
            hipCheck(hipMemcpyAsync(scal, &c.s, sizeof(StepScalars), hipMemcpyHostToDevice, stream), "copy");
            hipCheck(hipStreamSynchronize(stream), "sync");
            readScalars();
            primsStored = false;  // stored primitive fields (unfused FP64 form) are derived again from the state
        };
        bodyForce = true;
        forceHost.clear();
        std::vector<double> fx(nsteps + 1, 0.0);
        if (n0 == 0) {  // force of the initial state
            if (fp32)
                launchBodyForce<float>(stream, L, P, W, (const float*)U, cls, (const float*)rhoPre,
                                       (const float*)rhoEPre, ghostList, nGhost, partials, scal, forceRing, histCap);
            else
                launchBodyForce<double>(stream, L, P, W, (const double*)U, cls, (const double*)rhoPre,
                                        (const double*)rhoEPre, ghostList, nGhost, partials, scal, forceRing, histCap);
            readScalars();
            fx[0] = forceHost[forceHost.size() - 3];
        }
        for (int lv = 0; lv < nsteps; ++lv) {
            if (lv % ckpt == 0) save(cks[lv / ckpt]);
            tAt[lv] = t;
            dtAt[lv] = dt;
            if (fp32) stepOnce<float>();  // swaps U and A itself
            else stepOnce<double>();
            hipCheck(hipGetLastError(), "forward step");
            readScalars();
            fx[lv + 1] = forceHost[forceHost.size() - 3];
        }
        tAt[nsteps] = t;
        dtAt[nsteps] = dt;
        Ck last{};
        save(last);
        const double cJ = 1.0 / ((n1 - n0 + 1) * 0.5 * P.M_0 * P.M_0 * referenceArea());
        double Jv = 0;
        for (int lv = n0; lv <= n1; ++lv) Jv += cJ * fx[lv];
        if (J) *J = Jv;
        if (dtsOut) std::memcpy(dtsOut, dtAt.data(), (size_t)(nsteps + 1) * sizeof(double));
        bodyForce = forceWas;
        // recomputation reproduces the forward steps exactly at these step sizes
        const double* schedWas = dtSchedule;
        const int64_t schedLenWas = dtScheduleLen;
        std::vector<double> schedWasHost;
        if (schedWas) {
            schedWasHost.resize(schedLenWas);
            hipCheck(hipMemcpy(schedWasHost.data(), schedWas, schedLenWas * sizeof(double), hipMemcpyDeviceToHost), "copy");
        }

        // ---- adjoint buffers
        double* Ub = dalloc(fb);     // adjoint of the current state U^m
        double* Un = dalloc(fb);     // adjoint of U^{m-1} being formed
        double* Bb = dalloc(fb);     // working adjoint through one boundary application
        double* Sb[4] = {nullptr, dalloc(fb), dalloc(fb), dalloc(fb)};
        double* lam[4] = {dalloc(fb), dalloc(fb), dalloc(fb), dalloc(fb)};
        double* carryIn = dalloc(fb);
        double* carryOut = dalloc(fb);
        double* passA = dalloc(19 * L.fstride * sizeof(double));
        double* Pb = dalloc(wb);
        double* PEb = dalloc(wb);
        double* Ps[4], *PEs[4];
        for (int q = 0; q < 4; ++q) { Ps[q] = dalloc(wb); PEs[q] = dalloc(wb); }
        double* ftmpA = dalloc(4 * std::max<int64_t>(1, nFilter) * sizeof(double));
        // fixed-order sums of the ghost and immersed-filter transposes: the donors of every ghost's
        // interpolation box, and the filter nodes with their six neighbours
        auto dallocV = [&](size_t bytes) { return (void*)dalloc(bytes); };
        double* gcontrib = dalloc((size_t)std::max<int64_t>(1, nGhost) * 40 * sizeof(double));
        GatherPlan donors, filterTargets;
        if (8 * nGhost > INT32_MAX || 2 * nFilter > INT32_MAX)
            throw std::length_error("too many ghost or filter nodes for the adjoint's 32-bit gather lists");
        if (nGhost > 0) {
            std::vector<int64_t> corner((size_t)nGhost * 8);
            std::vector<uint8_t> gmask((size_t)nGhost);
            hipCheck(hipMemcpy(corner.data(), G.corner, corner.size() * sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
            hipCheck(hipMemcpy(gmask.data(), G.mask, gmask.size(), hipMemcpyDeviceToHost), "copy");
            std::vector<std::pair<int64_t, int32_t>> pairs;
            for (int64_t g = 0; g < nGhost; ++g)
                for (int c = 0; c < 8; ++c)
                    if (!((gmask[g] >> c) & 1)) pairs.emplace_back(corner[8 * g + c], (int32_t)(8 * g + c));
            donors = buildGatherPlan(std::move(pairs), dallocV);
        }
        if (nFilter > 0) {
            std::vector<int64_t> fl((size_t)nFilter);
            hipCheck(hipMemcpy(fl.data(), filterList, fl.size() * sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
            std::vector<std::pair<int64_t, int32_t>> pairs;
            const long sy = L.pitch, sz = (long)L.plane;
            for (int64_t q = 0; q < nFilter; ++q) {
                const int64_t m = fl[q];
                const int i = (int)(m % L.NI), j = (int)((m / L.NI) % L.NJ), k = (int)(m / ((int64_t)L.NI * L.NJ));
                const int64_t o = (int64_t)L.off(i, j, k);
                pairs.emplace_back(o, (int32_t)(2 * q + 1));
                const long nbs[6] = {Layout::xstep(i, 1), Layout::xstep(i, -1), sy, -sy, sz, -sz};
                for (long d : nbs) pairs.emplace_back(o + d, (int32_t)(2 * q));
            }
            filterTargets = buildGatherPlan(std::move(pairs), dallocV);
        }
        const int nre = adjRhsPartials(L);
        double* rePart = dalloc((size_t)nre * sizeof(double));
        double reSum = 0;
        std::vector<double> reHost(nre);
        int64_t planeN = 0;
        for (const EdgeFace& f : faces)
            if (f.type == NS_BC_INLET)
                planeN = f.axis == 0 ? (int64_t)L.NJ * L.NK : (f.axis == 1 ? (int64_t)L.NI * L.NK : (int64_t)L.NI * L.NJ);
        double* qb = planeN ? dalloc((size_t)planeN * sizeof(double)) : nullptr;
        int64_t inletN = 0;  // nodes of the largest inlet face (its in-plane filter's adjoint, per node)
        for (const EdgeFace& f : faces)
            if (f.type == NS_BC_INLET)
                inletN = std::max<int64_t>(inletN, (int64_t)std::max(0, f.i1 - f.i0 + 1) * std::max(0, f.j1 - f.j0 + 1) *
                                                       std::max(0, f.k1 - f.k0 + 1));
        double* wbuf = inletN ? dalloc((size_t)inletN * sizeof(double)) : nullptr;
        if (qb) zero(qb, (size_t)planeN * sizeof(double));
        const AdjConst ac = makeAdjConst(P);
        const double vol = P.hx * P.hy * P.hz;

        auto rhsT = [&](const double* S, const double* Pp, const double* PEp, const double* lm, double objCoef,
                        double* Sbar, double* Pbar, double* PEbar) {
            AdjArgs a;
            a.L = L;
            a.W = W;
            a.c = ac;
            a.S = S;
            a.cls = cls;
            a.Ppre = Pp;
            a.PEpre = PEp;
            a.lam = lm;
            a.objCoef = objCoef;
            a.vol = vol;
            a.passA = passA;
            a.Sbar = Sbar;
            a.Pbar = Pbar;
            a.PEbar = PEbar;
            launchAdjRhs(stream, a, rePart);
            hipCheck(hipMemcpyAsync(reHost.data(), rePart, (size_t)nre * sizeof(double), hipMemcpyDeviceToHost, stream),
                     "copy");
            hipCheck(hipStreamSynchronize(stream), "sync");
            for (double v : reHost) reSum += v;
        };
        // transposed boundary application (edges in application order, filter, ghosts) at time ts, for
        // the state S it produced; Pbar/PEbar: adjoint of the pre-filter tables it recorded
        auto bcT = [&](double* Bv, const double* S, const double* Pbar, const double* PEbar, double ts) {
            launchAdjGhost(stream, L, P, G, nGhost, Bv, gcontrib, donors);
            launchAdjFilter(stream, L, W, filterList, nFilter, Bv, ftmpA, Pbar, PEbar, filterTargets);
            for (int q = 5; q >= 0; --q) launchAdjEdge(stream, L, P, faces[q], Bv, S, ts, qb, wbuf);
        };
        auto objAt = [&](int lv) { return lv >= n0 && lv <= n1 ? cJ : 0.0; };

        // seed: objective at the final level
        zero(Ub, fb);
        zero(Pb, wb);
        zero(PEb, wb);
        zero(carryIn, fb);
        if (objAt(nsteps) != 0.0) {
            zero(lam[0], fb);
            rhsT(last.U64, last.P64, last.PE64, lam[0], objAt(nsteps), Ub, Pb, PEb);
        }

        // tape of one segment
        const int segMax = std::min(ckpt, nsteps);
        std::vector<StageTape> tape(segMax);
        for (auto& tp : tape)
            for (int q = 0; q < 4; ++q) { tp.S[q] = dalloc(fb); tp.P[q] = dalloc(wb); tp.PE[q] = dalloc(wb); }

        for (int seg = (nsteps - 1) / ckpt; seg >= 0; --seg) {
            const int a0 = seg * ckpt, b0 = std::min(nsteps, a0 + ckpt);
            restore(cks[seg]);
            setDtSchedule(dtAt.data(), nsteps + 1);
            for (int m = a0 + 1; m <= b0; ++m) stepTaped(&tape[m - a0 - 1]);
            for (int m = b0; m >= a0 + 1; --m) {
                const StageTape& tp = tape[m - a0 - 1];
                const double* endS = m == b0 ? (seg == (nsteps - 1) / ckpt ? last.U64 : cks[seg + 1].U64)
                                             : tape[m - a0].S[0];
                const double tt = tAt[m - 1], h = dtAt[m - 1];
                // final boundary application of step m (produced U^m and its tables)
                copy(Bb, Ub, fb);
                bcT(Bb, endS, Pb, PEb, tt + h);
                for (int q = 0; q < 4; ++q) zero(lam[q], fb);
                zero(Un, fb);
                for (int q = 1; q < 4; ++q) zero(Sb[q], fb);
                zero(carryOut, fb);
                {
                    double* ls[4] = {lam[0], lam[1], lam[2], lam[3]};
                    const double cf[4] = {h / 6, h / 3, h / 3, h / 6};
                    launchAdjRoute(stream, L, cls, Bb, Un, ls, cf, Sb[3]);
                }
                // stages 4, 3, 2: rhs at S_{s-1}, then its boundary application
                const double frac[4] = {0.5, 0.5, 1.0, 1.0};
                const double cst[4] = {0.5 * h, 0.5 * h, h, h / 6};
                for (int s = 4; s >= 2; --s) {
                    zero(Ps[s - 1], wb);
                    zero(PEs[s - 1], wb);
                    rhsT(tp.S[s - 1], tp.P[s - 1], tp.PE[s - 1], lam[s - 1], 0.0, Sb[s - 1], Ps[s - 1], PEs[s - 1]);
                    copy(Bb, Sb[s - 1], fb);
                    bcT(Bb, tp.S[s - 1], Ps[s - 1], PEs[s - 1], tt + frac[s - 2] * h);
                    double* ls[4] = {lam[s - 2], nullptr, nullptr, nullptr};
                    const double cf[4] = {cst[s - 2], 0, 0, 0};
                    launchAdjRoute(stream, L, cls, Bb, Un, ls, cf, s - 2 >= 1 ? Sb[s - 2] : carryOut);
                }
                // stage 1: rhs at U^{m-1} (plus the objective at level m-1)
                zero(Pb, wb);
                zero(PEb, wb);
                rhsT(tp.S[0], tp.P[0], tp.PE[0], lam[0], objAt(m - 1), Un, Pb, PEb);
                // the stage-1 copies of step m+1 came from U^{m-1}
                {
                    double* ls[4] = {nullptr, nullptr, nullptr, nullptr};
                    const double cf[4] = {0, 0, 0, 0};
                    launchAdjRoute(stream, L, cls, carryIn, Un, ls, cf, Un);
                }
                std::swap(Ub, Un);
                std::swap(carryIn, carryOut);
            }
        }
        // U^0 is the initial field as given (no boundary application before the first step); its
        // pre-filter tables are copies of it
        launchAdjTablesToState(stream, L, W, filterList, nFilter, Pb, PEb, Ub);
        hipCheck(hipStreamSynchronize(stream), "sync");
        if (gradU0) {
            std::vector<double> h(F5);
            hipCheck(hipMemcpy(h.data(), Ub, fb, hipMemcpyDeviceToHost), "copy");
            const size_t N = L.nodes();
            for (int q = 0; q < 5; ++q)
                for (int k = 0; k < L.NK; ++k)
                    for (int j = 0; j < L.NJ; ++j)
                        for (int i = 0; i < L.NI; ++i)
                            gradU0[q * N + ((size_t)k * L.NJ + j) * L.NI + i] = h[L.off(i, j, k) + q * L.fstride];
        }
        if (gradInflow && qb) hipCheck(hipMemcpy(gradInflow, qb, (size_t)planeN * sizeof(double), hipMemcpyDeviceToHost), "copy");
        if (gradRe) *gradRe = reSum;
        // leave the case at the end of the forward run, with the step-size rule it had
        restore(last);
        if (schedWas) setDtSchedule(schedWasHost.data(), schedLenWas);
        else { dtSchedule = nullptr; dtScheduleLen = 0; }
    } catch (...) {
        for (void* p : mine) (void)hipFree(p);
        useGraph = graphWas;
        bodyForce = forceWas;
        throw;
    }
    for (void* p : mine) (void)hipFree(p);
    useGraph = graphWas;
    bodyForce = forceWas;
}

}  // namespace nsfsi
