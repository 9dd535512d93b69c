/*
 * adjoint.h — launchers of the discrete-adjoint kernels (adjoint.hip), FP64 (the forward run they reverse
 * may be FP32: its states are widened first).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <utility>
#include <vector>

#include "kernels.h"
#include "ns_internal.h"

namespace nsfsi {

struct AdjConst {
    double i2d[3];  // 1 / (2 h)
    double c43[3];  // 2 / (3 h^2)
    double c1[3];   // 1 / (2 h^2)
    double gamma, gm1, H0, ScP1, Re, pf;
};
AdjConst makeAdjConst(const Params& P);

// Transposed right-hand side at one stage state.
struct AdjArgs {
    Layout L;
    Window W;
    AdjConst c;
    const double* S;       // stage state after its boundary application (5 fields)
    const uint8_t* cls;
    const double* Ppre;    // pre-filter density / energy of the filter nodes (window-dense)
    const double* PEpre;
    const double* lam;     // adjoint of the stage right-hand side (5 fields, zero at non-interior nodes)
    double objCoef;        // d J / d Fx at this state (0: no objective term)
    double vol;            // dx dy dz
    double* passA;         // scratch, 19 fields
    double* Sbar;          // += adjoint of S (5 fields)
    double* Pbar;          // += adjoint of the pre-filter tables (window-dense)
    double* PEbar;
};
int adjRhsPartials(const Layout& L);  // dJ/dRe partial slots written by launchAdjRhs
void launchAdjRhs(hipStream_t s, const AdjArgs& a, double* rePartial);

// Fixed-order sums of the transposed boundary pieces (each sum in list order, so results do not depend
// on thread scheduling): target t, at layout offset tgt[t], adds the entries src[start[t] .. start[t+1]-1].
struct GatherPlan {
    int64_t n = 0;
    int64_t* tgt = nullptr;
    int64_t* start = nullptr;
    int32_t* src = nullptr;
};
// From (target offset, entry) pairs, in (target, entry) order; device arrays from alloc.
GatherPlan buildGatherPlan(std::vector<std::pair<int64_t, int32_t>> pairs, const std::function<void*(size_t)>& alloc);

// Transposed boundary pieces. Sb is updated in place (it holds the adjoint of the state after the
// piece on entry and of the state before it on return).
// Ghosts: contrib holds 40 doubles per ghost; donors lists, per fluid donor node, the entries 8 g + c of
// the ghosts g whose interpolation box has it as corner c.
void launchAdjGhost(hipStream_t s, const Layout& L, const Params& P, const GhostTables& G, int64_t n, double* Sb,
                    double* contrib, const GatherPlan& donors);
// Immersed filter: tmp holds 4 doubles per filter node; targets lists, per filter node or neighbour of
// one, the entries 2 q + 1 (filter node q itself) and 2 q (a filter node q it neighbours).
void launchAdjFilter(hipStream_t s, const Layout& L, const Window& W, const int64_t* fn, int64_t n, double* Sb,
                     double* tmp, const double* Pb, const double* PEb, const GatherPlan& targets);
// wbuf: one double per node of an inlet face (unused for the other face types).
void launchAdjEdge(hipStream_t s, const Layout& L, const Params& P, const EdgeFace& f, double* Sb, const double* S,
                   double ts, double* qb, double* wbuf);
// Interior nodes: Ub += B, lam[r] += coef[r] * B (null entries skipped); other nodes: prev += B (if set).
void launchAdjRoute(hipStream_t s, const Layout& L, const uint8_t* cls, const double* B, double* Ub, double* const lam[4],
                    const double coef[4], double* prev);
void launchAxpyWindow(hipStream_t s, double* y, const double* x, size_t n);
void launchAdjTablesToState(hipStream_t s, const Layout& L, const Window& W, const int64_t* fn, int64_t n,
                            const double* Pb, const double* PEb, double* Sb);
// y[q] = x[q] for q < n: the stage states and tables of an FP32 forward step, widened for the FP64 reverse sweep.
void launchWiden(hipStream_t s, const float* x, double* y, size_t n);

}  // namespace nsfsi
