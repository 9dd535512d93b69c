/*
 * stage.h — launch wrappers of the RK4 stage kernels and of the end-of-step reductions.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ns_internal.h"

namespace nsfsi {

// Device-resident step scalars: read by every stage kernel, written only by the finalize kernel, so a
// sequence of steps runs without host synchronisation.
struct StepScalars {
    double t;
    double dt;
    uint64_t level;
    uint64_t histCount;
    uint64_t forceCount;  // entries written to the body-force ring
    int32_t nonFinite;
    int32_t filterNow;  // periodic density filter active in the current step
};

// Schedule of the periodic density filter (Mesh::checkFilterCondition, src/Mesh.cpp:85-103).
struct FilterSchedule {
    int n_intervals, n_change;
    uint32_t intervals[16];
    double change_times[16];
};

// ---- unfused stage (stage_plain.hip)
// Primitive and transport fields (7: u, v, w, p', T', mu, kappa) from a state. mode 0: every node, pre-filter
// density and energy at filter nodes; mode 1 (stored primitives, the reference's updatePrimitiveVariables):
// FluidInterior nodes only; mode 2: every node with the boundary conditions' conductivity mu / ((gamma-1) Pr).
template <class T>
void launchDerive(hipStream_t s, const Layout& L, const Params& P, const T* S, T* prim, const uint8_t* cls, const Window& W,
                  const T* rhoPre, const T* rhoEPre, int mode = 0);
// literal (FP64 only): the reference's own combine (slope sums in A, final U + dt/6 * sum), see stage_plain.hip.
template <class T>
void launchRhsPlain(hipStream_t s, int stage, const Layout& L, const Params& P, const T* Sin, const T* prim, const T* U,
                    T* Sout, T* A, const uint8_t* cls, const double* dtdev, const T* filteredRho,
                    const int32_t* filterFlag, int8_t* comp, bool literal = false);

// Periodic density filter (Mesh::applyFilter_ifAppropriate, src/Mesh.cpp:108-140), in three parts:
// the flag for this step from (t, timeLevel); the filtered interior density of U into `out`; after the
// stage-1 kernel, U's interior density := out and U's untouched solid density := A's.
void launchFilterFlag(hipStream_t s, const FilterSchedule& fs, StepScalars* st);
template <class T>
void launchRhoFilter(hipStream_t s, const Layout& L, const T* U, T* out, const uint8_t* cls, const int32_t* flag);
template <class T>
void launchRhoCommit(hipStream_t s, const Layout& L, T* U, const T* filtered, const T* A, const uint8_t* cls,
                     const int32_t* flag);

// ---- fused stage (stage_fused.hip): primitives + RHS + combine in one pass
// Tile shape of the fused stage: ty output rows per block (one of fusedTileHeights), kc z-planes per
// block (the z-chunk each block streams through).
struct FusedTile {
    int ty = 8;
    int kc = 64;
    // FP32: global accesses at per-field base pointers even when the state fits one buffer resource
    // (the form every FP32 state beyond 4 GiB uses; same arithmetic, so bit-identical results)
    bool globalAccess = false;
};
constexpr int kMinTileKc = 8;  // smallest z-chunk accepted (sizes the scan partial buffers)
// Stage 4 with partials != null also writes one partial per block of the end-of-step scans over
// interior, non-filter nodes (norms of change when norms != 0, dt maxima, non-finite count). Returns the
// number of blocks (= partial slots written).
// filteredRho / filterFlag (stage 1): when *filterFlag, interior nodes combine from filteredRho instead of
// the stage input's density, and non-interior nodes copy their density from the stage input.
// comp != null (FP32 only; both stage kernels): compensated combine. The accumulator A holds the step's
// increment at interior nodes (stages 1-3) and stage 4 forms U + increment with the carried rounding
// error; comp is 5 fields of 8-bit carries (carryEncode) with the layout's element offsets, zero before
// the first step.
// outlet != null: that outlet face (outletFusable) is applied to the stage output (Sout, A in stage 4) by
// the kernel itself, with the values the edge-face pass would give it when run first after the kernel.
template <class T>
int launchStageFused(hipStream_t s, int stage, const Layout& L, const Params& P, const Window& W, const T* Sin,
                     const T* U, T* Sout, T* A, const uint8_t* cls, const T* rhoPre, const T* rhoEPre,
                     const double* dtdev, FusedTile tile, double* partials, int norms, const T* filteredRho,
                     const int32_t* filterFlag, int8_t* comp, const EdgeFace* outlet = nullptr);
// Whether the fused stage kernel can apply edge face f itself: an outlet at x = NI-1 whose row's last
// x-segment holds at least three nodes (the node and the two it extrapolates from share a wavefront).
// The caller also keeps f out of the edge-face pass and checks that f is independent of every face
// applied before it.
inline bool outletFusable(const Layout& L, const EdgeFace& f) {
    const int tail = L.NI - Layout::kSegNodes * ((L.NI + Layout::kSegNodes - 1) / Layout::kSegNodes - 1);
    return f.type == NS_BC_OUTLET && f.axis == 0 && f.side == 1 && f.i0 == L.NI - 1 && f.i1 == L.NI - 1 && tail >= 3;
}
const int* fusedTileHeights(int* n);
const int* fusedChunkPlanes(int* n);  // z-chunk lengths the autotuners try
// Batched FP32 fused stage: fusedArgsBytes() bytes of arguments per case (buildFusedArgs fills them and
// returns the case's tile count); launchStageBatched runs stage `stage` of every case in one launch over
// a device table of those blocks, start = device prefix sums of the tile counts (ncases + 1 entries).
size_t fusedArgsBytes();
int buildFusedArgs(void* dst, int stage, const Layout& L, const Params& P, const Window& W, const float* Sin,
                   const float* U, float* Sout, float* A, const uint8_t* cls, const float* rhoPre, const float* rhoEPre,
                   const double* dtdev, FusedTile tile, double* partials, int norms, int8_t* comp,
                   const EdgeFace* outlet = nullptr, const float* filteredRho = nullptr,
                   const int32_t* filterFlag = nullptr);
// ctr != null: the persistent form (one block per resident slot pulling tiles through ctr, 9 unsigned
// counters that must be zero before the first launch and that each launch leaves zero).
void launchStageBatched(hipStream_t s, int stage, int ty, bool scan, bool kahan, const void* tab, const int* start,
                        int ncases, int total, unsigned* ctr = nullptr);
// Force on the body from the state S after a step's boundary application (FP64; the discrete momentum
// balance over the faces between fluid interior and ghost nodes): appends (timeLevel, Fx, Fy, Fz) to
// `ring` (cap entries of 4) and advances st->forceCount. partials: bodyForceBlocks(nGhost) x 3 doubles.
int bodyForceBlocks(int64_t nGhost);
template <class T>
void launchBodyForce(hipStream_t s, const Layout& L, const Params& P, const Window& W, const T* S, const uint8_t* cls,
                     const T* rhoPre, const T* rhoEPre, const int64_t* ghosts, int64_t nGhost, double* partials,
                     StepScalars* st, double* ring, int cap);
int fusedMaxBlocks(const Layout& L);  // largest block count over all accepted tile shapes

// End-of-step scans over what the stage-4 pass leaves out, after the boundary conditions: the owned
// nodes of the six faces, the ghost nodes and the filter nodes. Returns the partial slots written.
template <class T>
int launchBoundaryReduce(hipStream_t s, const Layout& L, const Params& P, const Window& W, const EdgeFace faces[6],
                         const int64_t* ghostNode, int64_t nGhost, const int64_t* filterList, int64_t nFilter,
                         const T* U, const T* A, const uint8_t* cls, const T* rhoPre, const T* rhoEPre, int norms,
                         double* partials);

// ---- end-of-step reductions (reduce.hip)
template <class T>
// prim != null: the dt maxima read the stored primitive fields at every node (the stored-primitive form).
void launchStepReduce(hipStream_t s, const Layout& L, const Params& P, const T* U, const T* A, const uint8_t* cls,
                      const Window& W, const T* rhoPre, const T* rhoEPre, int withNorms, double* partials, int nBlocks,
                      const T* prim = nullptr);
template <class T>
void launchInactiveReduce(hipStream_t s, const Layout& L, const Params& P, const T* A, const int64_t* nodes,
                          const double* icPrim, int64_t n, double* partial);
// dtSched: optional device array of prescribed step sizes indexed by time level (replaces the limit).
void launchFinalize(hipStream_t s, const Params& P, const double* partials, int nPartials, int mode, int convHistory,
                    int stopEndTime, double tEnd, double cfl, double sigma, StepScalars* st, double* hist, int histCap,
                    const double* dtSched = nullptr);
int reducePartialSlots();
int inactiveReduceBlocks(int64_t n);  // partial slots the inactive-node reduction writes (<= 64)

}  // namespace nsfsi
