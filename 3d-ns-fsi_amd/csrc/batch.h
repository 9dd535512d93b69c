/*
 * batch.h — device-side description of one case of a batch (many independent cases advanced together,
 * one launch per kernel type for all of them), and the index helpers of the batched kernels.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "ns_internal.h"
#include "stage.h"

namespace nsfsi {

// Everything the per-step kernels of one case read: built on the host from an initialised case, kept in
// device memory for the life of the batch. Buffer roles: 0 = U (the state at the start of the step),
// 1 = A (accumulator / next state), 2 = S1, 3 = S2; with parity 1 the U and A buffers are exchanged
// (the step swaps them, as the single-case step swaps its pointers).
template <class T>
struct CaseView {
    Layout L;
    Params P;
    Window W;
    EdgeFace faces[6];      // application order
    int64_t faceStart[7];   // prefix sums of the owned nodes of faces[0..5]
    const uint8_t* cls;
    T* buf[4];
    T* rhoPre;
    T* rhoEPre;
    T* ftmp;
    const int64_t* filterList;
    int64_t nFilter;
    GhostTables G;
    int64_t nGhost;
    const int64_t* inactiveList;
    const double* inactiveIC;
    int64_t nInactive;
    StepScalars* scal;
    double* partials;       // stage-4 block partials, then the boundary and inactive-node partials
    double* hist;
    int histCap;
    int convHistory, stopEndTime;
    double tEnd, cfl, sigma;
    int nStage4, nBoundaryBlocks, nInactiveBlocks;
    int norms;
    FilterSchedule fs;      // periodic density filter (src/Mesh.cpp:85-140), decided on the device
};

template <class T>
__device__ __forceinline__ T* roleBuffer(const CaseView<T>& v, int role, int parity) {
    return role <= 1 ? v.buf[role ^ parity] : v.buf[role];
}

// Last case c with start[c] <= q (start has ncases + 1 ascending entries, start[0] = 0).
template <class I>
__device__ __forceinline__ int findCase(const I* __restrict__ start, int ncases, I q) {
    int lo = 0, hi = ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- batched launchers (one launch for all cases; *Start arrays are device prefix sums over cases)
// Periodic density filter of the cases ids[0..n) (those with a filter interval): the schedule test, the
// filtered density into each case's S2 (nodes over the cases by the prefix sums start), its commit to U
// after stage 1 (stage_plain.hip).
void launchFilterFlagBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, int n);
void launchRhoFilterBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, const int64_t* start, int n,
                            int64_t total, int parity);
void launchRhoCommitBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, const int64_t* start, int n,
                            int64_t total, int parity);
// Faces of every case launched together (faces that are independent in every case): member m is face
// face[m] of each case, its nodes over the cases given by the device prefix sums start[m], and it takes
// the thread range [off[m], off[m+1]).
struct EdgeFaceListB {
    int face[6];
    const int64_t* start[6];
    int64_t off[7];
    int n;
};
template <class T>
void launchEdgeFacesBatched(hipStream_t s, const CaseView<T>* cv, const EdgeFaceListB& fl, int ncases, int role, int parity,
                            double frac);
template <class T>
void launchFilterComputeBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                                int role, int parity);
template <class T>
void launchFilterWriteBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                              int role, int parity);
template <class T>
void launchGhostApplyBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                             int role, int parity);
// End-of-step scans of the cases' boundary nodes and untouched solid nodes (after a stage-4 launch with
// in-kernel scans), then one finalising block per case: dt / t / timeLevel of every case.
template <class T>
void launchBoundaryReduceBatched(hipStream_t s, const CaseView<T>* cv, const int* blockStart, int totalBlocks,
                                 int ncases, int parity);
template <class T>
void launchInactiveReduceBatched(hipStream_t s, const CaseView<T>* cv, const int* blockStart, int totalBlocks,
                                 int ncases, int parity);
template <class T>
void launchFinalizeBatched(hipStream_t s, const CaseView<T>* cv, int ncases);

}  // namespace nsfsi
