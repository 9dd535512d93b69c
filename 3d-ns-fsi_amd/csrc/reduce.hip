/*
 * reduce.hip — the FP64 scans that close a time step: norms of change (Mesh::getNormOfChange,
 * src/Mesh.cpp:171-178), the finite check (Mesh::checkFinity, src/Mesh.h:48-61) and the two spectral
 * maxima of the time-step limits (src/Solver.cpp:91-126), then the dt/t/timeLevel update
 * (src/Solver.cpp:80-89, 164-166) on the device, so a step never waits for the host.
 *
 * Deterministic: fixed grid, per-block partials, one final block that combines them in a fixed order.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch.h"
#include "device_common.h"
#include "stage.h"

namespace nsfsi {

namespace {

constexpr int RB = 256;  // threads per reduction block

// partial slots: 0..3 maxima (x, y, z spectral radius, viscous factor), 4..8 sums of squared change,
// 9 non-finite count
constexpr int NP = 10;

__device__ __forceinline__ double blockSum(double v, double* sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r += sh[q];
    return r;
}
__device__ __forceinline__ double blockMax(double v, double* sh) {
    v = wave_max(v);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, sh[q]);
    return r;
}

// prim != null (stored primitives): every node's u, v, w, p' and mu come from the stored fields, as
// Solver::getInviscidTimeStepLimit / getViscousTimeStepLimit read them (src/Solver.cpp:91-126).
template <class T>
__global__ void k_step_reduce(Layout L, Params P, const T* U, const T* A, const uint8_t* cls, Window W, const T* rhoPre,
                              const T* rhoEPre, int withNorms, double* partials, const T* prim) {
    // one block per group of x-rows: no per-node index division
    __shared__ double sh[8];
    const size_t F = L.fstride;
    const int rows = L.NJ * L.NK;
    Eos<T> e = Eos<T>::make(P);
    const double vm = fmax(4. / 3, P.gamma / P.Pr);
    double mx[4] = {0, 0, 0, 0}, sm[5] = {0, 0, 0, 0, 0}, bad = 0;
    for (int rw = blockIdx.x; rw < rows; rw += gridDim.x) {
        const int j = rw % L.NJ, k = rw / L.NJ;
        const bool rowW = W.ni > 0 && j >= W.j0 && j < W.j0 + W.nj && k >= W.k0 && k < W.k0 + W.nk;
        for (int i = threadIdx.x; i < L.NI; i += blockDim.x) {
            const size_t o = L.off(i, j, k);
            const uint8_t c = cls[o];
            T a[5];
            for (int q = 0; q < 5; ++q) a[q] = A[o + q * F];
            for (int q = 0; q < 5; ++q)
                if (!isfinite(a[q])) bad += 1;
            if (withNorms && (c & CLS_TYPE_MASK) == NS_FLUID_INTERIOR)
                for (int q = 0; q < 5; ++q) {
                    double d = (double)U[o + q * F] - (double)a[q];
                    sm[q] += d * d;
                }
            if (prim) {
                const double cs = sqrt((1 + P.gamma * (double)prim[o + 3 * F]) / (1 + (double)a[0]));
                mx[0] = fmax(mx[0], cs + fabs((double)prim[o]));
                mx[1] = fmax(mx[1], cs + fabs((double)prim[o + F]));
                mx[2] = fmax(mx[2], cs + fabs((double)prim[o + 2 * F]));
                mx[3] = fmax(mx[3], (double)prim[o + 5 * F] / ((double)a[0] + 1) * vm);
                continue;
            }
            if (c & (CLS_DT_SKIP | CLS_INACTIVE_IC)) continue;
            T rho = a[0], rE = a[4];
            if (rowW && (c & CLS_FILTER)) {
                size_t w = W.idx(i, j, k);
                rho = rhoPre[w];
                rE = rhoEPre[w];
            }
            T u, v, ww, p, Tt, mu;
            derivePrims(e, rho, a[1], a[2], a[3], rE, u, v, ww, p, Tt, mu);
            double cs = sqrt((1 + P.gamma * (double)p) / (1 + (double)a[0]));
            mx[0] = fmax(mx[0], cs + fabs((double)u));
            mx[1] = fmax(mx[1], cs + fabs((double)v));
            mx[2] = fmax(mx[2], cs + fabs((double)ww));
            mx[3] = fmax(mx[3], (double)mu / ((double)a[0] + 1) * vm);
        }
    }
    double* out = partials + (size_t)blockIdx.x * NP;
    for (int q = 0; q < 4; ++q) {
        double v = blockMax(mx[q], sh);
        if (threadIdx.x == 0) out[q] = v;
    }
    for (int q = 0; q < 5; ++q) {
        double v = blockSum(sm[q], sh);
        if (threadIdx.x == 0) out[4 + q] = v;
    }
    double b = blockSum(bad, sh);
    if (threadIdx.x == 0) out[9] = b;
}

// SolidInactive nodes keep their initial primitive values (never written by the reference); their
// density is whatever the result buffer holds there.
// Block blk of nblk over the list; writes one partial at out.
template <class T>
__device__ __forceinline__ void inactiveReduceBlock(const Layout& L, const Params& P, const T* A, const int64_t* nodes,
                                                    const double* icPrim, int64_t n, int blk, int nblk, double* out,
                                                    double* sh) {
    const double vm = fmax(4. / 3, P.gamma / P.Pr);
    double mx[4] = {0, 0, 0, 0};
    for (int64_t q = (int64_t)blk * blockDim.x + threadIdx.x; q < n; q += (int64_t)nblk * blockDim.x) {
        int64_t m = nodes[q];
        int i = (int)(m % L.NI);
        int64_t r = m / L.NI;
        int j = (int)(r % L.NJ), k = (int)(r / L.NJ);
        double rho = (double)A[L.off(i, j, k)];
        const double* ic = icPrim + 5 * q;  // u, v, w, p, mu
        double cs = sqrt((1 + P.gamma * ic[3]) / (1 + rho));
        mx[0] = fmax(mx[0], cs + fabs(ic[0]));
        mx[1] = fmax(mx[1], cs + fabs(ic[1]));
        mx[2] = fmax(mx[2], cs + fabs(ic[2]));
        mx[3] = fmax(mx[3], ic[4] / (rho + 1) * vm);
    }
    for (int q = 0; q < 4; ++q) {
        double v = blockMax(mx[q], sh);
        if (threadIdx.x == 0) out[q] = v;
    }
    if (threadIdx.x == 0)
        for (int q = 4; q < NP; ++q) out[q] = 0;
}

template <class T>
__global__ void k_inactive_reduce(Layout L, Params P, const T* A, const int64_t* nodes, const double* icPrim, int64_t n,
                                  double* partial) {
    __shared__ double sh[8];
    inactiveReduceBlock(L, P, A, nodes, icPrim, n, blockIdx.x, gridDim.x, partial + (size_t)blockIdx.x * NP, sh);
}

// Combines the partials and advances t, timeLevel, dt. mode 0: initial dt only (Solver::initialize);
// mode 1: end of a step.
// dtSched (optional): prescribed step sizes by time level, used instead of the stability limit.
__device__ __forceinline__ void finalizeCase(const Params& P, const double* partials, int nPartials, int mode,
                                             int convHistory, int stopEndTime, double tEnd, double cfl, double sigma,
                                             StepScalars* st, double* hist, int histCap, double* sh,
                                             const double* dtSched) {
    double v[NP];
    for (int q = 0; q < NP; ++q) v[q] = 0;
    for (int b = threadIdx.x; b < nPartials; b += blockDim.x) {
        const double* p = partials + (size_t)b * NP;
        for (int q = 0; q < 4; ++q) v[q] = fmax(v[q], p[q]);
        for (int q = 4; q < NP; ++q) v[q] += p[q];
    }
    double red[NP];
    for (int q = 0; q < 4; ++q) red[q] = blockMax(v[q], sh);
    for (int q = 4; q < NP; ++q) red[q] = blockSum(v[q], sh);
    if (threadIdx.x != 0) return;
    if (red[9] > 0) st->nonFinite = 1;
    double t = st->t;
    if (mode == 1) {
        if (convHistory > 0) {
            int slot = (int)(st->histCount % (uint64_t)histCap);
            double vol = P.hx * P.hy * P.hz;
            for (int q = 0; q < 5; ++q) hist[5 * slot + q] = sqrt(vol * red[4 + q]);
            st->histCount += 1;
        }
        t = t + st->dt;
        st->t = t;
        st->level += 1;
    }
    // updateTimeStepSize, src/Solver.cpp:80-126
    double dtConv = fabs(cfl) / (red[0] / P.hx + red[1] / P.hy + red[2] / P.hz);
    double dx2 = P.hx * P.hx, dy2 = P.hy * P.hy, dz2 = P.hz * P.hz;
    double dtVisc = fabs(sigma) / (red[3] * (1 / dx2 + 1 / dy2 + 1 / dz2));
    double dt = fmin(dtConv, dtVisc);
    if (stopEndTime && t + dt > tEnd) dt = tEnd - t;
    if (dtSched) dt = dtSched[st->level];
    st->dt = dt;
}

__global__ void k_finalize(Params P, const double* partials, int nPartials, int mode, int convHistory, int stopEndTime,
                           double tEnd, double cfl, double sigma, StepScalars* st, double* hist, int histCap,
                           const double* dtSched) {
    __shared__ double sh[8];
    finalizeCase(P, partials, nPartials, mode, convHistory, stopEndTime, tEnd, cfl, sigma, st, hist, histCap, sh,
                 dtSched);
}

struct FaceSet {
    EdgeFace f[6];
    int64_t start[7];  // prefix sums of the owned-node counts
};

// Owned face nodes, ghosts and filter nodes after the final boundary application of a step.
template <class T>
__device__ __forceinline__ void boundaryReduceBlock(const Layout& L, const Params& P, const Window& W, const FaceSet& fs,
                                                    const int64_t* ghostNode, int64_t nGhost, const int64_t* filterList,
                                                    int64_t nFilter, const T* U, const T* A, const uint8_t* cls,
                                                    const T* rhoPre, const T* rhoEPre, int norms, int blk, int nblk,
                                                    double* out, double* sh) {
    const size_t F = L.fstride;
    Eos<T> e = Eos<T>::make(P);
    const double vm = fmax(4. / 3, P.gamma / P.Pr);
    const int64_t nFace = fs.start[6];
    const int64_t total = nFace + nGhost + nFilter;
    double mx[4] = {0, 0, 0, 0}, sm[5] = {0, 0, 0, 0, 0}, bad = 0;
    for (int64_t q = (int64_t)blk * blockDim.x + threadIdx.x; q < total; q += (int64_t)nblk * blockDim.x) {
        size_t o;
        bool filt = false;
        size_t w = 0;
        if (q < nFace) {
            int f = 0;
            while (q >= fs.start[f + 1]) ++f;
            const EdgeFace& ef = fs.f[f];
            int64_t r = q - fs.start[f];
            int lo[3] = {ef.i0, ef.j0, ef.k0}, hi[3] = {ef.i1, ef.j1, ef.k1};
            int ax1 = ef.axis == 0 ? 1 : 0, ax2 = ef.axis == 2 ? 1 : 2;
            int n1 = hi[ax1] - lo[ax1] + 1;
            int c3[3];
            c3[ef.axis] = lo[ef.axis];
            c3[ax1] = lo[ax1] + (int)(r % n1);
            c3[ax2] = lo[ax2] + (int)(r / n1);
            o = L.off(c3[0], c3[1], c3[2]);
            if (cls[o] & CLS_DT_SKIP) continue;
        } else if (q < nFace + nGhost) {
            o = (size_t)ghostNode[q - nFace];
        } else {
            int64_t m = filterList[q - nFace - nGhost];
            int i = (int)(m % L.NI);
            int64_t r = m / L.NI;
            int j = (int)(r % L.NJ), k = (int)(r / L.NJ);
            o = L.off(i, j, k);
            filt = true;
            w = W.idx(i, j, k);
        }
        double a[5];
        for (int c = 0; c < 5; ++c) a[c] = (double)A[o + c * F];
        for (int c = 0; c < 5; ++c)
            if (!isfinite(a[c])) bad += 1;
        T rho = A[o], rE = A[o + 4 * F];
        if (filt) {
            if (norms)
                for (int c = 0; c < 5; ++c) {
                    double d = (double)U[o + c * F] - a[c];
                    sm[c] += d * d;
                }
            rho = rhoPre[w];
            rE = rhoEPre[w];
        }
        T u, v, ww, p, Tt, mu;
        derivePrims(e, rho, A[o + F], A[o + 2 * F], A[o + 3 * F], rE, u, v, ww, p, Tt, mu);
        double cs = sqrt((1 + P.gamma * (double)p) / (1 + a[0]));
        mx[0] = fmax(mx[0], cs + fabs((double)u));
        mx[1] = fmax(mx[1], cs + fabs((double)v));
        mx[2] = fmax(mx[2], cs + fabs((double)ww));
        mx[3] = fmax(mx[3], (double)mu / (a[0] + 1) * vm);
    }
    for (int q = 0; q < 4; ++q) {
        double v = blockMax(mx[q], sh);
        if (threadIdx.x == 0) out[q] = v;
    }
    for (int q = 0; q < 5; ++q) {
        double v = blockSum(sm[q], sh);
        if (threadIdx.x == 0) out[4 + q] = v;
    }
    double b = blockSum(bad, sh);
    if (threadIdx.x == 0) out[9] = b;
}

template <class T>
__global__ void k_boundary_reduce(Layout L, Params P, Window W, FaceSet fs, const int64_t* ghostNode, int64_t nGhost,
                                  const int64_t* filterList, int64_t nFilter, const T* U, const T* A, const uint8_t* cls,
                                  const T* rhoPre, const T* rhoEPre, int norms, double* partials) {
    __shared__ double sh[8];
    boundaryReduceBlock(L, P, W, fs, ghostNode, nGhost, filterList, nFilter, U, A, cls, rhoPre, rhoEPre, norms,
                        blockIdx.x, gridDim.x, partials + (size_t)blockIdx.x * NP, sh);
}

// ---- batched forms: block b belongs to case findCase(blockStart, b); partial slots of a case follow its
// stage-4 block partials (boundary blocks first, then untouched-solid blocks)
template <class T>
__global__ void k_boundary_reduce_batched(const CaseView<T>* __restrict__ cv, const int* __restrict__ blockStart,
                                          int ncases, int parity) {
    __shared__ double sh[8];
    const int b = blockIdx.x;
    const int c = findCase(blockStart, ncases, b);
    const CaseView<T>& v = cv[c];
    FaceSet fs;
    for (int f = 0; f < 6; ++f) fs.f[f] = v.faces[f];
    for (int f = 0; f < 7; ++f) fs.start[f] = v.faceStart[f];
    const int local = b - blockStart[c];
    boundaryReduceBlock(v.L, v.P, v.W, fs, v.G.node, v.nGhost, v.filterList, v.nFilter,
                        (const T*)roleBuffer(v, 0, parity), (const T*)roleBuffer(v, 1, parity), v.cls,
                        (const T*)v.rhoPre, (const T*)v.rhoEPre, v.norms, local, v.nBoundaryBlocks,
                        v.partials + (size_t)(v.nStage4 + local) * NP, sh);
}

template <class T>
__global__ void k_inactive_reduce_batched(const CaseView<T>* __restrict__ cv, const int* __restrict__ blockStart,
                                          int ncases, int parity) {
    __shared__ double sh[8];
    const int b = blockIdx.x;
    const int c = findCase(blockStart, ncases, b);
    const CaseView<T>& v = cv[c];
    const int local = b - blockStart[c];
    inactiveReduceBlock(v.L, v.P, (const T*)roleBuffer(v, 1, parity), v.inactiveList, v.inactiveIC, v.nInactive, local,
                        v.nInactiveBlocks, v.partials + (size_t)(v.nStage4 + v.nBoundaryBlocks + local) * NP, sh);
}

template <class T>
__global__ void k_finalize_batched(const CaseView<T>* __restrict__ cv) {
    __shared__ double sh[8];
    const CaseView<T>& v = cv[blockIdx.x];
    finalizeCase(v.P, v.partials, v.nStage4 + v.nBoundaryBlocks + v.nInactiveBlocks, 1, v.convHistory, v.stopEndTime,
                 v.tEnd, v.cfl, v.sigma, v.scal, v.hist, v.histCap, sh, nullptr);
}

}  // namespace

template <class T>
void launchBoundaryReduceBatched(hipStream_t s, const CaseView<T>* cv, const int* blockStart, int totalBlocks, int ncases,
                                 int parity) {
    if (totalBlocks > 0)
        hipLaunchKernelGGL(k_boundary_reduce_batched<T>, dim3(totalBlocks), dim3(RB), 0, s, cv, blockStart, ncases, parity);
}

template <class T>
void launchInactiveReduceBatched(hipStream_t s, const CaseView<T>* cv, const int* blockStart, int totalBlocks, int ncases,
                                 int parity) {
    if (totalBlocks > 0)
        hipLaunchKernelGGL(k_inactive_reduce_batched<T>, dim3(totalBlocks), dim3(RB), 0, s, cv, blockStart, ncases, parity);
}

template <class T>
void launchFinalizeBatched(hipStream_t s, const CaseView<T>* cv, int ncases) {
    hipLaunchKernelGGL(k_finalize_batched<T>, dim3(ncases), dim3(RB), 0, s, cv);
}

template void launchBoundaryReduceBatched<float>(hipStream_t, const CaseView<float>*, const int*, int, int, int);
template void launchBoundaryReduceBatched<double>(hipStream_t, const CaseView<double>*, const int*, int, int, int);
template void launchInactiveReduceBatched<float>(hipStream_t, const CaseView<float>*, const int*, int, int, int);
template void launchInactiveReduceBatched<double>(hipStream_t, const CaseView<double>*, const int*, int, int, int);
template void launchFinalizeBatched<float>(hipStream_t, const CaseView<float>*, int);
template void launchFinalizeBatched<double>(hipStream_t, const CaseView<double>*, int);

template <class T>
int launchBoundaryReduce(hipStream_t s, const Layout& L, const Params& P, const Window& W, const EdgeFace faces[6],
                         const int64_t* ghostNode, int64_t nGhost, const int64_t* filterList, int64_t nFilter,
                         const T* U, const T* A, const uint8_t* cls, const T* rhoPre, const T* rhoEPre, int norms,
                         double* partials) {
    FaceSet fs;
    fs.start[0] = 0;
    for (int f = 0; f < 6; ++f) {
        fs.f[f] = faces[f];
        const EdgeFace& e = faces[f];
        int64_t n = (int64_t)(e.i1 - e.i0 + 1) * (e.j1 - e.j0 + 1) * (e.k1 - e.k0 + 1);
        fs.start[f + 1] = fs.start[f] + std::max<int64_t>(n, 0);
    }
    int64_t total = fs.start[6] + nGhost + nFilter;
    int nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, (total + RB - 1) / RB));
    hipLaunchKernelGGL(k_boundary_reduce<T>, dim3(nb), dim3(RB), 0, s, L, P, W, fs, ghostNode, nGhost, filterList, nFilter,
                       U, A, cls, rhoPre, rhoEPre, norms, partials);
    return nb;
}

template int launchBoundaryReduce<double>(hipStream_t, const Layout&, const Params&, const Window&, const EdgeFace[6],
                                          const int64_t*, int64_t, const int64_t*, int64_t, const double*, const double*,
                                          const uint8_t*, const double*, const double*, int, double*);
template int launchBoundaryReduce<float>(hipStream_t, const Layout&, const Params&, const Window&, const EdgeFace[6],
                                         const int64_t*, int64_t, const int64_t*, int64_t, const float*, const float*,
                                         const uint8_t*, const float*, const float*, int, double*);

template <class T>
void launchStepReduce(hipStream_t s, const Layout& L, const Params& P, const T* U, const T* A, const uint8_t* cls,
                      const Window& W, const T* rhoPre, const T* rhoEPre, int withNorms, double* partials, int nBlocks,
                      const T* prim) {
    hipLaunchKernelGGL(k_step_reduce<T>, dim3(nBlocks), dim3(RB), 0, s, L, P, U, A, cls, W, rhoPre, rhoEPre, withNorms,
                       partials, prim);
}

template <class T>
void launchInactiveReduce(hipStream_t s, const Layout& L, const Params& P, const T* A, const int64_t* nodes,
                          const double* icPrim, int64_t n, double* partial) {
    hipLaunchKernelGGL(k_inactive_reduce<T>, dim3(inactiveReduceBlocks(n)), dim3(RB), 0, s, L, P, A, nodes, icPrim, n,
                       partial);
}

void launchFinalize(hipStream_t s, const Params& P, const double* partials, int nPartials, int mode, int convHistory,
                    int stopEndTime, double tEnd, double cfl, double sigma, StepScalars* st, double* hist, int histCap,
                    const double* dtSched) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RB), 0, s, P, partials, nPartials, mode, convHistory, stopEndTime, tEnd,
                       cfl, sigma, st, hist, histCap, dtSched);
}

int reducePartialSlots() { return NP; }

int inactiveReduceBlocks(int64_t n) { return (int)std::min<int64_t>(64, std::max<int64_t>(1, (n + RB * 8 - 1) / (RB * 8))); }

template void launchStepReduce<double>(hipStream_t, const Layout&, const Params&, const double*, const double*,
                                       const uint8_t*, const Window&, const double*, const double*, int, double*, int,
                                       const double*);
template void launchStepReduce<float>(hipStream_t, const Layout&, const Params&, const float*, const float*,
                                      const uint8_t*, const Window&, const float*, const float*, int, double*, int,
                                      const float*);
template void launchInactiveReduce<double>(hipStream_t, const Layout&, const Params&, const double*, const int64_t*,
                                           const double*, int64_t, double*);
template void launchInactiveReduce<float>(hipStream_t, const Layout&, const Params&, const float*, const int64_t*,
                                          const double*, int64_t, double*);

}  // namespace nsfsi
