/*
 * solver.hip — host orchestration of one case on one GPU: the Mesh (classification, boundaries, device
 * fields) and the Solver (RK4 stages, boundary applications, dt), mirroring the reference's
 * Mesh / Solver split (src/Mesh.cpp, src/Solver.cpp). Kernels are launched on one HIP stream; a time
 * step never synchronises with the host.
 */
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "adjoint.h"
#include "case.h"
#include "device_common.h"
#include "kernels.h"
#include "stage.h"
#include "totals.h"

namespace nsfsi {

// ------------------------------------------------------------------------------------------------ errors

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

Layout makeLayout(int NI, int NJ, int NK, size_t eb) {
    Layout L;
    L.NI = NI;
    L.NJ = NJ;
    L.NK = NK;
    // FP32 rows hold whole segments (256 B), so the 8-bit compensation carry, stored at a quarter of the byte
    // offsets, covers 64 aligned bytes per segment; FP64 rows are padded to 128 B
    const size_t rowElems = eb == 4 ? (size_t)Layout::kSegSlots : 128 / eb;
    const size_t slots = (size_t)Layout::col(NI - 1) + 1;  // through the last node's slot
    L.pitch = (int)((slots + rowElems - 1) / rowElems * rowElems);
    L.plane = (size_t)NJ * L.pitch;
    size_t fe = 256 / eb;
    L.fstride = ((size_t)NK * L.plane + fe - 1) / fe * fe;
    return L;
}

namespace {

__global__ void k_flag_inactive(Layout L, uint8_t* cls, Window W, int perX, int perY, int perZ) {
    int i = W.i0 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = W.j0 + blockIdx.y, k = W.k0 + blockIdx.z;
    if (i >= W.i0 + W.ni) return;
    size_t o = L.off(i, j, k);
    if ((cls[o] & CLS_TYPE_MASK) != NS_SOLID_INACTIVE) return;
    bool onX = i == 0 || i == L.NI - 1, onY = j == 0 || j == L.NJ - 1, onZ = k == 0 || k == L.NK - 1;
    if (!onX && !onY && !onZ) { cls[o] |= CLS_INACTIVE_IC; return; }
    // A node on a periodic plane receives a copy of its source node; when that source is itself an
    // untouched inactive node, the copy adds nothing new to the dt maxima.
    int si = i, sj = j, sk = k;
    bool periodicOwner = false;
    if (onZ && perZ) { sk = k == 0 ? L.NK - 2 : 1; periodicOwner = true; }
    else if (!onZ && onY && perY) { sj = j == 0 ? L.NJ - 2 : 1; periodicOwner = true; }
    else if (!onZ && !onY && onX && perX) { si = i == 0 ? L.NI - 2 : 1; periodicOwner = true; }
    if (!periodicOwner) return;
    bool sOnX = si == 0 || si == L.NI - 1, sOnY = sj == 0 || sj == L.NJ - 1, sOnZ = sk == 0 || sk == L.NK - 1;
    if ((cls[L.off(si, sj, sk)] & CLS_TYPE_MASK) == NS_SOLID_INACTIVE && !sOnX && !sOnY && !sOnZ) cls[o] |= CLS_DT_SKIP;
}

template <class T>
__global__ void k_gather_ic(Layout L, Params P, const T* U, const int64_t* nodes, int64_t n, double* ic) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int64_t m = nodes[q];
    int i = (int)(m % L.NI);
    int64_t r = m / L.NI;
    int j = (int)(r % L.NJ), k = (int)(r / L.NJ);
    size_t o = L.off(i, j, k);
    Eos<T> e = Eos<T>::make(P);
    T u, v, w, p, Tt, mu;
    derivePrims(e, U[o], U[o + L.fstride], U[o + 2 * L.fstride], U[o + 3 * L.fstride], U[o + 4 * L.fstride], u, v, w, p, Tt,
                mu);
    ic[5 * q] = u; ic[5 * q + 1] = v; ic[5 * q + 2] = w; ic[5 * q + 3] = p; ic[5 * q + 4] = mu;
}

// Boundary nodes with constant conserved values (symmetry mirrors of untouched solid nodes).
template <class T>
__global__ void k_apply_fixed(Layout L, const int64_t* nodes, const double* vals, int64_t n, T* S) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const size_t o = (size_t)nodes[q];
    for (int c = 0; c < 5; ++c) S[o + (size_t)c * L.fstride] = (T)vals[5 * q + c];
}

// Primitive output for SolidInactive nodes: their initial values (never updated by the reference).
template <class T>
__global__ void k_patch_inactive(Layout L, Params P, T* prim, const int64_t* nodes, const double* ic, int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int64_t m = nodes[q];
    int i = (int)(m % L.NI);
    int64_t r = m / L.NI;
    int j = (int)(r % L.NJ), k = (int)(r / L.NJ);
    size_t o = L.off(i, j, k);
    const double* c = ic + 5 * q;
    prim[o] = (T)c[0];
    prim[o + L.fstride] = (T)c[1];
    prim[o + 2 * L.fstride] = (T)c[2];
    prim[o + 3 * L.fstride] = (T)c[3];
    // T' from the initial state: T' = (gamma p' - rho')/(1+rho') with rho' = 0 for every built-in IC
    prim[o + 5 * L.fstride] = (T)c[4];
    prim[o + 6 * L.fstride] = (T)(c[4] * (1 / ((P.gamma - 1) * P.Pr)));
}

template <class T>
__global__ void k_totals(Layout L, const T* S, const uint8_t* cls, double* partials) {
    auto offOf = [&](size_t n) {
        const int i = (int)(n % L.NI);
        const size_t r = n / L.NI;
        return L.off(i, (int)(r % L.NJ), (int)(r / L.NJ));
    };
    totalsBlock<T>(L.nodes(), offOf, L.fstride, S, cls, partials);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------------------------------------ Case

Case::Case(const ns_config& c) : cfg(c) {
    setDerived(cfg);
    std::string v = validateConfig(cfg);
    if (!v.empty()) throw ConfigError(v);
    fp32 = cfg.precision == NS_FP32;
    eb = fp32 ? sizeof(float) : sizeof(double);
    int dev = 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) throw HipError("no HIP device");
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw HipError(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    L = makeLayout(cfg.NI, cfg.NJ, cfg.NK, eb);
}

Case::~Case() {
    // a destructor cannot throw: work still queued on the stream is waited for, and a failed release is
    // reported on stderr (the device memory of a failed hipFree stays with the process)
    auto report = [](hipError_t e, const char* what) {
        if (e != hipSuccess) std::fprintf(stderr, "nsfsi: %s failed in the case destructor: %s\n", what, hipGetErrorString(e));
    };
    if (stream) report(hipStreamSynchronize(stream), "hipStreamSynchronize");
    for (int q = 0; q < 2; ++q) {
        if (graphExec[q]) report(hipGraphExecDestroy(graphExec[q]), "hipGraphExecDestroy");
        graphExec[q] = nullptr;
        graphState[q] = nullptr;
    }
    for (void* p : allocations) report(hipFree(p), "hipFree");
    if (stream) report(hipStreamDestroy(stream), "hipStreamDestroy");
}

void* Case::alloc(size_t bytes) {
    void* p = nullptr;
    hipCheck(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc");
    allocations.push_back(p);
    return p;
}

// Mesh::setupBoundaries, src/Mesh.cpp:25-58, with the face types and body of the configuration.
void Case::setupBoundaries() {
    double Ls[3] = {cfg.L_x, cfg.L_y, cfg.L_z};
    int Ns[3] = {cfg.NI, cfg.NJ, cfg.NK};
    double h[3], off[3];
    for (int a = 0; a < 3; ++a) {
        int tmin = cfg.bc[2 * a], tmax = cfg.bc[2 * a + 1];
        int gmin = (tmin == NS_BC_PERIODIC || tmin == NS_BC_SYMMETRY) ? 1 : 0;
        int gmax = tmax == NS_BC_SYMMETRY ? 1 : 0;
        h[a] = Ls[a] / (double)(Ns[a] - 1 - gmin - gmax);
        off[a] = gmin;
    }
    P.hx = h[0]; P.hy = h[1]; P.hz = h[2];
    P.ox = off[0]; P.oy = off[1]; P.oz = off[2];
    P.gamma = cfg.gamma;
    P.Pr = cfg.Pr;
    P.Re = cfg.Re;
    P.ScPlusOne = 1 + cfg.sutherlands_C2 / cfg.T_0;
    P.M_0 = cfg.M_0;
    P.rampTime = cfg.inlet_ramp_time;
    // identifyOwnedNodes in push order z-min, z-max, y-min, y-max, x-min, x-max; applied in reverse
    int ulo[3] = {0, 0, 0}, uhi[3] = {cfg.NI - 1, cfg.NJ - 1, cfg.NK - 1};
    const int order[6][2] = {{2, 0}, {2, 1}, {1, 0}, {1, 1}, {0, 0}, {0, 1}};
    EdgeFace pushed[6];
    for (int q = 0; q < 6; ++q) {
        int a = order[q][0], s = order[q][1];
        EdgeFace f;
        f.axis = a;
        f.side = s;
        f.type = cfg.bc[2 * a + s];
        int lo[3] = {ulo[0], ulo[1], ulo[2]}, hi[3] = {uhi[0], uhi[1], uhi[2]};
        if (s == 0) { hi[a] = lo[a]; ulo[a]++; }
        else { lo[a] = hi[a]; uhi[a]--; }
        f.i0 = lo[0]; f.i1 = hi[0]; f.j0 = lo[1]; f.j1 = hi[1]; f.k0 = lo[2]; f.k1 = hi[2];
        pushed[q] = f;
    }
    for (int q = 0; q < 6; ++q) faces[q] = pushed[5 - q];
    // The outlet face goes into the fused stage kernel when the kernel can evaluate it and it is independent
    // of every face applied before it: applying it first, inside the stage, then changes nothing.
    fusedOutlet = -1;
    for (int q = 0; q < 6; ++q) bcFaces[q] = faces[q];
    for (int q = 0; q < 6 && fusedOutlet < 0; ++q) {
        if (!outletFusable(L, faces[q])) continue;
        bool ok = true;
        for (int r = 0; r < q; ++r) ok = ok && edgeFacesIndependent(L, faces[r], faces[q]);
        if (ok) fusedOutlet = q;
    }
    if (fusedOutlet >= 0) {  // empty owned box: the edge-face pass skips it
        bcFaces[fusedOutlet].i0 = 1;
        bcFaces[fusedOutlet].i1 = 0;
    }
    // bodies, in the order of the reference's body list (src/Mesh.cpp:52-57)
    bodies = BodySet{};
    W = Window{0, 0, 0, 0, 0, 0};
    auto addBody = [&](int type, int axis, const double* centre, double r) {
        BodyDesc B{};
        B.type = type;
        B.axis = axis;
        B.cx = centre[0]; B.cy = centre[1]; B.cz = centre[2];
        B.r = r;
        B.mpb = cfg.machine_precision_buffer;
        const int layer = 2;
        if (B.type == NS_BODY_SPHERE) B.filterWidth = layer * std::max({h[0], h[1], h[2]});
        else if (B.axis == 0) B.filterWidth = layer * std::max(h[1], h[2]);
        else if (B.axis == 1) B.filterWidth = layer * std::max(h[0], h[2]);
        else B.filterWidth = layer * std::max(h[0], h[1]);
        // getSphereBoundingBox / getCylinderBoundingBox, src/Boundary.cpp:938-965, :1056-1072: centre index
        // round(c/h) without the origin offset, radius ceil(r/h) + 1 + layer
        double c[3] = {B.cx, B.cy, B.cz};
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            if (B.type == NS_BODY_CYLINDER && a == B.axis) { lo[a] = 0; hi[a] = Ns[a] - 1; continue; }
            long rad = (long)std::ceil(B.r / h[a]) + 1 + layer;
            double rc = std::round(c[a] / h[a]);
            if (!(rc >= 0)) throw GeometryError("body bounding box below index 0");
            long ci = (long)rc;
            if (ci - rad < 0 || ci + rad > Ns[a] - 1) throw GeometryError("body bounding box leaves the mesh");
            lo[a] = (int)(ci - rad);
            hi[a] = (int)(ci + rad);
        }
        B.box = Window{lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
        for (int q = 0; q < bodies.n; ++q) {
            const Window& o = bodies.b[q].box;
            bool apart = false;
            for (int a = 0; a < 3; ++a) {
                const int l0 = a == 0 ? o.i0 : a == 1 ? o.j0 : o.k0, n0 = a == 0 ? o.ni : a == 1 ? o.nj : o.nk;
                apart = apart || hi[a] < l0 || l0 + n0 - 1 < lo[a];
            }
            if (!apart) throw GeometryError("the bounding boxes of two immersed bodies overlap");
        }
        bodies.b[bodies.n++] = B;
    };
    if (cfg.body_type != NS_BODY_NONE) addBody(cfg.body_type, cfg.body_axis, cfg.body_centre, cfg.body_radius);
    for (int q = 0; q < cfg.n_extra_bodies; ++q)
        addBody(cfg.extra_bodies[q].type, cfg.extra_bodies[q].axis, cfg.extra_bodies[q].centre, cfg.extra_bodies[q].radius);
    // the window: the smallest box holding every body's box (side tables and ghost map are dense over it)
    if (bodies.n > 0) {
        int lo[3] = {Ns[0], Ns[1], Ns[2]}, hi[3] = {-1, -1, -1};
        for (int q = 0; q < bodies.n; ++q) {
            const Window& o = bodies.b[q].box;
            const int l[3] = {o.i0, o.j0, o.k0}, n[3] = {o.ni, o.nj, o.nk};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], l[a]);
                hi[a] = std::max(hi[a], l[a] + n[a] - 1);
            }
        }
        W = Window{lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    }
}

// Mesh::categorizeNodes, src/Mesh.cpp:61-82, on the device. The bodies' boxes are disjoint, so classifying
// them in one pass equals the reference's body-by-body classification; ghost augmentation, ghost and filter
// lists run body by body, each over its own box, so every list holds one body's entries after the previous
// body's.
void Case::categorizeNodes() {
    cls = (uint8_t*)alloc((size_t)L.NK * L.plane);
    hipCheck(hipMemsetAsync(cls, 0, (size_t)L.NK * L.plane, stream), "memset cls");
    int ft[6];
    for (int q = 0; q < 6; ++q) ft[q] = (cfg.bc[q] == NS_BC_INLET || cfg.bc[q] == NS_BC_OUTLET) ? NS_FLUID_EDGE : NS_FLUID_GHOST;
    launchClassify(stream, L, cls, ft, bodies, P);
    nGhost = nFilter = nInactive = 0;
    for (int q = 0; q <= NS_MAX_BODIES; ++q) ghostOff[q] = filterOff[q] = 0;
    if (bodies.n == 0) return;
    launchMarkGhosts(stream, L, cls, W);
    size_t cap = 0, nbMax = 0;  // list capacity: the sum of the body boxes
    for (int b = 0; b < bodies.n; ++b) {
        cap += bodies.b[b].box.size();
        nbMax = std::max(nbMax, (bodies.b[b].box.size() + 255) / 256);
    }
    int32_t* counts = (int32_t*)alloc(nbMax * sizeof(int32_t));
    int32_t* offsets = (int32_t*)alloc(nbMax * sizeof(int32_t));
    int32_t* flags = (int32_t*)alloc(2 * sizeof(int32_t));  // changed, error
    int64_t* list = (int64_t*)alloc(cap * sizeof(int64_t));
    double* bi = (double*)alloc(3 * cap * sizeof(double));
    double* ip = (double*)alloc(3 * cap * sizeof(double));
    auto readFlags = [&](int32_t* h) {
        hipCheck(hipMemcpyAsync(h, flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, stream), "copy flags");
        hipCheck(hipStreamSynchronize(stream), "sync");
    };
    // iterative ghost augmentation (CylinderBody::identifyRelatedNodes, src/Boundary.cpp:907-916), body by
    // body: repeat until no SolidInactive node remains in any of the body's image-point boxes
    for (int b = 0; b < bodies.n; ++b) {
        const BodyDesc& B = bodies.b[b];
        const int64_t base = ghostOff[b];
        for (int round = 0;; ++round) {
            int64_t n = compactWindow(stream, L, cls, B.box, CLS_TYPE_MASK, NS_SOLID_GHOST, list + base, counts, offsets);
            hipCheck(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), stream), "memset flags");
            launchImagePoints(stream, L, cls, P, B, list + base, n, bi + 3 * base, ip + 3 * base, flags, flags + 1);
            int32_t h[2];
            readFlags(h);
            if (h[1] & 2) throw GeometryError("an image-point box of one immersed body reaches another body's solid nodes");
            if (h[1]) throw GeometryError("image point box leaves the mesh");
            if (h[0] == 0) { ghostOff[b + 1] = base + n; break; }
            if (round > 1000) throw GeometryError("ghost augmentation does not converge");
        }
    }
    nGhost = ghostOff[bodies.n];
    ghostList = list;
    int32_t* gmap = (int32_t*)alloc(W.size() * sizeof(int32_t));
    hipCheck(hipMemsetAsync(gmap, 0xff, W.size() * sizeof(int32_t), stream), "memset gmap");
    launchFillGhostMap(stream, L, W, ghostList, nGhost, gmap);
    G.node = (int64_t*)alloc(nGhost * sizeof(int64_t));
    G.corner = (int64_t*)alloc(8 * nGhost * sizeof(int64_t));
    G.mask = (uint8_t*)alloc(nGhost);
    G.rel = (double*)alloc(3 * nGhost * sizeof(double));
    G.ip = ip;
    G.bi = bi;
    G.luD = (double*)alloc(64 * nGhost * sizeof(double));
    G.luN = (double*)alloc(64 * nGhost * sizeof(double));
    G.perm = (int8_t*)alloc(32 * nGhost);
    G.wD = (double*)alloc(8 * nGhost * sizeof(double));
    G.wN = (double*)alloc(8 * nGhost * sizeof(double));
    G.boxLo = (int32_t*)alloc(3 * nGhost * sizeof(int32_t));
    hipCheck(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), stream), "memset flags");
    for (int b = 0; b < bodies.n; ++b)
        launchGhostGeometry(stream, L, cls, P, W, ghostList + ghostOff[b], ghostOff[b], ghostOff[b + 1] - ghostOff[b],
                            gmap, ghostTablesFrom(G, ghostOff[b]), flags + 1);
    int32_t h[2];
    readFlags(h);
    if (h[1] & 4) throw GeometryError("Impossible situation. Found solid node around image point.");
    if (h[1] & 2) throw GeometryError("an image-point box holds a ghost node of another immersed body");
    if (h[1]) throw GeometryError("ghost geometry error");
    // filter nodes (ascending within each body)
    filterList = (int64_t*)alloc(std::max<size_t>(1, cap) * sizeof(int64_t));
    for (int b = 0; b < bodies.n; ++b)
        filterOff[b + 1] = filterOff[b] + compactWindow(stream, L, cls, bodies.b[b].box, CLS_FILTER, CLS_FILTER,
                                                        filterList + filterOff[b], counts, offsets);
    nFilter = filterOff[bodies.n];
    // SolidInactive bookkeeping for the dt scan
    size_t nbW = (W.size() + 255) / 256;
    int32_t* countsW = nbW > nbMax ? (int32_t*)alloc(nbW * sizeof(int32_t)) : counts;
    int32_t* offsetsW = nbW > nbMax ? (int32_t*)alloc(nbW * sizeof(int32_t)) : offsets;
    dim3 blk(64), grd(nblk(W.ni, 64), W.nj, W.nk);
    hipLaunchKernelGGL(k_flag_inactive, grd, blk, 0, stream, L, cls, W, cfg.bc[NS_XMIN] == NS_BC_PERIODIC,
                       cfg.bc[NS_YMIN] == NS_BC_PERIODIC, cfg.bc[NS_ZMIN] == NS_BC_PERIODIC);
    inactiveList = (int64_t*)alloc(W.size() * sizeof(int64_t));
    nInactive = compactWindow(stream, L, cls, W, CLS_INACTIVE_IC, CLS_INACTIVE_IC, inactiveList, countsW, offsetsW);
    inactiveIC = (double*)alloc(5 * std::max<int64_t>(1, nInactive) * sizeof(double));
    findFixedBoundary();
}

// The reference's symmetry face mirrors the stored primitive values of the node two in
// (src/Boundary.cpp:336-417); a periodic face copies every stored value of its source
// (src/Boundary.cpp:250-300). An untouched SolidInactive node keeps the primitive values of the initial
// state for good, while its conserved values alternate between the initial state and the zero-initialised
// second array (the Old/Cur swap, src/Mesh.cpp:144-151). A symmetry node whose source chain ends at such a
// node therefore holds constant conserved values, which the forms that derive primitives from the state
// cannot reproduce: they are found here, once, and written after every edge-face pass.
void Case::findFixedBoundary() {
    nFixed = 0;
    fixedHostNodes.clear();
    fixedSrc.clear();
    fixedMirror.clear();
    bool sym = false;
    for (int q = 0; q < 6; ++q) sym = sym || cfg.bc[q] == NS_BC_SYMMETRY;
    if (bodies.n == 0 || !sym || nInactive == 0) return;
    // a chain reaches at most three nodes in from a face: only windows that come that close matter
    const int Nn[3] = {L.NI, L.NJ, L.NK}, wl[3] = {W.i0, W.j0, W.k0}, wn[3] = {W.ni, W.nj, W.nk};
    bool near = false;
    for (int a = 0; a < 3; ++a) near = near || wl[a] <= 3 || wl[a] + wn[a] - 1 >= Nn[a] - 4;
    if (!near) return;
    std::vector<uint8_t> hc((size_t)L.NK * L.plane);
    hipCheck(hipMemcpyAsync(hc.data(), cls, hc.size(), hipMemcpyDeviceToHost, stream), "copy cls");
    hipCheck(hipStreamSynchronize(stream), "sync");
    auto owner = [&](int i, int j, int k) -> const EdgeFace* {
        for (int q = 0; q < 6; ++q) {
            const EdgeFace& f = faces[q];
            if (i >= f.i0 && i <= f.i1 && j >= f.j0 && j <= f.j1 && k >= f.k0 && k <= f.k1) return &f;
        }
        return nullptr;
    };
    struct Src { bool ok; int i, j, k; uint8_t mirror; };
    // where the stored primitives of (i,j,k) come from, when they are constant
    std::function<Src(int, int, int, int)> prims = [&](int i, int j, int k, int depth) -> Src {
        if (depth > 6) return Src{false, 0, 0, 0, 0};
        if (hc[L.off(i, j, k)] & CLS_INACTIVE_IC) return Src{true, i, j, k, 0};
        const EdgeFace* f = owner(i, j, k);
        if (!f) return Src{false, 0, 0, 0, 0};
        int c[3] = {i, j, k};
        if (f->type == NS_BC_SYMMETRY) {
            c[f->axis] += f->side == 0 ? 2 : -2;
            Src r = prims(c[0], c[1], c[2], depth + 1);
            if (r.ok) r.mirror ^= (uint8_t)(1u << f->axis);
            return r;
        }
        if (f->type == NS_BC_PERIODIC) {
            c[f->axis] = f->side == 0 ? Nn[f->axis] - 2 : 1;
            return prims(c[0], c[1], c[2], depth + 1);
        }
        return Src{false, 0, 0, 0, 0};
    };
    // nodes with constant conserved values: symmetry nodes with constant primitives, and periodic copies
    // of such nodes
    std::function<Src(int, int, int, int)> fixedCons = [&](int i, int j, int k, int depth) -> Src {
        if (depth > 6) return Src{false, 0, 0, 0, 0};
        const EdgeFace* f = owner(i, j, k);
        if (!f) return Src{false, 0, 0, 0, 0};
        if (f->type == NS_BC_SYMMETRY) return prims(i, j, k, depth);
        if (f->type == NS_BC_PERIODIC) {
            int c[3] = {i, j, k};
            c[f->axis] = f->side == 0 ? Nn[f->axis] - 2 : 1;
            return fixedCons(c[0], c[1], c[2], depth + 1);
        }
        return Src{false, 0, 0, 0, 0};
    };
    auto visit = [&](int i, int j, int k) {
        Src r = fixedCons(i, j, k, 0);
        if (!r.ok) return;
        fixedHostNodes.push_back((int64_t)L.off(i, j, k));
        fixedSrc.push_back((int64_t)L.off(r.i, r.j, r.k));
        fixedMirror.push_back(r.mirror);
    };
    for (int k = 0; k < L.NK; ++k)
        for (int j = 0; j < L.NJ; ++j) {
            const bool edgeRow = k == 0 || k == L.NK - 1 || j == 0 || j == L.NJ - 1;
            if (edgeRow) {
                for (int i = 0; i < L.NI; ++i) visit(i, j, k);
            } else {
                visit(0, j, k);
                visit(L.NI - 1, j, k);
            }
        }
    nFixed = (int64_t)fixedHostNodes.size();
    if (nFixed == 0) return;
    fixedNodes = (int64_t*)alloc(nFixed * sizeof(int64_t));
    fixedVals = (double*)alloc(5 * nFixed * sizeof(double));
    hipCheck(hipMemcpyAsync(fixedNodes, fixedHostNodes.data(), nFixed * sizeof(int64_t), hipMemcpyHostToDevice, stream),
             "copy fixed nodes");
    hipCheck(hipStreamSynchronize(stream), "sync");
}

// Values of the constant boundary nodes from the state being uploaded (its primitive values are the stored
// ones of the untouched solid nodes): the source's primitives, mirrored, turned into conserved values as
// the symmetry face does (src/FlowVariableGroupStructs.h:152-185).
template <class T>
void Case::computeFixedBoundary(const std::vector<T>& h) {
    if (nFixed == 0) return;
    const double g = cfg.gamma;
    std::vector<double> v(5 * nFixed);
    for (int64_t q = 0; q < nFixed; ++q) {
        const size_t o = (size_t)fixedSrc[q];
        const double rho = (double)h[o], ru = (double)h[o + L.fstride], rv = (double)h[o + 2 * L.fstride],
                     rw = (double)h[o + 3 * L.fstride], rE = (double)h[o + 4 * L.fstride];
        double u = ru / (1 + rho), vv = rv / (1 + rho), w = rw / (1 + rho);
        const double p = (g - 1) * (rE - (1 + rho) / 2 * (u * u + vv * vv + w * w));
        const double Tt = (g * p - rho) / (1 + rho);
        if (fixedMirror[q] & 1) u = -u;
        if (fixedMirror[q] & 2) vv = -vv;
        if (fixedMirror[q] & 4) w = -w;
        const double r2 = (g * p - Tt) / (1 + Tt);
        v[5 * q] = r2;
        v[5 * q + 1] = (1 + r2) * u;
        v[5 * q + 2] = (1 + r2) * vv;
        v[5 * q + 3] = (1 + r2) * w;
        v[5 * q + 4] = p / (g - 1) + (1 + r2) / 2 * (u * u + vv * vv + w * w);
    }
    hipCheck(hipMemcpyAsync(fixedVals, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, stream),
             "copy fixed values");
    hipCheck(hipStreamSynchronize(stream), "sync");
}

template <class T>
void Case::allocFields() {
    size_t fb = 5 * L.fstride * sizeof(T);
    U = alloc(fb);
    A = alloc(fb);
    S1 = alloc(fb);
    S2 = alloc(fb);
    prim = alloc(7 * L.fstride * sizeof(T));
    if (fp32 && cfg.kahan) comp = (int8_t*)alloc(5 * L.fstride * sizeof(int8_t));
    hipCheck(hipMemsetAsync(U, 0, fb, stream), "memset");
    hipCheck(hipMemsetAsync(A, 0, fb, stream), "memset");  // the reference's zero-initialised "Old" arrays
    hipCheck(hipMemsetAsync(S1, 0, fb, stream), "memset");
    hipCheck(hipMemsetAsync(S2, 0, fb, stream), "memset");
    size_t wn = std::max<size_t>(1, W.size());
    rhoPre = alloc(wn * sizeof(T));
    rhoEPre = alloc(wn * sizeof(T));
    ftmp = alloc(2 * std::max<int64_t>(1, nFilter) * sizeof(T));
    nPartials = 1024;
    size_t slots = (size_t)std::max(nPartials, fusedMaxBlocks(L)) + 512 + 64;
    partialSlots = slots;
    partials = (double*)alloc(slots * reducePartialSlots() * sizeof(double));
    scal = (StepScalars*)alloc(sizeof(StepScalars));
    hist = (double*)alloc((size_t)histCap * 5 * sizeof(double));
    forceRing = (double*)alloc((size_t)histCap * 4 * sizeof(double));
    StepScalars s0{};
    hipCheck(hipMemcpyAsync(scal, &s0, sizeof(s0), hipMemcpyHostToDevice, stream), "copy scalars");
}

// Initial condition on the host (applyStagnation_IC / applyUniformFlow_IC, src/Solver.cpp:31-75, or
// freestream plus a counter-based seeded perturbation), in the device layout.
template <class T>
void Case::hostInitialField(std::vector<T>& h) const {
    h.assign(5 * L.fstride, (T)0);
    const double g = cfg.gamma;
    double u0 = 0, v0 = 0, w0 = 0;
    bool perturb = false;
    if (cfg.ic_type == NS_IC_UNIFORM_OBLIQUE) { u0 = v0 = w0 = cfg.M_0 / std::sqrt(3); }
    else if (cfg.ic_type == NS_IC_FREESTREAM_GAUSSIAN || cfg.ic_type == NS_IC_FREESTREAM_UNIFORM) { u0 = cfg.M_0; perturb = true; }
    const bool gauss = cfg.ic_type == NS_IC_FREESTREAM_GAUSSIAN;
    const double a = cfg.ic_amplitude;
    const uint64_t seed = cfg.ic_seed;
    auto mix = [](uint64_t z) {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    };
    auto work = [&](int k0, int k1) {
        for (int k = k0; k < k1; ++k)
            for (int j = 0; j < L.NJ; ++j)
                for (int i = 0; i < L.NI; ++i) {
                    double u = u0, v = v0, w = w0, p = 0, Tt = 0;
                    if (perturb) {
                        uint64_t base = (((uint64_t)k * L.NJ + j) * L.NI + i) * 6;
                        double d[3];
                        for (int c = 0; c < 3; ++c) {
                            double r1 = (double)(mix(seed + 0x9E3779B97F4A7C15ULL * (base + 2 * c + 1)) >> 11) * 0x1.0p-53;
                            double r2 = (double)(mix(seed + 0x9E3779B97F4A7C15ULL * (base + 2 * c + 2)) >> 11) * 0x1.0p-53;
                            d[c] = gauss ? a * (std::sqrt(-2.0 * std::log(1.0 - r1)) * std::cos(6.283185307179586 * r2))
                                         : a * (2.0 * r1 - 1.0);
                        }
                        u = u0 + d[0];
                        v = v0 + d[1];
                        w = w0 + d[2];
                    }
                    double rho = (g * p - Tt) / (1 + Tt);
                    double m = 1 + rho;
                    double E = p / (g - 1) + m / 2 * (u * u + v * v + w * w);
                    size_t o = L.off(i, j, k);
                    h[o] = (T)rho;
                    h[o + L.fstride] = (T)(m * u);
                    h[o + 2 * L.fstride] = (T)(m * v);
                    h[o + 3 * L.fstride] = (T)(m * w);
                    h[o + 4 * L.fstride] = (T)E;
                }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if ((size_t)L.NK * L.NJ * L.NI < (1u << 20)) nt = 1;
    std::vector<std::thread> th;
    int chunk = (L.NK + (int)nt - 1) / (int)nt;
    for (unsigned q = 0; q < nt; ++q) {
        int k0 = (int)q * chunk, k1 = std::min(L.NK, k0 + chunk);
        if (k0 < k1) th.emplace_back(work, k0, k1);
    }
    for (auto& t : th) t.join();
}

template <class T>
void Case::uploadState(const std::vector<T>& h) {
    computeFixedBoundary(h);
    hipCheck(hipMemcpyAsync(U, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream), "upload IC");
    if (comp) hipCheck(hipMemsetAsync(comp, 0, 5 * L.fstride * sizeof(int8_t), stream), "memset");
    // pre-filter side tables start as the current values (the reference's primitives match the IC)
    launchFilterCompute<T>(stream, L, W, filterList, nFilter, (const T*)U, (T*)ftmp, (T*)rhoPre, (T*)rhoEPre);
    if (nInactive > 0)
        hipLaunchKernelGGL(k_gather_ic<T>, dim3(nblk(nInactive, 128)), dim3(128), 0, stream, L, P, (const T*)U,
                           inactiveList, nInactive, inactiveIC);
    // stored primitive and transport fields of the initial state (FP64; deriveTransportProperties' conductivity)
    if (!fp32) {
        launchDerive<T>(stream, L, P, (const T*)U, (T*)prim, cls, W, (const T*)rhoPre, (const T*)rhoEPre, 2);
        primsStored = true;
    }
    // updateTimeStepSize(0) on the initial state
    computeDt<T>(0);
    hipCheck(hipStreamSynchronize(stream), "sync");
    hipCheck(hipGetLastError(), "init kernels");
}

void Case::syncStoredPrims() {
    if (!literal() || primsStored) return;
    launchDerive<double>(stream, L, P, (const double*)U, (double*)prim, cls, W, (const double*)rhoPre,
                         (const double*)rhoEPre, 2);
    primsStored = true;
}

template <class T>
void Case::computeDt(int mode) {
    // the stored-primitive form scans the stored fields at every node (untouched solid nodes included)
    const bool stored = !fp32 && primsStored && (literal() || mode == 0);
    launchStepReduce<T>(stream, L, P, (const T*)U, mode == 0 ? (const T*)U : (const T*)A, cls, W, (const T*)rhoPre,
                        (const T*)rhoEPre, mode, partials, nPartials, stored ? (const T*)prim : nullptr);
    int npart = nPartials;
    if (nInactive > 0 && !stored) {
        launchInactiveReduce<T>(stream, L, P, mode == 0 ? (const T*)U : (const T*)A, inactiveList, inactiveIC, nInactive,
                                partials + (size_t)nPartials * reducePartialSlots());
        npart += inactiveReduceBlocks(nInactive);
    }
    launchFinalize(stream, P, partials, npart, mode, cfg.conv_history, cfg.stop_criterion == 1, cfg.t_end,
                   cfg.conv_stability_limit, cfg.visc_stability_limit, scal, hist, histCap, dtSchedule);
}

template <class T>
void Case::applyBCs(T* S, double frac, bool outletDone, T* storedPrims) {
    // Mesh::applyAllBoundaryConditions, src/Mesh.cpp:155-167: edge faces in reverse claim order, then the
    // immersed body (filter first, then ghosts; src/Boundary.cpp:431-456)
    const double* tdev = &scal->t;
    const double* dtdev = &scal->dt;
    launchEdgeFaces<T>(stream, L, P, outletDone ? bcFaces : faces, S, tdev, dtdev, frac, storedPrims);
    if (!storedPrims && nFixed > 0)  // the stored-primitive form mirrors the stored values itself
        hipLaunchKernelGGL(k_apply_fixed<T>, dim3(nblk(nFixed, 128)), dim3(128), 0, stream, L, (const int64_t*)fixedNodes,
                           (const double*)fixedVals, nFixed, S);
    // ImmersedBoundary::applyBoundaryCondition body by body, in list order (src/Mesh.cpp:165-166)
    for (int b = 0; b < bodies.n; ++b) {
        const int64_t f0 = filterOff[b], nf = filterOff[b + 1] - f0, g0 = ghostOff[b], ng = ghostOff[b + 1] - g0;
        launchFilterCompute<T>(stream, L, W, filterList + f0, nf, S, (T*)ftmp, (T*)rhoPre, (T*)rhoEPre);
        launchFilterWrite<T>(stream, L, filterList + f0, nf, S, (const T*)ftmp);
        launchGhostApply<T>(stream, L, P, ghostTablesFrom(G, g0), ng, S, storedPrims);
    }
}

// Solver::marchTimeStep, src/Solver.cpp:130-167 (without the rho filter, see ns_init)
template <class T>
void Case::stepOnce() {
    T* u = (T*)U;
    T* a = (T*)A;
    T* s1 = (T*)S1;
    T* s2 = (T*)S2;
    T* pr = (T*)prim;
    const double* dtdev = &scal->dt;
    T* in[4] = {u, s1, s2, s1};
    T* out[4] = {s1, s2, s1, a};
    const double frac[4] = {0.5, 0.5, 1.0, 1.0};
    // periodic density filter (src/Solver.cpp:134): decided on the device from t and timeLevel
    const bool rhoFilter = !filterDisabled();
    T* fRho = s2;  // S2's density field is free until stage 2
    const int32_t* fFlag = &scal->filterNow;
    if (rhoFilter) {
        FilterSchedule fs{};
        fs.n_intervals = cfg.n_filter_intervals;
        fs.n_change = cfg.n_filter_change_times;
        for (int q = 0; q < NS_MAX_FILTER; ++q) {
            fs.intervals[q] = cfg.filter_intervals[q];
            fs.change_times[q] = cfg.filter_change_times[q];
        }
        launchFilterFlag(stream, fs, scal);
        launchRhoFilter<T>(stream, L, u, fRho, cls, fFlag);
    }
    auto tapeCopy = [&](int slot, const T* S) {
        if (!tapeNow) return;
        const size_t fn = 5 * L.fstride, wn = std::max<size_t>(1, W.size());
        if constexpr (std::is_same<T, float>::value) {  // FP32 forward: the reverse sweep reads FP64 copies
            launchWiden(stream, S, tapeNow->S[slot], fn);
            launchWiden(stream, (const float*)rhoPre, tapeNow->P[slot], wn);
            launchWiden(stream, (const float*)rhoEPre, tapeNow->PE[slot], wn);
        } else {
            hipCheck(hipMemcpyAsync(tapeNow->S[slot], S, fn * sizeof(T), hipMemcpyDeviceToDevice, stream), "tape");
            hipCheck(hipMemcpyAsync(tapeNow->P[slot], rhoPre, wn * sizeof(T), hipMemcpyDeviceToDevice, stream), "tape");
            hipCheck(hipMemcpyAsync(tapeNow->PE[slot], rhoEPre, wn * sizeof(T), hipMemcpyDeviceToDevice, stream), "tape");
        }
    };
    tapeCopy(0, u);
    if (!literal()) primsStored = false;  // the other forms do not keep the stored primitive fields
    for (int st = 1; st <= 4; ++st) {
        if (stageKernel == 1) {
            mark(0, true);
            nStage4Partials = launchStageFused<T>(stream, st, L, P, W, in[st - 1], u, out[st - 1], a, cls,
                                                  (const T*)rhoPre, (const T*)rhoEPre, dtdev, tile,
                                                  fusedScans ? partials : nullptr, cfg.conv_history > 0,
                                                  rhoFilter ? fRho : nullptr, rhoFilter ? fFlag : nullptr, comp,
                                                  stageOutlet());
            mark(0, false);
        } else if (literal()) {
            // stored primitives: the slopes read the stored fields; after the combine the
            // interior primitives are updated (src/Solver.cpp:446-479) and the boundary conditions write theirs
            mark(0, true);
            launchRhsPlain<T>(stream, st, L, P, in[st - 1], pr, u, out[st - 1], a, cls, dtdev, rhoFilter ? fRho : nullptr,
                              rhoFilter ? fFlag : nullptr, nullptr, true);
            mark(0, false);
        } else {
            mark(1, true);
            launchDerive<T>(stream, L, P, in[st - 1], pr, cls, W, (const T*)rhoPre, (const T*)rhoEPre);
            mark(1, false);
            mark(0, true);
            launchRhsPlain<T>(stream, st, L, P, in[st - 1], pr, u, out[st - 1], a, cls, dtdev, rhoFilter ? fRho : nullptr,
                              rhoFilter ? fFlag : nullptr, comp);
            mark(0, false);
        }
        if (st == 1 && rhoFilter) launchRhoCommit<T>(stream, L, u, fRho, a, cls, fFlag);
        if (literal()) {
            mark(1, true);
            launchDerive<T>(stream, L, P, out[st - 1], pr, cls, W, (const T*)rhoPre, (const T*)rhoEPre, 1);
            mark(1, false);
        }
        mark(2, true);
        applyBCs<T>(out[st - 1], frac[st - 1], stageKernel == 1 && stageOutlet(), literal() ? pr : nullptr);
        mark(2, false);
        if (st < 4) tapeCopy(st, out[st - 1]);
    }
    mark(3, true);
    if (stageKernel == 1 && fusedScans) {
        // scans of the interior were done by the stage-4 kernel; add the faces, ghosts, filter nodes and
        // untouched solid nodes, then advance t / timeLevel / dt
        double* pb = partials + (size_t)nStage4Partials * reducePartialSlots();
        int nb = launchBoundaryReduce<T>(stream, L, P, W, faces, G.node, nGhost, filterList, nFilter, (const T*)U,
                                         (const T*)A, cls, (const T*)rhoPre, (const T*)rhoEPre, cfg.conv_history > 0,
                                         pb);
        int npart = nStage4Partials + nb;
        if (nInactive > 0) {
            launchInactiveReduce<T>(stream, L, P, (const T*)A, inactiveList, inactiveIC, nInactive,
                                    partials + (size_t)npart * reducePartialSlots());
            npart += inactiveReduceBlocks(nInactive);
        }
        launchFinalize(stream, P, partials, npart, 1, cfg.conv_history, cfg.stop_criterion == 1, cfg.t_end,
                       cfg.conv_stability_limit, cfg.visc_stability_limit, scal, hist, histCap, dtSchedule);
    } else {
        computeDt<T>(1);
    }
    if (bodyForce && nGhost > 0)  // the scan partials are consumed: reuse them
        launchBodyForce<T>(stream, L, P, W, (const T*)A, cls, (const T*)rhoPre, (const T*)rhoEPre, ghostList, nGhost,
                           partials, scal, forceRing, histCap);
    mark(3, false);
    std::swap(U, A);
}

void Case::setOption(const std::string& name, int64_t value) {
    dropGraphs();
    if (name == "stage_kernel") {
        if (value != 0 && value != 1) throw std::invalid_argument("stage_kernel must be 0 or 1");
        stageKernel = (int)value;
    } else if (name == "fused_scans") {
        fusedScans = value != 0;
    } else if (name == "body_force") {
        bodyForce = value != 0;
    } else if (name == "graph") {
        useGraph = value != 0;
    } else if (name == "tile_ty") {
        int n = 0;
        const int* v = fusedTileHeights(&n);
        bool ok = false;
        for (int q = 0; q < n; ++q) ok = ok || v[q] == value;
        if (!ok) throw std::invalid_argument("tile_ty must be one of the compiled tile heights");
        tile.ty = (int)value;
    } else if (name == "tile_kc") {
        // the partial-slot buffer of the in-kernel scans is sized for z-chunks of >= kMinTileKc planes
        if (value < kMinTileKc || value > 4096) throw std::invalid_argument("tile_kc out of range (8..4096)");
        tile.kc = (int)value;
    } else if (name == "global_access") {
        tile.globalAccess = value != 0;
    } else if (name == "fused_outlet") {
        fuseOutletOpt = value != 0;
    } else if (name == "autotune") {
        if (!initialized) throw StateError("autotune before init");
        autotune(value > 0 ? (int)value : 3);
    } else {
        throw std::invalid_argument("unknown option " + name);
    }
}

const EdgeFace* Case::stageOutlet() const {
    return fuseOutletOpt && fusedOutlet >= 0 ? &faces[fusedOutlet] : nullptr;
}

double Case::autotune(int reps) {
    // stages 1-3 of a step (the stage inputs and outputs of the step itself): they write S1, S2 and the
    // interior of A, which the next step rewrites in the same order before reading them; U, the other nodes
    // of A and the carry are untouched, so the state of the run is unchanged
    int n = 0;
    const int* tys = fusedTileHeights(&n);
    int nk = 0;
    const int* kcs = fusedChunkPlanes(&nk);
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "event");
    hipCheck(hipEventCreate(&e1), "event");
    double best = 1e30;
    FusedTile bestTile = tile;
    tuneTimes.clear();
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < nk; ++b) {
            FusedTile tl;
            tl.ty = tys[a];
            tl.kc = kcs[b];
            tl.globalAccess = tile.globalAccess;
            auto run = [&](int count) {
                for (int r = 0; r < count; ++r)
                    for (int st = 1; st <= 3; ++st) {
                        void* in = st == 1 ? U : (st == 2 ? S1 : S2);
                        void* out = st == 2 ? S2 : S1;
                        if (fp32)
                            launchStageFused<float>(stream, st, L, P, W, (const float*)in, (const float*)U, (float*)out,
                                                    (float*)A, cls, (const float*)rhoPre, (const float*)rhoEPre,
                                                    &scal->dt, tl, nullptr, 0, nullptr, nullptr, comp, stageOutlet());
                        else
                            launchStageFused<double>(stream, st, L, P, W, (const double*)in, (const double*)U,
                                                     (double*)out, (double*)A, cls, (const double*)rhoPre,
                                                     (const double*)rhoEPre, &scal->dt, tl, nullptr, 0, nullptr, nullptr,
                                                     nullptr, stageOutlet());
                    }
            };
            run(1);  // warm-up
            hipCheck(hipEventRecord(e0, stream), "record");
            run(reps);
            hipCheck(hipEventRecord(e1, stream), "record");
            hipCheck(hipEventSynchronize(e1), "sync");
            float ms = 0;
            hipCheck(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            double per = ms / reps;
            tuneTimes.push_back(per);
            if (per < best) { best = per; bestTile = tl; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hipCheck(hipGetLastError(), "autotune");
    tile = bestTile;
    return best;
}

void Case::mark(int group, bool begin) {
    if (!profiling) return;
    if (begin) {
        EventPair e{};
        hipCheck(hipEventCreate(&e.a), "event");
        hipCheck(hipEventCreate(&e.b), "event");
        e.group = group;
        hipCheck(hipEventRecord(e.a, stream), "record");
        events.push_back(e);
    } else {
        for (auto it = events.rbegin(); it != events.rend(); ++it)
            if (it->group == group) { hipCheck(hipEventRecord(it->b, stream), "record"); break; }
    }
}

void Case::profile(int nsteps, double ms[4], int launches[4]) {
    if (!initialized) throw StateError("profile before init");
    checkSchedule(nsteps);
    syncStoredPrims();
    double sum[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    auto collect = [&](bool keep) {
        for (auto& e : events) {
            float t = 0;
            if (keep && hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) {
                sum[e.group] += t;
                cnt[e.group] += 1;
            }
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        events.clear();
    };
    events.clear();
    // batches of at most histCap steps, each followed by a drain of the norm and force rings (readScalars),
    // as step() does, so no ring entry is overwritten before the host has it
    for (int done = 0; done < nsteps;) {
        const int batch = std::min(nsteps - done, histCap);
        profiling = true;
        try {
            for (int q = 0; q < batch; ++q) {
                if (fp32) stepOnce<float>();
                else stepOnce<double>();
            }
        } catch (...) {
            profiling = false;
            (void)hipStreamSynchronize(stream);
            collect(false);
            throw;
        }
        profiling = false;
        hipCheck(hipStreamSynchronize(stream), "sync");
        hipCheck(hipGetLastError(), "profiled steps");
        collect(true);
        readScalars();
        done += batch;
    }
    for (int g = 0; g < 4; ++g) {
        ms[g] = cnt[g] ? sum[g] / cnt[g] : 0.0;
        launches[g] = nsteps ? cnt[g] / nsteps : 0;
    }
}

void Case::init() {
    setupBoundaries();
    categorizeNodes();
    if (fp32) { allocFields<float>(); std::vector<float> h; hostInitialField(h); uploadState(h); }
    else { allocFields<double>(); std::vector<double> h; hostInitialField(h); uploadState(h); }
    readScalars();
    initialized = true;
}

void Case::setInitialField(const double* U5) {
    auto fill = [&](auto& h) {
        using T = typename std::decay_t<decltype(h)>::value_type;
        h.assign(5 * L.fstride, (T)0);
        size_t N = L.nodes();
        for (int q = 0; q < 5; ++q)
            for (int k = 0; k < L.NK; ++k)
                for (int j = 0; j < L.NJ; ++j)
                    for (int i = 0; i < L.NI; ++i)
                        h[L.off(i, j, k) + q * L.fstride] = (T)U5[q * N + ((size_t)k * L.NJ + j) * L.NI + i];
    };
    if (fp32) { std::vector<float> h; fill(h); uploadState(h); }
    else { std::vector<double> h; fill(h); uploadState(h); }
    readScalars();
}

void Case::setDtSchedule(const double* dts, int64_t n) {
    if (!initialized) throw StateError("dt schedule before init");
    dropGraphs();
    if (!dts || n <= 0) {
        dtSchedule = nullptr;
        dtScheduleLen = 0;
        return;
    }
    if ((int64_t)level >= n) throw std::invalid_argument("dt schedule shorter than the current time level");
    if (n > dtScheduleCap) {
        dtScheduleBuf = (double*)alloc((size_t)n * sizeof(double));
        dtScheduleCap = n;
    }
    hipCheck(hipMemcpy(dtScheduleBuf, dts, (size_t)n * sizeof(double), hipMemcpyHostToDevice), "copy dt schedule");
    dtSchedule = dtScheduleBuf;
    dtScheduleLen = n;
    // the next step already uses the schedule
    hipCheck(hipMemcpy(&scal->dt, dts + level, sizeof(double), hipMemcpyHostToDevice), "copy dt");
    readScalars();
}

void Case::setInflowProfile(const double* q, int64_t n) {
    if (!initialized) throw StateError("inflow profile before init");
    dropGraphs();
    int axis = -1;
    for (const EdgeFace& f : faces)
        if (f.type == NS_BC_INLET) {
            if (axis >= 0 && axis != f.axis) throw std::invalid_argument("inlets on different axes");
            axis = f.axis;
        }
    if (axis < 0) throw std::invalid_argument("no inlet boundary");
    const int64_t need = axis == 0 ? (int64_t)L.NJ * L.NK : (axis == 1 ? (int64_t)L.NI * L.NK : (int64_t)L.NI * L.NJ);
    if (!q) {
        P.inflow = nullptr;
        return;
    }
    if (n != need) throw std::invalid_argument("inflow profile size must equal the inlet plane's node count");
    if (!inflowBuf) inflowBuf = (double*)alloc((size_t)need * sizeof(double));
    hipCheck(hipMemcpy(inflowBuf, q, (size_t)need * sizeof(double), hipMemcpyHostToDevice), "copy inflow profile");
    P.inflow = inflowBuf;
}

void Case::readScalars() {
    StepScalars s;
    hipCheck(hipMemcpyAsync(&s, scal, sizeof(s), hipMemcpyDeviceToHost, stream), "copy scalars");
    hipCheck(hipStreamSynchronize(stream), "sync");
    t = s.t;
    dt = s.dt;
    level = s.level;
    nonFinite = s.nonFinite != 0;
    // drain the norm-history ring
    uint64_t have = normsHost[0].size();
    if (s.histCount > have) {
        std::vector<double> h((size_t)histCap * 5);
        hipCheck(hipMemcpy(h.data(), hist, h.size() * sizeof(double), hipMemcpyDeviceToHost), "copy hist");
        for (uint64_t q = have; q < s.histCount; ++q) {
            int slot = (int)(q % (uint64_t)histCap);
            for (int v = 0; v < 5; ++v) normsHost[v].push_back(h[5 * slot + v]);
        }
    }
    uint64_t haveF = forceHost.size() / 4;
    if (s.forceCount > haveF) {
        std::vector<double> h((size_t)histCap * 4);
        hipCheck(hipMemcpy(h.data(), forceRing, h.size() * sizeof(double), hipMemcpyDeviceToHost), "copy forces");
        for (uint64_t q = haveF; q < s.forceCount; ++q) {
            int slot = (int)(q % (uint64_t)histCap);
            for (int v = 0; v < 4; ++v) forceHost.push_back(h[4 * slot + v]);
        }
    }
}

void Case::dropGraphs() {
    if (graphExec[0] || graphExec[1]) hipCheck(hipStreamSynchronize(stream), "sync before dropping graphs");
    for (int q = 0; q < 2; ++q) {
        if (graphExec[q]) (void)hipGraphExecDestroy(graphExec[q]);
        graphExec[q] = nullptr;
        graphState[q] = nullptr;
    }
}

void Case::checkSchedule(int64_t nsteps) const {
    if (dtSchedule && (int64_t)level + nsteps >= dtScheduleLen)
        throw std::invalid_argument("dt schedule too short for the requested steps");
}

void Case::step(int nsteps) {
    if (!initialized) throw StateError("step before init");
    checkSchedule(nsteps);
    syncStoredPrims();
    int done = 0;
    while (done < nsteps) {
        int batch = std::min(nsteps - done, histCap);
        for (int q = 0; q < batch; ++q) {
            if (useGraph) {
                // the step reads dt and t from device memory and never synchronises, so one captured
                // step replays as is; the U/A swap alternates the buffers, hence one graph per parity
                int slot = graphState[0] == U ? 0 : (graphState[1] == U ? 1 : -1);
                if (slot < 0) {
                    slot = graphExec[0] ? 1 : 0;
                    void* start = U;
                    hipGraph_t g;
                    hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
                    if (fp32) stepOnce<float>();
                    else stepOnce<double>();
                    hipCheck(hipStreamEndCapture(stream, &g), "end capture");
                    std::swap(U, A);  // capture recorded the launches only; undo the host-side swap
                    hipCheck(hipGraphInstantiate(&graphExec[slot], g, nullptr, nullptr, 0), "instantiate");
                    (void)hipGraphDestroy(g);
                    graphState[slot] = start;
                }
                hipCheck(hipGraphLaunch(graphExec[slot], stream), "graph launch");
                std::swap(U, A);
            } else if (fp32) {
                stepOnce<float>();
            } else {
                stepOnce<double>();
            }
        }
        hipCheck(hipGetLastError(), "step kernels");
        readScalars();
        done += batch;
    }
}

void Case::stepTaped(StageTape* tape) {
    if (!initialized) throw StateError("step before init");
    checkSchedule(1);
    syncStoredPrims();
    tapeNow = tape;
    try {
        if (fp32) stepOnce<float>();  // swaps U and A itself
        else stepOnce<double>();
    } catch (...) {
        tapeNow = nullptr;
        throw;
    }
    tapeNow = nullptr;
    hipCheck(hipGetLastError(), "taped step");
    readScalars();
}

bool Case::filterDisabled() const {
    for (int q = 0; q < cfg.n_filter_intervals; ++q)
        if (cfg.filter_intervals[q] > 0) return false;
    return true;
}

void Case::getField(int which, double* out) {
    if (!initialized) throw StateError("get_field before init");
    size_t N = L.nodes();
    auto copyOut = [&](const void* base, int q) {
        if (fp32) {
            std::vector<float> h(L.fstride);
            hipCheck(hipMemcpy(h.data(), (const float*)base + (size_t)q * L.fstride, L.fstride * sizeof(float),
                               hipMemcpyDeviceToHost), "copy field");
            for (int k = 0; k < L.NK; ++k)
                for (int j = 0; j < L.NJ; ++j)
                    for (int i = 0; i < L.NI; ++i) out[((size_t)k * L.NJ + j) * L.NI + i] = h[L.off(i, j, k)];
        } else {
            std::vector<double> h(L.fstride);
            hipCheck(hipMemcpy(h.data(), (const double*)base + (size_t)q * L.fstride, L.fstride * sizeof(double),
                               hipMemcpyDeviceToHost), "copy field");
            for (int k = 0; k < L.NK; ++k)
                for (int j = 0; j < L.NJ; ++j)
                    for (int i = 0; i < L.NI; ++i) out[((size_t)k * L.NJ + j) * L.NI + i] = h[L.off(i, j, k)];
        }
        (void)N;
    };
    hipCheck(hipStreamSynchronize(stream), "sync");
    if (which < 5) { copyOut(U, which); return; }
    if (literal() && primsStored) {  // the stored fields are the reference's values
        copyOut(prim, which - 5);
        return;
    }
    primsStored = false;  // prim is rewritten below
    // primitive and transport values as the reference keeps them: derived from the state, pre-filter
    // values at filter nodes, initial values at untouched solid nodes
    if (fp32) {
        launchDerive<float>(stream, L, P, (const float*)U, (float*)prim, cls, W, (const float*)rhoPre, (const float*)rhoEPre);
        if (nInactive > 0)
            hipLaunchKernelGGL(k_patch_inactive<float>, dim3(nblk(nInactive, 128)), dim3(128), 0, stream, L, P, (float*)prim,
                               inactiveList, inactiveIC, nInactive);
    } else {
        launchDerive<double>(stream, L, P, (const double*)U, (double*)prim, cls, W, (const double*)rhoPre,
                             (const double*)rhoEPre);
        if (nInactive > 0)
            hipLaunchKernelGGL(k_patch_inactive<double>, dim3(nblk(nInactive, 128)), dim3(128), 0, stream, L, P,
                               (double*)prim, inactiveList, inactiveIC, nInactive);
    }
    hipCheck(hipStreamSynchronize(stream), "sync");
    copyOut(prim, which - 5);
}

// ------------------------------------------------------------------------------------------------ checkpoints

namespace {

constexpr uint64_t kCkptMagic = 0x31544b5049534e4eULL;  // "NNSIPKT1"
constexpr uint32_t kCkptVersion = 3;

// Configuration fields a checkpoint must agree on: the grid, boundaries, body, initial-state generator
// and physics (stopping, output and stability-limit settings may change between runs).
struct CkptKey {
    int32_t NI, NJ, NK, precision, kahan, bc[6], body_type, body_axis;
    double body_centre[3], body_radius, gamma, Pr, Re, sutherlands_C2, M_0, T_0, L_x, L_y, L_z, inlet_ramp_time;
    int32_t n_extra_bodies, pad;
    ns_body extra_bodies[NS_MAX_BODIES - 1];
};
CkptKey ckptKey(const ns_config& c) {
    CkptKey k;
    std::memset(&k, 0, sizeof(k));
    k.NI = c.NI; k.NJ = c.NJ; k.NK = c.NK; k.precision = c.precision; k.kahan = c.precision == NS_FP32 ? c.kahan : 0;
    for (int q = 0; q < 6; ++q) k.bc[q] = c.bc[q];
    k.body_type = c.body_type;
    k.body_axis = c.body_type == NS_BODY_CYLINDER ? c.body_axis : 0;
    for (int q = 0; q < 3; ++q) k.body_centre[q] = c.body_centre[q];
    k.body_radius = c.body_radius; k.gamma = c.gamma; k.Pr = c.Pr; k.Re = c.Re; k.sutherlands_C2 = c.sutherlands_C2;
    k.M_0 = c.M_0; k.T_0 = c.T_0; k.L_x = c.L_x; k.L_y = c.L_y; k.L_z = c.L_z; k.inlet_ramp_time = c.inlet_ramp_time;
    k.n_extra_bodies = c.n_extra_bodies;
    for (int q = 0; q < c.n_extra_bodies && q < NS_MAX_BODIES - 1; ++q) {
        k.extra_bodies[q] = c.extra_bodies[q];
        if (c.extra_bodies[q].type != NS_BODY_CYLINDER) k.extra_bodies[q].axis = 0;
    }
    return k;
}

struct CkptHeader {
    uint64_t magic;
    uint32_t version, primsStored;
    CkptKey key;
    uint64_t fstride, windowSize, nInactive, nNorms, nForces, nFixed;
    int32_t conv_history, pad;
};

}  // namespace

void Case::save(const std::string& path) {
    if (!initialized) throw StateError("save before init");
    readScalars();  // drains the history rings into normsHost / forceHost
    hipCheck(hipStreamSynchronize(stream), "sync");
    CkptHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    hd.magic = kCkptMagic;
    hd.version = kCkptVersion;
    hd.primsStored = primsStored ? 1 : 0;
    hd.key = ckptKey(cfg);
    hd.fstride = L.fstride;
    hd.windowSize = W.size();
    hd.nInactive = (uint64_t)nInactive;
    hd.nFixed = (uint64_t)nFixed;
    hd.nNorms = normsHost[0].size();
    hd.nForces = forceHost.size() / 4;
    hd.conv_history = cfg.conv_history;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::invalid_argument("cannot open checkpoint file for writing: " + path);
    auto put = [&](const void* p, size_t n) {
        f.write((const char*)p, (std::streamsize)n);
        if (!f) throw std::runtime_error("checkpoint write failed: " + path);
    };
    std::vector<char> host;
    auto putDevice = [&](const void* d, size_t n) {
        host.resize(std::max<size_t>(n, 1));
        if (n) hipCheck(hipMemcpy(host.data(), d, n, hipMemcpyDeviceToHost), "checkpoint copy");
        put(host.data(), n);
    };
    put(&hd, sizeof(hd));
    const size_t fb = 5 * L.fstride * eb, wb = std::max<size_t>(1, W.size()) * eb;
    putDevice(U, fb);
    putDevice(A, fb);
    if (!fp32) putDevice(prim, 7 * L.fstride * eb);
    putDevice(rhoPre, wb);
    putDevice(rhoEPre, wb);
    if (comp) putDevice(comp, 5 * L.fstride * sizeof(int8_t));
    putDevice(inactiveIC, 5 * std::max<int64_t>(1, nInactive) * sizeof(double));
    if (nFixed) putDevice(fixedVals, 5 * nFixed * sizeof(double));
    StepScalars st;
    hipCheck(hipMemcpy(&st, scal, sizeof(st), hipMemcpyDeviceToHost), "checkpoint scalars");
    put(&st, sizeof(st));
    for (int v = 0; v < 5; ++v) put(normsHost[v].data(), normsHost[v].size() * sizeof(double));
    put(forceHost.data(), forceHost.size() * sizeof(double));
}

void Case::restore(const std::string& path) {
    if (!initialized) throw StateError("restore before init");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::invalid_argument("cannot open checkpoint file: " + path);
    auto get = [&](void* p, size_t n) {
        f.read((char*)p, (std::streamsize)n);
        if (!f) throw std::invalid_argument("checkpoint file is truncated: " + path);
    };
    CkptHeader hd;
    get(&hd, sizeof(hd));
    if (hd.magic != kCkptMagic || hd.version != kCkptVersion) throw std::invalid_argument("not a checkpoint file: " + path);
    const CkptKey want = ckptKey(cfg);
    if (std::memcmp(&hd.key, &want, sizeof(want)) != 0)
        throw std::invalid_argument("checkpoint was written by a case with another grid, boundaries, body or physics");
    if (hd.fstride != L.fstride || hd.windowSize != W.size() || hd.nInactive != (uint64_t)nInactive ||
        hd.nFixed != (uint64_t)nFixed || hd.conv_history != cfg.conv_history)
        throw std::invalid_argument("checkpoint layout does not match this case");
    dropGraphs();
    std::vector<char> host;
    auto getDevice = [&](void* d, size_t n) {
        host.resize(std::max<size_t>(n, 1));
        get(host.data(), n);
        if (n) hipCheck(hipMemcpy(d, host.data(), n, hipMemcpyHostToDevice), "restore copy");
    };
    const size_t fb = 5 * L.fstride * eb, wb = std::max<size_t>(1, W.size()) * eb;
    getDevice(U, fb);
    getDevice(A, fb);
    if (!fp32) getDevice(prim, 7 * L.fstride * eb);
    getDevice(rhoPre, wb);
    getDevice(rhoEPre, wb);
    if (comp) getDevice(comp, 5 * L.fstride * sizeof(int8_t));
    getDevice(inactiveIC, 5 * std::max<int64_t>(1, nInactive) * sizeof(double));
    if (nFixed) getDevice(fixedVals, 5 * nFixed * sizeof(double));
    StepScalars st;
    get(&st, sizeof(st));
    // the device rings restart empty; the histories so far are kept on the host
    std::vector<double> norms[5], forces(hd.nForces * 4);
    for (int v = 0; v < 5; ++v) {
        norms[v].resize(hd.nNorms);
        get(norms[v].data(), hd.nNorms * sizeof(double));
    }
    get(forces.data(), forces.size() * sizeof(double));
    char extra;
    if (f.read(&extra, 1)) throw std::invalid_argument("checkpoint file has trailing data: " + path);
    hipCheck(hipMemcpy(scal, &st, sizeof(st), hipMemcpyHostToDevice), "restore scalars");
    for (int v = 0; v < 5; ++v) normsHost[v] = norms[v];
    forceHost = forces;
    primsStored = hd.primsStored != 0;
    readScalars();
}

void Case::nodeTypes(uint8_t* out) {
    if (!initialized) throw StateError("node types before init");
    std::vector<uint8_t> h((size_t)L.NK * L.plane);
    hipCheck(hipMemcpy(h.data(), cls, h.size(), hipMemcpyDeviceToHost), "copy cls");
    for (int k = 0; k < L.NK; ++k)
        for (int j = 0; j < L.NJ; ++j)
            for (int i = 0; i < L.NI; ++i) out[((size_t)k * L.NJ + j) * L.NI + i] = h[L.off(i, j, k)] & CLS_TYPE_MASK;
}

static std::vector<int64_t> sortedByReference(const std::vector<int64_t>& nodes, int NI, int NJ, int NK,
                                              std::vector<size_t>& order) {
    // reference index (i*NJ + j)*NK + k
    std::vector<int64_t> key(nodes.size());
    for (size_t q = 0; q < nodes.size(); ++q) {
        int64_t n = nodes[q];
        int64_t i = n % NI, r = n / NI, j = r % NJ, k = r / NJ;
        key[q] = (i * NJ + j) * NK + k;
    }
    order.resize(nodes.size());
    for (size_t q = 0; q < order.size(); ++q) order[q] = q;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
    return key;
}

void Case::ghosts(int32_t* ijk, double* bi, double* ip, int32_t* box) {
    if (!initialized) throw StateError("ghosts before init");
    std::vector<int64_t> nodes(nGhost);
    std::vector<double> hb(3 * nGhost), hp(3 * nGhost);
    std::vector<int32_t> hbox(3 * nGhost);
    if (nGhost) {
        hipCheck(hipMemcpy(nodes.data(), ghostList, nGhost * sizeof(int64_t), hipMemcpyDeviceToHost), "copy ghosts");
        hipCheck(hipMemcpy(hb.data(), G.bi, hb.size() * sizeof(double), hipMemcpyDeviceToHost), "copy bi");
        hipCheck(hipMemcpy(hp.data(), G.ip, hp.size() * sizeof(double), hipMemcpyDeviceToHost), "copy ip");
        hipCheck(hipMemcpy(hbox.data(), G.boxLo, hbox.size() * sizeof(int32_t), hipMemcpyDeviceToHost), "copy box");
    }
    std::vector<size_t> order;
    sortedByReference(nodes, L.NI, L.NJ, L.NK, order);
    for (size_t q = 0; q < order.size(); ++q) {
        size_t g = order[q];
        int64_t n = nodes[g];
        if (ijk) { ijk[3 * q] = (int32_t)(n % L.NI); ijk[3 * q + 1] = (int32_t)((n / L.NI) % L.NJ); ijk[3 * q + 2] = (int32_t)(n / L.NI / L.NJ); }
        for (int c = 0; c < 3; ++c) {
            if (bi) bi[3 * q + c] = hb[3 * g + c];
            if (ip) ip[3 * q + c] = hp[3 * g + c];
            if (box) box[3 * q + c] = hbox[3 * g + c];
        }
    }
}

void Case::filterNodes(int32_t* ijk) {
    if (!initialized) throw StateError("filter nodes before init");
    std::vector<int64_t> nodes(nFilter);
    if (nFilter) hipCheck(hipMemcpy(nodes.data(), filterList, nFilter * sizeof(int64_t), hipMemcpyDeviceToHost), "copy filter");
    std::vector<size_t> order;
    sortedByReference(nodes, L.NI, L.NJ, L.NK, order);
    for (size_t q = 0; q < order.size(); ++q) {
        int64_t n = nodes[order[q]];
        ijk[3 * q] = (int32_t)(n % L.NI);
        ijk[3 * q + 1] = (int32_t)((n / L.NI) % L.NJ);
        ijk[3 * q + 2] = (int32_t)(n / L.NI / L.NJ);
    }
}

int64_t Case::interiorCount() {
    std::vector<uint8_t> h((size_t)L.NK * L.plane);
    hipCheck(hipMemcpy(h.data(), cls, h.size(), hipMemcpyDeviceToHost), "copy cls");
    int64_t n = 0;
    for (int k = 0; k < L.NK; ++k)
        for (int j = 0; j < L.NJ; ++j)
            for (int i = 0; i < L.NI; ++i) n += (h[L.off(i, j, k)] & CLS_TYPE_MASK) == NS_FLUID_INTERIOR;
    return n;
}

void Case::totals(double out[5]) {
    const int nb = kTotalsBlocks;
    double* part = partials;  // reuse (1024 x 10 doubles >= 256 x 5)
    hipCheck(hipStreamSynchronize(stream), "sync");
    if (fp32)
        hipLaunchKernelGGL(k_totals<float>, dim3(nb), dim3(kTotalsThreads), 0, stream, L, (const float*)U, cls, part);
    else
        hipLaunchKernelGGL(k_totals<double>, dim3(nb), dim3(kTotalsThreads), 0, stream, L, (const double*)U, cls, part);
    std::vector<double> h((size_t)nb * 5);
    hipCheck(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "copy totals");
    hipCheck(hipStreamSynchronize(stream), "sync");
    double vol = P.hx * P.hy * P.hz;
    for (int q = 0; q < 5; ++q) {
        double s = 0;
        for (int b = 0; b < nb; ++b) s += h[(size_t)b * 5 + q];
        out[q] = s * vol;
    }
}

}  // namespace nsfsi
