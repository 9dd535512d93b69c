/*
 * mesh.hip — node classification, ghost-node geometry and the per-stage boundary conditions.
 *
 * Everything in this file is compiled without FMA contraction (-ffp-contract=off on its command line and
 * the pragma below), with IEEE division and square root, so that node classes, ghost sets, image points
 * and the 8x8 factorisations come out bit-identical to a serial FP64 evaluation of the same expressions
 * (the reference's g++ -O3 build emits no FMA on x86-64).
 *
 * Reference counterparts: Mesh::categorizeNodes (src/Mesh.cpp:61-82),
 * MeshEdgeBoundary::identifyOwnedNodes (src/Boundary.cpp:18-71), the body classes
 * (src/Boundary.cpp:461-615, 887-1103), the edge BCs (src/Boundary.cpp:142-417) and
 * ImmersedBoundary::applyBoundaryCondition (src/Boundary.cpp:427-457, 618-883).
 */
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.h"
#include "batch.h"
#include "kernels.h"

namespace nsfsi {

namespace {

__device__ __forceinline__ void nodePos(const Params& P, int i, int j, int k, double& x, double& y, double& z) {
    // getNodePosition, src/SmallVectors.h:56-63: (int index - offset) * spacing
    x = ((double)i - P.ox) * P.hx;
    y = ((double)j - P.oy) * P.hy;
    z = ((double)k - P.oz) * P.hz;
}

// getSurroundingNodesBox, src/SmallVectors.h:108-127. Returns false if the box leaves the mesh.
__device__ __forceinline__ bool surroundingBox(const Layout& L, const Params& P, double px, double py, double pz, int& i0,
                                               int& j0, int& k0) {
    double vx = ceil(px / P.hx) + P.ox, vy = ceil(py / P.hy) + P.oy, vz = ceil(pz / P.hz) + P.oz;
    long in = vx >= 1.0 ? (long)vx : 1, jn = vy >= 1.0 ? (long)vy : 1, kn = vz >= 1.0 ? (long)vz : 1;
    if (in > L.NI) in = L.NI;
    if (jn > L.NJ) jn = L.NJ;
    if (kn > L.NK) kn = L.NK;
    i0 = (int)in - 1;
    j0 = (int)jn - 1;
    k0 = (int)kn - 1;
    return in <= L.NI - 1 && jn <= L.NJ - 1 && kn <= L.NK - 1;
}

__device__ __forceinline__ double len3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// CylinderBody::getNormalProbe / SphereBody::getNormalProbe, src/Boundary.cpp:920-935, :1047-1053
__device__ __forceinline__ void normalProbe(const BodyDesc& B, double gx, double gy, double gz, double& px, double& py,
                                            double& pz) {
    double cx = gx - B.cx, cy = gy - B.cy, cz = gz - B.cz;
    if (B.type == NS_BODY_CYLINDER) {
        if (B.axis == 0) cx = 0;
        else if (B.axis == 1) cy = 0;
        else cz = 0;
    }
    double l = len3(cx, cy, cz);
    double lf = (B.r - l) / l;
    px = cx * lf;
    py = cy * lf;
    pz = cz * lf;
}

__global__ void k_classify(Layout L, uint8_t* cls, int t0, int t1, int t2, int t3, int t4, int t5, BodySet bodies,
                           Params P) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    int k = blockIdx.z;
    if (i >= L.NI) return;
    // Edge claims in push order z-min, z-max, y-min, y-max, x-min, x-max (src/Mesh.cpp:33-38, :67-69)
    int t = NS_FLUID_INTERIOR;
    if (k == 0) t = t4;
    else if (k == L.NK - 1) t = t5;
    else if (j == 0) t = t2;
    else if (j == L.NJ - 1) t = t3;
    else if (i == 0) t = t0;
    else if (i == L.NI - 1) t = t1;
    uint8_t flags = 0;
    // the body whose (disjoint) box holds the node; the others never look at it
    for (int b = 0; b < bodies.n; ++b) {
        const BodyDesc& B = bodies.b[b];
        if (!B.box.contains(i, j, k)) continue;
        double x, y, z;
        nodePos(P, i, j, k, x, y, z);
        double dx = x - B.cx, dy = y - B.cy, dz = z - B.cz, d;
        if (B.type == NS_BODY_SPHERE) d = sqrt(dx * dx + dy * dy + dz * dz);
        else if (B.axis == 0) d = sqrt(dy * dy + dz * dz);
        else if (B.axis == 1) d = sqrt(dx * dx + dz * dz);
        else d = sqrt(dx * dx + dy * dy);
        // getSolidAndFilterNodesIn{Cylinder,Sphere}, src/Boundary.cpp:1005-1012, :1094-1101
        if (d < B.r - B.mpb) t = NS_SOLID_INACTIVE;
        else if (d < B.r + B.filterWidth && t == NS_FLUID_INTERIOR) flags |= CLS_FILTER;
        break;
    }
    cls[L.off(i, j, k)] = (uint8_t)(t | flags);
}

// findGhostNodesWithFluidNeighbors, src/Boundary.cpp:461-562: a solid node with a FluidInterior node
// among its 18 face/edge neighbours becomes a ghost.
__global__ void k_mark_ghosts(Layout L, uint8_t* cls, Window W) {
    int i = W.i0 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = W.j0 + blockIdx.y;
    int k = W.k0 + blockIdx.z;
    if (i >= W.i0 + W.ni) return;
    size_t o = L.off(i, j, k);
    if ((cls[o] & CLS_TYPE_MASK) != NS_SOLID_INACTIVE) return;
    bool has = false;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                int nz = (di != 0) + (dj != 0) + (dk != 0);
                if (nz == 0 || nz == 3) continue;
                int ii = i + di, jj = j + dj, kk = k + dk;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= L.NI || jj >= L.NJ || kk >= L.NK) continue;
                if ((cls[L.off(ii, jj, kk)] & CLS_TYPE_MASK) == NS_FLUID_INTERIOR) has = true;
            }
    if (has) cls[o] = (uint8_t)((cls[o] & ~CLS_TYPE_MASK) | NS_SOLID_GHOST);
}

__device__ __forceinline__ void nodeIJK(const Layout& L, int64_t n, int& i, int& j, int& k) {
    i = (int)(n % L.NI);
    int64_t r = n / L.NI;
    j = (int)(r % L.NJ);
    k = (int)(r / L.NJ);
}

__global__ void k_count(Layout L, const uint8_t* cls, Window W, uint8_t mask, uint8_t value, int32_t* counts) {
    __shared__ int32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < W.size()) {
        int i = W.i0 + (int)(w % W.ni);
        size_t r = w / W.ni;
        int j = W.j0 + (int)(r % W.nj);
        int k = W.k0 + (int)(r / W.nj);
        if ((cls[L.off(i, j, k)] & mask) == value) atomicAdd(&s, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

__global__ void k_scatter(Layout L, const uint8_t* cls, Window W, uint8_t mask, uint8_t value, const int32_t* offsets,
                          int64_t* out) {
    __shared__ int32_t flags[256];
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0;
    int64_t node = 0;
    if (w < W.size()) {
        int i = W.i0 + (int)(w % W.ni);
        size_t r = w / W.ni;
        int j = W.j0 + (int)(r % W.nj);
        int k = W.k0 + (int)(r / W.nj);
        f = (cls[L.off(i, j, k)] & mask) == value;
        node = ((int64_t)k * L.NJ + j) * L.NI + i;
    }
    flags[threadIdx.x] = f;
    __syncthreads();
    if (f) {
        int pos = 0;
        for (int q = 0; q < (int)threadIdx.x; ++q) pos += flags[q];
        out[offsets[blockIdx.x] + pos] = node;
    }
}

// setImagePointPositions + checkIfSurroundingShouldBeGhost, src/Boundary.cpp:568-602
__global__ void k_image_points(Layout L, uint8_t* cls, Params P, BodyDesc B, const int64_t* ghosts, int64_t n, double* bi,
                               double* ip, int32_t* changed, int32_t* error) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int i, j, k;
    nodeIJK(L, ghosts[g], i, j, k);
    double gx, gy, gz, px, py, pz;
    nodePos(P, i, j, k, gx, gy, gz);
    normalProbe(B, gx, gy, gz, px, py, pz);
    double bx = gx + px, by = gy + py, bz = gz + pz;
    double qx = bx + px, qy = by + py, qz = bz + pz;
    bi[3 * g] = bx; bi[3 * g + 1] = by; bi[3 * g + 2] = bz;
    ip[3 * g] = qx; ip[3 * g + 1] = qy; ip[3 * g + 2] = qz;
    int i0, j0, k0;
    if (!surroundingBox(L, P, qx, qy, qz, i0, j0, k0)) { atomicOr(error, 1); return; }
    for (int c = 0; c < 8; ++c) {
        int ii = i0 + ((c >> 2) & 1), jj = j0 + ((c >> 1) & 1), kk = k0 + (c & 1);
        size_t o = L.off(ii, jj, kk);
        uint8_t v = cls[o];
        if ((v & CLS_TYPE_MASK) == NS_SOLID_INACTIVE) {
            if (!B.box.contains(ii, jj, kk)) { atomicOr(error, 2); continue; }
            cls[o] = (uint8_t)((v & ~CLS_TYPE_MASK) | NS_SOLID_GHOST);
            atomicAdd(changed, 1);
        }
    }
}

__global__ void k_fill_gmap(Layout L, Window W, const int64_t* ghosts, int64_t n, int32_t* gmap) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int i, j, k;
    nodeIJK(L, ghosts[g], i, j, k);
    gmap[W.idx(i, j, k)] = (int32_t)g;
}

// Textbook Gaussian elimination with complete pivoting on an 8x8 row-major matrix (first maximum in a
// column-major scan of the trailing block). Same algorithm and operation order as the serial FP64 path.
__device__ void lu8Factor(double* A, int8_t* rp, int8_t* cp) {
    for (int q = 0; q < 8; ++q) { rp[q] = (int8_t)q; cp[q] = (int8_t)q; }
    for (int s = 0; s < 8; ++s) {
        int pr = s, pc = s;
        double best = -1;
        for (int c = s; c < 8; ++c)
            for (int r = s; r < 8; ++r) {
                double m = fabs(A[r * 8 + c]);
                if (m > best) { best = m; pr = r; pc = c; }
            }
        if (pr != s) {
            for (int c = 0; c < 8; ++c) { double t = A[s * 8 + c]; A[s * 8 + c] = A[pr * 8 + c]; A[pr * 8 + c] = t; }
            int8_t t = rp[s]; rp[s] = rp[pr]; rp[pr] = t;
        }
        if (pc != s) {
            for (int r = 0; r < 8; ++r) { double t = A[r * 8 + s]; A[r * 8 + s] = A[r * 8 + pc]; A[r * 8 + pc] = t; }
            int8_t t = cp[s]; cp[s] = cp[pc]; cp[pc] = t;
        }
        double piv = A[s * 8 + s];
        if (piv == 0) continue;
        for (int r = s + 1; r < 8; ++r) {
            double l = A[r * 8 + s] / piv;
            A[r * 8 + s] = l;
            for (int c = s + 1; c < 8; ++c) A[r * 8 + c] -= l * A[s * 8 + c];
        }
    }
}

// Weights w with w . f = b . c for the solution c of V c = f (luSolveDot below): w = P^T L^-T U^-T Q^T b,
// the transposed substitutions in reverse order, with the same treatment of a zero pivot (that component of
// the solution is 0). The interpolated value is then one dot product with the 8 corner values.
__device__ void luWeights(const double* LU, const int8_t* rp, const int8_t* cp, const double* b, double* w) {
    double t[8], s[8], v[8];
    for (int r = 0; r < 8; ++r) t[r] = b[cp[r]];
    for (int r = 0; r < 8; ++r) {  // U^T s = t (forward)
        double acc = t[r];
        for (int q = 0; q < r; ++q) acc -= LU[q * 8 + r] * s[q];
        s[r] = LU[r * 8 + r] != 0 ? acc / LU[r * 8 + r] : 0.0;
    }
    for (int r = 7; r >= 0; --r) {  // L^T v = s (backward, unit diagonal)
        double acc = s[r];
        for (int q = r + 1; q < 8; ++q) acc -= LU[q * 8 + r] * v[q];
        v[r] = acc;
    }
    for (int r = 0; r < 8; ++r) w[rp[r]] = v[r];
}

// Image-point geometry, corner tables and both Vandermonde factorisations for every ghost
// (setInterpolationValues + populateVandermonde*, src/Boundary.cpp:655-728, :788-817).
__global__ void k_ghost_geometry(Layout L, const uint8_t* cls, Params P, Window W, const int64_t* ghosts, int64_t base,
                                 int64_t n, const int32_t* gmap, GhostTables G, int32_t* error) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int i, j, k;
    nodeIJK(L, ghosts[g], i, j, k);
    G.node[g] = (int64_t)L.off(i, j, k);
    double qx = G.ip[3 * g], qy = G.ip[3 * g + 1], qz = G.ip[3 * g + 2];
    int i0, j0, k0;
    if (!surroundingBox(L, P, qx, qy, qz, i0, j0, k0)) { atomicOr(error, 1); return; }
    G.boxLo[3 * g] = i0; G.boxLo[3 * g + 1] = j0; G.boxLo[3 * g + 2] = k0;
    double X[8], Y[8], Z[8], nx[8], ny[8], nz[8];
    uint8_t mask = 0;
    for (int c = 0; c < 8; ++c) {
        int ii = i0 + ((c >> 2) & 1), jj = j0 + ((c >> 1) & 1), kk = k0 + (c & 1);
        size_t o = L.off(ii, jj, kk);
        G.corner[8 * g + c] = (int64_t)o;
        int t = cls[o] & CLS_TYPE_MASK;
        nx[c] = ny[c] = nz[c] = 0;
        if (t == NS_FLUID_INTERIOR || t == NS_FLUID_EDGE || t == NS_FLUID_GHOST) {
            nodePos(P, ii, jj, kk, X[c], Y[c], Z[c]);
        } else if (t == NS_SOLID_GHOST) {
            mask |= (uint8_t)(1u << c);
            // ghostNodeMap.at (src/Boundary.cpp:681): only this body's ghosts
            int64_t og = W.contains(ii, jj, kk) ? (int64_t)gmap[W.idx(ii, jj, kk)] - base : -1;
            if (og < 0 || og >= n) { atomicOr(error, 2); return; }
            double bx = G.bi[3 * og], by = G.bi[3 * og + 1], bz = G.bi[3 * og + 2];
            double vx = G.ip[3 * og] - bx, vy = G.ip[3 * og + 1] - by, vz = G.ip[3 * og + 2] - bz;
            double l = len3(vx, vy, vz);
            nx[c] = vx / l; ny[c] = vy / l; nz[c] = vz / l;
            X[c] = bx; Y[c] = by; Z[c] = bz;
        } else {
            atomicOr(error, 4);  // "Impossible situation. Found solid node around image point."
            return;
        }
    }
    G.mask[g] = mask;
    double lx, ly, lz;
    nodePos(P, i0, j0, k0, lx, ly, lz);
    G.rel[3 * g] = qx - lx; G.rel[3 * g + 1] = qy - ly; G.rel[3 * g + 2] = qz - lz;
    if (mask == 0) return;
    double* D = G.luD + 64 * g;
    double* N = G.luN + 64 * g;
    for (int r = 0; r < 8; ++r) {
        double row[8] = {1, X[r], Y[r], Z[r], X[r] * Y[r], X[r] * Z[r], Y[r] * Z[r], X[r] * Y[r] * Z[r]};
        for (int q = 0; q < 8; ++q) D[r * 8 + q] = N[r * 8 + q] = row[q];
        if (mask & (1u << r)) {
            N[r * 8 + 0] = 0;
            N[r * 8 + 1] = nx[r];
            N[r * 8 + 2] = ny[r];
            N[r * 8 + 3] = nz[r];
            N[r * 8 + 4] = X[r] * ny[r] + Y[r] * nx[r];
            N[r * 8 + 5] = X[r] * nz[r] + Z[r] * nx[r];
            N[r * 8 + 6] = Y[r] * nz[r] + Z[r] * ny[r];
            N[r * 8 + 7] = X[r] * Y[r] * nz[r] + X[r] * Z[r] * ny[r] + Y[r] * Z[r] * nx[r];
        }
    }
    int8_t* pm = G.perm + 32 * g;
    lu8Factor(D, pm, pm + 8);
    lu8Factor(N, pm + 16, pm + 24);
    const double b[8] = {1, qx, qy, qz, qx * qy, qx * qz, qy * qz, qx * qy * qz};
    luWeights(D, pm, pm + 8, b, G.wD + 8 * g);
    luWeights(N, pm + 16, pm + 24, b, G.wN + 8 * g);
}

// ------------------------------------------------------------------------------------------------ BCs

template <class T>
struct Fields {
    T* f[5];
};
template <class T>
__device__ __forceinline__ Fields<T> fields(const Layout& L, T* S) {
    Fields<T> F;
    for (int q = 0; q < 5; ++q) F.f[q] = S + (size_t)q * L.fstride;
    return F;
}

// One edge face, all owned nodes (src/Boundary.cpp:250-417). Each thread handles one node of the face;
// the inlet's in-plane density filter recomputes its neighbours' extrapolated densities itself, so the
// face needs no second pass.
// Node (a, b) of the face's owned box (a along the first in-plane axis, b along the second); ts is the
// time of the stage the boundary values belong to.
// Stored primitive and transport values of node o (7 fields: u, v, w, p', T', mu, kappa).
template <class T>
__device__ __forceinline__ void storePrims(const Layout& L, T* prim, size_t o, T u, T v, T w, T p, T Tt, T mu, T ka) {
    prim[o] = u;
    prim[o + L.fstride] = v;
    prim[o + 2 * L.fstride] = w;
    prim[o + 3 * L.fstride] = p;
    prim[o + 4 * L.fstride] = Tt;
    prim[o + 5 * L.fstride] = mu;
    prim[o + 6 * L.fstride] = ka;
}

// prim != null: the face also writes the primitive and transport values the reference stores with the node
// (setFlowVariablesAtNode, src/FlowVariableGroupStructs.h:189-207), and a symmetry face mirrors the stored
// primitives of the node two in instead of deriving them from its conserved values.
template <class T>
__device__ __forceinline__ void edgeFaceNode(const Layout& L, const Params& P, const EdgeFace& f, T* S, double ts, int a,
                                             int b, T* prim = nullptr) {
    // in-plane axes in the reference's loop order: x-normal (j,k), y-normal (i,k), z-normal (i,j)
    int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
    int lo[3] = {f.i0, f.j0, f.k0}, hi[3] = {f.i1, f.j1, f.k1};
    int c[3];
    c[f.axis] = lo[f.axis];
    c[ax1] = lo[ax1] + a;
    c[ax2] = lo[ax2] + b;
    if (c[ax1] > hi[ax1] || c[ax2] > hi[ax2]) return;
    Fields<T> F = fields(L, S);
    Eos<T> e = Eos<T>::make(P);
    size_t o = L.off(c[0], c[1], c[2]);
    // offsets of the nodes one and two in from the face (x-steps depend on the column: padding slots)
    const int dn = f.side == 0 ? 1 : -1;  // towards the interior
    const long st = f.axis == 1 ? (long)L.pitch : (long)L.plane;
    const long s1 = f.axis == 0 ? Layout::xstep(c[0], dn) : dn * st;
    const long s2 = f.axis == 0 ? Layout::xstep(c[0], 2 * dn) : 2 * dn * st;
    if (f.type == NS_BC_INLET) {
        auto extrap = [&](size_t oo) { return (T)2 * F.f[0][oo + s1] - F.f[0][oo + s2]; };
        // filterInletDensity, src/Boundary.cpp:142-245: neighbours inside the owned box, order +a1,-a1,+a2,-a2
        T acc = 0;
        unsigned nb = 0;
        for (int ax : {ax1, ax2}) {
            const long up = ax == 0 ? Layout::xstep(c[0], 1) : (ax == 1 ? (long)L.pitch : (long)L.plane);
            const long down = ax == 0 ? Layout::xstep(c[0], -1) : -up;
            if (c[ax] < hi[ax]) { acc += extrap(o + up); ++nb; }
            if (c[ax] > lo[ax]) { acc += extrap(o + down); ++nb; }
        }
        acc += (T)nb * extrap(o);
        acc /= (T)(2 * nb);
        double vn = fmin(1., ts / P.rampTime) * P.M_0;
        if (P.inflow) vn *= P.inflow[(size_t)c[ax2] * (ax1 == 0 ? L.NI : L.NJ) + c[ax1]];
        T vel = (T)vn;
        if (f.side == 1) vel *= (T)-1;
        // component selection by selects, not a three-way branch: gfx950 code generation has dropped the
        // last arm of such a branch on a kernel argument (the z-face case) in this function
        const T uu = f.axis == 0 ? vel : (T)0, vv = f.axis == 1 ? vel : (T)0, ww = f.axis == 2 ? vel : (T)0;
        T Tt = 0;
        T p = (Tt * ((T)1 + acc) + acc) / e.gamma;
        T rho, ru, rv, rw, rE;
        deriveCons(e, uu, vv, ww, p, Tt, rho, ru, rv, rw, rE);
        F.f[0][o] = rho; F.f[1][o] = ru; F.f[2][o] = rv; F.f[3][o] = rw; F.f[4][o] = rE;
        if (prim) {
            T mu, ka;
            deriveTrans(e, Tt, mu, ka);
            storePrims(L, prim, o, uu, vv, ww, p, Tt, mu, ka);
        }
    } else if (f.type == NS_BC_OUTLET) {
        T a[4], b[4], v[5];
        for (int q = 0; q < 4; ++q) {
            a[q] = F.f[q][o + s1];
            b[q] = F.f[q][o + s2];
        }
        outletValues(e.gm1, a, b, v);
        for (int q = 0; q < 5; ++q) F.f[q][o] = v[q];
        if (prim) {  // src/Boundary.cpp:318-331
            const T r = v[0], p = 0;
            const T u = v[1] / ((T)1 + r), vv = v[2] / ((T)1 + r), w = v[3] / ((T)1 + r);
            const T Tt = (e.gamma * p - r) / ((T)1 + r);
            T mu, ka;
            deriveTrans(e, Tt, mu, ka);
            storePrims(L, prim, o, u, vv, w, p, Tt, mu, ka);
        }
    } else if (f.type == NS_BC_PERIODIC) {
        int cc[3] = {c[0], c[1], c[2]};
        int Nn = f.axis == 0 ? L.NI : (f.axis == 1 ? L.NJ : L.NK);
        cc[f.axis] = f.side == 0 ? Nn - 2 : 1;
        size_t src = L.off(cc[0], cc[1], cc[2]);
        for (int q = 0; q < 5; ++q) F.f[q][o] = F.f[q][src];
        if (prim)
            for (int q = 0; q < 7; ++q) prim[o + q * L.fstride] = prim[src + q * L.fstride];
    } else {  // symmetry: mirror the primitives of the node two in, negate the normal velocity
        size_t a2 = o + s2;
        T u, v, w, p, Tt, mu;
        if (prim) {
            u = prim[a2];
            v = prim[a2 + L.fstride];
            w = prim[a2 + 2 * L.fstride];
            p = prim[a2 + 3 * L.fstride];
            Tt = prim[a2 + 4 * L.fstride];
        } else {
            derivePrims(e, F.f[0][a2], F.f[1][a2], F.f[2][a2], F.f[3][a2], F.f[4][a2], u, v, w, p, Tt, mu);
        }
        u = f.axis == 0 ? -u : u;  // selects, see the inlet above
        v = f.axis == 1 ? -v : v;
        w = f.axis == 2 ? -w : w;
        T rho, ru, rv, rw, rE;
        deriveCons(e, u, v, w, p, Tt, rho, ru, rv, rw, rE);
        F.f[0][o] = rho; F.f[1][o] = ru; F.f[2][o] = rv; F.f[3][o] = rw; F.f[4][o] = rE;
        if (prim) {
            T ka;
            deriveTrans(e, Tt, mu, ka);
            storePrims(L, prim, o, u, v, w, p, Tt, mu, ka);
        }
    }
}

template <class T>
__global__ void k_edge_face(Layout L, Params P, EdgeFace f, T* S, const double* tdev, const double* dtdev, double frac,
                            T* prim) {
    const double ts = *tdev + frac * *dtdev;
    edgeFaceNode(L, P, f, S, ts, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, prim);
}

// Several edge faces in one launch (faces that neither write what another of them reads): face q takes
// the blocks start[q] .. start[q+1]-1, laid out as its own (ceil(n1/64), n2) grid.
// rowSlots: the face's first in-plane axis is x and it owns whole rows; its threads then run over the slots
// of a row, and the padding slots between x-segments are written too (zeros, never read), so the face's
// lines are written whole.
struct EdgeFaceGroup {
    EdgeFace f[6];
    int start[7];
    int nbx[6];
    int rowSlots[6];
    int n;
};

template <class T>
__global__ void k_edge_faces(Layout L, Params P, EdgeFaceGroup g, T* S, const double* tdev, const double* dtdev,
                             double frac, T* prim) {
    const int b = blockIdx.x;
    int q = 0;
    while (q + 1 < g.n && b >= g.start[q + 1]) ++q;
    const int lb = b - g.start[q];
    const int a = (lb % g.nbx[q]) * 64 + threadIdx.x, bb = lb / g.nbx[q];
    const double ts = *tdev + frac * *dtdev;
    const EdgeFace& f = g.f[q];
    if (g.rowSlots[q]) {
        if (a >= L.pitch) return;
        const int r = a % Layout::kSegSlots;
        const int i = (a / Layout::kSegSlots) * Layout::kSegNodes + r - 1;
        if (r == 0 || r == Layout::kSegSlots - 1 || i >= L.NI) {
            const int j = f.axis == 1 ? f.j0 : f.j0 + bb, k = f.axis == 1 ? f.k0 + bb : f.k0;
            const size_t o = ((size_t)k * L.NJ + j) * L.pitch + a;
            for (int c = 0; c < 5; ++c) S[o + c * L.fstride] = (T)0;
            return;
        }
        edgeFaceNode(L, P, f, S, ts, i, bb, prim);
    } else {
        edgeFaceNode(L, P, f, S, ts, a, bb, prim);
    }
}

// Several faces of every case in one launch: entries [off[m], off[m+1]) of the thread range belong to face
// face[m], whose nodes over the cases follow the prefix sums start[m].
template <class T>
__global__ void k_edge_faces_batched(const CaseView<T>* __restrict__ cv, EdgeFaceListB fl, int ncases, int role,
                                     int parity, double frac) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fl.off[fl.n]) return;
    int m = 0;
    while (m + 1 < fl.n && t >= fl.off[m + 1]) ++m;
    const int64_t* start = fl.start[m];
    const int64_t q = t - fl.off[m];
    const int c = findCase(start, ncases, q);
    const CaseView<T>& v = cv[c];
    const EdgeFace& f = v.faces[fl.face[m]];
    const int64_t r = q - start[c];
    const double ts = v.scal->t + frac * v.scal->dt;
    T* S = roleBuffer(v, role, parity);
    const int ax1 = f.axis == 0 ? 1 : 0;
    const int lo1 = ax1 == 0 ? f.i0 : f.j0, hi1 = ax1 == 0 ? f.i1 : f.j1;
    const int n1 = hi1 - lo1 + 1;
    edgeFaceNode(v.L, v.P, f, S, ts, (int)(r % n1), (int)(r / n1));
}

// filterClosestFluidNodes, src/Boundary.cpp:618-652, first phase: filtered rho and rho*E into tmp, and
// the pre-filter values into the side tables (window-dense), which the next stencil evaluation uses for
// the primitive variables of filter nodes (the reference keeps them pre-filter).
template <class T>
__device__ __forceinline__ void filterComputeNode(const Layout& L, const Window& W, const int64_t* fn, int64_t q, const T* S,
                                                  T* tmp, T* rhoPre, T* rhoEPre) {
    int i, j, k;
    nodeIJK(L, fn[q], i, j, k);
    size_t o = L.off(i, j, k);
    const T* r = S;
    const T* E = S + 4 * L.fstride;
    const long xp = Layout::xstep(i, 1), xm = Layout::xstep(i, -1), sy = L.pitch, sz = (long)L.plane;
    // order of the reference sum: P*6, i+1, i-1, j+1, j-1, k+1, k-1
    T fr = ((T)6 * r[o] + r[o + xp] + r[o + xm] + r[o + sy] + r[o - sy] + r[o + sz] + r[o - sz]) / (T)12;
    T fe = ((T)6 * E[o] + E[o + xp] + E[o + xm] + E[o + sy] + E[o - sy] + E[o + sz] + E[o - sz]) / (T)12;
    tmp[2 * q] = fr;
    tmp[2 * q + 1] = fe;
    size_t w = W.idx(i, j, k);
    rhoPre[w] = r[o];
    rhoEPre[w] = E[o];
}

template <class T>
__device__ __forceinline__ void filterWriteNode(const Layout& L, const int64_t* fn, int64_t q, T* S, const T* tmp) {
    int i, j, k;
    nodeIJK(L, fn[q], i, j, k);
    size_t o = L.off(i, j, k);
    S[o] = tmp[2 * q];
    S[o + 4 * L.fstride] = tmp[2 * q + 1];
}

template <class T>
__global__ void k_filter_compute(Layout L, Window W, const int64_t* fn, int64_t n, const T* S, T* tmp, T* rhoPre,
                                 T* rhoEPre) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) filterComputeNode(L, W, fn, q, S, tmp, rhoPre, rhoEPre);
}

template <class T>
__global__ void k_filter_write(Layout L, const int64_t* fn, int64_t n, T* S, const T* tmp) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) filterWriteNode(L, fn, q, S, tmp);
}

template <class T>
__global__ void k_filter_compute_batched(const CaseView<T>* __restrict__ cv, const int64_t* __restrict__ start,
                                         int ncases, int role, int parity) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= start[ncases]) return;
    const int c = findCase(start, ncases, q);
    const CaseView<T>& v = cv[c];
    filterComputeNode(v.L, v.W, v.filterList, q - start[c], (const T*)roleBuffer(v, role, parity), v.ftmp, v.rhoPre,
                      v.rhoEPre);
}

template <class T>
__global__ void k_filter_write_batched(const CaseView<T>* __restrict__ cv, const int64_t* __restrict__ start,
                                       int ncases, int role, int parity) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= start[ncases]) return;
    const int c = findCase(start, ncases, q);
    const CaseView<T>& v = cv[c];
    filterWriteNode(v.L, v.filterList, q - start[c], roleBuffer(v, role, parity), (const T*)v.ftmp);
}

__device__ __forceinline__ double luSolveDot(const double* LU, const int8_t* rp, const int8_t* cp, const double* f, double x,
                                             double y, double z) {
    double yv[8], zv[8], c[8];
    for (int r = 0; r < 8; ++r) yv[r] = f[rp[r]];
    for (int r = 1; r < 8; ++r)
        for (int q = 0; q < r; ++q) yv[r] -= LU[r * 8 + q] * yv[q];
    for (int r = 7; r >= 0; --r) {
        double acc = yv[r];
        for (int q = r + 1; q < 8; ++q) acc -= LU[r * 8 + q] * zv[q];
        zv[r] = LU[r * 8 + r] != 0 ? acc / LU[r * 8 + r] : 0.0;
    }
    for (int r = 0; r < 8; ++r) c[cp[r]] = zv[r];
    // trilinearInterpolation, src/Boundary.cpp:826-833
    return c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * x * z + c[6] * y * z + c[7] * x * y * z;
}

// Ghost update over the compacted ghost list (src/Boundary.cpp:432-456): gather the 8 image-point box
// values, interpolate (tensor-product lerp when all 8 are fluid, else the stored FP64 Vandermonde
// factors), mirror into the ghost. Interpolation runs in FP64 whatever the field precision. One thread
// per (ghost, field): the five fields of a ghost are independent (a ghost corner that is itself a ghost
// enters as 0 and is never read), so they run in parallel.
template <class T>
__device__ __forceinline__ void ghostApplyField(const Layout& L, const Params& P, const GhostTables& G, int64_t g, int q,
                                                T* S) {
    const uint8_t mask = G.mask[g];
    T* Sq = S + (size_t)q * L.fstride;
    double f[8];
    for (int c = 0; c < 8; ++c) {
        const size_t o = (size_t)G.corner[8 * g + c];
        f[c] = (mask >> c) & 1 ? 0.0 : (double)Sq[o];
    }
    double ipv;
    if (mask == 0) {  // simplifiedInterpolation, src/Boundary.cpp:732-757
        const double rx = G.rel[3 * g], ry = G.rel[3 * g + 1], rz = G.rel[3 * g + 2];
        const double a01 = f[0] + (f[1] - f[0]) * rz / P.hz;
        const double a23 = f[2] + (f[3] - f[2]) * rz / P.hz;
        const double a45 = f[4] + (f[5] - f[4]) * rz / P.hz;
        const double a67 = f[6] + (f[7] - f[6]) * rz / P.hz;
        const double a0123 = a01 + (a23 - a01) * ry / P.hy;
        const double a4567 = a45 + (a67 - a45) * ry / P.hy;
        ipv = a0123 + (a4567 - a0123) * rx / P.hx;
    } else if constexpr (sizeof(T) == 4) {
        // FP32 fields: the precomputed image-point weights (one dot product; the same linear map as the
        // substitutions, rounded differently)
        const double* w = (q == 0 || q == 4 ? G.wN : G.wD) + 8 * g;
        ipv = 0;
        for (int c = 0; c < 8; ++c) ipv += w[c] * f[c];
    } else {
        // density and energy: Neumann factors; momentum: Dirichlet
        const bool neu = q == 0 || q == 4;
        const int8_t* pm = G.perm + 32 * g + (neu ? 16 : 0);
        ipv = luSolveDot((neu ? G.luN : G.luD) + 64 * g, pm, pm + 8, f, G.ip[3 * g], G.ip[3 * g + 1], G.ip[3 * g + 2]);
    }
    // getGhostNodeBCVariables, src/Boundary.cpp:777-785: the momentum changes sign
    Sq[(size_t)G.node[g]] = (T)(q >= 1 && q <= 3 ? -ipv : ipv);
}

// All five fields of one ghost in one thread (FP32 fields: one dot product per field, so the corner offsets
// and tables are read once per ghost).
template <class T>
__device__ __forceinline__ void ghostApplyAll(const Layout& L, const Params& P, const GhostTables& G, int64_t g, T* S) {
    const uint8_t mask = G.mask[g];
    size_t o[8];
    for (int c = 0; c < 8; ++c) o[c] = (size_t)G.corner[8 * g + c];
    const size_t node = (size_t)G.node[g];
    double r[3], wD[8], wN[8];
    if (mask == 0) {
        for (int d = 0; d < 3; ++d) r[d] = G.rel[3 * g + d];
    } else {
        for (int c = 0; c < 8; ++c) {
            wD[c] = G.wD[8 * g + c];
            wN[c] = G.wN[8 * g + c];
        }
    }
    double out[5];
    for (int q = 0; q < 5; ++q) {
        const T* Sq = S + (size_t)q * L.fstride;
        double f[8];
        for (int c = 0; c < 8; ++c) f[c] = (mask >> c) & 1 ? 0.0 : (double)Sq[o[c]];
        double ipv;
        if (mask == 0) {  // simplifiedInterpolation, src/Boundary.cpp:732-757
            const double a01 = f[0] + (f[1] - f[0]) * r[2] / P.hz;
            const double a23 = f[2] + (f[3] - f[2]) * r[2] / P.hz;
            const double a45 = f[4] + (f[5] - f[4]) * r[2] / P.hz;
            const double a67 = f[6] + (f[7] - f[6]) * r[2] / P.hz;
            const double a0123 = a01 + (a23 - a01) * r[1] / P.hy;
            const double a4567 = a45 + (a67 - a45) * r[1] / P.hy;
            ipv = a0123 + (a4567 - a0123) * r[0] / P.hx;
        } else {
            const double* w = q == 0 || q == 4 ? wN : wD;
            ipv = 0;
            for (int c = 0; c < 8; ++c) ipv += w[c] * f[c];
        }
        out[q] = q >= 1 && q <= 3 ? -ipv : ipv;
    }
    // a ghost corner enters as its boundary value, never as its stored value, so the writes of other ghosts
    // of the same launch do not reach this ghost's inputs
    for (int q = 0; q < 5; ++q) S[(size_t)q * L.fstride + node] = (T)out[q];
}

// Stored primitive and transport values of the ghosts after their update (src/Boundary.cpp:447-455:
// derivePrimitiveVariables and deriveTransportProperties of the ghost's conserved values).
template <class T>
__global__ void k_ghost_prims(Layout L, Params P, GhostTables G, int64_t n, const T* S, T* prim) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t o = (size_t)G.node[g];
    Eos<T> e = Eos<T>::make(P);
    T u, v, w, p, Tt, mu;
    derivePrims(e, S[o], S[o + L.fstride], S[o + 2 * L.fstride], S[o + 3 * L.fstride], S[o + 4 * L.fstride], u, v, w, p,
                Tt, mu);
    storePrims(L, prim, o, u, v, w, p, Tt, mu, mu / e.gm1Pr);
}

// Ghosts per launch from which a thread per ghost (all five fields) beats a thread per (ghost, field) in
// FP32: below it the launch is latency-bound and five times the threads finish sooner.
constexpr int64_t kGhostPerThread = 65536;

template <class T, bool PER_GHOST>
__global__ void k_ghost_apply(Layout L, Params P, GhostTables G, int64_t n, T* S) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (PER_GHOST) {
        if (t < n) ghostApplyAll(L, P, G, t, S);
    } else {
        if (t < 5 * n) ghostApplyField(L, P, G, t / 5, (int)(t % 5), S);
    }
}

template <class T, bool PER_GHOST>
__global__ void k_ghost_apply_batched(const CaseView<T>* __restrict__ cv, const int64_t* __restrict__ start, int ncases,
                                      int role, int parity) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = PER_GHOST ? t : t / 5;
    if (q >= start[ncases]) return;
    const int c = findCase(start, ncases, q);
    const CaseView<T>& v = cv[c];
    if constexpr (PER_GHOST) ghostApplyAll(v.L, v.P, v.G, q - start[c], roleBuffer(v, role, parity));
    else ghostApplyField(v.L, v.P, v.G, q - start[c], (int)(t % 5), roleBuffer(v, role, parity));
}

inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------------------------------------ launchers

void launchClassify(hipStream_t s, const Layout& L, uint8_t* cls, const int ft[6], const BodySet& bodies, const Params& P) {
    dim3 blk(128), grd(blocks(L.NI, 128), L.NJ, L.NK);
    hipLaunchKernelGGL(k_classify, grd, blk, 0, s, L, cls, ft[0], ft[1], ft[2], ft[3], ft[4], ft[5], bodies, P);
}

void launchMarkGhosts(hipStream_t s, const Layout& L, uint8_t* cls, const Window& W) {
    if (W.size() == 0) return;
    dim3 blk(64), grd(blocks(W.ni, 64), W.nj, W.nk);
    hipLaunchKernelGGL(k_mark_ghosts, grd, blk, 0, s, L, cls, W);
}

int64_t compactWindow(hipStream_t s, const Layout& L, const uint8_t* cls, const Window& W, uint8_t mask, uint8_t value,
                      int64_t* out, int32_t* counts, int32_t* offsets) {
    if (W.size() == 0) return 0;
    unsigned nb = blocks((int64_t)W.size(), 256);
    hipLaunchKernelGGL(k_count, dim3(nb), dim3(256), 0, s, L, cls, W, mask, value, counts);
    std::vector<int32_t> h(nb);
    hipCheck(hipMemcpyAsync(h.data(), counts, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s), "copy compaction counts");
    hipCheck(hipStreamSynchronize(s), "sync compaction counts");
    int64_t total = 0;
    for (unsigned b = 0; b < nb; ++b) {
        int32_t c = h[b];
        h[b] = (int32_t)total;
        total += c;
    }
    hipCheck(hipMemcpyAsync(offsets, h.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, s), "copy compaction offsets");
    if (out && total > 0) hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, s, L, cls, W, mask, value, offsets, out);
    hipCheck(hipStreamSynchronize(s), "sync compaction");
    hipCheck(hipGetLastError(), "compaction kernels");
    return total;
}

void launchImagePoints(hipStream_t s, const Layout& L, uint8_t* cls, const Params& P, const BodyDesc& B,
                       const int64_t* ghosts, int64_t n, double* bi, double* ip, int32_t* changed, int32_t* error) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_image_points, dim3(blocks(n, 128)), dim3(128), 0, s, L, cls, P, B, ghosts, n, bi, ip, changed, error);
}

void launchFillGhostMap(hipStream_t s, const Layout& L, const Window& W, const int64_t* ghosts, int64_t n, int32_t* gmap) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_gmap, dim3(blocks(n, 128)), dim3(128), 0, s, L, W, ghosts, n, gmap);
}

void launchGhostGeometry(hipStream_t s, const Layout& L, const uint8_t* cls, const Params& P, const Window& W,
                         const int64_t* ghosts, int64_t base, int64_t n, const int32_t* gmap, GhostTables G,
                         int32_t* error) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_ghost_geometry, dim3(blocks(n, 64)), dim3(64), 0, s, L, cls, P, W, ghosts, base, n, gmap, G,
                       error);
}

GhostTables ghostTablesFrom(const GhostTables& G, int64_t b) {
    GhostTables o = G;
    o.node = G.node + b;
    o.corner = G.corner + 8 * b;
    o.mask = G.mask + b;
    o.rel = G.rel + 3 * b;
    o.ip = G.ip + 3 * b;
    o.bi = G.bi + 3 * b;
    o.luD = G.luD + 64 * b;
    o.luN = G.luN + 64 * b;
    o.perm = G.perm + 32 * b;
    o.boxLo = G.boxLo + 3 * b;
    o.wD = G.wD + 8 * b;
    o.wN = G.wN + 8 * b;
    return o;
}

template <class T>
void launchEdgeFace(hipStream_t s, const Layout& L, const Params& P, const EdgeFace& f, T* S, const double* tdev,
                    const double* dtdev, double frac) {
    int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
    int lo[3] = {f.i0, f.j0, f.k0}, hi[3] = {f.i1, f.j1, f.k1};
    int n1 = hi[ax1] - lo[ax1] + 1, n2 = hi[ax2] - lo[ax2] + 1;
    if (n1 <= 0 || n2 <= 0) return;
    hipLaunchKernelGGL(k_edge_face<T>, dim3(blocks(n1, 64), n2), dim3(64), 0, s, L, P, f, S, tdev, dtdev, frac, (T*)nullptr);
}

namespace {

// Inclusive index box.
struct Box {
    int lo[3], hi[3];
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
    bool meets(const Box& o) const {
        for (int a = 0; a < 3; ++a)
            if (hi[a] < o.lo[a] || o.hi[a] < lo[a]) return false;
        return true;
    }
};

Box ownedBox(const EdgeFace& f) { return Box{{f.i0, f.j0, f.k0}, {f.i1, f.j1, f.k1}}; }

// Nodes the face reads: the two planes in from it along its normal (inlet, outlet, symmetry; the inlet's
// in-plane neighbours lie inside its own box), or the source plane of a periodic copy.
Box readBox(const Layout& L, const EdgeFace& f) {
    Box r = ownedBox(f);
    const int a = f.axis, N = a == 0 ? L.NI : (a == 1 ? L.NJ : L.NK);
    if (f.type == NS_BC_PERIODIC) {
        r.lo[a] = r.hi[a] = f.side == 0 ? N - 2 : 1;
    } else if (f.side == 0) {
        r.lo[a] = r.lo[a] + 1;
        r.hi[a] = r.lo[a] + 1;
    } else {
        r.hi[a] = r.hi[a] - 1;
        r.lo[a] = r.hi[a] - 1;
    }
    return r;
}

}  // namespace

bool edgeFacesIndependent(const Layout& L, const EdgeFace& x, const EdgeFace& y) {
    return !readBox(L, x).meets(ownedBox(y)) && !readBox(L, y).meets(ownedBox(x)) && !ownedBox(x).meets(ownedBox(y));
}

template <class T>
void launchEdgeFaces(hipStream_t s, const Layout& L, const Params& P, const EdgeFace faces[6], T* S, const double* tdev,
                     const double* dtdev, double frac, T* prim) {
    // faces in application order, grouped greedily: a face joins the current launch when it is independent
    // of every face already in it; the groups run in order, so the result equals face-by-face application
    EdgeFaceGroup g{};
    auto flush = [&]() {
        if (g.n > 0 && g.start[g.n] > 0)
            hipLaunchKernelGGL(k_edge_faces<T>, dim3(g.start[g.n]), dim3(64), 0, s, L, P, g, S, tdev, dtdev, frac, prim);
        g = EdgeFaceGroup{};
    };
    for (int q = 0; q < 6; ++q) {
        const EdgeFace& f = faces[q];
        if (ownedBox(f).empty()) continue;
        bool ok = true;
        for (int r = 0; r < g.n; ++r) ok = ok && edgeFacesIndependent(L, g.f[r], f);
        if (!ok) flush();
        const int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
        const Box b = ownedBox(f);
        const bool rows = ax1 == 0 && b.lo[0] == 0 && b.hi[0] == L.NI - 1;
        const int n1 = rows ? L.pitch : b.hi[ax1] - b.lo[ax1] + 1, n2 = b.hi[ax2] - b.lo[ax2] + 1;
        g.f[g.n] = f;
        g.rowSlots[g.n] = rows;
        g.nbx[g.n] = (n1 + 63) / 64;
        g.start[g.n + 1] = g.start[g.n] + g.nbx[g.n] * n2;
        ++g.n;
    }
    flush();
}

template <class T>
void launchFilterCompute(hipStream_t s, const Layout& L, const Window& W, const int64_t* fn, int64_t n, const T* S, T* tmp,
                         T* rhoPre, T* rhoEPre) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_filter_compute<T>, dim3(blocks(n, 128)), dim3(128), 0, s, L, W, fn, n, S, tmp, rhoPre, rhoEPre);
}

template <class T>
void launchFilterWrite(hipStream_t s, const Layout& L, const int64_t* fn, int64_t n, T* S, const T* tmp) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_filter_write<T>, dim3(blocks(n, 128)), dim3(128), 0, s, L, fn, n, S, tmp);
}

template <class T>
void launchGhostApply(hipStream_t s, const Layout& L, const Params& P, const GhostTables& G, int64_t n, T* S, T* prim) {
    if (n == 0) return;
    if (sizeof(T) == 4 && n >= kGhostPerThread)
        hipLaunchKernelGGL((k_ghost_apply<T, true>), dim3(blocks(n, 64)), dim3(64), 0, s, L, P, G, n, S);
    else
        hipLaunchKernelGGL((k_ghost_apply<T, false>), dim3(blocks(5 * n, 64)), dim3(64), 0, s, L, P, G, n, S);
    if (prim) hipLaunchKernelGGL(k_ghost_prims<T>, dim3(blocks(n, 64)), dim3(64), 0, s, L, P, G, n, (const T*)S, prim);
}

template <class T>
void launchEdgeFacesBatched(hipStream_t s, const CaseView<T>* cv, const EdgeFaceListB& fl, int ncases, int role, int parity,
                            double frac) {
    if (fl.n == 0 || fl.off[fl.n] == 0) return;
    hipLaunchKernelGGL(k_edge_faces_batched<T>, dim3(blocks(fl.off[fl.n], 128)), dim3(128), 0, s, cv, fl, ncases, role,
                       parity, frac);
}

template <class T>
void launchFilterComputeBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                                int role, int parity) {
    if (total == 0) return;
    hipLaunchKernelGGL(k_filter_compute_batched<T>, dim3(blocks(total, 128)), dim3(128), 0, s, cv, start, ncases, role,
                       parity);
}

template <class T>
void launchFilterWriteBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                              int role, int parity) {
    if (total == 0) return;
    hipLaunchKernelGGL(k_filter_write_batched<T>, dim3(blocks(total, 128)), dim3(128), 0, s, cv, start, ncases, role,
                       parity);
}

template <class T>
void launchGhostApplyBatched(hipStream_t s, const CaseView<T>* cv, const int64_t* start, int64_t total, int ncases,
                             int role, int parity) {
    if (total == 0) return;
    if (sizeof(T) == 4 && total >= kGhostPerThread)
        hipLaunchKernelGGL((k_ghost_apply_batched<T, true>), dim3(blocks(total, 64)), dim3(64), 0, s, cv, start, ncases,
                           role, parity);
    else
        hipLaunchKernelGGL((k_ghost_apply_batched<T, false>), dim3(blocks(5 * total, 64)), dim3(64), 0, s, cv, start,
                           ncases, role, parity);
}

#define NS_INST(T)                                                                                                           \
    template void launchEdgeFacesBatched<T>(hipStream_t, const CaseView<T>*, const EdgeFaceListB&, int, int, int, double);   \
    template void launchFilterComputeBatched<T>(hipStream_t, const CaseView<T>*, const int64_t*, int64_t, int, int, int);    \
    template void launchFilterWriteBatched<T>(hipStream_t, const CaseView<T>*, const int64_t*, int64_t, int, int, int);      \
    template void launchGhostApplyBatched<T>(hipStream_t, const CaseView<T>*, const int64_t*, int64_t, int, int, int);       \
    template void launchEdgeFaces<T>(hipStream_t, const Layout&, const Params&, const EdgeFace[6], T*, const double*,         \
                                     const double*, double, T*);                                                            \
    template void launchEdgeFace<T>(hipStream_t, const Layout&, const Params&, const EdgeFace&, T*, const double*,            \
                                    const double*, double);                                                                   \
    template void launchFilterCompute<T>(hipStream_t, const Layout&, const Window&, const int64_t*, int64_t, const T*, T*,   \
                                         T*, T*);                                                                             \
    template void launchFilterWrite<T>(hipStream_t, const Layout&, const int64_t*, int64_t, T*, const T*);                  \
    template void launchGhostApply<T>(hipStream_t, const Layout&, const Params&, const GhostTables&, int64_t, T*, T*);
NS_INST(double)
NS_INST(float)
#undef NS_INST

}  // namespace nsfsi
