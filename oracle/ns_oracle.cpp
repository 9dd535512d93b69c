/*
 * ns_oracle.cpp — TEST INFRASTRUCTURE ONLY. Serial FP64 CPU restatement of the reference
 * (Frederik-Kris/3D-NS-FSI) explicit RK4 time step, used as the correctness checker and as the timed
 * single-core CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product path (the HIP library) never calls it.
 *
 * It keeps the reference's data layout (37 FP64 arrays, k fastest, src/Array3D.h:42), its loop
 * structure, its buffer choreography (Old/Cur swap, stale ghost reads, pre-filter primitives) and its
 * floating-point expression order, and must be built like the reference Release build: g++ -O3, no
 * -march, -ffp-contract=off (no FMA contraction).
 *
 * The class is a template over the scalar type of the state: on double it is the oracle; on Dual (value,
 * tangent) it is the oracle's tangent-linear form, the forward-mode derivative of the same step at given
 * step sizes (ns_oracle_tl_*), used to check the device adjoint's gradients.
 *
 * Where the reference calls Eigen's FullPivLU (src/Boundary.cpp:825), this file has its own textbook
 * full-pivot LU (lu8_*); agreement with Eigen's own rounding is parity unpinned. The sphere body uses
 * the cylinder's ghost-augmentation loop (src/Boundary.cpp:912-916): the sphere's own loop
 * (src/Boundary.cpp:1041-1042) iterates across two containers, which is undefined behaviour.
 * Index boxes that would wrap below zero or leave the mesh (undefined behaviour in the reference,
 * src/Boundary.cpp:945-963, :1061-1070) raise a geometry error instead.
 */
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/ns_fsi.h"

namespace {

struct GeometryError : std::runtime_error { using std::runtime_error::runtime_error; };

enum NodeType : uint8_t { FluidInterior = 0, FluidEdge = 1, FluidGhost = 2, SolidInactive = 3, SolidGhost = 4 };

struct V3 {
    double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, double f) { return {a.x * f, a.y * f, a.z * f}; }
static inline V3 operator/(V3 a, double f) { return {a.x / f, a.y / f, a.z / f}; }
// Vector3_d::length, src/SmallVectors.h:20 (pow(x,2) is exactly x*x)
static inline double length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

struct U3 {
    size_t i, j, k;
};

// IndexBoundingBox, src/SmallVectors.h:74-104
struct Box {
    size_t iMin, iMax, jMin, jMax, kMin, kMax;
};

template <class R> struct ConsT { R rho, ru, rv, rw, rE; };
template <class R> struct PrimT { R u, v, w, p, T; };
template <class R> struct TransT { R mu, kappa; };

struct Ghost {
    U3 idx;
    V3 bi, ip;
};

// SplitMix64 finaliser; counter-based so every node's perturbation is independent of traversal.
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline double uniform01(uint64_t seed, uint64_t counter) {
    return (double)(mix64(seed + 0x9E3779B97F4A7C15ULL * (counter + 1)) >> 11) * 0x1.0p-53;
}

// Value and tangent of one quantity: the oracle instantiated on this type is the tangent-linear form of
// the step (forward-mode differentiation). Each value is computed by exactly the operations the plain
// FP64 oracle applies to it, so the values of a tangent run equal the plain run bit for bit; the tangents
// follow the sum, product and quotient rules (and d(x^1.5) = 1.5 sqrt(x) dx). Step sizes, norms of change
// and the time stay plain values: the tangent model is that of the step at given step sizes.
struct Dual {
    double v = 0, d = 0;
    Dual() = default;
    Dual(double x) : v(x) {}
    Dual(double x, double dx) : v(x), d(dx) {}
    Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
    Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
    Dual& operator/=(const Dual& b) { const double q = v / b.v; d = (d - q * b.d) / b.v; v = q; return *this; }
};
static inline Dual operator-(const Dual& a) { return {-a.v, -a.d}; }
static inline Dual operator+(const Dual& a, const Dual& b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(const Dual& a, const Dual& b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator*(const Dual& a, const Dual& b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(const Dual& a, const Dual& b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
static inline double valueOf(double x) { return x; }
static inline double valueOf(const Dual& x) { return x.v; }

template <class R>
class OracleT {
public:
    using Cons = ConsT<R>;
    using Prim = PrimT<R>;
    using Trans = TransT<R>;

    explicit OracleT(const ns_config& c) : P(c) {
        NI = (size_t)P.NI; NJ = (size_t)P.NJ; NK = (size_t)P.NK;
        if (P.NI < 3 || P.NJ < 3 || P.NK < 3) throw std::invalid_argument("NI, NJ, NK must be >= 3");
        N = NI * NJ * NK;
        // Mesh::Mesh, src/Mesh.cpp:10-21: 37 zero-initialised arrays + node types
        for (auto* grp : {&cons, &consOld}) for (auto& a : *grp) a.assign(N, 0.0);
        for (auto& a : prim) a.assign(N, 0.0);
        for (auto& a : trans) a.assign(N, 0.0);
        for (auto& s : k) for (auto& a : s) a.assign(N, 0.0);
        nodeType.assign(N, FluidInterior);
    }

    // ---------------------------------------------------------------- index math
    size_t id(size_t i, size_t j, size_t kk) const { return (i * NJ + j) * NK + kk; }  // src/SmallVectors.h:46-50
    U3 ijk(size_t n) const {  // getIndices3D, src/SmallVectors.h:37-43
        return {n / (NJ * NK), n % (NJ * NK) / NK, n % (NJ * NK) % NK};
    }
    V3 pos(size_t i, size_t j, size_t kk) const {  // getNodePosition, src/SmallVectors.h:56-63
        return {((int)i - off.x) * h.x, ((int)j - off.y) * h.y, ((int)kk - off.z) * h.z};
    }
    // getSurroundingNodesBox, src/SmallVectors.h:108-127 (negative or NaN values clamp to 1 instead of UB)
    Box surrounding(V3 p) const {
        auto nextIdx = [](double q, double hh, double o, size_t nmax) {
            double v = std::ceil(q / hh) + o;
            size_t n = (v >= 1.0) ? (size_t)v : 1;
            n = std::max<size_t>(n, 1);
            n = std::min<size_t>(n, nmax);
            return n;
        };
        size_t in = nextIdx(p.x, h.x, off.x, NI), jn = nextIdx(p.y, h.y, off.y, NJ), kn = nextIdx(p.z, h.z, off.z, NK);
        if (in > NI - 1 || jn > NJ - 1 || kn > NK - 1)
            throw GeometryError("image point box leaves the mesh");
        return {in - 1, in, jn - 1, jn, kn - 1, kn};
    }
    // IndexBoundingBox::asIndexList, src/SmallVectors.h:91-103: k fastest, i slowest
    std::array<size_t, 8> corners(const Box& b) const {
        return {id(b.iMin, b.jMin, b.kMin), id(b.iMin, b.jMin, b.kMax), id(b.iMin, b.jMax, b.kMin), id(b.iMin, b.jMax, b.kMax),
                id(b.iMax, b.jMin, b.kMin), id(b.iMax, b.jMin, b.kMax), id(b.iMax, b.jMax, b.kMin), id(b.iMax, b.jMax, b.kMax)};
    }

    // ---------------------------------------------------------------- EOS, src/FlowVariableGroupStructs.h:152-185
    Cons deriveCons(const Prim& q) const {
        R rho = (P.gamma * q.p - q.T) / (1 + q.T);
        R ru = (1 + rho) * q.u, rv = (1 + rho) * q.v, rw = (1 + rho) * q.w;
        R rE = q.p / (P.gamma - 1) + (1 + rho) / 2 * (q.u * q.u + q.v * q.v + q.w * q.w);
        return {rho, ru, rv, rw, rE};
    }
    Prim derivePrim(const Cons& c) const {
        R u = c.ru / (1 + c.rho), v = c.rv / (1 + c.rho), w = c.rw / (1 + c.rho);
        R p = (P.gamma - 1) * (c.rE - (1 + c.rho) / 2 * (u * u + v * v + w * w));
        R T = (P.gamma * p - c.rho) / (1 + c.rho);
        return {u, v, w, p, T};
    }
    // x^1.5 (src/Solver.cpp:476, src/FlowVariableGroupStructs.h:183): the reference's std::pow by default;
    // with crPow the correctly rounded value (Dekker's double-length x*sqrt(x), as glibc's pow returned it
    // before version 2.28; the current glibc pow is within 0.52 ulp and differs from the correctly rounded
    // value by one ulp for about one input in a thousand near 1).
    static double pow15cr(double x) {
        const double s = std::sqrt(x);
        const double r = std::fma(-s, s, x);
        const double e = r / (2.0 * s);
        const double p = x * s;
        const double pe = std::fma(x, s, -p);
        return p + (pe + x * e);
    }
    // tangent of x^1.5: 1.5 sqrt(x) dx
    R pow15(const R& x) const {
        if constexpr (std::is_same<R, double>::value) return crPow ? pow15cr(x) : std::pow(x, 1.5);
        else return Dual(crPow ? pow15cr(x.v) : std::pow(x.v, 1.5), 1.5 * std::sqrt(x.v) * x.d);
    }
    Trans deriveTrans(const Prim& q) const {
        double ScPlusOne = 1 + P.sutherlands_C2 / P.T_0;
        R mu = pow15(1 + q.T) * ScPlusOne / (Re * (q.T + ScPlusOne));
        R kappa = mu / ((P.gamma - 1) * P.Pr);
        return {mu, kappa};
    }
    // setFlowVariablesAtNode, src/FlowVariableGroupStructs.h:189-207 (on the current arrays)
    void setNode(size_t n, const Cons& c, const Prim& q, const Trans& t) {
        cons[0][n] = c.rho; cons[1][n] = c.ru; cons[2][n] = c.rv; cons[3][n] = c.rw; cons[4][n] = c.rE;
        prim[0][n] = q.u; prim[1][n] = q.v; prim[2][n] = q.w; prim[3][n] = q.p; prim[4][n] = q.T;
        trans[0][n] = t.mu; trans[1][n] = t.kappa;
    }

    // ---------------------------------------------------------------- setup, src/Mesh.cpp:25-58
    struct Edge {
        int axis, side, type;  // side 0 = min, 1 = max
        Box owned;
    };
    std::vector<Edge> edges;  // push order; applied in reverse (src/Mesh.cpp:70)

    void setupBoundaries() {
        const int order[6][2] = {{2, 0}, {2, 1}, {1, 0}, {1, 1}, {0, 0}, {0, 1}};  // z-min, z-max, y-min, y-max, x-min, x-max
        for (auto& o : order) edges.push_back({o[0], o[1], P.bc[2 * o[0] + o[1]], {}});
        // Spacing and origin offset per axis (src/Mesh.cpp:40-49): in/out N-1 (offset 0), periodic N-2
        // (offset 1), symmetry N-3 (offset 1); a ghost layer on a face removes one interval.
        double L[3] = {P.L_x, P.L_y, P.L_z};
        size_t Nn[3] = {NI, NJ, NK};
        double hh[3], oo[3];
        for (int a = 0; a < 3; ++a) {
            int tmin = P.bc[2 * a], tmax = P.bc[2 * a + 1];
            if ((tmin == NS_BC_PERIODIC) != (tmax == NS_BC_PERIODIC))
                throw std::invalid_argument("periodic faces must come in pairs");
            int gmin = (tmin == NS_BC_PERIODIC || tmin == NS_BC_SYMMETRY) ? 1 : 0;
            int gmax = (tmax == NS_BC_SYMMETRY) ? 1 : 0;
            long intervals = (long)Nn[a] - 1 - gmin - gmax;
            if (intervals < 1) throw std::invalid_argument("too few nodes for the boundary types");
            hh[a] = L[a] / (double)intervals;
            oo[a] = gmin;
        }
        h = {hh[0], hh[1], hh[2]};
        off = {oo[0], oo[1], oo[2]};
    }

    // MeshEdgeBoundary::identifyOwnedNodes, src/Boundary.cpp:18-71
    void claim(Edge& e, Box& unclaimed) {
        e.owned = unclaimed;
        size_t* mn[3] = {&e.owned.iMin, &e.owned.jMin, &e.owned.kMin};
        size_t* mx[3] = {&e.owned.iMax, &e.owned.jMax, &e.owned.kMax};
        size_t* umn[3] = {&unclaimed.iMin, &unclaimed.jMin, &unclaimed.kMin};
        size_t* umx[3] = {&unclaimed.iMax, &unclaimed.jMax, &unclaimed.kMax};
        if (e.side == 0) { *mx[e.axis] = *mn[e.axis]; ++*umn[e.axis]; }
        else { *mn[e.axis] = *mx[e.axis]; --*umx[e.axis]; }
        uint8_t t = (e.type == NS_BC_INLET || e.type == NS_BC_OUTLET) ? FluidEdge : FluidGhost;
        for (size_t i = e.owned.iMin; i <= e.owned.iMax; ++i)
            for (size_t j = e.owned.jMin; j <= e.owned.jMax; ++j)
                for (size_t kk = e.owned.kMin; kk <= e.owned.kMax; ++kk) nodeType[id(i, j, kk)] = t;
    }

    // ---------------------------------------------------------------- immersed bodies
    // One ImmersedBoundary each (src/Boundary.h:171-193): its own ghost list, ghost map and filter nodes,
    // classified and applied in the order of the body list (src/Mesh.cpp:52-57, :73-74, :165-166).
    struct Body {
        int type, axis;
        V3 c;
        double r;
        std::vector<Ghost> ghosts;
        std::map<size_t, size_t> ghostMap;
        std::vector<size_t> filterNodes;
    };
    std::vector<Body> bodies;
    // every body's lists concatenated in body order (what the inspection calls report)
    std::vector<Ghost> ghosts;
    std::vector<size_t> filterNodes;

    void setupBodies() {
        bodies.clear();
        if (P.n_extra_bodies < 0 || P.n_extra_bodies > NS_MAX_BODIES - 1)
            throw std::invalid_argument("n_extra_bodies out of range");
        if (P.n_extra_bodies > 0 && P.body_type == NS_BODY_NONE)
            throw std::invalid_argument("extra bodies need a first body");
        auto add = [&](int type, int axis, const double* c, double r) {
            if (type != NS_BODY_CYLINDER && type != NS_BODY_SPHERE) throw std::invalid_argument("body type");
            if (type == NS_BODY_CYLINDER && (axis < 0 || axis > 2)) throw std::invalid_argument("cylinder axis");
            if (!(r > 0)) throw std::invalid_argument("body radius must be positive");
            Body b;
            b.type = type;
            b.axis = axis;
            b.c = {c[0], c[1], c[2]};
            b.r = r;
            bodies.push_back(b);
        };
        if (P.body_type != NS_BODY_NONE) add(P.body_type, P.body_axis, P.body_centre, P.body_radius);
        for (int q = 0; q < P.n_extra_bodies; ++q)
            add(P.extra_bodies[q].type, P.extra_bodies[q].axis, P.extra_bodies[q].centre, P.extra_bodies[q].radius);
    }

    // CylinderBody::getNormalProbe / SphereBody::getNormalProbe, src/Boundary.cpp:920-935, :1047-1053
    static V3 normalProbe(const Body& B, V3 g) {
        V3 c2g = g - B.c;
        if (B.type == NS_BODY_CYLINDER) {
            if (B.axis == 0) c2g.x = 0; else if (B.axis == 1) c2g.y = 0; else c2g.z = 0;
        }
        double lf = (B.r - length(c2g)) / length(c2g);
        return c2g * lf;
    }

    // getCylinderBoundingBox / getSphereBoundingBox, src/Boundary.cpp:938-965, :1056-1072
    Box bodyBox(const Body& B, size_t layer) const {
        double r = B.r;
        long rad[3] = {(long)std::ceil(r / h.x) + 1 + (long)layer, (long)std::ceil(r / h.y) + 1 + (long)layer,
                       (long)std::ceil(r / h.z) + 1 + (long)layer};
        double c[3] = {B.c.x, B.c.y, B.c.z};
        double hh[3] = {h.x, h.y, h.z};
        long lo[3], hi[3];
        long Nn[3] = {(long)NI, (long)NJ, (long)NK};
        for (int a = 0; a < 3; ++a) {
            if (B.type == NS_BODY_CYLINDER && a == B.axis) { lo[a] = 0; hi[a] = Nn[a] - 1; continue; }
            double rc = std::round(c[a] / hh[a]);  // centre index without the origin offset (reference quirk)
            if (!(rc >= 0)) throw GeometryError("body bounding box below index 0");
            long ci = (long)rc;
            lo[a] = ci - rad[a];
            hi[a] = ci + rad[a];
            if (lo[a] < 0 || hi[a] > Nn[a] - 1) throw GeometryError("body bounding box leaves the mesh");
        }
        return {(size_t)lo[0], (size_t)hi[0], (size_t)lo[1], (size_t)hi[1], (size_t)lo[2], (size_t)hi[2]};
    }

    // getSolidAndFilterNodesIn{Cylinder,Sphere}, src/Boundary.cpp:968-1014, :1075-1103
    void solidAndFilter(Body& B, const Box& b, size_t layer, std::vector<size_t>& solids) {
        V3 c = B.c;
        double fw;
        if (B.type == NS_BODY_SPHERE) fw = layer * std::max({h.x, h.y, h.z});
        else if (B.axis == 0) fw = layer * std::max(h.y, h.z);
        else if (B.axis == 1) fw = layer * std::max(h.x, h.z);
        else fw = layer * std::max(h.x, h.y);
        for (size_t i = b.iMin; i <= b.iMax; ++i)
            for (size_t j = b.jMin; j <= b.jMax; ++j)
                for (size_t kk = b.kMin; kk <= b.kMax; ++kk) {
                    V3 p = pos(i, j, kk);
                    double dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z, d;
                    if (B.type == NS_BODY_SPHERE) d = std::sqrt(dx * dx + dy * dy + dz * dz);
                    else if (B.axis == 0) d = std::sqrt(dy * dy + dz * dz);
                    else if (B.axis == 1) d = std::sqrt(dx * dx + dz * dz);
                    else d = std::sqrt(dx * dx + dy * dy);
                    size_t n = id(i, j, kk);
                    if (d < B.r - P.machine_precision_buffer) {
                        nodeType[n] = SolidInactive;
                        solids.push_back(n);
                    } else if (d < B.r + fw && nodeType[n] == FluidInterior) {
                        B.filterNodes.push_back(n);
                    }
                }
    }

    // findGhostNodesWithFluidNeighbors, src/Boundary.cpp:461-562: 18 neighbours (no corners)
    void findGhosts(Body& B, const std::vector<size_t>& solids) {
        for (size_t n : solids) {
            U3 s = ijk(n);
            bool has = false;
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    for (int dk = -1; dk <= 1; ++dk) {
                        int nz = (di != 0) + (dj != 0) + (dk != 0);
                        if (nz == 0 || nz == 3) continue;
                        long ii = (long)s.i + di, jj = (long)s.j + dj, kk = (long)s.k + dk;
                        if (ii < 0 || jj < 0 || kk < 0 || ii >= (long)NI || jj >= (long)NJ || kk >= (long)NK) continue;
                        if (nodeType[id(ii, jj, kk)] == FluidInterior) has = true;
                    }
            if (has) {
                B.ghostMap[n] = B.ghosts.size();
                B.ghosts.push_back({s, {}, {}});
                nodeType[n] = SolidGhost;
            }
        }
    }

    // setImagePointPositions + checkIfSurroundingShouldBeGhost, src/Boundary.cpp:568-602
    std::vector<Ghost> imagePoints(Body& B, size_t first) {
        std::vector<Ghost> fresh;
        for (size_t g = first; g < B.ghosts.size(); ++g) {
            Ghost& gh = B.ghosts[g];
            V3 gp = pos(gh.idx.i, gh.idx.j, gh.idx.k);
            V3 pr = normalProbe(B, gp);
            gh.bi = gp + pr;
            gh.ip = gh.bi + pr;
            Box b = surrounding(gh.ip);
            for (size_t n : corners(b))
                if (nodeType[n] == SolidInactive) {
                    fresh.push_back({ijk(n), {}, {}});
                    nodeType[n] = SolidGhost;
                }
        }
        return fresh;
    }

    // Mesh::categorizeNodes, src/Mesh.cpp:61-82 and CylinderBody::identifyRelatedNodes, src/Boundary.cpp:894-917
    void categorizeNodes() {
        std::fill(nodeType.begin(), nodeType.end(), FluidInterior);
        Box unclaimed{0, NI - 1, 0, NJ - 1, 0, NK - 1};
        for (auto& e : edges) claim(e, unclaimed);
        std::reverse(edges.begin(), edges.end());
        for (Body& B : bodies) {
            const size_t layer = 2;
            Box b = bodyBox(B, layer);
            std::vector<size_t> solids;
            solidAndFilter(B, b, layer, solids);
            findGhosts(B, solids);
            std::vector<Ghost> fresh = imagePoints(B, 0);
            while (!fresh.empty()) {
                size_t first = B.ghosts.size();
                for (auto& g : fresh) {  // appendGhostNodes, src/Boundary.cpp:605-615
                    B.ghostMap[id(g.idx.i, g.idx.j, g.idx.k)] = B.ghosts.size();
                    B.ghosts.push_back(g);
                }
                fresh = imagePoints(B, first);
            }
        }
        ghosts.clear();
        filterNodes.clear();
        for (const Body& B : bodies) {
            ghosts.insert(ghosts.end(), B.ghosts.begin(), B.ghosts.end());
            filterNodes.insert(filterNodes.end(), B.filterNodes.begin(), B.filterNodes.end());
        }
        interior.clear();
        solidGhost.clear();
        for (size_t n = 0; n < N; ++n) {
            if (nodeType[n] == FluidInterior) interior.push_back(n);
            else if (nodeType[n] == SolidGhost) solidGhost.push_back(n);
        }
    }

    // ---------------------------------------------------------------- initial conditions
    void setAllNodes(const Prim& q0, bool perturb) {
        const double a = P.ic_amplitude;
        const double twoPi = 6.283185307179586;
        for (size_t i = 0; i < NI; ++i)
            for (size_t j = 0; j < NJ; ++j)
                for (size_t kk = 0; kk < NK; ++kk) {
                    Prim q = q0;
                    if (perturb) {
                        uint64_t base = (uint64_t)((kk * NJ + j) * NI + i) * 6;
                        double d[3];
                        for (int c = 0; c < 3; ++c) {
                            double u1 = uniform01(P.ic_seed, base + 2 * c), u2 = uniform01(P.ic_seed, base + 2 * c + 1);
                            if (P.ic_type == NS_IC_FREESTREAM_GAUSSIAN)
                                d[c] = a * (std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(twoPi * u2));
                            else
                                d[c] = a * (2.0 * u1 - 1.0);
                        }
                        q.u = q0.u + d[0]; q.v = q0.v + d[1]; q.w = q0.w + d[2];
                    }
                    Cons c = deriveCons(q);
                    Trans t = deriveTrans(q);
                    setNode(id(i, j, kk), c, q, t);
                }
    }
    // applyStagnation_IC / applyUniformFlow_IC, src/Solver.cpp:31-75; only the current arrays are set
    void applyIC() {
        if (P.ic_type == NS_IC_STAGNATION) setAllNodes({0, 0, 0, 0, 0}, false);
        else if (P.ic_type == NS_IC_UNIFORM_OBLIQUE) {
            double vel = P.M_0 / std::sqrt(3);
            setAllNodes({vel, vel, vel, 0, 0}, false);
        } else setAllNodes({P.M_0, 0, 0, 0, 0}, true);
    }

    // ---------------------------------------------------------------- dt, src/Solver.cpp:80-127
    void updateTimeStepSize(double tnow) {
        double sx = 0, sy = 0, sz = 0;
        for (size_t n = 0; n < N; ++n) {
            double c = std::sqrt((1 + P.gamma * valueOf(prim[3][n])) / (1 + valueOf(cons[0][n])));
            sx = std::max(sx, c + std::fabs(valueOf(prim[0][n])));
            sy = std::max(sy, c + std::fabs(valueOf(prim[1][n])));
            sz = std::max(sz, c + std::fabs(valueOf(prim[2][n])));
        }
        double dtConv = std::fabs(P.conv_stability_limit) / (sx / h.x + sy / h.y + sz / h.z);
        double vm = std::max(4. / 3, P.gamma / P.Pr);
        double nuMax = 0;
        for (size_t n = 0; n < N; ++n) nuMax = std::max(nuMax, valueOf(trans[0][n]) / (valueOf(cons[0][n]) + 1) * vm);
        double dx2 = h.x * h.x, dy2 = h.y * h.y, dz2 = h.z * h.z;
        double dtVisc = std::fabs(P.visc_stability_limit) / (nuMax * (1 / dx2 + 1 / dy2 + 1 / dz2));
        dt = std::min(dtConv, dtVisc);
        if (P.stop_criterion == 1 && tnow + dt > P.t_end) dt = P.t_end - tnow;
        // prescribed step sizes (test hook, as the device's ns_set_dt_schedule): the step from time level n
        // takes dts[n]
        if (!dtSchedule.empty() && timeLevel < dtSchedule.size()) dt = dtSchedule[timeLevel];
    }
    std::vector<double> dtSchedule;

    // ---------------------------------------------------------------- RK4 slopes, src/Solver.cpp:170-418
    void slopes(const std::array<std::vector<R>, 5>& U, std::array<std::vector<R>, 5>& K) {
        const double dx = h.x, dy = h.y, dz = h.z;
        // node-invariant factors (src/Solver.cpp:202-211, 249-258, 296-305, 345-360)
        const double n2dx = -1. / (2 * dx), n2dy = -1. / (2 * dy), n2dz = -1. / (2 * dz);
        const double p23xx = 2. / (3 * dx * dx), p23yy = 2. / (3 * dy * dy), p23zz = 2. / (3 * dz * dz);
        const double p12xx = 1. / (2 * dx * dx), p12yy = 1. / (2 * dy * dy), p12zz = 1. / (2 * dz * dz);
        const double n16ydx = -1. / (6 * dy * dx), p14xdy = 1. / (4 * dx * dy), n16zdx = -1. / (6 * dz * dx), p14xdz = 1. / (4 * dx * dz);
        const double n16xdy = -1. / (6 * dx * dy), p14ydx = 1. / (4 * dy * dx), n16zdy = -1. / (6 * dz * dy), p14ydz = 1. / (4 * dy * dz);
        const double p14zdx = 1. / (4 * dz * dx), n16xdz = -1. / (6 * dx * dz), p14zdy = 1. / (4 * dz * dy), n16ydz = -1. / (6 * dy * dz);
        // energy-equation factor names (src/Solver.cpp:345-359)
        const double eA = -1. / (6 * dy * dx), eB = -1. / (6 * dz * dx), eC = 1. / (4 * dy * dx), eD = 1. / (4 * dz * dx);
        const double eE = -1. / (6 * dz * dy), eF = 1. / (4 * dz * dy);
        const double rhoH0 = 1. / (P.gamma - 1);
        const auto &ru = U[1], &rv = U[2], &rw = U[3], &E = U[4];
        const auto &u = prim[0], &v = prim[1], &w = prim[2], &p = prim[3], &T = prim[4], &mu = trans[0], &ka = trans[1];
        const size_t sI = NJ * NK, sJ = NK, sK = 1;
        for (size_t n : interior) {
            const size_t e = n + sI, wq = n - sI, nn = n + sJ, s = n - sJ, up = n + sK, d = n - sK;
            // continuity, src/Solver.cpp:191-194
            {
                R fx = -(ru[e] - ru[wq]) / (2 * dx);
                R fy = -(rv[nn] - rv[s]) / (2 * dy);
                R fz = -(rw[up] - rw[d]) / (2 * dz);
                K[0][n] = fx + fy + fz;
            }
            // x-momentum, src/Solver.cpp:225-241
            {
                R cX = n2dx * (ru[e] * u[e] + p[e] - ru[wq] * u[wq] - p[wq]);
                R cY = n2dy * (ru[nn] * v[nn] - ru[s] * v[s]);
                R cZ = n2dz * (ru[up] * w[up] - ru[d] * w[d]);
                R vXX = p23xx * ((mu[e] + mu[n]) * (u[e] - u[n]) - (mu[wq] + mu[n]) * (u[n] - u[wq]));
                R vYY = p12yy * ((mu[nn] + mu[n]) * (u[nn] - u[n]) - (mu[s] + mu[n]) * (u[n] - u[s]));
                R vZZ = p12zz * ((mu[up] + mu[n]) * (u[up] - u[n]) - (mu[d] + mu[n]) * (u[n] - u[d]));
                R vXY = p14xdy * (mu[nn] * (v[nn + sI] - v[nn - sI]) - mu[s] * (v[s + sI] - v[s - sI]));
                R vYX = n16ydx * (mu[e] * (v[e + sJ] - v[e - sJ]) - mu[wq] * (v[wq + sJ] - v[wq - sJ]));
                R vXZ = p14xdz * (mu[up] * (w[up + sI] - w[up - sI]) - mu[d] * (w[d + sI] - w[d - sI]));
                R vZX = n16zdx * (mu[e] * (w[e + sK] - w[e - sK]) - mu[wq] * (w[wq + sK] - w[wq - sK]));
                K[1][n] = cX + cY + cZ + vXX + vYY + vZZ + vXY + vYX + vXZ + vZX;
            }
            // y-momentum, src/Solver.cpp:272-288
            {
                R cX = n2dx * (rv[e] * u[e] - rv[wq] * u[wq]);
                R cY = n2dy * (rv[nn] * v[nn] + p[nn] - rv[s] * v[s] - p[s]);
                R cZ = n2dz * (rv[up] * w[up] - rv[d] * w[d]);
                R vXX = p12xx * ((mu[e] + mu[n]) * (v[e] - v[n]) - (mu[wq] + mu[n]) * (v[n] - v[wq]));
                R vYY = p23yy * ((mu[nn] + mu[n]) * (v[nn] - v[n]) - (mu[s] + mu[n]) * (v[n] - v[s]));
                R vZZ = p12zz * ((mu[up] + mu[n]) * (v[up] - v[n]) - (mu[d] + mu[n]) * (v[n] - v[d]));
                R vYX = p14ydx * (mu[e] * (u[e + sJ] - u[e - sJ]) - mu[wq] * (u[wq + sJ] - u[wq - sJ]));
                R vXY = n16xdy * (mu[nn] * (u[nn + sI] - u[nn - sI]) - mu[s] * (u[s + sI] - u[s - sI]));
                R vYZ = p14ydz * (mu[up] * (w[up + sJ] - w[up - sJ]) - mu[d] * (w[d + sJ] - w[d - sJ]));
                R vZY = n16zdy * (mu[nn] * (w[nn + sK] - w[nn - sK]) - mu[s] * (w[s + sK] - w[s - sK]));
                K[2][n] = cX + cY + cZ + vXX + vYY + vZZ + vYX + vXY + vYZ + vZY;
            }
            // z-momentum, src/Solver.cpp:320-337
            {
                R cX = n2dx * (rw[e] * u[e] - rw[wq] * u[wq]);
                R cY = n2dy * (rw[nn] * v[nn] - rw[s] * v[s]);
                R cZ = n2dz * (rw[up] * w[up] + p[up] - rw[d] * w[d] - p[d]);
                R vXX = p12xx * ((mu[e] + mu[n]) * (w[e] - w[n]) - (mu[wq] + mu[n]) * (w[n] - w[wq]));
                R vYY = p12yy * ((mu[nn] + mu[n]) * (w[nn] - w[n]) - (mu[s] + mu[n]) * (w[n] - w[s]));
                R vZZ = p23zz * ((mu[up] + mu[n]) * (w[up] - w[n]) - (mu[d] + mu[n]) * (w[n] - w[d]));
                R vZX = p14zdx * (mu[e] * (u[e + sK] - u[e - sK]) - mu[wq] * (u[wq + sK] - u[wq - sK]));
                R vXZ = n16xdz * (mu[up] * (u[up + sI] - u[up - sI]) - mu[d] * (u[d + sI] - u[d - sI]));
                R vZY = p14zdy * (mu[nn] * (v[nn + sK] - v[nn - sK]) - mu[s] * (v[s + sK] - v[s - sK]));
                R vYZ = n16ydz * (mu[up] * (v[up + sJ] - v[up - sJ]) - mu[d] * (v[d + sJ] - v[d - sJ]));
                K[3][n] = cX + cY + cZ + vXX + vYY + vZZ + vZX + vXZ + vZY + vYZ;
            }
            // energy, src/Solver.cpp:378-416
            {
                const R muP = mu[n], muW = mu[wq], muE = mu[e], muS = mu[s], muN = mu[nn], muD = mu[d], muU = mu[up];
                const R uP = u[n], uW = u[wq], uE = u[e], uS = u[s], uN = u[nn], uD = u[d], uU = u[up];
                const R vP = v[n], vW = v[wq], vE = v[e], vS = v[s], vN = v[nn], vD = v[d], vU = v[up];
                const R wP = w[n], wW = w[wq], wE = w[e], wS = w[s], wN = w[nn], wD = w[d], wU = w[up];
                const R kP = ka[n], TP = T[n];
                R inv = n2dx * ((rhoH0 + E[e] + p[e]) * uE - (rhoH0 + E[wq] + p[wq]) * uW)
                           + n2dy * ((rhoH0 + E[nn] + p[nn]) * vN - (rhoH0 + E[s] + p[s]) * vS)
                           + n2dz * ((rhoH0 + E[up] + p[up]) * wU - (rhoH0 + E[d] + p[d]) * wD);
                R vx = p23xx * ((muE * uE + muP * uP) * (uE - uP) - (muW * uW + muP * uP) * (uP - uW))
                          + p12xx * ((muE * vE + muP * vP) * (vE - vP) - (muW * vW + muP * vP) * (vP - vW))
                          + p12xx * ((muE * wE + muP * wP) * (wE - wP) - (muW * wW + muP * wP) * (wP - wW))
                          + eA * (muE * uE * (v[e + sJ] - v[e - sJ]) - muW * uW * (v[wq + sJ] - v[wq - sJ]))
                          + eB * (muE * uE * (w[e + sK] - w[e - sK]) - muW * uW * (w[wq + sK] - w[wq - sK]))
                          + eC * (muE * vE * (u[e + sJ] - u[e - sJ]) - muW * vW * (u[wq + sJ] - u[wq - sJ]))
                          + eD * (muE * wE * (u[e + sK] - u[e - sK]) - muW * wW * (u[wq + sK] - u[wq - sK]));
                R vy = p12yy * ((muN * uN + muP * uP) * (uN - uP) - (muS * uS + muP * uP) * (uP - uS))
                          + p23yy * ((muN * vN + muP * vP) * (vN - vP) - (muS * vS + muP * vP) * (vP - vS))
                          + p12yy * ((muN * wN + muP * wP) * (wN - wP) - (muS * wS + muP * wP) * (wP - wS))
                          + eA * (muN * vN * (u[nn + sI] - u[nn - sI]) - muS * vS * (u[s + sI] - u[s - sI]))
                          + eE * (muN * vN * (w[nn + sK] - w[nn - sK]) - muS * vS * (w[s + sK] - w[s - sK]))
                          + eC * (muN * uN * (v[nn + sI] - v[nn - sI]) - muS * uS * (v[s + sI] - v[s - sI]))
                          + eF * (muN * wN * (v[nn + sK] - v[nn - sK]) - muS * wS * (v[s + sK] - v[s - sK]));
                R vz = p12zz * ((muU * uU + muP * uP) * (uU - uP) - (muD * uD + muP * uP) * (uP - uD))
                          + p12zz * ((muU * vU + muP * vP) * (vU - vP) - (muD * vD + muP * vP) * (vP - vD))
                          + p23zz * ((muU * wU + muP * wP) * (wU - wP) - (muD * wD + muP * wP) * (wP - wD))
                          + eB * (muU * wU * (u[up + sI] - u[up - sI]) - muD * wD * (u[d + sI] - u[d - sI]))
                          + eE * (muU * wU * (v[up + sJ] - v[up - sJ]) - muD * wD * (v[d + sJ] - v[d - sJ]))
                          + eD * (muU * uU * (w[up + sI] - w[up - sI]) - muD * uD * (w[d + sI] - w[d - sI]))
                          + eF * (muU * vU * (w[up + sJ] - w[up - sJ]) - muD * vD * (w[d + sJ] - w[d - sJ]));
                R heat = p12xx * ((ka[e] + kP) * (T[e] - TP) - (ka[wq] + kP) * (TP - T[wq]))
                            + p12yy * ((ka[nn] + kP) * (T[nn] - TP) - (ka[s] + kP) * (TP - T[s]))
                            + p12zz * ((ka[up] + kP) * (T[up] - TP) - (ka[d] + kP) * (TP - T[d]));
                K[4][n] = inv + vx + vy + vz + heat;
            }
        }
    }

    // computeIntermediateSolution, src/Solver.cpp:435-442
    void intermediate(const std::array<std::vector<R>, 5>& Ks, double inc) {
        for (int q = 0; q < 5; ++q)
            for (size_t n : interior) cons[q][n] = consOld[q][n] + inc * Ks[q][n];
    }
    // compute_RK4_final_step, src/Solver.cpp:494-507
    void finalStep() {
        double f = dt / 6;
        for (int q = 0; q < 5; ++q)
            for (size_t n : interior) {
                R r = f * (k[0][q][n] + 2 * k[1][q][n] + 2 * k[2][q][n] + k[3][q][n]);
                cons[q][n] = consOld[q][n] + r;
            }
    }
    // updatePrimitiveVariables, src/Solver.cpp:446-479
    void updatePrimitives() {
        const double gm1 = P.gamma - 1, Sc = P.sutherlands_C2 / P.T_0, ScP1 = 1 + Sc, pf = 1 / (gm1 * P.Pr);
        for (size_t n : interior) {
            R rho = cons[0][n];
            R u = cons[1][n] / (rho + 1), v = cons[2][n] / (rho + 1), w = cons[3][n] / (rho + 1);
            R p = gm1 * (cons[4][n] - (rho + 1) / 2 * (u * u + v * v + w * w));
            R T = (P.gamma * p - rho) / (1 + rho);
            R mu = pow15(1 + T) * ScP1 / (Re * (T + ScP1));
            prim[0][n] = u; prim[1][n] = v; prim[2][n] = w; prim[3][n] = p; prim[4][n] = T;
            trans[0][n] = mu; trans[1][n] = mu * pf;
        }
    }

    // ---------------------------------------------------------------- edge BCs, src/Boundary.cpp:74-417
    void adjacent(const Edge& e, size_t n, size_t& a1, size_t& a2) const {
        size_t st = e.axis == 0 ? NJ * NK : (e.axis == 1 ? NK : 1);
        if (e.side == 0) { a1 = n + st; a2 = n + 2 * st; }
        else { a1 = n - st; a2 = n - 2 * st; }
    }
    size_t periodicSource(const Edge& e, size_t i, size_t j, size_t kk) const {
        size_t Nn = e.axis == 0 ? NI : (e.axis == 1 ? NJ : NK);
        size_t v = e.side == 0 ? Nn - 2 : 1;
        if (e.axis == 0) i = v; else if (e.axis == 1) j = v; else kk = v;
        return id(i, j, kk);
    }
    template <class F> void forOwned(const Box& b, F f) {
        for (size_t i = b.iMin; i <= b.iMax; ++i)
            for (size_t j = b.jMin; j <= b.jMax; ++j)
                for (size_t kk = b.kMin; kk <= b.kMax; ++kk) f(i, j, kk);
    }
    void applyEdge(const Edge& e, double tnow) {
        auto& rho = cons[0];
        const Box& b = e.owned;
        if (e.type == NS_BC_INLET) {
            forOwned(b, [&](size_t i, size_t j, size_t kk) {  // linear extrapolation, src/Boundary.cpp:254-264
                size_t n = id(i, j, kk), a1, a2;
                adjacent(e, n, a1, a2);
                rho[n] = 2 * rho[a1] - rho[a2];
            });
            // filterInletDensity, src/Boundary.cpp:142-245: in-plane neighbours inside the owned box
            std::vector<R> filt;
            std::vector<size_t> nodes;
            int t1 = e.axis == 0 ? 1 : 0, t2 = e.axis == 2 ? 1 : 2;  // the two in-plane axes, in loop order
            forOwned(b, [&](size_t i, size_t j, size_t kk) {
                size_t c[3] = {i, j, kk};
                size_t lo[3] = {b.iMin, b.jMin, b.kMin}, hi[3] = {b.iMax, b.jMax, b.kMax};
                R acc = 0;
                unsigned nb = 0;
                for (int ax : {t1, t2}) {
                    if (c[ax] < hi[ax]) { size_t cc[3] = {c[0], c[1], c[2]}; ++cc[ax]; acc += rho[id(cc[0], cc[1], cc[2])]; ++nb; }
                    if (c[ax] > lo[ax]) { size_t cc[3] = {c[0], c[1], c[2]}; --cc[ax]; acc += rho[id(cc[0], cc[1], cc[2])]; ++nb; }
                }
                acc += nb * rho[id(i, j, kk)];
                acc /= 2 * nb;
                filt.push_back(acc);
                nodes.push_back(id(i, j, kk));
            });
            for (size_t q = 0; q < nodes.size(); ++q) rho[nodes[q]] = filt[q];
            // ramped normal velocity, T' = 0, p' from rho (src/Boundary.cpp:267-295)
            double vel = std::min(1., tnow / P.inlet_ramp_time) * P.M_0;
            if (e.side == 1) vel *= -1;
            double uu = 0, vv = 0, ww = 0;
            if (e.axis == 0) uu = vel; else if (e.axis == 1) vv = vel; else ww = vel;
            forOwned(b, [&](size_t i, size_t j, size_t kk) {
                size_t n = id(i, j, kk);
                R r = rho[n], T = 0;
                R p = (T * (1 + r) + r) / P.gamma;
                Prim q{uu, vv, ww, p, T};
                setNode(n, deriveCons(q), q, deriveTrans(q));
            });
        } else if (e.type == NS_BC_OUTLET) {  // src/Boundary.cpp:304-336
            forOwned(b, [&](size_t i, size_t j, size_t kk) {
                size_t n = id(i, j, kk), a1, a2;
                adjacent(e, n, a1, a2);
                R r = 2 * cons[0][a1] - cons[0][a2];
                R ru = 2 * cons[1][a1] - cons[1][a2];
                R rv = 2 * cons[2][a1] - cons[2][a2];
                R rw = 2 * cons[3][a1] - cons[3][a2];
                R p = 0;
                R u = ru / (1 + r), v = rv / (1 + r), w = rw / (1 + r);
                R T = (P.gamma * p - r) / (1 + r);
                R rE = p / (P.gamma - 1) + (1 + r) / 2 * (u * u + v * v + w * w);
                Prim q{u, v, w, p, T};
                setNode(n, {r, ru, rv, rw, rE}, q, deriveTrans(q));
            });
        } else if (e.type == NS_BC_PERIODIC) {  // src/Boundary.cpp:344-367
            forOwned(b, [&](size_t i, size_t j, size_t kk) {
                size_t n = id(i, j, kk), o = periodicSource(e, i, j, kk);
                for (int q = 0; q < 5; ++q) cons[q][n] = cons[q][o];
                for (int q = 0; q < 5; ++q) prim[q][n] = prim[q][o];
                trans[0][n] = trans[0][o];
                trans[1][n] = trans[1][o];
            });
        } else {  // symmetry, src/Boundary.cpp:377-417
            forOwned(b, [&](size_t i, size_t j, size_t kk) {
                size_t n = id(i, j, kk), a1, a2;
                adjacent(e, n, a1, a2);
                R u = prim[0][a2], v = prim[1][a2], w = prim[2][a2];
                if (e.axis == 0) u = -u; else if (e.axis == 1) v = -v; else w = -w;
                Prim q{u, v, w, prim[3][a2], prim[4][a2]};
                setNode(n, deriveCons(q), q, deriveTrans(q));
            });
        }
    }

    // ---------------------------------------------------------------- 8x8 full-pivot LU (replaces Eigen FullPivLU)
    // Textbook Gaussian elimination with complete pivoting: at step s the entry of largest magnitude in the
    // trailing (8-s)x(8-s) block is moved to (s,s) (first maximum in column-major scan order); rows and
    // columns are swapped. Solve: c = Q * U^-1 * L^-1 * P * f.
    static void lu8Factor(double A[64], int rowPerm[8], int colPerm[8]) {
        for (int q = 0; q < 8; ++q) { rowPerm[q] = q; colPerm[q] = q; }
        for (int s = 0; s < 8; ++s) {
            int pr = s, pc = s;
            double best = -1;
            for (int c = s; c < 8; ++c)
                for (int r = s; r < 8; ++r) {
                    double m = std::fabs(A[r * 8 + c]);
                    if (m > best) { best = m; pr = r; pc = c; }
                }
            if (pr != s) { for (int c = 0; c < 8; ++c) std::swap(A[s * 8 + c], A[pr * 8 + c]); std::swap(rowPerm[s], rowPerm[pr]); }
            if (pc != s) { for (int r = 0; r < 8; ++r) std::swap(A[r * 8 + s], A[r * 8 + pc]); std::swap(colPerm[s], colPerm[pc]); }
            double piv = A[s * 8 + s];
            if (piv == 0) continue;  // singular trailing block: leave zeros (solution component set to 0)
            for (int r = s + 1; r < 8; ++r) {
                double l = A[r * 8 + s] / piv;
                A[r * 8 + s] = l;
                for (int c = s + 1; c < 8; ++c) A[r * 8 + c] -= l * A[s * 8 + c];
            }
        }
    }
    template <class S>
    static void lu8Solve(const double LU[64], const int rowPerm[8], const int colPerm[8], const S f[8], S x[8]) {
        S y[8];
        for (int r = 0; r < 8; ++r) y[r] = f[rowPerm[r]];
        for (int r = 1; r < 8; ++r)
            for (int c = 0; c < r; ++c) y[r] -= LU[r * 8 + c] * y[c];
        S z[8];
        for (int r = 7; r >= 0; --r) {
            S acc = y[r];
            for (int c = r + 1; c < 8; ++c) acc -= LU[r * 8 + c] * z[c];
            z[r] = LU[r * 8 + r] != 0 ? acc / LU[r * 8 + r] : 0.0;
        }
        for (int r = 0; r < 8; ++r) x[colPerm[r]] = z[r];
    }
    template <class S>
    static S lu8SolveDot(const double V[64], const S f[8], V3 ip) {
        double A[64];
        std::memcpy(A, V, sizeof(A));
        int rp[8], cp[8];
        lu8Factor(A, rp, cp);
        S c[8];
        lu8Solve(A, rp, cp, f, c);
        // trilinearInterpolation, src/Boundary.cpp:826-833
        return c[0] + c[1] * ip.x + c[2] * ip.y + c[3] * ip.z + c[4] * ip.x * ip.y + c[5] * ip.x * ip.z + c[6] * ip.y * ip.z
             + c[7] * ip.x * ip.y * ip.z;
    }

    // ---------------------------------------------------------------- immersed BC, src/Boundary.cpp:427-457, 618-883
    void applyImmersed(const Body& B) {
        const std::vector<size_t>& filterNodes = B.filterNodes;
        // filterClosestFluidNodes (two-phase), src/Boundary.cpp:618-652
        std::vector<R> fr, fe;
        fr.reserve(filterNodes.size());
        fe.reserve(filterNodes.size());
        const size_t sI = NJ * NK, sJ = NK;
        for (size_t n : filterNodes) {
            auto& r = cons[0];
            auto& E = cons[4];
            fr.push_back((6 * r[n] + r[n + sI] + r[n - sI] + r[n + sJ] + r[n - sJ] + r[n + 1] + r[n - 1]) / 12);
            fe.push_back((6 * E[n] + E[n + sI] + E[n - sI] + E[n + sJ] + E[n - sJ] + E[n + 1] + E[n - 1]) / 12);
        }
        for (size_t q = 0; q < filterNodes.size(); ++q) { cons[0][filterNodes[q]] = fr[q]; cons[4][filterNodes[q]] = fe[q]; }
        for (const Ghost& g : B.ghosts) {
            Box b = surrounding(g.ip);
            R val[5][8];
            double X[8], Y[8], Z[8];
            bool gflag[8] = {false, false, false, false, false, false, false, false};
            bool allFluid = true;
            V3 normals[8];
            int nNormals = 0;
            int c = 0;
            for (size_t n : corners(b)) {  // setInterpolationValues, src/Boundary.cpp:697-728
                uint8_t t = nodeType[n];
                if (t == FluidInterior || t == FluidEdge || t == FluidGhost) {
                    for (int q = 0; q < 5; ++q) val[q][c] = cons[q][n];
                    U3 s = ijk(n);
                    V3 p = pos(s.i, s.j, s.k);
                    X[c] = p.x; Y[c] = p.y; Z[c] = p.z;
                } else if (t == SolidGhost) {
                    gflag[c] = true;
                    allFluid = false;
                    // ghostNodeMap.at (src/Boundary.cpp:681) finds only this body's ghosts
                    auto it = B.ghostMap.find(n);
                    if (it == B.ghostMap.end()) throw GeometryError("image-point box holds a ghost of another body");
                    const Ghost& o = B.ghosts[it->second];
                    V3 nrm = o.ip - o.bi;
                    nrm = nrm / length(nrm);
                    normals[nNormals++] = nrm;
                    for (int q = 0; q < 5; ++q) val[q][c] = 0;
                    X[c] = o.bi.x; Y[c] = o.bi.y; Z[c] = o.bi.z;
                } else {
                    throw GeometryError("Impossible situation. Found solid node around image point.");
                }
                ++c;
            }
            R ipv[5];
            if (allFluid) {  // simplifiedInterpolation, src/Boundary.cpp:732-757
                V3 lo = pos(b.iMin, b.jMin, b.kMin);
                V3 rel = g.ip - lo;
                for (int q = 0; q < 5; ++q) {
                    const R* f = val[q];
                    R a01 = f[0] + (f[1] - f[0]) * rel.z / h.z;
                    R a23 = f[2] + (f[3] - f[2]) * rel.z / h.z;
                    R a45 = f[4] + (f[5] - f[4]) * rel.z / h.z;
                    R a67 = f[6] + (f[7] - f[6]) * rel.z / h.z;
                    R a0123 = a01 + (a23 - a01) * rel.y / h.y;
                    R a4567 = a45 + (a67 - a45) * rel.y / h.y;
                    ipv[q] = a0123 + (a4567 - a0123) * rel.x / h.x;
                }
            } else {  // Vandermonde, src/Boundary.cpp:788-849
                double VD[64], VN[64];
                for (int r = 0; r < 8; ++r) {
                    double row[8] = {1, X[r], Y[r], Z[r], X[r] * Y[r], X[r] * Z[r], Y[r] * Z[r], X[r] * Y[r] * Z[r]};
                    for (int q = 0; q < 8; ++q) VD[r * 8 + q] = VN[r * 8 + q] = row[q];
                }
                int cn = 0;
                for (int r = 0; r < 8; ++r)
                    if (gflag[r]) {
                        V3 nv = normals[cn++];
                        VN[r * 8 + 0] = 0;
                        VN[r * 8 + 1] = nv.x;
                        VN[r * 8 + 2] = nv.y;
                        VN[r * 8 + 3] = nv.z;
                        VN[r * 8 + 4] = X[r] * nv.y + Y[r] * nv.x;
                        VN[r * 8 + 5] = X[r] * nv.z + Z[r] * nv.x;
                        VN[r * 8 + 6] = Y[r] * nv.z + Z[r] * nv.y;
                        VN[r * 8 + 7] = X[r] * Y[r] * nv.z + X[r] * Z[r] * nv.y + Y[r] * Z[r] * nv.x;
                    }
                ipv[0] = lu8SolveDot(VN, val[0], g.ip);
                ipv[1] = lu8SolveDot(VD, val[1], g.ip);
                ipv[2] = lu8SolveDot(VD, val[2], g.ip);
                ipv[3] = lu8SolveDot(VD, val[3], g.ip);
                ipv[4] = lu8SolveDot(VN, val[4], g.ip);
            }
            Cons gc{ipv[0], -ipv[1], -ipv[2], -ipv[3], ipv[4]};  // getGhostNodeBCVariables, src/Boundary.cpp:777-785
            Prim gq = derivePrim(gc);
            Trans gt = deriveTrans(gq);
            setNode(id(g.idx.i, g.idx.j, g.idx.k), gc, gq, gt);
        }
    }

    // Mesh::applyAllBoundaryConditions, src/Mesh.cpp:155-167
    void applyAllBCs(double tnow) {
        for (const Edge& e : edges) applyEdge(e, tnow);
        for (const Body& B : bodies) applyImmersed(B);
    }

    // ---------------------------------------------------------------- filter, src/Mesh.cpp:85-140
    bool filterNow() const {
        size_t counter = 0;
        for (int q = 0; q < P.n_filter_change_times; ++q)
            if (t > P.filter_change_times[q]) ++counter;
        if (counter < (size_t)P.n_filter_intervals) {
            uint32_t iv = P.filter_intervals[counter];
            if (iv > 0 && (timeLevel + 1) % iv == 0) return true;
        }
        return false;
    }
    void applyRhoFilter() {
        if (!filterNow()) return;
        auto& fv = cons[0];
        auto& tmp = consOld[0];
        for (const Edge& e : edges) forOwned(e.owned, [&](size_t i, size_t j, size_t kk) { tmp[id(i, j, kk)] = fv[id(i, j, kk)]; });
        for (size_t n : solidGhost) tmp[n] = fv[n];
        const size_t sI = NJ * NK, sJ = NK;
        for (size_t n : interior)
            tmp[n] = (6 * fv[n] + fv[n + sI] + fv[n - sI] + fv[n + sJ] + fv[n - sJ] + fv[n + 1] + fv[n - 1]) / 12;
        std::swap(fv, tmp);
    }

    // ---------------------------------------------------------------- one step, src/Solver.cpp:130-167
    void march() {
        for (int q = 0; q < 5; ++q)  // checkFinity, src/Mesh.h:48-61 (prints only; recorded here)
            for (const R& x : cons[q]) if (!std::isfinite(valueOf(x))) { nonFinite = true; break; }
        applyRhoFilter();
        std::swap(cons, consOld);  // swapConservedVariables, src/Mesh.cpp:144-151
        slopes(consOld, k[0]);
        intermediate(k[0], dt / 2);
        updatePrimitives();
        applyAllBCs(t + dt / 2);
        slopes(cons, k[1]);
        intermediate(k[1], dt / 2);
        updatePrimitives();
        applyAllBCs(t + dt / 2);
        slopes(cons, k[2]);
        intermediate(k[2], dt);
        updatePrimitives();
        applyAllBCs(t + dt);
        slopes(cons, k[3]);
        finalStep();
        updatePrimitives();
        applyAllBCs(t + dt);
        storeNorms();
        t += dt;
        ++timeLevel;
        updateTimeStepSize(t);
    }
    // storeNormsOfChange / getNormOfChange, src/Solver.cpp:510-530, src/Mesh.cpp:171-178
    double normOfChange(int q) const {
        double s = 0;
        for (size_t n : interior) { double d = valueOf(consOld[q][n]) - valueOf(cons[q][n]); s += d * d; }
        return std::sqrt(h.x * h.y * h.z * s);
    }
    void storeNorms() {
        if (P.conv_history != 0) norms[0].push_back(normOfChange(0));
        if (P.conv_history == 2) for (int q = 1; q < 5; ++q) norms[q].push_back(normOfChange(q));
    }

    void initialize() {  // Solver::initialize, src/Solver.cpp:22-28
        setupBoundaries();
        setupBodies();
        categorizeNodes();
        applyIC();
        updateTimeStepSize(0);
    }

    const std::vector<R>& field(int which) const {
        if (which < 5) return cons[which];
        if (which < 10) return prim[which - 5];
        return trans[which - 10];
    }

    ns_config P;
    R Re = P.Re;  // the Reynolds number of the viscosity law (a tangent may be set on it)
    size_t NI, NJ, NK, N;
    V3 h{0, 0, 0}, off{0, 0, 0};
    std::array<std::vector<R>, 5> cons, consOld, prim;
    std::array<std::vector<R>, 2> trans;
    std::array<std::array<std::vector<R>, 5>, 4> k;
    std::vector<uint8_t> nodeType;
    std::vector<size_t> interior, solidGhost;
    std::array<std::vector<double>, 5> norms;
    double t = 0, dt = 0;
    uint64_t timeLevel = 0;
    bool nonFinite = false;
    bool crPow = false;
    std::string err;
};
using Oracle = OracleT<double>;
using TangentOracle = OracleT<Dual>;

int fail(Oracle* o, const std::exception& e, int code) {
    if (o) o->err = e.what();
    return code;
}

}  // namespace

extern "C" {

int ns_oracle_create(const ns_config* cfg, void** out) {
    if (!cfg || !out) return NS_ERR_ARGUMENT;
    try {
        *out = new Oracle(*cfg);
        return NS_OK;
    } catch (const std::exception&) {
        *out = nullptr;
        return NS_ERR_CONFIG;
    }
}

int ns_oracle_initialize(void* h) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try { o->initialize(); return NS_OK; }
    catch (const GeometryError& e) { return fail(o, e, NS_ERR_GEOMETRY); }
    catch (const std::exception& e) { return fail(o, e, NS_ERR_CONFIG); }
}

int ns_oracle_march(void* h, int nsteps) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try {
        for (int s = 0; s < nsteps; ++s) o->march();
        return o->nonFinite ? NS_ERR_NONFINITE : NS_OK;
    } catch (const GeometryError& e) { return fail(o, e, NS_ERR_GEOMETRY); }
    catch (const std::exception& e) { return fail(o, e, NS_ERR_STATE); }
}

/* Simulation::run loop (src/Simulation.cpp:19-52): march until the stopping criterion holds. */
int ns_oracle_run(void* h) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try {
        for (;;) {
            bool stop = o->P.stop_criterion == 0 ? !(o->timeLevel < o->P.stop_time_level) : !(o->t < o->P.t_end);
            if (stop) break;
            o->march();
        }
        return o->nonFinite ? NS_ERR_NONFINITE : NS_OK;
    } catch (const GeometryError& e) { return fail(o, e, NS_ERR_GEOMETRY); }
    catch (const std::exception& e) { return fail(o, e, NS_ERR_STATE); }
}

void ns_oracle_state(void* h, double* t, uint64_t* level, double* dt) {
    auto* o = (Oracle*)h;
    if (t) *t = o->t;
    if (level) *level = o->timeLevel;
    if (dt) *dt = o->dt;
}

void ns_oracle_geometry(void* h, double* spacing, double* offset) {
    auto* o = (Oracle*)h;
    spacing[0] = o->h.x; spacing[1] = o->h.y; spacing[2] = o->h.z;
    offset[0] = o->off.x; offset[1] = o->off.y; offset[2] = o->off.z;
}

/* which: 0-4 conserved (current), 5-9 u,v,w,p,T, 10-11 mu,kappa, 12-16 conserved "Old".
 * Output x-fastest: out[(k*NJ + j)*NI + i]. */
int ns_oracle_get_field(void* h, int which, double* out) {
    auto* o = (Oracle*)h;
    if (!o || which < 0 || which > 16) return NS_ERR_ARGUMENT;
    const std::vector<double>& f = which < 12 ? o->field(which) : o->consOld[which - 12];
    for (size_t i = 0; i < o->NI; ++i)
        for (size_t j = 0; j < o->NJ; ++j)
            for (size_t kk = 0; kk < o->NK; ++kk) out[(kk * o->NJ + j) * o->NI + i] = f[o->id(i, j, kk)];
    return NS_OK;
}

/* Overwrite one field (same ids and layout as ns_oracle_get_field). */
int ns_oracle_set_field(void* h, int which, const double* in) {
    auto* o = (Oracle*)h;
    if (!o || which < 0 || which > 16) return NS_ERR_ARGUMENT;
    std::vector<double>* f;
    if (which < 5) f = &o->cons[which];
    else if (which < 10) f = &o->prim[which - 5];
    else if (which < 12) f = &o->trans[which - 10];
    else f = &o->consOld[which - 12];
    for (size_t i = 0; i < o->NI; ++i)
        for (size_t j = 0; j < o->NJ; ++j)
            for (size_t kk = 0; kk < o->NK; ++kk) (*f)[o->id(i, j, kk)] = in[(kk * o->NJ + j) * o->NI + i];
    return NS_OK;
}

/* Derive primitives and transport at every node from the current conserved fields (deriveX,
 * src/FlowVariableGroupStructs.h:166-185) and recompute dt at t. Used after a custom field upload. */
int ns_oracle_derive_all(void* h) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    for (size_t n = 0; n < o->N; ++n) {
        Oracle::Cons c{o->cons[0][n], o->cons[1][n], o->cons[2][n], o->cons[3][n], o->cons[4][n]};
        Oracle::Prim q = o->derivePrim(c);
        Oracle::Trans tr = o->deriveTrans(q);
        o->setNode(n, c, q, tr);
    }
    o->updateTimeStepSize(o->t);
    return NS_OK;
}

/* RK4 slopes of the current conserved state with the current primitive arrays (computeRK4slopes,
 * src/Solver.cpp:170-179); out: 5 x-fastest arrays, zero at non-interior nodes. */
int ns_oracle_slopes(void* h, double* out) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    std::array<std::vector<double>, 5> K;
    for (auto& a : K) a.assign(o->N, 0.0);
    o->slopes(o->cons, K);
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < o->NI; ++i)
            for (size_t j = 0; j < o->NJ; ++j)
                for (size_t kk = 0; kk < o->NK; ++kk)
                    out[(size_t)q * o->N + (kk * o->NJ + j) * o->NI + i] = K[q][o->id(i, j, kk)];
    return NS_OK;
}

/* Apply all boundary conditions at time tnow to the current arrays (src/Mesh.cpp:155-167). */
int ns_oracle_apply_bcs(void* h, double tnow) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try { o->applyAllBCs(tnow); return NS_OK; }
    catch (const GeometryError& e) { return fail(o, e, NS_ERR_GEOMETRY); }
}

int ns_oracle_get_node_types(void* h, uint8_t* out) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    for (size_t i = 0; i < o->NI; ++i)
        for (size_t j = 0; j < o->NJ; ++j)
            for (size_t kk = 0; kk < o->NK; ++kk) out[(kk * o->NJ + j) * o->NI + i] = o->nodeType[o->id(i, j, kk)];
    return NS_OK;
}

void ns_oracle_counts(void* h, int64_t* nGhost, int64_t* nFilter, int64_t* nInterior) {
    auto* o = (Oracle*)h;
    if (nGhost) *nGhost = (int64_t)o->ghosts.size();
    if (nFilter) *nFilter = (int64_t)o->filterNodes.size();
    if (nInterior) *nInterior = (int64_t)o->interior.size();
}

/* Ghosts in list order (initial ghosts ascending, then augmentation order). */
int ns_oracle_get_ghosts(void* h, int32_t* ijk, double* bi, double* ip, int32_t* box) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try {
        for (size_t g = 0; g < o->ghosts.size(); ++g) {
            const Ghost& gh = o->ghosts[g];
            if (ijk) { ijk[3 * g] = (int32_t)gh.idx.i; ijk[3 * g + 1] = (int32_t)gh.idx.j; ijk[3 * g + 2] = (int32_t)gh.idx.k; }
            if (bi) { bi[3 * g] = gh.bi.x; bi[3 * g + 1] = gh.bi.y; bi[3 * g + 2] = gh.bi.z; }
            if (ip) { ip[3 * g] = gh.ip.x; ip[3 * g + 1] = gh.ip.y; ip[3 * g + 2] = gh.ip.z; }
            if (box) {
                Box b = o->surrounding(gh.ip);
                box[3 * g] = (int32_t)b.iMin; box[3 * g + 1] = (int32_t)b.jMin; box[3 * g + 2] = (int32_t)b.kMin;
            }
        }
        return NS_OK;
    } catch (const GeometryError& e) { return fail(o, e, NS_ERR_GEOMETRY); }
}

int ns_oracle_get_filter_nodes(void* h, int32_t* ijk) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    for (size_t q = 0; q < o->filterNodes.size(); ++q) {
        U3 s = o->ijk(o->filterNodes[q]);
        ijk[3 * q] = (int32_t)s.i; ijk[3 * q + 1] = (int32_t)s.j; ijk[3 * q + 2] = (int32_t)s.k;
    }
    return NS_OK;
}

int64_t ns_oracle_norm_history(void* h, int var, double* out, int64_t n) {
    auto* o = (Oracle*)h;
    if (!o || var < 0 || var > 4) return -1;
    const auto& v = o->norms[var];
    for (int64_t q = 0; q < n && q < (int64_t)v.size(); ++q) out[q] = v[q];
    return (int64_t)v.size();
}

const char* ns_oracle_last_error(void* h) { return h ? ((Oracle*)h)->err.c_str() : "null handle"; }

void ns_oracle_destroy(void* h) { delete (Oracle*)h; }

/* Unit hook: set t and timeLevel (the periodic filter's schedule reads both, src/Mesh.cpp:85-104). */
int ns_oracle_set_time(void* h, double t, uint64_t level) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    o->t = t;
    o->timeLevel = level;
    return NS_OK;
}

/* Unit hook: Mesh::checkFilterCondition (src/Mesh.cpp:85-104) at the current t and timeLevel. */
int ns_oracle_filter_now(void* h) {
    auto* o = (Oracle*)h;
    return o ? (o->filterNow() ? 1 : 0) : -1;
}

/* Unit hook: Mesh::applyFilter_ifAppropriate (src/Mesh.cpp:107-140) on the current density, as the
 * step does it before the buffer swap; returns 1 when the schedule fired. */
int ns_oracle_rho_filter(void* h) {
    auto* o = (Oracle*)h;
    if (!o) return -1;
    bool f = o->filterNow();
    o->applyRhoFilter();
    return f ? 1 : 0;
}

/* Unit hook: Mesh::getNormOfChange (src/Mesh.cpp:171-178) of variable q between the "Old" and the
 * current conserved arrays. */
double ns_oracle_norm_of_change(void* h, int q) {
    auto* o = (Oracle*)h;
    if (!o || q < 0 || q > 4) return -1;
    return o->normOfChange(q);
}

/* Unit hook: Solver::updateTimeStepSize (src/Solver.cpp:80-127) at tnow from the current arrays,
 * without deriving anything first; returns the new dt. */
double ns_oracle_update_dt(void* h, double tnow) {
    auto* o = (Oracle*)h;
    if (!o) return -1;
    o->updateTimeStepSize(tnow);
    return o->dt;
}

/* Selects x^1.5 in the viscosity law: 0 = std::pow (the reference's call), 1 = correctly rounded. */
int ns_oracle_set_cr_pow(void* h, int on) {
    auto* o = (Oracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    o->crPow = on != 0;
    return NS_OK;
}

/* Unit hook: x^1.5 as the oracle evaluates it (mode 0 std::pow, 1 correctly rounded). */
/* Prescribed step sizes: the step from time level n takes dts[n] (n < len); also sets the current dt. */
int ns_oracle_set_dt_schedule(void* h, const double* dts, int64_t n) {
    auto* o = (Oracle*)h;
    if (!o || n < 0) return NS_ERR_ARGUMENT;
    o->dtSchedule.assign(dts, dts + n);
    if (o->timeLevel < o->dtSchedule.size()) o->dt = o->dtSchedule[o->timeLevel];
    return NS_OK;
}

double ns_oracle_pow15(double x, int mode) { return mode ? Oracle::pow15cr(x) : std::pow(x, 1.5); }

/* Unit hook: 8x8 full-pivot LU solve of A x = b (A row-major). */
void ns_oracle_lu8_solve(const double* A, const double* b, double* x) {
    double LU[64];
    std::memcpy(LU, A, sizeof(LU));
    int rp[8], cp[8];
    Oracle::lu8Factor(LU, rp, cp);
    Oracle::lu8Solve(LU, rp, cp, b, x);
}

/* Unit hook: the perturbation generator (uniform01 of the IC). */
double ns_oracle_uniform01(uint64_t seed, uint64_t counter) { return uniform01(seed, counter); }

/* ---- tangent-linear form: the same set-up and step on (value, tangent) pairs (TangentOracle). Values
 * equal the plain oracle's bit for bit; tangents are the derivatives of the computed step with respect to
 * the state, at the step sizes the run takes (prescribe them with ns_oracle_tl_set_dt_schedule). */
int ns_oracle_tl_create(const ns_config* cfg, void** out) {
    if (!cfg || !out) return NS_ERR_ARGUMENT;
    try {
        *out = new TangentOracle(*cfg);
        return NS_OK;
    } catch (const std::exception&) {
        *out = nullptr;
        return NS_ERR_CONFIG;
    }
}

int ns_oracle_tl_initialize(void* h) {
    auto* o = (TangentOracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try { o->initialize(); return NS_OK; }
    catch (const GeometryError& e) { o->err = e.what(); return NS_ERR_GEOMETRY; }
    catch (const std::exception& e) { o->err = e.what(); return NS_ERR_CONFIG; }
}

/* Conserved state: values and tangents, 5 fields of N values each, x-fastest (as ns_oracle_get_field);
 * primitives and transport are derived at every node from them and dt recomputed (as
 * ns_oracle_set_field + ns_oracle_derive_all). tan may be null (zero tangents). */
int ns_oracle_tl_set_state(void* h, const double* val, const double* tan) {
    auto* o = (TangentOracle*)h;
    if (!o || !val) return NS_ERR_ARGUMENT;
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < o->NI; ++i)
            for (size_t j = 0; j < o->NJ; ++j)
                for (size_t kk = 0; kk < o->NK; ++kk) {
                    const size_t x = (size_t)q * o->N + (kk * o->NJ + j) * o->NI + i;
                    o->cons[q][o->id(i, j, kk)] = Dual(val[x], tan ? tan[x] : 0.0);
                }
    for (size_t n = 0; n < o->N; ++n) {
        TangentOracle::Cons c{o->cons[0][n], o->cons[1][n], o->cons[2][n], o->cons[3][n], o->cons[4][n]};
        TangentOracle::Prim q = o->derivePrim(c);
        o->setNode(n, c, q, o->deriveTrans(q));
    }
    o->updateTimeStepSize(o->t);
    return NS_OK;
}

/* Conserved state out: values and tangents (either may be null), layout as ns_oracle_tl_set_state. */
int ns_oracle_tl_get_state(void* h, double* val, double* tan) {
    auto* o = (TangentOracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < o->NI; ++i)
            for (size_t j = 0; j < o->NJ; ++j)
                for (size_t kk = 0; kk < o->NK; ++kk) {
                    const size_t x = (size_t)q * o->N + (kk * o->NJ + j) * o->NI + i;
                    const Dual& c = o->cons[q][o->id(i, j, kk)];
                    if (val) val[x] = c.v;
                    if (tan) tan[x] = c.d;
                }
    return NS_OK;
}

/* Tangent of the Reynolds number of the viscosity law (src/Solver.cpp:476): with zero state tangents,
 * the tangents of a run are then the derivatives with respect to Re. Set before ns_oracle_tl_set_state. */
int ns_oracle_tl_set_re_tangent(void* h, double dRe) {
    auto* o = (TangentOracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    o->Re = Dual(o->P.Re, dRe);
    return NS_OK;
}

int ns_oracle_tl_set_dt_schedule(void* h, const double* dts, int64_t n) {
    auto* o = (TangentOracle*)h;
    if (!o || n < 0) return NS_ERR_ARGUMENT;
    o->dtSchedule.assign(dts, dts + n);
    if (o->timeLevel < o->dtSchedule.size()) o->dt = o->dtSchedule[o->timeLevel];
    return NS_OK;
}

int ns_oracle_tl_march(void* h, int nsteps) {
    auto* o = (TangentOracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    try {
        for (int s = 0; s < nsteps; ++s) o->march();
        return o->nonFinite ? NS_ERR_NONFINITE : NS_OK;
    } catch (const GeometryError& e) { o->err = e.what(); return NS_ERR_GEOMETRY; }
    catch (const std::exception& e) { o->err = e.what(); return NS_ERR_STATE; }
}

/* RK4 slopes of the current state (as ns_oracle_slopes): values and tangents, 5 x-fastest arrays each. */
int ns_oracle_tl_slopes(void* h, double* val, double* tan) {
    auto* o = (TangentOracle*)h;
    if (!o) return NS_ERR_ARGUMENT;
    std::array<std::vector<Dual>, 5> K;
    for (auto& a : K) a.assign(o->N, Dual());
    o->slopes(o->cons, K);
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < o->NI; ++i)
            for (size_t j = 0; j < o->NJ; ++j)
                for (size_t kk = 0; kk < o->NK; ++kk) {
                    const size_t x = (size_t)q * o->N + (kk * o->NJ + j) * o->NI + i;
                    const Dual& c = K[q][o->id(i, j, kk)];
                    if (val) val[x] = c.v;
                    if (tan) tan[x] = c.d;
                }
    return NS_OK;
}

void ns_oracle_tl_state(void* h, double* t, uint64_t* level, double* dt) {
    auto* o = (TangentOracle*)h;
    if (t) *t = o->t;
    if (level) *level = o->timeLevel;
    if (dt) *dt = o->dt;
}

const char* ns_oracle_tl_last_error(void* h) { return h ? ((TangentOracle*)h)->err.c_str() : "null handle"; }

void ns_oracle_tl_destroy(void* h) { delete (TangentOracle*)h; }

}  // extern "C"
