/*
 * adjoint.hip — reverse-mode (discrete adjoint) kernels of the FP64 time step.
 *
 * The forward stage is  k = R(S, P)  (R = the flux-form right-hand side at FluidInterior nodes, S the
 * stage state after its boundary application, P the pre-filter density and energy at immersed-filter
 * nodes, from which the reference takes the primitive variables there) followed by the combine and the
 * boundary application  S' = BC(X)  (edge faces in application order, immersed filter, ghost update).
 * These kernels apply the transposed Jacobians:
 *
 *   R^T  (two passes, gather form, no atomics): pass A visits every node q, forms the adjoints of the six
 *        face fluxes around q (Phi_d(a,b) enters RHS(a) with + and RHS(b) with -, so its adjoint is
 *        lambda(a) - lambda(b), plus the objective's seed on faces between interior and ghost nodes) and
 *        back-propagates through the three nodal fluxes G_d(q) and q's side of the compact terms; it
 *        stores q's direct conserved adjoints, primitive adjoints and the adjoints of the nine velocity
 *        derivatives at q. Pass B gathers the derivative adjoints of q's six neighbours and
 *        back-propagates through the primitive-variable formula at q.
 *   BC^T: ghost update (linear: frozen trilinear or Vandermonde weights, transposed solve), immersed
 *        filter (linear, two-phase), edge faces in reverse application order (symmetry and periodic
 *        linear; outlet and inlet linearised at the forward values; inlet profile adjoint).
 *
 * The objective's face seed reproduces the force of k_body_force (stage_fused.hip): Fx on the body is
 * -dx*dy*dz * sum over (interior, ghost) faces of +-Phi_x-momentum.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <stdexcept>
#include <utility>
#include <vector>

#include "adjoint.h"
#include "device_common.h"

namespace nsfsi {

namespace {

constexpr int AB = 256;

struct APt {
    double rho, ru, rv, rw, E;  // stage values
    double rp, Ep;              // density and energy the primitives come from (pre-filter at filter nodes)
    double m, u, v, w, p, T, mu;
};

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ __forceinline__ void pointAt(const AdjArgs& a, int i, int j, int k, APt& q) {
    const Layout& L = a.L;
    i = clampi(i, 0, L.NI - 1);
    j = clampi(j, 0, L.NJ - 1);
    k = clampi(k, 0, L.NK - 1);
    const size_t o = L.off(i, j, k);
    const size_t F = L.fstride;
    q.rho = a.S[o];
    q.ru = a.S[o + F];
    q.rv = a.S[o + 2 * F];
    q.rw = a.S[o + 3 * F];
    q.E = a.S[o + 4 * F];
    q.rp = q.rho;
    q.Ep = q.E;
    if (a.W.contains(i, j, k) && (a.cls[o] & CLS_FILTER)) {
        const size_t w = a.W.idx(i, j, k);
        q.rp = a.Ppre[w];
        q.Ep = a.PEpre[w];
    }
    const AdjConst& c = a.c;
    q.m = 1.0 + q.rp;
    const double im = 1.0 / q.m;
    q.u = q.ru * im;
    q.v = q.rv * im;
    q.w = q.rw * im;
    q.p = c.gm1 * (q.Ep - q.m * 0.5 * (q.u * q.u + q.v * q.v + q.w * q.w));
    q.T = (c.gamma * q.p - q.rp) * im;
    const double aa = 1.0 + q.T;
    q.mu = aa * sqrt(aa) * c.ScP1 / (c.Re * (q.T + c.ScP1));
}

__device__ __forceinline__ double lam(const AdjArgs& a, int i, int j, int k, int f) {
    const Layout& L = a.L;
    if (i < 0 || i >= L.NI || j < 0 || j >= L.NJ || k < 0 || k >= L.NK) return 0.0;
    return a.lam[L.off(i, j, k) + (size_t)f * L.fstride];  // zero at non-interior nodes
}

__device__ __forceinline__ int typeAt(const AdjArgs& a, int i, int j, int k) {
    const Layout& L = a.L;
    if (i < 0 || i >= L.NI || j < 0 || j >= L.NJ || k < 0 || k >= L.NK) return -1;
    return a.cls[L.off(i, j, k)] & CLS_TYPE_MASK;
}

// Seed of the objective on the face (lower, upper) along d, x-momentum component only.
__device__ __forceinline__ double faceSeed(const AdjArgs& a, int tl, int tu) {
    if (a.objCoef == 0.0) return 0.0;
    if (tl == NS_FLUID_INTERIOR && tu == NS_SOLID_GHOST) return -a.vol * a.objCoef;  // interior is lower: +Phi
    if (tl == NS_SOLID_GHOST && tu == NS_FLUID_INTERIOR) return a.vol * a.objCoef;   // interior is upper: -Phi
    return 0.0;
}

// Pass A: per node, adjoints of the nodal fluxes and of q's side of the compact terms.
__global__ void k_adj_rhs_a(AdjArgs a) {
    const Layout& L = a.L;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.NI) return;
    const size_t o = L.off(i, j, k);
    const size_t F = L.fstride;
    const AdjConst& c = a.c;
    APt q, nb[6];  // neighbours: -x, +x, -y, +y, -z, +z
    pointAt(a, i, j, k, q);
    pointAt(a, i - 1, j, k, nb[0]);
    pointAt(a, i + 1, j, k, nb[1]);
    pointAt(a, i, j - 1, k, nb[2]);
    pointAt(a, i, j + 1, k, nb[3]);
    pointAt(a, i, j, k - 1, nb[4]);
    pointAt(a, i, j, k + 1, nb[5]);
    // velocity derivatives at q, as the forward evaluates them (clamped neighbours at the domain edge)
    const double dxu = (nb[1].u - nb[0].u) * c.i2d[0], dxv = (nb[1].v - nb[0].v) * c.i2d[0],
                 dxw = (nb[1].w - nb[0].w) * c.i2d[0];
    const double dyu = (nb[3].u - nb[2].u) * c.i2d[1], dyv = (nb[3].v - nb[2].v) * c.i2d[1],
                 dyw = (nb[3].w - nb[2].w) * c.i2d[1];
    const double dzu = (nb[5].u - nb[4].u) * c.i2d[2], dzv = (nb[5].v - nb[4].v) * c.i2d[2],
                 dzw = (nb[5].w - nb[4].w) * c.i2d[2];
    const int tq = typeAt(a, i, j, k);
    double lq[5];
    for (int f = 0; f < 5; ++f) lq[f] = lam(a, i, j, k, f);

    double bru = 0, brv = 0, brw = 0, bE = 0;          // direct conserved adjoints
    double bu = 0, bv = 0, bw = 0, bp = 0, bT = 0, bmu = 0;  // primitive adjoints at q
    double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};          // dxu dxv dxw dyu dyv dyw dzu dzv dzw
    const double h = c.H0 + q.E + q.p;
    const int di[3] = {1, 0, 0}, dj[3] = {0, 1, 0}, dk[3] = {0, 0, 1};
    for (int d = 0; d < 3; ++d) {
        const int im_ = i - di[d], jm = j - dj[d], km = k - dk[d];
        const int ip_ = i + di[d], jp = j + dj[d], kp = k + dk[d];
        const int tm = typeAt(a, im_, jm, km), tp = typeAt(a, ip_, jp, kp);
        double phLo[5], phHi[5], gb[5];
        for (int f = 0; f < 5; ++f) {
            phLo[f] = lam(a, im_, jm, km, f) - lq[f];
            phHi[f] = lq[f] - lam(a, ip_, jp, kp, f);
        }
        phLo[1] += faceSeed(a, tm, tq);
        phHi[1] += faceSeed(a, tq, tp);
        for (int f = 0; f < 5; ++f) gb[f] = c.i2d[d] * (phLo[f] + phHi[f]);
        // nodal flux G_d(q)
        const double vn = d == 0 ? q.u : (d == 1 ? q.v : q.w);
        double t0, t1, t2;
        if (d == 0) {
            t0 = -(2.0 / 3.0) * q.mu * (dyv + dzw); t1 = q.mu * dyu; t2 = q.mu * dzu;
        } else if (d == 1) {
            t0 = q.mu * dxv; t1 = -(2.0 / 3.0) * q.mu * (dxu + dzw); t2 = q.mu * dzv;
        } else {
            t0 = q.mu * dxw; t1 = q.mu * dyw; t2 = -(2.0 / 3.0) * q.mu * (dxu + dyv);
        }
        // g0 = -m_d; g1..g3 = -m_c vn (- p on the normal component) + t; g4 = -h vn + u t0 + v t1 + w t2
        if (d == 0) bru += -gb[0];
        else if (d == 1) brv += -gb[0];
        else brw += -gb[0];
        bru += -gb[1] * vn;
        brv += -gb[2] * vn;
        brw += -gb[3] * vn;
        const double bvn = -gb[1] * q.ru - gb[2] * q.rv - gb[3] * q.rw - gb[4] * h;
        if (d == 0) bu += bvn;
        else if (d == 1) bv += bvn;
        else bw += bvn;
        bp += -gb[1 + d];
        const double bh = -gb[4] * vn;
        bE += bh;
        bp += bh;
        bu += gb[4] * t0;
        bv += gb[4] * t1;
        bw += gb[4] * t2;
        const double b0 = gb[1] + gb[4] * q.u, b1 = gb[2] + gb[4] * q.v, b2 = gb[3] + gb[4] * q.w;
        const double f23 = -(2.0 / 3.0) * q.mu;
        if (d == 0) {
            bmu += b0 * (-(2.0 / 3.0) * (dyv + dzw)) + b1 * dyu + b2 * dzu;
            D[4] += b0 * f23; D[8] += b0 * f23; D[3] += b1 * q.mu; D[6] += b2 * q.mu;
        } else if (d == 1) {
            bmu += b0 * dxv + b1 * (-(2.0 / 3.0) * (dxu + dzw)) + b2 * dzv;
            D[1] += b0 * q.mu; D[0] += b1 * f23; D[8] += b1 * f23; D[7] += b2 * q.mu;
        } else {
            bmu += b0 * dxw + b1 * dyw + b2 * (-(2.0 / 3.0) * (dxu + dyv));
            D[2] += b0 * q.mu; D[5] += b1 * q.mu; D[0] += b2 * f23; D[4] += b2 * f23;
        }
        // compact terms of the two faces along d, q's side only
        const double kn = c.c43[d], kt = c.c1[d];
        const double ku = d == 0 ? kn : kt, kv = d == 1 ? kn : kt, kw = d == 2 ? kn : kt;
        for (int side = 0; side < 2; ++side) {
            const APt& A = side == 0 ? nb[2 * d] : q;      // lower node
            const APt& B = side == 0 ? q : nb[2 * d + 1];  // upper node
            const double* cb = side == 0 ? phLo : phHi;
            const double mm = A.mu + B.mu, du = B.u - A.u, dv = B.v - A.v, dw = B.w - A.w, dT = B.T - A.T;
            const double su = B.mu * B.u + A.mu * A.u, sv = B.mu * B.v + A.mu * A.v, sw = B.mu * B.w + A.mu * A.w;
            const double bmm = cb[1] * ku * du + cb[2] * kv * dv + cb[3] * kw * dw + cb[4] * kt * c.pf * dT;
            const double bdu = cb[1] * ku * mm + cb[4] * ku * su, bdv = cb[2] * kv * mm + cb[4] * kv * sv,
                         bdw = cb[3] * kw * mm + cb[4] * kw * sw, bdT = cb[4] * kt * c.pf * mm;
            const double bsu = cb[4] * ku * du, bsv = cb[4] * kv * dv, bsw = cb[4] * kw * dw;
            const double sg = side == 0 ? 1.0 : -1.0;  // q is the upper node (side 0) or the lower one
            const APt& Q = q;
            bmu += bmm + bsu * Q.u + bsv * Q.v + bsw * Q.w;
            bu += sg * bdu + bsu * Q.mu;
            bv += sg * bdv + bsv * Q.mu;
            bw += sg * bdw + bsw * Q.mu;
            bT += sg * bdT;
        }
    }
    double* out = a.passA;
    out[o] = bru;
    out[o + F] = brv;
    out[o + 2 * F] = brw;
    out[o + 3 * F] = bE;
    out[o + 4 * F] = bu;
    out[o + 5 * F] = bv;
    out[o + 6 * F] = bw;
    out[o + 7 * F] = bp;
    out[o + 8 * F] = bT;
    out[o + 9 * F] = bmu;
    for (int r = 0; r < 9; ++r) out[o + (10 + r) * F] = D[r];
}

__device__ __forceinline__ double passAt(const AdjArgs& a, int i, int j, int k, int f) {
    const Layout& L = a.L;
    if (i < 0 || i >= L.NI || j < 0 || j >= L.NJ || k < 0 || k >= L.NK) return 0.0;
    return a.passA[L.off(i, j, k) + (size_t)f * L.fstride];
}

// Pass B: derivative adjoints of the neighbours, then the primitive-variable formula at q. Adds into
// Sbar (5 fields) and, at filter nodes, into the pre-filter adjoint tables; block partial of dJ/dRe.
__global__ void k_adj_rhs_b(AdjArgs a, double* rePartial) {
    const Layout& L = a.L;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    double reb = 0;
    if (i < L.NI) {
        const size_t o = L.off(i, j, k);
        const size_t F = L.fstride;
        const AdjConst& c = a.c;
        APt q;
        pointAt(a, i, j, k, q);
        double bu = a.passA[o + 4 * F], bv = a.passA[o + 5 * F], bw = a.passA[o + 6 * F];
        double bp = a.passA[o + 7 * F], bT = a.passA[o + 8 * F], bmu = a.passA[o + 9 * F];
        // derivative at node r along d: (x(r + e_d) - x(r - e_d)) * i2d, neighbours clamped at the edge
        const int N[3] = {L.NI, L.NJ, L.NK};
        const int x[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
            const int e[3] = {d == 0, d == 1, d == 2};
            for (int comp = 0; comp < 3; ++comp) {
                const int f = 10 + 3 * d + comp;
                double acc = passAt(a, i - e[0], j - e[1], k - e[2], f) - passAt(a, i + e[0], j + e[1], k + e[2], f);
                // clamped stencils at the first and last node along d use the node itself
                if (x[d] == 0) acc -= a.passA[o + (size_t)f * F];
                if (x[d] == N[d] - 1) acc += a.passA[o + (size_t)f * F];
                // neighbour r with a clamped stencil that reaches q twice (N[d] == 2 never occurs here)
                const double v = acc * c.i2d[d];
                if (comp == 0) bu += v;
                else if (comp == 1) bv += v;
                else bw += v;
            }
        }
        // mu = a^1.5 ScP1 / (Re (T + ScP1)), a = 1 + T
        const double aa = 1.0 + q.T;
        bT += bmu * q.mu * (1.5 / aa - 1.0 / (q.T + c.ScP1));
        reb = bmu * (-q.mu / c.Re);
        // T = (gamma p - rp) / m
        const double im = 1.0 / q.m;
        bp += bT * c.gamma * im;
        double brp = -bT * im;
        double bim = bT * q.T * q.m;
        // p = gm1 (Ep - m K / 2)
        const double K = q.u * q.u + q.v * q.v + q.w * q.w;
        const double bEp = bp * c.gm1;
        double bm = -0.5 * c.gm1 * K * bp;
        bu += -c.gm1 * q.m * q.u * bp;
        bv += -c.gm1 * q.m * q.v * bp;
        bw += -c.gm1 * q.m * q.w * bp;
        // u = ru / m
        const double bru = a.passA[o] + bu * im, brv = a.passA[o + F] + bv * im, brw = a.passA[o + 2 * F] + bw * im;
        bim += bu * q.ru + bv * q.rv + bw * q.rw;
        bm += -bim * im * im;
        brp += bm;
        double* Sb = a.Sbar;
        Sb[o + F] += bru;
        Sb[o + 2 * F] += brv;
        Sb[o + 3 * F] += brw;
        Sb[o + 4 * F] += a.passA[o + 3 * F];
        if (a.W.contains(i, j, k) && (a.cls[o] & CLS_FILTER)) {
            const size_t w = a.W.idx(i, j, k);
            a.Pbar[w] += brp;
            a.PEbar[w] += bEp;
        } else {
            Sb[o] += brp;
            Sb[o + 4 * F] += bEp;
        }
    }
    // block partial of dJ/dRe
    __shared__ double sh[AB / 64];
    reb = wave_sum(reb);
    const int t = threadIdx.x;
    if ((t & 63) == 0) sh[t >> 6] = reb;
    __syncthreads();
    if (t == 0) {
        double s = 0;
        for (int r = 0; r < (int)(blockDim.x >> 6); ++r) s += sh[r];
        rePartial[(size_t)blockIdx.z * gridDim.y * gridDim.x + (size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

// ---- boundary application, transposed


// Ghost update: S[g] = +-(weights . corners); transposed: corners += weights * (+-Sbar[g]), Sbar[g] = 0.
// Phase 1 (here): every ghost writes its 8 corners' contributions (5 fields each) to contrib and clears
// its own adjoint; phase 2 (k_adj_gather5 over the donor plan) adds them up per donor node in a fixed
// order, so the result does not depend on the order the ghosts run in.
__global__ void k_adj_ghost(Layout L, Params P, GhostTables G, int64_t n, double* Sb, double* contrib) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t F = L.fstride;
    const size_t o = (size_t)G.node[g];
    double gb[5];
    for (int f = 0; f < 5; ++f) {
        gb[f] = Sb[o + f * F];
        Sb[o + f * F] = 0.0;
    }
    // adjoint of the interpolated image-point values (the ghost holds rho, -m, rhoE of the image point)
    const double ib[5] = {gb[0], -gb[1], -gb[2], -gb[3], gb[4]};
    const uint8_t mask = G.mask[g];
    double w[2][8];  // [0]: weights of the momentum interpolation, [1]: of density and energy
    if (mask == 0) {
        const double tx = G.rel[3 * g] / P.hx, ty = G.rel[3 * g + 1] / P.hy, tz = G.rel[3 * g + 2] / P.hz;
        for (int c = 0; c < 8; ++c) {
            const double wx = (c >> 2) & 1 ? tx : 1.0 - tx, wy = (c >> 1) & 1 ? ty : 1.0 - ty, wz = c & 1 ? tz : 1.0 - tz;
            w[0][c] = w[1][c] = wx * wy * wz;
        }
    } else {
        // value = b(IP) . c, c = solve(V, f) with V = P^T L U Q^T: weights = P^T L^-T U^-T Q^T b
        const double x = G.ip[3 * g], y = G.ip[3 * g + 1], z = G.ip[3 * g + 2];
        const double b[8] = {1.0, x, y, z, x * y, x * z, y * z, x * y * z};
        for (int s = 0; s < 2; ++s) {
            const double* LU = s == 0 ? G.luD + 64 * g : G.luN + 64 * g;
            const int8_t* rp = G.perm + 32 * g + (s == 0 ? 0 : 16);
            const int8_t* cp = rp + 8;
            double zb[8], yb[8];
            for (int r = 0; r < 8; ++r) zb[r] = b[cp[r]];
            // back substitution reversed (rows 0..7): zv[r] = d_r (yv[r] - sum_{q>r} U[r][q] zv[q])
            for (int r = 0; r < 8; ++r) {
                const double d = LU[r * 8 + r] != 0 ? 1.0 / LU[r * 8 + r] : 0.0;
                const double t = zb[r] * d;
                yb[r] = t;
                for (int qq = r + 1; qq < 8; ++qq) zb[qq] -= LU[r * 8 + qq] * t;
            }
            // forward substitution reversed (rows 7..1): yv[r] -= L[r][q] yv[q], q < r
            for (int r = 7; r >= 1; --r)
                for (int qq = 0; qq < r; ++qq) yb[qq] -= LU[r * 8 + qq] * yb[r];
            for (int r = 0; r < 8; ++r) w[s][rp[r]] = yb[r];
        }
    }
    for (int c = 0; c < 8; ++c) {
        double* out = contrib + (size_t)(8 * g + c) * 5;
        const bool solid = (mask >> c) & 1;  // ghost corners enter as zero (the donor plan leaves them out)
        out[0] = solid ? 0.0 : w[1][c] * ib[0];
        out[1] = solid ? 0.0 : w[0][c] * ib[1];
        out[2] = solid ? 0.0 : w[0][c] * ib[2];
        out[3] = solid ? 0.0 : w[0][c] * ib[3];
        out[4] = solid ? 0.0 : w[1][c] * ib[4];
    }
}

// Phase 2 of a transposed gather: target t (layout offset p.tgt[t]) adds the entries p.start[t] ..
// p.start[t+1]-1 of its list, five fields each, in list order.
__global__ void k_adj_gather5(Layout L, GatherPlan p, const double* contrib, double* Sb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.n) return;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t e = p.start[t]; e < p.start[t + 1]; ++e) {
        const double* c = contrib + (size_t)p.src[e] * 5;
        for (int f = 0; f < 5; ++f) acc[f] += c[f];
    }
    const size_t o = (size_t)p.tgt[t];
    for (int f = 0; f < 5; ++f) Sb[o + f * L.fstride] += acc[f];
}

// Immersed filter, (6 P + sum of 6 neighbours) / 12 transposed, plus the pre-filter table adjoints (the
// tables hold the pre-filter values of the filter nodes). Phase 1: every filter node takes the adjoint of
// its filtered density and energy, clears it, and writes what it gives its own node (6/12 of it plus the
// tables' adjoints) and each of its six neighbours (1/12): tmp[4 q .. 4 q + 3].
__global__ void k_adj_filter_take(Layout L, Window W, const int64_t* fn, int64_t n, double* Sb, double* tmp,
                                  const double* Pb, const double* PEb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t m = fn[q];
    const int i = (int)(m % L.NI), j = (int)((m / L.NI) % L.NJ), k = (int)(m / ((int64_t)L.NI * L.NJ));
    const size_t o = L.off(i, j, k);
    const size_t E4 = 4 * L.fstride;
    const double br = Sb[o] / 12.0, be = Sb[o + E4] / 12.0;
    const size_t w = W.idx(i, j, k);
    tmp[4 * q] = 6.0 * br + Pb[w];
    tmp[4 * q + 1] = 6.0 * be + PEb[w];
    tmp[4 * q + 2] = br;
    tmp[4 * q + 3] = be;
    Sb[o] = 0.0;
    Sb[o + E4] = 0.0;
}

// Phase 2: target t adds, in list order, the entries of its list: 2 q + 1 for filter node q itself,
// 2 q for a filter node q it neighbours.
__global__ void k_adj_filter_gather(Layout L, GatherPlan p, const double* tmp, double* Sb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.n) return;
    double r = 0, e = 0;
    for (int64_t x = p.start[t]; x < p.start[t + 1]; ++x) {
        const int64_t q = p.src[x] >> 1;
        const int own = p.src[x] & 1;
        r += tmp[4 * q + (own ? 0 : 2)];
        e += tmp[4 * q + (own ? 1 : 3)];
    }
    const size_t o = (size_t)p.tgt[t];
    Sb[o] += r;
    Sb[o + 4 * L.fstride] += e;
}

// One edge face, transposed (src/Boundary.cpp:250-417 forward). S holds the forward values after the
// boundary application (linearisation point of the inlet and outlet); qb the inlet profile adjoint. The
// periodic, symmetry and outlet faces map distinct face nodes to distinct source nodes, so each target
// has one contributor per launch; the inlet's in-plane filter spreads onto neighbouring columns and is
// summed by k_adj_inlet_gather.
__global__ void k_adj_edge(Layout L, Params P, EdgeFace f, double* Sb, const double* S, double ts, double* qb,
                           double* wbuf) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    const int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
    const int lo[3] = {f.i0, f.j0, f.k0}, hi[3] = {f.i1, f.j1, f.k1};
    int c[3];
    c[f.axis] = lo[f.axis];
    c[ax1] = lo[ax1] + a;
    c[ax2] = lo[ax2] + b;
    if (c[ax1] > hi[ax1] || c[ax2] > hi[ax2]) return;
    const size_t F = L.fstride;
    const size_t o = L.off(c[0], c[1], c[2]);
    // offsets of the nodes one and two in from the face (x-steps depend on the column: padding slots)
    const int dn = f.side == 0 ? 1 : -1;
    const long st = f.axis == 1 ? (long)L.pitch : (long)L.plane;
    const long s1 = f.axis == 0 ? Layout::xstep(c[0], dn) : dn * st;
    const long s2 = f.axis == 0 ? Layout::xstep(c[0], 2 * dn) : 2 * dn * st;
    double ob[5];
    for (int r = 0; r < 5; ++r) {
        ob[r] = Sb[o + r * F];
        Sb[o + r * F] = 0.0;
    }
    if (f.type == NS_BC_PERIODIC) {
        int cc[3] = {c[0], c[1], c[2]};
        const int Nn = f.axis == 0 ? L.NI : (f.axis == 1 ? L.NJ : L.NK);
        cc[f.axis] = f.side == 0 ? Nn - 2 : 1;
        const size_t src = L.off(cc[0], cc[1], cc[2]);
        for (int r = 0; r < 5; ++r) Sb[src + r * F] += ob[r];
    } else if (f.type == NS_BC_SYMMETRY) {
        // conserved mirror of the node two in: the normal momentum changes sign, all else is copied
        const size_t a2 = o + s2;
        for (int r = 0; r < 5; ++r) Sb[a2 + r * F] += r == 1 + f.axis ? -ob[r] : ob[r];
    } else if (f.type == NS_BC_OUTLET) {
        // r, m extrapolated (2 x1 - x2); rhoE = |m|^2 / (2 (1 + r)) (p' = 0)
        const double r = S[o], mu_ = S[o + F], mv = S[o + 2 * F], mw = S[o + 3 * F];
        const double ir = 1.0 / (1.0 + r);
        const double K = (mu_ * mu_ + mv * mv + mw * mw) * ir * ir;
        const double gr = ob[0] - 0.5 * K * ob[4];
        const double gm[3] = {ob[1] + ob[4] * mu_ * ir, ob[2] + ob[4] * mv * ir, ob[3] + ob[4] * mw * ir};
        const size_t x1 = o + s1, x2 = o + s2;
        Sb[x1] += 2.0 * gr;
        Sb[x2] += -gr;
        for (int r = 0; r < 3; ++r) {
            Sb[x1 + (1 + r) * F] += 2.0 * gm[r];
            Sb[x2 + (1 + r) * F] += -gm[r];
        }
    } else {  // inlet
        const double acc = S[o];  // the boundary density equals the filtered extrapolation
        double qf = 1.0;
        size_t qi = 0;
        if (P.inflow) {
            qi = (size_t)c[ax2] * (ax1 == 0 ? L.NI : L.NJ) + c[ax1];
            qf = P.inflow[qi];
        }
        const double ramp = fmin(1., ts / P.rampTime) * P.M_0 * (f.side == 1 ? -1.0 : 1.0);
        const double vel = ramp * qf;
        const double gammaGm1 = P.gamma * (P.gamma - 1);
        // rho = acc, m_n = (1 + acc) vel, rhoE = acc / (gamma (gamma - 1)) + (1 + acc) vel^2 / 2
        const double bacc = ob[0] + ob[1 + f.axis] * vel + ob[4] * (1.0 / gammaGm1 + 0.5 * vel * vel);
        const double bvel = ob[1 + f.axis] * (1.0 + acc) + ob[4] * (1.0 + acc) * vel;
        if (qb && P.inflow) qb[qi] += bvel * ramp;  // one face node per profile entry: no race
        else if (qb) qb[(size_t)c[ax2] * (ax1 == 0 ? L.NI : L.NJ) + c[ax1]] += bvel * ramp;
        // acc = (sum_nb extrap(nb) + n extrap(o)) / (2 n), extrap(x) = 2 rho(x + s1) - rho(x + 2 s1): the
        // columns of this node and its in-plane neighbours take their shares in k_adj_inlet_gather
        const int n1 = hi[ax1] - lo[ax1] + 1;
        wbuf[(size_t)(c[ax2] - lo[ax2]) * n1 + (c[ax1] - lo[ax1])] = bacc;
    }
}

// Inlet, phase 2: the column of face node o (the nodes one and two in) gathers, in a fixed order, the
// extrapolation adjoints that the in-plane filter of o and of each of its in-plane neighbours p spread
// onto it: bacc(o) n_o / (2 n_o) and bacc(p) / (2 n_p), n the node's in-plane neighbour count.
__global__ void k_adj_inlet_gather(Layout L, EdgeFace f, double* Sb, const double* wbuf) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    const int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
    const int lo[3] = {f.i0, f.j0, f.k0}, hi[3] = {f.i1, f.j1, f.k1};
    const int n1 = hi[ax1] - lo[ax1] + 1, n2 = hi[ax2] - lo[ax2] + 1;
    if (a >= n1 || b >= n2) return;
    auto count = [&](int u, int v) {
        return (u > 0) + (u < n1 - 1) + (v > 0) + (v < n2 - 1);
    };
    const int no = count(a, b);
    double total = wbuf[(size_t)b * n1 + a] * (double)no / (2.0 * no);
    const int du[4] = {1, -1, 0, 0}, dv[4] = {0, 0, 1, -1};
    for (int r = 0; r < 4; ++r) {
        const int u = a + du[r], v = b + dv[r];
        if (u < 0 || u >= n1 || v < 0 || v >= n2) continue;
        total += wbuf[(size_t)v * n1 + u] / (2.0 * count(u, v));
    }
    int c[3];
    c[f.axis] = lo[f.axis];
    c[ax1] = lo[ax1] + a;
    c[ax2] = lo[ax2] + b;
    const size_t o = L.off(c[0], c[1], c[2]);
    const int dn = f.side == 0 ? 1 : -1;
    const long st = f.axis == 1 ? (long)L.pitch : (long)L.plane;
    const long s1 = f.axis == 0 ? Layout::xstep(c[0], dn) : dn * st;
    const long s2 = f.axis == 0 ? Layout::xstep(c[0], 2 * dn) : 2 * dn * st;
    Sb[o + s1] += 2.0 * total;
    Sb[o + s2] += -total;
}

// Routing after a stage's transposed boundary application (Xbar in B): interior nodes feed the combine
// (Ubar += B, each listed stage adjoint += coef * B); non-interior nodes were copies of the previous
// stage state (prev += B).
__global__ void k_adj_route(Layout L, const uint8_t* cls, const double* B, double* Ub, double* l0, double c0, double* l1,
                            double c1, double* l2, double c2, double* l3, double c3, double* prev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.NI) return;
    const size_t o = L.off(i, j, k);
    const bool interior = (cls[o] & CLS_TYPE_MASK) == NS_FLUID_INTERIOR;
    for (int f = 0; f < 5; ++f) {
        const size_t x = o + (size_t)f * L.fstride;
        const double v = B[x];
        if (interior) {
            Ub[x] += v;
            if (l0) l0[x] += c0 * v;
            if (l1) l1[x] += c1 * v;
            if (l2) l2[x] += c2 * v;
            if (l3) l3[x] += c3 * v;
        } else if (prev) {
            prev[x] += v;
        }
    }
}

__global__ void k_axpy_window(double* y, const double* x, size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] += x[q];
}

// Adds a window table of pre-filter adjoints into a state adjoint at the filter nodes (used where the
// tables are plain copies of the state: the initial state).
__global__ void k_adj_tables_to_state(Layout L, Window W, const int64_t* fn, int64_t n, const double* Pb,
                                      const double* PEb, double* Sb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t m = fn[q];
    const int i = (int)(m % L.NI), j = (int)((m / L.NI) % L.NJ), k = (int)(m / ((int64_t)L.NI * L.NJ));
    const size_t o = L.off(i, j, k);
    const size_t w = W.idx(i, j, k);
    Sb[o] += Pb[w];
    Sb[o + 4 * L.fstride] += PEb[w];
}

__global__ void k_widen(const float* x, double* y, size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = (double)x[q];
}

inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

AdjConst makeAdjConst(const Params& P) {
    AdjConst c;
    const double h[3] = {P.hx, P.hy, P.hz};
    for (int d = 0; d < 3; ++d) {
        c.i2d[d] = 1 / (2 * h[d]);
        c.c43[d] = 2 / (3 * h[d] * h[d]);
        c.c1[d] = 1 / (2 * h[d] * h[d]);
    }
    c.gamma = P.gamma;
    c.gm1 = P.gamma - 1;
    c.H0 = 1 / (P.gamma - 1);
    c.ScP1 = P.ScPlusOne;
    c.Re = P.Re;
    c.pf = 1 / ((P.gamma - 1) * P.Pr);
    return c;
}

int adjRhsPartials(const Layout& L) { return (int)nb(L.NI, 64) * L.NJ * L.NK; }

void launchAdjRhs(hipStream_t s, const AdjArgs& a, double* rePartial) {
    const Layout& L = a.L;
    dim3 blk(64), grd(nb(L.NI, 64), L.NJ, L.NK);
    hipLaunchKernelGGL(k_adj_rhs_a, grd, blk, 0, s, a);
    hipLaunchKernelGGL(k_adj_rhs_b, grd, blk, 0, s, a, rePartial);
}

void launchAdjGhost(hipStream_t s, const Layout& L, const Params& P, const GhostTables& G, int64_t n, double* Sb,
                    double* contrib, const GatherPlan& donors) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_adj_ghost, dim3(nb(n, 64)), dim3(64), 0, s, L, P, G, n, Sb, contrib);
    if (donors.n > 0) hipLaunchKernelGGL(k_adj_gather5, dim3(nb(donors.n, 128)), dim3(128), 0, s, L, donors, contrib, Sb);
}

void launchAdjFilter(hipStream_t s, const Layout& L, const Window& W, const int64_t* fn, int64_t n, double* Sb,
                     double* tmp, const double* Pb, const double* PEb, const GatherPlan& targets) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_adj_filter_take, dim3(nb(n, 128)), dim3(128), 0, s, L, W, fn, n, Sb, tmp, Pb, PEb);
    hipLaunchKernelGGL(k_adj_filter_gather, dim3(nb(targets.n, 128)), dim3(128), 0, s, L, targets, tmp, Sb);
}

void launchAdjEdge(hipStream_t s, const Layout& L, const Params& P, const EdgeFace& f, double* Sb, const double* S,
                   double ts, double* qb, double* wbuf) {
    const int ax1 = f.axis == 0 ? 1 : 0, ax2 = f.axis == 2 ? 1 : 2;
    const int lo[3] = {f.i0, f.j0, f.k0}, hi[3] = {f.i1, f.j1, f.k1};
    const int n1 = hi[ax1] - lo[ax1] + 1, n2 = hi[ax2] - lo[ax2] + 1;
    if (n1 <= 0 || n2 <= 0) return;
    if (f.type == NS_BC_INLET && !wbuf) throw std::invalid_argument("inlet adjoint needs its face buffer");
    hipLaunchKernelGGL(k_adj_edge, dim3(nb(n1, 64), n2), dim3(64), 0, s, L, P, f, Sb, S, ts, qb, wbuf);
    if (f.type == NS_BC_INLET)
        hipLaunchKernelGGL(k_adj_inlet_gather, dim3(nb(n1, 64), n2), dim3(64), 0, s, L, f, Sb, (const double*)wbuf);
}

GatherPlan buildGatherPlan(std::vector<std::pair<int64_t, int32_t>> pairs, const std::function<void*(size_t)>& alloc) {
    std::sort(pairs.begin(), pairs.end());
    std::vector<int64_t> tgt, start;
    std::vector<int32_t> src(pairs.size());
    for (size_t e = 0; e < pairs.size(); ++e) {
        if (e == 0 || pairs[e].first != pairs[e - 1].first) {
            tgt.push_back(pairs[e].first);
            start.push_back((int64_t)e);
        }
        src[e] = pairs[e].second;
    }
    start.push_back((int64_t)pairs.size());
    GatherPlan p;
    p.n = (int64_t)tgt.size();
    p.tgt = (int64_t*)alloc(std::max<size_t>(1, tgt.size()) * sizeof(int64_t));
    p.start = (int64_t*)alloc(start.size() * sizeof(int64_t));
    p.src = (int32_t*)alloc(std::max<size_t>(1, src.size()) * sizeof(int32_t));
    if (!tgt.empty()) {
        if (hipMemcpy(p.tgt, tgt.data(), tgt.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p.src, src.data(), src.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("gather plan upload failed");
    }
    if (hipMemcpy(p.start, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("gather plan upload failed");
    return p;
}

void launchAdjRoute(hipStream_t s, const Layout& L, const uint8_t* cls, const double* B, double* Ub, double* const lam[4],
                    const double coef[4], double* prev) {
    dim3 blk(64), grd(nb(L.NI, 64), L.NJ, L.NK);
    hipLaunchKernelGGL(k_adj_route, grd, blk, 0, s, L, cls, B, Ub, lam[0], coef[0], lam[1], coef[1], lam[2], coef[2],
                       lam[3], coef[3], prev);
}

void launchAxpyWindow(hipStream_t s, double* y, const double* x, size_t n) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_window, dim3(nb((int64_t)n, 256)), dim3(256), 0, s, y, x, n);
}

void launchAdjTablesToState(hipStream_t s, const Layout& L, const Window& W, const int64_t* fn, int64_t n,
                            const double* Pb, const double* PEb, double* Sb) {
    if (n > 0)
        hipLaunchKernelGGL(k_adj_tables_to_state, dim3(nb(n, 128)), dim3(128), 0, s, L, W, fn, n, Pb, PEb, Sb);
}

void launchWiden(hipStream_t s, const float* x, double* y, size_t n) {
    if (n > 0) hipLaunchKernelGGL(k_widen, dim3(nb((int64_t)n, 256)), dim3(256), 0, s, x, y, n);
}

}  // namespace nsfsi
