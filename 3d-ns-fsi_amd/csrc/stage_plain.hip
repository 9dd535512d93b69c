/*
 * stage_plain.hip — the unfused RK4 stage: a primitive-variable pass (7 stored fields) followed by a
 * per-node 19-point right-hand side that follows the reference's flux expressions term by term
 * (compute_RK4_step_*, src/Solver.cpp:181-418) and the stage combine.
 *
 * This is the straightforward "plain" device path: one thread per node, neighbours through the cache
 * hierarchy. The fused LDS-tiled stage (stage_fused.hip) is checked against it.
 *
 * Stage combine (3-register form, src/Solver.cpp:422-507): with U the state at t, A the accumulator and
 * S the stage state,
 *   stage 1: S1 = U + dt/2 k1,  A = U + dt/6 k1
 *   stage 2: S2 = U + dt/2 k2,  A += dt/3 k2
 *   stage 3: S3 = U + dt k3,    A += dt/3 k3
 *   stage 4: A += dt/6 k4       (A becomes the state at t+dt)
 * Only FluidInterior nodes are combined; every other node of the output copies the buffer that the
 * reference's "current" array holds at that point (stage 1: A, which still holds the state of the
 * previous step; later stages: the stage input), so boundary kernels see the same stale values the
 * reference sees (ghost values two boundary applications old at stage 1).
 */
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "batch.h"
#include "device_common.h"
#include "stage.h"

namespace nsfsi {

namespace {

// Both kernels run one block of 64 threads per x-segment of the layout, thread t on slot t: the padding
// slots are written too (zeros, never read), so a segment's lines are written whole.
struct SegSlot {
    int i;        // node of the slot
    size_t o;     // its offset
    bool pad;     // padding slot (or past the last node)
    bool write;   // whether the slot lies in the row (padding slots included)
};
__device__ __forceinline__ SegSlot segSlot(const Layout& L, int j, int k) {
    SegSlot r;
    const int t = threadIdx.x;
    r.i = blockIdx.x * Layout::kSegNodes + t - 1;
    r.o = ((size_t)k * L.NJ + j) * L.pitch + (size_t)blockIdx.x * Layout::kSegSlots + t;
    r.pad = t == 0 || t == Layout::kSegSlots - 1 || r.i >= L.NI;
    r.write = t == 0 || r.i < L.NI;  // slot 63 exists when the next segment does
    return r;
}

// MODE 0: every node from S, pre-filter density and energy at filter nodes (the primitives of the stage input
// as the reference holds them there). MODE 1 (stored primitives): updatePrimitiveVariables, src/Solver.cpp:
// 446-479 — FluidInterior nodes only, from S as it leaves the stage combine; the other nodes keep what the
// boundary conditions wrote. MODE 2: every node from S with deriveTransportProperties' conductivity (the
// initial condition, src/Solver.cpp:31-75, and a state set from outside).
template <class T, int MODE>
__global__ void k_derive(Layout L, Params P, const T* S, T* prim, const uint8_t* cls, Window W, const T* rhoPre,
                         const T* rhoEPre) {
    const int j = blockIdx.y, k = blockIdx.z;
    const SegSlot sl = segSlot(L, j, k);
    if (sl.pad) {
        if (sl.write && MODE != 1)
            for (int q = 0; q < 7; ++q) prim[sl.o + q * L.fstride] = (T)0;
        return;
    }
    const int i = sl.i;
    const size_t o = sl.o;
    if (MODE == 1 && (cls[o] & CLS_TYPE_MASK) != NS_FLUID_INTERIOR) return;
    Eos<T> e = Eos<T>::make(P);
    T rho = S[o], rE = S[o + 4 * L.fstride];
    if (MODE == 0 && (cls[o] & CLS_FILTER)) {
        size_t w = W.idx(i, j, k);
        rho = rhoPre[w];
        rE = rhoEPre[w];
    }
    T u, v, w, p, Tt, mu;
    derivePrims(e, rho, S[o + L.fstride], S[o + 2 * L.fstride], S[o + 3 * L.fstride], rE, u, v, w, p, Tt, mu);
    prim[o] = u;
    prim[o + L.fstride] = v;
    prim[o + 2 * L.fstride] = w;
    prim[o + 3 * L.fstride] = p;
    prim[o + 4 * L.fstride] = Tt;
    prim[o + 5 * L.fstride] = mu;
    prim[o + 6 * L.fstride] = MODE == 2 ? mu / e.gm1Pr : mu * e.pf;
}

// LIT (FP64): the reference's own combine. A holds the slope sum k1, k1 + 2 k2, k1 + 2 k2 + 2 k3 at interior
// nodes through stages 1-3 and stage 4 forms U + dt/6 (that sum + k4) (compute_RK4_final_step,
// src/Solver.cpp:494-507), the same operations in the same order as the reference's four stored slopes.
template <class T, int STAGE, bool KAHAN, bool LIT>
__global__ void k_rhs_plain(Layout L, Params P, const T* Sin, const T* prim, const T* U, T* Sout, T* A, const uint8_t* cls,
                            const double* dtdev, const T* fRho, const int32_t* fFlag, int8_t* comp) {
    const bool rhoFiltered = STAGE == 1 && fFlag && *fFlag;
    const int j = blockIdx.y, k = blockIdx.z;
    const SegSlot sl = segSlot(L, j, k);
    const size_t F = L.fstride;
    T* out = STAGE == 4 ? A : Sout;
    if (sl.pad) {
        if (sl.write) {
            for (int q = 0; q < 5; ++q) {
                out[sl.o + q * F] = (T)0;
                A[sl.o + q * F] = (T)0;
                if (KAHAN && STAGE == 4) comp[sl.o + q * F] = 0;
            }
        }
        return;
    }
    const int i = sl.i;
    const size_t o = sl.o;
    if ((cls[o] & CLS_TYPE_MASK) != NS_FLUID_INTERIOR) {
        const T* src = STAGE == 1 ? A : Sin;
        for (int q = 0; q < 5; ++q) out[o + q * F] = src[o + q * F];
        if (rhoFiltered) out[o] = Sin[o];  // unfiltered density (see the fused kernel)
        return;
    }
    const T dx = (T)P.hx, dy = (T)P.hy, dz = (T)P.hz;
    const T n2dx = (T)-1 / ((T)2 * dx), n2dy = (T)-1 / ((T)2 * dy), n2dz = (T)-1 / ((T)2 * dz);
    const T p23xx = (T)2 / ((T)3 * dx * dx), p23yy = (T)2 / ((T)3 * dy * dy), p23zz = (T)2 / ((T)3 * dz * dz);
    const T p12xx = (T)1 / ((T)2 * dx * dx), p12yy = (T)1 / ((T)2 * dy * dy), p12zz = (T)1 / ((T)2 * dz * dz);
    // mixed-derivative factors with the reference's operand order per use (src/Solver.cpp:202-211, 249-258,
    // 296-305, 345-359): 1/(6 dy dx) and 1/(6 dx dy) can round differently
    const T n6ydx = (T)-1 / ((T)6 * dy * dx), n6zdx = (T)-1 / ((T)6 * dz * dx), n6xdy = (T)-1 / ((T)6 * dx * dy);
    const T n6zdy = (T)-1 / ((T)6 * dz * dy), n6xdz = (T)-1 / ((T)6 * dx * dz), n6ydz = (T)-1 / ((T)6 * dy * dz);
    const T p4xdy = (T)1 / ((T)4 * dx * dy), p4xdz = (T)1 / ((T)4 * dx * dz), p4ydx = (T)1 / ((T)4 * dy * dx);
    const T p4ydz = (T)1 / ((T)4 * dy * dz), p4zdx = (T)1 / ((T)4 * dz * dx), p4zdy = (T)1 / ((T)4 * dz * dy);
    const T H0 = (T)(1 / (P.gamma - 1));
    const T* ru = Sin + F;
    const T* rv = Sin + 2 * F;
    const T* rw = Sin + 3 * F;
    const T* E = Sin + 4 * F;
    const T* u = prim;
    const T* v = prim + F;
    const T* w = prim + 2 * F;
    const T* p = prim + 3 * F;
    const T* Tt = prim + 4 * F;
    const T* mu = prim + 5 * F;
    const T* ka = prim + 6 * F;
    // x-steps from column i (a step can cross the padding slots between segments)
    const long XP = Layout::xstep(i, 1), XM = Layout::xstep(i, -1), Y = L.pitch, Z = (long)L.plane;
    const size_t e_ = o + XP, w_ = o + XM, n_ = o + Y, s_ = o - Y, u_ = o + Z, d_ = o - Z;
    T kq[5];
    // continuity (src/Solver.cpp:191-194)
    kq[0] = -(ru[e_] - ru[w_]) / ((T)2 * dx) + -(rv[n_] - rv[s_]) / ((T)2 * dy) + -(rw[u_] - rw[d_]) / ((T)2 * dz);
    // momentum (src/Solver.cpp:225-241, 272-288, 320-337)
    {
        T cX = n2dx * (ru[e_] * u[e_] + p[e_] - ru[w_] * u[w_] - p[w_]);
        T cY = n2dy * (ru[n_] * v[n_] - ru[s_] * v[s_]);
        T cZ = n2dz * (ru[u_] * w[u_] - ru[d_] * w[d_]);
        T vXX = p23xx * ((mu[e_] + mu[o]) * (u[e_] - u[o]) - (mu[w_] + mu[o]) * (u[o] - u[w_]));
        T vYY = p12yy * ((mu[n_] + mu[o]) * (u[n_] - u[o]) - (mu[s_] + mu[o]) * (u[o] - u[s_]));
        T vZZ = p12zz * ((mu[u_] + mu[o]) * (u[u_] - u[o]) - (mu[d_] + mu[o]) * (u[o] - u[d_]));
        T vXY = p4xdy * (mu[n_] * (v[n_ + XP] - v[n_ + XM]) - mu[s_] * (v[s_ + XP] - v[s_ + XM]));
        T vYX = n6ydx * (mu[e_] * (v[e_ + Y] - v[e_ - Y]) - mu[w_] * (v[w_ + Y] - v[w_ - Y]));
        T vXZ = p4xdz * (mu[u_] * (w[u_ + XP] - w[u_ + XM]) - mu[d_] * (w[d_ + XP] - w[d_ + XM]));
        T vZX = n6zdx * (mu[e_] * (w[e_ + Z] - w[e_ - Z]) - mu[w_] * (w[w_ + Z] - w[w_ - Z]));
        kq[1] = cX + cY + cZ + vXX + vYY + vZZ + vXY + vYX + vXZ + vZX;
    }
    {
        T cX = n2dx * (rv[e_] * u[e_] - rv[w_] * u[w_]);
        T cY = n2dy * (rv[n_] * v[n_] + p[n_] - rv[s_] * v[s_] - p[s_]);
        T cZ = n2dz * (rv[u_] * w[u_] - rv[d_] * w[d_]);
        T vXX = p12xx * ((mu[e_] + mu[o]) * (v[e_] - v[o]) - (mu[w_] + mu[o]) * (v[o] - v[w_]));
        T vYY = p23yy * ((mu[n_] + mu[o]) * (v[n_] - v[o]) - (mu[s_] + mu[o]) * (v[o] - v[s_]));
        T vZZ = p12zz * ((mu[u_] + mu[o]) * (v[u_] - v[o]) - (mu[d_] + mu[o]) * (v[o] - v[d_]));
        T vYX = p4ydx * (mu[e_] * (u[e_ + Y] - u[e_ - Y]) - mu[w_] * (u[w_ + Y] - u[w_ - Y]));
        T vXY = n6xdy * (mu[n_] * (u[n_ + XP] - u[n_ + XM]) - mu[s_] * (u[s_ + XP] - u[s_ + XM]));
        T vYZ = p4ydz * (mu[u_] * (w[u_ + Y] - w[u_ - Y]) - mu[d_] * (w[d_ + Y] - w[d_ - Y]));
        T vZY = n6zdy * (mu[n_] * (w[n_ + Z] - w[n_ - Z]) - mu[s_] * (w[s_ + Z] - w[s_ - Z]));
        kq[2] = cX + cY + cZ + vXX + vYY + vZZ + vYX + vXY + vYZ + vZY;
    }
    {
        T cX = n2dx * (rw[e_] * u[e_] - rw[w_] * u[w_]);
        T cY = n2dy * (rw[n_] * v[n_] - rw[s_] * v[s_]);
        T cZ = n2dz * (rw[u_] * w[u_] + p[u_] - rw[d_] * w[d_] - p[d_]);
        T vXX = p12xx * ((mu[e_] + mu[o]) * (w[e_] - w[o]) - (mu[w_] + mu[o]) * (w[o] - w[w_]));
        T vYY = p12yy * ((mu[n_] + mu[o]) * (w[n_] - w[o]) - (mu[s_] + mu[o]) * (w[o] - w[s_]));
        T vZZ = p23zz * ((mu[u_] + mu[o]) * (w[u_] - w[o]) - (mu[d_] + mu[o]) * (w[o] - w[d_]));
        T vZX = p4zdx * (mu[e_] * (u[e_ + Z] - u[e_ - Z]) - mu[w_] * (u[w_ + Z] - u[w_ - Z]));
        T vXZ = n6xdz * (mu[u_] * (u[u_ + XP] - u[u_ + XM]) - mu[d_] * (u[d_ + XP] - u[d_ + XM]));
        T vZY = p4zdy * (mu[n_] * (v[n_ + Z] - v[n_ - Z]) - mu[s_] * (v[s_ + Z] - v[s_ - Z]));
        T vYZ = n6ydz * (mu[u_] * (v[u_ + Y] - v[u_ - Y]) - mu[d_] * (v[d_ + Y] - v[d_ - Y]));
        kq[3] = cX + cY + cZ + vXX + vYY + vZZ + vZX + vXZ + vZY + vYZ;
    }
    // energy (src/Solver.cpp:378-416)
    {
        const T mP = mu[o], mW = mu[w_], mE = mu[e_], mS = mu[s_], mN = mu[n_], mD = mu[d_], mU = mu[u_];
        const T uP = u[o], uW = u[w_], uE = u[e_], uS = u[s_], uN = u[n_], uD = u[d_], uU = u[u_];
        const T vP = v[o], vW = v[w_], vE = v[e_], vS = v[s_], vN = v[n_], vD = v[d_], vU = v[u_];
        const T wP = w[o], wW = w[w_], wE = w[e_], wS = w[s_], wN = w[n_], wD = w[d_], wU = w[u_];
        const T kP = ka[o], TP = Tt[o];
        T inv = n2dx * ((H0 + E[e_] + p[e_]) * uE - (H0 + E[w_] + p[w_]) * uW)
              + n2dy * ((H0 + E[n_] + p[n_]) * vN - (H0 + E[s_] + p[s_]) * vS)
              + n2dz * ((H0 + E[u_] + p[u_]) * wU - (H0 + E[d_] + p[d_]) * wD);
        T vx = p23xx * ((mE * uE + mP * uP) * (uE - uP) - (mW * uW + mP * uP) * (uP - uW))
             + p12xx * ((mE * vE + mP * vP) * (vE - vP) - (mW * vW + mP * vP) * (vP - vW))
             + p12xx * ((mE * wE + mP * wP) * (wE - wP) - (mW * wW + mP * wP) * (wP - wW))
             + n6ydx * (mE * uE * (v[e_ + Y] - v[e_ - Y]) - mW * uW * (v[w_ + Y] - v[w_ - Y]))
             + n6zdx * (mE * uE * (w[e_ + Z] - w[e_ - Z]) - mW * uW * (w[w_ + Z] - w[w_ - Z]))
             + p4ydx * (mE * vE * (u[e_ + Y] - u[e_ - Y]) - mW * vW * (u[w_ + Y] - u[w_ - Y]))
             + p4zdx * (mE * wE * (u[e_ + Z] - u[e_ - Z]) - mW * wW * (u[w_ + Z] - u[w_ - Z]));
        T vy = p12yy * ((mN * uN + mP * uP) * (uN - uP) - (mS * uS + mP * uP) * (uP - uS))
             + p23yy * ((mN * vN + mP * vP) * (vN - vP) - (mS * vS + mP * vP) * (vP - vS))
             + p12yy * ((mN * wN + mP * wP) * (wN - wP) - (mS * wS + mP * wP) * (wP - wS))
             + n6ydx * (mN * vN * (u[n_ + XP] - u[n_ + XM]) - mS * vS * (u[s_ + XP] - u[s_ + XM]))
             + n6zdy * (mN * vN * (w[n_ + Z] - w[n_ - Z]) - mS * vS * (w[s_ + Z] - w[s_ - Z]))
             + p4ydx * (mN * uN * (v[n_ + XP] - v[n_ + XM]) - mS * uS * (v[s_ + XP] - v[s_ + XM]))
             + p4zdy * (mN * wN * (v[n_ + Z] - v[n_ - Z]) - mS * wS * (v[s_ + Z] - v[s_ - Z]));
        T vz = p12zz * ((mU * uU + mP * uP) * (uU - uP) - (mD * uD + mP * uP) * (uP - uD))
             + p12zz * ((mU * vU + mP * vP) * (vU - vP) - (mD * vD + mP * vP) * (vP - vD))
             + p23zz * ((mU * wU + mP * wP) * (wU - wP) - (mD * wD + mP * wP) * (wP - wD))
             + n6zdx * (mU * wU * (u[u_ + XP] - u[u_ + XM]) - mD * wD * (u[d_ + XP] - u[d_ + XM]))
             + n6zdy * (mU * wU * (v[u_ + Y] - v[u_ - Y]) - mD * wD * (v[d_ + Y] - v[d_ - Y]))
             + p4zdx * (mU * uU * (w[u_ + XP] - w[u_ + XM]) - mD * uD * (w[d_ + XP] - w[d_ + XM]))
             + p4zdy * (mU * vU * (w[u_ + Y] - w[u_ - Y]) - mD * vD * (w[d_ + Y] - w[d_ - Y]));
        T heat = p12xx * ((ka[e_] + kP) * (Tt[e_] - TP) - (ka[w_] + kP) * (TP - Tt[w_]))
               + p12yy * ((ka[n_] + kP) * (Tt[n_] - TP) - (ka[s_] + kP) * (TP - Tt[s_]))
               + p12zz * ((ka[u_] + kP) * (Tt[u_] - TP) - (ka[d_] + kP) * (TP - Tt[d_]));
        kq[4] = inv + vx + vy + vz + heat;
    }
    const double dtd = *dtdev;
    if (STAGE == 1) {
        const T h2 = (T)(dtd / 2), h6 = (T)(dtd / 6);
        for (int q = 0; q < 5; ++q) {
            T u0 = (q == 0 && rhoFiltered) ? fRho[o] : U[o + q * F];
            Sout[o + q * F] = u0 + h2 * kq[q];
            A[o + q * F] = LIT ? kq[q] : (KAHAN ? h6 * kq[q] : u0 + h6 * kq[q]);
        }
    } else if (STAGE == 2 || STAGE == 3) {
        const T hs = (T)(STAGE == 2 ? dtd / 2 : dtd), h3 = (T)(dtd / 3);
        for (int q = 0; q < 5; ++q) {
            Sout[o + q * F] = U[o + q * F] + hs * kq[q];
            A[o + q * F] = LIT ? A[o + q * F] + (T)2 * kq[q] : A[o + q * F] + h3 * kq[q];
        }
    } else if (LIT) {
        const T f = (T)(dtd / 6);
        for (int q = 0; q < 5; ++q) A[o + q * F] = U[o + q * F] + f * (A[o + q * F] + kq[q]);
    } else {
        const T h6 = (T)(dtd / 6);
        for (int q = 0; q < 5; ++q) {
            T an = A[o + q * F] + h6 * kq[q];
            if constexpr (KAHAN) {
                // U^{n+1} = U^n + (D + e^{n-1}); the new rounding error is carried (device_common.h)
                T e;
                an = twoSum(U[o + q * F], an + carryDecode(comp[o + q * F], U[o + q * F]), e);
                comp[o + q * F] = carryEncode(e, an);
            }
            A[o + q * F] = an;
        }
    }
}

}  // namespace

template <class T>
void launchDerive(hipStream_t s, const Layout& L, const Params& P, const T* S, T* prim, const uint8_t* cls, const Window& W,
                  const T* rhoPre, const T* rhoEPre, int mode) {
    dim3 blk(Layout::kSegSlots), grd((L.NI + Layout::kSegNodes - 1) / Layout::kSegNodes, L.NJ, L.NK);
    if (mode == 1) hipLaunchKernelGGL((k_derive<T, 1>), grd, blk, 0, s, L, P, S, prim, cls, W, rhoPre, rhoEPre);
    else if (mode == 2) hipLaunchKernelGGL((k_derive<T, 2>), grd, blk, 0, s, L, P, S, prim, cls, W, rhoPre, rhoEPre);
    else hipLaunchKernelGGL((k_derive<T, 0>), grd, blk, 0, s, L, P, S, prim, cls, W, rhoPre, rhoEPre);
}

template <class T, bool K, bool LIT>
void launchRhsPlainK(hipStream_t s, int stage, const Layout& L, const Params& P, const T* Sin, const T* prim, const T* U,
                     T* Sout, T* A, const uint8_t* cls, const double* dtdev, const T* fRho, const int32_t* fFlag,
                     int8_t* comp) {
    dim3 blk(Layout::kSegSlots), grd((L.NI + Layout::kSegNodes - 1) / Layout::kSegNodes, L.NJ, L.NK);
    switch (stage) {
        case 1:
            hipLaunchKernelGGL((k_rhs_plain<T, 1, K, LIT>), grd, blk, 0, s, L, P, Sin, prim, U, Sout, A, cls, dtdev, fRho,
                               fFlag, comp);
            break;
        case 2:
            hipLaunchKernelGGL((k_rhs_plain<T, 2, K, LIT>), grd, blk, 0, s, L, P, Sin, prim, U, Sout, A, cls, dtdev,
                               nullptr, nullptr, comp);
            break;
        case 3:
            hipLaunchKernelGGL((k_rhs_plain<T, 3, K, LIT>), grd, blk, 0, s, L, P, Sin, prim, U, Sout, A, cls, dtdev,
                               nullptr, nullptr, comp);
            break;
        default:
            hipLaunchKernelGGL((k_rhs_plain<T, 4, K, LIT>), grd, blk, 0, s, L, P, Sin, prim, U, Sout, A, cls, dtdev,
                               nullptr, nullptr, comp);
            break;
    }
}

template <class T>
void launchRhsPlain(hipStream_t s, int stage, const Layout& L, const Params& P, const T* Sin, const T* prim, const T* U,
                    T* Sout, T* A, const uint8_t* cls, const double* dtdev, const T* fRho, const int32_t* fFlag,
                    int8_t* comp, bool literal) {
    if constexpr (sizeof(T) == 4) {
        if (literal) throw std::invalid_argument("the reference's own combine is the FP64 form");
        if (comp) {
            launchRhsPlainK<T, true, false>(s, stage, L, P, Sin, prim, U, Sout, A, cls, dtdev, fRho, fFlag, comp);
            return;
        }
        launchRhsPlainK<T, false, false>(s, stage, L, P, Sin, prim, U, Sout, A, cls, dtdev, fRho, fFlag, nullptr);
    } else {
        if (comp) throw std::invalid_argument("compensated combine needs FP32 fields");
        if (literal) launchRhsPlainK<T, false, true>(s, stage, L, P, Sin, prim, U, Sout, A, cls, dtdev, fRho, fFlag, nullptr);
        else launchRhsPlainK<T, false, false>(s, stage, L, P, Sin, prim, U, Sout, A, cls, dtdev, fRho, fFlag, nullptr);
    }
}

namespace {

__device__ __forceinline__ void filterFlagOf(const FilterSchedule& fs, StepScalars* st) {
    // Mesh::checkFilterCondition, src/Mesh.cpp:85-103, with t and timeLevel at the start of the step
    int counter = 0;
    for (int q = 0; q < fs.n_change; ++q)
        if (st->t > fs.change_times[q]) ++counter;
    int now = 0;
    if (counter < fs.n_intervals) {
        uint32_t iv = fs.intervals[counter];
        if (iv > 0 && (st->level + 1) % iv == 0) now = 1;
    }
    st->filterNow = now;
}

__global__ void k_filter_flag(FilterSchedule fs, StepScalars* st) { filterFlagOf(fs, st); }

template <class T>
__device__ __forceinline__ void rhoFilterNode(const Layout& L, const T* U, T* out, const uint8_t* cls, int i, int j, int k) {
    size_t o = L.off(i, j, k);
    if ((cls[o] & CLS_TYPE_MASK) != NS_FLUID_INTERIOR) return;
    const long XP = Layout::xstep(i, 1), XM = Layout::xstep(i, -1), Y = L.pitch, Z = (long)L.plane;
    // (6 rho_P + sum of the 6 face neighbours) / 12, in the reference's summation order
    out[o] = ((T)6 * U[o] + U[o + XP] + U[o + XM] + U[o + Y] + U[o - Y] + U[o + Z] + U[o - Z]) / (T)12;
}

template <class T>
__device__ __forceinline__ void rhoCommitNode(const Layout& L, T* U, const T* filtered, const T* A, const uint8_t* cls,
                                              int i, int j, int k) {
    size_t o = L.off(i, j, k);
    uint8_t c = cls[o];
    if ((c & CLS_TYPE_MASK) == NS_FLUID_INTERIOR) U[o] = filtered[o];
    else if (c & CLS_INACTIVE_IC) U[o] = A[o];  // the swapped-in scratch array keeps its solid values
}

template <class T>
__global__ void k_rho_filter(Layout L, const T* U, T* out, const uint8_t* cls, const int32_t* flag) {
    if (!*flag) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.NI) return;
    rhoFilterNode(L, U, out, cls, i, j, k);
}

template <class T>
__global__ void k_rho_commit(Layout L, T* U, const T* filtered, const T* A, const uint8_t* cls, const int32_t* flag) {
    if (!*flag) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= L.NI) return;
    rhoCommitNode(L, U, filtered, A, cls, i, j, k);
}

// ---- batched forms: only the cases whose filter is on (ids[m], nodes over them by the prefix sums start)
__global__ void k_filter_flag_batched(const CaseView<float>* __restrict__ cv, const int* __restrict__ ids, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) filterFlagOf(cv[ids[m]].fs, cv[ids[m]].scal);
}

__device__ __forceinline__ void batchNode(const CaseView<float>& v, int64_t r, int& i, int& j, int& k) {
    i = (int)(r % v.L.NI);
    const int64_t q = r / v.L.NI;
    j = (int)(q % v.L.NJ);
    k = (int)(q / v.L.NJ);
}

// the filtered density goes into S2's density field (free until stage 2), as in the single-case step
__global__ void k_rho_filter_batched(const CaseView<float>* __restrict__ cv, const int* __restrict__ ids,
                                     const int64_t* __restrict__ start, int n, int parity) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= start[n]) return;
    const int m = findCase(start, n, t);
    const CaseView<float>& v = cv[ids[m]];
    if (!v.scal->filterNow) return;
    int i, j, k;
    batchNode(v, t - start[m], i, j, k);
    rhoFilterNode(v.L, (const float*)roleBuffer(v, 0, parity), v.buf[3], v.cls, i, j, k);
}

__global__ void k_rho_commit_batched(const CaseView<float>* __restrict__ cv, const int* __restrict__ ids,
                                     const int64_t* __restrict__ start, int n, int parity) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= start[n]) return;
    const int m = findCase(start, n, t);
    const CaseView<float>& v = cv[ids[m]];
    if (!v.scal->filterNow) return;
    int i, j, k;
    batchNode(v, t - start[m], i, j, k);
    rhoCommitNode(v.L, roleBuffer(v, 0, parity), (const float*)v.buf[3], (const float*)roleBuffer(v, 1, parity),
                  v.cls, i, j, k);
}

}  // namespace

void launchFilterFlag(hipStream_t s, const FilterSchedule& fs, StepScalars* st) {
    hipLaunchKernelGGL(k_filter_flag, dim3(1), dim3(1), 0, s, fs, st);
}

template <class T>
void launchRhoFilter(hipStream_t s, const Layout& L, const T* U, T* out, const uint8_t* cls, const int32_t* flag) {
    dim3 blk(64), grd((L.NI + 63) / 64, L.NJ, L.NK);
    hipLaunchKernelGGL(k_rho_filter<T>, grd, blk, 0, s, L, U, out, cls, flag);
}

template <class T>
void launchRhoCommit(hipStream_t s, const Layout& L, T* U, const T* filtered, const T* A, const uint8_t* cls,
                     const int32_t* flag) {
    dim3 blk(64), grd((L.NI + 63) / 64, L.NJ, L.NK);
    hipLaunchKernelGGL(k_rho_commit<T>, grd, blk, 0, s, L, U, filtered, A, cls, flag);
}

void launchFilterFlagBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, int n) {
    if (n > 0) hipLaunchKernelGGL(k_filter_flag_batched, dim3((n + 63) / 64), dim3(64), 0, s, cv, ids, n);
}

void launchRhoFilterBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, const int64_t* start, int n,
                            int64_t total, int parity) {
    if (total > 0)
        hipLaunchKernelGGL(k_rho_filter_batched, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, cv, ids, start, n,
                           parity);
}

void launchRhoCommitBatched(hipStream_t s, const CaseView<float>* cv, const int* ids, const int64_t* start, int n,
                            int64_t total, int parity) {
    if (total > 0)
        hipLaunchKernelGGL(k_rho_commit_batched, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, cv, ids, start, n,
                           parity);
}

template void launchRhoFilter<double>(hipStream_t, const Layout&, const double*, double*, const uint8_t*, const int32_t*);
template void launchRhoFilter<float>(hipStream_t, const Layout&, const float*, float*, const uint8_t*, const int32_t*);
template void launchRhoCommit<double>(hipStream_t, const Layout&, double*, const double*, const double*, const uint8_t*,
                                      const int32_t*);
template void launchRhoCommit<float>(hipStream_t, const Layout&, float*, const float*, const float*, const uint8_t*,
                                     const int32_t*);

template void launchDerive<double>(hipStream_t, const Layout&, const Params&, const double*, double*, const uint8_t*,
                                   const Window&, const double*, const double*, int);
template void launchDerive<float>(hipStream_t, const Layout&, const Params&, const float*, float*, const uint8_t*,
                                  const Window&, const float*, const float*, int);
template void launchRhsPlain<double>(hipStream_t, int, const Layout&, const Params&, const double*, const double*,
                                     const double*, double*, double*, const uint8_t*, const double*, const double*,
                                     const int32_t*, int8_t*, bool);
template void launchRhsPlain<float>(hipStream_t, int, const Layout&, const Params&, const float*, const float*,
                                    const float*, float*, float*, const uint8_t*, const double*, const float*,
                                    const int32_t*, int8_t*, bool);

}  // namespace nsfsi
