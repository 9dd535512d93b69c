/*
 * device_common.h — device-side equation of state and small helpers, templated on the field type.
 *
 * Nondimensional perturbation variables as in the reference (src/FlowVariableGroupStructs.h:152-185):
 * rho' = rho/rho_0 - 1, velocities in units of c_0, p' and T' perturbations, rho*E' without the
 * constant 1/(gamma(gamma-1)).
 */
#pragma once

#include <hip/hip_runtime.h>

#include "ns_internal.h"

namespace nsfsi {

template <class T>
struct Eos {
    T gamma, gm1, ScPlusOne, Re, pf;  // pf = 1 / ((gamma-1) Pr)
    T invGm1;                         // 1 / (gamma-1)
    T gm1Pr;                          // (gamma-1) Pr
    __host__ __device__ static Eos make(const Params& P) {
        Eos e;
        e.gamma = (T)P.gamma;
        e.gm1 = (T)(P.gamma - 1);
        e.ScPlusOne = (T)P.ScPlusOne;
        e.Re = (T)P.Re;
        e.pf = (T)(1 / ((P.gamma - 1) * P.Pr));
        e.invGm1 = (T)(1 / (P.gamma - 1));
        e.gm1Pr = (T)((P.gamma - 1) * P.Pr);
        return e;
    }
};

// (1+T')^1.5 = x^1.5 (src/Solver.cpp:476, src/FlowVariableGroupStructs.h:183), correctly rounded in FP64:
// with s = sqrt(x) (correctly rounded), the residual x - s^2 is exact in one FMA, so sqrt(x) = s + e with
// e = (x - s^2)/(2s) to within 2^-53 |e|; x*s = p + pe exactly (FMA); x^1.5 = p + (pe + x*e) then carries
// one final rounding of an error of about 2^-104 |p| (Dekker's double-length products). FP32: x*sqrt(x).
__device__ __forceinline__ double pow15(double x) {
#pragma clang fp contract(off)
    const double s = sqrt(x);
    const double r = __builtin_fma(-s, s, x);
    const double e = r / (2.0 * s);
    const double p = x * s;
    const double pe = __builtin_fma(x, s, -p);
    return p + (pe + x * e);
}
__device__ __forceinline__ float pow15(float x) { return x * sqrtf(x); }

// Primitive variables and viscosity from conserved (updatePrimitiveVariables, src/Solver.cpp:446-479).
template <class T>
__device__ __forceinline__ void derivePrims(const Eos<T>& e, T rho, T ru, T rv, T rw, T rE, T& u, T& v, T& w, T& p,
                                            T& Tt, T& mu) {
    T m = rho + (T)1;
    u = ru / m;
    v = rv / m;
    w = rw / m;
    p = e.gm1 * (rE - m / (T)2 * (u * u + v * v + w * w));
    Tt = (e.gamma * p - rho) / ((T)1 + rho);
    mu = pow15((T)1 + Tt) * e.ScPlusOne / (e.Re * (Tt + e.ScPlusOne));
}

// Transport properties from the temperature as the boundary conditions and initial condition set them
// (deriveTransportProperties, src/FlowVariableGroupStructs.h:176-185: kappa = mu / ((gamma-1) Pr), a division,
// where the stage update of interior nodes multiplies by 1 / ((gamma-1) Pr), src/Solver.cpp:451, 478).
template <class T>
__device__ __forceinline__ void deriveTrans(const Eos<T>& e, T Tt, T& mu, T& kappa) {
    mu = pow15((T)1 + Tt) * e.ScPlusOne / (e.Re * (Tt + e.ScPlusOne));
    kappa = mu / e.gm1Pr;
}

// Conserved from primitive (deriveConservedVariables, src/FlowVariableGroupStructs.h:152-160).
template <class T>
__device__ __forceinline__ void deriveCons(const Eos<T>& e, T u, T v, T w, T p, T Tt, T& rho, T& ru, T& rv, T& rw, T& rE) {
    rho = (e.gamma * p - Tt) / ((T)1 + Tt);
    T m = (T)1 + rho;
    ru = m * u;
    rv = m * v;
    rw = m * w;
    rE = p / e.gm1 + m / (T)2 * (u * u + v * v + w * w);
}

// Outlet node values (OutletBoundary::applyBoundaryCondition, src/Boundary.cpp:304-333): density and
// momenta extrapolated linearly from the nodes one (a) and two (b) in from the face, zero pressure
// perturbation. Multiply-adds stay unfused here, so the edge-face kernel and the fused stage kernel, which
// both evaluate it, give the same bits.
template <class T>
__device__ __forceinline__ void outletValues(T gm1, const T a[4], const T b[4], T out[5]) {
#pragma clang fp contract(off)
    const T r = (T)2 * a[0] - b[0];
    const T ru = (T)2 * a[1] - b[1];
    const T rv = (T)2 * a[2] - b[2];
    const T rw = (T)2 * a[3] - b[3];
    const T p = 0;
    const T u = ru / ((T)1 + r), v = rv / ((T)1 + r), w = rw / ((T)1 + r);
    out[0] = r;
    out[1] = ru;
    out[2] = rv;
    out[3] = rw;
    out[4] = p / gm1 + ((T)1 + r) / (T)2 * (u * u + v * v + w * w);
}

// Compensated FP32 combine. The increment form keeps Uⁿ out of the stage accumulator (it holds
// D = dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4 only), so the one rounding that is relative to |U| happens
// in U + D at the end of the step. TwoSum (Knuth) gives that rounding error e exactly; it is carried to
// the next step's increment (Kahan's compensated summation over the time steps) as a signed 8-bit
// multiple of ulp(s)/254, s the rounded sum it belongs to (|e| <= ulp(s)/2, so the multiple is at most
// 127): the residual is at most ulp(s)/508, about 2^-9 of the FP32 rounding. Sums below 2^-103 in
// magnitude take the scale of 2^-103 (their error then encodes as 0). The next step decodes it with its
// own copy of s, the state it reads.
__device__ __forceinline__ int carryExp(float s) {
    const int E = (int)((__float_as_uint(s) >> 23) & 0xffu);
    return E < 24 ? 24 : E;
}
__device__ __forceinline__ int8_t carryEncode(float e, float s) {
    const float inv = __uint_as_float((uint32_t)(277 - carryExp(s)) << 23);  // 1 / ulp(s)
    float q = rintf(e * inv * 254.0f);
    q = fminf(fmaxf(q, -127.0f), 127.0f);
    return (int8_t)(int)q;
}
__device__ __forceinline__ float carryDecode(int8_t q, float s) {
    const float ulp = __uint_as_float((uint32_t)(carryExp(s) - 23) << 23);
    return (float)q * (ulp * (1.0f / 254.0f));
}
// s + e == u + d exactly (s = fl(u + d)); no operand order assumption.
__device__ __forceinline__ float twoSum(float u, float d, float& e) {
    float s = u + d;
    float bv = s - u;
    e = (u - (s - bv)) + (d - bv);
    return s;
}

// Order-preserving encoding of a non-negative double for atomicMax on the bit pattern.
__device__ __forceinline__ unsigned long long orderedBits(double x) { return (unsigned long long)__double_as_longlong(x); }

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace nsfsi
