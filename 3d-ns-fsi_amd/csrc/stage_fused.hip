/*
 * stage_fused.hip — the fused RK4 stage: one kernel computes primitive variables, the 19-point
 * convective + viscous + heat right-hand side (compute_RK4_step_*, src/Solver.cpp:181-418) and the
 * stage combine (src/Solver.cpp:422-507) in one pass over the grid.
 *
 * Flux form. The reference's per-node expression is regrouped into face fluxes: for every direction d,
 *   RHS(P) = sum_d [ Phi_d(P + h_d/2) - Phi_d(P - h_d/2) ],
 *   Phi_d(a, b) = (G_d(a) + G_d(b)) / (2 h_d) + C_d(a, b),
 * where G_d is the nodal central flux (inviscid flux plus the cross-derivative viscous stresses, e.g.
 * G_x of x-momentum = -(rho u u + p) - 2/3 mu (dv/dy + dw/dz)) and C_d the compact second-difference
 * terms, e.g. 2/(3 dx^2) (mu_a + mu_b)(u_b - u_a). Expanding the differences gives back every term of
 * the reference expression (same stencil, same coefficients); only the rounding differs. Each face flux
 * is computed once and shared by the two nodes of the face.
 *
 * Tiling. A block is 64 x (TY + 2) threads: one wavefront per x-row. The x-tiles are the layout's
 * x-segments (62 nodes in 64 slots): lanes 1..62 own the segment's nodes, lanes 0 and 63 are the x-halo
 * (they read the neighbouring segments' edge nodes); rows 1..TY own output rows, rows 0 and TY+1 the
 * y-halo. A row's last segment, when it holds at most 30 nodes, is tiled at half width: each half of the
 * wave takes it in one of two y-bands, so a 512-node row costs 8.5 waves instead of 9 and a 64-node row
 * 1.5 instead of 2. The block streams through a chunk of z-planes: the own column's neighbours in z live in
 * registers, x-neighbours come by cross-lane shuffle, y-neighbours through LDS (two buffers in FP32, so a
 * plane needs one barrier). Primitive variables are
 * recomputed from the conserved state; at immersed-filter nodes from the pre-filter density and energy
 * (the reference keeps primitives pre-filter there). Each output wave stores its whole segment, padding
 * slots included, with streaming stores, so every line it writes is written whole. Blocks are mapped so
 * that each XCD sweeps a contiguous slab (its L2 then holds the halo of its neighbours).
 *
 * Multiply-adds are fused only inside one source expression (contract(on), not the compiler's
 * cross-statement fusion), so the arithmetic of a node does not depend on which block, tile shape or
 * z-chunk computes it: results are bit-identical for every tile shape.
 */
#pragma clang fp contract(on)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>

#include "device_common.h"
#include "stage.h"

namespace nsfsi {

namespace {

template <class T>
struct Coef {
    T i2dx, i2dy, i2dz;      // 1 / (2 h)
    T c43x, c43y, c43z;      // 2 / (3 h^2)
    T c1x, c1y, c1z;         // 1 / (2 h^2)
    T gamma, gm1, H0, ScP1, Re, pf;
};

template <class T>
Coef<T> makeCoef(const Params& P) {
    Coef<T> c;
    const double dx = P.hx, dy = P.hy, dz = P.hz;
    c.i2dx = (T)(1 / (2 * dx));
    c.i2dy = (T)(1 / (2 * dy));
    c.i2dz = (T)(1 / (2 * dz));
    c.c43x = (T)(2 / (3 * dx * dx));
    c.c43y = (T)(2 / (3 * dy * dy));
    c.c43z = (T)(2 / (3 * dz * dz));
    c.c1x = (T)(1 / (2 * dx * dx));
    c.c1y = (T)(1 / (2 * dy * dy));
    c.c1z = (T)(1 / (2 * dz * dz));
    c.gamma = (T)P.gamma;
    c.gm1 = (T)(P.gamma - 1);
    c.H0 = (T)(1 / (P.gamma - 1));
    c.ScP1 = (T)P.ScPlusOne;
    c.Re = (T)P.Re;
    c.pf = (T)(1 / ((P.gamma - 1) * P.Pr));
    return c;
}

template <class T>
struct FusedArgs {
    Layout L;
    Window W;
    Coef<T> c;
    // one base pointer per conserved field: loads and stores then share one 32-bit byte offset per
    // node (scalar base + vector offset addressing, no per-access 64-bit address arithmetic)
    const T* Sin[5];
    const T* U[5];
    T* Sout[5];
    T* A[5];
    const uint8_t* cls;
    const T* rhoPre;
    const T* rhoEPre;
    const double* dtdev;
    int nbx, nby, nbz, kc;   // tiles in x, y, z and planes per z-chunk
    int tailHalf;            // the last x-segment holds at most 30 nodes: its tiles are half-width
    int nblocks;
    // outlet face at x = NI-1 applied here (its owned rows j0..j1 x k0..k1), instead of by the edge-face pass
    int fuseOutlet, oj0, oj1, ok0, ok1;
    // stage 4 only: per-block partials of the end-of-step scans over interior, non-filter nodes
    // (slots: 3 spectral-radius maxima, viscous maximum, 5 sums of squared change, non-finite count)
    double* part;
    int norms;
    T vm;  // max(4/3, gamma/Pr)
    // stage 1 of a step with the periodic density filter: combine base density and flag
    const T* fRho;
    const int32_t* fFlag;
    // compensated FP32 combine: per-field 8-bit rounding error of U carried between steps (the stage
    // accumulator then holds the increment only)
    int8_t* comp[5];
};

// Whole-wavefront shifts by one lane (DPP wave_shl:1 / wave_shr:1 of GFX9): shl1 gives each lane the
// value of lane + 1, shr1 the value of lane - 1 (the edge lane receives 0; edge lanes are halo lanes).
// bound_ctrl set: the lane without a source reads 0 in hardware (no move of an old value first).
__device__ __forceinline__ float shl1(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float shr1(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double shl1(double x) {
    int2 v = *reinterpret_cast<int2*>(&x);
    v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x130, 0xf, 0xf, true);
    v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x130, 0xf, 0xf, true);
    return *reinterpret_cast<double*>(&v);
}
__device__ __forceinline__ double shr1(double x) {
    int2 v = *reinterpret_cast<int2*>(&x);
    v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x138, 0xf, 0xf, true);
    v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x138, 0xf, 0xf, true);
    return *reinterpret_cast<double*>(&v);
}

template <class T>
__device__ __forceinline__ T ldAt(const T* base, uint32_t boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <class T>
__device__ __forceinline__ void stAt(T* base, uint32_t boff, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}
template <class T>
__device__ __forceinline__ void stNT(T* base, uint32_t boff, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff));
}

// Field accessors of one stage launch. BUF (FP32 states of at most 4 GiB): buffer instructions, one
// resource per 5-field state buffer, the node's byte offset in a VGPR and the field's in an SGPR, so no
// address arithmetic per access. Otherwise (FP64, or an FP32 state beyond 4 GiB): global accesses at
// per-field base pointers moved to the tile's first plane (kb0, plane kb - 1 or 0), with the node's 32-bit
// byte offset from there: a tile spans at most its z-chunk plus two planes, so fields of any size work.
template <class T, bool BUF>
struct Mem;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

// cache-policy operand of the buffer intrinsics: non-temporal (streaming)
constexpr int kNT = 2;

template <>
struct Mem<float, true> {
    __amdgpu_buffer_rsrc_t sin, u, a, sout, comp;
    uint32_t fsb;  // field stride in bytes (FP32)
    static constexpr int kb0 = 0;  // offsets from the fields' start
    __device__ Mem(const FusedArgs<float>& x, int) {
        fsb = (uint32_t)(x.L.fstride * sizeof(float));
        const uint32_t nb = 5 * fsb;
        sin = rsrc(x.Sin[0], nb);
        u = rsrc(x.U[0], nb);
        a = rsrc(x.A[0], nb);
        sout = rsrc(x.Sout[0], nb);
        comp = rsrc(x.comp[0], nb / 4);
    }
    __device__ float ld(__amdgpu_buffer_rsrc_t r, int q, uint32_t bo) const {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)bo, (int)(q * fsb), 0));
    }
    // stores are streaming (non-temporal): every output wave writes whole lines that this launch never
    // reads again
    __device__ void st(__amdgpu_buffer_rsrc_t r, int q, uint32_t bo, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)bo, (int)(q * fsb), kNT);
    }
    __device__ float ldSin(int q, uint32_t bo) const { return ld(sin, q, bo); }
    __device__ float ldU(int q, uint32_t bo) const { return ld(u, q, bo); }
    __device__ float ldA(int q, uint32_t bo) const { return ld(a, q, bo); }
    __device__ void stA(int q, uint32_t bo, float v) const { st(a, q, bo, v); }
    __device__ void stSout(int q, uint32_t bo, float v) const { st(sout, q, bo, v); }
    // compensation: 8-bit fields at a quarter of the byte offsets
    __device__ int8_t ldComp(int q, uint32_t bo) const {
        return (int8_t)__builtin_amdgcn_raw_buffer_load_b8(comp, (int)(bo >> 2), (int)(q * (fsb >> 2)), 0);
    }
    __device__ void stComp(int q, uint32_t bo, int8_t v) const {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, comp, (int)(bo >> 2), (int)(q * (fsb >> 2)), kNT);
    }
};

template <class T>
struct Mem<T, false> {
    const T* sin[5];
    const T* u[5];
    T* a[5];
    T* sout[5];
    int8_t* comp[5];
    int kb0;  // plane the offsets start from
    __device__ Mem(const FusedArgs<T>& x, int k0) : kb0(k0) {
        const size_t e = (size_t)k0 * x.L.plane;  // elements (and carry bytes) before plane k0
        for (int q = 0; q < 5; ++q) {
            sin[q] = x.Sin[q] + e;
            u[q] = x.U[q] + e;
            a[q] = x.A[q] + e;
            sout[q] = x.Sout[q] ? x.Sout[q] + e : nullptr;
            comp[q] = x.comp[q] ? x.comp[q] + e : nullptr;
        }
    }
    __device__ T ldSin(int q, uint32_t bo) const { return ldAt(sin[q], bo); }
    __device__ T ldU(int q, uint32_t bo) const { return ldAt(u[q], bo); }
    __device__ T ldA(int q, uint32_t bo) const { return ldAt((const T*)a[q], bo); }
    __device__ void stA(int q, uint32_t bo, T v) const { stNT(a[q], bo, v); }
    __device__ void stSout(int q, uint32_t bo, T v) const { stNT(sout[q], bo, v); }
    __device__ int8_t ldComp(int q, uint32_t bo) const {
        if constexpr (sizeof(T) == 4) return ldAt((const int8_t*)comp[q], bo >> 2);
        else return 0;
    }
    __device__ void stComp(int q, uint32_t bo, int8_t v) const {
        if constexpr (sizeof(T) == 4) stNT(comp[q], bo >> 2, v);
    }
};

// Whether an FP32 state buffer (five fields) fits one buffer resource.
inline bool stateFitsBuffer(const Layout& L) { return 5 * L.fstride * sizeof(float) <= 0xFFFFFFFFull; }

// Point state at one plane of the own column.
template <class T>
struct Pt {
    T ru, rv, rw, E;      // conserved (current buffer values)
    T rho;                // current density (for the combine / copies)
    T u, v, w, p, Tt, mu; // primitives (pre-filter at filter nodes)
    T dxu, dxv, dxw, dyu, dyv, dyw;  // in-plane central derivatives
};

// Reciprocal and square root: IEEE in FP64 (the parity form); in FP32 the hardware's 1-ulp
// v_rcp_f32 / v_sqrt_f32, which replace multi-instruction correctly-rounded sequences.
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double sqrtT(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_amdgcn_sqrtf(x); }

template <class T>
__device__ __forceinline__ void prims(const Coef<T>& c, T rho, T ru, T rv, T rw, T E, Pt<T>& q) {
    T m = (T)1 + rho;
    T im = rcp(m);
    q.u = ru * im;
    q.v = rv * im;
    q.w = rw * im;
    q.p = c.gm1 * (E - m * (T)0.5 * (q.u * q.u + q.v * q.v + q.w * q.w));
    q.Tt = (c.gamma * q.p - rho) * im;
    T a = (T)1 + q.Tt;
    q.mu = a * sqrtT(a) * c.ScP1 * rcp(c.Re * (q.Tt + c.ScP1));
}

// Compact face terms between a and b along one direction; k43 is the coefficient of the velocity
// component normal to the face (4/3 mu), k1 of the tangential ones and of the heat flux.
template <class T, int DIR>
__device__ __forceinline__ void compact(const Coef<T>& c, const Pt<T>& a, const Pt<T>& b, T out[5]) {
    const T k43 = DIR == 0 ? c.c43x : (DIR == 1 ? c.c43y : c.c43z);
    const T k1 = DIR == 0 ? c.c1x : (DIR == 1 ? c.c1y : c.c1z);
    T mm = a.mu + b.mu;
    T du = b.u - a.u, dv = b.v - a.v, dw = b.w - a.w, dT = b.Tt - a.Tt;
    T ku = DIR == 0 ? k43 : k1, kv = DIR == 1 ? k43 : k1, kw = DIR == 2 ? k43 : k1;
    out[0] = 0;
    out[1] = ku * mm * du;
    out[2] = kv * mm * dv;
    out[3] = kw * mm * dw;
    out[4] = ku * (b.mu * b.u + a.mu * a.u) * du + kv * (b.mu * b.v + a.mu * a.v) * dv +
             kw * (b.mu * b.w + a.mu * a.w) * dw + k1 * c.pf * mm * dT;
}

// Nodal central fluxes. dzu/dzv/dzw are the z-derivatives at the point (from the own column).
template <class T>
__device__ __forceinline__ void gx(const Coef<T>& c, const Pt<T>& q, T dzu, T dzw, T g[5]) {
    const T h = c.H0 + q.E + q.p;
    T t0 = -(T)(2.0 / 3.0) * q.mu * (q.dyv + dzw), t1 = q.mu * q.dyu, t2 = q.mu * dzu;
    g[0] = -q.ru;
    g[1] = -(q.ru * q.u + q.p) + t0;
    g[2] = -q.rv * q.u + t1;
    g[3] = -q.rw * q.u + t2;
    g[4] = -h * q.u + q.u * t0 + q.v * t1 + q.w * t2;
}
template <class T>
__device__ __forceinline__ void gy(const Coef<T>& c, const Pt<T>& q, T dzv, T dzw, T g[5]) {
    const T h = c.H0 + q.E + q.p;
    T t0 = q.mu * q.dxv, t1 = -(T)(2.0 / 3.0) * q.mu * (q.dxu + dzw), t2 = q.mu * dzv;
    g[0] = -q.rv;
    g[1] = -q.ru * q.v + t0;
    g[2] = -(q.rv * q.v + q.p) + t1;
    g[3] = -q.rw * q.v + t2;
    g[4] = -h * q.v + q.u * t0 + q.v * t1 + q.w * t2;
}
template <class T>
__device__ __forceinline__ void gz(const Coef<T>& c, const Pt<T>& q, T g[5]) {
    const T h = c.H0 + q.E + q.p;
    T t0 = q.mu * q.dxw, t1 = q.mu * q.dyw, t2 = -(T)(2.0 / 3.0) * q.mu * (q.dxu + q.dyv);
    g[0] = -q.rw;
    g[1] = -q.ru * q.w + t0;
    g[2] = -q.rv * q.w + t1;
    g[3] = -(q.rw * q.w + q.p) + t2;
    g[4] = -h * q.w + q.u * t0 + q.v * t1 + q.w * t2;
}

// LDS layout: [component][row][lane]
template <class T, int TY>
struct Smem {
    T uvw[3][TY + 2][64];
    T face[10][TY + 2][64];
};

// Raw conserved values of one point, fetched ahead of use (software pipelining). The pre-filter density
// and energy of a filter node are kept beside the raw values and selected at use, so no register copy of
// a load in flight is needed at fetch time (that copy would make the wave wait for the load there).
template <class T>
struct Raw {
    T rho, ru, rv, rw, E;  // current buffer values
    T rp, Ep;              // pre-filter density and energy (filter nodes only)
    bool filt;
};

// Byte offset of slot s of row j in a plane: 32-bit (a field stays below 4 GiB), the column part per
// thread and the plane part uniform.
template <class T>
__device__ __forceinline__ uint32_t slotBytes(const Layout& L, int s, int j) {
    return (uint32_t)(j * L.pitch + s) * (uint32_t)sizeof(T);
}

// Planes past `last` (the highest plane the tile's stencil reads: one past its last output plane) are not
// loaded: the software pipeline runs ahead of the loop's end and those values are never used.
template <class T, class M>
__device__ __forceinline__ Raw<T> fetch(const FusedArgs<T>& a, const M& m, int ic, int jc, uint32_t col, int k,
                                        int last) {
    const Layout& L = a.L;
    Raw<T> r;
    if (k > last) {  // uniform over the block
        r.rho = r.ru = r.rv = r.rw = r.E = r.rp = r.Ep = 0;
        r.filt = false;
        return r;
    }
    const int kc = min(max(k, 0), L.NK - 1);
    const uint32_t bo = col + (uint32_t)(kc - m.kb0) * (uint32_t)(L.plane * sizeof(T));
    r.rho = m.ldSin(0, bo);
    r.ru = m.ldSin(1, bo);
    r.rv = m.ldSin(2, bo);
    r.rw = m.ldSin(3, bo);
    r.E = m.ldSin(4, bo);
    r.rp = r.Ep = 0;
    r.filt = false;
    if (a.W.contains(ic, jc, kc) && (a.cls[L.off(ic, jc, kc)] & CLS_FILTER)) {
        size_t w = a.W.idx(ic, jc, kc);
        r.filt = true;
        r.rp = a.rhoPre[w];
        r.Ep = a.rhoEPre[w];
    }
    return r;
}

// Primitives and x-derivatives (by shuffle) of a fetched plane; the velocities are published in LDS for
// the y-derivatives, which planeDy reads after a barrier.
template <class T, int TY>
__device__ __forceinline__ void planePrims(const FusedArgs<T>& a, Smem<T, TY>& sm, int row, int lane, const Raw<T>& r,
                                           Pt<T>& q) {
    q.rho = r.rho;
    q.ru = r.ru;
    q.rv = r.rv;
    q.rw = r.rw;
    q.E = r.E;
    prims(a.c, r.filt ? r.rp : r.rho, r.ru, r.rv, r.rw, r.filt ? r.Ep : r.E, q);
    T uE = shl1(q.u), uW = shr1(q.u), vE = shl1(q.v), vW = shr1(q.v), wE = shl1(q.w), wW = shr1(q.w);
    q.dxu = (uE - uW) * a.c.i2dx;
    q.dxv = (vE - vW) * a.c.i2dx;
    q.dxw = (wE - wW) * a.c.i2dx;
    sm.uvw[0][row][lane] = q.u;
    sm.uvw[1][row][lane] = q.v;
    sm.uvw[2][row][lane] = q.w;
}

// y-derivatives from the published velocities. Halo rows never use them (they only supply G_y, which
// needs x- and z-derivatives).
template <class T, int TY>
__device__ __forceinline__ void planeDy(const FusedArgs<T>& a, const Smem<T, TY>& sm, int row, int lane, Pt<T>& q,
                                        bool needDy) {
    if (needDy) {
        q.dyu = (sm.uvw[0][row + 1][lane] - sm.uvw[0][row - 1][lane]) * a.c.i2dy;
        q.dyv = (sm.uvw[1][row + 1][lane] - sm.uvw[1][row - 1][lane]) * a.c.i2dy;
        q.dyw = (sm.uvw[2][row + 1][lane] - sm.uvw[2][row - 1][lane]) * a.c.i2dy;
    } else {
        q.dyu = q.dyv = q.dyw = 0;
    }
}

template <class T, int TY>
__device__ __forceinline__ void finishPlane(const FusedArgs<T>& a, Smem<T, TY>& sm, int row, int lane, const Raw<T>& r,
                                            Pt<T>& q, bool needDy) {
    planePrims<T, TY>(a, sm, row, lane, r, q);
    __syncthreads();
    planeDy<T, TY>(a, sm, row, lane, q, needDy);
}

// LDS buffers of the z-loop: two in FP32 (planes alternate, one barrier per plane; two blocks of 6 output
// rows still fit a CU), one in FP64 (a second barrier per plane guards its reuse).
template <class T>
constexpr int smemBuffers() { return sizeof(T) == 4 ? 2 : 1; }

template <class T, int TY>
__device__ __forceinline__ void readFaceNb(const Smem<T, TY>& sm, int r, int lane, Pt<T>& nb, T g[5]) {
    for (int q = 0; q < 5; ++q) g[q] = sm.face[q][r][lane];
    nb.mu = sm.face[5][r][lane];
    nb.u = sm.face[6][r][lane];
    nb.v = sm.face[7][r][lane];
    nb.w = sm.face[8][r][lane];
    nb.Tt = sm.face[9][r][lane];
}

// Position of a tile: x-segment, y-band (the first of two for a half-width tile), z-chunk.
struct TileIdx {
    int bx, by, bz;
    bool half;
};

// Tiles in x-segment-fastest order. With half-width tail tiles the y-bands go in pairs: the pair's full
// tiles (band 2m, then band 2m+1, segments 0..nbx-2 each), then one tile of the last segment holding both
// bands (the last pair of an odd band count has one band).
__device__ __forceinline__ TileIdx tileOf(int lb, int nbx, int nby, bool tailHalf) {
    TileIdx t;
    if (!tailHalf) {
        t.bx = lb % nbx;
        t.by = (lb / nbx) % nby;
        t.bz = lb / (nbx * nby);
        t.half = false;
        return t;
    }
    const int nf = nbx - 1, perPair = 2 * nf + 1, perChunk = nf * nby + (nby + 1) / 2;
    t.bz = lb / perChunk;
    const int r0 = lb % perChunk, m = r0 / perPair, r = r0 % perPair;
    const int nfull = (nby & 1) && m == nby / 2 ? nf : 2 * nf;
    if (r < nfull) {
        t.bx = r % nf;
        t.by = 2 * m + r / nf;
        t.half = false;
    } else {
        t.bx = nf;
        t.by = 2 * m;
        t.half = true;
    }
    return t;
}

// One tile (block lb of the case described by a): the whole stage for one x-segment x TY output rows over one
// z-chunk. When a row's last segment holds at most 30 nodes, its tiles are half-width: lanes 0..31 take the
// segment in y-band by, lanes 32..63 in band by+1, each half with its own halo lanes (sub-lane 0 and the
// sub-lanes past the last node), so a wave shift across the halves only moves values between halo lanes.
template <class T, int STAGE, int TY, bool SCAN, bool KAHAN, bool BUF>
__device__ __forceinline__ void stageTile(const FusedArgs<T>& a, const int lb, Smem<T, TY> (&smb)[smemBuffers<T>()],
                                          double (&wred)[TY + 2][10]) {
    constexpr int NB = smemBuffers<T>();
    const TileIdx tl = tileOf(lb, a.nbx, a.nby, a.tailHalf);
    const int bx = tl.bx, bz = tl.bz;
    const int lane = threadIdx.x, row = threadIdx.y;
    const int sl = tl.half ? lane & 31 : lane;            // lane within the (half-)tile
    const int by = tl.half ? tl.by + (lane >> 5) : tl.by;  // y-band of this lane
    const Layout& L = a.L;
    const int i = bx * 62 + sl - 1;
    const int j = by * TY + row - 1;
    const int kb = bz * a.kc, ke = min(L.NK, kb + a.kc);
    const bool outRow = row >= 1 && row <= TY && j < L.NJ;
    const bool outLane = sl >= 1 && sl <= 62 && i < L.NI;
    const Coef<T>& c = a.c;
    const double dtd = *a.dtdev;
    const bool rhoFiltered = STAGE == 1 && a.fFlag && *a.fFlag;
    static_assert(!KAHAN || sizeof(T) == 4, "the compensated combine is the FP32 form");
    static_assert(!BUF || sizeof(T) == 4, "buffer-resource access is the FP32 form");
    // FP32 stage 2 does not read U^n: stage 1 left S1 = U^n + (dt/2) k1 and the accumulator A = (dt/6) k1
    // (compensated combine) or U^n + (dt/6) k1 (plain), so U^n = S1 - 3 A, or (3 A - S1) / 2, one rounding
    // from the two values this stage loads anyway. Filter nodes load it: the boundary pass after stage 1
    // replaced their S1 density and energy. FP64 keeps the load (its parity form).
    constexpr bool rebuildU = sizeof(T) == 4 && STAGE == 2;
    const Mem<T, BUF> m(a, kb > 0 ? kb - 1 : 0);
    // clamped column of the stencil point (halo lanes and rows beyond the grid repeat the edge); the x-tile
    // is segment bx of the layout, so the halo lanes read the neighbouring segments
    const int ic = min(max(i, 0), L.NI - 1), jc = min(max(j, 0), L.NJ - 1);
    const uint32_t colc = slotBytes<T>(L, Layout::col(ic), jc);
    // the own slot: sub-lane l writes slot l of segment bx (sub-lanes 0 and 63 the padding slots)
    const uint32_t colo = slotBytes<T>(L, Layout::kSegSlots * bx + sl, j);
    const uint32_t planeB = (uint32_t)(L.plane * sizeof(T));
    // Slots without a node that this wave writes (values never read): the segment's padding slots and, in
    // FP32, in the row's last segment, the slots past the last node up to the end of its 128-byte line
    // (for the 8-bit carry, whose 64 aligned bytes span a whole segment, up to the segment's end; a
    // half-width tile writes the carry's upper half with a second store), so that no line (carry: no
    // aligned 64-byte half) of the row is written in part. (FP64 ran 0.5 % slower with its last line completed on 512-node rows, whose last line holds
    // one node and fifteen slots past it.)
    const int slot = Layout::kSegSlots * bx + sl, endSlot = Layout::col(L.NI - 1) + 1;
    const bool padLane = sl == 0 || sl == 63 || i >= L.NI;
    const bool padLine = padLane && slot < (sizeof(T) == 4 ? (endSlot + 31) / 32 * 32 : endSlot);
    const int segEnd = min(L.pitch, (endSlot + 63) / 64 * 64);
    const bool padSeg = padLane && slot < segEnd;
    const bool upperSeg = tl.half && slot + 32 < segEnd;  // carry slot + 32 of a half-width tile

    // prologue: planes kb-1 and kb, and the fetch of kb+1 in flight
    Pt<T> qm, q0, q1;
    const int last = ke;  // highest plane the stencil of output plane ke-1 reads
    Raw<T> rnext = fetch(a, m, ic, jc, colc, kb - 1, last);
    const bool inner = row >= 1 && row <= TY;  // output rows (halo rows are 0 and TY+1)
    finishPlane<T, TY>(a, smb[0], row, lane, rnext, qm, inner);
    rnext = fetch(a, m, ic, jc, colc, kb, last);
    if (NB == 1) __syncthreads();  // the readers of plane kb-1 are done before its buffer is rewritten
    finishPlane<T, TY>(a, smb[NB - 1], row, lane, rnext, q0, inner);
    rnext = fetch(a, m, ic, jc, colc, kb + 1, last);
    // The stencil input is fetched two planes ahead of its use, except in the FP32 stage 4 with global
    // addressing, whose 64-bit addresses, scans and compensation leave no registers for it (it spills).
    constexpr bool kDeep = STAGE != 4 || sizeof(T) == 8 || BUF;
    Raw<T> rnext2{};
    if constexpr (kDeep) rnext2 = fetch(a, m, ic, jc, colc, kb + 2, last);
    T gz0[5], phizm[5];
    if (inner) {
        T gzm[5], cz[5];
        gz(c, qm, gzm);
        gz(c, q0, gz0);
        compact<T, 2>(c, qm, q0, cz);
        for (int q = 0; q < 5; ++q) phizm[q] = (gzm[q] + gz0[q]) * c.i2dz + cz[q];
    } else {
        for (int q = 0; q < 5; ++q) gz0[q] = phizm[q] = 0;
    }
    T umz = qm.u, vmz = qm.v, wmz = qm.w;
    // per-thread scan accumulators in the field type (a thread sees at most one z-chunk of nodes); the
    // block partials are combined in FP64
    T red[10];
    for (int q = 0; q < 10; ++q) red[q] = 0;
    // One buffer: the prologue's readers of uvw are done before the loop rewrites it. Two buffers: plane k
    // uses buffer (k - kb) & 1, and whoever writes it again (plane k + 2) has passed the barrier of plane
    // k + 1, which every wave reaches after its reads of plane k; the prologue's planes use buffer 0
    // (plane kb-1, read before the barrier of plane kb) and buffer 1 (plane kb, read before the barrier of
    // the loop's first plane).
    if (NB == 1) __syncthreads();

    for (int k = kb; k < ke; ++k) {
        Smem<T, TY>& sm = smb[NB == 2 ? (k - kb) & 1 : 0];
        const Raw<T> rcur = rnext;
        if constexpr (kDeep) {
            rnext = rnext2;
            rnext2 = fetch(a, m, ic, jc, colc, k + 3, last);
        } else {
            rnext = fetch(a, m, ic, jc, colc, k + 2, last);
        }
        // combine inputs of the output node, fetched before the stencil work
        const bool doOut = outRow && outLane;
        const uint32_t bo = outRow ? colo + (uint32_t)(k - m.kb0) * planeB : 0;
        const size_t o = bo / sizeof(T) + (size_t)m.kb0 * L.plane;  // the node's layout offset
        bool interior = false, filterNode = false;
        T base[5], acc[5];
        (void)filterNode;
        if (doOut) {
            interior = i >= 1 && i <= L.NI - 2 && j >= 1 && j <= L.NJ - 2 && k >= 1 && k <= L.NK - 2;
            if (interior && a.W.contains(i, j, k)) {
                const uint8_t cl = a.cls[o];
                interior = (cl & CLS_TYPE_MASK) == NS_FLUID_INTERIOR;
                filterNode = (cl & CLS_FILTER) != 0;
            }
            if (STAGE == 2 || STAGE == 3 || (STAGE == 4 && KAHAN) || (SCAN && a.norms)) {
                if (interior && (!rebuildU || filterNode))
                    for (int q = 0; q < 5; ++q) base[q] = m.ldU(q, bo);
            }
            if (STAGE >= 2) {
                if (interior)
                    for (int q = 0; q < 5; ++q) acc[q] = m.ldA(q, bo);
            } else if (!interior) {
                for (int q = 0; q < 5; ++q) acc[q] = m.ldA(q, bo);
            }
        }
        // plane k+1: primitives and x-derivatives, velocities published for its y-derivatives
        planePrims<T, TY>(a, sm, row, lane, rcur, q1);
        const T dzu = (q1.u - umz) * c.i2dz, dzv = (q1.v - vmz) * c.i2dz, dzw = (q1.w - wmz) * c.i2dz;
        // y-faces of plane k: every row publishes G_y and its compact inputs; output rows evaluate both
        // faces (G_y needs no y-derivative, so it goes out with plane k+1's velocities under one barrier)
        T g[5];
        gy(c, q0, dzv, dzw, g);
        for (int q = 0; q < 5; ++q) sm.face[q][row][lane] = g[q];
        sm.face[5][row][lane] = q0.mu;
        sm.face[6][row][lane] = q0.u;
        sm.face[7][row][lane] = q0.v;
        sm.face[8][row][lane] = q0.w;
        sm.face[9][row][lane] = q0.Tt;
        __syncthreads();
        planeDy<T, TY>(a, sm, row, lane, q1, inner);
        // z-face above plane k (output rows only)
        T gz1[5], phiz[5];
        if (inner) {
            gz(c, q1, gz1);
            T cz[5];
            compact<T, 2>(c, q0, q1, cz);
            for (int q = 0; q < 5; ++q) phiz[q] = (gz0[q] + gz1[q]) * c.i2dz + cz[q];
        } else {
            for (int q = 0; q < 5; ++q) gz1[q] = phiz[q] = 0;
        }
        if (outRow) {
            T rhs[5];
            {
                Pt<T> nb;
                T gn[5], cy[5];
                readFaceNb<T, TY>(sm, row + 1, lane, nb, gn);
                compact<T, 1>(c, q0, nb, cy);
                for (int q = 0; q < 5; ++q) rhs[q] = (g[q] + gn[q]) * c.i2dy + cy[q];
                readFaceNb<T, TY>(sm, row - 1, lane, nb, gn);
                compact<T, 1>(c, nb, q0, cy);
                for (int q = 0; q < 5; ++q) rhs[q] -= (gn[q] + g[q]) * c.i2dy + cy[q];
            }
            // x part: faces between lanes by shuffle
            T gxv[5];
            gx(c, q0, dzu, dzw, gxv);
            Pt<T> e;
            e.mu = shl1(q0.mu);
            e.u = shl1(q0.u);
            e.v = shl1(q0.v);
            e.w = shl1(q0.w);
            e.Tt = shl1(q0.Tt);
            T cx[5];
            compact<T, 0>(c, q0, e, cx);
            for (int q = 0; q < 5; ++q) {
                T phix = (gxv[q] + shl1(gxv[q])) * c.i2dx + cx[q];
                rhs[q] += phix - shr1(phix);
                rhs[q] += phiz[q] - phizm[q];
            }
            // Values written at this slot. Each field goes out in one store for the whole wave: output
            // lanes with their results, the padding lanes with the node's current values (never read),
            // so a wave whose nodes all write covers its segment's lines in full.
            T cur[5] = {q0.rho, q0.ru, q0.rv, q0.rw, q0.E};
            T vS[5], vA[5];
            int8_t vC[5];
            bool wS = padLine, wA = padLine, wC = padSeg;
            for (int q = 0; q < 5; ++q) {
                vS[q] = vA[q] = cur[q];
                vC[q] = 0;
            }
            if (doOut) {
                if (interior) {
                    wS = wA = wC = true;
                    if (STAGE == 1) {
                        const T h2 = (T)(dtd / 2), h6 = (T)(dtd / 6);
                        // the reference filters the density of U before the step but keeps the
                        // primitives, so only the combine base changes (src/Solver.cpp:134-141)
                        const T c0 = cur[0];
                        cur[0] = rhoFiltered ? ldAt(a.fRho + (size_t)m.kb0 * L.plane, bo) : c0;
                        for (int q = 0; q < 5; ++q) {
                            vS[q] = cur[q] + h2 * rhs[q];
                            vA[q] = KAHAN ? h6 * rhs[q] : cur[q] + h6 * rhs[q];
                        }
                    } else if (STAGE == 2 || STAGE == 3) {
                        const T hs = (T)(STAGE == 2 ? dtd / 2 : dtd), h3 = (T)(dtd / 3);
                        if (rebuildU && !filterNode)
                            for (int q = 0; q < 5; ++q)
                                base[q] = KAHAN ? fmaf(-3.0f, (float)acc[q], (float)cur[q])
                                                : fmaf(1.5f, (float)acc[q], -0.5f * (float)cur[q]);
                        for (int q = 0; q < 5; ++q) {
                            vS[q] = base[q] + hs * rhs[q];
                            vA[q] = acc[q] + h3 * rhs[q];
                        }
                    } else {
                        const T h6 = (T)(dtd / 6);
                        T an[5];
                        for (int q = 0; q < 5; ++q) {
                            an[q] = acc[q] + h6 * rhs[q];
                            if constexpr (KAHAN) {
                                // U^{n+1} = U^n + (D + e^{n-1}), its rounding error kept for the next step
                                T e;
                                an[q] = twoSum(base[q], an[q] + carryDecode(m.ldComp(q, bo), base[q]), e);
                                vC[q] = carryEncode(e, an[q]);
                            }
                            vA[q] = an[q];
                        }
                        // end-of-step scans; filter nodes change once more in the boundary pass and are
                        // scanned there (src/Solver.cpp:91-126, src/Mesh.cpp:171-178, src/Mesh.h:48-61)
                        if (SCAN && !filterNode) {
                            if (a.norms)
                                for (int q = 0; q < 5; ++q) {
                                    T d = base[q] - an[q];
                                    red[4 + q] += d * d;
                                }
                            Pt<T> pn;
                            prims(c, an[0], an[1], an[2], an[3], an[4], pn);
                            T cs, nu;
                            if constexpr (sizeof(T) == 4) {
                                // FP32 scans only bound the next step size: hardware reciprocal and
                                // square root (1 ulp) instead of the correctly rounded sequences
                                const T im = rcp((T)1 + an[0]);
                                cs = sqrtT(((T)1 + c.gamma * pn.p) * im);
                                nu = pn.mu * im * a.vm;
                            } else {
                                cs = sqrt(((T)1 + c.gamma * pn.p) / ((T)1 + an[0]));
                                nu = pn.mu / (an[0] + (T)1) * a.vm;
                            }
                            red[0] = fmax(red[0], cs + fabs(pn.u));
                            red[1] = fmax(red[1], cs + fabs(pn.v));
                            red[2] = fmax(red[2], cs + fabs(pn.w));
                            red[3] = fmax(red[3], nu);
                            for (int q = 0; q < 5; ++q)
                                if (!isfinite(an[q])) red[9] += 1;
                        }
                    }
                } else {
                    if (STAGE == 1) {
                        // the reference's target array holds the unfiltered density of U after the
                        // filter's buffer swap (src/Mesh.cpp:133), the other fields of the step before;
                        // A keeps its values there
                        if (rhoFiltered) acc[0] = cur[0];
                        for (int q = 0; q < 5; ++q) vS[q] = acc[q];
                        wS = true;
                    } else if (STAGE == 4) {
                        wA = true;  // the node's current values
                    } else {
                        wS = true;
                    }
                }
            }
            if (a.fuseOutlet && bx == a.nbx - 1) {
                // the outlet node (NI-1, j, k) takes the outlet values from this stage's results at NI-2 and
                // NI-3, two and one lanes down in the same (half-)tile (the row's last segment holds at least
                // three nodes), exactly as the edge-face pass right after this kernel would
                T* v = STAGE == 4 ? vA : vS;
                T n1[4], n2[4], ov[5];
                for (int q = 0; q < 4; ++q) {
                    n1[q] = shr1(v[q]);
                    n2[q] = shr1(n1[q]);
                }
                if (i == L.NI - 1 && j >= a.oj0 && j <= a.oj1 && k >= a.ok0 && k <= a.ok1) {
                    outletValues(c.gm1, n1, n2, ov);
                    for (int q = 0; q < 5; ++q) v[q] = ov[q];
                }
            }
            if (STAGE != 4 && wS)
                for (int q = 0; q < 5; ++q) m.stSout(q, bo, vS[q]);
            if (wA)
                for (int q = 0; q < 5; ++q) m.stA(q, bo, vA[q]);
            if (STAGE == 4 && KAHAN && wC)
                for (int q = 0; q < 5; ++q) m.stComp(q, bo, vC[q]);
            if (STAGE == 4 && KAHAN && upperSeg)
                for (int q = 0; q < 5; ++q) m.stComp(q, bo + 32 * (uint32_t)sizeof(T), 0);
        }
        // rotate the column
        umz = q0.u;
        vmz = q0.v;
        wmz = q0.w;
        q0 = q1;
        for (int q = 0; q < 5; ++q) {
            gz0[q] = gz1[q];
            phizm[q] = phiz[q];
        }
        if (NB == 1) __syncthreads();  // one buffer: this plane's readers are done before it is rewritten
    }
    if (SCAN) {
        // block partials: wave reduction, then one value per wave through LDS
        for (int q = 0; q < 10; ++q) {
            double v = (double)red[q];
            if (q < 4) v = wave_max(v);
            else v = wave_sum(v);
            if (lane == 0) wred[row][q] = v;
        }
        __syncthreads();
        if (row == 0 && lane < 10) {
            double v = 0;
            for (int r = 0; r < TY + 2; ++r) v = lane < 4 ? fmax(v, wred[r][lane]) : v + wred[r][lane];
            a.part[(size_t)lb * 10 + lane] = v;
        }
    }
}

// FP32: at least 4 waves per SIMD (<= 128 VGPRs); FP64: at least 2.
template <class T, int STAGE, int TY, bool SCAN, bool KAHAN, bool BUF>
__global__ __launch_bounds__(64 * (TY + 2)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_stage_fused(FusedArgs<T> a) {
    __shared__ Smem<T, TY> sm[smemBuffers<T>()];
    __shared__ double wred[TY + 2][10];
    // XCD-aware mapping: hardware block p runs on XCD p % 8; give each XCD a contiguous range of tiles
    const int p = blockIdx.x;
    const int per = (a.nblocks + 7) / 8;
    const int lb = (p % 8) * per + p / 8;
    if (lb >= a.nblocks) return;
    stageTile<T, STAGE, TY, SCAN, KAHAN, BUF>(a, lb, sm, wred);
}

// Many independent cases in one launch: the tiles of case c are the global tiles start[c] ..
// start[c+1]-1 (XCD-aware order as above, so each XCD sweeps a contiguous run of cases and tiles);
// tab[c] holds the case's arguments for this stage (its own layout, coefficients, buffers, dt).
template <class T, int STAGE, int TY, bool SCAN, bool KAHAN>
__global__ __launch_bounds__(64 * (TY + 2)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_stage_batched(const FusedArgs<T>* __restrict__ tab, const int* __restrict__ start, int ncases, int total) {
    __shared__ Smem<T, TY> sm[smemBuffers<T>()];
    __shared__ double wred[TY + 2][10];
    const int p = blockIdx.x;
    const int per = (total + 7) / 8;
    const int g = (p % 8) * per + p / 8;
    if (g >= total) return;
    int lo = 0, hi = ncases - 1;  // last case whose first tile is <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    stageTile<T, STAGE, TY, SCAN, KAHAN, sizeof(T) == 4>(tab[lo], g - start[lo], sm, wred);
}

// The same stage as a persistent launch: gridDim.x = the blocks the GPU holds at once (a multiple of 8).
// The tiles are split into 8 contiguous ranges as above, one per XCD; a block takes the next tile of its
// own XCD's range (hardware block p runs on XCD p % 8) and, once that is exhausted, the next tiles of the
// other ranges, so no XCD idles while another still has tiles. ctr[x] counts the tiles taken from range x,
// ctr[8] the blocks finished; the last block to finish zeroes all nine for the next launch (no memset
// node in a captured step). Every block leaves the loop once all ranges are exhausted.
template <class T, int STAGE, int TY, bool SCAN, bool KAHAN>
__global__ __launch_bounds__(64 * (TY + 2)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_stage_batched_persistent(const FusedArgs<T>* __restrict__ tab, const int* __restrict__ start, int ncases,
                                int total, unsigned* __restrict__ ctr) {
    __shared__ Smem<T, TY> sm[smemBuffers<T>()];
    __shared__ double wred[TY + 2][10];
    __shared__ int taken;
    const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
    const int per = (total + 7) / 8;
    for (int r = 0; r < 8; ++r) {
        const int x = (int)((blockIdx.x + r) & 7);
        const int lo0 = x * per, n = min(per, total - lo0);
        while (n > 0) {
            if (leader) taken = (int)atomicAdd(&ctr[x], 1u);
            __syncthreads();
            const int t = taken;
            __syncthreads();  // every thread has read `taken` before the leader writes it again
            if (t >= n) break;
            const int g = lo0 + t;
            int lo = 0, hi = ncases - 1;  // last case whose first tile is <= g
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (start[mid] <= g) lo = mid;
                else hi = mid - 1;
            }
            stageTile<T, STAGE, TY, SCAN, KAHAN, sizeof(T) == 4>(tab[lo], g - start[lo], sm, wred);
            __syncthreads();  // this tile's LDS readers are done before the next tile's prologue writes it
        }
    }
    if (leader) {
        __threadfence();
        if (atomicAdd(&ctr[8], 1u) == gridDim.x - 1)
            for (int q = 0; q < 9; ++q) atomicExch(&ctr[q], 0u);
    }
}

// ---- force on the body: the discrete momentum balance of the flux form. Summed over all fluid interior
// nodes, the right-hand sides telescope to the face fluxes across the outer boundary and across the faces
// between interior nodes and ghost nodes; the latter are the momentum per unit time the body gives the
// fluid. The force on the body is minus that sum times the node volume. Evaluated in FP64 from the state
// after the step's last boundary application, with the stage kernel's own face-flux functions.
template <class T>
struct ForceArgs {
    Layout L;
    Window W;
    Coef<double> c;
    const T* S;
    const uint8_t* cls;
    const T* rhoPre;
    const T* rhoEPre;
};

template <class T>
__device__ void primsAt(const ForceArgs<T>& a, int i, int j, int k, Pt<double>& q) {
    const Layout& L = a.L;
    i = min(max(i, 0), L.NI - 1);
    j = min(max(j, 0), L.NJ - 1);
    k = min(max(k, 0), L.NK - 1);
    const size_t o = L.off(i, j, k);
    q.rho = (double)a.S[o];
    q.ru = (double)a.S[o + L.fstride];
    q.rv = (double)a.S[o + 2 * L.fstride];
    q.rw = (double)a.S[o + 3 * L.fstride];
    q.E = (double)a.S[o + 4 * L.fstride];
    double rp = q.rho, Ep = q.E;
    if (a.W.contains(i, j, k) && (a.cls[o] & CLS_FILTER)) {  // primitives stay pre-filter there
        const size_t w = a.W.idx(i, j, k);
        rp = (double)a.rhoPre[w];
        Ep = (double)a.rhoEPre[w];
    }
    prims(a.c, rp, q.ru, q.rv, q.rw, Ep, q);
}

// Point with all nine velocity derivatives (central differences; dz* returned separately).
template <class T>
__device__ void pointAt(const ForceArgs<T>& a, int i, int j, int k, Pt<double>& q, double dz[3]) {
    primsAt(a, i, j, k, q);
    Pt<double> m, p;
    primsAt(a, i - 1, j, k, m);
    primsAt(a, i + 1, j, k, p);
    q.dxu = (p.u - m.u) * a.c.i2dx;
    q.dxv = (p.v - m.v) * a.c.i2dx;
    q.dxw = (p.w - m.w) * a.c.i2dx;
    primsAt(a, i, j - 1, k, m);
    primsAt(a, i, j + 1, k, p);
    q.dyu = (p.u - m.u) * a.c.i2dy;
    q.dyv = (p.v - m.v) * a.c.i2dy;
    q.dyw = (p.w - m.w) * a.c.i2dy;
    primsAt(a, i, j, k - 1, m);
    primsAt(a, i, j, k + 1, p);
    dz[0] = (p.u - m.u) * a.c.i2dz;
    dz[1] = (p.v - m.v) * a.c.i2dz;
    dz[2] = (p.w - m.w) * a.c.i2dz;
}

template <class T>
__device__ void nodalFlux(const ForceArgs<T>& a, int d, int i, int j, int k, double g[5], Pt<double>& q) {
    double dz[3];
    pointAt(a, i, j, k, q, dz);
    if (d == 0) gx(a.c, q, dz[0], dz[2], g);
    else if (d == 1) gy(a.c, q, dz[1], dz[2], g);
    else gz(a.c, q, g);
}

// One block partial (3 force components on the fluid, before the volume factor) per block.
template <class T>
__global__ void k_body_force(ForceArgs<T> a, const int64_t* ghosts, int64_t n, double* partials) {
    double f[3] = {0, 0, 0};
    const Layout& L = a.L;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t id = ghosts[q];
        const int gi = (int)(id % L.NI), gj = (int)((id / L.NI) % L.NJ), gk = (int)(id / ((int64_t)L.NI * L.NJ));
        for (int d = 0; d < 3; ++d)
            for (int s = -1; s <= 1; s += 2) {
                const int pi = gi + (d == 0 ? s : 0), pj = gj + (d == 1 ? s : 0), pk = gk + (d == 2 ? s : 0);
                if (pi < 0 || pi >= L.NI || pj < 0 || pj >= L.NJ || pk < 0 || pk >= L.NK) continue;
                if ((a.cls[L.off(pi, pj, pk)] & CLS_TYPE_MASK) != NS_FLUID_INTERIOR) continue;
                // face between the lower node lo and the upper node up along d
                const bool pUp = s > 0;
                double glo[5], gup[5], cf[5];
                Pt<double> lo, up;
                nodalFlux(a, d, pUp ? gi : pi, pUp ? gj : pj, pUp ? gk : pk, glo, lo);
                nodalFlux(a, d, pUp ? pi : gi, pUp ? pj : gj, pUp ? pk : gk, gup, up);
                const double i2d = d == 0 ? a.c.i2dx : (d == 1 ? a.c.i2dy : a.c.i2dz);
                if (d == 0) compact<double, 0>(a.c, lo, up, cf);
                else if (d == 1) compact<double, 1>(a.c, lo, up, cf);
                else compact<double, 2>(a.c, lo, up, cf);
                // the face flux enters RHS(lower) with +, RHS(upper) with -
                const double sg = pUp ? -1.0 : 1.0;
                for (int c = 0; c < 3; ++c) f[c] += sg * ((glo[1 + c] + gup[1 + c]) * i2d + cf[1 + c]);
            }
    }
    __shared__ double red[3][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = 0; c < 3; ++c) {
        double v = wave_sum(f[c]);
        if (lane == 0) red[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[threadIdx.x][w];
        partials[(size_t)blockIdx.x * 3 + threadIdx.x] = v;
    }
}

__global__ void k_force_finalize(const double* partials, int nb, double vol, StepScalars* st, double* ring, int cap) {
    __shared__ double red[3][4];
    double f[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int c = 0; c < 3; ++c) f[c] += partials[(size_t)b * 3 + c];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = 0; c < 3; ++c) {
        double v = wave_sum(f[c]);
        if (lane == 0) red[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int slot = (int)(st->forceCount % (uint64_t)cap);
        ring[4 * slot] = (double)st->level;
        for (int c = 0; c < 3; ++c) {
            double v = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[c][w];
            ring[4 * slot + 1 + c] = -vol * v;  // force on the body
        }
        st->forceCount += 1;
    }
}

template <class T, int TY>
int fusedBlocks(const Layout& L, int kc) {
    return ((L.NI + 61) / 62) * ((L.NJ + TY - 1) / TY) * ((L.NK + kc - 1) / kc);
}

template <class T, int TY, bool K, bool B>
void launchTYK(hipStream_t s, int stage, const FusedArgs<T>& a) {
    const int grid = (a.nblocks + 7) / 8 * 8;
    dim3 blk(64, TY + 2);
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_stage_fused<T, 1, TY, false, K, B>), dim3(grid), blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_stage_fused<T, 2, TY, false, K, B>), dim3(grid), blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_stage_fused<T, 3, TY, false, K, B>), dim3(grid), blk, 0, s, a); break;
        default:
            if (a.part) hipLaunchKernelGGL((k_stage_fused<T, 4, TY, true, K, B>), dim3(grid), blk, 0, s, a);
            else hipLaunchKernelGGL((k_stage_fused<T, 4, TY, false, K, B>), dim3(grid), blk, 0, s, a);
            break;
    }
}

template <class T, int TY>
void launchTY(hipStream_t s, int stage, const FusedArgs<T>& a, bool buf) {
    if constexpr (sizeof(T) == 4) {
        if (a.comp[0]) buf ? launchTYK<T, TY, true, true>(s, stage, a) : launchTYK<T, TY, true, false>(s, stage, a);
        else buf ? launchTYK<T, TY, false, true>(s, stage, a) : launchTYK<T, TY, false, false>(s, stage, a);
    } else {
        launchTYK<T, TY, false, false>(s, stage, a);
    }
}

// Blocks of `kern` (threads per block) that the device holds at once, rounded down to a multiple of 8 (one
// per XCD): the grid of a persistent launch. Found once per kernel and device.
int residentBlocks(const void* kern, int threads) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    std::lock_guard<std::mutex> lock(mu);
    auto it = known.find({kern, dev});
    if (it != known.end()) return it->second;
    hipDeviceProp_t prop;
    int perCU = 0;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, threads, 0) != hipSuccess)
        throw std::runtime_error("occupancy query failed");
    const int n = std::max(8, perCU * prop.multiProcessorCount / 8 * 8);
    known[{kern, dev}] = n;
    return n;
}

template <class T, int TY, bool K>
void launchBatchedTYK(hipStream_t s, int stage, bool scan, const FusedArgs<T>* tab, const int* start, int ncases,
                      int total, unsigned* ctr) {
    dim3 blk(64, TY + 2);
    if (ctr) {
        // persistent: as many blocks as the GPU holds at once (found once per kernel), a multiple of 8
        auto launchP = [&](auto kern) {
            const int resident = residentBlocks(reinterpret_cast<const void*>(kern), 64 * (TY + 2));
            hipLaunchKernelGGL(kern, dim3(resident), blk, 0, s, tab, start, ncases, total, ctr);
        };
        switch (stage) {
            case 1: launchP(k_stage_batched_persistent<T, 1, TY, false, K>); break;
            case 2: launchP(k_stage_batched_persistent<T, 2, TY, false, K>); break;
            case 3: launchP(k_stage_batched_persistent<T, 3, TY, false, K>); break;
            default:
                if (scan) launchP(k_stage_batched_persistent<T, 4, TY, true, K>);
                else launchP(k_stage_batched_persistent<T, 4, TY, false, K>);
                break;
        }
        return;
    }
    const int grid = (total + 7) / 8 * 8;
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_stage_batched<T, 1, TY, false, K>), dim3(grid), blk, 0, s, tab, start, ncases, total); break;
        case 2: hipLaunchKernelGGL((k_stage_batched<T, 2, TY, false, K>), dim3(grid), blk, 0, s, tab, start, ncases, total); break;
        case 3: hipLaunchKernelGGL((k_stage_batched<T, 3, TY, false, K>), dim3(grid), blk, 0, s, tab, start, ncases, total); break;
        default:
            if (scan) hipLaunchKernelGGL((k_stage_batched<T, 4, TY, true, K>), dim3(grid), blk, 0, s, tab, start, ncases, total);
            else hipLaunchKernelGGL((k_stage_batched<T, 4, TY, false, K>), dim3(grid), blk, 0, s, tab, start, ncases, total);
            break;
    }
}

int normTY(int ty) { return ty == 4 || ty == 6 || ty == 14 ? ty : 8; }

// Every argument of one stage launch on one case, tile geometry included.
inline int chunkPlanes(const FusedTile& t) { return t.kc > 0 ? t.kc : 64; }

template <class T>
void fillArgs(FusedArgs<T>& a, int stage, const Layout& L, const Params& P, const Window& W, const T* Sin, const T* U,
              T* Sout, T* A, const uint8_t* cls, const T* rhoPre, const T* rhoEPre, const double* dtdev, FusedTile tile,
              double* partials, int norms, const T* filteredRho, const int32_t* filterFlag, int8_t* comp,
              const EdgeFace* outlet) {
    a = FusedArgs<T>{};
    if (comp && sizeof(T) != 4) throw std::invalid_argument("compensated combine needs FP32 fields");
    a.fRho = stage == 1 ? filteredRho : nullptr;
    a.fFlag = stage == 1 ? filterFlag : nullptr;
    a.part = stage == 4 ? partials : nullptr;
    a.norms = norms;
    a.vm = (T)std::max(4. / 3, P.gamma / P.Pr);
    a.L = L;
    a.W = W;
    a.c = makeCoef<T>(P);
    // byte offsets within a tile (its z-chunk, the plane below and the plane above) are 32-bit
    if ((size_t)(chunkPlanes(tile) + 2) * L.plane * sizeof(T) > 0xFFFFFFFFull)
        throw std::length_error("z-chunk of more than 4 GiB of planes: 32-bit tile offsets");
    for (int q = 0; q < 5; ++q) {
        a.Sin[q] = Sin + q * L.fstride;
        a.U[q] = U + q * L.fstride;
        a.Sout[q] = Sout ? Sout + q * L.fstride : nullptr;
        a.A[q] = A + q * L.fstride;
        a.comp[q] = comp ? comp + q * L.fstride : nullptr;
    }
    a.cls = cls;
    a.rhoPre = rhoPre;
    a.rhoEPre = rhoEPre;
    a.dtdev = dtdev;
    const int ty = normTY(tile.ty);
    a.kc = chunkPlanes(tile);
    a.nbx = (L.NI + 61) / 62;
    a.nby = (L.NJ + ty - 1) / ty;
    a.nbz = (L.NK + a.kc - 1) / a.kc;
    a.tailHalf = L.NI - Layout::kSegNodes * (a.nbx - 1) <= 30;
    a.nblocks = ((a.nbx - a.tailHalf) * a.nby + (a.tailHalf ? (a.nby + 1) / 2 : 0)) * a.nbz;
    if (outlet) {
        if (!outletFusable(L, *outlet)) throw std::invalid_argument("outlet face cannot be applied in the stage kernel");
        a.fuseOutlet = 1;
        a.oj0 = outlet->j0;
        a.oj1 = outlet->j1;
        a.ok0 = outlet->k0;
        a.ok1 = outlet->k1;
    }
}

}  // namespace

template <class T>
int launchStageFused(hipStream_t s, int stage, const Layout& L, const Params& P, const Window& W, const T* Sin,
                     const T* U, T* Sout, T* A, const uint8_t* cls, const T* rhoPre, const T* rhoEPre,
                     const double* dtdev, FusedTile tile, double* partials, int norms, const T* filteredRho,
                     const int32_t* filterFlag, int8_t* comp, const EdgeFace* outlet) {
    FusedArgs<T> a;
    fillArgs(a, stage, L, P, W, Sin, U, Sout, A, cls, rhoPre, rhoEPre, dtdev, tile, partials, norms, filteredRho,
             filterFlag, comp, outlet);
    const bool buf = !tile.globalAccess && stateFitsBuffer(L);
    switch (normTY(tile.ty)) {
        case 4: launchTY<T, 4>(s, stage, a, buf); break;
        case 6: launchTY<T, 6>(s, stage, a, buf); break;
        case 14: launchTY<T, 14>(s, stage, a, buf); break;
        default: launchTY<T, 8>(s, stage, a, buf); break;
    }
    return a.nblocks;
}

size_t fusedArgsBytes() { return sizeof(FusedArgs<float>); }

int buildFusedArgs(void* dst, int stage, const Layout& L, const Params& P, const Window& W, const float* Sin,
                   const float* U, float* Sout, float* A, const uint8_t* cls, const float* rhoPre, const float* rhoEPre,
                   const double* dtdev, FusedTile tile, double* partials, int norms, int8_t* comp,
                   const EdgeFace* outlet, const float* filteredRho, const int32_t* filterFlag) {
    if (!stateFitsBuffer(L)) throw std::length_error("batched case: state buffer beyond 4 GiB");
    FusedArgs<float> a;
    fillArgs<float>(a, stage, L, P, W, Sin, U, Sout, A, cls, rhoPre, rhoEPre, dtdev, tile, partials, norms, filteredRho,
                    filterFlag, comp, outlet);
    std::memcpy(dst, &a, sizeof a);
    return a.nblocks;
}

void launchStageBatched(hipStream_t s, int stage, int ty, bool scan, bool kahan, const void* tab, const int* start,
                        int ncases, int total, unsigned* ctr) {
    const FusedArgs<float>* t = static_cast<const FusedArgs<float>*>(tab);
    switch (normTY(ty)) {
        case 4: kahan ? launchBatchedTYK<float, 4, true>(s, stage, scan, t, start, ncases, total, ctr)
                      : launchBatchedTYK<float, 4, false>(s, stage, scan, t, start, ncases, total, ctr); break;
        case 6: kahan ? launchBatchedTYK<float, 6, true>(s, stage, scan, t, start, ncases, total, ctr)
                      : launchBatchedTYK<float, 6, false>(s, stage, scan, t, start, ncases, total, ctr); break;
        case 14: kahan ? launchBatchedTYK<float, 14, true>(s, stage, scan, t, start, ncases, total, ctr)
                       : launchBatchedTYK<float, 14, false>(s, stage, scan, t, start, ncases, total, ctr); break;
        default: kahan ? launchBatchedTYK<float, 8, true>(s, stage, scan, t, start, ncases, total, ctr)
                       : launchBatchedTYK<float, 8, false>(s, stage, scan, t, start, ncases, total, ctr); break;
    }
}

int fusedMaxBlocks(const Layout& L) { return fusedBlocks<float, 4>(L, kMinTileKc); }

int bodyForceBlocks(int64_t nGhost) { return (int)std::min<int64_t>(256, std::max<int64_t>(1, (nGhost + 255) / 256)); }

template <class T>
void launchBodyForce(hipStream_t s, const Layout& L, const Params& P, const Window& W, const T* S, const uint8_t* cls,
                     const T* rhoPre, const T* rhoEPre, const int64_t* ghosts, int64_t nGhost, double* partials,
                     StepScalars* st, double* ring, int cap) {
    ForceArgs<T> a;
    a.L = L;
    a.W = W;
    a.c = makeCoef<double>(P);
    a.S = S;
    a.cls = cls;
    a.rhoPre = rhoPre;
    a.rhoEPre = rhoEPre;
    const int nb = bodyForceBlocks(nGhost);
    hipLaunchKernelGGL(k_body_force<T>, dim3(nb), dim3(256), 0, s, a, ghosts, nGhost, partials);
    hipLaunchKernelGGL(k_force_finalize, dim3(1), dim3(256), 0, s, partials, nb, P.hx * P.hy * P.hz, st, ring, cap);
}

template void launchBodyForce<double>(hipStream_t, const Layout&, const Params&, const Window&, const double*,
                                      const uint8_t*, const double*, const double*, const int64_t*, int64_t, double*,
                                      StepScalars*, double*, int);
template void launchBodyForce<float>(hipStream_t, const Layout&, const Params&, const Window&, const float*,
                                     const uint8_t*, const float*, const float*, const int64_t*, int64_t, double*,
                                     StepScalars*, double*, int);

const int* fusedTileHeights(int* n) {
    static const int v[4] = {4, 6, 8, 14};
    *n = 4;
    return v;
}

// z-chunk lengths tried by the autotuners: not only powers of two, since the number of blocks per launch
// against the blocks the chip holds at once sets how idle the last round of blocks leaves it
const int* fusedChunkPlanes(int* n) {
    static const int v[10] = {12, 16, 20, 24, 32, 40, 48, 64, 96, 128};
    *n = 10;
    return v;
}

template int launchStageFused<double>(hipStream_t, int, const Layout&, const Params&, const Window&, const double*,
                                      const double*, double*, double*, const uint8_t*, const double*, const double*,
                                      const double*, FusedTile, double*, int, const double*, const int32_t*,
                                      int8_t*, const EdgeFace*);
template int launchStageFused<float>(hipStream_t, int, const Layout&, const Params&, const Window&, const float*,
                                     const float*, float*, float*, const uint8_t*, const float*, const float*,
                                     const double*, FusedTile, double*, int, const float*, const int32_t*,
                                     int8_t*, const EdgeFace*);

}  // namespace nsfsi
