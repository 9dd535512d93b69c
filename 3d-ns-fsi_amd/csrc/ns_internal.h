/*
 * ns_internal.h — shared host/device declarations of the MI355X Navier–Stokes step.
 *
 * Device field layout: structure of arrays, x fastest, x-rows in segments. The nodes of a row are cut into
 * segments of 62 consecutive nodes; segment s occupies the 64 slots 64 s .. 64 s + 63 of the row and node
 * i = 62 s + r sits in slot 64 s + r + 1, so slots 0 and 63 of every segment hold no node (padding). A
 * segment is then 256 B (FP32) or 512 B (FP64) and starts on such a boundary: the fused stage kernel,
 * whose x-tiles are the segments, writes whole aligned cache lines (a line written in part costs a
 * read-modify-write in the memory system). Every x-row is padded (pitch) to whole segments in FP32 (so the
 * 8-bit compensation carry, at a quarter of the byte offsets, takes 64 aligned bytes per segment) and to a
 * multiple of 128 bytes in FP64; every field starts on a 256-byte boundary;
 * offset(i,j,k) = (k*NJ + j)*pitch + col(i). The five
 * conserved fields of one state (rho', rho*u, rho*v, rho*w, rho*E') are consecutive fields of one
 * allocation.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/ns_fsi.h"

#if defined(__HIPCC__)
#define NS_HD __host__ __device__
#else
#define NS_HD
#endif

namespace nsfsi {

void setDerived(ns_config& c);
std::string validateConfig(const ns_config& c);

// Node class byte: low 3 bits = ns_node_type; flag bits below.
enum : uint8_t {
    CLS_TYPE_MASK = 0x07,
    CLS_FILTER = 0x08,     // FluidInterior node filtered by the immersed BC (src/Boundary.cpp:618-652)
    CLS_DT_SKIP = 0x10,    // node excluded from the dt scan on the device (its value is accounted elsewhere)
    CLS_INACTIVE_IC = 0x20 // SolidInactive node whose primitive values stay at their initial values
};

struct Layout {
    static constexpr int kSegNodes = 62;  // nodes per x-segment
    static constexpr int kSegSlots = 64;  // slots per x-segment (first and last are padding)
    int NI, NJ, NK;
    int pitch;        // slots per padded x-row
    size_t plane;     // NJ * pitch
    size_t fstride;   // elements between consecutive fields (>= NK * plane, 256-B aligned)
    NS_HD size_t nodes() const { return (size_t)NI * NJ * NK; }
    // slot of node i in its row (i >= -1; node -1 maps to the padding slot 0)
    NS_HD static int col(int i) { return i + 2 * (i / kSegNodes) + 1; }
    NS_HD size_t off(int i, int j, int k) const { return ((size_t)k * NJ + j) * pitch + col(i); }
    // off(i + d, j, k) - off(i, j, k): the x-step of d nodes from node i (crosses padding slots)
    NS_HD static long xstep(int i, int d) { return (long)col(i + d) - col(i); }
};

Layout makeLayout(int NI, int NJ, int NK, size_t elemBytes);

// Axis-aligned index window holding every solid, ghost and filter node of the body.
struct Window {
    int i0, j0, k0, ni, nj, nk;
    NS_HD bool contains(int i, int j, int k) const {
        return i >= i0 && i < i0 + ni && j >= j0 && j < j0 + nj && k >= k0 && k < k0 + nk;
    }
    NS_HD size_t idx(int i, int j, int k) const { return ((size_t)(k - k0) * nj + (j - j0)) * ni + (i - i0); }
    NS_HD size_t size() const { return (size_t)ni * nj * nk; }
};

// Scalars every kernel needs (spacings, physical constants), passed by value.
struct Params {
    double hx, hy, hz;         // grid spacing
    double ox, oy, oz;         // origin offset in nodes
    double gamma, Pr, Re, ScPlusOne, M_0;
    double rampTime;
    // optional inlet velocity profile: factor per node of the inlet plane (index a + n_a * b over the
    // plane's two in-plane axes in (i, j, k) order), null = uniform (the reference's inlet)
    const double* inflow;
};

// Owned node box of one edge boundary (src/Boundary.cpp:18-71), with its type and side.
struct EdgeFace {
    int axis, side, type;
    int i0, i1, j0, j1, k0, k1;  // inclusive
};

}  // namespace nsfsi
