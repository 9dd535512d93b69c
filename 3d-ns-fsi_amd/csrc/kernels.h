/*
 * kernels.h — launch wrappers of the device kernels (declarations). All launches go to the stream
 * passed in; none synchronises except compactWindow (set-up only), which reads its counts back.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ns_internal.h"

namespace nsfsi {

// Throws a HipError (mapped to NS_ERR_HIP at the C-ABI) when e is not hipSuccess.
void hipCheck(hipError_t e, const char* what);

// Immersed body description for classification (reference Body classes, src/Boundary.h:200-346).
struct BodyDesc {
    int type;   // ns_body_type
    int axis;   // cylinder axis
    double cx, cy, cz, r;
    double mpb;         // machine precision buffer
    double filterWidth; // layer * max spacing across the body's cross-section axes
    Window box;         // body bounding box (getSphereBoundingBox / getCylinderBoundingBox)
};

// The immersed bodies in the order of the reference's body list (src/Mesh.cpp:52-57); their boxes are
// pairwise disjoint.
struct BodySet {
    int n;
    BodyDesc b[NS_MAX_BODIES];
};

// Per-ghost data of the image-point interpolation (frozen geometry).
struct GhostTables {
    int64_t* node;      // layout offset of the ghost node
    int64_t* corner;    // 8 layout offsets per ghost (asIndexList order)
    uint8_t* mask;      // bit c set: corner c is a solid ghost
    double* rel;        // 3 per ghost: IP minus lower box corner (all-fluid interpolation)
    double* ip;         // 3 per ghost
    double* bi;         // 3 per ghost
    double* luD;        // 64 per ghost: LU factors of the Dirichlet Vandermonde matrix
    double* luN;        // 64 per ghost: LU factors of the Neumann Vandermonde matrix
    int8_t* perm;       // 32 per ghost: row/col permutations (D rows, D cols, N rows, N cols)
    double* wD;         // 8 per ghost: image-point weights of the Dirichlet interpolant, b(IP)^T V_D^-1
    double* wN;         // 8 per ghost: the same for the Neumann interpolant
    int32_t* boxLo;     // 3 per ghost: lower corner (i,j,k)
};

// ---- classification and ghost geometry (mesh.hip, compiled without FMA contraction)
void launchClassify(hipStream_t s, const Layout& L, uint8_t* cls, const int faceNodeType[6], const BodySet& bodies,
                    const Params& P);
void launchMarkGhosts(hipStream_t s, const Layout& L, uint8_t* cls, const Window& W);
// Compacts node ids n = (k*NJ+j)*NI+i of window nodes whose class satisfies (cls & mask) == value, in
// ascending order. blockCounts must hold ceil(W.size()/256) ints. Returns the count (synchronises).
int64_t compactWindow(hipStream_t s, const Layout& L, const uint8_t* cls, const Window& W, uint8_t mask, uint8_t value,
                      int64_t* out, int32_t* blockCounts, int32_t* blockOffsets);
// Image points of the ghosts of body B; a SolidInactive box corner becomes a ghost. A corner outside B's own
// box sets error bit 2 (the box reaches another body's solid nodes).
void launchImagePoints(hipStream_t s, const Layout& L, uint8_t* cls, const Params& P, const BodyDesc& B,
                       const int64_t* ghosts, int64_t n, double* bi, double* ip, int32_t* changed, int32_t* error);
void launchFillGhostMap(hipStream_t s, const Layout& L, const Window& W, const int64_t* ghosts, int64_t n, int32_t* gmap);
// Ghost geometry of one body's ghosts, entries [base, base + n) of the ghost list; G points at entry base.
// A ghost corner of another body sets error bit 2.
void launchGhostGeometry(hipStream_t s, const Layout& L, const uint8_t* cls, const Params& P, const Window& W,
                         const int64_t* ghosts, int64_t base, int64_t n, const int32_t* gmap, GhostTables G,
                         int32_t* error);
// The tables of entries [base, ...) of a ghost list.
GhostTables ghostTablesFrom(const GhostTables& G, int64_t base);

// ---- boundary conditions on a state buffer (mesh.hip)
// Whether two edge faces can be applied in one launch: disjoint owned boxes and neither reads a node the
// other writes (the planes one and two in along the normal, or a periodic face's source plane).
bool edgeFacesIndependent(const Layout& L, const EdgeFace& x, const EdgeFace& y);
// All six edge faces in application order, independent faces sharing one launch (same result as
// launchEdgeFace face by face).
// prim != null (the stored-primitive form): the faces also write the primitive and transport values the
// reference stores with each node, and symmetry faces mirror the stored primitives.
template <class T>
void launchEdgeFaces(hipStream_t s, const Layout& L, const Params& P, const EdgeFace faces[6], T* S, const double* tdev,
                     const double* dtdev, double stageFrac, T* prim = nullptr);
template <class T>
void launchEdgeFace(hipStream_t s, const Layout& L, const Params& P, const EdgeFace& f, T* S, const double* tdev,
                    const double* dtdev, double stageFrac);
template <class T>
void launchFilterCompute(hipStream_t s, const Layout& L, const Window& W, const int64_t* fnodes, int64_t n, const T* S,
                         T* tmp, T* rhoPre, T* rhoEPre);
template <class T>
void launchFilterWrite(hipStream_t s, const Layout& L, const int64_t* fnodes, int64_t n, T* S, const T* tmp);
// prim != null: also the ghosts' stored primitive and transport values, derived from their new state.
template <class T>
void launchGhostApply(hipStream_t s, const Layout& L, const Params& P, const GhostTables& G, int64_t n, T* S,
                      T* prim = nullptr);

}  // namespace nsfsi
