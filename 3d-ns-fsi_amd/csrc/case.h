/*
 * case.h — one simulation case on one GPU (host object behind an ns_handle).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"
#include "ns_internal.h"
#include "stage.h"

namespace nsfsi {

struct ConfigError : std::runtime_error { using std::runtime_error::runtime_error; };
struct GeometryError : std::runtime_error { using std::runtime_error::runtime_error; };
struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };
struct StateError : std::runtime_error { using std::runtime_error::runtime_error; };
struct Unsupported : std::runtime_error { using std::runtime_error::runtime_error; };

void hipCheck(hipError_t e, const char* what);

class Case {
public:
    explicit Case(const ns_config& c);
    ~Case();
    Case(const Case&) = delete;
    Case& operator=(const Case&) = delete;

    void init();
    void setInitialField(const double* U5);
    void step(int nsteps);
    void getField(int which, double* out);
    void nodeTypes(uint8_t* out);
    void ghosts(int32_t* ijk, double* bi, double* ip, int32_t* box);
    void filterNodes(int32_t* ijk);
    int64_t interiorCount();
    void totals(double out[5]);
    void readScalars();
    bool filterDisabled() const;
    // Runs nsteps steps with HIP events around each kernel group; ms[g] receives the mean duration of
    // one launch of group g (0 stage kernel, 1 primitive pass, 2 boundary kernels of one stage,
    // 3 end-of-step reductions) and launches[g] the launches per step.
    void profile(int nsteps, double ms[4], int launches[4]);
    // Run-time options: "stage_kernel" 0 = unfused (primitive pass + per-node RHS), 1 = fused.
    void setOption(const std::string& name, int64_t value);
    int stageKernel = 1;
    // The FP64 unfused stage keeps the reference's stored primitive and transport fields (`prim`): interior
    // nodes updated after each stage combine, the other nodes written by the boundary conditions, untouched
    // solid nodes at their initial values, and the reference's own final combine. Bit for bit the
    // reference's arithmetic, given a correctly rounded x^1.5.
    bool literal() const { return stageKernel == 0 && !fp32; }
    // `prim` holds the stored values (set by the initial state, kept by stored-primitive steps; other steps
    // and primitive read-outs in the other forms leave it stale, and the next stored-primitive step first
    // derives it from the current state)
    bool primsStored = false;
    void syncStoredPrims();
    bool fusedScans = true;  // end-of-step scans inside the stage-4 kernel (else a separate full pass)
    FusedTile tile{};
    // Times every fused-stage tile shape on this case (stage-2 launches into scratch outputs, state
    // untouched) and keeps the fastest; returns the chosen tile and its time per launch in ms.
    double autotune(int reps);
    std::vector<double> tuneTimes;  // ms per launch of each tried shape, in try order
    // HIP-graph replay of the time step: one graph per buffer parity (which allocation holds the state)
    bool useGraph = false;
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    void* graphState[2] = {nullptr, nullptr};
    void dropGraphs();

    ns_config cfg;
    bool fp32 = false;
    size_t eb = 8;
    bool initialized = false;
    hipStream_t stream = nullptr;
    Layout L{};
    Params P{};
    Window W{};
    BodySet bodies{};
    EdgeFace faces[6]{};  // application order (reverse of the claim order)
    // faces[fusedOutlet] (or -1): the outlet face that the fused stage kernel applies itself; bcFaces are the
    // faces with that one left out, for the edge-face pass after a fused stage
    int fusedOutlet = -1;
    EdgeFace bcFaces[6]{};
    bool fuseOutletOpt = true;  // option "fused_outlet"
    const EdgeFace* stageOutlet() const;  // the face the fused stage kernel applies, or null

    // classification results
    uint8_t* cls = nullptr;
    int64_t nGhost = 0, nFilter = 0, nInactive = 0;
    // body b's entries of the ghost and filter lists: [ghostOff[b], ghostOff[b+1]), [filterOff[b], filterOff[b+1])
    int64_t ghostOff[NS_MAX_BODIES + 1]{}, filterOff[NS_MAX_BODIES + 1]{};
    int64_t* ghostList = nullptr;
    int64_t* filterList = nullptr;
    int64_t* inactiveList = nullptr;
    double* inactiveIC = nullptr;
    // Boundary nodes whose conserved values are constant: symmetry mirrors (possibly through periodic copies
    // and further mirrors) of untouched SolidInactive nodes, whose stored primitives the reference keeps at
    // their initial values. fixedNodes: layout offsets; fixedSrc: layout offset of the untouched source node;
    // fixedMirror: bit a set = velocity component a negated; fixedVals: 5 conserved values per node.
    int64_t nFixed = 0;
    std::vector<int64_t> fixedHostNodes, fixedSrc;
    std::vector<uint8_t> fixedMirror;
    int64_t* fixedNodes = nullptr;
    double* fixedVals = nullptr;
    void findFixedBoundary();
    template <class T> void computeFixedBoundary(const std::vector<T>& h);
    GhostTables G{};

    // fields (element type float or double)
    void *U = nullptr, *A = nullptr, *S1 = nullptr, *S2 = nullptr, *prim = nullptr;
    void *rhoPre = nullptr, *rhoEPre = nullptr, *ftmp = nullptr;
    int8_t* comp = nullptr;  // FP32 compensated combine: carried rounding error, 5 fields of 8 bits

    // step scalars and reductions
    StepScalars* scal = nullptr;
    double* partials = nullptr;
    int nPartials = 0;
    size_t partialSlots = 0;  // capacity of `partials` in slots of reducePartialSlots() doubles
    int nStage4Partials = 0;
    double* hist = nullptr;
    int histCap = 1024;
    std::vector<double> normsHost[5];
    // force on the body after every step while the "body_force" option is on: (timeLevel, Fx, Fy, Fz)
    bool bodyForce = false;
    double* forceRing = nullptr;
    std::vector<double> forceHost;
    // prescribed step sizes (index = time level) instead of the stability limit; and the inlet profile
    void setDtSchedule(const double* dts, int64_t n);
    void setInflowProfile(const double* q, int64_t n);
    const double* dtSchedule = nullptr;
    int64_t dtScheduleLen = 0;
    // Every path that advances the state calls this first: with a prescribed schedule, the step sizes of
    // time levels level .. level + nsteps must all be in it (the finalize kernel reads entry level + 1
    // after each step). Throws std::invalid_argument otherwise, before any step runs.
    void checkSchedule(int64_t nsteps) const;
    double* dtScheduleBuf = nullptr;
    int64_t dtScheduleCap = 0;
    double* inflowBuf = nullptr;
    // Stage states of one step for the adjoint: S[0] = the step's input state and its pre-filter tables,
    // S[1..3] = the outputs of stages 1-3 after their boundary application, with the tables of each.
    struct StageTape {
        double* S[4];
        double* P[4];
        double* PE[4];
    };
    // One step (kernel by kernel) recording its stage states into tape, in FP64 (an FP32 case's widened).
    void stepTaped(StageTape* tape);
    // Discrete adjoint of J = mean over time levels n0..n1 of C_D after nsteps steps from the current
    // (initial) state, at the step sizes of the forward run (frozen); see ns_adjoint_drag.
    void adjointDrag(int nsteps, int n0, int n1, int ckpt, double* J, double* gradU0, double* gradInflow,
                     double* gradRe, double* dts);
    double referenceArea() const;
    // Checkpoint of everything a later step reads: both state buffers (the "Old" values of non-interior nodes
    // live in A), the stored primitive fields (FP64), the pre-filter tables, the FP32 compensation carry,
    // the untouched solid nodes' initial primitives, the step scalars and the norm and force histories.
    // restore() needs a case created from the same grid, boundaries, body and physics and initialised.
    void save(const std::string& path);
    void restore(const std::string& path);
    double t = 0, dt = 0;
    uint64_t level = 0;
    bool nonFinite = false;

private:
    void* alloc(size_t bytes);
    void setupBoundaries();
    void categorizeNodes();
    template <class T> void allocFields();
    template <class T> void hostInitialField(std::vector<T>& h) const;
    template <class T> void uploadState(const std::vector<T>& h);
    template <class T> void computeDt(int mode);
    template <class T> void applyBCs(T* S, double frac, bool outletDone, T* storedPrims = nullptr);
    template <class T> void stepOnce();
    std::vector<void*> allocations;
    StageTape* tapeNow = nullptr;
    struct EventPair { hipEvent_t a, b; int group; };
    std::vector<EventPair> events;
    bool profiling = false;
    void mark(int group, bool begin);
};

}  // namespace nsfsi
