/*
 * batch_host.h — host object behind an ns_batch handle (see batch.cpp).
 */
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "batch.h"
#include "case.h"

namespace nsfsi {

class Batch {
public:
    Batch(const ns_config* cfgs, int n);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    void init();
    void step(int nsteps);
    // Times the batched stage for every (rows, z-planes) tile shape, keeps the fastest; ms per launch.
    double autotune(int reps);
    void setOption(const std::string& name, int64_t value);

    std::vector<std::unique_ptr<Case>> cases;
    FusedTile tile{};
    std::vector<double> tuneTimes;
    bool useGraph = true;
    bool kahan = true;
    bool persistent = false;  // stage launches in the persistent form (tiles pulled through devCtr)
    int totalTiles = 0;

private:
    void* alloc(size_t bytes);
    void buildStageTables();
    void stepOnce();
    void applyBCs(int role, double frac);
    void advanceParity();
    void dropGraphs();

    hipStream_t stream = nullptr;
    bool initialized = false;
    int parity = 0;
    std::vector<CaseView<float>> hostViews;
    CaseView<float>* devViews = nullptr;
    void* devArgs[2][4] = {};
    int* devStageStart = nullptr;
    unsigned* devCtr = nullptr;  // tile counters of the persistent stage launches (zero between launches)
    int64_t* devFaceStart[6] = {};
    int64_t faceTotal[6] = {};
    std::vector<EdgeFaceListB> faceGroups;  // faces launched together, in application order
    int64_t* devFilterStart = nullptr;
    int64_t* devGhostStart = nullptr;
    int64_t filterTotal = 0, ghostTotal = 0;
    int* devBBStart = nullptr;
    int* devIBStart = nullptr;
    int totalBB = 0, totalIB = 0;
    int nFiltCases = 0;  // cases with a periodic density filter
    int* devFiltIds = nullptr;
    int64_t* devFiltStart = nullptr;
    int64_t filtNodeTotal = 0;
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    std::vector<void*> allocations;
};

}  // namespace nsfsi
