/*
 * ns_fsi.hpp — the reference's C++ host surface (Frederik-Kris/3D-NS-FSI) over the C-ABI of ns_fsi.h:
 * ConfigSettings, Mesh, Solver, OutputManager and Simulation with the reference's names, members and call
 * sequence, so code written against them (src/main.cpp, src/Simulation.cpp, src/OutputManager.cpp) runs on the
 * MI355X path. Header-only; link with libnsfsi.so.
 *
 * What each class replaces (paths relative to the reference root):
 *   ConfigSettings  src/ConfigSettings.h:33-80 (loadSettings through ns_config_parse; errorOccurred flag)
 *   Mesh            src/Mesh.h:79-89, read side: NI/NJ/NK, nNodesTotal, dx/dy/dz, positionOffset, nodeType,
 *                   conservedVariables, primitiveVariables, transportProperties. The arrays mirror the device
 *                   state: they are copied from the device on first access after each step.
 *   Solver          src/Solver.h:17-31: initialize(), marchTimeStep(t, timeLevel), getConvergenceHistory(),
 *                   public params, mesh and dt.
 *   OutputManager   src/OutputManager.h:19-60: output folder, legacy ASCII VTK STRUCTURED_POINTS snapshots
 *                   (Node_Flag, scalars, vectors, x fastest), times.dat, norm_*.dat, save schedule and the
 *                   wall-clock status report, in the reference's text formats.
 *   Simulation      src/Simulation.h:17-32: the run loop and stopping criterion.
 */
#ifndef NS_FSI_HPP_
#define NS_FSI_HPP_

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ns_fsi.h"

typedef unsigned long ulong;
typedef unsigned int uint;

enum class StopCriterionEnum { timesteps, end_time };
enum class ConvHistoryEnum { none, density, all };
// Node classes in the reference's order (src/Node.h:27-34); the values equal ns_node_type.
enum class NodeTypeEnum { FluidInterior, FluidEdge, FluidGhost, SolidInactive, SolidGhost };

struct Vector3_d {
    double x, y, z;
    Vector3_d(double x, double y, double z) : x{x}, y{y}, z{z} {}
    Vector3_d() = default;
};

// ConfigSettings::ConfigSettings / loadSettings (src/ConfigSettings.cpp:11-28): the ConfigFile is parsed by
// ns_config_parse (libconfig syntax, strict typing); on any error errorOccurred is set and the message goes to
// stderr. The extension keys of ns_config (boundary types, body, initial condition, precision) are read from
// the same file.
class ConfigSettings {
public:
    explicit ConfigSettings(std::string filename) { loadSettings(filename); }
    explicit ConfigSettings(const ns_config& c) : errorOccurred(false) { fromConfig(c); }

    void loadSettings(std::string filename) {
        errorOccurred = false;
        std::ifstream in(filename);
        if (!in) {
            std::cerr << "I/O error while reading file: " << filename << std::endl;
            errorOccurred = true;
            return;
        }
        std::stringstream ss;
        ss << in.rdbuf();
        ns_config c;
        char err[512] = {0};
        if (ns_config_parse(ss.str().c_str(), &c, err, sizeof(err)) != NS_OK) {
            std::cerr << err << std::endl;
            errorOccurred = true;
            return;
        }
        fromConfig(c);
    }

    size_t NI = 0, NJ = 0, NK = 0;
    StopCriterionEnum stopCriterion = StopCriterionEnum::timesteps;
    ulong stopTimeLevel = 0;
    double t_end = 0;
    double convStabilityLimit = 0, viscStabilityLimit = 0;
    std::vector<uint> filterIntervals;
    std::vector<double> filterIntervalChangeTimes;
    double statusReportInterval = 0;
    bool saveDensity = false, saveMomentum = false, saveEnergy = false;
    bool saveVelocity = false, savePressure = false, saveTemperature = false;
    bool saveViscosity = false, saveThermalCond = false;
    bool saveIC = false, saveFinal = false, saveIntervals = false;
    double savePeriod = 0, saveIntervalsStartTime = 0, saveIntervalsEndTime = 0;
    ConvHistoryEnum saveConvergenceHistory = ConvHistoryEnum::none;
    double Gamma = 0, Pr = 0, R = 0, Re = 0, sutherlands_C2 = 0, M_0 = 0, T_0 = 0, L_x = 0, L_y = 0, L_z = 0;
    bool errorOccurred = true;
    double machinePrecisionBuffer = 0;
    ns_config config{};  // every key, extension keys included, as the C-ABI takes it

private:
    void fromConfig(const ns_config& c) {
        config = c;
        NI = (size_t)c.NI; NJ = (size_t)c.NJ; NK = (size_t)c.NK;
        stopCriterion = c.stop_criterion == 0 ? StopCriterionEnum::timesteps : StopCriterionEnum::end_time;
        stopTimeLevel = (ulong)c.stop_time_level;
        t_end = c.t_end;
        convStabilityLimit = c.conv_stability_limit;
        viscStabilityLimit = c.visc_stability_limit;
        filterIntervals.assign(c.filter_intervals, c.filter_intervals + c.n_filter_intervals);
        filterIntervalChangeTimes.assign(c.filter_change_times, c.filter_change_times + c.n_filter_change_times);
        statusReportInterval = c.status_report_interval;
        saveDensity = c.save_density; saveMomentum = c.save_momentum; saveEnergy = c.save_energy;
        saveVelocity = c.save_velocity; savePressure = c.save_pressure; saveTemperature = c.save_temperature;
        saveViscosity = c.save_viscosity; saveThermalCond = c.save_thermal_cond;
        saveIC = c.save_ic; saveFinal = c.save_final; saveIntervals = c.save_intervals;
        savePeriod = c.save_period;
        saveIntervalsStartTime = c.save_intervals_start_time;
        saveIntervalsEndTime = c.save_intervals_end_time;
        saveConvergenceHistory = c.conv_history == 0 ? ConvHistoryEnum::none
                                 : (c.conv_history == 1 ? ConvHistoryEnum::density : ConvHistoryEnum::all);
        Gamma = c.gamma; Pr = c.Pr; R = c.R; Re = c.Re; sutherlands_C2 = c.sutherlands_C2; M_0 = c.M_0;
        T_0 = c.T_0; L_x = c.L_x; L_y = c.L_y; L_z = c.L_z;
        machinePrecisionBuffer = c.machine_precision_buffer;
    }
};

namespace nsfsi_host {

// Copy of one device array, refreshed when the state version it was read at is stale.
template <class V>
class Mirror {
public:
    Mirror() = default;
    void bind(size_t NI, size_t NJ, size_t NK, const uint64_t* version, std::function<void(V*)> load) {
        NI_ = NI; NJ_ = NJ; NK_ = NK;
        version_ = version;
        load_ = std::move(load);
        loadedAt_ = ~uint64_t(0);
    }
    // element (i, j, k); stored x fastest as the C-ABI delivers it
    const V& get(size_t i, size_t j, size_t k) const { ensure(); return data_[(k * NJ_ + j) * NI_ + i]; }
    // 1-D index in the reference's order (i*NJ + j)*NK + k (src/Array3D.h:42)
    const V& get1D(size_t n) const {
        const size_t i = n / (NJ_ * NK_), j = n % (NJ_ * NK_) / NK_, k = n % NK_;
        return get(i, j, k);
    }
    size_t size() const { return NI_ * NJ_ * NK_; }
    const std::vector<V>& raw() const { ensure(); return data_; }

private:
    void ensure() const {
        if (!version_ || loadedAt_ == *version_) return;
        data_.resize(size());
        load_(data_.data());
        loadedAt_ = *version_;
    }
    size_t NI_ = 0, NJ_ = 0, NK_ = 0;
    const uint64_t* version_ = nullptr;
    std::function<void(V*)> load_;
    mutable std::vector<V> data_;
    mutable uint64_t loadedAt_ = ~uint64_t(0);
};

inline void check(ns_handle* h, ns_status s, const char* what) {
    if (s != NS_OK) throw std::runtime_error(std::string(what) + ": " + ns_last_error(h));
}

}  // namespace nsfsi_host

// Array3D_d, read side (src/Array3D.h:20-99): the values of one field of the device state.
class Array3D_d {
public:
    double operator()(size_t i, size_t j, size_t k) const { return m.get(i, j, k); }
    double operator()(size_t index1D) const { return m.get1D(index1D); }
    double at(size_t i, size_t j, size_t k) const { return m.get(i, j, k); }
    double at(size_t index1D) const { return m.get1D(index1D); }
    size_t size() const { return m.size(); }
    bool allFinite() const {
        for (double d : m.raw())
            if (!std::isfinite(d)) return false;
        return true;
    }
    nsfsi_host::Mirror<double> m;
};

class Array3D_nodeType {
public:
    NodeTypeEnum operator()(size_t i, size_t j, size_t k) const { return (NodeTypeEnum)m.get(i, j, k); }
    NodeTypeEnum operator()(size_t index1D) const { return (NodeTypeEnum)m.get1D(index1D); }
    nsfsi_host::Mirror<uint8_t> m;
};

struct ConservedVariablesArrayGroup { Array3D_d rho, rho_u, rho_v, rho_w, rho_E; };
struct PrimitiveVariablesArrayGroup { Array3D_d u, v, w, p, T; };
struct TransportPropertiesArrayGroup { Array3D_d mu, kappa; };
struct ConservedVariablesVectorGroup { std::vector<double> rho, rho_u, rho_v, rho_w, rho_E; };

// Mesh, read side (src/Mesh.h:79-89).
class Mesh {
public:
    explicit Mesh(const ConfigSettings& params)
        : NI(params.NI), NJ(params.NJ), NK(params.NK), nNodesTotal(params.NI * params.NJ * params.NK) {}
    const size_t NI, NJ, NK;
    const size_t nNodesTotal;
    double dx = 0, dy = 0, dz = 0;
    ConservedVariablesArrayGroup conservedVariables;
    PrimitiveVariablesArrayGroup primitiveVariables;
    TransportPropertiesArrayGroup transportProperties;
    Array3D_nodeType nodeType;
    Vector3_d positionOffset{0, 0, 0};

    // Binds every array to the device state of handle h (called by Solver::initialize).
    void bindDevice(ns_handle* h, const uint64_t* version) {
        double sp[3], off[3];
        nsfsi_host::check(h, ns_get_geometry(h, sp, off), "ns_get_geometry");
        dx = sp[0]; dy = sp[1]; dz = sp[2];
        positionOffset = Vector3_d(off[0], off[1], off[2]);
        Array3D_d* f[12] = {&conservedVariables.rho, &conservedVariables.rho_u, &conservedVariables.rho_v,
                            &conservedVariables.rho_w, &conservedVariables.rho_E, &primitiveVariables.u,
                            &primitiveVariables.v, &primitiveVariables.w, &primitiveVariables.p, &primitiveVariables.T,
                            &transportProperties.mu, &transportProperties.kappa};
        for (int q = 0; q < 12; ++q)
            f[q]->m.bind(NI, NJ, NK, version, [h, q](double* out) {
                nsfsi_host::check(h, ns_get_field(h, q, out), "ns_get_field");
            });
        static const uint64_t never = 0;  // node classes do not change after initialisation
        nodeType.m.bind(NI, NJ, NK, &never, [h](uint8_t* out) {
            nsfsi_host::check(h, ns_get_node_types(h, out), "ns_get_node_types");
        });
    }
};

// Solver (src/Solver.h:17-31) on one MI355X through the C-ABI.
class Solver {
public:
    explicit Solver(const ConfigSettings& params) : params(params), mesh(params), dt(0) {
        ns_status s = ns_create(&params.config, &h);
        if (s != NS_OK) {
            std::string msg = std::string("ns_create: ") + ns_last_error(h);
            ns_destroy(h);
            throw std::runtime_error(msg);
        }
    }
    ~Solver() { ns_destroy(h); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // Solver::initialize (src/Solver.cpp:22-28): boundaries, classification, initial condition, first dt.
    void initialize() {
        nsfsi_host::check(h, ns_init(h), "ns_init");
        (void)ns_set_option(h, "graph", 1);  // each step replays one captured HIP graph
        double t;
        uint64_t lv;
        nsfsi_host::check(h, ns_get_time(h, &t, &lv, &dt), "ns_get_time");
        ++version;
        mesh.bindDevice(h, &version);
    }

    // Solver::marchTimeStep (src/Solver.cpp:130-167): one RK4 step; t and timeLevel advance, dt becomes the
    // size of the next step. A non-finite state is reported on stdout and the run carries on, as in the
    // reference (src/Mesh.h:48-61).
    void marchTimeStep(double& t, ulong& timeLevel) {
        uint64_t lv = 0;
        ns_status s = ns_step(h, 1, &t, &lv, &dt);
        ++version;
        if (s == NS_ERR_NONFINITE) {
            if (!nonFiniteReported) std::cout << "Non-finite value found in the conserved variables.\n";
            nonFiniteReported = true;
            nsfsi_host::check(h, ns_get_time(h, &t, &lv, &dt), "ns_get_time");
        } else {
            nsfsi_host::check(h, s, "ns_step");
        }
        timeLevel = (ulong)lv;
    }

    const ConservedVariablesVectorGroup& getConvergenceHistory() const {
        if (historyAt != version) {
            std::vector<double>* v[5] = {&history.rho, &history.rho_u, &history.rho_v, &history.rho_w, &history.rho_E};
            for (int q = 0; q < 5; ++q) {
                int64_t n = 0;
                nsfsi_host::check(h, ns_norm_history(h, q, nullptr, 0, &n), "ns_norm_history");
                v[q]->assign((size_t)n, 0.0);
                if (n) nsfsi_host::check(h, ns_norm_history(h, q, v[q]->data(), n, &n), "ns_norm_history");
            }
            historyAt = version;
        }
        return history;
    }

    // Checkpoint and restart (ns_save / ns_restore; no reference counterpart).
    void saveCheckpoint(const std::string& path) { nsfsi_host::check(h, ns_save(h, path.c_str()), "ns_save"); }
    void restoreCheckpoint(const std::string& path, double& t, ulong& timeLevel) {
        nsfsi_host::check(h, ns_restore(h, path.c_str()), "ns_restore");
        uint64_t lv = 0;
        nsfsi_host::check(h, ns_get_time(h, &t, &lv, &dt), "ns_get_time");
        timeLevel = (ulong)lv;
        ++version;
    }

    ns_handle* handle() const { return h; }

    const ConfigSettings params;
    Mesh mesh;
    double dt;

private:
    ns_handle* h = nullptr;
    uint64_t version = 0;  // bumped by every call that changes the device state
    bool nonFiniteReported = false;
    mutable ConservedVariablesVectorGroup history;
    mutable uint64_t historyAt = ~uint64_t(0);
};

// Wall-clock timer (the reference uses sf::Clock).
class Clock {
public:
    Clock() : start(std::chrono::steady_clock::now()) {}
    struct Time {
        double s;
        double asSeconds() const { return s; }
    };
    Time getElapsedTime() const {
        return {std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count()};
    }
    Time restart() {
        Time t = getElapsedTime();
        start = std::chrono::steady_clock::now();
        return t;
    }

private:
    std::chrono::steady_clock::time_point start;
};

// OutputManager (src/OutputManager.cpp:12-362): same files, same text, same schedule.
class OutputManager {
public:
    explicit OutputManager(const ConfigSettings& params) : params(params), savedSolutions{0} {}

    // Creates ./output, or empties it except for ParaView state files (src/OutputManager.cpp:19-45).
    void initialize() {
        namespace fs = std::filesystem;
        fs::path outputPath("./output");
        if (!(fs::exists(outputPath) && fs::is_directory(outputPath))) {
            if (!fs::create_directory(outputPath)) {
                std::cerr << "Output folder could not be created.\n";
                return;
            }
        }
        std::vector<fs::path> doomed;
        for (const fs::directory_entry& entry : fs::recursive_directory_iterator(outputPath))
            if (entry.path().generic_string().find(".pvsm") == std::string::npos) doomed.push_back(entry.path());
        for (const fs::path& p : doomed) {
            try {
                fs::remove_all(p);
            } catch (std::exception& ex) {
                std::cerr << "Could not remove " << p << std::endl << ex.what() << std::endl;
            }
        }
    }

    void processInitialOutput(const Mesh& mesh, double t) {
        if (params.saveIC) storeCurrentSolution(mesh, t);
    }

    // src/OutputManager.cpp:55-80: save when the next step would pass a multiple of savePeriod inside the
    // save window; status report every statusReportInterval seconds of wall-clock time.
    void processIntermediateOutput(const Mesh& mesh, double t, ulong timeLevel, double dt) {
        if (params.saveIntervals) {
            bool withinRelevantTimeRegime =
                t >= params.saveIntervalsStartTime && (t <= params.saveIntervalsEndTime || params.saveIntervalsEndTime <= 0);
            bool enoughTimeSinceSave = true;
            if (params.savePeriod > 0) {
                double timeSinceSave = std::fmod(t, params.savePeriod);
                enoughTimeSinceSave = timeSinceSave + dt >= params.savePeriod;
            }
            if (withinRelevantTimeRegime && enoughTimeSinceSave) storeCurrentSolution(mesh, t);
        }
        double timeSinceStatusReport = statusReportTimer.getElapsedTime().asSeconds();
        if (timeSinceStatusReport >= params.statusReportInterval) {
            writeStatusReport_toScreen(t, timeLevel, dt);
            statusReportTimer.restart();
        }
    }

    void processFinalOutput(const Mesh& mesh, double t, ulong timeLevel, double dt,
                            const ConservedVariablesVectorGroup& convergenceHistory) {
        if (params.saveFinal) storeCurrentSolution(mesh, t);
        writeStatusReport_toScreen(t, timeLevel, dt);
        writeOutputTimes();
        writeConvergenceHistoryFiles(convergenceHistory);
    }

    uint savedCount() const { return savedSolutions; }

private:
    void storeCurrentSolution(const Mesh& mesh, double t) {
        storeCurrentSolution_vtk(mesh, t);
        ++savedSolutions;
        outputTimes.push_back(t);
    }

    std::string getVtkHeader(const Mesh& mesh, double t) {
        using std::to_string;
        std::string h;
        h += "# vtk DataFile Version 5.1\n";
        h += "Output at time t = " + to_string(t) + "\n";
        h += "ASCII\n";
        h += "DATASET STRUCTURED_POINTS\n";
        h += "DIMENSIONS " + to_string(mesh.NI) + " " + to_string(mesh.NJ) + " " + to_string(mesh.NK) + "\n";
        h += "ORIGIN " + to_string(-mesh.positionOffset.x * mesh.dx) + " " + to_string(-mesh.positionOffset.y * mesh.dy) +
             " " + to_string(-mesh.positionOffset.z * mesh.dz) + "\n";
        h += "SPACING " + to_string(mesh.dx) + " " + to_string(mesh.dy) + " " + to_string(mesh.dz) + "\n";
        h += "POINT_DATA " + to_string(mesh.nNodesTotal) + "\n";
        return h;
    }

    // Node_Flag: 0 fluid interior, 1 fluid edge, 2 fluid ghost, 3 solid ghost, 4 solid inactive
    // (src/OutputManager.cpp:114-140; the two solid classes swap places against the enum order).
    void writeVtkNodeFlags(const Mesh& mesh, std::ofstream& out) {
        out << "SCALARS Node_Flag int 1\n";
        out << "LOOKUP_TABLE default\n";
        for (size_t k = 0; k < mesh.NK; ++k)
            for (size_t j = 0; j < mesh.NJ; ++j)
                for (size_t i = 0; i < mesh.NI; ++i) {
                    int flag = 0;
                    switch (mesh.nodeType(i, j, k)) {
                        case NodeTypeEnum::FluidInterior: flag = 0; break;
                        case NodeTypeEnum::FluidEdge: flag = 1; break;
                        case NodeTypeEnum::FluidGhost: flag = 2; break;
                        case NodeTypeEnum::SolidGhost: flag = 3; break;
                        case NodeTypeEnum::SolidInactive: flag = 4; break;
                        default: throw std::logic_error("Unexpected enum value.");
                    }
                    out << flag << "\n";
                }
    }

    void writeVtkScalarData(const Mesh& mesh, std::ofstream& out) {
        std::vector<const Array3D_d*> vars;
        std::vector<std::string> names;
        if (params.saveDensity) { vars.push_back(&mesh.conservedVariables.rho); names.push_back("Density"); }
        if (params.saveEnergy) {
            vars.push_back(&mesh.conservedVariables.rho_E);
            names.push_back("Total_Specific_Energy_per_Volume");
        }
        if (params.savePressure) { vars.push_back(&mesh.primitiveVariables.p); names.push_back("Pressure"); }
        if (params.saveTemperature) { vars.push_back(&mesh.primitiveVariables.T); names.push_back("Temperature"); }
        if (params.saveViscosity) { vars.push_back(&mesh.transportProperties.mu); names.push_back("Dynamic_Viscosity"); }
        if (params.saveThermalCond) {
            vars.push_back(&mesh.transportProperties.kappa);
            names.push_back("Thermal_Conductivity");
        }
        for (size_t q = 0; q < vars.size(); ++q) {
            out << "SCALARS " << names[q] << " double 1\n";
            out << "LOOKUP_TABLE default\n";
            for (size_t k = 0; k < mesh.NK; ++k)
                for (size_t j = 0; j < mesh.NJ; ++j)
                    for (size_t i = 0; i < mesh.NI; ++i) out << vars[q]->at(i, j, k) << "\n";
        }
    }

    void writeVtkVectorData(const Mesh& mesh, std::ofstream& out) {
        std::vector<std::array<const Array3D_d*, 3>> vars;
        std::vector<std::string> names;
        if (params.saveMomentum) {
            vars.push_back({&mesh.conservedVariables.rho_u, &mesh.conservedVariables.rho_v, &mesh.conservedVariables.rho_w});
            names.push_back("Momentum_Density");
        }
        if (params.saveVelocity) {
            vars.push_back({&mesh.primitiveVariables.u, &mesh.primitiveVariables.v, &mesh.primitiveVariables.w});
            names.push_back("Velocity");
        }
        for (size_t q = 0; q < vars.size(); ++q) {
            out << "VECTORS " << names[q] << " double\n";
            for (size_t k = 0; k < mesh.NK; ++k)
                for (size_t j = 0; j < mesh.NJ; ++j)
                    for (size_t i = 0; i < mesh.NI; ++i)
                        out << vars[q][0]->at(i, j, k) << " " << vars[q][1]->at(i, j, k) << " " << vars[q][2]->at(i, j, k)
                            << "\n";
        }
    }

    void storeCurrentSolution_vtk(const Mesh& mesh, double t) {
        std::ofstream out;
        std::string filename = "output/out.vtk." + std::to_string(savedSolutions);
        out.open(filename);
        if (!out) {
            std::cout << "Could not open file:  " + filename << std::endl << "Solution was not saved." << std::endl;
            return;
        }
        out << getVtkHeader(mesh, t);
        writeVtkNodeFlags(mesh, out);
        writeVtkScalarData(mesh, out);
        writeVtkVectorData(mesh, out);
        out.close();
    }

    // src/OutputManager.cpp:284-307
    void writeStatusReport_toScreen(double t, ulong timeLevel, double dt) {
        using std::cout;
        using std::endl;
        using std::setprecision;
        cout << "Simulated time: t = " << t;
        if (params.stopCriterion == StopCriterionEnum::end_time) {
            double progressPercentage = t / params.t_end * 100.;
            cout << " , t_end = " << params.t_end << " ( " << setprecision(3) << progressPercentage << setprecision(6)
                 << " % )" << endl;
        } else {
            cout << endl;
        }
        cout << "Time level: n = " << timeLevel;
        if (params.stopCriterion == StopCriterionEnum::timesteps) {
            double progressPercentage = (double)timeLevel / (double)params.stopTimeLevel * 100.;
            cout << " , n_max = " << params.stopTimeLevel << " ( " << setprecision(3) << progressPercentage
                 << setprecision(6) << " % )" << endl;
        } else {
            cout << endl;
        }
        cout << "Timestep size: dt = " << dt << endl;
        cout << "Wall clock time: " << setprecision(3) << wallClockTimer.getElapsedTime().asSeconds() << setprecision(6)
             << " sec" << endl
             << endl;
    }

    void writeOutputTimes() {
        std::ofstream timeFile("output/times.dat");
        if (!timeFile) {
            std::cout << "Could not open file:  output/times.dat" << std::endl;
            return;
        }
        for (double time : outputTimes) timeFile << time << std::endl;
    }

    void writeConvergenceHistoryFiles(const ConservedVariablesVectorGroup& h) {
        std::vector<const std::vector<double>*> hist;
        std::vector<std::string> names;
        if (params.saveConvergenceHistory != ConvHistoryEnum::none) { hist.push_back(&h.rho); names.push_back("rho"); }
        if (params.saveConvergenceHistory == ConvHistoryEnum::all) {
            hist.push_back(&h.rho_u); names.push_back("rho_u");
            hist.push_back(&h.rho_v); names.push_back("rho_v");
            hist.push_back(&h.rho_w); names.push_back("rho_w");
            hist.push_back(&h.rho_E); names.push_back("rho_E");
        }
        for (size_t q = 0; q < names.size(); ++q) {
            std::ofstream normFile("output/norm_" + names[q] + ".dat");
            if (!normFile) {
                std::cout << "Could not open convergence history file for " + names[q] + ". " << std::endl;
                continue;
            }
            for (double norm : *hist[q]) normFile << norm << std::endl;
        }
    }

    const ConfigSettings params;
    uint savedSolutions;
    std::vector<double> outputTimes;
    Clock wallClockTimer;
    Clock statusReportTimer;
};

// Simulation (src/Simulation.cpp:11-52).
class Simulation {
public:
    explicit Simulation(const ConfigSettings& params) : params(params), solver(params), output(params), t{0}, timeLevel{0} {}

    void run() {
        solver.initialize();
        output.initialize();
        output.processInitialOutput(solver.mesh, t);
        while (!checkStoppingCriterion()) {
            solver.marchTimeStep(t, timeLevel);
            output.processIntermediateOutput(solver.mesh, t, timeLevel, solver.dt);
        }
        output.processFinalOutput(solver.mesh, t, timeLevel, solver.dt, solver.getConvergenceHistory());
    }

private:
    bool checkStoppingCriterion() {
        if (params.stopCriterion == StopCriterionEnum::timesteps) return !(timeLevel < params.stopTimeLevel);
        return !(t < params.t_end);
    }

    const ConfigSettings params;
    Solver solver;
    OutputManager output;
    double t;
    ulong timeLevel;
};

#endif  // NS_FSI_HPP_
