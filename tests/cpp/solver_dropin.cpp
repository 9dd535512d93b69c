/*
 * solver_dropin.cpp — test program: drives the C++ drop-in surface (include/ns_fsi.hpp) exactly as the
 * reference's Simulation::run does (src/Simulation.cpp:19-30): Solver::initialize, then marchTimeStep(t,
 * timeLevel) until the stopping criterion, reading solver.dt and solver.mesh between steps, then
 * getConvergenceHistory(). Writes, as raw little-endian doubles to the file named by argv[2]:
 * t, timeLevel, dt, then the 12 fields of solver.mesh (rho .. kappa) read element by element through the
 * reference's accessors in (i,j,k) order with i fastest, then nodeType as doubles, then the 5 norm histories
 * (each preceded by its length).
 * Usage: solver_dropin <ConfigFile> <out.bin>
 */
#include <cstdio>
#include <iostream>

#include "../../include/ns_fsi.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: solver_dropin <ConfigFile> <out.bin>\n";
        return 2;
    }
    const ConfigSettings params(argv[1]);
    if (params.errorOccurred) return -1;
    try {
        Solver solver(params);
        double t = 0;
        ulong timeLevel = 0;
        solver.initialize();
        double lastDt = solver.dt;
        while (params.stopCriterion == StopCriterionEnum::timesteps ? timeLevel < params.stopTimeLevel : t < params.t_end) {
            solver.marchTimeStep(t, timeLevel);
            lastDt = solver.dt;  // public member, read between steps as OutputManager does
            if (timeLevel == 1) (void)solver.mesh.conservedVariables.rho.at(1, 1, 1);  // a mesh read between steps
        }
        const Mesh& mesh = solver.mesh;
        std::FILE* f = std::fopen(argv[2], "wb");
        if (!f) return 3;
        double head[3] = {t, (double)timeLevel, lastDt};
        std::fwrite(head, sizeof(double), 3, f);
        const Array3D_d* fields[12] = {&mesh.conservedVariables.rho, &mesh.conservedVariables.rho_u,
                                       &mesh.conservedVariables.rho_v, &mesh.conservedVariables.rho_w,
                                       &mesh.conservedVariables.rho_E, &mesh.primitiveVariables.u,
                                       &mesh.primitiveVariables.v, &mesh.primitiveVariables.w,
                                       &mesh.primitiveVariables.p, &mesh.primitiveVariables.T,
                                       &mesh.transportProperties.mu, &mesh.transportProperties.kappa};
        for (const Array3D_d* a : fields)
            for (size_t k = 0; k < mesh.NK; ++k)
                for (size_t j = 0; j < mesh.NJ; ++j)
                    for (size_t i = 0; i < mesh.NI; ++i) {
                        double v = a->at(i, j, k);
                        std::fwrite(&v, sizeof(double), 1, f);
                    }
        for (size_t k = 0; k < mesh.NK; ++k)
            for (size_t j = 0; j < mesh.NJ; ++j)
                for (size_t i = 0; i < mesh.NI; ++i) {
                    double v = (double)(int)mesh.nodeType(i, j, k);
                    std::fwrite(&v, sizeof(double), 1, f);
                }
        const ConservedVariablesVectorGroup& h = solver.getConvergenceHistory();
        for (const std::vector<double>* v : {&h.rho, &h.rho_u, &h.rho_v, &h.rho_w, &h.rho_E}) {
            double n = (double)v->size();
            std::fwrite(&n, sizeof(double), 1, f);
            if (!v->empty()) std::fwrite(v->data(), sizeof(double), v->size(), f);
        }
        std::fclose(f);
        std::cout << "t = " << t << " timeLevel = " << timeLevel << " dt = " << solver.dt << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
