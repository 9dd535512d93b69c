/*
 * main.cpp — command-line program of the MI355X path, as the reference's main (src/main.cpp:12-20): reads
 * ./ConfigFile, returns -1 when it cannot be read, otherwise runs the simulation with the output the
 * ConfigFile asks for (./output/out.vtk.N, times.dat, norm_*.dat, status reports on stdout).
 */
#include <exception>
#include <iostream>

#include "../../include/ns_fsi.hpp"

int main() {
    const ConfigSettings params("ConfigFile");  // Read parameters from config file.
    if (params.errorOccurred) return -1;
    try {
        Simulation sim(params);  // Initialize simulation with given parameters.
        sim.run();               // Run simulation. Produce output as specified.
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
