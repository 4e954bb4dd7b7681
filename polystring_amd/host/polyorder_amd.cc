// polyorder_amd.cc -- command-line driver: param.ini -> Model_AB -> scft::run()
// -> done.txt, the flow of the reference's scft/AB_ps_pbc.cc and
// scft/polyorder.cc (AB model only).
#include <chrono>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>

#include "Model_AB.h"
#include "scft.h"

int main(int argc, char *argv[]) {
    const std::string config_file = argc > 1 ? argv[1] : "param.ini";
    const auto t0 = std::chrono::steady_clock::now();
    Config cfg(config_file.c_str());
    Model_AB model(cfg);
    scft sim(config_file, &model);
    sim.run();
    const double cost = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::time_t raw = std::time(nullptr);
    std::ofstream f("done.txt");
    f << "Program terminates at: " << std::asctime(std::localtime(&raw));
    f << "Total time elapsed in second:\t" << cost << std::endl;
    f << "Iterations:\t" << sim.iterations() << std::endl;
    f.precision(16);
    f << "H:\t" << sim.last_H() << std::endl;
    f << "Residual:\t" << sim.last_residual() << std::endl;
    std::cout.precision(16);
    std::cout << "done: iterations " << sim.iterations() << " H " << sim.last_H() << " residual "
              << sim.last_residual() << std::endl;
    return 0;
}
