// main.cpp - example driver with the reference's flow (main.cpp:5-16): construct, solve, write
// the VTK file, report the wall time.  The case comes from the config.h on the include path.
#include <chrono>
#include <iostream>

#include "Solver.h"

int main() {
    auto start = std::chrono::high_resolution_clock::now();

    Solver solver{0, 0, 0, 0, 0};
    solver.solve();
    solver.writeVTK("test1.vtk");

    auto end = std::chrono::high_resolution_clock::now();
    auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(end - start);
    std::cout << "Elapsed time: " << ms.count() << " ms\n";
    return 0;
}
