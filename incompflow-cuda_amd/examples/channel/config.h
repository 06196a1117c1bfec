// Plain-method channel: 512 x 128 cells on [-0.0625, 0.0625] x [-0.0125, 0.0125], rho = 998.2,
// mu = 1.002e-3, west inlet u = 0.001, east outlet, north/south walls, alpha_u = 0.7, alpha_p = 0.3.
// Same header format as the reference's include/config.h.
#ifndef CONFIG_H
#define CONFIG_H

#include "types.h"

constexpr int dim = 2;

constexpr int nx = 512;
constexpr int ny = 128;
constexpr int nz = 1;

constexpr scalar xmin = -0.0625;
constexpr scalar xmax = 0.0625;
constexpr scalar ymin = -0.0125;
constexpr scalar ymax = 0.0125;
constexpr scalar zmin = 0.0;
constexpr scalar zmax = 1.0;

constexpr scalar thermalConductivity = 1;
constexpr scalar specificHeatCapacity = 1;
constexpr scalar density = 998.2;
constexpr scalar dynamicViscosity = 1.002e-3;
constexpr scalar thermalDiffusivity = thermalConductivity / specificHeatCapacity / density;

struct velBCs {
    static constexpr int type[6] = {2, 1, 0, 0, 0, 0};   // east outlet, west inlet, walls
    static constexpr scalar val[6][3] = {
        {0, 0, 0}, {0.001, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

struct TempBCs {
    static constexpr int type[6] = {1, 1, 0, 0, 1, 1};
    static constexpr scalar val[6] = {0.0, 0.0, 293.0, 373.0, 0.0, 0.0};
};

constexpr int numOuterIter = 3000;

constexpr int nIter_u = 10;
constexpr int nIter_v = 10;
constexpr int nIter_w = 10;
constexpr int nIter_p = 20;

constexpr scalar relax_u = 0.7;
constexpr scalar relax_v = 0.7;
constexpr scalar relax_w = 0.7;
constexpr scalar relax_p = 0.3;

constexpr scalar tol_u = 1e-5;
constexpr scalar tol_v = 1e-5;
constexpr scalar tol_w = 1e-5;
constexpr scalar tol_p = 1e-3;
constexpr scalar tol_mass = 1e-9;

#endif
