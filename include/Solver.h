// Solver.h - drop-in replacement for the reference's `Solver` class (include/Solver.h:8-72).
//
// Same public surface: Solver(u, v, w, p, temp), ~Solver(), solve(), writeVTK(name), built from
// the user's compile-time case header `config.h` (include/config.h:1-82 format: dim, nx..nz,
// domain, density, dynamicViscosity, velBCs, numOuterIter, nIter_*, relax_*, tol_*).  The work
// runs on the GPU through the C-ABI of incompflow.h; link with -lincompflow.
//
// Differences from the reference, all deliberate:
//   * move-only (the reference's implicit copy double-frees its device buffers);
//   * failures throw std::runtime_error with the library's message instead of being ignored;
//   * the smoother is red-black Gauss-Seidel (deterministic; the reference's in-place sweep
//     is order-dependent) and norms are reduced in FP64;
//   * solve() may be called again to continue iterating from the current state.
#ifndef INCOMPFLOW_SOLVER_H
#define INCOMPFLOW_SOLVER_H

#include <stdexcept>
#include <string>
#include <utility>

#include "config.h"       // the case: the reference's config.h format (pulls in types.h)
#include "incompflow.h"

namespace ifl_compat {

// ifl_config from the constexprs of config.h (include/config.h:6-81).
inline ifl_config config_from_header(scalar u, scalar v, scalar w, scalar p, scalar temp) {
    ifl_config c;
    ifl_config_default(&c);
    c.dim = dim;
    c.nx = nx; c.ny = ny; c.nz = nz;
    c.xmin = xmin; c.xmax = xmax; c.ymin = ymin; c.ymax = ymax; c.zmin = zmin; c.zmax = zmax;
    c.density = density;
    c.viscosity = dynamicViscosity;
    for (int f = 0; f < 6; ++f) {
        c.bc_type[f] = velBCs::type[f];
        for (int q = 0; q < 3; ++q) c.bc_val[f][q] = velBCs::val[f][q];
    }
    c.num_outer_iter = numOuterIter;
    c.n_iter[0] = nIter_u; c.n_iter[1] = nIter_v; c.n_iter[2] = nIter_w; c.n_iter[3] = nIter_p;
    c.relax[0] = relax_u; c.relax[1] = relax_v; c.relax[2] = relax_w; c.relax[3] = relax_p;
    c.tol[0] = tol_u; c.tol[1] = tol_v; c.tol[2] = tol_w; c.tol[3] = tol_p;
    c.tol_mass = tol_mass;
    c.init[0] = u; c.init[1] = v; c.init[2] = w; c.init[3] = p; c.init[4] = temp;
    // config.h:19-20, 45-63: thermal properties and TempBCs.  The reference never solves for
    // temperature; defining IFL_SOLVE_ENERGY before this header turns the equation on.
    c.conductivity = thermalConductivity;
    c.specific_heat = specificHeatCapacity;
    for (int f = 0; f < 6; ++f) {
        c.temp_bc_type[f] = TempBCs::type[f];
        c.temp_bc_val[f] = TempBCs::val[f];
    }
#ifdef IFL_SOLVE_ENERGY
    c.energy = 1;
#endif
    c.stop_rule = IFL_STOP_FAITHFUL;   // Solver.cu:237-268 as written
    c.log_every = 1000;                // Solver.cu:270
    return c;
}

inline void check(int st) {
    if (st != IFL_OK) throw std::runtime_error(std::string("incompflow: ") + ifl_last_error());
}

}  // namespace ifl_compat

class Solver {
public:
    // Solver.h:46: uniform initial fields.
    Solver(scalar u = 0.0, scalar v = 0.0, scalar w = 0.0, scalar p = 0.0, scalar temp = 273.0) {
        ifl_config c = ifl_compat::config_from_header(u, v, w, p, temp);
        ifl_compat::check(ifl_create(&c, &h_));
    }
    // Solver.h:48
    ~Solver() { ifl_destroy(h_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;
    Solver(Solver &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Solver &operator=(Solver &&o) noexcept {
        if (this != &o) {
            ifl_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }

    // Solver.h:50: outer iterations until the stop rule holds or numOuterIter ran; prints the
    // Solver.cu:270-285 log block at iteration 1, every 1000 iterations and at the end.
    void solve() {
        ifl_report r;
        ifl_compat::check(ifl_solve(h_, 0, &r, nullptr, 0));
    }

    // Solver.h:52: legacy VTK written to output/<filepath> (Solver.cu:303).
    void writeVTK(const std::string &filepath) const {
        ifl_compat::check(ifl_write_vtk(h_, ("output/" + filepath).c_str()));
    }

    // No reference counterpart: binary state files (ifl_save_state / ifl_load_state) to continue
    // a solve or warm start a case of the same grid; paths are used as given.
    void saveState(const std::string &path) const { ifl_compat::check(ifl_save_state(h_, path.c_str())); }
    void loadState(const std::string &path) { ifl_compat::check(ifl_load_state(h_, path.c_str())); }

    // The C-ABI handle, for field access (ifl_get_field) and the other entry points.
    ifl_solver *handle() const { return h_; }

private:
    ifl_solver *h_ = nullptr;
};

#endif  // INCOMPFLOW_SOLVER_H
