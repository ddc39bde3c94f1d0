/*
 * gpgmp_hip_solver.hpp -- C++ adapter with the reference solver's member names over
 * the C ABI of gpgmp_hip.h, so a host written against gpgmp::Solver
 * (exec/src/gpgmp/src/Solver.h:15-239) swaps its OpenCL solver for this one:
 *
 *   gpgmp_hip::HipSolver solver(desc, &runState);      // Solver(model, runState, ...)
 *   solver.doReactions(deltat, seed);                  // Solver.h:52, int &seed
 *   solver.diffuse(i, mindt, seed);                    // Solver.h:80
 *   std::vector<float> dt = solver.getNextDiffusionTime();  // Solver.h:68
 *   solver.synchronizeStateBufferFromDevice();         // Solver.h:85
 *
 * Errors become gpgmp_hip::Error exceptions, as the reference throws gpgmp::Error /
 * cl::Error (DiffusionModel.h:68-105).  Header-only; link libgpgmp_hip.so.
 */
#ifndef GPGMP_HIP_SOLVER_HPP
#define GPGMP_HIP_SOLVER_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "gpgmp_hip.h"

namespace gpgmp_hip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != GPGMP_OK) throw Error(rc, gpgmp_last_error());
}

typedef float Real;

class HipSolver {
  public:
    // Solver(DiffusionModel&, std::vector<int>* runState, std::vector<Real>* detState, ...)
    // (Solver.h:33-36): the host state vectors stay caller-owned; synchronize* copy the one
    // of the solver type (int32 counts for the stochastic types, fp32 concentrations for
    // the deterministic ones) to/from the device.  flags: GPGMP_SOLVER_*.
    HipSolver(const gpgmp_model_desc &desc, std::vector<int> *runState, std::vector<Real> *detState = nullptr,
              void *stream = nullptr, uint32_t flags = 0u)
        : m_runState(runState), m_detState(detState), m_numSpecies(desc.num_species),
          m_deterministic(desc.solver_type >= GPGMP_DETERMINISTIC_HOMOGENEOUS_RK4) {
        check(gpgmp_model_create(&desc, &m_model));
        int rc = gpgmp_solver_create(m_model, stream, flags, &m_solver);
        if (rc != GPGMP_OK) {
            gpgmp_model_destroy(m_model);
            check(rc);
        }
        if (!m_deterministic && m_runState && !m_runState->empty()) synchronizeStateBufferToDevice();
        if (m_deterministic && m_detState && !m_detState->empty()) synchronizeStateBufferToDevice();
    }
    ~HipSolver() {
        gpgmp_solver_destroy(m_solver);
        gpgmp_model_destroy(m_model);
    }
    HipSolver(const HipSolver &) = delete;
    HipSolver &operator=(const HipSolver &) = delete;

    // Solver::doReactions(Real deltat, int &seed) (Solver.h:52)
    void doReactions(Real deltat, int &seed) {
        int32_t s = seed;
        check(gpgmp_solver_do_reactions(m_solver, deltat, &s));
        seed = s;
    }
    // Solver::diffuse(int speciesIndex, Real currentTime, int &seed) (Solver.h:80)
    void diffuse(int speciesIndex, Real currentTime, int &seed) {
        int32_t s = seed;
        check(gpgmp_solver_diffuse(m_solver, speciesIndex, currentTime, &s));
        seed = s;
    }
    // Solver::getNextDiffusionTime() (Solver.h:68)
    std::vector<Real> getNextDiffusionTime() {
        std::vector<Real> out((size_t)m_numSpecies);
        check(gpgmp_solver_next_diffusion_time(m_solver, out.data()));
        return out;
    }
    // Solver::synchronizeStateBufferFromDevice / ToDevice (Solver.h:85-90)
    void synchronizeStateBufferFromDevice() {
        if (m_deterministic) check(gpgmp_solver_get_state_f32(m_solver, m_detState->data()));
        else check(gpgmp_solver_get_state(m_solver, m_runState->data()));
    }
    void synchronizeStateBufferToDevice() {
        if (m_deterministic) check(gpgmp_solver_set_state_f32(m_solver, m_detState->data()));
        else check(gpgmp_solver_set_state(m_solver, m_runState->data()));
    }
    // Solver::getIndividualSolver() (Solver.h:109): the particle handle of an individual
    // species attached with gpgmp_solver_attach_individuals (one per species here), NULL if none
    gpgmp_individuals *getIndividualSolver(int speciesIndex) {
        gpgmp_individuals *g = nullptr;
        check(gpgmp_solver_individuals(m_solver, speciesIndex, &g));
        return g;
    }
    // DiffusionModel::oclSolve's loop on the device (DiffusionModel.cpp:1467-1603)
    Real run(int steps, unsigned seed0) {
        float t = 0.f;
        check(gpgmp_solver_run(m_solver, steps, seed0, &t));
        return t;
    }
    gpgmp_solver *handle() { return m_solver; }

  private:
    std::vector<int> *m_runState;
    std::vector<Real> *m_detState;
    int m_numSpecies;
    bool m_deterministic;
    gpgmp_model *m_model = nullptr;
    gpgmp_solver *m_solver = nullptr;
};

}  // namespace gpgmp_hip

#endif
