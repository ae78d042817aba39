// The time loop (src/DVMBase.hpp:43-105) on a device-resident simulation
// (dvm_sim_t): blobs and sheet stay in HBM between steps; blobs() copies them
// out when asked.  Screen and file output of the reference are not part of
// this layer.
#ifndef DVMPP_AMD_DVMBASE_HPP
#define DVMPP_AMD_DVMBASE_HPP

#include <string>
#include <tuple>
#include <vector>

#include "VortexBlobs.hpp"
#include "dvm_hip.h"

class DVMBase {
public:
    DVMBase(const std::vector<double>& body_x, const std::vector<double>& body_z,
            const dvm_sim_params& params, unsigned steps);
    DVMBase(const std::string& body_file, const dvm_sim_params& params, unsigned steps);
    ~DVMBase();
    DVMBase(const DVMBase&) = delete;
    DVMBase& operator=(const DVMBase&) = delete;

    /// Runs all steps (src/DVMBase.cpp:67-84 without output).
    void solve();
    /// One increment_step + compute_step.
    void step();
    [[nodiscard]] unsigned current_step() const;
    [[nodiscard]] double time() const;
    [[nodiscard]] unsigned num_blobs() const;
    [[nodiscard]] VortexBlobs blobs() const;
    /// {fx, fz, cd, cl} of the last step.
    [[nodiscard]] std::tuple<double, double, double, double> loads() const;

private:
    void init(const std::vector<double>& x, const std::vector<double>& z);
    dvm_sim_t m_sim = nullptr;
    dvm_sim_params m_p;
    unsigned m_steps;
};

#endif
