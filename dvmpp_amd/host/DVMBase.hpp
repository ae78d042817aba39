// The time loop (src/DVMBase.hpp:43-105) on a device-resident simulation
// (dvm_sim_t): blobs and sheet stay in HBM between steps; blobs() copies them
// out when asked.  Built from an input file (the reference's constructor,
// src/DVMBase.cpp:8-66), solve() also prints the reference's screen lines and
// writes its per-step files through an asynchronous StepWriter.
#ifndef DVMPP_AMD_DVMBASE_HPP
#define DVMPP_AMD_DVMBASE_HPP

#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "VortexBlobs.hpp"
#include "XmlHandler.hpp"
#include "dvm_hip.h"

namespace dvm {
class StepWriter;
}

/// Choices an input file does not hold: the arithmetic of the pair sums, the
/// random stream and how often state is written.
struct RunOptions {
    int precision = DVM_F64;         ///< dvm_precision of the pair sums
    int rng_mode = DVM_RNG_REFERENCE;///< dvm_rng_mode
    unsigned output_every = 1;       ///< write every n-th step (the reference: 1)
    bool write_blobs = true;         ///< false: no per-blob rows in <stamp>_vortex.dat
    bool screen = true;              ///< per-step screen lines
};

class DVMBase {
public:
    DVMBase(const std::vector<double>& body_x, const std::vector<double>& body_z,
            const dvm_sim_params& params, unsigned steps);
    DVMBase(const std::string& body_file, const dvm_sim_params& params, unsigned steps);
    /// Reads the case from the input file and creates the output files
    /// <output_dir>/<timestamp>_{vortex,vortex_num,gamma,loads,probe}.dat.
    DVMBase(XmlHandler& xml, const std::string& timestamp);
    DVMBase(XmlHandler& xml, const std::string& timestamp, const RunOptions& opts);
    ~DVMBase();
    DVMBase(const DVMBase&) = delete;
    DVMBase& operator=(const DVMBase&) = delete;

    /// Runs the remaining steps up to the total (src/DVMBase.cpp:67-84); with
    /// an input file, also the per-step files and screen lines.
    void solve();
    /// The time loop up to step `last` (at most the total).
    void solve_until(unsigned last);
    /// One increment_step + compute_step.
    void step();
    [[nodiscard]] unsigned current_step() const;
    [[nodiscard]] double time() const;
    [[nodiscard]] unsigned num_blobs() const;
    [[nodiscard]] VortexBlobs blobs() const;
    /// {fx, fz, cd, cl} of the last step.
    [[nodiscard]] std::tuple<double, double, double, double> loads() const;
    [[nodiscard]] unsigned steps() const { return m_steps; }

    /// Writes the whole simulation state (dvm_sim_snapshot) to `file`.
    void save_snapshot(const std::string& file) const;
    /// Continues from a state saved by save_snapshot for the same case.
    void load_snapshot(const std::string& file);

private:
    void init(const std::vector<double>& x, const std::vector<double>& z);
    void from_xml(XmlHandler& xml, const std::string& timestamp, const RunOptions& opts);
    dvm_sim_t m_sim = nullptr;
    dvm_sim_params m_p;
    unsigned m_steps;
    RunOptions m_opts;
    std::unique_ptr<dvm::StepWriter> m_writer;
};

#endif
