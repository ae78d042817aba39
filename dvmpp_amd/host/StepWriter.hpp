// Per-step text output of the reference's time loop (DVMBase::write_outputs,
// src/DVMBase.cpp:151-158) for a device-resident simulation, written without
// stalling the stepping thread: each output step's state is copied into a
// pinned host frame on the simulation's stream (dvm_sim_copy_out + an event),
// and a writer thread waits on the event, formats the frame (blob rows split
// over several formatting threads) and appends it to the files.  Files,
// headers and row formats are the reference's:
//   <stamp>_vortex.dat      src/VortexBlobs.cpp:16-42, 310-331
//   <stamp>_vortex_num.dat  src/VortexBlobs.cpp:35-41, 327-330
//   <stamp>_gamma.dat       src/VortexSheet.cpp:38-46, 335-340
//   <stamp>_loads.dat       src/VortexSheet.cpp:48-53, 342
//   <stamp>_probe.dat       src/Probe.cpp:12-34, 42-52
// Numbers are printed as an std::ostream with default flags prints them
// (six significant digits, %g style).
#ifndef DVMPP_AMD_STEPWRITER_HPP
#define DVMPP_AMD_STEPWRITER_HPP

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dvm_hip.h"

namespace dvm {

struct OutputSpec {
    std::string dir;        ///< output directory, with a trailing '/'
    std::string stamp;      ///< file-name prefix
    double dt = 0.0;        ///< for the headers
    double steps = 0.0;     ///< for the headers
    std::vector<double> xc, zc, ds;  ///< panel collocation points and lengths
    std::vector<double> probe_x, probe_z;
    bool write_blobs = true;  ///< false: <stamp>_vortex.dat gets its header only
};

/// "%g" with six significant digits, as std::ostream prints a double.
std::string format_g(double v);

class StepWriter {
public:
    StepWriter(dvm_ctx_t ctx, dvm_sim_t sim, OutputSpec spec, int frames = 3);
    ~StepWriter();
    StepWriter(const StepWriter&) = delete;
    StepWriter& operator=(const StepWriter&) = delete;

    /// Stages the simulation's current state for output as step `step` at `time`.
    void push(unsigned step, double time);
    /// Blocks until every staged step is on disk; rethrows a writer error.
    void flush();

private:
    struct Frame {
        unsigned step = 0;
        double time = 0.0;
        int64_t n = 0;
        int64_t cap = 0;
        double *x = nullptr, *z = nullptr, *c = nullptr;  // pinned, cap each
        double *gamma = nullptr, *loads = nullptr;        // pinned
        std::vector<double> pu, pw;                        // probe velocities
        dvm_event_t ev = nullptr;
    };

    void writer_main();
    void write_frame(Frame& f);
    void release(Frame& f);
    void rethrow_locked();

    dvm_ctx_t m_ctx;
    dvm_sim_t m_sim;
    OutputSpec m_spec;
    std::vector<Frame> m_frames;
    std::deque<int> m_free, m_full;
    std::mutex m_mu;
    std::condition_variable m_cv;
    bool m_stop = false;
    int m_busy = 0;  ///< frames taken by the writer thread
    std::exception_ptr m_err;
    std::ofstream m_vortex, m_num, m_gamma, m_loads, m_probe;
    std::string m_vortex_path, m_num_path, m_gamma_path, m_loads_path, m_probe_path;
    void* m_probe_dev = nullptr;  // px pz u w uvs wvs on the device
    unsigned m_fmt_threads = 1;
    std::thread m_thread;
};

}  // namespace dvm

#endif
