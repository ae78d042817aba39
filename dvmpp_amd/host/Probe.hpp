// Velocity probe points (src/Probe.hpp:16-31): positions and the velocities
// VortexSheet::probe_velocities writes, and the reference's output file.
#ifndef DVMPP_AMD_PROBE_HPP
#define DVMPP_AMD_PROBE_HPP

#include <string>
#include <vector>

#include "BaseTypes.hpp"
#include "XmlHandler.hpp"

class Probe {
public:
    Vector m_x;    ///< x of the probe points
    Vector m_z;    ///< z of the probe points
    Vector m_u;    ///< x velocity
    Vector m_w;    ///< z velocity
    Vector m_uvs;  ///< x velocity induced by the vortex sheet
    Vector m_wvs;  ///< z velocity induced by the vortex sheet

    Probe() = default;
    Probe(const std::vector<double>& x, const std::vector<double>& z)
        : m_x(x.size()), m_z(z.size()), m_u(x.size()), m_w(x.size()), m_uvs(x.size()),
          m_wvs(x.size())
    {
        for (std::size_t i = 0; i < x.size(); ++i) {
            m_x(i) = x[i];
            m_z(i) = z[i];
        }
    }
    /// src/Probe.cpp:12-34: positions from <probe> x and z; creates
    /// <output_dir>/<timestamp>_probe.dat with its header.
    Probe(const XmlHandler& xml, const std::string& timestamp);
    [[nodiscard]] unsigned size() const { return (unsigned)m_x.size(); }
    /// Appends one row per probe (src/Probe.cpp:40-51).
    void write_step(double time);

private:
    std::string m_outfile;
};

#endif
