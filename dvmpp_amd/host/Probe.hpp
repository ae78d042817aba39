// Velocity probe points (src/Probe.hpp:16-31): positions and the velocities
// VortexSheet::probe_velocities writes.  The XML constructor and the output
// file of the reference are not part of this layer.
#ifndef DVMPP_AMD_PROBE_HPP
#define DVMPP_AMD_PROBE_HPP

#include <vector>

#include "BaseTypes.hpp"

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
    [[nodiscard]] unsigned size() const { return (unsigned)m_x.size(); }
};

#endif
