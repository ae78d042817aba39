// Background mesh for remeshing.  The reference's BackgroundMesh source is
// missing; its API is the one tests/unit/test_BackgroundMesh.cpp:14-18,105-111,
// 167-204 and src/DVMBase.cpp:160-165 use: a default-constructed mesh is
// disabled, never remeshes, has zero dimensions and returns a copy.
#ifndef DVMPP_AMD_BACKGROUNDMESH_HPP
#define DVMPP_AMD_BACKGROUNDMESH_HPP

#include "VortexBlobs.hpp"
#include "dvm_hip.h"

class BackgroundMesh {
public:
    BackgroundMesh();
    explicit BackgroundMesh(const dvm_remesh_params& params);
    [[nodiscard]] bool is_enabled() const { return m_enabled; }
    [[nodiscard]] bool should_remesh(unsigned step) const;
    [[nodiscard]] unsigned nx() const { return m_nx; }
    [[nodiscard]] unsigned nz() const { return m_nz; }
    [[nodiscard]] VortexBlobs remesh(const VortexBlobs& blobs) const;

private:
    dvm_remesh_params m_p{};
    bool m_enabled = false;
    unsigned m_nx = 0, m_nz = 0;
};

#endif
