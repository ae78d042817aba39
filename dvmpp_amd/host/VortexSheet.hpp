// Vortex sheet (src/VortexSheet.hpp:21-127) on the GPU.  The body is given as
// points (or the reference's two-column body file); the XML constructor and the
// output files of the reference are not part of this layer.
#ifndef DVMPP_AMD_VORTEXSHEET_HPP
#define DVMPP_AMD_VORTEXSHEET_HPP

#include <string>
#include <tuple>
#include <vector>

#include "BaseTypes.hpp"
#include "Probe.hpp"
#include "Random.hpp"
#include "VortexBlobs.hpp"
#include "dvm_hip.h"

class VortexSheet {
public:
    /// Body points (closed: last == first) and the constants of
    /// src/VortexSheet.cpp:12-54.
    VortexSheet(const std::vector<double>& x, const std::vector<double>& z,
                const dvm_sheet_params& params, int precision = DVM_F64);
    /// Body file with rows "x z" (read_input_coord, src/VortexSheet.cpp:56-76).
    VortexSheet(const std::string& body_file, const dvm_sheet_params& params,
                int precision = DVM_F64);
    ~VortexSheet();
    VortexSheet(const VortexSheet&) = delete;
    VortexSheet& operator=(const VortexSheet&) = delete;

    void vortexsheetbc(VortexBlobs& blobs);
    void solvevortexsheet(VortexBlobs& blobs);
    [[nodiscard]] VortexBlobs release_nascent_vortices_rw(Random& rand);
    void reflect(VortexBlobs& vortex);
    void delete_inside(VortexBlobs& vortex);
    void compute_loads(double Urel);
    /// src/VortexSheet.cpp:559-646 with the reference's quirks (never called
    /// there): fills probe.m_u, m_w, m_uvs, m_wvs.
    void probe_velocities(const VortexBlobs& blobs, Probe& probe);
    [[nodiscard]] unsigned size() const;
    [[nodiscard]] std::tuple<double, double> get_forces() const;
    /// Drag and lift coefficients of the last compute_loads (printed by the reference).
    [[nodiscard]] std::tuple<double, double> get_coefficients() const;
    [[nodiscard]] Vector gamma() const;
    void set_gamma(const Vector& g);
    [[nodiscard]] dvm_sheet_t handle() const { return m_sheet; }

private:
    void init(const std::vector<double>& x, const std::vector<double>& z);
    dvm_sheet_t m_sheet = nullptr;
    dvm_sheet_params m_params;
    int m_prec;
    unsigned m_nl = 0;
    double m_fx = 0.0, m_fz = 0.0, m_cd = 0.0, m_cl = 0.0;
};

#endif
