// Vortex sheet (src/VortexSheet.hpp:21-127) on the GPU.  The body comes from
// the input file (the reference's constructor, which also creates the gamma and
// loads files), from points, or from the reference's two-column body file.
#ifndef DVMPP_AMD_VORTEXSHEET_HPP
#define DVMPP_AMD_VORTEXSHEET_HPP

#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "BaseTypes.hpp"
#include "Probe.hpp"
#include "Random.hpp"
#include "VortexBlobs.hpp"
#include "XmlHandler.hpp"
#include "dvm_hip.h"

class VortexSheet {
public:
    /// No body (src/VortexSheet.cpp:7-10): size() is 0, get_forces() is
    /// (0, 0); the methods that need a body throw DVMException.
    VortexSheet();
    /// src/VortexSheet.cpp:12-54: constants from <constants>, <time>, <flow>;
    /// the body from <io><input_dir>/<domain_file>; creates
    /// <output_dir>/<timestamp>_gamma.dat and _loads.dat with their headers.
    VortexSheet(const XmlHandler& xml, const std::string& timestamp,
                int precision = DVM_F64);
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
    /// Appends gamma rows and the loads row (src/VortexSheet.cpp:335-343);
    /// needs the input-file constructor.
    void write_step(double time);
    void print_collocation() const;
    void print_unit_vectors() const;
    void print_gamma() const;
    [[nodiscard]] std::tuple<double, double> get_forces() const;
    /// Drag and lift coefficients of the last compute_loads (printed by the reference).
    [[nodiscard]] std::tuple<double, double> get_coefficients() const;
    [[nodiscard]] Vector gamma() const;
    void set_gamma(const Vector& g);
    [[nodiscard]] dvm_sheet_t handle() const { return m_sheet; }

private:
    void init(const std::vector<double>& x, const std::vector<double>& z);
    void need_body(const char* where) const;
    dvm_sheet_t m_sheet = nullptr;
    dvm_sheet_params m_params{};
    int m_prec = DVM_F64;
    unsigned m_nl = 0;
    double m_fx = 0.0, m_fz = 0.0, m_cd = 0.0, m_cl = 0.0;
    // host copy of the panel geometry (written and printed)
    Vector m_xc, m_zc, m_ds, m_enx, m_enz, m_etx, m_etz;
    std::ofstream m_gammafile;
    std::ofstream m_forcefile;
};

#endif
