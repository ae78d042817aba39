// Vortex blobs (src/VortexBlobs.hpp:21-109): public SoA fields in host memory;
// the compute methods run on the GPU through the C-ABI; the output files are
// the reference's.
#ifndef DVMPP_AMD_VORTEXBLOBS_HPP
#define DVMPP_AMD_VORTEXBLOBS_HPP

#include <memory>
#include <string>
#include <vector>

#include "BaseTypes.hpp"
#include "Exceptions.hpp"
#include "Random.hpp"
#include "XmlHandler.hpp"

namespace dvm {
class VelocityKernel;
}

class VortexBlobs {
public:
    std::vector<unsigned> m_ID;
    Vector m_x;
    Vector m_z;
    Vector m_circ;
    Vector m_sigma;
    Vector m_u;
    Vector m_w;
    Vector m_uvs;
    Vector m_wvs;

    VortexBlobs();
    /// src/VortexBlobs.cpp:16-42: creates <output_dir>/<timestamp>_vortex.dat
    /// and _vortex_num.dat with their headers (no blobs yet).
    VortexBlobs(const XmlHandler& xml, const std::string& timestamp);
    explicit VortexBlobs(unsigned N);
    ~VortexBlobs();
    VortexBlobs(VortexBlobs&& other) noexcept;
    VortexBlobs& operator=(VortexBlobs&& other) noexcept;
    VortexBlobs(const VortexBlobs& other);
    VortexBlobs& operator=(const VortexBlobs& other);

    /// src/VortexBlobs.cpp:211-227: appended IDs are Nold + the new blob's ID.
    void append_vortices(VortexBlobs& NewVortBlobs);
    /// src/VortexBlobs.cpp:229-249 through the selected velocity kernel.
    void biotsavart();
    /// src/VortexBlobs.cpp:251-265: draws R1, R2 per blob from `rand` in the
    /// reference's order; the displacements are applied on the GPU.
    void diffusion_random_walk(Random& rand, double nu, double dt);
    [[nodiscard]] double totalcirc() const;
    /// src/VortexBlobs.cpp:110-209 on the GPU (dvm_merge_blobs): the same
    /// merges and values; returns the number of merged pairs.
    unsigned merge_blobs(double threshold);
    /// Throws SizeMismatchException if the fields disagree (src/VortexBlobs.cpp:272-285).
    [[nodiscard]] unsigned size() const;
    void print_location() const;
    void print_velocity() const;
    void print_circulation() const;
    /// Appends the blob rows and the count row (src/VortexBlobs.cpp:310-331).
    void write_step(double time, unsigned step);
    void setKernelType(const std::string& type);
    [[nodiscard]] std::string kernelType() const;
    void copyOutputConfig(const VortexBlobs& other);

private:
    std::string m_blobsfile;
    std::string m_numfile;
    std::unique_ptr<dvm::VelocityKernel> m_kernel;
};

#endif
