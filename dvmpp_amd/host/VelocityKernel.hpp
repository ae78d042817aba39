// Velocity-kernel plugin interface of the reference (src/VelocityKernel.hpp:12-110),
// with HIP implementations.
//
//   create("direct")      -> DirectKernel: FP64 direct sum on the GPU, the
//                            reference's pair arithmetic and order
//   create("direct_f32c") -> DirectF32CKernel: FP32-compensated direct sum
//   create("fmm")         -> FMMKernel(10, 0.5): 2-D fast multipole sum on the
//                            GPU (the reference's FMMKernel, a stub there that
//                            calls the direct sum); targets use the direct sum
//   anything else         -> AlgorithmException
#ifndef DVMPP_AMD_VELOCITYKERNEL_HPP
#define DVMPP_AMD_VELOCITYKERNEL_HPP

#include <memory>
#include <string>
#include <vector>

#include "BaseTypes.hpp"

namespace dvm {

struct VelocityResult {
    double u = 0.0;
    double w = 0.0;
};

struct VortexParticle {
    double x;
    double z;
    double circ;
    double sigma;
};

class VelocityKernel {
public:
    virtual ~VelocityKernel() = default;
    /// Blob-on-blob velocities; `velocities` is resized to particles.size().
    virtual void computeSelfInduced(const std::vector<VortexParticle>& particles,
                                    std::vector<VelocityResult>& velocities) = 0;
    /// Velocities at (target_x, target_z); pairs with r^2 <= 1e-15 skipped.
    virtual void computeInducedAt(const std::vector<VortexParticle>& sources,
                                  const std::vector<double>& target_x,
                                  const std::vector<double>& target_z,
                                  std::vector<VelocityResult>& velocities) = 0;
    [[nodiscard]] static std::unique_ptr<VelocityKernel> create(const std::string& type);
    [[nodiscard]] virtual std::string name() const = 0;
};

class DirectKernel : public VelocityKernel {
public:
    void computeSelfInduced(const std::vector<VortexParticle>& particles,
                            std::vector<VelocityResult>& velocities) override;
    void computeInducedAt(const std::vector<VortexParticle>& sources,
                          const std::vector<double>& target_x,
                          const std::vector<double>& target_z,
                          std::vector<VelocityResult>& velocities) override;
    [[nodiscard]] std::string name() const override { return "direct"; }
    /// dvm_precision this kernel passes to the C-ABI (not part of the
    /// reference's abstract interface, so its subclasses need not define it).
    [[nodiscard]] virtual int precision() const;
};

class DirectF32CKernel : public DirectKernel {
public:
    [[nodiscard]] std::string name() const override { return "direct_f32c"; }
    [[nodiscard]] int precision() const override;
};

/// src/VelocityKernel.hpp:84-110: expansion order (terms, 1..32) and opening
/// criterion theta in (0.2, 1]; the error falls roughly as theta^order.
class FMMKernel : public DirectKernel {
public:
    explicit FMMKernel(int expansion_order = 10, double theta = 0.5);
    void computeSelfInduced(const std::vector<VortexParticle>& particles,
                            std::vector<VelocityResult>& velocities) override;
    [[nodiscard]] std::string name() const override { return "fmm"; }
    [[nodiscard]] int precision() const override;
    [[nodiscard]] int expansion_order() const { return m_expansion_order; }
    [[nodiscard]] double theta() const { return m_theta; }

private:
    int m_expansion_order;
    double m_theta;
};

}  // namespace dvm

#endif
