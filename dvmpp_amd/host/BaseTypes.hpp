// Base types of the drop-in layer: the subset of the reference's Armadillo
// aliases its callers use (src/BaseTypes.hpp:10-27), without Armadillo.
#ifndef DVMPP_AMD_BASETYPES_HPP
#define DVMPP_AMD_BASETYPES_HPP

#include <cstddef>
#include <initializer_list>
#include <vector>

namespace math {
inline constexpr double pi = 3.14159265358979323846;
inline constexpr double rpi2 = 1.0 / (2.0 * 3.14159265358979323846);
template <typename T>
[[nodiscard]] constexpr T sqr(T x) noexcept { return x * x; }
}  // namespace math

// Column vector of doubles with the element access and sizing members the
// reference's callers use on arma::vec (m_x(i), n_elem, size(), set_size).
class Vector {
public:
    Vector() = default;
    explicit Vector(std::size_t n) : v_(n, 0.0) {}
    Vector(std::initializer_list<double> l) : v_(l) {}
    double& operator()(std::size_t i) { return v_[i]; }
    double operator()(std::size_t i) const { return v_[i]; }
    double& operator[](std::size_t i) { return v_[i]; }
    double operator[](std::size_t i) const { return v_[i]; }
    [[nodiscard]] std::size_t size() const { return v_.size(); }
    [[nodiscard]] std::size_t n_elem() const { return v_.size(); }
    void set_size(std::size_t n) { v_.resize(n); }
    void zeros(std::size_t n) { v_.assign(n, 0.0); }
    void fill(double a) { v_.assign(v_.size(), a); }
    double* memptr() { return v_.data(); }
    const double* memptr() const { return v_.data(); }
    auto begin() { return v_.begin(); }
    auto end() { return v_.end(); }
    auto begin() const { return v_.begin(); }
    auto end() const { return v_.end(); }
    void insert_rows(std::size_t at, const Vector& o) { v_.insert(v_.begin() + at, o.v_.begin(), o.v_.end()); }

private:
    std::vector<double> v_;
};

#endif
