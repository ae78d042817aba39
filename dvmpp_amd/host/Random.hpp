// Random number stream of the reference (src/Random.hpp:10-27): std::mt19937
// with std::uniform_real_distribution<double>, seeded from std::random_device
// and reseeded by seed(s) when s >= 0.
#ifndef DVMPP_AMD_RANDOM_HPP
#define DVMPP_AMD_RANDOM_HPP

#include <random>

class Random {
public:
    Random() { m_re.seed(m_rd()); }
    /// A number in [0, 1).
    double rand() { return m_dist(m_re); }
    /// <0: keep the random-device seed; >= 0: reseed.
    void seed(int s)
    {
        if (s >= 0) m_re.seed(s);
    }

private:
    std::uniform_real_distribution<> m_dist;
    std::mt19937 m_re;
    std::random_device m_rd;
};

#endif
