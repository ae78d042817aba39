// C++ tests of the drop-in classes (dvmpp_amd/host), mirroring the reference's
// GoogleTest suites case by case (tests/unit, tests/integration,
// tests/regression) with a minimal runner, since GoogleTest is not available
// offline.  Needs a GPU: the classes run the HIP kernels.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <sstream>
#include <random>
#include <set>
#include <string>
#include <vector>

#include <unistd.h>

#include "BackgroundMesh.hpp"
#include "DVMBase.hpp"
#include "Device.hpp"
#include "Exceptions.hpp"
#include "Random.hpp"
#include "VelocityKernel.hpp"
#include "VortexBlobs.hpp"
#include "VortexSheet.hpp"
#include "XmlHandler.hpp"

namespace {
struct Case {
    const char* name;
    std::function<void()> fn;
};
std::vector<Case>& cases()
{
    static std::vector<Case> c;
    return c;
}
int g_fail = 0;
struct Reg {
    Reg(const char* n, std::function<void()> f) { cases().push_back({n, f}); }
};
}  // namespace

#define TEST(name)                          \
    static void name();                     \
    static Reg reg_##name(#name, name);     \
    static void name()
#define EXPECT(cond)                                                              \
    do {                                                                          \
        if (!(cond)) {                                                            \
            std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);        \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)
#define EXPECT_NEAR(a, b, tol) EXPECT(std::fabs((a) - (b)) <= (tol))
#define EXPECT_THROW(stmt, exc)             \
    do {                                    \
        bool thrown_ = false;               \
        try {                               \
            stmt;                           \
        } catch (const exc&) {              \
            thrown_ = true;                 \
        }                                   \
        EXPECT(thrown_);                    \
    } while (0)

static std::string g_golden;

// ---------------------------------------------------------------- VelocityKernel
// tests/unit/test_VelocityKernel.cpp:12-125

TEST(CreateDirectKernel)
{
    auto k = dvm::VelocityKernel::create("direct");
    EXPECT(k != nullptr);
    EXPECT(k->name() == "direct");
    EXPECT(dvm::VelocityKernel::create("direct_f32c")->name() == "direct_f32c");
}

TEST(CreateInvalidKernelThrows) { EXPECT_THROW((void)dvm::VelocityKernel::create("invalid"), dvm::AlgorithmException); }

// tests/unit/test_VelocityKernel.cpp:134-165 (the USE_FMM cases)
TEST(CreateFMMKernel)
{
    auto k = dvm::VelocityKernel::create("fmm");
    EXPECT(k != nullptr);
    EXPECT(k->name() == "fmm");
    EXPECT_THROW(dvm::FMMKernel(0, 0.5), dvm::AlgorithmException);
    EXPECT_THROW(dvm::FMMKernel(10, 1.5), dvm::AlgorithmException);
}

static std::vector<dvm::VortexParticle> random_particles(int n, double lo, double hi, double sigma)
{
    std::mt19937 rng(42);
    std::uniform_real_distribution<> pos(lo, hi);
    std::uniform_real_distribution<> circ(-1.0, 1.0);
    std::vector<dvm::VortexParticle> p(n);
    for (int i = 0; i < n; ++i) p[i] = {pos(rng), pos(rng), circ(rng), sigma};
    return p;
}

TEST(FMMKernelMatchesDirectForSmallN)
{
    auto direct = dvm::VelocityKernel::create("direct");
    auto fmm = dvm::VelocityKernel::create("fmm");
    const auto p = random_particles(100, -10.0, 10.0, 0.1);
    std::vector<dvm::VelocityResult> vd, vf;
    direct->computeSelfInduced(p, vd);
    fmm->computeSelfInduced(p, vf);
    for (int i = 0; i < 100; ++i) {
        EXPECT_NEAR(vf[i].u, vd[i].u, 1e-6);
        EXPECT_NEAR(vf[i].w, vd[i].w, 1e-6);
    }
}

TEST(FMMKernelMatchesDirectWithATree)
{
    // large enough for a multi-level tree; order 20 reaches FP64 rounding
    dvm::FMMKernel fmm(20, 0.5);
    auto direct = dvm::VelocityKernel::create("direct");
    const auto p = random_particles(20000, -10.0, 10.0, 0.02);
    std::vector<dvm::VelocityResult> vd, vf;
    direct->computeSelfInduced(p, vd);
    fmm.computeSelfInduced(p, vf);
    int level = 0;
    dvm::check(dvm_fmm_last_level(dvm::device_context(), &level), "level", nullptr);
    EXPECT(level >= 4);
    double scale = 0.0, err = 0.0;
    for (size_t i = 0; i < p.size(); ++i) {
        scale = std::max(scale, std::fabs(vd[i].u) + std::fabs(vd[i].w));
        err = std::max(err, std::fabs(vf[i].u - vd[i].u) + std::fabs(vf[i].w - vd[i].w));
    }
    EXPECT(err <= 1e-10 * scale);
}

TEST(DirectKernelComputesSelfInducedVelocity)
{
    auto k = dvm::VelocityKernel::create("direct");
    std::vector<dvm::VortexParticle> p = {{0.0, 0.0, 1.0, 0.1}, {1.0, 0.0, -1.0, 0.1}};
    std::vector<dvm::VelocityResult> v;
    k->computeSelfInduced(p, v);
    EXPECT(v.size() == 2u);
    EXPECT(v[0].w != 0.0 && v[1].w != 0.0);
    EXPECT(v[0].w * v[1].w > 0.0);
}

TEST(DirectKernelEmptyParticlesReturnsEmpty)
{
    auto k = dvm::VelocityKernel::create("direct");
    std::vector<dvm::VortexParticle> p;
    std::vector<dvm::VelocityResult> v(3);
    k->computeSelfInduced(p, v);
    EXPECT(v.empty());
}

TEST(DirectKernelSingleParticleHasZeroSelfVelocity)
{
    for (const char* name : {"direct", "direct_f32c"}) {
        auto k = dvm::VelocityKernel::create(name);
        std::vector<dvm::VortexParticle> p = {{0.0, 0.0, 1.0, 0.1}};
        std::vector<dvm::VelocityResult> v;
        k->computeSelfInduced(p, v);
        EXPECT(v.size() == 1u);
        EXPECT(v[0].u == 0.0 && v[0].w == 0.0);
    }
}

TEST(DirectKernelComputesInducedAtTargets)
{
    auto k = dvm::VelocityKernel::create("direct");
    std::vector<dvm::VortexParticle> s = {{0.0, 0.0, 1.0, 0.1}};
    std::vector<dvm::VelocityResult> v;
    k->computeInducedAt(s, {1.0}, {0.0}, v);
    EXPECT(v.size() == 1u);
    EXPECT_NEAR(v[0].u, 0.0, 1e-10);
    EXPECT_NEAR(v[0].w, 1.0 / (2.0 * math::pi), 0.01);
}

TEST(BiotSavartVortexPairAnalytical)
{
    for (const char* name : {"direct", "direct_f32c"}) {
        auto k = dvm::VelocityKernel::create(name);
        const double d = 2.0, G = 1.0;
        std::vector<dvm::VortexParticle> p = {{-d / 2, 0.0, G, 0.01}, {d / 2, 0.0, -G, 0.01}};
        std::vector<dvm::VelocityResult> v;
        k->computeSelfInduced(p, v);
        const double ew = G / (2.0 * math::pi * d);
        EXPECT_NEAR(v[0].w, ew, 0.01 * ew);
        EXPECT_NEAR(v[1].w, ew, 0.01 * ew);
        EXPECT_NEAR(v[0].u, 0.0, 1e-6);
        EXPECT_NEAR(v[1].u, 0.0, 1e-6);
    }
}

// ---------------------------------------------------------------- BiotSavart integration
// tests/integration/test_BiotSavart.cpp:15-233

static VortexBlobs pair(double d, double g0, double g1, double sigma)
{
    VortexBlobs b(2);
    b.m_ID[0] = 0; b.m_ID[1] = 1;
    b.m_x(0) = -d / 2; b.m_x(1) = d / 2;
    b.m_z(0) = 0.0; b.m_z(1) = 0.0;
    b.m_circ(0) = g0; b.m_circ(1) = g1;
    b.m_sigma(0) = sigma; b.m_sigma(1) = sigma;
    return b;
}

TEST(VortexPairTranslation)
{
    VortexBlobs b = pair(2.0, 1.0, -1.0, 0.01);
    b.biotsavart();
    const double ew = 1.0 / (2.0 * math::pi * 2.0);
    EXPECT(b.m_w(0) > 0.0 && b.m_w(1) > 0.0);
    EXPECT_NEAR(b.m_w(0), ew, 0.01 * ew);
    EXPECT_NEAR(b.m_w(1), ew, 0.01 * ew);
    EXPECT_NEAR(b.m_u(0), 0.0, 1e-10);
    EXPECT_NEAR(b.m_u(1), 0.0, 1e-10);
}

TEST(CoRotatingVortexPair)
{
    VortexBlobs b = pair(2.0, 1.0, 1.0, 0.01);
    b.biotsavart();
    const double ew = 1.0 / (2.0 * math::pi * 2.0);
    EXPECT_NEAR(b.m_w(0), -ew, 0.01 * ew);
    EXPECT_NEAR(b.m_w(1), ew, 0.01 * ew);
    EXPECT_NEAR(b.m_u(0), 0.0, 1e-10);
    EXPECT_NEAR(b.m_u(1), 0.0, 1e-10);
}

TEST(RegularizationPreventsInfinity)
{
    VortexBlobs b = pair(0.001, 1.0, -1.0, 0.1);
    b.biotsavart();
    for (unsigned i = 0; i < 2; ++i) EXPECT(std::isfinite(b.m_u(i)) && std::isfinite(b.m_w(i)));
}

TEST(FourVortexLeapfrog)
{
    VortexBlobs b(4);
    const double xs[4] = {-1, -1, 1, 1}, zs[4] = {0.5, -0.5, 0.5, -0.5}, gs[4] = {1, -1, 1, -1};
    for (unsigned i = 0; i < 4; ++i) {
        b.m_ID[i] = i; b.m_x(i) = xs[i]; b.m_z(i) = zs[i]; b.m_circ(i) = gs[i]; b.m_sigma(i) = 0.05;
    }
    b.biotsavart();
    for (unsigned i = 0; i < 4; ++i) EXPECT(std::isfinite(b.m_u(i)) && std::isfinite(b.m_w(i)));
    EXPECT_NEAR(b.m_u(0), b.m_u(2), 1e-10);
    EXPECT_NEAR(b.m_u(1), b.m_u(3), 1e-10);
}

TEST(CircularRingApproximation)
{
    const int N = 20;
    VortexBlobs b(N);
    for (int i = 0; i < N; ++i) {
        const double th = 2.0 * math::pi * i / N;
        b.m_ID[i] = i; b.m_x(i) = std::cos(th); b.m_z(i) = std::sin(th);
        b.m_circ(i) = 1.0 / N; b.m_sigma(i) = 0.05;
    }
    b.biotsavart();
    EXPECT_NEAR(b.totalcirc(), 1.0, 1e-10);
    for (int i = 0; i < N; ++i) EXPECT(std::isfinite(b.m_u(i)) && std::isfinite(b.m_w(i)));
}

TEST(KernelConsistencyWithVortexBlobs)
{
    VortexBlobs b(3);
    const double xs[3] = {0.0, 1.0, 0.5}, zs[3] = {0.0, 0.0, 1.0}, gs[3] = {1.0, -0.5, 0.3};
    for (unsigned i = 0; i < 3; ++i) {
        b.m_ID[i] = i; b.m_x(i) = xs[i]; b.m_z(i) = zs[i]; b.m_circ(i) = gs[i]; b.m_sigma(i) = 0.1;
    }
    b.biotsavart();
    auto k = dvm::VelocityKernel::create("direct");
    std::vector<dvm::VortexParticle> p(3);
    for (int i = 0; i < 3; ++i) p[i] = {b.m_x(i), b.m_z(i), b.m_circ(i), b.m_sigma(i)};
    std::vector<dvm::VelocityResult> v;
    k->computeSelfInduced(p, v);
    for (int i = 0; i < 3; ++i) EXPECT(b.m_u(i) == v[i].u && b.m_w(i) == v[i].w);
}

// ---------------------------------------------------------------- VortexBlobs unit
// tests/unit/test_VortexBlobs.cpp:15-141

TEST(VortexBlobsConstruction)
{
    VortexBlobs a;
    EXPECT(a.size() == 0u);
    VortexBlobs b(10);
    EXPECT(b.size() == 10u);
    EXPECT(b.m_x.size() == 10u && b.m_sigma.size() == 10u && b.m_ID.size() == 10u);
}

TEST(AppendVorticesIncreasesSize)
{
    VortexBlobs a(3), b(2);
    for (unsigned i = 0; i < 2; ++i) {
        b.m_ID[i] = i; b.m_x(i) = i * 1.0; b.m_z(i) = i * 2.0; b.m_circ(i) = 0.1; b.m_sigma(i) = 0.05;
    }
    a.append_vortices(b);
    EXPECT(a.size() == 5u);
    EXPECT(a.m_ID[3] == 3u && a.m_ID[4] == 4u);
}

TEST(BiotSavartTwoVorticesOppositeCirculation)
{
    VortexBlobs b = pair(1.0, 1.0, -1.0, 0.1);
    b.biotsavart();
    const double ew = 1.0 / (2.0 * math::pi);
    EXPECT_NEAR(b.m_w(0), ew, 0.1 * ew);
    EXPECT_NEAR(b.m_w(1), ew, 0.1 * ew);
}

TEST(BiotSavartEmptyBlobsDoesNotCrash)
{
    VortexBlobs b;
    b.biotsavart();
    EXPECT(b.size() == 0u);
}

TEST(TotalCirculation)
{
    VortexBlobs b(3);
    b.m_circ(0) = 1.0; b.m_circ(1) = 2.0; b.m_circ(2) = -0.5;
    EXPECT(b.totalcirc() == 2.5);
    VortexBlobs e;
    EXPECT(e.totalcirc() == 0.0);
}

TEST(CopyAndMove)
{
    VortexBlobs a(3);
    a.m_x(0) = 1.0; a.m_x(1) = 2.0; a.m_x(2) = 3.0;
    VortexBlobs c(a);
    EXPECT(c.size() == 3u && c.m_x(0) == 1.0);
    a.m_x(0) = 100.0;
    EXPECT(c.m_x(0) == 1.0);
    VortexBlobs m(std::move(c));
    EXPECT(m.size() == 3u && m.m_x(0) == 1.0);
}

TEST(KernelType)
{
    VortexBlobs b;
    b.setKernelType("direct");
    EXPECT_THROW(b.setKernelType("invalid"), dvm::AlgorithmException);
}

TEST(SizeMismatchThrows)
{
    VortexBlobs b(3);
    b.m_x.set_size(2);
    EXPECT_THROW((void)b.size(), dvm::SizeMismatchException);
}

// ---------------------------------------------------------------- CylinderFlow regression
// tests/regression/test_CylinderFlow.cpp:26-167

TEST(DiffusionRandomWalkStatistics)
{
    const int N = 1000;
    VortexBlobs b(N);
    Random rng;
    rng.seed(42);
    for (int i = 0; i < N; ++i) { b.m_ID[i] = i; b.m_circ(i) = 1.0; b.m_sigma(i) = 0.1; }
    b.diffusion_random_walk(rng, 0.001, 0.01);
    double mx = 0, mz = 0;
    for (int i = 0; i < N; ++i) { mx += b.m_x(i); mz += b.m_z(i); }
    mx /= N; mz /= N;
    EXPECT_NEAR(mx, 0.0, 0.1);
    EXPECT_NEAR(mz, 0.0, 0.1);
    double vx = 0, vz = 0;
    for (int i = 0; i < N; ++i) { vx += (b.m_x(i) - mx) * (b.m_x(i) - mx); vz += (b.m_z(i) - mz) * (b.m_z(i) - mz); }
    const double sx = std::sqrt(vx / (N - 1)), sz = std::sqrt(vz / (N - 1));
    const double spread = std::sqrt(4.0 * 0.001 * 0.01);
    EXPECT(sx > 0.5 * spread && sx < 2.0 * spread);
    EXPECT(sz > 0.5 * spread && sz < 2.0 * spread);
}

TEST(CirculationConservation)
{
    const int N = 50;
    VortexBlobs b(N);
    Random rng;
    rng.seed(123);
    double total = 0.0;
    for (int i = 0; i < N; ++i) {
        b.m_ID[i] = i; b.m_x(i) = rng.rand() * 10.0 - 5.0; b.m_z(i) = rng.rand() * 10.0 - 5.0;
        b.m_circ(i) = rng.rand() * 2.0 - 1.0; b.m_sigma(i) = 0.1; total += b.m_circ(i);
    }
    const double before = b.totalcirc();
    b.biotsavart();
    EXPECT(b.totalcirc() == before);
    EXPECT_NEAR(before, total, 1e-12);
}

TEST(VortexAppendPreservesData)
{
    VortexBlobs a(3), b(2);
    for (int i = 0; i < 3; ++i) { a.m_ID[i] = i; a.m_x(i) = i * 1.0; a.m_z(i) = i * 2.0; a.m_circ(i) = i * 0.1; a.m_sigma(i) = 0.1; }
    for (int i = 0; i < 2; ++i) { b.m_ID[i] = i; b.m_x(i) = 10.0 + i; b.m_z(i) = 20.0 + i; b.m_circ(i) = 1.0 + i * 0.1; b.m_sigma(i) = 0.2; }
    const double before = a.totalcirc() + b.totalcirc();
    a.append_vortices(b);
    EXPECT(a.size() == 5u);
    EXPECT_NEAR(a.totalcirc(), before, 1e-15);
    EXPECT(a.m_x(0) == 0.0 && a.m_x(1) == 1.0 && a.m_x(2) == 2.0 && a.m_x(3) == 10.0 && a.m_x(4) == 11.0);
}

TEST(ReproducibilityWithSeed)
{
    auto run = [](int seed) {
        VortexBlobs b(10);
        Random rng;
        rng.seed(seed);
        for (int i = 0; i < 10; ++i) {
            b.m_ID[i] = i; b.m_x(i) = rng.rand(); b.m_z(i) = rng.rand();
            b.m_circ(i) = rng.rand() - 0.5; b.m_sigma(i) = 0.1;
        }
        b.biotsavart();
        b.diffusion_random_walk(rng, 0.001, 0.01);
        return b.m_x(0);
    };
    EXPECT(run(42) == run(42));
    EXPECT(run(42) != run(123));
}

// ---------------------------------------------------------------- Random
// tests/unit/test_Random.cpp:10-78

TEST(RandomRangeAndSeeds)
{
    Random r;
    for (int i = 0; i < 1000; ++i) { const double v = r.rand(); EXPECT(v >= 0.0 && v < 1.0); }
    Random a, b;
    a.seed(42); b.seed(42);
    for (int i = 0; i < 100; ++i) EXPECT(a.rand() == b.rand());
    Random c, d;
    c.seed(42); d.seed(123);
    bool diff = false;
    for (int i = 0; i < 100; ++i) diff = diff || c.rand() != d.rand();
    EXPECT(diff);
    Random e;
    e.seed(42);
    std::set<double> vals;
    for (int i = 0; i < 100; ++i) vals.insert(e.rand());
    EXPECT(vals.size() > 90u);
    Random z;
    z.seed(0);
    (void)z.rand();
}

// ---------------------------------------------------------------- BackgroundMesh
// tests/unit/test_BackgroundMesh.cpp:14-18,105-111,167-204

TEST(DisabledMesh)
{
    BackgroundMesh m;
    EXPECT(!m.is_enabled());
    for (unsigned s : {0u, 1u, 10u, 100u}) EXPECT(!m.should_remesh(s));
    EXPECT(m.nx() == 0u && m.nz() == 0u);
    VortexBlobs empty;
    EXPECT(m.remesh(empty).size() == 0u);
    VortexBlobs p(2);
    p.m_x(0) = 1.0; p.m_x(1) = 2.0; p.m_z(0) = 1.0; p.m_z(1) = 2.0;
    p.m_circ(0) = 0.5; p.m_circ(1) = 0.3; p.m_sigma(0) = 0.1; p.m_sigma(1) = 0.1;
    VortexBlobs r = m.remesh(p);
    EXPECT(r.size() == p.size() && r.m_x(0) == p.m_x(0) && r.totalcirc() == p.totalcirc());
}

TEST(EnabledMeshConservesCirculation)
{
    dvm_remesh_params rp{5, 0, -1.0, 2.0, -1.0, 2.0, 0.25, 0.25, 0.05, 0.0};
    BackgroundMesh m(rp);
    EXPECT(m.is_enabled() && m.nx() == 13u && m.nz() == 13u);
    EXPECT(m.should_remesh(10) && !m.should_remesh(11));
    VortexBlobs p(4);
    const double xs[4] = {0.25, 0.75, 0.25, 0.75}, zs[4] = {0.25, 0.25, 0.75, 0.75}, gs[4] = {0.5, 0.3, -0.2, 0.4};
    for (unsigned i = 0; i < 4; ++i) { p.m_ID[i] = i; p.m_x(i) = xs[i]; p.m_z(i) = zs[i]; p.m_circ(i) = gs[i]; p.m_sigma(i) = 0.1; }
    EXPECT(p.totalcirc() == 1.0);
    VortexBlobs r = m.remesh(p);
    EXPECT_NEAR(r.totalcirc(), 1.0, 1e-14);
}

// ---------------------------------------------------------------- VortexSheet + DVMBase

static dvm_sheet_params sheet_params()
{
    // tests/fixtures/test_configs/cylinder_test.xml
    return dvm_sheet_params{1.0, 0.001, 0.005, 3.0, 0.0, 0.675, 3};
}

TEST(VortexSheetMethods)
{
    VortexSheet sh(g_golden + "/simple_body.txt", sheet_params());
    EXPECT(sh.size() == 20u);
    VortexBlobs none;
    sh.solvevortexsheet(none);
    Random rng;
    rng.seed(42);
    VortexBlobs nascent = sh.release_nascent_vortices_rw(rng);
    EXPECT(nascent.size() >= 20u);
    double released = nascent.totalcirc();
    EXPECT(std::isfinite(released));
    VortexBlobs v;
    v.append_vortices(nascent);
    v.biotsavart();
    sh.solvevortexsheet(v);
    sh.vortexsheetbc(v);
    for (unsigned i = 0; i < v.size(); ++i) EXPECT(std::isfinite(v.m_uvs(i)) && std::isfinite(v.m_wvs(i)));
    // push a blob inside, reflect it out; delete another inside one
    v.m_x(0) = 0.1; v.m_z(0) = 0.0;
    sh.reflect(v);
    EXPECT(std::hypot(v.m_x(0), v.m_z(0)) > 0.45);
    v.m_x(1) = 0.0; v.m_z(1) = 0.1;
    const unsigned n0 = v.size();
    sh.delete_inside(v);
    EXPECT(v.size() == n0 - 1u && v.m_ID[0] == 0u && v.m_ID[v.size() - 1] == v.size() - 1);
    sh.compute_loads(3.0);
    auto [fx, fz] = sh.get_forces();
    EXPECT(std::isfinite(fx) && std::isfinite(fz));
    // probe_velocities: the sheet part is independent of the blobs (only the
    // last panel's term survives) and the total adds the free stream
    Probe probe({2.0, 3.0, -2.0}, {0.0, 0.5, 1.0});
    sh.probe_velocities(v, probe);
    Probe alone({2.0, 3.0, -2.0}, {0.0, 0.5, 1.0});
    sh.probe_velocities(VortexBlobs(), alone);
    for (unsigned i = 0; i < probe.size(); ++i) {
        EXPECT(std::isfinite(probe.m_u(i)) && std::isfinite(probe.m_w(i)));
        EXPECT(probe.m_uvs(i) == alone.m_uvs(i) && probe.m_wvs(i) == alone.m_wvs(i));
        EXPECT(alone.m_u(i) == alone.m_uvs(i) + sheet_params().ux);
    }
}

TEST(DVMBaseRunsCylinder)
{
    dvm_sim_params p{};
    p.sheet = sheet_params();
    p.scheme = DVM_EULER;
    p.crossing = DVM_REFLECT;
    p.seed = 42;
    p.prec = DVM_F64;
    p.rng_mode = DVM_RNG_REFERENCE;
    DVMBase dvm(g_golden + "/simple_body.txt", p, 10);
    dvm.solve();
    EXPECT(dvm.current_step() == 10u);
    EXPECT(dvm.num_blobs() >= 200u);
    VortexBlobs b = dvm.blobs();
    EXPECT(b.size() == dvm.num_blobs());
    auto [fx, fz, cd, cl] = dvm.loads();
    EXPECT(std::isfinite(cd) && std::isfinite(cl) && std::isfinite(fx) && std::isfinite(fz));
    EXPECT_NEAR(dvm.time(), 0.05, 1e-15);
}

// tests/unit/test_VortexSheet.cpp:12-27
TEST(VortexSheetDefaultConstructorWorks)
{
    VortexSheet sheet;
    EXPECT(sheet.size() == 0u);
}

TEST(VortexSheetGetForcesReturnsZeroForDefaultSheet)
{
    VortexSheet sheet;
    auto [fx, fz] = sheet.get_forces();
    EXPECT(fx == 0.0);
    EXPECT(fz == 0.0);
    // a default sheet has no body: the methods that need one say so
    VortexBlobs none;
    EXPECT_THROW(sheet.solvevortexsheet(none), dvm::DVMException);
}

// tests/unit/test_Probe.cpp:10-20
TEST(ProbeDefaultConstructorCreatesEmptyProbe)
{
    Probe probe;
    EXPECT(probe.size() == 0u);
}

TEST(ProbeSizeReturnsCorrectCount)
{
    Probe probe;
    EXPECT(probe.size() == 0u);
}

// A kernel written against the reference's abstract interface
// (src/VelocityKernel.hpp:29-62: two compute methods and name()) compiles
// and is usable through the base class.
namespace {
class ZeroKernel : public dvm::VelocityKernel {
public:
    void computeSelfInduced(const std::vector<dvm::VortexParticle>& p,
                            std::vector<dvm::VelocityResult>& v) override
    {
        v.assign(p.size(), dvm::VelocityResult{});
    }
    void computeInducedAt(const std::vector<dvm::VortexParticle>&, const std::vector<double>& tx,
                          const std::vector<double>&, std::vector<dvm::VelocityResult>& v) override
    {
        v.assign(tx.size(), dvm::VelocityResult{});
    }
    [[nodiscard]] std::string name() const override { return "zero"; }
};
}  // namespace

TEST(ReferenceStyleKernelSubclass)
{
    std::unique_ptr<dvm::VelocityKernel> k = std::make_unique<ZeroKernel>();
    std::vector<dvm::VortexParticle> p{{0.0, 0.0, 1.0, 0.1}, {1.0, 0.0, -1.0, 0.1}};
    std::vector<dvm::VelocityResult> v;
    k->computeSelfInduced(p, v);
    EXPECT(v.size() == 2u && v[0].u == 0.0 && k->name() == "zero");
}

// The input-file constructors (src/VortexSheet.cpp:12-54,
// src/VortexBlobs.cpp:16-42, src/Probe.cpp:12-34): the body from
// <io><input_dir>/<domain_file>, the output files with the reference headers.
namespace {
std::string read_file(const std::string& f)
{
    std::ifstream in(f);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}
}  // namespace

TEST(InputFileConstructors)
{
    namespace fs = std::filesystem;
    const fs::path dir = fs::temp_directory_path() / ("dvmpp_dropin_" + std::to_string(::getpid()));
    fs::create_directories(dir);
    const std::string xml_text =
        "<?xml version=\"1.0\"?>\n<DVMpp>\n"
        "<algorithms><surface_crossing string=\"REFLECT\" /></algorithms>\n"
        "<constants><density val=\"1.0\" /><nu val=\"0.001\" /><max_NumPanelVort val=\"3\" />"
        "<cutoff_exp val=\"0.675\" /><seed val=\"42\" /></constants>\n"
        "<io><input_dir string=\"" + g_golden + "\" /><output_dir string=\"" + dir.string() +
        "\" /><domain_file string=\"simple_body.txt\" /></io>\n"
        "<flow><ux val=\"3.0\" /><uz val=\"0.0\" /></flow>\n"
        "<probe><x string=\"2.0 3.0\" /><z string=\"0.0 0.5\" /></probe>\n"
        "<time><scheme string=\"euler\" /><dt val=\"0.005\" /><steps val=\"10\" /></time>\n"
        "</DVMpp>\n";
    const std::string xml_file = (dir / "case.xml").string();
    std::ofstream(xml_file) << xml_text;
    XmlHandler xml(xml_file);
    unsigned nblobs = 0;
    {  // the sheet's files are flushed when it goes (as the reference's are)
    VortexSheet sheet(xml, "stamp");
    EXPECT(sheet.size() == 20u);
    VortexBlobs blobs(xml, "stamp");
    EXPECT(blobs.size() == 0u);
    Probe probe(xml, "stamp");
    EXPECT(probe.size() == 2u && probe.m_x(1) == 3.0 && probe.m_z(1) == 0.5);

    // one step of the reference's loop through the classes, then the writers
    sheet.solvevortexsheet(blobs);
    Random rng;
    rng.seed(42);
    VortexBlobs nascent = sheet.release_nascent_vortices_rw(rng);
    blobs.append_vortices(nascent);
    sheet.compute_loads(3.0);
    sheet.write_step(0.005);
    blobs.write_step(0.005, 1);
    sheet.probe_velocities(blobs, probe);
    probe.write_step(0.005);
    nblobs = blobs.size();
    // a copy keeps the output files (src/VortexBlobs.cpp:64-77)
    VortexBlobs copy(blobs);
    copy.write_step(0.01, 2);
    }

    const std::string gam = read_file((dir / "stamp_gamma.dat").string());
    EXPECT(gam.rfind("20 # Number of collocation points\n0.005 # Time Step\n10 # Steps\n"
                     "Time [s] Gamma - Vortex sheet strength\n0.005\t ", 0) == 0);
    EXPECT(std::count(gam.begin(), gam.end(), '\n') == 4 + 20);
    const std::string loads = read_file((dir / "stamp_loads.dat").string());
    EXPECT(loads.rfind("Time [s]\tF_x [-]\tF_z [s]\n0.005\t", 0) == 0);
    const std::string vort = read_file((dir / "stamp_vortex.dat").string());
    EXPECT(vort.rfind("0 # Number of nodes\n0.005 # Time step\n10 # Steps\n"
                      "Time [s]\tx-position [m]\tz-position [m]\tcirculation\n", 0) == 0);
    EXPECT(std::count(vort.begin(), vort.end(), '\n') == 4 + 2 * (long)nblobs);
    const std::string num = read_file((dir / "stamp_vortex_num.dat").string());
    EXPECT(num == "Time [s]\tNumber of vortices\n1\t" + std::to_string(nblobs) + "\n2\t" +
                      std::to_string(nblobs) + "\n");
    const std::string pr = read_file((dir / "stamp_probe.dat").string());
    EXPECT(pr.rfind("0.005 # Time step\n10 # Steps\nTime [s]\tu [m/s]\tw [m/s]\n0.005\t", 0) == 0);
    EXPECT(std::count(pr.begin(), pr.end(), '\n') == 3 + 2);
    VortexBlobs plain;
    EXPECT_THROW(plain.write_step(0.0, 0), dvm::FileIOException);
    fs::remove_all(dir);
}

int main(int argc, char** argv)
{
    g_golden = argc > 1 ? argv[1] : "tests/golden";
    int failed_cases = 0;
    for (auto& c : cases()) {
        const int before = g_fail;
        try {
            c.fn();
        } catch (const std::exception& e) {
            std::printf("  FAIL %s: unexpected exception: %s\n", c.name, e.what());
            ++g_fail;
        }
        const bool ok = g_fail == before;
        failed_cases += ok ? 0 : 1;
        std::printf("[%s] %s\n", ok ? "  OK  " : " FAIL ", c.name);
        std::fflush(stdout);
    }
    std::printf("%zu cases, %d failed\n", cases().size(), failed_cases);
    return failed_cases == 0 ? 0 : 1;
}
