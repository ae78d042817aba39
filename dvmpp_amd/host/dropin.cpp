// Drop-in C++17 classes over the C-ABI (include/dvm_hip.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <iterator>
#include <mutex>
#include <sstream>

#include "BackgroundMesh.hpp"
#include "DVMBase.hpp"
#include "Device.hpp"
#include "Exceptions.hpp"
#include "StepWriter.hpp"
#include "VelocityKernel.hpp"
#include "VortexBlobs.hpp"
#include "VortexSheet.hpp"

namespace dvm {

// ---------------------------------------------------------------- errors

void throw_status(dvm_status st, const std::string& where, dvm_ctx_t ctx)
{
    const std::string msg = where + ": " + (ctx ? dvm_last_error(ctx) : dvm_status_string(st));
    switch (st) {
    case DVM_E_INVALID: throw AlgorithmException(msg);
    case DVM_E_SIZE: throw SizeMismatchException(msg);
    case DVM_E_NAN: throw VortexException(msg);
    default: throw DVMException(msg);
    }
}

// ---------------------------------------------------------------- device

namespace {
std::mutex g_ctx_mutex;
dvm_ctx_t g_ctx = nullptr;  // lives until process exit (the OS reclaims it)
}  // namespace

dvm_ctx_t device_context()
{
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (!g_ctx) {
        const char* env = std::getenv("DVM_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        check(dvm_ctx_create(dev, nullptr, &g_ctx), "dvm_ctx_create", nullptr);
    }
    return g_ctx;
}

DeviceBuffer::~DeviceBuffer()
{
    if (ptr_) (void)dvm_free(device_context(), ptr_);
}

void DeviceBuffer::reset(std::size_t bytes)
{
    dvm_ctx_t ctx = device_context();
    if (ptr_) check(dvm_free(ctx, ptr_), "dvm_free", ctx);
    ptr_ = nullptr;
    bytes_ = bytes;
    if (bytes) check(dvm_malloc(ctx, bytes, &ptr_), "dvm_malloc", ctx);
}

void DeviceBuffer::upload(const void* src, std::size_t bytes, std::size_t offset)
{
    if (!bytes) return;
    dvm_ctx_t ctx = device_context();
    check(dvm_memcpy_h2d(ctx, static_cast<char*>(ptr_) + offset, src, bytes), "upload", ctx);
}

void DeviceBuffer::download(void* dst, std::size_t bytes, std::size_t offset) const
{
    if (!bytes) return;
    dvm_ctx_t ctx = device_context();
    check(dvm_memcpy_d2h(ctx, dst, static_cast<char*>(ptr_) + offset, bytes), "download", ctx);
}

// ---------------------------------------------------------------- VelocityKernel

std::unique_ptr<VelocityKernel> VelocityKernel::create(const std::string& type)
{
    if (type == "direct") return std::make_unique<DirectKernel>();
    if (type == "direct_f32c") return std::make_unique<DirectF32CKernel>();
    if (type == "fmm") return std::make_unique<FMMKernel>();
    throw AlgorithmException("Unknown velocity kernel type: " + type +
                             ". Available: direct, direct_f32c, fmm");
}

int DirectKernel::precision() const { return DVM_F64; }
int DirectF32CKernel::precision() const { return DVM_F32C; }
int FMMKernel::precision() const { return DVM_FMM; }

FMMKernel::FMMKernel(int expansion_order, double theta)
    : m_expansion_order(expansion_order), m_theta(theta)
{
    if (expansion_order < 1 || expansion_order > 32 || !(theta > 0.2 && theta <= 1.0))
        throw AlgorithmException("FMMKernel: expansion order must be 1..32 and theta in (0.2, 1]");
}

void FMMKernel::computeSelfInduced(const std::vector<VortexParticle>& particles,
                                   std::vector<VelocityResult>& velocities)
{
    dvm_ctx_t ctx = device_context();
    check(dvm_set_fmm(ctx, m_expansion_order, m_theta), "FMMKernel", ctx);
    DirectKernel::computeSelfInduced(particles, velocities);
}

static_assert(sizeof(VortexParticle) == 4 * sizeof(double), "AoS particle layout");
static_assert(sizeof(VelocityResult) == 2 * sizeof(double), "AoS result layout");

void DirectKernel::computeSelfInduced(const std::vector<VortexParticle>& particles,
                                      std::vector<VelocityResult>& velocities)
{
    velocities.resize(particles.size());
    if (particles.empty()) return;
    dvm_ctx_t ctx = device_context();
    check(dvm_bs_self_host(ctx, precision(), (int64_t)particles.size(),
                           reinterpret_cast<const double*>(particles.data()),
                           reinterpret_cast<double*>(velocities.data())),
          "computeSelfInduced", ctx);
}

void DirectKernel::computeInducedAt(const std::vector<VortexParticle>& sources,
                                    const std::vector<double>& target_x,
                                    const std::vector<double>& target_z,
                                    std::vector<VelocityResult>& velocities)
{
    if (target_x.size() != target_z.size()) throw SizeMismatchException("computeInducedAt targets");
    velocities.resize(target_x.size());
    if (target_x.empty()) return;
    dvm_ctx_t ctx = device_context();
    check(dvm_bs_targets_host(ctx, precision(), (int64_t)sources.size(),
                              reinterpret_cast<const double*>(sources.data()),
                              (int64_t)target_x.size(), target_x.data(), target_z.data(),
                              reinterpret_cast<double*>(velocities.data())),
          "computeInducedAt", ctx);
}

}  // namespace dvm

// ---------------------------------------------------------------- VortexBlobs

VortexBlobs::VortexBlobs() : m_kernel(dvm::VelocityKernel::create("direct")) {}

VortexBlobs::VortexBlobs(const XmlHandler& xml, const std::string& stamp) : VortexBlobs()
{
    const std::filesystem::path outpath(xml.getStringAttribute("io", "output_dir", true));
    const double dt = xml.getValueAttribute("time", "dt");
    const double steps = xml.getValueAttribute("time", "steps");
    m_blobsfile = (outpath / (stamp + "_vortex.dat")).string();
    std::ofstream bf(m_blobsfile);
    if (!bf) throw dvm::FileIOException::cannotOpen(m_blobsfile);
    bf << size() << " # Number of nodes\n"
       << dt << " # Time step\n"
       << steps << " # Steps\n"
       << "Time [s]\tx-position [m]\tz-position [m]\tcirculation\n";
    m_numfile = (outpath / (stamp + "_vortex_num.dat")).string();
    std::ofstream nf(m_numfile);
    if (!nf) throw dvm::FileIOException::cannotOpen(m_numfile);
    nf << "Time [s]\tNumber of vortices\n";
}

VortexBlobs::VortexBlobs(unsigned N) : VortexBlobs()
{
    m_ID.assign(N, 0u);
    for (Vector* v : {&m_x, &m_z, &m_circ, &m_sigma, &m_u, &m_w, &m_uvs, &m_wvs}) v->zeros(N);
}

VortexBlobs::~VortexBlobs() = default;
VortexBlobs::VortexBlobs(VortexBlobs&& other) noexcept = default;
VortexBlobs& VortexBlobs::operator=(VortexBlobs&& other) noexcept = default;

VortexBlobs::VortexBlobs(const VortexBlobs& o)
    : m_ID(o.m_ID), m_x(o.m_x), m_z(o.m_z), m_circ(o.m_circ), m_sigma(o.m_sigma), m_u(o.m_u),
      m_w(o.m_w), m_uvs(o.m_uvs), m_wvs(o.m_wvs), m_blobsfile(o.m_blobsfile),
      m_numfile(o.m_numfile), m_kernel(dvm::VelocityKernel::create(o.m_kernel ? o.m_kernel->name() : "direct"))
{
}

VortexBlobs& VortexBlobs::operator=(const VortexBlobs& o)
{
    if (this != &o) {
        m_ID = o.m_ID; m_x = o.m_x; m_z = o.m_z; m_circ = o.m_circ; m_sigma = o.m_sigma;
        m_u = o.m_u; m_w = o.m_w; m_uvs = o.m_uvs; m_wvs = o.m_wvs;
        m_blobsfile = o.m_blobsfile; m_numfile = o.m_numfile;
        m_kernel = dvm::VelocityKernel::create(o.m_kernel ? o.m_kernel->name() : "direct");
    }
    return *this;
}

void VortexBlobs::setKernelType(const std::string& type) { m_kernel = dvm::VelocityKernel::create(type); }

std::string VortexBlobs::kernelType() const { return m_kernel->name(); }

void VortexBlobs::copyOutputConfig(const VortexBlobs& other)
{
    m_blobsfile = other.m_blobsfile;
    m_numfile = other.m_numfile;
}

void VortexBlobs::print_location() const
{
    for (unsigned i = 0; i < size(); i++)
        std::cout << " The location of Vortex Blob " << m_ID[i] << " is (x,z) = (" << m_x(i) << ","
                  << m_z(i) << ")" << std::endl;
}

void VortexBlobs::print_velocity() const
{
    for (unsigned i = 0; i < size(); i++)
        std::cout << "(u,w) = (" << m_u(i) << "," << m_w(i) << ")" << std::endl;
}

void VortexBlobs::print_circulation() const
{
    for (unsigned i = 0; i < size(); i++) std::cout << " circ = " << m_circ(i) << std::endl;
}

void VortexBlobs::write_step(double time, unsigned step)
{
    if (m_blobsfile.empty() || m_numfile.empty())
        throw dvm::FileIOException("VortexBlobs::write_step: must set filenames before writing");
    std::ofstream bf(m_blobsfile, std::ios_base::app);
    if (!bf) throw dvm::FileIOException::cannotWrite(m_blobsfile);
    for (unsigned i = 0; i < size(); ++i)
        bf << time << "\t" << m_x(i) << "\t" << m_z(i) << "\t" << m_circ(i) << "\n";
    std::ofstream nf(m_numfile, std::ios_base::app);
    if (!nf) throw dvm::FileIOException::cannotWrite(m_numfile);
    nf << step << "\t" << size() << "\n";
}

// ---------------------------------------------------------------- Probe

Probe::Probe(const XmlHandler& xml, const std::string& stamp)
{
    const std::vector<double> x = xml.getList("probe", "x");
    const std::vector<double> z = xml.getList("probe", "z");
    m_x.zeros(x.size());
    m_z.zeros(z.size());
    for (std::size_t i = 0; i < x.size(); ++i) m_x(i) = x[i];
    for (std::size_t i = 0; i < z.size(); ++i) m_z(i) = z[i];
    for (Vector* v : {&m_u, &m_w, &m_uvs, &m_wvs}) v->zeros(x.size());
    const std::filesystem::path outpath(xml.getStringAttribute("io", "output_dir", true));
    m_outfile = (outpath / (stamp + "_probe.dat")).string();
    std::ofstream of(m_outfile);
    if (!of) throw dvm::FileIOException::cannotOpen(m_outfile);
    of << xml.getValueAttribute("time", "dt") << " # Time step\n"
       << xml.getValueAttribute("time", "steps") << " # Steps\n"
       << "Time [s]\tu [m/s]\tw [m/s]\n";
}

void Probe::write_step(double time)
{
    std::ofstream of(m_outfile, std::ios_base::app);
    if (!of) throw dvm::FileIOException::cannotWrite(m_outfile);
    for (unsigned i = 0; i < size(); ++i) of << time << "\t" << m_u(i) << "\t" << m_w(i) << "\n";
}

void VortexBlobs::append_vortices(VortexBlobs& nb)
{
    const unsigned nold = size();
    const unsigned nnew = nb.size();
    for (unsigned i = 0; i < nnew; i++) m_ID.push_back(nold + nb.m_ID[i]);
    m_x.insert_rows(m_x.size(), nb.m_x);
    m_z.insert_rows(m_z.size(), nb.m_z);
    m_circ.insert_rows(m_circ.size(), nb.m_circ);
    m_sigma.insert_rows(m_sigma.size(), nb.m_sigma);
    m_u.insert_rows(m_u.size(), nb.m_u);
    m_w.insert_rows(m_w.size(), nb.m_w);
    m_uvs.insert_rows(m_uvs.size(), nb.m_uvs);
    m_wvs.insert_rows(m_wvs.size(), nb.m_wvs);
}

void VortexBlobs::biotsavart()
{
    const unsigned N = size();
    if (N == 0) return;
    std::vector<dvm::VortexParticle> particles(N);
    for (unsigned i = 0; i < N; ++i) particles[i] = {m_x(i), m_z(i), m_circ(i), m_sigma(i)};
    std::vector<dvm::VelocityResult> velocities;
    m_kernel->computeSelfInduced(particles, velocities);
    for (unsigned i = 0; i < N; ++i) {
        m_u(i) = velocities[i].u;
        m_w(i) = velocities[i].w;
    }
}

void VortexBlobs::diffusion_random_walk(Random& rand, double nu, double dt)
{
    const unsigned N = size();
    if (N == 0) return;
    // The reference's draw order and Box-Muller transform; the increments are
    // applied to the positions on the device.
    std::vector<double> tape(2 * (size_t)N);
    for (unsigned i = 0; i < N; i++) {
        const double R1 = rand.rand();
        const double R2 = rand.rand();
        const double rrw = std::sqrt(4.0 * nu * dt * std::log(1.0 / R1));
        const double thetarw = 2.0 * math::pi * R2;
        tape[2 * i] = rrw * std::cos(thetarw);
        tape[2 * i + 1] = rrw * std::sin(thetarw);
    }
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(4 * b);
    buf.upload(tape.data(), 2 * b, 0);
    buf.upload(m_x.memptr(), b, 2 * b);
    buf.upload(m_z.memptr(), b, 3 * b);
    dvm_ctx_t ctx = dvm::device_context();
    double* d = buf.as<double>();
    dvm::check(dvm_diffuse(ctx, N, d, d + 2 * (size_t)N, d + 3 * (size_t)N), "diffusion_random_walk", ctx);
    buf.download(m_x.memptr(), b, 2 * b);
    buf.download(m_z.memptr(), b, 3 * b);
}

unsigned VortexBlobs::merge_blobs(double threshold)
{
    const unsigned N = size();
    if (N < 2) return 0;
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(8 * b + sizeof(uint32_t) * N);
    Vector* f[8] = {&m_x, &m_z, &m_circ, &m_sigma, &m_u, &m_w, &m_uvs, &m_wvs};
    for (int k = 0; k < 8; ++k) buf.upload(f[k]->memptr(), b, k * b);
    double* d = buf.as<double>();
    uint32_t* id = reinterpret_cast<uint32_t*>(buf.as<char>() + 8 * b);
    dvm_ctx_t ctx = dvm::device_context();
    int64_t merged = 0;
    dvm::check(dvm_merge_blobs(ctx, threshold, N, id, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N,
                               d + 4 * (size_t)N, d + 5 * (size_t)N, d + 6 * (size_t)N,
                               d + 7 * (size_t)N, &merged),
               "merge_blobs", ctx);
    if (merged == 0) return 0;
    const unsigned m = N - (unsigned)merged;
    for (int k = 0; k < 8; ++k) {
        f[k]->set_size(m);
        buf.download(f[k]->memptr(), sizeof(double) * m, k * b);
    }
    std::vector<uint32_t> ids(m);
    buf.download(ids.data(), sizeof(uint32_t) * m, 8 * b);
    m_ID.assign(ids.begin(), ids.end());
    return (unsigned)merged;
}

double VortexBlobs::totalcirc() const
{
    // Armadillo's accumulate order: two interleaved partial sums.
    double a1 = 0.0, a2 = 0.0;
    const size_t n = m_circ.size();
    size_t j;
    for (j = 1; j < n; j += 2) {
        a1 += m_circ(j - 1);
        a2 += m_circ(j);
    }
    if ((j - 1) < n) a1 += m_circ(j - 1);
    return a1 + a2;
}

unsigned VortexBlobs::size() const
{
    const size_t n = m_ID.size();
    if (n != m_x.size() || n != m_z.size() || n != m_circ.size() || n != m_sigma.size() ||
        n != m_u.size() || n != m_w.size() || n != m_uvs.size() || n != m_wvs.size())
        throw dvm::SizeMismatchException("VortexBlobs");
    return static_cast<unsigned>(n);
}

// ---------------------------------------------------------------- VortexSheet

static void read_body(const std::string& file, std::vector<double>& x, std::vector<double>& z)
{
    std::ifstream in(file);
    if (!in) throw dvm::FileIOException::cannotRead(file);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        double a, b;
        if (ss >> a >> b) {
            x.push_back(a);
            z.push_back(b);
        }
    }
    if (x.size() < 3) throw dvm::FileIOException::cannotRead(file);
}

VortexSheet::VortexSheet(const std::vector<double>& x, const std::vector<double>& z,
                         const dvm_sheet_params& params, int precision)
    : m_params(params), m_prec(precision)
{
    init(x, z);
}

VortexSheet::VortexSheet(const std::string& body_file, const dvm_sheet_params& params, int precision)
    : m_params(params), m_prec(precision)
{
    std::vector<double> x, z;
    read_body(body_file, x, z);
    init(x, z);
}

VortexSheet::VortexSheet() = default;

VortexSheet::VortexSheet(const XmlHandler& xml, const std::string& stamp, int precision)
    : m_prec(precision)
{
    m_params.rho = xml.getValueAttribute("constants", "density");
    m_params.nu = xml.getValueAttribute("constants", "nu");
    m_params.dt = xml.getValueAttribute("time", "dt");
    m_params.ux = xml.getValueAttribute("flow", "ux");
    m_params.uz = xml.getValueAttribute("flow", "uz");
    m_params.max_npv = (int)static_cast<unsigned>(xml.getValueAttribute("constants", "max_NumPanelVort"));
    m_params.cutoff_exp = xml.getValueAttribute("constants", "cutoff_exp");

    const std::filesystem::path domain = std::filesystem::path(xml.getStringAttribute("io", "input_dir")) /
                                         xml.getStringAttribute("io", "domain_file");
    std::vector<double> x, z;
    read_body(domain.string(), x, z);
    std::cout << "Successfully loaded coordinate file with " << x.size() << " points." << std::endl;
    std::cout << "A = " << *std::max_element(x.begin(), x.end()) - *std::min_element(x.begin(), x.end())
              << " (assuming cylinder)\n";
    init(x, z);
    std::cout << "Created vortex sheet of size " << m_nl << std::endl;

    const std::filesystem::path outpath(xml.getStringAttribute("io", "output_dir", true));
    const std::string gamma_path = (outpath / (stamp + "_gamma.dat")).string();
    m_gammafile.open(gamma_path);
    if (!m_gammafile) throw dvm::FileIOException::cannotOpen(gamma_path);
    m_gammafile << m_nl << " # Number of collocation points\n"
                << m_params.dt << " # Time Step\n"
                << xml.getValueAttribute("time", "steps") << " # Steps\n"
                << "Time [s] Gamma - Vortex sheet strength\n";
    const std::string force_path = (outpath / (stamp + "_loads.dat")).string();
    m_forcefile.open(force_path);
    if (!m_forcefile) throw dvm::FileIOException::cannotOpen(force_path);
    m_forcefile << "Time [s]\tF_x [-]\tF_z [s]\n";
}

void VortexSheet::init(const std::vector<double>& x, const std::vector<double>& z)
{
    if (x.size() != z.size()) throw dvm::SizeMismatchException("VortexSheet body");
    dvm_ctx_t ctx = dvm::device_context();
    dvm::check(dvm_sheet_create(ctx, (int)x.size(), x.data(), z.data(), &m_params, &m_sheet),
               "VortexSheet", ctx);
    int nl = 0;
    dvm::check(dvm_sheet_size(m_sheet, &nl), "VortexSheet::size", ctx);
    m_nl = (unsigned)nl;
    for (Vector* v : {&m_xc, &m_zc, &m_ds, &m_enx, &m_enz, &m_etx, &m_etz}) v->zeros(m_nl);
    double aref = 0.0;
    dvm::check(dvm_sheet_geometry(m_sheet, m_xc.memptr(), m_zc.memptr(), nullptr, m_ds.memptr(),
                                  m_enx.memptr(), m_enz.memptr(), m_etx.memptr(), m_etz.memptr(),
                                  &aref),
               "VortexSheet geometry", ctx);
}

void VortexSheet::need_body(const char* where) const
{
    if (!m_sheet)
        throw dvm::DVMException(std::string("VortexSheet::") + where +
                                ": the sheet has no body (default-constructed)");
}

void VortexSheet::write_step(double time)
{
    if (!m_gammafile.is_open() || !m_forcefile.is_open())
        throw dvm::FileIOException("VortexSheet::write_step: no output files (not built from an input file)");
    const Vector g = m_nl ? gamma() : Vector();
    for (unsigned i = 0; i < m_nl; ++i)
        m_gammafile << time << "\t " << m_xc(i) << "\t" << m_zc(i) << "\t" << g(i) << "\t" << m_ds(i)
                    << "\n";
    m_forcefile << time << "\t" << m_fx << "\t" << m_fz << "\n";
}

void VortexSheet::print_collocation() const
{
    for (unsigned i = 0; i < m_nl; i++)
        std::cout << "(xc,zc) = (" << m_xc(i) << "," << m_zc(i) << ")" << std::endl;
}

void VortexSheet::print_unit_vectors() const
{
    for (unsigned i = 0; i < m_nl; i++)
        std::cout << "For panel No " << i << " the normal vectors are en = (" << m_enx(i) << ","
                  << m_enz(i) << ") and et = (" << m_etx(i) << "," << m_etz(i) << ")" << std::endl;
}

void VortexSheet::print_gamma() const
{
    // one value per line (the reference prints the Armadillo column)
    const Vector g = m_nl ? gamma() : Vector();
    for (unsigned i = 0; i < m_nl; i++) std::cout << g(i) << std::endl;
}

VortexSheet::~VortexSheet()
{
    if (m_sheet) (void)dvm_sheet_destroy(m_sheet);
}

unsigned VortexSheet::size() const { return m_nl; }

std::tuple<double, double> VortexSheet::get_forces() const { return {m_fx, m_fz}; }
std::tuple<double, double> VortexSheet::get_coefficients() const { return {m_cd, m_cl}; }

Vector VortexSheet::gamma() const
{
    need_body("gamma");
    Vector g(m_nl);
    dvm::check(dvm_sheet_get_gamma(m_sheet, g.memptr()), "gamma", dvm::device_context());
    return g;
}

void VortexSheet::set_gamma(const Vector& g)
{
    need_body("set_gamma");
    if (g.size() != m_nl) throw dvm::SizeMismatchException("VortexSheet gamma");
    dvm::check(dvm_sheet_set_gamma(m_sheet, g.memptr()), "set_gamma", dvm::device_context());
}

void VortexSheet::solvevortexsheet(VortexBlobs& blobs)
{
    need_body("solvevortexsheet");
    const unsigned N = blobs.size();
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(4 * b + 8);
    buf.upload(blobs.m_x.memptr(), b, 0);
    buf.upload(blobs.m_z.memptr(), b, b);
    buf.upload(blobs.m_circ.memptr(), b, 2 * b);
    buf.upload(blobs.m_sigma.memptr(), b, 3 * b);
    double* d = buf.as<double>();
    dvm::check(dvm_sheet_solve(m_sheet, m_prec, N, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N,
                               nullptr),
               "solvevortexsheet", dvm::device_context());
    dvm::check(dvm_ctx_sync(dvm::device_context()), "solvevortexsheet", dvm::device_context());
}

void VortexSheet::vortexsheetbc(VortexBlobs& blobs)
{
    need_body("vortexsheetbc");
    const unsigned N = blobs.size();
    if (N == 0) return;
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(4 * b);
    buf.upload(blobs.m_x.memptr(), b, 0);
    buf.upload(blobs.m_z.memptr(), b, b);
    double* d = buf.as<double>();
    dvm::check(dvm_sheet_on_blobs(m_sheet, N, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N),
               "vortexsheetbc", dvm::device_context());
    buf.download(blobs.m_uvs.memptr(), b, 2 * b);
    buf.download(blobs.m_wvs.memptr(), b, 3 * b);
}

void VortexSheet::probe_velocities(const VortexBlobs& blobs, Probe& probe)
{
    need_body("probe_velocities");
    const unsigned N = blobs.size();
    const unsigned P = probe.size();
    if (probe.m_z.size() != P) throw dvm::SizeMismatchException("Probe");
    probe.m_u.zeros(P);
    probe.m_w.zeros(P);
    probe.m_uvs.zeros(P);
    probe.m_wvs.zeros(P);
    if (P == 0) return;
    const size_t b = sizeof(double) * N, pb = sizeof(double) * P;
    dvm::DeviceBuffer buf(4 * b + 6 * pb + 8);
    buf.upload(blobs.m_x.memptr(), b, 0);
    buf.upload(blobs.m_z.memptr(), b, b);
    buf.upload(blobs.m_circ.memptr(), b, 2 * b);
    buf.upload(blobs.m_sigma.memptr(), b, 3 * b);
    buf.upload(probe.m_x.memptr(), pb, 4 * b);
    buf.upload(probe.m_z.memptr(), pb, 4 * b + pb);
    double* d = buf.as<double>();
    double* p = d + 4 * (size_t)N;
    dvm::check(dvm_sheet_probe_quirk(m_sheet, m_prec, N, d, d + N, d + 2 * (size_t)N,
                                     d + 3 * (size_t)N, P, p, p + P, p + 2 * (size_t)P,
                                     p + 3 * (size_t)P, p + 4 * (size_t)P, p + 5 * (size_t)P),
               "probe_velocities", dvm::device_context());
    buf.download(probe.m_u.memptr(), pb, 4 * b + 2 * pb);
    buf.download(probe.m_w.memptr(), pb, 4 * b + 3 * pb);
    buf.download(probe.m_uvs.memptr(), pb, 4 * b + 4 * pb);
    buf.download(probe.m_wvs.memptr(), pb, 4 * b + 5 * pb);
}

VortexBlobs VortexSheet::release_nascent_vortices_rw(Random& rand)
{
    need_body("release_nascent_vortices_rw");
    int64_t nrv = 0;
    dvm::check(dvm_sheet_release_counts(m_sheet, &nrv), "VortexSheet::release_nascent_vortices_rw",
               dvm::device_context());
    VortexBlobs out((unsigned)nrv);
    if (nrv == 0) return out;
    std::vector<double> rrw((size_t)nrv);
    const double k = 4.0 * m_params.nu * m_params.dt;
    for (int64_t i = 0; i < nrv; ++i) {
        const double R = rand.rand();
        rrw[(size_t)i] = std::fabs(std::sqrt(k * std::log(1.0 / R)));
    }
    const size_t b = sizeof(double) * nrv;
    dvm::DeviceBuffer buf(9 * b + sizeof(uint32_t) * nrv);
    buf.upload(rrw.data(), b, 0);
    double* d = buf.as<double>();
    uint32_t* id = reinterpret_cast<uint32_t*>(d + 9 * nrv);
    dvm::check(dvm_sheet_release_emit(m_sheet, nrv, d, 0u, 0, d + nrv, d + 2 * nrv, d + 3 * nrv,
                                      d + 4 * nrv, d + 5 * nrv, d + 6 * nrv, d + 7 * nrv,
                                      d + 8 * nrv, id),
               "release_nascent_vortices_rw", dvm::device_context());
    Vector* f[8] = {&out.m_x, &out.m_z, &out.m_circ, &out.m_sigma, &out.m_u, &out.m_w, &out.m_uvs, &out.m_wvs};
    for (int k2 = 0; k2 < 8; ++k2) buf.download(f[k2]->memptr(), b, (k2 + 1) * b);
    return out;
}

void VortexSheet::reflect(VortexBlobs& v)
{
    need_body("reflect");
    const unsigned N = v.size();
    if (N == 0) return;
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(2 * b);
    buf.upload(v.m_x.memptr(), b, 0);
    buf.upload(v.m_z.memptr(), b, b);
    double* d = buf.as<double>();
    dvm::check(dvm_sheet_reflect(m_sheet, N, d, d + N), "reflect", dvm::device_context());
    buf.download(v.m_x.memptr(), b, 0);
    buf.download(v.m_z.memptr(), b, b);
}

void VortexSheet::delete_inside(VortexBlobs& v)
{
    need_body("delete_inside");
    const unsigned N = v.size();
    if (N == 0) return;
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer buf(8 * b + sizeof(uint32_t) * N);
    Vector* f[8] = {&v.m_x, &v.m_z, &v.m_circ, &v.m_sigma, &v.m_u, &v.m_w, &v.m_uvs, &v.m_wvs};
    for (int k = 0; k < 8; ++k) buf.upload(f[k]->memptr(), b, k * b);
    buf.upload(v.m_ID.data(), sizeof(uint32_t) * N, 8 * b);
    double* d = buf.as<double>();
    uint32_t* id = reinterpret_cast<uint32_t*>(d + 8 * (size_t)N);
    int64_t m = 0;
    dvm::check(dvm_sheet_delete_inside(m_sheet, N, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N,
                                       d + 4 * (size_t)N, d + 5 * (size_t)N, d + 6 * (size_t)N,
                                       d + 7 * (size_t)N, id, &m),
               "delete_inside", dvm::device_context());
    if (m == (int64_t)N) return;
    VortexBlobs nv((unsigned)m);
    nv.setKernelType(v.kernelType());
    Vector* g[8] = {&nv.m_x, &nv.m_z, &nv.m_circ, &nv.m_sigma, &nv.m_u, &nv.m_w, &nv.m_uvs, &nv.m_wvs};
    for (int k = 0; k < 8; ++k) buf.download(g[k]->memptr(), sizeof(double) * m, k * b);
    buf.download(nv.m_ID.data(), sizeof(uint32_t) * m, 8 * b);
    v = std::move(nv);
}

void VortexSheet::compute_loads(double Urel)
{
    need_body("compute_loads");
    double out[4];
    dvm::check(dvm_sheet_loads(m_sheet, Urel, out), "compute_loads", dvm::device_context());
    m_fx = out[0];
    m_fz = out[1];
    m_cd = out[2];
    m_cl = out[3];
}

// ---------------------------------------------------------------- BackgroundMesh

BackgroundMesh::BackgroundMesh() = default;

BackgroundMesh::BackgroundMesh(const dvm_remesh_params& p) : m_p(p), m_enabled(p.frequency > 0)
{
    if (m_enabled) {
        m_nx = (unsigned)(std::llround((p.xmax - p.xmin) / p.dx) + 1);
        m_nz = (unsigned)(std::llround((p.zmax - p.zmin) / p.dz) + 1);
    }
}

bool BackgroundMesh::should_remesh(unsigned step) const
{
    return m_enabled && step > 0 && step % (unsigned)m_p.frequency == 0;
}

VortexBlobs BackgroundMesh::remesh(const VortexBlobs& blobs) const
{
    if (!m_enabled) return blobs;
    const unsigned N = blobs.size();
    const size_t b = sizeof(double) * N;
    dvm::DeviceBuffer in(4 * b + 8);
    in.upload(blobs.m_x.memptr(), b, 0);
    in.upload(blobs.m_z.memptr(), b, b);
    in.upload(blobs.m_circ.memptr(), b, 2 * b);
    in.upload(blobs.m_sigma.memptr(), b, 3 * b);
    double* d = in.as<double>();
    dvm_ctx_t ctx = dvm::device_context();
    int64_t m = 0;
    dvm_status st = dvm_remesh(ctx, &m_p, N, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N, 0,
                               nullptr, nullptr, nullptr, nullptr, &m);
    if (st != DVM_OK && st != DVM_E_SIZE) dvm::throw_status(st, "remesh", ctx);
    VortexBlobs out((unsigned)m);
    out.setKernelType(blobs.kernelType());
    if (m == 0) return out;
    const size_t bm = sizeof(double) * m;
    dvm::DeviceBuffer o(4 * bm);
    double* od = o.as<double>();
    dvm::check(dvm_remesh(ctx, &m_p, N, d, d + N, d + 2 * (size_t)N, d + 3 * (size_t)N, m, od,
                          od + m, od + 2 * m, od + 3 * m, &m),
               "remesh", ctx);
    o.download(out.m_x.memptr(), bm, 0);
    o.download(out.m_z.memptr(), bm, bm);
    o.download(out.m_circ.memptr(), bm, 2 * bm);
    o.download(out.m_sigma.memptr(), bm, 3 * bm);
    for (unsigned i = 0; i < (unsigned)m; ++i) out.m_ID[i] = i;
    return out;
}

// ---------------------------------------------------------------- DVMBase

DVMBase::DVMBase(const std::vector<double>& x, const std::vector<double>& z,
                 const dvm_sim_params& params, unsigned steps)
    : m_p(params), m_steps(steps)
{
    init(x, z);
}

DVMBase::DVMBase(const std::string& body_file, const dvm_sim_params& params, unsigned steps)
    : m_p(params), m_steps(steps)
{
    std::vector<double> x, z;
    read_body(body_file, x, z);
    init(x, z);
}

void DVMBase::init(const std::vector<double>& x, const std::vector<double>& z)
{
    if (x.size() != z.size()) throw dvm::SizeMismatchException("DVMBase body");
    dvm_ctx_t ctx = dvm::device_context();
    dvm::check(dvm_sim_create(ctx, (int)x.size(), x.data(), z.data(), &m_p, &m_sim), "DVMBase", ctx);
}

DVMBase::DVMBase(XmlHandler& xml, const std::string& timestamp) : m_steps(0)
{
    from_xml(xml, timestamp, RunOptions{});
}

DVMBase::DVMBase(XmlHandler& xml, const std::string& timestamp, const RunOptions& opts) : m_steps(0)
{
    from_xml(xml, timestamp, opts);
}

// The reference's constructor (src/DVMBase.cpp:8-66) and those of its members
// (src/VortexBlobs.cpp:16-42, src/VortexSheet.cpp:12-54, src/Probe.cpp:12-34):
// the same entries, the same checks, the same files.
void DVMBase::from_xml(XmlHandler& xml, const std::string& timestamp, const RunOptions& opts)
{
    m_opts = opts;
    if (m_opts.output_every == 0) m_opts.output_every = 1;
    auto val = [&xml](const char* l, const char* p) { return xml.getValueAttribute(l, p); };
    auto str = [&xml](const char* l, const char* p) { return xml.getStringAttribute(l, p); };

    m_p = dvm_sim_params{};
    dvm_sheet_params& sp = m_p.sheet;
    sp.rho = val("constants", "density");
    sp.nu = val("constants", "nu");
    sp.dt = val("time", "dt");
    const double steps = val("time", "steps");
    m_steps = (unsigned)steps;
    sp.ux = val("flow", "ux");
    sp.uz = val("flow", "uz");
    sp.max_npv = (int)(unsigned)val("constants", "max_NumPanelVort");
    sp.cutoff_exp = val("constants", "cutoff_exp");
    m_p.seed = xml.getIntAttribute("constants", "seed");

    const std::string scheme = str("time", "scheme");
    if (scheme == "euler") m_p.scheme = DVM_EULER;
    else if (scheme == "RK3") m_p.scheme = DVM_RK3;
    else throw dvm::XMLParseException::invalidOption("time", "scheme", "euler RK3");
    const std::string cross = str("algorithms", "surface_crossing");
    if (cross == "DELETE") m_p.crossing = DVM_DELETE;
    else if (cross == "ABSORB") m_p.crossing = DVM_ABSORB;
    else if (cross == "REFLECT") m_p.crossing = DVM_REFLECT;
    else throw dvm::XMLParseException::invalidOption("algorithms", "surface_crossing",
                                                     "DELETE ABSORB REFLECT");

    // Vortex merging is optional in the input (src/DVMBase.cpp:49-60).
    double merge_threshold = 0.0;
    int merge_frequency = 0;
    try {
        merge_threshold = xml.getValueAttribute("algorithms", "merge_threshold");
        merge_frequency = xml.getIntAttribute("algorithms", "merge_frequency");
    } catch (const dvm::XMLParseException&) {
        merge_threshold = 0.0;
        merge_frequency = 0;
    }
    if (merge_frequency > 0 && merge_threshold > 0) {
        m_p.merge_threshold = merge_threshold;
        m_p.merge_frequency = merge_frequency;
        if (m_opts.screen)
            std::cout << "Vortex merging enabled: threshold = " << merge_threshold
                      << ", frequency = " << merge_frequency << std::endl;
    }

    // Remeshing is optional: no <remesh><frequency> means none.
    m_p.remesh = dvm_remesh_params{};
    if (xml.has("remesh", "frequency")) {
        dvm_remesh_params& r = m_p.remesh;
        r.frequency = xml.getIntAttribute("remesh", "frequency");
        if (r.frequency > 0) {
            r.xmin = val("remesh", "xmin");
            r.xmax = val("remesh", "xmax");
            r.zmin = val("remesh", "zmin");
            r.zmax = val("remesh", "zmax");
            r.dx = val("remesh", "dx");
            r.dz = val("remesh", "dz");
            r.sigma = val("remesh", "sigma");
            r.circ_threshold = val("remesh", "circ_threshold");
            const std::string k = xml.has("remesh", "kernel") ? str("remesh", "kernel") : "M4prime";
            if (k == "M4prime") r.kernel = 0;
            else if (k == "gaussian") r.kernel = 1;
            else if (k == "M6") r.kernel = 2;
            else throw dvm::XMLParseException::invalidOption("remesh", "kernel", "M4prime gaussian M6");
        } else {
            r.frequency = 0;
        }
    }
    m_p.prec = m_opts.precision;
    m_p.rng_mode = m_opts.rng_mode;
    // <algorithms><velocity_kernel> (optional; the reference lists it but its
    // blobs always use "direct"): "fmm" selects the fast multipole sum with
    // <fmm_expansion_order> and <fmm_theta> (defaults 10 and 0.5)
    if (xml.has("algorithms", "velocity_kernel") && str("algorithms", "velocity_kernel") == "fmm") {
        const int order = xml.has("algorithms", "fmm_expansion_order")
                              ? (int)val("algorithms", "fmm_expansion_order") : 10;
        const double theta = xml.has("algorithms", "fmm_theta") ? val("algorithms", "fmm_theta") : 0.5;
        dvm_ctx_t ctx = dvm::device_context();
        dvm::check(dvm_set_fmm(ctx, order, theta), "velocity_kernel fmm", ctx);
        m_p.prec = DVM_FMM;
    }

    // Output directory first: the blob files are created before the body is read.
    dvm::OutputSpec spec;
    spec.dir = xml.getStringAttribute("io", "output_dir", true);
    spec.stamp = timestamp;
    spec.dt = sp.dt;
    spec.steps = steps;
    spec.write_blobs = m_opts.write_blobs;

    const std::filesystem::path domain = std::filesystem::path(str("io", "input_dir")) /
                                         str("io", "domain_file");
    std::vector<double> x, z;
    read_body(domain.string(), x, z);
    if (m_opts.screen) {
        std::cout << "Successfully loaded coordinate file with " << x.size() << " points." << std::endl;
        std::cout << "A = " << *std::max_element(x.begin(), x.end()) - *std::min_element(x.begin(), x.end())
                  << " (assuming cylinder)\n";
    }
    const size_t nl = x.size() - 1;
    spec.xc.resize(nl);
    spec.zc.resize(nl);
    spec.ds.resize(nl);
    for (size_t i = 0; i < nl; ++i) {
        spec.xc[i] = 0.5 * (x[i] + x[i + 1]);
        spec.zc[i] = 0.5 * (z[i] + z[i + 1]);
        const double dx = x[i + 1] - x[i], dz = z[i + 1] - z[i];
        spec.ds[i] = std::sqrt(std::pow(dx, 2) + std::pow(dz, 2));
    }
    spec.probe_x = xml.getList("probe", "x");
    spec.probe_z = xml.getList("probe", "z");
    if (spec.probe_x.size() != spec.probe_z.size()) throw dvm::SizeMismatchException("probe x and z lists");

    init(x, z);
    m_writer = std::make_unique<dvm::StepWriter>(dvm::device_context(), m_sim, std::move(spec));
}

DVMBase::~DVMBase()
{
    try {
        m_writer.reset();
    } catch (...) {
    }
    if (m_sim) (void)dvm_sim_destroy(m_sim);
}

void DVMBase::solve() { solve_until(m_steps); }

void DVMBase::solve_until(unsigned last)
{
    dvm_ctx_t ctx = dvm::device_context();
    if (last > m_steps) last = m_steps;
    if (!m_writer) {
        const unsigned done = current_step();
        if (done < last) dvm::check(dvm_sim_run(m_sim, (int)(last - done)), "DVMBase::solve", ctx);
        return;
    }
    for (unsigned j = current_step() + 1; j <= last; ++j) {
        dvm::check(dvm_sim_step(m_sim), "DVMBase::solve", ctx);
        int64_t merged = 0;
        dvm::check(dvm_sim_merged(m_sim, &merged), "DVMBase::solve", ctx);
        if (merged > 0 && m_opts.screen) std::cout << "Merged " << merged << " vortex pairs" << std::endl;
        const double t = j * m_p.sheet.dt;
        if (j % m_opts.output_every == 0) m_writer->push(j, t);
        if (m_opts.screen) {
            std::cout << "Simulation time          = " << t << "\tStep " << j << "/" << m_steps << std::endl;
            std::cout << "Number of vortex blobs   = " << num_blobs() << std::endl;
        }
    }
    m_writer->flush();
}

void DVMBase::save_snapshot(const std::string& file) const
{
    dvm_ctx_t ctx = dvm::device_context();
    size_t bytes = 0;
    dvm::check(dvm_sim_snapshot_size(m_sim, &bytes), "snapshot", ctx);
    std::vector<char> buf(bytes);
    dvm::check(dvm_sim_snapshot(m_sim, buf.data(), bytes), "snapshot", ctx);
    std::ofstream out(file, std::ios::binary | std::ios::trunc);
    if (!out) throw dvm::FileIOException::cannotOpen(file);
    out.write(buf.data(), (std::streamsize)bytes);
    if (!out) throw dvm::FileIOException::cannotWrite(file);
}

void DVMBase::load_snapshot(const std::string& file)
{
    std::ifstream in(file, std::ios::binary);
    if (!in) throw dvm::FileIOException::cannotRead(file);
    std::vector<char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    dvm_ctx_t ctx = dvm::device_context();
    dvm::check(dvm_sim_restore(m_sim, buf.data(), buf.size()), "restore", ctx);
}

void DVMBase::step() { dvm::check(dvm_sim_step(m_sim), "DVMBase::step", dvm::device_context()); }

unsigned DVMBase::current_step() const
{
    int s = 0;
    dvm::check(dvm_sim_step_index(m_sim, &s), "step_index", dvm::device_context());
    return (unsigned)s;
}

double DVMBase::time() const { return current_step() * m_p.sheet.dt; }

unsigned DVMBase::num_blobs() const
{
    int64_t n = 0;
    dvm::check(dvm_sim_size(m_sim, &n), "size", dvm::device_context());
    return (unsigned)n;
}

VortexBlobs DVMBase::blobs() const
{
    const unsigned n = num_blobs();
    VortexBlobs v(n);
    if (n == 0) return v;
    dvm::check(dvm_sim_get_state(m_sim, v.m_ID.data(), v.m_x.memptr(), v.m_z.memptr(),
                                 v.m_circ.memptr(), v.m_sigma.memptr(), v.m_u.memptr(),
                                 v.m_w.memptr(), v.m_uvs.memptr(), v.m_wvs.memptr()),
               "blobs", dvm::device_context());
    return v;
}

std::tuple<double, double, double, double> DVMBase::loads() const
{
    double o[4];
    dvm::check(dvm_sim_loads(m_sim, o), "loads", dvm::device_context());
    return {o[0], o[1], o[2], o[3]};
}
