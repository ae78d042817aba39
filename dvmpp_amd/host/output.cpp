// StepWriter: asynchronous per-step output in the reference's file formats.
#include <algorithm>
#include <charconv>
#include <cstring>

#include "Exceptions.hpp"
#include "StepWriter.hpp"

namespace dvm {

namespace {

inline char* put_g(char* p, double v)
{
    // std::to_chars with a precision prints as printf("%.6g") does, which is
    // how an ostream with default flags and precision 6 prints a double.
    return std::to_chars(p, p + 40, v, std::chars_format::general, 6).ptr;
}

void open_new(std::ofstream& f, const std::string& path)
{
    f.open(path, std::ios::out | std::ios::trunc | std::ios::binary);
    if (!f) throw FileIOException::cannotOpen(path);
}

void check_written(std::ofstream& f, const std::string& path)
{
    if (!f) throw FileIOException::cannotWrite(path);
}

}  // namespace

std::string format_g(double v)
{
    char b[48];
    return std::string(b, put_g(b, v));
}

StepWriter::StepWriter(dvm_ctx_t ctx, dvm_sim_t sim, OutputSpec spec, int frames)
    : m_ctx(ctx), m_sim(sim), m_spec(std::move(spec))
{
    const std::string base = m_spec.dir + m_spec.stamp;
    m_vortex_path = base + "_vortex.dat";
    m_num_path = base + "_vortex_num.dat";
    m_gamma_path = base + "_gamma.dat";
    m_loads_path = base + "_loads.dat";
    m_probe_path = base + "_probe.dat";
    const std::string dt = format_g(m_spec.dt), steps = format_g(m_spec.steps);

    // Headers, in the order the reference's constructors write them.
    open_new(m_vortex, m_vortex_path);
    m_vortex << "0 # Number of nodes\n" << dt << " # Time step\n" << steps << " # Steps\n"
             << "Time [s]\tx-position [m]\tz-position [m]\tcirculation\n";
    open_new(m_num, m_num_path);
    m_num << "Time [s]\tNumber of vortices\n";
    open_new(m_gamma, m_gamma_path);
    m_gamma << m_spec.xc.size() << " # Number of collocation points\n" << dt << " # Time Step\n"
            << steps << " # Steps\n" << "Time [s] Gamma - Vortex sheet strength\n";
    open_new(m_loads, m_loads_path);
    m_loads << "Time [s]\tF_x [-]\tF_z [s]\n";
    open_new(m_probe, m_probe_path);
    m_probe << dt << " # Time step\n" << steps << " # Steps\n" << "Time [s]\tu [m/s]\tw [m/s]\n";
    for (auto* f : {&m_vortex, &m_num, &m_gamma, &m_loads, &m_probe}) f->flush();

    const size_t np = m_spec.probe_x.size();
    if (m_spec.probe_z.size() != np) throw SizeMismatchException("probe x and z lists");
    if (np) {
        check(dvm_malloc(ctx, sizeof(double) * 6 * np, &m_probe_dev), "probe buffers", ctx);
        double* d = static_cast<double*>(m_probe_dev);
        check(dvm_memcpy_h2d(ctx, d, m_spec.probe_x.data(), sizeof(double) * np), "probe x", ctx);
        check(dvm_memcpy_h2d(ctx, d + np, m_spec.probe_z.data(), sizeof(double) * np), "probe z", ctx);
    }

    const unsigned hw = std::thread::hardware_concurrency();
    m_fmt_threads = std::max(1u, std::min(8u, hw > 2 ? hw - 2 : 1u));
    m_frames.resize(std::max(frames, 2));
    const size_t nl = m_spec.xc.size();
    for (size_t i = 0; i < m_frames.size(); ++i) {
        Frame& f = m_frames[i];
        check(dvm_event_create(ctx, &f.ev), "event", ctx);
        void* p = nullptr;
        check(dvm_host_alloc(sizeof(double) * (nl + 4), &p), "pinned frame", ctx);
        f.gamma = static_cast<double*>(p);
        f.loads = f.gamma + nl;
        f.pu.resize(np);
        f.pw.resize(np);
        m_free.push_back((int)i);
    }
    m_thread = std::thread(&StepWriter::writer_main, this);
}

StepWriter::~StepWriter()
{
    {
        std::unique_lock<std::mutex> lk(m_mu);
        m_cv.wait(lk, [&] { return (m_full.empty() && m_busy == 0) || m_err; });
        m_stop = true;
    }
    m_cv.notify_all();
    if (m_thread.joinable()) m_thread.join();
    for (Frame& f : m_frames) release(f);
    if (m_probe_dev) (void)dvm_free(m_ctx, m_probe_dev);
}

void StepWriter::release(Frame& f)
{
    if (f.ev) (void)dvm_event_wait(f.ev);  // no copy may still target the buffers
    for (double* p : {f.x, f.z, f.c, f.gamma}) (void)dvm_host_free(p);
    f.x = f.z = f.c = f.gamma = f.loads = nullptr;
    if (f.ev) (void)dvm_event_destroy(f.ev);
    f.ev = nullptr;
}

void StepWriter::rethrow_locked()
{
    if (m_err) std::rethrow_exception(m_err);
}

void StepWriter::push(unsigned step, double time)
{
    int idx;
    {
        std::unique_lock<std::mutex> lk(m_mu);
        m_cv.wait(lk, [&] { return !m_free.empty() || m_err; });
        rethrow_locked();
        idx = m_free.front();
        m_free.pop_front();
    }
    Frame& f = m_frames[idx];
    f.step = step;
    f.time = time;
    int64_t n = 0;
    check(dvm_sim_size(m_sim, &n), "StepWriter size", m_ctx);
    if (m_spec.write_blobs && n > f.cap) {
        const int64_t cap = n + n / 4 + 1024;
        for (double** p : {&f.x, &f.z, &f.c}) {
            (void)dvm_host_free(*p);
            *p = nullptr;
        }
        f.cap = 0;
        for (double** p : {&f.x, &f.z, &f.c}) {
            void* q = nullptr;
            check(dvm_host_alloc(sizeof(double) * cap, &q), "pinned frame", m_ctx);
            *p = static_cast<double*>(q);
        }
        f.cap = cap;
    }
    const size_t np = m_spec.probe_x.size();
    if (np) {
        double* d = static_cast<double*>(m_probe_dev);
        check(dvm_sim_probe_quirk(m_sim, DVM_F64, (int64_t)np, d, d + np, d + 2 * np, d + 3 * np,
                                  d + 4 * np, d + 5 * np),
              "probe velocities", m_ctx);
        check(dvm_memcpy_d2h(m_ctx, f.pu.data(), d + 2 * np, sizeof(double) * np), "probe u", m_ctx);
        check(dvm_memcpy_d2h(m_ctx, f.pw.data(), d + 3 * np, sizeof(double) * np), "probe w", m_ctx);
    }
    const bool b = m_spec.write_blobs;
    check(dvm_sim_copy_out(m_sim, b ? f.cap : 0, b ? f.x : nullptr, b ? f.z : nullptr,
                           b ? f.c : nullptr, f.gamma, f.loads, &f.n),
          "StepWriter copy", m_ctx);
    check(dvm_event_record(m_ctx, f.ev), "StepWriter event", m_ctx);
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_full.push_back(idx);
    }
    m_cv.notify_all();
}

void StepWriter::flush()
{
    std::unique_lock<std::mutex> lk(m_mu);
    m_cv.wait(lk, [&] { return (m_full.empty() && m_busy == 0) || m_err; });
    rethrow_locked();
}

void StepWriter::writer_main()
{
    for (;;) {
        int idx;
        {
            std::unique_lock<std::mutex> lk(m_mu);
            m_cv.wait(lk, [&] { return !m_full.empty() || m_stop; });
            if (m_full.empty()) return;
            idx = m_full.front();
            m_full.pop_front();
            ++m_busy;
        }
        try {
            write_frame(m_frames[idx]);
        } catch (...) {
            std::lock_guard<std::mutex> lk(m_mu);
            if (!m_err) m_err = std::current_exception();
        }
        {
            std::lock_guard<std::mutex> lk(m_mu);
            --m_busy;
            m_free.push_back(idx);
        }
        m_cv.notify_all();
    }
}

void StepWriter::write_frame(Frame& f)
{
    if (dvm_event_wait(f.ev) != DVM_OK) throw DVMException("StepWriter: waiting for the copy failed");
    char tb[48];
    const std::string t(tb, put_g(tb, f.time));

    if (m_spec.write_blobs && f.n > 0) {
        // rows "time\tx\tz\tcirc\n", formatted in parallel chunks, written in order
        const int64_t n = std::min(f.n, f.cap);
        const unsigned k = n >= 32768 ? m_fmt_threads : 1u;
        std::vector<std::string> out(k);
        auto fmt = [&](unsigned part) {
            const int64_t b = n * part / k, e = n * (part + 1) / k;
            std::string& s = out[part];
            s.resize((size_t)(e - b) * (t.size() + 3 * 16 + 4));
            char* p = s.data();
            for (int64_t i = b; i < e; ++i) {
                std::memcpy(p, t.data(), t.size());
                p += t.size();
                *p++ = '\t';
                p = put_g(p, f.x[i]);
                *p++ = '\t';
                p = put_g(p, f.z[i]);
                *p++ = '\t';
                p = put_g(p, f.c[i]);
                *p++ = '\n';
            }
            s.resize((size_t)(p - s.data()));
        };
        if (k == 1) {
            fmt(0);
        } else {
            std::vector<std::thread> pool;
            for (unsigned i = 1; i < k; ++i) pool.emplace_back(fmt, i);
            fmt(0);
            for (auto& th : pool) th.join();
        }
        for (const std::string& s : out) m_vortex.write(s.data(), (std::streamsize)s.size());
        m_vortex.flush();
        check_written(m_vortex, m_vortex_path);
    }
    m_num << f.step << '\t' << f.n << '\n';
    m_num.flush();
    check_written(m_num, m_num_path);

    std::string g;
    const size_t nl = m_spec.xc.size();
    g.reserve(nl * (t.size() + 4 * 16 + 6));
    char b[48];
    for (size_t i = 0; i < nl; ++i) {
        g += t;
        g += "\t ";
        g.append(b, put_g(b, m_spec.xc[i]));
        g += '\t';
        g.append(b, put_g(b, m_spec.zc[i]));
        g += '\t';
        g.append(b, put_g(b, f.gamma[i]));
        g += '\t';
        g.append(b, put_g(b, m_spec.ds[i]));
        g += '\n';
    }
    m_gamma << g;
    m_gamma.flush();
    check_written(m_gamma, m_gamma_path);

    m_loads << t << '\t';
    m_loads.write(b, put_g(b, f.loads[0]) - b);
    m_loads << '\t';
    m_loads.write(b, put_g(b, f.loads[1]) - b);
    m_loads << '\n';
    m_loads.flush();
    check_written(m_loads, m_loads_path);

    std::string pr;
    for (size_t i = 0; i < f.pu.size(); ++i) {
        pr += t;
        pr += '\t';
        pr.append(b, put_g(b, f.pu[i]));
        pr += '\t';
        pr.append(b, put_g(b, f.pw[i]));
        pr += '\n';
    }
    m_probe << pr;
    m_probe.flush();
    check_written(m_probe, m_probe_path);
}

}  // namespace dvm
