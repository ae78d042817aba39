// The dvmpp command: the reference's front end (src/main.cpp:20-115) — the same
// options (-h/--help, -f/--input-file or a positional input file,
// --example-xml [file]), screen lines, time stamp, saved copy of the input and
// exit codes — over the device-resident time loop.  Options the reference does
// not have select this library's choices:
//   --precision f64|f32c     arithmetic of the pair sums (default f64)
//   --rng reference|device   random stream (default reference: std::mt19937)
//   --output-every N         write every N-th step (default 1, as the reference)
//   --no-blob-output         keep <stamp>_vortex.dat to its header
//   --snapshot-every N       save <stamp>_snapshot.bin every N steps and at the end
//   --restore FILE           continue from a snapshot of the same case
//   --stamp S                use S instead of the local-time stamp
//   --quiet                  no per-step screen lines
#include <chrono>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "DVMBase.hpp"
#include "Exceptions.hpp"
#include "XmlHandler.hpp"

namespace {

const char* kHelp =
    "DVMpp commandline options:\n"
    "  -h [ --help ]             prints this help message\n"
    "  -f [ --input-file ] arg   input xml file\n"
    "  --example-xml             prints an example xml to stdout or file\n"
    "  --precision arg           pair-sum arithmetic: f64 (default) or f32c\n"
    "  --rng arg                 random stream: reference (default) or device\n"
    "  --output-every arg        write every N-th step (default 1)\n"
    "  --no-blob-output          write no per-blob rows\n"
    "  --snapshot-every arg      save a snapshot every N steps and at the end\n"
    "  --restore arg             continue from a snapshot file\n"
    "  --stamp arg               output file prefix (default: local time)\n"
    "  --quiet                   no per-step screen lines\n";

struct Args {
    bool any = false, help = false, example = false, quiet = false, no_blobs = false;
    std::string input, precision = "f64", rng = "reference", restore, stamp;
    unsigned every = 1, snapshot_every = 0;
};

unsigned to_count(const std::string& opt, const std::string& v)
{
    char* end = nullptr;
    const long n = std::strtol(v.c_str(), &end, 10);
    if (v.empty() || *end != '\0' || n < 0 || n > 1000000000L)
        throw std::invalid_argument("the argument ('" + v + "') for option '" + opt + "' is invalid");
    return (unsigned)n;
}

Args parse(int argc, char** argv)
{
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i], v;
        const bool has_eq = k.rfind("--", 0) == 0 && k.find('=') != std::string::npos;
        if (has_eq) {
            v = k.substr(k.find('=') + 1);
            k = k.substr(0, k.find('='));
        }
        auto value = [&]() -> std::string {
            if (has_eq) return v;
            if (i + 1 >= argc) throw std::invalid_argument("the required argument for option '" + k + "' is missing");
            return argv[++i];
        };
        a.any = true;
        if (k == "-h" || k == "--help") a.help = true;
        else if (k == "--example-xml") a.example = true;
        else if (k == "-f" || k == "--input-file") a.input = value();
        else if (k == "--precision") a.precision = value();
        else if (k == "--rng") a.rng = value();
        else if (k == "--output-every") a.every = to_count(k, value());
        else if (k == "--snapshot-every") a.snapshot_every = to_count(k, value());
        else if (k == "--restore") a.restore = value();
        else if (k == "--stamp") a.stamp = value();
        else if (k == "--quiet") a.quiet = true;
        else if (k == "--no-blob-output") a.no_blobs = true;
        else if (!k.empty() && k[0] == '-' && k.size() > 1) throw std::invalid_argument("unrecognised option '" + k + "'");
        else if (a.input.empty()) a.input = k;
        else throw std::invalid_argument("too many positional options have been specified on the command line");
    }
    return a;
}

}  // namespace

int main(int argc, char* argv[])
{
    try {
        const Args a = parse(argc, argv);
        if (!a.any || a.help) {
            std::cout << kHelp << "\n";
            return a.help ? 0 : 1;
        }
        if (a.example) {
            XmlHandler xml;
            xml.writeExample(a.input.empty() ? "-" : a.input);
            return 0;
        }
        if (a.input.empty()) throw std::invalid_argument("no input file given");

        RunOptions opts;
        if (a.precision == "f64") opts.precision = DVM_F64;
        else if (a.precision == "f32c") opts.precision = DVM_F32C;
        else throw std::invalid_argument("--precision takes f64 or f32c");
        if (a.rng == "reference") opts.rng_mode = DVM_RNG_REFERENCE;
        else if (a.rng == "device") opts.rng_mode = DVM_RNG_DEVICE;
        else throw std::invalid_argument("--rng takes reference or device");
        opts.output_every = a.every ? a.every : 1;
        opts.write_blobs = !a.no_blobs;
        opts.screen = !a.quiet;

        const std::string rule(62, '=');
        std::cout << rule << std::endl << "Running DVMpp - 2D Discrete Vortex Method" << std::endl
                  << rule << std::endl;

        XmlHandler xml(a.input);
        std::string stamp = a.stamp;
        if (stamp.empty()) {
            const std::time_t now = std::time(nullptr);
            const std::tm local = *std::localtime(&now);
            std::ostringstream oss;
            oss << std::put_time(&local, "%Y_%m_%d_%H_%M_%S");
            stamp = oss.str();
        }
        std::cout << "File timestamp is " << stamp << std::endl;

        DVMBase dvm(xml, stamp, opts);
        std::string outdir = xml.getStringAttribute("io", "output_dir");
        if (outdir.empty() || outdir.back() != '/') outdir += '/';
        xml.save((std::filesystem::path(outdir) / (stamp + "_xml_in.xml")).string());
        if (!a.restore.empty()) {
            dvm.load_snapshot(a.restore);
            std::cout << "Restored " << a.restore << " at step " << dvm.current_step() << std::endl;
        }

        const auto start = std::chrono::system_clock::now();
        const std::string snap = outdir + stamp + "_snapshot.bin";
        if (a.snapshot_every == 0) {
            dvm.solve();
        } else {
            // the time loop in segments that end on snapshot steps
            while (dvm.current_step() < dvm.steps()) {
                const unsigned next = (dvm.current_step() / a.snapshot_every + 1) * a.snapshot_every;
                dvm.solve_until(next < dvm.steps() ? next : dvm.steps());
                dvm.save_snapshot(snap);
            }
        }
        std::cout << "\nTimestamp : " << stamp << "\n";
        const std::chrono::duration<double> elapsed = std::chrono::system_clock::now() - start;
        std::cout << "Runtime : " << elapsed.count() << "s\n";
        std::cout << "Completed DVM computations successfully" << std::endl;
    } catch (const dvm::DVMException& e) {
        std::cerr << "DVMpp Error: " << e.what() << std::endl;
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "Exception: " << e.what() << std::endl;
        return 1;
    } catch (...) {
        std::cerr << "Unknown exception occurred" << std::endl;
        return 1;
    }
    return 0;
}
