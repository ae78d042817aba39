// CPU tests of the input-file reader and the output number format
// (dvmpp_amd/host/xml.cpp, output.cpp): the cases of the reference's
// tests/unit/test_XmlHandler.cpp, the reader's error paths, and that the
// writer prints doubles exactly as an std::ostream with default flags does.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "Exceptions.hpp"
#include "StepWriter.hpp"
#include "XmlHandler.hpp"

namespace fs = std::filesystem;

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
std::string g_dir;

std::string write_file(const std::string& name, const std::string& text)
{
    const std::string p = g_dir + "/" + name;
    std::ofstream(p) << text;
    return p;
}

const char* kValid = R"(<?xml version="1.0"?>
<DVMpp>
    <algorithms>
        <surface_crossing string="REFLECT" />
    </algorithms>
    <constants>
        <density val="1.0" />
        <nu val="0.001" />
        <max_NumPanelVort val="3" />
        <cutoff_exp val="0.675" />
        <seed val="42" />
    </constants>
    <io>
        <input_dir string="./input/" />
        <output_dir string="./output/" />
        <domain_file string="body.txt" />
    </io>
    <flow>
        <ux val="3.0" />
        <uz val="0.0" />
    </flow>
    <probe>
        <x val="20" />
        <z val="1" />
    </probe>
    <time>
        <scheme string="euler" />
        <dt val="0.005" />
        <steps val="100" />
    </time>
</DVMpp>
)";
}  // namespace

#define TEST(name)                      \
    static void name();                 \
    static Reg reg_##name(#name, name); \
    static void name()
#define EXPECT(cond)                                                       \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)
#define EXPECT_THROW(stmt, exc)  \
    do {                         \
        bool thrown_ = false;    \
        try {                    \
            stmt;                \
        } catch (const exc&) {   \
            thrown_ = true;      \
        }                        \
        EXPECT(thrown_);         \
    } while (0)

// ---- tests/unit/test_XmlHandler.cpp:67-150

TEST(DefaultConstructorWorks) { XmlHandler xml; }

TEST(LoadValidXmlFile) { XmlHandler xml(write_file("valid.xml", kValid)); }

TEST(LoadNonExistentFileThrows)
{
    EXPECT_THROW(XmlHandler("/nonexistent/path/file.xml"), dvm::XMLParseException);
}

TEST(GetValueAttributeReturnsCorrectValue)
{
    XmlHandler xml(write_file("valid.xml", kValid));
    EXPECT(xml.getValueAttribute("constants", "density") == 1.0);
    EXPECT(xml.getValueAttribute("constants", "nu") == 0.001);
    EXPECT(xml.getValueAttribute("time", "dt") == 0.005);
}

TEST(GetStringAttributeReturnsCorrectValue)
{
    XmlHandler xml(write_file("valid.xml", kValid));
    EXPECT(xml.getStringAttribute("algorithms", "surface_crossing") == "REFLECT");
    EXPECT(xml.getStringAttribute("time", "scheme") == "euler");
}

TEST(GetIntAttributeReturnsCorrectValue)
{
    XmlHandler xml(write_file("valid.xml", kValid));
    EXPECT(xml.getIntAttribute("constants", "seed") == 42);
    EXPECT(xml.getIntAttribute("time", "steps") == 100);
}

TEST(GetStringAttributeWithCheckSepAddsTrailingSlash)
{
    XmlHandler xml(write_file("valid.xml", kValid));
    EXPECT(xml.getStringAttribute("io", "input_dir", true).back() == '/');
    EXPECT(xml.getStringAttribute("io", "domain_file", true) == "body.txt/");
}

TEST(MissingParameterThrows)
{
    XmlHandler xml(write_file("incomplete.xml",
                              "<?xml version=\"1.0\"?>\n<DVMpp>\n<constants>\n<density val=\"1.0\" />\n"
                              "</constants>\n</DVMpp>\n"));
    EXPECT_THROW((void)xml.getValueAttribute("constants", "nu"), dvm::XMLParseException);
}

TEST(WriteExampleCreatesFile)
{
    XmlHandler xml;
    const std::string p = g_dir + "/example_output.xml";
    xml.writeExample(p);
    EXPECT(fs::exists(p));
    // and it parses back with every option present
    XmlHandler back(p);
    EXPECT(back.has("remesh", "circ_threshold"));
    EXPECT(back.has("algorithms", "surface_crossing"));
    EXPECT(back.has("time", "steps"));
}

// ---- option checks (src/XmlHandler.cpp:143-183)

TEST(InvalidStringOptionThrows)
{
    std::string t = kValid;
    t.replace(t.find("REFLECT"), 7, "BOUNCE");
    XmlHandler xml(write_file("bad_option.xml", t));
    EXPECT_THROW((void)xml.getStringAttribute("algorithms", "surface_crossing"), dvm::XMLParseException);
}

TEST(NonPositiveValueThrows)
{
    std::string t = kValid;
    t.replace(t.find("0.005"), 5, "-0.01");
    XmlHandler xml(write_file("bad_value.xml", t));
    EXPECT_THROW((void)xml.getValueAttribute("time", "dt"), dvm::XMLParseException);
    // options without a requirement take any value
    EXPECT(xml.getValueAttribute("flow", "uz") == 0.0);
}

TEST(UnknownOptionThrows)
{
    XmlHandler xml(write_file("valid.xml", kValid));
    EXPECT_THROW((void)xml.getValueAttribute("constants", "max_Gamma"), dvm::XMLParseException);
    // an entry that is present but not in the option map
    XmlHandler extra(write_file("extra.xml",
                                "<DVMpp><constants><max_Gamma val=\"1\"/></constants></DVMpp>"));
    EXPECT_THROW((void)extra.getValueAttribute("constants", "max_Gamma"), dvm::XMLParseException);
}

TEST(ListsAndBools)
{
    XmlHandler xml(write_file("lists.xml",
                              "<DVMpp><probe><x string=\"20 1.5,-3\" /><z string=\"1, 0.5 2e-1\" />"
                              "</probe><io><input_dir string=\"true\"/></io></DVMpp>"));
    const std::vector<double> x = xml.getList("probe", "x"), z = xml.getList("probe", "z");
    EXPECT(x.size() == 3 && x[0] == 20 && x[1] == 1.5 && x[2] == -3);
    EXPECT(z.size() == 3 && z[0] == 1 && z[1] == 0.5 && z[2] == 0.2);
    EXPECT(xml.getBoolAttribute("io", "input_dir"));
    // the reference's probe lists read the "string" attribute: "val" gives none
    XmlHandler v(write_file("valid.xml", kValid));
    EXPECT(v.getList("probe", "x").empty());
}

TEST(IntegerPrefixAndMissingAttribute)
{
    XmlHandler xml(write_file("ints.xml",
                              "<DVMpp><constants><seed val=\"-17.9\"/></constants>"
                              "<time><steps val=\"0x10\"/><dt/></time></DVMpp>"));
    EXPECT(xml.getIntAttribute("constants", "seed") == -17);
    EXPECT(xml.getIntAttribute("time", "steps") == 16);
    // an absent "val" reads as 0, which dt's ">0" rejects
    EXPECT_THROW((void)xml.getValueAttribute("time", "dt"), dvm::XMLParseException);
}

TEST(CommentsEntitiesAndSave)
{
    const std::string t =
        "\xEF\xBB\xBF<?xml version=\"1.0\"?>\n<!-- leading comment -->\n<DVMpp>\n"
        "  <io><!-- inner --><output_dir string='a&amp;b &lt;c&gt; &#65;&#x42;' /></io>\n"
        "  <remesh>\n    <kernel string=\"M6\"></kernel>\n  </remesh>\n</DVMpp>\n";
    XmlHandler xml(write_file("entities.xml", t));
    EXPECT(xml.getStringAttribute("io", "output_dir") == "a&b <c> AB");
    EXPECT(xml.getStringAttribute("remesh", "kernel") == "M6");
    const std::string p = g_dir + "/saved.xml";
    xml.save(p);
    XmlHandler back(p);
    EXPECT(back.getStringAttribute("io", "output_dir") == "a&b <c> AB");
    EXPECT(back.getStringAttribute("remesh", "kernel") == "M6");
}

TEST(MalformedDocumentsThrow)
{
    const char* bad[] = {
        "<DVMpp><io></DVMpp>",                       // tags mismatch
        "<DVMpp><io><output_dir string=\"x\"></io>",  // unclosed
        "<DVMpp><io x=1/></DVMpp>",                   // unquoted value
        "<DVMpp>&bogus;</DVMpp>",                     // unknown entity
        "",                                            // no element
        "<DVMpp/><Other/>",                           // two roots
        "<!-- open comment <DVMpp/>",
    };
    int i = 0;
    for (const char* b : bad)
        EXPECT_THROW(XmlHandler(write_file("bad" + std::to_string(i++) + ".xml", b)), dvm::XMLParseException);
}

// ---- number format of the output files

TEST(FormatMatchesOstream)
{
    std::mt19937_64 g(7);
    std::vector<double> vals = {0.0, -0.0, 1.0, 0.005, 10.0, 100.0, 1e-5, 123456.0, 1234567.0,
                                0.000123456789, -2.5e-300, 1e300, 0.1 + 0.2, 3.0 / 7.0,
                                std::numeric_limits<double>::denorm_min(),
                                std::numeric_limits<double>::max(), 999999.5, 9999995.0,
                                std::numeric_limits<double>::infinity()};
    std::uniform_real_distribution<double> u(-1.0, 1.0), e(-320.0, 300.0);
    for (int i = 0; i < 200000; ++i) vals.push_back(u(g) * std::pow(10.0, e(g)));
    for (int i = 0; i < 20000; ++i) vals.push_back(std::round(u(g) * 1e7) * 0.005);
    int bad = 0;
    for (double v : vals) {
        std::ostringstream o;
        o << v;
        if (o.str() != dvm::format_g(v) && bad++ < 5)
            std::printf("  %s vs %s\n", o.str().c_str(), dvm::format_g(v).c_str());
    }
    EXPECT(bad == 0);
}

int main(int argc, char** argv)
{
    g_dir = argc > 1 ? argv[1] : "/tmp";
    fs::create_directories(g_dir);
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
