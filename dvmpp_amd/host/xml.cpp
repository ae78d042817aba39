// XmlHandler: option map, getters and checks with the reference's semantics
// (src/XmlHandler.cpp:8-258) over this library's own XML reader and writer.
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

#include "Exceptions.hpp"
#include "XmlHandler.hpp"

// ---------------------------------------------------------------- nodes

const XmlNode* XmlNode::child(const std::string& n) const
{
    for (const XmlNode& c : children)
        if (c.name == n) return &c;
    return nullptr;
}

XmlNode* XmlNode::child(const std::string& n)
{
    for (XmlNode& c : children)
        if (c.name == n) return &c;
    return nullptr;
}

const std::string* XmlNode::attribute(const std::string& n) const
{
    for (const auto& a : attrs)
        if (a.first == n) return &a.second;
    return nullptr;
}

// ---------------------------------------------------------------- reader

namespace {

class Reader {
public:
    Reader(const std::string& t, const std::string& src) : s_(t), src_(src) {}

    XmlNode document()
    {
        XmlNode doc;
        if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0) p_ = 3;
        bool root_seen = false;
        for (;;) {
            skip_ws();
            if (p_ >= s_.size()) break;
            if (!starts("<")) fail("Text outside the document element");
            if (skip_markup()) continue;
            if (root_seen) fail("More than one document element");
            doc.children.push_back(element());
            root_seen = true;
        }
        if (!root_seen) fail("No document element found");
        return doc;
    }

private:
    const std::string& s_;
    const std::string& src_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string& why) const
    {
        size_t line = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); ++i) line += s_[i] == '\n';
        throw dvm::XMLParseException::cannotLoad(src_, why + " (line " + std::to_string(line) + ")");
    }

    bool starts(const char* t) const { return s_.compare(p_, std::char_traits<char>::length(t), t) == 0; }

    void skip_ws()
    {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }

    void skip_past(const char* end, const char* what)
    {
        const size_t e = s_.find(end, p_);
        if (e == std::string::npos) fail(std::string("Unterminated ") + what);
        p_ = e + std::char_traits<char>::length(end);
    }

    // Comments, declarations and processing instructions; true if one was skipped.
    bool skip_markup()
    {
        if (starts("<!--")) {
            skip_past("-->", "comment");
            return true;
        }
        if (starts("<?")) {
            skip_past("?>", "declaration");
            return true;
        }
        if (starts("<!DOCTYPE")) {
            int depth = 0;
            for (; p_ < s_.size(); ++p_) {
                if (s_[p_] == '[') ++depth;
                if (s_[p_] == ']') --depth;
                if (s_[p_] == '>' && depth == 0) {
                    ++p_;
                    return true;
                }
            }
            fail("Unterminated document type declaration");
        }
        return false;
    }

    static bool name_char(char c)
    {
        return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':' ||
               (unsigned char)c >= 0x80;
    }

    std::string name()
    {
        const size_t b = p_;
        while (p_ < s_.size() && name_char(s_[p_])) ++p_;
        if (p_ == b) fail("Invalid element or attribute name");
        return s_.substr(b, p_ - b);
    }

    std::string decode(const std::string& raw)
    {
        std::string out;
        out.reserve(raw.size());
        for (size_t i = 0; i < raw.size(); ++i) {
            if (raw[i] != '&') {
                out += raw[i];
                continue;
            }
            const size_t e = raw.find(';', i);
            if (e == std::string::npos) fail("Unterminated character reference");
            const std::string ent = raw.substr(i + 1, e - i - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (ent.size() > 1 && ent[0] == '#') {
                const bool hex = ent[1] == 'x' || ent[1] == 'X';
                char* endp = nullptr;
                const long cp = std::strtol(ent.c_str() + (hex ? 2 : 1), &endp, hex ? 16 : 10);
                if (*endp != '\0' || cp <= 0 || cp > 0x10FFFF) fail("Invalid character reference");
                if (cp < 0x80) {
                    out += (char)cp;
                } else if (cp < 0x800) {
                    out += (char)(0xC0 | (cp >> 6));
                    out += (char)(0x80 | (cp & 0x3F));
                } else if (cp < 0x10000) {
                    out += (char)(0xE0 | (cp >> 12));
                    out += (char)(0x80 | ((cp >> 6) & 0x3F));
                    out += (char)(0x80 | (cp & 0x3F));
                } else {
                    out += (char)(0xF0 | (cp >> 18));
                    out += (char)(0x80 | ((cp >> 12) & 0x3F));
                    out += (char)(0x80 | ((cp >> 6) & 0x3F));
                    out += (char)(0x80 | (cp & 0x3F));
                }
            } else {
                fail("Unknown entity &" + ent + ";");
            }
            i = e;
        }
        return out;
    }

    XmlNode element()
    {
        ++p_;  // '<'
        XmlNode n;
        n.name = name();
        for (;;) {
            skip_ws();
            if (p_ >= s_.size()) fail("Unterminated start tag <" + n.name + ">");
            if (starts("/>")) {
                p_ += 2;
                return n;
            }
            if (s_[p_] == '>') {
                ++p_;
                break;
            }
            std::string an = name();
            skip_ws();
            if (p_ >= s_.size() || s_[p_] != '=') fail("Attribute " + an + " has no value");
            ++p_;
            skip_ws();
            if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) fail("Attribute value is not quoted");
            const char q = s_[p_++];
            const size_t e = s_.find(q, p_);
            if (e == std::string::npos) fail("Unterminated attribute value");
            n.attrs.emplace_back(an, decode(s_.substr(p_, e - p_)));
            p_ = e + 1;
        }
        // content
        std::string text;
        for (;;) {
            if (p_ >= s_.size()) fail("Start-end tags mismatch: <" + n.name + "> is not closed");
            if (s_[p_] != '<') {
                const size_t e = s_.find('<', p_);
                text += s_.substr(p_, (e == std::string::npos ? s_.size() : e) - p_);
                p_ = e == std::string::npos ? s_.size() : e;
                continue;
            }
            if (starts("<![CDATA[")) {
                p_ += 9;
                const size_t e = s_.find("]]>", p_);
                if (e == std::string::npos) fail("Unterminated CDATA section");
                n.text += s_.substr(p_, e - p_);
                p_ = e + 3;
                continue;
            }
            if (skip_markup()) continue;
            if (starts("</")) {
                p_ += 2;
                const std::string close = name();
                skip_ws();
                if (close != n.name || p_ >= s_.size() || s_[p_] != '>')
                    fail("Start-end tags mismatch: <" + n.name + "> closed by </" + close + ">");
                ++p_;
                break;
            }
            n.text += trimmed(decode(text));
            text.clear();
            n.children.push_back(element());
        }
        n.text += trimmed(decode(text));
        return n;
    }

    static std::string trimmed(const std::string& t)
    {
        size_t b = 0, e = t.size();
        while (b < e && std::isspace((unsigned char)t[b])) ++b;
        while (e > b && std::isspace((unsigned char)t[e - 1])) --e;
        return t.substr(b, e - b);
    }
};

std::string escape(const std::string& v, bool attr)
{
    std::string o;
    for (char c : v) {
        switch (c) {
        case '&': o += "&amp;"; break;
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        case '"':
            if (attr) {
                o += "&quot;";
                break;
            }
            o += c;
            break;
        default: o += c;
        }
    }
    return o;
}

void write_node(std::ostringstream& o, const XmlNode& n, int depth)
{
    const std::string ind(depth, '\t');
    o << ind << '<' << n.name;
    for (const auto& a : n.attrs) o << ' ' << a.first << "=\"" << escape(a.second, true) << '"';
    if (n.children.empty() && n.text.empty()) {
        o << " />\n";
        return;
    }
    o << '>';
    if (n.children.empty()) {
        o << escape(n.text, false) << "</" << n.name << ">\n";
        return;
    }
    o << '\n';
    if (!n.text.empty()) o << ind << '\t' << escape(n.text, false) << '\n';
    for (const XmlNode& c : n.children) write_node(o, c, depth + 1);
    o << ind << "</" << n.name << ">\n";
}

// The integer prefix of a "val" attribute: optional sign, decimal or 0x
// digits; 0 when there is none; clamped to int.
int parse_int(const std::string& v)
{
    const char* p = v.c_str();
    while (std::isspace((unsigned char)*p)) ++p;
    bool neg = false;
    if (*p == '-' || *p == '+') neg = *p++ == '-';
    int base = 10;
    if (p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) {
        base = 16;
        p += 2;
    }
    long long acc = 0;
    for (; *p; ++p) {
        int d;
        if (*p >= '0' && *p <= '9') d = *p - '0';
        else if (base == 16 && std::isxdigit((unsigned char)*p)) d = 10 + (std::tolower(*p) - 'a');
        else break;
        acc = acc * base + d;
        if (acc > (long long)INT_MAX + 1) {
            acc = (long long)INT_MAX + 1;
            break;
        }
    }
    if (neg) return acc > (long long)INT_MAX ? INT_MIN : -(int)acc;
    return acc > INT_MAX ? INT_MAX : (int)acc;
}

}  // namespace

XmlNode XmlHandler::parse(const std::string& text, const std::string& source)
{
    return Reader(text, source).document();
}

std::string XmlHandler::serialise(const XmlNode& root)
{
    std::ostringstream o;
    o << "<?xml version=\"1.0\"?>\n";
    write_node(o, root, 0);
    return o.str();
}

// ---------------------------------------------------------------- handler

XmlHandler::XmlHandler() { buildOptionMap(); }

XmlHandler::XmlHandler(std::string xml_file)
{
    std::ifstream in(xml_file, std::ios::binary);
    if (!in) throw dvm::XMLParseException::cannotLoad(xml_file, "File was not found");
    std::ostringstream ss;
    ss << in.rdbuf();
    m_doc = parse(ss.str(), xml_file);
    std::cout << "Loaded xml input: " << xml_file << "\n";
    buildOptionMap();
}

std::string XmlHandler::form_key(const char* level, const char* parameter)
{
    return std::string(level) + "_" + parameter;
}

bool XmlHandler::has(const char* level, const char* parameter) const
{
    const XmlNode* root = m_doc.child("DVMpp");
    const XmlNode* l = root ? root->child(level) : nullptr;
    return l && l->child(parameter);
}

const XmlNode& XmlHandler::getNode(const char* level, const char* parameter) const
{
    const auto it = m_index.find(form_key(level, parameter));
    const XmlNode* root = m_doc.child("DVMpp");
    const XmlNode* l = root ? root->child(level) : nullptr;
    const XmlNode* n = l ? l->child(parameter) : nullptr;
    if (!n) {
        std::ostringstream err;
        if (it != m_index.end()) {
            const opts& o = m_options[it->second];
            err << "xml error: <" << level << "><" << parameter << "> missing.\nAttribute: " << o.attr
                << "\nInfo: " << o.info << "\nOptions: " << o.reqs;
        } else {
            err << "xml error: Are you trying to add an option without updating "
                   "XmlHandler::m_option_map?\nRequested option: <"
                << level << "><" << parameter << ">";
        }
        throw dvm::XMLParseException(err.str());
    }
    if (it == m_index.end())
        throw dvm::XMLParseException("<" + std::string(level) + "><" + parameter +
                                     "> not found in the option map.");
    return *n;
}

std::string XmlHandler::getStringAttribute(const char* level, const char* parameter, bool checksep) const
{
    const std::string* a = getNode(level, parameter).attribute("string");
    std::string attr = a ? *a : std::string();
    check_string(level, parameter, attr);
    if (checksep && (attr.empty() || attr.back() != '/')) attr += '/';
    return attr;
}

std::vector<double> XmlHandler::getList(const char* level, const char* parameter) const
{
    const std::string* a = getNode(level, parameter).attribute("string");
    const std::string vals = a ? *a : std::string();
    check_string(level, parameter, vals);
    std::vector<double> out;
    std::istringstream in(vals);
    double v;
    while (in >> v) {
        out.push_back(v);
        if (in.peek() == ' ' || in.peek() == ',') in.ignore();
    }
    return out;
}

double XmlHandler::getValueAttribute(const char* level, const char* parameter) const
{
    const std::string* a = getNode(level, parameter).attribute("val");
    const double v = a ? std::strtod(a->c_str(), nullptr) : 0.0;
    check_value(level, parameter, v);
    return v;
}

int XmlHandler::getIntAttribute(const char* level, const char* parameter) const
{
    const std::string* a = getNode(level, parameter).attribute("val");
    return a ? parse_int(*a) : 0;
}

bool XmlHandler::getBoolAttribute(const char* level, const char* parameter) const
{
    const std::string* a = getNode(level, parameter).attribute("string");
    if (!a || a->empty()) return false;
    const char c = (*a)[0];
    return c == '1' || c == 't' || c == 'T' || c == 'y' || c == 'Y';
}

void XmlHandler::save(std::string filename) const
{
    std::ofstream out(filename, std::ios::binary);
    if (!out) return;  // like the reference, a failed save is not an error
    for (const XmlNode& top : m_doc.children) out << serialise(top);
}

void XmlHandler::check_string(const char* level, const char* parameter, const std::string& attr) const
{
    const std::string& req = m_options[m_index.at(form_key(level, parameter))].reqs;
    if (req == "none") return;
    // a substring test, as in the reference (an empty value passes)
    if (req.find(attr) == std::string::npos)
        throw dvm::XMLParseException::invalidOption(level, parameter, req);
}

void XmlHandler::check_value(const char* level, const char* parameter, double attr) const
{
    const std::string& req = m_options[m_index.at(form_key(level, parameter))].reqs;
    if (req == "none") return;
    if (req == ">0" && attr > 0) return;
    throw dvm::XMLParseException::invalidOption(level, parameter, req);
}

void XmlHandler::add(const char* level, const char* parameter, const char* attr, const char* info,
                     const char* reqs)
{
    m_index.emplace(form_key(level, parameter), m_options.size());
    m_options.push_back({level, parameter, attr, info, reqs});
}

// The options a DVMpp input may hold: the reference's table
// (src/XmlHandler.cpp:192-229), same keys, attribute kinds and allowed values.
void XmlHandler::buildOptionMap()
{
    add("algorithms", "surface_crossing", "string", "surface-crossing algorithm", "DELETE ABSORB REFLECT");
    add("algorithms", "velocity_kernel", "string", "velocity computation kernel", "direct fmm");
    add("algorithms", "fmm_expansion_order", "val", "FMM expansion order", ">0");
    add("algorithms", "fmm_theta", "val", "FMM opening angle criterion", "none");
    add("algorithms", "merge_threshold", "val", "vortex merge threshold (0=disabled)", "none");
    add("algorithms", "merge_frequency", "val", "merge every N steps (0=disabled)", "none");

    add("constants", "density", "val", "specific gravity of fluid", ">0");
    add("constants", "nu", "val", "kinematic viscosity [m^2 / s]", ">0");
    add("constants", "max_NumPanelVort", "val", "maximum number of panel vortices", ">0");
    add("constants", "cutoff_exp", "val", "cutoff exp", "none");
    add("constants", "seed", "val", "seed for rng. <0 random, >=0 seed", "none");

    add("io", "input_dir", "string", "input directory", "none");
    add("io", "output_dir", "string", "output directory", "none");
    add("io", "domain_file", "string", "body geometry file", "none");

    add("flow", "ux", "val", "flow in the x-direction", "none");
    add("flow", "uz", "val", "flow in the z-direction", "none");

    add("probe", "x", "val", "x-coordinate of the probe", "none");
    add("probe", "z", "val", "z-coordinate of the probe", "none");

    add("time", "scheme", "string", "time-stepping scheme", "euler RK3");
    add("time", "dt", "val", "time step [s]", ">0");
    add("time", "steps", "val", "total number of timesteps", ">0");

    add("remesh", "frequency", "val", "remesh every N steps (0=disabled)", "none");
    add("remesh", "xmin", "val", "mesh x-direction minimum bound", "none");
    add("remesh", "xmax", "val", "mesh x-direction maximum bound", "none");
    add("remesh", "zmin", "val", "mesh z-direction minimum bound", "none");
    add("remesh", "zmax", "val", "mesh z-direction maximum bound", "none");
    add("remesh", "dx", "val", "mesh spacing in x-direction", ">0");
    add("remesh", "dz", "val", "mesh spacing in z-direction", ">0");
    add("remesh", "kernel", "string", "interpolation kernel", "M4prime gaussian M6");
    add("remesh", "sigma", "val", "core size for new particles", ">0");
    add("remesh", "circ_threshold", "val", "minimum circulation for particle creation", "none");
}

void XmlHandler::writeExample(std::string file) const
{
    XmlNode root;
    root.name = "DVMpp";
    for (const opts& o : m_options) {
        XmlNode* l = root.child(o.level);
        if (!l) {
            root.children.push_back(XmlNode{o.level, {}, {}, {}});
            l = &root.children.back();
        }
        XmlNode p{o.parameter, {{o.attr, o.info + ". Options: " + o.reqs}}, {}, {}};
        l->children.push_back(p);
    }
    const std::string doc = serialise(root);
    if (file == "-") {
        std::cout << doc;
        return;
    }
    std::cout << "Saving example xml to " << file << std::endl;
    std::ofstream out(file, std::ios::binary);
    if (!out) throw dvm::FileIOException::cannotOpen(file);
    out << doc;
}
