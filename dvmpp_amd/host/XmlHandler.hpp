// Input-file reader of the reference (src/XmlHandler.hpp:15-100): the same
// option map, getters, error behaviour, save() and writeExample(), over a
// small XML reader of this library's own (elements, attributes, text,
// comments and declarations; no DTD or namespaces), so no third-party XML
// library is needed.  Values live in one layer of nesting under <DVMpp>.
#ifndef DVMPP_AMD_XMLHANDLER_HPP
#define DVMPP_AMD_XMLHANDLER_HPP

#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

/// One element of a parsed document.
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;
    std::string text;  ///< character data directly inside the element

    /// First child element called `n`, or nullptr.
    [[nodiscard]] const XmlNode* child(const std::string& n) const;
    XmlNode* child(const std::string& n);
    /// Value of attribute `n`; nullptr when absent.
    [[nodiscard]] const std::string* attribute(const std::string& n) const;
};

class XmlHandler {
public:
    /// Only builds the option map (for writeExample).
    XmlHandler();
    /// Parses `xml_file`; throws dvm::XMLParseException if it cannot.
    explicit XmlHandler(std::string xml_file);

    /// "string" attribute of <DVMpp><level><parameter>; checked against the
    /// option's allowed values; `checksep` appends a trailing '/'.
    [[nodiscard]] std::string getStringAttribute(const char* level, const char* parameter,
                                                 bool checksep = false) const;
    /// Space- or comma-separated numbers of the "string" attribute.
    [[nodiscard]] std::vector<double> getList(const char* level, const char* parameter) const;
    /// "val" attribute as a double, checked against the option's requirement.
    [[nodiscard]] double getValueAttribute(const char* level, const char* parameter) const;
    /// "val" attribute as an integer.
    [[nodiscard]] int getIntAttribute(const char* level, const char* parameter) const;
    /// "string" attribute as a bool (leading 1/t/T/y/Y is true).
    [[nodiscard]] bool getBoolAttribute(const char* level, const char* parameter) const;
    /// Writes the parsed document (indented, with an XML declaration).
    void save(std::string filename) const;
    /// Writes a document listing every option, to stdout for "-".
    void writeExample(std::string file) const;

    /// Whether <DVMpp><level><parameter> is present (no option-map check).
    [[nodiscard]] bool has(const char* level, const char* parameter) const;

    /// Parses XML text; throws dvm::XMLParseException naming `source`.
    static XmlNode parse(const std::string& text, const std::string& source);
    /// Serialises a document: declaration, then `root` indented by tabs.
    static std::string serialise(const XmlNode& root);

private:
    struct opts {
        std::string level;
        std::string parameter;
        std::string attr;
        std::string info;
        std::string reqs;
    };

    XmlNode m_doc;  ///< document node; its children are the top-level elements
    std::vector<opts> m_options;                      ///< in declaration order
    std::unordered_map<std::string, size_t> m_index;  ///< "level_parameter" -> m_options

    [[nodiscard]] const XmlNode& getNode(const char* level, const char* parameter) const;
    void check_string(const char* level, const char* parameter, const std::string& attr) const;
    void check_value(const char* level, const char* parameter, double attr) const;
    [[nodiscard]] static std::string form_key(const char* level, const char* parameter);
    void buildOptionMap();
    void add(const char* level, const char* parameter, const char* attr, const char* info,
             const char* reqs);
};

#endif
