// Exception hierarchy of the drop-in layer: the reference's names and
// messages (src/Exceptions.hpp:13-81), plus the mapping from dvm_status.
#ifndef DVMPP_AMD_EXCEPTIONS_HPP
#define DVMPP_AMD_EXCEPTIONS_HPP

#include <stdexcept>
#include <string>

#include "dvm_hip.h"

namespace dvm {

class DVMException : public std::runtime_error {
public:
    explicit DVMException(const std::string& m) : std::runtime_error(m) {}
};

class SizeMismatchException : public DVMException {
public:
    explicit SizeMismatchException(const std::string& context)
        : DVMException("Size mismatch in " + context) {}
};

class FileIOException : public DVMException {
public:
    explicit FileIOException(const std::string& m) : DVMException(m) {}
    static FileIOException cannotOpen(const std::string& f) { return FileIOException("Cannot open file: " + f); }
    static FileIOException cannotWrite(const std::string& f) { return FileIOException("Cannot write to file: " + f); }
    static FileIOException cannotRead(const std::string& f) { return FileIOException("Cannot read file: " + f); }
};

class XMLParseException : public DVMException {
public:
    explicit XMLParseException(const std::string& m) : DVMException(m) {}
    static XMLParseException cannotLoad(const std::string& f, const std::string& why)
    {
        return XMLParseException("Cannot load XML file '" + f + "': " + why);
    }
    static XMLParseException missingParameter(const std::string& level, const std::string& param)
    {
        return XMLParseException("Missing XML parameter: <" + level + "><" + param + ">");
    }
    static XMLParseException invalidOption(const std::string& level, const std::string& param,
                                           const std::string& options)
    {
        return XMLParseException("Invalid option for <" + level + "><" + param +
                                 ">. Valid options: " + options);
    }
};

class AlgorithmException : public DVMException {
public:
    explicit AlgorithmException(const std::string& m) : DVMException(m) {}
};

class VortexException : public DVMException {
public:
    explicit VortexException(const std::string& m) : DVMException(m) {}
};

// Throws the exception the reference raises for the given status.
void throw_status(dvm_status st, const std::string& where, dvm_ctx_t ctx);

inline void check(dvm_status st, const char* where, dvm_ctx_t ctx)
{
    if (st != DVM_OK) throw_status(st, where, ctx);
}

}  // namespace dvm

#endif
