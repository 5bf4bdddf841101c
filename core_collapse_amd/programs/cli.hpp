// cli.hpp -- command-line flags of the drop-in programs.
//
// The reference programs take no arguments: every setting is a default member of their struct
// Params.  Here each such member (and each IC constant) is a flag of the same name whose default
// is the reference value, so running a program with no arguments reproduces the reference run.
// Flags are written --name value or --name=value; booleans accept 0/1/true/false.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace ccprog {

class Cli {
   public:
    explicit Cli(std::string prog, std::string about) : prog_(std::move(prog)), about_(std::move(about)) {}

    void add(const char* name, double* v, const char* help) { flags_.push_back({name, kDouble, v, help}); }
    void add(const char* name, int* v, const char* help) { flags_.push_back({name, kInt, v, help}); }
    void add(const char* name, unsigned* v, const char* help) { flags_.push_back({name, kUnsigned, v, help}); }
    void add_bool(const char* name, int* v, const char* help) { flags_.push_back({name, kBool, v, help}); }
    void add(const char* name, std::string* v, const char* help) { flags_.push_back({name, kString, v, help}); }

    // Returns 0 to continue, 1 after --help, 2 on a bad flag (message on stderr).
    int parse(int argc, char** argv) {
        for (int a = 1; a < argc; a++) {
            std::string arg = argv[a];
            if (arg == "-h" || arg == "--help") {
                usage(stdout);
                return 1;
            }
            if (arg.rfind("--", 0) != 0) return fail("unexpected argument '" + arg + "'");
            std::string name = arg.substr(2), value;
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
            } else {
                if (a + 1 >= argc) return fail("flag --" + name + " needs a value");
                value = argv[++a];
            }
            Flag* f = find(name);
            if (!f) return fail("unknown flag --" + name);
            if (!set(*f, value)) return fail("bad value '" + value + "' for --" + name);
        }
        return 0;
    }

    void usage(FILE* out) const {
        std::fprintf(out, "usage: %s [--flag value ...]\n%s\n\nflags (default in brackets):\n", prog_.c_str(),
                     about_.c_str());
        for (const Flag& f : flags_) std::fprintf(out, "  --%-14s %s [%s]\n", f.name, f.help, show(f).c_str());
    }

   private:
    enum Kind { kDouble, kInt, kUnsigned, kBool, kString };
    struct Flag {
        const char* name;
        Kind kind;
        void* ptr;
        const char* help;
    };
    std::string prog_, about_;
    std::vector<Flag> flags_;

    Flag* find(const std::string& n) {
        for (Flag& f : flags_)
            if (n == f.name) return &f;
        return nullptr;
    }
    int fail(const std::string& msg) const {
        std::fprintf(stderr, "%s: %s (see --help)\n", prog_.c_str(), msg.c_str());
        return 2;
    }
    static bool set(Flag& f, const std::string& v) {
        char* end = nullptr;
        switch (f.kind) {
            case kDouble: {
                const double x = std::strtod(v.c_str(), &end);
                if (end == v.c_str() || *end) return false;
                *(double*)f.ptr = x;
                return true;
            }
            case kInt: {
                const long x = std::strtol(v.c_str(), &end, 10);
                if (end == v.c_str() || *end) return false;
                *(int*)f.ptr = (int)x;
                return true;
            }
            case kUnsigned: {
                const unsigned long x = std::strtoul(v.c_str(), &end, 10);
                if (end == v.c_str() || *end) return false;
                *(unsigned*)f.ptr = (unsigned)x;
                return true;
            }
            case kBool:
                if (v == "1" || v == "true") *(int*)f.ptr = 1;
                else if (v == "0" || v == "false") *(int*)f.ptr = 0;
                else return false;
                return true;
            case kString:
                *(std::string*)f.ptr = v;
                return true;
        }
        return false;
    }
    static std::string show(const Flag& f) {
        char b[64];
        switch (f.kind) {
            case kDouble: std::snprintf(b, sizeof b, "%.17g", *(const double*)f.ptr); return b;
            case kInt: return std::to_string(*(const int*)f.ptr);
            case kUnsigned: return std::to_string(*(const unsigned*)f.ptr);
            case kBool: return *(const int*)f.ptr ? "true" : "false";
            case kString: return *(const std::string*)f.ptr;
        }
        return "";
    }
};

// "fp64" / "mixed" -> cc_precision value, -1 otherwise.
inline int parse_precision(const std::string& s) {
    if (s == "fp64") return 0;
    if (s == "mixed") return 1;
    return -1;
}

}  // namespace ccprog
