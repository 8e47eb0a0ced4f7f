// config.cpp - parameter files (YAML) and the Middlebury calib.txt reader. Host code only.
//
// The reference reads its parameter files with yaml-cpp (src/config.cpp:17-72). Its files are
// a top-level block mapping of scalars, so this reader parses exactly that subset: comments,
// blank lines, document markers, `key: value` lines with plain or quoted scalars. Lines
// indented under a key make that key's value a nested node, which a scalar conversion rejects
// (yaml-cpp throws there, and the reference's loader returns false).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "lumen_config.h"
#include "lumen_sgbm.h"

namespace lumen {
void set_last_error(const std::string& msg);
}

namespace {

struct Scalar {
    std::string text;
    bool nested = false;  // the key has an indented block under it (not a scalar)
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

// Removes a comment: '#' at the start of the line or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
    char quote = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (quote) {
            if (c == quote) quote = 0;
        } else if (c == '"' || c == '\'') {
            quote = c;
        } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
            return s.substr(0, i);
        }
    }
    return s;
}

bool unquote(std::string& v) {
    if (v.size() >= 2 && (v[0] == '"' || v[0] == '\'') && v.back() == v[0]) {
        v = v.substr(1, v.size() - 2);
        return true;
    }
    return !(v.size() && (v[0] == '"' || v[0] == '\''));  // an unterminated quote is malformed
}

// Parses the top-level mapping; false (with `why`) when the text is not one.
bool parse_mapping(std::istream& in, std::map<std::string, Scalar>& out, std::string& why) {
    std::string line, lastKey;
    int lineNo = 0;
    while (std::getline(in, line)) {
        ++lineNo;
        std::string body = strip_comment(line);
        if (trim(body).empty()) continue;
        if (body.rfind("---", 0) == 0 || body.rfind("...", 0) == 0) continue;
        if (body[0] == ' ' || body[0] == '\t') {
            if (lastKey.empty()) {
                why = "line " + std::to_string(lineNo) + ": indented line outside a mapping";
                return false;
            }
            out[lastKey].nested = true;
            continue;
        }
        const size_t colon = body.find(':');
        if (colon == std::string::npos || colon == 0) {
            why = "line " + std::to_string(lineNo) + ": expected `key: value`";
            return false;
        }
        std::string key = trim(body.substr(0, colon));
        std::string val = trim(body.substr(colon + 1));
        if (!unquote(key)) {
            why = "line " + std::to_string(lineNo) + ": malformed key";
            return false;
        }
        const bool quoted = val.size() && (val[0] == '"' || val[0] == '\'');
        if (!unquote(val)) {
            why = "line " + std::to_string(lineNo) + ": unterminated quoted scalar";
            return false;
        }
        Scalar s;
        s.text = val;
        // An empty value is a null node, or the start of a nested block; flow collections
        // ([..] or {..}) are not scalars either.
        s.nested = !quoted && (val.empty() || val[0] == '[' || val[0] == '{');
        out[key] = s;
        lastKey = key;
    }
    return true;
}

// yaml-cpp's convert<int>, which reads the scalar with a stream whose base is detected from
// the prefix: optional sign, then 0x/0X hex, 0o octal, a leading 0 octal, or decimal digits,
// and nothing after them.
bool to_int(const std::string& s, int& v) {
    if (s.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (s[0] == '+' || s[0] == '-') {
        neg = s[0] == '-';
        i = 1;
    }
    int base = 10;
    if (s.size() > i + 2 && s[i] == '0' && (s[i + 1] == 'x' || s[i + 1] == 'X' || s[i + 1] == 'o')) {
        base = s[i + 1] == 'o' ? 8 : 16;
        i += 2;
    } else if (s.size() > i + 1 && s[i] == '0') {
        base = 8;
        i += 1;
    }
    if (i >= s.size()) return false;
    long long acc = 0;
    for (; i < s.size(); ++i) {
        const char c = s[i];
        int d;
        if (c >= '0' && c <= '9')
            d = c - '0';
        else if (base == 16 && c >= 'a' && c <= 'f')
            d = c - 'a' + 10;
        else if (base == 16 && c >= 'A' && c <= 'F')
            d = c - 'A' + 10;
        else
            return false;
        if (d >= base) return false;
        acc = acc * base + d;
        if (acc > 2147483648LL) return false;
    }
    if (!neg && acc > 2147483647LL) return false;
    v = static_cast<int>(neg ? -acc : acc);
    return true;
}

// yaml-cpp's convert<bool>.
bool to_bool(const std::string& s, bool& v) {
    static const char* yes[] = {"y", "Y", "yes", "Yes", "YES", "true", "True", "TRUE", "on", "On", "ON"};
    static const char* no[] = {"n", "N", "no", "No", "NO", "false", "False", "FALSE", "off", "Off", "OFF"};
    for (const char* t : yes)
        if (s == t) return (v = true, true);
    for (const char* t : no)
        if (s == t) return (v = false, true);
    return false;
}

// StereoParams::isValid (include/stereo/stereo_params.h:96-103).
bool params_valid(const lumen_stereo_params& p) {
    if (p.blockSize < 1 || p.blockSize % 2 == 0) return false;
    if (p.minDisparity < 0) return false;
    if (p.maxDisparity <= p.minDisparity) return false;
    if (p.P2 <= p.P1) return false;
    if (p.uniquenessRatio < 0 || p.uniquenessRatio > 100) return false;
    return true;
}

int fail(const std::string& msg) {
    lumen::set_last_error(msg);
    return LUMEN_EINVAL;
}

// "[a b c; d e f; g h i]" -> nine values (std::istringstream, as the reference reads them).
bool parse_matrix(const std::string& line, double (&v)[9]) {
    const size_t a = line.find('['), b = line.find(']');
    if (a == std::string::npos || b == std::string::npos || b < a) return false;
    std::string m = line.substr(a + 1, b - a - 1);
    for (char& c : m)
        if (c == ';') c = ' ';
    std::istringstream iss(m);
    for (double& x : v)
        if (!(iss >> x)) return false;
    return true;
}

bool parse_double(const std::string& s, double& v) {
    const char* p = s.c_str();
    char* end = nullptr;
    errno = 0;
    v = std::strtod(p, &end);
    return end != p && errno != ERANGE;
}

bool parse_int(const std::string& s, int& v) {
    const char* p = s.c_str();
    char* end = nullptr;
    errno = 0;
    const long r = std::strtol(p, &end, 10);
    if (end == p || errno == ERANGE || r < INT32_MIN || r > INT32_MAX) return false;
    v = static_cast<int>(r);
    return true;
}

}  // namespace

extern "C" int lumen_params_load_yaml(const char* path, lumen_stereo_params* params) {
    if (!path || !params) return fail("NULL argument");
    std::ifstream in(path);
    if (!in.is_open()) return fail(std::string("Failed to load stereo params: cannot open ") + path);
    std::map<std::string, Scalar> m;
    std::string why;
    if (!parse_mapping(in, m, why)) return fail("Failed to load stereo params: " + why);
    lumen_stereo_params p = *params;
    struct IntKey {
        const char* name;
        int32_t* field;
    } ints[] = {{"minDisparity", &p.minDisparity},   {"maxDisparity", &p.maxDisparity},
                {"blockSize", &p.blockSize},         {"P1", &p.P1},
                {"P2", &p.P2},                       {"uniquenessRatio", &p.uniquenessRatio},
                {"disp12MaxDiff", &p.disp12MaxDiff}, {"preFilterCap", &p.preFilterCap},
                {"numDirections", &p.numDirections}};
    for (const IntKey& k : ints) {
        auto it = m.find(k.name);
        if (it == m.end()) continue;
        int v;
        if (it->second.nested || !to_int(it->second.text, v))
            return fail(std::string("Failed to load stereo params: bad conversion of ") + k.name);
        *k.field = v;
    }
    if (auto it = m.find("useGPU"); it != m.end()) {
        bool b;
        if (it->second.nested || !to_bool(it->second.text, b))
            return fail("Failed to load stereo params: bad conversion of useGPU");
        p.useGPU = b ? 1 : 0;
    }
    if (auto it = m.find("matchingCostMode"); it != m.end()) {
        if (it->second.nested) return fail("Failed to load stereo params: bad conversion of matchingCostMode");
        const std::string& s = it->second.text;
        p.matchingCostMode = (s == "Census" || s == "census") ? 1 : 0;
    }
    *params = p;
    if (!params_valid(p)) return fail("Loaded stereo params are not valid (StereoParams::isValid)");
    return LUMEN_OK;
}

extern "C" int lumen_params_save_yaml(const char* path, const lumen_stereo_params* params) {
    if (!path || !params) return fail("NULL argument");
    const lumen_stereo_params& p = *params;
    std::ostringstream out;
    out << "minDisparity: " << p.minDisparity << "\n"
        << "maxDisparity: " << p.maxDisparity << "\n"
        << "blockSize: " << p.blockSize << "\n"
        << "P1: " << p.P1 << "\n"
        << "P2: " << p.P2 << "\n"
        << "preFilterCap: " << p.preFilterCap << "\n"
        << "uniquenessRatio: " << p.uniquenessRatio << "\n"
        << "disp12MaxDiff: " << p.disp12MaxDiff << "\n"
        << "useGPU: " << (p.useGPU ? "true" : "false") << "\n"
        << "numDirections: " << p.numDirections << "\n"
        << "matchingCostMode: " << (p.matchingCostMode == 1 ? "Census" : "SAD");
    std::ofstream f(path);
    if (!f.is_open()) return fail(std::string("Failed to save stereo params: cannot open ") + path);
    f << out.str();
    f.close();
    if (!f) return fail(std::string("Failed to save stereo params: write error on ") + path);
    return LUMEN_OK;
}

extern "C" int lumen_middlebury_calib_load(const char* path, lumen_camera_params* params) {
    if (!path || !params) return fail("NULL argument");
    std::ifstream file(path);
    if (!file.is_open()) return fail(std::string("Failed to open Middlebury calib file: ") + path);
    lumen_camera_params c = *params;
    static const double eye[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::memcpy(c.K0, eye, sizeof eye);
    std::memcpy(c.K1, eye, sizeof eye);
    std::memcpy(c.R, eye, sizeof eye);
    for (double& v : c.D0) v = 0;
    for (double& v : c.D1) v = 0;
    for (double& v : c.T) v = 0;
    double doffs = 0, baselineMm = 0;
    int ndisp = 0;
    std::string line;
    while (std::getline(file, line)) {
        bool ok = true;
        if (line.rfind("cam0=", 0) == 0 || line.rfind("cam1=", 0) == 0) {
            double v[9];
            ok = parse_matrix(line, v);
            double* K = line[3] == '0' ? c.K0 : c.K1;
            if (ok) {
                K[0] = v[0];
                K[2] = v[2];
                K[4] = v[4];
                K[5] = v[5];
            }
        } else if (line.rfind("doffs=", 0) == 0) {
            ok = parse_double(line.substr(6), doffs);
        } else if (line.rfind("baseline=", 0) == 0) {
            ok = parse_double(line.substr(9), baselineMm);
        } else if (line.rfind("width=", 0) == 0) {
            ok = parse_int(line.substr(6), c.width);
        } else if (line.rfind("height=", 0) == 0) {
            ok = parse_int(line.substr(7), c.height);
        } else if (line.rfind("ndisp=", 0) == 0) {
            ok = parse_int(line.substr(6), ndisp);
        }
        if (!ok) return fail("Malformed Middlebury calib line: " + trim(line));
    }
    c.baseline = baselineMm / 1000.0;
    c.disparityOffset = static_cast<int32_t>(doffs + 0.5);
    c.numDisparities = ndisp > 0 ? ndisp : 280;
    c.T[0] = -baselineMm / 1000.0;
    c.focalLength = c.K0[0];
    *params = c;
    if (!(c.width > 0 && c.height > 0)) return fail("Middlebury calib file has no image size");
    return LUMEN_OK;
}
