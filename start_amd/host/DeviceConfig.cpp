// Device options from the run's JSON configuration file: the reference's
// config (read by ConfigParser, ConfigParser.cpp:18-226) plus optional
// "Device ..." keys that the reference does not know and ignores.  A small
// JSON reader of its own (objects, arrays, strings, numbers, literals); the
// reference's own keys are skipped.
#include "StartProcessors.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace start_amd {

namespace {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;  // String; for Number, the text as written
    std::vector<JsonValue> items;
    std::vector<std::pair<std::string, JsonValue>> members;  // in file order
};

class JsonReader {
public:
    explicit JsonReader(const std::string& text) : s(text) {}

    JsonValue document()
    {
        JsonValue v = value();
        space();
        if (p != s.size()) fail("text after the top-level value");
        return v;
    }

private:
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void fail(const std::string& what) const
    {
        size_t line = 1;
        for (size_t i = 0; i < p && i < s.size(); ++i) line += s[i] == '\n';
        throw std::runtime_error("JSON line " + std::to_string(line) + ": " + what);
    }
    void space()
    {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    bool eat(char c)
    {
        space();
        if (p < s.size() && s[p] == c) {
            ++p;
            return true;
        }
        return false;
    }
    void expect(char c)
    {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }
    JsonValue value()
    {
        space();
        if (p >= s.size()) fail("unexpected end of text");
        const char c = s[p];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            JsonValue v;
            v.kind = JsonValue::String;
            v.str = string();
            return v;
        }
        if (c == '-' || std::isdigit((unsigned char)c)) return number();
        JsonValue v;
        if (s.compare(p, 4, "true") == 0) {
            p += 4;
            v.kind = JsonValue::Bool;
            v.b = true;
        } else if (s.compare(p, 5, "false") == 0) {
            p += 5;
            v.kind = JsonValue::Bool;
        } else if (s.compare(p, 4, "null") == 0) {
            p += 4;
        } else {
            fail("unexpected character");
        }
        return v;
    }
    JsonValue object()
    {
        JsonValue v;
        v.kind = JsonValue::Object;
        expect('{');
        if (eat('}')) return v;
        do {
            space();
            if (p >= s.size() || s[p] != '"') fail("expected a member name");
            std::string k = string();
            expect(':');
            v.members.emplace_back(std::move(k), value());
        } while (eat(','));
        expect('}');
        return v;
    }
    JsonValue array()
    {
        JsonValue v;
        v.kind = JsonValue::Array;
        expect('[');
        if (eat(']')) return v;
        do v.items.push_back(value());
        while (eat(','));
        expect(']');
        return v;
    }
    static void utf8(std::string& out, unsigned cp)
    {
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
    }
    unsigned hex4()
    {
        if (p + 4 > s.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string()
    {
        ++p;  // opening quote
        std::string out;
        while (true) {
            if (p >= s.size()) fail("unterminated string");
            const char c = s[p++];
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) fail("unterminated escape");
            const char e = s[p++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && p + 1 < s.size() && s[p] == '\\' && s[p + 1] == 'u') {
                    p += 2;
                    const unsigned lo = hex4();
                    if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
    }
    JsonValue number()
    {
        const size_t a = p;
        if (s[p] == '-') ++p;
        auto digits = [&] {
            const size_t d = p;
            while (p < s.size() && std::isdigit((unsigned char)s[p])) ++p;
            if (p == d) fail("malformed number");
        };
        digits();
        if (p < s.size() && s[p] == '.') {
            ++p;
            digits();
        }
        if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
            ++p;
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) ++p;
            digits();
        }
        JsonValue v;
        v.kind = JsonValue::Number;
        v.str = s.substr(a, p - a);
        v.num = std::strtod(v.str.c_str(), nullptr);
        return v;
    }
};

std::string lower(std::string v)
{
    for (char& c : v) c = (char)std::tolower((unsigned char)c);
    return v;
}

bool as_bool(const std::string& key, const JsonValue& v)
{
    if (v.kind != JsonValue::Bool) throw std::runtime_error("\"" + key + "\" must be true or false");
    return v.b;
}

// a non-negative integer given as a JSON number without fraction or exponent
uint64_t as_count(const std::string& key, const JsonValue& v)
{
    if (v.kind != JsonValue::Number || v.str.find_first_of(".eE-") != std::string::npos)
        throw std::runtime_error("\"" + key + "\" must be a non-negative integer");
    errno = 0;
    const unsigned long long n = std::strtoull(v.str.c_str(), nullptr, 10);
    if (errno) throw std::runtime_error("\"" + key + "\" is out of range");
    return (uint64_t)n;
}

std::map<ParticleType, double> per_type(const std::string& key, const JsonValue& v, double scale, bool positive)
{
    if (v.kind != JsonValue::Object) throw std::runtime_error("\"" + key + "\" must be an object of particle types");
    std::map<ParticleType, double> out;
    for (const auto& [name, x] : v.members) {
        const ParticleType t = typeFromName(name);
        if (t == Unknown) throw std::runtime_error("\"" + key + "\": unknown particle type \"" + name + "\"");
        if (x.kind != JsonValue::Number || !std::isfinite(x.num) || (positive && !(x.num > 0.0)))
            throw std::runtime_error("\"" + key + "\": the value for " + name + " must be a " +
                                     (positive ? "positive " : "") + "number");
        out[t] = x.num * scale;
    }
    return out;
}

}  // namespace

DeviceKernelOptions deviceOptionsFromConfig(const std::string& path, DeviceKernelOptions options)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open the configuration file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    JsonValue doc;
    try {
        doc = JsonReader(text).document();
    } catch (const std::exception& e) {
        throw std::runtime_error(path + ": " + e.what());
    }
    if (doc.kind != JsonValue::Object) throw std::runtime_error(path + ": the configuration is not a JSON object");
    for (const auto& [key, v] : doc.members) {
        if (key.rfind("Device ", 0) != 0) continue;  // the reference's own keys
        if (key == "Device Precision") {
            const std::string p = v.kind == JsonValue::String ? lower(v.str) : "";
            if (p == "fp64") options.precision = START_PREC_FP64;
            else if (p == "mixed") options.precision = START_PREC_MIXED;
            else throw std::runtime_error("\"Device Precision\" must be \"FP64\" or \"Mixed\"");
        } else if (key == "Device EM Push") {
            options.emPush = as_bool(key, v);
        } else if (key == "Device Seed") {
            options.seed = as_count(key, v);
        } else if (key == "Device Density Mode") {
            const std::string m = v.kind == JsonValue::String ? lower(v.str) : "";
            if (m == "barycentric") options.densityMode = START_DENSITY_BARYCENTRIC;
            else if (m == "tetaverage") options.densityMode = START_DENSITY_TET_AVERAGE;
            else throw std::runtime_error("\"Device Density Mode\" must be \"Barycentric\" or \"TetAverage\"");
        } else if (key == "Device Sort Interval") {
            const uint64_t n = as_count(key, v);
            if (n > 1000000) throw std::runtime_error("\"Device Sort Interval\" is out of range");
            options.sortInterval = (int)n;
        } else if (key == "Device Write Back Each Step") {
            options.writeBackEachStep = as_bool(key, v);
        } else if (key == "Device Movement Max Particles") {
            options.movementMaxParticles = (size_t)as_count(key, v);
        } else if (key == "Device Charges") {  // elementary charges by particle type
            for (const auto& [t, q] : per_type(key, v, 1.602176565e-19, false)) options.chargeOverride[t] = q;
        } else if (key == "Device Diameters") {  // collision diameters r_p + r_t in m by particle type
            for (const auto& [t, d] : per_type(key, v, 1.0, true)) options.diameterOverride[t] = d;
        } else {
            throw std::runtime_error(path + ": unknown device key \"" + key + "\"");
        }
    }
    return options;
}

}  // namespace start_amd
