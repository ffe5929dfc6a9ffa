#include "InputOutput/jsonlite.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace LatticeIO {
namespace {
struct Parser {
    const std::string &t;
    size_t i = 0;
    explicit Parser(const std::string &text) : t(text) {}

    [[noreturn]] void error(const char *what)
    {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) i++;
    }
    bool eat(char c)
    {
        ws();
        if (i < t.size() && t[i] == c) {
            i++;
            return true;
        }
        return false;
    }
    void expect(char c)
    {
        if (!eat(c)) error((std::string("expected '") + c + "'").c_str());
    }
    std::string str()
    {
        expect('"');
        std::string out;
        while (i < t.size() && t[i] != '"') {
            char c = t[i++];
            if (c == '\\') {
                if (i >= t.size()) error("bad escape");
                char e = t[i++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i + 4 > t.size()) error("bad \\u escape");
                    unsigned cp = unsigned(std::strtoul(t.substr(i, 4).c_str(), nullptr, 16));
                    i += 4;
                    if (cp < 0x80) out += char(cp);
                    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                    else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= t.size()) error("unterminated string");
        i++;
        return out;
    }
    JsonValue value()
    {
        ws();
        if (i >= t.size()) error("unexpected end");
        JsonValue v;
        const char c = t[i];
        if (c == '{') {
            i++;
            v.kind = JsonValue::Object;
            if (eat('}')) return v;
            do {
                ws();
                std::string k = str();
                expect(':');
                v.obj[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            i++;
            v.kind = JsonValue::Array;
            if (eat(']')) return v;
            do v.arr.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = JsonValue::String;
            v.str = str();
        } else if (t.compare(i, 4, "true") == 0) {
            i += 4;
            v.kind = JsonValue::Bool;
            v.b = true;
        } else if (t.compare(i, 5, "false") == 0) {
            i += 5;
            v.kind = JsonValue::Bool;
        } else if (t.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            char *end = nullptr;
            v.num = std::strtod(t.c_str() + i, &end);
            if (end == t.c_str() + i) error("unexpected character");
            i = size_t(end - t.c_str());
            v.kind = JsonValue::Number;
        }
        return v;
    }
};
}  // namespace

const JsonValue &JsonValue::operator[](const std::string &k) const
{
    if (kind != Object) throw std::runtime_error("JSON: \"" + k + "\" looked up in a non-object");
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("JSON: missing key \"" + k + "\"");
    return it->second;
}

double JsonValue::number() const
{
    if (kind != Number) throw std::runtime_error("JSON: expected a number");
    return num;
}

int JsonValue::integer() const
{
    const double d = number();
    if (d != std::floor(d)) throw std::runtime_error("JSON: expected an integer");
    return int(d);
}

const std::string &JsonValue::string() const
{
    if (kind != String) throw std::runtime_error("JSON: expected a string");
    return str;
}

JsonValue JsonValue::parse(const std::string &text)
{
    Parser p(text);
    JsonValue v = p.value();
    p.ws();
    if (p.i != text.size()) p.error("trailing characters");
    return v;
}

JsonValue JsonValue::parseFile(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}
}  // namespace LatticeIO
