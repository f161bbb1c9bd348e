#include "xf_json.h"
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace xfh {

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string &t) : s(t) {}

    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool fail(const char *m)
    {
        if (err.empty()) err = std::string(m) + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char *w)
    {
        size_t n = std::char_traits<char>::length(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    bool str(std::string &o)
    {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        i++;
        o.clear();
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\' && i < s.size()) {
                char e = s[i++];
                switch (e) {
                case 'n': o += '\n'; break;
                case 't': o += '\t'; break;
                case 'r': o += '\r'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'u': o += '?'; i += 4; break;
                default: o += e;
                }
            } else
                o += c;
        }
        if (i >= s.size()) return fail("unterminated string");
        i++;
        return true;
    }
    bool value(Json &v)
    {
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.kind = Json::Object;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') {
                i++;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                i++;
                Json e;
                if (!value(e)) return false;
                v.obj[k] = std::move(e);
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    ws();
                    if (i < s.size() && s[i] == '}') { // tolerate trailing comma
                        i++;
                        return true;
                    }
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    i++;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::Array;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') {
                i++;
                return true;
            }
            for (;;) {
                Json e;
                if (!value(e)) return false;
                v.arr.push_back(std::move(e));
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    ws();
                    if (i < s.size() && s[i] == ']') {
                        i++;
                        return true;
                    }
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    i++;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::String;
            return str(v.str);
        }
        if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = Json::Bool;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.kind = Json::Null;
            return true;
        }
        const char *b = s.c_str() + i;
        char *e = nullptr;
        double d = std::strtod(b, &e);
        if (e == b) return fail("unexpected character");
        i += (size_t)(e - b);
        v.kind = Json::Number;
        v.num = d;
        return true;
    }
};

} // namespace

bool json_parse(const std::string &text, Json &out, std::string &err)
{
    Parser p(text);
    out = Json();
    if (!p.value(out)) {
        err = p.err;
        return false;
    }
    return true;
}

bool json_read_settings(const std::string &path, Json &out, std::string &err)
{
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    std::string line;
    while (std::getline(in, line)) {
        size_t pos = line.find_first_of("//");
        ss << line.substr(0, pos) << '\n';
    }
    return json_parse(ss.str(), out, err);
}

} // namespace xfh
