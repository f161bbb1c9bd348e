// Minimal JSON reader for the settings/*.json case files.
//
// Comment handling follows the reference's ReadJson
// (src/read_ini/settings/read_json.cpp:150-178): every line is cut at its
// first '/' character (std::string::find_first_of("//") matches either
// slash), then the remainder is parsed as JSON.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace xfh {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool has(const std::string &k) const { return kind == Object && obj.count(k); }
    const Json *get(const std::string &k) const
    {
        if (kind != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
    // j.at(section).value(key, default) as the reference reads its keys
    double value(const std::string &k, double def) const
    {
        const Json *v = get(k);
        if (!v) return def;
        if (v->kind == Number) return v->num;
        if (v->kind == Bool) return v->b ? 1.0 : 0.0;
        return def;
    }
    std::string value(const std::string &k, const std::string &def) const
    {
        const Json *v = get(k);
        return (v && v->kind == String) ? v->str : def;
    }
    std::vector<double> values(const std::string &k, const std::vector<double> &def) const
    {
        const Json *v = get(k);
        if (!v || v->kind != Array) return def;
        std::vector<double> out;
        for (const Json &e : v->arr) out.push_back(e.kind == Bool ? (e.b ? 1.0 : 0.0) : e.num);
        return out;
    }
};

// Parse text (comments already stripped).  Returns false and sets err on failure.
bool json_parse(const std::string &text, Json &out, std::string &err);
// Read a settings file applying the reference's line-comment rule.
bool json_read_settings(const std::string &path, Json &out, std::string &err);

} // namespace xfh
