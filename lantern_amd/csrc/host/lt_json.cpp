/*
 * lt_json.cpp - recursive-descent JSON reader (see lt_json.h).
 */
#include "lt_json.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lt_json {

namespace {

const Value kNull;

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    int depth = 0;

    explicit Parser(const std::string &t) : s(t) {}

    bool fail(const char *what) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "JSON parse error at offset %zu: %s", i, what);
        err = buf;
        return false;
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char *w) {
        size_t n = std::char_traits<char>::length(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    static void utf8(std::string &o, unsigned cp) {
        if (cp < 0x80) o += (char)cp;
        else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
        else { o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
    }
    bool string(std::string &o) {
        if (i >= s.size() || s[i] != '"') return fail("expected '\"'");
        ++i;
        while (i < s.size()) {
            char c = s[i++];
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') { o += c; continue; }
            if (i >= s.size()) break;
            char e = s[i++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                if (i + 4 > s.size()) return fail("short \\u escape");
                unsigned cp = 0;
                for (int k = 0; k < 4; ++k) {
                    char h = s[i++];
                    cp <<= 4;
                    if (h >= '0' && h <= '9') cp |= (unsigned)(h - '0');
                    else if (h >= 'a' && h <= 'f') cp |= (unsigned)(h - 'a' + 10);
                    else if (h >= 'A' && h <= 'F') cp |= (unsigned)(h - 'A' + 10);
                    else return fail("bad \\u escape");
                }
                utf8(o, cp);
                break;
            }
            default: return fail("bad escape");
            }
        }
        return fail("unterminated string");
    }
    bool number(Value &v) {
        size_t st = i;
        if (i < s.size() && s[i] == '-') ++i;
        if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return fail("bad number");
        while (i < s.size() && ((s[i] >= '0' && s[i] <= '9') || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                                s[i] == '+' || s[i] == '-'))
            ++i;
        std::string t = s.substr(st, i - st);
        char *end = nullptr;
        errno = 0;
        double d = std::strtod(t.c_str(), &end);
        if (end != t.c_str() + t.size()) return fail("bad number");
        v.type = Value::Number;
        v.num = d;
        return true;
    }
    bool value(Value &v) {
        if (++depth > 256) return fail("nesting too deep");
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        bool ok;
        if (c == '{') ok = object(v);
        else if (c == '[') ok = array(v);
        else if (c == '"') { v.type = Value::String; ok = string(v.str); }
        else if (c == 't') { ok = lit("true"); v.type = Value::Bool; v.b = true; if (!ok) fail("bad literal"); }
        else if (c == 'f') { ok = lit("false"); v.type = Value::Bool; v.b = false; if (!ok) fail("bad literal"); }
        else if (c == 'n') { ok = lit("null"); v.type = Value::Null; if (!ok) fail("bad literal"); }
        else ok = number(v);
        --depth;
        return ok;
    }
    bool array(Value &v) {
        v.type = Value::Array;
        ++i;
        ws();
        if (i < s.size() && s[i] == ']') { ++i; return true; }
        while (true) {
            Value e;
            if (!value(e)) return false;
            v.arr.push_back(std::move(e));
            ws();
            if (i < s.size() && s[i] == ',') { ++i; continue; }
            if (i < s.size() && s[i] == ']') { ++i; return true; }
            return fail("expected ',' or ']'");
        }
    }
    bool object(Value &v) {
        v.type = Value::Object;
        ++i;
        ws();
        if (i < s.size() && s[i] == '}') { ++i; return true; }
        while (true) {
            ws();
            std::string key;
            if (!string(key)) return false;
            ws();
            if (i >= s.size() || s[i] != ':') return fail("expected ':'");
            ++i;
            Value e;
            if (!value(e)) return false;
            v.obj.emplace_back(std::move(key), std::move(e));
            ws();
            if (i < s.size() && s[i] == ',') { ++i; continue; }
            if (i < s.size() && s[i] == '}') { ++i; return true; }
            return fail("expected ',' or '}'");
        }
    }
};

} // namespace

size_t Value::count(const std::string &key) const {
    if (type != Object) return 0;
    for (const auto &kv : obj)
        if (kv.first == key) return 1;
    return 0;
}

const Value &Value::operator[](const std::string &key) const {
    if (type == Object)
        for (const auto &kv : obj)
            if (kv.first == key) return kv.second;
    return kNull;
}

const Value &Value::operator[](size_t k) const {
    if (type == Array && k < arr.size()) return arr[k];
    return kNull;
}

bool parse(const std::string &text, Value &out, std::string &err) {
    Parser p(text);
    Value v;
    if (!p.value(v)) {
        err = p.err;
        return false;
    }
    p.ws();
    if (p.i != text.size()) {
        p.fail("trailing characters");
        err = p.err;
        return false;
    }
    out = std::move(v);
    return true;
}

bool parse_file(const std::string &path, Value &out, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        err = "could not open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), out, err);
}

} // namespace lt_json
