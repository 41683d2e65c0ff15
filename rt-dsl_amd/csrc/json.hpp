// json.hpp -- a small JSON reader for the library's text inputs (shader-IR
// programs, schedule / wisdom files): objects, arrays, strings with the usual
// escapes, numbers, true / false / null.
#pragma once
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace gprt {

// ---------------------------------------------------------------------------
// a small JSON reader: objects, arrays, strings (with the usual escapes),
// numbers, true / false / null
// ---------------------------------------------------------------------------
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* key) const {
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* s;
    size_t n, i = 0;
    std::string err;
    JParser(const char* src, size_t len) : s(src), n(len) {}
    void ws() {
        while (i < n && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
    }
    bool fail_at(const char* what) {
        if (err.empty()) err = std::string(what) + " at byte " + std::to_string(i);
        return false;
    }
    bool lit(const char* w) {
        size_t m = strlen(w);
        if (i + m > n || strncmp(s + i, w, m) != 0) return false;
        i += m;
        return true;
    }
    bool string(std::string& out) {
        if (i >= n || s[i] != '"') return fail_at("expected a string");
        ++i;
        while (i < n && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= n) return fail_at("bad escape");
                char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case '"': case '\\': case '/': out += e; break;
                    case 'u': {
                        if (i + 4 > n) return fail_at("bad \\u escape");
                        unsigned v = (unsigned)strtoul(std::string(s + i, 4).c_str(), nullptr, 16);
                        i += 4;
                        if (v < 0x80) out += (char)v;
                        else return fail_at("non-ASCII \\u escape");
                        break;
                    }
                    default: return fail_at("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= n) return fail_at("unterminated string");
        ++i;
        return true;
    }
    bool value(JVal& v, int depth = 0) {
        if (depth > 256) return fail_at("nesting too deep");
        ws();
        if (i >= n) return fail_at("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.kind = JVal::Obj;
            ++i;
            ws();
            if (i < n && s[i] == '}') { ++i; return true; }
            while (true) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (i >= n || s[i] != ':') return fail_at("expected ':'");
                ++i;
                JVal item;
                if (!value(item, depth + 1)) return false;
                v.obj.emplace_back(std::move(key), std::move(item));
                ws();
                if (i < n && s[i] == ',') { ++i; continue; }
                if (i < n && s[i] == '}') { ++i; return true; }
                return fail_at("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::Arr;
            ++i;
            ws();
            if (i < n && s[i] == ']') { ++i; return true; }
            while (true) {
                JVal item;
                if (!value(item, depth + 1)) return false;
                v.arr.push_back(std::move(item));
                ws();
                if (i < n && s[i] == ',') { ++i; continue; }
                if (i < n && s[i] == ']') { ++i; return true; }
                return fail_at("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return string(v.str);
        }
        if (lit("true")) { v.kind = JVal::Bool; v.b = true; return true; }
        if (lit("false")) { v.kind = JVal::Bool; v.b = false; return true; }
        if (lit("null")) { v.kind = JVal::Null; return true; }
        if (c == '-' || (c >= '0' && c <= '9')) {
            char* end = nullptr;
            v.kind = JVal::Num;
            v.num = strtod(s + i, &end);
            if (end == s + i) return fail_at("bad number");
            i = (size_t)(end - s);
            return true;
        }
        return fail_at("unexpected character");
    }
};

}  // namespace gprt
