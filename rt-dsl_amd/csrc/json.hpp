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
                        unsigned v = 0;
                        for (int h = 0; h < 4; ++h) {
                            const char x = s[i + h];
                            const int dgt = x >= '0' && x <= '9' ? x - '0'
                                            : x >= 'a' && x <= 'f' ? x - 'a' + 10
                                            : x >= 'A' && x <= 'F' ? x - 'A' + 10 : -1;
                            if (dgt < 0) return fail_at("bad \\u escape");
                            v = v * 16 + (unsigned)dgt;
                        }
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
            // the token is copied out first: the input need not be NUL-terminated
            size_t j = i;
            while (j < n && (s[j] == '-' || s[j] == '+' || s[j] == '.' || s[j] == 'e' || s[j] == 'E' ||
                             (s[j] >= '0' && s[j] <= '9')))
                ++j;
            if (j - i > 64) return fail_at("number too long");
            char tok[72];
            memcpy(tok, s + i, j - i);
            tok[j - i] = 0;
            char* end = nullptr;
            v.kind = JVal::Num;
            v.num = strtod(tok, &end);
            if (end != tok + (j - i)) return fail_at("bad number");
            i = j;
            return true;
        }
        return fail_at("unexpected character");
    }
};

}  // namespace gprt
