// Small JSON document model for setup strings (host only), strict as
// serde_json: RFC 8259 number grammar, \u escapes (with surrogate pairs), no
// control characters in strings, duplicate object keys rejected (serde's
// derived structs reject a repeated field).  Integer literals (no fraction, no
// exponent) are marked, because serde reads usize fields only from them.
//
// Numbers go through strtod (correctly rounded).  serde_json's default float
// parser, which the reference uses, is not guaranteed correctly rounded; a
// decimal that serde_json rounds differently would differ by one ulp.
#pragma once

#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sb {
namespace json {

struct Node {
    enum Type { Null, Bool, Num, Str, Arr, Obj };
    Type t = Null;
    bool b = false;
    bool is_int = false;  // number written without fraction or exponent
    double n = 0.0;
    std::string s;
    std::vector<std::unique_ptr<Node>> items;
    std::vector<std::string> keys;  // parallel to items for objects

    const Node* find(const char* k) const {
        if (t != Obj) return nullptr;
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return items[i].get();
        return nullptr;
    }
};

class Reader {
public:
    explicit Reader(const char* src) : begin_(src), cur_(src) {}
    std::unique_ptr<Node> document() {
        auto n = node();
        skip();
        if (*cur_) error("unexpected trailing text");
        return n;
    }

private:
    const char* begin_;
    const char* cur_;

    [[noreturn]] void error(const std::string& m) {
        throw std::runtime_error("invalid JSON at byte " + std::to_string(cur_ - begin_) + ": " + m);
    }
    void skip() {
        while (*cur_ && strchr(" \t\r\n", *cur_)) ++cur_;
    }
    bool take(char c) {
        skip();
        if (*cur_ == c) { ++cur_; return true; }
        return false;
    }
    unsigned hex4() {
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = *cur_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else error("invalid \\u escape");
        }
        return v;
    }
    static void utf8(std::string& r, unsigned cp) {
        if (cp < 0x80) {
            r.push_back((char)cp);
        } else if (cp < 0x800) {
            r.push_back((char)(0xC0 | (cp >> 6)));
            r.push_back((char)(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            r.push_back((char)(0xE0 | (cp >> 12)));
            r.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            r.push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            r.push_back((char)(0xF0 | (cp >> 18)));
            r.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
            r.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            r.push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    std::string text() {
        if (*cur_ != '"') error("expected string");
        ++cur_;
        std::string r;
        for (;;) {
            char c = *cur_++;
            if (c == '\0') error("unterminated string");
            if ((unsigned char)c < 0x20) error("control character in string");
            if (c == '"') break;
            if (c == '\\') {
                char e = *cur_++;
                if (e == 'u') {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {  // high surrogate: a low one must follow
                        if (cur_[0] != '\\' || cur_[1] != 'u') error("lone surrogate in \\u escape");
                        cur_ += 2;
                        unsigned lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) error("invalid surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        error("lone surrogate in \\u escape");
                    }
                    utf8(r, cp);
                    continue;
                }
                const char* from = "\"\\/bfnrt";
                const char* to = "\"\\/\b\f\n\r\t";
                const char* hit = e ? strchr(from, e) : nullptr;
                if (!hit) error("invalid escape sequence");
                r.push_back(to[hit - from]);
            } else {
                r.push_back(c);
            }
        }
        return r;
    }
    // -?(0|[1-9][0-9]*)(.[0-9]+)?([eE][+-]?[0-9]+)?
    void number(Node& n) {
        const char* p = cur_;
        bool is_int = true;
        if (*p == '-') ++p;
        if (*p == '0') ++p;
        else if (*p >= '1' && *p <= '9') while (*p >= '0' && *p <= '9') ++p;
        else error("expected a value");
        if (*p == '.') {
            is_int = false;
            ++p;
            if (!(*p >= '0' && *p <= '9')) error("expected digits after the decimal point");
            while (*p >= '0' && *p <= '9') ++p;
        }
        if (*p == 'e' || *p == 'E') {
            is_int = false;
            ++p;
            if (*p == '+' || *p == '-') ++p;
            if (!(*p >= '0' && *p <= '9')) error("expected exponent digits");
            while (*p >= '0' && *p <= '9') ++p;
        }
        std::string lit(cur_, p);
        n.t = Node::Num;
        n.is_int = is_int;
        n.n = strtod(lit.c_str(), nullptr);
        cur_ = p;
    }
    std::unique_ptr<Node> node() {
        skip();
        auto n = std::make_unique<Node>();
        if (take('{')) {
            n->t = Node::Obj;
            if (take('}')) return n;
            do {
                skip();
                std::string k = text();
                for (const auto& prev : n->keys)
                    if (prev == k) error("duplicate field `" + k + "`");
                n->keys.push_back(k);
                if (!take(':')) error("expected ':'");
                n->items.push_back(node());
            } while (take(','));
            if (!take('}')) error("expected '}'");
        } else if (take('[')) {
            n->t = Node::Arr;
            if (take(']')) return n;
            do {
                n->items.push_back(node());
            } while (take(','));
            if (!take(']')) error("expected ']'");
        } else if (*cur_ == '"') {
            n->t = Node::Str;
            n->s = text();
        } else if (strncmp(cur_, "true", 4) == 0 || strncmp(cur_, "false", 5) == 0) {
            n->t = Node::Bool;
            n->b = *cur_ == 't';
            cur_ += n->b ? 4 : 5;
        } else if (strncmp(cur_, "null", 4) == 0) {
            cur_ += 4;
        } else {
            number(*n);
        }
        return n;
    }
};

inline std::unique_ptr<Node> parse(const char* s) { return Reader(s).document(); }

}  // namespace json
}  // namespace sb
