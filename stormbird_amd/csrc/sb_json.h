// Small JSON document model for setup strings (host only).
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
    std::string text() {
        if (*cur_ != '"') error("expected string");
        ++cur_;
        std::string r;
        for (;;) {
            char c = *cur_++;
            if (c == '\0') error("unterminated string");
            if (c == '"') break;
            if (c == '\\') {
                char e = *cur_++;
                const char* from = "\"\\/bfnrt";
                const char* to = "\"\\/\b\f\n\r\t";
                const char* hit = e ? strchr(from, e) : nullptr;
                if (!hit) error("unsupported escape sequence");
                r.push_back(to[hit - from]);
            } else {
                r.push_back(c);
            }
        }
        return r;
    }
    std::unique_ptr<Node> node() {
        skip();
        auto n = std::make_unique<Node>();
        if (take('{')) {
            n->t = Node::Obj;
            if (take('}')) return n;
            do {
                skip();
                n->keys.push_back(text());
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
            char* end = nullptr;
            n->t = Node::Num;
            n->n = strtod(cur_, &end);
            if (end == cur_) error("expected a value");
            cur_ = end;
        }
        return n;
    }
};

inline std::unique_ptr<Node> parse(const char* s) { return Reader(s).document(); }

}  // namespace json
}  // namespace sb
