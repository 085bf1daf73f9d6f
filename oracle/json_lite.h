// Minimal JSON reader for the oracle (test infrastructure only).
//
// Parses the subset of JSON that serde_json accepts for Stormbird's setup
// strings: objects, arrays, numbers (strtod, correctly rounded), strings,
// true/false/null.  Object key order is preserved.  Errors throw
// std::runtime_error with a position.
#pragma once

#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace orc_json {

struct Value;
using ValuePtr = std::shared_ptr<Value>;

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<ValuePtr> arr;
    std::vector<std::pair<std::string, ValuePtr>> obj;

    const Value* get(const std::string& key) const {
        for (auto& kv : obj)
            if (kv.first == key) return kv.second.get();
        return nullptr;
    }
};

class Parser {
public:
    explicit Parser(const char* text) : s_(text), p_(text) {}

    ValuePtr parse() {
        ValuePtr v = value();
        ws();
        if (*p_ != '\0') fail("trailing characters");
        return v;
    }

private:
    const char* s_;
    const char* p_;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("JSON parse error at offset ") +
                                 std::to_string(p_ - s_) + ": " + what);
    }
    void ws() {
        while (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r') ++p_;
    }
    ValuePtr value() {
        ws();
        auto v = std::make_shared<Value>();
        char c = *p_;
        if (c == '{') {
            v->kind = Value::Object;
            ++p_;
            ws();
            if (*p_ == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (*p_ != '"') fail("expected key");
                std::string k = string_lit();
                ws();
                if (*p_ != ':') fail("expected ':'");
                ++p_;
                v->obj.emplace_back(k, value());
                ws();
                if (*p_ == ',') { ++p_; continue; }
                if (*p_ == '}') { ++p_; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v->kind = Value::Array;
            ++p_;
            ws();
            if (*p_ == ']') { ++p_; return v; }
            for (;;) {
                v->arr.push_back(value());
                ws();
                if (*p_ == ',') { ++p_; continue; }
                if (*p_ == ']') { ++p_; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v->kind = Value::String;
            v->str = string_lit();
        } else if (!strncmp(p_, "true", 4)) {
            v->kind = Value::Bool; v->b = true; p_ += 4;
        } else if (!strncmp(p_, "false", 5)) {
            v->kind = Value::Bool; v->b = false; p_ += 5;
        } else if (!strncmp(p_, "null", 4)) {
            v->kind = Value::Null; p_ += 4;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            char* end = nullptr;
            v->kind = Value::Number;
            v->num = std::strtod(p_, &end);
            if (end == p_) fail("bad number");
            p_ = end;
        } else {
            fail("unexpected character");
        }
        return v;
    }
    std::string string_lit() {
        std::string out;
        ++p_;  // opening quote
        while (*p_ && *p_ != '"') {
            if (*p_ == '\\') {
                ++p_;
                switch (*p_) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    default: fail("unsupported escape");
                }
                ++p_;
            } else {
                out += *p_++;
            }
        }
        if (*p_ != '"') fail("unterminated string");
        ++p_;
        return out;
    }
};

inline ValuePtr parse(const char* text) { return Parser(text).parse(); }

}  // namespace orc_json
