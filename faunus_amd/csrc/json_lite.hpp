// Minimal JSON document model and parser used to read the reference's input
// format (the `energy:`, `atomlist:`, `moleculelist:` and `geometry:` blocks after
// YAML->JSON conversion).  Objects keep insertion order because the reference's
// energy lists and pair-potential lists are order dependent (terms are summed in
// the order they are given).
#pragma once
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace jsonlite {

class Value {
  public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Value() = default;
    static Value number(double v) { Value x; x.type_ = Type::Number; x.num_ = v; return x; }
    static Value boolean(bool v) { Value x; x.type_ = Type::Bool; x.bool_ = v; return x; }
    static Value string(std::string s) { Value x; x.type_ = Type::String; x.str_ = std::move(s); return x; }
    static Value array() { Value x; x.type_ = Type::Array; return x; }
    static Value object() { Value x; x.type_ = Type::Object; return x; }

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }

    double as_number() const {
        if (type_ == Type::Number) return num_;
        if (type_ == Type::Bool) return bool_ ? 1.0 : 0.0;
        throw std::runtime_error("json: number expected");
    }
    bool as_bool() const {
        if (type_ == Type::Bool) return bool_;
        if (type_ == Type::Number) return num_ != 0.0;
        throw std::runtime_error("json: bool expected");
    }
    const std::string& as_string() const {
        if (type_ != Type::String) throw std::runtime_error("json: string expected");
        return str_;
    }
    const std::vector<Value>& items() const {
        if (type_ != Type::Array) throw std::runtime_error("json: array expected");
        return arr_;
    }
    std::vector<Value>& items() {
        if (type_ != Type::Array) throw std::runtime_error("json: array expected");
        return arr_;
    }
    const std::vector<std::pair<std::string, Value>>& members() const {
        if (type_ != Type::Object) throw std::runtime_error("json: object expected");
        return obj_;
    }
    size_t size() const {
        if (type_ == Type::Array) return arr_.size();
        if (type_ == Type::Object) return obj_.size();
        return 0;
    }
    bool contains(const std::string& key) const {
        if (type_ != Type::Object) return false;
        for (auto& kv : obj_)
            if (kv.first == key) return true;
        return false;
    }
    const Value& at(const std::string& key) const {
        if (type_ != Type::Object) throw std::runtime_error("json: object expected for key '" + key + "'");
        for (auto& kv : obj_)
            if (kv.first == key) return kv.second;
        throw std::runtime_error("json: missing key '" + key + "'");
    }
    const Value& at(size_t i) const { return items().at(i); }
    double value(const std::string& key, double fallback) const {
        return contains(key) ? at(key).as_number() : fallback;
    }
    bool value(const std::string& key, bool fallback) const {
        return contains(key) ? at(key).as_bool() : fallback;
    }
    std::string value(const std::string& key, const char* fallback) const {
        return contains(key) ? at(key).as_string() : std::string(fallback);
    }
    void push_back(Value v) { items().push_back(std::move(v)); }
    void set(const std::string& key, Value v) {
        if (type_ != Type::Object) throw std::runtime_error("json: object expected");
        for (auto& kv : obj_)
            if (kv.first == key) { kv.second = std::move(v); return; }
        obj_.emplace_back(key, std::move(v));
    }

  private:
    Type type_ = Type::Null;
    bool bool_ = false;
    double num_ = 0.0;
    std::string str_;
    std::vector<Value> arr_;
    std::vector<std::pair<std::string, Value>> obj_;
};

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = parse_value();
        skip_ws();
        if (pos_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string& s_;
    size_t pos_ = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("json parse error at offset " + std::to_string(pos_) + ": " + what);
    }
    void skip_ws() {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\n' || s_[pos_] == '\t' || s_[pos_] == '\r'))
            ++pos_;
    }
    char peek() {
        skip_ws();
        if (pos_ >= s_.size()) fail("unexpected end");
        return s_[pos_];
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        ++pos_;
    }
    bool match_word(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(pos_, n, w) == 0) { pos_ += n; return true; }
        return false;
    }
    Value parse_value() {
        char c = peek();
        if (c == '{') return parse_object();
        if (c == '[') return parse_array();
        if (c == '"') return Value::string(parse_string());
        if (match_word("true")) return Value::boolean(true);
        if (match_word("false")) return Value::boolean(false);
        if (match_word("null")) return Value();
        if (match_word("Infinity")) return Value::number(INFINITY);
        if (match_word("-Infinity")) return Value::number(-INFINITY);
        if (match_word("NaN")) return Value::number(NAN);
        return parse_number();
    }
    Value parse_number() {
        const char* begin = s_.c_str() + pos_;
        char* end = nullptr;
        double v = std::strtod(begin, &end);
        if (end == begin) fail("invalid value");
        pos_ += static_cast<size_t>(end - begin);
        return Value::number(v);
    }
    std::string parse_string() {
        expect('"');
        std::string out;
        while (true) {
            if (pos_ >= s_.size()) fail("unterminated string");
            char c = s_[pos_++];
            if (c == '"') break;
            if (c == '\\') {
                if (pos_ >= s_.size()) fail("bad escape");
                char e = s_[pos_++];
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
                    if (pos_ + 4 > s_.size()) fail("bad unicode escape");
                    unsigned code = std::stoul(s_.substr(pos_, 4), nullptr, 16);
                    pos_ += 4;
                    if (code < 0x80) out += static_cast<char>(code);
                    else if (code < 0x800) {
                        out += static_cast<char>(0xC0 | (code >> 6));
                        out += static_cast<char>(0x80 | (code & 0x3F));
                    } else {
                        out += static_cast<char>(0xE0 | (code >> 12));
                        out += static_cast<char>(0x80 | ((code >> 6) & 0x3F));
                        out += static_cast<char>(0x80 | (code & 0x3F));
                    }
                    break;
                }
                default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    Value parse_array() {
        expect('[');
        Value v = Value::array();
        if (peek() == ']') { ++pos_; return v; }
        while (true) {
            v.push_back(parse_value());
            char c = peek();
            ++pos_;
            if (c == ']') break;
            if (c != ',') fail("expected ',' or ']'");
        }
        return v;
    }
    Value parse_object() {
        expect('{');
        Value v = Value::object();
        if (peek() == '}') { ++pos_; return v; }
        while (true) {
            if (peek() != '"') fail("expected key");
            std::string key = parse_string();
            expect(':');
            v.set(key, parse_value());
            char c = peek();
            ++pos_;
            if (c == '}') break;
            if (c != ',') fail("expected ',' or '}'");
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// The reference writes single-key objects such as {"lennardjones": {...}} inside
// lists; this returns that key/value pair (reference: aux/json_support jsonSingleItem).
inline const std::pair<std::string, Value>& single_item(const Value& v) {
    if (!v.is_object() || v.size() != 1) throw std::runtime_error("json: single-key object expected");
    return v.members().front();
}

} // namespace jsonlite
