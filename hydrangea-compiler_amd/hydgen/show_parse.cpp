// show_parse.cpp - a reader for the text GHC's derived Show instances print.
//
// The forms it accepts are the ones `showsPrec` of a derived instance, of
// lists, tuples, strings, Integer, Double and Data.Map produce:
//   Con a1 a2 ...            constructor application (arguments at precedence
//                            11, so compound ones come parenthesised)
//   Con {f1 = v1, f2 = v2}   record syntax
//   [v1,v2]  (v1,v2)  "str"  123  (-5)  1.0e-3  (-1.5)  Infinity  NaN
//   fromList [(k,v),...]     Data.Map
// Strings use Haskell escapes (\" \\ \n \t \r \& \' and \DDD decimal).
#include <ctype.h>
#include <stdio.h>

#include "hydgen.hpp"

namespace hydgen {

const Value& Value::field(const std::string& name) const {
    for (const auto& f : fields)
        if (f.first == name) return *f.second;
    throw Error("record " + text + " has no field " + name);
}

namespace {

struct Reader {
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void fail(const std::string& what) const {
        char buf[64];
        snprintf(buf, sizeof buf, " at offset %zu", p);
        std::string ctx = s.substr(p, 40);
        throw Error("show text: " + what + buf + " near \"" + ctx + "\"");
    }
    void ws() {
        while (p < s.size() && isspace((unsigned char)s[p])) ++p;
    }
    char peek() {
        ws();
        return p < s.size() ? s[p] : '\0';
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        ++p;
    }
    bool ident_start(char c) const { return isalpha((unsigned char)c) || c == '_'; }
    std::string ident() {
        ws();
        size_t a = p;
        while (p < s.size() && (isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '\'' || s[p] == '.')) ++p;
        if (a == p) fail("expected a name");
        return s.substr(a, p - a);
    }
    // Can the next token start a constructor argument?
    bool arg_start() {
        const char c = peek();
        return c == '(' || c == '[' || c == '"' || isdigit((unsigned char)c) || (ident_start(c) && isupper((unsigned char)c));
    }

    VPtr number(bool neg) {
        ws();
        size_t a = p;
        while (p < s.size() && (isdigit((unsigned char)s[p]) || s[p] == '.' || s[p] == 'e' || s[p] == 'E' ||
                                ((s[p] == '-' || s[p] == '+') && p > a && (s[p - 1] == 'e' || s[p - 1] == 'E'))))
            ++p;
        if (a == p) fail("expected a number");
        auto v = std::make_shared<Value>();
        v->kind = Value::Num;
        v->text = (neg ? "-" : "") + s.substr(a, p - a);
        return v;
    }

    VPtr string_lit() {
        expect('"');
        std::string out;
        while (true) {
            if (p >= s.size()) fail("unterminated string");
            char c = s[p++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) fail("bad escape");
            char e = s[p++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '\'': out += '\''; break;
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case '&': break;  // empty separator
                default:
                    if (isdigit((unsigned char)e)) {
                        int code = e - '0';
                        while (p < s.size() && isdigit((unsigned char)s[p])) code = code * 10 + (s[p++] - '0');
                        if (code > 255) fail("character code beyond one byte");
                        out += (char)code;
                    } else {
                        fail(std::string("unsupported escape \\") + e);
                    }
            }
        }
        auto v = std::make_shared<Value>();
        v->kind = Value::Str;
        v->text = out;
        return v;
    }

    // An argument: a nullary constructor, or a bracketed/literal value.
    VPtr argument() {
        const char c = peek();
        if (c == '(' || c == '[' || c == '"' || isdigit((unsigned char)c)) return value();
        auto v = std::make_shared<Value>();
        v->kind = Value::Con;
        v->text = ident();
        if (peek() == '{') return record(v->text);
        return v;
    }

    VPtr record(const std::string& name) {
        expect('{');
        auto v = std::make_shared<Value>();
        v->kind = Value::Record;
        v->text = name;
        if (peek() == '}') {
            ++p;
            return v;
        }
        while (true) {
            std::string f = ident();
            expect('=');
            v->fields.emplace_back(f, value());
            if (peek() == ',') {
                ++p;
                continue;
            }
            expect('}');
            return v;
        }
    }

    VPtr value() {
        const char c = peek();
        if (c == '(') {
            ++p;
            if (peek() == ')') {  // ()
                ++p;
                auto v = std::make_shared<Value>();
                v->kind = Value::Tuple;
                return v;
            }
            VPtr first = value();
            if (peek() == ')') {
                ++p;
                return first;
            }
            auto t = std::make_shared<Value>();
            t->kind = Value::Tuple;
            t->items.push_back(first);
            while (peek() == ',') {
                ++p;
                t->items.push_back(value());
            }
            expect(')');
            return t;
        }
        if (c == '[') {
            ++p;
            auto l = std::make_shared<Value>();
            l->kind = Value::List;
            if (peek() == ']') {
                ++p;
                return l;
            }
            while (true) {
                l->items.push_back(value());
                if (peek() == ',') {
                    ++p;
                    continue;
                }
                expect(']');
                return l;
            }
        }
        if (c == '"') return string_lit();
        if (c == '-') {
            ++p;
            if (ident_start(peek())) {  // -Infinity
                auto v = std::make_shared<Value>();
                v->kind = Value::Num;
                v->text = "-" + ident();
                return v;
            }
            return number(true);
        }
        if (isdigit((unsigned char)c)) return number(false);
        if (!ident_start(c)) fail("unexpected character");
        std::string name = ident();
        if (name == "Infinity" || name == "NaN") {
            auto v = std::make_shared<Value>();
            v->kind = Value::Num;
            v->text = name;
            return v;
        }
        if (peek() == '{') return record(name);
        auto v = std::make_shared<Value>();
        v->kind = Value::Con;
        v->text = name;
        while (arg_start()) v->items.push_back(argument());
        if (name == "fromList" && v->items.size() != 1) fail("fromList takes one list");
        return v;
    }
};

}  // namespace

VPtr parse_show(const std::string& text) {
    Reader r{text};
    VPtr v = r.value();
    r.ws();
    if (r.p != text.size()) r.fail("trailing text");
    return v;
}

std::vector<VPtr> parse_show_lines(const std::string& text) {
    std::vector<VPtr> out;
    size_t a = 0;
    while (a < text.size()) {
        size_t e = text.find('\n', a);
        if (e == std::string::npos) e = text.size();
        std::string line = text.substr(a, e - a);
        bool blank = true;
        for (char c : line)
            if (!isspace((unsigned char)c)) blank = false;
        if (!blank && line[line.find_first_not_of(" \t")] != '#') out.push_back(parse_show(line));
        a = e + 1;
    }
    return out;
}

}  // namespace hydgen
