// Minimal JSON reader for the run configuration files (objects, arrays,
// numbers, strings, true/false/null). Numbers are kept as their source text and
// converted on access, so integers and floats read exactly as written.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace asph::json
{

struct Value
{
    enum class Kind
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    };
    Kind kind = Kind::Null;
    bool boolean = false;
    std::string text;  // number source text or string contents
    std::vector<Value> items;
    std::map<std::string, Value> members;

    [[nodiscard]] bool isNull() const { return kind == Kind::Null; }
    [[nodiscard]] bool has(const std::string& k) const { return kind == Kind::Object && members.count(k) != 0; }
    [[nodiscard]] const Value& at(const std::string& k) const
    {
        const auto it = members.find(k);
        if (kind != Kind::Object || it == members.end())
        {
            throw std::runtime_error("missing key \"" + k + "\"");
        }
        return it->second;
    }
    [[nodiscard]] double number() const
    {
        if (kind != Kind::Number)
        {
            throw std::runtime_error("expected a number");
        }
        return std::strtod(text.c_str(), nullptr);
    }
    [[nodiscard]] float f32() const
    {
        if (kind != Kind::Number)
        {
            throw std::runtime_error("expected a number");
        }
        return std::strtof(text.c_str(), nullptr);
    }
    [[nodiscard]] unsigned long long u64() const
    {
        if (kind != Kind::Number)
        {
            throw std::runtime_error("expected a number");
        }
        return std::strtoull(text.c_str(), nullptr, 10);
    }
    [[nodiscard]] bool flag() const
    {
        if (kind != Kind::Bool)
        {
            throw std::runtime_error("expected true or false");
        }
        return boolean;
    }
    [[nodiscard]] const std::string& str() const
    {
        if (kind != Kind::String)
        {
            throw std::runtime_error("expected a string");
        }
        return text;
    }
};

class Parser
{
public:
    explicit Parser(std::string src) : s_(std::move(src)) {}

    Value parse()
    {
        Value v = value();
        ws();
        if (p_ != s_.size())
        {
            fail("trailing characters");
        }
        return v;
    }

private:
    [[noreturn]] void fail(const std::string& what) const
    {
        std::ostringstream m;
        m << "JSON parse error at offset " << p_ << ": " << what;
        throw std::runtime_error(m.str());
    }
    void ws()
    {
        while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_])) != 0)
        {
            ++p_;
        }
    }
    bool eat(char c)
    {
        ws();
        if (p_ < s_.size() && s_[p_] == c)
        {
            ++p_;
            return true;
        }
        return false;
    }
    void expect(char c)
    {
        if (!eat(c))
        {
            fail(std::string("expected '") + c + "'");
        }
    }
    bool word(const char* w)
    {
        const std::string t(w);
        if (s_.compare(p_, t.size(), t) == 0)
        {
            p_ += t.size();
            return true;
        }
        return false;
    }
    std::string string()
    {
        expect('"');
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"')
        {
            char c = s_[p_++];
            if (c == '\\')
            {
                if (p_ >= s_.size())
                {
                    fail("bad escape");
                }
                const char e = s_[p_++];
                switch (e)
                {
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case 'r': c = '\r'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'u':
                    if (p_ + 4 > s_.size())
                    {
                        fail("bad \\u escape");
                    }
                    p_ += 4;  // non-ASCII is not used by the configuration keys
                    c = '?';
                    break;
                default: c = e; break;
                }
            }
            out.push_back(c);
        }
        if (p_ >= s_.size())
        {
            fail("unterminated string");
        }
        ++p_;
        return out;
    }
    Value value()
    {
        ws();
        if (p_ >= s_.size())
        {
            fail("unexpected end");
        }
        Value v;
        const char c = s_[p_];
        if (c == '{')
        {
            ++p_;
            v.kind = Value::Kind::Object;
            if (eat('}'))
            {
                return v;
            }
            do
            {
                ws();
                const std::string k = string();
                expect(':');
                v.members[k] = value();
            } while (eat(','));
            expect('}');
        }
        else if (c == '[')
        {
            ++p_;
            v.kind = Value::Kind::Array;
            if (eat(']'))
            {
                return v;
            }
            do
            {
                v.items.push_back(value());
            } while (eat(','));
            expect(']');
        }
        else if (c == '"')
        {
            v.kind = Value::Kind::String;
            v.text = string();
        }
        else if (word("true"))
        {
            v.kind = Value::Kind::Bool;
            v.boolean = true;
        }
        else if (word("false"))
        {
            v.kind = Value::Kind::Bool;
        }
        else if (word("null"))
        {
            v.kind = Value::Kind::Null;
        }
        else
        {
            const size_t b = p_;
            while (p_ < s_.size() && (std::isdigit(static_cast<unsigned char>(s_[p_])) != 0 || s_[p_] == '-' ||
                                      s_[p_] == '+' || s_[p_] == '.' || s_[p_] == 'e' || s_[p_] == 'E'))
            {
                ++p_;
            }
            if (b == p_)
            {
                fail("unexpected character");
            }
            v.kind = Value::Kind::Number;
            v.text = s_.substr(b, p_ - b);
        }
        return v;
    }

    std::string s_;
    size_t p_ = 0;
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace asph::json
