// Minimal JSON reader for simulation_config.json and flow-checkpoint metadata:
// objects (insertion order kept), arrays, numbers, strings with the standard
// escapes, true/false/null.  Errors name the byte offset.
#pragma once

#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tfi_host {

struct Json {
	enum class Kind { Null, Bool, Number, String, Array, Object };
	Kind kind = Kind::Null;
	bool boolean = false;
	double number = 0.0;
	std::string string;
	std::vector<Json> array;
	std::vector<std::pair<std::string, Json>> object;

	const Json* Find(const std::string& key) const
	{
		for (const auto& item : object)
			if (item.first == key) return &item.second;
		return nullptr;
	}
};

class JsonReader {
public:
	explicit JsonReader(const std::string& text) : text_(text) {}

	Json Parse()
	{
		Json value = Value();
		Skip();
		if (pos_ != text_.size()) Fail("trailing characters");
		return value;
	}

private:
	const std::string& text_;
	std::size_t pos_ = 0;

	[[noreturn]] void Fail(const std::string& what) const
	{
		throw std::runtime_error("JSON: " + what + " at byte " + std::to_string(pos_));
	}
	void Skip()
	{
		while (pos_ < text_.size() && (text_[pos_] == ' ' || text_[pos_] == '\n' || text_[pos_] == '\r'
			|| text_[pos_] == '\t'))
			++pos_;
	}
	bool Take(char c)
	{
		Skip();
		if (pos_ < text_.size() && text_[pos_] == c) {
			++pos_;
			return true;
		}
		return false;
	}
	void Expect(char c)
	{
		if (!Take(c)) Fail(std::string("expected '") + c + "'");
	}
	bool Literal(const char* word)
	{
		const std::string w(word);
		if (text_.compare(pos_, w.size(), w) == 0) {
			pos_ += w.size();
			return true;
		}
		return false;
	}
	static void AppendUtf8(std::string& out, unsigned code)
	{
		if (code < 0x80) out.push_back(static_cast<char>(code));
		else if (code < 0x800) {
			out.push_back(static_cast<char>(0xC0 | (code >> 6)));
			out.push_back(static_cast<char>(0x80 | (code & 0x3F)));
		} else if (code < 0x10000) {
			out.push_back(static_cast<char>(0xE0 | (code >> 12)));
			out.push_back(static_cast<char>(0x80 | ((code >> 6) & 0x3F)));
			out.push_back(static_cast<char>(0x80 | (code & 0x3F)));
		} else {
			out.push_back(static_cast<char>(0xF0 | (code >> 18)));
			out.push_back(static_cast<char>(0x80 | ((code >> 12) & 0x3F)));
			out.push_back(static_cast<char>(0x80 | ((code >> 6) & 0x3F)));
			out.push_back(static_cast<char>(0x80 | (code & 0x3F)));
		}
	}
	unsigned Hex4()
	{
		if (pos_ + 4 > text_.size()) Fail("truncated \\u escape");
		unsigned v = 0;
		for (int i = 0; i < 4; ++i) {
			const char c = text_[pos_++];
			v <<= 4;
			if (c >= '0' && c <= '9') v |= static_cast<unsigned>(c - '0');
			else if (c >= 'a' && c <= 'f') v |= static_cast<unsigned>(c - 'a' + 10);
			else if (c >= 'A' && c <= 'F') v |= static_cast<unsigned>(c - 'A' + 10);
			else Fail("invalid \\u escape");
		}
		return v;
	}
	std::string String()
	{
		Expect('"');
		std::string out;
		while (true) {
			if (pos_ >= text_.size()) Fail("unterminated string");
			const char c = text_[pos_++];
			if (c == '"') break;
			if (static_cast<unsigned char>(c) < 0x20) Fail("control character in string");
			if (c != '\\') {
				out.push_back(c);
				continue;
			}
			if (pos_ >= text_.size()) Fail("unterminated escape");
			const char e = text_[pos_++];
			switch (e) {
			case '"': out.push_back('"'); break;
			case '\\': out.push_back('\\'); break;
			case '/': out.push_back('/'); break;
			case 'b': out.push_back('\b'); break;
			case 'f': out.push_back('\f'); break;
			case 'n': out.push_back('\n'); break;
			case 'r': out.push_back('\r'); break;
			case 't': out.push_back('\t'); break;
			case 'u': {
				unsigned code = Hex4();
				if (code >= 0xD800 && code < 0xDC00) {
					if (!(Literal("\\u"))) Fail("unpaired surrogate");
					const unsigned low = Hex4();
					if (low < 0xDC00 || low >= 0xE000) Fail("invalid surrogate pair");
					code = 0x10000 + ((code - 0xD800) << 10) + (low - 0xDC00);
				}
				AppendUtf8(out, code);
				break;
			}
			default: Fail("invalid escape");
			}
		}
		return out;
	}
	Json Value()
	{
		Skip();
		if (pos_ >= text_.size()) Fail("unexpected end of input");
		Json v;
		const char c = text_[pos_];
		if (c == '{') {
			++pos_;
			v.kind = Json::Kind::Object;
			std::set<std::string> keys;
			if (Take('}')) return v;
			do {
				Skip();
				std::string key = String();
				if (!keys.insert(key).second) Fail("duplicate key '" + key + "'");
				Expect(':');
				v.object.emplace_back(std::move(key), Value());
			} while (Take(','));
			Expect('}');
		} else if (c == '[') {
			++pos_;
			v.kind = Json::Kind::Array;
			if (Take(']')) return v;
			do v.array.push_back(Value());
			while (Take(','));
			Expect(']');
		} else if (c == '"') {
			v.kind = Json::Kind::String;
			v.string = String();
		} else if (Literal("true")) {
			v.kind = Json::Kind::Bool;
			v.boolean = true;
		} else if (Literal("false")) {
			v.kind = Json::Kind::Bool;
		} else if (Literal("null")) {
			v.kind = Json::Kind::Null;
		} else {
			const std::size_t start = pos_;
			if (text_[pos_] == '-') ++pos_;
			while (pos_ < text_.size() && (std::isdigit(static_cast<unsigned char>(text_[pos_])) || text_[pos_] == '.'
				|| text_[pos_] == 'e' || text_[pos_] == 'E' || text_[pos_] == '+' || text_[pos_] == '-'))
				++pos_;
			const std::string token = text_.substr(start, pos_ - start);
			if (token.empty() || token == "-") Fail("invalid value");
			char* end = nullptr;
			v.number = std::strtod(token.c_str(), &end);
			if (end != token.c_str() + token.size() || !std::isfinite(v.number)) Fail("invalid number '" + token + "'");
			v.kind = Json::Kind::Number;
		}
		return v;
	}
};

inline Json ParseJson(const std::string& text) { return JsonReader(text).Parse(); }

// typed accessors with the context path in every error
inline const Json& RequireObject(const Json& v, const std::string& ctx)
{
	if (v.kind != Json::Kind::Object) throw std::runtime_error(ctx + " must be an object");
	return v;
}
inline const Json& RequireArray(const Json& v, const std::string& ctx)
{
	if (v.kind != Json::Kind::Array) throw std::runtime_error(ctx + " must be an array");
	return v;
}
inline double RequireNumber(const Json& v, const std::string& ctx)
{
	if (v.kind != Json::Kind::Number) throw std::runtime_error(ctx + " must be a number");
	return v.number;
}
inline int RequireInteger(const Json& v, const std::string& ctx)
{
	const double d = RequireNumber(v, ctx);
	if (std::floor(d) != d || std::abs(d) > 2147483647.0) throw std::runtime_error(ctx + " must be an integer");
	return static_cast<int>(d);
}
inline const std::string& RequireString(const Json& v, const std::string& ctx)
{
	if (v.kind != Json::Kind::String) throw std::runtime_error(ctx + " must be a string");
	return v.string;
}
inline const Json& Required(const Json& obj, const std::string& key, const std::string& ctx)
{
	const Json* v = obj.Find(key);
	if (!v) throw std::runtime_error(ctx + " requires '" + key + "'");
	return *v;
}
inline void RequireKnownKeys(const Json& obj, const std::set<std::string>& keys, const std::string& ctx)
{
	for (const auto& item : obj.object)
		if (!keys.count(item.first)) throw std::runtime_error(ctx + " has unknown key '" + item.first + "'");
}

}  // namespace tfi_host
