// Minimal command-line option parser with the conventions of the reference's
// Boost.Program_options usage (common/main.cpp:53-70): "--name value",
// "--name=value", "-x value", "-xvalue", flags without values, required
// options, defaults, and a printable option table.
#pragma once

#include <cstdint>
#include <map>
#include <sstream>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

namespace nlc_cli {

struct OptionSpec {
	std::string name;  // long name ("" for short-only options such as -m)
	char letter = 0;   // short name (0 if none)
	bool takes_value = true;
	bool required = false;
	std::string default_value; // shown and used when absent ("" = none)
	std::string help;
};

class Options {
public:
	explicit Options(std::string caption) : caption_(std::move(caption)) {}

	Options &add(const std::string &name, char letter, bool takes_value, bool required, const std::string &help,
	             const std::string &default_value = "") {
		specs_.push_back({name, letter, takes_value, required, default_value, help});
		return *this;
	}

	// key of an option: its long name, or "-x" for short-only options
	static std::string key(const OptionSpec &s) { return s.name.empty() ? std::string("-") + s.letter : s.name; }

	void parse(int argc, char **argv) {
		for (int i = 1; i < argc; ++i) {
			std::string arg = argv[i];
			const OptionSpec *spec = nullptr;
			std::string value;
			bool has_value = false;
			if (arg.rfind("--", 0) == 0) {
				std::string name = arg.substr(2);
				auto eq = name.find('=');
				if (eq != std::string::npos) value = name.substr(eq + 1), name = name.substr(0, eq), has_value = true;
				for (auto &s : specs_)
					if (!s.name.empty() && s.name == name) spec = &s;
				if (!spec) throw std::invalid_argument("unrecognised option '" + arg + "'");
			} else if (arg.size() >= 2 && arg[0] == '-') {
				for (auto &s : specs_)
					if (s.letter == arg[1]) spec = &s;
				if (!spec) throw std::invalid_argument("unrecognised option '" + arg + "'");
				if (arg.size() > 2) value = arg.substr(2), has_value = true;
			} else {
				throw std::invalid_argument("unexpected positional argument '" + arg + "'");
			}
			const std::string k = key(*spec);
			if (spec->takes_value) {
				if (!has_value) {
					if (i + 1 >= argc) throw std::invalid_argument("the required argument for option '" + display(*spec) +
					                                               "' is missing");
					value = argv[++i];
				}
			} else if (has_value) {
				throw std::invalid_argument("option '" + display(*spec) + "' does not take any arguments");
			}
			if (values_.count(k)) throw std::invalid_argument("option '" + display(*spec) + "' cannot be specified more than once");
			values_[k] = value;
		}
		for (auto &s : specs_) {
			const std::string k = key(s);
			if (!values_.count(k) && !s.default_value.empty()) values_[k] = s.default_value, defaulted_[k] = true;
			if (s.required && !values_.count(k))
				throw std::invalid_argument("the option '" + display(s) + "' is required but missing");
		}
	}

	bool count(const std::string &k) const { return values_.count(k) && !defaulted_.count(k); }
	bool has(const std::string &k) const { return values_.count(k) != 0; }

	template <class T>
	T get(const std::string &k) const {
		auto it = values_.find(k);
		if (it == values_.end()) throw std::logic_error("option " + k + " not set");
		return convert<T>(k, it->second);
	}

	std::string usage() const {
		std::ostringstream o;
		o << caption_ << ":\n";
		for (auto &s : specs_) {
			std::string lhs = "  ";
			if (!s.name.empty()) {
				lhs += "--" + s.name;
				if (s.letter) lhs += std::string(" [ -") + s.letter + " ]";
			} else {
				lhs += std::string("-") + s.letter;
			}
			if (s.takes_value) lhs += " arg";
			if (!s.default_value.empty()) lhs += " (=" + s.default_value + ")";
			if (lhs.size() < 32) lhs.resize(32, ' ');
			else lhs += ' ';
			o << lhs << s.help << "\n";
		}
		return o.str();
	}

private:
	std::string display(const OptionSpec &s) const { return s.name.empty() ? std::string("-") + s.letter : "--" + s.name; }

	template <class T>
	T convert(const std::string &k, const std::string &v) const {
		std::istringstream in(v);
		T out{};
		if constexpr (std::is_unsigned<T>::value) {
			if (!v.empty() && v[0] == '-') bad(k, v);
		}
		in >> out;
		if (in.fail() || !in.eof()) bad(k, v);
		return out;
	}
	[[noreturn]] void bad(const std::string &k, const std::string &v) const {
		throw std::invalid_argument("the argument ('" + v + "') for option '" + k + "' is invalid");
	}

	std::string caption_;
	std::vector<OptionSpec> specs_;
	std::map<std::string, std::string> values_;
	std::map<std::string, bool> defaulted_;
};

template <>
inline std::string Options::convert<std::string>(const std::string &, const std::string &v) const {
	return v;
}

} // namespace nlc_cli
