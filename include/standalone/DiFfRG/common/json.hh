// DiFfRG/common/json.hh -- JSONValue: the "/section/key" parameter lookups the integrator and
// provider constructors read (reference common/json.hh: get_double/get_int/get_uint/...).
// This build keeps the lookup interface over a flat key -> value map; nested JSON documents are
// flattened into "/a/b" keys by the caller (or by JSONValue::from_string for "key=value" lists).
#pragma once

#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace DiFfRG
{
  class JSONValue
  {
  public:
    JSONValue() = default;
    JSONValue(std::initializer_list<std::pair<const std::string, std::string>> kv) : values(kv) {}

    // "key=value" entries separated by whitespace or ';', e.g. "/physical/T=0.1 /integration/cudathreadsperblock=256"
    static JSONValue from_string(const std::string &s)
    {
      JSONValue j;
      std::string tok;
      std::istringstream is(s);
      while (std::getline(is, tok, ';')) {
        std::istringstream ts(tok);
        std::string kv;
        while (ts >> kv) {
          const auto p = kv.find('=');
          if (p == std::string::npos) throw std::runtime_error("JSONValue: entry without '=': " + kv);
          j.values[kv.substr(0, p)] = kv.substr(p + 1);
        }
      }
      return j;
    }

    void set(const std::string &key, const std::string &value) { values[key] = value; }
    void set(const std::string &key, double value) { values[key] = std::to_string(value); }

    double get_double(const std::string &key) const { return std::stod(at(key)); }
    int get_int(const std::string &key) const { return std::stoi(at(key)); }
    unsigned int get_uint(const std::string &key) const { return (unsigned int)std::stoul(at(key)); }
    std::string get_string(const std::string &key) const { return at(key); }
    bool get_bool(const std::string &key) const
    {
      const std::string v = at(key);
      return v == "true" || v == "1";
    }
    double get_double(const std::string &key, const double def) const { return has(key) ? get_double(key) : def; }
    int get_int(const std::string &key, const int def) const { return has(key) ? get_int(key) : def; }
    unsigned int get_uint(const std::string &key, const unsigned int def) const
    {
      return has(key) ? get_uint(key) : def;
    }
    std::string get_string(const std::string &key, const std::string &def) const { return has(key) ? at(key) : def; }
    bool has(const std::string &key) const { return values.count(key) > 0; }

  private:
    const std::string &at(const std::string &key) const
    {
      const auto it = values.find(key);
      if (it == values.end()) throw std::runtime_error("JSONValue: key " + key + " not found");
      return it->second;
    }
    std::map<std::string, std::string> values;
  };
} // namespace DiFfRG
