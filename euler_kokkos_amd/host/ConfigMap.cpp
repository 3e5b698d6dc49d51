// ConfigMap.cpp -- see ConfigMap.h for the rules this parser follows.
#include "ConfigMap.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace euler_kokkos {

namespace {

constexpr int kMaxLine = 200;   // fgets chunk of the file parser (ini.cpp:21,95)
constexpr int kMaxSection = 50; // section / previous-name buffers (ini.cpp:22-23)
constexpr int kMaxName = 50;

char* rstrip(char* s)
{
  char* p = s + std::strlen(s);
  while (p > s && std::isspace((unsigned char)*--p))
    *p = '\0';
  return s;
}

char* lskip(char* s)
{
  while (*s && std::isspace((unsigned char)*s))
    s++;
  return s;
}

// first `c` or a ';' that follows whitespace, else the terminating NUL
char* find_char_or_comment(char* s, char c)
{
  bool was_space = false;
  while (*s && *s != c && !(was_space && *s == ';')) {
    was_space = std::isspace((unsigned char)*s) != 0;
    s++;
  }
  return s;
}

void copy_trunc(char* dst, const char* src, size_t size)
{
  std::strncpy(dst, src, size);
  dst[size - 1] = '\0';
}

} // namespace

std::string ConfigMap::makeKey(const std::string& section, const std::string& name)
{
  std::string key = section + "." + name;
  for (auto& ch : key)
    ch = (char)std::tolower((unsigned char)ch);
  return key;
}

void ConfigMap::handle(const char* section, const char* name, const char* value)
{
  m_values[makeKey(section, name)] = value;
}

int ConfigMap::parseLine(char* line, char* section, char* prev_name, int lineno, int error)
{
  char* start = lskip(rstrip(line));
  if (*prev_name && *start && start > line) {
    // indented continuation line: replaces the previous name's value
    handle(section, prev_name, start);
  } else if (*start == ';' || *start == '#') {
    // whole-line comment
  } else if (*start == '[') {
    char* end = find_char_or_comment(start + 1, ']');
    if (*end == ']') {
      *end = '\0';
      copy_trunc(section, start + 1, kMaxSection);
      *prev_name = '\0';
    } else if (!error) {
      error = lineno;
    }
  } else if (*start) {
    char* end = find_char_or_comment(start, '=');
    if (*end == '=') {
      *end = '\0';
      char* name = rstrip(start);
      char* value = lskip(end + 1);
      end = find_char_or_comment(value, '\0');
      if (*end == ';')
        *end = '\0';
      rstrip(value);
      copy_trunc(prev_name, name, kMaxName);
      handle(section, name, value);
    } else if (!error) {
      error = lineno;
    }
  }
  return error;
}

ConfigMap::ConfigMap(const std::string& filename)
{
  FILE* f = std::fopen(filename.c_str(), "r");
  if (!f) {
    m_error = -1;
    return;
  }
  char line[kMaxLine];
  char section[kMaxSection] = "";
  char prev_name[kMaxName] = "";
  int lineno = 0, error = 0;
  while (std::fgets(line, sizeof(line), f) != nullptr) {
    lineno++;
    error = parseLine(line, section, prev_name, lineno, error);
  }
  std::fclose(f);
  m_error = error;
}

ConfigMap::ConfigMap(const char* buffer, int buffer_size)
{
  std::istringstream in(std::string(buffer ? buffer : "", buffer ? (size_t)buffer_size : 0));
  char section[kMaxSection] = "";
  char prev_name[kMaxName] = "";
  int lineno = 0, error = 0;
  for (std::string s; std::getline(in, s);) {
    lineno++;
    std::string copy = s;
    error = parseLine(&copy[0], section, prev_name, lineno, error);
  }
  m_error = error;
}

std::string ConfigMap::getString(const std::string& section, const std::string& name,
                                 const std::string& default_value) const
{
  auto it = m_values.find(makeKey(section, name));
  return it == m_values.end() ? default_value : it->second;
}

void ConfigMap::setString(const std::string& section, const std::string& name, const std::string& value)
{
  m_values[makeKey(section, name)] = value;
}

long ConfigMap::getInteger(const std::string& section, const std::string& name, long default_value) const
{
  const std::string s = getString(section, name, "");
  const char* v = s.c_str();
  char* end;
  const long n = std::strtol(v, &end, 0);
  return end > v ? n : default_value;
}

void ConfigMap::setInteger(const std::string& section, const std::string& name, long value)
{
  setString(section, name, std::to_string(value));
}

float ConfigMap::getFloat(const std::string& section, const std::string& name, float default_value) const
{
  const std::string s = getString(section, name, "");
  const char* v = s.c_str();
  char* end;
  const float x = std::strtof(v, &end);
  return end > v ? x : default_value;
}

void ConfigMap::setFloat(const std::string& section, const std::string& name, float value)
{
  std::ostringstream ss;
  ss << value;
  setString(section, name, ss.str());
}

bool ConfigMap::getBool(const std::string& section, const std::string& name, bool default_value) const
{
  const std::string s = getString(section, name, "");
  bool val = default_value;
  if (s == "1" || s == "yes" || s == "true" || s == "on")
    val = true;
  if (s == "0" || s == "no" || s == "false" || s == "off")
    val = false;
  if (s.empty())
    val = default_value;
  return val;
}

void ConfigMap::setBool(const std::string& section, const std::string& name, bool value)
{
  setString(section, name, value ? "true" : "false");
}

std::ostream& operator<<(std::ostream& os, const ConfigMap& cfg)
{
  for (const auto& kv : cfg.m_values)
    os << kv.first << " = " << kv.second << "\n";
  return os;
}

ConfigMap broadcast_parameters(const std::string& filename)
{
  return ConfigMap(filename);
}

} // namespace euler_kokkos
