// ConfigMap.h -- .ini parameter store with the reference's parsing rules.
//
// Same interface and semantics as euler_kokkos' ConfigMap/INIReader
// (src/utils/config/ConfigMap.{h,cpp}, src/utils/config/inih/{INIReader.cpp,ini.cpp}):
//  - keys are "section.name", lower-cased, so lookups are case-insensitive;
//  - ';' starts an inline comment only after whitespace, '#'/';' comment whole lines;
//  - an indented non-empty line continues (replaces) the previous name's value;
//  - file lines are read in 200-byte chunks (longer lines split), buffers line by line;
//  - getFloat parses with strtof (float precision), getInteger with strtol(base 0).
#pragma once

#include <map>
#include <ostream>
#include <string>

namespace euler_kokkos {

class ConfigMap {
public:
  explicit ConfigMap(const std::string& filename);
  ConfigMap(const char* buffer, int buffer_size);
  virtual ~ConfigMap() = default;

  // 0 on success, line number of the first error, or -1 if the file could not be opened
  int ParseError() const { return m_error; }

  std::string getString(const std::string& section, const std::string& name,
                        const std::string& default_value) const;
  void setString(const std::string& section, const std::string& name, const std::string& value);
  long getInteger(const std::string& section, const std::string& name, long default_value) const;
  void setInteger(const std::string& section, const std::string& name, long value);
  float getFloat(const std::string& section, const std::string& name, float default_value) const;
  void setFloat(const std::string& section, const std::string& name, float value);
  bool getBool(const std::string& section, const std::string& name, bool default_value) const;
  void setBool(const std::string& section, const std::string& name, bool value);

  const std::map<std::string, std::string>& values() const { return m_values; }
  friend std::ostream& operator<<(std::ostream& os, const ConfigMap& cfg);

private:
  static std::string makeKey(const std::string& section, const std::string& name);
  void handle(const char* section, const char* name, const char* value);
  int parseLine(char* line, char* section, char* prev_name, int lineno, int error);

  int m_error = 0;
  std::map<std::string, std::string> m_values;
};

// ConfigMap from a parameter file (broadcast_parameters in a single-process run).
ConfigMap broadcast_parameters(const std::string& filename);

} // namespace euler_kokkos
