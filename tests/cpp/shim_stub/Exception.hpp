// Stand-in for the reference's Exception.hpp: tk::Exception and the Throw
// macro with the same call form (Exception.hpp:40-41).
#pragma once
#include <stdexcept>
#include <string>

namespace tk {

class Exception : public std::runtime_error {
 public:
  Exception(const std::string& msg, const char* file, const char* fn, int line)
      : std::runtime_error(msg), m_where(std::string(file) + ":" + std::to_string(line) +
                                         " in " + fn) {}
  const std::string& where() const { return m_where; }

 private:
  std::string m_where;
};

}  // namespace tk

#define Throw(...) throw tk::Exception(__VA_ARGS__, __FILE__, __PRETTY_FUNCTION__, __LINE__)
