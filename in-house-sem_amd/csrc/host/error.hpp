// Error type of the host library: the counterpart of insem::Error and
// INSEM_CHECK (reference: cpp/include/insem/core/error.hpp:9-25).
#pragma once

#include <sstream>
#include <stdexcept>
#include <string>

namespace insem_amd {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

}  // namespace insem_amd

#define INSEM_REQUIRE(cond, msg)                       \
  do {                                                 \
    if (!(cond)) {                                     \
      std::ostringstream insem_msg_;                   \
      insem_msg_ << msg;                               \
      throw ::insem_amd::Error(insem_msg_.str());      \
    }                                                  \
  } while (0)
