// lsmf_error.hpp -- thread-local last-error message for the C ABI.
// The reference reports errors by AssertThrow -> dealii::ExceptionBase caught
// in main (src/main.cc:131-156); across a C ABI the same information travels
// as a status code plus this message.
#pragma once

#include <string>

namespace lsmf {

std::string &last_error_string();

inline int fail(int code, const std::string &msg) {
  last_error_string() = msg;
  return code;
}

}  // namespace lsmf
