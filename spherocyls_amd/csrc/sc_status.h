// Exception-to-status mapping for the extern "C" entry points of the library.
// Host code inside the library reports failures by throwing; no exception may
// leave an extern "C" function, so every entry point ends in a catch (...)
// that calls current_exception_status().  Host-only: no HIP types here, so the
// mapping is testable on a machine without a GPU.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/spherocyls.h"

namespace sc {

// A failure with a status of its own (argument, state or capacity errors found
// by host code below the entry point).
struct StatusError {
  int status;
  std::string msg;
};

// The status and message of the exception being handled.  Call only from a
// catch block (it rethrows the in-flight exception to inspect it).
inline int current_exception_status(std::string* msg) {
  try {
    throw;
  } catch (const StatusError& e) {
    if (msg) *msg = e.msg;
    return e.status;
  } catch (const std::bad_alloc&) {
    if (msg) *msg = "host memory exhausted";
    return SC_ERR_HOST;
  } catch (const std::exception& e) {
    if (msg) *msg = std::string("internal error: ") + e.what();
    return SC_ERR_HOST;
  } catch (...) {
    if (msg) *msg = "internal error: unknown exception";
    return SC_ERR_HOST;
  }
}

}  // namespace sc
