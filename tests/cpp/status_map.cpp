// Every exception kind host code can raise maps to a status code (no exception
// may leave an extern "C" entry point).  Built and run by test_native_cpu.py.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../spherocyls_amd/csrc/sc_status.h"

template <class F>
static int guarded(F&& f, std::string* msg) {
  try {
    f();
    return SC_OK;
  } catch (...) {
    return sc::current_exception_status(msg);
  }
}

int main() {
  int bad = 0;
  auto expect = [&](const char* what, int got, int want, const std::string& msg) {
    std::printf("%s -> %d (%s)\n", what, got, msg.c_str());
    if (got != want || msg.empty()) ++bad;
  };
  std::string m;
  expect("bad_alloc", guarded([] { throw std::bad_alloc(); }, &m), SC_ERR_HOST, m);
  expect("length_error", guarded([] {
    std::vector<double> v;
    v.reserve(v.max_size() + 1);  // throws std::length_error
  }, &m), SC_ERR_HOST, m);
  expect("runtime_error", guarded([] { throw std::runtime_error("x"); }, &m), SC_ERR_HOST, m);
  expect("unknown", guarded([] { throw 42; }, &m), SC_ERR_HOST, m);
  expect("status", guarded([] { throw sc::StatusError{SC_ERR_STATE, "state"}; }, &m),
         SC_ERR_STATE, m);
  return bad;
}
