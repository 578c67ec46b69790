// Error reporting shared by the C ABI translation units: every extern "C"
// entry point runs its body through abi_guarded, which turns an exception
// into a nonzero return code and keeps the message for mfem_last_error().
#pragma once
#include <exception>
#include <string>

namespace mfem {

// Defined in capi.cpp (the thread's last error message).
void set_last_error(const std::string& msg);

template <typename F>
int abi_guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    set_last_error(e.what());
  } catch (...) {
    set_last_error("unknown error");
  }
  return 1;
}

}  // namespace mfem
