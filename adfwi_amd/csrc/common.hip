// ABI-wide entry points: version, error strings, fault injection.
#include "common.hpp"

namespace adfwi {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

int fault_injected(const char *entry) {
  const char *v = std::getenv("ADFWI_FAULT_INJECT");
  if (!v || !*v) return ADFWI_OK;
  if (std::strcmp(v, "all") == 0 || std::strcmp(v, entry) == 0) {
    set_error("fault injected into %s (ADFWI_FAULT_INJECT=%s)", entry, v);
    return ADFWI_E_FAULT_INJECTED;
  }
  return ADFWI_OK;
}

}  // namespace adfwi

extern "C" {

int adfwi_abi_version(void) { return ADFWI_ABI_VERSION; }

const char *adfwi_strerror(int code) {
  switch (code) {
    case ADFWI_OK: return "ok";
    case ADFWI_E_VERSION: return "ABI version or struct size mismatch";
    case ADFWI_E_ARG: return "invalid argument";
    case ADFWI_E_LAUNCH: return "HIP launch/runtime failure";
    case ADFWI_E_FAULT_INJECTED: return "fault injected";
    default: return "unknown error";
  }
}

const char *adfwi_last_error(void) { return adfwi::last_error(); }

}  // extern "C"
