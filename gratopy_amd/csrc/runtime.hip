// Error state and version of the C ABI.
#include <cstring>

#include "common.h"

namespace gt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

int check_launch(const char *what) {
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    return fail(GT_EHIP, std::string(what) + ": " + hipGetErrorString(err));
  }
  return GT_OK;
}

}  // namespace gt

extern "C" int gt_version(void) { return 10000; }

extern "C" size_t gt_last_error(char *buf, size_t n) {
  const std::string &msg = gt::g_last_error;
  if (buf && n) {
    const size_t k = msg.size() < n - 1 ? msg.size() : n - 1;
    std::memcpy(buf, msg.data(), k);
    buf[k] = '\0';
  }
  return msg.size();
}
