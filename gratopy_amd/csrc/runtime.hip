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

Tuning &tuning() {
  static Tuning t;
  return t;
}

}  // namespace gt

extern "C" int gt_version(void) { return 10000; }

extern "C" int gt_tune_set(int key, int value) {
  gt::Tuning &t = gt::tuning();
  switch (key) {
    case GT_TUNE_FWD_GROUP:
      if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
        return gt::fail(GT_EINVAL, "forward group size must be 1, 2, 4, 8, 16 or 32");
      t.fwd_group = value;
      return GT_OK;
    case GT_TUNE_BACK_ROWS:
      if (value != 1 && value != 2 && value != 4 && value != 8)
        return gt::fail(GT_EINVAL, "backprojection rows per lane must be 1, 2, 4 or 8");
      t.back_rows = value;
      return GT_OK;
    case GT_TUNE_FWD_BANDS:
      if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
        return gt::fail(GT_EINVAL, "forward row bands must be 1, 2, 4, 8, 16 or 32");
      t.fwd_bands = value;
      return GT_OK;
    case GT_TUNE_BACK_BANDS:
      if (value != 1 && value != 2 && value != 4 && value != 8)
        return gt::fail(GT_EINVAL, "backprojection angle bands must be 1, 2, 4 or 8");
      t.back_bands = value;
      return GT_OK;
    case GT_TUNE_BACK_STAGE:
      if (value != 0 && value != 8 && value != 16)
        return gt::fail(GT_EINVAL, "backprojection staging chunk must be 0, 8 or 16 angles");
      t.back_stage = value;
      return GT_OK;
    case GT_TUNE_AUTO:
      if (value != 0 && value != 1) return gt::fail(GT_EINVAL, "automatic launch shapes: 0 or 1");
      t.automatic = value;
      return GT_OK;
    case GT_TUNE_FWD_CHUNK:
      if (value != 1 && value != 4)
        return gt::fail(GT_EINVAL, "forward chunk must be 1 or 4");
      t.fwd_chunk = value;
      return GT_OK;
    case GT_TUNE_FWD_COMPACT:
      if (value != 0 && value != 1) return gt::fail(GT_EINVAL, "forward compaction: 0 or 1");
      t.fwd_compact = value;
      return GT_OK;
    case GT_TUNE_BACK_TILE:
      if (value != 16 && value != 32)
        return gt::fail(GT_EINVAL, "staged backprojection tile width must be 16 or 32");
      t.back_tile = value;
      return GT_OK;
    default:
      return gt::fail(GT_EINVAL, "unknown tuning key");
  }
}

extern "C" int gt_tune_get(int key) {
  const gt::Tuning &t = gt::tuning();
  switch (key) {
    case GT_TUNE_FWD_GROUP:
      return t.fwd_group;
    case GT_TUNE_BACK_ROWS:
      return t.back_rows;
    case GT_TUNE_FWD_BANDS:
      return t.fwd_bands;
    case GT_TUNE_BACK_BANDS:
      return t.back_bands;
    case GT_TUNE_BACK_STAGE:
      return t.back_stage;
    case GT_TUNE_BACK_TILE:
      return t.back_tile;
    case GT_TUNE_FWD_CHUNK:
      return t.fwd_chunk;
    case GT_TUNE_AUTO:
      return t.automatic;
    case GT_TUNE_FWD_COMPACT:
      return t.fwd_compact;
    default:
      return -1;
  }
}

extern "C" size_t gt_last_error(char *buf, size_t n) {
  const std::string &msg = gt::g_last_error;
  if (buf && n) {
    const size_t k = msg.size() < n - 1 ? msg.size() : n - 1;
    std::memcpy(buf, msg.data(), k);
    buf[k] = '\0';
  }
  return msg.size();
}
