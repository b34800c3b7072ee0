// Operator kernels compiled at run time: the HIP counterpart of the OpenCL
// program build and kernel invocation of the reference's operator interface
// (gratopy/operator/opencl.py:262-305 `_get_program` / `_get_projection_kernel`,
// 444-454 `_invoke_kernel`).
//
// A plugin source is ordinary HIP C++ whose kernels carry the ABI names
// `{base}[_ad]_{float|double}_{out_order}{in_order}` with C linkage.  hiprtc
// compiles it for the device's architecture into a code object; the code object
// is loaded as a module per device; kernels are looked up by name and launched
// with one pointer per argument.  Compilation needs no GPU.

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace {

// Options every plugin is compiled with, ahead of the caller's own: full
// optimisation, and packed FP32 instructions off (the same code-generation
// choice as the library's own kernels).
const char *const kBaseOptions[] = {"-O3", "-Xclang", "-target-feature", "-Xclang",
                                    "-packed-fp32-ops"};

int rtc_fail(hiprtcResult r, const char *what, const std::string &log = std::string()) {
  std::string msg = std::string(what) + ": " + hiprtcGetErrorString(r);
  if (!log.empty()) msg += "\n" + log;
  return gt::fail(GT_EHIP, msg);
}

int hip_fail(hipError_t e, const char *what) {
  return gt::fail(GT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// Compiles `source` for `arch` (e.g. "gfx950").  The code object is copied to
// `image` when `capacity` is large enough; `*image_size` always receives its size,
// so a first call with image = NULL sizes the buffer.  On a compile error the
// compiler log is part of the error message.
extern "C" int gt_plugin_compile(const char *source, const char *name, const char *arch,
                                 const char *const *options, int n_options, void *image,
                                 size_t capacity, size_t *image_size) {
  GT_REQUIRE(source && name && arch && image_size, "gt_plugin_compile: null argument");
  GT_REQUIRE(n_options >= 0 && (n_options == 0 || options), "gt_plugin_compile: bad option list");
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, name, 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) return rtc_fail(r, "hiprtcCreateProgram");
  const std::string arch_opt = std::string("--offload-arch=") + arch;
  std::vector<const char *> opts;
  opts.push_back(arch_opt.c_str());
  for (const char *o : kBaseOptions) opts.push_back(o);
  for (int i = 0; i < n_options; ++i) opts.push_back(options[i]);
  r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      hiprtcGetProgramLog(prog, &log[0]);
      log.resize(std::strlen(log.c_str()));
    }
    hiprtcDestroyProgram(&prog);
    return rtc_fail(r, "hiprtcCompileProgram", log);
  }
  size_t size = 0;
  r = hiprtcGetCodeSize(prog, &size);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return rtc_fail(r, "hiprtcGetCodeSize");
  }
  *image_size = size;
  if (image && capacity >= size) {
    r = hiprtcGetCode(prog, static_cast<char *>(image));
    if (r != HIPRTC_SUCCESS) {
      hiprtcDestroyProgram(&prog);
      return rtc_fail(r, "hiprtcGetCode");
    }
  }
  hiprtcDestroyProgram(&prog);
  return GT_OK;
}

// Architecture name of the current device (e.g. "gfx950:sramecc+:xnack-").
extern "C" int gt_plugin_device_arch(char *buf, size_t n) {
  GT_REQUIRE(buf && n > 0, "gt_plugin_device_arch: null buffer");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  const size_t k = std::strlen(prop.gcnArchName);
  GT_REQUIRE(k < n, "gt_plugin_device_arch: buffer too small");
  std::memcpy(buf, prop.gcnArchName, k + 1);
  return GT_OK;
}

// Loads a code object on the current device.
extern "C" int gt_plugin_module_load(const void *image, void **module) {
  GT_REQUIRE(image && module, "gt_plugin_module_load: null argument");
  hipModule_t m = nullptr;
  hipError_t e = hipModuleLoadData(&m, image);
  if (e != hipSuccess) return hip_fail(e, "hipModuleLoadData");
  *module = m;
  return GT_OK;
}

// Unloads a module after the device has finished all work (kernels of the
// module may still be queued on any stream).
extern "C" int gt_plugin_module_unload(void *module) {
  GT_REQUIRE(module, "gt_plugin_module_unload: null module");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  e = hipModuleUnload(static_cast<hipModule_t>(module));
  if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
  return GT_OK;
}

// Looks up a kernel of a loaded module by its (C-linkage) name.
extern "C" int gt_plugin_function(void *module, const char *name, void **function) {
  GT_REQUIRE(module && name && function, "gt_plugin_function: null argument");
  hipFunction_t f = nullptr;
  hipError_t e = hipModuleGetFunction(&f, static_cast<hipModule_t>(module), name);
  if (e != hipSuccess) return hip_fail(e, "hipModuleGetFunction");
  *function = f;
  return GT_OK;
}

// Launches `function` over a grid of grid[0..2] blocks of block[0..2] threads on
// `stream`; args[i] points to the value of kernel argument i.
extern "C" int gt_plugin_launch(gt_stream_t stream, void *function, const unsigned *grid,
                                const unsigned *block, void **args) {
  GT_REQUIRE(function && grid && block, "gt_plugin_launch: null argument");
  GT_REQUIRE(grid[0] && grid[1] && grid[2], "gt_plugin_launch: empty grid");
  GT_REQUIRE(block[0] && block[1] && block[2] && block[0] * block[1] * block[2] <= 1024,
             "gt_plugin_launch: block must hold 1..1024 threads");
  GT_REQUIRE(grid[1] <= 65535 && grid[2] <= 65535, "gt_plugin_launch: grid too large");
  hipError_t e = hipModuleLaunchKernel(static_cast<hipFunction_t>(function), grid[0], grid[1],
                                       grid[2], block[0], block[1], block[2], 0,
                                       static_cast<hipStream_t>(stream), args, nullptr);
  if (e != hipSuccess) return hip_fail(e, "hipModuleLaunchKernel");
  return GT_OK;
}
