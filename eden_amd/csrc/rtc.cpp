// rtc.cpp -- run-time compilation of the generic LEMS synapse and point-neuron kernels with
// hiprtc. The kernel template (lems_kernel.hip) and the device header it includes
// (device_model.hpp) are embedded in the library at build time (_embedded_sources.inc,
// written by eden_amd/build.py); the generated synapse and cell functions come from the caller. This replaces the reference's
// system("gcc ...") and dlopen of the emitted C (eden/Eden.cpp:6992-7216) with an in-process
// compile for the GPU.
#include "rtc.hpp"

#include <hip/hiprtc.h>

#include <map>
#include <mutex>
#include <vector>

#include "_embedded_sources.inc"

namespace edenhip {

namespace {

struct Compiled {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr, cell_fn = nullptr, ext_fn = nullptr;
};

}  // namespace

int rtc_lems_kernel(const std::string &generated, hipFunction_t *fn, hipFunction_t *cell_fn, hipFunction_t *ext_fn,
                    std::string &err) {
    static std::mutex mu;
    static std::map<std::string, Compiled> cache;   // key: device arch + '\n' + source
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        err = "hipGetDevice failed";
        return -1;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        err = "hipGetDeviceProperties failed";
        return -1;
    }
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    const std::string key = std::to_string(dev) + ":" + arch + "\n" + generated;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *fn = it->second.fn;
        *cell_fn = it->second.cell_fn;
        *ext_fn = it->second.ext_fn;
        return 0;
    }
    hiprtcProgram prog;
    // hiprtc declares the HIP runtime itself, and the fixed-width integer types in its
    // __hip_internal namespace: the two system headers device_model.hpp includes are stood in
    // for by a comment and by aliases of those types
    static const char *kRuntime = "// declared by hiprtc\n";
    static const char *kStdint =
        "#pragma once\n"
        "typedef __hip_internal::int8_t int8_t; typedef __hip_internal::uint8_t uint8_t;\n"
        "typedef __hip_internal::int16_t int16_t; typedef __hip_internal::uint16_t uint16_t;\n"
        "typedef __hip_internal::int32_t int32_t; typedef __hip_internal::uint32_t uint32_t;\n"
        "typedef __hip_internal::int64_t int64_t; typedef __hip_internal::uint64_t uint64_t;\n";
    const char *headers[] = {kDeviceModelHpp, generated.c_str(), kRuntime, kStdint};
    const char *names[] = {"device_model.hpp", "lems_synapse_functions.h", "hip/hip_runtime.h", "stdint.h"};
    if (hiprtcCreateProgram(&prog, kLemsKernelHip, "lems_kernel.hip", 4, headers, names) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return -1;
    }
    const std::string off = "--offload-arch=" + arch;
    // the same floating-point settings as the library's own kernels: no contraction, IEEE
    // division and square root
    const char *opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, 5, opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        err = "hiprtc could not compile the generated LEMS kernels: " + log;
        hiprtcDestroyProgram(&prog);
        return -1;
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    std::vector<char> code(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    Compiled c;
    if (hipModuleLoadData(&c.module, code.data()) != hipSuccess) {
        err = "hipModuleLoadData of the synapse kernel failed";
        return -1;
    }
    if (hipModuleGetFunction(&c.fn, c.module, "lems_syn_kernel") != hipSuccess ||
        hipModuleGetFunction(&c.cell_fn, c.module, "lems_cell_kernel") != hipSuccess ||
        hipModuleGetFunction(&c.ext_fn, c.module, "ext_channels_kernel_rtc") != hipSuccess) {
        err = "lems_syn_kernel / lems_cell_kernel / ext_channels_kernel_rtc not found in the compiled module";
        return -1;
    }
    cache[key] = c;
    *fn = c.fn;
    *cell_fn = c.cell_fn;
    *ext_fn = c.ext_fn;
    return 0;
}

}  // namespace edenhip
