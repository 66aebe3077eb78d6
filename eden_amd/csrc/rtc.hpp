// rtc.hpp -- run-time compilation of the generic LEMS synapse kernel (hiprtc).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace edenhip {

// Compile lems_kernel.hip with the generated `lems_syn`, `lems_cell` and `lems_rate` source for
// the current device and load it: `fn` the synapse kernel, `cell_fn` the point-neuron kernel,
// `ext_fn` the channel pre-pass with user rate types; returns 0, or -1 with `err` set. Compiled
// code objects are kept per (device architecture, source) for the process, so contexts of one
// model compile once.
int rtc_lems_kernel(const std::string &generated, hipFunction_t *fn, hipFunction_t *cell_fn, hipFunction_t *ext_fn,
                    std::string &err);

}  // namespace edenhip
