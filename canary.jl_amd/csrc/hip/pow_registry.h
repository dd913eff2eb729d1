// Host-side registry of the pow tables of crpow.h (one KappaTable per gas
// exponent kappa = cp/cv, at most kMaxKappa per device), kept in one device
// array of the library.  A table is built on the host (the same code as the
// device's accurate phase) and copied in the first time a kappa is asked
// for on a device, synchronously and outside graph capture; the kernels
// read it through PhysParams::ktab.  Without a table (capture in progress,
// kappa outside [1, 2], registry full) the kernels take the general
// correctly rounded path: the same bits, more instructions.
#pragma once
#include <hip/hip_runtime.h>

#include "crpow.h"

namespace canary {

constexpr int kMaxKappa = 16;

// The device table of kappa on the current device, or null.  `stream` is
// the caller's stream: while it is capturing, no table is created.
const crpow::KappaTable* kappa_table(double kappa, hipStream_t stream);

// The current device's registered tables (first `*n` entries of the array).
const crpow::KappaTable* kappa_tables(int* n);

}  // namespace canary
