// Device-side coupling between the spectral LES and the learned closure
// network (both in this library): physical fields are planar (n³ per
// component, x fastest), and all work is queued on the caller's stream.
#pragma once

#include "common.hpp"

struct dns_cnn;

namespace dns {

// τ (n³ × 6: xx, yy, zz, xy, yz, zx) predicted by the network from the
// physical velocity vel (n³ × 3), in physical units (scales of the dataset).
template <typename T> int cnn_stress(dns_cnn *h, const T *vel, T *tau, int n, hipStream_t st);

// Appends one (ū, τ) pair of physical fields to the network's device dataset.
template <typename T> int cnn_append(dns_cnn *h, const T *ubar, const T *tau, int n, hipStream_t st);

}  // namespace dns
