// ewg_cnn_internal.h -- what the Newton driver uses of the warm-start network.
#pragma once

#include <hip/hip_runtime.h>

struct ewg_cnn;

namespace ewg {
namespace cnn {

// xguess = xold + the network's update for the state xold (cells of the network's grid),
// queued on stream s
int predict_guess(ewg_cnn *net, const double *xold, const double *K, const double *phi,
                  double dt, double *xguess, hipStream_t s);

// the network takes samples of exactly this grid
bool fits_grid(const ewg_cnn *net, int nx, int ny, int nz);

}  // namespace cnn
}  // namespace ewg
