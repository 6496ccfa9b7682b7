// ewg_cnn_internal.h -- what the Newton driver uses of the warm-start network.
#pragma once

#include <hip/hip_runtime.h>

struct ewg_cnn;

namespace ewg {
namespace cnn {

// xguess = xold + the network's update for the state xold (cells of the network's grid).
// The network runs on the active region: the box around the cells with non-wetting phase
// and src_box (x0, x1, y0, y1, z0, z1; empty when x1 < x0: the cells next to boundaries that
// inject it), grown by the layers' reach; elsewhere xguess = xold. Synchronizes s once.
int predict_guess(ewg_cnn *net, const double *xold, const double *K, const double *phi,
                  double dt, double *xguess, hipStream_t s, const int *src_box);
// cells of the last prediction's region
int64_t last_inference_cells(const ewg_cnn *net);

// the network takes samples of exactly this grid
bool fits_grid(const ewg_cnn *net, int nx, int ny, int nz);

}  // namespace cnn
}  // namespace ewg
