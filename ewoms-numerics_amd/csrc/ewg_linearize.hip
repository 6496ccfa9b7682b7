// ewg_linearize.hip -- ECFV linearization of the immiscible two-phase model over the whole
// grid: one thread per cell, first the intensive quantities (k_iq), then the element
// context of every cell as focus DOF, by forward-mode AD (k_linearize) or finite
// differences (k_linearize_fd); the per-cell code, with the reference functions it
// follows, is in ewg_lin_device.h.
#include "ewg_lin_device.h"

namespace ewg {

namespace {

using namespace lin;

// With a device Newton state (a.ns) the kernels do nothing unless this iteration
// linearizes, and take dt and the Newton iteration (storage-cache semantics) from it.
__global__ void __launch_bounds__(256) k_iq(LinArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N || (a.ns && !a.ns->active)) return;
    iq_cell(a, i);
}

template <int DIM, class S>
__global__ void __launch_bounds__(256) k_linearize(LinArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    double dt = a.dt;
    int newton_iter = a.newton_iter;
    if (a.ns) {
        if (!a.ns->active) return;
        dt = a.ns->dt;
        newton_iter = a.ns->it;
    }
    linearize_cell<DIM, S>(a, i, dt, newton_iter);
}

template <int DIM>
__global__ void __launch_bounds__(256) k_linearize_fd(LinArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    double dt = a.dt;
    int newton_iter = a.newton_iter;
    if (a.ns) {
        if (!a.ns->active) return;
        dt = a.ns->dt;
        newton_iter = a.ns->it;
    }
    linearize_cell_fd<DIM>(a, i, dt, newton_iter, *a.pref, a.fd_method);
}

__global__ void __launch_bounds__(256) k_bc_aux(LinArgs a, double *aux) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    bc_aux_cell(a, i, aux);
}

__global__ void __launch_bounds__(256) k_storage(LinArgs a, const double *x, double *M) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    double2 xi = reinterpret_cast<const double2 *>(x)[i];
    double m[2];
    storage<double>(a, a.phi[i], xi.y, m);
    reinterpret_cast<double2 *>(M)[i] = make_double2(m[0], m[1]);
}

}  // namespace

void launch_linearize(const LinArgs &a, hipStream_t s) {
    int blocks = (a.g.N + 255) / 256;
    hipLaunchKernelGGL(k_iq, dim3(blocks), dim3(256), 0, s, a);
    if (a.fd) {
        if (a.g.dim == 3)
            hipLaunchKernelGGL(k_linearize_fd<3>, dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_linearize_fd<2>, dim3(blocks), dim3(256), 0, s, a);
        return;
    }
    if (a.flux_fp32) {
        if (a.g.dim == 3)
            hipLaunchKernelGGL((k_linearize<3, float>), dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_linearize<2, float>), dim3(blocks), dim3(256), 0, s, a);
    } else {
        if (a.g.dim == 3)
            hipLaunchKernelGGL((k_linearize<3, double>), dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_linearize<2, double>), dim3(blocks), dim3(256), 0, s, a);
    }
}

void launch_bc_aux(const LinArgs &a, double *aux, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_aux, dim3((a.g.N + 255) / 256), dim3(256), 0, s, a, aux);
}

void launch_storage(const LinArgs &a, const double *x, double *M, hipStream_t s) {
    int blocks = (a.g.N + 255) / 256;
    hipLaunchKernelGGL(k_storage, dim3(blocks), dim3(256), 0, s, a, x, M);
}

}  // namespace ewg
