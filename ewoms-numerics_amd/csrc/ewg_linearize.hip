// ewg_linearize.hip -- ECFV linearization of the immiscible two-phase model over the whole
// grid. 3D, automatic differentiation: one fused pass (k_linearize_zt) over 16x16 columns
// of cells marching in z, the intensive quantities of each z-plane (with a one-cell halo)
// evaluated into LDS and those of the planes below and above kept in registers, then the
// element context of every cell of the plane as focus DOF. 2D, and the finite-difference
// local linearizer: one thread per cell, first the intensive quantities (k_iq), then the
// element contexts (k_linearize, k_linearize_fd). The per-cell code, with the reference
// functions it follows, is in ewg_lin_device.h.
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

// ---------------------------------------------------------------------------------------
// Fused 3D pass: a workgroup of kZT x kZT threads owns that many z-columns of cells and
// kZChunk consecutive z-planes. For each plane it evaluates the intensive quantities of
// its columns as dual numbers (kept in registers for the column's own element context)
// and their values into an LDS plane, together with the values of the one-cell halo
// around the tile (18 x 18 - 16 x 16 = 68 columns, spread over the four waves); the
// values of the plane below come from the previous step's registers, those of the plane
// above are evaluated one plane ahead. One barrier per plane (double-buffered planes).
// The arithmetic of every element context is linearize_cell's: the neighbours' values
// are the value parts of the same evaluations (intensive<double> performs the value
// operations of intensive<Ev> one for one).
// ---------------------------------------------------------------------------------------
constexpr int kZT = 16;
constexpr int kZH = kZT + 2;
constexpr int kZHalo = kZH * kZH - kZT * kZT;  // 68
#ifndef EWG_ZCHUNK
#define EWG_ZCHUNK 6
#endif
constexpr int kZChunk = EWG_ZCHUNK;  // z-planes per workgroup

struct NbVal {
    double p[2], lam[2], K;
};

__device__ __forceinline__ NbVal nb_values(const LinArgs &a, int cell) {
    const double2 xv = reinterpret_cast<const double2 *>(a.x)[cell];
    const IQ<double> q = intensive<double>(a, cell, xv.x, xv.y);
    return NbVal{{q.p[0], q.p[1]}, {q.lam[0], q.lam[1]}, a.K[cell]};
}
// the column's own intensive quantities carried from one plane to the next: the primary
// variables as plain values (their dual numbers are rebuilt with literal derivative slots
// where they are used, so that the compiler folds the products with those constants) and
// the evaluated quantities as dual numbers
struct OwnIQ {
    double pw, Sw;
    Ev pn, lw, ln;
};
__device__ __forceinline__ OwnIQ own_iq(const LinArgs &a, int cell) {
    const double2 xv = reinterpret_cast<const double2 *>(a.x)[cell];
    const IQ<Ev> q = intensive<Ev>(a, cell, Ev{xv.x, 1.0, 0.0}, Ev{xv.y, 0.0, 1.0});
    return OwnIQ{xv.x, xv.y, q.p[1], q.lam[0], q.lam[1]};
}
__device__ __forceinline__ IQ<Ev> iq_dual(const OwnIQ &o) {
    IQ<Ev> q;
    q.p[0] = Ev{o.pw, 1.0, 0.0};
    q.Sw = Ev{o.Sw, 0.0, 1.0};
    q.p[1] = o.pn;
    q.lam[0] = o.lw;
    q.lam[1] = o.ln;
    return q;
}
__device__ __forceinline__ NbVal nb_of(const OwnIQ &q, double K) {
    return NbVal{{q.pw, q.pn.v}, {q.lw.v, q.ln.v}, K};
}
__device__ __forceinline__ IQ<double> iq_of(const NbVal &v) {
    IQ<double> q;
    q.p[0] = v.p[0];
    q.p[1] = v.p[1];
    q.Sw = 0.0;  // not used by the flux
    q.lam[0] = v.lam[0];
    q.lam[1] = v.lam[1];
    return q;
}

// three workgroups' waves per SIMD (168 registers; more occupancy spills)
#ifndef EWG_ZT_OCC
#define EWG_ZT_OCC 3
#endif
template <class S, bool CONST_GEO>
__global__ void __launch_bounds__(kZT * kZT, EWG_ZT_OCC) k_linearize_zt(LinArgs a) {
    double dt = a.dt;
    int newton_iter = a.newton_iter;
    if (a.ns) {
        if (!a.ns->active) return;
        dt = a.ns->dt;
        newton_iter = a.ns->it;
    }
    const Grid &g = a.g;
    __shared__ NbVal plane[2][kZH * kZH];
    const int ntx = (g.nx + kZT - 1) / kZT, nty = (g.ny + kZT - 1) / kZT;
    int b = blockIdx.x;
    const int tx = b % ntx;
    b /= ntx;
    const int ty = b % nty;
    const int z0 = (b / nty) * kZChunk;
    const int z1 = min(z0 + kZChunk, g.nz);
    const int t = threadIdx.x, lx = t % kZT, ly = t / kZT;
    const int x = tx * kZT + lx, y = ty * kZT + ly;
    const bool valid = x < g.nx && y < g.ny;
    const int nxny = g.nxny;
    const int col = x + g.nx * y;  // the column's cell at z is col + z * nxny
    const int me = (lx + 1) + (ly + 1) * kZH;
    // the halo column this thread evaluates, if any: 17 per wave, h = 0..67 around the tile
    // (row below, row above, column left, column right)
    int hcol = -1, hslot = 0;
    {
        const int w = t >> 6, l = t & 63;
        if (l < kZHalo / 4) {
            const int h = w * (kZHalo / 4) + l;
            int hx, hy;
            if (h < kZH) {
                hx = h - 1; hy = -1;
            } else if (h < 2 * kZH) {
                hx = h - kZH - 1; hy = kZT;
            } else if (h < 2 * kZH + kZT) {
                hx = -1; hy = h - 2 * kZH;
            } else {
                hx = kZT; hy = h - 2 * kZH - kZT;
            }
            const int gx = tx * kZT + hx, gy = ty * kZT + hy;
            if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) {
                hcol = gx + g.nx * gy;
                hslot = (hx + 1) + (hy + 1) * kZH;
            }
        }
    }
    OwnIQ cur;
    double Kc = 0.0;
    NbVal prevv{}, nxtv{};
    if (valid) {
        const int c0 = col + z0 * nxny;
        Kc = a.K[c0];
        cur = own_iq(a, c0);
        plane[0][me] = nb_of(cur, Kc);
        if (z0 > 0) prevv = nb_values(a, c0 - nxny);
    }
    if (hcol >= 0) plane[0][hslot] = nb_values(a, hcol + z0 * nxny);
    __syncthreads();
    int buf = 0;
    for (int z = z0; z < z1; ++z) {
        const int i = col + z * nxny;
        const bool more = z + 1 < z1;  // plane z + 1 belongs to this workgroup
        // values of the plane above (the column's own for the +z face, the tile's and its
        // halo's for the next step); the column's dual numbers there are evaluated after
        // this plane's element contexts, so that only one set of them is live at a time
        if (valid && z + 1 < g.nz) {
            nxtv = nb_values(a, i + nxny);
            if (more) plane[buf ^ 1][me] = nxtv;
        }
        if (more && hcol >= 0) plane[buf ^ 1][hslot] = nb_values(a, hcol + (z + 1) * nxny);
        // the three phases are scheduled one after the other (interleaved, their
        // independent chains would keep far more values live than the occupancy allows)
        __builtin_amdgcn_sched_barrier(0);
        if (valid) {
            // the coordinates from the cell index, as linearize_cell has them (given as
            // {x, y, z} the compiler evaluates the faces' branches eagerly and the kernel
            // needs twice the registers)
            int ijk[3];
            cell_coords(g, i, ijk);
            const NbVal *pl = plane[buf];
            auto nbf = [&](int dir, int, double &K) {
                NbVal v;
                switch (dir) {
                    case 0: v = pl[me - 1]; break;
                    case 1: v = pl[me + 1]; break;
                    case 2: v = pl[me - kZH]; break;
                    case 3: v = pl[me + kZH]; break;
                    case 4: v = prevv; break;
                    default: v = nxtv; break;
                }
                K = v.K;
                return iq_of(v);
            };
            linearize_context<3, S, decltype(nbf), CONST_GEO>(a, i, ijk, Kc, iq_dual(cur), nbf,
                                                              dt, newton_iter);
            __builtin_amdgcn_sched_barrier(0);
            prevv = nb_of(cur, Kc);
            if (more) {
                Kc = nxtv.K;
                cur = own_iq(a, i + nxny);
            }
        }
        __syncthreads();
        buf ^= 1;
    }
}

// the per-direction face geometry of an interior cell (the same for every cell when every
// spacing is a power of two; the caller checks that)
__global__ void k_face_geo(LinArgs a, FaceGeoSet *out) {
    const int dir = threadIdx.x;
    if (dir >= 2 * a.g.dim) return;
    const int ijk[3] = {1, 1, a.g.dim == 3 ? 1 : 0};
    if (a.g.dim == 3) {
        out->d[dir] = face_geo<3, double>(a, ijk, dir);
        out->f[dir] = face_geo<3, float>(a, ijk, dir);
    } else {
        out->d[dir] = face_geo<2, double>(a, ijk, dir);
        out->f[dir] = face_geo<2, float>(a, ijk, dir);
    }
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
    if (a.g.dim == 3 && !a.fd) {
        const int blocks = ((a.g.nx + kZT - 1) / kZT) * ((a.g.ny + kZT - 1) / kZT) *
                           ((a.g.nz + kZChunk - 1) / kZChunk);
        if (a.flux_fp32) {
            if (a.fgeo)
                hipLaunchKernelGGL((k_linearize_zt<float, true>), dim3(blocks), dim3(kZT * kZT), 0, s, a);
            else
                hipLaunchKernelGGL((k_linearize_zt<float, false>), dim3(blocks), dim3(kZT * kZT), 0, s, a);
        } else {
            if (a.fgeo)
                hipLaunchKernelGGL((k_linearize_zt<double, true>), dim3(blocks), dim3(kZT * kZT), 0, s, a);
            else
                hipLaunchKernelGGL((k_linearize_zt<double, false>), dim3(blocks), dim3(kZT * kZT), 0, s, a);
        }
        return;
    }
    int blocks = (a.g.N + 255) / 256;
    hipLaunchKernelGGL(k_iq, dim3(blocks), dim3(256), 0, s, a);
    if (a.fd) {
        if (a.g.dim == 3)
            hipLaunchKernelGGL(k_linearize_fd<3>, dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(k_linearize_fd<2>, dim3(blocks), dim3(256), 0, s, a);
        return;
    }
    if (a.flux_fp32)
        hipLaunchKernelGGL((k_linearize<2, float>), dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_linearize<2, double>), dim3(blocks), dim3(256), 0, s, a);
}

void launch_face_geo(const LinArgs &a, FaceGeoSet *out, hipStream_t s) {
    hipLaunchKernelGGL(k_face_geo, dim3(1), dim3(64), 0, s, a, out);
}

void launch_bc_aux(const LinArgs &a, double *aux, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_aux, dim3((a.g.N + 255) / 256), dim3(256), 0, s, a, aux);
}

void launch_storage(const LinArgs &a, const double *x, double *M, hipStream_t s) {
    int blocks = (a.g.N + 255) / 256;
    hipLaunchKernelGGL(k_storage, dim3(blocks), dim3(256), 0, s, a, x, M);
}

}  // namespace ewg
