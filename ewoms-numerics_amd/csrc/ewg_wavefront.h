// ewg_wavefront.h -- persistent tile-wavefront launches for the 3D ILU(0) recurrences
// (factorization and sweeps in ewg_sweep.hip, edge-buffer set-up in ewg_wavefront.hip).
#pragma once

#include <vector>

#include <hip/hip_runtime.h>

#include "ewg_internal.h"

namespace ewg {

constexpr int kWfTX = 8;   // tile extent in x (lanes)
constexpr int kWfTY = 8;   // tile extent in y
constexpr int kFactorPD = 3;    // prefetch depth of the wavefront-layout factorization
#ifndef EWG_SWEEP_PD
#define EWG_SWEEP_PD 8
#endif
constexpr int kSweepPD = EWG_SWEEP_PD;  // prefetch depth of the triangular sweeps' operands
constexpr int kSweepPE = 1;     // prefetch depth of their cross-tile words

constexpr int kWfMaxWords = 4;  // 64-bit words per exchanged value (FP64 2x2 block)
// local steps per tile in the wavefront layout: the tile's skew of 14 steps plus nz,
// rounded up to the sweeps' unroll of 16 (every step the sweeps run has its own row)
inline __host__ __device__ int wl_nsteps(int nz) { return (nz + 14 + 15) & ~15; }
// largest grid the tile-wavefront kernels take: their 32-bit buffer offsets (32 bytes per
// cell and array, plus a skew margin and the out-of-range base 2^31) must not wrap
constexpr long long kWfMaxCells = 1LL << 25;

// synchronisation state: each launch takes tile tickets from a counter that starts at
// zero (reset by a one-thread kernel in front of the launch, or by the BiCGStab vector
// update queued before a preconditioner apply); the edge buffers hold the sentinel
// between launches
struct WfSync {
    unsigned *ticket;          // two counters (the warp-specialized apply uses both)
    const int2 *order;         // [ntiles] tiles (tx, ty) in anti-diagonal order
    unsigned long long *ex;    // [ntiles][TY][nz][kWfMaxWords] last-column values
    unsigned long long *ey;    // [ntiles][TX][nz][kWfMaxWords] last-row values
    size_t edge_words;         // words in each of ex, ey
    int ntx, nty, ntiles;
    int *err;                  // set when a wait exceeded its spin limit
};

// fill both edge buffers with the sentinel (context creation, and after an error)
void wf_init_edges(const WfSync &sync, hipStream_t s);

// The ILU(0) factors in the wavefront layout of the 8x8 tiles of sync: entry
// ((tile * wl_nsteps(nz) + step) * 64 + lane) holds the 2x2 block of the cell that lane of
// that tile reaches at that local step (L in forward order, U and Dinv in backward
// order), so that each step of a tile reads one contiguous row of 64 blocks.
template <class T>
struct WlPrecond {
    T *L;  // [3][entries][4]
    T *U;  // [3][entries][4]
    T *D;  // [entries][4]
    T *w;  // [entries][2] forward-sweep result
};
size_t wl_entries(const Grid &g, const WfSync &sync);
// the factorization straight into the layout (FP64 arithmetic; FP32 rounds the stored
// factors and upper blocks); entries without a cell keep what they hold (zeros). Af:
// scratch for the operands in the forward layout, 7 x wl_entries x 4 doubles
// gate (optional): a device flag; both launches return at once while it is 0
// A32 (optional): the packing pass also stores every block of A rounded to FP32 there
// (slot-major like A), the operator of the mixed-precision BiCGStab
void wl_factor64(const Grid &g, const WfSync &sync, const double *A, double *Af,
                 const WlPrecond<double> &wl, int *err, hipStream_t s, const int *gate = nullptr,
                 float *A32 = nullptr);
void wl_factor32(const Grid &g, const WfSync &sync, const double *A, double *Af,
                 const WlPrecond<float> &wl, int *err, hipStream_t s, const int *gate = nullptr,
                 float *A32 = nullptr);
// v = (LU)^-1 d with one compute and one edge wavefront per tile. The forward sweep takes
// tickets from sync.ticket[0], the backward sweep from sync.ticket[1]; tickets_clear:
// a kernel queued on the stream after the last sweep that took tickets has set both to
// zero (otherwise a small kernel in front of the forward sweep does).
void wl_apply64(const Grid &g, const WfSync &sync, const WlPrecond<double> &wl, const double *d,
                double *v, const int *status, hipStream_t s, bool tickets_clear = false);
void wl_apply32(const Grid &g, const WfSync &sync, const WlPrecond<float> &wl, const double *d,
                double *v, const int *status, hipStream_t s, bool tickets_clear = false);

// Several realizations' preconditioner applications v_i = (L_i U_i)^-1 d_i in one pair
// of sweep launches (each realization with its own layout, tickets and edge buffers; its
// blocks take tickets of their own realization, so the schedule inside each is the
// single launch's). Realizations whose status is not running skip, and so do those whose
// precision flag (prec) differs from the batch's (want).
struct WlMember {
    Grid g;
    WfSync sync;
    const void *L, *U, *D;  // WlPrecond<T> arrays
    void *w;
    const double *d;
    double *out;
    const int *status;
    const double *A = nullptr;     // Jacobian: the batch also factorizes (wl_batch_factor)
    double *Af = nullptr;          // the realization's operand scratch (see wl_factor64)
    int *err = nullptr;            // the realization's factorization error flag
    const int *active = nullptr;   // factorize only when *active
    const int *prec = nullptr;     // sweep only when *prec == want (null: always)
    int want = 0;
};
void *wl_batch_create(const std::vector<WlMember> &members, bool fp32, hipStream_t s);
void wl_batch_apply(void *batch, hipStream_t s);
// the same over the tiles that tmem lists (tile -> realization, device array of ntiles
// entries: every tile of each listed realization, a realization's tiles together)
void wl_batch_apply_tiles(void *batch, const int *tmem, int ntiles, hipStream_t s);
void wl_batch_factor(void *batch, hipStream_t s);
void wl_batch_free(void *batch);

}  // namespace ewg
