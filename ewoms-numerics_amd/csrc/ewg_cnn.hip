// ewg_cnn.hip -- the warm-start network: 3x3x3 convolutions as implicit GEMMs on the
// gfx950 matrix cores (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation),
// forward, backward-data and weight-gradient kernels, Adam, and the feature / guess
// kernels used by the Newton driver.
//
// Layouts (cells of a batch stacked along z, x fastest):
//   activations  BF16 channels-last [cell][C], C = 8 (input features) or 32
//   gradients    FP32 channels-last [cell][32] between layers, BF16 for the MFMA operands
//   transposed   BF16 [C][cell] (output gradients) and [3][C][cell] (inputs shifted by
//                dx = -1, 0, +1 with zeros outside the grid) for the weight gradients
//   weights      per layer BF16 [28 taps][32 rows][K channels], rows = output channels
//                (forward) or input channels (backward data, taps mirrored); tap 27 and
//                absent rows / channels are zero
//
// Matrix-core operand layout of v_mfma_f32_32x32x16_bf16 (64 lanes, l = i + 32 g):
//   A[i][k]: lane i + 32 (k / 8), element k % 8;  B[k][j]: lane j + 32 (k / 8), element
//   k % 8;  D[i][j]: lane j + 32 ((i / 4) % 2), element 4 (i / 8) + i % 4.
// Convolution: D[out channel][cell] += W[out][(tap, in)] * act[cell + off(tap)][in]; K
// runs over (tap, input channel) 16 at a time. Weight gradient: D[out][in] +=
// gz[out][cell] * act[cell + off(tap)][in], K over cells 16 at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ewg_cnn.h"
#include "ewg_cnn_internal.h"

namespace ewg {
namespace cnn {

using bf16x8 = __attribute__((ext_vector_type(8))) __bf16;
using f32x16 = __attribute__((ext_vector_type(16))) float;
using u32x4 = __attribute__((ext_vector_type(4))) unsigned;
using u32x2 = __attribute__((ext_vector_type(2))) unsigned;

constexpr unsigned kFar = 0x80000000u;  // offset that reads as zero (out of range)
constexpr int kTapsPad = 28;
constexpr int kWgParts = 128;           // cell partitions of the weight gradient
constexpr int kWgRing = 4;              // weight-gradient chunks in flight per wavefront
constexpr int kBiasParts = 32;          // cell partitions of the bias gradient
constexpr int kConvPF = 4;              // convolution operand prefetch distance (K steps)
// inference on the active region: cells with S_n above kActiveEps, grown by kActiveMargin
// cells (the five 3x3x3 layers see 5 cells: the active cells and 3 more layers around them
// get the whole-grid prediction bit for bit)
constexpr double kActiveEps = 1e-8;
constexpr int kActiveMargin = 8;
// the crop of a prediction on the device: origin, extent, cells
enum { kGeoX0 = 0, kGeoY0, kGeoZ0, kGeoCx, kGeoCy, kGeoCz, kGeoN, kGeoLen };

struct Layer {
    int ci, co, cip, cipb;  // real in/out channels, K channels forward / backward data
    int woff, boff;         // offsets in the flat parameter array
};
constexpr Layer kLayers[5] = {
    {5, 32, 8, 32, 0, 4320},
    {32, 32, 32, 32, 4352, 32000},
    {32, 32, 32, 32, 32032, 59680},
    {32, 32, 32, 32, 59712, 87360},
    {32, 2, 32, 8, 87392, 89120},
};
static_assert(89120 + 2 == EWG_CNN_NUM_PARAMS, "parameter layout");

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, size_t bytes) {
    unsigned n = bytes > 0x7FFFFFF0u ? 0x7FFFFFF0u : static_cast<unsigned>(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(n),
                                             0x00020000);
}
__device__ __forceinline__ bf16x8 ld8(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// float -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) {
    return __uint_as_float(static_cast<unsigned>(b) << 16);
}

enum Mode { FWD_FIRST = 0, FWD_RES = 1, FWD_OUT = 2, BWD_DATA = 3 };

struct ConvArgs {
    int nx, ny, nz;        // one sample
    int ncells;            // cells of the batch
    const int *geo;        // null, or a crop on the device (CropGeo): nx, ny, nz and ncells
                           // from there (the launch covers the largest crop)
    const unsigned short *in;    // [ncells][CIP]
    const unsigned short *w;     // [28][32][CIP]
    const float *bias;           // [32] or nullptr
    const unsigned short *res;   // FWD_RES: previous hidden state [ncells][32]
    const float *gadd;           // BWD_DATA: added to the result [ncells][32] or nullptr
    unsigned short *out_h;       // FWD_FIRST / FWD_RES: [ncells][32]
    unsigned short *out_u;       // FWD_RES: relu branch [ncells][32] or nullptr
    float *out_f;                // FWD_OUT: [ncells][2]; BWD_DATA: [ncells][32]
};

// tap t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1); tap 27 is padding (never valid)
__host__ __device__ constexpr int tap_dx(int t) { return t % 3 - 1; }
__host__ __device__ constexpr int tap_dy(int t) { return (t / 3) % 3 - 1; }
__host__ __device__ constexpr int tap_dz(int t) { return t / 9 - 1; }
// neighbour bits a cell needs for tap t (bit 6: the cell exists; padding: unreachable)
__host__ __device__ constexpr unsigned tap_mask(int t) {
    return t >= 27 ? 0x80u
                   : (tap_dx(t) < 0 ? 1u : tap_dx(t) > 0 ? 2u : 0u) |
                         (tap_dy(t) < 0 ? 4u : tap_dy(t) > 0 ? 8u : 0u) |
                         (tap_dz(t) < 0 ? 16u : tap_dz(t) > 0 ? 32u : 0u) | 64u;
}

// One wavefront computes 64 cells (two 32-cell column blocks) x 32 output rows; the
// workgroup's four wavefronts share the layer's weights in LDS and stride over the grid.
template <int CIP, int MODE>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    constexpr int NSTEP = CIP == 32 ? 54 : 14;  // K = 27 x 32 or 28 x 8, 16 per step
    constexpr int WELEM = kTapsPad * 32 * CIP;
    __shared__ __attribute__((aligned(16))) unsigned short wl[WELEM];
    if (a.geo) {
        a.nx = a.geo[kGeoCx];
        a.ny = a.geo[kGeoCy];
        a.nz = a.geo[kGeoCz];
        a.ncells = a.geo[kGeoN];
    }
    // workgroups past the cells (a launch sized for the largest crop) do nothing
    if (blockIdx.x * 4 * 64 >= a.ncells) return;
    {
        const u32x4 *src = reinterpret_cast<const u32x4 *>(a.w);
        u32x4 *dst = reinterpret_cast<u32x4 *>(wl);
        for (int i = threadIdx.x; i < WELEM / 8; i += 256) dst[i] = src[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 31, g = lane >> 5;
    const int nx = a.nx, ny = a.ny, nz = a.nz, nxny = nx * ny;
    const int ntile = (a.ncells + 63) / 64;
    const __amdgpu_buffer_rsrc_t rin = rsrc(a.in, (size_t)a.ncells * CIP * 2);
    float bias[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int co = 8 * (it >> 2) + 4 * g + (it & 3);
        bias[it] = a.bias ? a.bias[co] : 0.0f;
    }
    // byte offset of each tap's neighbour relative to the cell (uniform over the wave)
    const int sx = CIP * 2, sy = nx * CIP * 2, sz = nxny * CIP * 2;
    // this lane's channel half (CIP = 32) or tap parity (CIP = 8), in bytes / elements
    const unsigned lane_b = CIP == 32 ? 16u * g : 0u;
    const int lane_w = CIP == 32 ? j * CIP + 8 * g : j * CIP + 256 * g;
    for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
        int cell[2];
        unsigned nbm[2];   // bit d: neighbour in direction d (-x,+x,-y,+y,-z,+z) exists
        unsigned base[2];  // byte offset of the lane's first channel at the cell
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            cell[t] = tile * 64 + t * 32 + j;
            const int x = cell[t] % nx, r = cell[t] / nx, y = r % ny, z = (r / ny) % nz;
            const bool in = cell[t] < a.ncells;
            nbm[t] = in ? ((x > 0) | (x + 1 < nx) << 1 | (y > 0) << 2 | (y + 1 < ny) << 3 |
                           (z > 0) << 4 | (z + 1 < nz) << 5 | 1u << 6)
                        : 0u;
            base[t] = (unsigned)(cell[t] * CIP) * 2u + lane_b;
        }
        // operand offset of step s (compile-time after unrolling): K = 16 s + 8 g runs over
        // (tap, channel); CIP = 32: tap s / 2 for both halves; CIP = 8: tap 2 s + g
        auto bofs = [&](int s, int t) -> unsigned {
            const int ta = CIP == 32 ? s / 2 : 2 * s, tb = CIP == 32 ? s / 2 : 2 * s + 1;
            const unsigned ma = tap_mask(ta), mb = tap_mask(tb);
            const int oa = tap_dx(ta) * sx + tap_dy(ta) * sy + tap_dz(ta) * sz;
            const int ob = tap_dx(tb) * sx + tap_dy(tb) * sy + tap_dz(tb) * sz;
            const unsigned m = CIP == 32 || g == 0 ? ma : mb;
            const int o = CIP == 32 || g == 0 ? oa : ob;
            const unsigned ch = CIP == 32 ? (unsigned)(s & 1) * 32u : 0u;
            return (nbm[t] & m) == m ? base[t] + (unsigned)o + ch : kFar;
        };
        f32x16 acc0 = {}, acc1 = {};
        // the operands of the next kConvPF steps are in flight while a step's products run
        bf16x8 r0[kConvPF], r1[kConvPF];
#pragma unroll
        for (int k = 0; k < kConvPF; ++k) {
            r0[k] = ld8(rin, bofs(k < NSTEP ? k : NSTEP - 1, 0));
            r1[k] = ld8(rin, bofs(k < NSTEP ? k : NSTEP - 1, 1));
        }
#pragma unroll
        for (int s = 0; s < NSTEP; ++s) {
            const bf16x8 b0 = r0[s % kConvPF], b1 = r1[s % kConvPF];
            const int sn = s + kConvPF;
            if (sn < NSTEP) {
                r0[s % kConvPF] = ld8(rin, bofs(sn, 0));
                r1[s % kConvPF] = ld8(rin, bofs(sn, 1));
            }
            const int ws = CIP == 32 ? (s / 2) * 32 * CIP + (s & 1) * 16 : s * 512;
            const bf16x8 wa = *reinterpret_cast<const bf16x8 *>(&wl[lane_w + ws]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, b1, acc1, 0, 0, 0);
            // keep the prefetch distance as written (unbounded hoisting costs occupancy)
            __builtin_amdgcn_sched_barrier(0);
        }
        // epilogue: lane holds cell j of each block, rows 8q + 4g + r in element 4q + r
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (cell[t] >= a.ncells) continue;
            const f32x16 &acc = t == 0 ? acc0 : acc1;
            const size_t c = (size_t)cell[t];
            if (MODE == FWD_OUT) {
                if (g == 0)
                    *reinterpret_cast<float2 *>(a.out_f + c * 2) =
                        make_float2(acc[0] + bias[0], acc[1] + bias[1]);
                continue;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const size_t e = c * 32 + 8 * q + 4 * g;
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = acc[4 * q + r];
                if (MODE == BWD_DATA) {
                    if (a.gadd) {
                        const float4 ad = *reinterpret_cast<const float4 *>(a.gadd + e);
                        v[0] += ad.x;
                        v[1] += ad.y;
                        v[2] += ad.z;
                        v[3] += ad.w;
                    }
                    *reinterpret_cast<float4 *>(a.out_f + e) = make_float4(v[0], v[1], v[2], v[3]);
                    continue;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = fmaxf(v[r] + bias[4 * q + r], 0.0f);
                unsigned short hb[4], ub[4];
                if (MODE == FWD_RES) {
                    const uint2 pr = *reinterpret_cast<const uint2 *>(a.res + e);
                    const unsigned short p[4] = {
                        static_cast<unsigned short>(pr.x & 0xffff), static_cast<unsigned short>(pr.x >> 16),
                        static_cast<unsigned short>(pr.y & 0xffff), static_cast<unsigned short>(pr.y >> 16)};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        ub[r] = f2bf(v[r]);
                        hb[r] = f2bf(bf2f(p[r]) + v[r]);
                    }
                    if (a.out_u)
                        *reinterpret_cast<uint2 *>(a.out_u + e) =
                            make_uint2(ub[0] | (unsigned)ub[1] << 16, ub[2] | (unsigned)ub[3] << 16);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) hb[r] = f2bf(v[r]);
                }
                *reinterpret_cast<uint2 *>(a.out_h + e) =
                    make_uint2(hb[0] | (unsigned)hb[1] << 16, hb[2] | (unsigned)hb[3] << 16);
            }
        }
    }
}

// Weight gradient of one (dy, dz) row of taps (dx = -1, 0, +1) over a partition of the
// cells: part[p][tap][out][in] = sum over the partition's 16-cell chunks of
// gzT[out][cell] * actT[dx + 1][in][cell + dy nx + dz nx ny]. One wavefront per block.
struct WgArgs {
    int nx, ny, nz, ncells;
    int cout;                      // rows of gzT that are real
    int cin;                       // input channels (rows of each actT shift)
    const unsigned short *gzT;     // [32][ncells]
    const unsigned short *actT;    // [3][cin][ncells]
    float *part;                   // [kWgParts][27][32][32]
};

__global__ void __launch_bounds__(64) k_wgrad(WgArgs a) {
    const int lane = threadIdx.x, i = lane & 31, g = lane >> 5;
    const int p = blockIdx.x, row = blockIdx.y;  // row = (dz + 1) * 3 + (dy + 1)
    const int dy = row % 3 - 1, dz = row / 3 - 1;
    const int nx = a.nx, ny = a.ny, nz = a.nz, nxny = nx * ny;
    const size_t nc = (size_t)a.ncells;
    const __amdgpu_buffer_rsrc_t rg = rsrc(a.gzT, 32 * nc * 2);
    const __amdgpu_buffer_rsrc_t ra = rsrc(a.actT, 3 * (size_t)a.cin * nc * 2);
    const int nchunk = a.ncells / 16;
    const int roff = dy * nx + dz * nxny;
    f32x16 acc[3] = {{}, {}, {}};
    // operands of chunk c (zeros past the last chunk or outside the grid)
    auto load = [&](int c, bf16x8 &A, bf16x8 *B) {
        const bool in = c < nchunk;
        const int v0 = 16 * c;  // the chunk lies in one x-row (nx is a multiple of 16)
        const int r = v0 / nx, y = r % ny, z = (r / ny) % nz;
        const bool rowok = in && (unsigned)(y + dy) < (unsigned)ny && (unsigned)(z + dz) < (unsigned)nz;
        const unsigned ga = in && i < a.cout ? (unsigned)((size_t)i * nc + v0 + 8 * g) * 2u : kFar;
        A = ld8(rg, ga);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const unsigned bo = rowok && i < a.cin
                                    ? (unsigned)(((size_t)s * a.cin + i) * nc + v0 + 8 * g + roff) * 2u
                                    : kFar;
            B[s] = ld8(ra, bo);
        }
    };
    // the block's chunks p, p + G, p + 2G, ... in that order, kWgRing of them in flight
    const int G = gridDim.x;
    bf16x8 Ar[kWgRing], Br[kWgRing][3];
#pragma unroll
    for (int k = 0; k < kWgRing; ++k) load(p + k * G, Ar[k], Br[k]);
    for (int c = p; c < nchunk; c += kWgRing * G) {
#pragma unroll
        for (int k = 0; k < kWgRing; ++k) {
            if (c + k * G < nchunk) {
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    acc[s] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ar[k], Br[k][s], acc[s], 0, 0, 0);
            }
            load(c + (k + kWgRing) * G, Ar[k], Br[k]);
        }
    }
    // D[out][in]: lane holds in = i, out = 8q + 4g + r in element 4q + r
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int tap = (dz + 1) * 9 + (dy + 1) * 3 + s;
        float *dst = a.part + ((size_t)p * 27 + tap) * 1024;
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int o = 8 * (it >> 2) + 4 * g + (it & 3);
            dst[o * 32 + i] = acc[s][it];
        }
    }
}

// grad[layer W][out][in][tap] = sum over partitions (fixed order); one thread per entry of
// a partition's [tap][out][in] block, so that the loads of a wave are contiguous
__global__ void k_wreduce(const float *part, int nparts, int co, int ci, float *gw) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 27 * 1024) return;
    const int inn = q & 31, out = (q >> 5) & 31, tap = q >> 10;
    if (inn >= ci || out >= co) return;
    float s = 0.0f;
#pragma unroll 8
    for (int p = 0; p < nparts; ++p) s += part[(size_t)p * 27 * 1024 + q];
    gw[((size_t)out * ci + inn) * 27 + tap] = s;
}

// db[out] = sum over cells of gzT[out][cell]: block (out, b) sums partition b of the
// cells into part[out][b], k_bias_final adds the partitions in order
__global__ void __launch_bounds__(256) k_bias_grad(const unsigned short *gzT, int ncells, float *part) {
    const int o = blockIdx.x, b = blockIdx.y;
    const int len = (ncells + kBiasParts - 1) / kBiasParts;
    const int c0 = b * len, c1 = min(ncells, c0 + len);
    const unsigned short *row = gzT + (size_t)o * ncells;
    float s = 0.0f;
    for (int c = c0 + threadIdx.x; c < c1; c += 256) s += bf2f(row[c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[o * kBiasParts + b] = red[0];
}

__global__ void k_bias_final(const float *part, int co, float *db) {
    const int o = threadIdx.x;
    if (o >= co) return;
    float s = 0.0f;
    for (int b = 0; b < kBiasParts; ++b) s += part[o * kBiasParts + b];
    db[o] = s;
}

// loss 1/(2 n) sum_c (y - t)^2 per block; gradient g = (y - t) / n as BF16 [cell][8]
// (channels 2..7 zero) and transposed rows [2][cell]
__global__ void __launch_bounds__(256) k_loss(const float *y, const float *t, int ncells,
                                               unsigned short *gy, unsigned short *gyT,
                                               double *part) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (c < ncells) {
        const float2 yv = reinterpret_cast<const float2 *>(y)[c];
        const float2 tv = reinterpret_cast<const float2 *>(t)[c];
        const float d0 = yv.x - tv.x, d1 = yv.y - tv.y;
        l = 0.5 * ((double)d0 * d0 + (double)d1 * d1);
        const float inv = 1.0f / (float)ncells;
        const unsigned short g0 = f2bf(d0 * inv), g1 = f2bf(d1 * inv);
        *reinterpret_cast<u32x4 *>(gy + (size_t)c * 8) = u32x4{g0 | (unsigned)g1 << 16, 0u, 0u, 0u};
        gyT[c] = g0;
        gyT[(size_t)ncells + c] = g1;
    }
    __shared__ double red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// gz = g * (m > 0) as BF16 [cell][32] and transposed [32][cell]; 64 cells x 4 channel
// groups per block
__global__ void __launch_bounds__(256) k_mask(const float *gin, const unsigned short *m,
                                               int ncells, unsigned short *gz,
                                               unsigned short *gzT) {
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), cg = threadIdx.x >> 6;
    if (c >= ncells) return;
    const size_t e = (size_t)c * 32 + 8 * cg;
    const float4 g0 = *reinterpret_cast<const float4 *>(gin + e);
    const float4 g1 = *reinterpret_cast<const float4 *>(gin + e + 4);
    const u32x4 mv = *reinterpret_cast<const u32x4 *>(m + e);
    const float gv[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    unsigned short o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const unsigned short mb = static_cast<unsigned short>((mv[k >> 1] >> (16 * (k & 1))) & 0xffff);
        // positive BF16: sign bit clear and not zero
        const bool pos = mb != 0 && (mb & 0x8000) == 0;
        o[k] = pos ? f2bf(gv[k]) : static_cast<unsigned short>(0);
        gzT[(size_t)(8 * cg + k) * ncells + c] = o[k];
    }
    *reinterpret_cast<u32x4 *>(gz + e) =
        u32x4{o[0] | (unsigned)o[1] << 16, o[2] | (unsigned)o[3] << 16,
              o[4] | (unsigned)o[5] << 16, o[6] | (unsigned)o[7] << 16};
}

// actT[s][ch][cell] = act[cell + s - 1][ch] where x + s - 1 lies in the grid, else 0
template <int C>
__global__ void __launch_bounds__(256) k_shiftT(const unsigned short *act, int ncells, int nx,
                                                 int cin, unsigned short *actT) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const int x = c % nx;
    for (int s = 0; s < 3; ++s) {
        const int xs = x + s - 1;
        const bool ok = xs >= 0 && xs < nx;
        const unsigned short *src = act + (size_t)(c + s - 1) * C;
        for (int ch0 = 0; ch0 < cin; ch0 += 8) {
            u32x4 v = ok ? *reinterpret_cast<const u32x4 *>(src + ch0) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (ch0 + k < cin)
                    actT[((size_t)s * cin + ch0 + k) * ncells + c] =
                        static_cast<unsigned short>((v[k >> 1] >> (16 * (k & 1))) & 0xffff);
        }
    }
}

// FP32 master parameters -> BF16 operand layouts of one layer
__global__ void k_pack(const float *prm, Layer L, unsigned short *wf, unsigned short *wb,
                       float *bias) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = kTapsPad * 32 * L.cip, nb = kTapsPad * 32 * L.cipb;
    if (e < nf) {  // wf[tap][out][in] = W[out][in][tap]
        const int in = e % L.cip, out = (e / L.cip) % 32, tap = e / (32 * L.cip);
        const bool ok = tap < 27 && out < L.co && in < L.ci;
        wf[e] = ok ? f2bf(prm[L.woff + (out * L.ci + in) * 27 + tap]) : 0;
    }
    if (wb && e < nb) {  // wb[tap][in][out] = W[out][in][26 - tap]
        const int out = e % L.cipb, in = (e / L.cipb) % 32, tap = e / (32 * L.cipb);
        const bool ok = tap < 27 && out < L.co && in < L.ci;
        wb[e] = ok ? f2bf(prm[L.woff + (out * L.ci + in) * 27 + (26 - tap)]) : 0;
    }
    if (e < 32) bias[e] = e < L.co ? prm[L.boff + e] : 0.0f;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr,
                       float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float ge = g[e];
    const float me = b1 * m[e] + (1.0f - b1) * ge;
    const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
}

__global__ void k_features(const double *x, const double *K, const double *phi, int n,
                           double dt, ewg_cnn_norm nm, unsigned short *feat) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double p = x[2 * c], sw = x[2 * c + 1];
    const float f0 = static_cast<float>((p - nm.p_ref) / nm.p_scale);
    const float f1 = static_cast<float>(1.0 - sw);
    const float f2 = static_cast<float>((log(K[c]) - nm.lnk_ref) / nm.lnk_scale);
    const float f3 = static_cast<float>((phi[c] - nm.phi_ref) / nm.phi_scale);
    const float f4 = static_cast<float>(dt / nm.dt_scale);
    *reinterpret_cast<u32x4 *>(feat + (size_t)c * 8) =
        u32x4{f2bf(f0) | (unsigned)f2bf(f1) << 16, f2bf(f2) | (unsigned)f2bf(f3) << 16,
              f2bf(f4), 0u};
}

// Inference on the active region: the bounding box (x0, x1, y0, y1, z0, z1) of the cells
// with a non-wetting saturation above eps, merged into box by min / max atomics (one set
// per block that has such a cell)
__global__ void k_box_init(int *box) {
    if (threadIdx.x < 6) box[threadIdx.x] = (threadIdx.x & 1) ? -1 : 0x7fffffff;
}

__global__ void __launch_bounds__(256)
k_active_box(const double *x, int nx, int ny, int n, double eps, int *box) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int v[6] = {0x7fffffff, -1, 0x7fffffff, -1, 0x7fffffff, -1};
    if (c < n && 1.0 - x[2 * (size_t)c + 1] > eps) {
        const int xi = c % nx, r = c / nx, yi = r % ny, zi = r / ny;
        v[0] = v[1] = xi;
        v[2] = v[3] = yi;
        v[4] = v[5] = zi;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int o = 32; o > 0; o >>= 1) {
            const int u = __shfl_xor(v[k], o);
            v[k] = (k & 1) ? max(v[k], u) : min(v[k], u);
        }
    __shared__ int red[4][6];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k) red[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            for (int k = 0; k < 6; ++k) v[k] = (k & 1) ? max(v[k], red[w][k]) : min(v[k], red[w][k]);
        if (v[1] >= 0)
            for (int k = 0; k < 6; ++k) {
                if (k & 1) atomicMax(box + k, v[k]);
                else atomicMin(box + k, v[k]);
            }
    }
}

// The crop of the prediction on the device (k_crop_geo): origin, extent, cells
struct SrcBox {
    int b[6];   // box of the cells next to boundaries that inject the non-wetting phase
    int valid;
};
// k_crop_geo: from the active box (x0, x1, y0, y1, z0, z1) merged with the injection
// box, grown by kActiveMargin and clipped, x extent a multiple of 16 (nx is one); no
// active cell: 0 cells (the guess is the old state); more than half of the grid: the
// whole grid
__global__ void k_crop_geo(const int *box, SrcBox src, int nx, int ny, int nz, int *geo) {
    if (threadIdx.x != 0) return;
    int b[6];
    for (int k = 0; k < 6; ++k) b[k] = box[k];
    if (src.valid)
        for (int k = 0; k < 6; ++k) b[k] = (k & 1) ? max(b[k], src.b[k]) : min(b[k], src.b[k]);
    if (b[1] < b[0]) {
        for (int k = 0; k < kGeoLen; ++k) geo[k] = 0;
        return;
    }
    const int M = kActiveMargin;
    int x0 = max(0, b[0] - M);
    const int x1 = min(nx - 1, b[1] + M);
    const int y0 = max(0, b[2] - M), y1 = min(ny - 1, b[3] + M);
    const int z0 = max(0, b[4] - M), z1 = min(nz - 1, b[5] + M);
    x0 = (x0 / 16) * 16;
    const int cx = ((x1 - x0 + 1 + 15) / 16) * 16;
    if (x0 + cx > nx) x0 = nx - cx;
    const int cy = y1 - y0 + 1, cz = z1 - z0 + 1;
    const long long nc = (long long)nx * ny * nz, ncrop = (long long)cx * cy * cz;
    if (2 * ncrop > nc) {
        geo[kGeoX0] = 0; geo[kGeoY0] = 0; geo[kGeoZ0] = 0;
        geo[kGeoCx] = nx; geo[kGeoCy] = ny; geo[kGeoCz] = nz;
        geo[kGeoN] = (int)nc;
        return;
    }
    geo[kGeoX0] = x0; geo[kGeoY0] = y0; geo[kGeoZ0] = z0;
    geo[kGeoCx] = cx; geo[kGeoCy] = cy; geo[kGeoCz] = cz;
    geo[kGeoN] = (int)ncrop;
}

// features of the cells of the crop [x0, x0 + cx) x [y0, y0 + cy) x [z0, ...) of the grid
// (the crop from geo; a launch sized for the largest crop)
__global__ void k_features_crop(const double *x, const double *K, const double *phi, int nx,
                                int ny, const int *geo, double dt, ewg_cnn_norm nm,
                                unsigned short *feat) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = geo[kGeoN];
    if (q >= n) return;
    const int x0 = geo[kGeoX0], y0 = geo[kGeoY0], z0 = geo[kGeoZ0], cx = geo[kGeoCx],
              cy = geo[kGeoCy];
    const int i = q % cx, r = q / cx, j = r % cy, k = r / cy;
    const size_t c = (size_t)(x0 + i) + (size_t)nx * ((size_t)(y0 + j) + (size_t)ny * (z0 + k));
    const double p = x[2 * c], sw = x[2 * c + 1];
    const float f0 = static_cast<float>((p - nm.p_ref) / nm.p_scale);
    const float f1 = static_cast<float>(1.0 - sw);
    const float f2 = static_cast<float>((log(K[c]) - nm.lnk_ref) / nm.lnk_scale);
    const float f3 = static_cast<float>((phi[c] - nm.phi_ref) / nm.phi_scale);
    const float f4 = static_cast<float>(dt / nm.dt_scale);
    *reinterpret_cast<u32x4 *>(feat + (size_t)q * 8) =
        u32x4{f2bf(f0) | (unsigned)f2bf(f1) << 16, f2bf(f2) | (unsigned)f2bf(f3) << 16,
              f2bf(f4), 0u};
}

// guess = old + (dp, -dS_n) from the network output at the crop's cells, saturation kept
// in [0, 1]; the others keep the old state
__global__ void k_guess_crop(const double *xold, const float *y, int nx, int ny, const int *geo,
                             ewg_cnn_norm nm, double *xg) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = geo[kGeoN];
    if (q >= n) return;
    const int x0 = geo[kGeoX0], y0 = geo[kGeoY0], z0 = geo[kGeoZ0], cx = geo[kGeoCx],
              cy = geo[kGeoCy];
    const int i = q % cx, r = q / cx, j = r % cy, k = r / cy;
    const size_t c = (size_t)(x0 + i) + (size_t)nx * ((size_t)(y0 + j) + (size_t)ny * (z0 + k));
    const float2 yv = reinterpret_cast<const float2 *>(y)[q];
    double dp = (double)yv.x * nm.dp_scale, dsn = (double)yv.y * nm.ds_scale;
    dp = fabs(dp) >= nm.dp_min ? dp : 0.0;
    dsn = fabs(dsn) >= nm.ds_min ? dsn : 0.0;
    const double p = xold[2 * c] + dp;
    double s = xold[2 * c + 1] - dsn;
    s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
    xg[2 * c] = p;
    xg[2 * c + 1] = s;
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace cnn
}  // namespace ewg

using namespace ewg::cnn;

struct ewg_cnn {
    int nx = 0, ny = 0, nz = 0, max_batch = 0;
    size_t ncell = 0;  // cells of one sample
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float *prm = nullptr, *grad = nullptr, *adm = nullptr, *adv = nullptr;
    int adam_t = 0;
    unsigned short *wf[5] = {}, *wb[5] = {};
    float *bias[5] = {};
    unsigned short *h[4] = {}, *u[4] = {};
    float *y = nullptr, *g[2] = {};
    unsigned short *gz = nullptr, *gzT = nullptr, *actT = nullptr, *feat = nullptr;
    unsigned short *bfeat = nullptr;  // gathered batch (training)
    int *box_d = nullptr;                    // active-region box (device)
    int *geo_d = nullptr, *geo_h = nullptr;  // the last prediction's crop (device, pinned host)
    hipEvent_t geo_ev = nullptr;             // recorded after the copy of geo_h
    bool geo_pending = false;
    float *bpart = nullptr;                   // bias-gradient partitions [32][kBiasParts]
    float *btarget = nullptr;
    float *part = nullptr;
    double *lpart = nullptr;
    std::vector<double> lpart_host;
    ewg_cnn_norm norm{0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0};
    double last_fwd_ms = 0.0;
    std::string err;
};

namespace {

int set_err(ewg_cnn *n, int code, const std::string &msg) {
    if (n) n->err = msg;
    return code;
}

#define CNN_TRY(net, expr)                                                                \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return set_err(net, EWG_DEVICE_ERROR,                                         \
                           std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

void free_net(ewg_cnn *n) {
    std::vector<void *> ptrs = {n->prm, n->grad, n->adm, n->adv, n->y, n->g[0], n->g[1],
                                n->gz, n->gzT, n->actT, n->feat, n->part, n->lpart,
                                n->bfeat, n->btarget};
    for (int l = 0; l < 5; ++l) {
        ptrs.push_back(n->wf[l]);
        ptrs.push_back(n->wb[l]);
        ptrs.push_back(n->bias[l]);
    }
    for (int l = 0; l < 4; ++l) {
        ptrs.push_back(n->h[l]);
        ptrs.push_back(n->u[l]);
    }
    ptrs.push_back(n->box_d);
    ptrs.push_back(n->geo_d);
    ptrs.push_back(n->bpart);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (n->geo_h) (void)hipHostFree(n->geo_h);
    if (n->geo_ev) (void)hipEventDestroy(n->geo_ev);
    if (n->ev0) (void)hipEventDestroy(n->ev0);
    if (n->ev1) (void)hipEventDestroy(n->ev1);
    if (n->stream) (void)hipStreamDestroy(n->stream);
}

int pack(ewg_cnn *n) {
    for (int l = 0; l < 5; ++l) {
        const Layer &L = kLayers[l];
        const int ne = kTapsPad * 32 * std::max(L.cip, L.cipb);
        hipLaunchKernelGGL(k_pack, dim3((ne + 255) / 256), dim3(256), 0, n->stream, n->prm, L,
                           n->wf[l], l > 0 ? n->wb[l] : nullptr, n->bias[l]);
    }
    CNN_TRY(n, hipGetLastError());
    return EWG_OK;
}

int conv_grid(int ncells) {
    const int tiles4 = (ncells + 255) / 256;
    return std::max(1, std::min(tiles4, 1024));
}

template <int CIP, int MODE>
void launch_conv(const ConvArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_conv<CIP, MODE>), dim3(conv_grid(a.ncells)), dim3(256), 0, s, a);
}

ConvArgs conv_args(const ewg_cnn *n, int ncells) {
    ConvArgs a{};
    a.nx = n->nx;
    a.ny = n->ny;
    a.nz = n->nz;
    a.ncells = ncells;
    return a;
}

// forward pass over `ncells` cells of feat; keeps h[0..3] (and u[1..3] when training)
void forward(ewg_cnn *n, const unsigned short *feat, int ncells, bool keep_u, hipStream_t s,
             bool timed, int gx = 0, int gy = 0, int gz = 0, const int *geo = nullptr) {
    ConvArgs a = conv_args(n, ncells);
    if (gx > 0) {  // samples of another grid (a crop of the network's grid)
        a.nx = gx;
        a.ny = gy;
        a.nz = gz;
    }
    a.geo = geo;  // a crop on the device (ncells: the largest crop)
    if (timed) (void)hipEventRecord(n->ev0, s);
    a.in = feat;
    a.w = n->wf[0];
    a.bias = n->bias[0];
    a.out_h = n->h[0];
    launch_conv<8, FWD_FIRST>(a, s);
    for (int l = 1; l <= 3; ++l) {
        a.in = n->h[l - 1];
        a.res = n->h[l - 1];
        a.w = n->wf[l];
        a.bias = n->bias[l];
        a.out_h = n->h[l];
        a.out_u = keep_u ? n->u[l] : nullptr;
        launch_conv<32, FWD_RES>(a, s);
    }
    a.in = n->h[3];
    a.res = nullptr;
    a.out_h = nullptr;
    a.out_u = nullptr;
    a.w = n->wf[4];
    a.bias = n->bias[4];
    a.out_f = n->y;
    launch_conv<32, FWD_OUT>(a, s);
    if (timed) (void)hipEventRecord(n->ev1, s);
}

// weight and bias gradient of layer l from gzT (rows `cout`) and the layer's input
void layer_grad(ewg_cnn *n, int l, const unsigned short *act, int ncells, hipStream_t s) {
    const Layer &L = kLayers[l];
    const int cin = L.ci;
    if (L.cip == 8)
        hipLaunchKernelGGL((k_shiftT<8>), dim3((ncells + 255) / 256), dim3(256), 0, s, act,
                           ncells, n->nx, cin, n->actT);
    else
        hipLaunchKernelGGL((k_shiftT<32>), dim3((ncells + 255) / 256), dim3(256), 0, s, act,
                           ncells, n->nx, cin, n->actT);
    WgArgs w{};
    w.nx = n->nx;
    w.ny = n->ny;
    w.nz = n->nz;
    w.ncells = ncells;
    w.cout = L.co;
    w.cin = cin;
    w.gzT = n->gzT;
    w.actT = n->actT;
    w.part = n->part;
    const int nparts = std::max(1, std::min(kWgParts, ncells / 16));
    hipLaunchKernelGGL(k_wgrad, dim3(nparts, 9), dim3(64), 0, s, w);
    hipLaunchKernelGGL(k_wreduce, dim3(27 * 1024 / 256), dim3(256), 0, s, n->part, nparts,
                       L.co, L.ci, n->grad + L.woff);
    hipLaunchKernelGGL(k_bias_grad, dim3(L.co, kBiasParts), dim3(256), 0, s, n->gzT, ncells,
                       n->bpart);
    hipLaunchKernelGGL(k_bias_final, dim3(1), dim3(64), 0, s, n->bpart, L.co, n->grad + L.boff);
}

// forward + backward: loss into *loss, gradient into n->grad
int backprop(ewg_cnn *n, const unsigned short *feat, const float *target, int ncells,
             double *loss) {
    hipStream_t s = n->stream;
    forward(n, feat, ncells, true, s, false);
    const int lblocks = (ncells + 255) / 256;
    if ((size_t)lblocks > n->lpart_host.size())
        return set_err(n, EWG_INVALID_ARGUMENT, "batch larger than the network was created for");
    // output layer: gy (BF16 [cell][8] in gz), rows 0..1 of gzT
    hipLaunchKernelGGL(k_loss, dim3(lblocks), dim3(256), 0, s, n->y, target, ncells, n->gz,
                       n->gzT, n->lpart);
    layer_grad(n, 4, n->h[3], ncells, s);
    ConvArgs a = conv_args(n, ncells);
    a.in = n->gz;
    a.w = n->wb[4];
    a.gadd = nullptr;
    a.out_f = n->g[0];
    launch_conv<8, BWD_DATA>(a, s);
    // residual layers: gh_{l-1} = gh_l + conv_l^T(gh_l * relu'(.))
    int cur = 0;
    for (int l = 3; l >= 1; --l) {
        hipLaunchKernelGGL(k_mask, dim3((ncells + 63) / 64), dim3(256), 0, s, n->g[cur],
                           n->u[l], ncells, n->gz, n->gzT);
        layer_grad(n, l, n->h[l - 1], ncells, s);
        ConvArgs b = conv_args(n, ncells);
        b.in = n->gz;
        b.w = n->wb[l];
        b.gadd = n->g[cur];
        b.out_f = n->g[cur ^ 1];
        launch_conv<32, BWD_DATA>(b, s);
        cur ^= 1;
    }
    // first layer: relu'(.) from h0 itself
    hipLaunchKernelGGL(k_mask, dim3((ncells + 63) / 64), dim3(256), 0, s, n->g[cur], n->h[0],
                       ncells, n->gz, n->gzT);
    layer_grad(n, 0, feat, ncells, s);
    CNN_TRY(n, hipGetLastError());
    CNN_TRY(n, hipMemcpyAsync(n->lpart_host.data(), n->lpart, lblocks * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    CNN_TRY(n, hipStreamSynchronize(s));
    double sum = 0.0;
    for (int b = 0; b < lblocks; ++b) sum += n->lpart_host[b];
    if (loss) *loss = sum / ncells;
    return EWG_OK;
}

int check_batch(ewg_cnn *n, int batch) {
    if (batch < 1 || batch > n->max_batch)
        return set_err(n, EWG_INVALID_ARGUMENT, "batch must be in 1..max_batch");
    return EWG_OK;
}

// the batch's samples in one contiguous pair of buffers: the caller's arrays when idx is
// null, else copies of samples idx[0..batch-1]
int gather(ewg_cnn *n, const unsigned short *feat, const float *target, int64_t nsamples,
           const int32_t *idx, int batch, const unsigned short **bf, const float **bt) {
    if (!idx) {
        if (batch > nsamples) return set_err(n, EWG_INVALID_ARGUMENT, "batch above nsamples");
        *bf = feat;
        *bt = target;
        return EWG_OK;
    }
    const size_t fb = n->ncell * 8 * sizeof(unsigned short), tb = n->ncell * 2 * sizeof(float);
    for (int b = 0; b < batch; ++b) {
        if (idx[b] < 0 || idx[b] >= nsamples)
            return set_err(n, EWG_INVALID_ARGUMENT, "sample index out of range");
        CNN_TRY(n, hipMemcpyAsync(reinterpret_cast<char *>(n->bfeat) + b * fb,
                                  reinterpret_cast<const char *>(feat) + (size_t)idx[b] * fb, fb,
                                  hipMemcpyDeviceToDevice, n->stream));
        CNN_TRY(n, hipMemcpyAsync(reinterpret_cast<char *>(n->btarget) + b * tb,
                                  reinterpret_cast<const char *>(target) + (size_t)idx[b] * tb,
                                  tb, hipMemcpyDeviceToDevice, n->stream));
    }
    *bf = n->bfeat;
    *bt = n->btarget;
    return EWG_OK;
}

}  // namespace

namespace ewg {
namespace cnn {

int predict_guess(ewg_cnn *n, const double *xold, const double *K, const double *phi,
                  double dt, double *xguess, hipStream_t s, const int *src_box) {
    const int nc = static_cast<int>(n->ncell);
    const int nx = n->nx, ny = n->ny;
    // the active region: cells with non-wetting phase, and the cells next to rate
    // boundaries that inject it; the crop is derived on the device, and the launches
    // after it are sized for the whole grid (their workgroups past the crop return at
    // once), so that the prediction is queued without a wait for the box
    SrcBox src{};
    src.valid = src_box && src_box[1] >= src_box[0];
    if (src.valid)
        for (int k = 0; k < 6; ++k) src.b[k] = src_box[k];
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, s, n->box_d);
    hipLaunchKernelGGL(k_active_box, dim3((nc + 255) / 256), dim3(256), 0, s, xold, nx, ny, nc,
                       kActiveEps, n->box_d);
    hipLaunchKernelGGL(k_crop_geo, dim3(1), dim3(64), 0, s, n->box_d, src, nx, ny, n->nz, n->geo_d);
    if (!n->geo_ev && hipEventCreateWithFlags(&n->geo_ev, hipEventDisableTiming) != hipSuccess)
        return EWG_DEVICE_ERROR;
    if (hipMemcpyAsync(n->geo_h, n->geo_d, kGeoLen * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(n->geo_ev, s) != hipSuccess)
        return EWG_DEVICE_ERROR;
    n->geo_pending = true;
    hipLaunchKernelGGL(k_features_crop, dim3((nc + 255) / 256), dim3(256), 0, s, xold, K, phi, nx,
                       ny, n->geo_d, dt, n->norm, n->feat);
    forward(n, n->feat, nc, false, s, false, 0, 0, 0, n->geo_d);
    if (hipMemcpyAsync(xguess, xold, 2 * n->ncell * sizeof(double), hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
        return EWG_DEVICE_ERROR;
    hipLaunchKernelGGL(k_guess_crop, dim3((nc + 255) / 256), dim3(256), 0, s, xold, n->y, nx, ny,
                       n->geo_d, n->norm, xguess);
    return hipGetLastError() == hipSuccess ? EWG_OK : EWG_DEVICE_ERROR;
}

// the cells of the last prediction's crop (waits for its copy to the host)
int64_t last_inference_cells(const ewg_cnn *n) {
    if (!n || !n->geo_pending) return 0;
    if (hipEventSynchronize(n->geo_ev) != hipSuccess) return 0;
    return n->geo_h[kGeoN];
}

bool fits_grid(const ewg_cnn *n, int nx, int ny, int nz) {
    return n && n->nx == nx && n->ny == ny && n->nz == nz && n->max_batch >= 1;
}

}  // namespace cnn
}  // namespace ewg

extern "C" {

int64_t ewg_cnn_last_inference_cells(const ewg_cnn *net) { return last_inference_cells(net); }

int ewg_cnn_create(int32_t nx, int32_t ny, int32_t nz, int32_t max_batch, ewg_cnn **out) {
    if (!out) return EWG_INVALID_ARGUMENT;
    *out = nullptr;
    if (nx < 16 || nx % 16 != 0 || ny < 1 || nz < 1 || max_batch < 1) return EWG_INVALID_ARGUMENT;
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t V = ncell * (size_t)max_batch;
    // 32-bit byte offsets of the 3-shift transposed copies and of the activations
    if (V * 3 * 32 * 2 >= 0x7FFFFFF0ull) return EWG_INVALID_ARGUMENT;
    auto *n = new ewg_cnn();
    n->nx = nx;
    n->ny = ny;
    n->nz = nz;
    n->max_batch = max_batch;
    n->ncell = ncell;
    auto fail = [&](hipError_t e) {
        free_net(n);
        delete n;
        (void)e;
        return EWG_DEVICE_ERROR;
    };
    hipError_t e;
#define CNN_ALLOC(p, cnt)                          \
    if ((e = dalloc(&(p), (cnt))) != hipSuccess) \
    return fail(e)
    CNN_ALLOC(n->prm, EWG_CNN_NUM_PARAMS);
    CNN_ALLOC(n->grad, EWG_CNN_NUM_PARAMS);
    CNN_ALLOC(n->adm, EWG_CNN_NUM_PARAMS);
    CNN_ALLOC(n->adv, EWG_CNN_NUM_PARAMS);
    for (int l = 0; l < 5; ++l) {
        CNN_ALLOC(n->wf[l], (size_t)kTapsPad * 32 * kLayers[l].cip);
        if (l > 0) CNN_ALLOC(n->wb[l], (size_t)kTapsPad * 32 * kLayers[l].cipb);
        CNN_ALLOC(n->bias[l], 32);
    }
    for (int l = 0; l < 4; ++l) {
        CNN_ALLOC(n->h[l], V * 32);
        if (l > 0) CNN_ALLOC(n->u[l], V * 32);
    }
    CNN_ALLOC(n->y, V * 2);
    CNN_ALLOC(n->g[0], V * 32);
    CNN_ALLOC(n->g[1], V * 32);
    CNN_ALLOC(n->gz, V * 32);
    CNN_ALLOC(n->gzT, V * 32);
    CNN_ALLOC(n->actT, V * 3 * 32);
    CNN_ALLOC(n->feat, ncell * 8);
    CNN_ALLOC(n->box_d, 6);
    CNN_ALLOC(n->geo_d, kGeoLen);
    CNN_ALLOC(n->bpart, 32 * kBiasParts);
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&n->geo_h), kGeoLen * sizeof(int))) != hipSuccess)
        return fail(e);
    for (int k = 0; k < kGeoLen; ++k) n->geo_h[k] = 0;
    CNN_ALLOC(n->part, (size_t)kWgParts * 27 * 1024);
    CNN_ALLOC(n->lpart, (V + 255) / 256);
    CNN_ALLOC(n->bfeat, V * 8);
    CNN_ALLOC(n->btarget, V * 2);
#undef CNN_ALLOC
    n->lpart_host.assign((V + 255) / 256, 0.0);
    if ((e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&n->ev0)) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&n->ev1)) != hipSuccess) return fail(e);
    std::vector<float> zero(EWG_CNN_NUM_PARAMS, 0.0f);
    if ((e = hipMemcpy(n->prm, zero.data(), zero.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(n->adm, zero.data(), zero.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(n->adv, zero.data(), zero.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e);
    // zero-filled transposed rows beyond a layer's real channels are never read; the
    // gradient rows of the output layer beyond 2 are masked by `cout`
    if (pack(n) != EWG_OK) return fail(hipErrorUnknown);
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return fail(e);
    *out = n;
    return EWG_OK;
}

void ewg_cnn_destroy(ewg_cnn *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    free_net(n);
    delete n;
}

const char *ewg_cnn_last_error(const ewg_cnn *n) { return n ? n->err.c_str() : "null network"; }

int ewg_cnn_set_params(ewg_cnn *n, const float *params) {
    if (!n || !params) return EWG_INVALID_ARGUMENT;
    CNN_TRY(n, hipMemcpyAsync(n->prm, params, EWG_CNN_NUM_PARAMS * 4, hipMemcpyHostToDevice,
                              n->stream));
    std::vector<float> zero(EWG_CNN_NUM_PARAMS, 0.0f);
    CNN_TRY(n, hipMemcpyAsync(n->adm, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, n->stream));
    CNN_TRY(n, hipMemcpyAsync(n->adv, zero.data(), zero.size() * 4, hipMemcpyHostToDevice, n->stream));
    n->adam_t = 0;
    int rc = pack(n);
    if (rc) return rc;
    CNN_TRY(n, hipStreamSynchronize(n->stream));
    return EWG_OK;
}

int ewg_cnn_get_params(const ewg_cnn *n, float *params) {
    if (!n || !params) return EWG_INVALID_ARGUMENT;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return EWG_DEVICE_ERROR;
    if (hipMemcpy(params, n->prm, EWG_CNN_NUM_PARAMS * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return EWG_DEVICE_ERROR;
    return EWG_OK;
}

int ewg_cnn_set_norm(ewg_cnn *n, const ewg_cnn_norm *norm) {
    if (!n || !norm) return EWG_INVALID_ARGUMENT;
    if (!(norm->p_scale != 0.0 && norm->lnk_scale != 0.0 && norm->phi_scale != 0.0 &&
          norm->dt_scale != 0.0))
        return set_err(n, EWG_INVALID_ARGUMENT, "feature scales must be nonzero");
    if (!(norm->dp_min >= 0.0 && norm->ds_min >= 0.0))
        return set_err(n, EWG_INVALID_ARGUMENT, "thresholds must be non-negative");
    n->norm = *norm;
    return EWG_OK;
}

int ewg_cnn_get_norm(const ewg_cnn *n, ewg_cnn_norm *norm) {
    if (!n || !norm) return EWG_INVALID_ARGUMENT;
    *norm = n->norm;
    return EWG_OK;
}

int ewg_cnn_features(ewg_cnn *n, const double *x, const double *K, const double *phi,
                     int64_t ncells, double dt, uint16_t *feat) {
    if (!n || !x || !K || !phi || !feat || ncells < 1 || ncells > 0x7FFFFFFF)
        return EWG_INVALID_ARGUMENT;
    const int nc = static_cast<int>(ncells);
    hipLaunchKernelGGL(k_features, dim3((nc + 255) / 256), dim3(256), 0, n->stream, x, K, phi,
                       nc, dt, n->norm, feat);
    CNN_TRY(n, hipGetLastError());
    CNN_TRY(n, hipStreamSynchronize(n->stream));
    return EWG_OK;
}

int ewg_cnn_forward(ewg_cnn *n, const uint16_t *feat, int32_t batch, float *y) {
    if (!n || !feat || !y) return EWG_INVALID_ARGUMENT;
    if (int rc = check_batch(n, batch)) return rc;
    const int nc = static_cast<int>(n->ncell * batch);
    forward(n, feat, nc, false, n->stream, true);
    CNN_TRY(n, hipGetLastError());
    CNN_TRY(n, hipMemcpyAsync(y, n->y, (size_t)nc * 2 * sizeof(float), hipMemcpyDeviceToDevice,
                              n->stream));
    CNN_TRY(n, hipStreamSynchronize(n->stream));
    float ms = 0.0f;
    CNN_TRY(n, hipEventElapsedTime(&ms, n->ev0, n->ev1));
    n->last_fwd_ms = ms;
    return EWG_OK;
}

int ewg_cnn_gradients(ewg_cnn *n, const uint16_t *feat, const float *target,
                      int64_t nsamples, const int32_t *idx, int32_t batch, double *loss,
                      float *grad) {
    if (!n || !feat || !target) return EWG_INVALID_ARGUMENT;
    if (int rc = check_batch(n, batch)) return rc;
    const unsigned short *bf;
    const float *bt;
    if (int rc = gather(n, feat, target, nsamples, idx, batch, &bf, &bt)) return rc;
    int rc = backprop(n, bf, bt, static_cast<int>(n->ncell * batch), loss);
    if (rc) return rc;
    if (grad)
        CNN_TRY(n, hipMemcpy(grad, n->grad, EWG_CNN_NUM_PARAMS * 4, hipMemcpyDeviceToHost));
    return EWG_OK;
}

int ewg_cnn_train_step(ewg_cnn *n, const uint16_t *feat, const float *target,
                       int64_t nsamples, const int32_t *idx, int32_t batch, float lr,
                       double *loss) {
    if (!n || !feat || !target || !(lr > 0.0f)) return EWG_INVALID_ARGUMENT;
    if (int rc = check_batch(n, batch)) return rc;
    const unsigned short *bf;
    const float *bt;
    if (int rc = gather(n, feat, target, nsamples, idx, batch, &bf, &bt)) return rc;
    int rc = backprop(n, bf, bt, static_cast<int>(n->ncell * batch), loss);
    if (rc) return rc;
    n->adam_t++;
    const float c1 = 1.0f - std::pow(0.9f, (float)n->adam_t);
    const float c2 = 1.0f - std::pow(0.999f, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((EWG_CNN_NUM_PARAMS + 255) / 256), dim3(256), 0, n->stream,
                       n->prm, n->grad, n->adm, n->adv, EWG_CNN_NUM_PARAMS, lr, c1, c2);
    if ((rc = pack(n))) return rc;
    CNN_TRY(n, hipStreamSynchronize(n->stream));
    return EWG_OK;
}

int ewg_device_alloc(int64_t bytes, void **ptr) {
    if (!ptr || bytes < 1) return EWG_INVALID_ARGUMENT;
    *ptr = nullptr;
    return hipMalloc(ptr, (size_t)bytes) == hipSuccess ? EWG_OK : EWG_DEVICE_ERROR;
}

int ewg_device_free(void *ptr) {
    if (!ptr) return EWG_OK;
    if (hipDeviceSynchronize() != hipSuccess) return EWG_DEVICE_ERROR;
    return hipFree(ptr) == hipSuccess ? EWG_OK : EWG_DEVICE_ERROR;
}

int ewg_device_upload(void *dst, const void *src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return EWG_INVALID_ARGUMENT;
    if (hipDeviceSynchronize() != hipSuccess) return EWG_DEVICE_ERROR;
    if (hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) != hipSuccess)
        return EWG_DEVICE_ERROR;
    return hipDeviceSynchronize() == hipSuccess ? EWG_OK : EWG_DEVICE_ERROR;
}

int ewg_device_download(void *dst, const void *src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return EWG_INVALID_ARGUMENT;
    if (hipDeviceSynchronize() != hipSuccess) return EWG_DEVICE_ERROR;
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess
               ? EWG_OK
               : EWG_DEVICE_ERROR;
}

int ewg_cnn_last_forward_ms(const ewg_cnn *n, double *ms) {
    if (!n || !ms) return EWG_INVALID_ARGUMENT;
    *ms = n->last_fwd_ms;
    return EWG_OK;
}

}  // extern "C"
