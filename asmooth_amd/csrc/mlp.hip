// Learned seed radius: a 27-64-64-1 multilayer perceptron (GELU) that maps
// ln(CIC density) over the 3x3x3 cells around a particle, on a 64^3 mesh of
// the tree cube, to ln(h / s), s the mean particle spacing of the set.  It
// has no counterpart in the reference, and the exact search still decides
// every result (a predicted radius only bounds the search, and a lane whose
// radius holds fewer than 33 particles is searched again without it).
//
// Training runs on the device with these kernels: a per-sample forward pass
// that keeps the pre-activations, a per-sample backward pass to the
// pre-activation gradients, weight gradients as sums over the batch (one
// thread per parameter, fixed summation order, no atomics), and Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/asmooth.h"
#include "common.hpp"
#include "kernels.hpp"

namespace asm_hip {
namespace {

constexpr int kIn = kMlpIn;
constexpr int kH = kMlpHidden;
// Parameter layout: W1[kIn][kH], b1[kH], W2[kH][kH], b2[kH], w3[kH], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
static_assert(kOffB3 + 1 == kMlpParams, "parameter layout");

constexpr float kGeluC = 0.7978845608028654f;   // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

// GELU in its tanh form, 0.5 z (1 + tanh(c (z + a z^3))), with
// tanh(u) = 1 - 2 / (exp(2u) + 1).
__device__ __forceinline__ float gelu_tanh_arg(float z, float& th)
{
    const float u = kGeluC * (z + kGeluA * z * z * z);
    th = 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * u) + 1.0f);
    return 0.5f * z * (1.0f + th);
}
__device__ __forceinline__ float gelu(float z)
{
    float th;
    return gelu_tanh_arg(z, th);
}
// d gelu / dz
__device__ __forceinline__ float gelu_grad(float z)
{
    float th;
    (void)gelu_tanh_arg(z, th);
    return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * kGeluC * (1.0f + 3.0f * kGeluA * z * z);
}

// Forward pass of one sample held by the calling lane; parameters in LDS.
// Optionally keeps the pre-activations z1, z2 (kH each) at the given rows.
__device__ __forceinline__ float mlp_eval(const float* __restrict__ P, const float (&x)[kIn], float* z1_out,
                                          float* z2_out)
{
    float h1[kH];
#pragma unroll
    for (int i = 0; i < kH; ++i) {
        float z = P[kOffB1 + i];
#pragma unroll
        for (int k = 0; k < kIn; ++k) z += x[k] * P[kOffW1 + k * kH + i];
        if (z1_out) z1_out[i] = z;
        h1[i] = gelu(z);
    }
    float y = P[kOffB3];
    for (int j = 0; j < kH; ++j) {
        float z = P[kOffB2 + j];
#pragma unroll
        for (int i = 0; i < kH; ++i) z += h1[i] * P[kOffW2 + i * kH + j];
        if (z2_out) z2_out[j] = z;
        y += P[kOffW3 + j] * gelu(z);
    }
    return y;
}

__device__ __forceinline__ void load_params(float* sp, const float* __restrict__ params)
{
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
}

// y[b] = f(x[idx[b]]) (idx may be null); keeps z1, z2 when non-null.
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                                          const int64_t* __restrict__ idx, int64_t nb,
                                                          float* __restrict__ y, float* __restrict__ z1,
                                                          float* __restrict__ z2)
{
    __shared__ float sp[kMlpParams];
    load_params(sp, params);
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t row = idx ? idx[b] : b;
    float xv[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) xv[k] = x[row * kIn + k];
    y[b] = mlp_eval(sp, xv, z1 ? z1 + b * kH : nullptr, z2 ? z2 + b * kH : nullptr);
}

// Per sample: dy = 2 (y - t) / nb (mean squared error), dz2 = dy w3 g'(z2),
// dz1 = (W2 dz2) g'(z1); z1, z2 are overwritten by g(z1), g(z2) (the
// activations the weight gradients need).  loss_part[b] = (y - t)^2.
__global__ __launch_bounds__(256) void mlp_backward_kernel(const float* __restrict__ params,
                                                           const float* __restrict__ t, const int64_t* __restrict__ idx,
                                                           int64_t nb, const float* __restrict__ y,
                                                           float* __restrict__ z1, float* __restrict__ z2,
                                                           float* __restrict__ dz1, float* __restrict__ dz2,
                                                           float* __restrict__ dy, float* __restrict__ loss_part)
{
    __shared__ float sp[kMlpParams];
    load_params(sp, params);
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t row = idx ? idx[b] : b;
    const float r = y[b] - t[row];
    loss_part[b] = r * r;
    const float g = 2.0f * r / static_cast<float>(nb);
    dy[b] = g;
    float d2[kH];
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        const float z = z2[b * kH + j];
        d2[j] = g * sp[kOffW3 + j] * gelu_grad(z);
        dz2[b * kH + j] = d2[j];
        z2[b * kH + j] = gelu(z);
    }
    for (int i = 0; i < kH; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < kH; ++j) s += sp[kOffW2 + i * kH + j] * d2[j];
        const float z = z1[b * kH + i];
        dz1[b * kH + i] = s * gelu_grad(z);
        z1[b * kH + i] = gelu(z);
    }
}

// Weight gradients as small GEMMs on the matrix cores.  Over the batch b:
//   dW1[k][i] = sum x[b][k] dz1[b][i],  db1[i] = sum dz1[b][i]      (rows k < 27, row 27 = ones)
//   dW2[i][j] = sum h1[b][i] dz2[b][j], db2[j] = sum dz2[b][j]      (rows i < 64, row 64 = ones)
//   dw3[j]    = sum h2[b][j] dy[b],     db3    = sum dy[b]          (rows j < 64, row 64 = ones)
// Output blocks of 32 x 32: 0-1 the first layer (one row block, two column
// blocks), 2-7 the second (three row blocks, two column blocks), 8-10 the
// output layer (three row blocks, one column block whose column 0 is dy).
// A block's batch is cut into kWgradSplit slices; one wavefront per (block,
// slice) runs v_mfma_f32_32x32x2_f32 over its slice, two samples per step,
// and stores its partial block; mlp_wgrad_reduce_kernel adds the slices in
// order (deterministic run to run).  Operands: lane l holds A[l % 32][l / 32]
// (row, sample) and B[l / 32][l % 32] (sample, column); lane l ends with
// D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32] in accumulator v.  h1, h2 are the
// activations (see mlp_backward_kernel).
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kWgradBlocks = 11;
constexpr int kWgradSplit = 32;
constexpr int kWgradPartial = kWgradBlocks * kWgradSplit * kWave * 16;   // floats

struct WgradBlock {
    int layer, mb, nbk;
};

__device__ __forceinline__ WgradBlock wgrad_block(int blk)
{
    WgradBlock w;
    w.layer = blk < 2 ? 1 : (blk < 8 ? 2 : 3);
    w.mb = w.layer == 1 ? 0 : (w.layer == 2 ? (blk - 2) / 2 : blk - 8);
    w.nbk = w.layer == 1 ? blk : (w.layer == 2 ? (blk - 2) % 2 : 0);
    return w;
}

__global__ __launch_bounds__(kWave) void mlp_wgrad_kernel(const float* __restrict__ x, const int64_t* __restrict__ idx,
                                                        int64_t nb, const float* __restrict__ h1,
                                                        const float* __restrict__ h2, const float* __restrict__ dz1,
                                                        const float* __restrict__ dz2, const float* __restrict__ dy,
                                                        float* __restrict__ partial)
{
    const int lane = threadIdx.x;
    const int blk = blockIdx.x / kWgradSplit;
    const int slice = blockIdx.x % kWgradSplit;
    const WgradBlock w = wgrad_block(blk);
    const int row = w.mb * 32 + lane % 32;   // this lane's A row
    const int col = w.nbk * 32 + lane % 32;  // this lane's B column
    const int ks = lane / 32;                // sample within the step
    const int nrows = w.layer == 1 ? kIn : kH;   // rows past nrows: the ones row, then zeros
    // Slice bounds: an even number of samples per slice.
    const int64_t per = ((nb + kWgradSplit - 1) / kWgradSplit + 1) & ~static_cast<int64_t>(1);
    const int64_t s0 = slice * per;
    const int64_t s1 = s0 + per < nb ? s0 + per : nb;
    v16f acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int64_t b0 = s0; b0 < s1; b0 += 2) {
        const int64_t b = b0 + ks;
        float av = 0.0f, bv = 0.0f;
        if (b < s1) {
            if (row < nrows) {
                av = w.layer == 1 ? x[(idx ? idx[b] : b) * kIn + row]
                                  : (w.layer == 2 ? h1[b * kH + row] : h2[b * kH + row]);
            } else if (row == nrows) {
                av = 1.0f;
            }
            if (w.layer == 1) bv = dz1[b * kH + col];
            else if (w.layer == 2) bv = dz2[b * kH + col];
            else bv = col == 0 ? dy[b] : 0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
    float* const out = partial + (static_cast<int64_t>(blk) * kWgradSplit + slice) * (kWave * 16) + lane * 16;
#pragma unroll
    for (int v = 0; v < 16; ++v) out[v] = acc[v];
}

// One thread per (block, lane, accumulator): the slices summed in order,
// stored at the parameter the element stands for.
__global__ void mlp_wgrad_reduce_kernel(const float* __restrict__ partial, float* __restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kWgradBlocks * kWave * 16) return;
    const int blk = e / (kWave * 16);
    const int lane = (e / 16) % kWave;
    const int v = e % 16;
    float sum = 0.0f;
    for (int sl = 0; sl < kWgradSplit; ++sl)
        sum += partial[(static_cast<int64_t>(blk) * kWgradSplit + sl) * (kWave * 16) + lane * 16 + v];
    const WgradBlock w = wgrad_block(blk);
    const int r = w.mb * 32 + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
    const int col = w.nbk * 32 + lane % 32;
    int p = -1;
    if (w.layer == 1) {
        if (r < kIn) p = kOffW1 + r * kH + col;
        else if (r == kIn) p = kOffB1 + col;
    } else if (w.layer == 2) {
        if (r < kH) p = kOffW2 + r * kH + col;
        else if (r == kH) p = kOffB2 + col;
    } else if (col == 0) {
        if (r < kH) p = kOffW3 + r;
        else if (r == kH) p = kOffB3;
    }
    if (p >= 0) grad[p] = sum;
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, float lr, float beta1, float beta2, float eps, float bc1,
                            float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float mp = beta1 * m[p] + (1.0f - beta1) * g;
    const float vp = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mp;
    v[p] = vp;
    params[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Features on the smoothing context's sorted particles.

// Cloud-in-cell counts on an nf^3 mesh of the cube [left, left + nf/inv).
// A workgroup takes kCicThreads x kCicRun Morton-consecutive particles.  It
// reads them coalesced into LDS (mesh coordinates minus half a cell, one
// array per axis), then each lane takes kCicRun consecutive ones: these
// mostly share the stencil's lower corner (the particles of one octant of a
// cell are consecutive in Morton order), so the lane sums their eight
// weights in registers and adds the sums only when the corner changes, the
// last sums of the lanes sharing a corner first summed across lanes.  The
// sums go into an LDS window of kCicWin^3 cells (FP64, ds_add_f64) placed at
// the lowest corner of the workgroup, whose non-zero cells are then added to
// the mesh (FP32 global atomics); stencils leaving the window go straight to
// the mesh.  Corners past the mesh are clamped onto its edge.
constexpr int kCicThreads = 256;
constexpr int kCicRun = 8;
constexpr int kCicWin = 12;
constexpr int kCicStage = kCicThreads * kCicRun + kCicThreads;   // one pad word per run: conflict-free reads

__device__ __forceinline__ int cic_slot(int j) { return j + j / kCicRun; }

__device__ __forceinline__ void cic_add(double* win, const int (&k0)[3], const int (&c)[3], const float (&acc)[8],
                                        int nf, float* __restrict__ grid)
{
    const bool inwin = c[0] >= k0[0] && c[1] >= k0[1] && c[2] >= k0[2] && c[0] + 1 - k0[0] < kCicWin &&
                       c[1] + 1 - k0[1] < kCicWin && c[2] + 1 - k0[2] < kCicWin;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = q & 1, dy = (q >> 1) & 1, dz = q >> 2;
        if (inwin) {
            atomicAdd(&win[((c[2] + dz - k0[2]) * kCicWin + (c[1] + dy - k0[1])) * kCicWin + (c[0] + dx - k0[0])],
                      static_cast<double>(acc[q]));
        } else {
            const int cx = min(max(c[0] + dx, 0), nf - 1);
            const int cy = min(max(c[1] + dy, 0), nf - 1);
            const int cz = min(max(c[2] + dz, 0), nf - 1);
            atomicAdd(&grid[(static_cast<int64_t>(cz) * nf + cy) * nf + cx], acc[q]);
        }
    }
}

__global__ __launch_bounds__(kCicThreads) void cic_kernel(const float4* __restrict__ spos, int64_t n, float left,
                                                          float inv_cell, int nf, float* __restrict__ grid)
{
    __shared__ double win[kCicWin * kCicWin * kCicWin];
    __shared__ float stage[3][kCicStage];
    __shared__ int corner[3];
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kCicThreads * kCicRun;
    const int64_t left_n = n - first;
    const int cnt = left_n < kCicThreads * kCicRun ? static_cast<int>(left_n) : kCicThreads * kCicRun;
    for (int c = threadIdx.x; c < kCicWin * kCicWin * kCicWin; c += kCicThreads) win[c] = 0.0;
    if (threadIdx.x < 3) corner[threadIdx.x] = 0x7fffffff;
    __syncthreads();
    // Indices clamped into the workgroup's particles and no branch, so that
    // the loads are in flight together; slots past the count hold copies of
    // the last particle and are never read.
    float4 pv[kCicRun];
#pragma unroll
    for (int r = 0; r < kCicRun; ++r) pv[r] = spos[first + min(r * kCicThreads + static_cast<int>(threadIdx.x), cnt - 1)];
    int m0[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
#pragma unroll
    for (int r = 0; r < kCicRun; ++r) {
        const int j = r * kCicThreads + threadIdx.x;
        const float4 p = pv[r];
        const float u[3] = {(p.x - left) * inv_cell - 0.5f, (p.y - left) * inv_cell - 0.5f,
                            (p.z - left) * inv_cell - 0.5f};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            stage[d][cic_slot(j)] = u[d];
            m0[d] = min(m0[d], static_cast<int>(floorf(u[d])));
        }
    }
    // Minimum over the wavefront by shuffles, then one LDS atomic per
    // wavefront (per-lane LDS atomics on one word serialise).
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int v = m0[d];
#pragma unroll
        for (int o = kWave / 2; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o));
        if (threadIdx.x % kWave == 0) atomicMin(&corner[d], v);
    }
    __syncthreads();
    const int k0[3] = {corner[0], corner[1], corner[2]};
    int cur[3] = {0x7fffffff, 0, 0};
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int r = 0; r < kCicRun; ++r) {
        const int j = threadIdx.x * kCicRun + r;
        if (j >= cnt) break;
        int c0[3];
        float w1[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float u = stage[d][cic_slot(j)];
            const float f = floorf(u);
            c0[d] = static_cast<int>(f);
            w1[d] = u - f;
        }
        if (c0[0] != cur[0] || c0[1] != cur[1] || c0[2] != cur[2]) {
            if (cur[0] != 0x7fffffff) cic_add(win, k0, cur, acc, nf, grid);
#pragma unroll
            for (int d = 0; d < 3; ++d) cur[d] = c0[d];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] = 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            acc[q] += ((q & 1) ? w1[0] : 1.0f - w1[0]) * ((q & 2) ? w1[1] : 1.0f - w1[1]) *
                      ((q & 4) ? w1[2] : 1.0f - w1[2]);
    }
    // Lanes whose runs end on the same corner are adjacent: sum each such
    // segment of lanes into its first lane (suffix sums within segments), so
    // one lane adds per corner instead of all of them onto the same words.
    const int lane = threadIdx.x % kWave;
    const bool head = lane == 0 || cur[0] != __shfl_up(cur[0], 1) || cur[1] != __shfl_up(cur[1], 1) ||
                      cur[2] != __shfl_up(cur[2], 1);
    const uint64_t heads = __ballot(head);
    const int seg = __popcll(heads & (lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1)));
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int so = __shfl_down(seg, o);
        const bool take = lane + o < kWave && so == seg;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float v = __shfl_down(acc[q], o);
            if (take) acc[q] += v;
        }
    }
    if (head && cur[0] != 0x7fffffff) cic_add(win, k0, cur, acc, nf, grid);
    __syncthreads();
    for (int c = threadIdx.x; c < kCicWin * kCicWin * kCicWin; c += kCicThreads) {
        const double v = win[c];
        if (v == 0.0) continue;
        const int cx = min(max(k0[0] + c % kCicWin, 0), nf - 1);
        const int cy = min(max(k0[1] + (c / kCicWin) % kCicWin, 0), nf - 1);
        const int cz = min(max(k0[2] + c / (kCicWin * kCicWin), 0), nf - 1);
        atomicAdd(&grid[(static_cast<int64_t>(cz) * nf + cy) * nf + cx], static_cast<float>(v));
    }
}

// grid <- ln(grid / mean + kLnFloor).
__global__ void ln_grid_kernel(float* __restrict__ grid, int64_t ncell, float inv_mean)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < ncell) grid[i] = logf(grid[i] * inv_mean + kMlpLnFloor);
}

__device__ __forceinline__ void gather_features(const float* __restrict__ lngrid, int nf, float left, float inv_cell,
                                                float px, float py, float pz, float (&x)[kIn])
{
    const int cx = min(max(static_cast<int>(floorf((px - left) * inv_cell)), 0), nf - 1);
    const int cy = min(max(static_cast<int>(floorf((py - left) * inv_cell)), 0), nf - 1);
    const int cz = min(max(static_cast<int>(floorf((pz - left) * inv_cell)), 0), nf - 1);
    int k = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int gx = min(max(cx + dx, 0), nf - 1);
                const int gy = min(max(cy + dy, 0), nf - 1);
                const int gz = min(max(cz + dz, 0), nf - 1);
                x[k++] = lngrid[(static_cast<int64_t>(gz) * nf + gy) * nf + gx];
            }
}

// Features of the particles at sorted positions sidx[0..m).
__global__ void features_kernel(const float4* __restrict__ spos, const int32_t* __restrict__ sidx, int64_t m,
                                const float* __restrict__ lngrid, int nf, float left, float inv_cell,
                                float* __restrict__ x)
{
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= m) return;
    const float4 p = spos[sidx[b]];
    float xv[kIn];
    gather_features(lngrid, nf, left, inv_cell, p.x, p.y, p.z, xv);
#pragma unroll
    for (int k = 0; k < kIn; ++k) x[b * kIn + k] = xv[k];
}

// The features of a particle are those of its cell, so the model is
// evaluated once per cell: rcell[c] = exp(f(x_c)) * scale.
__global__ __launch_bounds__(256) void seed_cell_kernel(const float* __restrict__ lngrid, int nf,
                                                        const float* __restrict__ params, float scale,
                                                        float* __restrict__ rcell)
{
    __shared__ float sp[kMlpParams];
    load_params(sp, params);
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncell = static_cast<int64_t>(nf) * nf * nf;
    if (c >= ncell) return;
    const int cx = static_cast<int>(c % nf), cy = static_cast<int>((c / nf) % nf), cz = static_cast<int>(c / nf / nf);
    float xv[kIn];
    int k = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int gx = min(max(cx + dx, 0), nf - 1);
                const int gy = min(max(cy + dy, 0), nf - 1);
                const int gz = min(max(cz + dz, 0), nf - 1);
                xv[k++] = lngrid[(static_cast<int64_t>(gz) * nf + gy) * nf + gx];
            }
    rcell[c] = __expf(mlp_eval(sp, xv, nullptr, nullptr)) * scale;
}

// tau_init[q] = rcell[cell of q]^2.
__global__ void seed_bound_kernel(const float4* __restrict__ spos, int64_t n, const float* __restrict__ rcell, int nf,
                                  float left, float inv_cell, float* __restrict__ tau_init)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float4 p = spos[q];
    const int cx = min(max(static_cast<int>(floorf((p.x - left) * inv_cell)), 0), nf - 1);
    const int cy = min(max(static_cast<int>(floorf((p.y - left) * inv_cell)), 0), nf - 1);
    const int cz = min(max(static_cast<int>(floorf((p.z - left) * inv_cell)), 0), nf - 1);
    const float r = rcell[(static_cast<int64_t>(cz) * nf + cy) * nf + cx];
    tau_init[q] = r * r;
}

// Training samples: sorted positions of caller indices, and ln(h / spacing).
__global__ void seed_index_kernel(const int32_t* __restrict__ inv_perm, const int64_t* __restrict__ idx, int64_t m,
                                  int32_t* __restrict__ sidx)
{
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b < m) sidx[b] = inv_perm[idx[b]];
}

__global__ void seed_target_kernel(const float* __restrict__ rk_sorted, const int32_t* __restrict__ sidx, int64_t m,
                                   float inv_spacing, float* __restrict__ target)
{
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b < m) target[b] = logf(rk_sorted[sidx[b]] * inv_spacing);
}

}  // namespace

hipError_t launch_seed_samples(const float4* spos, const float* rk_sorted, const int32_t* inv_perm,
                               const int64_t* idx, int64_t m, const float* lngrid, float left, float right,
                               float inv_spacing, int32_t* sidx, float* x, float* target, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((m + 255) / 256);
    seed_index_kernel<<<blocks, 256, 0, s>>>(inv_perm, idx, m, sidx);
    ASM_HIP_CHECK(hipGetLastError());
    ASM_HIP_CHECK(launch_seed_features(spos, sidx, m, lngrid, left, right, x, s));
    seed_target_kernel<<<blocks, 256, 0, s>>>(rk_sorted, sidx, m, inv_spacing, target);
    return hipGetLastError();
}

hipError_t launch_seed_grid(const float4* spos, int64_t n, float left, float right, float* grid, hipStream_t s)
{
    const int nf = kMlpGrid;
    const int64_t ncell = static_cast<int64_t>(nf) * nf * nf;
    ASM_HIP_CHECK(launch_fill_zero_f32(grid, ncell, s));
    const float inv_cell = static_cast<float>(nf) / (right - left);
    const int64_t per_block = static_cast<int64_t>(kCicThreads) * kCicRun;
    if (n > 0)
        cic_kernel<<<static_cast<unsigned>((n + per_block - 1) / per_block), kCicThreads, 0, s>>>(spos, n, left,
                                                                                                 inv_cell, nf, grid);
    ASM_HIP_CHECK(hipGetLastError());
    ln_grid_kernel<<<static_cast<unsigned>((ncell + 255) / 256), 256, 0, s>>>(grid, ncell,
                                                                               static_cast<float>(ncell) /
                                                                                   static_cast<float>(n > 0 ? n : 1));
    return hipGetLastError();
}

hipError_t launch_seed_features(const float4* spos, const int32_t* sidx, int64_t m, const float* lngrid, float left,
                                float right, float* x, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    const int nf = kMlpGrid;
    features_kernel<<<static_cast<unsigned>((m + 255) / 256), 256, 0, s>>>(spos, sidx, m, lngrid, nf, left,
                                                                           static_cast<float>(nf) / (right - left), x);
    return hipGetLastError();
}

hipError_t launch_seed_bound(const float4* spos, int64_t n, const float* lngrid, float left, float right,
                             const float* params, float scale, float* rcell, float* tau_init, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int nf = kMlpGrid;
    const int64_t ncell = static_cast<int64_t>(nf) * nf * nf;
    seed_cell_kernel<<<static_cast<unsigned>((ncell + 255) / 256), 256, 0, s>>>(lngrid, nf, params, scale, rcell);
    ASM_HIP_CHECK(hipGetLastError());
    seed_bound_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(
        spos, n, rcell, nf, left, static_cast<float>(nf) / (right - left), tau_init);
    return hipGetLastError();
}

}  // namespace asm_hip

using namespace asm_hip;

struct asm_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float* params = nullptr;   // device, kMlpParams
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
};

namespace {

int mlp_fail(asm_mlp* m, int code, const std::string& msg)
{
    if (m) m->err = msg;
    return code;
}

#define MLPC(expr)                                                                     \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) return mlp_fail(m, ASM_ERR_HIP, hipGetErrorString(_e));  \
    } while (0)

// Device scratch of one batch.
struct Batch {
    float *x = nullptr, *t = nullptr, *y = nullptr, *z1 = nullptr, *z2 = nullptr, *dz1 = nullptr, *dz2 = nullptr,
          *dy = nullptr, *loss = nullptr, *wpart = nullptr;
    int64_t* idx = nullptr;
    ~Batch()
    {
        for (void* p : {static_cast<void*>(x), static_cast<void*>(t), static_cast<void*>(y), static_cast<void*>(z1),
                        static_cast<void*>(z2), static_cast<void*>(dz1), static_cast<void*>(dz2),
                        static_cast<void*>(dy), static_cast<void*>(loss), static_cast<void*>(idx),
                        static_cast<void*>(wpart)})
            if (p) (void)hipFree(p);
    }
    hipError_t alloc(int64_t nrows, int64_t nb)
    {
        hipError_t e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&x), sizeof(float) * nrows * kIn)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&t), sizeof(float) * nrows)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&y), sizeof(float) * nb)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&z1), sizeof(float) * nb * kH)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&z2), sizeof(float) * nb * kH)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&dz1), sizeof(float) * nb * kH)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&dz2), sizeof(float) * nb * kH)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&dy), sizeof(float) * nb)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&loss), sizeof(float) * nb)) != hipSuccess) return e;
        if ((e = hipMalloc(reinterpret_cast<void**>(&wpart), sizeof(float) * kWgradPartial)) != hipSuccess) return e;
        return hipMalloc(reinterpret_cast<void**>(&idx), sizeof(int64_t) * nb);
    }
};

// Gradient of the mean squared error over rows idx[0..nb) (or 0..nb) into
// m->grad; per-sample squared errors into B.loss.
hipError_t batch_grad(asm_mlp* m, Batch& B, const int64_t* idx, int64_t nb)
{
    const unsigned blocks = static_cast<unsigned>((nb + 255) / 256);
    mlp_forward_kernel<<<blocks, 256, 0, m->stream>>>(m->params, B.x, idx, nb, B.y, B.z1, B.z2);
    ASM_HIP_CHECK(hipGetLastError());
    mlp_backward_kernel<<<blocks, 256, 0, m->stream>>>(m->params, B.t, idx, nb, B.y, B.z1, B.z2, B.dz1, B.dz2, B.dy,
                                                       B.loss);
    ASM_HIP_CHECK(hipGetLastError());
    mlp_wgrad_kernel<<<kWgradBlocks * kWgradSplit, kWave, 0, m->stream>>>(B.x, idx, nb, B.z1, B.z2, B.dz1, B.dz2,
                                                                         B.dy, B.wpart);
    ASM_HIP_CHECK(hipGetLastError());
    constexpr int kRed = kWgradBlocks * kWave * 16;
    mlp_wgrad_reduce_kernel<<<(kRed + 255) / 256, 256, 0, m->stream>>>(B.wpart, m->grad);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int asm_mlp_param_count(void)
{
    return kMlpParams;
}

int asm_mlp_create(asm_mlp** out, int device)
{
    if (!out) return ASM_ERR_ARG;
    *out = nullptr;
    asm_mlp* m = new asm_mlp();
    m->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&m->stream);
    for (float** p : {&m->params, &m->grad, &m->adam_m, &m->adam_v})
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), sizeof(float) * kMlpParams);
    for (float* p : {m->params, m->grad, m->adam_m, m->adam_v})
        if (e == hipSuccess) e = launch_fill_zero_f32(p, kMlpParams, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        asm_mlp_destroy(m);
        return ASM_ERR_HIP;
    }
    *out = m;
    return ASM_OK;
}

void asm_mlp_destroy(asm_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    for (float* p : {m->params, m->grad, m->adam_m, m->adam_v})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* asm_mlp_last_error(const asm_mlp* m)
{
    return m ? m->err.c_str() : "null model";
}

int asm_mlp_set_params(asm_mlp* m, const float* params)
{
    if (!m || !params) return ASM_ERR_ARG;
    MLPC(hipSetDevice(m->device));
    MLPC(hipMemcpyAsync(m->params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    MLPC(launch_fill_zero_f32(m->adam_m, kMlpParams, m->stream));
    MLPC(launch_fill_zero_f32(m->adam_v, kMlpParams, m->stream));
    m->adam_t = 0;
    MLPC(hipStreamSynchronize(m->stream));
    return ASM_OK;
}

int asm_mlp_get_params(asm_mlp* m, float* params)
{
    if (!m || !params) return ASM_ERR_ARG;
    MLPC(hipSetDevice(m->device));
    MLPC(hipMemcpyAsync(params, m->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    MLPC(hipStreamSynchronize(m->stream));
    return ASM_OK;
}

int asm_mlp_forward(asm_mlp* m, const float* x, int64_t n, float* y)
{
    if (!m || (n > 0 && (!x || !y)) || n < 0) return ASM_ERR_ARG;
    if (n == 0) return ASM_OK;
    MLPC(hipSetDevice(m->device));
    float *dx = nullptr, *dy = nullptr;
    MLPC(hipMalloc(reinterpret_cast<void**>(&dx), sizeof(float) * n * kIn));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dy), sizeof(float) * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        mlp_forward_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, m->stream>>>(m->params, dx, nullptr, n,
                                                                                         dy, nullptr, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return mlp_fail(m, ASM_ERR_HIP, hipGetErrorString(e));
    return ASM_OK;
}

int asm_mlp_loss_grad(asm_mlp* m, const float* x, const float* t, int64_t n, double* loss, float* grad)
{
    if (!m || !x || !t || n <= 0) return ASM_ERR_ARG;
    MLPC(hipSetDevice(m->device));
    Batch B;
    MLPC(B.alloc(n, n));
    MLPC(hipMemcpyAsync(B.x, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice, m->stream));
    MLPC(hipMemcpyAsync(B.t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    MLPC(batch_grad(m, B, nullptr, n));
    std::vector<float> lp(static_cast<size_t>(n));
    MLPC(hipMemcpyAsync(lp.data(), B.loss, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    if (grad) MLPC(hipMemcpyAsync(grad, m->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    MLPC(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = s / static_cast<double>(n);
    }
    return ASM_OK;
}

int asm_mlp_train(asm_mlp* m, const float* x, const float* t, int64_t n, int epochs, int batch, float lr,
                  uint64_t seed, float* epoch_loss)
{
    if (!m || !x || !t || n <= 0 || epochs <= 0 || batch <= 0) return ASM_ERR_ARG;
    MLPC(hipSetDevice(m->device));
    Batch B;
    const int64_t nb = batch < n ? batch : n;
    MLPC(B.alloc(n, nb));
    MLPC(hipMemcpyAsync(B.x, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice, m->stream));
    MLPC(hipMemcpyAsync(B.t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    std::vector<int64_t> perm(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) perm[static_cast<size_t>(i)] = i;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&st]() {   // xorshift64*
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return st * 0x2545F4914F6CDD1Dull;
    };
    std::vector<float> lp(static_cast<size_t>(nb));
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[static_cast<size_t>(i)], perm[rnd() % (i + 1)]);
        double esum = 0.0;
        int64_t ecount = 0;
        for (int64_t b0 = 0; b0 + nb <= n; b0 += nb) {
            MLPC(hipMemcpyAsync(B.idx, perm.data() + b0, sizeof(int64_t) * nb, hipMemcpyHostToDevice, m->stream));
            MLPC(batch_grad(m, B, B.idx, nb));
            ++m->adam_t;
            const float bc1 = 1.0f - powf(beta1, static_cast<float>(m->adam_t));
            const float bc2 = 1.0f - powf(beta2, static_cast<float>(m->adam_t));
            adam_kernel<<<(kMlpParams + 255) / 256, 256, 0, m->stream>>>(m->params, m->grad, m->adam_m, m->adam_v, lr,
                                                                         beta1, beta2, eps, bc1, bc2);
            MLPC(hipGetLastError());
            MLPC(hipMemcpyAsync(lp.data(), B.loss, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream));
            MLPC(hipStreamSynchronize(m->stream));
            for (float v : lp) esum += v;
            ecount += nb;
        }
        if (epoch_loss) epoch_loss[ep] = static_cast<float>(esum / static_cast<double>(ecount > 0 ? ecount : 1));
    }
    return ASM_OK;
}

}  // extern "C"
