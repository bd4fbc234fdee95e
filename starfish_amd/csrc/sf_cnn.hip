// Learned initial guess for the potential solve: a 3-level encoder-decoder CNN
// (32/64/128 channels, 3x3 convolutions, BF16 weights and activations, FP32 accumulation on
// the gfx950 matrix cores) mapping (rho, Dirichlet mask, boundary phi) to phi.  Training runs
// on the GPU with the kernels below (forward, backward-data, backward-weights, Adam).  The
// network only supplies the starting point of the Newton/PCG solve; every solve still runs to
// the requested relative residual, so plain PCG remains the fallback and the judge.
//
// Layout: activations are NHWC BF16 tensors addressed through views (pointer, channel stride,
// channel offset) so that skip connections concatenate without copies.  3x3 convolutions are
// implicit GEMMs on v_mfma_f32_16x16x32_bf16: M = 16 pixels of a row, N = 16 output channels,
// K = 32 input channels of one tap.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "sf_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace sf {

__device__ __host__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);   // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                     // round to nearest even
    return (uint16_t)(u >> 16);
}
// two floats to packed BF16 (round to nearest even, as f2bf for finite values) with the
// hardware conversion
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    const __bf16 x = (__bf16)a, y = (__bf16)b;
    uint16_t ux, uy;
    memcpy(&ux, &x, 2);
    memcpy(&uy, &y, 2);
    return (uint32_t)ux | ((uint32_t)uy << 16);
}
__device__ __host__ __forceinline__ float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ReLU backward of 4 channels: g where the BF16 activation (4 values packed in a) is positive, else 0
__device__ __forceinline__ float4 relu_mask4(const float4 g, const uint2 a) {
    return make_float4(bf2f((uint16_t)(a.x & 0xFFFFu)) > 0.f ? g.x : 0.f, bf2f((uint16_t)(a.x >> 16)) > 0.f ? g.y : 0.f,
                       bf2f((uint16_t)(a.y & 0xFFFFu)) > 0.f ? g.z : 0.f, bf2f((uint16_t)(a.y >> 16)) > 0.f ? g.w : 0.f);
}
__device__ __forceinline__ uint2 pack_bf16x4(const float4 e) {
    return make_uint2(pack_bf16x2(e.x, e.y), pack_bf16x2(e.z, e.w));
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    bf16x8 r;
    memcpy(&r, &v, 16);
    return r;
}
// u where ok, else zero: a select on the values (a select between the addresses of two locals
// would keep them in scratch memory)
__device__ __forceinline__ uint4 sel4(bool ok, const uint4 u) {
    return make_uint4(ok ? u.x : 0u, ok ? u.y : 0u, ok ? u.z : 0u, ok ? u.w : 0u);
}
__device__ __forceinline__ bf16x8 zero8() {
    uint4 v = make_uint4(0, 0, 0, 0);
    bf16x8 r;
    memcpy(&r, &v, 16);
    return r;
}

// ------------------------------------------------------------------------------------------
// 3x3 convolution as an implicit GEMM (stride 1, zero padding 1).
//   forward:       out[p][n] = sum_tap sum_k in[p + d(tap)][k] * w[tap][n][k]   (+ bias, ReLU)
//   backward data: din[p][k'] = sum_tap sum_n' g[p - d(tap)][n'] * w[tap][n'][k'] with the weight
//                  stored transposed as wt[tap][k'][n'] -> the same kernel with sgn = -1.
// MODE 0: BF16 output with bias and optional ReLU; MODE 1: FP32 output (optionally accumulated);
// MODE 2: the ReLU backward of the previous layer fused: BF16 output act > 0 ? din : 0, with act
// that layer's BF16 activation.
// Requirements (checked on the host): K multiple of 32, N multiple of 16, channel strides and
// offsets multiples of 8 (16-byte operand loads).
// ------------------------------------------------------------------------------------------
struct View {
    uint16_t* p;
    int cs;   // channel stride (elements per pixel)
    int co;   // channel offset
};
struct FView {
    float* p;
    int cs, co;
};

// The operands are staged in LDS.  One workgroup: R rows of CONV_TP pixels x TN output channels
// (TN = 64: waves 2 x 2 over pixels x channels; TN = 32: 4 x 1).  For each chunk of 32 input
// channels the R + 2 input rows of CONV_TP + 2 pixels and the 9 x TN x 32 weights go to LDS once,
// and every wave runs its 16 x 16 tiles of all R rows over the 9 taps from there: a weight fragment
// is read once per tap for R x PT pixel tiles, an input fragment once for two channel tiles.  The
// next chunk's operands are loaded into registers while the current one is multiplied.  The
// weights are the MFMA's A operand (rows = output channels) and the pixels its B operand, so a
// lane's result holds 4 consecutive channels of one pixel and leaves in one 8-byte (BF16) or
// 16-byte (FP32) store.
constexpr int CONV_TP = 128;
template <int MODE, int TN, int R>
__global__ void __launch_bounds__(256, 2) conv3x3_lds(const uint16_t* __restrict__ in, int in_cs, int in_co, int K,
                                                   const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                   int N, uint16_t* __restrict__ out, float* __restrict__ fout,
                                                   int out_cs, int out_co, int H, int Wd, int relu, int sgn,
                                                   int accumulate, const uint16_t* __restrict__ act, int act_cs,
                                                   int act_co, int nrt) {
    constexpr int WN = TN / 32;                 // waves along the channels (two 16-channel tiles each)
    constexpr int WP = 4 / WN;                  // waves along the pixels
    constexpr int PT = CONV_TP / 16 / WP;       // 16-pixel tiles per wave and row
    constexpr int IW = CONV_TP + 2;             // staged pixels per row
    constexpr int NIN = (R + 2) * IW * 4;       // 16-byte groups of the input stage (8 channels each)
    constexpr int NWT = 9 * TN * 4;             // 16-byte groups of the weight stage
    constexpr int RIN = (NIN + 255) / 256, RWT = (NWT + 255) / 256;
    __shared__ uint4 s_in[NIN];                 // [row][pixel][group]
    __shared__ uint4 s_w[NWT];                  // [tap][channel][group]
    const int ntn = (N + TN - 1) / TN;
    // the workgroup walks nrt tiles of R rows down its column; step it is chunk it % nkc of row
    // tile it / nkc, so the next tile's first chunk is loaded while the current tile is multiplied
    const int x0 = blockIdx.x * CONV_TP, y_first = blockIdx.y * nrt * R;
    const int nkc = K / 32;
    const int steps = min(nrt, (H - y_first + R - 1) / R) * nkc;
    const int b = blockIdx.z / ntn, n0 = (blockIdx.z % ntn) * TN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wp = wave % WP, wn = wave / WP;
    const int prow = lane & 15, kq = lane >> 4;
    uint4 rin[RIN], rwt[RWT];
    auto fetch = [&](int it) __attribute__((always_inline)) {
        const int y0 = y_first + (it / nkc) * R, k0 = (it % nkc) * 32;
#pragma unroll
        for (int q = 0; q < RIN; q++) {
            const int e = threadIdx.x + 256 * q;
            const int g = e & 3, px = (e >> 2) % IW, r = (e >> 2) / IW;
            const int yy = y0 - 1 + r, xx = x0 - 1 + px;
            const bool ok = e < NIN && yy >= 0 && yy < H && xx >= 0 && xx < Wd;
            const size_t o = ok ? (((size_t)b * H + yy) * Wd + xx) * in_cs + in_co + k0 + g * 8 : 0;
            rin[q] = sel4(ok, *reinterpret_cast<const uint4*>(in + o));
        }
#pragma unroll
        for (int q = 0; q < RWT; q++) {
            const int e = threadIdx.x + 256 * q;
            const int g = e & 3, n = (e >> 2) % TN, tap = (e >> 2) / TN;
            const bool ok = e < NWT && n0 + n < N;
            const size_t o = ok ? ((size_t)tap * N + n0 + n) * K + k0 + g * 8 : 0;
            rwt[q] = sel4(ok, *reinterpret_cast<const uint4*>(w + o));
        }
    };
    f32x4 acc[R][PT][2];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int p = 0; p < PT; p++) acc[r][p][0] = acc[r][p][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (steps > 0) fetch(0);
    for (int it = 0; it < steps; it++) {
        __syncthreads();   // the previous chunk's fragments are read
#pragma unroll
        for (int q = 0; q < RIN; q++)
            if (threadIdx.x + 256 * q < NIN) s_in[threadIdx.x + 256 * q] = rin[q];
#pragma unroll
        for (int q = 0; q < RWT; q++)
            if (threadIdx.x + 256 * q < NWT) s_w[threadIdx.x + 256 * q] = rwt[q];
        __syncthreads();
        if (it + 1 < steps) fetch(it + 1);
        // (taps not unrolled: the R x PT x 2 tiles of a tap are, and the accumulators stay in
        // registers)
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dy = sgn * (tap / 3 - 1), dx = sgn * (tap % 3 - 1);
            bf16x8 bw[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const uint4 v = s_w[(tap * TN + wn * 32 + t * 16 + prow) * 4 + kq];
                memcpy(&bw[t], &v, 16);
            }
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int p = 0; p < PT; p++) {
                    const int px = (wp * PT + p) * 16 + prow;
                    const uint4 v = s_in[((r + 1 + dy) * IW + px + 1 + dx) * 4 + kq];
                    bf16x8 a;
                    memcpy(&a, &v, 16);
                    acc[r][p][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[0], a, acc[r][p][0], 0, 0, 0);
                    acc[r][p][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[1], a, acc[r][p][1], 0, 0, 0);
                }
        }
        if (it % nkc != nkc - 1) continue;
        // the row tile's last chunk: store its outputs and clear the accumulators
        const int y0 = y_first + (it / nkc) * R;
        // D layout: lane holds rows (lane/16)*4 + i (channels), column lane%16 (pixel)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = y0 + r;
            if (y >= H) continue;
#pragma unroll
            for (int p = 0; p < PT; p++) {
                const int xo = x0 + (wp * PT + p) * 16 + prow;
                if (xo >= Wd) continue;
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int n = n0 + wn * 32 + t * 16 + kq * 4;   // this lane's 4 channels n .. n + 3
                    if (n >= N) continue;
                    const size_t o = (((size_t)b * H + y) * Wd + xo) * out_cs + out_co + n;
                    const f32x4 v = acc[r][p][t];
                    if (MODE == 0) {
                        float e[4];
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            e[i] = v[i] + bias[n + i];
                            if (relu) e[i] = fmaxf(e[i], 0.f);
                        }
                        *reinterpret_cast<uint2*>(out + o) = make_uint2(pack_bf16x2(e[0], e[1]), pack_bf16x2(e[2], e[3]));
                    } else if (MODE == 2) {
                        const uint2 a = *reinterpret_cast<const uint2*>(
                            act + (((size_t)b * H + y) * Wd + xo) * act_cs + act_co + n);
                        *reinterpret_cast<uint2*>(out + o) = pack_bf16x4(relu_mask4(make_float4(v[0], v[1], v[2], v[3]), a));
                    } else {
                        float4* f = reinterpret_cast<float4*>(fout + o);
                        if (accumulate) {
                            const float4 g = *f;
                            *f = make_float4(g.x + v[0], g.y + v[1], g.z + v[2], g.w + v[3]);
                        } else {
                            *f = make_float4(v[0], v[1], v[2], v[3]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int p = 0; p < PT; p++) acc[r][p][0] = acc[r][p][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

// Weight gradient with the operands staged in LDS.  One workgroup: 32 input x 32 output channels
// for all 9 taps over a run of chunks, a chunk being 64 pixels of one row; the chunks of a run go
// down the rows of a 64-pixel column band, so the three input rows a chunk needs (y - 1, y, y + 1)
// stay in a ring of three LDS rows and each chunk stages one new input row (three at the start of
// a run or a band) and its gradient row.  Rows are staged transposed to [channel][pixel], each in
// three copies shifted by one pixel (dx = -1, 0, 1), so an operand fragment (8 consecutive pixels
// of one channel) is one 16-byte LDS read; a row holds 16 four-pixel entries plus 2 of padding, so
// the 16 channels of a fragment read start on distinct banks.  Wave w takes the pixels
// 32 (w & 1) .. + 31 of every chunk (one MFMA reduction step) and input channels 16 (w >> 1) .. + 15
// against all 32 output channels: 9 input and 2 gradient fragments per 18 MFMAs.  The next
// chunk's rows are loaded into registers while the current one is multiplied.  The waves of the
// first input-channel tile also sum the gradient over the pixels (the bias gradient) with one
// more MFMA per output tile against an all-ones operand.  At the end the two pixel halves are
// added through LDS and the workgroup stores its sums as a partial (WG_PART floats: [tap][k][n],
// then the 32 bias sums); wgrad_reduce adds the partials of a channel tile pair in a fixed order,
// which is deterministic and avoids thousands of workgroups' atomics on the same addresses.
// The kernel is held to 168 registers (three waves per SIMD: three workgroups per CU, as many as
// its 46 KB of LDS allow); with 174 it ran two and took 243 instead of 141 us per launch.
// Requirements: K and N multiples of 32, strides and offsets multiples of 8.
constexpr int WG_PX = 64;           // pixels per chunk (two MFMA reduction steps)
constexpr int WG_Q = WG_PX / 4 + 2; // four-pixel entries per staged row (with padding)
constexpr int WG_PART = 9 * 32 * 32 + 32;   // floats of a workgroup's partial
constexpr int WG_BLOCKS = 768;              // target workgroups of a weight-gradient launch (3 per CU)
__device__ __forceinline__ uint32_t bf_at(const uint4& v, int c) {
    const uint32_t w = c < 2 ? v.x : (c < 4 ? v.y : (c < 6 ? v.z : v.w));
    return (c & 1) ? (w >> 16) : (w & 0xFFFFu);
}
__global__ void __launch_bounds__(256, 3) conv3x3_wgrad_lds(const uint16_t* __restrict__ in, int in_cs, int in_co,
                                                         int K, const uint16_t* __restrict__ g, int g_cs, int N,
                                                         float* __restrict__ part, int H, int Wd, int B,
                                                         int chunks_per_block) {
    __shared__ uint2 s_in[3][3][32][WG_Q];   // [dx + 1][row % 3][input channel][pixel / 4]
    __shared__ uint2 s_g[32][WG_Q];          // [output channel][pixel / 4]
    const int nkt = K / 32;
    const int k0 = (blockIdx.y % nkt) * 32, n0 = (blockIdx.y / nkt) * 32;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int ks = wave & 1, tk = wave >> 1;   // pixel half (reduction step), input-channel half
    const int prow = lane & 15, kq = lane >> 4;
    const int cpr = (Wd + WG_PX - 1) / WG_PX;
    const int nchunks = B * H * cpr;
    const int c_begin = blockIdx.x * chunks_per_block;
    const int c_end = min(c_begin + chunks_per_block, nchunks);
    // chunk c: row y = c % H of column band (c / H) % cpr of image c / (H cpr)
    // staging items: waves 0..2 an input row each (row y - 1 + r; in a run's steady state only
    // r = 2, the new row y + 1), wave 3 the gradient row y; a thread's item is the pixel quad q
    // (pixels 4q - 1 .. 4q + 4 of 8 channels for the input, 4q .. 4q + 3 for the gradient) of
    // channel group gq.  The quad is the fastest index, so the 16 lanes of an LDS write phase cover
    // all banks.
    const bool is_in = wave < 3;
    const int r = wave, q = lane & 15, gq = lane >> 4;
    uint4 v[6];
    // (the same loads for both kinds of item, so that v stays in registers)
    const uint16_t* const src = is_in ? in : g;
    const int scs = is_in ? in_cs : g_cs, sco = is_in ? in_co + k0 + gq * 8 : n0 + gq * 8;
    const int sdx = is_in ? -1 : 0, sj = is_in ? 6 : 4;
    // rows of chunk c this thread stages: the run's first chunk and the first row of a band
    // stage all three input rows
    auto stages = [&](int c) { return !is_in || r == 2 || c == c_begin || c % H == 0; };
    auto fetch = [&](int c) {
        const int y = c % H, band = c / H;
        const int b = band / cpr, x0 = (band % cpr) * WG_PX;
        const int yy = is_in ? y - 1 + r : y;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const int xx = x0 + 4 * q + sdx + j;
            const bool ok = j < sj && yy >= 0 && yy < H && xx >= 0 && xx < Wd;
            const size_t o = ok ? (((size_t)b * H + yy) * Wd + xx) * scs + sco : 0;
            v[j] = sel4(ok, *reinterpret_cast<const uint4*>(src + o));
        }
    };
    f32x4 acc[9][2];
#pragma unroll
    for (int tap = 0; tap < 9; tap++) acc[tap][0] = acc[tap][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    f32x4 accb[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
    const bool do_bias = k0 == 0 && tk == 0;   // uniform per wave
    bf16x8 ones;
    {
        const uint32_t o2 = 0x3F803F80u;   // two BF16 1.0
        const uint4 u = make_uint4(o2, o2, o2, o2);
        memcpy(&ones, &u, 16);
    }
    if (c_begin < c_end && stages(c_begin)) fetch(c_begin);
    for (int c = c_begin; c < c_end; c++) {
        const int y = c % H;
        __syncthreads();   // the previous chunk's fragments are read
        if (stages(c)) {
            if (is_in) {
                const int slot = (y - 1 + r + 3) % 3;
#pragma unroll
                for (int d = 0; d < 3; d++)   // copy d holds pixels 4q + d - 1 + i: v[d + i]
#pragma unroll
                    for (int ch = 0; ch < 8; ch++)
                        s_in[d][slot][gq * 8 + ch][q] = make_uint2(bf_at(v[d], ch) | (bf_at(v[d + 1], ch) << 16),
                                                                   bf_at(v[d + 2], ch) | (bf_at(v[d + 3], ch) << 16));
            } else {
#pragma unroll
                for (int ch = 0; ch < 8; ch++)
                    s_g[gq * 8 + ch][q] = make_uint2(bf_at(v[0], ch) | (bf_at(v[1], ch) << 16),
                                                     bf_at(v[2], ch) | (bf_at(v[3], ch) << 16));
            }
        }
        __syncthreads();
        if (c + 1 < c_end && stages(c + 1)) fetch(c + 1);
        bf16x8 bg[2];
#pragma unroll
        for (int n = 0; n < 2; n++) {
            const uint2 lo = s_g[n * 16 + prow][8 * ks + 2 * kq], hi = s_g[n * 16 + prow][8 * ks + 2 * kq + 1];
            const uint4 u = make_uint4(lo.x, lo.y, hi.x, hi.y);
            memcpy(&bg[n], &u, 16);
            if (do_bias) accb[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, bg[n], accb[n], 0, 0, 0);
        }
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;   // copy indices (offset + 1)
            const int slot = (y - 1 + dy + 3) % 3;
            const uint2 lo = s_in[dx][slot][tk * 16 + prow][8 * ks + 2 * kq];
            const uint2 hi = s_in[dx][slot][tk * 16 + prow][8 * ks + 2 * kq + 1];
            const uint4 u = make_uint4(lo.x, lo.y, hi.x, hi.y);
            bf16x8 a;
            memcpy(&a, &u, 16);
            acc[tap][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bg[0], acc[tap][0], 0, 0, 0);
            acc[tap][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bg[1], acc[tap][1], 0, 0, 0);
        }
    }
    // the second pixel half's sums go through LDS (the input stage, 9 x 2 x 4 floats per lane and
    // wave, then the bias sums) to the first half's waves, which add them and store the partial
    __syncthreads();
    float* const red = reinterpret_cast<float*>(&s_in[0][0][0][0]);   // 2 x 80 x 64 floats
    static_assert(sizeof(s_in) >= 2 * 80 * 64 * sizeof(float), "reduction space");
    if (ks == 1) {
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int i = 0; i < 4; i++) red[((tk * 80) + (tap * 2 + n) * 4 + i) * 64 + lane] = acc[tap][n][i];
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int i = 0; i < 4; i++) red[((tk * 80) + 72 + n * 4 + i) * 64 + lane] = accb[n][i];
    }
    __syncthreads();
    if (ks == 1) return;
    // D rows (lane/16)*4+i = input channel, column lane%16 = output channel
    float* const pb = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * WG_PART;
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                pb[(tap * 32 + tk * 16 + kq * 4 + i) * 32 + n * 16 + prow] =
                    acc[tap][n][i] + red[((tk * 80) + (tap * 2 + n) * 4 + i) * 64 + lane];
    // every row of accb holds the column sums: lanes 0..15 store row 0
    if (tk == 0 && kq == 0)
#pragma unroll
        for (int n = 0; n < 2; n++)
            pb[9 * 32 * 32 + n * 16 + prow] = do_bias ? accb[n][0] + red[(72 + n * 4) * 64 + lane] : 0.f;
}

// Sum of the weight-gradient partials of channel tile pair y: a workgroup takes 32 entries e;
// slice j of its threads adds the partials b = j, j + 8, ... in order, and the 8 slice sums are
// added in order (a fixed order, so the result does not depend on timing).  Entry e < 9 * 32 * 32
// ([tap][k][n] of the tile) goes to dw, the rest (the bias sums, first input-channel tile only)
// to db.
__global__ void __launch_bounds__(256) wgrad_reduce(const float* __restrict__ part, int nblk, int K, int N,
                                                    float* __restrict__ dw, float* __restrict__ db) {
    __shared__ float red[8][32];
    const int nkt = K / 32;
    const int k0 = (blockIdx.y % nkt) * 32, n0 = (blockIdx.y / nkt) * 32;
    const int el = threadIdx.x & 31, slice = threadIdx.x >> 5;
    const int e = blockIdx.x * 32 + el;
    float s = 0.f;
    if (e < WG_PART) {
        const float* p = part + (size_t)blockIdx.y * nblk * WG_PART + e;
        for (int b = slice; b < nblk; b += 8) s += p[(size_t)b * WG_PART];
    }
    red[slice][el] = s;
    __syncthreads();
    if (slice != 0 || e >= WG_PART) return;
    s = red[0][el];
    for (int j = 1; j < 8; j++) s += red[j][el];
    if (e < 9 * 32 * 32) {
        const int tap = e / 1024, k = (e / 32) % 32, n = e % 32;
        dw[((size_t)tap * N + n0 + n) * K + k0 + k] += s;
    } else if (k0 == 0) {
        db[n0 + e - 9 * 32 * 32] += s;
    }
}



// 2x2 average pooling (forward BF16, backward FP32 accumulate)
// (8 channels per thread: C, strides and offsets multiples of 8)
__global__ void avgpool_fwd(const uint16_t* __restrict__ in, int in_cs, int in_co, int C, int H, int Wd, int B,
                            uint16_t* __restrict__ out, int out_cs, int out_co) {
    const int Ho = H / 2, Wo = Wd / 2, G = C / 8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)B * Ho * Wo * G;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = 8 * (int)(i % G);
        const int64_t q = i / G;
        const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), b = (int)(q / ((int64_t)Wo * Ho));
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; a++)
            for (int e = 0; e < 2; e++) {
                const uint4 v = *reinterpret_cast<const uint4*>(
                    in + (((size_t)b * H + 2 * yo + a) * Wd + 2 * xo + e) * in_cs + in_co + c);
#pragma unroll
                for (int k = 0; k < 8; k++) s[k] += bf2f((uint16_t)bf_at(v, k));
            }
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = (uint32_t)f2bf(0.25f * s[2 * k]) | ((uint32_t)f2bf(0.25f * s[2 * k + 1]) << 16);
        *reinterpret_cast<uint4*>(out + (((size_t)b * Ho + yo) * Wo + xo) * out_cs + out_co + c) =
            make_uint4(o[0], o[1], o[2], o[3]);
    }
}
// Backward of the 2x2 average pooling fused with the ReLU backward of the pooled activation: the
// gradient of a full-resolution pixel is its skip-path gradient (the view gskip) plus a quarter of
// its parent's (dout, [pixel][C]), masked by the activation (view act) and rounded to BF16 into
// g [pixel][C].  4 channels per thread; C, strides and offsets multiples of 4.
__global__ void avgpool_relu_bwd(const float* __restrict__ dout, int C, int H, int Wd, int B,
                                 const float* __restrict__ gskip, int s_cs, int s_co, const uint16_t* __restrict__ act,
                                 int a_cs, int a_co, uint16_t* __restrict__ g) {
    const int Ho = H / 2, Wo = Wd / 2, G = C / 4;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)B * H * Wd * G;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = 4 * (int)(i % G);
        const int64_t q = i / G;
        const int x = (int)(q % Wd), y = (int)((q / Wd) % H), b = (int)(q / ((int64_t)Wd * H));
        const float4 d = *reinterpret_cast<const float4*>(dout + (((size_t)b * Ho + y / 2) * Wo + x / 2) * C + c);
        const float4 sk = *reinterpret_cast<const float4*>(gskip + q * s_cs + s_co + c);
        const uint2 a = *reinterpret_cast<const uint2*>(act + q * a_cs + a_co + c);
        const float4 e = make_float4(sk.x + 0.25f * d.x, sk.y + 0.25f * d.y, sk.z + 0.25f * d.z, sk.w + 0.25f * d.w);
        *reinterpret_cast<uint2*>(g + q * C + c) = pack_bf16x4(relu_mask4(e, a));
    }
}
// nearest 2x upsampling (forward BF16 into a view, 8 channels per 16-byte access; backward FP32
// sum of the four children).  Forward: C, strides and offset multiples of 8.
__global__ void upsample_fwd(const uint16_t* __restrict__ in, int in_cs, int C, int Hi, int Wi, int B,
                             uint16_t* __restrict__ out, int out_cs, int out_co) {
    const int H = 2 * Hi, Wd = 2 * Wi, G = C / 8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)B * H * Wd * G;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(i % G);
        const int64_t q = i / G;
        const int x = (int)(q % Wd), y = (int)((q / Wd) % H), b = (int)(q / ((int64_t)Wd * H));
        *reinterpret_cast<uint4*>(out + (((size_t)b * H + y) * Wd + x) * out_cs + out_co + 8 * g) =
            *reinterpret_cast<const uint4*>(in + (((size_t)b * Hi + y / 2) * Wi + x / 2) * in_cs + 8 * g);
    }
}
// Backward of the nearest 2x upsampling fused with the ReLU backward of the upsampled activation:
// the sum of the four children's gradients (view dout, d_cs channels per pixel), masked by the
// activation (act, [pixel][C] at the low resolution) and rounded to BF16 into g [pixel][C].  4
// channels per thread; C and d_cs multiples of 4.
__global__ void upsample_relu_bwd(const float* __restrict__ dout, int d_cs, int C, int Hi, int Wi, int B,
                                  const uint16_t* __restrict__ act, uint16_t* __restrict__ g) {
    const int H = 2 * Hi, Wd = 2 * Wi, G = C / 4;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)B * Hi * Wi * G;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = 4 * (int)(i % G);
        const int64_t q = i / G;
        const int x = (int)(q % Wi), y = (int)((q / Wi) % Hi), b = (int)(q / ((int64_t)Wi * Hi));
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int a = 0; a < 2; a++)
            for (int e = 0; e < 2; e++) {
                const float4 v =
                    *reinterpret_cast<const float4*>(dout + (((size_t)b * H + 2 * y + a) * Wd + 2 * x + e) * d_cs + c);
                s = make_float4(s.x + v.x, s.y + v.y, s.z + v.z, s.w + v.w);
            }
        const uint2 av = *reinterpret_cast<const uint2*>(act + q * C + c);
        *reinterpret_cast<uint2*>(g + q * C + c) = pack_bf16x4(relu_mask4(s, av));
    }
}
// 1x1 head C -> 1 in FP32 on top of a base channel of the input: y = base + sum_c w[c] act[c] + b
constexpr int HEAD_C = 32;   // channels into the 1x1 head
__global__ void __launch_bounds__(256) head_fwd(const uint16_t* __restrict__ act, int64_t npix,
                                                const float* __restrict__ w, const float* __restrict__ bias,
                                                const uint16_t* __restrict__ base, int base_cs, int base_co,
                                                float* __restrict__ y) {
    float wv[HEAD_C];
#pragma unroll
    for (int c = 0; c < HEAD_C; c++) wv[c] = w[c];
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
        float s = bias[0] + bf2f(base[p * base_cs + base_co]);
        const uint4* a4 = reinterpret_cast<const uint4*>(act + p * HEAD_C);
#pragma unroll
        for (int q = 0; q < HEAD_C / 8; q++) {
            const uint4 u = a4[q];
#pragma unroll
            for (int k = 0; k < 8; k++) s += wv[8 * q + k] * bf2f((uint16_t)bf_at(u, k));
        }
        y[p] = s;
    }
}
// masked MSE loss: L = mean over counted pixels of (y - t)^2; dy = 2 (y - t) / count
__global__ void mse_loss(const float* __restrict__ y, const float* __restrict__ t, const uint8_t* __restrict__ m,
                         int64_t npix, float inv_count, float* __restrict__ dy, float* __restrict__ loss) {
    float acc = 0;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
        float e = m[p] ? (y[p] - t[p]) : 0.f;
        dy[p] = 2.f * e * inv_count;
        acc += e * e;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ float red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
        atomicAdd(loss, s * inv_count);   // one atomic per block (launched with at most 256 blocks)
    }
}
// head backward (C = HEAD_C channels): dw[c] += sum_p dy[p] act[p][c]; db += sum dy;
// gact[p][c] = act[p][c] > 0 ? dy[p] w[c] : 0 in BF16 (the ReLU backward of act fused).  Each
// thread keeps its channel sums in registers; the block adds them with shuffles and one atomic per
// channel.
__global__ void __launch_bounds__(256) head_bwd(const uint16_t* __restrict__ act, int64_t npix,
                                                const float* __restrict__ w, const float* __restrict__ dy,
                                                float* __restrict__ dw, float* __restrict__ db,
                                                uint16_t* __restrict__ gact) {
    __shared__ float red[HEAD_C + 1][4];
    float sw[HEAD_C], sdb = 0.f, wv[HEAD_C];
#pragma unroll
    for (int c = 0; c < HEAD_C; c++) {
        sw[c] = 0.f;
        wv[c] = w[c];
    }
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
        const float d = dy[p];
        const uint4* a4 = reinterpret_cast<const uint4*>(act + p * HEAD_C);
        uint2* o4 = reinterpret_cast<uint2*>(gact + p * HEAD_C);
#pragma unroll
        for (int q = 0; q < HEAD_C / 8; q++) {
            const uint4 u = a4[q];
            const uint32_t uw[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                sw[8 * q + 2 * e] += d * bf2f((uint16_t)(uw[e] & 0xFFFFu));
                sw[8 * q + 2 * e + 1] += d * bf2f((uint16_t)(uw[e] >> 16));
            }
            o4[2 * q] = pack_bf16x4(relu_mask4(
                make_float4(d * wv[8 * q], d * wv[8 * q + 1], d * wv[8 * q + 2], d * wv[8 * q + 3]), make_uint2(u.x, u.y)));
            o4[2 * q + 1] = pack_bf16x4(relu_mask4(
                make_float4(d * wv[8 * q + 4], d * wv[8 * q + 5], d * wv[8 * q + 6], d * wv[8 * q + 7]),
                make_uint2(u.z, u.w)));
        }
        sdb += d;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c <= HEAD_C; c++) {
        float v = c < HEAD_C ? sw[c] : sdb;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x <= HEAD_C) {
        const float v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
        atomicAdd(threadIdx.x < HEAD_C ? &dw[threadIdx.x] : db, v);
    }
}
// Adam (Kingma & Ba 2015) on FP32 master weights
// The parameter tensors as one index space for the optimizer: segment k holds elements
// [begin, begin + n) with its parameter, gradient and moment arrays.  One launch zeroes every
// gradient and one runs Adam on every parameter (instead of one memset and one Adam launch per
// tensor); each element finds its segment by binary search over the begins.
struct AdamSeg {
    float *p, *g, *m, *v;
    int64_t begin, n;
};
constexpr int ADAM_MAXSEG = 32;
__device__ __forceinline__ int adam_seg_of(const AdamSeg* s, int nseg, int64_t i) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s[mid].begin <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__global__ void __launch_bounds__(256) zero_grads(const AdamSeg* __restrict__ segs, int nseg, int64_t total) {
    __shared__ AdamSeg s[ADAM_MAXSEG];
    if (threadIdx.x < nseg) s[threadIdx.x] = segs[threadIdx.x];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int k = adam_seg_of(s, nseg, i);
        s[k].g[i - s[k].begin] = 0.f;
    }
}
__global__ void __launch_bounds__(256) adam_segs(const AdamSeg* __restrict__ segs, int nseg, int64_t total, float lr,
                                                 float b1, float b2, float eps, float c1, float c2) {
    __shared__ AdamSeg s[ADAM_MAXSEG];
    if (threadIdx.x < nseg) s[threadIdx.x] = segs[threadIdx.x];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int k = adam_seg_of(s, nseg, i);
        const int64_t j = i - s[k].begin;
        const float gi = s[k].g[j];
        const float mi = b1 * s[k].m[j] + (1 - b1) * gi;
        const float vi = b2 * s[k].v[j] + (1 - b2) * gi * gi;
        s[k].m[j] = mi;
        s[k].v[j] = vi;
        s[k].p[j] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}
// master FP32 weights [tap][n][k] (k < Kreal used, padded to K) -> BF16 forward [tap][n][K] and
// transposed backward [tap][k][n] copies
__global__ void pack_weights(const float* __restrict__ wm, int N, int K, uint16_t* __restrict__ wf,
                             uint16_t* __restrict__ wt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 9LL * N * K; i += (int64_t)gridDim.x * blockDim.x) {
        int k = (int)(i % K);
        int64_t q = i / K;
        int n = (int)(q % N), tap = (int)(q / N);
        uint16_t h = f2bf(wm[i]);
        wf[i] = h;
        wt[((size_t)tap * K + k) * N + n] = h;
    }
}

}  // namespace sf

// ==========================================================================================
// host: the network, its training loop and the C-ABI
// ==========================================================================================
namespace sf {

struct ConvL {
    int K = 0, N = 0, Kreal = 0;       // padded input channels, output channels, real inputs
    float *wm = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    uint16_t *wf = nullptr, *wt = nullptr;
    size_t nw() const { return (size_t)9 * N * K; }
};

}  // namespace sf

struct sf_cnn {
    sf_ctx* ctx = nullptr;
    int ni = 0, nj = 0;       // mesh
    int H = 0, W = 0;         // network grid (mesh padded to multiples of 4)
    int Bmax = 0;
    int C1 = sf::HEAD_C, C2 = 64, C3 = 128;
    sf::ConvL c1a, c1b, c2a, c2b, c3a, c3b, c4a, c4b, c5a, c5b;
    float *hw = nullptr, *hb = nullptr, *ghw = nullptr, *ghb = nullptr;
    float *mhw = nullptr, *vhw = nullptr, *mhb = nullptr, *vhb = nullptr;
    sf::AdamSeg* segs = nullptr;      // the optimizer's segment table (device)
    // activations (BF16 NHWC)
    uint16_t *x0 = nullptr, *e1a = nullptr, *u1 = nullptr, *p1 = nullptr, *e2a = nullptr, *u2 = nullptr,
             *p2 = nullptr, *e3a = nullptr, *e3 = nullptr, *d2a = nullptr, *d2 = nullptr, *d1a = nullptr,
             *d1 = nullptr;
    uint16_t *gop = nullptr, *gop2 = nullptr;   // BF16 operand gradients [npix][channels] (backward)
    float* y = nullptr;
    // FP32 gradients where two paths add
    float *g_u1 = nullptr, *g_p1 = nullptr, *g_u2 = nullptr, *g_p2 = nullptr, *dy = nullptr, *loss = nullptr;
    // dataset
    std::vector<uint16_t*> s_in;      // [H][W][4] BF16: rho, mask, phi_b, 0
    std::vector<float*> s_t;          // [H][W] target phi / phi_ref
    std::vector<uint8_t*> s_m;        // [H][W] 1 on free nodes (loss mask)
    uint8_t* bmask = nullptr;         // batch loss mask
    float* bt = nullptr;              // batch target
    double rho_ref = sf::QE * 1e16, phi_ref = 100.0;
    int adam_t = 0;
    std::mt19937_64 rng{1};
    double* phi_save = nullptr;       // [ni*nj] stash of a context's potential
    float* wpart = nullptr;           // weight-gradient partials: (WG_BLOCKS + 128) x WG_PART
    std::vector<void*> allocs;
};

namespace sf {

static int grid1(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

template <class T>
static int cnn_alloc(sf_cnn* c, T** p, size_t n) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return sf_fail(c->ctx, SF_ERR_OOM, "cnn allocation of %zu bytes failed", n * sizeof(T));
    hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T));
    c->allocs.push_back((void*)*p);
    return SF_OK;
}

static int conv_init(sf_cnn* c, ConvL& L, int Kreal, int N, std::mt19937_64& rng) {
    L.Kreal = Kreal;
    L.K = (Kreal + 31) / 32 * 32;
    L.N = N;
    size_t nw = L.nw();
    int rc = 0;
    rc |= cnn_alloc(c, &L.wm, nw);
    rc |= cnn_alloc(c, &L.gw, nw);
    rc |= cnn_alloc(c, &L.mw, nw);
    rc |= cnn_alloc(c, &L.vw, nw);
    rc |= cnn_alloc(c, &L.b, N);
    rc |= cnn_alloc(c, &L.gb, N);
    rc |= cnn_alloc(c, &L.mb, N);
    rc |= cnn_alloc(c, &L.vb, N);
    rc |= cnn_alloc(c, &L.wf, nw);
    rc |= cnn_alloc(c, &L.wt, nw);
    if (rc) return SF_ERR_OOM;
    // He initialisation over the real fan-in; padded input channels stay zero
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * Kreal)));
    std::vector<float> h(nw, 0.f);
    for (int tap = 0; tap < 9; tap++)
        for (int n = 0; n < N; n++)
            for (int k = 0; k < Kreal; k++) h[((size_t)tap * N + n) * L.K + k] = (float)nd(rng);
    if (hipMemcpy(L.wm, h.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess) return SF_ERR_HIP;
    return SF_OK;
}

static void conv_pack(sf_cnn* c, ConvL& L) {
    hipLaunchKernelGGL(pack_weights, dim3(grid1(L.nw())), dim3(256), 0, c->ctx->stream, (const float*)L.wm, L.N, L.K,
                       L.wf, L.wt);
}

// 3x3 convolution launch (conv3x3_lds): 64-channel workgroup tiles when N allows, else 32
template <int MODE>
static void conv_launch(sf_cnn* c, const uint16_t* in, int in_cs, int in_co, int K, const uint16_t* w,
                        const float* bias, int N, uint16_t* out, float* fout, int out_cs, int out_co, int H, int W,
                        int B, int relu, int sgn, int accumulate, const uint16_t* act = nullptr, int act_cs = 0,
                        int act_co = 0) {
    // rows per workgroup: 2 with 64-channel tiles, 4 with 32 (the same accumulator count)
    const bool t64 = N % 64 == 0;
    const int tn = t64 ? 64 : 32, rows = t64 ? 2 : 4;
    // row tiles per workgroup: up to 4 while the launch keeps at least 1024 workgroups
    const int64_t tiles = (int64_t)((W + CONV_TP - 1) / CONV_TP) * ((H + rows - 1) / rows) * B * ((N + tn - 1) / tn);
    const int nrt = (int)std::max<int64_t>(1, std::min<int64_t>(4, tiles / 1024));
    const dim3 g((W + CONV_TP - 1) / CONV_TP, (H + rows * nrt - 1) / (rows * nrt), B * ((N + tn - 1) / tn));
    if (t64)
        hipLaunchKernelGGL((conv3x3_lds<MODE, 64, 2>), g, dim3(256), 0, c->ctx->stream, in, in_cs, in_co, K, w, bias,
                           N, out, fout, out_cs, out_co, H, W, relu, sgn, accumulate, act, act_cs, act_co, nrt);
    else
        hipLaunchKernelGGL((conv3x3_lds<MODE, 32, 4>), g, dim3(256), 0, c->ctx->stream, in, in_cs, in_co, K, w, bias,
                           N, out, fout, out_cs, out_co, H, W, relu, sgn, accumulate, act, act_cs, act_co, nrt);
}
// forward conv: in view (cs, co, K channels) -> out view, bias + ReLU
static void conv_fwd(sf_cnn* c, const ConvL& L, const uint16_t* in, int in_cs, int in_co, uint16_t* out, int out_cs,
                     int out_co, int H, int W, int B) {
    conv_launch<0>(c, in, in_cs, in_co, L.K, (const uint16_t*)L.wf, (const float*)L.b, L.N, out, (float*)nullptr,
                   out_cs, out_co, H, W, B, 1, 1, 0);
}
// backward of one conv: g (BF16 [npix][N]) -> weight / bias gradients and, if din != null, the FP32
// input gradient into the din view; if gout != null, the input gradient through the ReLU of the
// layer that produced the input (activation view act) as BF16 [npix][K] into gout
static void conv_bwd(sf_cnn* c, const ConvL& L, const uint16_t* in, int in_cs, int in_co, const uint16_t* g,
                     float* din, int din_cs, int din_co, int H, int W, int B, uint16_t* gout = nullptr,
                     const uint16_t* act = nullptr, int act_cs = 0, int act_co = 0) {
    // weight gradient: about WG_BLOCKS workgroups over the (32 x 32 channel tile, pixel band)
    // pairs, their partials summed in workgroup order
    const int nchunks = B * H * ((W + WG_PX - 1) / WG_PX);
    const int pairs = (L.K / 32) * (L.N / 32);
    int bx = std::max(1, std::min(nchunks, (WG_BLOCKS + pairs - 1) / pairs));
    const int cpb = (nchunks + bx - 1) / bx;
    bx = (nchunks + cpb - 1) / cpb;
    hipLaunchKernelGGL(conv3x3_wgrad_lds, dim3(bx, pairs), dim3(256), 0, c->ctx->stream, in, in_cs, in_co, L.K, g, L.N,
                       L.N, c->wpart, H, W, B, cpb);
    hipLaunchKernelGGL(wgrad_reduce, dim3((WG_PART + 31) / 32, pairs), dim3(256), 0, c->ctx->stream,
                       (const float*)c->wpart, bx, L.K, L.N, L.gw, L.gb);
    if (din)
        conv_launch<1>(c, g, L.N, 0, L.N, (const uint16_t*)L.wt, (const float*)nullptr, L.K, (uint16_t*)nullptr, din,
                       din_cs, din_co, H, W, B, 0, -1, 0);
    if (gout)
        conv_launch<2>(c, g, L.N, 0, L.N, (const uint16_t*)L.wt, (const float*)nullptr, L.K, gout, (float*)nullptr,
                       L.K, 0, H, W, B, 0, -1, 0, act, act_cs, act_co);
}

static void forward(sf_cnn* c, int B) {
    const int H = c->H, W = c->W, C1 = c->C1, C2 = c->C2, C3 = c->C3;
    const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    hipStream_t st = c->ctx->stream;
    for (ConvL* L : {&c->c1a, &c->c1b, &c->c2a, &c->c2b, &c->c3a, &c->c3b, &c->c4a, &c->c4b, &c->c5a, &c->c5b})
        conv_pack(c, *L);
    conv_fwd(c, c->c1a, c->x0, 32, 0, c->e1a, C1, 0, H, W, B);
    conv_fwd(c, c->c1b, c->e1a, C1, 0, c->u1, C2 + C1, C2, H, W, B);                  // e1 -> u1[C2:]
    hipLaunchKernelGGL(avgpool_fwd, dim3(grid1((int64_t)B * H2 * W2 * C1 / 8)), dim3(256), 0, st, (const uint16_t*)c->u1,
                       C2 + C1, C2, C1, H, W, B, c->p1, C1, 0);
    conv_fwd(c, c->c2a, c->p1, C1, 0, c->e2a, C2, 0, H2, W2, B);
    conv_fwd(c, c->c2b, c->e2a, C2, 0, c->u2, C3 + C2, C3, H2, W2, B);                // e2 -> u2[C3:]
    hipLaunchKernelGGL(avgpool_fwd, dim3(grid1((int64_t)B * H4 * W4 * C2 / 8)), dim3(256), 0, st, (const uint16_t*)c->u2,
                       C3 + C2, C3, C2, H2, W2, B, c->p2, C2, 0);
    conv_fwd(c, c->c3a, c->p2, C2, 0, c->e3a, C3, 0, H4, W4, B);
    conv_fwd(c, c->c3b, c->e3a, C3, 0, c->e3, C3, 0, H4, W4, B);
    hipLaunchKernelGGL(upsample_fwd, dim3(grid1((int64_t)B * H2 * W2 * C3 / 8)), dim3(256), 0, st, (const uint16_t*)c->e3,
                       C3, C3, H4, W4, B, c->u2, C3 + C2, 0);
    conv_fwd(c, c->c4a, c->u2, C3 + C2, 0, c->d2a, C2, 0, H2, W2, B);
    conv_fwd(c, c->c4b, c->d2a, C2, 0, c->d2, C2, 0, H2, W2, B);
    hipLaunchKernelGGL(upsample_fwd, dim3(grid1((int64_t)B * H * W * C2 / 8)), dim3(256), 0, st, (const uint16_t*)c->d2, C2,
                       C2, H2, W2, B, c->u1, C2 + C1, 0);
    conv_fwd(c, c->c5a, c->u1, C2 + C1, 0, c->d1a, C1, 0, H, W, B);
    conv_fwd(c, c->c5b, c->d1a, C1, 0, c->d1, C1, 0, H, W, B);
    hipLaunchKernelGGL(head_fwd, dim3(grid1((int64_t)B * H * W)), dim3(256), 0, st, (const uint16_t*)c->d1,
                       (int64_t)B * H * W, (const float*)c->hw, (const float*)c->hb, (const uint16_t*)c->x0, 32, 3, c->y);
}

// The BF16 operand gradients alternate between gop and gop2 (a layer's backward reads one and
// writes the next into the other); every ReLU backward is fused into the kernel that produces the
// gradient of the activation.  FP32 gradients remain only where two paths add: the inputs of the
// decoder convolutions (upsampled part plus skip part) and the pooled activations.
static void backward(sf_cnn* c, int B) {
    const int H = c->H, W = c->W, C1 = c->C1, C2 = c->C2, C3 = c->C3;
    const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
    const int64_t n1 = (int64_t)B * H * W, n2 = (int64_t)B * H2 * W2, n4 = (int64_t)B * H4 * W4;
    hipStream_t st = c->ctx->stream;
    uint16_t* ga = c->gop;
    uint16_t* gb = c->gop2;
    hipLaunchKernelGGL(head_bwd, dim3(1024), dim3(256), 0, st, (const uint16_t*)c->d1, n1, (const float*)c->hw,
                       (const float*)c->dy, c->ghw, c->ghb, ga);
    conv_bwd(c, c->c5b, c->d1a, C1, 0, ga, nullptr, 0, 0, H, W, B, gb, c->d1a, C1, 0);
    conv_bwd(c, c->c5a, c->u1, C2 + C1, 0, gb, c->g_u1, C2 + C1, 0, H, W, B);
    hipLaunchKernelGGL(upsample_relu_bwd, dim3(grid1(n2 * C2 / 4)), dim3(256), 0, st, (const float*)c->g_u1, C2 + C1,
                       C2, H2, W2, B, (const uint16_t*)c->d2, ga);
    conv_bwd(c, c->c4b, c->d2a, C2, 0, ga, nullptr, 0, 0, H2, W2, B, gb, c->d2a, C2, 0);
    conv_bwd(c, c->c4a, c->u2, C3 + C2, 0, gb, c->g_u2, C3 + C2, 0, H2, W2, B);
    hipLaunchKernelGGL(upsample_relu_bwd, dim3(grid1(n4 * C3 / 4)), dim3(256), 0, st, (const float*)c->g_u2, C3 + C2,
                       C3, H4, W4, B, (const uint16_t*)c->e3, ga);
    conv_bwd(c, c->c3b, c->e3a, C3, 0, ga, nullptr, 0, 0, H4, W4, B, gb, c->e3a, C3, 0);
    conv_bwd(c, c->c3a, c->p2, C2, 0, gb, c->g_p2, C2, 0, H4, W4, B);
    hipLaunchKernelGGL(avgpool_relu_bwd, dim3(grid1(n2 * C2 / 4)), dim3(256), 0, st, (const float*)c->g_p2, C2, H2, W2,
                       B, (const float*)c->g_u2, C3 + C2, C3, (const uint16_t*)c->u2, C3 + C2, C3, ga);
    conv_bwd(c, c->c2b, c->e2a, C2, 0, ga, nullptr, 0, 0, H2, W2, B, gb, c->e2a, C2, 0);
    conv_bwd(c, c->c2a, c->p1, C1, 0, gb, c->g_p1, C1, 0, H2, W2, B);
    hipLaunchKernelGGL(avgpool_relu_bwd, dim3(grid1(n1 * C1 / 4)), dim3(256), 0, st, (const float*)c->g_p1, C1, H, W, B,
                       (const float*)c->g_u1, C2 + C1, C2, (const uint16_t*)c->u1, C2 + C1, C2, ga);
    conv_bwd(c, c->c1b, c->e1a, C1, 0, ga, nullptr, 0, 0, H, W, B, gb, c->e1a, C1, 0);
    conv_bwd(c, c->c1a, c->x0, 32, 0, gb, nullptr, 0, 0, H, W, B);
}

static std::vector<std::pair<float*, size_t>> grads(sf_cnn* c) {
    std::vector<std::pair<float*, size_t>> g;
    for (ConvL* L : {&c->c1a, &c->c1b, &c->c2a, &c->c2b, &c->c3a, &c->c3b, &c->c4a, &c->c4b, &c->c5a, &c->c5b}) {
        g.push_back({L->gw, L->nw()});
        g.push_back({L->gb, (size_t)L->N});
    }
    g.push_back({c->ghw, (size_t)c->C1});
    g.push_back({c->ghb, 1});
    return g;
}
static std::vector<std::pair<float*, size_t>> params(sf_cnn* c) {
    std::vector<std::pair<float*, size_t>> g;
    for (ConvL* L : {&c->c1a, &c->c1b, &c->c2a, &c->c2b, &c->c3a, &c->c3b, &c->c4a, &c->c4b, &c->c5a, &c->c5b}) {
        g.push_back({L->wm, L->nw()});
        g.push_back({L->b, (size_t)L->N});
    }
    g.push_back({c->hw, (size_t)c->C1});
    g.push_back({c->hb, 1});
    return g;
}

// place sample s into batch slot b of x0 / target / mask
__global__ void assemble_batch(const uint16_t* __restrict__ sin, const float* __restrict__ st,
                               const uint8_t* __restrict__ sm, int64_t npix, uint16_t* __restrict__ x0,
                               float* __restrict__ bt, uint8_t* __restrict__ bm) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 4; c++) x0[p * 32 + c] = sin[p * 4 + c];
        for (int c = 4; c < 32; c++) x0[p * 32 + c] = 0;
        bt[p] = st[p];
        bm[p] = sm[p];
    }
}
// Snapshot of a context on the padded grid: input channels (rho / rho_ref, Dirichlet mask,
// boundary phi / phi_ref, Boltzmann inversion (phi0 + kTe ln max(rho / (e n0), 1)) / phi_ref),
// target phi / phi_ref and the free-node mask.  Pixels outside the mesh are zero.
__global__ void snapshot_kernel(int ni, int nj, int H, int W, const double* __restrict__ rho,
                                const double* __restrict__ phi, const int32_t* __restrict__ src,
                                const uint8_t* __restrict__ cls, const double* __restrict__ fixval, double inv_rho,
                                double inv_phi, double inv_en0, double kTe, double phi0,
                                uint16_t* __restrict__ sin, float* __restrict__ st, uint8_t* __restrict__ sm) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < (int64_t)H * W; p += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(p % W), y = (int)(p / W);
        uint16_t r = 0, mk = 0, pb = 0, bz = 0;
        float t = 0;
        uint8_t m = 0;
        if (x < ni && y < nj) {
            size_t u = (size_t)y * ni + x;
            bool fixed = src[u] < 0;
            r = f2bf((float)(rho[u] * inv_rho));
            mk = f2bf(fixed ? 1.f : 0.f);
            pb = f2bf(fixed ? (float)(fixval[u] * inv_phi) : 0.f);
            if (kTe > 0) bz = f2bf((float)((phi0 + kTe * log(fmax(rho[u] * inv_en0, 1.0))) * inv_phi));
            if (phi) t = (float)(phi[u] * inv_phi);
            m = cls[u] != 0xFF ? 1 : 0;
        }
        sin[p * 4 + 0] = r;
        sin[p * 4 + 1] = mk;
        sin[p * 4 + 2] = pb;
        sin[p * 4 + 3] = bz;
        if (st) st[p] = t;
        if (sm) sm[p] = m;
    }
}
// network output -> phi on the free nodes of the context (fixed / face nodes set by the solver)
__global__ void guess_to_phi(int ni, int nj, int W, const float* __restrict__ y, double phi_ref,
                             double* __restrict__ phi) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < (int64_t)ni * nj; u += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(u % ni), yy = (int)(u / ni);
        phi[u] = (double)y[(size_t)yy * W + x] * phi_ref;
    }
}

}  // namespace sf

// accessors used by the solver (defined in sf_poisson.hip)
namespace sf {
const int32_t* poisson_src(sf_ctx* ctx);
const uint8_t* poisson_cls(sf_ctx* ctx);
const double* poisson_fixval(sf_ctx* ctx);
const sf_poisson_desc* poisson_desc(sf_ctx* ctx);
int poisson_rel_residual(sf_ctx* ctx, double* rel);

// parameters of the Boltzmann-inversion input channel (zero channel for a linear solver)
static double feat_kTe(sf_ctx* t) { return poisson_desc(t)->linear ? 0.0 : poisson_desc(t)->Te0; }
static double feat_inv_en0(sf_ctx* t) {
    double n0 = poisson_desc(t)->n0;
    return n0 > 0 ? 1.0 / (QE * n0) : 0.0;
}
}

using namespace sf;

extern "C" {

int sf_cnn_create(sf_ctx* ctx, int ni, int nj, int max_batch, uint64_t seed, sf_cnn** out) {
    SF_CHECK_CTX(ctx);
    if (!out || max_batch < 1 || max_batch > 64) return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_create: bad argument");
    if (ni <= 0 || nj <= 0) {
        if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_cnn_create: no size given and no mesh");
        ni = ctx->mp.ni;
        nj = ctx->mp.nj;
    }
    if (ni > 16384 || nj > 16384) return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_create: grid too large");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    sf_cnn* c = new sf_cnn();
    c->ctx = ctx;
    c->ni = ni;
    c->nj = nj;
    c->W = (c->ni + 3) / 4 * 4;
    c->H = (c->nj + 3) / 4 * 4;
    c->Bmax = max_batch;
    c->rng.seed(seed);
    const int C1 = c->C1, C2 = c->C2, C3 = c->C3;
    const int64_t n1 = (int64_t)max_batch * c->H * c->W, n2 = n1 / 4, n4 = n1 / 16;
    int rc = 0;
    rc |= conv_init(c, c->c1a, 4, C1, c->rng);
    rc |= conv_init(c, c->c1b, C1, C1, c->rng);
    rc |= conv_init(c, c->c2a, C1, C2, c->rng);
    rc |= conv_init(c, c->c2b, C2, C2, c->rng);
    rc |= conv_init(c, c->c3a, C2, C3, c->rng);
    rc |= conv_init(c, c->c3b, C3, C3, c->rng);
    rc |= conv_init(c, c->c4a, C3 + C2, C2, c->rng);
    rc |= conv_init(c, c->c4b, C2, C2, c->rng);
    rc |= conv_init(c, c->c5a, C2 + C1, C1, c->rng);
    rc |= conv_init(c, c->c5b, C1, C1, c->rng);
    // (a launch has at most ceil(WG_BLOCKS / pairs) * pairs <= WG_BLOCKS + pairs - 1 workgroups;
    // the network's layers have at most 16 channel tile pairs)
    rc |= cnn_alloc(c, &c->wpart, (size_t)(WG_BLOCKS + 128) * WG_PART);
    rc |= cnn_alloc(c, &c->hw, C1);
    rc |= cnn_alloc(c, &c->hb, 1);
    rc |= cnn_alloc(c, &c->ghw, C1);
    rc |= cnn_alloc(c, &c->ghb, 1);
    rc |= cnn_alloc(c, &c->mhw, C1);
    rc |= cnn_alloc(c, &c->vhw, C1);
    rc |= cnn_alloc(c, &c->mhb, 1);
    rc |= cnn_alloc(c, &c->vhb, 1);
    rc |= cnn_alloc(c, &c->segs, ADAM_MAXSEG);
    rc |= cnn_alloc(c, &c->x0, n1 * 32);
    rc |= cnn_alloc(c, &c->e1a, n1 * C1);
    rc |= cnn_alloc(c, &c->u1, n1 * (C2 + C1));
    rc |= cnn_alloc(c, &c->p1, n2 * C1);
    rc |= cnn_alloc(c, &c->e2a, n2 * C2);
    rc |= cnn_alloc(c, &c->u2, n2 * (C3 + C2));
    rc |= cnn_alloc(c, &c->p2, n4 * C2);
    rc |= cnn_alloc(c, &c->e3a, n4 * C3);
    rc |= cnn_alloc(c, &c->e3, n4 * C3);
    rc |= cnn_alloc(c, &c->d2a, n2 * C2);
    rc |= cnn_alloc(c, &c->d2, n2 * C2);
    rc |= cnn_alloc(c, &c->d1a, n1 * C1);
    rc |= cnn_alloc(c, &c->d1, n1 * C1);
    rc |= cnn_alloc(c, &c->gop, n1 * (C2 + C1));
    rc |= cnn_alloc(c, &c->gop2, n1 * C1);
    rc |= cnn_alloc(c, &c->y, n1);
    rc |= cnn_alloc(c, &c->g_u1, n1 * (C2 + C1));
    rc |= cnn_alloc(c, &c->g_p1, n2 * C1);
    rc |= cnn_alloc(c, &c->g_u2, n2 * (C3 + C2));
    rc |= cnn_alloc(c, &c->g_p2, n4 * C2);
    rc |= cnn_alloc(c, &c->dy, n1);
    rc |= cnn_alloc(c, &c->loss, 4);
    rc |= cnn_alloc(c, &c->bt, n1);
    rc |= cnn_alloc(c, &c->bmask, n1);
    rc |= cnn_alloc(c, &c->phi_save, (size_t)ni * nj);
    if (rc) {
        for (void* p : c->allocs) hipFree(p);
        delete c;
        return SF_ERR_OOM;
    }
    std::normal_distribution<double> nd(0.0, std::sqrt(1.0 / C1));
    std::vector<float> h(C1);
    for (auto& v : h) v = (float)nd(c->rng);
    SF_HIP(ctx, hipMemcpy(c->hw, h.data(), C1 * 4, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipDeviceSynchronize());
    *out = c;
    return SF_OK;
}

int sf_cnn_destroy(sf_cnn* c) {
    if (!c) return SF_ERR_ARG;
    hipSetDevice(c->ctx->device);
    hipStreamSynchronize(c->ctx->stream);
    for (void* p : c->allocs) hipFree(p);
    for (auto* p : c->s_in) hipFree(p);
    for (auto* p : c->s_t) hipFree(p);
    for (auto* p : c->s_m) hipFree(p);
    delete c;
    return SF_OK;
}

int sf_cnn_num_samples(const sf_cnn* c) { return c ? (int)c->s_in.size() : SF_ERR_ARG; }

int sf_cnn_dims(const sf_cnn* c, int* H, int* W, int64_t* nparams) {
    if (!c) return SF_ERR_ARG;
    if (H) *H = c->H;
    if (W) *W = c->W;
    if (nparams) {
        int64_t n = 0;
        for (auto& p : params(const_cast<sf_cnn*>(c))) n += (int64_t)p.second;
        *nparams = n;
    }
    return SF_OK;
}

// Adds a training sample given on the host: x4 [H][W][4] FP32 input channels (stored as BF16),
// t [H][W] FP32 target, m [H][W] loss mask.
int sf_cnn_add_sample(sf_cnn* c, const float* x4, const float* t, const uint8_t* m) {
    if (!c || !x4 || !t || !m) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t np = (int64_t)c->H * c->W;
    std::vector<uint16_t> h(np * 4);
    for (int64_t i = 0; i < np * 4; i++) h[i] = f2bf(x4[i]);
    uint16_t* sin = nullptr;
    float* st = nullptr;
    uint8_t* sm = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&sin, np * 8));
    SF_HIP(ctx, hipMalloc((void**)&st, np * 4));
    SF_HIP(ctx, hipMalloc((void**)&sm, np));
    SF_HIP(ctx, hipMemcpy(sin, h.data(), np * 8, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(st, t, np * 4, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(sm, m, np, hipMemcpyHostToDevice));
    c->s_in.push_back(sin);
    c->s_t.push_back(st);
    c->s_m.push_back(sm);
    return SF_OK;
}

// Reads back sample idx (input channels as stored, i.e. BF16-rounded, widened to FP32).
int sf_cnn_get_sample(sf_cnn* c, int idx, float* x4, float* t, uint8_t* m) {
    if (!c || idx < 0 || idx >= (int)c->s_in.size()) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t np = (int64_t)c->H * c->W;
    if (x4) {
        std::vector<uint16_t> h(np * 4);
        SF_HIP(ctx, hipMemcpy(h.data(), c->s_in[idx], np * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < np * 4; i++) x4[i] = bf2f(h[i]);
    }
    if (t) SF_HIP(ctx, hipMemcpy(t, c->s_t[idx], np * 4, hipMemcpyDeviceToHost));
    if (m) SF_HIP(ctx, hipMemcpy(m, c->s_m[idx], np, hipMemcpyDeviceToHost));
    return SF_OK;
}

// Adds the current (rho, Dirichlet mask / values, phi) of `src_ctx` to the training set.
int sf_cnn_snapshot(sf_cnn* c, sf_ctx* src_ctx) {
    if (!c || !src_ctx) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (src_ctx->mp.ni != c->ni || src_ctx->mp.nj != c->nj || !src_ctx->poisson)
        return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_snapshot: context has another mesh size or no solver");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(src_ctx->stream));
    const int64_t np = (int64_t)c->H * c->W;
    uint16_t* sin = nullptr;
    float* st = nullptr;
    uint8_t* sm = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&sin, np * 8));
    SF_HIP(ctx, hipMalloc((void**)&st, np * 4));
    SF_HIP(ctx, hipMalloc((void**)&sm, np));
    hipLaunchKernelGGL(snapshot_kernel, dim3(grid1(np)), dim3(256), 0, ctx->stream, c->ni, c->nj, c->H, c->W,
                       (const double*)src_ctx->d_rho, (const double*)src_ctx->d_phi, poisson_src(src_ctx),
                       poisson_cls(src_ctx), poisson_fixval(src_ctx), 1.0 / c->rho_ref, 1.0 / c->phi_ref,
                       feat_inv_en0(src_ctx), feat_kTe(src_ctx), poisson_desc(src_ctx)->phi0, sin, st, sm);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    c->s_in.push_back(sin);
    c->s_t.push_back(st);
    c->s_m.push_back(sm);
    return SF_OK;
}

// Adam training on random mini-batches of the snapshots; returns the mean loss of the last
// `steps/10` steps in *loss_out (normalised MSE of phi / 100 V on free nodes).
int sf_cnn_train(sf_cnn* c, int steps, int batch, double lr, double* loss_out) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (c->s_in.empty()) return sf_fail(ctx, SF_ERR_STATE, "sf_cnn_train: no snapshots");
    if (batch < 1 || batch > c->Bmax || steps < 0) return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_train: bad batch/steps");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t np = (int64_t)c->H * c->W;
    std::uniform_int_distribution<int> pick(0, (int)c->s_in.size() - 1);
    double tail = 0;
    int ntail = 0;
    auto P = params(c);
    auto G = grads(c);
    std::vector<std::pair<float*, float*>> mv;
    for (ConvL* L : {&c->c1a, &c->c1b, &c->c2a, &c->c2b, &c->c3a, &c->c3b, &c->c4a, &c->c4b, &c->c5a, &c->c5b}) {
        mv.push_back({L->mw, L->vw});
        mv.push_back({L->mb, L->vb});
    }
    mv.push_back({c->mhw, c->vhw});
    mv.push_back({c->mhb, c->vhb});
    std::vector<AdamSeg> segs;
    int64_t total = 0;
    for (size_t q = 0; q < P.size(); q++) {
        segs.push_back(AdamSeg{P[q].first, G[q].first, mv[q].first, mv[q].second, total, (int64_t)P[q].second});
        total += (int64_t)P[q].second;
    }
    if (segs.size() > (size_t)ADAM_MAXSEG) return sf_fail(ctx, SF_ERR_STATE, "sf_cnn_train: too many tensors");
    SF_HIP(ctx, hipMemcpy(c->segs, segs.data(), segs.size() * sizeof(AdamSeg), hipMemcpyHostToDevice));
    const int nseg = (int)segs.size();
    for (int s = 0; s < steps; s++) {
        int64_t counted = 0;
        for (int b = 0; b < batch; b++) {
            int k = pick(c->rng);
            hipLaunchKernelGGL(assemble_batch, dim3(grid1(np)), dim3(256), 0, st, (const uint16_t*)c->s_in[k],
                               (const float*)c->s_t[k], (const uint8_t*)c->s_m[k], np, c->x0 + b * np * 32,
                               c->bt + b * np, c->bmask + b * np);
            counted += np;
        }
        hipLaunchKernelGGL(zero_grads, dim3(grid1(total)), dim3(256), 0, st, (const AdamSeg*)c->segs, nseg, total);
        SF_HIP(ctx, hipMemsetAsync(c->loss, 0, 4, st));
        forward(c, batch);
        hipLaunchKernelGGL(mse_loss, dim3(std::min(grid1((int64_t)batch * np), 256)), dim3(256), 0, st, (const float*)c->y,
                           (const float*)c->bt, (const uint8_t*)c->bmask, (int64_t)batch * np, 1.0f / (float)counted,
                           c->dy, c->loss);
        backward(c, batch);
        c->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)c->adam_t), c2 = 1.f - powf(b2, (float)c->adam_t);
        hipLaunchKernelGGL(adam_segs, dim3(grid1(total)), dim3(256), 0, st, (const AdamSeg*)c->segs, nseg, total,
                           (float)lr, b1, b2, 1e-8f, c1, c2);
        SF_HIP(ctx, hipGetLastError());
        if (s >= steps - std::max(1, steps / 10)) {
            float l = 0;
            SF_HIP(ctx, hipMemcpyAsync(&l, c->loss, 4, hipMemcpyDeviceToHost, st));
            SF_HIP(ctx, hipStreamSynchronize(st));
            tail += l;
            ntail++;
        }
    }
    SF_HIP(ctx, hipStreamSynchronize(st));
    if (loss_out) *loss_out = ntail ? tail / ntail : 0.0;
    return SF_OK;
}

// Loss and all parameter gradients of one sample (for gradient checks).  grads_out receives the
// gradients in the order of sf_cnn_params; n_out the total count.
int sf_cnn_loss_and_grads(sf_cnn* c, int sample, double* loss_out, float* y_out, float* grads_out,
                          int64_t* n_out) {
    if (!c || sample < 0 || sample >= (int)c->s_in.size()) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t np = (int64_t)c->H * c->W;
    hipLaunchKernelGGL(assemble_batch, dim3(grid1(np)), dim3(256), 0, st, (const uint16_t*)c->s_in[sample],
                       (const float*)c->s_t[sample], (const uint8_t*)c->s_m[sample], np, c->x0, c->bt, c->bmask);
    auto G = grads(c);
    for (auto& g : G) SF_HIP(ctx, hipMemsetAsync(g.first, 0, g.second * 4, st));
    SF_HIP(ctx, hipMemsetAsync(c->loss, 0, 4, st));
    forward(c, 1);
    hipLaunchKernelGGL(mse_loss, dim3(std::min(grid1(np), 256)), dim3(256), 0, st, (const float*)c->y, (const float*)c->bt,
                       (const uint8_t*)c->bmask, np, 1.0f / (float)np, c->dy, c->loss);
    backward(c, 1);
    SF_HIP(ctx, hipGetLastError());
    float l = 0;
    SF_HIP(ctx, hipMemcpyAsync(&l, c->loss, 4, hipMemcpyDeviceToHost, st));
    SF_HIP(ctx, hipStreamSynchronize(st));
    if (loss_out) *loss_out = l;
    if (y_out) SF_HIP(ctx, hipMemcpy(y_out, c->y, np * 4, hipMemcpyDeviceToHost));
    int64_t off = 0;
    for (auto& g : G) {
        if (grads_out) SF_HIP(ctx, hipMemcpy(grads_out + off, g.first, g.second * 4, hipMemcpyDeviceToHost));
        off += (int64_t)g.second;
    }
    if (n_out) *n_out = off;
    return SF_OK;
}

// Diagnostic read-back of an activation of batch slot 0 after the last forward pass, widened to
// FP32: which = 0 x0 [H][W][32], 1 e1a, 2 u1 [.][96], 3 p1, 4 e2a, 5 u2 [.][192], 6 p2, 7 e3a,
// 8 e3, 9 d2a, 10 d2, 11 d1a, 12 d1, 13 y [H][W].  *count receives the number of values.
int sf_cnn_activation(sf_cnn* c, int which, float* out, int64_t* count) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    const int64_t n1 = (int64_t)c->H * c->W, n2 = n1 / 4, n4 = n1 / 16;
    const int C1 = c->C1, C2 = c->C2, C3 = c->C3;
    const uint16_t* src = nullptr;
    int64_t n = 0;
    switch (which) {
        case 0: src = c->x0; n = n1 * 32; break;
        case 1: src = c->e1a; n = n1 * C1; break;
        case 2: src = c->u1; n = n1 * (C2 + C1); break;
        case 3: src = c->p1; n = n2 * C1; break;
        case 4: src = c->e2a; n = n2 * C2; break;
        case 5: src = c->u2; n = n2 * (C3 + C2); break;
        case 6: src = c->p2; n = n4 * C2; break;
        case 7: src = c->e3a; n = n4 * C3; break;
        case 8: src = c->e3; n = n4 * C3; break;
        case 9: src = c->d2a; n = n2 * C2; break;
        case 10: src = c->d2; n = n2 * C2; break;
        case 11: src = c->d1a; n = n1 * C1; break;
        case 12: src = c->d1; n = n1 * C1; break;
        case 13: n = n1; break;
        default: return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_activation: unknown tensor %d", which);
    }
    if (count) *count = n;
    if (!out) return SF_OK;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (which == 13) {
        SF_HIP(ctx, hipMemcpy(out, c->y, n * 4, hipMemcpyDeviceToHost));
        return SF_OK;
    }
    std::vector<uint16_t> h(n);
    SF_HIP(ctx, hipMemcpy(h.data(), src, n * 2, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) out[i] = bf2f(h[i]);
    return SF_OK;
}

// parameters (FP32 master copies) in a fixed order: for each of the 10 convolutions
// W[9][N][K] then b[N]; then the head w[C1], b[1]. get = 1 downloads, get = 0 uploads.
int sf_cnn_params(sf_cnn* c, int get, float* buf, int64_t* n_out) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t off = 0;
    for (auto& p : params(c)) {
        if (buf) {
            if (get) SF_HIP(ctx, hipMemcpy(buf + off, p.first, p.second * 4, hipMemcpyDeviceToHost));
            else SF_HIP(ctx, hipMemcpy(p.first, buf + off, p.second * 4, hipMemcpyHostToDevice));
        }
        off += (int64_t)p.second;
    }
    if (n_out) *n_out = off;
    return SF_OK;
}

int sf_cnn_predict(sf_cnn* c, sf_ctx* target);

// Potential solve started from the network's guess.  policy 0: always start from the guess;
// policy 1: start from whichever of the current potential and the guess has the smaller
// nonlinear residual (the guess is discarded when it does not help).  info->newton_iterations
// etc. describe the solve; *used_guess (may be NULL) tells which start was taken.
int sf_poisson_solve_learned(sf_ctx* ctx, sf_cnn* c, int policy, sf_solve_info* info, int* used_guess) {
    SF_CHECK_CTX(ctx);
    if (!c || (policy != 0 && policy != 1)) return sf_fail(ctx, SF_ERR_ARG, "sf_poisson_solve_learned: bad argument");
    if (!ctx->poisson) return sf_fail(ctx, SF_ERR_STATE, "sf_poisson_solve_learned: solver not set up");
    if (ctx->mp.ni != c->ni || ctx->mp.nj != c->nj)
        return sf_fail(ctx, SF_ERR_ARG, "sf_poisson_solve_learned: network grid differs from the mesh");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)c->ni * c->nj;
    double r_prev = 0, r_guess = 0;
    if (policy == 1) {
        int rc = poisson_rel_residual(ctx, &r_prev);
        if (rc < 0) return rc;
    }
    SF_HIP(ctx, hipMemcpyAsync(c->phi_save, ctx->d_phi, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
    int rc = sf_cnn_predict(c, ctx);
    if (rc < 0) return sf_fail(ctx, rc, "sf_poisson_solve_learned: prediction failed: %s", c->ctx->err.c_str());
    // The guess is only a starting point: one whose nonlinear residual is not finite (exp of a
    // wild potential), or worse than the previous potential's under policy 1, is replaced by the
    // previous potential, and a solve that diverges from the guess is repeated from it.
    rc = poisson_rel_residual(ctx, &r_guess);
    if (rc < 0) return rc;
    bool use = std::isfinite(r_guess) && (policy == 0 || r_guess < r_prev);
    auto restore = [&]() -> int {
        SF_HIP(ctx, hipMemcpyAsync(ctx->d_phi, c->phi_save, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
        return SF_OK;
    };
    if (!use && (rc = restore()) < 0) return rc;
    rc = poisson_solve(ctx, info);
    if (rc < 0 && use) {
        use = false;
        if ((rc = restore()) < 0) return rc;
        rc = poisson_solve(ctx, info);
    }
    if (used_guess) *used_guess = use ? 1 : 0;
    return rc;
}

// Network output for the current state of `target` (rho, Dirichlet values) written as phi.
int sf_cnn_predict(sf_cnn* c, sf_ctx* target) {
    if (!c || !target) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (target->mp.ni != c->ni || target->mp.nj != c->nj || !target->poisson)
        return sf_fail(ctx, SF_ERR_ARG, "sf_cnn_predict: context has another mesh size or no solver");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(target->stream));
    const int64_t np = (int64_t)c->H * c->W;
    uint16_t* sin = c->gop;   // scratch [np][4]
    hipLaunchKernelGGL(snapshot_kernel, dim3(grid1(np)), dim3(256), 0, ctx->stream, c->ni, c->nj, c->H, c->W,
                       (const double*)target->d_rho, (const double*)nullptr, poisson_src(target), poisson_cls(target),
                       poisson_fixval(target), 1.0 / c->rho_ref, 1.0 / c->phi_ref, feat_inv_en0(target),
                       feat_kTe(target), poisson_desc(target)->phi0, sin, (float*)nullptr, (uint8_t*)nullptr);
    hipLaunchKernelGGL(assemble_batch, dim3(grid1(np)), dim3(256), 0, ctx->stream, (const uint16_t*)sin,
                       (const float*)c->bt, (const uint8_t*)c->bmask, np, c->x0, c->bt, c->bmask);
    forward(c, 1);
    hipLaunchKernelGGL(guess_to_phi, dim3(grid1((int64_t)c->ni * c->nj)), dim3(256), 0, ctx->stream, c->ni, c->nj,
                       c->W, (const float*)c->y, c->phi_ref, target->d_phi);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SF_OK;
}

// MFMA lane-layout probe: D = A(16x32) * B(32x16) with the operand layout the convolution
// kernels assume; host compares against a plain product.
__global__ void mfma_probe_kernel(const uint16_t* a, const uint16_t* b, float* d) {
    const int lane = threadIdx.x;
    bf16x8 av, bv;
    uint16_t ta[8], tb[8];
    for (int j = 0; j < 8; j++) {
        ta[j] = a[(lane & 15) * 32 + (lane >> 4) * 8 + j];   // A[row][k]
        tb[j] = b[((lane >> 4) * 8 + j) * 16 + (lane & 15)];  // B[k][col]
    }
    memcpy(&av, ta, 16);
    memcpy(&bv, tb, 16);
    f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    for (int i = 0; i < 4; i++) d[((lane >> 4) * 4 + i) * 16 + (lane & 15)] = acc[i];
}

int sf_cnn_mfma_probe(sf_ctx* ctx, const float* a, const float* b, float* d) {
    SF_CHECK_CTX(ctx);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<uint16_t> ha(16 * 32), hb(32 * 16);
    for (int i = 0; i < 16 * 32; i++) ha[i] = f2bf(a[i]);
    for (int i = 0; i < 32 * 16; i++) hb[i] = f2bf(b[i]);
    uint16_t *da = nullptr, *db = nullptr;
    float* dd = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&da, 16 * 32 * 2));
    SF_HIP(ctx, hipMalloc((void**)&db, 32 * 16 * 2));
    SF_HIP(ctx, hipMalloc((void**)&dd, 16 * 16 * 4));
    SF_HIP(ctx, hipMemcpy(da, ha.data(), 16 * 32 * 2, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(db, hb.data(), 32 * 16 * 2, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, ctx->stream, (const uint16_t*)da, (const uint16_t*)db,
                       dd);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(d, dd, 16 * 16 * 4, hipMemcpyDeviceToHost));
    hipFree(da);
    hipFree(db);
    hipFree(dd);
    return SF_OK;
}

}  // extern "C"
