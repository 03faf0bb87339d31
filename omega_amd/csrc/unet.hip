// unet.hip -- kernels of the 3-D U-Net that supplies the OSEM-RDP warm start: BF16 activations and
// weights, FP32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16), FP32 master weights.
//
// Layout: activations are channels-last, act[voxel][channel] with voxel = (z * H + y) * W + x, channel
// counts multiples of 16; conv weights are w[tap][cout][cin] with tap = (dz * 3 + dy) * 3 + dx.
// A 3x3x3 convolution (zero padding 1) is an implicit GEMM: Y[v][co] = sum_{tap, ci} W[tap][co][ci] *
// X[v + off(tap)][ci]. One wave computes 4 tiles of 16 voxels x all output channels; per (tap, 16-channel
// block) each lane feeds the matrix core 4 consecutive input channels of one voxel per tile (8-byte
// loads) and 4 consecutive input channels of one weight row, reused by the 4 tiles. Lane l of a 16x16x16 MFMA holds A[l % 16][4 (l / 16) + j],
// B[4 (l / 16) + j][l % 16] and receives D[4 (l / 16) + j][l % 16], j = 0..3.
//
// Backward: dX is the same convolution of dY with the flipped, transposed weights (k_conv_flipT);
// dW[tap] = sum_v dY[v]^T X[v + off(tap)] is a GEMM over voxels with partial sums per wave
// (k_conv3d_wgrad), reduced with float atomics; ReLU, pooling, upsampling, the 1x1 output layer,
// the loss and Adam are elementwise kernels.
#include "common.hpp"

#include <algorithm>

namespace omg {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even; NaN stays NaN
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct Vol { int D, H, W; };

__device__ __forceinline__ bool neighbour(const Vol& g, int v, int tap, int& nb) {
    const int x = v % g.W, t = v / g.W, y = t % g.H, z = t / g.H;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    nb = (zz * g.H + yy) * g.W + xx;
    return zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
}

__device__ __forceinline__ bf4 load4(const uint16_t* p) { return *reinterpret_cast<const bf4*>(p); }

// Y = act(conv(X) + b), X = [X1 | X2] (channel concatenation), MT = Cout / 16. A wave computes NT tiles of
// 16 consecutive voxels x all output channels: per (tap, 16-channel block) it loads NT activation
// fragments and, for each of the MT output blocks, one weight fragment that feeds NT matrix-core
// operations (weights reused NT times from registers). 4 waves per block.
constexpr int kConvNT = 4;

template <int MT>
__global__ __launch_bounds__(256) void k_conv3d_fwd(const uint16_t* __restrict__ X1, int C1,
                                                    const uint16_t* __restrict__ X2, int C2,
                                                    const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                    uint16_t* __restrict__ Y, Vol g, int act) {
    constexpr int NT = kConvNT;
    const int V = g.D * g.H * g.W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = (blockIdx.x * 4 + wave) * 16 * NT;
    if (base >= V) return;
    const int Cin = C1 + C2, Cout = MT * 16;
    const int col = lane & 15, kq = (lane >> 4) * 4;
    int v[NT];
    bool vin[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        v[t] = base + t * 16 + col;
        vin[t] = v[t] < V;
    }
    f4v acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f4v{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; tap++) {
        int nb[NT];
        bool ok[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            nb[t] = 0;
            ok[t] = vin[t] && neighbour(g, v[t], tap, nb[t]);
        }
        for (int cb = 0; cb < Cin; cb += 16) {
            const int c = cb + kq;
            bf4 b[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                b[t] = bf4{0, 0, 0, 0};
                if (ok[t]) b[t] = (c < C1) ? load4(X1 + (size_t)nb[t] * C1 + c) : load4(X2 + (size_t)nb[t] * C2 + (c - C1));
            }
            const uint16_t* wrow = Wt + ((size_t)tap * Cout + col) * Cin + c;
#pragma unroll
            for (int m = 0; m < MT; m++) {
                const bf4 a = load4(wrow + (size_t)m * 16 * Cin);
#pragma unroll
                for (int t = 0; t < NT; t++)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc[m][t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (!vin[t]) continue;
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = m * 16 + kq;
            uint16_t o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float val = acc[m][t][j] + (bias ? bias[co + j] : 0.f);
                if (act == 1) val = fmaxf(val, 0.f);
                o[j] = f2bf(val);
            }
            uint2 pk;
            pk.x = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
            pk.y = (uint32_t)o[2] | ((uint32_t)o[3] << 16);
            *reinterpret_cast<uint2*>(Y + (size_t)v[t] * Cout + co) = pk;
        }
    }
}

// The same convolution on gfx950's K = 32 matrix-core operation (v_mfma_f32_16x16x32_bf16) with the input
// staged in LDS. A workgroup of WAVES waves owns a block of 16 (x) x TY (y) x TZ (z) output voxels and all
// Cout channels, TY * TZ = NT * WAVES rows of 16 voxels; it first copies the block's input halo,
// 18 x (TY + 2) x (TZ + 2) voxels x CIN channels (zero outside the volume), into LDS (all of a lane's loads
// issued before its stores), so every input voxel is read from memory once per workgroup instead of once per
// tap. The reduction runs over K = 27 CIN in chunks of 32, fully unrolled: lane l takes the 8 consecutive k
// of its quarter, k = 32 kc + 8 (l / 16) + j, i.e. tap k / CIN and channels k % CIN .. +7 (CIN % 8 == 0, so the
// 8 never straddle a tap), reads them as one 16-byte LDS load of its voxel's shifted neighbour (B[k][voxel])
// and one 16-byte load of weight row co = 16 m + l % 16 (A[co][k]), the latter issued two chunks ahead; the
// result D[co][voxel] holds output channels 16 m + 4 (l / 16) + j of voxel l % 16. Wave w computes rows
// NT w .. NT w + NT - 1 for all MT output blocks: every weight fragment feeds NT operations and every
// activation fragment MT. LDS slots are CIN elements apart (measured faster than padded strides).
// Tile shapes by volume (measured on 128 x 128 x 64, 64 x 64 x 32 and 32 x 32 x 16): 8 waves on 16 x 8 x 4
// where the volume gives at least 512 such blocks, 4 waves on 16 x 4 x 4 where it gives 256, otherwise 4
// waves on 16 x 4 x 1 rows (one row per wave), so that small volumes still occupy the chip.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int CIN, int WAVES, int NT, int TY, int MW = 1>
constexpr size_t conv_tile_lds() {
    return (size_t)18 * (TY + 2) * (NT * WAVES / MW / TY + 2) * CIN * sizeof(uint16_t);
}

// MW: the block's waves split the output channels MW ways (wave w takes channel group w % MW and row group
// w / MW), so a wave streams 1 / MW of the weights and reuses each weight fragment over its NT rows; the
// halo in LDS is shared by all of them.
template <int CIN, int MT, int WAVES, int NT, int TY, int MW = 1>
__global__ __launch_bounds__(64 * WAVES) void k_conv3d_fwd_tile(const uint16_t* __restrict__ X1, int C1,
                                                                const uint16_t* __restrict__ X2,
                                                                const uint16_t* __restrict__ Wt,
                                                                const float* __restrict__ bias,
                                                                uint16_t* __restrict__ Y, Vol g, int act, int tiles_x,
                                                                int tiles_y) {
    constexpr int RW = WAVES / MW, TX = 16, TZ = NT * RW / TY, HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
    constexpr int HS = HX * HY * HZ;
    constexpr int S = CIN, Cout = MT * 16 * MW, K = 27 * CIN, KC = (K + 31) / 32, NTH = 64 * WAVES, PF = 2;
    static_assert(WAVES % MW == 0 && TZ * TY == NT * RW, "rows of the block");
    extern __shared__ uint16_t halo[];
    const int bx = blockIdx.x % tiles_x, rest = blockIdx.x / tiles_x;
    const int x0 = bx * TX, y0 = (rest % tiles_y) * TY, z0 = (rest / tiles_y) * TZ;
    const int C2 = CIN - C1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    const int cg = wave % MW, rw = wave / MW;
    const bf8 zero8 = {};
    auto wload = [&](int kc, int m) -> bf8 {
        const int k = kc * 32 + kg * 8;
        const bool in = kc < KC && ((K % 32 == 0) || k < K);
        const int tap = in ? k / CIN : 0, ci = k - tap * CIN;
        return in ? *reinterpret_cast<const bf8*>(Wt + ((size_t)tap * Cout + (cg * MT + m) * 16 + col) * CIN + ci)
                  : zero8;
    };
    bf8 wq[PF][MT];
#pragma unroll
    for (int p = 0; p < PF; p++)
#pragma unroll
        for (int m = 0; m < MT; m++) wq[p][m] = wload(p, m);
    // halo fill, in batches of up to 8 16-byte loads per lane, all issued before their stores
    constexpr int CH = CIN / 8, FILL = (HS * CH + NTH - 1) / NTH, FB = FILL < 8 ? FILL : 8;
#pragma unroll
    for (int it0 = 0; it0 < FILL; it0 += FB) {
        uint4 buf[FB];
#pragma unroll
        for (int j = 0; j < FB; j++) {
            const int i = threadIdx.x + (it0 + j) * NTH;
            buf[j] = uint4{0u, 0u, 0u, 0u};
            if (it0 + j < FILL && i < HS * CH) {
                const int slot = i / CH, c8 = (i - slot * CH) * 8;
                const int hx = slot % HX, r = slot / HX, hy = r % HY, hz = r / HY;
                const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
                if (x >= 0 && x < g.W && y >= 0 && y < g.H && z >= 0 && z < g.D) {
                    const size_t v = ((size_t)z * g.H + y) * g.W + x;
                    buf[j] = (c8 < C1) ? *reinterpret_cast<const uint4*>(X1 + v * C1 + c8)
                                       : *reinterpret_cast<const uint4*>(X2 + v * C2 + (c8 - C1));
                }
            }
        }
#pragma unroll
        for (int j = 0; j < FB; j++) {
            const int i = threadIdx.x + (it0 + j) * NTH;
            if (it0 + j < FILL && i < HS * CH) {
                const int slot = i / CH, c8 = (i - slot * CH) * 8;
                *reinterpret_cast<uint4*>(halo + slot * S + c8) = buf[j];
            }
        }
    }
    __syncthreads();
    f4v acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        const int k = kc * 32 + kg * 8;
        const bool kin = (K % 32 == 0) || k < K;
        const int tap = kin ? k / CIN : 0, ci = k - tap * CIN;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;   // halo coordinates start one voxel before the block
        bf8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            a[m] = wq[kc % PF][m];
            wq[kc % PF][m] = wload(kc + PF, m);
        }
        bf8 b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int row = rw * NT + t, yy = row % TY, zz = row / TY;
            const int slot = ((zz + dz) * HY + (yy + dy)) * HX + col + dx;
            b[t] = kin ? *reinterpret_cast<const bf8*>(halo + slot * S + ci) : zero8;
        }
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[t], acc[m][t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int row = rw * NT + t;
        const int x = x0 + col, y = y0 + row % TY, z = z0 + row / TY;
        if (x >= g.W || y >= g.H || z >= g.D) continue;
        const size_t v = ((size_t)z * g.H + y) * g.W + x;
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = (cg * MT + m) * 16 + kg * 4;
            uint16_t o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float val = acc[m][t][j] + (bias ? bias[co + j] : 0.f);
                if (act == 1) val = fmaxf(val, 0.f);
                o[j] = f2bf(val);
            }
            uint2 pk;
            pk.x = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
            pk.y = (uint32_t)o[2] | ((uint32_t)o[3] << 16);
            *reinterpret_cast<uint2*>(Y + v * Cout + co) = pk;
        }
    }
}

// dW[tap][co][ci] += sum over this wave's voxels of dY[v][co] * X[v + off(tap)][ci]. grid = (chunks, 27).
// ROWS (W % 16 == 0, vox_per_wave % 16 == 0): a 16-voxel K block lies in one image row, so the row's
// (y, z) and the tap's row offset are found once per block and each voxel's neighbour is v + offset with
// only its x checked (instead of a division-based neighbour() per voxel).
template <int MT, int NT, bool ROWS>
__global__ __launch_bounds__(256) void k_conv3d_wgrad(const uint16_t* __restrict__ dY,
                                                      const uint16_t* __restrict__ X1, int C1,
                                                      const uint16_t* __restrict__ X2, int C2,
                                                      float* __restrict__ dW, Vol g, int vox_per_wave) {
    const int V = g.D * g.H * g.W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.y;
    const int Cout = MT * 16, Cin = NT * 16;
    const int v0 = (blockIdx.x * 4 + wave) * vox_per_wave;
    if (v0 >= V) return;
    const int v1 = min(V, v0 + vox_per_wave);
    const int col = lane & 15, kq = (lane >> 4) * 4;
    f4v acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = f4v{0.f, 0.f, 0.f, 0.f};
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int toff = (dz * g.H + dy) * g.W + dx;
    for (int kb = v0; kb < v1; kb += 16) {
        // this lane's 4 voxels of the K block
        int nbv[4];
        bool okv[4];
        if (ROWS) {
            const int row = kb / g.W, y = row % g.H, z = row / g.H;
            const bool row_ok = z + dz >= 0 && z + dz < g.D && y + dy >= 0 && y + dy < g.H;
            const int x0 = kb - row * g.W + kq + dx;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int v = kb + kq + j;
                nbv[j] = v + toff;
                okv[j] = row_ok && v < v1 && x0 + j >= 0 && x0 + j < g.W;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int v = kb + kq + j;
                okv[j] = v < v1 && neighbour(g, v, tap, nbv[j]);
            }
        }
        bf4 a[MT];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = m * 16 + col;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int v = kb + kq + j;
                a[m][j] = (v < v1) ? (short)dY[(size_t)v * Cout + co] : (short)0;
            }
        }
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const int ci = n * 16 + col;
            bf4 b;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint16_t xv = 0;
                if (okv[j]) xv = (ci < C1) ? X1[(size_t)nbv[j] * C1 + ci] : X2[(size_t)nbv[j] * C2 + (ci - C1)];
                b[j] = (short)xv;
            }
#pragma unroll
            for (int m = 0; m < MT; m++) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc[m][n], 0, 0, 0);
        }
    }
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int co = m * 16 + kq + j, ci = n * 16 + col;
                unsafeAtomicAdd(dW + ((size_t)tap * Cout + co) * Cin + ci, acc[m][n][j]);
            }
}

// ---- weight gradient on the K = 32 BF16 matrix cores ---------------------------------------------------
// dW[tap][co][ci] = sum_v dY[v][co] X[v + off(tap)][ci] is a product whose K dimension is the voxels. The
// MFMA's lanes hold 8 consecutive K values of one row (A) or column (B), so both operands are staged
// transposed in LDS: dYt[co][voxel] for the block's 32 x 4 x 2 voxels and, for one 16-channel slice of X
// (blockIdx.y), Xt[dx + 1][ci][halo row][x] -- three copies of the (2 + 2) x (4 + 2) halo rows shifted by
// dx = -1, 0, 1, so that every operand is one aligned 16-byte LDS read. The 8 waves split the 27 taps
// (wave w: w, w + 8, w + 16, w + 24) and reuse each row's dY operand across theirs. Blocks persist over
// the tiles (blockIdx.x, step gridDim.x) and store their sums as one partial per block, added into dW by
// k_wgrad_reduce (no atomics).
constexpr int kWgTX = 32, kWgTY = 4, kWgTZ = 2, kWgRows = kWgTY * kWgTZ, kWgVox = kWgTX * kWgRows;
constexpr int kWgHY = kWgTY + 2, kWgHZ = kWgTZ + 2, kWgHRows = kWgHY * kWgHZ, kWgThreads = 512;
// channel-row strides of the LDS operands, padded by 16 B so that the 16 lanes of a read phase (one
// channel each, 16 B apart otherwise a multiple of the 256-byte bank cycle) fall on distinct banks
constexpr int kWgDS = kWgVox + 8, kWgXS = kWgHRows * kWgTX + 8;

// u[e] holds 8 channels of voxel e; channel j's 4 voxels go to dst + j * stride as one 8-byte write.
__device__ __forceinline__ void wg_store4(uint16_t* dst, size_t stride, const uint4 (&u)[4]) {
    const uint32_t w[4][4] = {{u[0].x, u[0].y, u[0].z, u[0].w}, {u[1].x, u[1].y, u[1].z, u[1].w},
                              {u[2].x, u[2].y, u[2].z, u[2].w}, {u[3].x, u[3].y, u[3].z, u[3].w}};
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int sh = 16 * (j & 1), k = j >> 1;
        const uint32_t lo = ((w[0][k] >> sh) & 0xffffu) | (((w[1][k] >> sh) & 0xffffu) << 16);
        const uint32_t hi = ((w[2][k] >> sh) & 0xffffu) | (((w[3][k] >> sh) & 0xffffu) << 16);
        *reinterpret_cast<uint2*>(dst + (size_t)j * stride) = uint2{lo, hi};
    }
}

template <int MT>
constexpr size_t wgrad_tile_lds() { return sizeof(uint16_t) * ((size_t)MT * 16 * kWgDS + 3 * 16 * (size_t)kWgXS); }

template <int MT>
__global__ __launch_bounds__(kWgThreads) void k_conv3d_wgrad_tile(const uint16_t* __restrict__ dY,
                                                                  const uint16_t* __restrict__ X1, int C1,
                                                                  const uint16_t* __restrict__ X2, int C2,
                                                                  float* __restrict__ part, Vol g, int tiles_x,
                                                                  int tiles_y, int n_tiles) {
    constexpr int Cout = MT * 16;
    extern __shared__ uint16_t wl[];
    uint16_t* dYt = wl;                              // [Cout][kWgVox], row stride kWgDS
    uint16_t* Xt = wl + (size_t)Cout * kWgDS;        // [3][16][kWgHRows][kWgTX], channel stride kWgXS
    const int Cin = C1 + C2;
    const int n = blockIdx.y;                        // input-channel slice [16 n, 16 n + 16)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4;
    const int ntap = wave < 3 ? 4 : 3;
    f4v acc[4][MT];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int m = 0; m < MT; m++) acc[t][m] = f4v{0.f, 0.f, 0.f, 0.f};
    const uint16_t* Xs = (n * 16 < C1) ? X1 : X2;    // the slice lies in one input (C1 % 16 == 0)
    const int Cs = (n * 16 < C1) ? C1 : C2;
    const int cs0 = (n * 16 < C1) ? n * 16 : n * 16 - C1;
    for (int ti = blockIdx.x; ti < n_tiles; ti += gridDim.x) {
        const int bx = ti % tiles_x, rest = ti / tiles_x;
        const int x0 = bx * kWgTX, y0 = (rest % tiles_y) * kWgTY, z0 = (rest / tiles_y) * kWgTZ;
        // Transposing fills: an item is 4 consecutive voxels of 8 channels (four 16-byte loads), stored as one
        // 8-byte LDS write per channel. dY tile:
        for (int i = threadIdx.x; i < (kWgVox / 4) * (Cout / 8); i += kWgThreads) {
            const int quad = i / (Cout / 8), c8 = (i - quad * (Cout / 8)) * 8;
            const int vox = quad * 4;
            const int vx = vox % kWgTX, r = vox / kWgTX, vy = r % kWgTY, vz = r / kWgTY;
            const int y = y0 + vy, z = z0 + vz;
            uint4 u[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int x = x0 + vx + e;
                u[e] = uint4{0u, 0u, 0u, 0u};
                if (x < g.W && y < g.H && z < g.D)
                    u[e] = *reinterpret_cast<const uint4*>(dY + (((size_t)z * g.H + y) * g.W + x) * Cout + c8);
            }
            wg_store4(dYt + (size_t)c8 * kWgDS + vox, kWgDS, u);
        }
        // X slice: halo rows, and in copy d the voxels x0 + p + d - 1 at p = 0..31 (x from x0 - 1 to x0 + 32)
        for (int i = threadIdx.x; i < kWgHRows * (kWgTX / 4) * 2 * 3; i += kWgThreads) {
            const int d = i % 3, r2 = i / 3;
            const int c8 = (r2 & 1) * 8, r3 = r2 >> 1;
            const int xq = r3 % (kWgTX / 4), hrow = r3 / (kWgTX / 4);
            const int hy = hrow % kWgHY, hz = hrow / kWgHY;
            const int y = y0 + hy - 1, z = z0 + hz - 1;
            const bool row_in = y >= 0 && y < g.H && z >= 0 && z < g.D;
            uint4 u[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int x = x0 + xq * 4 + e + d - 1;
                u[e] = uint4{0u, 0u, 0u, 0u};
                if (row_in && x >= 0 && x < g.W)
                    u[e] = *reinterpret_cast<const uint4*>(Xs + (((size_t)z * g.H + y) * g.W + x) * Cs + cs0 + c8);
            }
            wg_store4(Xt + (size_t)(d * 16 + c8) * kWgXS + hrow * kWgTX + xq * 4, kWgXS, u);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kWgRows; r++) {
            const int ry = r % kWgTY, rz = r / kWgTY;
            bf8 a[MT];
#pragma unroll
            for (int m = 0; m < MT; m++)
                a[m] = *reinterpret_cast<const bf8*>(dYt + (size_t)(m * 16 + col) * kWgDS + r * kWgTX + 8 * q);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t >= ntap) break;
                const int tap = wave + 8 * t;
                const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;   // 0..2: offset + 1
                const int hrow = (rz + dz) * kWgHY + (ry + dy);
                const bf8 b = *reinterpret_cast<const bf8*>(Xt + (size_t)(dx * 16 + col) * kWgXS + hrow * kWgTX + 8 * q);
#pragma unroll
                for (int m = 0; m < MT; m++) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[t][m], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // this block's sums: part[block][tap][co][ci]
    const size_t per = (size_t)27 * Cout * Cin;
    float* pb = part + (size_t)blockIdx.x * per;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (t >= ntap) break;
        const int tap = wave + 8 * t;
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                pb[((size_t)tap * Cout + m * 16 + 4 * q + j) * Cin + n * 16 + col] = acc[t][m][j];
    }
}

// dW[i] += sum over the P partials of part[p][i]: a block takes 64 consecutive outputs (lane l: i0 + l,
// coalesced), its 4 waves the partials p = wave, wave + 4, ..., combined through LDS.
__global__ __launch_bounds__(1024) void k_wgrad_reduce(const float* __restrict__ part, float* __restrict__ dW,
                                                       size_t per, int P) {
    // 16 waves per 64 outputs, wave w summing partials w, w + 16, ...; folded in a fixed order
    __shared__ float red[16][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t i = (size_t)blockIdx.x * 64 + lane;
    float s = 0.f;
    if (i < per) {
#pragma unroll 4
        for (int p = wave; p < P; p += 16) s += part[(size_t)p * per + i];
    }
    red[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && i < per) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < 16; w++) t += red[w][lane];
        dW[i] += t;
    }
}

// Wf[tap][ci][co] = W[26 - tap][co][ci]: dX = conv(dY) with these weights.
__global__ void k_conv_flipT(const uint16_t* __restrict__ W, uint16_t* __restrict__ Wf, int Cout, int Cin) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 27 * Cout * Cin) return;
    const int ci = i % Cin, t = i / Cin, co = t % Cout, tap = t / Cout;
    Wf[((size_t)(26 - tap) * Cin + ci) * Cout + co] = W[i];
}

// db[c] += sum_v dY[v][c]: per-thread partial sums, reduced over the block's rows in LDS, then one
// atomic per channel per block (a float atomic per thread onto C addresses serialises).
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dY, float* __restrict__ db, int V, int C) {
    __shared__ float part[256];
    const int c = threadIdx.x % C;
    const int lanes = 256 / C;     // C divides 256 for C in {16, 32, 64}
    const int r = threadIdx.x / C;
    float s = 0.f;
    for (int v = blockIdx.x * lanes + r; v < V; v += gridDim.x * lanes) s += bf2f(dY[(size_t)v * C + c]);
    part[threadIdx.x] = s;
    __syncthreads();
    if (r == 0) {
        for (int k = 1; k < lanes; k++) s += part[k * C + c];
        unsafeAtomicAdd(db + c, s);
    }
}

// k_bias_grad reading 8 channels (16 bytes) per thread: the block's 256 / (C / 8) rows keep 8 partial sums
// each, folded in two LDS passes (8 rows per thread, then 256 / C per channel), then one atomic per channel.
__global__ __launch_bounds__(256) void k_bias_grad8(const uint16_t* __restrict__ dY, float* __restrict__ db, int V,
                                                    int C) {
    __shared__ float part[2048];
    const int g8 = C / 8;                // 16-byte groups per voxel row
    const int rows = 256 / g8;
    const int r = threadIdx.x / g8, c8 = (threadIdx.x - r * g8) * 8;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int64_t v = (int64_t)blockIdx.x * rows + r; v < V; v += (int64_t)gridDim.x * rows) {
        const uint4 u = *reinterpret_cast<const uint4*>(dY + (size_t)v * C + c8);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += bf2f((uint16_t)(w[j >> 1] >> (16 * (j & 1))));
    }
#pragma unroll
    for (int j = 0; j < 8; j++) part[r * C + c8 + j] = s[j];   // rows * C == 2048
    __syncthreads();
    const int c = threadIdx.x % C, q = threadIdx.x / C, step = 256 / C;
    float t = 0.f;
    for (int k = q; k < rows; k += step) t += part[k * C + c];
    __syncthreads();
    part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x < C) {
        float u = 0.f;
        for (int k = 0; k < step; k++) u += part[k * C + threadIdx.x];
        unsafeAtomicAdd(db + threadIdx.x, u);
    }
}

// dZ = (Y > 0) ? dY : 0 over n elements (in place allowed)
__global__ void k_relu_bwd(const uint16_t* __restrict__ Y, uint16_t* dY, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !(bf2f(Y[i]) > 0.f)) dY[i] = 0;
}

// The same, 8 elements (16 bytes) per thread; n8 = n / 8.
__global__ void k_relu_bwd8(const uint4* __restrict__ Y, uint4* dY, size_t n8) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n8) return;
    const uint4 y = Y[i];
    uint4 d = dY[i];
    const uint32_t yw[4] = {y.x, y.y, y.z, y.w};
    uint32_t dw[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!(bf2f((uint16_t)(yw[k] & 0xffffu)) > 0.f)) dw[k] &= 0xffff0000u;
        if (!(bf2f((uint16_t)(yw[k] >> 16)) > 0.f)) dw[k] &= 0x0000ffffu;
    }
    dY[i] = uint4{dw[0], dw[1], dw[2], dw[3]};
}

// Nearest 2x upsampling, 8 channels (16 bytes) per thread: grid = (ceil(W * C / 8 / 256), H, D) of the fine grid.
__global__ void k_up_fwd8(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, Vol g, int C) {
    const int c8n = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.W * c8n) return;
    const int x = i / c8n, c8 = (i - x * c8n) * 8, y = blockIdx.y, z = blockIdx.z;
    const int Ho = g.H / 2, Wo = g.W / 2;
    const size_t vo = ((size_t)(z / 2) * Ho + y / 2) * Wo + x / 2;
    const size_t v = ((size_t)z * g.H + y) * g.W + x;
    *reinterpret_cast<uint4*>(Y + v * C + c8) = *reinterpret_cast<const uint4*>(X + vo * C + c8);
}

// k_up_bwd, 8 channels per thread: grid = (ceil(W * C / 8 / 256), H, D) of the coarse grid g.
__global__ void k_up_bwd8(const uint16_t* __restrict__ dY, int Cs, int off, uint16_t* __restrict__ dX, Vol g, int C) {
    const int c8n = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.W * c8n) return;
    const int xo = i / c8n, c8 = (i - xo * c8n) * 8, yo = blockIdx.y, zo = blockIdx.z;
    const int H2 = g.H * 2, W2 = g.W * 2;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 8; d++) {   // children in the order of k_up_bwd
        const int z = 2 * zo + (d >> 2), y = 2 * yo + ((d >> 1) & 1), x = 2 * xo + (d & 1);
        const uint4 u = *reinterpret_cast<const uint4*>(dY + ((size_t)(z * H2 + y) * W2 + x) * Cs + off + c8);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += bf2f((uint16_t)(w[j >> 1] >> (16 * (j & 1))));
    }
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (uint32_t)f2bf(s[2 * k]) | ((uint32_t)f2bf(s[2 * k + 1]) << 16);
    const size_t vo = ((size_t)zo * g.H + yo) * g.W + xo;
    *reinterpret_cast<uint4*>(dX + vo * C + c8) = uint4{o[0], o[1], o[2], o[3]};
}

// 2x2x2 average pooling of X[D][H][W][C] (even D, H, W)
__global__ void k_pool_fwd(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, Vol g, int C) {
    const int Ho = g.H / 2, Wo = g.W / 2, Do = g.D / 2;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)Do * Ho * Wo * C) return;
    const int c = (int)(i % C);
    const size_t vo = i / C;
    const int xo = (int)(vo % Wo), yo = (int)((vo / Wo) % Ho), zo = (int)(vo / ((size_t)Wo * Ho));
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int z = 2 * zo + (d >> 2), y = 2 * yo + ((d >> 1) & 1), x = 2 * xo + (d & 1);
        s += bf2f(X[((size_t)(z * g.H + y) * g.W + x) * C + c]);
    }
    Y[i] = f2bf(s * 0.125f);
}

// dX (fine grid, C channels, stride/offset into dst) (+)= dY(coarse)[..][c] / 8
__global__ void k_pool_bwd(const uint16_t* __restrict__ dY, uint16_t* __restrict__ dX, Vol g, int C, int accumulate) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)g.D * g.H * g.W * C) return;
    const int c = (int)(i % C);
    const size_t v = i / C;
    const int x = (int)(v % g.W), y = (int)((v / g.W) % g.H), z = (int)(v / ((size_t)g.W * g.H));
    const int Ho = g.H / 2, Wo = g.W / 2;
    const float val = bf2f(dY[((size_t)((z / 2) * Ho + y / 2) * Wo + x / 2) * C + c]) * 0.125f;
    dX[i] = f2bf(accumulate ? bf2f(dX[i]) + val : val);
}

// nearest-neighbour x2 upsampling: Y(fine)[v][c] = X(coarse)[parent(v)][c]; g = fine grid
__global__ void k_up_fwd(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, Vol g, int C) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)g.D * g.H * g.W * C) return;
    const int c = (int)(i % C);
    const size_t v = i / C;
    const int x = (int)(v % g.W), y = (int)((v / g.W) % g.H), z = (int)(v / ((size_t)g.W * g.H));
    const int Ho = g.H / 2, Wo = g.W / 2;
    Y[i] = X[((size_t)((z / 2) * Ho + y / 2) * Wo + x / 2) * C + c];
}

// dX(coarse)[vo][c] = sum of the 8 children dY(fine)[v][off + c] (row stride Cs); g = coarse grid
__global__ void k_up_bwd(const uint16_t* __restrict__ dY, int Cs, int off, uint16_t* __restrict__ dX, Vol g, int C) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)g.D * g.H * g.W * C) return;
    const int c = (int)(i % C);
    const size_t vo = i / C;
    const int xo = (int)(vo % g.W), yo = (int)((vo / g.W) % g.H), zo = (int)(vo / ((size_t)g.W * g.H));
    const int H2 = g.H * 2, W2 = g.W * 2;
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int z = 2 * zo + (d >> 2), y = 2 * yo + ((d >> 1) & 1), x = 2 * xo + (d & 1);
        s += bf2f(dY[((size_t)(z * H2 + y) * W2 + x) * Cs + off + c]);
    }
    dX[i] = f2bf(s);
}

// dst[v][c] (+)= src[v][off + c] (row stride Cs), C channels
__global__ void k_take_channels(const uint16_t* __restrict__ src, int Cs, int off, uint16_t* __restrict__ dst,
                                size_t V, int C, int accumulate) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * C) return;
    const size_t v = i / C;
    const int c = (int)(i % C);
    const float val = bf2f(src[v * Cs + off + c]);
    dst[i] = f2bf(accumulate ? bf2f(dst[i]) + val : val);
}

// Output layer: out[v] = base[v] + b + sum_c w[c] * X[v][c]  (FP32)
__global__ void k_head_fwd(const uint16_t* __restrict__ X, const float* __restrict__ w, const float* __restrict__ b,
                           const float* __restrict__ base, float* __restrict__ out, int V, int C) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    float s = b[0];
    for (int c = 0; c < C; c++) s += w[c] * bf2f(X[(size_t)v * C + c]);
    out[v] = base[v] + s;
}

// Head backward: dX[v][c] = g[v] w[c] (BF16); dw[c] += sum g[v] X[v][c]; db += sum g[v]
__global__ void k_head_bwd(const uint16_t* __restrict__ X, const float* __restrict__ w, const float* __restrict__ gout,
                           uint16_t* __restrict__ dX, float* __restrict__ dw, float* __restrict__ db, int V, int C) {
    __shared__ float red[64][17];
    const int v = blockIdx.x * 256 + threadIdx.x;
    const float gv = v < V ? gout[v] : 0.f;
    float part[64];
    for (int c = 0; c < C; c++) {
        const float xv = v < V ? bf2f(X[(size_t)v * C + c]) : 0.f;
        part[c] = gv * xv;
        if (v < V) dX[(size_t)v * C + c] = f2bf(gv * w[c]);
    }
    // reduce part[c] over the block: per wave with shuffles, then atomics
    for (int c = 0; c < C; c++) {
        float s = part[c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = s;
    }
    float sg = gv;
    for (int o = 32; o > 0; o >>= 1) sg += __shfl_xor(sg, o, 64);
    if ((threadIdx.x & 63) == 0) red[C][threadIdx.x >> 6] = sg;  // C < 64: row C free
    __syncthreads();
    if (threadIdx.x < C) {
        const float s = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
        unsafeAtomicAdd(dw + threadIdx.x, s);
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(db, red[C][0] + red[C][1] + red[C][2] + red[C][3]);
}

// Masked mean squared error: loss += sum m (out - t)^2 / n_valid (FP64 accumulation); g = 2 m (out - t) / n_valid.
// Grid-stride over a capped grid with the block's sum reduced in LDS: one FP64 atomic per block (atomics
// on the one loss address serialise, so one per wave costs more than the whole pass).
constexpr int kMseBlocks = 512;
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out, const float* __restrict__ t,
                                             const float* __restrict__ mask, float* __restrict__ g,
                                             double* __restrict__ loss, int V, float inv_n) {
    __shared__ double part[4];
    double e2 = 0.0;
    for (int v = blockIdx.x * 256 + threadIdx.x; v < V; v += gridDim.x * 256) {
        const float m = mask ? mask[v] : 1.f;
        const float e = (out[v] - t[v]) * m;
        g[v] = 2.f * e * inv_n;
        e2 += (double)e * e * inv_n;
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// Adam on FP32 master weights p with moments m, v; writes the BF16 copy q (q may be null).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ q, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (q) q[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

// Network input: image [Nz][Ny][Nx] * scale into channel 0 of a 16-channel BF16 volume [D][H][W][16]
// (D >= Nz, H = Ny, W = Nx; planes z >= Nz and channels 1..15 are zero).
__global__ void k_embed(const float* __restrict__ img, uint16_t* __restrict__ X, Vol g, int Nz, float scale) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t V = (size_t)g.D * g.H * g.W;
    if (i >= V * 16) return;
    const int c = (int)(i & 15);
    const size_t v = i >> 4;
    const int z = (int)(v / ((size_t)g.H * g.W));
    X[i] = (c == 0 && z < Nz) ? f2bf(img[v] * scale) : (uint16_t)0;
}

// Unpacks 8 BF16 values from 16 bytes.
__device__ inline void bf8_unpack(const uint4 u, float (&f)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int j = 0; j < 8; j++) f[j] = bf2f((uint16_t)(w[j >> 1] >> (16 * (j & 1))));
}

__device__ inline uint4 bf8_pack(const float (&f)[8]) {
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    return uint4{o[0], o[1], o[2], o[3]};
}

// k_pool_bwd, 8 channels per thread: grid = (ceil(W * C / 8 / 256), H, D) of the fine grid g.
__global__ void k_pool_bwd8(const uint16_t* __restrict__ dY, uint16_t* __restrict__ dX, Vol g, int C, int accumulate) {
    const int c8n = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.W * c8n) return;
    const int x = i / c8n, c8 = (i - x * c8n) * 8, y = blockIdx.y, z = blockIdx.z;
    const int Ho = g.H / 2, Wo = g.W / 2;
    float a[8];
    bf8_unpack(*reinterpret_cast<const uint4*>(dY + ((size_t)(z / 2) * Ho + y / 2) * Wo * C + (size_t)(x / 2) * C + c8), a);
    uint4* d = reinterpret_cast<uint4*>(dX + (((size_t)z * g.H + y) * g.W + x) * C + c8);
    float o[8];
    if (accumulate) bf8_unpack(*d, o);
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = accumulate ? o[j] + a[j] * 0.125f : a[j] * 0.125f;
    *d = bf8_pack(o);
}

// k_head_fwd with 16-byte channel loads (C a multiple of 8); channels summed in the same order.
__global__ void k_head_fwd8(const uint16_t* __restrict__ X, const float* __restrict__ w, const float* __restrict__ b,
                            const float* __restrict__ base, float* __restrict__ out, int V, int C) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    float s = b[0];
    for (int c8 = 0; c8 < C; c8 += 8) {
        float f[8];
        bf8_unpack(*reinterpret_cast<const uint4*>(X + (size_t)v * C + c8), f);
#pragma unroll
        for (int j = 0; j < 8; j++) s += w[c8 + j] * f[j];
    }
    out[v] = base[v] + s;
}

// k_head_bwd, 8 channels per thread over a capped grid: dX rows written as 16-byte stores, the weight and
// bias gradients kept as per-thread partial sums, folded in LDS, one atomic per channel per block.
__global__ __launch_bounds__(256) void k_head_bwd8(const uint16_t* __restrict__ X, const float* __restrict__ w,
                                                   const float* __restrict__ gout, uint16_t* __restrict__ dX,
                                                   float* __restrict__ dw, float* __restrict__ db, int V, int C) {
    __shared__ float part[2048];
    __shared__ float gpart[4];
    const int g8 = C / 8;
    const int rows = 256 / g8;
    const int r = threadIdx.x / g8, c8 = (threadIdx.x - r * g8) * 8;
    float wc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) wc[j] = w[c8 + j];
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float sg = 0.f;
    for (int64_t v = (int64_t)blockIdx.x * rows + r; v < V; v += (int64_t)gridDim.x * rows) {
        const float gv = gout[v];
        float f[8], d[8];
        bf8_unpack(*reinterpret_cast<const uint4*>(X + (size_t)v * C + c8), f);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            s[j] += gv * f[j];
            d[j] = gv * wc[j];
        }
        *reinterpret_cast<uint4*>(dX + (size_t)v * C + c8) = bf8_pack(d);
        if (c8 == 0) sg += gv;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) part[r * C + c8 + j] = s[j];   // rows * C == 2048
    for (int o = 32; o > 0; o >>= 1) sg += __shfl_xor(sg, o, 64);
    if ((threadIdx.x & 63) == 0) gpart[threadIdx.x >> 6] = sg;
    __syncthreads();
    const int c = threadIdx.x % C, q = threadIdx.x / C, step = 256 / C;
    float t = 0.f;
    for (int k = q; k < rows; k += step) t += part[k * C + c];
    __syncthreads();
    part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x < C) {
        float u = 0.f;
        for (int k = 0; k < step; k++) u += part[k * C + threadIdx.x];
        unsafeAtomicAdd(dw + threadIdx.x, u);
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(db, gpart[0] + gpart[1] + gpart[2] + gpart[3]);
}

// k_embed, one voxel (two 16-byte rows of 8 channels) per thread.
__global__ void k_embed16(const float* __restrict__ img, uint16_t* __restrict__ X, Vol g, int Nz, float scale) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= g.D * g.H * g.W) return;
    const int z = v / (g.H * g.W);
    uint32_t c0 = 0u;
    if (z < Nz) c0 = f2bf(img[v] * scale);
    uint4* d = reinterpret_cast<uint4*>(X + (size_t)v * 16);
    d[0] = uint4{c0, 0u, 0u, 0u};
    d[1] = uint4{0u, 0u, 0u, 0u};
}

// k_take_channels, 8 channels (16 bytes) per thread (C, Cs and off multiples of 8).
__global__ void k_take_channels8(const uint16_t* __restrict__ src, int Cs, int off, uint16_t* __restrict__ dst,
                                 size_t V, int C, int accumulate) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int c8n = C / 8;
    if (i >= V * c8n) return;
    const size_t v = i / c8n;
    const int c8 = (int)(i - v * c8n) * 8;
    const uint4 u = *reinterpret_cast<const uint4*>(src + v * Cs + off + c8);
    uint4* d = reinterpret_cast<uint4*>(dst + v * C + c8);
    if (!accumulate) {
        *d = u;
        return;
    }
    float a[8], o[8];
    bf8_unpack(u, a);
    bf8_unpack(*d, o);
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] += a[j];
    *d = bf8_pack(o);
}

static inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

template <int MT>
static void launch_fwd(const uint16_t* X1, int C1, const uint16_t* X2, int C2, const uint16_t* W, const float* b,
                       uint16_t* Y, Vol g, int act, hipStream_t s) {
    const int V = g.D * g.H * g.W;
    const int per_block = 4 * 16 * kConvNT;
    hipLaunchKernelGGL(k_conv3d_fwd<MT>, dim3((unsigned)((V + per_block - 1) / per_block)), dim3(256), 0, s, X1, C1, X2,
                       C2, W, b, Y, g, act);
}

// MT: output-channel tiles of 16 per wave; MW: channel groups across the block's waves (Cout = 16 MT MW).
template <int CIN, int MT, int WAVES, int NT, int TY, int MW = 1>
static int launch_fwd_tile(const uint16_t* X1, int C1, const uint16_t* X2, const uint16_t* W, const float* b,
                           uint16_t* Y, Vol g, int act, hipStream_t s) {
    constexpr int TZ = NT * WAVES / MW / TY;
    const int tx = (g.W + 15) / 16, ty = (g.H + TY - 1) / TY, tz = (g.D + TZ - 1) / TZ;
    const size_t lds = conv_tile_lds<CIN, WAVES, NT, TY, MW>();
    static bool attr = false;   // above 64 KiB of LDS the kernel must opt in once
    if (!attr) {
        OMG_CHECK_HIP(hipFuncSetAttribute(
            reinterpret_cast<const void*>(k_conv3d_fwd_tile<CIN, MT, WAVES, NT, TY, MW>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    hipLaunchKernelGGL((k_conv3d_fwd_tile<CIN, MT, WAVES, NT, TY, MW>), dim3((unsigned)(tx * ty * tz)),
                       dim3(64 * WAVES), lds, s, X1, C1, X2, W, b, Y, g, act, tx, ty);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

// Tile shape by the number of blocks the volume gives (see k_conv3d_fwd_tile). (Splitting the output
// channels over blocks as well, one 16-channel slice per block, was measured slower on 5 of the 7 layers
// with more than 16 of them: each slice block refills the halo.)
template <int CIN, int MT>
static int launch_fwd_shape(const uint16_t* X1, int C1, const uint16_t* X2, const uint16_t* W, const float* b,
                            uint16_t* Y, Vol g, int act, hipStream_t s) {
    const long bx = (g.W + 15) / 16;
    constexpr bool big_fits = conv_tile_lds<CIN, 8, 4, 8>() <= 160 * 1024;   // LDS of one CU
    if (big_fits && bx * ((g.H + 7) / 8) * ((g.D + 3) / 4) >= 512)
        return launch_fwd_tile<CIN, MT, big_fits ? 8 : 4, 4, big_fits ? 8 : 4>(X1, C1, X2, W, b, Y, g, act, s);
    // with several output-channel tiles the waves split them (each streams 1 / MW of the weights, and reuses
    // each fragment over more rows); one 16-voxel row per wave would stream every weight once per row
    if (bx * ((g.H + 3) / 4) * ((g.D + 3) / 4) >= 256) {
        if constexpr (MT % 2 == 0) return launch_fwd_tile<CIN, MT / 2, 4, 8, 4, 2>(X1, C1, X2, W, b, Y, g, act, s);
        return launch_fwd_tile<CIN, MT, 4, 4, 4>(X1, C1, X2, W, b, Y, g, act, s);
    }
    if constexpr (MT % 4 == 0) return launch_fwd_tile<CIN, MT / 4, 4, 4, 4, 4>(X1, C1, X2, W, b, Y, g, act, s);
    return launch_fwd_tile<CIN, MT, 4, 1, 4>(X1, C1, X2, W, b, Y, g, act, s);
}

template <int CIN>
static int dispatch_fwd_tile(int MT, const uint16_t* X1, int C1, const uint16_t* X2, const uint16_t* W, const float* b,
                             uint16_t* Y, Vol g, int act, hipStream_t s) {
    switch (MT) {
        case 1: return launch_fwd_shape<CIN, 1>(X1, C1, X2, W, b, Y, g, act, s);
        case 2: return launch_fwd_shape<CIN, 2>(X1, C1, X2, W, b, Y, g, act, s);
        case 3: return launch_fwd_shape<CIN, 3>(X1, C1, X2, W, b, Y, g, act, s);
        case 4: return launch_fwd_shape<CIN, 4>(X1, C1, X2, W, b, Y, g, act, s);
        case 6:   // 96 x 96 would spill: it takes the register kernel
            if constexpr (CIN != 96) return launch_fwd_shape<CIN, 6>(X1, C1, X2, W, b, Y, g, act, s);
            break;
    }
    return OMG_ERR_UNSUPPORTED;
}

// Partial sums of the tiled weight gradient: a library buffer grown on demand (kept for the process).
static float* g_wg_part = nullptr;
static size_t g_wg_floats = 0;

template <int MT>
static int launch_wgrad_tile(const uint16_t* dY, const uint16_t* X1, int C1, const uint16_t* X2, int C2, float* dW,
                             Vol g, hipStream_t s) {
    const int Cin = C1 + C2, NT = Cin / 16;
    const int tiles_x = (g.W + kWgTX - 1) / kWgTX, tiles_y = (g.H + kWgTY - 1) / kWgTY;
    const int n_tiles = tiles_x * tiles_y * ((g.D + kWgTZ - 1) / kWgTZ);
    const int P = std::max(1, std::min(n_tiles, 256));   // persistent blocks per channel slice
    const size_t per = (size_t)27 * MT * 16 * Cin;
    if ((size_t)P * per > g_wg_floats) {
        OMG_CHECK_HIP(hipStreamSynchronize(s));
        OMG_CHECK_HIP(hipFree(g_wg_part));
        g_wg_part = nullptr;
        g_wg_floats = 0;
        OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g_wg_part), (size_t)P * per * sizeof(float)));
        g_wg_floats = (size_t)P * per;
    }
    const size_t lds = wgrad_tile_lds<MT>();
    static bool attr = false;
    if (!attr) {
        OMG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_conv3d_wgrad_tile<MT>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    hipLaunchKernelGGL(k_conv3d_wgrad_tile<MT>, dim3((unsigned)P, (unsigned)NT), dim3(kWgThreads), lds, s, dY, X1, C1,
                       X2, C2, g_wg_part, g, tiles_x, tiles_y, n_tiles);
    OMG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((per + 63) / 64)), dim3(1024), 0, s, g_wg_part, dW, per, P);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

template <int MT, int NT>
static void launch_wgrad(const uint16_t* dY, const uint16_t* X1, int C1, const uint16_t* X2, int C2, float* dW, Vol g,
                         hipStream_t s) {
    const int V = g.D * g.H * g.W;
    const int vpw = 1024;
    const dim3 grid((unsigned)((V + 4 * vpw - 1) / (4 * vpw)), 27);
    if (g.W % 16 == 0)
        hipLaunchKernelGGL((k_conv3d_wgrad<MT, NT, true>), grid, dim3(256), 0, s, dY, X1, C1, X2, C2, dW, g, vpw);
    else
        hipLaunchKernelGGL((k_conv3d_wgrad<MT, NT, false>), grid, dim3(256), 0, s, dY, X1, C1, X2, C2, dW, g, vpw);
}

}  // namespace omg

using namespace omg;

static int vol_ok(int D, int H, int W) { return D > 0 && H > 0 && W > 0 && (int64_t)D * H * W < (1ll << 31); }

extern "C" int omg_conv3d_forward(const uint16_t* X1, int32_t C1, const uint16_t* X2, int32_t C2, const uint16_t* W,
                                  const float* bias, uint16_t* Y, int32_t Cout, int32_t D, int32_t H, int32_t Wd,
                                  int32_t relu, void* stream) {
    OMG_REQUIRE(vol_ok(D, H, Wd), "bad volume size");
    OMG_REQUIRE(X1 && W && Y && C1 > 0 && C1 % 16 == 0 && C2 >= 0 && C2 % 16 == 0 && (C2 == 0 || X2),
                "inputs must have a multiple of 16 channels");
    const Vol g{D, H, Wd};
    hipStream_t s = as_stream(stream);
    OMG_REQUIRE(Cout == 16 || Cout == 32 || Cout == 48 || Cout == 64 || Cout == 96,
                "Cout %d is not one of 16, 32, 48, 64, 96", Cout);
    // the LDS-tiled K = 32 kernel (16-byte operand loads) for the input widths the network uses; other widths
    // and pointers that are not 16-byte aligned take the register kernel
    const bool al16 = ((uintptr_t)X1 % 16 == 0) && (C2 == 0 || (uintptr_t)X2 % 16 == 0) && ((uintptr_t)W % 16 == 0) &&
                      ((uintptr_t)Y % 8 == 0);
    int rc = OMG_ERR_UNSUPPORTED;
    switch (al16 ? C1 + C2 : 0) {
        case 16: rc = dispatch_fwd_tile<16>(Cout / 16, X1, C1, X2, W, bias, Y, g, relu, s); break;
        case 32: rc = dispatch_fwd_tile<32>(Cout / 16, X1, C1, X2, W, bias, Y, g, relu, s); break;
        case 48: rc = dispatch_fwd_tile<48>(Cout / 16, X1, C1, X2, W, bias, Y, g, relu, s); break;
        case 64: rc = dispatch_fwd_tile<64>(Cout / 16, X1, C1, X2, W, bias, Y, g, relu, s); break;
        case 96: rc = dispatch_fwd_tile<96>(Cout / 16, X1, C1, X2, W, bias, Y, g, relu, s); break;
        default: break;
    }
    if (rc != OMG_ERR_UNSUPPORTED) return rc;
    switch (Cout) {
        case 16: launch_fwd<1>(X1, C1, X2, C2, W, bias, Y, g, relu, s); break;
        case 32: launch_fwd<2>(X1, C1, X2, C2, W, bias, Y, g, relu, s); break;
        case 48: launch_fwd<3>(X1, C1, X2, C2, W, bias, Y, g, relu, s); break;
        case 64: launch_fwd<4>(X1, C1, X2, C2, W, bias, Y, g, relu, s); break;
        case 96: launch_fwd<6>(X1, C1, X2, C2, W, bias, Y, g, relu, s); break;
        default: OMG_REQUIRE(false, "Cout %d is not one of 16, 32, 48, 64, 96", Cout);
    }
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_conv3d_wgrad(const uint16_t* dY, int32_t Cout, const uint16_t* X1, int32_t C1, const uint16_t* X2,
                                int32_t C2, float* dW, int32_t D, int32_t H, int32_t Wd, void* stream) {
    OMG_REQUIRE(vol_ok(D, H, Wd), "bad volume size");
    OMG_REQUIRE(dY && X1 && dW && C1 % 16 == 0 && C2 % 16 == 0 && (C2 == 0 || X2), "bad arguments");
    const Vol g{D, H, Wd};
    hipStream_t s = as_stream(stream);
    const int MT = Cout / 16, NT = (C1 + C2) / 16;
    // the tiled matrix-core kernel for 16-byte aligned operands (channel-last rows of 16 B multiples)
    const bool al16 = ((uintptr_t)dY % 16 == 0) && ((uintptr_t)X1 % 16 == 0) && (C2 == 0 || (uintptr_t)X2 % 16 == 0);
    if (al16 && (MT == 1 || MT == 2 || MT == 4) && NT >= 1) {
        if (MT == 1) return launch_wgrad_tile<1>(dY, X1, C1, X2, C2, dW, g, s);
        if (MT == 2) return launch_wgrad_tile<2>(dY, X1, C1, X2, C2, dW, g, s);
        return launch_wgrad_tile<4>(dY, X1, C1, X2, C2, dW, g, s);
    }
    if (MT == 1 && NT == 1) launch_wgrad<1, 1>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 2 && NT == 1) launch_wgrad<2, 1>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 2 && NT == 2) launch_wgrad<2, 2>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 4 && NT == 2) launch_wgrad<4, 2>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 4 && NT == 4) launch_wgrad<4, 4>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 2 && NT == 6) launch_wgrad<2, 6>(dY, X1, C1, X2, C2, dW, g, s);
    else if (MT == 1 && NT == 3) launch_wgrad<1, 3>(dY, X1, C1, X2, C2, dW, g, s);
    else OMG_REQUIRE(false, "(Cout, Cin) = (%d, %d) is not a layer shape of the network", Cout, C1 + C2);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_conv3d_flipT(const uint16_t* W, uint16_t* Wf, int32_t Cout, int32_t Cin, void* stream) {
    OMG_REQUIRE(W && Wf && Cout > 0 && Cin > 0, "bad arguments");
    hipLaunchKernelGGL(k_conv_flipT, dim3(blocks((size_t)27 * Cout * Cin)), dim3(256), 0, as_stream(stream), W, Wf,
                       Cout, Cin);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_bias_grad(const uint16_t* dY, float* db, int32_t V, int32_t C, void* stream) {
    OMG_REQUIRE(dY && db && V > 0 && (C == 16 || C == 32 || C == 64), "bias gradient needs V > 0, C in {16, 32, 64}");
    if ((uintptr_t)dY % 16 == 0)
        hipLaunchKernelGGL(k_bias_grad8, dim3(512), dim3(256), 0, as_stream(stream), dY, db, V, C);
    else
        hipLaunchKernelGGL(k_bias_grad, dim3(512), dim3(256), 0, as_stream(stream), dY, db, V, C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_relu_backward(const uint16_t* Y, uint16_t* dY, int64_t n, void* stream) {
    OMG_REQUIRE(Y && dY && n >= 0, "bad arguments");
    if (n && n % 8 == 0 && (uintptr_t)Y % 16 == 0 && (uintptr_t)dY % 16 == 0)
        hipLaunchKernelGGL(k_relu_bwd8, dim3(blocks((size_t)n / 8)), dim3(256), 0, as_stream(stream),
                           reinterpret_cast<const uint4*>(Y), reinterpret_cast<uint4*>(dY), (size_t)n / 8);
    else if (n)
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks((size_t)n)), dim3(256), 0, as_stream(stream), Y, dY, (size_t)n);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_pool_forward(const uint16_t* X, uint16_t* Y, int32_t D, int32_t H, int32_t W, int32_t C,
                                void* stream) {
    OMG_REQUIRE(X && Y && C > 0 && vol_ok(D, H, W) && D % 2 == 0 && H % 2 == 0 && W % 2 == 0,
                "pooling needs even positive dimensions");
    hipLaunchKernelGGL(k_pool_fwd, dim3(blocks((size_t)D * H * W * C / 8)), dim3(256), 0, as_stream(stream), X, Y,
                       Vol{D, H, W}, C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_pool_backward(const uint16_t* dY, uint16_t* dX, int32_t D, int32_t H, int32_t W, int32_t C,
                                 int32_t accumulate, void* stream) {
    OMG_REQUIRE(dY && dX && C > 0 && vol_ok(D, H, W) && D % 2 == 0 && H % 2 == 0 && W % 2 == 0,
                "pooling needs even positive dimensions");
    if (H <= 65535 && D <= 65535 && C % 8 == 0 && (uintptr_t)dY % 16 == 0 && (uintptr_t)dX % 16 == 0)
        hipLaunchKernelGGL(k_pool_bwd8, dim3(blocks((size_t)W * (C / 8)), (unsigned)H, (unsigned)D), dim3(256), 0,
                           as_stream(stream), dY, dX, Vol{D, H, W}, C, accumulate);
    else
        hipLaunchKernelGGL(k_pool_bwd, dim3(blocks((size_t)D * H * W * C)), dim3(256), 0, as_stream(stream), dY, dX,
                           Vol{D, H, W}, C, accumulate);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_upsample_forward(const uint16_t* X, uint16_t* Y, int32_t D, int32_t H, int32_t W, int32_t C,
                                    void* stream) {
    OMG_REQUIRE(X && Y && C > 0 && vol_ok(D, H, W) && D % 2 == 0 && H % 2 == 0 && W % 2 == 0,
                "upsampling needs an even positive fine grid");
    if (H <= 65535 && D <= 65535 && C % 8 == 0 && (uintptr_t)X % 16 == 0 && (uintptr_t)Y % 16 == 0)
        hipLaunchKernelGGL(k_up_fwd8, dim3(blocks((size_t)W * (C / 8)), (unsigned)H, (unsigned)D), dim3(256), 0,
                           as_stream(stream), X, Y, Vol{D, H, W}, C);
    else
        hipLaunchKernelGGL(k_up_fwd, dim3(blocks((size_t)D * H * W * C)), dim3(256), 0, as_stream(stream), X, Y,
                           Vol{D, H, W}, C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_upsample_backward(const uint16_t* dY, int32_t Cs, int32_t off, uint16_t* dX, int32_t D, int32_t H,
                                     int32_t W, int32_t C, void* stream) {
    OMG_REQUIRE(dY && dX && vol_ok(D, H, W) && C > 0 && off >= 0 && off + C <= Cs, "bad channel window");
    if (H <= 65535 && D <= 65535 && C % 8 == 0 && Cs % 8 == 0 && off % 8 == 0 && (uintptr_t)dY % 16 == 0 &&
        (uintptr_t)dX % 16 == 0)
        hipLaunchKernelGGL(k_up_bwd8, dim3(blocks((size_t)W * (C / 8)), (unsigned)H, (unsigned)D), dim3(256), 0,
                           as_stream(stream), dY, Cs, off, dX, Vol{D, H, W}, C);
    else
        hipLaunchKernelGGL(k_up_bwd, dim3(blocks((size_t)D * H * W * C)), dim3(256), 0, as_stream(stream), dY, Cs, off,
                           dX, Vol{D, H, W}, C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_take_channels(const uint16_t* src, int32_t Cs, int32_t off, uint16_t* dst, int64_t V, int32_t C,
                                 int32_t accumulate, void* stream) {
    OMG_REQUIRE(src && dst && V > 0 && C > 0 && off >= 0 && off + C <= Cs, "bad channel window");
    if (C % 8 == 0 && Cs % 8 == 0 && off % 8 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0)
        hipLaunchKernelGGL(k_take_channels8, dim3(blocks((size_t)V * (C / 8))), dim3(256), 0, as_stream(stream), src,
                           Cs, off, dst, (size_t)V, C, accumulate);
    else
        hipLaunchKernelGGL(k_take_channels, dim3(blocks((size_t)V * C)), dim3(256), 0, as_stream(stream), src, Cs, off,
                           dst, (size_t)V, C, accumulate);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_head_forward(const uint16_t* X, const float* w, const float* b, const float* base, float* out,
                                int32_t V, int32_t C, void* stream) {
    OMG_REQUIRE(X && w && b && base && out && V > 0 && C > 0 && C < 64, "bad arguments");
    if (C % 8 == 0 && (uintptr_t)X % 16 == 0)
        hipLaunchKernelGGL(k_head_fwd8, dim3(blocks((size_t)V)), dim3(256), 0, as_stream(stream), X, w, b, base, out, V,
                           C);
    else
        hipLaunchKernelGGL(k_head_fwd, dim3(blocks((size_t)V)), dim3(256), 0, as_stream(stream), X, w, b, base, out, V,
                           C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_head_backward(const uint16_t* X, const float* w, const float* g, uint16_t* dX, float* dw, float* db,
                                 int32_t V, int32_t C, void* stream) {
    OMG_REQUIRE(X && w && g && dX && dw && db && V > 0 && C > 0 && C < 64, "bad arguments");
    if ((C == 8 || C == 16 || C == 32) && (uintptr_t)X % 16 == 0 && (uintptr_t)dX % 16 == 0)
        hipLaunchKernelGGL(k_head_bwd8, dim3(512), dim3(256), 0, as_stream(stream), X, w, g, dX, dw, db, V, C);
    else
        hipLaunchKernelGGL(k_head_bwd, dim3(blocks((size_t)V)), dim3(256), 0, as_stream(stream), X, w, g, dX, dw, db, V,
                           C);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_mse(const float* out, const float* t, const float* mask, float* g, double* loss, int32_t V,
                       float inv_n, void* stream) {
    OMG_REQUIRE(out && t && g && loss && V > 0, "bad arguments");
    const unsigned nb = (unsigned)std::min<size_t>(blocks((size_t)V), kMseBlocks);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, as_stream(stream), out, t, mask, g, loss, V, inv_n);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_adam(float* p, const float* g, float* m, float* v, uint16_t* q, int32_t n, float lr, float b1,
                        float b2, float eps, int32_t step, void* stream) {
    OMG_REQUIRE(p && g && m && v && n >= 0 && step >= 1, "bad arguments");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    if (n) hipLaunchKernelGGL(k_adam, dim3(blocks((size_t)n)), dim3(256), 0, as_stream(stream), p, g, m, v, q, n, lr,
                              b1, b2, eps, c1, c2);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_to_bf16(const float* src, uint16_t* dst, int64_t n, void* stream) {
    OMG_REQUIRE(src && dst && n >= 0, "bad arguments");
    if (n) hipLaunchKernelGGL(k_to_bf16, dim3(blocks((size_t)n)), dim3(256), 0, as_stream(stream), src, dst, (size_t)n);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

extern "C" int omg_embed(const float* img, uint16_t* X, int32_t D, int32_t H, int32_t W, int32_t Nz, float scale,
                         void* stream) {
    OMG_REQUIRE(img && X && vol_ok(D, H, W) && Nz >= 1 && Nz <= D, "bad arguments");
    if ((uintptr_t)X % 16 == 0)
        hipLaunchKernelGGL(k_embed16, dim3(blocks((size_t)D * H * W)), dim3(256), 0, as_stream(stream), img, X,
                           Vol{D, H, W}, Nz, scale);
    else
        hipLaunchKernelGGL(k_embed, dim3(blocks((size_t)D * H * W * 16)), dim3(256), 0, as_stream(stream), img, X,
                           Vol{D, H, W}, Nz, scale);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
