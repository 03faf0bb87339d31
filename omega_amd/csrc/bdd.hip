// bdd.hip -- branchless distance-driven projector (projector type 5), CT flat-panel data.
//
// Semantics of projectorType5Forward / projectorType5Backward (source/opencl/projectorType5.cl:64-337,
// :381-756): the forward projection averages the volume over each detector pixel's footprint on every
// slab (x- or y-slabs, whichever the source direction is closer to) with four taps of the slab's
// summed-area table; the backprojection is voxel-driven and takes four taps of each projection's
// summed-area table over the voxel's footprint. Summed-area tables are built here on the device
// (prefix along z / columns first, then along the fast axis), and every tap is an exact FP32 bilinear
// interpolation in normalised coordinates with clamp-to-edge, where the reference uses texture
// filtering; the slab or projection coordinate selects its table exactly.
//
// gfx950 mapping: FP is one lane per (detector row, group of 8 detector columns, projection) --
// the reference's grouping, which lets consecutive columns reuse two of the four taps -- with rows
// fastest so a wave's lanes read neighbouring texels of one slab table. BP is one lane per voxel
// with x fastest; all lanes walk the projections in the same order, so the projection's table
// (e.g. 513 x 385 floats) stays resident in L2. The per-projection back-projection geometry is
// precomputed once per geometry handle (computeGeom5, source/Python/omegatomo/projector/init.py:6-37)
// and read through scalar loads.
#include "geom.hpp"

#include <cmath>
#include <vector>

namespace omg {

struct BddArgs {
    uint32_t Nx, Ny, Nz, nRows, nCols, nProj;   // nProj: projections in this call
    float dx, dy, dz, bx, by, bz, pitchX, pitchY;
    float scaleX, scaleY, scaleZ;                // 1 / (d * (N + 1)), computeProjectorScalingValues
    const float* xyz;                            // 6 floats per projection (subset-offset)
    const float* uv;                             // 2 floats per projection (subset-offset)
    const float* geom5;                          // 16 floats per projection (subset-offset)
};

__device__ __forceinline__ float tex2(const float* __restrict__ T, uint32_t W, uint32_t H, float u, float v) {
    const float cx = u * (float)W - 0.5f, cy = v * (float)H - 0.5f;
    const float fx = floorf(cx), fy = floorf(cy);
    const float a = cx - fx, b = cy - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = min(max(x0, 0), (int)W - 1); x1 = min(max(x1, 0), (int)W - 1);
    y0 = min(max(y0, 0), (int)H - 1); y1 = min(max(y1, 0), (int)H - 1);
    const float r0 = (1.f - a) * T[(size_t)y0 * W + x0] + a * T[(size_t)y0 * W + x1];
    const float r1 = (1.f - a) * T[(size_t)y1 * W + x0] + a * T[(size_t)y1 * W + x1];
    return (1.f - b) * r0 + b * r1;
}

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// ---- summed-area tables ----------------------------------------------------------------------------
// IY: Ny slabs of (Nz+1) x (Nx+1); IX: Nx slabs of (Nz+1) x (Ny+1); first row/column zero.
__global__ __launch_bounds__(256) void k_sat_vol_z(BddArgs a, const float* __restrict__ vol, float* __restrict__ IY,
                                                   float* __restrict__ IX) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (x >= a.Nx) return;
    const size_t WX = a.Nx + 1, WY = a.Ny + 1, H = a.Nz + 1;
    float* cy = IY + (size_t)j * WX * H;
    float* cx = IX + (size_t)x * WY * H;
    cy[x + 1] = 0.f;
    cx[j + 1] = 0.f;
    if (x == 0) cy[0] = 0.f;
    if (j == 0) cx[0] = 0.f;
    float acc = 0.f;
    for (uint32_t z = 0; z < a.Nz; z++) {
        acc += vol[x + (size_t)j * a.Nx + (size_t)z * a.Nx * a.Ny];
        cy[(size_t)(z + 1) * WX + x + 1] = acc;
        cx[(size_t)(z + 1) * WY + j + 1] = acc;
    }
    if (x == 0)
        for (uint32_t z = 1; z <= a.Nz; z++) cy[(size_t)z * WX] = 0.f;
    if (j == 0)
        for (uint32_t z = 1; z <= a.Nz; z++) cx[(size_t)z * WY] = 0.f;
}

// prefix along the fast axis of each row of `rows` tables of W x H (row 0 untouched)
__global__ __launch_bounds__(256) void k_sat_rows(float* __restrict__ T, uint32_t W, uint32_t H, uint32_t tables) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= (H - 1) * tables) return;
    const uint32_t tab = t / (H - 1), row = t - tab * (H - 1) + 1;
    float* p = T + (size_t)tab * W * H + (size_t)row * W;
    float acc = 0.f;
    for (uint32_t i = 1; i < W; i++) {
        acc += p[i];
        p[i] = acc;
    }
}

// sinogram tables: nProj tables of (nCols+1) x (nRows+1); prefix along columns per row first
__global__ __launch_bounds__(256) void k_sat_sino_cols(BddArgs a, const float* __restrict__ y, float* __restrict__ S) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    if (r >= a.nRows) return;
    const size_t W = a.nRows + 1, H = a.nCols + 1;
    float* T = S + (size_t)k * W * H;
    T[r + 1] = 0.f;
    if (r == 0) {
        T[0] = 0.f;
        for (uint32_t c = 1; c <= a.nCols; c++) T[(size_t)c * W] = 0.f;
    }
    float acc = 0.f;
    for (uint32_t c = 0; c < a.nCols; c++) {
        acc += y[r + (size_t)c * a.nRows + (size_t)k * a.nRows * a.nCols];
        T[(size_t)(c + 1) * W + r + 1] = acc;
    }
}

// ---- forward projection ----------------------------------------------------------------------------
constexpr int kCols = 8;

__global__ __launch_bounds__(256) void k_bdd_fp(BddArgs a, const float* __restrict__ IY, const float* __restrict__ IX,
                                                float* __restrict__ out) {
    const uint32_t groups = (a.nCols + kCols - 1) / kCols;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)a.nRows * groups * a.nProj) return;
    const uint32_t r = (uint32_t)(t % a.nRows);
    const uint32_t c0 = (uint32_t)((t / a.nRows) % groups) * kCols;
    const uint32_t k = (uint32_t)(t / ((uint64_t)a.nRows * groups));
    const float* P = a.xyz + (size_t)k * 6;
    const float sx = P[0], sy = P[1], sz = P[2];
    float dxp = P[3], dyp = P[4], dzp = P[5];
    const float i1 = (float)r - (float)a.nRows / 2.f + .5f, i2 = (float)c0 - (float)a.nCols / 2.f + .5f;
    const float ux = a.uv[(size_t)k * 2], uy = a.uv[(size_t)k * 2 + 1];
    dxp += i1 * ux;
    dyp += i1 * uy;
    dzp += i2 * a.pitchY;
    const float dRx = dxp - ux * 0.5f, dRy = dyp - uy * 0.5f;
    const float dLx = dxp + ux * 0.5f, dLy = dyp + uy * 0.5f;
    const float dUz = dzp + a.pitchY * 0.5f, dDz = dzp - a.pitchY * 0.5f;
    const float ex = dxp - sx, ey = dyp - sy, ez = dzp - sz;
    const float inv = 1.f / len3(ex, ey, ez);
    float kerroin = a.dx * a.dy * a.dz;
    const float apuZ = (a.bz - sz - a.dz / 2.f);
    const int maxZZ = min(kCols, (int)a.nCols - (int)c0);
    float temp[kCols];
#pragma unroll
    for (int q = 0; q < kCols; q++) temp[q] = 0.f;
    const bool ydir = fabsf(sx) <= fabsf(sy);
    // slab axis: y-slabs sample IY in (x, z); x-slabs sample IX in (y, z)
    const float L0 = ydir ? (dLx - sx) / (dLy - sy) : (dLy - sy) / (dLx - sx);
    const float L1 = ydir ? (dRx - sx) / (dRy - sy) : (dRy - sy) / (dRx - sx);
    const float Z0 = ydir ? (dUz - sz) / (dyp - sy) : (dUz - sz) / (dxp - sx);
    const float Z1 = ydir ? (dDz - sz) / (dyp - sy) : (dDz - sz) / (dxp - sx);
    kerroin /= ydir ? fabsf(ey * inv) : fabsf(ex * inv);
    const float apuT = ydir ? (a.bx - sx - a.dx / 2.f) : (a.by - sy - a.dy / 2.f);
    const float apuV = ydir ? a.by - sy : a.bx - sx;
    const float dslab = ydir ? a.dy : a.dx;
    const float scaleT = ydir ? a.scaleX : a.scaleY;
    const uint32_t W = ydir ? a.Nx + 1 : a.Ny + 1, H = a.Nz + 1;
    const uint32_t nslab = ydir ? a.Ny : a.Nx;
    const float* tabs = ydir ? IY : IX;
    float half = dslab / 2.f;
    for (uint32_t jj = 0; jj < nslab; jj++) {
        const float v = half + apuV;
        half += dslab;
        const float a0 = v * L0 - apuT, a1 = v * L1 - apuT;
        const float b0 = v * Z0 - apuZ, b1 = v * Z1 - apuZ;
        float tl = fminf(a0, a1), tr = fmaxf(a0, a1);
        float zu = fminf(b0, b1), zd = fmaxf(b0, b1);
        const float invArea = 1.f / fabsf((tl - tr) * (zu - zd));
        tl *= scaleT; tr *= scaleT; zu *= a.scaleZ; zd *= a.scaleZ;
        const float* T = tabs + (size_t)jj * W * H;
        float A = tex2(T, W, H, tr, zu), B = tex2(T, W, H, tl, zd);
        float C = tex2(T, W, H, tr, zd), D = tex2(T, W, H, tl, zu);
        temp[0] += (C + D - A - B) * invArea;
        const float step = fabsf(zu - zd);
#pragma unroll
        for (int q = 1; q < kCols; q++) {
            if (q < maxZZ) {
                D = B; A = C;
                zd += step;
                B = tex2(T, W, H, tl, zd);
                C = tex2(T, W, H, tr, zd);
                temp[q] += (C + D - A - B) * invArea;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kCols; q++)
        if (q < maxZZ) out[r + (size_t)(c0 + q) * a.nRows + (size_t)k * a.nRows * a.nCols] = temp[q] * kerroin;
}

// ---- backprojection (voxel-driven) -------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bdd_bp(BddArgs a, const float* __restrict__ S, float* __restrict__ vol,
                                                float* __restrict__ sens) {
    const uint32_t n = blockIdx.x * 256 + threadIdx.x;
    const uint32_t Nxy = a.Nx * a.Ny;
    if (n >= Nxy * a.Nz) return;
    const uint32_t kz = n / Nxy, rem = n - kz * Nxy, j = rem / a.Nx, i = rem - j * a.Nx;
    const float vx = (float)i * a.dx + a.dx / 2.f + a.bx;
    const float vy = (float)j * a.dy + a.dy / 2.f + a.by;
    const float vz = (float)kz * a.dz + a.dz / 2.f + a.bz;
    const float kokoX = (float)(a.nRows + 1) * a.pitchX, kokoY = (float)(a.nCols + 1) * a.pitchY;
    const float invKx = 1.f / kokoX, invKy = 1.f / kokoY;
    const uint32_t W = a.nRows + 1, H = a.nCols + 1;
    float temp = 0.f, ws = 0.f;
    for (uint32_t kk = 0; kk < a.nProj; kk++) {
        const float* o = a.geom5 + (size_t)kk * 16;
        const float s0 = o[0], s1 = o[1], s2 = o[2];
        const bool sxDim = fabsf(s0) <= fabsf(s1);
        float lu0, lu1, rd0, rd1;
        if (sxDim) {
            const bool pos = s1 > 0.f;
            lu0 = pos ? vx - a.dx / 2.f : vx + a.dx / 2.f;
            rd0 = pos ? vx + a.dx / 2.f : vx - a.dx / 2.f;
            lu1 = vy; rd1 = vy;
        } else {
            const bool pos = s0 > 0.f;
            lu1 = pos ? vy + a.dy / 2.f : vy - a.dy / 2.f;
            rd1 = pos ? vy - a.dy / 2.f : vy + a.dy / 2.f;
            lu0 = vx; rd0 = vx;
        }
        float lu2 = vz - a.dz / 2.f, rd2 = vz + a.dz / 2.f;
        lu0 -= s0; lu1 -= s1; lu2 -= s2;
        rd0 -= s0; rd1 -= s1; rd2 -= s2;
        const float tLU = o[15] / (-lu0 * o[12] + -lu1 * o[13] + -lu2 * o[14]);
        const float tRD = o[15] / (-rd0 * o[12] + -rd1 * o[13] + -rd2 * o[14]);
        const float pl0 = lu0 * tLU + s0 - o[3], pl1 = lu1 * tLU + s1 - o[4], pl2 = lu2 * tLU + s2 - o[5];
        const float pr0 = rd0 * tRD + s0 - o[3], pr1 = rd1 * tRD + s1 - o[4], pr2 = rd2 * tRD + s2 - o[5];
        float Ax = pl0 * o[6] + pl1 * o[7] + pl2 * o[8] + a.pitchX / 2.f;
        float Ay = pl0 * o[9] + pl1 * o[10] + pl2 * o[11] + a.pitchY / 2.f;
        float Dx = pr0 * o[6] + pr1 * o[7] + pr2 * o[8] + a.pitchX / 2.f;
        float Dy = pr0 * o[9] + pr1 * o[10] + pr2 * o[11] + a.pitchY / 2.f;
        if (Dx > Ax) { const float t = Ax; Ax = Dx; Dx = t; }
        if (Dy > Ay) { const float t = Ay; Ay = Dy; Dy = t; }
        const float* T = S + (size_t)kk * W * H;
        const float AxN = Ax * invKx, DxN = Dx * invKx, AyN = Ay * invKy, DyN = Dy * invKy;
        const float A = tex2(T, W, H, AxN, AyN), B = tex2(T, W, H, AxN, DyN);
        const float C = tex2(T, W, H, DxN, AyN), D = tex2(T, W, H, DxN, DyN);
        const float vdx = vx - s0, vdy = vy - s1, vdz = vz - s2;
        const float kerroin = sxDim ? a.dy * len3(vdx, vdy, vdz) / fabsf(vdy) : a.dx * len3(vdx, vdy, vdz) / fabsf(vdx);
        const float apu = A + D - C - B;
        if (apu != 0.f) {
            temp += apu * kerroin;
            ws += kerroin;
        }
    }
    vol[n] += temp;
    if (sens) sens[n] += ws;
}

static BddArgs make_bdd_args(const omg_geom* g, int64_t first, uint32_t nproj) {
    BddArgs a;
    a.Nx = g->d.Nx; a.Ny = g->d.Ny; a.Nz = g->d.Nz;
    a.nRows = g->d.size_x; a.nCols = g->d.size_y; a.nProj = nproj;
    a.dx = g->d.dx; a.dy = g->d.dy; a.dz = g->d.dz;
    a.bx = g->d.bx; a.by = g->d.by; a.bz = g->d.bz;
    a.pitchX = g->d.dPitchX; a.pitchY = g->d.dPitchY;
    a.scaleX = 1.f / (g->d.dx * (float)(g->d.Nx + 1));
    a.scaleY = 1.f / (g->d.dy * (float)(g->d.Ny + 1));
    a.scaleZ = 1.f / (g->d.dz * (float)(g->d.Nz + 1));
    a.xyz = g->x_dev + first * 6;
    a.uv = g->z_dev + first * 2;
    a.geom5 = g->bdd_geom5 + first * 16;
    return a;
}

}  // namespace omg

using namespace omg;

// Host: per-projection back-projection geometry, computeGeom5 (init.py:6-37) in float arithmetic.
int bdd_prepare(omg_geom* g) {
    OMG_REQUIRE(!g->d.ct_pitch, "projector_type 5 supports the uV (2 floats per projection) panel description");
    OMG_REQUIRE(g->d.dPitchX > 0.f && g->d.dPitchY > 0.f, "projector_type 5 needs dPitchX and dPitchY");
    const uint32_t P = g->d.n_proj;
    std::vector<float> x(P * 6), uv(P * 2), out((size_t)P * 16);
    OMG_CHECK_HIP(hipMemcpy(x.data(), g->x_dev, x.size() * sizeof(float), hipMemcpyDeviceToHost));
    OMG_CHECK_HIP(hipMemcpy(uv.data(), g->z_dev, uv.size() * sizeof(float), hipMemcpyDeviceToHost));
    const float indX = (float)g->d.size_x / 2.f, indY = (float)g->d.size_y / 2.f;
    for (uint32_t k = 0; k < P; k++) {
        const float* p = &x[(size_t)k * 6];
        const float s[3] = {p[0], p[1], p[2]}, d[3] = {p[3], p[4], p[5]};
        const float ax[3] = {uv[2 * k] * indX, uv[2 * k + 1] * indX, 0.f};
        const float ay[3] = {0.f, 0.f, indY * g->d.dPitchY};
        float d3[3], d2[3], e[3];
        for (int c = 0; c < 3; c++) { d3[c] = d[c] - ax[c] - ay[c]; d2[c] = ax[c] - ay[c]; }
        for (int c = 0; c < 3; c++) e[c] = d3[c] - d[c];
        const float lx = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        const float ly = std::sqrt(ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2]);
        const float cp[3] = {d2[1] * e[2] - d2[2] * e[1], d2[2] * e[0] - d2[0] * e[2], d2[0] * e[1] - d2[1] * e[0]};
        float* o = &out[(size_t)k * 16];
        for (int c = 0; c < 3; c++) {
            o[c] = s[c]; o[3 + c] = d3[c]; o[6 + c] = ax[c] / lx; o[9 + c] = ay[c] / ly; o[12 + c] = cp[c];
        }
        o[15] = cp[0] * (s[0] - d[0]) + cp[1] * (s[1] - d[1]) + cp[2] * (s[2] - d[2]);
    }
    OMG_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(&g->bdd_geom5), out.size() * sizeof(float)));
    OMG_CHECK_HIP(hipMemcpy(g->bdd_geom5, out.data(), out.size() * sizeof(float), hipMemcpyHostToDevice));
    return OMG_OK;
}

void bdd_release(omg_geom* g) {
    hipFree(g->bdd_geom5);
    g->bdd_geom5 = nullptr;
}

int bdd_forward(const omg_geom* g, int64_t first, int64_t n, const float* img, float* y, hipStream_t s) {
    const int64_t per = (int64_t)g->d.size_x * g->d.size_y;
    const uint32_t nproj = (uint32_t)(n / per);
    BddArgs a = make_bdd_args(g, first, nproj);
    const size_t WX = a.Nx + 1, WY = a.Ny + 1, H = a.Nz + 1;
    const size_t need = WX * H * a.Ny + WY * H * a.Nx;
    int rc = scratch_reserve(const_cast<omg_geom*>(g), need);
    if (rc) return rc;
    float* IY = scratch_ptr(g);
    float* IX = scratch_ptr(g) + WX * H * a.Ny;
    hipLaunchKernelGGL(k_sat_vol_z, dim3(grid_for(a.Nx, 256), a.Ny), dim3(256), 0, s, a, img, IY, IX);
    OMG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sat_rows, dim3(grid_for((int64_t)(H - 1) * a.Ny, 256)), dim3(256), 0, s, IY, (uint32_t)WX,
                       (uint32_t)H, a.Ny);
    OMG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sat_rows, dim3(grid_for((int64_t)(H - 1) * a.Nx, 256)), dim3(256), 0, s, IX, (uint32_t)WY,
                       (uint32_t)H, a.Nx);
    OMG_CHECK_LAUNCH();
    const uint64_t threads = (uint64_t)a.nRows * ((a.nCols + kCols - 1) / kCols) * nproj;
    hipLaunchKernelGGL(k_bdd_fp, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, a, IY, IX, y);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}

int bdd_backward(const omg_geom* g, int64_t first, int64_t n, const float* yin, float* rhs, float* sens, hipStream_t s) {
    const int64_t per = (int64_t)g->d.size_x * g->d.size_y;
    const uint32_t nproj = (uint32_t)(n / per);
    BddArgs a = make_bdd_args(g, first, nproj);
    const size_t W = a.nRows + 1, H = a.nCols + 1;
    int rc = scratch_reserve(const_cast<omg_geom*>(g), W * H * nproj);
    if (rc) return rc;
    float* S = scratch_ptr(g);
    hipLaunchKernelGGL(k_sat_sino_cols, dim3(grid_for(a.nRows, 256), nproj), dim3(256), 0, s, a, yin, S);
    OMG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sat_rows, dim3(grid_for((int64_t)(H - 1) * nproj, 256)), dim3(256), 0, s, S, (uint32_t)W,
                       (uint32_t)H, nproj);
    OMG_CHECK_LAUNCH();
    const uint64_t N = (uint64_t)a.Nx * a.Ny * a.Nz;
    hipLaunchKernelGGL(k_bdd_bp, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a, S, rhs, sens);
    OMG_CHECK_LAUNCH();
    return OMG_OK;
}
