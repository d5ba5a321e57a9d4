// Stage (a) and the run-length half of (b): decimation, blur/sharpen,
// 4x4 tile min/max, the adaptive threshold and the row runs.
//
// Every launch is 2D: blockIdx.y selects the frame of a batch, so one
// launch serves frames of different sizes (blocks past a frame's work
// exit at once). All work here is byte arithmetic, HBM-bound: rows are
// read coalesced (one 64-lane wave per row segment), the threshold's
// dilated tile row is staged in LDS, and the run extraction uses wave
// ballots instead of a per-pixel scan.
#include <hip/hip_runtime.h>

#include "atg_internal.h"

namespace atg {

// ---------------------------------------------------------------- decimate

// One pixel of the quad image (fast/image.c:26-46 x2 with the double
// rounding for the first qw - qw % 32 columns, :50-62 x1.5, :83-87
// integer point sampling; decimate <= 1 copies, fast/detector.c:283).
__device__ __forceinline__ uint8_t dec_px(const Frame &F, float f, int x, int y) {
    const uint8_t *src = F.img;
    int64_t s = F.stride;
    if (!(f > 1.0f)) return src[(int64_t)y * s + x];
    if (f == 1.5f) {
        const uint8_t *r0 = src + (int64_t)(y / 2) * 3 * s, *r1 = r0 + s, *r2 = r1 + s;
        int k = x >> 1, c3 = 3 * k;
        int a = r0[c3], b = r0[c3 + 1], c = r0[c3 + 2];
        int d = r1[c3], e = r1[c3 + 1], ff = r1[c3 + 2];
        if ((y & 1) == 0) return (uint8_t)((x & 1) ? (4 * c + 2 * b + 2 * ff + e) / 9 : (4 * a + 2 * b + 2 * d + e) / 9);
        int g = r2[c3], h = r2[c3 + 1], i = r2[c3 + 2];
        return (uint8_t)((x & 1) ? (4 * i + 2 * ff + 2 * h + e) / 9 : (4 * g + 2 * d + 2 * h + e) / 9);
    }
    int k = (int)f;
    if (k == 2) {
        const uint8_t *a = src + (int64_t)(2 * y) * s + 2 * x, *b = a + s;
        int p = a[0], q = a[1], r = b[0], t = b[1];
        int nv = F.qw - F.qw % 32;
        if (x < nv) return (uint8_t)((((p + r + 1) >> 1) + ((q + t + 1) >> 1) + 1) >> 1);
        return (uint8_t)((p + q + r + t + 2) >> 2);
    }
    return src[(int64_t)y * k * s + (int64_t)x * k];
}

// One thread per 4x4 tile slot (partial slots at the right and bottom
// edges hold quad-image pixels that belong to no tile). Writes the quad
// image and, unless a blur follows, the tile max/min (fast/threshold.c:15-33).
__global__ void k_decimate_tiles(const Frame *frames, const DevParams *P, int with_tiles) {
    const Frame F = frames[blockIdx.y];
    int sw = (F.qw + 3) >> 2, sh = (F.qh + 3) >> 2;
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= sw * sh) return;
    int tx = slot % sw, ty = slot / sw;
    float f = P->decimate;
    uint8_t hi = 0, lo = 255;
    for (int dy = 0; dy < 4; dy++) {
        int y = 4 * ty + dy;
        if (y >= F.qh) break;
        uint32_t word = 0;
        int nx = min(4, F.qw - 4 * tx);
        for (int dx = 0; dx < nx; dx++) {
            uint8_t v = dec_px(F, f, 4 * tx + dx, y);
            word |= (uint32_t)v << (8 * dx);
            hi = max(hi, v);
            lo = min(lo, v);
        }
        uint8_t *o = F.qim + (int64_t)y * F.qw + 4 * tx;
        if (nx == 4 && ((reinterpret_cast<uintptr_t>(o) & 3) == 0)) {
            *reinterpret_cast<uint32_t *>(o) = word;
        } else {
            for (int dx = 0; dx < nx; dx++) o[dx] = (uint8_t)(word >> (8 * dx));
        }
    }
    if (with_tiles && tx < F.tw && ty < F.th) {
        F.tmx[ty * F.tw + tx] = hi;
        F.tmn[ty * F.tw + tx] = lo;
    }
}

// Tile max/min of the quad image (after a blur).
__global__ void k_tiles(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.tw * F.th) return;
    int tx = i % F.tw, ty = i / F.tw;
    uint8_t hi = 0, lo = 255;
    for (int dy = 0; dy < 4; dy++)
        for (int dx = 0; dx < 4; dx++) {
            uint8_t v = F.qim[(int64_t)(4 * ty + dy) * F.qw + 4 * tx + dx];
            hi = max(hi, v);
            lo = min(lo, v);
        }
    F.tmx[i] = hi;
    F.tmn[i] = lo;
}

// ---------------------------------------------------------------- blur

// One output of the C library's 1D convolution (fast/image.c:93-108):
// copies at the head and the tail, the kernel in between; `par` is the
// variant for images of 65536 pixels or more (one more output).
__device__ __forceinline__ uint8_t conv_at(const uint8_t *x, int64_t step, int o, int sz, const uint8_t *k, int ksz,
                                           bool par) {
    int half = ksz / 2;
    int n = par ? sz - ksz + 1 : sz - ksz;
    if (o < half || o - half >= n) return x[(int64_t)o * step];
    int i = o - half;
    uint32_t acc = 0;
    for (int j = 0; j < ksz; j++) acc += (uint32_t)k[j] * x[(int64_t)(i + j) * step];
    return (uint8_t)(acc >> 8);
}

// Row pass of gaussian_blur (fast/image.c:128-132): qim -> tbuf.
__global__ void k_blur_rows(const Frame *frames, const DevParams *P) {
    const Frame F = frames[blockIdx.y];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)F.qw * F.qh) return;
    int x = (int)(i % F.qw), y = (int)(i / F.qw);
    bool par = (int64_t)F.qw * F.qh >= 65536;
    F.tbuf[i] = conv_at(F.qim + (int64_t)y * F.qw, 1, x, F.qw, P->blur_k, P->blur_ksz, par);
}

// Column pass (fast/image.c:133-137), then for a negative sigma the
// sharpen clamp(2 orig - blurred) (fast/image.c:149-154): tbuf -> qim.
__global__ void k_blur_cols(const Frame *frames, const DevParams *P) {
    const Frame F = frames[blockIdx.y];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)F.qw * F.qh) return;
    int x = (int)(i % F.qw), y = (int)(i / F.qw);
    bool par = (int64_t)F.qw * F.qh >= 65536;
    uint8_t b = conv_at(F.tbuf + x, F.qw, y, F.qh, P->blur_k, P->blur_ksz, par);
    if (P->sigma < 0.0f) {
        int v = 2 * (int)F.qim[i] - b;
        b = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
    F.qim[i] = b;
}

// ---------------------------------------------------------------- threshold + runs

// The runs of one thresholded row, one wave: changes as ballots over 64
// pixels, a run written by the lane at its end (fast/threshold.c:213-266:
// columns [0, qw-2], runs of 127 dropped). tfun(x) gives t[x].
template <class TF>
__device__ __forceinline__ void row_runs(const Frame &F, int y, TF tfun) {
    const int lane = threadIdx.x & 63;
    const int xmax = F.qw - 2;
    uint32_t *out = F.runs + (int64_t)y * F.qw;
    int carry_start = 0, count = 0;
    uint32_t carry_t = 0;
    for (int x0 = 0; x0 <= xmax; x0 += 64) {
        int x = x0 + lane;
        bool valid = x <= xmax;
        uint32_t tv = valid ? tfun(x) : 0u;
        const uint32_t tp = wave_from_left(tv, carry_t);
        uint32_t fill = 0;
        if (lane == 63 && x + 1 <= xmax) fill = tfun(x + 1);
        const uint32_t tn = wave_from_right(tv, fill);
        bool is_start = valid && (x == 0 || tv != tp);
        bool is_end = valid && (x == xmax || tn != tv);
        uint64_t smask = __ballot(is_start);
        bool store = is_end && tv != 127u;
        uint64_t emask = __ballot(store);
        uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        if (store) {
            uint64_t m = smask & le;
            int start = m ? x0 + 63 - __clzll(m) : carry_start;
            int idx = count + __popcll(emask & (le >> 1));
            out[idx] = (uint32_t)start | ((uint32_t)x << 15) | ((tv == 255u ? 1u : 0u) << 30);
        }
        if (smask) carry_start = x0 + 63 - __clzll(smask);
        count += __popcll(emask);
        carry_t = (uint32_t)__builtin_amdgcn_readlane((int)tv, 63);
    }
    if (lane == 0) F.rcnt[y] = (uint32_t)count;
}

// One block (4 waves) per `trows` tile rows: the 3x3 max/min of the tile
// grid around each tile of those rows (clipped at the grid edge) into LDS,
// then each wave thresholds pixel rows (fast/threshold.c:71-150) and, with
// do_rle, extracts their runs. The rows below the last whole tile row use
// tile row th-1 and are never flat (fast/threshold.c:144-148).
// Eight waves a SIMD (64 VGPRs, two of them spilled to scratch): the rows'
// loads in flight cover more of the latency than at six (80 VGPRs, no
// spills): 0.56 -> 0.49 ms for 256 frames of 1080p.
// Each wave first copies its pixel row into LDS (rboff: the row buffers'
// offset, rbstride bytes a wave), its loads issued eight steps at a time:
// the threshold and the run extraction then read the row from LDS.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_threshold(const Frame *frames, const DevParams *P, int do_rle, int trows,
                                                                                          int rboff, int rbstride) {
    extern __shared__ uint8_t lds[];
    const Frame F = frames[blockIdx.y];
    const int t0 = (int)blockIdx.x * trows;
    if (t0 > F.th) return;
    const int t1 = min(t0 + trows, F.th + 1);  // tile rows [t0, t1), th = the tail
    const int tw = F.tw;
    for (int j = 0; j < t1 - t0; j++) {
        const int tyd = min(t0 + j, F.th - 1);
        uint8_t *hx = lds + 2 * tw * j, *hn = hx + tw;
        const int up = max(tyd - 1, 0), dn = min(tyd + 1, F.th - 1);
        for (int tx = threadIdx.x; tx < tw; tx += blockDim.x) {
            int l = max(tx - 1, 0), r = min(tx + 1, tw - 1);
            uint8_t hi = 0, lo = 255;
            for (int yy = up; yy <= dn; yy++)
                for (int xx = l; xx <= r; xx++) {
                    hi = max(hi, F.tmx[yy * tw + xx]);
                    lo = min(lo, F.tmn[yy * tw + xx]);
                }
            hx[tx] = hi;
            hn[tx] = lo;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y1 = min(4 * t1, F.qh);
    const int diff = P->mwbd;
    for (int y = 4 * t0 + wave; y < y1; y += 4) {
        const int ty = y >> 2;
        const bool tail = ty == F.th;
        const uint8_t *hx = lds + 2 * tw * (ty - t0), *hn = hx + tw;
        const uint8_t *grow = F.qim + (int64_t)y * F.qw;
        uint8_t *trow = F.t + (int64_t)y * F.qw;
        uint8_t *row = lds + rboff + wave * rbstride;
        for (int x0 = 0; x0 < F.qw; x0 += 512) {
            uint32_t v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int x = x0 + 64 * q + lane;
                v[q] = x < F.qw ? grow[x] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int x = x0 + 64 * q + lane;
                if (x < F.qw) row[x] = (uint8_t)v[q];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        auto tfun = [&](int x) -> uint32_t {
            bool xin = x < 4 * tw;
            int tx = xin ? x >> 2 : tw - 1;
            int hi = hx[tx], lo = hn[tx];
            int thr = lo + (hi - lo) / 2;
            bool flat = !tail && xin && hi - lo < diff;
            return flat ? 127u : (row[x] > thr ? 255u : 0u);
        };
        for (int x = lane; x < F.qw; x += 64) trow[x] = (uint8_t)tfun(x);
        if (do_rle) row_runs(F, y, tfun);
    }
}

// Runs of the thresholded image (after deglitch), one wave per row.
__global__ __launch_bounds__(64) void k_rle(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int y = blockIdx.x;
    if (y >= F.qh) return;
    const uint8_t *trow = F.t + (int64_t)y * F.qw;
    row_runs(F, y, [&](int x) -> uint32_t { return trow[x]; });
}

// deglitch (fast/threshold.c:173-193): 3x3 max into tbuf (border 0), then
// the 3x3 min of tbuf back into the interior of t.
__global__ void k_deglitch_max(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)F.qw * F.qh) return;
    int x = (int)(i % F.qw), y = (int)(i / F.qw);
    uint8_t m = 0;
    if (x > 0 && y > 0 && x < F.qw - 1 && y < F.qh - 1)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) m = max(m, F.t[(int64_t)(y + dy) * F.qw + x + dx]);
    F.tbuf[i] = m;
}

__global__ void k_deglitch_min(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)F.qw * F.qh) return;
    int x = (int)(i % F.qw), y = (int)(i / F.qw);
    if (!(x > 0 && y > 0 && x < F.qw - 1 && y < F.qh - 1)) return;
    uint8_t m = 255;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) m = min(m, F.tbuf[(int64_t)(y + dy) * F.qw + x + dx]);
    F.t[i] = m;
}

// A plain copy (16 bytes a thread, grid-stride): the measurements' known-size
// reference for HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(uint4 *dst, const uint4 *src, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy(hipStream_t st, void *dst, const void *src, int64_t bytes) {
    int64_t n16 = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, static_cast<uint4 *>(dst), static_cast<const uint4 *>(src),
                       n16);
}

// ---------------------------------------------------------------- launchers

static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void launch_image_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                        int max_qh, bool blur, bool deglitch, const Tune &tu, Marks *mk) {
    int64_t maxpx = (int64_t)max_qw * max_qh;
    int64_t slots = (int64_t)((max_qw + 3) / 4) * ((max_qh + 3) / 4);
    dim3 g1(cdiv(slots, 256), nframes);
    hipLaunchKernelGGL(k_decimate_tiles, g1, dim3(256), 0, st, dframes, dparams, blur ? 0 : 1);
    mk->mark(st, "decimate");
    if (blur) {
        dim3 gp(cdiv(maxpx, 256), nframes);
        hipLaunchKernelGGL(k_blur_rows, gp, dim3(256), 0, st, dframes, dparams);
        hipLaunchKernelGGL(k_blur_cols, gp, dim3(256), 0, st, dframes, dparams);
        dim3 gt(cdiv((int64_t)(max_qw / 4) * (max_qh / 4), 256), nframes);
        hipLaunchKernelGGL(k_tiles, gt, dim3(256), 0, st, dframes);
        mk->mark(st, "blur");
    }
    int max_th = max_qh / 4, trows = tu.thr_tile_rows;
    // the tile extrema of trows tile rows, then a pixel row a wave
    const int rboff = (int)(((size_t)2 * (max_qw / 4) * trows + 16 + 15) & ~(size_t)15);
    const int rbstride = (max_qw + 15) & ~15;
    size_t lds = (size_t)rboff + (size_t)4 * rbstride;
    dim3 g2(cdiv(max_th + 1, trows), nframes);
    hipLaunchKernelGGL(k_threshold, g2, dim3(256), lds, st, dframes, dparams, deglitch ? 0 : 1, trows, rboff, rbstride);
    if (deglitch) {
        dim3 gp(cdiv(maxpx, 256), nframes);
        hipLaunchKernelGGL(k_deglitch_max, gp, dim3(256), 0, st, dframes);
        hipLaunchKernelGGL(k_deglitch_min, gp, dim3(256), 0, st, dframes);
        hipLaunchKernelGGL(k_rle, dim3(max_qh, nframes), dim3(64), 0, st, dframes);
    }
    mk->mark(st, "threshold_rle");
}

__global__ void k_stamp(uint64_t *slot) {
    if (threadIdx.x == 0) *slot = wall_clock64();
}

void launch_stamp(hipStream_t st, uint64_t *slot) { hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, slot); }

}  // namespace atg
