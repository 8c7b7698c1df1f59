// k_cache.hip — radiance-cache network kernels.
//
//   kCacheInfer  fully fused forward pass: weights in LDS, each wave runs 32
//                queries at a time through all six layers in registers (MFMA 16x16x16
//                FP16, FP32 accumulate); writes y or adds beta * max(y, 0) to
//                the path's result
//   kCacheTrain  fused forward + backward for 128 queries per block: layer
//                inputs kept in LDS ([feature][query] FP16), loss gradient of
//                the relative L2 loss, weight gradients as MFMA products over
//                the block's queries, written as one partial per block
//   kCacheAdam   sums the block partials in a fixed order (eight runs of
//                consecutive blocks, then the runs; deterministic), then Adam,
//                FP32 master weights and the two FP16 layouts
#include "cache.hpp"
#include "launch_cache.hpp"

namespace prr {

constexpr int kCacheBlock = 256;
constexpr int kTrainQ = 128;       // queries per training block (4 waves x 32)
constexpr int kXsStride = 132;     // halves per LDS row: 128 queries + 4 pad
constexpr int kXsRows = kCacheIn + 5 * kCacheH;
PRR_HD constexpr int xsRow(int l) { return l == 0 ? 0 : kCacheIn + (l - 1) * kCacheH; }

PRR_DEV void loadQuery(const float *__restrict__ q, int64_t qi, int64_t n, float *qq) {
    if (qi < n) {
#pragma unroll
        for (int k = 0; k < kCacheQ; k++) qq[k] = q[qi * kCacheQ + k];
    } else {
#pragma unroll
        for (int k = 0; k < kCacheQ; k++) qq[k] = 0.f;
        qq[4] = 1.f;
        qq[7] = 1.f;
    }
}

PRR_DEV h4 ldh4(const _Float16 *p) { return *(const h4 *) p; }

// One hidden layer (64 outputs) on QB query blocks: x[qb][0..KB) -> x[qb][0..4).
template <int QB, int KB>
PRR_DEV void hiddenLayer(const _Float16 *wT, int stride, int g, int c, h4 (&x)[QB][6]) {
    f4 acc[QB][4];
#pragma unroll
    for (int qb = 0; qb < QB; qb++)
#pragma unroll
        for (int ob = 0; ob < 4; ob++) acc[qb][ob] = f4{0.f, 0.f, 0.f, 0.f};
    static_assert(KB % 2 == 0, "K-blocks are taken in pairs");
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
        for (int kb = 0; kb < KB; kb += 2) {
            const _Float16 *row = wT + (ob * 16 + c) * stride + 4 * g;
            const h8 a = cat8(ldh4(row + kb * 16), ldh4(row + kb * 16 + 16));
#pragma unroll
            for (int qb = 0; qb < QB; qb++) acc[qb][ob] = mfma32(a, cat8(x[qb][kb], x[qb][kb + 1]), acc[qb][ob]);
        }
#pragma unroll
    for (int qb = 0; qb < QB; qb++)
#pragma unroll
        for (int ob = 0; ob < 4; ob++) x[qb][ob] = reluHalf(acc[qb][ob]);
}

template <int QB>
PRR_DEV void outputLayer(const _Float16 *wT, int g, int c, const h4 (&x)[QB][6], f4 (&y)[QB]) {
#pragma unroll
    for (int qb = 0; qb < QB; qb++) y[qb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb += 2) {
        const _Float16 *row = wT + c * cacheStrideT(5) + 4 * g;
        const h8 a = cat8(ldh4(row + kb * 16), ldh4(row + kb * 16 + 16));
#pragma unroll
        for (int qb = 0; qb < QB; qb++) y[qb] = mfma32(a, cat8(x[qb][kb], x[qb][kb + 1]), y[qb]);
    }
}

// ---------------------------------------------------------------------------
constexpr int kInferQB = 2;  // query blocks of 16 per wave and trip

// With result != null the outputs are path contributions: result[slot] +=
// beta * max(y, 0), rescaled to mean clampInd when brighter (the renderer's
// indirect clamp); n is then read from nDev when nDev != null (capped by n).
__global__ __launch_bounds__(kCacheBlock) void kCacheInfer(CacheView cv, int64_t n, const uint32_t *__restrict__ nDev,
                                                           const float *__restrict__ q, float *__restrict__ out,
                                                           const uint32_t *__restrict__ slot,
                                                           const float4 *__restrict__ beta,
                                                           float4 *__restrict__ result, float clampInd) {
    if (nDev) n = min(n, (int64_t) *nDev);
    __shared__ __align__(16) _Float16 w[kCacheHalfT];
    {
        const uint4 *src = (const uint4 *) cv.wt;
        uint4 *dst = (uint4 *) w;
        for (int i = threadIdx.x; i < kCacheHalfT / 8; i += kCacheBlock) dst[i] = src[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int64_t wave = (int64_t) blockIdx.x * (kCacheBlock / 64) + (threadIdx.x >> 6);
    const int64_t nWaves = (int64_t) gridDim.x * (kCacheBlock / 64);
    constexpr int QB = kInferQB;
    for (int64_t base = wave * 16 * QB; base < n; base += nWaves * 16 * QB) {
        // compiler-only barrier: weight fragments are re-read from LDS each
        // trip instead of being hoisted into ~190 registers
        asm volatile("" ::: "memory");
        h4 x[QB][6];
#pragma unroll
        for (int qb = 0; qb < QB; qb++) {
            float qq[kCacheQ];
            loadQuery(q, base + qb * 16 + c, n, qq);
            const CacheQuery cq = cacheQuery(cv, qq, g);
#pragma unroll
            for (int kb = 0; kb < 6; kb++) x[qb][kb] = cacheInputFrag(cq, kb, g);
        }
        hiddenLayer<QB, 6>(w + cacheOffT(0), cacheStrideT(0), g, c, x);
#pragma unroll
        for (int l = 1; l < 5; l++) hiddenLayer<QB, 4>(w + cacheOffT(l), cacheStrideT(l), g, c, x);
        f4 y[QB];
        outputLayer<QB>(w + cacheOffT(5), g, c, x, y);
        if (g == 0) {
#pragma unroll
            for (int qb = 0; qb < QB; qb++) {
                const int64_t qi = base + qb * 16 + c;
                if (qi >= n) continue;
                if (result) {
                    const uint32_t s = slot[qi];
                    const float4 b = beta[qi];
                    float cx = b.x * fmaxf(y[qb][0], 0.f), cy = b.y * fmaxf(y[qb][1], 0.f),
                          cz = b.z * fmaxf(y[qb][2], 0.f);
                    const float lum = (cx + cy + cz) / 3;
                    if (clampInd > 0.f && lum > clampInd) {
                        const float k = clampInd / lum;
                        cx *= k, cy *= k, cz *= k;
                    }
                    float4 r = result[s];
                    if (isfinite(cx + cy + cz)) {
                        r.x += cx;
                        r.y += cy;
                        r.z += cz;
                    }
                    result[s] = r;
                } else {
                    out[3 * qi] = y[qb][0];
                    out[3 * qi + 1] = y[qb][1];
                    out[3 * qi + 2] = y[qb][2];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Training: layer inputs of the block's 128 queries as [feature][query] FP16.
template <int KB>
PRR_DEV void storeX(_Float16 *xs, int l, int g, int c, int wv, const h4 (&x)[2][6]) {
#pragma unroll
    for (int qb = 0; qb < 2; qb++)
#pragma unroll
        for (int kb = 0; kb < KB; kb++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                xs[(xsRow(l) + kb * 16 + 4 * g + i) * kXsStride + wv * 32 + qb * 16 + c] = x[qb][kb][i];
}

// Backward step of layer L: d holds dLoss/d(output of layer L) for this
// wave's queries (D layout). Writes the block's dW of layer L to part and
// leaves dLoss/d(output of layer L-1) in d.
template <int L>
PRR_DEV void backwardLayer(const CacheView &cv, _Float16 *xs, _Float16 *ds, float *part, int g, int c, int wv,
                           h4 (&d)[2][4]) {
    constexpr int OB = L == 5 ? 1 : 4;
    constexpr int IB = L == 0 ? 6 : 4;
#pragma unroll
    for (int qb = 0; qb < 2; qb++)
#pragma unroll
        for (int ob = 0; ob < OB; ob++)
#pragma unroll
            for (int i = 0; i < 4; i++) ds[(ob * 16 + 4 * g + i) * kXsStride + wv * 32 + qb * 16 + c] = d[qb][ob][i];
    __syncthreads();
    // dW^T[out][in] = sum over the block's queries of delta[out][q] x[in][q]
    for (int t = wv; t < OB * IB; t += kCacheBlock / 64) {
        const int ob = t / IB, ib = t - ob * IB;
        f4 acc = f4{0.f, 0.f, 0.f, 0.f};
        const _Float16 *ar = ds + (ob * 16 + c) * kXsStride + 4 * g;
        const _Float16 *br = xs + (xsRow(L) + ib * 16 + c) * kXsStride + 4 * g;
#pragma unroll
        for (int k = 0; k < kTrainQ / 16; k += 2)
            acc = mfma32(cat8(ldh4(ar + k * 16), ldh4(ar + k * 16 + 16)), cat8(ldh4(br + k * 16), ldh4(br + k * 16 + 16)),
                         acc);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = ob * 16 + 4 * g + i;
            if (o < cacheOut(L)) part[cacheOff(L) + o * cacheIn(L) + ib * 16 + c] = acc[i];
        }
    }
    if constexpr (L > 0) {
        // delta of the layer's input: W delta, times ReLU' of that input
        const _Float16 *wn = cv.wn + cacheOffN(L);
        constexpr int sN = cacheStrideN(L);
        h4 nd[2][4];
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            f4 acc[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
            const _Float16 *row = wn + (ib * 16 + c) * sN + 4 * g;
            if constexpr (OB == 1) {
                const h4 a = ldh4(row);
#pragma unroll
                for (int qb = 0; qb < 2; qb++) acc[qb] = mfma16(a, d[qb][0], acc[qb]);
            } else {
#pragma unroll
                for (int ob = 0; ob < OB; ob += 2) {
                    const h8 a = cat8(ldh4(row + ob * 16), ldh4(row + ob * 16 + 16));
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) acc[qb] = mfma32(a, cat8(d[qb][ob], d[qb][ob + 1]), acc[qb]);
                }
            }
#pragma unroll
            for (int qb = 0; qb < 2; qb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float xv = (float) xs[(xsRow(L) + ib * 16 + 4 * g + i) * kXsStride + wv * 32 + qb * 16 + c];
                    nd[qb][ib][i] = (_Float16) (xv > 0.f ? acc[qb][i] : 0.f);
                }
        }
#pragma unroll
        for (int qb = 0; qb < 2; qb++)
#pragma unroll
            for (int ib = 0; ib < 4; ib++) d[qb][ib] = nd[qb][ib];
    }
    __syncthreads();
}

// Relative L2 loss (Mueller et al. 2021): (y - t)^2 / (sg(y)^2 + 0.01) per
// channel; dL/dy = 2 (y - t) / (y^2 + 0.01), clipped to +-clip (the mean's
// 1/(3n) is applied to the summed gradient).
// Sample j < n of the batch is record (start + j * mult) mod total of q /
// target: mult = 1 takes a contiguous slice, a mult coprime to total spreads
// the batch over the whole record set (mult < 2^32, start < total).
__global__ __launch_bounds__(kCacheBlock) void kCacheTrain(CacheView cv, int64_t n, const float *__restrict__ q,
                                                           const float *__restrict__ target,
                                                           float *__restrict__ partial,
                                                           float *__restrict__ lossPart, float clip, int64_t start,
                                                           int64_t total, int64_t mult) {
    __shared__ __align__(16) _Float16 xs[kXsRows * kXsStride];
    __shared__ __align__(16) _Float16 ds[kCacheH * kXsStride];
    __shared__ float lossW[kCacheBlock / 64];
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wv = threadIdx.x >> 6;
    const int64_t base = (int64_t) blockIdx.x * kTrainQ + wv * 32;
    const _Float16 *wt = cv.wt;

    // j < 2^31 and mult < 2^32: the product fits in 64 bits
    auto rec = [&](int64_t j) { return (int64_t) (((uint64_t) j * (uint64_t) mult + (uint64_t) start) % (uint64_t) total); };
    // forward: this wave's 32 queries
    h4 x[2][6];
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        float qq[kCacheQ];
        const int64_t j = base + qb * 16 + c;
        if (j < n) {
            const int64_t r = rec(j);
#pragma unroll
            for (int k = 0; k < kCacheQ; k++) qq[k] = q[r * kCacheQ + k];
        } else {
            loadQuery(q, 0, 0, qq);
        }
        const CacheQuery cq = cacheQuery(cv, qq, g);
#pragma unroll
        for (int kb = 0; kb < 6; kb++) x[qb][kb] = cacheInputFrag(cq, kb, g);
    }
    storeX<6>(xs, 0, g, c, wv, x);
    hiddenLayer<2, 6>(wt + cacheOffT(0), cacheStrideT(0), g, c, x);
#pragma unroll
    for (int l = 1; l < 5; l++) {
        storeX<4>(xs, l, g, c, wv, x);
        hiddenLayer<2, 4>(wt + cacheOffT(l), cacheStrideT(l), g, c, x);
    }
    storeX<4>(xs, 5, g, c, wv, x);
    f4 y[2];
    outputLayer<2>(wt + cacheOffT(5), g, c, x, y);

    // loss gradient at the outputs (lanes g == 0 hold channels 0..2)
    h4 d[2][4];
    float loss = 0.f;
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const int64_t qi = base + qb * 16 + c;
        f4 dv = f4{0.f, 0.f, 0.f, 0.f};
        if (g == 0 && qi < n) {
            const int64_t r = rec(qi);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float yv = y[qb][k], t = target[3 * r + k];
                const float den = yv * yv + 0.01f;
                loss += (yv - t) * (yv - t) / den;
                dv[k] = fminf(clip, fmaxf(-clip, 2.f * (yv - t) / den));
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) d[qb][0][i] = (_Float16) dv[i];
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) lossW[wv] = loss;

    float *part = partial + (int64_t) blockIdx.x * kCacheWeights;
    backwardLayer<5>(cv, xs, ds, part, g, c, wv, d);
    backwardLayer<4>(cv, xs, ds, part, g, c, wv, d);
    backwardLayer<3>(cv, xs, ds, part, g, c, wv, d);
    backwardLayer<2>(cv, xs, ds, part, g, c, wv, d);
    backwardLayer<1>(cv, xs, ds, part, g, c, wv, d);
    backwardLayer<0>(cv, xs, ds, part, g, c, wv, d);
    if (threadIdx.x == 0) lossPart[blockIdx.x] = lossW[0] + lossW[1] + lossW[2] + lossW[3];
}

// ---------------------------------------------------------------------------
// Index of master weight w in the FP16 layouts.
PRR_DEV void cacheIndex(int w, int &l, int &o, int &i) {
    l = 0;
    while (l < kCacheLayers - 1 && w >= cacheOff(l + 1)) l++;
    const int r = w - cacheOff(l);
    o = r / cacheIn(l);
    i = r - o * cacheIn(l);
}

// Block of 256 threads = 32 weights x 8 groups: group g sums the partials of
// blocks [g nb / 8, (g + 1) nb / 8) in block order, then the group sums are
// added in group order (deterministic; spreads the 4 * nblocks * 22720-byte
// read over many CUs instead of one thread per weight).
constexpr int kAdamW = 32, kAdamG = 8;
__global__ __launch_bounds__(kAdamW * kAdamG) void kCacheAdam(float *__restrict__ master, float *__restrict__ m1,
                                                               float *__restrict__ m2, const float *__restrict__ partial,
                                                               int nblocks, CacheOpt opt, _Float16 *__restrict__ wt,
                                                               _Float16 *__restrict__ wn, float *__restrict__ gradOut) {
    __shared__ float gpart[kAdamG][kAdamW];
    const int wl = threadIdx.x % kAdamW, g = threadIdx.x / kAdamW;
    const int w = blockIdx.x * kAdamW + wl;
    if (partial) {
        float gsum = 0.f;
        if (w < kCacheWeights) {
            const int b0 = (int) ((int64_t) g * nblocks / kAdamG), b1 = (int) ((int64_t) (g + 1) * nblocks / kAdamG);
            // loads of 8 blocks in flight at once; the adds stay in block order
#pragma unroll 8
            for (int b = b0; b < b1; b++) gsum += partial[(int64_t) b * kCacheWeights + w];
        }
        gpart[g][wl] = gsum;
    }
    __syncthreads();
    if (g != 0 || w >= kCacheWeights) return;
    float wv = master[w];
    if (partial) {
        float gsum = 0.f;
#pragma unroll
        for (int k = 0; k < kAdamG; k++) gsum += gpart[k][wl];
        const float gr = gsum * opt.gscale;
        if (gradOut) {
            gradOut[w] = gr;
            return;
        }
        const float a = opt.beta1 * m1[w] + (1.f - opt.beta1) * gr;
        const float v = opt.beta2 * m2[w] + (1.f - opt.beta2) * gr * gr;
        m1[w] = a;
        m2[w] = v;
        wv -= opt.lr * (a / opt.bias1) / (sqrtf(v / opt.bias2) + opt.eps);
        master[w] = wv;
    }
    int l, o, i;
    cacheIndex(w, l, o, i);
    const _Float16 h = (_Float16) wv;
    wt[cacheOffT(l) + o * cacheStrideT(l) + i] = h;
    if (l > 0) wn[cacheOffN(l) + i * cacheStrideN(l) + o] = h;
}

// One wave: lane l sums blocks l, l + 64, ... in order, then a fixed
// butterfly adds the lanes' sums (the same order on every run).
__global__ void kCacheLossSum(const float *lossPart, int nblocks, float *out) {
    const int l = threadIdx.x;
    float s = 0.f;
    for (int b = l; b < nblocks; b += 64) s += lossPart[b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) *out = s;
}

// ---------------------------------------------------------------------------
int cacheTrainBlocks(int64_t n) { return (int) ((n + kTrainQ - 1) / kTrainQ); }

hipError_t launchCacheInfer(const CacheView &cv, int64_t n, const uint32_t *nDev, const float *q, float *out,
                            const uint32_t *slot, const float4 *beta, float4 *result, float clampInd, int grid,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t need = (n + 4 * 16 * kInferQB - 1) / (4 * 16 * kInferQB);
    const int blocks = (int) (need < grid ? need : grid);
    hipLaunchKernelGGL(kCacheInfer, dim3(blocks > 0 ? blocks : 1), dim3(kCacheBlock), 0, st, cv, n, nDev, q, out,
                       slot, beta, result, clampInd);
    return hipGetLastError();
}

hipError_t launchCacheTrain(const CacheView &cv, int64_t n, const float *q, const float *target, float *partial,
                            float *lossPart, float clip, int64_t start, int64_t total, int64_t mult,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kCacheTrain, dim3(cacheTrainBlocks(n)), dim3(kCacheBlock), 0, st, cv, n, q, target, partial,
                       lossPart, clip, start, total, mult);
    return hipGetLastError();
}

hipError_t launchCacheAdam(float *master, float *m1, float *m2, const float *partial, int nblocks, const CacheOpt &opt,
                           _Float16 *wt, _Float16 *wn, float *gradOut, hipStream_t st) {
    hipLaunchKernelGGL(kCacheAdam, dim3((kCacheWeights + kAdamW - 1) / kAdamW), dim3(kAdamW * kAdamG), 0, st, master,
                       m1, m2, partial, nblocks, opt, wt, wn, gradOut);
    return hipGetLastError();
}

hipError_t launchCacheLossSum(const float *lossPart, int nblocks, float *out, hipStream_t st) {
    hipLaunchKernelGGL(kCacheLossSum, dim3(1), dim3(64), 0, st, lossPart, nblocks, out);
    return hipGetLastError();
}

}  // namespace prr
