// non_silent_region.hip -- rppt_non_silent_region_detection_gpu
// (api/rppt_tensor_audio_augmentations.h:86).
//
// Semantics of the HOST kernel (src/modules/tensor/cpu/kernel/non_silent_region_detection.cpp:93-191):
// the moving mean square mms[p] = (sum of x[i]^2, p-W < i <= p) / W is a running FP32 sum that
// is restarted every resetInterval outputs (from a fresh sum of the W-1 samples before the
// block); the region runs from the first to the last p with mms[p] >= base * 10^(cutOffDB/10),
// base = max(mms) (referencePower == 0) or referencePower, and its begin is moved W-1 samples
// earlier (clamped at 0) unless it is 0.
//
// The running sum is evaluated in the HOST order (add the square, scale, subtract the square
// leaving the window; products rounded before the add, as the reference's -mno-fma build does),
// so mms is bit-identical to that sequence and the detected indices match it exactly.  The
// reference's HIP kernel uses per-tile prefix sums instead, whose rounding differs from HOST.
//
// Parallelism is the reset blocks: one lane runs one block's recurrence, or on 16-byte aligned rows
// four lanes share one block's loads and squares and run its recurrence together.  The recurrence's whole
// state is the running sum, so pass 1 records it every kSeg outputs together with the largest sum
// of those outputs; pass 2, one workgroup per clip, reduces the records to the cut-off (unless the
// reference power fixes it) and resumes the recurrence from the recorded sums over only the two
// kSeg-output spans holding the region's ends.
#include <cmath>
#include <type_traits>

#include "common.hpp"

namespace rppa {
namespace {

constexpr int kChunk = 32;          // samples per lane per prefetch step (three steps in flight)
constexpr int kVecChunk = 16;       // the same with 16-byte loads
constexpr int kVecSteps = 6;        // 16-byte-load chunks in flight per lane (3: 8% slower)
constexpr int kSeg = 192;           // outputs between checkpoints of the running sum
constexpr int kPass1Threads = 64;   // one wave per workgroup spreads the few lanes over CUs
// rows (reset blocks) per pass-1 wave: 32 of the 64 lanes.  Measured on workload 1 (5,378 rows of up
// to 10,240 samples over 512 x 39 lanes, block-major): 64 rows per wave 228.6 us, 32 219.4 us, 16 331 us, 8 515 us; each wave's
// instruction stream is the same for any number of active lanes, and below 32 rows per wave the
// waves outnumber what the compute units run at one wave per SIMD
constexpr int kPass1Rows = 32;
constexpr int kPass2Threads = 128;
static_assert(kSeg % (3 * kChunk) == 0 && kSeg % (kVecSteps * kVecChunk) == 0, "checkpoints fall on loop steps");

struct NsrArgs {
    const float *src;
    long long nStride;
    const int *len;
    int W;
    float meanFactor;
    int R;              // reset length; <= 0 means the whole clip is one block
    int maxBlocks;
    int segsPerBlock;   // checkpoints per reset block
    int maxSegs;        // checkpoints per clip row
    float cutOff;       // 10^(cutOffDB/10)
    float refPower;     // 0: per-clip maximum
    float *segMax;      // [B][maxSegs] maximum running sum over each checkpoint's outputs
    float *segSum;      // [B][maxSegs] running sum before each checkpoint's first output
    float *blockMax;    // [B][maxBlocks] maximum running sum over each reset block
    int *begin;
    int *length;
    int vec;            // rows 16-byte aligned, nStride and R multiples of 4, batch under 2 GiB:
                        // 16-byte buffer loads
    int srcBytes;       // B * nStride * 4 when vec
};

__device__ __forceinline__ int clip_len(const NsrArgs &a, int b)
{
    long long L = a.len[b];
    if (L < 0) L = 0;
    if (L > a.nStride) L = a.nStride;
    return static_cast<int>(L);
}

__device__ __forceinline__ int blocks_of(int L, int R) { return L <= 0 ? 0 : (R <= 0 ? 1 : (L + R - 1) / R); }

// x[i] for i clamped into [0, n): chunk loads may run past either end of a stream; the values
// loaded there are never used.
__device__ __forceinline__ float ld(const float *x, int i, int n) { return x[min(max(i, 0), n - 1)]; }

// Fresh sum of the W-1 squares before output `start`, in order (HOST :129-130).
__device__ __forceinline__ float fresh_sum(const float *x, int n, int start, int W)
{
    float sum = 0.0f;
    const int f0 = max(start - W + 1, 0);
    float b0[kChunk], b1[kChunk], b2[kChunk];
    auto load = [&](float (&v)[kChunk], int i0) {
#pragma unroll
        for (int u = 0; u < kChunk; u++) v[u] = ld(x, i0 + u, n);
    };
    auto add = [&](const float (&v)[kChunk], int i0) {
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            float q = v[u] * v[u];
            sum = sum + (i0 + u < start ? q : 0.0f);   // adding +0 keeps the sum
        }
    };
    if (f0 < start) {
        load(b0, f0);
        load(b1, f0 + kChunk);
        load(b2, f0 + 2 * kChunk);
        for (int i0 = f0; i0 < start; i0 += 3 * kChunk) {
            add(b0, i0);
            load(b0, i0 + 3 * kChunk);
            add(b1, i0 + kChunk);
            load(b1, i0 + 4 * kChunk);
            add(b2, i0 + 2 * kChunk);
            load(b2, i0 + 5 * kChunk);
        }
    }
    return sum;
}

// Runs the HOST recurrence over outputs [start, end) of one clip of length n from the running
// sum `sum` and calls visit(p, sum_p, p < end) in order, where mms[p] = sum_p * (1/W) rounded;
// mark(off, sum) is called before the outputs start + off of every loop step (off a multiple of
// 3 kChunk) with the sum the recurrence resumes from there.  Branch-free: the recurrence runs on
// to the end of the last chunk (values past `end` are flagged invalid and never used).  The
// recurrence is sequential (its rounding is the HOST's), so the loads run ahead of it: three
// chunks are in flight while one is consumed (the buffers rotate by name in a loop unrolled
// three times, so no values are copied).
template <class Visit, class Mark>
__device__ __forceinline__ void mms_main(const float *x, int n, int start, int end, int W, float sum,
                                         Visit &visit, Mark &mark)
{
    float a0[kChunk], a1[kChunk], a2[kChunk], s0[kChunk], s1[kChunk], s2[kChunk];
    auto load2 = [&](float (&av)[kChunk], float (&sv)[kChunk], int p0) {
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            av[i] = ld(x, p0 + i, n);
            sv[i] = ld(x, p0 - W + 1 + i, n);
        }
    };
    auto run = [&](const float (&av)[kChunk], const float (&sv)[kChunk], int p0) {
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            const int p = p0 + i;
            float q = av[i] * av[i];
            sum = sum + q;
            visit(p, sum, p < end);
            float r = sv[i] * sv[i];
            sum = sum - (p - W + 1 >= 0 ? r : 0.0f);   // x - 0 is x: the HOST's skipped subtraction
        }
    };
    load2(a0, s0, start);
    load2(a1, s1, start + kChunk);
    load2(a2, s2, start + 2 * kChunk);
    for (int p0 = start; p0 < end; p0 += 3 * kChunk) {
        mark(p0 - start, sum);
        run(a0, s0, p0);
        load2(a0, s0, p0 + 3 * kChunk);
        run(a1, s1, p0 + kChunk);
        load2(a1, s1, p0 + 4 * kChunk);
        run(a2, s2, p0 + 2 * kChunk);
        load2(a2, s2, p0 + 5 * kChunk);
    }
}

// 16-byte loads of a clip's row through a buffer resource over the whole batch (under 2 GiB):
// loads past the batch's end, or before its start (a negative offset wraps above 2^31), return
// zeros instead of faulting, and loads before the row's start or past its end read neighbouring
// rows; neither kind of value reaches an output.
struct Row4 {
    __amdgpu_buffer_rsrc_t rsrc;
    unsigned rowOff;   // byte offset of the row

    __device__ __forceinline__ float4 operator()(int i) const
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, rowOff + static_cast<unsigned>(i) * 4u, 0, 0);
        return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
    }
};

// Two samples of the recurrence, s += q0 (visited), s -= r0, s += q1 (visited), s -= r1, m = max of m
// and the two visited sums, with the squares of two later samples (x, y -> nq, nr) formed between
// the dependent operations: a lone wave issues in order, and each dependent FP32 operation waits
// about 8 cycles for its operand, so the products fill those slots (the compiler schedules each
// product next to the addition that uses it).  The operations and their rounding are the C++ ones.
__device__ __forceinline__ void mms_pair(float &s, float &m, float q0, float q1, float r0, float r1, float x0, float x1,
                                         float y0, float y1, float &nq0, float &nq1, float &nr0, float &nr1)
{
    float a0, a1;
    __asm__("v_add_f32 %[a0], %[s], %[q0]\n\t"
            "v_mul_f32 %[nq0], %[x0], %[x0]\n\t"
            "v_sub_f32 %[s], %[a0], %[r0]\n\t"
            "v_mul_f32 %[nr0], %[y0], %[y0]\n\t"
            "v_add_f32 %[a1], %[s], %[q1]\n\t"
            "v_mul_f32 %[nq1], %[x1], %[x1]\n\t"
            "v_sub_f32 %[s], %[a1], %[r1]\n\t"
            "v_mul_f32 %[nr1], %[y1], %[y1]\n\t"
            "v_max3_f32 %[m], %[m], %[a0], %[a1]"
            : [s] "+v"(s), [m] "+v"(m), [a0] "=&v"(a0), [a1] "=&v"(a1), [nq0] "=&v"(nq0), [nq1] "=&v"(nq1),
              [nr0] "=&v"(nr0), [nr1] "=&v"(nr1)
            : [q0] "v"(q0), [q1] "v"(q1), [r0] "v"(r0), [r1] "v"(r1), [x0] "v"(x0), [x1] "v"(x1), [y0] "v"(y0),
              [y1] "v"(y1));
}

// Two terms of a running sum, s += q0, s += q1, with the squares of two later samples formed between.
__device__ __forceinline__ void add_pair(float &s, float q0, float q1, float x0, float x1, float &nq0, float &nq1)
{
    __asm__("v_add_f32 %[s], %[s], %[q0]\n\t"
            "v_mul_f32 %[nq0], %[x0], %[x0]\n\t"
            "v_add_f32 %[s], %[s], %[q1]\n\t"
            "v_mul_f32 %[nq1], %[x1], %[x1]"
            : [s] "+v"(s), [nq0] "=&v"(nq0), [nq1] "=&v"(nq1)
            : [q0] "v"(q0), [q1] "v"(q1), [x0] "v"(x0), [x1] "v"(x1));
}

__device__ __forceinline__ float comp(const float4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// The running maximum of the visited sums (the pass-1 visitor; the unmasked loop steps update m
// directly).
struct RunMax {
    float m = -INFINITY;
    __device__ __forceinline__ void operator()(int, float s, bool valid) { m = fmaxf(m, valid ? s : -INFINITY); }
};

// The same recurrence on 16-byte loads, for rows that are 16-byte aligned with a start that is a
// multiple of 4: the operations and their order are mms_main's, only the loads differ (the
// trailing samples x[p - W + 1] are elements M .. M+3 of two aligned groups, M = (start - W + 1) & 3).
// Full loop steps run unmasked; the last, partial step flags the outputs past `end`.  Clamp
// (start < W - 1) keeps the HOST's skipped subtraction for outputs p < W - 1.
template <int M, bool Clamp, class Visit, class Mark>
__device__ __forceinline__ float mms_main_vec(const Row4 &ld4, int start, int end, int W, float sum,
                                              Visit &visit, Mark &mark)
{
    constexpr int G = kVecChunk / 4;                       // 16-byte groups per chunk
    constexpr int NS = kVecSteps;                          // chunks in flight
    const int tb0 = start - W + 1 - M;                     // a multiple of 4
    float4 A[NS][G], T[NS][G];
    auto load = [&](int c, int p0) {
        const int t0 = tb0 + (p0 - start);
#pragma unroll
        for (int g = 0; g < G; g++) {
            A[c][g] = ld4(p0 + 4 * g);
            T[c][g] = ld4(t0 + 4 * g);
        }
    };
    auto run = [&](int c, const float4 &Tn, int p0, auto masked) {
        // the chunk's squares first (independent of the running sum), held behind an empty asm so
        // they are not scheduled next to the additions: a lone wave would wait out each product's
        // latency inside the dependent chain
        float qv[kVecChunk], rv[kVecChunk];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const float4 t1 = g + 1 < G ? T[c][g + 1] : Tn;
            const float a[4] = {A[c][g].x, A[c][g].y, A[c][g].z, A[c][g].w};
            const float t8[8] = {T[c][g].x, T[c][g].y, T[c][g].z, T[c][g].w, t1.x, t1.y, t1.z, t1.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                qv[4 * g + u] = a[u] * a[u];
                rv[4 * g + u] = t8[M + u] * t8[M + u];
            }
        }
#pragma unroll
        for (int i = 0; i < kVecChunk; i += 8)
            __asm__ volatile("" : "+v"(qv[i]), "+v"(qv[i + 1]), "+v"(qv[i + 2]), "+v"(qv[i + 3]), "+v"(qv[i + 4]),
                                  "+v"(qv[i + 5]), "+v"(qv[i + 6]), "+v"(qv[i + 7]), "+v"(rv[i]), "+v"(rv[i + 1]),
                                  "+v"(rv[i + 2]), "+v"(rv[i + 3]), "+v"(rv[i + 4]), "+v"(rv[i + 5]), "+v"(rv[i + 6]),
                                  "+v"(rv[i + 7]));
#pragma unroll
        for (int i = 0; i < kVecChunk; i++) {
            const int p = p0 + i;
            sum = sum + qv[i];
            visit(p, sum, decltype(masked)::value ? p < end : true);
            sum = sum - ((!Clamp || p - W + 1 >= 0) ? rv[i] : 0.0f);   // x - 0 is x
        }
    };
    // unclamped full steps: sample i's value and trailing value (chunk c's last group reads the first
    // trailing group of chunk c + 1; for the last chunk that is chunk 0's buffer, already reloaded
    // with the next step's first chunk); the squares of the next two samples are carried in q, r
    auto xval = [&](int c, int i) { return comp(A[c][i >> 2], i & 3); };
    auto yval = [&](int c, int i) {
        const int g = i >> 2, k = M + (i & 3);
        return comp(k < 4 ? T[c][g] : (g + 1 < G ? T[c][g + 1] : T[(c + 1) % NS][0]), k & 3);
    };
    float q[2], r[2];
    auto run_fast = [&](int c) {
        const int cn = (c + 1) % NS;
#pragma unroll
        for (int i = 0; i < kVecChunk; i += 2) {
            const float x0 = i + 2 < kVecChunk ? xval(c, i + 2) : xval(cn, i + 2 - kVecChunk);
            const float x1 = i + 3 < kVecChunk ? xval(c, i + 3) : xval(cn, i + 3 - kVecChunk);
            const float y0 = i + 2 < kVecChunk ? yval(c, i + 2) : yval(cn, i + 2 - kVecChunk);
            const float y1 = i + 3 < kVecChunk ? yval(c, i + 3) : yval(cn, i + 3 - kVecChunk);
            float nq0, nq1, nr0, nr1;
            mms_pair(sum, visit.m, q[0], q[1], r[0], r[1], x0, x1, y0, y1, nq0, nq1, nr0, nr1);
            q[0] = nq0;
            q[1] = nq1;
            r[0] = nr0;
            r[1] = nr1;
        }
    };
#pragma unroll
    for (int c = 0; c < NS; c++) load(c, start + c * kVecChunk);
    int p0 = start;
    if (!Clamp && p0 + NS * kVecChunk <= end) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            q[u] = xval(0, u) * xval(0, u);
            r[u] = yval(0, u) * yval(0, u);
        }
    }
    for (; p0 + NS * kVecChunk <= end; p0 += NS * kVecChunk) {
        mark(p0 - start, sum);
#pragma unroll
        for (int c = 0; c < NS; c++) {
            if (Clamp) run(c, T[(c + 1) % NS][0], p0 + c * kVecChunk, std::false_type{});
            else run_fast(c);
            load(c, p0 + (NS + c) * kVecChunk);
        }
    }
    if (p0 < end) {
        mark(p0 - start, sum);
        const float4 tn = ld4(tb0 + (p0 - start) + NS * kVecChunk);   // the group after the last chunk
#pragma unroll
        for (int c = 0; c < NS; c++) run(c, c + 1 < NS ? T[c + 1][0] : tn, p0 + c * kVecChunk, std::true_type{});
    }
    return sum;
}

// fresh_sum on aligned 16-byte groups: the first group is masked below f0, the full loop steps
// add unmasked, the last, partial step is masked at `start`.
__device__ __forceinline__ float fresh_sum_vec(const Row4 &ld4, int start, int W)
{
    constexpr int G = kVecChunk / 4;
    constexpr int NS = kVecSteps;
    float sum = 0.0f;
    const int f0 = max(start - W + 1, 0);
    if (f0 >= start) return sum;
    const int a0 = f0 & ~3;
    float4 bv[NS][G];
    auto load = [&](int c, int i0) {
#pragma unroll
        for (int g = 0; g < G; g++) bv[c][g] = ld4(i0 + 4 * g);
    };
    auto add = [&](int c, int i0, auto masked) {
        float qv[kVecChunk];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const float e[4] = {bv[c][g].x, bv[c][g].y, bv[c][g].z, bv[c][g].w};
#pragma unroll
            for (int u = 0; u < 4; u++) qv[4 * g + u] = e[u] * e[u];
        }
#pragma unroll
        for (int i = 0; i < kVecChunk; i += 8)
            __asm__ volatile("" : "+v"(qv[i]), "+v"(qv[i + 1]), "+v"(qv[i + 2]), "+v"(qv[i + 3]), "+v"(qv[i + 4]),
                                  "+v"(qv[i + 5]), "+v"(qv[i + 6]), "+v"(qv[i + 7]));
#pragma unroll
        for (int i = 0; i < kVecChunk; i++)
            sum = sum + ((!decltype(masked)::value || i0 + i < start) ? qv[i] : 0.0f);   // +0 keeps the sum
    };
    {
        const float4 v = ld4(a0);
        const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float q = e[u] * e[u];
            sum = sum + ((a0 + u >= f0 && a0 + u < start) ? q : 0.0f);
        }
    }
    int i0 = a0 + 4;
#pragma unroll
    for (int c = 0; c < NS; c++) load(c, i0 + c * kVecChunk);
    // full steps: the squares of the next two samples carried in q (see add_pair)
    float q[2];
    auto xval = [&](int c, int i) { return comp(bv[c][i >> 2], i & 3); };
    if (i0 + NS * kVecChunk <= start) {
        q[0] = xval(0, 0) * xval(0, 0);
        q[1] = xval(0, 1) * xval(0, 1);
    }
    for (; i0 + NS * kVecChunk <= start; i0 += NS * kVecChunk) {
#pragma unroll
        for (int c = 0; c < NS; c++) {
            const int cn = (c + 1) % NS;
#pragma unroll
            for (int i = 0; i < kVecChunk; i += 2) {
                const float x0 = i + 2 < kVecChunk ? xval(c, i + 2) : xval(cn, i + 2 - kVecChunk);
                const float x1 = i + 3 < kVecChunk ? xval(c, i + 3) : xval(cn, i + 3 - kVecChunk);
                float nq0, nq1;
                add_pair(sum, q[0], q[1], x0, x1, nq0, nq1);
                q[0] = nq0;
                q[1] = nq1;
            }
            load(c, i0 + (NS + c) * kVecChunk);
        }
    }
    if (i0 < start) {
#pragma unroll
        for (int c = 0; c < NS; c++) add(c, i0 + c * kVecChunk, std::true_type{});
    }
    return sum;
}

__device__ __forceinline__ Row4 row4(const NsrArgs &a, int b)
{
    return Row4{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(a.src), 0, a.srcBytes, 0x00020000),
                static_cast<unsigned>(b) * static_cast<unsigned>(a.nStride) * 4u};
}

__device__ __forceinline__ float fresh(const NsrArgs &a, const float *x, int b, int L, int start)
{
    return a.vec ? fresh_sum_vec(row4(a, b), start, a.W) : fresh_sum(x, L, start, a.W);
}

template <int M, class Visit, class Mark>
__device__ __forceinline__ void mms_vec(const Row4 &ld4, int start, int end, int W, float sum, Visit &visit, Mark &mark)
{
    if (start - W + 1 < 0) {
        // the clamped form only up to the first checkpoint at or past output W - 1 (a multiple of
        // kSeg from start, so the checkpoints and the trailing alignment M carry on unchanged)
        const int split = start + (W - 1 - start + kSeg - 1) / kSeg * kSeg;
        sum = mms_main_vec<M, true>(ld4, start, min(split, end), W, sum, visit, mark);
        if (split < end) mms_main_vec<M, false>(ld4, split, end, W, sum, visit, mark);
    } else {
        mms_main_vec<M, false>(ld4, start, end, W, sum, visit, mark);
    }
}

template <class Visit, class Mark>
__device__ __forceinline__ void mms_run(const NsrArgs &a, const float *x, int b, int L, int start, int end, float sum,
                                        Visit &visit, Mark &mark)
{
    if (a.vec) {
        const Row4 ld4 = row4(a, b);
        switch ((start - a.W + 1) & 3) {
        case 0: mms_vec<0>(ld4, start, end, a.W, sum, visit, mark); break;
        case 1: mms_vec<1>(ld4, start, end, a.W, sum, visit, mark); break;
        case 2: mms_vec<2>(ld4, start, end, a.W, sum, visit, mark); break;
        default: mms_vec<3>(ld4, start, end, a.W, sum, visit, mark); break;
        }
    } else {
        mms_main(x, L, start, end, a.W, sum, visit, mark);
    }
}

// ---- Pass 1 on aligned rows, four lanes per reset block ----------------------------------------
// Lane q of a quad loads samples 4q .. 4q+3 of each 16-sample chunk of its block, and of the trailing
// samples W - 1 behind, and squares them; every lane of the quad runs the block's recurrence on the
// squares taken from the lanes that hold them: each 16-sample chunk's 32 squares are first broadcast
// across the quad (DPP quad permutes, independent of one another), then the chain runs in plain
// additions and subtractions, so no permute waits on the chain or the chain on a permute (31 cycles
// per sample with the permutes folded into the chain's operations, 166 us per detection on workload 1;
// 151 us this way).  A block's loads run four times as far ahead for the same registers as with one
// lane per block, and each square is formed once.  The operations and their order are mms_main's.
constexpr int kQuadRows = 16;                  // reset blocks per wave
constexpr int kQuadChunks = 24;                // 16-sample chunks of each stream in flight per block
constexpr int kQuadWaves = 2;                  // waves per workgroup (1: 152.6 us per detection, 2: 151.3 us)
static_assert((kQuadChunks * kVecChunk) % kSeg == 0 && kSeg % kVecChunk == 0, "checkpoints fall on chunks");

__device__ __forceinline__ float4 sq4(const float4 &v) { return make_float4(v.x * v.x, v.y * v.y, v.z * v.z, v.w * v.w); }
__device__ __forceinline__ float el(const float4 &v, int e) { return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w)); }

// Broadcast of quad lane K's value to the quad with bound_ctrl set (no read of the old destination).
template <int K>
__device__ __forceinline__ float qbb(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), K * 0x55, 0xf, 0xf, true));
}
// Two steps of the recurrence with plain operations: s += q0 (a0), s -= r0, s += q1 (a1), s -= r1,
// m = max(m, a0, a1) -- mms_pair without the products.
__device__ __forceinline__ void qpair(float &s, float &m, float q0, float q1, float r0, float r1)
{
    float a0, a1;
    __asm__ volatile("v_add_f32 %[a0], %[s], %[q0]\n\t"
                     "v_sub_f32 %[s], %[a0], %[r0]\n\t"
                     "v_add_f32 %[a1], %[s], %[q1]\n\t"
                     "v_sub_f32 %[s], %[a1], %[r1]\n\t"
                     "v_max3_f32 %[m], %[m], %[a0], %[a1]"
                     : [s] "+v"(s), [m] "+v"(m), [a0] "=&v"(a0), [a1] "=&v"(a1)
                     : [q0] "v"(q0), [q1] "v"(q1), [r0] "v"(r0), [r1] "v"(r1));
}

// One 16-sample chunk of the recurrence: each sample's squares are first broadcast from the lanes that
// hold them (DPP, off the dependent chain), then the chain runs in plain operations.  q4/rv are this
// chunk's squares (this lane's group), rn the next trailing chunk's (trailing values past this chunk's
// end).  MASK: the tail chunk, whose samples from `lim` on are not visited.
template <int M, bool MASK>
__device__ __forceinline__ void qchunk(float &s, float &m, const float4 &q4, const float4 &rv, const float4 &rn, int lim)
{
    float bq[16], br[16];
#define RPPA_QBR(U)                                                                                                 \
    {                                                                                                               \
        constexpr int t_ = (U) + M;                                                                                 \
        bq[U] = qbb<((U) >> 2)>(el(q4, (U) & 3));                                                                   \
        br[U] = qbb<((t_ < 16) ? (t_ >> 2) : 0)>((t_ < 16) ? el(rv, t_ & 3) : el(rn, t_ & 3));                      \
    }
    RPPA_QBR(0) RPPA_QBR(1) RPPA_QBR(2) RPPA_QBR(3) RPPA_QBR(4) RPPA_QBR(5) RPPA_QBR(6) RPPA_QBR(7)
    RPPA_QBR(8) RPPA_QBR(9) RPPA_QBR(10) RPPA_QBR(11) RPPA_QBR(12) RPPA_QBR(13) RPPA_QBR(14) RPPA_QBR(15)
#undef RPPA_QBR
    if constexpr (!MASK) {
#pragma unroll
        for (int i = 0; i < 16; i += 2) qpair(s, m, bq[i], bq[i + 1], br[i], br[i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            s = s + bq[i];
            m = (i < lim) ? fmaxf(m, s) : m;
            s = s - br[i];
        }
    }
}

// The recurrence over outputs [start, end) of one block from `sum` (quad-cooperative; see above):
// calls mark(off, sum, m) at every checkpoint and keeps the running maximum in m.
template <int M, class Mark>
__device__ __forceinline__ void mms_quad(const Row4 &ld4, int q4i, int start, int end, int W, float &sum, float &m,
                                         Mark &mark)
{
    constexpr int NC = kQuadChunks;
    const int tb0 = start - W + 1 - M;                     // a multiple of 4
    float4 X[NC], T[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        X[c] = ld4(start + 16 * c + 4 * q4i);
        T[c] = ld4(tb0 + 16 * c + 4 * q4i);
    }
    // trailing squares: positions before the clip's start subtract 0 (the HOST's skipped subtraction)
    auto rsq = [&](const float4 &v, int pos) {
        float4 r = sq4(v);
        if (pos < 0) {
            r.x = pos + 0 >= 0 ? r.x : 0.0f;
            r.y = pos + 1 >= 0 ? r.y : 0.0f;
            r.z = pos + 2 >= 0 ? r.z : 0.0f;
            r.w = pos + 3 >= 0 ? r.w : 0.0f;
        }
        return r;
    };
    float4 rv = rsq(T[0], tb0 + 4 * q4i);
    int p0 = start;
    for (; p0 + NC * 16 <= end; p0 += NC * 16) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if ((c * 16) % kSeg == 0) mark(p0 + 16 * c - start, sum, m);
            const float4 q4 = sq4(X[c]);
            const float4 rn = rsq(T[(c + 1) % NC], tb0 + (p0 - start) + 16 * (c + 1) + 4 * q4i);
            qchunk<M, false>(sum, m, q4, rv, rn, 16);
            rv = rn;
            X[c] = ld4(p0 + 16 * (NC + c) + 4 * q4i);
            T[c] = ld4(tb0 + (p0 - start) + 16 * (NC + c) + 4 * q4i);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (p0 + 16 * c < end) {
            if ((c * 16) % kSeg == 0) mark(p0 + 16 * c - start, sum, m);
            const float4 q4 = sq4(X[c]);
            const float4 rn = rsq(T[(c + 1) % NC], tb0 + (p0 - start) + 16 * (c + 1) + 4 * q4i);
            qchunk<M, true>(sum, m, q4, rv, rn, end - (p0 + 16 * c));
            rv = rn;
            T[c] = ld4(tb0 + (p0 - start) + 16 * (NC + c) + 4 * q4i);   // the last chunk reads chunk NC's first group
        }
    }
}

// The fresh sum of the W - 1 squares before `start`, in order, quad-cooperative.
__device__ __forceinline__ float fresh_quad(const Row4 &ld4, int q4i, int start, int W)
{
    constexpr int NC = kQuadChunks;
    float sum = 0.0f;
    const int f0 = max(start - W + 1, 0);
    if (f0 >= start) return sum;
    const int a0 = f0 & ~3;
    float4 X[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) X[c] = ld4(a0 + 16 * c + 4 * q4i);
    auto msq = [&](const float4 &v, int pos) {
        float4 r = sq4(v);
        r.x = (pos + 0 >= f0 && pos + 0 < start) ? r.x : 0.0f;   // adding +0 keeps the sum
        r.y = (pos + 1 >= f0 && pos + 1 < start) ? r.y : 0.0f;
        r.z = (pos + 2 >= f0 && pos + 2 < start) ? r.z : 0.0f;
        r.w = (pos + 3 >= f0 && pos + 3 < start) ? r.w : 0.0f;
        return r;
    };
    // Whole loop steps, branch-free: the chunks of the last step past `start` add masked zeros, which
    // leave the sum unchanged (it is never negative), and their loads read ahead harmlessly.  A branch
    // per chunk, or the scheduler gathering every chunk's squares ahead of the adds, made the loop
    // wait for all of its loads once a step.
    for (int i0 = a0; i0 < start; i0 += NC * 16) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const float4 q4 = msq(X[c], i0 + 16 * c + 4 * q4i);
            float bq[16];
#define RPPA_QF(U) bq[U] = qbb<((U) >> 2)>(el(q4, (U) & 3));
            RPPA_QF(0) RPPA_QF(1) RPPA_QF(2) RPPA_QF(3) RPPA_QF(4) RPPA_QF(5) RPPA_QF(6) RPPA_QF(7)
            RPPA_QF(8) RPPA_QF(9) RPPA_QF(10) RPPA_QF(11) RPPA_QF(12) RPPA_QF(13) RPPA_QF(14) RPPA_QF(15)
#pragma unroll
            for (int u = 0; u < 16; u++) sum = sum + bq[u];
            X[c] = ld4(i0 + 16 * (NC + c) + 4 * q4i);
            __builtin_amdgcn_sched_barrier(0);   // chunk by chunk: each waits only for its own load
        }
    }
#undef RPPA_QF
    return sum;
}

// The blocks are numbered clip by clip (the prefix of each clip's block count, in LDS) and dealt to
// the quads in that order, so every wave but the last holds 16 blocks: the waves are as few as the
// blocks allow, one per SIMD.
__global__ void __launch_bounds__(64 * kQuadWaves) nsr_blocks_quad_kernel(NsrArgs a, int B)
{
    extern __shared__ int pre[];   // [B + 1]
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int wv = static_cast<int>(threadIdx.x) >> 6;
    int carry = 0;
    if (wv == 0)
    for (int base = 0; base < B; base += 64) {
        const int bb = base + lane;
        const int n = bb < B ? blocks_of(clip_len(a, bb), a.R) : 0;
        int v = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o);
            if (lane >= o) v += u;
        }
        if (bb < B) pre[bb] = carry + v - n;
        carry += __shfl(v, 63);
    }
    if (threadIdx.x == 0) pre[B] = carry;
    __syncthreads();
    const int q4 = lane & 3;
    const int row = (static_cast<int>(blockIdx.x) * kQuadWaves + wv) * kQuadRows + (lane >> 2);
    if (row >= pre[B]) return;   // a whole quad returns together
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= row) lo = mid;
        else hi = mid;
    }
    const int b = lo;
    const int j = row - pre[b];
    const int L = clip_len(a, b);
    const int R = a.R <= 0 ? L : a.R;
    const int start = j * R;
    const int end = static_cast<int>(min(static_cast<long long>(L), static_cast<long long>(start) + R));
    const Row4 ld4 = row4(a, b);
    float sum = fresh_quad(ld4, q4, start, a.W);
    float *segMax = a.segMax + static_cast<long long>(b) * a.maxSegs + static_cast<long long>(j) * a.segsPerBlock;
    float *segSum = a.segSum + static_cast<long long>(b) * a.maxSegs + static_cast<long long>(j) * a.segsPerBlock;
    float mb = -INFINITY, m = -INFINITY;
    int k = -1;
    auto mark = [&](int off, float s, float &mm) {
        if (k >= 0) {
            if (q4 == 0) segMax[k] = mm;
            mb = fmaxf(mb, mm);
        }
        ++k;
        mm = -INFINITY;
        if (q4 == 0) segSum[k] = s;
    };
    switch ((start - a.W + 1) & 3) {
    case 0: mms_quad<0>(ld4, q4, start, end, a.W, sum, m, mark); break;
    case 1: mms_quad<1>(ld4, q4, start, end, a.W, sum, m, mark); break;
    case 2: mms_quad<2>(ld4, q4, start, end, a.W, sum, m, mark); break;
    default: mms_quad<3>(ld4, q4, start, end, a.W, sum, m, mark); break;
    }
    if (q4 == 0) {
        segMax[k] = m;   // start < end: the first chunk marked checkpoint 0
        a.blockMax[static_cast<long long>(b) * a.maxBlocks + j] = fmaxf(mb, m);
    }
}

// Pass 1: one lane per reset block records, every kSeg outputs, the running sum the recurrence
// resumes from and the largest running sum of those outputs, and the block's largest running sum
// (mms = sum * (1/W) rounded is monotone in sum, so the largest mms is the largest sum times 1/W,
// rounded, and a block or span holds an mms >= cut exactly when its largest sum times 1/W does).
__global__ void __launch_bounds__(kPass1Threads) nsr_blocks_kernel(NsrArgs a, int B)
{
    if (threadIdx.x >= kPass1Rows) return;
    const long long lane = static_cast<long long>(blockIdx.x) * kPass1Rows + threadIdx.x;
    if (lane >= static_cast<long long>(B) * a.maxBlocks) return;
    // consecutive lanes take the same block of consecutive clips: the lanes of a wave then read
    // rows a clip stride apart rather than blocks a reset interval apart (a power-of-two stride
    // that sends every lane's load to the same memory channel)
    const int b = static_cast<int>(lane % B);
    const int j = static_cast<int>(lane / B);
    const int L = clip_len(a, b);
    if (j >= blocks_of(L, a.R)) return;
    const int R = a.R <= 0 ? L : a.R;
    const int start = j * R;
    const int end = static_cast<int>(min(static_cast<long long>(L), static_cast<long long>(start) + R));
    const float *x = a.src + b * a.nStride;
    const float sum0 = fresh(a, x, b, L, start);
    float *segMax = a.segMax + static_cast<long long>(b) * a.maxSegs + static_cast<long long>(j) * a.segsPerBlock;
    float *segSum = a.segSum + static_cast<long long>(b) * a.maxSegs + static_cast<long long>(j) * a.segsPerBlock;
    float mb = -INFINITY;
    int k = -1;
    RunMax visit;
    auto mark = [&](int off, float s) {
        if (off % kSeg == 0) {
            if (k >= 0) {
                segMax[k] = visit.m;
                mb = fmaxf(mb, visit.m);
            }
            ++k;
            visit.m = -INFINITY;
            segSum[k] = s;
        }
    };
    mms_run(a, x, b, L, start, end, sum0, visit, mark);
    segMax[k] = visit.m;   // start < end: the first loop step marked checkpoint 0
    a.blockMax[static_cast<long long>(b) * a.maxBlocks + j] = fmaxf(mb, visit.m);
}

// Pass 2: one workgroup per clip reduces pass 1's records to the cut-off (unless the reference
// power fixes it), finds the first and the last reset block holding a hit, in them the first and
// the last checkpoint span holding one, and reruns the recurrence over only those two spans from
// their recorded sums.
__global__ void __launch_bounds__(kPass2Threads) nsr_region_kernel(NsrArgs a)
{
    const int b = blockIdx.x;
    const int L = clip_len(a, b);
    const int nb = blocks_of(L, a.R);
    __shared__ float sMax[kPass2Threads];
    __shared__ int sFirstBlk, sLastBlk, sFirst, sLast, sBegin, sEnd;
    if (threadIdx.x == 0) {
        sFirstBlk = 0x7fffffff;
        sLastBlk = -1;
        sFirst = 0x7fffffff;
        sLast = -1;
        sBegin = L;
        sEnd = -1;
    }
    __syncthreads();

    const int R = a.R <= 0 ? L : a.R;
    const float mf = a.meanFactor;
    const float *bm = a.blockMax + static_cast<long long>(b) * a.maxBlocks;
    float m = -INFINITY;
    for (int j = threadIdx.x; j < nb; j += kPass2Threads) m = fmaxf(m, bm[j]);
    sMax[threadIdx.x] = m;
    __syncthreads();
    for (int s = kPass2Threads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sMax[threadIdx.x] = fmaxf(sMax[threadIdx.x], sMax[threadIdx.x + s]);
        __syncthreads();
    }
    // the reference power, else the largest mms, times the cut-off ratio
    const float cut = a.refPower != 0.0f ? a.refPower * a.cutOff : (sMax[0] * mf) * a.cutOff;
    // the first and the last reset block holding a hit, then the first checkpoint span of the
    // one and the last of the other holding a hit
    for (int j = threadIdx.x; j < nb; j += kPass2Threads) {
        if (bm[j] * mf >= cut) {
            atomicMin(&sFirstBlk, j);
            atomicMax(&sLastBlk, j);
        }
    }
    __syncthreads();
    if (sLastBlk >= 0) {
        const float *sm = a.segMax + static_cast<long long>(b) * a.maxSegs;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int j = h == 0 ? sFirstBlk : sLastBlk;
            const long long blockEnd = min(static_cast<long long>(L), static_cast<long long>(j) * R + R);
            const int segs = static_cast<int>((blockEnd - static_cast<long long>(j) * R + kSeg - 1) / kSeg);
            for (int k = threadIdx.x; k < segs; k += kPass2Threads) {
                const int g = j * a.segsPerBlock + k;
                if (sm[g] * mf >= cut) {
                    if (h == 0) atomicMin(&sFirst, g);
                    else atomicMax(&sLast, g);
                }
            }
        }
        __syncthreads();
        // the squares of the two spans' samples and trailing samples x[p - W + 1] go to LDS
        // (0 where p < W - 1: x - 0 is x, the HOST's skipped subtraction); one lane per span
        // runs the additions in order and keeps every running sum; then all threads search
        __shared__ float sQ[2][kSeg], sR[2][kSeg], sS[2][kSeg];
        const float *x = a.src + b * a.nStride;
        int s0[2], n[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int g = h == 0 ? sFirst : sLast;
            const int j = g / a.segsPerBlock;
            s0[h] = j * R + (g - j * a.segsPerBlock) * kSeg;
            const long long e0 = min(static_cast<long long>(L),
                                     min(static_cast<long long>(s0[h]) + kSeg, static_cast<long long>(j) * R + R));
            n[h] = static_cast<int>(e0 - s0[h]);
            for (int i = threadIdx.x; i < n[h]; i += kPass2Threads) {
                const int t = s0[h] + i - a.W + 1;
                const float v = x[s0[h] + i];
                const float w = t >= 0 ? x[t] : 0.0f;
                sQ[h][i] = v * v;
                sR[h][i] = w * w;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 || threadIdx.x == kWave) {
            const int h = threadIdx.x == 0 ? 0 : 1;
            float sum = a.segSum[static_cast<long long>(b) * a.maxSegs + (h == 0 ? sFirst : sLast)];
            const int nh = n[h];
#pragma unroll 8
            for (int i = 0; i < nh; i++) {
                sum = sum + sQ[h][i];
                sS[h][i] = sum;
                sum = sum - sR[h][i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            for (int i = threadIdx.x; i < n[h]; i += kPass2Threads) {
                if (sS[h][i] * mf >= cut) {
                    if (h == 0) atomicMin(&sBegin, s0[h] + i);
                    else atomicMax(&sEnd, s0[h] + i);
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int begin = 0, length = 0;
        if (L > 0 && sEnd >= 0 && sBegin <= sEnd) {
            begin = sBegin;
            length = sEnd - sBegin + 1;
            if (begin != 0) {  // HOST :179-184
                const int nbgn = max(begin - (a.W - 1), 0);
                length += begin - nbgn;
                begin = nbgn;
            }
        }
        a.begin[b] = begin;
        a.length[b] = length;
    }
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:332-369; windowLength < 1 and a
// resetInterval of 0 or below -1 (which make the HOST loop divide by zero or never advance) are
// rejected with RPP_ERROR_INVALID_ARGUMENTS.
RPPA_EXPORT RppStatus rppt_non_silent_region_detection_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr,
                                                           Rpp32s *srcLengthTensor, Rpp32s *detectedIndexTensor,
                                                           Rpp32s *detectionLengthTensor, Rpp32f cutOffDB,
                                                           Rpp32s windowLength, Rpp32f referencePower,
                                                           Rpp32s resetInterval, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->numDims - 1 != 1) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !srcLengthTensor || !detectedIndexTensor || !detectionLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    if (windowLength < 1 || resetInterval == 0 || resetInterval < -1) return RPP_ERROR_INVALID_ARGUMENTS;
    const int B = static_cast<int>(srcDescPtr->n);
    if (B == 0) return RPP_SUCCESS;
    const long long nStride = srcDescPtr->strides.nStride;
    if (nStride < 1 || nStride > 0x7fffffffLL) return RPP_ERROR_INVALID_ARGUMENTS;

    rppa::NsrArgs a{};
    a.src = static_cast<const float *>(srcPtr);
    a.nStride = nStride;
    a.len = srcLengthTensor;
    a.W = windowLength;
    a.meanFactor = 1.0f / static_cast<float>(windowLength);
    a.R = resetInterval;
    a.maxBlocks = resetInterval <= 0 ? 1 : static_cast<int>((nStride + resetInterval - 1) / resetInterval);
    a.cutOff = std::pow(10.0f, cutOffDB * 0.1f);
    a.refPower = referencePower;
    a.begin = detectedIndexTensor;
    a.length = detectionLengthTensor;
    a.vec = (reinterpret_cast<uintptr_t>(srcPtr) & 15) == 0 && nStride % 4 == 0 && (resetInterval <= 0 || resetInterval % 4 == 0) &&
            static_cast<long long>(B) * nStride * 4 < (1LL << 31) ? 1 : 0;
    a.srcBytes = a.vec ? static_cast<int>(static_cast<long long>(B) * nStride * 4) : 0;
    a.segsPerBlock = static_cast<int>(((resetInterval <= 0 ? nStride : resetInterval) + rppa::kSeg - 1) / rppa::kSeg);
    a.maxSegs = resetInterval <= 0 ? a.segsPerBlock : a.maxBlocks * a.segsPerBlock;
    const size_t cells = static_cast<size_t>(B) * a.maxBlocks;
    const size_t segCells = static_cast<size_t>(B) * a.maxSegs;
    char *ws = static_cast<char *>(rppa::workspace(rppHandle, segCells * 8 + cells * 4));
    if (!ws) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    a.segMax = reinterpret_cast<float *>(ws);
    a.segSum = reinterpret_cast<float *>(ws + segCells * 4);
    a.blockMax = reinterpret_cast<float *>(ws + segCells * 8);

    if (a.vec && B <= 16383) {
        const unsigned g1 = static_cast<unsigned>((cells + rppa::kQuadRows * rppa::kQuadWaves - 1) / (rppa::kQuadRows * rppa::kQuadWaves));
        hipLaunchKernelGGL(rppa::nsr_blocks_quad_kernel, dim3(g1), dim3(64 * rppa::kQuadWaves), sizeof(int) * (B + 1), rppHandle->stream, a, B);
    } else {
        const unsigned g1 = static_cast<unsigned>((cells + rppa::kPass1Rows - 1) / rppa::kPass1Rows);
        hipLaunchKernelGGL(rppa::nsr_blocks_kernel, dim3(g1), dim3(rppa::kPass1Threads), 0, rppHandle->stream, a, B);
    }
    RPPA_HIP(hipGetLastError());
    hipLaunchKernelGGL(rppa::nsr_region_kernel, dim3(B), dim3(rppa::kPass2Threads), 0, rppHandle->stream, a);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
