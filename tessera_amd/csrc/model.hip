// A synthetic integer transformer for served proving: 4 layers, d_model 256, 4 heads of 64,
// d_ff 1024, vocabulary 4096, int8 weights uniform in [-127, 127], int32 accumulation,
// exact greedy decoding. There is no counterpart in the reference (its models come from
// crates/tessera-witness); this one only has to be deterministic and exact, so that a request's
// witness, and hence its commitment and proof, are the same whenever and beside whatever it runs.
//
// Per token t, with x the int32 residual stream (x = Embed[token] at the start):
//   for each layer: a = rq(x); q, k, v = rq8(W a) (q, k, v int8, k and v cached);
//     per head: s_u = relu((q . k_u) >> 6) for u <= t; o = (sum_u s_u v_u) / (1 + sum_u s_u)
//     (C division); att = clamp(o); x += Wo att; m = rq(x); h = clamp(relu((W1 m) >> 8), 0, 127);
//     x += (W2 h) >> 8
//   f = rq(x); logits = Wout f; next = argmax (lowest index on ties)
// with rq(x) = clamp(x >> 2, -127, 127) and rq8(y) = clamp(y >> 8, -127, 127) (arithmetic shifts).
// Integer sums are exact in any order, so the products are split over the inputs ("split-K")
// and the partial sums added in LDS. Bounds: |q . k| <= 64 * 127^2, so s_u <= 16129 and
// |sum_u s_u v_u| <= 512 * 16129 * 127 < 2^31 (int32 is exact).
//
// The witness of a token is written as Goldilocks elements (Goldilocks::from_i64,
// crates/tessera-field/src/goldilocks.rs:374: x >= 0 -> x, x < 0 -> p - |x|), WIT_STRIDE values:
// per layer a, q, k, v, att, m (256 each), h (1024), x (256); then f (256) and the next token.
//
// Decoding: one workgroup of 1024 threads per group of R requests (R = DECODE_GROUP), whose
// matrix products read each weight word once for the group; a launch advances every request of
// a task table by its own range of positions, so a serving loop can let requests join and leave
// between launches (iteration-level batching) without changing any request's arithmetic.
// Weight reads are packed (model.h) so each wavefront load is 256 contiguous bytes; the 7.3 MB
// of weights are shared by all requests and stay in the L2 / last-level cache.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "field.h"
#include "model.h"

namespace tsr {

constexpr int MT = 1024;  // threads of a decode workgroup

__device__ __forceinline__ int32_t clamp127(int32_t v) { return v < -127 ? -127 : (v > 127 ? 127 : v); }

__device__ __forceinline__ uint64_t from_i64(int64_t v) {
    return v >= 0 ? (uint64_t)v : gl::P - (uint64_t)(-v);
}

// SplitMix64 of a counter, for the weights: w = (z mod 255) - 127.
__device__ __forceinline__ int8_t weight_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (int8_t)((int)(z % 255) - 127);
}

// Offset, within its matrix, of byte r of a packed [in / 4][out] matrix: the logical o * in + i.
__device__ __forceinline__ uint64_t unpack(uint64_t r, uint64_t in, uint64_t out, uint64_t *o_out) {
    const uint64_t w = r >> 2, i = 4 * (w / out) + (r & 3), o = w % out;
    *o_out = o;
    return i;
}

// The generator index of device byte b (model.h: device layout -> logical layout).
__device__ uint64_t logical_index(uint64_t b) {
    uint64_t o, i;
    if (b < MODEL_OFF_LAYER) return b;
    if (b >= MODEL_OFF_OUT) {
        i = unpack(b - MODEL_OFF_OUT, D_MODEL, VOCAB, &o);
        return MODEL_OFF_OUT + o * D_MODEL + i;
    }
    const uint64_t l = (b - MODEL_OFF_LAYER) / LAYER_BYTES, r = (b - MODEL_OFF_LAYER) % LAYER_BYTES;
    const uint64_t base = MODEL_OFF_LAYER + l * LAYER_BYTES;
    if (r < DOFF_O) {
        i = unpack(r, D_MODEL, 3 * D_MODEL, &o);
        return base + LOFF_Q + (o / D_MODEL) * D_MODEL * D_MODEL + (o % D_MODEL) * D_MODEL + i;
    }
    if (r < DOFF_W1) {
        i = unpack(r - DOFF_O, D_MODEL, D_MODEL, &o);
        return base + LOFF_O + o * D_MODEL + i;
    }
    if (r < DOFF_W2) {
        i = unpack(r - DOFF_W1, D_MODEL, D_FF, &o);
        return base + LOFF_W1 + o * D_MODEL + i;
    }
    i = unpack(r - DOFF_W2, D_FF, D_MODEL, &o);
    return base + LOFF_W2 + o * D_FF + i;
}

__global__ void k_model_weights(int8_t *__restrict__ w, uint64_t n, uint64_t seed) {
    uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) w[b] = weight_at(seed, logical_index(b));
}

// part[j][s * O + o] = sum over the s-th 1/S of the input words of w(o, .) . act[j], for a packed
// [I4][O] matrix and R activations of I4 packed words in LDS: each weight word is read once for
// the R requests of the workgroup. O * S is a multiple of MT.
template <int R, int O, int I4, int S>
__device__ __forceinline__ void matvec_partials(const int32_t *__restrict__ Wp, const int32_t (*act)[D_MODEL],
                                                int32_t (*part)[4096], int tid) {
    constexpr int WPS = I4 / S;
    static_assert((O * S) % MT == 0 && I4 % S == 0 && O * S <= 4096, "matvec split");
#pragma unroll 1
    for (int idx = tid; idx < O * S; idx += MT) {
        const int o = idx % O, s = idx / O;
        const int32_t *col = Wp + (uint64_t)(s * WPS) * O + o;
        int32_t acc[R];
#pragma unroll
        for (int j = 0; j < R; j++) acc[j] = 0;
#pragma unroll 16
        for (int w = 0; w < WPS; w++) {
            const int32_t word = col[(uint64_t)w * O];
#pragma unroll
            for (int j = 0; j < R; j++) acc[j] = __builtin_amdgcn_sdot4(word, act[j][s * WPS + w], acc[j], false);
        }
#pragma unroll
        for (int j = 0; j < R; j++) part[j][idx] = acc[j];
    }
}

// (value, index) pairs: the larger value wins, the lower index on equal values (the first maximum).
__device__ __forceinline__ void argmax_merge(int32_t &v, uint32_t &i, int32_t ov, uint32_t oi) {
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// One workgroup advances the R tasks tasks[R b .. R b + R) (those below `count`), each by its own
// range of positions, in lockstep: step s is position start + s of every task that has one. The
// weight reads of every matrix product are shared by the R tasks; attention, the KV caches and the
// witness are each task's own. A task's arithmetic is the same whatever R and its neighbours are.
template <int R>
__global__ void __launch_bounds__(MT)
k_decode(const int8_t *__restrict__ W, const DecodeTask *__restrict__ tasks, uint32_t count) {
    const int tid = threadIdx.x;
    const int32_t *embed = (const int32_t *)(W + MODEL_OFF_EMBED);
    const int32_t *wout = (const int32_t *)(W + MODEL_OFF_OUT);

    __shared__ int32_t x[R][D_MODEL];
    // packed int8 activations: a (also m and f), q, k, att of D_MODEL / 4 words; h of D_FF / 4
    __shared__ int32_t a_pk[R][D_MODEL], q_pk[R][D_MODEL / 4], k_pk[R][D_MODEL / 4], att_pk[R][D_MODEL];
    __shared__ int32_t h_pk[R][D_MODEL];
    __shared__ int32_t part[R][4096], partn[MT];
    __shared__ int32_t scores[R][N_HEADS][MAX_TOKENS];
    __shared__ int32_t red_v[R][MT / 64];
    __shared__ uint32_t red_i[R][MT / 64];
    __shared__ uint32_t cur[R];

    // the tasks and each step's facts live in LDS (uniform reads), not in every lane's registers
    __shared__ DecodeTask task[R];
    __shared__ uint32_t steps_s, t[R];
    __shared__ int on[R];
    if (tid < R) {
        const uint32_t at = blockIdx.x * R + tid;
        DecodeTask tk = tasks[at < count ? at : 0];
        if (at >= count) tk.end = tk.start;  // no position: never on
        task[tid] = tk;
        cur[tid] = tk.end == tk.start ? 0
                   : tk.start < tk.n_prompt ? tk.prompt[tk.start]
                                            : (uint32_t)tk.wit[(uint64_t)(tk.start - 1) * WIT_STRIDE + WIT_NEXT];
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t m = 0;
        for (int j = 0; j < R; j++) m = task[j].end - task[j].start > m ? task[j].end - task[j].start : m;
        steps_s = m;
    }
    __syncthreads();
    const uint32_t steps = steps_s;
    for (uint32_t st = 0; st < steps; st++) {
        if (tid < R) {
            t[tid] = task[tid].start + st;
            on[tid] = t[tid] < task[tid].end;
        }
        __syncthreads();
        auto wit_at = [&](int j) { return task[j].wit + (uint64_t)t[j] * WIT_STRIDE; };
        for (int e = tid; e < R * (D_MODEL / 4); e += MT) {
            const int j = e / (D_MODEL / 4), w = e % (D_MODEL / 4);
            if (!on[j]) continue;
            const int32_t v = embed[(uint64_t)cur[j] * (D_MODEL / 4) + w];
            for (int b = 0; b < 4; b++) x[j][4 * w + b] = (int8_t)(v >> (8 * b));
            if (w == 0) task[j].tokens[t[j]] = cur[j];
        }
        __syncthreads();
        for (int l = 0; l < N_LAYERS; l++) {
            const int8_t *Lw = W + MODEL_OFF_LAYER + (uint64_t)l * LAYER_BYTES;
            // a = rq(x)
            for (int e = tid; e < R * D_MODEL; e += MT) {
                const int j = e / D_MODEL, i = e % D_MODEL;
                if (!on[j]) continue;
                const int32_t av = clamp127(x[j][i] >> 2);
                ((int8_t *)a_pk[j])[i] = (int8_t)av;
                wit_at(j)[(uint64_t)l * WIT_LAYER + i] = from_i64(av);
            }
            __syncthreads();
            // q, k, v
            matvec_partials<R, 3 * D_MODEL, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_QKV), a_pk, part, tid);
            __syncthreads();
            for (int e = tid; e < R * 3 * D_MODEL; e += MT) {
                const int j = e / (3 * D_MODEL), o = e % (3 * D_MODEL);
                if (!on[j]) continue;
                const int32_t y = part[j][o] + part[j][768 + o] + part[j][1536 + o] + part[j][2304 + o];
                const int32_t v = clamp127(y >> 8);
                wit_at(j)[(uint64_t)l * WIT_LAYER + D_MODEL + o] = from_i64(v);  // q, k, v in a row
                if (o < D_MODEL) {
                    ((int8_t *)q_pk[j])[o] = (int8_t)v;
                } else if (o < 2 * D_MODEL) {
                    ((int8_t *)k_pk[j])[o - D_MODEL] = (int8_t)v;
                } else {
                    const uint32_t tcap = kv_tcap(task[j].total);
                    int32_t *vpl = task[j].kv + (uint64_t)N_LAYERS * 64 * tcap + (uint64_t)l * (tcap / 4) * D_MODEL;
                    const int i = o - 2 * D_MODEL;
                    ((int8_t *)vpl)[((uint64_t)(t[j] / 4) * D_MODEL + i) * 4 + (t[j] & 3)] = (int8_t)v;
                }
            }
            __syncthreads();
            for (int e = tid; e < R * 64; e += MT) {
                const int j = e / 64, w = e % 64;
                if (!on[j]) continue;
                const uint32_t tcap = kv_tcap(task[j].total);
                task[j].kv[((uint64_t)l * 64 + w) * tcap + t[j]] = k_pk[j][w];
            }
            __threadfence_block();
            __syncthreads();
            // scores s_u = relu((q . k_u) >> 6), u <= t, per head; zeros up to a multiple of 4
#pragma unroll
            for (int j = 0; j < R; j++) {
                if (!on[j]) continue;
                const uint32_t tcap = kv_tcap(task[j].total), n = t[j] + 1, n4 = (n + 3) & ~3u;
                const int32_t *ktl = task[j].kv + (uint64_t)l * 64 * tcap;
                for (uint32_t e = tid; e < N_HEADS * n; e += MT) {
                    const uint32_t hd = e / n, u = e % n;
                    const int32_t *kc = ktl + (uint64_t)(hd * 16) * tcap + u;
                    int32_t acc = 0;
#pragma unroll
                    for (int w = 0; w < 16; w++) acc = __builtin_amdgcn_sdot4(q_pk[j][hd * 16 + w], kc[(uint64_t)w * tcap], acc, false);
                    const int32_t sc = acc >> 6;
                    scores[j][hd][u] = sc > 0 ? sc : 0;
                }
                if (tid < 16) {
                    const uint32_t u = n + (tid & 3);
                    if (u < n4) scores[j][tid >> 2][u] = 0;
                }
            }
            __syncthreads();
            // o_i = (sum_u s_u v_u(i)) / (1 + sum_u s_u): element i = tid % 256, quarter q = tid / 256
#pragma unroll
            for (int j = 0; j < R; j++) {
                if (!on[j]) continue;
                const uint32_t tcap = kv_tcap(task[j].total), n4 = (t[j] + 1 + 3) & ~3u;
                const int32_t *vpl = task[j].kv + (uint64_t)N_LAYERS * 64 * tcap + (uint64_t)l * (tcap / 4) * D_MODEL;
                const int i = tid & (D_MODEL - 1), q = tid >> 8, hd = i / D_HEAD;
                const int32_t *vc = vpl + i;
                int32_t num = 0, nsum = 0;
                // four cache words in flight per thread: a request of 512 positions decodes in
                // 36 ms instead of 41 (unrolled 8 or 16, or with the score loop unrolled too: no
                // faster)
#pragma unroll 4
                for (uint32_t wd = q; wd < n4 / 4; wd += 4) {
                    const int32_t word = vc[(uint64_t)wd * D_MODEL];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int32_t sc = scores[j][hd][4 * wd + b];
                        num += sc * (int32_t)(int8_t)(word >> (8 * b));
                        nsum += sc;
                    }
                }
                part[0][tid] = num;
                partn[tid] = nsum;
                __syncthreads();
                if (tid < D_MODEL) {
                    const int32_t nm = part[0][tid] + part[0][256 + tid] + part[0][512 + tid] + part[0][768 + tid];
                    const int32_t norm = 1 + partn[tid] + partn[256 + tid] + partn[512 + tid] + partn[768 + tid];
                    const int32_t o = clamp127(nm / norm);
                    ((int8_t *)att_pk[j])[tid] = (int8_t)o;
                    wit_at(j)[(uint64_t)l * WIT_LAYER + 4 * D_MODEL + tid] = from_i64(o);
                }
                __syncthreads();
            }
            // x += Wo att; m = rq(x)
            matvec_partials<R, D_MODEL, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_O), att_pk, part, tid);
            __syncthreads();
            for (int e = tid; e < R * D_MODEL; e += MT) {
                const int j = e / D_MODEL, i = e % D_MODEL;
                if (!on[j]) continue;
                x[j][i] += part[j][i] + part[j][256 + i] + part[j][512 + i] + part[j][768 + i];
                const int32_t mv = clamp127(x[j][i] >> 2);
                ((int8_t *)a_pk[j])[i] = (int8_t)mv;
                wit_at(j)[(uint64_t)l * WIT_LAYER + 5 * D_MODEL + i] = from_i64(mv);
            }
            __syncthreads();
            // h = clamp(relu((W1 m) >> 8), 0, 127)
            matvec_partials<R, D_FF, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_W1), a_pk, part, tid);
            __syncthreads();
            for (int e = tid; e < R * D_FF; e += MT) {
                const int j = e / D_FF, i = e % D_FF;
                if (!on[j]) continue;
                int32_t hv = (part[j][i] + part[j][1024 + i] + part[j][2048 + i] + part[j][3072 + i]) >> 8;
                hv = hv < 0 ? 0 : (hv > 127 ? 127 : hv);
                ((int8_t *)h_pk[j])[i] = (int8_t)hv;
                wit_at(j)[(uint64_t)l * WIT_LAYER + 6 * D_MODEL + i] = from_i64(hv);
            }
            __syncthreads();
            // x += (W2 h) >> 8
            matvec_partials<R, D_MODEL, D_FF / 4, 16>((const int32_t *)(Lw + DOFF_W2), h_pk, part, tid);
            __syncthreads();
            for (int e = tid; e < R * D_MODEL; e += MT) {
                const int j = e / D_MODEL, i = e % D_MODEL;
                if (!on[j]) continue;
                int32_t y = 0;
#pragma unroll
                for (int q = 0; q < 16; q++) y += part[j][q * 256 + i];
                x[j][i] += y >> 8;
                wit_at(j)[(uint64_t)l * WIT_LAYER + 6 * D_MODEL + D_FF + i] = from_i64(x[j][i]);
            }
            __syncthreads();
        }
        // f = rq(x); logits and the greedy choice
        for (int e = tid; e < R * D_MODEL; e += MT) {
            const int j = e / D_MODEL, i = e % D_MODEL;
            if (!on[j]) continue;
            const int32_t fv = clamp127(x[j][i] >> 2);
            ((int8_t *)a_pk[j])[i] = (int8_t)fv;
            wit_at(j)[N_LAYERS * WIT_LAYER + i] = from_i64(fv);
        }
        __syncthreads();
        int32_t best[R];
        uint32_t best_i[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            best[j] = INT32_MIN;
            best_i[j] = 0;
        }
#pragma unroll 1
        for (int v = tid; v < VOCAB; v += MT) {  // v increases: a thread keeps its first maximum
            const int32_t *col = wout + v;
            int32_t acc[R];
#pragma unroll
            for (int j = 0; j < R; j++) acc[j] = 0;
#pragma unroll 16
            for (int w = 0; w < D_MODEL / 4; w++) {
                const int32_t word = col[(uint64_t)w * VOCAB];
#pragma unroll
                for (int j = 0; j < R; j++) acc[j] = __builtin_amdgcn_sdot4(word, a_pk[j][w], acc[j], false);
            }
#pragma unroll
            for (int j = 0; j < R; j++)
                if (acc[j] > best[j]) {
                    best[j] = acc[j];
                    best_i[j] = v;
                }
        }
        const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
        for (int j = 0; j < R; j++) {
            int32_t bv = best[j];
            uint32_t bi = best_i[j];
            for (int m = 32; m >= 1; m >>= 1) argmax_merge(bv, bi, __shfl_xor(bv, m, 64), __shfl_xor(bi, m, 64));
            if (lane == 0) {
                red_v[j][wave] = bv;
                red_i[j][wave] = bi;
            }
        }
        __syncthreads();
        if (tid < R && on[tid]) {
            const int j = tid;
            int32_t bv = red_v[j][0];
            uint32_t bi = red_i[j][0];
            for (int w = 1; w < MT / 64; w++) argmax_merge(bv, bi, red_v[j][w], red_i[j][w]);
            wit_at(j)[WIT_NEXT] = bi;
            cur[j] = (t[j] + 1 < task[j].n_prompt) ? task[j].prompt[t[j] + 1] : bi;
        }
        __syncthreads();
    }
}

Status model_weights(Ctx *ctx, int8_t *w, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_model_weights, dim3((unsigned)((MODEL_BYTES + 255) / 256)), dim3(256), 0, s,
                       w, (uint64_t)MODEL_BYTES, seed);
    return check_launch("model weights");
}

// Requests per decode workgroup. One: measured on the served workload (256 requests, up to 73
// in flight), four requests per workgroup took a request-token from 69 to 110 us and served
// capacity from 2014 to 591 proofs/s. A token's time is the chain of dependent weight loads
// from the last-level cache and the barriers between the steps, which grow with the requests a
// workgroup carries, while the shared weight reads saved little; and four per workgroup leave
// three quarters of the CUs idle when fewer requests than CUs are decoding.
constexpr uint32_t DECODE_GROUP = 1;

Status model_decode(Ctx *ctx, const int8_t *w, const DecodeTask *tasks, uint32_t count,
                    hipStream_t s) {
    if (count == 0) return Status::ok();
    KernelTimer timer(ctx, "k_decode", s);
    hipLaunchKernelGGL(k_decode<DECODE_GROUP>, dim3((count + DECODE_GROUP - 1) / DECODE_GROUP), dim3(MT), 0, s,
                       w, tasks, count);
    return check_launch("decode");
}

}  // namespace tsr
