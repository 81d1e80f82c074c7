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
// Decoding: one workgroup of 1024 threads per request; a launch advances every request of a
// task table by its own range of positions, so a serving loop can let requests join and leave
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

// part[s * O + o] = sum over the s-th 1/S of the input words of w(o, .) . act, for a packed
// [I4][O] matrix and an activation of I4 packed words in LDS. O * S is a multiple of MT.
template <int O, int I4, int S>
__device__ __forceinline__ void matvec_partials(const int32_t *__restrict__ Wp, const int32_t *act,
                                                int32_t *part, int tid) {
    constexpr int WPS = I4 / S;
    static_assert((O * S) % MT == 0 && I4 % S == 0, "matvec split");
#pragma unroll 1
    for (int idx = tid; idx < O * S; idx += MT) {
        const int o = idx % O, s = idx / O;
        const int32_t *col = Wp + (uint64_t)(s * WPS) * O + o;
        const int32_t *a = act + s * WPS;
        int32_t acc = 0;
#pragma unroll 16
        for (int w = 0; w < WPS; w++) acc = __builtin_amdgcn_sdot4(col[(uint64_t)w * O], a[w], acc, false);
        part[idx] = acc;
    }
}

__global__ void __launch_bounds__(MT)
k_decode(const int8_t *__restrict__ W, const DecodeTask *__restrict__ tasks) {
    const DecodeTask task = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const uint32_t tcap = (task.total + 3) & ~3u;
    const int32_t *embed = (const int32_t *)(W + MODEL_OFF_EMBED);
    const int32_t *wout = (const int32_t *)(W + MODEL_OFF_OUT);
    int32_t *kt = task.kv;                                     // [L][64][tcap]
    int32_t *vp = task.kv + (uint64_t)N_LAYERS * 64 * tcap;    // [L][tcap / 4][256]

    __shared__ int32_t x[D_MODEL];
    __shared__ int32_t a_pk[D_MODEL / 4], q_pk[D_MODEL / 4], k_pk[D_MODEL / 4], att_pk[D_MODEL / 4];
    __shared__ int32_t h_pk[D_FF / 4];
    __shared__ int32_t part[4096], partn[MT];
    __shared__ int32_t scores[N_HEADS][MAX_TOKENS];
    __shared__ int32_t red_v[MT];
    __shared__ uint32_t red_i[MT];
    __shared__ uint32_t cur;
    int8_t *a8 = (int8_t *)a_pk, *q8 = (int8_t *)q_pk, *k8 = (int8_t *)k_pk, *att8 = (int8_t *)att_pk,
           *h8 = (int8_t *)h_pk;

    if (tid == 0)
        cur = task.start < task.n_prompt ? task.prompt[task.start]
                                         : (uint32_t)task.wit[(uint64_t)(task.start - 1) * WIT_STRIDE + WIT_NEXT];
    __syncthreads();
    for (uint32_t t = task.start; t < task.end; t++) {
        const uint32_t tok = cur;
        uint64_t *wt = task.wit + (uint64_t)t * WIT_STRIDE;
        if (tid == 0) task.tokens[t] = tok;
        if (tid < D_MODEL / 4) {
            const int32_t e = embed[(uint64_t)tok * (D_MODEL / 4) + tid];
            for (int b = 0; b < 4; b++) x[4 * tid + b] = (int8_t)(e >> (8 * b));
        }
        __syncthreads();
        for (int l = 0; l < N_LAYERS; l++) {
            const int8_t *Lw = W + MODEL_OFF_LAYER + (uint64_t)l * LAYER_BYTES;
            uint64_t *wl = wt + (uint64_t)l * WIT_LAYER;
            int32_t *ktl = kt + (uint64_t)l * 64 * tcap;
            int32_t *vpl = vp + (uint64_t)l * (tcap / 4) * D_MODEL;
            // a = rq(x)
            if (tid < D_MODEL) {
                const int32_t av = clamp127(x[tid] >> 2);
                a8[tid] = (int8_t)av;
                wl[tid] = from_i64(av);
            }
            __syncthreads();
            // q, k, v
            matvec_partials<3 * D_MODEL, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_QKV), a_pk, part, tid);
            __syncthreads();
            if (tid < 3 * D_MODEL) {
                const int32_t y = part[tid] + part[768 + tid] + part[1536 + tid] + part[2304 + tid];
                const int32_t v = clamp127(y >> 8);
                wl[D_MODEL + tid] = from_i64(v);  // q, k, v are consecutive in the witness
                if (tid < D_MODEL) {
                    q8[tid] = (int8_t)v;
                } else if (tid < 2 * D_MODEL) {
                    k8[tid - D_MODEL] = (int8_t)v;
                } else {
                    const int i = tid - 2 * D_MODEL;
                    ((int8_t *)vpl)[((uint64_t)(t / 4) * D_MODEL + i) * 4 + (t & 3)] = (int8_t)v;
                }
            }
            __syncthreads();
            if (tid < 64) ktl[(uint64_t)tid * tcap + t] = k_pk[tid];
            __threadfence_block();
            __syncthreads();
            // scores s_u = relu((q . k_u) >> 6), u <= t, per head; zeros up to a multiple of 4
            const uint32_t n = t + 1, n4 = (n + 3) & ~3u;
            for (uint32_t e = tid; e < N_HEADS * n; e += MT) {
                const uint32_t hd = e / n, u = e % n;
                const int32_t *kc = ktl + (uint64_t)(hd * 16) * tcap + u;
                int32_t acc = 0;
#pragma unroll
                for (int w = 0; w < 16; w++) acc = __builtin_amdgcn_sdot4(q_pk[hd * 16 + w], kc[(uint64_t)w * tcap], acc, false);
                const int32_t s = acc >> 6;
                scores[hd][u] = s > 0 ? s : 0;
            }
            if (tid < 16) {
                const uint32_t u = n + (tid & 3);
                if (u < n4) scores[tid >> 2][u] = 0;
            }
            __syncthreads();
            // o_i = (sum_u s_u v_u(i)) / (1 + sum_u s_u): element i = tid % 256, quarter s = tid / 256
            {
                const int i = tid & (D_MODEL - 1), s = tid >> 8, hd = i / D_HEAD;
                const int32_t *vc = vpl + i;
                int32_t num = 0, nsum = 0;
                for (uint32_t wd = s; wd < n4 / 4; wd += 4) {
                    const int32_t word = vc[(uint64_t)wd * D_MODEL];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int32_t sc = scores[hd][4 * wd + b];
                        num += sc * (int32_t)(int8_t)(word >> (8 * b));
                        nsum += sc;
                    }
                }
                part[tid] = num;
                partn[tid] = nsum;
            }
            __syncthreads();
            if (tid < D_MODEL) {
                const int32_t num = part[tid] + part[256 + tid] + part[512 + tid] + part[768 + tid];
                const int32_t norm = 1 + partn[tid] + partn[256 + tid] + partn[512 + tid] + partn[768 + tid];
                const int32_t o = clamp127(num / norm);
                att8[tid] = (int8_t)o;
                wl[4 * D_MODEL + tid] = from_i64(o);
            }
            __syncthreads();
            // x += Wo att; m = rq(x)
            matvec_partials<D_MODEL, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_O), att_pk, part, tid);
            __syncthreads();
            if (tid < D_MODEL) {
                x[tid] += part[tid] + part[256 + tid] + part[512 + tid] + part[768 + tid];
                const int32_t mv = clamp127(x[tid] >> 2);
                a8[tid] = (int8_t)mv;
                wl[5 * D_MODEL + tid] = from_i64(mv);
            }
            __syncthreads();
            // h = clamp(relu((W1 m) >> 8), 0, 127)
            matvec_partials<D_FF, D_MODEL / 4, 4>((const int32_t *)(Lw + DOFF_W1), a_pk, part, tid);
            __syncthreads();
            {
                int32_t hv = (part[tid] + part[1024 + tid] + part[2048 + tid] + part[3072 + tid]) >> 8;
                hv = hv < 0 ? 0 : (hv > 127 ? 127 : hv);
                h8[tid] = (int8_t)hv;
                wl[6 * D_MODEL + tid] = from_i64(hv);
            }
            __syncthreads();
            // x += (W2 h) >> 8
            matvec_partials<D_MODEL, D_FF / 4, 16>((const int32_t *)(Lw + DOFF_W2), h_pk, part, tid);
            __syncthreads();
            if (tid < D_MODEL) {
                int32_t y = 0;
#pragma unroll
                for (int s = 0; s < 16; s++) y += part[s * 256 + tid];
                x[tid] += y >> 8;
                wl[6 * D_MODEL + D_FF + tid] = from_i64(x[tid]);
            }
            __syncthreads();
        }
        // f = rq(x); logits and the greedy choice
        if (tid < D_MODEL) {
            const int32_t fv = clamp127(x[tid] >> 2);
            a8[tid] = (int8_t)fv;
            wt[N_LAYERS * WIT_LAYER + tid] = from_i64(fv);
        }
        __syncthreads();
        int32_t best = INT32_MIN;
        uint32_t best_i = 0;
#pragma unroll 1
        for (int j = tid; j < VOCAB; j += MT) {  // j increases: the first maximum is kept
            const int32_t *col = wout + j;
            int32_t acc = 0;
#pragma unroll 16
            for (int w = 0; w < D_MODEL / 4; w++) acc = __builtin_amdgcn_sdot4(col[(uint64_t)w * VOCAB], a_pk[w], acc, false);
            if (acc > best) {
                best = acc;
                best_i = j;
            }
        }
        red_v[tid] = best;
        red_i[tid] = best_i;
        __syncthreads();
        for (int s = MT / 2; s > 0; s >>= 1) {
            if (tid < s) {
                const int32_t ov = red_v[tid + s];
                const uint32_t oi = red_i[tid + s];
                if (ov > red_v[tid] || (ov == red_v[tid] && oi < red_i[tid])) {
                    red_v[tid] = ov;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            const uint32_t next = red_i[0];
            wt[WIT_NEXT] = next;
            cur = (t + 1 < task.n_prompt) ? task.prompt[t + 1] : next;
        }
        __syncthreads();
    }
}

Status model_weights(Ctx *ctx, int8_t *w, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_model_weights, dim3((unsigned)((MODEL_BYTES + 255) / 256)), dim3(256), 0, s,
                       w, (uint64_t)MODEL_BYTES, seed);
    return check_launch("model weights");
}

Status model_decode(Ctx *ctx, const int8_t *w, const DecodeTask *tasks, uint32_t count,
                    hipStream_t s) {
    if (count == 0) return Status::ok();
    KernelTimer timer(ctx, "k_decode", s);
    hipLaunchKernelGGL(k_decode, dim3(count), dim3(MT), 0, s, w, tasks);
    return check_launch("decode");
}

}  // namespace tsr
