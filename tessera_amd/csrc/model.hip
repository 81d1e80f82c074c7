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
// Every dot product of int8 vectors is v_dot4_i32_i8 on packed bytes.
//
// The witness of a token is written as Goldilocks elements (Goldilocks::from_i64,
// crates/tessera-field/src/goldilocks.rs:374: x >= 0 -> x, x < 0 -> p - |x|), WIT_STRIDE values:
// per layer a, q, k, v, att, m (256 each), h (1024), x (256); then f (256) and the next token.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "field.h"
#include "model.h"

namespace tsr {

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

__global__ void k_model_weights(int8_t *__restrict__ w, uint64_t n, uint64_t seed) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = weight_at(seed, i);
}

// dot of a 256-byte int8 weight row with a 256-byte int8 vector in LDS (64 packed words).
__device__ __forceinline__ int32_t dot256(const int8_t *__restrict__ row, const int32_t *packed) {
    const int32_t *r = (const int32_t *)row;
    int32_t acc = 0;
#pragma unroll 16
    for (int i = 0; i < 64; i++) acc = __builtin_amdgcn_sdot4(r[i], packed[i], acc, false);
    return acc;
}

// dot of a 1024-byte row with a 1024-byte vector (256 packed words).
__device__ __forceinline__ int32_t dot1024(const int8_t *__restrict__ row, const int32_t *packed) {
    const int32_t *r = (const int32_t *)row;
    int32_t acc = 0;
#pragma unroll 16
    for (int i = 0; i < 256; i++) acc = __builtin_amdgcn_sdot4(r[i], packed[i], acc, false);
    return acc;
}

constexpr int MT = 256;  // threads of a decode workgroup

// One request's whole decode: prompt tokens prompt[0..n_prompt), generation up to `total`
// tokens; tokens[t] and the witness of every position written out. The KV cache (layers x total
// x 2 x 256 bytes) lives at kv.
__global__ void __launch_bounds__(MT)
k_decode(const int8_t *__restrict__ W, const uint32_t *__restrict__ prompt, uint32_t n_prompt,
         uint32_t total, uint32_t *__restrict__ tokens, int8_t *__restrict__ kv,
         uint64_t *__restrict__ wit) {
    const int8_t *embed = W + MODEL_OFF_EMBED;
    const int8_t *wout = W + MODEL_OFF_OUT;
    __shared__ int32_t x[D_MODEL];
    __shared__ int32_t a_pk[D_MODEL / 4], att_pk[D_MODEL / 4], h_pk[D_FF / 4];
    __shared__ int8_t a8[D_MODEL], att8[D_MODEL], h8[D_FF];
    __shared__ int8_t q8[D_MODEL];
    __shared__ int32_t scores[N_HEADS][MAX_TOKENS];
    __shared__ int32_t red_v[MT];
    __shared__ uint32_t red_i[MT];
    __shared__ uint32_t cur;
    const int tid = threadIdx.x;
    if (tid == 0) cur = prompt[0];
    __syncthreads();
    for (uint32_t t = 0; t < total; t++) {
        const uint32_t tok = cur;
        if (tid == 0) tokens[t] = tok;
        uint64_t *wt = wit + (uint64_t)t * WIT_STRIDE;
        x[tid] = embed[(uint64_t)tok * D_MODEL + tid];
        __syncthreads();
        for (int l = 0; l < N_LAYERS; l++) {
            const int8_t *L = W + MODEL_OFF_LAYER + (uint64_t)l * LAYER_BYTES;
            uint64_t *wl = wt + (uint64_t)l * WIT_LAYER;
            int8_t *kvl = kv + ((uint64_t)l * total) * 2 * D_MODEL;
            // a = rq(x)
            int32_t av = clamp127(x[tid] >> 2);
            a8[tid] = (int8_t)av;
            wl[0 * D_MODEL + tid] = from_i64(av);
            __syncthreads();
            if (tid < D_MODEL / 4) a_pk[tid] = ((const int32_t *)a8)[tid];
            __syncthreads();
            int32_t qv = clamp127(dot256(L + LOFF_Q + (uint64_t)tid * D_MODEL, a_pk) >> 8);
            int32_t kvv = clamp127(dot256(L + LOFF_K + (uint64_t)tid * D_MODEL, a_pk) >> 8);
            int32_t vv = clamp127(dot256(L + LOFF_V + (uint64_t)tid * D_MODEL, a_pk) >> 8);
            q8[tid] = (int8_t)qv;
            kvl[(uint64_t)t * 2 * D_MODEL + tid] = (int8_t)kvv;
            kvl[(uint64_t)t * 2 * D_MODEL + D_MODEL + tid] = (int8_t)vv;
            wl[1 * D_MODEL + tid] = from_i64(qv);
            wl[2 * D_MODEL + tid] = from_i64(kvv);
            wl[3 * D_MODEL + tid] = from_i64(vv);
            __threadfence_block();
            __syncthreads();
            // scores s_u = relu((q . k_u) >> 6) for u <= t, per head
            for (uint32_t e = tid; e < N_HEADS * (t + 1); e += MT) {
                const uint32_t hd = e / (t + 1), u = e % (t + 1);
                const int8_t *kr = kvl + (uint64_t)u * 2 * D_MODEL + hd * D_HEAD;
                const int32_t *kp = (const int32_t *)kr;
                const int32_t *qp = (const int32_t *)(q8 + hd * D_HEAD);
                int32_t acc = 0;
#pragma unroll
                for (int i = 0; i < D_HEAD / 4; i++) acc = __builtin_amdgcn_sdot4(qp[i], kp[i], acc, false);
                int32_t s = acc >> 6;
                scores[hd][u] = s > 0 ? s : 0;
            }
            __syncthreads();
            {
                // element tid of the attention output: head tid / 64
                const uint32_t hd = tid / D_HEAD;
                int64_t num = 0, norm = 1;
                for (uint32_t u = 0; u <= t; u++) {
                    const int64_t s = scores[hd][u];
                    num += s * (int64_t)kvl[(uint64_t)u * 2 * D_MODEL + D_MODEL + tid];
                    norm += s;
                }
                int32_t o = clamp127((int32_t)(num / norm));
                att8[tid] = (int8_t)o;
                wl[4 * D_MODEL + tid] = from_i64(o);
            }
            __syncthreads();
            if (tid < D_MODEL / 4) att_pk[tid] = ((const int32_t *)att8)[tid];
            __syncthreads();
            x[tid] += dot256(L + LOFF_O + (uint64_t)tid * D_MODEL, att_pk);
            __syncthreads();
            int32_t mv = clamp127(x[tid] >> 2);
            a8[tid] = (int8_t)mv;
            wl[5 * D_MODEL + tid] = from_i64(mv);
            __syncthreads();
            if (tid < D_MODEL / 4) a_pk[tid] = ((const int32_t *)a8)[tid];
            __syncthreads();
            for (int j = tid; j < D_FF; j += MT) {
                int32_t hv = dot256(L + LOFF_W1 + (uint64_t)j * D_MODEL, a_pk) >> 8;
                hv = hv < 0 ? 0 : (hv > 127 ? 127 : hv);
                h8[j] = (int8_t)hv;
                wl[6 * D_MODEL + j] = from_i64(hv);
            }
            __syncthreads();
            for (int j = tid; j < D_FF / 4; j += MT) h_pk[j] = ((const int32_t *)h8)[j];
            __syncthreads();
            x[tid] += dot1024(L + LOFF_W2 + (uint64_t)tid * D_FF, h_pk) >> 8;
            wl[6 * D_MODEL + D_FF + tid] = from_i64(x[tid]);
            __syncthreads();
        }
        // logits and the greedy choice
        int32_t fv = clamp127(x[tid] >> 2);
        a8[tid] = (int8_t)fv;
        wt[N_LAYERS * WIT_LAYER + tid] = from_i64(fv);
        __syncthreads();
        if (tid < D_MODEL / 4) a_pk[tid] = ((const int32_t *)a8)[tid];
        __syncthreads();
        int32_t best = INT32_MIN;
        uint32_t best_i = 0;
        for (int j = tid; j < VOCAB; j += MT) {
            int32_t lg = dot256(wout + (uint64_t)j * D_MODEL, a_pk);
            if (lg > best) {  // j increases: the first maximum is kept
                best = lg;
                best_i = j;
            }
        }
        red_v[tid] = best;
        red_i[tid] = best_i;
        __syncthreads();
        for (int s = MT / 2; s > 0; s >>= 1) {
            if (tid < s) {
                int32_t ov = red_v[tid + s];
                uint32_t oi = red_i[tid + s];
                if (ov > red_v[tid] || (ov == red_v[tid] && oi < red_i[tid])) {
                    red_v[tid] = ov;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            const uint32_t next = red_i[0];
            wt[N_LAYERS * WIT_LAYER + D_MODEL] = next;
            cur = (t + 1 < n_prompt) ? prompt[t + 1] : next;
        }
        __syncthreads();
    }
}

Status model_weights(Ctx *ctx, int8_t *w, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_model_weights, dim3((unsigned)((MODEL_BYTES + 255) / 256)), dim3(256), 0, s,
                       w, (uint64_t)MODEL_BYTES, seed);
    return check_launch("model weights");
}

Status model_decode(Ctx *ctx, const int8_t *w, const uint32_t *prompt, uint32_t n_prompt,
                    uint32_t total, uint32_t *tokens, int8_t *kv, uint64_t *wit, hipStream_t s) {
    KernelTimer timer(ctx, "k_decode", s);
    hipLaunchKernelGGL(k_decode, dim3(1), dim3(MT), 0, s, w, prompt, n_prompt, total, tokens, kv, wit);
    return check_launch("decode");
}

}  // namespace tsr
