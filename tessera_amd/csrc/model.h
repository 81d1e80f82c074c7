// The synthetic integer transformer of served proving (see model.hip): shapes, the logical
// weight layout (row-major [out][in] int8, the order the weights are generated in), the device
// layout the decoder reads, the KV cache layout and the witness layout.
#pragma once
#include <stdint.h>
#include "common.h"

namespace tsr {

constexpr int N_LAYERS = 4, D_MODEL = 256, N_HEADS = 4, D_HEAD = 64, D_FF = 1024, VOCAB = 4096;
constexpr int MAX_TOKENS = 512;

// Logical layout: weight w(o, i) of a matrix at `base` is generator word base + o * in + i.
constexpr uint64_t LOFF_Q = 0, LOFF_K = LOFF_Q + D_MODEL * D_MODEL, LOFF_V = LOFF_K + D_MODEL * D_MODEL,
                   LOFF_O = LOFF_V + D_MODEL * D_MODEL, LOFF_W1 = LOFF_O + D_MODEL * D_MODEL,
                   LOFF_W2 = LOFF_W1 + (uint64_t)D_FF * D_MODEL,
                   LAYER_BYTES = LOFF_W2 + (uint64_t)D_MODEL * D_FF;
constexpr uint64_t MODEL_OFF_EMBED = 0, MODEL_OFF_LAYER = (uint64_t)VOCAB * D_MODEL,
                   MODEL_OFF_OUT = MODEL_OFF_LAYER + N_LAYERS * LAYER_BYTES,
                   MODEL_BYTES = MODEL_OFF_OUT + (uint64_t)VOCAB * D_MODEL;

// Device layout (same total size): the embedding stays row-major; every other matrix is
// "packed": 32-bit words [in / 4][out], word (i4, o) holding w(o, 4 i4 + b) in byte b, so the
// lanes of a wavefront computing consecutive outputs read consecutive words. Per layer: Q, K and
// V as one packed [64][768] matrix (outputs q 0..255, k 256..511, v 512..767), then O [64][256],
// W1 [64][1024], W2 [256][256]; then the output projection [64][4096].
constexpr uint64_t DOFF_QKV = 0, DOFF_O = DOFF_QKV + 3ull * D_MODEL * D_MODEL,
                   DOFF_W1 = DOFF_O + (uint64_t)D_MODEL * D_MODEL,
                   DOFF_W2 = DOFF_W1 + (uint64_t)D_FF * D_MODEL;
static_assert(DOFF_W2 + (uint64_t)D_MODEL * D_FF == LAYER_BYTES, "device layer layout");

// KV cache of one request of `total` tokens, tcap = total rounded up to 4, int32 words:
// K transposed, [layer][64 words of 4 int8][tcap] (word w of position u at (l * 64 + w) * tcap + u);
// V packed by 4 positions, [layer][tcap / 4][256] (v(u, i) in byte u % 4 of word
// (l * tcap / 4 + u / 4) * 256 + i).
inline __host__ __device__ uint32_t kv_tcap(uint32_t total) { return (total + 3) & ~3u; }
inline uint64_t kv_words(uint32_t total) {
    return (uint64_t)N_LAYERS * (64ull * kv_tcap(total) + (uint64_t)kv_tcap(total) / 4 * D_MODEL);
}

// Witness values per token: per layer a, q, k, v, att, m (D_MODEL each), h (D_FF), x (D_MODEL);
// then f (D_MODEL) and the chosen next token.
constexpr uint64_t WIT_LAYER = 7 * D_MODEL + D_FF;
constexpr uint64_t WIT_STRIDE = N_LAYERS * WIT_LAYER + D_MODEL + 1;
constexpr uint64_t WIT_A0 = 0, WIT_M0 = 5 * D_MODEL;  // layer 0's a and m, for the sumcheck
constexpr uint64_t WIT_NEXT = N_LAYERS * WIT_LAYER + D_MODEL;

// One request's share of a decode launch: positions [start, end) of a request whose buffers
// were allocated for `total` tokens. The token at `start` is prompt[start] inside the prompt,
// else the next token the witness recorded at start - 1.
struct DecodeTask {
    const uint32_t *prompt;
    uint32_t *tokens;
    int32_t *kv;
    uint64_t *wit;
    uint32_t n_prompt, total, start, end;
};

Status model_weights(Ctx *ctx, int8_t *w, uint64_t seed, hipStream_t s);
// One launch: workgroup b advances tasks[b] (a device array of `count` tasks).
Status model_decode(Ctx *ctx, const int8_t *w, const DecodeTask *tasks, uint32_t count,
                    hipStream_t s);

}  // namespace tsr
