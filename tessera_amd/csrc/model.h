// The synthetic integer transformer of served proving (see model.hip): shapes, the weight
// layout (row-major [out][in] int8), and the witness layout.
#pragma once
#include <stdint.h>
#include "common.h"

namespace tsr {

constexpr int N_LAYERS = 4, D_MODEL = 256, N_HEADS = 4, D_HEAD = 64, D_FF = 1024, VOCAB = 4096;
constexpr int MAX_TOKENS = 512;

constexpr uint64_t LOFF_Q = 0, LOFF_K = LOFF_Q + D_MODEL * D_MODEL, LOFF_V = LOFF_K + D_MODEL * D_MODEL,
                   LOFF_O = LOFF_V + D_MODEL * D_MODEL, LOFF_W1 = LOFF_O + D_MODEL * D_MODEL,
                   LOFF_W2 = LOFF_W1 + (uint64_t)D_FF * D_MODEL,
                   LAYER_BYTES = LOFF_W2 + (uint64_t)D_MODEL * D_FF;
constexpr uint64_t MODEL_OFF_EMBED = 0, MODEL_OFF_LAYER = (uint64_t)VOCAB * D_MODEL,
                   MODEL_OFF_OUT = MODEL_OFF_LAYER + N_LAYERS * LAYER_BYTES,
                   MODEL_BYTES = MODEL_OFF_OUT + (uint64_t)VOCAB * D_MODEL;

// Witness values per token: per layer a, q, k, v, att, m (D_MODEL each), h (D_FF), x (D_MODEL);
// then f (D_MODEL) and the chosen next token.
constexpr uint64_t WIT_LAYER = 7 * D_MODEL + D_FF;
constexpr uint64_t WIT_STRIDE = N_LAYERS * WIT_LAYER + D_MODEL + 1;
constexpr uint64_t WIT_A0 = 0, WIT_M0 = 5 * D_MODEL;  // layer 0's a and m, for the sumcheck

Status model_weights(Ctx *ctx, int8_t *w, uint64_t seed, hipStream_t s);
Status model_decode(Ctx *ctx, const int8_t *w, const uint32_t *prompt, uint32_t n_prompt,
                    uint32_t total, uint32_t *tokens, int8_t *kv, uint64_t *wit, hipStream_t s);

}  // namespace tsr
