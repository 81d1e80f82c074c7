/*
 * Test infrastructure (oracle): the synthetic integer transformer of served proving, restated
 * on the CPU from its definition (no reference counterpart: the reference's models come from
 * crates/tessera-witness). 4 layers, d_model 256, 4 heads of 64, d_ff 1024, vocabulary 4096,
 * int8 weights (SplitMix64 of the index mod 255, minus 127), int32 arithmetic:
 *   a = rq(x); q, k, v = rq8(W a); s_u = relu((q . k_u) >> 6), u <= t; o = (sum s_u v_u) /
 *   (1 + sum s_u) (C division); x += Wo clamp(o); m = rq(x); h = clamp(relu((W1 m) >> 8), 0, 127);
 *   x += (W2 h) >> 8; after the layers f = rq(x), next = argmax(Wout f) (lowest index on ties);
 * rq(x) = clamp(x >> 2), rq8(y) = clamp(y >> 8), clamp to [-127, 127], arithmetic shifts.
 * The witness of a token: per layer a, q, k, v, att, m, h, x, then f and the next token, as
 * Goldilocks elements (x >= 0 -> x, x < 0 -> p - |x|).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

enum { L = 4, D = 256, H = 4, DH = 64, FF = 1024, V = 4096 };
#define LOFF_Q 0
#define LOFF_K (LOFF_Q + D * D)
#define LOFF_V (LOFF_K + D * D)
#define LOFF_O (LOFF_V + D * D)
#define LOFF_W1 (LOFF_O + D * D)
#define LOFF_W2 (LOFF_W1 + FF * D)
#define LAYER_BYTES (LOFF_W2 + D * FF)
#define OFF_LAYER ((uint64_t)V * D)
#define OFF_OUT (OFF_LAYER + (uint64_t)L * LAYER_BYTES)
#define MODEL_BYTES (OFF_OUT + (uint64_t)V * D)
#define WIT_LAYER (7 * D + FF)
#define WIT_STRIDE (L * WIT_LAYER + D + 1)

uint64_t orc_model_bytes(void) { return MODEL_BYTES; }
uint64_t orc_model_witness_stride(void) { return WIT_STRIDE; }

void orc_model_weights(uint64_t seed, int8_t *w) {
    for (uint64_t i = 0; i < MODEL_BYTES; i++) {
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        w[i] = (int8_t)((int)(z % 255) - 127);
    }
}

static int32_t clamp127(int32_t v) { return v < -127 ? -127 : (v > 127 ? 127 : v); }
static uint64_t emb(int64_t v) { return v >= 0 ? (uint64_t)v : ORC_P - (uint64_t)(-v); }
static int32_t dot(const int8_t *w, const int8_t *x, int n) {
    int32_t acc = 0;
    for (int i = 0; i < n; i++) acc += (int32_t)w[i] * (int32_t)x[i];
    return acc;
}

void orc_model_decode(const int8_t *W, const uint32_t *prompt, uint32_t n_prompt, uint32_t total,
                      uint32_t *tokens, uint64_t *wit) {
    int8_t *kv = calloc((size_t)L * total * 2 * D, 1);
    int32_t x[D];
    int8_t a8[D], q8[D], att8[D], h8[FF];
    int32_t *scores = malloc(sizeof(int32_t) * total);
    uint32_t cur = prompt[0];
    for (uint32_t t = 0; t < total; t++) {
        tokens[t] = cur;
        uint64_t *wt = wit + (uint64_t)t * WIT_STRIDE;
        for (int i = 0; i < D; i++) x[i] = W[(uint64_t)cur * D + i];
        for (int l = 0; l < L; l++) {
            const int8_t *Lw = W + OFF_LAYER + (uint64_t)l * LAYER_BYTES;
            uint64_t *wl = wt + (uint64_t)l * WIT_LAYER;
            int8_t *kvl = kv + (uint64_t)l * total * 2 * D;
            for (int i = 0; i < D; i++) {
                a8[i] = (int8_t)clamp127(x[i] >> 2);
                wl[i] = emb(a8[i]);
            }
            for (int i = 0; i < D; i++) {
                int32_t qv = clamp127(dot(Lw + LOFF_Q + (uint64_t)i * D, a8, D) >> 8);
                int32_t kv_ = clamp127(dot(Lw + LOFF_K + (uint64_t)i * D, a8, D) >> 8);
                int32_t vv = clamp127(dot(Lw + LOFF_V + (uint64_t)i * D, a8, D) >> 8);
                q8[i] = (int8_t)qv;
                kvl[(uint64_t)t * 2 * D + i] = (int8_t)kv_;
                kvl[(uint64_t)t * 2 * D + D + i] = (int8_t)vv;
                wl[D + i] = emb(qv);
                wl[2 * D + i] = emb(kv_);
                wl[3 * D + i] = emb(vv);
            }
            for (int h = 0; h < H; h++) {
                for (uint32_t u = 0; u <= t; u++) {
                    int32_t s = dot(q8 + h * DH, kvl + (uint64_t)u * 2 * D + h * DH, DH) >> 6;
                    scores[u] = s > 0 ? s : 0;
                }
                for (int c = 0; c < DH; c++) {
                    int i = h * DH + c;
                    int64_t num = 0, norm = 1;
                    for (uint32_t u = 0; u <= t; u++) {
                        num += (int64_t)scores[u] * kvl[(uint64_t)u * 2 * D + D + i];
                        norm += scores[u];
                    }
                    att8[i] = (int8_t)clamp127((int32_t)(num / norm));
                    wl[4 * D + i] = emb(att8[i]);
                }
            }
            int32_t xo[D];
            for (int i = 0; i < D; i++) xo[i] = x[i] + dot(Lw + LOFF_O + (uint64_t)i * D, att8, D);
            memcpy(x, xo, sizeof x);
            for (int i = 0; i < D; i++) {
                a8[i] = (int8_t)clamp127(x[i] >> 2);
                wl[5 * D + i] = emb(a8[i]);
            }
            for (int j = 0; j < FF; j++) {
                int32_t hv = dot(Lw + LOFF_W1 + (uint64_t)j * D, a8, D) >> 8;
                hv = hv < 0 ? 0 : (hv > 127 ? 127 : hv);
                h8[j] = (int8_t)hv;
                wl[6 * D + j] = emb(hv);
            }
            for (int i = 0; i < D; i++) {
                x[i] += dot(Lw + LOFF_W2 + (uint64_t)i * FF, h8, FF) >> 8;
                wl[6 * D + FF + i] = emb(x[i]);
            }
        }
        for (int i = 0; i < D; i++) {
            a8[i] = (int8_t)clamp127(x[i] >> 2);
            wt[L * WIT_LAYER + i] = emb(a8[i]);
        }
        int32_t best = 0;
        uint32_t best_i = 0;
        for (int j = 0; j < V; j++) {
            int32_t lg = dot(W + OFF_OUT + (uint64_t)j * D, a8, D);
            if (j == 0 || lg > best) {
                best = lg;
                best_i = (uint32_t)j;
            }
        }
        wt[L * WIT_LAYER + D] = best_i;
        cur = (t + 1 < n_prompt) ? prompt[t + 1] : best_i;
    }
    free(kv);
    free(scores);
}
