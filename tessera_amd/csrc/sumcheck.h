// Batched sumchecks (see sumcheck.hip): the product relation of prove_batch and the
// fraction-tree level relation, run by one device-resident round loop.
#pragma once
#include <string>
#include <vector>
#include "common.h"
#include "field.h"
#include "transcript.h"

namespace tsr {

enum { REL_PRODUCT = 0, REL_FRACTION = 1, REL_TERMS = 2 };

struct SumcheckOut {
    gl::fp2 gamma;
    std::vector<gl::fp2> rounds;      // n_max x degree
    std::vector<gl::fp2> challenges;  // n_max
    std::vector<gl::fp2> finals;
};

// One table of a batch instance, on the device. Entry x of the original table is
// src[(x << shift) + off]; src == nullptr is the constant 1 (no storage).
struct GenTable {
    const void *src = nullptr;
    bool base = false;
    bool is_public = false;
    uint8_t shift = 0, off = 0;
};

// A term of a REL_TERMS instance: weight times the product of its factors, with mult[k] the
// number of times table k appears among them.
struct GenTerm {
    gl::fp2 weight{1, 0};
    std::vector<uint8_t> mult;
};

struct GenInst {
    uint32_t vars = 0;
    int rel = REL_PRODUCT;  // REL_FRACTION: tables N0, N1, D0, D1 with the eq factor (degree 3)
    std::vector<GenTable> tabs;
    std::vector<GenTerm> terms;  // REL_TERMS: sum_t w_t prod_k T_k^mult[t][k]
    bool claim_given = false;  // padding rounds use `claim` (a tree level's) instead of S_i
    gl::fp2 claim{0, 0};
    // eq(eq_point, x) multiplies the relation without being a table (degree 3, product or
    // fraction relations with a claim): the rounds weight each pair by the eq of the variables
    // not yet bound and fold the bound ones into a running factor.
    bool eq_factor = false;
    std::vector<gl::fp2> eq_point;  // vars coordinates, bit i of the index is coordinate i
    // REL_FRACTION with stored numerators: the values of the level the instance proves, entry x
    // = (parent_n[x], parent_d[x]) = (N0 D1 + N1 D0, D0 D1) at x (nullptr: not available). The
    // first round's sums at z = 0, 1 then take them instead of the tables' products.
    const gl::fp2 *parent_n = nullptr, *parent_d = nullptr;
};

// The rounds and final evaluations of a batch whose gamma (and lambda) the caller squeezed.
Status run_batch(Ctx *ctx, Transcript &tr, const std::string &prefix, uint32_t degree,
                 std::vector<GenInst> &insts, gl::fp2 gamma, gl::fp2 lambda, SumcheckOut &out);

// prove_batch of product instances (host tables are uploaded into `uploads`).
Status prove_batch(Ctx *ctx, Transcript &tr, const char *prefix, uint32_t degree,
                   const tsr_instance *instances, uint32_t count, SumcheckOut &out,
                   std::vector<DevBuf<uint64_t>> &uploads);

// eq(r, .) over n variables from a device point of n Fp2 (tessera-poly eq.rs:23).
Status eq_table(Ctx *ctx, const gl::fp2 *point, uint32_t n, gl::fp2 *out, hipStream_t s);

}  // namespace tsr
