// The batched product sumcheck (see sumcheck.hip).
#pragma once
#include <vector>
#include "common.h"
#include "field.h"
#include "transcript.h"

namespace tsr {

struct SumcheckOut {
    gl::fp2 gamma;
    std::vector<gl::fp2> rounds;      // n_max x degree
    std::vector<gl::fp2> challenges;  // n_max
    std::vector<gl::fp2> finals;
};

Status prove_batch(Ctx *ctx, Transcript &tr, const char *prefix, uint32_t degree,
                   const tsr_instance *instances, uint32_t count, SumcheckOut &out);

}  // namespace tsr
