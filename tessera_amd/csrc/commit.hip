// Column chains, Merkle trees and the unmasked commitment on gfx950.
//
// Semantics (bit for bit): crates/tessera-hash/src/columns.rs:72-148 (d_j[0] = H(column-init,
// id || u64LE(j)); d_j <- H(column-step, d_j || u64LE(symbols of the block's rows))),
// crates/tessera-hash/src/merkle.rs:10-225 (H(merkle-node, l || r), levels leaves-first,
// canonical multiproofs) and crates/tessera-pcs/src/commitment.rs:343-598 (commit_unmasked,
// answer). One lane per column: a wave reads 64 consecutive symbols of a row (512 B, coalesced)
// per load, and every chain step is a single-chunk BLAKE3 kept in registers.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "blake3.h"
#include "commit.h"
#include "field.h"

namespace tsr {

struct Key8 {
    uint32_t w[8];
};

static Key8 key_of(const Ctx *ctx, int d) {
    Key8 k;
    memcpy(k.w, ctx->keys[d], 32);
    return k;
}

// d <- H(column-step, d || u64LE(col[0]) || u64LE(col[stride]) || ...), count <= 124 symbols.
__device__ __forceinline__ void chain_step(uint32_t d[8], const uint64_t *__restrict__ col,
                                           uint64_t stride, uint32_t count, const Key8 &key) {
    const uint32_t total = 32 + 8 * count;  // bytes
    const uint32_t nblocks = (total + 63) / 64;
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = key.w[i];
    for (uint32_t blk = 0; blk < nblocks; blk++) {
        uint32_t m[16];
        if (blk == 0) {
#pragma unroll
            for (int i = 0; i < 8; i++) m[i] = d[i];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint64_t v = (uint32_t)i < count ? col[(uint64_t)i * stride] : 0;
                m[8 + 2 * i] = (uint32_t)v;
                m[9 + 2 * i] = (uint32_t)(v >> 32);
            }
        } else {
            const uint32_t s0 = 8 * blk - 4;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                uint32_t s = s0 + i;
                uint64_t v = s < count ? col[(uint64_t)s * stride] : 0;
                m[2 * i] = (uint32_t)v;
                m[2 * i + 1] = (uint32_t)(v >> 32);
            }
        }
        uint32_t flags = b3::DERIVE_KEY_MATERIAL;
        if (blk == 0) flags |= b3::CHUNK_START;
        uint32_t len = 64;
        if (blk == nblocks - 1) {
            flags |= b3::CHUNK_END | b3::ROOT;
            len = total - 64 * blk;
        }
        b3::compress(cv, m, 0, len, flags, cv);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = cv[i];
}

// Every column's whole chain over `rows` rows (blocks of `block`, the last may be short), the
// last digest written as the column's leaf.
__global__ void __launch_bounds__(256)
k_chain_leaves(const uint64_t *__restrict__ cw, uint64_t n, uint64_t rows, uint32_t block,
               Key8 init_key, Key8 step_key, Key8 id, uint32_t *__restrict__ leaves) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t d[8];
    b3::hash_id_index(init_key.w, id.w, j, d);
    for (uint64_t r0 = 0; r0 < rows; r0 += block) {
        uint32_t count = (uint32_t)(rows - r0 < block ? rows - r0 : block);
        chain_step(d, cw + r0 * n + j, n, count, step_key);
    }
    uint4 *o = (uint4 *)(leaves + 8 * j);
    o[0] = make_uint4(d[0], d[1], d[2], d[3]);
    o[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

__global__ void __launch_bounds__(256)
k_chain_init(uint64_t n, Key8 init_key, Key8 id, uint32_t *__restrict__ chains) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t d[8];
    b3::hash_id_index(init_key.w, id.w, j, d);
#pragma unroll
    for (int i = 0; i < 8; i++) chains[8 * j + i] = d[i];
}

__global__ void __launch_bounds__(256)
k_chain_absorb(const uint64_t *__restrict__ rows, uint64_t n, uint32_t count, Key8 step_key,
               uint32_t *__restrict__ chains) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t d[8];
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = chains[8 * j + i];
    chain_step(d, rows + j, n, count, step_key);
#pragma unroll
    for (int i = 0; i < 8; i++) chains[8 * j + i] = d[i];
}

// One Merkle level: out[i] = H(merkle-node, in[2i] || in[2i+1]).
__global__ void __launch_bounds__(256)
k_merkle_level(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t count,
               Key8 key) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4 *l = (const uint4 *)(in + 16 * i);
    uint4 a = l[0], b = l[1], c = l[2], e = l[3];
    uint32_t L[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t R[8] = {c.x, c.y, c.z, c.w, e.x, e.y, e.z, e.w};
    uint32_t d[8];
    b3::hash64(key.w, L, R, d);
    uint4 *o = (uint4 *)(out + 8 * i);
    o[0] = make_uint4(d[0], d[1], d[2], d[3]);
    o[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

__global__ void k_gather_columns(const uint64_t *__restrict__ cw, uint64_t n, uint64_t rows,
                                 const uint64_t *__restrict__ idx, uint64_t q,
                                 uint64_t *__restrict__ out) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q * rows) return;
    uint64_t c = t / rows, r = t % rows;
    out[t] = cw[r * n + idx[c]];
}

__global__ void k_gather_digests(const uint32_t *__restrict__ nodes,
                                 const uint64_t *__restrict__ idx, uint64_t q,
                                 uint32_t *__restrict__ out) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q * 8) return;
    out[t] = nodes[8 * idx[t / 8] + t % 8];
}

// ---- many matrices per launch (batched commitments) ----

// Every column chain of every matrix in one launch: block b covers columns
// refs[b].first .. +255 of matrix refs[b].mat.
__global__ void __launch_bounds__(256)
k_chain_leaves_multi(const MatDesc *__restrict__ mats, const BlockRef *__restrict__ refs,
                     Key8 init_key, Key8 step_key) {
    const BlockRef ref = refs[blockIdx.x];
    const MatDesc &m = mats[ref.mat];
    const uint64_t j = ref.first + threadIdx.x;
    if (j >= m.n) return;
    Key8 id;
#pragma unroll
    for (int i = 0; i < 8; i++) id.w[i] = m.id[i];
    uint32_t d[8];
    b3::hash_id_index(init_key.w, id.w, j, d);
    for (uint64_t r0 = 0; r0 < m.rows; r0 += m.block) {
        uint32_t count = (uint32_t)(m.rows - r0 < m.block ? m.rows - r0 : m.block);
        chain_step(d, m.cw + r0 * m.n + j, m.n, count, step_key);
    }
    uint4 *o = (uint4 *)(m.nodes + 8 * j);
    o[0] = make_uint4(d[0], d[1], d[2], d[3]);
    o[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

// One Merkle level of every matrix that still has one: block b computes parents
// refs[b].first .. +255 of level `level` of matrix refs[b].mat.
__global__ void __launch_bounds__(256)
k_merkle_level_multi(const MatDesc *__restrict__ mats, const BlockRef *__restrict__ refs,
                     uint32_t level, Key8 key) {
    const BlockRef ref = refs[blockIdx.x];
    const MatDesc &m = mats[ref.mat];
    const uint64_t width_in = m.n >> (level - 1), count = m.n >> level;
    const uint64_t i = ref.first + threadIdx.x;
    if (i >= count) return;
    // level l starts at node sum_{t<l} n >> t = 2n - 2 (n >> (l-1)) ... = 2n - 2 width_in
    const uint64_t start_in = 2 * m.n - 2 * width_in, start_out = start_in + width_in;
    const uint4 *l = (const uint4 *)(m.nodes + 8 * (start_in + 2 * i));
    uint4 a = l[0], b = l[1], c = l[2], e = l[3];
    uint32_t L[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t R[8] = {c.x, c.y, c.z, c.w, e.x, e.y, e.z, e.w};
    uint32_t d[8];
    b3::hash64(key.w, L, R, d);
    uint4 *o = (uint4 *)(m.nodes + 8 * (start_out + i));
    o[0] = make_uint4(d[0], d[1], d[2], d[3]);
    o[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

// Every matrix's root (node 2n - 2) into roots[mat].
__global__ void k_gather_roots(const MatDesc *__restrict__ mats, uint32_t count,
                               uint32_t *__restrict__ roots) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 8) return;
    const MatDesc &m = mats[t / 8];
    roots[t] = m.nodes[8 * (2 * m.n - 2) + t % 8];
}

void launch_chain_leaves_multi(Ctx *ctx, const MatDesc *mats, const BlockRef *refs,
                               uint32_t nblocks, hipStream_t s) {
    if (!nblocks) return;
    hipLaunchKernelGGL(k_chain_leaves_multi, dim3(nblocks), dim3(256), 0, s, mats, refs,
                       key_of(ctx, DOM_COLUMN_INIT), key_of(ctx, DOM_COLUMN_STEP));
}

void launch_merkle_level_multi(Ctx *ctx, const MatDesc *mats, const BlockRef *refs,
                               uint32_t nblocks, uint32_t level, hipStream_t s) {
    if (!nblocks) return;
    hipLaunchKernelGGL(k_merkle_level_multi, dim3(nblocks), dim3(256), 0, s, mats, refs, level,
                       key_of(ctx, DOM_MERKLE_NODE));
}

void launch_gather_roots(const MatDesc *mats, uint32_t count, uint32_t *roots, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_roots, dim3((count * 8 + 255) / 256), dim3(256), 0, s, mats,
                       count, roots);
}

static unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

// The top of a Merkle tree in one workgroup: from the level of `width` (<= 2 MERKLE_TOP) nodes at
// node index `start`, every level above it up to the root, each appended after the one below
// as k_merkle_level writes them; the level being built stays in LDS.
constexpr uint32_t MERKLE_TOP = 1024;
__global__ void __launch_bounds__(MERKLE_TOP)
k_merkle_top(uint32_t *__restrict__ nodes, uint64_t start, uint64_t width, Key8 key) {
    __shared__ uint32_t d[MERKLE_TOP][8];
    const uint32_t t = threadIdx.x;
    bool from_global = true;
    while (width > 1) {
        const uint64_t count = width / 2;
        uint32_t out[8];
        if (t < count) {
            uint32_t L[8], R[8];
            if (from_global) {
                const uint4 *l = (const uint4 *)(nodes + 8 * (start + 2 * t));
                const uint4 a = l[0], b = l[1], c = l[2], e = l[3];
                const uint32_t l8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                const uint32_t r8[8] = {c.x, c.y, c.z, c.w, e.x, e.y, e.z, e.w};
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    L[i] = l8[i];
                    R[i] = r8[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    L[i] = d[2 * t][i];
                    R[i] = d[2 * t + 1][i];
                }
            }
            b3::hash64(key.w, L, R, out);
        }
        __syncthreads();  // the level below is read before it is overwritten
        if (t < count) {
#pragma unroll
            for (int i = 0; i < 8; i++) d[t][i] = out[i];
            uint4 *o = (uint4 *)(nodes + 8 * (start + width + t));
            o[0] = make_uint4(out[0], out[1], out[2], out[3]);
            o[1] = make_uint4(out[4], out[5], out[6], out[7]);
        }
        __syncthreads();
        start += width;
        width = count;
        from_global = false;
    }
}

// Nine Merkle levels in one launch: workgroup b hashes the 512 nodes [512 b, 512 b + 512) of the
// level of `width` nodes at node index `start` into their 256 parents, those into 128, ..., down
// to their one common ancestor, each level in LDS and written out where k_merkle_level would
// write it. width is a multiple of 512.
constexpr uint32_t MERKLE_SUB_LEVELS = 9, MERKLE_SUB_IN = 1u << MERKLE_SUB_LEVELS;
__global__ void __launch_bounds__(MERKLE_SUB_IN / 2)
k_merkle_subtree(uint32_t *__restrict__ nodes, uint64_t start, uint64_t width, Key8 key) {
    __shared__ uint32_t d[MERKLE_SUB_IN / 2][8];
    const uint32_t t = threadIdx.x;
    const uint64_t b = blockIdx.x;
    uint32_t out[8];
    {
        const uint4 *l = (const uint4 *)(nodes + 8 * (start + b * MERKLE_SUB_IN + 2 * t));
        const uint4 a = l[0], c = l[1], e = l[2], f = l[3];
        const uint32_t L[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
        const uint32_t R[8] = {e.x, e.y, e.z, e.w, f.x, f.y, f.z, f.w};
        b3::hash64(key.w, L, R, out);
    }
    uint32_t cnt = MERKLE_SUB_IN / 2;  // this workgroup's nodes on the level just hashed
    for (uint32_t lv = 0;; lv++) {
        start += width;
        width >>= 1;
        if (t < cnt) {
#pragma unroll
            for (int i = 0; i < 8; i++) d[t][i] = out[i];
            uint4 *o = (uint4 *)(nodes + 8 * (start + b * cnt + t));
            o[0] = make_uint4(out[0], out[1], out[2], out[3]);
            o[1] = make_uint4(out[4], out[5], out[6], out[7]);
        }
        if (lv + 1 == MERKLE_SUB_LEVELS) break;
        __syncthreads();
        cnt >>= 1;
        if (t < cnt) {
            uint32_t L[8], R[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                L[i] = d[2 * t][i];
                R[i] = d[2 * t + 1][i];
            }
            b3::hash64(key.w, L, R, out);
        }
        __syncthreads();  // the level below is read before it is overwritten
    }
}

Status merkle_build(Ctx *ctx, uint32_t *nodes, uint64_t n, hipStream_t stream, bool timed) {
    Key8 key = key_of(ctx, DOM_MERKLE_NODE);
    uint64_t start = 0, width = n;
    KernelTimer timer(timed ? ctx : nullptr, "k_merkle_level", stream);
    // Levels of more than 2^20 nodes one launch each (their hashes keep the GPU busy), smaller ones
    // nine levels per launch while a level has more than 2 MERKLE_TOP nodes (widths are powers of
    // two), then the top at once. (Nine levels per launch from 2^22 leaves up: 0.19 ms against
    // 0.15 ms for the single levels; the 2^16 leaves of a 256 x 2^14 commit: 28 us less.)
    while (width > 2 * MERKLE_TOP) {
        if (width > (1ull << 20)) {
            const uint64_t count = width / 2;
            hipLaunchKernelGGL(k_merkle_level, dim3(blocks_for(count, 256)), dim3(256), 0, stream,
                               nodes + 8 * start, nodes + 8 * (start + width), count, key);
            start += width;
            width /= 2;
            continue;
        }
        hipLaunchKernelGGL(k_merkle_subtree, dim3((unsigned)(width / MERKLE_SUB_IN)), dim3(MERKLE_SUB_IN / 2),
                           0, stream, nodes, start, width, key);
        for (uint32_t lv = 0; lv < MERKLE_SUB_LEVELS; lv++) {
            start += width;
            width /= 2;
        }
    }
    if (width > 1)
        hipLaunchKernelGGL(k_merkle_top, dim3(1), dim3(MERKLE_TOP), 0, stream, nodes, start, width, key);
    return check_launch("merkle levels");
}

Status chain_leaves(Ctx *ctx, const uint64_t *cw, uint64_t n, uint64_t rows, uint32_t block,
                    const uint8_t id[32], uint32_t *leaves, hipStream_t stream) {
    Key8 idk;
    memcpy(idk.w, id, 32);
    KernelTimer timer(ctx, "k_chain_leaves", stream);
    hipLaunchKernelGGL(k_chain_leaves, dim3(blocks_for(n, 256)), dim3(256), 0, stream, cw, n,
                       rows, block, key_of(ctx, DOM_COLUMN_INIT), key_of(ctx, DOM_COLUMN_STEP),
                       idk, leaves);
    return check_launch("column chains");
}

// leaf_j = H(column-merge, d_j^(1) || ... || d_j^(S)) (commitment.rs:434-446): BLAKE3 in
// derive-key mode over the S digests of column j (segment s's at dig[(s n + j) 8 ..]), 1024-byte
// chunks of 32 digests, chunk chaining values merged into the BLAKE3 tree when S > 32.
__global__ void k_merge_leaves(const uint32_t *__restrict__ dig, uint32_t S, uint64_t n, Key8 key,
                               uint32_t *__restrict__ leaves) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t chunks = (S + 31) / 32;
    uint32_t stack[32][8];
    int depth = 0;
    uint32_t out[8];
    for (uint32_t c = 0; c < chunks; c++) {
        uint32_t cv[8];
        for (int i = 0; i < 8; i++) cv[i] = key.w[i];
        const uint32_t d0 = 32 * c, dn = S - d0 < 32 ? S - d0 : 32, nb = (dn + 1) / 2;
        for (uint32_t b = 0; b < nb; b++) {
            uint32_t m[16];
            const uint32_t *a = dig + ((uint64_t)(d0 + 2 * b) * n + j) * 8;
            const bool pair = 2 * b + 1 < dn;
            const uint32_t *bb = dig + ((uint64_t)(d0 + 2 * b + 1) * n + j) * 8;
            for (int i = 0; i < 8; i++) {
                m[i] = a[i];
                m[8 + i] = pair ? bb[i] : 0;
            }
            uint32_t flags = b3::DERIVE_KEY_MATERIAL;
            if (b == 0) flags |= b3::CHUNK_START;
            if (b == nb - 1) flags |= b3::CHUNK_END;
            if (chunks == 1 && b == nb - 1) flags |= b3::ROOT;
            b3::compress(cv, m, c, pair ? 64 : 32, flags, cv);
        }
        if (c + 1 < chunks) {
            // a chunk followed by more input: push its chaining value, merging complete subtrees
            uint32_t total = c + 1;
            while ((total & 1) == 0) {
                uint32_t m[16];
                depth--;
                for (int i = 0; i < 8; i++) {
                    m[i] = stack[depth][i];
                    m[8 + i] = cv[i];
                }
                b3::compress(key.w, m, 0, 64, b3::DERIVE_KEY_MATERIAL | b3::PARENT, cv);
                total >>= 1;
            }
            for (int i = 0; i < 8; i++) stack[depth][i] = cv[i];
            depth++;
        } else {
            for (int i = 0; i < 8; i++) out[i] = cv[i];
        }
    }
    while (depth > 0) {
        uint32_t m[16];
        depth--;
        for (int i = 0; i < 8; i++) {
            m[i] = stack[depth][i];
            m[8 + i] = out[i];
        }
        b3::compress(key.w, m, 0, 64, b3::DERIVE_KEY_MATERIAL | b3::PARENT | (depth == 0 ? b3::ROOT : 0u), out);
    }
    for (int i = 0; i < 8; i++) leaves[j * 8 + i] = out[i];
}

Status merge_leaves(Ctx *ctx, const uint32_t *digests, uint32_t segments, uint64_t n,
                    uint32_t *leaves, hipStream_t stream) {
    KernelTimer timer(ctx, "k_merge_leaves", stream);
    hipLaunchKernelGGL(k_merge_leaves, dim3(blocks_for(n, 256)), dim3(256), 0, stream, digests, segments, n,
                       key_of(ctx, DOM_COLUMN_MERGE), leaves);
    return check_launch("column merge");
}

Status chains_init(Ctx *ctx, const uint8_t id[32], uint64_t n, uint32_t *chains,
                   hipStream_t stream) {
    Key8 idk;
    memcpy(idk.w, id, 32);
    hipLaunchKernelGGL(k_chain_init, dim3(blocks_for(n, 256)), dim3(256), 0, stream, n,
                       key_of(ctx, DOM_COLUMN_INIT), idk, chains);
    return check_launch("column init");
}

Status chains_absorb(Ctx *ctx, const uint64_t *rows, uint64_t n, uint32_t count,
                     uint32_t *chains, hipStream_t stream) {
    hipLaunchKernelGGL(k_chain_absorb, dim3(blocks_for(n, 256)), dim3(256), 0, stream, rows, n,
                       count, key_of(ctx, DOM_COLUMN_STEP), chains);
    return check_launch("column absorb");
}

Status gather_columns(const uint64_t *cw, uint64_t n, uint64_t rows, const uint64_t *idx,
                      uint64_t q, uint64_t *out, hipStream_t stream) {
    if (q * rows == 0) return Status::ok();
    hipLaunchKernelGGL(k_gather_columns, dim3(blocks_for(q * rows, 256)), dim3(256), 0, stream,
                       cw, n, rows, idx, q, out);
    return check_launch("gather columns");
}

Status gather_digests(const uint32_t *nodes, const uint64_t *idx, uint64_t q, uint32_t *out,
                      hipStream_t stream) {
    if (q == 0) return Status::ok();
    hipLaunchKernelGGL(k_gather_digests, dim3(blocks_for(q * 8, 256)), dim3(256), 0, stream,
                       nodes, idx, q, out);
    return check_launch("gather digests");
}

// Canonical multiproof sibling positions (merkle.rs:126-169), as indices into the node array.
std::vector<uint64_t> multiproof_positions(uint64_t n, std::vector<uint64_t> known) {
    std::vector<uint64_t> out;
    std::sort(known.begin(), known.end());
    known.erase(std::unique(known.begin(), known.end()), known.end());
    if (known.size() == n) return out;
    uint64_t start = 0, width = n;
    while (width > 1) {
        std::vector<uint64_t> parents;
        size_t i = 0;
        while (i < known.size()) {
            uint64_t node = known[i];
            if ((node & 1) == 0 && i + 1 < known.size() && known[i + 1] == node + 1) {
                i += 2;
            } else {
                out.push_back(start + (node ^ 1));
                i += 1;
            }
            parents.push_back(node / 2);
        }
        known.swap(parents);
        start += width;
        width /= 2;
    }
    return out;
}

}  // namespace tsr
