// Storage of the boundary records that OWL's gradient saves every forward
// step and reads back in the reverse pass: FP32 as OWL keeps them, or 16-bit
// values (FP16 or BF16) of v / s with one scale s per (shot, step) record.
// s = amax 2^-15, amax the record's largest magnitude, so the stored values
// lie in [-2^15, 2^15] (inside FP16's range, where FP16 keeps 11 bits) and
// multiplying by s restores the magnitude exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "owl_amd.h"

namespace owl {

struct RecStore {
    void* base;      // float (FP32) or 16-bit elements, the strip layout's indices
    unsigned* amax;  // per (shot, step) record: bits of the largest |v| (non-negative floats order as their bits)
    int kind;        // OWL_STORE_F32, OWL_STORE_F16 or OWL_STORE_BF16
    int nt;          // records per shot
};

// Modes of a strip kernel: the record restores the fields, the fields are
// saved into the record, or (16-bit storage, before saving) the record's
// largest magnitude is taken.
enum RecMode : int { kRecRestore = 0, kRecSave = 1, kRecAmax = 2 };

__device__ inline uint16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ inline float bf16_value(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Scale of record `rec` (1 for FP32 storage or an all-zero record).
__device__ inline float rec_scale(const RecStore& r, int rec) {
    if (r.kind == OWL_STORE_F32) return 1.0f;
    const float a = __uint_as_float(r.amax[rec]);
    return a > 0.0f ? a * 0x1p-15f : 1.0f;
}

__device__ inline void rec_put(const RecStore& r, long long e, float v, float inv) {
    if (r.kind == OWL_STORE_F32)
        static_cast<float*>(r.base)[e] = v;
    else if (r.kind == OWL_STORE_F16)
        static_cast<_Float16*>(r.base)[e] = static_cast<_Float16>(v * inv);
    else
        static_cast<uint16_t*>(r.base)[e] = bf16_rne(v * inv);
}

__device__ inline float rec_get(const RecStore& r, long long e, float s) {
    if (r.kind == OWL_STORE_F32) return static_cast<const float*>(r.base)[e];
    if (r.kind == OWL_STORE_F16) return float(static_cast<const _Float16*>(r.base)[e]) * s;
    return bf16_value(static_cast<const uint16_t*>(r.base)[e]) * s;
}

// Largest magnitude over the block's threads into the record's amax: one
// atomic per wavefront.  Every lane of the wavefront must call it.
__device__ inline void rec_amax(const RecStore& r, int rec, float m) {
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(r.amax + rec, __float_as_uint(m));
}

// Buffers for n record values of a batch: 16-bit storage packs two values
// per float element; the (nshot x nt) largest magnitudes start at zero.
inline void alloc_records(DevBuf<float>& strips, DevBuf<unsigned>& amax, size_t n, int kind, size_t nshot, int nt,
                          hipStream_t st) {
    strips.alloc(std::max<size_t>(kind == OWL_STORE_F32 ? n : (n + 1) / 2, 1));
    if (kind != OWL_STORE_F32) {
        amax.alloc(std::max<size_t>(nshot * size_t(nt), 1));
        amax.zero(st);
    }
}

// After a forward that kept 16-bit records: a non-finite largest magnitude
// (its bits at or above those of +inf) means the run diverged, and the
// scaled records would hold nothing usable.
inline void check_records_finite(const DevBuf<unsigned>& amax, int kind, size_t n) {
    if (kind == OWL_STORE_F32 || n == 0) return;
    std::vector<unsigned> h(n);
    OWL_HIP(hipMemcpy(h.data(), amax.get(), n * sizeof(unsigned), hipMemcpyDeviceToHost));
    for (unsigned v : h)
        if (v >= 0x7f800000u)
            fail(OWL_ERR_NONFINITE, "the wavefield became NaN or infinite (the run diverged): 16-bit boundary "
                                    "records cannot hold it");
}

inline void check_record_kind(int kind) {
    if (kind != OWL_STORE_F32 && kind != OWL_STORE_F16 && kind != OWL_STORE_BF16)
        fail(OWL_ERR_ARG, "record storage must be OWL_STORE_F32, OWL_STORE_F16 or OWL_STORE_BF16");
}

}  // namespace owl
