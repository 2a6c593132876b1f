// sx_ctx.h -- the context behind the C ABI (one staged trajectory block on
// one GPU) and host helpers shared by the single-trajectory and the batched
// entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <map>
#include <string>

#include "../../include/sassena_amd.h"
#include "sx_internal.h"

namespace sxh {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr size_t kStageChunkFloats = size_t(16) << 20;   // 64 MB staging pieces
constexpr size_t kAmpBudgetBytes = size_t(24) << 30;      // amplitude buffer per batch

// A bump region of device memory the caller owns (the allocation-free entry
// point's scratch): DevBufs bound to it take slices instead of allocating.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool exhausted = false;
};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;  // non-null: slices of this arena, never hipMalloc/hipFree
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (arena) {
            const size_t off = (arena->used + 255) / 256 * 256;
            if (off + bytes > arena->cap) {
                arena->exhausted = true;
                return hipErrorOutOfMemory;
            }
            p = arena->base + off;
            cap = bytes;
            arena->used = off + bytes;
            return hipSuccess;
        }
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p && !arena) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

float elapsed(hipEvent_t a, hipEvent_t b);

}  // namespace sxh

struct sx_ctx {
    sx_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    size_t NA = 0, NF = 0, NAp = 0, NFp = 0;
    bool staged = false, have_b = false;
    sx::FftPlan plan;
    bool plan_ok = false;
    sxh::DevBuf coords, b, twL, twL64, twp, twp64, tw64, w, aidx, apair, tail;
    std::vector<int> aidx_h;  // "self": atoms with a nonzero length in the current call
    std::vector<float> b_h;   // host copy of the float factors of sx_set_factors
    sxh::DevBuf qrev, qlen, uvec, brow, bshell, amp, ppart, apart, fqt, fq, fq2, stage, stage2, mpmom, mpq, mpsph, mpj,
        mpjt, qres, redbuf;
    sx_timing timing{};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // stepping mode of the oriented shells of the current call (sx::shell_progression:
    // 0 none, 1 exact progression, 2 progression with first-order residual corrections);
    // origin and step in |q| / 2 pi
    int q_mode = 0;
    double q_step = 0.0, q_origin = 0.0;
    sx::ShellGrid q_grid;              // the call's shells (mode 2): grid and exact runs
    std::vector<double> q_turns;       // and their |q| / 2 pi
    // host copies of the tile lines and residuals last queued to the device
    // (alive until the stream has consumed them)
    std::vector<sx::TileLine> tl_lines;
    std::vector<float> tl_res;
    // largest |r| of the staged coordinates (Angstrom): bounds the residual phases
    double rmax = 0.0;
    // autotuned oriented tile shape per (NFp, NAp, NM, shells, per-shell b, progression)
    std::map<std::array<size_t, 6>, int> tuned;
    // set during an allocation-free call: work buffers are arena slices, no
    // autotuning (default or fixed tile shapes), atom ranges only as many as
    // needed to fill the GPU (no round-filling extra ranges)
    bool noalloc = false;
    void *scratch = nullptr;  // device address of this device's scratch array
};
