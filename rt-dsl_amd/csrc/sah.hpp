// sah.hpp -- the split decisions of the SAH tree build (sah.hip) as
// __host__ __device__ code, shared with its host restatement (host_walk.cpp) so
// both make bit-identical choices.
//
// A segment of sorted points [lo, hi) with its exact point box is split by
// binned SAH: along each axis the box is cut into SAH_BINS equal bins, every
// point counted in the bin of its coordinate, each bin keeping the exact box of
// its points (min / max as order-preserving integer encodings, so the device's
// atomics and the host's loop give the same bits).  A split "bins 0..j left of
// axis a" costs  n_left * area(box_left) + n_right * area(box_right)  in FP32
// (area = dx*dy + dy*dz + dz*dx of the box's edges); the cheapest split with
// both sides non-empty is taken, the lowest axis and then the lowest j on ties.
// Bins are monotone in the coordinate, so the two sides are separated by a
// plane and their boxes do not overlap along that axis.
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

constexpr int SAH_BINS = 8;
constexpr int SAH_WORDS = 3 * SAH_BINS * 7;  // counts, then encoded lo, then encoded hi

// order-preserving encoding of a float as uint32 (-0 below +0)
GPRT_HD uint32_t sah_enc(float f) {
    const uint32_t u = (uint32_t)as_int(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
GPRT_HD float sah_dec(uint32_t u) { return as_float((int)((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u)); }

// bins per unit length of an axis of extent [lo, hi]; 0 for a flat axis (every point in bin 0)
GPRT_HD float sah_scale(float lo, float hi) {
    const float e = hi - lo;
    return e > 0.0f ? (float)SAH_BINS / e : 0.0f;
}
GPRT_HD int sah_bin(float x, float lo, float scale) {
    const int b = (int)((x - lo) * scale);
    return b < 0 ? 0 : (b >= SAH_BINS ? SAH_BINS - 1 : b);
}

// Bin data of one segment, as 32-bit words: cnt[a][b] at a*B + b, encoded lo
// coordinate c of bin (a, b) at 3B + (a*B + b)*3 + c, encoded hi at 3B + 9B + ...
GPRT_HD int sah_cnt_word(int a, int b) { return a * SAH_BINS + b; }
GPRT_HD int sah_lo_word(int a, int b, int c) { return 3 * SAH_BINS + (a * SAH_BINS + b) * 3 + c; }
GPRT_HD int sah_hi_word(int a, int b, int c) { return 12 * SAH_BINS + (a * SAH_BINS + b) * 3 + c; }
// the empty state of a segment's bin words
GPRT_HD uint32_t sah_init_word(int w) { return w < 3 * SAH_BINS ? 0u : (w < 12 * SAH_BINS ? 0xFFFFFFFFu : 0u); }

GPRT_HD float sah_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float xy = dx * dy, yz = dy * dz, zx = dz * dx;
    return (xy + yz) + zx;
}

struct SahSplit {
    int axis, bin;     // points with sah_bin(coord[axis]) <= bin go left
    uint32_t nleft;    // points on the left
    float llo[3], lhi[3], rlo[3], rhi[3];  // exact boxes of the two sides
};

// The cheapest split of the binned segment; false when no split has both sides
// non-empty (every axis flat: all points equal).
GPRT_HD bool sah_choose(const uint32_t* w, SahSplit& out) {
    float best = INFINITY;
    int ba = -1, bb = -1;
    for (int a = 0; a < 3; ++a) {
        // prefix / suffix counts and encoded boxes, areas decoded
        uint32_t pc[SAH_BINS], sc[SAH_BINS];
        float pa[SAH_BINS], sa[SAH_BINS];
        uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
        uint32_t c = 0;
        for (int b = 0; b < SAH_BINS; ++b) {
            const uint32_t n = w[sah_cnt_word(a, b)];
            if (n) {
                c += n;
                for (int k = 0; k < 3; ++k) {
                    const uint32_t l = w[sah_lo_word(a, b, k)], h = w[sah_hi_word(a, b, k)];
                    lo[k] = l < lo[k] ? l : lo[k];
                    hi[k] = h > hi[k] ? h : hi[k];
                }
            }
            pc[b] = c;
            float fl[3], fh[3];
            for (int k = 0; k < 3; ++k) { fl[k] = sah_dec(lo[k]); fh[k] = sah_dec(hi[k]); }
            pa[b] = c ? sah_area(fl, fh) : 0.0f;
        }
        for (int k = 0; k < 3; ++k) { lo[k] = 0xFFFFFFFFu; hi[k] = 0u; }
        c = 0;
        for (int b = SAH_BINS - 1; b >= 0; --b) {
            const uint32_t n = w[sah_cnt_word(a, b)];
            if (n) {
                c += n;
                for (int k = 0; k < 3; ++k) {
                    const uint32_t l = w[sah_lo_word(a, b, k)], h = w[sah_hi_word(a, b, k)];
                    lo[k] = l < lo[k] ? l : lo[k];
                    hi[k] = h > hi[k] ? h : hi[k];
                }
            }
            sc[b] = c;
            float fl[3], fh[3];
            for (int k = 0; k < 3; ++k) { fl[k] = sah_dec(lo[k]); fh[k] = sah_dec(hi[k]); }
            sa[b] = c ? sah_area(fl, fh) : 0.0f;
        }
        for (int b = 0; b + 1 < SAH_BINS; ++b) {
            if (pc[b] == 0 || sc[b + 1] == 0) continue;
            const float cost = (float)pc[b] * pa[b] + (float)sc[b + 1] * sa[b + 1];
            if (cost < best) {
                best = cost;
                ba = a;
                bb = b;
            }
        }
    }
    if (ba < 0) return false;
    out.axis = ba;
    out.bin = bb;
    uint32_t llo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, lhi[3] = {0u, 0u, 0u};
    uint32_t rlo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, rhi[3] = {0u, 0u, 0u};
    uint32_t nl = 0;
    for (int b = 0; b < SAH_BINS; ++b) {
        const uint32_t n = w[sah_cnt_word(ba, b)];
        if (!n) continue;
        uint32_t* L = b <= bb ? llo : rlo;
        uint32_t* H = b <= bb ? lhi : rhi;
        if (b <= bb) nl += n;
        for (int k = 0; k < 3; ++k) {
            const uint32_t l = w[sah_lo_word(ba, b, k)], h = w[sah_hi_word(ba, b, k)];
            L[k] = l < L[k] ? l : L[k];
            H[k] = h > H[k] ? h : H[k];
        }
    }
    out.nleft = nl;
    for (int k = 0; k < 3; ++k) {
        out.llo[k] = sah_dec(llo[k]);
        out.lhi[k] = sah_dec(lhi[k]);
        out.rlo[k] = sah_dec(rlo[k]);
        out.rhi[k] = sah_dec(rhi[k]);
    }
    return true;
}

// Node ids follow the Karras numbering, which holds for any binary tree over
// contiguous ranges of the sorted points: a cloud's root is its first point,
// an internal child A is its last point and an internal child B its first
// point.  The children of a node split at mid are then mid - 1 and mid, every
// internal node gets a distinct id in [first, last) of its cloud, and the
// walks find a child's first point as they do in the LBVH (A: id - count + 1,
// B: id).  slot: (parent << 1 | side), or negative for a cloud's root.
GPRT_HD int sah_seg_id(int lo, int hi, int slot) { return slot < 0 ? lo : ((slot & 1) ? lo : hi - 1); }

// below this depth segments are split by SAH, deeper ones at their index
// median, so no tree is deeper than SAH_MAX_DEPTH + 31 levels above its
// buckets (the walks' stacks hold 96 entries)
constexpr int SAH_MAX_DEPTH = 48;

}  // namespace gprt
