// fp_tile_types.hpp — the host-built records that the tiled kernels read: string
// groups over coset tiles, their chunks and per-string records, and the
// per-string EV matrix's chunks.  Plain data with no HIP dependency, so that the
// host-side table builder also compiles as ordinary C++ (and under the address
// and undefined-behaviour sanitizers in its unit test).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define FP_HD __host__ __device__
#else
#define FP_HD
#endif

namespace fpt
{

constexpr int kMaxDh = 13; // string-span dimensions per tile (B table of 2^13 entries, read once per block)

struct TileGroup
{
    uint32_t n_chunks;
    uint32_t chunk_first;
    uint32_t n_piv;    // dh: pivots (high dimensions of V)
    uint32_t btab_off; // first entry of this group's B table (2^dh entries)
    uint64_t piv_lo;   // pivot bit positions 0..7, ascending, 8 bits each
    uint64_t piv_hi;   // pivot bit positions 8..15
    uint32_t hoist;     // bit hoist_bit(W, U1, r): every wave of a block has at most one
                        // unit and it takes the one-string fast path (expectation tables)
    uint32_t shared;    // bit shared_bit(W, c64, rr): the group runs the shared-leader
                        // form (expectation tables)
    uint32_t str_first; // first ShStr of the group (expectation tables)
    uint32_t n_str;     // strings of the group
    uint32_t n_wht;     // per-string EV matrix tables: the first n_wht chunks are transform classes
    uint32_t pad;
};

// Strings of a shared-leader group: every local x mask has bit k set, so the
// leader rows (local bit k clear) are common to all of them.  At most
// max_shared(W) strings for the instance with W waves per block (its per-string
// registers), kMaxShared over all instances.
constexpr int kMaxShared = 10;
FP_HD constexpr int max_shared(uint32_t W)
{
    return W == 16 ? 10 : 9;
}

struct ShStr
{
    uint64_t zc;   // z compressed onto the coset-index bits: coset sign = parity(q & zc)
    uint32_t A;    // local x mask (bit k set)
    uint32_t Z;    // local z mask
    uint32_t imag; // 1: Im w (odd nY), 0: Re w
    uint32_t pad;
    double sre; // scale c_s * kappa_s
    double sim;
};

// Bit of TileGroup::shared for W waves per block, complex64 (two states per
// lane) or not, and 2^rr leader rows per pass.
FP_HD constexpr uint32_t shared_bit(uint32_t W, bool c64, uint32_t rr)
{
    return (W == 16 ? 16u : 0u) + (c64 ? 8u : 0u) + rr;
}

// Bit of TileGroup::hoist for W waves per block, U1 passes per fast-path step and
// 2^r leader rows per pass (r = 6 - log2 S_blk).
FP_HD constexpr uint32_t hoist_bit(uint32_t W, uint32_t U1, uint32_t r)
{
    return (W == 16 ? 16u : 0u) + (U1 == 8 ? 8u : 0u) + r;
}

struct TileChunk
{
    uint32_t A;     // local x mask; 0: every row, kappa 1
    uint32_t h;     // top bit of A
    uint32_t count; // strings in this chunk (1..kChunk), all with local mask A
    uint32_t first; // first TileStr
    uint32_t need;  // bit 0: some string uses Re w, bit 1: some string uses Im w
    uint32_t pad[3];
};

struct TileStr
{
    uint64_t zc;   // z compressed onto the coset-index bits: coset sign = parity(q & zc)
    uint32_t zt;   // local z with bit h removed: row sign of leader p = parity(p & zt)
    uint32_t imag; // 1: Im w (odd nY), 0: Re w
    double sre;    // scale c_s * kappa_s
    double sim;
};

constexpr int kChunk = 4;

} // namespace fpt

namespace fpa
{

struct ApplyStr
{
    uint64_t zc;  // z compressed onto the coset-index bits: coset sign = parity(q & zc)
    uint32_t A;   // local x mask
    uint32_t Z;   // local z mask
    double cr;    // folded coefficient c_s (-i)^{nY_s}, complex128
    double ci;
    float fr;     // the same, complex64 (rounded as the streaming kernel's table)
    float fi;
    uint32_t s;   // the string's index in the operator
    uint32_t pad;
};

} // namespace fpa

namespace fpl
{

constexpr int kChunk = 8;

// A class of strings with equal x and equal half of w is evaluated by one
// Walsh-Hadamard transform of G over the leader rows when it holds at least
// kWhtMin strings (split into classes of at most kWhtMax):
//     ev_s = scale_s * sum_leaders (-1)^{parity(a & Z_s)} G(a) = scale_s * W[compress(Z_s)],
// one pass over the rows and (d-1) butterfly levels instead of one pass per string.
// One wave transforms a class on its own, in parts of kWhtPartBytes of G held in
// a private LDS slice (no block barrier); the parts are combined per string by
// their top index bits.
constexpr uint32_t kWhtMin = 12;
constexpr uint32_t kWhtMax = 256;        // 4 strings per lane
constexpr uint32_t kWhtMinRowsLog2 = 10; // columns shorter than this keep the per-string loop
constexpr uint32_t kWhtPartBytes = 2048;
constexpr uint32_t kWhtWaves = 15;       // waves with a slice (the last wave takes per-string chunks only)
constexpr uint32_t kWhtLdsBytes = kWhtWaves * kWhtPartBytes;

enum : uint32_t
{
    kAllRows = 0,
    kPairs = 1,
    kPairsLane = 2
};

struct EvChunk
{
    uint32_t x;     // x mask (n <= 14: fits 32 bits)
    uint32_t h;     // highest set bit of x (pair modes)
    uint32_t mode;  // kAllRows / kPairs / kPairsLane
    uint32_t count; // strings in this chunk (1..kChunk)
    uint32_t first; // index of the first EvStr
    uint32_t imag;  // 1: every string of the chunk takes Im w (odd nY), 0: Re w
    uint32_t pad[2];
};

// Mode of a string with x mask x (highest set bit h) over `rows` rows.
FP_HD constexpr uint32_t ev_mode(uint64_t x, uint32_t h, uint64_t rows)
{
    return x == 0 ? kAllRows : (h >= 6 ? kPairs : (rows >= 128 ? kPairsLane : kAllRows));
}

struct EvStr
{
    uint32_t z;
    uint32_t imag; // 1: G = Im w, 0: G = Re w
    uint32_t out;  // output row (original string index)
    uint32_t zc;   // coset kernel: z compressed onto the coset-index bits (0 otherwise)
    double scale;  // eps (kAllRows) or 2 eps (kPairs)
};

} // namespace fpl
