// fp_tiles_host.hpp — the host-side builder of the coset-tile tables (plain C++,
// no HIP): tile geometry, and the packed image of string groups, chunks and
// per-string records that the tiled kernels read.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_tile_types.hpp"

namespace fprt
{

// Records a printf-style message as the calling thread's last error and returns `code`.
fp_status fail(fp_status code, char const *fmt, ...) __attribute__((format(printf, 2, 3)));

// Coset-tile geometry of the tiled kernels: 2^lsb states per block,
// tiles of 2^d rows whose low k bits are consecutive rows and whose dh = d - k
// high dimensions come from the string group's span.
struct TileGeom
{
    uint32_t lsb, k, d, dh, tile_kb;
};

enum
{
    kTilesOpExpect = 0, // fpt::TileChunk / fpt::TileStr records (PauliOp expectation)
    kTilesEvMatrix = 1, // fpl::EvChunk / fpl::EvStr records (per-string EV matrix on cosets)
    kTilesOpApply = 2,  // fpa::ApplyStr records, groups of consecutive strings (PauliOp apply)
    kTilesSummedApply = 3, // the same for SummedPauliOp::apply: StringTable::cf128 / cf64 hold the
                           // folded row sums of the coefficient matrix
    kTilesWeightedApply = 4 // the same for SummedPauliOp::apply_weighted: StringTable::cf128 / cf64
                            // hold (-i)^{nY} (the per-state weights are applied in the kernel)
};

// The apply kinds: string groups of consecutive strings, one ApplyStr per string.
inline bool apply_kind(int kind)
{
    return kind == kTilesOpApply || kind == kTilesSummedApply || kind == kTilesWeightedApply;
}

// `expect`: tiles for the expectation kernel (complex128 takes 128-byte row slices), else apply
TileGeom tile_geom(int dtype, size_t n_qubits, size_t n_states, bool expect);

// The strings a table is cut from: host arrays of an operator set, string s at index s.
struct StringTable
{
    size_t n_qubits;
    size_t n_strings;
    uint64_t const *x;
    uint64_t const *z;
    uint32_t const *ny;     // number of Y letters
    double const *coeffs;   // c_s, complex128 (re, im)
    double const *cf128;    // c_s (-i)^{nY_s}, complex128
    float const *cf64;      // the same rounded to complex64
};

// Packed tables of one kind and geometry: [groups | btab | chunks | strs | sh],
// each part at a 256-byte aligned offset of `h`.
struct TileImage
{
    std::vector<unsigned char> h;
    size_t o_btab = 0, o_chunks = 0, o_strs = 0, o_sh = 0;
    uint32_t n_groups = 0;   // fpt::TileGroup records at offset 0
    uint32_t max_chunks = 0; // the most chunks (apply: strings) of any group
};

// Groups the strings into subspaces of dimension k + dh (the low k row bits plus
// dh reduced basis vectors) and writes the kind's records.  Apply groups are runs
// of consecutive strings in the operator's order; the others follow x order.
// kTilesEvMatrix: classes of at least `wht_min` strings with equal local x and
// equal half of w become transform classes (listed first in each group).
fp_status build_tile_image(StringTable const &st, int kind, uint32_t k, uint32_t dh, uint32_t wht_min,
                           TileImage &out);

} // namespace fprt
