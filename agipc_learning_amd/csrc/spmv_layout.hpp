// Limits and entry column words of the FP32 SpMV layout, shared by the host
// builder (plan.cpp) and the SpMV kernel.
//
// Entry column word:
//   bit 31 clear: a column outside the tile, in bits 0..29;
//   bit 31 set:   an entry inside the tile, tpos << 16 | r_local << 8 | c_local,
//                 where tpos (bits 16..29) is the slot of its A^T x_r product
//                 among the tile's transposed products (unused for the
//                 diagonal block);
//   bit 30:       the entry is the first of its row.
// The transposed products of row c occupy slots [tptr[c] - tptr[r0],
// tptr[c+1] - tptr[r0]), in increasing r.  Every tile's first entry is a
// multiple of kSpmvAlign, so a thread's run of consecutive entries loads as
// whole vectors; the entry arrays and value planes end with kSpmvSlack spare
// entries so that the last run of the last tile stays inside them.
#pragma once

namespace masx
{

constexpr int      kSpmvTileRows = 256;          // one thread per row
constexpr unsigned kSpmvInTile   = 0x80000000u;
constexpr unsigned kSpmvRowStart = 0x40000000u;
constexpr unsigned kSpmvColMask  = 0x3fffffffu;
constexpr int      kSpmvMaxCap   = 1 << 14;      // tpos has 14 bits
constexpr int      kSpmvAlign    = 4;            // entries (one float4 of a value plane)
constexpr int      kSpmvSlack    = 8;            // >= the longest run of one thread
constexpr int      kSpmvPadSrc   = 0x7fffffff;   // source code of a padding entry: zero values

}  // namespace masx
